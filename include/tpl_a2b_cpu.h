/*
 * tpl_a2b_cpu.h -- C ABI of the host (CPU) Angles2Backbone library
 * (torchproteinlibrary_amd/libtpl_a2b_cpu.so), FP64.
 *
 * The reference layer sends CPU tensors to Angles2BackboneCPUFunction
 * (TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:54-93,
 * :101-104), which binds Angles2BackboneCPU_forward / _backward
 * (Layers/ReducedModel/Angles2Backbone/angles2backbone_interface.cpp:60-111)
 * over cpu_computeCoordinatesBackbone, cpu_computeDerivativesBackbone and
 * cpu_backwardFromCoordsBackbone (Layers/ReducedModel/cBackboneProteinCPUKernels.cpp:14-145).
 * These two entry points replace them.  Same layouts as the GPU library
 * (include/tpl_a2b.h), all arrays float64:
 *   angles [batch][3][angles_stride], coords [batch][9*angles_stride],
 *   grad_angles [batch][3][angles_stride], lengths int32 [batch].
 * The reference computes only in float64 on the CPU; so does this library.
 *
 * The forward is the serial FP64 recurrence per chain; the backward is the
 * O(L) adjoint from the forward's coordinates,
 *   dL/dtheta_j = u_j . sum_{i>=j} (x_i - x_j) x g_i,
 * where the reference materialises a [batch][3][9 L^2] derivative tensor
 * (276 MB at batch 32, L 200) and reduces it.  Chains are spread over
 * `threads` host threads (0: one per hardware thread); each call starts and
 * joins its own threads and allocates nothing else.
 *
 * Outputs past each chain are written 0; the gradients of psi[len-1] and
 * omega[len-1] are exactly 0 (the reference reads past the chain there).
 * Lengths outside [0, angles_stride] are clamped and reported in *status
 * (TPL_A2B_STATUS_LENGTH); a NaN output sets TPL_A2B_STATUS_NAN.  status may
 * be NULL.  Return value: TPL_A2B_OK or TPL_A2B_ERR_ARG / TPL_A2B_ERR_SIZE
 * (include/tpl_a2b.h).
 */
#ifndef TPL_A2B_CPU_H
#define TPL_A2B_CPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Angles2BackboneCPU_forward (angles2backbone_interface.cpp:60-80)
 * and cpu_computeCoordinatesBackbone (cBackboneProteinCPUKernels.cpp:14-60). */
int tpl_a2b_cpu_forward(const double *angles, double *coords, const int32_t *lengths,
                        uint32_t *status, int batch, int angles_stride, int threads);

/* Replaces Angles2BackboneCPU_backward (angles2backbone_interface.cpp:83-111)
 * and cpu_computeDerivativesBackbone + cpu_backwardFromCoordsBackbone
 * (cBackboneProteinCPUKernels.cpp:62-145).  `coords` is the forward's output
 * for the same angles and lengths (it stands where the reference's frame
 * buffer A stood). */
int tpl_a2b_cpu_backward(double *grad_angles, const double *grad_coords, const double *coords,
                         const int32_t *lengths, uint32_t *status, int batch, int angles_stride,
                         int threads);

#ifdef __cplusplus
}
#endif

#endif /* TPL_A2B_CPU_H */
