/*
 * tpl_rmsd.h -- C ABI of the MI355X (gfx950) Coords2RMSD library
 * (torchproteinlibrary_amd/libtpl_rmsd.so).
 *
 * Drop-in boundary for TorchProteinLibrary's RMSD.Coords2RMSD: the minimal
 * RMSD over rigid rotations between two batches of structures, and its
 * gradient with respect to the first.  The reference splits the layer over
 * three Python autograd functions: Coords2Center and CoordsTranslate
 * (FullAtomModel/CoordsTransform/CoordsTransform.py:46-98,158-213), then
 * Coords2RMSDFunction (RMSD/Coords2RMSD/Coords2RMSD.py:12-62).  Its only native
 * entry point is Coords2RMSDGPU_forward (Layers/RMSD/main.cpp:4-7,
 * Layers/RMSD/Coords2RMSD/coords2rmsd_interface.cpp:52-111).  That entry
 * point computes the correlation matrix on the GPU, copies it to the host,
 * and solves one 4x4 eigenproblem per structure there.  Here both directions
 * are single device kernels: centring, moments, the eigenproblem, the RMSD
 * and the gradient stay on the GPU.
 *
 * Conventions (same meaning as the reference):
 *   input, target  [batch][3*atoms_stride]  xyz interleaved; atoms >= num_atoms[b] ignored
 *   num_atoms      int32 [batch]            atoms per structure, <= atoms_stride
 *   rmsd           [batch]                  sqrt((R2 - 2*lambda_max + 1e-10) / n)
 *   ut             [batch][3][3]            U^T, U the rotation taking the centred
 *                                           input onto the centred target
 *                                           (coords2rmsd_interface.cpp:90-93)
 *   grad_input     [batch][3*atoms_stride]  (x_c - U^T y_c) * grad_rmsd / (rmsd * n);
 *                                           0 for atoms >= num_atoms
 * The layer centres both structures itself, so already centred inputs (the
 * reference Function's contract) and raw ones (its Module's) give the same
 * result.
 *
 * dtype is TPL_RMSD_F32 or TPL_RMSD_F64 for every array except num_atoms.
 * The kernels compute in FP64.  Streams, status words, return codes and
 * buffer rules are those of tpl_a2b.h: nothing allocates or synchronises,
 * everything is capture-safe, and `state` starts on a 256-byte boundary.
 */
#ifndef TPL_RMSD_H
#define TPL_RMSD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *tpl_rmsd_stream_t; /* a hipStream_t */

enum tpl_rmsd_dtype { TPL_RMSD_F32 = 0, TPL_RMSD_F64 = 1 };

enum tpl_rmsd_error {
    TPL_RMSD_OK = 0,
    TPL_RMSD_ERR_ARG = 1,    /* null pointer, negative size or unknown dtype */
    TPL_RMSD_ERR_LAUNCH = 3, /* a kernel launch failed */
    TPL_RMSD_ERR_SIZE = 5    /* batch * atoms_stride too large */
};

/* Bits OR-ed into *status by the kernels. */
#define TPL_RMSD_STATUS_NAN 1u    /* an output element is NaN */
#define TPL_RMSD_STATUS_COUNT 2u  /* some num_atoms[b] < 0 or > atoms_stride (clamped) */

const char *tpl_rmsd_strerror(int code);

/* Bytes of the forward->backward state: per structure the two centres, U and
 * the RMSD in FP64 (16 doubles).  It replaces the reference's saved centred
 * copies of both inputs (Coords2RMSD.py:37). */
size_t tpl_rmsd_state_bytes(int batch);

/* Forward: rmsd and ut per structure; state for the backward (may be NULL).
 * Replaces Coords2RMSDGPU_forward (coords2rmsd_interface.cpp:52-111) and the
 * centring ops in front of it (CoordsTransform.py:158-213, 46-98). */
int tpl_rmsd_forward(int dtype, const void *input, const void *target, const int32_t *num_atoms,
                     void *rmsd, void *ut, void *state, uint32_t *status, int batch,
                     int atoms_stride, tpl_rmsd_stream_t stream);

/* Backward: d(loss)/d(input) from d(loss)/d(rmsd), with the forward's state.
 * Replaces Coords2RMSDFunction.backward (Coords2RMSD.py:40-62: a CoordsRotate
 * launch, a per-structure host loop and a host NaN check).  grad_input is
 * fully overwritten. */
int tpl_rmsd_backward(int dtype, void *grad_input, const void *grad_rmsd, const void *input,
                      const void *target, const int32_t *num_atoms, const void *state,
                      uint32_t *status, int batch, int atoms_stride, tpl_rmsd_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TPL_RMSD_H */
