/*
 * tpl_xform.h -- C ABI of the MI355X (gfx950) coordinate transforms
 * (torchproteinlibrary_amd/libtpl_xform.so).
 *
 * Drop-in boundary for TorchProteinLibrary's FullAtomModel.CoordsTransform GPU
 * ops (FullAtomModel/CoordsTransform/CoordsTransform.py:46-213), whose native
 * entry points are CoordsTranslateGPU_forward/backward, CoordsRotateGPU_forward/
 * backward and Coords2CenterGPU_forward/backward
 * (Layers/FullAtomModel/main.cpp:23-36,
 * Layers/FullAtomModel/CoordsTransform/coordsTransformGPU_interface.cpp:5-146,
 * kernels Layers/FullAtomModel/TransformCUDAKernels.cu:3-108).
 *
 * Conventions (same meaning as the reference):
 *   coords      [batch][3*atoms_stride]  xyz interleaved; atoms >= num_atoms[b] ignored
 *   num_atoms   int32 [batch]            atoms per structure, <= atoms_stride
 *   T           [batch][3]               translation
 *   R           [batch][3][3]            rotation, applied as R x (row-major)
 *   center      [batch][3]               mean of the first num_atoms atoms
 * Every output slot of an atom >= num_atoms is written 0 (the reference's
 * outputs are zero-initialised by its Python layer).
 *
 * dtype is TPL_XFORM_F32 or TPL_XFORM_F64 for every array except num_atoms.
 * Sums (the centre) are taken in FP64.  Streams, status words, return codes:
 * as in tpl_a2b.h; nothing allocates or synchronises.
 */
#ifndef TPL_XFORM_H
#define TPL_XFORM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *tpl_xform_stream_t; /* a hipStream_t */

enum tpl_xform_dtype { TPL_XFORM_F32 = 0, TPL_XFORM_F64 = 1 };

enum tpl_xform_error {
    TPL_XFORM_OK = 0,
    TPL_XFORM_ERR_ARG = 1,    /* null pointer, negative size or unknown dtype */
    TPL_XFORM_ERR_LAUNCH = 3, /* a kernel launch failed */
    TPL_XFORM_ERR_SIZE = 5    /* sizes too large */
};

#define TPL_XFORM_STATUS_COUNT 2u /* some num_atoms[b] < 0 or > atoms_stride (clamped) */

const char *tpl_xform_strerror(int code);

/* out = in + T[b] per atom.  Replaces CoordsTranslateGPU_forward
 * (coordsTransformGPU_interface.cpp:5-25, TransformCUDAKernels.cu:3-18). */
int tpl_xform_translate(int dtype, void *out, const void *in, const void *T,
                        const int32_t *num_atoms, uint32_t *status, int batch, int atoms_stride,
                        tpl_xform_stream_t stream);

/* grad_in = grad_out per atom (the translation gets no gradient, as in the
 * reference).  Replaces CoordsTranslateGPU_backward (:27-46, TransformCUDAKernels.cu:21-37). */
int tpl_xform_translate_backward(int dtype, void *grad_in, const void *grad_out,
                                 const int32_t *num_atoms, uint32_t *status, int batch,
                                 int atoms_stride, tpl_xform_stream_t stream);

/* out = R[b] in (transpose = 0) or R[b]^T in (transpose = 1, the backward of
 * the rotation; R gets no gradient, as in the reference).  Replaces
 * CoordsRotateGPU_forward/backward (:48-104, TransformCUDAKernels.cu:40-75). */
int tpl_xform_rotate(int dtype, void *out, const void *in, const void *R, int transpose,
                     const int32_t *num_atoms, uint32_t *status, int batch, int atoms_stride,
                     tpl_xform_stream_t stream);

/* center[b] = mean of the first num_atoms[b] atoms (0/0 -> NaN, as the
 * reference's 1/n).  Replaces Coords2CenterGPU_forward (:106-126,
 * TransformCUDAKernels.cu:77-86: one thread per structure there). */
int tpl_xform_center(int dtype, void *center, const void *in, const int32_t *num_atoms,
                     uint32_t *status, int batch, int atoms_stride, tpl_xform_stream_t stream);

/* grad_in = grad_center[b] / num_atoms[b] per atom.  Replaces
 * Coords2CenterGPU_backward (:128-146, TransformCUDAKernels.cu:88-98). */
int tpl_xform_center_backward(int dtype, void *grad_in, const void *grad_center,
                              const int32_t *num_atoms, uint32_t *status, int batch,
                              int atoms_stride, tpl_xform_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TPL_XFORM_H */
