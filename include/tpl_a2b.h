/*
 * tpl_a2b.h -- C ABI of the MI355X (gfx950) Angles2Backbone library
 * (torchproteinlibrary_amd/libtpl_a2b.so).
 *
 * This is the drop-in boundary for TorchProteinLibrary's
 * ReducedModel.Angles2Backbone forward and backward.  The reference binds four
 * torch-typed functions through pybind11 (Layers/ReducedModel/main.cpp:4-9,
 * declared in Layers/ReducedModel/Angles2Backbone/angles2backbone_interface.h:3-29).
 * Their only caller is the Python layer
 * TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:25,47.
 * This header replaces the two GPU ones with plain-pointer entry points that
 * take a stream, allocate nothing and never synchronise the host.
 *
 * Conventions (same meaning as the reference):
 *   angles       [batch][3][angles_stride]  rows phi, psi, omega (radians)
 *   lengths      int32 [batch]              residues per chain, <= angles_stride
 *   coords       [batch][9*angles_stride]   atoms N0,CA0,C0,N1,... xyz interleaved
 *   grad_angles  [batch][3][angles_stride]
 * Every output slot outside a chain (residue >= length) is written 0.
 * The gradients of psi[len-1] and omega[len-1] are exactly 0: those angles move
 * no atom.  The reference instead reads frames past the chain there.
 *
 * dtype is TPL_A2B_F32 or TPL_A2B_F64 for the angle/coordinate/gradient
 * arrays.  The scan kernels always compute in FP64.
 *
 * Streams: `stream` is a hipStream_t (NULL = the legacy default stream).  All
 * work is enqueued on it in stream order.  No call synchronises or allocates,
 * and every call is safe under hipStreamBeginCapture.
 *
 * `state` and `sync` buffers must start on a 256-byte boundary (hipMalloc and
 * torch allocations do); a misaligned one returns TPL_A2B_ERR_ARG.
 *
 * Return value: TPL_A2B_OK (0) or an error code; tpl_a2b_strerror() names it.
 * Device-side conditions go to the optional `status` word as bits, because a
 * kernel cannot return an error.  Pass NULL to skip them.  The word is OR-ed
 * into, so the caller zeroes it before the call.
 */
#ifndef TPL_A2B_H
#define TPL_A2B_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *tpl_stream_t; /* a hipStream_t */

enum tpl_a2b_dtype { TPL_A2B_F32 = 0, TPL_A2B_F64 = 1 };

enum tpl_a2b_error {
    TPL_A2B_OK = 0,
    TPL_A2B_ERR_ARG = 1,    /* null pointer, negative size or unknown dtype */
    TPL_A2B_ERR_CONFIG = 2, /* unsupported tile configuration */
    TPL_A2B_ERR_LAUNCH = 3, /* a kernel launch failed */
    TPL_A2B_ERR_HIP = 4,    /* another HIP runtime call failed */
    TPL_A2B_ERR_SIZE = 5    /* batch * angles_stride too large */
};

/* Bits OR-ed into *status by the kernels. */
#define TPL_A2B_STATUS_NAN 1u    /* an output element is NaN */
#define TPL_A2B_STATUS_LENGTH 2u /* some lengths[b] < 0 or > angles_stride (clamped) */
#define TPL_A2B_STATUS_TIMEOUT 4u /* a look-back wait gave up (result invalid) */

/* Launch shape of the scan kernels.  A forward tile is 64*waves*res_per_lane
 * consecutive residues of one chain; one workgroup scans one tile, and chains
 * longer than a tile are chained across workgroups by decoupled look-back.
 * Every lane keeps its residues' inputs and outputs in registers, which bounds
 * the tile: at most 1024 residues for TPL_A2B_F32 and 512 for TPL_A2B_F64
 * (larger tiles return TPL_A2B_ERR_CONFIG).  The backward holds more per
 * residue and runs one residue per lane with the same waves. */
typedef struct tpl_a2b_config {
    int32_t waves;        /* 64-lane waves per workgroup: 1, 2 or 4 */
    int32_t res_per_lane; /* residues composed serially per lane: 1, 2, 4 or 8 */
    int32_t max_wgs;      /* forward grid cap (persistent workgroups walk the
                             tiles; with one tile per chain each loads its next
                             chain during the current one); 0 = one workgroup
                             per tile.  The backward runs one workgroup per tile. */
    int32_t flags;        /* reserved, must be 0 */
} tpl_a2b_config;

/* Human-readable name of an error code. */
const char *tpl_a2b_strerror(int code);

/* A configuration chosen from (batch, angles_stride) alone; used when no tuned
 * configuration is known. */
int tpl_a2b_default_config(int batch, int angles_stride, tpl_a2b_config *cfg);

/* Residues per tile of cfg, or -1 if cfg is invalid. */
int64_t tpl_a2b_tile_residues(const tpl_a2b_config *cfg);

/* Bytes of the forward->backward state for this dtype and config.  For
 * TPL_A2B_F32 it holds, per output coordinate, the rounding residual of the
 * forward's FP64 value as a scaled half (2 bytes), so that the backward
 * rebuilds FP64-grade coordinates from the FP32 ones; when chains span
 * several backward tiles it also holds the backward's look-back words.  The
 * forward writes all of it; the backward leaves its look-back words as it
 * found them, so one state can serve repeated backward calls.  It replaces
 * the reference's per-atom 4x4 frame buffer A (Angles2Backbone.py:21) and is
 * 0 for TPL_A2B_F64 when every chain fits in one backward tile. */
size_t tpl_a2b_state_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config *cfg);

/* Bytes of the forward's look-back scratch for one call (0 when every chain
 * fits in one forward tile).  The library clears what must start at zero
 * itself, with a kernel on `stream`; contents need not survive the call. */
size_t tpl_a2b_sync_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config *cfg);

/* Forward: angles -> coords.  Replaces Angles2BackboneGPU_forward
 * (Layers/ReducedModel/Angles2Backbone/angles2backbone_interface.cpp:7-26) and
 * its kernel computeCoordinatesBackbone
 * (Layers/ReducedModel/cBackboneProteinCUDAKernels.cu:6-41,164-167).
 * Writes `state` (tpl_a2b_state_bytes) for the backward; with state == NULL
 * (a forward that no backward follows) it skips that work. */
int tpl_a2b_forward(int dtype, const void *angles, void *coords, const int32_t *lengths,
                    void *state, void *sync, uint32_t *status, int batch, int angles_stride,
                    const tpl_a2b_config *cfg, tpl_stream_t stream);

/* Backward: d(loss)/d(coords) -> d(loss)/d(angles), O(L) reverse-scan
 * adjoint over the forward's coordinates.  Replaces Angles2BackboneGPU_backward
 * (angles2backbone_interface.cpp:27-56) and its kernels
 * computeGradientsOptimizedBackbone + backwardFromCoordinatesBackbone
 * (cBackboneProteinCUDAKernels.cu:84-163,169-178).  Where the reference passes
 * the input angles and the saved frames A, this takes the forward's output
 * `coords` and `state` as the forward wrote them for the same lengths and cfg.
 * grad_angles is fully overwritten. */
int tpl_a2b_backward(int dtype, void *grad_angles, const void *grad_coords, const void *coords,
                     const int32_t *lengths, void *state, uint32_t *status, int batch,
                     int angles_stride, const tpl_a2b_config *cfg, tpl_stream_t stream);

/* ---- plain form: the reference algorithm written as straightforward HIP ----
 * One thread per chain runs the serial recurrence and stores every atom's 3x4
 * frame.  The backward uses one thread per (chain, angle) with the per-pair
 * formula A_{m-1} dB_m A_m^{-1} x_i summed over downstream atoms (O(L^2) per
 * chain), without the dr_dangle tensor.  It computes in the I/O dtype.  It is
 * the comparison baseline for the scan kernels and an independent GPU check. */

/* Bytes of the per-atom frame buffer of the plain form. */
size_t tpl_a2b_plain_frames_bytes(int dtype, int batch, int angles_stride);

/* Plain forward (cBackboneProteinCUDAKernels.cu:6-41, on a grid instead of
 * <<<1,B>>>). */
int tpl_a2b_forward_plain(int dtype, const void *angles, void *coords, const int32_t *lengths,
                          void *frames, uint32_t *status, int batch, int angles_stride,
                          tpl_stream_t stream);

/* Plain backward (cBackboneProteinCUDAKernels.cu:84-163, without the
 * derivative tensor). */
int tpl_a2b_backward_plain(int dtype, void *grad_angles, const void *grad_coords,
                           const void *angles, const int32_t *lengths, const void *frames,
                           uint32_t *status, int batch, int angles_stride,
                           tpl_stream_t stream);

/* Autotuner: times forward+backward for each candidate on the given buffers
 * (median of `reps` runs after one warm-up, hipEvents on `stream`).  It
 * writes the fastest candidate to *best and each candidate's ms to
 * ms_per_candidate (may be NULL; invalid candidates get -1).  state and sync
 * must be sized for the largest candidate.  It synchronises the stream, so
 * it must not be called under stream capture.  It has no reference
 * counterpart: the reference has a single fixed launch shape. */
int tpl_a2b_tune(int dtype, const void *angles, const int32_t *lengths,
                 const void *grad_coords, void *coords, void *grad_angles, void *state,
                 void *sync, int batch, int angles_stride, const tpl_a2b_config *candidates,
                 int n_candidates, int reps, tpl_a2b_config *best, float *ms_per_candidate,
                 tpl_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TPL_A2B_H */
