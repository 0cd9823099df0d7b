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
 * arrays.  The chunk kernels always compute in FP64.
 *
 * Streams: `stream` is a hipStream_t (NULL = the legacy default stream).  All
 * work is enqueued on it in stream order.  No call synchronises or allocates,
 * and every call is safe under hipStreamBeginCapture.
 *
 * `state` and `sync` buffers must start on a 256-byte boundary (hipMalloc and
 * torch allocations do); a misaligned one returns TPL_A2B_ERR_ARG.
 *
 * Lengths outside [0, angles_stride] are clamped on the device and reported
 * in *status (the reference writes out of bounds there).
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
    TPL_A2B_ERR_CONFIG = 2, /* unsupported chunk configuration */
    TPL_A2B_ERR_LAUNCH = 3, /* a kernel launch failed */
    TPL_A2B_ERR_HIP = 4,    /* another HIP runtime call failed */
    TPL_A2B_ERR_SIZE = 5,   /* batch * angles_stride too large */
    TPL_A2B_ERR_SCRATCH = 6 /* too large for tpl_a2b_forward_noalloc's scratch */
};

/* Bits OR-ed into *status by the kernels. */
#define TPL_A2B_STATUS_NAN 1u    /* an output element is NaN */
#define TPL_A2B_STATUS_LENGTH 2u /* some lengths[b] < 0 or > angles_stride (clamped) */
#define TPL_A2B_STATUS_TIMEOUT 4u /* a look-back wait gave up (result invalid; clear the
                                     sync buffer before its next use) */
#define TPL_A2B_STATUS_SYNC 8u    /* the sync buffer was not all zero at the call (result
                                     invalid; clear it before its next use) */

/* Launch shape of the chunk kernels.  A forward chunk is 64*fwd_res_per_lane
 * consecutive residues of one chain and a backward chunk 64*bwd_res_per_lane;
 * one 64-lane wave takes one chunk, and the chunks of one chain are chained
 * by a decoupled look-back.  When a chain has at most 4 chunks and every
 * chain's workgroup fits on the device at once, one workgroup takes a whole
 * chain instead (a wave per chunk, the chunks meeting in LDS; the sync words
 * go unused).  The backward chunk is also the checkpoint stride: the forward
 * saves the FP64 frame at every backward-chunk boundary, so forward and
 * backward of one step must use the same config. */
typedef struct tpl_a2b_config {
    int32_t fwd_res_per_lane; /* 1, 2, 4 or 8 (TPL_A2B_F64: up to 4) */
    int32_t bwd_res_per_lane; /* 1, 2 or 4 (TPL_A2B_F64: up to 2) */
    int32_t flags;            /* reserved, must be 0 */
    int32_t reserved;         /* reserved, must be 0 */
} tpl_a2b_config;

/* Human-readable name of an error code. */
const char *tpl_a2b_strerror(int code);

/* A configuration chosen from (batch, angles_stride) alone; used when no tuned
 * configuration is known. */
int tpl_a2b_default_config(int batch, int angles_stride, tpl_a2b_config *cfg);

/* Residues per forward (backward = 0) or backward (backward != 0) chunk of
 * cfg, or -1 if cfg is invalid. */
int64_t tpl_a2b_chunk_residues(const tpl_a2b_config *cfg, int backward);

/* Bytes of the forward->backward state: one FP64 3x4 frame per backward chunk
 * of every chain (96 bytes per 64*bwd_res_per_lane residues), the checkpoints
 * from which the backward recomputes its frames.  It replaces the reference's
 * per-atom 4x4 frame buffer A (Angles2Backbone.py:21, 192 bytes per residue
 * in FP32). */
size_t tpl_a2b_state_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config *cfg);

/* Bytes of the look-back words of forward and backward (per-chain tickets
 * and done counters, then one record of 64-bit words per chunk, each word
 * marking itself written; laid out so that no 256-byte block holds words of
 * chains that run on different XCDs).  The buffer must be
 * all zero before its first use (tpl_a2b_sync_clear, or any zero fill); every
 * call that ends without TPL_A2B_STATUS_TIMEOUT returns all it wrote to zero,
 * so one buffer serves every later call on the same stream, of any shape and
 * config that fits in it.  Calls that overlap in time (different streams)
 * need different buffers.  A call that finds a ticket word out of range
 * reports TPL_A2B_STATUS_SYNC and leaves that chain's outputs unwritten. */
size_t tpl_a2b_sync_bytes(int dtype, int batch, int angles_stride, const tpl_a2b_config *cfg);

/* Zeroes `bytes` of `sync` with a kernel on `stream` (capture-safe: no
 * memset node).  bytes must be a multiple of 4. */
int tpl_a2b_sync_clear(void *sync, size_t bytes, tpl_stream_t stream);

/* Forward: angles -> coords.  Replaces Angles2BackboneGPU_forward
 * (Layers/ReducedModel/Angles2Backbone/angles2backbone_interface.cpp:7-26) and
 * its kernel computeCoordinatesBackbone
 * (Layers/ReducedModel/cBackboneProteinCUDAKernels.cu:6-41,164-167).
 * Writes the checkpoints into `state` (tpl_a2b_state_bytes) for the backward;
 * with state == NULL (a forward that no backward follows) it skips them. */
int tpl_a2b_forward(int dtype, const void *angles, void *coords, const int32_t *lengths,
                    void *state, void *sync, uint32_t *status, int batch, int angles_stride,
                    const tpl_a2b_config *cfg, tpl_stream_t stream);

/* Backward: d(loss)/d(coords) -> d(loss)/d(angles), the O(L) reverse-scan
 * adjoint.  Replaces Angles2BackboneGPU_backward
 * (angles2backbone_interface.cpp:27-56: gradInput, gradOutput, input_angles,
 * angles_length, A, dr_dangle) and its kernels
 * computeGradientsOptimizedBackbone + backwardFromCoordinatesBackbone
 * (cBackboneProteinCUDAKernels.cu:84-163,169-178), argument for argument:
 * `state` (the checkpoints the forward wrote for the same angles, lengths and
 * cfg) stands where A stood, `sync` where the O(L^2) dr_dangle stood.
 * grad_angles is fully overwritten. */
int tpl_a2b_backward(int dtype, void *grad_angles, const void *grad_coords, const void *angles,
                     const int32_t *lengths, const void *state, void *sync, uint32_t *status,
                     int batch, int angles_stride, const tpl_a2b_config *cfg,
                     tpl_stream_t stream);

/* Forward that allocates nothing and takes no workspace.  It has no
 * counterpart in the reference.  It is for callers that cannot allocate or
 * own a workspace: code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same as tpl_a2b_forward with state == NULL (no
 * backward follows), but its look-back words live in a device array inside
 * this library, tpl_a2b_noalloc_max_bytes() bytes (4 MiB) per GPU, held until
 * the process ends.  A problem whose forward look-back words exceed it
 * returns TPL_A2B_ERR_SCRATCH.  Those words are the tickets and the done
 * counters, 2 * 8 * ceil(ceil(batch / 8) / 64) * 256 bytes, then the
 * records, batch * ceil(12 * nc / 32) * 256 bytes with
 * nc = ceil(angles_stride / (64 * fwd_res_per_lane)); e.g. batch 1024 and
 * angles_stride 1000 take 0.53 MB with 4 residues per lane.  (Where one
 * workgroup takes a whole chain, as in tpl_a2b_forward, no look-back words
 * are used.)  Every call on
 * one GPU shares that array, so calls that overlap in time on one GPU
 * (different streams, or different threads without ordering) race and must
 * not be made.  Each call first
 * clears the words it uses, with a kernel on `stream`. */
int tpl_a2b_forward_noalloc(int dtype, const void *angles, void *coords, const int32_t *lengths,
                            uint32_t *status, int batch, int angles_stride,
                            const tpl_a2b_config *cfg, tpl_stream_t stream);

/* Bytes of tpl_a2b_forward_noalloc's built-in scratch. */
size_t tpl_a2b_noalloc_max_bytes(void);

/* ---- plain form: the reference algorithm written as straightforward HIP ----
 * One thread per chain runs the serial recurrence and stores every atom's 3x4
 * frame.  The backward uses one thread per (chain, angle) with the per-pair
 * formula A_{m-1} dB_m A_m^{-1} x_i summed over downstream atoms (O(L^2) per
 * chain), without the dr_dangle tensor.  It computes in the I/O dtype.  It is
 * the comparison baseline for the chunk kernels and an independent GPU check. */

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
 * (hipEvents on `stream` around 6 back-to-back steps; the candidates take
 * turns for `reps` passes after one warm-up pass, and each gets the median
 * of its passes, in ms per step).  It
 * writes the fastest candidate to *best and each candidate's ms to
 * ms_per_candidate (may be NULL; invalid candidates get -1).  state and sync
 * must be sized for the largest candidate (sync zeroed).  It synchronises the stream, so
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
