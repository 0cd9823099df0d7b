/*
 * tpl_a2c.h -- C ABI of the MI355X (gfx950) full-atom protein builder
 * (torchproteinlibrary_amd/libtpl_a2c.so).
 *
 * Drop-in boundary for TorchProteinLibrary's FullAtomModel.Angles2Coords
 * (FullAtomModel/Angles2Coords/Angles2Coords.py:34-87), whose native entry
 * points are Angles2Coords_forward / Angles2Coords_backward / getSeqNumAtoms
 * (Layers/FullAtomModel/main.cpp:10-14,
 * Layers/FullAtomModel/Angles2Coords/angles2coords_interface.cpp:8-144).
 * The reference runs them on the CPU, one cConformation tree per structure
 * (Layers/FullAtomModel/cConformation.cpp:29-195); here they run on the GPU.
 *
 * Conventions (same meaning as the reference):
 *   sequences   uint8 [batch][seq_stride]   one-letter codes, 0-terminated
 *                                           (or filling the row)
 *   angles      [batch][8][angles_stride]   rows phi, psi, omega, xi1..xi5 in
 *                                           radians; residue i's N hangs on
 *                                           omega of residue i-1
 *   coords      [batch][3*atoms_stride]     xyz per atom in the reference's
 *                                           atom order, 0 past the structure
 *   res_names,  uint8 [batch][atoms_stride][4]  0-padded names per atom
 *   atom_names                                  ("ALA", "CB"), 0 past the structure
 *   grad_angles [batch][8][angles_stride]   every entry written: 0 for angles
 *                                           the structure does not use
 *                                           (FP32 coordinates give gradients
 *                                           within about 1e-5 of FP64 ones)
 *
 * dtype is TPL_A2C_F32 or TPL_A2C_F64 for angles, coords and gradients; the
 * frames and sums are FP64 inside.  Streams, status words, return codes: as in
 * tpl_a2b.h; nothing allocates or synchronises.  A structure flagged in the
 * status word (unknown letter, too many residues or atoms for the strides)
 * gets all-zero outputs.
 */
#ifndef TPL_A2C_H
#define TPL_A2C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *tpl_a2c_stream_t; /* a hipStream_t */

enum tpl_a2c_dtype { TPL_A2C_F32 = 0, TPL_A2C_F64 = 1 };

enum tpl_a2c_error {
    TPL_A2C_OK = 0,
    TPL_A2C_ERR_ARG = 1,    /* null pointer, negative size or unknown dtype */
    TPL_A2C_ERR_LAUNCH = 3, /* a kernel launch failed */
    TPL_A2C_ERR_SIZE = 5    /* sizes too large */
};

#define TPL_A2C_STATUS_SEQ 1u    /* a sequence holds a letter outside the 20 amino acids */
#define TPL_A2C_STATUS_LENGTH 2u /* a sequence is longer than angles_stride */
#define TPL_A2C_STATUS_ATOMS 4u  /* a structure has more atoms than atoms_stride */
#define TPL_A2C_STATUS_ORDER 8u  /* an order entry outside 0..batch-1 (that workgroup did nothing) */

const char *tpl_a2c_strerror(int code);

/* Atoms of a 0-terminated sequence, or -1 if it holds a letter outside the
 * 20 amino acids.  Replaces getSeqNumAtoms (angles2coords_interface.cpp:
 * 138-143, nUtil.cpp:174-190). */
int tpl_a2c_num_atoms(const char *sequence);

/* Coordinates (and, where the pointers are not null, residue and atom names)
 * of each structure.  Replaces Angles2Coords_forward
 * (angles2coords_interface.cpp:8-69). */
int tpl_a2c_forward(int dtype, const uint8_t *sequences, int seq_stride, const void *angles,
                    int angles_stride, void *coords, uint8_t *res_names, uint8_t *atom_names,
                    int atoms_stride, uint32_t *status, int batch, tpl_a2c_stream_t stream);

/* grad_angles = d(sum grad_coords . coords) / d angles, from the forward's
 * coordinates (coords, as tpl_a2c_forward wrote them for the same sequences
 * and angles): each bound angle turns its subtree about a bond whose
 * direction and origin are atoms of coords, so no angles are needed.
 * Replaces Angles2Coords_backward (angles2coords_interface.cpp:71-111, which
 * rebuilds the conformation from the angles; cConformation.cpp:160-195:
 * O(atoms x depth) there, O(atoms) here). */
int tpl_a2c_backward(int dtype, const uint8_t *sequences, int seq_stride, const void *coords,
                     const void *grad_coords, int atoms_stride, void *grad_angles, int angles_stride,
                     uint32_t *status, int batch, tpl_a2c_stream_t stream);

/* Backward kernels.  TPL_A2C_BWD_RUNS: each thread walks a run of residues
 * (fastest when the batch is small against the GPU).  TPL_A2C_BWD_TILES:
 * the structure is walked in tiles of 1024 consecutive atoms, one per
 * thread pair, each atom read once with coalesced loads (fastest when the
 * batch fills the GPU; angles_stride at most 2048, else TPL_A2C_ERR_SIZE).
 * TPL_A2C_BWD_AUTO picks TILES for batch * angles_stride >= 100000 within
 * that limit.  Both give the same gradients up to rounding. */
#define TPL_A2C_BWD_AUTO 0
#define TPL_A2C_BWD_RUNS 1
#define TPL_A2C_BWD_TILES 2

/* Scheduling order.  One workgroup builds one structure, and workgroups
 * start in index order; with mixed lengths the long structures that happen
 * to come last keep the GPU busy after the rest is done.  With `order` (a
 * device int32 array holding a permutation of 0..batch-1, or NULL for
 * 0..batch-1) workgroup w takes structure order[w]; the outputs are the same
 * bit for bit.  Longest first (the atom counts sorted in descending order,
 * which the caller knows from the sequences) is the intended use.  An entry
 * outside 0..batch-1 sets TPL_A2C_STATUS_ORDER and its workgroup does
 * nothing; a repeated entry leaves the structures it displaces unwritten.
 * No counterpart in the reference (its CPU loop has no such order). */
int tpl_a2c_forward_ordered(const int32_t *order, int dtype, const uint8_t *sequences, int seq_stride,
                            const void *angles, int angles_stride, void *coords, uint8_t *res_names,
                            uint8_t *atom_names, int atoms_stride, uint32_t *status, int batch,
                            tpl_a2c_stream_t stream);

/* tpl_a2c_backward with the kernel chosen by `form` and the scheduling
 * order above; tpl_a2c_backward is TPL_A2C_BWD_AUTO without an order.  No
 * counterpart in the reference: for callers that pin a kernel or pass an
 * order (the Python layer, tests comparing the kernels, timing). */
int tpl_a2c_backward_form(int form, const int32_t *order, int dtype, const uint8_t *sequences,
                          int seq_stride, const void *coords, const void *grad_coords, int atoms_stride,
                          void *grad_angles, int angles_stride, uint32_t *status, int batch,
                          tpl_a2c_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TPL_A2C_H */
