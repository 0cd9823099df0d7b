/*
 * tpl_pdb.h -- C ABI of the PDB reader (torchproteinlibrary_amd/libtpl_pdb.so,
 * host code: parsing text is CPU work and feeds target coordinates to the GPU
 * layers).
 *
 * Drop-in boundary for TorchProteinLibrary's FullAtomModel.PDB2CoordsUnordered
 * (FullAtomModel/PDB2Coords/PDB2Coords.py:76-89), whose native entry point is
 * PDB2CoordsUnordered (Layers/FullAtomModel/main.cpp:21,
 * Layers/FullAtomModel/PDB2Coords/pdb2coords_interface.cpp:65-128) over
 * cPDBLoader (Layers/FullAtomModel/cPDBLoader.cpp:23-53).
 *
 * Records read, as there: lines whose first four characters are "ATOM", whose
 * atom name (columns 13-16, trimmed) starts with C, N, O or S; x, y, z from
 * columns 31-38, 39-46, 47-54 (parsed in single precision, then widened, as
 * the reference's std::stof does); residue name columns 18-20, chain column
 * 22, residue number columns 23-26.  Names are stored 0-padded in 4 bytes.
 */
#ifndef TPL_PDB_H
#define TPL_PDB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tpl_pdb_error {
    TPL_PDB_ERR_OPEN = -1,     /* the file cannot be opened */
    TPL_PDB_ERR_RECORD = -2,   /* a heavy-atom ATOM record is too short or unparsable */
    TPL_PDB_ERR_ARG = -3       /* null pointer or negative size */
};

const char *tpl_pdb_strerror(int code);

/* Number of heavy-atom ATOM records of a file, or a negative error code. */
int tpl_pdb_count(const char *filename);

/* The first max_atoms heavy-atom ATOM records of a file: coords [3*max_atoms]
 * (xyz per atom), chain_names / res_names / atom_names [max_atoms][4],
 * res_nums [max_atoms].  Returns the number of records written, or a
 * negative error code.  Replaces cPDBLoader's constructor plus the copy loop
 * of PDB2CoordsUnordered (pdb2coords_interface.cpp:103-125). */
int tpl_pdb_read(const char *filename, double *coords, uint8_t *chain_names, uint8_t *res_names,
                 int32_t *res_nums, uint8_t *atom_names, int max_atoms);

#ifdef __cplusplus
}
#endif

#endif /* TPL_PDB_H */
