// Host build of the residue templates in torchproteinlibrary_amd/csrc/
// a2c_table.h, for tests/test_a2c_table.py: prints every residue's atoms
// (name, local position) and side-chain nodes so that the test can compare
// them with the oracle's restatement.
#include <cstdio>

#include "a2c_table.h"

using namespace a2c;

static const Residue kRes[kTypes] = A2C_RESIDUES;
static const Side kSide[kTypes][kMaxSide] = A2C_SIDES;
static const Atom kAtom[kTypes][kMaxAtoms] = A2C_ATOMS;

int main() {
    for (int t = 0; t < kTypes; ++t) {
        const Residue& r = kRes[t];
        std::printf("R %c %s %d %d %d\n", kCodes[t], r.name3, r.natoms, r.nca, r.nside);
        for (int k = 0; k < r.natoms; ++k) {
            double v[3];
            local_position(kAtom[t][k], v);
            std::printf("A %c %d %s %.17g %.17g %.17g\n", kCodes[t], k, kAtom[t][k].name, v[0], v[1],
                        v[2]);
        }
        for (int k = 0; k < r.nside; ++k) {
            const Side& s = kSide[t][k];
            std::printf("S %c %d %d %.17g %.17g %.17g %d %d\n", kCodes[t], k, s.row, s.alpha, s.beta,
                        s.d, s.first, s.count);
        }
        if (residue_type((uint8_t)kCodes[t]) != t) std::printf("E type %c\n", kCodes[t]);
    }
    return 0;
}
