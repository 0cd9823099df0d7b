// ref_a2c_shim.cpp -- C-ABI entry points over the reference's own CPU
// full-atom conformation code (FullAtomModel.Angles2Coords).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's sources, compiled where they lie under /root/reference
// (Layers/FullAtomModel/cConformation.cpp, cConformationAA.cpp,
// cRigidGroup.cpp, cGeometry.cpp and Math/*.cpp; Math/nUtil.cpp is compiled
// against the installed torch headers and libraries, which the reference's
// cConformation.cpp needs for its PDB writer), into oracle/_ref/.  No
// reference source is copied here: the class declarations come from the
// reference's headers via -I.
//
// The wrappers do what Layers/FullAtomModel/Angles2Coords/
// angles2coords_interface.cpp:30-68 and :72-112 do per structure after their
// tensor checks: build a cConformation<double> over raw pointers, read its
// atom names, or run its backward.
#include <cstdint>
#include <cstring>
#include <string>

#include "cConformation.h"

extern "C" {

// Coordinates of one sequence: angles [8][stride] (rows phi, psi, omega,
// xi1..xi5), coords [3 * num_atoms].  Returns the number of atoms.
int tplref_a2c_forward(const char* seq, double* angles, int stride, double* coords) {
    std::string aa(seq);
    double* dummy = new double[8 * (size_t)stride]();
    int n = 0;
    {
        cConformation<double> conf(aa, angles, dummy, (uint)stride, coords);
        n = (int)conf.num_atoms;
    }
    delete[] dummy;
    return n;
}

// Atom names (4 bytes each, 0-padded) and one-letter residue names of one
// sequence, in output order.  Returns the number of atoms.
int tplref_a2c_names(const char* seq, int stride, char* atom_names, char* res_names) {
    std::string aa(seq);
    double* angles = new double[8 * (size_t)stride]();
    double* dummy = new double[8 * (size_t)stride]();
    double* coords = new double[3 * 16 * (size_t)(aa.size() + 1)]();
    int n = 0;
    {
        cConformation<double> conf(aa, angles, dummy, (uint)stride, coords);
        n = (int)conf.num_atoms;
        for (size_t g = 0; g < conf.groups.size(); ++g)
            for (size_t k = 0; k < conf.groups[g]->atomNames.size(); ++k) {
                const int idx = (int)conf.groups[g]->atomIndexes[k];
                std::strncpy(atom_names + 4 * idx, conf.groups[g]->atomNames[k].c_str(), 4);
                res_names[idx] = conf.groups[g]->residueName;
            }
    }
    delete[] angles;
    delete[] dummy;
    delete[] coords;
    return n;
}

// Gradient of sum(grad_coords * coords) over angles [8][stride] into
// grad_angles [8][stride] (entries the conformation does not bind keep
// their value; the caller zero-fills).
void tplref_a2c_backward(const char* seq, double* angles, int stride, double* grad_coords,
                         double* grad_angles) {
    std::string aa(seq);
    double* coords = new double[3 * 16 * (size_t)(aa.size() + 1)]();
    {
        cConformation<double> conf(aa, angles, grad_angles, (uint)stride, coords);
        conf.backward(conf.root, grad_coords);
    }
    delete[] coords;
}

}  // extern "C"
