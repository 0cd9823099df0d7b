// ref_shim.cpp -- C-ABI entry points over the reference's own CPU kernels.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's CPU sources, compiled where they lie under /root/reference
// (Layers/ReducedModel/cBackboneProteinCPUKernels.cpp, Math/cMatrix44.cpp,
// Math/cMatrix33.cpp, Math/cVector3.cpp), into oracle/_ref/.  No reference
// source is copied here: the template declarations come from the reference's
// own header Layers/ReducedModel/cBackboneProteinCPUKernels.hpp via -I.
//
// The wrappers do exactly what Layers/ReducedModel/Angles2Backbone/
// angles2backbone_interface.cpp:60-111 does after its tensor checks: call the
// FP64 template instantiations on raw pointers.  The caller owns every buffer
// and zero-fills A and dr_dangle as TorchProteinLibrary/ReducedModel/
// Angles2Backbone/Angles2Backbone.py:64-65,83-84 does.
#include <cstdint>
#include "cBackboneProteinCPUKernels.hpp"

extern "C" {

// Angles2BackboneCPU_forward (angles2backbone_interface.cpp:60-79).
// angles [B][3][stride], coords [B][9*stride], A [B][48*stride] (+16 spare).
void tplref_forward(double* angles, double* coords, double* A, int32_t* lengths,
                    int batch, int stride) {
    cpu_computeCoordinatesBackbone<double>(angles, coords, A, lengths, batch, stride);
}

// Angles2BackboneCPU_backward (angles2backbone_interface.cpp:82-112).
// grad_angles [B][3][stride] (zeroed), grad_coords [B][9*stride],
// dr_dangle [B][3][9*stride*stride] (zeroed).
void tplref_backward(double* grad_angles, double* grad_coords, double* angles,
                     int32_t* lengths, double* A, double* dr_dangle, int batch, int stride) {
    cpu_computeDerivativesBackbone<double>(angles, dr_dangle, A, lengths, batch, stride);
    cpu_backwardFromCoordsBackbone<double>(grad_angles, grad_coords, dr_dangle, lengths,
                                           batch, stride);
}

}  // extern "C"
