// Adjoint (path replay) of the path tracer: device state and launch.
#pragma once
#include "pt_device.hpp"

namespace mfh {

constexpr int kAdjMaxVertices = 32;  // longest path the replay records

struct AdjointState {
    const float* delta;    // dLoss/dRadiance, 3 floats per pixel
    int iterations;        // iterations of the primal image being differentiated
    uint64_t* rngs;        // replay RNG states (seeded as a reset seeds them)
    float* const* texGrad; // per texture id: gradient buffer (4 floats per texel) or null
    float* emissionGrad;   // 3 floats per material: d loss / d emission scale
    int fixedDecisions;    // 0: derivative of the expected image; 1: of this realization (MFH_ADJOINT_*)
};

void launch_pt_adjoint(const PtScene& S, const AdjointState& A, int seed, hipStream_t stream);

}  // namespace mfh
