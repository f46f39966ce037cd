// Mixed-precision form of the energy/force path:
//   * symmetry functions in FP32 with compensated (Neumaier) pair sums,
//     scaling in FP64;
//   * atomic NNs as BF16 MFMA GEMMs with FP32 accumulation over all atoms
//     of an element, adjoint dE/dG by the transposed GEMMs;
//   * a fused force kernel that recomputes the radial and angular terms,
//     contracts them with dE/dG on the fly (no dG/dr tensor) and
//     accumulates forces in FP64;
//   * per-structure energies reduced in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

#include "kernels.hpp"
#include "model.hpp"

namespace nnpx {

// FP32 copy of one group member's constants.
struct MemberF {
    float eta, rs, lambda, zeta;
    float factorNorm, factorDeriv, zetaLambda, gnorm;
    float sfac;
    int sf, calcExp, zetaInt;
    int useIntPow, slot1, slot2, pad;
};

struct MixedModel {
    int ldK = 0;                // padded input width (multiple of 32)
    int ldH = 0;                // padded hidden width (multiple of 32)
    MemberF* d_mem = nullptr;   // [nmembers]
    uint16_t* d_w = nullptr;    // BF16 weights per element/layer, padded tiles
    float* d_b = nullptr;       // FP32 biases per element/layer, padded
    int woff[kMaxElem][kMaxLayers] = {};  // offsets into d_w (elements of BF16)
    int boff[kMaxElem][kMaxLayers] = {};  // offsets into d_b
    int kin[kMaxElem][kMaxLayers] = {};   // padded in-width of layer l
    int kout[kMaxElem][kMaxLayers] = {};  // padded out-width of layer l
    bool ready = false;
};

struct MixedBuffers {
    int ldG;
    double* G;         // [N*ldG] scaled SFs (FP64 scaling of FP32 sums)
    uint16_t* Gbf;     // [N*ldK] BF16 NN input
    double* dEdG;      // [N*ldG] dE/dG (FP64 copy of the FP32 adjoint)
    float* dEdGf;      // [N*ldG] dE/dG * scaling factor, FP32, for the force kernel
    double* Eatom;     // [N]
    double* E;         // [nstruct]
    double* F;         // [3N]
    double* Fpart;     // [3*N*cap] per-entry force contributions
    const int* elem_atoms;
    const int* atom_off;
    int elem_off[kMaxElem + 1];
    int nstruct;
    int cap;
};

using TimedFn = std::function<void(const char*, const std::function<void()>&)>;

void setup_mixed(MixedModel& mm, const HostModel& hm);
void free_mixed(MixedModel& mm);
void run_mixed(const MixedModel& mm, const ModelT& t, const SysView& s, const MixedBuffers& b,
               bool forces, bool push, hipStream_t stream, const TimedFn& timed);

}  // namespace nnpx

struct nnpx_ctx;
nnpx::MixedBuffers mixed_buffers(nnpx_ctx* c);
