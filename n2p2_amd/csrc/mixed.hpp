// Mixed-precision form of the energy/force path:
//   * symmetry functions in FP32 with compensated (Kahan) pair sums per lane,
//     lanes combined and scaled in FP64;
//   * atomic NNs as BF16 MFMA GEMMs (v_mfma_f32_16x16x16_bf16) with FP32
//     accumulation over all atoms of an element; the adjoint dE/dG runs as
//     the transposed GEMMs.  Operands are split into bf16 hi + lo parts and
//     the three leading products are accumulated (1 term = plain BF16);
//   * a fused force kernel that recomputes the radial and angular terms and
//     contracts them with dE/dG on the fly -- per triplet three scalars
//     replace the per-member dG/dr -- with FP64 force accumulation;
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
    float gnorm;       // angular: 2^(1-zeta) (G normalisation)
    float twoEta;      // 2 eta
    float zetaLambda;  // zeta * lambda
    float etaL;        // eta * log2(e): exponent factor for exp2
    int sf, calcExp, sidx;  // sidx: index of lambda in the group's lambda table
    int beta;               // power path: index of the member's (eta, lambda, zeta) term, else -1
};

// Per-group table of the distinct lambda values (members index it by sidx).
// Power path (angular groups whose zeta are powers of two up to 16, with at
// most two distinct eta and two lambda): the group's terms are
// exp(-eta_h sum r^2) * (1 + lambda_s cos)^(2^b), indexed
// beta = s * NB0 + b for h = 0 and 2 NB0 + s * NB1 + b for h = 1.
struct GroupMixF {
    float lam[4];
    int nlam;
    int used;          // power path: bit beta set for every term a member uses
    float etaL[2];     // power path: eta_h * log2(e)
    float twoEta[2];   // power path: 2 eta_h
    int pad[2];
};

// FP64 constants of a power-path angular group for the plain form's fused
// kernels (same term indexing as GroupMixF).
struct GroupPow64 {
    double eta[2];     // eta_h
    double twoEta[2];  // 2 eta_h
    double lam[2];     // lambda_s
    int used;          // bit beta set for every term a member uses
    int pad;
};

// LDS staging of the mixed kernels: float4 slots of the per-group constant
// area (member constants, two per member, on the member paths; one per term
// on the power path) ...
__host__ __device__ constexpr int stage_mp(int nb0, int nb1)
{
    return nb0 > 0 ? 2 * (nb0 + nb1) : 2 * kMaxMembersExp;
}
// ... and bytes of the whole staging area: per slot geometry, cutoffs, two
// list entries and the element; the triplet buffer (trip entries and up to a
// wave's worth carried past them); the constants; the element offsets (capL
// is a multiple of 32 and trip of 64, so the byte arrays end 16-byte
// aligned, as does the offset table)
__host__ __device__ constexpr size_t stage_bytes(int capL, int trip, int mpn)
{
    return (size_t)capL * (32 + 2 + 2 + 1) + (size_t)(trip + kWave) * 4 + (size_t)mpn * 16 + 48;
}

// Each atom's compacted angular pairs in device memory, walk by walk (the
// walks of its element in group order, a merged walk counting as one): the
// SF kernel stores them as it tests the candidates, the force kernel reads
// them back instead of testing again.  A walk whose pairs do not fit keeps
// count -1 and the force kernel tests its candidates itself, which visits
// the same pairs in the same buffers.  Without lists the force kernel may
// instead walk rows, whose order suits some models better (the tuning
// decides).
struct TripList {
    unsigned* list = nullptr;  // [natoms][cap] pairs j | k << 16 (nullptr: none kept)
    int* counts = nullptr;     // [natoms][ld] pairs of each walk, -1: not kept
    int cap = 0, ld = 0;
    bool count = false;        // record the largest per-atom pair count in flags[3]
    bool rows = false;         // the force kernel walks rows (no lists kept)
};

// Each atom's LDS staging kept in device memory by the SF kernel for the
// force kernel of the same evaluation, which loads it (coalesced) instead of
// gathering and sorting the neighbour list again: slots in staging order
// with per-atom stride ld.
struct StageKeep {
    float4* g = nullptr;         // [natoms][ld] x, y, z, |dr| (FP32)
    float* rinv = nullptr;       // [natoms][ld] 1/|dr|
    int* nbr = nullptr;          // [natoms][ld] the entry's neighbour atom
    unsigned char* el = nullptr; // [natoms][ld] its element
    int* off = nullptr;          // [natoms][kMaxElem + 1] first slot of each element, then the slot count
    int ld = 0;
};

// Angular groups of one element walked together (power path): the groups
// share type, cutoff and term constants and differ only in their neighbour
// element pair and members, so one triplet walk over all in-cutoff
// neighbours serves them all, each triplet's terms going to the group of
// its pair.  ng = 0: the element's groups are walked one by one.
constexpr int kMaxMerge = 16;
struct AngMergeF {
    int ng;                                // merged groups
    int used;                              // union of the groups' term masks
    int grp[kMaxMerge];                    // their ModelT::grp indices
    signed char pair[kMaxElem * kMaxElem]; // (element j, element k) -> merged group, -1: none
};

// Launch shapes of the mixed-form kernels, set by the autotuner
// (nnpx_autotune) or left at these defaults.
struct Tuning {
    int trip_sf = 384;       // triplet buffer of the SF kernel (multiple of 64)
    int trip_force = 384;    // ... of the force kernel
    int trip_tangent = 384;  // ... of the tangent kernel
    int grad_atoms = 32;     // atoms per block of the gradient kernel
    int sf_gather = -1;      // SF kernel partner gather: -1 auto, 0 off, 1 on
    int merge_sf = -1;       // merged angular walk in the SF kernel: -1 auto, 0 off, 1 on
    int merge_force = -1;    // ... in the force kernel
    int pair_lists = -1;     // angular pair lists from the SF to the force kernel: -1 auto (off until tuned), 0 off, 1 on
};

struct MixedModel {
    int ldK = 0;                // padded NN input width (multiple of 16)
    int maxpad = 0;             // max padded layer width
    int maxpadh = 0;            // max padded width of the hidden and output layers
    int terms = 3;              // bf16 product terms (1 or 3)
    MemberF* d_mem = nullptr;   // [nmembers], same order as ModelT::mem
    GroupMixF* d_gmix = nullptr; // [ngroups]
    GroupPow64* d_gpow64 = nullptr;  // [ngroups] FP64 power-path constants
    int maxS = 2;               // max distinct lambda values per angular group (2 or 4)
    bool hasRs = false;         // some angular member has rs > 0
    int fastM = 0;              // exact angular member count of the branch-free path (0: general path)
    int powNB[2] = {0, 0};      // power path shape (NB0, NB1); {0, 0}: not used
    AngMergeF* d_amerge = nullptr;  // [nelem] merged angular walks (power path only)
    int merge_ng = 0;           // most merged groups of an element (0: no element merges)
    int max_walks = 0;          // most angular walks of an element (unmerged)
    double merge_avg = 0.0;     // mean angular groups per merged element
    float* d_sffac = nullptr;   // [nelem*kMaxSF] force weight factor per SF
    uint16_t* d_w = nullptr;    // bf16 weights: per element/layer W hi, W lo, WT hi, WT lo
    float* d_b = nullptr;       // FP32 biases, padded
    int wW[kMaxElem][kMaxLayers] = {};   // offset of W [out_pad][in_pad] (hi; lo follows)
    int wT[kMaxElem][kMaxLayers] = {};   // offset of WT [in_pad][out_pad] (hi; lo follows)
    int bo[kMaxElem][kMaxLayers] = {};   // offset of the bias
    int pin[kMaxElem][kMaxLayers] = {};  // padded in-width of layer l
    int pout[kMaxElem][kMaxLayers] = {}; // padded out-width of layer l
    Tuning tune;
    bool ready = false;
};

struct MixedBuffers {
    int ldG;
    double* G;         // [N*ldG] scaled SFs (FP64 scaling of FP32 sums)
    float* Gin;        // [N*ldK] FP32 NN input (split into bf16 operands by the NN kernel)
    double* dEdG;      // [N*ldG] dE/dG (FP64 copy of the FP32 adjoint)
    float* dEdGf;      // [N*ldG] dE/dG times the SF's force factor, FP32
    double* Eatom;     // [N]
    double* E;         // [nstruct]
    double* F;         // [3N]
    const int* elem_atoms;
    const int* atom_off;
    int elem_off[kMaxElem + 1];
    int nstruct;
    int cap;
    TripList trips;    // the angular pair lists shared by the SF and force kernels
    StageKeep keep;    // the staging shared by the SF and force kernels
};

using TimedFn = std::function<void(const char*, const std::function<void()>&)>;

void setup_mixed(MixedModel& mm, const HostModel& hm);
void free_mixed(MixedModel& mm);
void run_mixed(const MixedModel& mm, const ModelT& t, const ModelT* d_model, const SysView& s,
               const MixedBuffers& b, bool forces, bool push, hipStream_t stream, const TimedFn& timed);
// Tangent of the scaled SFs along a displacement field u [3N]:
// Gdot[i*ldG + q] = sum_a dG_iq/dx_a . u_a, recomputed in FP32.
void launch_tangent32(const MixedModel& mm, const ModelT& t, const ModelT* d_model, const SysView& s,
                      const double* u, double* Gdot, int ldG, hipStream_t stream);
// Checks the MFMA lane layout assumed by the NN kernels on a known product;
// returns the max abs error.
double mfma_selftest();

}  // namespace nnpx

struct nnpx_ctx;
nnpx::MixedBuffers mixed_buffers(nnpx_ctx* c);
