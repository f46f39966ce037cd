// Host-side launchers of the gfx950 kernels (all asynchronous on `stream`).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_common.hpp"
#include "model.hpp"

namespace nnpx {

// Device view of the structures and the neighbour list.
struct SysView {
    int natoms;
    int nstruct;
    const double* pos;     // [3N]
    const int* elem;       // [N]
    const int* sid;        // [N] structure of each atom
    const StructT* st;     // [nstruct]
    // neighbour list, fixed capacity per atom, entries sorted by (j, image)
    int cap;
    int stage_cap;         // LDS staging slots of the mixed kernels (<= cap)
    int* nb_n;             // [N]
    int* nb_j;             // [N*cap]
    int* nb_img;           // [N*cap]
    int* nb_rev;           // [N*cap] entry (j*cap + k) holding (i, -image)
    int* flags;            // [0] overflow bits, [1] max list count, [2] max staged count
    // cell list
    int ncells;
    int* atom_cell;        // [N]
    int* cell_count;       // [ncells]
    int* cell_rank;        // [N] atom's place among its cell's atoms
    int* cell_start;       // [ncells+1]
    int* cell_atoms;       // [N]
    double* cell_pos;      // [3N] positions in cell order
    // when non-null, the list kernels run only if *gate != 0 (device-side
    // rebuild decision inside captured MD steps)
    const int* gate;
    // caller-provided neighbour lists (the LAMMPS path): distance vector
    // r_i - r_j of every entry [N*cap*3]; null when lists are built here
    const double* nb_dr;
    // force entries (natoms, or local + ghost atoms with external lists)
    int nforce;
};

// Distance vector r_i - r_j of entry e (= i*cap + k) of atom i: from the
// positions and the image shift, or as given with an external list.
__device__ inline void entry_vector(const SysView& s, int i, size_t e, const StructT& S, double* dr)
{
    if (s.nb_dr) {
        dr[0] = s.nb_dr[3 * e];
        dr[1] = s.nb_dr[3 * e + 1];
        dr[2] = s.nb_dr[3 * e + 2];
        return;
    }
    int b0, b1, b2;
    unpack_image(s.nb_img[e], b0, b1, b2);
    pair_vector(s.pos + 3 * (size_t)i, s.pos + 3 * (size_t)s.nb_j[e], S.box, b0, b1, b2, dr);
}

void fill_int(int* p, int v, size_t n, hipStream_t stream, const int* gate = nullptr);

// Neighbour list (Structure::calculateNeighborList, Structure.cpp:287-398):
// cell assignment and per-atom build (brute force or cell list, same set
// and order).
// clear_flags: reset the overflow flags first (MD keeps them across rebuilds).
void launch_neighbors(const SysView& s, double rc, bool any_cell, hipStream_t stream, bool clear_flags = true);
// Reverse entries: for (i -> j, image) the slot of (j -> i, -image).
void launch_reverse(const SysView& s, hipStream_t stream);

// FP64 symmetry functions; with deriv also dG/dr in the reference layout.
struct SF64Out {
    double* G;          // [N*ldG]
    double* dGself;     // [N*ldG*3]
    double* dGnbr;      // [N*cap*S*3]
    int ldG, S;
};
void launch_sf64(const ModelT* m, const ModelT& hm, const SysView& s, const SF64Out& o,
                 bool deriv, int max_nb, hipStream_t stream);
// The compact-function groups (types 20, 21, 22) into the same outputs.
void launch_sfcomp64(const ModelT* m, const ModelT& hm, const SysView& s, const SF64Out& o,
                     bool deriv, int max_nb, hipStream_t stream);
// The compact groups in the mixed form (FP32 terms): values (scaled G and its
// FP32 copy) or, with forces, the terms' derivatives contracted with the
// weights fw = dE/dG times the scaling factor and added to F.  Stages
// s.stage_cap neighbours (a longer list sets flags[0] bit 4).
void launch_sfcomp32(const ModelT* m, const ModelT& hm, const SysView& s, double* G, float* Gin, int ldG, int ldK,
                     const float* fw, double* F, bool push, bool forces, hipStream_t stream);
// Tangent Gdot = sum_a dG/dx_a . u_a of the compact groups (FP32 terms,
// FP64 sums; the scaling factor included), written to Gdot [N][ldG].
void launch_tangentcomp32(const ModelT* m, const ModelT& hm, const SysView& s, const double* u, double* Gdot,
                          int ldG, hipStream_t stream);

// FP64 atomic NNs over per-element atom lists; dE/dG by the adjoint.
struct NNIO {
    const double* G;       // [N*ldG] scaled SFs
    double* energy;        // [N]
    double* dEdG;          // [N*ldG] (may be null)
    int ldG;
};
void launch_nn64(const ModelT* m, const ModelT& hm, const double* w, const int* elem_atoms,
                 const int* elem_count_host, const int* elem_off_host, const NNIO& io,
                 hipStream_t stream);

// Per-structure energy sums (fixed-order tree reduction).
void launch_energy(const double* atomE, const int* atom_off, int nstruct, double* E,
                   hipStream_t stream);

// FP64 push-form forces from stored dG/dr scattered onto every neighbour
// entry (InterfaceLammps::getForces, InterfaceLammps.cpp:679-740): works
// with external lists whose ghost atoms have no lists of their own.
void launch_force64_push(const ModelT* m, const SysView& s, const double* dEdG, int ldG, const double* dGself,
                         const double* dGnbr, int S, double* f, hipStream_t stream);

// FP64 force contraction from stored dG/dr.
void launch_force64(const ModelT* m, const SysView& s, const double* dEdG, int ldG,
                    const double* dGself, const double* dGnbr, int S, int push,
                    double* f, hipStream_t stream);

// Plain-form forces without stored dG/dr: every SF term recomputed in FP64
// and contracted with w = dE/dG times fac (the member's derivative scale per
// element and SF, [kMaxElem*kMaxSF]); w [N*ldG] is scratch.  F [3*nforce] is
// overwritten.  The compact groups go through their kernel's FP64 force mode.
// mm supplies the power-path tables (MixedModel::powNB, d_mem, d_gpow64).
struct MixedModel;
void launch_force64_fused(const ModelT* m, const ModelT& hm, const MixedModel& mm, const SysView& s,
                          const double* dEdG, const double* fac, double* w, int ldG, double* F, bool push,
                          hipStream_t stream);
void launch_sfcomp64_force(const ModelT* m, const ModelT& hm, const SysView& s, const double* w, int ldG, double* F,
                           bool push, hipStream_t stream);
// Scaled SF values of the Gaussian groups in FP64 (no derivatives) on the
// fused kernels' staging and compacted triplet walk.
void launch_values64(const ModelT* m, const ModelT& hm, const MixedModel& mm, const SysView& s, double* G, int ldG,
                     hipStream_t stream);
// LDS bytes of the fused FP64 force kernel at a list capacity (<= 65536 to run).
size_t fused64_lds(int cap);

// Extrapolation warnings (Element::updateSymmetryFunctionStatistics): every
// (atom, SF) whose unscaled G lies outside [Gmin, Gmax] by more than 1000 eps
// increments *count; the first cap are written (atom, SF, unscaled value).
void launch_extrapolation(const ModelT* m, int natoms, const int* elem, const double* G, int ldG,
                          unsigned long long* count, long long cap, int* atom, int* sf, double* val,
                          hipStream_t stream);
// Low-neighbour atoms (Mode.cpp:1667-1679): fewer than nmin[e] entries with
// |dr| <= rmin[e]; the first cap are written (atom, count).
void launch_low_neighbors(const ModelT* m, const SysView& s, const double* rmin, const int* nmin,
                          unsigned long long* count, long long cap, int* atom, int* num, hipStream_t stream);

// Kalman filter pieces (KalmanFilter::update): P = diag I; X = P H and
// A = H^T X; then K = X Ai, P = (P - K X^T) scale + q I (symmetric, from the
// lower triangle) and dw = K xi.  P [n][n], H/X/K [n][m], A/Ai [m][m].
void kf_init(double* P, int n, double diag, hipStream_t stream);
void kf_xa(const double* P, const double* H, int n, int m, double* X, double* A, hipStream_t stream);
void kf_finish(double* P, const double* X, const double* Ai, const double* xi, int n, int m, double scale, double q,
               double* K, double* dw, hipStream_t stream);

// Algorithmic work counters (pairs, triplets, member evaluations).
void launch_work_counts(const ModelT* m, const SysView& s, unsigned long long* out5,
                        hipStream_t stream);

}  // namespace nnpx
