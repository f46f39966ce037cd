// Compact symmetry-function groups (types 20-25).
//
// Computes, per central atom, what SymGrpCompRad::calculate
// (SymGrpCompRad.cpp:129-199), SymGrpCompAngn::calculate
// (SymGrpCompAngn.cpp:128-390) and SymGrpCompAngw::calculate
// (SymGrpCompAngw.cpp:117-341) compute, for the groups of types 20, 21 and 22
// and of their element-weighted forms 23, 24 and 25 (SymGrpCompRadWeighted,
// SymGrpCompAng{n,w}Weighted); the Gaussian groups are the other kernels'.
// Three modes share the term code:
//  * plain (FP64 terms): scaled G, the atom's own dG/dr and every neighbour
//    entry's dG/dr in the compact slot order (the stored-derivative form);
//  * values (FP32 terms, FP64 sums): scaled G and its FP32 copy, the mixed
//    form's network input;
//  * force (FP32 terms): the terms' derivatives contracted with dE/dG on the
//    fly (members summed per triplet in FP32, neighbours accumulated in LDS
//    FP64) and added to the forces, the mixed form's fused backward;
//  * tangent (FP32 terms, FP64 sums): Gdot = sum_a dG/dx_a . u_a for a field
//    u, the mixed training gradient's tangent (Training::collectDGdxia
//    contracted with u, Training.cpp:3586-3625).
//
// Mapping: one 256-thread workgroup per atom with the neighbour geometry
// staged in LDS.  Radial groups: threads over neighbours, members in
// sequence.  Angular groups: threads over the unordered neighbour pairs of
// the group's element pair.  Plain and values modes take the members in
// chunks of CH, each chunk re-walking the pairs (plain: dG/dr contributions
// accumulated by LDS FP64 atomics [CH][3][cap] and written once); compact
// groups are large (tens of members with disjoint radial and angular
// windows), so a chunk keeps the accumulators bounded whatever the group size.
#include <type_traits>

#include "kernels.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kCThreads = 256;
constexpr int kCWaves = kCThreads / kWave;
constexpr int kRedSlots = 16;  // per-wave reduction slots: members of a chunk (and 4 radial sums)

struct CLds {
    double* dr;   // [3][cap]
    double* d;    // [cap]
    double* red;  // [kCWaves][kRedSlots]
    double* acc;  // [CH][3][cap] (neighbour index fastest)
    int* el;      // [cap]
    int* la;      // [cap] neighbours of element e1 inside the group range
    int* lb;      // [cap] neighbours of element e2
};

__device__ inline void block_sum_slot(double v, double* red, int slot)
{
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) red[(threadIdx.x / kWave) * kRedSlots + slot] = v;
}

__device__ inline double red_total(const double* red, int slot)
{
    double v = 0.0;
    for (int w = 0; w < kCWaves; ++w) v += red[w * kRedSlots + slot];
    return v;
}

// Ordered compaction of the neighbours of element `elem` (any element if
// negative) with rmin < |dr| < rmax (one wave, so the list keeps neighbour
// order).
__device__ inline int build_list(const CLds& L, int n, int elem, double rmin, double rmax, int* out)
{
    __shared__ int cnt_sh;
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        int cnt = 0;
        for (int b = 0; b < n; b += kWave) {
            const int k = b + lane;
            const bool ok = k < n && (elem < 0 || L.el[k] == elem) && L.d[k] < rmax && L.d[k] > rmin;
            const int incl = wave_incl_scan(ok ? 1 : 0);
            if (ok) out[cnt + incl - 1] = k;
            cnt += lane_value(incl, kWave - 1);
        }
        if (lane == 0) cnt_sh = cnt;
    }
    __syncthreads();
    const int c = cnt_sh;
    __syncthreads();
    return c;
}

enum CompMode { CM_PLAIN = 0, CM_VALUES = 1, CM_FORCE = 2, CM_TANGENT = 3 };

// Member windows and centres in the term precision.
template <typename T>
__device__ inline void radial_fdf(const MemberT& m, T r, T& f, T& df)
{
    compact_fdf(m.core, m.asym, (T)m.cen, (T)m.scl, r, f, df);
}

// The r_ij factor of member m (zero outside its window), as the reference's
// per-j precomputation radij[l] holds it.
template <typename T>
__device__ inline T radial_ij(const MemberT& m, T rij, T& dradij)
{
    T f = T(0);
    dradij = T(0);
    if (rij > (T)m.rl && rij < (T)m.rcm) radial_fdf(m, rij, f, dradij);
    return f;
}

// The reference's break condition of member m for one triplet.
template <typename T>
__device__ inline bool member_skipped(const MemberT& m, T radij, T rik, T rjk, bool narrow, T th)
{
    return radij == T(0) || rik <= (T)m.rl || rik >= (T)m.rcm || (narrow && (rjk <= (T)m.rl || rjk >= (T)m.rcm)) ||
           th <= (T)m.al || th >= (T)m.ar;
}

// Outputs of the mixed modes.
struct CompMix {
    double* G;        // [N*ldG] scaled values (values mode)
    float* Gin;       // [N*ldK] their FP32 copy, the network input
    int ldG, ldK;
    const float* fw;  // [N*ldG] dE/dG times the scaling factor (force mode)
    double* F;        // forces (force mode)
    int push;         // keep contributions to the atom's own images
    const double* u;  // [3N] field of the tangent mode (its output goes to G)
    const double* fw64;  // [N*ldG] FP64 weights of the FP64 force mode (plain form)
};

template <typename T, int MODE, int CH>
__global__ __launch_bounds__(kCThreads) void k_sfcomp(const ModelT* __restrict__ M, SysView s, SF64Out o,
                                                       CompMix c, int deriv, int capL)
{
    constexpr bool PLAIN = MODE == CM_PLAIN, FORCE = MODE == CM_FORCE, TANGENT = MODE == CM_TANGENT;
    constexpr bool DER = PLAIN || TANGENT;  // derivatives times the scaling factor
    static_assert(CH <= kRedSlots, "a member chunk needs one reduction slot per member");
    // [3][cap] sets in LDS: plain dG/dr accumulators, force sums, tangent field
    constexpr int ACC = PLAIN ? CH : (FORCE || TANGENT) ? 1 : 0;
    extern __shared__ double smem[];
    CLds L;
    L.dr = smem;
    L.d = L.dr + 3 * capL;
    L.red = L.d + capL;
    L.acc = L.red + kCWaves * kRedSlots;
    L.el = reinterpret_cast<int*>(L.acc + (size_t)ACC * 3 * capL);
    L.la = L.el + capL;
    L.lb = L.la + capL;

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const StructT& S = s.st[s.sid[i]];
    const int n = min(s.nb_n[i], capL);
    if (!PLAIN && tid == 0 && s.nb_n[i] > capL) atomicOr(&s.flags[0], 4);  // the host repeats at full capacity
    const size_t ebase = (size_t)i * s.cap;

    for (int k = tid; k < n; k += kCThreads) {
        const int j = s.nb_j[ebase + k];
        double dr[3];
        entry_vector(s, i, ebase + k, S, dr);
        L.dr[k] = dr[0];
        L.dr[capL + k] = dr[1];
        L.dr[2 * capL + k] = dr[2];
        L.d[k] = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        L.el[k] = s.elem[j];
        if constexpr (FORCE) {
            L.acc[k] = 0.0;
            L.acc[capL + k] = 0.0;
            L.acc[2 * capL + k] = 0.0;
        }
        if constexpr (TANGENT) {  // u of the neighbour minus u of the atom
            for (int cc = 0; cc < 3; ++cc) L.acc[cc * capL + k] = c.u[3 * (size_t)j + cc] - c.u[3 * (size_t)i + cc];
        }
    }
    __syncthreads();

    double* Gi = PLAIN ? o.G + (size_t)i * o.ldG : c.G + (size_t)i * c.ldG;
    double* dGi = o.dGself + (size_t)i * o.ldG * 3;
    double* dGn = o.dGnbr + ebase * o.S * 3;
    float* Ii = c.Gin + (size_t)i * c.ldK;
    // force weights: FP32 in the mixed form, FP64 in the plain form's fused force
    auto weight = [&](int sf) -> T {
        if constexpr (std::is_same<T, double>::value) return c.fw64[(size_t)i * c.ldG + sf];
        else return (T)c.fw[(size_t)i * c.ldG + sf];
    };
    double fself[3] = {0.0, 0.0, 0.0};  // force mode: this thread's part of the atom's own force

    // Writes member mb's value v (values and plain modes) or tangent.
    auto put = [&](const MemberT& mb, double v) {
        if (TANGENT) {
            Gi[mb.sf] = v;
            return;
        }
        const double gs = scale_sf(M->scale[E.sf0 + mb.sf], v);
        Gi[mb.sf] = gs;
        if (MODE == CM_VALUES) Ii[mb.sf] = (float)gs;
    };

    // ---------------------------------------------------------- radial groups
    for (int gi = 0; gi < E.nrad; ++gi) {
        const GroupT& G = M->grp[E.grp0 + gi];
        if (G.type != 20 && G.type != 23) continue;
        // weighted (SymGrpCompRadWeighted.cpp:120-200): every neighbour
        // element, each term times its atomic number
        const bool wt = G.type == 23;
        if constexpr (FORCE) {
            // threads own neighbours: sum_q w_q dG_q/dr over the members
            for (int k = tid; k < n; k += kCThreads) {
                if (!wt && L.el[k] != G.e1) continue;
                const T rij = (T)L.d[k];
                if (!(rij < (T)G.rc && rij > (T)G.rmin)) continue;
                const T Z = wt ? (T)M->Zel[L.el[k]] : T(1);
                T P = T(0);
                for (int q = 0; q < G.nmem; ++q) {
                    const MemberT& mb = M->mem[G.mem0 + q];
                    if (!(rij > (T)mb.rl && rij < (T)mb.rcm)) continue;
                    T rad, drad;
                    radial_fdf(mb, rij, rad, drad);
                    P += weight(mb.sf) * drad;
                }
                P *= Z / rij;
                for (int cc = 0; cc < 3; ++cc) {
                    const double v = (double)(P * (T)L.dr[cc * capL + k]);
                    fself[cc] -= v;
                    L.acc[cc * capL + k] += v;
                }
            }
            continue;
        }
        for (int q = 0; q < G.nmem; ++q) {
            const MemberT& mb = M->mem[G.mem0 + q];
            double gs = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (int k = tid; k < n; k += kCThreads) {
                if (!wt && L.el[k] != G.e1) continue;
                double px = 0.0, py = 0.0, pz = 0.0;
                const T rij = (T)L.d[k];
                if (rij < (T)G.rc && rij > (T)G.rmin && rij > (T)mb.rl && rij < (T)mb.rcm) {
                    T rad, drad;
                    radial_fdf(mb, rij, rad, drad);
                    const T Z = (T)M->Zel[L.el[k]];
                    if (!TANGENT) {
                        if (wt) gs += (double)(Z * rad);
                        else gs += (double)rad;
                    }
                    if (DER && drad != T(0)) {
                        if (wt) drad *= Z * mb.sfac / rij;
                        else drad *= mb.sfac / rij;
                        px = drad * L.dr[k];
                        py = drad * L.dr[capL + k];
                        pz = drad * L.dr[2 * capL + k];
                        sx += px; sy += py; sz += pz;
                        // tangent: entry k's dG/dr is -p, times u_k - u_i
                        if (TANGENT) gs -= px * L.acc[k] + py * L.acc[capL + k] + pz * L.acc[2 * capL + k];
                    }
                }
                if (PLAIN && deriv) {
                    double* d = dGn + ((size_t)k * o.S + (wt ? mb.slotE[L.el[k]] : mb.slot1)) * 3;
                    d[0] = -px; d[1] = -py; d[2] = -pz;
                }
            }
            block_sum_slot(gs, L.red, 0);
            if (PLAIN) {
                block_sum_slot(sx, L.red, 1);
                block_sum_slot(sy, L.red, 2);
                block_sum_slot(sz, L.red, 3);
            }
            __syncthreads();
            if (tid < 4) {
                const double v = red_total(L.red, tid);
                if (tid == 0) put(mb, v);
                else if (PLAIN && deriv) dGi[mb.sf * 3 + tid - 1] = v;
            }
            __syncthreads();
        }
    }

    // --------------------------------------------------------- angular groups
    for (int gi = E.nrad; gi < E.ngrp; ++gi) {
        const GroupT& G = M->grp[E.grp0 + gi];
        if (G.type < 21 || G.type == 23) continue;
        const bool narrow = G.type == 21 || G.type == 24;
        // weighted (types 24, 25; SymGrpCompAng{n,w}Weighted.cpp): all
        // neighbour pairs, the angular factor times Z_j Z_k
        const bool wt = G.type >= 24;
        const bool same = G.e1 == G.e2;
        const int nA = build_list(L, n, G.e1, G.rmin, G.rc, L.la);
        const int nB = same ? nA : build_list(L, n, G.e2, G.rmin, G.rc, L.lb);
        const int* lb = same ? L.la : L.lb;
        int D = 0, full = 0, total;
        if (same) {
            D = (nA - 1) / 2;
            full = nA * D;
            total = full + ((nA % 2 == 0) ? nA / 2 : 0);
        } else {
            total = nA * nB;
        }
        const T r2min = (T)(G.rmin * G.rmin), r2max = (T)G.rc2;
        const MemberT* mem = M->mem + G.mem0;
        // force mode: one pass over all members; plain/values: chunks of CH
        const int step = FORCE ? G.nmem : CH;

        for (int q0 = 0; q0 < G.nmem; q0 += step) {
            const int nq = min(step, G.nmem - q0);
            if (PLAIN && deriv)
                for (int x = tid; x < nq * 3 * capL; x += kCThreads) L.acc[x] = 0.0;
            __syncthreads();
            double gacc[FORCE ? 1 : CH];
#pragma unroll
            for (int q = 0; q < (FORCE ? 1 : CH); ++q) gacc[q] = 0.0;

            for (int t = tid; t < total; t += kCThreads) {
                int a, b;
                if (same) {
                    if (t < full) {
                        a = t % nA;
                        b = (a + t / nA + 1) % nA;
                    } else {
                        a = t - full;
                        b = a + nA / 2;
                    }
                } else {
                    a = t % nA;
                    b = t / nA;
                }
                int j = L.la[a], k = lb[b];
                if (j > k) { const int x = j; j = k; k = x; }
                const T dr1[3] = {(T)L.dr[j], (T)L.dr[capL + j], (T)L.dr[2 * capL + j]};
                const T dr2[3] = {(T)L.dr[k], (T)L.dr[capL + k], (T)L.dr[2 * capL + k]};
                const T dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
                T rjk = T(0);
                if (narrow) {
                    rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
                    if (rjk >= r2max || rjk <= r2min) continue;
                    rjk = sqrt(rjk);
                }
                const T rij = (T)L.d[j], rik = (T)L.d[k];
                const T rinvijik = T(1) / (rij * rik);
                const T costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
                T th, dacos;
                if constexpr (PLAIN || std::is_same<T, double>::value) {
                    // FP64 (the plain form's values and forces): the
                    // reference's acos (SymGrpCompAngn.cpp)
                    // the angular polynomial vanishes at 0 and 180 degrees
                    if (costijk <= T(-1) || costijk >= T(1)) continue;
                    // (FP32 below: the angle from the sine and cosine instead)
                    th = acos(costijk);
                    dacos = T(-1) / sqrt(T(1) - costijk * costijk);
                } else {
                    // FP32: the angle from the sine (cross product) and cosine,
                    // accurate at both ends where acos(cos) is not
                    const T cx = dr1[1] * dr2[2] - dr1[2] * dr2[1];
                    const T cy = dr1[2] * dr2[0] - dr1[0] * dr2[2];
                    const T cz = dr1[0] * dr2[1] - dr1[1] * dr2[0];
                    const T sint = sqrt(cx * cx + cy * cy + cz * cz) * rinvijik;
                    if (!(sint > T(0))) continue;
                    th = atan2f(sint, costijk);
                    dacos = T(-1) / sint;
                }
                const T rinvij = rinvijik * rik;
                const T rinvik = rinvijik * rij;
                const T rinvjk = narrow ? T(1) / rjk : T(0);
                const T phiijik0 = rinvij * (rinvik - rinvij * costijk);
                const T phiikij0 = rinvik * (rinvij - rinvik * costijk);
                T P1 = T(0), P2 = T(0), P3 = T(0);  // force mode: sum_q w_q p_q
                // tangent mode: dr1.w_j, dr2.w_k, dr3.(w_j - w_k) with w = u_nbr - u_i
                T a1 = T(0), a2 = T(0), a3 = T(0);
                if constexpr (TANGENT) {
                    for (int cc = 0; cc < 3; ++cc) {
                        const T wj = (T)L.acc[cc * capL + j], wk = (T)L.acc[cc * capL + k];
                        a1 += dr1[cc] * wj;
                        a2 += dr2[cc] * wk;
                        a3 += dr3[cc] * (wj - wk);
                    }
                }
                auto member = [&](int q) {
                    const int l = q0 + q;
                    const MemberT& mb = mem[l];
                    T dradij;
                    const T radij = radial_ij(mb, rij, dradij);
                    if (member_skipped(mb, radij, rik, rjk, narrow, th)) return;
                    T radik, dradik, radjk = T(1), dradjk = T(0), ang, dang;
                    radial_fdf(mb, rik, radik, dradik);
                    if (narrow) {
                        // r_jk factor: this member's, or the one the
                        // reference carries over from earlier members
                        int src = l;
                        if (mb.walk) {
                            while (!mem[src].calcComp) {
                                T dp;
                                const T rp = radial_ij(mem[src - 1], rij, dp);
                                if (member_skipped(mem[src - 1], rp, rik, rjk, true, th)) break;
                                --src;
                            }
                        }
                        radial_fdf(mem[src], rjk, radjk, dradjk);
                    }
                    compact_fdf(mb.core, 0, (T)mb.acen, (T)mb.ascl, th, ang, dang);
                    T z = T(1);
                    if (wt) {
                        z = (T)(M->Zel[L.el[j]] * M->Zel[L.el[k]]);
                        ang *= z;
                    }
                    T rad = narrow ? radij * radik * radjk : radij * radik;
                    if constexpr (!FORCE && !TANGENT) {
                        gacc[q] += (double)(rad * ang);
                        if (!PLAIN || !deriv) return;
                    }
                    dang *= dacos;
                    const T phiijik = phiijik0 * dang;
                    const T phiikij = phiikij0 * dang;
                    const T psiijik = rinvijik * dang;
                    T chiij, chiik, chijk = T(0);
                    if (narrow) {
                        chiij = rinvij * dradij * radik * radjk;
                        chiik = rinvik * radij * dradik * radjk;
                        chijk = -rinvjk * radij * radik * dradjk;
                    } else {
                        // the weighted group multiplies in the other order
                        // (SymGrpCompAngwWeighted.cpp:257)
                        chiij = wt ? rinvij * dradij * radik : rinvij * radik * dradij;
                        chiik = rinvik * radij * dradik;
                    }
                    // plain: times the scaling factor here; force mode: it
                    // is part of the weight w_q
                    const T sf = DER ? (T)mb.sfac : T(1);
                    T p1, p2, p3;
                    if (dang != T(0) && rad != T(0)) {
                        if (wt) rad *= sf * z;
                        else rad *= sf;
                        ang *= sf;
                        p1 = rad * phiijik + ang * chiij;
                        p2 = rad * phiikij + ang * chiik;
                        p3 = narrow ? rad * psiijik + ang * chijk : rad * psiijik;
                    } else if (ang != T(0)) {
                        ang *= sf;
                        p1 = ang * chiij;
                        p2 = ang * chiik;
                        p3 = narrow ? ang * chijk : T(0);
                    } else {
                        return;
                    }
                    if constexpr (FORCE) {
                        const T w = weight(mb.sf);
                        P1 += w * p1;
                        P2 += w * p2;
                        P3 += w * p3;
                    } else if constexpr (TANGENT) {
                        // entries j, k hold -(p1 dr1 + p3 dr3), -(p2 dr2 - p3 dr3)
                        gacc[q] -= (double)(p1 * a1 + p2 * a2 + p3 * a3);
                    } else {
                        for (int cc = 0; cc < 3; ++cc) {
                            const double p1c = p1 * dr1[cc], p2c = p2 * dr2[cc], p3c = p3 * dr3[cc];
                            double* acc = L.acc + (size_t)(q * 3 + cc) * capL;
                            lds_add(acc + j, -(p1c + p3c));
                            lds_add(acc + k, -(p2c - p3c));
                        }
                    }
                };
                if constexpr (FORCE) {
                    for (int q = 0; q < nq; ++q) member(q);
                } else {
#pragma unroll
                    for (int q = 0; q < CH; ++q)
                        if (q < nq) member(q);
                }
                if constexpr (FORCE) {
                    // dG/dr contracted with dE/dG: entry j gets -(P1 dr1 + P3 dr3),
                    // entry k -(P2 dr2 - P3 dr3), the atom minus their sum;
                    // forces are minus those
                    for (int cc = 0; cc < 3; ++cc) {
                        const T gj = -(P1 * dr1[cc] + P3 * dr3[cc]);
                        const T gk = -(P2 * dr2[cc] - P3 * dr3[cc]);
                        lds_add(L.acc + cc * capL + j, -(double)gj);
                        lds_add(L.acc + cc * capL + k, -(double)gk);
                        fself[cc] += (double)(gj + gk);
                    }
                }
            }
            if constexpr (!FORCE) {
                // G of the chunk's members.
#pragma unroll
                for (int q = 0; q < CH; ++q)
                    if (q < nq) block_sum_slot(gacc[q], L.red, q);
                __syncthreads();
                if (tid < nq) put(mem[q0 + tid], red_total(L.red, tid));
            }
            if (PLAIN && deriv) {
                // Neighbour entries of the chunk's slots (zero where no triplet).
                for (int x = tid; x < n * nq; x += kCThreads) {
                    const int k = x / nq, q = x % nq;
                    const int ek = L.el[k];
                    if (!wt && ek != G.e1 && ek != G.e2) continue;
                    const MemberT& mb = mem[q0 + q];
                    const int slot = wt ? mb.slotE[ek] : ek == G.e1 ? mb.slot1 : mb.slot2;
                    double* d = dGn + ((size_t)k * o.S + slot) * 3;
                    const double* a = L.acc + (size_t)q * 3 * capL + k;
                    d[0] = a[0]; d[1] = a[capL]; d[2] = a[2 * capL];
                }
                // The atom's own dG/dr: minus the sum over its neighbour entries.
                if (tid < nq * 3) {
                    const int q = tid / 3, cc = tid % 3;
                    double v = 0.0;
                    for (int k = 0; k < n; ++k) v += L.acc[(size_t)(q * 3 + cc) * capL + k];
                    dGi[mem[q0 + q].sf * 3 + cc] = -v;
                }
            }
            __syncthreads();
        }
    }

    if constexpr (FORCE) {
        // the atom's own force, then its neighbours' (the pull form drops
        // the atom's own periodic images, Mode.cpp:1957)
        for (int cc = 0; cc < 3; ++cc) block_sum_slot(fself[cc], L.red, cc);
        __syncthreads();
        if (tid < 3) atomicAdd(c.F + 3 * (size_t)i + tid, red_total(L.red, tid));
        for (int k = tid; k < n; k += kCThreads) {
            const int j = s.nb_j[ebase + k];
            if (j == i && !c.push) continue;
            atomicAdd(c.F + 3 * (size_t)j, L.acc[k]);
            atomicAdd(c.F + 3 * (size_t)j + 1, L.acc[capL + k]);
            atomicAdd(c.F + 3 * (size_t)j + 2, L.acc[2 * capL + k]);
        }
    }
}

size_t comp_lds(int capL, int acc)
{
    return sizeof(double) * ((size_t)capL * 4 + kCWaves * kRedSlots + (size_t)acc * 3 * capL) +
           sizeof(int) * (size_t)capL * 3;
}

}  // namespace

void launch_sfcomp64(const ModelT* m, const ModelT& hm, const SysView& s, const SF64Out& o, bool deriv,
                     int max_nb, hipStream_t stream)
{
    if (s.natoms == 0 || !hm.has_comp) return;
    const int capL = ((std::max(max_nb, 1) + 31) / 32) * 32;
    int ch = 8;
    while (ch > 1 && comp_lds(capL, ch) > 65536) ch /= 2;
    const size_t lds = comp_lds(capL, ch);
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the FP64 compact SF kernel's LDS staging"};
    const dim3 grid(s.natoms), block(kCThreads);
    const int d = deriv ? 1 : 0;
    const CompMix c{};
    switch (ch) {
    case 8: hipLaunchKernelGGL((k_sfcomp<double, CM_PLAIN, 8>), grid, block, lds, stream, m, s, o, c, d, capL); break;
    case 4: hipLaunchKernelGGL((k_sfcomp<double, CM_PLAIN, 4>), grid, block, lds, stream, m, s, o, c, d, capL); break;
    case 2: hipLaunchKernelGGL((k_sfcomp<double, CM_PLAIN, 2>), grid, block, lds, stream, m, s, o, c, d, capL); break;
    default: hipLaunchKernelGGL((k_sfcomp<double, CM_PLAIN, 1>), grid, block, lds, stream, m, s, o, c, d, capL); break;
    }
}

void launch_sfcomp64_force(const ModelT* m, const ModelT& hm, const SysView& s, const double* w, int ldG, double* F,
                           bool push, hipStream_t stream)
{
    if (s.natoms == 0 || !hm.has_comp) return;
    const int capL = ((std::max(s.cap, 1) + 31) / 32) * 32;
    const size_t lds = comp_lds(capL, 1);
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the FP64 compact force kernel's LDS staging"};
    const SF64Out o{};
    CompMix c{};
    c.ldG = ldG;
    c.F = F;
    c.push = push ? 1 : 0;
    c.fw64 = w;
    hipLaunchKernelGGL((k_sfcomp<double, CM_FORCE, 1>), dim3(s.natoms), dim3(kCThreads), lds, stream, m, s, o, c, 0,
                       capL);
}

void launch_tangentcomp32(const ModelT* m, const ModelT& hm, const SysView& s, const double* u, double* Gdot,
                          int ldG, hipStream_t stream)
{
    if (s.natoms == 0 || !hm.has_comp) return;
    const int capL = s.stage_cap;
    const size_t lds = comp_lds(capL, 1);
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed compact SF kernel's LDS staging"};
    const SF64Out o{};
    const CompMix c{Gdot, nullptr, ldG, 0, nullptr, nullptr, 0, u};
    hipLaunchKernelGGL((k_sfcomp<float, CM_TANGENT, 16>), dim3(s.natoms), dim3(kCThreads), lds, stream, m, s, o, c, 0,
                       capL);
}

void launch_sfcomp32(const ModelT* m, const ModelT& hm, const SysView& s, double* G, float* Gin, int ldG, int ldK,
                     const float* fw, double* F, bool push, bool forces, hipStream_t stream)
{
    if (s.natoms == 0 || !hm.has_comp) return;
    const int capL = s.stage_cap;
    const size_t lds = comp_lds(capL, forces ? 1 : 0);
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed compact SF kernel's LDS staging"};
    const dim3 grid(s.natoms), block(kCThreads);
    const SF64Out o{};
    const CompMix c{G, Gin, ldG, ldK, fw, F, push ? 1 : 0, nullptr};
    if (forces) hipLaunchKernelGGL((k_sfcomp<float, CM_FORCE, 1>), grid, block, lds, stream, m, s, o, c, 0, capL);
    else hipLaunchKernelGGL((k_sfcomp<float, CM_VALUES, 16>), grid, block, lds, stream, m, s, o, c, 0, capL);
}

}  // namespace nnpx
