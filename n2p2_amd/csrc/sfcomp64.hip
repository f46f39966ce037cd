// FP64 compact symmetry-function groups with stored dG/dr (the plain form).
//
// Computes, per central atom, what SymGrpCompRad::calculate
// (SymGrpCompRad.cpp:129-199), SymGrpCompAngn::calculate
// (SymGrpCompAngn.cpp:128-390) and SymGrpCompAngw::calculate
// (SymGrpCompAngw.cpp:117-341) compute: scaled G, the atom's own dG/dr and
// every neighbour entry's dG/dr in the compact slot order, for the groups of
// types 20, 21 and 22, and of their element-weighted forms 23, 24 and 25
// (SymGrpCompRadWeighted, SymGrpCompAng{n,w}Weighted); the Gaussian groups
// are the FP64 SF kernel's.
//
// Mapping: one 256-thread workgroup per atom with the neighbour geometry
// staged in LDS.  Radial groups: threads over neighbours, members in
// sequence.  Angular groups: threads over the unordered neighbour pairs of
// the group's element pair; the members are taken in chunks of CH, each
// chunk re-walking the pairs with its dG/dr contributions accumulated by LDS
// FP64 atomics ([CH][3][cap]) and written once.  Compact groups are large
// (tens of members with disjoint radial and angular windows), so a chunk
// keeps the LDS accumulator bounded whatever the group size.
#include "kernels.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kCThreads = 256;
constexpr int kCWaves = kCThreads / kWave;
constexpr int kRedSlots = 8;

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
            cnt += __shfl(incl, kWave - 1, kWave);
        }
        if (lane == 0) cnt_sh = cnt;
    }
    __syncthreads();
    const int c = cnt_sh;
    __syncthreads();
    return c;
}

__device__ inline void radial_fdf(const MemberT& m, double r, double& f, double& df)
{
    compact_fdf(m.core, m.asym, m.cen, m.scl, r, f, df);
}

// The r_ij factor of member m (zero outside its window), as the reference's
// per-j precomputation radij[l] holds it.
__device__ inline double radial_ij(const MemberT& m, double rij, double& dradij)
{
    double f = 0.0;
    dradij = 0.0;
    if (rij > m.rl && rij < m.rcm) radial_fdf(m, rij, f, dradij);
    return f;
}

// The reference's break condition of member m for one triplet.
__device__ inline bool member_skipped(const MemberT& m, double radij, double rik, double rjk, bool narrow, double th)
{
    return radij == 0.0 || rik <= m.rl || rik >= m.rcm || (narrow && (rjk <= m.rl || rjk >= m.rcm)) ||
           th <= m.al || th >= m.ar;
}

template <int CH>
__global__ __launch_bounds__(kCThreads) void k_sfcomp64(const ModelT* __restrict__ M, SysView s, SF64Out o,
                                                         int deriv, int capL)
{
    extern __shared__ double smem[];
    CLds L;
    L.dr = smem;
    L.d = L.dr + 3 * capL;
    L.red = L.d + capL;
    L.acc = L.red + kCWaves * kRedSlots;
    L.el = reinterpret_cast<int*>(L.acc + (size_t)CH * 3 * capL);
    L.la = L.el + capL;
    L.lb = L.la + capL;

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const StructT& S = s.st[s.sid[i]];
    const int n = min(s.nb_n[i], capL);
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
    }
    __syncthreads();

    double* Gi = o.G + (size_t)i * o.ldG;
    double* dGi = o.dGself + (size_t)i * o.ldG * 3;
    double* dGn = o.dGnbr + ebase * o.S * 3;

    // ---------------------------------------------------------- radial groups
    for (int gi = 0; gi < E.nrad; ++gi) {
        const GroupT& G = M->grp[E.grp0 + gi];
        if (G.type != 20 && G.type != 23) continue;
        // weighted (SymGrpCompRadWeighted.cpp:120-200): every neighbour
        // element, each term times its atomic number
        const bool wt = G.type == 23;
        for (int q = 0; q < G.nmem; ++q) {
            const MemberT& mb = M->mem[G.mem0 + q];
            double gs = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (int k = tid; k < n; k += kCThreads) {
                if (!wt && L.el[k] != G.e1) continue;
                double px = 0.0, py = 0.0, pz = 0.0;
                const double rij = L.d[k];
                if (rij < G.rc && rij > G.rmin && rij > mb.rl && rij < mb.rcm) {
                    double rad, drad;
                    radial_fdf(mb, rij, rad, drad);
                    const double Z = M->Zel[L.el[k]];
                    if (wt) gs += Z * rad;
                    else gs += rad;
                    if (drad != 0.0) {
                        if (wt) drad *= Z * mb.sfac / rij;
                        else drad *= mb.sfac / rij;
                        px = drad * L.dr[k];
                        py = drad * L.dr[capL + k];
                        pz = drad * L.dr[2 * capL + k];
                        sx += px; sy += py; sz += pz;
                    }
                }
                if (deriv) {
                    double* d = dGn + ((size_t)k * o.S + (wt ? mb.slotE[L.el[k]] : mb.slot1)) * 3;
                    d[0] = -px; d[1] = -py; d[2] = -pz;
                }
            }
            block_sum_slot(gs, L.red, 0);
            block_sum_slot(sx, L.red, 1);
            block_sum_slot(sy, L.red, 2);
            block_sum_slot(sz, L.red, 3);
            __syncthreads();
            if (tid < 4) {
                const double v = red_total(L.red, tid);
                if (tid == 0) Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], v);
                else if (deriv) dGi[mb.sf * 3 + tid - 1] = v;
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
        const double r2min = G.rmin * G.rmin, r2max = G.rc2;
        const MemberT* mem = M->mem + G.mem0;

        for (int q0 = 0; q0 < G.nmem; q0 += CH) {
            const int nq = min(CH, G.nmem - q0);
            if (deriv)
                for (int x = tid; x < nq * 3 * capL; x += kCThreads) L.acc[x] = 0.0;
            __syncthreads();
            double gacc[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) gacc[q] = 0.0;

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
                const double dr1[3] = {L.dr[j], L.dr[capL + j], L.dr[2 * capL + j]};
                const double dr2[3] = {L.dr[k], L.dr[capL + k], L.dr[2 * capL + k]};
                const double dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1], dr2[2] - dr1[2]};
                double rjk = 0.0;
                if (narrow) {
                    rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
                    if (rjk >= r2max || rjk <= r2min) continue;
                    rjk = sqrt(rjk);
                }
                const double rij = L.d[j], rik = L.d[k];
                const double rinvijik = 1.0 / (rij * rik);
                const double costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
                // the angular polynomial vanishes at 0 and 180 degrees
                if (costijk <= -1.0 || costijk >= 1.0) continue;
                const double th = acos(costijk);
                const double rinvij = rinvijik * rik;
                const double rinvik = rinvijik * rij;
                const double rinvjk = narrow ? 1.0 / rjk : 0.0;
                const double phiijik0 = rinvij * (rinvik - rinvij * costijk);
                const double phiikij0 = rinvik * (rinvij - rinvik * costijk);
                const double dacos = -1.0 / sqrt(1.0 - costijk * costijk);
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    if (q >= nq) break;
                    const int l = q0 + q;
                    const MemberT& mb = mem[l];
                    double dradij;
                    const double radij = radial_ij(mb, rij, dradij);
                    if (member_skipped(mb, radij, rik, rjk, narrow, th)) continue;
                    double radik, dradik, radjk = 1.0, dradjk = 0.0, ang, dang;
                    radial_fdf(mb, rik, radik, dradik);
                    if (narrow) {
                        // r_jk factor: this member's, or the one the
                        // reference carries over from earlier members
                        int src = l;
                        if (mb.walk) {
                            while (!mem[src].calcComp) {
                                double dp;
                                const double rp = radial_ij(mem[src - 1], rij, dp);
                                if (member_skipped(mem[src - 1], rp, rik, rjk, true, th)) break;
                                --src;
                            }
                        }
                        radial_fdf(mem[src], rjk, radjk, dradjk);
                    }
                    compact_fdf(mb.core, 0, mb.acen, mb.ascl, th, ang, dang);
                    double z = 1.0;
                    if (wt) {
                        z = M->Zel[L.el[j]] * M->Zel[L.el[k]];
                        ang *= z;
                    }
                    double rad = narrow ? radij * radik * radjk : radij * radik;
                    gacc[q] += rad * ang;
                    if (!deriv) continue;
                    dang *= dacos;
                    const double phiijik = phiijik0 * dang;
                    const double phiikij = phiikij0 * dang;
                    const double psiijik = rinvijik * dang;
                    double chiij, chiik, chijk = 0.0;
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
                    double p1, p2, p3;
                    if (dang != 0.0 && rad != 0.0) {
                        if (wt) rad *= mb.sfac * z;
                        else rad *= mb.sfac;
                        ang *= mb.sfac;
                        p1 = rad * phiijik + ang * chiij;
                        p2 = rad * phiikij + ang * chiik;
                        p3 = narrow ? rad * psiijik + ang * chijk : rad * psiijik;
                    } else if (ang != 0.0) {
                        ang *= mb.sfac;
                        p1 = ang * chiij;
                        p2 = ang * chiik;
                        p3 = narrow ? ang * chijk : 0.0;
                    } else {
                        continue;
                    }
                    for (int c = 0; c < 3; ++c) {
                        const double p1c = p1 * dr1[c], p2c = p2 * dr2[c], p3c = p3 * dr3[c];
                        double* acc = L.acc + (size_t)(q * 3 + c) * capL;
                        lds_add(acc + j, -(p1c + p3c));
                        lds_add(acc + k, -(p2c - p3c));
                    }
                }
            }
            // G of the chunk's members.
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (q < nq) block_sum_slot(gacc[q], L.red, q);
            __syncthreads();
            if (tid < nq) {
                const MemberT& mb = mem[q0 + tid];
                Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], red_total(L.red, tid));
            }
            if (deriv) {
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
                    const int q = tid / 3, c = tid % 3;
                    double v = 0.0;
                    for (int k = 0; k < n; ++k) v += L.acc[(size_t)(q * 3 + c) * capL + k];
                    dGi[mem[q0 + q].sf * 3 + c] = -v;
                }
            }
            __syncthreads();
        }
    }
}

size_t comp_lds(int capL, int ch)
{
    return sizeof(double) * ((size_t)capL * 4 + kCWaves * kRedSlots + (size_t)ch * 3 * capL) +
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
    switch (ch) {
    case 8: hipLaunchKernelGGL(k_sfcomp64<8>, grid, block, lds, stream, m, s, o, d, capL); break;
    case 4: hipLaunchKernelGGL(k_sfcomp64<4>, grid, block, lds, stream, m, s, o, d, capL); break;
    case 2: hipLaunchKernelGGL(k_sfcomp64<2>, grid, block, lds, stream, m, s, o, d, capL); break;
    default: hipLaunchKernelGGL(k_sfcomp64<1>, grid, block, lds, stream, m, s, o, d, capL); break;
    }
}

}  // namespace nnpx
