// FP64 symmetry-function groups with stored dG/dr (the plain form).
//
// Computes, per central atom, what SymGrpExpRad::calculate
// (SymGrpExpRad.cpp:144-239) and SymGrpExpAngn::calculate
// (SymGrpExpAngn.cpp:156-403) compute: scaled G, the atom's own dG/dr
// (Atom::dGdr) and every neighbour entry's dG/dr in the compact slot order
// of Element::setupSymmetryFunctionMemory (Atom::Neighbor::dGdr).
//
// Mapping: one 256-thread workgroup per atom.  Neighbour geometry (dr, |dr|,
// element, cutoff values) is staged in LDS.  Radial groups: threads over
// neighbours, members in sequence.  Angular groups: threads over the
// unordered neighbour pairs of the group's element pair (a balanced
// "folded" enumeration for like elements), members unrolled into registers
// for G; the j/k dG/dr contributions are accumulated with LDS FP64 atomics
// and written once per group.  The atom's own dG/dr is the negative sum of
// its neighbours' entries (each term adds +p to the atom and -p to
// neighbours), which removes one accumulator set per member.
#include "kernels.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kSFThreads = 256;
constexpr int kSFWaves = kSFThreads / kWave;

struct Lds {
    double* dr;    // [3][cap]
    double* d;     // [cap]
    double* fc;    // [cap]
    double* dfc;   // [cap]
    int* el;       // [cap]
    int* la;       // [cap] list of neighbours with element e1 inside rc
    int* lb;       // [cap] list with element e2
    double* acc;   // [nmem][3][cap] (neighbour index fastest: conflict-free atomics)
    double* red;   // [kSFWaves][4*kMaxMembersExp]
};

__device__ inline double block_sum(double v, double* red, int slot)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[w * 4 * kMaxMembersExp + slot] = v;
    return v;
}

// Ordered compaction of neighbours of element `elem` (any element if
// negative) inside rc into a list (a single wave does it so the list keeps
// neighbour order).
__device__ inline int build_list(const Lds& L, int n, int elem, double rc, int* out)
{
    __shared__ int cnt_sh;
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        int cnt = 0;
        for (int b = 0; b < n; b += kWave) {
            const int k = b + lane;
            const bool ok = k < n && (elem < 0 || L.el[k] == elem) && L.d[k] < rc;
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

template <int MAXM>
__global__ __launch_bounds__(kSFThreads) void k_sf64(const ModelT* __restrict__ M, SysView s,
                                                      SF64Out o, int deriv, int capL, int accM)
{
    extern __shared__ double smem[];
    Lds L;
    L.dr = smem;
    L.d = L.dr + 3 * capL;
    L.fc = L.d + capL;
    L.dfc = L.fc + capL;
    L.red = L.dfc + capL;
    L.acc = L.red + kSFWaves * 4 * kMaxMembersExp;
    L.el = reinterpret_cast<int*>(L.acc + (size_t)capL * accM * 3);
    L.la = L.el + capL;
    L.lb = L.la + capL;

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const double tanhPre = M->tanhPre;
    const StructT& S = s.st[s.sid[i]];
    const int n = min(s.nb_n[i], capL);
    const size_t ebase = (size_t)i * s.cap;

    for (int k = tid; k < n; k += kSFThreads) {
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
        if (G.type != 2 && G.type != 12) continue;  // compact groups: k_sfcomp64
        // weighted (SymGrpExpRadWeighted.cpp:137-231): every neighbour
        // element, each term times its atomic number
        const bool wt = G.type == 12;
        for (int k = tid; k < n; k += kSFThreads) {
            double f = 0.0, df = 0.0;
            if ((wt || L.el[k] == G.e1) && L.d[k] < G.rc) cutoff_fdf(ctype, G, tanhPre, L.d[k], f, df);
            L.fc[k] = f;
            L.dfc[k] = df;
        }
        __syncthreads();
        for (int q = 0; q < G.nmem; ++q) {
            const MemberT& mb = M->mem[G.mem0 + q];
            double gs = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (int k = tid; k < n; k += kSFThreads) {
                if (!wt && L.el[k] != G.e1) continue;
                double px = 0.0, py = 0.0, pz = 0.0;
                const double rij = L.d[k];
                if (rij < G.rc) {
                    const double pfc = L.fc[k], pdfc = L.dfc[k];
                    double pexp = exp(-mb.eta * (rij - mb.rs) * (rij - mb.rs));
                    if (wt) pexp = M->Zel[L.el[k]] * pexp;
                    gs += pexp * pfc;
                    const double p1 = mb.sfac * (pdfc - 2.0 * mb.eta * (rij - mb.rs) * pfc) * pexp / rij;
                    px = p1 * L.dr[k];
                    py = p1 * L.dr[capL + k];
                    pz = p1 * L.dr[2 * capL + k];
                    sx += px; sy += py; sz += pz;
                }
                if (deriv) {
                    double* d = dGn + ((size_t)k * o.S + (wt ? mb.slotE[L.el[k]] : mb.slot1)) * 3;
                    d[0] = -px; d[1] = -py; d[2] = -pz;
                }
            }
            block_sum(gs, L.red, 0);
            block_sum(sx, L.red, 1);
            block_sum(sy, L.red, 2);
            block_sum(sz, L.red, 3);
            __syncthreads();
            if (tid < 4) {
                double v = 0.0;
                for (int w = 0; w < kSFWaves; ++w) v += L.red[w * 4 * kMaxMembersExp + tid];
                if (tid == 0) Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], v);
                else if (deriv) dGi[mb.sf * 3 + tid - 1] = v;
            }
            __syncthreads();
        }
    }

    // --------------------------------------------------------- angular groups
    for (int gi = E.nrad; gi < E.ngrp; ++gi) {
        const GroupT& G = M->grp[E.grp0 + gi];
        if (G.type != 3 && G.type != 9 && G.type != 13) continue;  // compact groups: k_sfcomp64
        const int nmem = G.nmem;
        const bool same = G.e1 == G.e2;
        // wide angular groups (SymGrpExpAngw.cpp:156-380): no r_jk cutoff and
        // no r_jk term in the exponent
        const bool wide = G.type == 9;
        // weighted (SymGrpExpAngnWeighted.cpp:165-395): all neighbour pairs,
        // cutoff product times Z_j Z_k, exponent always in shifted form
        const bool wt = G.type == 13;
        const int nA = build_list(L, n, G.e1, G.rc, L.la);
        const int nB = same ? nA : build_list(L, n, G.e2, G.rc, L.lb);
        const int* lb = same ? L.la : L.lb;
        for (int k = tid; k < n; k += kSFThreads) {
            double f = 0.0, df = 0.0;
            if ((wt || L.el[k] == G.e1 || L.el[k] == G.e2) && L.d[k] < G.rc)
                cutoff_fdf(ctype, G, tanhPre, L.d[k], f, df);
            L.fc[k] = f;
            L.dfc[k] = df;
        }
        if (deriv)
            for (int x = tid; x < nmem * 3 * capL; x += kSFThreads) L.acc[x] = 0.0;
        __syncthreads();

        double gacc[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; ++q) gacc[q] = 0.0;

        int D = 0, full = 0, total;
        if (same) {
            D = (nA - 1) / 2;
            full = nA * D;
            total = full + ((nA % 2 == 0) ? nA / 2 : 0);
        } else {
            total = nA * nB;
        }
        const double rc2 = G.rc2;
        for (int t = tid; t < total; t += kSFThreads) {
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
            double rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
            if (!wide && !(rjk < rc2)) continue;
            const double rij = L.d[j], rik = L.d[k];
            const double pfcij = L.fc[j], pdfcij = L.dfc[j];
            const double pfcik = L.fc[k], pdfcik = L.dfc[k];
            const double r2ij = rij * rij;
            const double rinvijik = 1.0 / rij / rik;
            const double costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1] + dr1[2] * dr2[2]) * rinvijik;
            const double r2ik = rik * rik;
            double pfc, r2sum, pr1, pr2, pr3 = 0.0;
            if (wide) {
                pfc = pfcij * pfcik;
                r2sum = r2ij + r2ik;
                pr1 = pfcik * pdfcij / rij;
                pr2 = pfcij * pdfcik / rik;
            } else {
                rjk = sqrt(rjk);
                double pfcjk, pdfcjk;
                cutoff_fdf(ctype, G, tanhPre, rjk, pfcjk, pdfcjk);
                r2sum = r2ij + r2ik + rjk * rjk;
                if (wt) {
                    const double z = M->Zel[L.el[j]] * M->Zel[L.el[k]];
                    pfc = z * pfcij * pfcik * pfcjk;
                    pr1 = z * pfcik * pfcjk * pdfcij / rij;
                    pr2 = z * pfcij * pfcjk * pdfcik / rik;
                    pr3 = z * pfcij * pfcik * pdfcjk / rjk;
                } else {
                    pfc = pfcij * pfcik * pfcjk;
                    pr1 = pfcik * pfcjk * pdfcij / rij;
                    pr2 = pfcij * pfcjk * pdfcik / rik;
                    pr3 = pfcij * pfcik * pdfcjk / rjk;
                }
            }
            double vexp = 0.0, rijs = 0.0, riks = 0.0, rjks = 0.0;
#pragma unroll
            for (int q = 0; q < MAXM; ++q) {
                if (q < nmem) {
                const MemberT& mb = M->mem[G.mem0 + q];
                if (mb.calcExp) {
                    if (mb.rs > 0.0 || wt) {
                        rijs = rij - mb.rs;
                        riks = rik - mb.rs;
                        if (wide) {
                            vexp = exp(-mb.eta * (rijs * rijs + riks * riks));
                        } else {
                            rjks = rjk - mb.rs;
                            vexp = exp(-mb.eta * (rijs * rijs + riks * riks + rjks * rjks));
                        }
                    } else {
                        vexp = exp(-mb.eta * r2sum);
                    }
                }
                const double plambda = 1.0 + mb.lambda * costijk;
                double fg = vexp;
                if (plambda <= 0.0) fg = 0.0;
                else if (mb.useIntPow) fg *= pow_int(plambda, mb.zetaInt - 1);
                else fg *= pow(plambda, mb.zeta - 1.0);
                gacc[q] += fg * plambda * pfc;
                if (deriv) {
                fg *= mb.factorNorm;
                const double pfczl = pfc * mb.zetaLambda;
                const double p2etapl = plambda * mb.factorDeriv;
                double p1, p2, p3;
                if (mb.rs > 0.0 || wt) {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl * rijs / rij) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl * riks / rik) + pr2 * plambda);
                    p3 = wide ? fg * pfczl * rinvijik
                              : fg * (pfczl * (rinvijik + p2etapl * rjks / rjk) - pr3 * plambda);
                } else {
                    p1 = fg * (pfczl * (rinvijik - costijk / r2ij - p2etapl) + pr1 * plambda);
                    p2 = fg * (pfczl * (rinvijik - costijk / r2ik - p2etapl) + pr2 * plambda);
                    p3 = wide ? fg * pfczl * rinvijik : fg * (pfczl * (rinvijik + p2etapl) - pr3 * plambda);
                }
                for (int c = 0; c < 3; ++c) {
                    const double p1c = p1 * dr1[c], p2c = p2 * dr2[c], p3c = p3 * dr3[c];
                    double* a = L.acc + (size_t)(q * 3 + c) * capL;
                    lds_add(a + j, -(p1c + p3c));
                    lds_add(a + k, -(p2c - p3c));
                }
                }
                }
            }
        }
        // G of the group's members.
#pragma unroll
        for (int q = 0; q < MAXM; ++q)
            if (q < nmem) block_sum(gacc[q], L.red, q);
        __syncthreads();
        if (tid < nmem) {
            const MemberT& mb = M->mem[G.mem0 + tid];
            double v = 0.0;
            for (int w = 0; w < kSFWaves; ++w) v += L.red[w * 4 * kMaxMembersExp + tid];
            v *= mb.gnorm;
            Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], v);
        }
        if (deriv) {
            // Neighbour entries of this group's slots (zero where no triplet).
            for (int x = tid; x < n * nmem; x += kSFThreads) {
                const int k = x / nmem, q = x % nmem;
                const int ek = L.el[k];
                if (!wt && ek != G.e1 && ek != G.e2) continue;
                const MemberT& mb = M->mem[G.mem0 + q];
                const int slot = wt ? mb.slotE[ek] : ek == G.e1 ? mb.slot1 : mb.slot2;
                double* d = dGn + ((size_t)k * o.S + slot) * 3;
                const double* a = L.acc + (size_t)q * 3 * capL + k;
                d[0] = a[0]; d[1] = a[capL]; d[2] = a[2 * capL];
            }
            // The atom's own dG/dr: minus the sum over its neighbour entries.
            if (tid < nmem * 3) {
                const int q = tid / 3, c = tid % 3;
                double v = 0.0;
                for (int k = 0; k < n; ++k) v += L.acc[(size_t)(q * 3 + c) * capL + k];
                dGi[M->mem[G.mem0 + q].sf * 3 + c] = -v;
            }
        }
        __syncthreads();
    }
}

}  // namespace

void launch_sf64(const ModelT* m, const ModelT& hm, const SysView& s, const SF64Out& o,
                 bool deriv, int max_nb, hipStream_t stream)
{
    if (s.natoms == 0 || !hm.has_exp) return;
    const int capL = ((std::max(max_nb, 1) + 31) / 32) * 32;
    const int mam = std::max(hm.max_mem_ang, 1);
    const int MAXM = mam <= 4 ? 4 : mam <= 8 ? 8 : mam <= 16 ? 16 : 32;
    // the dG/dr accumulators only when derivatives are stored
    const int accM = deriv ? mam : 0;
    const size_t lds = sizeof(double) * ((size_t)capL * 6 + kSFWaves * 4 * kMaxMembersExp +
                                         (size_t)capL * accM * 3) +
                       sizeof(int) * (size_t)capL * 3;
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the FP64 SF kernel's LDS staging"};
    const dim3 grid(s.natoms), block(kSFThreads);
    const int d = deriv ? 1 : 0;
    switch (MAXM) {
    case 4: hipLaunchKernelGGL(k_sf64<4>, grid, block, lds, stream, m, s, o, d, capL, accM); break;
    case 8: hipLaunchKernelGGL(k_sf64<8>, grid, block, lds, stream, m, s, o, d, capL, accM); break;
    case 16: hipLaunchKernelGGL(k_sf64<16>, grid, block, lds, stream, m, s, o, d, capL, accM); break;
    default: hipLaunchKernelGGL(k_sf64<32>, grid, block, lds, stream, m, s, o, d, capL, accM); break;
    }
}

}  // namespace nnpx
