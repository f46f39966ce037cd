// FP64 fused force backward of the plain form.
//
// The forces Mode::calculateForces computes from the stored Atom::dGdr and
// Atom::Neighbor::dGdr (Mode.cpp:1922-1984, Atom.cpp:371-401), without
// storing them: every radial pair term and angular triplet term of
// SymGrpExpRad::calculate / SymGrpExpAngn::calculate (SymGrpExpRad.cpp:
// 144-239, SymGrpExpAngn.cpp:156-403, and the wide and element-weighted
// groups SymGrpExpAngw.cpp:156-380, SymGrpExpRadWeighted.cpp:137-231,
// SymGrpExpAngnWeighted.cpp:165-395) is recomputed in FP64 and contracted on
// the fly with w = dE/dG times the member's derivative scale (angular
// 2^(1-zeta) times the scaling factor, radial the scaling factor).  Per
// triplet the members reduce to three sums, A = sum u zeta lambda,
// B = sum u (1 + lambda cos), C = sum u 2 eta (1 + lambda cos) with
// u = w exp(...) (1 + lambda cos)^(zeta-1), which give the reference's
// p1, p2, p3 summed over the members:
//   P1 = pfc (A (1/(rij rik) - cos/rij^2) - C) + B pr1
//   P2 = pfc (A (1/(rij rik) - cos/rik^2) - C) + B pr2
//   P3 = pfc (A /(rij rik) + C) - B pr3
// (shifted and weighted members carry (r - rs)/r factors into C per
// distance), and the atom receives -(P1 r_ij + P2 r_ik), neighbour j
// +(P1 r_ij + P3 r_jk), neighbour k +(P2 r_ik - P3 r_jk).
//
// Mapping: one wave per atom, neighbour geometry and cutoff values staged in
// LDS in FP64.  Radial groups: lanes own neighbours.  Angular groups: the
// group's neighbour pairs are enumerated row-major from a linear index,
// tested (r_jk^2 < rc^2 in FP64), compacted through LDS, and each lane takes
// a contiguous chunk, so it sees runs of one first atom j: j's geometry and
// its force contribution stay in registers over the run, the atom's own
// contribution too, and only partner k's goes through an LDS FP64 atomic.
// Neighbour sums are pushed to the force array with FP64 device atomics;
// the pull form drops the atom's own periodic images (Mode.cpp:1957).
#include "kernels.hpp"
#include "mixed.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kFill64 = 4;   // candidate tests per lane per fill step
constexpr int kTrip64 = 384; // compacted triplet buffer (multiple of kFill64 * 64)

struct Stage64 {
    double4* g;          // x, y, z, |dr|
    double4* c;          // 1/|dr|, fc, dfc (current group's cutoff), atomic number
    double* ax;          // [capL] force sums per neighbour entry
    double* ay;
    double* az;
    unsigned* tjk;       // compacted pairs (j | k << 16)
    unsigned short* la;  // group lists
    unsigned short* lb;
    unsigned char* el;   // neighbour element
    unsigned short* idx; // list entry of each staged slot
    int* eoff;           // [kMaxElem + 1] first staged slot of each neighbour element
    double2* pw;         // [kPw64] power-path term weights (force) or term sums (values)
    int trip;
};

// A group's neighbour list: staged slots base .. base + n - 1 (contig: the
// group's cutoff covers every staged neighbour of its element) or the n
// slots listed in S.la / S.lb.
struct List64 {
    int base, n;
    bool contig;
};

template <bool CONTIG>
__device__ inline int slot64(const unsigned short* l, int base, int a)
{
    if constexpr (CONTIG) return base + a;
    else return l[a];
}

constexpr int kPw64 = 32;

__host__ __device__ constexpr size_t stage64_bytes(int capL, int trip)
{
    return (size_t)kPw64 * 16 + (size_t)capL * (32 + 32 + 24 + 2 + 2 + 1 + 2) + (size_t)trip * 4 + 48;
}

__device__ inline Stage64 carve64(double4* base, int capL, int trip)
{
    Stage64 S;
    S.trip = trip;
    S.pw = reinterpret_cast<double2*>(base);
    S.g = base + kPw64 / 2;
    S.c = S.g + capL;
    S.ax = reinterpret_cast<double*>(S.c + capL);
    S.ay = S.ax + capL;
    S.az = S.ay + capL;
    S.tjk = reinterpret_cast<unsigned*>(S.az + capL);
    S.la = reinterpret_cast<unsigned short*>(S.tjk + trip);
    S.lb = S.la + capL;
    S.el = reinterpret_cast<unsigned char*>(S.lb + capL);
    S.idx = reinterpret_cast<unsigned short*>(S.el + capL);
    S.eoff = reinterpret_cast<int*>(S.idx + capL);
    return S;
}

// Stages the atom's list entries closer than the model's largest cutoff
// (entries beyond it contribute nothing; MD lists built with a skin carry
// them), grouped by neighbour element and in list order within an element,
// so element e's neighbours are the slots eoff[e] .. eoff[e + 1] - 1; each
// slot's list entry in idx and, for the force kernel, zeroed neighbour force
// sums.  More such entries than capL: only slots below capL are staged, flag
// bit 4 is set (the caller repeats with the full capacity) and flags[2]
// keeps the largest count met.
__device__ inline int stage64(const ModelT* __restrict__ M, const SysView& s, int i, int capL, const Stage64& S,
                              bool zero_sums)
{
    const int lane = threadIdx.x;
    const StructT& St = s.st[s.sid[i]];
    const int nl = s.nb_n[i];
    const int ne = M->nelem;
    const size_t base = (size_t)i * s.cap;
    const double rmax = M->maxCutoff;
    auto entry = [&](int k, double* dr, double& d) {
        int el = -1;
        d = 0.0;
        dr[0] = dr[1] = dr[2] = 0.0;
        if (k < nl) {
            entry_vector(s, i, base + k, St, dr);
            d = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
            if (d < rmax) el = s.elem[s.nb_j[base + k]];
        }
        return el;
    };
    // pass 1: in-cutoff entries per element (lane e counts element e)
    int cnt = 0;
    for (int b = 0; b < nl; b += kWave) {
        double dr[3], d;
        const int el = entry(b + lane, dr, d);
        for (int e = 0; e < ne; ++e) {
            const int c = __popcll(__ballot(el == e));
            if (lane == e) cnt += c;
        }
    }
    const int incl = wave_incl_scan(lane < ne ? cnt : 0);
    const int total = lane_value(incl, kWave - 1);
    int next = incl - (lane < ne ? cnt : 0);  // lane e: next slot of element e
    if (lane <= ne) S.eoff[lane] = min(lane < ne ? next : total, capL);
    // pass 2: each entry to its element's next slot
    for (int b = 0; b < nl; b += kWave) {
        const int k = b + lane;
        double dr[3], d;
        const int el = entry(k, dr, d);
        int pos = -1;
        for (int e = 0; e < ne; ++e) {
            const unsigned long long m = __ballot(el == e);
            const int at = lane_value(next, e);
            if (el == e)
                pos = at + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (lane == e) next += __popcll(m);
        }
        if (pos >= 0 && pos < capL) {
            S.g[pos] = make_double4(dr[0], dr[1], dr[2], d);
            S.c[pos] = make_double4(1.0 / d, 0.0, 0.0, M->Zel[el]);
            S.el[pos] = (unsigned char)el;
            S.idx[pos] = (unsigned short)k;
            if (zero_sums) {
                S.ax[pos] = 0.0;
                S.ay[pos] = 0.0;
                S.az[pos] = 0.0;
            }
        }
    }
    if (lane == 0) {
        if (total > capL) atomicOr(&s.flags[0], 4);
        if (total > __hip_atomic_load(&s.flags[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&s.flags[2], total);
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(min(total, capL));
}

// Cutoff value and derivative; the cosine cutoff through sincospi (exact
// argument reduction for x in [0, 1]), the others as CutoffFunction::fdf.
__device__ inline void cutoff_fdf64(int type, const GroupT& g, double tanhPre, double r, double& fc, double& dfc)
{
    if (type == 1 && r < g.rc && r >= g.rci) {
        const double x = (r - g.rci) * g.iw;
        double sn, cs;
        sincospi(x, &sn, &cs);
        fc = 0.5 * (cs + 1.0);
        dfc = -0.5 * g.iw * M_PI * sn;
        return;
    }
    cutoff_fdf(type, g, tanhPre, r, fc, dfc);
}

// Neighbours of element `elem` (all elements: elem < 0) inside the group's
// cutoff: the element's slot range when the group's cutoff is the model's
// largest, else the qualifying slots listed in `out`.  Ends with a barrier.
__device__ inline List64 make_list64(const ModelT* __restrict__ M, const Stage64& S, int n, const GroupT& G, int elem,
                                     unsigned short* out)
{
    const int lane = threadIdx.x;
    const int lo = __builtin_amdgcn_readfirstlane(elem < 0 ? 0 : S.eoff[elem]);
    const int hi = __builtin_amdgcn_readfirstlane(elem < 0 ? n : S.eoff[elem + 1]);
    List64 L{lo, hi - lo, true};
    if (!(G.rc >= M->maxCutoff)) {
        int cnt = 0;
        for (int b = lo; b < hi; b += kWave) {
            const int k = b + lane;
            const bool ok = k < hi && S.g[k].w < G.rc;
            const int incl = wave_incl_scan(ok ? 1 : 0);
            if (ok) out[cnt + incl - 1] = (unsigned short)k;
            cnt += lane_value(incl, kWave - 1);
        }
        L = List64{0, __builtin_amdgcn_readfirstlane(cnt), false};
    }
    __syncthreads();
    return L;
}

// Calls work(j, k) for every unordered pair of the group's lists with
// r_jk^2 < rc2 (like elements: row a pairs with b = a + c + 1 mod nA,
// c < (nA - 1) / 2, plus b = a + nA / 2 for even nA on rows a < nA / 2;
// unlike elements: the full product), rows consecutive in the compacted
// buffer; flush() after each buffer.
template <bool CONTIG, typename Work, typename Flush>
__device__ inline void triplets64(const Stage64& S, const List64& A, const List64& B, bool same, double rc2,
                                 Work&& work, Flush&& flush)
{
    const int lane = threadIdx.x;
    const int nA = A.n, nB = B.n;
    const int bB = same ? A.base : B.base;
    int full, total, R;
    if (same) {
        R = (nA - 1) / 2;
        full = nA * R;
        total = full + ((nA % 2 == 0) ? nA / 2 : 0);
    } else {
        R = nB;
        full = nA * nB;
        total = full;
    }
    const unsigned short* lb = same ? S.la : S.lb;
    const float inv = R > 0 ? 1.f / (float)R : 0.f;
    int t0 = 0;
    while (t0 < total) {
        int cnt = 0;
        while (t0 < total && cnt <= S.trip - kFill64 * kWave) {
            bool ok[kFill64];
            unsigned jk[kFill64];
#pragma unroll
            for (int f = 0; f < kFill64; ++f) {
                const int t = t0 + f * kWave + lane;
                ok[f] = false;
                jk[f] = 0;
                if (t < total) {
                    int a, b;
                    if (t < full) {
                        int q = (int)((float)t * inv);
                        int r = t - q * R;
                        if (r < 0) { --q; r += R; }
                        else if (r >= R) { ++q; r -= R; }
                        a = q;
                        if (same) { b = a + r + 1; if (b >= nA) b -= nA; }
                        else b = r;
                    } else {
                        a = t - full;
                        b = a + nA / 2;
                    }
                    const int j = slot64<CONTIG>(S.la, A.base, a), k = slot64<CONTIG>(lb, bB, b);
                    const double4 gj = S.g[j], gk = S.g[k];
                    const double dx = gk.x - gj.x, dy = gk.y - gj.y, dz = gk.z - gj.z;
                    ok[f] = dx * dx + dy * dy + dz * dz < rc2;
                    jk[f] = (unsigned)j | ((unsigned)k << 16);
                }
            }
#pragma unroll
            for (int f = 0; f < kFill64; ++f) {
                const unsigned long long m = __ballot(ok[f]);
                if (ok[f]) {
                    const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                    S.tjk[pos] = jk[f];
                }
                cnt += __popcll(m);
            }
            t0 += kFill64 * kWave;
        }
        __syncthreads();
        const int per = (cnt + kWave - 1) / kWave;
        const int x0 = lane * per, x1 = min(cnt, x0 + per);
        for (int x = x0; x < x1; ++x) work(S.tjk[x] & 0xffffu, S.tjk[x] >> 16);
        flush();
        __syncthreads();
    }
}

template <typename Work, typename Flush>
__device__ inline void for_each_triplet64(const Stage64& S, const List64& A, const List64& B, bool same, double rc2,
                                          Work&& work, Flush&& flush)
{
    if (A.contig) triplets64<true>(S, A, B, same, rc2, work, flush);
    else triplets64<false>(S, A, B, same, rc2, work, flush);
}

// NB0 > 0: the power path (angular groups whose members are
// exp(-eta_h sum r^2) (1 + lambda_s cos)^(2^b), b < NB0 resp. NB1, rs = 0):
// per triplet one exponential per eta and the powers by repeated squaring;
// the members' weights are folded per term (wb = sum w, wa = sum zeta lambda
// w), and (1 + lambda cos)^(2^b - 1) is carried beside the power so A needs
// no division.
template <int MAXM, int NB0, int NB1>
__global__ __launch_bounds__(64) void k_force64f(const ModelT* __restrict__ M, const MemberF* __restrict__ MF,
                                                  const GroupPow64* __restrict__ GP, SysView s,
                                                  const double* __restrict__ fw, int ldG, double* __restrict__ F,
                                                  int push, int capL, int trip)
{
    extern __shared__ double4 sm64[];
    const Stage64 S = carve64(sm64, capL, trip);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const double tanhPre = M->tanhPre;
    const int n = stage64(M, s, i, capL, S, true);
    const size_t base = (size_t)i * s.cap;
    const double* wi = fw + (size_t)i * ldG;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    bool cut_done = false;  // the cutoffs of a uniform-cutoff element are staged

    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& G = M->grp[E.grp0 + gi];
        const int type = G.type;
        if (type != 2 && type != 12 && type != 3 && type != 9 && type != 13) continue;  // compact: k_sfcomp
        const bool radial = type == 2 || type == 12;
        const bool wt = type == 12 || type == 13;
        // cutoffs of the group's neighbours (every slot once per atom when
        // all the element's groups share one radius)
        if (!E.ucut || !cut_done) {
            for (int k = lane; k < n; k += kWave) {
                double f = 0.0, df = 0.0;
                const int ek = S.el[k];
                if ((E.ucut || wt || ek == G.e1 || ek == G.e2) && S.g[k].w < G.rc)
                    cutoff_fdf64(ctype, G, tanhPre, S.g[k].w, f, df);
                S.c[k].y = f;
                S.c[k].z = df;
            }
            __syncthreads();
            cut_done = true;
        }
        const int nmem = G.nmem;
        const MemberT* mem = M->mem + G.mem0;
        if (radial) {
            // lanes own neighbours: sum_q w_q sfac_q (dfc - 2 eta (r - rs) fc) exp(-eta (r - rs)^2);
            // type 2: the slots of element e1, type 12: every slot
            const int lo = wt ? 0 : S.eoff[G.e1], hi = wt ? n : S.eoff[G.e1 + 1];
            for (int k = lo + lane; k < hi; k += kWave) {
                const double4 gk = S.g[k];
                if (!(gk.w < G.rc)) continue;
                const double4 ck = S.c[k];
                double P = 0.0;
                for (int q = 0; q < nmem; ++q) {
                    const MemberT& mb = mem[q];
                    const double dr = gk.w - mb.rs;
                    P += wi[mb.sf] * (ck.z - 2.0 * mb.eta * dr * ck.y) * exp(-mb.eta * dr * dr);
                }
                P *= wt ? ck.x * ck.w : ck.x;
                const double px = P * gk.x, py = P * gk.y, pz = P * gk.z;
                fx -= px; fy -= py; fz -= pz;
                S.ax[k] += px; S.ay[k] += py; S.az[k] += pz;
            }
            __syncthreads();
            continue;
        }
        const bool same = G.e1 == G.e2;
        const bool wide = type == 9;
        const List64 A = make_list64(M, S, n, G, G.e1, S.la);
        const List64 B = same ? A : make_list64(M, S, n, G, G.e2, S.lb);
        GroupPow64 gp{};
        if constexpr (NB0 > 0) {
            gp = GP[E.grp0 + gi];
            // term weights: wb = sum of the members' w, wa = sum of zeta lambda w
            if (lane < 2 * (NB0 + NB1)) {
                double sb = 0.0, sa = 0.0;
                for (int q = 0; q < nmem; ++q)
                    if (MF[G.mem0 + q].beta == lane) {
                        const double wq = wi[mem[q].sf];
                        sb += wq;
                        sa += wq * mem[q].zetaLambda;
                    }
                S.pw[lane] = make_double2(sb, sa);
            }
            __syncthreads();
        }
        int cur = -1;
        double rx = 0.0, ry = 0.0, rz = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
        double4 gj = make_double4(0.0, 0.0, 0.0, 0.0), cj = gj;
        auto flush = [&]() {
            if (cur >= 0) {
                lds_add(S.ax + cur, rx);
                lds_add(S.ay + cur, ry);
                lds_add(S.az + cur, rz);
                fx -= sx;
                fy -= sy;
                fz -= sz;
                cur = -1;
            }
        };
        for_each_triplet64(S, A, B, same, wide ? INFINITY : G.rc2, [&](int j, int k) {
            if (j != cur) {
                flush();
                cur = j;
                rx = 0.0; ry = 0.0; rz = 0.0;
                sx = 0.0; sy = 0.0; sz = 0.0;
                gj = S.g[j];
                cj = S.c[j];
            }
            const double4 gk = S.g[k];
            const double4 ck = S.c[k];
            const double x3 = gk.x - gj.x, y3 = gk.y - gj.y, z3 = gk.z - gj.z;
            const double r2jk = x3 * x3 + y3 * y3 + z3 * z3;
            const double rij = gj.w, rik = gk.w;
            double rjk = 0.0, ijk = 0.0, fcjk = 1.0, dfcjk = 0.0;
            if (!wide) {
                rjk = sqrt(r2jk);
                ijk = 1.0 / rjk;
                cutoff_fdf64(ctype, G, tanhPre, rjk, fcjk, dfcjk);
            }
            const double rinvijik = cj.x * ck.x;
            const double cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * rinvijik;
            const double zw = wt ? cj.w * ck.w : 1.0;
            const double pfc = zw * cj.y * ck.y * fcjk;
            const double pr1 = zw * ck.y * fcjk * cj.z * cj.x;
            const double pr2 = zw * cj.y * fcjk * ck.z * ck.x;
            const double pr3 = zw * cj.y * ck.y * dfcjk * ijk;
            const double r2sum = rij * rij + rik * rik + (wide ? 0.0 : r2jk);
            double A = 0.0, B = 0.0, C0 = 0.0, C1 = 0.0, C2 = 0.0, C3 = 0.0;
            double vexp = 0.0, fij = 0.0, fik = 0.0, fjk = 0.0;
            bool shifted = false;
            if constexpr (NB0 > 0) {
                constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
                const double E0 = exp(-gp.eta[0] * r2sum);
                const double E1 = NB1 > 0 ? exp(-gp.eta[1] * r2sum) : 0.0;
                double Q0 = 0.0, Q1 = 0.0;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const double pl = 1.0 + gp.lam[sl] * cost;
                    const bool pos = pl > 0.0;
                    double pp = pos ? pl : 0.0;   // pl^(2^b)
                    double qq = pos ? 1.0 : 0.0;  // pl^(2^b - 1)
                    double qa0 = 0.0, qa1 = 0.0;
#pragma unroll
                    for (int b = 0; b < NBM; ++b) {
                        if (b < NB0) {
                            const double2 wv = S.pw[sl * NB0 + b];
                            Q0 += wv.x * pp;
                            qa0 += wv.y * qq;
                        }
                        if (b < NB1) {
                            const double2 wv = S.pw[2 * NB0 + sl * NB1 + b];
                            Q1 += wv.x * pp;
                            qa1 += wv.y * qq;
                        }
                        qq *= pp;
                        pp *= pp;
                    }
                    A += E0 * qa0 + E1 * qa1;
                }
                const double B0 = E0 * Q0, B1 = E1 * Q1;
                B = B0 + B1;
                C0 = gp.twoEta[0] * B0 + gp.twoEta[1] * B1;
            } else {
#pragma unroll
            for (int q = 0; q < MAXM; ++q) {
                if (q < nmem) {
                    const MemberT& mb = mem[q];
                    if (mb.calcExp) {
                        shifted = mb.rs > 0.0 || wt;
                        if (shifted) {
                            const double a = rij - mb.rs, b = rik - mb.rs, c = rjk - mb.rs;
                            vexp = exp(-mb.eta * (a * a + b * b + (wide ? 0.0 : c * c)));
                            fij = a * cj.x;
                            fik = b * ck.x;
                            fjk = wide ? 0.0 : c * ijk;
                        } else {
                            vexp = exp(-mb.eta * r2sum);
                        }
                    }
                    const double pl = 1.0 + mb.lambda * cost;
                    if (pl > 0.0) {
                        const double pw = mb.useIntPow ? pow_int(pl, mb.zetaInt - 1) : pow(pl, mb.zeta - 1.0);
                        const double u = wi[mb.sf] * vexp * pw;
                        A += u * mb.zetaLambda;
                        const double t = u * pl;
                        B += t;
                        const double c2 = 2.0 * mb.eta * t;
                        if (shifted) {
                            C1 += c2 * fij;
                            C2 += c2 * fik;
                            C3 += c2 * fjk;
                        } else {
                            C0 += c2;
                        }
                    }
                }
            }
            }
            C1 += C0;
            C2 += C0;
            C3 += C0;
            const double P1 = pfc * (A * (rinvijik - cost * cj.x * cj.x) - C1) + B * pr1;
            const double P2 = pfc * (A * (rinvijik - cost * ck.x * ck.x) - C2) + B * pr2;
            const double P3 = wide ? pfc * A * rinvijik : pfc * (A * rinvijik + C3) - B * pr3;
            const double p1x = P1 * gj.x, p1y = P1 * gj.y, p1z = P1 * gj.z;
            const double p2x = P2 * gk.x, p2y = P2 * gk.y, p2z = P2 * gk.z;
            const double p3x = P3 * x3, p3y = P3 * y3, p3z = P3 * z3;
            sx += p1x + p2x;
            sy += p1y + p2y;
            sz += p1z + p2z;
            rx += p1x + p3x;
            ry += p1y + p3y;
            rz += p1z + p3z;
            lds_add(S.ax + k, p2x - p3x);
            lds_add(S.ay + k, p2y - p3y);
            lds_add(S.az + k, p2z - p3z);
        }, flush);
        flush();
        __syncthreads();
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
        atomicAdd(F + 3 * (size_t)i, fx);
        atomicAdd(F + 3 * (size_t)i + 1, fy);
        atomicAdd(F + 3 * (size_t)i + 2, fz);
    }
    for (int k = lane; k < n; k += kWave) {
        const int j = s.nb_j[base + S.idx[k]];
        if (j == i && !push) continue;
        atomicAdd(F + 3 * (size_t)j, S.ax[k]);
        atomicAdd(F + 3 * (size_t)j + 1, S.ay[k]);
        atomicAdd(F + 3 * (size_t)j + 2, S.az[k]);
    }
}

// Scaled SF values of the plain form (what k_sf64 computes without
// derivatives) on the same staging and triplet walk as the force kernel:
// radial members summed by the lanes owning the neighbours, angular members
// per lane over its chunk of compacted triplets (no idle lanes on rejected
// candidates), one wave reduction per member.  Terms follow the
// reference's expressions (SymGrpExpRad.cpp:190-196, SymGrpExpAngn.cpp:
// 260-300): exp(-eta (r - rs)^2) fc, and vexp (1 + lambda cos)^(zeta - 1)
// (1 + lambda cos) fc_ij fc_ik fc_jk with the exponential shared by members
// of equal (eta, rs).
template <int MAXM, int NB0, int NB1>
__global__ __launch_bounds__(64) void k_values64(const ModelT* __restrict__ M, const MemberF* __restrict__ MF,
                                                  const GroupPow64* __restrict__ GP, SysView s,
                                                  double* __restrict__ G, int ldG, int capL, int trip)
{
    extern __shared__ double4 sm64[];
    const Stage64 S = carve64(sm64, capL, trip);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const double tanhPre = M->tanhPre;
    const int n = stage64(M, s, i, capL, S, false);
    double* Gi = G + (size_t)i * ldG;
    bool cut_done = false;  // the cutoffs of a uniform-cutoff element are staged

    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& GR = M->grp[E.grp0 + gi];
        const int type = GR.type;
        if (type != 2 && type != 12 && type != 3 && type != 9 && type != 13) continue;  // compact: k_sfcomp
        const bool radial = type == 2 || type == 12;
        const bool wt = type == 12 || type == 13;
        if (!E.ucut || !cut_done) {
            for (int k = lane; k < n; k += kWave) {
                double f = 0.0, df = 0.0;
                const int ek = S.el[k];
                if ((E.ucut || wt || ek == GR.e1 || ek == GR.e2) && S.g[k].w < GR.rc)
                    cutoff_fdf(ctype, GR, tanhPre, S.g[k].w, f, df);
                S.c[k].y = f;
            }
            __syncthreads();
            cut_done = true;
        }
        const int nmem = GR.nmem;
        const MemberT* mem = M->mem + GR.mem0;
        if (radial) {
            constexpr int CH = 8;
            const int lo = wt ? 0 : S.eoff[GR.e1], hi = wt ? n : S.eoff[GR.e1 + 1];
            for (int q0 = 0; q0 < nmem; q0 += CH) {
                double acc[CH];
#pragma unroll
                for (int q = 0; q < CH; ++q) acc[q] = 0.0;
                for (int k = lo + lane; k < hi; k += kWave) {
                    const double r = S.g[k].w;
                    if (!(r < GR.rc)) continue;
                    const double fc = S.c[k].y;
#pragma unroll
                    for (int q = 0; q < CH; ++q)
                        if (q0 + q < nmem) {
                            const MemberT& mb = mem[q0 + q];
                            double pexp = exp(-mb.eta * (r - mb.rs) * (r - mb.rs));
                            if (wt) pexp = S.c[k].w * pexp;
                            acc[q] += pexp * fc;
                        }
                }
#pragma unroll
                for (int q = 0; q < CH; ++q)
                    if (q0 + q < nmem) {
                        const double v = wave_sum(acc[q]);
                        if (lane == 0) {
                            const MemberT& mb = mem[q0 + q];
                            Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], v);
                        }
                    }
            }
            __syncthreads();
            continue;
        }
        const bool same = GR.e1 == GR.e2;
        const bool wide = type == 9;
        const List64 A = make_list64(M, S, n, GR, GR.e1, S.la);
        const List64 B = same ? A : make_list64(M, S, n, GR, GR.e2, S.lb);
        if constexpr (NB0 > 0) {
            // power path: term sums pfc E_h (1 + lambda_s cos)^(2^b) per lane,
            // one wave reduction per used term, members take their term
            constexpr int NBAS = 2 * (NB0 + NB1);
            constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
            const GroupPow64 gp = GP[E.grp0 + gi];
            double part[NBAS];
#pragma unroll
            for (int b = 0; b < NBAS; ++b) part[b] = 0.0;
            for_each_triplet64(S, A, B, same, wide ? INFINITY : GR.rc2, [&](int j, int k) {
                const double4 gj = S.g[j], gk = S.g[k];
                const double4 cj = S.c[j], ck = S.c[k];
                const double x3 = gk.x - gj.x, y3 = gk.y - gj.y, z3 = gk.z - gj.z;
                const double r2jk = x3 * x3 + y3 * y3 + z3 * z3;
                double fcjk = 1.0, dfcjk = 0.0;
                if (!wide) cutoff_fdf(ctype, GR, tanhPre, sqrt(r2jk), fcjk, dfcjk);
                const double costijk = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * (cj.x * ck.x);
                const double pfc = (wt ? cj.w * ck.w : 1.0) * cj.y * ck.y * fcjk;
                const double r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.0 : r2jk);
                const double E0 = exp(-gp.eta[0] * r2sum) * pfc;
                const double E1 = NB1 > 0 ? exp(-gp.eta[1] * r2sum) * pfc : 0.0;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const double pl = 1.0 + gp.lam[sl] * costijk;
                    double p = pl > 0.0 ? pl : 0.0;
#pragma unroll
                    for (int b = 0; b < NBM; ++b) {
                        if (b < NB0) part[sl * NB0 + b] += E0 * p;
                        if (b < NB1) part[2 * NB0 + sl * NB1 + b] += E1 * p;
                        p *= p;
                    }
                }
            }, [] {});
#pragma unroll
            for (int b = 0; b < NBAS; ++b)
                if ((gp.used >> b) & 1) {
                    const double v = wave_sum(part[b]);
                    if (lane == 0) S.pw[b].x = v;
                }
            __syncthreads();
            for (int q = lane; q < nmem; q += kWave) {
                const MemberT& mb = mem[q];
                Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], S.pw[MF[GR.mem0 + q].beta].x * mb.gnorm);
            }
            __syncthreads();
            continue;
        }
        double acc[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; ++q) acc[q] = 0.0;
        for_each_triplet64(S, A, B, same, wide ? INFINITY : GR.rc2, [&](int j, int k) {
            const double4 gj = S.g[j], gk = S.g[k];
            const double4 cj = S.c[j], ck = S.c[k];
            const double x3 = gk.x - gj.x, y3 = gk.y - gj.y, z3 = gk.z - gj.z;
            const double r2jk = x3 * x3 + y3 * y3 + z3 * z3;
            const double rij = gj.w, rik = gk.w;
            double rjk = 0.0, fcjk = 1.0, dfcjk = 0.0;
            if (!wide) {
                rjk = sqrt(r2jk);
                cutoff_fdf(ctype, GR, tanhPre, rjk, fcjk, dfcjk);
            }
            const double costijk = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * (cj.x * ck.x);
            const double pfc = (wt ? cj.w * ck.w : 1.0) * cj.y * ck.y * fcjk;
            const double r2sum = rij * rij + rik * rik + (wide ? 0.0 : r2jk);
            double vexp = 0.0;
#pragma unroll
            for (int q = 0; q < MAXM; ++q)
                if (q < nmem) {
                    const MemberT& mb = mem[q];
                    if (mb.calcExp) {
                        if (mb.rs > 0.0 || wt) {
                            const double a = rij - mb.rs, b = rik - mb.rs, c = rjk - mb.rs;
                            vexp = exp(-mb.eta * (a * a + b * b + (wide ? 0.0 : c * c)));
                        } else {
                            vexp = exp(-mb.eta * r2sum);
                        }
                    }
                    const double plambda = 1.0 + mb.lambda * costijk;
                    if (plambda > 0.0) {
                        const double fg =
                            vexp * (mb.useIntPow ? pow_int(plambda, mb.zetaInt - 1) : pow(plambda, mb.zeta - 1.0));
                        acc[q] += fg * plambda * pfc;
                    }
                }
        }, [] {});
#pragma unroll
        for (int q = 0; q < MAXM; ++q)
            if (q < nmem) {
                const double v = wave_sum(acc[q]);
                if (lane == 0) {
                    const MemberT& mb = mem[q];
                    Gi[mb.sf] = scale_sf(M->scale[E.sf0 + mb.sf], v * mb.gnorm);
                }
            }
        __syncthreads();
    }
}

// w[i][q] = dE/dG[i][q] times the member's derivative scale.
__global__ void k_weights64(const int* __restrict__ elem, const double* __restrict__ dEdG,
                            const double* __restrict__ fac, int ldG, size_t total, double* __restrict__ w)
{
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const size_t i = x / ldG;
        const int q = (int)(x - i * ldG);
        w[x] = dEdG[x] * fac[elem[i] * kMaxSF + q];
    }
}

__global__ void k_zero_f64(double* p, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.0;
}

}  // namespace

// Kernel instance by the model's angular shape: the power path (NB0, NB1)
// when every angular group qualifies (MixedModel::powNB), else the member
// loop unrolled to the largest group.
template <typename F>
void dispatch64(const ModelT& hm, const MixedModel& mm, F&& f)
{
    using std::integral_constant;
    using I0 = integral_constant<int, 0>;
    using I1 = integral_constant<int, 1>;
    using I3 = integral_constant<int, 3>;
    using I4 = integral_constant<int, 4>;
    using I5 = integral_constant<int, 5>;
    if (mm.powNB[0] > 0) {
        switch (mm.powNB[0] * 10 + mm.powNB[1]) {
        case 30: f(I4{}, I3{}, I0{}); return;
        case 33: f(I4{}, I3{}, I3{}); return;
        case 50: f(I4{}, I5{}, I0{}); return;
        case 51: f(I4{}, I5{}, I1{}); return;
        case 55: f(I4{}, I5{}, I5{}); return;
        default: break;
        }
    }
    const int mam = std::max(hm.max_mem_ang, 1);
    if (mam <= 4) f(integral_constant<int, 4>{}, I0{}, I0{});
    else if (mam <= 8) f(integral_constant<int, 8>{}, I0{}, I0{});
    else if (mam <= 16) f(integral_constant<int, 16>{}, I0{}, I0{});
    else f(integral_constant<int, 32>{}, I0{}, I0{});
}

void launch_values64(const ModelT* m, const ModelT& hm, const MixedModel& mm, const SysView& s, double* G, int ldG,
                     hipStream_t stream)
{
    if (s.natoms == 0) return;
    if (!G || !s.nb_n || !s.nb_j) throw Error{NNPX_ERR_STATE, "FP64 SF values: a buffer is missing"};
    const int capL = ((std::max(s.stage_cap, 1) + 31) / 32) * 32;
    const size_t lds = stage64_bytes(capL, kTrip64);
    if (lds > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the FP64 SF value kernel's LDS"};
    if (!hm.has_exp) return;
    dispatch64(hm, mm, [&](auto mx, auto n0, auto n1) {
        hipLaunchKernelGGL((k_values64<decltype(mx)::value, decltype(n0)::value, decltype(n1)::value>),
                           dim3(s.natoms), dim3(kWave), lds, stream, m, mm.d_mem, mm.d_gpow64, s, G, ldG, capL,
                           kTrip64);
    });
}

size_t fused64_lds(int cap) { return stage64_bytes(((std::max(cap, 1) + 31) / 32) * 32, kTrip64); }

void launch_force64_fused(const ModelT* m, const ModelT& hm, const MixedModel& mm, const SysView& s,
                          const double* dEdG, const double* fac, double* w, int ldG, double* F, bool push,
                          hipStream_t stream)
{
    if (s.natoms == 0) return;
    if (!dEdG || !fac || !w || !F || !s.nb_n || !s.nb_j)
        throw Error{NNPX_ERR_STATE, "fused FP64 forces: a buffer is missing"};
    const int capL = ((std::max(s.stage_cap, 1) + 31) / 32) * 32;
    const size_t lds = stage64_bytes(capL, kTrip64);
    if (lds > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the FP64 fused force kernel's LDS"};
    const size_t total = (size_t)s.natoms * ldG;
    hipLaunchKernelGGL(k_weights64, dim3((unsigned)std::min<size_t>((total + 255) / 256, 4096)), dim3(256), 0, stream,
                       s.elem, dEdG, fac, ldG, total, w);
    const size_t n3 = (size_t)3 * s.nforce;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)std::min<size_t>((n3 + 255) / 256, 4096)), dim3(256), 0, stream, F,
                       n3);
    const int ps = push ? 1 : 0;
    if (hm.has_exp)
        dispatch64(hm, mm, [&](auto mx, auto n0, auto n1) {
            hipLaunchKernelGGL((k_force64f<decltype(mx)::value, decltype(n0)::value, decltype(n1)::value>),
                               dim3(s.natoms), dim3(kWave), lds, stream, m, mm.d_mem, mm.d_gpow64, s, w, ldG, F, ps,
                               capL, kTrip64);
        });
    launch_sfcomp64_force(m, hm, s, w, ldG, F, push, stream);
}

}  // namespace nnpx
