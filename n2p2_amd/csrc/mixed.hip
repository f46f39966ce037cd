// Mixed-precision form of the energy/force path (see mixed.hpp).
//
// k_sf32     symmetry-function values in FP32, Kahan sums per lane, lanes
//            combined in FP64, FP64 scaling (SymGrpExpRad::calculate /
//            SymGrpExpAngn::calculate without derivatives,
//            SymGrpExpRad.cpp:144-239, SymGrpExpAngn.cpp:156-403); keeps
//            each atom's staging, and optionally its angular pairs, in
//            device memory for the force kernel of the same evaluation.
// k_nnbf16   NeuralNetwork::propagate (NeuralNetwork.cpp:761-896) for a
//            16-atom tile per block (four waves) as v_mfma_f32_16x16x16_bf16
//            tiles, and the adjoint dE/dG as the transposed GEMMs (replacing
//            the reference's forward-mode NeuralNetwork::calculateDEdG,
//            NeuralNetwork.cpp:396-442).
// k_force32  fused dG/dr backward: recomputes each pair/triplet term and
//            contracts it with dE/dG (Mode::calculateForces semantics,
//            Mode.cpp:1922-1984) without storing dG/dr; FP64 accumulation
//            in LDS per neighbour entry, pushed with FP64 atomics.
#include <algorithm>
#include <array>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "mixed.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kNNRows = 16;  // atoms per tile of the NN kernel
constexpr int kNNWaves = 4;  // waves per tile

// float -> bf16, round to nearest even
__host__ __device__ inline uint16_t to_bf16(float f)
{
    uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
    u = __float_as_uint(f);
#else
    std::memcpy(&u, &f, 4);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(u);
#else
    float f;
    std::memcpy(&f, &u, 4);
    return f;
#endif
}

// Cutoff function in FP32, branch-light.  CT_COS uses the hardware sine and
// cosine on x in [0, 1] (arguments in revolutions) and sin(pi x) for the
// derivative (the reference's sqrt(1 - cos^2) form cancels in single
// precision).
__device__ inline void cutoff32(int type, float rc, float rci, float iw, float rcinv, float tanhPre,
                                float r, float& fc, float& dfc)
{
    if (type == 1) {
        const float x = fminf(fmaxf((r - rci) * iw, 0.f), 1.f);
        const float c = __builtin_amdgcn_cosf(0.5f * x);
        const float s = __builtin_amdgcn_sinf(0.5f * x);
        const bool in = r < rc;
        fc = in ? 0.5f * (c + 1.f) : 0.f;
        dfc = in ? -0.5f * iw * (float)M_PI * s : 0.f;
        return;
    }
    if (r >= rc) { fc = 0.f; dfc = 0.f; return; }
    if (type == 2 || type == 3) {
        const float t = tanhf(1.f - r * rcinv);
        const float t2 = t * t;
        const float pre = type == 3 ? tanhPre : 1.f;
        fc = pre * t * t2;
        dfc = 3.f * pre * t2 * (t2 - 1.f) * rcinv;
        return;
    }
    if (type == 0) { fc = 1.f; dfc = 0.f; return; }
    if (r < rci) { fc = 1.f; dfc = 0.f; return; }
    const float x = (r - rci) * iw;
    if (type == 4) {
        const float t = 1.f / (x * x - 1.f);
        const float t2 = __expf(t);
        fc = (float)M_E * t2;
        dfc = -2.f * (float)M_E * x * t * t * t2;
    } else if (type == 5) {
        fc = (2.f * x - 3.f) * x * x + 1.f;
        dfc = x * (6.f * x - 6.f);
    } else if (type == 6) {
        const float x2 = x * x;
        fc = ((15.f - 6.f * x) * x - 10.f) * x * x2 + 1.f;
        dfc = x2 * ((60.f - 30.f * x) * x - 30.f);
    } else if (type == 7) {
        const float x2 = x * x;
        fc = (x * (x * (20.f * x - 70.f) + 84.f) - 35.f) * x2 * x2 + 1.f;
        dfc = x2 * x * (x * (x * (140.f * x - 420.f) + 420.f) - 140.f);
    } else {
        float x4 = x * x;
        x4 *= x4;
        fc = (x * (x * ((315.f - 70.f * x) * x - 540.f) + 420.f) - 126.f) * x * x4 + 1.f;
        dfc = x4 * (x * (x * ((2520.f - 630.f * x) * x - 3780.f) + 2520.f) - 630.f);
    }
    // The polynomials vanish at x = 1; in FP32 they can round to tiny
    // negative values just below it, which the log2 form of the angular
    // terms must not see.
    fc = fmaxf(fc, 0.f);
    dfc *= iw;
}

struct GroupF {
    float rc, rc2, rci, iw, rcinv;
};

__device__ inline GroupF group_f(const GroupT& g)
{
    return GroupF{(float)g.rc, (float)g.rc2, (float)g.rci, (float)g.iw, (float)g.rcinv};
}

constexpr int kRadChunk = 8;  // radial members accumulated per pass
constexpr int kRadBatch = 32;  // radial sums scaled per batch (fits the smallest triplet buffer)

// LDS staging of one atom's neighbours (FP32 geometry) and the compacted
// triplet buffer.
struct Stage {
    float4* g;       // x, y, z, |dr|
    float4* c;       // 1/|dr|, fc, dfc (cutoff of the current group), unused
    unsigned char* el;  // neighbour element
    unsigned short* la; unsigned short* lb;  // group lists
    unsigned* tjk;   // [trip + kWave] compacted pairs (j | k << 16) and up to
                     // a wave's worth past them; r_jk^2 is recomputed from g
    float4* mp;      // [mpn] per-group constants (member constants or power-path terms)
    int* eoff;       // [kMaxElem + 1] first staged slot of each neighbour element
    int trip;        // capacity of the triplet buffer
};

// A group's neighbour list: the staged slots base .. base + n - 1 (contig:
// every staged neighbour of the element lies inside the group's cutoff) or
// the n slots listed in S.la / S.lb.
struct List {
    int base, n;
    bool contig;
};

// A wave-uniform value held in a scalar register (values read from LDS or
// shuffled are uniform but not known to be by the compiler).
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Slot of list entry a (CONTIG: base + a, else l[a]).
template <bool CONTIG>
__device__ inline int slot(const unsigned short* l, int base, int a)
{
    if constexpr (CONTIG) return base + a;
    else return l[a];
}


__device__ inline Stage carve(float* base, int capL, int trip, int mpn)
{
    Stage S;
    S.trip = trip;
    S.g = reinterpret_cast<float4*>(base);
    S.c = S.g + capL;
    S.tjk = reinterpret_cast<unsigned*>(S.c + capL);
    S.mp = reinterpret_cast<float4*>(S.tjk + trip + kWave);
    S.la = reinterpret_cast<unsigned short*>(S.mp + mpn);
    S.lb = S.la + capL;
    S.el = reinterpret_cast<unsigned char*>(S.lb + capL);
    S.eoff = reinterpret_cast<int*>(S.el + capL);
    return S;
}

// |b - a|^2 of two staged positions, one fixed operation order for the
// candidate test and the terms
__device__ inline float dist2(const float4& a, const float4& b)
{
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Stages the atom's list entries closer than the model's largest cutoff
// (entries beyond it contribute nothing; lists built with a skin carry
// them), grouped by neighbour element and in list order within an element,
// so an element's neighbours are the slots eoff[e] .. eoff[e + 1] - 1.  idx,
// when given, receives each staged slot's list entry.  More such entries than
// capL: only slots below capL are staged, flag bit 4 is set (the caller
// repeats with the full capacity) and flags[2] keeps the largest count met,
// from which callers size capL.
__device__ inline int stage_atom(const ModelT* __restrict__ M, const SysView& s, int i, int capL, const Stage& S,
                                 unsigned short* idx = nullptr, const StageKeep* keep = nullptr)
{
    const int lane = threadIdx.x;
    const StructT& St = s.st[s.sid[i]];
    const int nl = s.nb_n[i];
    const int ne = M->nelem;
    const size_t base = (size_t)i * s.cap;
    const double rmax = M->maxCutoff;
    // the entry of this lane in pass b: its vector, distance and element
    // (-1: none or beyond the cutoff)
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
    // The entries of the first kKeep passes are kept in registers between
    // the counting pass and the placing pass (lists up to kKeep * 64
    // entries); later ones are recomputed.
    constexpr int kKeep = 4;
    float kx[kKeep], ky[kKeep], kz[kKeep], kd[kKeep], ki[kKeep];
    int kel[kKeep];
    // pass 1: in-cutoff entries per element (lane e counts element e)
    int cnt = 0;
    auto count = [&](int el) {
        for (int e = 0; e < ne; ++e) {
            const int c = __popcll(__ballot(el == e));
            if (lane == e) cnt += c;
        }
    };
#pragma unroll
    for (int p = 0; p < kKeep; ++p) {
        kel[p] = -1;
        kx[p] = ky[p] = kz[p] = kd[p] = ki[p] = 0.f;
        if (p * kWave < nl) {
            double dr[3], d;
            kel[p] = entry(p * kWave + lane, dr, d);
            kx[p] = (float)dr[0];
            ky[p] = (float)dr[1];
            kz[p] = (float)dr[2];
            kd[p] = (float)d;
            ki[p] = (float)(1.0 / d);
            count(kel[p]);
        }
    }
    for (int b = kKeep * kWave; b < nl; b += kWave) {
        double dr[3], d;
        count(entry(b + lane, dr, d));
    }
    const int incl = wave_incl_scan(lane < ne ? cnt : 0);
    const int total = lane_value(incl, kWave - 1);
    int next = incl - (lane < ne ? cnt : 0);  // lane e: next slot of element e
    if (lane <= ne) {
        S.eoff[lane] = min(lane < ne ? next : total, capL);
        if (keep) keep->off[(size_t)i * (kMaxElem + 1) + lane] = min(lane < ne ? next : total, capL);
    }
    // pass 2: each entry to its element's next slot
    auto place = [&](int k, int el, float x, float y, float z, float d, float id) {
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
            S.g[pos] = make_float4(x, y, z, d);
            // .w: the neighbour's atomic number, the weight of types 12 and 13
            S.c[pos] = make_float4(id, 0.f, 0.f, (float)M->Zel[el]);
            S.el[pos] = (unsigned char)el;
            if (idx) idx[pos] = (unsigned short)k;
            if (keep) {
                const size_t q = (size_t)i * keep->ld + pos;
                keep->g[q] = make_float4(x, y, z, d);
                keep->rinv[q] = id;
                keep->nbr[q] = s.nb_j[base + k];
                keep->el[q] = (unsigned char)el;
            }
        }
    };
#pragma unroll
    for (int p = 0; p < kKeep; ++p)
        if (p * kWave < nl) place(p * kWave + lane, kel[p], kx[p], ky[p], kz[p], kd[p], ki[p]);
    for (int b = kKeep * kWave; b < nl; b += kWave) {
        double dr[3], d;
        const int el = entry(b + lane, dr, d);
        place(b + lane, el, (float)dr[0], (float)dr[1], (float)dr[2], (float)d, (float)(1.0 / d));
    }
    if (lane == 0) {
        if (total > capL) atomicOr(&s.flags[0], 4);
        if (total > __hip_atomic_load(&s.flags[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&s.flags[2], total);
    }
    __syncthreads();
    return uni(min(total, capL));
}

// The staging of atom i as the SF kernel kept it (same slots, same values
// as stage_atom gives); returns the slot count.
__device__ inline int load_stage(const ModelT* __restrict__ M, const StageKeep& K, int i, const Stage& S)
{
    const int lane = threadIdx.x;
    const int ne = M->nelem;
    const int* off = K.off + (size_t)i * (kMaxElem + 1);
    if (lane <= ne) S.eoff[lane] = off[lane];
    const int n = uni(off[ne]);
    const size_t b = (size_t)i * K.ld;
    for (int k = lane; k < n; k += kWave) {
        const int el = K.el[b + k];
        S.g[k] = K.g[b + k];
        S.c[k] = make_float4(K.rinv[b + k], 0.f, 0.f, (float)M->Zel[el]);
        S.el[k] = (unsigned char)el;
    }
    __syncthreads();
    return n;
}

// Neighbours of element `elem` (all elements: elem < 0) inside the group's
// cutoff rc.  full: rc is the model's largest cutoff, so every staged
// neighbour of the element qualifies and the list is its slot range;
// otherwise the qualifying slots are listed in `out`.  Ends with a barrier.
__device__ inline List make_list(const Stage& S, int n, int elem, float rc, bool full, unsigned short* out)
{
    const int lane = threadIdx.x;
    const int lo = uni(elem < 0 ? 0 : S.eoff[elem]), hi = uni(elem < 0 ? n : S.eoff[elem + 1]);
    List L{lo, hi - lo, true};
    if (!full) {
        int cnt = 0;
        for (int b = lo; b < hi; b += kWave) {
            const int k = b + lane;
            const bool ok = k < hi && S.g[k].w < rc;
            const int incl = wave_incl_scan(ok ? 1 : 0);
            if (ok) out[cnt + incl - 1] = (unsigned short)k;
            cnt += lane_value(incl, kWave - 1);
        }
        L = List{0, uni(cnt), false};
    }
    __syncthreads();
    return L;
}

// Whether a group's cutoff covers every staged neighbour.
__device__ inline bool full_cutoff(const ModelT* __restrict__ M, const GroupT& G) { return G.rc >= M->maxCutoff; }

// Cutoff values of every staged neighbour for one radius (elements whose
// groups all share it compute them once per atom).
__device__ inline void all_cutoffs(const Stage& S, int n, int ctype, const GroupF& g, float tanhPre)
{
    for (int k = threadIdx.x; k < n; k += kWave) {
        float f = 0.f, df = 0.f;
        cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, S.g[k].w, f, df);
        S.c[k].y = f;
        S.c[k].z = df;
    }
    __syncthreads();
}

__device__ inline void group_cutoffs(const Stage& S, int n, int e1, int e2, int ctype, const GroupF& g,
                                     float tanhPre)
{
    for (int k = threadIdx.x; k < n; k += kWave) {
        float f = 0.f, df = 0.f;
        if (e1 < 0 || S.el[k] == e1 || S.el[k] == e2)
            cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, S.g[k].w, f, df);
        S.c[k].y = f;
        S.c[k].z = df;
    }
    __syncthreads();
}

__device__ inline float powi(float x, int n) { return pow_int<float>(x, n); }

// Compensated summation step (Fast2Sum error term accumulated in c; the
// terms are non-negative, so the running sum dominates after the first).
__device__ inline void kahan(float& s, float& c, float x)
{
    const float t = s + x;
    c += (s - t) + x;
    s = t;
}

// Value of a per-triplet table at a wave-uniform index.
template <int MAXS>
__device__ inline float pick(const float* v, int s)
{
    float r = v[0];
    if (MAXS > 1) r = s == 1 ? v[1] : r;
    if (MAXS > 2) r = s == 2 ? v[2] : r;
    if (MAXS > 3) r = s == 3 ? v[3] : r;
    return r;
}

template <int MAXS>
__device__ inline float pickb(const float* v, unsigned b1, unsigned b2)
{
    if (MAXS <= 2) return b1 ? v[1] : v[0];
    const float lo = b1 ? v[1] : v[0];
    const float hi = b1 ? v[3] : v[2];
    return b2 ? hi : lo;
}

// Member constants of one group: flags as wave-uniform bit masks, values
// staged in LDS (mp[q] = zeta, eta log2 e, rs; mp[32+q] = the force-pass
// weights w zeta lambda, w, 2 eta w), read as broadcasts in the triplet loops.
struct MemberMasks {
    unsigned calc, s1, s2;
};

__device__ inline MemberMasks load_members(const Stage& S, const MemberF* __restrict__ mf, int nmem,
                                           const float* w)
{
    const int lane = threadIdx.x;
    if (lane < nmem) {
        const MemberF m = mf[lane];
        const float wq = w ? w[m.sf] : 0.f;
        S.mp[lane] = make_float4(m.zeta, m.etaL, m.rs, (float)(m.sidx & 1));
        S.mp[kMaxMembersExp + lane] = make_float4(wq * m.zetaLambda, wq, wq * m.twoEta, (float)((m.sidx >> 1) & 1));
    }
    MemberMasks R;
    R.calc = (unsigned)__ballot(lane < nmem && mf[min(lane, nmem - 1)].calcExp);
    R.s1 = (unsigned)__ballot(lane < nmem && (mf[min(lane, nmem - 1)].sidx & 1));
    R.s2 = (unsigned)__ballot(lane < nmem && (mf[min(lane, nmem - 1)].sidx & 2));
    __syncthreads();
    return R;
}

__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }

// Calls work(j, k, r2jk) for every triplet (j < k in neighbour order) of the
// group's element pair with r_jk^2 < rc^2.  Pairs are enumerated without
// divergence (like elements: the folded order (a, a+d mod n), a balanced
// rectangle; unlike elements: the full product), tested, and the passing
// ones compacted through LDS so the member loops run with full waves.
struct NoFlush {
    __device__ void operator()() const {}
};

// `flush` runs after each buffer of triplets (a few per lane), so callers can
// keep short FP32 partial sums and fold them into their long accumulators.
template <bool CONTIG, typename Work, typename Flush = NoFlush>
__device__ inline void for_each_triplet(const Stage& S, const List& A, const List& B, bool same, float rc2,
                                        Work&& work, Flush&& flush = Flush{})
{
    const int lane = threadIdx.x;
    const int nA = A.n, nB = B.n;
    int full, total;
    if (same) {
        full = nA * ((nA - 1) / 2);
        total = full + ((nA % 2 == 0) ? nA / 2 : 0);
    } else {
        full = 0;
        total = nA * nB;
    }
    const unsigned short* lb = same ? S.la : S.lb;
    const int bB = same ? A.base : B.base;
    const float inv = nA > 0 ? 1.f / (float)nA : 0.f;
    int t0 = 0;
    while (t0 < total) {
        // fill the buffer with passing pairs ...
        int cnt = 0;
        while (t0 < total && cnt <= S.trip - kWave) {
            const int t = t0 + lane;
            bool ok = false;
            unsigned jk = 0;
            if (t < total) {
                const int u = same && t >= full ? t - full : t;
                int q = (int)((float)u * inv);
                int r = u - q * nA;
                if (r < 0) { --q; r += nA; }
                else if (r >= nA) { ++q; r -= nA; }
                int a, b;
                if (!same) { a = r; b = q; }
                else if (t < full) { a = r; b = r + q + 1; if (b >= nA) b -= nA; }
                else { a = u; b = u + nA / 2; }
                int j = slot<CONTIG>(S.la, A.base, a), k = slot<CONTIG>(lb, bB, b);
                if (j > k) { const int x = j; j = k; k = x; }
                ok = dist2(S.g[j], S.g[k]) < rc2;
                jk = (unsigned)j | ((unsigned)k << 16);
            }
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                S.tjk[pos] = jk;
            }
            cnt += __popcll(m);
            t0 += kWave;
        }
        __syncthreads();
        // ... then run the member loops with full waves
        for (int x = lane; x < cnt; x += kWave) work(S.tjk[x] & 0xffffu, S.tjk[x] >> 16);
        flush();
        __syncthreads();
    }
}

// Runs `work` over the first cnt buffered pairs, each lane a contiguous
// chunk, so a lane sees runs of one first atom j (j's geometry stays in
// registers over the run and the force kernel keeps j's contribution there).
template <typename Work>
__device__ inline void run_buffer(const Stage& S, int cnt, Work&& work)
{
    const int per = (cnt + kWave - 1) / kWave;
    const int x0 = threadIdx.x * per, x1 = min(cnt, x0 + per);
    for (int x = x0; x < x1; ++x) work(S.tjk[x] & 0xffffu, S.tjk[x] >> 16);
}

// The angular walk: each row a of list A pairs with its columns c -- for one
// list b = (a + c + 1) mod nA with c < (nA - 1) / 2, and for even nA one
// more column (b = a + nA / 2) on the rows a < nA / 2, so every unordered
// pair appears once; for two lists b = c, c < nB.  A step tests a block of
// rows x columns (lane = row * width + column, fixed per group), so a
// candidate's indices cost a few adds, and the row's geometry stays in
// registers.  Passing pairs are compacted through LDS in block order; the
// pairs run in buffers of exactly S.trip (the last one shorter), the pairs
// a step adds past a full buffer being carried to the next, so the buffers
// depend only on the pair sequence.  `flush` runs after each buffer.
// GATHER: gB holds the partners' geometry in list order (gB[b] =
// S.g[lb[b]]), so a candidate's index and position load together.
// src: the walk's nsrc pairs as stored by an earlier walk (same sequence,
// same buffers), read instead of tested.  dst: the pairs are stored there
// too while they fit in `room` entries.  Returns the walk's pair count.
template <bool CONTIG, bool GATHER, typename Work, typename Flush = NoFlush>
__device__ inline int for_each_triplet_blocks(const Stage& S, const float4* gB, const List& A, const List& B,
                                              bool same, float rc2, const unsigned* src, int nsrc, unsigned* dst,
                                              int room, Work&& work, Flush&& flush = Flush{})
{
    const int lane = threadIdx.x;
    const int nA = A.n, nB = B.n;
    const int R = same ? (nA - 1) / 2 : nB;          // regular columns per row
    const int Rx = same && nA % 2 == 0 ? R + 1 : R;  // with the half-way column (rows a < nA / 2 only)
    const int half = nA / 2;
    // a row's columns split into nearly equal blocks of at most 8: the
    // blocks fill the wave's lanes closely, and the passing partners of one
    // row still form runs in the buffer
    const int nblk = (Rx + 7) >> 3;
    const int cw = nblk > 0 ? (Rx + nblk - 1) / nblk : 1, rh = kWave / cw;
    const int lr = lane / cw, lc = lane - lr * cw;  // lanes with lr >= rh stay idle
    const bool live = lr < rh;
    const unsigned short* lb = same ? S.la : S.lb;
    const int bB = same ? A.base : B.base;
    // contiguous partners are already in list order
    const float4* gp = CONTIG ? S.g + bB : gB;
    // no candidates: the fill loop never runs
    int a0 = src || nA == 0 || Rx == 0 ? nA : 0, c0 = 0;
    // the lane's row atom and its geometry, reloaded when the block moves down
    int ja = slot<CONTIG>(S.la, A.base, lr < min(nA, rh) ? lr : 0);
    float4 ga = a0 < nA ? S.g[ja] : make_float4(0.f, 0.f, 0.f, 0.f);
    int cnt = 0, done = 0;  // pairs in the buffer, pairs run so far
    // stored pairs are loaded a buffer ahead (buffers of up to kAhead per
    // lane), so the loads overlap the previous buffer's work
    constexpr int kAhead = 6;
    unsigned ahead[kAhead];
    const bool pre = src && S.trip <= kAhead * kWave;
    auto load_ahead = [&](int from) {
        const int m = min(S.trip, nsrc - from);
#pragma unroll
        for (int q = 0; q < kAhead; ++q) {
            const int x = q * kWave + lane;
            ahead[q] = x < m ? src[from + x] : 0u;
        }
    };
    if (pre) load_ahead(0);
    for (;;) {
        if (src) {
            cnt = min(S.trip, nsrc - done);
            if (pre) {
#pragma unroll
                for (int q = 0; q < kAhead; ++q)
                    if (q * kWave + lane < cnt) S.tjk[q * kWave + lane] = ahead[q];
                if (done + cnt < nsrc) load_ahead(done + cnt);
            } else {
                for (int x = lane; x < cnt; x += kWave) S.tjk[x] = src[done + x];
            }
        } else {
            while (a0 < nA && cnt < S.trip) {
                // branch-free: lanes outside the block test column 0 and fail
                const int a = a0 + lr, c = c0 + lc;
                const bool valid = live && a < nA && c < Rx && (c < R || a < half);
                int b = same ? a + c + 1 : c;
                b = same && b >= nA ? b - nA : b;
                b = valid ? b : 0;
                const int k = slot<CONTIG>(lb, bB, b);
                const float d2 = dist2(ga, (GATHER || CONTIG) ? gp[b] : S.g[k]);  // in bounds for every lane
                const bool ok = valid & (d2 < rc2);
                const unsigned long long m = __ballot(ok);
                // passing pairs to the next entries (at most a wave past the buffer)
                if (ok)
                    S.tjk[cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] =
                        (unsigned)ja | ((unsigned)k << 16);
                cnt += __popcll(m);
                c0 += cw;
                if (c0 >= Rx) {
                    c0 = 0;
                    a0 += rh;
                    const int an = a0 + lr;
                    ja = slot<CONTIG>(S.la, A.base, an < nA ? an : 0);
                    ga = S.g[ja];
                }
            }
        }
        const int m = min(cnt, S.trip);
        if (m <= 0) break;
        __syncthreads();
        if (dst && done + m <= room)
            for (int x = lane; x < m; x += kWave) dst[done + x] = S.tjk[x];
        run_buffer(S, m, work);
        flush();
        done += m;
        __syncthreads();
        // pairs past the buffer move to its front (disjoint ranges)
        const int rest = cnt - m;
        if (lane < rest) S.tjk[lane] = S.tjk[S.trip + lane];
        cnt = rest;
        __syncthreads();
    }
    return done;
}

// Row-major walk (the force kernel without pair lists): candidates are
// enumerated a-major from a linear index (all partners of la[a]
// consecutively), compacted the same way, and each lane takes a contiguous
// chunk of the compacted buffer.  Its partner atoms k differ more across the
// lanes of one step than in the block walk, which keeps the force kernel's
// LDS atomics on k from colliding; the block walk, whose lists the SF kernel
// can keep, tests candidates more cheaply.  `flush` runs after each buffer.
template <bool CONTIG, typename Work, typename Flush = NoFlush>
__device__ inline void for_each_triplet_rows(const Stage& S, const List& A, const List& B, bool same, float rc2,
                                             Work&& work, Flush&& flush = Flush{})
{
    const int lane = threadIdx.x;
    const int nA = A.n, nB = B.n;
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
    if (total == 0) return;
    const unsigned short* lb = same ? S.la : S.lb;
    const int bB = same ? A.base : B.base;
    // the lane's linear index t = q R + r, advanced by kWave = dq R + dr per
    // step (decoded once per buffer, so only t0 lives across the work)
    const int dq = R > 0 ? kWave / R : 0, dr = R > 0 ? kWave - dq * R : kWave;
    int t0 = 0;
    while (t0 < total) {
        int cnt = 0;
        int t = t0 + lane, q = 0, r = t;
        if (R > 0) {
            q = t / R;
            r = t - q * R;
        }
        while (t0 < total && cnt <= S.trip - kWave) {
            // branch-free: rows q (t < full) or the half-way column of row
            // t - full; lanes past the end test slot 0 and fail
            const bool valid = t < total, tail = t >= full;
            int b = same ? q + r + 1 : r;
            b = same && b >= nA ? b - nA : b;
            const int a = valid ? (tail ? t - full : q) : 0;
            b = valid ? (tail ? a + (nA >> 1) : b) : 0;
            const int j = slot<CONTIG>(S.la, A.base, a), k = slot<CONTIG>(lb, bB, b);
            const float d2 = dist2(S.g[j], S.g[k]);  // in bounds for every lane
            const bool ok = valid & (d2 < rc2);
            const unsigned long long m = __ballot(ok);
            if (ok)
                S.tjk[cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] =
                    (unsigned)j | ((unsigned)k << 16);
            cnt += __popcll(m);
            t0 += kWave;
            t += kWave;
            q += dq;
            r += dr;
            if (r >= R) { r -= R; ++q; }
        }
        __syncthreads();
        run_buffer(S, cnt, work);
        flush();
        __syncthreads();
    }
}

// Copies the partner list's geometry into list order for the block walk
// (listed partners only: a slot range is in list order already).
__device__ inline void gather_partners(const Stage& S, const List& A, const List& B, bool same, float4* gB)
{
    if (A.contig) return;
    const unsigned short* lb = same ? S.la : S.lb;
    const int nB = same ? A.n : B.n;
    for (int b = threadIdx.x; b < nB; b += kWave) gB[b] = S.g[lb[b]];
    __syncthreads();
}

template <typename Work, typename Flush = NoFlush>
__device__ inline void walk_rows(const Stage& S, const List& A, const List& B, bool same, float rc2, Work&& work,
                                 Flush&& flush = Flush{})
{
    if (A.contig) for_each_triplet_rows<true>(S, A, B, same, rc2, work, flush);
    else for_each_triplet_rows<false>(S, A, B, same, rc2, work, flush);
}

// Per-atom cursor over the stored pair lists of its walks.
struct TripCursor {
    unsigned* p;   // the next walk's pairs
    int room;      // entries left for them (0: none kept)
    int* counts;   // this atom's counts (nullptr: no list)
    int ld;        // walks with a count
    int w;         // walks so far
    int total;     // pairs of all walks so far
    bool rows;     // force kernel: the row-major walk (no lists)
};

__device__ inline TripCursor trip_cursor(const TripList& T, int i)
{
    TripCursor c{nullptr, 0, nullptr, 0, 0, 0, T.rows};
    if (T.list) {
        c.p = T.list + (size_t)i * T.cap;
        c.room = T.cap;
        c.counts = T.counts + (size_t)i * T.ld;
        c.ld = T.ld;
    }
    return c;
}

// The walk for slot-range lists and for listed slots.  Storing side (the SF
// kernel, store): the pairs go to the cursor's list while they fit and the
// walk's count is recorded.  Reading side (the force kernel, read): a walk
// with a recorded count reads its pairs back instead of testing candidates.
template <bool GATHER, typename Work, typename Flush = NoFlush>
__device__ inline void walk_blocks(const Stage& S, const float4* gB, const List& A, const List& B, bool same,
                                   float rc2, TripCursor& tc, bool store, Work&& work, Flush&& flush = Flush{})
{
    const int w = tc.w++;
    const bool listed = tc.counts && w < tc.ld;
    const unsigned* src = nullptr;
    int nsrc = 0;
    if (listed && !store) {
        nsrc = uni(tc.counts[w]);
        if (nsrc >= 0) {
            src = tc.p;
            tc.p += nsrc;
        }
    }
    unsigned* dst = listed && store ? tc.p : nullptr;
    const int room = dst ? tc.room : 0;
    const int n = A.contig ? for_each_triplet_blocks<true, GATHER>(S, gB, A, B, same, rc2, src, nsrc, dst, room, work, flush)
                           : for_each_triplet_blocks<false, GATHER>(S, gB, A, B, same, rc2, src, nsrc, dst, room, work, flush);
    tc.total += n;
    if (dst) {
        const bool kept = n <= room;
        if (threadIdx.x == 0) tc.counts[w] = kept ? n : -1;
        if (kept) {
            tc.p += n;
            tc.room -= n;
        }
    }
}

template <typename Work, typename Flush = NoFlush>
__device__ inline void walk_all(const Stage& S, const List& A, const List& B, bool same, float rc2, Work&& work,
                                Flush&& flush = Flush{})
{
    if (A.contig) for_each_triplet<true>(S, A, B, same, rc2, work, flush);
    else for_each_triplet<false>(S, A, B, same, rc2, work, flush);
}

// ------------------------------------------------------------ power path --
// Angular groups whose members are exp(-eta_h sum r^2) (1 + lambda_s cos)^zeta
// with zeta = 2^b (b <= 4), at most two distinct eta and two lambda, rs = 0:
// per triplet one exponential per eta and the powers p, p^2, p^4, ... of the
// two bases by repeated squaring (the reference's integer-zeta pow_int,
// utility.h, also multiplies) stand in for one exp2/log2 per member.  Terms
// are indexed beta = s NB0 + b (first eta) and 2 NB0 + s NB1 + b (second).

// SF values of one power-path group: FP32 partial sums per term over a
// triplet buffer folded into compensated sums, one FP64 wave reduction per
// used term, then each member takes its term (times 2^(1-zeta)) and is scaled.
template <int NB0, int NB1, bool GATHER>
__device__ inline void sf32_power_group(const ModelT* __restrict__ M, const Stage& S, float4* gB,
                                        const MemberF* __restrict__ mf, const GroupMixF& gm, const GroupT& GT,
                                        const GroupF& g, int n, int ctype, float tanhPre, int sf0,
                                        double* __restrict__ Gi, float* __restrict__ Ii, TripCursor& tc)
{
    constexpr int NBAS = 2 * (NB0 + NB1);
    constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
    const int lane = threadIdx.x;
    const bool same = GT.e1 == GT.e2;
    const bool wide = GT.type == 9;  // no r_jk cutoff or exponent term (SymGrpExpAngw.cpp:156-380)
    const bool wt = GT.type == 13;
    const bool full = full_cutoff(M, GT);
    const List A = make_list(S, n, GT.e1, g.rc, full, S.la);
    const List B = same ? A : make_list(S, n, GT.e2, g.rc, full, S.lb);
    if constexpr (GATHER) gather_partners(S, A, B, same, gB);
    float acc[NBAS], cmp[NBAS], part[NBAS];
#pragma unroll
    for (int b = 0; b < NBAS; ++b) { acc[b] = 0.f; cmp[b] = 0.f; part[b] = 0.f; }
    const float ce0 = -gm.etaL[0], ce1 = -gm.etaL[1];
    const float lam0 = gm.lam[0], lam1 = gm.lam[1];
    int jc = -1;  // j of the cached geometry (runs of one j per lane)
    float4 gj = make_float4(0.f, 0.f, 0.f, 0.f), cj = gj;
    walk_blocks<GATHER>(S, gB, A, B, same, wide ? INFINITY : g.rc2, tc, true, [&](int j, int k) {
        if (j != jc) {
            gj = S.g[j];
            cj = S.c[j];
            jc = j;
        }
        const float4 gk = S.g[k];
        const float4 ck = S.c[k];
        const float r2jk = dist2(gj, gk);
        float fcjk = 1.f, dfcjk = 0.f;
        if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, __builtin_amdgcn_sqrtf(r2jk), fcjk, dfcjk);
        const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * cj.x * ck.x;
        // type 13: times Z_j Z_k (SymGrpExpAngnWeighted.cpp)
        const float pfc = cj.y * ck.y * fcjk * (wt ? cj.w * ck.w : 1.f);
        const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
        const float E0 = fexp2(ce0 * r2sum) * pfc;
        const float E1 = NB1 > 0 ? fexp2(ce1 * r2sum) * pfc : 0.f;
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            const float pl = fmaf(sl ? lam1 : lam0, cost, 1.f);
            float p = pl > 0.f ? pl : 0.f;  // the reference's zero term for 1 + lambda cos <= 0
#pragma unroll
            for (int b = 0; b < NBM; ++b) {
                if (b < NB0) part[sl * NB0 + b] = fmaf(E0, p, part[sl * NB0 + b]);
                if (b < NB1) part[2 * NB0 + sl * NB1 + b] = fmaf(E1, p, part[2 * NB0 + sl * NB1 + b]);
                p *= p;
            }
        }
    }, [&]() {
#pragma unroll
        for (int b = 0; b < NBAS; ++b) {
            kahan(acc[b], cmp[b], part[b]);
            part[b] = 0.f;
        }
    });
    double* red = reinterpret_cast<double*>(S.mp);
    double v[NBAS];
#pragma unroll
    for (int b = 0; b < NBAS; ++b) v[b] = (gm.used >> b) & 1 ? wave_sum((double)acc[b] + (double)cmp[b]) : 0.0;
    if (lane == 0) {
#pragma unroll
        for (int b = 0; b < NBAS; ++b) red[b] = v[b];
    }
    __syncthreads();
    if (lane < GT.nmem) {
        const MemberT& mt = M->mem[GT.mem0 + lane];
        const double gs = scale_sf(M->scale[sf0 + mt.sf], red[mf[lane].beta] * mt.gnorm);
        Gi[mt.sf] = gs;
        Ii[mt.sf] = (float)gs;
    }
    __syncthreads();
}

// Force weights of one power-path group: wb[beta] = sum of the members'
// dE/dG weights w on term beta, wa[beta] = sum of zeta lambda w.
template <int NBAS>
__device__ inline void power_weights(const Stage& S, const MemberF* __restrict__ mf, int nmem,
                                     const float* __restrict__ w, float* wb, float* wa)
{
    const int lane = threadIdx.x;
    // one member per lane (nmem <= kMaxMembersExp < kWave), so its constants
    // and its weight are loaded once, all members at the same time
    int beta = -1;
    float wq = 0.f, wz = 0.f;
    if (lane < nmem) {
        const MemberF m = mf[lane];
        beta = m.beta;
        wq = w[m.sf];
        wz = wq * m.zetaLambda;
    }
    // lane b sums the weights of term b in member order
    float sb = 0.f, sa = 0.f;
    for (int q = 0; q < nmem; ++q) {
        const int bq = lane_value(beta, q);
        const float x = lane_value(wq, q), y = lane_value(wz, q);
        if (lane == bq) {
            sb += x;
            sa += y;
        }
    }
    if (lane < NBAS) S.mp[lane] = make_float4(sb, sa, 0.f, 0.f);
    __syncthreads();
    // uniform: held in scalar registers
#pragma unroll
    for (int b = 0; b < NBAS; ++b) {
        const float4 v = S.mp[b];
        wb[b] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v.x)));
        wa[b] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v.y)));
    }
}

// LDS of a merged walk: the pair map, then per (merged group, term) one
// 8-byte entry (the FP64 term sum, or the FP32 weights wb, wa).
__host__ __device__ constexpr size_t merge_bytes(int ng, int nbas)
{
    return (size_t)kMaxElem * kMaxElem + (size_t)ng * nbas * 8;
}

// SF values of all power-path angular groups of an element in one triplet
// walk over the element-ordered in-cutoff list: FP32 partial sums per term
// over a lane's run of one group, folded into FP64 LDS sums per (group,
// term) when the group changes and after each buffer; then every member
// takes its group's term (times 2^(1-zeta)) and is scaled.
template <int NB0, int NB1, bool GATHER>
__device__ inline void sf32_power_merged(const ModelT* __restrict__ M, const Stage& S, float4* gB,
                                         const MemberF* __restrict__ MF, const GroupMixF* __restrict__ GM,
                                         const AngMergeF& am, const ElementT& E, int n, int ctype, float tanhPre,
                                         char* mbase, double* __restrict__ Gi, float* __restrict__ Ii,
                                         TripCursor& tc)
{
    constexpr int NBAS = 2 * (NB0 + NB1);
    constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
    const int lane = threadIdx.x;
    const GroupT& G0 = M->grp[am.grp[0]];
    const GroupF g = group_f(G0);
    const GroupMixF gm = GM[am.grp[0]];
    const bool wide = G0.type == 9;
    const int ng = am.ng, used = am.used;
    signed char* pmap = reinterpret_cast<signed char*>(mbase);
    double* red = reinterpret_cast<double*>(mbase + kMaxElem * kMaxElem);
    for (int x = lane; x < kMaxElem * kMaxElem; x += kWave) pmap[x] = am.pair[x];
    for (int x = lane; x < ng * NBAS; x += kWave) red[x] = 0.0;
    // all elements' in-cutoff neighbours, element by element (staging order)
    const List A = make_list(S, n, -1, g.rc, full_cutoff(M, G0), S.la);  // its barrier orders the fills above
    if constexpr (GATHER) gather_partners(S, A, A, true, gB);
    float part[NBAS];
#pragma unroll
    for (int b = 0; b < NBAS; ++b) part[b] = 0.f;
    int gc = -1;  // group of the partial sums
    auto flush = [&]() {
        if (gc >= 0) {
#pragma unroll
            for (int b = 0; b < NBAS; ++b)
                if ((used >> b) & 1) lds_add(red + gc * NBAS + b, (double)part[b]);
        }
#pragma unroll
        for (int b = 0; b < NBAS; ++b) part[b] = 0.f;
        gc = -1;
    };
    const float ce0 = -gm.etaL[0], ce1 = -gm.etaL[1];
    const float lam0 = gm.lam[0], lam1 = gm.lam[1];
    int jc = -1, rowj = 0;
    float4 gj = make_float4(0.f, 0.f, 0.f, 0.f), cj = gj;
    walk_blocks<GATHER>(S, gB, A, A, true, wide ? INFINITY : g.rc2, tc, true, [&](int j, int k) {
        if (j != jc) {
            gj = S.g[j];
            cj = S.c[j];
            rowj = S.el[j] * kMaxElem;
            jc = j;
        }
        const int q = pmap[rowj + S.el[k]];
        if (q != gc) {
            flush();
            gc = q;
        }
        if (q < 0) return;
        const float4 gk = S.g[k];
        const float4 ck = S.c[k];
        const float r2jk = dist2(gj, gk);
        float fcjk = 1.f, dfcjk = 0.f;
        if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, __builtin_amdgcn_sqrtf(r2jk), fcjk, dfcjk);
        const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * cj.x * ck.x;
        const float pfc = cj.y * ck.y * fcjk;
        const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
        const float E0 = fexp2(ce0 * r2sum) * pfc;
        const float E1 = NB1 > 0 ? fexp2(ce1 * r2sum) * pfc : 0.f;
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            const float pl = fmaf(sl ? lam1 : lam0, cost, 1.f);
            float p = pl > 0.f ? pl : 0.f;
#pragma unroll
            for (int b = 0; b < NBM; ++b) {
                if (b < NB0) part[sl * NB0 + b] = fmaf(E0, p, part[sl * NB0 + b]);
                if (b < NB1) part[2 * NB0 + sl * NB1 + b] = fmaf(E1, p, part[2 * NB0 + sl * NB1 + b]);
                p *= p;
            }
        }
    }, flush);
    __syncthreads();
    for (int q = 0; q < ng; ++q) {
        const GroupT& GT = M->grp[am.grp[q]];
        for (int x = lane; x < GT.nmem; x += kWave) {
            const MemberT& mt = M->mem[GT.mem0 + x];
            const double gs = scale_sf(M->scale[E.sf0 + mt.sf], red[q * NBAS + MF[GT.mem0 + x].beta] * mt.gnorm);
            Gi[mt.sf] = gs;
            Ii[mt.sf] = (float)gs;
        }
    }
    __syncthreads();
}

// SF values of an element's radial groups when all its groups share one
// cutoff (staged cutoff values in S.c): all groups at once, each member on
// L = 64 / members lanes (at least one) that sum every L-th neighbour of its
// group in FP32 with compensation; the L partial sums are added in FP64 in
// lane order through LDS (the idle triplet buffer), and the member's first
// lane scales and stores the value.
__device__ inline void radial_lanes(const ModelT* __restrict__ M, const Stage& S, const MemberF* __restrict__ MF,
                                    const ElementT& E, int n, double* __restrict__ Gi, float* __restrict__ Ii)
{
    const int lane = threadIdx.x;
    int nrad = 0;
    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& GT = M->grp[E.grp0 + gi];
        if (GT.type == 2 || GT.type == 12) nrad += GT.nmem;
    }
    const int L = nrad > 0 && nrad < kWave ? kWave / nrad : 1;  // lanes per member
    const int sub = lane % L;
    double* part = reinterpret_cast<double*>(S.tjk);  // kWave doubles fit the smallest buffer
    for (int base = 0; base < nrad; base += kWave / L) {
        // the lane's member: the (base + lane / L)-th radial member of the element
        int rem = base + lane / L, gsel = -1;
        if (lane / L < kWave / L)
            for (int gi = 0; gi < E.ngrp && gsel < 0; ++gi) {
                const GroupT& GT = M->grp[E.grp0 + gi];
                if (GT.type != 2 && GT.type != 12) continue;
                if (rem < GT.nmem) gsel = gi;
                else rem -= GT.nmem;
            }
        float acc = 0.f, cmp = 0.f;
        if (gsel >= 0) {
            const GroupT& GT = M->grp[E.grp0 + gsel];
            const MemberF mq = MF[GT.mem0 + rem];
            // type 2: the slots of neighbour element e1; type 12: every slot,
            // each term times its atomic number (SymGrpExpRadWeighted.cpp)
            const bool wt = GT.type == 12;
            const float rc = (float)GT.rc;
            const int lo = wt ? 0 : S.eoff[GT.e1], hi = wt ? n : S.eoff[GT.e1 + 1];
            for (int k = lo + sub; k < hi; k += L) {
                const float r = S.g[k].w;
                if (!(r < rc)) continue;
                const float4 ck = S.c[k];
                const float dr = r - mq.rs;
                kahan(acc, cmp, __expf(-mq.eta * dr * dr) * (wt ? ck.y * ck.w : ck.y));
            }
        }
        double v = (double)acc + (double)cmp;
        if (L > 1) {
            part[lane] = v;
            __syncthreads();
            if (sub == 0 && gsel >= 0)
                for (int t = 1; t < L; ++t) v += part[lane + t];
            __syncthreads();
        }
        if (gsel >= 0 && sub == 0) {
            const GroupT& GT = M->grp[E.grp0 + gsel];
            const MemberT& mt = M->mem[GT.mem0 + rem];
            const double gs = scale_sf(M->scale[E.sf0 + mt.sf], v);
            Gi[mt.sf] = gs;
            Ii[mt.sf] = (float)gs;
        }
    }
}

// ------------------------------------------------------------- SF values --
// Angular terms: vexp * pfc * pl^zeta evaluated as one exp2 of
// zeta * log2(pl) + log2(pfc) - eta log2(e) * sum(r - rs)^2, with log2(pl)
// shared by all members with the same lambda (pl <= 0 gives log2 = -inf, the
// reference's zero term).
// FAST: every angular group has exactly MAXM members, at most two distinct
// lambda and rs = 0 -- the member loop is then branch-free.  Otherwise the
// general loop handles shifted Gaussians and up to four lambda values.
// MERGE: the element's angular groups go through one merged walk when it
// has one (AngMergeF::ng > 0).
template <int MAXM, bool FAST, int NB0, int NB1, bool GATHER, bool MERGE>
__global__ __launch_bounds__(64) void k_sf32(const ModelT* __restrict__ M, const MemberF* __restrict__ MF,
                                             const GroupMixF* __restrict__ GM, const AngMergeF* __restrict__ AM,
                                             SysView s, double* __restrict__ G, float* __restrict__ Gin, int ldG,
                                             int ldK, int capL, int trip, TripList T, StageKeep K)
{
    extern __shared__ float sm[];
    const Stage S = carve(sm, capL, trip, stage_mp(NB0, NB1));
    float4* gB = reinterpret_cast<float4*>(reinterpret_cast<char*>(sm) + stage_bytes(capL, trip, stage_mp(NB0, NB1)));
    char* mbase = reinterpret_cast<char*>(gB) + (GATHER ? sizeof(float4) * (size_t)capL : 0);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const float tanhPre = (float)M->tanhPre;
    const int n = stage_atom(M, s, i, capL, S, nullptr, K.g ? &K : nullptr);
    TripCursor tc = trip_cursor(T, i);
    double* Gi = G + (size_t)i * ldG;
    float* Ii = Gin + (size_t)i * ldK;
    for (int q = E.nsf + lane; q < ldK; q += kWave) Ii[q] = 0.f;
    const bool merged = MERGE && AM[e].ng > 0;
    // one cutoff for all of the element's groups: the cutoff values once,
    // and every radial member at once, one lane each
    if (E.ucut && E.ngrp > 0) {
        all_cutoffs(S, n, ctype, group_f(M->grp[E.grp0]), tanhPre);
        radial_lanes(M, S, MF, E, n, Gi, Ii);
    }

    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& GT = M->grp[E.grp0 + gi];
        const GroupF g = group_f(GT);
        const int nmem = GT.nmem;
        const MemberF* mf = MF + GT.mem0;
        if (!E.ucut) group_cutoffs(S, n, GT.e1, GT.type == 2 ? GT.e1 : GT.e2, ctype, g, tanhPre);
        if (sf_compact(GT.type)) continue;  // the compact kernel's
        if (merged && GT.type != 2 && GT.type != 12) continue;  // the merged walk's
        if (GT.type == 2 || GT.type == 12) {
            if (E.ucut) continue;  // radial_lanes did it
            // radial members in chunks of kRadChunk (any member count);
            // type 12: every neighbour, each term times its atomic number
            // (SymGrpExpRadWeighted.cpp)
            const bool wt = GT.type == 12;
            // the wave sums go through LDS (the triplet buffer, >= kRadBatch
            // doubles) so a batch of members is scaled and stored by one lane
            // each, not one after another by lane 0
            double* red = reinterpret_cast<double*>(S.tjk);
            for (int q0 = 0; q0 < nmem; q0 += kRadChunk) {
                float acc[kRadChunk], cmp[kRadChunk];
#pragma unroll
                for (int q = 0; q < kRadChunk; ++q) { acc[q] = 0.f; cmp[q] = 0.f; }
                // type 2: the slots of neighbour element e1; type 12: every slot
                const int lo = wt ? 0 : S.eoff[GT.e1], hi = wt ? n : S.eoff[GT.e1 + 1];
                for (int k = lo + lane; k < hi; k += kWave) {
                    const float r = S.g[k].w;
                    if (!(r < g.rc)) continue;
                    const float fc = wt ? S.c[k].y * S.c[k].w : S.c[k].y;
#pragma unroll
                    for (int q = 0; q < kRadChunk; ++q)
                        if (q0 + q < nmem) {
                            const float dr = r - mf[q0 + q].rs;
                            kahan(acc[q], cmp[q], __expf(-mf[q0 + q].eta * dr * dr) * fc);
                        }
                }
                double v[kRadChunk];
#pragma unroll
                for (int q = 0; q < kRadChunk; ++q)
                    v[q] = q0 + q < nmem ? wave_sum((double)acc[q] + (double)cmp[q]) : 0.0;
                if (lane == 0) {
#pragma unroll
                    for (int q = 0; q < kRadChunk; ++q) red[(q0 + q) % kRadBatch] = v[q];
                }
                const int q1 = q0 + kRadChunk;
                if (q1 % kRadBatch == 0 || q1 >= nmem) {
                    __syncthreads();
                    const int qm = q0 - q0 % kRadBatch + lane;  // this batch's members
                    if (lane < kRadBatch && qm < min(q1, nmem)) {
                        const MemberT& mt = M->mem[GT.mem0 + qm];
                        const double gs = scale_sf(M->scale[E.sf0 + mt.sf], red[qm % kRadBatch]);
                        Gi[mt.sf] = gs;
                        Ii[mt.sf] = (float)gs;
                    }
                    __syncthreads();
                }
            }
            continue;
        }
        if constexpr (NB0 > 0) {
            sf32_power_group<NB0, NB1, GATHER>(M, S, gB, mf, GM[E.grp0 + gi], GT, g, n, ctype, tanhPre, E.sf0, Gi, Ii,
                                               tc);
            continue;
        }
        float acc[MAXM], cmp[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; ++q) { acc[q] = 0.f; cmp[q] = 0.f; }
        {
            const bool same = GT.e1 == GT.e2;
            const bool wide = GT.type == 9;  // no r_jk cutoff or exponent term (SymGrpExpAngw.cpp:156-380)
            const bool wt = GT.type == 13;
            const bool full = full_cutoff(M, GT);
            const List A = make_list(S, n, GT.e1, g.rc, full, S.la);
            const List B = same ? A : make_list(S, n, GT.e2, g.rc, full, S.lb);
            if constexpr (GATHER) gather_partners(S, A, B, same, gB);
            const GroupMixF gm = GM[E.grp0 + gi];
            const MemberMasks R = load_members(S, mf, nmem, nullptr);
            // FAST: zeta and -eta log2 e in registers for the whole group;
            // terms summed in FP32 over one triplet buffer (a few per lane),
            // then folded into the compensated sums
            constexpr int NC = FAST ? MAXM : 1;
            float cz[NC], ce[NC], part[NC];
            if constexpr (FAST) {
#pragma unroll
                for (int q = 0; q < MAXM; ++q) {
                    cz[q] = S.mp[q].x;
                    ce[q] = -S.mp[q].y;
                    part[q] = 0.f;
                }
            }
            int jc = -1;  // j of the cached geometry (runs of one j per lane)
            float4 gj = make_float4(0.f, 0.f, 0.f, 0.f), cj = gj;
            walk_blocks<GATHER>(S, gB, A, B, same, wide ? INFINITY : g.rc2, tc, true, [&](int j, int k) {
                if (j != jc) {
                    gj = S.g[j];
                    cj = S.c[j];
                    jc = j;
                }
                const float4 gk = S.g[k];
                const float4 ck = S.c[k];
                const float r2jk = dist2(gj, gk);
                const float rjk = __builtin_amdgcn_sqrtf(r2jk);
                float fcjk, dfcjk;
                fcjk = 1.f;
                dfcjk = 0.f;
                if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, rjk, fcjk, dfcjk);
                const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * cj.x * ck.x;
                const float lpfc = flog2(cj.y * ck.y * fcjk * (wt ? cj.w * ck.w : 1.f));  // type 13: Z_j Z_k
                const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
                constexpr int MAXS = FAST ? 2 : 4;
                float L[MAXS];
#pragma unroll
                for (int t = 0; t < MAXS; ++t) {
                    const float pl = 1.f + gm.lam[t] * cost;
                    L[t] = pl > 0.f ? flog2(pl) : -INFINITY;
                }
                if constexpr (FAST) {
#pragma unroll
                    for (int q = 0; q < MAXM; ++q) {
                        const bool s1 = (R.s1 >> q) & 1u;
                        const float ec = fmaf(ce[q], r2sum, lpfc);
                        part[q] += fexp2(fmaf(cz[q], s1 ? L[1] : L[0], ec));
                    }
                } else {
                    float ec = 0.f;
#pragma unroll
                    for (int q = 0; q < MAXM; ++q)
                        if (q < nmem) {
                            const float4 mq = S.mp[q];
                            const float4 mw = S.mp[kMaxMembersExp + q];
                            if ((R.calc >> q) & 1u) {
                                float arg = r2sum;
                                if (mq.z > 0.f) {
                                    const float a = gj.w - mq.z, b = gk.w - mq.z, c = rjk - mq.z;
                                    arg = a * a + b * b + (wide ? 0.f : c * c);
                                }
                                ec = fmaf(-mq.y, arg, lpfc);
                            }
                            const float lo = mq.w != 0.f ? L[1] : L[0];
                            const float hi = mq.w != 0.f ? L[3] : L[2];
                            const float Ls = mw.w != 0.f ? hi : lo;
                            kahan(acc[q], cmp[q], fexp2(fmaf(mq.x, Ls, ec)));
                        }
                }
            }, [&]() {
                if constexpr (FAST) {
#pragma unroll
                    for (int q = 0; q < MAXM; ++q) {
                        kahan(acc[q], cmp[q], part[q]);
                        part[q] = 0.f;
                    }
                }
            });
        }
        // wave sums through LDS, then one lane per member scales and stores
        double* red = reinterpret_cast<double*>(S.tjk);
        static_assert(MAXM <= 32, "member sums must fit the smallest triplet buffer");
        double v[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; ++q) v[q] = q < nmem ? wave_sum((double)acc[q] + (double)cmp[q]) : 0.0;
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < MAXM; ++q) red[q] = v[q];
        }
        __syncthreads();
        if (lane < nmem) {
            const MemberT& mt = M->mem[GT.mem0 + lane];
            const double gs = scale_sf(M->scale[E.sf0 + mt.sf], red[lane] * mt.gnorm);
            Gi[mt.sf] = gs;
            Ii[mt.sf] = (float)gs;
        }
        __syncthreads();
    }
    if constexpr (MERGE && NB0 > 0) {
        if (merged)
            sf32_power_merged<NB0, NB1, GATHER>(M, S, gB, MF, GM, AM[e], E, n, ctype, tanhPre, mbase, Gi, Ii, tc);
    }
    // the largest pair count per atom sizes the lists
    if (T.count && lane == 0 &&
        tc.total > __hip_atomic_load(&s.flags[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&s.flags[3], tc.total);
}

// ------------------------------------------------------------ BF16 MFMA NN --
struct NNMix {
    const uint16_t* w;
    const float* b;
    const float* sffac;
    const int* atoms;
    int blk_off[kMaxElem + 1];
    int atom_off[kMaxElem + 1];
    int wW[kMaxElem][kMaxLayers], wT[kMaxElem][kMaxLayers], bo[kMaxElem][kMaxLayers];
    int pin[kMaxElem][kMaxLayers], pout[kMaxElem][kMaxLayers];
    int nelem, ldK, ldG, terms;
    int stride_in;  // LDS row length of the input tile (padded input width + 4)
    int stride;     // ... of the hidden/output tiles (padded layer width + 4)
    double* zero;   // cleared by the blocks in a grid-stride loop (the force
    size_t nzero;   // array of the force kernel that follows), nzero doubles
};

__device__ inline float act32(int a, float x, float& d)
{
    float v;
    switch (a) {
    case AF_IDENTITY: d = 1.f; return x;
    case AF_TANH: v = tanhf(x); d = 1.f - v * v; return v;
    case AF_LOGISTIC:
        if (x > 35.f) { d = 0.f; return 1.f; }
        if (x < -35.f) { d = 0.f; return 0.f; }
        v = 1.f / (1.f + __expf(-x)); d = v * (1.f - v); return v;
    case AF_SOFTPLUS:
        if (x > 35.f) { d = 1.f; return x; }
        if (x < -35.f) { d = 0.f; return 0.f; }
        { const float t = __expf(x); d = 1.f / (1.f + 1.f / t); return log1pf(t); }
    case AF_RELU: if (x > 0.f) { d = 1.f; return x; } d = 0.f; return 0.f;
    case AF_GAUSSIAN: { const float t = __expf(-0.5f * x * x); d = -x * t; return t; }
    case AF_COS: d = -sinf(x); return cosf(x);
    case AF_REVLOGISTIC: v = 1.f / (1.f + __expf(-x)); d = v * (v - 1.f); return 1.f - v;
    case AF_EXP: v = __expf(-x); d = -v; return v;
    case AF_HARMONIC: d = 2.f * x; return x * x;
    }
    d = 0.f;
    return 0.f;
}

// A fragment: rows = atoms (lane % 16), k = k0 + 4*(lane/16) + 0..3 from an
// LDS row-major [16][stride] FP32 tile, split into bf16 hi and lo.
__device__ inline void a_frag(const float* tile, int stride, int k0, s4& hi, s4& lo)
{
    const int lane = threadIdx.x & (kWave - 1);
    const float* p = tile + (lane & 15) * stride + k0 + 4 * (lane >> 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const uint16_t h = to_bf16(p[v]);
        hi[v] = (short)h;
        lo[v] = (short)to_bf16(p[v] - from_bf16(h));
    }
}

// B fragment: B[k][n] = Wt[n][k] for n = n0 + lane % 16, k = k0 + 4*(lane/16)
// + 0..3, from a row-major bf16 matrix with row length ld (hi; lo at +lo_off).
__device__ inline void b_frag(const uint16_t* __restrict__ w, int ld, int lo_off, int n0, int k0, s4& hi, s4& lo)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint16_t* p = w + (size_t)(n0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4);
    const uint2 h = *reinterpret_cast<const uint2*>(p);
    const uint2 l = *reinterpret_cast<const uint2*>(p + lo_off);
    hi = __builtin_bit_cast(s4, h);
    lo = __builtin_bit_cast(s4, l);
}

__device__ inline f4 mma(f4 c, const s4& ah, const s4& al, const s4& bh, const s4& bl, int terms)
{
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
    if (terms > 1) {
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    }
    return c;
}

// kNNWaves waves share one tile of 16 atoms: each takes every kNNWaves-th
// 16-column block of a layer, so the chain of dependent MFMA steps per wave
// is that much shorter (the network kernel is latency-bound: a few waves
// per CU).
__global__ __launch_bounds__(kWave * kNNWaves) void k_nnbf16(const ModelT* __restrict__ M, NNMix a,
                                                            const float* __restrict__ Gin,
                                                            double* __restrict__ Eatom, double* __restrict__ dEdG,
                                                            float* __restrict__ dEdGf)
{
    extern __shared__ float sm[];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.nzero; q += (size_t)gridDim.x * blockDim.x)
        a.zero[q] = 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    int e = 0;
    while (e + 1 < a.nelem && (int)blockIdx.x >= a.blk_off[e + 1]) ++e;
    const NNT& N = M->el[e].nn;
    const int nl = N.nl;
    const int st = a.stride;
    float* bufI = sm;                           // [16][stride_in] network input
    float* bufA = bufI + kNNRows * a.stride_in;  // hidden/output tiles [16][st]
    float* bufB = bufA + kNNRows * st;
    float* dfd = bufB + kNNRows * st;  // [layer][16][st]
    const int tile = (int)blockIdx.x - a.blk_off[e];
    const int first = a.atom_off[e] + tile * kNNRows;
    const int cnt = min(kNNRows, a.atom_off[e + 1] - first);

    // layer-0 input rows
    const int pin0 = a.pin[e][1];
    for (int x = threadIdx.x; x < kNNRows * pin0; x += kWave * kNNWaves) {
        const int r = x / pin0, k = x % pin0;
        bufI[r * a.stride_in + k] = r < cnt ? Gin[(size_t)a.atoms[first + r] * a.ldK + k] : 0.f;
    }
    __syncthreads();
    float* in = bufI;
    float* out = bufA;
    int sin = a.stride_in;
    for (int l = 1; l < nl; ++l) {
        const int pi = a.pin[e][l], po = a.pout[e][l];
        const uint16_t* W = a.w + a.wW[e][l];
        const float* bias = a.b + a.bo[e][l];
        const float norm = N.normalize ? 1.f / N.n[l - 1] : 1.f;
        for (int n0 = 16 * wv; n0 < po; n0 += 16 * kNNWaves) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            // the next weight fragment is loaded while this one multiplies
            s4 bh, bl;
            b_frag(W, pi, po * pi, n0, 0, bh, bl);
            for (int k0 = 0; k0 < pi; k0 += 16) {
                s4 ah, al, nh = bh, nl_ = bl;
                if (k0 + 16 < pi) b_frag(W, pi, po * pi, n0, k0 + 16, nh, nl_);
                a_frag(in, sin, k0, ah, al);
                c = mma(c, ah, al, bh, bl, a.terms);
                bh = nh;
                bl = nl_;
            }
            const int n = n0 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * (lane >> 4) + v;
                float d;
                const float x = (c[v] + bias[n]) * norm;
                const float y = act32(N.act[l], x, d);
                out[r * st + n] = n < N.n[l] ? y : 0.f;
                dfd[(l * kNNRows + r) * st + n] = n < N.n[l] ? d : 0.f;
            }
        }
        __syncthreads();
        // the hidden and output tiles alternate between bufA and bufB
        in = out;
        out = in == bufA ? bufB : bufA;
        sin = st;
    }
    if (threadIdx.x < cnt) Eatom[a.atoms[first + threadIdx.x]] = (double)in[threadIdx.x * st];
    // adjoint: delta of the output layer is f'_out for the energy neuron
    float* dl = out;
    const int pl = a.pout[e][nl - 1];
    for (int x = threadIdx.x; x < kNNRows * pl; x += kWave * kNNWaves) {
        const int r = x / pl, n = x % pl;
        dl[r * st + n] = n == 0 ? dfd[((nl - 1) * kNNRows + r) * st] : 0.f;
    }
    __syncthreads();
    float* dn = in;
    for (int l = nl - 1; l >= 1; --l) {
        const int pi = a.pin[e][l], po = a.pout[e][l];
        const uint16_t* WT = a.w + a.wT[e][l];
        const float norm = N.normalize ? 1.f / N.n[l - 1] : 1.f;
        for (int m0 = 16 * wv; m0 < pi; m0 += 16 * kNNWaves) {
            f4 c = {0.f, 0.f, 0.f, 0.f};
            s4 bh, bl;
            b_frag(WT, po, po * pi, m0, 0, bh, bl);
            for (int k0 = 0; k0 < po; k0 += 16) {
                s4 ah, al, nh = bh, nl_ = bl;
                if (k0 + 16 < po) b_frag(WT, po, po * pi, m0, k0 + 16, nh, nl_);
                a_frag(dl, st, k0, ah, al);
                c = mma(c, ah, al, bh, bl, a.terms);
                bh = nh;
                bl = nl_;
            }
            const int m = m0 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * (lane >> 4) + v;
                float g = c[v] * norm;
                if (l > 1) {
                    dn[r * st + m] = g * dfd[((l - 1) * kNNRows + r) * st + m];
                } else if (r < cnt && m < N.n[0]) {
                    const size_t at = (size_t)a.atoms[first + r];
                    dEdG[at * a.ldG + m] = (double)g;
                    dEdGf[at * a.ldG + m] = g * a.sffac[e * kMaxSF + m];
                }
            }
        }
        __syncthreads();
        float* t = dl; dl = dn; dn = t;
    }
}

// --------------------------------------------------- fused force backward --
// One triplet's force terms, P1 r_ij + P2 r_ik + P3 (r_ik - r_ij) in the
// form (P1 - P3) r_ij + P3 r_ik for j and (P2 - P3) r_ik + P3 r_ij for k:
// the atom's own and j's parts summed in FP32 over the lane's run of j, k's
// part added to its FP64 LDS sum.
__device__ inline void force_terms(float P1, float P2, float P3, const float4& gj, const float4& gk, float& sx,
                                   float& sy, float& sz, float& rx, float& ry, float& rz, double* kx, double* ky,
                                   double* kz)
{
    const float d1 = P1 - P3, d2 = P2 - P3;
    sx = fmaf(P1, gj.x, fmaf(P2, gk.x, sx));
    sy = fmaf(P1, gj.y, fmaf(P2, gk.y, sy));
    sz = fmaf(P1, gj.z, fmaf(P2, gk.z, sz));
    rx = fmaf(d1, gj.x, fmaf(P3, gk.x, rx));
    ry = fmaf(d1, gj.y, fmaf(P3, gk.y, ry));
    rz = fmaf(d1, gj.z, fmaf(P3, gk.z, rz));
    lds_add(kx, (double)fmaf(d2, gk.x, P3 * gj.x));
    lds_add(ky, (double)fmaf(d2, gk.y, P3 * gj.y));
    lds_add(kz, (double)fmaf(d2, gk.z, P3 * gj.z));
}

// Forces of all power-path angular groups of an element in one triplet walk
// over the element-ordered in-cutoff list: each triplet takes the term
// weights (wb, wa) of its pair's group from LDS; the geometry and the
// accumulation are those of the per-group walk in k_force32.
template <int NB0, int NB1>
__device__ inline void force32_power_merged(const ModelT* __restrict__ M, const Stage& S,
                                            const MemberF* __restrict__ MF, const GroupMixF* __restrict__ GM,
                                            const AngMergeF& am, int n, int ctype, float tanhPre, char* mbase,
                                            const float* __restrict__ wi, double* ax, double* ay, double* az,
                                            double& fx, double& fy, double& fz, TripCursor& tc)
{
    constexpr int NBAS = 2 * (NB0 + NB1);
    constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
    const int lane = threadIdx.x;
    const GroupT& G0 = M->grp[am.grp[0]];
    const GroupF g = group_f(G0);
    const GroupMixF gm = GM[am.grp[0]];
    const bool wide = G0.type == 9;
    const int ng = am.ng;
    signed char* pmap = reinterpret_cast<signed char*>(mbase);
    float2* wtab = reinterpret_cast<float2*>(mbase + kMaxElem * kMaxElem);
    for (int x = lane; x < kMaxElem * kMaxElem; x += kWave) pmap[x] = am.pair[x];
    // wb = sum of the members' w on term b, wa = sum of zeta lambda w
    for (int x = lane; x < ng * NBAS; x += kWave) {
        const int q = x / NBAS, b = x - q * NBAS;
        const GroupT& GT = M->grp[am.grp[q]];
        const MemberF* mf = MF + GT.mem0;
        float sb = 0.f, sa = 0.f;
        for (int m = 0; m < GT.nmem; ++m)
            if (mf[m].beta == b) {
                const float wq = wi[mf[m].sf];
                sb += wq;
                sa += wq * mf[m].zetaLambda;
            }
        wtab[x] = make_float2(sb, sa);
    }
    // all elements' in-cutoff neighbours, element by element (staging order)
    const List A = make_list(S, n, -1, g.rc, full_cutoff(M, G0), S.la);  // its barrier orders the tables above
    int cur = -1, rowj = 0;
    float rx = 0.f, ry = 0.f, rz = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    auto flush = [&]() {
        if (cur >= 0) {
            lds_add(ax + cur, (double)rx);
            lds_add(ay + cur, (double)ry);
            lds_add(az + cur, (double)rz);
            fx -= (double)sx;
            fy -= (double)sy;
            fz -= (double)sz;
        }
        cur = -1;
    };
    float4 gj = make_float4(0.f, 0.f, 0.f, 0.f), cj = gj;
    auto work = [&](int j, int k) {
        if (j != cur) {
            flush();
            cur = j;
            rx = 0.f; ry = 0.f; rz = 0.f;
            sx = 0.f; sy = 0.f; sz = 0.f;
            gj = S.g[j];
            cj = S.c[j];
            rowj = S.el[j] * kMaxElem;
        }
        const int q = pmap[rowj + S.el[k]];
        if (q < 0) return;
        const float2* wq = wtab + q * NBAS;
        const float4 gk = S.g[k];
        const float4 ck = S.c[k];
        const float r2jk = dist2(gj, gk);
        const float ijk = __builtin_amdgcn_rsqf(r2jk);
        const float rjk = r2jk * ijk;
        float fcjk = 1.f, dfcjk = 0.f;
        if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, rjk, fcjk, dfcjk);
        const float rinvijik = cj.x * ck.x;
        const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * rinvijik;
        const float pfc = cj.y * ck.y * fcjk;
        const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
        const float pr1 = ck.y * fcjk * cj.z * cj.x;
        const float pr2 = cj.y * fcjk * ck.z * ck.x;
        const float pr3 = cj.y * ck.y * dfcjk * ijk;
        const float E0 = fexp2(-gm.etaL[0] * r2sum);
        const float E1 = NB1 > 0 ? fexp2(-gm.etaL[1] * r2sum) : 0.f;
        // (1 + lambda cos)^(2^b - 1) carried beside the power: no division
        float Q0 = 0.f, Q1 = 0.f, qa0 = 0.f, qa1 = 0.f;
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            const float pl = fmaf(gm.lam[sl], cost, 1.f);
            const bool pos = pl > 0.f;
            float p = pos ? pl : 0.f, pq = pos ? 1.f : 0.f;
#pragma unroll
            for (int b = 0; b < NBM; ++b) {
                if (b < NB0) {
                    const float2 w = wq[sl * NB0 + b];
                    Q0 = fmaf(w.x, p, Q0);
                    qa0 = fmaf(w.y, pq, qa0);
                }
                if (b < NB1) {
                    const float2 w = wq[2 * NB0 + sl * NB1 + b];
                    Q1 = fmaf(w.x, p, Q1);
                    qa1 = fmaf(w.y, pq, qa1);
                }
                pq *= p;
                p *= p;
            }
        }
        const float A = E0 * qa0 + E1 * qa1;
        const float B0 = E0 * Q0, B1 = E1 * Q1;
        const float B = B0 + B1;
        const float C0 = gm.twoEta[0] * B0 + gm.twoEta[1] * B1;
        const float C3 = wide ? 0.f : C0;
        const float P1 = pfc * (A * (rinvijik - cost * cj.x * cj.x) - C0) + B * pr1;
        const float P2 = pfc * (A * (rinvijik - cost * ck.x * ck.x) - C0) + B * pr2;
        const float P3 = pfc * (A * rinvijik + C3) - B * pr3;
        force_terms(P1, P2, P3, gj, gk, sx, sy, sz, rx, ry, rz, ax + k, ay + k, az + k);
    };
    if (tc.rows) walk_rows(S, A, A, true, wide ? INFINITY : g.rc2, work, flush);
    else walk_blocks<false>(S, nullptr, A, A, true, wide ? INFINITY : g.rc2, tc, false, work, flush);
    flush();
    __syncthreads();
}

// Forces on the atom's neighbours are accumulated per neighbour entry in LDS
// (FP64) and pushed to the force array with FP64 atomics; contributions to
// the atom's own periodic images are dropped in the pull form
// (Mode::calculateForces skips them, Mode.cpp:1957) and kept in the push form.
template <int MAXM, bool FAST, int NB0, int NB1, bool MERGE>
__global__ __launch_bounds__(64) void k_force32(const ModelT* __restrict__ M, const MemberF* __restrict__ MF,
                                                const GroupMixF* __restrict__ GM, const AngMergeF* __restrict__ AM,
                                                SysView s, const float* __restrict__ fw, int ldG,
                                                double* __restrict__ F, int push, int capL, int trip, TripList T,
                                                StageKeep K)
{
    extern __shared__ float sm[];
    const Stage S = carve(sm, capL, trip, stage_mp(NB0, NB1));
    double* ax = reinterpret_cast<double*>(reinterpret_cast<char*>(sm) + stage_bytes(capL, trip, stage_mp(NB0, NB1)));
    double* ay = ax + capL;
    double* az = ay + capL;
    unsigned short* idx = reinterpret_cast<unsigned short*>(az + capL);
    char* mbase = reinterpret_cast<char*>(idx + capL);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const float tanhPre = (float)M->tanhPre;
    const int n = K.g ? load_stage(M, K, i, S) : stage_atom(M, s, i, capL, S, idx);
    TripCursor tc = trip_cursor(T, i);
    for (int k = lane; k < n; k += kWave) { ax[k] = 0.0; ay[k] = 0.0; az[k] = 0.0; }
    const float* wi = fw + (size_t)i * ldG;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const bool merged = MERGE && AM[e].ng > 0;
    __syncthreads();

    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& GT = M->grp[E.grp0 + gi];
        const GroupF g = group_f(GT);
        const int nmem = GT.nmem;
        const MemberF* mf = MF + GT.mem0;
        if (E.ucut) {
            if (gi == 0) all_cutoffs(S, n, ctype, g, tanhPre);
        } else {
            group_cutoffs(S, n, GT.e1, GT.type == 2 ? GT.e1 : GT.e2, ctype, g, tanhPre);
        }
        if (sf_compact(GT.type)) continue;  // the compact kernel's
        if (merged && GT.type != 2 && GT.type != 12) continue;  // the merged walk's
        if (GT.type == 2 || GT.type == 12) {
            // radial: each lane owns its neighbours, no conflicts; type 12
            // weights every neighbour by its atomic number
            const bool wt = GT.type == 12;
            // type 2: the slots of neighbour element e1; type 12: every slot
            const int lo = wt ? 0 : S.eoff[GT.e1], hi = wt ? n : S.eoff[GT.e1 + 1];
            for (int k = lo + lane; k < hi; k += kWave) {
                const float4 gk = S.g[k];
                if (!(gk.w < g.rc)) continue;
                const float4 ck = S.c[k];
                float P = 0.f;
#pragma unroll 2
                for (int q = 0; q < nmem; ++q) {
                    const float dr = gk.w - mf[q].rs;
                    const float ex = __expf(-mf[q].eta * dr * dr);
                    P += wi[mf[q].sf] * (ck.z - mf[q].twoEta * dr * ck.y) * ex;
                }
                P *= wt ? ck.x * ck.w : ck.x;
                const double vx = (double)(P * gk.x), vy = (double)(P * gk.y), vz = (double)(P * gk.z);
                fx -= vx; fy -= vy; fz -= vz;
                ax[k] += vx; ay[k] += vy; az[k] += vz;
            }
            __syncthreads();
            continue;
        }
        const bool same = GT.e1 == GT.e2;
        const bool wide = GT.type == 9;  // no r_jk cutoff or exponent term (SymGrpExpAngw.cpp:156-380)
        const bool wt = GT.type == 13;    // Z_j Z_k weights (SymGrpExpAngnWeighted.cpp)
        const bool full = full_cutoff(M, GT);
        const List A = make_list(S, n, GT.e1, g.rc, full, S.la);
        const List B = same ? A : make_list(S, n, GT.e2, g.rc, full, S.lb);
        const GroupMixF gm = GM[E.grp0 + gi];
        constexpr bool POW = NB0 > 0;
        constexpr int NBAS = POW ? 2 * (NB0 + NB1) : 1;
        constexpr int NBM = NB0 > NB1 ? NB0 : NB1;
        float pwb[NBAS], pwa[NBAS];
        MemberMasks R{0u, 0u, 0u};
        if constexpr (POW) power_weights<NBAS>(S, mf, nmem, wi, pwb, pwa);
        else R = load_members(S, mf, nmem, wi);
        // FAST: the member constants stay in registers for the whole group
        // (zeta, -eta log2 e, the weights of B and C, and the weight of A
        // split by lambda index so A needs no per-member select)
        constexpr int NC = FAST && !POW ? MAXM : 1;
        float cz[NC], ce[NC], cA0[NC], cA1[NC], cB[NC], cC[NC];
        if constexpr (FAST && !POW) {
#pragma unroll
            for (int q = 0; q < MAXM; ++q) {
                const float4 mq = S.mp[q];
                const float4 wq = S.mp[kMaxMembersExp + q];
                const bool s1 = (R.s1 >> q) & 1u;
                cz[q] = mq.x;
                ce[q] = -mq.y;
                cB[q] = wq.y;
                cC[q] = wq.z;
                cA0[q] = s1 ? 0.f : wq.x;
                cA1[q] = s1 ? wq.x : 0.f;
            }
        }
        // Contributions to the row atom j and to the atom itself are summed
        // in FP32 over the lane's run of equal j (a few triplets) and added
        // in FP64 when j changes.
        int cur = -1;
        float rx = 0.f, ry = 0.f, rz = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
        auto flush = [&]() {
            if (cur >= 0) {
                lds_add(ax + cur, (double)rx);
                lds_add(ay + cur, (double)ry);
                lds_add(az + cur, (double)rz);
                fx -= (double)sx;
                fy -= (double)sy;
                fz -= (double)sz;
            }
        };
        float4 gj = make_float4(0.f, 0.f, 0.f, 0.f), cj = gj;  // geometry of the run's j
        auto work = [&](int j, int k) {
            if (j != cur) {
                flush();
                cur = j;
                rx = 0.f; ry = 0.f; rz = 0.f;
                sx = 0.f; sy = 0.f; sz = 0.f;
                gj = S.g[j];
                cj = S.c[j];
            }
            const float4 gk = S.g[k];
            const float4 ck = S.c[k];
            const float r2jk = dist2(gj, gk);
            const float ijk = __builtin_amdgcn_rsqf(r2jk);
            const float rjk = r2jk * ijk;
            float fcjk, dfcjk;
            fcjk = 1.f;
            dfcjk = 0.f;
            if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, rjk, fcjk, dfcjk);
            const float rinvijik = cj.x * ck.x;
            const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * rinvijik;
            const float pfc = cj.y * ck.y * fcjk;
            const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
            const float pr1 = ck.y * fcjk * cj.z * cj.x;
            const float pr2 = cj.y * fcjk * ck.z * ck.x;
            const float pr3 = cj.y * ck.y * dfcjk * ijk;
            // per distinct lambda: log2(pl) (-inf where pl <= 0) and 1/pl
            constexpr int MAXS = FAST ? 2 : 4;
            float L[MAXS], Rp[MAXS];
#pragma unroll
            for (int t = 0; t < MAXS; ++t) {
                const float pl = 1.f + gm.lam[t] * cost;
                const bool pos = pl > 0.f;
                L[t] = pos ? flog2(pl) : -INFINITY;
                Rp[t] = pos ? __builtin_amdgcn_rcpf(pl) : 0.f;
            }
            // t_q = vexp * pl^zeta: u_q * pl = w t_q and u_q = w t_q / pl,
            // u_q = w vexp pl^(zeta-1) being the reference's fg times w
            float A = 0.f, B = 0.f, C0 = 0.f, C1 = 0.f, C2 = 0.f, C3 = 0.f;
            if constexpr (POW) {
                // B_h = E_h sum wb P, A_s = sum_h E_h sum wa P (powers of lambda s)
                const float E0 = fexp2(-gm.etaL[0] * r2sum);
                const float E1 = NB1 > 0 ? fexp2(-gm.etaL[1] * r2sum) : 0.f;
                // (1 + lambda cos)^(2^b - 1) is carried beside the power, so
                // A needs no division
                float Q0 = 0.f, Q1 = 0.f, qa0 = 0.f, qa1 = 0.f;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const float pl = fmaf(gm.lam[sl], cost, 1.f);
                    const bool pos = pl > 0.f;
                    float p = pos ? pl : 0.f, pq = pos ? 1.f : 0.f;
#pragma unroll
                    for (int b = 0; b < NBM; ++b) {
                        if (b < NB0) {
                            Q0 = fmaf(pwb[sl * NB0 + b], p, Q0);
                            qa0 = fmaf(pwa[sl * NB0 + b], pq, qa0);
                        }
                        if (b < NB1) {
                            Q1 = fmaf(pwb[2 * NB0 + sl * NB1 + b], p, Q1);
                            qa1 = fmaf(pwa[2 * NB0 + sl * NB1 + b], pq, qa1);
                        }
                        pq *= p;
                        p *= p;
                    }
                }
                const float B0 = E0 * Q0, B1 = E1 * Q1;
                B = B0 + B1;
                C0 = gm.twoEta[0] * B0 + gm.twoEta[1] * B1;
                A = E0 * qa0 + E1 * qa1;
            } else if constexpr (FAST) {
                float A0 = 0.f, A1 = 0.f;
#pragma unroll
                for (int q = 0; q < MAXM; ++q) {
                    const bool s1 = (R.s1 >> q) & 1u;
                    const float t = fexp2(fmaf(cz[q], s1 ? L[1] : L[0], ce[q] * r2sum));
                    B = fmaf(cB[q], t, B);
                    C0 = fmaf(cC[q], t, C0);
                    A0 = fmaf(cA0[q], t, A0);
                    A1 = fmaf(cA1[q], t, A1);
                }
                A = A0 * Rp[0] + A1 * Rp[1];
            } else {
                float ec = 0.f, fij = 1.f, fik = 1.f, fjk = 1.f;
                bool shifted = false;
#pragma unroll
                for (int q = 0; q < MAXM; ++q)
                    if (q < nmem) {
                        const float4 mq = S.mp[q];
                        const float4 wq = S.mp[kMaxMembersExp + q];
                        if ((R.calc >> q) & 1u) {
                            float arg = r2sum;
                            shifted = mq.z > 0.f;
                            if (shifted) {
                                const float a = gj.w - mq.z, b = gk.w - mq.z, c = rjk - mq.z;
                                arg = a * a + b * b + (wide ? 0.f : c * c);
                                fij = a * cj.x;
                                fik = b * ck.x;
                                fjk = wide ? 0.f : c * ijk;
                            }
                            ec = -mq.y * arg;
                        }
                        const bool s1 = mq.w != 0.f, s2 = wq.w != 0.f;
                        const float Ls = s2 ? (s1 ? L[3] : L[2]) : (s1 ? L[1] : L[0]);
                        const float Rs = s2 ? (s1 ? Rp[3] : Rp[2]) : (s1 ? Rp[1] : Rp[0]);
                        const float t = fexp2(fmaf(mq.x, Ls, ec));
                        B = fmaf(wq.y, t, B);
                        const float c = wq.z * t;
                        if (shifted) {
                            C1 = fmaf(c, fij, C1);
                            C2 = fmaf(c, fik, C2);
                            C3 = fmaf(c, fjk, C3);
                        } else {
                            C0 += c;
                        }
                        A = fmaf(wq.x, t * Rs, A);
                    }
            }
            C1 += C0; C2 += C0;
            if (!wide) C3 += C0;
            const float zw = wt ? cj.w * ck.w : 1.f;
            const float P1 = zw * (pfc * (A * (rinvijik - cost * cj.x * cj.x) - C1) + B * pr1);
            const float P2 = zw * (pfc * (A * (rinvijik - cost * ck.x * ck.x) - C2) + B * pr2);
            const float P3 = zw * (pfc * (A * rinvijik + C3) - B * pr3);
            force_terms(P1, P2, P3, gj, gk, sx, sy, sz, rx, ry, rz, ax + k, ay + k, az + k);
        };
        if (tc.rows) walk_rows(S, A, B, same, wide ? INFINITY : g.rc2, work);
        else walk_blocks<false>(S, nullptr, A, B, same, wide ? INFINITY : g.rc2, tc, false, work);
        flush();
        __syncthreads();
    }
    if constexpr (MERGE && NB0 > 0) {
        if (merged)
            force32_power_merged<NB0, NB1>(M, S, MF, GM, AM[e], n, ctype, tanhPre, mbase, wi, ax, ay, az, fx, fy, fz,
                                           tc);
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
        atomicAdd(F + 3 * (size_t)i, fx);
        atomicAdd(F + 3 * (size_t)i + 1, fy);
        atomicAdd(F + 3 * (size_t)i + 2, fz);
    }
    const size_t base = (size_t)i * s.cap;
    for (int k = lane; k < n; k += kWave) {
        const int j = K.g ? K.nbr[(size_t)i * K.ld + k] : s.nb_j[base + idx[k]];
        if (j == i && !push) continue;
        atomicAdd(F + 3 * (size_t)j, ax[k]);
        atomicAdd(F + 3 * (size_t)j + 1, ay[k]);
        atomicAdd(F + 3 * (size_t)j + 2, az[k]);
    }
}


// ------------------------------------------- tangent of G along a field --
// dG_i/dx . u for every SF of every atom (the contraction of
// Training::collectDGdxia with a displacement field u), recomputing the
// radial and angular terms: per triplet the three geometric scalars
// s1 = r_ij.(u_i - u_j), s2 = r_ik.(u_i - u_k), s3 = r_jk.(u_k - u_j) are
// combined into three coefficients, and each member adds t_q times their
// member-weighted sum.  FP32 compensated sums per lane, FP64 across lanes,
// scaled like the SFs themselves (scaling factor and angular normalisation).
template <int MAXM, bool FAST>
__global__ __launch_bounds__(64) void k_tangent32(const ModelT* __restrict__ M, const MemberF* __restrict__ MF,
                                                  const GroupMixF* __restrict__ GM, SysView s,
                                                  const double* __restrict__ u, const float* __restrict__ sffac,
                                                  double* __restrict__ Gdot, int ldG, int capL, int trip)
{
    extern __shared__ float sm[];
    const Stage S = carve(sm, capL, trip, stage_mp(0, 0));
    float4* du = reinterpret_cast<float4*>(reinterpret_cast<char*>(sm) + stage_bytes(capL, trip, stage_mp(0, 0)));
    unsigned short* idx = reinterpret_cast<unsigned short*>(du + capL);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const int ctype = M->cutoffType;
    const float tanhPre = (float)M->tanhPre;
    const int n = stage_atom(M, s, i, capL, S, idx);
    {
        const double ux = u[3 * (size_t)i], uy = u[3 * (size_t)i + 1], uz = u[3 * (size_t)i + 2];
        const size_t base = (size_t)i * s.cap;
        for (int k = lane; k < n; k += kWave) {
            const int j = s.nb_j[base + idx[k]];
            const float dx = (float)(ux - u[3 * (size_t)j]);
            const float dy = (float)(uy - u[3 * (size_t)j + 1]);
            const float dz = (float)(uz - u[3 * (size_t)j + 2]);
            const float4 g = S.g[k];
            du[k] = make_float4(dx, dy, dz, g.x * dx + g.y * dy + g.z * dz);
        }
    }
    double* Gi = Gdot + (size_t)i * ldG;
    __syncthreads();

    for (int gi = 0; gi < E.ngrp; ++gi) {
        const GroupT& GT = M->grp[E.grp0 + gi];
        const GroupF g = group_f(GT);
        const int nmem = GT.nmem;
        const MemberF* mf = MF + GT.mem0;
        if (E.ucut) {
            if (gi == 0) all_cutoffs(S, n, ctype, g, tanhPre);
        } else {
            group_cutoffs(S, n, GT.e1, GT.type == 2 ? GT.e1 : GT.e2, ctype, g, tanhPre);
        }
        if (sf_compact(GT.type)) continue;  // the compact kernel's
        if (GT.type == 2 || GT.type == 12) {
            const bool wt = GT.type == 12;  // every neighbour, times its atomic number
            for (int q0 = 0; q0 < nmem; q0 += kRadChunk) {
                float acc[kRadChunk], cmp[kRadChunk];
#pragma unroll
                for (int q = 0; q < kRadChunk; ++q) { acc[q] = 0.f; cmp[q] = 0.f; }
                // type 2: the slots of neighbour element e1; type 12: every slot
                const int lo = wt ? 0 : S.eoff[GT.e1], hi = wt ? n : S.eoff[GT.e1 + 1];
                for (int k = lo + lane; k < hi; k += kWave) {
                    const float4 gk = S.g[k];
                    if (!(gk.w < g.rc)) continue;
                    const float4 ck = S.c[k];
                    const float sk = du[k].w * (wt ? ck.x * ck.w : ck.x);  // (r_ij / r).(u_i - u_j)
#pragma unroll
                    for (int q = 0; q < kRadChunk; ++q)
                        if (q0 + q < nmem) {
                            const float dr = gk.w - mf[q0 + q].rs;
                            const float ex = __expf(-mf[q0 + q].eta * dr * dr);
                            kahan(acc[q], cmp[q], (ck.z - mf[q0 + q].twoEta * dr * ck.y) * ex * sk);
                        }
                }
#pragma unroll
                for (int q = 0; q < kRadChunk; ++q)
                    if (q0 + q < nmem) {
                        const double v = wave_sum((double)acc[q] + (double)cmp[q]);
                        if (lane == 0) {
                            const int sf = mf[q0 + q].sf;
                            Gi[sf] = v * (double)sffac[e * kMaxSF + sf];
                        }
                    }
            }
            __syncthreads();
            continue;
        }
        float acc[MAXM], cmp[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; ++q) { acc[q] = 0.f; cmp[q] = 0.f; }
        const bool same = GT.e1 == GT.e2;
        const bool wide = GT.type == 9;  // no r_jk cutoff or exponent term (SymGrpExpAngw.cpp:156-380)
        const bool wt = GT.type == 13;    // Z_j Z_k weights (SymGrpExpAngnWeighted.cpp)
        const bool full = full_cutoff(M, GT);
        const List A = make_list(S, n, GT.e1, g.rc, full, S.la);
        const List B = same ? A : make_list(S, n, GT.e2, g.rc, full, S.lb);
        const GroupMixF gm = GM[E.grp0 + gi];
        const MemberMasks R = load_members(S, mf, nmem, nullptr);
        walk_all(S, A, B, same, wide ? INFINITY : g.rc2, [&](int j, int k) {
            const float4 gj = S.g[j], gk = S.g[k];
            const float4 cj = S.c[j], ck = S.c[k];
            const float r2jk = dist2(gj, gk);
            const float4 uj = du[j], uk = du[k];
            const float ijk = __builtin_amdgcn_rsqf(r2jk);
            const float rjk = r2jk * ijk;
            float fcjk, dfcjk;
            fcjk = 1.f;
            dfcjk = 0.f;
            if (!wide) cutoff32(ctype, g.rc, g.rci, g.iw, g.rcinv, tanhPre, rjk, fcjk, dfcjk);
            const float rinvijik = cj.x * ck.x;
            const float cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z) * rinvijik;
            // type 13: every term and derivative times Z_j Z_k
            const float zw = wt ? cj.w * ck.w : 1.f;
            const float pfc = cj.y * ck.y * fcjk * zw;
            const float r2sum = gj.w * gj.w + gk.w * gk.w + (wide ? 0.f : r2jk);
            const float pr1 = ck.y * fcjk * cj.z * cj.x * zw;
            const float pr2 = cj.y * fcjk * ck.z * ck.x * zw;
            const float pr3 = cj.y * ck.y * dfcjk * ijk * zw;
            // geometric scalars of the field
            const float s1 = uj.w, s2 = uk.w;
            const float s3 = (gk.x - gj.x) * (uj.x - uk.x) + (gk.y - gj.y) * (uj.y - uk.y) +
                             (gk.z - gj.z) * (uj.z - uk.z);
            const float al = pfc * ((rinvijik - cost * cj.x * cj.x) * s1 + (rinvijik - cost * ck.x * ck.x) * s2 +
                                    rinvijik * s3);
            const float be = pr1 * s1 + pr2 * s2 - pr3 * s3;
            const float ga = pfc * ((wide ? 0.f : s3) - s1 - s2);
            constexpr int MAXS = FAST ? 2 : 4;
            float L[MAXS], Rp[MAXS];
#pragma unroll
            for (int t = 0; t < MAXS; ++t) {
                const float pl = 1.f + gm.lam[t] * cost;
                const bool pos = pl > 0.f;
                L[t] = pos ? flog2(pl) : -INFINITY;
                Rp[t] = pos ? __builtin_amdgcn_rcpf(pl) : 0.f;
            }
            if constexpr (FAST) {
#pragma unroll
                for (int q = 0; q < MAXM; ++q) {
                    const float4 mq = S.mp[q];
                    const bool s1b = mq.w != 0.f;
                    const float t = fexp2(fmaf(mq.x, s1b ? L[1] : L[0], -mq.y * r2sum));
                    const float zl = mq.x * gm.lam[s1b ? 1 : 0];
                    const float c = fmaf(zl * (s1b ? Rp[1] : Rp[0]), al, be);
                    kahan(acc[q], cmp[q], t * fmaf(2.f * mq.y * 0.69314718f, ga, c));
                }
            } else {
                float ec = 0.f, fij = 1.f, fik = 1.f, fjk = 1.f;
                bool shifted = false;
#pragma unroll
                for (int q = 0; q < MAXM; ++q)
                    if (q < nmem) {
                        const float4 mq = S.mp[q];
                        if ((R.calc >> q) & 1u) {
                            float arg = r2sum;
                            shifted = mq.z > 0.f;
                            if (shifted) {
                                const float a = gj.w - mq.z, b = gk.w - mq.z, c = rjk - mq.z;
                                arg = a * a + b * b + (wide ? 0.f : c * c);
                                fij = a * cj.x;
                                fik = b * ck.x;
                                fjk = wide ? 0.f : c * ijk;
                            }
                            ec = -mq.y * arg;
                        }
                        const unsigned b1 = (R.s1 >> q) & 1u, b2 = (R.s2 >> q) & 1u;
                        const int sidx = (int)b1 | ((int)b2 << 1);
                        const float Ls = pickb<MAXS>(L, b1, b2);
                        const float Rs = pickb<MAXS>(Rp, b1, b2);
                        const float t = fexp2(fmaf(mq.x, Ls, ec));
                        const float zl = mq.x * pick<MAXS>(gm.lam, sidx);
                        const float twoEta = 2.f * mq.y * 0.69314718f;
                        float v = fmaf(zl * Rs, al, be);
                        if (shifted) v += twoEta * pfc * (fjk * s3 - fij * s1 - fik * s2);
                        else v += twoEta * ga;
                        kahan(acc[q], cmp[q], t * v);
                    }
            }
        });
#pragma unroll
        for (int q = 0; q < MAXM; ++q)
            if (q < nmem) {
                const double v = wave_sum((double)acc[q] + (double)cmp[q]);
                if (lane == 0) {
                    const int sf = mf[q].sf;
                    Gi[sf] = v * (double)sffac[e * kMaxSF + sf];
                }
            }
        __syncthreads();
    }
}

__global__ void k_zero(double* p, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.0;
}

__global__ void k_mfma_probe(const float* A, const float* B, float* D)
{
    // A, B: 16x16 row-major; D = A * B (16x16 row-major)
    const int lane = threadIdx.x;
    s4 ah, al, bh, bl;
    a_frag(A, 16, 0, ah, al);
    // B fragment from B^T (row n, column k)
    float bt[4];
    for (int v = 0; v < 4; ++v) bt[v] = B[(4 * (lane >> 4) + v) * 16 + (lane & 15)];
    for (int v = 0; v < 4; ++v) {
        const uint16_t h = to_bf16(bt[v]);
        bh[v] = (short)h;
        bl[v] = (short)to_bf16(bt[v] - from_bf16(h));
    }
    f4 c = {0.f, 0.f, 0.f, 0.f};
    c = mma(c, ah, al, bh, bl, 3);
    for (int v = 0; v < 4; ++v) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = c[v];
}

// Calls f(integral_constant MAXM, bool_constant FAST) for the kernel
// instance matching the model: exact member counts {4, 8, 10, 16} on the fast
// path, powers of two up to 32 on the general one.
template <typename F>
void dispatch_members(int exact, int maxm, F&& f)
{
    using std::integral_constant;
    switch (exact) {
    case 4: f(integral_constant<int, 4>{}, std::true_type{}); return;
    case 8: f(integral_constant<int, 8>{}, std::true_type{}); return;
    case 10: f(integral_constant<int, 10>{}, std::true_type{}); return;
    case 16: f(integral_constant<int, 16>{}, std::true_type{}); return;
    default: break;
    }
    if (maxm <= 8) f(integral_constant<int, 8>{}, std::false_type{});
    else if (maxm <= 16) f(integral_constant<int, 16>{}, std::false_type{});
    else f(integral_constant<int, 32>{}, std::false_type{});
}

// Power-path shapes (NB0, NB1) with kernels built for them: terms per
// lambda of the first and second eta.
constexpr int kPowShapes[][2] = {{3, 0}, {3, 3}, {5, 0}, {5, 1}, {5, 5}};
constexpr int kNumPowShapes = sizeof(kPowShapes) / sizeof(kPowShapes[0]);

// b with zeta == 2^b, 0 <= b <= 4; -1 otherwise.
int pow2_exponent(double zeta)
{
    for (int b = 0; b <= 4; ++b)
        if (zeta == (double)(1 << b)) return b;
    return -1;
}

// Calls f(MAXM, FAST, NB0, NB1) for the SF/force kernel instance of the
// model: the power path when it applies, else as dispatch_members.
template <typename F>
void dispatch_angular(const MixedModel& mm, int maxm, F&& f)
{
    using std::integral_constant;
    using I0 = integral_constant<int, 0>;
    using I1 = integral_constant<int, 1>;
    using I3 = integral_constant<int, 3>;
    using I4 = integral_constant<int, 4>;
    using I5 = integral_constant<int, 5>;
    if (mm.powNB[0] > 0) {
        const std::true_type fast{};
        switch (mm.powNB[0] * 10 + mm.powNB[1]) {
        case 30: f(I4{}, fast, I3{}, I0{}); return;
        case 33: f(I4{}, fast, I3{}, I3{}); return;
        case 50: f(I4{}, fast, I5{}, I0{}); return;
        case 51: f(I4{}, fast, I5{}, I1{}); return;
        case 55: f(I4{}, fast, I5{}, I5{}); return;
        default: throw Error{NNPX_ERR_STATE, "no kernel built for this power-path shape"};
        }
    }
    dispatch_members(mm.fastM, maxm, [&](auto m, auto fs) { f(m, fs, I0{}, I0{}); });
}

int pow2_members(const ModelT& t)
{
    const int m = std::max(std::max(t.max_mem_ang, t.max_mem_rad), 1);
    return m <= 4 ? 4 : m <= 8 ? 8 : m <= 16 ? 16 : 32;
}

}  // namespace

void setup_mixed(MixedModel& mm, const HostModel& hm)
{
    free_mixed(mm);
    const ModelT& t = hm.t;
    std::vector<MemberF> mf(std::max(t.nmembers, 1));
    std::vector<GroupMixF> gmix(std::max(t.ngroups, 1));
    std::vector<GroupPow64> gp64(std::max(t.ngroups, 1), GroupPow64{});
    std::vector<float> sffac((size_t)kMaxElem * kMaxSF, 0.f);
    mm.maxS = 1;
    mm.hasRs = false;
    for (int e = 0; e < t.nelem; ++e) {
        const ElementT& E = t.el[e];
        for (int g = E.grp0; g < E.grp0 + E.ngrp; ++g) {
            const GroupT& G = t.grp[g];
            GroupMixF& gm = gmix[g];
            gm = GroupMixF{};
            for (int q = G.mem0; q < G.mem0 + G.nmem; ++q) {
                const MemberT& M = t.mem[q];
                MemberF& f = mf[q];
                f.eta = (float)M.eta;
                f.rs = (float)M.rs;
                f.lambda = (float)M.lambda;
                f.zeta = (float)M.zeta;
                f.gnorm = (G.type == 3 || G.type == 9 || G.type == 13) ? (float)M.gnorm : 1.f;
                f.twoEta = (float)(2.0 * M.eta);
                f.zetaLambda = (float)M.zetaLambda;
                f.etaL = (float)(M.eta * 1.4426950408889634);  // eta * log2(e)
                f.sf = M.sf;
                f.calcExp = M.calcExp;
                f.sidx = 0;
                f.beta = -1;
                if (G.type == 3 || G.type == 9 || G.type == 13) {
                    if (M.rs > 0.0) mm.hasRs = true;
                    int k = 0;
                    while (k < gm.nlam && gm.lam[k] != f.lambda) ++k;
                    if (k == gm.nlam) {
                        if (gm.nlam == 4) throw Error{NNPX_ERR_UNSUPPORTED, "more than 4 distinct lambda per SF group"};
                        gm.lam[gm.nlam++] = f.lambda;
                    }
                    f.sidx = k;
                }
                sffac[(size_t)e * kMaxSF + M.sf] =
                    (float)((G.type == 3 || G.type == 9 || G.type == 13) ? M.factorNorm : M.sfac);
            }
            mm.maxS = std::max(mm.maxS, gm.nlam);
            // unused table slots repeat the first lambda (evaluated, never picked)
            for (int k = gm.nlam; k < 4; ++k) gm.lam[k] = gm.nlam ? gm.lam[0] : 0.f;
        }
    }
    mm.maxS = mm.maxS <= 2 ? 2 : 4;
    // fast path: all angular groups share one member count, <= 2 lambda, rs = 0
    mm.fastM = 0;
    if (mm.maxS == 2 && !mm.hasRs) {
        int cnt = -1;
        bool same = true;
        for (int g = 0; g < t.ngroups; ++g)
            if (t.grp[g].type == 3 || t.grp[g].type == 9 || t.grp[g].type == 13) {
                if (cnt < 0) cnt = t.grp[g].nmem;
                same = same && t.grp[g].nmem == cnt;
            }
        if (same && (cnt == 4 || cnt == 8 || cnt == 10 || cnt == 16)) mm.fastM = cnt;
    }
    // power path: every angular group has zeta in {1, 2, 4, 8, 16}, at most
    // two distinct eta and two lambda, rs = 0
    mm.powNB[0] = mm.powNB[1] = 0;
    if (mm.maxS == 2 && !mm.hasRs) {
        bool ok = true;
        int need[2] = {0, 0};
        std::vector<std::array<double, 2>> heta(std::max(t.ngroups, 1));
        std::vector<std::array<int, 2>> hnb(std::max(t.ngroups, 1));
        for (int g = 0; g < t.ngroups && ok; ++g) {
            const GroupT& G = t.grp[g];
            if (G.type != 3 && G.type != 9 && G.type != 13) continue;
            int nh = 0;
            int nb[2] = {0, 0};
            double eta[2] = {0.0, 0.0};
            for (int q = G.mem0; q < G.mem0 + G.nmem && ok; ++q) {
                const MemberT& M = t.mem[q];
                int h = 0;
                while (h < nh && eta[h] != M.eta) ++h;
                if (h == nh) {
                    if (nh == 2) { ok = false; break; }
                    eta[nh++] = M.eta;
                }
                const int b = pow2_exponent(M.zeta);
                if (b < 0) { ok = false; break; }
                nb[h] = std::max(nb[h], b + 1);
            }
            if (!ok) break;
            // the eta with the higher powers takes the larger block
            if (nb[1] > nb[0]) {
                std::swap(nb[0], nb[1]);
                std::swap(eta[0], eta[1]);
            }
            heta[g] = {eta[0], eta[1]};
            hnb[g] = {nb[0], nb[1]};
            need[0] = std::max(need[0], nb[0]);
            need[1] = std::max(need[1], nb[1]);
        }
        int best = -1;
        for (int k = 0; k < kNumPowShapes; ++k)
            if (kPowShapes[k][0] >= need[0] && kPowShapes[k][1] >= need[1] &&
                (best < 0 || kPowShapes[k][0] + kPowShapes[k][1] < kPowShapes[best][0] + kPowShapes[best][1]))
                best = k;
        if (ok && need[0] > 0 && best >= 0) {
            const int NB0 = kPowShapes[best][0], NB1 = kPowShapes[best][1];
            mm.powNB[0] = NB0;
            mm.powNB[1] = NB1;
            for (int g = 0; g < t.ngroups; ++g) {
                const GroupT& G = t.grp[g];
                if (G.type != 3 && G.type != 9 && G.type != 13) continue;
                GroupMixF& gm = gmix[g];
                GroupPow64& gp = gp64[g];
                for (int h = 0; h < 2; ++h) {
                    gm.etaL[h] = (float)(heta[g][h] * 1.4426950408889634);
                    gm.twoEta[h] = (float)(2.0 * heta[g][h]);
                    gp.eta[h] = heta[g][h];
                    gp.twoEta[h] = 2.0 * heta[g][h];
                }
                gm.used = 0;
                for (int q = G.mem0; q < G.mem0 + G.nmem; ++q) {
                    const MemberT& M = t.mem[q];
                    const int h = M.eta == heta[g][0] ? 0 : 1;
                    const int b = pow2_exponent(M.zeta);
                    mf[q].beta = h == 0 ? mf[q].sidx * NB0 + b : 2 * NB0 + mf[q].sidx * NB1 + b;
                    gm.used |= 1 << mf[q].beta;
                    gp.lam[mf[q].sidx] = M.lambda;
                }
                gp.used = gm.used;
            }
        }
    }
    // merged angular walks: per element, every angular group on the power
    // path, one type (3 or 9), the element's single cutoff radius, identical
    // term constants and distinct neighbour element pairs
    std::vector<AngMergeF> amerge(std::max(t.nelem, 1));
    mm.merge_ng = 0;
    mm.merge_avg = 0.0;
    int merged_elements = 0;
    for (int e = 0; e < t.nelem && mm.powNB[0] > 0; ++e) {
        AngMergeF& am = amerge[e];
        std::memset(&am, 0, sizeof(am));
        std::memset(am.pair, -1, sizeof(am.pair));
        const ElementT& E = t.el[e];
        std::vector<int> ang;
        for (int gi = 0; gi < E.ngrp; ++gi) {
            const int g = E.grp0 + gi;
            const int ty = t.grp[g].type;
            if (ty == 3 || ty == 9 || ty == 13) ang.push_back(g);
        }
        bool ok = E.ucut && ang.size() >= 2 && (int)ang.size() <= kMaxMerge;
        for (size_t q = 0; q < ang.size() && ok; ++q) {
            const GroupT& G = t.grp[ang[q]];
            const GroupT& G0 = t.grp[ang[0]];
            const GroupMixF& a = gmix[ang[q]];
            const GroupMixF& a0 = gmix[ang[0]];
            ok = G.type != 13 && G.type == G0.type && G.rc == G0.rc && a.nlam == a0.nlam && a.lam[0] == a0.lam[0] &&
                 a.lam[1] == a0.lam[1] && a.etaL[0] == a0.etaL[0] && a.etaL[1] == a0.etaL[1] &&
                 a.twoEta[0] == a0.twoEta[0] && a.twoEta[1] == a0.twoEta[1] &&
                 am.pair[G.e1 * kMaxElem + G.e2] < 0;
            if (!ok) break;
            am.grp[q] = ang[q];
            am.used |= a.used;
            am.pair[G.e1 * kMaxElem + G.e2] = (signed char)q;
            am.pair[G.e2 * kMaxElem + G.e1] = (signed char)q;
        }
        if (!ok) {
            std::memset(&am, 0, sizeof(am));
            std::memset(am.pair, -1, sizeof(am.pair));
            continue;
        }
        am.ng = (int)ang.size();
        mm.merge_ng = std::max(mm.merge_ng, am.ng);
        mm.merge_avg += am.ng;
        ++merged_elements;
    }
    if (merged_elements) mm.merge_avg /= merged_elements;
    // angular walks per element (an upper bound: every non-radial group, and
    // the merged walk)
    mm.max_walks = 0;
    for (int e = 0; e < t.nelem; ++e) {
        int nw = 1;
        for (int gi = 0; gi < t.el[e].ngrp; ++gi) {
            const int ty = t.grp[t.el[e].grp0 + gi].type;
            if (ty != 2 && ty != 12) ++nw;
        }
        mm.max_walks = std::max(mm.max_walks, nw);
    }
    // bf16 weights, padded to 16: W [out][in] and WT [in][out], hi then lo.
    std::vector<uint16_t> w;
    std::vector<float> b;
    auto pad16 = [](int x) { return (x + 15) / 16 * 16; };
    mm.maxpad = 16;
    mm.maxpadh = 16;
    mm.ldK = 16;
    for (int e = 0; e < t.nelem; ++e) {
        const NNT& N = t.el[e].nn;
        mm.ldK = std::max(mm.ldK, pad16(N.n[0]));
        for (int l = 1; l < N.nl; ++l) {
            const int nin = N.n[l - 1], nout = N.n[l];
            const int pi = pad16(nin), po = pad16(nout);
            mm.pin[e][l] = pi;
            mm.pout[e][l] = po;
            mm.maxpad = std::max(mm.maxpad, std::max(pi, po));
            mm.maxpadh = std::max(mm.maxpadh, po);
            std::vector<float> W((size_t)po * pi, 0.f);
            for (int k = 0; k < nout; ++k)
                for (int j = 0; j < nin; ++j) W[(size_t)k * pi + j] = (float)hm.weights[N.woff[l] + (size_t)k * nin + j];
            mm.wW[e][l] = (int)w.size();
            std::vector<uint16_t> hi(W.size()), lo(W.size());
            for (size_t q = 0; q < W.size(); ++q) {
                hi[q] = to_bf16(W[q]);
                lo[q] = to_bf16(W[q] - from_bf16(hi[q]));
            }
            w.insert(w.end(), hi.begin(), hi.end());
            w.insert(w.end(), lo.begin(), lo.end());
            mm.wT[e][l] = (int)w.size();
            for (int j = 0; j < pi; ++j)
                for (int k = 0; k < po; ++k) hi[(size_t)j * po + k] = to_bf16(W[(size_t)k * pi + j]);
            for (int j = 0; j < pi; ++j)
                for (int k = 0; k < po; ++k) {
                    const float x = W[(size_t)k * pi + j];
                    lo[(size_t)j * po + k] = to_bf16(x - from_bf16(to_bf16(x)));
                }
            w.insert(w.end(), hi.begin(), hi.end());
            w.insert(w.end(), lo.begin(), lo.end());
            mm.bo[e][l] = (int)b.size();
            for (int k = 0; k < po; ++k) b.push_back(k < nout ? (float)hm.weights[N.boff[l] + k] : 0.f);
        }
    }
    auto up = [](void** dst, const void* src, size_t bytes) {
        if (hipMalloc(dst, std::max<size_t>(bytes, 16)) != hipSuccess) throw Error{NNPX_ERR_HIP, "hipMalloc mixed"};
        if (hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) throw Error{NNPX_ERR_HIP, "hipMemcpy"};
    };
    up((void**)&mm.d_mem, mf.data(), sizeof(MemberF) * mf.size());
    up((void**)&mm.d_gmix, gmix.data(), sizeof(GroupMixF) * gmix.size());
    up((void**)&mm.d_gpow64, gp64.data(), sizeof(GroupPow64) * gp64.size());
    up((void**)&mm.d_amerge, amerge.data(), sizeof(AngMergeF) * amerge.size());
    up((void**)&mm.d_sffac, sffac.data(), sizeof(float) * sffac.size());
    up((void**)&mm.d_w, w.data(), sizeof(uint16_t) * w.size());
    up((void**)&mm.d_b, b.data(), sizeof(float) * b.size());
    mm.ready = true;
}

void free_mixed(MixedModel& mm)
{
    if (mm.d_mem) (void)hipFree(mm.d_mem);
    if (mm.d_gmix) (void)hipFree(mm.d_gmix);
    if (mm.d_gpow64) (void)hipFree(mm.d_gpow64);
    if (mm.d_amerge) (void)hipFree(mm.d_amerge);
    if (mm.d_sffac) (void)hipFree(mm.d_sffac);
    if (mm.d_w) (void)hipFree(mm.d_w);
    if (mm.d_b) (void)hipFree(mm.d_b);
    const int terms = mm.terms;
    const Tuning tune = mm.tune;
    mm = MixedModel{};
    mm.terms = terms;
    mm.tune = tune;
}

// Resident blocks per CU of a kernel instance at a dynamic LDS size, from the
// HIP occupancy calculator, cached per (instance, size).
static int blocks_per_cu(const void* kernel, size_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kernel, lds});
    if (it != cache.end()) return it->second;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kWave, lds) != hipSuccess)
        throw Error{NNPX_ERR_HIP, "hipOccupancyMaxActiveBlocksPerMultiprocessor"};
    cache[{kernel, lds}] = occ;
    return occ;
}

// Merged angular walk on or off: a tuning value of 0 or 1 decides, -1 takes
// it where the model merges at least four groups per element on average
// (fewer groups have enough triplets each to fill the per-group walks).
static bool use_merge(const MixedModel& mm, int want)
{
    if (mm.merge_ng == 0 || mm.powNB[0] == 0) return false;
    if (want >= 0) return want == 1;
    return mm.merge_avg >= 4.0;
}

void run_mixed(const MixedModel& mm, const ModelT& t, const ModelT* d_model, const SysView& s,
               const MixedBuffers& b, bool forces, bool push, hipStream_t stream, const TimedFn& timed)
{
    if (!mm.ready) throw Error{NNPX_ERR_STATE, "mixed-precision tables not set up"};
    if (s.natoms == 0) return;
    const int capL = s.stage_cap;
    const int P = pow2_members(t);
    const int tsf = mm.tune.trip_sf, tf = mm.tune.trip_force;
    // LDS of the largest kernel instances (the member paths' constant area)
    const size_t lds_sf_max = stage_bytes(capL, tsf, stage_mp(0, 0));
    const size_t lds_f_max = stage_bytes(capL, tf, stage_mp(0, 0)) + (sizeof(double) * 3 + 2) * (size_t)capL;
    if (lds_f_max > 65536 || lds_sf_max > 65536)
        throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed kernels' LDS"};
    const bool merge_sf = use_merge(mm, mm.tune.merge_sf), merge_f = use_merge(mm, mm.tune.merge_force);
    // with pair lists on, the angular pairs pass from the SF to the force
    // kernel when both make the same walks (the SF kernel also counts them,
    // which sizes the lists of later evaluations); off, the force kernel
    // walks rows
    const bool lists = mm.tune.pair_lists == 1;
    TripList T{};
    if (forces && lists && merge_sf == merge_f) {
        T = b.trips;
        T.count = true;
    }
    T.rows = !lists;
    // the staging passes from the SF to the force kernel (identical values)
    const StageKeep KS = forces && b.keep.ld >= capL ? b.keep : StageKeep{};
    timed("sf32", [&] {
        dispatch_angular(mm, P, [&](auto m, auto fast, auto nb0, auto nb1) {
            constexpr int cm = decltype(m)::value, c0 = decltype(nb0)::value, c1 = decltype(nb1)::value;
            constexpr bool cf = decltype(fast)::value;
            constexpr bool cmg = c0 > 0;  // merged instances exist on the power path only
            const bool mg = cmg && merge_sf;
            const size_t extra = mg ? merge_bytes(mm.merge_ng, 2 * (c0 + c1)) : 0;
            // the partner gather costs 16 B of LDS per slot: taken only when
            // it leaves the blocks per CU unchanged
            const size_t lds0 = stage_bytes(capL, tsf, stage_mp(c0, c1)) + extra;
            const size_t lds1 = lds0 + sizeof(float4) * (size_t)capL;
            if (lds0 > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed kernels' LDS"};
            auto launch = [&](auto gather, auto merge) {
                constexpr bool cg = decltype(gather)::value, cmm = decltype(merge)::value;
                hipLaunchKernelGGL((k_sf32<cm, cf, c0, c1, cg, cmm>), dim3(s.natoms), dim3(kWave), cg ? lds1 : lds0,
                                   stream, d_model, mm.d_mem, mm.d_gmix, mm.d_amerge, s, b.G, b.Gin, b.ldG, mm.ldK,
                                   capL, tsf, T, KS);
            };
            auto pick = [&](auto merge) {
                constexpr bool cmm = decltype(merge)::value;
                int occ0 = 0, occ1 = 0;
                const int want = mm.tune.sf_gather;
                if (lds1 <= 65536 && want != 0) {
                    occ0 = blocks_per_cu(reinterpret_cast<const void*>(&k_sf32<cm, cf, c0, c1, false, cmm>), lds0);
                    occ1 = blocks_per_cu(reinterpret_cast<const void*>(&k_sf32<cm, cf, c0, c1, true, cmm>), lds1);
                }
                if (occ1 > 0 && (occ1 >= occ0 || want == 1))
                    launch(std::true_type{}, merge);
                else
                    launch(std::false_type{}, merge);
            };
            if (mg) pick(std::integral_constant<bool, cmg>{});
            else pick(std::false_type{});
        });
        launch_sfcomp32(d_model, t, s, b.G, b.Gin, b.ldG, mm.ldK, nullptr, nullptr, false, false, stream);
    });
    NNMix a{};
    a.w = mm.d_w;
    a.b = mm.d_b;
    a.sffac = mm.d_sffac;
    a.atoms = b.elem_atoms;
    a.nelem = t.nelem;
    a.ldK = mm.ldK;
    a.ldG = b.ldG;
    a.stride = mm.maxpadh + 4;
    a.stride_in = mm.ldK + 4;
    a.terms = mm.terms;
    int blocks = 0;
    for (int e = 0; e < t.nelem; ++e) {
        a.blk_off[e] = blocks;
        a.atom_off[e] = b.elem_off[e];
        blocks += (b.elem_off[e + 1] - b.elem_off[e] + kNNRows - 1) / kNNRows;
        for (int l = 0; l < kMaxLayers; ++l) {
            a.wW[e][l] = mm.wW[e][l];
            a.wT[e][l] = mm.wT[e][l];
            a.bo[e][l] = mm.bo[e][l];
            a.pin[e][l] = mm.pin[e][l];
            a.pout[e][l] = mm.pout[e][l];
        }
    }
    a.blk_off[t.nelem] = blocks;
    a.atom_off[t.nelem] = b.elem_off[t.nelem];
    const size_t lds_nn = sizeof(float) * (size_t)kNNRows * (a.stride_in + (size_t)a.stride * (2 + kMaxLayers));
    if (lds_nn > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "network too wide for the mixed NN kernel's LDS"};
    // the NN blocks clear the forces of the force kernel (one launch fewer)
    const size_t n3 = forces ? (size_t)3 * s.nforce : 0;
    const bool nn_zero = blocks > 0;
    a.zero = nn_zero ? b.F : nullptr;
    a.nzero = nn_zero ? n3 : 0;
    timed("nnbf16", [&] {
        hipLaunchKernelGGL(k_nnbf16, dim3(blocks), dim3(kWave * kNNWaves), lds_nn, stream, d_model, a, b.Gin, b.Eatom,
                           b.dEdG, b.dEdGf);
    });
    timed("energy", [&] { launch_energy(b.Eatom, b.atom_off, b.nstruct, b.E, stream); });
    if (!forces) return;
    timed("force32", [&] {
        if (!nn_zero && n3 > 0)
            hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>((n3 + 255) / 256, 4096)), dim3(256), 0,
                               stream, b.F, n3);
        const int ps = push ? 1 : 0;
        dispatch_angular(mm, P, [&](auto m, auto fast, auto nb0, auto nb1) {
            constexpr int cm = decltype(m)::value, c0 = decltype(nb0)::value, c1 = decltype(nb1)::value;
            constexpr bool cf = decltype(fast)::value;
            const bool mg = c0 > 0 && merge_f;
            const size_t lds_f = stage_bytes(capL, tf, stage_mp(c0, c1)) + (sizeof(double) * 3 + 2) * (size_t)capL +
                                 (mg ? merge_bytes(mm.merge_ng, 2 * (c0 + c1)) : 0);
            if (lds_f > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed kernels' LDS"};
            if (mg)
                hipLaunchKernelGGL((k_force32<cm, cf, c0, c1, (c0 > 0)>), dim3(s.natoms), dim3(kWave), lds_f, stream,
                                   d_model, mm.d_mem, mm.d_gmix, mm.d_amerge, s, b.dEdGf, b.ldG, b.F, ps, capL, tf, T,
                                   KS);
            else
                hipLaunchKernelGGL((k_force32<cm, cf, c0, c1, false>), dim3(s.natoms), dim3(kWave), lds_f, stream,
                                   d_model, mm.d_mem, mm.d_gmix, mm.d_amerge, s, b.dEdGf, b.ldG, b.F, ps, capL, tf, T,
                                   KS);
        });
        launch_sfcomp32(d_model, t, s, nullptr, nullptr, b.ldG, mm.ldK, b.dEdGf, b.F, push, true, stream);
    });
}

void launch_tangent32(const MixedModel& mm, const ModelT& t, const ModelT* d_model, const SysView& s,
                      const double* u, double* Gdot, int ldG, hipStream_t stream)
{
    if (!mm.ready) throw Error{NNPX_ERR_STATE, "mixed-precision tables not set up"};
    if (s.natoms == 0) return;
    const int capL = s.stage_cap;
    const int tt = mm.tune.trip_tangent;
    const size_t lds = stage_bytes(capL, tt, stage_mp(0, 0)) + (sizeof(float4) + 2) * (size_t)capL;
    if (lds > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "neighbour count too large for the mixed kernels' LDS"};
    dispatch_members(mm.fastM, pow2_members(t), [&](auto m, auto fast) {
        hipLaunchKernelGGL((k_tangent32<decltype(m)::value, decltype(fast)::value>), dim3(s.natoms), dim3(kWave), lds,
                           stream, d_model, mm.d_mem, mm.d_gmix, s, u, mm.d_sffac, Gdot, ldG, capL, tt);
    });
    launch_tangentcomp32(d_model, t, s, u, Gdot, ldG, stream);
}

double mfma_selftest()
{
    std::vector<float> A(256), B(256), D(256, 0.f);
    for (int i = 0; i < 16; ++i)
        for (int k = 0; k < 16; ++k) {
            A[i * 16 + k] = (float)((i * 7 + k * 3) % 11) - 5.f;
            B[i * 16 + k] = (float)((i * 5 + k * 13) % 9) - 4.f;
        }
    float *dA, *dB, *dD;
    if (hipMalloc(&dA, 1024) != hipSuccess || hipMalloc(&dB, 1024) != hipSuccess || hipMalloc(&dD, 1024) != hipSuccess)
        throw Error{NNPX_ERR_HIP, "hipMalloc"};
    (void)hipMemcpy(dA, A.data(), 1024, hipMemcpyHostToDevice);
    (void)hipMemcpy(dB, B.data(), 1024, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dA, dB, dD);
    (void)hipMemcpy(D.data(), dD, 1024, hipMemcpyDeviceToHost);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dD);
    double err = 0.0;
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            double ref = 0.0;
            for (int k = 0; k < 16; ++k) ref += (double)A[i * 16 + k] * B[k * 16 + j];
            err = std::max(err, std::fabs(ref - D[i * 16 + j]));
        }
    return err;
}

}  // namespace nnpx
