// Device-side helpers shared by the gfx950 kernels: cutoff functions,
// integer powers, SF scaling and wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "model.hpp"

namespace nnpx {

constexpr int kWave = 64;

// Per-structure geometry, device copy.
struct StructT {
    double box[9];     // lattice vectors a, b, c as rows
    double invbox[9];  // Structure::calculateInverseBox (Structure.cpp:518-544)
    int pbc[3];        // Structure::calculatePbcCopies (Structure.cpp:485-516)
    int periodic;
    int nc[3];         // cells per axis for the cell list
    int algo;          // 0 brute force, 1 cell list
    int atom0, natoms;
    int cell0;         // first cell of this structure in the global cell arrays
    int inside;        // every atom has fractional coordinates in [0, 1)
};

// Image shift packing: 7 bits per axis, biased by 64, so that the packed
// value orders like (bc0, bc1, bc2) lexicographically.
__host__ __device__ inline int pack_image(int b0, int b1, int b2)
{
    return ((b0 + 64) << 14) | ((b1 + 64) << 7) | (b2 + 64);
}
__host__ __device__ inline void unpack_image(int c, int& b0, int& b1, int& b2)
{
    b0 = ((c >> 14) & 127) - 64;
    b1 = ((c >> 7) & 127) - 64;
    b2 = (c & 127) - 64;
}
__host__ __device__ inline int negate_image(int c)
{
    int b0, b1, b2;
    unpack_image(c, b0, b1, b2);
    return pack_image(-b0, -b1, -b2);
}

// Distance vector of neighbour (j, image) of atom i in the reference's
// evaluation order ((r_i - r_j) + bc0*a) + bc1*b + bc2*c
// (Structure.cpp:322-325).  Contraction is disabled so the result is
// bit-identical to the host expression.
__device__ inline void pair_vector(const double* __restrict__ ri, const double* __restrict__ rj,
                                   const double* __restrict__ box, int b0, int b1, int b2,
                                   double* dr)
{
#pragma clang fp contract(off)
    const double f0 = (double)b0, f1 = (double)b1, f2 = (double)b2;
    for (int k = 0; k < 3; ++k)
        dr[k] = ri[k] - rj[k] + f0 * box[k] + f1 * box[3 + k] + f2 * box[6 + k];
}

// Core functions on x in [0, 1) (CoreFunction.h:214-319): 0 exponential
// e*exp(1/(x^2 - 1)), 1..4 the polynomials of increasing order.
__device__ inline void core_fdf(int core, double x, double& f, double& df)
{
    if (core == 0) {
        double const t = 1.0 / (x * x - 1.0);
        double const t2 = exp(t);
        f = M_E * t2;
        df = -2.0 * M_E * x * t * t * t2;
    } else if (core == 1) {
        f = (2.0 * x - 3.0) * x * x + 1.0;
        df = x * (6.0 * x - 6.0);
    } else if (core == 2) {
        double const x2 = x * x;
        f = ((15.0 - 6.0 * x) * x - 10.0) * x * x2 + 1.0;
        df = x2 * ((60.0 - 30.0 * x) * x - 30.0);
    } else if (core == 3) {
        double const x2 = x * x;
        f = (x * (x * (20.0 * x - 70.0) + 84.0) - 35.0) * x2 * x2 + 1.0;
        df = x2 * x * (x * (x * (140.0 * x - 420.0) + 420.0) - 140.0);
    } else {
        double x4 = x * x;
        x4 *= x4;
        f = (x * (x * ((315.0 - 70.0 * x) * x - 540.0) + 420.0) - 126.0) * x * x4 + 1.0;
        df = x4 * (x * (x * ((2520.0 - 630.0 * x) * x - 3780.0) + 2520.0) - 630.0);
    }
}

// Cutoff function value and derivative (CutoffFunction::fdf,
// CutoffFunction.h:210-220, CutoffFunction.cpp:114-213, CoreFunction.h).
__device__ inline void cutoff_fdf(int type, const GroupT& g, double tanhPre, double r,
                                  double& fc, double& dfc)
{
    if (r >= g.rc) { fc = 0.0; dfc = 0.0; return; }
    switch (type) {
    case 0: fc = 1.0; dfc = 0.0; return;
    case 1: {
        if (r < g.rci) { fc = 1.0; dfc = 0.0; return; }
        double const x = (r - g.rci) * g.iw;
        double const t = cos(M_PI * x);
        fc = 0.5 * (t + 1.0);
        dfc = -0.5 * g.iw * M_PI * sqrt(1.0 - t * t);
        return;
    }
    case 2: {
        double const t = tanh(1.0 - r * g.rcinv);
        double const t2 = t * t;
        fc = t * t2;
        dfc = 3.0 * t2 * (t2 - 1.0) * g.rcinv;
        return;
    }
    case 3: {
        double const t = tanh(1.0 - r * g.rcinv);
        double const t2 = t * t;
        fc = tanhPre * t * t2;
        dfc = 3.0 * tanhPre * t2 * (t2 - 1.0) * g.rcinv;
        return;
    }
    default: break;
    }
    if (r < g.rci) { fc = 1.0; dfc = 0.0; return; }
    double const x = (r - g.rci) * g.iw;
    core_fdf(type - 4, x, fc, dfc);
    dfc *= g.iw;
}

// Compact function (CompactFunction::fdf, CompactFunction.h:150-172): the
// core at |a - cen| * scl, through x -> (2 - x) x when asymmetric
// (CoreFunction.h:197-212).
__device__ inline void compact_fdf(int core, int asym, double cen, double scl, double a, double& fa,
                                   double& dfa)
{
    a = (a - cen) * scl;
    double const x = fabs(a);
    if (asym) {
        core_fdf(core, (2.0 - x) * x, fa, dfa);
        dfa *= 2.0 - 2.0 * x;
    } else {
        core_fdf(core, x, fa, dfa);
    }
    dfa *= copysign(scl, a);
}

// FP32 forms of core_fdf / compact_fdf for the mixed form's compact terms.
__device__ inline void core_fdf(int core, float x, float& f, float& df)
{
    if (core == 0) {
        float const t = 1.f / (x * x - 1.f);
        float const t2 = __expf(t);
        f = (float)M_E * t2;
        df = -2.f * (float)M_E * x * t * t * t2;
    } else if (core == 1) {
        f = (2.f * x - 3.f) * x * x + 1.f;
        df = x * (6.f * x - 6.f);
    } else if (core == 2) {
        float const x2 = x * x;
        f = ((15.f - 6.f * x) * x - 10.f) * x * x2 + 1.f;
        df = x2 * ((60.f - 30.f * x) * x - 30.f);
    } else if (core == 3) {
        float const x2 = x * x;
        f = (x * (x * (20.f * x - 70.f) + 84.f) - 35.f) * x2 * x2 + 1.f;
        df = x2 * x * (x * (x * (140.f * x - 420.f) + 420.f) - 140.f);
    } else {
        float x4 = x * x;
        x4 *= x4;
        f = (x * (x * ((315.f - 70.f * x) * x - 540.f) + 420.f) - 126.f) * x * x4 + 1.f;
        df = x4 * (x * (x * ((2520.f - 630.f * x) * x - 3780.f) + 2520.f) - 630.f);
    }
}

__device__ inline void compact_fdf(int core, int asym, float cen, float scl, float a, float& fa, float& dfa)
{
    a = (a - cen) * scl;
    float const x = fabsf(a);
    if (asym) {
        core_fdf(core, (2.f - x) * x, fa, dfa);
        dfa *= 2.f - 2.f * x;
    } else {
        core_fdf(core, x, fa, dfa);
    }
    dfa *= copysignf(scl, a);
}

// Exponentiation by squaring (utility.cpp:292-316).
template <typename T>
__device__ inline T pow_int(T x, int n)
{
    unsigned m;
    if (n < 0) { x = T(1) / x; m = (unsigned)(-n); }
    else m = (unsigned)n;
    T r = T(1);
    do {
        if (m & 1u) r *= x;
        m >>= 1;
        x *= x;
    } while (m);
    return r;
}

// SymFnc::scale (SymFnc.cpp:169-195).
__device__ inline double scale_sf(const ScaleT& s, double v)
{
    switch (s.type) {
    case ST_NONE: return v;
    case ST_SCALE: return s.Smin + s.factor * (v - s.Gmin);
    case ST_CENTER: return v - s.Gmean;
    default: return s.Smin + s.factor * (v - s.Gmean);
    }
}

// Cross-lane moves through DPP (data-parallel primitives: operand swizzles
// inside rows of 16 lanes, no LDS traffic) and readlane, for 4- and 8-byte
// values (8-byte values move as two 32-bit halves).  Every lane of the wave
// must be active.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7 - i within each 8
constexpr int kDppMirror = 0x140;     // lane i <-> 15 - i within each 16
constexpr int kDppShr = 0x110;        // row_shr:n = kDppShr + n (zero where no source)
constexpr int kDppBcast15 = 0x142;    // lane 15 of each row to the next row
constexpr int kDppBcast31 = 0x143;    // lane 31 to rows 2 and 3

template <int CTRL, int ROWS = 0xF, bool ZERO = false, typename T>
__device__ inline T dpp_move(T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        const int r = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xF, ZERO);
        return __builtin_bit_cast(T, r);
    } else {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xF, ZERO);
        const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, ZERO);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// Value of lane l (l uniform), in a scalar register.
template <typename T>
__device__ inline T lane_value(T v, int l)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    } else {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// Sum over the wave, received by every lane: pairwise sums inside rows of 16
// (quad swaps, half-row and row mirrors), then the four row sums added in
// row order.
template <typename T>
__device__ inline T wave_sum(T v)
{
    v += dpp_move<kDppXor1>(v);
    v += dpp_move<kDppXor2>(v);
    v += dpp_move<kDppHalfMirror>(v);
    v += dpp_move<kDppMirror>(v);
    return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}

// Inclusive prefix sum over the wave: shifts by 1, 2, 4, 8 inside each row,
// then lane 15 of a row added to the next and lane 31 to the upper half.
__device__ inline int wave_incl_scan(int v)
{
    v += dpp_move<kDppShr + 1, 0xF, true>(v);
    v += dpp_move<kDppShr + 2, 0xF, true>(v);
    v += dpp_move<kDppShr + 4, 0xF, true>(v);
    v += dpp_move<kDppShr + 8, 0xF, true>(v);
    v += dpp_move<kDppBcast15, 0xA>(v);
    v += dpp_move<kDppBcast31, 0xC>(v);
    return v;
}

__device__ inline void lds_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_add(float* p, float v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

}  // namespace nnpx
