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

template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Inclusive prefix sum over the wave.
__device__ inline int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
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
