// Per-evaluation checks of the reference's symmetry-function driver
// (Mode::calculateSymmetryFunctionGroups, Mode.cpp:1630-1709):
//  * extrapolation warnings, Element::updateSymmetryFunctionStatistics
//    (Element.cpp:462-541): an SF whose unscaled value (SymFnc::unscale,
//    SymFnc.cpp:197-223) lies below Gmin or above Gmax of scaling.data by more
//    than 1000 eps;
//  * the low-neighbour check (Mode.cpp:1667-1679): fewer neighbours within
//    the element's smallest SF cutoff (Element::getMinCutoffRadius) than its
//    SFs need (Element::getMinNeighbors: 1 radial, 2 angular).
// One thread per (atom, SF) or per atom; hits are appended through an atomic
// counter (pass 1 counts, pass 2 fills a buffer of that size) and ordered on
// the host.
#include <cfloat>

#include "kernels.hpp"

namespace nnpx {
namespace {

__device__ inline double unscale_sf(const ScaleT& s, double v)
{
    switch (s.type) {
    case ST_NONE: return v;
    case ST_SCALE: return (v - s.Smin) / s.factor + s.Gmin;
    case ST_CENTER: return v + s.Gmean;
    default: return (v - s.Smin) / s.factor + s.Gmean;
    }
}

__global__ void k_extrapolation(const ModelT* __restrict__ M, int natoms, const int* __restrict__ elem,
                                const double* __restrict__ G, int ldG, unsigned long long* __restrict__ count,
                                long long cap, int* __restrict__ atom, int* __restrict__ sf, double* __restrict__ val)
{
    const size_t total = (size_t)natoms * ldG;
    const double eps = 1000.0 * DBL_EPSILON;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(x / ldG);
        const int q = (int)(x - (size_t)i * ldG);
        const ElementT& E = M->el[elem[i]];
        if (q >= E.nsf) continue;
        const ScaleT& S = M->scale[E.sf0 + q];
        const double v = unscale_sf(S, G[x]);
        if (!(v + eps < S.Gmin || v - eps > S.Gmax)) continue;
        const unsigned long long p = atomicAdd(count, 1ull);
        if ((long long)p < cap) {
            atom[p] = i;
            sf[p] = q;
            val[p] = v;
        }
    }
}

__global__ __launch_bounds__(64) void k_low_neighbors(const ModelT* __restrict__ M, SysView s,
                                                      const double* __restrict__ rmin, const int* __restrict__ nmin,
                                                      unsigned long long* __restrict__ count, long long cap,
                                                      int* __restrict__ atom, int* __restrict__ num)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const StructT& St = s.st[s.sid[i]];
    const size_t base = (size_t)i * s.cap;
    int c = 0;
    for (int k = lane; k < s.nb_n[i]; k += kWave) {
        double dr[3];
        entry_vector(s, i, base + k, St, dr);
        if (sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]) <= rmin[e]) ++c;
    }
    c = wave_sum(c);
    if (lane == 0 && c < nmin[e]) {
        const unsigned long long p = atomicAdd(count, 1ull);
        if ((long long)p < cap) {
            atom[p] = i;
            num[p] = c;
        }
    }
}

}  // namespace

void launch_extrapolation(const ModelT* m, int natoms, const int* elem, const double* G, int ldG,
                          unsigned long long* count, long long cap, int* atom, int* sf, double* val, hipStream_t stream)
{
    if (natoms == 0) return;
    const size_t total = (size_t)natoms * ldG;
    hipLaunchKernelGGL(k_extrapolation, dim3((unsigned)std::min<size_t>((total + 255) / 256, 8192)), dim3(256), 0,
                       stream, m, natoms, elem, G, ldG, count, cap, atom, sf, val);
}

void launch_low_neighbors(const ModelT* m, const SysView& s, const double* rmin, const int* nmin,
                          unsigned long long* count, long long cap, int* atom, int* num, hipStream_t stream)
{
    if (s.natoms == 0) return;
    hipLaunchKernelGGL(k_low_neighbors, dim3(s.natoms), dim3(kWave), 0, stream, m, s, rmin, nmin, count, cap, atom,
                       num);
}

}  // namespace nnpx
