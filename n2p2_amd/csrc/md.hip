// Velocity-Verlet molecular dynamics around the energy/force path: the
// integrator a driver such as LAMMPS would run (fix nve), with the
// neighbour list rebuilt on the device when an atom has moved more than half
// the skin.  Every step is a fixed sequence of kernels whose rebuild kernels
// are gated by a device flag, so a run of steps can be captured in a HIP
// graph and replayed without host decisions.
#include <hip/hip_runtime.h>

#include "md.hpp"

namespace nnpx {
namespace {

constexpr int kBlock = 256;

// v += hk_e F ; x += dt v   (first half kick and drift)
__global__ void k_kick_drift(double* __restrict__ pos, double* __restrict__ vel, const double* __restrict__ F,
                             const int* __restrict__ elem, MDConst c, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hk = c.hk[elem[i]];
    for (int k = 0; k < 3; ++k) {
        const size_t q = 3 * (size_t)i + k;
        const double v = vel[q] + hk * F[q];
        vel[q] = v;
        pos[q] += c.dt * v;
    }
}

// Rebuild decision: any atom displaced by more than half the skin since the
// last build sets the gate (every writer stores the same value).
__global__ void k_check_displacement(const double* __restrict__ pos, const double* __restrict__ ref, int n,
                                     double lim2, int* gate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double d = pos[3 * (size_t)i + k] - ref[3 * (size_t)i + k];
        d2 += d * d;
    }
    if (d2 > lim2) *gate = 1;
}

// Gated: wrap periodic positions back into the cell (the rules of
// Structure::remap, Structure.cpp:1224-1248) and record the reference
// positions of the new list.
__global__ void k_wrap(double* __restrict__ pos, double* __restrict__ ref, const int* __restrict__ sid,
                       const StructT* __restrict__ st, int n, const int* gate)
{
#pragma clang fp contract(off)
    if (*gate == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* r = pos + 3 * (size_t)i;
    const StructT& S = st[sid[i]];
    if (S.periodic) {
        double f[3];
        for (int k = 0; k < 3; ++k) {
            double x = r[0] * S.invbox[k] + r[1] * S.invbox[3 + k] + r[2] * S.invbox[6 + k];
            if (x > 1.0) x -= (int)x;
            if (x < 0.0) x += 1.0 - (int)x;
            if (x == 1.0) x = 0.0;
            f[k] = x;
        }
        for (int k = 0; k < 3; ++k) r[k] = f[0] * S.box[k] + f[1] * S.box[3 + k] + f[2] * S.box[6 + k];
    }
    for (int k = 0; k < 3; ++k) ref[3 * (size_t)i + k] = r[k];
}

// v += hk_e F (second half kick) and the kinetic energy of the step,
// accumulated in FP64 into the step's slot.
__global__ __launch_bounds__(kBlock) void k_kick(double* __restrict__ vel, const double* __restrict__ F,
                                                 const int* __restrict__ elem, MDConst c, int n,
                                                 const int* __restrict__ counters, double* __restrict__ ekin)
{
    __shared__ double part[kBlock / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (i < n) {
        const int e = elem[i];
        const double hk = c.hk[e];
        double v2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const size_t q = 3 * (size_t)i + k;
            const double v = vel[q] + hk * F[q];
            vel[q] = v;
            v2 += v * v;
        }
        ke = c.half_m[e] * v2;
    }
    ke = wave_sum(ke);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ke;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += part[w];
        atomicAdd(ekin + (counters[0] & (kMDHistory - 1)), t);
    }
}

// End of step (one thread): potential energy of the step, rebuild count,
// gate reset, next kinetic slot cleared, step counter advanced.
__global__ void k_step_end(int* gate, int* counters, const double* __restrict__ E, double* epot, double* ekin)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int step = counters[0];
    epot[step & (kMDHistory - 1)] = E[0];
    if (*gate) {
        counters[1] += 1;
        *gate = 0;
    }
    ekin[(step + 1) & (kMDHistory - 1)] = 0.0;
    counters[0] = step + 1;
}

__global__ void k_md_reset(int* gate, int* counters, double* ekin)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *gate = 0;
    counters[0] = 0;
    counters[1] = 0;
    ekin[0] = 0.0;
}

__global__ void k_set_gate(int* gate, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *gate = v;
}

}  // namespace

void md_kick_drift(double* pos, double* vel, const double* F, const int* elem, const MDConst& c, int n,
                   hipStream_t stream)
{
    hipLaunchKernelGGL(k_kick_drift, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, pos, vel, F, elem, c,
                       n);
}

void md_check(const double* pos, const double* ref, int n, double lim2, int* gate, hipStream_t stream)
{
    hipLaunchKernelGGL(k_check_displacement, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, pos, ref, n,
                       lim2, gate);
}

void md_wrap(double* pos, double* ref, const int* sid, const StructT* st, int n, const int* gate,
             hipStream_t stream)
{
    hipLaunchKernelGGL(k_wrap, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, pos, ref, sid, st, n,
                       gate);
}

void md_kick(double* vel, const double* F, const int* elem, const MDConst& c, int n, const int* counters,
             double* ekin, hipStream_t stream)
{
    hipLaunchKernelGGL(k_kick, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, vel, F, elem, c, n,
                       counters, ekin);
}

void md_step_end(int* gate, int* counters, const double* E, double* epot, double* ekin, hipStream_t stream)
{
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, stream, gate, counters, E, epot, ekin);
}

void md_reset(int* gate, int* counters, double* ekin, hipStream_t stream)
{
    hipLaunchKernelGGL(k_md_reset, dim3(1), dim3(64), 0, stream, gate, counters, ekin);
}

void md_set_gate(int* gate, int v, hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_gate, dim3(1), dim3(64), 0, stream, gate, v);
}

}  // namespace nnpx
