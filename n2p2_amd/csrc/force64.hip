// FP64 forces from stored dG/dr (plain form).
//
// Mode::calculateForces (Mode.cpp:1922-1984) in gather form: the force on i
// is its self term -sum_s dE/dG_i[s] dG_i/dr_i[s] (Atom::calculateSelfForceShort,
// Atom.cpp:371-379) plus, for every neighbour entry (j, image) of i, the
// contribution stored on the reverse entry (i, -image) of j's list
// (Atom::calculatePairForceShort, Atom.cpp:381-401).  Entries with j == i
// (the atom's own periodic images) are skipped in the pull form, as the
// reference's loop over neighborsUnique+1 does, and included in the push
// form (InterfaceLammps::getForces, InterfaceLammps.cpp:679-740).
// One wave per atom; no atomics, fixed summation order.
#include <algorithm>

#include "kernels.hpp"

namespace nnpx {
namespace {

__global__ __launch_bounds__(64) void k_force64(const ModelT* __restrict__ M, SysView s,
                                                 const double* __restrict__ dEdG, int ldG,
                                                 const double* __restrict__ dGself,
                                                 const double* __restrict__ dGnbr, int S,
                                                 int push, double* __restrict__ f)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int ei = s.elem[i];
    const int nsf = M->el[ei].nsf;
    const double* dE = dEdG + (size_t)i * ldG;
    const double* ds = dGself + (size_t)i * ldG * 3;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int q = lane; q < nsf; q += kWave) {
        const double w = dE[q];
        fx -= w * ds[3 * q];
        fy -= w * ds[3 * q + 1];
        fz -= w * ds[3 * q + 2];
    }
    const int n = s.nb_n[i];
    const size_t base = (size_t)i * s.cap;
    for (int k = lane; k < n; k += kWave) {
        const int j = s.nb_j[base + k];
        if (j == i && !push) continue;
        const int r = s.nb_rev[base + k];
        if (r < 0) continue;
        const ElementT& Ej = M->el[s.elem[j]];
        const int* tab = M->table + Ej.tab0[ei];
        const int nt = Ej.nsfd[ei];
        const double* dEj = dEdG + (size_t)j * ldG;
        const double* d = dGnbr + (size_t)r * S * 3;
        double px = 0.0, py = 0.0, pz = 0.0;
        for (int t = 0; t < nt; ++t) {
            const double w = dEj[tab[t]];
            px -= w * d[3 * t];
            py -= w * d[3 * t + 1];
            pz -= w * d[3 * t + 2];
        }
        fx += px;
        fy += py;
        fz += pz;
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
        f[3 * (size_t)i] = fx;
        f[3 * (size_t)i + 1] = fy;
        f[3 * (size_t)i + 2] = fz;
    }
}

// Work counters for the roofline: per atom P (pairs within rc), P_m (radial
// member evaluations), T (triplets with r_jk < rc of an existing angular
// group), T_m (angular member evaluations), T_e (exponentials).
__global__ __launch_bounds__(64) void k_work(const ModelT* __restrict__ M, SysView s,
                                              unsigned long long* out)
{
    __shared__ double sd[1024];
    __shared__ double sx[3][1024];
    __shared__ int se[1024];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const StructT& S = s.st[s.sid[i]];
    const int n = min(s.nb_n[i], 1024);
    for (int k = lane; k < n; k += kWave) {
        const int j = s.nb_j[(size_t)i * s.cap + k];
        double dr[3];
        entry_vector(s, i, (size_t)i * s.cap + k, S, dr);
        sx[0][k] = dr[0]; sx[1][k] = dr[1]; sx[2][k] = dr[2];
        sd[k] = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        se[k] = s.elem[j];
    }
    __syncthreads();
    unsigned long long P = 0, Pm = 0, T = 0, Tm = 0, Te = 0;
    for (int k = lane; k < n; k += kWave) {
        if (sd[k] < M->maxCutoff) ++P;
        for (int g = 0; g < E.nrad; ++g) {
            const GroupT& G = M->grp[E.grp0 + g];
            if ((G.e1 < 0 || se[k] == G.e1) && sd[k] < G.rc) Pm += G.nmem;
        }
    }
    for (int g = E.nrad; g < E.ngrp; ++g) {
        const GroupT& G = M->grp[E.grp0 + g];
        int nexp = 0;
        for (int q = 0; q < G.nmem; ++q) nexp += M->mem[G.mem0 + q].calcExp;
        for (int j = lane; j < n; j += kWave) {
            const bool any = G.e1 < 0;  // element-weighted group
            if (!((any || se[j] == G.e1 || se[j] == G.e2) && sd[j] < G.rc)) continue;
            for (int k = j + 1; k < n; ++k) {
                if (!(any || (se[j] == G.e1 && se[k] == G.e2) || (se[j] == G.e2 && se[k] == G.e1))) continue;
                if (!(sd[k] < G.rc)) continue;
                const double a = sx[0][k] - sx[0][j], b = sx[1][k] - sx[1][j], c = sx[2][k] - sx[2][j];
                if (G.type == 9 || G.type == 22 || G.type == 25 || a * a + b * b + c * c < G.rc2) {
                    ++T;
                    Tm += G.nmem;
                    Te += nexp;
                }
            }
        }
    }
    P = wave_sum(P); Pm = wave_sum(Pm); T = wave_sum(T); Tm = wave_sum(Tm); Te = wave_sum(Te);
    if (lane == 0) {
        atomicAdd(out + 0, P);
        atomicAdd(out + 1, T);
        atomicAdd(out + 2, Pm);
        atomicAdd(out + 3, Tm);
        atomicAdd(out + 4, Te);
    }
}

// Push form (InterfaceLammps::getForces, InterfaceLammps.cpp:679-740): atom i
// scatters -dE_i/dG . dG_i/dr_j onto every neighbour entry j (ghosts of an
// external list included) and -dE_i/dG . dG_i/dr_i onto itself.
__global__ __launch_bounds__(64) void k_force64_push(const ModelT* __restrict__ M, SysView s,
                                                      const double* __restrict__ dEdG, int ldG,
                                                      const double* __restrict__ dGself,
                                                      const double* __restrict__ dGnbr, int S, double* f)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.elem[i];
    const ElementT& E = M->el[e];
    const double* dE = dEdG + (size_t)i * ldG;
    const double* ds = dGself + (size_t)i * ldG * 3;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int q = lane; q < E.nsf; q += kWave) {
        fx -= dE[q] * ds[3 * q];
        fy -= dE[q] * ds[3 * q + 1];
        fz -= dE[q] * ds[3 * q + 2];
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
        atomicAdd(f + 3 * (size_t)i, fx);
        atomicAdd(f + 3 * (size_t)i + 1, fy);
        atomicAdd(f + 3 * (size_t)i + 2, fz);
    }
    const int n = s.nb_n[i];
    const size_t base = (size_t)i * s.cap;
    for (int k = lane; k < n; k += kWave) {
        const int j = s.nb_j[base + k];
        const int ej = s.elem[j];
        const int* tab = M->table + E.tab0[ej];
        const int nt = E.nsfd[ej];
        const double* d = dGnbr + (base + k) * S * 3;
        double px = 0.0, py = 0.0, pz = 0.0;
        for (int t = 0; t < nt; ++t) {
            const double w = dE[tab[t]];
            px -= w * d[3 * t];
            py -= w * d[3 * t + 1];
            pz -= w * d[3 * t + 2];
        }
        atomicAdd(f + 3 * (size_t)j, px);
        atomicAdd(f + 3 * (size_t)j + 1, py);
        atomicAdd(f + 3 * (size_t)j + 2, pz);
    }
}

__global__ void k_zero64(double* p, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.0;
}

}  // namespace

void launch_force64_push(const ModelT* m, const SysView& s, const double* dEdG, int ldG, const double* dGself,
                         const double* dGnbr, int S, double* f, hipStream_t stream)
{
    if (s.natoms == 0) return;
    const size_t n3 = (size_t)3 * s.nforce;
    hipLaunchKernelGGL(k_zero64, dim3((unsigned)std::min<size_t>((n3 + 255) / 256, 4096)), dim3(256), 0, stream, f,
                       n3);
    hipLaunchKernelGGL(k_force64_push, dim3(s.natoms), dim3(kWave), 0, stream, m, s, dEdG, ldG, dGself, dGnbr, S, f);
}

void launch_force64(const ModelT* m, const SysView& s, const double* dEdG, int ldG,
                    const double* dGself, const double* dGnbr, int S, int push, double* f,
                    hipStream_t stream)
{
    if (s.natoms == 0) return;
    hipLaunchKernelGGL(k_force64, dim3(s.natoms), dim3(kWave), 0, stream, m, s, dEdG, ldG,
                       dGself, dGnbr, S, push, f);
}

void launch_work_counts(const ModelT* m, const SysView& s, unsigned long long* out5,
                        hipStream_t stream)
{
    if (s.natoms == 0) return;
    hipLaunchKernelGGL(k_work, dim3(s.natoms), dim3(kWave), 0, stream, m, s, out5);
}

}  // namespace nnpx
