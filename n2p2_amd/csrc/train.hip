// Training gradient of the energy+force loss with respect to the network
// connections:
//   L = sum_s (E_s - E_ref,s)^2 / N_s + wF sum_{a,c} (F_ac - F_ref,ac)^2.
// The force term is differentiated through the forces' own derivative
// (second order): with u = dL/dF, sum_a u_a . dF_a/dw = -sum_i d/dw [dE_i/dG_i
// . Gdot_i], where Gdot_i = sum_a dG_i/dx_a . u_a is the tangent of atom i's
// SFs along u (the contraction of Training::collectDGdxia,
// Training.cpp:3586-3625).  Per atom, the network is run forward with that
// tangent and reversed once for both terms (the reference's calculateDEdc and
// calculateDFdc, NeuralNetwork.cpp:444-527, compute the same quantities per
// force component).
#include <hip/hip_runtime.h>

#include "train.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kBlock = 256;

// Activation, first and second derivative (NeuralNetwork::propagateLayer,
// NeuralNetwork.cpp:761-896).
template <typename T>
__device__ inline T act2(int a, T x, T& d1, T& d2)
{
    const T EXP_LIMIT = (T)35.0;
    T v;
    switch (a) {
    case AF_IDENTITY: d1 = 1; d2 = 0; return x;
    case AF_TANH: v = tanh(x); d1 = 1 - v * v; d2 = -2 * v * (1 - v * v); return v;
    case AF_LOGISTIC:
        if (x > EXP_LIMIT) { d1 = 0; d2 = 0; return 1; }
        if (x < -EXP_LIMIT) { d1 = 0; d2 = 0; return 0; }
        v = 1 / (1 + exp(-x)); d1 = v * (1 - v); d2 = v * (1 - v) * (1 - 2 * v); return v;
    case AF_SOFTPLUS:
        if (x > EXP_LIMIT) { d1 = 1; d2 = 0; return x; }
        if (x < -EXP_LIMIT) { d1 = 0; d2 = 0; return 0; }
        {
            const T t = exp(x);
            const T s = 1 / (1 + 1 / t);
            d1 = s; d2 = s * (1 - s);
            return log(1 + t);
        }
    case AF_RELU:
        d2 = 0;
        if (x > 0) { d1 = 1; return x; }
        d1 = 0; return 0;
    case AF_GAUSSIAN: { const T t = exp(-(T)0.5 * x * x); d1 = -x * t; d2 = (x * x - 1) * t; return t; }
    case AF_COS: { const T t = cos(x); d1 = -sin(x); d2 = -t; return t; }
    case AF_REVLOGISTIC:
        v = 1 / (1 + exp(-x)); d1 = v * (v - 1); d2 = v * (v - 1) * (1 - 2 * v); return 1 - v;
    case AF_EXP: v = exp(-x); d1 = -v; d2 = v; return v;
    case AF_HARMONIC: d1 = 2 * x; d2 = 2; return x * x;
    }
    d1 = 0; d2 = 0;
    return 0;
}

// u = 2 wF (F - F_ref); cE_s = 2 (E_s - E_ref,s) / N_s; loss terms.
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ F, const double* __restrict__ Fref,
                                                     const double* __restrict__ E, const double* __restrict__ Eref,
                                                     const int* __restrict__ atom_off, int natoms, int nstruct,
                                                     double wF, double* __restrict__ u, double* __restrict__ cE,
                                                     double* __restrict__ loss)
{
    __shared__ double part[kBlock / 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (t < natoms)
        for (int k = 0; k < 3; ++k) {
            const size_t q = 3 * (size_t)t + k;
            const double d = F[q] - Fref[q];
            u[q] = 2.0 * wF * d;
            sq += d * d;
        }
    sq = wave_sum(sq);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += part[w];
        atomicAdd(loss + 2, wF * s);
    }
    if (t < nstruct) {
        const double n = (double)(atom_off[t + 1] - atom_off[t]);
        const double r = E[t] - Eref[t];
        cE[t] = n > 0 ? 2.0 * r / n : 0.0;
        if (n > 0) atomicAdd(loss + 1, r * r / n);
    }
}

// Tangent from the stored dG/dr: Gdot_iq = dGself_iq . u_i + sum over i's
// neighbour entries (self images included) of dGnbr . u_j.  One wave per
// atom, lanes over SFs.
__global__ __launch_bounds__(64) void k_tangent64(const ModelT* __restrict__ M, SysView s,
                                                  const double* __restrict__ dGself, const double* __restrict__ dGnbr,
                                                  int S, const int* __restrict__ islot, const double* __restrict__ u,
                                                  double* __restrict__ Gdot, int ldG)
{
    const int i = blockIdx.x;
    const int e = s.elem[i];
    const int nsf = M->el[e].nsf;
    const int n = s.nb_n[i];
    const size_t base = (size_t)i * s.cap;
    const double* ui = u + 3 * (size_t)i;
    for (int q = threadIdx.x; q < nsf; q += kWave) {
        const double* ds = dGself + ((size_t)i * ldG + q) * 3;
        double acc = ds[0] * ui[0] + ds[1] * ui[1] + ds[2] * ui[2];
        for (int k = 0; k < n; ++k) {
            const int j = s.nb_j[base + k];
            const int t = islot[(e * kMaxElem + s.elem[j]) * kMaxSF + q];
            if (t < 0) continue;
            const double* d = dGnbr + ((base + k) * S + t) * 3;
            const double* uj = u + 3 * (size_t)j;
            acc += d[0] * uj[0] + d[1] * uj[1] + d[2] * uj[2];
        }
        Gdot[(size_t)i * ldG + q] = acc;
    }
}

// Network forward with the tangent and one reverse pass for
// Phi_i = cE_s E_i - dE_i/dG . Gdot_i; dPhi/dw accumulated per block in LDS
// (connection order) and added to grad.  One wave per block, a.per_block
// atoms of one element per block.  GATOM: networks whose connection gradient
// does not fit in LDS add every atom's dPhi/dw straight into its structure's
// row with FP64 global atomics (only the activations are staged in LDS).
template <typename T, bool GATOM>
__global__ __launch_bounds__(64) void k_grad(const ModelT* __restrict__ M, GradArgs a)
{
    extern __shared__ double gsm[];
    const int lane = threadIdx.x;
    int e = 0;
    while (e + 1 < a.nelem && (int)blockIdx.x >= a.blk_off[e + 1]) ++e;
    const NNT& N = M->el[e].nn;
    const int nl = N.nl;
    const int W = a.maxw;
    // LDS: gradient accumulator (connection order), then per layer h, hdot,
    // f', f'', adot; then the two adjoint buffers of the current layer and
    // the two of the layer below.
    double* acc = gsm;
    T* act = reinterpret_cast<T*>(GATOM ? gsm : acc + a.nconn_max);
    T* H = act;                       // [nl][W]
    T* Hd = H + kMaxLayers * W;       // tangent activations
    T* D1 = Hd + kMaxLayers * W;
    T* D2 = D1 + kMaxLayers * W;
    T* Ad = D2 + kMaxLayers * W;      // tangent pre-activations
    T* hb = Ad + kMaxLayers * W;      // d Phi / d h_l
    T* hd = hb + W;                   // d Phi / d hdot_l
    T* hb2 = hd + W;
    T* hd2 = hb2 + W;
    T* ab = hd2 + W;                  // d Phi / d a_l
    T* ad = ab + W;                   // d Phi / d adot_l
    const int nconn = a.nconn[e];
    if (!GATOM)
        for (int x = lane; x < nconn; x += kWave) acc[x] = 0.0;
    // connection-order offsets (NeuralNetwork::setConnections, NeuralNetwork.cpp:147-169)
    int cw[kMaxLayers], cb[kMaxLayers];
    {
        int o = 0;
        for (int l = 1; l < nl; ++l) {
            cw[l] = o;
            o += N.n[l - 1] * N.n[l];
            cb[l] = o;
            o += N.n[l];
        }
    }
    const double* wt = a.w + a.nweights;  // transposed copy [in][out]
    const int first = a.atom_off[e] + ((int)blockIdx.x - a.blk_off[e]) * a.per_block;
    const int last = min(first + a.per_block, a.atom_off[e + 1]);
    __syncthreads();
    // per-structure sums: flushed when the structure changes (the block's
    // atoms are in index order, so each structure's atoms are contiguous)
    int cur = -1;
    for (int p = first; p < last; ++p) {
        const int i = a.atoms[p];
        const int si = a.sid[i];
        double* grow = a.grad + (a.ldgrad ? (size_t)si * a.ldgrad : 0) + a.coff[e];
        if (!GATOM && a.ldgrad && si != cur) {
            if (cur >= 0) {
                double* g = a.grad + (size_t)cur * a.ldgrad + a.coff[e];
                for (int x = lane; x < nconn; x += kWave) {
                    atomicAdd(g + x, acc[x]);
                    acc[x] = 0.0;
                }
            }
            cur = si;
        }
        const T cE = (T)a.cE[si];
        for (int k = lane; k < N.n[0]; k += kWave) {
            H[k] = (T)a.G[(size_t)i * a.ldG + k];
            Hd[k] = (T)a.Gdot[(size_t)i * a.ldG + k];
        }
        __syncthreads();
        for (int l = 1; l < nl; ++l) {
            const int nin = N.n[l - 1], nout = N.n[l];
            const double* w = wt + N.woff[l];
            const double* b = a.w + N.boff[l];
            const T* hin = H + (l - 1) * W;
            const T* hdin = Hd + (l - 1) * W;
            for (int k = lane; k < nout; k += kWave) {
                T x = 0, xd = 0;
                for (int j = 0; j < nin; ++j) {
                    const T wjk = (T)w[(size_t)j * nout + k];
                    x += wjk * hin[j];
                    xd += wjk * hdin[j];
                }
                x += (T)b[k];
                if (N.normalize) { x /= (T)nin; xd /= (T)nin; }
                T d1, d2;
                const T v = act2<T>(N.act[l], x, d1, d2);
                H[l * W + k] = v;
                D1[l * W + k] = d1;
                D2[l * W + k] = d2;
                Ad[l * W + k] = xd;
                Hd[l * W + k] = d1 * xd;
            }
            __syncthreads();
        }
        // seeds: dPhi/dE = cE, dPhi/d(dE/dG . Gdot) = -1
        if (lane == 0) { hb[0] = cE; hd[0] = (T)-1; }
        __syncthreads();
        for (int l = nl - 1; l >= 1; --l) {
            const int nin = N.n[l - 1], nout = N.n[l];
            const T inv = N.normalize ? (T)1 / (T)nin : (T)1;
            for (int k = lane; k < nout; k += kWave) {
                const T d1 = D1[l * W + k];
                ab[k] = (hb[k] * d1 + hd[k] * D2[l * W + k] * Ad[l * W + k]) * inv;
                ad[k] = hd[k] * d1 * inv;
            }
            __syncthreads();
            const T* hin = H + (l - 1) * W;
            const T* hdin = Hd + (l - 1) * W;
            for (int x = lane; x < nin * nout; x += kWave) {
                const int j = x / nout, k = x - j * nout;
                const double v = (double)(ab[k] * hin[j] + ad[k] * hdin[j]);
                if (GATOM) atomicAdd(grow + cw[l] + x, v);
                else acc[cw[l] + x] += v;
            }
            for (int k = lane; k < nout; k += kWave) {
                if (GATOM) atomicAdd(grow + cb[l] + k, (double)ab[k]);
                else acc[cb[l] + k] += (double)ab[k];
            }
            if (l > 1) {
                const double* w = a.w + N.woff[l];  // [out][in]
                for (int j = lane; j < nin; j += kWave) {
                    T sb = 0, sd = 0;
                    for (int k = 0; k < nout; ++k) {
                        const T wkj = (T)w[(size_t)k * nin + j];
                        sb += wkj * ab[k];
                        sd += wkj * ad[k];
                    }
                    hb2[j] = sb;
                    hd2[j] = sd;
                }
            }
            __syncthreads();
            if (l > 1) {
                for (int j = lane; j < nin; j += kWave) {
                    hb[j] = hb2[j];
                    hd[j] = hd2[j];
                }
                __syncthreads();
            }
        }
    }
    if (GATOM) return;
    __syncthreads();
    if (a.ldgrad && cur < 0) return;
    double* g = a.grad + (a.ldgrad ? (size_t)cur * a.ldgrad : 0) + a.coff[e];
    for (int x = lane; x < nconn; x += kWave) atomicAdd(g + x, acc[x]);
}

}  // namespace

void launch_residual(const double* F, const double* Fref, const double* E, const double* Eref, const int* atom_off,
                     int natoms, int nstruct, double wF, double* u, double* cE, double* loss, hipStream_t stream)
{
    const int n = std::max(natoms, nstruct);
    hipLaunchKernelGGL(k_residual, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, F, Fref, E, Eref,
                       atom_off, natoms, nstruct, wF, u, cE, loss);
}

void launch_tangent64(const ModelT* m, const SysView& s, const double* dGself, const double* dGnbr, int S,
                      const int* islot, const double* u, double* Gdot, int ldG, hipStream_t stream)
{
    if (s.natoms == 0) return;
    hipLaunchKernelGGL(k_tangent64, dim3(s.natoms), dim3(kWave), 0, stream, m, s, dGself, dGnbr, S, islot, u, Gdot,
                       ldG);
}

size_t grad_act_bytes(const GradArgs& a, bool fp32)
{
    const size_t t = fp32 ? sizeof(float) : sizeof(double);
    return t * ((size_t)5 * kMaxLayers * a.maxw + 6 * (size_t)a.maxw);
}

size_t grad_lds_bytes(const GradArgs& a, bool fp32)
{
    return sizeof(double) * (size_t)a.nconn_max + grad_act_bytes(a, fp32);
}

void launch_grad(const ModelT* m, const GradArgs& a, int blocks, bool fp32, hipStream_t stream)
{
    if (blocks == 0) return;
    size_t lds = grad_lds_bytes(a, fp32);
    const bool gatom = lds > 65536;
    if (gatom) lds = grad_act_bytes(a, fp32);
    if (lds > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "network too wide for the gradient kernel's LDS staging"};
    if (fp32) {
        if (gatom) hipLaunchKernelGGL((k_grad<float, true>), dim3(blocks), dim3(kWave), lds, stream, m, a);
        else hipLaunchKernelGGL((k_grad<float, false>), dim3(blocks), dim3(kWave), lds, stream, m, a);
    } else {
        if (gatom) hipLaunchKernelGGL((k_grad<double, true>), dim3(blocks), dim3(kWave), lds, stream, m, a);
        else hipLaunchKernelGGL((k_grad<double, false>), dim3(blocks), dim3(kWave), lds, stream, m, a);
    }
}

}  // namespace nnpx
