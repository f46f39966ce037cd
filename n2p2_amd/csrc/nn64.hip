// FP64 atomic neural networks and per-structure energies (plain form).
//
// NeuralNetwork::propagate / propagateLayer (NeuralNetwork.cpp:386-394,
// 761-896) for every atom of an element, and dE/dG by reverse-mode
// (adjoint) accumulation instead of the reference's forward-mode
// NeuralNetwork::calculateDEdG (NeuralNetwork.cpp:396-442); both give
// dE/dG_k = sum over paths of W * f' products, the adjoint in one backward
// sweep.  One wave per atom, lanes over neurons; the layer input sits in LDS.
// Weights are read in two layouts: W^T ([in][out]) so lanes read consecutive
// outputs in the forward sweep, W ([out][in]) for the backward sweep.
#include "kernels.hpp"
#include "nnpx.h"

namespace nnpx {
namespace {

constexpr int kNNThreads = 256;
constexpr int kNNWaves = kNNThreads / kWave;

__device__ inline double act_fwd(int a, double x, double& dfdx)
{
    const double EXP_LIMIT = 35.0;  // NeuralNetwork.cpp EXP_LIMIT
    double v;
    switch (a) {
    case AF_IDENTITY: dfdx = 1.0; return x;
    case AF_TANH: v = tanh(x); dfdx = 1.0 - v * v; return v;
    case AF_LOGISTIC:
        if (x > EXP_LIMIT) { dfdx = 0.0; return 1.0; }
        if (x < -EXP_LIMIT) { dfdx = 0.0; return 0.0; }
        v = 1.0 / (1.0 + exp(-x)); dfdx = v * (1.0 - v); return v;
    case AF_SOFTPLUS:
        if (x > EXP_LIMIT) { dfdx = 1.0; return x; }
        if (x < -EXP_LIMIT) { dfdx = 0.0; return 0.0; }
        { const double t = exp(x); dfdx = 1.0 / (1.0 + 1.0 / t); return log(1.0 + t); }
    case AF_RELU:
        if (x > 0.0) { dfdx = 1.0; return x; }
        dfdx = 0.0; return 0.0;
    case AF_GAUSSIAN: { const double t = exp(-0.5 * x * x); dfdx = -x * t; return t; }
    case AF_COS: dfdx = -sin(x); return cos(x);
    case AF_REVLOGISTIC: v = 1.0 / (1.0 + exp(-x)); dfdx = v * (v - 1.0); return 1.0 - v;
    case AF_EXP: v = exp(-x); dfdx = -v; return v;
    case AF_HARMONIC: dfdx = 2.0 * x; return x * x;
    }
    dfdx = 0.0;
    return 0.0;
}

struct NNArgs {
    const double* w;    // [o][i] weights + biases (model layout)
    const double* wt;   // [i][o] transposed weights
    const int* atoms;   // atom indices sorted by element
    int blk_off[kMaxElem + 1];  // first block of each element
    int atom_off[kMaxElem + 1]; // first atom (in atoms[]) of each element
    int nelem;
};

__global__ __launch_bounds__(kNNThreads) void k_nn64(const ModelT* __restrict__ M, NNArgs a,
                                                      NNIO io, int maxw, int maxh)
{
    extern __shared__ double sm[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    int e = 0;
    while (e + 1 < a.nelem && (int)blockIdx.x >= a.blk_off[e + 1]) ++e;
    const NNT& N = M->el[e].nn;
    const int nl = N.nl;
    // per-wave LDS: two activation buffers (any layer, input included) and
    // f' of every hidden/output layer
    double* base = sm + (size_t)wave * (2 * maxw + kMaxLayers * maxh);
    double* va = base;
    double* vb = base + maxw;
    double* dfd = base + 2 * maxw;  // [layer][maxh]
    const int slot = ((int)blockIdx.x - a.blk_off[e]) * kNNWaves + wave;
    const int pos = a.atom_off[e] + slot;
    if (pos >= a.atom_off[e + 1]) return;
    const int i = a.atoms[pos];
    const double* g = io.G + (size_t)i * io.ldG;
    for (int k = lane; k < N.n[0]; k += kWave) va[k] = g[k];
    __builtin_amdgcn_wave_barrier();
    double* vin = va;
    double* vout = vb;
    for (int l = 1; l < nl; ++l) {
        const int nin = N.n[l - 1], nout = N.n[l];
        const double* wt = a.wt + N.woff[l];
        const double* b = a.w + N.boff[l];
        for (int k = lane; k < nout; k += kWave) {
            double x = 0.0;
            for (int j = 0; j < nin; ++j) x += wt[(size_t)j * nout + k] * vin[j];
            x += b[k];
            if (N.normalize) x /= nin;
            double df;
            vout[k] = act_fwd(N.act[l], x, df);
            dfd[l * maxh + k] = df;
        }
        __builtin_amdgcn_wave_barrier();
        double* t = vin; vin = vout; vout = t;
    }
    if (lane == 0) io.energy[i] = vin[0];
    if (!io.dEdG) return;
    // Backward: delta_l = dE/dx_l (pre-activation), delta_{L-1} = f'_out.
    double* din = vout;  // free buffer
    double* dout = vin;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) din[0] = dfd[(nl - 1) * maxh];
    __builtin_amdgcn_wave_barrier();
    for (int l = nl - 1; l >= 1; --l) {
        const int nin = N.n[l - 1], nout = N.n[l];
        const double* w = a.w + N.woff[l];
        const double norm = N.normalize ? 1.0 / nin : 1.0;
        for (int j = lane; j < nin; j += kWave) {
            double s = 0.0;
            for (int k = 0; k < nout; ++k) s += w[(size_t)k * nin + j] * din[k];
            if (N.normalize) s *= norm;
            if (l - 1 >= 1) s *= dfd[(l - 1) * maxh + j];
            dout[j] = s;
        }
        __builtin_amdgcn_wave_barrier();
        double* t = din; din = dout; dout = t;
    }
    double* dE = io.dEdG + (size_t)i * io.ldG;
    for (int k = lane; k < N.n[0]; k += kWave) dE[k] = din[k];
}

// One block per structure, fixed-order tree sum.
// Per-structure sum of the atomic energies, one block per structure: four
// independent strided partial sums per thread keep several loads in flight,
// then a fixed-order tree in LDS (deterministic from run to run).
constexpr int kEnergyThreads = 1024;
__global__ __launch_bounds__(kEnergyThreads) void k_energy(const double* atomE, const int* off, double* E)
{
    __shared__ double red[kEnergyThreads];
    const int st = blockIdx.x;
    const int b = off[st], e = off[st + 1];
    const int t = (int)threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int q = b + t;
    for (; q + 3 * kEnergyThreads < e; q += 4 * kEnergyThreads) {
        s0 += atomE[q];
        s1 += atomE[q + kEnergyThreads];
        s2 += atomE[q + 2 * kEnergyThreads];
        s3 += atomE[q + 3 * kEnergyThreads];
    }
    for (; q < e; q += kEnergyThreads) s0 += atomE[q];
    red[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int o = kEnergyThreads / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) E[st] = red[0];
}

}  // namespace

void launch_nn64(const ModelT* m, const ModelT& hm, const double* w, const int* elem_atoms,
                 const int* elem_count_host, const int* elem_off_host, const NNIO& io,
                 hipStream_t stream)
{
    NNArgs a{};
    a.w = w;
    a.wt = w + hm.nweights;
    a.atoms = elem_atoms;
    a.nelem = hm.nelem;
    int blocks = 0;
    for (int e = 0; e < hm.nelem; ++e) {
        a.blk_off[e] = blocks;
        a.atom_off[e] = elem_off_host[e];
        blocks += (elem_count_host[e] + kNNWaves - 1) / kNNWaves;
    }
    a.blk_off[hm.nelem] = blocks;
    a.atom_off[hm.nelem] = elem_off_host[hm.nelem];
    if (blocks == 0) return;
    const int maxw = hm.max_width, maxh = hm.max_hidden;
    const size_t lds = sizeof(double) * (size_t)kNNWaves * (2 * maxw + kMaxLayers * maxh);
    if (lds > 65536) throw Error{NNPX_ERR_UNSUPPORTED, "network too wide for the FP64 NN kernel's LDS staging"};
    hipLaunchKernelGGL(k_nn64, dim3(blocks), dim3(kNNThreads), lds, stream, m, a, io, maxw, maxh);
}

void launch_energy(const double* atomE, const int* atom_off, int nstruct, double* E,
                   hipStream_t stream)
{
    if (nstruct == 0) return;
    hipLaunchKernelGGL(k_energy, dim3(nstruct), dim3(kEnergyThreads), 0, stream, atomE, atom_off, E);
}

}  // namespace nnpx
