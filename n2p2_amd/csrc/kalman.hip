// Kalman filter update of the network weights on the device
// (KalmanFilter::update, KalmanFilter.cpp:119-185), FP64:
//   X = P H          one wave per row of P, lanes over columns (coalesced)
//   A = H^T X        one block per entry, reduced over the state
//   K = X A^-1       (A^-1 of the m x m innovation matrix comes from the host)
//   P = (P - K X^T) / lambda + q I
//                    32 x 32 tiles of the lower triangle (read there, as
//                    selfAdjointView<Lower> reads P), written to the tile and,
//                    transposed through LDS, to its mirror: P stays symmetric
//   dw = K xi
// P is row-major n x n; H, X, K are [n][m] (observation index fastest).
#include "kernels.hpp"

namespace nnpx {
namespace {

constexpr int kKFCols = 8;  // observation columns per pass of the X kernel
constexpr int kTile = 32;

__global__ __launch_bounds__(256) void k_kf_x(const double* __restrict__ P, const double* __restrict__ H, int n, int m,
                                              int c0, double* __restrict__ X)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (row >= n) return;
    const int nc = min(kKFCols, m - c0);
    double acc[kKFCols];
#pragma unroll
    for (int c = 0; c < kKFCols; ++c) acc[c] = 0.0;
    const double* Pr = P + (size_t)row * n;
    for (int j = lane; j < n; j += kWave) {
        const double p = Pr[j];
#pragma unroll
        for (int c = 0; c < kKFCols; ++c)
            if (c < nc) acc[c] += p * H[(size_t)j * m + c0 + c];
    }
#pragma unroll
    for (int c = 0; c < kKFCols; ++c)
        if (c < nc) {
            const double v = wave_sum(acc[c]);
            if (lane == 0) X[(size_t)row * m + c0 + c] = v;
        }
}

__global__ __launch_bounds__(256) void k_kf_a(const double* __restrict__ H, const double* __restrict__ X, int n, int m,
                                              double* __restrict__ A)
{
    __shared__ double red[256 / kWave];
    const int a = blockIdx.x / m, b = blockIdx.x % m;
    double v = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) v += H[(size_t)j * m + a] * X[(size_t)j * m + b];
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += red[w];
        A[(size_t)a * m + b] = t;
    }
}

__global__ void k_kf_k(const double* __restrict__ X, const double* __restrict__ Ai, int n, int m, double* __restrict__ K)
{
    const size_t total = (size_t)n * m;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const size_t i = x / m;
        const int b = (int)(x - i * m);
        double v = 0.0;
        for (int a = 0; a < m; ++a) v += X[i * m + a] * Ai[(size_t)a * m + b];
        K[x] = v;
    }
}

// Lower-triangle tile (I, J), I >= J, from the linear block index.
__device__ inline void tri_tile(int t, int& I, int& J)
{
    I = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    J = t - I * (I + 1) / 2;
}

__global__ __launch_bounds__(256) void k_kf_p(double* __restrict__ P, const double* __restrict__ K,
                                              const double* __restrict__ X, int n, int m, double scale, double q)
{
    __shared__ double tile[kTile][kTile + 1];
    int I, J;
    tri_tile(blockIdx.x, I, J);
    const int c = threadIdx.x & (kTile - 1);
    const int r0 = threadIdx.x / kTile;  // 8 row groups
    for (int r = r0; r < kTile; r += 256 / kTile) {
        const int i = I * kTile + r, j = J * kTile + c;
        double v = 0.0;
        if (i < n && j < n && i >= j) {
            double s = 0.0;
            for (int b = 0; b < m; ++b) s += K[(size_t)i * m + b] * X[(size_t)j * m + b];
            v = (P[(size_t)i * n + j] - s) * scale;
            if (i == j) v += q;
            P[(size_t)i * n + j] = v;
        }
        tile[r][c] = v;
    }
    __syncthreads();
    // mirror: element (j, i) of the upper triangle takes (i, j)
    for (int r = r0; r < kTile; r += 256 / kTile) {
        const int i2 = J * kTile + r, j2 = I * kTile + c;  // target row, column
        if (i2 < n && j2 < n && j2 > i2) P[(size_t)i2 * n + j2] = tile[c][r];
    }
}

__global__ void k_kf_dw(const double* __restrict__ K, const double* __restrict__ xi, int n, int m, double* __restrict__ dw)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int b = 0; b < m; ++b) v += K[(size_t)i * m + b] * xi[b];
        dw[i] = v;
    }
}

__global__ void k_kf_init(double* P, int n, double diag)
{
    const size_t total = (size_t)n * n;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x)
        P[x] = (x / n == x % n) ? diag : 0.0;
}

unsigned grid_for(size_t work, unsigned block) { return (unsigned)std::min<size_t>((work + block - 1) / block, 16384); }

}  // namespace

void kf_init(double* P, int n, double diag, hipStream_t stream)
{
    hipLaunchKernelGGL(k_kf_init, dim3(grid_for((size_t)n * n, 256)), dim3(256), 0, stream, P, n, diag);
}

void kf_xa(const double* P, const double* H, int n, int m, double* X, double* A, hipStream_t stream)
{
    const int rows_per_block = 256 / kWave;
    for (int c0 = 0; c0 < m; c0 += kKFCols)
        hipLaunchKernelGGL(k_kf_x, dim3((n + rows_per_block - 1) / rows_per_block), dim3(256), 0, stream, P, H, n, m,
                           c0, X);
    hipLaunchKernelGGL(k_kf_a, dim3(m * m), dim3(256), 0, stream, H, X, n, m, A);
}

void kf_finish(double* P, const double* X, const double* Ai, const double* xi, int n, int m, double scale, double q,
               double* K, double* dw, hipStream_t stream)
{
    hipLaunchKernelGGL(k_kf_k, dim3(grid_for((size_t)n * m, 256)), dim3(256), 0, stream, X, Ai, n, m, K);
    const int nt = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_kf_p, dim3(nt * (nt + 1) / 2), dim3(256), 0, stream, P, K, X, n, m, scale, q);
    hipLaunchKernelGGL(k_kf_dw, dim3(grid_for((size_t)n, 256)), dim3(256), 0, stream, K, xi, n, m, dw);
}

}  // namespace nnpx
