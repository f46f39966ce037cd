// Neighbour list on gfx950, set- and order-identical to
// Structure::calculateNeighborList(rc, false) (Structure.cpp:287-398):
// for atom i every (j, bc0, bc1, bc2) with |dr|^2 <= rc^2, ordered by j, then
// bc0, bc1, bc2, with dr evaluated by the reference's expression.
//
// Two builders share the output layout (fixed capacity per atom):
//  * brute force -- one wave per atom, lanes over j, serial images; hits are
//    kept as a bit per image and emitted in loop order after a wave prefix
//    sum (no sort needed);
//  * cell list  -- for structures whose cell holds >= 3 cells of width >= rc
//    on every axis: atoms binned by fractional coordinate, candidates from
//    the 27 surrounding cells (less those wholly beyond the cutoff), hits
//    collected in LDS and bitonic-sorted by (j, image) so the order equals
//    the brute-force loop order.
#include "kernels.hpp"

#include <algorithm>

namespace nnpx {
namespace {

constexpr int kNbMaxLds = 1024;  // max entries sorted in LDS per atom
constexpr int kCandLds = 2048;   // candidates whose cell is tabulated in LDS

// Sort buffer of the cell-list builder: the power of two at or above the
// list capacity (the builder's LDS, and so the waves per CU, follow it).
inline int sort_slots(int cap)
{
    int np = 64;
    while (np < cap && np < kNbMaxLds) np <<= 1;
    return np;
}
constexpr int kCandUnroll = 4;   // candidates per lane and step

__device__ inline bool gated_off(const SysView& s) { return s.gate && *s.gate == 0; }

// Cell of atom i (-1: its structure is built by brute force).
__device__ inline int atom_cell_of(const SysView& s, int i)
{
    const StructT& S = s.st[s.sid[i]];
    if (S.algo != 1) return -1;
    const double* r = s.pos + 3 * (size_t)i;
    int c[3];
    for (int k = 0; k < 3; ++k) {
        // fractional coordinate as in Structure::remap (Structure.cpp:1226-1228)
        double f = r[0] * S.invbox[k] + r[1] * S.invbox[3 + k] + r[2] * S.invbox[6 + k];
        int q = (int)(f * S.nc[k]);
        c[k] = q < 0 ? 0 : (q >= S.nc[k] ? S.nc[k] - 1 : q);
    }
    return S.cell0 + (c[0] * S.nc[1] + c[1]) * S.nc[2] + c[2];
}

// Atom i into its cell's cell-ordered arrays at position p.
__device__ inline void cell_place(const SysView& s, int i, int p)
{
    s.cell_atoms[p] = i;
    // cell-ordered copy of the positions: candidate loads become contiguous
    s.cell_pos[3 * (size_t)p] = s.pos[3 * (size_t)i];
    s.cell_pos[3 * (size_t)p + 1] = s.pos[3 * (size_t)i + 1];
    s.cell_pos[3 * (size_t)p + 2] = s.pos[3 * (size_t)i + 2];
}

__global__ void k_cell_count(SysView s)
{
    if (gated_off(s)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.natoms) return;
    const int cell = atom_cell_of(s, i);
    s.atom_cell[i] = cell;
    // the count before this atom is its place in the cell's atom range
    if (cell >= 0) s.cell_rank[i] = atomicAdd(&s.cell_count[cell], 1);
}

// Exclusive scan of cell counts (single block; ncells is small).
__global__ __launch_bounds__(1024) void k_scan_cells(SysView s)
{
    if (gated_off(s)) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int n = s.ncells;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int sum = 0;
    for (int q = b; q < e; ++q) sum += s.cell_count[q];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int q = b; q < e; ++q) {
        s.cell_start[q] = run;
        run += s.cell_count[q];
    }
    if (t == 1023) s.cell_start[n] = part[1023];
}

__global__ void k_cell_fill(SysView s)
{
    if (gated_off(s)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.natoms) return;
    const int c = s.atom_cell[i];
    if (c < 0) return;
    cell_place(s, i, s.cell_start[c] + s.cell_rank[i]);
}

__device__ inline void emit(const SysView& s, int i, int slot, int j, int img)
{
    if (slot < s.cap) {
        s.nb_j[(size_t)i * s.cap + slot] = j;
        s.nb_img[(size_t)i * s.cap + slot] = img;
    }
}

// Whether every atom of the neighbour cell at offset (d0, d1, d2) = (o / 9,
// o / 3 % 3, o % 3) - 1 from atom i's cell (c0, c1, c2) lies beyond the
// cutoff, so the cell need not be searched.  Along lattice direction k the
// cell lies at least gap_k cell widths from atom i (measured perpendicular
// to the cell faces, gap_k = u_k for d_k = -1, 1 - u_k for d_k = 1, with u_k
// atom i's position inside its cell); the gaps combine as a sum of squares
// in a cell with orthogonal axes and by their maximum otherwise.  The gaps
// are shortened by 1e-6 cell widths, far beyond any rounding of the
// fractional coordinates, so no candidate within the cutoff is dropped.
__device__ inline bool cell_beyond_cutoff(const StructT& S, const double* ri, int c0, int c1, int c2, int o,
                                          double rc2)
{
    const int d[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
    const int c[3] = {c0, c1, c2};
    double g[3][3], w[3], sum2 = 0.0, max2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        g[k][0] = S.invbox[k];
        g[k][1] = S.invbox[3 + k];
        g[k][2] = S.invbox[6 + k];
        // cell width perpendicular to the faces of direction k
        w[k] = 1.0 / (S.nc[k] * sqrt(g[k][0] * g[k][0] + g[k][1] * g[k][1] + g[k][2] * g[k][2]));
    }
    for (int k = 0; k < 3; ++k) {
        if (d[k] == 0) continue;
        const double f = ri[0] * g[k][0] + ri[1] * g[k][1] + ri[2] * g[k][2];
        const double u = f * S.nc[k] - c[k];
        const double gap = fmax(0.0, (d[k] < 0 ? u : 1.0 - u) - 1e-6) * w[k];
        sum2 += gap * gap;
        max2 = fmax(max2, gap * gap);
    }
    bool ortho = true;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            const double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
            const double na = g[a][0] * g[a][0] + g[a][1] * g[a][1] + g[a][2] * g[a][2];
            const double nb = g[b][0] * g[b][0] + g[b][1] * g[b][1] + g[b][2] * g[b][2];
            if (dot * dot > 1e-24 * na * nb) ortho = false;
        }
    return (ortho ? sum2 : max2) > rc2;
}

// One wave per atom.  Dynamic LDS (cell-list structures only): the sort
// buffer of `slots` keys, then the candidate-cell table.
__global__ __launch_bounds__(64) void k_build(SysView s, double rc2, int slots)
{
    if (gated_off(s)) return;
    extern __shared__ long long keys[];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const StructT& S = s.st[s.sid[i]];
    const double* ri = s.pos + 3 * (size_t)i;
    int count = 0;
    if (S.algo != 1) {
        // Image ranges: all of -p..p, or, when every atom lies inside the
        // cell, only the images whose slab along each lattice direction
        // comes within rc of atom i (slab distance gap * w, w the
        // perpendicular width), with a margin; the loops keep the
        // reference's (j, bc0, bc1, bc2) order.
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        double rk[3] = {0.0, 0.0, 0.0};  // rc in fractional units of direction k, with a margin
        if (S.periodic) {
            const double rc = sqrt(rc2);
            for (int k = 0; k < 3; ++k) {
                lo[k] = -S.pbc[k];
                hi[k] = S.pbc[k];
                const double a = S.invbox[k], b = S.invbox[3 + k], c = S.invbox[6 + k];
                const double r = rc * sqrt(a * a + b * b + c * c) * (1.0 + 1e-9) + 1e-9;
                rk[k] = r;
                if (S.inside) {
                    const double f = ri[0] * a + ri[1] * b + ri[2] * c;
                    lo[k] = max(lo[k], -(int)floor(f + r));
                    hi[k] = min(hi[k], (int)floor(1.0 - f + r));
                }
            }
        }
        const int n12 = (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
        const int nimg = (hi[0] - lo[0] + 1) * n12;
        if (nimg <= kWave) {
            // One pass: each lane records its hits as a bit per image (in loop
            // order) and emits them from the mask after the wave's prefix sum.
            // A lane tests only the images of its pair (i, j) whose slab
            // along every lattice direction k comes within rc, i.e.
            // |(r_i - r_j) . g_k + b_k| <= rc |g_k| (g_k the reciprocal
            // vectors, so the left side over |g_k| is the perpendicular
            // distance); with the margin in rk the images skipped hold no
            // neighbour.
            __shared__ int img_tab[kWave];
            if (lane < nimg) {
                const int a0 = lane / n12, a1 = (lane - a0 * n12) / (hi[2] - lo[2] + 1);
                const int a2 = lane - a0 * n12 - a1 * (hi[2] - lo[2] + 1);
                img_tab[lane] = pack_image(lo[0] + a0, lo[1] + a1, lo[2] + a2);
            }
            __syncthreads();
            for (int jb = 0; jb < S.natoms; jb += kWave) {
                const int jl = jb + lane;
                const int j = S.atom0 + jl;
                unsigned long long hm = 0;
                if (jl < S.natoms) {
                    const double* rj = s.pos + 3 * (size_t)j;
                    int l[3] = {lo[0], lo[1], lo[2]}, h[3] = {hi[0], hi[1], hi[2]};
                    if (S.periodic)
                        for (int k = 0; k < 3; ++k) {
                            const double df = (ri[0] - rj[0]) * S.invbox[k] + (ri[1] - rj[1]) * S.invbox[3 + k] +
                                              (ri[2] - rj[2]) * S.invbox[6 + k];
                            l[k] = max(l[k], (int)ceil(-df - rk[k]));
                            h[k] = min(h[k], (int)floor(-df + rk[k]));
                        }
                    const int n2 = hi[2] - lo[2] + 1, n1 = hi[1] - lo[1] + 1;
                    for (int b0 = l[0]; b0 <= h[0]; ++b0)
                        for (int b1 = l[1]; b1 <= h[1]; ++b1)
                            for (int b2 = l[2]; b2 <= h[2]; ++b2) {
                                if (i == j && b0 == 0 && b1 == 0 && b2 == 0) continue;
                                double dr[3];
                                pair_vector(ri, rj, S.box, b0, b1, b2, dr);
                                if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2)
                                    hm |= 1ull << (((b0 - lo[0]) * n1 + (b1 - lo[1])) * n2 + (b2 - lo[2]));
                            }
                }
                const int hits = __popcll(hm);
                const int incl = wave_incl_scan(hits);
                int slot = count + incl - hits;
                while (hm) {
                    emit(s, i, slot++, j, img_tab[__builtin_ctzll(hm)]);
                    hm &= hm - 1;
                }
                count += lane_value(incl, kWave - 1);
            }
        } else {
        // More images than lanes: count the hits, then test again to emit.
        for (int jb = 0; jb < S.natoms; jb += kWave) {
            const int jl = jb + lane;
            const int j = S.atom0 + jl;
            int hits = 0;
            if (jl < S.natoms) {
                const double* rj = s.pos + 3 * (size_t)j;
                for (int b0 = lo[0]; b0 <= hi[0]; ++b0)
                    for (int b1 = lo[1]; b1 <= hi[1]; ++b1)
                        for (int b2 = lo[2]; b2 <= hi[2]; ++b2) {
                            if (i == j && b0 == 0 && b1 == 0 && b2 == 0) continue;
                            double dr[3];
                            pair_vector(ri, rj, S.box, b0, b1, b2, dr);
                            if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2) ++hits;
                        }
            }
            const int incl = wave_incl_scan(hits);
            int slot = count + incl - hits;
            if (hits) {
                const double* rj = s.pos + 3 * (size_t)j;
                for (int b0 = lo[0]; b0 <= hi[0]; ++b0)
                    for (int b1 = lo[1]; b1 <= hi[1]; ++b1)
                        for (int b2 = lo[2]; b2 <= hi[2]; ++b2) {
                            if (i == j && b0 == 0 && b1 == 0 && b2 == 0) continue;
                            double dr[3];
                            pair_vector(ri, rj, S.box, b0, b1, b2, dr);
                            if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2)
                                emit(s, i, slot++, j, pack_image(b0, b1, b2));
                        }
            }
            count += lane_value(incl, kWave - 1);
        }
        }
    } else {
        // The 27 surrounding cells are flattened into one candidate range so
        // the wave tests 64 candidates per step with contiguous loads.
        __shared__ int cs[27], co[28], cim[27];
        unsigned char* ccell = reinterpret_cast<unsigned char*>(keys + slots);
        const int c = s.atom_cell[i] - S.cell0;
        const int c0 = c / (S.nc[1] * S.nc[2]), c1 = (c / S.nc[2]) % S.nc[1], c2 = c % S.nc[2];
        int start = 0, len = 0, img = 0;
        if (lane < 27) {
            int q[3] = {c0 + lane / 9 - 1, c1 + (lane / 3) % 3 - 1, c2 + lane % 3 - 1};
            int b[3];
            for (int k = 0; k < 3; ++k) {
                // wrapped cell: neighbour image r_j + m*box, i.e. bc = -m
                const int m = q[k] < 0 ? -1 : (q[k] >= S.nc[k] ? 1 : 0);
                q[k] -= m * S.nc[k];
                b[k] = -m;
            }
            const int cell = S.cell0 + (q[0] * S.nc[1] + q[1]) * S.nc[2] + q[2];
            start = s.cell_start[cell];
            len = s.cell_start[cell + 1] - start;
            img = pack_image(b[0], b[1], b[2]);
            if (cell_beyond_cutoff(S, ri, c0, c1, c2, lane, rc2)) len = 0;
        }
        const int incl = wave_incl_scan(len);
        if (lane < 27) {
            cs[lane] = start;
            co[lane] = incl - len;
            cim[lane] = img;
        }
        if (lane == 26) co[27] = incl;
        __syncthreads();
        const int total = co[27];
        // cell of every candidate (one LDS read per candidate in the loop
        // below instead of a search over the 27 offsets)
        const bool table = total <= kCandLds;
        if (table && lane < 27)
            for (int t = incl - len; t < incl; ++t) ccell[t] = (unsigned char)lane;
        __syncthreads();
        const int self_img = pack_image(0, 0, 0);
        // kCandUnroll candidates per lane and step: their loads are
        // independent, so the wave keeps several in flight; the order of the
        // collected keys does not matter (they are sorted below).
        for (int b0 = 0; b0 < total; b0 += kWave * kCandUnroll) {
            bool hit[kCandUnroll];
            long long key[kCandUnroll];
            int nh = 0;
#pragma unroll
            for (int u = 0; u < kCandUnroll; ++u) {
                const int x = b0 + u * kWave + lane;
                hit[u] = false;
                key[u] = 0;
                if (x < total) {
                    int lo = 0;
                    if (table) {
                        lo = ccell[x];
                    } else {
                        int hi = 26;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (co[mid] <= x) lo = mid;
                            else hi = mid - 1;
                        }
                    }
                    const int p = cs[lo] + x - co[lo];
                    const int j = s.cell_atoms[p];
                    const int im = cim[lo];
                    if (!(i == j && im == self_img)) {
                        int q0, q1, q2;
                        unpack_image(im, q0, q1, q2);
                        double dr[3];
                        pair_vector(ri, s.cell_pos + 3 * (size_t)p, S.box, q0, q1, q2, dr);
                        hit[u] = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] <= rc2;
                        key[u] = ((long long)j << 21) | im;
                    }
                }
                nh += hit[u] ? 1 : 0;
            }
            const int inc = wave_incl_scan(nh);
            int slot = count + inc - nh;
#pragma unroll
            for (int u = 0; u < kCandUnroll; ++u)
                if (hit[u]) {
                    if (slot < slots) keys[slot] = key[u];
                    ++slot;
                }
            count += lane_value(inc, kWave - 1);
        }
        // Bitonic sort of the collected keys (padded to a power of two).
        const int n = min(count, slots);
        int np = 1;
        while (np < n) np <<= 1;
        for (int q = n + lane; q < np; q += kWave) keys[q] = 0x7fffffffffffffffLL;
        __syncthreads();
        for (int k = 2; k <= np; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int x = lane; x < np; x += kWave) {
                    const int y = x ^ jj;
                    if (y > x) {
                        const long long a = keys[x], bb = keys[y];
                        const bool up = (x & k) == 0;
                        if ((a > bb) == up) { keys[x] = bb; keys[y] = a; }
                    }
                }
                __syncthreads();
            }
        for (int q = lane; q < n; q += kWave)
            emit(s, i, q, (int)(keys[q] >> 21), (int)(keys[q] & ((1 << 21) - 1)));
        if (count > slots && lane == 0) atomicOr(&s.flags[0], 1);
    }
    if (lane == 0) {
        s.nb_n[i] = min(count, s.cap);
        if (count > s.cap) atomicOr(&s.flags[0], 1);
        // every wave would otherwise hit the one address; a plain read first
        // skips the atomic once the running maximum is at least this count
        if (count > __hip_atomic_load(&s.flags[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&s.flags[1], count);
    }
}

// Reverse entry: (i -> j, image) is stored in j's list as (i, -image).
__global__ void k_reverse(SysView s)
{
    if (gated_off(s)) return;
    const int i = blockIdx.x;
    const int n = s.nb_n[i];
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const size_t e = (size_t)i * s.cap + k;
        const int j = s.nb_j[e];
        const long long want = ((long long)i << 21) | negate_image(s.nb_img[e]);
        int lo = 0, hi = s.nb_n[j] - 1, found = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const size_t m = (size_t)j * s.cap + mid;
            const long long key = ((long long)s.nb_j[m] << 21) | s.nb_img[m];
            if (key == want) { found = mid; break; }
            if (key < want) lo = mid + 1;
            else hi = mid - 1;
        }
        s.nb_rev[e] = found < 0 ? -1 : (int)((size_t)j * s.cap + found);
        if (found < 0) atomicOr(&s.flags[0], 2);
    }
}

__global__ void k_fill_int(int* p, int v, size_t n, const int* gate)
{
    if (gate && *gate == 0) return;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = v;
}

// The flags (na entries) and the cell counts cleared by one launch.
__global__ void k_clear_two(int* a, int na, int* b, int nb, const int* gate)
{
    if (gate && *gate == 0) return;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < na + nb; q += gridDim.x * blockDim.x) {
        if (q < na) a[q] = 0;
        else b[q - na] = 0;
    }
}

}  // namespace

// Memory clears use a fill kernel (not hipMemsetAsync) so the sequence can be
// captured in a HIP graph.
void fill_int(int* p, int v, size_t n, hipStream_t stream, const int* gate)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_int, dim3(nb), dim3(256), 0, stream, p, v, n, gate);
}

void launch_neighbors(const SysView& s, double rc, bool any_cell, hipStream_t stream, bool clear_flags)
{
    if (s.natoms == 0) return;
    if (any_cell) {
        const int nf = clear_flags ? 4 : 0;
        const int nc = (int)std::max<long long>(1, std::min<long long>((nf + (long long)s.ncells + 255) / 256, 4096));
        hipLaunchKernelGGL(k_clear_two, dim3(nc), dim3(256), 0, stream, s.flags, nf, s.cell_count, s.ncells, s.gate);
        const int nb = (s.natoms + 255) / 256;
        hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, stream, s);
        hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, s);
        hipLaunchKernelGGL(k_cell_fill, dim3(nb), dim3(256), 0, stream, s);
    } else if (clear_flags) {
        fill_int(s.flags, 0, 4, stream, s.gate);
    }
    const int slots = sort_slots(s.cap);
    const size_t lds = any_cell ? sizeof(long long) * (size_t)slots + kCandLds : 0;
    hipLaunchKernelGGL(k_build, dim3(s.natoms), dim3(kWave), lds, stream, s, rc * rc, slots);
}

void launch_reverse(const SysView& s, hipStream_t stream)
{
    if (s.natoms == 0) return;
    hipLaunchKernelGGL(k_reverse, dim3(s.natoms), dim3(kWave), 0, stream, s);
}

}  // namespace nnpx
