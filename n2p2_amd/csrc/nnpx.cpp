// C ABI of libnnpx (include/nnpx.h): context, model and structure state, and
// the nnp::Mode compute sequence on one GPU.
#include "nnpx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "md.hpp"
#include "mixed.hpp"
#include "train.hpp"
#include "model.hpp"

using namespace nnpx;

// Device memory of nnpx_predict_noalloc: every GPU the library runs on
// holds its own copy of this array until the process ends.  Host code never
// passes its name to a kernel or a copy: it uses the device address
// hipGetSymbolAddress returns for the current device (scratch_address).
// kScratchBytes is the most device memory one such call may use.
constexpr size_t kScratchBytes = (size_t)64 << 20;
// device memory the mixed form's angular pair lists may take
constexpr size_t kTripBudget = (size_t)4 << 30;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_create_error;

// Device address of `scratch` on each device, looked up once per device
// (with that device current) when the first context on it is created.
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;

hipError_t scratch_address(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    *out = static_cast<unsigned char*>(p);
    return hipSuccess;
}

struct Timer {
    std::string name;
    hipEvent_t a, b;
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct nnpx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // options
    int force_mode = NNPX_FORCES_PULL;
    bool timing = false;
    int nbr_algo = 0;
    // model
    bool has_model = false;
    HostModel hm;
    ModelT* d_model = nullptr;
    double* d_w = nullptr;       // FP64 weights, then the transposed copy
    MixedModel mixed;            // BF16/FP32 tables of the mixed form
    // structures
    int N = 0, S = 0;
    std::vector<int64_t> atom_off;
    std::vector<int> elem_host;
    std::vector<StructT> st_host;
    std::vector<int> elem_count, elem_off;  // per element, into elem_atoms
    bool any_cell = false;
    int ncells = 0;
    // device buffers
    std::map<std::string, DevBuf> bufs;
    // neighbour list state
    int cap = 0;
    int max_nb = 0;
    // LDS staging capacity of the mixed kernels: the last checked build's
    // maximum count (0: the list capacity); a larger count at run time is
    // flagged (bit 4) and the evaluation repeated at full capacity
    int stage_cap = 0;
    int trip_cap = 0;            // entries per atom of the mixed form's angular pair lists (0: none)
    int md_stage_cap = 0;        // MD: staging fitted to the in-cutoff entries (0: the list capacity)
    int md_stage_req = -1;       // NNPX_OPT_MD_STAGING
    bool nbr_valid = false;
    bool rev_valid = false;
    // results state
    bool sf_valid = false, sfd_valid = false, nn_valid = false, dedg_valid = false;
    bool e_valid = false, f_valid = false;
    int last_precision = NNPX_PLAIN_FP64;
    bool last_stored = false;    // the last plain evaluation stored dG/dr
    int plain_forces = 0;        // NNPX_OPT_PLAIN_FORCES: 0 fused recompute, 1 stored dG/dr
    // MD (velocity Verlet with a skin-buffered list)
    double skin = 0.0;           // extra list cutoff, model length units
    bool md_ready = false;
    int md_precision = NNPX_MIXED;
    MDConst mdc{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    // caller-provided neighbour lists (LAMMPS path): forces on ntotal atoms
    bool external = false;
    int ntotal = 0;
    // no-allocation calls: named buffers carved from the device's file-scope
    // scratch array instead of hipMalloc (see nnpx_predict_noalloc)
    unsigned char* scratch = nullptr;  // this device's copy of the array
    bool arena = false;
    size_t arena_used = 0;
    std::map<std::string, DevBuf> abufs;
    // Kalman filter weight updaters (one over all connections, or one per
    // element): state size, first connection, scalars; P on the device
    struct KF {
        int n = 0, off = 0;
        double eta = 0, etatau = 0, etamax = 0, lambda = 0, nu = 0, gamma = 1, q = 0, qtau = 0, qmin = 0;
        int64_t updates = 0;
    };
    int kf_type = -1;  // -1: not set up
    std::vector<KF> kf;
    // timing
    std::vector<Timer> pending;
    std::map<std::string, std::pair<double, int64_t>> ktime;
    std::vector<hipEvent_t> event_pool;
};

namespace {

struct Fail {
    nnpx_status st;
    std::string msg;
};

[[noreturn]] void fail(nnpx_status st, const std::string& m) { throw Fail{st, m}; }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(NNPX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
nnpx_status guard(nnpx_ctx* c, F&& f)
{
    if (!c) return NNPX_ERR_ARGUMENT;
    try {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        f();
        return NNPX_OK;
    } catch (const Fail& x) {
        c->err = x.msg;
        return x.st;
    } catch (const Error& x) {
        c->err = x.msg;
        return (nnpx_status)x.status;
    } catch (const std::exception& x) {
        c->err = x.what();
        return NNPX_ERR_ARGUMENT;
    }
}

void drop_graph(nnpx_ctx* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_steps = 0;
}

// Named device buffer of at least count elements; growing one invalidates a
// captured MD graph (its kernels hold the old addresses).
template <typename T>
T* buf(nnpx_ctx* c, const std::string& name, size_t count)
{
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (c->arena) {
        // bump allocation in the scratch array; a buffer that grows within a
        // call moves to fresh space (the old one is not reused before the
        // next call)
        DevBuf& a = c->abufs[name];
        if (a.bytes < bytes) {
            const size_t off = (c->arena_used + 255) & ~(size_t)255;
            if (off > kScratchBytes || bytes > kScratchBytes - off)
                fail(NNPX_ERR_RANGE, "input too large for the no-allocation scratch buffer (" +
                                         std::to_string(kScratchBytes >> 20) + " MiB): buffer " + name + " needs " +
                                         std::to_string(bytes) + " bytes at offset " + std::to_string(off));
            a.p = c->scratch + off;
            a.bytes = bytes;
            c->arena_used = off + bytes;
        }
        return static_cast<T*>(a.p);
    }
    DevBuf& b = c->bufs[name];
    if (b.bytes < bytes) {
        drop_graph(c);
        if (b.p) hip_check(hipFree(b.p), "hipFree");
        b.p = nullptr;
        b.bytes = 0;
        hip_check(hipMalloc(&b.p, bytes), ("hipMalloc " + name).c_str());
        b.bytes = bytes;
    }
    return static_cast<T*>(b.p);
}

// Model-level tables (set up with the model, read by the kernels of every
// evaluation): always the context's own buffers.
template <typename T>
T* model_table(nnpx_ctx* c, const std::string& name)
{
    auto it = c->bufs.find(name);
    if (it == c->bufs.end() || !it->second.p) fail(NNPX_ERR_STATE, "model table " + name + " missing");
    return static_cast<T*>(it->second.p);
}

template <typename T>
T* existing(nnpx_ctx* c, const std::string& name)
{
    auto& m = c->arena ? c->abufs : c->bufs;
    auto it = m.find(name);
    return it == m.end() ? nullptr : static_cast<T*>(it->second.p);
}

hipEvent_t get_event(nnpx_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    // timing-only markers: no system-scope fence (cache write-back and
    // invalidate) when they complete, which delays the next dispatch; the
    // times are read after a stream synchronization
    hipEvent_t e;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreateWithFlags");
    return e;
}

// Brackets one launcher call with events when timing is enabled.
// Every launch group is checked: a rejected launch (a dynamic LDS size the
// host guards missed, say) fails the call instead of leaving a previous
// evaluation's results in the buffers.
template <typename F>
void timed(nnpx_ctx* c, const char* name, F&& f)
{
    if (!c->timing) {
        f();
        hip_check(hipGetLastError(), name);
        return;
    }
    Timer t{name, get_event(c), get_event(c)};
    hip_check(hipEventRecord(t.a, c->stream), "hipEventRecord");
    f();
    hip_check(hipGetLastError(), name);
    hip_check(hipEventRecord(t.b, c->stream), "hipEventRecord");
    c->pending.push_back(t);
}

void collect_times(nnpx_ctx* c)
{
    if (c->pending.empty()) return;
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (auto& t : c->pending) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, t.a, t.b), "hipEventElapsedTime");
        auto& r = c->ktime[t.name];
        r.first += ms;
        r.second += 1;
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

SysView sysview(nnpx_ctx* c)
{
    SysView s{};
    s.natoms = c->N;
    s.nstruct = c->S;
    s.pos = existing<double>(c, "pos");
    s.elem = existing<int>(c, "elem");
    s.sid = existing<int>(c, "sid");
    s.st = existing<StructT>(c, "st");
    s.cap = c->cap;
    // MD lists carry a skin: the mixed kernels stage only the entries inside
    // the cutoff, in a capacity fitted to them at nnpx_md_init
    const int fit = c->md_ready ? c->md_stage_cap : c->stage_cap;
    s.stage_cap = fit > 0 && fit < c->cap ? fit : c->cap;
    s.nb_n = existing<int>(c, "nb_n");
    s.nb_j = existing<int>(c, "nb_j");
    s.nb_img = existing<int>(c, "nb_img");
    s.nb_rev = existing<int>(c, "nb_rev");
    s.flags = existing<int>(c, "flags");
    s.ncells = c->ncells;
    s.atom_cell = existing<int>(c, "atom_cell");
    s.cell_count = existing<int>(c, "cell_count");
    s.cell_rank = existing<int>(c, "cell_rank");
    s.cell_start = existing<int>(c, "cell_start");
    s.cell_atoms = existing<int>(c, "cell_atoms");
    s.cell_pos = existing<double>(c, "cell_pos");
    s.nb_dr = c->external ? existing<double>(c, "nb_dr") : nullptr;
    s.nforce = c->external ? c->ntotal : c->N;
    return s;
}

int force_entries(const nnpx_ctx* c) { return c->external ? c->ntotal : c->N; }

void need_model(nnpx_ctx* c)
{
    if (!c->has_model) fail(NNPX_ERR_STATE, "no model loaded");
}
void need_structures(nnpx_ctx* c)
{
    need_model(c);
    if (c->N <= 0) fail(NNPX_ERR_STATE, "no structures set");
}

void invalidate(nnpx_ctx* c)
{
    c->sf_valid = c->sfd_valid = c->nn_valid = c->dedg_valid = c->e_valid = c->f_valid = false;
}

double list_cutoff(const nnpx_ctx* c) { return c->hm.t.maxCutoff + c->skin; }

// Geometry tables of one structure: inverse box (Structure.cpp:518-544),
// image range (Structure::calculatePbcCopies, Structure.cpp:485-516) and the
// cell grid of the cell list.
void setup_struct(StructT& t, const double* box, bool periodic, double rc)
{
    std::memset(&t, 0, sizeof t);
    t.periodic = periodic ? 1 : 0;
    if (!periodic) return;
    for (int q = 0; q < 9; ++q) t.box[q] = box[q];
    const double(*b)[3] = reinterpret_cast<const double(*)[3]>(t.box);
    double ib[3][3];
    double invdet = b[0][0] * b[1][1] * b[2][2] + b[1][0] * b[2][1] * b[0][2] +
                    b[2][0] * b[0][1] * b[1][2] - b[2][0] * b[1][1] * b[0][2] -
                    b[0][0] * b[2][1] * b[1][2] - b[1][0] * b[0][1] * b[2][2];
    if (invdet == 0.0) fail(NNPX_ERR_ARGUMENT, "singular box");
    invdet = 1.0 / invdet;
    ib[0][0] = b[1][1] * b[2][2] - b[2][1] * b[1][2];
    ib[0][1] = b[2][1] * b[0][2] - b[0][1] * b[2][2];
    ib[0][2] = b[0][1] * b[1][2] - b[1][1] * b[0][2];
    ib[1][0] = b[2][0] * b[1][2] - b[1][0] * b[2][2];
    ib[1][1] = b[0][0] * b[2][2] - b[2][0] * b[0][2];
    ib[1][2] = b[1][0] * b[0][2] - b[0][0] * b[1][2];
    ib[2][0] = b[1][0] * b[2][1] - b[2][0] * b[1][1];
    ib[2][1] = b[2][0] * b[0][1] - b[0][0] * b[2][1];
    ib[2][2] = b[0][0] * b[1][1] - b[1][0] * b[0][1];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) t.invbox[3 * i + k] = ib[i][k] * invdet;
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    auto unit = [](double* w) {
        const double n = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        w[0] /= n; w[1] /= n; w[2] /= n;
    };
    auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double axb[3], axc[3], bxc[3];
    cross(b[0], b[1], axb); unit(axb);
    cross(b[0], b[2], axc); unit(axc);
    cross(b[1], b[2], bxc); unit(bxc);
    const double proj[3] = {std::fabs(dot(b[0], bxc)), std::fabs(dot(b[1], axc)), std::fabs(dot(b[2], axb))};
    for (int k = 0; k < 3; ++k) {
        t.pbc[k] = 0;
        while (t.pbc[k] * proj[k] <= rc) t.pbc[k]++;
        if (t.pbc[k] > 60) fail(NNPX_ERR_UNSUPPORTED, "cell too small for the cutoff (image range > 60)");
        // cells of perpendicular width >= rc (with a relative margin)
        t.nc[k] = (int)std::floor(proj[k] / (rc * (1.0 + 1e-9) + 1e-12));
    }
}

// Exact fractional coordinates as Structure::remap evaluates them.
bool inside_cell(const StructT& t, const double* r)
{
    for (int k = 0; k < 3; ++k) {
        const double f = r[0] * t.invbox[k] + r[1] * t.invbox[3 + k] + r[2] * t.invbox[6 + k];
        if (!(f >= 0.0 && f < 1.0)) return false;
    }
    return true;
}

void choose_algorithms(nnpx_ctx* c, const double* pos)
{
    c->ncells = 0;
    c->any_cell = false;
    for (int s = 0; s < c->S; ++s) {
        StructT& t = c->st_host[s];
        t.algo = 0;
        t.cell0 = 0;
        t.inside = 0;
        if (t.periodic) {
            t.inside = 1;
            for (int64_t i = c->atom_off[s]; i < c->atom_off[s + 1] && t.inside; ++i)
                t.inside = inside_cell(t, pos + 3 * i) ? 1 : 0;
        }
        const bool cells = t.inside && t.nc[0] >= 3 && t.nc[1] >= 3 && t.nc[2] >= 3 && c->nbr_algo != 1;
        if (cells) {
            t.algo = 1;
            t.cell0 = c->ncells;
            c->ncells += t.nc[0] * t.nc[1] * t.nc[2];
            c->any_cell = true;
        }
    }
}

int estimate_cap(nnpx_ctx* c)
{
    const double rc = list_cutoff(c);
    int cap = 32;
    for (int s = 0; s < c->S; ++s) {
        const StructT& t = c->st_host[s];
        const int64_t n = c->atom_off[s + 1] - c->atom_off[s];
        double est;
        if (t.periodic) {
            const double* b = t.box;
            const double vol = std::fabs(b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                                         b[2] * (b[3] * b[7] - b[4] * b[6]));
            est = 4.0 / 3.0 * M_PI * rc * rc * rc * n / vol;
            const double maxn = (double)n * (2 * t.pbc[0] + 1) * (2 * t.pbc[1] + 1) * (2 * t.pbc[2] + 1);
            est = std::min(maxn, 1.5 * est + 32.0);
        } else {
            est = (double)n;
        }
        cap = std::max(cap, (int)est);
    }
    return ((cap + 31) / 32) * 32;
}

void alloc_neighbors(nnpx_ctx* c)
{
    const size_t n = (size_t)c->N * c->cap;
    buf<int>(c, "nb_n", c->N);
    buf<int>(c, "nb_j", n);
    buf<int>(c, "nb_img", n);
    buf<int>(c, "nb_rev", n);
    buf<int>(c, "flags", 4);
    buf<int>(c, "atom_cell", c->N);
    buf<int>(c, "cell_count", std::max(c->ncells, 1));
    buf<int>(c, "cell_rank", c->N);
    buf<int>(c, "cell_start", c->ncells + 1);
    buf<int>(c, "cell_atoms", c->N);
    buf<double>(c, "cell_pos", (size_t)3 * c->N);
}

// The four device flags: overflow bits, most list entries, most staged
// entries, most angular pairs per atom.
void read_flags(nnpx_ctx* c, int (&f)[4])
{
    hip_check(hipMemcpyAsync(f, existing<int>(c, "flags"), 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// Sizes the angular pair lists for the most pairs per atom of the last
// evaluation (an eighth and 64 entries of headroom) within kTripBudget.  Not
// while an MD run holds captured kernels (their list addresses stay) nor in
// the no-allocation scratch.
void fit_trips(nnpx_ctx* c, int most)
{
    if (c->md_ready || c->arena || most <= c->trip_cap) return;
    const int cap = ((most + most / 8 + 64 + 63) / 64) * 64;
    if ((size_t)c->N * cap * sizeof(unsigned) <= kTripBudget) c->trip_cap = cap;
}

// Builds the list; grows the capacity and rebuilds when an atom overflowed.
void build_neighbors(nnpx_ctx* c, bool check)
{
    if (c->external) {  // the caller's list stays
        c->nbr_valid = true;
        return;
    }
    for (int attempt = 0; attempt < 3; ++attempt) {
        alloc_neighbors(c);
        SysView s = sysview(c);
        timed(c, "nbr", [&] { launch_neighbors(s, list_cutoff(c), c->any_cell, c->stream); });
        c->rev_valid = false;
        if (!check) break;
        int f[4];
        read_flags(c, f);
        c->max_nb = f[1];
        c->stage_cap = ((f[1] + 31) / 32) * 32;
        if (!(f[0] & 1)) break;
        if (f[1] > 1024) fail(NNPX_ERR_UNSUPPORTED, "more than 1024 neighbours per atom");
        c->cap = ((f[1] + f[1] / 4 + 31) / 32) * 32;
    }
    c->nbr_valid = true;
}

// FP64 weights in model layout followed by a transposed copy ([in][out]),
// and the BF16/FP32 tables of the mixed form.
void upload_weights(nnpx_ctx* c)
{
    const ModelT& t = c->hm.t;
    std::vector<double> w2(2 * (size_t)t.nweights, 0.0);
    std::copy(c->hm.weights.begin(), c->hm.weights.end(), w2.begin());
    for (int e = 0; e < t.nelem; ++e) {
        const NNT& N = t.el[e].nn;
        for (int l = 1; l < N.nl; ++l)
            for (int k = 0; k < N.n[l]; ++k)
                for (int j = 0; j < N.n[l - 1]; ++j)
                    w2[t.nweights + N.woff[l] + (size_t)j * N.n[l] + k] =
                        c->hm.weights[N.woff[l] + (size_t)k * N.n[l - 1] + j];
    }
    if (!c->d_w) hip_check(hipMalloc(&c->d_w, sizeof(double) * w2.size()), "hipMalloc weights");
    hip_check(hipMemcpy(c->d_w, w2.data(), sizeof(double) * w2.size(), hipMemcpyHostToDevice), "hipMemcpy");
    setup_mixed(c->mixed, c->hm);
}

// Connections per element in NeuralNetwork::setConnections order
// (NeuralNetwork.cpp:147-169): per layer the weights j-major, then biases.
int num_connections(const NNT& N)
{
    int n = 0;
    for (int l = 1; l < N.nl; ++l) n += N.n[l - 1] * N.n[l] + N.n[l];
    return n;
}

void compute_sf64(nnpx_ctx* c, bool deriv)
{
    const ModelT& t = c->hm.t;
    SF64Out o{};
    o.ldG = t.max_sf;
    o.S = std::max(t.max_sfd, 1);
    o.G = buf<double>(c, "G", (size_t)c->N * o.ldG);
    if (deriv) {
        o.dGself = buf<double>(c, "dGself", (size_t)c->N * o.ldG * 3);
        o.dGnbr = buf<double>(c, "dGnbr", (size_t)c->N * c->cap * o.S * 3);
    }
    SysView s = sysview(c);
    if (!deriv && c->plain_forces == 0 && fused64_lds(c->cap) <= 65536) {
        // values only: the compacted triplet walk of the fused kernels
        timed(c, "sf64v", [&] {
            launch_values64(c->d_model, t, c->mixed, s, o.G, o.ldG, c->stream);
            launch_sfcomp64(c->d_model, t, s, o, false, c->cap, c->stream);
        });
    } else {
        timed(c, "sf64", [&] {
            launch_sf64(c->d_model, t, s, o, deriv, c->cap, c->stream);
            launch_sfcomp64(c->d_model, t, s, o, deriv, c->cap, c->stream);
        });
    }
    c->sf_valid = true;
    c->sfd_valid = deriv;
}

void compute_nn64(nnpx_ctx* c, bool deriv)
{
    const ModelT& t = c->hm.t;
    NNIO io{};
    io.ldG = t.max_sf;
    io.G = existing<double>(c, "G");
    io.energy = buf<double>(c, "Eatom", c->N);
    io.dEdG = deriv ? buf<double>(c, "dEdG", (size_t)c->N * io.ldG) : nullptr;
    const int* atoms = existing<int>(c, "elem_atoms");
    timed(c, "nn64", [&] {
        launch_nn64(c->d_model, t, c->d_w, atoms, c->elem_count.data(), c->elem_off.data(), io, c->stream);
    });
    c->nn_valid = true;
    c->dedg_valid = deriv;
}

// Plain-form forces without stored dG/dr (the fused FP64 kernel); needs the
// values G (compute_sf64 without derivatives) and dE/dG.
void compute_force64_fused(nnpx_ctx* c)
{
    const ModelT& t = c->hm.t;
    double* f = buf<double>(c, "F", (size_t)3 * force_entries(c));
    double* w = buf<double>(c, "w64", (size_t)c->N * t.max_sf);
    SysView s = sysview(c);
    timed(c, "force64f", [&] {
        launch_force64_fused(c->d_model, t, c->mixed, s, existing<double>(c, "dEdG"), model_table<double>(c, "fac64"), w, t.max_sf, f,
                             c->external || c->force_mode == NNPX_FORCES_PUSH, c->stream);
    });
    c->f_valid = true;
}

// The fused FP64 force kernel stages the whole list in LDS.
bool fused_fits(const nnpx_ctx* c) { return c->plain_forces == 0 && fused64_lds(c->cap) <= 65536; }

void compute_energy(nnpx_ctx* c)
{
    double* E = buf<double>(c, "E", c->S);
    timed(c, "energy", [&] { launch_energy(existing<double>(c, "Eatom"), existing<int>(c, "atom_off"), c->S, E, c->stream); });
    c->e_valid = true;
}

void compute_force64(nnpx_ctx* c)
{
    const ModelT& t = c->hm.t;
    double* f = buf<double>(c, "F", (size_t)3 * force_entries(c));
    SysView s = sysview(c);
    if (c->external) {
        timed(c, "force64", [&] {
            launch_force64_push(c->d_model, s, existing<double>(c, "dEdG"), t.max_sf, existing<double>(c, "dGself"),
                                existing<double>(c, "dGnbr"), std::max(t.max_sfd, 1), f, c->stream);
        });
        c->f_valid = true;
        return;
    }
    if (!c->rev_valid) {
        timed(c, "rev", [&] { launch_reverse(s, c->stream); });
        c->rev_valid = true;
    }
    timed(c, "force64", [&] {
        launch_force64(c->d_model, s, existing<double>(c, "dEdG"), t.max_sf, existing<double>(c, "dGself"),
                       existing<double>(c, "dGnbr"), std::max(t.max_sfd, 1),
                       c->force_mode == NNPX_FORCES_PUSH ? 1 : 0, f, c->stream);
    });
    c->f_valid = true;
}

// store: the plain form keeps the stored dG/dr (the gradient kernels read it)
// instead of the fused force kernel.
void evaluate(nnpx_ctx* c, int precision, unsigned flags, bool check, bool store = false)
{
    const bool forces = flags & NNPX_EVAL_FORCES;
    invalidate(c);
    if (c->md_ready) {
        // MD leaves positions unwrapped between rebuilds: wrap them (and
        // reset the displacement reference) before a full list build.
        int* gate = existing<int>(c, "md_gate");
        md_set_gate(gate, 1, c->stream);
        md_wrap(existing<double>(c, "pos"), existing<double>(c, "md_ref"), existing<int>(c, "sid"),
                existing<StructT>(c, "st"), c->N, gate, c->stream);
        md_set_gate(gate, 0, c->stream);
    }
    build_neighbors(c, check);
    // an external list has no build kernel to clear the flags the mixed
    // kernels set
    if (c->external) fill_int(existing<int>(c, "flags"), 0, 4, c->stream);
    if (precision == NNPX_MIXED) {
        MixedBuffers mb = mixed_buffers(c);
        run_mixed(c->mixed, c->hm.t, c->d_model, sysview(c), mb, forces,
                  c->external || c->force_mode == NNPX_FORCES_PUSH, c->stream,
                  [&](const char* name, const std::function<void()>& f) { timed(c, name, f); });
        c->sf_valid = true;
        c->nn_valid = true;
        c->dedg_valid = true;
        c->e_valid = true;
        c->f_valid = forces;
    } else {
        const bool fused = forces && !store && fused_fits(c);
        compute_sf64(c, forces && !fused);
        compute_nn64(c, true);
        compute_energy(c);
        if (forces) {
            if (fused) compute_force64_fused(c);
            else compute_force64(c);
        }
        c->last_stored = forces && !fused;
    }
    c->last_precision = precision;
}

// Results are read only after the overflow flag of the last neighbour build
// was checked; an overflow re-runs the evaluation with a larger capacity.
void settle(nnpx_ctx* c)
{
    int f[4];
    read_flags(c, f);
    fit_trips(c, f[3]);
    if (f[0] & 2) fail(NNPX_ERR_HIP, "neighbour list is not symmetric");
    if (f[0] & 1) {
        c->cap = ((f[1] + f[1] / 4 + 31) / 32) * 32;
        evaluate(c, c->last_precision, c->f_valid ? NNPX_EVAL_FORCES : 0u, true, c->last_stored);
    } else if (f[0] & 4) {
        // an atom had more neighbours than the mixed kernels staged: the
        // checked rebuild refits the staging capacity to this list
        evaluate(c, c->last_precision, c->f_valid ? NNPX_EVAL_FORCES : 0u, true, c->last_stored);
    } else {
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        return;
    }
    read_flags(c, f);
    if (f[0] & 7) fail(NNPX_ERR_HIP, "neighbour list or staging overflow persists after a refit");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// Structures of nnpx_set_structures (copied to device buffers; the first
// neighbour build sizes the list capacity).
void set_structures(nnpx_ctx* c, int ns, const int64_t* off, const double* pos, const int* elem, const double* box,
                    const int* periodic)
{
    need_model(c);
    if (ns < 1 || !off || !pos || !elem || !periodic) fail(NNPX_ERR_ARGUMENT, "null or empty structure input");
    if (off[0] != 0) fail(NNPX_ERR_ARGUMENT, "atom_offsets[0] must be 0");
    for (int s = 0; s < ns; ++s)
        if (off[s + 1] < off[s]) fail(NNPX_ERR_ARGUMENT, "atom_offsets must be non-decreasing");
    const int64_t N = off[ns];
    if (N <= 0 || N > (int64_t)1 << 30) fail(NNPX_ERR_ARGUMENT, "bad atom count");
    // a new structure set ends any MD set-up (skin, graph) and any
    // caller-provided neighbour list
    c->external = false;
    c->ntotal = 0;
    c->skin = 0.0;
    c->md_ready = false;
    drop_graph(c);
    for (int64_t i = 0; i < N; ++i)
        if (elem[i] < 0 || elem[i] >= c->hm.t.nelem) fail(NNPX_ERR_RANGE, "element index out of range");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->N = (int)N;
    c->S = ns;
    c->atom_off.assign(off, off + ns + 1);
    c->elem_host.assign(elem, elem + N);
    c->st_host.assign(ns, StructT{});
    for (int s = 0; s < ns; ++s) {
        if (periodic[s] && !box) fail(NNPX_ERR_ARGUMENT, "periodic structure without box");
        setup_struct(c->st_host[s], periodic[s] ? box + 9 * s : nullptr, periodic[s] != 0, list_cutoff(c));
        c->st_host[s].atom0 = (int)off[s];
        c->st_host[s].natoms = (int)(off[s + 1] - off[s]);
    }
    choose_algorithms(c, pos);
    std::vector<int> sid(N);
    for (int s = 0; s < ns; ++s)
        for (int64_t i = off[s]; i < off[s + 1]; ++i) sid[i] = s;
    const int ne = c->hm.t.nelem;
    c->elem_count.assign(ne, 0);
    c->elem_off.assign(ne + 1, 0);
    for (int64_t i = 0; i < N; ++i) c->elem_count[elem[i]]++;
    for (int e = 0; e < ne; ++e) c->elem_off[e + 1] = c->elem_off[e] + c->elem_count[e];
    std::vector<int> atoms(N);
    {
        std::vector<int> fillp(c->elem_off.begin(), c->elem_off.end() - 1);
        for (int64_t i = 0; i < N; ++i) atoms[fillp[elem[i]]++] = (int)i;
    }
    std::vector<int> off32(ns + 1);
    for (int s = 0; s <= ns; ++s) off32[s] = (int)off[s];
    auto up = [&](const std::string& name, const void* src, size_t bytes) {
        void* d = buf<char>(c, name, bytes);
        hip_check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    };
    up("pos", pos, sizeof(double) * 3 * N);
    up("elem", elem, sizeof(int) * N);
    up("sid", sid.data(), sizeof(int) * N);
    up("st", c->st_host.data(), sizeof(StructT) * ns);
    up("elem_atoms", atoms.data(), sizeof(int) * N);
    up("atom_off", off32.data(), sizeof(int) * (ns + 1));
    c->cap = estimate_cap(c);
    c->nbr_valid = false;
    invalidate(c);
    // First build sizes the capacity to the actual maximum.
    build_neighbors(c, true);
    const int want = ((c->max_nb + c->max_nb / 8 + 8 + 31) / 32) * 32;
    if (want < c->cap) {
        c->cap = std::max(want, 32);
        build_neighbors(c, true);
    }
}

void copy_out(nnpx_ctx* c, void* dst, const void* src, size_t bytes)
{
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // namespace

MixedBuffers mixed_buffers(nnpx_ctx* c)
{
    const ModelT& t = c->hm.t;
    MixedBuffers m{};
    m.ldG = t.max_sf;
    m.G = buf<double>(c, "G", (size_t)c->N * m.ldG);
    m.Gin = buf<float>(c, "Gin", (size_t)c->N * c->mixed.ldK);
    m.dEdG = buf<double>(c, "dEdG", (size_t)c->N * m.ldG);
    m.dEdGf = buf<float>(c, "dEdGf", (size_t)c->N * m.ldG);
    m.Eatom = buf<double>(c, "Eatom", c->N);
    m.E = buf<double>(c, "E", c->S);
    m.F = buf<double>(c, "F", (size_t)3 * force_entries(c));
    m.elem_atoms = existing<int>(c, "elem_atoms");
    m.atom_off = existing<int>(c, "atom_off");
    for (int e = 0; e <= t.nelem; ++e) m.elem_off[e] = c->elem_off[e];
    m.nstruct = c->S;
    m.cap = c->cap;
    // the staging shared by the SF and force kernels, one slot per list entry
    const size_t ks = (size_t)c->N * c->cap;
    if (!c->arena && ks * (sizeof(float4) + sizeof(float) + sizeof(int) + 1) <= kTripBudget) {
        m.keep.ld = c->cap;
        m.keep.g = buf<float4>(c, "keep_g", ks);
        m.keep.rinv = buf<float>(c, "keep_r", ks);
        m.keep.nbr = buf<int>(c, "keep_j", ks);
        m.keep.el = buf<unsigned char>(c, "keep_e", ks);
        m.keep.off = buf<int>(c, "keep_o", (size_t)c->N * (kMaxElem + 1));
    }
    const size_t tl = (size_t)c->N * c->trip_cap;
    if (c->trip_cap > 0 && c->mixed.tune.pair_lists == 1 && !c->arena && tl * sizeof(unsigned) <= kTripBudget) {
        m.trips.cap = c->trip_cap;
        m.trips.ld = c->mixed.max_walks;
        m.trips.list = buf<unsigned>(c, "trips", tl);
        m.trips.counts = buf<int>(c, "trip_n", (size_t)c->N * m.trips.ld);
    }
    return m;
}

// ======================================================================== ABI
extern "C" {

nnpx_status nnpx_create(nnpx_ctx** ctx, int device)
{
    if (!ctx) return NNPX_ERR_ARGUMENT;
    *ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_create_error = "no HIP device available";
        return NNPX_ERR_HIP;
    }
    if (device < 0 || device >= n) {
        g_create_error = "device index out of range";
        return NNPX_ERR_ARGUMENT;
    }
    nnpx_ctx* c = new nnpx_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "could not create a HIP stream";
        delete c;
        return NNPX_ERR_HIP;
    }
    const hipError_t se = scratch_address(device, &c->scratch);
    if (se != hipSuccess) {
        g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NNPX_ERR_HIP;
    }
    *ctx = c;
    return NNPX_OK;
}

nnpx_status nnpx_destroy(nnpx_ctx* c)
{
    if (!c) return NNPX_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    if (c->d_model) (void)hipFree(c->d_model);
    if (c->d_w) (void)hipFree(c->d_w);
    free_mixed(c->mixed);
    drop_graph(c);
    for (auto& t : c->pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NNPX_OK;
}

const char* nnpx_last_error(const nnpx_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

nnpx_status nnpx_set_option(nnpx_ctx* c, int option, int64_t value)
{
    return guard(c, [&] {
        switch (option) {
        case NNPX_OPT_FORCE_MODE:
            if (value != NNPX_FORCES_PULL && value != NNPX_FORCES_PUSH) fail(NNPX_ERR_ARGUMENT, "bad force mode");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->force_mode = (int)value;
            c->f_valid = false;
            drop_graph(c);
            break;
        case NNPX_OPT_TIMING: c->timing = value != 0; break;
        case NNPX_OPT_NEIGHBOR_ALGO:
            if (value < 0 || value > 2) fail(NNPX_ERR_ARGUMENT, "bad neighbour algorithm");
            c->nbr_algo = (int)value;
            break;
        case NNPX_OPT_BF16_TERMS:
            if (value != 1 && value != 3) fail(NNPX_ERR_ARGUMENT, "bf16 terms must be 1 or 3");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->mixed.terms = (int)value;
            drop_graph(c);  // captured MD steps hold the old value
            invalidate(c);
            break;
        case NNPX_OPT_PLAIN_FORCES:
            if (value != 0 && value != 1) fail(NNPX_ERR_ARGUMENT, "plain forces must be 0 or 1");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->plain_forces = (int)value;
            drop_graph(c);
            invalidate(c);
            break;
        case NNPX_OPT_SF_GATHER:
            if (value < -1 || value > 1) fail(NNPX_ERR_ARGUMENT, "sf gather must be -1, 0 or 1");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->mixed.tune.sf_gather = (int)value;
            drop_graph(c);
            break;
        case NNPX_OPT_MD_STAGING:
            if (value < -1 || value > 65535) fail(NNPX_ERR_ARGUMENT, "MD staging must be -1, 0 or a slot count");
            c->md_stage_req = (int)value;
            break;
        case NNPX_OPT_PAIR_LISTS:
            if (value < -1 || value > 1) fail(NNPX_ERR_ARGUMENT, "pair lists must be -1, 0 or 1");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            c->mixed.tune.pair_lists = (int)value;
            drop_graph(c);
            break;
        case NNPX_OPT_ANGULAR_MERGE_SF:
        case NNPX_OPT_ANGULAR_MERGE_FORCE:
            if (value < -1 || value > 1) fail(NNPX_ERR_ARGUMENT, "angular merge must be -1, 0 or 1");
            hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            (option == NNPX_OPT_ANGULAR_MERGE_SF ? c->mixed.tune.merge_sf : c->mixed.tune.merge_force) = (int)value;
            drop_graph(c);
            break;
        default: fail(NNPX_ERR_ARGUMENT, "unknown option");
        }
    });
}

nnpx_status nnpx_get_option(const nnpx_ctx* cc, int option, int64_t* value)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        if (!value) fail(NNPX_ERR_ARGUMENT, "null output");
        switch (option) {
        case NNPX_OPT_FORCE_MODE: *value = c->force_mode; break;
        case NNPX_OPT_TIMING: *value = c->timing ? 1 : 0; break;
        case NNPX_OPT_NEIGHBOR_ALGO: *value = c->nbr_algo; break;
        case NNPX_OPT_BF16_TERMS: *value = c->mixed.terms; break;
        case NNPX_OPT_SF_GATHER: *value = c->mixed.tune.sf_gather; break;
        case NNPX_OPT_PLAIN_FORCES: *value = c->plain_forces; break;
        case NNPX_OPT_ANGULAR_MERGE_SF: *value = c->mixed.tune.merge_sf; break;
        case NNPX_OPT_ANGULAR_MERGE_FORCE: *value = c->mixed.tune.merge_force; break;
        case NNPX_OPT_MD_STAGING: *value = c->md_stage_req; break;
        case NNPX_OPT_PAIR_LISTS: *value = c->mixed.tune.pair_lists; break;
        default: fail(NNPX_ERR_ARGUMENT, "unknown option");
        }
    });
}

nnpx_status nnpx_synchronize(nnpx_ctx* c)
{
    return guard(c, [&] { hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); });
}

nnpx_status nnpx_load_model(nnpx_ctx* c, const char* dir)
{
    return guard(c, [&] {
        if (!dir) fail(NNPX_ERR_ARGUMENT, "null model directory");
        HostModel hm;
        load_model(dir, hm);
        c->hm = std::move(hm);
        c->skin = 0.0;
        c->md_ready = false;
        drop_graph(c);
        const ModelT& t = c->hm.t;
        if (!c->d_model) hip_check(hipMalloc(&c->d_model, sizeof(ModelT)), "hipMalloc model");
        hip_check(hipMemcpy(c->d_model, &t, sizeof(ModelT), hipMemcpyHostToDevice), "hipMemcpy model");
        if (c->d_w) hip_check(hipFree(c->d_w), "hipFree");
        c->d_w = nullptr;
        upload_weights(c);
        // SF -> derivative slot per (centre element, neighbour element)
        std::vector<int> islot((size_t)kMaxElem * kMaxElem * kMaxSF, -1);
        for (int e = 0; e < t.nelem; ++e)
            for (int n = 0; n < t.nelem; ++n)
                for (int q = 0; q < t.el[e].nsfd[n]; ++q)
                    islot[((size_t)e * kMaxElem + n) * kMaxSF + t.table[t.el[e].tab0[n] + q]] = q;
        int* di = buf<int>(c, "islot", islot.size());
        hip_check(hipMemcpy(di, islot.data(), sizeof(int) * islot.size(), hipMemcpyHostToDevice), "hipMemcpy");
        // derivative scale of every SF for the fused FP64 forces: Gaussian
        // angular members 2^(1-zeta) times the scaling factor, all others
        // the scaling factor
        std::vector<double> fac((size_t)kMaxElem * kMaxSF, 0.0);
        for (int e = 0; e < t.nelem; ++e)
            for (int g = t.el[e].grp0; g < t.el[e].grp0 + t.el[e].ngrp; ++g) {
                const GroupT& G = t.grp[g];
                const bool ang = G.type == 3 || G.type == 9 || G.type == 13;
                for (int q = G.mem0; q < G.mem0 + G.nmem; ++q)
                    fac[(size_t)e * kMaxSF + t.mem[q].sf] = ang ? t.mem[q].factorNorm : t.mem[q].sfac;
            }
        double* dfac = buf<double>(c, "fac64", fac.size());
        hip_check(hipMemcpy(dfac, fac.data(), sizeof(double) * fac.size(), hipMemcpyHostToDevice), "hipMemcpy");
        c->has_model = true;
        c->N = c->S = 0;
        c->nbr_valid = false;
        invalidate(c);
    });
}

nnpx_status nnpx_model_info(const nnpx_ctx* cc, int* ne, int* max_sf, double* rc)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (ne) *ne = c->hm.t.nelem;
        if (max_sf) *max_sf = c->hm.t.max_sf;
        if (rc) *rc = c->hm.t.maxCutoff;
    });
}

nnpx_status nnpx_num_sf(const nnpx_ctx* cc, int e, int* n)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (e < 0 || e >= c->hm.t.nelem || !n) fail(NNPX_ERR_ARGUMENT, "bad element index");
        *n = c->hm.t.el[e].nsf;
    });
}

nnpx_status nnpx_atomic_number(const nnpx_ctx* cc, int e, int* z)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (e < 0 || e >= c->hm.t.nelem || !z) fail(NNPX_ERR_ARGUMENT, "bad element index");
        *z = c->hm.Z[e];
    });
}

nnpx_status nnpx_element_index(const nnpx_ctx* cc, const char* sym, int* e)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (!sym || !e) fail(NNPX_ERR_ARGUMENT, "null argument");
        for (int q = 0; q < c->hm.t.nelem; ++q)
            if (c->hm.symbols[q] == sym) { *e = q; return; }
        fail(NNPX_ERR_RANGE, std::string("element not in model: ") + sym);
    });
}

nnpx_status nnpx_num_sf_derivatives(const nnpx_ctx* cc, int e, int e_nbr, int* n)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        const int ne = c->hm.t.nelem;
        if (e < 0 || e >= ne || e_nbr < 0 || e_nbr >= ne || !n) fail(NNPX_ERR_ARGUMENT, "bad element index");
        *n = c->hm.t.el[e].nsfd[e_nbr];
    });
}

nnpx_status nnpx_element_symbol(const nnpx_ctx* cc, int e, char* buf, int len)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (e < 0 || e >= c->hm.t.nelem || !buf || len <= 0) fail(NNPX_ERR_ARGUMENT, "bad element index or buffer");
        const std::string& s = c->hm.symbols[e];
        if ((int)s.size() >= len) fail(NNPX_ERR_RANGE, "symbol buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

nnpx_status nnpx_model_units(const nnpx_ctx* cc, int* normalize, double* mean, double* ce, double* cl,
                             double* offsets)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (normalize) *normalize = c->hm.normalize ? 1 : 0;
        if (mean) *mean = c->hm.meanEnergy;
        if (ce) *ce = c->hm.convEnergy;
        if (cl) *cl = c->hm.convLength;
        if (offsets)
            for (int e = 0; e < c->hm.t.nelem; ++e) offsets[e] = c->hm.energyOffset[e];
    });
}

nnpx_status nnpx_set_structures(nnpx_ctx* c, int ns, const int64_t* off, const double* pos, const int* elem,
                                const double* box, const int* periodic)
{
    return guard(c, [&] { set_structures(c, ns, off, pos, elem, box, periodic); });
}

nnpx_status nnpx_set_neighbors_external(nnpx_ctx* c, int nlocal, const int* elem, int ntotal,
                                        const int64_t* nb_offsets, const int* nb_index, const int* nb_elem,
                                        const double* nb_dr)
{
    return guard(c, [&] {
        need_model(c);
        if (nlocal <= 0 || ntotal < nlocal || !elem || !nb_offsets) fail(NNPX_ERR_ARGUMENT, "bad atom counts");
        const int ne = c->hm.t.nelem;
        const int64_t P = nb_offsets[nlocal];
        if (nb_offsets[0] != 0 || P < 0 || (P > 0 && (!nb_index || !nb_elem || !nb_dr)))
            fail(NNPX_ERR_ARGUMENT, "bad neighbour offsets or null arrays");
        std::vector<int> el_all(ntotal, -1);
        for (int i = 0; i < nlocal; ++i) {
            if (elem[i] < 0 || elem[i] >= ne) fail(NNPX_ERR_RANGE, "element index out of range");
            el_all[i] = elem[i];
        }
        int maxn = 0;
        for (int i = 0; i < nlocal; ++i) {
            const int64_t a = nb_offsets[i], b = nb_offsets[i + 1];
            if (b < a) fail(NNPX_ERR_ARGUMENT, "neighbour offsets must be non-decreasing");
            maxn = std::max<int64_t>(maxn, b - a);
            for (int64_t p = a; p < b; ++p) {
                const int j = nb_index[p];
                if (j < 0 || j >= ntotal) fail(NNPX_ERR_RANGE, "neighbour index out of range");
                if (nb_elem[p] < 0 || nb_elem[p] >= ne) fail(NNPX_ERR_RANGE, "neighbour element out of range");
                if (el_all[j] >= 0 && el_all[j] != nb_elem[p]) fail(NNPX_ERR_ARGUMENT, "inconsistent neighbour element");
                el_all[j] = nb_elem[p];
            }
        }
        if (maxn > 1024) fail(NNPX_ERR_UNSUPPORTED, "more than 1024 neighbours per atom");
        for (auto& x : el_all) x = std::max(x, 0);
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        drop_graph(c);
        c->md_ready = false;
        c->skin = 0.0;
        c->external = true;
        c->ntotal = ntotal;
        c->N = nlocal;
        c->S = 1;
        c->atom_off.assign({0, (int64_t)nlocal});
        c->elem_host = el_all;  // local atoms, then ghosts
        c->st_host.assign(1, StructT{});
        c->st_host[0].atom0 = 0;
        c->st_host[0].natoms = nlocal;
        c->any_cell = false;
        c->ncells = 0;
        c->elem_count.assign(ne, 0);
        c->elem_off.assign(ne + 1, 0);
        for (int i = 0; i < nlocal; ++i) c->elem_count[elem[i]]++;
        for (int e = 0; e < ne; ++e) c->elem_off[e + 1] = c->elem_off[e] + c->elem_count[e];
        std::vector<int> atoms(nlocal);
        {
            std::vector<int> fillp(c->elem_off.begin(), c->elem_off.end() - 1);
            for (int i = 0; i < nlocal; ++i) atoms[fillp[elem[i]]++] = i;
        }
        c->cap = std::max(32, ((maxn + 31) / 32) * 32);
        c->max_nb = maxn;
        c->stage_cap = 0;  // stage the whole list (a fit from earlier structures does not apply)
        const size_t slots = (size_t)nlocal * c->cap;
        std::vector<int> nn(nlocal), jj(slots, 0), im(slots, pack_image(0, 0, 0));
        std::vector<double> dr(3 * slots, 0.0);
        for (int i = 0; i < nlocal; ++i) {
            nn[i] = (int)(nb_offsets[i + 1] - nb_offsets[i]);
            for (int k = 0; k < nn[i]; ++k) {
                const int64_t p = nb_offsets[i] + k;
                jj[(size_t)i * c->cap + k] = nb_index[p];
                for (int x = 0; x < 3; ++x) dr[3 * ((size_t)i * c->cap + k) + x] = nb_dr[3 * p + x];
            }
        }
        std::vector<int> sid(nlocal, 0), off32 = {0, nlocal}, flags(4, 0);
        std::vector<double> pos(3 * (size_t)nlocal, 0.0);
        auto up = [&](const std::string& name, const void* src, size_t bytes) {
            void* d = buf<char>(c, name, bytes);
            hip_check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
        };
        up("pos", pos.data(), sizeof(double) * pos.size());
        up("elem", el_all.data(), sizeof(int) * ntotal);
        up("sid", sid.data(), sizeof(int) * nlocal);
        up("st", c->st_host.data(), sizeof(StructT));
        up("elem_atoms", atoms.data(), sizeof(int) * nlocal);
        up("atom_off", off32.data(), sizeof(int) * 2);
        up("nb_n", nn.data(), sizeof(int) * nlocal);
        up("nb_j", jj.data(), sizeof(int) * slots);
        up("nb_img", im.data(), sizeof(int) * slots);
        up("nb_dr", dr.data(), sizeof(double) * dr.size());
        up("flags", flags.data(), sizeof(int) * flags.size());
        buf<int>(c, "nb_rev", slots);
        invalidate(c);
        c->nbr_valid = true;
        c->rev_valid = true;
    });
}

nnpx_status nnpx_update_positions(nnpx_ctx* c, const double* pos)
{
    return guard(c, [&] {
        need_structures(c);
        if (!pos) fail(NNPX_ERR_ARGUMENT, "null positions");
        if (c->external) fail(NNPX_ERR_STATE, "positions are not used with an external neighbour list");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        bool changed = false;
        for (int s = 0; s < c->S; ++s) {
            StructT& t = c->st_host[s];
            if (!t.periodic) continue;
            int inside = 1;
            for (int64_t i = c->atom_off[s]; i < c->atom_off[s + 1] && inside; ++i)
                inside = inside_cell(t, pos + 3 * i) ? 1 : 0;
            if (t.algo == 1 && !inside) fail(NNPX_ERR_ARGUMENT, "position outside the cell (remap first)");
            // the brute-force image ranges are pruned only while every atom
            // is inside the cell
            if (inside != t.inside) {
                t.inside = inside;
                changed = true;
            }
        }
        if (changed)
            hip_check(hipMemcpy(existing<StructT>(c, "st"), c->st_host.data(), sizeof(StructT) * c->S,
                                hipMemcpyHostToDevice),
                      "hipMemcpy");
        hip_check(hipMemcpy(existing<double>(c, "pos"), pos, sizeof(double) * 3 * c->N, hipMemcpyHostToDevice),
                  "hipMemcpy");
        c->nbr_valid = false;
        invalidate(c);
    });
}

nnpx_status nnpx_calculate_neighbor_list(nnpx_ctx* c)
{
    return guard(c, [&] {
        need_structures(c);
        invalidate(c);
        build_neighbors(c, true);
    });
}

nnpx_status nnpx_calculate_symmetry_function_groups(nnpx_ctx* c, int derivatives)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->nbr_valid) build_neighbors(c, true);
        // Mode.cpp:1634-1635: skip when results are already there
        if (c->sfd_valid || (c->sf_valid && !derivatives)) return;
        compute_sf64(c, derivatives != 0);
        c->nn_valid = c->dedg_valid = c->e_valid = c->f_valid = false;
    });
}

nnpx_status nnpx_calculate_atomic_neural_networks(nnpx_ctx* c, int derivatives)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->sf_valid) fail(NNPX_ERR_STATE, "symmetry functions not calculated");
        compute_nn64(c, derivatives != 0);
        c->e_valid = c->f_valid = false;
    });
}

nnpx_status nnpx_calculate_energy(nnpx_ctx* c)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->nn_valid) fail(NNPX_ERR_STATE, "atomic neural networks not calculated");
        compute_energy(c);
    });
}

nnpx_status nnpx_calculate_forces(nnpx_ctx* c)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->sfd_valid) fail(NNPX_ERR_STATE, "symmetry function derivatives not calculated");
        if (!c->dedg_valid) fail(NNPX_ERR_STATE, "dE/dG not calculated");
        compute_force64(c);
    });
}

nnpx_status nnpx_evaluate(nnpx_ctx* c, int precision, unsigned flags)
{
    return guard(c, [&] {
        need_structures(c);
        if (precision != NNPX_PLAIN_FP64 && precision != NNPX_MIXED) fail(NNPX_ERR_ARGUMENT, "bad precision");
        evaluate(c, precision, flags, false);
    });
}

nnpx_status nnpx_get_energy(nnpx_ctx* c, double* E)
{
    return guard(c, [&] {
        need_structures(c);
        if (!E) fail(NNPX_ERR_ARGUMENT, "null output");
        if (!c->e_valid) fail(NNPX_ERR_STATE, "energy not calculated");
        settle(c);
        copy_out(c, E, existing<double>(c, "E"), sizeof(double) * c->S);
    });
}

nnpx_status nnpx_get_atomic_energies(nnpx_ctx* c, double* e)
{
    return guard(c, [&] {
        need_structures(c);
        if (!e) fail(NNPX_ERR_ARGUMENT, "null output");
        if (!c->nn_valid) fail(NNPX_ERR_STATE, "atomic neural networks not calculated");
        settle(c);
        copy_out(c, e, existing<double>(c, "Eatom"), sizeof(double) * c->N);
    });
}

nnpx_status nnpx_get_forces(nnpx_ctx* c, double* f)
{
    return guard(c, [&] {
        need_structures(c);
        if (!f) fail(NNPX_ERR_ARGUMENT, "null output");
        if (!c->f_valid) fail(NNPX_ERR_STATE, "forces not calculated");
        settle(c);
        copy_out(c, f, existing<double>(c, "F"), sizeof(double) * 3 * force_entries(c));
    });
}

static void get_rows(nnpx_ctx* c, double* out, int ld, const char* name)
{
    const int w = c->hm.t.max_sf;
    if (ld < w) fail(NNPX_ERR_ARGUMENT, "row stride smaller than the maximum SF count");
    std::vector<double> tmp((size_t)c->N * w);
    copy_out(c, tmp.data(), existing<double>(c, name), sizeof(double) * tmp.size());
    for (int i = 0; i < c->N; ++i) {
        const int ns = c->hm.t.el[c->elem_host[i]].nsf;
        for (int q = 0; q < ld; ++q) out[(size_t)i * ld + q] = q < ns ? tmp[(size_t)i * w + q] : 0.0;
    }
}

nnpx_status nnpx_get_G(nnpx_ctx* c, double* G, int ld)
{
    return guard(c, [&] {
        need_structures(c);
        if (!G) fail(NNPX_ERR_ARGUMENT, "null output");
        if (!c->sf_valid) fail(NNPX_ERR_STATE, "symmetry functions not calculated");
        settle(c);
        get_rows(c, G, ld, "G");
    });
}

nnpx_status nnpx_get_dEdG(nnpx_ctx* c, double* d, int ld)
{
    return guard(c, [&] {
        need_structures(c);
        if (!d) fail(NNPX_ERR_ARGUMENT, "null output");
        if (!c->dedg_valid) fail(NNPX_ERR_STATE, "dE/dG not calculated");
        settle(c);
        get_rows(c, d, ld, "dEdG");
    });
}

nnpx_status nnpx_get_neighbors(nnpx_ctx* c, int64_t* total, int* counts, int* j, int* shift)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->nbr_valid) build_neighbors(c, true);
        settle(c);
        std::vector<int> n(c->N);
        copy_out(c, n.data(), existing<int>(c, "nb_n"), sizeof(int) * c->N);
        int64_t tot = 0;
        for (int v : n) tot += v;
        if (total) *total = tot;
        if (counts) std::copy(n.begin(), n.end(), counts);
        if (j || shift) {
            std::vector<int> jj((size_t)c->N * c->cap), im((size_t)c->N * c->cap);
            copy_out(c, jj.data(), existing<int>(c, "nb_j"), sizeof(int) * jj.size());
            copy_out(c, im.data(), existing<int>(c, "nb_img"), sizeof(int) * im.size());
            int64_t p = 0;
            for (int i = 0; i < c->N; ++i)
                for (int k = 0; k < n[i]; ++k, ++p) {
                    const size_t e = (size_t)i * c->cap + k;
                    if (j) j[p] = jj[e];
                    if (shift) unpack_image(im[e], shift[3 * p], shift[3 * p + 1], shift[3 * p + 2]);
                }
        }
    });
}

nnpx_status nnpx_get_dGdr(nnpx_ctx* c, double* self, int ld_self, double* nbr, int ld_nbr)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->sfd_valid || c->last_precision != NNPX_PLAIN_FP64)
            fail(NNPX_ERR_STATE, "stored dG/dr not available (plain form with derivatives required)");
        settle(c);
        const ModelT& t = c->hm.t;
        if (self) {
            if (ld_self < t.max_sf) fail(NNPX_ERR_ARGUMENT, "ld_self too small");
            std::vector<double> tmp((size_t)c->N * t.max_sf * 3);
            copy_out(c, tmp.data(), existing<double>(c, "dGself"), sizeof(double) * tmp.size());
            for (int i = 0; i < c->N; ++i)
                for (int q = 0; q < ld_self; ++q)
                    for (int k = 0; k < 3; ++k)
                        self[((size_t)i * ld_self + q) * 3 + k] =
                            q < t.el[c->elem_host[i]].nsf ? tmp[((size_t)i * t.max_sf + q) * 3 + k] : 0.0;
        }
        if (nbr) {
            const int S = std::max(t.max_sfd, 1);
            if (ld_nbr < S) fail(NNPX_ERR_ARGUMENT, "ld_nbr too small");
            std::vector<int> n(c->N), jj((size_t)c->N * c->cap);
            copy_out(c, n.data(), existing<int>(c, "nb_n"), sizeof(int) * c->N);
            copy_out(c, jj.data(), existing<int>(c, "nb_j"), sizeof(int) * jj.size());
            std::vector<double> tmp((size_t)c->N * c->cap * S * 3);
            copy_out(c, tmp.data(), existing<double>(c, "dGnbr"), sizeof(double) * tmp.size());
            int64_t p = 0;
            for (int i = 0; i < c->N; ++i) {
                const ElementT& E = t.el[c->elem_host[i]];
                for (int k = 0; k < n[i]; ++k, ++p) {
                    const int ns = E.nsfd[c->elem_host[jj[(size_t)i * c->cap + k]]];
                    for (int q = 0; q < ld_nbr; ++q)
                        for (int x = 0; x < 3; ++x)
                            nbr[((size_t)p * ld_nbr + q) * 3 + x] =
                                q < ns ? tmp[(((size_t)i * c->cap + k) * S + q) * 3 + x] : 0.0;
                }
            }
        }
    });
}

}  // extern "C"

namespace {

// Two passes of an append kernel: the first counts, the second fills buffers
// of that size; returns the count.
template <typename Launch>
int64_t append_records(nnpx_ctx* c, Launch&& launch, int64_t want)
{
    unsigned long long* cnt = buf<unsigned long long>(c, "rec_count", 1);
    hip_check(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream), "hipMemsetAsync");
    launch(cnt, 0ll);
    unsigned long long n = 0;
    copy_out(c, &n, cnt, sizeof n);
    if (n > 0 && want > 0) {
        hip_check(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->stream), "hipMemsetAsync");
        launch(cnt, (long long)n);
        hip_check(hipGetLastError(), "append kernel");
    }
    return (int64_t)n;
}

}  // namespace

extern "C" {

nnpx_status nnpx_extrapolation_warnings(nnpx_ctx* c, int64_t max, int64_t* count, int* atom, int* sf, int* type,
                                        double* value, double* gmin, double* gmax)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->sf_valid) fail(NNPX_ERR_STATE, "symmetry functions not calculated");
        if (max < 0 || (max > 0 && !atom)) fail(NNPX_ERR_ARGUMENT, "bad record buffers");
        settle(c);
        const ModelT& t = c->hm.t;
        const size_t total = (size_t)c->N * t.max_sf;
        int* ra = buf<int>(c, "rec_atom", total);
        int* rs = buf<int>(c, "rec_sf", total);
        double* rv = buf<double>(c, "rec_val", total);
        const int64_t n = append_records(c, [&](unsigned long long* cnt, long long cap) {
            launch_extrapolation(c->d_model, c->N, existing<int>(c, "elem"), existing<double>(c, "G"), t.max_sf, cnt,
                                 cap, ra, rs, rv, c->stream);
        }, max);
        if (count) *count = n;
        if (n == 0 || max == 0) return;
        std::vector<int> a(n), q(n);
        std::vector<double> v(n);
        copy_out(c, a.data(), ra, sizeof(int) * n);
        copy_out(c, q.data(), rs, sizeof(int) * n);
        copy_out(c, v.data(), rv, sizeof(double) * n);
        std::vector<int64_t> order(n);
        for (int64_t k = 0; k < n; ++k) order[k] = k;
        std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
            return a[x] != a[y] ? a[x] < a[y] : q[x] < q[y];
        });
        // SF type of every (element, SF)
        std::vector<int> sft((size_t)kMaxElem * kMaxSF, 0);
        for (int e = 0; e < t.nelem; ++e)
            for (int g = t.el[e].grp0; g < t.el[e].grp0 + t.el[e].ngrp; ++g)
                for (int m = t.grp[g].mem0; m < t.grp[g].mem0 + t.grp[g].nmem; ++m)
                    sft[(size_t)e * kMaxSF + t.mem[m].sf] = t.grp[g].type;
        const int64_t k = std::min<int64_t>(max, n);
        for (int64_t r = 0; r < k; ++r) {
            const int64_t x = order[r];
            const int e = c->elem_host[a[x]];
            const ScaleT& S = t.scale[t.el[e].sf0 + q[x]];
            atom[r] = a[x];
            if (sf) sf[r] = q[x];
            if (type) type[r] = sft[(size_t)e * kMaxSF + q[x]];
            if (value) value[r] = v[x];
            if (gmin) gmin[r] = S.Gmin;
            if (gmax) gmax[r] = S.Gmax;
        }
    });
}

nnpx_status nnpx_low_neighbor_atoms(nnpx_ctx* c, int64_t max, int64_t* count, int* atom, int* num_neighbors)
{
    return guard(c, [&] {
        need_structures(c);
        if (max < 0 || (max > 0 && !atom)) fail(NNPX_ERR_ARGUMENT, "bad record buffers");
        if (!c->nbr_valid) build_neighbors(c, true);
        settle(c);
        const ModelT& t = c->hm.t;
        // Element::getMinCutoffRadius / getMinNeighbors (Element.cpp:384-411)
        std::vector<double> rmin(kMaxElem, 0.0);
        std::vector<int> nmin(kMaxElem, 0);
        for (int e = 0; e < t.nelem; ++e) {
            double r = 1e300;
            for (int g = t.el[e].grp0; g < t.el[e].grp0 + t.el[e].ngrp; ++g) {
                const int ty = t.grp[g].type;
                const bool radial = ty == 2 || ty == 12 || ty == 20 || ty == 23;
                nmin[e] = std::max(nmin[e], radial ? 1 : 2);
                r = std::min(r, t.grp[g].rc);
            }
            rmin[e] = t.el[e].ngrp ? r : 0.0;
        }
        double* drmin = buf<double>(c, "low_rmin", kMaxElem);
        int* dnmin = buf<int>(c, "low_nmin", kMaxElem);
        hip_check(hipMemcpyAsync(drmin, rmin.data(), sizeof(double) * kMaxElem, hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(dnmin, nmin.data(), sizeof(int) * kMaxElem, hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync");
        int* ra = buf<int>(c, "rec_atom", c->N);
        int* rn = buf<int>(c, "rec_sf", c->N);
        SysView s = sysview(c);
        const int64_t n = append_records(c, [&](unsigned long long* cnt, long long cap) {
            launch_low_neighbors(c->d_model, s, drmin, dnmin, cnt, cap, ra, rn, c->stream);
        }, max);
        if (count) *count = n;
        if (n == 0 || max == 0) return;
        std::vector<int> a(n), m(n);
        copy_out(c, a.data(), ra, sizeof(int) * n);
        copy_out(c, m.data(), rn, sizeof(int) * n);
        std::vector<int64_t> order(n);
        for (int64_t k = 0; k < n; ++k) order[k] = k;
        std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return a[x] < a[y]; });
        const int64_t k = std::min<int64_t>(max, n);
        for (int64_t r = 0; r < k; ++r) {
            atom[r] = a[order[r]];
            if (num_neighbors) num_neighbors[r] = m[order[r]];
        }
    });
}

nnpx_status nnpx_num_connections(const nnpx_ctx* cc, int element, int* n)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (!n || element < -1 || element >= c->hm.t.nelem) fail(NNPX_ERR_ARGUMENT, "bad element or null output");
        int total = 0;
        for (int e = 0; e < c->hm.t.nelem; ++e)
            if (element < 0 || element == e) total += num_connections(c->hm.t.el[e].nn);
        *n = total;
    });
}

}  // extern "C"

namespace {

// The connections of all networks in NeuralNetwork::setConnections order
// (NeuralNetwork.cpp:147-169), elements in order.
void get_connections(const nnpx_ctx* c, double* w)
{
    size_t p = 0;
    for (int e = 0; e < c->hm.t.nelem; ++e) {
        const NNT& N = c->hm.t.el[e].nn;
        for (int l = 1; l < N.nl; ++l) {
            for (int j = 0; j < N.n[l - 1]; ++j)
                for (int k = 0; k < N.n[l]; ++k) w[p++] = c->hm.weights[N.woff[l] + (size_t)k * N.n[l - 1] + j];
            for (int k = 0; k < N.n[l]; ++k) w[p++] = c->hm.weights[N.boff[l] + k];
        }
    }
}

void set_connections(nnpx_ctx* c, const double* w)
{
    hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    size_t p = 0;
    for (int e = 0; e < c->hm.t.nelem; ++e) {
        const NNT& N = c->hm.t.el[e].nn;
        for (int l = 1; l < N.nl; ++l) {
            for (int j = 0; j < N.n[l - 1]; ++j)
                for (int k = 0; k < N.n[l]; ++k) c->hm.weights[N.woff[l] + (size_t)k * N.n[l - 1] + j] = w[p++];
            for (int k = 0; k < N.n[l]; ++k) c->hm.weights[N.boff[l] + k] = w[p++];
        }
    }
    upload_weights(c);
    drop_graph(c);
    invalidate(c);
}

}  // namespace

extern "C" {

nnpx_status nnpx_get_weights(const nnpx_ctx* cc, double* w)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        need_model(c);
        if (!w) fail(NNPX_ERR_ARGUMENT, "null output");
        get_connections(c, w);
    });
}

nnpx_status nnpx_set_weights(nnpx_ctx* c, const double* w)
{
    return guard(c, [&] {
        need_model(c);
        if (!w) fail(NNPX_ERR_ARGUMENT, "null weights");
        set_connections(c, w);
    });
}

namespace {

// Weight vector-Jacobian product of the evaluated structures: with the seeds
// u (3N, device) on the forces and cE (S, device) on the energies,
//   grad += sum_s cE_s dE_s/dc + sum_{a,c} u_ac dF_ac/dc,
// the tangent Gdot = sum_ac u_ac dG/dx_ac (Training::collectDGdxia,
// Training.cpp:3586-3625) followed by the network's forward tangent and one
// reverse pass (NeuralNetwork::calculateDEdc / calculateDFdc,
// NeuralNetwork.cpp:444-527).  ldgrad = 0 sums over structures; otherwise
// structure s adds into gd + s * ldgrad.  gd must be cleared by the caller.
void weight_vjp(nnpx_ctx* c, int precision, const double* u, const double* cE, double* gd, int ldgrad)
{
    const ModelT& t = c->hm.t;
    const int N = c->N, ldG = t.max_sf;
    double* Gdot = buf<double>(c, "t_Gdot", (size_t)N * ldG);
    hip_check(hipMemsetAsync(Gdot, 0, sizeof(double) * (size_t)N * ldG, c->stream), "hipMemsetAsync");
    SysView s = sysview(c);
    if (precision == NNPX_PLAIN_FP64) {
        timed(c, "tangent64", [&] {
            launch_tangent64(c->d_model, s, existing<double>(c, "dGself"), existing<double>(c, "dGnbr"),
                             std::max(t.max_sfd, 1), model_table<int>(c, "islot"), u, Gdot, ldG, c->stream);
        });
    } else {
        timed(c, "tangent32", [&] { launch_tangent32(c->mixed, t, c->d_model, s, u, Gdot, ldG, c->stream); });
    }
    GradArgs a{};
    int total = 0;
    for (int e = 0; e < t.nelem; ++e) {
        a.coff[e] = total;
        a.nconn[e] = num_connections(t.el[e].nn);
        a.nconn_max = std::max(a.nconn_max, a.nconn[e]);
        total += a.nconn[e];
    }
    a.w = c->d_w;
    a.nweights = t.nweights;
    a.G = existing<double>(c, "G");
    a.Gdot = Gdot;
    a.ldG = ldG;
    a.cE = cE;
    a.sid = existing<int>(c, "sid");
    a.atoms = existing<int>(c, "elem_atoms");
    a.nelem = t.nelem;
    a.maxw = std::max(t.max_width, t.max_sf);
    a.per_block = c->mixed.tune.grad_atoms;
    a.grad = gd;
    a.ldgrad = ldgrad;
    int blocks = 0;
    for (int e = 0; e < t.nelem; ++e) {
        a.blk_off[e] = blocks;
        a.atom_off[e] = c->elem_off[e];
        blocks += (c->elem_count[e] + a.per_block - 1) / a.per_block;
    }
    a.blk_off[t.nelem] = blocks;
    a.atom_off[t.nelem] = c->elem_off[t.nelem];
    if (grad_act_bytes(a, precision == NNPX_MIXED) > 65536)
        fail(NNPX_ERR_UNSUPPORTED, "network too wide for the gradient kernel's LDS staging");
    timed(c, precision == NNPX_MIXED ? "grad32" : "grad64",
          [&] { launch_grad(c->d_model, a, blocks, precision == NNPX_MIXED, c->stream); });
}

int total_connections(const ModelT& t)
{
    int total = 0;
    for (int e = 0; e < t.nelem; ++e) total += num_connections(t.el[e].nn);
    return total;
}

void gradient_checks(nnpx_ctx* c, int precision)
{
    need_structures(c);
    if (precision != NNPX_PLAIN_FP64 && precision != NNPX_MIXED) fail(NNPX_ERR_ARGUMENT, "bad precision");
    const ModelT& t = c->hm.t;
    for (int e = 0; e < t.nelem; ++e)
        if (t.el[e].nn.normalize) fail(NNPX_ERR_UNSUPPORTED, "weight derivatives with normalize_nodes");
}

}  // namespace

nnpx_status nnpx_loss_gradient(nnpx_ctx* c, int precision, const double* E_ref, const double* F_ref, double beta,
                               double* loss, double* grad)
{
    return guard(c, [&] {
        gradient_checks(c, precision);
        if (!E_ref || !F_ref || !loss || !grad) fail(NNPX_ERR_ARGUMENT, "null argument");
        if (!(beta >= 0.0)) fail(NNPX_ERR_ARGUMENT, "beta must be >= 0");
        evaluate(c, precision, NNPX_EVAL_FORCES, true, precision == NNPX_PLAIN_FP64);
        settle(c);
        const int N = c->N, S = c->S;
        double* Eref = buf<double>(c, "t_Eref", S);
        double* Fref = buf<double>(c, "t_Fref", (size_t)3 * N);
        hip_check(hipMemcpyAsync(Eref, E_ref, sizeof(double) * S, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(Fref, F_ref, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream),
                  "hipMemcpyAsync");
        double* u = buf<double>(c, "t_u", (size_t)3 * N);
        double* cE = buf<double>(c, "t_cE", S);
        double* lossd = buf<double>(c, "t_loss", 4);
        const int total = total_connections(c->hm.t);
        double* gd = buf<double>(c, "t_grad", total);
        hip_check(hipMemsetAsync(lossd, 0, 4 * sizeof(double), c->stream), "hipMemsetAsync");
        hip_check(hipMemsetAsync(gd, 0, sizeof(double) * total, c->stream), "hipMemsetAsync");
        timed(c, "residual", [&] {
            launch_residual(existing<double>(c, "F"), Fref, existing<double>(c, "E"), Eref, existing<int>(c, "atom_off"),
                            N, S, beta / (3.0 * N), u, cE, lossd, c->stream);
        });
        weight_vjp(c, precision, u, cE, gd, 0);
        double l[4];
        copy_out(c, l, lossd, sizeof l);
        loss[1] = l[1];
        loss[2] = l[2];
        loss[0] = l[1] + l[2];
        copy_out(c, grad, gd, sizeof(double) * total);
        collect_times(c);
    });
}

nnpx_status nnpx_weight_vjp(nnpx_ctx* c, int precision, const double* vE, const double* vF, double* grad)
{
    return guard(c, [&] {
        gradient_checks(c, precision);
        if (!vE || !vF || !grad) fail(NNPX_ERR_ARGUMENT, "null argument");
        evaluate(c, precision, NNPX_EVAL_FORCES, true, precision == NNPX_PLAIN_FP64);
        settle(c);
        const int N = c->N, S = c->S;
        const int total = total_connections(c->hm.t);
        double* u = buf<double>(c, "t_u", (size_t)3 * N);
        double* cE = buf<double>(c, "t_cE", S);
        double* gd = buf<double>(c, "t_gradS", (size_t)S * total);
        hip_check(hipMemcpyAsync(cE, vE, sizeof(double) * S, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(u, vF, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
        hip_check(hipMemsetAsync(gd, 0, sizeof(double) * (size_t)S * total, c->stream), "hipMemsetAsync");
        weight_vjp(c, precision, u, cE, gd, total);
        copy_out(c, grad, gd, sizeof(double) * (size_t)S * total);
        collect_times(c);
    });
}

nnpx_status nnpx_kalman_init(nnpx_ctx* c, int type, int per_element, const double* par)
{
    return guard(c, [&] {
        need_model(c);
        if (type != 0 && type != 1) fail(NNPX_ERR_ARGUMENT, "kalman type must be 0 (standard) or 1 (fading memory)");
        if (!par) fail(NNPX_ERR_ARGUMENT, "null parameters");
        if (!(par[0] > 0.0)) fail(NNPX_ERR_ARGUMENT, "kalman epsilon must be > 0");
        const ModelT& t = c->hm.t;
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->kf.clear();
        int off = 0;
        for (int e = 0; e < t.nelem; ++e) {
            const int n = num_connections(t.el[e].nn);
            if (per_element || e == 0) {
                nnpx_ctx::KF k;
                k.off = off;
                c->kf.push_back(k);
            }
            c->kf.back().n += n;
            off += n;
        }
        for (size_t u = 0; u < c->kf.size(); ++u) {
            nnpx_ctx::KF& k = c->kf[u];
            // KalmanFilter::setParametersStandard / setParametersFadingMemory
            k.q = par[1];
            k.qtau = par[2];
            k.qmin = par[3];
            if (type == 0) {
                k.eta = par[4];
                k.etatau = par[5];
                k.etamax = par[6];
            } else {
                k.lambda = par[4];
                k.nu = par[5];
                k.gamma = 1.0;
            }
            double* P = buf<double>(c, "kf_P" + std::to_string(u), (size_t)k.n * k.n);
            kf_init(P, k.n, 1.0 / par[0], c->stream);  // P = I / epsilon
        }
        hip_check(hipGetLastError(), "kf_init");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->kf_type = type;
    });
}

nnpx_status nnpx_kalman_state(nnpx_ctx* c, int updater, double* P, double* scalars)
{
    return guard(c, [&] {
        if (c->kf_type < 0) fail(NNPX_ERR_STATE, "Kalman filter not set up (nnpx_kalman_init)");
        if (updater < 0 || updater >= (int)c->kf.size()) fail(NNPX_ERR_ARGUMENT, "bad updater index");
        const nnpx_ctx::KF& k = c->kf[updater];
        if (P) copy_out(c, P, existing<double>(c, "kf_P" + std::to_string(updater)), sizeof(double) * k.n * k.n);
        if (scalars) {
            scalars[0] = k.n;
            scalars[1] = k.off;
            scalars[2] = k.eta;
            scalars[3] = k.lambda;
            scalars[4] = k.gamma;
            scalars[5] = k.q;
            scalars[6] = (double)k.updates;
        }
    });
}

nnpx_status nnpx_kalman_update(nnpx_ctx* c, int precision, int property, int batch, const int* structure,
                               const int* atom, const int* component, const double* E_ref, const double* F_ref,
                               double force_weight, int jacobian_full, double* error)
{
    return guard(c, [&] {
        gradient_checks(c, precision);
        if (c->kf_type < 0) fail(NNPX_ERR_STATE, "Kalman filter not set up (nnpx_kalman_init)");
        if (property != 0 && property != 1) fail(NNPX_ERR_ARGUMENT, "property must be 0 (energy) or 1 (force)");
        if (batch < 1 || !structure || (property == 0 && !E_ref) || (property == 1 && (!F_ref || !atom || !component)))
            fail(NNPX_ERR_ARGUMENT, "bad update candidates");
        const int N = c->N, S = c->S;
        for (int b = 0; b < batch; ++b) {
            if (structure[b] < 0 || structure[b] >= S) fail(NNPX_ERR_RANGE, "candidate structure out of range");
            if (property == 1) {
                if (atom[b] < c->atom_off[structure[b]] || atom[b] >= c->atom_off[structure[b] + 1])
                    fail(NNPX_ERR_RANGE, "candidate atom not in its structure");
                if (component[b] < 0 || component[b] > 2) fail(NNPX_ERR_RANGE, "candidate component out of range");
            }
        }
        const ModelT& t = c->hm.t;
        const int total = total_connections(t);
        const int m = jacobian_full ? batch : 1;
        const bool force = property == 1;
        const double fw = force ? force_weight : 1.0;
        // Errors of the current weights (Training::update PART 2)
        evaluate(c, precision, NNPX_EVAL_FORCES, true, precision == NNPX_PLAIN_FP64);
        settle(c);
        std::vector<double> E(S), F((size_t)3 * N);
        copy_out(c, E.data(), existing<double>(c, "E"), sizeof(double) * S);
        if (force) copy_out(c, F.data(), existing<double>(c, "F"), sizeof(double) * 3 * N);
        std::vector<double> err(batch);
        for (int b = 0; b < batch; ++b)
            err[b] = force ? F_ref[3 * atom[b] + component[b]] - F[3 * (size_t)atom[b] + component[b]]
                           : E_ref[structure[b]] - E[structure[b]];
        if (error)
            for (int b = 0; b < batch; ++b) error[b] = err[b];
        // Per-element error fractions of US_ELEMENT: energy N_e / N of the
        // structure, force the atom's neighbours of element e over all of
        // them (Training.cpp:2881-2906)
        const int nu = (int)c->kf.size();
        std::vector<double> frac((size_t)batch * nu, 1.0);
        if (nu > 1) {
            std::vector<int> cnt, jj;
            if (force) {
                cnt.resize(N);
                copy_out(c, cnt.data(), existing<int>(c, "nb_n"), sizeof(int) * N);
                jj.resize((size_t)N * c->cap);
                copy_out(c, jj.data(), existing<int>(c, "nb_j"), sizeof(int) * jj.size());
            }
            for (int b = 0; b < batch; ++b) {
                std::vector<double> ne(t.nelem, 0.0);
                double all = 0.0;
                if (force) {
                    const int a = atom[b];
                    for (int k = 0; k < cnt[a]; ++k) ne[c->elem_host[jj[(size_t)a * c->cap + k]]] += 1.0;
                    all = cnt[a];
                } else {
                    for (int64_t i = c->atom_off[structure[b]]; i < c->atom_off[structure[b] + 1]; ++i)
                        ne[c->elem_host[i]] += 1.0;
                    all = (double)(c->atom_off[structure[b] + 1] - c->atom_off[structure[b]]);
                }
                for (int u = 0; u < nu; ++u) frac[(size_t)b * nu + u] = all > 0.0 ? ne[u] / all : 0.0;
            }
        }
        // Jacobian rows (Training::dPdc contracted with unit seeds): one
        // summed row, or one row per candidate in rounds of candidates on
        // different structures
        std::vector<double> rows((size_t)m * total, 0.0);
        double* u = buf<double>(c, "t_u", (size_t)3 * N);
        double* cE = buf<double>(c, "t_cE", S);
        double* gd = buf<double>(c, "t_gradS", (size_t)S * total);
        std::vector<int> done(batch, 0);
        int left = batch;
        while (left > 0) {
            std::vector<double> vE(S, 0.0), vF((size_t)3 * N, 0.0);
            std::vector<int> take, used(S, 0);
            for (int b = 0; b < batch; ++b) {
                if (done[b] || (jacobian_full && used[structure[b]])) continue;
                used[structure[b]] = 1;
                take.push_back(b);
                if (force) vF[3 * (size_t)atom[b] + component[b]] += 1.0;
                else vE[structure[b]] += 1.0;
            }
            hip_check(hipMemcpyAsync(cE, vE.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync");
            hip_check(hipMemcpyAsync(u, vF.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync");
            const int ld = jacobian_full ? total : 0;
            hip_check(hipMemsetAsync(gd, 0, sizeof(double) * (jacobian_full ? (size_t)S * total : total), c->stream),
                      "hipMemsetAsync");
            weight_vjp(c, precision, u, cE, gd, ld);
            if (jacobian_full) {
                std::vector<double> r((size_t)S * total);
                copy_out(c, r.data(), gd, sizeof(double) * r.size());
                for (int b : take)
                    std::copy(r.begin() + (size_t)structure[b] * total, r.begin() + (size_t)(structure[b] + 1) * total,
                              rows.begin() + (size_t)b * total);
            } else {
                copy_out(c, rows.data(), gd, sizeof(double) * total);
            }
            for (int b : take) done[b] = 1;
            left -= (int)take.size();
        }
        // The filters (Training::update PART 4, KalmanFilter::update)
        std::vector<double> w(total);
        get_connections(c, w.data());
        for (int ui = 0; ui < nu; ++ui) {
            nnpx_ctx::KF& k = c->kf[ui];
            const int n = k.n;
            std::vector<double> H((size_t)n * m), xi(m, 0.0);
            for (int j = 0; j < n; ++j)
                for (int col = 0; col < m; ++col) H[(size_t)j * m + col] = rows[(size_t)col * total + k.off + j] * fw;
            for (int b = 0; b < batch; ++b) xi[jacobian_full ? b : 0] += err[b] * fw * frac[(size_t)b * nu + ui];
            double* dH = buf<double>(c, "kf_H", (size_t)n * m);
            double* dX = buf<double>(c, "kf_X", (size_t)n * m);
            double* dK = buf<double>(c, "kf_K", (size_t)n * m);
            double* dA = buf<double>(c, "kf_A", (size_t)m * m);
            double* dxi = buf<double>(c, "kf_xi", m);
            double* ddw = buf<double>(c, "kf_dw", n);
            double* P = existing<double>(c, "kf_P" + std::to_string(ui));
            hip_check(hipMemcpyAsync(dH, H.data(), sizeof(double) * H.size(), hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync");
            hip_check(hipMemcpyAsync(dxi, xi.data(), sizeof(double) * m, hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync");
            timed(c, "kalman", [&] { kf_xa(P, dH, n, m, dX, dA, c->stream); });
            std::vector<double> A((size_t)m * m), Ai((size_t)m * m, 0.0);
            copy_out(c, A.data(), dA, sizeof(double) * A.size());
            if (c->kf_type == 0 && k.eta < k.etamax) k.eta *= std::exp(k.etatau);
            for (int a = 0; a < m; ++a) A[(size_t)a * m + a] += c->kf_type == 0 ? 1.0 / k.eta : k.lambda;
            // A^-1 (Gauss-Jordan, partial pivoting)
            for (int a = 0; a < m; ++a) Ai[(size_t)a * m + a] = 1.0;
            for (int col = 0; col < m; ++col) {
                int piv = col;
                for (int r = col + 1; r < m; ++r)
                    if (std::fabs(A[(size_t)r * m + col]) > std::fabs(A[(size_t)piv * m + col])) piv = r;
                if (A[(size_t)piv * m + col] == 0.0) fail(NNPX_ERR_RANGE, "Kalman filter: singular innovation matrix");
                if (piv != col)
                    for (int q = 0; q < m; ++q) {
                        std::swap(A[(size_t)piv * m + q], A[(size_t)col * m + q]);
                        std::swap(Ai[(size_t)piv * m + q], Ai[(size_t)col * m + q]);
                    }
                const double d = A[(size_t)col * m + col];
                for (int q = 0; q < m; ++q) {
                    A[(size_t)col * m + q] /= d;
                    Ai[(size_t)col * m + q] /= d;
                }
                for (int r = 0; r < m; ++r)
                    if (r != col) {
                        const double f = A[(size_t)r * m + col];
                        for (int q = 0; q < m; ++q) {
                            A[(size_t)r * m + q] -= f * A[(size_t)col * m + q];
                            Ai[(size_t)r * m + q] -= f * Ai[(size_t)col * m + q];
                        }
                    }
            }
            double* dAi = buf<double>(c, "kf_Ai", (size_t)m * m);
            hip_check(hipMemcpyAsync(dAi, Ai.data(), sizeof(double) * Ai.size(), hipMemcpyHostToDevice, c->stream),
                      "hipMemcpyAsync");
            const double scale = c->kf_type == 1 ? 1.0 / k.lambda : 1.0;
            timed(c, "kalman", [&] { kf_finish(P, dX, dAi, dxi, n, m, scale, k.q, dK, ddw, c->stream); });
            std::vector<double> dw(n);
            copy_out(c, dw.data(), ddw, sizeof(double) * n);
            for (int j = 0; j < n; ++j) w[k.off + j] += dw[j];
            if (k.q > k.qmin) k.q *= std::exp(-k.qtau);
            if (c->kf_type == 1) {
                k.lambda = k.nu * k.lambda + 1.0 - k.nu;
                k.gamma = 1.0 / (1.0 + k.lambda / k.gamma);
            }
            k.updates++;
        }
        set_connections(c, w.data());
        collect_times(c);
    });
}

namespace {

// Mean event time (ms) of `name` over the launches since the last reset.
double mean_time(nnpx_ctx* c, const char* name)
{
    collect_times(c);
    auto it = c->ktime.find(name);
    return it == c->ktime.end() || it->second.second == 0 ? 1e30 : it->second.first / it->second.second;
}

}  // namespace

nnpx_status nnpx_autotune(nnpx_ctx* c)
{
    return guard(c, [&] {
        need_structures(c);
        if (c->md_ready) fail(NNPX_ERR_STATE, "autotune before nnpx_md_init");
        const bool timing = c->timing;
        Tuning& tu = c->mixed.tune;
        const Tuning keep = tu;
        const int capL = c->cap;
        const ModelT& t = c->hm.t;
        auto fits = [&](int trip, size_t extra) {
            return stage_bytes(capL, trip, stage_mp(0, 0)) + extra <= 65536;
        };
        c->timing = true;
        double best_t = 1e30;
        Tuning best = keep;
        const int ldG = t.max_sf;
        double* u = buf<double>(c, "t_u", (size_t)3 * c->N);
        double* Gdot = buf<double>(c, "t_Gdot", (size_t)c->N * ldG);
        hip_check(hipMemsetAsync(u, 0, sizeof(double) * 3 * c->N, c->stream), "hipMemsetAsync");
        try {
            // triplet buffers per kernel, with and without the pair lists
            // (the force kernel walking blocks and reading the SF kernel's
            // pairs, or walking rows): the mode of the faster pair of kernels
            double mode_time[2] = {1e30, 1e30};
            Tuning mode_best[2] = {keep, keep};
            for (int pl = 0; pl < 2; ++pl) {
                double bsf = 1e30, bf = 1e30;
                Tuning& mb = mode_best[pl];
                mb.pair_lists = pl;
                for (int trip : {128, 256, 384, 512, 768, 1024}) {
                    if (!fits(trip, (sizeof(double) * 3 + 2) * (size_t)capL + kMaxElem * kMaxElem +
                                        (size_t)kMaxMerge * 20 * 8))
                        continue;
                    tu.trip_sf = tu.trip_force = tu.trip_tangent = trip;
                    tu.pair_lists = pl;
                    // warm-up; its settle() sizes the lists, the second run fills them
                    for (int r = 0; r < 2; ++r) {
                        evaluate(c, NNPX_MIXED, NNPX_EVAL_FORCES, true);
                        settle(c);
                    }
                    c->ktime.clear();
                    for (int r = 0; r < 3; ++r) evaluate(c, NNPX_MIXED, NNPX_EVAL_FORCES, false);
                    const double a = mean_time(c, "sf32"), b = mean_time(c, "force32");
                    if (a < bsf) { bsf = a; mb.trip_sf = trip; }
                    if (b < bf) { bf = b; mb.trip_force = trip; }
                    if (pl == 0) {
                        SysView s = sysview(c);
                        for (int r = 0; r < 3; ++r)
                            timed(c, "tangent32",
                                  [&] { launch_tangent32(c->mixed, t, c->d_model, s, u, Gdot, ldG, c->stream); });
                        const double d = mean_time(c, "tangent32");
                        if (d < best_t) { best_t = d; best.trip_tangent = trip; }
                    }
                }
                mode_time[pl] = bsf + bf;
            }
            const int pl = mode_time[1] < mode_time[0] ? 1 : 0;
            best.trip_sf = mode_best[pl].trip_sf;
            best.trip_force = mode_best[pl].trip_force;
            best.pair_lists = pl;
            tu = best;
            // merged angular walks (models that have them), per kernel
            if (c->mixed.merge_ng > 0) {
                double t_sf[2], t_f[2];
                for (int mg = 0; mg < 2; ++mg) {
                    tu.merge_sf = tu.merge_force = mg;
                    evaluate(c, NNPX_MIXED, NNPX_EVAL_FORCES, true);
                    settle(c);
                    c->ktime.clear();
                    for (int r = 0; r < 3; ++r) evaluate(c, NNPX_MIXED, NNPX_EVAL_FORCES, false);
                    t_sf[mg] = mean_time(c, "sf32");
                    t_f[mg] = mean_time(c, "force32");
                }
                // one choice for both kernels: the force kernel reads the SF
                // kernel's pair lists only when both make the same walks
                tu.merge_sf = tu.merge_force = t_sf[1] + t_f[1] < t_sf[0] + t_f[0] ? 1 : 0;
            }
            // atoms per block of the gradient kernel (FP32 network pass on the
            // current G with a zero tangent)
            GradArgs a{};
            int total = 0;
            for (int e = 0; e < t.nelem; ++e) {
                a.coff[e] = total;
                a.nconn[e] = num_connections(t.el[e].nn);
                a.nconn_max = std::max(a.nconn_max, a.nconn[e]);
                total += a.nconn[e];
            }
            double* cE = buf<double>(c, "t_cE", c->S);
            double* gd = buf<double>(c, "t_grad", total);
            hip_check(hipMemsetAsync(cE, 0, sizeof(double) * c->S, c->stream), "hipMemsetAsync");
            hip_check(hipMemsetAsync(Gdot, 0, sizeof(double) * (size_t)c->N * ldG, c->stream), "hipMemsetAsync");
            a.w = c->d_w;
            a.nweights = t.nweights;
            a.G = existing<double>(c, "G");
            a.Gdot = Gdot;
            a.ldG = ldG;
            a.cE = cE;
            a.sid = existing<int>(c, "sid");
            a.atoms = existing<int>(c, "elem_atoms");
            a.nelem = t.nelem;
            a.maxw = std::max(t.max_width, t.max_sf);
            a.grad = gd;
            double best_g = 1e30;
            if (grad_lds_bytes(a, true) <= 65536) {
                for (int per : {8, 16, 32, 64, 128}) {
                    a.per_block = per;
                    int blocks = 0;
                    for (int e = 0; e < t.nelem; ++e) {
                        a.blk_off[e] = blocks;
                        a.atom_off[e] = c->elem_off[e];
                        blocks += (c->elem_count[e] + per - 1) / per;
                    }
                    a.blk_off[t.nelem] = blocks;
                    a.atom_off[t.nelem] = c->elem_off[t.nelem];
                    c->ktime.clear();
                    launch_grad(c->d_model, a, blocks, true, c->stream);
                    for (int r = 0; r < 3; ++r) timed(c, "grad32", [&] { launch_grad(c->d_model, a, blocks, true, c->stream); });
                    const double g = mean_time(c, "grad32");
                    if (g < best_g) { best_g = g; tu.grad_atoms = per; }
                }
            }
        } catch (...) {
            tu = keep;
            c->timing = timing;
            throw;
        }
        c->ktime.clear();
        c->timing = timing;
        invalidate(c);
    });
}

nnpx_status nnpx_get_tuning(const nnpx_ctx* cc, int* v)
{
    nnpx_ctx* c = const_cast<nnpx_ctx*>(cc);
    return guard(c, [&] {
        if (!v) fail(NNPX_ERR_ARGUMENT, "null output");
        const Tuning& tu = c->mixed.tune;
        v[0] = tu.trip_sf;
        v[1] = tu.trip_force;
        v[2] = tu.trip_tangent;
        v[3] = tu.grad_atoms;
    });
}

nnpx_status nnpx_set_tuning(nnpx_ctx* c, const int* v)
{
    return guard(c, [&] {
        if (!v) fail(NNPX_ERR_ARGUMENT, "null input");
        for (int k = 0; k < 3; ++k)
            if (v[k] < 64 || v[k] > 4096 || v[k] % 64) fail(NNPX_ERR_ARGUMENT, "triplet buffer: multiple of 64 in [64, 4096]");
        if (v[3] < 1 || v[3] > 1024) fail(NNPX_ERR_ARGUMENT, "gradient atoms per block out of range");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        drop_graph(c);
        Tuning& tu = c->mixed.tune;
        tu.trip_sf = v[0];
        tu.trip_force = v[1];
        tu.trip_tangent = v[2];
        tu.grad_atoms = v[3];
    });
}

namespace {

// One velocity-Verlet step, enqueued without host synchronisation: half kick
// and drift, displacement check, gated wrap and list rebuild, forces, second
// half kick with the kinetic energy, step bookkeeping.
void md_step(nnpx_ctx* c)
{
    double* pos = existing<double>(c, "pos");
    double* vel = existing<double>(c, "md_vel");
    double* ref = existing<double>(c, "md_ref");
    int* gate = existing<int>(c, "md_gate");
    int* counters = existing<int>(c, "md_counters");
    const double* F = existing<double>(c, "F");
    const int* elem = existing<int>(c, "elem");
    const double half = 0.5 * c->skin;
    md_kick_drift(pos, vel, F, elem, c->mdc, c->N, c->stream);
    md_check(pos, ref, c->N, half * half, gate, c->stream);
    md_wrap(pos, ref, existing<int>(c, "sid"), existing<StructT>(c, "st"), c->N, gate, c->stream);
    SysView s = sysview(c);
    s.gate = gate;
    launch_neighbors(s, list_cutoff(c), c->any_cell, c->stream, false);
    if (c->md_precision == NNPX_MIXED) {
        MixedBuffers mb = mixed_buffers(c);
        run_mixed(c->mixed, c->hm.t, c->d_model, sysview(c), mb, true, c->force_mode == NNPX_FORCES_PUSH, c->stream,
                  [&](const char* name, const std::function<void()>& f) { timed(c, name, f); });
    } else if (fused_fits(c)) {
        compute_sf64(c, false);
        compute_nn64(c, true);
        compute_energy(c);
        compute_force64_fused(c);
    } else {
        launch_reverse(s, c->stream);
        c->rev_valid = true;
        compute_sf64(c, true);
        compute_nn64(c, true);
        compute_energy(c);
        compute_force64(c);
    }
    md_kick(vel, existing<double>(c, "F"), elem, c->mdc, c->N, counters, existing<double>(c, "md_ekin"), c->stream);
    md_step_end(gate, counters, existing<double>(c, "E"), existing<double>(c, "md_epot"),
                existing<double>(c, "md_ekin"), c->stream);
    hip_check(hipGetLastError(), "MD step launch");
}

void need_md(nnpx_ctx* c)
{
    need_structures(c);
    if (!c->md_ready) fail(NNPX_ERR_STATE, "MD not initialised (nnpx_md_init)");
}

}  // namespace

nnpx_status nnpx_md_init(nnpx_ctx* c, int precision, const double* masses, const double* vel, double dt,
                         double skin)
{
    return guard(c, [&] {
        need_structures(c);
        if (c->S != 1) fail(NNPX_ERR_UNSUPPORTED, "MD runs one structure per context");
        if (c->external) fail(NNPX_ERR_STATE, "MD needs positions (not an external neighbour list)");
        if (precision != NNPX_PLAIN_FP64 && precision != NNPX_MIXED) fail(NNPX_ERR_ARGUMENT, "bad precision");
        if (!masses || !vel) fail(NNPX_ERR_ARGUMENT, "null masses or velocities");
        if (!(dt > 0.0) || !(skin >= 0.0)) fail(NNPX_ERR_ARGUMENT, "dt must be > 0 and skin >= 0");
        const int ne = c->hm.t.nelem;
        for (int e = 0; e < ne; ++e)
            if (!(masses[e] > 0.0)) fail(NNPX_ERR_ARGUMENT, "masses must be > 0");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        drop_graph(c);
        c->md_ready = false;
        // Skin-extended list: image ranges and cell grid for rc + skin.
        const double cl = c->hm.normalize ? c->hm.convLength : 1.0;
        const double ce = c->hm.normalize ? c->hm.convEnergy : 1.0;
        c->skin = skin * cl;
        std::vector<double> pos((size_t)3 * c->N);
        copy_out(c, pos.data(), existing<double>(c, "pos"), sizeof(double) * pos.size());
        StructT& t = c->st_host[0];
        const int a0 = t.atom0, na = t.natoms;
        double box[9];
        std::memcpy(box, t.box, sizeof box);
        setup_struct(t, box, t.periodic != 0, list_cutoff(c));
        t.atom0 = a0;
        t.natoms = na;
        choose_algorithms(c, pos.data());
        hip_check(hipMemcpy(existing<StructT>(c, "st"), c->st_host.data(), sizeof(StructT), hipMemcpyHostToDevice),
                  "hipMemcpy");
        c->cap = estimate_cap(c);
        invalidate(c);
        build_neighbors(c, true);
        // headroom for density fluctuations between rebuilds
        c->cap = std::max(32, ((c->max_nb + c->max_nb / 4 + 32 + 31) / 32) * 32);
        build_neighbors(c, true);
        // Velocity-Verlet constants: a = F/m with F in eV/A, m in amu, t in
        // fs gives 9.64853321e-3 A/fs^2 per eV/(A amu); converted to the
        // model's length and energy units.
        const double kappa = 9.64853321e-3;
        for (int e = 0; e < ne; ++e) {
            c->mdc.hk[e] = 0.5 * dt * kappa * cl * cl / (ce * masses[e]);
            c->mdc.half_m[e] = 0.5 * masses[e] * ce / (kappa * cl * cl);
        }
        c->mdc.dt = dt;
        std::vector<double> v((size_t)3 * c->N);
        for (size_t q = 0; q < v.size(); ++q) v[q] = vel[q] * cl;
        double* dv = buf<double>(c, "md_vel", v.size());
        hip_check(hipMemcpy(dv, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
        double* ref = buf<double>(c, "md_ref", (size_t)3 * c->N);
        hip_check(hipMemcpy(ref, existing<double>(c, "pos"), sizeof(double) * 3 * c->N, hipMemcpyDeviceToDevice),
                  "hipMemcpy");
        int* gate = buf<int>(c, "md_gate", 16);
        int* counters = buf<int>(c, "md_counters", 16);
        buf<double>(c, "md_epot", kMDHistory);
        double* ekin = buf<double>(c, "md_ekin", kMDHistory);
        c->md_precision = precision;
        c->md_stage_cap = 0;
        // forces of the initial positions (allocates every buffer a step uses)
        evaluate(c, precision, NNPX_EVAL_FORCES, true);
        settle(c);
        if (c->md_stage_req != 0) {
            // staging for the most in-cutoff entries met plus an eighth (at
            // least 16), or as requested: a run that meets more repeats at
            // the list capacity
            int f[3];
            copy_out(c, f, existing<int>(c, "flags"), sizeof f);
            const int m = c->md_stage_req > 0 ? c->md_stage_req : f[2] > 0 ? f[2] + std::max(16, f[2] / 8) : 0;
            const int fitted = ((m + 31) / 32) * 32;
            c->md_stage_cap = m > 0 && fitted < c->cap ? fitted : 0;
        }
        // the pair lists at the size settle() fitted, before any step is captured
        if (precision == NNPX_MIXED) (void)mixed_buffers(c);
        md_reset(gate, counters, ekin, c->stream);
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->md_ready = true;
    });
}

namespace {

// State a run of MD steps changes, copied before a run with fitted staging
// so a run that overflowed it can be repeated at the list capacity.
const char* const kMDSaved[] = {"pos", "md_vel", "md_ref", "F", "md_epot", "md_ekin"};

void md_backup(nnpx_ctx* c, bool restore)
{
    for (const char* name : kMDSaved) {
        const DevBuf& b = c->bufs.at(name);
        const std::string bak = std::string("bak_") + name;
        void* d = restore ? b.p : buf<char>(c, bak, b.bytes);
        const void* src = restore ? existing<char>(c, bak) : b.p;
        hip_check(hipMemcpyAsync(d, src, b.bytes, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    }
    int* cnt = existing<int>(c, "md_counters");
    int* bak = restore ? cnt : buf<int>(c, "bak_md_counters", 16);
    hip_check(hipMemcpyAsync(bak, restore ? existing<int>(c, "bak_md_counters") : cnt, 16 * sizeof(int),
                             hipMemcpyDeviceToDevice, c->stream),
              "hipMemcpyAsync");
}

void md_steps(nnpx_ctx* c, int nsteps, int graph_steps);

}  // namespace

nnpx_status nnpx_md_run(nnpx_ctx* c, int nsteps, int graph_steps)
{
    return guard(c, [&] {
        need_md(c);
        if (nsteps < 0 || graph_steps < 0) fail(NNPX_ERR_ARGUMENT, "negative step count");
        const bool fitted = c->md_stage_cap > 0;
        int* flags = existing<int>(c, "flags");
        if (fitted) {
            int f[3];
            copy_out(c, f, flags, sizeof f);
            if (f[0] & 4) {
                f[0] &= ~4;
                hip_check(hipMemcpy(flags, f, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
            }
            md_backup(c, false);
        }
        md_steps(c, nsteps, graph_steps);
        if (fitted) {
            int f[3];
            copy_out(c, f, flags, sizeof f);
            if (f[0] & 4) {
                // an atom met more in-cutoff entries than the fitted staging:
                // back to the state before the run, a fresh list, and the run
                // again with the whole list capacity staged
                md_backup(c, true);
                int* gate = existing<int>(c, "md_gate");
                md_set_gate(gate, 1, c->stream);
                md_wrap(existing<double>(c, "pos"), existing<double>(c, "md_ref"), existing<int>(c, "sid"),
                        existing<StructT>(c, "st"), c->N, gate, c->stream);
                SysView s = sysview(c);
                s.gate = gate;
                launch_neighbors(s, list_cutoff(c), c->any_cell, c->stream, false);
                md_set_gate(gate, 0, c->stream);
                f[0] &= ~4;
                hip_check(hipMemcpyAsync(flags, f, sizeof(int), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync");
                hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
                c->md_stage_cap = 0;
                drop_graph(c);
                md_steps(c, nsteps, graph_steps);
            }
        }
    });
}

namespace {

void md_steps(nnpx_ctx* c, int nsteps, int graph_steps)
{
    {
        const bool timing = c->timing;
        if (graph_steps > 0 && (!c->gexec || c->graph_steps != graph_steps)) {
            drop_graph(c);
            c->timing = false;  // no event records inside the captured sequence
            hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                for (int k = 0; k < graph_steps; ++k) md_step(c);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                c->timing = timing;
                throw;
            }
            c->timing = timing;
            hip_check(hipStreamEndCapture(c->stream, &c->graph), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            c->graph_steps = graph_steps;
        }
        c->timing = false;
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = get_event(c);
            b = get_event(c);
            hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
        }
        int done = 0;
        if (graph_steps > 0)
            for (; done + graph_steps <= nsteps; done += graph_steps)
                hip_check(hipGraphLaunch(c->gexec, c->stream), "hipGraphLaunch");
        for (; done < nsteps; ++done) md_step(c);
        if (timing) {
            hip_check(hipEventRecord(b, c->stream), "hipEventRecord");
            c->pending.push_back(Timer{"md", a, b});
        }
        c->timing = timing;
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        int f[4];
        read_flags(c, f);
        if (f[0] & 1) fail(NNPX_ERR_RANGE, "neighbour capacity exceeded during MD (raise the skin or re-initialise)");
        c->sf_valid = c->nn_valid = c->dedg_valid = c->e_valid = c->f_valid = true;
        c->sfd_valid = c->md_precision == NNPX_PLAIN_FP64 && !fused_fits(c);
        c->last_stored = c->sfd_valid;
        c->last_precision = c->md_precision;
    }
}

}  // namespace

nnpx_status nnpx_md_state(nnpx_ctx* c, double* pos, double* vel)
{
    return guard(c, [&] {
        need_md(c);
        if (pos) copy_out(c, pos, existing<double>(c, "pos"), sizeof(double) * 3 * c->N);
        if (vel) {
            copy_out(c, vel, existing<double>(c, "md_vel"), sizeof(double) * 3 * c->N);
            const double cl = c->hm.normalize ? c->hm.convLength : 1.0;
            for (int64_t q = 0; q < 3 * (int64_t)c->N; ++q) vel[q] /= cl;
        }
    });
}

nnpx_status nnpx_md_energies(nnpx_ctx* c, double* epot, double* ekin, int64_t max, int64_t* nsteps,
                             int64_t* rebuilds)
{
    return guard(c, [&] {
        need_md(c);
        int cnt[2];
        copy_out(c, cnt, existing<int>(c, "md_counters"), sizeof cnt);
        if (nsteps) *nsteps = cnt[0];
        if (rebuilds) *rebuilds = cnt[1];
        const int64_t n = std::min<int64_t>(std::min<int64_t>(cnt[0], kMDHistory), std::max<int64_t>(max, 0));
        if (n > 0 && cnt[0] > kMDHistory) fail(NNPX_ERR_RANGE, "more steps than the kept energy history");
        if (epot && n > 0) copy_out(c, epot, existing<double>(c, "md_epot"), sizeof(double) * n);
        if (ekin && n > 0) copy_out(c, ekin, existing<double>(c, "md_ekin"), sizeof(double) * n);
    });
}

nnpx_status nnpx_predict_noalloc(nnpx_ctx* c, int precision, unsigned flags, int n_struct,
                                 const int64_t* atom_offsets, const double* pos, const int* elem,
                                 const double* box, const int* periodic, double* energy, double* forces)
{
    return guard(c, [&] {
        need_model(c);
        if (precision != NNPX_PLAIN_FP64 && precision != NNPX_MIXED) fail(NNPX_ERR_ARGUMENT, "bad precision");
        if (!energy || ((flags & NNPX_EVAL_FORCES) && !forces)) fail(NNPX_ERR_ARGUMENT, "null output");
        if (c->md_ready || c->gexec) fail(NNPX_ERR_STATE, "no-allocation call on a context set up for MD");
        if (!c->scratch) fail(NNPX_ERR_STATE, "scratch buffer address unknown");
        hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        // The call works in the scratch array; afterwards the context holds
        // no structures (its own buffers were not touched).
        const bool timing = c->timing;
        struct Scope {
            nnpx_ctx* c;
            bool timing;
            ~Scope()
            {
                (void)hipStreamSynchronize(c->stream);
                c->arena = false;
                c->abufs.clear();
                c->timing = timing;
                c->N = c->S = 0;
                c->external = false;
                c->nbr_valid = c->rev_valid = false;
                c->stage_cap = 0;
                invalidate(c);
            }
        } scope{c, timing};
        c->timing = false;  // no event objects are created
        c->arena = true;
        c->arena_used = 0;
        c->abufs.clear();
        set_structures(c, n_struct, atom_offsets, pos, elem, box, periodic);
        evaluate(c, precision, flags, false);
        settle(c);
        copy_out(c, energy, existing<double>(c, "E"), sizeof(double) * c->S);
        if (flags & NNPX_EVAL_FORCES) copy_out(c, forces, existing<double>(c, "F"), sizeof(double) * 3 * c->N);
    });
}

nnpx_status nnpx_mfma_selftest(nnpx_ctx* c, double* err)
{
    return guard(c, [&] {
        if (!err) fail(NNPX_ERR_ARGUMENT, "null output");
        *err = mfma_selftest();
    });
}

nnpx_status nnpx_kernel_time(nnpx_ctx* c, const char* name, double* ms, int64_t* launches)
{
    return guard(c, [&] {
        if (!name) fail(NNPX_ERR_ARGUMENT, "null name");
        collect_times(c);
        auto it = c->ktime.find(name);
        if (ms) *ms = it == c->ktime.end() ? 0.0 : it->second.first;
        if (launches) *launches = it == c->ktime.end() ? 0 : it->second.second;
    });
}

nnpx_status nnpx_reset_kernel_times(nnpx_ctx* c)
{
    return guard(c, [&] {
        collect_times(c);
        c->ktime.clear();
    });
}

nnpx_status nnpx_work_counts(nnpx_ctx* c, int64_t* pairs, int64_t* triplets, int64_t* pm, int64_t* tm,
                             int64_t* te)
{
    return guard(c, [&] {
        need_structures(c);
        if (!c->nbr_valid) build_neighbors(c, true);
        settle(c);
        unsigned long long* d = buf<unsigned long long>(c, "work", 5);
        hip_check(hipMemsetAsync(d, 0, 5 * sizeof(unsigned long long), c->stream), "hipMemsetAsync");
        launch_work_counts(c->d_model, sysview(c), d, c->stream);
        unsigned long long h[5];
        copy_out(c, h, d, sizeof h);
        if (pairs) *pairs = (int64_t)h[0];
        if (triplets) *triplets = (int64_t)h[1];
        if (pm) *pm = (int64_t)h[2];
        if (tm) *tm = (int64_t)h[3];
        if (te) *te = (int64_t)h[4];
    });
}

}  // extern "C"
