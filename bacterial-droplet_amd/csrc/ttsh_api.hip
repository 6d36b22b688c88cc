// ttsh_api.hip -- C ABI of libttsh3d: device memory, step sequencing, HIP
// graphs and batches of droplets.
//
// An Engine owns a set of droplet slots on one device and one stream.  One
// step of the engine is eight kernel kinds, each launched once per step (the
// FFT passes once per grid size present), each covering every active
// droplet through a work table (see ttsh_dev.h):
//   K_rhs, K_div (with k_proj_stats, the per-droplet sums behind it), FFT x
//   fwd, FFT y fwd, FFT z (fwd, -1/k^2, inv), FFT y inv, FFT x inv, K_corr
// standing in for the reference's EEuler (TTSH3D.cu:347-359: 107 kernel
// launches and 4 cuFFT transforms per step).
//
// ttsh_sim is an engine with one slot stepped with plain launches.  A
// ttsh_batch may capture K steps in a HIP graph and replay it; the kernels
// then run with fixed grids and read work counts from device memory, so
// admitting or retiring droplets rewrites tables without re-capturing
// (unless a table must grow or a grid size appears that the graph lacks).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ttsh3d.h"
#include "ttsh_dev.h"
#include "ttsh_internal.h"
#include "ttsh_kernels.h"

namespace ttsh {

thread_local LaunchEvents g_launch_ev;

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(TTSH_EHIP, std::string(#expr " failed: ") + hipGetErrorString(e_));           \
    } while (0)

// Kernel kinds of one step, in launch order.  K_fft_xy is the x pass, or the
// fused x+y pass for sizes N <= 128 of which the batch holds at least two
// z-planes per CU (those sizes then have no work in K_fft_y, K_fft_y_inv).
enum { KK_RHS = 0, KK_DIV, KK_FX, KK_FY, KK_FZ, KK_FYI, KK_FXI, KK_CORR, KK_N };
static const char* const kKindNames[KK_N] = {"K_rhs", "K_div", "K_fft_xy", "K_fft_y", "K_fft_z_poisson",
                                             "K_fft_y_inv", "K_fft_xy_inv", "K_corr"};

// Work tables: rhs, div/corr (shared), one item per droplet (the projection's
// sums), and per grid size the x/y-pass table, the z-pass table and the
// z-plane table of the fused x+y pass.
enum { T_RHS = 0, T_PROJ = 1, T_STAT = 2, T_CORR = 3, T_FFT0 = 4 };
constexpr int kNTables = T_FFT0 + 3 * kNumGridSizes;
static_assert(kNumGridSizes <= 32, "sizes are kept as bits of a 32-bit mask");
static inline int t_xy(int s) { return T_FFT0 + 3 * s; }      // line groups of x or y passes
static inline int t_z(int s) { return T_FFT0 + 3 * s + 1; }   // line groups of the z pass
static inline int t_pl(int s) { return T_FFT0 + 3 * s + 2; }  // z-planes of the fused x+y pass

static int size_index(int N)
{
    for (int s = 0; s < kNumGridSizes; ++s)
        if (kGridSizes[s] == N) return s;
    return -1;
}

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Upper bound of divergence work items of one droplet (z-depth >= 1).
static int proj_items_max(int N)
{
    return ((N + kProjTX - 1) / kProjTX) * ((N + kProjTY - 1) / kProjTY) * N;
}

size_t droplet_bytes(int N)
{
    const size_t n = (size_t)N * N * N;
    return 20 * align256(n * sizeof(float)) + align256(n * sizeof(float2)) +
           align256((size_t)proj_items_max(N) * kPartSums * sizeof(double)) + align256(2 * sizeof(double)) +
           align256(sizeof(float)) + align256(n * sizeof(float));
}

static int default_rhs_variant(int N) { return (N % 32 == 0 && N >= 64) ? RHS_32x8 : RHS_16x16; }

// A K_rhs tile needs N >= TX+2 and N >= TY+2; 16x16 fits every supported N.
static bool variant_fits(int v, int N)
{
    int tx, ty;
    rhs_tile(v, &tx, &ty);
    return N >= tx + 2 && N >= ty + 2;
}

struct Slot {
    bool used = false;
    ttsh_params prm{};
    int N = 0;
    void* mem = nullptr;
    size_t bytes = 0;
    Drop drop{};
    int q = 0;              // parity of the droplet's buffers: phi current = phi[q]
    int64_t steps = 0;      // steps since set_state / admission
    int64_t limit = -1;     // max steps, -1 = none
    double T = 0.0;         // T accumulated in double as the reference (TTSH3D.cu:342)
    int rhs_tz = 16, proj_tz = 8;
    bool rhs_tz_set = false;  // rhs_tz came from the parameters, not the tuning table
    int64_t fin_call = -1;    // the step call that took the droplet's last step (limit reached)
    bool published = false;   // in the device tables since its allocation
};

struct Table {
    Work* d = nullptr;
    int cap = 0;
    int n = 0;
};

struct Engine {
    int device = 0;
    hipStream_t st = nullptr;
    int max_slots = 1;
    size_t budget = 0, used = 0;
    bool use_graph = false;
    int graph_steps = 16;
    int rhs_variant = -1;  // engine-wide K_rhs tile variant (-1: from the first droplet)
    int rhs_active = -1;   // the variant the work tables and launches use (see rebuild)
    int cus = 256;

    std::vector<Slot> slots;
    std::map<int, float2*> tw;  // twiddle table per N

    Drop* d_drops = nullptr;
    std::vector<Drop> h_drops;
    Table tab[kNTables];
    int h_counts[kNTables] = {};
    int* d_counts = nullptr;

    double* d_redpart = nullptr;
    double* d_redout = nullptr;
    static constexpr int kRedBlocks = 256;
    // diagnostics: centroid partial sums and result, one slice buffer
    double* d_cpart = nullptr;
    double* d_csum = nullptr;
    float* d_slice = nullptr;
    int slice_N = 0;

    // graphs, one per starting parity; the sizes present at capture
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int ggrid[kNTables] = {};
    unsigned gsizes = 0;
    // sizes that run the fused x+y pass (current tables, and at capture)
    unsigned xy_mask = 0, gxy_mask = 0;
    int engine_parity = 0;
    int last_chunk = -1;
    bool tables_dirty = true;

    // Side stream for uploads, and for reads of droplets that reached their
    // step limit: it never waits for the step kernels on st except through
    // events.  Uploads are published to the step kernels at the next table
    // rebuild, which makes st wait for up_ev.
    hipStream_t us = nullptr;
    hipEvent_t up_ev = nullptr, st_ev = nullptr;
    bool up_pending = false;
    // One event per step call (ttsh_batch_step / ttsh_step), recorded on st
    // after the call's launches, in a ring.
    static constexpr int kCallEv = 16;
    hipEvent_t call_ev[kCallEv] = {};
    int64_t ncalls = 0;
    // Device memory of retired droplets, kept for reuse by droplets of the
    // same size; ev marks the end of the work queued before the retirement.
    struct Pooled {
        int N;
        void* mem;
        size_t bytes;
        hipEvent_t ev;
    };
    std::vector<Pooled> pool;
    size_t pooled = 0;
    Drop* d_rdrop = nullptr;  // descriptor of the droplet being reduced

    hipStream_t side() const { return us ? us : st; }

    // Work queued on the side stream so far (fills, uploads) is waited for by
    // st at the next table rebuild, or before any other use on st.
    hipError_t publish_side()
    {
        if (!us) return hipSuccess;
        up_pending = true;
        return hipEventRecord(up_ev, us);
    }
    hipError_t uploads_into_st()
    {
        return up_pending ? hipStreamWaitEvent(st, up_ev, 0) : hipSuccess;
    }

    // For a droplet that reached its step limit, the event after which no
    // step kernel touches it (nullptr: none, or not recorded yet).
    hipEvent_t fin_event(const Slot& s) const
    {
        if (!call_ev[0] || s.limit < 0 || s.steps < s.limit || s.fin_call < 0 || s.fin_call >= ncalls) return nullptr;
        // a later call's event in the ring also marks the end of this one
        const int64_t k = std::max<int64_t>(s.fin_call, ncalls - kCallEv);
        return call_ev[k % kCallEv];
    }

    // Wait until every step call but the last `keep` is done on the device.
    int wait_calls(int keep)
    {
        HIPCHK(hipSetDevice(device));
        if (keep <= 0 || !call_ev[0]) {
            HIPCHK(hipStreamSynchronize(st));
            HIPCHK(hipStreamSynchronize(side()));
            return TTSH_OK;
        }
        if (keep >= kCallEv) return fail(TTSH_EINVAL, "keep must be below " + std::to_string(kCallEv));
        const int64_t k = ncalls - 1 - keep;
        if (k >= 0) HIPCHK(hipEventSynchronize(call_ev[k % kCallEv]));
        return TTSH_OK;
    }

    // External arena (ttsh_step_noalloc): every device buffer is carved from
    // [arena, arena + arena_cap) instead of hipMalloc, nothing is freed, and
    // the stream is the caller's.
    bool external = false;
    char* arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;

    ~Engine() { release(); }

    hipError_t dev_alloc(void** p, size_t bytes)
    {
        if (!external) return hipMalloc(p, bytes);
        const size_t b = align256(bytes);
        if (arena_off + b > arena_cap) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        *p = arena + arena_off;
        arena_off += b;
        return hipSuccess;
    }
    template <typename T>
    hipError_t dev_alloc(T** p, size_t bytes)
    {
        void* v = nullptr;
        const hipError_t e = dev_alloc(&v, bytes);
        *p = static_cast<T*>(v);
        return e;
    }
    void dev_free(void* p)
    {
        if (!external && p) (void)hipFree(p);
    }

    void drop_graphs()
    {
        for (int p = 0; p < 2; ++p)
            if (gexec[p]) { (void)hipGraphExecDestroy(gexec[p]); gexec[p] = nullptr; }
        gsizes = 0;
        gxy_mask = 0;
    }

    void release()
    {
        if (st || external) (void)hipSetDevice(device);
        if (st || external) (void)hipStreamSynchronize(st);
        drop_graphs();
        for (auto& s : slots)
            if (s.mem) { dev_free(s.mem); s.mem = nullptr; }
        for (auto& kv : tw) dev_free(kv.second);
        tw.clear();
        dev_free(d_drops);
        for (auto& t : tab)
            if (t.d) { dev_free(t.d); t.d = nullptr; t.cap = 0; }
        dev_free(d_counts);
        dev_free(d_redpart);
        dev_free(d_redout);
        dev_free(d_cpart);
        dev_free(d_csum);
        dev_free(d_slice);
        dev_free(d_rdrop);
        d_rdrop = nullptr;
        if (us) (void)hipStreamSynchronize(us);
        for (auto& b : pool) {
            dev_free(b.mem);
            if (b.ev) (void)hipEventDestroy(b.ev);
        }
        pool.clear();
        pooled = 0;
        if (up_ev) (void)hipEventDestroy(up_ev);
        if (st_ev) (void)hipEventDestroy(st_ev);
        up_ev = st_ev = nullptr;
        for (auto& e : call_ev)
            if (e) { (void)hipEventDestroy(e); e = nullptr; }
        if (us) (void)hipStreamDestroy(us);
        us = nullptr;
        d_drops = nullptr; d_counts = nullptr; d_redpart = nullptr; d_redout = nullptr;
        d_cpart = nullptr; d_csum = nullptr; d_slice = nullptr; slice_N = 0;
        if (st && !external) (void)hipStreamDestroy(st);
        st = nullptr;
    }

    int init(int dev, int nslots, size_t bud, bool graph)
    {
        device = dev;
        max_slots = nslots;
        budget = bud;
        use_graph = graph;
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (dev < 0 || dev >= ndev) return fail(TTSH_EHIP, "device " + std::to_string(dev) + " not present");
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        if (!external) {
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HIPCHK(hipStreamCreateWithFlags(&us, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&up_ev, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&st_ev, hipEventDisableTiming));
            for (auto& e : call_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        slots.resize(nslots);
        HIPCHK(dev_alloc(&d_rdrop, sizeof(Drop)));
        HIPCHK(dev_alloc(&d_drops, sizeof(Drop) * nslots));
        h_drops.resize(nslots);
        HIPCHK(dev_alloc(&d_counts, sizeof(int) * kNTables));
        HIPCHK(dev_alloc(&d_redpart, sizeof(double) * 3 * kRedBlocks));
        HIPCHK(dev_alloc(&d_redout, sizeof(double) * 3));
        HIPCHK(dev_alloc(&d_cpart, sizeof(double) * 4 * kRedBlocks));
        HIPCHK(dev_alloc(&d_csum, sizeof(double) * 4));
        return TTSH_OK;
    }

    // An engine whose device memory is the arena [base, base + cap) and whose
    // stream is the caller's (ttsh_step_noalloc).
    int init_external(int dev, hipStream_t stream, void* base, size_t cap)
    {
        external = true;
        arena = static_cast<char*>(base);
        arena_cap = cap;
        arena_off = 0;
        st = stream;
        return init(dev, 1, 0, false);
    }

    // Diagnostic launches on the engine stream for slot si: the centroid sums
    // into sums (4 doubles), then (slice != null) plane z (< 0: the centroid
    // plane) of phi, v, P and the vorticities into slice (9*N*N+1 floats).
    int launch_diag(int si, double* sums, int z, float* slice)
    {
        Slot& sl = slots[si];
        HIPCHK(uploads_into_st());
        const float* phi = sl.drop.phi[sl.q];
        HIPCHK(launch_centroid(phi, sl.N, d_cpart, kRedBlocks, sums, st));
        if (slice) {
            const float* v[3] = {sl.drop.v[0], sl.drop.v[1], sl.drop.v[2]};
            const float* P[3] = {sl.drop.P[0], sl.drop.P[1], sl.drop.P[2]};
            HIPCHK(launch_slice(phi, v, P, sl.N, z, sums, (float)sl.prm.h, slice, st));
        }
        return TTSH_OK;
    }

    int centroid(int si, ttsh_centroid* out)
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        HIPCHK(hipSetDevice(device));
        int rc = launch_diag(si, d_csum, 0, nullptr);
        if (rc) return rc;
        double h[4];
        HIPCHK(hipMemcpyAsync(h, d_csum, sizeof(h), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        centroid_from_sums(h, out);
        return TTSH_OK;
    }

    static void centroid_from_sums(const double h[4], ttsh_centroid* out)
    {
        // sums are (phi, phi*y, phi*x, phi*z); the snapshot axes are (y, x, z)
        out->mass = h[0];
        for (int a = 0; a < 3; ++a) {
            out->c[a] = h[1 + a] / h[0];
            out->idx[a] = std::isfinite(out->c[a]) ? (int32_t)out->c[a] : 0;  // truncation, as Python int()
        }
    }

    int slice(int si, int z, float* out, int* z_used)
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        if (z >= s.N) return fail(TTSH_EINVAL, "plane outside the box");
        HIPCHK(hipSetDevice(device));
        if (slice_N != s.N) {
            dev_free(d_slice);
            d_slice = nullptr;
            HIPCHK(dev_alloc(&d_slice, sizeof(float) * (9 * (size_t)s.N * s.N + 1)));
            slice_N = s.N;
        }
        int rc = launch_diag(si, d_csum, z, d_slice);
        if (rc) return rc;
        const size_t nn = (size_t)s.N * s.N;
        HIPCHK(hipMemcpyAsync(out, d_slice, sizeof(float) * 9 * nn, hipMemcpyDeviceToHost, st));
        float zf = (float)z;
        if (z < 0) HIPCHK(hipMemcpyAsync(&zf, d_slice + 9 * nn, sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (z_used) *z_used = (int)zf;
        return TTSH_OK;
    }

    int twiddles(int N, const float2** out)
    {
        auto it = tw.find(N);
        if (it != tw.end()) { *out = it->second; return TTSH_OK; }
        std::vector<float2> h(N);
        const double twopi = 6.283185307179586476925286766559;
        for (int q = 0; q < N; ++q) {
            const double a = twopi * (double)q / (double)N;
            h[q] = make_float2((float)std::cos(a), (float)-std::sin(a));
        }
        float2* d = nullptr;
        HIPCHK(dev_alloc(&d, sizeof(float2) * N));
        HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
        tw[N] = d;
        *out = d;
        return TTSH_OK;
    }

    static void make_coef(const ttsh_params& p, Coef& co)
    {
        const int N = p.Nx;
        const double h = p.h;
        co.d1 = (float)(1.0 / (2.0 * h));
        co.sw0 = (float)(4.0 / 6.0);
        co.sw1 = (float)(1.0 / 6.0);
        co.e0 = (float)(1.0 / (2.0 * h) * (4.0 / 6.0));
        co.e1 = (float)(1.0 / (2.0 * h) * (1.0 / 6.0));
        co.f00 = (float)(1.0 / (2.0 * h) * (16.0 / 36.0));
        co.f01 = (float)(1.0 / (2.0 * h) * (4.0 / 36.0));
        co.f11 = (float)(1.0 / (2.0 * h) * (1.0 / 36.0));
        co.lcs = (float)(6.0 / (12.0 * h * h));
        co.lcf = (float)(-4.0 / (12.0 * h * h));
        co.l0s = (float)(6.0 * 2.0 / (3.0 * h * h));
        co.l0f = (float)(-14.0 / (3.0 * h * h) - 4.0 * 2.0 / (3.0 * h * h));
        co.lc = (float)(1.0 / (12.0 * h * h));
        co.lf = (float)(2.0 / (3.0 * h * h));
        co.l0 = (float)(-14.0 / (3.0 * h * h));
        co.Dt = (float)p.dt0;
        co.inv_dt0 = (float)(1.0 / p.dt0);
        co.mu = (float)p.mu;
        co.cA = (float)p.cA;
        co.alpha = (float)p.alpha;
        co.beta = (float)p.beta;
        co.eta = (float)p.eta;
        co.La2 = (float)(p.La * p.La);
        co.lambda0 = (float)p.lambda0;
        co.aPhi2 = (float)(p.aPhi * 2.0);
        co.kPhi = (float)p.kPhi;
        co.gamma = (float)p.gamma;
        co.A = (float)p.A;
        const double kk = 6.283185307179586476925286766559 / (N * h);
        co.k2s = (float)(kk * kk);
        co.inv_n3 = (float)(1.0 / ((double)N * N * N));
        co.anchoring = (p.A != 0.0) ? 1 : 0;
    }

    int alloc_slot(int si, const ttsh_params& p)
    {
        Slot& s = slots[si];
        const int N = p.Nx;
        if (size_index(N) < 0) return fail(TTSH_EUNSUPPORTED, "no FFT plan for N=" + std::to_string(N));
        const size_t bytes = droplet_bytes(N);
        if (budget && used + bytes > budget)
            return fail(TTSH_EBUDGET, "droplet of N=" + std::to_string(N) + " needs " + std::to_string(bytes) +
                                          " bytes; budget has " + std::to_string(budget - used) + " left");
        // tuned choices for N (autotuner table), unless the parameters fix them
        int tv = -1, ttz = 0, tpz = 0;
        if (!tune_lookup(N, &tv, &ttz, &tpz) || tv < 0 || tv >= RHS_NVARIANTS) { tv = -1; ttz = 0; tpz = 0; }
        const int want = p.rhs_variant >= 0 ? p.rhs_variant : (tv >= 0 ? tv : default_rhs_variant(N));
        int v = rhs_variant < 0 ? want : rhs_variant;
        if (!variant_fits(v, N)) {
            if (p.rhs_variant >= 0) return fail(TTSH_EUNSUPPORTED, "grid smaller than the requested K_rhs tile");
            v = RHS_16x16;  // engine-wide: every droplet fits it
        }
        HIPCHK(hipSetDevice(device));
        void* mem = nullptr;
        // a retired droplet's memory of the same size, once the work queued
        // before its retirement is done (the side stream waits for it)
        for (size_t k = 0; k < pool.size(); ++k)
            if (pool[k].N == N) {
                mem = pool[k].mem;
                HIPCHK(hipStreamWaitEvent(side(), pool[k].ev, 0));
                HIPCHK(hipEventDestroy(pool[k].ev));
                pooled -= pool[k].bytes;
                pool.erase(pool.begin() + (long)k);
                break;
            }
        if (!mem) {
            if (budget && used + pooled + bytes > budget) {
                int rc = release_pool();
                if (rc) return rc;
            }
            hipError_t e = dev_alloc(&mem, bytes);
            if (e != hipSuccess && !pool.empty()) {
                (void)hipGetLastError();
                int rc = release_pool();
                if (rc) return rc;
                e = dev_alloc(&mem, bytes);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                return fail(TTSH_ENOMEM, std::string("hipMalloc of droplet state failed: ") + hipGetErrorString(e));
            }
        }
        if (v != rhs_variant) {
            rhs_variant = v;
            drop_graphs();
        }
        s.used = true;
        s.prm = p;
        s.N = N;
        s.mem = mem;
        s.bytes = bytes;
        used += bytes;
        const size_t n = (size_t)N * N * N;
        char* b = static_cast<char*>(mem);
        auto take = [&](size_t sz) {
            char* r = b;
            b += align256(sz);
            return r;
        };
        Drop& D = s.drop;
        std::memset(&D, 0, sizeof(D));
        D.N = N;
        D.nst = fft_plan(N, D.rad, kMaxStages);
        const float2* t = nullptr;
        int rc = twiddles(N, &t);
        if (rc) return rc;
        D.tw = t;
        D.phi[0] = (float*)take(n * 4);
        D.phi[1] = (float*)take(n * 4);
        for (int a = 0; a < 3; ++a) D.P[a] = (float*)take(n * 4);
        for (int a = 0; a < 3; ++a) D.v[a] = (float*)take(n * 4);
        for (int q = 0; q < 2; ++q)
            for (int a = 0; a < 3; ++a) D.vs[q][a] = (float*)take(n * 4);
        for (int q = 0; q < 2; ++q)
            for (int a = 0; a < 3; ++a) D.ps[q][a] = (float*)take(n * 4);
        D.c = (float2*)take(n * 8);
        D.part = (double*)take((size_t)proj_items_max(N) * kPartSums * sizeof(double));
        D.stat = (double*)take(2 * sizeof(double));
        D.sigma = (float*)take(sizeof(float));
        // last: the divergence addresses it from the predictor buffers with
        // a non-negative 32-bit byte offset
        D.edge = (float*)take(n * 4);
        D.eslot[0] = 0;
        D.eslot[1] = 1;
        D.faithful = p.faithful_halo ? 1 : 0;
        make_coef(p, D.co);
        // zero state, sigma = 1 until the first step sets it, on the side
        // stream (fill kernels: no memset node, should a caller capture an
        // external engine); the upload that follows is ordered after them
        HIPCHK(launch_fill_u32(mem, 0u, bytes / 4, side()));
        HIPCHK(launch_fill_u32(D.sigma, 0x3f800000u, 1, side()));
        HIPCHK(publish_side());
        s.q = engine_parity;
        s.steps = 0;
        s.limit = -1;
        s.T = p.T0;
        s.rhs_tz = std::min(p.rhs_tz > 0 ? p.rhs_tz : (ttz > 0 ? ttz : 16), N);
        s.rhs_tz_set = p.rhs_tz > 0;
        s.proj_tz = std::min(p.corr_tz > 0 ? p.corr_tz : (tpz > 0 ? tpz : 8), N);
        tables_dirty = true;
        return TTSH_OK;
    }

    int free_slot(int si)
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        HIPCHK(hipSetDevice(device));
        if (us) {
            // keep the memory for the next droplet of this size; it may be
            // reused once the work queued so far is done
            hipEvent_t ev = nullptr;
            HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            HIPCHK(hipEventRecord(ev, st));
            pool.push_back(Pooled{s.N, s.mem, s.bytes, ev});
            pooled += s.bytes;
        } else {
            HIPCHK(hipStreamSynchronize(st));
            dev_free(s.mem);
        }
        used -= s.bytes;
        s = Slot();
        tables_dirty = true;
        return TTSH_OK;
    }

    // Free the pooled memory of retired droplets (after the queued work).
    int release_pool()
    {
        if (pool.empty()) return TTSH_OK;
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipStreamSynchronize(side()));
        for (auto& b : pool) {
            dev_free(b.mem);
            HIPCHK(hipEventDestroy(b.ev));
        }
        pool.clear();
        pooled = 0;
        return TTSH_OK;
    }

    // Make every droplet's buffer parity equal the engine parity by swapping
    // its ping-pong pointers (no data moves).
    void normalise_parity()
    {
        for (auto& s : slots) {
            if (!s.used || s.q == engine_parity) continue;
            std::swap(s.drop.phi[0], s.drop.phi[1]);
            std::swap(s.drop.eslot[0], s.drop.eslot[1]);
            for (int a = 0; a < 3; ++a) {
                std::swap(s.drop.vs[0][a], s.drop.vs[1][a]);
                std::swap(s.drop.ps[0][a], s.drop.ps[1][a]);
            }
            s.q = engine_parity;
            tables_dirty = true;
        }
    }

    int64_t remaining(const Slot& s) const { return s.limit < 0 ? INT64_MAX : std::max<int64_t>(0, s.limit - s.steps); }

    // Layout of a batch that fills the device.  The tuned tile and z-depth of
    // a size are chosen for that droplet alone, with enough z-chunks to fill
    // the device; each chunk re-reads two warm-up planes.  A batch that fills
    // the device by itself runs every droplet's K_rhs with one tile (the
    // batch entry of the tuning table; by default 32x8, three groups per CU)
    // and chunks of at least batch_rhs_tz planes, which cuts the warm-up share
    // and gives items of nearly equal cost.  It applies when at least two
    // droplets are active, none asks for a tile, every one fits the tile, and
    // the items make at least two rounds of resident groups.  K_rhs has no
    // reductions, so a droplet's result does not depend on this layout.
    // K_corr (no reductions either) and K_div take chunks of about
    // batch_corr_tz planes in the same batches.  K_div's sums, which set the
    // k = 0 coefficient and sigma, stay per tuned chunk (Drop::part), so they
    // do not change either.
    int batch_rhs_variant = RHS_32x8, batch_rhs_tz = 32, batch_corr_tz = 32;
    int tune_gen_seen = -1;

    // Re-read the batch entry of the tuning table after a reload; a running
    // batch is laid out again at its next rebuild.
    void refresh_tuning()
    {
        const int g = tune_generation();
        if (g == tune_gen_seen) return;
        tune_gen_seen = g;
        int v = RHS_32x8, rtz = 32, ctz = 32;
        if (!tune_batch_lookup(&v, &rtz, &ctz) || v < 0 || v >= RHS_NVARIANTS) {
            v = RHS_32x8;
            rtz = 32;
            ctz = 32;
        }
        batch_rhs_variant = v;
        batch_rhs_tz = std::max(1, rtz);
        batch_corr_tz = std::max(1, ctz);
        tables_dirty = true;
    }

    bool batch_layout() const
    {
        int active = 0;
        int64_t items = 0;
        int tx, ty;
        rhs_tile(batch_rhs_variant, &tx, &ty);
        for (const Slot& s : slots) {
            if (!s.used || remaining(s) <= 0) continue;
            if (s.prm.rhs_variant >= 0 || !variant_fits(batch_rhs_variant, s.N)) return false;
            ++active;
            const int tz = std::min(std::max(s.rhs_tz, batch_rhs_tz), s.N);
            items += (int64_t)((s.N + tx - 1) / tx) * ((s.N + ty - 1) / ty) * ((s.N + tz - 1) / tz);
        }
        return active >= 2 && items >= 2 * (int64_t)rhs_groups_per_cu(batch_rhs_variant) * cus;
    }

    // Rebuild descriptor and work tables from the active slots.
    int rebuild(int steps_this_call)
    {
        // droplets filled or uploaded on the side stream join the step
        // kernels from here on
        HIPCHK(uploads_into_st());
        up_pending = false;
        std::vector<Work> w[kNTables];
        const bool batch = batch_layout();
        const int rv = batch ? batch_rhs_variant : (rhs_variant < 0 ? 0 : rhs_variant);
        if (rv != rhs_active) {
            HIPCHK(hipStreamSynchronize(st));
            rhs_active = rv;
            drop_graphs();  // captured launches name the tile's kernel
        }
        int tx, ty;
        rhs_tile(rhs_active, &tx, &ty);
        for (int si = 0; si < (int)slots.size(); ++si) {
            Slot& s = slots[si];
            h_drops[si] = s.drop;
            if (!s.used) continue;
            s.published = true;
            const int64_t rem = remaining(s);
            // a droplet without a step limit never stops inside a launch
            // sequence, so its descriptor serves any chunk length
            h_drops[si].nsteps = s.limit < 0 ? kNoLimit : (int)std::min<int64_t>(rem, (int64_t)steps_this_call);
            if (rem <= 0) continue;
            const int N = s.N;
            // z is split into ceil(N/tz) chunks of near-equal depth
            const int rtz = batch && !s.rhs_tz_set ? std::min(std::max(s.rhs_tz, batch_rhs_tz), N) : s.rhs_tz;
            const int nzr = (N + rtz - 1) / rtz;
            for (int k = 0; k < nzr; ++k) {
                const int zb = k * N / nzr, ze = (k + 1) * N / nzr;
                for (int by = 0; by < (N + ty - 1) / ty; ++by)
                    for (int bx = 0; bx < (N + tx - 1) / tx; ++bx)
                        w[T_RHS].push_back(Work{si, bx, by, zb | (ze << 16)});
            }
            // divergence: the droplet's tuned z-chunks are its reduction blocks
            // (Drop::part); an item covers g consecutive blocks of one tile
            // (one, except in a device-filling batch)
            const int ntile = ((N + kProjTX - 1) / kProjTX) * ((N + kProjTY - 1) / kProjTY);
            const int nzp = (N + s.proj_tz - 1) / s.proj_tz;
            const int g = batch && s.prm.corr_tz <= 0 ? std::max(1, std::min(nzp, batch_corr_tz * nzp / N)) : 1;
            for (int k0 = 0; k0 < nzp; k0 += g) {
                const int k1 = std::min(nzp, k0 + g);
                const int zb = k0 * N / nzp, ze = k1 * N / nzp;
                for (int t = 0; t < ntile; ++t) w[T_PROJ].push_back(Work{si, t, zb | (ze << 16), k0 * ntile + t});
            }
            const int ctz = batch && s.prm.corr_tz <= 0 ? std::min(std::max(s.proj_tz, batch_corr_tz), N) : s.proj_tz;
            const int nzc = (N + ctz - 1) / ctz;
            for (int k = 0; k < nzc; ++k) {
                const int zb = k * N / nzc, ze = (k + 1) * N / nzc;
                for (int t = 0; t < ntile; ++t) w[T_CORR].push_back(Work{si, t, zb | (ze << 16), 0});
            }
            // k_proj_stats sums the blocks' FP64 partial sums in block order
            h_drops[si].npart = nzp * ntile;
            h_drops[si].nzp = nzp;
            w[T_STAT].push_back(Work{si, 0, 0, 0});
            const int sidx = size_index(N);
            const int L = fft_lines_per_group(N);
            const bool planes = fft_xy_fused(N);
            for (int b = 0; b < N; ++b) {
                if (planes) w[t_pl(sidx)].push_back(Work{si, 0, b, 0});
                for (int a = 0; a < N / L; ++a) {
                    w[t_xy(sidx)].push_back(Work{si, a, b, 0});
                    w[t_z(sidx)].push_back(Work{si, a, b, 0});
                }
            }
        }
        for (int k = 0; k < kNTables; ++k) {
            const int cnt = (int)w[k].size();
            if (cnt > tab[k].cap) {
                HIPCHK(hipStreamSynchronize(st));
                dev_free(tab[k].d);
                tab[k].d = nullptr;
                const int cap = std::max(cnt, 256);
                HIPCHK(dev_alloc(&tab[k].d, sizeof(Work) * cap));
                tab[k].cap = cap;
                drop_graphs();  // captured launches hold the old table
            }
            if (cnt) HIPCHK(hipMemcpyAsync(tab[k].d, w[k].data(), sizeof(Work) * cnt, hipMemcpyHostToDevice, st));
            tab[k].n = cnt;
            h_counts[k] = cnt;
        }
        // The fused x+y pass puts one work-group on each z-plane; it pays only
        // when a size has enough planes to give every CU two of them (a
        // batch), otherwise the per-plane transform work leaves CUs idle and
        // the single-axis passes are faster.
        xy_mask = 0;
        for (int sz = 0; sz < kNumGridSizes; ++sz)
            if (tab[t_pl(sz)].n >= 2 * cus) xy_mask |= 1u << sz;
        HIPCHK(hipMemcpyAsync(d_counts, h_counts, sizeof(int) * kNTables, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_drops, h_drops.data(), sizeof(Drop) * slots.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));  // host tables must outlive the copies
        tables_dirty = false;
        last_chunk = steps_this_call;
        return TTSH_OK;
    }

    unsigned sizes_present() const
    {
        unsigned m = 0;
        for (int s = 0; s < kNumGridSizes; ++s)
            if (tab[t_z(s)].n > 0) m |= 1u << s;
        return m;
    }

    // One step's launches.  graph: use the captured fixed grids.  ev (2*KK_N
    // events, may be null): kind k's first launch records ev[2k] at its start
    // and its last launch ev[2k+1] at its end, from the dispatch itself
    // (hipExtLaunchKernel), so the events time the kernels as a kernel trace
    // does; launched[k] tells whether kind k launched at all.
    int launch_step(int parity, int kstep, bool graph, hipEvent_t* ev = nullptr, bool* launched = nullptr)
    {
        // plain launches: one work-group per visit of the XCD-blocked order
        auto grid = [&](int t) { return graph ? ggrid[t] : ((tab[t].n + kXcds - 1) / kXcds) * kXcds; };
        auto on = [&](int s) { return graph ? ((gsizes >> s) & 1u) != 0 : tab[t_z(s)].n > 0; };
        const int* cnt = d_counts;
        // the launches of the step in order, with their kind
        struct L {
            int kind, what, s, t;
        };
        enum { W_RHS, W_DIV, W_STAT, W_FFT, W_FFT_XY, W_CORR };
        L list[4 + 5 * kNumGridSizes];
        int nl = 0;
        auto add = [&](int kind, int what, int s, int t) {
            if (grid(t) > 0) list[nl++] = L{kind, what, s, t};
        };
        add(KK_RHS, W_RHS, 0, T_RHS);
        add(KK_DIV, W_DIV, 0, T_PROJ);
        add(KK_DIV, W_STAT, 0, T_STAT);
        // x, y, z-Poisson, y inverse, x inverse; sizes with the fused x+y
        // pass do x+y in the x slots and nothing in the y slots
        const int axes[5] = {0, 1, 2, 1, 0};
        const int modes[5] = {FFT_MODE_FWD, FFT_MODE_FWD, FFT_MODE_POISSON, FFT_MODE_INV, FFT_MODE_INV};
        for (int f = 0; f < 5; ++f)
            for (int s = 0; s < kNumGridSizes; ++s) {
                if (!on(s)) continue;
                const bool fused = (((graph ? gxy_mask : xy_mask) >> s) & 1u) != 0;
                if (axes[f] != 2 && fused) {
                    if (axes[f] == 0) add(KK_FX + f, W_FFT_XY, s, t_pl(s));
                    continue;
                }
                add(KK_FX + f, W_FFT, s, axes[f] == 2 ? t_z(s) : t_xy(s));
            }
        add(KK_CORR, W_CORR, 0, T_CORR);
        if (launched)
            for (int k = 0; k < KK_N; ++k) launched[k] = false;
        for (int i = 0; i < nl; ++i) {
            const L& l = list[i];
            if (ev) {
                g_launch_ev.start = (i == 0 || list[i - 1].kind != l.kind) ? ev[2 * l.kind] : nullptr;
                g_launch_ev.stop = (i == nl - 1 || list[i + 1].kind != l.kind) ? ev[2 * l.kind + 1] : nullptr;
                if (launched) launched[l.kind] = true;
            }
            const int N = kGridSizes[l.s];
            const int f = l.kind - KK_FX;
            hipError_t e = hipSuccess;
            switch (l.what) {
            case W_RHS: e = launch_rhs(rhs_active, d_drops, tab[T_RHS].d, cnt + T_RHS, parity, kstep, grid(T_RHS), st); break;
            case W_DIV: e = launch_div(d_drops, tab[T_PROJ].d, cnt + T_PROJ, parity, kstep, grid(T_PROJ), st); break;
            case W_STAT: e = launch_proj_stats(d_drops, tab[T_STAT].d, cnt + T_STAT, kstep, grid(T_STAT), st); break;
            case W_FFT_XY: e = launch_fft_xy(N, modes[f], d_drops, tab[l.t].d, cnt + l.t, kstep, grid(l.t), st); break;
            case W_FFT: e = launch_fft(N, axes[f], modes[f], d_drops, tab[l.t].d, cnt + l.t, kstep, grid(l.t), st); break;
            default: e = launch_corr(d_drops, tab[T_CORR].d, cnt + T_CORR, parity, kstep, grid(T_CORR), st); break;
            }
            g_launch_ev = LaunchEvents{};
            HIPCHK(e);
        }
        return TTSH_OK;
    }

    int capture(int parity)
    {
        // fixed grids: enough work-groups to fill the device, never more than
        // a table can hold
        for (int t = 0; t < kNTables; ++t) ggrid[t] = std::max(0, std::min(tab[t].cap, cus * 8));
        // the stencil kernels' items differ in cost (z-depths, partial tiles):
        // one work-group per table slot leaves their balance to the hardware
        // dispatcher instead of a fixed stride over the items
        for (int t : {T_RHS, T_PROJ, T_CORR}) ggrid[t] = ((tab[t].cap + kXcds - 1) / kXcds) * kXcds;
        gsizes = sizes_present();
        gxy_mask = xy_mask;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = TTSH_OK;
        for (int k = 0; k < graph_steps && rc == TTSH_OK; ++k) rc = launch_step(parity ^ (k & 1), k, true);
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(TTSH_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&gexec[parity], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(TTSH_EHIP, std::string("graph instantiate failed: ") + hipGetErrorString(e));
        return TTSH_OK;
    }

    int any_active() const
    {
        int a = 0;
        for (auto& s : slots)
            if (s.used && remaining(s) > 0) ++a;
        return a;
    }

    static constexpr int kNoLimit = 1 << 30;

    // Limited droplets store min(remaining, chunk) as their step count, so
    // their descriptors must be rebuilt when the chunk length changes.
    bool any_limited() const
    {
        for (auto& s : slots)
            if (s.used && s.limit >= 0) return true;
        return false;
    }

    bool needs_rebuild(int64_t chunk) const
    {
        return tables_dirty || limits_bind(chunk) || (chunk != last_chunk && any_limited());
    }

    bool limits_bind(int64_t chunk) const
    {
        for (auto& s : slots)
            if (s.used && remaining(s) < chunk) return true;
        return false;
    }

    void advance(int64_t k)
    {
        for (auto& s : slots) {
            if (!s.used) continue;
            const int64_t d = std::min<int64_t>(remaining(s), k);
            s.steps += d;
            if (d > 0 && s.limit >= 0 && s.steps >= s.limit) s.fin_call = ncalls;
            for (int64_t i = 0; i < d; ++i) s.T = s.T + s.prm.dt0;  // T = T + Dt (TTSH3D.cu:342)
            if (d & 1) s.q ^= 1;
        }
        if (k & 1) engine_parity ^= 1;
    }

    int step(int64_t n, bool graphs)
    {
        HIPCHK(hipSetDevice(device));
        refresh_tuning();
        const int rc = step_chunks(n, graphs);
        if (call_ev[0]) HIPCHK(hipEventRecord(call_ev[ncalls % kCallEv], st));
        ++ncalls;
        return rc;
    }

    int step_chunks(int64_t n, bool graphs)
    {
        while (n > 0) {
            if (!any_active()) return TTSH_OK;
            normalise_parity();
            const bool g = graphs && use_graph && n >= graph_steps;
            const int64_t chunk = g ? graph_steps : std::min<int64_t>(n, 1 << 20);
            // per-droplet step counts of this chunk live in the descriptors
            if (needs_rebuild(chunk)) {
                int rc = rebuild((int)chunk);
                if (rc) return rc;
            }
            if (g) {
                if ((sizes_present() & ~gsizes) || xy_mask != gxy_mask) drop_graphs();
                if (!gexec[engine_parity]) {
                    const unsigned keep = gsizes;
                    int rc = capture(engine_parity);
                    if (rc) return rc;
                    (void)keep;
                }
                HIPCHK(hipGraphLaunch(gexec[engine_parity], st));
            } else {
                for (int64_t k = 0; k < chunk; ++k) {
                    int rc = launch_step(engine_parity ^ (int)(k & 1), (int)k, false);
                    if (rc) return rc;
                }
            }
            advance(chunk);
            n -= chunk;
        }
        return TTSH_OK;
    }

    // n plain steps with a HIP event at every kernel-kind boundary; ms[k]
    // gets the summed duration of kind k over the n steps.
    int profile(int64_t n, double ms[KK_N])
    {
        HIPCHK(hipSetDevice(device));
        refresh_tuning();
        for (int k = 0; k < KK_N; ++k) ms[k] = 0.0;
        constexpr int64_t kChunk = 256;
        constexpr int E = 2 * KK_N;
        std::vector<hipEvent_t> ev(E * kChunk);
        std::vector<char> went(KK_N * kChunk);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        int rc = TTSH_OK;
        while (n > 0 && rc == TTSH_OK) {
            const int64_t m = std::min(n, kChunk);
            normalise_parity();
            if (needs_rebuild(m)) rc = rebuild((int)m);
            bool launched[KK_N];
            for (int64_t s = 0; s < m && rc == TTSH_OK; ++s) {
                rc = launch_step(engine_parity ^ (int)(s & 1), (int)s, false, &ev[E * s], launched);
                for (int k = 0; k < KK_N; ++k) went[KK_N * s + k] = launched[k];
            }
            if (rc) break;
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) { rc = fail(TTSH_EHIP, hipGetErrorString(e)); break; }
            for (int64_t s = 0; s < m && rc == TTSH_OK; ++s)
                for (int k = 0; k < KK_N; ++k) {
                    if (!went[KK_N * s + k]) continue;
                    float t = 0.f;
                    e = hipEventElapsedTime(&t, ev[E * s + 2 * k], ev[E * s + 2 * k + 1]);
                    if (e != hipSuccess) { rc = fail(TTSH_EHIP, hipGetErrorString(e)); break; }
                    ms[k] += t;
                }
            advance(m);
            n -= m;
        }
        if (call_ev[0]) (void)hipEventRecord(call_ev[ncalls % kCallEv], st);
        ++ncalls;
        for (auto& e : ev) (void)hipEventDestroy(e);
        return rc;
    }

    // State in from host arrays, on the side stream: the step kernels queued
    // on st keep running; the droplet joins them at the next rebuild.
    int upload_state(int si, int dtype, const void* phi, const void* P, const void* v, const void* vsp,
                     const void* psp)
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        if (dtype != TTSH_F32 && dtype != TTSH_F64) return fail(TTSH_EINVAL, "dtype must be TTSH_F32 or TTSH_F64");
        if (!phi || !P || !v) return fail(TTSH_EINVAL, "phi, P and v are required");
        HIPCHK(hipSetDevice(device));
        normalise_parity();
        const size_t n = (size_t)s.N * s.N * s.N;
        const hipStream_t ss = side();
        if (s.published && us) {
            // steps already queued may still use this droplet's buffers
            HIPCHK(hipEventRecord(st_ev, st));
            HIPCHK(hipStreamWaitEvent(ss, st_ev, 0));
        }
        std::vector<float> tmp;
        auto put = [&](float* dst, const void* src, size_t off) -> int {
            const void* p = nullptr;
            if (dtype == TTSH_F32) {
                p = static_cast<const float*>(src) + off;
            } else {
                tmp.resize(n);
                const double* d = static_cast<const double*>(src) + off;
                for (size_t i = 0; i < n; ++i) tmp[i] = (float)d[i];
                p = tmp.data();
            }
            HIPCHK(hipMemcpyAsync(dst, p, n * sizeof(float), hipMemcpyHostToDevice, ss));
            HIPCHK(hipStreamSynchronize(ss));  // the host buffer is reused or released next
            return TTSH_OK;
        };
        const int q = s.q;
        int rc = put(s.drop.phi[q], phi, 0);
        for (int a = 0; a < 3 && !rc; ++a) rc = put(s.drop.P[a], P, a * n);
        for (int a = 0; a < 3 && !rc; ++a) rc = put(s.drop.v[a], v, a * n);
        // the previous step's predictors sit in the other parity
        for (int a = 0; a < 3 && !rc; ++a) rc = put(s.drop.vs[q ^ 1][a], vsp ? vsp : v, a * n);
        for (int a = 0; a < 3 && !rc; ++a) rc = put(s.drop.ps[q ^ 1][a], psp ? psp : P, a * n);
        if (rc) return rc;
        {
            const float* vs[3] = {s.drop.vs[q ^ 1][0], s.drop.vs[q ^ 1][1], s.drop.vs[q ^ 1][2]};
            const float* ps[3] = {s.drop.ps[q ^ 1][0], s.drop.ps[q ^ 1][1], s.drop.ps[q ^ 1][2]};
            HIPCHK(launch_fill_edges(vs, ps, s.drop.edge, s.drop.eslot[q ^ 1], s.N, ss));
            HIPCHK(publish_side());
        }
        s.steps = 0;
        s.fin_call = -1;
        s.T = s.prm.T0;
        tables_dirty = true;
        return TTSH_OK;
    }

    // State in from device memory (FP32, same layouts as upload_state): phi,
    // P, v and the previous step's predictors vsp, psp, copied on the
    // engine's stream.
    int upload_device(int si, const float* phi, const float* P, const float* v, const float* vsp, const float* psp)
    {
        Slot& s = slots[si];
        HIPCHK(hipSetDevice(device));
        normalise_parity();
        const size_t n = (size_t)s.N * s.N * s.N, b = n * sizeof(float);
        const int q = s.q;
        HIPCHK(hipMemcpyAsync(s.drop.phi[q], phi, b, hipMemcpyDeviceToDevice, st));
        for (int a = 0; a < 3; ++a) {
            HIPCHK(hipMemcpyAsync(s.drop.P[a], P + a * n, b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(s.drop.v[a], v + a * n, b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(s.drop.vs[q ^ 1][a], vsp + a * n, b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(s.drop.ps[q ^ 1][a], psp + a * n, b, hipMemcpyDeviceToDevice, st));
        }
        const float* vs[3] = {s.drop.vs[q ^ 1][0], s.drop.vs[q ^ 1][1], s.drop.vs[q ^ 1][2]};
        const float* ps[3] = {s.drop.ps[q ^ 1][0], s.drop.ps[q ^ 1][1], s.drop.ps[q ^ 1][2]};
        HIPCHK(launch_fill_edges(vs, ps, s.drop.edge, s.drop.eslot[q ^ 1], s.N, st));
        s.steps = 0;
        s.T = s.prm.T0;
        tables_dirty = true;
        return TTSH_OK;
    }

    // State out to device memory: phi, P, v and the last step's predictors.
    int download_device(int si, float* phi, float* P, float* v, float* vs_out, float* ps_out)
    {
        Slot& s = slots[si];
        HIPCHK(hipSetDevice(device));
        const size_t n = (size_t)s.N * s.N * s.N, b = n * sizeof(float);
        const int last = s.q ^ 1;  // parity of the last step's predictors
        HIPCHK(hipMemcpyAsync(phi, s.drop.phi[s.q], b, hipMemcpyDeviceToDevice, st));
        for (int a = 0; a < 3; ++a) {
            HIPCHK(hipMemcpyAsync(P + a * n, s.drop.P[a], b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(v + a * n, s.drop.v[a], b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(vs_out + a * n, s.drop.vs[last][a], b, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(ps_out + a * n, s.drop.ps[last][a], b, hipMemcpyDeviceToDevice, st));
        }
        return TTSH_OK;
    }

    int download(int si, int dtype, void* phi, void* P, void* v)
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        if (dtype != TTSH_F32 && dtype != TTSH_F64) return fail(TTSH_EINVAL, "dtype must be TTSH_F32 or TTSH_F64");
        HIPCHK(hipSetDevice(device));
        // a droplet past its step limit is read once its last step is done,
        // without waiting for steps queued after it; otherwise everything
        // queued so far completes first
        if (hipEvent_t ev = fin_event(s)) {
            HIPCHK(hipEventSynchronize(ev));
        } else {
            HIPCHK(uploads_into_st());
            HIPCHK(hipStreamSynchronize(st));
        }
        HIPCHK(hipStreamSynchronize(side()));
        const size_t n = (size_t)s.N * s.N * s.N;
        std::vector<float> tmp(n);
        auto get = [&](const float* src, void* dst, size_t off) -> int {
            if (!dst) return TTSH_OK;
            if (dtype == TTSH_F32) {
                HIPCHK(hipMemcpy(static_cast<float*>(dst) + off, src, n * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                HIPCHK(hipMemcpy(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
                double* d = static_cast<double*>(dst) + off;
                for (size_t i = 0; i < n; ++i) d[i] = tmp[i];
            }
            return TTSH_OK;
        };
        int rc = get(s.drop.phi[s.q], phi, 0);
        for (int a = 0; a < 3 && !rc; ++a) rc = get(s.drop.P[a], P, a * n);
        for (int a = 0; a < 3 && !rc; ++a) rc = get(s.drop.v[a], v, a * n);
        return rc;
    }

    int reductions(int si, double out[3])
    {
        Slot& s = slots[si];
        if (!s.used) return fail(TTSH_EINVAL, "slot not in use");
        HIPCHK(hipSetDevice(device));
        // a droplet past its step limit is reduced on the side stream after
        // its last step; otherwise on st after everything queued
        hipStream_t rs = st;
        if (hipEvent_t ev = fin_event(s)) {
            rs = side();
            HIPCHK(hipStreamWaitEvent(rs, ev, 0));
        } else {
            HIPCHK(uploads_into_st());
        }
        Drop d = s.drop;
        // the reduce kernel reads phi[parity]; point it at the current buffer
        d.phi[0] = s.drop.phi[s.q];
        HIPCHK(hipMemcpyAsync(d_rdrop, &d, sizeof(Drop), hipMemcpyHostToDevice, rs));
        HIPCHK(launch_reduce(d_rdrop, 1, 0, d_redpart, kRedBlocks, d_redout, rs));
        double raw[3];
        HIPCHK(hipMemcpyAsync(raw, d_redout, sizeof(raw), hipMemcpyDeviceToHost, rs));
        HIPCHK(hipStreamSynchronize(rs));
        const double h = s.prm.h, h3 = h * h * h;
        const double n = (double)s.N * s.N * s.N;
        out[0] = raw[0] * h3;
        out[1] = raw[1] / n;
        out[2] = 0.5 * raw[2] * h3;
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(out[k])) return fail(TTSH_ENONFINITE, "non-finite reduction: the droplet diverged");
        return TTSH_OK;
    }
};

// Exports of ttsh_run_ex.  At an export the state (and, for diagnostics, the
// centroid sums and the centroid slice) is copied into one of two device
// snapshot buffers on the compute stream; a second stream copies it to
// pinned host memory once that copy is done; a host thread formats and
// writes the files.  Stepping goes on meanwhile.  With both buffers in use
// the next export waits for the older one.
class SnapWriter {
public:
    ~SnapWriter() { (void)finish(); release(); }

    int open(Engine* eng, int slot, const std::string& run_dir, int fl)
    {
        e = eng;
        si = slot;
        dir = run_dir;
        flags = fl;
        const Slot& sl = e->slots[si];
        N = sl.N;
        n = (size_t)N * N * N;
        expo_data = sl.prm.ExpoData;
        HIPCHK(hipSetDevice(e->device));
        HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        const size_t nn = (size_t)N * N;
        for (int b = 0; b < 2; ++b) {
            if (flags & (TTSH_RUN_TEXT | TTSH_RUN_BINARY)) {
                HIPCHK(hipMalloc(&dsnap[b], sizeof(float) * 7 * n));
                HIPCHK(hipHostMalloc((void**)&hsnap[b], sizeof(float) * 7 * n, hipHostMallocDefault));
            }
            if (flags & TTSH_RUN_DIAG) {
                HIPCHK(hipMalloc(&dslice[b], sizeof(float) * (9 * nn + 1)));
                HIPCHK(hipHostMalloc((void**)&hslice[b], sizeof(float) * (9 * nn + 1), hipHostMallocDefault));
                HIPCHK(hipMalloc(&dsum[b], sizeof(double) * 4));
                HIPCHK(hipHostMalloc((void**)&hsum[b], sizeof(double) * 4, hipHostMallocDefault));
            }
            HIPCHK(hipEventCreateWithFlags(&snap_ev[b], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&done_ev[b], hipEventDisableTiming));
        }
        if (flags & TTSH_RUN_DIAG) {
            fdiag = std::fopen((dir + "/diag.dat").c_str(), "w");
            far = std::fopen((dir + "/a_R.txt").c_str(), "w");
            if (!fdiag || !far) return fail(TTSH_EIO, "cannot write diag.dat / a_R.txt in " + dir);
        }
        th = std::thread([this] { work(); });
        return TTSH_OK;
    }

    // Export m of the state at time T after `steps` steps (main thread).
    int submit(int m, double T, int64_t steps)
    {
        const int b = next;
        next ^= 1;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !busy[b] || err; });
            if (err) return fail(err, errmsg);
        }
        const Slot& sl = e->slots[si];
        hipStream_t st = e->st;
        HIPCHK(e->uploads_into_st());
        if (dsnap[b]) {
            const float* src[7] = {sl.drop.v[0], sl.drop.v[1], sl.drop.v[2], sl.drop.P[0],
                                   sl.drop.P[1], sl.drop.P[2], sl.drop.phi[sl.q]};
            for (int q = 0; q < 7; ++q)
                HIPCHK(hipMemcpyAsync(dsnap[b] + q * n, src[q], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
        }
        if (dslice[b]) {
            int rc = e->launch_diag(si, dsum[b], -1, dslice[b]);
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(snap_ev[b], st));
        HIPCHK(hipStreamWaitEvent(cs, snap_ev[b], 0));
        if (dsnap[b]) HIPCHK(hipMemcpyAsync(hsnap[b], dsnap[b], sizeof(float) * 7 * n, hipMemcpyDeviceToHost, cs));
        if (dslice[b]) {
            HIPCHK(hipMemcpyAsync(hslice[b], dslice[b], sizeof(float) * (9 * (size_t)N * N + 1), hipMemcpyDeviceToHost, cs));
            HIPCHK(hipMemcpyAsync(hsum[b], dsum[b], sizeof(double) * 4, hipMemcpyDeviceToHost, cs));
        }
        HIPCHK(hipEventRecord(done_ev[b], cs));
        {
            std::lock_guard<std::mutex> lk(mu);
            busy[b] = true;
            q.push_back(Job{b, m, T, steps});
        }
        cv.notify_all();
        if (flags & TTSH_RUN_SYNC) {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !busy[b] || err; });
            if (err) return fail(err, errmsg);
        }
        return TTSH_OK;
    }

    // Wait for every export to be written; the first error, if any.
    int finish()
    {
        if (th.joinable()) {
            {
                std::lock_guard<std::mutex> lk(mu);
                quit = true;
            }
            cv.notify_all();
            th.join();
        }
        if (fdiag) { std::fclose(fdiag); fdiag = nullptr; }
        if (far) { std::fclose(far); far = nullptr; }
        if (err) return fail(err, errmsg);
        return TTSH_OK;
    }

private:
    struct Job {
        int b, m;
        double T;
        int64_t steps;
    };

    void work()
    {
        (void)hipSetDevice(e->device);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return quit || !q.empty(); });
                if (q.empty()) break;
                j = q.front();
                q.pop_front();
            }
            int rc = write_job(j);
            {
                std::lock_guard<std::mutex> lk(mu);
                busy[j.b] = false;
                if (rc && !err) {
                    err = rc;
                    errmsg = g_last_error;
                }
            }
            cv.notify_all();
        }
    }

    int write_job(const Job& j)
    {
        const hipError_t he = hipEventSynchronize(done_ev[j.b]);
        if (he != hipSuccess) return fail(TTSH_EHIP, std::string("snapshot copy failed: ") + hipGetErrorString(he));
        const std::string stem = dir + "/conf_" + std::to_string(j.m);
        int rc = TTSH_OK;
        if (flags & TTSH_RUN_TEXT) {
            const float* h = hsnap[j.b];
            rc = write_conf_host((stem + ".dat").c_str(), N, expo_data, h + 6 * n, h + 3 * n, h);
            if (rc) return rc;
        }
        if (flags & TTSH_RUN_BINARY) {
            const float* h = hsnap[j.b];
            const float* f[7] = {h, h + n, h + 2 * n, h + 3 * n, h + 4 * n, h + 5 * n, h + 6 * n};
            rc = write_conf_bin_host((stem + ".bin").c_str(), N, j.m, j.T, j.steps, f);
            if (rc) return rc;
        }
        if (flags & TTSH_RUN_DIAG) {
            ttsh_centroid c;
            Engine::centroid_from_sums(hsum[j.b], &c);
            double R[120], ctr[2];
            rc = contour_radius(N, hslice[j.b], 0.5, R, ctr);  // phi is field 0 of the slice
            if (rc) return rc;
            std::fprintf(fdiag, "%d %.15g %d %d %d %.17g %.17g %.17g %.17g\n", j.m, j.T, c.idx[0], c.idx[1], c.idx[2],
                         c.c[0], c.c[1], c.c[2], c.mass);
            for (int a = 0; a < 120; ++a) std::fprintf(far, a ? " %.18e" : "%.18e", R[a]);
            std::fputc('\n', far);
            if (std::fflush(fdiag) || std::fflush(far)) return fail(TTSH_EIO, "write failed: diag.dat / a_R.txt");
        }
        return TTSH_OK;
    }

    void release()
    {
        if (!e) return;
        (void)hipSetDevice(e->device);
        if (cs) (void)hipStreamSynchronize(cs);
        for (int b = 0; b < 2; ++b) {
            if (dsnap[b]) (void)hipFree(dsnap[b]);
            if (hsnap[b]) (void)hipHostFree(hsnap[b]);
            if (dslice[b]) (void)hipFree(dslice[b]);
            if (hslice[b]) (void)hipHostFree(hslice[b]);
            if (dsum[b]) (void)hipFree(dsum[b]);
            if (hsum[b]) (void)hipHostFree(hsum[b]);
            if (snap_ev[b]) (void)hipEventDestroy(snap_ev[b]);
            if (done_ev[b]) (void)hipEventDestroy(done_ev[b]);
            dsnap[b] = hsnap[b] = dslice[b] = hslice[b] = nullptr;
            dsum[b] = hsum[b] = nullptr;
            snap_ev[b] = done_ev[b] = nullptr;
        }
        if (cs) (void)hipStreamDestroy(cs);
        cs = nullptr;
    }

    Engine* e = nullptr;
    int si = 0, N = 0, flags = 0, expo_data = 1, next = 0;
    size_t n = 0;
    std::string dir;
    hipStream_t cs = nullptr;
    float* dsnap[2] = {nullptr, nullptr};
    float* hsnap[2] = {nullptr, nullptr};
    float* dslice[2] = {nullptr, nullptr};
    float* hslice[2] = {nullptr, nullptr};
    double* dsum[2] = {nullptr, nullptr};
    double* hsum[2] = {nullptr, nullptr};
    hipEvent_t snap_ev[2] = {nullptr, nullptr}, done_ev[2] = {nullptr, nullptr};
    FILE* fdiag = nullptr;
    FILE* far = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> q;
    bool busy[2] = {false, false};
    bool quit = false;
    int err = TTSH_OK;
    std::string errmsg;
};

}  // namespace ttsh

using namespace ttsh;

struct ttsh_sim {
    Engine eng;
};

struct ttsh_batch {
    Engine eng;
};

// ---- allocation-free stepping (ttsh_step_noalloc) ---------------------------
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) step a droplet whose
// state they own in device memory.  Every buffer the step needs beyond that
// state -- the phi and predictor ping-pong partners, the complex projection
// buffer, edge columns, divergence sums, sigma, twiddles, descriptor and work
// tables -- lives in one file-scope device array per GPU, held until the
// process ends.  The kernels address a droplet's fields as 32-bit offsets
// from its lowest buffer, so the whole droplet sits in this one array and the
// caller's state is copied in and out around the steps.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocMaxN = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev`: taken once per
// device with hipGetSymbolAddress (each GPU holds its own copy of the array
// at its own address; the array's name in host code is not that address).
int scratch_address(int dev, void** out)
{
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(TTSH_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return TTSH_OK;
}
}  // namespace

extern "C" {

size_t ttsh_droplet_bytes(int N) { return droplet_bytes(N); }

int ttsh_create(const ttsh_params* p, int device, ttsh_sim** out)
{
    if (!out) return fail(TTSH_EINVAL, "null output handle");
    *out = nullptr;
    int rc = ttsh_validate_params(p);
    if (rc) return rc;
    ttsh_sim* s = new ttsh_sim();
    rc = s->eng.init(device, 1, 0, false);
    if (!rc) rc = s->eng.alloc_slot(0, *p);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return TTSH_OK;
}

int ttsh_destroy(ttsh_sim* s)
{
    delete s;
    return TTSH_OK;
}

int ttsh_set_state(ttsh_sim* s, int dtype, const void* phi, const void* P, const void* v, const void* vstar_prev,
                   const void* pstar_prev)
{
    if (!s) return fail(TTSH_EINVAL, "null handle");
    return s->eng.upload_state(0, dtype, phi, P, v, vstar_prev, pstar_prev);
}

int ttsh_get_state(ttsh_sim* s, int dtype, void* phi, void* P, void* v)
{
    if (!s) return fail(TTSH_EINVAL, "null handle");
    return s->eng.download(0, dtype, phi, P, v);
}

int ttsh_step(ttsh_sim* s, int64_t n)
{
    if (!s) return fail(TTSH_EINVAL, "null handle");
    if (n < 0) return fail(TTSH_EINVAL, "negative step count");
    return s->eng.step(n, false);
}

int ttsh_synchronize(ttsh_sim* s)
{
    if (!s) return fail(TTSH_EINVAL, "null handle");
    return s->eng.wait_calls(0);
}

int ttsh_reductions(ttsh_sim* s, double out[3])
{
    if (!s || !out) return fail(TTSH_EINVAL, "null argument");
    return s->eng.reductions(0, out);
}

int ttsh_profile(ttsh_sim* s, int64_t n, double kernel_ms[8])
{
    if (!s || !kernel_ms) return fail(TTSH_EINVAL, "null argument");
    if (n < 0) return fail(TTSH_EINVAL, "negative step count");
    return s->eng.profile(n, kernel_ms);
}

const char* ttsh_kernel_name(int k) { return (k >= 0 && k < KK_N) ? kKindNames[k] : ""; }

int ttsh_get_clock(ttsh_sim* s, int64_t* steps, double* T)
{
    if (!s) return fail(TTSH_EINVAL, "null handle");
    if (steps) *steps = s->eng.slots[0].steps;
    if (T) *T = s->eng.slots[0].T;
    return TTSH_OK;
}

int ttsh_get_aux(ttsh_sim* s, int which, float* out)
{
    if (!s || !out) return fail(TTSH_EINVAL, "null argument");
    Engine& e = s->eng;
    Slot& sl = e.slots[0];
    HIPCHK(hipSetDevice(e.device));
    HIPCHK(hipStreamSynchronize(e.st));
    HIPCHK(hipStreamSynchronize(e.side()));
    const size_t n = (size_t)sl.N * sl.N * sl.N;
    const int last = sl.q ^ 1;  // parity of the last step's predictor
    if (which == 0 || which == 1) {
        for (int a = 0; a < 3; ++a)
            HIPCHK(hipMemcpy(out + a * n, which == 0 ? sl.drop.vs[last][a] : sl.drop.ps[last][a], n * sizeof(float),
                             hipMemcpyDeviceToHost));
        return TTSH_OK;
    }
    if (which == 2) {
        HIPCHK(hipMemcpy(out, sl.drop.c, n * sizeof(float2), hipMemcpyDeviceToHost));
        // p_p = sigma * Im, with the last step's power-of-two scale (exact)
        float sig = 1.f;
        HIPCHK(hipMemcpy(&sig, sl.drop.sigma, sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) out[2 * i + 1] *= sig;
        return TTSH_OK;
    }
    return fail(TTSH_EINVAL, "which must be 0, 1 or 2");
}

int ttsh_write_conf(ttsh_sim* s, const char* path)
{
    if (!s || !path) return fail(TTSH_EINVAL, "null argument");
    const int N = s->eng.slots[0].N;
    const size_t n = (size_t)N * N * N;
    std::vector<float> phi(n), P(3 * n), v(3 * n);
    int rc = s->eng.download(0, TTSH_F32, phi.data(), P.data(), v.data());
    if (rc) return rc;
    return write_conf_host(path, N, s->eng.slots[0].prm.ExpoData, phi.data(), P.data(), v.data());
}

int ttsh_run(ttsh_sim* s, const char* run_dir, int64_t* steps_done)
{
    return ttsh_run_ex(s, run_dir, TTSH_RUN_TEXT, steps_done);
}

int ttsh_run_ex(ttsh_sim* s, const char* run_dir, int flags, int64_t* steps_done)
{
    if (!s || !run_dir) return fail(TTSH_EINVAL, "null argument");
    if (flags & ~(TTSH_RUN_TEXT | TTSH_RUN_BINARY | TTSH_RUN_DIAG | TTSH_RUN_SYNC))
        return fail(TTSH_EINVAL, "unknown run flag");
    Slot& sl = s->eng.slots[0];
    const ttsh_params& p = sl.prm;
    SnapWriter w;
    const bool out = (flags & (TTSH_RUN_TEXT | TTSH_RUN_BINARY | TTSH_RUN_DIAG)) != 0;
    int rc;
    if (out) {
        rc = w.open(&s->eng, 0, run_dir, flags);
        if (rc) return rc;
    }
    if (p.T0 == 0 && out) {
        rc = w.submit(0, sl.T, sl.steps);
        if (rc) return rc;
    }
    // evolve() (TTSH3D.cu:311-331): step, T += Dt, export on a dte crossing,
    // stop once T > Ts.  The export schedule is replayed on the host in the
    // same double arithmetic, so steps between exports run as one batch.
    double T = sl.T;
    const double Dt = p.dt0;
    int64_t total = 0, pending = 0;
    for (;;) {
        T = T + Dt;
        ++pending;
        const bool expo = std::floor(T / p.dte) > std::floor((T - Dt) / p.dte);
        const bool stop = T > p.Ts;
        if (expo || stop) {
            rc = ttsh_step(s, pending);
            if (rc) return rc;
            total += pending;
            pending = 0;
            if (expo && out) {
                rc = w.submit((int)std::floor(T / p.dte), sl.T, sl.steps);
                if (rc) return rc;
            }
            if (stop) break;
        }
    }
    if (steps_done) *steps_done = total;
    rc = ttsh_synchronize(s);
    if (rc) return rc;
    return w.finish();
}

int ttsh_get_centroid(ttsh_sim* s, ttsh_centroid* out)
{
    if (!s || !out) return fail(TTSH_EINVAL, "null argument");
    return s->eng.centroid(0, out);
}

int ttsh_get_slice(ttsh_sim* s, int z, float* out, int* z_used)
{
    if (!s || !out) return fail(TTSH_EINVAL, "null argument");
    return s->eng.slice(0, z, out, z_used);
}

// ---- batches -----------------------------------------------------------------

int ttsh_batch_create(int device, int max_slots, size_t hbm_budget, int use_graph, ttsh_batch** out)
{
    if (!out) return fail(TTSH_EINVAL, "null output handle");
    *out = nullptr;
    if (max_slots < 1 || max_slots > 4096) return fail(TTSH_EINVAL, "max_slots must be in [1, 4096]");
    ttsh_batch* b = new ttsh_batch();
    int rc = b->eng.init(device, max_slots, hbm_budget, use_graph != 0);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return TTSH_OK;
}

int ttsh_batch_destroy(ttsh_batch* b)
{
    delete b;
    return TTSH_OK;
}

int ttsh_batch_admit(ttsh_batch* b, const ttsh_params* p, int dtype, const void* phi, const void* P, const void* v,
                     const void* vstar_prev, const void* pstar_prev, int* slot)
{
    if (!b || !slot) return fail(TTSH_EINVAL, "null argument");
    int rc = ttsh_validate_params(p);
    if (rc) return rc;
    Engine& e = b->eng;
    int si = -1;
    for (int i = 0; i < (int)e.slots.size(); ++i)
        if (!e.slots[i].used) { si = i; break; }
    if (si < 0) return fail(TTSH_EFULL, "no free batch slot");
    rc = e.alloc_slot(si, *p);
    if (rc) {
        if (e.slots[si].used) (void)e.free_slot(si);
        return rc;
    }
    rc = e.upload_state(si, dtype, phi, P, v, vstar_prev, pstar_prev);
    if (rc) {
        (void)e.free_slot(si);
        return rc;
    }
    *slot = si;
    return TTSH_OK;
}

static int check_slot(ttsh_batch* b, int slot)
{
    if (!b) return fail(TTSH_EINVAL, "null handle");
    if (slot < 0 || slot >= (int)b->eng.slots.size() || !b->eng.slots[slot].used)
        return fail(TTSH_EINVAL, "invalid slot " + std::to_string(slot));
    return TTSH_OK;
}

int ttsh_batch_step(ttsh_batch* b, int64_t n)
{
    if (!b) return fail(TTSH_EINVAL, "null handle");
    if (n < 0) return fail(TTSH_EINVAL, "negative step count");
    return b->eng.step(n, true);
}

int ttsh_batch_set_limit(ttsh_batch* b, int slot, int64_t max_steps)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    b->eng.slots[slot].limit = max_steps < 0 ? -1 : max_steps;
    b->eng.tables_dirty = true;
    return TTSH_OK;
}

int ttsh_batch_steps_done(ttsh_batch* b, int slot, int64_t* steps)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (steps) *steps = b->eng.slots[slot].steps;
    return TTSH_OK;
}

int ttsh_batch_get_state(ttsh_batch* b, int slot, int dtype, void* phi, void* P, void* v)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    return b->eng.download(slot, dtype, phi, P, v);
}

int ttsh_batch_reductions(ttsh_batch* b, int slot, double out[3])
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!out) return fail(TTSH_EINVAL, "null output");
    return b->eng.reductions(slot, out);
}

int ttsh_batch_retire(ttsh_batch* b, int slot)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    return b->eng.free_slot(slot);
}

int ttsh_batch_usage(ttsh_batch* b, size_t* bytes_used, int* active)
{
    if (!b) return fail(TTSH_EINVAL, "null handle");
    if (bytes_used) *bytes_used = b->eng.used;
    if (active) {
        int a = 0;
        for (auto& s : b->eng.slots)
            if (s.used) ++a;
        *active = a;
    }
    return TTSH_OK;
}

int ttsh_batch_synchronize(ttsh_batch* b)
{
    if (!b) return fail(TTSH_EINVAL, "null handle");
    return b->eng.wait_calls(0);
}

int ttsh_batch_wait(ttsh_batch* b, int keep)
{
    if (!b) return fail(TTSH_EINVAL, "null handle");
    if (keep < 0) return fail(TTSH_EINVAL, "keep must be >= 0");
    return b->eng.wait_calls(keep);
}

int ttsh_batch_ready(ttsh_batch* b, int slot, int* ready)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!ready) return fail(TTSH_EINVAL, "null output");
    HIPCHK(hipSetDevice(b->eng.device));
    *ready = 0;
    if (hipEvent_t ev = b->eng.fin_event(b->eng.slots[slot])) {
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) *ready = 1;
        else if (e != hipErrorNotReady) return fail(TTSH_EHIP, std::string("hipEventQuery failed: ") + hipGetErrorString(e));
    }
    return TTSH_OK;
}

int ttsh_step_noalloc(const ttsh_params* p, float* phi, float* P, float* v, float* vstar_prev, float* pstar_prev,
                      int64_t n, void* stream)
{
    if (!p || !phi || !P || !v || !vstar_prev || !pstar_prev) return fail(TTSH_EINVAL, "null argument");
    if (n < 0) return fail(TTSH_EINVAL, "negative step count");
    int rc = ttsh_validate_params(p);
    if (rc) return rc;
    if (p->Nx > kNoAllocMaxN)
        return fail(TTSH_EINVAL, "ttsh_step_noalloc: N=" + std::to_string(p->Nx) + " exceeds its maximum N=" +
                                     std::to_string(kNoAllocMaxN) + " (the scratch array holds " +
                                     std::to_string(kScratchBytes >> 20) + " MiB)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TTSH_EINVAL, "device index out of range");
    // calls on one device share its scratch array: serialise them
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    void* base = nullptr;
    rc = scratch_address(dev, &base);
    if (rc) return rc;
    Engine e;
    rc = e.init_external(dev, static_cast<hipStream_t>(stream), base, kScratchBytes);
    if (!rc) rc = e.alloc_slot(0, *p);
    if (!rc) rc = e.upload_device(0, phi, P, v, vstar_prev, pstar_prev);
    if (!rc) rc = e.step(n, false);
    if (!rc) rc = e.download_device(0, phi, P, v, vstar_prev, pstar_prev);
    if (!rc) HIPCHK(hipStreamSynchronize(e.st));
    return rc;
}

}  // extern "C"
