// ttsh_host.cpp -- host-only parts of libttsh3d: parameters, input.dat I/O,
// the conf_<m>.dat snapshot writer and the seeded initial conditions.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ttsh3d.h"
#include "ttsh_internal.h"

namespace ttsh {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

// Factor N for the device FFT: radix 4 first, then 2, then 3.  Returns the
// number of stages, or 0 when N has another prime factor.
int fft_plan(int N, int* rad, int maxst)
{
    int m = N, n = 0;
    while (m % 4 == 0 && n < maxst) { rad[n++] = 4; m /= 4; }
    while (m % 2 == 0 && n < maxst) { rad[n++] = 2; m /= 2; }
    while (m % 3 == 0 && n < maxst) { rad[n++] = 3; m /= 3; }
    while (m % 5 == 0 && n < maxst) { rad[n++] = 5; m /= 5; }
    return (m == 1) ? n : 0;
}

bool supported_N(int N, std::string* why)
{
    for (int s : kGridSizes)
        if (s == N) return true;
    if (why) {
        std::string list;
        for (int s : kGridSizes) list += (list.empty() ? "" : ", ") + std::to_string(s);
        *why = "no FFT plan for N = " + std::to_string(N) + " (sizes: " + list + ")";
    }
    return false;
}

// ---------------------------------------------------------------------------
// Tuning table: per grid size N, the K_rhs tile variant, its z-depth and the
// z-depth of the divergence/corrector stencils, as text lines
// "N rhs_variant rhs_tz proj_tz"; and the layout of a batch that fills the
// device, a line "batch rhs_variant rhs_tz proj_tz" (the K_rhs tile every
// droplet then runs, the minimum K_rhs z-depth, the divergence/corrector
// z-depth).  Written by the autotuner (ttsh3d/autotune.py); read from
// $TTSH_TUNE_FILE or tuned.txt beside libttsh3d.so.
struct TuneEntry {
    int rhs_variant, rhs_tz, proj_tz;
};
static std::mutex g_tune_mu;
static std::map<int, TuneEntry> g_tune;
static TuneEntry g_tune_batch{-1, 0, 0};
static bool g_tune_loaded = false;
static std::string g_tune_path;
static std::atomic<int> g_tune_gen{0};

static std::string default_tune_path()
{
    const char* env = std::getenv("TTSH_TUNE_FILE");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&default_tune_path), &info) && info.dli_fname) {
        std::string so(info.dli_fname);
        const size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuned.txt";
    }
    return "tuned.txt";
}

static int load_tune_locked(const std::string& path)
{
    g_tune.clear();
    g_tune_batch = TuneEntry{-1, 0, 0};
    g_tune_loaded = true;
    g_tune_path = path;
    g_tune_gen.fetch_add(1);
    std::ifstream in(path);
    if (!in.is_open()) return TTSH_EIO;
    std::string line;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string key;
        if (!(ss >> key)) continue;
        int v, rtz, ptz;
        if (!(ss >> v >> rtz >> ptz)) continue;
        if (key == "batch") {
            if (v >= 0 && rtz > 0 && ptz > 0) g_tune_batch = TuneEntry{v, rtz, ptz};
        } else {
            char* end = nullptr;
            const long N = std::strtol(key.c_str(), &end, 10);
            if (end && *end == 0 && N > 0) g_tune[(int)N] = TuneEntry{v, rtz, ptz};
        }
    }
    return TTSH_OK;
}

bool tune_batch_lookup(int* rhs_variant, int* rhs_tz, int* proj_tz)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_loaded) (void)load_tune_locked(default_tune_path());
    if (g_tune_batch.rhs_variant < 0) return false;
    *rhs_variant = g_tune_batch.rhs_variant;
    *rhs_tz = g_tune_batch.rhs_tz;
    *proj_tz = g_tune_batch.proj_tz;
    return true;
}

int tune_generation() { return g_tune_gen.load(); }

bool tune_lookup(int N, int* rhs_variant, int* rhs_tz, int* proj_tz)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_loaded) (void)load_tune_locked(default_tune_path());
    auto it = g_tune.find(N);
    if (it == g_tune.end()) return false;
    *rhs_variant = it->second.rhs_variant;
    *rhs_tz = it->second.rhs_tz;
    *proj_tz = it->second.proj_tz;
    return true;
}

// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): a counter-based stream.
static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in (0,1) for (seed, cell, stream).
static inline double unif(uint64_t seed, uint64_t cell, uint64_t k)
{
    const uint64_t u = splitmix64(splitmix64(seed) ^ (cell * 64ull + k));
    return ((double)(u >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Standard normal by Box-Muller from streams (k, k+1).
static inline double normal(uint64_t seed, uint64_t cell, uint64_t k)
{
    const double u1 = unif(seed, cell, k), u2 = unif(seed, cell, k + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
}

}  // namespace ttsh

using namespace ttsh;

extern "C" {

const char* ttsh_last_error(void) { return g_last_error.c_str(); }

const char* ttsh_version(void) { return "ttsh3d-mi355x 0.1"; }

int ttsh_tuning_reload(const char* path)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const int rc = load_tune_locked(path ? std::string(path) : default_tune_path());
    return rc == TTSH_OK ? TTSH_OK : fail(rc, "cannot read tuning table " + g_tune_path);
}

int ttsh_tuning_lookup_batch(int* rhs_variant, int* rhs_tz, int* proj_tz)
{
    if (!rhs_variant || !rhs_tz || !proj_tz) return fail(TTSH_EINVAL, "null argument");
    if (!tune_batch_lookup(rhs_variant, rhs_tz, proj_tz)) return fail(TTSH_EINVAL, "no batch entry in the tuning table");
    return TTSH_OK;
}

int ttsh_tuning_lookup(int N, int* rhs_variant, int* rhs_tz, int* proj_tz)
{
    if (!rhs_variant || !rhs_tz || !proj_tz) return fail(TTSH_EINVAL, "null argument");
    if (!tune_lookup(N, rhs_variant, rhs_tz, proj_tz))
        return fail(TTSH_EINVAL, "no tuning entry for N=" + std::to_string(N));
    return TTSH_OK;
}

int ttsh_params_default(ttsh_params* p)
{
    if (!p) return fail(TTSH_EINVAL, "null params");
    std::memset(p, 0, sizeof(*p));
    // run_cuda.pl:7-44
    p->TimeScheme = 1;
    p->RKStage = 16;
    p->ExplImpi = 1;
    p->AdapTime = 0;
    p->ExpoData = 1;
    p->InitCond = 1;
    p->isSimplified = 1;
    p->checkUnStable = 0;
    p->Nx = p->Ny = p->Nz = 64;
    p->Nb = 3;
    p->h = 1.0;
    p->dt0 = 0.002;
    p->T0 = 0.0;
    p->Ts = 1000.0;
    p->dte = 10.0;
    p->dtVarC = 0.01;
    p->dtMax = 0.000005;
    p->mu = 1.0;
    p->cA = -1.0;
    p->alpha = -0.02;
    p->beta = 4.366;
    p->eta = -0.187;
    p->La = 1.0;
    p->kappa = -0.364;
    p->lambda0 = 9.0;
    p->aPhi = 1.0;
    p->kPhi = 1.0;
    p->gamma = 1.0;
    p->A = 0.0;
    p->init0 = 0.0001;
    p->ReadSuccess = -8848;
    p->faithful_halo = 1;
    p->rhs_variant = -1;
    p->rhs_tz = 0;
    p->corr_tz = 0;
    return TTSH_OK;
}

int ttsh_parse_input_dat(const char* path, ttsh_params* p)
{
    if (!path || !p) return fail(TTSH_EINVAL, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) return fail(TTSH_EIO, std::string("cannot open ") + path);
    ttsh_params q = *p;
    q.ReadSuccess = 0;
    // Read order of GetInput, TTSH3D.cu:702-739.
    in >> q.TimeScheme >> q.RKStage >> q.ExplImpi >> q.AdapTime >> q.ExpoData >> q.InitCond >> q.isSimplified >>
        q.checkUnStable >> q.Nx >> q.Ny >> q.Nz >> q.Nb >> q.h >> q.dt0 >> q.T0 >> q.Ts >> q.dte >> q.dtVarC >>
        q.dtMax >> q.mu >> q.cA >> q.alpha >> q.beta >> q.eta >> q.La >> q.kappa >> q.lambda0 >> q.aPhi >> q.kPhi >>
        q.gamma >> q.A >> q.init0;
    if (!in) return fail(TTSH_EIO, std::string("input file is short or malformed: ") + path);
    in >> q.ReadSuccess;
    if (!in || q.ReadSuccess != -8848) {
        *p = q;
        return fail(TTSH_ESENTINEL, "Error while reading the input file!");
    }
    *p = q;
    return TTSH_OK;
}

int ttsh_write_input_dat(const char* path, const ttsh_params* p)
{
    if (!path || !p) return fail(TTSH_EINVAL, "null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(TTSH_EIO, std::string("cannot write ") + path);
    // Perl interpolates numbers with %.15g; write order run_cuda.pl:54-86.
    std::fprintf(f, "%d\n%d\n%d\n%d\n%d\n%d\n%d\n%d\n", p->TimeScheme, p->RKStage, p->ExplImpi, p->AdapTime,
                 p->ExpoData, p->InitCond, p->isSimplified, p->checkUnStable);
    std::fprintf(f, "%d\n%d\n%d\n%d\n", p->Nx, p->Ny, p->Nz, p->Nb);
    const double d[] = {p->h,  p->dt0,   p->T0,    p->Ts,    p->dte,   p->dtVarC, p->dtMax, p->mu,
                        p->cA, p->alpha, p->beta,  p->eta,   p->La,    p->kappa,  p->lambda0, p->aPhi,
                        p->kPhi, p->gamma, p->A, p->init0};
    for (double v : d) std::fprintf(f, "%.15g\n", v);
    std::fprintf(f, "%d\n", p->ReadSuccess);
    const bool ok = std::fclose(f) == 0;
    return ok ? TTSH_OK : fail(TTSH_EIO, std::string("write failed: ") + path);
}

int ttsh_validate_params(const ttsh_params* p)
{
    if (!p) return fail(TTSH_EINVAL, "null params");
    if (p->ReadSuccess != -8848) return fail(TTSH_ESENTINEL, "ReadSuccess sentinel is not -8848");
    if (p->TimeScheme != 1)
        return fail(TTSH_EUNSUPPORTED, "only TimeScheme 1 (explicit Euler) evolves the fields (TTSH3D.cu:338-341)");
    if (p->Nx != p->Ny || p->Nx != p->Nz)
        return fail(TTSH_EUNSUPPORTED, "grid must be cubic (Nx == Ny == Nz)");
    std::string why;
    if (!supported_N(p->Nx, &why)) return fail(TTSH_EUNSUPPORTED, why);
    if (p->Nb < 1) return fail(TTSH_EINVAL, "Nb must be >= 1");
    if (!(p->h > 0.0) || !(p->dt0 > 0.0)) return fail(TTSH_EINVAL, "h and dt0 must be positive");
    if (!(p->dte > 0.0)) return fail(TTSH_EINVAL, "dte must be positive");
    if (p->rhs_variant < -1 || p->rhs_variant >= kRhsVariants) return fail(TTSH_EINVAL, "rhs_variant out of range");
    if (p->rhs_tz < 0 || p->corr_tz < 0) return fail(TTSH_EINVAL, "tile depths must be >= 0");
    return TTSH_OK;
}

int ttsh_init_reference(const ttsh_params* p, uint64_t seed, double* phi, double* P, double* v, double* vstar,
                        double* pstar)
{
    if (!p || !phi || !P || !v) return fail(TTSH_EINVAL, "null argument");
    if (p->Nx != p->Ny || p->Nx != p->Nz) return fail(TTSH_EUNSUPPORTED, "grid must be cubic");
    const int N = p->Nx;
    const size_t n = (size_t)N * N * N;
    const double a = p->init0;
    // InitConf, InitCond == 1 (TTSH3D.cu:798-825): i is x, j is y, k is z.
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t idx = ((size_t)k * N + j) * N + i;
                const uint64_t c = idx;
                v[idx] = 0 + a * (unif(seed, c, 0) - 0.5);
                v[n + idx] = 0 + a * (unif(seed, c, 1) - 0.5);
                v[2 * n + idx] = 0 + a * (unif(seed, c, 2) - 0.5);
                if (vstar) {
                    vstar[idx] = 0 + a * (unif(seed, c, 3) - 0.5);
                    vstar[n + idx] = 0 + a * (unif(seed, c, 4) - 0.5);
                    vstar[2 * n + idx] = 0 + a * (unif(seed, c, 5) - 0.5);
                }
                P[idx] = 0 + a * (unif(seed, c, 6) - 0.5);
                P[n + idx] = 0 + a * (unif(seed, c, 7) - 0.5);
                P[2 * n + idx] = 0 + a * (unif(seed, c, 8) - 0.5);
                if (pstar) {
                    pstar[idx] = 0 + a * (unif(seed, c, 9) - 0.5);
                    pstar[n + idx] = 0 + a * (unif(seed, c, 10) - 0.5);
                    pstar[2 * n + idx] = 0 + a * (unif(seed, c, 11) - 0.5);
                }
                const double dx = (i - p->Nx / 2.);
                const double dy = (j - p->Ny / 2.);
                const double dz = (k - p->Nz / 2.);
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                phi[idx] = (1 - std::tanh((r - p->Nx / 4.) * 2)) / 2;
            }
    return TTSH_OK;
}

int ttsh_init_droplet(int N, double R, double xi, int pmode, uint64_t seed, double* phi, double* P, double* v)
{
    if (!phi || !P || !v) return fail(TTSH_EINVAL, "null argument");
    if (N <= 0 || !(xi > 0.0) || !(R > 0.0)) return fail(TTSH_EINVAL, "need N > 0, R > 0, xi > 0");
    if (pmode != 0 && pmode != 1) return fail(TTSH_EINVAL, "pmode must be 0 (random unit) or 1 (aster)");
    const size_t n = (size_t)N * N * N;
    const double c0 = N / 2.0;
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const size_t idx = ((size_t)z * N + y) * N + x;
                const double dx = x - c0, dy = y - c0, dz = z - c0;
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                const double ph = 0.5 * (1.0 - std::tanh((r - R) / xi));
                phi[idx] = ph;
                v[idx] = v[n + idx] = v[2 * n + idx] = 0.0;
                double px = 0, py = 0, pz = 0;
                if (ph > 0.5) {
                    const double g0 = normal(seed, idx, 0), g1 = normal(seed, idx, 2), g2 = normal(seed, idx, 4);
                    if (pmode == 0) {
                        px = g0; py = g1; pz = g2;
                    } else {
                        const double inv = r > 0 ? 1.0 / r : 0.0;
                        px = dx * inv + 0.1 * g0;
                        py = dy * inv + 0.1 * g1;
                        pz = dz * inv + 0.1 * g2;
                    }
                    const double nrm = std::sqrt(px * px + py * py + pz * pz);
                    if (nrm > 0) { px /= nrm; py /= nrm; pz /= nrm; }
                }
                P[idx] = px;
                P[n + idx] = py;
                P[2 * n + idx] = pz;
            }
    return TTSH_OK;
}

}  // extern "C"

namespace ttsh {

// An FP32 value as std::fixed << setprecision(15) prints the double it
// converts to (printf "%.15f": the exact binary value rounded half to even
// at the 15th decimal).  The value is M * 2^e with M < 2^24, so
// |v| * 10^15 = M * 10^15 * 2^e is rounded exactly in 128-bit integers and
// printed digit by digit; NaN, infinities and |v| >= 2^40 go to snprintf.
static int fmt15(char* out, float v)
{
    uint32_t b;
    std::memcpy(&b, &v, 4);
    const int ex = (int)((b >> 23) & 0xff);
    if (ex == 0xff || ex >= 127 + 40) return std::snprintf(out, 64, "%.15f", (double)v);
    const uint64_t M = ex ? ((b & 0x7fffffu) | 0x800000u) : (b & 0x7fffffu);
    const int e = ex ? ex - 150 : -149;  // v = +-M * 2^e
    constexpr uint64_t P15 = 1000000000000000ull;
    unsigned __int128 q;
    if (e >= 0) {
        q = ((unsigned __int128)M * P15) << e;  // < 2^90 for e < 17
    } else {
        const unsigned __int128 num = (unsigned __int128)M * P15;  // < 2^74
        const int sh = -e;
        if (sh > 76) {
            q = 0;  // below half a unit of the 15th decimal
        } else {
            q = num >> sh;
            const unsigned __int128 rem = num - (q << sh), half = (unsigned __int128)1 << (sh - 1);
            if (rem > half || (rem == half && (q & 1))) ++q;
        }
    }
    uint64_t ip = (uint64_t)(q / P15), fp = (uint64_t)(q % P15);
    char* o = out;
    if (b >> 31) *o++ = '-';
    char buf[24];
    int nd = 0;
    do {
        buf[nd++] = (char)('0' + ip % 10);
        ip /= 10;
    } while (ip);
    while (nd) *o++ = buf[--nd];
    *o++ = '.';
    for (int d = 14; d >= 0; --d) {
        o[d] = (char)('0' + fp % 10);
        fp /= 10;
    }
    o += 15;
    *o = '\0';
    return (int)(o - out);
}

// Text of one ExpoConf row i (y): the N*N lines of x, then z, in order.
// Returns the row's length in out (the buffer keeps its size between rows).
static size_t format_row(std::string& out, int N, int expo_data, int i, const float* phi, const float* P,
                         const float* v)
{
    const size_t n = (size_t)N * N * N;
    // a value takes at most 64 bytes (snprintf's NaN/inf/large-value text);
    // the buffer grows only if a line might not fit
    constexpr size_t kLine = 7 * 64;
    if (out.size() < (size_t)N * N * 7 * 20 + kLine) out.resize((size_t)N * N * 7 * 20 + kLine);
    size_t pos = 0;
    for (int j = 0; j < N; ++j) {
        for (int k = 0; k < N; ++k) {
            if (out.size() - pos < kLine) out.resize(2 * out.size());
            char* o = &out[pos];
            const size_t idx = ((size_t)k * N + i) * N + j;
            if (expo_data == 1) {
                const float vals[7] = {v[idx], v[n + idx], v[2 * n + idx], P[idx], P[n + idx], P[2 * n + idx], phi[idx]};
                for (int q = 0; q < 7; ++q) {
                    o += fmt15(o, vals[q]);
                    *o++ = q < 6 ? ' ' : '\n';
                }
            } else if (expo_data == -1) {
                const float vals[2] = {v[idx], v[n + idx]};
                for (int q = 0; q < 2; ++q) {
                    o += fmt15(o, vals[q]);
                    *o++ = ' ';
                }
            }
            pos = (size_t)(o - &out[0]);
        }
    }
    return pos;
}

int write_conf_host(const char* path, int N, int expo_data, const float* phi, const float* P, const float* v)
{
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(TTSH_EIO, std::string("cannot write ") + path);
    // ExpoConf loop order (TTSH3D.cu:975-981): i < Ny (y), j < Nx (x), k < Nz
    // (z).  Rows i are formatted a group of nt at a time, by nt threads, into
    // one of two buffer sets; while this thread writes a group, the next one
    // is being formatted into the other set.
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::string> rows[2] = {std::vector<std::string>(nt), std::vector<std::string>(nt)};
    std::vector<size_t> len[2] = {std::vector<size_t>(nt, 0), std::vector<size_t>(nt, 0)};
    auto format_group = [&](int b, int i0) {
        const int cnt = std::min(nt, N - i0);
        std::vector<std::thread> th;
        for (int t = 1; t < cnt; ++t)
            th.emplace_back([&, b, t, i0] { len[b][t] = format_row(rows[b][t], N, expo_data, i0 + t, phi, P, v); });
        len[b][0] = format_row(rows[b][0], N, expo_data, i0, phi, P, v);
        for (auto& x : th) x.join();
    };
    format_group(0, 0);
    bool ok = true;
    for (int i0 = 0, b = 0; i0 < N; i0 += nt, b ^= 1) {
        std::thread next;
        if (i0 + nt < N) next = std::thread(format_group, b ^ 1, i0 + nt);
        const int cnt = std::min(nt, N - i0);
        for (int t = 0; t < cnt && ok; ++t)
            if (len[b][t] && std::fwrite(rows[b][t].data(), 1, len[b][t], f) != len[b][t]) ok = false;
        if (next.joinable()) next.join();
        if (!ok) break;
    }
    if (!ok) {
        std::fclose(f);
        return fail(TTSH_EIO, std::string("write failed: ") + path);
    }
    if (std::fclose(f) != 0) return fail(TTSH_EIO, std::string("close failed: ") + path);
    return TTSH_OK;
}

int write_conf_bin_host(const char* path, int N, int m, double T, int64_t steps, const float* const fields[7])
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(TTSH_EIO, std::string("cannot write ") + path);
    const size_t n = (size_t)N * N * N;
    const int32_t hdr_i[2] = {N, m};
    bool ok = std::fwrite("TTSH3DB1", 1, 8, f) == 8 && std::fwrite(hdr_i, sizeof(int32_t), 2, f) == 2 &&
              std::fwrite(&T, sizeof(double), 1, f) == 1 && std::fwrite(&steps, sizeof(int64_t), 1, f) == 1;
    for (int q = 0; q < 7 && ok; ++q) ok = std::fwrite(fields[q], sizeof(float), n, f) == n;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) return fail(TTSH_EIO, std::string("write failed: ") + path);
    return TTSH_OK;
}

// Level-contour radius R(theta) (characteristicVelocity.py:75-122 on the
// centroid slice).  The reference calls plt.contour(phi, levels=[0.5]) and
// takes the vertices of every contour path, in path order, with the centre
// at the mean of the vertices of the LAST path (:79-87).  The paths here are
// traced by marching squares with the conventions of matplotlib's default
// contour generator (contourpy's "mpl2014"), so the vertex list is the same:
//   - a point is above the level when its value is > level; an edge is
//     crossed when its ends differ, at x0 + (level - z0)/(z1 - z0);
//   - lines run with the values above the level on their left;
//   - lines that enter through the array border come first, found by
//     scanning quads row by row and, in a quad, its bottom, left, right and
//     top edges; then closed loops, found by scanning points row by row and,
//     at a point, the edge up before the edge to the right;
//   - a closed loop repeats its first vertex at its end;
//   - a saddle quad (corners alternating) joins its above-level corners when
//     the mean of its four corners is above the level.
// The theta scan then takes, for theta = -180, -177, ..., 177 degrees, the
// radius of the first vertex whose polar angle about the centre is nearest.
namespace {

struct ContourTracer {
    int nx, ny;
    const float* z;
    double level;
    std::vector<unsigned char> hi, vis_h, vis_v;  // point above level; edge visited
    std::vector<double> px, py;                    // vertices of all lines, in order
    size_t last_begin = 0;                         // first vertex of the last line

    double Z(int x, int y) const { return (double)z[(size_t)y * nx + x]; }
    bool H(int x, int y) const { return hi[(size_t)y * nx + x] != 0; }
    // edge: kind 0 = horizontal (x,y)-(x+1,y), 1 = vertical (x,y)-(x,y+1)
    bool crossed(int k, int x, int y) const { return k == 0 ? H(x, y) != H(x + 1, y) : H(x, y) != H(x, y + 1); }
    unsigned char& visited(int k, int x, int y) { return (k == 0 ? vis_h : vis_v)[(size_t)y * nx + x]; }
    void emit(int k, int x, int y)
    {
        if (k == 0) {
            const double z0 = Z(x, y), z1 = Z(x + 1, y);
            px.push_back(x + (level - z0) / (z1 - z0));
            py.push_back(y);
        } else {
            const double z0 = Z(x, y), z1 = Z(x, y + 1);
            px.push_back(x);
            py.push_back(y + (level - z0) / (z1 - z0));
        }
    }
    // the quad a line crossing edge e enters, with the above-level side on its left
    void forward(int k, int x, int y, int& qx, int& qy) const
    {
        if (k == 0) {
            qx = x;
            qy = H(x, y) ? y : y - 1;
        } else {
            qx = H(x, y + 1) ? x : x - 1;
            qy = y;
        }
    }
    bool inside(int qx, int qy) const { return qx >= 0 && qy >= 0 && qx < nx - 1 && qy < ny - 1; }
    // quad edges by side: 0 bottom, 1 top, 2 left, 3 right
    static void side(int qx, int qy, int s, int& k, int& x, int& y)
    {
        switch (s) {
        case 0: k = 0; x = qx; y = qy; break;
        case 1: k = 0; x = qx; y = qy + 1; break;
        case 2: k = 1; x = qx; y = qy; break;
        default: k = 1; x = qx + 1; y = qy; break;
        }
    }
    // exit side of quad q for a line entering through side s_in
    int exit_side(int qx, int qy, int s_in) const
    {
        int cr[4], n = 0;
        for (int s = 0; s < 4; ++s) {
            if (s == s_in) continue;
            int k, x, y;
            side(qx, qy, s, k, x, y);
            if (crossed(k, x, y)) cr[n++] = s;
        }
        if (n == 1) return cr[0];
        // saddle: turn to the side on the left or the right of the entry
        static const int left_of[4] = {2, 3, 1, 0}, right_of[4] = {3, 2, 0, 1};
        const double zmid = 0.25 * (Z(qx, qy) + Z(qx + 1, qy) + Z(qx, qy + 1) + Z(qx + 1, qy + 1));
        return zmid > level ? right_of[s_in] : left_of[s_in];
    }
    static int side_of(int qx, int qy, int k, int x, int y)
    {
        if (k == 0) return y == qy ? 0 : 1;
        return x == qx ? 2 : 3;
    }
    void follow(int k0, int x0, int y0)
    {
        last_begin = px.size();
        emit(k0, x0, y0);
        visited(k0, x0, y0) = 1;
        int k = k0, x = x0, y = y0, qx, qy;
        forward(k, x, y, qx, qy);
        while (inside(qx, qy)) {
            const int s = exit_side(qx, qy, side_of(qx, qy, k, x, y));
            side(qx, qy, s, k, x, y);
            if (k == k0 && x == x0 && y == y0) {  // back at the start: close the loop
                px.push_back(px[last_begin]);
                py.push_back(py[last_begin]);
                return;
            }
            visited(k, x, y) = 1;
            emit(k, x, y);
            forward(k, x, y, qx, qy);
        }
    }
    void run()
    {
        const size_t n = (size_t)nx * ny;
        hi.assign(n, 0);
        vis_h.assign(n, 0);
        vis_v.assign(n, 0);
        for (size_t q = 0; q < n; ++q) hi[q] = (double)z[q] > level;
        // lines entering through the border, by quad, sides bottom, left, right, top
        static const int order[4] = {0, 2, 3, 1};
        for (int qy = 0; qy < ny - 1; ++qy)
            for (int qx = 0; qx < nx - 1; ++qx)
                for (int o = 0; o < 4; ++o) {
                    int k, x, y;
                    side(qx, qy, order[o], k, x, y);
                    const bool border = k == 0 ? (y == 0 || y == ny - 1) : (x == 0 || x == nx - 1);
                    if (!border || visited(k, x, y) || !crossed(k, x, y)) continue;
                    int fx, fy;
                    forward(k, x, y, fx, fy);
                    if (fx == qx && fy == qy) follow(k, x, y);
                }
        // closed loops, by point, the edge up before the edge to the right
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (y + 1 < ny && !visited(1, x, y) && crossed(1, x, y)) follow(1, x, y);
                if (x + 1 < nx && !visited(0, x, y) && crossed(0, x, y)) follow(0, x, y);
            }
    }
};

}  // namespace

int contour_radius(int N, const float* field, double level, double* R, double* center)
{
    ContourTracer t;
    t.nx = t.ny = N;
    t.z = field;
    t.level = level;
    t.run();
    const std::vector<double>& px = t.px;
    const std::vector<double>& py = t.py;
    if (px.empty()) return fail(TTSH_EINVAL, "the level is not crossed");
    double cx = 0.0, cy = 0.0;
    for (size_t q = t.last_begin; q < px.size(); ++q) {
        cx += px[q];
        cy += py[q];
    }
    const double cnt = (double)(px.size() - t.last_begin);
    cx /= cnt;
    cy /= cnt;
    std::vector<double> pr(px.size()), pt(px.size());
    for (size_t q = 0; q < px.size(); ++q) {
        const double dx = px[q] - cx, dy = py[q] - cy;
        pr[q] = std::sqrt(dx * dx + dy * dy);
        pt[q] = std::atan2(dy, dx);
    }
    const double pi = 3.14159265358979323846;
    int a = 0;
    for (int deg = -180; deg < 180; deg += 3, ++a) {
        const double th = deg / 180.0 * pi;
        size_t best = 0;
        double bd = std::fabs(pt[0] - th);
        for (size_t q = 1; q < pt.size(); ++q) {
            const double d = std::fabs(pt[q] - th);
            if (d < bd) {
                bd = d;
                best = q;
            }
        }
        R[a] = pr[best];
    }
    if (center) {
        center[0] = cx;
        center[1] = cy;
    }
    return TTSH_OK;
}

}  // namespace ttsh

extern "C" int ttsh_contour_radius(int N, const float* field, double level, double* R, double* center)
{
    if (!field || !R || N < 2) return ttsh::fail(TTSH_EINVAL, "bad argument");
    return ttsh::contour_radius(N, field, level, R, center);
}

extern "C" int ttsh_write_conf_arrays(const char* path, int N, int ExpoData, const float* phi, const float* P,
                                      const float* v)
{
    if (!path || !phi || !P || !v || N <= 0) return ttsh::fail(TTSH_EINVAL, "bad argument");
    return ttsh::write_conf_host(path, N, ExpoData, phi, P, v);
}
