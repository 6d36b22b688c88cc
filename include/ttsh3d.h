/*
 * ttsh3d.h -- C ABI of the MI355X-native TTSH3D time step (libttsh3d.so).
 *
 * The reference (Sherlocked-L/Bacterial-Droplet, bacterial_droplet_12-16/
 * TTSH3D.cu) has no plugin or FFI layer: its step EEuler() works on process
 * globals and is reached through the process contract of run_cuda.pl
 * (input.dat in, conf_<m>.dat out).  This header is the boundary that
 * replaces both: every entry point below names the reference code it stands
 * in for.  Plain pointers and sizes only; no exceptions cross it.
 *
 * Conventions
 *   - Every function returns an int status: TTSH_OK (0) or a negative code;
 *     ttsh_last_error() gives a message for the calling thread.
 *   - Fields are host arrays of N^3 values, index (z*N + y)*N + x (x fastest,
 *     z slowest): the interior ordering of the reference (TTSH3D.cu:477).
 *     Vector fields are three such blocks (x, y, z components).
 *   - A ttsh_sim or ttsh_batch owns its device memory and one HIP stream.
 *     Handles are not thread-safe; use one handle per thread.
 *   - Grids are cubic (Nx = Ny = Nz = N, as the reference's cuFFT plan
 *     requires, TTSH3D.cu:1123) with N in {32, 36, 40, 48, 50, 60, 64, 72,
 *     80, 90, 96, 100, 120, 128, 144, 150, 160, 180, 192, 200, 240, 256}:
 *     the sizes with a compile-time FFT plan (even, prime factors 2, 3 and
 *     5, no factor 27 or 125).  The reference's default and swept N is 64
 *     (run_cuda.pl:14-16, forEach.pl:30); any other N its cuFFT plan accepts
 *     -- below 32, above 256, odd, or with a factor 7, 27 or 125 (e.g. 16, 54,
 *     70, 250) -- is refused here with TTSH_EUNSUPPORTED.
 */
#ifndef TTSH3D_H
#define TTSH3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TTSH_OK = 0,
    TTSH_EINVAL = -1,       /* bad argument or parameter */
    TTSH_EIO = -2,          /* file could not be read or written */
    TTSH_ESENTINEL = -3,    /* input.dat sentinel is not -8848 (TTSH3D.cu:270,748-751) */
    TTSH_EUNSUPPORTED = -4, /* TimeScheme != 1, non-cubic grid, unsupported N */
    TTSH_EHIP = -5,         /* HIP runtime error (no device, launch failure) */
    TTSH_ENOMEM = -6,       /* device allocation failed */
    TTSH_EBUDGET = -7,      /* request exceeds the batch's HBM budget */
    TTSH_ENONFINITE = -8,   /* a reduction found NaN/Inf: the droplet diverged */
    TTSH_EFULL = -9         /* no free batch slot */
};

enum { TTSH_F32 = 0, TTSH_F64 = 1 };

/* The 33 positional values of input.dat in file order (read by GetInput,
 * TTSH3D.cu:702-739; written by run_cuda.pl:54-86), then build options that
 * are not part of input.dat. */
typedef struct ttsh_params {
    int32_t TimeScheme, RKStage, ExplImpi, AdapTime, ExpoData, InitCond, isSimplified, checkUnStable;
    int32_t Nx, Ny, Nz, Nb;
    double h, dt0, T0, Ts, dte, dtVarC, dtMax;
    double mu, cA, alpha, beta, eta, La, kappa, lambda0, aPhi, kPhi, gamma, A, init0;
    int32_t ReadSuccess;
    /* --- options --- */
    int32_t faithful_halo; /* 1 (default): div(v*), div(P*) read the previous step's
                              v*, P* across the periodic boundary, as the reference's
                              stale halo does (TTSH3D.cu:366-371,405-407,665-667) */
    int32_t rhs_variant;   /* K_rhs tile: -1 = tuned default, else 0..6 */
    int32_t rhs_tz;        /* K_rhs z-range per work item, 0 = tuned default */
    int32_t corr_tz;       /* divergence and corrector z-range per work item,
                              0 = tuned default.  With the defaults, a batch
                              whose droplets fill the device runs K_rhs with
                              the 32x8 tile and ranges of at least 32, and the
                              divergence and corrector with ranges of about 32;
                              results do not change. */
} ttsh_params;

typedef struct ttsh_sim ttsh_sim;
typedef struct ttsh_batch ttsh_batch;

/* ---- parameters and files ------------------------------------------------ */

/* Defaults of run_cuda.pl:7-44 (dt0 0.002, Ts 1000, dte 10, N 64, mu 1, cA -1,
 * ...), ReadSuccess = -8848, faithful_halo = 1. */
int ttsh_params_default(ttsh_params* p);

/* Replaces GetInput (TTSH3D.cu:696-752): reads the 33 whitespace-separated
 * values of input.dat.  Returns TTSH_ESENTINEL when the last value is not
 * -8848 and TTSH_EIO when the file is missing or short.  Options keep the
 * values already in *p (call ttsh_params_default first). */
int ttsh_parse_input_dat(const char* path, ttsh_params* p);

/* Replaces the input.dat writer of run_cuda.pl:53-87 (one value per line). */
int ttsh_write_input_dat(const char* path, const ttsh_params* p);

/* Checks what the step supports: TimeScheme == 1 (TTSH3D.cu:338-341),
 * Nx == Ny == Nz, supported N, Nb >= 1, h > 0, dt0 > 0. */
int ttsh_validate_params(const ttsh_params* p);

/* ---- one droplet ------------------------------------------------------------ */

/* Replaces MemAlloc + initPoissonSolver (TTSH3D.cu:1127-1272, :1088-1124):
 * allocates the droplet's device state on HIP device `device` and its stream.
 * The state starts at zero. */
int ttsh_create(const ttsh_params* p, int device, ttsh_sim** out);

/* Replaces MemFree + cufftDestroy (TTSH3D.cu:1275-1419, :280). */
int ttsh_destroy(ttsh_sim* s);

/* Replaces the host-to-device copies of InitConf (TTSH3D.cu:834-893).
 * dtype TTSH_F32 or TTSH_F64.  phi: N^3; P, v: 3*N^3.  vstar_prev and
 * pstar_prev (3*N^3 each, may be NULL = v and P) are the predictor values the
 * first step's divergence reads across the boundary (the reference's initial
 * v*, P*, TTSH3D.cu:810-818).  Resets the step counter and T to T0. */
int ttsh_set_state(ttsh_sim* s, int dtype, const void* phi, const void* P, const void* v,
                   const void* vstar_prev, const void* pstar_prev);

/* Replaces the device-to-host copies of ExpoConf (TTSH3D.cu:904-963).  Any
 * pointer may be NULL.  Synchronises the stream. */
int ttsh_get_state(ttsh_sim* s, int dtype, void* phi, void* P, void* v);

/* Replaces n calls of EEuler (TTSH3D.cu:347-359) with T += Dt after each
 * (TTSH3D.cu:336-343).  Asynchronous on the handle's stream. */
int ttsh_step(ttsh_sim* s, int64_t n);

/* Waits for the handle's stream. */
int ttsh_synchronize(ttsh_sim* s);

/* New (no reference counterpart): FP64 reductions of the current state,
 * out[0] = sum(phi) h^3, out[1] = mean |P|, out[2] = 1/2 sum |v|^2 h^3.
 * Returns TTSH_ENONFINITE (out still written) when any is NaN or Inf. */
int ttsh_reductions(ttsh_sim* s, double out[3]);

/* Replaces ExpoConf's file output (TTSH3D.cu:966-1013): one line per cell in
 * y, x, z loop order, "vx vy vz px py pz phi" at fixed 15 decimals
 * (ExpoData == 1).  ExpoData == -1 writes "vx vy " per cell with no newline,
 * as the reference does (:1005-1009). */
int ttsh_write_conf(ttsh_sim* s, const char* path);

/* The same writer for host arrays (FP32, layout as above); no device needed. */
int ttsh_write_conf_arrays(const char* path, int N, int ExpoData, const float* phi, const float* P,
                           const float* v);

/* Replaces evolve()/run() (TTSH3D.cu:290-343): writes <run_dir>/conf_0.dat if
 * T0 == 0, then steps while T <= Ts, accumulating T += Dt in double and
 * writing <run_dir>/conf_<floor(T/dte)>.dat whenever floor(T/dte) increases
 * (at least one step is taken).  *steps_done (may be NULL) receives the
 * number of steps taken. */
int ttsh_run(ttsh_sim* s, const char* run_dir, int64_t* steps_done);

/* New (the reference only prints host CPU time, TTSH3D.cu:266,284-286): runs n
 * steps like ttsh_step, each kernel kind's first launch recording a HIP
 * event at its start and its last launch one at its end, from the kernel
 * dispatch itself (hipExtLaunchKernel: no marker between kernels), on the
 * handle's stream; kernel_ms[k] receives the summed milliseconds of kernel
 * kind k (8 kinds, names from ttsh_kernel_name: K_rhs, K_div (with the
 * per-droplet reduction of its sums that follows it), the five FFT passes,
 * K_corr).  Synchronises once per 256 steps. */
int ttsh_profile(ttsh_sim* s, int64_t n, double kernel_ms[8]);

/* Name of kernel kind k of ttsh_profile ("" outside 0..7). */
const char* ttsh_kernel_name(int k);

/* Steps taken since the last ttsh_set_state, and the time T. */
int ttsh_get_clock(ttsh_sim* s, int64_t* steps, double* T);

/* Debug/verification access to the last step's intermediates (FP32):
 * which = 0: v* (3*N^3), 1: P* (3*N^3), 2: the pressures p_v + i p_p
 * (2*N^3 interleaved re/im) that the last corrector used. */
int ttsh_get_aux(ttsh_sim* s, int which, float* out);

/* ---- diagnostics (replace the post-processing reads of full snapshots) ---- */

/* New: the phase-field centroid as readDraw.py:30-52 and
 * characteristicVelocity.py:16-39 compute it from a snapshot: the sums of
 * phi*i, phi*j, phi*k over the snapshot array axes (i, j, k) = (y, x, z),
 * divided by sum phi.  idx[] holds them truncated toward zero (Python int()),
 * c[] before truncation, mass = sum phi.  FP64 sums on the device. */
typedef struct {
    int32_t idx[3]; /* (y, x, z) */
    double c[3];    /* (y, x, z) */
    double mass;
} ttsh_centroid;
int ttsh_get_centroid(ttsh_sim* s, ttsh_centroid* out);

/* New: plane z (z < 0: the centroid plane idx[2], readDraw.py:220-235) of
 * phi, vx, vy, vz, px, py, pz and the vorticities d_x vy - d_y vx and
 * d_x py - d_y px (readDraw.py:236-237) with the 4th-order periodic
 * differences of op.py:3-13.  out receives 9*N*N floats, field-major, each
 * field [y][x]; *z_used (may be NULL) the plane. */
int ttsh_get_slice(ttsh_sim* s, int z, float* out, int* z_used);

/* New, host only: radius R(theta) of the level contour of an N x N field,
 * as characteristicVelocity.py:75-122 extracts it from the centroid slice:
 * the vertices of every path of plt.contour(field, levels=[level]) in path
 * order (traced here with the conventions of matplotlib's default contour
 * generator, mpl2014: linear interpolation on grid edges, no periodic wrap,
 * lines through the border first, closed loops repeating their first
 * vertex), the centre at the mean of the vertices of the last path (:85),
 * and for theta = -180, -177, ..., 177 degrees the distance of the vertex
 * whose polar angle about the centre is nearest (the first on ties).
 * field is [row][col] with x = col, y = row; R receives 120 values, center
 * (may be NULL) the centre (x, y).  TTSH_EINVAL if the level is never
 * crossed. */
int ttsh_contour_radius(int N, const float* field, double level, double* R, double* center);

/* Output options of ttsh_run_ex (bitwise or). */
#define TTSH_RUN_TEXT 1   /* conf_<m>.dat, the reference's text format */
#define TTSH_RUN_BINARY 2 /* conf_<m>.bin: "TTSH3DB1", int32 N, int32 m, double T,
                             int64 steps, then vx vy vz px py pz phi as FP32 N^3
                             each, x fastest */
#define TTSH_RUN_DIAG 4   /* per export: a line "m T iy ix iz cy cx cz mass" in
                             diag.dat and the 120 radii of the 0.5 contour of the
                             centroid slice of phi in a_R.txt (np.savetxt layout,
                             characteristicVelocity.py:125-135) */
#define TTSH_RUN_SYNC 8   /* finish each export before stepping on */

/* New: ttsh_run with a choice of outputs.  Exports are written by a host
 * thread from a device snapshot copied out on a second stream, so stepping
 * continues while a snapshot is formatted and written (two snapshots in
 * flight at most); ttsh_run_ex returns once every file is written.
 * ttsh_run(s, d, n) is ttsh_run_ex(s, d, TTSH_RUN_TEXT, n). */
int ttsh_run_ex(ttsh_sim* s, const char* run_dir, int flags, int64_t* steps_done);

/* ---- initial conditions (host, FP64, seeded and reproducible) --------------- */

/* InitCond == 1 of InitConf (TTSH3D.cu:798-825): uniform noise of amplitude
 * init0 in v, v*, P, P* (and p, which the step overwrites) and the droplet
 * phi = (1 - tanh((r - Nx/4)*2))/2 centred at (Nx/2, Ny/2, Nz/2).  The
 * reference seeds std::default_random_engine with time(0) (TTSH3D.cu:40-41);
 * here the noise comes from a counter-based SplitMix64 stream of `seed`.
 * vstar, pstar may be NULL. */
int ttsh_init_reference(const ttsh_params* p, uint64_t seed, double* phi, double* P, double* v,
                        double* vstar, double* pstar);

/* The droplet initial conditions of the benchmark workloads:
 * phi = (1 - tanh((r - R)/xi))/2 about the box centre (N/2, N/2, N/2), v = 0,
 * and P inside phi > 0.5 (zero outside):
 *   pmode 0: a unit vector uniform on the sphere (three normals, normalised);
 *   pmode 1: aster, r_hat + 0.1 N(0,1) per component, renormalised. */
int ttsh_init_droplet(int N, double R, double xi, int pmode, uint64_t seed, double* phi, double* P,
                      double* v);

/* ---- batches of droplets ------------------------------------------------------ */
/* New (the reference runs one droplet per process; forEach.pl:33-133 sweeps
 * serially).  A batch steps all its active droplets together: one launch per
 * kernel covers every droplet, whatever its N. */

/* Device bytes one droplet of size N needs in a batch. */
size_t ttsh_droplet_bytes(int N);

/* Create a batch on `device` that may hold up to max_slots droplets and
 * hbm_budget bytes of droplet state (0 = no budget).  use_graph: capture the
 * step sequence in a HIP graph and replay it. */
int ttsh_batch_create(int device, int max_slots, size_t hbm_budget, int use_graph, ttsh_batch** out);
int ttsh_batch_destroy(ttsh_batch* b);

/* Admit a droplet with parameters p and state (dtype as ttsh_set_state;
 * vstar_prev/pstar_prev may be NULL).  Returns TTSH_EBUDGET when its memory
 * does not fit the remaining budget and TTSH_EFULL when no slot is free.
 * *slot receives its slot id. */
int ttsh_batch_admit(ttsh_batch* b, const ttsh_params* p, int dtype, const void* phi, const void* P,
                     const void* v, const void* vstar_prev, const void* pstar_prev, int* slot);

/* Step every active droplet by n steps; a droplet whose step limit (see
 * ttsh_batch_set_limit) is reached stops there. */
int ttsh_batch_step(ttsh_batch* b, int64_t n);

/* Limit the total steps of a droplet (counted from its admission); < 0 = none. */
int ttsh_batch_set_limit(ttsh_batch* b, int slot, int64_t max_steps);

/* Steps done by a droplet since admission. */
int ttsh_batch_steps_done(ttsh_batch* b, int slot, int64_t* steps);

/* Copy a droplet's state out (synchronises). */
int ttsh_batch_get_state(ttsh_batch* b, int slot, int dtype, void* phi, void* P, void* v);

/* FP64 reductions of one droplet (as ttsh_reductions). */
int ttsh_batch_reductions(ttsh_batch* b, int slot, double out[3]);

/* Free a droplet's slot and memory. */
int ttsh_batch_retire(ttsh_batch* b, int slot);

/* Bytes of the budget in use, and number of active droplets. */
int ttsh_batch_usage(ttsh_batch* b, size_t* bytes_used, int* active);

int ttsh_batch_synchronize(ttsh_batch* b);

/* New: wait until every ttsh_batch_step call but the last `keep` (0 <= keep
 * < 16) is done on the device; keep = 0 is ttsh_batch_synchronize.  A
 * request loop queues the next steps and then waits for the previous ones,
 * so the device never idles while the host retires and admits droplets. */
int ttsh_batch_wait(ttsh_batch* b, int keep);

/* New: *ready = 1 when the droplet has reached its step limit and those
 * steps are done on the device.  Its state and reductions are then read
 * after its own last step, without waiting for steps queued later (which do
 * not touch it).  Admission likewise uploads on a second stream without
 * waiting for queued steps; the droplet joins them at the next step call,
 * and a retired droplet's memory is kept for the next droplet of its size. */
int ttsh_batch_ready(ttsh_batch* b, int slot, int* ready);

/* ---- stepping without allocation ----------------------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: steps one droplet whose state lives in caller-owned
 * device memory, and allocates nothing.  All pointers are device pointers to
 * FP32 arrays in the layouts above: phi (N^3), P and v (3*N^3) in/out;
 * vstar_prev, pstar_prev (3*N^3 each) in: the predictors the first step's
 * stale-halo divergence reads (the reference's initial v*, P*; pass copies
 * of v and P for a fresh state), out: the last step's predictors, so
 * consecutive calls continue one run exactly.  n steps as ttsh_step with
 * the same kernels, tiles and results; stream is a hipStream_t (NULL: the
 * null stream); the call returns when the steps are done.
 * Every other buffer of the step lives in one device array of 32 MiB per
 * GPU, a library-wide static held until the process ends: N is limited to
 * 64 (TTSH_EINVAL beyond).  Calls on one GPU share that array and keep the
 * last call's contents (the droplet's region is cleared at every call).
 * Calls that overlapped on one GPU would race on it, so the library
 * serialises them per device: a call waits for the previous one to return,
 * and a call returns only when its steps are done. */
int ttsh_step_noalloc(const ttsh_params* p, float* phi, float* P, float* v, float* vstar_prev,
                      float* pstar_prev, int64_t n, void* stream);

/* ---- misc ------------------------------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char* ttsh_last_error(void);

/* Library version string. */
const char* ttsh_version(void);

/* New: the tuning table the autotuner writes (lines "N rhs_variant rhs_tz
 * corr_tz", and a "batch" line, below), read from $TTSH_TUNE_FILE or
 * tuned.txt beside libttsh3d.so and applied when ttsh_params leaves
 * rhs_variant = -1, rhs_tz = 0, corr_tz = 0.  reload: re-read from path
 * (NULL = the default location). */
int ttsh_tuning_reload(const char* path);
int ttsh_tuning_lookup(int N, int* rhs_variant, int* rhs_tz, int* corr_tz);
/* New: the layout of a batch whose droplets fill the device (table line
 * "batch rhs_variant rhs_tz corr_tz"): the K_rhs tile every droplet then
 * runs, the minimum K_rhs z-range per work item and the divergence and
 * corrector z-range.  Without the line the engine uses 32x8, 32 and 32.  A
 * reload applies to running batches at their next step call. */
int ttsh_tuning_lookup_batch(int* rhs_variant, int* rhs_tz, int* corr_tz);

#ifdef __cplusplus
}
#endif

#endif /* TTSH3D_H */
