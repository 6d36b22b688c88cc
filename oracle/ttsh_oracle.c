/*
 * ttsh_oracle.c -- FP64 CPU restatement of the TTSH3D explicit-Euler time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker.  The product (libttsh3d.so) never
 * links, loads or calls it.
 *
 * It is a plain transcription of the reference step in bacterial_droplet_12-16/
 * TTSH3D.cu (cited below as TTSH3D.cu:N), in FP64, on an unpadded periodic grid
 * (periodic index wrap replaces the Nb-wide halo copies of BounPeriU,
 * TTSH3D.cu:1422-1607).  Expressions follow the source's evaluation order so
 * that rounding matches the reference's FP64 arithmetic up to the FFT.
 *
 * Reference quirks that are reproduced on purpose:
 *   - integer-division zeros (TTSH3D.cu:579-602): 1/2*(..), 1/28*lap(..) and
 *     1/3*dPhi2 are evaluated literally, so DP=OP=0, Sigma=phi*G, A=DphiDphi;
 *   - stale halo of v* and P* (TTSH3D.cu:349-356,366-371,405-407,665-667):
 *     with faithful=1, div(v*) and div(P*) read, at any neighbour that lies
 *     across the periodic boundary, the v* and P* of the previous step;
 *   - k2[0]=1 (TTSH3D.cu:1120): the k=0 mode of p is -s_hat(0).
 *
 * Layout: idx = (z*N + y)*N + x  (x fastest, z slowest), as the reference's
 * interior ordering (TTSH3D.cu:477 without Nb).  Cubic grids only (the cuFFT
 * plan of TTSH3D.cu:1123 is only right for Nx=Ny=Nz).
 *
 * FFT: self-contained mixed-radix Stockham (radix 4,2,3,5, generic), FP64,
 * unnormalised, forward e^{-2 pi i jk/N}, inverse e^{+2 pi i jk/N} -- the DFT
 * definition that cufftExecZ2Z(FORWARD/INVERSE) computes (TTSH3D.cu:638,640).
 *
 * Parity pins (see tests/): stencil Fourier symbols, brute-force DFT and numpy
 * fixtures, single-mode projection formula, uniform-state closed forms,
 * shift/permutation invariance, kappa independence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int N;
    double h, dt0;
    double mu, cA, alpha, beta, eta, La, kappa, lambda0, aPhi, kPhi, gamma, A;
    int faithful;
} orc_params;

typedef struct {
    orc_params p;
    size_t n;
    /* state */
    double *phi, *px, *py, *pz, *vx, *vy, *vz;
    /* predictor (current) and previous-step predictor (stale halo source) */
    double *vxs, *vys, *vzs, *pxs, *pys, *pzs;
    double *vxs_o, *vys_o, *vzs_o, *pxs_o, *pys_o, *pzs_o;
    /* temporaries of getTempVari_1/_2 and getDtF */
    double *G[9]; /* dxPx dyPx dzPx dxPy dyPy dzPy dxPz dyPz dzPz */
    double *dPhi2, *mu0, *P2, *mux, *muy, *muz;
    double *DPx, *DPy, *DPz, *OPx, *OPy, *OPz;
    double *Sig[9]; /* Sigxx Sigxy Sigxz Sigyx Sigyy Sigyz Sigzx Sigzy Sigzz */
    double *Axx, *Axy, *Axz, *Ayy, *Ayz, *Azz;
    double *dtvx, *dtvy, *dtvz, *dtpx, *dtpy, *dtpz, *dtphi;
    double *p_v, *p_p;
    double *fre, *fim; /* Fstream */
    double *sdiv_v, *sdiv_p; /* getStream type-0 output (div/dt0) before the FFT */
    double *k2;
} orc_state;

#define IDX(N, x, y, z) ((((size_t)(z)) * (size_t)(N) + (size_t)(y)) * (size_t)(N) + (size_t)(x))

/* --------------------------------------------------------------------------
 * Neighbourhood: nb[(dz+1)*9+(dy+1)*3+(dx+1)] is the index of (x+dx,y+dy,z+dz)
 * with periodic wrap; bit of the same position in *wrap is set when that
 * neighbour lies across the boundary (it is a halo cell in the reference).
 * ------------------------------------------------------------------------ */
static inline void neighbourhood(int N, int x, int y, int z, size_t nb[27], uint32_t *wrap)
{
    uint32_t m = 0;
    for (int dz = -1; dz <= 1; dz++) {
        int zz = z + dz, wz = (zz < 0 || zz >= N);
        zz = (zz + N) % N;
        for (int dy = -1; dy <= 1; dy++) {
            int yy = y + dy, wy = (yy < 0 || yy >= N);
            yy = (yy + N) % N;
            for (int dx = -1; dx <= 1; dx++) {
                int xx = x + dx, wx = (xx < 0 || xx >= N);
                xx = (xx + N) % N;
                int pos = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                nb[pos] = IDX(N, xx, yy, zz);
                if (wx | wy | wz) m |= (1u << pos);
            }
        }
    }
    if (wrap) *wrap = m;
}

/* F(dx,dy,dz): value of f at the neighbour.  The reference names x as j (dj=1),
 * y as i (di=Nx+2Nb) and z as k (dk), TTSH3D.cu:3283-3285. */
#define NBP(dx, dy, dz) (((dz) + 1) * 9 + ((dy) + 1) * 3 + ((dx) + 1))
#define F(dx, dy, dz) f[nb[NBP(dx, dy, dz)]]

/* d1xCO2I3D, TTSH3D.cu:3280-3294 */
static inline double d1x(const double *f, const size_t *nb, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (F(1, 0, 0) - F(-1, 0, 0))
        + 4.0 / 36 * (F(1, 0, 1) + F(1, 0, -1) + F(1, 1, 0) + F(1, -1, 0)
                      - F(-1, 0, 1) - F(-1, 0, -1) - F(-1, 1, 0) - F(-1, -1, 0))
        + 1.0 / 36 * (F(1, 1, 1) + F(1, -1, 1) + F(1, 1, -1) + F(1, -1, -1)
                      - F(-1, 1, 1) - F(-1, -1, 1) - F(-1, 1, -1) - F(-1, -1, -1)));
}

/* d1yCO2I3D, TTSH3D.cu:3296-3311 */
static inline double d1y(const double *f, const size_t *nb, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (F(0, 1, 0) - F(0, -1, 0))
        + 4.0 / 36 * (F(0, 1, 1) + F(0, 1, -1) + F(1, 1, 0) + F(-1, 1, 0)
                      - F(0, -1, 1) - F(0, -1, -1) - F(1, -1, 0) - F(-1, -1, 0))
        + 1.0 / 36 * (F(1, 1, 1) + F(-1, 1, 1) + F(1, 1, -1) + F(-1, 1, -1)
                      - F(1, -1, 1) - F(-1, -1, 1) - F(1, -1, -1) - F(-1, -1, -1)));
}

/* d1zCO2I3D, TTSH3D.cu:3313-3328 */
static inline double d1z(const double *f, const size_t *nb, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (F(0, 0, 1) - F(0, 0, -1))
        + 4.0 / 36 * (F(1, 0, 1) + F(-1, 0, 1) + F(0, 1, 1) + F(0, -1, 1)
                      - F(1, 0, -1) - F(-1, 0, -1) - F(0, 1, -1) - F(0, -1, -1))
        + 1.0 / 36 * (F(1, 1, 1) + F(1, -1, 1) + F(-1, 1, 1) + F(-1, -1, 1)
                      - F(1, 1, -1) - F(1, -1, -1) - F(-1, 1, -1) - F(-1, -1, -1)));
}

/* laplaceCO2I3D, TTSH3D.cu:3331-3343 */
static inline double lap(const double *f, const size_t *nb, double h)
{
    return 1.0 / (12 * h * h) * (F(1, 1, -1) + F(1, -1, -1) + F(-1, 1, -1) + F(-1, -1, -1)
                                 + F(1, 1, 1) + F(1, -1, 1) + F(-1, 1, 1) + F(-1, -1, 1))
        + 2.0 / (3 * h * h) * (F(0, 0, -1) + F(0, 0, 1) + F(-1, 0, 0) + F(1, 0, 0) + F(0, -1, 0) + F(0, 1, 0))
        - 14.0 / (3 * h * h) * F(0, 0, 0);
}
#undef F

/* Stale-halo variants for getStream type 0 (TTSH3D.cu:665-667): neighbours
 * whose wrap bit is set read the previous-step array g instead of f. */
#define G2(dx, dy, dz) ((wrap >> NBP(dx, dy, dz)) & 1u ? g[nb[NBP(dx, dy, dz)]] : f[nb[NBP(dx, dy, dz)]])
static inline double d1x_h(const double *f, const double *g, const size_t *nb, uint32_t wrap, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (G2(1, 0, 0) - G2(-1, 0, 0))
        + 4.0 / 36 * (G2(1, 0, 1) + G2(1, 0, -1) + G2(1, 1, 0) + G2(1, -1, 0)
                      - G2(-1, 0, 1) - G2(-1, 0, -1) - G2(-1, 1, 0) - G2(-1, -1, 0))
        + 1.0 / 36 * (G2(1, 1, 1) + G2(1, -1, 1) + G2(1, 1, -1) + G2(1, -1, -1)
                      - G2(-1, 1, 1) - G2(-1, -1, 1) - G2(-1, 1, -1) - G2(-1, -1, -1)));
}
static inline double d1y_h(const double *f, const double *g, const size_t *nb, uint32_t wrap, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (G2(0, 1, 0) - G2(0, -1, 0))
        + 4.0 / 36 * (G2(0, 1, 1) + G2(0, 1, -1) + G2(1, 1, 0) + G2(-1, 1, 0)
                      - G2(0, -1, 1) - G2(0, -1, -1) - G2(1, -1, 0) - G2(-1, -1, 0))
        + 1.0 / 36 * (G2(1, 1, 1) + G2(-1, 1, 1) + G2(1, 1, -1) + G2(-1, 1, -1)
                      - G2(1, -1, 1) - G2(-1, -1, 1) - G2(1, -1, -1) - G2(-1, -1, -1)));
}
static inline double d1z_h(const double *f, const double *g, const size_t *nb, uint32_t wrap, double h)
{
    return 1.0 / (2.0 * h) * (
        16.0 / 36 * (G2(0, 0, 1) - G2(0, 0, -1))
        + 4.0 / 36 * (G2(1, 0, 1) + G2(-1, 0, 1) + G2(0, 1, 1) + G2(0, -1, 1)
                      - G2(1, 0, -1) - G2(-1, 0, -1) - G2(0, 1, -1) - G2(0, -1, -1))
        + 1.0 / 36 * (G2(1, 1, 1) + G2(1, -1, 1) + G2(-1, 1, 1) + G2(-1, -1, 1)
                      - G2(1, 1, -1) - G2(1, -1, -1) - G2(-1, 1, -1) - G2(-1, -1, -1)));
}
#undef G2

/* ==========================================================================
 * FFT: Stockham autosort, mixed radix (Stockham 1966; Cochran et al. 1967;
 * Van Loan 1992), FP64.  sign=-1 forward, +1 inverse; unnormalised.
 * ======================================================================== */
typedef struct {
    int N, nst, rad[32];
    double *wr, *wi; /* exp(-2 pi i q/N), q in [0,N) */
} orc_fftplan;

static void fftplan_init(orc_fftplan *pl, int N)
{
    pl->N = N;
    pl->nst = 0;
    int m = N;
    while (m % 4 == 0) { pl->rad[pl->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->rad[pl->nst++] = 2; m /= 2; }
    while (m % 3 == 0) { pl->rad[pl->nst++] = 3; m /= 3; }
    while (m % 5 == 0) { pl->rad[pl->nst++] = 5; m /= 5; }
    if (m > 1) pl->rad[pl->nst++] = m; /* generic O(r^2) DFT for a leftover factor */
    pl->wr = (double *)malloc(sizeof(double) * N);
    pl->wi = (double *)malloc(sizeof(double) * N);
    const double twopi = 6.283185307179586476925286766559;
    for (int q = 0; q < N; q++) {
        double a = twopi * (double)q / (double)N;
        pl->wr[q] = cos(a);
        pl->wi[q] = -sin(a);
    }
}

static void fftplan_free(orc_fftplan *pl)
{
    free(pl->wr);
    free(pl->wi);
}

/* One 1D transform of a contiguous line (xr,xi) using scratch (yr,yi). */
static void fft1d_line(const orc_fftplan *pl, double *xr, double *xi, double *yr, double *yi, int sign)
{
    const int N = pl->N;
    double *ar = xr, *ai = xi, *br = yr, *bi = yi;
    int Ns = 1;
    for (int s = 0; s < pl->nst; s++) {
        const int R = pl->rad[s];
        const int NR = N / R;
        const int M = Ns * R;       /* length of the sub-transforms after this stage */
        const int tstep = N / M;    /* table stride for W_M */
        double vr[64], vi[64], ur[64], ui[64];
        for (int j = 0; j < NR; j++) {
            const int k = j % Ns;
            for (int r = 0; r < R; r++) {
                double xr_ = ar[j + r * NR], xi_ = ai[j + r * NR];
                int q = (int)(((long)k * r * tstep) % N);
                double wr = pl->wr[q], wi = sign < 0 ? pl->wi[q] : -pl->wi[q];
                vr[r] = xr_ * wr - xi_ * wi;
                vi[r] = xr_ * wi + xi_ * wr;
            }
            /* radix-R DFT of v */
            if (R == 2) {
                ur[0] = vr[0] + vr[1]; ui[0] = vi[0] + vi[1];
                ur[1] = vr[0] - vr[1]; ui[1] = vi[0] - vi[1];
            } else if (R == 4) {
                double s0r = vr[0] + vr[2], s0i = vi[0] + vi[2];
                double d0r = vr[0] - vr[2], d0i = vi[0] - vi[2];
                double s1r = vr[1] + vr[3], s1i = vi[1] + vi[3];
                double d1r = vr[1] - vr[3], d1i = vi[1] - vi[3];
                /* multiply d1 by sign*i */
                double jr = sign < 0 ? d1i : -d1i, ji = sign < 0 ? -d1r : d1r;
                ur[0] = s0r + s1r; ui[0] = s0i + s1i;
                ur[2] = s0r - s1r; ui[2] = s0i - s1i;
                ur[1] = d0r + jr;  ui[1] = d0i + ji;
                ur[3] = d0r - jr;  ui[3] = d0i - ji;
            } else {
                for (int o = 0; o < R; o++) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; r++) {
                        int q = (int)(((long)o * r % R) * (N / R));
                        double wr = pl->wr[q], wi = sign < 0 ? pl->wi[q] : -pl->wi[q];
                        sr += vr[r] * wr - vi[r] * wi;
                        si += vr[r] * wi + vi[r] * wr;
                    }
                    ur[o] = sr; ui[o] = si;
                }
            }
            const int idxD = (j / Ns) * M + k;
            for (int r = 0; r < R; r++) {
                br[idxD + r * Ns] = ur[r];
                bi[idxD + r * Ns] = ui[r];
            }
        }
        double *t;
        t = ar; ar = br; br = t;
        t = ai; ai = bi; bi = t;
        Ns = M;
    }
    if (ar != xr) {
        memcpy(xr, ar, sizeof(double) * N);
        memcpy(xi, ai, sizeof(double) * N);
    }
}

/* Public 1D FFT of one line. */
void orc_fft1d(int N, double *re, double *im, int sign)
{
    orc_fftplan pl;
    fftplan_init(&pl, N);
    double *yr = (double *)malloc(sizeof(double) * N), *yi = (double *)malloc(sizeof(double) * N);
    fft1d_line(&pl, re, im, yr, yi, sign);
    free(yr);
    free(yi);
    fftplan_free(&pl);
}

/* 3D FFT of an N^3 complex array, layout (z*N+y)*N+x; all three axes. */
static void fft3d_plan(const orc_fftplan *pl, double *re, double *im, int sign)
{
    const int N = pl->N;
    for (int axis = 0; axis < 3; axis++) {
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)N : (size_t)N * N);
#pragma omp parallel
        {
            double *lr = (double *)malloc(sizeof(double) * 4 * N);
            double *li = lr + N, *sr = lr + 2 * N, *si = lr + 3 * N;
#pragma omp for schedule(static)
            for (long line = 0; line < (long)N * N; line++) {
                int a = (int)(line % N), b = (int)(line / N);
                size_t base;
                if (axis == 0) base = IDX(N, 0, a, b);
                else if (axis == 1) base = IDX(N, a, 0, b);
                else base = IDX(N, a, b, 0);
                for (int t = 0; t < N; t++) { lr[t] = re[base + t * stride]; li[t] = im[base + t * stride]; }
                fft1d_line(pl, lr, li, sr, si, sign);
                for (int t = 0; t < N; t++) { re[base + t * stride] = lr[t]; im[base + t * stride] = li[t]; }
            }
            free(lr);
        }
    }
}

void orc_fft3d(int N, double *re, double *im, int sign)
{
    orc_fftplan pl;
    fftplan_init(&pl, N);
    fft3d_plan(&pl, re, im, sign);
    fftplan_free(&pl);
}

/* ==========================================================================
 * Stencil application on a whole field (for the Fourier-symbol KATs).
 * which: 0 d1x, 1 d1y, 2 d1z, 3 laplace.
 * ======================================================================== */
void orc_stencil(int which, int N, double h, const double *f, double *out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                double r;
                switch (which) {
                case 0: r = d1x(f, nb, h); break;
                case 1: r = d1y(f, nb, h); break;
                case 2: r = d1z(f, nb, h); break;
                default: r = lap(f, nb, h); break;
                }
                out[IDX(N, x, y, z)] = r;
            }
        }
}

/* ==========================================================================
 * State management
 * ======================================================================== */
static double *zalloc(size_t n)
{
    double *p = (double *)calloc(n, sizeof(double));
    return p;
}

orc_state *orc_create(const orc_params *prm)
{
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *prm;
    const int N = prm->N;
    const size_t n = (size_t)N * N * N;
    s->n = n;
    double **all[] = {&s->phi, &s->px, &s->py, &s->pz, &s->vx, &s->vy, &s->vz,
                      &s->vxs, &s->vys, &s->vzs, &s->pxs, &s->pys, &s->pzs,
                      &s->vxs_o, &s->vys_o, &s->vzs_o, &s->pxs_o, &s->pys_o, &s->pzs_o,
                      &s->dPhi2, &s->mu0, &s->P2, &s->mux, &s->muy, &s->muz,
                      &s->DPx, &s->DPy, &s->DPz, &s->OPx, &s->OPy, &s->OPz,
                      &s->Axx, &s->Axy, &s->Axz, &s->Ayy, &s->Ayz, &s->Azz,
                      &s->dtvx, &s->dtvy, &s->dtvz, &s->dtpx, &s->dtpy, &s->dtpz, &s->dtphi,
                      &s->p_v, &s->p_p, &s->fre, &s->fim, &s->k2, &s->sdiv_v, &s->sdiv_p};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) *all[i] = zalloc(n);
    for (int i = 0; i < 9; i++) { s->G[i] = zalloc(n); s->Sig[i] = zalloc(n); }

    /* initPoissonSolver, TTSH3D.cu:1088-1124 (Nx=Ny=Nz=N) */
    const double Pi = 3.1415926535897932384626433832795;
    const double h = prm->h;
    for (int i = 0; i < N; i++) {
        double ky = 2 * Pi * i / (N * h + 0.0);
        if (i >= N / 2) ky = 2 * Pi * (i - N) / (N * h + 0.0);
        for (int j = 0; j < N; j++) {
            double kx = 2 * Pi * j / (N * h + 0.0);
            if (j >= N / 2) kx = 2 * Pi * (j - N) / (N * h + 0.0);
            for (int k = 0; k < N; k++) {
                double kz = 2 * Pi * k / (N * h + 0.0);
                if (k >= N / 2) kz = 2 * Pi * (k - N) / (N * h + 0.0);
                size_t idx = (size_t)(N * N) * k + (size_t)i * N + j;
                s->k2[idx] = kx * kx + ky * ky + kz * kz;
            }
        }
    }
    s->k2[0] = 1;
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    double *all[] = {s->phi, s->px, s->py, s->pz, s->vx, s->vy, s->vz,
                     s->vxs, s->vys, s->vzs, s->pxs, s->pys, s->pzs,
                     s->vxs_o, s->vys_o, s->vzs_o, s->pxs_o, s->pys_o, s->pzs_o,
                     s->dPhi2, s->mu0, s->P2, s->mux, s->muy, s->muz,
                     s->DPx, s->DPy, s->DPz, s->OPx, s->OPy, s->OPz,
                     s->Axx, s->Axy, s->Axz, s->Ayy, s->Ayz, s->Azz,
                     s->dtvx, s->dtvy, s->dtvz, s->dtpx, s->dtpy, s->dtpz, s->dtphi,
                     s->p_v, s->p_p, s->fre, s->fim, s->k2, s->sdiv_v, s->sdiv_p};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) free(all[i]);
    for (int i = 0; i < 9; i++) { free(s->G[i]); free(s->Sig[i]); }
    free(s);
}

/* phi[n], P[3n] (x,y,z planes), v[3n]; vs_prev/ps_prev may be NULL (then v, P). */
void orc_set_state(orc_state *s, const double *phi, const double *P, const double *v,
                   const double *vs_prev, const double *ps_prev)
{
    const size_t n = s->n, b = n * sizeof(double);
    memcpy(s->phi, phi, b);
    memcpy(s->px, P, b); memcpy(s->py, P + n, b); memcpy(s->pz, P + 2 * n, b);
    memcpy(s->vx, v, b); memcpy(s->vy, v + n, b); memcpy(s->vz, v + 2 * n, b);
    const double *vp = vs_prev ? vs_prev : v, *pp = ps_prev ? ps_prev : P;
    /* the "current" predictor slot holds v*_{n-1} until the next step */
    memcpy(s->vxs, vp, b); memcpy(s->vys, vp + n, b); memcpy(s->vzs, vp + 2 * n, b);
    memcpy(s->pxs, pp, b); memcpy(s->pys, pp + n, b); memcpy(s->pzs, pp + 2 * n, b);
}

void orc_get_state(const orc_state *s, double *phi, double *P, double *v)
{
    const size_t n = s->n, b = n * sizeof(double);
    if (phi) memcpy(phi, s->phi, b);
    if (P) { memcpy(P, s->px, b); memcpy(P + n, s->py, b); memcpy(P + 2 * n, s->pz, b); }
    if (v) { memcpy(v, s->vx, b); memcpy(v + n, s->vy, b); memcpy(v + 2 * n, s->vz, b); }
}

/* which: 0 v*(3n) 1 P*(3n) 2 p_v 3 p_p 4 dtv(3n) 5 dtP(3n) 6 dtphi 7 mu0 8 P2
 *        9 mu(3n) 10 Sigma(9n, xx xy xz yx yy yz zx zy zz) 11 G(9n)
 *        12 div(v*)/dt0 13 div(P*)/dt0 (before the FFT) */
void orc_get_aux(const orc_state *s, int which, double *out)
{
    const size_t n = s->n, b = n * sizeof(double);
    switch (which) {
    case 0: memcpy(out, s->vxs, b); memcpy(out + n, s->vys, b); memcpy(out + 2 * n, s->vzs, b); break;
    case 1: memcpy(out, s->pxs, b); memcpy(out + n, s->pys, b); memcpy(out + 2 * n, s->pzs, b); break;
    case 2: memcpy(out, s->p_v, b); break;
    case 3: memcpy(out, s->p_p, b); break;
    case 4: memcpy(out, s->dtvx, b); memcpy(out + n, s->dtvy, b); memcpy(out + 2 * n, s->dtvz, b); break;
    case 5: memcpy(out, s->dtpx, b); memcpy(out + n, s->dtpy, b); memcpy(out + 2 * n, s->dtpz, b); break;
    case 6: memcpy(out, s->dtphi, b); break;
    case 7: memcpy(out, s->mu0, b); break;
    case 8: memcpy(out, s->P2, b); break;
    case 9: memcpy(out, s->mux, b); memcpy(out + n, s->muy, b); memcpy(out + 2 * n, s->muz, b); break;
    case 10: for (int i = 0; i < 9; i++) memcpy(out + i * n, s->Sig[i], b); break;
    case 11: for (int i = 0; i < 9; i++) memcpy(out + i * n, s->G[i], b); break;
    case 12: memcpy(out, s->sdiv_v, b); break;
    case 13: memcpy(out, s->sdiv_p, b); break;
    default: break;
    }
}

/* ==========================================================================
 * The step
 * ======================================================================== */

/* getTempVari_1, TTSH3D.cu:537-558 */
static void temp_vari_1(orc_state *s)
{
    const int N = s->p.N;
    const double h = s->p.h;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                const size_t idx = nb[13];
                s->G[0][idx] = d1x(s->px, nb, h);
                s->G[1][idx] = d1y(s->px, nb, h);
                s->G[2][idx] = d1z(s->px, nb, h);
                s->G[3][idx] = d1x(s->py, nb, h);
                s->G[4][idx] = d1y(s->py, nb, h);
                s->G[5][idx] = d1z(s->py, nb, h);
                s->G[6][idx] = d1x(s->pz, nb, h);
                s->G[7][idx] = d1y(s->pz, nb, h);
                s->G[8][idx] = d1z(s->pz, nb, h);
                const double ax = d1x(s->phi, nb, h), ay = d1y(s->phi, nb, h), az = d1z(s->phi, nb, h);
                s->dPhi2[idx] = ax * ax + ay * ay + az * az;
            }
        }
}

/* getTempVari_2, TTSH3D.cu:561-605 */
static void temp_vari_2(orc_state *s)
{
    const int N = s->p.N;
    const orc_params P = s->p;
    const double h = P.h;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                const size_t idx = nb[13];
                const double phi = s->phi[idx], px = s->px[idx], py = s->py[idx], pz = s->pz[idx];
                s->mu0[idx] = P.aPhi * 2 * (phi - 3 * phi * phi + 2 * phi * phi * phi) - P.kPhi * lap(s->phi, nb, h);
                s->P2[idx] = px * px + py * py + pz * pz;
                s->mux[idx] = phi * px + P.La * P.La * lap(s->px, nb, h);
                s->muy[idx] = phi * py + P.La * P.La * lap(s->py, nb, h);
                s->muz[idx] = phi * pz + P.La * P.La * lap(s->pz, nb, h);

                const double dxvx = d1x(s->vx, nb, h), dyvx = d1y(s->vx, nb, h), dzvx = d1z(s->vx, nb, h);
                const double dxvy = d1x(s->vy, nb, h), dyvy = d1y(s->vy, nb, h), dzvy = d1z(s->vy, nb, h);
                const double dxvz = d1x(s->vz, nb, h), dyvz = d1y(s->vz, nb, h), dzvz = d1z(s->vz, nb, h);
                /* integer 1/2 == 0 in the reference (TTSH3D.cu:579-585) */
                s->DPx[idx] = 1 / 2 * (px * 2 * dxvx + py * (dxvy + dyvx) + pz * (dxvz + dzvx));
                s->DPy[idx] = 1 / 2 * (px * (dyvx + dxvy) + py * 2 * dyvy + pz * (dyvz + dzvy));
                s->DPz[idx] = 1 / 2 * (px * (dzvx + dxvz) + py * (dzvy + dyvz) + pz * 2 * dzvz);
                s->OPx[idx] = 1 / 2 * (py * (dxvy - dyvx) + pz * (dxvz - dzvx));
                s->OPy[idx] = 1 / 2 * (pz * (dyvz - dzvy) + px * (dyvx - dxvy));
                s->OPz[idx] = 1 / 2 * (px * (dzvx - dxvz) + py * (dzvy - dyvz));

                /* Sig_ab = phi*(daPb + 1/28*lap(daPb)), 1/28 == 0 (TTSH3D.cu:587-595).
                 * Sig order xx xy xz yx yy yz zx zy zz uses G index 3*b+a. */
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        const double *g = s->G[3 * b + a];
                        s->Sig[3 * a + b][idx] = phi * (g[idx] + 1 / 28 * lap(g, nb, h));
                    }

                const double ax = d1x(s->phi, nb, h), ay = d1y(s->phi, nb, h), az = d1z(s->phi, nb, h);
                const double dP2 = s->dPhi2[idx];
                s->Axx[idx] = ax * ax - 1 / 3 * dP2;
                s->Axy[idx] = ax * ay;
                s->Axz[idx] = ax * az;
                s->Ayy[idx] = ay * ay - 1 / 3 * dP2;
                s->Ayz[idx] = ay * az;
                s->Azz[idx] = az * az - 1 / 3 * dP2;
            }
        }
}

/* getDtF, TTSH3D.cu:608-632 */
static void get_dtf(orc_state *s)
{
    const int N = s->p.N;
    const orc_params P = s->p;
    const double h = P.h;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                const size_t idx = nb[13];
                const double vx = s->vx[idx], vy = s->vy[idx], vz = s->vz[idx];
                const double px = s->px[idx], py = s->py[idx], pz = s->pz[idx], phi = s->phi[idx];
                const double mu0 = s->mu0[idx], P2 = s->P2[idx], dP2 = s->dPhi2[idx];
                const double ax = d1x(s->phi, nb, h), ay = d1y(s->phi, nb, h), az = d1z(s->phi, nb, h);

                s->dtvx[idx] = -vx * d1x(s->vx, nb, h) - vy * d1y(s->vx, nb, h) - vz * d1z(s->vx, nb, h)
                    + P.mu * lap(s->vx, nb, h) + mu0 * ax
                    + P.cA * (d1x(s->Sig[0], nb, h) + d1y(s->Sig[3], nb, h) + d1z(s->Sig[6], nb, h));
                s->dtvy[idx] = -vx * d1x(s->vy, nb, h) - vy * d1y(s->vy, nb, h) - vz * d1z(s->vy, nb, h)
                    + P.mu * lap(s->vy, nb, h) + mu0 * ay
                    + P.cA * (d1x(s->Sig[1], nb, h) + d1y(s->Sig[4], nb, h) + d1z(s->Sig[7], nb, h));
                s->dtvz[idx] = -vx * d1x(s->vz, nb, h) - vy * d1y(s->vz, nb, h) - vz * d1z(s->vz, nb, h)
                    + P.mu * lap(s->vz, nb, h) + mu0 * az
                    + P.cA * (d1x(s->Sig[2], nb, h) + d1y(s->Sig[5], nb, h) + d1z(s->Sig[8], nb, h));

                const double dxpx = d1x(s->px, nb, h), dypx = d1y(s->px, nb, h), dzpx = d1z(s->px, nb, h);
                const double dxpy = d1x(s->py, nb, h), dypy = d1y(s->py, nb, h), dzpy = d1z(s->py, nb, h);
                const double dxpz = d1x(s->pz, nb, h), dypz = d1y(s->pz, nb, h), dzpz = d1z(s->pz, nb, h);
                const double sP2 = sqrt(P2);

                s->dtpx[idx] = -vx * dxpx - vy * dypx - vz * dzpx - P.alpha * phi * px - P.beta * P2 * px
                    + P.eta * lap(s->mux, nb, h) + s->OPx[idx] + P.kappa * phi * s->DPx[idx]
                    - P.lambda0 * (px * dxpx + py * dypx + pz * dzpx)
                    - P.A * (dP2 * (px * px * px + px * py * py + px * pz * pz)
                             + sP2 * (s->Axx[idx] * px + s->Axy[idx] * py + s->Axz[idx] * pz));
                s->dtpy[idx] = -vx * dxpy - vy * dypy - vz * dzpy - P.alpha * phi * py - P.beta * P2 * py
                    + P.eta * lap(s->muy, nb, h) + s->OPy[idx] + P.kappa * phi * s->DPy[idx]
                    - P.lambda0 * (px * dxpy + py * dypy + pz * dzpy)
                    - P.A * (dP2 * (px * py * px + py * py * py + py * pz * pz)
                             + sP2 * (s->Axy[idx] * px + s->Ayy[idx] * py + s->Ayz[idx] * pz));
                s->dtpz[idx] = -vx * dxpz - vy * dypz - vz * dzpz - P.alpha * phi * pz - P.beta * P2 * pz
                    + P.eta * lap(s->muz, nb, h) + s->OPz[idx] + P.kappa * phi * s->DPz[idx]
                    - P.lambda0 * (px * dxpz + py * dypz + pz * dzpz)
                    - P.A * (dP2 * (px * pz * px + py * pz * py + pz * pz * pz)
                             + sP2 * (s->Axz[idx] * px + s->Ayz[idx] * py + s->Azz[idx] * pz));

                s->dtphi[idx] = -vx * ax - vy * ay - vz * az + P.gamma * (lap(s->mu0, nb, h));
            }
        }
}

/* getVStar, TTSH3D.cu:471-491 */
static void get_vstar(orc_state *s, double ddt)
{
    const size_t n = s->n;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        s->vxs[i] = s->vx[i] + ddt * s->dtvx[i];
        s->vys[i] = s->vy[i] + ddt * s->dtvy[i];
        s->vzs[i] = s->vz[i] + ddt * s->dtvz[i];
        s->pxs[i] = s->px[i] + ddt * s->dtpx[i];
        s->pys[i] = s->py[i] + ddt * s->dtpy[i];
        s->pzs[i] = s->pz[i] + ddt * s->dtpz[i];
    }
}

/* getP + getStream (types 0,1,2), TTSH3D.cu:635-691 */
static void get_p(orc_state *s, const orc_fftplan *pl, int fieldType)
{
    const int N = s->p.N;
    const double h = s->p.h, dt0 = s->p.dt0;
    const double *fx = fieldType == 0 ? s->vxs : s->pxs;
    const double *fy = fieldType == 0 ? s->vys : s->pys;
    const double *fz = fieldType == 0 ? s->vzs : s->pzs;
    const int stale = s->p.faithful;
    const double *gx = stale ? (fieldType == 0 ? s->vxs_o : s->pxs_o) : fx;
    const double *gy = stale ? (fieldType == 0 ? s->vys_o : s->pys_o) : fy;
    const double *gz = stale ? (fieldType == 0 ? s->vzs_o : s->pzs_o) : fz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            uint32_t wrap;
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, &wrap);
                const size_t idx = nb[13];
                s->fre[idx] = (d1x_h(fx, gx, nb, wrap, h) + d1y_h(fy, gy, nb, wrap, h) + d1z_h(fz, gz, nb, wrap, h)) / dt0;
                s->fim[idx] = 0;
            }
        }
    const size_t n = s->n;
    memcpy(fieldType == 0 ? s->sdiv_v : s->sdiv_p, s->fre, n * sizeof(double));
    fft3d_plan(pl, s->fre, s->fim, -1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        if (s->k2[i] == 0) {
            s->fre[i] = 0;
            s->fim[i] = 0;
        } else {
            s->fre[i] = -s->fre[i] / (s->k2[i]);
            s->fim[i] = -s->fim[i] / (s->k2[i]);
        }
    }
    fft3d_plan(pl, s->fre, s->fim, +1);
    double *out = fieldType == 0 ? s->p_v : s->p_p;
    const int NNN = N * N * N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) out[i] = s->fre[i] / (NNN);
}

/* getVNew, TTSH3D.cu:494-516 */
static void get_vnew(orc_state *s, double ddt)
{
    const int N = s->p.N;
    const double h = s->p.h;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                const size_t idx = nb[13];
                s->vx[idx] = s->vxs[idx] - ddt * d1x(s->p_v, nb, h);
                s->vy[idx] = s->vys[idx] - ddt * d1y(s->p_v, nb, h);
                s->vz[idx] = s->vzs[idx] - ddt * d1z(s->p_v, nb, h);
                s->px[idx] = s->pxs[idx] - ddt * d1x(s->p_p, nb, h);
                s->py[idx] = s->pys[idx] - ddt * d1y(s->p_p, nb, h);
                s->pz[idx] = s->pzs[idx] - ddt * d1z(s->p_p, nb, h);
                s->phi[idx] = s->phi[idx] + ddt * s->dtphi[idx];
            }
        }
}

static void swap_ptr(double **a, double **b)
{
    double *t = *a;
    *a = *b;
    *b = t;
}

/* EEuler, TTSH3D.cu:347-359; Dt = dt0 (TTSH3D.cu:759) */
void orc_step(orc_state *s, int nsteps)
{
    orc_fftplan pl;
    fftplan_init(&pl, s->p.N);
    const double Dt = s->p.dt0;
    for (int it = 0; it < nsteps; it++) {
        /* The halo exchange at the top of getRHS (TTSH3D.cu:369-371,405-407)
         * freezes the v* and P* of the previous step into the halo. */
        swap_ptr(&s->vxs, &s->vxs_o); swap_ptr(&s->vys, &s->vys_o); swap_ptr(&s->vzs, &s->vzs_o);
        swap_ptr(&s->pxs, &s->pxs_o); swap_ptr(&s->pys, &s->pys_o); swap_ptr(&s->pzs, &s->pzs_o);
        temp_vari_1(s);
        temp_vari_2(s);
        get_dtf(s);
        get_vstar(s, Dt);
        get_p(s, &pl, 0);
        get_p(s, &pl, 1);
        get_vnew(s, Dt);
    }
    fftplan_free(&pl);
}

/* The projection alone, clean halo: p from getP (TTSH3D.cu:635-691) applied to
 * the given predictor Xs (3n), then the corrector X = Xs - Dt D p
 * (TTSH3D.cu:505-511).  Writes p (n) and X (3n). */
void orc_projection(const orc_params *prm, const double *Xs, double *p_out, double *X_out)
{
    orc_params q = *prm;
    q.faithful = 0;
    orc_state *s = orc_create(&q);
    const size_t n = s->n, b = n * sizeof(double);
    memcpy(s->vxs, Xs, b);
    memcpy(s->vys, Xs + n, b);
    memcpy(s->vzs, Xs + 2 * n, b);
    orc_fftplan pl;
    fftplan_init(&pl, q.N);
    get_p(s, &pl, 0);
    fftplan_free(&pl);
    memcpy(p_out, s->p_v, b);
    const int N = q.N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++) {
            size_t nb[27];
            for (int x = 0; x < N; x++) {
                neighbourhood(N, x, y, z, nb, NULL);
                const size_t idx = nb[13];
                X_out[idx] = s->vxs[idx] - q.dt0 * d1x(s->p_v, nb, q.h);
                X_out[n + idx] = s->vys[idx] - q.dt0 * d1y(s->p_v, nb, q.h);
                X_out[2 * n + idx] = s->vzs[idx] - q.dt0 * d1z(s->p_v, nb, q.h);
            }
        }
    orc_destroy(s);
}

/* FP64 diagnostics: mass = sum(phi) h^3, mean |P|, kinetic energy = 1/2 sum |v|^2 h^3 */
void orc_reductions(const orc_state *s, double out[3])
{
    const size_t n = s->n;
    const double h3 = s->p.h * s->p.h * s->p.h;
    double m = 0, pm = 0, ke = 0;
#pragma omp parallel for reduction(+ : m, pm, ke) schedule(static)
    for (size_t i = 0; i < n; i++) {
        m += s->phi[i];
        pm += sqrt(s->px[i] * s->px[i] + s->py[i] * s->py[i] + s->pz[i] * s->pz[i]);
        ke += s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i];
    }
    out[0] = m * h3;
    out[1] = pm / (double)n;
    out[2] = 0.5 * ke * h3;
}

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

size_t orc_sizeof_params(void) { return sizeof(orc_params); }
