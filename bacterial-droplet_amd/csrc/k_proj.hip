// k_proj.hip -- divergence (K_div) and corrector (K_corr) stencils of the
// projection, and the FP64 reductions.
//
// Reference: getStream type 0 (TTSH3D.cu:654-669), getVNew (TTSH3D.cu:494-516).
// The reference projects v* and P* with two complex transforms each; since
// -1/k^2 is real and even, packing div v* into the real part and div P* into
// the imaginary part of one complex field gives IFFT(-FFT(a + i b)/k^2) =
// p_v + i p_p exactly, so K_div writes s = (D.v* + i D.P*)/dt0 and K_corr
// reads p_v + i p_p after the inverse transforms (k_fft.hip).
//
// In FP32 the rounding of the larger half of s leaks into the smaller one.
// An aster's div P* can exceed div v* by 10^2 and its P by 10^3 its v; left
// unscaled, v then loses three digits in the first step.  k_proj_stats
// picks a power of two sigma per droplet and step; the first forward pass
// divides the imaginary part by it and K_corr multiplies p_p back (both
// exact).  Cross-talk into v scales with |div P*|/(sigma |div v*|) relative
// to |P|/|v|, into P with the inverse, so
//   sigma^4 = (sum (div P*)^2 sum |P*|^2) / (sum (div v*)^2 sum |v*|^2)
// gives both corrected fields the same relative share, rounded to 2^e,
// |e| <= 30 (e = 0 when s has no imaginary part or a sum is undefined).
//
// Both are z-marching stencils: a work-group owns a TX x TY column and walks
// a z-range, loading one plane (with a one-cell xy apron) per iteration into
// LDS and adding the plane's contribution to the three output planes it
// touches.  With D_a = 1/(2h) d_a (x) S (x) S, S = [1,4,1]/6 (TTSH3D.cu:3280-3328):
//   D_x f (z) = sum_dz S(dz) [1/(2h) S_y d_x f](z+dz)
//   D_z f (z) = [1/(2h) S_x S_y f](z+1) - [..](z-1)
//
// K_div reproduces the reference's stale halo: an apron or wrapped-plane
// position, i.e. any position outside the box that the reference reads from
// its halo, is taken from the previous step's v*, P* (TTSH3D.cu:366-371,
// 405-407: the halo was filled before getVStar overwrote the interior).
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {


template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_div(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                 const int* __restrict__ nwork_p, int parity, int kstep)
{
    constexpr int NT = TX * TY, W1 = TX + 2, A1 = W1 * (TY + 2);
    __shared__ float sF[6][A1];
    __shared__ double red[kPartSums * (NT / 64)];
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nwork = *nwork_p;
    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int N = D.N;
        const Coef co = D.co;
        const int nbx = (N + TX - 1) / TX;
        const int x0 = (wk.a % nbx) * TX, y0 = (wk.a / nbx) * TY;
        const int zb = wk.b & 0xffff, ze = wk.b >> 16;
        const int pp = D.faithful ? (parity ^ 1) : parity;
        gptr<const float> cur[6], prv[6];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            cur[f] = glb<const float>(D.vs[parity][f]);
            cur[3 + f] = glb<const float>(D.ps[parity][f]);
            prv[f] = glb<const float>(D.vs[pp][f]);
            prv[3 + f] = glb<const float>(D.ps[pp][f]);
        }
        const int gx = x0 + tx, gy = y0 + ty;
        const bool own = gx < N && gy < N;
        const size_t NN = (size_t)N * N;
        float am[2] = {0.f, 0.f}, ac[2] = {0.f, 0.f}, ap[2] = {0.f, 0.f};
        // sums over this item's cells (see Drop::part): s in FP64 (the k = 0
        // coefficient), the squares, which only set sigma, per thread in FP32
        double q[kPartSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        float q2 = 0.f, q3 = 0.f, q4 = 0.f, q5 = 0.f;
        // Per field, the current and the previous-step buffer are addressed
        // from the lower of the two (uniform base lo[f]) with a per-lane byte
        // offset that adds the distance dh[f] to reach the higher one.  Every
        // load is then one scalar-base load; the buffers of a droplet share
        // one allocation, so the distance fits 32 bits.
        gptr<const float> lo[6];
        unsigned dh[6];
        bool prv_hi[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            prv_hi[f] = prv[f] > cur[f];
            lo[f] = prv_hi[f] ? cur[f] : prv[f];
            dh[f] = 4u * (unsigned)(prv_hi[f] ? prv[f] - cur[f] : cur[f] - prv[f]);
        }
        // In the stale-halo mode, positions outside the box in x read the
        // previous step's x = 0 / x = N-1 columns from the edge buffer (same
        // z-stride as a field, so the same scalar base serves; K_corr writes
        // each step's columns).
        const bool xedge = D.faithful != 0;
        const int eprev = D.eslot[pp];
        // Loads: the thread's own cell, one per field (a wave loads two whole
        // aligned rows), and the one-cell ring around the tile (top row,
        // bottom row, left and right column) as RING x 6 (position, field)
        // pairs spread over two loads per thread.  Ring loads address every
        // field from gb, the lowest of the droplet's predictor buffers (one
        // allocation, with the edge buffer after them: 32-bit offsets).
        constexpr int RING = 2 * W1 + 2 * TY;
        static_assert(A1 - NT == RING && 6 * RING <= 2 * NT, "two ring loads per thread");
        gptr<const float> gb = lo[0];
#pragma unroll
        for (int f = 1; f < 6; ++f) gb = lo[f] < gb ? lo[f] : gb;
        auto pos = [&](int i, bool& sxy, int& ex, int& ey) {
            ex = x0 - 1 + i % W1;
            ey = y0 - 1 + i / W1;
            sxy = ex < 0 || ex >= N || ey < 0 || ey >= N;
            return 4u * ((unsigned)wrapi(ey, N) * (unsigned)N + (unsigned)wrapi(ex, N));
        };
        // byte offset of field f's value at position (ex, ey) from base b:
        // the current or previous buffer (stale: outside the box, or the
        // whole plane when wz), or the edge buffer for x outside the box
        auto foff = [&](int f, gptr<const float> b, unsigned o, bool stale, int ex, int ey) {
            if (xedge && (ex < 0 || ex >= N)) {
                const int col = ((eprev * 6 + f) * 2 + (ex < 0 ? 1 : 0)) * N + wrapi(ey, N);
                return 4u * (unsigned)(glb<const float>(D.edge) - b) + 4u * (unsigned)col;
            }
            return 4u * (unsigned)(lo[f] - b) + o + ((stale == prv_hi[f]) ? dh[f] : 0u);
        };
        unsigned own_o[6], own_w[6];  // own cell: normal plane, whole-plane-stale plane
        bool own_in;
        {
            bool sxy;
            int ex, ey;
            const unsigned o = pos((ty + 1) * W1 + (tx + 1), sxy, ex, ey);
            own_in = !sxy;
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                own_o[f] = foff(f, lo[f], o, sxy, ex, ey);
                own_w[f] = foff(f, lo[f], o, true, ex, ey);
            }
        }
        unsigned ring_o[2], ring_w[2];
        int ring_l[2];  // flat LDS index f*A1 + i (-1: past the last pair)
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int j = min(tid + m * NT, 6 * RING - 1), f = j / RING, r = j % RING;
            int i;
            if (r < W1) i = r;                                      // row y0-1
            else if (r < 2 * W1) i = (TY + 1) * W1 + (r - W1);      // row y0+TY
            else if (r < 2 * W1 + TY) i = (r - 2 * W1 + 1) * W1;    // column x0-1
            else i = (r - 2 * W1 - TY + 1) * W1 + (W1 - 1);         // column x0+TX
            bool sxy;
            int ex, ey;
            const unsigned o = pos(i, sxy, ex, ey);
            ring_o[m] = foff(f, gb, o, sxy, ex, ey);
            ring_w[m] = foff(f, gb, o, true, ex, ey);
            ring_l[m] = tid + m * NT < 6 * RING ? f * A1 + i : -1;
        }
        // planes are loaded two ahead, into two register slots; a plane
        // outside the box (wrapped in z) is read wholly from the previous step
        float pre[2][8];
        auto fetch = [&](int slot, int zi) {
            const size_t zo = (size_t)wrapi(zi, N) * NN;
            if (zi >= 0 && zi < N) {
#pragma unroll
                for (int f = 0; f < 6; ++f) pre[slot][f] = ldb(lo[f] + zo, own_o[f]);
#pragma unroll
                for (int m = 0; m < 2; ++m) pre[slot][6 + m] = ldb(gb + zo, ring_o[m]);
            } else {
#pragma unroll
                for (int f = 0; f < 6; ++f) pre[slot][f] = ldb(lo[f] + zo, own_w[f]);
#pragma unroll
                for (int m = 0; m < 2; ++m) pre[slot][6 + m] = ldb(gb + zo, ring_w[m]);
            }
        };
        // The sums are written per reduction block (see Drop::part): the
        // item starts at block wk.c, and at the last plane of each block the
        // work-group reduces them in a fixed order and starts the next.  The
        // sums of a block are then the same whichever item covers it.
        const int ntile = nbx * ((N + TY - 1) / TY);
        int ordb = wk.c;
        int zend = (wk.c / ntile + 1) * N / D.nzp;
        auto flush = [&]() {
            q[2] = q2;
            q[3] = q3;
            q[4] = q4;
            q[5] = q5;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                for (int e = 0; e < kPartSums; ++e) q[e] += __shfl_down(q[e], off, 64);
            }
            __syncthreads();
            if ((tid & 63) == 0) {
#pragma unroll
                for (int e = 0; e < kPartSums; ++e) red[kPartSums * (tid >> 6) + e] = q[e];
            }
            __syncthreads();
            if (tid < kPartSums) {
                double a = 0.0;
                for (int i = 0; i < NT / 64; ++i) a += red[kPartSums * i + tid];
                glb(D.part)[(size_t)kPartSums * ordb + tid] = a;
            }
#pragma unroll
            for (int e = 0; e < kPartSums; ++e) q[e] = 0.0;
            q2 = q3 = q4 = q5 = 0.f;
            ordb += ntile;
            zend = (ordb / ntile + 1) * N / D.nzp;
        };
        auto body = [&](auto ph, int zi) {
            constexpr int S = decltype(ph)::value;
            // squares of this plane's inputs, added after the block check below
            float sq4 = 0.f, sq5 = 0.f;
            const bool sq = zi >= zb && zi < ze && own_in;
            if (sq) {
                const float* u = pre[S];
                sq4 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                sq5 = u[3] * u[3] + u[4] * u[4] + u[5] * u[5];
            }
            __syncthreads();
            {
                const int i = (ty + 1) * W1 + (tx + 1);
#pragma unroll
                for (int f = 0; f < 6; ++f) sF[f][i] = pre[S][f];
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    if (ring_l[m] >= 0) (&sF[0][0])[ring_l[m]] = pre[S][6 + m];
            }
            fetch(S, min(zi + 2, ze));  // past ze: a redundant reload of plane ze
            __syncthreads();
            const int o = (ty + 1) * W1 + (tx + 1);
#pragma unroll
            for (int F = 0; F < 2; ++F) {
                const float* fx = sF[3 * F];
                const float* fy = sF[3 * F + 1];
                const float* fz = sF[3 * F + 2];
                const float ax = co.sw0 * (fx[o + 1] - fx[o - 1])
                               + co.sw1 * ((fx[o - W1 + 1] - fx[o - W1 - 1]) + (fx[o + W1 + 1] - fx[o + W1 - 1]));
                const float ay = co.sw0 * (fy[o + W1] - fy[o - W1])
                               + co.sw1 * ((fy[o + W1 - 1] - fy[o - W1 - 1]) + (fy[o + W1 + 1] - fy[o - W1 + 1]));
                const float r0 = co.sw0 * fz[o] + co.sw1 * (fz[o - 1] + fz[o + 1]);
                const float rm = co.sw0 * fz[o - W1] + co.sw1 * (fz[o - W1 - 1] + fz[o - W1 + 1]);
                const float rp = co.sw0 * fz[o + W1] + co.sw1 * (fz[o + W1 - 1] + fz[o + W1 + 1]);
                const float A = co.d1 * (ax + ay);
                const float B = co.d1 * (co.sw0 * r0 + co.sw1 * (rm + rp));
                am[F] += co.sw1 * A + B;
                ac[F] += co.sw0 * A;
                ap[F] += co.sw1 * A - B;
            }
            const int zo = zi - 1;
            if (zo >= zb && own) {
                const float2 sv = make_float2(am[0] * co.inv_dt0, am[1] * co.inv_dt0);
                st2(glb(D.c), gidx(N, gx, gy, zo), sv);
                q[0] += (double)sv.x;
                q[1] += (double)sv.y;
                q2 += sv.x * sv.x;
                q3 += sv.y * sv.y;
            }
            if (zo >= zb && zo == zend - 1) flush();  // uniform across the work-group
            if (sq) {
                q4 += sq4;
                q5 += sq5;
            }
#pragma unroll
            for (int F = 0; F < 2; ++F) { am[F] = ac[F]; ac[F] = ap[F]; ap[F] = 0.f; }
        };
        fetch(0, zb - 1);
        fetch(1, zb);
        wait_vmem();
        for (int zi = zb - 1;;) {
            body(Ph<0>{}, zi);
            if (++zi > ze) break;
            body(Ph<1>{}, zi);
            if (++zi > ze) break;
        }
        // the item ends at a block end, so its last block was written above
    }
}

// One work-group per droplet (work item (d, -, -)): the sums of the
// divergence items in item order, then the k = 0 coefficient of s/sigma and
// this step's sigma (derivation at the top of this file).
constexpr int kStatsThreads = 256;

// log2 x for x >= 0 (-inf at 0, +inf at inf): exponent plus the FP32 log2 of
// the mantissa, enough for a rounded power of two
__device__ __forceinline__ double log2_coarse(double x)
{
    if (!(x > 0.0)) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int ex;
    const double m = frexp(x, &ex);
    return (double)ex + (double)log2f((float)m);
}

__global__ __launch_bounds__(kStatsThreads) void k_proj_stats(const Drop* __restrict__ drops,
                                                              const Work* __restrict__ work,
                                                              const int* __restrict__ nwork_p, int kstep)
{
    constexpr int NW = kStatsThreads / 64;
    __shared__ double red[kPartSums * NW];
    __shared__ double tot[kPartSums];
    const int tid = threadIdx.x;
    const int nwork = *nwork_p;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Drop& D = drops[work[w].d];
        if (kstep >= D.nsteps) continue;
        const int np = D.npart;
        const gptr<const double> pt = glb<const double>(D.part);
        double q[kPartSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int j = tid; j < np; j += kStatsThreads) {
#pragma unroll
            for (int e = 0; e < kPartSums; ++e) q[e] += pt[(size_t)kPartSums * j + e];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int e = 0; e < kPartSums; ++e) q[e] += __shfl_down(q[e], off, 64);
        }
        __syncthreads();
        if ((tid & 63) == 0) {
#pragma unroll
            for (int e = 0; e < kPartSums; ++e) red[kPartSums * (tid >> 6) + e] = q[e];
        }
        __syncthreads();
        if (tid < kPartSums) {
            double a = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) a += red[kPartSums * i + tid];
            tot[tid] = a;
        }
        __syncthreads();
        if (tid == 0) {
            double t[kPartSums];
#pragma unroll
            for (int e = 0; e < kPartSums; ++e) t[e] = tot[e];
            // log2 of 0 is -inf, of inf +inf; -inf + inf is NaN
            const double l = 0.25 * ((log2_coarse(t[3]) + log2_coarse(t[5])) - (log2_coarse(t[2]) + log2_coarse(t[4])));
            int e = 0;
            if (t[3] > 0.0 && l == l) e = (int)rint(fmin(fmax(l, -30.0), 30.0));
            const float sig = ldexpf(1.f, e);
            glb(D.sigma)[0] = sig;
            glb(D.stat)[0] = t[0];
            glb(D.stat)[1] = ldexp(t[1], -e);
        }
    }
}

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_corr(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                  const int* __restrict__ nwork_p, int parity, int kstep)
{
    constexpr int NT = TX * TY, W1 = TX + 2, A1 = W1 * (TY + 2);
    __shared__ float2 sp[A1];
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nwork = *nwork_p;
    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int N = D.N;
        const Coef co = D.co;
        const int nbx = (N + TX - 1) / TX;
        const int x0 = (wk.a % nbx) * TX, y0 = (wk.a / nbx) * TY;
        const int zb = wk.b & 0xffff, ze = wk.b >> 16;
        const gptr<const float2> c = glb<const float2>(D.c);
        // p_p = sigma * Im(p): the power-of-two scale folded into Dt
        const float dts = co.Dt * glb(D.sigma)[0];
        gptr<const float> src[6];
        gptr<float> dst[6];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            src[f] = glb<const float>(D.vs[parity][f]);
            src[3 + f] = glb<const float>(D.ps[parity][f]);
            dst[f] = glb(D.v[f]);
            dst[3 + f] = glb(D.P[f]);
        }
        const int gx = x0 + tx, gy = y0 + ty;
        const bool own = gx < N && gy < N;
        // accumulators of D_x p, D_y p, D_z p (complex) for planes zi-1, zi, zi+1
        float2 xm = {0.f, 0.f}, xc = {0.f, 0.f}, xp = {0.f, 0.f};
        float2 ym = {0.f, 0.f}, yc = {0.f, 0.f}, yp = {0.f, 0.f};
        float2 zmv = {0.f, 0.f}, zcv = {0.f, 0.f}, zpv = {0.f, 0.f};
        // planes of p are loaded two ahead into two register slots, with the
        // v*, P* of the own cell at the plane finished by the slot's
        // iteration; loads are unconditional (clamped positions and planes)
        // so no register needs a merge copy
        // slot 0: the own cell (whole aligned rows per wave); slot 1: the
        // one-cell ring around the tile, clamped so every thread loads it
        static_assert(A1 - NT == 2 * W1 + 2 * TY, "slot 1 holds the ring");
        constexpr int MAXP = 2;
        unsigned off[MAXP];
        int lds_i[MAXP];  // the position's LDS index (-1: nothing to store)
#pragma unroll
        for (int k = 0; k < MAXP; ++k) {
            int i;
            if (k == 0) {
                i = (ty + 1) * W1 + (tx + 1);
                lds_i[k] = i;
            } else {
                const int r = min(tid, 2 * W1 + 2 * TY - 1);
                if (r < W1) i = r;                                      // row y0-1
                else if (r < 2 * W1) i = (TY + 1) * W1 + (r - W1);      // row y0+TY
                else if (r < 2 * W1 + TY) i = (r - 2 * W1 + 1) * W1;    // column x0-1
                else i = (r - 2 * W1 - TY + 1) * W1 + (W1 - 1);         // column x0+TX
                lds_i[k] = tid < 2 * W1 + 2 * TY ? i : -1;
            }
            off[k] = (unsigned)wrapi(y0 - 1 + i / W1, N) * (unsigned)N + (unsigned)wrapi(x0 - 1 + i % W1, N);
        }
        const unsigned own_off = 4u * ((unsigned)min(gy, N - 1) * (unsigned)N + (unsigned)min(gx, N - 1));
        const size_t NN = (size_t)N * N;
        // In the stale-halo mode the cells at x = 0 and x = N-1 also store
        // this step's v*, P* in the edge buffer: the next step's divergence
        // reads them there (the layout of Drop::edge).
        const bool ewr = D.faithful != 0 && own && (gx == 0 || gx == N - 1);
        const unsigned ewo = 4u * (unsigned)(((D.eslot[parity] * 6) * 2 + (gx == 0 ? 0 : 1)) * N + min(gy, N - 1));
        float2 pre[2][MAXP];
        float sv[2][6];
        auto fetch = [&](int slot, int zi) {
            const gptr<const float2> cz = c + (size_t)wrapi(zi, N) * NN;
#pragma unroll
            for (int k = 0; k < MAXP; ++k) pre[slot][k] = ld2(cz, off[k]);
            // own-cell predictors of plane zi-1, the output of iteration zi
            const size_t zo = (size_t)wrapi(max(zi - 1, zb), N) * NN;
#pragma unroll
            for (int f = 0; f < 6; ++f) sv[slot][f] = ldb(src[f] + zo, own_off);
        };
        auto body = [&](auto ph, int zi) {
            constexpr int S = decltype(ph)::value;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < MAXP; ++k) {
                const int i = lds_i[k];
                if (i >= 0) sp[i] = pre[S][k];
            }
            // this iteration's own-cell predictors, before the slot is reused
            float svc[6];
#pragma unroll
            for (int f = 0; f < 6; ++f) svc[f] = sv[S][f];
            fetch(S, min(zi + 2, ze));  // past ze: a redundant reload of plane ze
            __syncthreads();
            const int o = (ty + 1) * W1 + (tx + 1);
            float2 ax, ay, bz;
#define SPX(k) sp[k].x
#define SPY(k) sp[k].y
            ax.x = co.sw0 * (SPX(o + 1) - SPX(o - 1)) + co.sw1 * ((SPX(o - W1 + 1) - SPX(o - W1 - 1)) + (SPX(o + W1 + 1) - SPX(o + W1 - 1)));
            ax.y = co.sw0 * (SPY(o + 1) - SPY(o - 1)) + co.sw1 * ((SPY(o - W1 + 1) - SPY(o - W1 - 1)) + (SPY(o + W1 + 1) - SPY(o + W1 - 1)));
            ay.x = co.sw0 * (SPX(o + W1) - SPX(o - W1)) + co.sw1 * ((SPX(o + W1 - 1) - SPX(o - W1 - 1)) + (SPX(o + W1 + 1) - SPX(o - W1 + 1)));
            ay.y = co.sw0 * (SPY(o + W1) - SPY(o - W1)) + co.sw1 * ((SPY(o + W1 - 1) - SPY(o - W1 - 1)) + (SPY(o + W1 + 1) - SPY(o - W1 + 1)));
            {
                const float r0x = co.sw0 * SPX(o) + co.sw1 * (SPX(o - 1) + SPX(o + 1));
                const float rmx = co.sw0 * SPX(o - W1) + co.sw1 * (SPX(o - W1 - 1) + SPX(o - W1 + 1));
                const float rpx = co.sw0 * SPX(o + W1) + co.sw1 * (SPX(o + W1 - 1) + SPX(o + W1 + 1));
                const float r0y = co.sw0 * SPY(o) + co.sw1 * (SPY(o - 1) + SPY(o + 1));
                const float rmy = co.sw0 * SPY(o - W1) + co.sw1 * (SPY(o - W1 - 1) + SPY(o - W1 + 1));
                const float rpy = co.sw0 * SPY(o + W1) + co.sw1 * (SPY(o + W1 - 1) + SPY(o + W1 + 1));
                bz.x = co.d1 * (co.sw0 * r0x + co.sw1 * (rmx + rpx));
                bz.y = co.d1 * (co.sw0 * r0y + co.sw1 * (rmy + rpy));
            }
#undef SPX
#undef SPY
            ax.x *= co.d1; ax.y *= co.d1; ay.x *= co.d1; ay.y *= co.d1;
            xm.x += co.sw1 * ax.x; xm.y += co.sw1 * ax.y;
            xc.x += co.sw0 * ax.x; xc.y += co.sw0 * ax.y;
            xp.x += co.sw1 * ax.x; xp.y += co.sw1 * ax.y;
            ym.x += co.sw1 * ay.x; ym.y += co.sw1 * ay.y;
            yc.x += co.sw0 * ay.x; yc.y += co.sw0 * ay.y;
            yp.x += co.sw1 * ay.x; yp.y += co.sw1 * ay.y;
            zmv.x += bz.x; zmv.y += bz.y;  // plane zi is the z+1 plane of output zi-1
            zpv.x -= bz.x; zpv.y -= bz.y;  // and the z-1 plane of output zi+1
            const int zo = zi - 1;
            if (zo >= zb && own) {
                const size_t g = gidx(N, gx, gy, zo);
                dst[0][g] = svc[0] - co.Dt * xm.x;
                dst[1][g] = svc[1] - co.Dt * ym.x;
                dst[2][g] = svc[2] - co.Dt * zmv.x;
                dst[3][g] = svc[3] - dts * xm.y;
                dst[4][g] = svc[4] - dts * ym.y;
                dst[5][g] = svc[5] - dts * zmv.y;
                if (ewr) {
                    const gptr<float> e = glb(D.edge) + (size_t)zo * NN;
#pragma unroll
                    for (int f = 0; f < 6; ++f) stb(e, ewo + 4u * (unsigned)(f * 2 * N), svc[f]);
                }
            }
            xm = xc; xc = xp; xp = make_float2(0.f, 0.f);
            ym = yc; yc = yp; yp = make_float2(0.f, 0.f);
            zmv = zcv; zcv = zpv; zpv = make_float2(0.f, 0.f);
        };
        fetch(0, zb - 1);
        fetch(1, zb);
        wait_vmem();
        for (int zi = zb - 1;;) {
            body(Ph<0>{}, zi);
            if (++zi > ze) break;
            body(Ph<1>{}, zi);
            if (++zi > ze) break;
        }
    }
}

// ===========================================================================
// FP64 reductions
// ===========================================================================
__global__ __launch_bounds__(256) void k_reduce_part(const Drop* __restrict__ drops, int parity, double* part, int nblk)
{
    __shared__ double red[3 * 4];
    const int d = blockIdx.y;
    const Drop& D = drops[d];
    const size_t n = (size_t)D.N * D.N * D.N;
    const float* phi = D.phi[parity];
    double m = 0.0, pm = 0.0, ke = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)nblk * blockDim.x) {
        const double px = D.P[0][i], py = D.P[1][i], pz = D.P[2][i];
        const double vx = D.v[0][i], vy = D.v[1][i], vz = D.v[2][i];
        m += (double)phi[i];
        pm += sqrt(px * px + py * py + pz * pz);
        ke += vx * vx + vy * vy + vz * vz;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m += __shfl_down(m, off, 64);
        pm += __shfl_down(pm, off, 64);
        ke += __shfl_down(ke, off, 64);
    }
    if (lane == 0) { red[3 * wid] = m; red[3 * wid + 1] = pm; red[3 * wid + 2] = ke; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int w = 0; w < 4; ++w) { a += red[3 * w]; b += red[3 * w + 1]; c += red[3 * w + 2]; }
        double* o = part + 3 * ((size_t)d * nblk + blockIdx.x);
        o[0] = a; o[1] = b; o[2] = c;
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* part, int nblk, double* out)
{
    __shared__ double red[3 * 4];
    const int d = blockIdx.x;
    double m = 0.0, pm = 0.0, ke = 0.0;
    for (int j = threadIdx.x; j < nblk; j += blockDim.x) {
        const double* p = part + 3 * ((size_t)d * nblk + j);
        m += p[0]; pm += p[1]; ke += p[2];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m += __shfl_down(m, off, 64);
        pm += __shfl_down(pm, off, 64);
        ke += __shfl_down(ke, off, 64);
    }
    if (lane == 0) { red[3 * wid] = m; red[3 * wid + 1] = pm; red[3 * wid + 2] = ke; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int w = 0; w < 4; ++w) { a += red[3 * w]; b += red[3 * w + 1]; c += red[3 * w + 2]; }
        out[3 * d] = a; out[3 * d + 1] = b; out[3 * d + 2] = c;
    }
}

// ===========================================================================
// launchers
// ===========================================================================
hipError_t launch_div(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int grid,
                      hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return launch_kernel(k_div<kProjTX, kProjTY>, dim3(grid), dim3(kProjTX * kProjTY), st, drops, work, nwork_p,
                         parity, kstep);
}

hipError_t launch_proj_stats(const Drop* drops, const Work* work, const int* nwork_p, int kstep, int grid,
                             hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return launch_kernel(k_proj_stats, dim3(grid), dim3(kStatsThreads), st, drops, work, nwork_p, kstep);
}

// One thread per (z, y): the two end columns of the six predictor fields.
__global__ __launch_bounds__(256) void k_fill_edges(const float* __restrict__ v0, const float* __restrict__ v1,
                                                    const float* __restrict__ v2, const float* __restrict__ p0,
                                                    const float* __restrict__ p1, const float* __restrict__ p2,
                                                    float* __restrict__ edge, int slot, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * N) return;
    const int y = i % N, z = i / N;
    const size_t row = ((size_t)z * N + y) * N, NN = (size_t)N * N;
    const float* src[6] = {v0, v1, v2, p0, p1, p2};
    const gptr<float> e = glb(edge) + (size_t)z * NN;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        e[((slot * 6 + f) * 2 + 0) * N + y] = glb(src[f])[row];
        e[((slot * 6 + f) * 2 + 1) * N + y] = glb(src[f])[row + N - 1];
    }
}

hipError_t launch_fill_edges(const float* const vs[3], const float* const ps[3], float* edge, int slot, int N,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_edges, dim3((N * N + 255) / 256), dim3(256), 0, st, vs[0], vs[1], vs[2], ps[0], ps[1],
                       ps[2], edge, slot, N);
    return hipGetLastError();
}

// Fill n 32-bit words with a value (vector stores; grid-stride).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) glb(p)[i] = value;
}

hipError_t launch_fill_u32(void* p, uint32_t value, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<uint32_t*>(p), value, n);
    return hipGetLastError();
}

hipError_t launch_corr(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int grid,
                       hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return launch_kernel(k_corr<kProjTX, kProjTY>, dim3(grid), dim3(kProjTX * kProjTY), st, drops, work, nwork_p,
                         parity, kstep);
}

hipError_t launch_reduce(const Drop* drops, int ndrops, int parity, double* part, int nblk, double* out,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_part, dim3(nblk, ndrops), dim3(256), 0, st, drops, parity, part, nblk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(ndrops), dim3(256), 0, st, part, nblk, out);
    return hipGetLastError();
}

}  // namespace ttsh
