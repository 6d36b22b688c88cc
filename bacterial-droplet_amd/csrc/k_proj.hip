// k_proj.hip -- the projection of the TTSH3D step: divergence, 3D FFT Poisson
// solve and corrector, in four kernels per step.
//
// Reference: getP/getStream (TTSH3D.cu:635-691), the cuFFT Z2Z plan and
// transforms (:1123, :638, :640), initPoissonSolver (:1088-1124) and getVNew
// (:494-516).  The reference solves for p_v and p_p with two complex
// transforms each; because -1/k^2 is real and even, packing div v* into the
// real part and div P* into the imaginary part of one complex field gives
// IFFT(-FFT(a + i b)/k^2) = p_v + i p_p exactly, so one forward and one
// inverse 3D transform serve both.
//
//   K_divx  s = (D.v* + i D.P*)/dt0 for a block of x-rows, then FFT along x.
//           Neighbours across the periodic boundary read the previous step's
//           v*, P* (the reference's stale halo, TTSH3D.cu:366-371,405-407,665-667).
//   K_ffty  FFT along y (forward, and again inverse after K_fftz).
//   K_fftz  FFT along z, multiply by -1/(k^2 N^3) with k2[0]=1 (TTSH3D.cu:1120,
//           :680-681, :687), inverse FFT along z.  The k=0 coefficient comes
//           from an FP64 sum of s.
//   K_corrx inverse FFT along x of the rows a block of outputs needs, then
//           v = v* - Dt D p_v, P = P* - Dt D p_p.
//
// The FFT is a Stockham autosort transform in LDS (Stockham 1966; Cochran et
// al. 1967; the work-group formulation of Govindaraju et al., SC 2008) with
// radix 4, 2 and 3 stages and FP32 twiddles computed in FP64 on the host.
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {

__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Small DFT of radix R in place; fs = +1 forward (e^{-i}), -1 inverse.
__device__ __forceinline__ void butterfly(float2* v, int R, float fs)
{
    if (R == 4) {
        const float2 s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        const float2 s1 = cadd(v[1], v[3]), d1 = csub(v[1], v[3]);
        const float2 j = make_float2(fs * d1.y, -fs * d1.x);  // -i*fs*d1
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, j);
        v[3] = csub(d0, j);
    } else if (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {  // R == 3
        const float c3 = 0.866025403784438646763723170752936f;  // sqrt(3)/2
        const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
        const float2 q = make_float2(fs * c3 * t2.y, -fs * c3 * t2.x);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, q);
        v[2] = csub(m, q);
    }
}

// Stockham FFT of nl lines of length N held in LDS; element j of line l is at
// buf[l*ls + j*es].  INTERLEAVED maps consecutive lanes to consecutive lines
// (for es > 1), otherwise to consecutive elements of a line.  Starts and ends
// with all threads synchronised.
template <bool INTERLEAVED>
__device__ void lds_fft(float2* buf, int nl, int ls, int es, int N, int nst, const int* __restrict__ rad,
                        const float2* tw, float fs)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    int Ns = 1;
    for (int s = 0; s < nst; ++s) {
        const int R = rad[s];
        const int NR = N / R;
        const int M = Ns * R;
        const int tstep = N / M;
        const int nb = nl * NR;
        float2 u[kFftMaxB][4];
        int dst[kFftMaxB];
#pragma unroll
        for (int c = 0; c < kFftMaxB; ++c) {
            const int b = tid + c * nt;
            dst[c] = -1;
            if (b < nb) {
                int l, j;
                if (INTERLEAVED) { j = b / nl; l = b - j * nl; }
                else { l = b / NR; j = b - l * NR; }
                const int k = j % Ns;
                const float2* src = buf + l * ls;
                float2 v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (r < R) {
                        float2 a = src[(j + r * NR) * es];
                        if (r > 0) {
                            float2 w = tw[k * r * tstep];
                            w.y *= fs;
                            a = cmul(a, w);
                        }
                        v[r] = a;
                    } else {
                        v[r] = make_float2(0.f, 0.f);
                    }
                }
                butterfly(v, R, fs);
#pragma unroll
                for (int r = 0; r < 4; ++r) u[c][r] = v[r];
                dst[c] = l * ls + ((j / Ns) * M + k) * es;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kFftMaxB; ++c) {
            if (dst[c] >= 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (r < R) buf[dst[c] + r * Ns * es] = u[c][r];
            }
        }
        __syncthreads();
        Ns = M;
    }
}

// Fixed-order block sum of two doubles (deterministic); result valid in thread 0.
__device__ void block_sum2(double& a, double& b, double* red)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    __syncthreads();
    if (lane == 0) { red[2 * wid] = a; red[2 * wid + 1] = b; }
    __syncthreads();
    if (tid == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < (nt + 63) / 64; ++w) { sa += red[2 * w]; sb += red[2 * w + 1]; }
        a = sa;
        b = sb;
    }
}

// ===========================================================================
// K_divx
// ===========================================================================
__global__ __launch_bounds__(kFftThreads) void k_divx(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                      const int* __restrict__ nwork_p, int parity, int kstep)
{
    extern __shared__ float4 lds4[];
    __shared__ double red[2 * (kFftThreads / 64)];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nwork = *nwork_p;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int faithful = D.faithful;
        const int N = D.N;
        const Coef co = D.co;
        const int y0 = wk.a * kDivRows, z = wk.b;
        const int WE = N + 2, HE = kDivRows + 2, AE = WE * HE;
        float2* stw = reinterpret_cast<float2*>(lds4);
        float2* sl = stw + N;
        float* sg = reinterpret_cast<float*>(sl + kDivRows * N);

        __syncthreads();
        for (int i = tid; i < N; i += nt) stw[i] = D.tw[i];
        const float* cur[6] = {D.vs[parity][0], D.vs[parity][1], D.vs[parity][2],
                               D.ps[parity][0], D.ps[parity][1], D.ps[parity][2]};
        const int pp = faithful ? (parity ^ 1) : parity;
        const float* prv[6] = {D.vs[pp][0], D.vs[pp][1], D.vs[pp][2], D.ps[pp][0], D.ps[pp][1], D.ps[pp][2]};
        const int zm = wrapi(z - 1, N), zp = wrapi(z + 1, N);
        const bool wzm = (z == 0), wzp = (z == N - 1);
        for (int i = tid; i < AE; i += nt) {
            const int ex = i % WE - 1;
            const int ey = y0 - 1 + i / WE;
            const bool sxy = (ex < 0) || (ex >= N) || (ey < 0) || (ey >= N);
            const int xw = wrapi(ex, N), yw = wrapi(ey, N);
            const size_t gm = gidx(N, xw, yw, zm), g0 = gidx(N, xw, yw, z), gp = gidx(N, xw, yw, zp);
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const float* a0 = sxy ? prv[f] : cur[f];
                const float* am = (sxy || wzm) ? prv[f] : cur[f];
                const float* ap = (sxy || wzp) ? prv[f] : cur[f];
                const float vm = am[gm], v0 = a0[g0], vp = ap[gp];
                // components x, y: S_z smoothing; component z: z-difference
                sg[f * AE + i] = ((f % 3) == 2) ? (vp - vm) : (co.sw0 * v0 + co.sw1 * (vm + vp));
            }
        }
        __syncthreads();

        double accr = 0.0, acci = 0.0;
        for (int i = tid; i < kDivRows * N; i += nt) {
            const int ly = i / N, x = i - ly * N;
            const int o = (ly + 1) * WE + (x + 1);
            float dv[2];
#pragma unroll
            for (int F = 0; F < 2; ++F) {
                const float* gx = sg + (3 * F + 0) * AE;
                const float* gy = sg + (3 * F + 1) * AE;
                const float* gz = sg + (3 * F + 2) * AE;
                const float ax = co.sw0 * (gx[o + 1] - gx[o - 1])
                               + co.sw1 * ((gx[o - WE + 1] - gx[o - WE - 1]) + (gx[o + WE + 1] - gx[o + WE - 1]));
                const float ay = co.sw0 * (gy[o + WE] - gy[o - WE])
                               + co.sw1 * ((gy[o + WE - 1] - gy[o - WE - 1]) + (gy[o + WE + 1] - gy[o - WE + 1]));
                const float r0 = co.sw0 * gz[o] + co.sw1 * (gz[o - 1] + gz[o + 1]);
                const float rm = co.sw0 * gz[o - WE] + co.sw1 * (gz[o - WE - 1] + gz[o - WE + 1]);
                const float rp = co.sw0 * gz[o + WE] + co.sw1 * (gz[o + WE - 1] + gz[o + WE + 1]);
                const float az = co.sw0 * r0 + co.sw1 * (rm + rp);
                dv[F] = co.d1 * (ax + ay + az);
            }
            const float2 s = make_float2(dv[0] * co.inv_dt0, dv[1] * co.inv_dt0);
            sl[i] = s;
            accr += (double)s.x;
            acci += (double)s.y;
        }
        __syncthreads();
        lds_fft<false>(sl, kDivRows, N, 1, N, D.nst, D.rad, stw, 1.f);
        for (int i = tid; i < kDivRows * N; i += nt) {
            const int ly = i / N, x = i - ly * N;
            D.c[gidx(N, x, y0 + ly, z)] = sl[i];
        }
        block_sum2(accr, acci, red);
        if (tid == 0) {
            const int item = z * (N / kDivRows) + wk.a;
            D.part[2 * item] = accr;
            D.part[2 * item + 1] = acci;
        }
    }
}

// ===========================================================================
// K_ffty
// ===========================================================================
__global__ __launch_bounds__(kFftThreads) void k_ffty(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                      const int* __restrict__ nwork_p, int inverse, int kstep)
{
    extern __shared__ float4 lds4[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const float fs = inverse ? -1.f : 1.f;
    const int nwork = *nwork_p;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int N = D.N;
        const int x0 = wk.a * kColTile, z = wk.b;
        float2* stw = reinterpret_cast<float2*>(lds4);
        float2* sb = stw + N;
        __syncthreads();
        for (int i = tid; i < N; i += nt) stw[i] = D.tw[i];
        for (int i = tid; i < N * kColTile; i += nt) {
            const int y = i / kColTile, xc = i - y * kColTile;
            sb[i] = D.c[gidx(N, x0 + xc, y, z)];
        }
        __syncthreads();
        lds_fft<true>(sb, kColTile, 1, kColTile, N, D.nst, D.rad, stw, fs);
        for (int i = tid; i < N * kColTile; i += nt) {
            const int y = i / kColTile, xc = i - y * kColTile;
            D.c[gidx(N, x0 + xc, y, z)] = sb[i];
        }
    }
}

// ===========================================================================
// K_fftz
// ===========================================================================
__global__ __launch_bounds__(kFftThreads) void k_fftz(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                      const int* __restrict__ nwork_p, int kstep)
{
    extern __shared__ float4 lds4[];
    __shared__ double red[2 * (kFftThreads / 64)];
    __shared__ double k0[2];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nwork = *nwork_p;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int N = D.N;
        const Coef co = D.co;
        const int x0 = wk.a * kColTile, y = wk.b;
        const size_t NN = (size_t)N * N;
        float2* stw = reinterpret_cast<float2*>(lds4);
        float2* sb = stw + N;
        __syncthreads();
        for (int i = tid; i < N; i += nt) stw[i] = D.tw[i];
        const size_t base = (size_t)y * N + x0;
        for (int i = tid; i < N * kColTile; i += nt) {
            const int zz = i / kColTile, xc = i - zz * kColTile;
            sb[i] = D.c[base + (size_t)zz * NN + xc];
        }
        const bool has0 = (x0 == 0) && (y == 0);
        if (has0) {
            // FP64 sum of s over the whole droplet: the k=0 coefficient
            double a = 0.0, b = 0.0;
            for (int j = tid; j < D.npart; j += nt) {
                a += D.part[2 * j];
                b += D.part[2 * j + 1];
            }
            block_sum2(a, b, red);
            if (tid == 0) { k0[0] = a; k0[1] = b; }
        }
        __syncthreads();
        lds_fft<true>(sb, kColTile, 1, kColTile, N, D.nst, D.rad, stw, 1.f);
        const int my = (y < N / 2) ? y : y - N;
        for (int i = tid; i < N * kColTile; i += nt) {
            const int zz = i / kColTile, xc = i - zz * kColTile;
            const int kx = x0 + xc;
            const int mx = (kx < N / 2) ? kx : kx - N;
            const int mz = (zz < N / 2) ? zz : zz - N;
            float2 s = sb[i];
            float k2 = co.k2s * (float)(mx * mx + my * my + mz * mz);
            if (has0 && kx == 0 && zz == 0) {
                s = make_float2((float)k0[0], (float)k0[1]);
                k2 = 1.f;  // k2[0] = 1 (TTSH3D.cu:1120)
            }
            const float m = -co.inv_n3 / k2;
            sb[i] = make_float2(s.x * m, s.y * m);
        }
        __syncthreads();
        lds_fft<true>(sb, kColTile, 1, kColTile, N, D.nst, D.rad, stw, -1.f);
        for (int i = tid; i < N * kColTile; i += nt) {
            const int zz = i / kColTile, xc = i - zz * kColTile;
            D.c[base + (size_t)zz * NN + xc] = sb[i];
        }
    }
}

// ===========================================================================
// K_corrx
// ===========================================================================
__global__ __launch_bounds__(kFftThreads) void k_corrx(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                       const int* __restrict__ nwork_p, int parity, int kstep)
{
    extern __shared__ float4 lds4[];
    const int tid = threadIdx.x, nt = blockDim.x;
    constexpr int NR = kCorrRows + 2;
    const int nwork = *nwork_p;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const int N = D.N;
        const Coef co = D.co;
        const int y0 = wk.a * kCorrRows;
        const int zb = wk.b & 0xffff, ze = wk.b >> 16;
        float2* stw = reinterpret_cast<float2*>(lds4);
        float2* sp = stw + N;  // 3 slots x NR rows x N
        const int slot = NR * N;
        __syncthreads();
        for (int i = tid; i < N; i += nt) stw[i] = D.tw[i];
        int s_m = 0, s_c = 1, s_p = 2;
        for (int zp = zb - 1; zp <= ze; ++zp) {
            __syncthreads();
            {
                const int gz = wrapi(zp, N);
                float2* dst = sp + s_p * slot;
                for (int i = tid; i < NR * N; i += nt) {
                    const int r = i / N, x = i - r * N;
                    dst[i] = D.c[gidx(N, x, wrapi(y0 - 1 + r, N), gz)];
                }
            }
            __syncthreads();
            lds_fft<false>(sp + s_p * slot, NR, N, 1, N, D.nst, D.rad, stw, -1.f);
            const int z = zp - 1;
            if (z >= zb) {
                const float2* pm = sp + s_m * slot;
                const float2* pc = sp + s_c * slot;
                const float2* pq = sp + s_p * slot;
                for (int i = tid; i < kCorrRows * N; i += nt) {
                    const int ly = i / N, x = i - ly * N;
                    const int xm = (x == 0) ? N - 1 : x - 1, xp = (x == N - 1) ? 0 : x + 1;
                    // 27-point neighbourhood, rows ly..ly+2 of the slots
                    float2 q[27];
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy) {
                        const int ro = (ly + dy) * N;
                        q[0 * 9 + dy * 3 + 0] = pm[ro + xm];
                        q[0 * 9 + dy * 3 + 1] = pm[ro + x];
                        q[0 * 9 + dy * 3 + 2] = pm[ro + xp];
                        q[1 * 9 + dy * 3 + 0] = pc[ro + xm];
                        q[1 * 9 + dy * 3 + 1] = pc[ro + x];
                        q[1 * 9 + dy * 3 + 2] = pc[ro + xp];
                        q[2 * 9 + dy * 3 + 0] = pq[ro + xm];
                        q[2 * 9 + dy * 3 + 1] = pq[ro + x];
                        q[2 * 9 + dy * 3 + 2] = pq[ro + xp];
                    }
#define Q(dx, dy, dz) q[((dz) + 1) * 9 + ((dy) + 1) * 3 + ((dx) + 1)]
                    float2 a[3], b[3], c[3];
#pragma unroll
                    for (int k = -1; k <= 1; ++k) {
                        a[k + 1].x = co.sw0 * (Q(1, 0, k).x - Q(-1, 0, k).x)
                                   + co.sw1 * ((Q(1, -1, k).x - Q(-1, -1, k).x) + (Q(1, 1, k).x - Q(-1, 1, k).x));
                        a[k + 1].y = co.sw0 * (Q(1, 0, k).y - Q(-1, 0, k).y)
                                   + co.sw1 * ((Q(1, -1, k).y - Q(-1, -1, k).y) + (Q(1, 1, k).y - Q(-1, 1, k).y));
                        b[k + 1].x = co.sw0 * (Q(0, 1, k).x - Q(0, -1, k).x)
                                   + co.sw1 * ((Q(-1, 1, k).x - Q(-1, -1, k).x) + (Q(1, 1, k).x - Q(1, -1, k).x));
                        b[k + 1].y = co.sw0 * (Q(0, 1, k).y - Q(0, -1, k).y)
                                   + co.sw1 * ((Q(-1, 1, k).y - Q(-1, -1, k).y) + (Q(1, 1, k).y - Q(1, -1, k).y));
                        c[k + 1].x = co.sw0 * (Q(0, k, 1).x - Q(0, k, -1).x)
                                   + co.sw1 * ((Q(-1, k, 1).x - Q(-1, k, -1).x) + (Q(1, k, 1).x - Q(1, k, -1).x));
                        c[k + 1].y = co.sw0 * (Q(0, k, 1).y - Q(0, k, -1).y)
                                   + co.sw1 * ((Q(-1, k, 1).y - Q(-1, k, -1).y) + (Q(1, k, 1).y - Q(1, k, -1).y));
                    }
#undef Q
                    const float ddx_r = co.d1 * (co.sw0 * a[1].x + co.sw1 * (a[0].x + a[2].x));
                    const float ddx_i = co.d1 * (co.sw0 * a[1].y + co.sw1 * (a[0].y + a[2].y));
                    const float ddy_r = co.d1 * (co.sw0 * b[1].x + co.sw1 * (b[0].x + b[2].x));
                    const float ddy_i = co.d1 * (co.sw0 * b[1].y + co.sw1 * (b[0].y + b[2].y));
                    const float ddz_r = co.d1 * (co.sw0 * c[1].x + co.sw1 * (c[0].x + c[2].x));
                    const float ddz_i = co.d1 * (co.sw0 * c[1].y + co.sw1 * (c[0].y + c[2].y));
                    const size_t g = gidx(N, x, y0 + ly, z);
                    D.v[0][g] = D.vs[parity][0][g] - co.Dt * ddx_r;
                    D.v[1][g] = D.vs[parity][1][g] - co.Dt * ddy_r;
                    D.v[2][g] = D.vs[parity][2][g] - co.Dt * ddz_r;
                    D.P[0][g] = D.ps[parity][0][g] - co.Dt * ddx_i;
                    D.P[1][g] = D.ps[parity][1][g] - co.Dt * ddy_i;
                    D.P[2][g] = D.ps[parity][2][g] - co.Dt * ddz_i;
                }
            }
            const int t = s_m; s_m = s_c; s_c = s_p; s_p = t;
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
size_t divx_lds_bytes(int N)
{
    return sizeof(float2) * (N + kDivRows * N) + sizeof(float) * 6 * (N + 2) * (kDivRows + 2);
}
size_t ffty_lds_bytes(int N) { return sizeof(float2) * (N + (size_t)N * kColTile); }
size_t fftz_lds_bytes(int N) { return ffty_lds_bytes(N); }
size_t corrx_lds_bytes(int N) { return sizeof(float2) * (N + 3 * (size_t)(kCorrRows + 2) * N); }

hipError_t launch_divx(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int maxN,
                       int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_divx, dim3(grid), dim3(kFftThreads), divx_lds_bytes(maxN), st, drops, work, nwork_p, parity,
                       kstep);
    return hipGetLastError();
}

hipError_t launch_ffty(const Drop* drops, const Work* work, const int* nwork_p, int inverse, int kstep, int maxN,
                       int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ffty, dim3(grid), dim3(kFftThreads), ffty_lds_bytes(maxN), st, drops, work, nwork_p, inverse,
                       kstep);
    return hipGetLastError();
}

hipError_t launch_fftz(const Drop* drops, const Work* work, const int* nwork_p, int kstep, int maxN, int grid,
                       hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fftz, dim3(grid), dim3(kFftThreads), fftz_lds_bytes(maxN), st, drops, work, nwork_p, kstep);
    return hipGetLastError();
}

hipError_t launch_corrx(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int maxN,
                        int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_corrx, dim3(grid), dim3(kFftThreads), corrx_lds_bytes(maxN), st, drops, work, nwork_p,
                       parity, kstep);
    return hipGetLastError();
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
