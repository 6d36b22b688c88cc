// k_fft.hip -- batched 1D FFT passes along x, y or z of a droplet's complex
// projection buffer, with the grid size N fixed at compile time.
//
// These replace the reference's cufftExecZ2Z FORWARD/INVERSE on its
// (Ny, Nx, Nz) plan (TTSH3D.cu:638, :640, :1123) and the pointwise getStream
// type 1 and 2 kernels (-1/k2 with k2[0]=1, and the 1/N^3 scaling,
// TTSH3D.cu:671-690), which are folded into the z pass.
//
// Each line of length N = R1*R2 is transformed by the four-step method
// (Gentleman & Sande 1966; Bailey 1990): T = max(R1, R2) threads per line
//   A: thread n1 < R1 loads x[n1 + R1 n2] (n2 < R2), does a length-R2 DFT in
//      registers, multiplies by W_N^{n1 k2} and stores to LDS;
//   B: thread k2 < R2 loads the R1 values of its k2, does a length-R1 DFT in
//      registers and writes X[k2 + R2 k1].
// Register DFTs of length 2, 3, 4, 5 are written out; longer ones (6, 8, 10,
// 12, 15, 16, 20) are composed from them with compile-time twiddles.  The
// plans cover the grid sizes of kGridSizes.  Sign convention:
// forward e^{-2 pi i jk/N}, inverse e^{+2 pi i jk/N}, both unnormalised, as
// cufftExecZ2Z.
//
// The first forward pass (x, or the fused x+y) divides the imaginary part of
// s by the droplet's power-of-two sigma and the z pass takes the k = 0
// coefficient from the FP64 sum of s; both come from k_proj_stats
// (Drop::sigma, Drop::stat).
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

// No multiply-add contraction in this file: every pass then performs the
// same IEEE operations whichever kernel runs a line (single-axis passes or
// the fused x+y pass), so a droplet's result does not depend on the batch
// it runs in.
#pragma clang fp contract(off)

namespace ttsh {

// cos and sin of 2 pi r/48 for r in [0, 12): the first quadrant of the 48th
// roots of unity, which contain every root a DFT of length 2..16 needs.
__device__ constexpr float kCos48[12] = {
    1.0f, 0.99144486137381038f, 0.96592582628906831f, 0.92387953251128674f,
    0.86602540378443865f, 0.79335334029123517f, 0.70710678118654752f, 0.60876142900872066f,
    0.5f, 0.38268343236508977f, 0.25881904510252076f, 0.13052619222005159f};
__device__ constexpr float kSin48[12] = {
    0.0f, 0.13052619222005159f, 0.25881904510252076f, 0.38268343236508977f,
    0.5f, 0.60876142900872066f, 0.70710678118654752f, 0.79335334029123517f,
    0.86602540378443865f, 0.92387953251128674f, 0.96592582628906831f, 0.99144486137381038f};

__device__ __forceinline__ float2 f2(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul2(float2 a, float2 w) { return f2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x); }

// a * exp(i 2 pi m/48) for a compile-time m (quarter turns are exact swaps).
template <int M>
__device__ __forceinline__ float2 rot48(float2 a)
{
    constexpr int m = ((M % 48) + 48) % 48;
    constexpr int q = m / 12, r = m % 12;
    float2 b = a;
    if constexpr (r != 0) b = mul2(a, f2(kCos48[r], kSin48[r]));
    if constexpr (q == 1) return f2(-b.y, b.x);
    else if constexpr (q == 2) return f2(-b.x, -b.y);
    else if constexpr (q == 3) return f2(b.y, -b.x);
    else return b;
}

// cos(2 pi r/240) for r in [0, 60]; sin(2 pi r/240) = cos(2 pi (60 - r)/240).
// The first quadrant of the 240th roots of unity, for the register DFTs
// whose length divides 240 but not 48 (5, 10, 15, 20).
__device__ constexpr float kCos240[61] = {
    1.0f, 0.99965732497555728004f, 0.99862953475457387378f, 0.9969173337331279762f,
    0.99452189536827333692f, 0.99144486137381041114f, 0.98768834059513772619f, 0.98325490756395458455f,
    0.97814760073380563793f, 0.97236992039767660183f, 0.96592582628906828675f, 0.95881973486819304976f,
    0.95105651629515357212f, 0.94264149109217839478f, 0.93358042649720174899f, 0.92387953251128675613f,
    0.9135454576426008955f, 0.90258528434986060676f, 0.89100652418836786236f, 0.87881711266196537413f,
    0.86602540378443864676f, 0.85264016435409222152f, 0.83867056794542402964f, 0.82412618862201566173f,
    0.8090169943749474241f, 0.79335334029123516458f, 0.77714596145697087998f, 0.76040596560003093817f,
    0.74314482547739423501f, 0.72537437101228763799f, 0.7071067811865475244f, 0.68835457569375398439f,
    0.66913060635885821383f, 0.64944804833018365573f, 0.62932039104983745271f, 0.60876142900872063942f,
    0.58778525229247312917f, 0.56640623692483283182f, 0.54463903501502708222f, 0.52249856471594886499f,
    0.5f, 0.47715876025960841505f, 0.45399049973954679156f, 0.43051109680829514438f,
    0.40673664307580020775f, 0.38268343236508977173f, 0.35836794954530027348f, 0.33380685923377092883f,
    0.3090169943749474241f, 0.28401534470392261744f, 0.25881904510252076235f, 0.23344536385590541177f,
    0.2079116908177593371f, 0.1822355254921474566f, 0.15643446504023086901f, 0.13052619222005159155f,
    0.1045284632676534714f, 0.078459095727844945033f, 0.052335956242943832722f, 0.026176948307873152611f,
    8.478427660368899644e-32f};

// a * exp(i 2 pi m/240) for a compile-time m (quarter turns are exact swaps).
template <int M>
__device__ __forceinline__ float2 rot240(float2 a)
{
    constexpr int m = ((M % 240) + 240) % 240;
    constexpr int q = m / 60, r = m % 60;
    float2 b = a;
    if constexpr (r != 0) b = mul2(a, f2(kCos240[r], kCos240[60 - r]));
    if constexpr (q == 1) return f2(-b.y, b.x);
    else if constexpr (q == 2) return f2(-b.x, -b.y);
    else if constexpr (q == 3) return f2(b.y, -b.x);
    else return b;
}

// a * W_R^{DIR*e}, W_R = exp(2 pi i/R) with DIR = -1 forward, +1 inverse.
// Lengths dividing 48 take the 48th-root table, the others the 240th-root one.
template <int R, int DIR, int E>
__device__ __forceinline__ float2 twid(float2 a)
{
    if constexpr (48 % R == 0) {
        return rot48<DIR * E * (48 / R)>(a);
    } else {
        static_assert(240 % R == 0, "register DFT length must divide 48 or 240");
        return rot240<DIR * E * (240 / R)>(a);
    }
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
    __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float2 a = v[0], b = v[1];
        v[0] = add2(a, b);
        v[1] = sub2(a, b);
    }
};

template <int DIR>
struct Dft<3, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        constexpr float c3 = 0.86602540378443864676f;  // sin(2 pi/3)
        const float2 t1 = add2(v[1], v[2]), t2 = sub2(v[1], v[2]);
        const float2 m = f2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
        // DIR * i * sin(2pi/3) * t2
        const float2 q = f2(-(float)DIR * c3 * t2.y, (float)DIR * c3 * t2.x);
        v[0] = add2(v[0], t1);
        v[1] = add2(m, q);
        v[2] = sub2(m, q);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float2 s0 = add2(v[0], v[2]), d0 = sub2(v[0], v[2]);
        const float2 s1 = add2(v[1], v[3]), d1 = sub2(v[1], v[3]);
        const float2 j = f2(-(float)DIR * d1.y, (float)DIR * d1.x);  // DIR * i * d1
        v[0] = add2(s0, s1);
        v[2] = sub2(s0, s1);
        v[1] = add2(d0, j);
        v[3] = sub2(d0, j);
    }
};

// Length 5: with t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3 and
// c_k = cos(2 pi k/5), s_k = sin(2 pi k/5):
//   X0 = x0 + t1 + t2,
//   X1, X4 = x0 + c1 t1 + c2 t2 -+ DIR i (s1 t3 + s2 t4),
//   X2, X3 = x0 + c2 t1 + c1 t2 -+ DIR i (s2 t3 - s1 t4)   (DIR = -1 forward).
template <int DIR>
struct Dft<5, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        constexpr float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        constexpr float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 t1 = add2(v[1], v[4]), t2 = add2(v[2], v[3]);
        const float2 t3 = sub2(v[1], v[4]), t4 = sub2(v[2], v[3]);
        const float2 a1 = f2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const float2 a2 = f2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        const float2 b1 = f2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        const float2 b2 = f2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        // DIR * i * b
        const float2 j1 = f2(-(float)DIR * b1.y, (float)DIR * b1.x);
        const float2 j2 = f2(-(float)DIR * b2.y, (float)DIR * b2.x);
        v[0] = add2(v[0], add2(t1, t2));
        v[1] = add2(a1, j1);
        v[4] = sub2(a1, j1);
        v[2] = add2(a2, j2);
        v[3] = sub2(a2, j2);
    }
};

// Composite length R = A*B: B-point DFTs over n2 for each n1, twiddle
// W_R^{n1 k2}, then A-point DFTs over n1; output index k2 + B k1.
template <int A, int B, int DIR>
struct DftComp {
    template <int N1, int K2>
    __device__ __forceinline__ static void tw_one(float2* t)
    {
        t[N1 * B + K2] = twid<A * B, DIR, N1 * K2>(t[N1 * B + K2]);
    }
    template <int N1, int... K2s>
    __device__ __forceinline__ static void tw_row(float2* t, std::integer_sequence<int, K2s...>)
    {
        (tw_one<N1, K2s>(t), ...);
    }
    template <int... N1s>
    __device__ __forceinline__ static void tw_all(float2* t, std::integer_sequence<int, N1s...>)
    {
        (tw_row<N1s>(t, std::make_integer_sequence<int, B>{}), ...);
    }
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 t[A * B];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) {
            float2 u[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) u[n2] = v[n1 + A * n2];
            Dft<B, DIR>::run(u);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) t[n1 * B + k2] = u[k2];
        }
        tw_all(t, std::make_integer_sequence<int, A>{});
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            float2 u[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) u[n1] = t[n1 * B + k2];
            Dft<A, DIR>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = u[k1];
        }
    }
};

template <int DIR> struct Dft<6, DIR> : DftComp<2, 3, DIR> {};
template <int DIR> struct Dft<8, DIR> : DftComp<2, 4, DIR> {};
template <int DIR> struct Dft<12, DIR> : DftComp<4, 3, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftComp<4, 4, DIR> {};
template <int DIR> struct Dft<10, DIR> : DftComp<2, 5, DIR> {};
template <int DIR> struct Dft<15, DIR> : DftComp<3, 5, DIR> {};
template <int DIR> struct Dft<20, DIR> : DftComp<4, 5, DIR> {};

// Line plan: N = R1*R2, T = max(R1,R2) threads per line, L lines per work-group.
template <int N>
struct Plan;
template <> struct Plan<32>  { static constexpr int R1 = 4,  R2 = 8,  L = 32; };
template <> struct Plan<36>  { static constexpr int R1 = 6,  R2 = 6,  L = 18; };
template <> struct Plan<40>  { static constexpr int R1 = 8,  R2 = 5,  L = 20; };
template <> struct Plan<48>  { static constexpr int R1 = 6,  R2 = 8,  L = 16; };
template <> struct Plan<50>  { static constexpr int R1 = 10, R2 = 5,  L = 25; };
template <> struct Plan<60>  { static constexpr int R1 = 12, R2 = 5,  L = 20; };
template <> struct Plan<64>  { static constexpr int R1 = 8,  R2 = 8,  L = 32; };
template <> struct Plan<72>  { static constexpr int R1 = 12, R2 = 6,  L = 18; };
template <> struct Plan<80>  { static constexpr int R1 = 10, R2 = 8,  L = 16; };
template <> struct Plan<90>  { static constexpr int R1 = 15, R2 = 6,  L = 15; };
template <> struct Plan<96>  { static constexpr int R1 = 8,  R2 = 12, L = 16; };
template <> struct Plan<100> { static constexpr int R1 = 10, R2 = 10, L = 25; };
template <> struct Plan<120> { static constexpr int R1 = 12, R2 = 10, L = 20; };
template <> struct Plan<128> { static constexpr int R1 = 16, R2 = 8,  L = 16; };
template <> struct Plan<144> { static constexpr int R1 = 12, R2 = 12, L = 16; };
template <> struct Plan<150> { static constexpr int R1 = 15, R2 = 10, L = 15; };
template <> struct Plan<160> { static constexpr int R1 = 10, R2 = 16, L = 16; };
template <> struct Plan<180> { static constexpr int R1 = 15, R2 = 12, L = 15; };
template <> struct Plan<192> { static constexpr int R1 = 16, R2 = 12, L = 16; };
template <> struct Plan<200> { static constexpr int R1 = 20, R2 = 10, L = 10; };
template <> struct Plan<240> { static constexpr int R1 = 16, R2 = 15, L = 16; };
template <> struct Plan<256> { static constexpr int R1 = 16, R2 = 16, L = 16; };

template <int N>
struct PlanT : Plan<N> {
    static constexpr int R1 = Plan<N>::R1, R2 = Plan<N>::R2, L = Plan<N>::L;
    static constexpr int T = R1 > R2 ? R1 : R2;
    static constexpr int NT = T * L;
    static_assert(R1 * R2 == N, "plan");
    static constexpr int NB = (NT + 63) / 64 * 64;  // work-group size
    static_assert(NT <= 1024, "work-group size");
    static_assert(N % L == 0, "lines per work-group must divide N");
};

// Pass kinds.
enum { FFT_FWD = 0, FFT_INV = 1, FFT_POISSON = 2 };

// LDS layout of the exchange buffer (L lines x R1 x R2 complex).  For the x
// pass lanes run along a line (index k2 fastest, padded); for y/z passes
// lanes run across lines (line fastest).
template <int N, int AXIS>
struct Xch {
    using P = PlanT<N>;
    static constexpr int R1 = P::R1, R2 = P::R2, L = P::L;
    static constexpr int SIZE = AXIS == 0 ? L * R1 * (R2 + 1) : L * R1 * R2;
    __device__ __forceinline__ static int at(int l, int n1, int k2)
    {
        if constexpr (AXIS == 0) return (l * R1 + n1) * (R2 + 1) + k2;
        else return (n1 * R2 + k2) * L + l;
    }
};

// -1/(k^2 N^3) at integer wavenumbers (fftfreq order), k2[0] = 1.
__device__ __forceinline__ float poisson_factor(const Coef& co, int N, int kx, int ky, int kz)
{
    const int mx = kx < N / 2 ? kx : kx - N;
    const int my = ky < N / 2 ? ky : ky - N;
    const int mz = kz < N / 2 ? kz : kz - N;
    const int m2 = mx * mx + my * my + mz * mz;
    const float k2 = m2 == 0 ? 1.f : co.k2s * (float)m2;
    return -co.inv_n3 / k2;
}

// Work item (d, a, b): droplet d; the pass's lines are the L consecutive
// lines starting at line-block a of the plane b (x pass: rows y of plane z=b;
// y pass: columns x of plane z=b; z pass: columns x of row y=b).
template <int N, int AXIS, int MODE>
__global__ __launch_bounds__(PlanT<N>::NB) void k_fft(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                              const int* __restrict__ nwork_p, int kstep)
{
    using P = PlanT<N>;
    constexpr int R1 = P::R1, R2 = P::R2, L = P::L, T = P::T, NT = P::NT;
    using X = Xch<N, AXIS>;
    __shared__ float2 sx[X::SIZE];
    __shared__ float2 stw[N];

    const int tid = threadIdx.x;
    const bool active = tid < NT;
    // lane -> (line l, thread-in-line t)
    int l, t;
    if constexpr (AXIS == 0) { l = tid / T; t = tid % T; }
    else { t = tid / L; l = tid % L; }
    const int nwork = *nwork_p;
    constexpr size_t es = AXIS == 0 ? 1 : (AXIS == 1 ? (size_t)N : (size_t)N * N);

    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        __syncthreads();  // the previous item has read sx and stw
        // first element of line l
        size_t base;
        int lx = 0, ly = 0;  // for the Poisson factor: the line's fixed indices
        if constexpr (AXIS == 0) {
            ly = wk.a * L + l;
            base = ((size_t)wk.b * N + ly) * N;
        } else if constexpr (AXIS == 1) {
            lx = wk.a * L + l;
            base = (size_t)wk.b * N * N + lx;
        } else {
            lx = wk.a * L + l;
            ly = wk.b;
            base = (size_t)ly * N + lx;
        }
        const gptr<float2> c = glb(D.c);

        // the Poisson pass's work item holding kx = ky = 0 sets the k = 0
        // coefficient from the FP64 sum of s (k_proj_stats)
        const bool has0 = MODE == FFT_POISSON && wk.a == 0 && wk.b == 0;
        // the first forward pass divides the imaginary part by sigma (a
        // power of two: exact)
        float isg = 1.f;
        if constexpr (AXIS == 0 && MODE == FFT_FWD) isg = 1.f / glb(D.sigma)[0];

        constexpr int DIR0 = MODE == FFT_INV ? 1 : -1;
        float2 u[R1 > R2 ? R1 : R2];
        // ---- step A: R1 threads, length-R2 DFTs -----------------------------
        // The line's loads are issued before the twiddle table's, so both
        // are in flight together; the table's LDS stores wait for all of them.
        if (active && t < R1) {
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) u[n2] = ld2(c, base + (size_t)(t + R1 * n2) * es);
        }
        for (int i = tid; i < N; i += blockDim.x) stw[i] = ld2(glb(D.tw), i);
        if (active && t < R1) {
            if constexpr (AXIS == 0 && MODE == FFT_FWD) {
#pragma unroll
                for (int n2 = 0; n2 < R2; ++n2) u[n2].y *= isg;
            }
            Dft<R2, DIR0>::run(u);
        }
        __syncthreads();  // stw
        if (active && t < R1) {
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const float2 wv = stw[(t * k2) % N];
                const float2 w2 = f2(wv.x, DIR0 == -1 ? wv.y : -wv.y);
                sx[X::at(l, t, k2)] = k2 == 0 ? u[0] : mul2(u[k2], w2);
            }
        }
        __syncthreads();
        // ---- step B: R2 threads, length-R1 DFTs -----------------------------
        if (active && t < R2) {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) u[n1] = sx[X::at(l, n1, t)];
            Dft<R1, DIR0>::run(u);
        }
        if constexpr (MODE != FFT_POISSON) {
            if (active && t < R2) {
#pragma unroll
                for (int k1 = 0; k1 < R1; ++k1) st2(c, base + (size_t)(t + R2 * k1) * es, u[k1]);
            }
        } else {
            // multiply by -1/(k^2 N^3); element index kz = t + R2 k1
            if (active && t < R2) {
                const Coef co = D.co;
#pragma unroll
                for (int k1 = 0; k1 < R1; ++k1) {
                    const int kz = t + R2 * k1;
                    float2 s = u[k1];
                    if (has0 && l == 0 && kz == 0) s = f2((float)glb(D.stat)[0], (float)glb(D.stat)[1]);
                    const float m = poisson_factor(co, N, lx, ly, kz);
                    u[k1] = f2(s.x * m, s.y * m);
                }
                // inverse, roles swapped: n' = t + R2 k1, thread t = n1' < R2,
                // length-R1 DFT over k1 in registers
                Dft<R1, 1>::run(u);
            }
            __syncthreads();  // every thread has read sx in step B
            if (active && t < R2) {
#pragma unroll
                for (int m2 = 0; m2 < R1; ++m2) {
                    const float2 wv = stw[(t * m2) % N];
                    sx[X::at(l, m2, t)] = m2 == 0 ? u[0] : mul2(u[m2], f2(wv.x, -wv.y));
                }
            }
            __syncthreads();
            if (active && t < R1) {
#pragma unroll
                for (int n1 = 0; n1 < R2; ++n1) u[n1] = sx[X::at(l, t, n1)];
                Dft<R2, 1>::run(u);
#pragma unroll
                for (int m1 = 0; m1 < R2; ++m1) st2(c, base + (size_t)(t + R1 * m1) * es, u[m1]);
            }
        }
    }
}

// ===========================================================================
// Fused x+y pass for N <= 128: one work-group per z-plane holds the N x N
// complex plane in LDS (row pitch N+1) and transforms all rows and all
// columns there, so the plane crosses HBM once per direction instead of
// twice.  Forward: x then y (e^{-i...}); inverse: y then x (e^{+i...});
// unnormalised like the single-axis passes.
// ===========================================================================
template <int N>
struct XyCfg {
    static constexpr int NT = N >= 96 ? 1024 : 256;
    static constexpr int P = N + 1;  // LDS row pitch in complex elements
};

typedef float fvec4 __attribute__((ext_vector_type(4)));

// In-place DFT of the N lines of the plane along x (AX 0: line l is row y=l)
// or y (AX 1: line l is column x=l), four-step as the single-axis passes.
template <int N, int AX, int DIR, int NT>
__device__ __forceinline__ void plane_dft(float2* __restrict__ pl, const float2* __restrict__ stw, int tid)
{
    using PL = PlanT<N>;
    constexpr int R1 = PL::R1, R2 = PL::R2, P = XyCfg<N>::P;
    auto at = [](int l, int j) { return AX == 0 ? l * P + j : j * P + l; };
    // A: task (l, n1): length-R2 DFT of x[n1 + R1 n2], twiddle W_N^{n1 k2},
    // back to the same positions.  Lanes run along a row (x lines) or across
    // columns (y lines), contiguous in LDS either way.
    constexpr int TA = N * R1;
    for (int task = tid; task < TA; task += NT) {
        const int l = AX == 0 ? task / R1 : task % N;
        const int n1 = AX == 0 ? task % R1 : task / N;
        float2 u[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) u[n2] = pl[at(l, n1 + R1 * n2)];
        Dft<R2, DIR>::run(u);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) {
            const float2 wv = stw[(n1 * k2) % N];
            const float2 w2 = f2(wv.x, DIR == -1 ? wv.y : -wv.y);
            pl[at(l, n1 + R1 * k2)] = k2 == 0 ? u[0] : mul2(u[k2], w2);
        }
    }
    __syncthreads();
    // B: task (l, k2): length-R1 DFT over n1, result X[k2 + R2 k1]; all reads
    // precede all writes (same plane)
    constexpr int TB = N * R2, KB = (TB + NT - 1) / NT;
    float2 v[KB][R1];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int task = tid + k * NT;
        if (task < TB) {
            const int l = AX == 0 ? task / R2 : task % N;
            const int k2 = AX == 0 ? task % R2 : task / N;
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) v[k][n1] = pl[at(l, n1 + R1 * k2)];
            Dft<R1, DIR>::run(v[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int task = tid + k * NT;
        if (task < TB) {
            const int l = AX == 0 ? task / R2 : task % N;
            const int k2 = AX == 0 ? task % R2 : task / N;
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) pl[at(l, k2 + R2 * k1)] = v[k][k1];
        }
    }
    __syncthreads();
}

// Work item (d, -, z): plane z of droplet d.
template <int N, int MODE>
__global__ __launch_bounds__(XyCfg<N>::NT) void k_fft_xy(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                        const int* __restrict__ nwork_p, int kstep)
{
    constexpr int NT = XyCfg<N>::NT, P = XyCfg<N>::P;
    constexpr int DIR = MODE == FFT_INV ? 1 : -1;
    __shared__ float2 pl[N * P];
    __shared__ float2 stw[N];
    const int tid = threadIdx.x;
    const int nwork = *nwork_p;
    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const gptr<fvec4> cp = (gptr<fvec4>)(glb(D.c) + (size_t)wk.b * N * N);
        __syncthreads();  // the previous item's stores have read the plane
        // load: two complex elements of one row per 16-byte access, all of
        // the thread's accesses issued before the twiddle table's and sigma's
        // so that they are in flight together; the forward pass divides the
        // imaginary part by sigma (a power of two)
        constexpr int NQ = (N * N / 2 + NT - 1) / NT;
        fvec4 q[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int i = tid + k * NT;
            if (i < N * N / 2) q[k] = cp[i];
        }
        for (int i = tid; i < N; i += NT) stw[i] = ld2(glb(D.tw), i);
        const float isg = MODE == FFT_FWD ? 1.f / glb(D.sigma)[0] : 1.f;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int i = tid + k * NT;
            if (i < N * N / 2) {
                if constexpr (MODE == FFT_FWD) {
                    q[k].y *= isg;
                    q[k].w *= isg;
                }
                const int y = (2 * i) / N, x = (2 * i) % N;
                pl[y * P + x] = f2(q[k].x, q[k].y);
                pl[y * P + x + 1] = f2(q[k].z, q[k].w);
            }
        }
        __syncthreads();
        if constexpr (MODE == FFT_INV) {
            plane_dft<N, 1, DIR, NT>(pl, stw, tid);
            plane_dft<N, 0, DIR, NT>(pl, stw, tid);
        } else {
            plane_dft<N, 0, DIR, NT>(pl, stw, tid);
            plane_dft<N, 1, DIR, NT>(pl, stw, tid);
        }
        for (int i = tid; i < N * N / 2; i += NT) {
            const int y = (2 * i) / N, x = (2 * i) % N;
            const float2 a = pl[y * P + x], b = pl[y * P + x + 1];
            fvec4 q;
            q.x = a.x;
            q.y = a.y;
            q.z = b.x;
            q.w = b.y;
            cp[i] = q;
        }
    }
}

template <int N>
static hipError_t launch_fft_xy_n(int mode, const Drop* drops, const Work* work, const int* nwork_p, int kstep,
                                  int grid, hipStream_t st)
{
    const dim3 blk(XyCfg<N>::NT);
    if (mode == FFT_FWD) return launch_kernel(k_fft_xy<N, FFT_FWD>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    if (mode == FFT_INV) return launch_kernel(k_fft_xy<N, FFT_INV>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    return hipErrorInvalidValue;
}

bool fft_xy_fused(int N) { return N <= kMaxFusedXy; }

// f(std::integral_constant<int, N>) for the grid size N of the size list.
template <int I = 0, class F>
static hipError_t with_size(int N, F&& f)
{
    if constexpr (I == kNumGridSizes) {
        return hipErrorInvalidValue;
    } else {
        if (N == kGridSizes[I]) return f(std::integral_constant<int, kGridSizes[I]>{});
        return with_size<I + 1>(N, f);
    }
}

hipError_t launch_fft_xy(int N, int mode, const Drop* drops, const Work* work, const int* nwork_p, int kstep, int grid,
                         hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return with_size(N, [&](auto n) {
        constexpr int NS = decltype(n)::value;
        if constexpr (NS <= kMaxFusedXy) return launch_fft_xy_n<NS>(mode, drops, work, nwork_p, kstep, grid, st);
        else return hipErrorInvalidValue;
    });
}

template <int N>
static hipError_t launch_fft_n(int axis, int mode, const Drop* drops, const Work* work, const int* nwork_p,
                               int kstep, int grid, hipStream_t st)
{
    const dim3 blk(PlanT<N>::NB);
    if (axis == 0 && mode == FFT_FWD) return launch_kernel(k_fft<N, 0, FFT_FWD>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    if (axis == 0 && mode == FFT_INV) return launch_kernel(k_fft<N, 0, FFT_INV>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    if (axis == 1 && mode == FFT_FWD) return launch_kernel(k_fft<N, 1, FFT_FWD>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    if (axis == 1 && mode == FFT_INV) return launch_kernel(k_fft<N, 1, FFT_INV>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    if (axis == 2 && mode == FFT_POISSON)
        return launch_kernel(k_fft<N, 2, FFT_POISSON>, dim3(grid), blk, st, drops, work, nwork_p, kstep);
    return hipErrorInvalidValue;
}

hipError_t launch_fft(int N, int axis, int mode, const Drop* drops, const Work* work, const int* nwork_p, int kstep,
                      int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return with_size(N, [&](auto n) {
        return launch_fft_n<decltype(n)::value>(axis, mode, drops, work, nwork_p, kstep, grid, st);
    });
}

int fft_lines_per_group(int N)
{
    int L = 0;
    (void)with_size(N, [&](auto n) {
        L = PlanT<decltype(n)::value>::L;
        return hipSuccess;
    });
    return L;
}

}  // namespace ttsh
