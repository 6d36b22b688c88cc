// k_rhs.hip -- K_rhs: the fused right-hand side and predictor of one step.
//
// Replaces the reference's getRHS (TTSH3D.cu:363-468: 94 halo fills plus
// getTempVari_1 :537-558, getTempVari_2 :561-605, getDtF :608-632) and
// getVStar (:471-491) with one LDS-tiled pass that reads phi, P, v once and
// writes phi^{n+1}, v*, P*:
//
//   G_ab  = D_a P_b                    mu0 = 2 aPhi (phi - 3phi^2 + 2phi^3) - kPhi L phi
//   mu_b  = phi P_b + La^2 L P_b       Sig_ab = phi G_ab
//   dtv_b = -v.D v_b + mu L v_b + mu0 D_b phi + cA sum_a D_a Sig_ab
//   dtP_b = -v.D P_b - alpha phi P_b - beta |P|^2 P_b + eta L mu_b - lambda0 P.D P_b
//           - A (|D phi|^2 |P|^2 P_b + |P| sum_c D_b phi D_c phi P_c)
//   dtphi = -v.D phi + gamma L mu0
//   v* = v + Dt dtv,  P* = P + Dt dtP,  phi^{n+1} = phi + Dt dtphi   (getVNew :513)
//
// The reference's integer-division zeros (TTSH3D.cu:579-602: DP = OP = 0, the
// 1/28 correction of Sigma and the 1/3 trace of A) are terms that are exactly
// zero for finite fields, so they are not evaluated; kappa therefore has no
// effect, as in the reference.
//
// Stencils (TTSH3D.cu:3280-3343) in factored form, S = [1,4,1]/6:
//   D_x f = 1/(2h) sum_dy S(dy) [S_z f(x+1, y+dy) - S_z f(x-1, y+dy)]
//   D_z f = 1/(2h) sum_dx,dy S(dx) S(dy) [f(z+1) - f(z-1)](x+dx, y+dy)
//   L f   = lc sum_{dx,dy=+-1} [f(z-1)+f(z+1)](x+dx,y+dy)
//         + lf ([f(z-1)+f(z+1)] + f(x+-1) + f(y+-1)) + l0 f
// so each plane is first "z-prepared" per xy position (f, f(z-1)+f(z+1),
// f(z+1)-f(z-1), S_z f) and every derivative becomes a 2D stencil over those
// arrays, sharing the partial sums between neighbouring cells.
//
// Schedule: a work-group owns a TX x TY column of cells and marches through a
// z-range.  Each thread keeps the three z-planes of its xy positions in
// registers (phi, P over a 2-cell apron, v over a 1-cell apron) and loads the
// next plane while the current one is used.  Per iteration: z-prepare the
// plane behind the newest into LDS; compute the 13 intermediates (Sigma, mu,
// mu0) over the 1-cell apron into LDS; add their contribution to the D.Sigma,
// L mu, L mu0 of the three output planes they touch (registers); finish the
// output plane two behind the newest input plane.
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {

// The three z-prepared arrays of one field: f, f(z+1)-f(z-1), S_z f.
enum { ZF = 0, ZD = 1, ZS = 2, NZ = 3 };

struct Der4 {
    float dx, dy, dz, l;
};

template <int W>
__device__ __forceinline__ float lap_zp(const float* __restrict__ zf, const float* __restrict__ zs, int o,
                                        const Coef& co)
{
    return co.lcs * ((zs[o - W - 1] + zs[o - W + 1]) + (zs[o + W - 1] + zs[o + W + 1]))
         + co.lcf * ((zf[o - W - 1] + zf[o - W + 1]) + (zf[o + W - 1] + zf[o + W + 1]))
         + co.lf * ((zf[o - 1] + zf[o + 1]) + (zf[o - W] + zf[o + W])) + co.l0s * zs[o] + co.l0f * zf[o];
}

// D_x, D_y, D_z, L at position o of a plane with row stride W from the
// z-prepared arrays (each of the plane's size).
template <int W>
__device__ __forceinline__ Der4 der_zp(const float* __restrict__ zf, const float* __restrict__ zd,
                                       const float* __restrict__ zs, int o, const Coef& co)
{
    Der4 r;
    r.dx = co.e0 * (zs[o + 1] - zs[o - 1])
         + co.e1 * ((zs[o - W + 1] - zs[o - W - 1]) + (zs[o + W + 1] - zs[o + W - 1]));
    r.dy = co.e0 * (zs[o + W] - zs[o - W])
         + co.e1 * ((zs[o + W - 1] - zs[o - W - 1]) + (zs[o + W + 1] - zs[o - W + 1]));
    r.dz = co.f00 * zd[o] + co.f01 * ((zd[o - 1] + zd[o + 1]) + (zd[o - W] + zd[o + W]))
         + co.f11 * ((zd[o - W - 1] + zd[o - W + 1]) + (zd[o + W - 1] + zd[o + W + 1]));
    r.l = lap_zp<W>(zf, zs, o, co);
    return r;
}

// Values of the thread's own column at one plane, kept for the output stage.
struct Centre {
    float G[9];  // G[3a+b] = D_a P_b
    float mu0, phi, P[3], Dphi[3];
};

template <int TX, int TY, int MINB>
__global__ __launch_bounds__(TX * TY, MINB) void k_rhs(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                 const int* __restrict__ nwork_p, int parity, int kstep)
{
    constexpr int NT = TX * TY;
    constexpr int W2 = TX + 4, H2 = TY + 4, A2 = W2 * H2;  // inputs phi, P: 2-cell apron
    constexpr int W1 = TX + 2, H1 = TY + 2, A1 = W1 * H1;  // v and intermediates: 1-cell apron
    constexpr int NRING = A1 - NT;
    constexpr int MP2 = (A2 + NT - 1) / NT;  // input positions per thread
    constexpr int MP1 = (A1 + NT - 1) / NT;  // v positions per thread
    __shared__ float zp[4][NZ][A2];  // phi, Px, Py, Pz x (f, dz, sz)
    __shared__ float zv[3][NZ][A1];  // vx, vy, vz x (f, dz, sz)
    __shared__ float sM[13][A1];    // Sigma(9, [3a+b]), mu(3), mu0 of one plane

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int nwork = *nwork_p;

    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const Coef co = D.co;
        const int N = D.N;
        const int x0 = wk.a * TX, y0 = wk.b * TY;
        const int zb = wk.c & 0xffff;
        const int ze = wk.c >> 16;
        const float* __restrict__ gin[4] = {D.phi[parity], D.P[0], D.P[1], D.P[2]};
        const float* __restrict__ gv[3] = {D.v[0], D.v[1], D.v[2]};
        float* __restrict__ ophi = D.phi[parity ^ 1];
        float* __restrict__ ovs[3] = {D.vs[parity][0], D.vs[parity][1], D.vs[parity][2]};
        float* __restrict__ ops[3] = {D.ps[parity][0], D.ps[parity][1], D.ps[parity][2]};

        const int gx = x0 + tx, gy = y0 + ty;
        const bool own = (gx < N) && (gy < N);

        // xy offsets (plane index) of this thread's input and v positions
        size_t off2[MP2], off1[MP1];
#pragma unroll
        for (int k = 0; k < MP2; ++k) {
            const int i = min(tid + k * NT, A2 - 1);
            off2[k] = (size_t)wrapi(y0 - 2 + i / W2, N) * N + wrapi(x0 - 2 + i % W2, N);
        }
#pragma unroll
        for (int k = 0; k < MP1; ++k) {
            const int i = min(tid + k * NT, A1 - 1);
            off1[k] = (size_t)wrapi(y0 - 1 + i / W1, N) * N + wrapi(x0 - 1 + i % W1, N);
        }
        const size_t NN = (size_t)N * N;
        // three rolling planes in registers: [0] oldest ... [2] newest
        float in[MP2][4][3];
        float vin[MP1][3][3];
        auto load_in = [&](int slot, int z) {
            const size_t zo = (size_t)wrapi(z, N) * NN;
#pragma unroll
            for (int k = 0; k < MP2; ++k)
#pragma unroll
                for (int f = 0; f < 4; ++f) in[k][f][slot] = gin[f][zo + off2[k]];
        };
        auto load_v = [&](int slot, int z) {
            const size_t zo = (size_t)wrapi(z, N) * NN;
#pragma unroll
            for (int k = 0; k < MP1; ++k)
#pragma unroll
                for (int f = 0; f < 3; ++f) vin[k][f][slot] = gv[f][zo + off1[k]];
        };
        load_in(0, zb - 2);
        load_in(1, zb - 1);
        load_in(2, zb);
        load_v(0, zb - 3);
        load_v(1, zb - 2);
        load_v(2, zb - 1);

        // accumulators for output planes zm-1 (m), zm (c), zm+1 (p)
        float aSm[3] = {0.f, 0.f, 0.f}, aSc[3] = {0.f, 0.f, 0.f}, aSp[3] = {0.f, 0.f, 0.f};
        float aLm[3] = {0.f, 0.f, 0.f}, aLc[3] = {0.f, 0.f, 0.f}, aLp[3] = {0.f, 0.f, 0.f};
        float aL0m = 0.f, aL0c = 0.f, aL0p = 0.f;
        Centre cur, nxt;

        for (int zi = zb; zi <= ze + 1; ++zi) {
            // ---- 1. z-prepare plane zm = zi-1 (phi, P) and zo = zi-2 (v) ---
            __syncthreads();
#pragma unroll
            for (int k = 0; k < MP2; ++k) {
                const int i = tid + k * NT;
                if (i < A2) {
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        const float a = in[k][f][0], b = in[k][f][1], c = in[k][f][2];
                        const float cz = a + c;
                        zp[f][ZF][i] = b;
                        zp[f][ZD][i] = c - a;
                        zp[f][ZS][i] = co.sw0 * b + co.sw1 * cz;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < MP1; ++k) {
                const int i = tid + k * NT;
                if (i < A1) {
#pragma unroll
                    for (int f = 0; f < 3; ++f) {
                        const float a = vin[k][f][0], b = vin[k][f][1], c = vin[k][f][2];
                        const float cz = a + c;
                        zv[f][ZF][i] = b;
                        zv[f][ZD][i] = c - a;
                        zv[f][ZS][i] = co.sw0 * b + co.sw1 * cz;
                    }
                }
            }
            // roll and prefetch the next planes (zi+1 inputs, zi v)
#pragma unroll
            for (int k = 0; k < MP2; ++k)
#pragma unroll
                for (int f = 0; f < 4; ++f) { in[k][f][0] = in[k][f][1]; in[k][f][1] = in[k][f][2]; }
#pragma unroll
            for (int k = 0; k < MP1; ++k)
#pragma unroll
                for (int f = 0; f < 3; ++f) { vin[k][f][0] = vin[k][f][1]; vin[k][f][1] = vin[k][f][2]; }
            if (zi <= ze) {
                load_in(2, zi + 1);
                load_v(2, zi);
            }
            __syncthreads();

            // ---- 2. intermediates of plane zm over the 1-cell apron --------
            {
                // the thread's own cell first (keeps its centre values)
                const int o2 = (ty + 2) * W2 + (tx + 2);
                const int o1 = (ty + 1) * W1 + (tx + 1);
                const Der4 dphi = der_zp<W2>(zp[0][ZF], zp[0][ZD], zp[0][ZS], o2, co);
                const float ph = zp[0][ZF][o2];
                nxt.phi = ph;
                nxt.Dphi[0] = dphi.dx; nxt.Dphi[1] = dphi.dy; nxt.Dphi[2] = dphi.dz;
                const float ph2 = ph * ph;
                nxt.mu0 = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * dphi.l;
                sM[12][o1] = nxt.mu0;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const Der4 d = der_zp<W2>(zp[1 + b][ZF], zp[1 + b][ZD], zp[1 + b][ZS], o2, co);
                    const float pb = zp[1 + b][ZF][o2];
                    nxt.P[b] = pb;
                    nxt.G[0 * 3 + b] = d.dx;
                    nxt.G[1 * 3 + b] = d.dy;
                    nxt.G[2 * 3 + b] = d.dz;
                    sM[9 + b][o1] = ph * pb + co.La2 * d.l;
                    sM[0 * 3 + b][o1] = ph * d.dx;
                    sM[1 * 3 + b][o1] = ph * d.dy;
                    sM[2 * 3 + b][o1] = ph * d.dz;
                }
            }
            for (int r = tid; r < NRING; r += NT) {
                int lx1, ly1;
                if (r < W1) { lx1 = r; ly1 = 0; }
                else if (r < 2 * W1) { lx1 = r - W1; ly1 = H1 - 1; }
                else if (r < 2 * W1 + TY) { lx1 = 0; ly1 = 1 + (r - 2 * W1); }
                else { lx1 = W1 - 1; ly1 = 1 + (r - 2 * W1 - TY); }
                const int o2 = (ly1 + 1) * W2 + (lx1 + 1);
                const int o1 = ly1 * W1 + lx1;
                const float ph = zp[0][ZF][o2];
                const float ph2 = ph * ph;
                sM[12][o1] = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * lap_zp<W2>(zp[0][ZF], zp[0][ZS], o2, co);
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const Der4 d = der_zp<W2>(zp[1 + b][ZF], zp[1 + b][ZD], zp[1 + b][ZS], o2, co);
                    sM[9 + b][o1] = ph * zp[1 + b][ZF][o2] + co.La2 * d.l;
                    sM[0 * 3 + b][o1] = ph * d.dx;
                    sM[1 * 3 + b][o1] = ph * d.dy;
                    sM[2 * 3 + b][o1] = ph * d.dz;
                }
            }
            __syncthreads();

            // ---- 3. contribution of plane zm to outputs zm-1, zm, zm+1 ------
            {
                const int o = (ty + 1) * W1 + (tx + 1);
#define SM(f, dx, dy) sM[f][o + (dy) * W1 + (dx)]
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int fx = 0 * 3 + b, fy = 1 * 3 + b, fz = 2 * 3 + b;
                    const float ax = co.e0 * (SM(fx, 1, 0) - SM(fx, -1, 0))
                                   + co.e1 * ((SM(fx, 1, -1) - SM(fx, -1, -1)) + (SM(fx, 1, 1) - SM(fx, -1, 1)));
                    const float ay = co.e0 * (SM(fy, 0, 1) - SM(fy, 0, -1))
                                   + co.e1 * ((SM(fy, -1, 1) - SM(fy, -1, -1)) + (SM(fy, 1, 1) - SM(fy, 1, -1)));
                    const float A = ax + ay;
                    const float Bz = co.f00 * SM(fz, 0, 0)
                                   + co.f01 * ((SM(fz, -1, 0) + SM(fz, 1, 0)) + (SM(fz, 0, -1) + SM(fz, 0, 1)))
                                   + co.f11 * ((SM(fz, -1, -1) + SM(fz, 1, -1)) + (SM(fz, -1, 1) + SM(fz, 1, 1)));
                    aSm[b] += co.sw1 * A + Bz;
                    aSc[b] += co.sw0 * A;
                    aSp[b] += co.sw1 * A - Bz;
                    const int fm = 9 + b;
                    const float cen = SM(fm, 0, 0);
                    const float E = co.lc * ((SM(fm, -1, -1) + SM(fm, 1, -1)) + (SM(fm, -1, 1) + SM(fm, 1, 1))) + co.lf * cen;
                    const float F = co.lf * ((SM(fm, -1, 0) + SM(fm, 1, 0)) + (SM(fm, 0, -1) + SM(fm, 0, 1))) + co.l0 * cen;
                    aLm[b] += E;
                    aLc[b] += F;
                    aLp[b] += E;
                }
                {
                    const float cen = SM(12, 0, 0);
                    const float E = co.lc * ((SM(12, -1, -1) + SM(12, 1, -1)) + (SM(12, -1, 1) + SM(12, 1, 1))) + co.lf * cen;
                    const float F = co.lf * ((SM(12, -1, 0) + SM(12, 1, 0)) + (SM(12, 0, -1) + SM(12, 0, 1))) + co.l0 * cen;
                    aL0m += E;
                    aL0c += F;
                    aL0p += E;
                }
#undef SM
            }

            // ---- 4. finish output plane zo = zi-2 ---------------------------
            const int zo = zi - 2;
            if (zo >= zb && own) {
                const int o = (ty + 1) * W1 + (tx + 1);
                float vc[3], Dv[3][3], Lv[3];  // Dv[c][b] = D_c v_b
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const Der4 d = der_zp<W1>(zv[b][ZF], zv[b][ZD], zv[b][ZS], o, co);
                    vc[b] = zv[b][ZF][o];
                    Dv[0][b] = d.dx; Dv[1][b] = d.dy; Dv[2][b] = d.dz;
                    Lv[b] = d.l;
                }
                const float ph = cur.phi;
                const float P2 = cur.P[0] * cur.P[0] + cur.P[1] * cur.P[1] + cur.P[2] * cur.P[2];
                const size_t g = gidx(N, gx, gy, zo);
                const float dtphi = -(vc[0] * cur.Dphi[0] + vc[1] * cur.Dphi[1] + vc[2] * cur.Dphi[2]) + co.gamma * aL0m;
                ophi[g] = ph + co.Dt * dtphi;
                float dtp[3];
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const float dtv = -(vc[0] * Dv[0][b] + vc[1] * Dv[1][b] + vc[2] * Dv[2][b]) + co.mu * Lv[b]
                                    + cur.mu0 * cur.Dphi[b] + co.cA * aSm[b];
                    ovs[b][g] = vc[b] + co.Dt * dtv;
                    const float vgp = vc[0] * cur.G[0 * 3 + b] + vc[1] * cur.G[1 * 3 + b] + vc[2] * cur.G[2 * 3 + b];
                    const float pgp = cur.P[0] * cur.G[0 * 3 + b] + cur.P[1] * cur.G[1 * 3 + b] + cur.P[2] * cur.G[2 * 3 + b];
                    dtp[b] = -vgp - co.alpha * ph * cur.P[b] - co.beta * P2 * cur.P[b] + co.eta * aLm[b] - co.lambda0 * pgp;
                }
                if (co.anchoring) {
                    const float dP2 = cur.Dphi[0] * cur.Dphi[0] + cur.Dphi[1] * cur.Dphi[1] + cur.Dphi[2] * cur.Dphi[2];
                    const float sP = sqrtf(P2);
                    const float dpP = cur.Dphi[0] * cur.P[0] + cur.Dphi[1] * cur.P[1] + cur.Dphi[2] * cur.P[2];
#pragma unroll
                    for (int b = 0; b < 3; ++b) dtp[b] -= co.A * (dP2 * P2 * cur.P[b] + sP * cur.Dphi[b] * dpP);
                }
#pragma unroll
                for (int b = 0; b < 3; ++b) ops[b][g] = cur.P[b] + co.Dt * dtp[b];
            }

            // ---- rotate accumulators and held centre values -----------------
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                aSm[b] = aSc[b]; aSc[b] = aSp[b]; aSp[b] = 0.f;
                aLm[b] = aLc[b]; aLc[b] = aLp[b]; aLp[b] = 0.f;
            }
            aL0m = aL0c; aL0c = aL0p; aL0p = 0.f;
            cur = nxt;
        }
    }
}

// ---------------------------------------------------------------------------
template <int TX, int TY, int MINB>
static hipError_t launch_rhs_t(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                               int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_rhs<TX, TY, MINB>), dim3(grid), dim3(TX * TY), 0, st, drops, work, nwork_p, parity, kstep);
    return hipGetLastError();
}

// Tile variants: TX x TY cells per work-group and the work-groups per CU the
// register allocation targets.
struct RhsVar {
    int tx, ty, minb;
};
static const RhsVar kRhsVars[RHS_NVARIANTS] = {
    {32, 8, 2}, {16, 16, 2}, {64, 4, 2}, {16, 8, 2}, {32, 8, 3}, {32, 16, 1}, {64, 8, 1},
};

hipError_t launch_rhs(int variant, const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                      int grid, hipStream_t st)
{
    switch (variant) {
    case 0: return launch_rhs_t<32, 8, 2>(drops, work, nwork_p, parity, kstep, grid, st);
    case 1: return launch_rhs_t<16, 16, 2>(drops, work, nwork_p, parity, kstep, grid, st);
    case 2: return launch_rhs_t<64, 4, 2>(drops, work, nwork_p, parity, kstep, grid, st);
    case 3: return launch_rhs_t<16, 8, 2>(drops, work, nwork_p, parity, kstep, grid, st);
    case 4: return launch_rhs_t<32, 8, 3>(drops, work, nwork_p, parity, kstep, grid, st);
    case 5: return launch_rhs_t<32, 16, 1>(drops, work, nwork_p, parity, kstep, grid, st);
    case 6: return launch_rhs_t<64, 8, 1>(drops, work, nwork_p, parity, kstep, grid, st);
    default: return hipErrorInvalidValue;
    }
}

void rhs_tile(int variant, int* tx, int* ty)
{
    if (variant >= 0 && variant < RHS_NVARIANTS) {
        *tx = kRhsVars[variant].tx;
        *ty = kRhsVars[variant].ty;
    } else {
        *tx = 0;
        *ty = 0;
    }
}

}  // namespace ttsh
