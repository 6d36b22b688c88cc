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
// Stencils (TTSH3D.cu:3280-3343) are used in factored form:
//   D_x = 1/(2h) dx (x) S_y (x) S_z with S = [1,4,1]/6, and the 15-point L.
//
// Schedule: a work-group owns a TX x TY column of cells and marches through a
// z-range.  Each iteration loads one new input plane (phi, P with a 2-cell xy
// apron; v with a 1-cell apron) into rolling LDS slots, computes the 13
// intermediates (Sigma, mu, mu0) of the plane behind it over the 1-cell apron
// into LDS, adds that plane's contribution to the D.Sigma, L mu and L mu0 of
// the three output planes it touches (kept in registers), and finishes the
// output plane two behind the input plane.
#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {

#define V27(v, dx, dy, dz) v[((dz) + 1) * 9 + ((dy) + 1) * 3 + ((dx) + 1)]

struct Der4 {
    float dx, dy, dz, l;
};

// D_x, D_y, D_z and L of one 27-point neighbourhood.
__device__ __forceinline__ Der4 der27(const float* v, const Coef& co)
{
    Der4 r;
    float a[3], b[3], c[3];
#pragma unroll
    for (int k = -1; k <= 1; ++k) {
        a[k + 1] = co.sw0 * (V27(v, 1, 0, k) - V27(v, -1, 0, k))
                 + co.sw1 * ((V27(v, 1, -1, k) - V27(v, -1, -1, k)) + (V27(v, 1, 1, k) - V27(v, -1, 1, k)));
        b[k + 1] = co.sw0 * (V27(v, 0, 1, k) - V27(v, 0, -1, k))
                 + co.sw1 * ((V27(v, -1, 1, k) - V27(v, -1, -1, k)) + (V27(v, 1, 1, k) - V27(v, 1, -1, k)));
        c[k + 1] = co.sw0 * (V27(v, 0, k, 1) - V27(v, 0, k, -1))
                 + co.sw1 * ((V27(v, -1, k, 1) - V27(v, -1, k, -1)) + (V27(v, 1, k, 1) - V27(v, 1, k, -1)));
    }
    r.dx = co.d1 * (co.sw0 * a[1] + co.sw1 * (a[0] + a[2]));
    r.dy = co.d1 * (co.sw0 * b[1] + co.sw1 * (b[0] + b[2]));
    r.dz = co.d1 * (co.sw0 * c[1] + co.sw1 * (c[0] + c[2]));
    const float corners = ((V27(v, 1, 1, -1) + V27(v, 1, -1, -1)) + (V27(v, -1, 1, -1) + V27(v, -1, -1, -1)))
                        + ((V27(v, 1, 1, 1) + V27(v, 1, -1, 1)) + (V27(v, -1, 1, 1) + V27(v, -1, -1, 1)));
    const float faces = (V27(v, 0, 0, -1) + V27(v, 0, 0, 1)) + (V27(v, -1, 0, 0) + V27(v, 1, 0, 0))
                      + (V27(v, 0, -1, 0) + V27(v, 0, 1, 0));
    r.l = co.lc * corners + co.lf * faces + co.l0 * V27(v, 0, 0, 0);
    return r;
}

// L only (15 points).
__device__ __forceinline__ float lap27(const float* v, const Coef& co)
{
    const float corners = ((V27(v, 1, 1, -1) + V27(v, 1, -1, -1)) + (V27(v, -1, 1, -1) + V27(v, -1, -1, -1)))
                        + ((V27(v, 1, 1, 1) + V27(v, 1, -1, 1)) + (V27(v, -1, 1, 1) + V27(v, -1, -1, 1)));
    const float faces = (V27(v, 0, 0, -1) + V27(v, 0, 0, 1)) + (V27(v, -1, 0, 0) + V27(v, 1, 0, 0))
                      + (V27(v, 0, -1, 0) + V27(v, 0, 1, 0));
    return co.lc * corners + co.lf * faces + co.l0 * V27(v, 0, 0, 0);
}

template <int W>
__device__ __forceinline__ void load27(const float* __restrict__ zm, const float* __restrict__ z0,
                                       const float* __restrict__ zp, int o, float* v)
{
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            V27(v, dx, dy, -1) = zm[o + dy * W + dx];
            V27(v, dx, dy, 0) = z0[o + dy * W + dx];
            V27(v, dx, dy, 1) = zp[o + dy * W + dx];
        }
}

// Values of the thread's own column at one plane, kept for the output stage.
struct Centre {
    float G[9];  // G[3a+b] = D_a P_b
    float mu0, phi, P[3], Dphi[3];
};

template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_rhs(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                 const int* __restrict__ nwork_p, int parity, int kstep)
{
    constexpr int NT = TX * TY;
    constexpr int W2 = TX + 4, H2 = TY + 4, A2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, A1 = W1 * H1;
    constexpr int NRING = A1 - NT;
    __shared__ float sIn[3][4][A2];  // phi, Px, Py, Pz; 3 rolling planes
    __shared__ float sV[3][3][A1];   // vx, vy, vz; 3 rolling planes
    __shared__ float sM[13][A1];     // Sigma(9, [3a+b]), mu(3), mu0 of one plane

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int nwork = *nwork_p;

    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const Coef co = D.co;
        const int N = D.N;
        const int x0 = wk.a * TX, y0 = wk.b * TY;
        const int zb = wk.c & 0xffff;
        const int ze = wk.c >> 16;
        const float* __restrict__ gphi = D.phi[parity];
        const float* __restrict__ gP0 = D.P[0];
        const float* __restrict__ gP1 = D.P[1];
        const float* __restrict__ gP2 = D.P[2];
        const float* __restrict__ gv0 = D.v[0];
        const float* __restrict__ gv1 = D.v[1];
        const float* __restrict__ gv2 = D.v[2];
        float* __restrict__ ophi = D.phi[parity ^ 1];
        float* __restrict__ ovs0 = D.vs[parity][0];
        float* __restrict__ ovs1 = D.vs[parity][1];
        float* __restrict__ ovs2 = D.vs[parity][2];
        float* __restrict__ ops0 = D.ps[parity][0];
        float* __restrict__ ops1 = D.ps[parity][1];
        float* __restrict__ ops2 = D.ps[parity][2];

        const int gx = x0 + tx, gy = y0 + ty;
        const bool own = (gx < N) && (gy < N);

        // accumulators for output planes zm-1 (m), zm (c), zm+1 (p)
        float aSm[3] = {0.f, 0.f, 0.f}, aSc[3] = {0.f, 0.f, 0.f}, aSp[3] = {0.f, 0.f, 0.f};
        float aLm[3] = {0.f, 0.f, 0.f}, aLc[3] = {0.f, 0.f, 0.f}, aLp[3] = {0.f, 0.f, 0.f};
        float aL0m = 0.f, aL0c = 0.f, aL0p = 0.f;
        Centre cur, nxt;

        int s_m2 = 0, s_m1 = 1, s_0 = 2;  // sIn slots of planes zi-2, zi-1, zi
        int v_m2 = 0, v_m1 = 1, v_0 = 2;  // sV slots of planes zi-3, zi-2, zi-1

        for (int zi = zb - 2; zi <= ze + 1; ++zi) {
            // ---- 1. load input plane zi and v plane zi-1 -------------------
            __syncthreads();
            {
                const int gz = wrapi(zi, N);
                for (int i = tid; i < A2; i += NT) {
                    const int lx = i % W2, ly = i / W2;
                    const size_t g = gidx(N, wrapi(x0 - 2 + lx, N), wrapi(y0 - 2 + ly, N), gz);
                    sIn[s_0][0][i] = gphi[g];
                    sIn[s_0][1][i] = gP0[g];
                    sIn[s_0][2][i] = gP1[g];
                    sIn[s_0][3][i] = gP2[g];
                }
                if (zi >= zb) {
                    const int gzv = wrapi(zi - 1, N);
                    for (int i = tid; i < A1; i += NT) {
                        const int lx = i % W1, ly = i / W1;
                        const size_t g = gidx(N, wrapi(x0 - 1 + lx, N), wrapi(y0 - 1 + ly, N), gzv);
                        sV[v_0][0][i] = gv0[g];
                        sV[v_0][1][i] = gv1[g];
                        sV[v_0][2][i] = gv2[g];
                    }
                }
            }
            __syncthreads();

            if (zi >= zb) {
                // ---- 2. intermediates of plane zm = zi-1 over the 1-cell apron
                // pass over the thread's own cell (keeps its centre values)
                {
                    const int o2 = (ty + 2) * W2 + (tx + 2);
                    const int o1 = (ty + 1) * W1 + (tx + 1);
                    float v[27];
                    load27<W2>(sIn[s_m2][0], sIn[s_m1][0], sIn[s_0][0], o2, v);
                    const Der4 dphi = der27(v, co);
                    const float ph = V27(v, 0, 0, 0);
                    nxt.phi = ph;
                    nxt.Dphi[0] = dphi.dx; nxt.Dphi[1] = dphi.dy; nxt.Dphi[2] = dphi.dz;
                    const float ph2 = ph * ph;
                    nxt.mu0 = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * dphi.l;
                    sM[12][o1] = nxt.mu0;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        load27<W2>(sIn[s_m2][1 + b], sIn[s_m1][1 + b], sIn[s_0][1 + b], o2, v);
                        const Der4 d = der27(v, co);
                        const float pb = V27(v, 0, 0, 0);
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
                // pass over the apron ring
                for (int r = tid; r < NRING; r += NT) {
                    int lx1, ly1;
                    if (r < W1) { lx1 = r; ly1 = 0; }
                    else if (r < 2 * W1) { lx1 = r - W1; ly1 = H1 - 1; }
                    else if (r < 2 * W1 + TY) { lx1 = 0; ly1 = 1 + (r - 2 * W1); }
                    else { lx1 = W1 - 1; ly1 = 1 + (r - 2 * W1 - TY); }
                    const int o2 = (ly1 + 1) * W2 + (lx1 + 1);
                    const int o1 = ly1 * W1 + lx1;
                    float v[27];
                    load27<W2>(sIn[s_m2][0], sIn[s_m1][0], sIn[s_0][0], o2, v);
                    const float ph = V27(v, 0, 0, 0);
                    const float ph2 = ph * ph;
                    sM[12][o1] = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * lap27(v, co);
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        load27<W2>(sIn[s_m2][1 + b], sIn[s_m1][1 + b], sIn[s_0][1 + b], o2, v);
                        const Der4 d = der27(v, co);
                        sM[9 + b][o1] = ph * V27(v, 0, 0, 0) + co.La2 * d.l;
                        sM[0 * 3 + b][o1] = ph * d.dx;
                        sM[1 * 3 + b][o1] = ph * d.dy;
                        sM[2 * 3 + b][o1] = ph * d.dz;
                    }
                }
                __syncthreads();

                // ---- 3. contribution of plane zm to outputs zm-1, zm, zm+1 ---
                {
                    const int o = (ty + 1) * W1 + (tx + 1);
#define SM(f, dx, dy) sM[f][o + (dy) * W1 + (dx)]
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const int fx = 0 * 3 + b, fy = 1 * 3 + b, fz = 2 * 3 + b;
                        const float ax = co.sw0 * (SM(fx, 1, 0) - SM(fx, -1, 0))
                                       + co.sw1 * ((SM(fx, 1, -1) - SM(fx, -1, -1)) + (SM(fx, 1, 1) - SM(fx, -1, 1)));
                        const float ay = co.sw0 * (SM(fy, 0, 1) - SM(fy, 0, -1))
                                       + co.sw1 * ((SM(fy, -1, 1) - SM(fy, -1, -1)) + (SM(fy, 1, 1) - SM(fy, 1, -1)));
                        const float rz0 = co.sw0 * SM(fz, 0, 0) + co.sw1 * (SM(fz, -1, 0) + SM(fz, 1, 0));
                        const float rzm = co.sw0 * SM(fz, 0, -1) + co.sw1 * (SM(fz, -1, -1) + SM(fz, 1, -1));
                        const float rzp = co.sw0 * SM(fz, 0, 1) + co.sw1 * (SM(fz, -1, 1) + SM(fz, 1, 1));
                        const float A = co.d1 * (ax + ay);
                        const float Bz = co.d1 * (co.sw0 * rz0 + co.sw1 * (rzm + rzp));
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

                // ---- 4. finish output plane zo = zi-2 -------------------------
                const int zo = zi - 2;
                if (zo >= zb && own) {
                    const int o = (ty + 1) * W1 + (tx + 1);
                    float v[27];
                    float vc[3], Dv[3][3], Lv[3];  // Dv[c][b] = D_c v_b
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        load27<W1>(sV[v_m2][b], sV[v_m1][b], sV[v_0][b], o, v);
                        const Der4 d = der27(v, co);
                        vc[b] = V27(v, 0, 0, 0);
                        Dv[0][b] = d.dx; Dv[1][b] = d.dy; Dv[2][b] = d.dz;
                        Lv[b] = d.l;
                    }
                    const float ph = cur.phi;
                    const float P2 = cur.P[0] * cur.P[0] + cur.P[1] * cur.P[1] + cur.P[2] * cur.P[2];
                    const size_t g = gidx(N, gx, gy, zo);
                    const float dtphi = -(vc[0] * cur.Dphi[0] + vc[1] * cur.Dphi[1] + vc[2] * cur.Dphi[2]) + co.gamma * aL0m;
                    ophi[g] = ph + co.Dt * dtphi;
                    float dtv[3], dtp[3];
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        dtv[b] = -(vc[0] * Dv[0][b] + vc[1] * Dv[1][b] + vc[2] * Dv[2][b]) + co.mu * Lv[b]
                               + cur.mu0 * cur.Dphi[b] + co.cA * aSm[b];
                        const float vgp = vc[0] * cur.G[0 * 3 + b] + vc[1] * cur.G[1 * 3 + b] + vc[2] * cur.G[2 * 3 + b];
                        const float pgp = cur.P[0] * cur.G[0 * 3 + b] + cur.P[1] * cur.G[1 * 3 + b] + cur.P[2] * cur.G[2 * 3 + b];
                        dtp[b] = -vgp - co.alpha * ph * cur.P[b] - co.beta * P2 * cur.P[b] + co.eta * aLm[b] - co.lambda0 * pgp;
                    }
                    if (co.anchoring) {
                        const float dP2 = cur.Dphi[0] * cur.Dphi[0] + cur.Dphi[1] * cur.Dphi[1] + cur.Dphi[2] * cur.Dphi[2];
                        const float sP = sqrtf(P2);
                        const float dpP = cur.Dphi[0] * cur.P[0] + cur.Dphi[1] * cur.P[1] + cur.Dphi[2] * cur.P[2];
#pragma unroll
                        for (int b = 0; b < 3; ++b)
                            dtp[b] -= co.A * (dP2 * P2 * cur.P[b] + sP * cur.Dphi[b] * dpP);
                    }
                    ovs0[g] = vc[0] + co.Dt * dtv[0];
                    ovs1[g] = vc[1] + co.Dt * dtv[1];
                    ovs2[g] = vc[2] + co.Dt * dtv[2];
                    ops0[g] = cur.P[0] + co.Dt * dtp[0];
                    ops1[g] = cur.P[1] + co.Dt * dtp[1];
                    ops2[g] = cur.P[2] + co.Dt * dtp[2];
                }

                // ---- rotate accumulators and held centre values ---------------
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    aSm[b] = aSc[b]; aSc[b] = aSp[b]; aSp[b] = 0.f;
                    aLm[b] = aLc[b]; aLc[b] = aLp[b]; aLp[b] = 0.f;
                }
                aL0m = aL0c; aL0c = aL0p; aL0p = 0.f;
                cur = nxt;
                const int tv = v_m2; v_m2 = v_m1; v_m1 = v_0; v_0 = tv;
            }
            const int t = s_m2; s_m2 = s_m1; s_m1 = s_0; s_0 = t;
        }
    }
}

// ---------------------------------------------------------------------------
template <int TX, int TY>
static hipError_t launch_rhs_t(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                               int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_rhs<TX, TY>), dim3(grid), dim3(TX * TY), 0, st, drops, work, nwork_p, parity, kstep);
    return hipGetLastError();
}

hipError_t launch_rhs(int variant, const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                      int grid, hipStream_t st)
{
    switch (variant) {
    case RHS_32x8: return launch_rhs_t<32, 8>(drops, work, nwork_p, parity, kstep, grid, st);
    case RHS_16x16: return launch_rhs_t<16, 16>(drops, work, nwork_p, parity, kstep, grid, st);
    case RHS_64x4: return launch_rhs_t<64, 4>(drops, work, nwork_p, parity, kstep, grid, st);
    case RHS_16x8: return launch_rhs_t<16, 8>(drops, work, nwork_p, parity, kstep, grid, st);
    default: return hipErrorInvalidValue;
    }
}

void rhs_tile(int variant, int* tx, int* ty)
{
    switch (variant) {
    case RHS_32x8: *tx = 32; *ty = 8; break;
    case RHS_16x16: *tx = 16; *ty = 16; break;
    case RHS_64x4: *tx = 64; *ty = 4; break;
    case RHS_16x8: *tx = 16; *ty = 8; break;
    default: *tx = 0; *ty = 0; break;
    }
}

}  // namespace ttsh
