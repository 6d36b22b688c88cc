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
#include <type_traits>

#include "ttsh_dev.h"
#include "ttsh_kernels.h"

namespace ttsh {

// The three z-prepared arrays of one field: f, f(z+1)-f(z-1), S_z f.
enum { ZF = 0, ZD = 1, ZS = 2, NZ = 3 };

struct Der4 {
    float dx, dy, dz, l;
};

// LDS tiles hold the arrays of one position together (array-of-structs) with
// an odd stride S in floats, so lanes at consecutive positions hit distinct
// banks.  A stencil reads three rows through three row pointers, each at the
// x-1 neighbour of its row: every value of a row then lies within 1 KiB of
// the row pointer, the reach of a paired LDS read's immediate offsets, so no
// per-pair base addresses are needed.
using lcptr = __attribute__((address_space(3))) const float*;

struct Rows {
    lcptr m;  // row y-1, position x-1
    lcptr c;  // row y
    lcptr p;  // row y+1
};

// Row pointers of position pos in a tile with row stride W and position
// stride S.  The LDS byte addresses, array base included, are made opaque to
// the compiler so that it does not re-derive them from one far base.
template <int W, int S>
__device__ __forceinline__ Rows rows_at(const float* base, int pos)
{
    const unsigned b0 = (unsigned)(size_t)(lcptr)base;
    unsigned a = b0 + 4u * (unsigned)((pos - W - 1) * S);
    unsigned b = b0 + 4u * (unsigned)((pos - 1) * S);
    unsigned c = b0 + 4u * (unsigned)((pos + W - 1) * S);
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c));
    return {(lcptr)(size_t)a, (lcptr)(size_t)b, (lcptr)(size_t)c};
}

// Laplacian from the f and S_z f arrays (field offset fo within a position).
template <int S>
__device__ __forceinline__ float lap_rows(const Rows& r, int fo, const Coef& co)
{
    auto F = [&](lcptr q, int dx) { return q[(dx + 1) * S + fo + ZF]; };
    auto Z = [&](lcptr q, int dx) { return q[(dx + 1) * S + fo + ZS]; };
    return co.lcs * ((Z(r.m, -1) + Z(r.m, 1)) + (Z(r.p, -1) + Z(r.p, 1)))
         + co.lcf * ((F(r.m, -1) + F(r.m, 1)) + (F(r.p, -1) + F(r.p, 1)))
         + co.lf * ((F(r.c, -1) + F(r.c, 1)) + (F(r.m, 0) + F(r.p, 0))) + co.l0s * Z(r.c, 0) + co.l0f * F(r.c, 0);
}

// D_x, D_y, D_z, L of the field at offset fo.
template <int S>
__device__ __forceinline__ Der4 der_rows(const Rows& r, int fo, const Coef& co)
{
    auto Z = [&](lcptr q, int dx) { return q[(dx + 1) * S + fo + ZS]; };
    auto D = [&](lcptr q, int dx) { return q[(dx + 1) * S + fo + ZD]; };
    Der4 d;
    d.dx = co.e0 * (Z(r.c, 1) - Z(r.c, -1)) + co.e1 * ((Z(r.m, 1) - Z(r.m, -1)) + (Z(r.p, 1) - Z(r.p, -1)));
    d.dy = co.e0 * (Z(r.p, 0) - Z(r.m, 0)) + co.e1 * ((Z(r.p, -1) - Z(r.m, -1)) + (Z(r.p, 1) - Z(r.m, 1)));
    d.dz = co.f00 * D(r.c, 0) + co.f01 * ((D(r.c, -1) + D(r.c, 1)) + (D(r.m, 0) + D(r.p, 0)))
         + co.f11 * ((D(r.m, -1) + D(r.m, 1)) + (D(r.p, -1) + D(r.p, 1)));
    d.l = lap_rows<S>(r, fo, co);
    return d;
}


template <int TX, int TY, int WPS>
__global__ __launch_bounds__(TX * TY, WPS) void k_rhs(const Drop* __restrict__ drops, const Work* __restrict__ work,
                                                 const int* __restrict__ nwork_p, int parity, int kstep)
{
    constexpr int NT = TX * TY;
    constexpr int W2 = TX + 4, H2 = TY + 4, A2 = W2 * H2;  // phi, P, v planes: 2-cell apron layout
    constexpr int W1 = TX + 2, H1 = TY + 2, A1 = W1 * H1;  // v and intermediates: 1-cell apron
    constexpr int NRING = A1 - NT;
    static_assert(NRING <= NT, "one ring position per thread at most");
    // apron rings around the tile and the loads per thread that cover them
    // as (position, field) pairs
    constexpr int R2 = A2 - NT, R1 = A1 - NT;
    constexpr int NR2 = (4 * R2 + NT - 1) / NT, NR1 = (3 * R1 + NT - 1) / NT;
    // per position: phi, Px, Py, Pz x (f, dz, sz) | vx, vy, vz x (f, dz, sz) |
    // Sigma(9, [3a+b]), mu(3), mu0 of one plane
    constexpr int SP = 13, SV = 9, SS = 13;
    // Tiles of which LDS fits one work-group per CU (WPS 1: two waves per
    // SIMD at most) can spend registers on unrolling the z-march by its
    // 3-slot rotation, which removes the register moves of the rotation.
    constexpr bool kUnroll3 = WPS == 1;
    __shared__ float zp[A2 * SP];
    __shared__ float zv[A1 * SV];
    __shared__ float sM[A1 * SS];

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int o2c = (ty + 2) * W2 + (tx + 2);  // own cell, 2-cell-apron layout
    const int o1c = (ty + 1) * W1 + (tx + 1);  // own cell, 1-cell-apron layout
    const int nwork = *nwork_p;

    for (int v = blockIdx.x; v < xcd_span(nwork); v += gridDim.x) {
        const int w = xcd_item(v, nwork);
        if (w >= nwork) continue;
        const Work wk = work[w];
        const Drop& D = drops[wk.d];
        if (kstep >= D.nsteps) continue;
        const Coef co = D.co;
        const int N = D.N;
        const unsigned NN = (unsigned)N * (unsigned)N;
        const int x0 = wk.a * TX, y0 = wk.b * TY;
        const int zb = wk.c & 0xffff;
        const int ze = wk.c >> 16;
        const gptr<const float> gin[4] = {glb<const float>(D.phi[parity]), glb<const float>(D.P[0]),
                                          glb<const float>(D.P[1]), glb<const float>(D.P[2])};
        const gptr<const float> gv[3] = {glb<const float>(D.v[0]), glb<const float>(D.v[1]),
                                         glb<const float>(D.v[2])};
        const gptr<float> ophi = glb(D.phi[parity ^ 1]);
        const gptr<float> ovs[3] = {glb(D.vs[parity][0]), glb(D.vs[parity][1]), glb(D.vs[parity][2])};
        const gptr<float> ops[3] = {glb(D.ps[parity][0]), glb(D.ps[parity][1]), glb(D.ps[parity][2])};

        const int gx = x0 + tx, gy = y0 + ty;
        const bool own = (gx < N) && (gy < N);
        const unsigned gown = 4u * ((unsigned)gy * (unsigned)N + (unsigned)gx);

        // Loads of a plane: the thread's own cell (a wave loads whole aligned
        // rows), one per field, and the apron ring around the tile as
        // (position, field) pairs spread over NR2 (phi, P) and NR1 (v) loads
        // per thread.  Ring loads address the fields from the lowest field
        // pointer of their group (one allocation: 32-bit offsets).
        gptr<const float> gb2 = gin[0], gb1 = gv[0];
#pragma unroll
        for (int f = 1; f < 4; ++f) gb2 = gin[f] < gb2 ? gin[f] : gb2;
#pragma unroll
        for (int f = 1; f < 3; ++f) gb1 = gv[f] < gb1 ? gv[f] : gb1;
        const unsigned own2 = 4u * ((unsigned)wrapi(gy, N) * (unsigned)N + (unsigned)wrapi(gx, N));
        const unsigned own1 = own2;
        unsigned roff2[NR2], roff1[NR1];
        int rl2[NR2], rl1[NR1];  // LDS float index of the pair's f array (-1: past the last pair)
#pragma unroll
        for (int m = 0; m < NR2; ++m) {
            const int j = min(tid + m * NT, 4 * R2 - 1), f = j / R2, r = j % R2;
            int i;
            if (r < 2 * W2) i = r;                                                  // rows y0-2, y0-1
            else if (r < 4 * W2) i = (TY + 2) * W2 + (r - 2 * W2);                  // rows y0+TY, +TY+1
            else if (r < 4 * W2 + 2 * TY) i = (2 + (r - 4 * W2) / 2) * W2 + (r - 4 * W2) % 2;
            else i = (2 + (r - 4 * W2 - 2 * TY) / 2) * W2 + W2 - 2 + (r - 4 * W2 - 2 * TY) % 2;
            roff2[m] = 4u * (unsigned)(gin[f] - gb2)
                     + 4u * ((unsigned)wrapi(y0 - 2 + i / W2, N) * (unsigned)N + (unsigned)wrapi(x0 - 2 + i % W2, N));
            rl2[m] = tid + m * NT < 4 * R2 ? i * SP + f * NZ : -1;
        }
#pragma unroll
        for (int m = 0; m < NR1; ++m) {
            const int j = min(tid + m * NT, 3 * R1 - 1), f = j / R1, r = j % R1;
            int i;
            if (r < W1) i = r;                                       // row y0-1
            else if (r < 2 * W1) i = (TY + 1) * W1 + (r - W1);       // row y0+TY
            else if (r < 2 * W1 + TY) i = (r - 2 * W1 + 1) * W1;     // column x0-1
            else i = (r - 2 * W1 - TY + 1) * W1 + (W1 - 1);          // column x0+TX
            roff1[m] = 4u * (unsigned)(gv[f] - gb1)
                     + 4u * ((unsigned)wrapi(y0 - 1 + i / W1, N) * (unsigned)N + (unsigned)wrapi(x0 - 1 + i % W1, N));
            rl1[m] = tid + m * NT < 3 * R1 ? i * SV + f * NZ : -1;
        }
        // three z-planes per value in registers, in slots rotating with the
        // z-march: in phase K the slots K, K+1, K+2 (mod 3) hold planes zi-2,
        // zi-1, zi
        float in[4][3], inr[NR2][3];
        float vin[3][3], vinr[NR1][3];
        auto load_plane = [&](int slot, int z, unsigned o2, const unsigned* r2, unsigned o1, const unsigned* r1) {
            const size_t zo = (size_t)wrapi(z, N) * NN;
#pragma unroll
            for (int f = 0; f < 4; ++f) in[f][slot] = ldb(gin[f] + zo, o2);
#pragma unroll
            for (int m = 0; m < NR2; ++m) inr[m][slot] = ldb(gb2 + zo, r2[m]);
#pragma unroll
            for (int f = 0; f < 3; ++f) vin[f][slot] = ldb(gv[f] + zo, o1);
#pragma unroll
            for (int m = 0; m < NR1; ++m) vinr[m][slot] = ldb(gb1 + zo, r1[m]);
        };
        load_plane(0, zb - 2, own2, roff2, own1, roff1);
        load_plane(1, zb - 1, own2, roff2, own1, roff1);
        load_plane(2, zb, own2, roff2, own1, roff1);
        wait_vmem();

        // D.Sigma, L mu, L mu0 accumulated per output plane, slot-rotated like
        // the inputs: in phase K slot K is plane zm-1, K+1 is zm, K+2 is zm+1
        float aS[3][3], aL[3][3], aL0[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            aL0[q] = 0.f;
#pragma unroll
            for (int b = 0; b < 3; ++b) { aS[q][b] = 0.f; aL[q][b] = 0.f; }
        }
        // pointwise part of the update of the own cell, per plane:
        //   [b]   v_b + Dt (-v.D v_b + mu L v_b + mu0 D_b phi)
        //   [3+b] P_b + Dt (-v.D P_b - alpha phi P_b - beta |P|^2 P_b - lambda0 P.D P_b - anchoring_b)
        //   [6]   phi + Dt (-v.D phi)
        float cur[7], nxt[7];

        auto body = [&](auto ph, int zi) {
            constexpr int S0 = decltype(ph)::value, S1 = (S0 + 1) % 3, S2 = (S0 + 2) % 3;
            // LDS positions made opaque per plane: the compiler then folds each
            // neighbour's constant offset into the LDS instruction instead of
            // hoisting dozens of loop-invariant addresses into registers
            // (the stencils' row pointers are opaque the same way)
            int ti = tid;
            asm volatile("" : "+v"(ti));
            ti &= 0x7fff;  // known non-negative again (a no-op)
            // likewise the in-plane global byte offsets, so that loads and
            // stores take a scalar plane base plus this 32-bit offset
            unsigned go = gown, o2 = own2, o1 = own1, q2[NR2], q1[NR1];
            asm volatile("" : "+v"(go), "+v"(o2), "+v"(o1));
#pragma unroll
            for (int m = 0; m < NR2; ++m) {
                q2[m] = roff2[m];
                asm volatile("" : "+v"(q2[m]));
            }
#pragma unroll
            for (int m = 0; m < NR1; ++m) {
                q1[m] = roff1[m];
                asm volatile("" : "+v"(q1[m]));
            }
            // ---- 1. z-prepare plane zm = zi-1 ------------------------------
            // Two barriers per plane suffice: zp and zv are read only in stage
            // 2, which every thread ends before the previous plane's barrier
            // ahead of stage 3; sM is written in stage 2 and read in stage 3,
            // and the barrier after stage 1 orders them across planes.
            // f, f(z+1)-f(z-1), S_z f of one value into LDS at float index q
            auto zprep = [&](float* q, const float* t) {
                const float a = t[S0], b = t[S1], c = t[S2];
                q[ZF] = b;
                q[ZD] = c - a;
                q[ZS] = co.sw0 * b + co.sw1 * (a + c);
            };
            {
                int i2 = o2c * SP, i1 = o1c * SV;
                asm volatile("" : "+v"(i2), "+v"(i1));
#pragma unroll
                for (int f = 0; f < 4; ++f) zprep(zp + i2 + f * NZ, in[f]);
#pragma unroll
                for (int f = 0; f < 3; ++f) zprep(zv + i1 + f * NZ, vin[f]);
            }
#pragma unroll
            for (int m = 0; m < NR2; ++m) {
                int i = rl2[m];
                asm volatile("" : "+v"(i));
                if (i >= 0) zprep(zp + i, inr[m]);
            }
#pragma unroll
            for (int m = 0; m < NR1; ++m) {
                int i = rl1[m];
                asm volatile("" : "+v"(i));
                if (i >= 0) zprep(zv + i, vinr[m]);
            }
            if constexpr (kUnroll3) {
                // prefetch plane zi+1 into the slot of plane zi-2
                if (zi <= ze) load_plane(S0, zi + 1, o2, q2, o1, q1);
            } else {
                // roll the planes and prefetch plane zi+1 into the newest slot
                auto roll = [](float* t) { t[0] = t[1]; t[1] = t[2]; };
#pragma unroll
                for (int f = 0; f < 4; ++f) roll(in[f]);
#pragma unroll
                for (int m = 0; m < NR2; ++m) roll(inr[m]);
#pragma unroll
                for (int f = 0; f < 3; ++f) roll(vin[f]);
#pragma unroll
                for (int m = 0; m < NR1; ++m) roll(vinr[m]);
                if (zi <= ze) load_plane(2, zi + 1, o2, q2, o1, q1);
            }
            __syncthreads();

            // ---- 2. intermediates of plane zm over the 1-cell apron --------
            {
                // own cell: intermediates and the pointwise update terms
                const Rows rp = rows_at<W2, SP>(zp, o2c);
                const Rows rv = rows_at<W1, SV>(zv, o1c);
                float* const mo = sM + (size_t)o1c * SS;
                const Der4 dph = der_rows<SP>(rp, 0, co);
                const float ph = rp.c[SP + ZF];
                const float ph2 = ph * ph;
                const float mu0 = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * dph.l;
                mo[12] = mu0;
                __builtin_amdgcn_sched_barrier(0);
                float Pc[3], vc[3];
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    Pc[b] = rp.c[SP + (1 + b) * NZ + ZF];
                    vc[b] = rv.c[SV + b * NZ + ZF];
                }
                const float P2 = Pc[0] * Pc[0] + Pc[1] * Pc[1] + Pc[2] * Pc[2];
                float an[3] = {0.f, 0.f, 0.f};
                if (co.anchoring) {
                    const float dP2 = dph.dx * dph.dx + dph.dy * dph.dy + dph.dz * dph.dz;
                    const float sP = sqrtf(P2);
                    const float dpP = dph.dx * Pc[0] + dph.dy * Pc[1] + dph.dz * Pc[2];
                    const float Dp[3] = {dph.dx, dph.dy, dph.dz};
#pragma unroll
                    for (int b = 0; b < 3; ++b) an[b] = co.A * (dP2 * P2 * Pc[b] + sP * Dp[b] * dpP);
                }
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const Der4 d = der_rows<SP>(rp, (1 + b) * NZ, co);
                    mo[9 + b] = ph * Pc[b] + co.La2 * d.l;
                    mo[0 * 3 + b] = ph * d.dx;
                    mo[1 * 3 + b] = ph * d.dy;
                    mo[2 * 3 + b] = ph * d.dz;
                    const float vgp = vc[0] * d.dx + vc[1] * d.dy + vc[2] * d.dz;
                    const float pgp = Pc[0] * d.dx + Pc[1] * d.dy + Pc[2] * d.dz;
                    const float dtp = -vgp - co.alpha * ph * Pc[b] - co.beta * P2 * Pc[b] - co.lambda0 * pgp - an[b];
                    nxt[3 + b] = Pc[b] + co.Dt * dtp;
                    __builtin_amdgcn_sched_barrier(0);
                }
                const float Dp[3] = {dph.dx, dph.dy, dph.dz};
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const Der4 d = der_rows<SV>(rv, b * NZ, co);
                    const float vdv = vc[0] * d.dx + vc[1] * d.dy + vc[2] * d.dz;
                    nxt[b] = vc[b] + co.Dt * (-vdv + co.mu * d.l + mu0 * Dp[b]);
                    __builtin_amdgcn_sched_barrier(0);
                }
                nxt[6] = ph + co.Dt * (-(vc[0] * dph.dx + vc[1] * dph.dy + vc[2] * dph.dz));
                // materialise the carried terms here: sunk past the barrier,
                // their inputs would stay live through stage 3
#pragma unroll
                for (int j = 0; j < 7; ++j) asm volatile("" : "+v"(nxt[j]));
            }
            // Intermediates of the apron ring, spread over every wave as
            // (position, part) pairs when one pass of the tile's threads covers
            // them: four parts (mu0, then the intermediates of P_x, P_y, P_z),
            // or three (those of P_b, with mu0 in part 0).  Otherwise each of
            // the first NRING threads computes all 13 of one position.  All use
            // the own cell's expressions.
            // two waves with a ring of at most 64 positions: one part per wave
            constexpr bool kWaveParts = !(3 * NRING <= NT) && NT == 128 && NRING <= 64;
            constexpr int NPART = 4 * NRING <= NT ? 4 : 3 * NRING <= NT ? 3 : kWaveParts ? 2 : 1;
            constexpr int PSTR = kWaveParts ? 64 : NRING;  // lane stride between parts
            if (ti % PSTR < NRING && ti < NPART * PSTR) {
                const int f = ti / PSTR, r = ti % PSTR;
                int lx1, ly1;
                if (r < W1) { lx1 = r; ly1 = 0; }
                else if (r < 2 * W1) { lx1 = r - W1; ly1 = H1 - 1; }
                else if (r < 2 * W1 + TY) { lx1 = 0; ly1 = 1 + (r - 2 * W1); }
                else { lx1 = W1 - 1; ly1 = 1 + (r - 2 * W1 - TY); }
                const int q2 = (ly1 + 1) * W2 + (lx1 + 1);
                const int q1 = ly1 * W1 + lx1;
                const Rows rq = rows_at<W2, SP>(zp, q2);
                const float ph = rq.c[SP + ZF];
                float* m = sM + q1 * SS;
                // intermediates of P_b at this position
                auto pint = [&](int b, int fo) {
                    const Der4 d = der_rows<SP>(rq, fo, co);
                    m[9 + b] = ph * rq.c[SP + fo + ZF] + co.La2 * d.l;
                    m[0 * 3 + b] = ph * d.dx;
                    m[1 * 3 + b] = ph * d.dy;
                    m[2 * 3 + b] = ph * d.dz;
                };
                auto mu0 = [&]() {
                    const float ph2 = ph * ph;
                    m[12] = co.aPhi2 * (ph - 3.f * ph2 + 2.f * ph2 * ph) - co.kPhi * lap_rows<SP>(rq, 0, co);
                };
                if constexpr (NPART == 4) {
                    if (f == 0) mu0();
                    else pint(f - 1, f * NZ);
                } else if constexpr (NPART == 3) {
                    if (f == 0) mu0();
                    pint(f, (1 + f) * NZ);
                } else if constexpr (NPART == 2) {
                    if (f == 0) {  // wave 0: mu0 and P_x
                        mu0();
                        pint(0, 1 * NZ);
                    } else {  // wave 1: P_y and P_z
                        pint(1, 2 * NZ);
                        __builtin_amdgcn_sched_barrier(0);
                        pint(2, 3 * NZ);
                    }
                } else {
                    mu0();
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        pint(b, (1 + b) * NZ);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }
            __syncthreads();

            // ---- 3. contribution of plane zm to outputs zm-1, zm, zm+1 ------
            {
                const Rows rs = rows_at<W1, SS>(sM, o1c);
#define SM(f, dx, dy) ((dy) < 0 ? rs.m : (dy) > 0 ? rs.p : rs.c)[((dx) + 1) * SS + (f)]
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
                    aS[S0][b] += co.sw1 * A + Bz;
                    aS[S1][b] += co.sw0 * A;
                    aS[S2][b] += co.sw1 * A - Bz;
                    const int fm = 9 + b;
                    const float cen = SM(fm, 0, 0);
                    const float E = co.lc * ((SM(fm, -1, -1) + SM(fm, 1, -1)) + (SM(fm, -1, 1) + SM(fm, 1, 1))) + co.lf * cen;
                    const float F = co.lf * ((SM(fm, -1, 0) + SM(fm, 1, 0)) + (SM(fm, 0, -1) + SM(fm, 0, 1))) + co.l0 * cen;
                    aL[S0][b] += E;
                    aL[S1][b] += F;
                    aL[S2][b] += E;
                    __builtin_amdgcn_sched_barrier(0);
                }
                {
                    const float cen = SM(12, 0, 0);
                    const float E = co.lc * ((SM(12, -1, -1) + SM(12, 1, -1)) + (SM(12, -1, 1) + SM(12, 1, 1))) + co.lf * cen;
                    const float F = co.lf * ((SM(12, -1, 0) + SM(12, 1, 0)) + (SM(12, 0, -1) + SM(12, 0, 1))) + co.l0 * cen;
                    aL0[S0] += E;
                    aL0[S1] += F;
                    aL0[S2] += E;
                }
#undef SM
            }

            // ---- 4. finish output plane zo = zi-2 ---------------------------
            const int zo = zi - 2;
            if (zo >= zb && own) {
                const size_t pz = (size_t)zo * NN;
                stb(ophi + pz, go, cur[6] + co.Dt * (co.gamma * aL0[S0]));
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    stb(ovs[b] + pz, go, cur[b] + co.Dt * (co.cA * aS[S0][b]));
                    stb(ops[b] + pz, go, cur[3 + b] + co.Dt * (co.eta * aL[S0][b]));
                }
            }
            if constexpr (kUnroll3) {
                // slot S0 becomes plane zm+2 in the next phase
                aL0[S0] = 0.f;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    aS[S0][b] = 0.f;
                    aL[S0][b] = 0.f;
                }
            } else {
                // rotate the output planes: zm becomes the next zm-1
                aL0[0] = aL0[1];
                aL0[1] = aL0[2];
                aL0[2] = 0.f;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    aS[0][b] = aS[1][b];
                    aS[1][b] = aS[2][b];
                    aS[2][b] = 0.f;
                    aL[0][b] = aL[1][b];
                    aL[1][b] = aL[2][b];
                    aL[2][b] = 0.f;
                }
            }
#pragma unroll
            for (int j = 0; j < 7; ++j) cur[j] = nxt[j];
        };

        // z-march over zi = zb .. ze+1
        if constexpr (kUnroll3) {
            for (int zi = zb;;) {
                body(Ph<0>{}, zi);
                if (++zi > ze + 1) break;
                body(Ph<1>{}, zi);
                if (++zi > ze + 1) break;
                body(Ph<2>{}, zi);
                if (++zi > ze + 1) break;
            }
        } else {
            for (int zi = zb; zi <= ze + 1; ++zi) body(Ph<0>{}, zi);
        }
    }
}

// ---------------------------------------------------------------------------
template <int TX, int TY, int WPS>
static hipError_t launch_rhs_t(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                               int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    return launch_kernel(k_rhs<TX, TY, WPS>, dim3(grid), dim3(TX * TY), st, drops, work, nwork_p, parity, kstep);
}

// Tile variants: TX x TY cells per work-group and the waves per SIMD the
// register allocation targets (the second __launch_bounds__ argument: 2 allows
// 256 VGPRs, 3 allows 168, 1 allows 512).
struct RhsVar {
    int tx, ty, wps;
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

// Work-groups per CU that the LDS of a tile allows (160 KiB per CU): the
// phi/P planes over a 2-cell apron (13 floats a position), v and the
// intermediates over a 1-cell apron (9 and 13).
int rhs_groups_per_cu(int variant)
{
    int tx, ty;
    rhs_tile(variant, &tx, &ty);
    if (tx <= 0) return 1;
    const int bytes = 4 * ((tx + 4) * (ty + 4) * 13 + (tx + 2) * (ty + 2) * (9 + 13));
    return std::max(1, (160 * 1024) / bytes);
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
