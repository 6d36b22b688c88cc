// ttsh_dev.h -- device-side data layout shared by the TTSH3D step kernels.
//
// One droplet lives in HBM as unpadded FP32 structure-of-arrays, x fastest and
// z slowest: idx = (z*N + y)*N + x, the interior ordering of the reference's
// padded fields (TTSH3D.cu:477 without the Nb halo).  Periodic boundaries are
// index wraps inside the tile loaders; there are no halo arrays and no halo
// kernels (the reference's BounPeriU, TTSH3D.cu:1422-1607, has no counterpart).
//
// Buffers per droplet (N^3 each unless stated):
//   phi[2]           phase field, ping-pong (K_rhs reads one, writes the other)
//   P[3], v[3]       polarization and flow (updated in place by K_corr)
//   vs[2][3],ps[2][3] predictors v*, P*, ping-pong by step parity; the other
//                    parity holds the previous step's v*, P*, which the
//                    divergence reads across the periodic boundary exactly as
//                    the reference's stale halo does (TTSH3D.cu:366-371,405-407,665-667)
//   c                complex FP32 N^3: (div v* + i div P*)/dt0, transformed in place;
//                    the first forward pass scales the imaginary part by 1/sigma
//
// All kernels take a device table of droplet descriptors and a work table of
// (droplet, tile...) entries, so one launch serves one droplet or a batch of
// droplets of different sizes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ttsh {

constexpr int kMaxStages = 8;
constexpr int kMaxN = 256;

// FP32 coefficients of one droplet, computed in FP64 on the host.
struct Coef {
    float d1;         // 1/(2h): first-derivative prefactor (TTSH3D.cu:3287)
    float sw0, sw1;   // transverse smoothing S=[sw1,sw0,sw1] = [1,4,1]/6
    float lc, lf, l0; // Laplacian: 8 corners 1/(12h^2), 6 faces 2/(3h^2), centre -14/(3h^2)
    float Dt;         // dt0 (Dt never changes in the reference, TTSH3D.cu:759)
    float inv_dt0;    // 1/dt0, the getStream type-0 divisor (TTSH3D.cu:665)
    float mu, cA, alpha, beta, eta, La2, lambda0, aPhi2, kPhi, gamma, A;
    float e0, e1;     // d1*sw0, d1*sw1: weights of a transverse-smoothed difference
    float f00, f01, f11;  // d1*sw0^2, d1*sw0*sw1, d1*sw1^2: weights of S (x) S (x) difference
    // Laplacian from f and S_z f = (4 f + f(z-1) + f(z+1))/6 of the plane:
    // f(z-1)+f(z+1) = 6 S_z f - 4 f, so L = lcs*sum_corners S_z f + lcf*sum_corners f
    // + lf*sum_faces f + l0s*S_z f + l0f*f
    float lcs, lcf, l0s, l0f;
    float k2s;        // (2 pi/(N h))^2: k2 = k2s*(mx^2+my^2+mz^2) (TTSH3D.cu:1094-1114)
    float inv_n3;     // 1/N^3, the type-2 divisor (TTSH3D.cu:687)
    int   anchoring;  // A != 0: evaluate the anchoring term
};

constexpr int kPartSums = 6;

struct Drop {
    int N;
    int nst;                 // FFT stages
    int rad[kMaxStages];     // radices (4, 2 or 3), product N
    const float2* tw;        // exp(-2 pi i q/N), q in [0,N)
    float* phi[2];
    float* P[3];
    float* v[3];
    float* vs[2][3];
    float* ps[2][3];
    float2* c;
    // FP64 sums per reduction block, kPartSums each: s.x, s.y, s.x^2, s.y^2
    // over the block's cells and |v*|^2, |P*|^2 over the same cells.  A block
    // is one divergence tile over one of nzp z-ranges (the droplet's tuned
    // z-chunks); npart = tiles * nzp, ordered by z-range, then tile.  A
    // divergence work item covers one or more whole blocks of one tile.
    double* part;
    int npart;
    int nzp;
    // Per step, from the sums of all items (k_proj_stats): stat[0..1] the
    // k = 0 coefficient of s/sigma (FP64); sigma the power-of-two scale of
    // the imaginary part.  The packed transform carries div v* and div P*/
    // sigma, and K_corr multiplies p_p back by sigma: with sigma chosen so
    // that the FP32 rounding of either half disturbs the other by the same
    // relative amount in the corrected v and P, a droplet whose div P* is
    // orders of magnitude above its div v* keeps the accuracy of v.
    double* stat;
    float* sigma;
    // x = 0 and x = N-1 columns of v*, P* per parity slot, laid out with a
    // field's z-stride: plane z holds at ((slot*6 + f)*2 + side)*N + y the
    // value of field f (vx..vz, Px..Pz) at x = side ? N-1 : 0, row y.  The
    // stale-halo divergence reads the previous step's columns here instead
    // of fetching a whole cache line of the previous predictor per row end.
    // eslot[p] is the slot of the predictors in vs[p], ps[p].
    float* edge;
    int eslot[2];
    int faithful;            // 1: stale-halo divergence (the reference's semantics)
    int nsteps;              // steps this droplet takes in the current launch sequence
    Coef co;
};

// Work item: droplet index and up to three tile coordinates.  Kernels skip an
// item whose droplet has nsteps <= kstep (the step index within a launch
// sequence), so droplets with different step budgets share one sequence.
struct Work {
    int d, a, b, c;
};

// Work-groups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Visit v of the grid-stride loop maps to work item
//   (v mod 8) * chunk + v / 8,   chunk = ceil(nwork / 8),
// so each XCD processes one contiguous, spatially ordered block of the work
// table and tiles that share halo cache lines share an L2.  The loop runs v
// over [0, 8*chunk) and skips items >= nwork.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_span(int nwork) { return kXcds * ((nwork + kXcds - 1) / kXcds); }
__device__ __forceinline__ int xcd_item(int v, int nwork)
{
    const int chunk = (nwork + kXcds - 1) / kXcds;
    return (v % kXcds) * chunk + v / kXcds;
}

// Global-memory views of the descriptor's pointers.  Pointers read from a
// table in memory are generic to the compiler, which then emits flat loads and
// stores; those count against the LDS wait counter too, so the first LDS wait
// after a prefetch would also wait for the prefetch's HBM loads.  A cast to
// the global address space gives global loads, tracked by the vector-memory
// counter alone.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glb(T* p)
{
    return (gptr<T>)p;
}

// float2 elements of global memory go through a plain two-float vector: the
// HIP float2 class has no assignment operators for an address-space-qualified
// object.
typedef float fvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld2(gptr<const float2> p, size_t i)
{
    const fvec2 t = *(gptr<const fvec2>)(p + i);
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ void st2(gptr<float2> p, size_t i, float2 v)
{
    fvec2 t;
    t.x = v.x;
    t.y = v.y;
    *(gptr<fvec2>)(p + i) = t;
}

// Byte-offset access to a plane: a uniform base plus a 32-bit per-lane byte
// offset, the form of a global load or store with a scalar base address.
__device__ __forceinline__ float ldb(gptr<const float> base, unsigned boff)
{
    return *(gptr<const float>)((gptr<const char>)base + boff);
}
__device__ __forceinline__ void stb(gptr<float> base, unsigned boff, float v)
{
    *(gptr<float>)((gptr<char>)base + boff) = v;
}

// Compile-time phase of an unrolled z-march (register slot rotation).
template <int K>
using Ph = std::integral_constant<int, K>;

// Waits for every outstanding vector-memory load (s_waitcnt vmcnt(0), the
// export and LDS counters left at their maxima).  Issued at the end of a
// prologue whose loads fill a loop's rolling registers, so the compiler's
// wait analysis enters the loop with no load pending and does not fence the
// loop's own prefetch of the next plane behind later LDS traffic.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ int wrapi(int i, int N)
{
    i += (i < 0) ? N : 0;
    i -= (i >= N) ? N : 0;
    return i;
}

__device__ __forceinline__ size_t gidx(int N, int x, int y, int z)
{
    return ((size_t)z * (size_t)N + (size_t)y) * (size_t)N + (size_t)x;
}

}  // namespace ttsh
