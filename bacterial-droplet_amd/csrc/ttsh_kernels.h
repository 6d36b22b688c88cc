// ttsh_kernels.h -- host-side launchers of the step kernels (library internal).
//
// Every launcher takes a device work table, a device work count (nwork_p) and
// a grid size; the kernels loop over work items with a grid stride and skip
// items whose droplet has no step left at index kstep.
#pragma once
#include <hip/hip_ext.h>

#include "ttsh_dev.h"
#include "ttsh_internal.h"

namespace ttsh {

// Events that the next step-kernel launch on this thread records from its
// own dispatch (start) and completion (stop), as a profiler's kernel trace
// times it, with no marker of its own between kernels.  A launch clears them.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchEvents g_launch_ev;

template <typename K, typename... Args>
inline hipError_t launch_kernel(K kernel, dim3 grid, dim3 block, hipStream_t st, Args... args)
{
    if (g_launch_ev.start || g_launch_ev.stop) {
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, g_launch_ev.start, g_launch_ev.stop, 0, args...);
        g_launch_ev = LaunchEvents{};
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    }
    return hipGetLastError();
}

// K_rhs tile variants (TX x TY cells per work-group, one cell per thread, and
// the waves per SIMD the register allocation targets); see kRhsVars in
// k_rhs.hip.
enum RhsVariant { RHS_32x8 = 0, RHS_16x16 = 1, RHS_64x4 = 2, RHS_16x8 = 3, RHS_32x8_O3 = 4, RHS_32x16 = 5,
                  RHS_64x8 = 6, RHS_NVARIANTS = 7 };
static_assert(RHS_NVARIANTS == kRhsVariants, "tile variant count");

// Tile of the divergence and corrector stencils.
constexpr int kProjTX = 32;
constexpr int kProjTY = 8;

// FFT pass kinds (k_fft.hip).
enum { FFT_MODE_FWD = 0, FFT_MODE_INV = 1, FFT_MODE_POISSON = 2 };

hipError_t launch_rhs(int variant, const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                      int grid, hipStream_t st);
void rhs_tile(int variant, int* tx, int* ty);
int rhs_groups_per_cu(int variant);

// s = (D.v* + i D.P*)/dt0 with the stale-halo rule; FP64 partial sums of s.
hipError_t launch_div(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int grid,
                      hipStream_t st);
// Per droplet (work item (d, -, -)): the partial sums of the divergence items
// in item order -> stat (k = 0 coefficient) and sigma of this step.
hipError_t launch_proj_stats(const Drop* drops, const Work* work, const int* nwork_p, int kstep, int grid,
                             hipStream_t st);
// The x = 0 and x = N-1 columns of vs[3], ps[3] into edge slot `slot`
// (the layout of Drop::edge).
hipError_t launch_fill_edges(const float* const vs[3], const float* const ps[3], float* edge, int slot, int N,
                             hipStream_t st);
// n 32-bit words at p set to value.
hipError_t launch_fill_u32(void* p, uint32_t value, size_t n, hipStream_t st);
// v = v* - Dt D p_v, P = P* - Dt D p_p from p_v + i p_p.
hipError_t launch_corr(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int grid,
                       hipStream_t st);
// One FFT pass along axis (0 x, 1 y, 2 z) for droplets of size N.  mode
// FFT_MODE_POISSON (z only): forward, times -1/(k^2 N^3), inverse.
hipError_t launch_fft(int N, int axis, int mode, const Drop* drops, const Work* work, const int* nwork_p, int kstep,
                      int grid, hipStream_t st);
// Lines per FFT work item for size N (0 if N has no plan).
int fft_lines_per_group(int N);
// Sizes up to this have the fused x+y pass (one work-group per z-plane,
// the plane in LDS).
constexpr int kMaxFusedXy = 128;
bool fft_xy_fused(int N);
// Fused x and y passes over every z-plane (work item (d, -, z)): mode
// FFT_MODE_FWD transforms x then y, FFT_MODE_INV y then x.
hipError_t launch_fft_xy(int N, int mode, const Drop* drops, const Work* work, const int* nwork_p, int kstep, int grid,
                         hipStream_t st);
// Diagnostics (k_diag.hip): FP64 sums of phi, phi*y, phi*x, phi*z -> out[4]
// (part holds 4*nblk doubles); plane z of phi, v, P and the 4th-order
// vorticities of v and P -> out[9][N][N] (out has 9*N*N+1 floats; z < 0 takes
// the centroid plane from sums and stores it in out[9*N*N]).
hipError_t launch_centroid(const float* phi, int N, double* part, int nblk, double* out, hipStream_t st);
hipError_t launch_slice(const float* phi, const float* const v[3], const float* const P[3], int N, int z,
                        const double* sums, float h, float* out, hipStream_t st);
// FP64 reductions: per droplet sum(phi), sum|P|, sum|v|^2 -> out[3*d..].
hipError_t launch_reduce(const Drop* drops, int ndrops, int parity, double* part, int nblk, double* out,
                         hipStream_t st);

}  // namespace ttsh
