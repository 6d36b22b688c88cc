// ttsh_kernels.h -- host-side launchers of the step kernels (library internal).
#pragma once
#include "ttsh_dev.h"

namespace ttsh {

// K_rhs tile variants (TX x TY cells per work-group, one cell per thread).
enum RhsVariant { RHS_32x8 = 0, RHS_16x16 = 1, RHS_64x4 = 2, RHS_16x8 = 3, RHS_NVARIANTS = 4 };

// Tile constants of the projection kernels.
constexpr int kDivRows = 4;    // rows (y) per divergence work item
constexpr int kColTile = 16;   // x-columns per Y/Z FFT work item
constexpr int kCorrRows = 4;   // rows (y) per corrector work item
constexpr int kFftThreads = 256;
constexpr int kFftMaxB = 4;    // butterflies per thread per stage

// Every launcher takes the device work count (nwork_p) and a grid size; the
// kernels loop over work items with a grid stride.
hipError_t launch_rhs(int variant, const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep,
                      int grid, hipStream_t st);
void rhs_tile(int variant, int* tx, int* ty);

// Divergence of v* (real) and P* (imaginary) / dt0, forward FFT along x.
hipError_t launch_divx(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int maxN,
                       int grid, hipStream_t st);
// FFT along y (forward if inverse==0).
hipError_t launch_ffty(const Drop* drops, const Work* work, const int* nwork_p, int inverse, int kstep, int maxN,
                       int grid, hipStream_t st);
// Forward FFT along z, multiply by -1/(k^2 N^3) (k2[0]=1), inverse FFT along z.
hipError_t launch_fftz(const Drop* drops, const Work* work, const int* nwork_p, int kstep, int maxN, int grid,
                       hipStream_t st);
// Inverse FFT along x and the corrector v = v* - Dt D p_v, P = P* - Dt D p_p.
hipError_t launch_corrx(const Drop* drops, const Work* work, const int* nwork_p, int parity, int kstep, int maxN,
                        int grid, hipStream_t st);
// FP64 reductions: per droplet sum(phi), sum|P|, sum|v|^2 -> out[3*d..].
hipError_t launch_reduce(const Drop* drops, int ndrops, int parity, double* part, int nblk, double* out,
                         hipStream_t st);

size_t divx_lds_bytes(int maxN);
size_t ffty_lds_bytes(int maxN);
size_t fftz_lds_bytes(int maxN);
size_t corrx_lds_bytes(int maxN);

}  // namespace ttsh
