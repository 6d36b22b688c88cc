// ttsh_internal.h -- library-internal helpers shared by the host translation units.
#pragma once
#include <cstdint>
#include <string>

namespace ttsh {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int fft_plan(int N, int* rad, int maxst);
bool supported_N(int N, std::string* why);
int write_conf_host(const char* path, int N, int expo_data, const float* phi, const float* P, const float* v);
// Binary sidecar: "TTSH3DB1", int32 N, int32 m, double T, int64 steps, then
// fields[0..6] = vx vy vz px py pz phi as FP32 N^3 each.
int write_conf_bin_host(const char* path, int N, int m, double T, int64_t steps, const float* const fields[7]);
// Level-contour radius R(theta) at 120 angles of an N x N [row][col] field.
int contour_radius(int N, const float* field, double level, double* R, double* center);
// Grid sizes with a compile-time FFT plan (k_fft.hip): the even N in
// [32, 256] whose prime factors are 2, 3 and 5 and that split as R1 x R2 with
// register DFT lengths R1, R2 in {2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20}:
// every such N but those with a factor 27 or 125 (54, 108, 162, 216, 250).
constexpr int kNumGridSizes = 22;
constexpr int kGridSizes[kNumGridSizes] = {32, 36, 40, 48, 50, 60, 64, 72, 80, 90, 96,
                                           100, 120, 128, 144, 150, 160, 180, 192, 200, 240, 256};
// Number of K_rhs tile variants (RhsVariant in ttsh_kernels.h).
constexpr int kRhsVariants = 7;
// Tuned K_rhs variant, K_rhs z-depth and projection z-depth for size N.
bool tune_lookup(int N, int* rhs_variant, int* rhs_tz, int* proj_tz);
// Tuned layout of a batch that fills the device: K_rhs variant, minimum
// K_rhs z-depth, projection z-depth (false: no entry).
bool tune_batch_lookup(int* rhs_variant, int* rhs_tz, int* proj_tz);
// Incremented whenever the tuning table is (re)loaded.
int tune_generation();

}  // namespace ttsh
