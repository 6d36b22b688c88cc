// ttsh_internal.h -- library-internal helpers shared by the host translation units.
#pragma once
#include <string>

namespace ttsh {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int fft_plan(int N, int* rad, int maxst);
bool supported_N(int N, std::string* why);
int write_conf_host(const char* path, int N, int expo_data, const float* phi, const float* P, const float* v);
// Tuned K_rhs variant, K_rhs z-depth and projection z-depth for size N.
bool tune_lookup(int N, int* rhs_variant, int* rhs_tz, int* proj_tz);

}  // namespace ttsh
