// ttsh3d -- command-line drop-in for the reference's TTSH3D.out.
//
// Process contract of the reference (TTSH3D.cu:263-287, run_cuda.pl:89-146):
// read data_<gpu_node>/input.dat, initialise, run the time loop writing
// data_<gpu_node>/conf_<m>.dat, print "Simulation finished." and the time.
// gpu_node is a compile-time literal in the reference (TTSH3D.cu:42, patched
// by run_cuda.pl:98-110); here it is the first argument or TTSH_GPU_NODE
// (default 1).  As with CUDA_VISIBLE_DEVICES in run_cuda.pl:136, the process
// uses the first visible device unless --device is given.
//
// Differences that are deliberate:
//   - a sentinel failure exits with status 1 (the reference prints the error
//     and exits 0, TTSH3D.cu:270,285);
//   - InitCond == 1 noise comes from a seeded SplitMix64 stream (--seed or
//     TTSH_SEED; default time(0) as the reference's seed, TTSH3D.cu:40);
//   - InitCond == 0 (a legacy 2D file read of vx, vy, TTSH3D.cu:768-797) is
//     rejected.
// Output options (none changes the default, the reference's text snapshots):
//   --binary    also write conf_<m>.bin (raw FP32 fields)
//   --diag      also write diag.dat (centroid) and a_R.txt (contour radii)
//   --no-text   skip the text snapshots
// Snapshots are written by a host thread while the time loop continues.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/ttsh3d.h"

static int usage()
{
    std::fprintf(stderr,
                 "usage: ttsh3d [gpu_node] [--device D] [--seed S] [--exact-halo 0|1] [--binary] [--diag] "
                 "[--no-text]\n");
    return 2;
}

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    int gpu_node = 1, device = 0, faithful = 1, flags = TTSH_RUN_TEXT;
    const char* env_node = std::getenv("TTSH_GPU_NODE");
    if (env_node) gpu_node = std::atoi(env_node);
    unsigned long long seed = (unsigned long long)std::time(nullptr);
    const char* env_seed = std::getenv("TTSH_SEED");
    if (env_seed) seed = std::strtoull(env_seed, nullptr, 10);
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--exact-halo" && i + 1 < argc) faithful = std::atoi(argv[++i]);
        else if (a == "--binary") flags |= TTSH_RUN_BINARY;
        else if (a == "--diag") flags |= TTSH_RUN_DIAG;
        else if (a == "--no-text") flags &= ~TTSH_RUN_TEXT;
        else if (a == "-h" || a == "--help") return usage();
        else if (!a.empty() && a[0] != '-') gpu_node = std::atoi(a.c_str());
        else return usage();
    }
    const std::string dir = "data_" + std::to_string(gpu_node);
    ttsh_params p;
    ttsh_params_default(&p);
    int rc = ttsh_parse_input_dat((dir + "/input.dat").c_str(), &p);
    std::printf("hhhhhhhhhhh:%d\n", p.ReadSuccess);
    if (rc != TTSH_OK) {
        std::printf("Error while reading the input file!\n");
        std::fprintf(stderr, "ttsh3d: %s\n", ttsh_last_error());
        return 1;
    }
    p.faithful_halo = faithful;
    rc = ttsh_validate_params(&p);
    if (rc != TTSH_OK) {
        std::fprintf(stderr, "ttsh3d: %s\n", ttsh_last_error());
        return 1;
    }
    if (p.InitCond != 1) {
        std::fprintf(stderr, "ttsh3d: only InitCond 1 is supported (got %d)\n", p.InitCond);
        return 1;
    }
    const int N = p.Nx;
    const size_t n = (size_t)N * N * N;
    std::vector<double> phi(n), P(3 * n), v(3 * n), vs(3 * n), ps(3 * n);
    ttsh_init_reference(&p, seed, phi.data(), P.data(), v.data(), vs.data(), ps.data());
    ttsh_sim* sim = nullptr;
    rc = ttsh_create(&p, device, &sim);
    if (rc == TTSH_OK) rc = ttsh_set_state(sim, TTSH_F64, phi.data(), P.data(), v.data(), vs.data(), ps.data());
    int64_t steps = 0;
    if (rc == TTSH_OK) rc = ttsh_run_ex(sim, dir.c_str(), flags, &steps);
    if (rc != TTSH_OK) {
        std::fprintf(stderr, "ttsh3d: %s\n", ttsh_last_error());
        ttsh_destroy(sim);
        return 1;
    }
    ttsh_destroy(sim);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation finished. Wall time = %g sec (%lld steps)\n", wall, (long long)steps);
    return 0;
}
