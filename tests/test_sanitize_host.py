"""AddressSanitizer and UndefinedBehaviorSanitizer over the CPU code: the FP64
oracle's step, projection, FFT and stencils, and the host side of the C ABI
(input.dat parsing and writing, snapshot text, initial conditions, contour
radius).  Both are built here with gcc/g++ into small drivers and run at 16^3 and 32^3;
any sanitizer report fails the test.  No GPU code is built (the device side
is checked by the GPU tests' bounds on the host and by reading the kernels)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]

ORACLE_DRIVER = r"""
#include "ttsh_oracle.c"
#include <stdio.h>
int main(void)
{
    const int N = 16;
    const size_t n = (size_t)N * N * N;
    orc_params p;
    memset(&p, 0, sizeof p);
    p.N = N; p.h = 1.1; p.dt0 = 0.01; p.mu = 1; p.cA = -1.5; p.alpha = -0.02; p.beta = 4.366;
    p.eta = -0.187; p.La = 1.2; p.kappa = -0.364; p.lambda0 = 9; p.aPhi = 1; p.kPhi = 1; p.gamma = 1;
    p.A = 0.3; p.faithful = 1;
    double *phi = malloc(n * 8), *P = malloc(3 * n * 8), *v = malloc(3 * n * 8), *vs = malloc(3 * n * 8);
    double *ps = malloc(3 * n * 8), *aux = malloc(9 * n * 8), red[3];
    uint64_t s = 12345;
    for (size_t i = 0; i < 3 * n; ++i) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        const double r = (double)(s >> 11) / 9007199254740992.0;
        if (i < n) phi[i] = r;
        P[i] = r - 0.5; v[i] = 0.01 * (r - 0.5); vs[i] = v[i] * 1.01; ps[i] = P[i] * 0.99;
    }
    orc_state *st = orc_create(&p);
    if (!st) return 2;
    orc_set_state(st, phi, P, v, vs, ps);
    orc_step(st, 3);
    for (int w = 0; w < 14; ++w) orc_get_aux(st, w, aux);
    orc_reductions(st, red);
    orc_get_state(st, phi, P, v);
    orc_destroy(st);
    p.faithful = 0;
    double *pp = malloc(n * 8), *X = malloc(3 * n * 8);
    orc_projection(&p, v, pp, X);
    orc_fft3d(N, phi, pp, -1);
    orc_fft1d(N, phi, pp, 1);
    for (int w = 0; w < 4; ++w) orc_stencil(w, N, 1.0, phi, pp);
    printf("ok %g %g %g\n", red[0], red[1], red[2]);
    free(phi); free(P); free(v); free(vs); free(ps); free(aux); free(pp); free(X);
    return 0;
}
"""

HOST_DRIVER = r"""
#include "ttsh3d.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
int main(int argc, char** argv)
{
    const std::string dir = argv[1];
    ttsh_params p;
    if (ttsh_params_default(&p)) return 2;
    p.Nx = p.Ny = p.Nz = 32;  // the smallest size with an FFT plan
    const std::string in = dir + "/input.dat";
    if (ttsh_write_input_dat(in.c_str(), &p)) return 3;
    ttsh_params q;
    ttsh_params_default(&q);  // parsing keeps the fields input.dat does not hold
    if (ttsh_parse_input_dat(in.c_str(), &q)) return 4;
    if (ttsh_validate_params(&q)) return 5;
    if (ttsh_parse_input_dat((dir + "/missing.dat").c_str(), &q) == 0) return 6;
    const int N = 32;
    const size_t n = (size_t)N * N * N;
    std::vector<double> phi(n), P(3 * n), v(3 * n), vs(3 * n), ps(3 * n);
    if (ttsh_init_reference(&p, 7, phi.data(), P.data(), v.data(), vs.data(), ps.data())) return 7;
    for (int mode = 0; mode < 2; ++mode)
        if (ttsh_init_droplet(N, 10.0, 2.0, mode, 9, phi.data(), P.data(), v.data())) return 8;
    std::vector<float> f(n), F(3 * n), V(3 * n);
    for (size_t i = 0; i < n; ++i) f[i] = (float)phi[i];
    for (size_t i = 0; i < 3 * n; ++i) { F[i] = (float)P[i]; V[i] = (float)v[i]; }
    F[0] = std::numeric_limits<float>::quiet_NaN(); F[1] = -INFINITY; F[2] = 1e30f; V[3] = -0.0f; V[4] = 1e-45f;
    for (int e : {1, -1})
        if (ttsh_write_conf_arrays((dir + "/conf_" + std::to_string(e + 1) + ".dat").c_str(), N, e, f.data(),
                                   F.data(), V.data()))
            return 9;
    double R[120], c[2];
    std::vector<float> plane(f.begin() + 8 * N * N, f.begin() + 9 * N * N);
    if (ttsh_contour_radius(N, plane.data(), 0.5, R, c)) return 10;
    std::vector<float> zero(N * N, 0.f);
    if (ttsh_contour_radius(N, zero.data(), 0.5, R, c) == 0) return 11;
    std::printf("ok %g %g\n", R[0], c[0]);
    return 0;
}
"""


def _tool(name):
    path = shutil.which(name)
    if path is None:
        pytest.skip(f"{name} not available")
    return path


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    return r.returncode, r.stdout + r.stderr


def test_oracle_under_sanitizers(tmp_path):
    gcc = _tool("gcc")
    src = tmp_path / "oracle_driver.c"
    src.write_text(ORACLE_DRIVER)
    exe = tmp_path / "oracle_driver"
    rc, out = _run([gcc, "-std=c11", *SAN, "-I", os.path.join(ROOT, "oracle"), str(src), "-lm", "-o", str(exe)])
    assert rc == 0, out
    rc, out = _run([str(exe)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert rc == 0 and "runtime error" not in out and out.startswith("ok"), out


def test_host_abi_under_sanitizers(tmp_path):
    gxx = _tool("g++")
    src = tmp_path / "host_driver.cpp"
    src.write_text(HOST_DRIVER)
    exe = tmp_path / "host_driver"
    rc, out = _run([gxx, "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), str(src),
                    os.path.join(ROOT, "bacterial-droplet_amd", "csrc", "ttsh_host.cpp"), "-ldl", "-lpthread",
                    "-o", str(exe)])
    assert rc == 0, out
    rc, out = _run([str(exe), str(tmp_path)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert rc == 0 and "runtime error" not in out and out.startswith("ok"), out
