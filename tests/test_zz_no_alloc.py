"""ttsh_step_noalloc, the entry point that allocates nothing, against
ttsh_step on the same states: bitwise equal state and last predictors, on
two calls in a row (the second with a smaller droplet, so it runs on what
the first left in the shared scratch array), and a continued run across
calls.  The caller's device memory is allocated here with hipMalloc of the
HIP runtime the library itself loads (through ctypes); the library's C ABI
is called through ctypes too."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip(ttsh):
    ttsh.lib()  # loads libamdhip64.so.7 as a dependency; this is the same runtime
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


class DevArray:
    """A caller-owned device copy of a float32 host array."""

    def __init__(self, H, a):
        self.H = H
        self.host = np.ascontiguousarray(a, dtype=np.float32)
        self.ptr = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.ptr), self.host.nbytes) == 0
        assert H.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.H.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def run_noalloc(ttsh, p, state, prev, n):
    H = hip(ttsh)
    arrs = [DevArray(H, x) for x in (*state, *prev)]
    try:
        assert H.hipDeviceSynchronize() == 0
        ttsh.step_noalloc(p, *(a.ptr.value for a in arrs), n)
        out = [a.get() for a in arrs]
    finally:
        for a in arrs:
            a.free()
    return tuple(out[:3]), (out[3], out[4])


def run_step(ttsh, p, state, prev, n):
    with ttsh.Simulation(p) as s:
        s.set_state(*state, *prev, dtype=ttsh.F32)
        s.step(n)
        out = s.get_state()
        last = (s.aux("vstar"), s.aux("pstar"))
    return out, last


def same(a, b, what):
    for name, x, y in zip(("phi", "P", "v"), a, b):
        x, y = np.asarray(x).reshape(-1), np.asarray(y).reshape(-1)
        log(what, name, "max |diff|", float(np.max(np.abs(x - y))))
        assert np.array_equal(x, y), f"{what}: {name}"


def test_noalloc_equals_step_on_two_calls(ttsh):
    cases = [(64, 1, -2.0, 6, 777), (32, 0, -1.0, 5, 778)]  # second call: a smaller droplet
    for N, pmode, cA, n, seed in cases:
        log(f"[noalloc N={N} pmode={pmode} n={n}]")
        p = ttsh.default_params(N=N, dt0=0.01, cA=cA)
        phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, pmode, seed)
        state = tuple(np.asarray(x, np.float32) for x in (phi, P, v))
        prev = (state[2].copy(), state[1].copy())  # a fresh state: v*, P* = v, P
        got, got_last = run_noalloc(ttsh, p, state, prev, n)
        ref, ref_last = run_step(ttsh, p, state, prev, n)
        same(got, ref, f"N={N} state")
        assert np.array_equal(got_last[0].reshape(-1), ref_last[0].reshape(-1)), "v* out"
        assert np.array_equal(got_last[1].reshape(-1), ref_last[1].reshape(-1)), "P* out"
        assert all(np.isfinite(x).all() for x in got)


def test_noalloc_continues_a_run(ttsh):
    """Two calls of k steps, passing the returned state and predictors on,
    equal one ttsh_step run of 2k steps (stale halo on)."""
    N, k = 48, 7
    log(f"[noalloc continue N={N} k={k}]")
    p = ttsh.default_params(N=N, dt0=0.01)
    phi, P, v = ttsh.init_droplet(N, 14.0, 2.0, 1, 779)
    state = tuple(np.asarray(x, np.float32) for x in (phi, P, v))
    prev = (state[2].copy(), state[1].copy())
    mid, mid_last = run_noalloc(ttsh, p, state, prev, k)
    got, _ = run_noalloc(ttsh, p, mid, mid_last, k)
    ref, _ = run_step(ttsh, p, state, prev, 2 * k)
    same(got, ref, "continued")


def test_noalloc_rejects_large_grid(ttsh):
    p = ttsh.default_params(N=96)
    with pytest.raises(ttsh.TTSHError, match="exceeds its maximum"):
        ttsh.step_noalloc(p, 1, 1, 1, 1, 1, 1)
