"""Python binding of libttsh3d, the MI355X-native TTSH3D time step.

This is the host-side mirror of the reference's interface
(bacterial_droplet_12-16/TTSH3D.cu and run_cuda.pl): the input.dat parameter
block (GetInput, TTSH3D.cu:696-752), one droplet stepped by EEuler
(TTSH3D.cu:347-359), the conf_<m>.dat snapshots (ExpoConf, TTSH3D.cu:899-1014)
and the evolve() time loop (TTSH3D.cu:290-343), plus batches of droplets.
Every call goes through the C ABI declared in include/ttsh3d.h; the compute
runs in the HIP kernels of libttsh3d.so.  There is no CPU fallback: when the
library is missing, importing the binding raises.

Arrays are numpy, shape (N, N, N) indexed [z, y, x] (x fastest) for scalars
and (3, N, N, N) for vectors.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libttsh3d.so")

OK, EINVAL, EIO, ESENTINEL, EUNSUPPORTED, EHIP, ENOMEM, EBUDGET, ENONFINITE, EFULL = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)
F32, F64 = 0, 1

# input.dat order (TTSH3D.cu:702-739); readDraw.py:74-77 names them the same way.
INPUT_FIELDS = ["TimeScheme", "RKStage", "ExplImpi", "AdapTime", "ExpoData", "InitCond",
                "isSimplified", "checkUnStable", "Nx", "Ny", "Nz", "Nb", "h", "dt0", "T0", "Ts",
                "dte", "dtVarC", "dtMax", "mu", "cA", "alpha", "beta", "eta", "La", "kappa",
                "lambda0", "aPhi", "kPhi", "gamma", "A", "init0", "ReadSuccess"]


class TTSHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ttsh3d error {code}: {msg}")
        self.code = code


class Centroid(C.Structure):
    """ttsh_centroid: idx and c are (y, x, z), the snapshot array axes."""
    _fields_ = [("idx", C.c_int32 * 3), ("c", C.c_double * 3), ("mass", C.c_double)]


# ttsh_run_ex output flags
RUN_TEXT, RUN_BINARY, RUN_DIAG, RUN_SYNC = 1, 2, 4, 8
# fields of ttsh_get_slice, in order
SLICE_FIELDS = ("phi", "vx", "vy", "vz", "px", "py", "pz", "vort_v", "vort_p")


class Params(C.Structure):
    """Mirror of struct ttsh_params (include/ttsh3d.h)."""
    _fields_ = ([(n, C.c_int32) for n in INPUT_FIELDS[:12]]
                + [(n, C.c_double) for n in INPUT_FIELDS[12:32]]
                + [("ReadSuccess", C.c_int32), ("faithful_halo", C.c_int32),
                   ("rhs_variant", C.c_int32), ("rhs_tz", C.c_int32), ("corr_tz", C.c_int32)])

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


_lib = None


def lib():
    """Load libttsh3d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C bacterial-droplet_amd` "
                          "or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(Params)
    vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "ttsh_params_default": (C.c_int, [P]),
        "ttsh_parse_input_dat": (C.c_int, [C.c_char_p, P]),
        "ttsh_write_input_dat": (C.c_int, [C.c_char_p, P]),
        "ttsh_validate_params": (C.c_int, [P]),
        "ttsh_create": (C.c_int, [P, C.c_int, C.POINTER(vp)]),
        "ttsh_destroy": (C.c_int, [vp]),
        "ttsh_set_state": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
        "ttsh_get_state": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "ttsh_step": (C.c_int, [vp, i64]),
        "ttsh_synchronize": (C.c_int, [vp]),
        "ttsh_reductions": (C.c_int, [vp, dp]),
        "ttsh_write_conf": (C.c_int, [vp, C.c_char_p]),
        "ttsh_write_conf_arrays": (C.c_int, [C.c_char_p, C.c_int, C.c_int, vp, vp, vp]),
        "ttsh_run": (C.c_int, [vp, C.c_char_p, C.POINTER(i64)]),
        "ttsh_run_ex": (C.c_int, [vp, C.c_char_p, C.c_int, C.POINTER(i64)]),
        "ttsh_get_centroid": (C.c_int, [vp, C.POINTER(Centroid)]),
        "ttsh_get_slice": (C.c_int, [vp, C.c_int, vp, C.POINTER(C.c_int)]),
        "ttsh_contour_radius": (C.c_int, [C.c_int, vp, C.c_double, dp, dp]),
        "ttsh_get_clock": (C.c_int, [vp, C.POINTER(i64), dp]),
        "ttsh_profile": (C.c_int, [vp, i64, dp]),
        "ttsh_kernel_name": (C.c_char_p, [C.c_int]),
        "ttsh_get_aux": (C.c_int, [vp, C.c_int, vp]),
        "ttsh_init_reference": (C.c_int, [P, C.c_uint64, vp, vp, vp, vp, vp]),
        "ttsh_init_droplet": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, C.c_uint64, vp, vp, vp]),
        "ttsh_droplet_bytes": (C.c_size_t, [C.c_int]),
        "ttsh_batch_create": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "ttsh_batch_destroy": (C.c_int, [vp]),
        "ttsh_batch_admit": (C.c_int, [vp, P, C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_int)]),
        "ttsh_batch_step": (C.c_int, [vp, i64]),
        "ttsh_batch_set_limit": (C.c_int, [vp, C.c_int, i64]),
        "ttsh_batch_steps_done": (C.c_int, [vp, C.c_int, C.POINTER(i64)]),
        "ttsh_batch_get_state": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp]),
        "ttsh_batch_reductions": (C.c_int, [vp, C.c_int, dp]),
        "ttsh_batch_retire": (C.c_int, [vp, C.c_int]),
        "ttsh_batch_usage": (C.c_int, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]),
        "ttsh_batch_synchronize": (C.c_int, [vp]),
        "ttsh_batch_wait": (C.c_int, [vp, C.c_int]),
        "ttsh_batch_ready": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
        "ttsh_step_noalloc": (C.c_int, [P, vp, vp, vp, vp, vp, i64, vp]),
        "ttsh_tuning_reload": (C.c_int, [C.c_char_p]),
        "ttsh_tuning_lookup": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ttsh_tuning_lookup_batch": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ttsh_last_error": (C.c_char_p, []),
        "ttsh_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names of the C entry points this binding binds."""
    lib()
    return [n for n in dir(_lib) if n.startswith("ttsh_")]


def _check(rc: int):
    if rc != OK:
        raise TTSHError(rc, lib().ttsh_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def default_params(**kw) -> Params:
    """run_cuda.pl:7-44 defaults; keyword overrides (N sets Nx=Ny=Nz)."""
    p = Params()
    _check(lib().ttsh_params_default(C.byref(p)))
    if "N" in kw:
        n = int(kw.pop("N"))
        p.Nx = p.Ny = p.Nz = n
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def parse_input_dat(path: str, base: Params | None = None) -> Params:
    """GetInput (TTSH3D.cu:696-752). Raises TTSHError(ESENTINEL) on a bad sentinel."""
    p = base.copy() if base is not None else default_params()
    _check(lib().ttsh_parse_input_dat(os.fsencode(path), C.byref(p)))
    return p


def write_input_dat(path: str, p: Params) -> None:
    _check(lib().ttsh_write_input_dat(os.fsencode(path), C.byref(p)))


def validate(p: Params) -> None:
    _check(lib().ttsh_validate_params(C.byref(p)))


def write_conf_arrays(path: str, phi, P, v, expo_data: int = 1) -> None:
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    P = np.ascontiguousarray(P, dtype=np.float32)
    v = np.ascontiguousarray(v, dtype=np.float32)
    N = phi.shape[0]
    _check(lib().ttsh_write_conf_arrays(os.fsencode(path), N, expo_data, _ptr(phi), _ptr(P), _ptr(v)))


def init_reference(p: Params, seed: int):
    """InitCond == 1 initial state (TTSH3D.cu:798-825) from a seeded stream.
    Returns phi, P, v, vstar, pstar (float64)."""
    N = p.Nx
    phi = np.empty((N, N, N)); P = np.empty((3, N, N, N)); v = np.empty((3, N, N, N))
    vs = np.empty((3, N, N, N)); ps = np.empty((3, N, N, N))
    _check(lib().ttsh_init_reference(C.byref(p), C.c_uint64(seed), _ptr(phi), _ptr(P), _ptr(v), _ptr(vs), _ptr(ps)))
    return phi, P, v, vs, ps


def init_droplet(N: int, R: float, xi: float, pmode: int = 0, seed: int = 0):
    """Benchmark droplet: tanh profile, random-unit (pmode 0) or aster (pmode 1) P, v = 0."""
    phi = np.empty((N, N, N)); P = np.empty((3, N, N, N)); v = np.empty((3, N, N, N))
    _check(lib().ttsh_init_droplet(int(N), float(R), float(xi), int(pmode), C.c_uint64(seed),
                                   _ptr(phi), _ptr(P), _ptr(v)))
    return phi, P, v


def tuning_reload(path: str | None = None) -> None:
    """Re-read the tuning table (None: $TTSH_TUNE_FILE or tuned.txt beside the library)."""
    _check(lib().ttsh_tuning_reload(path.encode() if path else None))


def tuning_lookup(N: int):
    """(rhs_variant, rhs_tz, corr_tz) of size N, or None without an entry."""
    v, a, b = C.c_int(), C.c_int(), C.c_int()
    if lib().ttsh_tuning_lookup(int(N), C.byref(v), C.byref(a), C.byref(b)) != OK:
        return None
    return v.value, a.value, b.value


def tuning_lookup_batch():
    """(rhs_variant, rhs_tz, corr_tz) of a device-filling batch, or None without an entry."""
    v, a, b = C.c_int(), C.c_int(), C.c_int()
    if lib().ttsh_tuning_lookup_batch(C.byref(v), C.byref(a), C.byref(b)) != OK:
        return None
    return v.value, a.value, b.value


def step_noalloc(params: "Params", phi_ptr: int, P_ptr: int, v_ptr: int, vstar_prev_ptr: int, pstar_prev_ptr: int,
                 n: int, stream: int = 0) -> None:
    """ttsh_step_noalloc: n steps of a droplet whose FP32 state lives in
    caller-owned device memory (raw device addresses; phi N^3, the others
    3*N^3), with no device allocation by the library.  On return P_ptr..
    hold the new state and the *_prev buffers the last step's predictors."""
    _check(lib().ttsh_step_noalloc(C.byref(params), C.c_void_p(phi_ptr), C.c_void_p(P_ptr), C.c_void_p(v_ptr),
                                   C.c_void_p(vstar_prev_ptr), C.c_void_p(pstar_prev_ptr), int(n),
                                   C.c_void_p(stream) if stream else None))


def contour_radius(field, level: float = 0.5):
    """(R[120], (cx, cy)): ttsh_contour_radius of a [row][col] field."""
    f = np.ascontiguousarray(field, np.float32)
    assert f.ndim == 2 and f.shape[0] == f.shape[1]
    R = np.empty(120, np.float64)
    c = np.empty(2, np.float64)
    _check(lib().ttsh_contour_radius(f.shape[0], _ptr(f), level, R.ctypes.data_as(C.POINTER(C.c_double)),
                                     c.ctypes.data_as(C.POINTER(C.c_double))))
    return R, (c[0], c[1])


def read_conf_bin(path: str):
    """Fields of a conf_<m>.bin sidecar: dict of vx..phi as (N, N, N) [z][y][x], and m, T, steps."""
    with open(path, "rb") as f:
        raw = f.read()
    assert raw[:8] == b"TTSH3DB1", "not a TTSH3D binary snapshot"
    N, m = np.frombuffer(raw, np.int32, 2, 8)
    T = np.frombuffer(raw, np.float64, 1, 16)[0]
    steps = np.frombuffer(raw, np.int64, 1, 24)[0]
    n = int(N) ** 3
    data = np.frombuffer(raw, np.float32, 7 * n, 32).reshape(7, N, N, N)
    names = ("vx", "vy", "vz", "px", "py", "pz", "phi")
    return {k: data[i] for i, k in enumerate(names)}, int(m), float(T), int(steps)


def droplet_bytes(N: int) -> int:
    return int(lib().ttsh_droplet_bytes(int(N)))


def _arr(a, dt, shape):
    a = np.ascontiguousarray(a, dtype=np.float32 if dt == F32 else np.float64)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape} values, got {a.shape}")
    return a


class Simulation:
    """One droplet on one GPU (ttsh_create ... ttsh_destroy)."""

    def __init__(self, params: Params, device: int = 0):
        self.params = params.copy()
        self.N = params.Nx
        h = C.c_void_p()
        _check(lib().ttsh_create(C.byref(self.params), int(device), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ttsh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, phi, P, v, vstar_prev=None, pstar_prev=None, dtype: int = F64):
        N = self.N
        s1, s3 = (N, N, N), (3, N, N, N)
        a = [_arr(phi, dtype, s1), _arr(P, dtype, s3), _arr(v, dtype, s3),
             None if vstar_prev is None else _arr(vstar_prev, dtype, s3),
             None if pstar_prev is None else _arr(pstar_prev, dtype, s3)]
        _check(lib().ttsh_set_state(self._h, dtype, *[_ptr(x) for x in a]))

    def get_state(self, dtype: int = F32):
        N = self.N
        t = np.float32 if dtype == F32 else np.float64
        phi = np.empty((N, N, N), t); P = np.empty((3, N, N, N), t); v = np.empty((3, N, N, N), t)
        _check(lib().ttsh_get_state(self._h, dtype, _ptr(phi), _ptr(P), _ptr(v)))
        return phi, P, v

    def step(self, n: int = 1):
        _check(lib().ttsh_step(self._h, int(n)))

    def synchronize(self):
        _check(lib().ttsh_synchronize(self._h))

    def reductions(self):
        out = (C.c_double * 3)()
        rc = lib().ttsh_reductions(self._h, out)
        if rc not in (OK, ENONFINITE):
            _check(rc)
        return np.array(out[:]), rc == OK

    def profile(self, n: int):
        """Run n steps with HIP events between kernel kinds; returns {kernel: total ms}."""
        out = (C.c_double * 8)()
        _check(lib().ttsh_profile(self._h, int(n), out))
        return {lib().ttsh_kernel_name(k).decode(): out[k] for k in range(8)}

    def clock(self):
        s = C.c_int64(); T = C.c_double()
        _check(lib().ttsh_get_clock(self._h, C.byref(s), C.byref(T)))
        return s.value, T.value

    def aux(self, which: str):
        N = self.N
        if which in ("vstar", "pstar"):
            out = np.empty((3, N, N, N), np.float32)
            _check(lib().ttsh_get_aux(self._h, 0 if which == "vstar" else 1, _ptr(out)))
            return out
        if which == "p":
            c = np.empty((N, N, N, 2), np.float32)
            _check(lib().ttsh_get_aux(self._h, 2, _ptr(c)))
            return c[..., 0].astype(np.float64), c[..., 1].astype(np.float64)
        raise KeyError(which)

    def write_conf(self, path: str):
        _check(lib().ttsh_write_conf(self._h, os.fsencode(path)))

    def run(self, run_dir: str, text: bool = True, binary: bool = False, diag: bool = False,
            sync: bool = False) -> int:
        """The reference's time loop with exports (ttsh_run_ex); returns the steps taken."""
        flags = (RUN_TEXT if text else 0) | (RUN_BINARY if binary else 0) | (RUN_DIAG if diag else 0) \
            | (RUN_SYNC if sync else 0)
        n = C.c_int64()
        _check(lib().ttsh_run_ex(self._h, os.fsencode(run_dir), flags, C.byref(n)))
        return n.value

    def centroid(self):
        """(idx, c, mass): the phi-weighted centroid over the (y, x, z) axes."""
        c = Centroid()
        _check(lib().ttsh_get_centroid(self._h, C.byref(c)))
        return tuple(c.idx), tuple(c.c), c.mass

    def slice(self, z: int = -1):
        """(fields, z): plane z (-1: the centroid plane) of SLICE_FIELDS, each [y][x]."""
        N = self.N
        out = np.empty((9, N, N), np.float32)
        zu = C.c_int()
        _check(lib().ttsh_get_slice(self._h, int(z), _ptr(out), C.byref(zu)))
        return {k: out[i] for i, k in enumerate(SLICE_FIELDS)}, zu.value


class Batch:
    """Many droplets stepped together on one GPU (ttsh_batch_*)."""

    def __init__(self, device: int = 0, max_slots: int = 64, hbm_budget: int = 0, use_graph: bool = True):
        h = C.c_void_p()
        _check(lib().ttsh_batch_create(int(device), int(max_slots), C.c_size_t(int(hbm_budget)),
                                       1 if use_graph else 0, C.byref(h)))
        self._h = h
        self.N = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().ttsh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def admit(self, params: Params, phi, P, v, vstar_prev=None, pstar_prev=None, dtype: int = F64,
              limit: int | None = None) -> int:
        N = params.Nx
        s1, s3 = (N, N, N), (3, N, N, N)
        a = [_arr(phi, dtype, s1), _arr(P, dtype, s3), _arr(v, dtype, s3),
             None if vstar_prev is None else _arr(vstar_prev, dtype, s3),
             None if pstar_prev is None else _arr(pstar_prev, dtype, s3)]
        slot = C.c_int()
        _check(lib().ttsh_batch_admit(self._h, C.byref(params), dtype, *[_ptr(x) for x in a], C.byref(slot)))
        self.N[slot.value] = N
        if limit is not None:
            self.set_limit(slot.value, limit)
        return slot.value

    def try_admit(self, *a, **kw):
        """admit(), but returns None when the budget or the slots are exhausted."""
        try:
            return self.admit(*a, **kw)
        except TTSHError as e:
            if e.code in (EBUDGET, EFULL):
                return None
            raise

    def step(self, n: int):
        _check(lib().ttsh_batch_step(self._h, int(n)))

    def set_limit(self, slot: int, max_steps: int):
        _check(lib().ttsh_batch_set_limit(self._h, int(slot), int(max_steps)))

    def steps_done(self, slot: int) -> int:
        s = C.c_int64()
        _check(lib().ttsh_batch_steps_done(self._h, int(slot), C.byref(s)))
        return s.value

    def get_state(self, slot: int, dtype: int = F32):
        N = self.N[slot]
        t = np.float32 if dtype == F32 else np.float64
        phi = np.empty((N, N, N), t); P = np.empty((3, N, N, N), t); v = np.empty((3, N, N, N), t)
        _check(lib().ttsh_batch_get_state(self._h, int(slot), dtype, _ptr(phi), _ptr(P), _ptr(v)))
        return phi, P, v

    def reductions(self, slot: int):
        out = (C.c_double * 3)()
        rc = lib().ttsh_batch_reductions(self._h, int(slot), out)
        if rc not in (OK, ENONFINITE):
            _check(rc)
        return np.array(out[:]), rc == OK

    def wait(self, keep: int = 0):
        """Wait for every step call but the last `keep`."""
        _check(lib().ttsh_batch_wait(self._h, int(keep)))

    def ready(self, slot: int) -> bool:
        """True when the droplet has reached its limit and those steps are done."""
        r = C.c_int()
        _check(lib().ttsh_batch_ready(self._h, int(slot), C.byref(r)))
        return bool(r.value)

    def retire(self, slot: int):
        _check(lib().ttsh_batch_retire(self._h, int(slot)))
        self.N.pop(slot, None)

    def usage(self):
        b = C.c_size_t(); a = C.c_int()
        _check(lib().ttsh_batch_usage(self._h, C.byref(b), C.byref(a)))
        return b.value, a.value

    def synchronize(self):
        _check(lib().ttsh_batch_synchronize(self._h))
