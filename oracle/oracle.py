"""ctypes wrapper around libttsh_oracle.so, the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.

The oracle restates the explicit-Euler step EEuler() of
bacterial_droplet_12-16/TTSH3D.cu:347-359 in FP64 (see ttsh_oracle.c).
Arrays are numpy float64, shape (N, N, N) indexed [z, y, x] (x fastest), the
reference's interior ordering (TTSH3D.cu:477 without the Nb halo).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libttsh_oracle.so")
_lib = None


class OrcParams(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("h", C.c_double), ("dt0", C.c_double),
        ("mu", C.c_double), ("cA", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
        ("eta", C.c_double), ("La", C.c_double), ("kappa", C.c_double), ("lambda0", C.c_double),
        ("aPhi", C.c_double), ("kPhi", C.c_double), ("gamma", C.c_double), ("A", C.c_double),
        ("faithful", C.c_int),
    ]


def build() -> str:
    """Compile the oracle in place (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p, dp, dp, dp, C.c_void_p, C.c_void_p]
        L.orc_get_state.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_get_aux.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_reductions.argtypes = [C.c_void_p, dp]
        L.orc_stencil.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_fft3d.argtypes = [C.c_int, dp, dp, C.c_int]
        L.orc_fft1d.argtypes = [C.c_int, dp, dp, C.c_int]
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_sizeof_params.restype = C.c_size_t
        L.orc_projection.argtypes = [C.POINTER(OrcParams), dp, dp, dp]
        assert L.orc_sizeof_params() == C.sizeof(OrcParams)
        _lib = L
    return _lib


# Physical defaults of run_cuda.pl:30-44 (mu, cA, A, eta, La, alpha, beta,
# lambda_0, kappa, aPhi, kPhi, gamma).
DEFAULTS = dict(h=1.0, dt0=0.01, mu=1.0, cA=-1.0, alpha=-0.02, beta=4.366, eta=-0.187,
                La=1.0, kappa=-0.364, lambda0=9.0, aPhi=1.0, kPhi=1.0, gamma=1.0, A=0.0)


def make_params(N: int, faithful: bool = True, **kw) -> OrcParams:
    d = dict(DEFAULTS)
    d.update(kw)
    unknown = set(d) - {f[0] for f in OrcParams._fields_}
    if unknown:
        raise KeyError(f"unknown oracle parameters {sorted(unknown)}")
    p = OrcParams()
    p.N = int(N)
    for k, v in d.items():
        setattr(p, k, float(v))
    p.faithful = 1 if faithful else 0
    return p


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One FP64 droplet state stepped by the oracle."""

    def __init__(self, N: int, faithful: bool = True, **params):
        self.N = int(N)
        self.params = make_params(N, faithful, **params)
        self._L = lib()
        self._h = self._L.orc_create(C.byref(self.params))
        if not self._h:
            raise MemoryError("orc_create failed")

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, phi, P, v, vs_prev=None, ps_prev=None):
        N = self.N
        phi = _c(phi).reshape(N, N, N)
        P = _c(P).reshape(3, N, N, N)
        v = _c(v).reshape(3, N, N, N)
        keep = []
        def opt(a):
            if a is None:
                return None
            a = _c(a).reshape(3, N, N, N)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)
        self._L.orc_set_state(self._h, phi, P, v, opt(vs_prev), opt(ps_prev))

    def get_state(self):
        N = self.N
        phi = np.empty((N, N, N)); P = np.empty((3, N, N, N)); v = np.empty((3, N, N, N))
        self._L.orc_get_state(self._h, phi, P, v)
        return phi, P, v

    _AUX = {"vstar": (0, 3), "pstar": (1, 3), "p_v": (2, 1), "p_p": (3, 1), "dtv": (4, 3),
            "dtP": (5, 3), "dtphi": (6, 1), "mu0": (7, 1), "P2": (8, 1), "mu": (9, 3),
            "Sigma": (10, 9), "G": (11, 9), "sdiv_v": (12, 1), "sdiv_p": (13, 1)}

    def aux(self, name: str):
        which, k = self._AUX[name]
        N = self.N
        out = np.empty((k, N, N, N))
        self._L.orc_get_aux(self._h, which, out)
        return out[0] if k == 1 else out

    def step(self, n: int = 1):
        self._L.orc_step(self._h, int(n))

    def reductions(self):
        out = np.empty(3)
        self._L.orc_reductions(self._h, out)
        return out


def stencil(which: str, f, h: float = 1.0):
    """Apply d1x/d1y/d1z/lap (TTSH3D.cu:3280-3343) to a periodic (N,N,N) field."""
    f = _c(f)
    N = f.shape[0]
    out = np.empty_like(f)
    lib().orc_stencil({"d1x": 0, "d1y": 1, "d1z": 2, "lap": 3}[which], N, float(h), f, out)
    return out


def projection(Xs, **params):
    """getP + corrector on a given predictor (3,N,N,N), clean halo: returns (p, X)."""
    Xs = _c(Xs)
    N = Xs.shape[1]
    prm = make_params(N, False, **params)
    p = np.empty(Xs.shape[1:])
    X = np.empty_like(Xs)
    lib().orc_projection(C.byref(prm), Xs, p, X)
    return p, X


def fft3d(a, inverse: bool = False):
    """Unnormalised 3D DFT of a complex (N,N,N) array (e^{-i} forward, e^{+i} inverse)."""
    a = np.asarray(a, dtype=np.complex128)
    N = a.shape[0]
    re = np.ascontiguousarray(a.real); im = np.ascontiguousarray(a.imag)
    lib().orc_fft3d(N, re, im, 1 if inverse else -1)
    return re + 1j * im


def fft1d(a, inverse: bool = False):
    a = np.asarray(a, dtype=np.complex128)
    re = np.ascontiguousarray(a.real); im = np.ascontiguousarray(a.imag)
    lib().orc_fft1d(a.shape[0], re, im, 1 if inverse else -1)
    return re + 1j * im


def threads() -> int:
    return lib().orc_threads()


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))
