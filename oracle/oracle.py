"""ctypes wrapper of the CPU FP64 oracle (oracle/ctpwa_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctpwa_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE, "libctpwa_oracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                 _dp, ctypes.c_int, _dp,
                                 _dp, ctypes.c_int, _dp,
                                 _dp, ctypes.c_int, _dp]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        for f in ("orc_nll_grad",):
            getattr(L, f).argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_amplitudes.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.orc_intensity.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.orc_slamp.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_dims.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 4
        L.orc_cg.restype = ctypes.c_double
        L.orc_cg.argtypes = [ctypes.c_int] * 6
        L.orc_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_dfunc.restype = ctypes.c_double
        L.orc_dfunc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_factorial.restype = ctypes.c_double
        L.orc_factorial.argtypes = [ctypes.c_double]
        L.orc_bf.restype = ctypes.c_double
        L.orc_bf.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_bwr.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_flatte.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.orc_gs.argtypes = [ctypes.c_double] * 6 + [_dp]
        L.orc_hist.restype = ctypes.c_double
        L.orc_hist.argtypes = [_dp, ctypes.c_int, ctypes.c_double]
        L.orc_interp.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_massive_trans.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.orc_massless_trans.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.orc_node_tensor.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return ctypes.cast(None, _dp)
    return a.ctypes.data_as(_dp)


def _prep(m):
    if m is None:
        return None, 0
    m = np.ascontiguousarray(m, dtype=np.float64)
    return m, m.shape[0]


class Oracle:
    """FP64 reference likelihood on the CPU for one plan and one set of samples."""

    def __init__(self, plan_text: str, nfinal: int, phsp, data, bkg=None,
                 w_phsp=None, w_data=None, w_bkg=None, threads: int = 0):
        L = lib()
        if threads:
            L.orc_set_threads(threads)
        self._keep = []
        args = []
        for m, w in ((phsp, w_phsp), (data, w_data), (bkg, w_bkg)):
            m, n = _prep(m)
            w = None if w is None else np.ascontiguousarray(np.broadcast_to(w, (n,)), dtype=np.float64)
            self._keep += [m, w]
            args += [_ptr(m), n, _ptr(w)]
        self.h = L.orc_create(plan_text.encode(), nfinal, *args)
        if not self.h:
            raise RuntimeError("oracle construction failed")
        a, f, t, p = (ctypes.c_int() for _ in range(4))
        L.orc_dims(self.h, ctypes.byref(a), ctypes.byref(f), ctypes.byref(t), ctypes.byref(p))
        self.n_amp, self.n_free, self.n_theta, self.n_pol = a.value, f.value, t.value, p.value
        self.n_events = [0 if x is None else len(x) for x in (phsp, data, bkg)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def n_params(self):
        return 2 * self.n_free + self.n_theta

    def nll_grad(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        nll = ctypes.c_double()
        g = np.zeros(self.n_params)
        lib().orc_nll_grad(self.h, _ptr(p), ctypes.byref(nll), _ptr(g))
        return nll.value, g

    def hessian(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        nll = ctypes.c_double()
        g = np.zeros(self.n_params)
        H = np.zeros((self.n_params, self.n_params))
        lib().orc_hessian(self.h, _ptr(p), ctypes.byref(nll), _ptr(g), _ptr(H))
        return nll.value, g, H

    def amplitudes(self, params, cls: int):
        """A[e, p, a] complex128 for class cls (0 phsp, 1 data, 2 bkg)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        n = self.n_events[cls]
        out = np.zeros((n, self.n_pol, self.n_amp, 2))
        lib().orc_amplitudes(self.h, _ptr(p), cls, _ptr(out))
        return out[..., 0] + 1j * out[..., 1]

    def slamp(self, cls: int, topo: int, n_slc: int):
        """Contracted node tensors of one topology, before lineshapes and barriers:
        S[c, e, p] complex128 (SL combination c, event e, polarisation p)."""
        n = self.n_events[cls]
        out = np.zeros((n_slc, n, self.n_pol, 2))
        if lib().orc_slamp(self.h, cls, topo, _ptr(out)) != 0:
            raise ValueError("no topology %d" % topo)
        return out[..., 0] + 1j * out[..., 1]

    def intensity(self, params, cls: int):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(self.n_events[cls])
        lib().orc_intensity(self.h, _ptr(p), cls, _ptr(out))
        return out


# ---- component functions (pinned by tests against independent references) ----
def cg(dj1, m1x2, dj2, m2x2, dJ, Mx2):
    return lib().orc_cg(dj1, m1x2, dj2, m2x2, dJ, Mx2)


def ylm(l, m, theta, phi):
    out = np.zeros(2)
    lib().orc_ylm(l, m, theta, phi, _ptr(out))
    return complex(out[0], out[1])


def dfunc(dim, m1x2, m2x2, beta):
    return lib().orc_dfunc(dim, m1x2, m2x2, beta)


def factorial(n):
    return lib().orc_factorial(n)


def bf(L, q, q0, d):
    return lib().orc_bf(L, q, q0, d)


def bwr(m, m0, g0, L, q, q0, d):
    out = np.zeros(2)
    lib().orc_bwr(m, m0, g0, L, q, q0, d, _ptr(out))
    return complex(out[0], out[1])


def flatte(m, m0, g, ch):
    g = np.ascontiguousarray(g, dtype=np.float64)
    ch = np.ascontiguousarray(ch, dtype=np.float64).reshape(-1)
    out = np.zeros(2)
    lib().orc_flatte(m, m0, len(g), _ptr(g), _ptr(ch), _ptr(out))
    return complex(out[0], out[1])


def gs(m, m0, g0, q, q0, mpi):
    """Gounaris-Sakurai propagator (without the barrier factor)."""
    out = np.zeros(2)
    lib().orc_gs(m, m0, g0, q, q0, mpi, _ptr(out))
    return complex(out[0], out[1])


def node_tensor(p1, dj1, p2, dj2, dj, dS, L):
    p1 = np.ascontiguousarray(p1, dtype=np.float64)
    p2 = np.ascontiguousarray(p2, dtype=np.float64)
    out = np.zeros((dj, dj1, dj2, 2))
    lib().orc_node_tensor(_ptr(p1), dj1, _ptr(p2), dj2, dj, dS, L, _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def _trans(fn, P, p, dim):
    P = np.ascontiguousarray(P, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros((dim, dim, 2))
    fn(_ptr(P), _ptr(p), dim, _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def massive_trans(P, p, dim):
    """Spin-transform matrix of a massive daughter with four-momentum p of a mother with
    four-momentum P (both (E, px, py, pz) in one frame); rows and columns ordered
    m = j, j-1, ..., -j."""
    return _trans(lib().orc_massive_trans, P, p, dim)


def massless_trans(P, p, dim):
    """The diagonal phase matrix of a massless daughter (before the helicity rotation)."""
    return _trans(lib().orc_massless_trans, P, p, dim)


def hist(table, m):
    """Hist lineshape value at mass m for a table [m_min, m_max, n, values...]."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    return lib().orc_hist(_ptr(t), len(t), float(m))


def interp(table, x):
    """Interp lineshape (linear) at x for a table [1, N, x0, dx, re, im] or [-2, N, xs, re, im]."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    out = np.zeros(2)
    lib().orc_interp(_ptr(t), len(t), float(x), _ptr(out))
    return complex(out[0], out[1])
