"""ctypes binding of the native library ``libctpwa_amd.so`` (C ABI in include/ctpwa_amd.h).

The library is built in-tree by ``ctpwa_amd.build`` (or ``__graft_entry__.build()``).
There is no fallback: if the library is missing, importing a GPU entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctpwa_amd.so")
TORCH_EXT_PATH = os.path.join(HERE, "_ctpwa_torch.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_ll = ctypes.c_longlong

# every symbol include/ctpwa_amd.h declares
EXPORTS = [
    "ctpwa_create", "ctpwa_destroy", "ctpwa_dims", "ctpwa_param_info", "ctpwa_nll_grad",
    "ctpwa_hessian", "ctpwa_intensity", "ctpwa_amplitudes", "ctpwa_time_eval_kernel",
    "ctpwa_footprint_bytes", "ctpwa_hessian_workspace_bytes", "ctpwa_layout", "ctpwa_kernel_source", "ctpwa_generate_source",
    "ctpwa_generate_hessian_source",
    "ctpwa_compile_check", "ctpwa_last_error", "ctpwa_version",
    "ctpwa_batch_create", "ctpwa_batch_destroy", "ctpwa_batch_nll_grad",
    "ctpwa_amplitude_gram", "ctpwa_partial_intensities", "ctpwa_partial_amplitudes", "ctpwa_last_theta",
    "ctpwa_nll_grad_noalloc",
]

_lib = None
_ops = None


class NativeError(RuntimeError):
    pass


def torch_ops():
    """The PyTorch extension over the C ABI (``_ctpwa_torch.so``): ``nll(params, handle,
    stream, n_params)``, the likelihood as a C++ autograd function.  It links the same
    ``libctpwa_amd.so`` that :func:`lib` loads (one copy in the process)."""
    global _ops
    if _ops is not None:
        return _ops
    if not os.path.exists(TORCH_EXT_PATH):
        raise NativeError("torch extension %s is not built; run `python -m ctpwa_amd.build`" % TORCH_EXT_PATH)
    import torch  # noqa: F401  (the extension resolves torch's libraries)
    lib()
    from . import _ctpwa_torch
    _ops = _ctpwa_torch
    return _ops


def current_stream(device_index: int) -> int:
    """The current torch stream of a device, as an integer handle for the C ABI (torch's raw
    accessor, about 0.2 us, where it exists; the public stream object otherwise)."""
    import torch
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(device_index)
    return torch.cuda.current_stream(device_index).cuda_stream


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError("native library %s is not built; run `python -m ctpwa_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.ctpwa_create.restype = ctypes.c_int
    L.ctpwa_create.argtypes = [ctypes.c_char_p, _vp, _ll, _vp, _vp, _ll, _vp, _vp, _ll, _vp,
                               ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.ctpwa_destroy.argtypes = [_vp]
    L.ctpwa_destroy.restype = None
    L.ctpwa_dims.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_int)] * 4
    L.ctpwa_param_info.argtypes = [_vp, _dp, _dp, _dp]
    for f in ("ctpwa_nll_grad", "ctpwa_hessian", "ctpwa_nll_grad_noalloc"):
        getattr(L, f).argtypes = [_vp, _vp, _vp, _vp]
    for f in ("ctpwa_intensity", "ctpwa_amplitudes", "ctpwa_amplitude_gram", "ctpwa_partial_intensities",
              "ctpwa_partial_amplitudes"):
        getattr(L, f).argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
    L.ctpwa_last_theta.argtypes = [_vp, _dp, _vp]
    L.ctpwa_time_eval_kernel.argtypes = [_vp, _vp, ctypes.c_int, _vp, _dp]
    L.ctpwa_footprint_bytes.restype = _ll
    L.ctpwa_footprint_bytes.argtypes = [ctypes.c_char_p, _ll, _ll, _ll, ctypes.c_int]
    L.ctpwa_hessian_workspace_bytes.restype = _ll
    L.ctpwa_hessian_workspace_bytes.argtypes = [ctypes.c_char_p, _ll, _ll, _ll]
    L.ctpwa_layout.argtypes = [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_ll), ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.ctpwa_kernel_source.restype = ctypes.c_char_p
    L.ctpwa_kernel_source.argtypes = [_vp]
    L.ctpwa_generate_source.restype = _ll
    L.ctpwa_generate_source.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, _ll]
    L.ctpwa_generate_hessian_source.restype = _ll
    L.ctpwa_generate_hessian_source.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, _ll]
    L.ctpwa_compile_check.restype = _ll
    L.ctpwa_compile_check.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.ctpwa_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.ctpwa_batch_destroy.argtypes = [_vp]
    L.ctpwa_batch_destroy.restype = None
    L.ctpwa_batch_nll_grad.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                       ctypes.c_int, _vp]
    L.ctpwa_last_error.restype = ctypes.c_char_p
    L.ctpwa_version.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error() -> str:
    return lib().ctpwa_last_error().decode(errors="replace")


def check(rc: int, what: str):
    if rc != 0:
        raise NativeError("%s failed: %s" % (what, last_error()))


# storage precision of the per-event inputs (arithmetic is FP64 in both)
FP64 = 0
MIXED = 1
PRECISIONS = {"fp64": FP64, "double": FP64, "mixed": MIXED, "float": MIXED}


def precision_code(precision) -> int:
    if isinstance(precision, int) and precision in (FP64, MIXED):
        return precision
    if isinstance(precision, str) and precision.lower() in PRECISIONS:
        return PRECISIONS[precision.lower()]
    raise ValueError("precision must be 'fp64' or 'mixed', got %r" % (precision,))


def generate_source(plan_text: str, precision=FP64) -> str:
    pc = precision_code(precision)
    n = lib().ctpwa_generate_source(plan_text.encode(), pc, None, 0)
    if n < 0:
        raise NativeError(last_error())
    buf = ctypes.create_string_buffer(int(n))
    lib().ctpwa_generate_source(plan_text.encode(), pc, buf, n)
    return buf.value.decode()


def generate_hessian_source(plan_text: str, precision=FP64) -> str:
    pc = precision_code(precision)
    n = lib().ctpwa_generate_hessian_source(plan_text.encode(), pc, None, 0)
    if n < 0:
        raise NativeError(last_error())
    buf = ctypes.create_string_buffer(int(n))
    lib().ctpwa_generate_hessian_source(plan_text.encode(), pc, buf, n)
    return buf.value.decode()


def compile_check(plan_text: str, precision=FP64) -> int:
    n = lib().ctpwa_compile_check(plan_text.encode(), precision_code(precision))
    if n < 0:
        raise NativeError(last_error())
    return int(n)


def footprint_bytes(plan_text: str, n_phsp: int, n_data: int, n_bkg: int, precision=FP64) -> int:
    n = lib().ctpwa_footprint_bytes(plan_text.encode(), n_phsp, n_data, n_bkg, precision_code(precision))
    if n < 0:
        raise NativeError(last_error())
    return int(n)


def hessian_workspace_bytes(plan_text: str, n_phsp: int, n_data: int, n_bkg: int) -> int:
    n = lib().ctpwa_hessian_workspace_bytes(plan_text.encode(), n_phsp, n_data, n_bkg)
    if n < 0:
        raise NativeError(last_error())
    return int(n)


def _addr(a):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


class Handle:
    """Owns one native likelihood (device buffers, generated kernel)."""

    def __init__(self, plan_text: str, phsp, data, bkg=None, w_phsp=None, w_data=None, w_bkg=None,
                 device: int = 0, precision=FP64):
        import numpy as np
        L = lib()
        arrs = []
        for a in (phsp, w_phsp, data, w_data, bkg, w_bkg):
            arrs.append(None if a is None else np.ascontiguousarray(a, dtype=np.float64))
        phsp, w_phsp, data, w_data, bkg, w_bkg = arrs
        n = [0 if x is None else x.shape[0] for x in (phsp, data, bkg)]
        for w, k in ((w_phsp, 0), (w_data, 1), (w_bkg, 2)):
            if w is not None and w.shape[0] != n[k]:
                raise ValueError("weight array length does not match its sample")
        h = ctypes.c_void_p()
        pc = precision_code(precision)
        rc = L.ctpwa_create(plan_text.encode(), _addr(phsp), n[0], _addr(w_phsp), _addr(data), n[1],
                            _addr(w_data), _addr(bkg), n[2], _addr(w_bkg), pc, device, ctypes.byref(h))
        check(rc, "ctpwa_create")
        self.h = h
        self.device = device
        self.precision = pc
        a, f, t, p = (ctypes.c_int() for _ in range(4))
        check(L.ctpwa_dims(h, ctypes.byref(a), ctypes.byref(f), ctypes.byref(t), ctypes.byref(p)), "ctpwa_dims")
        self.n_amp, self.n_free, self.n_theta, self.n_pol = a.value, f.value, t.value, p.value
        self.n_events = n

    def close(self):
        if getattr(self, "h", None):
            lib().ctpwa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return 2 * self.n_free + self.n_theta

    def param_info(self):
        import numpy as np
        init, lo, hi = (np.zeros(self.n_theta) for _ in range(3))
        check(lib().ctpwa_param_info(self.h, init.ctypes.data_as(_dp), lo.ctypes.data_as(_dp),
                                     hi.ctypes.data_as(_dp)), "ctpwa_param_info")
        return init, lo, hi

    def nll_grad(self, params_ptr: int, out_ptr: int, stream: int):
        check(lib().ctpwa_nll_grad(self.h, ctypes.c_void_p(params_ptr), ctypes.c_void_p(out_ptr),
                                   ctypes.c_void_p(stream)), "ctpwa_nll_grad")

    def nll_grad_noalloc(self, params_ptr: int, out_ptr: int, stream: int):
        check(lib().ctpwa_nll_grad_noalloc(self.h, ctypes.c_void_p(params_ptr), ctypes.c_void_p(out_ptr),
                                           ctypes.c_void_p(stream)), "ctpwa_nll_grad_noalloc")

    def hessian(self, params_ptr: int, out_ptr: int, stream: int):
        check(lib().ctpwa_hessian(self.h, ctypes.c_void_p(params_ptr), ctypes.c_void_p(out_ptr),
                                  ctypes.c_void_p(stream)), "ctpwa_hessian")

    def intensity(self, params_ptr: int, cls: int, out_ptr: int, stream: int):
        check(lib().ctpwa_intensity(self.h, ctypes.c_void_p(params_ptr), cls, ctypes.c_void_p(out_ptr),
                                    ctypes.c_void_p(stream)), "ctpwa_intensity")

    def amplitudes(self, params_ptr: int, cls: int, out_ptr: int, stream: int):
        check(lib().ctpwa_amplitudes(self.h, ctypes.c_void_p(params_ptr), cls, ctypes.c_void_p(out_ptr),
                                     ctypes.c_void_p(stream)), "ctpwa_amplitudes")

    def amplitude_gram(self, params_ptr: int, cls: int, out_ptr: int, stream: int):
        """complex [n_amp][n_amp] sum over the class of conj(A_l) A_k (no couplings)"""
        check(lib().ctpwa_amplitude_gram(self.h, ctypes.c_void_p(params_ptr), cls, ctypes.c_void_p(out_ptr),
                                         ctypes.c_void_p(stream)), "ctpwa_amplitude_gram")

    def partial_intensities(self, params_ptr: int, cls: int, out_ptr: int, stream: int):
        """[n_wave + 1][n_cls]: per-event partial-wave intensities, then the total"""
        check(lib().ctpwa_partial_intensities(self.h, ctypes.c_void_p(params_ptr), cls, ctypes.c_void_p(out_ptr),
                                              ctypes.c_void_p(stream)), "ctpwa_partial_intensities")

    def last_theta(self, stream: int):
        """theta of the most recent evaluation (host float64 [ntheta])"""
        import numpy as np
        out = np.zeros(max(self.n_theta, 1))
        check(lib().ctpwa_last_theta(self.h, out.ctypes.data_as(_dp), ctypes.c_void_p(stream)), "ctpwa_last_theta")
        return out[:self.n_theta].copy()

    def partial_amplitudes(self, params_ptr: int, cls: int, out_ptr: int, stream: int):
        """[n_wave][n_cls][n_pol][2]: per-event coherent partial-wave amplitudes"""
        check(lib().ctpwa_partial_amplitudes(self.h, ctypes.c_void_p(params_ptr), cls, ctypes.c_void_p(out_ptr),
                                              ctypes.c_void_p(stream)), "ctpwa_partial_amplitudes")

    def time_eval_kernel(self, params_ptr: int, iters: int, stream: int) -> float:
        ms = ctypes.c_double()
        check(lib().ctpwa_time_eval_kernel(self.h, ctypes.c_void_p(params_ptr), iters, ctypes.c_void_p(stream),
                                           ctypes.byref(ms)), "ctpwa_time_eval_kernel")
        return ms.value

    def layout(self):
        n3 = (_ll * 3)()
        b = _ll()
        nu, ni, nw, nps = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().ctpwa_layout(self.h, n3, ctypes.byref(b), ctypes.byref(nu), ctypes.byref(ni),
                                 ctypes.byref(nw), ctypes.byref(nps)), "ctpwa_layout")
        return dict(n_events=list(n3), bytes_per_eval=b.value, n_sl_tensors=nu.value,
                    n_invariants=ni.value, n_workgroups=nw.value, n_pol_stored=nps.value)

    def kernel_source(self) -> str:
        return lib().ctpwa_kernel_source(self.h).decode()


class Batch:
    """Evaluates many handles of one model in one set of launches (ctpwa_batch_nll_grad)."""

    def __init__(self, max_jobs: int, device: int = 0):
        b = ctypes.c_void_p()
        check(lib().ctpwa_batch_create(device, max_jobs, ctypes.byref(b)), "ctpwa_batch_create")
        self.b = b
        self.max_jobs = max_jobs
        self._hs = (_vp * max_jobs)()
        self._ps = (_vp * max_jobs)()
        self._os = (_vp * max_jobs)()

    def nll_grad(self, handles, params_ptrs, out_ptrs, stream: int):
        n = len(handles)
        if n > self.max_jobs:
            raise ValueError("batch holds at most %d likelihoods" % self.max_jobs)
        for j in range(n):
            self._hs[j] = handles[j].h.value
            self._ps[j] = params_ptrs[j]
            self._os[j] = out_ptrs[j]
        check(lib().ctpwa_batch_nll_grad(self.b, self._hs, self._ps, self._os, n, ctypes.c_void_p(stream)),
              "ctpwa_batch_nll_grad")

    def close(self):
        if getattr(self, "b", None):
            lib().ctpwa_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
