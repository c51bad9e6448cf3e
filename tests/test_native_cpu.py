"""The C-ABI library on a machine without a GPU: it loads, exports every symbol the
header declares, generates and compiles (hipRTC, gfx950) the fused kernel for the
workloads, and reports errors instead of crashing."""
import ctypes
import os
import re

import pytest

from helpers import plan_of

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "ctpwa_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ctpwa_[a-z_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from ctpwa_amd import _native
    L = _native.lib()
    names = _header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_native.EXPORTS)
    assert L.ctpwa_version().decode().startswith("ctpwa_amd")


def test_torch_extension_loads():
    """The PyTorch extension over the C ABI loads beside the library and exposes the
    likelihood's autograd function; a call with a bad handle reports the library's error
    instead of crashing (the handle is checked before any device work)."""
    import torch
    from ctpwa_amd import _native
    ops = _native.torch_ops()
    assert callable(ops.nll)
    p = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    with pytest.raises(RuntimeError, match="ctpwa_nll_grad"):
        ops.nll(p, 0, 0, 3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["jpsi_rho", "jpsi_3pi", "jpsi_kkpipi"])
def test_generated_kernel_compiles_for_gfx950(name, precision):
    from ctpwa_amd import _native
    plan = plan_of(name)
    src = _native.generate_source(plan.to_text(), precision)
    assert "ctpwa_eval" in src and "__launch_bounds__(128" in src
    # every amplitude coefficient appears
    for a in range(plan.n_amp):
        assert "c%dr" % a in src
    # storage types: FP32 SL tensors only in the mixed form
    assert ("const float2* __restrict__ sl" in src) == (precision == "mixed")
    assert _native.compile_check(plan.to_text(), precision) > 1000


def test_equal_denominator_coefficients_are_loaded_once():
    """The three-pion model carries rho(770), rho(1450) and rho(1700) in each of its three
    chains: every rho mass term m0^2 is loaded once (6 aliases), and the BWR width terms of
    the two pi+- pi0 chains (equal daughter masses) are shared (3 aliases) while those of
    the pi+ pi- chain (other daughter masses) are not.  An alias always names a coefficient
    loaded earlier.  The single-chain model has nothing to share."""
    import re
    from ctpwa_amd import _native
    src = _native.generate_source(plan_of("jpsi_3pi").to_text(), "mixed")
    body = src[src.index("eval_body("):]
    aliases = re.findall(r"const double (k\d+_\d+) = (k\d+_\d+);", body)
    loads = re.findall(r"const double (k\d+_\d+) = \(UL \? uld", body)
    assert len(aliases) == 9 and len(loads) == 11
    for dst, srcv in aliases:
        assert srcv in loads and dst not in loads
        assert srcv.endswith(dst[-2:])            # same coefficient index (m0^2 or width term)
    masses = {d: s_ for d, s_ in aliases if d.endswith("_0")}
    assert sorted(set(masses.values())) == ["k0_0", "k1_0", "k2_0"]
    one = _native.generate_source(plan_of("jpsi_rho").to_text(), "fp64")
    assert not re.findall(r"const double k\d+_\d+ = k\d+_\d+;", one)


def test_footprint_and_layout_formula():
    from ctpwa_amd import _native
    plan = plan_of("jpsi_3pi")
    n = _native.footprint_bytes(plan.to_text(), 2_500_000, 500_000, 50_000)
    # 4 SL tensors x 2 polarisations x 16 B + 7 invariants x 8 B per event
    per_event = 4 * 2 * 16 + 7 * 8
    assert n >= 3_050_000 * per_event
    assert n < 3_050_000 * per_event * 1.2
    # mixed: 8 B per complex SL value, 3 squared masses in FP64, 4 other invariants in FP32
    # (the upload staging is shared per device, not part of a handle)
    m = _native.footprint_bytes(plan.to_text(), 2_500_000, 500_000, 50_000, "mixed")
    per_event_m = 4 * 2 * 8 + 3 * 8 + 4 * 4
    assert 3_050_000 * per_event_m <= m < 3_050_000 * per_event_m + (1 << 20)
    with pytest.raises(ValueError):
        _native.footprint_bytes(plan.to_text(), 1, 1, 0, "half")


def test_bad_plan_reports_error():
    from ctpwa_amd import _native
    with pytest.raises(_native.NativeError):
        _native.generate_source("not a plan")
    assert "CTPWA_PLAN" in _native.last_error()
    with pytest.raises(_native.NativeError):
        _native.generate_source("CTPWA_PLAN 1\nnfinal 3\n")


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    import numpy as np
    from ctpwa_amd import _native
    plan = plan_of("jpsi_rho")
    ev = np.zeros((4, 3, 4))
    with pytest.raises(_native.NativeError):
        _native.Handle(plan.to_text(), ev, ev)


def test_package_exposes_analysis_class():
    """``import ctpwa_amd as ctpwa; ctpwa.analysis(...)`` names the class (src/main.cu:90-131)."""
    import ctpwa_amd as ctpwa
    assert isinstance(ctpwa.analysis, type)
    assert ctpwa.analysis.__name__ == "analysis"
    for m in ("getNLL", "getHessian", "getNVector", "getNFreeTheta"):
        assert callable(getattr(ctpwa.analysis, m))


def test_wide_cascade_plan_generates():
    """The 149-amplitude four-body cascade passes the native layout checks (node tensors
    of at most 64 entries, thresholds) and yields both generated kernels' sources (its
    hipRTC compile takes tens of seconds, so the GPU tests cover that)."""
    from ctpwa_amd import _native
    plan = plan_of("jpsi_kkpipi_wide")
    assert (plan.n_amp, plan.n_theta) == (149, 24)
    src = _native.generate_source(plan.to_text(), "fp64")
    hsrc = _native.generate_hessian_source(plan.to_text(), "mixed")
    assert "ctpwa_eval" in src and "ctpwa_hess_rows" in hsrc and "float* __restrict__ J" in hsrc


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus above 1 exits with a message before any device work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
