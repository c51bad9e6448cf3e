"""The generated Custom-lineshape factor kernels (first order, ctpwa_custom; second order,
ctpwa_custom2) compiled as host C++ and run on the CPU, as a check of the emitted code
itself rather than of its GPU build:

* built with every automatic variable initialised to a bit pattern and to zero
  (-ftrivial-auto-var-init), the outputs are identical bit for bit, so no output depends
  on a variable read before it is written;
* the first derivatives equal central differences of the values, and the second
  derivatives central differences of the first;
* the build has no -Wuninitialized / -Wsometimes-uninitialized diagnostics.

Inputs are synthetic (per-event invariants drawn in [0.3, 2.5]); the kernels only need
finite values."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import plan_of
from test_custom_cpu import _custom

CLANG = next((c for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or "") if c and os.path.exists(c)), None)

EXPRS = [
    "BWR(p0, p1, 1, 3.0) * Bf(1, q, q0(p0), 3.0)",
    ("exp(-p1*m) * (cos(p0*m) + i*sin(p1)) / sqrt(p0^2 + m^2) * abs(p0 - i*p1) "
     "+ log(p0) * conj(m + i*p1)^2 - Im(p0 * (q + 1j)) + Re(csqrt(p1 - q0))"),
    "Flatte(p0, p1, 0.13957, 0.13957) * (1 + 0.3*p1costheta - 0.2*p2p + 0.1*sin(p1phi - p2phi) + 0.05*p1e)",
]

SHIM = r'''
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__ __restrict
struct D3 { unsigned x, y, z; };
static D3 blockIdx, threadIdx, gridDim;
using std::exp; using std::cos; using std::sin; using std::log; using std::sqrt; using std::atan2;
'''

MAIN = r'''
// rows: value (r, i), NCV first derivatives (r then i), [NH second derivatives (r then i)]
static void run(const std::vector<double>& tab, const std::vector<double>& inv, long long ld, std::vector<double>& out) {
    gridDim = {1, 1, 1};
    blockIdx = {0, 0, 0};
    for (unsigned t = 0; t < 256; ++t) { threadIdx = {t, 0, 0}; KERNEL(tab.data(), inv.data(), nullptr, ld, out.data()); }
}
int main(int argc, char** argv) {
    const long long ld = 512;
    const int ninv = atoi(argv[1]), nf = atoi(argv[2]), nrow = atoi(argv[3]), second = atoi(argv[4]);
    std::vector<double> inv((size_t)ninv * ld);
    unsigned long long st = 99;
    for (auto& v : inv) { st = st * 6364136223846793005ULL + 1442695040888963407ULL; v = 0.3 + 2.2 * (double)(st >> 11) / 9007199254740992.0; }
    std::vector<double> tab(OFFT + 16, 0.0);
    const double p[2] = {0.775, 0.149};
    for (int k = 0; k < NCV; ++k) tab[OFFT + k] = p[k % 2] * (1.0 + 0.1 * k);
    std::vector<double> o0((size_t)nf * nrow * ld, 0.0);
    run(tab, inv, ld, o0);
    fwrite(o0.data(), 8, o0.size(), stdout);
    for (int k = 0; k < NCV; ++k) {
        const double h = 1e-5 * tab[OFFT + k];
        std::vector<double> tp = tab, tm = tab, op(o0.size(), 0.0), om(o0.size(), 0.0);
        tp[OFFT + k] += h; tm[OFFT + k] -= h;
        run(tp, inv, ld, op); run(tm, inv, ld, om);
        fwrite(&h, 8, 1, stdout);
        fwrite(op.data(), 8, op.size(), stdout);
        fwrite(om.data(), 8, om.size(), stdout);
    }
    return 0;
}
'''


def _module(src, kernel):
    lines = src.split("\n")
    k0 = next(i for i, l in enumerate(lines) if "__global__" in l and (kernel + "(") in l)
    k1 = next(i for i in range(k0 + 1, len(lines)) if lines[i] == "}")
    # the preamble: everything before the first __global__ kernel, without the fused pass
    first = next(i for i, l in enumerate(lines) if "__global__" in l or "eval_body(" in l or "double frcp" in l)
    pre = "\n".join(lines[:first])
    return pre, "\n".join(lines[k0:k1 + 1]).replace('extern "C" ', "")


def _build(tmp_path, pre, kern, kernel, init):
    cpp = tmp_path / ("%s_%s.cpp" % (kernel, init))
    cpp.write_text(SHIM + pre + "\n" + kern + "\n#define KERNEL " + kernel + "\n" + MAIN)
    exe = str(cpp)[:-4]
    r = subprocess.run([CLANG, "-std=c++17", "-O2", "-Wall", "-Wuninitialized", "-Wsometimes-uninitialized",
                        "-Wconditional-uninitialized", "-Wno-unused-variable", "-Wno-unused-parameter",
                        "-ftrivial-auto-var-init=" + init, str(cpp), "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "uninitialized" not in r.stderr, r.stderr
    return exe


@pytest.mark.skipif(CLANG is None, reason="no clang++ for the host build")
@pytest.mark.parametrize("k", range(len(EXPRS)))
@pytest.mark.parametrize("order", [1, 2])
def test_generated_custom_kernel_on_host(tmp_path, k, order):
    from ctpwa_amd import _native
    plan = plan_of(_custom(EXPRS[k]))
    if order == 1:
        src = _native.generate_source(plan.to_text(), "fp64")
        kernel = "ctpwa_custom"
        ncf_m = re.search(r"#define NCF (\d+)", src)
        nrow = int(ncf_m.group(1))
    else:
        src = _native.generate_hessian_source(plan.to_text(), "fp64").split("// ---- module: ctpwa_custom2 ----")[1]
        kernel = "ctpwa_custom2"
    ncv = int(re.search(r"#define NCV (\d+)", src).group(1))
    nh = ncv * (ncv + 1) // 2
    if order == 2:
        nrow = 2 + 2 * ncv + 2 * nh
    pre, kern = _module(src, kernel)
    rows = [int(x) for x in re.findall(r"inv\[(\d+)LL \* ld", kern)]
    ninv = max(rows) + 1
    nf = len(re.findall(r"const (?:CD|CH) F = ", kern))
    outs = {}
    for init in ("pattern", "zero"):
        exe = _build(tmp_path, pre, kern, kernel, init)
        raw = subprocess.run([exe, str(ninv), str(nf), str(nrow), str(order - 1)], capture_output=True,
                             timeout=300, check=True).stdout
        outs[init] = raw
    assert outs["pattern"] == outs["zero"], "an output depends on an uninitialised variable"
    ld = 512
    a = np.frombuffer(outs["zero"], dtype=np.float64)
    size = nf * nrow * ld
    o0 = a[:size].reshape(nf, nrow, ld)
    off = size
    for j in range(ncv):
        h = a[off]
        op = a[off + 1:off + 1 + size].reshape(nf, nrow, ld)
        om = a[off + 1 + size:off + 1 + 2 * size].reshape(nf, nrow, ld)
        off += 1 + 2 * size
        fd = (op - om) / (2 * h)
        ok = np.isfinite(o0).all(axis=1)
        for part in (0, 1):                     # real, imaginary
            an = o0[:, 2 + part * ncv + j, :]
            num = fd[:, part, :]
            scale = np.abs(o0[:, 0, :]) + np.abs(o0[:, 1, :]) + np.abs(an) + 1e-300
            err = np.abs(num - an) / scale
            assert np.max(err[ok]) <= 1e-6, (k, order, j, part, np.max(err[ok]))
        if order == 2:
            # the packed second derivatives (l <= m) against differences of the first
            for m in range(ncv):
                lo, hi = min(j, m), max(j, m)
                hidx = lo * ncv - lo * (lo - 1) // 2 + (hi - lo)
                for part in (0, 1):
                    an = o0[:, 2 + 2 * ncv + part * nh + hidx, :]
                    num = fd[:, 2 + part * ncv + m, :]
                    scale = np.abs(o0[:, 2 + part * ncv + m, :]) + np.abs(an) + np.abs(o0[:, 0, :]) + 1e-300
                    err = np.abs(num - an) / scale
                    assert np.max(err[ok]) <= 1e-5, (k, order, j, m, part, np.max(err[ok]))
