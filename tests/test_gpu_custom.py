"""Custom lineshapes (hipRTC-compiled expressions with forward-mode derivatives) on the
GPU against the oracle's interpreter: NLL, gradient, Hessian (second-order factors),
amplitudes and intensities, in both precision forms, and the post-fit integrals; the
Custom spelling of the built-in BWR gives the built-in's NLL."""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from test_custom_cpu import _custom
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

EXPRS = [
    "BWR(p0, p1, 1, 3.0) * Bf(1, q, q0(p0), 3.0)",
    "p0^2/(p0^2 - m^2 - i*p0*p1)",
    ("exp(-p1*m) * (cos(p0*m) + i*sin(p1)) / sqrt(p0^2 + m^2) * abs(p0 - i*p1) "
     "+ log(p0) * conj(m + i*p1)^2 - Im(p0 * (q + 1j)) + Re(csqrt(p1 - q0))"),
    # Flatte() expanded in the expression and the daughters' momentum variables
    "Flatte(p0, p1, 0.13957, 0.13957) * (1 + 0.3*p1costheta - 0.2*p2p + 0.1*sin(p1phi - p2phi) + 0.05*p1e)",
]


def _run(h, p):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[0], o[1:]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", range(len(EXPRS)))
def test_custom_parity(k, precision, oracle_lib):
    from ctpwa_amd import _native
    plan = plan_of(_custom(EXPRS[k]))
    ph, da, bk = samples(20000, 4000, 1000, 61 + k)
    wb = np.full(bk.shape[0], 0.8)
    print("case custom", k, precision, flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision=precision)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    p = perturbed(W.nominal_params(plan, 1), plan, 30 + k)
    nll_o, g_o = O.nll_grad(p)
    nll_g, g_g = _run(h, p)
    # mixed: the FP32-stored SL tensors carry 2^-24 relative rounding per entry; over
    # 4k data events the NLL moves by a few 1e-9 relative (the |dNLL| <= 1e-4 gate of the
    # mixed form at the full size is far looser)
    tol_n, tol_g, tol_a = (1e-9, 1e-7, 1e-10) if precision == "fp64" else (1e-8, 1e-5, 1e-6)
    print(nll_o, nll_g, rel(g_g, g_o), flush=True)
    assert abs(nll_g - nll_o) <= tol_n * abs(nll_o)
    assert rel(g_g, g_o) <= tol_g
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    n = h.n_events[1]
    A = torch.empty((n, plan.npol, plan.n_amp, 2), dtype=torch.float64, device="cuda")
    I = torch.empty(n, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    h.amplitudes(t.data_ptr(), 1, A.data_ptr(), s)
    h.intensity(t.data_ptr(), 1, I.data_ptr(), s)
    torch.cuda.synchronize()
    A = A.cpu().numpy()
    A = A[..., 0] + 1j * A[..., 1]
    Ao = O.amplitudes(p, 1)
    assert np.max(np.abs(A - Ao)) <= tol_a * np.max(np.abs(Ao))
    assert rel(I.cpu().numpy(), O.intensity(p, 1)) <= 10 * tol_a
    # the Hessian: second-order Custom factors from their own generated kernel
    H = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
    h.hessian(t.data_ptr(), H.data_ptr(), s)
    torch.cuda.synchronize()
    _, _, H_o = O.hessian(p)
    print("hessian", rel(H.cpu().numpy(), H_o), flush=True)
    # mixed: the FP32-stored SL tensors' rounding (2^-24 per entry) reaches the Hessian
    # through sums with large cancelling terms; the expression with every function sits
    # near 2e-5
    assert rel(H.cpu().numpy(), H_o) <= (1e-6 if precision == "fp64" else 5e-5)


def test_custom_post_fit_outputs(oracle_lib):
    """Amplitude Gram and per-event partial-wave amplitudes of a model with Custom
    lineshapes against the oracle's amplitudes."""
    from ctpwa_amd import _native
    from ctpwa_amd import results as R
    plan = plan_of(_custom(EXPRS[2]))
    ph, da, _ = samples(8000, 2000, 0, 77)
    print("case custom post-fit", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    p = perturbed(W.nominal_params(plan, 1), plan, 8)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    Q = torch.empty((plan.n_amp, plan.n_amp, 2), dtype=torch.float64, device="cuda")
    h.amplitude_gram(t.data_ptr(), 0, Q.data_ptr(), s)
    nw = len(plan.waves)
    Pw = torch.empty((nw, ph.shape[0], plan.npol, 2), dtype=torch.float64, device="cuda")
    h.partial_amplitudes(t.data_ptr(), 0, Pw.data_ptr(), s)
    torch.cuda.synchronize()
    A = O.amplitudes(p, 0)
    Qo = np.einsum("epa,epb->ab", np.conj(A), A)
    Qg = Q.cpu().numpy()
    assert np.abs(Qg[..., 0] + 1j * Qg[..., 1] - Qo).max() <= 1e-10 * np.abs(Qo).max()
    v = R.coupling_vector(plan, p)
    Po = np.stack([A[:, :, B.site:B.site + B.nsl] @ v[B.site:B.site + B.nsl] for B in plan.blocks])
    Pg = Pw.cpu().numpy()
    assert np.abs(Pg[..., 0] + 1j * Pg[..., 1] - Po).max() <= 1e-10 * np.abs(Po).max()


def test_custom_bwr_matches_builtin_on_gpu():
    from ctpwa_amd import _native
    ph, da, bk = samples(20000, 4000, 1000, 71)
    wb = np.full(bk.shape[0], 0.8)
    pc, pb = plan_of(_custom(EXPRS[0])), plan_of(W.JPSI_3PI)
    hc = _native.Handle(pc.to_text(), ph, da, bk, None, None, wb)
    hb = _native.Handle(pb.to_text(), ph, da, bk, None, None, wb)
    p = perturbed(W.nominal_params(pb, 1), pb, 33)
    nc, gc = _run(hc, p)
    nb, gb = _run(hb, p)
    assert abs(nc - nb) <= 1e-11 * abs(nb)
    assert rel(gc, gb) <= 1e-9


def test_custom_every_handle_and_batch(oracle_lib):
    """The most demanding expression on several handles of one process (created before and
    after other device allocations), with repeated evaluations and through the batched
    path: every gradient equals the oracle's and repeated evaluations are bit-identical."""
    from ctpwa_amd import _native
    plan = plan_of(_custom(EXPRS[2]))
    ph, da, bk = samples(20000, 4000, 1000, 75)
    wb = np.full(bk.shape[0], 0.8)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    p = perturbed(W.nominal_params(plan, 1), plan, 36)
    nll_o, g_o = O.nll_grad(p)
    print("case custom handles", flush=True)
    hs = []
    for k in range(3):
        pad = torch.empty(int(1e6) * (k + 1), dtype=torch.float64, device="cuda")   # shifts later allocations
        hs.append(_native.Handle(plan.to_text(), ph, da, bk, None, None, wb))
        first = None
        for rep in range(3):
            nll_g, g_g = _run(hs[-1], p)
            assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
            assert rel(g_g, g_o) <= 1e-7, (k, rep)
            if first is None:
                first = g_g
            assert np.array_equal(g_g, first)
        del pad
    b = _native.Batch(4)
    P = torch.tensor(np.stack([p] * 3), dtype=torch.float64, device="cuda")
    Ob = torch.zeros((3, 1 + hs[0].n_params), dtype=torch.float64, device="cuda")
    b.nll_grad(hs, [P[j].data_ptr() for j in range(3)], [Ob[j].data_ptr() for j in range(3)],
               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    G = Ob.cpu().numpy()
    for j in range(3):
        assert rel(G[j, 1:], g_o) <= 1e-7
    b.close()


@pytest.mark.parametrize("which", ["nnbar", "b1_identical"])
def test_custom_per_wave_L(oracle_lib, which):
    """A Custom expression taking each wave's L (a block split into runs of one L; with
    identical particles also their permuted copies) against the oracle of the built-in
    model it spells: NLL 1e-9, gradient 1e-7, Hessian 1e-6, and per-event partial-wave
    intensities equal to the built-in plan's on the device."""
    from test_custom_cpu import PER_WAVE_L, PER_WAVE_MASSES, X_BWR, X_CUSTOM, _b1_case
    from test_identical_cpu import _sample
    from ctpwa_amd import _native
    if which == "nnbar":
        tb, tc, masses = PER_WAVE_L, PER_WAVE_L.replace(X_BWR, X_CUSTOM), PER_WAVE_MASSES
    else:
        tb, tc, masses = _b1_case()
    pb, pc = plan_of(tb), plan_of(tc)
    ph, da = _sample(20000, 81, masses), _sample(4000, 82, masses)
    O = oracle_lib.Oracle(pb.to_text(), 3, ph, da)
    print("case custom per-wave L", which, flush=True)
    hc = _native.Handle(pc.to_text(), ph, da, None)
    hb = _native.Handle(pb.to_text(), ph, da, None)
    s = torch.cuda.current_stream().cuda_stream
    for k, p in enumerate([W.nominal_params(pb, 1), perturbed(W.nominal_params(pb, 1), pb, 7, 0.1)]):
        nll_o, g_o, H_o = O.hessian(p)
        t = torch.tensor(p, dtype=torch.float64, device="cuda")
        out = torch.empty(1 + hc.n_params, dtype=torch.float64, device="cuda")
        Ht = torch.empty((hc.n_params, hc.n_params), dtype=torch.float64, device="cuda")
        hc.nll_grad(t.data_ptr(), out.data_ptr(), s)
        hc.hessian(t.data_ptr(), Ht.data_ptr(), s)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        print("per-wave L", k, nll_o, o[0], rel(o[1:], g_o), rel(Ht.cpu().numpy(), H_o), flush=True)
        assert abs(o[0] - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(o[1:], g_o) <= 1e-7
        assert rel(Ht.cpu().numpy(), H_o) <= 1e-6
        for cls in (0, 1):
            n = hc.n_events[cls]
            pw = [torch.empty((len(pb.waves) + 1, n), dtype=torch.float64, device="cuda") for _ in range(2)]
            hc.partial_intensities(t.data_ptr(), cls, pw[0].data_ptr(), s)
            hb.partial_intensities(t.data_ptr(), cls, pw[1].data_ptr(), s)
            torch.cuda.synchronize()
            assert rel(pw[0].cpu().numpy(), pw[1].cpu().numpy()) <= 1e-10
    hc.close()
    hb.close()
