"""``ctpwa.analysis`` drop-in: the reference's Python boundary over the native library.

Method names, argument meaning, layouts and error behaviour follow the reference's
pybind11 module (src/main.cu:90-131) and ``class analysis`` (src/Analysis.cu:1268-4832):

* ``getNLL(params)``: params is a 1-D float64 tensor [2n + P] on the analysis device,
  laid out [Re p_0..n-1, Im p_0..n-1, theta_0..P-1]; returns a 0-dim float64 tensor
  whose backward gives the analytic gradient (NLLFunction, src/Analysis.cu:625-1261).
  A wrong dtype or device raises RuntimeError (TORCH_CHECK, src/Analysis.cu:1296-1301).
* ``getHessian(params)``: [(2n+P), (2n+P)] float64 on the device, grouped layout.
* ``analysis(config, precision="mixed")`` (or ``precision: float`` in the config) stores
  the per-event amplitude tensors in FP32 and keeps every arithmetic step and sum in FP64.
* ``getNVector() / getNFreeTheta() / getNParams() / getParamNames() /
  getFreeResParams() / getFitMode() / setFitMode() / getNPolarizations() /
  getAmplitudeNames() / isValid() / reCalcAmp(theta)``.
* ``getPhspTensor() / getDataTensor() / getBkgTensor() / getBkgWeightsTensor() /
  getSLVectors() / getConstraintsIndex() / getConstraintsValues()``; ``getSLAmpsTensor``
  and ``saveSLAmps`` raise (the SL tensors are stored with barrier factors folded in).
* ``getFitFractions(vector, hessian)`` and ``writeResult(params, filename,
  is_saved_weight, waves)``: post-fit outputs from the amplitude Gram
  (ctpwa_amd/results.py); ``writeResult`` writes a NumPy ``.npz`` instead of a ROOT file.

All device work runs on the current torch stream; nothing synchronises the host
except reading results.
"""
from __future__ import annotations

import os
from typing import Dict, Optional

import numpy as np
import torch

from . import _native
from .config import ConfigParser
from .model import Plan, build_plan
from .phasespace import load_sample


def _read_weights(entry, base_dir: str, n: int) -> Optional[np.ndarray]:
    """``*_weights``: ``[x]`` a constant, ``[dat, file]`` one value per line
    (src/Analysis.cu:292-495)."""
    if entry is None:
        return None
    if len(entry) == 1:
        return np.full(n, float(entry[0]))
    path = entry[1]
    if not os.path.isabs(path):
        path = os.path.join(base_dir, path)
    if entry[0].lower() == "npy":
        w = np.load(path, allow_pickle=False).astype(np.float64).reshape(-1)
    else:
        with open(path) as f:
            w = np.array(f.read().split(), dtype=np.float64)
    if w.shape[0] < n:
        raise ValueError("%s holds %d weights for %d events" % (path, w.shape[0], n))
    return w[:n]


class NLLFunction:
    """The reference's NLLFunction (src/Analysis.cu:625-1261): ``apply(params, ana)`` gives
    the NLL as a 0-dim tensor whose backward is the analytic gradient.  The autograd
    function itself is C++ (``ctpwa_amd/csrc/torch_ops.cpp``): one fused NLL+gradient call
    in forward, gradient times grad_out in backward."""

    @staticmethod
    def apply(params, ana):
        return ana.getNLL(params)


class analysis:
    """One likelihood on one device.

    ``samples`` (an extension for callers that hold events in memory) maps ``phsp``,
    ``data``, ``bkg`` to [N, nfinal, 4] float64 arrays and ``*_weights`` to [N] arrays;
    when given, the ``Data`` section's files are not read.
    """

    def __init__(self, config_file: str = "config.yml", samples: Optional[Dict[str, np.ndarray]] = None,
                 device: Optional[int] = None, precision: Optional[str] = None):
        self._valid = False
        self._fit_mode = 1
        self.config_file = config_file
        self.cfg = ConfigParser(config_file)
        self.plan: Plan = build_plan(self.cfg, self._fit_mode)
        nf = len(self.plan.final_order)
        base = self.cfg.base_dir
        df = self.cfg.data_files
        s = dict(samples or {})
        for key in ("phsp", "data", "bkg"):
            if key not in s and key in df:
                s[key] = load_sample(df[key], base, nf)
        if s.get("phsp") is None:
            raise RuntimeError("config has no phase-space sample")
        if s.get("data") is None:
            raise RuntimeError("config has no data sample")
        n = {k: (0 if s.get(k) is None else s[k].shape[0]) for k in ("phsp", "data", "bkg")}
        w = {}
        for key in ("phsp", "data", "bkg"):
            wk = key + "_weights"
            if wk in s and s[wk] is not None:
                w[key] = np.broadcast_to(np.asarray(s[wk], dtype=np.float64), (n[key],))
            elif wk in df and n[key] > 0:
                w[key] = _read_weights(df[wk], base, n[key])
            else:
                w[key] = None
        # config `precision` (src/Analysis.cu:3925-3937): "double" and "auto" give the FP64
        # form, "float" the mixed form (FP32 amplitude storage, FP64 arithmetic and sums);
        # the `precision` argument ("fp64" | "mixed") overrides the config
        if self.cfg.precision not in ("auto", "double", "float"):
            raise RuntimeError("invalid precision in config (only float | double | auto)")
        if precision is None:
            precision = "mixed" if self.cfg.precision == "float" else "fp64"
        self.precision = precision
        if device is None:
            device = torch.cuda.current_device() if torch.cuda.is_available() else 0
        self.device = torch.device("cuda", device)
        self.n_events = n
        self._samples = s
        self._weights = w
        self._truth = None
        # theta of the stored amplitudes: a host copy (initial values, reCalcAmp) or, after
        # getNLL / getHessian / writeResult, the theta the device evaluated last (the
        # handle keeps it in its own slot, which intensities, amplitudes and Grams leave;
        # read back only on demand, so evaluations never wait for the host)
        self._theta_host = np.asarray([t.init for t in self.plan.theta], dtype=np.float64)
        self._theta_on_host = True
        self._h = _native.Handle(self.plan.to_text(), s["phsp"], s["data"], s.get("bkg"),
                                 w["phsp"], w["data"], w["bkg"], device=device, precision=precision)
        self._ops = _native.torch_ops()
        self._hptr = self._h.h.value
        self._dev_index = device
        self._n_params = self.plan.n_params()
        self._valid = True

    # ---- reference API -------------------------------------------------------
    def isValid(self) -> bool:
        return self._valid

    def getNLL(self, params: torch.Tensor) -> torch.Tensor:
        self._check_params(params)
        self._remember_theta(params)
        return self._ops.nll(params, self._hptr, _native.current_stream(self._dev_index), self._n_params)

    def getHessian(self, params: torch.Tensor) -> torch.Tensor:
        self._check_params(params)
        p = params.detach().contiguous()
        self._remember_theta(p)
        n = self.getNParams()
        H = torch.empty((n, n), dtype=torch.float64, device=self.device)
        self._h.hessian(p.data_ptr(), H.data_ptr(), _native.current_stream(self._dev_index))
        return H

    def getNVector(self) -> int:
        return self.plan.n_free

    def getNFreeTheta(self) -> int:
        return self.plan.n_theta

    def getNParams(self) -> int:
        return self.plan.n_params()

    def getParamNames(self):
        return list(self.plan.param_names)

    def getAmplitudeNames(self):
        return [self.plan.amp_names[a] for a in self.plan.ref_amplitudes()]

    def getNPolarizations(self) -> int:
        return self.plan.npol

    def getFreeResParams(self) -> torch.Tensor:
        """[3, P] float64 on the CPU: initial value, lower and upper bound."""
        init, lo, hi = self._h.param_info()
        return torch.tensor(np.stack([init, lo, hi]), dtype=torch.float64)

    def getFitMode(self) -> int:
        return self._fit_mode

    def setFitMode(self, mode: int) -> None:
        """0 (chain x step) or 1 (vspace).  The reference builds its parameter mapping in
        vspace at construction (src/Analysis.cu:4382-4478) and the mode read afterwards
        changes no count or layout of that mapping (src/Analysis.cu:1309-1342), so both
        values keep the vspace layout here."""
        if mode not in (0, 1):
            raise RuntimeError("fit mode must be 0 (chain x step) or 1 (vspace)")
        self._fit_mode = int(mode)

    def reCalcAmp(self, theta) -> None:
        """Recompute the stored amplitudes at ``theta`` (src/AmpGen.cu:2344-2726).  Every
        likelihood call evaluates the amplitudes from the theta part of its parameter
        vector; this sets the theta that the amplitude getters (getDataTensor, ...) and
        getFitFractions use until the next getNLL / getHessian."""
        t = torch.as_tensor(theta, dtype=torch.float64).detach().cpu().reshape(-1)
        if t.numel() != self.plan.n_theta:
            raise RuntimeError("theta must have %d entries" % self.plan.n_theta)
        self._theta_host = t.numpy().copy()
        self._theta_on_host = True

    def getSLVectors(self) -> torch.Tensor:
        """Number of SL amplitudes of every partial wave, int32 on the device
        (src/Analysis.cu:1346-1351, src/Info.cu:288)."""
        n = [0] * len(self.plan.waves)
        for w, blk in zip(self.plan.wave_of(), self.plan.blocks):
            if blk.sigma == 0:
                n[w] += blk.nsl
        return torch.tensor(n, dtype=torch.int32,
                            device=self.device)

    def getConstraintsIndex(self):
        """Amplitude index groups of the legacy constraint mode.  The reference always
        builds its coupling matrix when the model has amplitudes (src/Info.cu:322-336),
        which leaves these empty (src/Analysis.cu:4352-4378)."""
        return []

    def getConstraintsValues(self):
        return []

    def getPhspTensor(self) -> torch.Tensor:
        """Amplitudes of the phase-space sample at the current theta without couplings,
        flattened [event][pol][amp] complex128 on the device (src/Analysis.cu:3641-3676);
        identical-particle permutations are summed into their amplitude."""
        return self._stored_amplitudes(0)

    def getDataTensor(self) -> torch.Tensor:
        return self._stored_amplitudes(1)

    def getBkgTensor(self) -> torch.Tensor:
        if self.n_events["bkg"] == 0:
            return torch.empty(0, dtype=torch.complex128, device=self.device)
        return self._stored_amplitudes(2)

    def getBkgWeightsTensor(self) -> torch.Tensor:
        """Background weights as float32 on the device (empty without weights)."""
        w = self._weights["bkg"]
        if self.n_events["bkg"] == 0 or w is None:
            return torch.empty(0, dtype=torch.float32, device=self.device)
        return torch.tensor(np.asarray(w, dtype=np.float32), device=self.device)

    def getSLAmpsTensor(self):
        raise RuntimeError("getSLAmpsTensor is not built: the per-event SL tensors are stored with the "
                           "event parts of the barrier factors folded in, not as the bare angular tensors")

    def saveSLAmps(self, filename: str):
        raise RuntimeError("saveSLAmps is not built (see getSLAmpsTensor)")

    def getFitFractions(self, vector: torch.Tensor, hessian: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Fit fractions FF_b = integral |P_b|^2 / sum_j integral |P_j|^2 over the
        ``phsp_truth`` sample, with errors sqrt(diag(J C J^T)) (src/Analysis.cu:3527-3638).

        ``vector``: the free couplings as a complex128 tensor [n_free] on the device (the
        resonance parameters are those of the last getNLL/getHessian call, else their
        initial values), or the whole float64 parameter vector.  ``hessian``: optional
        full Hessian (its coupling block gives C); computed at the point when omitted.
        Returns [n_partials, 2] float64 on the CPU: fraction, error."""
        from . import results as R
        params = self._params_from_vector(vector)
        Q = self._truth_gram(params)
        pn = params.detach().cpu().numpy()
        ff = R.fit_fractions(Q, self.plan, pn)
        if hessian is None:
            hessian = self.getHessian(params)
        J = R.fit_fraction_jacobian(Q, self.plan, pn)
        err = R.fit_fraction_errors(J, hessian.detach().cpu().numpy(), self.plan.n_free)
        return torch.tensor(np.stack([ff, err], axis=1), dtype=torch.float64)

    def writeResult(self, params: torch.Tensor, filename: str, is_saved_weight: int = 0, waves=()) -> str:
        """Fit result over the phase-space sample as a NumPy .npz (the reference writes ROOT,
        src/Analysis.cu:1355-1700, src/ComputeResults.cu:49-185):

        * ``legends`` (one label per partial wave);
        * ``interference`` (symmetric: sum |P_b|^2 / total on the diagonal,
          sum 2 Re(conj(P_b) P_b') / total off it);
        * ``phsp_integral`` and ``data_integral``.

        ``waves``, a subset of partial-wave indices, replaces the per-event total by
        sum_pol |sum_{w in waves} P_w|^2 and the phase-space integral by its sum, as the
        reference does; the interference matrix keeps the full total.  With
        ``is_saved_weight=1`` the file also holds, normalised to the data yield by
        data_integral / phsp_integral:
        * the per-event ``totalweight`` and ``weight_<k>`` (k-th partial wave);
        * ``interf_<i>_<j>`` (i < j), sum_pol 2 Re(P_i conj P_j), so that without a
          subset totalweight = sum_k weight_k + sum_{i<j} interf_i_j;
        * the final-state momenta ``<particle>_{px,py,pz,E}``.
        Returns the file written."""
        from . import results as R
        self._check_params(params)
        nw = len(self.plan.waves)
        sel = [int(w) for w in waves]
        for w in sel:
            if not 0 <= w < nw:
                raise RuntimeError("waves index out of range: %d (partials = %d)" % (w, nw))
        p = params.detach().contiguous()
        # writeResult moves the stored amplitudes to its theta (host copy: the Grams and
        # per-event outputs below do not store a last-evaluated theta)
        self._theta_host = p[2 * self.plan.n_free:].cpu().numpy().astype(np.float64)
        self._theta_on_host = True
        Q = self._gram(self._h, p, 0)
        X, total = R.interference_matrix(Q, self.plan, p.cpu().numpy())
        nd = float(np.sum(self._weights["data"])) if self._weights["data"] is not None else float(self.n_events["data"])
        di = R.data_integral(nd, self._weights["bkg"], self.n_events["bkg"])
        integral = total
        P = None
        if sel or is_saved_weight == 1:
            P = self.getPartialAmplitudes(p, 0)                       # [nw, N, nPol]
            S = P[sel].sum(dim=0) if sel else P.sum(dim=0)
            ev_total = (S.real ** 2 + S.imag ** 2).sum(dim=1)
            if sel:
                integral = float(ev_total.sum().item())
        out = {"legends": np.array(R.partial_names(self.plan)), "interference": X,
               "phsp_integral": np.float64(integral), "data_integral": np.float64(di)}
        if is_saved_weight == 1:
            norm = di / (integral if integral > 0 else 1e-30)
            out["totalweight"] = (ev_total * norm).cpu().numpy()
            for k in range(nw):
                out["weight_%d" % k] = ((P[k].real ** 2 + P[k].imag ** 2).sum(dim=1) * norm).cpu().numpy()
            for a in range(nw):
                for b in range(a + 1, nw):
                    t = 2.0 * (P[a].real * P[b].real + P[a].imag * P[b].imag).sum(dim=1) * norm
                    out["interf_%d_%d" % (a, b)] = t.cpu().numpy()
            mom = self._samples["phsp"]
            for i, name in enumerate(self.plan.final_order):       # events hold (E, px, py, pz)
                for c, comp in ((1, "px"), (2, "py"), (3, "pz"), (0, "E")):
                    out["%s_%s" % (name, comp)] = np.ascontiguousarray(mom[:, i, c])
        path = filename if filename.endswith(".npz") else os.path.splitext(filename)[0] + ".npz"
        np.savez(path, **out)
        return path

    # ---- extensions ----------------------------------------------------------
    def getPartialIntensities(self, params: torch.Tensor, cls: int = 0) -> torch.Tensor:
        """[n_partials + 1, N] float64 on the device: per event sum_pol |P_b|^2 for every
        partial wave b, then the total intensity."""
        self._check_params(params)
        n = self.n_events[("phsp", "data", "bkg")[cls]]
        out = torch.empty((len(self.plan.waves) + 1, n), dtype=torch.float64, device=self.device)
        if n:
            self._h.partial_intensities(params.detach().contiguous().data_ptr(), cls, out.data_ptr(),
                                        _native.current_stream(self._dev_index))
        return out

    def getPartialAmplitudes(self, params: torch.Tensor, cls: int = 0) -> torch.Tensor:
        """[n_partials, N, nPol] complex128 on the device: per event and polarisation the
        coherent amplitude P_b of every partial wave (couplings applied)."""
        self._check_params(params)
        n = self.n_events[("phsp", "data", "bkg")[cls]]
        out = torch.empty((len(self.plan.waves), n, self.plan.npol, 2), dtype=torch.float64, device=self.device)
        if n:
            self._h.partial_amplitudes(params.detach().contiguous().data_ptr(), cls, out.data_ptr(),
                                       _native.current_stream(self._dev_index))
        return torch.view_as_complex(out)

    def getAmplitudeGram(self, params: torch.Tensor, cls: int = 0) -> np.ndarray:
        """complex [n_amp, n_amp] sum over a sample of conj(A_a) A_a' (no couplings)."""
        self._check_params(params)
        return self._gram(self._h, params.detach().contiguous(), cls)

    def getIntensity(self, params: torch.Tensor, cls: int = 1) -> torch.Tensor:
        """Per-event intensity of a sample (0 phsp, 1 data, 2 bkg)."""
        self._check_params(params)
        n = self.n_events[("phsp", "data", "bkg")[cls]]
        out = torch.empty(n, dtype=torch.float64, device=self.device)
        if n:
            self._h.intensity(params.detach().contiguous().data_ptr(), cls, out.data_ptr(),
                              _native.current_stream(self._dev_index))
        return out

    def getAmplitudes(self, params: torch.Tensor, cls: int = 1) -> torch.Tensor:
        """Per-event amplitudes [N, nPol, nAmp] complex128 (couplings not applied)."""
        self._check_params(params)
        n = self.n_events[("phsp", "data", "bkg")[cls]]
        out = torch.empty((n, self.plan.npol, self.plan.n_amp, 2), dtype=torch.float64, device=self.device)
        if n:
            self._h.amplitudes(params.detach().contiguous().data_ptr(), cls, out.data_ptr(),
                               _native.current_stream(self._dev_index))
        return torch.view_as_complex(out)

    def kernelTimeMs(self, params: torch.Tensor, iters: int = 20) -> float:
        self._check_params(params)
        return self._h.time_eval_kernel(params.detach().contiguous().data_ptr(), iters,
                                        _native.current_stream(self._dev_index))

    def layout(self):
        return self._h.layout()

    # ---- internals -----------------------------------------------------------
    def _check_params(self, params: torch.Tensor):
        if not isinstance(params, torch.Tensor):
            raise RuntimeError("params must be a torch.Tensor")
        if params.dtype != torch.float64:
            raise RuntimeError("params must be float64")
        if not params.is_cuda or params.device != self.device:
            raise RuntimeError("params must be on %s" % self.device)
        if params.dim() != 1 or params.numel() != self._n_params:
            raise RuntimeError("params must be 1-D with %d entries" % self._n_params)

    def _stored_amplitudes(self, cls: int) -> torch.Tensor:
        plan = self.plan
        p = torch.cat([torch.zeros(2 * plan.n_free, dtype=torch.float64),
                       torch.tensor(self._theta, dtype=torch.float64)]).to(self.device)
        A = self.getAmplitudes(p, cls)
        ref = plan.ref_amplitudes()
        if len(ref) != plan.n_amp:
            out = A[:, :, ref].clone()
            for blk in plan.blocks:
                if blk.sigma == 0:
                    continue
                orig = plan.blocks[blk.src]
                for c in range(blk.nsl):
                    a, a0 = blk.site + c, orig.site + c
                    out[:, :, ref.index(a0)] += (plan.amp_ratio[a] / plan.amp_ratio[a0]) * A[:, :, a]
            A = out
        return A.reshape(-1)

    def _remember_theta(self, p: torch.Tensor):
        # the native call that follows stores p's theta as the last evaluated one
        self._theta_on_host = False

    @property
    def _theta(self) -> np.ndarray:
        if not self._theta_on_host:
            self._theta_host = self._h.last_theta(_native.current_stream(self._dev_index))
            self._theta_on_host = True
        return self._theta_host

    def _params_from_vector(self, vector: torch.Tensor) -> torch.Tensor:
        if not isinstance(vector, torch.Tensor) or not vector.is_cuda or vector.device != self.device:
            raise RuntimeError("vector must be a tensor on %s" % self.device)
        if vector.dtype == torch.float64 and vector.dim() == 1 and vector.numel() == self.getNParams():
            return vector.detach().contiguous()
        if vector.dtype != torch.complex128 or vector.numel() != self.plan.n_free:
            raise RuntimeError("vector must be complex128 with %d entries (or the float64 parameters)"
                               % self.plan.n_free)
        v = vector.detach().reshape(-1)
        th = torch.tensor(self._theta, dtype=torch.float64, device=self.device)
        return torch.cat([v.real.to(torch.float64), v.imag.to(torch.float64), th]).contiguous()

    def _gram(self, handle, p: torch.Tensor, cls: int) -> np.ndarray:
        na = self.plan.n_amp
        out = torch.empty((na, na, 2), dtype=torch.float64, device=self.device)
        handle.amplitude_gram(p.data_ptr(), cls, out.data_ptr(), _native.current_stream(self._dev_index))
        o = out.cpu().numpy()
        return o[..., 0] + 1j * o[..., 1]

    def _truth_gram(self, p: torch.Tensor) -> np.ndarray:
        """Amplitude Gram over the phsp_truth sample (config ``Data: phsp_truth`` or
        ``samples["phsp_truth"]``); its handle is built on first use."""
        if self._truth is None:
            truth = self._samples.get("phsp_truth")
            df = self.cfg.data_files
            if truth is None and "phsp_truth" in df:
                truth = load_sample(df["phsp_truth"], self.cfg.base_dir, len(self.plan.final_order))
            if truth is None:
                raise RuntimeError("No phsp_truth in config")
            self._truth = _native.Handle(self.plan.to_text(), truth, truth[:1], None, device=self.device.index,
                                         precision=self.precision)
        return self._gram(self._truth, p, 0)

