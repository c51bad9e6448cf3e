"""Post-fit quantities from the amplitude Gram: partial-wave integrals, fit fractions with
errors, the interference matrix and per-event weights.

Every quantity the reference computes after a fit (``getFitFractions``,
src/Analysis.cu:3527-3638; ``writeResult``, src/Analysis.cu:1355-1700 with
src/ComputeResults.cu:22-190) is a function of the couplings and of one theta-dependent
object, the complex Gram of the amplitudes over a sample,

    Q[a][a'] = sum_events sum_pol conj(A_a) A_a'        (ctpwa_amplitude_gram)

with a partial wave b (one chain with one resonance set: the reference's nSLvectors_
groups, src/Info.cu:240-300, together with its identical-particle permuted copies) the
coherent sum P_b = sum_{a in b} v_a A_a.  Then

    M[b][b'] = sum conj(P_b) P_b' = sum_{a in b, a' in b'} conj(v_a) v_a' Q[a][a']

gives the partial integrals M[b][b], the total sum_{b,b'} M[b][b'], fit fractions
FF_b = M[b][b] / sum_j M[j][j] (the reference's definition: partial integrals over their
sum, so sum_b FF_b = 1) and the interference matrix (diagonal M[b][b] / total,
off-diagonal 2 Re M[b][b'] / total).  The couplings enter M quadratically, so the
Jacobian of the fit fractions in the fit parameters is exact here, where the reference
differentiates numerically (relative step 5e-6); the errors are
sqrt(diag(J C J^T)) with C the inverse of the coupling block of the Hessian without the
fixed first coupling, used only when that block is positive definite (as in the
reference).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def coupling_vector(plan, params: np.ndarray) -> np.ndarray:
    """Amplitude couplings v_a = r_a p_f(a) from the grouped fit parameters."""
    n = plan.n_free
    p = np.asarray(params[:n], dtype=np.float64) + 1j * np.asarray(params[n:2 * n], dtype=np.float64)
    return np.array([plan.amp_ratio[a] * p[plan.amp_free[a]] for a in range(plan.n_amp)], dtype=np.complex128)


def partial_weights_matrix(plan, v: np.ndarray) -> np.ndarray:
    """W[w][a] = v_a for the amplitudes a of partial wave w (a block and its
    identical-particle permuted copies), else 0."""
    wave = plan.wave_of()
    W = np.zeros((len(plan.waves), plan.n_amp), dtype=np.complex128)
    for b, B in enumerate(plan.blocks):
        W[wave[b], B.site:B.site + B.nsl] = v[B.site:B.site + B.nsl]
    return W


def partial_matrix(Q: np.ndarray, W: np.ndarray) -> np.ndarray:
    """M[b][b'] = sum_{a, a'} conj(W[b][a]) Q[a][a'] W[b'][a']."""
    return np.conj(W) @ Q @ W.T


def fit_fractions(Q: np.ndarray, plan, params: np.ndarray) -> np.ndarray:
    M = partial_matrix(Q, partial_weights_matrix(plan, coupling_vector(plan, params)))
    d = np.real(np.diag(M))
    s = d.sum()
    return d / (s if s > 0 else 1e-30)


def fit_fraction_jacobian(Q: np.ndarray, plan, params: np.ndarray) -> np.ndarray:
    """dFF_b / dx for the 2 n_free coupling parameters x = [Re p, Im p] (grouped)."""
    n = plan.n_free
    v = coupling_vector(plan, params)
    W = partial_weights_matrix(plan, v)
    QW = Q @ W.T                                 # [a][b] = (Q W_b)_a
    d = np.real(np.einsum("ba,ab->b", np.conj(W), QW))
    S = d.sum() if d.sum() > 0 else 1e-30
    wave = plan.wave_of()
    D = np.zeros((len(plan.waves), 2 * n))       # dM[w][w] / dx
    for b, B in enumerate(plan.blocks):
        w = wave[b]
        for a in range(B.site, B.site + B.nsl):
            f, r = plan.amp_free[a], plan.amp_ratio[a]
            D[w, f] += 2.0 * r * np.real(QW[a, w])          # dv_a / dRe p_f = r_a
            D[w, n + f] += 2.0 * r * np.imag(QW[a, w])      # dv_a / dIm p_f = i r_a
    return (D * S - np.outer(d, D.sum(axis=0))) / (S * S)


def fit_fraction_errors(J: np.ndarray, hessian: np.ndarray, n_free: int) -> np.ndarray:
    """sqrt(diag(J C J^T)), C = inverse of the coupling block of the Hessian without the
    fixed first coupling (indices 0 and n_free); zeros unless that block is positive
    definite (smallest eigenvalue above 1e-8), as the reference does."""
    n2 = 2 * n_free
    free = [j for j in range(n2) if j != 0 and j != n_free]
    Hc = np.asarray(hessian, dtype=np.float64)[:n2, :n2][np.ix_(free, free)]
    if len(free) == 0 or np.linalg.eigvalsh(Hc)[0] <= 1e-8:
        return np.zeros(J.shape[0])
    C = np.linalg.inv(Hc)
    Jf = J[:, free]
    return np.sqrt(np.maximum(np.einsum("bi,ij,bj->b", Jf, C, Jf), 0.0))


def interference_matrix(Q: np.ndarray, plan, params: np.ndarray) -> Tuple[np.ndarray, float]:
    """writeResult's interference matrix (symmetric) and the total integral:
    diagonal sum |P_b|^2 / total, off-diagonal sum 2 Re(conj(P_b) P_b') / total."""
    M = partial_matrix(Q, partial_weights_matrix(plan, coupling_vector(plan, params)))
    total = float(np.real(M.sum()))
    t = total if total > 0 else 1e-30
    X = (np.real(M) + np.real(M).T) / t          # 2 Re M[b][b'] (M is Hermitian), exactly symmetric
    np.fill_diagonal(X, np.real(np.diag(M)) / t)
    return X, total


def partial_names(plan) -> List[str]:
    """One label per partial wave: its first amplitude's name without the SL labels, or
    the chain's resonances for a partial wave without allowed waves."""
    import re
    out = []
    for b in plan.waves:
        B = plan.blocks[b]
        if B.nsl > 0:
            out.append(re.sub(r"_SL\(\d+,\d+\)", "", plan.amp_names[B.site]))
        else:
            out.append("topology%d[%s]" % (B.topo, ",".join(r.name for r in B.res)))
    return out


def data_integral(n_data_weight: float, bkg_weights: Sequence[float], n_bkg: int) -> float:
    """Signal yield behind the weights: data weight minus background (the background
    weights' sum, or the background count when they do not add up to a positive value),
    src/Analysis.cu:1559-1584."""
    if n_bkg <= 0:
        return float(n_data_weight)
    sb = float(np.sum(bkg_weights)) if bkg_weights is not None else 0.0
    return float(n_data_weight) - (sb if sb > 0 else float(n_bkg))
