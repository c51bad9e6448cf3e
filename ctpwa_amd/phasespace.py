"""Event samples: phase-space generation and the reference's text event format.

* ``genbod``: uniform n-body phase space by the method of F. James, "Monte Carlo
  phase space", CERN 68-15 (1968) -- ordered uniform invariant masses, weight equal to
  the product of two-body breakup momenta, unweighted by accept-reject against the
  analytic weight bound.  Written from the published method.
* ``read_dat`` / ``write_dat``: one line ``E px py pz`` per particle, particles of an
  event in ``Data.order`` (src/Analysis.cu:58-290); the initial state is the sum of the
  final state in that order (src/Analysis.cu:148-160).
"""
from __future__ import annotations

import numpy as np


def _pdk(a, b, c):
    """Two-body breakup momentum of a -> b c (0 below threshold)."""
    x = (a - b - c) * (a + b + c) * (a - b + c) * (a + b - c)
    return np.sqrt(np.maximum(x, 0.0)) / (2.0 * a)


def _boost(p, beta):
    """Boost four-vectors p[..., 4] by velocity beta[..., 3] (active boost)."""
    b2 = np.sum(beta * beta, axis=-1, keepdims=True)
    gamma = 1.0 / np.sqrt(1.0 - b2)
    bp = np.sum(beta * p[..., 1:], axis=-1, keepdims=True)
    coef = np.where(b2 > 0, (gamma - 1.0) / np.where(b2 > 0, b2, 1.0), 0.0)
    E = p[..., :1]
    out = np.empty_like(p)
    out[..., 1:] = p[..., 1:] + coef * bp * beta + gamma * beta * E
    out[..., :1] = gamma * (E + bp)
    return out


def _weights_and_masses(rng, n, M, masses):
    k = len(masses)
    T = M - np.sum(masses)
    if T <= 0:
        raise ValueError("decay is kinematically forbidden")
    r = np.sort(rng.random((n, k - 2)), axis=1) if k > 2 else np.zeros((n, 0))
    R = np.concatenate([np.zeros((n, 1)), r, np.ones((n, 1))], axis=1)
    cum = np.cumsum(masses)
    Mk = cum[None, :] + R * T               # Mk[:, i] = invariant mass of particles 0..i
    pk = np.empty((n, k - 1))
    for i in range(1, k):
        pk[:, i - 1] = _pdk(Mk[:, i], Mk[:, i - 1], masses[i])
    w = np.prod(pk, axis=1)
    return w, Mk, pk


def _wtmax(M, masses):
    T = M - np.sum(masses)
    emmax = T + masses[0]
    emmin = 0.0
    wt = 1.0
    for i in range(1, len(masses)):
        emmin += masses[i - 1]
        emmax += masses[i]
        wt *= float(_pdk(np.float64(emmax), np.float64(emmin), np.float64(masses[i])))
    return wt


def _build_momenta(rng, Mk, pk, masses):
    n, k = Mk.shape
    p = np.zeros((n, k, 4))

    def rand_dir():
        cth = 2.0 * rng.random(n) - 1.0
        sth = np.sqrt(np.maximum(0.0, 1.0 - cth * cth))
        phi = 2.0 * np.pi * rng.random(n)
        return np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], axis=1)

    # particles 0 and 1 back to back in the rest frame of (0, 1)
    d = rand_dir()
    q = pk[:, 0:1]
    p[:, 1, 1:] = q * d
    p[:, 1, 0] = np.sqrt(q[:, 0] ** 2 + masses[1] ** 2)
    p[:, 0, 1:] = -q * d
    p[:, 0, 0] = np.sqrt(q[:, 0] ** 2 + masses[0] ** 2)
    for i in range(2, k):
        d = rand_dir()
        q = pk[:, i - 1:i]
        # particle i in the rest frame of (0..i); subsystem (0..i-1) recoils
        sub_E = np.sqrt(q[:, 0] ** 2 + Mk[:, i - 1] ** 2)
        beta = (-q * d) / sub_E[:, None]
        p[:, :i, :] = _boost(p[:, :i, :], beta[:, None, :])
        p[:, i, 1:] = q * d
        p[:, i, 0] = np.sqrt(q[:, 0] ** 2 + masses[i] ** 2)
    return p


def genbod(n_events: int, M: float, masses, seed: int = 0, batch: int = 1 << 20) -> np.ndarray:
    """Unweighted uniform phase-space events [n_events, k, 4] (E, px, py, pz), float64,
    in the rest frame of the decaying system of mass ``M``."""
    masses = np.asarray(masses, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    wmax = _wtmax(M, masses)
    out = []
    have = 0
    while have < n_events:
        m = max(1024, min(batch, int((n_events - have) * 2.5) + 1024))
        w, Mk, pk = _weights_and_masses(rng, m, M, masses)
        u = rng.random(m)
        keep = w > u * wmax
        mom = _build_momenta(rng, Mk[keep], pk[keep], masses)
        out.append(mom)
        have += mom.shape[0]
    return np.concatenate(out, axis=0)[:n_events]


def accept_reject(candidates: np.ndarray, intensity: np.ndarray, n_keep: int, seed: int,
                  i_max: float = None) -> np.ndarray:
    """Unweight ``candidates`` by their model intensity (keep with probability I/I_max)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if i_max is None:
        i_max = float(np.max(intensity)) * 1.05
    u = rng.random(intensity.shape[0]) * i_max
    sel = candidates[intensity > u]
    return sel[:n_keep]


def write_dat(path: str, events: np.ndarray) -> None:
    """Write [N, k, 4] momenta as text, one particle per line."""
    flat = np.ascontiguousarray(events, dtype=np.float64).reshape(-1, 4)
    np.savetxt(path, flat, fmt="%.17g")


def read_dat(path: str, n_particles: int) -> np.ndarray:
    """Read text momenta into [N, k, 4]; blank lines are ignored."""
    with open(path, "r") as f:
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size % (4 * n_particles) != 0:
        raise ValueError("%s: %d values do not make whole events of %d particles"
                         % (path, vals.size, n_particles))
    return vals.reshape(-1, n_particles, 4)


def load_sample(entry, base_dir: str, n_particles: int) -> np.ndarray:
    """Load a ``Data`` entry: ``[dat, file]`` or ``[npy, file]`` (npy holds [N, k, 4])."""
    import os
    kind = entry[0].lower()
    path = entry[1]
    if not os.path.isabs(path):
        path = os.path.join(base_dir, path)
    if kind == "dat":
        return read_dat(path, n_particles)
    if kind == "npy":
        arr = np.load(path, allow_pickle=False)
        if arr.ndim != 3 or arr.shape[1] != n_particles or arr.shape[2] != 4:
            raise ValueError("%s: expected [N, %d, 4]" % (path, n_particles))
        return np.ascontiguousarray(arr, dtype=np.float64)
    raise ValueError("unsupported sample format '%s' (dat and npy are supported)" % entry[0])
