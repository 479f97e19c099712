"""Seed-controlled synthetic inputs for the bubble finder.

The reference reads its inputs from upstream stages (``delta/deltanl_*`` from
get_nldensity.c:169-175, ``SFR/sfrd_*`` from get_SFR.c:134-138).  For
benchmarks and tests the boxes are synthesised instead, with this recipe:

* ``g``: Gaussian random field with P(k) = A k^slope (k = 0 mode zeroed), A
  fixed by rescaling so the per-cell standard deviation is ``sigma``;
* ``1 + delta = exp(g) / <exp(g)>`` (so delta > -1 and <delta> = 0);
* emissivity: ``n_src`` point sources in cells drawn with probability
  proportional to (1+delta)^2, each of lognormal strength (mu=0, sigma=1);
  several sources may share a cell (their strengths add);
* recombination box ``rec = (1+delta)^2`` (direct mode).

Boxes are float32 in the reference layout i*N^2 + j*N + p.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft


def gaussian_field(n: int, slope: float = -2.5, sigma: float = 1.2, seed: int = 0,
                   dtype=np.float32) -> np.ndarray:
    """Real n^3 Gaussian field with power-law spectrum |k|^slope, std = sigma."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n), dtype=np.float32 if dtype == np.float32 else np.float64)
    spec = sfft.rfftn(white, workers=-1)
    kx = np.fft.fftfreq(n) * n
    kz = np.arange(n // 2 + 1, dtype=np.float64)
    k2 = (kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (slope / 4.0), 0.0)
    spec *= amp.astype(spec.real.dtype)
    del amp, k2
    g = sfft.irfftn(spec, s=(n, n, n), workers=-1)
    del spec
    g = g.astype(np.float64)
    g -= g.mean()
    g *= sigma / g.std()
    return g.astype(dtype)


def lognormal_delta(g: np.ndarray) -> np.ndarray:
    """delta = exp(g)/<exp(g)> - 1 as float32 (FP64 normalisation)."""
    e = np.exp(g.astype(np.float64))
    e /= e.mean()
    return (e - 1.0).astype(np.float32)


def point_sources(delta: np.ndarray, n_src: int, seed: int = 0) -> np.ndarray:
    """Unit-efficiency emissivity box: n_src sources, P(cell) ~ (1+delta)^2,
    strengths lognormal(0, 1)."""
    rng = np.random.default_rng(seed + 7919)
    w = (1.0 + delta.astype(np.float64).ravel()) ** 2
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cells = np.searchsorted(cdf, rng.random(n_src), side="right")
    cells = np.minimum(cells, delta.size - 1)
    strength = rng.lognormal(0.0, 1.0, n_src)
    emis = np.zeros(delta.size, np.float64)
    np.add.at(emis, cells, strength)
    return emis.reshape(delta.shape).astype(np.float32)


def geometric_ladder(rmax: float, factor: float, dx: float) -> list[float]:
    """Radii rmax, rmax*factor, ... while R >= dx (repeated FP64 multiply)."""
    radii = []
    r = float(rmax)
    while r >= dx:
        radii.append(r)
        r *= factor
    return radii


def direct_boxes(n: int, side: float, n_src: int, seed: int, slope: float = -2.5,
                 sigma: float = 1.2, growth: float | None = None):
    """Direct-mode inputs of the configs: returns (delta, emis_unit, rec).

    ``growth``: optional linear-growth rescaling of g (slice recipe g_s =
    growth * g) before the lognormal map."""
    g = gaussian_field(n, slope, sigma, seed)
    if growth is not None:
        g = (g * np.float32(growth)).astype(np.float32)
    delta = lognormal_delta(g)
    emis = point_sources(delta, n_src, seed)
    rec = ((1.0 + delta.astype(np.float64)) ** 2).astype(np.float32)
    return delta, emis, rec


def mean_xhii_direct(xhii: np.ndarray) -> float:
    return float(np.mean(xhii, dtype=np.float64))


def d2max(R: float, dx: float) -> int:
    """Largest integer m with sqrt(m)*dx <= R (the ball test of
    get_HIIbubbles.c:366, :443, :447)."""
    m = int(math.floor((R / dx) ** 2))
    while m > 0 and not (math.sqrt(m) * dx <= R):
        m -= 1
    while math.sqrt(m + 1) * dx <= R:
        m += 1
    return m
