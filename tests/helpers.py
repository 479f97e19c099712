"""Shared test helpers: synthetic cases and the flag-parity check."""
import numpy as np

from simfast21_amd import synth

PI = 3.1415926535897932385


def direct_case(n, zeta=1.0, seed=1001, side=None, rmax=None):
    """Config-1 recipe scaled to n^3 (side 200 Mpc at 128^3, radii 60 Mpc * 0.9^k)."""
    side = side if side is not None else 200.0 * n / 128
    rmax = rmax if rmax is not None else 60.0 * n / 128
    n_src = max(8, int(round(4000 * (n / 128) ** 3)))
    delta, emis, rec = synth.direct_boxes(n, side, n_src, seed)
    ion = (emis * np.float32(zeta)).astype(np.float32)
    radii = synth.geometric_ladder(rmax, 0.9, side / n)
    return dict(n=n, side=side, radii=radii, rec=rec, ion=ion, delta=delta, emis=emis)


def weights_from_first_centre(fc, radii, dx):
    """Per-cell ball bound of the largest radius at which a cell is a centre."""
    d2 = np.array([synth.d2max(R, dx) for R in radii], np.int64)
    w = np.full(fc.shape, -1, np.int64)
    m = fc >= 0
    w[m] = d2[fc[m]]
    return w.astype(np.int32)
