"""The dilation kernels on given ball-bound maps (sf21_hii_dilate_dev):
the xHII box equals, bit for bit, the quick fill merged with the oracle's
exact periodic dilation of the same map (brute-force ball painting, or the
oracle's separable max-plus transform at 600^3), and the neutral fraction
is its FP64 mean.

The maps are chosen to reach what the finder's maps at the bench's
ionisation levels do not: bounds up to 1.8 (n/2)^2, so the line passes run
at their reach cap n/2 and meet periodic images from both sides (a ball of
bound 1.8 (n/2)^2 leaves the far corners of the box uncovered), centres
on the box's corners and edges, equal bounds (ties of the maximiser), bound
0 (a centre covering only itself), dense small bounds, and the empty map.
Sides: 64, 80, 96, 120, 128 and 600 -- 1, 2, 2, 3, 2 and 15 cells per lane
in the line passes.  The context's ladder reaches R = n cells, so every
bound up to n^2 is accepted; larger bounds and values below -1 are refused.
"""
import numpy as np
import pytest

import oracle_bind as orc
import parity
import simfast21_amd as sf
from simfast21_amd import synth

pytestmark = pytest.mark.gpu


def maps(n, rng):
    h2 = (n // 2) ** 2
    n3 = n ** 3
    out = []
    w = np.full(n3, -1, np.int32)
    w[rng.integers(n3)] = int(0.6 * 3 * h2)
    out.append(("one wide ball", w))
    w = np.full(n3, -1, np.int32)
    w[0] = h2 + 7
    w[(n - 1) * n * n + (n - 1)] = 40
    out.append(("corner balls", w))
    w = np.full(n3, -1, np.int32)
    cells = rng.choice(n3, 4, replace=False)
    w[cells] = rng.integers(h2 // 2, 3 * h2 // 2 + 1, 4)
    out.append(("sparse, bounds to 1.5 (n/2)^2", w))
    w = np.full(n3, -1, np.int32)
    m = rng.random(n3) < 0.02
    w[m] = rng.integers(0, 9, int(m.sum()))
    out.append(("dense, small bounds", w))
    w = np.full(n3, -1, np.int32)
    i0, j0 = rng.integers(n, size=2)
    w[(i0 * n + j0) * n + np.arange(0, n, 5)] = h2 // 3  # equal bounds along one line
    w[rng.choice(n3, 50, replace=False)] = 0              # centres covering only themselves
    out.append(("ties and zero bounds", w))
    out.append(("empty", np.full(n3, -1, np.int32)))
    return out


@pytest.mark.parametrize("n", [64, 80, 96, 120, 128])
def test_dilation_on_given_maps(torch_cuda, n):
    torch = torch_cuda
    rng = np.random.default_rng(1000 + n)
    radii = synth.geometric_ladder(float(n), 0.9, 1.0)  # side n: dx = 1, largest bound n^2
    ctx = sf.HiiContext(n, float(n), radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    rec = np.ones(n ** 3, np.float32)
    ion = np.zeros(n ** 3, np.float32)
    part = rng.random(n ** 3) < 0.1
    ion[part] = rng.random(int(part.sum())).astype(np.float32)  # quick fill in (0, 1)
    a = torch.from_numpy(rec).cuda()
    b = torch.from_numpy(ion).cuda()
    for name, w in maps(n, rng):
        x, neu = ctx.dilate(w, a, b)
        x = x.cpu().numpy()
        x_ref, neu_ref = parity.box_from_map(n, w, rec, ion, orc.dilate_brute)
        covered = int(orc.dilate_brute(n, w).sum())
        print(f"  n={n} {name}: covered {covered} of {n ** 3}, neutral gpu {neu:.12f} "
              f"oracle {neu_ref:.12f}", flush=True)
        assert np.array_equal(x, x_ref), (n, name)
        assert abs(neu - neu_ref) < 1e-12, (n, name)
    # bounds beyond the ladder's largest (the passes' reach cap) and values
    # below -1 are refused
    w = np.full(n ** 3, -1, np.int32)
    w[0] = synth.d2max(float(n), 1.0) + 1
    with pytest.raises(sf.Sf21Error):
        ctx.dilate(w, a, b)
    w[0] = -2
    with pytest.raises(sf.Sf21Error):
        ctx.dilate(w, a, b)
    ctx.close()


def test_dilation_on_given_maps_600(torch_cuda):
    """600^3 (15 cells per lane): a sparse map with bounds to 1.5 (n/2)^2 and
    a dense one with small bounds, against the oracle's max-plus dilation."""
    torch = torch_cuda
    n = 600
    rng = np.random.default_rng(1600)
    radii = synth.geometric_ladder(float(n), 0.9, 1.0)
    ctx = sf.HiiContext(n, float(n), radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    n3 = n ** 3
    rec = np.ones(n3, np.float32)
    ion = np.zeros(n3, np.float32)
    a = torch.from_numpy(rec).cuda()
    b = torch.from_numpy(ion).cuda()
    h2 = (n // 2) ** 2
    w1 = np.full(n3, -1, np.int32)
    w1[rng.choice(n3, 3, replace=False)] = rng.integers(h2 // 2, 3 * h2 // 2 + 1, 3)
    w2 = np.full(n3, -1, np.int32)
    m = rng.random(n3) < 4e-4
    w2[m] = rng.integers(0, 101, int(m.sum()))
    for name, w in (("sparse, bounds to 1.5 (n/2)^2", w1), ("dense, small bounds", w2)):
        x, neu = ctx.dilate(w, a, b)
        cov = orc.dilate_maxplus(n, w)
        ok = np.array_equal(x.cpu().numpy(), cov.astype(np.float32))
        neu_ref = 1.0 - cov.sum() / n3
        print(f"  n={n} {name}: covered {int(cov.sum())} of {n3}, neutral gpu {neu:.12f} "
              f"oracle {neu_ref:.12f}", flush=True)
        assert ok, name
        assert abs(neu - neu_ref) < 1e-12, name
    ctx.close()
