"""Pins of the oracle (the CPU restatement used as the checker) against what
fixes it independently: numpy's FFT, scipy's spherical Bessel function, the
photoionisation-equilibrium identity behind XHI, special values of the fits,
brute-force dilation and lattice-point counts."""
import ctypes
import json
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.special import spherical_jn

import oracle_bind as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ladders.json")))


@pytest.mark.parametrize("n", [8, 12, 16, 30, 32, 48, 64, 96, 128])
def test_oracle_fft_vs_numpy(n):
    L = orc.lib()
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n))
    out = np.zeros((n, n, n // 2 + 1), np.complex128)
    L.orc_r2c3d_d(n, x.ctypes.data, out.ctypes.data)
    ref = np.fft.rfftn(x)
    assert abs(out - ref).max() / abs(ref).max() < 1e-13
    back = np.zeros((n, n, n))
    L.orc_c2r3d_d(n, out.ctypes.data, back.ctypes.data)
    assert abs(back / n ** 3 - x).max() < 1e-12
    xf = x.astype(np.float32)
    of = np.zeros((n, n, n // 2 + 1), np.complex64)
    L.orc_r2c3d_f(n, xf.ctypes.data, of.ctypes.data)
    assert abs(of - np.fft.rfftn(xf.astype(np.float64))).max() / abs(ref).max() < 2e-6


def test_w_filter_vs_bessel():
    """W(x) = 3 j1(x)/x (auxiliary.c:350-359); the reference switches to the
    two-term series below x = 0.03, accurate there to O(x^4/280)."""
    L = orc.lib()
    xs = np.concatenate([np.linspace(0.03, 60, 2000), [0.031, 1.0, 4.4934, 7.7253]])
    for x in xs:
        w = L.orc_W_filter(float(x))
        ref = 3 * spherical_jn(1, x) / x
        assert abs(w - ref) < 1e-10 * max(1.0, abs(ref)) + 1e-12, x
    for x in np.linspace(0, 0.0299, 50):
        assert abs(L.orc_W_filter(float(x)) - (1 - x * x / 10)) < 1e-15
        ref = 3 * spherical_jn(1, x) / x if x > 0 else 1.0
        assert abs(L.orc_W_filter(float(x)) - ref) < 3e-9


def test_xhi_equilibrium_identity():
    """XHI(r) is the neutral fraction x solving x = r (1-x)^2 with x <= 1
    (user_functions.c:101-106 is the smaller root of r x^2 - (2r+1) x + r)."""
    L = orc.lib()
    assert L.orc_XHI(0.0) == 0.0
    for r in np.geomspace(1e-6, 1e3, 60):
        x = L.orc_XHI(float(r))
        assert 0 < x < 1
        # the reference's expression cancels for small r: up to ~1e-16/r^2 relative
        tol = 1e-12 + 4e-16 / (r * r)
        assert abs(x - r * (1 - x) ** 2) <= tol * max(x, 1e-300) + 1e-300, r


def test_fit_special_values():
    """Special values that fix each fitted formula's structure:
    Rrec(D = r2) = K/16 and Rrec(D -> inf) -> K with K = 1e-24 r1 (1+z)^r4
    (L_h 3.0857e24/n/h/(1+z))^3 / 1e45; Qion(z) = 1e-45 at a z + b = 45
    (10^45/V_norm = 1); G_H(0) = 4.19232273531e-13 * 10^12.08875."""
    L = orc.lib()
    z, Lh, n, h = 6.0, 100.0, 64, 0.7
    K = 1e-24 * 9.84815696 * (1 + z) ** 5.06773134 * (Lh * 3.0857e24 / n / h / (1 + z)) ** 3 / 1e45
    assert math.isclose(L.orc_Rrec(ctypes.c_float(1.76105133).value, z, Lh, n, h), K / 16, rel_tol=1e-6)
    assert math.isclose(L.orc_Rrec(1e30, z, Lh, n, h), K, rel_tol=1e-6)
    zq = (45.0 - 5.29944413e+01) / 2.05707274e-02
    assert math.isclose(L.orc_Qion(zq), 1.0, rel_tol=1e-12)
    assert math.isclose(L.orc_G_H(0.0), 4.19232273531e-13 * 10 ** 12.08875, rel_tol=1e-12)
    # Rion's exponential cutoff: Rion(c2) = c1 c2 (1+z)^c4 e^-1 / V_norm
    c1, c2, c4 = 1.575e39, 9.367608e+07, 2.27044695
    mf = ctypes.c_float(c2).value
    ref = c1 * mf * (1 + z) ** c4 * (mf / c2) ** 0.44 * math.exp(-(c2 / mf) ** 3) / 1e45
    assert math.isclose(L.orc_Rion(mf, z), ref, rel_tol=1e-12)


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
def test_ladder_reference_formula(case):
    dx = case["side"] / case["n"]
    nb = round(math.log(case["rmax"] / dx) / math.log(1 / 0.9))
    assert nb == case["nbins"]
    radii, nm = orc.ladder(case["rmax"], dx, nb)
    assert nm == case["n_main"] and len(radii) == case["n_total"]
    bf = 10 ** (math.log10(case["rmax"] / dx) / nb)
    assert abs(bf - case["bfactor"]) < 5e-7
    assert all(r >= 2 * dx for r in radii[:nm]) and all(dx <= r < 2 * dx for r in radii[nm:])


def test_d2max_bruteforce():
    for dx in [0.78125, 1.171875, 1.5625, 0.3]:
        for R in np.linspace(dx, 40 * dx, 157):
            m = orc.d2max(float(R), dx)
            assert math.sqrt(m) * dx <= R and not (math.sqrt(m + 1) * dx <= R)


def test_lattice_ball_counts():
    """One centre of weight m paints exactly the lattice points |d|^2 <= m."""
    lb = GOLD["lattice_ball_counts"]
    n = 16
    for m, cnt in zip(lb["m"], lb["count"]):
        w = np.full(n ** 3, -1, np.int32)
        w[(5 * n + 7) * n + 9] = m
        assert orc.dilate_brute(n, w).sum() == cnt


@pytest.mark.parametrize("n,seed", [(16, 1), (20, 2), (24, 3)])
def test_edt_dilation_equals_brute_force(n, seed):
    """The distance-transform dilation used at full size equals brute force
    over centres on random weighted maps, bounds up to past the half box
    (periodic images) and a centre on every face."""
    rng = np.random.default_rng(seed)
    w = np.full(n ** 3, -1, np.int32)
    cells = rng.choice(n ** 3, 12, replace=False)
    w[cells] = rng.choice([0, 1, 2, 3, 5, 9, 14, 30, (n // 2) ** 2, 3 * (n // 2) ** 2], 12)
    for face in (0, n - 1):
        w[(face * n + n // 3) * n + 2] = 4
        w[(3 * n + face) * n + n // 2] = 6
        w[(n // 2 * n + 1) * n + face] = 2
    brute = orc.dilate_brute(n, w)
    assert np.array_equal(orc.dilate_edt(n, w), brute)
    assert np.array_equal(orc.dilate_maxplus(n, w), brute)
    empty = np.full(n ** 3, -1, np.int32)
    assert orc.dilate_edt(n, empty).sum() == 0 and orc.dilate_maxplus(n, empty).sum() == 0


@pytest.mark.parametrize("n,seed", [(18, 5), (30, 9)])
def test_maxplus_dilation_equals_brute_force(n, seed):
    """The max-plus dilation (one pass per axis, any number of distinct
    bounds) against brute force on 20 random maps each, odd half sides and
    bounds past the half box included."""
    rng = np.random.default_rng(seed)
    for _ in range(20):
        w = np.full(n ** 3, -1, np.int32)
        k = int(rng.integers(1, 40))
        cells = rng.choice(n ** 3, k, replace=False)
        w[cells] = rng.choice([0, 1, 2, 3, 5, 7, 9, 14, 30, 50, (n // 2) ** 2, 3 * (n // 2) ** 2], k)
        assert np.array_equal(orc.dilate_maxplus(n, w), orc.dilate_brute(n, w))


def test_oracle_uniform_known_answers():
    """Uniform fields: every filtered field equals the input (W(0) = 1, no
    other mode).  ion >= rec everywhere ionises the box; ion = c*rec with
    c < 1 leaves the quick-fill value c in every cell."""
    n, side = 16, 32.0
    radii = [8.0, 5.0, 3.0, 2.5]
    rec = np.full(n ** 3, 2.0, np.float32)
    for prec in (32, 64):
        for dil in (0, 1):
            x, neu, fc = orc.hii_slice(n, side, radii, rec, np.full(n ** 3, 2.5, np.float32), None, prec, dil)
            assert np.all(x == 1.0) and neu == 0.0 and np.all(fc == 0)
            x, neu, fc = orc.hii_slice(n, side, radii, rec, np.full(n ** 3, 0.5, np.float32), None, prec, dil)
            assert np.allclose(x, 0.25) and abs(neu - 0.75) < 1e-7 and np.all(fc == -1)


def test_oracle_clamp_by_fresid():
    n, side = 16, 32.0
    rec = np.full(n ** 3, 1.0, np.float32)
    fr = np.full(n ** 3, 0.125, np.float32)
    x, neu, _ = orc.hii_slice(n, side, [8.0, 3.0], rec, np.full(n ** 3, 4.0, np.float32), fr, 64, 1)
    assert np.all(x == np.float32(0.875)) and abs(neu - 0.125) < 1e-12


@pytest.mark.parametrize("seed", [1, 2])
def test_exact_dilation_equals_fp64_reference_dataflow(seed):
    """The FP64 run of the reference's dataflow (ball FFT convolution with
    count > 0.01, stencil below 2dx) equals the exact ball dilation, and both
    equal brute-force dilation of the first-centre weights."""
    from helpers import direct_case, weights_from_first_centre
    case = direct_case(32, 1.0, seed=seed)
    n, side, radii = case["n"], case["side"], case["radii"]
    x0, neu0, fc0 = orc.hii_slice(n, side, radii, case["rec"], case["ion"], None, 64, 0)
    x1, neu1, fc1 = orc.hii_slice(n, side, radii, case["rec"], case["ion"], None, 64, 1)
    assert np.array_equal(fc0, fc1)
    assert np.array_equal(x0, x1) and neu0 == neu1
    w = weights_from_first_centre(fc1, radii, side / n)
    dil = orc.dilate_brute(n, w).astype(bool)
    rec, ion = case["rec"].ravel(), case["ion"].ravel()
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(rec > 0, np.minimum(ion.astype(np.float64) / rec, 1.0), 1.0)
    qf = np.where(ion > 0, q, 0.0).astype(np.float32)
    assert np.array_equal(x1, np.where(dil, np.float32(1), qf))


def test_filtered_vs_numpy_independent():
    """The oracle's filter (get_HIIbubbles.c:305-325) against numpy's FFT with
    numpy.fft.fftfreq frequencies and scipy's spherical Bessel window."""
    from helpers import direct_case
    case = direct_case(32, 1.0)
    n, side = case["n"], case["side"]
    R = case["radii"][10]
    rr, ii = orc.filtered(n, side, R, case["rec"], case["ion"], 64)
    k1 = np.fft.fftfreq(n, d=side / n) * 2 * np.pi
    kz = np.fft.rfftfreq(n, d=side / n) * 2 * np.pi
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    x = kk * R
    with np.errstate(divide="ignore", invalid="ignore"):
        W = np.where(x > 0, 3 * spherical_jn(1, x) / np.where(x > 0, x, 1), 1.0)
    for box, got in ((case["rec"], rr), (case["ion"], ii)):
        ref = np.fft.irfftn(np.fft.rfftn(box.reshape(n, n, n).astype(np.float64)) * W,
                            s=(n, n, n), axes=(0, 1, 2)) * n ** 3
        assert abs(got.reshape(n, n, n) - ref).max() / abs(ref).max() < 1e-9


def test_cic_and_smooth_box_known_answers():
    """auxiliary.c:761-808 and :643-682: an integer position deposits the
    whole value in one cell, a half-integer position 1/8 in each of 8 cells,
    mass is conserved with periodic wrap, a negative remainder is mirrored
    (the reference's x1 < 0 branch), and smooth_box averages sm^3 blocks."""
    L = orc.lib()
    L.orc_cic.argtypes = [ctypes.c_float] * 4 + [ctypes.c_void_p, ctypes.c_long]
    L.orc_smooth_box.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_long]
    n = 8
    g = np.zeros(n ** 3, np.float32)
    L.orc_cic(3.0, 4.0, 5.0, 2.0, g.ctypes.data, n)
    assert g[(3 * n + 4) * n + 5] == 2.0 and g.sum() == 2.0
    g[:] = 0
    L.orc_cic(7.5, 0.5, 3.5, 8.0, g.ctypes.data, n)
    nz = np.nonzero(g)[0]
    assert len(nz) == 8 and np.allclose(g[nz], 1.0)
    assert g[(7 * n + 0) * n + 3] == 1.0 and g[(0 * n + 1) * n + 4] == 1.0  # wraps 7 -> 0
    g[:] = 0
    L.orc_cic(-0.25, 2.0, 2.0, 1.0, g.ctypes.data, n)
    # trunc(-0.25) = 0, remainder -0.25 -> cell -1 (= 7) with 1 - 0.25, cell 0 with 0.25
    assert abs(g[(7 * n + 2) * n + 2] - 0.75) < 1e-7 and abs(g[(0 * n + 2) * n + 2] - 0.25) < 1e-7
    box = np.arange(16 ** 3, dtype=np.float32)
    out = np.zeros(8 ** 3, np.float32)
    L.orc_smooth_box(box.ctypes.data, out.ctypes.data, 16, 8)
    ref = box.reshape(8, 2, 8, 2, 8, 2).mean(axis=(1, 3, 5)).ravel()
    assert np.allclose(out, ref)


def test_oracle_cli_ion_cutoff(tmp_path):
    """The oracle CLI keeps the reference's cutoff contract
    (get_HIIbubbles.c:206, :495-513, :515-528): after a first slice whose
    neutral fraction reaches Ion_cutoff, no further input is read, every
    higher z gets an all-zero box and an x_HI line of 1.0."""
    from test_gpu_modes import check_cutoff_outputs, cutoff_base_dir
    d = cutoff_base_dir(tmp_path, "oracle", 32, 0.2)
    env = dict(os.environ, ORC_PRECISION="64", ORC_DILATION="1")
    r = subprocess.run([orc.CLI, str(d)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stdout
    assert check_cutoff_outputs(d, r.stdout, 32) >= 0.2


@pytest.mark.parametrize("zeta", [0.3, 3.0])
def test_parity_helper_restates_oracle_slice(zeta):
    """The GPU parity check rebuilds the FP64 and FP32-typing results from
    per-radius centre flags (ball bound of the largest radius at which a cell
    is a centre, exact dilation over the quick fill): that reconstruction
    equals the oracle's own slice bit for bit, and its FP64 mean to 1e-12."""
    import parity
    from helpers import direct_case, weights_from_first_centre
    c = direct_case(48, zeta)
    n, dx = c["n"], c["side"] / c["n"]
    maps = {64: np.full(n ** 3, -1, np.int32), 32: np.full(n ** 3, -1, np.int32)}
    for R in c["radii"]:
        for prec, w in maps.items():
            rr, ii = orc.filtered(n, c["side"], R, c["rec"], c["ion"], precision=prec)
            w[(w < 0) & parity.centre_flags(rr.astype(np.float64), ii.astype(np.float64))] = \
                orc.d2max(R, dx)
    for prec, w in maps.items():
        xo, neu, fc = orc.hii_slice(n, c["side"], c["radii"], c["rec"], c["ion"], precision=prec,
                                    dilation=1)
        assert np.array_equal(weights_from_first_centre(fc, c["radii"], dx), w)
        x, neu2 = parity.box_from_map(n, w, c["rec"], c["ion"], orc.dilate_brute)
        assert np.array_equal(x, xo) and abs(neu - neu2) < 1e-12


def test_fp32_typing_mean_depends_on_memory_order():
    """The noise floor behind the mean bar: the reference's FP32 typing on
    one 64^3 input, stored as given, transposed or rolled (the same
    arithmetic in another order), lands at different distances from FP64,
    more than 1e-5 away in some orders -- so a single FP32 realisation's
    distance is a noisy allowance, and the parity bar falls back on the
    typing's changed-cell count."""
    from helpers import direct_case
    n = 64
    c = direct_case(n, 0.3)
    rec, ion = c["rec"].reshape(n, n, n), c["ion"].reshape(n, n, n)
    _, o64, _ = orc.hii_slice(n, c["side"], c["radii"], rec, ion, precision=64, dilation=1)
    dev = []
    for fw in (lambda a: a, lambda a: a.transpose(1, 0, 2), lambda a: a.transpose(2, 1, 0),
               lambda a: np.roll(a, (1, 2, 3), (0, 1, 2))):
        r, i = np.ascontiguousarray(fw(rec)), np.ascontiguousarray(fw(ion))
        _, o32, _ = orc.hii_slice(n, c["side"], c["radii"], r, i, precision=32, dilation=1)
        dev.append(abs(o32 - o64))
    assert max(dev) > 1e-5 and max(dev) > 1.5 * min(dev), dev
