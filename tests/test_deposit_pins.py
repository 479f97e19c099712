"""CPU pins of the oracle's restatements of the steps that make the bubble
finder's inputs: SFRD (get_SFR.c:98-148) and the non-linear density
(get_nldensity.c:95-166)."""
import math

import numpy as np

import oracle_bind as orc


def test_sfrd_single_halo_closed_form():
    nh, n, L = 16, 8, 50.0
    dx_h = L / nh
    z = 8.0
    M = 3.0e9
    halos = np.array([[M, 5.0, 7.0, 9.0]], np.float32)  # integer position: one cell
    out, mean = orc.sfrd(halos, z, nh, n, dx_h)
    s = orc.lib().orc_Rion(np.float32(M), z) / orc.lib().orc_Qion(z)
    sfr = float(np.float32(s))
    sm = nh // n
    cell = (5 // sm * n + 7 // sm) * n + 9 // sm
    assert math.isclose(out[cell], sfr / sm ** 3 / dx_h ** 3, rel_tol=1e-6)
    assert np.count_nonzero(out) == 1
    # mean SFRD = total SFR / box volume
    assert math.isclose(mean, sfr / L ** 3, rel_tol=1e-6)


def test_nldensity_zero_velocity_keeps_mass_in_place():
    nh, n = 16, 8
    rng = np.random.default_rng(2)
    d0 = rng.normal(0, 0.3, nh ** 3).astype(np.float32)
    v0 = np.zeros(nh ** 3, np.float32)
    g_in = 0.2
    box, out = orc.nldensity(nh, n, 100.0, 0.7, g_in, 0.5, v0, v0, v0, d0)
    mass = (d0.astype(np.float64) * g_in + 1.0).astype(np.float32)
    assert np.array_equal(box, (np.float32(-1.0) + mass))
    blocks = (mass.astype(np.float64) - 1).reshape(n, 2, n, 2, n, 2).mean(axis=(1, 3, 5)).ravel()
    assert np.allclose(out, blocks, atol=1e-6)


def test_nldensity_uniform_shift_and_mass_conservation():
    nh, n, L_h, h = 16, 8, 100.0, 0.7
    rng = np.random.default_rng(4)
    d0 = rng.normal(0, 0.3, nh ** 3).astype(np.float32)
    # displacement of exactly 2 cells along i: v * (nh/L_h * h) * growth = 2
    growth = 0.5
    v = np.full(nh ** 3, 2.0 / growth / (nh / L_h * h), np.float32)
    zero = np.zeros(nh ** 3, np.float32)
    box, _ = orc.nldensity(nh, n, L_h, h, 1.0, growth, v, zero, zero, d0)
    mass = (d0.astype(np.float64) + 1.0).reshape(nh, nh, nh)
    assert np.allclose(box.reshape(nh, nh, nh), np.roll(mass, 2, axis=0) - 1.0, atol=2e-5)
    # random displacements: every cell's mass lands somewhere
    vr = [rng.normal(0, 3.0, nh ** 3).astype(np.float32) for _ in range(3)]
    box2, _ = orc.nldensity(nh, n, L_h, h, 1.0, growth, vr[0], vr[1], vr[2], d0)
    assert math.isclose(float(np.sum(box2.astype(np.float64) + 1.0)), float(mass.sum()), rel_tol=1e-6)
