"""CPU pins of the 21-cm step (t21.c:116-252) and its background cosmology
(auxiliary.c:224-292): the oracle's and the library's host restatements
against closed forms, and the oracle's t21 against known answers."""
import math

import numpy as np
import pytest

import oracle_bind as orc
import simfast21_amd as sf

H0 = 3.24078e-18  # Input_variables.h:108


def eds_growth(z):
    """Einstein-de Sitter (Om = 1, OL = 0): H ~ (1+z)^1.5 and the growth
    integral of (1+z')^-3.5 over [z, z+100) in closed form."""
    F = lambda zz: ((1 + zz) ** -2.5 - (101 + zz) ** -2.5) / 2.5
    return (1 + z) ** 1.5 * F(z) / F(0.0)


def eds_dgrowth_dz(z):
    F = lambda zz: ((1 + zz) ** -2.5 - (101 + zz) ** -2.5) / 2.5
    dF = -((1 + z) ** -3.5 - (101 + z) ** -3.5)
    return (1.5 * (1 + z) ** 0.5 * F(z) + (1 + z) ** 1.5 * dF) / F(0.0)


@pytest.mark.parametrize("z", [0.0, 6.0, 9.5, 15.0])
def test_cosmology_closed_forms(z):
    h = 0.7
    c = orc.cosmo(h, 1.0, 0.045, 0.0)
    Hz_lib, D_lib, dDdt_lib = sf.cosmology(z, h, 1.0, 0.0, 0.045)
    Hz = h * H0 * (1 + z) ** 1.5
    assert math.isclose(orc.hubble_z(c, z), Hz, rel_tol=1e-14)
    assert math.isclose(Hz_lib, Hz, rel_tol=1e-14)
    # midpoint sum of a smooth integrand with step 1e-3: ~1e-8 relative
    assert math.isclose(orc.growth(c, z), eds_growth(z), rel_tol=1e-6)
    assert math.isclose(D_lib, eds_growth(z), rel_tol=1e-6)
    # dD/dt = dD/dz dz/dt, dz/dt = -H (1+z); central difference of step 1e-3
    ref = eds_dgrowth_dz(z) * (-Hz * (1 + z))
    assert math.isclose(orc.dgrowth_dt(c, z), ref, rel_tol=1e-5)
    assert math.isclose(dDdt_lib, ref, rel_tol=1e-5)
    if z == 0.0:
        assert orc.growth(c, 0.0) == 1.0 and D_lib == 1.0


def test_library_and_oracle_cosmology_agree():
    c = orc.cosmo(0.7, 0.3, 0.045, 0.7)
    for z in (0.0, 7.0, 11.3, 20.0):
        H, D, dD = sf.cosmology(z, 0.7, 0.3, 0.7, 0.045)
        assert math.isclose(H, orc.hubble_z(c, z), rel_tol=1e-15)
        assert math.isclose(D, orc.growth(c, z), rel_tol=1e-13)
        assert math.isclose(dD, orc.dgrowth_dt(c, z), rel_tol=1e-10)
    # LCDM growth is suppressed against EdS at low z and matches it at high z
    assert orc.growth(c, 0.0) == 1.0
    assert orc.growth(c, 20.0) * 21.0 > 1.1


def c21(z, h=0.7, om=0.3, ob=0.045):
    return 23e-3 * (0.7 / h) * (ob * h * h / 0.02) * math.sqrt((0.15 / om / h / h) * (1 + z) / 10)


def test_t21_known_answers():
    n, z = 32, 9.0
    c = orc.cosmo(0.7, 0.3, 0.045, 0.7)
    vfac = -orc.dgrowth_dt(c, z) / orc.hubble_z(c, z) / orc.growth(c, z)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, n ** 3).astype(np.float32)
    # delta = 0: no velocity term, t21 = c21 (1 - x)
    out, mt, _ = orc.t21(n, c, 3, z, np.zeros(n ** 3, np.float32), x)
    assert np.allclose(out, c21(z) * (1 - x.astype(np.float64)), rtol=1e-6, atol=0)
    assert math.isclose(mt, c21(z) * np.mean(1 - x.astype(np.float64)), rel_tol=1e-12)
    # plane wave along p: the (k_p/k)^2 filter is 1 for vel_comp = 3 (p axis)
    # and 0 for vel_comp = 1 (i axis)
    p = np.arange(n)
    d = (0.3 * np.cos(2 * np.pi * 3 * p / n))[None, None, :] * np.ones((n, n, 1))
    d32 = d.astype(np.float32).ravel()
    zero = np.zeros(n ** 3, np.float32)
    out_p, _, _ = orc.t21(n, c, 3, z, d32, zero)
    dv = np.clip(vfac * d32.astype(np.float64), -0.5, 1.0)
    assert np.allclose(out_p, c21(z) * (1 + d32) / (1 + dv), rtol=1e-6, atol=0)
    out_i, _, _ = orc.t21(n, c, 1, z, d32, zero)
    assert np.allclose(out_i, c21(z) * (1 + d32), rtol=1e-6, atol=0)
    # spin branch with uniform boxes: S = x/(1+x) (1 - Tcmb/TK), TS = Tcmb/(1 - S)
    xc = np.full(n ** 3, 0.4, np.float32)
    xa = np.full(n ** 3, 1.1, np.float32)
    tk = np.full(n ** 3, 60.0, np.float32)
    out_s, mt_s, ts = orc.t21(n, c, 3, z, zero, zero, xc, xa, tk)
    tcmb = 2.725 * (1 + z)
    xt = 0.4 + 1.1
    S = xt / (1 + xt) * (1 - tcmb / 60.0)
    assert np.allclose(out_s, c21(z) * S, rtol=1e-6)
    assert math.isclose(ts, tcmb / (1 - S), rel_tol=1e-6)


def test_params_read_t21_keys(tmp_path):
    (tmp_path / "simfast21.ini").write_text(
        "N_smoothed = 64\nN_halo = 128\nsim_length = 100\nhubble = 0.7\nzmax = 12.0\nzmin = 6.0\n"
        "dz = 0.5\nomega_matter = 0.3\nomega_baryon = 0.045\nomega_lambda = 0.7\nVel_comp = 2\n"
        "use_Lya_xrays = T\nZminsfr = 9.2\n")
    p = sf.read_params(tmp_path)
    assert p["omega_m"] == 0.3 and p["omega_b"] == 0.045 and p["omega_lambda"] == 0.7
    assert p["vel_comp"] == 2 and p["use_lya_xrays"] == 1
    # Input_variables.c:137-139: snapped down onto the z grid from zmax
    assert p["zmin_sfr"] == 12.0 - 0.5 * math.ceil((12.0 - 9.2) / 0.5)
