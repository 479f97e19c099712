"""Host logic of the product (no GPU): the parameter file, the radius ladder,
the ball bound, the exported C-ABI, and the drop-in CLI's argv/error paths."""
import ctypes
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as orc
import simfast21_amd as sf

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ladders.json")))

INI = """# test parameter file
nthreads = 4
zmax = 9 # maximum redshift
zmin = 6.7
dz = 1.0
sim_length = 100   # Mpc
hubble = 0.7
N_halo = 64
N_smoothed = 32
fesc = 0.3
Ion_cutoff = 0.99
bubble_Rmax = 80
bubble_Nbins = 20
use_SFR = T
some_unknown_key = 17
"""


def write_ini(d, text=INI):
    (d / "simfast21.ini").write_text(text)
    return d


def test_library_exports_header_symbols():
    lib = sf._lib()
    names = sf.header_symbols()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("radii", [[2.0, 3.0, 5.0], [5.0, 2.0, 3.0], [5.0, 0.0], [5.0, float("nan")]])
def test_create_rejects_unordered_ladder(radii):
    """The ladder must be descending and positive (get_HIIbubbles.c:298-461
    divides R each step); sf21_hii_create refuses anything else with
    SF21_ERR_ARG before it touches a device."""
    for path in (sf.PATH_FUSED, sf.PATH_PLAIN):
        with pytest.raises(sf.Sf21Error) as e:
            sf.HiiContext(64, 100.0, radii, sf.INPUT_DIRECT, path)
        assert e.value.code == sf.SF21_ERR_ARG


def test_create_accepts_equal_radii_past_argument_checks():
    """Repeated radii are non-increasing and pass the argument checks: with
    no device the call fails only at the device lookup."""
    import torch
    try:
        ctx = sf.HiiContext(64, 100.0, [5.0, 5.0, 3.0], sf.INPUT_DIRECT, sf.PATH_FUSED)
    except sf.Sf21Error as e:
        assert not torch.cuda.is_available() and e.code == sf.SF21_ERR_HIP
    else:
        ctx.close()


def test_params_derived_values(tmp_path):
    p = sf.read_params(write_ini(tmp_path))
    h = 0.7
    assert p["n"] == 32 and p["n_halo"] == 64 and p["nthreads"] == 4 and p["use_sfr"] == 1
    assert p["box_len"] == pytest.approx(100 * h)                 # Input_variables.c:121
    assert p["bubble_rmax"] == pytest.approx(100 * h / 2)         # capped at L/2 (:124)
    assert p["dx"] == pytest.approx(100 * h / 32)
    assert p["dk"] == pytest.approx(2 * math.pi / (100 * h))
    assert p["zmin"] == pytest.approx(9 - 1.0 * math.ceil((9 - 6.7) / 1.0))  # :133 -> 6.0
    assert p["fesc"] == 0.3 and p["ion_cutoff"] == 0.99


def test_params_errors(tmp_path):
    with pytest.raises(sf.Sf21Error) as e:
        sf.read_params(tmp_path)
    assert e.value.code == sf.SF21_ERR_IO and "cannot be open" in e.value.message
    write_ini(tmp_path, INI + "this line has no equals sign\n")
    with pytest.raises(sf.Sf21Error) as e:
        sf.read_params(tmp_path)
    assert e.value.code == sf.SF21_ERR_FORMAT and "Wrong format" in e.value.message
    write_ini(tmp_path, INI + "key_without_value =\n")
    with pytest.raises(sf.Sf21Error) as e:
        sf.read_params(tmp_path)
    assert e.value.code == sf.SF21_ERR_FORMAT


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
def test_product_ladder_matches_golden_and_oracle(case):
    dx = case["side"] / case["n"]
    radii, nm = sf.ladder(case["rmax"], dx, case["nbins"])
    assert nm == case["n_main"] and len(radii) == case["n_total"]
    oradii, onm = orc.ladder(case["rmax"], dx, case["nbins"])
    assert radii == oradii and nm == onm


@pytest.mark.parametrize("case", GOLD["geometric_0p9"], ids=lambda c: c["name"])
def test_geometric_ladders(case):
    from simfast21_amd import synth
    dx = case["side"] / case["n"]
    radii = synth.geometric_ladder(case["rmax"], 0.9, dx)
    assert len(radii) == case["n_total"]
    assert sum(r >= 2 * dx for r in radii) == case["n_main"]


def test_ball_d2max_product_vs_oracle():
    for dx in [0.78125, 1.5625, 0.7 * 200 / 128]:
        for R in np.linspace(0.5 * dx, 60 * dx, 301):
            assert sf.ball_d2max(float(R), dx) == orc.d2max(float(R), dx)


def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_create_without_gpu_fails_cleanly():
    if _have_gpu():
        pytest.skip("a HIP device is present")
    radii = (ctypes.c_double * 2)(4.0, 2.0)
    cfg = sf.HiiConfig(64, 100.0, 2, radii, sf.INPUT_DIRECT, sf.PATH_FUSED, 0.7, 1.0, 0, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    assert sf._lib().sf21_hii_create(ctypes.byref(cfg), ctypes.byref(h)) == sf.SF21_ERR_HIP
    cfg.n = 100  # unsupported size is rejected before any device work
    assert sf._lib().sf21_hii_create(ctypes.byref(cfg), ctypes.byref(h)) == sf.SF21_ERR_ARG


def _fused_side_ok(n):
    """The fused finder's side rule restated: a register-plan side, or a
    multiple of 8 in 16..768 with prime factors 2, 3, 5 and some q <= 16 with
    n/q a multiple of 8 and at most 64 (the dilation's lanes)."""
    if n in (32, 64, 96, 128, 192, 256, 512):
        return True
    if n < 16 or n > 768 or n % 8:
        return False
    m = n
    for f in (2, 3, 5):
        while m % f == 0:
            m //= f
    return m == 1 and any(n % q == 0 and n // q <= 64 and (n // q) % 8 == 0 for q in range(1, 17))


def test_size_rule_over_all_sides():
    """Every side 8..1032 in steps of 8 (and a few odd ones): accepted by
    the library exactly when the rule above holds."""
    if _have_gpu():
        pytest.skip("a HIP device is present")
    radii = (ctypes.c_double * 2)(4.0, 2.0)
    h = ctypes.c_void_p()
    accepted = []
    for n in list(range(8, 1040, 8)) + [60, 75, 100, 150]:
        cfg = sf.HiiConfig(n, 100.0, 2, radii, sf.INPUT_DIRECT, sf.PATH_FUSED, 0.7, 1.0, 0, 0, 0, 0, 0)
        rc = sf._lib().sf21_hii_create(ctypes.byref(cfg), ctypes.byref(h))
        assert rc in (sf.SF21_ERR_HIP, sf.SF21_ERR_ARG)
        assert (rc == sf.SF21_ERR_HIP) == _fused_side_ok(n), n
        if rc == sf.SF21_ERR_HIP:
            accepted.append(n)
    assert 600 in accepted and 768 in accepted and 1024 not in accepted


@pytest.mark.parametrize("n,path,ok", [
    (600, "FUSED", True), (80, "FUSED", True), (120, "FUSED", True), (200, "FUSED", True),
    (360, "FUSED", True), (768, "FUSED", True), (600, "PLAIN", False), (56, "FUSED", False),
    (100, "FUSED", False), (1024, "FUSED", False), (1000, "FUSED", False), (8, "FUSED", False)])
def test_size_acceptance(n, path, ok):
    """Sides the fused finder takes: the register-plan sides and any other
    multiple of 8 from 16 to 768 with prime factors 2, 3 and 5 whose
    dilation lanes tile it.  The plain path keeps the plan sides.
    Without a GPU an accepted size fails later, at device selection."""
    if _have_gpu():
        pytest.skip("a HIP device is present")
    radii = (ctypes.c_double * 2)(4.0, 2.0)
    cfg = sf.HiiConfig(n, 100.0, 2, radii, sf.INPUT_DIRECT, getattr(sf, "PATH_" + path), 0.7, 1.0,
                       0, 0, 0, 0, 0)
    h = ctypes.c_void_p()
    rc = sf._lib().sf21_hii_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == (sf.SF21_ERR_HIP if ok else sf.SF21_ERR_ARG)


@pytest.mark.parametrize("cli", ["product", "oracle"])
def test_cli_usage_and_errors(tmp_path, cli):
    exe = str(sf.CLI_PATH) if cli == "product" else orc.CLI
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage: get_HIIbubbles base_dir" in r.stdout
    r = subprocess.run([exe, "a", "b"], capture_output=True, text=True)
    assert r.returncode == 1
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot be open" in r.stdout
    write_ini(tmp_path)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 1 and "Error reading deltanl file" in r.stdout
    assert (tmp_path / "Ionization").is_dir() and (tmp_path / "Output_text_files").is_dir()
    assert "Bubble radius ratio (bfactor)" in r.stdout and "z = 6.000000" in r.stdout


@pytest.mark.parametrize("name", ["get_HIIbubbles", "get_SFR", "get_nldensity", "t21"])
def test_cli_reports_thread_count(tmp_path, name):
    """The reference programs are built with -D_OMPTHREAD_ (makefile:14) and
    print "Using %d threads" with the ini's nthreads right after reading the
    parameters (get_HIIbubbles.c:74-79, get_SFR.c:59-62,
    get_nldensity.c:51-54, t21.c:60-65); the drop-in CLIs print the same line
    before they touch any input (here there is none, so they then fail)."""
    write_ini(tmp_path)
    exe = os.path.join(os.path.dirname(str(sf.CLI_PATH)), name)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert "Using 4 threads\n" in r.stdout, r.stdout
    assert r.returncode == 1
