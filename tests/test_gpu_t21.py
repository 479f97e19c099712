"""21-cm brightness temperature on the GPU (sf21_t21_dev and the drop-in
t21 CLI) against the oracle's FP64 restatement of t21.c:116-252.  Sides 80
and 120 have no register plan: complex transforms of the real box through
the generic line passes."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as orc
import simfast21_amd as sf

pytestmark = pytest.mark.gpu

COSMO = dict(hubble=0.7, omega_m=0.3, omega_b=0.045, omega_lambda=0.7)


def fields(n, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n, n))
    # smooth lognormal density (a few cells correlation), mean 0
    k = np.fft.fftfreq(n)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    g = np.real(np.fft.ifftn(np.fft.fftn(g) * np.exp(-k2 * 40)))
    g *= 1.0 / g.std()
    e = np.exp(g)
    delta = (e / e.mean() - 1).astype(np.float32).ravel()
    x = rng.uniform(0, 1, n ** 3).astype(np.float32)
    xc = rng.uniform(0.1, 1.0, n ** 3).astype(np.float32)
    xa = rng.uniform(0.5, 2.0, n ** 3).astype(np.float32)
    tk = rng.uniform(20.0, 300.0, n ** 3).astype(np.float32)
    return delta, x, xc, xa, tk


@pytest.mark.parametrize("n,vel_comp,spin", [(64, 3, False), (64, 1, True), (128, 2, True),
                                             (80, 3, False), (120, 1, True)])
def test_t21_vs_oracle(torch_cuda, n, vel_comp, spin):
    torch = torch_cuda
    z = 9.0
    delta, x, xc, xa, tk = fields(n, 600 + n + vel_comp)
    L_h = 140.0
    ctx = sf.HiiContext(n, L_h, [L_h / n], sf.INPUT_DIRECT, sf.PATH_FUSED, radius_batch=1)
    dev = lambda a: torch.from_numpy(a).cuda()
    params = dict(COSMO, vel_comp=vel_comp, n=n, box_len=L_h)
    out, mt, ms = ctx.t21(params, z, dev(delta), dev(x), *(map(dev, (xc, xa, tk)) if spin else (None,) * 3))
    ctx.close()
    c = orc.cosmo(COSMO["hubble"], COSMO["omega_m"], COSMO["omega_b"], COSMO["omega_lambda"])
    ref, rmt, rms = orc.t21(n, c, vel_comp, z, delta, x, *((xc, xa, tk) if spin else (None,) * 3))
    got = out.cpu().numpy()
    err = np.abs(got.astype(np.float64) - ref) / np.abs(ref).max()
    print(f"\nt21 n={n} axis={vel_comp} spin={spin}: max |err|/max {err.max():.2e}, "
          f"mean gpu {mt:.8e} oracle {rmt:.8e}, TS {ms:.6e} / {rms:.6e}", flush=True)
    assert err.max() < 2e-5
    assert abs(mt - rmt) <= 1e-5 * abs(rmt)
    if spin:
        assert abs(ms - rms) <= 1e-5 * abs(rms)
    else:
        assert ms == 0.0


INI = """# 21-cm step test
N_smoothed = 64
N_halo = 128
sim_length = 100.0
hubble = 0.7
omega_matter = 0.3
omega_baryon = 0.045
omega_lambda = 0.7
zmax = 10.0
zmin = 9.5
dz = 0.5
Vel_comp = 3
use_Lya_xrays = T
Zminsfr = 10.0
"""


def test_t21_cli_vs_oracle(tmp_path, torch_cuda):
    """The drop-in CLI on a base_dir with two redshifts (z = 10 in the spin
    branch, z = 9.5 without): boxes against the oracle, the text tables, and
    the skip of existing outputs on a second run."""
    n, lmpc = 64, 100.0
    base = tmp_path / "run"
    for sub in ("delta", "Ionization", "x_c", "Lya", "xrays", "Output_text_files"):
        (base / sub).mkdir(parents=True)
    (base / "simfast21.ini").write_text(INI)
    data = {}
    for k, z in enumerate((10.0, 9.5)):
        delta, x, xc, xa, tk = fields(n, 700 + k)
        data[z] = (delta, x, xc, xa, tk)
        tag = f"z{z:.3f}_N{n}_L{lmpc:.1f}.dat"
        delta.tofile(base / "delta" / f"deltanl_{tag}")
        x.tofile(base / "Ionization" / f"xHII_{tag}")
        xc.tofile(base / "x_c" / f"xc_{tag}")
        xa.tofile(base / "Lya" / f"xalpha_{tag}")
        tk.tofile(base / "xrays" / f"TempX_{tag}")
    cli = str(sf.PKG_DIR / "_build" / "t21")
    r = subprocess.run([cli, str(base)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "T21 - z = 10.000000" in r.stdout and "T21 - z = 9.500000" in r.stdout
    c = orc.cosmo(0.7, 0.3, 0.045, 0.7)
    av = (base / "Output_text_files" / f"t21_av_N{n}_L{lmpc:.1f}.dat").read_text().split("\n")
    ts = (base / "Output_text_files" / f"TS_av_N{n}_L{lmpc:.1f}.dat").read_text().split("\n")
    assert av[0].startswith("10.000000 ") and av[1].startswith("9.500000 ")
    assert len([l for l in ts if l]) == 1 and ts[0].startswith("10.000000 ")
    for k, z in enumerate((10.0, 9.5)):
        delta, x, xc, xa, tk = data[z]
        spin = z == 10.0
        ref, rmt, rms = orc.t21(n, c, 3, z, delta, x, *((xc, xa, tk) if spin else (None,) * 3))
        got = np.fromfile(base / "deltaTb" / f"deltaTb_z{z:.3f}_N{n}_L{lmpc:.1f}.dat", np.float32)
        assert got.size == n ** 3
        assert np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max() < 2e-5
        assert abs(float(av[k].split()[1]) - rmt) <= 1e-5 * abs(rmt)
        if spin:
            assert abs(float(ts[0].split()[1]) - rms) <= 1e-5 * abs(rms)
    r2 = subprocess.run([cli, str(base)], capture_output=True, text=True)
    assert r2.returncode == 0 and r2.stdout.count("already exists - skipping") == 2
    assert (base / "Output_text_files" / f"t21_av_N{n}_L{lmpc:.1f}.dat").read_text().count("\n") == 2
    r3 = subprocess.run([cli], capture_output=True, text=True)
    assert r3.returncode == 1 and r3.stdout == "Usage : t21 base_dir\n"
