"""The bubble finder's inputs on the GPU: SFRD from halo catalogs
(sf21_sfrd_dev, get_SFR CLI) and the non-linear density (sf21_nldensity_dev,
get_nldensity CLI) against the oracle's restatements."""
import subprocess

import numpy as np
import pytest

import oracle_bind as orc
import simfast21_amd as sf

pytestmark = pytest.mark.gpu


def catalog(nh, count, seed):
    rng = np.random.default_rng(seed)
    m = 10 ** rng.uniform(7.5, 10.5, count)
    pos = rng.uniform(0, nh, (count, 3))
    pos[:3] = [[0.0, 0.0, 0.0], [nh - 0.25, 3.5, 7.0], [-0.3, 2.2, 1.0]]
    return np.column_stack([m, pos]).astype(np.float32)


def smooth_field(n, seed, amp):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    g = np.real(np.fft.ifftn(np.fft.fftn(g) * np.exp(-k2 * 60)))
    return (amp * g / g.std()).astype(np.float32).ravel()


def test_sfrd_vs_oracle(torch_cuda):
    torch = torch_cuda
    n, nh, L_h, z = 64, 128, 70.0, 8.0
    cat = catalog(nh, 20000, 11)
    ctx = sf.HiiContext(n, L_h, [L_h / n], sf.INPUT_HALO, sf.PATH_FUSED, n_halo=nh, radius_batch=1)
    out, mean = ctx.sfrd(torch.from_numpy(cat).cuda(), z)
    ctx.close()
    ref, rmean = orc.sfrd(cat, z, nh, n, L_h / nh)
    err = np.abs(out.cpu().numpy() - ref) / ref.max()
    print(f"\nsfrd: max |err|/max {err.max():.2e}, mean {mean:.8e} / {rmean:.8e}", flush=True)
    assert err.max() < 2e-5 and abs(mean - rmean) <= 1e-5 * abs(rmean)


def test_nldensity_vs_oracle(torch_cuda):
    torch = torch_cuda
    n, nh, L_h, h, z = 64, 128, 140.0, 0.7, 9.0
    p = dict(hubble=h, omega_m=0.3, omega_b=0.045, omega_lambda=0.7, box_len=L_h, n=n, n_halo=nh)
    d0 = smooth_field(nh, 21, 2.0)
    v = [smooth_field(nh, 22 + k, 8.0) for k in range(3)]  # Mpc: a few cells of displacement
    dev = lambda a: torch.from_numpy(a).cuda()
    ctx = sf.HiiContext(n, L_h, [L_h / n], sf.INPUT_HALO, sf.PATH_FUSED, n_halo=nh, radius_batch=1)
    box = torch.empty(nh ** 3, dtype=torch.float32, device="cuda")
    out, _ = ctx.nldensity(p, z, dev(v[0]), dev(v[1]), dev(v[2]), dev(d0), box_halo=box)
    ctx.close()
    c = orc.cosmo(h, 0.3, 0.045, 0.7)
    g_in = orc.growth(c, 300.0)
    g = orc.growth(c, z) - g_in
    rbox, rout = orc.nldensity(nh, n, L_h, h, g_in, g, v[0], v[1], v[2], d0)
    # FP32 positions (the reference rounds FP64 ones) move CIC weights by
    # ~1e-5 of a cell's mass near i ~ N_halo
    eb = np.abs(box.cpu().numpy() - rbox).max()
    eo = np.abs(out.cpu().numpy() - rout).max()
    print(f"\nnldensity: max |err| halo grid {eb:.2e}, smoothed {eo:.2e}", flush=True)
    assert eb < 2e-4 and eo < 5e-5
    assert abs(float(np.sum(box.cpu().numpy().astype(np.float64) + 1.0)) -
               float(np.sum(d0.astype(np.float64) * g_in + 1.0))) < 1e-4 * nh ** 3


INI = """N_smoothed = {n}
N_halo = {nh}
sim_length = 100.0
hubble = 0.7
omega_matter = 0.3
omega_baryon = 0.045
omega_lambda = 0.7
zmax = 10.0
zmin = 9.5
dz = 0.5
use_SFR = T
Original_nldensity_box = T
"""


@pytest.mark.parametrize("n,nh", [(64, 128), (60, 120)])
def test_input_step_clis_vs_oracle(tmp_path, torch_cuda, n, nh):
    """The deposit CLIs on a base_dir; (60, 120) has the reference sample
    ini's N_halo/N_smoothed = 1200/600 shape, sizes with no transform plan
    (the deposits need none)."""
    lmpc = 100.0
    L_h = lmpc * 0.7
    base = tmp_path / "run"
    for sub in ("Halos", "Velocity", "delta"):
        (base / sub).mkdir(parents=True)
    (base / "simfast21.ini").write_text(INI.format(n=n, nh=nh))
    cats = {}
    for k, z in enumerate((10.0, 9.5)):
        cat = catalog(nh, 5000 + 1000 * k, 30 + k)
        cats[z] = cat
        with open(base / "Halos" / f"halonl_z{z:.3f}_N{nh}_L{lmpc:.1f}.dat.catalog", "wb") as f:
            np.array([len(cat)], np.int64).tofile(f)
            cat.tofile(f)
    d0 = smooth_field(nh, 40, 2.0)
    v = [smooth_field(nh, 41 + k, 6.0) for k in range(3)]
    d0.tofile(base / "delta" / f"delta_z0_N{nh}_L{lmpc:.1f}.dat")
    for k, ax in enumerate("xyz"):
        v[k].tofile(base / "Velocity" / f"vel_{ax}_z0_N{nh}_L{lmpc:.1f}.dat")
    bdir = sf.PKG_DIR / "_build"
    r = subprocess.run([str(bdir / "get_SFR"), str(base)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Redshift: 10.000000. Reading 5000 halos..." in r.stdout
    table = (base / "Output_text_files" / f"sfrd_av_N{n}_L{lmpc:.1f}.dat").read_text().split("\n")
    for k, z in enumerate((10.0, 9.5)):
        ref, rmean = orc.sfrd(cats[z], z, nh, n, L_h / nh)
        got = np.fromfile(base / "SFR" / f"sfrd_z{z:.3f}_N{n}_L{lmpc:.1f}.dat", np.float32)
        assert np.abs(got - ref).max() / ref.max() < 2e-5
        zz, mm = table[k].split()
        assert zz == f"{z:f}" and abs(float(mm) - rmean) <= 1e-5 * rmean
    r = subprocess.run([str(bdir / "get_nldensity"), str(base)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    c = orc.cosmo(0.7, 0.3, 0.045, 0.7)
    g_in = orc.growth(c, 300.0)
    for z in (10.0, 9.5):
        rbox, rout = orc.nldensity(nh, n, L_h, 0.7, g_in, orc.growth(c, z) - g_in, v[0], v[1], v[2], d0)
        got = np.fromfile(base / "delta" / f"deltanl_z{z:.3f}_N{n}_L{lmpc:.1f}.dat", np.float32)
        goto = np.fromfile(base / "delta" / f"deltanl_o_z{z:.3f}_N{nh}_L{lmpc:.1f}.dat", np.float32)
        assert np.abs(got - rout).max() < 5e-5 and np.abs(goto - rbox).max() < 2e-4
    r2 = subprocess.run([str(bdir / "get_nldensity"), str(base)], capture_output=True, text=True)
    assert r2.stdout.count("already exists - skipping...") == 2


@pytest.mark.parametrize("n,nh", [(60, 120), (75, 150)])
def test_deposit_context_any_size(torch_cuda, n, nh):
    """SF21_PATH_DEPOSIT: sizes without a transform plan, SFRD and the
    non-linear density against the oracle; the finder's entry points refuse
    such a context."""
    torch = torch_cuda
    L_h, z, h = 90.0, 8.5, 0.7
    ctx = sf.HiiContext(n, L_h, [], sf.INPUT_HALO, sf.PATH_DEPOSIT, n_halo=nh)
    cat = catalog(nh, 8000, 50 + n)
    out, mean = ctx.sfrd(torch.from_numpy(cat).cuda(), z)
    ref, rmean = orc.sfrd(cat, z, nh, n, L_h / nh)
    err = np.abs(out.cpu().numpy() - ref) / ref.max()
    assert err.max() < 2e-5 and abs(mean - rmean) <= 1e-5 * abs(rmean)
    p = dict(hubble=h, omega_m=0.3, omega_b=0.045, omega_lambda=0.7, box_len=L_h, n=n, n_halo=nh)
    d0 = smooth_field(nh, 60 + n, 2.0)
    v = [smooth_field(nh, 61 + n + k, 6.0) for k in range(3)]
    dev = lambda a: torch.from_numpy(a).cuda()
    box = torch.empty(nh ** 3, dtype=torch.float32, device="cuda")
    got, _ = ctx.nldensity(p, z, dev(v[0]), dev(v[1]), dev(v[2]), dev(d0), box_halo=box)
    c = orc.cosmo(h, 0.3, 0.045, 0.7)
    g_in = orc.growth(c, 300.0)
    rbox, rout = orc.nldensity(nh, n, L_h, h, g_in, orc.growth(c, z) - g_in, v[0], v[1], v[2], d0)
    print(f"\ndeposit context {n}/{nh}: sfrd {err.max():.2e}, nldensity "
          f"{np.abs(box.cpu().numpy() - rbox).max():.2e} / {np.abs(got.cpu().numpy() - rout).max():.2e}",
          flush=True)
    assert np.abs(box.cpu().numpy() - rbox).max() < 2e-4
    assert np.abs(got.cpu().numpy() - rout).max() < 5e-5
    x = torch.zeros(n ** 3, dtype=torch.float32, device="cuda")
    with pytest.raises(sf.Sf21Error):
        ctx.run_slice(x, x)
    ctx.close()
