"""Fused finder on box sides without a register transform plan (multiples
of 8 with prime factors 2, 3 and 5): mixed-radix transforms through LDS,
one radius per step, then the same dilation and final pass as the plan
sides.

  * 80 (8*2*5), 120 (8*3*5), 200 (8*5*5), 240 (8*2*3*5: a second middle
    stage): the fused parity bar against the FP64 oracle (the xHII box
    equals the exact dilation of the GPU's own centre map merged with the
    quick fill, bit for bit; centre-map differences only at near-threshold
    cells; mean x_HII);
  * 600 (8*3*5*5, the N_smoothed of the reference's sample simfast21.ini,
    with kernels specialised for it; dilation lanes of 15 cells): a plane
    wave through the filtered-field hook equals W(|k| R) times the wave
    (closed form), and the xHII box of a synthetic slice equals the exact
    dilation of the GPU's own centre map merged with the quick fill, bit for
    bit, with the neutral fraction its FP64 mean;
  * the autotuner's choice on these sides (the line passes' XCD block
    order): cached, reloaded, and bit-identical for every mask."""
import numpy as np
import pytest

import simfast21_amd as sf
from helpers import direct_case
import oracle_bind as orc
from test_gpu_fused import check_vs_oracle, quick_fill_np


@pytest.mark.gpu
@pytest.mark.parametrize("n,zeta", [(80, 1.0), (120, 1.0), (120, 3.0), (200, 1.0), (240, 1.0)])
def test_generic_vs_oracle(torch_cuda, n, zeta):
    case = direct_case(n, zeta)
    ctx = sf.HiiContext(n, case["side"], case["radii"], sf.INPUT_DIRECT, sf.PATH_FUSED)
    assert ctx.radius_batch == 1
    ctx.close()
    check_vs_oracle(torch_cuda, case, dilate=orc.dilate_maxplus, label=f"generic n={n} zeta={zeta}")


def w_tophat64(x):
    x = np.asarray(x, np.float64)
    return np.where(x < 1e-3, 1.0 - x * x / 10.0, 3.0 * (np.sin(x) - x * np.cos(x)) / x ** 3)


@pytest.mark.gpu
def test_generic_600_plane_wave(torch_cuda):
    """rec = 1 + 0.5 cos(2 pi k1.x / n) + 0.3 cos(2 pi k3.x / n), ion = 0.25
    + 0.4 cos(2 pi k2.x / n): the filtered pair at radius R is n^3 (1 + 0.5
    W(|k1| dk R) cos1 + 0.3 W(|k3| dk R) cos3, 0.25 + 0.4 W(|k2| dk R) cos2),
    W(0) = 1.  k2 = (597, 4, 590) folds to (-3, 4, -10); k3 = (250, 300, 7)
    holds the Nyquist index 300.  Within 1e-5 of the field maximum."""
    torch = torch_cuda
    n, side = 600, 600.0
    radii = [40.0, 9.0, 2.5]
    i = np.arange(n, dtype=np.float64)
    k1, k2, k3 = (3, 7, 2), (597, 4, 590), (250, 300, 7)

    def wave(k):
        ph = (2 * np.pi / n) * ((k[0] * i)[:, None, None] + (k[1] * i)[None, :, None]
                                + (k[2] * i)[None, None, :])
        return np.cos(ph)

    c1, c2, c3 = wave(k1), wave(k2), wave(k3)
    rec = (1.0 + 0.5 * c1 + 0.3 * c3).astype(np.float32)
    ion = (0.25 + 0.4 * c2).astype(np.float32)
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    a = torch.from_numpy(rec.ravel()).cuda()
    b = torch.from_numpy(ion.ravel()).cuda()
    dk = 2 * np.pi / side
    kk1 = np.sqrt(sum(min(q, n - q) ** 2 for q in k1)) * dk
    kk2 = np.sqrt(sum(min(q, n - q) ** 2 for q in k2)) * dk
    kk3 = np.sqrt(sum(min(q, n - q) ** 2 for q in k3)) * dk
    for r, R in enumerate(radii):
        g = ctx.filtered(a, b, radius_index=r).cpu().numpy().astype(np.float64) / n ** 3
        er = 1.0 + 0.5 * w_tophat64(kk1 * R) * c1.ravel() + 0.3 * w_tophat64(kk3 * R) * c3.ravel()
        ei = 0.25 + 0.4 * w_tophat64(kk2 * R) * c2.ravel()
        err_r = np.abs(g[:, 0] - er).max() / np.abs(er).max()
        err_i = np.abs(g[:, 1] - ei).max() / np.abs(ei).max()
        print(f"R={R}: W1 {w_tophat64(kk1 * R):.6f} W2 {w_tophat64(kk2 * R):.6f} "
              f"W3 {w_tophat64(kk3 * R):.2e} "
              f"rel err rec {err_r:.2e} ion {err_i:.2e}", flush=True)
        assert err_r < 1e-5 and err_i < 1e-5
    ctx.close()


@pytest.mark.gpu
def test_generic_600_dilation_and_mean(torch_cuda):
    torch = torch_cuda
    n, side = 600, 937.5
    delta, emis, rec = sf.synth.direct_boxes(n, side, 20000, 7)
    ion = emis.astype(np.float32)
    radii = sf.synth.geometric_ladder(281.25, 0.9, side / n)
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    a = torch.from_numpy(rec.ravel()).cuda()
    b = torch.from_numpy(ion.ravel()).cuda()
    x, neu = ctx.run_slice(a, b)
    torch.cuda.synchronize()
    x = x.cpu().numpy()
    w = ctx.weight_map()
    ctx.close()
    ncentre = int(np.count_nonzero(w >= 0))
    print(f"n=600: {len(radii)} radii, centres {ncentre}, mean x_HII {1 - neu:.6f}", flush=True)
    assert 0 < ncentre < n ** 3
    dil = orc.dilate_maxplus(n, w).astype(bool)
    expect = np.where(dil, np.float32(1.0), quick_fill_np(rec.ravel(), ion.ravel()))
    assert np.array_equal(x, expect)
    assert abs((1.0 - x.astype(np.float64).mean()) - neu) < 1e-9


@pytest.mark.gpu
def test_generic_autotune_records_and_reloads(torch_cuda, tmp_path):
    """On a side without a register plan the autotuner times the four
    XCD-order masks of the line passes (block order only): the winner is
    recorded in the cache file, a new context created with automatic
    settings picks it up, and every mask gives the untuned box bit for bit."""
    import os
    torch = torch_cuda
    n = 80
    case = direct_case(n, 1.0)
    side, radii = case["side"], case["radii"]
    a = torch.from_numpy(case["rec"].ravel()).cuda()
    b = torch.from_numpy(case["ion"].ravel()).cuda()
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    assert ctx.line_order == 4
    x0, neu0 = ctx.run_slice(a, b)
    x0 = x0.cpu().numpy().copy()
    cache = tmp_path / "tune.txt"
    ms = ctx.autotune(a, b, cache_path=cache)
    m = ctx.line_order
    print(f"\ngeneric autotune {n}^3: line order {m}, {ms:.3f} ms/slice", flush=True)
    assert ms > 0 and m in (0, 2, 4, 6)
    lines = cache.read_text().strip().splitlines()
    assert len(lines) == 1 and f"n={n} L={side:.17g} radii={len(radii)} ladder=" in lines[0]
    assert f"| gx={m} |" in lines[0]
    x1, neu1 = ctx.run_slice(a, b)
    assert np.array_equal(x1.cpu().numpy(), x0) and neu1 == neu0
    key = lines[0].split("|")[0].strip()
    ctx.close()
    old = os.environ.get("SF21_TUNE_CACHE")
    os.environ["SF21_TUNE_CACHE"] = str(cache)
    try:
        ctx2 = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
        assert ctx2.line_order == m
        ctx2.close()
        # every mask through the cache (the last matching line wins)
        for mask in (0, 2, 4, 6):
            with open(cache, "a") as f:
                f.write(f"{key} | gx={mask} | 1.000 ms\n")
            c3 = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
            assert c3.line_order == mask
            x3, neu3 = c3.run_slice(a, b)
            assert np.array_equal(x3.cpu().numpy(), x0) and neu3 == neu0, mask
            c3.close()
    finally:
        if old is None:
            del os.environ["SF21_TUNE_CACHE"]
        else:
            os.environ["SF21_TUNE_CACHE"] = old
    # plan sides have no line order
    ctx4 = sf.HiiContext(64, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    with pytest.raises(sf.Sf21Error):
        ctx4.line_order
    ctx4.close()


@pytest.mark.gpu
def test_generic_batch_graph_equals_single_runs(torch_cuda):
    """Several slices of a generic side in one hipGraph (sf21_hii_run_batch_dev):
    each equals its single run bit for bit, on the first launch and on
    replays."""
    torch = torch_cuda
    n = 80
    cases = [direct_case(n, zeta=0.6 + 0.5 * k, seed=7100 + k) for k in range(3)]
    ctx = sf.HiiContext(n, cases[0]["side"], cases[0]["radii"], sf.INPUT_DIRECT, sf.PATH_FUSED)
    a = [torch.from_numpy(c["rec"].ravel()).cuda() for c in cases]
    b = [torch.from_numpy(c["ion"].ravel()).cuda() for c in cases]
    single = []
    for k in range(3):
        x, neu = ctx.run_slice(a[k], b[k])
        single.append((x.cpu().numpy(), neu))
    outs = [torch.empty_like(x) for x in a]
    for rep in range(3):
        _, neus = ctx.run_batch(a, b, out_list=outs)
        for k in range(3):
            assert np.array_equal(outs[k].cpu().numpy(), single[k][0]), (rep, k)
            assert neus[k] == single[k][1], (rep, k)
    ctx.close()


SAMPLE_INI = """# the reference's sample simfast21.ini, bubble-finder keys, one redshift
nthreads = 40
zmax = 15
zmin = 15
dz = 1.0
sim_length = 150
hubble = 0.71
omega_matter = 0.27
omega_baryon = 0.046
omega_lambda = 0.73
N_halo = 1200
N_smoothed = 600
fesc = 0.25
Ion_cutoff = 0.999
bubble_Rmax = 60
bubble_Nbins = 50
use_SFR = T
"""


@pytest.mark.gpu
def test_cli_sample_ini_600(tmp_path, torch_cuda):
    """The drop-in CLI on the reference's sample configuration (N_smoothed =
    600, sim_length 150, h 0.71, bubble_Rmax 60 with 50 bins, SFR mode) at
    z = 15 with synthetic deltanl and SFRD boxes: it runs, writes the xHII box
    and the text tables, and its box equals, bit for bit, the library's run
    on the same inputs with the ini's derived parameters (box side, ladder);
    the x_HI table holds the box's FP64 mean neutral fraction."""
    import subprocess
    from test_gpu_modes import sfr_case
    torch = torch_cuda
    n, z = 600, 15.0
    (tmp_path / "delta").mkdir()
    (tmp_path / "SFR").mkdir()
    (tmp_path / "simfast21.ini").write_text(SAMPLE_INI)
    c = sfr_case(n=n, z=z, seed=77, zeta=2.0, hubble=0.71, side_mpc=150.0, fesc=0.25)
    c["delta"].astype(np.float32).tofile(tmp_path / "delta" / f"deltanl_z{z:.3f}_N600_L150.0.dat")
    c["sfrd"].astype(np.float32).tofile(tmp_path / "SFR" / f"sfrd_z{z:.3f}_N600_L150.0.dat")
    r = subprocess.run([str(sf.CLI_PATH), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("neutral fraction=") == 1
    x_cli = np.fromfile(tmp_path / "Ionization" / f"xHII_z{z:.3f}_N600_L150.0.dat", np.float32)
    assert x_cli.size == n ** 3
    assert x_cli.min() >= 0.0 and x_cli.max() <= 1.0
    xhi = np.array((tmp_path / "Output_text_files" / "x_HI_N600_L150.0.dat").read_text().split(), float)
    assert xhi.size == 1
    mean_neutral = 1.0 - x_cli.astype(np.float64).mean()
    assert abs(xhi[0] - mean_neutral) < 2e-6
    assert 0.0 < mean_neutral < 1.0
    # the library with the ini's derived values
    p = sf.read_params(tmp_path)
    radii, _ = sf.ladder(p["bubble_rmax"], p["dx"], p["bubble_nbins"])
    ctx = sf.HiiContext(n, p["box_len"], radii, sf.INPUT_SFR, sf.PATH_FUSED, hubble=p["hubble"],
                        fesc=p["fesc"])
    x_lib, neu = ctx.run_slice(torch.from_numpy(c["delta"].astype(np.float32)).cuda(),
                               torch.from_numpy(c["sfrd"]).cuda(), z=z)
    ctx.close()
    print(f"sample ini at 600: {len(radii)} radii, mean x_HII {1 - neu:.6f}", flush=True)
    assert np.array_equal(x_lib.cpu().numpy(), x_cli)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 24, 40, 48, 72, 144, 160, 216, 320, 432])
def test_generic_plans_filtered_and_dilation(torch_cuda, n):
    """Runtime plans of many shapes -- radix 8 first, then middles of 2, 3,
    4, 5, 8, two to five stages (16 = 8*2, 24 = 8*3, 40 = 8*5, 48 = 8*2*3,
    72 = 8*3*3, 144 = 8*2*3*3, 160 = 8*4*5, 216 = 8*3*3*3, 320 = 8*8*5, 432 =
    8*2*3*3*3) and dilation lanes of 1 to 9 cells: the filtered pair at two
    radii within 1e-5 of the FP64 oracle's (relative to the field maximum),
    and the xHII box equal to the exact dilation of the GPU's own centre map
    merged with the quick fill."""
    torch = torch_cuda
    side = 1.5 * n
    rng = np.random.default_rng(900 + n)
    rec = rng.lognormal(0.0, 0.8, n ** 3).astype(np.float32)
    ion = (rng.lognormal(-0.3, 1.2, n ** 3) * (rng.random(n ** 3) < 0.3)).astype(np.float32)
    radii = [0.25 * side, 6.0, 2.0]
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    a = torch.from_numpy(rec).cuda()
    b = torch.from_numpy(ion).cuda()
    for k in (0, 2):
        g = ctx.filtered(a, b, radius_index=k).cpu().numpy().astype(np.float64)
        rr, ii = orc.filtered(n, side, radii[k], rec, ion, precision=64)
        er = np.abs(g[:, 0] - rr).max() / np.abs(rr).max()
        ei = np.abs(g[:, 1] - ii).max() / np.abs(ii).max()
        assert er < 1e-5 and ei < 1e-5, (n, k, er, ei)
    x, neu = ctx.run_slice(a, b)
    torch.cuda.synchronize()
    w = ctx.weight_map()
    ctx.close()
    dil = orc.dilate_maxplus(n, w).astype(bool)
    expect = np.where(dil, np.float32(1.0), quick_fill_np(rec, ion))
    assert np.array_equal(x.cpu().numpy(), expect)
    print(f"n={n}: rel err ok, centres {int((w >= 0).sum())}, mean x_HII {1 - neu:.5f}", flush=True)
