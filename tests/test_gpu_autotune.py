"""Autotuner of the fused path (sf21_hii_autotune): the winner is one of the
candidates, it is recorded in the cache file, a new context created with
automatic settings picks it up from $SF21_TUNE_CACHE, and the tuned context
still matches the oracle."""
import os
import struct

import numpy as np
import pytest

import oracle_bind as orc
import simfast21_amd as sf
from helpers import direct_case
from test_gpu_fused import check_vs_oracle

pytestmark = pytest.mark.gpu


def test_autotune_records_and_reloads(torch_cuda, tmp_path):
    torch = torch_cuda
    case = direct_case(96, zeta=1.0, seed=8100)
    n, side, radii = case["n"], case["side"], case["radii"]
    a = torch.from_numpy(case["rec"].ravel()).cuda()
    b = torch.from_numpy(case["ion"].ravel()).cuda()
    cache = tmp_path / "tune.txt"
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    ms = ctx.autotune(a, b, cache_path=cache)
    t = ctx.tuning()
    print(f"\nautotune {n}^3, {len(radii)} radii: {ms:.3f} ms/slice with {t}", flush=True)
    assert ms > 0
    # the weight mode is a configuration setting, never a tuned choice
    assert t["filter_tile"] in (8, 16) and t["filter_w"] == 1 and t["filter_pair"] in (0, 1)
    assert (t["cols_tile"], t["cols_wide"]) in ((8, 0), (16, 0), (16, 2), (16, 1), (16, 3), (32, 1),
                                                (32, 3))
    assert t["radius_batch"] in (8, 12, 16, 24, 32, len(radii))
    lines = cache.read_text().strip().splitlines()
    assert len(lines) == 1 and f"n={n} L={side:.17g} radii={len(radii)} ladder=" in lines[0]
    assert " w=" not in lines[0]
    x, neu = ctx.run_slice(a, b)
    # the tuned context under the full parity bar
    check_vs_oracle(torch, case, ctx=ctx, label=f"tuned n={n}")
    ctx.close()
    old = os.environ.get("SF21_TUNE_CACHE")
    os.environ["SF21_TUNE_CACHE"] = str(cache)
    try:
        ctx2 = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
        assert ctx2.tuning() == t
        x2, neu2 = ctx2.run_slice(a, b)
        assert np.array_equal(x2.cpu().numpy(), x.cpu().numpy()) and neu2 == neu
        ctx2.close()
        # explicit settings are not overridden by the cache
        ctx3 = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED, radius_batch=4)
        assert ctx3.tuning()["radius_batch"] == 4
        ctx3.close()
    finally:
        if old is None:
            del os.environ["SF21_TUNE_CACHE"]
        else:
            os.environ["SF21_TUNE_CACHE"] = old


def tune_key(arch, n, side, radii):
    """The cache key sf21_hii_autotune writes: arch, n, box side, ladder
    length and the FNV-1a hash of the radii's bytes."""
    h = 1469598103934665603
    for R in radii:
        for byte in struct.pack("<d", R):
            h = ((h ^ byte) * 1099511628211) & (2 ** 64 - 1)
    return f"{arch} n={n} L={side:.17g} radii={len(radii)} ladder={h:016x}"


@pytest.mark.parametrize("cols,rows_wide,fp,tile,K", [
    ((16, 1), 1, 0, 16, 16), ((32, 1), 0, 1, 16, 16), ((8, 0), 1, 0, 16, 16),
    ((16, 2), 0, 1, 16, 16), ((16, 3), 1, 1, 16, 16), ((32, 3), 1, 0, 16, 16),
    ((16, 0), 0, 0, 8, 16), ((16, 0), 0, 1, 8, 12), ((16, 0), 0, 0, 16, 8)])
def test_wide_variants_are_bit_identical(torch_cuda, tmp_path, cols, rows_wide, fp, tile, K):
    """Every choice the autotuner searches leaves the result unchanged bit
    for bit: 16 values per thread run the same butterflies on the same values
    as 8, the prefetching column blocks (cw bit 1) the same transforms on the
    same tiles, the filter blocks over mirror line pairs (fp) the same
    weights (same |k|^2) and transforms, and the filter tile and the radius
    batch only regroup the same per-line work (choices forced through a
    cache line for this context's key)."""
    torch = torch_cuda
    case = direct_case(128, zeta=1.0, seed=8200)
    n, side, radii = case["n"], case["side"], case["radii"]
    a = torch.from_numpy(case["rec"].ravel()).cuda()
    b = torch.from_numpy(case["ion"].ravel()).cuda()
    ref = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED, radius_batch=16, filter_tile=16)
    x0, neu0 = ref.run_slice(a, b)
    ref.close()
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    cache = tmp_path / "tune.txt"
    cache.write_text(f"{tune_key(arch, n, side, radii)} | tile={tile} K={K} ct={cols[0]} cw={cols[1]} "
                     f"rw={rows_wide} fp={fp} | 0.0 ms\n")
    old = os.environ.get("SF21_TUNE_CACHE")
    os.environ["SF21_TUNE_CACHE"] = str(cache)
    try:
        ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
        t = ctx.tuning()
        assert (t["cols_tile"], t["cols_wide"], t["rows_wide"], t["filter_pair"],
                t["filter_tile"], t["radius_batch"]) == (cols[0], cols[1], rows_wide, fp, tile, K)
        x1, neu1 = ctx.run_slice(a, b)
        ctx.close()
    finally:
        if old is None:
            del os.environ["SF21_TUNE_CACHE"]
        else:
            os.environ["SF21_TUNE_CACHE"] = old
    assert torch.equal(x0, x1) and neu0 == neu1
