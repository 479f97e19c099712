"""The batched workload at its stated size: 512^3 float32, side 400 Mpc,
sigma 1.2, 2.5e5 point sources per slice, 10 slices whose fields are one
Gaussian field rescaled by linear growth factors 0.10-0.25, radii 100 Mpc by
0.9 down to the cell size (47 radii) -- the benchmark's own inputs.

  * the 10-slice hipGraph batch equals the 10 slices run one at a time, bit
    for bit, on capture and on replay;
  * one slice (the last, growth 0.25) against the oracle: at seven ladder
    radii spread over both loops (get_HIIbubbles.c:298-398 and :400-462),
    the FP32 filtered fields against FP64 (max error / max value < 1e-5)
    and the centre flags over the whole box under the FP32-typing allowance
    (tests/parity.py); and the xHII box equals, bit for bit, the oracle's
    max-plus dilation of the GPU's own centre map over the quick fill, so
    the reported neutral fraction is that box's FP64 mean."""
import numpy as np
import pytest

import oracle_bind as orc
import parity
import simfast21_amd as sf

pytestmark = pytest.mark.gpu

N, SIDE, SLICES, SOURCES = 512, 400.0, 10, 250000


@pytest.fixture(scope="module")
def c3(torch_cuda):
    import bench
    torch = torch_cuda
    radii = sf.synth.geometric_ladder(100.0, 0.9, SIDE / N)
    recs, ions = bench.make_inputs(torch, N, SIDE, SLICES, SOURCES, seed=3001)
    ctx = sf.HiiContext(N, SIDE, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    yield dict(torch=torch, radii=radii, recs=recs, ions=ions, ctx=ctx)
    ctx.close()
    del recs, ions
    torch.cuda.empty_cache()


def test_c3_batch_graph_equals_single_slices(c3):
    torch, ctx = c3["torch"], c3["ctx"]
    assert len(c3["radii"]) == 47
    outs = [torch.empty_like(x) for x in c3["recs"]]
    for rep in range(2):  # capture, then replay
        _, neus = ctx.run_batch(c3["recs"], c3["ions"], out_list=outs)
        torch.cuda.synchronize()
        for s in range(SLICES):
            x1, n1 = ctx.run_slice(c3["recs"][s], c3["ions"][s])
            assert torch.equal(outs[s], x1) and neus[s] == n1, (rep, s)
    print(f"c3 batch: mean x_HII per slice {[round(1 - v, 4) for v in neus]}", flush=True)


def test_c3_slice_vs_oracle(c3):
    torch, ctx, radii = c3["torch"], c3["ctx"], c3["radii"]
    a, b = c3["recs"][-1], c3["ions"][-1]
    x, neu = ctx.run_slice(a, b)
    x = x.cpu().numpy()
    w = ctx.weight_map()
    case = dict(n=N, side=SIDE, radii=radii, rec=a.cpu().numpy(), ion=b.cpu().numpy())
    x_own, neu_own = parity.box_from_map(N, w, case["rec"], case["ion"], orc.dilate_maxplus)
    assert np.array_equal(x, x_own)
    assert abs(neu - neu_own) < 1e-12
    del x, x_own, w
    n_main = sum(1 for R in radii if R >= 2 * SIDE / N)
    idx = [0, 10, 20, 30, n_main - 1, n_main + 2, len(radii) - 1]
    print(f"c3 slice 10: mean x_HII {1 - neu:.6f}; radii checked {idx} of {len(radii)} "
          f"({n_main} in the main loop)", flush=True)
    parity.check_slice(ctx, a, b, case, label="c3 512", radii_idx=idx)
