"""Uneven-size batch in one hipGraph launch (sf21_hii_run_multi_dev): every
box equals the same box run alone through its context, bit for bit, on the
first launch, on graph replays, and after the inputs change in place."""
import numpy as np
import pytest

import simfast21_amd as sf
from helpers import direct_case
import oracle_bind as orc

pytestmark = pytest.mark.gpu

SIZES = [64, 96, 128, 64, 96, 128, 64]
# with sides that have no register plan (LDS line transforms, one radius per step)
SIZES_GENERIC = [64, 80, 120, 80, 96]


@pytest.mark.parametrize("sizes", [SIZES, SIZES_GENERIC])
def test_multi_box_graph_equals_single_runs(torch_cuda, sizes):
    torch = torch_cuda
    cases = [direct_case(n, zeta=0.5 + 0.3 * k, seed=7000 + k) for k, n in enumerate(sizes)]
    ctxs = {}
    for c in cases:
        if c["n"] not in ctxs:
            ctxs[c["n"]] = sf.HiiContext(c["n"], c["side"], c["radii"], sf.INPUT_DIRECT, sf.PATH_FUSED)
    boxes_ctx = [ctxs[c["n"]] for c in cases]
    a = [torch.from_numpy(c["rec"].ravel()).cuda() for c in cases]
    b = [torch.from_numpy(c["ion"].ravel()).cuda() for c in cases]
    outs = [torch.empty_like(x) for x in a]

    single = []
    for k, c in enumerate(cases):
        x, neu = boxes_ctx[k].run_slice(a[k], b[k])
        single.append((x.cpu().numpy(), neu))

    for rep in range(3):
        _, neus = sf.run_multi(boxes_ctx, a, b, out_list=outs)
        for k in range(len(cases)):
            xk = outs[k].cpu().numpy()
            assert np.array_equal(xk, single[k][0]), (rep, k, sizes[k])
            assert neus[k] == single[k][1], (rep, k)
        print(f"replay {rep}: means {[round(1 - v, 5) for v in neus]}", flush=True)

    # new data in the same buffers: the cached graph sees it
    b[2].mul_(1.7)
    torch.cuda.synchronize()
    _, neus = sf.run_multi(boxes_ctx, a, b, out_list=outs)
    x2, neu2 = boxes_ctx[2].run_slice(a[2], b[2])
    assert np.array_equal(outs[2].cpu().numpy(), x2.cpu().numpy())
    assert neus[2] == neu2 and neu2 != single[2][1]
    for ctx in ctxs.values():
        ctx.close()


def test_multi_graph_recaptured_after_context_change(torch_cuda):
    """A multi-box graph cached on the first context is replayed only while
    every context it uses is unchanged: after the second context is
    autotuned (its batch buffer and kernel choices change), and after it is
    destroyed and created again (possibly at the same address), the next
    call captures a new graph, and every box still equals its single run."""
    torch = torch_cuda
    cases = [direct_case(64, zeta=0.8, seed=7101), direct_case(96, zeta=1.1, seed=7102)]
    a = [torch.from_numpy(c["rec"].ravel()).cuda() for c in cases]
    b = [torch.from_numpy(c["ion"].ravel()).cuda() for c in cases]
    outs = [torch.empty_like(x) for x in a]

    def make(c):
        return sf.HiiContext(c["n"], c["side"], c["radii"], sf.INPUT_DIRECT, sf.PATH_FUSED)

    def check(ctxs, label):
        _, neus = sf.run_multi(ctxs, a, b, out_list=outs)
        for k in range(2):
            x1, n1 = ctxs[k].run_slice(a[k], b[k])
            assert np.array_equal(outs[k].cpu().numpy(), x1.cpu().numpy()), (label, k)
            assert neus[k] == n1, (label, k)

    c0, c1 = make(cases[0]), make(cases[1])
    check([c0, c1], "first")
    check([c0, c1], "replay")
    assert c0.graph_captures == 1
    c1.autotune(a[1], b[1])
    check([c0, c1], "after autotune")
    assert c0.graph_captures == 2
    c1.close()
    c1 = make(cases[1])
    check([c0, c1], "after recreate")
    assert c0.graph_captures == 3
    print(f"captures on the first context: {c0.graph_captures}", flush=True)
    c0.close()
    c1.close()


def test_c4_parameter_sweep_graph(torch_cuda):
    """Config c4 at its stated workload: 64 boxes with N drawn uniformly from
    {64, 96, 128, 192, 256} at a fixed cell size (sides 75-300 Mpc), P(k)
    slopes in [-2.8, -2.2], lognormal efficiencies, a radius ladder per size
    -- the benchmark's own inputs -- in one multi-box hipGraph.  Every box
    equals its single run bit for bit (192^3 and 256^3 boxes included), on
    capture and on replay; the single run of a 192^3 box and of a 256^3 box
    is held to the fused parity bar against the oracle (per-radius centre
    flags over the whole box, mean x_HII)."""
    import bench
    from test_gpu_fused import check_vs_oracle
    torch = torch_cuda
    sizes, ladders, a, b = bench.uneven_inputs(torch, sf.synth, 64, 4000)
    assert 192 in sizes and 256 in sizes
    dx = bench.UNEVEN_DX
    ctxs = {n: sf.HiiContext(n, n * dx, ladders[n], sf.INPUT_DIRECT, sf.PATH_FUSED)
            for n in sorted(set(sizes))}
    boxes = [ctxs[n] for n in sizes]
    outs = [torch.empty_like(x) for x in a]
    for rep in range(2):
        _, neus = sf.run_multi(boxes, a, b, out_list=outs)
        for k, n in enumerate(sizes):
            x1, n1 = boxes[k].run_slice(a[k], b[k])
            assert torch.equal(outs[k], x1) and neus[k] == n1, (rep, k, n)
    print(f"c4: 64 boxes {sorted(set(sizes))} x counts "
          f"{[sizes.count(v) for v in sorted(set(sizes))]}, mean x_HII "
          f"{1 - max(neus):.4f}..{1 - min(neus):.4f}", flush=True)
    for n in (192, 256):
        k = sizes.index(n)
        case = dict(n=n, side=n * dx, radii=ladders[n], rec=a[k].cpu().numpy(), ion=b[k].cpu().numpy())
        check_vs_oracle(torch, case, dilate=orc.dilate_edt, label=f"c4 box {k} n={n}", ctx=ctxs[n])
    for c in ctxs.values():
        c.close()
