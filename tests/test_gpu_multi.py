"""Uneven-size batch in one hipGraph launch (sf21_hii_run_multi_dev): every
box equals the same box run alone through its context, bit for bit, on the
first launch, on graph replays, and after the inputs change in place."""
import numpy as np
import pytest

import simfast21_amd as sf
from helpers import direct_case

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
