"""GPU checks of the differentiable mode on sides without a register
transform plan: 80^3 (80 = 8 * 2 * 5) and 120^3 (8 * 3 * 5, a radix-3
stage), side 200 Mpc * n/128, the ladder 60 Mpc * n/128 down by 0.9 to the
cell size (31 and 35 radii), zeta = 50.  The backward runs
through the line transforms of the generic sides, one radius pair at a time,
always recomputing the per-radius fields:

  * the relaxed adjoint against autograd of the FP64 torch reference
    (tests/diff_reference.py) at tau in {0.05, 0.01}, with the bars of the
    plan sides (tests/test_gpu_diff.py) -- where the gradient misses 2e-2
    (120^3 at tau 0.01), against 3x the error of the same model evaluated in
    FP32 on the CPU, the GPU's error confined to a few cells;
  * through torch.autograd.Function (tests/autograd_harness.py): the zeta
    gradient against central differences of the FP64 relaxed model;
  * the exact forward: <x> against the FP64 oracle's bubble finder on the
    same fields (FP32-undecidable flips allowed through the mean) and P_b of
    the GPU's own x against FP64 on the host;
  * field storage: the storing modes are refused, the recomputing ones
    accepted.
"""
import numpy as np
import pytest
import torch

import autograd_harness as ah
import diff_reference as dr
import oracle_bind as orc
import parity
from simfast21_amd import INPUT_DIRECT, PATH_FUSED, HiiContext, Sf21Error, synth

pytestmark = pytest.mark.gpu

ZETA = 50.0
NB = 20


@pytest.fixture(scope="module", params=[80, 120])
def setup(torch_cuda, request):
    N = request.param
    side = 200.0 * N / 128
    g = synth.gaussian_field(N, -2.5, 1.2, 7).astype(np.float64)
    delta = np.exp(g) / np.exp(g).mean() - 1
    E = synth.point_sources(delta.astype(np.float32), int(round(4000 * (N / 128) ** 3)), 7)
    radii = synth.geometric_ladder(60.0 * N / 128, 0.9, side / N)
    ctx = HiiContext(N, side, radii, INPUT_DIRECT, PATH_FUSED)
    g32 = g.astype(np.float32)
    gd = torch.from_numpy(g32.ravel()).cuda()
    ed = torch.from_numpy(E.ravel()).cuda()
    ref = dr.Setup(N, side, radii, E.astype(np.float64))
    yield dict(N=N, ctx=ctx, side=side, radii=radii, g=g32.astype(np.float64), E=E, gd=gd, ed=ed,
               ref=ref)
    ctx.close()


def _weights(ref, g, seed=11):
    """dL/dP_b scaled by the bin values so every bin weighs O(1)."""
    N = ref.n
    with torch.no_grad():
        _, P, _ = dr.relaxed(ref, torch.from_numpy(g.reshape(N, N, N)), ZETA, 0.05)
    rng = np.random.default_rng(seed)
    return rng.standard_normal(NB) / P.numpy()


@pytest.mark.parametrize("tau", [0.05, 0.01])
def test_generic_adjoint_matches_fp64_reference(setup, tau):
    s = setup
    N = s["N"]
    cp = _weights(s["ref"], s["g"])
    cx = 2.0
    grad, gz, mx, pk = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, tau, cx, cp, NB)
    grad = grad.cpu().numpy().astype(np.float64)
    _, gg, gzr = dr.loss_and_grad(s["ref"], s["g"], ZETA, tau, cx, cp, NB)
    with torch.no_grad():
        X, P, _ = dr.relaxed(s["ref"], torch.from_numpy(s["g"].reshape(N, N, N)), ZETA, tau, NB)
    X, P = float(X), P.numpy()
    rel_g = np.linalg.norm(grad - gg) / np.linalg.norm(gg)
    cos = float(grad @ gg / (np.linalg.norm(grad) * np.linalg.norm(gg)))
    print(f"\n{N}^3 tau={tau}: |dg|/|g| {rel_g:.2e} cos {cos:.8f}  zeta' gpu {gz:.6e} "
          f"ref {gzr:.6e}  <x~> gpu {mx:.8f} ref {X:.8f}  "
          f"P max rel {np.max(np.abs(pk / P - 1)):.2e}", flush=True)
    assert abs(mx - X) < 5e-5
    assert np.max(np.abs(pk / P - 1)) < 2e-4
    assert abs(gz - gzr) <= 5e-2 * abs(gzr) + 1e-5
    if rel_g < 2e-2:
        assert cos > 0.9999
    else:
        # At tau = 0.01 a handful of cells whose s_R sits at the threshold or
        # at the floor rho carry gradients that single precision cannot
        # resolve: the same model evaluated in FP32 on the CPU misses FP64 by
        # a like amount there (120^3: 6.7e-2, its error in the same few
        # cells).  The GPU is then held to 3x that FP32 evaluation's error,
        # with at least 95% of its squared error in at most 100 cells and the
        # rest of the box within 2e-2.
        _, g32, _ = dr.loss_and_grad(dr.single(s["ref"]), s["g"], ZETA, tau, cx, cp, NB,
                                     dtype=torch.float32)
        rel32 = np.linalg.norm(g32.astype(np.float64) - gg) / np.linalg.norm(gg)
        err = grad - gg
        top = np.argsort(-np.abs(err))[:100]
        share = float(np.sum(err[top] ** 2) / np.sum(err ** 2))
        rel_rest = np.linalg.norm(np.delete(err, top)) / np.linalg.norm(gg)
        print(f"  FP32 model |dg|/|g| {rel32:.2e}; GPU error in its 100 largest cells "
              f"{share:.3f} of the total, elsewhere {rel_rest:.2e}", flush=True)
        assert rel_g <= 3 * rel32 and share >= 0.95 and rel_rest < 2e-2
        assert cos > 0.99
    # Delta = exp(g)/<exp(g)> is invariant under a uniform shift of g
    assert abs(grad.sum()) <= 1e-3 * np.abs(grad).sum()


def test_generic_zeta_gradient_vs_central_differences(setup):
    """loss.backward() through torch.autograd.Function (relaxed mode) against
    the FP64 model's central difference along zeta, step 1e-3 relative."""
    s = setup
    N = s["N"]
    tau = 0.05
    _, _, _, p0 = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, tau, 1.0, None, NB)
    cp = torch.from_numpy(1.0 / (NB * p0))
    z = torch.tensor(ZETA, dtype=torch.float64, requires_grad=True)
    mx, pk = ah.observables(s["ctx"], s["ed"], s["gd"], z, tau, NB, "relaxed")
    (mx + (cp * pk).sum()).backward()
    gz = float(z.grad)
    gt = torch.from_numpy(s["g"].reshape(N, N, N))

    def loss_ref(zv):
        with torch.no_grad():
            X, P, _ = dr.relaxed(s["ref"], gt, zv, tau, NB)
        return float(X) + float((cp * P).sum())

    h = 1e-3 * ZETA
    fd = (loss_ref(ZETA + h) - loss_ref(ZETA - h)) / (2 * h)
    print(f"\n{N}^3 zeta: adjoint {gz:.6e}  central difference FP64 {fd:.6e}", flush=True)
    assert abs(fd - gz) <= 1e-2 * abs(gz)


def test_generic_exact_forward_matches_oracle(setup):
    s = setup
    N = s["N"]
    x, mx, pk = s["ctx"].diff_forward(s["gd"], s["ed"], ZETA, NB)
    x = x.cpu().numpy().astype(np.float64)
    d1 = np.exp(s["g"]) / np.exp(s["g"]).mean()
    rec = (d1 ** 2).astype(np.float32)
    ion = (ZETA * s["E"].astype(np.float64)).astype(np.float32)
    xo, neu, _ = orc.hii_slice(N, s["side"], s["radii"], rec, ion.ravel(), None, 64, 1)
    xo32, neu32, _ = orc.hii_slice(N, s["side"], s["radii"], rec, ion.ravel(), None, 32, 1)
    print(f"\n{N}^3 exact <x>: gpu {mx:.8f} oracle {1 - neu:.8f}; cells differing "
          f"{int(np.sum(np.abs(x - xo) > 1e-4))}", flush=True)
    assert abs(mx - float(np.mean(x))) < 1e-7
    parity.check_mean(1 - mx, neu, neu32, label=f"diff exact forward {N}^3",
                      d32=float(np.mean(xo32 != xo)))
    T = torch.from_numpy(((1 - x) * d1.ravel()).reshape(N, N, N))
    P = dr.power_spectrum(T, s["side"], NB).numpy()
    assert np.max(np.abs(pk / P - 1)) < 1e-4


def test_generic_field_storage_modes(setup):
    ctx = setup["ctx"]
    for mode in (1, 3):
        with pytest.raises(Sf21Error):
            ctx.diff_store_fields(mode)
    ctx.diff_store_fields(0)
    ctx.diff_store_fields(2)
