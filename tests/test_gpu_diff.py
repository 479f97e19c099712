"""GPU checks of the differentiable mode (sf21_hii_diff_forward_dev /
sf21_hii_diff_backward_dev) at 64^3, side 100 Mpc, 29 radii (60 Mpc * n/128
down by 0.9 to the cell size), zeta = 50 (<x> ~ 0.5):

  * the relaxed adjoint against autograd of the FP64 torch reference
    (tests/diff_reference.py) at tau in {0.05, 0.02, 0.01}: gradient in g
    (relative L2 error), gradient in zeta, relaxed <x~> and P_b;
  * through torch.autograd.Function (tests/autograd_harness.py): exact
    forward values, loss.backward() equal to the C-ABI adjoint bit for bit;
  * central differences along zeta: the FP64 reference's (step 1e-3) against
    the GPU adjoint, the GPU's own relaxed loss (step 1e-2, above its FP32
    noise floor) against the FP64 one; along g, the FP64 model's (step 1e-6)
    against the gradient from loss.backward();
  * the exact forward: <x> and P_b against the FP64 oracle's bubble finder
    on the same fields (FP32-undecidable flips allowed through the mean);
  * tau -> 0: the relaxed <x~> tends to the exact <x>.
"""
import numpy as np
import pytest
import torch

import autograd_harness as ah
import diff_reference as dr
import oracle_bind as orc
import parity
from simfast21_amd import INPUT_DIRECT, PATH_FUSED, HiiContext, synth

pytestmark = pytest.mark.gpu

N = 64
ZETA = 50.0
NB = 20


def _case(n=N, seed=5):
    side = 200.0 * n / 128
    g = synth.gaussian_field(n, -2.5, 1.2, seed).astype(np.float64)
    delta = np.exp(g) / np.exp(g).mean() - 1
    n_src = int(round(4000 * (n / 128) ** 3))
    E = synth.point_sources(delta.astype(np.float32), n_src, seed)
    radii = synth.geometric_ladder(60.0 * n / 128, 0.9, side / n)
    return side, radii, g, E


@pytest.fixture(scope="module")
def setup(torch_cuda):
    side, radii, g, E = _case()
    ctx = HiiContext(N, side, radii, INPUT_DIRECT, PATH_FUSED)
    g32 = g.astype(np.float32)
    gd = torch.from_numpy(g32.ravel()).cuda()
    ed = torch.from_numpy(E.ravel()).cuda()
    # the reference sees exactly the float32 inputs the GPU sees
    ref = dr.Setup(N, side, radii, E.astype(np.float64))
    yield dict(ctx=ctx, side=side, radii=radii, g=g32.astype(np.float64), E=E, gd=gd, ed=ed, ref=ref)
    ctx.close()


def _weights(ref, g, seed=11):
    """dL/dP_b scaled by the bin values so every bin weighs O(1)."""
    with torch.no_grad():
        _, P, _ = dr.relaxed(ref, torch.from_numpy(g.reshape(N, N, N)), ZETA, 0.05)
    rng = np.random.default_rng(seed)
    return rng.standard_normal(NB) / P.numpy()


@pytest.mark.parametrize("tau", [0.05, 0.02, 0.01])
def test_adjoint_matches_fp64_reference(setup, tau):
    s = setup
    cp = _weights(s["ref"], s["g"])
    cx = 2.0
    grad, gz, mx, pk = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, tau, cx, cp, NB)
    grad = grad.cpu().numpy().astype(np.float64)
    L, gg, gzr = dr.loss_and_grad(s["ref"], s["g"], ZETA, tau, cx, cp, NB)
    with torch.no_grad():
        X, P, _ = dr.relaxed(s["ref"], torch.from_numpy(s["g"].reshape(N, N, N)), ZETA, tau, NB)
    X, P = float(X), P.numpy()
    rel_g = np.linalg.norm(grad - gg) / np.linalg.norm(gg)
    cos = float(grad @ gg / (np.linalg.norm(grad) * np.linalg.norm(gg)))
    print(f"\ntau={tau}: |dg|/|g| {rel_g:.2e} cos {cos:.8f}  zeta' gpu {gz:.6e} ref {gzr:.6e}  "
          f"<x~> gpu {mx:.8f} ref {X:.8f}  P max rel {np.max(np.abs(pk / P - 1)):.2e}", flush=True)
    assert abs(mx - X) < 5e-5
    assert np.max(np.abs(pk / P - 1)) < 2e-4
    assert abs(gz - gzr) <= 5e-2 * abs(gzr) + 1e-5
    assert rel_g < 2e-2
    assert cos > 0.9999
    # Delta = exp(g)/<exp(g)> is invariant under a uniform shift of g
    assert abs(grad.sum()) <= 1e-3 * np.abs(grad).sum()


def test_autograd_function_exact_forward_relaxed_backward(setup):
    """Through torch.autograd.Function (tests/autograd_harness.py): the
    forward values are the exact finder's, loss.backward() delivers the
    relaxed adjoint of the C-ABI bit for bit, and that gradient matches FP64
    autograd of the relaxed model."""
    s = setup
    tau, cx = 0.02, 2.0
    cp = _weights(s["ref"], s["g"])
    g = s["gd"].clone().requires_grad_(True)
    z = torch.tensor(ZETA, dtype=torch.float64, requires_grad=True)
    mx, pk = ah.observables(s["ctx"], s["ed"], g, z, tau, NB, "exact")
    _, mx0, pk0 = s["ctx"].diff_forward(s["gd"], s["ed"], ZETA, NB)
    assert float(mx.detach()) == mx0 and np.allclose(pk.detach().numpy(), pk0, rtol=1e-12, atol=0)
    L = cx * mx + (torch.from_numpy(cp) * pk).sum()
    L.backward()
    grad_d, gz_d, _, _ = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, tau, cx, cp, NB)
    assert torch.equal(g.grad, grad_d) and float(z.grad) == gz_d
    _, gg, gzr = dr.loss_and_grad(s["ref"], s["g"], ZETA, tau, cx, cp, NB)
    gr = g.grad.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(gr - gg) / np.linalg.norm(gg)
    print(f"\nautograd.Function: |dg|/|g| vs FP64 {rel:.2e}, dL/dzeta {float(z.grad):.6e} "
          f"FP64 {gzr:.6e}", flush=True)
    assert rel < 2e-2 and abs(float(z.grad) - gzr) <= 5e-2 * abs(gzr) + 1e-5


def test_zeta_gradient_vs_central_differences(setup):
    """Central differences along zeta through torch.autograd.Function in
    relaxed mode (zeta is a scale, so steps are relative).  Loss: <x~> plus
    the mean relative change of P_b.
      * FP64 reference, step 1e-3: pins the GPU adjoint (within 1e-2);
      * the GPU's own relaxed loss carries ~1e-5 of FP32 evaluation noise,
        decorrelated between zeta +- h: at step 1e-3 that is ~20% of this
        derivative (~1e-3), so its central difference is taken at step 1e-2
        and compared with the FP64 one at the same step (within 5e-2)."""
    s = setup
    tau = 0.05
    ctx = s["ctx"]
    _, _, _, p0 = ctx.diff_backward(s["gd"], s["ed"], ZETA, tau, 1.0, None, NB)
    cp = torch.from_numpy(1.0 / (NB * p0))
    gt = torch.from_numpy(s["g"].reshape(N, N, N))

    def loss_gpu(zv, grad=False):
        z = torch.tensor(zv, dtype=torch.float64, requires_grad=grad)
        mx, pk = ah.observables(ctx, s["ed"], s["gd"], z, tau, NB, "relaxed")
        L = mx + (cp * pk).sum()
        if grad:
            L.backward()
            return float(z.grad)
        return float(L)

    def loss_ref(zv):
        with torch.no_grad():
            X, P, _ = dr.relaxed(s["ref"], gt, zv, tau, NB)
        return float(X) + float((cp * P).sum())

    def cd(f, rel):
        h = rel * ZETA
        return (f(ZETA + h) - f(ZETA - h)) / (2 * h)

    gz = loss_gpu(ZETA, grad=True)
    fd_ref3 = cd(loss_ref, 1e-3)
    fd_ref2 = cd(loss_ref, 1e-2)
    fd_gpu2 = cd(loss_gpu, 1e-2)
    print(f"\nzeta: adjoint {gz:.6e}  central difference fp64 step 1e-3 {fd_ref3:.6e}, "
          f"step 1e-2 fp64 {fd_ref2:.6e} gpu {fd_gpu2:.6e}", flush=True)
    assert abs(fd_ref3 - gz) <= 1e-2 * abs(gz)
    assert abs(fd_gpu2 - fd_ref2) <= 5e-2 * abs(fd_ref2)


def test_g_gradient_vs_central_differences(setup):
    """The gradient in g from loss.backward() through torch.autograd.Function
    against central differences of the FP64 relaxed model along smooth
    unit-RMS directions and single cells, step 1e-6 (the relaxed model at
    tau = 0.02 is too curved for 1e-3 along g; at 1e-6 the GPU's own FP32
    loss noise would swamp the difference, so the differences are taken on
    the FP64 model, which at this step agrees with FP64 autograd to ~1e-5)."""
    s = setup
    tau, cx = 0.02, 1.0
    cp = _weights(s["ref"], s["g"])
    g = s["gd"].clone().requires_grad_(True)
    z = torch.tensor(ZETA, dtype=torch.float64)
    mx, pk = ah.observables(s["ctx"], s["ed"], g, z, tau, NB, "exact")
    (cx * mx + (torch.from_numpy(cp) * pk).sum()).backward()
    grad = g.grad.cpu().numpy().astype(np.float64)
    gflat = s["g"].ravel()

    def loss_ref(gv):
        with torch.no_grad():
            X, P, _ = dr.relaxed(s["ref"], torch.from_numpy(gv.reshape(N, N, N)), ZETA, tau, NB)
        return cx * float(X) + float(np.dot(cp, P.numpy()))

    rng = np.random.default_rng(3)
    dirs = []
    for k in range(3):
        d = synth.gaussian_field(N, -4.0, 1.0, seed=200 + k).astype(np.float64).ravel()
        dirs.append(("smooth", d / d.std()))
    # single cells where the gradient is large (near bubble edges)
    for cell in np.argsort(-np.abs(grad))[rng.integers(0, 50, 2)]:
        e = np.zeros(N ** 3)
        e[cell] = 1.0
        dirs.append((f"cell {cell}", e))
    h = 1e-6
    for name, v in dirs:
        fd = (loss_ref(gflat + h * v) - loss_ref(gflat - h * v)) / (2 * h)
        an = float(grad @ v)
        print(f"  g direction {name}: adjoint {an:.6e} central difference FP64 {fd:.6e}", flush=True)
        assert abs(fd - an) <= 3e-2 * max(abs(an), 1e-2 * np.abs(grad).max()), (name, fd, an)


def test_exact_forward_matches_oracle(setup):
    s = setup
    x, mx, pk = s["ctx"].diff_forward(s["gd"], s["ed"], ZETA, NB)
    x = x.cpu().numpy().astype(np.float64)
    d1 = np.exp(s["g"]) / np.exp(s["g"]).mean()
    rec = (d1 ** 2).astype(np.float32)
    ion = (ZETA * s["E"].astype(np.float64)).astype(np.float32)
    xo, neu, _ = orc.hii_slice(N, s["side"], s["radii"], rec, ion.ravel(), None, 64, 1)
    xo32, neu32, _ = orc.hii_slice(N, s["side"], s["radii"], rec, ion.ravel(), None, 32, 1)
    ndiff = int(np.sum(np.abs(x - xo) > 1e-4))
    print(f"\nexact <x>: gpu {mx:.8f} oracle {1 - neu:.8f}; cells differing {ndiff}", flush=True)
    assert abs(mx - float(np.mean(x))) < 1e-7
    # the host forms rec and ion in FP64 from g, the GPU in FP32 (last-bit
    # differences in the inputs); the bar is the fused finder's
    parity.check_mean(1 - mx, neu, neu32, label="diff exact forward",
                      d32=float(np.mean(xo32 != xo)))
    # P(k) of T = (1 - x) Delta from the GPU's own x, FP64 on the host
    T = torch.from_numpy(((1 - x) * d1.ravel()).reshape(N, N, N))
    P = dr.power_spectrum(T, s["side"], NB).numpy()
    assert np.max(np.abs(pk / P - 1)) < 1e-4


def test_relaxed_tends_to_exact(setup):
    s = setup
    _, mx_exact, _ = s["ctx"].diff_forward(s["gd"], s["ed"], ZETA, NB)
    _, _, mx, _ = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, 1e-3, 1.0, None, NB)
    print(f"\n<x~>(tau=1e-3) {mx:.6f}  exact <x> {mx_exact:.6f}", flush=True)
    assert abs(mx - mx_exact) < 2e-3


def test_field_storage_mode_gives_the_same_gradient(setup):
    """sf21_hii_diff_store_fields: the second sweep reads the first sweep's
    filtered pairs instead of recomputing them.  The first sweep's phi pass
    and the second sweep's seed pass run the same line transform in
    different kernels, whose FP32 contractions (fused multiply-adds) may
    differ in the last bit, so the two paths are held to the bars of the
    FP64 comparison, like the forward-kept fields below."""
    s = setup
    cp = _weights(s["ref"], s["g"])
    s["ctx"].diff_store_fields(0)
    try:
        g1, z1, m1, p1 = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, 0.02, 1.5, cp, NB)
        g1 = g1.clone()
        s["ctx"].diff_store_fields(1)
        g2, z2, m2, p2 = s["ctx"].diff_backward(s["gd"], s["ed"], ZETA, 0.02, 1.5, cp, NB)
    finally:
        s["ctx"].diff_store_fields(2)
    rel = (torch.linalg.vector_norm(g2.double() - g1.double()) /
           torch.linalg.vector_norm(g1.double())).item()
    print(f"\nstored fields: |dg|/|g| {rel:.2e}, dL/dzeta {z2:.8e} vs {z1:.8e}, "
          f"<x~> {m2:.8f} vs {m1:.8f}", flush=True)
    assert rel < 1e-3
    assert abs(z2 - z1) <= 2e-2 * abs(z1) + 1e-5 and abs(m2 - m1) < 5e-5
    assert np.allclose(p1, p2, rtol=2e-4, atol=0)


def test_forward_kept_fields_give_the_same_gradient(setup):
    """Mode 3: the exact forward keeps the filtered pairs and the next
    backward on the same inputs reads them instead of recomputing its first
    sweep.  The forward's centre-test pass and the backward's phi pass run
    the same line transform in different kernels, whose FP32 contractions
    (fused multiply-adds) may differ in the last bit.  The relaxed model
    amplifies last-bit field differences in cells whose rec_R is near the
    floor rho (s_R = (I_R - R_R)/rho, divided by tau), so the two paths are
    held to the bars of the FP64 comparison (relative L2 1e-3 on grad g,
    2e-2 on dL/dzeta, 5e-5 on <x~>, 2e-4 on P_b); the forward's own
    outputs do not change.  An input changed in place between forward and
    backward is caught by the fingerprints and the backward recomputes (as
    mode 1 does), within the same bars of mode 0 on the changed input."""
    s = setup
    ctx = s["ctx"]
    cp = _weights(s["ref"], s["g"])
    gx = s["gd"].clone()
    g_changed = s["gd"] * 1.01
    ctx.diff_store_fields(0)
    try:
        x0, mx0, pk0 = ctx.diff_forward(s["gd"], s["ed"], ZETA, NB)
        x0 = x0.clone()
        g0, z0, m0, p0 = ctx.diff_backward(s["gd"], s["ed"], ZETA, 0.02, 1.5, cp, NB)
        g0 = g0.clone()
        gc, zc, mc, _ = ctx.diff_backward(g_changed, s["ed"], ZETA, 0.02, 1.5, cp, NB)
        gc = gc.clone()
        ctx.diff_store_fields(3)
        x1, mx1, pk1 = ctx.diff_forward(gx, s["ed"], ZETA, NB)
        assert torch.equal(x0, x1) and mx0 == mx1
        assert np.allclose(pk0, pk1, rtol=1e-12, atol=0)  # FP64 LDS atomics: last bits vary
        g1, z1, m1, p1 = ctx.diff_backward(gx, s["ed"], ZETA, 0.02, 1.5, cp, NB)
        g1 = g1.clone()
        assert ctx.diff_last_kept
        rel = (torch.linalg.vector_norm(g1.double() - g0.double()) /
               torch.linalg.vector_norm(g0.double())).item()
        print(f"\nforward-kept fields: |dg|/|g| {rel:.2e}, dL/dzeta {z1:.8e} vs {z0:.8e}, "
              f"<x~> {m1:.8f} vs {m0:.8f}", flush=True)
        # the bars of the FP64 comparison (test_adjoint_matches_fp64_reference):
        # both paths are FP32 renderings of the same fields
        assert rel < 1e-3
        assert abs(z1 - z0) <= 2e-2 * abs(z0) + 1e-5 and abs(m1 - m0) < 5e-5
        assert np.allclose(p0, p1, rtol=2e-4, atol=0)
        # a second backward on the same inputs (another tau) reuses them too
        ctx.diff_backward(gx, s["ed"], ZETA, 0.05, 1.5, cp, NB)
        assert ctx.diff_last_kept
        # in-place change after the forward: detected, recomputed
        gx.mul_(1.01)
        g2, z2, m2, _ = ctx.diff_backward(gx, s["ed"], ZETA, 0.02, 1.5, cp, NB)
        g2 = g2.clone()
        assert not ctx.diff_last_kept
        rel2 = (torch.linalg.vector_norm(g2.double() - gc.double()) /
                torch.linalg.vector_norm(gc.double())).item()
        assert rel2 < 1e-3 and abs(z2 - zc) <= 2e-2 * abs(zc) + 1e-5 and abs(m2 - mc) < 5e-5
        # the recomputing backward left its input's pairs: the next one keeps
        ctx.diff_backward(gx, s["ed"], ZETA, 0.02, 1.5, cp, NB)
        assert ctx.diff_last_kept
        # another zeta: recomputed
        ctx.diff_backward(gx, s["ed"], ZETA * 1.1, 0.02, 1.5, cp, NB)
        assert not ctx.diff_last_kept
        ctx.diff_backward(gx, s["ed"], ZETA * 1.1, 0.02, 1.5, cp, NB)
        assert ctx.diff_last_kept
        # two cells 8191 apart swapped in place (a change that equal-weight
        # sums of the values would miss): detected by the position-mixed hash
        a, b = 5, 5 + 8191
        assert float(gx[a]) != float(gx[b])
        va, vb = gx[a].clone(), gx[b].clone()
        gx[a], gx[b] = vb, va
        torch.cuda.synchronize()
        ctx.diff_backward(gx, s["ed"], ZETA * 1.1, 0.02, 1.5, cp, NB)
        assert not ctx.diff_last_kept
    finally:
        ctx.diff_store_fields(2)
