"""3D transforms of the C-ABI (sf21_fft_*) against numpy's FP64 FFT.

These replace fftwf_plan_dft_{r2c,c2r}_3d + fftwf_execute
(get_HIIbubbles.c:110-183, :290-291, :324-325): unnormalised, forward sign
-1, half layout n x n x (n/2+1).  Tolerance: max |error| / max |reference|
below 2e-6 (FP32 transform, about 10-20 ulp of the largest value)."""
import numpy as np
import pytest

import simfast21_amd as sf

SIZES = [64, 96, 128, 192, 256]


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_r2c_c2r_vs_numpy(torch_cuda, n):
    torch = torch_cuda
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    spec = torch.empty((n, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    sf.FftPlan(n, sf.FFT_R2C).execute(xd, spec)
    torch.cuda.synchronize()
    ref = np.fft.rfftn(x.astype(np.float64))
    got = spec.cpu().numpy()
    assert abs(got - ref).max() / abs(ref).max() < 2e-6
    back = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    sf.FftPlan(n, sf.FFT_C2R).execute(spec, back)
    torch.cuda.synchronize()
    rb = back.cpu().numpy() / n ** 3
    assert abs(rb - x).max() / abs(x).max() < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 64, 96, 128, 256])
def test_c2c_vs_numpy(torch_cuda, n):
    torch = torch_cuda
    rng = np.random.default_rng(n + 1)
    z = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))).astype(np.complex64)
    zd = torch.from_numpy(z).cuda()
    out = torch.empty_like(zd)
    torch.cuda.synchronize()
    sf.FftPlan(n, sf.FFT_C2C_FORWARD).execute(zd, out)
    torch.cuda.synchronize()
    ref = np.fft.fftn(z.astype(np.complex128))
    assert abs(out.cpu().numpy() - ref).max() / abs(ref).max() < 2e-6
    sf.FftPlan(n, sf.FFT_C2C_BACKWARD).execute(out, out)
    torch.cuda.synchronize()
    assert abs(out.cpu().numpy() / n ** 3 - z).max() / abs(z).max() < 2e-6


@pytest.mark.gpu
def test_r2c_delta_and_constant(torch_cuda):
    """Closed forms: a unit impulse at the origin transforms to all ones; a
    constant c to c*n^3 at k=0 and zero elsewhere."""
    torch = torch_cuda
    n = 64
    x = np.zeros((n, n, n), np.float32)
    x[0, 0, 0] = 1.0
    spec = torch.empty((n, n, n // 2 + 1), dtype=torch.complex64, device="cuda")
    xd = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    sf.FftPlan(n, sf.FFT_R2C).execute(xd, spec)
    torch.cuda.synchronize()
    assert np.allclose(spec.cpu().numpy(), 1.0, atol=1e-6)
    xd = torch.full((n, n, n), 2.5, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    sf.FftPlan(n, sf.FFT_R2C).execute(xd, spec)
    torch.cuda.synchronize()
    s = spec.cpu().numpy()
    assert abs(s[0, 0, 0] - 2.5 * n ** 3) < 1e-3 * n ** 3
    s[0, 0, 0] = 0
    assert abs(s).max() < 1e-6 * 2.5 * n ** 3
