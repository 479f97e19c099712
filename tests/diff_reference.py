"""FP64 PyTorch reference of the differentiable bubble finder (test
infrastructure: the checker of the GPU adjoint; never part of the product).

Model (one slice, direct inputs):
    Delta   = exp(g) / <exp(g)>                      (1 + delta, lognormal map)
    rec     = Delta^2,  ion = zeta * E               (E: fixed emissivity pattern)
    rec_R, ion_R = W_R * rec, W_R * ion              (k-space top hat, get_HIIbubbles.c:305-325)
    s_R     = min(ion_R - rec_R, ion_R) / max(rec_R, rho)
                                                     (criterion margin, :339-349: its sign is
                                                      the exact test ion_R > 0 and ion_R >= rec_R;
                                                      it is the ratio test ion_R/rec_R - 1
                                                      wherever rec_R >= rho = 1e-3 <rec>;
                                                      continuous, no blow-up)
    q       = min(ion/rec, 1) (0 if ion <= 0)        (quick fill, :281-288)
exact forward:   x = 1 where a ball B_R around a centre (s_R >= 0) covers the
                 cell, else q                       (:358-393, :428-458)
relaxed backward (temperature tau):
    phi_R   = softplus(u) - softplus(u - c), u = s_R / tau, c = PHI_CAP = 20
                                                     (a softplus bounded by c: one centre
                                                      covering a cell already gives
                                                      1 - exp(-c) = 1 - 2e-9; the bound keeps
                                                      the FP32 ball convolution on the GPU
                                                      accurate where A is small)
    A       = sum_R B_R (*) phi_R                    ((*) periodic convolution,
                                                       B_R the exact ball |d|^2 <= d2max(R))
    D       = 1 - exp(-A)                            (soft union of the balls)
    x~      = q + (1 - q) D                          (-> the exact x as tau -> 0)
observables: X = <x>,  T = (1 - x) Delta,  P_b = binned power spectrum of T
(tools/power3d.c normalisation, nbins linear bins of |k|/dk over [1, n/2]).
The reference evaluates everything with x~ (the relaxed model) so that
autograd gives the gradient the GPU adjoint computes.
"""
import copy
import math

import numpy as np
import torch

PI = 3.1415926535897932385
PHI_CAP = 20.0


def d2max(R, dx):
    m = int(math.floor((R / dx) ** 2))
    while m > 0 and not (math.sqrt(m) * dx <= R):
        m -= 1
    while math.sqrt(m + 1) * dx <= R:
        m += 1
    return m


def w_filter(x):
    """auxiliary.c:350-359 in FP64 (numpy)."""
    x = np.asarray(x, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        big = -(np.cos(x) - np.sin(x) / x) / x / x
    return 3.0 * np.where(x < 0.03, 1.0 / 3.0 - x * x / 30.0, big)


def freqs(n):
    """Integer DFT frequencies in numpy's order (fftfreq(n) * n rounds off
    the integers at sides that are not powers of two, which moves cells with
    |d|^2 exactly d2max out of the ball)."""
    return np.rint(np.fft.fftfreq(n) * n).astype(np.int64)


def k2_full(n):
    f = freqs(n)
    return f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2


def ball_hat(n, d2):
    f = freqs(n)
    a = np.minimum(np.abs(f), n - np.abs(f))
    m = a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2
    ball = (m <= d2).astype(np.float64)
    return np.fft.fftn(ball).real


def pk_bins(n, nbins):
    kx = freqs(n)
    kq = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + np.arange(n // 2 + 1)[None, None, :] ** 2)
    b = np.full(kq.shape, -1, np.int64)
    sel = (kq >= 1) & (kq <= n // 2)
    b[sel] = np.minimum(np.floor((kq[sel] - 1) / ((n // 2 - 1) / nbins)).astype(np.int64), nbins - 1)
    return b


def power_spectrum(T, side, nbins):
    n = T.shape[0]
    F = torch.fft.rfftn(T)
    b = torch.from_numpy(pk_bins(n, nbins))
    sel = b >= 0
    p2 = (F.real ** 2 + F.imag ** 2)[sel]
    bb = b[sel]
    sums = torch.zeros(nbins, dtype=T.dtype).index_add(0, bb, p2)
    cnt = torch.bincount(bb, minlength=nbins).to(T.dtype)
    return sums / cnt * ((side / n) ** 3 / n ** 3)


class Setup:
    def __init__(self, n, side, radii, E):
        self.n, self.side, self.radii = n, side, list(radii)
        dx = side / n
        dk = 2 * PI / side
        k2 = k2_full(n)
        self.W = [torch.from_numpy(w_filter(dk * np.sqrt(k2) * R)) for R in radii]
        self.B = [torch.from_numpy(ball_hat(n, d2max(R, dx))) for R in radii]
        self.E = torch.as_tensor(np.asarray(E, np.float64).reshape(n, n, n))


def relaxed(setup, g, zeta, tau, nbins=20, exact_x=None):
    """Returns (X, P, x~) of the relaxed model (differentiable in g, zeta)."""
    e = torch.exp(g)
    delta1 = e / e.mean()
    rec = delta1 ** 2
    ion = zeta * setup.E
    Z = torch.fft.fftn(torch.complex(rec, ion))
    A = torch.zeros_like(rec)
    rho = 1e-3 * rec.mean()
    for W, B in zip(setup.W, setup.B):
        zr = torch.fft.ifftn(W * Z)
        rec_r, ion_r = zr.real, zr.imag
        s = torch.minimum(ion_r - rec_r, ion_r) / torch.maximum(rec_r, rho)
        u = s / tau
        phi = torch.nn.functional.softplus(u) - torch.nn.functional.softplus(u - PHI_CAP)
        A = A + torch.fft.ifftn(B * torch.fft.fftn(phi)).real
    D = 1.0 - torch.exp(-A)
    pos = (ion > 0) & (rec > 0)
    q = torch.where(pos, torch.minimum(ion / torch.where(rec > 0, rec, torch.ones_like(rec)),
                                       torch.ones_like(rec)),
                    torch.where(ion > 0, torch.ones_like(rec), torch.zeros_like(rec)))
    x = q + (1.0 - q) * D
    X = x.mean()
    T = (1.0 - x) * delta1
    P = power_spectrum(T, setup.side, nbins)
    return X, P, x


def single(setup):
    """The same setup with float32 tensors: with float32 inputs, relaxed()
    and loss_and_grad() evaluate the model in single precision (the error an
    FP32 evaluation of it carries)."""
    s = copy.copy(setup)
    s.W = [w.float() for w in setup.W]
    s.B = [b.float() for b in setup.B]
    s.E = setup.E.float()
    return s


def loss_and_grad(setup, g_np, zeta, tau, cx, cp, nbins=20, dtype=torch.float64):
    """L = cx * X + sum_b cp[b] * P_b of the relaxed model, its gradient in
    g (numpy box) and zeta (float), autograd in dtype (FP64 by default)."""
    g = torch.tensor(np.asarray(g_np, np.float64).reshape(setup.n, setup.n, setup.n), dtype=dtype,
                     requires_grad=True)
    z = torch.tensor(float(zeta), dtype=dtype, requires_grad=True)
    X, P, _ = relaxed(setup, g, z, tau, nbins)
    L = cx * X + (torch.as_tensor(np.asarray(cp, np.float64)) * P).sum()
    L.backward()
    return float(L.detach()), g.grad.numpy().ravel(), float(z.grad)
