"""torch.autograd.Function around the differentiable mode of the C-ABI
(sf21_hii_diff_forward_dev / sf21_hii_diff_backward_dev).

Test harness only: the product exposes the adjoint through the C-ABI; this
wrapper lets the checks drive it through torch autograd (loss.backward()),
as a training loop would.

BubbleFinderObservables.apply(g, zeta, hii, emis, tau, nbins, mode) maps a
device box g (float32, n^3) and a scalar zeta (0-d float64 tensor) to
(<x>, P_b) of the bubble finder on rec = Delta^2, ion = zeta emis with
Delta = exp(g)/<exp(g)>:
  * mode "exact": the forward values are the exact finder's
    (sf21_hii_diff_forward_dev); the backward is the tau-relaxed adjoint --
    the pairing the differentiable mode is built for;
  * mode "relaxed": the forward values are the relaxed model's own (the
    by-products of sf21_hii_diff_backward_dev), so that forward and backward
    belong to one smooth function, as finite differences need.
The backward passes dL/d<x> and dL/dP_b to sf21_hii_diff_backward_dev and
returns its gradients in g (device) and zeta.
"""
import numpy as np
import torch


class BubbleFinderObservables(torch.autograd.Function):
    @staticmethod
    def forward(fctx, g, zeta, hii, emis, tau, nbins, mode):
        z = float(zeta)
        if mode == "exact":
            _, mx, pk = hii.diff_forward(g.detach(), emis, z, nbins)
        elif mode == "relaxed":
            _, _, mx, pk = hii.diff_backward(g.detach(), emis, z, tau, 0.0, None, nbins)
        else:
            raise ValueError(mode)
        fctx.save_for_backward(g)
        fctx.hii, fctx.emis, fctx.z, fctx.tau, fctx.nbins = hii, emis, z, tau, nbins
        return (torch.tensor(mx, dtype=torch.float64),
                torch.from_numpy(np.array(pk, np.float64)))

    @staticmethod
    def backward(fctx, d_mx, d_pk):
        (g,) = fctx.saved_tensors
        grad_g, gz, _, _ = fctx.hii.diff_backward(
            g.detach(), fctx.emis, fctx.z, fctx.tau, float(d_mx),
            d_pk.detach().to(torch.float64).cpu().numpy(), fctx.nbins)
        return grad_g, torch.tensor(gz, dtype=torch.float64), None, None, None, None, None


def observables(hii, emis, g, zeta, tau, nbins, mode="exact"):
    return BubbleFinderObservables.apply(g, zeta, hii, emis, tau, nbins, mode)
