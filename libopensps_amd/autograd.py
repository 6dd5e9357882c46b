"""A torch.autograd.Function over the batch and its reverse pass (the thin PyTorch-ROCm harness of the gradients).
forward() runs the batch (opensps_hip_csp_batch) and keeps its inputs; backward() runs opensps_hip_csp_batch_grad
with the incoming cotangent of the magnitudes (its forward half recomputes the batch's age sums, as the reverse pass
recomputes the transmissions rather than storing them).

    mags = CSPMags.apply(dctx, prm, weights_z, precision)
    loss = ((mags - obs) ** 2).sum(); loss.backward()   # prm.grad (NPAR, n) in the parameter slots, weights_z.grad

With weights None the batch forms the SFH weights from the SFH parameters (sfh 0, 1, 4, 5) and prm.grad also holds
their rows (tage, tau, const, fburst, tburst, sf_start, sf_trunc, sf_slope: the SFH tangents).

The device buffers stay torch tensors on the GPU; the kernels are libopensps_hip's (no torch arithmetic on the
path). prm.grad holds the rows of params_bar that the batch computes, mapped to the parameter slots
(dust1, dust2, dust3, dust_index, dust1_index, frac_obrun, frac_nodust, logzsol, zred, sigma_smooth, igm_factor,
and the SFH parameters without given weights); other slots get 0.
"""
from __future__ import annotations

import torch

from . import hip as H
from .params import NPAR, SLOT

# params_bar rows (the reference's BGRAD_* - 1) -> parameter slots
ROW_SLOT = {1: "dust1", 2: "dust2", 3: "dust3", 4: "dust_index", 5: "dust1_index", 6: "frac_obrun",
            7: "frac_nodust", 15: "logzsol", 22: "zred", 23: "sigma_smooth", 32: "igm_factor",
            16: "tage", 17: "tau", 18: "const", 19: "fburst", 20: "tburst", 21: "sf_start", 24: "sf_trunc",
            25: "sf_slope"}


class CSPMags(torch.autograd.Function):
    """mags (n, nbands) of the batch as a function of prm (NPAR, n) and of the weights over ages and metallicities
    (n, nz, nt), the weights over ages (n, nt) or (None) the SFH parameters."""

    @staticmethod
    def forward(ctx, dctx: H.DeviceContext, prm: torch.Tensor, weights, precision: int = H.MIXED):
        prm = prm.detach().contiguous()
        weights = weights.detach().contiguous() if weights is not None else None
        n = prm.shape[1]
        over_z = weights is not None and weights.dim() == 3
        out = dctx.alloc_outputs(n, precision, want=("mags",))
        dctx.launch(prm, out, precision, split_parts=1, ssp_weights=None if over_z else weights,
                    ssp_weights_z=weights if over_z else None)
        dctx.sync()
        ctx.dctx, ctx.precision, ctx.over_z = dctx, precision, over_z
        ctx.has_w = weights is not None
        ctx.save_for_backward(prm, weights if weights is not None else prm.new_zeros(0))
        return out["mags"]

    @staticmethod
    def backward(ctx, gmags: torch.Tensor):
        prm, weights = ctx.saved_tensors
        if not ctx.has_w:
            weights = None
        n = prm.shape[1]
        g = ctx.dctx.grid
        gm = torch.nan_to_num(gmags.detach().to(torch.float64).contiguous(), nan=0.0)
        out = ctx.dctx.alloc_outputs(n, ctx.precision, want=("mags",))
        grads = {"params_bar": torch.zeros((n, H.NBGRAD), dtype=torch.float64, device=prm.device)}
        if ctx.over_z:
            grads["weights_z_bar"] = torch.zeros((n, g.nz, g.nt), dtype=torch.float64, device=prm.device)
        else:
            grads["weights_bar"] = torch.zeros((n, g.nt), dtype=torch.float64, device=prm.device)
        ctx.dctx.launch_grad(prm, out, {"mags": gm}, grads, ctx.precision, split_parts=1,
                             ssp_weights=None if ctx.over_z else weights,
                             ssp_weights_z=weights if ctx.over_z else None)
        ctx.dctx.sync()
        gprm = torch.zeros((NPAR, n), dtype=torch.float64, device=prm.device)
        for r, name in ROW_SLOT.items():
            gprm[SLOT[name]] = grads["params_bar"][:, r]
        gw = (grads["weights_z_bar"] if ctx.over_z else grads["weights_bar"]) if ctx.has_w else None
        return None, gprm, gw, None
