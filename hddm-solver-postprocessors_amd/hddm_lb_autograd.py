"""torch.autograd.Function over the engine's C ABI (forward + adjoint).

interpolate(ctx, x, surplus, istate) returns y = sum_k B_k(x) S_k for the
state uploaded as `istate` and makes x and surplus differentiable: backward
calls hddm_lb_adjoint, which recomputes the basis (nothing is stored between
forward and backward). `surplus` must hold the values uploaded for the state
(the engine reads its own copy); it is the tensor that receives dL/dS.
"""
import torch

import hddm_lb


class _LinearBasis(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, surplus, hctx, istate, precision):
        n, dofs = x.shape[0], surplus.shape[1]
        y = torch.empty((n, dofs), dtype=torch.float64, device=x.device)
        sid = None
        if istate != 0:
            sid = torch.full((n,), istate, dtype=torch.int32, device=x.device)
        hctx.forward(x, y, sid, precision)
        fctx.save_for_backward(x)
        fctx.hctx, fctx.istate, fctx.sid, fctx.shape = hctx, istate, sid, tuple(surplus.shape)
        return y

    @staticmethod
    def backward(fctx, gy):
        (x,) = fctx.saved_tensors
        hctx = fctx.hctx
        need_x, need_s = fctx.needs_input_grad[0], fctx.needs_input_grad[1]
        g = gy.contiguous().to(torch.float64)
        dx = torch.empty_like(x) if need_x else None
        ds_full = None
        if need_s:
            off, total = hctx.surplus_offset(fctx.istate)
            ds_full = torch.empty((max(total, 1),), dtype=torch.float64, device=x.device)
        hctx.adjoint(x, g, dx, ds_full, fctx.sid, hddm_lb.FP64)
        ds = None
        if need_s:
            nno, dofs = fctx.shape
            ds = ds_full[off:off + nno * dofs].view(nno, dofs)
        return dx, ds, None, None, None


def interpolate(hctx, x, surplus, istate=0, precision=hddm_lb.FP64):
    """Differentiable forward of state `istate` of engine context `hctx`.

    x: [n, dim] FP64 on the context's device; surplus: [nno, dofs] FP64 (the
    uploaded values, in upload order). The backward is the FP64 adjoint."""
    return _LinearBasis.apply(x, surplus, hctx, istate, precision)
