"""Helpers for the GPU parity tests (imported by tests only)."""
import numpy as np

import hddm_lb

TOL = {hddm_lb.FP64: 1e-12, hddm_lb.F16X2: 1e-6, hddm_lb.BF16X3: 1e-6, hddm_lb.F32: 1e-6}
ALL_PRECISIONS = [hddm_lb.FP64, hddm_lb.F16X2, hddm_lb.BF16X3, hddm_lb.F32]


def normwise_rel_err(y, yref):
    """max over dof columns of max_p |y - yref| / max_p |yref| (columns with yref == 0 use abs error)."""
    y = np.asarray(y, np.float64)
    yref = np.asarray(yref, np.float64)
    den = np.abs(yref).max(axis=0)
    num = np.abs(y - yref).max(axis=0)
    den = np.where(den > 0, den, 1.0)
    return float((num / den).max()) if y.size else 0.0


def run_forward(ctx, X, sid=None, precision=hddm_lb.FP64, ldy=None, fill=np.nan):
    import torch
    n = X.shape[0]
    dofs = max(v[2] for v in ctx.states.values())
    ldy = ldy or dofs
    x = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    y = torch.full((n, ldy), fill, dtype=torch.float64, device="cuda")
    s = None if sid is None else torch.from_numpy(np.ascontiguousarray(sid, np.int32)).cuda()
    ctx.forward(x, y, s, precision)
    torch.cuda.synchronize()
    return y.cpu().numpy()
