"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline. The
product (hddm-solver-postprocessors_amd) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("liboracle.so is not built (%s); run make -C oracle" % LIB_PATH)
lib = ctypes.CDLL(LIB_PATH)
_vp, _i32, _i64, _dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
lib.oracle_phi.restype = _dbl
lib.oracle_phi.argtypes = [_dbl, _i32, _i32]
lib.oracle_forward_point.restype = None
lib.oracle_forward_point.argtypes = [_i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64, _vp, _vp]
lib.oracle_forward_batch.restype = None
lib.oracle_forward_batch.argtypes = [_i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i32]
lib.oracle_forward_multistate.restype = None
lib.oracle_forward_multistate.argtypes = [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
lib.oracle_adjoint_batch.restype = None
lib.oracle_adjoint_batch.argtypes = [_i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64,
                                     _vp, _i64, _vp, _i64, _i32]
lib.oracle_x86_build.restype = _vp
lib.oracle_x86_build.argtypes = [_i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64]
lib.oracle_x86_free.restype = None
lib.oracle_x86_free.argtypes = [_vp]
lib.oracle_x86_nfreqs.restype = _i32
lib.oracle_x86_nfreqs.argtypes = [_vp]
lib.oracle_x86_nxps.restype = _i32
lib.oracle_x86_nxps.argtypes = [_vp]
lib.oracle_x86_forward_batch.restype = None
lib.oracle_x86_forward_batch.argtypes = [_vp, _i64, _vp, _i64, _vp, _i64, _i32]


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def phi(x, i, j):
    return lib.oracle_phi(float(x), int(i), int(j))


def forward(index_i, index_j, surplus, x, nthreads=0):
    """Gold forward of every row of x against one state: y [n, dofs] FP64."""
    ii, jj, ss, xx = _c(index_i, np.int32), _c(index_j, np.int32), _c(surplus, np.float64), _c(x, np.float64)
    nno, dim = ii.shape
    dofs = ss.shape[1]
    xx = xx.reshape(-1, dim)
    y = np.zeros((xx.shape[0], dofs), np.float64)
    lib.oracle_forward_batch(dim, nno, dofs, ii.ctypes.data, jj.ctypes.data, dim, ss.ctypes.data, dofs,
                             xx.shape[0], xx.ctypes.data, dim, y.ctypes.data, dofs, nthreads)
    return y


def forward_point(index_i, index_j, surplus, x):
    ii, jj, ss, xx = _c(index_i, np.int32), _c(index_j, np.int32), _c(surplus, np.float64), _c(x, np.float64)
    nno, dim = ii.shape
    y = np.zeros(ss.shape[1], np.float64)
    lib.oracle_forward_point(dim, nno, ss.shape[1], ii.ctypes.data, jj.ctypes.data, dim, ss.ctypes.data,
                             ss.shape[1], xx.ctypes.data, y.ctypes.data)
    return y


def forward_multistate(states, xs):
    """states: list of (index_i, index_j, surplus); xs: list of one point per state."""
    count = len(states)
    keep = []
    arr = lambda vals, t: (t * count)(*vals)
    ii = [_c(s[0], np.int32) for s in states]
    jj = [_c(s[1], np.int32) for s in states]
    ss = [_c(s[2], np.float64) for s in states]
    xx = [_c(x, np.float64) for x in xs]
    ys = [np.zeros(s.shape[1], np.float64) for s in ss]
    keep += [ii, jj, ss, xx, ys]
    dim = ii[0].shape[1]
    nno = arr([a.shape[0] for a in ii], ctypes.c_int)
    dofs = arr([a.shape[1] for a in ss], ctypes.c_int)
    pi = arr([a.ctypes.data for a in ii], ctypes.c_void_p)
    pj = arr([a.ctypes.data for a in jj], ctypes.c_void_p)
    ldi = arr([dim] * count, ctypes.c_int64)
    ps = arr([a.ctypes.data for a in ss], ctypes.c_void_p)
    lds = arr([a.shape[1] for a in ss], ctypes.c_int64)
    px = arr([a.ctypes.data for a in xx], ctypes.c_void_p)
    py = arr([a.ctypes.data for a in ys], ctypes.c_void_p)
    lib.oracle_forward_multistate(dim, count, ctypes.addressof(nno), ctypes.addressof(dofs), ctypes.addressof(pi),
                                  ctypes.addressof(pj), ctypes.addressof(ldi), ctypes.addressof(ps),
                                  ctypes.addressof(lds), ctypes.addressof(px), ctypes.addressof(py))
    return ys


def adjoint(index_i, index_j, surplus, x, gbar, nthreads=0):
    """FP64 analytic adjoint: (dx [n, dim], dS [nno, dofs])."""
    ii, jj, ss = _c(index_i, np.int32), _c(index_j, np.int32), _c(surplus, np.float64)
    xx, gg = _c(x, np.float64), _c(gbar, np.float64)
    nno, dim = ii.shape
    dofs = ss.shape[1]
    n = xx.shape[0]
    dx = np.zeros((n, dim), np.float64)
    ds = np.zeros((nno, dofs), np.float64)
    lib.oracle_adjoint_batch(dim, nno, dofs, ii.ctypes.data, jj.ctypes.data, dim, ss.ctypes.data, dofs, n,
                             xx.ctypes.data, dim, gg.ctypes.data, dofs, dx.ctypes.data, dim, ds.ctypes.data, dofs,
                             nthreads)
    return dx, ds


class X86State:
    """The reference's xps/chains CPU algorithm (CPU baseline)."""

    def __init__(self, index_i, index_j, surplus):
        self.ii, self.jj, self.ss = _c(index_i, np.int32), _c(index_j, np.int32), _c(surplus, np.float64)
        nno, dim = self.ii.shape
        self.dim, self.dofs = dim, self.ss.shape[1]
        self._h = lib.oracle_x86_build(dim, nno, self.dofs, self.ii.ctypes.data, self.jj.ctypes.data, dim,
                                       self.ss.ctypes.data, self.dofs)
        self.nfreqs = lib.oracle_x86_nfreqs(self._h)
        self.nxps = lib.oracle_x86_nxps(self._h)

    def forward(self, x, nthreads=0):
        xx = _c(x, np.float64).reshape(-1, self.dim)
        y = np.zeros((xx.shape[0], self.dofs), np.float64)
        lib.oracle_x86_forward_batch(self._h, xx.shape[0], xx.ctypes.data, self.dim, y.ctypes.data, self.dofs,
                                     nthreads)
        return y

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            lib.oracle_x86_free(self._h)
            self._h = None
