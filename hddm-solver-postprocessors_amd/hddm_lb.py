"""Python binding of the MI355X LinearBasis engine (ctypes over include/hddm_lb.h).

Device buffers come from torch (ROCm build): torch is plumbing for device
memory and streams here; all arithmetic runs in the gfx950 kernels of
lib/libhddm_lb.so. Importing this module fails loudly when that library is
missing: there is no CPU fallback.

Mirrors the reference's operator surface for the hot path:
  * Context.forward     <- Interpolator::interpolate (include/interpolator.h:31-37),
                           batched over points, one state id per point
  * Context.upload_state <- DataDense::load + device upload (LinearBasis/gold/src/Data.cpp:322-595,
                           LinearBasis/cuda/src/Data.cu:944-948)
  * read_plt / write_plt_* <- the .plt surplus file format
Errors raise HddmLbError carrying the library's status code and message,
where the reference aborts the process.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhddm_lb.so")

FP64, F16X2, BF16X3, F32 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "f16x2": F16X2, "bf16x3": BF16X3, "f32": F32}

OPT_AUTOTUNE, OPT_SEGMENT_NODES, OPT_WAVES_PER_WG, OPT_SORT_POINTS = 1, 2, 3, 4
OPT_TIMING, OPT_SLAB_TILES, OPT_STATS, OPT_WALK, OPT_FORMULATION, OPT_COALESCE = 5, 6, 8, 9, 10, 11
OPT_DEVICE_LAYOUT = 12
OPT_ADJOINT_OWNED = 13
# state arrays readable with Context.state_array (hddm_lb_state_array)
STATE_ARRAYS = ("ops", "skip", "order", "s64", "s16", "sbf", "s32", "s32r", "iscale", "lomask", "himask", "ftab")
STATUS_WAIT_LIMIT = 1

SUCCESS = 0
ERROR_NAMES = {1: "INVALID_VALUE", 2: "NO_DEVICE", 3: "OUT_OF_MEMORY", 4: "HIP",
               5: "STATE", 6: "IO", 7: "FORMAT", 8: "UNSUPPORTED"}


class HddmLbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hddm_lb error %d (%s): %s" % (code, ERROR_NAMES.get(code, "?"), msg))
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError("libhddm_lb.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)


def _load_hip_runtime():
    """Makes one HIP runtime global before the engine loads (the engine has no
    NEEDED entry for one): torch's bundled runtime when torch is importable, so
    engine kernels and torch tensors share a runtime; else the system ROCm."""
    try:
        import torch
        cand = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        if os.path.exists(cand):
            return ctypes.CDLL(cand, mode=ctypes.RTLD_GLOBAL)
    except ImportError:
        pass
    return ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)


HIP_RUNTIME = _load_hip_runtime()
lib = ctypes.CDLL(LIB_PATH)

_c = ctypes
_vp, _i32, _i64, _u64, _dbl = _c.c_void_p, _c.c_int, _c.c_int64, _c.c_uint64, _c.c_double
_I32P = _c.POINTER(_c.c_int32)
_DP = _c.POINTER(_c.c_double)

_SIGS = {
    "hddm_lb_device_count": (_i32, [_c.POINTER(_i32)]),
    "hddm_lb_create": (_i32, [_i32, _c.POINTER(_vp)]),
    "hddm_lb_destroy": (_i32, [_vp]),
    "hddm_lb_last_error": (_c.c_char_p, []),
    "hddm_lb_device_status": (_i32, [_vp, _c.POINTER(_c.c_uint32), _i32]),
    "hddm_lb_set_option": (_i32, [_vp, _i32, _i64]),
    "hddm_lb_upload_state": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64]),
    "hddm_lb_state_info": (_i32, [_vp, _i32, _c.POINTER(_i32), _c.POINTER(_i32), _c.POINTER(_i32), _c.POINTER(_i32)]),
    "hddm_lb_state_array": (_i32, [_vp, _i32, _i32, _vp, _c.POINTER(_c.c_uint64)]),
    "hddm_lb_forward": (_i32, [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _i32, _vp]),
    "hddm_lb_forward_host": (_i32, [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _i32]),
    "hddm_lb_forward_noalloc": (_i32, [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _i32, _vp]),
    "hddm_lb_adjoint": (_i32, [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i32, _vp]),
    "hddm_lb_surplus_offset": (_i32, [_vp, _i32, _c.POINTER(_i64), _c.POINTER(_i64)]),
    "hddm_lb_plan_create": (_i32, [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _i32, _vp, _c.POINTER(_vp)]),
    "hddm_lb_plan_launch": (_i32, [_vp, _vp]),
    "hddm_lb_plan_destroy": (_i32, [_vp]),
    "hddm_lb_last_sweep_ms": (_dbl, [_vp]),
    "hddm_lb_last_stats": (_i32, [_vp, _vp]),
    "hddm_lb_last_config": (_i32, [_vp, _vp]),
    "hddm_lb_last_adjoint_config": (_i32, [_vp, _vp]),
    "hddm_lb_host_call_stats": (_i32, [_vp, _c.POINTER(_i64), _c.POINTER(_i64)]),
    "hddm_lb_last_adjoint_stats": (_i32, [_vp, _vp, _i32]),
    "hddm_lb_last_stats_n": (_i32, [_vp, _vp, _i32]),
    "hddm_lb_sweep_time_summary": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "hddm_lb_kernel_time_summary": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "hddm_lb_plt_read": (_i32, [_c.c_char_p, _i32, _c.POINTER(_i32), _c.POINTER(_i32), _c.POINTER(_i32),
                                _c.POINTER(_i32), _c.POINTER(_I32P), _c.POINTER(_I32P), _c.POINTER(_DP)]),
    "hddm_lb_plt_write_compressed": (_i32, [_c.c_char_p, _i32, _i32, _i32, _i32, _vp, _vp, _vp]),
    "hddm_lb_plt_write_uncompressed": (_i32, [_c.c_char_p, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp]),
    "hddm_lb_free": (None, [_vp]),
    "hddm_lb_grid_generate": (_i32, [_i32, _i32, _i32, _dbl, _u64, _i64, _c.POINTER(_vp)]),
    "hddm_lb_grid_nno": (_i32, [_vp, _c.POINTER(_i64)]),
    "hddm_lb_grid_index": (_i32, [_vp, _vp, _vp]),
    "hddm_lb_grid_levels": (_i32, [_vp, _vp, _vp]),
    "hddm_lb_grid_free": (_i32, [_vp]),
    "hddm_lb_layout_build": (_i32, [_i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64, _c.POINTER(_vp)]),
    "hddm_lb_layout_info": (_i32, [_vp] + [_c.POINTER(_i32)] * 4),
    "hddm_lb_layout_ops": (_i32, [_vp, _vp, _vp, _vp]),
    "hddm_lb_layout_order": (_i32, [_vp, _vp]),
    "hddm_lb_layout_factors": (_i32, [_vp, _vp, _c.POINTER(_i32)]),
    "hddm_lb_layout_free": (_i32, [_vp]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error():
    return lib.hddm_lb_last_error().decode(errors="replace")


def _check(rc):
    if rc != SUCCESS:
        raise HddmLbError(rc, last_error())


def _ptr(a):
    """Address of a numpy array, a torch tensor, or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()


def device_count():
    n = _i32(0)
    _check(lib.hddm_lb_device_count(_c.byref(n)))
    return n.value


# ---------------------------------------------------------------- grids

def generate_grid(dim, start_sum, max_level, p_refine=0.3, seed=1, target_nno=0):
    """Seeded sparse grid; returns (index_i, index_j, levels, indices), each int32 [nno, dim]."""
    g = _vp()
    _check(lib.hddm_lb_grid_generate(dim, start_sum, max_level, float(p_refine), seed, int(target_nno), _c.byref(g)))
    try:
        n = _i64(0)
        _check(lib.hddm_lb_grid_nno(g, _c.byref(n)))
        shape = (n.value, dim)
        ii = np.zeros(shape, np.int32)
        jj = np.zeros(shape, np.int32)
        lv = np.zeros(shape, np.int32)
        ix = np.zeros(shape, np.int32)
        _check(lib.hddm_lb_grid_index(g, _ptr(ii), _ptr(jj)))
        _check(lib.hddm_lb_grid_levels(g, _ptr(lv), _ptr(ix)))
    finally:
        lib.hddm_lb_grid_free(g)
    return ii, jj, lv, ix


def synthetic_surplus(levels, dofs, seed):
    """S[k][c] = U[-1, 1] * 2^-(sum_d l_d), FP64, seeded (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    scale = np.ldexp(1.0, -levels.sum(axis=1).astype(np.int64))
    return (rng.uniform(-1.0, 1.0, size=(levels.shape[0], dofs)) * scale[:, None]).astype(np.float64)


def synthetic_points(n, dim, seed, lo=0.0, hi=1.0, kink_level=None, kink_gap=1e-3):
    """x ~ U[lo, hi]^dim; with kink_level L, every coordinate is moved at least
    kink_gap away from the multiples of 2^-(L-1) (the kinks of all hats up to level L)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, size=(n, dim))
    if kink_level is not None:
        h = 2.0 ** -(kink_level - 1)
        k = np.round(x / h) * h
        d = x - k
        close = np.abs(d) < kink_gap
        x = np.where(close, k + np.where(d >= 0, kink_gap, -kink_gap), x)
        x = np.clip(x, lo, hi)
    return np.ascontiguousarray(x)


# ---------------------------------------------------------------- .plt files

def read_plt(path, nagents):
    dim, nno, dofs, level = _i32(), _i32(), _i32(), _i32()
    pi, pj, ps = _I32P(), _I32P(), _DP()
    _check(lib.hddm_lb_plt_read(os.fsencode(path), nagents, _c.byref(dim), _c.byref(nno), _c.byref(dofs),
                                _c.byref(level), _c.byref(pi), _c.byref(pj), _c.byref(ps)))
    try:
        n, d, f = nno.value, dim.value, dofs.value
        ii = np.ctypeslib.as_array(pi, shape=(max(n * d, 1),))[: n * d].reshape(n, d).copy()
        jj = np.ctypeslib.as_array(pj, shape=(max(n * d, 1),))[: n * d].reshape(n, d).copy()
        ss = np.ctypeslib.as_array(ps, shape=(max(n * f, 1),))[: n * f].reshape(n, f).copy()
    finally:
        lib.hddm_lb_free(_c.cast(pi, _vp))
        lib.hddm_lb_free(_c.cast(pj, _vp))
        lib.hddm_lb_free(_c.cast(ps, _vp))
    return dict(dim=dim.value, nno=nno.value, dofs=dofs.value, level=level.value, index_i=ii, index_j=jj, surplus=ss)


def write_plt_compressed(path, index_i, index_j, surplus, level=0):
    ii = np.ascontiguousarray(index_i, np.int32)
    jj = np.ascontiguousarray(index_j, np.int32)
    ss = np.ascontiguousarray(surplus, np.float64)
    _check(lib.hddm_lb_plt_write_compressed(os.fsencode(path), ii.shape[1], ii.shape[0], ss.shape[1], level,
                                            _ptr(ii), _ptr(jj), _ptr(ss)))


def write_plt_uncompressed(path, levels, indices, surplus, binary=True, level=0):
    lv = np.ascontiguousarray(levels, np.int32)
    ix = np.ascontiguousarray(indices, np.int32)
    ss = np.ascontiguousarray(surplus, np.float64)
    _check(lib.hddm_lb_plt_write_uncompressed(os.fsencode(path), int(binary), lv.shape[1], lv.shape[0],
                                              ss.shape[1], level, _ptr(lv), _ptr(ix), _ptr(ss)))


# ---------------------------------------------------------------- layout

class Layout:
    """Host-side prefix-trie layout of a state (no device needed): nodes in
    trie order and the op stream of the walk (see include/hddm_lb.h)."""

    def __init__(self, index_i, index_j, surplus):
        ii = np.ascontiguousarray(index_i, np.int32)
        jj = np.ascontiguousarray(index_j, np.int32)
        ss = np.ascontiguousarray(surplus, np.float64)
        nno, dim = ii.shape
        self.dim, self.nno, self.dofs = dim, nno, ss.shape[1]
        h = _vp()
        _check(lib.hddm_lb_layout_build(dim, nno, self.dofs, _ptr(ii), _ptr(jj), dim, _ptr(ss), self.dofs, _c.byref(h)))
        self._h = h
        vals = [_i32() for _ in range(4)]
        _check(lib.hddm_lb_layout_info(h, *[_c.byref(v) for v in vals]))
        self.fmax, self.nops, self.nchunks, self.dofs_pad = [v.value for v in vals]
        ops = np.zeros((self.nops, 4), np.uint32)
        self.skip = np.zeros(self.nops, np.uint32)
        self.opstart = np.zeros(nno + 1, np.int32)
        _check(lib.hddm_lb_layout_ops(h, _ptr(ops), _ptr(self.skip), _ptr(self.opstart)))
        self.op_dim = ((ops[:, 0] & 0xff) >> 1).astype(np.int64)
        self.op_relskip = (ops[:, 0] >> 8).astype(np.int64)
        ctrl = ops[:, 1]
        self.op_depth = (ctrl & 15).astype(np.int64)
        self.op_last = (ctrl >> 4) & 1 == 1
        self.op_rel = ((ctrl >> 5) & 3).astype(np.int64)
        self.op_kids = (ctrl >> 7) & 1 == 1
        self.op_node = (ctrl >> 8).astype(np.int64)
        self.op_i = (ops[:, 2].astype(np.uint64) << np.uint64(32)).view(np.float64)
        self.op_negj = (ops[:, 3].astype(np.uint64) << np.uint64(32)).view(np.float64)
        self.order = np.zeros(nno, np.int32)
        _check(lib.hddm_lb_layout_order(h, _ptr(self.order)))
        fs = _i32()
        _check(lib.hddm_lb_layout_factors(h, None, _c.byref(fs)))
        self.fstride = fs.value
        ft = np.zeros(((nno + 31) // 32 * 32, self.fstride, 4), np.uint32)
        _check(lib.hddm_lb_layout_factors(h, _ptr(ft), None))
        self.ftab = ft

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:
            lib.hddm_lb_layout_free(h)
        self._h = None


# ---------------------------------------------------------------- context

def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


class Context:
    """One device, any number of states sharing dim."""

    def __init__(self, device=0):
        h = _vp()
        _check(lib.hddm_lb_create(device, _c.byref(h)))
        self._h = h
        self.device = device
        self.states = {}

    def close(self):
        if getattr(self, "_h", None) and lib is not None:
            lib.hddm_lb_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def set_option(self, option, value):
        _check(lib.hddm_lb_set_option(self._h, option, int(value)))

    def upload_state(self, istate, index_i, index_j, surplus):
        ii = np.ascontiguousarray(index_i, np.int32)
        jj = np.ascontiguousarray(index_j, np.int32)
        ss = np.ascontiguousarray(surplus, np.float64)
        nno, dim = ii.shape
        _check(lib.hddm_lb_upload_state(self._h, istate, dim, nno, ss.shape[1], _ptr(ii), _ptr(jj), dim,
                                        _ptr(ss), ss.shape[1]))
        self.states[istate] = (nno, dim, ss.shape[1])

    def state_info(self, istate):
        v = [_i32() for _ in range(4)]
        _check(lib.hddm_lb_state_info(self._h, istate, *[_c.byref(a) for a in v]))
        return dict(zip(("dim", "nno", "dofs", "max_factors"), [a.value for a in v]))

    def state_array(self, istate, name):
        """Bytes (uint8 array) of one device array of an uploaded state's layout."""
        which = STATE_ARRAYS.index(name)
        n = _c.c_uint64(0)
        _check(lib.hddm_lb_state_array(self._h, istate, which, None, _c.byref(n)))
        out = np.zeros(n.value, np.uint8)
        _check(lib.hddm_lb_state_array(self._h, istate, which, _ptr(out), _c.byref(n)))
        return out

    def forward(self, x, y, state_ids=None, precision=FP64, stream=None):
        """x: device FP64 [n, >=dim]; y: device FP64 [n, >=dofs]; state_ids: device int32 [n] or None."""
        n = x.shape[0]
        _check(lib.hddm_lb_forward(self._h, n, _ptr(x), x.stride(0), _ptr(state_ids), _ptr(y), y.stride(0),
                                   precision, _stream_ptr(stream)))
        return y

    def forward_host(self, x, y, state_ids=None, precision=FP64):
        """Host numpy arrays; synchronous."""
        x = np.ascontiguousarray(x, np.float64)
        sid = None if state_ids is None else np.ascontiguousarray(state_ids, np.int32)
        _check(lib.hddm_lb_forward_host(self._h, x.shape[0], _ptr(x), x.shape[1], _ptr(sid), _ptr(y), y.shape[1],
                                        precision))
        return y

    def adjoint(self, x, gbar, dx=None, dsurplus=None, state_ids=None, precision=FP64, stream=None):
        """x: device FP64 [n, >=dim]; gbar: device FP64 [n, >=dofs]; dx: device FP64 [n, >=dim] or None;
        dsurplus: device FP64 [total] (state s at surplus_offset(s)) or None. Both outputs are overwritten."""
        n = x.shape[0]
        _check(lib.hddm_lb_adjoint(self._h, n, _ptr(x), x.stride(0), _ptr(state_ids), _ptr(gbar), gbar.stride(0),
                                   _ptr(dx), dx.stride(0) if dx is not None else 0, _ptr(dsurplus), precision,
                                   _stream_ptr(stream)))

    def surplus_offset(self, istate):
        off, tot = _i64(), _i64()
        _check(lib.hddm_lb_surplus_offset(self._h, istate, _c.byref(off), _c.byref(tot)))
        return off.value, tot.value

    def last_sweep_ms(self):
        return lib.hddm_lb_last_sweep_ms(self._h)

    def sweep_time_summary(self, kind=0):
        """(calls, total ms, min ms, max ms) of the kernels of timed calls since the last call of
        this kind: 0 the forwards' sweep kernels, 1 the wave-specialized adjoint kernels."""
        n = ctypes.c_int64(0)
        tot, mn, mx = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        _check(lib.hddm_lb_kernel_time_summary(self._h, kind, ctypes.byref(n), ctypes.byref(tot), ctypes.byref(mn),
                                               ctypes.byref(mx)))
        return n.value, tot.value, mn.value, mx.value

    def last_adjoint_stats(self, n=9):
        """Counters of the last wave-specialized adjoint (statistics mode), see include/hddm_lb.h."""
        st = np.zeros(n, np.uint64)
        _check(lib.hddm_lb_last_adjoint_stats(self._h, _ptr(st), n))
        return tuple(int(v) for v in st)

    def host_call_stats(self):
        """(requests, batches) of host-pointer calls served by coalescing."""
        r, b = _i64(), _i64()
        _check(lib.hddm_lb_host_call_stats(self._h, _c.byref(r), _c.byref(b)))
        return r.value, b.value

    def last_config(self):
        """Sweep configuration of the last forward: dict(formulation, waves, tiles, tuned, cache_hits)."""
        v = np.zeros(5, np.int32)
        _check(lib.hddm_lb_last_config(self._h, _ptr(v)))
        return dict(zip(("formulation", "waves", "tiles", "tuned", "cache_hits"), (int(a) for a in v)))

    def last_adjoint_config(self):
        """Configuration of the last adjoint: dict(kernels (1 per-tile, 2 wave-specialized),
        owned_copies (0: dS by atomics), ds_copies (summed after), max_segments)."""
        v = np.zeros(4, np.int32)
        _check(lib.hddm_lb_last_adjoint_config(self._h, _ptr(v)))
        return dict(zip(("kernels", "owned_copies", "ds_copies", "max_segments"), (int(a) for a in v)))

    def device_status(self, clear=True):
        """Status bits raised by this context's kernels since the last clearing call (0: none)."""
        v = _c.c_uint32(0)
        _check(lib.hddm_lb_device_status(self._h, _c.byref(v), int(clear)))
        return v.value

    def last_stats(self, n=4):
        """(ops evaluated, subtrees pruned, chunks contracted, work items[, MFMA
        instructions issued[, non-zero (point, node) pairs]]) of the last
        forward: the first n (at most 6)."""
        st = np.zeros(n, np.uint64)
        _check(lib.hddm_lb_last_stats_n(self._h, _ptr(st), n))
        return tuple(int(v) for v in st)

    def plan(self, x, y, state_ids=None, precision=FP64, stream=None):
        return Plan(self, x, y, state_ids, precision, stream)


class Plan:
    """A forward captured into a hipGraph; replay with launch()."""

    def __init__(self, ctx, x, y, state_ids, precision, stream):
        h = _vp()
        _check(lib.hddm_lb_plan_create(ctx._h, x.shape[0], _ptr(x), x.stride(0), _ptr(state_ids), _ptr(y),
                                       y.stride(0), precision, _stream_ptr(stream), _c.byref(h)))
        self._h = h
        self._keep = (ctx, x, y, state_ids)

    def launch(self, stream=None):
        _check(lib.hddm_lb_plan_launch(self._h, _stream_ptr(stream)))

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            lib.hddm_lb_plan_destroy(self._h)
        self._h = None
