"""gnurbs -- Python binding of libgnurbs, matrix-form NURBS evaluation and inversion on MI355X.

The binding is a thin ctypes layer over the C ABI declared in include/gnurbs.h; every call goes
straight to the HIP kernels in libgnurbs.so (built in-tree next to this file).  torch supplies
device memory and the current HIP stream; no computation happens in Python and there is no CPU
fallback: if the library is missing or no GPU is visible, calls raise.

Method names follow the reference's entry points (Qiang-Zou/GPU-NURBS): ``readNu``
(src/main.cu:45-113), ``findSpan`` (src/NURBS/Evaluate.cu:504-533), ``GetMatrixN``
(src/NURBS/Evaluate.cu:3-88), ``EvaluateRandom``/``Evaluate`` (src/NURBS/Evaluate.cu:1023-1038 and
the derivative functions 535-930), ``BuildTree`` (src/kernel/kernel.cu:375-431, 544-621),
``NewInverse`` (src/kernel/kernel.cu:433-469 with GaussNewton 151-249).
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgnurbs.so")

GN_OK, GN_EINVAL, GN_ENOMEM, GN_EDEVICE, GN_EIO, GN_ESTATE, GN_EBUDGET, GN_ENOTFOUND, GN_EAGAIN = range(9)
GN_FP64, GN_FP32, GN_MIXED = 0, 1, 2
PT_CONVERGED, PT_CONVERGED_DIST, PT_MAXITER, PT_SINGULAR, PT_NONFINITE, PT_NOSEED = range(6)
SEED_NEAREST_LEAF, SEED_REF_FIRST_HIT, SEED_GIVEN, SEED_GREEDY_LEAF, SEED_GLOBAL, SEED_GREEDY_STEP = 0, 1, 2, 3, 4, 5
SEED_GRID_STEP = 6
TREE_REF_PARENT = 1
NEWTON_FP32, NEWTON_FP64 = 1, 2

_c_int, _c_double, _c_size, _c_void = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p


class InvertOpts(ctypes.Structure):
    """gn_invert_opts (include/gnurbs.h)."""

    _fields_ = [
        ("tol_dist", _c_double),
        ("tol_step", _c_double),
        ("max_iter", _c_int),
        ("clamp", _c_int),
        ("precision", _c_int),
        ("fp32_iters", _c_int),
        ("fp64_tail", _c_int),
        ("tol_step32", _c_double),
        ("seed_mode", _c_int),
        ("newton", _c_int),
        ("tol_quad", _c_double),
        ("fallback", _c_int),
    ]


class GnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gnurbs error {code}: {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every C ABI export this binding uses
_SIGS = [
    ("gn_version", ctypes.c_char_p, []),
    ("gn_last_error", ctypes.c_char_p, []),
    ("gn_invert_opts_default", None, [ctypes.POINTER(InvertOpts)]),
    ("gn_surface_create", _c_int, [_c_int, _c_int, _c_int, _c_int, _c_void, _c_void, _c_void, _c_void, _c_void, ctypes.POINTER(_c_void)]),
    ("gn_surface_from_nu", _c_int, [ctypes.c_char_p, _c_void, ctypes.POINTER(_c_void)]),
    ("gn_surface_to_nu", _c_int, [_c_void, ctypes.c_char_p]),
    ("gn_surface_destroy", _c_int, [_c_void]),
    ("gn_surface_info", _c_int, [_c_void, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    ("gn_surface_device_bytes", _c_size, [_c_void]),
    ("gn_find_span", _c_int, [_c_void, _c_int, _c_void, _c_size, _c_void, _c_void]),
    ("gn_build_span_table", _c_int, [_c_void, _c_void]),
    ("gn_get_span_matrix", _c_int, [_c_void, _c_int, _c_int, _c_void]),
    ("gn_evaluate", _c_int, [_c_void, _c_void, _c_void, _c_size, _c_int, _c_void, _c_void]),
    ("gn_evaluate_f32", _c_int, [_c_void, _c_void, _c_void, _c_size, _c_int, _c_void, _c_void]),
    ("gn_evaluate_sorted_f32", _c_int, [_c_void, _c_void, _c_void, _c_size, _c_int, _c_void, _c_void, _c_void]),
    ("gn_evaluate_grid_f32", _c_int, [_c_void, _c_void, _c_int, _c_void, _c_int, _c_int, _c_void, _c_void]),
    ("gn_build_tree", _c_int, [_c_void, _c_int, _c_int, _c_void]),
    ("gn_get_tree", _c_int, [_c_void, _c_void, _c_void, _c_void, _c_size]),
    ("gn_get_normal_cones", _c_int, [_c_void, _c_void, _c_size]),
    ("gn_build_seed_grid", _c_int, [_c_void, _c_int, _c_void]),
    ("gn_get_seed_grid", _c_int, [_c_void, _c_void, _c_size, _c_void, _c_void, _c_void]),
    ("gn_seed", _c_int, [_c_void, _c_int, _c_void, _c_void, _c_void, _c_size, _c_void, _c_void, _c_void, _c_void]),
    ("gn_invert", _c_int, [_c_void, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, _c_void]),
    ("gn_invert_noalloc", _c_int, [_c_void, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, _c_void]),
    ("gn_invert_iters", _c_int, [_c_void, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, _c_void, _c_void, _c_void]),
    ("gn_pcurve", _c_int, [_c_void, _c_void, _c_size, _c_int, _c_int, ctypes.POINTER(InvertOpts), _c_void, _c_void,
                           ctypes.POINTER(_c_double), ctypes.POINTER(_c_size), _c_void]),
    ("gn_batch_create", _c_int, [_c_void, _c_int, _c_void, ctypes.POINTER(_c_void)]),
    ("gn_batch_destroy", _c_int, [_c_void]),
    ("gn_batch_device_bytes", _c_size, [_c_void]),
    ("gn_batch_set_shape", _c_int, [_c_void, _c_int, _c_int]),
    ("gn_batch_get_shape", _c_int, [_c_void, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    ("gn_batch_invert", _c_int, [_c_void, _c_void, _c_int, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, _c_void]),
    ("gn_batch_tune", _c_int, [_c_void, _c_void, _c_int, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, ctypes.c_char_p, _c_void]),
    ("gn_batch_graph_create", _c_int, [_c_void, _c_void, _c_int, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void, _c_void, ctypes.POINTER(_c_void)]),
    ("gn_graph_launch", _c_int, [_c_void, _c_void]),
    ("gn_graph_destroy", _c_int, [_c_void]),
    ("gn_curve_create", _c_int, [_c_int, _c_int, _c_void, _c_void, _c_void, _c_void, ctypes.POINTER(_c_void)]),
    ("gn_curve_destroy", _c_int, [_c_void]),
    ("gn_curve_info", _c_int, [_c_void, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    ("gn_curve_device_bytes", _c_size, [_c_void]),
    ("gn_curve_evaluate", _c_int, [_c_void, _c_void, _c_size, _c_int, _c_void, _c_void]),
    ("gn_curve_invert", _c_int, [_c_void, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(InvertOpts), _c_void, _c_void, _c_void]),
    ("gn_server_create", _c_int, [_c_void, _c_void, _c_int, ctypes.POINTER(InvertOpts), _c_size, _c_int, ctypes.POINTER(_c_void)]),
    ("gn_server_submit", _c_int, [_c_void, _c_int, _c_void, _c_void, _c_void, _c_size, ctypes.POINTER(ctypes.c_int64)]),
    ("gn_server_poll", _c_int, [_c_void, ctypes.c_int64, _c_void]),
    ("gn_server_wait", _c_int, [_c_void, ctypes.c_int64, _c_int, _c_void]),
    ("gn_server_release", _c_int, [_c_void, ctypes.c_int64]),
    ("gn_server_stats_get", _c_int, [_c_void, _c_void]),
    ("gn_server_destroy", _c_int, [_c_void]),
    ("gn_now_ns", ctypes.c_int64, []),
]

SEGMENT = np.dtype([("obj", np.int32), ("count", np.int32), ("start", np.int64)])


def lib():
    """Load libgnurbs.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # torch ships its own libamdhip64.so.7 (same SONAME as /opt/rocm's).  Loading torch first
        # makes libgnurbs bind to that already-loaded runtime, so the process has one HIP runtime;
        # loading ours first would pair the system runtime with torch's HSA library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != GN_OK:
        raise GnError(rc, lib().gn_last_error().decode())


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("gnurbs needs a visible GPU (HIP device); none is available")
    return torch


def stream_handle(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def to_device(x, dtype):
    """numpy array / tensor -> contiguous device tensor of `dtype` (a torch dtype)."""
    torch = _torch()
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    return x.to(device="cuda", dtype=dtype).contiguous()


def default_opts(**kw) -> InvertOpts:
    o = InvertOpts()
    lib().gn_invert_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def tree_count(levels: int) -> int:
    return sum(64 ** l for l in range(levels))


_live = weakref.WeakSet()


@atexit.register
def _release_all():
    # free device objects while the HIP runtime is certainly still up
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass


class Surface:
    """A rational B-spline surface resident in HBM (gn_surface).

    Construction mirrors the reference's surface tuple (degree_u, degree_v, cp_u, cp_v, knots_u,
    knots_v, control_points, weights) passed to GetMatrixN (src/NURBS/Evaluate.cu:3-4); control
    points are u-major, shape (cp_u, cp_v, 3).
    """

    def __init__(self, degree_u, degree_v, knots_u, knots_v, control_points, weights, stream=None, _handle=None):
        self._h = None
        if _handle is not None:
            self._h = _handle
        else:
            cp = np.ascontiguousarray(control_points, dtype=np.float64)
            w = np.ascontiguousarray(weights, dtype=np.float64)
            ku = np.ascontiguousarray(knots_u, dtype=np.float64)
            kv = np.ascontiguousarray(knots_v, dtype=np.float64)
            nu, nv = w.shape
            h = _c_void()
            _check(lib().gn_surface_create(int(degree_u), int(degree_v), nu, nv, ku.ctypes.data, kv.ctypes.data,
                                           cp.ctypes.data, w.ctypes.data, _c_void(stream_handle(stream)), ctypes.byref(h)))
            self._h = h
        p, q, nu, nv = _c_int(), _c_int(), _c_int(), _c_int()
        _check(lib().gn_surface_info(self._h, ctypes.byref(p), ctypes.byref(q), ctypes.byref(nu), ctypes.byref(nv)))
        self.p, self.q, self.nu, self.nv = p.value, q.value, nu.value, nv.value
        self.tree_levels = 0
        _live.add(self)

    @classmethod
    def readNu(cls, path: str, stream=None) -> "Surface":
        h = _c_void()
        _check(lib().gn_surface_from_nu(path.encode(), _c_void(stream_handle(stream)), ctypes.byref(h)))
        return cls(0, 0, None, None, None, None, _handle=h)

    def writeNu(self, path: str):
        """Writes the definition as a .nu file that readNu reads back identically."""
        _check(lib().gn_surface_to_nu(self._h, path.encode()))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().gn_surface_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self) -> int:
        return int(lib().gn_surface_device_bytes(self._h))

    # ---- span lookup / span table -------------------------------------------------------
    def findSpan(self, direction: int, u, stream=None):
        torch = _torch()
        ud = to_device(u, torch.float64)
        out = torch.empty(ud.numel(), dtype=torch.int32, device="cuda")
        _check(lib().gn_find_span(self._h, direction, ud.data_ptr(), ud.numel(), out.data_ptr(), _c_void(stream_handle(stream))))
        return out

    def GetMatrixN(self, i: int, j: int) -> np.ndarray:
        """Span matrix of knot spans (i, j) in the reference order [c][r][s] (FP64)."""
        out = np.zeros((4, self.p + 1, self.q + 1), dtype=np.float64)
        _check(lib().gn_get_span_matrix(self._h, i, j, out.ctypes.data))
        return out

    # ---- evaluation ----------------------------------------------------------------------
    def Evaluate(self, u, v, order: int = 0, precision: int = GN_FP64, out=None, stream=None):
        """Point and derivatives; returns a device tensor of shape (3*(1, 3, 6, 10)[order], n)."""
        torch = _torch()
        nblk = (3, 9, 18, 30)[order] if 0 <= order <= 3 else 3  # bad orders are rejected by the C ABI
        if precision == GN_FP64:
            ud, vd = to_device(u, torch.float64), to_device(v, torch.float64)
            n = ud.numel()
            out = out if out is not None else torch.empty((nblk, n), dtype=torch.float64, device="cuda")
            _check(lib().gn_evaluate(self._h, ud.data_ptr(), vd.data_ptr(), n, order, out.data_ptr(), _c_void(stream_handle(stream))))
        else:
            ud, vd = to_device(u, torch.float32), to_device(v, torch.float32)
            n = ud.numel()
            out = out if out is not None else torch.empty((nblk, n), dtype=torch.float32, device="cuda")
            _check(lib().gn_evaluate_f32(self._h, ud.data_ptr(), vd.data_ptr(), n, order, out.data_ptr(), _c_void(stream_handle(stream))))
        return out

    EvaluateRandom = Evaluate

    def EvaluateSorted(self, u, v, order: int = 0, out=None, perm=None, stream=None):
        """FP32 span-GEMM evaluation of points in any order (gn_evaluate_sorted_f32): returns
        (out, perm) with column k of out the fields at input point perm[k]."""
        torch = _torch()
        ud, vd = to_device(u, torch.float32), to_device(v, torch.float32)
        n = ud.numel()
        nblk = (3, 9, 18)[order]
        if out is None:
            out = torch.empty((nblk, n), dtype=torch.float32, device="cuda")
        if perm is None:
            perm = torch.empty(n, dtype=torch.int32, device="cuda")
        _check(lib().gn_evaluate_sorted_f32(self._h, ud.data_ptr(), vd.data_ptr(), n, order, out.data_ptr(),
                                            perm.data_ptr(), _c_void(stream_handle(stream))))
        return out, perm

    def EvaluateGrid(self, us, vs, order: int = 0, out=None, stream=None):
        """FP32 MFMA span-GEMM evaluation on the grid us x vs; returns (nblk, len(us)*len(vs))."""
        torch = _torch()
        ud, vd = to_device(us, torch.float32), to_device(vs, torch.float32)
        nblk = (3, 9, 18)[order] if 0 <= order <= 2 else 3
        out = out if out is not None else torch.empty((nblk, ud.numel() * vd.numel()), dtype=torch.float32, device="cuda")
        _check(lib().gn_evaluate_grid_f32(self._h, ud.data_ptr(), ud.numel(), vd.data_ptr(), vd.numel(), order,
                                          out.data_ptr(), _c_void(stream_handle(stream))))
        return out

    # ---- tree / seeds / inversion ----------------------------------------------------------
    def BuildTree(self, levels: int = 3, flags: int = 0, stream=None):
        _check(lib().gn_build_tree(self._h, levels, flags, _c_void(stream_handle(stream))))
        self.tree_levels = levels

    def GetTree(self):
        cnt = tree_count(self.tree_levels)
        nodes = np.zeros((cnt, 13), dtype=np.float64)
        valid = np.zeros(cnt, dtype=np.uint8)
        divided = np.zeros(cnt, dtype=np.uint8)
        _check(lib().gn_get_tree(self._h, nodes.ctypes.data, valid.ctypes.data, divided.ctypes.data, cnt))
        return nodes, valid, divided

    def GetNormalCones(self):
        """(leaves, 4) float32: unit axis and half-angle of each finest-level leaf (index bx*g + by);
        half-angle 4 marks a leaf with no bounded cone."""
        g = 8 ** (self.tree_levels - 1)
        out = np.zeros((g * g, 4), dtype=np.float32)
        _check(lib().gn_get_normal_cones(self._h, out.ctypes.data, g * g))
        return out

    def BuildSeedGrid(self, resolution: int = 0, stream=None):
        """Seed grid for GN_SEED_GRID_STEP (after BuildTree; 0 picks the default resolution)."""
        _check(lib().gn_build_seed_grid(self._h, resolution, _c_void(stream_handle(stream))))

    def GetSeedGrid(self):
        """(dims int32[3], origin float32[3], cell size, cells int32[dims] of leaf reference
        indices), the layout of the oracle's seed_grid."""
        dims = np.zeros(3, dtype=np.int32)
        org = np.zeros(3, dtype=np.float32)
        cell = np.zeros(1, dtype=np.float32)
        _check(lib().gn_get_seed_grid(self._h, None, 0, dims.ctypes.data, org.ctypes.data, cell.ctypes.data))
        cells = np.zeros(tuple(int(d) for d in dims), dtype=np.int32)
        _check(lib().gn_get_seed_grid(self._h, cells.ctypes.data, cells.size, None, None, None))
        return dims, org, cell[0], cells

    def Seed(self, x, y, z, exact=True, stream=None):
        """Seed search: exact=True the exact nearest leaf, False the greedy leaf, "step" the greedy
        descent's leaf plus one Gauss-Newton step (GN_SEED_GREEDY_STEP), "grid" the seed grid's
        leaf plus the same step (GN_SEED_GRID_STEP).  Returns (leaf, u, v)."""
        torch = _torch()
        mode = 2 if exact == "step" else 3 if exact == "grid" else (1 if exact else 0)
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        n = xd.numel()
        leaf = torch.empty(n, dtype=torch.int32, device="cuda")
        su = torch.empty(n, dtype=torch.float64, device="cuda")
        sv = torch.empty(n, dtype=torch.float64, device="cuda")
        _check(lib().gn_seed(self._h, mode, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n,
                             leaf.data_ptr(), su.data_ptr(),
                             sv.data_ptr(), _c_void(stream_handle(stream))))
        return leaf, su, sv

    def NewInverse(self, x, y, z, opts: Optional[InvertOpts] = None, seeds=None, out=None, stream=None,
                   iters=None):
        """Invert points (x, y, z); returns (u, v, status) device tensors, plus the per-point FP32
        and FP64 iteration counts when iters=True."""
        torch = _torch()
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        n = xd.numel()
        opts = opts if opts is not None else default_opts()
        if out is not None:
            u, v, st = out
        else:
            u = torch.empty(n, dtype=torch.float64, device="cuda")
            v = torch.empty(n, dtype=torch.float64, device="cuda")
            st = torch.empty(n, dtype=torch.uint8, device="cuda")
        if seeds is not None:
            u.copy_(to_device(seeds[0], torch.float64))
            v.copy_(to_device(seeds[1], torch.float64))
            opts.seed_mode = SEED_GIVEN
        if iters is None:
            _check(lib().gn_invert(self._h, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n, ctypes.byref(opts),
                                   u.data_ptr(), v.data_ptr(), st.data_ptr(), _c_void(stream_handle(stream))))
            return u, v, st
        it32 = torch.empty(n, dtype=torch.uint8, device="cuda")
        it64 = torch.empty(n, dtype=torch.uint8, device="cuda")
        _check(lib().gn_invert_iters(self._h, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n, ctypes.byref(opts),
                                     u.data_ptr(), v.data_ptr(), st.data_ptr(), it32.data_ptr(), it64.data_ptr(),
                                     _c_void(stream_handle(stream))))
        return u, v, st, it32, it64

    Invert = NewInverse

    def NewInverseNoAlloc(self, x, y, z, opts: Optional[InvertOpts] = None, seeds=None, out=None, stream=None):
        """gn_invert_noalloc: NewInverse with the library's scratch in a fixed device array
        instead of allocated memory (calls that overlap on one GPU race on it); no fallback.
        Allocates nothing itself when x, y, z are device tensors and out is given."""
        torch = _torch()
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        n = xd.numel()
        opts = opts if opts is not None else default_opts()
        if out is not None:
            u, v, st = out
        else:
            u = torch.empty(n, dtype=torch.float64, device="cuda")
            v = torch.empty(n, dtype=torch.float64, device="cuda")
            st = torch.empty(n, dtype=torch.uint8, device="cuda")
        if seeds is not None:
            u.copy_(to_device(seeds[0], torch.float64))
            v.copy_(to_device(seeds[1], torch.float64))
            opts.seed_mode = SEED_GIVEN
        _check(lib().gn_invert_noalloc(self._h, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n, ctypes.byref(opts),
                                       u.data_ptr(), v.data_ptr(), st.data_ptr(), _c_void(stream_handle(stream))))
        return u, v, st

    def PCurve(self, curve, samples: int, degree: int = 3, ncp: int = 16, opts: Optional[InvertOpts] = None,
               stream=None):
        """p-curve of `curve` on this surface (gn_pcurve): the clamped uniform B-spline (u(t), v(t))
        fitted to the inversions of `samples` curve points.  Returns (control points (ncp, 2),
        knots, largest 3-D deviation over the samples, samples that did not converge)."""
        import numpy as np

        opts = opts if opts is not None else default_opts()
        cp = np.zeros(2 * ncp)
        kn = np.zeros(ncp + degree + 1)
        dev, nf = _c_double(), _c_size()
        _check(lib().gn_pcurve(self._h, curve.handle, samples, degree, ncp, ctypes.byref(opts), cp.ctypes.data,
                               kn.ctypes.data, ctypes.byref(dev), ctypes.byref(nf), _c_void(stream_handle(stream))))
        return cp.reshape(ncp, 2), kn, dev.value, nf.value


class Curve:
    """A rational B-spline curve resident in HBM (gn_curve); no counterpart in the reference.

    Construction: (degree, knots, control_points (n, 3), weights (n,)).
    """

    def __init__(self, degree, knots, control_points, weights, stream=None):
        self._h = None
        cp = np.ascontiguousarray(control_points, dtype=np.float64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        k = np.ascontiguousarray(knots, dtype=np.float64)
        h = _c_void()
        _check(lib().gn_curve_create(int(degree), int(w.size), k.ctypes.data, cp.ctypes.data, w.ctypes.data,
                                     _c_void(stream_handle(stream)), ctypes.byref(h)))
        self._h = h
        p, n = _c_int(), _c_int()
        _check(lib().gn_curve_info(self._h, ctypes.byref(p), ctypes.byref(n)))
        self.p, self.n = p.value, n.value
        _live.add(self)

    @classmethod
    def from_def(cls, cd, stream=None) -> "Curve":
        return cls(cd.p, cd.k, cd.cp, cd.w, stream)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().gn_curve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self) -> int:
        return int(lib().gn_curve_device_bytes(self._h))

    def Evaluate(self, t, order: int = 0, out=None, stream=None):
        """C (and C', C'') at t; returns a device tensor of shape (3*(order+1), n)."""
        torch = _torch()
        td = to_device(t, torch.float64)
        n = td.numel()
        nblk = 3 * (order + 1) if 0 <= order <= 2 else 3
        out = out if out is not None else torch.empty((nblk, n), dtype=torch.float64, device="cuda")
        _check(lib().gn_curve_evaluate(self._h, td.data_ptr(), n, order, out.data_ptr(), _c_void(stream_handle(stream))))
        return out

    def Invert(self, x, y, z, opts: Optional[InvertOpts] = None, seeds=None, out=None, stream=None):
        """Point inversion onto the curve; returns (t, status) device tensors."""
        torch = _torch()
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        n = xd.numel()
        opts = opts if opts is not None else default_opts()
        if out is not None:
            t, st = out
        else:
            t = torch.empty(n, dtype=torch.float64, device="cuda")
            st = torch.empty(n, dtype=torch.uint8, device="cuda")
        if seeds is not None:
            t.copy_(to_device(seeds, torch.float64))
            opts.seed_mode = SEED_GIVEN
        _check(lib().gn_curve_invert(self._h, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n, ctypes.byref(opts),
                                     t.data_ptr(), st.data_ptr(), _c_void(stream_handle(stream))))
        return t, st


class RequestInfo(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_int64), ("admit_ns", ctypes.c_int64), ("launch_ns", ctypes.c_int64),
                ("done_ns", ctypes.c_int64), ("u", _c_void), ("v", _c_void), ("status", _c_void),
                ("n", _c_size), ("obj", _c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("steps", ctypes.c_int64),
                ("points", ctypes.c_int64), ("max_step_requests", ctypes.c_int64), ("arena_bytes", _c_size),
                ("arena_peak", _c_size), ("launch_ns_total", ctypes.c_int64), ("launch_ns_max", ctypes.c_int64)]


class _DeviceView:
    """Zero-copy view of a device array for torch.as_tensor (CUDA array interface)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (int(n),), "typestr": typestr, "data": (int(ptr or 0), False),
                                         "version": 2, "strides": None}


def now_ns() -> int:
    """CLOCK_MONOTONIC in ns: the clock of the server's request timestamps."""
    return int(lib().gn_now_ns())


class Server:
    """Continuous-batching inversion server (gn_server): objects 0..nsurf-1 are the batch's
    surfaces, then the curves."""

    def __init__(self, batch=None, curves=(), opts: Optional[InvertOpts] = None, hbm_budget: int = 1 << 30,
                 streams: int = 2):
        self._h = None
        self._batch, self._curves = batch, list(curves)  # must outlive the server
        arr = (_c_void * max(1, len(self._curves)))(*[c.handle for c in self._curves])
        opts = opts if opts is not None else default_opts()
        h = _c_void()
        _check(lib().gn_server_create(batch.handle if batch is not None else None, arr, len(self._curves),
                                      ctypes.byref(opts), int(hbm_budget), int(streams), ctypes.byref(h)))
        self._h = h
        self._keep = {}
        _live.add(self)

    def close(self):
        if self._h:
            lib().gn_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, obj: int, x, y, z) -> int:
        """x, y, z: device float64 tensors, kept alive until release."""
        torch = _torch()
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        rid = ctypes.c_int64()
        _check(lib().gn_server_submit(self._h, int(obj), xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), xd.numel(),
                                      ctypes.byref(rid)))
        self._keep[rid.value] = (xd, yd, zd)
        return rid.value

    def poll(self, rid: int) -> Optional[RequestInfo]:
        info = RequestInfo()
        rc = lib().gn_server_poll(self._h, rid, ctypes.byref(info))
        if rc == GN_EAGAIN:
            return None
        _check(rc)
        return info

    def wait(self, rid: int, timeout_ms: int = -1) -> Optional[RequestInfo]:
        info = RequestInfo()
        rc = lib().gn_server_wait(self._h, rid, int(timeout_ms), ctypes.byref(info))
        if rc == GN_EAGAIN:
            return None
        _check(rc)
        return info

    @staticmethod
    def results(info: RequestInfo):
        """Copies of a finished request's results: (u, v, status) for surfaces, (t, None, status)
        for curves, as device tensors."""
        torch = _torch()
        n = info.n
        u = torch.as_tensor(_DeviceView(info.u, n, "<f8"), device="cuda").clone()
        v = torch.as_tensor(_DeviceView(info.v, n, "<f8"), device="cuda").clone() if info.v else None
        st = torch.as_tensor(_DeviceView(info.status, n, "|u1"), device="cuda").clone()
        return u, v, st

    def release(self, rid: int):
        _check(lib().gn_server_release(self._h, rid))
        self._keep.pop(rid, None)

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib().gn_server_stats_get(self._h, ctypes.byref(s)))
        return s


def segments(objs, counts, starts=None) -> np.ndarray:
    """gn_segment array: points starts[i] .. starts[i]+counts[i]-1 belong to object objs[i]
    (starts default to the running sum of counts)."""
    counts = np.asarray(counts, dtype=np.int64)
    if starts is None:
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    s = np.zeros(len(counts), dtype=SEGMENT)
    s["obj"], s["count"], s["start"] = objs, counts, starts
    return s


class Graph:
    """An instantiated hipGraph of one batched inversion (gn_graph)."""

    def __init__(self, handle, keep):
        self._h, self._keep = handle, keep

    def launch(self, stream=None):
        _check(lib().gn_graph_launch(self._h, _c_void(stream_handle(stream))))

    def close(self):
        if self._h:
            lib().gn_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many surfaces packed for one persistent inversion launch (gn_batch).  Every surface must
    have its tree built."""

    def __init__(self, surfaces, stream=None):
        arr = (_c_void * len(surfaces))(*[s.handle for s in surfaces])
        h = _c_void()
        _check(lib().gn_batch_create(arr, len(surfaces), _c_void(stream_handle(stream)), ctypes.byref(h)))
        self._h = h
        self.count = len(surfaces)
        _live.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().gn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self) -> int:
        return int(lib().gn_batch_device_bytes(self._h))

    def set_shape(self, tile_points, blocks_per_cu=0):
        _check(lib().gn_batch_set_shape(self._h, tile_points, blocks_per_cu))

    def shape(self):
        t, b = _c_int(), _c_int()
        _check(lib().gn_batch_get_shape(self._h, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def _bufs(self, x, y, z, out):
        torch = _torch()
        xd, yd, zd = (to_device(a, torch.float64) for a in (x, y, z))
        n = xd.numel()
        if out is None:
            out = (torch.empty(n, dtype=torch.float64, device="cuda"), torch.empty(n, dtype=torch.float64, device="cuda"),
                   torch.empty(n, dtype=torch.uint8, device="cuda"))
        return xd, yd, zd, n, out

    def invert(self, segs, x, y, z, opts: Optional[InvertOpts] = None, out=None, stream=None):
        segs = np.ascontiguousarray(segs, dtype=SEGMENT)
        xd, yd, zd, n, (u, v, st) = self._bufs(x, y, z, out)
        opts = opts if opts is not None else default_opts()
        _check(lib().gn_batch_invert(self._h, segs.ctypes.data, len(segs), xd.data_ptr(), yd.data_ptr(), zd.data_ptr(),
                                     n, ctypes.byref(opts), u.data_ptr(), v.data_ptr(), st.data_ptr(),
                                     _c_void(stream_handle(stream))))
        return u, v, st

    def tune(self, segs, x, y, z, opts: Optional[InvertOpts] = None, cache_path: Optional[str] = None, out=None,
             stream=None):
        segs = np.ascontiguousarray(segs, dtype=SEGMENT)
        xd, yd, zd, n, (u, v, st) = self._bufs(x, y, z, out)
        opts = opts if opts is not None else default_opts()
        _check(lib().gn_batch_tune(self._h, segs.ctypes.data, len(segs), xd.data_ptr(), yd.data_ptr(), zd.data_ptr(),
                                   n, ctypes.byref(opts), u.data_ptr(), v.data_ptr(), st.data_ptr(),
                                   cache_path.encode() if cache_path else None, _c_void(stream_handle(stream))))
        return self.shape()

    def graph(self, segs, x, y, z, u, v, status, opts: Optional[InvertOpts] = None, stream=None) -> Graph:
        """Captures one batched inversion on these device tensors; stream must be a created
        stream (torch.cuda.Stream)."""
        segs = np.ascontiguousarray(segs, dtype=SEGMENT)
        opts = opts if opts is not None else default_opts()
        h = _c_void()
        _check(lib().gn_batch_graph_create(self._h, segs.ctypes.data, len(segs), x.data_ptr(), y.data_ptr(),
                                           z.data_ptr(), x.numel(), ctypes.byref(opts), u.data_ptr(), v.data_ptr(),
                                           status.data_ptr(), _c_void(stream_handle(stream)), ctypes.byref(h)))
        return Graph(h, (x, y, z, u, v, status))
