"""ctypes wrapper of liboracle.so, the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only
as the checker.  Functions take plain numpy arrays; see gnurbs_oracle.c for the reference
file:line each one restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_l = None
_P = ctypes.c_void_p
_I, _D, _L = ctypes.c_int, ctypes.c_double, ctypes.c_long

CONV_STEP, CONV_DIST, MAXITER, SINGULAR, NONFINITE, NOSEED = range(6)
TREE_REF_PARENT = 1


def lib():
    global _l
    if _l is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        L.or_find_span.restype = _I
        L.or_find_span.argtypes = [_I, _I, _P, _D]
        L.or_nrc.restype = _D
        L.or_nrc.argtypes = [_P, _I, _I, _I, _I, _I, _I]
        L.or_nabla.restype = _D
        L.or_nabla.argtypes = [_P, _I, _I, _I]
        L.or_matrix_n_position.restype = _I
        L.or_matrix_n_position.argtypes = [_I] * 7
        L.or_basis_matrix_nrc.argtypes = [_I, _P, _I, _P]
        L.or_basis_matrix_cdb.argtypes = [_I, _P, _I, _P]
        L.or_get_matrix_n.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P]
        L.or_eval.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _L, _I, _P, _I]
        L.or_eval_homogeneous.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _L, _P]
        L.or_gauss_newton.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _L, _D, _D, _I, _I, _I,
                                      _D, _P, _P, _P, _P, _P, _P, _I]
        L.or_tree_base.restype = _L
        L.or_tree_base.argtypes = [_I]
        L.or_need_divide.restype = _I
        L.or_need_divide.argtypes = [_P, _P]
        L.or_build_tree.argtypes = [_I, _I, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P]
        L.or_invert_first_hit.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _L, _P, _P, _P, _L,
                                          _D, _D, _I, _I, _P, _P, _P, _P, _I]
        L.or_seed_nearest_leaf.argtypes = [_P, _I, _P, _P, _P, _L, _P, _P, _P]
        L.or_leaf_cones.argtypes = [_I, _I, _I, _I, _P, _P, _P, _I, _P]
        L.or_normal_ders.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _L, _P]
        L.or_project_global.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _L, _I, _I, _D,
                                        _P, _P, _P, _I]
        L.or_seed_greedy_leaf.argtypes = [_P, _I, _P, _P, _P, _L, _P, _P, _P, _I]
        L.or_seed_greedy_descent.argtypes = [_P, _I, _P, _P, _P, _L, _P, _P, _P, _I]
        L.or_seed_greedy_step.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _L, _P, _P, _P, _I]
        L.or_seed_grid_geometry.restype = _I
        L.or_seed_grid_geometry.argtypes = [_P, _I, _I, _P, _P, _P]
        L.or_seed_grid_build.argtypes = [_P, _I, _P, _P, ctypes.c_float, _P, _I]
        L.or_seed_grid_step.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, ctypes.c_float, _P,
                                        _P, _P, _P, _L, _P, _P, _P, _I]
        L.or_curve_eval.argtypes = [_I, _I, _P, _P, _P, _P, _L, _I, _P, _I]
        L.or_curve_seed.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P, _L, _P]
        L.or_curve_invert.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P, _L, _D, _D, _I, _I, _I, _P, _P, _P,
                                      _P, _I]
        _l = L
    return _l


def _a(x, dt=np.float64):
    return np.ascontiguousarray(x, dtype=dt)


def find_span(degree, cp, knots, u) -> int:
    k = _a(knots)
    return lib().or_find_span(degree, cp, k.ctypes.data, float(u))


def nrc(knots, m, r, c, j, i, d) -> float:
    k = _a(knots)
    return lib().or_nrc(k.ctypes.data, m, r, c, j, i, d)


def basis_matrix(degree, knots, span, method="nrc") -> np.ndarray:
    k = _a(knots)
    M = np.zeros((degree + 1) * (degree + 1))
    f = lib().or_basis_matrix_nrc if method == "nrc" else lib().or_basis_matrix_cdb
    f(degree, k.ctypes.data, span, M.ctypes.data)
    return M.reshape(degree + 1, degree + 1)


class OSurf:
    """Surface definition plus its oracle span table (reference padded MatrixN layout)."""

    def __init__(self, p, q, ku, kv, cp, w, use_cdb=False):
        self.p, self.q = int(p), int(q)
        self.ku, self.kv = _a(ku), _a(kv)
        self.cp, self.w = _a(cp), _a(w)
        self.nu, self.nv = self.w.shape
        slots = (self.p + self.nu) * (self.q + self.nv)
        self.MatrixN = np.zeros(slots * 4 * (self.p + 1) * (self.q + 1))
        self.valid = np.zeros(slots, dtype=np.uint8)
        lib().or_get_matrix_n(self.p, self.q, self.nu, self.nv, self.ku.ctypes.data, self.kv.ctypes.data,
                              self.cp.ctypes.data, self.w.ctypes.data, 1 if use_cdb else 0,
                              self.MatrixN.ctypes.data, self.valid.ctypes.data)

    @classmethod
    def from_def(cls, sd, use_cdb=False):
        return cls(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w, use_cdb)

    def span_matrix(self, i, j):
        k1, k2 = self.p + 1, self.q + 1
        off = lib().or_matrix_n_position(i, j, self.p, self.q, self.nu, self.nv, 0)
        return self.MatrixN[off: off + 4 * k1 * k2].reshape(4, k1, k2)

    def _args(self):
        return (self.p, self.q, self.nu, self.nv, self.ku.ctypes.data, self.kv.ctypes.data, self.MatrixN.ctypes.data)

    def eval(self, u, v, order=0, nthreads=0) -> np.ndarray:
        u, v = _a(u), _a(v)
        n = u.size
        nblk = (3, 9, 18, 30)[order]
        out = np.zeros((nblk, n))
        lib().or_eval(*self._args(), u.ctypes.data, v.ctypes.data, n, order, out.ctypes.data, nthreads)
        return out

    def eval_homogeneous(self, u, v) -> np.ndarray:
        u, v = _a(u), _a(v)
        out = np.zeros((u.size, 4))
        lib().or_eval_homogeneous(*self._args(), u.ctypes.data, v.ctypes.data, u.size, out.ctypes.data)
        return out

    def gauss_newton(self, x, y, z, seed_u, seed_v, tol_dist=0.0, tol_step=1e-10, max_iter=20, clamp=1,
                     newton=0, nthreads=0, tol_quad=0.0):
        x, y, z, su, sv = (_a(a) for a in (x, y, z, seed_u, seed_v))
        n = x.size
        u, v = np.zeros(n), np.zeros(n)
        st = np.zeros(n, dtype=np.uint8)
        it = np.zeros(n, dtype=np.int32)
        lib().or_gauss_newton(*self._args(), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, tol_dist, tol_step,
                              max_iter, clamp, newton, tol_quad, su.ctypes.data, sv.ctypes.data, u.ctypes.data,
                              v.ctypes.data, st.ctypes.data, it.ctypes.data, nthreads)
        return u, v, st, it

    def project_global(self, nodes, levels, x, y, z, max_iter=20, leaf_iter=8, tol_step=1e-12, nthreads=0):
        """Exhaustive global projection: Newton from every leaf (clamped to its uv cell), best kept,
        then polished over the whole domain."""
        x, y, z = (_a(a) for a in (x, y, z))
        nodes = _a(nodes)
        n = x.size
        u, v = np.zeros(n), np.zeros(n)
        st = np.zeros(n, dtype=np.uint8)
        lib().or_project_global(*self._args(), nodes.ctypes.data, levels, x.ctypes.data, y.ctypes.data,
                                z.ctypes.data, n, max_iter, leaf_iter, tol_step, u.ctypes.data, v.ctypes.data,
                                st.ctypes.data, nthreads)
        return u, v, st

    def seed_greedy_step(self, nodes, levels, x, y, z, nthreads=0):
        """Greedy leaf, then one Gauss-Newton step with the Jacobian at the leaf's seed (the
        library's GN_SEED_GREEDY_STEP)."""
        nodes = _a(nodes)
        x, y, z = (_a(a) for a in (x, y, z))
        n = x.size
        leaf = np.zeros(n, dtype=np.int32)
        su, sv = np.zeros(n), np.zeros(n)
        lib().or_seed_greedy_step(*self._args(), nodes.ctypes.data, levels, x.ctypes.data, y.ctypes.data,
                                  z.ctypes.data, n, leaf.ctypes.data, su.ctypes.data, sv.ctypes.data, nthreads)
        return leaf, su, sv

    def seed_grid_step(self, nodes, levels, grid, x, y, z, nthreads=0):
        """The library's GN_SEED_GRID_STEP: the leaf of P's cell in `grid` (seed_grid's result),
        then the step of seed_greedy_step from it.  Returns (leaf, su, sv)."""
        dims, org, cell, cells = grid
        nodes = _a(nodes)
        x, y, z = (_a(a) for a in (x, y, z))
        n = x.size
        leaf = np.zeros(n, dtype=np.int32)
        su, sv = np.zeros(n), np.zeros(n)
        d, o, c = _a(dims, np.int32), _a(org, np.float32), _a(cells.ravel(), np.int32)
        lib().or_seed_grid_step(*self._args(), nodes.ctypes.data, levels, d.ctypes.data, o.ctypes.data,
                                float(cell), c.ctypes.data, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                n, leaf.ctypes.data, su.ctypes.data, sv.ctypes.data, nthreads)
        return leaf, su, sv

    def normal_ders(self, u, v):
        """(n, 4, 3): N = S_u x S_v, N_uu, N_uv, N_vv at (u, v)."""
        u, v = _a(u), _a(v)
        out = np.zeros((u.size, 4, 3))
        lib().or_normal_ders(*self._args(), u.ctypes.data, v.ctypes.data, u.size, out.ctypes.data)
        return out

    def leaf_cones(self, levels):
        """Normal cones of the leaves (level levels-1, index bx*g + by): (leaves, 4) array of unit
        axis and half-angle (4.0: unbounded)."""
        g = 8 ** (levels - 1)
        out = np.zeros((g * g, 4))
        lib().or_leaf_cones(*self._args(), levels, out.ctypes.data)
        return out

    def build_tree(self, levels, flags=0):
        cnt = lib().or_tree_base(levels)
        nodes = np.zeros((cnt, 13))
        valid = np.zeros(cnt, dtype=np.uint8)
        divided = np.zeros(cnt, dtype=np.uint8)
        lib().or_build_tree(*self._args(), levels, flags, nodes.ctypes.data, valid.ctypes.data,
                            divided.ctypes.data)
        return nodes, valid, divided

    def invert_first_hit(self, nodes, valid, x, y, z, tol_dist=0.01, tol_step=0.0, max_iter=10, clamp=0,
                         nthreads=0):
        x, y, z = (_a(a) for a in (x, y, z))
        nodes, valid = _a(nodes), _a(valid, np.uint8)
        n = x.size
        u, v = np.zeros(n), np.zeros(n)
        st = np.zeros(n, dtype=np.uint8)
        node = np.zeros(n, dtype=np.int32)
        lib().or_invert_first_hit(*self._args(), nodes.ctypes.data, valid.ctypes.data, nodes.shape[0],
                                  x.ctypes.data, y.ctypes.data, z.ctypes.data, n, tol_dist, tol_step, max_iter,
                                  clamp, u.ctypes.data, v.ctypes.data, st.ctypes.data, node.ctypes.data, nthreads)
        return u, v, st, node


def tree_base(level) -> int:
    return lib().or_tree_base(level)


def need_divide(lb, rt) -> int:
    lb, rt = _a(lb), _a(rt)
    return lib().or_need_divide(lb.ctypes.data, rt.ctypes.data)


def seed_nearest_leaf(nodes, levels, x, y, z):
    nodes = _a(nodes)
    x, y, z = (_a(a) for a in (x, y, z))
    n = x.size
    leaf = np.zeros(n, dtype=np.int32)
    su, sv = np.zeros(n), np.zeros(n)
    lib().or_seed_nearest_leaf(nodes.ctypes.data, levels, x.ctypes.data, y.ctypes.data, z.ctypes.data, n,
                               leaf.ctypes.data, su.ctypes.data, sv.ctypes.data)
    return leaf, su, sv


def seed_greedy_leaf(nodes, levels, x, y, z, nthreads=0):
    nodes = _a(nodes)
    x, y, z = (_a(a) for a in (x, y, z))
    n = x.size
    leaf = np.zeros(n, dtype=np.int32)
    su, sv = np.zeros(n), np.zeros(n)
    lib().or_seed_greedy_leaf(nodes.ctypes.data, levels, x.ctypes.data, y.ctypes.data, z.ctypes.data, n,
                              leaf.ctypes.data, su.ctypes.data, sv.ctypes.data, nthreads)
    return leaf, su, sv


def seed_grid(nodes, levels, resolution=0, nthreads=0, build=True):
    """The library's seed grid (gn_build_seed_grid): (dims int32[3], origin float32[3], cell size
    (numpy float32), cells int32[dims] of leaf reference indices; None when build is False), or
    None where the library rejects the grid."""
    nodes = _a(nodes)
    dims = np.zeros(3, dtype=np.int32)
    org = np.zeros(3, dtype=np.float32)
    cell = np.zeros(1, dtype=np.float32)
    if lib().or_seed_grid_geometry(nodes.ctypes.data, levels, resolution, dims.ctypes.data, org.ctypes.data,
                                   cell.ctypes.data) != 0:
        return None
    if not build:
        return dims, org, cell[0], None
    cells = np.zeros(tuple(int(d) for d in dims), dtype=np.int32)
    lib().or_seed_grid_build(nodes.ctypes.data, levels, dims.ctypes.data, org.ctypes.data, float(cell[0]),
                             cells.ctypes.data, nthreads)
    return dims, org, cell[0], cells


def seed_greedy_descent(nodes, levels, x, y, z, nthreads=0):
    """The greedy descent's leaf and seed, without the 3 x 3 refinement (GN_SEED_GREEDY_STEP's start)."""
    nodes = _a(nodes)
    x, y, z = (_a(a) for a in (x, y, z))
    n = x.size
    leaf = np.zeros(n, dtype=np.int32)
    su, sv = np.zeros(n), np.zeros(n)
    lib().or_seed_greedy_descent(nodes.ctypes.data, levels, x.ctypes.data, y.ctypes.data, z.ctypes.data, n,
                                 leaf.ctypes.data, su.ctypes.data, sv.ctypes.data, nthreads)
    return leaf, su, sv


class OCurve:
    """Rational B-spline curve evaluated by the textbook de Boor form (no matrix form)."""

    def __init__(self, p, k, cp, w):
        self.p = int(p)
        self.k, self.cp, self.w = _a(k), _a(cp), _a(w)
        self.n = self.w.size

    @classmethod
    def from_def(cls, cd):
        return cls(cd.p, cd.k, cd.cp, cd.w)

    def _args(self):
        return (self.p, self.n, self.k.ctypes.data, self.cp.ctypes.data, self.w.ctypes.data)

    def eval(self, t, order=0, nthreads=0) -> np.ndarray:
        t = _a(t)
        out = np.zeros((3 * (order + 1), t.size))
        lib().or_curve_eval(*self._args(), t.ctypes.data, t.size, order, out.ctypes.data, nthreads)
        return out

    def seed(self, x, y, z) -> np.ndarray:
        x, y, z = _a(x), _a(y), _a(z)
        t = np.zeros(x.size)
        lib().or_curve_seed(*self._args(), x.ctypes.data, y.ctypes.data, z.ctypes.data, x.size, t.ctypes.data)
        return t

    def invert(self, x, y, z, seed, tol_dist=0.0, tol_step=1e-10, max_iter=20, clamp=1, newton=0, nthreads=0):
        x, y, z, sd = (_a(a) for a in (x, y, z, seed))
        n = x.size
        t = np.zeros(n)
        st = np.zeros(n, dtype=np.uint8)
        it = np.zeros(n, dtype=np.int32)
        lib().or_curve_invert(*self._args(), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, tol_dist, tol_step,
                              max_iter, clamp, newton, sd.ctypes.data, t.ctypes.data, st.ctypes.data,
                              it.ctypes.data, nthreads)
        return t, st, it
