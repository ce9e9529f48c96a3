"""Test helpers: an independent .nu reader and an independent NURBS evaluator built on scipy's
B-spline basis (tensor-product sum over the control net, not the matrix form)."""
from __future__ import annotations

import numpy as np
from scipy.interpolate import BSpline


def read_nu(path):
    """Parse the tagged .nu text format (NU p q / CP nu nv / KNu / KNv / DP x y z w, u-major)."""
    p = q = nu = nv = None
    ku = kv = None
    pts, ws = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            tag = tok[0]
            if tag == "NU":
                p, q = int(tok[1]), int(tok[2])
            elif tag == "CP":
                nu, nv = int(tok[1]), int(tok[2])
            elif tag == "KNu":
                ku = np.array([float(t) for t in tok[1:1 + nu + p + 1]])
            elif tag == "KNv":
                kv = np.array([float(t) for t in tok[1:1 + nv + q + 1]])
            elif tag == "DP":
                pts.append([float(t) for t in tok[1:4]])
                ws.append(float(tok[4]))
    cp = np.array(pts[: nu * nv]).reshape(nu, nv, 3)
    w = np.array(ws[: nu * nv]).reshape(nu, nv)
    return p, q, ku, kv, cp, w


def _basis(knots, deg, n, x, der):
    """Values (der = 0) or derivatives of the n B-spline basis functions of degree deg.
    Derivatives use N'_{i,p} = p N_{i,p-1}/(t_{i+p}-t_i) - p N_{i+1,p-1}/(t_{i+p+1}-t_{i+1})
    (0/0 = 0) on scipy's lower-degree basis, so repeated interior knots are fine."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    if der > deg:
        return np.zeros((x.size, n))
    if der == 0:
        bs = BSpline(knots, np.eye(n), deg, extrapolate=False)
        return np.nan_to_num(bs(x))
    B = _basis(knots, deg - 1, n + 1, x, der - 1)
    t = np.asarray(knots, dtype=np.float64)
    out = np.zeros((x.size, n))
    for i in range(n):
        d1 = t[i + deg] - t[i]
        d2 = t[i + deg + 1] - t[i + 1]
        if d1 != 0:
            out[:, i] += deg * B[:, i] / d1
        if d2 != 0:
            out[:, i] -= deg * B[:, i + 1] / d2
    return out


DER_NAMES = ["S", "Su", "Sv", "Suu", "Suv", "Svv", "Suuu", "Suuv", "Suvv", "Svvv"]
DER_INDEX = {"S": (0, 0), "Su": (1, 0), "Sv": (0, 1), "Suu": (2, 0), "Suv": (1, 1), "Svv": (0, 2),
             "Suuu": (3, 0), "Suuv": (2, 1), "Suvv": (1, 2), "Svvv": (0, 3)}


def scipy_eval_general(p, q, ku, kv, cp, w, u, v, order=3, return_hom=False):
    """Independent rational evaluation of every derivative S^(k,l), k + l <= order, by the generic
    quotient rule of The NURBS Book eq. 4.20 (Piegl & Tiller, algorithm A4.4): with A = w S,
    S^(k,l) = (A^(k,l) - sum_{(i,j) != (0,0)} C(k,i) C(l,j) w^(i,j) S^(k-i,l-j)) / w.
    Returns a dict keyed by DER_NAMES, each (3, n); with return_hom also the homogeneous
    derivatives A[k, l] (3, n) and W[k, l] (n)."""
    from math import comb
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    nu, nv = w.shape
    Bu = [_basis(ku, p, nu, u, d) for d in range(order + 1)]
    Bv = [_basis(kv, q, nv, v, d) for d in range(order + 1)]
    Pw = cp * w[..., None]
    A, W, S = {}, {}, {}
    for k in range(order + 1):
        for l in range(order + 1 - k):
            A[k, l] = np.einsum("ni,ijc,nj->cn", Bu[k], Pw, Bv[l])
            W[k, l] = np.einsum("ni,ij,nj->n", Bu[k], w, Bv[l])
    for tot in range(order + 1):
        for k in range(tot, -1, -1):
            l = tot - k
            acc = A[k, l].copy()
            for i in range(k + 1):
                for j in range(l + 1):
                    if i == 0 and j == 0:
                        continue
                    acc -= comb(k, i) * comb(l, j) * W[i, j] * S[k - i, l - j]
            S[k, l] = acc / W[0, 0]
    res = {name: S[idx] for name, idx in DER_INDEX.items() if sum(idx) <= order}
    return (res, A, W) if return_hom else res


def scipy_eval(p, q, ku, kv, cp, w, u, v, order=2):
    """Independent rational evaluation: returns dict of S, Su, Sv, Suu, Suv, Svv, each (3, n)."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    nu, nv = w.shape
    Bu = [_basis(ku, p, nu, u, d) for d in range(3)]
    Bv = [_basis(kv, q, nv, v, d) for d in range(3)]
    Pw = cp * w[..., None]  # (nu, nv, 3)

    def hom(du, dv):
        A = np.einsum("ni,ijc,nj->cn", Bu[du], Pw, Bv[dv])
        W = np.einsum("ni,ij,nj->n", Bu[du], w, Bv[dv])
        return A, W

    A, W = hom(0, 0)
    S = A / W
    res = {"S": S}
    if order >= 1:
        Au, Wu = hom(1, 0)
        Av, Wv = hom(0, 1)
        Su = (Au - Wu * S) / W
        Sv = (Av - Wv * S) / W
        res.update(Su=Su, Sv=Sv)
    if order >= 2:
        Auu, Wuu = hom(2, 0)
        Auv, Wuv = hom(1, 1)
        Avv, Wvv = hom(0, 2)
        res["Suu"] = (Auu - 2 * Wu * Su - Wuu * S) / W
        res["Suv"] = (Auv - Wu * Sv - Wv * Su - Wuv * S) / W
        res["Svv"] = (Avv - 2 * Wv * Sv - Wvv * S) / W
    return res


def blocks(out):
    """Split an SoA evaluation array (nblk, n) into named (3, n) blocks."""
    return {DER_NAMES[i]: out[3 * i: 3 * i + 3] for i in range(out.shape[0] // 3)}


def field_close(a, b, rtol, floor=0.0):
    """Per-component |a-b| <= rtol*|b| + rtol*scale, scale = max(max|b| over the field, floor).
    The floor keeps a field that is zero up to rounding (e.g. S_uu of a ruled surface) from being
    compared relative to its own noise; callers pass a fraction of the point magnitude."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)) if b.size else 0.0, floor)
    err = np.abs(a - b)
    ok = err <= rtol * np.abs(b) + rtol * scale
    return bool(np.all(ok)), float(np.max(err / (np.abs(b) + scale + 1e-300))) if b.size else 0.0
