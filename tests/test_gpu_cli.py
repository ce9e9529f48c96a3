"""The C++ interface (include/gnurbs.hpp) driving the reference's INVERSION_TEST sequence
(src/main.cu:277-327) on the reference's own model, in a separate process."""
import json
import os
import subprocess

import numpy as np
import pytest

import gnurbs as G
import oracle as O
from helpers import read_nu

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(G.LIB_PATH), "inversion_test")


def test_cpp_inversion_driver(golden_dir):
    path = os.path.join(golden_dir, "center_bk.nu")
    out = subprocess.run([CLI, path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    # known answer of src/kernel/kernel.cu:453-454
    assert np.max(np.abs(np.array(r["S(0.25,0.25)"]) - [0.0, 0.200049, 0.616])) < 5e-7
    assert r["status"] == G.PT_CONVERGED_DIST and r["same_status"] == r["n"]
    # same result as the oracle's first-hit chain
    p, q, ku, kv, cp, w = read_nu(path)
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    nodes, valid, _ = os_.build_tree(2)
    ru, rv, rst, _ = os_.invert_first_hit(nodes, valid, [0.0], [0.200049], [0.616], tol_dist=0.01, max_iter=10)
    assert abs(r["u"] - ru[0]) < 1e-12 and abs(r["v"] - rv[0]) < 1e-12


def test_cpp_pcurve_example(tmp_path):
    """NurbsCurve + NurbsSurface::PCurve from C++ on a smooth rational surface written by writeNu:
    the p-curve of an iso-curve is the parameter line u = u0, v = t (control points u0 and the
    Greville abscissae).  (The reference's models fold back on themselves, so their iso-curves can
    have other pre-images.)"""
    from test_gpu_pcurve import smooth_surface

    sd = smooth_surface()
    path = str(tmp_path / "smooth.nu")
    G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w).writeNu(path)
    exe = os.path.join(os.path.dirname(G.LIB_PATH), "pcurve_example")
    out = subprocess.run([exe, path, "0.37", "3", "12", "20000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    cp = np.array(r["cp_uv"]).reshape(-1, 2)
    kn = np.array(r["knots"])
    grev = np.array([kn[i + 1:i + 4].mean() for i in range(12)])
    assert r["failed"] == 0 and r["max_dev"] < 1e-9
    assert np.max(np.abs(cp[:, 0] - 0.37)) < 1e-9 and np.max(np.abs(cp[:, 1] - grev)) < 1e-9
