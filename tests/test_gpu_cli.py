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
