"""CPU checks of the drop-in boundary: libgnurbs.so loads without a GPU and exports every
function declared in include/*.h; host-side argument checks fail cleanly without a device."""
import ctypes
import os
import re
import subprocess

import pytest

import gnurbs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        src = open(os.path.join(inc, f)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(gn_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("gn_surface_create", "gn_find_span", "gn_evaluate", "gn_build_tree", "gn_invert"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gn_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_version_and_defaults_without_gpu():
    assert b"gfx950" in G.lib().gn_version()
    o = G.default_opts()
    assert o.max_iter == 20 and o.precision == G.GN_FP64 and o.seed_mode == G.SEED_GREEDY_LEAF
    assert o.fp64_tail == 3 and o.newton == 0 and abs(o.tol_step - 1e-10) < 1e-25


def test_argument_checks_fail_cleanly_without_gpu():
    L = G.lib()
    h = ctypes.c_void_p()
    # degree out of range: rejected on the host before any device call
    import numpy as np

    k = np.zeros(20)
    rc = L.gn_surface_create(9, 3, 4, 4, k.ctypes.data, k.ctypes.data, k.ctypes.data, k.ctypes.data, None,
                             ctypes.byref(h))
    assert rc == G.GN_EINVAL and b"degree" in L.gn_last_error()
    # decreasing knots
    ku = np.array([0, 0, 0, 0, 1, 0.5, 1, 1.0])
    rc = L.gn_surface_create(3, 3, 4, 4, ku.ctypes.data, ku.ctypes.data, np.zeros(48).ctypes.data,
                             np.ones(16).ctypes.data, None, ctypes.byref(h))
    assert rc == G.GN_EINVAL and b"non-decreasing" in L.gn_last_error()
    # non-positive weights
    kk = np.array([0, 0, 0, 0, 1, 1, 1, 1.0])
    rc = L.gn_surface_create(3, 3, 4, 4, kk.ctypes.data, kk.ctypes.data, np.zeros(48).ctypes.data,
                             np.zeros(16).ctypes.data, None, ctypes.byref(h))
    assert rc == G.GN_EINVAL and b"weights" in L.gn_last_error()
    # null surface handles
    assert L.gn_evaluate(None, None, None, 0, 0, None, None) == G.GN_EINVAL
    assert L.gn_invert(None, None, None, None, 0, None, None, None, None, None) == G.GN_EINVAL
    assert L.gn_surface_from_nu(b"/nonexistent/file.nu", None, ctypes.byref(h)) == G.GN_EIO


def test_curve_argument_checks_fail_cleanly_without_gpu():
    import numpy as np

    L = G.lib()
    h = ctypes.c_void_p()
    k = np.array([0, 0, 0, 0, 1, 1, 1, 1.0])
    cp, w = np.zeros(12), np.ones(4)
    assert L.gn_curve_create(9, 4, k.ctypes.data, cp.ctypes.data, w.ctypes.data, None, ctypes.byref(h)) == G.GN_EINVAL
    assert b"degree" in L.gn_last_error()
    assert L.gn_curve_create(3, 3, k.ctypes.data, cp.ctypes.data, w.ctypes.data, None, ctypes.byref(h)) == G.GN_EINVAL
    kd = np.array([0, 0, 0, 0, 1, 0.5, 1, 1.0])
    assert L.gn_curve_create(3, 4, kd.ctypes.data, cp.ctypes.data, w.ctypes.data, None, ctypes.byref(h)) == G.GN_EINVAL
    assert b"non-decreasing" in L.gn_last_error()
    w0 = np.zeros(4)
    assert L.gn_curve_create(3, 4, k.ctypes.data, cp.ctypes.data, w0.ctypes.data, None, ctypes.byref(h)) == G.GN_EINVAL
    assert b"weights" in L.gn_last_error()
    assert L.gn_curve_evaluate(None, None, 0, 0, None, None) == G.GN_EINVAL
    assert L.gn_curve_invert(None, None, None, None, 0, None, None, None, None) == G.GN_EINVAL


def test_server_argument_checks_fail_cleanly_without_gpu():
    L = G.lib()
    h = ctypes.c_void_p()
    o = G.default_opts()
    assert L.gn_server_create(None, None, 0, ctypes.byref(o), 1 << 30, 2, ctypes.byref(h)) == G.GN_EINVAL
    assert b"objects" in L.gn_last_error()
    fake = (ctypes.c_void_p * 1)(None)
    assert L.gn_server_create(None, fake, 1, ctypes.byref(o), 1 << 30, 0, ctypes.byref(h)) == G.GN_EINVAL
    assert b"streams" in L.gn_last_error()
    o.seed_mode = G.SEED_GIVEN
    assert L.gn_server_create(None, fake, 1, ctypes.byref(o), 1 << 30, 2, ctypes.byref(h)) == G.GN_EINVAL
    o = G.default_opts()
    assert L.gn_server_create(None, fake, 1, ctypes.byref(o), 1 << 30, 2, ctypes.byref(h)) == G.GN_EINVAL
    assert b"null curve" in L.gn_last_error()
    rid = ctypes.c_int64()
    assert L.gn_server_submit(None, 0, None, None, None, 0, ctypes.byref(rid)) == G.GN_EINVAL
    assert L.gn_server_poll(None, 1, None) == G.GN_EINVAL
    assert L.gn_server_release(None, 1) == G.GN_EINVAL
    assert L.gn_server_destroy(None) == G.GN_OK
    assert L.gn_now_ns() > 0
