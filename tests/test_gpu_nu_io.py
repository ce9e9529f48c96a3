"""The .nu writer (gn_surface_to_nu): every reference model file round-trips exactly -- the
written file parses (with the independent test reader) to the same knots, points and weights, and
the surface read back evaluates bit-identically."""
import os

import numpy as np
import pytest

import gnurbs as G
from gnurbs import synth
from helpers import read_nu

pytestmark = pytest.mark.gpu

MODELS = ["center_bk.nu", "part3_0.nu", "part3_1.nu", "part3_2.nu", "part3_3.nu", "test.nu",
          "test_face1.nu", "test_face2.nu"]


@pytest.mark.parametrize("fname", MODELS)
def test_reference_models_round_trip(golden_dir, tmp_path, fname):
    src = os.path.join(golden_dir, fname)
    a = G.Surface.readNu(src)
    out = str(tmp_path / fname)
    a.writeNu(out)
    for x, y in zip(read_nu(src), read_nu(out)):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    b = G.Surface.readNu(out)
    u, v = synth.uniform_uv(4000, seed=5)
    ea = a.Evaluate(u, v, 2).cpu().numpy()
    eb = b.Evaluate(u, v, 2).cpu().numpy()
    assert np.array_equal(ea, eb)
    a.close()
    b.close()


def test_generated_surface_round_trips_all_digits(tmp_path):
    sd = synth.surface_G(9, 7, 3, 2, seed=4, jitter=0.3)  # jittered knots: non-terminating decimals
    a = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    out = str(tmp_path / "g.nu")
    a.writeNu(out)
    p, q, ku, kv, cp, w = read_nu(out)
    assert (p, q) == (sd.p, sd.q)
    assert np.array_equal(ku, sd.ku) and np.array_equal(kv, sd.kv)
    assert np.array_equal(cp, sd.cp) and np.array_equal(w, sd.w)


def test_write_errors_are_reported(tmp_path):
    sd = synth.surface_G(5, 5, 2, 2, seed=1)
    a = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    with pytest.raises(G.GnError) as e:
        a.writeNu(str(tmp_path / "no" / "such" / "dir" / "x.nu"))
    assert e.value.code == G.GN_EIO
