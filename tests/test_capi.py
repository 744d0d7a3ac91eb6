"""CPU checks of the C-ABI library: it loads, exports every entry point the
public header declares, and its host-side pieces (scene synthesis, camera
helper, libswscale filter table) agree bit for bit with the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(nes_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol(nes):
    names = declared_symbols()
    assert len(names) >= 15
    L = nes.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_declares_every_symbol(nes):
    # the ctypes binding types every entry point of the header
    assert not [n for n in declared_symbols() if n not in nes._SIGS]


def test_strerror(nes):
    L = nes.lib()
    assert L.nes_strerror(0) == b"ok"
    assert L.nes_strerror(-22) == b"invalid argument"


def test_scene_synthesis_matches_oracle(nes, oracle, scene1):
    t, w, o = nes.scene_synthesize(1)
    t0, w0, o0 = scene1
    assert np.array_equal(t.view(np.uint32), t0.view(np.uint32))
    assert np.array_equal(w.view(np.uint32), w0.view(np.uint32))
    assert np.array_equal(o, o0)
    t1, _, _ = nes.scene_synthesize(4, table_init=1)
    t2, _, _ = oracle.scene(4, table_init=1, weights=False, occ=False)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert np.abs(t1).max() <= 1e-4


def test_orbit_camera_matches_oracle(nes, oracle):
    for az, el, r, e in [(30, 20, 2, 0), (0, 0, 2, 0.032), (217.5, -33, 1.5, -0.032)]:
        assert np.array_equal(nes.orbit_camera(az, el, r, e), oracle.orbit_c2w(az, el, r, e))


def test_vchroma_filter_matches_oracle(nes, oracle):
    for h in list(range(2, 64, 2)) + [480, 720, 1080, 1600, 2208]:
        c1, p1 = nes.vchroma_filter(h)
        c0, p0 = oracle.vchroma_filter(h)
        assert np.array_equal(c1, c0) and np.array_equal(p1, p0), h


def test_bad_arguments_rejected_without_gpu(nes):
    # odd sizes are rejected before any device work (camera_manager.cc:38-44)
    img = np.zeros((5, 4, 3), np.uint8)
    with pytest.raises(nes.NesError) as e:
        nes.frameprep(img)
    assert e.value.code == -22
