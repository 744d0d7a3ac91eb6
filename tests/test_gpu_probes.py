"""The measurement probes the bench's ceilings leg reports: they run, return
rates in a plausible range for one MI355X, and reject bad arguments."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gather_rate_probe(nes):
    l1, ms = nes.gather_rate(16 << 10)
    assert ms > 0 and 1e10 < l1 < 1e14
    L = nes.lib()
    out_ms, n = C.c_float(), C.c_uint64()
    assert L.nes_device_gather_rate(3000, C.byref(out_ms), C.byref(n)) == -22  # not a power of two


@pytest.mark.parametrize("kind", ["f32", "f64", "u32", "f32_cas", "f32x2_cas"])
def test_lds_atomic_rate_probe(nes, kind):
    rate = nes.lds_atomic_rate(kind)
    # 256 workgroups x 1024 lanes x 1024 adds; anything from 10 G to 100 T per second
    assert np.isfinite(rate) and 1e10 < rate < 1e14


def test_lds_atomic_rate_rejects_unknown_kind(nes):
    out_ms, n = C.c_float(), C.c_uint64()
    assert nes.lib().nes_device_lds_atomic_rate(5, C.byref(out_ms), C.byref(n)) == -22
    assert nes.lib().nes_device_lds_atomic_rate(-1, C.byref(out_ms), C.byref(n)) == -22
