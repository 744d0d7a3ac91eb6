"""The allocation-free frame prep (nes_frameprep_device_static), called
through the C-ABI, against the allocating entry point nes_frameprep_device:
bit-identical planes on two calls in a row, the second on a smaller input
(fewer chroma rows than the table the first call left in the fixed scratch),
and a height beyond the scratch's limit refused before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def device_inputs(nes, w, h, n, seed):
    rng = np.random.default_rng(seed)
    host = (rng.integers(0, 256, (n, h, w, 3), dtype=np.uint8),
            rng.uniform(0.1, 10.0, (n, h, w)).astype(np.float32),
            rng.integers(0, 256, (n, h, w, 3), dtype=np.uint8),
            rng.uniform(0.1, 10.0, (n, h, w)).astype(np.float32))
    bufs = [nes.DeviceBuffer(x.nbytes) for x in host]
    for b, x in zip(bufs, host):
        b.upload(x)
    return bufs


def planes(nes, w, h, n, ys, uvs):
    return [nes.DeviceBuffer(n * ys * h), nes.DeviceBuffer(n * uvs * h // 2),
            nes.DeviceBuffer(n * uvs * h // 2)]


def download(bufs, w, h, n, ys, uvs):
    y = bufs[0].download(np.empty((n, h, ys), np.uint8))
    u = bufs[1].download(np.empty((n, h // 2, uvs), np.uint8))
    v = bufs[2].download(np.empty((n, h // 2, uvs), np.uint8))
    return y[:, :, :w], u[:, :, :w // 2], v[:, :, :w // 2]


def check_case(nes, w, h, n, mode, composite, seed, ys=None, uvs=None):
    ys, uvs = ys or w, uvs or w // 2
    a, da, b, db = device_inputs(nes, w, h, n, seed)
    if not composite:
        b = db = None
    ref, got = planes(nes, w, h, n, ys, uvs), planes(nes, w, h, n, ys, uvs)
    nes.frameprep_device(a, da, b, db, w, h, n, *ref, mode=mode, y_stride=ys, uv_stride=uvs)
    nes.frameprep_device_static(a, da, b, db, w, h, n, *got, mode=mode, y_stride=ys,
                                uv_stride=uvs)
    for r, g in zip(download(ref, w, h, n, ys, uvs), download(got, w, h, n, ys, uvs)):
        assert np.array_equal(r, g)
    for buf in [a, da] + ([b, db] if composite else []) + ref + got:
        buf.free()


@pytest.mark.parametrize("mode", [0, 1])
def test_static_frameprep_matches_allocating_entry_point(nes, mode):
    check_case(nes, 2064, 2208, 2, mode, True, 10 + mode, ys=2080, uvs=1056)  # first call
    check_case(nes, 640, 480, 1, mode, False, 20 + mode)                       # smaller, second


def test_static_frameprep_refuses_heights_beyond_scratch(nes):
    a = nes.DeviceBuffer(3 * 4 * 8194)
    y = nes.DeviceBuffer(4 * 8194)
    u = nes.DeviceBuffer(2 * 4097)
    with pytest.raises(nes.NesError) as e:
        nes.frameprep_device_static(a, None, None, None, 4, 8194, 1, y, u, u)
    assert e.value.code == -34
    for buf in (a, y, u):
        buf.free()
