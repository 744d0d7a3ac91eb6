"""K5 frame prep on the GPU: bit-exact against the libswscale restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def inputs(w, h, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    b = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    da = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    db = rng.uniform(0.1, 10.0, (h, w)).astype(np.float32)
    return a, da, b, db


def assert_planes_equal(got, ref):
    for name, g, r in zip("yuv", got, ref):
        assert g.shape == r.shape
        bad = np.argwhere(g != r)
        assert len(bad) == 0, f"plane {name}: {len(bad)} mismatches, first {bad[:3].tolist()}"


@pytest.mark.parametrize("mode", [0, 1])
def test_two_720p_inputs_composite_bit_exact(nes, oracle, mode):
    a, da, b, db = inputs(1280, 720, 0)
    assert_planes_equal(nes.frameprep(a, da, b, db, mode=mode),
                        oracle.frameprep(a, da, b, db, mode=mode))


@pytest.mark.parametrize("mode", [0, 1])
def test_single_input_conversion_bit_exact(nes, oracle, mode):
    a, _, _, _ = inputs(1280, 720, 1)
    assert_planes_equal(nes.frameprep(a, mode=mode), oracle.rgb24_to_yuv420p(a, mode))


@pytest.mark.parametrize("w,h", [(2, 2), (4, 2), (2, 8), (6, 4), (8, 10), (10, 12), (130, 98),
                                 (258, 66), (640, 480), (1440, 1600), (2064, 2208)])
def test_sizes_bit_exact(nes, oracle, w, h):
    a, da, b, db = inputs(w, h, w * 7 + h)
    for mode in (0, 1):
        assert_planes_equal(nes.frameprep(a, da, b, db, mode=mode),
                            oracle.frameprep(a, da, b, db, mode=mode))


def test_depth_ties_nan_and_extremes(nes, oracle):
    w, h = 96, 64
    rng = np.random.default_rng(5)
    a = np.zeros((h, w, 3), np.uint8)
    b = np.full((h, w, 3), 255, np.uint8)
    levels = np.array([0.0, -0.0, 1.0, np.inf, -np.inf, np.nan, 5.0], np.float32)
    da = levels[rng.integers(0, len(levels), (h, w))]
    db = levels[rng.integers(0, len(levels), (h, w))]
    for mode in (0, 1):
        assert_planes_equal(nes.frameprep(a, da, b, db, mode=mode),
                            oracle.frameprep(a, da, b, db, mode=mode))


def test_gray8_depth_stream_bit_exact(nes, oracle):
    rng = np.random.default_rng(6)
    for w, h in [(1280, 720), (66, 10)]:
        g = rng.integers(0, 256, (h, w), dtype=np.uint8)
        for mode in (0, 1):
            assert_planes_equal(nes.gray8_to_yuv420p(g, mode), oracle.gray8_to_yuv420p(g, mode))


def test_device_batch_bit_exact(nes, oracle):
    w, h, n = 640, 480, 6
    frames = [inputs(w, h, 100 + i) for i in range(n)]
    A = np.stack([f[0] for f in frames]); DA = np.stack([f[1] for f in frames])
    B = np.stack([f[2] for f in frames]); DB = np.stack([f[3] for f in frames])
    bufs = [nes.DeviceBuffer(x.nbytes) for x in (A, DA, B, DB)]
    for buf, x in zip(bufs, (A, DA, B, DB)):
        buf.upload(x)
    Y = nes.DeviceBuffer(n * w * h); U = nes.DeviceBuffer(n * w * h // 4)
    V = nes.DeviceBuffer(n * w * h // 4)
    ms = nes.frameprep_device(*bufs, w, h, n, Y, U, V, mode=0)
    assert ms > 0
    y = Y.download(np.empty((n, h, w), np.uint8))
    u = U.download(np.empty((n, h // 2, w // 2), np.uint8))
    v = V.download(np.empty((n, h // 2, w // 2), np.uint8))
    for i, f in enumerate(frames):
        assert_planes_equal((y[i], u[i], v[i]), oracle.frameprep(*f, mode=0))
    for buf in bufs + [Y, U, V]:
        buf.free()
