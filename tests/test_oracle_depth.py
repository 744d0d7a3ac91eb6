"""The GRAY8 depth plane of a RenderedFrame (proto/nes.proto:24): the
oracle's quantiser g = clamp(floor(fma(d, 255 / far, 0.5)), 0, 255), pinned
by closed forms that a dropped offset, a swapped scale or a wrong rounding
direction would fail."""
import numpy as np


def test_end_points_and_clamping(oracle):
    for far in (1.0, 2.5, 4.0, 10.0, 255.0):
        g = oracle.depth_gray8(np.array([0.0, far, 2 * far, -1.0, np.inf, -np.inf, np.nan],
                                        np.float32), far)
        assert g.tolist() == [0, 255, 255, 0, 255, 0, 0], far


def test_unit_scale_rounds_half_up(oracle):
    # far = 255 makes the scale exactly 1: g = floor(d + 0.5), so k + 0.5
    # rounds up and the float just below it rounds down (from k = 1: below
    # 0.5 the sum 1 - 2^-25 itself rounds to 1.0 in FP32, a tie to even)
    k = np.arange(0, 255, dtype=np.float32)
    up = oracle.depth_gray8(k + np.float32(0.5), 255.0)
    assert (up == np.minimum(k + 1, 255)).all()
    below = np.nextafter(k + np.float32(0.5), np.float32(0)).astype(np.float32)
    g = oracle.depth_gray8(below, 255.0)
    assert (g[1:] == k[1:]).all() and g[0] == 1
    assert oracle.depth_gray8(np.array([0.49999994], np.float32), 255.0)[0] == 0
    assert (oracle.depth_gray8(k, 255.0) == k).all()


def test_matches_real_rounding_away_from_ties(oracle):
    # far = 4: scale 63.75 is exact in FP32; wherever d * 63.75 + 0.5 is not
    # within FP32 rounding of an integer, the FP32 rule equals the exact one
    rng = np.random.default_rng(0)
    d = rng.uniform(-0.5, 4.5, 200000).astype(np.float32)
    exact = d.astype(np.float64) * 63.75 + 0.5
    clear = np.abs(exact - np.round(exact)) > 1e-4
    want = np.clip(np.floor(exact), 0, 255).astype(np.uint8)
    got = oracle.depth_gray8(d, 4.0)
    assert clear.mean() > 0.99
    assert (got[clear] == want[clear]).all()


def test_monotone_in_depth(oracle):
    d = np.sort(np.random.default_rng(1).uniform(0, 5, 100000).astype(np.float32))
    g = oracle.depth_gray8(d, 4.0).astype(int)
    assert (np.diff(g) >= 0).all() and g[0] == 0 and g[-1] == 255
