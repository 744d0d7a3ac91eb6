"""Pins of the FP64 render oracle against closed forms, textbook routines and
invariants (the reference ships no tests or golden vectors for this path)."""
import math

import numpy as np
import pytest

HASH_MASK = (1 << 19) - 1


def test_splitmix64_reference_outputs(oracle):
    L = oracle.lib()
    # splitmix64 seeded with 0: its first two outputs (published reference values)
    assert L.orc_mix64(0) == 0xE220A8397B1DCDAF
    assert L.orc_mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_rng_moments(oracle):
    L = oracle.lib()
    u = np.array([L.orc_uniform(7, 3, i) for i in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    n = np.array([L.orc_normal(7, 4, i) for i in range(20000)])
    assert abs(n.mean()) < 0.03 and abs(n.std() - 1.0) < 0.03


def test_level_resolutions_and_sizes(oracle):
    L = oracle.lib()
    for l in range(16):
        # N_l = floor(16 * b^l), b = 2^(7/15); l = 15 is exactly 2048
        exact = 2048 if l == 15 else math.floor(16 * 2 ** (7 * l / 15))
        assert L.orc_level_res(l) == exact
    sizes = [L.orc_level_size(l) for l in range(16)]
    assert sizes[:5] == [(L.orc_level_res(l) + 1) ** 3 for l in range(5)]
    assert sizes[5:] == [1 << 19] * 11
    assert sum(sizes) == oracle.N_TABLE == 6098925
    assert 2 * oracle.N_TABLE + oracle.N_WEIGHTS == 12207194


@pytest.mark.parametrize("xyz,expect", [((1, 1, 1), 339493), ((2, 3, 5), 383224),
                                        ((100, 200, 300), 110768)])
def test_hash_known_answers(oracle, xyz, expect):
    # (x*1 ^ y*2654435761 ^ z*805459861) mod 2^19, worked by hand
    for level in range(5, 16):
        assert oracle.lib().orc_grid_index(level, *xyz) == expect


def test_dense_levels_injective(oracle):
    L = oracle.lib()
    for level in range(5):
        n1 = L.orc_level_res(level) + 1
        seen = set()
        for z in range(0, n1, max(1, n1 // 7)):
            for y in range(0, n1, max(1, n1 // 7)):
                for x in range(n1):
                    i = L.orc_grid_index(level, x, y, z)
                    assert i < n1 ** 3 and i not in seen
                    seen.add(i)


def _dense_coords(oracle, level):
    n1 = oracle.lib().orc_level_res(level) + 1
    idx = np.arange(n1 ** 3)
    return idx % n1, (idx // n1) % n1, idx // (n1 * n1), n1 - 1


def test_trilinear_reproduces_affine_field(oracle):
    """On a dense level, a table affine in the vertex coordinates interpolates
    to the same affine function of the position."""
    table = np.zeros(oracle.N_TABLE * 2, np.float32)
    a, b, c, d0 = 0.25, -0.5, 0.125, 0.75
    for level in (0, 3):
        off = oracle.lib().orc_level_offset(level)
        x, y, z, n = _dense_coords(oracle, level)
        f = d0 + (a * x + b * y + c * z) / n
        table[2 * (off + np.arange(len(f)))] = f.astype(np.float32)
    rng = np.random.default_rng(0)
    for p in rng.uniform(0, 1, (50, 3)):
        feat = oracle.encode(table, p)
        # exactly representable grid values; the interpolation is affine
        for level in (0, 3):
            n = oracle.lib().orc_level_res(level)
            want = d0 + a * p[0] + b * p[1] + c * p[2]
            # vertices are x/N with N the level resolution
            assert abs(feat[2 * level] - want) < 1e-6, (level, feat[2 * level], want)


def test_constant_table_gives_constant_features(oracle):
    table = np.full(oracle.N_TABLE * 2, 0.375, np.float32)
    rng = np.random.default_rng(1)
    for p in np.concatenate([rng.uniform(0, 1, (20, 3)), [[0, 0, 0], [1, 1, 1], [1, 0, 0.5]]]):
        assert np.allclose(oracle.encode(table, p), 0.375, atol=1e-12)


def test_sh_orthonormal(oracle):
    """Real SH of bands 0-3: Gauss-Legendre x uniform-phi quadrature of
    Y_i Y_j over the sphere gives the identity."""
    nt, nphi = 24, 48
    ct, wt = np.polynomial.legendre.leggauss(nt)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    gram = np.zeros((16, 16))
    for c, w in zip(ct, wt):
        s = math.sqrt(1 - c * c)
        for ph in phi:
            y = oracle.sh4([s * math.cos(ph), s * math.sin(ph), c])
            gram += w * (2 * np.pi / nphi) * np.outer(y, y)
    assert np.allclose(gram, np.eye(16), atol=1e-10)


def _np_mlp(weights, feat, sh):
    w = weights.astype(np.float64)
    W1 = w[0:2048].reshape(64, 32)
    W2 = w[2048:3072].reshape(16, 64)
    W3 = w[3072:5056].reshape(64, 31)
    W4 = w[5056:9152].reshape(64, 64)
    W5 = w[9152:9344].reshape(3, 64)
    h1 = np.maximum(W1 @ feat, 0)
    o2 = W2 @ h1
    h3 = np.maximum(W3 @ np.concatenate([sh, o2[1:]]), 0)
    h4 = np.maximum(W4 @ h3, 0)
    return math.exp(o2[0]), 1 / (1 + np.exp(-(W5 @ h4)))


def test_mlp_matches_numpy(oracle, scene1):
    _, w, _ = scene1
    rng = np.random.default_rng(2)
    for _ in range(20):
        feat = rng.normal(0, 0.3, 32)
        d = rng.normal(size=3)
        sh = oracle.sh4(d / np.linalg.norm(d))
        s, rgb = oracle.mlp(w, feat, sh)
        s2, rgb2 = _np_mlp(w, feat, sh)
        assert abs(s - s2) <= 1e-12 * max(1, s2) and np.allclose(rgb, rgb2, atol=1e-13)


def test_scene_statistics(oracle, scene1):
    t, w, o = scene1
    frac = oracle.lib().orc_occupancy_fraction(oracle.ptr(o, oracle.C.c_uint32))
    assert 0.28 < frac < 0.32
    assert abs(float(t[:200000].std()) - 0.1) < 0.002
    a1 = math.sqrt(6 / 96)
    assert np.abs(w[:2048]).max() <= a1 and np.abs(w[:2048]).max() > 0.99 * a1


# ---------------------------------------------------------------- march and composite

def _const_sigma_net(sigma):
    """Constant table 1.0 with W1 = a, W2[0] = b gives sigma = exp(64 b * 32 a)."""
    w = np.zeros(9344, np.float32)
    a = 1.0 / 32.0
    w[0:2048] = a
    b = math.log(sigma) / 64.0
    w[2048:2048 + 64] = b
    return w


def _axis_ray(y=0.5, z=0.5):
    return np.array([-1.0, y, z], np.float32), np.array([1.0, 0.0, 0.0], np.float32)


def _ray(oracle, table, w, occ, o, d):
    return oracle.render_ray(table, w, occ, o, d)


def test_empty_occupancy_gives_background(oracle, scene1):
    t, w, _ = scene1
    occ = np.zeros(65536, np.uint32)
    o, d = _axis_ray()
    rgb, depth, T, n = _ray(oracle, t, w, occ, o, d)
    assert n == 0 and T == 1.0 and depth == 0.0 and np.all(rgb == 0.0)


def test_full_occupancy_step_counts(oracle, scene1):
    t, w, _ = scene1
    occ = np.full(65536, 0xFFFFFFFF, np.uint32)
    # axis-aligned chord of length 1: steps k with (k + 0.5) delta < 1 -> 74
    o, d = _axis_ray()
    assert _ray(oracle, t, w, occ, o, d)[3] == int(1 / float(oracle.DELTA) + 0.5)
    # main diagonal chord sqrt(3) = 128 delta -> the 128-step cap
    o = np.array([-0.5, -0.5, -0.5], np.float32)
    d = (np.ones(3) / math.sqrt(3)).astype(np.float32)
    assert _ray(oracle, t, w, occ, o, d)[3] == 128


def test_uniform_slab_closed_form(oracle):
    """Zero network: sigma = 1, colour 0.5 everywhere.  A full-occupancy slab of
    n samples gives T = exp(-n delta), C = 0.5 (1 - T) and the geometric-series
    expected depth."""
    table = np.zeros(oracle.N_TABLE * 2, np.float32)
    w = np.zeros(9344, np.float32)
    occ = np.full(65536, 0xFFFFFFFF, np.uint32)
    o, d = _axis_ray()
    rgb, depth, T, n = _ray(oracle, table, w, occ, o, d)
    dl = float(oracle.DELTA)
    assert n == 74
    assert abs(T - math.exp(-n * dl)) < 1e-14
    assert np.allclose(rgb, 0.5 * (1 - math.exp(-n * dl)), atol=1e-14)
    # sum_i r^i (1-r) t_i, t_i = t0 + (i + 1/2) delta, r = e^-delta, t0 = 1
    r = math.exp(-dl)
    a = 1.0 + 0.5 * dl
    s0 = (1 - r ** n) / (1 - r)
    s1 = (r - n * r ** n + (n - 1) * r ** (n + 1)) / (1 - r) ** 2
    assert abs(depth - (1 - r) * (a * s0 + dl * s1)) < 1e-5


def test_early_termination_count(oracle):
    """Constant sigma: T < 1e-4 first holds after ceil(ln(1e4) / (sigma delta)) samples."""
    table = np.full(oracle.N_TABLE * 2, 1.0, np.float32)
    occ = np.full(65536, 0xFFFFFFFF, np.uint32)
    o, d = _axis_ray()
    dl = float(oracle.DELTA)
    for sigma in (300.0, 1000.0, 5000.0):
        w = _const_sigma_net(sigma)
        rgb, depth, T, n = _ray(oracle, table, w, occ, o, d)
        s_eff = math.exp(64 * float(w[2048]) * 32 * float(w[0]))
        assert n == math.floor(math.log(1e4) / (s_eff * dl)) + 1
        assert T < 1e-4


def test_single_opaque_sample(oracle):
    """One occupied cell and an opaque density: C = colour, D = t of that sample."""
    table = np.full(oracle.N_TABLE * 2, 1.0, np.float32)
    w = _const_sigma_net(1e6)
    occ = np.zeros(65536, np.uint32)
    cell = 64 + 128 * (64 + 128 * 64)
    occ[cell >> 5] |= np.uint32(1 << (cell & 31))
    o, d = _axis_ray(y=64.5 / 128, z=64.5 / 128)
    rgb, depth, T, n = _ray(oracle, table, w, occ, o, d)
    assert n == 1 and np.allclose(rgb, 0.5, atol=1e-12)
    # the sample is the first step whose position falls in cell x = 64
    dl = float(oracle.DELTA)
    k = next(k for k in range(128) if math.floor((-1.0 + 1.0 + (k + 0.5) * dl) * 128) == 64)
    assert abs(depth - (1.0 + (k + 0.5) * dl)) < 1e-6


def test_camera_conventions(oracle):
    """The initial reference camera (camera_manager.h:18-19) sits at the cube
    centre looking along -z; the centre pixel's ray is the forward column."""
    c2w = [1, 0, 0, .5, 0, -1, 0, .5, 0, 0, -1, .5]
    cam = oracle.camera(640, 480, c2w)
    o, d = oracle.camera_rays(cam, [240 * 640 + 320])
    # pixel (320, 240) centre is half a pixel off the optical axis
    assert np.allclose(o[0], [0.5, 0.5, 0.5])
    assert d[0][2] < -0.9999 and d[0][0] > 0 and d[0][1] < 0
    # focal length: vertical FOV 60 degrees
    f = oracle.lib().orc_focal(480, 60.0)
    assert abs(f - 240 / math.tan(math.radians(30))) < 1e-3
    # orbit camera looks at the cube centre
    m = oracle.orbit_c2w(30, 20).reshape(3, 4)
    fwd, pos = m[:, 2], m[:, 3]
    assert np.allclose(pos + 2.0 * fwd, 0.5, atol=1e-6)
    assert abs(np.linalg.norm(pos - 0.5) - 2.0) < 1e-6
    assert np.allclose(m[:, :3].T @ m[:, :3], np.eye(3), atol=1e-6)


def test_adam_one_step_closed_form(oracle):
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 2.0])
    m = np.zeros(3)
    v = np.zeros(3)
    oracle.adam_step(p, g, m, v, 1, lr=1e-2, eps=1e-15)
    # bias-corrected first step is lr * sign(g)
    assert np.allclose(p, [1.0 - 1e-2, -2.0 + 1e-2, 0.5 - 1e-2], atol=1e-12)
