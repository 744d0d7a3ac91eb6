"""The oracle's analytic gradient (compositing, MLPs, hash-grid scatter)
against central finite differences in FP64."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def batch(oracle, scene1):
    t, w, o = scene1
    ro, rd, _, _ = oracle.training_rays(0, 0, 4000)
    keep = [i for i in range(len(ro)) if oracle.render_ray(t, w, o, ro[i], rd[i])[3] > 5][:40]
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (len(keep), 3)).astype(np.float32)
    return ro[keep], rd[keep], target


def test_gradient_matches_finite_differences(oracle, scene1, batch):
    t, w, o = scene1
    ro, rd, target = batch
    loss, gt, gw = oracle.loss_grad(t, w, o, ro, rd, target)
    assert loss > 0
    rng = np.random.default_rng(4)
    touched = np.flatnonzero(gt)
    assert len(touched) > 1000
    tab_idx = rng.choice(touched, 48, replace=False)
    # MLP weights from every layer
    w_idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False) for a, b in
                            [(0, 2048), (2048, 3072), (3072, 5056), (5056, 9152), (9152, 9344)]])
    h = 1e-5
    for i in tab_idx:
        fd = (oracle.loss_perturbed(t, w, o, ro, rd, target, int(i), -1, h) -
              oracle.loss_perturbed(t, w, o, ro, rd, target, int(i), -1, -h)) / (2 * h)
        assert abs(fd - gt[i]) <= 1e-4 * abs(gt[i]) + 2e-11, (i, fd, gt[i])
    for i in w_idx:
        fd = (oracle.loss_perturbed(t, w, o, ro, rd, target, -1, int(i), h) -
              oracle.loss_perturbed(t, w, o, ro, rd, target, -1, int(i), -h)) / (2 * h)
        assert abs(fd - gw[i]) <= 1e-4 * abs(gw[i]) + 2e-11, (i, fd, gw[i])
    # untouched table entries have exactly zero gradient
    untouched = rng.choice(np.flatnonzero(gt == 0), 8, replace=False)
    for i in untouched:
        assert oracle.loss_perturbed(t, w, o, ro, rd, target, int(i), -1, 1e-3) == loss


def test_training_rays_are_the_sampler(oracle):
    """Rays of a step come from the seeded poses: origin on the radius-2 orbit,
    unit directions, same rays for the same (seed, step)."""
    ro, rd, pose, pix = oracle.training_rays(0, 5, 512)
    ro2, rd2, _, _ = oracle.training_rays(0, 5, 512)
    assert np.array_equal(ro, ro2) and np.array_equal(rd, rd2)
    assert np.allclose(np.linalg.norm(ro - 0.5, axis=1), 2.0, atol=1e-5)
    assert np.allclose(np.linalg.norm(rd, axis=1), 1.0, atol=1e-6)
    assert pose.max() < 16 and pix.max() < 640 * 480


def test_relu_margin(oracle, scene1):
    """The kink margin is +inf for a ray that misses the cube and never above
    the smallest |layer-1 pre-activation| of the ray's first sample."""
    t, w, o = scene1
    miss = oracle.ray_relu_margin(t, w, o, np.array([5.0, 5.0, 5.0], np.float32),
                                  np.array([1.0, 0.0, 0.0], np.float32))
    assert miss == np.inf
    ro, rd, _, _ = oracle.training_rays(0, 0, 400)
    checked = 0
    for i in range(len(ro)):
        if oracle.render_ray(t, w, o, ro[i], rd[i])[3] == 0:
            continue
        m = oracle.ray_relu_margin(t, w, o, ro[i], rd[i])
        # first occupied sample position, by marching like the oracle
        x = first_sample(oracle, o, ro[i], rd[i])
        d1 = w[:2048].astype(np.float64).reshape(64, 32) @ oracle.encode(t, np.clip(x, 0, 1))
        assert 0 < m <= np.abs(d1).min()
        checked += 1
        if checked == 5:
            break
    assert checked == 5


def first_sample(oracle, occ, o, d):
    import ctypes as C
    L = oracle.lib()
    f32 = C.POINTER(C.c_float)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    ts, te = C.c_float(), C.c_float()
    n = L.orc_ray_box(o.ctypes.data_as(f32), d.ctypes.data_as(f32), C.byref(ts), C.byref(te))
    for k in range(n):
        t = L.orc_step_t(k, ts.value)
        x = np.zeros(3, np.float32)
        step_pos(L, o, d, t, x)
        if L.orc_occupied(occ.ctypes.data_as(C.POINTER(C.c_uint32)), x.ctypes.data_as(f32)):
            return x.astype(np.float64)
    raise AssertionError("ray has no sample")


def step_pos(L, o, d, t, x):
    import ctypes as C
    f32 = C.POINTER(C.c_float)
    L.orc_step_pos.restype = None
    L.orc_step_pos.argtypes = [f32, f32, C.c_float, f32]
    L.orc_step_pos(o.ctypes.data_as(f32), d.ctypes.data_as(f32), t, x.ctypes.data_as(f32))
