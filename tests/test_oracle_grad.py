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
