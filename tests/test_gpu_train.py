"""Differentiable render on the GPU: forward and gradients against the FP64
oracle (itself pinned by finite differences), the autograd binding, Adam's
closed-form first step and a loss that falls when training on 2^18-ray
batches from 16 poses.

Gradient tolerances: a ReLU whose pre-activation lies closer to zero than the
FP32 forward's rounding error can switch sides, and the FP32 gradient then
differs from the FP64 one by that unit's whole term for that sample (measured:
the two outlier rays of 60 had margins 2e-7 and 1.2e-6).  Element-wise
comparisons therefore use rays whose smallest |pre-activation| over every
evaluated sample (the oracle's kink margin) exceeds 1e-5, a thousand times the
FP32 error of the pre-activations; whole-batch comparisons use relative L2
norms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


KINK_MARGIN = 1e-5


@pytest.fixture(scope="module")
def rays(oracle, scene1):
    """3000 training rays; keep = those with samples and a kink margin above
    KINK_MARGIN."""
    t, w, o = scene1
    ro, rd, _, _ = oracle.training_rays(0, 0, 3000)
    ns = np.array([oracle.render_ray(t, w, o, ro[i], rd[i])[3] for i in range(len(ro))])
    hit = np.flatnonzero(ns > 0)
    margin = np.array([oracle.ray_relu_margin(t, w, o, ro[i], rd[i]) for i in hit])
    keep = hit[margin > KINK_MARGIN]
    assert len(keep) >= 100
    return ro, rd, keep


@pytest.fixture(scope="module")
def trainer(nes, scene1):
    t, w, o = scene1
    with nes.Trainer(t, w, o, rays_per_batch=4096) as tr:
        yield tr


def test_forward_matches_oracle(oracle, scene1, trainer, rays):
    t, w, o = scene1
    trainer.set_params(t, w)  # whatever ran before on the shared trainer
    ro, rd, _ = rays
    rgb = trainer.forward(ro, rd)
    ref = np.array([oracle.render_ray(t, w, o, ro[i], rd[i])[0] for i in range(len(ro))])
    assert np.abs(rgb - ref).max() < 2e-5


def test_backward_matches_oracle_gradient(oracle, scene1, trainer, rays):
    t, w, o = scene1
    trainer.set_params(t, w)
    ro, rd, keep = rays
    ro, rd = ro[keep], rd[keep]
    target = np.random.default_rng(5).uniform(0, 1, (len(ro), 3)).astype(np.float32)
    loss, gt_ref, gw_ref = oracle.loss_grad(t, w, o, ro, rd, target)
    rgb = trainer.forward(ro, rd)
    d_rgb = 2.0 * (rgb - target) / (3 * len(ro))
    gt, gw = trainer.backward(ro, rd, d_rgb)
    layers = [(0, 2048), (2048, 3072), (3072, 5056), (5056, 9152), (9152, 9344)]
    for a, b in layers + [(None, None)]:
        x, r = (gw[a:b], gw_ref[a:b]) if a is not None else (gt, gt_ref)
        rel = np.linalg.norm(x - r) / np.linalg.norm(r)
        mx = np.abs(x - r).max() / np.abs(r).max()
        print(f"grad {a}:{b} rel-l2 {rel:.2e} max/scale {mx:.2e}", flush=True)
    # MLP weights: every layer within 1e-4 of its largest oracle gradient
    for a, b in layers:
        scale = np.abs(gw_ref[a:b]).max()
        assert np.abs(gw[a:b] - gw_ref[a:b]).max() <= 1e-4 * scale, (a, b)
    # table: outside the oracle's support only FP32 cancellation residue, and
    # values within 1e-4 of the largest gradient
    scale = np.abs(gt_ref).max()
    assert np.abs(gt[gt_ref == 0]).max() <= 1e-6 * scale
    assert np.abs(gt - gt_ref).max() <= 1e-4 * scale
    assert np.linalg.norm(gt - gt_ref) <= 1e-4 * np.linalg.norm(gt_ref)


def test_autograd_binding(nes, oracle, scene1, trainer, rays):
    torch = pytest.importorskip("torch")
    from nes_amd.autograd import render_rays
    t, w, o = scene1
    ro, rd, keep = rays
    ro, rd = ro[keep[:64]], rd[keep[:64]]
    table = torch.tensor(t, requires_grad=True)
    weights = torch.tensor(w, requires_grad=True)
    target = torch.rand(len(ro), 3, generator=torch.Generator().manual_seed(0))
    rgb = render_rays(table, weights, trainer, ro, rd)
    loss = ((rgb - target) ** 2).mean()
    loss.backward()
    _, gt_ref, gw_ref = oracle.loss_grad(t, w, o, ro, rd, target.numpy())
    assert np.abs(weights.grad.numpy() - gw_ref).max() <= 1e-3 * np.abs(gw_ref).max()
    assert np.abs(table.grad.numpy() - gt_ref).max() <= 1e-3 * np.abs(gt_ref).max()
    # central differences through the GPU forward on the largest weight gradients
    base = weights.detach().clone()
    for i in np.argsort(-np.abs(gw_ref))[:4]:
        h = 1e-2 * max(abs(float(base[i])), 1e-2)
        vals = []
        for sgn in (1, -1):
            wp = base.clone()
            wp[i] += sgn * h
            with torch.no_grad():
                vals.append(float(((render_rays(table, wp, trainer, ro, rd) - target) ** 2).mean()))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - gw_ref[i]) <= 0.05 * abs(gw_ref[i]), (i, fd, gw_ref[i])


def test_rays_without_samples(oracle, scene1, trainer, rays):
    """Rays that miss the cube (and one that grazes outside it) evaluate no
    sample: black, zero gradients everywhere, and the next backward on real
    rays is unaffected (the bucket sort and slice scatter of an empty step
    leave nothing behind)."""
    t, w, o = scene1
    trainer.set_params(t, w)  # the autograd tests leave perturbed weights behind
    n = 37
    ro = np.tile(np.float32([[3.0, 3.0, 3.0]]), (n, 1))
    rd = np.tile(np.float32([[1.0, 0.0, 0.0]]), (n, 1))  # pointing away from the cube
    rd[-1] = [0.0, 0.0, -1.0]                              # passes beside it
    for i in range(n):
        assert oracle.render_ray(t, w, o, ro[i], rd[i])[3] == 0
    assert not trainer.forward(ro, rd).any()
    gt, gw = trainer.backward(ro, rd, np.ones((n, 3), np.float32))
    assert not gt.any() and not gw.any()
    ro2, rd2, keep = rays
    ro2, rd2 = ro2[keep[:50]], rd2[keep[:50]]
    target = np.random.default_rng(9).uniform(0, 1, (len(ro2), 3)).astype(np.float32)
    _, gt_ref, gw_ref = oracle.loss_grad(t, w, o, ro2, rd2, target)
    d_rgb = 2.0 * (trainer.forward(ro2, rd2) - target) / (3 * len(ro2))
    gt, gw = trainer.backward(ro2, rd2, d_rgb)
    assert np.abs(gw - gw_ref).max() <= 1e-4 * np.abs(gw_ref).max()
    assert np.abs(gt - gt_ref).max() <= 1e-4 * np.abs(gt_ref).max()


SLICE_ENTRIES = 16384  # entries per LDS slice of the table-gradient scatter
DENSE_RES = [16, 22, 30, 42, 58]
DENSE_OFFSET = [0, 4913, 17080, 46871, 126378]  # packed level offsets


def test_dense_slice_boundaries(nes, oracle, scene1):
    """A dense level's corner pair whose x corner is the last entry of one
    LDS slice of the scatter and whose x + 1 corner is the first of the next:
    each slice's workgroup adds one of them.  Rays through every such cell of
    levels 2-4 (occupancy only there, so a ray has a few samples and clears
    the kink margin), gradients against the oracle at those entries and
    everywhere."""
    t, w, _ = scene1
    cells = []
    for lv in (2, 3, 4):
        n1 = DENSE_RES[lv] + 1
        for k in range(1, n1 ** 3 // SLICE_ENTRIES + 1):
            z, rem = divmod(k * SLICE_ENTRIES - 1, n1 * n1)
            y, x = divmod(rem, n1)
            if max(x, y, z) < DENSE_RES[lv]:  # pair (y, z) of cell (x, y, z)
                cells.append((lv, x, y, z))
    assert len(cells) >= 10
    occ_bits = np.zeros(128 ** 3, bool)
    for lv, x, y, z in cells:
        lo = [int(np.floor(v / DENSE_RES[lv] * 128)) for v in (x, y, z)]
        hi = [int(np.ceil((v + 1) / DENSE_RES[lv] * 128)) for v in (x, y, z)]
        for cz in range(lo[2], hi[2]):
            for cy in range(lo[1], hi[1]):
                row = (cz * 128 + cy) * 128
                occ_bits[row + lo[0]:row + hi[0]] = True
    occ = np.packbits(occ_bits, bitorder="little").view(np.uint32).copy()
    rng = np.random.default_rng(11)
    ro, rd = [], []
    for lv, x, y, z in cells:
        for _ in range(6):  # a chord through the cell's middle is longer than a step
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p = (np.array([x, y, z]) + rng.uniform(0.3, 0.7, 3)) / DENSE_RES[lv]
            ro.append(p - 2.0 * d)
            rd.append(d)
    ro, rd = np.array(ro, np.float32), np.array(rd, np.float32)
    margin = np.array([oracle.ray_relu_margin(t, w, occ, ro[i], rd[i]) for i in range(len(ro))])
    keep = margin > KINK_MARGIN
    assert keep.sum() >= 3 * len(cells)
    ro, rd = ro[keep], rd[keep]
    target = np.random.default_rng(6).uniform(0, 1, (len(ro), 3)).astype(np.float32)
    _, gt_ref, gw_ref = oracle.loss_grad(t, w, occ, ro, rd, target)
    with nes.Trainer(t, w, occ, rays_per_batch=1024) as tr:
        rgb = tr.forward(ro, rd)
        gt, gw = tr.backward(ro, rd, 2.0 * (rgb - target) / (3 * len(ro)))
    scale = np.abs(gt_ref).max()
    assert np.abs(gt - gt_ref).max() <= 1e-4 * scale
    assert np.abs(gw - gw_ref).max() <= 1e-4 * np.abs(gw_ref).max()
    # the straddling entries themselves carry gradient, and it matches
    hit = 0
    for lv, x, y, z in cells:
        n1 = DENSE_RES[lv] + 1
        e0 = DENSE_OFFSET[lv] + x + (y + z * n1) * n1
        idx = [2 * e0, 2 * e0 + 1, 2 * e0 + 2, 2 * e0 + 3]  # both features of e0 and e0 + 1
        if np.abs(gt_ref[idx]).min() > 0:
            hit += 1
            assert np.abs(gt[idx] - gt_ref[idx]).max() <= 1e-4 * scale, (lv, x, y, z)
    assert hit >= len(cells) - 2


def poses_and_targets(nes, oracle, scene1, w=640, h=480):
    t, wt, o = scene1
    poses = np.stack([oracle.training_pose(0, i) for i in range(16)])
    with nes.Context(t, wt, o) as ctx:
        outs, _ = ctx.render([nes.camera(w, h, p) for p in poses], precision=nes.PRECISION_FP32)
    return poses, np.stack([x["rgb"] for x in outs])


def test_batch_loss_matches_oracle_and_adam_first_step(nes, oracle, scene1):
    t, w, o = scene1
    poses, targets = poses_and_targets(nes, oracle, scene1)
    ts, ws, _ = oracle.scene(4, table_init=1, occ=False)
    n = 2048
    with nes.Trainer(ts, ws, o, poses=poses, targets=targets, width=640, height=480,
                     rays_per_batch=n, seed=0) as tr:
        loss, gt, gw = tr.batch_grad(0)
        ro, rd, pose, pix = oracle.training_rays(0, 0, n)
        tgt = targets.reshape(16, -1, 3)[pose, pix]
        loss_ref, _, gw_ref = oracle.loss_grad(ts, ws, o, ro, rd, tgt)
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        # whole batch, kinks included: relative L2 per layer.  The student's
        # layer-1 gradient is a sum of ~1e5 terms of random sign, so one
        # ReLU pre-activation on the other side of zero in FP32 than in FP64
        # would move it by ~1/sqrt(1e5) = 3e-3 of its norm; the forward is
        # deterministic, and on this batch none does (measured: <= 3e-6 per
        # layer, the atomics' summation order), so 1e-3 still catches one.
        for a, b in [(0, 2048), (2048, 3072), (3072, 5056), (5056, 9152), (9152, 9344)]:
            rel = np.linalg.norm(gw[a:b] - gw_ref[a:b]) / np.linalg.norm(gw_ref[a:b])
            print(f"batch grad {a}:{b} rel-l2 {rel:.2e}", flush=True)
            assert rel <= 1e-3, (a, b, rel)
        l0, st = tr.step()
        assert abs(l0 - loss) <= 1e-5 * loss and st.step == 1
        t1, w1 = tr.get_params()
    # first Adam step from zero moments: p -= lr * g / (|g| + eps).  The step
    # re-evaluates the gradient (atomics in another order), and where |g| is
    # near eps the update is steep in g, so those entries are only held to a
    # step no longer than lr.
    for p0, p1, g in ((ts, t1, gt), (ws, w1, gw)):
        want = p0 - 1e-2 * g / (np.abs(g) + 1e-15)
        firm = np.abs(g) > 1e-12
        assert np.abs(p1 - want)[firm].max() <= 1e-6
        assert np.abs(p1 - p0).max() <= 1e-2 * (1 + 1e-5)


def test_training_2e18_rays_16_poses_reduces_loss(nes, oracle, scene1):
    t, w, o = scene1
    poses, targets = poses_and_targets(nes, oracle, scene1)
    ts, ws, _ = oracle.scene(4, table_init=1, occ=False)
    with nes.Trainer(ts, ws, o, poses=poses, targets=targets, width=640, height=480,
                     rays_per_batch=1 << 18, seed=0) as tr:
        losses, ms = [], []
        for _ in range(40):
            l, st = tr.step()
            losses.append(l)
            ms.append(st.ms)
    print(f"loss {losses[0]:.5f} -> {losses[-1]:.5f}, {np.median(ms):.2f} ms/step, "
          f"{st.samples} samples/batch")
    assert np.isfinite(losses).all()
    assert np.mean(losses[-5:]) < 0.5 * np.mean(losses[:3])


def test_two_render_calls_in_one_graph(nes, oracle, scene1, trainer, rays):
    """Gradients through two render_rays calls of one autograd graph (and a
    forward of other rays between them and the backward) equal the oracle's
    gradient of the summed loss: each backward recomputes its own rays."""
    torch = pytest.importorskip("torch")
    from nes_amd.autograd import render_rays
    t, w, o = scene1
    ro, rd, keep = rays
    ia, ib = keep[:40], keep[40:100]
    table = torch.tensor(t, requires_grad=True)
    weights = torch.tensor(w, requires_grad=True)
    ta = torch.rand(len(ia), 3, generator=torch.Generator().manual_seed(1))
    tb = torch.rand(len(ib), 3, generator=torch.Generator().manual_seed(2))
    la = ((render_rays(table, weights, trainer, ro[ia], rd[ia]) - ta) ** 2).mean()
    lb = ((render_rays(table, weights, trainer, ro[ib], rd[ib]) - tb) ** 2).mean()
    trainer.forward(ro[keep[100:130]], rd[keep[100:130]])  # unrelated work in between
    (la + 2.0 * lb).backward()
    _, gta, gwa = oracle.loss_grad(t, w, o, ro[ia], rd[ia], ta.numpy())
    _, gtb, gwb = oracle.loss_grad(t, w, o, ro[ib], rd[ib], tb.numpy())
    gw_ref, gt_ref = gwa + 2.0 * gwb, gta + 2.0 * gtb
    assert np.abs(weights.grad.numpy() - gw_ref).max() <= 1e-3 * np.abs(gw_ref).max()
    assert np.abs(table.grad.numpy() - gt_ref).max() <= 1e-3 * np.abs(gt_ref).max()


def test_backward_rejects_mismatched_shapes(nes, trainer, rays):
    ro, rd, keep = rays
    with pytest.raises(ValueError):
        trainer.backward(ro[keep[:10]], rd[keep[:10]], np.zeros((9, 3), np.float32))


def test_checkpoint_round_trip(nes, oracle, scene1, tmp_path):
    """Save after some steps, load into a fresh trainer: parameters, Adam
    moments and the step count continue exactly (the next step's update from
    the restored state equals the uninterrupted one's up to the atomics'
    summation order); a truncated file is refused and changes nothing."""
    t, w, o = scene1
    poses, targets = poses_and_targets(nes, oracle, scene1)
    ts, ws, _ = oracle.scene(4, table_init=1, occ=False)
    kw = dict(poses=poses, targets=targets, width=640, height=480, rays_per_batch=8192, seed=0)
    path = str(tmp_path / "ckpt.bin")
    with nes.Trainer(ts, ws, o, **kw) as a:
        for _ in range(3):
            a.step()
        a.save(path)
        ta, wa = a.get_params()
        la, st_a = a.step()
        ta4, wa4 = a.get_params()
    with nes.Trainer(ts, ws, o, **kw) as b:
        b.load(path)
        tb, wb = b.get_params()
        assert np.array_equal(ta, tb) and np.array_equal(wa, wb)
        lb, st_b = b.step()
        tb4, wb4 = b.get_params()
        assert st_b.step == st_a.step == 4
        # same parameters, so the same batch loss (to the loss reduction's
        # atomic order); the updates agree except where a gradient near zero
        # changes with the scatter's summation order (then within 2 lr)
        assert abs(lb - la) <= 1e-5 * la
        for x, y in ((tb4, ta4), (wb4, wa4)):
            d = np.abs(x - y)
            assert (d > 1e-6).mean() < 1e-3 and d.max() <= 2.01e-2
        with open(path, "rb") as f:
            head = f.read(1000)
        bad = str(tmp_path / "short.bin")
        with open(bad, "wb") as f:
            f.write(head)
        with pytest.raises(nes.NesError) as e:
            b.load(bad)
        assert e.value.code == -22
        tc, wc = b.get_params()
        assert np.array_equal(tc, tb4) and np.array_equal(wc, wb4)
