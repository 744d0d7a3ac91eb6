"""Continuous-batching server: frames are identical alone and in a mix,
equal to render + frame prep called directly, and superseding follows the
latest-pose-wins rule."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def stereo(nes, w, h, az, el=20.0, fov=60.0, index=0):
    return [nes.camera(w, h, nes.orbit_camera(az, el, 2.0, e), fov=fov, index=index,
                       is_left=int(e < 0)) for e in (-0.032, 0.032)]


def test_alone_equals_mix_equals_direct(nes, gpu_ctx):
    target = stereo(nes, 320, 240, 30.0)
    others = [stereo(nes, 640, 480, 100.0), stereo(nes, 144, 160, 200.0),
              stereo(nes, 206, 220, 300.0)]
    with nes.Server(gpu_ctx, max_inflight=1) as srv:
        rc, alone, _ = srv.wait(srv.submit(0, target))
        assert rc == 0
        # a mix: other clients queue up behind a busy slot, then the target
        tickets = [srv.submit(1 + i, v) for i, v in enumerate(others)]
        t = srv.submit(0, target)
        for tk in tickets:
            assert srv.wait(tk)[0] == 0
        rc, mixed, res = srv.wait(t)
        assert rc == 0 and res.rays == 2 * 320 * 240
    for (ya, ua, va), (yb, ub, vb) in zip(alone, mixed):
        assert np.array_equal(ya, yb) and np.array_equal(ua, ub) and np.array_equal(va, vb)
    outs, _ = gpu_ctx.render(target, precision=nes.PRECISION_MIXED)
    for o, (y, u, v) in zip(outs, alone):
        y2, u2, v2 = nes.frameprep(o["rgb24"])
        assert np.array_equal(y, y2) and np.array_equal(u, u2) and np.array_equal(v, v2)


def test_latest_pose_wins(nes, gpu_ctx):
    """A client's newer pose replaces its waiting one (admission held, so the
    first is certainly still waiting); other clients' requests are kept."""
    with nes.Server(gpu_ctx, max_inflight=1) as srv:
        srv.hold(True)
        busy = srv.submit(0, stereo(nes, 2064, 2208, 10.0))
        first = srv.submit(1, stereo(nes, 640, 480, 50.0, index=1))
        second = srv.submit(1, stereo(nes, 640, 480, 60.0, index=2))
        srv.hold(False)
        rc1, _, _ = srv.wait(first)
        rc2, frames, res = srv.wait(second)
        assert srv.wait(busy)[0] == 0
        st = srv.stats()
    assert rc2 == 0 and res.index == 2 and len(frames) == 2
    assert rc1 == nes.NES_ECANCELED
    assert st.superseded == 1 and st.completed == 2


def test_invalid_requests_rejected(nes, gpu_ctx):
    with nes.Server(gpu_ctx) as srv:
        with pytest.raises(nes.NesError):
            srv.submit(0, [nes.camera(641, 480, nes.orbit_camera(0, 0))])


def test_poisson_clients_short_run(nes, gpu_ctx):
    with nes.Server(gpu_ctx) as srv:
        st = srv.simulate(n_clients=8, rate_hz=72.0, duration_s=0.5, seed=7)
    print(st.as_dict())
    assert st.completed > 0 and st.failed == 0
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
    assert st.mrays_per_s > 0


def slot_bytes(rays, precision_fp32=False, depth=False):
    """Device bytes of one batch slot: per ray slot RayData 32 + RayAcc 32 +
    counts/offsets 8 + FP16 SH 32 + single-pass record 8 + outputs 21 + YUV 2
    (4 with the depth stream); per sample (4 per ray slot, at least 2^20)
    record 4 + position 16 + MLP output 16 + 8 feature planes (8 B FP16,
    16 B FP32); 1/8 buffer growth slack; 1 MiB."""
    cap = max(4 * rays, 1 << 20)
    per_sample = 4 + 16 + 16 + (16 if precision_fp32 else 8) * 8
    per_ray = 32 + 32 + 8 + 32 + 8 + 21 + (4 if depth else 2)
    return (rays * per_ray + cap * per_sample) // 8 * 9 + (1 << 20)


def test_hbm_budget_decides_batch_slots(nes, gpu_ctx):
    """Admission by memory: the number of concurrent batch slots is what the
    budget affords, and a budget below one slot is refused."""
    small_rays = 1 << 20
    with nes.Server(gpu_ctx, max_batch_rays=small_rays, max_inflight=8) as srv:
        assert srv.stats().slots == 8
    per_slot = slot_bytes(small_rays)
    with nes.Server(gpu_ctx, max_batch_rays=small_rays, max_inflight=8,
                    hbm_budget=2 * per_slot + 1) as srv:
        assert srv.stats().slots == 2
        t = srv.submit(0, stereo(nes, 320, 240, 40.0))
        assert srv.wait(t)[0] == 0
    with nes.Server(gpu_ctx, max_batch_rays=small_rays, max_inflight=8,
                    hbm_budget=2 * slot_bytes(small_rays, True) + 1,
                    precision=nes.PRECISION_FP32) as srv:
        assert srv.stats().slots == 2
    with pytest.raises(nes.NesError) as e:
        nes.Server(gpu_ctx, max_batch_rays=small_rays, hbm_budget=per_slot - 1)
    assert e.value.code == -12


def test_alone_equals_mix_at_headset_resolutions(nes, gpu_ctx):
    """Config-4 per-eye sizes: the target request is rendered alone, then in
    one batch with two other clients' requests (admission held until all
    three wait, so the batch composition is certain); its frames are
    bit-identical both times."""
    target = stereo(nes, 1440, 1600, 30.0)
    others = [stereo(nes, 640, 480, 100.0), stereo(nes, 2064, 2208, 200.0)]
    with nes.Server(gpu_ctx, max_inflight=1) as srv:
        rc, alone, r0 = srv.wait(srv.submit(0, target))
        assert rc == 0 and r0.batch_requests == 1
        srv.hold(True)
        tickets = [srv.submit(1 + i, v) for i, v in enumerate(others)]
        t = srv.submit(0, target)
        srv.hold(False)
        rest = [srv.wait(tk) for tk in tickets]
        rc, mixed, res = srv.wait(t)
    assert rc == 0 and all(r[0] == 0 for r in rest)
    assert res.batch_requests == 3 and all(r[2].batch == res.batch for r in rest)
    assert res.batch != r0.batch
    for (ya, ua, va), (yb, ub, vb) in zip(alone, mixed):
        assert np.array_equal(ya, yb) and np.array_equal(ua, ub) and np.array_equal(va, vb)


def test_superseded_latency_reported(nes, gpu_ctx):
    """A superseded request counts in the all-requests latency with the time
    to the frame that replaced it."""
    with nes.Server(gpu_ctx, max_inflight=1) as srv:
        srv.hold(True)
        a = srv.submit(1, stereo(nes, 320, 240, 10.0, index=1))
        b = srv.submit(1, stereo(nes, 320, 240, 20.0, index=2))
        srv.hold(False)
        rca = srv.wait(a)[0]
        rcb, _, res = srv.wait(b)
        st = srv.stats()
    assert rca == nes.NES_ECANCELED and rcb == 0 and res.index == 2
    assert st.superseded == 1 and st.completed == 1
    assert st.p99_all_ms >= st.p99_ms > 0


def test_depth_stream_into_encoder_planes(nes, oracle, gpu_ctx):
    """Depth-stream server at 2064 x 2208: scene and depth YUV420P per eye,
    written into caller-owned pinned planes with the encoder's linesizes
    (2080 / 1056) and into the server's own planes; both equal the direct
    render's frame prep and the oracle's GRAY8 depth conversion."""
    w, h = 2064, 2208
    views = stereo(nes, w, h, 40.0)
    P = [nes.HostPlanes(w, h) for _ in views]
    D = [nes.HostPlanes(w, h) for _ in views]
    try:
        with nes.Server(gpu_ctx, depth_stream=True, depth_far=3.0) as srv:
            rc, scene, depth, res = srv.wait(srv.submit(0, views), depth=True)
            rc2, scene2, depth2, res2 = srv.wait(srv.submit_into(0, views, P, D), depth=True)
        assert rc == 0 and rc2 == 0
        assert list(res2.linesize[0]) == [2080, 1056, 1056]
        assert list(res2.depth_linesize[1]) == [2080, 1056, 1056]
        outs, _ = gpu_ctx.render(views, precision=nes.PRECISION_MIXED, depth_far=3.0)
        for i, o in enumerate(outs):
            want = nes.frameprep(o["rgb24"])
            dwant = oracle.gray8_to_yuv420p(oracle.depth_gray8(o["depth"], 3.0))
            for got in (scene[i], scene2[i], P[i].yuv()):
                assert all(np.array_equal(g, r) for g, r in zip(got, want))
            for got in (depth[i], depth2[i], D[i].yuv()):
                assert all(np.array_equal(g, r) for g, r in zip(got, dwant))
            assert P[i].padding_intact() and D[i].padding_intact()
    finally:
        for p in P + D:
            p.free()


def _dense_scene():
    """Full occupancy, sigma = 1: about 100 samples per ray, none terminated."""
    table = np.full(6098925 * 2, 1.0, np.float32)
    w = np.zeros(9344, np.float32)
    w[0:2048] = 1.0 / 32.0
    occ = np.full(65536, 0xFFFFFFFF, np.uint32)
    return table, w, occ


def test_dense_scene_stays_within_slot_budget(nes):
    """A batch that marches ~100 samples per ray on a slot whose share of the
    budget holds 4: it is rendered in bounded passes (no regrowth past the
    budget, no failure) and its frames equal the direct render's."""
    scene = _dense_scene()
    views = stereo(nes, 320, 240, 25.0)
    rays = 1 << 18
    with nes.Context(*scene) as ctx:
        outs, st0 = ctx.render(views)
        with nes.Server(ctx, max_batch_rays=rays, max_inflight=1,
                        hbm_budget=slot_bytes(rays) + 1) as srv:
            rc, planes, res = srv.wait(srv.submit(0, views))
            st = srv.stats()
    print(f"direct {st0.samples} samples; served in {res.batch_passes} bounded passes")
    assert rc == 0 and st.failed == 0 and st.bounded_batches == 1 and res.batch_passes > 2
    for o, (y, u, v) in zip(outs, planes):
        y2, u2, v2 = nes.frameprep(o["rgb24"])
        assert np.array_equal(y, y2) and np.array_equal(u, u2) and np.array_equal(v, v2)


def test_scene_update_under_load_never_mixes_scenes(nes):
    """Scene updates while batches are in flight: every frame equals one
    whole scene's frame, and requests submitted after an update returns see
    the new scene."""
    A = nes.scene_synthesize(1)
    B = nes.scene_synthesize(2)
    reqs = [stereo(nes, 256, 192, 15.0 * i, index=i) for i in range(12)]

    def frames(scene):
        out = []
        with nes.Context(*scene) as c:
            for v in reqs:
                o, _ = c.render(v)
                out.append([nes.frameprep(x["rgb24"]) for x in o])
        return out

    ref = {"A": frames(A), "B": frames(B)}
    with nes.Context(*A) as ctx:
        with nes.Server(ctx, max_inflight=3) as srv:
            tickets, after = [], {}
            for i, v in enumerate(reqs):
                if i == 4:
                    ctx.update_scene(*B)
                    after.update({j: "B" for j in range(4, 8)})
                if i == 8:
                    ctx.update_scene(*A)
                    after.update({j: "A" for j in range(8, 12)})
                tickets.append(srv.submit(i, v))
            got = [srv.wait(t) for t in tickets]
    for i, (rc, planes, _) in enumerate(got):
        assert rc == 0
        same = {k: all(all(np.array_equal(g, r) for g, r in zip(p, q))
                       for p, q in zip(planes, ref[k][i])) for k in ("A", "B")}
        assert same["A"] or same["B"], i
        if i in after:
            assert same[after[i]], (i, after[i])
