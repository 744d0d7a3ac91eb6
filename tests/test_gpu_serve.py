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
    with nes.Server(gpu_ctx, max_inflight=1) as srv:
        busy = srv.submit(0, stereo(nes, 2064, 2208, 10.0))
        first = srv.submit(1, stereo(nes, 640, 480, 50.0, index=1))
        second = srv.submit(1, stereo(nes, 640, 480, 60.0, index=2))
        rc1, _, _ = srv.wait(first)
        rc2, frames, res = srv.wait(second)
        assert srv.wait(busy)[0] == 0
        st = srv.stats()
    assert rc2 == 0 and res.index == 2 and len(frames) == 2
    assert rc1 in (0, nes.NES_ECANCELED)
    assert st.superseded == (1 if rc1 == nes.NES_ECANCELED else 0)


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


def slot_bytes(rays, precision_fp32=False):
    """Device bytes of one batch slot: per ray slot RayData 32 + RayAcc 32 +
    counts/offsets 8 + FP16 SH 32 + single-pass record 8 + outputs 21 + YUV 2;
    per sample (4 per ray slot, at
    least 2^20) record 4 + position 16 + MLP output 16 + 8 feature planes
    (8 B FP16, 16 B FP32); 1/8 buffer growth slack; 1 MiB."""
    cap = max(4 * rays, 1 << 20)
    per_sample = 4 + 16 + 16 + (16 if precision_fp32 else 8) * 8
    return (rays * (32 + 32 + 8 + 32 + 8 + 21 + 2) + cap * per_sample) // 8 * 9 + (1 << 20)


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
