"""Serving loop: admission by HBM budget with idle eviction (host logic, CPU) and an end-to-end run
of the continuous-batching loop on the GPU."""
import numpy as np
import torch
import pytest

from xdn.serving import Admission, StreamReq, config3, config4, onoff_arrivals, poisson_arrivals


def test_admission_fits_and_waits():
    a = Admission(budget=100)
    assert a.try_admit(1, 40, 0.0, set()) == (True, [])
    assert a.try_admit(2, 40, 0.0, set()) == (True, [])
    # no room and nothing idle long enough -> waits
    assert a.try_admit(3, 40, 0.01, set()) == (False, [])
    assert a.used == 80
    # larger than the whole budget never fits
    assert a.try_admit(4, 200, 1.0, set()) == (False, [])


def test_admission_evicts_idle_oldest_first_but_not_busy():
    a = Admission(budget=100)
    a.try_admit(1, 40, 0.0, set())
    a.try_admit(2, 40, 0.5, set())
    a.touch(1, 0.1)
    # stream 1 is the oldest idle one, but it is busy -> stream 2 is evicted instead
    ok, ev = a.try_admit(3, 40, 1.0, busy={1}, idle_after=0.05)
    assert ok and ev == [2] and a.used == 80 and set(a.resident) == {1, 3}
    # both idle: evict the least recently active first, only as many as needed
    a.touch(1, 2.0)
    a.touch(3, 1.5)
    ok, ev = a.try_admit(4, 30, 3.0, busy=set())
    assert ok and ev == [3] and set(a.resident) == {1, 4}


def test_arrival_processes():
    rng = np.random.default_rng(0)
    t = poisson_arrivals(60.0, 4000, rng=rng)
    gaps = np.diff([0.0] + t)
    assert abs(np.mean(gaps) - 1 / 60) < 1e-3 and np.all(gaps > 0)
    b = onoff_arrivals(120.0, 500, period_s=0.2, rng=rng)
    assert np.all(np.diff(b) >= 0)
    assert all(int(x // 0.2) % 2 == 0 for x in b)  # only during 'on' periods
    c3 = config3(frames_per_stream=8)
    assert len(c3) == 16 and sum(r.height == 1080 for r in c3) == 6
    c4 = config4(frames_per_stream=8)
    assert len(c4) == 16 and sum(r.height == 2160 for r in c4) == 4 and min(c4[1].arrivals) >= 0.25


@pytest.mark.gpu
def test_serving_loop_runs_all_frames(dev):
    from xdn.serving import Server
    rng = np.random.default_rng(1)
    reqs = [StreamReq(i, h, w, 6, poisson_arrivals(120.0, 6, rng=rng), seed=i)
            for i, (h, w) in enumerate([(64, 64), (48, 96), (96, 64), (64, 64)])]
    srv = Server(dev, hbm_budget=1 << 40, pool_frames=2)
    res = srv.run(reqs)
    assert res["frames"] == 24 and np.isfinite(res["p99_ms"]) and res["p99_ms"] > 0
    # a budget for two slots forces waits and evictions but every frame is still served
    from xdn import pipeline
    b2 = 2 * pipeline.slot_bytes(96, 64) + 1
    srv2 = Server(dev, hbm_budget=b2, pool_frames=2, idle_after=0.0)
    res2 = srv2.run(reqs)
    assert res2["frames"] == 24 and res2["evictions"] > 0


@pytest.mark.gpu
def test_serving_loop_with_unet_streams_under_eviction(dev):
    """U-Net streams: batched serving with a two-slot budget evicts and readmits (network and Adam
    state go to the host and back) and serves every frame with finite outputs."""
    from xdn import pipeline
    from xdn.serving import Server
    rng = np.random.default_rng(2)
    reqs = [StreamReq(i, h, w, 5, poisson_arrivals(150.0, 5, rng=rng), seed=i)
            for i, (h, w) in enumerate([(64, 64), (48, 96), (64, 64)])]
    b2 = 2 * pipeline.slot_bytes(64, 64, "unet") + 1
    srv = Server(dev, hbm_budget=b2, pool_frames=2, idle_after=0.0, network="unet")
    res = srv.run(reqs)
    assert res["frames"] == 15 and res["evictions"] > 0 and np.isfinite(res["p99_ms"])
    for d in srv.denoisers.values():
        assert d.network == "unet" and torch.isfinite(d.params).all()
