"""Serving loop: admission by HBM budget with idle eviction (host logic, CPU) and an end-to-end run
of the continuous-batching loop on the GPU."""
import numpy as np
import torch
import pytest

from xdn.serving import Admission, Server, StreamReq, config3, config4, onoff_arrivals, poisson_arrivals


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


class _Done:
    def query(self):
        return True


class HostServer(Server):
    """The serving loop's scheduling with the device work stubbed out: every slot costs 100 bytes
    and a launched step completes at once."""

    def _init_device(self, device):
        self.launched = []

    def _slot_bytes(self, r):
        return 100

    def _sync(self):
        pass

    def _frame_pool(self, req):
        return list(range(self.pool_frames))

    def _admit(self, r):
        return object()

    def _evict(self, sid):
        self.saved[sid] = self.denoisers.pop(sid)

    def _launch(self, slot, sids, frames):
        for s in sids:
            assert s in self.denoisers, f"stream {s} launched without a slot"
        self.launched.append(list(sids))
        return _Done()


def test_serving_never_evicts_a_stream_picked_for_the_batch():
    """Three streams with frames pending at t = 0 and room for two slots: the third stream must wait
    for the first two to go idle instead of evicting a stream already picked for the batch."""
    reqs = [StreamReq(i, 64, 64, 3, [0.0, 0.0, 0.0], seed=i) for i in range(3)]
    srv = HostServer(None, hbm_budget=200, pool_frames=2, idle_after=0.0)
    res = srv.run(reqs)
    assert res["frames"] == 9
    assert srv.launched[0] == [0, 1]
    for sid in range(3):  # every frame of every stream ran, in order of arrival
        assert sum(b.count(sid) for b in srv.launched) == 3
    assert res["evictions"] >= 1 and [s for s, _ in srv.admit_log][:3] == [0, 1, 2]


def test_min_batch_launches_a_light_load_one_stream_at_a_time():
    """Below min_batch ready streams each is launched on its own; at min_batch or more they share
    one launch (up to max_batch)."""
    light = [StreamReq(i, 64, 64, 2, [0.0, 0.0], seed=i) for i in range(2)]
    srv = HostServer(None, hbm_budget=1000, pool_frames=2, min_batch=3)
    assert srv.run(light)["frames"] == 4
    assert all(len(b) == 1 for b in srv.launched) and len(srv.launched) == 4
    heavy = [StreamReq(i, 64, 64, 2, [0.0, 0.0], seed=i) for i in range(5)]
    srv = HostServer(None, hbm_budget=1000, pool_frames=2, min_batch=3, max_batch=4)
    assert srv.run(heavy)["frames"] == 10
    assert srv.launched[0] == [0, 1, 2, 3] and max(len(b) for b in srv.launched) == 4
    # without batching min_batch changes nothing: one stream per launch
    srv = HostServer(None, hbm_budget=1000, pool_frames=2, min_batch=3, batching=False)
    assert srv.run(heavy)["frames"] == 10 and all(len(b) == 1 for b in srv.launched)


def test_max_batch_px_fills_a_batch_oldest_first_up_to_the_pixel_limit():
    """A batch takes streams oldest pending frame first while their pixels fit the limit; a stream
    that does not fit waits for the next batch, a frame larger than the limit runs alone."""
    # equal arrival times: the streams are taken in their order in the request list
    reqs = [StreamReq(0, 128, 128, 1, [0.0], seed=0)] + [StreamReq(i, 64, 64, 1, [0.0], seed=i) for i in (1, 2, 3)]
    srv = HostServer(None, hbm_budget=1000, pool_frames=2, max_batch_px=2 * 64 * 64)
    res = srv.run(reqs)
    assert res["frames"] == 4
    assert srv.launched == [[0], [1, 2], [3]]
    srv = HostServer(None, hbm_budget=1000, pool_frames=2)  # no limit: up to max_batch streams
    srv.run(reqs)
    assert max(len(b) for b in srv.launched) >= 2


def _replay_alone(dev, req, pool, admits, network, lr=1e-3):
    """Steps one stream frame by frame through Denoiser.step on the server's own frame buffers,
    starting a new slot at each frame index in `admits` with the network and Adam state carried
    over and the history restarted (what the server does on readmission)."""
    from xdn import pipeline
    d = None
    for k in range(req.n_frames):
        if k in admits:
            saved = None
            if d is not None:
                saved = (d.params.clone(), d.adam_m.clone(), d.adam_v.clone(), d.t)
                d.close()
            d = pipeline.Denoiser(req.height, req.width, device=dev, seed=1234 + req.sid, lr=lr, network=network)
            if saved is not None:
                for t, src in zip((d.params, d.adam_m, d.adam_v), saved[:3]):
                    t.copy_(src)
                d.set_step(saved[3])
                d.repack()
        d.step(pool[k % len(pool)])
    torch.cuda.synchronize()
    return d


def check_against_alone(dev, srv, reqs, network="cnn5"):
    """Each stream's outputs and network state after serving equal the same stream stepped alone."""
    for r in reqs:
        admits = {k for s, k in srv.admit_log if s == r.sid}
        alone = _replay_alone(dev, r, srv.pools[r.sid], admits, network)
        if r.sid in srv.denoisers:
            d = srv.denoisers[r.sid]
            for k in ("out", "out_log", "ols1", "ols2", "loss"):
                assert torch.equal(d.outputs()[k], alone.outputs()[k]), (r.sid, k)
            state = (d.params, d.adam_m, d.adam_v)
        else:
            state = srv.saved[r.sid][:3]
        for got, want in zip(state, (alone.params, alone.adam_m, alone.adam_v)):
            assert torch.equal(got.cpu(), want.cpu()), r.sid
        alone.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "batched_graphs", "graphs"])
def test_serving_loop_runs_all_frames(dev, mode):
    from xdn.serving import make_server
    rng = np.random.default_rng(1)
    reqs = [StreamReq(i, h, w, 6, poisson_arrivals(120.0, 6, rng=rng), seed=i)
            for i, (h, w) in enumerate([(64, 64), (48, 96), (96, 64), (64, 64)])]
    srv = make_server(mode, dev, hbm_budget=1 << 40, pool_frames=2)
    res = srv.run(reqs)
    assert res["frames"] == 24 and np.isfinite(res["p99_ms"]) and res["p99_ms"] > 0
    check_against_alone(dev, srv, reqs)
    srv.close()
    # a budget for two slots forces waits and evictions but every frame is still served, and an
    # evicted and readmitted stream continues with its network and Adam state
    from xdn import pipeline
    b2 = 2 * pipeline.slot_bytes(96, 64) + 1
    srv2 = make_server(mode, dev, hbm_budget=b2, pool_frames=2, idle_after=0.0)
    res2 = srv2.run(reqs)
    assert res2["frames"] == 24 and res2["evictions"] > 0
    check_against_alone(dev, srv2, reqs)
    srv2.close()


@pytest.mark.gpu
def test_batched_graphs_are_replayed(dev):
    """Streams whose frames always arrive together form the same batch every time: from the second
    such batch on, the captured graph of the batched step is replayed."""
    from xdn.serving import make_server
    reqs = [StreamReq(i, h, w, 6, [0.002 * k for k in range(6)], seed=i)
            for i, (h, w) in enumerate([(64, 64), (48, 96)])]
    srv = make_server("batched_graphs", dev, hbm_budget=1 << 40, pool_frames=2, min_batch=2)
    res = srv.run(reqs)
    assert res["frames"] == 12 and res["graphs"]["replayed"] > 0 and res["mean_batch"] > 1
    check_against_alone(dev, srv, reqs)
    srv.close()


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
    check_against_alone(dev, srv, reqs, network="unet")
    srv.close()
