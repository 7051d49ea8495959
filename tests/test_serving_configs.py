"""The serving loop on the two serving workloads at their real stream sizes (GPU).

  - 16 streams: 8 x 1280x720 + 6 x 1920x1080 + 2 x 2560x1440, Poisson arrivals at 60 fps, three frames
    per stream through the continuous-batching server (eager batches and captured batch graphs):
    every frame is served and every stream's outputs, weights and Adam moments are bit-identical to
    the same stream stepped alone.
  - 16 streams: 4 x 3840x2160 + 12 x 1920x1080, bursty arrivals, half of them joining at 0.25 s, under
    a 12 GB budget (two 3840x2160 slots and a few 1920x1080 ones) that forces evictions and
    readmissions. Each stream's last frame comes one on/off cycle (0.4 s) after its burst, so every
    stream goes idle with a frame still to come: the joiners evict the idle early streams, whose last
    frames then readmit them. Restored network weights, Adam moments and step counters equal the
    saved ones, and every stream again matches its alone run (a new slot at each readmission, with
    the history restarted).
"""
import numpy as np
import pytest
import torch

from test_serving import check_against_alone
from xdn.serving import Server, config3, config4, make_server

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["batched", "batched_graphs"])
def test_config3_mix_matches_streams_alone(dev, mode):
    reqs = config3(frames_per_stream=3, fps=60.0, seed=5)
    srv = make_server(mode, dev, hbm_budget=64 << 30, pool_frames=2)
    res = srv.run(reqs)
    print(f"  config 3 ({mode}): {res['frames']} frames, {res['fps']:.0f} frames/s, p99 {res['p99_ms']:.1f} ms, "
          f"mean batch {res['mean_batch']:.1f}, graphs {res['graphs']}", flush=True)
    assert res["frames"] == 48 and res["evictions"] == 0 and not res["rejected_streams"]
    assert sorted({(d.H, d.W) for d in srv.denoisers.values()}) == [(720, 1280), (1080, 1920), (1440, 2560)]
    check_against_alone(dev, srv, reqs)
    srv.close()


class RestoreChecked(Server):
    """Checks each readmission: the new slot holds exactly the saved network, Adam moments and step."""

    restores = 0

    def _admit(self, r):
        saved = self.saved.get(r.sid)
        d = super()._admit(r)
        if saved is not None:
            torch.cuda.synchronize()
            for got, want in zip((d.params, d.adam_m, d.adam_v), saved[:3]):
                assert torch.equal(got.cpu(), want), r.sid
            assert d.t == saved[3] and int(d.adam_t.item()) == saved[3]
            self.restores += 1
        return d


def test_config4_mix_under_eviction(dev):
    reqs = config4(frames_per_stream=3, seed=6)
    for r in reqs:  # the last frame one on/off cycle later: it arrives in the stream's next burst
        r.arrivals = r.arrivals[:2] + [r.arrivals[2] + 0.4]
    srv = RestoreChecked(dev, hbm_budget=12 << 30, pool_frames=2, idle_after=0.0)
    res = srv.run(reqs)
    print(f"  config 4: {res['frames']} frames, {res['evictions']} evictions, {srv.restores} restores, "
          f"p99 {res['p99_ms']:.1f} ms", flush=True)
    assert res["frames"] == 48 and not res["rejected_streams"]
    assert res["evictions"] > 0 and srv.restores > 0
    check_against_alone(dev, srv, reqs)
    srv.close()
