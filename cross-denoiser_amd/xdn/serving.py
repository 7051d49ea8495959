"""Multi-stream serving loop for the per-frame online denoiser.

Render streams arrive (and may join mid-run) with frames at their own rates; each admitted stream
owns a slot whose history, network, Adam state and scratch stay resident in HBM. The scheduler does
continuous batching: whenever the GPU can take work, every stream that has an arrived frame and no
frame in flight joins the next batch, which is ONE batched step (one launch per stage over all its
streams, ``xdn_denoise_step``). Batches alternate between two HIP streams so a new batch overlaps the
tail of the previous one on the device. A batch whose composition repeats (under load: every stream,
each at its frame buffers and history parity) is captured once as a HIP graph of the batched step and
replayed. The comparison modes are the plain loop (one frame at a time, eager launches), per-stream
captured step graphs (``pipeline.StepGraph``) and eager batches. Streams are admitted against an HBM
budget; idle streams are evicted (their slot memory released) to admit waiting ones. Per-frame latency = completion (HIP
event) - arrival; frames/s and p50/p99 are reported.

The admission/eviction policy (``Admission``) is plain host logic, testable without a GPU.
"""
from __future__ import annotations

import math
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np


# ------------------------------------------------------------------------------------------------
# Workload description
# ------------------------------------------------------------------------------------------------
@dataclass
class StreamReq:
    sid: int
    height: int
    width: int
    n_frames: int
    arrivals: list            # arrival time of each frame, seconds from the start
    seed: int = 0
    pan_max: float = 3.0
    disocclusion: float = 0.1


def poisson_arrivals(rate_hz: float, n: int, start: float = 0.0, rng=None) -> list:
    rng = rng or np.random.default_rng(0)
    gaps = rng.exponential(1.0 / rate_hz, size=n)
    return list(start + np.cumsum(gaps))


def onoff_arrivals(peak_hz: float, n: int, period_s: float = 0.2, start: float = 0.0, rng=None) -> list:
    """Bursty arrivals: Poisson at peak_hz during 'on' periods, nothing during 'off' periods."""
    rng = rng or np.random.default_rng(0)
    out, t = [], start
    while len(out) < n:
        t += rng.exponential(1.0 / peak_hz)
        phase = ((t - start) // period_s) % 2
        if phase == 1:  # off: jump to the next on period
            t = start + (math.floor((t - start) / period_s) + 1) * period_s
            continue
        out.append(t)
    return out


def config3(frames_per_stream=64, fps=60.0, seed=0) -> list:
    """16 streams: 8x1280x720, 6x1920x1080, 2x2560x1440, Poisson at `fps`."""
    rng = np.random.default_rng(seed)
    sizes = [(720, 1280)] * 8 + [(1080, 1920)] * 6 + [(1440, 2560)] * 2
    return [StreamReq(i, h, w, frames_per_stream, poisson_arrivals(fps, frames_per_stream, rng=rng), seed=seed + i)
            for i, (h, w) in enumerate(sizes)]


def config4(frames_per_stream=64, seed=0, join_at=0.25) -> list:
    """4x3840x2160 + 12x1920x1080, bursty (on/off 200 ms at 120 fps), pans U[-8,8], 20% disocclusion;
    half of the streams join mid-run."""
    rng = np.random.default_rng(seed)
    sizes = [(2160, 3840)] * 4 + [(1080, 1920)] * 12
    out = []
    for i, (h, w) in enumerate(sizes):
        start = join_at if i % 2 else 0.0
        out.append(StreamReq(i, h, w, frames_per_stream, onoff_arrivals(120.0, frames_per_stream, start=start, rng=rng),
                             seed=seed + i, pan_max=8.0, disocclusion=0.2))
    return out


# ------------------------------------------------------------------------------------------------
# Admission by memory budget with idle eviction (host logic)
# ------------------------------------------------------------------------------------------------
@dataclass
class Admission:
    budget: int
    used: int = 0
    resident: dict = field(default_factory=dict)   # sid -> bytes
    last_active: dict = field(default_factory=dict)

    def try_admit(self, sid: int, nbytes: int, now: float, busy: set, idle_after: float = 0.05):
        """Admit `sid` if it fits, evicting idle resident streams (oldest activity first) if needed.

        Returns (admitted, evicted_sids)."""
        if sid in self.resident:
            return True, []
        if nbytes > self.budget:
            return False, []
        evicted = []
        if self.used + nbytes > self.budget:
            idle = sorted((t, s) for s, t in self.last_active.items()
                          if s not in busy and now - t >= idle_after)
            freed = 0
            plan = []
            for _, s in idle:
                if self.used - freed + nbytes <= self.budget:
                    break
                plan.append(s)
                freed += self.resident[s]
            if self.used - freed + nbytes > self.budget:
                return False, []
            for s in plan:
                self.release(s)
                evicted.append(s)
        self.resident[sid] = nbytes
        self.used += nbytes
        self.last_active[sid] = now
        return True, evicted

    def touch(self, sid: int, now: float):
        self.last_active[sid] = now

    def release(self, sid: int):
        if sid in self.resident:
            self.used -= self.resident.pop(sid)
            self.last_active.pop(sid, None)


# ------------------------------------------------------------------------------------------------
# GPU serving loop
# ------------------------------------------------------------------------------------------------
class Server:
    def __init__(self, device="cuda", hbm_budget=64 << 30, max_batch=16, lr=1e-3, pool_frames=4, batching=True,
                 idle_after=0.05, graphs=False, max_inflight=2, network="cnn5", graph_cache=64, graph_after=2,
                 min_batch=1, max_batch_px=None):
        """batching: same-stage kernels of all ready streams in one launch (xdn_denoise_step over
        many slots); otherwise one frame at a time. graphs: with batching off, each frame is the
        stream's captured step graph (one per frame buffers and history parity); with batching on,
        a batch whose composition (streams, their frame buffers and history parities) has been
        launched `graph_after` - 1 times before is captured as one graph of the batched step and
        replayed from then on (at most `graph_cache` graphs, least recently used dropped; other
        batches run eagerly). max_inflight: batches on the device at once (each on its own HIP
        stream). network: every stream's bandwidth network ("cnn5" or "unet"). min_batch: with
        batching, fewer ready streams than this are launched one at a time (a batch of one each,
        up to max_inflight on the device), so a light load keeps the per-frame latency of the
        one-frame loop and only a backlog is batched. max_batch_px: a batch takes streams oldest
        pending frame first while their frames' pixels add up to at most this (the first stream
        always fits; None: no limit), so one batch's duration stays near one frame of that size
        and a stream with a backlog, which advances one frame per batch, is not held back by
        long batches."""
        self.adm = Admission(hbm_budget)
        self.max_batch = max_batch if batching else 1
        self.min_batch = max(1, int(min_batch))
        self.max_batch_px = max_batch_px
        self.batching = batching
        self.graphs = graphs
        self.lr = lr
        self.network = network
        self.pool_frames = pool_frames
        self.idle_after = idle_after
        self.max_inflight = max(1, int(max_inflight))
        self.graph_cache = max(1, int(graph_cache))
        self.graph_after = max(1, int(graph_after))
        self.batch_graphs = OrderedDict()   # composition -> pipeline.StepGraph (LRU order)
        self.batch_seen = {}                # composition -> launches so far
        self.graph_stats = {"captured": 0, "replayed": 0, "eager": 0}
        self.denoisers = {}
        self.epoch = {}    # sid -> admission serial of its current slot
        self.pools = {}
        self.saved = {}    # evicted streams: (params, adam m, adam v, step) on the host
        self.admit_log = []  # (sid, first frame index run by the new slot)
        self._serial = 0
        self._init_device(device)

    # -- device hooks (overridden by host-only tests) ---------------------------------------------
    def _init_device(self, device):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.hip_streams = [torch.cuda.Stream(self.device) for _ in range(self.max_inflight)]
        self._cap_stream = None

    def _slot_bytes(self, r: StreamReq) -> int:
        from . import pipeline
        return pipeline.slot_bytes(r.height, r.width, self.network)

    def _sync(self):
        self.torch.cuda.synchronize()

    def _frame_pool(self, req: StreamReq):
        """A few distinct frames of the stream, resident in HBM (the renderer's output buffers)."""
        from . import synth
        spec = synth.StreamSpec(req.height, req.width, seed=req.seed, stream=req.sid, pan=None, pan_max=req.pan_max,
                                disocclusion=req.disocclusion)
        return [synth.make_frame(spec, t, device=self.device) for t in range(self.pool_frames)]

    def run(self, reqs: list, log=None) -> dict:
        rejected = [r for r in reqs if self._slot_bytes(r) > self.adm.budget]
        reqs = [r for r in reqs if r not in rejected]
        for r in reqs:
            self.pools[r.sid] = self._frame_pool(r)
        self._sync()
        events = []   # (arrival, sid, frame index)
        for r in reqs:
            for k, t in enumerate(r.arrivals):
                events.append((t, r.sid, k))
        events.sort()
        by_sid = {r.sid: r for r in reqs}
        pending = {r.sid: [] for r in reqs}   # arrived, not started: list of (arrival, k)
        next_ev = 0
        inflight = []     # (completion handle, [(sid, k, arrival)])
        busy = set()
        latencies, done_frames, evictions, admit_wait = [], 0, 0, 0
        total = len(events)
        t0 = time.perf_counter()
        bi = 0
        while done_frames < total:
            now = time.perf_counter() - t0
            while next_ev < total and events[next_ev][0] <= now:
                a, sid, k = events[next_ev]
                pending[sid].append((a, k))
                next_ev += 1
            # completions
            still = []
            for ev, items in inflight:
                if ev.query():
                    tdone = time.perf_counter() - t0
                    for sid, k, a in items:
                        latencies.append(tdone - a)
                        busy.discard(sid)
                        self.adm.touch(sid, tdone)
                    done_frames += len(items)
                else:
                    still.append((ev, items))
            inflight = still
            # admission of streams with pending frames; a stream already picked for this batch
            # (or on the device) is never evicted to make room for another
            ready, ready_px = [], 0
            for sid in sorted(pending, key=lambda s: pending[s][0][0] if pending[s] else 1e30):
                if not pending[sid] or sid in busy:
                    continue
                px = by_sid[sid].height * by_sid[sid].width
                if ready and self.max_batch_px is not None and ready_px + px > self.max_batch_px:
                    continue
                if sid not in self.denoisers:
                    r = by_sid[sid]
                    ok, ev_s = self.adm.try_admit(sid, self._slot_bytes(r), now, busy | set(ready), self.idle_after)
                    for s in ev_s:
                        self._evict(s)
                        evictions += 1
                    if not ok:
                        admit_wait += 1
                        continue
                    self.denoisers[sid] = self._admit(r)
                    self._serial += 1
                    self.epoch[sid] = self._serial
                    self.admit_log.append((sid, pending[sid][0][1]))
                ready.append(sid)
                ready_px += px
                if len(ready) >= self.max_batch:
                    break
            if self.batching and len(ready) < self.min_batch:
                ready = ready[:1]  # light load: one stream per launch, the others take the next slot
            if ready and len(inflight) < self.max_inflight:
                slot = bi % self.max_inflight
                bi += 1
                items, sids, frs = [], [], []
                for sid in ready:
                    a, k = pending[sid].pop(0)
                    sids.append(sid)
                    frs.append(self.pools[sid][k % self.pool_frames])
                    items.append((sid, k, a))
                    busy.add(sid)
                # no cross-batch wait: a stream joins a batch only after the host has seen its
                # previous frame complete, and streams share no state
                inflight.append((self._launch(slot, sids, frs), items))
            elif not inflight and next_ev < total:
                # idle until the next arrival
                wait = events[next_ev][0] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        self._sync()
        wall = time.perf_counter() - t0
        lat = np.array(latencies) * 1e3
        return {"frames": done_frames, "wall_s": wall, "fps": done_frames / wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "max_ms": float(lat.max()), "evictions": evictions, "admission_waits": admit_wait,
                "batches": bi, "mean_batch": done_frames / max(bi, 1), "rejected_streams": [r.sid for r in rejected],
                "graphs": dict(self.graph_stats)}

    def _launch(self, slot: int, sids: list, frames: list):
        """Enqueue one step of streams `sids` on frames `frames` on HIP stream `slot`; returns an
        event recorded after it."""
        from . import pipeline
        torch = self.torch
        st = self.hip_streams[slot]
        dens = [self.denoisers[s] for s in sids]
        with torch.cuda.stream(st):
            # slot set-up and restores of newly admitted streams ran on the default stream
            st.wait_stream(torch.cuda.default_stream(self.device))
            if self.graphs and not self.batching:
                for d, f in zip(dens, frames):
                    d._graph_for(f).launch(st.cuda_stream)
                    d.advance(f)
            elif self.graphs:
                key = tuple((s, self.epoch[s], d.graph_key(f)) for s, d, f in zip(sids, dens, frames))
                g = self.batch_graphs.get(key)
                if g is None:
                    seen = self.batch_seen.get(key, 0) + 1
                    self.batch_seen[key] = seen
                    if seen >= self.graph_after:
                        if self._cap_stream is None:
                            self._cap_stream = torch.cuda.Stream(self.device)
                        g = pipeline.StepGraph([d.make_slot(f) for d, f in zip(dens, frames)], self.lr,
                                               self._cap_stream.cuda_stream)
                        self.batch_graphs[key] = g
                        self.batch_seen.pop(key, None)
                        self.graph_stats["captured"] += 1
                        while len(self.batch_graphs) > self.graph_cache:
                            self.batch_graphs.popitem(last=False)[1].destroy()
                else:
                    self.batch_graphs.move_to_end(key)
                    self.graph_stats["replayed"] += 1
                if g is not None:
                    g.launch(st.cuda_stream)
                    for d, f in zip(dens, frames):
                        d.advance(f)
                else:
                    self.graph_stats["eager"] += 1
                    pipeline.denoise_step_multi(dens, frames, stream=st.cuda_stream)
            else:
                pipeline.denoise_step_multi(dens, frames, stream=st.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(st)
        return ev

    def _drop_graphs_of(self, sid: int):
        """Batched graphs hold the slot buffers of their streams: drop those of an evicted stream."""
        for key in [k for k in self.batch_graphs if any(e[0] == sid for e in k)]:
            self.batch_graphs.pop(key).destroy()
        for key in [k for k in self.batch_seen if any(e[0] == sid for e in k)]:
            self.batch_seen.pop(key)

    def _admit(self, r: StreamReq):
        """A slot for stream r; an evicted stream gets its network and Adam state back (its frame
        history restarts from zero, as at frame 0)."""
        from . import pipeline
        d = pipeline.Denoiser(r.height, r.width, device=self.device, seed=1234 + r.sid, lr=self.lr,
                              use_graphs=self.graphs and not self.batching, network=self.network)
        saved = self.saved.pop(r.sid, None)
        if saved is not None:
            for t, src in zip((d.params, d.adam_m, d.adam_v), saved[:3]):
                t.copy_(src)
            d.set_step(saved[3])
            d.repack()
        return d

    def _evict(self, sid: int):
        """Release the slot of an idle stream (no frame of it on the device): its network and Adam
        state go to the host."""
        self._drop_graphs_of(sid)
        d = self.denoisers.pop(sid, None)
        if d is not None:
            self.saved[sid] = (d.params.cpu(), d.adam_m.cpu(), d.adam_v.cpu(), d.t)
            d.close()

    def close(self):
        for g in self.batch_graphs.values():
            g.destroy()
        self.batch_graphs.clear()
        for d in self.denoisers.values():
            d.close()


# ------------------------------------------------------------------------------------------------
# Command line: python -m xdn.serving --config 3 --mode all
# ------------------------------------------------------------------------------------------------
def warm_up(device, sizes, lr=1e-3, network="cnn5"):
    """First launches of every kernel at every resolution (module load, code object upload)."""
    import torch

    from . import pipeline, synth
    for (h, w) in sorted(set(sizes)):
        spec = synth.StreamSpec(h, w, seed=99)
        fr = [synth.make_frame(spec, t, device=device) for t in range(2)]
        d = pipeline.Denoiser(h, w, device=device, lr=lr, network=network)
        d.step(fr[0])
        d.step(fr[1])
        pipeline.denoise_step_multi([d], [fr[0]])
        torch.cuda.synchronize()
        d.close()
        del d, fr
    torch.cuda.empty_cache()


MODES = ("plain", "graphs", "batched", "batched_graphs")


def make_server(mode: str, device, **kw) -> Server:
    """plain: the one-frame-at-a-time loop (one frame on the device at a time, eager launches);
    graphs: one frame at a time, each the stream's captured step graph, two frames in flight;
    batched: continuous batching, eager batched steps; batched_graphs: continuous batching with
    repeating batches replayed as captured graphs."""
    if mode not in MODES:
        raise ValueError(f"unknown serving mode {mode!r} (one of {MODES})")
    if mode.startswith("batched"):
        # batches of at most one 1920x1080 frame's pixels (e.g. two 1280x720 frames): with larger
        # batches a stream with a backlog advances one frame per long batch and the p99 grows
        kw.setdefault("max_batch_px", 1920 * 1080)
    return Server(device, batching=mode.startswith("batched"), graphs=mode.endswith("graphs"),
                  max_inflight=1 if mode == "plain" else 2, **kw)


def main(argv=None):
    import argparse
    import json
    import sys

    import torch
    ap = argparse.ArgumentParser(description="multi-stream serving loop of the online denoiser")
    ap.add_argument("--config", type=int, choices=(3, 4), default=3)
    ap.add_argument("--frames", type=int, default=64, help="frames per stream")
    ap.add_argument("--mode", choices=MODES + ("all",), default="all")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--fps", type=float, default=60.0, help="mean arrival rate per stream (configuration 3)")
    ap.add_argument("--network", choices=("cnn5", "unet"), default="cnn5")
    ap.add_argument("--min-batch", type=int, default=None, help="batched modes: ready streams below which "
                    "each is launched alone (default: the mode's default)")
    ap.add_argument("--max-batch-px", type=float, default=None, help="batched modes: pixel limit of one batch, "
                    "in 1920x1080 frames (default: the mode's default; 0: no limit)")
    a = ap.parse_args(argv)
    kw = {}
    if a.min_batch is not None:
        kw["min_batch"] = a.min_batch
    if a.max_batch_px is not None:
        kw["max_batch_px"] = int(a.max_batch_px * 1920 * 1080) if a.max_batch_px > 0 else None
    dev = torch.device("cuda:0")
    reqs = config3(a.frames, fps=a.fps, seed=a.seed) if a.config == 3 else config4(a.frames, seed=a.seed)
    warm_up(dev, [(r.height, r.width) for r in reqs], network=a.network)
    modes = MODES if a.mode == "all" else (a.mode,)
    for m in modes:
        print(f"[serving] config {a.config}, mode {m}, {len(reqs)} streams x {a.frames} frames", file=sys.stderr,
              flush=True)
        srv = make_server(m, dev, hbm_budget=int(a.budget_gb * (1 << 30)), network=a.network, **kw)
        res = srv.run(reqs)
        res.update({"config": a.config, "mode": m, "network": a.network, "streams": len(reqs), "frames_per_stream": a.frames,
                    "min_batch": srv.min_batch, "max_batch_px": srv.max_batch_px,
                    "budget_gb": a.budget_gb, "arrival_fps_per_stream": a.fps if a.config == 3 else 120.0,
                    "stream_sizes": sorted({f"{r.width}x{r.height}" for r in reqs})})
        print(json.dumps(res), flush=True)
        srv.close()
        del srv
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
