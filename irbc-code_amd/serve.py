"""Request loop: continuous batching of sweep requests on one GPU.

A request asks for one replica (box, density, mesh size, knobs, rates,
shear: a dict of ensemble.replica) to be run for a number of steps, and
returns its rolling time series (membrane centroid, centroid velocity,
bonded knobs every `series_every` steps) and its final membrane state.

The loop
  * keeps at most `max_active` micro-batches on the GPU (each has its own
    pair of HIP streams; more would share the device's few hardware queues
    and serialise);
  * when one of those places is free, admits every arrived request, in
    arrival order, while their device bytes (irbc_estimate_bytes) fit the
    HBM budget left, into one context (a micro-batch: one launch sequence
    for all of them), so requests that arrive while the GPU is busy are
    batched together.  With a batching `window` > 0 it forms a micro-batch
    only once the oldest waiting request has waited that long (or a full
    group is waiting): a few hundred milliseconds of waiting turn a trickle
    of one-replica micro-batches into a few large ones, which is where the
    GPU's throughput is;
  * keeps every admitted replica's state on the device until it is done;
  * advances every active context by `chunk` steps at a time, each on its
    own pair of HIP streams, so micro-batches admitted at different times
    run on the GPU together;
  * collects a micro-batch when its last chunk has finished, frees its
    memory and admits waiting requests into the space.

A replica's result does not depend on what it was batched with or when:
every per-replica computation is independent of the rest of its batch and
of how its steps are split into calls.
"""
import concurrent.futures as cf
import time
from collections import deque

import numpy as np

import ensemble as en
import irbc_amd as ia


class Request:
    def __init__(self, rid, replica, steps, series_every=100, arrival=0.0):
        self.id = rid
        self.replica = replica
        self.steps = int(steps)
        self.series_every = int(series_every)
        self.arrival = float(arrival)  # seconds after the loop starts
        self.spec = en.spec(replica)
        self.bytes = ia.estimate_bytes([self.spec])
        # filled by the loop
        self.admitted = None
        self.finished = None
        self.series = None
        self.final = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Group:
    """One context holding the requests admitted together."""

    def __init__(self, reqs, now, graphs=False, options=None):
        self.reqs = reqs
        self.batch = ia.Batch([r.spec for r in reqs])
        # list updates (and the near-mask rebuilds after them) beside the
        # collisions, whatever each replica's cadence: every chunk of a
        # collision period then has the same shape
        self.batch.set_option(ia.OPT_ALIGN_LISTS, 1)
        if options is not None and len(reqs) >= 2:
            # a launch-option index (bit i = irbc_amd.OPTIONS[i]) for
            # micro-batches of several replicas
            for i, (_, opt) in enumerate(ia.OPTIONS):
                self.batch.set_option(opt, (options >> i) & 1)
        else:
            self.batch.set_option(ia.OPT_GRAPHS, 1 if graphs else 0)
        for i, r in enumerate(reqs):
            en.load(self.batch, i, r.replica, device_solvent=True)
            if r.series_every > 0:
                self.batch.set_series(i, r.series_every, max(1, r.steps // r.series_every))
        en.start(self.batch)
        if len(reqs) >= 8 and options is None:
            # a large micro-batch streams four particles per lane, as the
            # autotuner picks for the 64-replica batch
            self.batch.set_option(ia.OPT_STREAM4, 1)
        self.steps = reqs[0].steps
        self.done_steps = 0
        self.bytes = self.batch.device_bytes()
        for r in reqs:
            r.admitted = now


class Server:
    def __init__(self, budget_bytes, chunk=250, max_group=64, max_active=2, window=0.0, builders=1, graphs=False,
                 options=None):
        self.budget = int(budget_bytes)
        self.chunk = int(chunk)
        self.max_group = int(max_group)
        self.max_active = int(max_active)
        self.window = float(window)
        self.builders = int(builders)
        self.graphs = bool(graphs)
        self.options = options
        self.used = 0
        self.active = []
        self.building = []  # (future of a _Group, its reserved bytes)
        self.stats = {"groups": 0, "max_active_groups": 0, "max_resident_replicas": 0}

    def _admit(self, queue, now):
        """Arrived requests, in arrival order, while they fit the budget.
        Requests of one group share a step count (the group advances
        together); a request that does not fit waits for memory."""
        while queue and queue[0].arrival <= now and len(self.active) + len(self.building) < self.max_active:
            head = queue[0]
            if now - head.arrival < self.window:
                arrived = 0
                for r in queue:
                    if r.arrival > now or arrived >= self.max_group:
                        break
                    arrived += 1
                if arrived < self.max_group:
                    return  # wait for more arrivals to batch with
            if head.bytes > self.budget:
                raise ValueError("request %s needs %d bytes, more than the whole budget %d"
                                 % (head.id, head.bytes, self.budget))
            group, need = [], 0
            while (queue and queue[0].arrival <= now and len(group) < self.max_group
                   and queue[0].steps == head.steps):
                specs = [r.spec for r in group] + [queue[0].spec]
                b = ia.estimate_bytes(specs)
                if self.used + b > self.budget:
                    break
                group.append(queue.popleft())
                need = b
            if not group:
                return
            # the context is built and its replicas set up (host-side state,
            # uploads, classification, lists) on a builder thread, while this
            # loop keeps the other micro-batches' chunks going
            self.used += need
            self.building.append((self.pool.submit(_Group, group, now, self.graphs, self.options), need))
            self.stats["groups"] += 1

    def _collect(self, g, now):
        b = g.batch
        b.sync()
        for i, r in enumerate(g.reqs):
            r.series = b.series(i) if r.series_every > 0 else None
            v = b.vertices(i)
            r.final = {k: v[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ix", "iy", "bondi")}
            r.finished = time.perf_counter() - self.t0
        # freeing device memory can wait for the whole device: the builder
        # thread does it, before any later build allocates
        self.pool.submit(b.close)
        self.used -= g.reserved

    def run(self, requests, idle_sleep=2e-4):
        """Serve `requests` (their arrival times are offsets from now);
        returns them with results and latencies filled in."""
        queue = deque(sorted(requests, key=lambda r: r.arrival))
        self.pool = cf.ThreadPoolExecutor(max_workers=self.builders)
        self.t0 = time.perf_counter()
        try:
            self._loop(queue, idle_sleep)
        finally:
            self.pool.shutdown(wait=True)
        self.makespan = time.perf_counter() - self.t0
        return requests

    def _loop(self, queue, idle_sleep):
        while queue or self.active or self.building:
            now = time.perf_counter() - self.t0
            self._admit(queue, now)
            progressed = False
            for item in list(self.building):
                fut, need = item
                if fut.done():
                    g = fut.result()  # a failed build raises here
                    g.reserved = max(need, g.bytes)
                    self.used += g.reserved - need
                    self.active.append(g)
                    self.building.remove(item)
                    progressed = True
            for g in list(self.active):
                if not g.batch.done():
                    continue
                if g.done_steps < g.steps:
                    n = min(self.chunk, g.steps - g.done_steps)
                    g.batch.md_steps(n)
                    g.done_steps += n
                    progressed = True
                else:
                    self._collect(g, now)
                    self.active.remove(g)
                    progressed = True
            self.stats["max_active_groups"] = max(self.stats["max_active_groups"], len(self.active))
            self.stats["max_resident_replicas"] = max(self.stats["max_resident_replicas"],
                                                     sum(len(g.reqs) for g in self.active))
            if not progressed:
                time.sleep(idle_sleep)


def poisson_requests(n, rate, steps, seed=0, series_every=100, pool=None):
    """n requests with exponential inter-arrival times at `rate` per second,
    replicas drawn like the uneven ensemble."""
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, size=n))
    pool = en.uneven(n, seed0=3000 + seed) if pool is None else pool
    return [Request(i, pool[i % len(pool)], steps, series_every, float(t[i])) for i in range(n)]


def summary(reqs, makespan):
    lat = np.array([r.latency for r in reqs])
    steps = sum(r.steps for r in reqs)
    return {"requests": len(reqs), "makespan_s": makespan, "replica_steps_per_s": steps / makespan,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max())}
