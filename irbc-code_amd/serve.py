"""Request loop: continuous batching of sweep requests on one GPU.

Two schedules share the request type, admission by HBM budget and the
batching window:

re-packing (the default): every running request lives in ONE context, so
each step is one launch sequence over all of them, whatever their sizes.
An arrived request is set up at once by one of a few stager threads in a
one-replica context of its own (mesh, knobs, solvent drawn on the device,
inside flags, lists) while the GPU runs.  When staged requests have waited
out the batching window, or finished requests hold enough of the running
context, a builder thread creates the next context, sized for the
still-running requests plus the new ones.  At the running context's next
chunk boundary, with the GPU idle, every slot is filled device to device
(irbc_copy_replica: state, mesh and series never leave the device) and the
new context takes over.  Requests of different step counts finish at chunk
boundaries (a chunk is cut short at the nearest finish); a finished
request's results are read at once and its slot idles until the next
re-pack.

micro-batches (repack=False): the requests admitted together form one
context of their own, and up to `max_active` such contexts run on the GPU
at once, each on its own HIP streams.

A request asks for one replica (box, density, mesh size, knobs, rates,
shear: a dict of ensemble.replica) to be run for a number of steps, and
returns its rolling time series (membrane centroid, centroid velocity,
bonded knobs every `series_every` steps) and its final membrane state.

The micro-batch loop
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
        self.staged = None  # re-packing: future of the request's staging context
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


def _configure(b, n, graphs, options, bin_span=1):
    """Launch options of a context of n replicas: list updates beside the
    collisions (every chunk of a collision period then has one shape), and
    either a launch-option index (bit i = irbc_amd.OPTIONS[i]) for contexts
    of several replicas, with the bin-sort tile span bin_span, or the
    library's defaults with graphs on or off."""
    b.set_option(ia.OPT_ALIGN_LISTS, 1)
    if options is not None and n >= 2:
        for i, (_, opt) in enumerate(ia.OPTIONS):
            b.set_option(opt, (options >> i) & 1)
        b.set_option(ia.OPT_BIN_SPAN, bin_span)
    else:
        b.set_option(ia.OPT_GRAPHS, 1 if graphs else 0)
        if n >= 8:
            # a large context streams four particles per lane, as the
            # autotuner picks for the 64-replica batch
            b.set_option(ia.OPT_STREAM4, 1)


def _stage(reqs):
    """The requests' initial state (maincode.cpp:56-391 for each) in a
    context of their own, to be copied into the running one."""
    b = ia.Batch([r.spec for r in reqs])
    try:
        b.set_option(ia.OPT_ALIGN_LISTS, 1)
        for i, r in enumerate(reqs):
            en.load(b, i, r.replica, device_solvent=True)
            if r.series_every > 0:
                b.set_series(i, r.series_every, max(1, r.steps // r.series_every))
        en.start(b)
    except BaseException:
        b.close()
        raise
    return b


class _Slots:
    """The running context of the re-packing loop: its requests by slot and
    the steps each has left (0: finished, its slot idles).  Created and
    configured on the builder thread; every slot's state (mesh, solvent,
    membrane, bonds, lists, series) is copied in device to device at the
    switch: the still-running requests' from the context before, the new
    ones' from their staging contexts."""

    def __init__(self, cont, new, graphs, options, bin_span=1):
        self.reqs = list(cont) + list(new)
        self.ncont = len(cont)
        self.left = [0] * len(self.reqs)
        self.read = [False] * len(self.reqs)
        t = time.perf_counter()
        b = ia.Batch([r.spec for r in self.reqs])
        try:
            _configure(b, len(self.reqs), graphs, options, bin_span)
        except BaseException:
            b.close()
            raise
        self.times = {"create": time.perf_counter() - t}
        self.batch = b
        self.bytes = b.device_bytes()
        self.np = [r.spec.n_solvent for r in self.reqs]

    def live(self):
        return [i for i, n in enumerate(self.left) if n > 0]


class Server:
    def __init__(self, budget_bytes, chunk=250, max_group=64, max_active=2, window=0.0, builders=1, graphs=False,
                 options=None, repack=True, repack_idle=0.25, stagers=4, bin_span=1):
        """budget_bytes: device bytes the loop may hold (re-packing holds the
        running context, the next one and the staged requests at a switch);
        chunk: steps per irbc_md_steps call; max_group: replicas per
        context; window: seconds the oldest arrival waits for others;
        options: launch-option index for contexts of several replicas;
        repack_idle: re-pack once finished requests hold this share of the
        running context's particles; stagers: threads setting up arrived
        requests, each in a context of its own, while the GPU runs;
        bin_span: the bin-sort tile span of contexts of several replicas
        that take `options`."""
        self.bin_span = int(bin_span)
        self.repack = bool(repack)
        self.repack_idle = float(repack_idle)
        self.stagers = int(stagers)
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
        self.log = []  # re-packing: one record per context built (when planned and switched to, timings)
        self.idle = 0.0  # re-packing: seconds with no chunk on the GPU while a context was being built
        # re-packing: per chunk [launched at, seconds in md_steps, seen done at, steps, slots, running]
        self.chunks = []

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
        self.stage_pool = cf.ThreadPoolExecutor(max_workers=max(1, self.stagers)) if self.repack else None
        self.t0 = time.perf_counter()
        try:
            if self.repack:
                self._loop_repack(queue, idle_sleep)
            else:
                self._loop(queue, idle_sleep)
        finally:
            if self.stage_pool is not None:
                self.stage_pool.shutdown(wait=True)
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


    # ---- re-packing: one running context ----
    def _stage_arrivals(self, queue, staged, now):
        """Arrived requests go to the stager threads, in arrival order, while
        the budget holds their staging contexts and their slots in the next
        running context (two request sizes each)."""
        n = 0
        while queue and queue[0].arrival <= now:
            r = queue[0]
            if 2 * r.bytes > self.budget:
                raise ValueError("request %s needs %d bytes and as many to stage, more than the whole budget %d"
                                 % (r.id, r.bytes, self.budget))
            if self.used + 2 * r.bytes > self.budget:
                break
            queue.popleft()
            self.used += 2 * r.bytes
            r.staged = self.stage_pool.submit(_stage, [r])
            staged.append(r)
            n += 1
        return n

    def _plan(self, staged, cur, now):
        """Start building the next context when staged requests have waited
        out the window since they arrived (or fill the context), or finished
        requests hold repack_idle of the running one's particles; None when
        there is nothing to do or no room in the budget."""
        live = cur.live() if cur is not None else []
        ready = []
        for r in staged:
            if not r.staged.done():
                break  # admitted in arrival order
            ready.append(r)
        room = self.max_group - len(live)
        ready = ready[:max(room, 0)]
        want_new = bool(ready) and (now - ready[0].arrival >= self.window or len(ready) >= room)
        idle_share = 0.0
        if cur is not None and live:
            idle_share = 1.0 - sum(cur.np[i] for i in live) / float(sum(cur.np))
        if not want_new and idle_share < self.repack_idle:
            return None
        cont = [cur.reqs[i] for i in live]
        new = []
        need = 0
        while want_new and len(new) < len(ready):
            # the new requests' slots were reserved when they were staged
            take = ready[:len(new) + 1]
            b = ia.estimate_bytes([r.spec for r in cont + take]) - sum(r.bytes for r in take)
            if self.used + b > self.budget:
                break
            new.append(ready[len(new)])
            need = b
        if not new:
            if idle_share < self.repack_idle or not cont:
                return None
            need = ia.estimate_bytes([r.spec for r in cont])
            if self.used + need > self.budget:
                return None
        for r in new:
            staged.remove(r)
            r.admitted = now
        self.used += need
        self.stats["groups"] += 1
        self.log.append({"planned": round(now, 4), "cont": len(cont), "new": len(new)})
        return (self.pool.submit(_Slots, cont, new, self.graphs, self.options, self.bin_span), cont, new, need)

    def _read(self, ctx, i):
        r = ctx.reqs[i]
        b = ctx.batch
        r.series = b.series(i) if r.series_every > 0 else None
        v = b.vertices(i)
        r.final = {k: v[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ix", "iy", "bondi")}
        r.finished = time.perf_counter() - self.t0
        ctx.read[i] = True

    def _switch(self, cur, pend):
        """The built context takes over, with the GPU idle between chunks:
        the still-running requests' state is copied across from cur and the
        new ones' from their staging contexts, all device to device."""
        fut, cont, new, need = pend
        nxt = fut.result()  # a failed build raises here
        t = time.perf_counter()
        if cont:
            slot = {id(r): i for i, r in enumerate(cur.reqs)}
            for j, r in enumerate(cont):
                i = slot[id(r)]
                nxt.batch.copy_from(j, cur.batch, i)
                nxt.left[j] = cur.left[i]
                nxt.read[j] = cur.read[i]
        for j, r in enumerate(new):
            st = r.staged.result()  # a failed staging raises here
            nxt.batch.copy_from(nxt.ncont + j, st, 0)
            nxt.left[nxt.ncont + j] = r.steps
            self.pool.submit(st.close)
            self.used -= r.bytes  # the staging context; its slot stays reserved
            r.staged = None
        held = need + sum(r.bytes for r in new)
        nxt.reserved = max(held, nxt.bytes)
        self.used += nxt.reserved - held
        e = self.log[-1]
        e.update({k: round(v, 4) for k, v in nxt.times.items()})
        e["switched"] = round(time.perf_counter() - self.t0, 4)
        e["copy"] = round(time.perf_counter() - t, 4)
        return nxt

    def _retire(self, ctx):
        # freeing device memory can wait for the whole device: the builder
        # thread does it, before any later build allocates
        self.pool.submit(ctx.batch.close)
        self.used -= ctx.reserved

    def _loop_repack(self, queue, idle_sleep):
        cur, pend, staged = None, None, []
        try:
            while queue or staged or cur is not None or pend is not None:
                now = time.perf_counter() - self.t0
                progressed = self._stage_arrivals(queue, staged, now) > 0
                if pend is None:
                    pend = self._plan(staged, cur, now)
                    progressed = progressed or pend is not None
                if cur is None:
                    if pend is not None and pend[0].done():
                        cur = self._switch(None, pend)
                        pend = None
                        progressed = True
                elif cur.batch.done():
                    for i, n in enumerate(cur.left):
                        if n == 0 and not cur.read[i]:
                            self._read(cur, i)
                    if pend is not None and pend[0].done():
                        nxt = self._switch(cur, pend)
                        self._retire(cur)
                        cur, pend = nxt, None
                        for i, n in enumerate(cur.left):
                            if n == 0 and not cur.read[i]:
                                self._read(cur, i)
                    live = cur.live()
                    if self.chunks and self.chunks[-1][2] is None:
                        self.chunks[-1][2] = round(time.perf_counter() - self.t0, 4)
                    if live:
                        n = min(self.chunk, min(cur.left[i] for i in live))
                        t = time.perf_counter()
                        cur.batch.md_steps(n)
                        self.chunks.append([round(t - self.t0, 4), round(time.perf_counter() - t, 4), None, n,
                                            len(cur.reqs), len(live)])
                        for i in live:
                            cur.left[i] -= n
                        progressed = True
                    elif pend is None:
                        self._retire(cur)
                        cur = None
                        progressed = True
                    else:
                        # nothing to run until the pending context is ready
                        t = time.perf_counter()
                        pend[0].result()
                        self.idle += time.perf_counter() - t
                        progressed = True
                    self.stats["max_resident_replicas"] = max(self.stats["max_resident_replicas"], len(live))
                self.stats["max_active_groups"] = max(self.stats["max_active_groups"],
                                                     (cur is not None) + (pend is not None))
                if not progressed:
                    time.sleep(idle_sleep)
        finally:
            # staged requests never admitted (an error ended the loop)
            for r in staged:
                if r.staged is not None:
                    self.pool.submit(lambda f=r.staged: f.result().close())


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
