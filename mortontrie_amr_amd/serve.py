"""Continuous-batching server for flow cases.

Requests (case dictionaries with a solid boundary) arrive over time.  The
server keeps every running request as a member of one packed case, so all of
them share each level sub-step's launch.  It advances the packed case in
chunks of `chunk` coarse steps; between chunks it
  * marks requests that finished (max_steps reached, or the drag averaged
    over the last chunk moved by less than `tol` from the previous chunk's),
  * every `admit_every` chunks (or at once when nothing is left running)
    admits arrived requests while their device footprint fits `hbm_budget`
    and repacks, dropping the finished members and carrying each running
    member's state over on the device (mt_case_copy_member), so a member
    computes exactly what it would compute alone.  Between repacks a finished
    member keeps stepping with the batch; its result is already final.
  * scans every member for non-finite populations on the device
    (mt_scan_members, the reference's NaN checks per member); a member that has
    any, or whose drag is not finite, finishes as diverged (`failed`) at once.
  * repacks to evict the finished and diverged members, whatever the admission
    interval, when that pays: their share of the batch's footprint times the
    chunks the batch may still run (the most any running member has left)
    reaches `repack_chunks`, the cost of a repack in chunks of the batch.
Latency is completion time minus arrival time, on the server's wall clock.
"""
from __future__ import annotations

import time
from collections import deque

import numpy as np

from . import amr, cases as C


def footprint_bytes(case, precision):
    """Device bytes of one case: per slot 3Q populations (f and two fcol), two
    rho/u buffers and the tables; slots counted as cells rounded up to bricks
    (x1.3 for partly filled bricks)."""
    pb = {amr.FP64: 8, amr.FP32_SHIFTED: 4, amr.FP16_STORAGE: 2}[precision]
    rb = 8 if precision == amr.FP64 else 4
    cells = sum(lv["keys"].size for lv in case["levels"])
    return int(cells * 1.3 * (3 * case["Q"] * pb + 8 * rb + 24))


# Evicting pays when the retired share x the chunks left reaches this many
# chunks.  A repack (pack, build the packed case's tables, copy the running
# members' state, capture the new case's graphs on its first chunk) costs
# about 30 ms on one MI355X against about 38 ms per 100-step chunk of a full
# batch of served requests; thresholds of 0.75, 1.5 and 3 chunks served the
# bench's stream alike within the run-to-run spread.
REPACK_CHUNKS = 1.5


class Request:
    def __init__(self, rid, case, arrival, max_steps=2000, tol=1e-4):
        self.id, self.case, self.arrival = rid, case, arrival
        self.max_steps, self.tol = max_steps, tol
        self.steps = 0
        self.cd = []
        self.admitted = None
        self.done = None
        self.failed = False

    def cd_of(self, fx_per_call):
        cd = self.case["cd"]
        return 2.0 * fx_per_call / (cd["U"] ** 2 * np.pi * cd["d"] ** 2 / 4.0)

    def finished(self):
        if self.steps >= self.max_steps:
            return True
        return len(self.cd) >= 2 and abs(self.cd[-1] - self.cd[-2]) < self.tol


class Server:
    """chunk_seconds: None serves in wall-clock time; a number makes time virtual
    (each chunk takes that long, idle waits are skipped) for reproducible tests."""

    def __init__(self, precision=amr.FP32_SHIFTED, chunk=100, hbm_budget=200 << 30, chunk_seconds=None,
                 slot_dims=(100, 48, 48), admit_every=1, repack_chunks=REPACK_CHUNKS):
        self.precision, self.chunk, self.budget = precision, chunk, hbm_budget
        self.admit_every = admit_every
        self.repack_chunks = repack_chunks
        self.chunk_seconds = chunk_seconds
        self.vtime = 0.0
        # every running request keeps one slot (a fixed box along x) for its life,
        # so its packed keys are encoded once and repacks move no state
        self.slot_dims = slot_dims
        self.slot_of = {}
        self.batch = None      # packed AmrCase of the running requests
        self.members = []      # running requests, in member order
        self.repacks = 0
        self.chunks = 0
        self.repack_log = []   # (server time, member ids) of every repack

    def _now(self, t0):
        return self.vtime if self.chunk_seconds is not None else time.perf_counter() - t0

    def _wait(self, dt):
        if self.chunk_seconds is not None:
            self.vtime += dt
        else:
            time.sleep(dt)

    def _slot(self, r):
        if id(r) not in self.slot_of:
            used = set(self.slot_of.values())
            k = 0
            while k in used:
                k += 1
            self.slot_of[id(r)] = k
        return self.slot_of[id(r)]

    def _evict_due(self):
        """Whether evicting the finished members pays for a repack."""
        done = [r for r in self.members if r.done is not None]
        active = [r for r in self.members if r.done is None]
        if not done or not active:
            return False
        fp = [footprint_bytes(r.case, self.precision) for r in done]
        total = sum(fp) + sum(footprint_bytes(r.case, self.precision) for r in active)
        left = max(-(-(r.max_steps - r.steps) // self.chunk) for r in active)
        return sum(fp) / total * left >= self.repack_chunks

    def _repack(self, running):
        live = {id(r) for r in running}
        self.slot_of = {k: v for k, v in self.slot_of.items() if k in live}
        for r in running:
            d = r.case["dims0"]
            if any(d[a] > self.slot_dims[a] - (4 if a == 0 else 0) for a in range(3)):
                raise ValueError(f"request {r.id} does not fit a serving slot {self.slot_dims}")
        slots = [self._slot(r) for r in running]
        sx = self.slot_dims[0]
        origins = [(k * sx, 0, 0) for k in slots]
        dims0 = ((max(slots) + 1) * sx, self.slot_dims[1], self.slot_dims[2])
        packed = C.pack([r.case for r in running], origins=origins, dims0=dims0)
        g = amr.AmrCase(packed, precision=self.precision, graphs=True)
        if self.batch is not None:
            old = {id(r): i for i, r in enumerate(self.members)}
            for j, r in enumerate(running):
                if id(r) in old:
                    g.copy_member_from(j, self.batch, old[id(r)])
            self.batch.close()
        self.batch, self.members = g, list(running)
        self.repacks += 1

    def run(self, requests):
        """Serve `requests` (sorted by arrival, seconds from now); returns them finished."""
        t0 = time.perf_counter()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        running, done = [], []     # running: the batch's members, finished or not
        since = self.admit_every   # chunks since the last repack
        while pending or any(r.done is None for r in running):
            now = self._now(t0)
            active = [r for r in running if r.done is None]
            arrived = bool(pending) and pending[0].arrival <= now
            if (arrived and (since >= self.admit_every or not active)) or (running and not active) or self._evict_due():
                used = sum(footprint_bytes(r.case, self.precision) for r in active)
                admitted = []
                while pending and pending[0].arrival <= now:
                    need = footprint_bytes(pending[0].case, self.precision)
                    if (active or admitted) and used + need > self.budget:
                        break
                    r = pending.popleft()
                    r.admitted = now
                    admitted.append(r)
                    used += need
                running = active + admitted
                if running:
                    self._repack(running)
                    self.repack_log.append((now, [r.id for r in running]))
                elif self.batch is not None:
                    self.batch.close()
                    self.batch, self.members = None, []
                since = 0
            if not running:
                self._wait(max(0.0, pending[0].arrival - now))
                continue
            self.batch.step(self.chunk)
            since += 1
            sums, n = self.batch.ib_force_members(reset=True)
            bad = self.batch.scan_members()
            self.chunks += 1
            if self.chunk_seconds is not None:
                self.vtime += self.chunk_seconds
            now = self._now(t0)
            for i, r in enumerate(self.members):
                if r.done is not None:
                    continue
                r.steps += self.chunk
                r.cd.append(r.cd_of(sums[i, 0] / max(n, 1)))
                if bad[i] > 0 or not np.isfinite(r.cd[-1]):
                    r.failed = True
                if r.failed or r.finished():
                    r.done = now
                    done.append(r)
        if self.batch is not None:
            self.batch.close()
            self.batch, self.members = None, []
        return done


def run_alone(req, precision=amr.FP64, chunk=100):
    """The plain path: one request on its own case, the same chunked rule."""
    g = amr.AmrCase(req.case, precision=precision, graphs=True)
    r = Request(req.id, req.case, 0.0, req.max_steps, req.tol)
    while not r.finished():
        g.step(chunk)
        s, n = g.ib_force(reset=True)
        r.steps += chunk
        r.cd.append(r.cd_of(s[0] / max(n, 1)))
        if g.scan_members()[0] > 0 or not np.isfinite(r.cd[-1]):
            r.failed = True
            break
    g.close()
    return r


def poisson_requests(n, rate=20.0, seed=1, max_steps=2000, tol=1e-4):
    """n requests with exponential inter-arrival times (mean 1/rate s)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        out.append(Request(i, C.sphere_request(rng), t, max_steps, tol))
    return out


# request statuses of mt_server_poll (include/mt_amr.h)
REQ_FINISHED, REQ_DIVERGED = 0, 1


class NativeServer:
    """The same serving loop inside the library (mt_server_*): submit request
    cases, call step() until nothing runs or waits, collect with poll()."""

    def __init__(self, precision=amr.FP32_SHIFTED, chunk=100, admit_every=1, hbm_budget=200 << 30):
        import ctypes as C
        self._C = C
        L = amr.load()
        h = C.c_void_p()
        amr._check(L.mt_server_create(precision, chunk, admit_every, hbm_budget, C.byref(h)))
        self.h = h
        self.precision = precision

    def submit(self, rid, case, max_steps=2000, tol=1e-4):
        L = amr.load()
        req = amr.AmrCase(case, precision=self.precision, finalize=False).detach()
        cd = case["cd"]
        amr._check(L.mt_server_submit(self.h, req, rid, max_steps, tol, cd["U"], cd["d"]))

    def step(self):
        C = self._C
        nr, nw = C.c_int32(), C.c_int32()
        amr._check(amr.load().mt_server_step(self.h, C.byref(nr), C.byref(nw)))
        return nr.value, nw.value

    def poll(self, max_results=256):
        """Finished requests: (id, steps, last chunk's Cd, diverged) in completion order."""
        C = self._C
        ids = (C.c_int64 * max_results)()
        steps = (C.c_int32 * max_results)()
        cd = (C.c_double * max_results)()
        st = (C.c_int32 * max_results)()
        n = C.c_int32()
        amr._check(amr.load().mt_server_poll(self.h, ids, steps, cd, st, max_results, C.byref(n)))
        return [(ids[i], steps[i], cd[i], st[i] == REQ_DIVERGED) for i in range(n.value)]

    def close(self):
        if getattr(self, "h", None) is not None and amr._lib is not None:
            amr._lib.mt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

