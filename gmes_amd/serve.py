"""A request loop that serves FDTD solves on one GPU with continuous batching.

Requests arrive over time.  Each is a callable that builds an initialised
simulation and a result monitor, plus a step count.  The loop:

* admits queued requests while the sum of their estimated HBM footprints
  fits the budget and fewer than ``max_active`` run: in arrival order, or
  with ``policy="shortest"`` the queued request of least estimated work
  first (lower median latency when many wait; a request that does not fit
  is passed over only while a smaller one does);
* builds admitted requests on worker threads (host tables, device
  allocation and uploads), so setup overlaps with the GPU running others;
* advances every active simulation together, as one batch: the main,
  source, ghost, ADE and DFT kernels of all of them run as one launch each
  (``gmes_amd.batch``), in chunks of at most ``chunk`` steps that end
  where the next active request finishes;
* between chunks, retires finished requests (reads their result, frees
  their device memory) and lets newly built ones join, so requests that
  arrive while others run are admitted as slots and memory free up and
  their work overlaps on the device;
* isolates failures: a request whose build raises, whose kernels report a
  deferred error (the DM2 corrector's non-convergence, src/pw_dm2.hh:
  398-402) or whose result cannot be read is retired with ``error`` set
  at the next retirement point, and the others keep running.

Each simulation keeps its own fields and clock, and batching does not
change its arithmetic: a result is the same alone and in any mix.  Latency
is completion time minus arrival time.  There is no counterpart in the
reference (one FDTD per Python process).
"""

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .batch import Batch


class Request(object):
    def __init__(self, rid, build, steps, arrival=0.0, footprint=0, work=None):
        """build() -> (sim, result_fn): an initialised FDTD and a function
        returning the request's result once it has run ``steps`` steps.
        footprint: estimated device bytes, used for admission; work: an
        estimate of its cost (cell-updates), used by the "shortest" policy
        (default footprint x steps)."""
        self.rid = rid
        self.build = build
        self.steps = int(steps)
        self.arrival = float(arrival)
        self.footprint = int(footprint)
        self.work = float(work) if work is not None else float(self.footprint) * self.steps
        self.sim = None
        self.result_fn = None
        self.result = None
        self.remaining = self.steps
        self.t_admit = self.t_start = self.t_done = None
        self.error = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class Server(object):
    def __init__(self, hbm_budget, max_active=32, chunk=250, workers=4, autotune=False, clock=None,
                 policy="arrival"):
        if policy not in ("arrival", "shortest"):
            raise ValueError("policy must be 'arrival' or 'shortest'")
        self.policy = policy
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.chunk = int(chunk)
        self.workers = int(workers)
        self.autotune = autotune
        self.clock = clock or time.perf_counter
        self.peak_reserved = 0
        self.peak_active = 0
        self.rounds = 0
        # host-side accounting (seconds): waiting with nothing to run,
        # forming batches, enqueueing rounds, waiting for finished rounds
        self.t_idle = self.t_batch = self.t_advance = self.t_sync = 0.0
        self.rebatches = 0

    def run(self, requests):
        """Serve ``requests`` (arrival times relative to the start of the
        call) to completion; returns them with results and timings."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, building, active, done = [], {}, [], []
        reserved = 0
        batch = None
        changed = False
        t0 = self.clock()
        now = lambda: self.clock() - t0
        pool = ThreadPoolExecutor(max_workers=self.workers)
        try:
            while pending or queue or building or active:
                t = now()
                while pending and pending[0].arrival <= t:
                    r = pending.pop(0)
                    if r.footprint > self.hbm_budget:  # can never be admitted
                        r.error = "footprint %d exceeds the HBM budget" % r.footprint
                        r.t_done = now()
                        done.append(r)
                    else:
                        queue.append(r)
                # admission by estimated footprint, in policy order
                if self.policy == "shortest":
                    queue.sort(key=lambda q: (q.work, q.arrival))
                while queue and len(active) + len(building) < self.max_active:
                    fits = [q for q in queue if reserved + q.footprint <= self.hbm_budget]
                    if not fits or (self.policy == "arrival" and fits[0] is not queue[0]):
                        break
                    r = fits[0]
                    queue.remove(r)
                    reserved += r.footprint
                    r.t_admit = now()
                    building[r] = pool.submit(r.build)
                self.peak_reserved = max(self.peak_reserved, reserved)
                for r, fut in list(building.items()):
                    if fut.done():
                        del building[r]
                        try:
                            r.sim, r.result_fn = fut.result()
                        except Exception as e:  # a failed build is reported, not served
                            r.error = repr(e)
                            r.t_done = now()
                            reserved -= r.footprint
                            done.append(r)
                            continue
                        r.t_start = now()
                        active.append(r)
                        changed = True
                if not active:
                    ti = now()
                    if building:
                        next(iter(building.values())).exception()  # wait for a build
                    elif pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
                    self.t_idle += now() - ti
                    continue
                if changed:
                    tb = now()
                    if batch is not None:
                        batch.close()
                    batch = Batch([r.sim for r in active], autotune=self.autotune)
                    changed = False
                    self.rebatches += 1
                    self.t_batch += now() - tb
                self.peak_active = max(self.peak_active, len(active))
                n = min(self.chunk, min(r.remaining for r in active))
                ta = now()
                batch.advance(n)
                self.t_advance += now() - ta
                self.rounds += 1
                for r in active:
                    r.remaining -= n
                if any(r.remaining == 0 for r in active):
                    ts = now()
                    # a member's device error retires that request alone
                    errors = batch.member_errors()
                    self.t_sync += now() - ts
                    for i, msg in errors.items():
                        r = next(q for q in active if q.sim is batch.sims[i])
                        r.error = msg
                    retired = [r for r in active if r.remaining == 0 or r.error is not None]
                    for r in retired:
                        if r.error is None:
                            try:
                                r.result = r.result_fn()
                            except Exception as e:  # reported on this request only
                                r.error = repr(e)
                        r.t_done = now()
                        active.remove(r)
                        reserved -= r.footprint
                        done.append(r)
                    batch.close()
                    batch = None
                    for r in retired:
                        r.sim.close()
                        r.sim = None
                    changed = True
        finally:
            if batch is not None:
                batch.close()
            pool.shutdown(wait=True)
        self.elapsed = now()
        return done


def latency_stats(done):
    lat = np.array([r.latency for r in done if r.error is None])
    if not len(lat):
        return dict(p50=None, p99=None, mean=None, max=None)
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                mean=float(lat.mean()), max=float(lat.max()))
