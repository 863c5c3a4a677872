"""Continuous-batching request loop over the batched engine.

Requests (one case each: TGV or isentropic vortex, with its own size, Mach,
Reynolds number and step count) arrive over time.  The loop

  * admits queued requests in arrival order while their HBM footprint (16^3
    blocks x three FP64 state copies) fits in the engine's pool, letting a
    later, smaller request pass a head request that does not fit yet
    (backfill), so the device stays busy;
  * keeps every admitted case resident on the device and advances all of them
    together, a few SSP-RK3 steps per round (at most max_chunk while requests
    are still due to arrive, so a new one waits at most that long; once all
    have arrived, as many as the shortest remaining case allows), so newly
    admitted cases join the running ones and finished cases leave between
    rounds;
  * retires a case when it has done its step count: its final conserved
    fields and its per-step kinetic-energy / enstrophy history are copied back
    and its slots are freed for the queue;
  * retires a case early, with an error, once its kinetic energy is no longer
    finite (the engine's status(id); the reference's driver stops on a NaN
    wave speed, development/cuda-tgv/main.cc:250-259): its slots go back to
    the queue and the other cases go on, unaffected (cases never read each
    other's blocks);
  * stops backfilling once the head request has waited max_head_wait seconds:
    freed slots are then kept for it, so a stream of small requests cannot
    starve a large one.

Per-case arithmetic does not depend on which other cases share the device
(the work list only changes which blocks a kernel visits, never how a block
is computed; per-case reductions run in the case's own block order), so a
case's result is bit-for-bit the same alone and in any mix.

The engine is anything with fits(spec), add(spec) -> id, step(k), download(id),
diag(id), remove(id); spade_amd.Batch is the device one.
"""
import collections
import time

import numpy as np


class Request:
    def __init__(self, rid, spec, arrival):
        self.rid, self.spec, self.arrival = rid, spec, float(arrival)


class Result:
    def __init__(self, request, admitted, completed, state, ke, enstrophy, error=None):
        self.request = request
        self.admitted, self.completed = admitted, completed
        self.state, self.ke, self.enstrophy = state, ke, enstrophy
        self.error = error          # None, or why the case was retired before its step count

    @property
    def latency(self):
        return self.completed - self.request.arrival


def poisson_requests(n, rate, make_spec, seed=0):
    """n requests with exponential inter-arrival times of mean 1/rate seconds."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        out.append(Request(rid, make_spec(rng), t))
    return out


class Server:
    def __init__(self, engine, max_chunk=8, backfill=True, keep_state=True, clock=time.perf_counter,
                 sleep=time.sleep, max_head_wait=2.0):
        self.engine = engine
        self.max_chunk = int(max_chunk)
        self.backfill = backfill
        self.max_head_wait = float(max_head_wait)
        self.keep_state = keep_state
        self.clock, self.sleep = clock, sleep
        self.rounds = 0
        self.cell_updates = 0

    def run(self, requests):
        """Serve requests (arrival times relative to the start); returns results in completion order."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        active = {}                      # case id -> [request, remaining steps, admitted time]
        results = []
        t0 = self.clock()
        self.start_time = t0
        while pending or queue or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            self._admit(queue, active, now)
            if not active:
                if pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        self.sleep(min(wait, 0.05))
                    continue
                if queue:
                    raise RuntimeError(f"request {queue[0].rid} can never fit in the engine's pool")
                break
            k = min(a[1] for a in active.values())
            if pending:
                k = min(k, self.max_chunk)
            self.engine.step(k)
            self.rounds += 1
            done_at = self.clock() - t0
            status = getattr(self.engine, "status", None)
            for cid in list(active):
                a = active[cid]
                a[1] -= k
                self.cell_updates += k * a[0].spec.num_cells
                bad = status(cid) if status is not None else -1
                if a[1] == 0 or bad >= 0:
                    state = self.engine.download(cid) if self.keep_state else None
                    ke, en = self.engine.diag(cid)
                    self.engine.remove(cid)
                    err = None if bad < 0 else f"diverged: non-finite kinetic energy at step {bad}"
                    results.append(Result(a[0], a[2], done_at, state, ke, en, err))
                    del active[cid]
        self.makespan = self.clock() - t0
        return results

    def _admit(self, queue, active, now):
        i = 0
        while i < len(queue):
            req = queue[i]
            if self.engine.fits(req.spec):
                cid = self.engine.add(req.spec)
                active[cid] = [req, req.spec.steps, now]
                del queue[i]
                continue
            # the head keeps the freed slots once it has waited too long
            if not self.backfill or (i == 0 and now - req.arrival > self.max_head_wait):
                break
            i += 1


def summarize(results, makespan, cell_updates):
    lat = np.array([r.latency for r in results])
    return {
        "requests": len(results),
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "makespan_s": makespan,
        "cell_updates_per_s": cell_updates / makespan if makespan > 0 else None,
    }
