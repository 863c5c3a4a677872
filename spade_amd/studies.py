"""Workloads of the batched / served FP32-flux form and of its precision study.

* ``ensemble``: independent periodic cases (TGV or isentropic vortex by a
  seeded coin flip, cells per axis from {32, 48, 64, 96}, Mach ~ U[0.1, 0.5],
  Re log-uniform in [100, 5000]) advanced together by the batched engine,
  compared with the plain FP64 path stepping the same cases one at a time.
* ``serving``: a Poisson request stream (cells per axis from {32, 64, 96, 128},
  steps ~ U[100, 2000]) through spade_amd.serve.Server: p50 / p99 completion
  latency and aggregate cell-updates/s.
* ``precision``: one TGV run twice from the same initial state, FP32 fluxes
  with FP64 state (batched engine) against the FP64 path; L2 and Linf error
  per conserved variable and the relative error of kinetic energy and of its
  dissipation rate -dKE/dt every ``every`` steps.

A cell-update is one cell advanced by one full SSP-RK3 step (three right-hand
sides and three stage combinations).
"""
import math
import time

import numpy as np

from . import (Batch, CartesianGrid, CaseSpec, GridArray, IdealGas, Integrator, compose, cent_keep,
               create_exchange, diagnostics, power_law, visc_lr)
from .serve import Server, poisson_requests, summarize

BYTES_PER_BLOCK = Batch.BYTES_PER_BLOCK


def case_constants(spec):
    """Gas constant, viscosity, reference temperature and domain lengths the
    batched engine uses for a case (Batch::setup_case, batch_host.cpp)."""
    g = spec.gamma
    if spec.kind == "tgv":
        c0 = 1.0 / spec.mach
        R = c0 * c0 / g                     # p0 / (rho0 T0), rho0 = T0 = u0 = 1
        mu = 1.0 / spec.reynolds
        L = (2.0 * math.pi,) * 3
    else:
        R = 1.0 / g
        mu = spec.mach * 1.0 / spec.reynolds    # rho0 u0 / Re, u0 = Mach c0, c0 = 1
        L = (10.0,) * 3
    return {"gamma": g, "R": R, "mu": mu, "T_ref": 1.0, "L": L}


def cons_to_prim(cons, gamma, R):
    """(blocks, 5, 16, 16, 16) conserved -> (blocks, 16, 16, 16, 5) primitives."""
    rho, E, mu, mv, mw = (cons[:, v] for v in range(5))
    u, v, w = mu / rho, mv / rho, mw / rho
    p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v + w * w))
    T = p / (R * rho)
    return np.stack([p, T, u, v, w], axis=-1)


def prim_to_cons(prim, gamma, R):
    p, T, u, v, w = (prim[..., v] for v in range(5))
    rho = p / (R * T)
    E = 0.5 * rho * (u * u + v * v + w * w) + p / (gamma - 1.0)
    return np.stack([rho, E, rho * u, rho * v, rho * w], axis=1)


class FP64Twin:
    """The plain FP64 path (grid_array + flux_div + ssprk3 integrator) on the
    initial state and time step of a batched case."""

    def __init__(self, spec, cons0, dt):
        k = case_constants(spec)
        self.gamma, self.R = k["gamma"], k["R"]
        nb = tuple(c // 16 for c in spec.cells)
        self.grid = CartesianGrid((16, 16, 16), nb, tuple((0.0, L) for L in k["L"]), num_exch=2)
        host = np.zeros(self.grid.host_shape)
        host[:, 2:-2, 2:-2, 2:-2] = cons_to_prim(cons0, self.gamma, self.R)
        self.q = GridArray(self.grid).upload(host)
        del host
        create_exchange(self.q).exchange(self.q)
        self.gas = IdealGas(self.gamma, self.R)
        law = power_law(k["mu"], k["T_ref"], spec.power, spec.prandtl)
        self.flux = compose(cent_keep(2, self.gas), visc_lr(law, self.gas))
        self.it = Integrator(self.q, self.flux, dt)

    def prims(self):
        return self.q.download()[:, 2:-2, 2:-2, 2:-2]

    def close(self):
        self.it.close()
        self.q.close()
        self.grid.close()


def ensemble_specs(n=300, steps=500, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = "tgv" if rng.random() < 0.5 else "vortex"
        cells = tuple(int(c) for c in rng.choice([32, 48, 64, 96], size=3))
        mach = float(rng.uniform(0.1, 0.5))
        re = float(math.exp(rng.uniform(math.log(100.0), math.log(5000.0))))
        out.append(CaseSpec(kind, cells, mach=mach, reynolds=re, steps=steps))
    return out


def run_ensemble(specs, steps, tune=True):
    """All cases resident at once, advanced together ``steps`` SSP-RK3 steps."""
    blocks = sum(s.blocks for s in specs)
    b = Batch(int(blocks * BYTES_PER_BLOCK * 1.02) + (64 << 20))
    ids = [b.add(s) for s in specs]
    tlog = []
    if tune:
        b.tune()
        tlog = b.tune_log()
    shape = b.shape()
    b.step(1)
    b.ctx.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.ctx.sync()
    dt = time.perf_counter() - t0
    cells = sum(s.num_cells for s in specs)
    # algorithmic bytes of one step per cell: stage 1 reads u0 and writes stage 1
    # (80 B), stages 2 and 3 read their input and u0 and write (120 B each)
    alg_bytes = cells * steps * (80.0 + 120.0 + 120.0)
    out = {"cases": len(specs), "blocks": blocks, "cells": cells, "steps": steps, "seconds": dt,
           "cell_updates_per_s": cells * steps / dt, "mcell_rhs_per_s": 3.0 * cells * steps / dt * 1e-6,
           "hbm_algorithmic_gb_s": alg_bytes / dt * 1e-9, "hbm_frac": alg_bytes / dt / 8e12, "tuned_shape": shape,
           "tune_log": tlog}
    b.close()
    del ids
    return out


def fp64_rate(spec, steps, warmup=3):
    """Cell-updates/s of the plain FP64 path on one case (initial state from the engine),
    timed over `steps` SSP-RK3 steps after `warmup` untimed ones; returns (cell-updates,
    seconds)."""
    b = Batch(spec.blocks * BYTES_PER_BLOCK + (16 << 20))
    cid = b.add(spec)
    cons0, dt = b.download(cid), b.info(cid)["dt"]
    b.close()
    tw = FP64Twin(spec, cons0, dt)
    tw.it.advance(warmup)
    tw.grid.ctx.sync()
    tw.grid.ctx.timer_start()
    tw.it.advance(steps)
    ms = tw.grid.ctx.timer_stop()
    tw.close()
    return spec.num_cells * steps, ms * 1e-3


def serving_spec(rng):
    """One request of the served stream: TGV or isentropic vortex by coin flip,
    cells per axis from {32, 64, 96, 128}, Mach ~ U[0.1, 0.5], Re log-uniform
    in [100, 5000], steps uniform in [100, 2000]."""
    kind = "tgv" if rng.random() < 0.5 else "vortex"
    cells = tuple(int(c) for c in rng.choice([32, 64, 96, 128], size=3))
    mach = float(rng.uniform(0.1, 0.5))
    re = float(math.exp(rng.uniform(math.log(100.0), math.log(5000.0))))
    steps = int(rng.integers(100, 2001))
    return CaseSpec(kind, cells, mach=mach, reynolds=re, steps=steps)


def serving(n=60, rate=8.0, seed=7, max_chunk=8, budget_gb=64):
    reqs = poisson_requests(n, rate, serving_spec, seed=seed)
    b = Batch(int(budget_gb) << 30)
    srv = Server(b, max_chunk=max_chunk, keep_state=False)
    res = srv.run(reqs)
    out = summarize(res, srv.makespan, srv.cell_updates)
    out.update({"rate_per_s": rate, "max_chunk": max_chunk, "rounds": srv.rounds})
    b.close()
    return out, reqs


def precision(cells=256, steps=1000, every=100, mach=0.1, reynolds=1600.0, log=None):
    """FP32-flux batched run against the FP64 path from the same TGV state."""
    spec = CaseSpec("tgv", (cells,) * 3, mach=mach, reynolds=reynolds, steps=steps)
    b = Batch(spec.blocks * BYTES_PER_BLOCK + (64 << 20))
    cid = b.add(spec)
    cons0, dt = b.download(cid), b.info(cid)["dt"]
    tw = FP64Twin(spec, cons0, dt)
    del cons0
    rows = []
    done = 0
    ke64 = {}
    for m in range(1, steps // every + 1):
        target = m * every
        # FP64: up to `target`, kinetic energy at target and target + 1
        tw.it.advance(target - done)
        ke64[target] = diagnostics(tw.q, tw.gas)[0]
        ref = tw.prims()
        tw.it.advance(1)
        ke64[target + 1] = diagnostics(tw.q, tw.gas)[0]
        # FP32-flux engine: to target, state; then two more steps for the KE history
        b.step(target - b.info(cid)["steps_done"])
        got = cons_to_prim(b.download(cid), tw.gamma, tw.R)
        b.step(2)
        ke32, _ = b.diag(cid)
        n32 = b.info(cid)["steps_done"]
        k32 = {n32 - len(ke32) + s: ke32[s] for s in range(len(ke32))}
        done = target + 1
        c_ref = prim_to_cons(ref, tw.gamma, tw.R)
        c_got = prim_to_cons(got, tw.gamma, tw.R)
        del ref, got
        row = {"step": target}
        for v, name in enumerate(["rho", "rhoE", "rhou", "rhov", "rhow"]):
            d = c_got[:, v] - c_ref[:, v]
            s = c_ref[:, v]
            row[f"{name}_l2"] = float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(s * s)), 1e-300))
            row[f"{name}_linf"] = float(np.max(np.abs(d)) / max(np.max(np.abs(s)), 1e-300))
        del c_ref, c_got
        row["ke_rel"] = abs(k32[target] - ke64[target]) / ke64[target]
        eps64 = -(ke64[target + 1] - ke64[target]) / dt
        eps32 = -(k32[target + 1] - k32[target]) / dt
        row["dissipation_fp64"] = eps64
        row["dissipation_rel"] = abs(eps32 - eps64) / abs(eps64)
        rows.append(row)
        if log:
            log(row)
    b.close()
    tw.close()
    return {"cells": cells, "steps": steps, "dt": dt, "rows": rows}


def channel(seed=1234, nb=(12, 8, 6), beta=1.5, mach=0.2, re_bulk=5600.0, pert=0.05):
    """Wall-bounded channel on stretched coordinates (192 x 128 x 96 cells by
    default): x in [0, 4 pi], y in [-1, 1] with the tanh map of factor ``beta``
    clustering cells at the walls, z in [0, 4 pi / 3]; periodic in x and z,
    isothermal no-slip walls (T_wall = 1) at y = -1 and 1.  Laminar Poiseuille
    profile u = 1.5 U_b (1 - y^2) with U_b = 1, plus seeded uniform random
    velocity perturbations of amplitude ``pert`` U_b.  Mach number of U_b at
    the wall temperature, bulk Reynolds number 2 delta U_b / nu (5600 gives
    Re_tau near 180).

    Returns a dict with the grid arguments, the host primitive array (ghosts
    zero), gas, viscous law and the time step for CFL 0.5."""
    gamma, T_w = 1.4, 1.0
    c_w = 1.0 / mach
    R = c_w * c_w / (gamma * T_w)
    p0 = 1.0 * R * T_w
    mu = 2.0 * 1.0 / re_bulk
    lo, hi = (0.0, -1.0, 0.0), (4.0 * math.pi, 1.0, 4.0 * math.pi / 3.0)
    nc = (16, 16, 16)
    n = [nb[d] * nc[d] for d in range(3)]
    dxi = [(hi[d] - lo[d]) / n[d] for d in range(3)]
    tb = math.tanh(beta)
    xi_y = lo[1] + (np.arange(n[1]) + 0.5) * dxi[1]
    y = np.tanh(beta * xi_y) / tb
    rng = np.random.default_rng(seed)
    prim = np.zeros((n[2], n[1], n[0], 5))
    prim[..., 0] = p0
    prim[..., 1] = T_w
    prim[..., 2] = (1.5 * (1.0 - y * y))[None, :, None]
    prim[..., 2:5] += pert * rng.uniform(-1.0, 1.0, size=(n[2], n[1], n[0], 3))
    # to blocks (x fastest), with 2 ghost layers
    ng = 2
    host = np.zeros((nb[0] * nb[1] * nb[2], nc[2] + 2 * ng, nc[1] + 2 * ng, nc[0] + 2 * ng, 5))
    lb = 0
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                host[lb, ng:-ng, ng:-ng, ng:-ng] = prim[bz * 16:(bz + 1) * 16, by * 16:(by + 1) * 16,
                                                        bx * 16:(bx + 1) * 16]
                lb += 1
    dy_min = dxi[1] * beta * (1.0 - math.tanh(beta * (1.0 - 0.5 * dxi[1])) ** 2) / tb
    umax = float(np.max(np.abs(prim[..., 2:5])))
    dt = 0.5 * min(dxi[0], dy_min, dxi[2]) / (umax + math.sqrt(gamma * R * T_w))
    return {"nb": nb, "nc": nc, "ng": ng, "lo": lo, "hi": hi, "tanh_y": (-1.0, 1.0, beta),
            "host": host, "gamma": gamma, "R": R, "mu": mu, "T_wall": T_w, "dt": dt,
            "power": 0.76, "prandtl": 0.72, "pdiss": (0.5, 0.1, 0.1)}
