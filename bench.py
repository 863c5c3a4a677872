"""Benchmark of the SPADE right-hand-side hot path on one MI355X.

Default workload: FP64 flux_div (cent_keep<2> + visc_lr with the power law)
on the 256^3 Taylor-Green vortex, 16x16x16 blocks of 16^3 cells with 2 ghost
layers, inputs resident in HBM.  One step = one flux_div over the whole grid.
Prints one JSON line (value = Mcell-RHS/s).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--settle S] [--no-cpu-baseline]

Before the W warmup launches, S further untimed launches (default 200, about
0.1 s) bring the kernel to its steady state after the other rows: in a
kernel trace the first ~50 flux_div launches after them take up to 30%
longer (540-630 us, falling to a steady 480 us), and the timed launches are
meant to measure the steady state.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_RHS = 80.0   # read 5 x 8 B primitives + write 5 x 8 B residual
# MI355X VALU issue: a wave64 instruction every 4 cycles per SIMD (FP64 at
# full rate), 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock
VALU_CYCLES_PER_WAVE_INST = 4
SIMDS = 1024
CLOCK_GHZ = 2.4


def tgv_interior(nb, nc, L, p0):
    """Analytic TGV primitives, shape (blocks, nz, ny, nx, 5), blocks x-fastest."""
    n = [nb[d] * nc[d] for d in range(3)]
    dx = [L / n[d] for d in range(3)]
    x = (np.arange(n[0]) + 0.5) * dx[0]
    y = (np.arange(n[1]) + 0.5) * dx[1]
    z = (np.arange(n[2]) + 0.5) * dx[2]
    # (bz, k, by, j, bx, i) views of the global axes
    X = x.reshape(1, 1, 1, 1, nb[0], nc[0])
    Y = y.reshape(1, 1, nb[1], nc[1], 1, 1)
    Z = z.reshape(nb[2], nc[2], 1, 1, 1, 1)
    shape = (nb[2], nc[2], nb[1], nc[1], nb[0], nc[0])
    out = np.empty((nb[2], nb[1], nb[0], nc[2], nc[1], nc[0], 5))

    def put(v, field):
        f = np.broadcast_to(field, shape)
        out[..., v] = f.transpose(0, 2, 4, 1, 3, 5)

    put(0, p0 + (1.0 / 16.0) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0))
    put(1, np.ones((1, 1, 1, 1, 1, 1)))
    put(2, np.sin(X) * np.cos(Y) * np.cos(Z))
    put(3, -np.cos(X) * np.sin(Y) * np.cos(Z))
    put(4, np.zeros((1, 1, 1, 1, 1, 1)))
    return out.reshape(nb[0] * nb[1] * nb[2], nc[2], nc[1], nc[0], 5)


def cpu_baseline(sample_blocks=4, reps=5, max_threads=16, full_blocks=16):
    """The CPU path timed on this host: the oracle's restatement of the
    reference's flux_div_basic (pde-algs/flux-div/flux_div_basic.h:17-76 --
    AoS with the variable fastest, six face evaluations per cell, blocks ->
    k -> j -> i, FP64, -O3), on a bounded sample of the same workload (TGV,
    blocks of 16^3, the same flux), median of reps after one untimed call:
    on one core, and on every core of this host's share (at most max_threads)
    with OpenMP over blocks, as the reference runs one rank per core over its
    share of the blocks (grid/partition.h:27-84).  Then one flux_div of the
    whole full_blocks^3-block grid on every core.  The compiled reference's
    own one-core rate on the same sample, measured where its sources are
    (not on this host), is kept in tests/golden/manifest.txt."""
    from oracle import pyoracle as po
    L, p0 = 2.0 * np.pi, 100.0 / 1.4
    fl = po.make_flux(po.make_term(po.TERM_KEEP, order=2),
                      po.make_term(po.TERM_VISC, law=po.LAW_POWER, mu=1 / 1600, T_ref=1.0, power=0.76, prandtl=0.72))
    gas = po.Gas(1.4, p0)

    def state(b):
        nb, nc, ng = (b,) * 3, (16, 16, 16), 2
        g = po.make_grid(nb, nc, ng, (0, 0, 0), (L, L, L))
        q = np.zeros(po.shape_of(g))
        q[:, ng:-ng, ng:-ng, ng:-ng] = tgv_interior(nb, nc, L, p0)
        po.exchange(g, q)
        return g, q, np.prod(nb) * 4096

    def rate(g, q, cells, n):
        rhs = np.zeros_like(q)
        po.flux_div(g, gas, fl, q, rhs)
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            po.flux_div(g, gas, fl, q, rhs)
            ts.append(time.perf_counter() - t0)
        return cells / sorted(ts)[len(ts) // 2] * 1e-6

    nthreads = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    g, q, cells = state(sample_blocks)
    po.set_threads(1)
    one = rate(g, q, cells, reps)
    used = po.set_threads(nthreads)
    allc = rate(g, q, cells, reps)
    out = {"value": round(allc, 4), "unit": "Mcell-RHS/s", "cores": used, "kind": "port",
           "sample": f"TGV {sample_blocks}^3 blocks of 16^3 ({cells} cells), cent_keep<2> + visc_lr power law, "
                     f"oracle flux_div (AoS, 6 faces per cell), OpenMP over blocks, median of {reps}",
           "cpu": _cpu_model(),
           "one_core": {"value": round(one, 4), "cores": 1, "sample": "same sample, one thread"}}
    del q
    if full_blocks:
        g, q, cells = state(full_blocks)
        t0 = time.perf_counter()
        po.flux_div(g, gas, fl, q)
        sec = time.perf_counter() - t0
        out["full_grid"] = {"value": round(cells / sec * 1e-6, 4), "cores": used, "seconds": round(sec, 3),
                            "sample": f"one flux_div of the whole TGV {full_blocks * 16}^3 grid"}
        del q
    po.set_threads(nthreads)
    return out


def rows_measure(sa, g, q, exch, gas, flux, reps=10):
    """The path's other rows on the same state, through the library's API: the
    stand-alone ghost exchange, one transform-aware SSP-RK3 step, the CFL and
    kinetic-energy / enstrophy reductions, the wall fill and the binary
    checkpoint.  Device rows are timed with HIP events on the library's
    stream (mean of reps calls after one untimed call; a reduction's time
    includes copying its partial results back); algorithmic bytes per call
    over that time give GB/s.  The checkpoint crosses PCIe and host memory and
    is timed on the host.  Leaves q changed (RK steps, wall ghosts)."""
    nc, nbs, ng = g.cells_in_block, g.num_blocks, g.num_exch
    nblk = nbs[0] * nbs[1] * nbs[2]
    stored = (nc[0] + 2 * ng) * (nc[1] + 2 * ng) * (nc[2] + 2 * ng)
    interior = nc[0] * nc[1] * nc[2]
    cells = interior * nblk

    def timed(fn, n=reps):
        fn()
        g.ctx.sync()
        g.ctx.timer_start()
        for _ in range(n):
            fn()
        return g.ctx.timer_stop() / n

    def entry(ms, nbytes, what):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"ms": round(ms, 4), "bytes": int(nbytes), "gb_s": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_are": what}

    out = {}
    out["exchange"] = entry(timed(lambda: exch.exchange(q)), 80.0 * (stored - interior) * nblk,
                            "each ghost cell's 5 doubles read from its source block and written")
    out["max_wavespeed"] = entry(timed(lambda: sa.max_wavespeed(q, gas)), 32.0 * cells,
                                 "T, u, v, w of every interior cell")
    out["diagnostics"] = entry(timed(lambda: sa.diagnostics(q, gas)), 40.0 * cells,
                               "5 doubles of every interior cell (neighbours from cache)")
    nwall = (nbs[0] * nbs[2]) * (nc[0] + 2 * ng) * ng * (nc[2] + 2 * ng)
    out["wall_fill"] = entry(timed(lambda: sa.boundary_fill_wall(q, 2, 1.0)), 80.0 * nwall,
                             "each y-min ghost cell's 5 doubles read from its mirror cell and written")
    it = sa.Integrator(q, flux, 1e-4, scheme=sa.ssprk3)
    ms = timed(lambda: it.advance(1), n=5)
    how = ("3 launches of flux_div fused with the transform-aware stage update; halos read from the "
           "neighbour blocks, ghost cells written by the last stage of each advance() call") if it.fused else \
        "3 flux_div + 3 transform-aware stage updates writing the exchanged ghosts"
    out["ssprk3_step"] = {"ms": round(ms, 4), "cell_updates_per_s": round(cells / (ms * 1e-3), 1),
                          "fused": it.fused, "note": "advance(1) per call: " + how}
    ms10 = timed(lambda: it.advance(10), n=2) / 10.0
    out["ssprk3_step_in_10"] = {"ms": round(ms10, 4), "cell_updates_per_s": round(cells / (ms10 * 1e-3), 1),
                                "note": "advance(10) per call, per step"}
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        path = os.path.join(tmp, "checkpoint.bin")
        t0 = time.perf_counter()
        sa.binary_write(path, q)
        tw = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        sa.binary_read(path, q)
        tr = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out["binary_write"] = {"ms": round(tw * 1e3, 1), "gb_s": round(size / tw / 1e9, 2),
                           "note": "host image of the array (reference layout) to a file in host memory"}
    out["binary_read"] = {"ms": round(tr * 1e3, 1), "gb_s": round(size / tr / 1e9, 2),
                          "note": "the same file back into the device array"}
    return out


def compositions_measure(sa, blocks=8, reps=10):
    """flux_div of the compositions outside the LDS-staged kernel's set on the
    TGV (blocks^3 blocks of 16^3 cells, identity coordinates): the
    compile-time specialised per-cell kernel against the generic one with
    run-time term dispatch (SPD_DISABLE_SPEC=1), mean of reps launches after
    three untimed ones.  The two give the same bits."""
    nb, nc, ng = (blocks,) * 3, (16, 16, 16), 2
    L, gamma, R = 2.0 * np.pi, 1.4, 100.0 / 1.4
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    host = np.zeros(g.host_shape)
    host[:, ng:-ng, ng:-ng, ng:-ng] = tgv_interior(nb, nc, L, R)
    q = sa.GridArray(g).upload(host)
    sa.create_exchange(q).exchange(q)
    rhs = sa.GridArray(g, fill=0.0)
    gas = sa.IdealGas(gamma, R)
    visc = sa.visc_lr(sa.constant_viscosity(1e-3, 0.72), gas)
    cases = {"weno_rusanov": sa.weno(gas), "fweno": sa.fweno(gas), "first_order": sa.first_order(gas),
             "totani_lr": sa.totani_lr(gas), "cent_keep4+visc_lr": sa.compose(sa.cent_keep(4, gas), visc),
             "weno_rusanov+visc_lr": sa.compose(sa.weno(gas), visc)}
    out = {"workload": f"TGV {blocks * 16}^3, identity coordinates, FP64"}
    for name, flux in cases.items():
        row = {}
        for generic in (False, True):
            if generic:
                os.environ["SPD_DISABLE_SPEC"] = "1"
            try:
                kind = sa.flux_div_kernel(q, flux)
                for _ in range(3):
                    sa.flux_div(q, rhs, flux, increment=False)
                g.ctx.sync()
                g.ctx.timer_start()
                for _ in range(reps):
                    sa.flux_div(q, rhs, flux, increment=False)
                ms = g.ctx.timer_stop() / reps
            finally:
                os.environ.pop("SPD_DISABLE_SPEC", None)
            row[kind] = {"ms": round(ms, 4), "mcell_rhs_per_s": round(g.num_cells / (ms * 1e-3) * 1e-6, 1)}
        row["speedup"] = round(row["generic"]["ms"] / row["specialized"]["ms"], 2)
        out[name] = row
    return out


def channel_measure(sa, reps=20, steps=200):
    """flux_div on the stretched wall-bounded channel (192 x 128 x 96 cells,
    12 x 8 x 6 blocks of 16^3, tanh wall-normal map, cent_keep<2> +
    pressure_diss_lr + visc_lr power law): the curvilinear path, with the
    kernel the library selects for it, timed like the headline after the
    exchange and both wall fills it needs.  Bytes per cell-RHS as for the
    headline (the prims read once, the residual written once).  Then the
    channel case as it runs: ``steps`` SSP-RK3 steps (200, the case's length)
    with the wall fills, timed as one advance() call."""
    from spade_amd import studies
    c = studies.channel()
    g = sa.CartesianGrid(c["nc"], c["nb"], tuple(zip(c["lo"], c["hi"])), num_exch=c["ng"],
                         periodic=(True, False, True), maps=(None, sa.tanh_map(*c["tanh_y"]), None))
    gas = sa.IdealGas(c["gamma"], c["R"])
    a, ep, et = c["pdiss"]
    flux = sa.compose(sa.cent_keep(2, gas), sa.pressure_diss_lr(gas, sa.const_sensor(a), ep, et),
                      sa.visc_lr(sa.power_law(c["mu"], 1.0, c["power"], c["prandtl"]), gas))
    q = sa.GridArray(g).upload(c["host"])
    rhs = sa.GridArray(g, fill=0.0)
    sa.create_exchange(q).exchange(q)
    sa.boundary_fill_wall(q, 2, c["T_wall"])
    sa.boundary_fill_wall(q, 3, c["T_wall"])
    for _ in range(5):
        sa.flux_div(q, rhs, flux, increment=False)
    g.ctx.sync()
    g.ctx.timer_start()
    for _ in range(reps):
        sa.flux_div(q, rhs, flux, increment=False)
    ms = g.ctx.timer_stop() / reps
    cells = g.num_cells
    gbs = BYTES_PER_CELL_RHS * cells / (ms * 1e-3) / 1e9
    out = {"ms": round(ms, 4), "mcell_rhs_per_s": round(cells / (ms * 1e-3) * 1e-6, 1), "cells": cells,
           "kernel": sa.flux_div_kernel(q, flux), "gb_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
           "note": "192x128x96 stretched channel, cent_keep<2> + pressure_diss_lr + visc_lr power law, FP64"}
    # the channel case as run: SSP-RK3 at CFL 0.5 with both isothermal walls,
    # 200 steps in one advance() call after 5 untimed
    it = sa.Integrator(q, flux, c["dt"], scheme=sa.ssprk3, walls=(2, 3), t_wall=c["T_wall"])
    it.advance(5)
    g.ctx.sync()
    g.ctx.timer_start()
    it.advance(steps)
    ms_step = g.ctx.timer_stop() / steps
    out["ssprk3_steps"] = {"steps": steps, "ms_per_step": round(ms_step, 4),
                           "cell_updates_per_s": round(cells / (ms_step * 1e-3), 1), "fused": it.fused,
                           "finite": bool(np.isfinite(q.download()).all()),
                           "note": "transform-aware SSP-RK3, dt for CFL 0.5; after the first stage of the call the "
                                   "fused stages read their halo from the neighbour blocks and form the wall ghosts, "
                                   "the last one writes the ghosts and fills the walls"}
    it.close()
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _traffic(workload):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (profiles/traffic.json), or None when not measured."""
    return _profiled(workload)


def _profiled(key):
    """A per-launch figure measured with rocprofv3 counters and kept in
    profiles/traffic.json, or None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def _valu_issue(workload, ms):
    """FP64 VALU issue beside the HBM roofline: the kernel's SQ_INSTS_VALU wave
    instructions (profiles/traffic.json) at 4 cycles each on 1024 SIMDs at the
    2.4 GHz peak clock, against the measured time per launch."""
    insts = _profiled(workload + "_valu_wave_insts")
    if insts is None:
        return None
    issue_us = insts * VALU_CYCLES_PER_WAVE_INST / (SIMDS * CLOCK_GHZ * 1e3)
    return {"wave_insts": insts, "issue_us": round(issue_us, 1), "frac": round(issue_us / (ms * 1e3), 4),
            "note": "VALU issue time (wave instructions x 4 cycles / 1024 SIMDs / 2.4 GHz) over the "
                    "measured time per flux_div"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 timed launches (about 55 ms) after 10 untimed ones: a short run
    # still carries the clock ramp of the first launches (20 after 3 measured
    # about 4% slower than the steady state)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle", type=int, default=200,
                    help="untimed flux_div launches before the warmup (the kernel reaches its steady state after the rows)")
    ap.add_argument("--blocks", type=int, default=16, help="blocks per axis (16 -> 256^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the timings of the exchange, RK step, reductions, wall fill, checkpoint and "
                         "the stretched-channel flux_div")
    ap.add_argument("--no-chosen-form", action="store_true",
                    help="skip the batched / served FP32-flux measurements")
    ap.add_argument("--fp64-steps", type=int, default=30)
    ap.add_argument("--fp64-sample", type=int, default=12, help="ensemble cases timed on the FP64 path")
    ap.add_argument("--ensemble-cases", type=int, default=300)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=8.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-rate-saturated", type=float, default=64.0,
                    help="arrival rate of the second served run (above the engine's capacity)")
    ap.add_argument("--copy-check", type=int, default=0,
                    help="after the timed loop, copy the primitive array N times (a plain device copy of "
                         "known size, used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--precision-study", action="store_true",
                    help="instead of the benchmark: FP32-flux vs FP64 TGV run (errors every 100 steps)")
    ap.add_argument("--study-cells", type=int, default=256)
    ap.add_argument("--study-steps", type=int, default=1000)
    ap.add_argument("--channel", action="store_true",
                    help="instead of the benchmark: time flux_div on the 192x128x96 stretched channel only")
    ap.add_argument("--compositions", action="store_true",
                    help="instead of the benchmark: other flux compositions on the 128^3 TGV, specialised "
                         "vs generic per-cell kernel")
    args = ap.parse_args()
    if args.channel:
        import spade_amd as sa
        sa.load()
        print(json.dumps(channel_measure(sa)), flush=True)
        return
    if args.compositions:
        import spade_amd as sa
        sa.load()
        print(json.dumps(compositions_measure(sa)), flush=True)
        return
    if args.precision_study:
        import spade_amd as sa
        from spade_amd import studies
        sa.load()
        res = studies.precision(args.study_cells, args.study_steps, 100,
                                log=lambda r: print(json.dumps(r), file=sys.stderr, flush=True))
        print(json.dumps(res), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by blocks "
              "across GPUs but no multi-GPU run is part of this build", file=sys.stderr)
        sys.exit(2)

    import spade_amd as sa
    sa.load()
    nb, nc, ng = (args.blocks,) * 3, (16, 16, 16), 2
    L = 2.0 * np.pi
    gamma, R = 1.4, 100.0 / 1.4
    p0 = R
    t0 = time.perf_counter()
    g = sa.CartesianGrid(nc, nb, ((0.0, L),) * 3, num_exch=ng)
    host = np.zeros(g.host_shape)
    host[:, ng:-ng, ng:-ng, ng:-ng] = tgv_interior(nb, nc, L, p0)
    q = sa.GridArray(g).upload(host)
    exch = sa.create_exchange(q)
    exch.exchange(q)
    rhs = sa.GridArray(g, fill=0.0)
    gas = sa.IdealGas(gamma, R)
    flux = sa.compose(sa.cent_keep(2, gas), sa.visc_lr(sa.power_law(1.0 / 1600.0, 1.0, 0.76, 0.72), gas))
    g.ctx.sync()
    print(f"setup {time.perf_counter() - t0:.1f}s, {g.num_cells} cells", file=sys.stderr, flush=True)

    # the path's other rows first (they change q: RK steps, wall ghosts), then
    # the headline on the analytic state again; the GPU has then left its idle
    # clocks, so the timed launches measure the kernel's steady state
    rows = None
    if not args.no_rows:
        rows = rows_measure(sa, g, q, exch, gas, flux)
        rows["channel_flux_div"] = channel_measure(sa)
        q.upload(host)
        exch.exchange(q)
        g.ctx.sync()

    for _ in range(args.settle + args.warmup):
        sa.flux_div(q, rhs, flux, increment=False)
    g.ctx.sync()
    g.ctx.timer_start()
    for _ in range(args.steps):
        sa.flux_div(q, rhs, flux, increment=False)
    ms_total = g.ctx.timer_stop()
    g.ctx.sync()
    ms = ms_total / args.steps
    cells = g.num_cells
    value = cells / (ms * 1e-3) * 1e-6
    achieved = BYTES_PER_CELL_RHS * cells / (ms * 1e-3) / 1e9
    workload = f"tgv{nb[0] * nc[0]}_fp64_flux_div"
    traffic = _traffic(workload)
    line = {
        "metric": "Mcell-RHS/s (FP64 flux_div, cent_keep<2> + visc_lr power law)",
        "value": round(value, 2),
        "unit": "Mcell-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle": args.settle,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Taylor-Green initial state)",
        "config": {"workload": f"TGV {nb[0] * nc[0]}^3, {nb[0]}^3 blocks of 16^3, ng=2, one flux_div per step",
                   "cells": cells},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "valu_issue": _valu_issue(workload, ms),
    }
    # drop-in use from host memory: the reference-layout host arrays cross the
    # boundary (layout conversion + PCIe) around one exchange + flux_div
    t0 = time.perf_counter()
    q.upload(host)
    exch.exchange(q)
    sa.flux_div(q, rhs, flux, increment=False)
    out = rhs.download()
    pcie_s = time.perf_counter() - t0
    del host, out
    line["host_arrays_inclusive"] = {"ms": round(pcie_s * 1e3, 1), "mcell_rhs_per_s": round(cells / pcie_s * 1e-6, 1),
                                     "note": "upload of the primitive array + exchange + flux_div + download of the residual, "
                                             "reference (AoS) layout on the host"}
    if args.copy_check:
        for _ in range(args.copy_check):
            rhs.copy_from(q)
        g.ctx.sync()
        # the runtime copies the whole device allocation
        line["copy_check_bytes"] = 8 * g.device_size
    if rows is not None:
        line["rows"] = rows
    del q, rhs
    if not args.no_chosen_form:
        line["chosen_form"] = chosen_form(args, log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


def chosen_form(args, log):
    """Batched and served FP32-flux form against the plain FP64 path (cell-updates/s)."""
    from spade_amd import studies
    out = {}
    # plain FP64 single-case path: 256^3 TGV, ssprk3 steps (3 exchanges, 3 flux_div, 3 stage updates)
    from spade_amd import CaseSpec
    big = CaseSpec("tgv", (256,) * 3, mach=0.1, reynolds=1600.0)
    cu, sec = studies.fp64_rate(big, args.fp64_steps)
    out["fp64_single_case"] = {"workload": "TGV 256^3, one case, FP64 flux_div + transform-aware SSP-RK3",
                               "cell_updates_per_s": cu / sec, "ms_per_step": sec / args.fp64_steps * 1e3}
    log(f"fp64 single case {cu / sec / 1e9:.3f} Gcell-updates/s")
    # ensemble: 300 cases, all resident, persistent stage kernels + hipGraph + tuned launch shape
    specs = studies.ensemble_specs(args.ensemble_cases, args.ensemble_steps)
    ens = studies.run_ensemble(specs, args.ensemble_steps)
    log(f"ensemble {ens['cell_updates_per_s'] / 1e9:.3f} Gcell-updates/s")
    # the FP64 path on a sample of the same cases, one at a time
    tot_cu, tot_s = 0, 0.0
    for sp in specs[:args.fp64_sample]:
        cu, sec = studies.fp64_rate(sp, args.fp64_steps)
        tot_cu += cu
        tot_s += sec
    fp64_ens = tot_cu / tot_s
    ens.update({"fp64_same_cases_cell_updates_per_s": fp64_ens, "fp64_sample_cases": args.fp64_sample,
                "speedup_vs_fp64_same_cases": ens["cell_updates_per_s"] / fp64_ens,
                "speedup_vs_fp64_256": ens["cell_updates_per_s"] / out["fp64_single_case"]["cell_updates_per_s"]})
    out["ensemble"] = ens
    log(f"fp64 on ensemble cases {fp64_ens / 1e9:.3f} Gcell-updates/s")
    # serving: Poisson arrivals, continuous batching (--serve-requests 0 leaves it out)
    if args.serve_requests <= 0:
        return out
    fp64_256 = out["fp64_single_case"]["cell_updates_per_s"]
    for key, rate in (("served", args.serve_rate), ("served_saturated", args.serve_rate_saturated)):
        # light load (latency) and an arrival rate above the engine's capacity (throughput)
        srv, _ = studies.serving(args.serve_requests, rate)
        srv["speedup_vs_fp64_same_cases"] = srv["cell_updates_per_s"] / fp64_ens
        srv["speedup_vs_fp64_256"] = srv["cell_updates_per_s"] / fp64_256
        out[key] = srv
        log(f"{key} {srv['cell_updates_per_s'] / 1e9:.3f} Gcell-updates/s at {rate:g}/s, "
            f"p50 {srv['p50_latency_s']:.3f}s p99 {srv['p99_latency_s']:.3f}s")
    return out


if __name__ == "__main__":
    main()
