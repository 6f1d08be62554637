#!/usr/bin/env python3
"""Throughput of the RPA response chain chi(q, omega) on one MI355X.

One step = one pass of the hot path for one q over all frequencies of the
workload: transition bookkeeping, vertices, S0 accumulation, Kramers-Kronig,
Coulomb, Dyson/W and the loss epilogue (Sloss semantics), inputs resident in
HBM, run as one replay of the captured chain (every kernel re-runs; the
replay's result is checked bitwise against an eager run).
value = (no-1) output frequencies per step / seconds per step.  The q-loop
workloads run eagerly (each q changes shapes).

Default workload (production size): 64x64x1 k-grid (4096 k), 20 bands
(8 valence / 12 conduction), ~1500 plane waves per k, 51 local fields,
701 frequencies over 0-30 eV, eta = 50 meV, q = one k-grid step.  value is
the mixed precision chain (complex FP32 vertices, on this workload as BF16
hi/lo split products into FP32; BF16 hi/lo split products into FP32 for S0;
FP64 weights, chunk reduction, KK and Dyson/W); the FP64
chain is timed on the same inputs next to it, and both are compared with the
FP64 oracle on a k-subset.

Prints one JSON line (the last line of stdout).
"""

from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def host_threads():
    """OpenMP threads for the CPU baseline: every core this process may run
    on (its affinity set), or OMP_NUM_THREADS where the environment sets it
    (a shared host states the cores granted to the process there).  Threads
    are bound close."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else n


# OpenMP reads these when the oracle's runtime starts (its first parallel region)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "ω-pts/s for χ(q,ω), mixed vs FP64 GPU at ε⁻¹ rel err ≤1e-6; % HBM/MFMA peak"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16X3_PEAK_TFLOPS = round(2500.0 / 3.0, 1)  # dense BF16 peak over three split products
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sys_, r, threads, target_s=12.0):
    """The oracle (FP64 restatement of the reference loops) on a k-subset,
    scaled to the full k-grid; the tail (KK and Dyson/W, or KK and the Pi
    assembly) at full size."""
    from oracle import oracle as orc
    cfg, rc = sys_.cfg, r.rc
    H = 27.2113846
    eta = cfg.eta_ev / H
    optabs = rc.program == "optabs"

    def problem(nk):
        return orc.Problem(sys_, r.K1[:nk], r.K2[:nk], r.iG0[:nk], r.ktot[:nk], r.q, r.lf_idx,
                           map_mode=0, ngd=r.ngd, E_hartree=r.E)

    def s0(pb):
        if optabs:
            pol = importlib.import_module("2dmaterialsopticalprops_amd._abi").POL[rc.pol]
            return orc.optabs_s0(pb, cfg.nval, cfg.no, pol, r.domega, eta, rc.lor_cut, rc.df_cut,
                                 rc.efermi_ev / H, rc.temp_ev / H, sys_.vcell, sys_.gcar, nthreads=threads)
        return orc.sloss_s0(pb, cfg.nval, cfg.no, r.domega, eta, sys_.vcell, nthreads=threads), None

    nk_full = len(r.K1)
    nk = min(nk_full, 16 if optabs else 64)
    t0 = time.perf_counter()
    s0(problem(nk))
    t_probe = time.perf_counter() - t0
    per_k = t_probe / nk
    # three timed repetitions of the sample fit the target; the best counts
    nk_s = int(min(nk_full, max(nk, target_s / 3.0 / max(per_k, 1e-9))))
    nk_s = max(8, (nk_s // 8) * 8)
    pb = problem(nk_s)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        S0, Q = s0(pb)
        times.append(time.perf_counter() - t0)
    t_s0 = min(times)
    t0 = time.perf_counter()
    out = {"S0": S0}
    if optabs:
        out["chi0"] = orc.kk_optabs(S0, r.domega)
        out["pi11"], out["si"], out["sa"] = orc.pi_assembly(out["chi0"], Q, r.domega, rc.eta_intra_ev / H, cfg.c)
    else:
        out["chi0"] = orc.kk_sloss(S0, r.domega)
        V = orc.genV(rc.eps, r.q[0] * sys_.gcar, r.q[1] * sys_.gcar, cfg.c, r.parg, r.glf, sys_.gcar)
        outs = [orc.dyson_w(out["chi0"][f], V) for f in range(cfg.no - 1)]
        out["epsinv"] = np.array([o[0] for o in outs])
        out["W"] = np.array([o[2] for o in outs])
    t_tail = time.perf_counter() - t0
    t_full = t_s0 * nk_full / nk_s + t_tail
    out.update({"t_full_s": t_full, "nk_sample": nk_s, "t_s0_sample": t_s0, "t_tail": t_tail,
                "t_s0_runs": times})
    return out


def hbm_calibration(mib):
    """A plain copy of known size (read n bytes, write n bytes) through the
    package's reference copy kernel (libhbm_copy.so, 16-byte loads and
    stores), three times, so HBM counters can be scaled against it."""
    import ctypes
    path = os.path.join(ROOT, "2dmaterialsopticalprops_amd", "libhbm_copy.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    rc = ctypes.CDLL(path).hbm_copy_calibration(0, mib)
    if rc != 0:
        raise RuntimeError(f"hbm_copy_calibration failed with HIP error {rc}")
    log(f"[bench] hbm calibration: 3 x k_hbm_copy of {mib} MiB")


def time_chain(r, steps, warmup):
    """Warmup, then `steps` runs of the chain as rpa_run (the timed steps:
    the library's entry point, and what the drivers call) and `steps` graph
    replays of the captured chain (rpa_run_graph), each in two halves that
    alternate so that clock drift favours neither; then `steps` runs with
    stage events for the per-stage breakdown.  The graph replays must
    reproduce rpa_run's result bit for bit."""
    host = importlib.import_module("2dmaterialsopticalprops_amd.host")
    ctx = r.ctx
    out_name = "LOSS" if r.rc.program == "sloss" else "PI11"
    for _ in range(max(warmup, 1)):
        ctx.run()
    eager_out = ctx.result(out_name)
    for _ in range(max(warmup, 1)):
        ctx.run(graph=True)
    graph_same = bool(np.array_equal(ctx.result(out_name), eager_out))
    halves = [steps // 2, steps - steps // 2]
    wall = gwall = 0.0
    for n in halves:
        t0 = time.perf_counter()
        for _ in range(n):
            ctx.run()
        wall += time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(n):
            ctx.run(graph=True)
        gwall += time.perf_counter() - t0
    wall, gwall = wall * 1e3 / steps, gwall * 1e3 / steps
    ctx.set_profiling(True)
    stages = {k: 0.0 for k in host._abi.STAGES}
    dev_ms = 0.0
    for _ in range(steps):
        dev_ms += ctx.run()
        for k, v in ctx.stage_ms().items():
            stages[k] += v
    ctx.set_profiling(False)
    return {"ms_step": wall, "graph_ms": gwall, "device_ms": dev_ms / steps, "graph_same": graph_same,
            "stages": {k: v / steps for k, v in stages.items()}}


KERNEL_NAME = {  # (precision, stage) -> substrings of the kernel name in the rocprof tables
    ("mixed", "vertices"): ("k_vertex_col_dma<",), ("fp64", "vertices"): ("k_vertex_col<", "ColF64"),
    ("mixed", "accumulate"): ("k_accum_hp<",), ("fp64", "accumulate"): ("k_accum<",),
}


def mixed_s0_form(r):
    """The mixed S0 form the library takes: home panels when every
    Lorentzian window fits 8 row blocks of 16 from its home tile (the
    window's half width hw: |Lor| >= cut <=> dE^2 <= eta/cut - eta^2), else
    the tile form."""
    eta = r.cfg.eta_ev / importlib.import_module("2dmaterialsopticalprops_amd.host").HARTREE
    cut = (r.rc.lor_cut if r.rc.program == "optabs" else 1.0e-3) / eta
    hw2 = eta / cut - eta * eta
    rows = 31 + math.floor(2.0 * math.sqrt(max(hw2, 0.0)) / r.domega) + 6
    return "k_accum_hp" if rows <= 128 else "k_accum_mx"


def hbm_traffic(workload, prec, stage):
    """HBM bytes per launch of a kernel from profiles/hbm_traffic.txt (two
    rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy),
    if that table was measured on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    rows, wl = {}, None
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] == "workload":
            wl = " ".join(f[1:])
        else:
            rows[f[0]] = float(f[1]) + float(f[2])
    if wl != workload:
        return None
    parts = KERNEL_NAME[(prec, stage)]
    hits = [v for k, v in rows.items() if all(p in k for p in parts)]
    return hits[0] if len(hits) == 1 else None


def vertex_kernel(r, prec, lf_idx=None):
    """Name of the vertex kernel the library picks for the runner's local
    fields: the column form when the in-plane classes hold at least 12 fields
    on average (current vertices: one class only; mixed charge vertices on
    the BF16 hi/lo split products, k+q rows staged by LDS DMA), else the gather-direct FP32 form (Sloss)
    or the staged form."""
    m = r.sys.mill_g[np.asarray(r.lf_idx if lf_idx is None else lf_idx) - 1]
    ncls = len({(int(x[0]), int(x[1])) for x in m})
    optabs = r.rc.program == "optabs"
    if r.rc.vertex_map == "direct" and len(m) >= 12 * ncls and not (optabs and ncls > 1):
        if prec == "fp64":
            return "k_vertex_col<ColF64>"
        return "k_vertex_col<ColF32>" if optabs else "k_vertex_col_dma"
    if prec == "mixed" and not optabs:
        nc = r.cfg.nband - r.cfg.nval  # Sloss band block: 8 valence rows x up to 16 conduction columns
        return "k_vertex_gd<PrecF32> (two local fields per wave)" if nc <= 8 else "k_vertex_gd<PrecF32>"
    return "k_vertex<PrecF32> (staged)" if prec == "mixed" else "k_vertex<PrecF64> (staged)"


def vertex_kernels(r, prec, plans):
    """The vertex kernels of a q loop, with the number of q points each serves."""
    names = {}
    for pt in plans:
        n = vertex_kernel(r, prec, pt["lf_idx"])
        names[n] = names.get(n, 0) + 1
    return " + ".join(f"{n} on {c} q" for n, c in names.items())


def vertex_operand_bytes(r, prec):
    """Bytes the vertex kernel has to read at least once per k (direct map):
    C1 of the NGd plane waves in the bands n, and, in the bands m, the k+q
    rows that a term uses: the distinct targets G(i) + Glf(G) + G0 whose
    global index lies within ngk(K2) (the reference adds a term only then,
    Sloss.f90:928-948), summed over the FBZ k."""
    mg = r.sys.mill_g
    index = {tuple(x): i for i, x in enumerate(mg.tolist())}
    m = mg[:r.ngd]
    lf = mg[np.asarray(r.lf_idx) - 1]
    base = (m[:, None, :] + lf[None, :, :]).reshape(-1, 3)
    ngk2 = r.sys.ngk[np.asarray(r.K2) - 1]
    rows = 0
    for g0 in np.unique(r.iG0):
        t = np.unique(base + mg[g0 - 1][None, :], axis=0)
        idx = np.sort(np.array([index.get(tuple(x), 1 << 30) for x in t.tolist()]))
        sel = np.asarray(r.iG0) == g0
        rows += int(np.searchsorted(idx, ngk2[sel], side="left").sum())
    nv, nb = r.cfg.nval, r.cfg.nband
    nn, nm = (nb, nb) if r.rc.program == "optabs" else (nv, nb - nv)
    es = 8 if prec == "mixed" else 16
    return float(len(r.K1) * r.ngd * nn + rows * nm) * es


def roofline_of(res, prec, flops_acc, flops_vtx, workload="", vname=None, vbytes=None, sform="k_accum_hp"):
    """Roofline of the dominant matrix-core kernel of one run."""
    acc_ms, vtx_ms = res["stages"]["accumulate"], res["stages"]["vertices"]
    if prec == "mixed":
        vname = vname or "k_vertex_col_dma"
        split = "col_dma" in vname or "col_split" in vname
        names = (f"{sform} (S0, BF16 hi/lo x3 into FP32, with its weight prepass)",
                 f"{vname} (vertices, {'BF16 hi/lo x3 into FP32' if split else 'FP32'})")
        peaks = (BF16X3_PEAK_TFLOPS, BF16X3_PEAK_TFLOPS if split else FP32_PEAK_TFLOPS)
    else:
        names = ("k_accum (S0, FP64)", f"{vname or 'k_vertex_col<ColF64>'} (vertices, FP64)")
        peaks = (FP64_PEAK_TFLOPS, FP64_PEAK_TFLOPS)
    if acc_ms >= vtx_ms:
        dom, fl, ms, peak, stage = names[0], flops_acc, acc_ms, peaks[0], "accumulate"
    else:
        dom, fl, ms, peak, stage = names[1], flops_vtx, vtx_ms, peaks[1], "vertices"
    achieved = fl / (ms * 1e-3) / 1e12
    traffic = hbm_traffic(workload, prec, stage)
    return {
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
        "frac": round(achieved / peak, 4), "traffic": traffic,
        "traffic_note": ("HBM bytes per launch, rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy "
                         "(profiles/hbm_traffic.txt)") if traffic else "not measured for this workload",
        "kernel": dom, "kernel_ms": round(ms, 4),
        "algorithmic_flop": fl,
        **({"algorithmic_bytes": vbytes,
            "traffic_over_algorithmic_bytes": round(traffic / vbytes, 3) if traffic else None}
           if stage == "vertices" and vbytes else {}),
        "other": {
            names[1]: {"ms": round(vtx_ms, 4), "flop": flops_vtx, "peak": peaks[1],
                       "tflops": round(flops_vtx / (vtx_ms * 1e-3) / 1e12, 3) if vtx_ms > 0 else None},
            names[0]: {"ms": round(acc_ms, 4), "flop": flops_acc, "peak": peaks[0],
                       "tflops": round(flops_acc / (acc_ms * 1e-3) / 1e12, 3) if acc_ms > 0 else None},
        },
    }


def work_model(r, host, cfg, K1, K2, nlf, ntr):
    """Algorithmic work of the two matrix-core kernels for one q: nonzero
    (transition, bin) weights, S0 flops (2 N_G^2 per real Sloss weight, 8 N_G^2
    per complex Optabs weight) and vertex flops (8 N_G NGd per transition and
    vertex: rho for Sloss, j for Optabs, j of both polarisations for yz --
    one complex product per term, as the reference's loop)."""
    nv, nb = cfg.nval, cfg.nband
    eta = cfg.eta_ev / host.HARTREE
    if r.rc.program == "optabs":
        nnz = host.nonzero_weights_optabs(r.E, K1, K2, nv, nb, cfg.no, r.domega, eta, r.rc.lor_cut)
        nvtx = 2 if r.rc.pol == "yz" else 1
        return nnz, 8.0 * nlf * nlf * nnz, 8.0 * nlf * r.ngd * ntr * nvtx
    nnz = host.nonzero_weights(r.E, K1, K2, nv, nb, cfg.no, r.domega, eta)
    return nnz, 2.0 * nlf * nlf * nnz, 8.0 * nlf * r.ngd * ntr


def time_q_loop(r, plans, steps, warmup, eager=False):
    """The q loop as one batch (rpa_batch_set once, then one replay of the
    captured graph of every q's chain per pass; vertex forms from the
    autotuner): warmup passes, then `steps` timed passes; with eager=True
    instead per q the context takes the q point and runs the chain.  Then
    `steps` eager passes with stage events for the breakdown."""
    from importlib import import_module
    host = import_module("2dmaterialsopticalprops_amd.host")
    ctx = r.ctx
    if eager:
        for _ in range(max(warmup, 1)):
            for pt in plans:
                r.set_q_point(pt)
                ctx.run()
        t0 = time.perf_counter()
        for _ in range(steps):
            for pt in plans:
                r.set_q_point(pt)
                ctx.run()
        wall = time.perf_counter() - t0
    else:
        tb = time.perf_counter()
        ctx.batch_set(plans, r.ktot)
        ctx.batch_run()  # prepares every q (autotuner), captures and replays once
        log(f"[bench] batch of {len(plans)} q set up and captured in {time.perf_counter() - tb:.1f}s")
        for _ in range(max(warmup, 1) - 1):
            ctx.batch_run()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.batch_run()
        wall = time.perf_counter() - t0
        ctx.set_autotune(True)  # the stage breakdown below on the forms the batch chose
    ctx.set_profiling(True)
    stages = {k: 0.0 for k in host._abi.STAGES}
    dev_ms = 0.0
    for _ in range(steps):
        for pt in plans:
            r.set_q_point(pt)
            dev_ms += ctx.run()
            for k, v in ctx.stage_ms().items():
                stages[k] += v
    ctx.set_profiling(False)
    return {"ms_step": wall * 1e3 / steps, "device_ms": dev_ms / steps,
            "stages": {k: v / steps for k, v in stages.items()}}


def bench_batched(args, s, cfg, synth, host):
    """Throughput over a q loop with per-q local-field counts: (q, omega)
    W-solves/s for the Sloss chain, omega-points/s for the Optabs route."""
    spec = synth.batched_q_points(args.config, args.nq)
    program = synth.BATCHED_Q[args.config]["program"]
    optabs = program == "optabs"
    unit = "omega-points/s" if optabs else "(q,omega) W-solves/s"
    precs = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    runs, runners, plans = {}, {}, {}
    for prec in precs:
        t0 = time.perf_counter()
        r = host.Runner(s, host.RunConfig(program=program, vertex_map=args.vertex_map, precision=prec,
                                          iq=spec[0][0], pol=args.pol))
        t1 = time.perf_counter()
        plans[prec] = [r.q_point(iq, nlf) for iq, nlf in spec]
        log(f"[bench] {prec}: setup+upload {t1 - t0:.1f}s, {len(spec)} q points planned "
            f"({time.perf_counter() - t1:.1f}s), N_G {spec[0][1]}..{spec[-1][1]}")
        res = time_q_loop(r, plans[prec], args.steps, args.warmup, eager=args.eager_q_loop)
        log(f"[bench] {prec}: {res['ms_step']:.3f} ms per q loop (device {res['device_ms']:.3f} ms); "
            f"stages {json.dumps({k: round(v, 4) for k, v in res['stages'].items()})}")
        runs[prec], runners[prec] = res, r
    head = precs[0]
    r = runners[head]
    nsolve = len(spec) * (cfg.no - 1)
    value = nsolve / (runs[head]["ms_step"] * 1e-3)

    # algorithmic work summed over the q points
    nv, nb = cfg.nval, cfg.nband
    ntr = r.ntransitions()
    flops_acc = flops_vtx = 0.0
    nnz_total = 0
    for pt, (iq, nlf) in zip(plans[head], spec):
        nnz, fa, fv = work_model(r, host, cfg, pt["K1"], pt["K2"], nlf, ntr)
        nnz_total += nnz
        flops_acc += fa
        flops_vtx += fv
    workload_key = f"{args.config} batched {program} {args.vertex_map}"
    roofline = roofline_of(runs[head], head, flops_acc, flops_vtx, workload_key,
                           vertex_kernels(r, head, plans[head]), sform=mixed_s0_form(r))
    roofline["other"]["nnz_weights"] = nnz_total
    dyson_ms = runs[head]["stages"]["dyson_or_pi"]
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": unit, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(runs[head]["ms_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("fp32 vertices + bf16 hi/lo x3 S0 into fp32, fp64 weights/reduction/KK/Dyson"
                  if head == "mixed" else "f64"),
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: Optabs spectral function + KK, pol {args.pol}, q loop "
                         f"iq={spec[0][0]}..{spec[-1][0]} ({len(spec)} q), {len(s.ngk)} k ({cfg.nk}x{cfg.nk}x1), "
                         f"{nb} bands ({nv}v/{nb - nv}c), ngk {s.ngk.min()}-{s.ngk.max()}, "
                         f"N_G {spec[0][1]}-{spec[-1][1]} by |q+G|, {2 * cfg.no + 1} bins "
                         f"+-{cfg.omax_ev:g} eV -> KK to {cfg.no - 1} omega, eta={cfg.eta_ev * 1e3:g} meV")
            if optabs else
                        (f"{args.config}: batched W(q, omega), Sloss q loop iq={spec[0][0]}..{spec[-1][0]} "
                         f"({len(spec)} q), {len(s.ngk)} k ({cfg.nk}x{cfg.nk}x1), {nb} bands "
                         f"({nv}v/{nb - nv}c), ngk {s.ngk.min()}-{s.ngk.max()}, N_G {spec[0][1]}-{spec[-1][1]} "
                         f"by |q+G|, {cfg.no} omega 0-{cfg.omax_ev:g} eV, eta={cfg.eta_ev * 1e3:g} meV"),
            "precision": head, "q_points": len(spec),
            "timing": ("eager q loop, q by q (rpa_set_local_fields, rpa_set_q, rpa_run)" if args.eager_q_loop else
                       "one replay per step of the captured graph of all q (rpa_batch_run; vertex forms by the "
                       "autotuner)"),
            ("omega_points_per_step" if optabs else "w_solves_per_step"): nsolve, "transitions_per_q": ntr,
            "device_ms_per_step": round(runs[head]["device_ms"], 4),
            "stage_ms": {k: round(v, 4) for k, v in runs[head]["stages"].items()},
        },
        "roofline": roofline,
    }
    if not optabs and dyson_ms > 0:
        out["config"]["dyson_w_solves_per_s"] = round(nsolve / (dyson_ms * 1e-3), 1)
    if "fp64" in runs and head != "fp64":
        f = runs["fp64"]
        fv = nsolve / (f["ms_step"] * 1e-3)
        out["fp64_gpu"] = {
            "value": round(fv, 1), "unit": unit, "ms_per_step": round(f["ms_step"], 4),
            "stage_ms": {k: round(v, 4) for k, v in f["stages"].items()},
            "roofline": roofline_of(f, "fp64", flops_acc, flops_vtx, workload_key,
                                    vertex_kernels(runners["fp64"], "fp64", plans["fp64"])),
        }
        out["mixed_vs_fp64"] = round(value / fv, 2)
    if not args.no_cpu:
        threads = host_threads()
        try:
            # 5 q points spanning the N_G range, each timed on a k-subset,
            # then scaled by sum N_G^2 over all q
            sub = sorted({int(round(j * (len(spec) - 1) / 4)) for j in range(5)})
            t_sub, t_meas, acc = 0.0, 0.0, {p: [] for p in precs}
            notes = []
            for j in sub:
                for p in precs:
                    runners[p].set_q_point(plans[p][j])
                cb = cpu_baseline(s, runners[head], threads, args.cpu_seconds / len(sub))
                t_sub += cb["t_full_s"]
                t_meas += cb["t_s0_sample"] + cb["t_tail"]
                notes.append(f"iq={spec[j][0]} N_G={spec[j][1]}: {cb['nk_sample']} k")
                for p in precs:
                    acc[p].append(dict(accuracy(s, runners[p], cb), iq=spec[j][0], N_G=spec[j][1]))
            w_all = sum(n * n for _, n in spec)
            w_sub = sum(spec[j][1] ** 2 for j in sub)
            t_full = t_sub * w_all / w_sub
            out["cpu_baseline"] = {
                "value": round(nsolve / t_full, 2), "unit": unit, "cores": threads,
                "kind": "port",
                "sample": f"oracle chain on {len(sub)} of {len(spec)} q ({'; '.join(notes)}), S0 on a k-subset "
                          f"scaled to {len(r.K1)} k, total scaled by sum N_G^2 over all q "
                          f"(x{w_all / w_sub:.2f}); {t_meas:.1f}s measured, {t_full:.1f}s estimated for the loop",
            }
            out["gpu_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
            out["accuracy_vs_fp64_oracle"] = acc
        except Exception as e:  # the baseline is reported, never silently replaced
            out["cpu_baseline"] = {"error": repr(e)}
    for rr in runners.values():
        rr.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--program", default="sloss", choices=["sloss", "optabs"])
    ap.add_argument("--vertex-map", default="direct", choices=["direct", "igk"])
    ap.add_argument("--pol", default="xx", choices=["xx", "yy", "zz", "yz"],
                    help="Optabs: polarisation of the current vertices")
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"],
                    help="'both': value is the mixed path, timed next to the FP64 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and accuracy check")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--nq", type=int, default=None,
                    help="batched-W workloads: use this many of the q points (spread over the N_G range)")
    ap.add_argument("--eager-q-loop", action="store_true",
                    help="batched workloads: run the q loop eagerly, q by q, instead of one graph of all q")
    ap.add_argument("--hbm-calibration-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (x1.0 elementwise kernel), "
                         "the reference point for FETCH_SIZE / WRITE_SIZE counter scaling")
    ap.add_argument("--interpolate", action="store_true",
                    help="time interpolate.x (rpa_interpolate) at No=2001, Nlf=65, No_interp=10 instead")
    ap.add_argument("--photon", action="store_true",
                    help="time Photon.x (rpa_photon) at No=2001, Nlf=33, 10 q lines instead")
    ap.add_argument("--gamma", action="store_true",
                    help="time the GW propagators (rpa_gamma_propagator): 701 omega, 26010 series")
    ap.add_argument("--ladder", action="store_true",
                    help="time the ladder frequency loop (rpa_ladder): 512 k, 401 omega, and the Fock kernel "
                         "(rpa_fock_kernel) of the same 512 k")
    args = ap.parse_args()
    if args.interpolate:
        return bench_interpolate(args)
    if args.gamma:
        return bench_gamma(args)
    if args.ladder:
        return bench_ladder(args)
    if args.photon:
        return bench_photon(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(q points would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.hbm_calibration_mib > 0:
        hbm_calibration(args.hbm_calibration_mib)
    synth = importlib.import_module("2dmaterialsopticalprops_amd.synth")
    host = importlib.import_module("2dmaterialsopticalprops_amd.host")
    cfg = synth.CONFIGS[args.config]
    t0 = time.perf_counter()
    s = synth.make_system(cfg)
    log(f"[bench] synthetic {args.config}: nk={len(s.ngk)} nband={cfg.nband} ngk={s.ngk.min()}-{s.ngk.max()} "
        f"({time.perf_counter() - t0:.1f}s)")
    if args.config in synth.BATCHED_Q:
        return bench_batched(args, s, cfg, synth, host)
    precs = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    runs, runners = {}, {}
    for prec in precs:
        t0 = time.perf_counter()
        r = host.Runner(s, host.RunConfig(program=args.program, vertex_map=args.vertex_map, precision=prec,
                                          pol=args.pol))
        log(f"[bench] {prec}: setup+upload {time.perf_counter() - t0:.1f}s nlf={r.nlf} no={cfg.no} "
            f"ntrans={r.ntransitions()}")
        res = time_chain(r, args.steps, args.warmup)
        log(f"[bench] {prec}: {res['ms_step']:.3f} ms/step (graph replays {res['graph_ms']:.3f} ms, "
            f"device {res['device_ms']:.3f} ms with stage events, graph = eager bitwise: {res['graph_same']}); "
            f"stages {json.dumps({k: round(v, 4) for k, v in res['stages'].items()})}")
        runs[prec], runners[prec] = res, r
    head = precs[0]
    r = runners[head]
    nfreq = cfg.no - 1
    value = nfreq / (runs[head]["ms_step"] * 1e-3)

    # algorithmic work of the two matrix-core kernels (nonzero weights only)
    nv, nb = cfg.nval, cfg.nband
    ntr = r.ntransitions()
    nnz, flops_acc, flops_vtx = work_model(r, host, cfg, r.K1, r.K2, r.nlf, ntr)
    workload_key = f"{args.config} {args.program} {args.vertex_map}"
    vbytes = vertex_operand_bytes(r, head) if args.vertex_map == "direct" else None
    roofline = roofline_of(runs[head], head, flops_acc, flops_vtx, workload_key, vertex_kernel(r, head), vbytes,
                           sform=mixed_s0_form(r))
    roofline["other"]["nnz_weights"] = nnz

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "omega-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(runs[head]["ms_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("fp32 vertices + bf16 hi/lo x3 S0 into fp32, fp64 weights/reduction/KK/Dyson"
                  if head == "mixed" else "f64"),
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: {program_name(args.program)} chain, {len(s.ngk)} k "
                        f"({cfg.nk}x{cfg.nk}x1), {nb} bands ({nv}v/{nb - nv}c), ngk {s.ngk.min()}-{s.ngk.max()}, "
                        f"N_G={r.nlf}, {cfg.no} omega 0-{cfg.omax_ev:g} eV, eta={cfg.eta_ev * 1e3:g} meV, "
                        f"q=k-grid step, vertex map {args.vertex_map}",
            "precision": head, "omega_points_per_step": nfreq, "transitions": ntr,
            "timing": "rpa_run per step (the chain's kernels on one stream, one synchronisation)",
            "graph_replay_ms_per_step": round(runs[head]["graph_ms"], 4),
            "graph_matches_eager_bitwise": runs[head]["graph_same"],
            "device_ms_per_step": round(runs[head]["device_ms"], 4),
            "stage_ms": {k: round(v, 4) for k, v in runs[head]["stages"].items()},
        },
        "roofline": roofline,
    }
    if "fp64" in runs and head != "fp64":
        f = runs["fp64"]
        fv = nfreq / (f["ms_step"] * 1e-3)
        out["fp64_gpu"] = {
            "value": round(fv, 1), "unit": "omega-points/s", "ms_per_step": round(f["ms_step"], 4),
            "stage_ms": {k: round(v, 4) for k, v in f["stages"].items()},
            "roofline": roofline_of(f, "fp64", flops_acc, flops_vtx, workload_key,
                                    vertex_kernel(runners["fp64"], "fp64")),
        }
        out["mixed_vs_fp64"] = round(value / fv, 2)
    if not args.no_cpu:
        threads = host_threads()
        try:
            cb = cpu_baseline(s, r, threads, args.cpu_seconds)
            out["cpu_baseline"] = {
                "value": round(nfreq / cb["t_full_s"], 2), "unit": "omega-points/s", "cores": threads,
                "kind": "port",
                "sample": f"oracle S0 on {cb['nk_sample']} of {len(r.K1)} k-points "
                          f"({cb['t_s0_sample']:.1f}s, scaled x{len(r.K1) / cb['nk_sample']:.2f}) "
                          f"+ full KK/Dyson/W ({cb['t_tail']:.1f}s)",
            }
            out["gpu_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
            out["accuracy_vs_fp64_oracle"] = {p: accuracy(s, runners[p], cb) for p in precs}
        except Exception as e:  # the baseline is reported, never silently replaced
            out["cpu_baseline"] = {"error": repr(e)}
    for rr in runners.values():
        rr.close()
    print(json.dumps(out))


def bench_interpolate(args):
    """interpolate.x (Chi_RPA/interpolate.f90) at the size its own comments
    name (No = 2001 coarse frequencies, No_interp = 10, Nlf = 65, every
    (iG, jG) pair written): one step = one rpa_interpolate call, host S0 and
    Qeff in, the dense Pi_tot_interp(17991, 65, 65) back on the host (the
    call's copies and device buffers included; the kernels' own times come
    from the kernel trace)."""
    abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
    no, nlf, nint = 2001, 65, 10
    hartree = 2.0 * 13.6056923
    omax = 50.0 / hartree
    dw, dwt, eta = omax / (no - 1), omax / ((nint - 1) * (no - 2)), 0.025 / hartree
    rng = np.random.default_rng(1)
    shape = (2 * no + 1, nlf, nlf)
    s0 = np.asfortranarray(1e-2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)))
    qeff = np.asfortranarray(1e-2 * (rng.standard_normal((nlf, nlf)) + 1j * rng.standard_normal((nlf, nlf))))
    no_tot = (nint - 1) * (no - 2)
    for _ in range(args.warmup):
        abi.interpolate(s0, qeff, nint, dw, dwt, eta)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, _ = abi.interpolate(s0, qeff, nint, dw, dwt, eta)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    vals = no_tot * nlf * nlf
    res = {"metric": "interpolate.x dense Pi values per second (host in, host out)",
           "value": round(vals / (ms * 1e-3), 1), "unit": "(point, pair) values/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"interpolate.x: No={no}, No_interp={nint} ({no_tot} dense points), Nlf={nlf}, "
                                  f"all pairs", "output_bytes": vals * 16}}
    if not args.no_cpu:
        orc = importlib.import_module("oracle.oracle")
        k = 3  # local fields of the sample: k*k of the nlf*nlf independent pairs
        t0 = time.perf_counter()
        pi0 = orc.kk_optabs(np.asfortranarray(s0[:, :k, :k]), dw)
        orc.interpolate(pi0, np.asfortranarray(qeff[:k, :k]), nint, dwt, eta)
        t = (time.perf_counter() - t0) * (nlf * nlf) / (k * k)
        res["cpu_baseline"] = {"value": round(vals / t, 1), "unit": "(point, pair) values/s",
                               "cores": min(k * k, host_threads()), "kind": "port",
                               "sample": f"oracle on {k * k} of {nlf * nlf} pairs, scaled"}
        ref = orc.interpolate(pi0, np.asfortranarray(qeff[:k, :k]), nint, dwt, eta)[0]
        den = np.abs(ref).max()
        res["max_rel_err_vs_oracle_sample"] = float(np.abs(out[:, :k, :k] - ref).max() / den)
    print(json.dumps(res))


def bench_photon(args):
    """Photon.x (BSE/Photonew.f90) at the size of its own inputs: No = 2001
    frequencies of Pi0_xx/yy/zz (0.5-50 eV), Nlf = 33 local fields along z
    (the reference's 5 Ha graphene count), 10 q lines: 20000 (q, omega)
    points, each an N_G and a 2 N_G complex inversion with their products.
    One step = one rpa_photon call, host Pi0 in, the per-point outputs back."""
    abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
    no, nlf, nq, c0 = 2001, 33, 10, 32.5411
    hartree = 2.0 * 13.6056923
    omin = 0.5 / hartree
    dw = (50.0 / hartree) / (no - 1)
    eta = 0.05 / hartree
    ms = np.array([0] + [(k + 1) // 2 * (1 if k % 2 else -1) for k in range(1, nlf)])
    gz = 2 * np.pi * ms / c0
    par = np.where(ms % 2 == 0, 1, -1).astype(np.int32)
    rng = np.random.default_rng(2)
    comps = []
    for _ in range(3):
        a = (0.02 / nlf) * (rng.standard_normal((no, nlf, nlf)) + 1j * rng.standard_normal((no, nlf, nlf)))
        comps.append(np.asfortranarray(a + np.swapaxes(a, 1, 2)))
    lines = [(iq * 0.002, 0.0, 0) for iq in range(nq)]
    for _ in range(args.warmup):
        abi.photon(comps[0], comps[1], comps[2], gz, par, lines, omin, dw, eta, c0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = abi.photon(comps[0], comps[1], comps[2], gz, par, lines, omin, dw, eta, c0)
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    npts = (no - 1) * nq
    res = {"metric": "Photon.x (q, omega) points per second (host in, host out)",
           "value": round(npts / (ms_step * 1e-3), 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"Photon.x: No={no}, Nlf={nlf} (TS {nlf}x{nlf}, TP {2 * nlf}x{2 * nlf}), "
                                  f"{nq} q lines, {npts} points"}}
    if not args.no_cpu:
        orc = importlib.import_module("oracle.oracle")
        nsample = 200
        pts = [(l, f) for l in range(nq) for f in range(0, no - 1, (no - 1) * nq // nsample)][:nsample]
        t0 = time.perf_counter()
        worst = 0.0
        for l, f in pts:
            pi0 = np.zeros((nlf, nlf, 5), dtype=np.complex128, order="F")
            pi0[:, :, 0], pi0[:, :, 1], pi0[:, :, 4] = comps[0][f], comps[1][f], comps[2][f]
            _, sums, spec, _ = orc.photon_point(lines[l][0], omin + f * dw, eta, c0, 0.0, par, gz, pi0)
            worst = max(worst, abs(out["psum"][f, l, 1] - sums[1]) / max(abs(sums[1]), 1e-300))
        t = (time.perf_counter() - t0) * npts / len(pts)
        res["cpu_baseline"] = {"value": round(npts / t, 1), "unit": "points/s", "cores": 1, "kind": "port",
                               "sample": f"oracle on {len(pts)} of {npts} points, scaled"}
        res["max_rel_err_vs_oracle_sample_sPiyy"] = float(worst)
    print(json.dumps(res))


def bench_gamma(args):
    """GW propagators (Modules/propagators.f90 genGammaPropagator) for the
    screened interaction of ten q at the default workload's shape: 701
    frequencies, N_G = 51, so 10 x 51^2 = 26010 series per call."""
    abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
    no, nser, dw = 701, 26010, (30.0 / (2.0 * 13.6056923)) / 700
    rng = np.random.default_rng(4)
    D = np.asfortranarray(rng.standard_normal((no, nser)) + 1j * rng.standard_normal((no, nser)))
    for _ in range(args.warmup):
        abi.gamma_propagator(D, dw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gp, gm, kks, skk = abi.gamma_propagator(D, dw)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    res = {"metric": "GW propagator series per second (host in, host out)", "value": round(nser / (ms * 1e-3), 1),
           "unit": "series/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"genGammaPropagator: No={no}, {nser} series (10 q x 51^2 pairs)"}}
    if not args.no_cpu:
        orc = importlib.import_module("oracle.oracle")
        ns = 100
        t0 = time.perf_counter()
        worst = 0.0
        for s in range(ns):
            wp, _, _, _ = orc.gamma_propagator(D[:, s], dw)
            worst = max(worst, float(np.abs(gp[:, s] - wp).max() / np.abs(wp).max()))
        t = (time.perf_counter() - t0) * nser / ns
        res["cpu_baseline"] = {"value": round(nser / t, 1), "unit": "series/s", "cores": 1, "kind": "port",
                               "sample": f"oracle on {ns} of {nser} series, scaled"}
        res["max_rel_err_vs_oracle_sample"] = worst
    print(json.dumps(res))


def bench_ladder(args):
    """The ladder frequency loop of Chi_ladd (Chi_ladd_up/BSE.f90:1116-1176):
    a 512 x 512 Fock kernel, 401 frequencies (the reference's no), one
    inversion of I - L0 F / NV per frequency."""
    abi = importlib.import_module("2dmaterialsopticalprops_amd._abi")
    nk, no, ntot, vcell, valley = 512, 401, 2601, 922.0586, 2.0
    omin, dw, eta = 0.0, (4.0 / (2.0 * 13.6056923)) / 400, 0.05 / (2.0 * 13.6056923)
    nv = ntot * vcell
    rng = np.random.default_rng(5)
    F = np.asfortranarray(0.3 * nv / np.sqrt(nk) * 0.05 * (rng.standard_normal((nk, nk))
                                                           + 1j * rng.standard_normal((nk, nk))))
    delta = rng.uniform(0.02, 0.1, nk)
    cur = np.asfortranarray(rng.standard_normal((2, nk)) + 1j * rng.standard_normal((2, nk)))
    for _ in range(args.warmup):
        abi.ladder(F, delta, cur, no, omin, dw, eta, ntot, vcell, valley)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cl, cp = abi.ladder(F, delta, cur, no, omin, dw, eta, ntot, vcell, valley)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    res = {"metric": "ladder frequencies per second (host in, host out)", "value": round(no / (ms * 1e-3), 2),
           "unit": "omega/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Chi_ladd frequency loop: nkbz={nk}, no={no}"}}
    if not args.no_cpu:
        orc = importlib.import_module("oracle.oracle")
        nsmp = 2
        t0 = time.perf_counter()
        wl, _ = orc.ladder(F, delta, cur, nsmp, omin, dw, eta, ntot, vcell, valley)
        t = (time.perf_counter() - t0) * no / nsmp
        res["cpu_baseline"] = {"value": round(no / t, 3), "unit": "omega/s", "cores": 1, "kind": "port",
                               "sample": f"oracle on {nsmp} of {no} frequencies, scaled"}
        res["max_rel_err_vs_oracle_sample"] = float(np.abs(cl[:, :nsmp] - wl).max() / np.abs(wl).max())
    # the Fock kernel of the same k-points (Chi_ladd_up/BSE.f90:804-816): 512^2 pairs,
    # N_G = 33, 121 W tables
    nlf, nq, qmin = 33, 121, 0.0025
    kbz = np.asfortranarray(np.vstack([rng.uniform(0, 0.05, nk), rng.uniform(0, 0.05, nk), np.zeros(nk)]))
    wt = np.asfortranarray(rng.standard_normal((nq, nlf, nlf)) + 1j * rng.standard_normal((nq, nlf, nlf)))
    wt[1, 0, 0] = abs(wt[1, 0, 0]) + 100.0
    mnm = np.asfortranarray(rng.standard_normal((2, nlf, nk, nk)) + 1j * rng.standard_normal((2, nlf, nk, nk)))
    gcar = 2 * np.pi / 5.9715
    abi.fock_kernel(kbz, wt, mnm, ntot, gcar, qmin, 29.8575, vcell)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fk = abi.fock_kernel(kbz, wt, mnm, ntot, gcar, qmin, 29.8575, vcell)
    res["fock_kernel"] = {"ms_per_call": round((time.perf_counter() - t0) * 1e3 / args.steps, 3),
                          "workload": f"{nk}^2 pairs, N_G={nlf}, {nq} W tables (host in, host out)"}
    if not args.no_cpu:
        t0 = time.perf_counter()
        sub = np.asfortranarray(mnm[:, :, :64, :64])
        wf = orc.fock_kernel(kbz[:, :64], wt, sub, nq, ntot, gcar, qmin, 29.8575, vcell)
        res["fock_kernel"]["cpu_oracle_ms_scaled"] = round((time.perf_counter() - t0) * 1e3 * (nk / 64) ** 2, 1)
        res["fock_kernel"]["max_rel_err_vs_oracle_64x64"] = float(np.abs(fk[:64, :64] - wf).max() / np.abs(wf).max())
    print(json.dumps(res))


def program_name(p):
    return "Sloss (chi0 -> eps^-1 -> chi -> W -> loss)" if p == "sloss" else "Optabs (Pi0 -> Pi, sigma)"


def accuracy(s, r, cb):
    """The device chain on the oracle's k-subset against the oracle: S0, the
    KK output and the chain's end product (eps^-1, W, loss for Sloss; Pi_11
    and the conductivities for Optabs)."""
    nk = cb["nk_sample"]
    ctx = r.ctx
    ctx.set_q(r.K1[:nk], r.K2[:nk], r.iG0[:nk], r.ktot[:nk], r.q)
    ctx.vertices()
    S0, _ = ctx.accumulate_s0()
    chi0 = ctx.kk()

    def fro(a, b):
        return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

    res = {"sample": f"{nk} of {len(r.K1)} k-points, full bands/plane waves/N_G/omega",
           "S0_rel_fro": fro(S0, cb["S0"]), "chi0_rel_fro": fro(chi0, cb["chi0"])}
    if r.rc.program == "optabs":
        _, pi11, sig = ctx.epilogue()
        res.update({"pi11_rel_fro": fro(pi11, cb["pi11"]), "sigma_inter_rel_fro": fro(sig[0], cb["si"]),
                    "sigma_peak_shift_bins": int(np.argmax(sig[0]) - np.argmax(cb["si"]))})
        return res
    ctx.coulomb()
    epsinv, chi, W = ctx.dyson_w()
    loss_g = -epsinv[:, 0, 0].imag
    loss_o = -cb["epsinv"][:, 0, 0].imag
    res.update({"epsinv_rel_fro": fro(epsinv, cb["epsinv"]), "W_rel_fro": fro(W, cb["W"]),
                "loss_peak_shift_bins": int(np.argmax(loss_g) - np.argmax(loss_o))})
    return res


if __name__ == "__main__":
    main()
