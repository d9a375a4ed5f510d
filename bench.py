#!/usr/bin/env python3
"""Throughput of the field -> lognormal -> sources path on one MI355X.

One step = Gaussian modes + 3-D c2r of delta and phi (with sigma^2), the
lognormal transform fused with the FP64 per-shell normalisation of every
tracer, and Poisson sources + positions for every population, ending with a
complete device-resident catalogue.  Workload: 1024^3 FP32 grid, z in
[0.05, 1.6], three populations (b=0.95/D(z) exponential, b=1.5
exp-truncated, b=2.2 truncated), seed 3003; inputs are synthetic (EH98
no-wiggle P(k), Gaussian n(z)).

Prints one JSON line.  `value` is grid cells per second of the whole step.
`roofline` is for the kernel with the largest device time per launch over
both streams (phi's passes on the second stream included): algorithmic HBM
bytes per launch over its mean duration (HIP events on the stream it runs
on), against 8 TB/s.  `roofline_critical_stream` is the same for the largest
kernel on the library's own stream, the one that sets the step time.  `cpu_baseline` times the FP64 CPU oracle on the same
configuration at 768^3 on the host cores (about 20 s of CPU work).
"""
import argparse
import math
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid cells/sec (field→lognormal→sources) at 1024³ FP32; % HBM roofline"
HBM_PEAK = 8.0e12


def kernel_bytes(N, nsrc, npop):
    """Algorithmic HBM bytes per launch of each timed kernel.

    k_gen_z  writes both half spectra (delta and phi) after the z pass;
    k_col_y  reads and writes the delta half spectrum (y pass);
    k_c2r_x  reads the delta half spectrum and writes its padded real grid;
    (phi)    the same two passes for phi, on a second stream: its y pass
             overlaps delta's x pass and the stages after it, so its duration
             includes the time those kernels hold the chip;
    k_shells reads delta_G and writes the lognormal field in place (cells only);
    k_count  reads the density of every cell once and writes one count byte
             per cell and population;
    k_fill+k_write  read the count bytes, write and read one 16-byte
             descriptor per source, gather the 6-point phi stencil and write
             pos + ipix.
    """
    nh = N // 2 + 1
    half = N * N * nh * 8          # one complex half spectrum
    real = N * N * (N + 2) * 4     # one padded real grid
    cells = N ** 3
    return {
        "k_gen_z": 2 * half,
        "k_col_y": 2 * half,
        "k_c2r_x": half + real,
        "k_col_y (phi)": 2 * half,
        "k_c2r_x (phi)": half + real,
        "k_shells": 2 * cells * 4,
        "k_count": cells * (4 + npop),
        "k_fill+k_write": cells * npop + nsrc * (16 + 16 + 16 + 4 + 6 * 4),
    }


# stage_times() key holding each kernel's device time
KERNEL_STAGE = {"k_gen_z": "gen", "k_col_y": "fft_yz", "k_c2r_x": "fft_x", "k_col_y (phi)": "phi_y",
                "k_c2r_x (phi)": "phi_x", "k_shells": "shells", "k_count": "src_pass", "k_fill+k_write": "src_write"}
# kernels on the second stream (phi's passes)
SECOND_STREAM = ("k_col_y (phi)", "k_c2r_x (phi)")


def traffic_per_launch(kernel, config, N):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE
    and WRITE_SIZE (each scaled by the same counter on a plain 1 GiB device
    copy), from profiles/hbm_traffic_<N>_<config>.txt; None when that
    workload has no such summary."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{N}_{config}.txt")
    if not os.path.exists(path):
        return None
    key = f"hbm_bytes_per_launch {kernel}:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    return None


def time_path_to_kappa(run, ctx, kap, step, nsteps, nsrc, npop):
    """Gaussian -> lognormal -> sources -> kappa: the timed step followed by
    the convergence maps of the workload's source planes (kappa_get_beam_
    properties, the maps copied at the end of the call into one page-locked
    host buffer that every realisation reuses), wall time per realisation over
    nsteps; B_alg adds one read of phi inside the lightcone sphere, (pi/6) 4p
    bytes per cell, and the maps' bytes."""
    import numpy as np
    from colore_amd import lib
    host = lib.load_host()
    N = int(run.scalar("n_grid"))
    zs = [z for z in kap["z_out"] if z <= run.scalar("z_max")]
    r_src = np.array([host.clr_run_r_of_z(run.h, z) for z in zs])
    nside = kap["nside"]
    maps = lib.pinned_empty((len(zs), 12 * nside * nside))
    ctx.kappa(nside, r_src, out=maps)   # warm: the tidal records' buffer
    kms = []
    t0 = time.perf_counter()
    for _ in range(nsteps):
        step(False)
        ctx.kappa(nside, r_src, out=maps)
        kms.append(ctx.stage_times()["kappa"])
    dt = (time.perf_counter() - t0) / nsteps
    cells = float(N) ** 3
    p = (N + 2.0) / N
    b_alg = 36 * p + 20.0 * nsrc / cells + (math.pi / 6) * 4 * p + len(zs) * 12 * nside * nside * 4 / cells
    return {"ms_per_step": dt * 1e3, "cells_per_s": cells / dt, "kappa_ms": float(np.median(kms)), "nside": nside,
            "planes": len(zs), "b_alg_per_cell": b_alg, "achieved": cells / dt * b_alg, "frac": cells / dt * b_alg / HBM_PEAK}


def time_maps(run, ctx, config, sigma2):
    """kappa and ISW at the workload's source planes, the first population's
    line-of-sight pass, the smallest population's skewers, the 21cm shells of
    the first intensity species and a custom map, on the last step's grids,
    then a 1LPT (CIC) density of a new realisation: device ms per call
    (median of 3; one call for skewers, best of 2 for 1LPT) and work rates."""
    import numpy as np
    from colore_amd import lib
    out = {}
    N = int(run.scalar("n_grid"))
    for which in ("kappa", "isw"):
        n = int(run.scalar(f"n_{which}"))
        if n == 0:
            continue
        nside = int(run.scalar(f"nside_{which}"))
        from colore_amd import inputs
        zs = [z for z in inputs.CONFIGS[config][which]["z_out"] if z <= run.scalar("z_max")]
        r_src = np.array([lib.load_host().clr_run_r_of_z(run.h, z) for z in zs])
        ts = []
        for _ in range(3):
            (ctx.kappa if which == "kappa" else ctx.isw)(nside, r_src)
            ts.append(ctx.stage_times()[which])
        ms = sorted(ts)[1]
        out[which] = {"ms": ms, "nside": nside, "planes": n,
                      "pixel_steps_per_s": 12 * nside * nside * (N // 2) / (ms * 1e-3)}
    # the sources' line-of-sight pass (velocity + lensing) on the first
    # population's catalogue, re-sampled alone from the same density
    n0 = ctx.sample_sources(0, int(run.scalar("seed")), 2)
    a2r = run.r_a2r()
    ts = []
    for _ in range(3):
        ctx.source_beams(0, n0, a2r, True)
        ts.append(ctx.stage_times()["beams"])
    ms = sorted(ts)[1]
    out["src_beams"] = {"ms": ms, "sources": n0, "sources_per_s": n0 / (ms * 1e-3)}
    # skewers of the smallest population (N/2 density and velocity values per source)
    counts = [ctx.sample_sources(i, int(run.scalar("seed")), 2) for i in range(int(run.scalar("n_srcs")))]
    ism = int(min(range(len(counts)), key=lambda i: counts[i]))
    ctx.source_skewers(ism, counts[ism], False)
    ms = ctx.stage_times()["skewers"]
    out["src_skewers"] = {"ms": ms, "population": ism, "sources": counts[ism],
                          "values_per_s": counts[ism] * (N // 2) / (ms * 1e-3)}
    if int(run.scalar("n_imap")) > 0:
        # the intensity species joins the shell sums of the last step's density
        amp, bz, model = run.tracer(1, 0)
        ctx.set_tracer(1, 0, model, amp, bz)
        ctx.shell_norms(lognormal_fused=False, sigma2=sigma2, ntracers=int(run.scalar("n_srcs")) + 1)
        nside, r0, rf = run.imap_shells(0)
        ts = []
        for _ in range(3):
            ctx.imap(0, nside, r0, rf)
            ts.append(ctx.stage_times()["imap"])
        ms = sorted(ts)[1]
        out["imap"] = {"ms": ms, "nside": nside, "shells": len(r0), "cells_per_s": N ** 3 / (ms * 1e-3)}
    # a custom projected map at nside 512 whose kernel K(r) is the first population's n(r)
    amp, bz, model = run.tracer(0, 0)
    ctx.set_tracer(2, 0, model, amp, bz)
    ctx.shell_norms(lognormal_fused=False, sigma2=sigma2,
                    ntracers=int(run.scalar("n_srcs")) + int(run.scalar("n_imap")) + 1)
    ts = []
    for _ in range(3):
        ctx.custom_map(0, 512)
        ts.append(ctx.stage_times()["cstm"])
    ms = sorted(ts)[1]
    out["custom_map"] = {"ms": ms, "nside": 512, "pixels_per_s": 12 * 512 * 512 / (ms * 1e-3)}
    # 1LPT density (CIC) of a fresh realisation, in place of the lognormal transform
    ts = []
    for _ in range(2):
        ctx.gaussian_fields(int(run.scalar("seed")), run.scalar("r_smooth"), True)
        ctx.lpt_density(1)
        ts.append(ctx.stage_times()["density"])
    ms = min(ts)
    out["lpt1_cic"] = {"ms": ms, "cells_per_s": N ** 3 / (ms * 1e-3)}
    return out


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size, for scaling the PMC byte
    counters (run under rocprofv3 --pmc next to the bench)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def stream_probe(nbytes=4 << 30, reps=5):
    """What plain streaming kernels reach on this device (torch's copy, in-place
    add, sum and fill over 4 GiB buffers; best of `reps`, HIP events): the
    practical ceiling next to the 8 TB/s published peak."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    ops = {"copy (read + write)": (lambda: b.copy_(a), 2), "in-place add (read + write)": (lambda: a.add_(1.0), 2),
           "sum (read)": (lambda: a.sum(), 1), "fill (write)": (lambda: b.fill_(2.0), 1)}
    out = {}
    for name, (fn, passes) in ops.items():
        best = 1e30
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name] = {"ms": best, "TB/s": passes * nbytes / (best * 1e-3) / 1e12}
    print(json.dumps({"stream_probe_bytes": nbytes, "reps": reps, "results": out}))


def cpu_baseline(cfg, n_cpu_grid):
    """FP64 oracle, same config at n_cpu_grid^3, full step; returns cells/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from colore_amd import inputs, lib
    from oracle_lib import Oracle
    d = tempfile.mkdtemp(prefix="colore_cpu_")
    run = lib.Run(inputs.write_config(cfg, d, n_grid=n_cpu_grid, with_maps=False))
    orc = Oracle(run.table_arrays())
    N, L = n_cpu_grid, run.scalar("l_box")
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    npop = int(run.scalar("n_srcs"))
    t0 = time.perf_counter()
    dens, npot, s2 = orc.fields(N, L, seed, rs, True)
    dens = orc.density(N, L, dens, s2, 0)
    tr = [(run.tracer(0, i)[1], run.tracer(0, i)[2]) for i in range(npop)]
    nz, zmax, cnt, zs, sums = orc.shell_sums(N, L, dens, tr)
    nsrc = 0
    for i in range(npop):
        tab, n0, nf = orc.norm_table(nz, zmax, cnt, zs, sums[i])
        amp, bz, model = run.tracer(0, i)
        pop = orc.pop(N, L, amp, bz, tab, n0, nf, model, i, seed)
        pos, _ = orc.sources(pop, dens, npot, 2)
        nsrc += len(pos)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": N ** 3 / dt, "unit": "cells/s", "cores": threads, "kind": "port", "cpu_model": model,
            "sample": f"FP64 OpenMP oracle, same cosmology/populations at {N}^3 (full step, {nsrc} sources, "
                      f"{dt:.1f} s)"}


def bench_ensemble(args):
    """The batched-ensemble workload: 64 realisations (grids of 128^3 ..
    384^3, 4 Mpc/h cells, one population, a kappa map at the box edge,
    nside 128) as one captured graph; FFT tiles autotuned per size first.
    Times `steps` replays of the graph after one launch that captures it."""
    from colore_amd import inputs, lib
    real = inputs.c4_realisations(64, 4000)
    d = tempfile.mkdtemp(prefix="colore_c4_")
    runs, ctxs = [], []
    for i, r in enumerate(real):
        pops = [(("gauss", r["nz_mean"], 0.1, 5e-4), ("1+z",), "exponential")]
        run = lib.Run(inputs.write_config("C4", os.path.join(d, str(i)), n_grid=r["n_grid"], z_max=r["z_max"],
                                          seed=r["seed"], with_maps=False, populations=pops))
        ctx = lib.Context(r["n_grid"], run.scalar("l_box"))
        ctx.upload_tables(run.tables())
        amp, bz, model = run.tracer(0, 0)
        ctx.set_tracer(0, 0, model, amp, bz)
        runs.append(run)
        ctxs.append(ctx)
    tuned = {}
    for r, ctx in zip(real, ctxs):
        if r["n_grid"] not in tuned:
            best, ms = ctx.fft_autotune(3)
            tuned[r["n_grid"]] = {"choice": best, "ms": [float(x) for x in ms]}
    ens = lib.Ensemble(ctxs, src_cap=2_000_000, nside_kappa=128,
                       r_kappa=np.array([run.scalar("r_max") for run in runs]), nstreams=8)
    seeds = np.array([r["seed"] for r in real], dtype=np.uint32)
    rsm = np.array([run.scalar("r_smooth") for run in runs])
    t0 = time.perf_counter()
    bad, ms_first = ens.run(seeds, rsm)
    t_capture = time.perf_counter() - t0
    for _ in range(args.warmup):
        ens.run(seeds, rsm)
    t0 = time.perf_counter()
    dev = []
    for _ in range(args.steps):
        b, ms = ens.run(seeds, rsm)
        bad = max(bad, b)
        dev.append(ms)
    dt = (time.perf_counter() - t0) / args.steps
    cells = float(sum(r["n_grid"] ** 3 for r in real))
    nsrc = sum(c.catalogue_size(0) for c in ctxs)
    out = {"metric": "grid cells/sec (batched ensemble: fields, lognormal + shell norms, sources, kappa)",
           "value": cells / dt, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (EH98 no-wiggle P(k)); Philox streams",
           "config": {"workload": "64 realisations, seeds 4000..4063, N in {128,192,256,384}, cell 4 Mpc/h, "
                                  "b=1+z, kappa nside 128 at the box edge, one captured HIP graph (8 streams)",
                      "grids": {str(k): sum(1 for r in real if r["n_grid"] == k) for k in (128, 192, 256, 384)},
                      "cells": cells, "sources": nsrc, "invalid_realisations": bad},
           "realisations_per_s": len(real) / dt, "graph_device_ms": float(np.median(dev)),
           "first_launch_incl_capture_s": t_capture, "fft_autotune": tuned}
    ens.close()
    for c in ctxs:
        c.close()
    print(json.dumps(out))


def bench_surrogate(args):
    """Surrogate-fused sampling against the exact two-pass path on the learned-
    surrogate workload (2048^3, z in [0.05, 2.0], the two populations of the
    512^3 precision workload, seed 5005): the network is trained first (its
    own device kernels, shells of 256^3 .. 1024^3 realisations); then per step
    the two-pass sampling (shell sums + sources) and the fused path (the
    read-only shell-moments pass, closure + inference, norm tables, sources)
    are timed on the same Gaussian field, and every shell's norm error
    against the exact reduction is reported."""
    from colore_amd import inputs, lib, surrogate
    t0 = time.perf_counter()
    x, y, groups = surrogate.training_set(with_groups=True)
    net, losses = surrogate.train(x, y, groups=groups)
    t_train = time.perf_counter() - t0
    d = tempfile.mkdtemp(prefix="colore_c5_")
    run = lib.Run(inputs.write_config("C5", d, n_grid=args.n_grid, with_maps=False))
    N, L = int(run.scalar("n_grid")), run.scalar("l_box")
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    npop = int(run.scalar("n_srcs"))
    ctx = lib.Context(N, L)
    ctx.upload_tables(run.tables())
    for i in range(npop):
        amp, bz, model = run.tracer(0, i)
        ctx.set_tracer(0, i, model, amp, bz)
    if not args.no_autotune:
        ctx.fft_autotune(3)
    two, fused = [], []
    for it in range(args.warmup + args.steps):
        # each path starts from the same finished Gaussian field (phi's passes
        # done, so neither waits for them)
        s2 = ctx.gaussian_fields(seed, rs, True)
        ctx.synchronize()
        t = time.perf_counter()
        ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
        n2 = ctx.sample_sources(-1, seed, 2)
        t2 = time.perf_counter() - t
        s2 = ctx.gaussian_fields(seed, rs, True)
        ctx.synchronize()
        t = time.perf_counter()
        nf = ctx.sample_sources_surrogate(net, -1, seed, 2)
        tf = time.perf_counter() - t
        if it >= args.warmup:
            two.append(t2)
            fused.append(tf)
    err, _ = surrogate.shell_errors(ctx, run, net, seed, npop)
    summ = surrogate.summarize(err)
    t2, tf = float(np.median(two)), float(np.median(fused))
    out = {"metric": "surrogate-fused sampling speedup over the exact two-pass sampling (shell sums + sources)",
           "value": t2 / tf, "unit": "x", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": tf * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (EH98 no-wiggle P(k)); Philox streams",
           "config": {"workload": f"{N}^3, z in [0.05, 2.0], populations b=1+z (5e-4) and b=2 (1e-4), seed {seed}",
                      "n_grid": N, "sources_two_pass": n2, "sources_fused": nf},
           "two_pass_ms": t2 * 1e3, "fused_ms": tf * 1e3,
           "fused_cells_per_s": N ** 3 / tf, "two_pass_cells_per_s": N ** 3 / t2,
           "per_shell_norm_error": summ, "shells": int(len(err["network"])),
           "training": {"shells": int(len(x)), "seconds": t_train, "loss_first": float(losses[0]),
                        "loss_last": float(losses[-1])}}
    ctx.close()
    net.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--n-grid", type=int, default=None)
    ap.add_argument("--cpu-grid", type=int, default=768)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--maps", action="store_true",
                    help="also time the kappa/ISW line-of-sight maps and the 21cm painting of the workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (PMC counter calibration)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="time plain streaming kernels (copy, in-place add, sum, fill) over 4 GiB and exit")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default FFT tile shapes (no autotuning before the timed steps)")
    ap.add_argument("--kappa-steps", type=int, default=3,
                    help="realisations timed through the kappa maps after the main steps (0: none)")
    ap.add_argument("--surrogate", action="store_true",
                    help="surrogate-fused sampling against the two-pass path (learned-surrogate workload)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.stream_probe:
        stream_probe()
        return
    if args.gpus == 1 and args.config == "C4":
        bench_ensemble(args)
        return
    if args.gpus == 1 and args.surrogate:
        bench_surrogate(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from colore_amd import inputs, lib
    d = tempfile.mkdtemp(prefix="colore_bench_")
    run = lib.Run(inputs.write_config(args.config, d, n_grid=args.n_grid, with_maps=args.maps))
    N, L = int(run.scalar("n_grid")), run.scalar("l_box")
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    npop = int(run.scalar("n_srcs"))
    ctx = lib.Context(N, L)
    ctx.upload_tables(run.tables())
    for i in range(npop):
        amp, bz, model = run.tracer(0, i)
        ctx.set_tracer(0, i, model, amp, bz)
    # FFT tile shapes: the autotuner times every candidate for this N (setup,
    # not timed)
    tuned = None
    if not args.no_autotune:
        best, tms = ctx.fft_autotune(3)
        tuned = {"choice": best, "candidate_ms": [float(x) for x in tms]}

    acc = {}

    def step(record):
        s2 = ctx.gaussian_fields(seed, rs, True)
        st = ctx.stage_times()
        ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
        st_n = ctx.stage_times()
        st["norm"], st["shells"] = st_n["norm"], st_n["shells"]
        ns = [ctx.sample_sources(-1, seed, 2)]
        st_s = ctx.stage_times()
        for k in ("srcs", "src_pass", "src_write", "fft_phi", "phi_y", "phi_x"):
            st[k] = st_s[k]
        if record:
            for k in ("gen", "fft_yz", "fft_x", "fft_phi", "phi_y", "phi_x", "norm", "shells", "srcs", "src_pass",
                      "src_write"):
                acc[k] = acc.get(k, 0.0) + st[k]
        return ns, s2

    for _ in range(args.warmup):
        step(False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ns, s2 = step(True)
    dt = (time.perf_counter() - t0) / args.steps
    cells = float(N) ** 3
    value = cells / dt

    nsrc_tot = sum(ns)
    kb = kernel_bytes(N, nsrc_tot, npop)
    kt = {k: acc[st] / args.steps for k, st in KERNEL_STAGE.items()}
    def roofline(k):
        a = kb[k] / (kt[k] * 1e-3)
        return {"kernel": k, "bound": "hbm", "achieved": a / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": a / HBM_PEAK, "traffic": traffic_per_launch(k, args.config, N), "bytes_per_launch": kb[k],
                "ms_per_launch": kt[k]}

    dom = max(kt, key=kt.get)
    dom_critical = max((k for k in kt if k not in SECOND_STREAM), key=kt.get)
    p = (N + 2.0) / N
    b_alg = 36 * p + 20.0 * nsrc_tot / cells

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (EH98 no-wiggle P(k), Gaussian n(z)); Philox streams",
        "config": {"workload": f"{args.config}: {N}^3 FP32 field->lognormal->sources, {npop} populations, "
                               f"seed {seed}, FP64 shell sums",
                   "n_grid": N, "l_box": L, "populations": npop, "sources": nsrc_tot, "sigma2": s2,
                   "fft_autotune": tuned},
        "path_roofline": {"b_alg_per_cell": b_alg, "achieved": value * b_alg, "peak": HBM_PEAK,
                          "frac": value * b_alg / HBM_PEAK},
        "stage_ms": {k: v / args.steps for k, v in acc.items()},
        "roofline": roofline(dom),
        "roofline_critical_stream": roofline(dom_critical),
        "kernels": {k: {"ms": kt[k], "GB/s": kb[k] / (kt[k] * 1e-3) / 1e9 if kt[k] > 0 else None}
                    for k in kt},
    }
    kap = inputs.CONFIGS[args.config].get("kappa")
    if kap and args.kappa_steps > 0:
        out["path_to_kappa"] = time_path_to_kappa(run, ctx, kap, step, args.kappa_steps, nsrc_tot, npop)
    if args.maps:
        out["maps"] = time_maps(run, ctx, args.config, s2)
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_grid)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
