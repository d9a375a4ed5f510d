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
`roofline` is for the kernel with the largest device time: algorithmic HBM
bytes per launch over its mean duration (HIP events on the library's stream),
against 8 TB/s.  `cpu_baseline` times the FP64 CPU oracle on the same
configuration at 512^3 on the host cores.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid cells/sec (field→lognormal→sources) at 1024³ FP32; % HBM roofline"
HBM_PEAK = 8.0e12


def kernel_bytes(N, nsrc, npop):
    """Algorithmic HBM bytes per launch of each timed kernel.

    k_gen_z  writes both half spectra (delta and phi) after the z pass;
    k_col_y  reads and writes the delta half spectrum (y pass);
    k_c2r_x  reads the delta half spectrum and writes its padded real grid;
    phi y+x  the same two passes for phi, on a second stream that overlaps
             the passes and stages after them (not a candidate for the
             dominant kernel: its interval includes waiting for shared units);
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
        "phi y+x": 3 * half + real,
        "k_shells": 2 * cells * 4,
        "k_count": cells * (4 + npop),
        "k_fill+k_write": cells * npop + nsrc * (16 + 16 + 16 + 4 + 6 * 4),
    }


# stage_times() key holding each kernel's device time
KERNEL_STAGE = {"k_gen_z": "gen", "k_col_y": "fft_yz", "k_c2r_x": "fft_x", "phi y+x": "fft_phi", "k_shells": "shells",
                "k_count": "src_pass", "k_fill+k_write": "src_write"}


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
    return {"value": N ** 3 / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"FP64 OpenMP oracle, same cosmology/populations at {N}^3 (full step, {nsrc} sources, "
                      f"{dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C3")
    ap.add_argument("--n-grid", type=int, default=None)
    ap.add_argument("--cpu-grid", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--maps", action="store_true",
                    help="also time the kappa/ISW line-of-sight maps and the 21cm painting of the workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (PMC counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
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

    acc = {}

    def step(record):
        s2 = ctx.gaussian_fields(seed, rs, True)
        st = ctx.stage_times()
        ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
        st_n = ctx.stage_times()
        st["norm"], st["shells"] = st_n["norm"], st_n["shells"]
        ns = [ctx.sample_sources(-1, seed, 2)]
        st_s = ctx.stage_times()
        for k in ("srcs", "src_pass", "src_write", "fft_phi"):
            st[k] = st_s[k]
        if record:
            for k in ("gen", "fft_yz", "fft_x", "fft_phi", "norm", "shells", "srcs", "src_pass", "src_write"):
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
    dom = max((k for k in kt if k != "phi y+x"), key=kt.get)
    achieved = kb[dom] / (kt[dom] * 1e-3)
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
                   "n_grid": N, "l_box": L, "populations": npop, "sources": nsrc_tot, "sigma2": s2},
        "path_roofline": {"b_alg_per_cell": b_alg, "achieved": value * b_alg, "peak": HBM_PEAK,
                          "frac": value * b_alg / HBM_PEAK},
        "stage_ms": {k: v / args.steps for k, v in acc.items()},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic_per_launch(dom, args.config, N),
                     "bytes_per_launch": kb[dom], "ms_per_launch": kt[dom]},
        "kernels": {k: {"ms": kt[k], "GB/s": kb[k] / (kt[k] * 1e-3) / 1e9 if kt[k] > 0 else None}
                    for k in kt},
    }
    if args.maps:
        out["maps"] = time_maps(run, ctx, args.config, s2)
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_grid)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
