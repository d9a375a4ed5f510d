"""Parity at the configured grid sizes, through the C ABI.

Every FFT size has its own kernel instantiation (tile shape and radices are
template arguments), so each size the workloads use is checked on its own:

  * single modes (N = 256, 384, 1024, 2048): an injected stream that is zero
    except for four modes -- a general one, one on the non-Hermitian kx = 0
    plane, one on the kx = N/2 plane and one at the y and z Nyquist
    frequencies -- must give the closed-form sum of cosines on sampled planes
    of both grids (RMS error <= 2e-6 of the field RMS, max <= 5e-5);
  * Parseval at 1024^3 on the benchmark realisation (1024^3, z in
    [0.05, 1.6], seed 3003, Philox streams): the device variances of delta
    and phi against the oracle's sums over the modes, within 2e-6;
  * the count pass at 1024^3: every cell of four whole planes (4.2e6 cells,
    three populations) equals the oracle's FP64 count from the device's own
    density and norm tables; the per-cell counts add up to the catalogue size,
    and each sampled plane's sources sit inside their cells;
  * the 512^3 precision workload (z in [0.05, 1.2], two populations, kappa at
    z = 1.0, nside 256) against the FP64 oracle run on its own from the same
    seed: max and RMS error per output, written to a report
    (precision_C2_512.txt in $CLR_REPORT_DIR, default gpurun_out/ when it
    exists); the catalogues bit-exact given the same density and norm tables.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import make_run

pytestmark = pytest.mark.gpu


def _ctx_for(run, n):
    from colore_amd import lib
    ctx = lib.Context(n, run.scalar("l_box"))
    ctx.upload_tables(run.tables())
    for i in range(int(run.scalar("n_srcs"))):
        amp, bz, model = run.tracer(0, i)
        ctx.set_tracer(0, i, model, amp, bz)
    return ctx


def _oracle(run):
    from oracle_lib import Oracle
    return Oracle(run.table_arrays())


# --------------------------------------------------------------- single modes
# (kx, ky, kz) as half-spectrum indices; phase and modulus draws
def _modes(N):
    return [((3, 5, N - 7), 0x2545F491, 0x9E3779B9),
            ((0, 2, 9), 0x6A09E667, 0xBB67AE85),
            ((N // 2, 1, 0), 0x3C6EF372, 0xA54FF53A),
            ((7, N // 2, N // 2), 0x510E527F, 0x1F83D9AB)]


def _mode_amplitudes(orc, run, N, kx, ky, kz):
    """(delta, phi) amplitudes sqrt(P(k)/dk^3) e^{-r^2 k^2/2} and -prefac/k^2
    times it, from the oracle's P(k) interpolation (fourier.c:341-350)."""
    import ctypes
    L = run.scalar("l_box")
    dk = 2 * math.pi / L
    f = lambda i: i if 2 * i <= N else i - N
    k2 = (kx * dk) ** 2 + (f(ky) * dk) ** 2 + (f(kz) * dk) ** 2
    pk = orc.lib.orc_pk_linear0(ctypes.byref(orc.pk), 0.5 * math.log10(k2))
    rs = run.scalar("r_smooth")
    a = math.sqrt(pk / dk ** 3) * math.exp(-0.5 * rs * rs * k2)
    return a, -orc.ta["prefac_lensing"] * a / k2


def _expected_plane(orc, run, N, iz, which):
    """Real grid on plane iz of the injected modes: 2 Re(X e^{i theta}) for
    0 < kx < N/2, Re(X e^{i theta}) on the kx = 0 and N/2 planes (the c2r keeps
    only the real part there), times (sqrt(2 pi)/L)^3."""
    L = run.scalar("l_box")
    norm = (math.sqrt(2 * math.pi) / L) ** 3
    ix = np.arange(N)[None, :]
    iy = np.arange(N)[:, None]
    out = np.zeros((N, N))
    for (kx, ky, kz), x0, x1 in _modes(N):
        ad, ap = _mode_amplitudes(orc, run, N, kx, ky, kz)
        w = math.sqrt(-math.log(1 - x1 / 2 ** 32))
        ph = 2 * math.pi * x0 / 2 ** 32
        amp = (ad if which == 0 else ap) * w
        theta = 2 * math.pi * ((kx * ix + ky * iy + kz * iz) % N) / N + ph
        fac = 1.0 if kx in (0, N // 2) else 2.0
        out += fac * amp * np.cos(theta)
    return out * norm


@pytest.mark.parametrize("N", [128, 192, 256, 384, 1024, 2048])
def test_single_modes_each_fft_size(workdir, N):
    """Every FFT tile-shape candidate of the size (the autotuner may pick any
    of them for a run), one after the other."""
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    nh = N // 2 + 1
    inj = np.zeros(2 * N * N * nh, dtype=np.uint32)
    for (kx, ky, kz), x0, x1 in _modes(N):
        m = kx + nh * (ky + N * kz)
        inj[2 * m], inj[2 * m + 1] = x0, x1
    ctx = _ctx_for(run, N)
    orc = _oracle(run)
    ncand = ctx.fft_candidates()
    assert ncand >= 1
    try:
        for k in range(ncand):
            ctx.fft_select(k)
            print(f"single modes N={N} candidate {k}/{ncand}", flush=True)
            s2 = ctx.gaussian_fields(1, run.scalar("r_smooth"), True, injected=inj)
            if k == ncand - 1:
                inj = None   # 34 GB at 2048^3
            _check_single_modes(ctx, orc, run, N, s2)
    finally:
        ctx.fft_select(0)
        ctx.close()


def _check_single_modes(ctx, orc, run, N, s2):
    for which in (0, 1):
        rms_tot, n_tot, emax, emax_rel = 0.0, 0, 0.0, 0.0
        for iz in (0, 1, N // 2, N - 3):
            g = ctx.fetch_grid_slab(which, iz, 1)[0, :, :N].astype(np.float64)
            want = _expected_plane(orc, run, N, iz, which)
            scale = np.sqrt(np.mean(want ** 2))
            e = g - want
            rms_tot += np.sum(e * e) / scale ** 2
            n_tot += e.size
            emax_rel = max(emax_rel, np.abs(e).max() / scale)
        rms_rel = math.sqrt(rms_tot / n_tot)
        assert rms_rel < 2e-6 and emax_rel < 5e-5, (N, which, rms_rel, emax_rel)
    # variance of the sum of cosines: (fac amp)^2 / 2 per mode
    want_s2 = 0.0
    for (kx, ky, kz), x0, x1 in _modes(N):
        ad, _ = _mode_amplitudes(orc, run, N, kx, ky, kz)
        a = ad * math.sqrt(-math.log(1 - x1 / 2 ** 32)) * (math.sqrt(2 * math.pi) / run.scalar("l_box")) ** 3
        fac = 1.0 if kx in (0, N // 2) else 2.0
        want_s2 += (fac * a) ** 2 / 2
    assert abs(s2 / want_s2 - 1) < 1e-5, (s2, want_s2)


# ------------------------------------------------------ 1024^3 workload checks
@pytest.fixture(scope="module")
def bench_1024(workdir):
    """The benchmark step at 1024^3 (fields, fused lognormal + shell sums,
    sources of all three populations) on the device."""
    path, run = make_run("C3", 1024, workdir, with_maps=False)
    ctx = _ctx_for(run, 1024)
    seed = int(run.scalar("seed"))
    print("bench realisation 1024^3", flush=True)
    s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
    # phi's mean square before the density stage (whole grid, in slabs)
    N = 1024
    acc = 0.0
    for iz in range(0, N, 128):
        g = ctx.fetch_grid_slab(1, iz, 128)[:, :, :N].astype(np.float64)
        acc += np.sum(g * g)
    phi2 = acc / N ** 3
    npop = int(run.scalar("n_srcs"))
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
    nsrc = ctx.sample_sources(-1, seed, 2)
    yield run, ctx, s2, phi2, nsrc
    ctx.close()


def test_parseval_bench_realisation_1024(bench_1024):
    run, ctx, s2, phi2, _ = bench_1024
    vd, vp = _oracle(run).kspace_variance(1024, run.scalar("l_box"), int(run.scalar("seed")), run.scalar("r_smooth"),
                                          True)
    assert abs(s2 / vd - 1) < 2e-6, (s2, vd)
    assert abs(phi2 / vp - 1) < 2e-6, (phi2, vp)


def test_counts_sampled_planes_1024(bench_1024):
    run, ctx, s2, _, nsrc = bench_1024
    N = 1024
    L = run.scalar("l_box")
    seed = int(run.scalar("seed"))
    orc = _oracle(run)
    npop = int(run.scalar("n_srcs"))
    planes = [40, 300, N // 2, N - 77]
    # per-cell counts add up to the catalogue sizes; plane offsets for below
    total = 0
    offsets = []
    for ip in range(npop):
        per_plane = np.zeros(N, dtype=np.int64)
        for iz in range(0, N, 128):
            c = ctx.fetch_cell_counts(ip, iz, 128)
            per_plane[iz:iz + 128] = c.reshape(128, -1).sum(axis=1)
        total += int(per_plane.sum())
        offsets.append(np.concatenate([[0], np.cumsum(per_plane)]))
    assert total == nsrc, (total, nsrc)
    dx = L / N
    for ip in range(npop):
        amp, bz, model = run.tracer(0, ip)
        tab, n0, nf = ctx.norm_table(0, ip)
        pop = orc.pop(N, L, amp, bz, tab, n0, nf, model, ip, seed)
        n_ip = int(offsets[ip][-1])
        pos, _ = ctx.fetch_sources(ip, n_ip)
        for iz in planes:
            dens = ctx.fetch_grid_slab(0, iz, 1).astype(np.float64)
            got = ctx.fetch_cell_counts(ip, iz, 1)
            want = orc.counts_planes(pop, dens, iz)
            bad = np.count_nonzero(got != want)
            assert bad == 0, (ip, iz, bad, int(got.sum()), int(want.sum()))
            # the plane's sources, in cell order, inside their cells
            a, b = int(offsets[ip][iz]), int(offsets[ip][iz + 1])
            assert b - a == int(want.sum())
            if b > a:
                cells = np.repeat(np.arange(N * N), want.reshape(-1))
                cx, cy = cells % N, cells // N
                p = pos[a:b].astype(np.float64)
                for ax, ci in ((0, cx), (1, cy)):
                    c0 = ci * dx - 0.5 * L
                    assert np.all(np.abs(p[:, ax] - c0) <= 0.5 * dx + 1e-3), (ip, iz, ax)
                assert np.all(np.abs(p[:, 2] - (iz * dx - 0.5 * L)) <= 0.5 * dx + 1e-3)


# ------------------------------------------------ 512^3 precision workload
def _err(a, b):
    """(RMS, max) of a - b relative to the RMS of b."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = math.sqrt(np.mean(b * b))
    e = a - b
    return math.sqrt(np.mean(e * e)) / scale, float(np.abs(e).max()) / scale


def _report_path():
    d = os.environ.get("CLR_REPORT_DIR")
    if not d:
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        d = os.path.join(root, "gpurun_out") if os.path.isdir(os.path.join(root, "gpurun_out")) else None
    return os.path.join(d, "precision_C2_512.txt") if d else None


def test_precision_report_c2_512(workdir):
    N = 512
    path, run = make_run("C2", N, workdir)
    L = run.scalar("l_box")
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    npop = int(run.scalar("n_srcs"))
    assert npop == 2
    ctx = _ctx_for(run, N)
    orc = _oracle(run)
    rep = {"workload": f"{N}^3, z in [0.05, 1.2], seed {seed}, 2 populations, kappa z=1.0 nside 256",
           "reference": "FP64 oracle run independently from the same seed (fields, density, shell sums, norms, "
                        "counts, kappa)", "outputs": {}}
    out = rep["outputs"]
    print("C2 fields", flush=True)
    s2 = ctx.gaussian_fields(seed, rs, True)
    gd = ctx.fetch_grid(0)
    gp = ctx.fetch_grid(1)
    od, op, os2 = orc.fields(N, L, seed, rs, True)
    for name, g, o in (("delta_G", gd, od), ("phi", gp, op)):
        r, m = _err(g[:, :, :N], o[:, :, :N])
        out[name] = {"rms_rel": r, "max_rel": m}
        assert r < 2e-6 and m < 5e-5, (name, r, m)
    out["sigma2_G"] = {"rel": abs(s2 / os2 - 1)}
    assert abs(s2 / os2 - 1) < 1e-5
    del gp
    # density: the device's fused lognormal + shell sums
    print("C2 shells", flush=True)
    nz, cnt, zs, sums = ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
    gl = ctx.fetch_grid(0)
    odl = orc.density(N, L, od, os2, 0)                      # FP64 end to end
    same_in = orc.density(N, L, gd.astype(np.float64), s2, 0)  # FP64 from the device's Gaussian field
    del od
    r, m = _err(gl[:, :, :N], odl[:, :, :N])
    out["delta_lognormal"] = {"rms_rel": r, "max_rel": m}
    assert r < 1e-5, r
    r2, m2 = _err(gl[:, :, :N], same_in[:, :, :N])
    out["delta_lognormal_same_input"] = {"rms_rel": r2, "max_rel": m2}
    del same_in
    tracers = [(run.tracer(0, i)[1], run.tracer(0, i)[2]) for i in range(npop)]
    onz, zmax, ocnt, ozs, osums = orc.shell_sums(N, L, odl, tracers)
    assert nz == onz and np.array_equal(cnt, ocnt)
    filled = ocnt > 0   # bin 0 is never filled (density.c:1168)
    out["shell_zsum"] = {"max_rel": float(np.max(np.abs(zs[filled] / ozs[filled] - 1)))}
    for i in range(npop):
        rel = np.abs(sums[i][filled] / osums[i][filled] - 1)
        out[f"shell_bias_sum_pop{i}"] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2)))}
        tab, n0, nf = ctx.norm_table(0, i)
        otab, on0, onf = orc.norm_table(onz, zmax, ocnt, ozs, osums[i])
        rel = np.abs(tab / otab - 1)
        out[f"norm_table_pop{i}"] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2)))}
        assert rel.max() < 1e-5
    # catalogues: bit-exact against the oracle on the device's density and norm
    # tables; end to end, the cells whose count differs from the independent run
    print("C2 sources", flush=True)
    gdl = gl.astype(np.float64)
    del gl
    npot = ctx.fetch_grid(1).astype(np.float64)
    for i in range(npop):
        amp, bz, model = run.tracer(0, i)
        n = ctx.sample_sources(i, seed, 2)
        pos, ipix = ctx.fetch_sources(i, n)
        tab, n0, nf = ctx.norm_table(0, i)
        pop = orc.pop(N, L, amp, bz, tab, n0, nf, model, i, seed)
        tot, cgpu = orc.counts(pop, gdl)
        assert tot == n
        opos, oipix = orc.sources(pop, gdl, npot, 2, cap=n + 1)
        assert np.array_equal(pos[:, :3], opos[:, :3]) and np.array_equal(ipix, oipix)
        r, m = _err(pos[:, 3], opos[:, 3])
        otab, on0, onf = orc.norm_table(onz, zmax, ocnt, ozs, osums[i])
        pop_ind = orc.pop(N, L, amp, bz, otab, on0, onf, model, i, seed)
        tot_ind, cind = orc.counts(pop_ind, odl)
        diff = int(np.count_nonzero(cind != cgpu))
        out[f"catalogue_pop{i}"] = {"sources": int(n), "bit_exact_same_input": True, "dz_rsd_rms_rel": r,
                                    "dz_rsd_max_rel": m, "independent_oracle_sources": int(tot_ind),
                                    "cells_with_other_count": diff, "cells": N ** 3}
        assert diff <= 1e-5 * N ** 3, diff
    del gdl, odl
    # kappa at z = 1.0, nside 256
    print("C2 kappa", flush=True)
    from colore_amd import lib
    r_src = np.array([lib.load_host().clr_run_r_of_z(run.h, 1.0)])
    kg = ctx.kappa(256, r_src)
    ko = orc.kappa_isw(0, N, L, op, 256, r_src)
    r, m = _err(kg, ko)
    out["kappa_z1_nside256"] = {"rms_rel": r, "max_rel": m}
    assert r < 1e-5 and m < 1e-4, (r, m)
    ks = orc.kappa_isw(0, N, L, npot, 256, r_src)
    r2, m2 = _err(kg, ks)
    out["kappa_z1_nside256_same_input"] = {"rms_rel": r2, "max_rel": m2}
    ctx.close()
    p = _report_path()
    text = json.dumps(rep, indent=1)
    print(text)
    if p:
        with open(p, "w") as f:
            f.write(text + "\n")
