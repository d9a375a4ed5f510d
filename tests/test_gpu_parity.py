"""Stage-by-stage parity of the HIP path (through the C ABI) against the FP64
oracle, on the same inputs.

Tolerances (FP32 device grids against the FP64 oracle):
  * Gaussian fields: RMS error <= 2e-6 of the field RMS, max error <= 5e-5 of
    it; sigma^2 within 1e-5 relative.
  * lognormal / clipped density from the same Gaussian grid: |err| <= 1e-6 (1+|delta|).
  * shell sums: counts exact; z and bias sums (FP32 per-cell terms summed
    in FP64) 2e-6 relative; norm tables from the same sums 1e-12 relative.
  * Poisson counts and the catalogue (positions, base pixels): bit-exact given
    the same density grid and norm tables; dz_rsd within 1e-6 relative.
  * kappa / ISW maps from the same potential: 1e-5 of the map RMS.
"""
import os

import numpy as np
import pytest

from conftest import make_run

pytestmark = pytest.mark.gpu


def _ctx_for(run, n):
    from colore_amd import lib
    ctx = lib.Context(n, run.scalar("l_box"))
    ctx.upload_tables(run.tables())
    ntr = [int(run.scalar(k)) for k in ("n_srcs", "n_imap", "n_cstm")]
    for kind in range(3):
        for i in range(ntr[kind]):
            amp, bz, model = run.tracer(kind, i)
            ctx.set_tracer(kind, i, model, amp, bz)
    return ctx


def _oracle(run):
    from oracle_lib import Oracle
    return Oracle(run.table_arrays())


def _field_err(g, o, N):
    a = g[:, :, :N].astype(np.float64)
    b = o[:, :, :N]
    rms = np.sqrt(np.mean(b * b))
    e = a - b
    return np.sqrt(np.mean(e * e)) / rms, np.abs(e).max() / rms


@pytest.mark.parametrize("N", [32, 64, 128, 192])
def test_gaussian_fields_philox(workdir, N):
    path, run = make_run("C1", N, workdir)
    ctx = _ctx_for(run, N)
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    s2 = ctx.gaussian_fields(seed, rs, True)
    d, p = ctx.fetch_grid(0), ctx.fetch_grid(1)
    od, op, os2 = _oracle(run).fields(N, run.scalar("l_box"), seed, rs, True)
    for g, o in ((d, od), (p, op)):
        rms_e, max_e = _field_err(g, o, N)
        assert rms_e < 2e-6 and max_e < 5e-5, (N, rms_e, max_e)
    assert abs(s2 / os2 - 1) < 1e-5, (s2, os2)


@pytest.mark.parametrize("N", [32, 192])
def test_grid_copies_round_trip(workdir, N):
    """Device rows are padded to 128-byte lines; uploads and fetches convert
    to and from the host layout float[N][N][2(N/2+1)] exactly."""
    from colore_amd import lib
    ctx = lib.Context(N, 1000.0)
    rng = np.random.default_rng(N)
    for which in (0, 1):
        g = rng.standard_normal(ctx.padded_shape).astype(np.float32)
        ctx.upload_grid(which, g)
        assert np.array_equal(ctx.fetch_grid(which), g)


def test_gaussian_fields_unsmoothed_potential(workdir):
    """smooth_potential = false: the potential's amplitudes come from their
    own table (the density alone is smoothed)."""
    N = 64
    path, run = make_run("C1", N, workdir)
    ctx = _ctx_for(run, N)
    seed, rs = int(run.scalar("seed")), run.scalar("r_smooth")
    assert rs > 0
    s2 = ctx.gaussian_fields(seed, rs, False)
    d, p = ctx.fetch_grid(0), ctx.fetch_grid(1)
    od, op, os2 = _oracle(run).fields(N, run.scalar("l_box"), seed, rs, False)
    for g, o in ((d, od), (p, op)):
        rms_e, max_e = _field_err(g, o, N)
        assert rms_e < 2e-6 and max_e < 5e-5, (rms_e, max_e)
    assert abs(s2 / os2 - 1) < 1e-5, (s2, os2)


def test_gaussian_fields_reference_stream(workdir):
    """Fed the reference's per-thread MT19937 mode stream (8 threads), the
    device spectra transform to the oracle's grids."""
    from oracle_lib import load, p
    import ctypes
    N = 64
    path, run = make_run("C1", N, workdir)
    lib = load()
    nmodes = N * N * (N // 2 + 1)
    inj = np.zeros(2 * nmodes, dtype=np.uint32)
    lib.orc_ref_mode_stream(N, 1001, 8, p(inj, ctypes.c_uint32))
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(1001, run.scalar("r_smooth"), True, injected=inj)
    od, op, os2 = _oracle(run).fields(N, run.scalar("l_box"), 1001, run.scalar("r_smooth"), True, injected=inj)
    for g, o in ((ctx.fetch_grid(0), od), (ctx.fetch_grid(1), op)):
        rms_e, max_e = _field_err(g, o, N)
        assert rms_e < 2e-6 and max_e < 5e-5, (rms_e, max_e)
    assert abs(s2 / os2 - 1) < 1e-5


@pytest.mark.parametrize("dens_type", [0, 3])
def test_density_transform(workdir, dens_type):
    N = 64
    path, run = make_run("C1", N, workdir)
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(7, run.scalar("r_smooth"), True)
    g0 = ctx.fetch_grid(0).astype(np.float64)
    ctx.density(dens_type, s2)
    g1 = ctx.fetch_grid(0)
    o = _oracle(run).density(N, run.scalar("l_box"), g0, s2, dens_type)
    e = np.abs(g1[:, :, :N] - o[:, :, :N]) / (1 + np.abs(o[:, :, :N]))
    assert e.max() < 1e-6, e.max()


@pytest.mark.parametrize("N", [64, 128])
def test_shell_norms(workdir, N):
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(3003, run.scalar("r_smooth"), True)
    g0 = ctx.fetch_grid(0).astype(np.float64)
    ntr = int(run.scalar("n_srcs"))
    nz, cnt, zs, sums = ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=ntr)
    gl = ctx.fetch_grid(0)
    orc = _oracle(run)
    L = run.scalar("l_box")
    ol = orc.density(N, L, g0, s2, 0)
    assert np.abs(gl[:, :, :N] - ol[:, :, :N]).max() / (1 + np.abs(ol).max()) < 1e-6
    tracers = [(run.tracer(0, i)[1], run.tracer(0, i)[2]) for i in range(ntr)]
    onz, zmax, ocnt, ozs, osums = orc.shell_sums(N, L, gl.astype(np.float64), tracers)
    assert nz == onz
    assert np.array_equal(cnt, ocnt)
    np.testing.assert_allclose(zs, ozs, rtol=2e-6)
    np.testing.assert_allclose(sums, osums, rtol=2e-6)
    for i in range(ntr):
        tab, n0, nf = ctx.norm_table(0, i)
        otab, on0, onf = orc.norm_table(nz, zmax, cnt, zs, sums[i])
        np.testing.assert_allclose(tab, otab, rtol=1e-12)
        assert n0 == on0 and nf == onf


def _sources_vs_oracle(run, N, ipop, seed, edit=None):
    """Samples population ipop on the device and checks the catalogue against
    the oracle's; returns (context, oracle, positions, oracle per-cell counts).
    edit(amp, bz, model) -> (amp, bz, model) replaces the population's tables
    on both sides."""
    ctx = _ctx_for(run, N)
    amp, bz, model = run.tracer(0, ipop)
    if edit is not None:
        amp, bz, model = edit(amp.copy(), bz.copy(), model)
        ctx.set_tracer(0, ipop, model, amp, bz)
    s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
    ctx.sigma2_gauss = s2
    ntr = int(run.scalar("n_srcs"))
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=ntr)
    dens = ctx.fetch_grid(0).astype(np.float64)
    npot = ctx.fetch_grid(1).astype(np.float64)
    n = ctx.sample_sources(ipop, seed, 2)
    pos, ipix = ctx.fetch_sources(ipop, n)
    tab, n0, nf = ctx.norm_table(0, ipop)
    orc = _oracle(run)
    pop = orc.pop(N, run.scalar("l_box"), amp, bz, tab, n0, nf, model, ipop, seed)
    tot, counts = orc.counts(pop, dens)
    assert n == tot and n > 0
    opos, oipix = orc.sources(pop, dens, npot, 2, cap=n + 1)
    assert np.array_equal(pos[:, :3], opos[:, :3])
    assert np.array_equal(ipix, oipix)
    np.testing.assert_allclose(pos[:, 3], opos[:, 3], rtol=1e-6, atol=1e-12)
    return ctx, orc, pos, counts


@pytest.mark.parametrize("N", [32, 64])
def test_sources_bit_exact_large_lambda(workdir, N):
    """One population at 1e-3 (Mpc/h)^-3 on coarse grids: lambda per cell
    reaches the hundreds, past the FP32 search (lambda >= 64) and, at N=32,
    past 500 where the count is a sum of inversions."""
    path, run = make_run("C1", N, workdir, with_maps=False)
    counts = _sources_vs_oracle(run, N, 0, 1001)[3]
    assert counts.max() > (500 if N == 32 else 64), counts.max()


def test_sources_bit_exact_192(workdir):
    """A 3*2^n grid: 1024-cell blocks span several rows."""
    path, run = make_run("C3", 192, workdir, z_max=1.6, with_maps=False)
    _sources_vs_oracle(run, 192, 1, 3003)


@pytest.mark.parametrize("model", [0, 2])
def test_sources_negative_bias(workdir, model):
    """b(z) < 0: the lane bound does not apply (the bias grows toward voids),
    so every cell inside the sphere takes the per-cell path.  (Model 1 with
    b < 0 is left out: exp(b d/(1+d)) is unbounded as d -> -1.)"""
    path, run = make_run("C3", 64, workdir, z_max=1.6, with_maps=False)

    def edit(amp, bz, m):
        return amp, np.full_like(bz, -0.6), model

    _sources_vs_oracle(run, 64, 0, 3003, edit)


def test_sources_mean_out_of_range_is_an_error(workdir):
    """A cell whose Poisson mean is not finite or above 2^20 gets no count:
    the call fails with CLR_ERR_ARG naming the cell (here bias model 1 with
    b = -0.6 and one cell's delta just above -1, so exp(b d/(1+d)) overflows),
    and the context samples normally once the grid is sane again."""
    N = 64
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx = _ctx_for(run, N)
    amp, bz, model = run.tracer(0, 0)
    ctx.set_tracer(0, 0, 1, amp, np.full_like(bz, -0.6))
    s2 = ctx.gaussian_fields(3003, run.scalar("r_smooth"), True)
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=int(run.scalar("n_srcs")))
    g = ctx.fetch_grid(0)
    cell = (N // 2, N // 2, N // 2 + 11)   # z, y, x: r = 11 cells ~ 1.1 Gpc/h, inside the n(z) of population 0
    orig = g[cell]
    g[cell] = np.float32(-0.9999999)
    ctx.upload_grid(0, g)
    print("sampling with one cell at delta = -0.9999999", flush=True)
    with pytest.raises(RuntimeError, match="Poisson mean"):
        ctx.sample_sources(0, 3003, 2)
    g[cell] = orig
    ctx.upload_grid(0, g)
    assert ctx.sample_sources(0, 3003, 2) > 0
    ctx.close()


def test_sources_gaps_in_nz(workdir):
    """n(z) with empty stretches, a single nonzero node and a negative node:
    zero bins, bins the FP64 path decides and coarse bounds of 0 next to
    populated ones."""
    path, run = make_run("C3", 128, workdir, z_max=1.6, with_maps=False)

    def edit(amp, bz, m):
        na = len(amp)
        amp[na // 5: na // 4] = 0.0
        amp[na // 3] = 0.0
        amp[na // 3 + 40] = -amp[na // 3 + 41]
        amp[na // 2: na // 2 + 300] = 0.0
        amp[na // 2 + 150] = 3 * amp[na // 2 + 301]
        return amp, bz, m

    _sources_vs_oracle(run, 128, 0, 3003, edit)


@pytest.mark.parametrize("N,ipop", [(64, 0), (128, 0), (128, 1), (128, 2)])
def test_sources_bit_exact(workdir, N, ipop):
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx, orc, pos, _ = _sources_vs_oracle(run, N, ipop, 3003)
    props = ctx.source_properties(ipop, len(pos))
    oprops = orc.source_props(pos)
    np.testing.assert_allclose(props, oprops, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("N,ipop,lensing", [(64, 0, True), (128, 1, True), (64, 2, False)])
def test_source_beams(workdir, N, ipop, lensing):
    """The sources' line-of-sight pass on the sampled catalogue and phi: the
    CIC radial velocity (dz_rsd with the V1(r_of_z(z0)) scaling) and the
    nearest-cell march (e1, e2, kappa, dra, ddec) against the oracle's FP64
    restatement: every value within 1e-5 of its column's RMS."""
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx, orc, pos, _ = _sources_vs_oracle(run, N, ipop, 3003)
    npot = ctx.fetch_grid(1).astype(np.float64)
    a2r = run.r_a2r()
    dz, lens = ctx.source_beams(ipop, len(pos), a2r, lensing)
    odz, olens = orc.src_beams(N, run.scalar("l_box"), npot, pos, a2r, lensing)
    cols = [(dz, odz)] + ([(lens[:, k], olens[:, k]) for k in range(5)] if lensing else [])
    for k, (g, o) in enumerate(cols):
        rms = np.sqrt(np.mean(o * o))
        assert rms > 0, k
        err = np.abs(g.astype(np.float64) - o).max()
        assert err <= 1e-5 * rms, (k, err, rms)
    # the pass replaces the catalogue's dz_rsd, it does not add to it
    assert not np.allclose(dz, pos[:, 3])


@pytest.mark.parametrize("gaussian", [False, True])
def test_source_skewers(workdir, gaussian):
    """Skewers of every source (CIC density or Gaussian field, scaled radial
    velocity, N/2 steps) against the oracle: within 1e-5 of the RMS, zeros
    past each source's last step."""
    N = 64
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx, orc, pos, _ = _sources_vs_oracle(run, N, 0, 3003)
    dens = ctx.fetch_grid(0).astype(np.float64)
    npot = ctx.fetch_grid(1).astype(np.float64)
    dg, vs = ctx.source_skewers(0, len(pos), gaussian, ctx.sigma2_gauss)
    odg, ovs = orc.src_skewers(N, run.scalar("l_box"), dens, npot, pos, gaussian, ctx.sigma2_gauss)
    for g, o in ((dg, odg), (vs, ovs)):
        rms = np.sqrt(np.mean(o * o))
        assert rms > 0
        err = np.abs(g.astype(np.float64) - o).max()
        assert err <= 1e-5 * rms, (err, rms)
        assert np.array_equal(g == 0, o == 0)


@pytest.mark.parametrize("fused", [True, False])
def test_custom_map(workdir, fused):
    """A custom projected map (Gaussian K(z) at z = 0.5, b = 1.2, nside 16)
    from the device density against the oracle's march on the same grid and
    norm table: within 1e-5 of the map's RMS."""
    N, nside = 64, 16
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False, custom=((0.5, 0.1, nside, 1.2),))
    assert run.scalar("n_cstm") == 1
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(int(run.scalar("seed")), run.scalar("r_smooth"), True)
    if not fused:
        ctx.density(0, s2)
    ctx.shell_norms(lognormal_fused=fused, sigma2=s2, ntracers=int(run.scalar("n_srcs")) + 1)
    dens = ctx.fetch_grid(0).astype(np.float64)
    out = ctx.custom_map(0, nside)
    kz, bz, model = run.tracer(2, 0)
    tab, n0, nf = ctx.norm_table(2, 0)
    o = _oracle(run).cstm(N, run.scalar("l_box"), dens, kz, bz, tab, n0, nf, model, nside)
    rms = np.sqrt(np.mean(o * o))
    assert rms > 0
    assert np.abs(out - o).max() <= 1e-5 * rms, (np.abs(out - o).max(), rms)


def test_line_of_sight_passes_with_no_sources(workdir):
    """A population with n(z) = 0 samples no source; its line-of-sight pass and
    skewers then return nothing, without error (empty catalogue)."""
    N = 64
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx = _ctx_for(run, N)
    amp, bz, model = run.tracer(0, 1)
    ctx.set_tracer(0, 1, model, np.zeros_like(amp), bz)
    s2 = ctx.gaussian_fields(3003, run.scalar("r_smooth"), True)
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=int(run.scalar("n_srcs")))
    assert ctx.sample_sources(1, 3003, 2) == 0
    dz, lens = ctx.source_beams(1, 0, run.r_a2r(), True)
    assert dz.shape == (0,) and lens.shape == (0, 5)
    dg, vs = ctx.source_skewers(1, 0, True, s2)
    assert dg.shape == (0, N // 2) and vs.shape == (0, N // 2)
    # and a wrong count is refused
    with pytest.raises(RuntimeError):
        ctx.source_beams(1, 5, run.r_a2r(), False)


def test_source_beams_192(workdir):
    """The line-of-sight pass on a 3*2^n grid (192^3)."""
    N = 192
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx, orc, pos, _ = _sources_vs_oracle(run, N, 2, 3003)
    npot = ctx.fetch_grid(1).astype(np.float64)
    a2r = run.r_a2r()
    dz, lens = ctx.source_beams(2, len(pos), a2r, True)
    odz, olens = orc.src_beams(N, run.scalar("l_box"), npot, pos, a2r, True)
    for k, (g, o) in enumerate([(dz, odz)] + [(lens[:, j], olens[:, j]) for j in range(5)]):
        rms = np.sqrt(np.mean(o * o))
        assert np.abs(g.astype(np.float64) - o).max() <= 1e-5 * rms, k


@pytest.mark.parametrize("order,interp,N", [(1, 0, 64), (1, 1, 64), (1, 2, 64), (2, 1, 64), (2, 2, 64),
                                            (1, 1, 128), (1, 2, 128), (2, 1, 192), (1, 1, 256)])
def test_lpt_density(workdir, order, interp, N):
    """1LPT and 2LPT densities from a Gaussian field on the device against the
    oracle's FP64 transforms and deposit of the same field: CIC and TSC within 1e-4
    up to 128^3 (FP32 FFT and atomics), NGP counts equal but in at most 1e-4 of the cells
    (a particle within rounding of a cell boundary), the same mass.  The
    field is the generated one times 25, so that displacements cross cells.
    At 128^3 the LDS-staged deposit has interior tiles as well as wrapped ones;
    192^3 and 256^3 run other radix lists of the project's r2c and c2r (max
    5e-4 there, RMS 4e-5 everywhere)."""
    path, run = make_run("C1", N, workdir, with_maps=False)
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(int(run.scalar("seed")), run.scalar("r_smooth"), True)
    g = ctx.fetch_grid(0) * np.float32(25)
    ctx.upload_grid(0, g)
    ctx.declare_grids(s2, False)
    dg = g.astype(np.float64)
    pos = ctx.lpt_density(interp, order, positions=True)
    d = ctx.fetch_grid(0)[:, :, :N].astype(np.float64)
    ta = run.table_arrays()
    o, opos = _oracle(run).lpt1(N, run.scalar("l_box"), dg, interp, with_pos=True, order=order, d2=ta["d2"])
    o = o[:, :, :N]
    # particle positions (cell order), as write_lpt dumps them: FP32 FFTs
    dp = np.abs(pos - opos.reshape(-1, 3))
    dp = np.minimum(dp, run.scalar("l_box") - dp)   # across the periodic edge
    assert dp.max() < 1e-3 * run.scalar("l_box") / N, dp.max()
    assert abs(d.sum()) < 1e-3 * N ** 3 and abs(o.sum()) < 1e-6 * N ** 3
    assert np.abs(o).max() > 0.1
    if interp == 0:
        diff = np.abs(d - o)
        assert np.all((diff < 1e-6) | (np.abs(diff - np.round(diff)) < 1e-6))
        assert np.count_nonzero(diff > 0.5) <= 1e-4 * N ** 3 + 2
    else:
        # FP32 transforms: the density error grows with the grid (log N
        # passes, more particles per cell edge); a wrong index or twiddle
        # gives errors of order the field itself
        e = np.abs(d - o)
        rms = float(np.sqrt(np.mean(e * e)))
        print(f"lpt order {order} interp {interp} N {N}: max {e.max():.3g} rms {rms:.3g}", flush=True)
        assert e.max() < (1e-4 if N <= 128 else 5e-4), e.max()
        assert rms < 4e-5, rms


def test_sources_all_populations_one_pass(workdir):
    """clr_sample_sources(ipop=-1) gives the catalogues of the per-population
    calls (same streams, same order)."""
    N = 128
    path, run = make_run("C3", N, workdir, z_max=1.6, with_maps=False)
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(3003, run.scalar("r_smooth"), True)
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=3)
    per = []
    for i in range(3):
        n = ctx.sample_sources(i, 3003, 2)
        per.append(ctx.fetch_sources(i, n))
    tot = ctx.sample_sources(-1, 3003, 2)
    assert tot == sum(len(p[0]) for p in per)
    for i in range(3):
        pos, ipix = ctx.fetch_sources(i, len(per[i][0]))
        assert np.array_equal(pos, per[i][0]) and np.array_equal(ipix, per[i][1])


@pytest.mark.parametrize("which", ["kappa", "isw"])
def test_los_maps(workdir, which):
    N = 64
    path, run = make_run("C2", N, workdir, z_max=1.2)
    ctx = _ctx_for(run, N)
    ctx.gaussian_fields(2002, run.scalar("r_smooth"), True)
    npot = ctx.fetch_grid(1).astype(np.float64)
    from colore_amd import lib
    host = lib.load_host()
    import ctypes
    r_src = np.array([run.scalar("r_max") * 0.5, run.scalar("r_max") * 0.95])
    nside = 16
    call = ctx.kappa if which == "kappa" else ctx.isw
    out = call(nside, r_src)
    o = _oracle(run).kappa_isw(0 if which == "kappa" else 1, N, run.scalar("l_box"), npot, nside, r_src)
    rms = np.sqrt(np.mean(o ** 2))
    assert np.abs(out - o).max() < 1e-5 * rms + 1e-30, (np.abs(out - o).max(), rms)
    # the same maps into a page-locked buffer, twice (the second call
    # overwrites the first's values)
    pin = lib.pinned_empty(out.shape)
    pin[:] = np.nan
    for _ in range(2):
        assert call(nside, r_src, out=pin) is pin
        assert np.array_equal(pin, out)


@pytest.mark.parametrize("N,nside,bands", [(64, 16, 8), (128, 64, 16), (192, 64, 12), (256, 128, 32)])
def test_imap_vs_oracle(workdir, N, nside, bands):
    """21cm intensity shells painted on the device against the oracle's walk
    over the same grids: identical hit counts per pixel, temperatures summed
    in FP32 atomics within 1e-5 of the FP64 sums.  The larger case paints
    up to 2e8 sub-points, so that many of them fall near a pixel or shell edge,
    where the device's shortened pixel evaluation defers to the full one."""
    path, run = make_run("C3", N, workdir, z_max=1.6, imap_bands=bands, imap_nside=nside)
    ctx = _ctx_for(run, N)
    s2 = ctx.gaussian_fields(3003, run.scalar("r_smooth"), True)
    ntr = int(run.scalar("n_srcs")) + int(run.scalar("n_imap"))
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=ntr)
    dens = ctx.fetch_grid(0).astype(np.float64)
    npot = ctx.fetch_grid(1).astype(np.float64)
    nside, r0, rf = run.imap_shells(0)
    data, nadd = ctx.imap(0, nside, r0, rf)
    tz, bz, model = run.tracer(1, 0)
    tab, n0, nf = ctx.norm_table(1, 0)
    od, on = _oracle(run).imap(N, run.scalar("l_box"), dens, npot, tz, bz, tab, n0, nf, model, nside, r0, rf)
    assert nadd.sum() > 0
    assert np.array_equal(nadd, on)
    scale = np.abs(od).max()
    assert np.abs(data - od).max() <= 1e-5 * scale, (np.abs(data - od).max(), scale)


def test_cli_end_to_end(tmp_path):
    """`CoLoRe param.cfg` on the device: every output file of the run, the
    catalogues holding exactly the sources the library samples for the same
    seed and grids."""
    import glob
    import subprocess
    from conftest import ROOT
    from colore_amd import inputs, lib
    N = 64
    p = inputs.write_config("C3", str(tmp_path), n_grid=N, z_max=1.6, output_format="FITS", output_density=True,
                            imap_bands=4, imap_nside=16, kappa_nside=16)
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    files = sorted(os.path.basename(f) for f in glob.glob(str(tmp_path / "out*")))
    for want in ("out_dens_gaussian_0.dat", "out_dens_lightcone_0.dat", "out_srcs_s1_0.fits",
                 "out_srcs_s3_0.fits", "out_kappa_z000.fits", "out_isw_z002.fits", "out_imap_s1_nu003.fits"):
        assert want in files, (want, files)
    run = lib.Run(p)
    ctx = _ctx_for(run, N)
    seed = int(run.scalar("seed"))
    s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
    ctx.density(0, s2)
    ctx.shell_norms(lognormal_fused=False, sigma2=s2, ntracers=int(run.scalar("n_srcs")) + int(run.scalar("n_imap")))
    for i in range(int(run.scalar("n_srcs"))):
        n = ctx.sample_sources(i, seed, int(run.scalar("nside_base")))
        with open(tmp_path / f"out_srcs_s{i + 1}_0.fits", "rb") as f:
            head = f.read(2880 * 4).decode("ascii", "replace")
        k = head.index("NAXIS2")
        assert int(head[k + 10:k + 30].strip()) == n, (i, head[k:k + 30], n)


def test_cli_lensing_catalogue(tmp_path):
    """`CoLoRe param.cfg` with include_lensing on the first population: its
    FITS catalogue carries E1 E2 KAPPA DRA DDEC, and every column equals the
    library's source properties and line-of-sight pass for the same seed; the
    other populations' DZ_RSD is the line-of-sight value too (the reference
    replaces it whenever it marches lines of sight), and their files carry the
    density (second) and Gaussian (third) skewers of the library; the custom
    map file holds the library's map in RING order."""
    import subprocess
    from conftest import ROOT
    from colore_amd import inputs, lib
    from test_host import _read_fits_table
    N = 64
    p = inputs.write_config("C3", str(tmp_path), n_grid=N, z_max=1.6, output_format="FITS", with_maps=False,
                            lensing=(0,), skewers={1: "density", 2: "gaussian"}, custom=((0.6, 0.15, 8, 1.0),))
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    run = lib.Run(p)
    assert run.scalar("n_srcs") == 3
    ctx = _ctx_for(run, N)
    seed = int(run.scalar("seed"))
    s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
    ctx.density(0, s2)
    ctx.shell_norms(lognormal_fused=False, sigma2=s2, ntracers=4)
    a2r = run.r_a2r()
    # the custom map, written in RING order
    cm = ctx.custom_map(0, 8)
    (_, _), (h1, tab) = _read_fits_table(str(tmp_path / "out_custom_s1.fits"))
    assert h1["ORDERING"] == "RING" and h1["NSIDE"] == "8"
    host = lib.load_host()
    ring = np.zeros_like(cm)
    ring[[host.clr_hp_nest2ring(8, i) for i in range(len(cm))]] = cm
    np.testing.assert_array_equal(np.frombuffer(tab, dtype=">f4"), ring)
    for i in range(3):
        n = ctx.sample_sources(i, seed, int(run.scalar("nside_base")))
        props = ctx.source_properties(i, n)
        dz, lens = ctx.source_beams(i, n, a2r, lensing=(i == 0))
        hdus = _read_fits_table(str(tmp_path / f"out_srcs_s{i + 1}_0.fits"))
        assert len(hdus) == (2 if i == 0 else 5)
        (_, _), (h1, tab) = hdus[:2]
        nf = 10 if i == 0 else 5
        assert h1["TFIELDS"] == str(nf) and int(h1["NAXIS2"]) == n
        rec = np.frombuffer(tab, dtype=[("t", ">i4")] + [(f"c{k}", ">f4") for k in range(nf - 1)])
        assert np.all(rec["t"] == i)
        np.testing.assert_array_equal(np.stack([rec[f"c{k}"] for k in range(3)], 1), props[:, :3])
        np.testing.assert_array_equal(rec["c3"], dz)
        if i == 0:
            np.testing.assert_array_equal(np.stack([rec[f"c{k}"] for k in range(4, 9)], 1), lens)
        else:
            dg, vs = ctx.source_skewers(i, n, gaussian=(i == 2), sigma2=s2)
            (hd, dd), (hv, dv), (hb, _) = hdus[2:]
            assert hd["CONTENTS"] == ("gaussian skewers" if i == 2 else "density skewers")
            np.testing.assert_array_equal(np.frombuffer(dd, dtype=">f4").reshape(n, N // 2), dg)
            np.testing.assert_array_equal(np.frombuffer(dv, dtype=">f4").reshape(n, N // 2), vs)
            assert abs(float(hb["SIGMA_G"]) - np.sqrt(s2)) < 1e-6 * np.sqrt(s2)


def test_cli_lpt(tmp_path):
    """`CoLoRe param.cfg` with dens_type = 1 (1LPT, CIC) and output_lpt: the run
    completes, its lightcone density dump holds the library's 1LPT density of
    the same seed (FP32 atomics: equal to 1e-5) and its Gadget snapshot the
    library's particle positions."""
    import subprocess
    from conftest import ROOT
    from colore_amd import inputs, lib
    N = 64
    p = inputs.write_config("C1", str(tmp_path), n_grid=N, dens_type=1, output_density=True, with_maps=False,
                            output_lpt=True)
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "1LPT" in r.stdout
    raw = open(tmp_path / "out_dens_lightcone_0.dat", "rb").read()
    hdr = np.frombuffer(raw[:8], dtype=np.int32)
    assert hdr[1] == 4
    n_grid, nz = np.frombuffer(raw[16:28], dtype=np.int32)[:2]
    assert n_grid == N and nz == N
    grid = np.frombuffer(raw[28:], dtype=np.float32).reshape(N, N, N)
    run = lib.Run(p)
    ctx = _ctx_for(run, N)
    ctx.gaussian_fields(int(run.scalar("seed")), run.scalar("r_smooth"), True)
    pos = ctx.lpt_density(1, positions=True)
    d = ctx.fetch_grid(0)[:, :, :N]
    np.testing.assert_allclose(grid, d, rtol=0, atol=1e-5)
    assert np.abs(d).max() > 0.01
    # output_lpt: the Gadget snapshot's position block
    snap = open(tmp_path / "out_lpt_out.0", "rb").read()
    o = 8 + 256
    assert np.frombuffer(snap[o:o + 4], dtype="<i4")[0] == 12 * N ** 3
    np.testing.assert_allclose(np.frombuffer(snap[o + 4:o + 4 + 12 * N ** 3], dtype="<f4").reshape(-1, 3), pos,
                               rtol=1e-6)


def test_cli_reference_example_layout(tmp_path):
    """A param file with the sections and switches of the reference's own
    example configuration (1LPT with CIC, lensing and skewers on the first
    population, skewers on the second, an intensity species, kappa and ISW
    at z = 0.4, predictions every 0.1 in z, the fast-lensing section that a
    build without it ignores), on synthetic inputs at 64^3: the run completes
    and writes every file the reference writes for it."""
    import glob
    import subprocess
    from conftest import ROOT
    from colore_amd import inputs
    from test_host import _read_fits_table
    N = 64
    p = inputs.write_config("C3", str(tmp_path), n_grid=N, z_max=0.45, dens_type=1, output_format="FITS",
                            output_density=True, imap_bands=4, imap_nside=16, lensing=(0,),
                            skewers={0: "density", 1: "density"}, pred_dz=0.1, just_pred=False)
    with open(p, "a") as f:
        f.write("kappa:\n{\n  z_out= [0.4]\n  nside= 16\n}\nisw:\n{\n  z_out= [0.4]\n  nside= 16\n}\n"
                'lensing:\n{\n  nside= 256\n  n_lensing= 11\n  spacing_type= "r"\n  write=false\n}\n')
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "1LPT" in r.stdout and "*** Writing predictions" in r.stdout
    files = sorted(os.path.basename(f) for f in glob.glob(str(tmp_path / "out*")))
    for want in ("out_gbias.txt", "out_pk_srcs_pop0_z0.100.txt", "out_xi_imap_pop0_z0.400.txt",
                 "out_dens_gaussian_0.dat", "out_dens_lightcone_0.dat", "out_srcs_s1_0.fits", "out_srcs_s2_0.fits",
                 "out_srcs_s3_0.fits", "out_kappa_z000.fits", "out_isw_z000.fits", "out_imap_s1_nu000.fits",
                 "out_imap_s1_nu003.fits"):
        assert want in files, (want, files)
    shapes = []
    for i in range(3):
        hdus = _read_fits_table(str(tmp_path / f"out_srcs_s{i + 1}_0.fits"))
        shapes.append((len(hdus), hdus[1][0]["TFIELDS"]))
    assert shapes == [(5, "10"), (5, "5"), (2, "5")], shapes


@pytest.mark.parametrize("N,nthr,ipop", [(128, 8, 0), (64, 13, 0)])
def test_sources_reference_stream(workdir, N, nthr, ipop):
    """Reference-stream mode: fed the reference's per-thread MT19937 streams
    (seed + t + nthr ipop, srcs.c:150; libgomp's static plane split), the
    device's Poisson counts, catalogue order, positions and base pixels equal
    the oracle's restatement of srcs.c:120-283 with gsl_ran_poisson bit for
    bit, on the device's density and norm tables; dz_rsd within 1e-6.  At
    128^3 (z in [0.05, 0.5], n = 1e-3 peak) lambda reaches ~9 (1+delta), so
    both of gsl_ran_poisson's branches run."""
    from oracle_lib import ref_streams, static_split
    path, run = make_run("C1", N, workdir, with_maps=False)
    ctx = _ctx_for(run, N)
    seed = int(run.scalar("seed"))
    s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=int(run.scalar("n_srcs")))
    dens = ctx.fetch_grid(0).astype(np.float64)
    npot = ctx.fetch_grid(1).astype(np.float64)
    izb = static_split(N, nthr)
    amp, bz, model = run.tracer(0, ipop)
    tab, n0, nf = ctx.norm_table(0, ipop)
    orc = _oracle(run)
    pop = orc.pop(N, run.scalar("l_box"), amp, bz, tab, n0, nf, model, ipop, seed)
    # stream length: the oracle's walk tells whether it suffices (doubling)
    per_thread = int(10 * N ** 3 / nthr) + 100000
    while True:
        st = ref_streams(seed, ipop, nthr, per_thread)
        try:
            opos, oipix, ocounts = orc.sources_ref_stream(pop, dens, npot, nthr, izb, st)
            break
        except AssertionError:
            per_thread *= 2
            assert per_thread < 1 << 26
    n = ctx.sample_sources_stream(ipop, nthr, izb, st)
    pos, ipix = ctx.fetch_sources(ipop, n)
    counts = ctx.fetch_cell_counts(ipop, 0, N).reshape(-1)
    assert n == len(opos) and n > 0
    np.testing.assert_array_equal(counts, ocounts)
    assert np.array_equal(pos[:, :3], opos[:, :3])
    assert np.array_equal(ipix, oipix)
    np.testing.assert_allclose(pos[:, 3], opos[:, 3], rtol=1e-6, atol=1e-12)
    lam_peak = counts.max()
    if N == 128:
        assert lam_peak > 20, lam_peak
    # a stream too short for the counting pass is refused
    with pytest.raises(RuntimeError):
        ctx.sample_sources_stream(ipop, nthr, izb, st[:, :1000])
    ctx.close()
