"""Batched ensembles in one captured HIP graph, through the C ABI.

Realisations of the batched-ensemble workload (grids of 128^3 .. 384^3 with
4 Mpc/h cells, one b = 1+z population, a kappa map at the box edge, nside
128) run as one graph after the FFT tile autotuner has picked each size's
tile shapes.  Every realisation must equal a single-context run of the same
stages (clr_gaussian_fields, clr_shell_norms with the fused lognormal,
clr_sample_sources, clr_kappa) bit for bit: sigma^2, norm tables, catalogue
(positions, RSD, base pixels) and kappa map -- after the first launch, after
replays with host work in between, and after a re-capture with new seeds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _contexts(realisations, workdir):
    from colore_amd import inputs, lib
    import os
    runs, ctxs = [], []
    for i, r in enumerate(realisations):
        d = os.path.join(workdir, f"ens_{i}_{r['n_grid']}_{r['seed']}")
        pops = [(("gauss", r["nz_mean"], 0.1, 5e-4), ("1+z",), "exponential")]
        run = lib.Run(inputs.write_config("C4", d, n_grid=r["n_grid"], z_max=r["z_max"], seed=r["seed"],
                                          with_maps=False, populations=pops))
        assert abs(run.scalar("l_box") / r["l_box"] - 1) < 1e-3, (run.scalar("l_box"), r["l_box"])
        ctx = lib.Context(r["n_grid"], run.scalar("l_box"))
        ctx.upload_tables(run.tables())
        amp, bz, model = run.tracer(0, 0)
        ctx.set_tracer(0, 0, model, amp, bz)
        runs.append(run)
        ctxs.append(ctx)
    return runs, ctxs


def _single(run, r, nside):
    """The same realisation through the single calls on a fresh context."""
    from colore_amd import lib
    ctx = lib.Context(r["n_grid"], run.scalar("l_box"))
    ctx.upload_tables(run.tables())
    amp, bz, model = run.tracer(0, 0)
    ctx.set_tracer(0, 0, model, amp, bz)
    s2 = ctx.gaussian_fields(r["seed"], run.scalar("r_smooth"), True)
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=1)
    n = ctx.sample_sources(-1, r["seed"], 2)
    pos, ipix = ctx.fetch_sources(0, n)
    tab = ctx.norm_table(0, 0)
    kap = ctx.kappa(nside, np.array([run.scalar("r_max")]))[0]
    ctx.close()
    return s2, pos, ipix, tab, kap


def _check(ens, ctxs, i, want):
    s2, flag, kap = ens.result(i)
    assert flag == 0
    ws2, wpos, wipix, wtab, wkap = want
    assert s2 == ws2, (i, s2, ws2)
    n = ctxs[i].catalogue_size(0)
    assert n == len(wpos), (i, n, len(wpos))
    pos, ipix = ctxs[i].fetch_sources(0, n)
    assert np.array_equal(pos, wpos) and np.array_equal(ipix, wipix), i
    tab = ctxs[i].norm_table(0, 0)
    assert np.array_equal(tab[0], wtab[0]) and tab[1:] == wtab[1:], i
    assert np.array_equal(kap, wkap), i


def test_ensemble_graph_equals_single_runs(workdir):
    from colore_amd import inputs, lib
    real = inputs.c4_realisations(64, 4000)
    # one realisation of each size plus two more, to keep the test short
    pick, seen = [], set()
    for i, r in enumerate(real):
        if r["n_grid"] not in seen:
            seen.add(r["n_grid"])
            pick.append(i)
    pick += [i for i in range(len(real)) if i not in pick][:2]
    rs = [real[i] for i in pick]
    assert {r["n_grid"] for r in rs} == {128, 192, 256, 384}
    runs, ctxs = _contexts(rs, workdir)
    # the FFT tile autotuner, once per size
    tuned = {}
    for r, ctx in zip(rs, ctxs):
        if r["n_grid"] not in tuned:
            best, ms = ctx.fft_autotune(3)
            assert len(ms) >= 1 and 0 <= best < len(ms) and np.all(ms > 0)
            tuned[r["n_grid"]] = (best, ms)
    print("autotuned", {k: (v[0], [round(float(x), 4) for x in v[1]]) for k, v in tuned.items()}, flush=True)
    nside = 128
    rk = np.array([run.scalar("r_max") for run in runs])
    ens = lib.Ensemble(ctxs, src_cap=4_000_000, nside_kappa=nside, r_kappa=rk, nstreams=4)
    seeds = np.array([r["seed"] for r in rs], dtype=np.uint32)
    rsm = np.array([run.scalar("r_smooth") for run in runs])
    bad, ms = ens.run(seeds, rsm)
    assert bad == 0 and ms > 0
    wants = [_single(run, r, nside) for run, r in zip(runs, rs)]
    assert sum(len(w[1]) > 0 for w in wants) >= 3   # n(z) of some realisations lies past their box
    for i in range(len(rs)):
        _check(ens, ctxs, i, wants[i])
    # replays of the same graph, with host work and other device work between
    for rep in range(2):
        ctxs[0].fetch_grid(0)
        bad, ms2 = ens.run(seeds, rsm)
        assert bad == 0
        for i in range(len(rs)):
            _check(ens, ctxs, i, wants[i])
    # new seeds: the graph is captured again
    seeds2 = seeds + 1000
    bad, _ = ens.run(seeds2, rsm)
    assert bad == 0
    r1 = dict(rs[1], seed=int(seeds2[1]))
    _check(ens, ctxs, 1, _single(runs[1], r1, nside))
    ens.close()
    for c in ctxs:
        c.close()
