"""clr_shell_norms_noalloc against clr_shell_norms, through the C ABI.

The no-allocation variant keeps its partial sums, final sums and packed
tables in one static device array per GPU; it must give the same shell
counts, z sums, bias sums, norm tables and (with the fused lognormal) the same
density grid as the allocating entry point, bit for bit (same kernels, same
fixed reduction order), on two calls in a row -- the second on a smaller grid
-- and refuse an input beyond the array's limit with a clear error.
"""
import numpy as np
import pytest

from conftest import make_run

pytestmark = pytest.mark.gpu


def _ctx(run, n, ntr=None):
    from colore_amd import lib
    ctx = lib.Context(n, run.scalar("l_box"))
    ctx.upload_tables(run.tables())
    npop = int(run.scalar("n_srcs")) if ntr is None else ntr
    for i in range(npop):
        amp, bz, model = run.tracer(0, i % int(run.scalar("n_srcs")))
        ctx.set_tracer(0, i, model, amp, bz)
    return ctx, npop


def _both(run, n, seed):
    """(allocating, no-allocation) results of the same realisation."""
    out = []
    for noalloc in (False, True):
        ctx, npop = _ctx(run, n)
        s2 = ctx.gaussian_fields(seed, run.scalar("r_smooth"), True)
        nz, cnt, zs, sums = ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop, noalloc=noalloc)
        tabs = [ctx.norm_table(0, i) for i in range(npop)]
        out.append((nz, cnt, zs, sums, tabs, ctx.fetch_grid(0)))
        ctx.close()
    return out


def test_noalloc_shell_norms_match_two_calls(workdir):
    from colore_amd import lib
    assert 0 < lib.load_amd().clr_noalloc_max_bytes() < 64 << 20
    for n, seed in ((128, 3003), (64, 17)):   # the second call on a smaller grid
        path, run = make_run("C3", n, workdir, z_max=1.6, with_maps=False)
        (nz, cnt, zs, sums, tabs, g), (nz2, cnt2, zs2, sums2, tabs2, g2) = _both(run, n, seed)
        assert nz == nz2 and nz > 2
        assert np.array_equal(cnt, cnt2) and cnt.sum() > 0
        assert np.array_equal(zs, zs2)
        assert np.array_equal(sums, sums2)
        for (t, n0, nf), (t2, n02, nf2) in zip(tabs, tabs2):
            assert np.array_equal(t, t2) and n0 == n02 and nf == nf2
        assert np.array_equal(g, g2)


def test_noalloc_rejects_beyond_limit(workdir):
    """34 shells x (2 + 8 tracers) = 340 values exceed the 320 of the array."""
    path, run = make_run("C3", 64, workdir, z_max=1.6, with_maps=False)
    ctx, npop = _ctx(run, 64, ntr=8)
    s2 = ctx.gaussian_fields(1, run.scalar("r_smooth"), True)
    with pytest.raises(RuntimeError, match="static scratch"):
        ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop, noalloc=True)
    # the allocating entry point takes it
    ctx.shell_norms(lognormal_fused=True, sigma2=s2, ntracers=npop)
    ctx.close()
