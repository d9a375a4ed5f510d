"""Pins of the CPU oracle: published known-answer vectors, library routines
available here (numpy FFT / MT19937, scipy), closed forms and brute force.
Each test is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it."""
import ctypes
import math

import numpy as np
import pytest
from scipy import stats

from oracle_lib import Oracle, load, p


@pytest.fixture(scope="module")
def L():
    return load()


# ---------------------------------------------------------------- RNG ----
def philox(L, ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    L.orc_philox4x32_10(p(c, ctypes.c_uint32), p(k, ctypes.c_uint32), p(o, ctypes.c_uint32))
    return [int(x) for x in o]


def test_philox_known_answers(L):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert philox(L, [0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox(L, [0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox(L, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def mt_stream(seed, n):
    g = np.random.MT19937(0)
    g._legacy_seeding(seed)
    return g.random_raw(n).astype(np.uint32)


def test_reference_mode_stream_matches_mt19937(L):
    """Per-thread MT19937 streams (init_genrand seeding, GSL's mt19937_set)
    consumed (phase, modulus) per mode over each thread's static block of kz
    planes; the k=0 mode draws nothing."""
    N, seed, nth = 8, 1001, 3
    nh = N // 2 + 1
    out = np.zeros(2 * N * N * nh, dtype=np.uint32)
    L.orc_ref_mode_stream(N, seed, nth, p(out, ctypes.c_uint32))
    out = out.reshape(N, N, nh, 2)
    q, rem = divmod(N, nth)
    start = 0
    for t in range(nth):
        cnt = q + (1 if t < rem else 0)
        blk = out[start:start + cnt].reshape(-1, 2)
        if start == 0:
            assert blk[0, 0] == 0 and blk[0, 1] == 0
            blk = blk[1:]
        ref = mt_stream(seed + t, 2 * len(blk)).reshape(-1, 2)
        assert np.array_equal(blk, ref)
        start += cnt


def test_gsl_seed_zero_maps_to_4357(L):
    out = np.zeros(2 * 2 * 2 * 2, dtype=np.uint32)
    L.orc_ref_mode_stream(2, 0, 1, p(out, ctypes.c_uint32))
    ref = mt_stream(4357, 14)
    assert np.array_equal(out.reshape(-1, 2)[1:].ravel(), ref)


# ---------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 16, 24, 48, 64, 96, 128, 384, 7, 11, 1024])
def test_cfft_matches_numpy(L, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    d = np.empty(2 * n)
    d[0::2], d[1::2] = x.real, x.imag
    L.orc_cfft_inv(n, p(d))
    ref = np.fft.ifft(x) * n
    np.testing.assert_allclose(d[0::2] + 1j * d[1::2], ref, atol=1e-12 * n)


def _half(N, seed):
    rng = np.random.default_rng(seed)
    nh = N // 2 + 1
    h = rng.normal(size=(N, N, nh)) + 1j * rng.normal(size=(N, N, nh))
    a = np.empty((N, N, nh, 2))
    a[..., 0], a[..., 1] = h.real, h.imag
    return h, a.ravel().copy()


@pytest.mark.parametrize("N", [4, 6, 8, 12])
def test_c2r_matches_brute_force(L, N):
    """Non-Hermitian kx=0 and kx=N/2 planes included (as the reference draws)."""
    h, a = _half(N, N)
    out = np.zeros(N * N * 2 * (N // 2 + 1))
    ref = np.zeros_like(out)
    L.orc_c2r_3d(N, p(a), p(out))
    L.orc_c2r_3d_brute(N, p(a), p(ref))
    np.testing.assert_allclose(out, ref, atol=1e-10 * N ** 3)


@pytest.mark.parametrize("N", [16, 24, 32, 48])
def test_c2r_matches_numpy_irfftn(L, N):
    h, a = _half(N, 100 + N)
    out = np.zeros(N * N * 2 * (N // 2 + 1))
    L.orc_c2r_3d(N, p(a), p(out))
    ref = np.fft.irfftn(h, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    got = out.reshape(N, N, -1)[:, :, :N]
    np.testing.assert_allclose(got, ref, atol=1e-9 * np.abs(ref).max())
    # the imaginary parts of the kx=0 and kx=N/2 terms after the (z,y) transforms are ignored
    h2 = h.copy()
    h2[:, :, 0] += 1j * 3.0 * (np.arange(N)[:, None] == 0) * (np.arange(N)[None, :] == 0)
    a2 = np.stack([h2.real, h2.imag], -1).ravel().copy()
    out2 = np.zeros_like(out)
    L.orc_c2r_3d(N, p(a2), p(out2))
    np.testing.assert_allclose(out2, out, atol=1e-9 * np.abs(ref).max())


# -------------------------------------------------------- modes / fields ----
def _tables(nk=200):
    """Small synthetic tables: r grid, z = r/3000, D = 1/(1+z), P(k) power law."""
    na = 5001
    r = np.arange(na) * (10000.0 / (na - 1))
    z = r / 3000.0
    ta = dict(r=r, z=z, d1=1 / (1 + z), v1=np.ones(na), pd=np.full(na, 0.5), ih=np.full(na, 3000.0),
              glob_idr=(na - 1) / 10000.0, ns=0.96, fgrowth_0=0.5, hubble_0=1 / 2997.92458, omega_m=0.3,
              prefac_lensing=1.5 * 0.3 / 2997.92458 ** 2, r_max=800.0)
    lk = np.linspace(-4, 2, nk)
    ta.update(logk=lk, pk=1e4 * 10 ** (-1.5 * lk), logkmin=-4.0, logkmax=2.0, idlogk=(nk - 1) / 6.0, numk=nk)
    return ta


def test_modes_closed_form_and_injection():
    ta = _tables()
    orc = Oracle(ta)
    N, Lb, rs = 8, 500.0, 5.0
    nh = N // 2 + 1
    rng = np.random.default_rng(3)
    inj = rng.integers(0, 2 ** 32, size=2 * N * N * nh, dtype=np.uint64).astype(np.uint32)
    d, ph = orc.modes(N, Lb, 1, rs, True, injected=inj)
    d = d.reshape(N, N, nh, 2)
    ph = ph.reshape(N, N, nh, 2)
    dk = 2 * math.pi / Lb
    assert np.all(d[0, 0, 0] == 0) and np.all(ph[0, 0, 0] == 0)
    for (i, j, k) in [(1, 0, 0), (0, 3, 2), (5, 7, 4), (4, 4, 4), (7, 1, 3)]:
        kz = (i if 2 * i <= N else i - N) * dk
        ky = (j if 2 * j <= N else j - N) * dk
        kx = k * dk
        k2 = kx * kx + ky * ky + kz * kz
        lgk = 0.5 * math.log10(k2)
        ik = int((lgk + 4) * ta["idlogk"])
        pk = ta["pk"][ik] + (lgk - ta["logk"][ik]) * (ta["pk"][ik + 1] - ta["pk"][ik]) * ta["idlogk"]
        s2 = pk / dk ** 3
        m = k + nh * (j + N * i)
        u1, u2 = inj[2 * m] / 2 ** 32, inj[2 * m + 1] / 2 ** 32
        amp = math.sqrt(-s2 * math.log(1 - u2))
        sm = math.exp(-0.5 * rs * rs * k2)
        want = amp * np.exp(2j * math.pi * u1) * sm
        got = d[i, j, k, 0] + 1j * d[i, j, k, 1]
        assert abs(got - want) <= 1e-12 * abs(want)
        wphi = -ta["prefac_lensing"] * amp * np.exp(2j * math.pi * u1) / k2 * sm
        gphi = ph[i, j, k, 0] + 1j * ph[i, j, k, 1]
        assert abs(gphi - wphi) <= 1e-12 * abs(wphi)


def test_mode_power_statistics():
    """<|delta_k|^2> = P(k)/dk^3 and uniform phases with the Philox stream."""
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 32, 800.0
    d, _ = orc.modes(N, Lb, 42, 0.0, False)
    d = d.reshape(N, N, N // 2 + 1, 2)
    c = d[..., 0] + 1j * d[..., 1]
    dk = 2 * math.pi / Lb
    f = np.fft.fftfreq(N, 1.0 / N)
    kx = np.arange(N // 2 + 1)
    kk = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + kx[None, None, :] ** 2) * dk
    kk[0, 0, 0] = 1
    pk = np.interp(np.log10(kk), ta["logk"], ta["pk"])
    ratio = (np.abs(c) ** 2 / (pk / dk ** 3))[kk > 2 * dk]
    assert abs(ratio.mean() - 1) < 5 / math.sqrt(ratio.size)
    ph = np.angle(c[kk > 2 * dk])
    assert stats.kstest((ph / (2 * math.pi)) % 1.0, "uniform").pvalue > 1e-3


def test_normalize_and_sigma2(L):
    N, Lb = 8, 100.0
    rng = np.random.default_rng(0)
    g = rng.normal(size=N * N * 2 * (N // 2 + 1))
    g2 = g.copy()
    out = np.zeros(2)
    s2 = L.orc_normalize_sigma2(N, Lb, p(g2), None, p(out))
    norm = (math.sqrt(2 * math.pi) / Lb) ** 3
    cells = (g * norm).reshape(N, N, -1)[:, :, :N]
    assert abs(s2 - cells.var()) < 1e-12 * cells.var()
    np.testing.assert_allclose(g2, g * norm, rtol=1e-15)


def test_lognormal_and_clip_closed_form():
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 8, 300.0
    rng = np.random.default_rng(1)
    g = rng.normal(size=(N, N, 2 * (N // 2 + 1))) * 0.5
    s2 = 0.3
    ln = orc.density(N, Lb, g, s2, 0)
    cl = orc.density(N, Lb, g, s2, 3)
    dx = Lb / N
    for (iz, iy, ix) in [(0, 0, 0), (4, 4, 4), (7, 2, 5), (3, 6, 1)]:
        r = math.sqrt(((ix * dx - Lb / 2) ** 2 + (iy * dx - Lb / 2) ** 2 + (iz * dx - Lb / 2) ** 2))
        D = np.interp(r, ta["r"], ta["d1"]) if r > 0 else 1.0
        dg = g[iz, iy, ix]
        assert abs(ln[iz, iy, ix] - (math.exp(D * (dg - 0.5 * D * s2)) - 1)) < 1e-12
        assert abs(cl[iz, iy, ix] - (max(1 + D * dg, 0) - 1)) < 1e-12


def test_bias_models_closed_form(L):
    bm = L.orc_bias_model
    for d in (-1.0, -1.5):
        for m in range(3):
            assert bm(d, 1.3, m) == 0.0
    assert abs(bm(0.5, 2.0, 0) - 2.25) < 1e-15
    assert abs(bm(-0.5, 2.0, 1) - math.exp(2.0 * -0.5 / 0.5)) < 1e-15
    assert abs(bm(0.5, 2.0, 1) - 2.0) < 1e-15
    assert bm(-0.6, 2.0, 2) == 0.0 and abs(bm(-0.4, 2.0, 2) - 0.2) < 1e-15
    # lognormal moment: <(1+delta)^b> = exp(b(b-1) s^2/2) for delta = exp(G - s^2/2) - 1
    rng = np.random.default_rng(5)
    s = 0.4
    G = rng.normal(0, s, 400000)
    dd = np.exp(G - s * s / 2) - 1
    for b in (0.5, 1.0, 2.0):
        m = np.mean([bm(x, b, 0) for x in dd[:20000]])
        assert abs(m / math.exp(b * (b - 1) * s * s / 2) - 1) < 0.02


def test_shell_sums_invariants():
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 16, 1200.0
    zero = np.zeros((N, N, 2 * (N // 2 + 1)))
    b = np.full(5001, 1.7)
    nz, zmax, cnt, zs, sums = orc.shell_sums(N, Lb, zero, [(b, 0), (b, 1), (b, 2)])
    assert nz == int(zmax / 0.05) + 2
    # delta = 0: every bias model gives 1 -> the sums equal the counts, norm = 1
    for t in range(3):
        np.testing.assert_array_equal(sums[t], cnt.astype(float))
    # count: all cells whose shell index is in range
    dx = Lb / N
    ii = np.arange(N) * dx - Lb / 2
    r = np.sqrt(ii[:, None, None] ** 2 + ii[None, :, None] ** 2 + ii[None, None, :] ** 2)
    z = np.interp(r, ta["r"], ta["z"])
    ind = (z * ((nz - 2) / zmax)).astype(int) + 1
    want = np.bincount(ind[ind < nz], minlength=nz)
    np.testing.assert_array_equal(cnt, want)
    idz = (nz - 2) / zmax
    for s in range(1, nz):
        if cnt[s]:
            zm = zs[s] / cnt[s]
            assert (s - 1) / idz - 1e-12 <= zm < s / idz + 1e-12
    tab, n0, nf = orc.norm_table(nz, zmax, cnt, zs, sums[0])
    np.testing.assert_allclose(tab, 1.0, rtol=1e-15)
    # b = 0 makes the power-law bias 1 for any delta
    rng = np.random.default_rng(2)
    g = np.clip(rng.normal(size=zero.shape) * 0.3, -0.95, None)
    b0 = np.zeros(5001)
    _, _, cnt2, _, sums2 = orc.shell_sums(N, Lb, g, [(b0, 0)])
    np.testing.assert_array_equal(sums2[0], cnt2.astype(float))
    # cells with delta <= -1 contribute nothing (bias_model's first branch)
    g[:, :, :N] = -1.0
    _, _, cnt3, _, sums3 = orc.shell_sums(N, Lb, g, [(b0, 0), (b, 1), (b, 2)])
    assert cnt3.sum() > 0 and np.all(sums3 == 0)


def test_norm_table_interpolates_shell_means():
    ta = _tables()
    orc = Oracle(ta)
    nz, zmax = 6, 0.2
    cnt = np.array([0, 10, 20, 30, 40, 50], dtype=np.uint64)
    zs = np.array([0, 0.2, 1.5, 3.75, 6.8, 9.5])
    sums = np.array([0, 8.0, 25.0, 27.0, 44.0, 40.0])
    tab, n0, nf = orc.norm_table(nz, zmax, cnt, zs, sums)
    zarr = zs / np.maximum(cnt, 1)
    nrm = np.where(cnt > 0, cnt / np.where(sums > 0, sums, 1), 0)
    zarr[0], zarr[-1] = 0, zmax
    nrm[0], nrm[-1] = nrm[1], nrm[-2]
    z = np.interp(ta["r"], ta["r"], ta["z"])
    want = np.interp(z, zarr, nrm)
    want[z >= zmax] = nrm[-1]
    assert n0 == nrm[0] and nf == nrm[-1]
    np.testing.assert_allclose(tab, want, rtol=1e-12)


# ------------------------------------------------------------- Poisson ----
@pytest.mark.parametrize("lam", [0.05, 1.0, 9.2, 37.0, 420.0, 1300.0])
def test_poisson_distribution(L, lam):
    n = 60000
    k = np.array([L.orc_poisson(lam, c, 0x100, 77) for c in range(n)])
    assert abs(k.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(k.var() / lam - 1) < 6 * math.sqrt(2.0 / n) + 0.02
    if lam < 50:
        vals = np.arange(0, int(lam + 8 * math.sqrt(lam) + 5))
        obs = np.array([(k == v).sum() for v in vals])
        exp = stats.poisson.pmf(vals, lam) * n
        m = exp > 5
        chi2 = ((obs[m] - exp[m]) ** 2 / exp[m]).sum()
        assert stats.chi2.sf(chi2, m.sum() - 1) > 1e-4
    assert L.orc_poisson(0.0, 5, 0x100, 1) == 0 and L.orc_poisson(float("nan"), 5, 0x100, 1) == 0


def test_source_positions_inside_cells():
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 16, 400.0
    g = np.zeros((N, N, 2 * (N // 2 + 1)))
    amp = np.full(5001, 2e-3)
    pop = orc.pop(N, Lb, amp, np.ones(5001), np.ones(5001), 1.0, 1.0, 0, 0, 9)
    pos, ipix = orc.sources(pop, g, g, 2)
    dx = Lb / N
    assert len(pos) > 1000
    cell = np.floor((pos[:, :3] + Lb / 2) / dx + 0.5)
    off = pos[:, :3] + Lb / 2 - cell * dx
    assert np.all(off >= -dx / 2 - 1e-4) and np.all(off < dx / 2 + 1e-4)
    L_ = load()
    for i in range(0, len(pos), 97):
        v = pos[i, :3].astype(np.float64)
        assert ipix[i] == L_.orc_ring2nest(2, L_.orc_vec2pix_ring(2, p(v)))
    # z-major order
    lin = cell[:, 0] + N * (cell[:, 1] + N * cell[:, 2])
    assert np.all(np.diff(lin) >= 0)


# ------------------------------------------------------------- HEALPix ----
def test_healpix_nside1_centres(L):
    v = np.zeros(3)
    for f in range(12):
        L.orc_pix2vec_nest(1, f, p(v))
        z = 2 / 3 if f < 4 else (0.0 if f < 8 else -2 / 3)
        phi = (math.pi / 4 + (f % 4) * math.pi / 2) if f < 4 or f >= 8 else (f - 4) * math.pi / 2
        assert abs(v[2] - z) < 1e-14
        assert abs(math.atan2(v[1], v[0]) % (2 * math.pi) - phi % (2 * math.pi)) < 1e-12


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64])
def test_healpix_bijections_and_roundtrip(L, nside):
    npix = 12 * nside * nside
    ring = np.array([L.orc_nest2ring(nside, i) for i in range(npix)])
    assert sorted(ring) == list(range(npix))
    assert all(L.orc_ring2nest(nside, int(r)) == i for i, r in enumerate(ring))
    v = np.zeros(3)
    for i in range(0, npix, max(1, npix // 500)):
        L.orc_pix2vec_nest(nside, i, p(v))
        assert L.orc_ring2nest(nside, L.orc_vec2pix_ring(nside, p(v))) == i


def test_healpix_equal_area(L):
    rng = np.random.default_rng(4)
    nside, n = 4, 192000
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    pix = np.array([L.orc_ang2pix_ring(nside, a, b) for a, b in zip(z, ph)])
    c = np.bincount(pix, minlength=12 * nside * nside)
    e = n / (12 * nside * nside)
    assert stats.chisquare(c).pvalue > 1e-4 and abs(c.mean() - e) < 1e-9


# --------------------------------------------------- lightcone painters ----
def test_kappa_closed_form():
    """phi = A/2 |x|^2 and D(1+z) = 1: the transverse trace of the tidal
    tensor is 2A, so kappa = 2A sum (rm - rm^2/r_s) dr over the march."""
    ta = _tables()
    ta["z"] = np.zeros(5001)
    ta["d1"] = np.ones(5001)
    orc = Oracle(ta)
    N, Lb = 32, 1000.0
    ta["r_max"] = Lb / (2 * (1 + 2.0 / N))
    orc = Oracle(ta)
    dx = Lb / N
    i = (np.arange(N) - N // 2) * dx
    A = 3e-7
    phi = np.zeros((N, N, 2 * (N // 2 + 1)))
    phi[:, :, :N] = 0.5 * A * (i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2)
    r_s = 0.7 * ta["r_max"]
    out = orc.kappa_isw(0, N, Lb, phi, 2, [r_s])
    nr = N // 2
    dr = ta["r_max"] / nr
    ih = int(r_s / dr + 0.5)
    rm = (np.arange(min(ih, nr - 1) + 1) + 0.5) * dr
    want = 2 * A * np.sum((rm - rm * rm / (ih * dr)) * dr)
    np.testing.assert_allclose(out[0], want, rtol=1e-9)
    # ISW of a constant potential: 2 phi sum PD dr
    phic = np.full_like(phi, 0.25)
    iw = orc.kappa_isw(1, N, Lb, phic, 2, [r_s])
    want = 0.25 * 2 * 0.5 * dr * (min(ih, nr - 1) + 1)
    np.testing.assert_allclose(iw[0], want, rtol=1e-12)


def _beam_positions(rng, n, rlo, rhi):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    u = u[np.abs(u[:, 2]) < 0.95][: n // 2]
    r = rng.uniform(rlo, rhi, len(u))
    pos = np.zeros((len(u), 4), dtype=np.float32)
    pos[:, :3] = u * r[:, None]
    return pos


def test_source_beams_closed_form():
    """phi = 1/2 x.M.x + g.x about the observer, D(1+z) = 1 (to 1e-7) and
    z = r/3000.  The CIC central difference of phi at a source is 2 dx (M x +
    g) exactly, so dz = (x.M.u + g.u) V1(r) (-f0 / (1.5 H0 Om)).  The tidal
    tensor is M dx^2 exactly, so kappa, e1 and e2 are (M_tt + M_pp),
    (M_tt - M_pp) and 2 M_tp, the tensor projected on the unit vectors
    e_theta, e_phi, times S = sum dr (rm - rm^2/r) over the march; with M = 0
    the deflections are (e_theta.g, e_phi.g) times sum 2 dr (1 - rm/r), in
    degrees."""
    ta = _tables()
    N, Lb = 32, 1000.0
    dx = Lb / N
    ta["r_max"] = Lb / (2 * (1 + 2.0 / N))
    ta["v1"] = 1 + ta["r"] / 1000.0
    orc = Oracle(ta)
    na = len(ta["r"])
    a = np.arange(na) / (na - 1.0)
    r_a2r = np.empty(na)
    r_a2r[1:] = 3000.0 * (1 / a[1:] - 1)
    r_a2r[0] = 1e9
    rng = np.random.default_rng(21)
    M = rng.normal(size=(3, 3)) * 1e-7
    M = M + M.T
    g = rng.normal(size=3) * 1e-5
    i = (np.arange(N) - N // 2) * dx

    def field(M, g):
        X = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1)   # [z][y][x] -> (z, y, x) coordinates
        X = X[..., ::-1]                                          # (x, y, z)
        phi = np.zeros((N, N, 2 * (N // 2 + 1)))
        phi[:, :, :N] = 0.5 * np.einsum("...i,ij,...j->...", X, M, X) + X @ g
        return phi

    pos = _beam_positions(rng, 400, 3 * dx, 12 * dx)
    x = pos[:, :3].astype(np.float64)
    r = np.linalg.norm(x, axis=1)
    u = x / r[:, None]
    fv = -ta["fgrowth_0"] / (1.5 * ta["hubble_0"] * ta["omega_m"])
    dz, lens = orc.src_beams(N, Lb, field(M, g), pos, r_a2r, lensing=True)
    want = (np.einsum("ni,ij,nj->n", x, M, u) + u @ g) * (1 + r / 1000.0) * fv
    np.testing.assert_allclose(dz, want, rtol=2e-6, atol=1e-6 * np.abs(want).max())
    cth = u[:, 2]
    sth = np.sqrt(1 - cth ** 2)
    cph, sph = u[:, 0] / sth, u[:, 1] / sth
    et = np.stack([cth * cph, cth * sph, -sth], 1)
    ep = np.stack([-sph, cph, np.zeros_like(sph)], 1)
    Mtt, Mpp, Mtp = (np.einsum("ni,ij,nj->n", e, M, f) for e, f in ((et, et), (ep, ep), (et, ep)))
    nr = N // 2
    dr = ta["r_max"] / nr
    S, D = np.zeros(len(r)), np.zeros(len(r))
    for k in range(len(r)):
        rm = (np.arange(min(int(r[k] / dr + 0.5), nr - 1) + 1) + 0.5) * dr
        S[k] = np.sum(dr * (rm - rm * rm / r[k]))
        D[k] = np.sum(2 * dr * (1 - rm / r[k]))
    for col, w in ((0, (Mtt - Mpp) * S), (1, 2 * Mtp * S), (2, (Mtt + Mpp) * S)):
        np.testing.assert_allclose(lens[:, col], w, rtol=1e-6, atol=1e-6 * np.abs(w).max())
    _, lens0 = orc.src_beams(N, Lb, field(np.zeros((3, 3)), g), pos, r_a2r, lensing=True)
    RTOD = 57.2957795
    np.testing.assert_allclose(lens0[:, 3], (ep @ g) * D * RTOD, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(lens0[:, 4], (et @ g) * D * RTOD, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(lens0[:, :3], 0, atol=1e-12)


def test_source_skewers_closed_form():
    """Linear fields about the observer, phi = g.x and ln(1+delta) = a + b.x,
    D = 1/2 (1 at r = 0, as get_bg returns there): the CIC reads are exact, so along the skewer of a source at r,
    step j (rm = (j+1/2) dr, j <= min(int(r/dr+1/2), nr-1)) holds
    delta = exp(a + b.u rm) - 1 to CIC accuracy, the Gaussian field
    2 (a + b.u rm) + sigma2/4, and the velocity (g.u) V1(rm) (-f0/(1.5 H0 Om));
    later steps are 0."""
    ta = _tables()
    N, Lb = 32, 1000.0
    dx = Lb / N
    ta["r_max"] = Lb / (2 * (1 + 2.0 / N))
    ta["v1"] = 1 + ta["r"] / 1000.0
    ta["d1"] = np.full_like(ta["r"], 0.5)
    orc = Oracle(ta)
    rng = np.random.default_rng(5)
    g = rng.normal(size=3) * 1e-5
    a, b = 0.1, rng.normal(size=3) * 2e-3
    i = (np.arange(N) - N // 2) * dx
    X = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1)[..., ::-1]
    phi = np.zeros((N, N, 2 * (N // 2 + 1)))
    phi[:, :, :N] = X @ g
    dens = np.zeros_like(phi)
    dens[:, :, :N] = np.expm1(a + X @ b)
    dens = dens.astype(np.float32).astype(np.float64)
    pos = _beam_positions(rng, 200, 3 * dx, 12 * dx)
    x = pos[:, :3].astype(np.float64)
    r = np.linalg.norm(x, axis=1)
    u = x / r[:, None]
    fv = -ta["fgrowth_0"] / (1.5 * ta["hubble_0"] * ta["omega_m"])
    nr = N // 2
    dr = ta["r_max"] / nr
    rm = (np.arange(nr) + 0.5) * dr
    live = np.arange(nr)[None, :] <= np.minimum((r / dr + 0.5).astype(int), nr - 1)[:, None]
    lin = a + (u @ b)[:, None] * rm[None, :]
    s2 = 0.37
    dg, vs = orc.src_skewers(N, Lb, dens, phi, pos, gaussian=True, sigma2=s2)
    want_g = np.where(live, 2 * lin + 0.25 * s2, 0)
    # get_bg(r <= 0, BG_D1) is 1, not the table: skip the steps whose CIC
    # corners include the observer's cell
    far = rm > 2 * dx
    np.testing.assert_allclose(dg[:, far], want_g[:, far], rtol=0, atol=1e-6)
    assert np.all(np.abs(dg[:, 0] - want_g[:, 0])[live[:, 0]] > 1e-4)
    want_v = np.where(live, (u @ g)[:, None] * (1 + rm / 1000.0)[None, :] * fv, 0)
    np.testing.assert_allclose(vs, want_v, rtol=1e-6, atol=1e-9 * np.abs(want_v).max())
    dd, vs2 = orc.src_skewers(N, Lb, dens, phi, pos, gaussian=False)
    np.testing.assert_array_equal(vs2, vs)
    # CIC of exp(linear): within the trilinear error (dx |b|)^2 / 8 of exp
    want_d = np.where(live, np.expm1(lin), 0)
    np.testing.assert_allclose(dd, want_d, rtol=0, atol=3 * (dx * np.abs(b).max()) ** 2 / 8 * np.exp(lin.max()))
    assert np.abs(dd - want_d).max() > 0   # nonlinear field: CIC is not exact here


def test_custom_map_closed_form(L):
    """delta = c.x about the observer, the truncated model with b = 1.5 and
    norm = 1: CIC is exact on linear fields, so the map at pixel u is
    b (c.u) sum K(rm)/IH rm dr over the steps where |K/IH| exceeds 1e-4 of its
    maximum (np.interp of the tables; IH = 3000)."""
    ta = _tables()
    N, Lb = 32, 1000.0
    ta["r_max"] = Lb / (2 * (1 + 2.0 / N))
    orc = Oracle(ta)
    dx = Lb / N
    rng = np.random.default_rng(8)
    c = rng.normal(size=3) * 3e-4
    i = (np.arange(N) - N // 2) * dx
    X = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1)[..., ::-1]
    dens = np.zeros((N, N, 2 * (N // 2 + 1)))
    dens[:, :, :N] = X @ c
    r = ta["r"]
    kz = np.exp(-((r - 250.0) / 40.0) ** 2)
    bz = np.full_like(r, 1.5)
    norm = np.ones_like(r)
    nside = 4
    out = orc.cstm(N, Lb, dens, kz, bz, norm, 1.0, 1.0, 2, nside)
    nr = N // 2
    dr = ta["r_max"] / nr
    rm = (np.arange(nr) + 0.5) * dr
    k = np.interp(rm, r, kz) / 3000.0
    on = np.nonzero(np.abs(k) > 1e-4 * np.abs(k).max())[0]
    sel = np.arange(on[0], on[-1] + 1)
    u = np.zeros((12 * nside * nside, 3))
    for ip in range(len(u)):
        L.orc_pix2vec_nest(nside, ip, p(u[ip]))
    want = 1.5 * (u @ c) * np.sum(k[sel] * rm[sel]) * dr
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-6 * np.abs(want).max())
    # a kernel that is 0 everywhere keeps the default edges (the whole march) and gives 0
    assert np.all(orc.cstm(N, Lb, dens, np.zeros_like(r), bz, norm, 1.0, 1.0, 2, nside) == 0)


def test_r2c_matches_numpy(L):
    N = 12
    rng = np.random.default_rng(4)
    g = np.zeros((N, N, 2 * (N // 2 + 1)))
    g[:, :, :N] = rng.normal(size=(N, N, N))
    out = np.zeros(2 * N * N * (N // 2 + 1))
    L.orc_r2c_3d(N, p(g), p(out))
    ref = np.fft.rfftn(g[:, :, :N])
    np.testing.assert_allclose(out[0::2] + 1j * out[1::2], ref.ravel(), atol=1e-10)


def _deposit_numpy(N, pos, L, interp):
    """NGP / CIC / TSC assignment of unit masses at pos (float32 [n, 3]) from
    the kernels' definitions, with np.add.at."""
    x = pos.astype(np.float32) * np.float32(N / L)
    rho = np.zeros((N, N, N))
    if interp == 0:
        i = np.floor(x.astype(np.float64) + 0.5).astype(int) % N
        np.add.at(rho, (i[:, 2], i[:, 1], i[:, 0]), 1.0)
        return rho
    if interp == 1:
        i0 = np.floor(x).astype(int)
        t = (x - i0).astype(np.float64)
        offs, w = [0, 1], [1 - t, t]
    else:
        i0 = np.floor(x.astype(np.float64) + 0.5).astype(int)
        t = (x - i0).astype(np.float64)
        offs, w = [-1, 0, 1], [0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2]
    for a, oz in enumerate(offs):
        for b, oy in enumerate(offs):
            for c, ox in enumerate(offs):
                np.add.at(rho, ((i0[:, 2] + oz) % N, (i0[:, 1] + oy) % N, (i0[:, 0] + ox) % N),
                          w[c][:, 0] * w[b][:, 1] * w[a][:, 2])
    return rho


@pytest.mark.parametrize("interp", [0, 1, 2])
def test_lpt1_closed_form(interp):
    """delta_G = A cos(2 pi m x / L) and D = 1: the 1LPT displacement is
    psi_x = -(A L / (2 pi m)) sin(2 pi m x / L), psi_y = psi_z = 0, so the
    particle of cell i sits at i dx + psi; with A = 0 every scheme returns
    delta = 0 exactly; with A != 0 the density is the NGP/CIC/TSC assignment of
    those positions (np.add.at from the kernels' definitions) minus 1, and the
    mass sums to N^3 (to the float weights' rounding)."""
    ta = _tables()
    ta["d1"] = np.ones_like(ta["d1"])
    orc = Oracle(ta)
    N, Lb, m = 16, 400.0, 2
    dx = Lb / N
    g = np.zeros((N, N, 2 * (N // 2 + 1)))
    zero, _ = orc.lpt1(N, Lb, g, interp)
    assert np.all(zero == 0)
    A = 1.0
    ix = np.arange(N)
    g[:, :, :N] = A * np.cos(2 * np.pi * m * ix / N)[None, None, :]
    d, pos = orc.lpt1(N, Lb, g, interp, with_pos=True)
    k0 = 2 * np.pi * m / Lb
    want_x = ix * dx - (A / k0) * np.sin(k0 * ix * dx)
    np.testing.assert_allclose(pos[:, :, :, 0], np.broadcast_to(want_x % Lb, (N, N, N)), atol=2e-4 * dx)
    iz, iy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    np.testing.assert_allclose(pos[:, :, :, 1], np.broadcast_to((iy * dx)[:, :, None], (N, N, N)), atol=1e-4)
    np.testing.assert_allclose(pos[:, :, :, 2], np.broadcast_to((iz * dx)[:, :, None], (N, N, N)), atol=1e-4)
    assert abs(d[:, :, :N].sum()) < 1e-6 * N ** 3   # float weights
    rho = _deposit_numpy(N, pos.reshape(-1, 3), Lb, interp)
    np.testing.assert_allclose(d[:, :, :N], rho - 1, atol=1e-6)
    assert np.abs(d[:, :, :N]).max() > 0.05


@pytest.mark.parametrize("interp", [1, 2])
def test_lpt2_closed_form(interp):
    """delta_G = A cos(k1 x) + B cos(k2 y), D1 = 1, D2 = c: the digradients
    are -A cos(k1 x) (xx) and -B cos(k2 y) (yy), the second-order source is
    A B cos(k1 x) cos(k2 y), and psi2 = -grad S / (k1^2 + k2^2), so the
    particle of cell (x, y, z) sits at x - (A/k1) sin(k1 x) + c k1 A B
    sin(k1 x) cos(k2 y) / (k1^2 + k2^2), y - (B/k2) sin(k2 y) + c k2 A B
    cos(k1 x) sin(k2 y) / (k1^2 + k2^2), z; the deposit of those positions is
    the density.  A single plane wave has S = 0: 2LPT = 1LPT."""
    ta = _tables()
    ta["d1"] = np.ones_like(ta["d1"])
    orc = Oracle(ta)
    N, Lb = 16, 400.0
    dx = Lb / N
    c2 = -0.4
    d2 = np.full_like(ta["r"], c2)
    A, B, m1, m2 = 0.8, 0.6, 1, 2
    k1, k2 = 2 * np.pi * m1 / Lb, 2 * np.pi * m2 / Lb
    i = np.arange(N) * dx
    Z, Y, X = np.meshgrid(i, i, i, indexing="ij")
    g = np.zeros((N, N, 2 * (N // 2 + 1)))
    g[:, :, :N] = A * np.cos(k1 * X) + B * np.cos(k2 * Y)
    d, pos = orc.lpt1(N, Lb, g, interp, with_pos=True, order=2, d2=d2)
    kk = k1 * k1 + k2 * k2
    wx = X - (A / k1) * np.sin(k1 * X) + c2 * k1 * A * B * np.sin(k1 * X) * np.cos(k2 * Y) / kk
    wy = Y - (B / k2) * np.sin(k2 * Y) + c2 * k2 * A * B * np.cos(k1 * X) * np.sin(k2 * Y) / kk
    np.testing.assert_allclose(pos[..., 0], wx % Lb, atol=5e-4 * dx)
    np.testing.assert_allclose(pos[..., 1], wy % Lb, atol=5e-4 * dx)
    np.testing.assert_allclose(pos[..., 2], Z, atol=1e-4)
    rho = _deposit_numpy(N, pos.reshape(-1, 3), Lb, interp)
    np.testing.assert_allclose(d[:, :, :N], rho - 1, atol=1e-6)
    g1 = np.zeros_like(g)
    g1[:, :, :N] = A * np.cos(k1 * X)
    d1o, p1 = orc.lpt1(N, Lb, g1, interp, with_pos=True)
    d2o, p2 = orc.lpt1(N, Lb, g1, interp, with_pos=True, order=2, d2=d2)
    np.testing.assert_allclose(p2, p1, atol=1e-4)


def test_imap_constant_temperature():
    """T = const, b = 0, norm = 1, no RSD: every painted pixel averages to T."""
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 16, 600.0
    g = np.random.default_rng(1).normal(size=(N, N, 2 * (N // 2 + 1))) * 0.2
    zero = np.zeros_like(g)
    T = 0.37
    r0 = np.array([100.0, 160.0, 230.0])
    rf = np.array([160.0, 230.0, 280.0])
    data, nadd = orc.imap(N, Lb, g, zero, np.full(5001, T), np.zeros(5001), np.ones(5001), 1.0, 1.0, 0, 4, r0, rf)
    hit = nadd > 0
    assert hit.sum() > 100
    np.testing.assert_allclose(data[hit] / nadd[hit], T, rtol=1e-12)


@pytest.mark.parametrize("N,inj", [(12, False), (16, True), (32, False)])
def test_kspace_variance_matches_grid(N, inj):
    """Parseval of the c2r with its kx = 0, N/2 planes: the variances the
    oracle computes from the modes alone equal the normalised grids' mean
    squares (the c2r itself is pinned against numpy's irfftn), for the Philox
    stream and for an injected stream whose kx = 0 and N/2 planes are not
    Hermitian.  Dropping the conjugate partner on those planes changes the
    result by O(1)."""
    ta = _tables()
    orc = Oracle(ta)
    Lb, rs = 700.0, 3.0
    nh = N // 2 + 1
    injected = None
    if inj:
        rng = np.random.default_rng(11)
        injected = rng.integers(0, 2 ** 32, size=2 * N * N * nh, dtype=np.uint64).astype(np.uint32)
    d, ph, s2 = orc.fields(N, Lb, 9, rs, True, injected=injected)
    vd, vp = orc.kspace_variance(N, Lb, 9, rs, True, injected=injected)
    md = np.mean(d[:, :, :N] ** 2)
    mp = np.mean(ph[:, :, :N] ** 2)
    assert abs(vd / md - 1) < 1e-12 and abs(vp / mp - 1) < 1e-12, (vd, md, vp, mp)
    # the k = 0 mode is zero, so the mean vanishes and sigma^2 = <delta^2>
    assert abs(s2 / md - 1) < 1e-9


# ------------------------------------------- lambda, RSD, shell sums, imap ----
def _bias_np(d, b, model):
    """bias_model (common.h:414-431) written from its definition."""
    d = np.asarray(d, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    alive = d > -1
    with np.errstate(all="ignore"):
        if model == 0:
            v = np.power(np.where(alive, 1 + d, 1.0), b)
        elif model == 1:
            v = np.where(d < 0, np.exp(b * d / np.where(alive, 1 + d, 1.0)), 1 + b * d)
        else:
            v = np.maximum(1 + b * d, 0.0)
    return np.where(alive, v, 0.0)


def _radii(N, Lb):
    dx = Lb / N
    c = np.arange(N) * dx - Lb / 2
    return np.sqrt(c[None, None, :] ** 2 + c[None, :, None] ** 2 + c[:, None, None] ** 2), c


@pytest.mark.parametrize("model", [0, 1, 2])
def test_cell_lambda_and_shell_counts(model):
    """lambda = n(r) dx^3 bias_model(delta, b(r)) norm(r) (srcs.c:158-176)
    against numpy on every cell, with n, b and norm all varying with r; then
    the counts per radial shell against the summed lambda within 5 Poisson
    sigma.  A dx^2, a missing norm, or b read at the wrong radius fails the
    first check."""
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 32, 1200.0
    r_t = ta["r"]
    amp = 6e-4 * (1 + r_t / 1000.0)
    bz = 0.5 + r_t / 1500.0
    norm = 0.8 + 0.3 * np.sin(r_t / 300.0)
    rng = np.random.default_rng(model + 20)
    g = np.exp(rng.normal(0, 0.6, size=(N, N, 2 * (N // 2 + 1)))) - 1.1
    pop = orc.pop(N, Lb, amp, bz, norm, float(norm[0]), float(norm[-1]), model, 0, 4242)
    r, _ = _radii(N, Lb)
    dx = Lb / N
    want = np.interp(r, r_t, amp) * dx ** 3 * _bias_np(g[:, :, :N], np.interp(r, r_t, bz), model) * \
        np.interp(r, r_t, norm)
    want[~(r < Lb / 2 + 20)] = 0
    want[r == 0] = 0   # get_bg at r <= 0 returns f0, which is 0 for n(r) (cosmo.c:30-86)
    gd = np.ascontiguousarray(g)
    got = np.array([[[orc.lib.orc_cell_lambda(ctypes.byref(pop), p(gd), ix, iy, iz) for ix in range(N)]
                     for iy in range(N)] for iz in range(N)])
    assert np.all(np.abs(got - want) <= 1e-12 * np.maximum(np.abs(want), 1e-300))
    assert np.count_nonzero(want == 0) > 0 and want.max() > 1
    tot, counts = orc.counts(pop, gd)
    counts = counts.reshape(N, N, N)
    edges = np.linspace(0, r.max() + 1, 9)
    for a, b in zip(edges[:-1], edges[1:]):
        m = (r >= a) & (r < b)
        lam = want[m].sum()
        if lam > 50:
            assert abs(counts[m].sum() - lam) < 5 * math.sqrt(lam), (a, b, counts[m].sum(), lam)


def test_sources_rsd_linear_potential():
    """Pass-2 RSD (srcs.c:87-118, 235-237): phi = g.x on the grid (values
    exact in FP32, so the central differences are exact) gives
    dz_rsd = f_v (g.u) V1(r), f_v = -f0/(1.5 H0 Om), u the cell's unit vector,
    for every source of cells away from the periodic edges; V1 varies with r."""
    ta = _tables()
    ta["v1"] = 1.0 + ta["r"] / 5000.0
    orc = Oracle(ta)
    N, Lb = 32, 800.0
    dx = Lb / N
    gvec = np.array([0.5, -0.25, 0.125])
    i = np.arange(N) * dx
    phi = np.zeros((N, N, 2 * (N // 2 + 1)))
    phi[:, :, :N] = gvec[0] * i[None, None, :] + gvec[1] * i[None, :, None] + gvec[2] * i[:, None, None]
    assert np.array_equal(phi.astype(np.float32).astype(np.float64), phi)
    r_t = ta["r"]
    amp = np.where(r_t < 0.35 * Lb, 4e-3, 0.0)   # only cells with 1 <= index <= N-2
    pop = orc.pop(N, Lb, amp, np.ones(5001), np.ones(5001), 1.0, 1.0, 0, 0, 77)
    dens = np.zeros_like(phi)
    pos, _ = orc.sources(pop, dens, phi, 2)
    assert len(pos) > 2000
    cell = np.floor((pos[:, :3].astype(np.float64) + Lb / 2) / dx + 0.5)
    c0 = cell * dx - Lb / 2
    rr = np.sqrt((c0 ** 2).sum(axis=1))
    fv = -ta["fgrowth_0"] / (1.5 * ta["hubble_0"] * ta["omega_m"])
    want = fv * (c0 @ gvec) / rr * np.interp(rr, r_t, ta["v1"])
    np.testing.assert_allclose(pos[:, 3], want, rtol=3e-7, atol=1e-12 * np.abs(want).max())
    # the sign: a potential rising outward pulls sources toward the observer
    assert np.corrcoef(pos[:, 3], c0 @ gvec / rr)[0, 1] < -0.99 if fv > 0 else True


def test_shell_sums_bias_varying_all_models():
    """collect_density_normalization_from_grid (density.c:1128-1211) with a
    random density and b(r) varying (one tracer per model, plus negative b)
    against a numpy brute force: counts exact, z and bias sums to 1e-12."""
    ta = _tables()
    orc = Oracle(ta)
    N, Lb = 16, 1500.0
    rng = np.random.default_rng(8)
    g = rng.uniform(-1.3, 3.0, size=(N, N, 2 * (N // 2 + 1)))
    r_t = ta["r"]
    b_pos = 0.3 + r_t / 800.0
    b_neg = -0.6 + r_t / 4000.0
    tracers = [(b_pos, 0), (b_pos, 1), (b_pos, 2), (b_neg, 0), (b_neg, 2)]
    nz, zmax, cnt, zs, sums = orc.shell_sums(N, Lb, g, tracers)
    r, _ = _radii(N, Lb)
    z = np.interp(r, r_t, ta["z"])
    idz = (nz - 2) / zmax
    s = (z * idz).astype(int) + 1
    m = s < nz
    np.testing.assert_array_equal(cnt, np.bincount(s[m], minlength=nz))
    np.testing.assert_allclose(zs, np.bincount(s[m], weights=z[m], minlength=nz), rtol=1e-12)
    d = g[:, :, :N]
    for t, (b, model) in enumerate(tracers):
        bb = np.interp(r, r_t, b)
        want = np.bincount(s[m], weights=_bias_np(d, bb, model)[m], minlength=nz)
        np.testing.assert_allclose(sums[t], want, rtol=1e-12, atol=1e-300)
    assert np.all(sums[3][1:] != sums[0][1:])   # bin 0 is never filled (density.c:1168)


@pytest.mark.parametrize("model", [0, 2])
def test_imap_rsd_bias_brute_force(model):
    """imap_set_cartesian_single (imap.c:135-245) with phi = g.x (exact FP32
    differences), T(r), b(r), norm(r) and V1(r) all varying, against a numpy
    enumeration of every cell's nsub^3 sub-points: the RSD shift
    dr = f_v (g.u) V1 IH moves sub-points between shells; the pixel sums
    (T bias_model(delta, b) norm per sub-point) agree to 1e-12 and the hit
    counts exactly.  A wrong sign of the shift or a missing IH factor moves
    hundreds of hits."""
    ta = _tables()
    ta["v1"] = 1.0 + ta["r"] / 5000.0
    ta["ih"] = 3000.0 * (1.0 + ta["r"] / 8000.0)
    orc = Oracle(ta)
    N, Lb, nside = 16, 600.0, 4
    dx = Lb / N
    gvec = np.array([2.0e-2, -1.0e-2, 1.5e-2]) / 64.0
    i = np.arange(N) * dx
    phi = np.zeros((N, N, 2 * (N // 2 + 1)))
    phi[:, :, :N] = gvec[0] * i[None, None, :] + gvec[1] * i[None, :, None] + gvec[2] * i[:, None, None]
    assert np.array_equal(phi.astype(np.float32).astype(np.float64), phi)
    rng = np.random.default_rng(model)
    dens = np.zeros_like(phi)
    dens[:, :, :N] = np.exp(rng.normal(0, 0.5, size=(N, N, N))) - 1
    r_t = ta["r"]
    tz = 0.2 * (1 + r_t / 1000.0)
    bz = 1.3 + r_t / 2000.0
    norm = 0.9 + r_t / 5000.0
    r0 = np.array([100.0, 150.0, 200.0])
    rf = np.array([150.0, 200.0, 250.0])
    data, nadd = orc.imap(N, Lb, dens, phi, tz, bz, norm, float(norm[0]), float(norm[-1]), model, nside, r0, rf)
    # numpy enumeration
    fv = -ta["fgrowth_0"] / (1.5 * ta["hubble_0"] * ta["omega_m"])
    npix = 12 * nside * nside
    pix_size = math.sqrt(4 * math.pi / npix)
    vol = (rf ** 3 - r0 ** 3) * pix_size ** 2 / 3
    sc = np.minimum(np.power(vol, 0.333333), np.minimum(r0 * pix_size, rf - r0))
    nsub = (dx / sc + 0.5).astype(int) + 1
    want_d = np.zeros((3, npix))
    want_n = np.zeros((3, npix), dtype=np.int64)
    c = np.arange(N) * dx - Lb / 2
    L_ = load()
    moved = 0
    for iz in range(N):
        for iy in range(N):
            for ix in range(N):
                x0 = np.array([c[ix], c[iy], c[iz]])
                rr = math.sqrt((x0 ** 2).sum())
                if not (r0[0] - 20 <= rr <= rf[-1] + 20):
                    continue
                vr = fv * (x0 @ gvec) / rr
                drs = vr * np.interp(rr, r_t, ta["v1"]) * np.interp(rr, r_t, ta["ih"])
                temp = np.interp(rr, r_t, tz) * _bias_np(dens[iz, iy, ix], np.interp(rr, r_t, bz), model) * \
                    np.interp(rr, r_t, norm)
                sh = np.searchsorted(rf, rr, side="right")
                ns = nsub[min(max(sh, 0), 2)] if rr >= r0[0] else nsub[0]
                off = (np.arange(ns) + 0.5) * dx / ns
                X, Y, Z = np.meshgrid(x0[0] + off, x0[1] + off, x0[2] + off, indexing="ij")
                R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
                shell = np.searchsorted(rf, R + drs, side="right")
                ok = (R + drs >= r0[0]) & (shell < 3)
                moved += np.count_nonzero(np.searchsorted(rf, R, side="right") != shell)
                for xs, ys, zs_, rs_, s_ in zip(X[ok], Y[ok], Z[ok], R[ok], shell[ok]):
                    ph = math.atan2(ys / rs_, xs / rs_)
                    if ph < 0:
                        ph += 2 * math.pi
                    pix = L_.orc_ang2pix_ring(nside, zs_ / rs_, ph)
                    want_d[s_, pix] += temp
                    want_n[s_, pix] += 1
    assert moved > 100
    np.testing.assert_array_equal(nadd, want_n)
    np.testing.assert_allclose(data, want_d, rtol=1e-12, atol=1e-300)


# ------------------------------------------- the reference's own streams ----
def _ref_poisson(L, stream, mu):
    used = ctypes.c_longlong()
    s = np.ascontiguousarray(stream, dtype=np.uint32)
    k = L.orc_ref_poisson(p(s, ctypes.c_uint32), len(s), mu, ctypes.byref(used))
    return k, used.value


def test_ref_poisson_multiplication_method(L):
    """mu <= 10 (gsl_ran_poisson's final loop): k = number of leading partial
    products of the uniforms above e^-mu, k + 1 uniforms used; mu = 0 takes
    one uniform and gives 0 (the reference calls the sampler for every cell
    with n(r) > 0, whatever its lambda)."""
    rng = np.random.default_rng(4)
    for mu in [0.0, 1e-3, 0.37, 1.0, 2.5, 5.0, 9.2, 10.0]:
        for _ in range(40):
            st = rng.integers(1, 2 ** 32, size=64, dtype=np.uint64).astype(np.uint32)
            u = st.astype(np.float64) / 2 ** 32
            cp = np.cumprod(u)
            kk = int(np.argmax(cp <= math.exp(-mu)))
            assert _ref_poisson(L, st, mu) == (kk, kk + 1), (mu, kk)


def test_ref_poisson_gamma_product_branch(L):
    """10 < mu < 13.72: one reduction step with gamma_int(m), m = (int)(7 mu/8)
    < 12, as -log of m uniforms; then either binomial(mu/X, m-1) by inversion
    (k the first value whose CDF exceeds the next uniform, scipy's binomial CDF;
    p > 1/2 counts from the other end) or k = m plus the multiplication method
    on mu - X."""
    rng = np.random.default_rng(6)
    branches = set()
    for it in range(400):
        mu = 10.0 + 3.7 * rng.random()
        st = rng.integers(1, 2 ** 32, size=80, dtype=np.uint64).astype(np.uint32)
        m = int(mu * 7.0 / 8.0)
        if it % 4 == 0:
            # small uniforms: X > 2 mu, so p = mu/X < 1/2 (no flip)
            st[:m] = rng.integers(1, 2 ** 32 // 20, size=m, dtype=np.uint64).astype(np.uint32)
        u = st.astype(np.float64) / 2 ** 32
        X = -math.log(np.prod(u[:m]))
        if X >= mu:
            pb, n = mu / X, m - 1
            flip = pb > 0.5
            pp = 1 - pb if flip else pb
            cdf = stats.binom.cdf(np.arange(n + 1), n, pp)
            ix = int(np.argmax(u[m] < cdf))
            want = (n - ix if flip else ix, m + 1)
            branches.add("binomial" + ("-flip" if flip else ""))
        else:
            rest = mu - X
            cp = np.cumprod(u[m:])
            kk = int(np.argmax(cp <= math.exp(-rest)))
            want = (m + kk, m + kk + 1)
            branches.add("reduce")
        got = _ref_poisson(L, st, mu)
        if got != want and X >= mu:
            # inversion boundary: the running subtraction of the reference
            # against the CDF; only within rounding of a CDF value
            assert np.min(np.abs(u[m] - cdf)) < 1e-12, (mu, got, want)
            continue
        assert got == want, (mu, got, want)
    assert branches == {"binomial", "binomial-flip", "reduce"}


@pytest.mark.parametrize("mu", [9.2, 25.0, 80.0])
def test_ref_poisson_distribution(L, mu):
    """Across the branches (gamma_large for m >= 12): Poisson mean, variance
    and chi^2 on MT19937 streams."""
    from oracle_lib import ref_streams
    st = ref_streams(1234, 0, 1, 4_000_000)[0]
    ks, pos = [], 0
    while len(ks) < 20000:
        k, used = _ref_poisson(L, st[pos:], mu)
        assert used > 0
        pos += used
        ks.append(k)
    k = np.array(ks)
    n = len(k)
    assert abs(k.mean() - mu) < 5 * math.sqrt(mu / n)
    assert abs(k.var() / mu - 1) < 6 * math.sqrt(2.0 / n) + 0.02
    vals = np.arange(max(0, int(mu - 6 * math.sqrt(mu))), int(mu + 6 * math.sqrt(mu)) + 1)
    obs = np.array([(k == v).sum() for v in vals])
    exp = stats.poisson.pmf(vals, mu) * n
    msk = exp > 5
    chi2 = ((obs[msk] - exp[msk]) ** 2 / exp[msk]).sum()
    assert stats.chi2.sf(chi2, msk.sum() - 1) > 1e-4


def test_sources_ref_stream_walk():
    """orc_sources_ref_stream against a driver of its own: every thread walks
    its planes in z, y, x order and calls the restated gsl_ran_poisson on its
    stream for each cell with r < L/2 + 20 and n(r) > 0; pass 2 restarts the
    stream (srcs.c:231), the k-th source of a thread taking uniforms 3k..3k+2
    for x, y, z.  MT19937 streams seeded seed + t + nthr ipop."""
    from oracle_lib import ref_streams, static_split
    ta = _tables()
    orc = Oracle(ta)
    N, Lb, nthr, ipop, seed = 16, 900.0, 3, 1, 555
    r_t = ta["r"]
    amp = np.where(r_t < 380, 6e-5, 0.0)   # lambda ~ 11 (1+delta)^1.4: both Poisson branches
    rng = np.random.default_rng(1)
    g = np.zeros((N, N, 2 * (N // 2 + 1)))
    g[:, :, :N] = np.exp(rng.normal(0, 0.7, size=(N, N, N))) - 1.05
    pop = orc.pop(N, Lb, amp, np.full(5001, 1.4), np.ones(5001), 1.0, 1.0, 0, ipop, seed)
    izb = static_split(N, nthr)
    assert list(izb) == [0, 6, 11, 16]
    st = ref_streams(seed, ipop, nthr, 400000)
    pos, ipix, counts = orc.sources_ref_stream(pop, g, g * 0, nthr, izb, st)
    dx = Lb / N
    want = np.zeros(N ** 3, dtype=np.int32)
    src_cells = []
    for t in range(nthr):
        at = 0
        for iz in range(izb[t], izb[t + 1]):
            for iy in range(N):
                for ix in range(N):
                    r = math.sqrt(sum(((i * dx) - Lb / 2) ** 2 for i in (ix, iy, iz)))
                    if r < Lb / 2 + 20 and r > 0 and np.interp(r, r_t, amp) > 0:
                        lam = orc.lib.orc_cell_lambda(ctypes.byref(pop), p(np.ascontiguousarray(g)), ix, iy, iz)
                        k, used = _ref_poisson(orc.lib, st[t][at:], lam)
                        at += used
                        c = ix + N * (iy + N * iz)
                        want[c] = k
                        src_cells += [(t, ix, iy, iz)] * k
    np.testing.assert_array_equal(counts, want)
    assert len(pos) == want.sum() > 500
    # pass 2: positions from the start of each thread's stream
    kt = {t: 0 for t in range(nthr)}
    for i, (t, ix, iy, iz) in enumerate(src_cells):
        k = kt[t]
        kt[t] += 1
        u = st[t][3 * k: 3 * k + 3].astype(np.float64) / 2 ** 32
        c0 = np.array([ix, iy, iz]) * dx - Lb / 2
        np.testing.assert_array_equal(pos[i, :3], (c0 + dx * (u - 0.5)).astype(np.float32))


# ------------------------------------------------------ surrogate network ----
def _mlp_case(seed=0, n=37):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 0.3, 4610)
    sc = (np.array([0.6, 1.8, 0.9, 2.5]), np.array([0.2, 0.7, 0.4, 1.1]), np.array([-0.3, 0.3]),
          np.array([0.5, 0.4]))
    x = np.column_stack([rng.uniform(0.3, 1, n), rng.uniform(0.5, 3.5, n), rng.uniform(0.2, 1.8, n),
                         rng.uniform(1, 4, n)])
    y = rng.normal(0, 0.5, (n, 2))
    return P, sc, x, y


def test_mlp_forward_matches_numpy():
    """The oracle network against numpy's matrix products (W[out][in])."""
    from oracle_lib import mlp_forward
    P, sc, x, _ = _mlp_case()
    W1, b1 = P[:256].reshape(64, 4), P[256:320]
    W2, b2 = P[320:4416].reshape(64, 64), P[4416:4480]
    W3, b3 = P[4480:4608].reshape(2, 64), P[4608:4610]
    xn = (x - sc[0]) / sc[1]
    h = np.tanh(np.tanh(xn @ W1.T + b1) @ W2.T + b2)
    want = sc[2] + sc[3] * (h @ W3.T + b3)
    np.testing.assert_allclose(mlp_forward(P, sc, x), want, rtol=1e-13, atol=1e-13)


def test_mlp_gradient_finite_differences():
    """Analytic backward pass against central differences of the loss for
    every parameter block (a transposed W2 or a dropped tanh' fails)."""
    from oracle_lib import mlp_loss_grad
    P, sc, x, y = _mlp_case(1)
    loss, g = mlp_loss_grad(P, sc, x, y)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
                          for a, b in [(0, 256), (256, 320), (320, 4416), (4416, 4480), (4480, 4608), (4608, 4610)]])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp_loss_grad(Pp, sc, x, y)[0] - mlp_loss_grad(Pm, sc, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert loss > 0
