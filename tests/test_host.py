"""Host library (param parsing, cosmology tables, writers, HEALPix) and the
C ABI surface, on the CPU."""
import ctypes
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest
from scipy import integrate, interpolate

from colore_amd import inputs, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def host():
    return lib.load_host()


def csm(host, om=0.3, ol=0.7, ob=0.05, h=0.7, w=-1.0):
    c = lib.Csm()
    err = ctypes.create_string_buffer(256)
    assert host.clr_csm_set(ctypes.byref(c), om, ol, ob, w, 0.0, h, err, 256) == 0, err.value
    return c


def test_cosmology_known_answers_from_reference_log(host):
    """example_CoLoRe.ipynb:298-326: Om=0.3, OL=0.7, Ob=0.3, h=0.7, 256^3, z in [1e-4, 0.15]."""
    c = csm(host, 0.3, 0.7, 0.3, 0.7)
    assert f"{c.a_equality:.5f}" == "0.75416"
    assert f"{c.phorizon:.3E}" == "9.908E+03"
    assert f"{c.growth0:.3f}" == "0.779"
    rmin = host.clr_csm_radial_comoving_distance(ctypes.byref(c), 1 / 1.0001)
    rmax = host.clr_csm_radial_comoving_distance(ctypes.byref(c), 1 / 1.15)
    lbox = 2 * rmax * (1 + 2.0 / 256)
    assert f"{rmin:.3f}" == "0.300" and f"{rmax:.3f}" == "434.105"
    assert f"{lbox:.3f}" == "874.993"
    assert f"{2 * math.pi / lbox:.3E}" == "7.181E-03" and f"{lbox / 256:.3E}" == "3.418E+00"


def test_background_against_quadrature(host):
    c = csm(host)
    om, ol = 0.3, 0.7
    E = lambda a: math.sqrt(om / a ** 3 + ol)
    for z in (0.1, 0.5, 1.0, 2.0):
        a = 1 / (1 + z)
        chi = 2997.92458 * integrate.quad(lambda x: 1 / (x * x * E(x)), a, 1, epsabs=0, epsrel=1e-12)[0]
        assert abs(host.clr_csm_radial_comoving_distance(ctypes.byref(c), a) / chi - 1) < 1e-6
        assert abs(host.clr_csm_hubble(ctypes.byref(c), a) - E(a) / 2997.92458) < 1e-15
        # growth: D(a) ~ (5/2) Om E(a) int_0^a da' / (a' E(a'))^3, normalised at a=1
        g = lambda aa: 2.5 * om * E(aa) * integrate.quad(lambda x: 1 / (x * E(x)) ** 3, 0, aa, epsrel=1e-12)[0]
        D = host.clr_csm_growth_factor(ctypes.byref(c), a) / host.clr_csm_growth_factor(ctypes.byref(c), 1.0)
        assert abs(D / (g(a) / g(1.0)) - 1) < 1e-6
        # f = dlnD/dlna by finite differences of the same D(a)
        eps = 1e-4
        f_fd = (math.log(g(a * (1 + eps))) - math.log(g(a * (1 - eps)))) / (math.log(1 + eps) - math.log(1 - eps))
        assert abs(host.clr_csm_f_growth(ctypes.byref(c), a) / f_fd - 1) < 1e-5


def test_natural_cubic_spline(host):
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(0, 10, 40))
    y = np.sin(x) + 0.1 * x
    xq = np.concatenate([np.linspace(x[0], x[-1], 500), [x[0] - 1e-9, x[-1] + 0.1]])
    out = np.zeros_like(xq)
    host.clr_cspline_eval(len(x), lib.ptr(x), lib.ptr(y), len(xq), lib.ptr(xq), lib.ptr(out))
    ref = interpolate.CubicSpline(x, y, bc_type="natural")(xq[:-2])
    np.testing.assert_allclose(out[:-2], ref, atol=1e-12)
    assert np.isnan(out[-2]) and np.isnan(out[-1])


def test_sigma2_tophat_and_gauss_against_quad(host):
    lk = np.linspace(-4, 2, 300)
    k = 10 ** lk
    pk = inputs.eh_nowiggle_pk(k) * 1e4
    idlogk = (len(lk) - 1) / 6.0
    ns = 0.96

    def pk0(l):
        ik = int((l + 4) * idlogk)
        if ik < 0:
            return pk[0] * 10 ** (ns * (l + 4))
        if ik < len(lk) - 1:
            return pk[ik] + (l - lk[ik]) * (pk[ik + 1] - pk[ik]) * idlogk
        return pk[-1] * 10 ** (-3 * (l - 2))

    def w_th(x):   # TopHat window with the series below x = 0.1 (cosmo.c:251-289)
        if x >= 0.1:
            return 3 * (math.sin(x) - x * math.cos(x)) / x ** 3
        return 1. - 0.1 * x * x + 0.003571429 * x ** 4 - 6.61376E-5 * x ** 6 + 7.51563E-7 * x ** 8

    # P is linear between table nodes: integrate node to node
    edges = np.concatenate([np.linspace(-16, -4, 25)[:-1], lk])
    for R, wf, w in ((8.0, 0, w_th), (3.0, 1, lambda x: math.exp(-0.5 * x * x))):
        f = lambda l: math.log(10) / (2 * math.pi ** 2) * pk0(l) * (10 ** l) ** 3 * w(10 ** l * R) ** 2
        ref = sum(integrate.quad(f, a, b, epsrel=1e-13, epsabs=0, limit=200)[0]
                  for a, b in zip(edges[:-1], edges[1:]))
        got = host.clr_sigma2_window(len(lk), lib.ptr(lk), lib.ptr(pk), -4.0, 2.0, idlogk, ns, R, wf)
        assert abs(got / ref - 1) < 1e-8


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_param_file_parse_and_tables(tmp_path):
    p = inputs.write_config("C1", str(tmp_path), n_grid=64)
    run = lib.Run(p)
    assert run.scalar("n_grid") == 64 and run.scalar("seed") == 1001
    assert run.scalar("n_srcs") == 1 and run.scalar("nside_base") == 2
    rmax = run.scalar("r_max")
    assert abs(run.scalar("l_box") - 2 * rmax * (1 + 2 / 64)) < 1e-9
    ta = run.table_arrays()
    host = run.host
    # tables: z(r(z)) = z, D(0) = 1, V1(0) = 1 (velocity growth normalised at z=0)
    for z in (0.05, 0.3, 0.5):
        r = host.clr_run_r_of_z(run.h, z)
        assert abs(host.clr_run_get_bg(run.h, r, 0, 0) - z) < 2e-5
    assert ta["d1"][0] == 1.0 and abs(ta["v1"][0] - 1) < 1e-12
    # sigma8 normalisation: sigma8(TopHat 8) of the normalised table is sigma_8
    s8 = host.clr_sigma2_window(ta["numk"], lib.ptr(ta["logk"]), lib.ptr(ta["pk"]), ta["logkmin"], ta["logkmax"],
                                ta["idlogk"], ta["ns"], 8.0, 0)
    assert abs(math.sqrt(s8) - 0.8) < 1e-9
    # n(z) in deg^-2 round trip: the table holds the comoving density written
    amp, bz, model = run.tracer(0, 0)
    assert model == 0 and np.all(bz[np.isfinite(bz)] == 1.0)
    r = host.clr_run_r_of_z(run.h, 0.3)
    i = int(r * ta["glob_idr"])
    nbar = amp[i] + (amp[i + 1] - amp[i]) * (r - ta["r"][i]) * ta["glob_idr"]
    assert abs(nbar / 1e-3 - 1) < 2e-3


def test_param_errors(tmp_path):
    base = inputs.write_config("C1", str(tmp_path), n_grid=32)
    text = open(base).read()

    def err_of(t):
        q = os.path.join(str(tmp_path), "bad.cfg")
        _write(q, t)
        with pytest.raises(RuntimeError) as e:
            lib.Run(q)
        return str(e.value)

    assert "Couldn't read variable global.seed" in err_of(re.sub(r"\s*seed = \d+;", "", text))
    assert "HDF5 format not supported" in err_of(text.replace('"ASCII"', '"HDF5"'))
    assert "Unrecognized format" in err_of(text.replace('"ASCII"', '"XML"'))
    assert "Unknown bias model" in err_of(text.replace('"exponential"', '"quadratic"'))
    assert "Density type 4 not supported" in err_of(text.replace("dens_type = 0", "dens_type = 4"))
    assert "Wrong interpolation type" in err_of(text.replace("dens_type = 0", "dens_type = 1")
                                                .replace("lpt_interp_type = 1", "lpt_interp_type = 3"))
    assert "Source plane 1 outside redshift range" in err_of(text + "kappa: { z_out = [0.9]; nside = 8; };\n")
    assert "Couldn't open file" in err_of(text.replace("pk.txt", "missing.txt"))
    # n(z) table that does not reach z_max
    nz = os.path.join(str(tmp_path), "nz1.txt")
    d = np.loadtxt(nz)
    np.savetxt(os.path.join(str(tmp_path), "nz_short.txt"), d[d[:, 0] < 0.3])
    assert "N(z) z-range is too small" in err_of(text.replace("nz1.txt", "nz_short.txt"))


def test_libconfig_syntax(tmp_path):
    """Comments, '=' and ':' separators, optional ';', int/float autoconvert,
    arrays, lists, string concatenation, hex integers, case-insensitive bools."""
    base = inputs.write_config("C1", str(tmp_path), n_grid=32)
    d = str(tmp_path)
    t = f"""
    # comment
    global : {{ prefix_out = "{d}/" "out" ; output_format="ASCII"
      output_density = FALSE, pk_filename: "{d}/pk.txt"  // trailing comment
      z_min = 5e-2; z_max = 0.5; seed = 0x3E9; write_pred = false; just_write_pred = false; }};
    /* block
       comment */
    field_par = {{ r_smooth = 2; smooth_potential = True; n_grid = 32.0; dens_type = 0;
      lpt_buffer_fraction = 0.6; lpt_interp_type = 1; output_lpt = 0 }}
    cosmo_par: {{ omega_M = 0.3; omega_L = 0.7; omega_B = 0.05; h = 0.7; w = -1; ns = 0.96; sigma_8 = 0.8; }}
    srcs1: {{ nz_filename = "{d}/nz1.txt"; bias_filename = "{d}/bz1.txt"; include_lensing = false;
      store_skewers = false; bias_model = 0; extra_list = ( 1, "two", [3.0] ) }}
    kappa: {{ z_out = [0.3, 0.45]; nside = 16; }}
    """
    q = os.path.join(d, "syn.cfg")
    _write(q, t)
    run = lib.Run(q)
    assert run.scalar("seed") == 1001 and run.scalar("n_grid") == 32 and run.scalar("r_smooth") == 2.0
    assert run.scalar("n_kappa") == 2 and run.scalar("nside_kappa") == 16 and run.scalar("output_density") == 0


def _read_fits_table(path):
    data = open(path, "rb").read()
    assert len(data) % 2880 == 0
    hdus, pos = [], 0
    while pos < len(data):
        hdr = {}
        while True:
            blk = data[pos:pos + 2880]
            pos += 2880
            done = False
            for i in range(36):
                c = blk[80 * i:80 * (i + 1)].decode()
                if c.startswith("END"):
                    done = True
                    break
                if c[8:10] == "= ":
                    v = c[10:].split(" / ")[0].strip()
                    hdr[c[:8].strip()] = v.strip("'").strip() if v.startswith("'") else v
            if done:
                break
        size = int(hdr.get("NAXIS1", 0)) * int(hdr.get("NAXIS2", 0)) if hdr.get("NAXIS") == "2" else 0
        size *= abs(int(hdr.get("BITPIX", 8))) // 8
        hdus.append((hdr, data[pos:pos + size]))
        pos += (size + 2879) // 2880 * 2880
    return hdus


def test_fits_writers(tmp_path, host):
    f = str(tmp_path / "m.fits")
    nside = 4
    m = np.arange(12 * nside * nside, dtype=np.float32) * 0.5
    assert host.clr_write_healpix_map(f.encode(), nside, lib.ptr(m, ctypes.c_float)) == 0
    (h0, _), (h1, tab) = _read_fits_table(f)
    assert h0["SIMPLE"] == "T" and h1["XTENSION"] == "BINTABLE" and h1["TFORM1"] == "1E"
    assert h1["PIXTYPE"] == "HEALPIX" and h1["ORDERING"] == "RING" and h1["NSIDE"] == "4"
    assert h1["COORDSYS"] == "G" and h1["EXTNAME"] == "BINTABLE" and h1["TTYPE1"] == "map 1"
    np.testing.assert_array_equal(np.frombuffer(tab, dtype=">f4"), m)
    c = str(tmp_path / "c.fits")
    rdz = np.random.default_rng(0).random((100, 4)).astype(np.float32)
    assert host.clr_write_catalog_fits(c.encode(), 2, 100, lib.ptr(rdz, ctypes.c_float), None) == 0
    (_, _), (h1, tab) = _read_fits_table(c)
    assert [h1[f"TTYPE{i}"] for i in range(1, 6)] == ["TYPE", "RA", "DEC", "Z_COSMO", "DZ_RSD"]
    assert [h1[f"TFORM{i}"] for i in range(1, 6)] == ["1J", "1E", "1E", "1E", "1E"]
    assert h1["TUNIT2"] == "DEG" and h1["CONTENTS"] == "Source catalog"
    rec = np.frombuffer(tab, dtype=[("t", ">i4"), ("ra", ">f4"), ("dec", ">f4"), ("z", ">f4"), ("dz", ">f4")])
    assert np.all(rec["t"] == 2)
    np.testing.assert_array_equal(np.stack([rec["ra"], rec["dec"], rec["z"], rec["dz"]], 1), rdz)
    # with lensing columns (has_lensing): 10 fields
    lens = np.random.default_rng(1).random((100, 5)).astype(np.float32)
    assert host.clr_write_catalog_fits(c.encode(), 1, 100, lib.ptr(rdz, ctypes.c_float),
                                       lib.ptr(lens, ctypes.c_float)) == 0
    (_, _), (h1, tab) = _read_fits_table(c)
    assert [h1[f"TTYPE{i}"] for i in range(1, 11)] == ["TYPE", "RA", "DEC", "Z_COSMO", "DZ_RSD", "E1", "E2",
                                                        "KAPPA", "DRA", "DDEC"]
    assert h1["TFIELDS"] == "10" and h1["TUNIT9"] == "DEG" and h1["TUNIT10"] == "DEG"
    rec = np.frombuffer(tab, dtype=[("t", ">i4")] + [(f"c{k}", ">f4") for k in range(9)])
    assert np.all(rec["t"] == 1)
    np.testing.assert_array_equal(np.stack([rec[f"c{k}"] for k in range(4)], 1), rdz)
    np.testing.assert_array_equal(np.stack([rec[f"c{k}"] for k in range(4, 9)], 1), lens)
    a = str(tmp_path / "c.txt")
    assert host.clr_write_catalog_ascii(a.encode(), 1, 100, lib.ptr(rdz, ctypes.c_float),
                                        lib.ptr(lens, ctypes.c_float)) == 0
    with open(a) as fh:
        head = fh.readline()
    assert head.startswith("#[1]type") and "[10]d_dec" in head
    rows = np.loadtxt(a)
    assert rows.shape == (100, 10)
    np.testing.assert_allclose(rows[:, 5:], lens, rtol=1e-6)


def test_skewer_hdus(tmp_path, host):
    """clr_append_skewers_fits: two n x nr images and the R Z D V table with
    SIGMA_G, after the catalogue table (io.c:1139-1187)."""
    c = str(tmp_path / "s.fits")
    rng = np.random.default_rng(2)
    n, nr = 7, 16
    rdz = rng.random((n, 4)).astype(np.float32)
    assert host.clr_write_catalog_fits(c.encode(), 0, n, lib.ptr(rdz, ctypes.c_float), None) == 0
    d, v = rng.random((n, nr)).astype(np.float32), rng.random((n, nr)).astype(np.float32)
    bg4 = rng.random((4, nr)).astype(np.float32)
    assert host.clr_append_skewers_fits(c.encode(), 1, n, nr, lib.ptr(d, ctypes.c_float), lib.ptr(v, ctypes.c_float),
                                        lib.ptr(bg4, ctypes.c_float), ctypes.c_float(0.75)) == 0
    hdus = _read_fits_table(c)
    assert len(hdus) == 5
    (hi, di), (hv, dv), (hb, db) = hdus[2:]
    assert hi["XTENSION"] == "IMAGE" and hi["BITPIX"] == "-32" and hi["NAXIS1"] == str(nr) and hi["NAXIS2"] == str(n)
    assert hi["CONTENTS"] == "gaussian skewers" and hv["CONTENTS"] == "velocity skewers"
    np.testing.assert_array_equal(np.frombuffer(di, dtype=">f4").reshape(n, nr), d)
    np.testing.assert_array_equal(np.frombuffer(dv, dtype=">f4").reshape(n, nr), v)
    assert hb["CONTENTS"] == "Background cosmology" and float(hb["SIGMA_G"]) == 0.75
    assert [hb[f"TTYPE{i}"] for i in range(1, 5)] == ["R", "Z", "D", "V"] and hb["TUNIT1"] == "MPC_H"
    rec = np.frombuffer(db, dtype=">f4").reshape(nr, 4)
    np.testing.assert_array_equal(rec.T, bg4)


def test_density_grid_writer(tmp_path, host):
    N, L = 8, 123.5
    g = np.random.default_rng(1).random((N, N, N + 2)).astype(np.float32)
    f = str(tmp_path / "d.dat")
    assert host.clr_write_density_grid(f.encode(), N, L, lib.ptr(g, ctypes.c_float)) == 0
    raw = open(f, "rb").read()
    nn, sz = struct.unpack("ii", raw[:8])
    (lb,) = struct.unpack("d", raw[8:16])
    n, nz, iz0 = struct.unpack("iii", raw[16:28])
    assert (nn, sz, n, nz, iz0) == (1, 4, N, N, 0) and lb == L
    np.testing.assert_array_equal(np.frombuffer(raw[28:], dtype=np.float32).reshape(N, N, N), g[:, :, :N])


def test_host_healpix_matches_oracle(host):
    from oracle_lib import load
    o = load()
    v, w = np.zeros(3), np.zeros(3)
    for nside in (1, 2, 8, 32):
        for p in range(0, 12 * nside * nside, max(1, nside)):
            assert host.clr_hp_nest2ring(nside, p) == o.orc_nest2ring(nside, p)
            assert host.clr_hp_ring2nest(nside, p) == o.orc_ring2nest(nside, p)
            host.clr_hp_pix2vec_nest(nside, p, lib.ptr(v))
            o.orc_pix2vec_nest(nside, p, lib.ptr(w))
            assert np.array_equal(v, w)
            assert host.clr_hp_vec2pix_ring(nside, lib.ptr(v)) == host.clr_hp_nest2ring(nside, p)


def test_c_abi_library_exports_every_declared_symbol():
    import ctypes as ct
    for hdr, so in (("colore_amd.h", lib.AMD_SO), ("colore_host.h", lib.HOST_SO)):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names = set(re.findall(r"\b(clr_[a-z0-9_]+)\s*\(", text))
        names -= {"clr_ctx"}
        assert len(names) > 10
        so_lib = ct.CDLL(so)
        missing = [n for n in sorted(names) if not hasattr(so_lib, n)]
        assert not missing, (hdr, missing)


def test_fft_candidate_entry_points():
    """clr_fft_candidates / clr_fft_select (host-side bookkeeping, no device
    call): every grid size with an FFT configuration has at least one tile
    shape, other sizes none; selection accepts 0..n-1 and rejects the rest."""
    import ctypes as ct
    so = ct.CDLL(lib.AMD_SO)
    so.clr_fft_candidates.restype = ct.c_int
    so.clr_fft_select.restype = ct.c_int
    for n in (32, 64, 128, 192, 256, 384, 512, 1024, 2048):
        k = so.clr_fft_candidates(n)
        assert k >= 1, n
        assert so.clr_fft_select(n, k) != 0 and so.clr_fft_select(n, -1) != 0
        assert so.clr_fft_select(n, k - 1) == 0 and so.clr_fft_select(n, 0) == 0
    assert so.clr_fft_candidates(1024) >= 5
    for n in (0, 100, 96, 4096):
        assert so.clr_fft_candidates(n) == 0
        assert so.clr_fft_select(n, 0) != 0


def test_cli_usage_and_test_memory(tmp_path):
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert "Usage: ./CoLoRe file_name" in r.stderr
    p = inputs.write_config("C1", str(tmp_path), n_grid=64)
    r = subprocess.run([exe, "--test-memory", p], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "L_box = 2726.703 Mpc/h, N_grid = 64" in r.stdout   # 2 r_max (1 + 2/64), r_max = chi(0.5)
    r = subprocess.run([exe, str(tmp_path / "missing.cfg")], capture_output=True, text=True)
    assert r.returncode == 1 and "Fatal: Couldn't open file" in r.stderr


# ------------------------------------------------------------ predictions ----
def _fftlog(host, name, x, y):
    n = len(x)
    xo, yo = np.zeros(n), np.zeros(n)
    fn = getattr(host, name)
    fn.restype = None
    fn.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 4
    x, y = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
    fn(n, lib.ptr(x), lib.ptr(y), lib.ptr(xo), lib.ptr(yo))
    return xo, yo


def _fht_numpy(r, a, mu):
    """fht (fftlog.c, q = 0, kcrc = 1, low-ringing kcrc) from numpy's FFT and
    scipy's complex log-Gamma."""
    from scipy.special import loggamma
    N = len(r)
    L = np.log(r[-1] / r[0]) * N / (N - 1.0)
    x, kcrc = (mu + 1) / 2, 1.0
    arg = np.log(2 / kcrc) * N / L + 2 * loggamma(x + 1j * np.pi * N / (2 * L)).imag / np.pi
    if arg != np.round(arg):
        kcrc *= np.exp((arg - np.round(arg)) * L / N)
    y, k0r0 = np.pi / L, kcrc * np.exp(-L)
    m = np.arange(N)
    mm = np.where(m <= N // 2, m, N - m)
    u = np.exp(1j * (mm * (-2 * y * np.log(k0r0 / 2)) + 2 * loggamma(x + 1j * mm * y).imag))
    u = np.where(m <= N // 2, u, np.conj(u))
    if N % 2 == 0:
        u[N // 2] = u[N // 2].real
    b = np.fft.ifft(np.fft.fft(a) * u)[::-1]
    return k0r0 / r[0] * np.exp(np.arange(N) * L / N), b


@pytest.mark.parametrize("n", [4096, 10000, 3 * 7 * 97])
def test_fftlog_matches_library_routines(host, n):
    """The host's FFTLog (its own mixed-radix FFT for lengths with factors
    2, 3, 5, 7, 97 and a Stirling-series log-Gamma) against the same
    transform built on numpy.fft and scipy.special.loggamma."""
    R = 10.0
    k = 1e-4 * (1e6) ** (np.arange(n) / (n - 1))
    pk = (2 * np.pi * R * R) ** 1.5 * np.exp(-0.5 * (k * R) ** 2) * (1 + 0.3 * np.sin(3 * np.log(k)))
    r, xi = _fftlog(host, "clr_pk2xi", k, pk)
    r2, b = _fht_numpy(k, k ** 1.5 * pk, 0.5)
    np.testing.assert_allclose(r, r2, rtol=1e-13)
    want = ((2 * np.pi * r2) ** -1.5 * b).real
    np.testing.assert_allclose(xi, want, rtol=0, atol=1e-9 * np.abs(want).max())
    k3, pk3 = _fftlog(host, "clr_xi2pk", r, xi)
    r4, b4 = _fht_numpy(r, r ** 1.5 * xi, 0.5)
    want = 8 * np.pi ** 3 * ((2 * np.pi * r4) ** -1.5 * b4).real
    np.testing.assert_allclose(k3, r4, rtol=1e-13)
    np.testing.assert_allclose(pk3, want, rtol=0, atol=1e-9 * np.abs(want).max())


def test_fftlog_gaussian_pair(host):
    """P(k) = (2 pi R^2)^(3/2) exp(-k^2 R^2 / 2) and xi(r) = exp(-r^2 / (2 R^2))
    are a 3-D Fourier pair.  The reference's fht labels its output one grid
    step low (the reversed inverse FFT is the forward one at n+1, while
    k[n] = k0 e^(n L/N)), so value n is the transform at r_n e^D, D = L/N,
    times e^(1.5 D) from the prefactor at r_n: the closed form holds in that
    form to 1e-3 and misses by ~1.5 D without the shift."""
    R, n = 10.0, 4096
    k = 1e-4 * (1e6) ** (np.arange(n) / (n - 1))
    pk = (2 * np.pi * R * R) ** 1.5 * np.exp(-0.5 * (k * R) ** 2)
    r, xi = _fftlog(host, "clr_pk2xi", k, pk)
    D = np.log(r[1] / r[0])
    mid = (r > 0.5) & (r < 60)
    exact = np.exp(-0.5 * (r[mid] * np.exp(D) / R) ** 2)
    assert np.abs(xi[mid] * np.exp(-1.5 * D) - exact).max() < 1e-3
    assert np.abs(xi[mid] - np.exp(-0.5 * (r[mid] / R) ** 2)).max() > 0.5 * 1.5 * D


def test_write_predictions(tmp_path):
    """just_write_pred: the driver writes <prefix>_gbias.txt (z, r(z), D(z),
    one bias per tracer, every pred_dz up to z_max) and per tracer and z the
    P(k) and xi(r) files; P_tl / P_ll = b exp(-(rsm_gm - rsm_mm) k^2) and, on
    linear scales, P_tt -> b^2 P_ll exp(-(rsm_gg - rsm_mm) k^2)."""
    from colore_amd import inputs
    N = 64
    p = inputs.write_config("C3", str(tmp_path), n_grid=N, z_max=1.6, with_maps=False, pred_dz=0.5, just_pred=True)
    exe = os.path.join(ROOT, "colore_amd", "CoLoRe")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "*** Writing predictions" in r.stdout
    assert not os.path.exists(tmp_path / "out_srcs_s1_0.txt")   # just_write_pred stops there
    run = lib.Run(p)
    gb = np.loadtxt(tmp_path / "out_gbias.txt")
    assert gb.shape == (4, 6)
    np.testing.assert_allclose(gb[:, 0], [0, 0.5, 1.0, 1.5])
    for row in gb:
        rr = run.host.clr_run_r_of_z(run.h, row[0])
        assert abs(row[1] - rr) <= 1e-5 * max(rr, 1)
        assert abs(row[2] - run.host.clr_run_get_bg(run.h, rr, 1, 0)) < 1e-5
        for ip in range(3):
            assert abs(row[3 + ip] - run.host.clr_run_get_bg(run.h, rr, 5, ip)) <= 1e-5 * abs(row[3 + ip])
    L, dx2 = run.scalar("l_box"), (run.scalar("l_box") / N) ** 2
    r2 = run.scalar("r_smooth") ** 2
    gg, gm, mm = r2 + 0.9 * dx2 / 12, r2 + 2.0 * dx2 / 6, r2 + 1.88 * dx2 / 6
    for ip in range(3):
        b = gb[2, 3 + ip]
        pk = np.loadtxt(tmp_path / f"out_pk_srcs_pop{ip}_z1.000.txt")
        xi = np.loadtxt(tmp_path / f"out_xi_srcs_pop{ip}_z1.000.txt")
        assert xi[:, 0].min() >= 0.5 and xi[:, 0].max() <= 300
        k = pk[:, 0]
        ok = (pk[:, 2] > 1e-200) & (pk[:, 3] > 1e-200)   # the grid smoothing underflows at high k
        assert ok.sum() > 1000
        np.testing.assert_allclose(pk[ok, 2] / pk[ok, 3], b * np.exp(-(gm - mm) * k[ok] ** 2), rtol=1e-4)
        lin = (k > 0.01) & (k < 0.03)
        ratio = pk[lin, 1] / (b * b * pk[lin, 3] * np.exp(-(gg - mm) * k[lin] ** 2))
        assert np.all(np.abs(ratio - 1) < 0.05), ratio
        np.testing.assert_allclose(xi[:, 2], b * xi[:, 3], rtol=1e-4)


def test_lpt_gadget_writer(tmp_path, host):
    """write_lpt's Gadget-2 snapshot: a 256-byte header (np[1], mass[1] =
    27.7455 Om L^3 / n, time 1, np_total[1], one file, box, Om, OL, h, the
    format flag word), then positions, zero velocities and 64-bit IDs, each
    block between its byte-count words."""
    import struct
    n, L = 27, 300.0
    xyz = np.random.default_rng(3).random((n, 3)).astype(np.float32) * L
    f = str(tmp_path / "out_lpt_out.0")
    assert host.clr_write_lpt_gadget(f.encode(), n, L, 0.3, 0.7, 0.7, lib.ptr(xyz, ctypes.c_float)) == 0
    raw = open(f, "rb").read()
    assert len(raw) == 8 + 256 + 3 * 8 + n * (12 + 12 + 8)
    assert struct.unpack("<i", raw[:4])[0] == 256 and struct.unpack("<i", raw[260:264])[0] == 256
    h = raw[4:260]
    assert struct.unpack("<6i", h[:24]) == (0, n, 0, 0, 0, 0)
    mass = struct.unpack("<6d", h[24:72])
    assert mass[1] == pytest.approx(27.7455 * 0.3 * L ** 3 / n, rel=1e-15) and mass[0] == 0
    assert struct.unpack("<2d", h[72:88]) == (1.0, 0.0)
    assert struct.unpack("<6I", h[96:120])[1] == n and struct.unpack("<i", h[124:128])[0] == 1
    assert struct.unpack("<4d", h[128:160]) == (L, 0.3, 0.7, 0.7)
    assert struct.unpack("<6I", h[168:192])[1] == 0 and struct.unpack("<i", h[196:200])[0] == 1
    o = 264
    assert struct.unpack("<i", raw[o:o + 4])[0] == 12 * n
    np.testing.assert_array_equal(np.frombuffer(raw[o + 4:o + 4 + 12 * n], dtype="<f4").reshape(n, 3), xyz)
    o += 8 + 12 * n
    assert not np.any(np.frombuffer(raw[o + 4:o + 4 + 12 * n], dtype="<f4"))
    o += 8 + 12 * n
    assert struct.unpack("<i", raw[o:o + 4])[0] == 8 * n
    np.testing.assert_array_equal(np.frombuffer(raw[o + 4:o + 4 + 8 * n], dtype="<u8"), np.arange(n))


@pytest.mark.parametrize("typ", [0, 2])
def test_limber_cl_against_quadrature(tmp_path, typ):
    """compute_cl (cosmo.c:114-224) restated: against scipy's quad of the
    same Limber integral written from the windows' definitions on the run's
    tables (np.interp, the oracle's P(k) interpolation), for the kappa and
    ISW windows, a source plane inside the box, several l and window pairs."""
    import ctypes
    from scipy import integrate
    from colore_amd import inputs
    from oracle_lib import Oracle
    p = inputs.write_config("C2", str(tmp_path), n_grid=64, with_maps=False)
    run = lib.Run(p)
    ta = run.table_arrays()
    orc = Oracle(ta)
    rs = run.scalar("r_smooth")
    chis = run.host.clr_run_r_of_z(run.h, 1.0)
    rmax = run.scalar("r_max")
    pre = ta["prefac_lensing"]

    def window(l, k, c0, cf):
        chi = (l + 0.5) / k
        if chi <= 0 or chi < c0 or chi > cf:
            return 0.0
        if typ == 0:
            gf = np.interp(chi, ta["r"], ta["d1"])
            aa = 1 / (1 + np.interp(chi, ta["r"], ta["z"]))
            return pre * l * (l + 1.0) * gf * (chis - chi) / (k * k * chis * chi * aa)
        return -2 * pre * np.interp(chi, ta["r"], ta["pd"]) / (k * k)

    for l in (2, 10, 60, 300):
        for (a0, af, b0, bf) in ((0, chis, 0, chis), (0, 0.5 * chis, 0.3 * chis, chis), (0, rmax, 0, rmax)):
            got = run.host.clr_limber_cl(run.h, l, a0, af, b0, bf, chis, typ)

            def f(lk):
                k = 10 ** lk
                pk = orc.lib.orc_pk_linear0(ctypes.byref(orc.pk), lk) * math.exp(-k * k * rs * rs)
                return k * window(l, k, a0, af) * window(l, k, b0, bf) * pk

            cmin, cmax = min(a0, af, b0, bf), max(a0, af, b0, bf)
            lkmax = 2.0 if cmin <= 0 else math.log10((l + 0.5) / cmin)
            lkmin = math.log10((l + 0.5) / cmax)
            pts = [math.log10((l + 0.5) / c) for c in (a0, af, b0, bf) if c > 0]
            pts += [lkmin + 0.5, lkmin + 1, lkmin + 2]
            val = integrate.quad(f, lkmin, lkmax, points=sorted(q for q in pts if lkmin < q < lkmax), limit=4000,
                                 epsrel=1e-12)[0]
            want = math.log(10) * val / (l + 0.5)
            assert abs(got - want) <= 2e-5 * abs(want) + 1e-30, (l, a0, af, b0, bf, got, want)
            if typ == 0 and (a0, af) == (0, chis):
                assert got > 0


def _bias_model_py(d, b, model):
    """bias_model (common.h:414-431) in Python floats."""
    if d <= -1:
        return 0.0
    if model == 1:
        return math.exp(b * d / (1 + d)) if d < 0 else 1 + b * d
    if model == 2:
        return 1 + b * d if 1 + b * d > 0 else 0.0
    return (1 + d) ** b


def test_bias_moment_against_quadrature_and_closed_forms():
    """clr_bias_moment (host code of the device library, no device call).
    One-cell case (mu = 0, var = s^2, k3 = 0): the exponential model equals
    exp(b(b-1) s^2/2) and scipy's quad; both truncated models match quad of
    the same expectation over G ~ N(0, s^2) (kink points given to quad); the
    truncated model also matches its closed form (1-b) Phi(-g0/s) + b
    Phi((s^2-g0)/s), g0 = ln(1-1/b) + s^2/2, for b > 1, and = 1 exactly when
    0 <= b <= 1 (no truncation, <delta> = 0).  Shifted case (mu, var, k3 of a
    shell): the exponential model equals exp(b mu + b^2 var/2 + b^3 k3/6 -
    b sc^2/2); the truncated models match quad over the Edgeworth density
    phi(t) (1 + k3/(6 var^1.5) He_3(t)), and with k3 = 0 quad over
    N(mu, var)."""
    from scipy import stats
    amd = lib.load_amd()

    def quad(model, b, sc, mu=0.0, var=None, k3=0.0):
        var = sc * sc if var is None else var
        sd = math.sqrt(var)
        g1 = k3 / sd ** 3

        def f(t):
            w = stats.norm.pdf(t) * (1 + g1 / 6 * (t ** 3 - 3 * t))
            return w * _bias_model_py(math.expm1(mu + sd * t - 0.5 * sc * sc), b, model)
        pts = []
        if model == 2 and (b > 1 or b < 0):
            pts = [(math.log(1 - 1 / b) + 0.5 * sc * sc - mu) / sd]
        if model == 1:
            pts = [(0.5 * sc * sc - mu) / sd]
        pts = [p for p in pts if -14 < p < 14]
        return integrate.quad(f, -14, 14, points=pts or None, limit=400, epsabs=0, epsrel=1e-13)[0]

    for model in (0, 1, 2):
        for b in (-0.5, 0.3, 1.0, 1.5, 2.5, 4.0):
            if model == 1 and b < 0:
                continue   # exp(b delta/(1+delta)) diverges as delta -> -1
            for s in (0.05, 0.3, 0.8, 1.5):
                got = amd.clr_bias_moment(model, b, s, 0.0, s * s, 0.0)
                assert got == pytest.approx(quad(model, b, s), rel=1e-10), (model, b, s)
    for b in (1.2, 2.5, 4.0):
        for s in (0.2, 0.9):
            g0 = math.log(1 - 1 / b) + 0.5 * s * s
            cf = (1 - b) * stats.norm.sf(g0 / s) + b * stats.norm.sf((g0 - s * s) / s)
            assert amd.clr_bias_moment(2, b, s, 0.0, s * s, 0.0) == pytest.approx(cf, rel=1e-12)
    for b in (0.0, 0.5, 1.0):
        assert amd.clr_bias_moment(2, b, 0.7, 0.0, 0.49, 0.0) == pytest.approx(1.0, abs=1e-13)
    assert amd.clr_bias_moment(0, 2.0, 0.6, 0.0, 0.36, 0.0) == pytest.approx(math.exp(0.36), rel=1e-15)
    assert amd.clr_bias_moment(1, 2.0, 0.0, 0.0, 0.0, 0.0) == 1.0
    # a shell's measured cumulants
    for (b, sc, mu, var, k3) in [(1.7, 0.8, 0.03, 0.61, 0.02), (2.4, 1.1, -0.05, 1.3, -0.04), (0.8, 0.5, 0.01, 0.24, 0.0)]:
        want0 = math.exp(b * mu + 0.5 * b * b * var + b ** 3 * k3 / 6 - 0.5 * b * sc * sc)
        assert amd.clr_bias_moment(0, b, sc, mu, var, k3) == pytest.approx(want0, rel=1e-14)
        for model in (1, 2):
            got = amd.clr_bias_moment(model, b, sc, mu, var, k3)
            assert got == pytest.approx(quad(model, b, sc, mu, var, k3), rel=1e-10), (model, b, sc, mu, var, k3)


def test_imap_azimuth_polynomial_within_its_bound():
    """The 21cm painter accepts a pixel from its shortened evaluation only
    when every floor lies further than the azimuth's error bound kAtanErr
    from an integer.  The bound must hold for the polynomial the kernel
    evaluates: read its coefficients and the bound from the source and
    check atan on [0, 1] densely (Abramowitz & Stegun 4.4.49 quote 2e-8).
    A slipped coefficient or sign moves the error far past the bound."""
    src = open(os.path.join(ROOT, "colore_amd", "csrc", "clr_imap.hip")).read()
    body = src[src.index("inline double atan01"):]
    body = body[:body.index("return t * p;")]
    c0 = float(re.search(r"double p = ([0-9.eE+-]+);", body).group(1))
    steps = re.findall(r"p = p \* t2 ([+-]) ([0-9.eE+-]+);", body)
    assert len(steps) == 8
    bound = float(re.search(r"constexpr double kAtanErr = ([0-9.eE+-]+);", src).group(1))
    x = np.linspace(0.0, 1.0, 2_000_001)
    t2 = x * x
    p = np.full_like(x, c0)
    for sign, c in steps:
        p = p * t2 + (float(c) if sign == "+" else -float(c))
    err = np.abs(x * p - np.arctan(x))
    assert err.max() < 0.75 * bound, err.max()
    assert bound <= 2e-8


def test_map_output_buffer_checks():
    """The map calls write into a caller's buffer only when it has the maps'
    exact layout (C-contiguous float32, planes x pixels); no device call."""
    assert lib._map_out(None, 2, 48).shape == (2, 48)
    ok = np.zeros((2, 48), dtype=np.float32)
    assert lib._map_out(ok, 2, 48) is ok
    for bad in (np.zeros((2, 48)), np.zeros((3, 48), dtype=np.float32),
                np.zeros((48, 2), dtype=np.float32).T):
        with pytest.raises(ValueError):
            lib._map_out(bad, 2, 48)
