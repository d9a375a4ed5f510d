"""ctypes binding of the CPU FP64 oracle (oracle/oracle.h).  Test
infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU baseline only."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.POINTER(ctypes.c_ulonglong)


class Bg(ctypes.Structure):
    _fields_ = [("na", ctypes.c_int), ("idr", ctypes.c_double)] + [(n, _d) for n in ("r", "z", "d1", "v1", "pd", "ih")]


class Pk(ctypes.Structure):
    _fields_ = [("numk", ctypes.c_int)] + [(n, ctypes.c_double) for n in ("logkmin", "logkmax", "idlogk", "ns")] + \
               [("logk", _d), ("pk", _d)]


class Pop(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("L", ctypes.c_double), ("obs", ctypes.c_double * 3), ("bg", ctypes.POINTER(Bg)),
                ("amp", _d), ("bz", _d), ("norm", _d), ("bz0", ctypes.c_double), ("norm0", ctypes.c_double),
                ("normf", ctypes.c_double), ("model", ctypes.c_int), ("ipop", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("fgrowth_0", ctypes.c_double), ("hubble_0", ctypes.c_double), ("omega_m", ctypes.c_double)]


def p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


_LIB = None


def load(build=True):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(SO) and build:
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = ctypes.CDLL(SO)
    ci, cd, cu = ctypes.c_int, ctypes.c_double, ctypes.c_uint32
    _sig(lib, "orc_philox4x32_10", None, [_u32, _u32, _u32])
    _sig(lib, "orc_ref_mode_stream", None, [ci, cu, ci, _u32])
    _sig(lib, "orc_f_of_r", cd, [ctypes.POINTER(Bg), _d, cd, cd, cd])
    _sig(lib, "orc_pk_linear0", cd, [ctypes.POINTER(Pk), cd])
    _sig(lib, "orc_bias_model", cd, [cd, cd, ci])
    _sig(lib, "orc_gen_modes", None, [ci, cd, ctypes.POINTER(Pk), cu, _u32, cd, ci, cd, _d, _d])
    _sig(lib, "orc_c2r_3d", None, [ci, _d, _d])
    _sig(lib, "orc_kspace_variance", None, [ci, cd, ctypes.POINTER(Pk), cu, _u32, cd, ci, cd, _d])
    _sig(lib, "orc_c2r_3d_brute", None, [ci, _d, _d])
    _sig(lib, "orc_cfft_inv", None, [ci, _d])
    _sig(lib, "orc_normalize_sigma2", cd, [ci, cd, _d, _d, _d])
    _sig(lib, "orc_density", None, [ci, cd, _d, ctypes.POINTER(Bg), cd, ci, _d])
    _sig(lib, "orc_shell_sums", ci, [ci, cd, _d, ctypes.POINTER(Bg), _d, ci, ctypes.POINTER(_d), _d, _i, ci, _d,
                                     _u64, _d, _d])
    _sig(lib, "orc_norm_table", None, [ctypes.POINTER(Bg), ci, cd, _u64, _d, _d, _d, _d, _d])
    _sig(lib, "orc_cell_count", ci, [ctypes.POINTER(Pop), _d, ci, ci, ci])
    _sig(lib, "orc_cell_lambda", cd, [ctypes.POINTER(Pop), _d, ci, ci, ci])
    _sig(lib, "orc_counts", ctypes.c_longlong, [ctypes.POINTER(Pop), _d, _i])
    _sig(lib, "orc_counts_planes", ctypes.c_longlong, [ctypes.POINTER(Pop), _d, ci, ci, _i])
    _sig(lib, "orc_sources", ctypes.c_longlong, [ctypes.POINTER(Pop), _d, _d, ci, ctypes.c_longlong, _f, _i])
    _sig(lib, "orc_source_props", None, [ctypes.POINTER(Bg), ctypes.c_longlong, _f, _f])
    _sig(lib, "orc_poisson", ci, [cd, ctypes.c_uint64, cu, cu])
    _sig(lib, "orc_ref_poisson", ci, [_u32, ctypes.c_longlong, cd, ctypes.POINTER(ctypes.c_longlong)])
    _sig(lib, "orc_mt_fill", None, [cu, ctypes.c_longlong, _u32])
    _sig(lib, "orc_sources_ref_stream", ctypes.c_longlong, [ctypes.POINTER(Pop), _d, _d, ci, ci, _i, _u32,
                                                              ctypes.c_longlong, ctypes.c_longlong, _f, _i, _i])
    _sig(lib, "orc_kappa_isw", None, [ci, ci, cd, _d, cd, ctypes.POINTER(Bg), _d, ci, ci, _d, _d])
    _sig(lib, "orc_src_skewers", None, [ci, cd, _d, cd, ctypes.POINTER(Bg), cd, cd, cd, _d, _d, ci, cd,
                                         ctypes.c_longlong, _f, _d, _d])
    _sig(lib, "orc_cstm", ci, [ci, cd, _d, cd, ctypes.POINTER(Bg), _d, _d, _d, _d, cd, cd, ci, ci, _d])
    _sig(lib, "orc_r2c_3d", None, [ci, _d, _d])
    _sig(lib, "orc_lpt", ci, [ci, ci, cd, _d, ctypes.POINTER(Bg), _d, _d, ci, _d, _f])
    _sig(lib, "orc_src_beams", None, [ci, cd, _d, cd, ctypes.POINTER(Bg), _d, cd, cd, cd, _d, ctypes.c_longlong,
                                       _f, _d, _d])
    _sig(lib, "orc_imap", None, [ci, cd, _d, ctypes.POINTER(Bg), _d, _d, _d, _d, cd, _d, cd, cd, ci, cd, cd, cd,
                                 ci, ci, _d, _d, _d, _i])
    _sig(lib, "orc_ang2pix_ring", ctypes.c_longlong, [ctypes.c_longlong, cd, cd])
    _sig(lib, "orc_mlp_forward", None, [_d, _d, _d, _d, _d, _d, ci, _d])
    _sig(lib, "orc_mlp_loss_grad", cd, [_d, _d, _d, _d, _d, _d, _d, ci, _d])
    _sig(lib, "orc_vec2pix_ring", ctypes.c_longlong, [ctypes.c_longlong, _d])
    _sig(lib, "orc_ring2nest", ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong])
    _sig(lib, "orc_nest2ring", ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong])
    _sig(lib, "orc_pix2vec_nest", None, [ctypes.c_longlong, ctypes.c_longlong, _d])
    _LIB = lib
    return lib


class Oracle:
    """Oracle bound to one set of tables (dict from colore_amd.lib.Run.table_arrays)."""

    def __init__(self, ta):
        self.lib = load()
        self.ta = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in ta.items()}
        t = self.ta
        self.bg = Bg(len(t["r"]), t["glob_idr"], p(t["r"]), p(t["z"]), p(t["d1"]), p(t["v1"]), p(t["pd"]), p(t["ih"]))
        self.pk = Pk(len(t["pk"]), t["logkmin"], t["logkmax"], t["idlogk"], t["ns"], p(t["logk"]), p(t["pk"]))

    def modes(self, N, L, seed, r_smooth, smooth_potential=True, injected=None):
        nh = N // 2 + 1
        dk = np.zeros(2 * N * N * nh)
        pk = np.zeros(2 * N * N * nh)
        inj = None if injected is None else p(np.ascontiguousarray(injected, dtype=np.uint32), ctypes.c_uint32)
        self.lib.orc_gen_modes(N, L, ctypes.byref(self.pk), seed, inj, r_smooth, int(smooth_potential),
                               self.ta["prefac_lensing"], p(dk), p(pk))
        return dk, pk

    def kspace_variance(self, N, L, seed, r_smooth, smooth_potential=True, injected=None):
        """(<delta^2>, <phi^2>) of the normalised grids, from the modes alone."""
        out = np.zeros(2)
        inj = None if injected is None else p(np.ascontiguousarray(injected, dtype=np.uint32), ctypes.c_uint32)
        self.lib.orc_kspace_variance(N, L, ctypes.byref(self.pk), seed, inj, r_smooth, int(smooth_potential),
                                     self.ta["prefac_lensing"], p(out))
        return out[0], out[1]

    def c2r(self, N, hk):
        out = np.zeros(N * N * 2 * (N // 2 + 1))
        self.lib.orc_c2r_3d(N, p(np.ascontiguousarray(hk)), p(out))
        return out

    def fields(self, N, L, seed, r_smooth, smooth_potential=True, injected=None):
        dk, pk = self.modes(N, L, seed, r_smooth, smooth_potential, injected)
        d, ph = self.c2r(N, dk), self.c2r(N, pk)
        s2 = self.lib.orc_normalize_sigma2(N, L, p(d), p(ph), None)
        shp = (N, N, 2 * (N // 2 + 1))
        return d.reshape(shp), ph.reshape(shp), s2

    def density(self, N, L, grid, sigma2, dens_type=0):
        g = np.ascontiguousarray(grid, dtype=np.float64).copy()
        obs = np.full(3, 0.5 * L)
        self.lib.orc_density(N, L, p(obs), ctypes.byref(self.bg), sigma2, dens_type, p(g))
        return g

    def shell_sums(self, N, L, grid, tracers, cap=256):
        """tracers: list of (bz array, model)."""
        g = np.ascontiguousarray(grid, dtype=np.float64)
        obs = np.full(3, 0.5 * L)
        bzs = [np.ascontiguousarray(b, dtype=np.float64) for b, _ in tracers]
        arr = (_d * max(len(bzs), 1))(*[p(b) for b in bzs])
        bz0 = np.array([b[0] for b in bzs] + [0.0])
        models = np.array([m for _, m in tracers] + [0], dtype=np.int32)
        cnt = np.zeros(cap, dtype=np.uint64)
        zs = np.zeros(cap)
        sums = np.zeros(max(len(bzs), 1) * cap)
        zmax = ctypes.c_double()
        nz = self.lib.orc_shell_sums(N, L, p(obs), ctypes.byref(self.bg), p(g), len(bzs), arr, p(bz0),
                                     p(models, ctypes.c_int), cap, ctypes.byref(zmax), p(cnt, ctypes.c_ulonglong),
                                     p(zs), p(sums))
        # sums are laid out [t][nz]
        return nz, zmax.value, cnt[:nz].copy(), zs[:nz].copy(), sums[:len(bzs) * nz].reshape(len(bzs), nz).copy()

    def norm_table(self, nz, zmax, cnt, zsum, sums_t):
        out = np.zeros(len(self.ta["r"]))
        n0, nf = ctypes.c_double(), ctypes.c_double()
        self.lib.orc_norm_table(ctypes.byref(self.bg), nz, zmax, p(np.ascontiguousarray(cnt, dtype=np.uint64), ctypes.c_ulonglong),
                                p(np.ascontiguousarray(zsum)), p(np.ascontiguousarray(sums_t)), p(out),
                                ctypes.byref(n0), ctypes.byref(nf))
        return out, n0.value, nf.value

    def pop(self, N, L, amp, bz, norm, norm0, normf, model, ipop, seed):
        self._keep = [np.ascontiguousarray(amp, dtype=np.float64), np.ascontiguousarray(bz, dtype=np.float64),
                      np.ascontiguousarray(norm, dtype=np.float64)]
        t = self.ta
        return Pop(N, L, (ctypes.c_double * 3)(0.5 * L, 0.5 * L, 0.5 * L), ctypes.pointer(self.bg), p(self._keep[0]),
                   p(self._keep[1]), p(self._keep[2]), float(self._keep[1][0]), norm0, normf, model, ipop, seed,
                   t["fgrowth_0"], t["hubble_0"], t["omega_m"])

    def counts(self, pop, dens):
        N = pop.N
        c = np.zeros(N ** 3, dtype=np.int32)
        tot = self.lib.orc_counts(ctypes.byref(pop), p(np.ascontiguousarray(dens, dtype=np.float64)), p(c, ctypes.c_int))
        return tot, c

    def counts_planes(self, pop, dens_slab, iz0):
        """Counts of the planes iz0 .. iz0+nz-1 held (padded) in dens_slab,
        int32[nz][N][N]."""
        N = pop.N
        nz = dens_slab.shape[0]
        c = np.zeros((nz, N, N), dtype=np.int32)
        self.lib.orc_counts_planes(ctypes.byref(pop), p(np.ascontiguousarray(dens_slab, dtype=np.float64)), iz0, nz,
                                   p(c, ctypes.c_int))
        return c

    def sources(self, pop, dens, npot, nside_base=2, cap=None):
        if cap is None:
            cap = int(self.counts(pop, dens)[0]) + 1
        pos = np.zeros((cap, 4), dtype=np.float32)
        ipix = np.zeros(cap, dtype=np.int32)
        n = self.lib.orc_sources(ctypes.byref(pop), p(np.ascontiguousarray(dens, dtype=np.float64)),
                                 p(np.ascontiguousarray(npot, dtype=np.float64)), nside_base, cap,
                                 p(pos, ctypes.c_float), p(ipix, ctypes.c_int))
        assert n >= 0
        return pos[:n], ipix[:n]

    def sources_ref_stream(self, pop, dens, npot, nthr, iz_begin, streams, nside_base=2, cap=None):
        """(pos4, ipix, counts) of srcs.c:120-283 on nthr threads fed the
        streams [nthr, len] (uint32 MT19937 outputs)."""
        N = pop.N
        streams = np.ascontiguousarray(streams, dtype=np.uint32)
        izb = np.ascontiguousarray(iz_begin, dtype=np.int32)
        counts = np.zeros(N ** 3, dtype=np.int32)
        dd = np.ascontiguousarray(dens, dtype=np.float64)
        pp = np.ascontiguousarray(npot, dtype=np.float64)
        if cap is None:
            n = self.lib.orc_sources_ref_stream(ctypes.byref(pop), p(dd), p(pp), nside_base, nthr, p(izb, ctypes.c_int),
                                                p(streams, ctypes.c_uint32), streams.shape[1], 0, None, None,
                                                p(counts, ctypes.c_int))
            assert n != -2, "stream too short"
            cap = int(counts.sum())
        pos = np.zeros((max(cap, 1), 4), dtype=np.float32)
        ipix = np.zeros(max(cap, 1), dtype=np.int32)
        n = self.lib.orc_sources_ref_stream(ctypes.byref(pop), p(dd), p(pp), nside_base, nthr, p(izb, ctypes.c_int),
                                            p(streams, ctypes.c_uint32), streams.shape[1], cap, p(pos, ctypes.c_float),
                                            p(ipix, ctypes.c_int), p(counts, ctypes.c_int))
        assert n >= 0, n
        return pos[:n], ipix[:n], counts

    def source_props(self, pos4):
        out = np.zeros_like(pos4)
        self.lib.orc_source_props(ctypes.byref(self.bg), len(pos4), p(np.ascontiguousarray(pos4), ctypes.c_float),
                                  p(out, ctypes.c_float))
        return out

    def kappa_isw(self, which, N, L, npot, nside, r_src):
        r = np.ascontiguousarray(r_src, dtype=np.float64)
        out = np.zeros(len(r) * 12 * nside * nside)
        obs = np.full(3, 0.5 * L)
        self.lib.orc_kappa_isw(which, N, L, p(obs), self.ta["r_max"], ctypes.byref(self.bg),
                               p(np.ascontiguousarray(npot, dtype=np.float64)), nside, len(r), p(r), p(out))
        return out.reshape(len(r), -1)

    def src_beams(self, N, L, npot, pos4, r_a2r, lensing=True):
        """(dz[n], lens[n, 5] or None) of srcs.c:425-737 for sources at pos4."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        n = len(pos4)
        dz = np.zeros(max(n, 1))
        lens = np.zeros((max(n, 1), 5)) if lensing else None
        obs = np.full(3, 0.5 * L)
        t = self.ta
        self.lib.orc_src_beams(N, L, p(obs), t["r_max"], ctypes.byref(self.bg),
                               p(np.ascontiguousarray(r_a2r, dtype=np.float64)), t["fgrowth_0"], t["hubble_0"],
                               t["omega_m"], p(np.ascontiguousarray(npot, dtype=np.float64)), n, p(pos4, ctypes.c_float),
                               p(dz), p(lens) if lensing else None)
        return dz[:n], (lens[:n] if lensing else None)

    def src_skewers(self, N, L, dens, npot, pos4, gaussian=False, sigma2=0.0):
        """(dg[n, N/2], v[n, N/2]) of the skewers (srcs.c:507-529, 725-733)."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        n = len(pos4)
        nr = N // 2
        dg = np.zeros((max(n, 1), nr))
        vs = np.zeros((max(n, 1), nr))
        obs = np.full(3, 0.5 * L)
        t = self.ta
        self.lib.orc_src_skewers(N, L, p(obs), t["r_max"], ctypes.byref(self.bg), t["fgrowth_0"], t["hubble_0"],
                                 t["omega_m"], p(np.ascontiguousarray(dens, dtype=np.float64)),
                                 p(np.ascontiguousarray(npot, dtype=np.float64)), int(gaussian), float(sigma2), n,
                                 p(pos4, ctypes.c_float), p(dg), p(vs))
        return dg[:n], vs[:n]

    def cstm(self, N, L, dens, kz, bz, norm, norm0, normf, model, nside):
        """One custom map (NEST) of cstm.c:39-150; None when the kernel has no support."""
        out = np.zeros(12 * nside * nside)
        obs = np.full(3, 0.5 * L)
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (dens, kz, bz, norm)]
        rc = self.lib.orc_cstm(N, L, p(obs), self.ta["r_max"], ctypes.byref(self.bg), p(keep[0]), p(keep[1]),
                               p(keep[2]), p(keep[3]), norm0, normf, model, nside, p(out))
        return None if rc else out

    def lpt1(self, N, L, dg, interp, with_pos=False, order=1, d2=None):
        """(delta[N][N][2(N/2+1)], positions[N][N][N][3] or None) of lpt_1
        (order 2: lpt_2 with the D2 table d2)."""
        out = np.zeros((N, N, 2 * (N // 2 + 1)))
        pos = np.zeros((N, N, N, 3), dtype=np.float32) if with_pos else None
        obs = np.full(3, 0.5 * L)
        dg = np.ascontiguousarray(dg, dtype=np.float64)
        d2a = np.ascontiguousarray(d2, dtype=np.float64) if d2 is not None else None
        rc = self.lib.orc_lpt(order, N, L, p(obs), ctypes.byref(self.bg), p(d2a) if d2a is not None else None, p(dg),
                              interp, p(out), p(pos, ctypes.c_float) if with_pos else None)
        assert rc == 0
        return out, pos

    def imap(self, N, L, dens, npot, tz, bz, norm, norm0, normf, model, nside, r0, rf):
        nr = len(r0)
        npix = 12 * nside * nside
        data = np.zeros(nr * npix)
        nadd = np.zeros(nr * npix, dtype=np.int32)
        obs = np.full(3, 0.5 * L)
        t = self.ta
        bz = np.ascontiguousarray(bz, dtype=np.float64)
        self.lib.orc_imap(N, L, p(obs), ctypes.byref(self.bg), p(np.ascontiguousarray(dens, dtype=np.float64)),
                          p(np.ascontiguousarray(npot, dtype=np.float64)), p(np.ascontiguousarray(tz, dtype=np.float64)),
                          p(bz), float(bz[0]), p(np.ascontiguousarray(norm, dtype=np.float64)), norm0, normf, model,
                          t["fgrowth_0"], t["hubble_0"], t["omega_m"], nside, nr,
                          p(np.ascontiguousarray(r0, dtype=np.float64)), p(np.ascontiguousarray(rf, dtype=np.float64)),
                          p(data), p(nadd, ctypes.c_int))
        return data.reshape(nr, npix), nadd.reshape(nr, npix)


def static_split(n, nthr):
    """First iteration of each thread (and n) for libgomp's schedule(static)
    over n iterations: the first n % nthr threads take one more."""
    q, r = divmod(n, nthr)
    b = [t * q + min(t, r) for t in range(nthr)] + [n]
    return np.array(b, dtype=np.int32)


def ref_streams(seed, ipop, nthr, length, n_srcs=1):
    """The reference's per-thread MT19937 output streams of the source
    sampler: thread t seeded with seed + t + nthr (ipop + n_srcs IThread0),
    IThread0 = 0 on one rank (srcs.c:150)."""
    lib = load()
    out = np.zeros((nthr, length), dtype=np.uint32)
    for t in range(nthr):
        row = np.zeros(length, dtype=np.uint32)
        lib.orc_mt_fill((seed + t + nthr * ipop) & 0xFFFFFFFF, length, p(row, ctypes.c_uint32))
        out[t] = row
    return out


def mlp_forward(params, scaling, x):
    """The surrogate network in FP64: y (n x 2) from x (n x 4); scaling =
    (x_mean, x_scale, y_mean, y_scale)."""
    lib = load()
    P = np.ascontiguousarray(params, dtype=np.float64)
    sc = [np.ascontiguousarray(a, dtype=np.float64) for a in scaling]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros((len(x), 2))
    lib.orc_mlp_forward(p(P), p(sc[0]), p(sc[1]), p(sc[2]), p(sc[3]), p(x), len(x), p(y))
    return y


def mlp_loss_grad(params, scaling, x, y):
    """(loss, gradient[4610]) of the surrogate's mean squared error in FP64."""
    lib = load()
    P = np.ascontiguousarray(params, dtype=np.float64)
    sc = [np.ascontiguousarray(a, dtype=np.float64) for a in scaling]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    g = np.zeros(4610)
    loss = lib.orc_mlp_loss_grad(p(P), p(sc[0]), p(sc[1]), p(sc[2]), p(sc[3]), p(x), p(y), len(x), p(g))
    return loss, g
