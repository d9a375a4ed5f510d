"""ctypes bindings of the C ABI (include/colore_amd.h, include/colore_host.h).

The HIP library is required: importing this module on a machine where
libcolore_amd.so was not built raises immediately, so no caller can fall back
to a CPU path by accident.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
AMD_SO = os.path.join(_HERE, "libcolore_amd.so")
HOST_SO = os.path.join(_HERE, "libcolore_host.so")

CLR_NA = 5001
CLR_GRID_DENS, CLR_GRID_NPOT = 0, 1
CLR_TRACER_SRCS, CLR_TRACER_IMAP, CLR_TRACER_CSTM = 0, 1, 2
CLR_BIAS_EXPONENTIAL, CLR_BIAS_EXP_TRUNCATED, CLR_BIAS_TRUNCATED = 0, 1, 2
CLR_DENS_LOGNORMAL, CLR_DENS_CLIP = 0, 3
CLR_RNG_PHILOX, CLR_RNG_INJECTED = 0, 1
CLR_CL_KAPPA, CLR_CL_ISW = 0, 2
STAGES = ["gen", "fft_yz", "fft_x", "density", "norm", "srcs", "kappa", "isw", "imap", "src_pass", "src_write",
          "fft_phi", "shells", "beams", "cstm", "skewers", "phi_y", "phi_x"]

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.POINTER(ctypes.c_ulonglong)
_vp = ctypes.c_void_p


class Tables(ctypes.Structure):
    """clr_tables (include/colore_amd.h)."""
    _fields_ = [
        ("na", ctypes.c_int), ("glob_idr", ctypes.c_double),
        ("r_arr_r2z", _d), ("z_arr_r2z", _d), ("growth_d_arr", _d), ("growth_v_arr", _d),
        ("growth_pd_arr", _d), ("ihub_arr", _d),
        ("numk", ctypes.c_int), ("logkmin", ctypes.c_double), ("logkmax", ctypes.c_double),
        ("idlogk", ctypes.c_double), ("logkarr", _d), ("pkarr", _d), ("n_scal", ctypes.c_double),
        ("fgrowth_0", ctypes.c_double), ("hubble_0", ctypes.c_double), ("omega_m", ctypes.c_double),
        ("prefac_lensing", ctypes.c_double), ("r_max", ctypes.c_double), ("growth_d2_arr", _d),
    ]


class Csm(ctypes.Structure):
    """clr_csm (include/colore_host.h)."""
    _fields_ = [(n, ctypes.c_double) for n in ("OM", "OB", "OL", "OK", "h", "w0", "wa")] + \
               [(n, ctypes.c_int) for n in ("ksign", "normalDE", "constantw")] + \
               [(n, ctypes.c_double) for n in ("phorizon", "a_equality", "growth0")]


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


def load_amd():
    if not os.path.exists(AMD_SO):
        raise RuntimeError(f"HIP library {AMD_SO} is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(AMD_SO)
    ci = ctypes.c_int
    pp = ctypes.POINTER(_vp)
    _sig(lib, "clr_create", ci, [pp, ci, ctypes.c_double, _d, ci])
    _sig(lib, "clr_destroy", None, [_vp])
    _sig(lib, "clr_last_error", ctypes.c_char_p, [_vp])
    _sig(lib, "clr_version", ci, [])
    _sig(lib, "clr_device_bytes", ctypes.c_longlong, [_vp])
    _sig(lib, "clr_upload_tables", ci, [_vp, ctypes.POINTER(Tables)])
    _sig(lib, "clr_set_tracer", ci, [_vp, ci, ci, ci, _d, _d])
    _sig(lib, "clr_gaussian_fields", ci, [_vp, ctypes.c_uint32, ctypes.c_double, ci, ci, _u32, _d])
    _sig(lib, "clr_density", ci, [_vp, ci, ctypes.c_double])
    _sig(lib, "clr_shell_norms", ci, [_vp, ci, ctypes.c_double, ci, ctypes.POINTER(ci), _u64, _d, _d])
    _sig(lib, "clr_shell_norms_noalloc", ci, [_vp, ci, ctypes.c_double, ci, ctypes.POINTER(ci), _u64, _d, _d])
    _sig(lib, "clr_noalloc_max_bytes", ctypes.c_longlong, [])
    _sig(lib, "clr_get_norm_table", ci, [_vp, ci, ci, _d, _d, _d])
    _sig(lib, "clr_set_norm_table", ci, [_vp, ci, ci, _d, ctypes.c_double, ctypes.c_double])
    _sig(lib, "clr_sample_sources", ci, [_vp, ci, ctypes.c_uint32, ci, ctypes.POINTER(ctypes.c_longlong)])
    _sig(lib, "clr_fetch_sources", ci, [_vp, ci, _f, _i32, ctypes.c_longlong])
    _sig(lib, "clr_sample_sources_stream", ci, [_vp, ci, ci, _i32, _u32, ctypes.c_longlong, ci,
                                                ctypes.POINTER(ctypes.c_longlong)])
    _sig(lib, "clr_source_properties", ci, [_vp, ci, _f, ctypes.c_longlong])
    _sig(lib, "clr_catalogue_size", ci, [_vp, ci, ctypes.POINTER(ctypes.c_longlong)])
    _sig(lib, "clr_kappa", ci, [_vp, ci, ci, _d, _f])
    _sig(lib, "clr_isw", ci, [_vp, ci, ci, _d, _f])
    _sig(lib, "clr_source_beams", ci, [_vp, ci, _d, _f, _f, ctypes.c_longlong])
    _sig(lib, "clr_custom_map", ci, [_vp, ci, ci, _f])
    _sig(lib, "clr_lpt_density", ci, [_vp, ci, ci, _f])
    _sig(lib, "clr_source_skewers", ci, [_vp, ci, ci, ctypes.c_double, _f, _f, ctypes.c_longlong])
    _sig(lib, "clr_imap", ci, [_vp, ci, ci, ci, _d, _d, _f, _i32])
    _sig(lib, "clr_fetch_grid", ci, [_vp, ci, _f])
    _sig(lib, "clr_upload_grid", ci, [_vp, ci, _f])
    _sig(lib, "clr_fetch_grid_slab", ci, [_vp, ci, ci, ci, _f])
    _sig(lib, "clr_fetch_cell_counts", ci, [_vp, ci, ci, ci, _i32])
    _sig(lib, "clr_declare_grids", ci, [_vp, ctypes.c_double, ci])
    _sig(lib, "clr_stage_times", ci, [_vp, _d, ci])
    _sig(lib, "clr_mlp_create", ci, [ctypes.POINTER(_vp), ci, ctypes.c_uint32])
    _sig(lib, "clr_mlp_destroy", None, [_vp])
    _sig(lib, "clr_mlp_last_error", ctypes.c_char_p, [_vp])
    _sig(lib, "clr_mlp_set_scaling", ci, [_vp, _f, _f, _f, _f])
    _sig(lib, "clr_mlp_set_params", ci, [_vp, _f])
    _sig(lib, "clr_mlp_get_params", ci, [_vp, _f])
    _sig(lib, "clr_mlp_grad", ci, [_vp, _f, _f, ci, _f, _d])
    _sig(lib, "clr_mlp_train", ci, [_vp, _f, _f, ci, ci, ctypes.c_float, ctypes.c_float, _d])
    _sig(lib, "clr_mlp_predict", ci, [_vp, _f, ci, _f, ci])
    _sig(lib, "clr_synchronize", ci, [_vp])
    _sig(lib, "clr_host_alloc", _vp, [ctypes.c_size_t])
    _sig(lib, "clr_host_free", ci, [_vp])
    _sig(lib, "clr_bias_moment", ctypes.c_double, [ci] + [ctypes.c_double] * 5)
    _sig(lib, "clr_ensemble_create", ci, [ctypes.POINTER(_vp), ci, ctypes.POINTER(_vp), ci, ci, _d,
                                          ctypes.POINTER(ctypes.c_longlong), ci])
    _sig(lib, "clr_ensemble_destroy", None, [_vp])
    _sig(lib, "clr_ensemble_last_error", ctypes.c_char_p, [_vp])
    _sig(lib, "clr_ensemble_run", ci, [_vp, _u32, _d, ci, ctypes.POINTER(ci), _f])
    _sig(lib, "clr_ensemble_result", ci, [_vp, ci, _d, ctypes.POINTER(ci), _f])
    _sig(lib, "clr_fft_autotune", ci, [_vp, ci, ci, _f, ctypes.POINTER(ci), ctypes.POINTER(ci)])
    _sig(lib, "clr_fft_candidates", ci, [ci])
    _sig(lib, "clr_fft_select", ci, [ci, ci])
    _sig(lib, "clr_surrogate_inputs", ci, [_vp, ci, ci, _f, _d, _i32, ctypes.POINTER(ci)])
    _sig(lib, "clr_sample_sources_surrogate", ci, [_vp, _vp, ci, ctypes.c_uint32, ci, ci,
                                                   ctypes.POINTER(ctypes.c_longlong)])
    return lib


def load_host():
    if not os.path.exists(HOST_SO):
        raise RuntimeError(f"host library {HOST_SO} is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(HOST_SO)
    ci, cd = ctypes.c_int, ctypes.c_double
    _sig(lib, "clr_csm_set", ci, [ctypes.POINTER(Csm), cd, cd, cd, cd, cd, cd, ctypes.c_char_p, ci])
    for n in ("clr_csm_hubble", "clr_csm_omega_m", "clr_csm_particle_horizon",
              "clr_csm_radial_comoving_distance", "clr_csm_growth_factor", "clr_csm_f_growth"):
        _sig(lib, n, cd, [ctypes.POINTER(Csm), cd])
    _sig(lib, "clr_cspline_eval", None, [ci, _d, _d, ci, _d, _d])
    _sig(lib, "clr_sigma2_window", cd, [ci, _d, _d, cd, cd, cd, cd, cd, ci])
    _sig(lib, "clr_run_read", _vp, [ctypes.c_char_p, ctypes.c_char_p, ci])
    _sig(lib, "clr_run_free", None, [_vp])
    _sig(lib, "clr_run_scalar", cd, [_vp, ctypes.c_char_p])
    _sig(lib, "clr_run_tables", ci, [_vp, ctypes.POINTER(Tables)])
    _sig(lib, "clr_run_tracer", ci, [_vp, ci, ci, ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(ci)])
    _sig(lib, "clr_run_imap_shells", ci, [_vp, ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(_d),
                                          ctypes.POINTER(_d)])
    _sig(lib, "clr_run_get_bg", cd, [_vp, cd, ci, ci])
    _sig(lib, "clr_run_r_of_z", cd, [_vp, cd])
    _sig(lib, "clr_run_a2r", _d, [_vp])
    _sig(lib, "clr_hp_ang2pix_ring", ctypes.c_longlong, [ctypes.c_longlong, cd, cd])
    _sig(lib, "clr_hp_vec2pix_ring", ctypes.c_longlong, [ctypes.c_longlong, _d])
    _sig(lib, "clr_hp_ring2nest", ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong])
    _sig(lib, "clr_hp_nest2ring", ctypes.c_longlong, [ctypes.c_longlong, ctypes.c_longlong])
    _sig(lib, "clr_hp_pix2vec_nest", None, [ctypes.c_longlong, ctypes.c_longlong, _d])
    _sig(lib, "clr_write_density_grid", ci, [ctypes.c_char_p, ci, cd, _f])
    _sig(lib, "clr_write_healpix_map", ci, [ctypes.c_char_p, ctypes.c_longlong, _f])
    _sig(lib, "clr_write_catalog_fits", ci, [ctypes.c_char_p, ci, ctypes.c_longlong, _f, _f])
    _sig(lib, "clr_write_catalog_ascii", ci, [ctypes.c_char_p, ci, ctypes.c_longlong, _f, _f])
    _sig(lib, "clr_append_skewers_fits", ci, [ctypes.c_char_p, ci, ctypes.c_longlong, ci, _f, _f, _f, ctypes.c_float])
    _sig(lib, "clr_limber_cl", cd, [_vp, ci, cd, cd, cd, cd, cd, ci])
    _sig(lib, "clr_write_lpt_gadget", ci, [ctypes.c_char_p, ctypes.c_longlong, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _f])
    return lib


def ptr(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def pinned_empty(shape, dtype=np.float32):
    """An uninitialised array in page-locked host memory (clr_host_alloc),
    freed when the array and every view of it are gone.  Device-to-host
    copies into it skip the runtime's pageable staging: pass it as `out` to
    the map calls that are repeated (one allocation, many realisations)."""
    lib = load_amd()
    dt = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dt.itemsize
    p = lib.clr_host_alloc(max(nbytes, 1))
    if not p:
        raise MemoryError(f"clr_host_alloc({nbytes}) failed")
    buf = (ctypes.c_char * max(nbytes, 1)).from_address(p)
    weakref.finalize(buf, lib.clr_host_free, p)
    return np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)


def _map_out(out, nplanes, npix):
    """The caller's map buffer (C-contiguous float32 of nplanes x npix) or a
    new one."""
    if out is None:
        return np.zeros((nplanes, npix), dtype=np.float32)
    if out.dtype != np.float32 or out.shape != (nplanes, npix) or not out.flags.c_contiguous:
        raise ValueError(f"out must be a C-contiguous float32 array of shape {(nplanes, npix)}")
    return out


class Run:
    """A parsed param file with its FP64 tables (the reference's ParamCoLoRe)."""

    def __init__(self, param_file, host=None):
        self.host = host or load_host()
        err = ctypes.create_string_buffer(1024)
        self.h = self.host.clr_run_read(param_file.encode(), err, 1024)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.host.clr_run_free(self.h)
            self.h = None

    def scalar(self, name):
        return self.host.clr_run_scalar(self.h, name.encode())

    def tables(self):
        t = Tables()
        self.host.clr_run_tables(self.h, ctypes.byref(t))
        return t

    def table_arrays(self):
        t = self.tables()
        na = t.na
        g = lambda p, n=na: np.ctypeslib.as_array(p, shape=(n,)).copy()
        return dict(r=g(t.r_arr_r2z), z=g(t.z_arr_r2z), d1=g(t.growth_d_arr), v1=g(t.growth_v_arr),
                    pd=g(t.growth_pd_arr), ih=g(t.ihub_arr), logk=g(t.logkarr, t.numk),
                    pk=g(t.pkarr, t.numk), glob_idr=t.glob_idr, numk=t.numk, logkmin=t.logkmin,
                    logkmax=t.logkmax, idlogk=t.idlogk, ns=t.n_scal, fgrowth_0=t.fgrowth_0,
                    hubble_0=t.hubble_0, omega_m=t.omega_m, prefac_lensing=t.prefac_lensing,
                    r_max=t.r_max, d2=g(t.growth_d2_arr))

    def r_a2r(self):
        """r(a) at a = i/(NA-1): the table r_of_z interpolates (cosmo.c:101-112)."""
        return np.ctypeslib.as_array(self.host.clr_run_a2r(self.h), shape=(CLR_NA,)).copy()

    def imap_shells(self, ipop):
        """(nside, r0, rf) of intensity species ipop."""
        ns, nr, r0, rf = ctypes.c_int(), ctypes.c_int(), _d(), _d()
        if self.host.clr_run_imap_shells(self.h, ipop, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(r0),
                                         ctypes.byref(rf)):
            raise IndexError(ipop)
        n = nr.value
        return (ns.value, np.ctypeslib.as_array(r0, shape=(n,)).copy(),
                np.ctypeslib.as_array(rf, shape=(n,)).copy())

    def tracer(self, kind, ipop):
        a, b, m = _d(), _d(), ctypes.c_int()
        if self.host.clr_run_tracer(self.h, kind, ipop, ctypes.byref(a), ctypes.byref(b), ctypes.byref(m)):
            raise IndexError((kind, ipop))
        return (np.ctypeslib.as_array(a, shape=(CLR_NA,)).copy(),
                np.ctypeslib.as_array(b, shape=(CLR_NA,)).copy(), m.value)


class Context:
    """Device context (clr_ctx) with one method per reference stage."""

    def __init__(self, n_grid, l_box, pos_obs=None, device=0, amd=None):
        self.lib = amd or load_amd()
        self.N = int(n_grid)
        self.L = float(l_box)
        obs = np.ascontiguousarray(pos_obs if pos_obs is not None else [0.5 * l_box] * 3, dtype=np.float64)
        h = _vp()
        rc = self.lib.clr_create(ctypes.byref(h), self.N, self.L, ptr(obs), device)
        self.h = h
        if rc:
            msg = self.lib.clr_last_error(h).decode() if h else "clr_create failed"
            self.close()
            raise RuntimeError(msg)
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.clr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise RuntimeError(f"clr error {rc}: {self.lib.clr_last_error(self.h).decode()}")

    @property
    def padded_shape(self):
        return (self.N, self.N, 2 * (self.N // 2 + 1))

    def upload_tables(self, t):
        self._keep.append(t)
        self._chk(self.lib.clr_upload_tables(self.h, ctypes.byref(t)))

    def upload_table_arrays(self, ta):
        arrs = {k: np.ascontiguousarray(ta[k], dtype=np.float64) for k in ("r", "z", "d1", "v1", "pd", "ih", "logk", "pk")}
        self._keep.append(arrs)
        t = Tables(na=len(arrs["r"]), glob_idr=ta["glob_idr"], r_arr_r2z=ptr(arrs["r"]), z_arr_r2z=ptr(arrs["z"]),
                   growth_d_arr=ptr(arrs["d1"]), growth_v_arr=ptr(arrs["v1"]), growth_pd_arr=ptr(arrs["pd"]),
                   ihub_arr=ptr(arrs["ih"]), numk=len(arrs["pk"]), logkmin=ta["logkmin"], logkmax=ta["logkmax"],
                   idlogk=ta["idlogk"], logkarr=ptr(arrs["logk"]), pkarr=ptr(arrs["pk"]), n_scal=ta["ns"],
                   fgrowth_0=ta["fgrowth_0"], hubble_0=ta["hubble_0"], omega_m=ta["omega_m"],
                   prefac_lensing=ta["prefac_lensing"], r_max=ta["r_max"])
        self.upload_tables(t)

    def set_tracer(self, kind, ipop, model, amp, bz):
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        bz = np.ascontiguousarray(bz, dtype=np.float64)
        self._chk(self.lib.clr_set_tracer(self.h, kind, ipop, model, ptr(amp), ptr(bz)))

    def gaussian_fields(self, seed, r_smooth, smooth_potential=True, injected=None):
        s2 = ctypes.c_double()
        if injected is not None:
            inj = np.ascontiguousarray(injected, dtype=np.uint32)
            self._chk(self.lib.clr_gaussian_fields(self.h, seed, r_smooth, int(smooth_potential), CLR_RNG_INJECTED,
                                                   ptr(inj, ctypes.c_uint32), ctypes.byref(s2)))
        else:
            self._chk(self.lib.clr_gaussian_fields(self.h, seed, r_smooth, int(smooth_potential), CLR_RNG_PHILOX,
                                                   None, ctypes.byref(s2)))
        return s2.value

    def density(self, dens_type, sigma2):
        self._chk(self.lib.clr_density(self.h, dens_type, sigma2))

    def shell_norms(self, lognormal_fused=False, sigma2=0.0, ntracers=0, cap=256, noalloc=False):
        nz = ctypes.c_int()
        cnt = np.zeros(cap, dtype=np.uint64)
        zs = np.zeros(cap)
        sums = np.zeros(max(ntracers, 1) * cap)
        fn = self.lib.clr_shell_norms_noalloc if noalloc else self.lib.clr_shell_norms
        self._chk(fn(self.h, int(lognormal_fused), sigma2, cap, ctypes.byref(nz), ptr(cnt, ctypes.c_ulonglong),
                     ptr(zs), ptr(sums)))
        n = nz.value
        return n, cnt[:n].copy(), zs[:n].copy(), sums[:ntracers * n].reshape(ntracers, n).copy()

    def norm_table(self, kind, ipop):
        arr = np.zeros(CLR_NA)
        n0, nf = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.clr_get_norm_table(self.h, kind, ipop, ptr(arr), ctypes.byref(n0), ctypes.byref(nf)))
        return arr, n0.value, nf.value

    def set_norm_table(self, kind, ipop, arr, n0, nf):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.lib.clr_set_norm_table(self.h, kind, ipop, ptr(arr), n0, nf))

    def sample_sources(self, ipop, seed, nside_base=2):
        n = ctypes.c_longlong()
        self._chk(self.lib.clr_sample_sources(self.h, ipop, seed, nside_base, ctypes.byref(n)))
        return n.value

    def sample_sources_stream(self, ipop, nthr, iz_begin, streams, nside_base=2):
        """Reference-stream mode: streams is uint32[nthr, len], the raw
        outputs of each OpenMP thread's MT19937; iz_begin its plane split."""
        streams = np.ascontiguousarray(streams, dtype=np.uint32)
        izb = np.ascontiguousarray(iz_begin, dtype=np.int32)
        assert streams.shape[0] == nthr and len(izb) == nthr + 1
        n = ctypes.c_longlong()
        self._chk(self.lib.clr_sample_sources_stream(self.h, ipop, nthr, ptr(izb, ctypes.c_int32),
                                                     ptr(streams, ctypes.c_uint32), streams.shape[1], nside_base,
                                                     ctypes.byref(n)))
        return n.value

    def surrogate_inputs(self, ipop):
        """(x[n, 4], base[n], shell[n]): the surrogate's inputs and the
        cumulant closure's ln norm per normalisation shell."""
        n = ctypes.c_int()
        self._chk(self.lib.clr_surrogate_inputs(self.h, ipop, 0, None, None, None, ctypes.byref(n)))
        x = np.zeros((max(n.value, 1), 4), dtype=np.float32)
        base = np.zeros(max(n.value, 1))
        sh = np.zeros(max(n.value, 1), dtype=np.int32)
        self._chk(self.lib.clr_surrogate_inputs(self.h, ipop, n.value, ptr(x, ctypes.c_float), ptr(base),
                                                ptr(sh, ctypes.c_int32), ctypes.byref(n)))
        return x[:n.value], base[:n.value], sh[:n.value]

    def sample_sources_surrogate(self, net, ipop, seed, nside_base=2, precision=1):
        n = ctypes.c_longlong()
        self._chk(self.lib.clr_sample_sources_surrogate(self.h, net.h if net is not None else None, ipop, seed,
                                                        nside_base, precision,
                                                        ctypes.byref(n)))
        return n.value

    def fft_autotune(self, reps=3):
        """(best candidate, median ms per candidate) of the FFT tile autotuner."""
        ms = np.zeros(16, dtype=np.float32)
        nc, best = ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.clr_fft_autotune(self.h, reps, 16, ptr(ms, ctypes.c_float), ctypes.byref(nc),
                                            ctypes.byref(best)))
        return best.value, ms[:nc.value].copy()

    def fft_candidates(self):
        """Number of FFT tile-shape candidates for this context's grid size."""
        return int(self.lib.clr_fft_candidates(self.N))

    def fft_select(self, k):
        """Use FFT tile-shape candidate k for this grid size (process-wide)."""
        if self.lib.clr_fft_select(self.N, k) != 0:
            raise ValueError(f"no FFT candidate {k} for n_grid={self.N}")

    def catalogue_size(self, ipop):
        n = ctypes.c_longlong()
        self._chk(self.lib.clr_catalogue_size(self.h, ipop, ctypes.byref(n)))
        return n.value

    def fetch_sources(self, ipop, n):
        pos = np.zeros((n, 4), dtype=np.float32)
        ipix = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.clr_fetch_sources(self.h, ipop, ptr(pos, ctypes.c_float), ptr(ipix, ctypes.c_int32), n))
        return pos, ipix

    def source_properties(self, ipop, n):
        out = np.zeros((n, 4), dtype=np.float32)
        self._chk(self.lib.clr_source_properties(self.h, ipop, ptr(out, ctypes.c_float), n))
        return out

    def source_beams(self, ipop, n, r_a2r, lensing=True):
        """(dz_rsd[n], lens[n, 5] or None): the sources' line-of-sight pass."""
        a2r = np.ascontiguousarray(r_a2r, dtype=np.float64)
        dz = np.zeros(max(n, 1), dtype=np.float32)
        lens = np.zeros((max(n, 1), 5), dtype=np.float32) if lensing else None
        self._chk(self.lib.clr_source_beams(self.h, ipop, ptr(a2r), ptr(dz, ctypes.c_float),
                                            ptr(lens, ctypes.c_float) if lensing else None, n))
        return dz[:n], (lens[:n] if lensing else None)

    def source_skewers(self, ipop, n, gaussian=False, sigma2=0.0):
        """(dens_or_gauss[n, N/2], v[n, N/2]): the sources' skewers."""
        nr = self.N // 2
        dg = np.zeros((max(n, 1), nr), dtype=np.float32)
        vs = np.zeros((max(n, 1), nr), dtype=np.float32)
        self._chk(self.lib.clr_source_skewers(self.h, ipop, int(gaussian), float(sigma2), ptr(dg, ctypes.c_float),
                                              ptr(vs, ctypes.c_float), n))
        return dg[:n], vs[:n]

    def lpt_density(self, interp, order=1, positions=False):
        """1LPT / 2LPT density (rho - 1) in the density grid, replacing
        clr_density; with positions, returns the particles' float32 (N^3, 3)."""
        pos = np.zeros((self.N ** 3, 3), dtype=np.float32) if positions else None
        self._chk(self.lib.clr_lpt_density(self.h, order, interp, ptr(pos, ctypes.c_float) if positions else None))
        return pos

    def custom_map(self, ipop, nside):
        """Custom projected map ipop (NEST, float32[12 nside^2])."""
        out = np.zeros(12 * nside * nside, dtype=np.float32)
        self._chk(self.lib.clr_custom_map(self.h, ipop, nside, ptr(out, ctypes.c_float)))
        return out

    def kappa(self, nside, r_src, out=None):
        """Convergence maps at the source radii r_src, one row per plane, into
        `out` when given (e.g. a pinned_empty array reused across calls)."""
        r = np.ascontiguousarray(r_src, dtype=np.float64)
        out = _map_out(out, len(r), 12 * nside * nside)
        self._chk(self.lib.clr_kappa(self.h, nside, len(r), ptr(r), ptr(out, ctypes.c_float)))
        return out

    def isw(self, nside, r_src, out=None):
        """ISW maps at the source radii r_src; `out` as for kappa."""
        r = np.ascontiguousarray(r_src, dtype=np.float64)
        out = _map_out(out, len(r), 12 * nside * nside)
        self._chk(self.lib.clr_isw(self.h, nside, len(r), ptr(r), ptr(out, ctypes.c_float)))
        return out

    def imap(self, ipop, nside, r0, rf):
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        rf = np.ascontiguousarray(rf, dtype=np.float64)
        npix = 12 * nside * nside
        data = np.zeros((len(r0), npix), dtype=np.float32)
        nadd = np.zeros((len(r0), npix), dtype=np.int32)
        self._chk(self.lib.clr_imap(self.h, ipop, nside, len(r0), ptr(r0), ptr(rf), ptr(data, ctypes.c_float),
                                    ptr(nadd, ctypes.c_int32)))
        return data, nadd

    def fetch_grid(self, which):
        g = np.empty(self.padded_shape, dtype=np.float32)
        self._chk(self.lib.clr_fetch_grid(self.h, which, ptr(g, ctypes.c_float)))
        return g

    def fetch_grid_slab(self, which, iz0, nz):
        """Planes iz0 .. iz0+nz-1 of a grid, float32[nz][N][2(N/2+1)]."""
        g = np.empty((nz,) + self.padded_shape[1:], dtype=np.float32)
        self._chk(self.lib.clr_fetch_grid_slab(self.h, which, iz0, nz, ptr(g, ctypes.c_float)))
        return g

    def fetch_cell_counts(self, ipop, iz0, nz):
        """Per-cell source counts of population ipop for planes iz0 .. iz0+nz-1,
        int32[nz][N][N] (z, y, x)."""
        c = np.empty((nz, self.N, self.N), dtype=np.int32)
        self._chk(self.lib.clr_fetch_cell_counts(self.h, ipop, iz0, nz, ptr(c, ctypes.c_int32)))
        return c

    def upload_grid(self, which, g):
        g = np.ascontiguousarray(g, dtype=np.float32)
        assert g.shape == self.padded_shape, (g.shape, self.padded_shape)
        self._chk(self.lib.clr_upload_grid(self.h, which, ptr(g, ctypes.c_float)))

    def declare_grids(self, sigma2, density_transformed):
        self._chk(self.lib.clr_declare_grids(self.h, sigma2, int(density_transformed)))

    def synchronize(self):
        """Waits for the context's device work (phi's passes included)."""
        self._chk(self.lib.clr_synchronize(self.h))

    def stage_times(self):
        ms = np.zeros(len(STAGES))
        n = self.lib.clr_stage_times(self.h, ptr(ms), len(STAGES))
        return dict(zip(STAGES, ms[:n]))


CLR_MLP_FP32, CLR_MLP_BF16_MFMA = 0, 1


class Mlp:
    """The surrogate network (clr_mlp): 4 -> 64 -> 64 -> 2, trained and run on
    the device."""

    def __init__(self, seed=0, device=0, amd=None):
        self.lib = amd or load_amd()
        h = _vp()
        rc = self.lib.clr_mlp_create(ctypes.byref(h), device, seed)
        self.h = h
        if rc:
            msg = self.lib.clr_mlp_last_error(h).decode() if h else "clr_mlp_create failed"
            self.close()
            raise RuntimeError(msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.clr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise RuntimeError(f"clr_mlp error {rc}: {self.lib.clr_mlp_last_error(self.h).decode()}")

    def set_scaling(self, x_mean, x_scale, y_mean, y_scale):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x_mean, x_scale, y_mean, y_scale)]
        self.scaling = tuple(np.asarray(v, dtype=np.float64) for v in a)
        self._chk(self.lib.clr_mlp_set_scaling(self.h, *[ptr(v, ctypes.c_float) for v in a]))

    def set_params(self, params):
        P = np.ascontiguousarray(params, dtype=np.float32)
        assert P.shape == (4610,)
        self._chk(self.lib.clr_mlp_set_params(self.h, ptr(P, ctypes.c_float)))

    def params(self):
        P = np.zeros(4610, dtype=np.float32)
        self._chk(self.lib.clr_mlp_get_params(self.h, ptr(P, ctypes.c_float)))
        return P

    def grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(4610, dtype=np.float32)
        loss = ctypes.c_double()
        self._chk(self.lib.clr_mlp_grad(self.h, ptr(x, ctypes.c_float), ptr(y, ctypes.c_float), len(x),
                                        ptr(g, ctypes.c_float), ctypes.byref(loss)))
        return loss.value, g

    def train(self, x, y, steps, lr=3e-3, lr_final=3e-4):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.zeros(max(steps, 1))
        self._chk(self.lib.clr_mlp_train(self.h, ptr(x, ctypes.c_float), ptr(y, ctypes.c_float), len(x), steps,
                                         lr, lr_final, ptr(losses)))
        return losses[:steps]

    def predict(self, x, precision=CLR_MLP_BF16_MFMA):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 2), dtype=np.float32)
        self._chk(self.lib.clr_mlp_predict(self.h, ptr(x, ctypes.c_float), len(x), ptr(y, ctypes.c_float),
                                           precision))
        return y


class Ensemble:
    """Realisations (contexts) run as one captured HIP graph (clr_ensemble)."""

    def __init__(self, ctxs, src_cap, nside_kappa=0, r_kappa=None, nside_base=2, nstreams=4):
        self.lib = ctxs[0].lib
        self.ctxs = list(ctxs)
        n = len(ctxs)
        arr = (_vp * n)(*[c.h for c in ctxs])
        caps = np.ascontiguousarray(np.broadcast_to(np.asarray(src_cap, dtype=np.int64), (n,)))
        rk = np.ascontiguousarray(r_kappa if r_kappa is not None else np.zeros(n), dtype=np.float64)
        h = _vp()
        rc = self.lib.clr_ensemble_create(ctypes.byref(h), n, arr, nside_base, nside_kappa, ptr(rk),
                                          caps.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), nstreams)
        self.h = h
        self.nside_kappa = nside_kappa
        if rc:
            msg = self.lib.clr_ensemble_last_error(h).decode() if h else "clr_ensemble_create failed"
            self.close()
            raise RuntimeError(msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.clr_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, seeds, r_smooth, smooth_potential=True):
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(r_smooth, dtype=np.float64), (len(s),)))
        bad, ms = ctypes.c_int(), ctypes.c_float()
        rc = self.lib.clr_ensemble_run(self.h, ptr(s, ctypes.c_uint32), ptr(r), int(smooth_potential),
                                       ctypes.byref(bad), ctypes.byref(ms))
        if rc:
            raise RuntimeError(f"clr_ensemble error {rc}: {self.lib.clr_ensemble_last_error(self.h).decode()}")
        return bad.value, ms.value

    def result(self, i):
        s2, flag = ctypes.c_double(), ctypes.c_int()
        kap = np.zeros(12 * self.nside_kappa ** 2, dtype=np.float32) if self.nside_kappa else None
        rc = self.lib.clr_ensemble_result(self.h, i, ctypes.byref(s2), ctypes.byref(flag),
                                          ptr(kap, ctypes.c_float) if kap is not None else None)
        if rc:
            raise RuntimeError(self.lib.clr_ensemble_last_error(self.h).decode())
        return s2.value, flag.value, kap
