"""Synthetic inputs and param files for the benchmark configurations.

Every config is flat LCDM (Om=0.3, Ob=0.05, h=0.7, sigma8=0.8, ns=0.96) with a
linear P(k) from the Eisenstein & Hu (1998, ApJ 496, 605) no-wiggle transfer
function (their eqs. 26, 28-31), 1000 log-spaced k in [1e-4, 1e2] h/Mpc,
written in the two-column CAMB format the reference reads (cosmo.c:459-463);
its amplitude is arbitrary since the run renormalises to sigma8.

n(z) files hold dN/dz dOmega in deg^-2 as the reference expects
(cosmo.c:575-578): dN/dzdOmega = nbar(z) chi(z)^2 / H(z) / RTOD^2, where nbar
is the requested comoving density ("normalised to X (Mpc/h)^-3" is read as the
peak density).  Every z-table extends past the redshift of the box corner,
because the reference's spline returns NaN outside the tabulated range.
"""
import math
import os

import numpy as np

RTOD = 57.2957795
COSMO = dict(omega_M=0.3, omega_L=0.7, omega_B=0.05, h=0.7, w=-1.0, ns=0.96, sigma_8=0.8)


def eh_nowiggle_pk(k_hmpc, om=0.3, ob=0.05, h=0.7, ns=0.96, tcmb=2.7255):
    """Unnormalised P(k) = k^ns T0(k)^2 with the EH98 no-wiggle T0."""
    k = np.asarray(k_hmpc, dtype=np.float64)
    theta = tcmb / 2.7
    omh2, obh2 = om * h * h, ob * h * h
    fb = ob / om
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1 + 10 * obh2 ** 0.75)                 # eq. 26, Mpc
    alpha = 1 - 0.328 * np.log(431 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb * fb  # eq. 31
    kmpc = k * h
    gamma = om * h * (alpha + (1 - alpha) / (1 + (0.43 * kmpc * s) ** 4))               # eq. 30
    q = k * theta * theta / gamma                                                        # eq. 28
    l0 = np.log(2 * math.e + 1.8 * q)
    c0 = 14.2 + 731.0 / (1 + 62.5 * q)                                                   # eq. 29
    t0 = l0 / (l0 + c0 * q * q)
    return k ** ns * t0 * t0


def _hub(a, om=0.3):
    return math.sqrt(om / a ** 3 + (1 - om)) / 2997.92458


def _chi(z, om=0.3, n=4000):
    zz = np.linspace(0, z, n + 1)
    f = 1.0 / np.sqrt(om * (1 + zz) ** 3 + (1 - om))
    return 2997.92458 * np.trapezoid(f, zz) if z > 0 else 0.0


def _growth(z, om=0.3):
    """Linear growth normalised to 1 at z=0 (integral solution for LCDM)."""
    def d(a):
        aa = np.linspace(1e-6, a, 4001)
        e = np.sqrt(om / aa ** 3 + (1 - om))
        return 2.5 * om * np.sqrt(om / a ** 3 + 1 - om) * np.trapezoid(1 / (aa * e) ** 3, aa)
    return d(1 / (1 + z)) / d(1.0)


def _z_of_chi(chi, om=0.3):
    lo, hi = 0.0, 20.0
    for _ in range(80):
        m = 0.5 * (lo + hi)
        if _chi(m, om) < chi:
            lo = m
        else:
            hi = m
    return hi


def z_table_top(z_max, n_grid):
    r_max = _chi(z_max)
    lbox = 2 * r_max * (1 + 2.0 / n_grid)
    return _z_of_chi(math.sqrt(3) * 0.5 * lbox * 1.02) + 0.1


def write_two_col(path, x, y):
    np.savetxt(path, np.column_stack([x, y]), fmt="%.10e")


def _nz_file(path, zs, nbar):
    chi = np.array([_chi(z) for z in zs])
    hz = np.array([_hub(1 / (1 + z)) for z in zs])
    write_two_col(path, zs, nbar * chi ** 2 / hz / RTOD ** 2)


def gaussian_nbar(zs, mean, sigma, peak):
    return peak * np.exp(-0.5 * ((zs - mean) / sigma) ** 2)


def flat_nbar(zs, z0, z1, level, edge=0.005):
    return level / ((1 + np.exp(-(zs - z0) / edge)) * (1 + np.exp((zs - z1) / edge)))


# populations: (nbar(z) spec, bias spec, bias model name)
CONFIGS = {
    "C1": dict(n_grid=128, z_min=0.05, z_max=0.5, r_smooth=2.0, seed=1001,
               srcs=[(("gauss", 0.3, 0.08, 1e-3), ("const", 1.0), "exponential")]),
    "C2": dict(n_grid=512, z_min=0.05, z_max=1.2, r_smooth=2.0, seed=2002,
               srcs=[(("gauss", 0.5, 0.15, 5e-4), ("1+z",), "exponential"),
                     (("flat", 0.8, 1.1, 1e-4), ("const", 2.0), "exponential")],
               kappa=dict(z_out=[1.0], nside=256)),
    "C3": dict(n_grid=1024, z_min=0.05, z_max=1.6, r_smooth=1.0, seed=3003,
               srcs=[(("gauss", 0.4, 0.1, 1e-3), ("0.95/D",), "exponential"),
                     (("gauss", 0.9, 0.15, 3e-4), ("const", 1.5), "exp-truncated"),
                     (("gauss", 1.3, 0.2, 1e-4), ("const", 2.2), "truncated")],
               kappa=dict(z_out=[0.5, 1.0, 1.5], nside=512), isw=dict(z_out=[0.5, 1.0, 1.5], nside=512),
               imap=dict(nside=512, nbands=64, t0=0.2)),
    "C4": dict(n_grid=128, z_min=0.0, z_max=0.1, r_smooth=2.0, seed=4000,
               srcs=[(("gauss", 0.5, 0.1, 5e-4), ("1+z",), "exponential")]),
    "C5": dict(n_grid=2048, z_min=0.05, z_max=2.0, r_smooth=1.0, seed=5005,
               srcs=[(("gauss", 0.5, 0.15, 5e-4), ("1+z",), "exponential"),
                     (("flat", 0.8, 1.1, 1e-4), ("const", 2.0), "exponential")]),
}


def c4_realisations(n=64, seed0=4000):
    """The batched-ensemble workload: n realisations with seeds seed0.. and grid
    sizes drawn uniformly from {128, 192, 256, 384} (NumPy generator seeded
    with seed0); the box is L = 4 N Mpc/h (cell 4 Mpc/h), so z_max is the
    redshift of L/2/(1+2/N) (cosmo.c:544 inverted); one population with
    b = 1+z and a Gaussian n(z), mean drawn in [0.3, 0.8], sigma 0.1, peak
    5e-4 (Mpc/h)^-3.  The kappa plane of z = 0.8 lies beyond every such box
    (z_max 0.086 .. 0.27; the reference stops there, cosmo.c:767-768): it is
    placed at the box edge r_max instead.  Returns a list of dicts."""
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n):
        N = int(rng.choice([128, 192, 256, 384]))
        mean = float(rng.uniform(0.3, 0.8))
        lbox = 4.0 * N
        zmax = _z_of_chi(lbox / 2 / (1 + 2.0 / N))
        out.append(dict(n_grid=N, seed=seed0 + i, z_max=zmax, nz_mean=mean, l_box=lbox))
    return out


def write_config(name, outdir, n_grid=None, z_max=None, dens_type=0, output_format="ASCII",
                 output_density=False, with_maps=True, seed=None, imap_bands=None, imap_nside=None,
                 kappa_nside=None, lensing=(), skewers=(), custom=(), pred_dz=None, just_pred=False,
                 output_lpt=False, sigma8=None, r_smooth=None, populations=None):
    """Writes P(k), n(z), b(z) (+T(z), nu table) and the param file of config
    `name` into outdir; returns the param-file path.  n_grid/z_max override
    the config's values (small variants for tests); `lensing` lists the
    source populations with include_lensing, `skewers` maps populations to
    store_skewers ("density" or "gaussian"); `custom` lists (z0, sigma_z,
    nside, bias) of custom maps with a Gaussian kernel K(z); pred_dz turns on
    write_pred (just_pred: just_write_pred)."""
    cfg = dict(CONFIGS[name])
    if r_smooth is not None:
        cfg["r_smooth"] = r_smooth
    if populations is not None:
        cfg["srcs"] = list(populations)
    cosmo = dict(COSMO)
    if sigma8 is not None:
        cosmo["sigma_8"] = sigma8
    os.makedirs(outdir, exist_ok=True)
    n = int(n_grid or cfg["n_grid"])
    zmax = float(z_max or cfg["z_max"])
    zmin = cfg["z_min"]
    k = np.logspace(-4, 2, 1000)
    write_two_col(os.path.join(outdir, "pk.txt"), k, eh_nowiggle_pk(k))
    ztop = z_table_top(zmax, n)
    zs = np.linspace(0.0, ztop, 1500)
    lines = ["global: {", f'  prefix_out = "{outdir}/out";', f'  output_format = "{output_format}";',
             f"  output_density = {'true' if output_density else 'false'};",
             f'  pk_filename = "{outdir}/pk.txt";', f"  z_min = {zmin};", f"  z_max = {zmax};",
             f"  seed = {int(seed if seed is not None else cfg['seed'])};"] + \
            ([f"  write_pred = true;", f"  just_write_pred = {'true' if just_pred else 'false'};",
              f"  pred_dz = {pred_dz};"] if pred_dz else ["  write_pred = false;", "  just_write_pred = false;"]) + \
            ["};", "field_par: {", f"  r_smooth = {cfg['r_smooth']};",
             "  smooth_potential = true;", f"  n_grid = {n};", f"  dens_type = {dens_type};",
             "  lpt_buffer_fraction = 0.6;", "  lpt_interp_type = 1;", f"  output_lpt = {int(output_lpt)};", "};",
             "cosmo_par: {"] + [f"  {kk} = {vv};" for kk, vv in cosmo.items()] + ["};"]
    for i, (nzs, bzs, model) in enumerate(cfg["srcs"]):
        if nzs[0] == "gauss":
            nbar = gaussian_nbar(zs, nzs[1], nzs[2], nzs[3])
        else:
            nbar = flat_nbar(zs, nzs[1], nzs[2], nzs[3])
        _nz_file(os.path.join(outdir, f"nz{i + 1}.txt"), zs, nbar)
        if bzs[0] == "const":
            b = np.full_like(zs, bzs[1])
        elif bzs[0] == "1+z":
            b = 1 + zs
        elif bzs[0] == "lin":
            b = bzs[1] + bzs[2] * zs
        else:
            b = 0.95 / np.array([_growth(z) for z in zs])
        write_two_col(os.path.join(outdir, f"bz{i + 1}.txt"), zs, b)
        lines += [f"srcs{i + 1}: {{", f'  nz_filename = "{outdir}/nz{i + 1}.txt";',
                  f'  bias_filename = "{outdir}/bz{i + 1}.txt";',
                  f"  include_lensing = {'true' if i in lensing else 'false'};",
                  f"  store_skewers = {'true' if i in skewers else 'false'};"]
        if i in skewers:
            lines.append(f"  gaussian_skewers = {'true' if skewers[i] == 'gaussian' else 'false'};")
        lines += [f'  bias_model = "{model}";', "};"]
    for i, (z0, sz, ns, bc) in enumerate(custom):
        write_two_col(os.path.join(outdir, f"kz_cstm{i + 1}.txt"), zs, np.exp(-0.5 * ((zs - z0) / sz) ** 2))
        write_two_col(os.path.join(outdir, f"bz_cstm{i + 1}.txt"), zs, np.full_like(zs, bc))
        lines += [f"custom{i + 1}: {{", f'  kz_filename = "{outdir}/kz_cstm{i + 1}.txt";',
                  f'  bias_filename = "{outdir}/bz_cstm{i + 1}.txt";', f"  nside = {ns};",
                  '  bias_model = "exponential";', "};"]
    if with_maps and "imap" in cfg:
        im = cfg["imap"]
        nb = imap_bands or im["nbands"]
        nu_rest = 1420.405
        nu_lo, nu_hi = nu_rest / (1 + zmax), nu_rest / (1 + zmin)
        edges = np.linspace(nu_lo, nu_hi, nb + 1)
        write_two_col(os.path.join(outdir, "nu.txt"), edges[:-1], edges[1:])
        write_two_col(os.path.join(outdir, "tz.txt"), zs, im["t0"] * np.sqrt(1 + zs))
        write_two_col(os.path.join(outdir, "bz_imap.txt"), zs, np.ones_like(zs))
        lines += ["imap1: {", f'  tbak_filename = "{outdir}/tz.txt";', f'  bias_filename = "{outdir}/bz_imap.txt";',
                  f'  freq_list = "{outdir}/nu.txt";', f"  freq_rest = {nu_rest};",
                  f"  nside = {imap_nside or im['nside']};", '  bias_model = "exponential";', "};"]
    for sec in ("kappa", "isw"):
        if with_maps and sec in cfg:
            zo = [z for z in cfg[sec]["z_out"] if z <= zmax]
            if zo:
                ns = kappa_nside or cfg[sec]["nside"]
                lines += [f"{sec}: {{", f"  z_out = [{', '.join(str(z) for z in zo)}];", f"  nside = {ns};", "};"]
    path = os.path.join(outdir, "params.cfg")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
