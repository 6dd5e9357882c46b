"""ctypes front end of the FP64 CPU oracle (csp_oracle.c).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline leg of the benchmark use
it as the checker of the GPU path; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_find_interval.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_find_interval.restype = C.c_int
        L.orc_expi.argtypes = [C.c_double]
        L.orc_expi.restype = C.c_double
        L.orc_sfh_weights.argtypes = [_dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_sfh_weights.restype = C.c_int
        L.orc_set_sfh_tab.argtypes = [C.c_int, _dp, _dp]
        L.orc_set_sfh_tab.restype = C.c_int
        L.orc_sfh_weights_tangent.argtypes = [_dp, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_sfh_weights_tangent.restype = C.c_int
        L.orc_sfh_gradients.argtypes = [_dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_sfh_gradients.restype = C.c_int
        L.orc_igm_tau.argtypes = [C.c_double, C.c_double]
        L.orc_igm_tau.restype = C.c_double
        L.orc_igm_pixel_terms.argtypes = [C.c_double, _dp]
        L.orc_igm_z_terms.argtypes = [C.c_double, _dp]
        L.orc_lum_distance.argtypes = [C.c_double] * 4
        L.orc_lum_distance.restype = C.c_double
        L.orc_cosmology_table.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_distance_term.argtypes = [_dp, _dp, C.c_int, C.c_double]
        L.orc_distance_term.restype = C.c_double
        L.orc_velocity_kernel.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, _ip, _ip, _ip, _dp,
                                          C.c_int]
        L.orc_velocity_kernel.restype = C.c_int
        L.orc_observed_frame.argtypes = [_dp, C.c_int, C.c_double, _dp, _dp]
        L.orc_band_supports.argtypes = [_dp, C.c_int, C.c_int, _ip, _ip]
        L.orc_band_fluxes_col.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _ip, _dp, _dp, _dp]
        L.orc_filter_band.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_filter_band.restype = C.c_int
        L.orc_csp_batch.restype = C.c_int
        L.orc_csp_batch_grad.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_powerlaw_bin.argtypes = [C.c_double] * 4
        L.orc_powerlaw_bin.restype = C.c_double
        L.orc_imf_integral.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_int]
        L.orc_imf_integral.restype = C.c_double
        L.orc_imf_bin_logs.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_imf_bin_weights_tab.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.orc_remnant_mass.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_remnant_mass.restype = C.c_double
        _lib = L
    return _lib


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def find_interval(a, v) -> int:
    a = _d(a)
    return lib().orc_find_interval(a, a.size, float(v))


def expi(x) -> float:
    return lib().orc_expi(float(x))


def sfh_weights(time_full, p, tiny_logt=0.0):
    tf = _d(time_full)
    w = np.zeros(tf.size)
    t1 = np.zeros(tf.size)
    t2 = np.zeros(tf.size)
    st = lib().orc_sfh_weights(tf, tf.size, float(tiny_logt), _d(p), w, t1, t2)
    return w, st


def set_sfh_tab(age_gyr, sfr):
    """The oracle's SFH table for sfh = 2, 3 (forward ages in Gyr, SFR in M_sun / yr; stored in years)."""
    a = _d(np.asarray(age_gyr, float) * 1.0e9)
    st = lib().orc_set_sfh_tab(a.size, a, _d(sfr))
    assert st == 0


# the SFH tangents' parameters in the order of sfh_weights_tangent's q = 1 .. 8
SFHT_PARAMS = ["tage", "tau", "const", "fburst", "tburst", "sf_start", "sf_trunc", "sf_slope"]


def sfh_weights_tangent(time_full, p, q, tiny_logt=0.0):
    """The SFH weights and their derivative with respect to SFHT_PARAMS[q - 1] (per Gyr). Returns w, dw, status."""
    tf = _d(time_full)
    w = np.zeros(tf.size)
    dw = np.zeros(tf.size)
    s = np.zeros(4 * tf.size)
    st = lib().orc_sfh_weights_tangent(tf, tf.size, float(tiny_logt), _d(p), int(q), w, dw, s)
    return w, dw, st


def sfh_gradients(time_full, p, wbar, tiny_logt=0.0):
    """params_bar's SFH rows (NBGRAD, zero elsewhere) of one model from d L / d w. Returns (pb, status)."""
    tf = _d(time_full)
    pb = np.zeros(NBGRAD)
    s = np.zeros(6 * tf.size)
    st = lib().orc_sfh_gradients(tf, tf.size, float(tiny_logt), _d(p), _d(wbar), pb, s)
    return pb, st


def igm_tau(w, opz) -> float:
    return lib().orc_igm_tau(float(w), float(opz))


def igm_pixel_terms(w):
    a = np.zeros(6)
    lib().orc_igm_pixel_terms(float(w), a)
    return a


def igm_z_terms(opz):
    b = np.zeros(6)
    lib().orc_igm_z_terms(float(opz), b)
    return b


def lum_distance(h0, om0, ol0, z) -> float:
    return lib().orc_lum_distance(float(h0), float(om0), float(ol0), float(z))


def cosmology_table(h0=72.0, om0=0.27, ol0=0.73):
    cz = np.zeros(500)
    cdl = np.zeros(500)
    lib().orc_cosmology_table(float(h0), float(om0), float(ol0), cz, cdl)
    return cz, cdl


def distance_term(cz, cdl, z) -> float:
    cz = _d(cz)
    return lib().orc_distance_term(cz, _d(cdl), cz.size, float(z))


def velocity_kernel(lam, sigma, minl=1.0e3, maxl=1.0e4):
    lam = _d(lam)
    n = lam.size
    lo = np.zeros(n, np.int32)
    hi = np.zeros(n, np.int32)
    off = np.zeros(n, np.int32)
    cap = n * 64
    coef = np.zeros(cap)
    nc = lib().orc_velocity_kernel(lam, n, float(sigma), float(minl), float(maxl), lo, hi, off, coef, cap)
    if nc < 0:
        raise RuntimeError("velocity kernel larger than its buffer")
    return lo, hi, off, coef[:nc]


def observed_frame(lam, z, total):
    lam = _d(lam)
    f = np.zeros(lam.size)
    lib().orc_observed_frame(lam, lam.size, float(z), _d(total), f)
    return f


def band_supports(bands):
    """bands: (nbands, nspec) C order (= Fortran (nspec, nbands))."""
    b = _d(bands)
    nb, ns = b.shape
    jlo = np.zeros(nb, np.int32)
    jhi = np.zeros(nb, np.int32)
    lib().orc_band_supports(b, ns, nb, jlo, jhi)
    return jlo, jhi


def band_fluxes_col(bands_t, jlo, jhi, do_band, wt, fobs):
    """bands_t: (nspec, nbands) C order (= Fortran bands_t(nbands, nspec))."""
    bt = _d(bands_t)
    ns, nb = bt.shape
    out = np.zeros(nb)
    lib().orc_band_fluxes_col(ns, nb, bt, _i(jlo), _i(jhi), _i(do_band), _d(wt), _d(fobs), out)
    return out


def filter_band(lc, rc, grid, mode=0, detector=0):
    g = _d(grid)
    band = np.zeros(g.size)
    lc = _d(lc)
    st = lib().orc_filter_band(lc, _d(rc), lc.size, g, g.size, int(mode), int(detector), band)
    return band, st


def build_bands(fset, grid, mode=0):
    """(nbands, nspec) normalised bands of a synth.FilterSet on grid."""
    out = np.zeros((len(fset.lam), len(grid)))
    for b, (lc, rc) in enumerate(zip(fset.lam, fset.resp)):
        out[b], _ = filter_band(lc, rc, grid, mode, fset.detector)
    return out


# ---- fixed attenuation curves (numpy restatement from the papers) ----
def calzetti_curve(lam):
    """A_lambda / A_V of Calzetti et al. (2000) eq. 4 (k' = 2.659, R_V = 4.05), the UV polynomial below 0.63 um
    and the linear law above, both carried beyond 0.12-2.2 um, k clipped at 0 (opensps_dust.f90:567-594)."""
    x = 1.0e4 / np.asarray(lam, float)
    k = np.where(np.asarray(lam) > 6300.0, 2.659 * (-1.857 + 1.040 * x) + 4.05,
                 4.05 + 2.659 * (-2.156 + 1.509 * x - 0.198 * x ** 2 + 0.011 * x ** 3))
    return np.maximum(k, 0.0) / 4.05


def reddy_curve(lam):
    """A_lambda / A_V of Reddy et al. (2015) (R_V = 2.505): their UV law below 0.60 um (x clipped at 1/0.15 um), the
    optical law to 2.85 um with the reference's continuity offset -0.036221981, zero beyond
    (opensps_dust.f90:630-662)."""
    lam = np.asarray(lam, float)
    x = 1.0e4 / lam
    xu = np.minimum(x, 1.0e4 / 1500.0)
    uv = -5.726 + 4.004 * xu - 0.525 * xu ** 2 + 0.029 * xu ** 3 + 2.505
    op = -2.672 - 0.010 * x + 1.532 * x ** 2 - 0.412 * x ** 3 + 2.505 - 0.036221981
    return np.where(lam < 6000.0, uv, np.where(lam < 28500.0, op, 0.0)) / 2.505


# ---- the reference spectra and the Vega system (numpy restatement) ----
TINY = np.finfo(np.float64).tiny


def ref_spectrum_on_grid(lt, ft, flambda, grid):
    """load_standard_sed (src/io/opensps_io.f90:1516-1582): log10(flux + tiny) linear in log10(lambda) with the
    end intervals' slopes beyond the table (interpolate_linear, src/math/opensps_interpolation.f90:87-130), then
    10^; an F_lambda table times lambda^2; tiny redward of the table."""
    x, y = np.log10(np.asarray(lt, float)), np.log10(np.asarray(ft, float) + TINY)
    g = np.asarray(grid, float)
    xo = np.log10(g)
    i = np.clip(np.searchsorted(x, xo, side="right") - 1, 0, x.size - 2)
    i = np.where(xo == x[-1], x.size - 2, i)
    slope = (y[i + 1] - y[i]) / (x[i + 1] - x[i])
    spec = 10.0 ** (y[i] + slope * (xo - x[i]))
    if flambda:
        spec = spec * g ** 2
    return np.where(g > np.max(lt), TINY, spec)


def zero_points(grid, spec, bands):
    """filters_zero_points (src/spectra/opensps_filters.f90:536-556): -2.5 log10(trapezoid(f_nu T / lambda)) - 48.60
    per band (bands (nbands, nspec)), 99 without flux."""
    g = np.asarray(grid, float)
    out = np.empty(len(bands))
    for b, t in enumerate(bands):
        y = spec * t / g
        v = float(np.sum(0.5 * (g[1:] - g[:-1]) * (y[1:] + y[:-1])))
        out[b] = -2.5 * np.log10(v) - 48.60 if v > TINY else 99.0
    return out


def sun_to_vega(magsun, magvega, vref):
    """sun_to_vega_system (src/spectra/opensps_context_filters.f90:185-198)."""
    ms = np.array(magsun, float)
    s0 = ms[vref - 1] if vref > 0 else 0.0
    for i in range(ms.size):
        if ms[i] == 99.0 or magvega[i] == 99.0:
            continue
        ms[i] = (ms[i] - s0) - (magvega[i] - magvega[vref - 1]) + s0 if vref > 0 else ms[i] - magvega[i]
    return ms


def vega_mags(mags_ab, magvega, vref):
    """The Vega system on AB magnitudes (n, nbands) (magnitudes_from_fluxes, src/spectra/opensps_batch_photometry.f90:
    185-232): vref = 0: AB - magvega; vref = k: band k stays AB, the others AB - magvega + magvega(k), NaN where the
    model's band k is NaN."""
    m = np.array(mags_ab, float)
    mv = np.asarray(magvega, float)
    if vref == 0:
        return m - mv[None, :]
    out = m - mv[None, :] + mv[vref - 1]
    out[~np.isfinite(m[:, vref - 1])] = np.nan
    out[:, vref - 1] = m[:, vref - 1]
    return out


# The packed IMF parameters of the oracle (csp_oracle.c ORC_IMF_*)
IMF_SALP, IMF_ALPHA, IMF_VDMC, IMF_MDAVE, IMF_LOWER_LIMIT, IMF_UPPER_LIMIT, IMF_LOWER_BOUND, IMF_MLIM_BH, \
    IMF_MLIM_NS, IMF_NPAR = 0, 1, 4, 5, 6, 7, 8, 9, 10, 11


def imf_vector(slopes=(1.3, 2.3, 2.3), lower=0.08, upper=120.0, salp=2.35, mdave=0.5, mlim_bh=40.0, mlim_ns=8.5):
    p = np.zeros(IMF_NPAR)
    p[IMF_SALP] = salp
    p[IMF_ALPHA:IMF_ALPHA + 3] = slopes
    p[IMF_MDAVE] = mdave
    p[IMF_LOWER_LIMIT] = lower
    p[IMF_UPPER_LIMIT] = upper
    p[IMF_LOWER_BOUND] = lower
    p[IMF_MLIM_BH] = mlim_bh
    p[IMF_MLIM_NS] = mlim_ns
    return p


def powerlaw_bin(plo, lo, hi, e) -> float:
    return lib().orc_powerlaw_bin(float(plo), float(lo), float(hi), float(e))


def imf_integral(imf_type, p, m1, m2, mass_weighted) -> float:
    return lib().orc_imf_integral(int(imf_type), _d(p), float(m1), float(m2), int(mass_weighted))


def imf_bin_logs(m1, m2):
    a, b = C.c_double(), C.c_double()
    lib().orc_imf_bin_logs(float(m1), float(m2), C.byref(a), C.byref(b))
    return a.value, b.value


def imf_bin_weights_tab(imf_type, p, total, mini, m1, m2, first):
    """Weights of one run of bins (logs formed by imf_bin_logs)."""
    mini, m1, m2 = _d(mini), _d(m1), _d(m2)
    llo = np.zeros(mini.size)
    dl = np.zeros(mini.size)
    for k in range(mini.size):
        llo[k], dl[k] = imf_bin_logs(m1[k], m2[k])
    w = np.zeros(mini.size)
    lib().orc_imf_bin_weights_tab(int(imf_type), _d(p), float(total), mini.size, mini, m1, m2, llo, dl, _i(first), w)
    return w


def remnant_mass(imf_type, p, max_living_mass) -> float:
    return lib().orc_remnant_mass(int(imf_type), _d(p), float(max_living_mass))


class _Ops(C.Structure):
    _fields_ = [("nz", C.c_int), ("nt", C.c_int), ("npts", C.c_int), ("ntpl", C.c_int), ("nent", C.c_int),
                ("age_ptr", C.c_void_p), ("mini", C.c_void_p), ("m1", C.c_void_p), ("m2", C.c_void_p),
                ("first", C.c_void_p), ("cm", C.c_void_p), ("cl", C.c_void_p), ("q", C.c_void_p),
                ("age_mmax", C.c_void_p), ("ent_tpl", C.c_void_p), ("ent_val", C.c_void_p), ("etpl", C.c_void_p)]


class _Ctx(C.Structure):
    _fields_ = [("nspec", C.c_int), ("nt", C.c_int), ("nz", C.c_int), ("nbands", C.c_int),
                ("lambda_", C.c_void_p), ("time_full", C.c_void_p), ("logz", C.c_void_p),
                ("basis", C.c_void_p), ("bmass", C.c_void_p), ("blbol", C.c_void_p), ("bands", C.c_void_p),
                ("mag_mask", C.c_void_p), ("ncosmo", C.c_int), ("cz", C.c_void_p), ("cdl", C.c_void_p),
                ("add_igm", C.c_int), ("smooth_fast", C.c_int), ("tiny_logt", C.c_double),
                ("ops", C.c_void_p), ("imf_type", C.c_int), ("remnants", C.c_int), ("salp_ind", C.c_double),
                ("mdave", C.c_double), ("mlim_bh", C.c_double), ("mlim_ns", C.c_double), ("op_count", C.c_void_p),
                ("ssp_weights_z", C.c_void_p), ("dcurve", C.c_void_p), ("zred_mags", C.c_void_p),
                ("add_dust_emission", C.c_int), ("add_agn_dust", C.c_int), ("nqpah", C.c_int), ("numin", C.c_int),
                ("ntau", C.c_int), ("qpah", C.c_void_p), ("umin", C.c_void_p), ("dustem", C.c_void_p),
                ("agn_tau", C.c_void_p), ("agn", C.c_void_p), ("mdust_out", C.c_void_p),
                ("add_neb", C.c_int), ("neb_cont_on", C.c_int), ("neb_lines_in_spec", C.c_int), ("nnz", C.c_int),
                ("nna", C.c_int), ("nnu", C.c_int), ("nline", C.c_int), ("neb_logz", C.c_void_p),
                ("neb_lage", C.c_void_p), ("neb_logu", C.c_void_p), ("neb_cont", C.c_void_p), ("neb_line", C.c_void_p),
                ("neb_lpos", C.c_void_p), ("emlines_out", C.c_void_p)]


def _context(grid, bands, add_igm, smooth_fast, mag_mask, cosmo, ptr):
    cz, cdl = cosmo if cosmo is not None else cosmology_table()
    ctx = _Ctx()
    ctx.nspec, ctx.nt, ctx.nz, ctx.nbands = grid.nspec, grid.nt, grid.nz, bands.shape[0]
    ctx.lambda_ = ptr(_d(grid.lam))
    ctx.time_full = ptr(_d(grid.log_age))
    ctx.logz = ptr(_d(grid.logz))
    ctx.basis = ptr(_d(grid.spec))
    ctx.bmass = ptr(_d(grid.mass))
    ctx.blbol = ptr(_d(grid.lbol))
    ctx.bands = ptr(bands)
    ctx.mag_mask = ptr(_i(mag_mask)) if mag_mask is not None else None
    ctx.ncosmo = len(cz)
    ctx.cz = ptr(_d(cz))
    ctx.cdl = ptr(_d(cdl))
    ctx.add_igm = int(add_igm)
    ctx.smooth_fast = int(smooth_fast)
    ctx.tiny_logt = 0.0
    return ctx


def _set_ops(ctx, ops, ptr, keep, imf_type, remnants, salp_ind, mdave, mlim_bh, mlim_ns):
    """The operator set (synth.SSPOperator) and the IMF settings on an oracle context."""
    o = _Ops()
    o.nz, o.nt, o.npts, o.ntpl, o.nent = ops.nz, ops.nt, ops.npts, ops.ntpl, ops.nent
    for k in ("age_ptr", "first", "ent_tpl"):
        setattr(o, k, ptr(_i(getattr(ops, k))))
    for k in ("mini", "m1", "m2", "cm", "cl", "q", "age_mmax", "ent_val", "etpl"):
        setattr(o, k, ptr(_d(getattr(ops, k))))
    keep.append(o)
    ctx.ops = C.cast(C.pointer(o), C.c_void_p)
    ctx.imf_type, ctx.remnants = int(imf_type), int(remnants)
    ctx.salp_ind, ctx.mdave, ctx.mlim_bh, ctx.mlim_ns = salp_ind, mdave, mlim_bh, mlim_ns


class Emission:
    """Dust emission and the AGN torus for the oracle: the Draine & Li templates on the grid (dustem (nqpah, 2 numin,
    nspec) C order = the reference's dustem2_dustem (nspec, nqpah, 2 numin)), q_PAH and U_min; the AGN templates
    (agn (ntau, nspec)) and their optical depths; add_dust_emission, add_agn_dust."""

    def __init__(self, qpah=None, umin=None, dustem=None, agn_tau=None, agn=None, add_dust_emission=1,
                 add_agn_dust=1):
        self.qpah, self.umin, self.agn_tau = qpah, umin, agn_tau
        self.dustem = None if dustem is None else np.ascontiguousarray(dustem, dtype=np.float64)
        self.agn = None if agn is None else np.ascontiguousarray(agn, dtype=np.float64)
        self.add_dust_emission, self.add_agn_dust = add_dust_emission, add_agn_dust


def _find_interval_asc(a, v):
    """find_interval (src/math/opensps_interpolation.f90:149-198) on an ascending table, vectorised: the 1-based i
    with a(i) <= v < a(i+1), 0 below, n above, 1 at a(1), n-1 at a(n)."""
    a = np.asarray(a, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    i = np.searchsorted(a, v, side="right")
    i = np.where(v == a[0], 1, i)
    return np.where(v == a[-1], a.size - 1, i)


def _interp_extrap(x, y, xo):
    """interpolate_linear_scalar (src/math/opensps_interpolation.f90:45-72): the bracketing interval clamped to the
    table, so that the ends extrapolate linearly."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    i = np.clip(_find_interval_asc(x, xo), 1, x.size - 1)
    return y[i - 1] + (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (xo - x[i - 1])


def emission_on_grid(lam, dust=None, agn=None, add_dust_emission=1, add_agn_dust=1):
    """The libraries of datalib.dust_emission (qpah, umin, wavelength, spectra (nqpah, numin, 2, nlam)) and
    datalib.agn_torus (tau, wavelength, spectra (ntau, nlam)) placed on the grid lam as the reference loads them:
    load_dust_emission_table (src/io/opensps_io.f90:1049-1059: each column linear in wavelength from the pixel of
    1 micron on, extrapolated past the library's ends, zero below) and load_agn_dust_models (:1144-1154: log10 of
    flux + tiny linear in log10 wavelength over the pixels from the library's first to its last wavelength, zero
    elsewhere). Returns an Emission."""
    lam = np.asarray(lam, dtype=np.float64)
    tiny = np.finfo(np.float64).tiny
    qp = um = dt = at = tb = None
    if dust is not None:
        qp, um, dl, ds = dust
        i0 = max(int(_find_interval_asc(lam, 1.0e4)), 1)
        nq, nu = len(qp), len(um)
        dt = np.zeros((nq, 2 * nu, lam.size))
        for k in range(nq):
            for j in range(nu):
                for comp in range(2):
                    dt[k, 2 * j + comp, i0 - 1:] = _interp_extrap(dl, ds[k, j, comp], lam[i0 - 1:])
    if agn is not None:
        at, al, asp = agn
        i1 = max(int(_find_interval_asc(lam, al[0])), 1)
        i2 = max(int(_find_interval_asc(lam, al[-1])), 1)
        tb = np.zeros((len(at), lam.size))
        for k in range(len(at)):
            tb[k, i1 - 1:i2] = 10.0 ** _interp_extrap(np.log10(al), np.log10(asp[k] + tiny),
                                                      np.log10(lam[i1 - 1:i2])) - tiny
    return Emission(qp, um, dt, at, tb, add_dust_emission=add_dust_emission, add_agn_dust=add_agn_dust)


NEBULAR_FLOOR = 1.0e-95   # load_nebular_grid's floor before the logarithm (src/io/opensps_io.f90:763)


class Nebular:
    """The CLOUDY tables for the oracle and the device, as load_nebular_grid leaves them (src/io/opensps_io.f90:749-823):
    logz (nnz), lage (nna, log10 yr), logu (nnu) ascending; cont (nnu, nna, nnz, nspec) C order = the reference's
    nebem_cont (nspec, nnz, nna, nnu): log10(L_nu + 1e-95) per unit ionizing photon rate, linear in wavelength onto
    the grid; lines (nnu, nna, nnz, nline) C order = nebem_line (nline, nnz, nna, nnu): log10(L + 1e-95); lpos
    (nline) the lines' wavelengths; add_neb_continuum and nebemlineinspec."""

    def __init__(self, logz, lage, logu, cont, lines, lpos, cont_on=1, lines_in_spec=1):
        self.logz, self.lage, self.logu = (np.ascontiguousarray(a, dtype=np.float64) for a in (logz, lage, logu))
        self.cont = np.ascontiguousarray(cont, dtype=np.float64)
        self.lines = np.ascontiguousarray(lines, dtype=np.float64)
        self.lpos = np.ascontiguousarray(lpos, dtype=np.float64)
        self.cont_on, self.lines_in_spec = int(cont_on), int(lines_in_spec)


def nebular_on_grid(lam, logz, age, logu, lib_lam, cont, line_lam, lines, cont_on=1, lines_in_spec=1):
    """The library of datalib.nebular (axes logZ, age (yr), logU; wavelength (nlam); continuum (nnz, nna, nnu, nlam)
    L_sun/Hz per unit Q; line wavelengths (nline); lines (nnz, nna, nnu, nline) L_sun per unit Q) on the grid lam, as
    load_nebular_grid places it: log10(x + 1e-95), the continuum linear in wavelength (the ends extrapolated, as
    interpolate_linear_scalar does), the age axis to log10."""
    cont = np.asarray(cont, dtype=np.float64)
    lines = np.asarray(lines, dtype=np.float64)
    nnz, nna, nnu = cont.shape[:3]
    lc = np.log10(cont + NEBULAR_FLOOR)
    out = np.empty((nnu, nna, nnz, len(lam)))
    for i in range(nnz):
        for j in range(nna):
            for k in range(nnu):
                out[k, j, i] = _interp_extrap(lib_lam, lc[i, j, k], lam)
    ll = np.log10(lines + NEBULAR_FLOOR).transpose(2, 1, 0, 3)   # (nnu, nna, nnz, nline)
    return Nebular(logz, np.log10(np.asarray(age, dtype=np.float64)), logu, out, ll, line_lam, cont_on, lines_in_spec)


def _set_nebular(ctx, nb, ptr):
    ctx.add_neb, ctx.neb_cont_on, ctx.neb_lines_in_spec = 1, nb.cont_on, nb.lines_in_spec
    ctx.nnz, ctx.nna, ctx.nnu, ctx.nline = len(nb.logz), len(nb.lage), len(nb.logu), len(nb.lpos)
    ctx.neb_logz, ctx.neb_lage, ctx.neb_logu = ptr(nb.logz), ptr(nb.lage), ptr(nb.logu)
    ctx.neb_cont, ctx.neb_line, ctx.neb_lpos = ptr(nb.cont), ptr(nb.lines), ptr(nb.lpos)


def _set_emission(ctx, e, ptr):
    ctx.add_dust_emission, ctx.add_agn_dust = int(e.add_dust_emission), int(e.add_agn_dust)
    if e.dustem is not None:
        ctx.nqpah, ctx.numin = len(e.qpah), len(e.umin)
        assert e.dustem.shape[:2] == (ctx.nqpah, 2 * ctx.numin)
        # (nqpah, 2 numin, nspec) C order -> (2 numin, nqpah, nspec) C order = (nspec, nqpah, 2 numin) column-major
        ctx.dustem = ptr(np.ascontiguousarray(np.transpose(e.dustem, (1, 0, 2))))
        ctx.qpah, ctx.umin = ptr(_d(e.qpah)), ptr(_d(e.umin))
    if e.agn is not None:
        ctx.ntau = len(e.agn_tau)
        ctx.agn, ctx.agn_tau = ptr(e.agn), ptr(_d(e.agn_tau))


NBGRAD = 44
# params_bar rows (0-based) this build computes: the reference's BGRAD_* - 1 (opensps_csp_batch.f90:166-213)
GRAD_ROWS = {"dust1": 1, "dust2": 2, "dust3": 3, "dust_index": 4, "dust1_index": 5, "frac_obrun": 6,
             "frac_nodust": 7, "logzsol": 15, "zred": 22, "sigma_smooth": 23, "igm_factor": 32,
             "tage": 16, "tau": 17, "const": 18, "fburst": 19, "tburst": 20, "sf_start": 21, "sf_trunc": 24,
             "sf_slope": 25, "imf1": 26, "imf2": 27, "imf3": 28, "mdave": 30}


def csp_batch_grad(grid, bands, prm, ssp_weights=None, ssp_weights_z=None, mags_bar=None, spec_bar=None,
                   mass_bar=None, lbol_bar=None, add_igm=1, smooth_fast=1, mag_mask=None, cosmo=None, ops=None,
                   imf_type=2, remnants=1, salp_ind=2.35, mdave=0.5, mlim_bh=40.0, mlim_ns=8.5, zred_mags=None,
                   nebular=None):
    """The oracle's reverse pass (FP64). prm (NPAR, n); ssp_weights (n, nt); ssp_weights_z (n, nz, nt);
    cotangents mags_bar (n, nbands), spec_bar (n, nspec), mass_bar, lbol_bar (n); ops: a synth.SSPOperator (the
    IMF per model and its rows imf1, imf2 and imf3 or mdave). Returns mags (n, nbands), params_bar (n, NBGRAD),
    weights_bar (n, nt), weights_z_bar (n, nz, nt) and the status."""
    keep = []

    def ptr(a):
        keep.append(a)
        return a.ctypes.data

    n = prm.shape[1]
    bands = _d(bands)
    ctx = _context(grid, bands, add_igm, smooth_fast, mag_mask, cosmo, ptr)
    if ssp_weights_z is not None:
        ctx.ssp_weights_z = ptr(_d(ssp_weights_z))
    if ops is not None:
        _set_ops(ctx, ops, ptr, keep, imf_type, remnants, salp_ind, mdave, mlim_bh, mlim_ns)
    if zred_mags is not None:
        ctx.zred_mags = ptr(_d(zred_mags))
    if nebular is not None:
        _set_nebular(ctx, nebular, ptr)
    out = {"mags": np.zeros((n, bands.shape[0])), "params_bar": np.zeros((n, NBGRAD)),
           "weights_bar": np.zeros((n, grid.nt)), "weights_z_bar": np.zeros((n, grid.nz, grid.nt))}
    vp = C.c_void_p

    def opt(a):
        return vp(ptr(_d(a)) if a is not None else None)

    st = lib().orc_csp_batch_grad(C.byref(ctx), C.c_int(n), vp(ptr(_d(prm))), opt(ssp_weights), opt(mags_bar),
                                  opt(spec_bar), opt(mass_bar), opt(lbol_bar), vp(out["mags"].ctypes.data),
                                  vp(out["params_bar"].ctypes.data), vp(out["weights_bar"].ctypes.data),
                                  vp(out["weights_z_bar"].ctypes.data))
    out["status"] = st
    return out


def csp_batch(grid, bands, prm, ssp_weights=None, add_igm=1, smooth_fast=1, mag_mask=None, cosmo=None,
              want=("mags", "spec", "mass", "lbol"), ops=None, imf_type=2, remnants=1, salp_ind=2.35, mdave=0.5,
              mlim_bh=40.0, mlim_ns=8.5, ssp_weights_z=None, dust_curve=None, zred_mags=None, emission=None,
              nebular=None):
    """Runs the oracle's batch. grid: synth.SSPGrid; bands (nbands, nspec); prm (NPAR, n); ops: a
    synth.SSPOperator (the IMF per model; the grid's bases are then unused); ssp_weights_z (n, nz, nt): weights
    over ages and metallicities; emission: an Emission (dust emission and AGN templates on the grid, with the two
    settings); nebular: a Nebular (add_neb_emission = 1; "emlines" in want: the attenuated line luminosities).
    Returns a dict of outputs (C-order arrays: mags (n, nbands), spec (n, nspec), ...) and the status."""
    keep = []

    def ptr(a):
        keep.append(a)
        return a.ctypes.data

    n = prm.shape[1]
    bands = _d(bands)
    nb = bands.shape[0]
    cz, cdl = cosmo if cosmo is not None else cosmology_table()
    ctx = _Ctx()
    if ssp_weights_z is not None:
        ctx.ssp_weights_z = ptr(_d(ssp_weights_z))
    if dust_curve is not None:
        ctx.dcurve = ptr(_d(dust_curve))
    if zred_mags is not None:
        ctx.zred_mags = ptr(_d(zred_mags))
    ctx.nspec, ctx.nt, ctx.nz, ctx.nbands = grid.nspec, grid.nt, grid.nz, nb
    ctx.lambda_ = ptr(_d(grid.lam))
    ctx.time_full = ptr(_d(grid.log_age))
    ctx.logz = ptr(_d(grid.logz))
    ctx.basis = ptr(_d(grid.spec))
    ctx.bmass = ptr(_d(grid.mass))
    ctx.blbol = ptr(_d(grid.lbol))
    ctx.bands = ptr(bands)
    ctx.mag_mask = ptr(_i(mag_mask)) if mag_mask is not None else None
    ctx.ncosmo = len(cz)
    ctx.cz = ptr(_d(cz))
    ctx.cdl = ptr(_d(cdl))
    ctx.add_igm = int(add_igm)
    ctx.smooth_fast = int(smooth_fast)
    ctx.tiny_logt = 0.0
    if emission is not None:
        _set_emission(ctx, emission, ptr)
    if nebular is not None:
        _set_nebular(ctx, nebular, ptr)
    if ops is not None:
        _set_ops(ctx, ops, ptr, keep, imf_type, remnants, salp_ind, mdave, mlim_bh, mlim_ns)
        if "op_count" in want:
            cnt = np.zeros((n, 2), np.int32)
            keep.append(cnt)
            ctx.op_count = cnt.ctypes.data
    out = {"mags": np.zeros((n, nb))}
    for k, shape in (("spec", (n, grid.nspec)), ("mass", (n,)), ("lbol", (n,)), ("young", (n, grid.nspec)),
                     ("old", (n, grid.nspec)), ("w", (n, grid.nt)), ("fband", (n, nb)), ("mdust", (n,))):
        if k in want:
            out[k] = np.zeros(shape)
    if "mdust" in out:
        ctx.mdust_out = out["mdust"].ctypes.data
    if nebular is not None and "emlines" in want:
        out["emlines"] = np.zeros((n, len(nebular.lpos)))
        ctx.emlines_out = out["emlines"].ctypes.data
    idx = np.zeros((n, 3), np.int32) if "idx" in want else None
    if idx is not None:
        out["idx"] = idx
    wts = _d(ssp_weights) if ssp_weights is not None else None
    p = _d(prm)
    f = lib().orc_csp_batch
    vp = C.c_void_p
    st = f(C.byref(ctx), C.c_int(n), vp(p.ctypes.data), vp(wts.ctypes.data if wts is not None else None),
           vp(out["mags"].ctypes.data), *[vp(out[k].ctypes.data) if k in out else vp(None)
                                          for k in ("spec", "mass", "lbol", "young", "old")],
           vp(idx.ctypes.data if idx is not None else None), vp(out["w"].ctypes.data if "w" in out else None),
           vp(out["fband"].ctypes.data if "fband" in out else None))
    out["status"] = st
    if ops is not None and "op_count" in want:
        out["op_count"] = cnt
    return out


def num_threads() -> int:
    return lib().orc_num_threads()
