"""Seeded synthetic inputs of the batched composite-population workloads.

Every input is drawn from numpy's PCG64 generator with a recorded seed, so the same call gives the
same arrays on any machine with this image. The SSP grids are sums of blackbodies (no isochrones or
stellar libraries are needed): per (age, metallicity) three components with temperatures in
3e3-5e4 K and log-uniform amplitudes in 1e-3-1, times a lognormal pixel noise of sigma 0.05, in
L_sun/Hz per solar mass (peak values ~1e-13, inside the FP32 range). With several metallicities the
temperatures of a component are shifted by +5% per dex in log Z/Z_sun.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .params import NPAR, PARAM_DEFAULTS, SLOT

ZSOL = 0.0142
HC_OVER_K = 1.4387769e8  # h c / k in Angstrom Kelvin


@dataclass
class SSPGrid:
    """A synthetic SSP library: wavelengths, ages, metallicities and the spectra over them."""
    lam: np.ndarray           # (nspec,) Angstrom, ascending
    log_age: np.ndarray       # (nt,) log10 yr, ascending
    logz: np.ndarray          # (nz,) log10 Z / Z_sun, ascending
    spec: np.ndarray          # (nz, nt, nspec) L_sun/Hz per solar mass (C order = Fortran (nspec, nt, nz))
    mass: np.ndarray          # (nz, nt) surviving stellar mass per solar mass formed
    lbol: np.ndarray          # (nz, nt) log10 L_bol [L_sun]
    seed: int = 0

    @property
    def nspec(self) -> int:
        return self.lam.size

    @property
    def nt(self) -> int:
        return self.log_age.size

    @property
    def nz(self) -> int:
        return self.logz.size

    @property
    def zlegend(self) -> np.ndarray:
        return ZSOL * 10.0 ** self.logz


def wavelength_grid(nspec: int = 5994, lmin: float = 91.0, lmax: float = 1.0e8) -> np.ndarray:
    return np.exp(np.linspace(np.log(lmin), np.log(lmax), nspec))


def age_grid(nt: int = 107, lo: float = 5.5, hi: float = 10.15) -> np.ndarray:
    return np.linspace(lo, hi, nt)


def _bb_shape(lam: np.ndarray, temp: np.ndarray) -> np.ndarray:
    """nu^3 / (exp(h nu / k T) - 1) in units where its peak is ~1.42 (x^3 / expm1(x), x = hc / lambda k T)."""
    x = HC_OVER_K / (lam[None, :] * temp[:, None])
    with np.errstate(over="ignore"):
        return x ** 3 / np.expm1(x)


def ssp_grid(seed: int = 1, nspec: int = 5994, nt: int = 107, nz: int = 1, logz_range=(-2.2, 0.5),
             lmin: float = 91.0, lmax: float = 1.0e8) -> SSPGrid:
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = wavelength_grid(nspec, lmin, lmax)
    log_age = age_grid(nt)
    logz = np.array([0.0]) if nz == 1 else np.linspace(logz_range[0], logz_range[1], nz)
    temps = rng.uniform(3.0e3, 5.0e4, size=(nt, 3))
    amps = 10.0 ** rng.uniform(-3.0, 0.0, size=(nt, 3))
    spec = np.empty((nz, nt, nspec))
    for z in range(nz):
        shift = 1.0 + 0.05 * (logz[z] if nz > 1 else 0.0)
        noise = np.exp(0.05 * rng.standard_normal(size=(nt, nspec)))
        for t in range(nt):
            comp = _bb_shape(lam, temps[t] * shift)
            spec[z, t] = 1.0e-13 * (amps[t] @ comp) * noise[t]
    frac = (log_age - log_age[0]) / (log_age[-1] - log_age[0])
    mass = np.empty((nz, nt))
    lbol = np.empty((nz, nt))
    for z in range(nz):
        mass[z] = 1.0 - (0.25 + 0.02 * logz[z]) * frac
        lbol[z] = 1.5 - 3.5 * frac + 0.1 * logz[z] + 0.02 * np.sin(7.0 * frac)
    return SSPGrid(lam=lam, log_age=log_age, logz=logz, spec=spec, mass=mass, lbol=lbol, seed=seed)


@dataclass
class FilterSet:
    """Piecewise-linear transmission curves (one per filter) on their own wavelength samples."""
    lam: list = field(default_factory=list)
    resp: list = field(default_factory=list)
    names: list = field(default_factory=list)
    detector: int = 0


def gaussian_filters(nbands: int, seed: int = 7, lo: float = 1500.0, hi: float = 5.0e4, npts: int = 201,
                     fwhm_frac: float = 0.1, nsig: float = 4.0) -> FilterSet:
    """Gaussian filters with centres log-uniform in [lo, hi] and FWHM = fwhm_frac * centre."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = np.sort(10.0 ** rng.uniform(np.log10(lo), np.log10(hi), size=nbands))
    fs = FilterSet()
    for b, c in enumerate(centres):
        sig = fwhm_frac * c / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        lam = np.linspace(c - nsig * sig, c + nsig * sig, npts)
        fs.lam.append(lam)
        fs.resp.append(np.exp(-0.5 * ((lam - c) / sig) ** 2))
        fs.names.append(f"gauss_{b:03d}_{c:.0f}")
    return fs


def default_params(n: int) -> np.ndarray:
    """(NPAR, n) array of the PARAMS_CORE defaults."""
    return np.repeat(np.asarray(PARAM_DEFAULTS, dtype=np.float64)[:, None], n, axis=1)


def dirichlet_weights(n: int, nt: int, seed: int = 3, alpha: float = 0.5) -> np.ndarray:
    """(n, nt) nonnegative SSP weights, Dirichlet(alpha) scaled to 1 solar mass per model."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.dirichlet(np.full(nt, alpha), size=n)


def parametric_params(n: int, seed: int = 5, zred=(0.0, 4.0), sigma=(50.0, 300.0), dust=True,
                      logzsol=(-2.0, 0.3)) -> np.ndarray:
    """(NPAR, n) parameters of the parametric workload: delayed-tau SFH with tage U(0.1, 13.7) Gyr and
    tau U(0.1, 10) Gyr, logzsol U(-2, 0.3), power-law dust tau_V U(0, 2) with slope U(-1.2, -0.4),
    redshift U(0, 4) and a Gaussian velocity dispersion U(50, 300) km/s."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = default_params(n)
    p[SLOT["sfh"]] = 4.0
    p[SLOT["tage"]] = rng.uniform(0.1, 13.7, n)
    p[SLOT["tau"]] = rng.uniform(0.1, 10.0, n)
    p[SLOT["logzsol"]] = rng.uniform(logzsol[0], logzsol[1], n)
    if dust:
        p[SLOT["dust2"]] = rng.uniform(0.0, 2.0, n)
        p[SLOT["dust_index"]] = rng.uniform(-1.2, -0.4, n)
    p[SLOT["zred"]] = rng.uniform(zred[0], zred[1], n)
    p[SLOT["sigma_smooth"]] = rng.uniform(sigma[0], sigma[1], n)
    return p


def lognormal_weights_z(n: int, log_age: np.ndarray, logz: np.ndarray, seed: int = 17, sigma_dex: float = 1.0,
                        zwidth: float = 0.3) -> np.ndarray:
    """(n, nz, nt) weights over ages and metallicities of the adjoint workload: per model a log-normal SFH over the
    ages (sigma = 1 dex in age, peak log-uniform over the grid) and, per age, an enrichment history over the
    metallicities (a Gaussian in log Z of width zwidth dex, centred on a metallicity that rises from the model's
    initial value at the oldest age to its final value at the youngest), scaled to 1 solar mass formed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    la = np.asarray(log_age, dtype=np.float64)
    lz = np.asarray(logz, dtype=np.float64)
    peak = rng.uniform(la[0], la[-1], n)
    w_t = np.exp(-0.5 * ((la[None, :] - peak[:, None]) / sigma_dex) ** 2)
    w_t /= w_t.sum(axis=1, keepdims=True)
    z0 = rng.uniform(lz[0], lz[0] + 0.5 * (lz[-1] - lz[0]), n)
    z1 = rng.uniform(lz[0] + 0.5 * (lz[-1] - lz[0]), lz[-1], n)
    frac = (la[-1] - la) / (la[-1] - la[0])                        # 1 at the youngest age, 0 at the oldest
    zc = z0[:, None] + (z1 - z0)[:, None] * frac[None, :]           # (n, nt)
    e = np.exp(-0.5 * ((lz[None, :, None] - zc[:, None, :]) / zwidth) ** 2)   # (n, nz, nt)
    e /= e.sum(axis=1, keepdims=True)
    return e * w_t[:, None, :]


def adjoint_params(n: int, seed: int = 19, birth_cloud: bool = False) -> np.ndarray:
    """(NPAR, n) physical parameters of the adjoint workload, uniform in the parametric workload's ranges: dust2
    U(0, 2), dust_index U(-1.2, -0.4), redshift U(0, 4), sigma_smooth U(50, 300) km/s; with birth_cloud also
    dust1 U(0, 1), dust1_index U(-1.5, -0.5), dust3 U(0, 0.5), frac_obrun U(0, 0.5) and frac_nodust U(0, 0.5)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = default_params(n)
    p[SLOT["dust2"]] = rng.uniform(0.0, 2.0, n)
    p[SLOT["dust_index"]] = rng.uniform(-1.2, -0.4, n)
    p[SLOT["zred"]] = rng.uniform(0.0, 4.0, n)
    p[SLOT["sigma_smooth"]] = rng.uniform(50.0, 300.0, n)
    if birth_cloud:
        p[SLOT["dust1"]] = rng.uniform(0.0, 1.0, n)
        p[SLOT["dust1_index"]] = rng.uniform(-1.5, -0.5, n)
        p[SLOT["dust3"]] = rng.uniform(0.0, 0.5, n)
        p[SLOT["frac_obrun"]] = rng.uniform(0.0, 0.5, n)
        p[SLOT["frac_nodust"]] = rng.uniform(0.0, 0.5, n)
    return p


def chi2_cotangent(mags: np.ndarray, seed: int = 23, frac_err: float = 0.05, detect_range: float = 40.0):
    """The adjoint workload's loss: chi^2 of the magnitudes against synthetic photometry (mags perturbed by 5 %
    Gaussian flux errors drawn from a held-out seed). Returns (loss, d loss / d mags). NaN bands, and bands more
    than detect_range magnitudes fainter than the model's brightest (non-detections: below the Lyman limit at high
    redshift the IGM leaves ~1e-30 of the flux), are left out."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sig = 2.5 / np.log(10.0) * frac_err
    obs = mags + sig * rng.standard_normal(mags.shape) + 0.1
    fin = np.isfinite(mags)
    bright = np.min(np.where(fin, mags, np.inf), axis=1, keepdims=True)
    ok = fin & (mags - bright <= detect_range)
    r = np.where(ok, (mags - np.where(ok, obs, 0.0)) / sig, 0.0)
    return float(np.sum(r * r)), np.where(ok, 2.0 * r / sig, 0.0)


def varied_names() -> list:
    """The parameters the parametric workload varies per model (the names[] of its batch call)."""
    return ["sfh", "tage", "tau", "logzsol", "dust2", "dust_index", "zred", "sigma_smooth"]


assert NPAR == len(PARAM_DEFAULTS)


# ------------------------------------------------------------------------------------------------
# The SSP operator of the IMF-per-model workload
# ------------------------------------------------------------------------------------------------
@dataclass
class SSPOperator:
    """Isochrone points of every (metallicity, age), their template entries and the template library.

    Points of (iz, t) are [age_ptr[iz * nt + t], age_ptr[iz * nt + t + 1]), sorted by initial mass; their mass bins
    m1 / m2 are the midpoints to the neighbours, consecutive bins sharing their edge (the first point of an age takes
    the IMF's lower bound, flag first = 1, m1 = 0; the last bin ends at its own mass). cm: surviving mass, cl: L_bol (L_sun), q: ionizing photons (1/s)
    per star. Each point has nent entries (template, value): its L_bol shared between the two library templates
    that bracket its Teff in log Teff. The library (ntpl, nspec) holds blackbody spectra per unit L_bol (1/Hz),
    ordered by the first age that uses them, so k_age[t] (templates first used at ages <= t) is increasing.
    """
    nz: int
    nt: int
    age_ptr: np.ndarray     # (nz * nt + 1,) int32
    mini: np.ndarray        # (npts,)
    m1: np.ndarray
    m2: np.ndarray
    first: np.ndarray       # (npts,) int32
    cm: np.ndarray
    cl: np.ndarray
    q: np.ndarray
    age_mmax: np.ndarray    # (nz * nt,) largest living initial mass of each isochrone
    ent_tpl: np.ndarray     # (npts, nent) int32, -1: no entry
    ent_val: np.ndarray     # (npts, nent)
    etpl: np.ndarray        # (ntpl, nspec) per unit L_bol
    teff: np.ndarray        # (ntpl,) the templates' temperatures
    k_age: np.ndarray       # (nt,) int32
    seed: int = 0

    @property
    def npts(self) -> int:
        return self.mini.size

    @property
    def ntpl(self) -> int:
        return self.teff.size

    @property
    def nent(self) -> int:
        return self.ent_tpl.shape[1]


def ssp_operator(lam: np.ndarray, log_age: np.ndarray, nz: int, npts: int = 1500, ntpl: int = 10000,
                 seed: int = 11, tmin: float = 2500.0, tmax: float = 5.0e4) -> SSPOperator:
    """The workload's synthetic operator: per (metallicity, age) npts initial masses log-uniform between 0.08 M_sun
    and the isochrone's largest living mass (a main-sequence lifetime of 1e10 yr m^-2.5, capped at 120 M_sun),
    sorted; log Teff linear in log m from tmin at 0.08 M_sun to tmax at 120 M_sun (0.01 dex scatter, -0.003 dex per
    metallicity step), L ~ m^3.5 (0.05 dex scatter); ntpl blackbody templates log-spaced in [tmin, tmax]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nt = log_age.size
    ages = 10.0 ** log_age
    mmax_t = np.clip((1.0e10 / ages) ** (1.0 / 2.5), 0.9, 120.0)
    lteff_tpl = np.linspace(np.log10(tmin), np.log10(tmax), ntpl)
    teff_tpl = 10.0 ** lteff_tpl
    # blackbody per unit bolometric luminosity: B_nu / integral(B_nu dnu) in 1/Hz
    nu = 2.99792458e18 / lam
    x = HC_OVER_K / (lam[None, :] * teff_tpl[:, None])
    with np.errstate(over="ignore"):
        bnu = nu[None, :] ** 3 / np.expm1(x)
    # integral of nu^3 / (exp(h nu / kT) - 1) dnu = (pi^4 / 15) (k T / h)^4
    kt_h = teff_tpl * 2.99792458e18 / HC_OVER_K
    etpl = bnu / ((np.pi ** 4 / 15.0) * kt_h[:, None] ** 4)
    per = nz * nt
    mini = np.empty(per * npts)
    teff = np.empty(per * npts)
    lum = np.empty(per * npts)
    age_mmax = np.empty(per)
    for iz in range(nz):
        for t in range(nt):
            a = iz * nt + t
            mm = mmax_t[t]
            m = np.sort(10.0 ** rng.uniform(np.log10(0.08), np.log10(mm), npts))
            sl = slice(a * npts, (a + 1) * npts)
            mini[sl] = m
            frac = (np.log10(m) - np.log10(0.08)) / (np.log10(120.0) - np.log10(0.08))
            lt = (np.log10(tmin) + frac * np.log10(tmax / tmin) - 0.003 * iz
                  + 0.01 * rng.standard_normal(npts))
            teff[sl] = 10.0 ** np.clip(lt, np.log10(tmin), np.log10(tmax))
            lum[sl] = m ** 3.5 * 10.0 ** (0.05 * rng.standard_normal(npts))
            age_mmax[a] = m[-1]
    age_ptr = (np.arange(per + 1) * npts).astype(np.int32)
    m1 = np.empty_like(mini)
    m2 = np.empty_like(mini)
    first = np.zeros(mini.size, np.int32)
    for a in range(per):
        m = mini[a * npts:(a + 1) * npts]
        # bins share their edges exactly, as the reference's operator build (g_m1(k) = g_m2(k - 1),
        # src/physics/opensps_ssp_operator.f90:414-423); the first bin's lower edge is 0 (the IMF's lower bound)
        hi = np.empty(npts)
        hi[:-1] = m[:-1] + 0.5 * (m[1:] - m[:-1])
        hi[-1] = m[-1]
        lo = np.empty(npts)
        lo[1:] = hi[:-1]
        lo[0] = 0.0
        m1[a * npts:(a + 1) * npts] = lo
        m2[a * npts:(a + 1) * npts] = hi
        first[a * npts] = 1
    cm = mini * rng.uniform(0.9, 1.0, mini.size)
    q = lum * 1.0e44 * (teff / tmax) ** 4
    # entries: the two bracketing templates in log Teff, linear weights, times L_bol
    lt = np.log10(teff)
    j = np.clip(np.searchsorted(lteff_tpl, lt), 1, ntpl - 1)
    f = (lt - lteff_tpl[j - 1]) / (lteff_tpl[j] - lteff_tpl[j - 1])
    ent_tpl = np.stack([j - 1, j], axis=1).astype(np.int64)
    ent_val = np.stack([(1.0 - f) * lum, f * lum], axis=1)
    # order the library by the first age that uses a template (any metallicity)
    age_of_pt = np.repeat(np.tile(np.arange(nt), nz), npts)
    first_age = np.full(ntpl, nt, np.int64)
    np.minimum.at(first_age, ent_tpl[:, 0], age_of_pt)
    np.minimum.at(first_age, ent_tpl[:, 1], age_of_pt)
    order = np.lexsort((np.arange(ntpl), first_age))
    relabel = np.empty(ntpl, np.int64)
    relabel[order] = np.arange(ntpl)
    ent_tpl = relabel[ent_tpl].astype(np.int32)
    etpl = etpl[order]
    teff_tpl = teff_tpl[order]
    fa = first_age[order]
    k_age = np.array([np.searchsorted(fa, t, side="right") for t in range(nt)], np.int32)
    return SSPOperator(nz=nz, nt=nt, age_ptr=age_ptr, mini=mini, m1=m1, m2=m2, first=first, cm=cm, cl=lum, q=q,
                       age_mmax=age_mmax, ent_tpl=ent_tpl, ent_val=ent_val, etpl=np.ascontiguousarray(etpl),
                       teff=teff_tpl, k_age=k_age, seed=seed)


def imf_params(n: int, seed: int = 13, slopes=(0.3, 3.3), upper=(60.0, 120.0), p=None) -> np.ndarray:
    """(NPAR, n) parameters with a per-model Kroupa IMF: slopes imf1, imf2, imf3 U(slopes), imf_upper_limit
    U(upper); the rest from p (or the defaults)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = default_params(n) if p is None else p.copy()
    out[SLOT["imf1"]] = rng.uniform(slopes[0], slopes[1], n)
    out[SLOT["imf2"]] = rng.uniform(slopes[0], slopes[1], n)
    out[SLOT["imf3"]] = rng.uniform(slopes[0], slopes[1], n)
    out[SLOT["imf_upper_limit"]] = rng.uniform(upper[0], upper[1], n)
    return out


DL07_QPAH = np.array([0.47, 1.12, 1.77, 2.5, 3.19, 3.9, 4.58])
DL07_UMIN = np.array([0.1, 0.15, 0.2, 0.3, 0.4, 0.5, 0.7, 0.8, 1.0, 1.2, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0, 7.0, 8.0,
                      12.0, 15.0, 20.0, 25.0])
AGN_TAU = np.array([5.0, 10.0, 20.0, 30.0, 40.0, 60.0, 80.0, 100.0, 150.0])


def dust_emission_templates(lam: np.ndarray, qpah=DL07_QPAH, umin=DL07_UMIN) -> np.ndarray:
    """Synthetic dust emission templates on the grid, (nqpah, 2 numin, nspec), in the layout of the Draine & Li
    tables as the reference loads them (zero below 1 micron): for each (q_PAH, U_min) a modified blackbody of
    T = 18 K U^(1/6) (beta = 2) for U_max = U_min, then a warmer mixture with PAH bands (Gaussians at 3.3, 6.2, 7.7,
    8.6, 11.3 micron, strength ~ q_PAH) for the power-law U distribution."""
    nu = 2.99792458e18 / lam
    out = np.zeros((len(qpah), 2 * len(umin), lam.size))
    on = lam >= 1.0e4
    um = lam * 1e-4
    pah = sum(a * np.exp(-0.5 * ((um - c) / (0.03 * c)) ** 2)
              for a, c in ((0.3, 3.3), (1.0, 6.2), (1.6, 7.7), (0.6, 8.6), (0.9, 11.3)))
    for k, q in enumerate(qpah):
        for j, u in enumerate(umin):
            for comp in range(2):
                t = 18.0 * u ** (1.0 / 6.0) * (1.0 if comp == 0 else 1.6)
                x = np.minimum(HC_OVER_K / (lam * t), 700.0)
                mbb = nu ** 5 / np.expm1(x)
                mbb = mbb / mbb[on].max()
                v = mbb + (0.02 * q * pah if comp == 1 else 0.005 * q * pah)
                out[k, 2 * j + comp] = np.where(on, 1.0e-24 * v, 0.0)
    return out


def agn_templates(lam: np.ndarray, tau=AGN_TAU) -> np.ndarray:
    """Synthetic AGN torus templates on the grid, (ntau, nspec), zero outside 0.1 - 1000 micron: a broken power law in
    nu L_nu peaking between 8 and 30 micron with the optical depth, per unit bolometric luminosity (1/Hz)."""
    nu = 2.99792458e18 / lam
    um = lam * 1e-4
    out = np.zeros((len(tau), lam.size))
    on = (um >= 0.1) & (um <= 1000.0)
    for k, t in enumerate(tau):
        peak = 8.0 * (t / 5.0) ** 0.4
        x = um / peak
        nulnu = x ** 1.5 / (1.0 + x ** 4.0)
        f = np.where(on, nulnu / nu, 0.0)
        dnu = np.abs(np.gradient(nu))
        out[k] = f / np.sum(f * dnu)
    return out


C_LIGHT_A = 2.99792458e18   # speed of light, A/s
NEB_LOGZ = np.linspace(-2.0, 0.2, 11)
NEB_AGE = np.array([0.5e6, 1.0e6, 2.0e6, 3.0e6, 4.0e6, 5.0e6, 6.0e6, 8.0e6, 1.0e7, 2.0e7])
NEB_LOGU = np.linspace(-4.0, -1.0, 7)


def nebular_library(grid: SSPGrid, nline: int = 166, seed: int = 31, cont_frac: float = 0.3):
    """A synthetic CLOUDY library in the layout datalib.nebular returns (logz, age (yr), logu, wavelength (nlam),
    continuum (nnz, nna, nnu, nlam) L_sun/Hz per unit ionizing photon rate, line wavelengths (nline), lines (nnz,
    nna, nnu, nline) L_sun per unit ionizing photon rate), scaled to the grid: per unit Q of the grid's youngest
    solar-metallicity SSP, its continuum is about cont_frac of that SSP's optical L_nu, smooth in wavelength (a power
    law whose slope follows log U, a Balmer-like step at 3646 A, 1e-6 of it below the Lyman limit), fading with age
    and rising with Z; the lines (log-uniform wavelengths 923 A - 6 micron) have equivalent widths of 1 - 300 A
    against that continuum, varying with Z, U and age."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = np.geomspace(91.0, 1.0e8, 200)
    llam = np.sort(np.exp(rng.uniform(np.log(923.0), np.log(6.0e4), nline)))
    iz = int(np.argmin(np.abs(grid.logz)))
    b = grid.spec[iz, 0]
    nu = C_LIGHT_A / grid.lam
    wl = int(np.searchsorted(grid.lam, 912.0, side="right"))
    q = np.sum(0.5 * np.abs(np.diff(nu[:wl])) * (b[:wl - 1] / nu[:wl - 1] + b[1:wl] / nu[1:wl])) / 6.62607015e-27 * 3.828e33
    opt = (grid.lam > 3000.0) & (grid.lam < 7000.0)
    lref = float(np.median(b[opt]))
    base = cont_frac * lref / max(q, 1e-300)
    nz, na, nu_ = NEB_LOGZ.size, NEB_AGE.size, NEB_LOGU.size
    cont = np.empty((nz, na, nu_, lam.size))
    lines = np.empty((nz, na, nu_, nline))
    ew = 10.0 ** rng.uniform(0.0, 2.5, nline)
    zs, us = rng.uniform(-0.3, 0.3, nline), rng.uniform(-0.4, 0.4, nline)
    for i, lz in enumerate(NEB_LOGZ):
        for j, age in enumerate(NEB_AGE):
            fade = np.exp(-age / 6.0e6)
            for k, lu in enumerate(NEB_LOGU):
                shape = (lam / 5000.0) ** (-0.3 + 0.1 * lu) * np.where(lam < 3646.0, 0.5, 1.0)
                shape = np.where(lam < 912.0, 1.0e-6 * shape, shape)
                cont[i, j, k] = base * shape * fade * (1.0 + 0.15 * lz)
                cl = base * (llam / 5000.0) ** (-0.3 + 0.1 * lu) * fade * (1.0 + 0.15 * lz)
                lines[i, j, k] = ew * (10.0 ** (zs * lz + us * (lu + 2.5))) * cl * C_LIGHT_A / llam ** 2
    return NEB_LOGZ.copy(), NEB_AGE.copy(), NEB_LOGU.copy(), lam, cont, llam, lines
