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


def varied_names() -> list:
    """The parameters the parametric workload varies per model (the names[] of its batch call)."""
    return ["sfh", "tage", "tau", "logzsol", "dust2", "dust_index", "zred", "sigma_smooth"]


assert NPAR == len(PARAM_DEFAULTS)
