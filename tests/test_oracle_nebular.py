"""Pins of the oracle's nebular stage (the CLOUDY grids: nebular_setup, nebular_kernels, line_profiles and
line_attenuated of src/physics/opensps_csp_batch.f90:1630-1770, 2612-3160, 3440-3473 in the reference), each chosen so
that a plausible mistake fails it:

* the ionizing photons: a basis of Planck spectra, whose photon rate below 912 A is a closed-form integral, seen
  through a constant continuum table (a wrong unit, frequency weight or age cut fails);
* the grid interpolation: tables linear in log age, gas log Z and log U are reproduced exactly by the
  interpolation (a wrong index, weight or axis fails), clamped outside the axes;
* the lines: a line's profile integrates over frequency to its luminosity, which equals the line table times the
  photons (a wrong normalisation or lambda^2/c fails), and emlines carries it;
* the escaping fraction: below the Lyman limit the nebular ages' stellar light scales by frac_obrun and the
  nebular light by 1 - frac_obrun;
* the lines' attenuation: the diffuse and birth-cloud transmissions at the line's wavelength;
* the loader: the test release's nebular library (a fixture) and its placement on the grid."""
import os

import numpy as np
import pytest
from scipy import integrate

from helpers import cosmo
from libopensps_amd import datalib, synth
from libopensps_amd.params import SLOT

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TESTDATA = os.path.join(HERE, "golden", "opensps-test-data")
C_A = 2.99792458e18
H_PLANCK, L_SOL = 6.62607015e-27, 3.828e33
K_B = 1.380649e-16


def _planck_grid(nt=24, nspec=6000):
    """One metallicity, Planck spectra (peak 1) at temperatures falling with age; L_sun/Hz per solar mass."""
    lam = np.geomspace(10.0, 2.0e5, nspec)
    log_age = np.linspace(5.5, 10.15, nt)
    temps = np.geomspace(8.0e4, 4.0e3, nt)
    nu = C_A / lam
    spec = np.empty((1, nt, nspec))
    for t in range(nt):
        x = H_PLANCK * nu / (K_B * temps[t])
        b = nu ** 3 / np.expm1(np.minimum(x, 700.0))
        spec[0, t] = b / b.max()
    g = synth.SSPGrid(lam=lam, log_age=log_age, logz=np.array([0.0]), spec=spec, mass=np.ones((1, nt)),
                      lbol=np.zeros((1, nt)))
    return g, temps


def _tables(g, cont_fn, line_fn=None, nline=3, lines_in_spec=1, cont_on=1, lpos=None):
    """A Nebular on grid g whose log10 tables are given as functions of (log Z, log10 age, log U)."""
    lz, age, lu = synth.NEB_LOGZ, synth.NEB_AGE, synth.NEB_LOGU
    la = np.log10(age)
    cont = np.empty((lu.size, la.size, lz.size, g.nspec))
    lines = np.empty((lu.size, la.size, lz.size, nline))
    for k, u in enumerate(lu):
        for j, a in enumerate(la):
            for i, z in enumerate(lz):
                cont[k, j, i] = cont_fn(z, a, u)
                lines[k, j, i] = line_fn(z, a, u) if line_fn is not None else -95.0
    if lpos is None:
        lpos = np.array([3000.0, 6564.6, 20000.0])[:nline]
    return O.Nebular(lz, la, lu, cont, lines, lpos, cont_on, lines_in_spec)


def _run(g, prm, neb, weights=None, want=("spec", "w", "mags")):
    b = np.ones((1, g.nspec)) / g.nspec
    return O.csp_batch(g, b, prm, ssp_weights=weights, add_igm=0, cosmo=cosmo(), want=want, nebular=neb)


def _prm(n, **kv):
    p = synth.default_params(n)
    p[SLOT["sfh"]] = 0.0
    p[SLOT["tage"]] = 1.0
    for k, v in kv.items():
        p[SLOT[k]] = v
    return p


def _onehot(g, t):
    w = np.zeros((1, g.nt))
    w[0, t] = 1.0
    return w


def test_ionizing_photons_closed_form():
    g, temps = _planck_grid()
    c0 = -40.0
    neb = _tables(g, lambda z, a, u: np.full(g.nspec, c0), lines_in_spec=0)
    prm = _prm(1)
    nt_neb = int(np.sum(g.log_age <= np.log10(synth.NEB_AGE[-1])))
    wl = int(np.sum(g.lam <= 912.0))
    numax, numin = C_A / g.lam[0], C_A / g.lam[wl - 1]
    for t in (0, nt_neb - 1, nt_neb):
        w = _onehot(g, t)
        r0 = _run(g, prm, None, w)
        r1 = _run(g, prm, neb, w)
        assert r0["status"] == 0 and r1["status"] == 0
        d = (r1["spec"] - r0["spec"])[0]
        above = g.lam > 1000.0
        if t >= nt_neb:      # no nebular emission past the CLOUDY ages
            assert np.all(d == 0.0)
            continue
        x = lambda nu: (nu ** 2 / np.expm1(H_PLANCK * nu / (K_B * temps[t])))
        peak = (C_A / g.lam) ** 3 / np.expm1(np.minimum(H_PLANCK * (C_A / g.lam) / (K_B * temps[t]), 700.0))
        q, _ = integrate.quad(x, numin, numax, limit=500, epsrel=1e-10)
        q = q / peak.max() / H_PLANCK * L_SOL
        want = 10.0 ** c0 * q
        np.testing.assert_allclose(d[above], want, rtol=2e-4)


def test_interpolation_exact_for_linear_tables():
    g, _ = _planck_grid()
    c0 = -40.0
    lin = lambda z, a, u: -45.0 + 0.8 * (a - 6.5) + 0.5 * z - 0.3 * u
    nt_neb = int(np.sum(g.log_age <= np.log10(synth.NEB_AGE[-1])))
    cases = [(-1.1, -2.7), (0.05, -3.9), (-2.5, -0.5)]          # the last outside both axes: clamped
    for gz, gu in cases:
        prm = _prm(1, gas_logz=gz, gas_logu=gu)
        zc = min(max(gz, synth.NEB_LOGZ[0]), synth.NEB_LOGZ[-1])
        uc = min(max(gu, synth.NEB_LOGU[0]), synth.NEB_LOGU[-1])
        for t in (0, nt_neb // 2, nt_neb - 1):
            w = _onehot(g, t)
            ra = _run(g, prm, _tables(g, lambda z, a, u: np.full(g.nspec, c0), lines_in_spec=0), w)
            rb = _run(g, prm, _tables(g, lambda z, a, u: np.full(g.nspec, lin(z, a, u)), lines_in_spec=0), w)
            r0 = _run(g, prm, None, w)
            above = g.lam > 1000.0
            da = (ra["spec"] - r0["spec"])[0, above]
            db = (rb["spec"] - r0["spec"])[0, above]
            ac = min(max(g.log_age[t], np.log10(synth.NEB_AGE[0])), np.log10(synth.NEB_AGE[-1]))
            want = 10.0 ** (lin(zc, ac, uc) - c0)
            np.testing.assert_allclose(db / da, want, rtol=1e-10)


def test_line_profile_and_luminosity():
    g, _ = _planck_grid()
    c0, l0 = -40.0, -30.0
    lpos = np.array([6564.6])
    neb_c = _tables(g, lambda z, a, u: np.full(g.nspec, c0), lines_in_spec=0, nline=1, lpos=lpos)
    neb_l = _tables(g, lambda z, a, u: np.full(g.nspec, -95.0), lambda z, a, u: np.array([l0]), nline=1,
                    cont_on=0, lpos=lpos)
    for sigma in (0.0, 150.0):
        prm = _prm(1, sigma_smooth=sigma)
        w = _onehot(g, 2)
        r0 = _run(g, prm, None, w)
        ra = _run(g, prm, neb_c, w)
        rl = _run(g, prm, neb_l, w, want=("spec", "w", "mags", "emlines"))
        wq = (ra["spec"] - r0["spec"])[0, g.lam > 1000.0][0] / 10.0 ** c0     # w_t Q_t
        lum = rl["emlines"][0, 0]
        np.testing.assert_allclose(lum, wq * 10.0 ** l0, rtol=1e-12)
        d = (rl["spec"] - r0["spec"])[0]
        nu = C_A / g.lam
        e = -np.trapezoid(d, nu)
        np.testing.assert_allclose(e, lum, rtol=2e-3)
        # the profile's width: the velocity width or two pixels, whichever is wider
        jc = np.argmax(d)
        assert abs(g.lam[jc] - lpos[0]) < 2 * (g.lam[jc + 1] - g.lam[jc])


def test_escaping_fraction():
    g, _ = _planck_grid()
    c0 = -40.0
    neb = _tables(g, lambda z, a, u: np.where(g.lam < 912.0, -95.0, c0), lines_in_spec=0)
    w = _onehot(g, 1)
    r0 = _run(g, _prm(1, frac_obrun=0.3), None, w)
    ra = _run(g, _prm(1, frac_obrun=0.0), neb, w)
    rb = _run(g, _prm(1, frac_obrun=0.3), neb, w)
    euv = g.lam <= 912.0
    # (the table's floor, 1e-95 per unit Q, leaves ~1e-35 there)
    np.testing.assert_allclose(ra["spec"][0, euv], 0.0, atol=1e-30)
    np.testing.assert_allclose(rb["spec"][0, euv], 0.3 * r0["spec"][0, euv], rtol=1e-12, atol=1e-30)
    above = g.lam > 1000.0
    np.testing.assert_allclose((rb["spec"] - r0["spec"])[0, above], 0.7 * (ra["spec"] - r0["spec"])[0, above],
                               rtol=1e-10)


def test_line_attenuation():
    g, _ = _planck_grid()
    l0 = -30.0
    lpos = np.array([4862.7, 6564.6])
    neb = _tables(g, lambda z, a, u: np.full(g.nspec, -95.0), lambda z, a, u: np.full(2, l0), nline=2, cont_on=0,
                  lpos=lpos)
    w = _onehot(g, 2)   # a young age (it_tesc: dust_tesc 7 -> 1e7 yr)
    clean = _run(g, _prm(1), neb, w, want=("spec", "emlines"))["emlines"][0]
    r = _run(g, _prm(1, dust2=0.5, dust_index=-0.7, dust1=0.8, dust1_index=-1.0), neb, w,
             want=("spec", "emlines"))["emlines"][0]
    want = clean * np.exp(-0.5 * (lpos / 5500.0) ** -0.7) * np.exp(-0.8 * (lpos / 5500.0) ** -1.0)
    np.testing.assert_allclose(r, want, rtol=1e-5)
    # an old age: the diffuse dust only
    w = _onehot(g, 20)
    assert g.log_age[20] > 7.0
    neb_old = _tables(g, lambda z, a, u: np.full(g.nspec, -95.0), lambda z, a, u: np.full(2, l0), nline=2,
                      cont_on=0, lpos=lpos)
    assert _run(g, _prm(1), neb_old, w, want=("emlines",))["emlines"].max() == 0.0   # past the CLOUDY ages


def test_statuses():
    g, _ = _planck_grid(nt=12, nspec=800)
    neb = _tables(g, lambda z, a, u: np.full(g.nspec, -40.0), lambda z, a, u: np.full(3, -30.0))
    prm = _prm(2)
    prm[SLOT["sigma_smooth"]] = [100.0, 120.0]   # lines in the spectrum: one profile width a batch
    assert _run(g, prm, neb)["status"] == 2
    neb.lines_in_spec = 0
    assert _run(g, prm, neb)["status"] == 0
    r = O.csp_batch_grad(g, np.ones((1, g.nspec)) / g.nspec, _prm(2), ssp_weights=np.ones((2, g.nt)) / g.nt,
                         mags_bar=np.ones((2, 1)), cosmo=cosmo(), nebular=neb)
    assert r["status"] == 2


def test_test_release_library():
    logz, age, logu, lam, cont, llam, lines = datalib.nebular(TESTDATA, "mist_nodust")
    assert cont.shape == (11, 10, 7, lam.size) and lines.shape == (11, 10, 7, 166) and llam.size == 166
    assert np.all(np.diff(logz) > 0) and np.all(np.diff(age) > 0) and np.all(np.diff(logu) > 0)
    g = synth.ssp_grid(seed=1, nspec=400, nt=20, nz=1)
    neb = O.nebular_on_grid(g.lam, logz, age, logu, lam, cont, llam, lines)
    np.testing.assert_allclose(neb.lage, np.log10(age))
    # the tables' floor and layout: nebem_line(:, iz, ia, iu) = log10(lines(iz, ia, iu, :) + 1e-95)
    np.testing.assert_array_equal(neb.lines[3, 5, 8], np.log10(lines[8, 5, 3] + O.NEBULAR_FLOOR))
    # a grid pixel on a library wavelength takes that wavelength's value
    j = np.searchsorted(g.lam, lam[100])
    lc = np.log10(cont[2, 4, 6] + O.NEBULAR_FLOOR)
    want = np.interp(g.lam[j], lam, lc)
    np.testing.assert_allclose(neb.cont[6, 4, 2, j], want, rtol=1e-12)
