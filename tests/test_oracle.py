"""Pins of the FP64 oracle (oracle/csp_oracle.c) against what does not depend on it: brute force on tiny
grids, closed forms, independent library routines (scipy, numpy) and the invariants the reference's own
tests use. Each test names the oracle function it pins and a mistake it would catch."""
import numpy as np
import pytest
from scipy import integrate, special

from helpers import bands, cosmo, grid
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O


# ---- find_interval (opensps_interpolation.f90:149-198) ---------------------------------------------------------
def _brute_interval(a, v):
    n = len(a)
    if v == a[0]:
        return 1
    if v == a[-1]:
        return n - 1
    return int(np.sum(a <= v)) if a[-1] >= a[0] else int(np.sum(a > v))


@pytest.mark.parametrize("desc", [False, True])
def test_find_interval_brute_force(desc):
    rng = np.random.default_rng(0)
    for trial in range(200):
        n = int(rng.integers(2, 12))
        a = np.sort(rng.uniform(-5, 5, n))
        if desc:
            a = a[::-1].copy()
        probes = list(a) + list(rng.uniform(-7, 7, 20)) + [a.min() - 1, a.max() + 1]
        for v in probes:
            assert O.find_interval(a, v) == _brute_interval(a, v), (a, v)


# ---- expi (opensps_special_functions.f90:50-122): series below 40, asymptotic above ----------------------------
def test_expi_against_scipy():
    xs = np.concatenate([np.geomspace(1e-10, 39.99, 300), [40.0, 40.0001], np.geomspace(40.01, 700, 200)])
    for x in xs:
        ref = special.expi(x)
        assert abs(O.expi(x) - ref) <= 2e-13 * abs(ref) + 1e-300, x
    assert O.expi(0.0) == -np.inf and np.isnan(O.expi(-1.0))


# ---- SFH weights (opensps_sfh.f90:97-257, 1077-1294) against quadrature of SFR x hat kernel -------------------------
def _hat_weights_quad(tf, sfr, t_lo, t_hi):
    """Weight of each node: integral over lookback time t in [t_lo, t_hi] of SFR(t) times the hat function in log10 t
    centred on the node (nodes tf, log10 yr); below tf[0] everything goes to the first node. Normalised to sum 1."""
    nt = len(tf)
    w = np.zeros(nt)
    lt_lo, lt_hi = np.log10(max(t_lo, 1.0)), np.log10(t_hi)
    # below the first node
    a, b = max(t_lo, 1.0), min(t_hi, 10 ** tf[0])
    if b > a:
        w[0] += integrate.quad(sfr, a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
    for j in range(nt - 1):
        x0, x1 = max(tf[j], lt_lo), min(tf[j + 1], lt_hi)
        if x1 <= x0:
            continue
        dx = tf[j + 1] - tf[j]
        f_lo = lambda t: sfr(t) * (tf[j + 1] - np.log10(t)) / dx
        f_hi = lambda t: sfr(t) * (np.log10(t) - tf[j]) / dx
        w[j] += integrate.quad(f_lo, 10 ** x0, 10 ** x1, epsabs=0, epsrel=1e-13, limit=200)[0]
        w[j + 1] += integrate.quad(f_hi, 10 ** x0, 10 ** x1, epsabs=0, epsrel=1e-13, limit=200)[0]
    return w / w.sum()


@pytest.mark.parametrize("sfh", [1, 4])
def test_sfh_weights_against_quadrature(sfh):
    tf = synth.age_grid(107)
    rng = np.random.default_rng(3)
    for _ in range(12):
        tage, tau = rng.uniform(0.1, 13.7), rng.uniform(0.1, 10.0)
        p = synth.default_params(1)[:, 0]
        p[SLOT["sfh"]], p[SLOT["tage"]], p[SLOT["tau"]] = sfh, tage, tau
        w, st = O.sfh_weights(tf, p)
        assert st == 0
        T, tu = tage * 1e9, tau * 1e9
        if sfh == 4:
            sfr = lambda t: (T - t) * np.exp((t - T) / tu)
        else:
            sfr = lambda t: np.exp((t - T) / tu)
        ref = _hat_weights_quad(tf, sfr, 0.0, T)
        assert abs(w.sum() - 1.0) < 1e-12
        # the moment method differences indefinite integrals of size ~tau exp(t / tau): for tau >> tage they cancel,
        # and the reference's FP64 algorithm keeps ~1e-9 of the unit total there (quadrature is exact to 1e-13)
        assert np.max(np.abs(w - ref)) < 1e-8, (tage, tau, np.max(np.abs(w - ref)))


def test_sfh_weights_simha_against_quadrature():
    """sfh 5 (Simha): delayed tau up to the truncation, then linear from the truncation's SFR with slope sf_slope
    (per Gyr) until it reaches zero; the weights are the hat-kernel integrals of that continuous SFR, normalised (the
    reference mixes its two parts by the linear part's mass fraction, get_sfh_properties_at_age)."""
    tf = synth.age_grid(107)
    rng = np.random.default_rng(17)
    checked = 0
    for slope in (-0.8, -0.15, 0.0, 0.3):
        for _ in range(4):
            tage, tau = rng.uniform(2.0, 13.5), rng.uniform(0.3, 8.0)
            start = rng.choice([0.0, rng.uniform(0.05, 0.5)])
            trunc = start + (tage - start) * rng.uniform(0.3, 0.9)
            p = synth.default_params(1)[:, 0]
            p[SLOT["sfh"]], p[SLOT["tage"]], p[SLOT["tau"]] = 5, tage, tau
            p[SLOT["sf_start"]], p[SLOT["sf_trunc"]], p[SLOT["sf_slope"]] = start, trunc, slope
            w, st = O.sfh_weights(tf, p)
            assert st == 0
            T, tr, tu, sm = (tage - start) * 1e9, (trunc - start) * 1e9, tau * 1e9, -slope * 1e-9
            s_tr = (tr / tu) * np.exp(-tr / tu)

            def sfr(t):   # lookback time t (yr); forward time since the start x = T - t
                x = T - t
                return (x / tu) * np.exp(-x / tu) if x <= tr else s_tr * max(1.0 - sm * (x - tr), 0.0)
            ref = _hat_weights_quad(tf, sfr, 0.0, T)
            assert abs(w.sum() - 1.0) < 1e-12
            assert np.max(np.abs(w - ref)) < 1e-8, (tage, tau, start, trunc, slope, np.max(np.abs(w - ref)))
            checked += 1
    assert checked == 16


@pytest.mark.parametrize("case", ["past", "on_a_node", "straddles_now"])
def test_sfh_weights_tabular_against_quadrature(case):
    """sfh 2, 3: the table's SFR, linear in forward time between its ages, weighted by the hat kernel of each node and
    integrated (M_sun: the weights are not normalised). The bin holding the model's age is the reference's own
    approximation (its shape is taken from the clamped end), so it is checked apart: that bin only moves weights
    among the youngest nodes and keeps the bin's mass."""
    tf = synth.age_grid(107)
    rng = np.random.default_rng(23)
    age = np.sort(np.concatenate([[0.0], rng.uniform(0.05, 12.0, 30)]))
    sfr = rng.uniform(0.5, 20.0, age.size)
    O.set_sfh_tab(age, sfr)
    try:
        tage = {"past": 12.5, "on_a_node": age[-6], "straddles_now": 0.5 * (age[-8] + age[-7])}[case]
        p = synth.default_params(1)[:, 0]
        p[SLOT["sfh"]], p[SLOT["tage"]] = 3, tage
        w, st = O.sfh_weights(tf, p)
        assert st == 0
        p[SLOT["sfh"]] = 2
        w2, _ = O.sfh_weights(tf, p)
        np.testing.assert_array_equal(w, w2)
        T = tage * 1e9
        A, S = age * 1e9, sfr

        def sfr_lb(t):   # lookback t (yr) -> forward time T - t
            return float(np.interp(T - t, A, S)) if 0.0 <= T - t <= A[-1] else 0.0
        past = A[A <= T]
        ref = _hat_weights_quad_mass(tf, sfr_lb, T - past[-1], T, breaks=T - A)
        if case == "straddles_now":
            # the current bin: its mass is the reference's formula, on nodes below its older edge's cell
            j = np.searchsorted(A, T) - 1
            m_bin = S[j + 1] * (1.0 + (-(S[j + 1] - S[j]) / (A[j + 1] - A[j]) / S[j + 1]) / 2.0 *
                                ((T - A[j]) + 1.0 - 2.0 * (T - A[j + 1]))) * ((T - A[j]) - 1.0)
            assert abs(w.sum() - (ref.sum() + m_bin)) < 1e-8 * w.sum()
            ref_old = _hat_weights_quad_mass(tf, sfr_lb, T - A[j], T, breaks=T - A)
            k = np.searchsorted(tf, np.log10(T - A[j])) - 1   # nodes above the current bin's cell are the old bins'
            np.testing.assert_allclose(w[k + 2:], ref_old[k + 2:], rtol=1e-8, atol=1e-8 * w.max())
        else:
            np.testing.assert_allclose(w, ref, rtol=1e-8, atol=1e-9 * w.max())
    finally:
        O.set_sfh_tab([], [])


def _hat_weights_quad_mass(tf, sfr, t_lo, t_hi, breaks=()):
    """_hat_weights_quad without the normalisation, with the SFR's kinks given to the integrator."""
    nt = len(tf)
    w = np.zeros(nt)
    lt_lo, lt_hi = np.log10(max(t_lo, 1.0)), np.log10(t_hi)
    bk = np.asarray(breaks, float)
    a, b = max(t_lo, 1.0), min(t_hi, 10 ** tf[0])
    if b > a:
        pts = bk[(bk > a) & (bk < b)]
        w[0] += integrate.quad(sfr, a, b, epsabs=0, epsrel=1e-12, limit=400, points=pts if pts.size else None)[0]
    for j in range(nt - 1):
        x0, x1 = max(tf[j], lt_lo), min(tf[j + 1], lt_hi)
        if x1 <= x0:
            continue
        dx = tf[j + 1] - tf[j]
        a, b = 10 ** x0, 10 ** x1
        pts = bk[(bk > a) & (bk < b)]
        pts = pts if pts.size else None
        f_lo = lambda t: sfr(t) * (tf[j + 1] - np.log10(t)) / dx
        f_hi = lambda t: sfr(t) * (np.log10(t) - tf[j]) / dx
        w[j] += integrate.quad(f_lo, a, b, epsabs=0, epsrel=1e-12, limit=400, points=pts)[0]
        w[j + 1] += integrate.quad(f_hi, a, b, epsabs=0, epsrel=1e-12, limit=400, points=pts)[0]
    return w


def test_sfh_weights_constant_and_burst():
    """sfh 1 with const and fburst: (1 - C - B) tau part + C constant part + B burst at tburst (two nodes)."""
    tf = synth.age_grid(107)
    p = synth.default_params(1)[:, 0]
    p[SLOT["sfh"]], p[SLOT["tage"]], p[SLOT["tau"]] = 1, 10.0, 2.0
    p[SLOT["const"]], p[SLOT["fburst"]], p[SLOT["tburst"]] = 0.2, 0.3, 7.0
    w, st = O.sfh_weights(tf, p)
    assert st == 0 and abs(w.sum() - 1.0) < 1e-12
    T = 10.0e9
    tau_part = _hat_weights_quad(tf, lambda t: np.exp((t - T) / 2e9), 0.0, T)
    const_part = _hat_weights_quad(tf, lambda t: 1.0, 0.0, T)
    burst = np.zeros_like(tf)
    lb = np.log10(T - 7.0e9)
    j = np.searchsorted(tf, lb) - 1
    burst[j] = (tf[j + 1] - lb) / (tf[j + 1] - tf[j])
    burst[j + 1] = (lb - tf[j]) / (tf[j + 1] - tf[j])
    ref = 0.5 * tau_part + 0.2 * const_part + 0.3 * burst
    assert np.max(np.abs(w - ref)) < 1e-9


def test_sfh_ssp_burst():
    """sfh 0: all mass at lookback tage, split linearly in log age between the two bracketing nodes."""
    tf = synth.age_grid(107)
    p = synth.default_params(1)[:, 0]
    p[SLOT["sfh"]], p[SLOT["tage"]] = 0, 3.3
    w, _ = O.sfh_weights(tf, p)
    lb = np.log10(3.3e9)
    j = np.searchsorted(tf, lb) - 1
    assert np.count_nonzero(w) == 2
    assert abs(w[j] - (tf[j + 1] - lb) / (tf[j + 1] - tf[j])) < 1e-14
    assert abs(w.sum() - 1.0) < 1e-14


# ---- the batch against linear algebra and closed forms ----------------------------------------------------------
def _small(nz=1, nspec=600, nt=40):
    g = synth.ssp_grid(seed=9, nspec=nspec, nt=nt, nz=nz, lmin=900.0, lmax=2.0e5)
    fs = synth.gaussian_filters(8, seed=2)
    return g, O.build_bands(fs, g.lam)


def test_no_dust_spectrum_is_weighted_basis_sum():
    """compute_csp_batch with a weight matrix, no dust, no IGM, z = 0, no smoothing: spec = basis . w (the
    reference's identity test, tests/test_csp_batch.f90:1965), mass = sum w M, L_bol = log10 sum w 10^lbol."""
    g, b = _small(1)
    n = 5
    prm = synth.default_params(n)
    w = synth.dirichlet_weights(n, g.nt, seed=4)
    r = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=0, cosmo=cosmo())
    assert r["status"] == 0
    ref = w @ g.spec[0]
    np.testing.assert_allclose(r["spec"], ref, rtol=1e-13)
    np.testing.assert_allclose(r["mass"], w @ g.mass[0], rtol=1e-13)
    np.testing.assert_allclose(r["lbol"], np.log10(w @ 10 ** g.lbol[0]), rtol=1e-13)


def test_metallicity_blend_at_nodes_and_between():
    """logzsol on a grid node gives that slab; between two nodes the linear blend of the two (not clamped)."""
    g, b = _small(4)
    n = 4
    prm = synth.default_params(n)
    prm[SLOT["logzsol"]] = [g.logz[1], g.logz[2], 0.5 * (g.logz[1] + g.logz[2]), g.logz[-1] + 0.3]
    w = synth.dirichlet_weights(n, g.nt, seed=5)
    r = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=0, cosmo=cosmo(), want=("mags", "spec", "idx"))
    np.testing.assert_allclose(r["spec"][0], w[0] @ g.spec[1], rtol=1e-13)
    np.testing.assert_allclose(r["spec"][1], w[1] @ g.spec[2], rtol=1e-13)
    np.testing.assert_allclose(r["spec"][2], w[2] @ (0.5 * g.spec[1] + 0.5 * g.spec[2]), rtol=1e-12)
    d = (g.logz[-1] + 0.3 - g.logz[-2]) / (g.logz[-1] - g.logz[-2])
    np.testing.assert_allclose(r["spec"][3], w[3] @ ((1 - d) * g.spec[-2] + d * g.spec[-1]), rtol=1e-11)
    assert list(r["idx"][:, 1]) == [2, 3, 2, g.nz - 1]


def test_linearity_in_weights():
    g, b = _small(1)
    prm = synth.default_params(2)
    w = synth.dirichlet_weights(1, g.nt, seed=6)
    r = O.csp_batch(g, b, prm, ssp_weights=np.vstack([w, 2.0 * w]), add_igm=0, cosmo=cosmo())
    assert np.array_equal(r["spec"][1], 2.0 * r["spec"][0])
    np.testing.assert_allclose(r["mags"][1] - r["mags"][0], -2.5 * np.log10(2.0), atol=1e-12)


def test_power_law_dust_closed_form():
    """dust_type 0: the diffuse transmission exp(-dust2 (lambda / 5500)^dust_index); birth cloud on the young ages
    exp(-dust1 (lambda / 5500)^dust1_index) with frac_obrun escaping; frac_nodust unattenuated."""
    g, b = _small(1)
    n = 3
    prm = synth.default_params(n)
    prm[SLOT["dust2"]] = [0.7, 0.7, 0.7]
    prm[SLOT["dust_index"]] = [-0.9, -0.9, -0.9]
    prm[SLOT["dust1"]] = [0.0, 1.1, 1.1]
    prm[SLOT["frac_obrun"]] = [0.0, 0.0, 0.25]
    prm[SLOT["frac_nodust"]] = [0.0, 0.0, 0.1]
    w = synth.dirichlet_weights(n, g.nt, seed=7)
    r = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=0, cosmo=cosmo(), want=("mags", "spec", "young", "old"))
    x = g.lam / 5500.0
    td = np.exp(-0.7 * x ** -0.9)
    tb = np.exp(-1.1 * x ** -1.0)
    it = np.sum(g.log_age <= 7.0)     # dust_tesc = 7: ages <= 10^7 yr are young
    young = w[:, :it] @ g.spec[0][:it]
    old = w[:, it:] @ g.spec[0][it:]
    np.testing.assert_allclose(r["young"], young, rtol=1e-13)
    np.testing.assert_allclose(r["old"], old, rtol=1e-13)
    np.testing.assert_allclose(r["spec"][0], (young[0] + old[0]) * td, rtol=1e-12)
    np.testing.assert_allclose(r["spec"][1], (young[1] * tb + old[1]) * td, rtol=1e-12)
    ref2 = (young[2] * (tb * 0.75 + 0.25) + old[2]) * (td * 0.9 + 0.1)
    np.testing.assert_allclose(r["spec"][2], ref2, rtol=1e-12)


# ---- IGM (cosmology.f90:333-507) against Madau (1995) written from the paper -----------------------------------
def _madau_direct(lam, z):
    """Madau (1995) eqs. 12-16: Lyman-series blanketing (17 lines, Ly-alpha metal term) and the Lyman continuum."""
    lw = np.array([1215.67, 1025.72, 972.537, 949.743, 937.803, 930.748, 926.226, 923.150, 920.963, 919.352,
                   918.129, 917.181, 916.429, 915.824, 915.329, 914.919, 914.576])
    lc = np.array([0.0036, 0.0017, 0.0011846, 0.0009410, 0.0007960, 0.0006967, 0.0006236, 0.0005665, 0.0005200,
                   0.0004817, 0.0004487, 0.0004200, 0.0003947, 0.000372, 0.000352, 0.0003334, 0.00031644])
    obs = lam * (1 + z)
    tau = 0.0
    for w0, c in zip(lw, lc):
        if lam < w0:
            tau += c * (obs / w0) ** 3.46
    if lam < lw[0]:
        tau += 0.0017 * (obs / lw[0]) ** 1.68
    if lam < 911.75:
        xc = obs / 911.75
        opz = 1 + z
        tau += (0.25 * xc ** 3 * (opz ** 0.46 - xc ** 0.46) + 9.4 * xc ** 1.5 * (opz ** 0.18 - xc ** 0.18)
                - 0.7 * xc ** 3 * (xc ** -1.32 - opz ** -1.32) - 0.023 * (opz ** 1.68 - xc ** 1.68))
    return tau


def test_igm_terms_against_madau():
    for z in (0.3, 1.0, 2.5, 4.0, 6.0):
        b6 = O.igm_z_terms(1 + z)
        for lam in np.geomspace(300.0, 1300.0, 97):
            ref = _madau_direct(lam, z)
            direct = O.igm_tau(lam, 1 + z)
            fact = float(np.dot(O.igm_pixel_terms(lam), b6))
            assert abs(direct - ref) <= 1e-13 * max(1.0, abs(ref))
            assert abs(fact - ref) <= 2e-13 * max(1.0, abs(ref))
            if lam >= 1215.67:
                assert direct == 0.0 and fact == 0.0


def test_igm_in_batch_cap_and_transparency():
    """Per-model IGM: pixels redward of Ly-alpha untouched; below, exp(-tau) with tau capped at its maximum for the
    pixels blueward of the maximum; z = 0 transparent."""
    g = synth.ssp_grid(seed=9, nspec=800, nt=30, nz=1, lmin=91.0, lmax=2.0e4)
    b = O.build_bands(synth.gaussian_filters(4, seed=2), g.lam)
    prm = synth.default_params(3)
    prm[SLOT["zred"]] = [0.0, 3.0, 5.0]
    w = synth.dirichlet_weights(3, g.nt, seed=8)
    r = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=1, cosmo=cosmo())
    r0 = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=0, cosmo=cosmo())
    base = r0["spec"]
    np.testing.assert_array_equal(r["spec"][0], base[0])
    red = g.lam >= 1215.67
    np.testing.assert_array_equal(r["spec"][:, red], base[:, red])
    for m, z in ((1, 3.0), (2, 5.0)):
        tau = np.array([_madau_direct(l, z) if l < 1215.67 else 0.0 for l in g.lam])
        jt = np.sum(g.lam < 1215.67)
        jmax = int(np.argmax(tau[:jt]))
        tau[:jmax + 1] = tau[jmax]
        np.testing.assert_allclose(r["spec"][m], base[m] * np.exp(-tau), rtol=1e-10)


# ---- velocity smoothing (opensps_smoothing.f90:553-633) against a direct convolution ------------------------------
def _convolve_direct(lam, spec, sigma, minl, maxl):
    """convolve_variable_resolution in velocity: per pixel, Gaussian weights in ln lambda within +-4 sigma divided by
    lambda, normalised, both integrals by the trapezoid rule over the window."""
    c = 2.99792458e5
    sb = sigma / c
    out = spec.copy()
    for i, li in enumerate(lam):
        if li < minl or li > maxl:
            continue
        lo, hi = li * np.exp(-4 * sb), li * np.exp(4 * sb)
        il = max(0, np.searchsorted(lam, lo, side="right") - 1)
        ih = min(len(lam) - 1, np.searchsorted(lam, hi, side="left"))
        if ih - il + 1 < 2:
            continue
        x = lam[il:ih + 1]
        wgt = np.exp(-0.5 * ((np.log(li) - np.log(x)) / sb) ** 2) / x
        out[i] = np.trapezoid(wgt * spec[il:ih + 1], x) / np.trapezoid(wgt, x)
    return out


def test_velocity_kernel_equals_direct_convolution():
    lam = np.exp(np.linspace(np.log(800.0), np.log(2.0e4), 900))
    rng = np.random.default_rng(5)
    spec = 1.0 + rng.random(lam.size)
    for sigma in (50.0, 150.0, 300.0, 1200.0):
        lo, hi, off, coef = O.velocity_kernel(lam, sigma, 1.0e3, 1.0e4)
        sm = spec.copy()
        for i in range(lam.size):
            if lo[i]:
                k = np.arange(lo[i], hi[i] + 1)
                sm[i] = np.dot(coef[off[i] - 1:off[i] - 1 + k.size], spec[k - 1])
        ref = _convolve_direct(lam, spec, sigma, 1.0e3, 1.0e4)
        np.testing.assert_allclose(sm, ref, rtol=1e-12)
        # a flat spectrum stays flat (tests/spectra/test_opensps_smoothing.f90:384)
        for i in range(lam.size):
            if lo[i]:
                assert abs(coef[off[i] - 1:off[i] - 1 + hi[i] - lo[i] + 1].sum() - 1.0) < 1e-14
    lo, hi, off, coef = O.velocity_kernel(lam, 0.0)
    assert not lo.any() and coef.size == 0


# ---- photometry: observed frame, band integrals, AB zero point, distance ------------------------------------------
def test_observed_frame_against_numpy_interp():
    lam = np.exp(np.linspace(np.log(900.0), np.log(3.0e4), 400))
    spec = 1.0 + np.sin(lam / 700.0) ** 2
    for z in (0.0, 0.37, 1.9):
        f = O.observed_frame(lam, z, spec)
        ref = np.interp(lam / (1 + z), lam, spec) / lam
        inside = lam / (1 + z) >= lam[0]
        np.testing.assert_allclose(f[inside], ref[inside], rtol=1e-13)


def test_band_fluxes_against_trapezoid():
    lam = np.exp(np.linspace(np.log(900.0), np.log(3.0e4), 400))
    fs = synth.gaussian_filters(5, seed=3, lo=2000.0, hi=1.0e4)
    bnd = O.build_bands(fs, lam)
    jlo, jhi = O.band_supports(bnd)
    wt = np.append(0.5 * np.abs(np.diff(lam)), 0.0)
    f = 1.0 + np.cos(lam / 900.0) ** 2
    fb = O.band_fluxes_col(np.ascontiguousarray(bnd.T), jlo, jhi, np.ones(5, np.int32), wt, f)
    for b in range(5):
        np.testing.assert_allclose(fb[b], np.trapezoid(f * bnd[b], lam), rtol=1e-13)
        # normalisation of filters_build: integral T / lambda d lambda = 1 (opensps_filters.f90:474-476)
        assert abs(np.trapezoid(bnd[b] / lam, lam) - 1.0) < 1e-12


def test_ab_zero_point_flat_3631_jy():
    """A flat f_nu = 3631 Jy source at 10 pc has M_AB = 0 (tests/test_physics_oracles.f90:288-320): with FSPS's
    48.60 zero point, 3631 Jy is at -5e-5 mag."""
    g = synth.ssp_grid(seed=9, nspec=500, nt=20, nz=1, lmin=1000.0, lmax=3.0e4)
    fs = synth.gaussian_filters(6, seed=4, lo=2000.0, hi=2.0e4)
    b = O.build_bands(fs, g.lam)
    pc = (648000.0 / np.pi) * 1.495978707e13
    lnu = 3631.0e-23 * 4 * np.pi * (10 * pc) ** 2 / 3.828e33    # L_sun / Hz
    g2 = synth.SSPGrid(lam=g.lam, log_age=g.log_age, logz=g.logz, spec=np.full_like(g.spec, lnu), mass=g.mass,
                       lbol=g.lbol)
    prm = synth.default_params(1)
    w = np.zeros((1, g.nt))
    w[0, 3] = 1.0
    r = O.csp_batch(g2, b, prm, ssp_weights=w, add_igm=0, cosmo=cosmo())
    np.testing.assert_allclose(r["mags"][0], -2.5 * np.log10(3631.0 / 3630.780547701), atol=2e-6)


def test_distance_term_against_quadrature():
    h0, om0, ol0 = 72.0, 0.27, 0.73
    cz, cdl = O.cosmology_table(h0, om0, ol0)
    assert np.all(np.diff(cz) < 0) and abs(cz[0] - 1000.0) < 1e-9 and abs(cz[-1]) < 1e-12
    for z in (0.01, 0.5, 1.0, 3.0, 6.0):
        dc = integrate.quad(lambda x: 1.0 / np.sqrt(om0 * (1 + x) ** 3 + ol0), 0, z, epsrel=1e-13)[0]
        dl = (1 + z) * 2.99792458e5 / h0 * 1e6 * dc
        assert abs(O.lum_distance(h0, om0, ol0, z) - dl) < 1e-8 * dl
        ref = 5 * np.log10(dl / 10.0) - 2.5 * np.log10(1 + z)
        assert abs(O.distance_term(cz, cdl, z) - ref) < 2e-3   # the 500-row table, linear in z
    assert O.distance_term(cz, cdl, 0.0) == 0.0


def test_filter_resampling_modes():
    """Point interpolation (resample_point) and hat projection (resample_hat): both normalised; the hat projection
    conserves integral T d lambda of a curve resolved by the grid."""
    lam = np.exp(np.linspace(np.log(900.0), np.log(3.0e4), 2000))
    fs = synth.gaussian_filters(3, seed=5, lo=2000.0, hi=1.0e4)
    for mode in (0, 1):
        for lc, rc in zip(fs.lam, fs.resp):
            band, st = O.filter_band(lc, rc, lam, mode=mode)
            assert st == 0
            assert abs(np.trapezoid(band / lam, lam) - 1.0) < 1e-12
    lc, rc = fs.lam[0], fs.resp[0]
    hat, _ = O.filter_band(lc, rc, lam, mode=1)
    pt, _ = O.filter_band(lc, rc, lam, mode=0)
    np.testing.assert_allclose(hat[pt > 0.05 * pt.max()], pt[pt > 0.05 * pt.max()], rtol=2e-3)


def test_fixed_attenuation_curves():
    """The oracle's fixed curves against values worked from the papers: Calzetti et al. (2000) k(0.55 um) / 4.05 and
    its two laws at 0.63 um (they differ by 0.7 %: the published law is not continuous there), zero beyond the linear
    law's root (~3.1 um); Reddy et al. (2015) at 0.55 um, continuous at 0.60 um with the 0.036 offset, zero beyond
    2.85 um (unclipped below it)."""
    c = O.calzetti_curve(np.array([5500.0, 6299.999, 6300.001, 2.0e4, 5.0e4, 1500.0]))
    k55 = 4.05 + 2.659 * (-2.156 + 1.509 / 0.55 - 0.198 / 0.55 ** 2 + 0.011 / 0.55 ** 3)
    assert abs(c[0] - k55 / 4.05) < 1e-14 and abs(c[0] - 0.9995) < 1e-3
    assert abs(c[1] * 4.05 - 3.4764) < 1e-3 and abs(c[2] * 4.05 - 3.5018) < 1e-3
    assert c[3] > 0.0 and c[4] == 0.0 and c[5] > 2.0
    r = O.reddy_curve(np.array([5500.0, 5999.9999, 6000.0, 28499.0, 28500.0]))
    assert abs(r[0] * 2.505 - (-5.726 + 4.004 / 0.55 - 0.525 / 0.55 ** 2 + 0.029 / 0.55 ** 3 + 2.505)) < 1e-13
    assert abs(r[1] - r[2]) < 1e-6 and r[4] == 0.0
    assert -0.02 < r[3] < 0.0   # (the optical law dips just below zero before 2.85 um; unclipped, as the reference)
