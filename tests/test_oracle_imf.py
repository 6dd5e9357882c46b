"""Pins of the oracle's IMF integrals and SSP operator (no GPU).

* The reference's own 60-digit table of power-law bins (tests/golden/imf_powerlaw_bins.txt, copied from the
  reference's tests/data): bins down to a relative width of 1e-14, slopes within 1e-6 of 1 and exactly 1, either
  side of the series switch. imf_bin_weights_tab and imf_integral must be within 1e-14 of it in every category, as
  the reference's test_powerlaw_bin_worst_cases requires (tests/physics/test_opensps_imf.f90:528-600, 686-745).
* The Kroupa integrals (breaks 0.08, 0.5, 1 M_sun) against scipy's adaptive quadrature of the broken power law.
* Bin weights: the segment classification and the limits (which points weigh nothing: bit-exact), the first bin's
  lower bound, bins across breaks.
* Remnant masses against quadrature of the remnant-mass law.
* One model's operator age sums against an independent numpy statement of the same sums (numpy's expm1 and
  log1p, dense matrix products).
"""
import os

import numpy as np
import pytest
from scipy import integrate

from helpers import grid, operator
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "imf_powerlaw_bins.txt")


def _table():
    rows = []
    with open(GOLDEN) as f:
        f.readline()
        n = int(f.readline())
        for _ in range(n):
            a, lo, hi, ref = f.readline().split()
            rows.append((float(a), float(lo), float(hi), float(ref)))
    return rows


def _category(i, n, alpha, lo, hi):
    rw = (hi - lo) / lo
    if i >= n - 12:
        return 6
    if abs(1.0 - alpha) < 1.0e-5:
        return 5
    if rw < 1.0e-11:
        return 1
    if rw < 1.0e-5:
        return 2
    if rw < 0.5:
        return 3
    return 4


def test_powerlaw_bins_against_60_digit_table():
    rows = _table()
    assert len(rows) == 444
    worst_w = np.zeros(7)
    worst_i = np.zeros(7)
    count = np.zeros(7, int)
    for i, (alpha, lo, hi, ref) in enumerate(rows):
        cat = _category(i, len(rows), alpha, lo, hi)
        p = O.imf_vector(salp=alpha, lower=0.0, upper=1.0e5)
        w = O.imf_bin_weights_tab(0, p, 1.0, [0.5 * (lo + hi)], [lo], [hi], [0])[0]
        v = O.imf_integral(0, p, lo, hi, 0)
        worst_w[cat] = max(worst_w[cat], abs(w / ref - 1.0))
        worst_i[cat] = max(worst_i[cat], abs(v / ref - 1.0))
        count[cat] += 1
    assert np.all(count[1:] > 0)
    assert worst_w.max() <= 1.0e-14, worst_w
    assert worst_i.max() <= 1.0e-14, worst_i


def _kroupa_pdf(m, a, mass_weighted):
    a1, a2, a3 = a
    c2 = 0.5 ** (-a1 + a2)
    if m < 0.08:
        v = 0.0
    elif m < 0.5:
        v = m ** (-a1)
    elif m < 1.0:
        v = c2 * m ** (-a2)
    else:
        v = c2 * m ** (-a3)
    return v * (m if mass_weighted else 1.0)


@pytest.mark.parametrize("seed", range(4))
def test_kroupa_integral_against_quadrature(seed):
    rng = np.random.default_rng(seed)
    for _ in range(40):
        a = rng.uniform(0.3, 3.3, 3)
        lo, hi = np.sort(10.0 ** rng.uniform(np.log10(0.05), np.log10(150.0), 2))
        p = O.imf_vector(slopes=a)
        for mw in (0, 1):
            pts = [x for x in (0.08, 0.5, 1.0) if lo < x < hi]
            ref = integrate.quad(_kroupa_pdf, lo, hi, args=(a, mw), points=pts or None, epsabs=0, epsrel=1e-13,
                                 limit=200)[0]
            got = O.imf_integral(2, p, lo, hi, mw)
            assert abs(got - ref) <= 1e-11 * abs(ref) + 1e-300, (a, lo, hi, mw, got, ref)


def test_bin_weights_classification_and_limits():
    """Which bins weigh nothing is exact: outside [lower, upper] by initial mass, or empty; the first bin of a run
    starts at the lower bound; bins across the breaks equal the general integral; the rest against quadrature."""
    rng = np.random.default_rng(5)
    m = np.sort(10.0 ** rng.uniform(np.log10(0.05), np.log10(130.0), 400))
    m1 = np.empty_like(m)
    m2 = np.empty_like(m)
    m1[1:] = m[1:] - 0.5 * (m[1:] - m[:-1])
    m1[0] = m[0]
    m2[:-1] = m[:-1] + 0.5 * (m[1:] - m[:-1])
    m2[-1] = m[-1]
    first = np.zeros(m.size, np.int32)
    first[0] = 1
    a = (0.7, 1.9, 2.6)
    p = O.imf_vector(slopes=a, lower=0.1, upper=100.0)
    tot = O.imf_integral(2, p, 0.1, 100.0, 1)
    w = O.imf_bin_weights_tab(2, p, tot, m, m1, m2, first)
    outside = (m < 0.1) | (m > 100.0)
    assert np.all(w[outside] == 0.0)
    assert np.all(w[~outside] > 0.0)
    for k in np.flatnonzero(~outside):
        lo = 0.1 if first[k] else m1[k]
        ref = integrate.quad(_kroupa_pdf, lo, m2[k], args=(a, 0), epsabs=0, epsrel=1e-13,
                             points=[x for x in (0.08, 0.5, 1.0) if lo < x < m2[k]] or None)[0] / tot
        assert abs(w[k] - ref) <= 1e-11 * ref, (k, w[k], ref)


def _remnant_ref(a, lo, up, mmax, mbh=40.0, mns=8.5):
    f = lambda x, mw: _kroupa_pdf(x, a, mw)
    q = lambda x0, x1, mw: integrate.quad(f, x0, x1, args=(mw,), epsabs=0, epsrel=1e-13, limit=200,
                                          points=[x for x in (0.5, 1.0) if x0 < x < x1] or None)[0] if x1 > x0 else 0.0
    norm = q(lo, up, 1)
    add = 0.0
    x = min(max(mbh, mmax), up)
    if x < up:
        add += 0.5 * q(x, up, 1) / norm
    if mmax <= mbh:
        x0, x1 = min(max(mns, mmax), up), min(mbh, up)
        add += 1.4 * q(x0, x1, 0) / norm
    if mmax <= mns:
        x0, x1 = min(mmax, up), min(mns, up)
        add += (0.48 * q(x0, x1, 0) + 0.077 * q(x0, x1, 1)) / norm
    return add


@pytest.mark.parametrize("mmax", [120.0, 60.0, 25.0, 8.0, 2.0, 0.9])
def test_remnant_mass_against_quadrature(mmax):
    a = (1.3, 2.3, 2.7)
    p = O.imf_vector(slopes=a, lower=0.08, upper=100.0)
    got = O.remnant_mass(2, p, mmax)
    ref = _remnant_ref(a, 0.08, 100.0, mmax)
    assert abs(got - ref) <= 1e-11 * max(ref, 1e-300)


def _np_bin_weights(a, lower, upper, mini, m1, m2, first):
    """Kroupa bin integrals per unit initial mass with numpy's expm1/log1p, per segment overlap."""
    a1, a2, a3 = a
    c2 = 0.5 ** (-a1 + a2)
    segs = [(0.08, 0.5, a1, 1.0), (0.5, 1.0, a2, c2), (1.0, np.inf, a3, c2)]

    def integ(lo, hi, mw):
        v = np.zeros_like(lo)
        for b0, b1, al, c in segs:
            x0 = np.maximum(lo, b0)
            x1 = np.minimum(hi, b1)
            ok = x1 > x0
            e = (2.0 if mw else 1.0) - al
            x0s = np.where(ok, x0, 1.0)
            x1s = np.where(ok, x1, 2.0)
            dl = np.log1p((x1s - x0s) / x0s)
            term = np.where(abs(e) > 0, x0s ** e * np.expm1(e * dl) / np.where(e == 0, 1.0, e), dl)
            v += np.where(ok, c * term, 0.0)
        return v

    tot = integ(np.array([lower]), np.array([upper]), True)[0]
    lo = np.where(first == 1, lower, m1)
    w = integ(lo, m2, False) / tot
    w[(mini < lower) | (mini > upper) | (m2 <= lo)] = 0.0
    return w, tot


def test_operator_model_sums_against_numpy():
    """The oracle batch with the IMF per model (no dust, z = 0, no smoothing, given SFH weights) against a numpy
    statement of the operator: IMF weights per point, per-age sums with remnants, the fold split at it_tesc
    (tested through young / old), the template products, the metallicity blend over two passes."""
    op = operator(3, 40, 300)
    g = grid(3)
    n = 6
    prm = synth.imf_params(n, seed=21)
    prm[SLOT["logzsol"]] = [-2.1, -1.3, -0.6, 0.0, 0.35, -0.9]
    prm[SLOT["dust_tesc"]] = 7.0
    w = synth.dirichlet_weights(n, g.nt, seed=9)
    b = np.zeros((1, g.nspec))
    b[0, 3000:3010] = 1.0
    r = O.csp_batch(g, b, prm, ssp_weights=w, add_igm=0, want=("mags", "spec", "mass", "lbol", "young", "old", "idx"),
                    ops=op, imf_type=2, remnants=1)
    assert r["status"] == 0
    nt = g.nt
    for m in range(n):
        it, zlo = r["idx"][m, 0], r["idx"][m, 1]
        dz = (prm[SLOT["logzsol"], m] - g.logz[zlo - 1]) / (g.logz[zlo] - g.logz[zlo - 1])
        a = prm[[SLOT["imf1"], SLOT["imf2"], SLOT["imf3"]], m]
        lower, upper = prm[SLOT["imf_lower_limit"], m], prm[SLOT["imf_upper_limit"], m]
        pv = O.imf_vector(slopes=a, lower=lower, upper=upper)
        vy = np.zeros(op.ntpl)
        vo = np.zeros(op.ntpl)
        smass = np.zeros(nt)
        lsum = np.zeros(nt)
        for iz, cz in ((zlo - 1, 1.0 - dz), (zlo, dz)):
            for t in range(nt):
                k0, k1 = op.age_ptr[iz * nt + t], op.age_ptr[iz * nt + t + 1]
                sl = slice(k0, k1)
                wp, tot = _np_bin_weights(a, lower, upper, op.mini[sl], op.m1[sl], op.m2[sl], op.first[sl])
                sm = np.sum(op.cm[sl] * wp) + _remnant_ref(a, lower, upper, op.age_mmax[iz * nt + t])
                smass[t] += cz * sm
                lsum[t] += cz * np.log10(max(np.sum(op.cl[sl] * wp), np.finfo(float).tiny))
                x = wp * w[m, t] * cz
                v = vy if t + 1 <= it else vo
                for j in range(op.nent):
                    np.add.at(v, op.ent_tpl[sl, j], op.ent_val[sl, j] * x)
        young = vy @ op.etpl
        old = vo @ op.etpl
        np.testing.assert_allclose(r["young"][m], young, rtol=1e-10, atol=1e-12 * young.max())
        np.testing.assert_allclose(r["old"][m], old, rtol=1e-10, atol=1e-12 * old.max())
        np.testing.assert_allclose(r["mass"][m], np.sum(w[m] * smass), rtol=1e-10)
        np.testing.assert_allclose(r["lbol"][m], np.log10(np.sum(w[m] * 10.0 ** lsum)), rtol=0, atol=1e-10)
        np.testing.assert_allclose(r["spec"][m], young + old, rtol=1e-10, atol=1e-12 * (young + old).max())
