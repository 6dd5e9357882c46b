"""Pins of the oracle's reverse pass (orc_csp_batch_grad, the restatement of batch_backward,
src/physics/opensps_csp_batch.f90:4713-5136) by central finite differences of the oracle's forward batch
(orc_csp_batch), h = 1e-6 relative, on a small grid: every params_bar row this build computes, entries of
weights_bar and weights_z_bar, and the mass / L_bol cotangents. A dropped term (the birth cloud's or the old
stars' screen, the IGM's dz, the distance term, the smoothing's sigma), a wrong sign or a transposed operand
(basis^T y_bar) fails one of them."""
import numpy as np
import pytest

from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

NSPEC, NT, NZ, NB = 900, 24, 4, 12


def _params(n, seed, birth):
    """The adjoint workload's parameters with a velocity dispersion the coarse test grid resolves (its pixels are
    ~4600 km/s wide; the workload's 50-300 km/s would smooth nothing here)."""
    prm = synth.adjoint_params(n, seed=seed, birth_cloud=birth)
    prm[SLOT["sigma_smooth"]] = np.linspace(3000.0, 9000.0, n)
    return prm


def _setup(nz=NZ):
    g = synth.ssp_grid(seed=3, nspec=NSPEC, nt=NT, nz=nz, )
    fs = synth.gaussian_filters(NB, seed=7)
    return g, O.build_bands(fs, g.lam), O.cosmology_table(72.0, 0.27, 0.73)


def _loss(fwd, cot):
    L = 0.0
    for k in ("mags", "spec", "mass", "lbol"):
        if k in cot:
            v = np.where(np.isfinite(fwd[k]), fwd[k], 0.0)
            L = L + np.sum(cot[k] * v, axis=tuple(range(1, v.ndim)))
    return L   # per model


def _cot(n, g, seed=5):
    rng = np.random.default_rng(seed)
    return {"mags": rng.standard_normal((n, NB)), "spec": rng.standard_normal((n, NSPEC)) * 1e3,
            "mass": rng.standard_normal(n), "lbol": rng.standard_normal(n)}


def _fwd(g, b, cz, prm, **kw):
    r = O.csp_batch(g, b, prm, cosmo=cz, want=("mags", "spec", "mass", "lbol"), **kw)
    assert r["status"] == 0
    return r


def _grad(g, b, cz, prm, cot, **kw):
    r = O.csp_batch_grad(g, b, prm, cosmo=cz, mags_bar=cot["mags"], spec_bar=cot["spec"], mass_bar=cot["mass"],
                         lbol_bar=cot["lbol"], **kw)
    assert r["status"] == 0
    return r


def _fd_param(g, b, cz, prm, cot, slot, rel=1e-6, **kw):
    x = prm[slot].copy()
    h = np.where(x != 0.0, np.abs(x) * rel, rel)
    pp, pm = prm.copy(), prm.copy()
    pp[slot] = x + h
    pm[slot] = x - h
    return (_loss(_fwd(g, b, cz, pp, **kw), cot) - _loss(_fwd(g, b, cz, pm, **kw), cot)) / (2 * h)


def _check(fd, an, tol, what):
    scale = np.maximum(np.abs(fd), 1e-4 * np.max(np.abs(fd)) + 1e-300)   # (entries below FD resolution)
    err = np.abs(an - fd) / scale
    assert np.all(err < tol), f"{what}: fd {fd} analytic {an} err {err}"


@pytest.mark.parametrize("birth", [False, True])
def test_params_bar_weights_z(birth):
    g, b, cz = _setup()
    n = 6
    prm = _params(n, 29, birth)
    prm[SLOT["zred"]] = np.array([0.0, 0.01, 0.8, 1.7, 2.9, 3.9])   # z = 0: no IGM, no distance term
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=31)
    cot = _cot(n, g)
    an = _grad(g, b, cz, prm, cot, ssp_weights_z=wz)["params_bar"]
    rows = ["dust2", "dust_index", "zred", "sigma_smooth", "igm_factor"]
    if birth:
        rows += ["dust1", "dust1_index", "dust3", "frac_obrun", "frac_nodust"]
    for name in rows:
        fd = _fd_param(g, b, cz, prm, cot, SLOT[name], ssp_weights_z=wz)
        a = an[:, O.GRAD_ROWS[name]]
        if name == "zred":
            # z = 0 is the rest frame (absolute magnitudes, no IGM): the reference's derivative there is zero
            # (distance_term_dz), and the forward is not continuous across it
            assert a[0] == 0.0
            fd, a = fd[1:], a[1:]
        _check(fd, a, 2e-4, name)
    # rows without a dependence in this batch (the SFH parameters, logzsol with weights over metallicities)
    assert np.all(an[:, O.GRAD_ROWS["logzsol"]] == 0.0)


def test_weights_z_bar_entries():
    g, b, cz = _setup()
    n = 4
    prm = _params(n, 37, True)
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=41)
    cot = _cot(n, g, seed=9)
    an = _grad(g, b, cz, prm, cot, ssp_weights_z=wz)["weights_z_bar"]
    # the entries where the models hold the most mass (an entry far from every model's enrichment history
    # holds weights too small to difference)
    order = np.argsort(-np.median(wz, axis=0).ravel())
    checked = 0
    for e in order[:12:2]:
        zz, t = divmod(int(e), NT)
        h = 1e-6 * wz[:].sum(axis=(1, 2))   # 1e-6 of the mass formed (the weights enter the sums linearly)
        wp, wm = wz.copy(), wz.copy()
        wp[:, zz, t] += h
        wm[:, zz, t] -= h
        fd = (_loss(_fwd(g, b, cz, prm, ssp_weights_z=wp), cot) - _loss(_fwd(g, b, cz, prm, ssp_weights_z=wm), cot)) / (2 * h)
        # (a weight within h of zero meets the forward's clamp of weights <= tiny: one-sided there, as the
        # reference's adjoint, which is the derivative from above)
        ok = wz[:, zz, t] > 2 * h
        if ok.any():
            _check(fd[ok], an[ok, zz, t], 1e-5, f"weights_z_bar[{zz},{t}]")
            checked += int(ok.sum())
    assert checked >= 10


def test_blended_weights_and_logzsol():
    g, b, cz = _setup()
    n = 5
    prm = _params(n, 43, True)
    prm[SLOT["logzsol"]] = np.linspace(-1.9, 0.2, n)
    w = synth.dirichlet_weights(n, NT, seed=47, alpha=2.0)
    cot = _cot(n, g, seed=13)
    r = _grad(g, b, cz, prm, cot, ssp_weights=w)
    fd = _fd_param(g, b, cz, prm, cot, SLOT["logzsol"], ssp_weights=w)
    _check(fd, r["params_bar"][:, O.GRAD_ROWS["logzsol"]], 1e-5, "logzsol")
    for t in (0, 5, NT - 1):
        h = 1e-6 * w.sum(axis=1)
        wp, wm = w.copy(), w.copy()
        wp[:, t] += h
        wm[:, t] -= h
        fd = (_loss(_fwd(g, b, cz, prm, ssp_weights=wp), cot) - _loss(_fwd(g, b, cz, prm, ssp_weights=wm), cot)) / (2 * h)
        _check(fd, r["weights_bar"][:, t], 1e-5, f"weights_bar[{t}]")


def test_grad_forward_matches_batch():
    """The reverse pass's own forward is the batch's (its magnitudes)."""
    g, b, cz = _setup()
    n = 5
    prm = _params(n, 53, True)
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=59)
    r = _grad(g, b, cz, prm, _cot(n, g), ssp_weights_z=wz)
    f = _fwd(g, b, cz, prm, ssp_weights_z=wz)
    np.testing.assert_allclose(r["mags"], f["mags"], rtol=0, atol=1e-9)


def _sfh_models(sfh, g):
    """Six models of SFH type sfh with const, a burst, a start time and truncation, their window ends and burst
    ages away from the coarse test grid's ages (where the weights have kinks)."""
    rng = np.random.default_rng(97 + sfh)
    prm = _params(6, 61 + sfh, True)
    prm[SLOT["sfh"]] = sfh
    for m in range(6):
        while True:
            tage = rng.uniform(0.3, 12.0)
            start = tage * rng.uniform(0.05, 0.3) if sfh != 0 and m % 2 else 0.0
            trunc = start + (tage - start) * rng.uniform(0.2, 0.8) if sfh != 0 and m % 3 else 0.0
            tburst = tage * rng.uniform(0.1, 0.9)
            slope = rng.uniform(-1.5, 1.0) if sfh == 5 else 0.0
            ends = [(tage - start) * 1e9, tage * 1e9, (tage - tburst) * 1e9]
            if trunc > 0.0:
                ends.append((tage - trunc) * 1e9)
                if slope < 0.0 and (tage - trunc) * 1e9 + 1e9 / slope > 0.0:   # Simha's zero crossing t0
                    ends.append((tage - trunc) * 1e9 + 1e9 / slope)
            if all(np.min(np.abs(g.log_age - np.log10(e))) > 2e-3 for e in ends):
                break
        prm[SLOT["tage"], m] = tage
        prm[SLOT["tau"], m] = rng.uniform(0.3, 8.0)
        prm[SLOT["const"], m] = rng.uniform(0.05, 0.3) if m % 2 == 0 else 0.0
        prm[SLOT["fburst"], m] = rng.uniform(0.05, 0.3) if m < 4 else 0.0
        prm[SLOT["tburst"], m] = tburst
        prm[SLOT["sf_start"], m] = start
        prm[SLOT["sf_trunc"], m] = trunc
        prm[SLOT["sf_slope"], m] = slope
    return prm


@pytest.mark.parametrize("sfh", [0, 1, 4, 5])
def test_sfh_rows(sfh):
    """params_bar's SFH rows (batch_sfh_gradients, opensps_csp_batch.f90:5198-5217) against central differences of
    the oracle's forward batch (h = 1e-4 relative: the SFH weights round at ~1e-10); const and fburst at zero against
    the upward difference; sf_slope's row is zero (Simha's linear part only)."""
    g, b, cz = _setup()
    prm = _sfh_models(sfh, g)
    n = prm.shape[1]
    cot = _cot(n, g, seed=19)
    an = _grad(g, b, cz, prm, cot)["params_bar"]
    names = ["tage"] if sfh == 0 else (["tage", "tau", "sf_start", "sf_trunc", "sf_slope"] if sfh == 5 else
                                       ["tage", "tau", "const", "fburst", "tburst", "sf_start", "sf_trunc"])
    for name in names:
        s = SLOT[name]
        x = prm[s].copy()
        h = np.where(x != 0.0, np.abs(x) * 1e-4, 1e-5)
        pp, pm = prm.copy(), prm.copy()
        pp[s] = x + h
        # upward at zero: const and fburst switch the mixture on there, a negative sf_start turns truncation on
        pm[s] = np.where(x != 0.0, x - h, x)
        fd = (_loss(_fwd(g, b, cz, pp), cot) - _loss(_fwd(g, b, cz, pm), cot)) / np.where(x != 0.0, 2 * h, h)
        a = an[:, O.GRAD_ROWS[name]]
        ok = (x != 0.0) | (name != "sf_trunc")
        if name == "sf_trunc":   # zero at zero: no truncation (a small positive sf_trunc truncates almost all)
            assert np.all(a[x == 0.0] == 0.0), name
        _check(fd[ok], a[ok], 2e-3 if np.any(x[ok] == 0.0) else 1e-4, name)
    zero = {0: ("sf_slope", "tau", "const", "fburst", "tburst", "sf_start", "sf_trunc"),
            5: ("const", "fburst", "tburst")}.get(sfh, ("sf_slope",))
    for name in zero:
        assert np.all(an[:, O.GRAD_ROWS[name]] == 0.0), name


def test_sfh_rows_not_with_given_weights():
    """With weights given (ssp_weights or ssp_weights_z) the SFH rows stay zero (opensps_csp_batch.f90:1163-1166)."""
    g, b, cz = _setup()
    prm = _sfh_models(4, g)
    n = prm.shape[1]
    w = synth.dirichlet_weights(n, NT, seed=3, alpha=2.0)
    an = _grad(g, b, cz, prm, _cot(n, g, seed=19), ssp_weights=w)["params_bar"]
    for name in O.SFHT_PARAMS:
        assert np.all(an[:, O.GRAD_ROWS[name]] == 0.0), name


@pytest.mark.parametrize("imf_type", [2, 4])
def test_params_bar_operator(imf_type):
    """The operator's reverse pass (the IMF per model): the IMF rows (imf1, imf2 and imf3 for Kroupa, mdave for Dave,
    a context setting differenced there), logzsol through both metallicities' operators and a dust row, against
    central differences of the oracle's operator batch with every cotangent (magnitudes, spectrum, mass, L_bol) in
    the loss. (The SFH rows come from w_bar, pinned below, by the path the resident bases' tests pin.)"""
    from helpers import bands, cosmo, grid, operator
    nz = 3
    op, g, b, cz = operator(nz, 40, 300), grid(nz), bands(NB), cosmo()
    n = 4
    prm = synth.parametric_params(n, seed=61, logzsol=(-2.1, 0.45))
    prm = synth.imf_params(n, seed=62, p=prm)
    prm[SLOT["dust1"]] = 0.5   # a birth-cloud screen: the young and old sums apart
    kw = dict(ops=op, imf_type=imf_type, remnants=1, mdave=0.6)
    rng = np.random.default_rng(63)
    cot = {"mags": rng.standard_normal((n, NB)), "spec": rng.standard_normal((n, g.nspec)) * 1e3,
           "mass": rng.standard_normal(n), "lbol": rng.standard_normal(n)}
    pb = _grad(g, b, cz, prm, cot, **kw)["params_bar"]
    names = ["imf1", "imf2"] + (["imf3"] if imf_type == 2 else []) + ["logzsol", "dust2"]
    for name in names:
        fd = _fd_param(g, b, cz, prm, cot, SLOT[name], **kw)
        _check(fd, pb[:, O.GRAD_ROWS[name]], 2e-5, name)
    if imf_type == 4:
        h = 0.6e-6
        lp = _loss(_fwd(g, b, cz, prm, **{**kw, "mdave": 0.6 + h}), cot)
        lm = _loss(_fwd(g, b, cz, prm, **{**kw, "mdave": 0.6 - h}), cot)
        _check((lp - lm) / (2 * h), pb[:, O.GRAD_ROWS["mdave"]], 2e-5, "mdave")
    else:
        assert np.all(pb[:, O.GRAD_ROWS["mdave"]] == 0.0)


def test_weights_bar_operator():
    """The operator's w_bar (the fold's transpose plus the per-age mass and L_bol terms) with given SSP weights,
    against central differences along single weights (h = 1e-4 relative: the loss is smooth in a weight, and a
    smaller step meets the loss's rounding)."""
    from helpers import bands, cosmo, grid, operator
    nz = 3
    op, g, b, cz = operator(nz, 40, 300), grid(nz), bands(NB), cosmo()
    n = 3
    prm = synth.parametric_params(n, seed=71, logzsol=(-2.1, 0.45))
    prm = synth.imf_params(n, seed=72, p=prm)
    w = synth.dirichlet_weights(n, g.nt, seed=73)
    kw = dict(ops=op, imf_type=2, remnants=1)
    rng = np.random.default_rng(74)
    cot = {"mags": rng.standard_normal((n, NB)), "spec": rng.standard_normal((n, g.nspec)) * 1e3,
           "mass": rng.standard_normal(n), "lbol": rng.standard_normal(n)}
    wb = _grad(g, b, cz, prm, cot, ssp_weights=w, **kw)["weights_bar"]
    for t in (3, 40, 90):
        h = 1e-4 * np.maximum(w[:, t], 1e-3)
        wp, wm = w.copy(), w.copy()
        wp[:, t] += h
        wm[:, t] -= h
        fd = (_loss(_fwd(g, b, cz, prm, ssp_weights=wp, **kw), cot) -
              _loss(_fwd(g, b, cz, prm, ssp_weights=wm, **kw), cot)) / (2 * h)
        _check(fd, wb[:, t], 2e-5, f"weights_bar[{t}]")


def test_zred_mags():
    """Magnitudes at a second redshift (zred_mags): without the IGM they equal the batch run at that redshift; with
    it, zred's row is the IGM's part alone (the observed frame and distance follow zred_mags), against central
    differences in zred with zred_mags held; the other rows against differences as usual."""
    g, b, cz = _setup()
    n = 5
    prm = _params(n, 91, True)
    prm[SLOT["zred"]] = np.array([0.0, 0.4, 1.3, 2.6, 3.7])
    zm = np.array([0.2, 0.0, 1.9, 2.1, 3.7])
    w = synth.dirichlet_weights(n, g.nt, seed=95)
    base = prm.copy()
    base[SLOT["zred"]] = zm
    a = _fwd(g, b, cz, prm, ssp_weights=w, zred_mags=zm, add_igm=0)
    r = _fwd(g, b, cz, base, ssp_weights=w, add_igm=0)
    np.testing.assert_array_equal(a["mags"], r["mags"])
    np.testing.assert_array_equal(a["spec"], r["spec"])
    cot = _cot(n, g, seed=93)
    pb = _grad(g, b, cz, prm, cot, ssp_weights=w, zred_mags=zm)["params_bar"]
    for name in ("zred", "dust2", "sigma_smooth", "igm_factor"):
        fd = _fd_param(g, b, cz, prm, cot, SLOT[name], ssp_weights=w, zred_mags=zm)
        k = slice(1, None) if name == "zred" else slice(None)   # (the IGM switches on above z = 0: a step there)
        _check(fd[k], pb[k, O.GRAD_ROWS[name]], 2e-4, name)
    assert pb[0, O.GRAD_ROWS["zred"]] == 0.0   # z = 0: no IGM, and the magnitudes do not follow zred
