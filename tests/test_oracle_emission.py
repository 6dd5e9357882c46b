"""Pins of the oracle's dust emission and AGN torus stages (orc_csp_batch with an Emission: dl07_indices /
dl07_shape, the absorbed light of attenuate_batch, dust_emission_reductions, add_dust_emission, add_agn_batch;
src/physics/opensps_csp_batch.f90:1569-1626, 3640-3656, 3963-4122, 4342-4433), at z = 0 without IGM or smoothing:

* energy balance: the emission added to the spectrum integrates over frequency (the batch's trapezoid) to the light
  the dust absorbs, which is the intrinsic spectrum (young + old) minus the attenuated one: a wrong absorbed-light
  term, a dropped birth-cloud or old-star screen, or a normalisation by the wrong sum fails it;
* at a template node (q_PAH, U_min on the grid, gamma 0 or 1) the emission's shape is that column of the table
  times the diffuse transmission; the mix in gamma is linear;
* the dust mass is 3.21e-3 / (4 pi) times the absorbed light over the template's integral: it doubles with the
  stellar light and halves with the templates doubled;
* the AGN torus adds fagn L_bol times the agn_tau template times the diffuse transmission, linear between nodes."""
import numpy as np
import pytest

from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

NSPEC, NT, NB = 900, 24, 6


def _setup():
    g = synth.ssp_grid(seed=3, nspec=NSPEC, nt=NT, nz=1)
    fs = synth.gaussian_filters(NB, seed=7)
    return g, O.build_bands(fs, g.lam)


def _prm(n, seed=5, birth=True):
    prm = synth.parametric_params(n, seed=seed, zred=(0.0, 0.0), sigma=(0.0, 0.0), logzsol=(0.0, 0.0))
    rng = np.random.default_rng(seed)
    prm[SLOT["dust2"]] = rng.uniform(0.1, 1.5, n)
    if birth:
        prm[SLOT["dust1"]] = rng.uniform(0.2, 1.0, n)
        prm[SLOT["frac_obrun"]] = rng.uniform(0.0, 0.4, n)
        prm[SLOT["dust3"]] = rng.uniform(0.0, 0.5, n)
        prm[SLOT["frac_nodust"]] = rng.uniform(0.0, 0.3, n)
    prm[SLOT["duste_qpah"]] = rng.uniform(0.5, 4.5, n)
    prm[SLOT["duste_umin"]] = rng.uniform(0.1, 25.0, n)
    prm[SLOT["duste_gamma"]] = rng.uniform(0.0, 0.5, n)
    return prm


def _run(g, b, prm, em, want=("spec", "young", "old", "mdust", "lbol")):
    r = O.csp_batch(g, b, prm, add_igm=0, want=want, emission=em)
    assert r["status"] == 0
    return r


def _nu_integral(lam, f):
    """The batch's frequency trapezoid: sum_i wt(i) (f(i) + f(i+1)), wt(i) = |c/l(i+1) - c/l(i)| / 2."""
    nu = 2.99792458e18 / lam
    wt = 0.5 * np.abs(np.diff(nu))
    return np.sum(wt * (f[..., :-1] + f[..., 1:]), axis=-1)


def _emission(g, **kw):
    return O.Emission(synth.DL07_QPAH, synth.DL07_UMIN, synth.dust_emission_templates(g.lam), synth.AGN_TAU,
                      synth.agn_templates(g.lam), **kw)


@pytest.mark.parametrize("birth", [False, True])
def test_energy_balance(birth):
    g, b = _setup()
    n = 8
    prm = _prm(n, 11, birth)
    on = _run(g, b, prm, _emission(g, add_agn_dust=0))
    off = _run(g, b, prm, _emission(g, add_dust_emission=0, add_agn_dust=0))
    np.testing.assert_array_equal(on["young"], off["young"])
    emitted = _nu_integral(g.lam, on["spec"] - off["spec"])
    absorbed = _nu_integral(g.lam, on["young"] + on["old"] - off["spec"])
    np.testing.assert_allclose(emitted, absorbed, rtol=1e-10)
    assert np.all(emitted > 0.0)
    # no dust, no emission
    p0 = prm.copy()
    p0[SLOT["dust1"]] = 0.0
    p0[SLOT["dust2"]] = 0.0
    r = _run(g, b, p0, _emission(g, add_agn_dust=0))
    r0 = _run(g, b, p0, _emission(g, add_dust_emission=0, add_agn_dust=0))
    np.testing.assert_array_equal(r["spec"], r0["spec"])
    assert np.all(r["mdust"] == np.finfo(np.float64).tiny)


def test_template_nodes_and_gamma():
    g, b = _setup()
    n = 3
    prm = _prm(n, 13, birth=False)
    tm = synth.dust_emission_templates(g.lam)
    prm[SLOT["duste_qpah"]] = synth.DL07_QPAH[[1, 3, 5]]
    prm[SLOT["duste_umin"]] = synth.DL07_UMIN[[2, 9, 17]]
    em = _emission(g, add_agn_dust=0)
    off = _run(g, b, prm, _emission(g, add_dust_emission=0, add_agn_dust=0))
    td = np.exp(-prm[SLOT["dust2"]][:, None] * (g.lam / 5500.0)[None, :] ** prm[SLOT["dust_index"]][:, None])
    shapes = {}
    for gam in (0.0, 1.0, 0.3):
        prm[SLOT["duste_gamma"]] = gam
        d = _run(g, b, prm, em)["spec"] - off["spec"]
        shapes[gam] = d / td
    ir = g.lam >= 2.0e4
    for m, (k, j) in enumerate(((1, 2), (3, 9), (5, 17))):
        for gam, comp in ((0.0, 0), (1.0, 1)):
            s = shapes[gam][m, ir]
            t = tm[k, 2 * j + comp, ir]
            ok = t > 1e-8 * t.max()
            ratio = s[ok] / t[ok]
            np.testing.assert_allclose(ratio, ratio[0], rtol=1e-9)
    # gamma mixes the two linearly before the normalisation: the normalised shapes are a convex combination
    for m in range(n):
        a0, a1, a3 = (shapes[x][m, ir] / _nu_integral(g.lam[ir], shapes[x][m, ir]) for x in (0.0, 1.0, 0.3))
        lam_ = a3 - a0
        coef = np.dot(lam_, a1 - a0) / np.dot(a1 - a0, a1 - a0)
        np.testing.assert_allclose(a3, a0 + coef * (a1 - a0), rtol=1e-8, atol=1e-12 * np.max(np.abs(a3)))


def test_mdust_scaling():
    g, b = _setup()
    n = 5
    prm = _prm(n, 17)
    em = _emission(g, add_agn_dust=0)
    r = _run(g, b, prm, em)
    g2 = synth.ssp_grid(seed=3, nspec=NSPEC, nt=NT, nz=1)
    g2.spec = 2.0 * g2.spec
    r2 = _run(g2, b, prm, em)
    np.testing.assert_allclose(r2["mdust"], 2.0 * r["mdust"], rtol=1e-12)
    em2 = O.Emission(synth.DL07_QPAH, synth.DL07_UMIN, 2.0 * synth.dust_emission_templates(g.lam),
                     add_agn_dust=0)
    r3 = _run(g, b, prm, em2)
    np.testing.assert_allclose(r3["mdust"], 0.5 * r["mdust"], rtol=1e-12)
    np.testing.assert_allclose(r3["spec"], r["spec"], rtol=1e-12)   # the shape is normalised away


def test_agn_torus():
    g, b = _setup()
    n = 4
    prm = _prm(n, 19, birth=False)
    prm[SLOT["fagn"]] = np.array([0.05, 0.2, 0.0, 1.0])
    prm[SLOT["agn_tau"]] = np.array([10.0, 42.0, 20.0, 400.0])   # a node, between nodes, off, beyond the grid
    on = _run(g, b, prm, _emission(g, add_dust_emission=0), want=("spec", "lbol"))
    off = _run(g, b, prm, _emission(g, add_dust_emission=0, add_agn_dust=0), want=("spec", "lbol"))
    tm = synth.agn_templates(g.lam)
    td = np.exp(-prm[SLOT["dust2"]][:, None] * (g.lam / 5500.0)[None, :] ** prm[SLOT["dust_index"]][:, None])
    lb = 10.0 ** on["lbol"]
    tmpl = np.stack([tm[1], (1.0 - 0.1) * tm[4] + 0.1 * tm[5], tm[2], tm[8]])
    want = prm[SLOT["fagn"]][:, None] * lb[:, None] * tmpl * td
    np.testing.assert_allclose(on["spec"] - off["spec"], want, rtol=1e-9, atol=1e-12 * np.max(np.abs(want)))
    np.testing.assert_array_equal(on["spec"][2], off["spec"][2])


def test_templates_on_the_grid():
    """emission_on_grid (the host's placement of the libraries, load_dust_emission_table / load_agn_dust_models):
    at the libraries' own wavelengths the table's values, below the pixel of 1 micron zero; a library linear in
    wavelength (a power law in log-log for the AGN) comes back exactly, beyond its ends too (extrapolated)."""
    g, _ = _setup()
    lam = g.lam
    q, u = synth.DL07_QPAH[:3], synth.DL07_UMIN[:4]
    dl = np.geomspace(2.0e4, 1.0e6, 40)
    a, b = 3.0e-25, 2.0e-31
    ds = np.zeros((q.size, u.size, 2, dl.size))
    for k in range(q.size):
        for j in range(u.size):
            for c in range(2):
                ds[k, j, c] = (k + 1) * a + (j + 1) * b * dl * (1 + c)
    at = synth.AGN_TAU[:3]
    al = np.geomspace(1.0e3, 1.0e7, 50)
    asp = np.stack([1e-14 * (al / 1e4) ** (-1.5 - 0.1 * k) for k in range(at.size)])
    em = O.emission_on_grid(lam, (q, u, dl, ds), (at, al, asp))
    i0 = max(int(O._find_interval_asc(lam, 1.0e4)), 1)
    assert np.all(em.dustem[:, :, :i0 - 1] == 0.0)
    x = lam[i0 - 1:]
    for k in range(q.size):
        for j in range(u.size):
            for c in range(2):
                np.testing.assert_allclose(em.dustem[k, 2 * j + c, i0 - 1:], (k + 1) * a + (j + 1) * b * x * (1 + c),
                                           rtol=1e-12)
    on = (lam >= al[0]) & (lam < al[-1])
    for k in range(at.size):
        np.testing.assert_allclose(em.agn[k, on], 1e-14 * (lam[on] / 1e4) ** (-1.5 - 0.1 * k), rtol=1e-10)
    i1, i2 = (max(int(O._find_interval_asc(lam, v)), 1) for v in (al[0], al[-1]))
    assert np.all(em.agn[:, :i1 - 1] == 0.0) and np.all(em.agn[:, i2:] == 0.0)
    # the reference's find_interval: at a node the interval that starts there, at the last node the last interval
    assert int(O._find_interval_asc(dl, dl[0])) == 1 and int(O._find_interval_asc(dl, dl[-1])) == dl.size - 1
    assert int(O._find_interval_asc(dl, dl[5])) == 6 and int(O._find_interval_asc(dl, 0.5 * dl[0])) == 0
