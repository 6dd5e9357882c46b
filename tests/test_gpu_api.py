"""The public C API on the GPU (libopensps.so -> libopensps_hip.so) against the oracle: the drop-in path a
reference caller takes (names[] / values[] in, host arrays out), as the reference's tests/test_c_batch.c does."""
import numpy as np
import pytest

from helpers import grid, mag_err, spec_rel_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

TOL = {"double": (1e-9, 1e-9), "mixed": (1e-4, 1e-3)}


def _ctx(g, fs, precision, igm=1):
    from libopensps_amd import api
    ctx = api.Context()
    ctx.set_ssp_basis(g)
    ctx.set_filters(fs)
    ctx.set_int("add_dust_emission", 0)
    ctx.set_int("add_igm_absorption", igm)
    ctx.set_str("kernel_precision", precision)
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_weight_matrix(precision):
    g = grid(1)
    fs = synth.gaussian_filters(20)
    n = 48
    w = synth.dirichlet_weights(n, g.nt, seed=21)
    print(f"case: api weight matrix n={n} {precision}", flush=True)
    with _ctx(g, fs, precision, igm=0) as ctx:
        r = ctx.compute_csp_batch({}, n=n, ssp_weights=w, want=("mags", "spec", "mass", "lbol", "sfr"))
        bands = ctx.bands()
    ref = O.csp_batch(g, bands, synth.default_params(n), ssp_weights=w, add_igm=0, cosmo=O.cosmology_table())
    ts, tm = TOL[precision]
    assert spec_rel_err(r["spec"], ref["spec"]) < ts
    assert mag_err(r["mags"], ref["mags"]) < tm
    np.testing.assert_allclose(r["mass"], ref["mass"], rtol=1e-12)
    assert np.all(np.isnan(r["sfr"]))    # SFR is not defined with a weight matrix


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_parametric(precision):
    g = grid(22)
    fs = synth.gaussian_filters(100)
    n = 96
    prm = synth.parametric_params(n, seed=23)
    print(f"case: api parametric n={n} {precision}", flush=True)
    with _ctx(g, fs, precision) as ctx:
        ctx.set_int("keep_maggies", 1)
        r = ctx.compute_csp_batch_params(prm, synth.varied_names(), want=("mags", "spec", "mass", "lbol", "sfr"))
        bands = ctx.bands()
        mg = ctx.batch_maggies()
        p, nspec, nn, nbytes = ctx.batch_spec_device()
        assert p and nspec == g.nspec and nn == n and nbytes == (8 if precision == "double" else 4)
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table())
    ts, tm = TOL[precision]
    assert spec_rel_err(r["spec"], ref["spec"]) < ts
    assert mag_err(r["mags"], ref["mags"]) < tm
    # maggies = 10^(-0.4 m) of the AB magnitudes, distance included
    ok = np.isfinite(r["mags"])
    np.testing.assert_allclose(mg[ok], 10.0 ** (-0.4 * r["mags"][ok]), rtol=2e-3)
    assert np.all(r["sfr"] > 0)


@pytest.mark.gpu
def test_api_parts_and_masses():
    """compute_csp_batch_parts: young / old spectra and the masses and L_bol of the two parts."""
    g = grid(22)
    fs = synth.gaussian_filters(20)
    n = 32
    prm = synth.parametric_params(n, seed=25)
    prm[SLOT["dust1"]] = 0.8
    names = synth.varied_names() + ["dust1"]
    with _ctx(g, fs, "double") as ctx:
        r = ctx.compute_csp_batch_params(prm, names, parts=True,
                                         want=("mags", "spec", "mass", "lbol", "mass_young", "mass_old", "lbol_young",
                                               "lbol_old"))
        bands = ctx.bands()
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table(),
                      want=("mags", "spec", "mass", "young", "old"))
    assert spec_rel_err(r["young"], ref["young"]) < 1e-9
    assert spec_rel_err(r["old"], ref["old"]) < 1e-9
    np.testing.assert_allclose(r["mass_young"] + r["mass_old"], r["mass"], rtol=1e-12)
    lsum = 10 ** r["lbol_young"] + 10 ** r["lbol_old"]
    np.testing.assert_allclose(np.log10(lsum), r["lbol"], atol=1e-12)


@pytest.mark.gpu
def test_api_device_resident_batch():
    """opensps_context_compute_csp_batch_device: the same numbers as the host call, on torch device buffers."""
    import torch
    g = grid(22)
    fs = synth.gaussian_filters(50)
    n = 64
    prm = synth.parametric_params(n, seed=27)
    with _ctx(g, fs, "mixed") as ctx:
        r = ctx.compute_csp_batch_params(prm, synth.varied_names())
        d_prm = torch.as_tensor(prm, device="cuda")
        d_mags = torch.empty((n, ctx.nbands), dtype=torch.float64, device="cuda")
        d_spec = torch.empty((n, g.nspec), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        ctx.compute_csp_batch_device(d_prm, d_mags, d_spec=d_spec)
        np.testing.assert_array_equal(d_mags.cpu().numpy(), r["mags"])
        np.testing.assert_array_equal(d_spec.cpu().numpy().astype(np.float64), r["spec"])


@pytest.mark.gpu
def test_api_device_buffers_checked():
    """A device buffer the batch would overrun is refused on the host (status 3) before anything is launched: a float
    spectrum buffer for an FP64 batch, too few magnitudes, a parameter array of the wrong shape."""
    import torch
    from libopensps_amd import api
    g = grid(3)
    fs = synth.gaussian_filters(5)
    n = 8
    prm = synth.parametric_params(n, seed=29)
    with _ctx(g, fs, "mixed") as ctx:
        d_prm = torch.as_tensor(prm, device="cuda")
        d_mags = torch.empty((n, ctx.nbands), dtype=torch.float64, device="cuda")
        f32 = torch.empty((n, g.nspec), dtype=torch.float32, device="cuda")
        cases = [dict(d_spec=f32, precision=0),
                 dict(d_spec=f32.double(), precision=1),
                 dict(d_spec=f32, d_mass=torch.empty(n - 1, dtype=torch.float64, device="cuda"))]
        for kw in cases:
            with pytest.raises(api.OpenSPSError) as e:
                ctx.compute_csp_batch_device(d_prm, d_mags, **kw)
            assert e.value.status == 3
        with pytest.raises(api.OpenSPSError):
            ctx.compute_csp_batch_device(d_prm, d_mags[:-1])
        with pytest.raises(api.OpenSPSError):
            ctx.compute_csp_batch_device(d_prm[:-1].contiguous(), d_mags, d_spec=f32)
        ctx.compute_csp_batch_device(d_prm, d_mags, d_spec=f32)   # the matching buffers still run
        torch.cuda.synchronize()
        assert np.isfinite(d_mags.cpu().numpy()).any()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_vega_system(precision):
    """compute_vega_mags = 1 with the reference spectra of the test data release (tests/golden/opensps-test-data):
    vref = 0 gives AB - magvega in every band, vref = 3 keeps band 3 in AB (computed even where the mask leaves it
    out) and shifts the others relative to it (the reference's tests/test_filters.f90:575-593 and
    test_c_filters.c:120-128); the linear fluxes stay AB; the gradients' forward carries the Vega magnitudes."""
    import os
    from libopensps_amd import datalib
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "opensps-test-data")
    vega, sun = datalib.reference_spectrum(root, "vega"), datalib.reference_spectrum(root, "sun")
    g = grid(1)
    fs = synth.gaussian_filters(12, seed=9)
    n = 40
    prm = synth.parametric_params(n, seed=29)
    names = synth.varied_names()
    print(f"case: api vega n={n} {precision}", flush=True)
    with _ctx(g, fs, precision) as ctx:
        ctx.set_reference_spectra(vega, sun)
        ctx.set_int("keep_maggies", 1)
        ab = ctx.compute_csp_batch_params(prm, names, want=("mags",))["mags"]
        mg_ab = ctx.batch_maggies()
        _, mv, _ = ctx.filter_data()
        assert np.all(np.abs(mv) < 99.0)
        ctx.set_int("compute_vega_mags", 1)
        v0 = ctx.compute_csp_batch_params(prm, names, want=("mags",))["mags"]
        mg_v = ctx.batch_maggies()
        np.testing.assert_allclose(v0, O.vega_mags(ab, mv, 0), rtol=0, atol=1e-12)
        np.testing.assert_allclose(mg_v, mg_ab, rtol=1e-14, atol=0)
        ctx.set_filters(fs, vega_ref_band=3)
        mask = np.ones(12, dtype=np.int32)
        mask[2] = 0
        mask[5] = 0
        ctx.set_mag_compute(mask)
        v3 = ctx.compute_csp_batch_params(prm, names, want=("mags",))["mags"]
        ref3 = O.vega_mags(ab, mv, 3)
        ref3[:, 5] = np.nan
        np.testing.assert_array_equal(np.isnan(v3), np.isnan(ref3))
        ok = ~np.isnan(ref3)
        np.testing.assert_allclose(v3[ok], ref3[ok], rtol=0, atol=1e-12)
        # the reverse pass's forward half: the same Vega magnitudes (its young and old sums are always apart: with
        # FP32 spectra they round differently from the single sum), finite gradients
        r = ctx.compute_csp_batch_grad({k: prm[SLOT[k]] for k in names}, n=n, mags_bar=np.ones((n, 12)),
                                       want=("mags",), grads=("params_bar",))
        np.testing.assert_allclose(r["mags"][ok], ref3[ok], rtol=0, atol=1e-12 if precision == "double" else 1e-6)
        assert np.all(np.isfinite(r["params_bar"]))


@pytest.mark.gpu
def test_api_tabular_sfr():
    """sfr of a tabular model: the table interpolated linearly at its age, M_sun / yr, at least 0."""
    g = grid(1)
    fs = synth.gaussian_filters(4)
    age = np.array([0.0, 1.0, 3.0, 6.0, 10.0])
    sfr = np.array([0.0, 5.0, 2.0, 8.0, 1.0])
    tage = np.array([0.5, 2.0, 3.0, 7.5, 9.9])
    print("case: api tabular sfr", flush=True)
    with _ctx(g, fs, "double", igm=0) as ctx:
        ctx.set_sfh_tab(age, sfr)
        r = ctx.compute_csp_batch({"sfh": np.full(5, 3.0), "tage": tage}, n=5, want=("mags", "sfr", "mass"))
    np.testing.assert_allclose(r["sfr"], np.interp(tage, age, sfr), rtol=1e-13)
    assert np.all(r["mass"] > 0)


@pytest.mark.gpu
def test_api_sfr_output():
    """The sfr output (batch_model_indices, opensps_csp_batch.f90:1556-1560: get_sfh_properties_at_age's SFR over the
    fraction formed) is the SFR at the model's age per unit mass formed by then, SFR(t') / integral_0^t' SFR, per
    year: against quadrature of the SFR for tau (1), delayed tau (4) and Simha (5: truncation, slopes of both signs)."""
    from scipy import integrate
    g = grid(1)
    fs = synth.gaussian_filters(4)
    cases = [(1, 6.0, 2.0, 0.0, 0.0), (4, 9.0, 3.0, 0.0, 0.0), (4, 0.5, 7.0, 0.0, 0.0), (5, 10.0, 2.5, 6.0, -0.2),
             (5, 12.0, 1.5, 4.0, 0.3), (5, 8.0, 4.0, 7.5, -2.0), (5, 5.0, 3.0, 0.0, -0.5)]
    n = len(cases)
    vals = {k: np.array([c[i] for c in cases], dtype=float)
            for i, k in enumerate(("sfh", "tage", "tau", "sf_trunc", "sf_slope"))}
    print(f"case: api sfr n={n}", flush=True)
    with _ctx(g, fs, "double", igm=0) as ctx:
        r = ctx.compute_csp_batch(vals, n=n, want=("mags", "sfr"))
    for m, (sfh, tage, tau, trunc, slope) in enumerate(cases):
        on = trunc > 0.0 and trunc < tage
        s_tr = (trunc / tau) * np.exp(-trunc / tau)

        def sfr(x):   # forward time since the start, Gyr
            if sfh == 1:
                return np.exp(-x / tau)
            if sfh == 4 or not on or x <= trunc:
                return (x / tau) * np.exp(-x / tau)
            return s_tr * max(1.0 - (-slope) * (x - trunc), 0.0)
        pts = [trunc] if on else None
        mass = integrate.quad(sfr, 0.0, tage, epsabs=0, epsrel=1e-12, limit=200, points=pts)[0]
        want = sfr(tage) / mass / 1e9
        assert abs(r["sfr"][m] - want) <= 1e-9 * abs(want) + 1e-30, (cases[m], r["sfr"][m], want)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_filter_library_batch(precision):
    """The parametric batch through the 159 filters of the test release's allfilters (narrow and wide, calibrated IR
    bands, supports from the UV to the far IR), against the oracle with the host's bands."""
    import os
    from libopensps_amd import datalib
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "opensps-test-data")
    g = grid(22)
    fl = datalib.filter_library(root, "allfilters")
    n = 64
    prm = synth.parametric_params(n, seed=37)
    print(f"case: api allfilters n={n} {precision}", flush=True)
    with _ctx(g, synth.gaussian_filters(4), precision) as ctx:
        ctx.set_filters(fl)
        r = ctx.compute_csp_batch_params(prm, synth.varied_names(), want=("mags", "spec"))
        bands = ctx.bands()
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table())
    ts, tm = TOL[precision]
    assert spec_rel_err(r["spec"], ref["spec"]) < ts
    assert mag_err(r["mags"], ref["mags"]) < tm


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("dust_type", [2, 6])
def test_api_fixed_dust_curves(precision, dust_type):
    """dust_type 2 (Calzetti) and 6 (Reddy): the host's curve on the grid, the chain's diffuse and old-star screens,
    against the oracle with the papers' curves; the reverse pass with such a curve is status 2."""
    g = grid(22)
    fs = synth.gaussian_filters(30)
    n = 40
    prm = synth.parametric_params(n, seed=41)
    prm[SLOT["dust3"]] = np.linspace(0.0, 0.4, n)
    names = synth.varied_names() + ["dust3"]
    print(f"case: api dust_type {dust_type} {precision}", flush=True)
    with _ctx(g, fs, precision) as ctx:
        ctx.set_int("dust_type", dust_type)
        r = ctx.compute_csp_batch_params(prm, names, want=("mags", "spec"))
        bands = ctx.bands()
        rg = ctx.compute_csp_batch_grad({k: prm[SLOT[k]] for k in names}, n=n, mags_bar=np.ones((n, 30)),
                                        grads=("params_bar",), raw_status=True)
        assert rg["status"] == 2
    curve = O.calzetti_curve(g.lam) if dust_type == 2 else O.reddy_curve(g.lam)
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table(), dust_curve=curve)
    ts, tm = TOL[precision]
    assert spec_rel_err(r["spec"], ref["spec"]) < ts
    if precision == "double":
        assert mag_err(r["mags"], ref["mags"]) < tm
    else:   # FP32 spectra: bands within 40 mag of each model's brightest (fainter far-UV ones can underflow to 0)
        q = ref["mags"]
        bright = np.nanmin(np.where(np.isfinite(q), q, np.nan), axis=1, keepdims=True)
        ok = np.isfinite(q) & (q - bright <= 40.0)
        assert np.all(np.isfinite(r["mags"][ok]))
        assert float(np.max(np.abs(r["mags"][ok] - q[ok]))) < tm



def _imf_models(n, nz, seed):
    prm = synth.parametric_params(n, seed=seed, logzsol=(-2.1, 0.45))
    prm = synth.imf_params(n, seed=seed + 1, p=prm)
    prm[SLOT["imf_lower_limit"]] = np.linspace(0.08, 0.6, n)
    return prm, synth.varied_names() + ["imf1", "imf2", "imf3", "imf_upper_limit", "imf_lower_limit"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_imf_per_model(precision):
    """The IMF per model through the public C API: an operator given by opensps_context_set_ssp_ops, Kroupa slopes
    and limits per model as names[] / values[] (the reference's IMF_PARAMS with ssp_ops), against the oracle's
    operator path; without the operator the same call is status 2."""
    from helpers import operator
    nz = 3
    g = grid(nz)
    op = operator(nz, 40, 300)
    fs = synth.gaussian_filters(20)
    n = 64
    prm, names = _imf_models(n, nz, 51)
    print(f"case: api imf per model n={n} {precision}", flush=True)
    with _ctx(g, fs, precision) as ctx:
        r0 = ctx.compute_csp_batch_params(prm, names, raw_status=True)
        assert r0["status"] == 2
        ctx.set_ssp_ops(op)
        assert ctx.has_ssp_ops()
        r = ctx.compute_csp_batch_params(prm, names, want=("mags", "spec", "mass", "lbol"))
        bands = ctx.bands()
        # an SSP parameter of the context changed: the operator is dropped (the reference's invalidate_ssps)
        ctx.set_int("imf_type", 0)
        assert not ctx.has_ssp_ops()
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table(), ops=op, imf_type=2, remnants=1)
    ts, tm = TOL[precision]
    es, em = spec_rel_err(r["spec"], ref["spec"]), mag_err(r["mags"], ref["mags"])
    print(f"api imf per model: spectra {es:.2e}, mags {em:.2e}")
    assert es < ts and em < tm
    np.testing.assert_allclose(r["mass"], ref["mass"], rtol=1e-9)
    np.testing.assert_allclose(r["lbol"], ref["lbol"], rtol=0, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_imf_grad(precision):
    """The batch with gradients and the IMF per model through the public C API (compute_csp_batch_grad with an
    operator on the context): params_bar's IMF rows (imf1 .. imf3), logzsol and the dust and SFH rows against the
    oracle's operator reverse pass; magnitude cotangents on the bands within 30 mag of each model's brightest."""
    from helpers import operator
    nz = 3
    g = grid(nz)
    op = operator(nz, 40, 300)
    fs = synth.gaussian_filters(20)
    n = 32
    prm, names = _imf_models(n, nz, 57)
    print(f"case: api imf grad n={n} {precision}", flush=True)
    rng = np.random.default_rng(59)
    with _ctx(g, fs, precision) as ctx:
        ctx.set_ssp_ops(op)
        bands = ctx.bands()
        fwd = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table(), ops=op, imf_type=2, remnants=1)
        mb = np.where(fwd["mags"] <= np.nanmin(fwd["mags"], axis=1, keepdims=True) + 30.0,
                      rng.standard_normal((n, 20)), 0.0)
        r = ctx.compute_csp_batch_grad({k: prm[SLOT[k]] for k in names}, n=n, mags_bar=mb, grads=("params_bar",))
    ref = O.csp_batch_grad(g, bands, prm, mags_bar=mb, cosmo=O.cosmology_table(), ops=op, imf_type=2, remnants=1)
    assert ref["status"] == 0
    tol = 1e-8 if precision == "double" else 1e-4
    for name in ("imf1", "imf2", "imf3", "logzsol", "dust2", "tage", "tau", "zred"):
        row = O.GRAD_ROWS[name]
        e = float(np.max(np.abs(r["params_bar"][:, row] - ref["params_bar"][:, row]))) / (
            np.max(np.abs(ref["params_bar"][:, row])) + 1e-300)
        print(f"  {name:8s} {e:.2e}")
        assert e < tol, name


@pytest.mark.gpu
def test_api_zred_mags():
    """compute_csp_batch_grad with zred_mags through the public C API (redshifted colours): the magnitudes at the
    second redshift with the IGM at zred, and zred's row (the IGM's part), against the oracle."""
    g = grid(22)
    fs = synth.gaussian_filters(12)
    n = 24
    prm = synth.parametric_params(n, seed=131)
    zm = np.random.default_rng(133).uniform(0.0, 3.0, n)
    names = synth.varied_names()
    print("case: api zred_mags", flush=True)
    with _ctx(g, fs, "double") as ctx:
        ctx.set_int("redshift_colors", 1)
        mb = np.ones((n, 12))
        r = ctx.compute_csp_batch_grad({k: prm[SLOT[k]] for k in names}, n=n, mags_bar=mb, grads=("params_bar",),
                                       zred_mags=zm)
        bands = ctx.bands()
    ref = O.csp_batch_grad(g, bands, prm, mags_bar=mb, cosmo=O.cosmology_table(), zred_mags=zm)
    assert ref["status"] == 0
    em = mag_err(r["mags"], ref["mags"])
    row = O.GRAD_ROWS["zred"]
    ez = float(np.max(np.abs(r["params_bar"][:, row] - ref["params_bar"][:, row]))) / (
        np.max(np.abs(ref["params_bar"][:, row])) + 1e-300)
    print(f"api zred_mags: mags {em:.2e}, zred row {ez:.2e}")
    assert em < 1e-7 and ez < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("imf_type", [2, 0])
def test_api_evaluate_ssp_op(imf_type):
    """opensps_context_evaluate_ssp_op: the SSP of every age for one IMF (ssp_operator_evaluate), at the context's
    metallicity zmet, against the oracle's operator path with a unit weight at each age; imf_params NULL takes the
    context's slopes."""
    from helpers import operator
    nz = 3
    g = grid(nz)
    op = operator(nz, 40, 300)
    fs = synth.gaussian_filters(8)
    slopes = (0.9, 1.8, 2.7)
    print(f"case: api evaluate_ssp_op imf_type={imf_type}", flush=True)
    with _ctx(g, fs, "mixed") as ctx:     # evaluate runs in FP64 whatever the batch precision
        ctx.set_int("imf_type", imf_type)
        ctx.set_ssp_ops(op)
        ctx.set_int("zmet", 2)
        mass, lbol, spec = ctx.evaluate_ssp_op(imf_type, (*slopes, 0.08, 0.5))
        for k, v in zip(("imf1", "imf2", "imf3"), slopes):
            ctx.set_float(k, v)          # (drops the operator: given again)
        ctx.set_ssp_ops(op)
        mass2, lbol2, spec2 = ctx.evaluate_ssp_op(-1, None)
        bands = ctx.bands()
    nt = g.nt
    prm = synth.default_params(nt)
    prm[SLOT["logzsol"]] = g.logz[1]
    for k, v in zip(("imf1", "imf2", "imf3"), slopes):
        prm[SLOT[k]] = v
    ref = O.csp_batch(g, bands, prm, ssp_weights=np.eye(nt), add_igm=1, cosmo=O.cosmology_table(), ops=op,
                      imf_type=imf_type, remnants=1)
    es = spec_rel_err(spec, ref["spec"])
    print(f"evaluate_ssp_op: spectra {es:.2e}")
    assert es < 1e-9
    np.testing.assert_allclose(mass, ref["mass"], rtol=1e-11)
    np.testing.assert_allclose(lbol, ref["lbol"], rtol=0, atol=1e-11)
    np.testing.assert_array_equal(mass2, mass)
    np.testing.assert_array_equal(spec2, spec)


@pytest.mark.gpu
def test_api_batch_ext_equals_grad_rows():
    """The header-only struct entry point (opensps_context_compute_csp_batch_ext, compiled from C by the tests'
    shim) gives what opensps_context_compute_csp_batch_grad_rows gives for the same call: forward outputs and
    params_bar, weights_z_bar, bit for bit."""
    import ctypes as C
    from libopensps_amd import api, build
    g = grid(4)
    fs = synth.gaussian_filters(12)
    n = 16
    prm = synth.parametric_params(n, seed=61)
    names = ["dust2", "dust_index", "zred", "sigma_smooth"]
    wz = np.random.Generator(np.random.PCG64(62)).dirichlet(np.full(g.nt * g.nz, 0.5), n).reshape(n, g.nz, g.nt)
    mb = np.random.Generator(np.random.PCG64(63)).normal(size=(n, 12))
    print("case: api batch_ext vs grad_rows", flush=True)
    with _ctx(g, fs, "double") as ctx:
        vals = {k: prm[SLOT[k]] for k in names}
        ref = ctx.compute_csp_batch_grad(vals, n=n, ssp_weights_z=wz, mags_bar=mb, want=("mags", "spec", "mass"),
                                         grads=("params_bar", "weights_z_bar"), params_bar_rows=api.NBGRAD)
        keep = [np.ascontiguousarray(vals[k]) for k in names]
        cn = (C.c_char_p * len(names))(*[k.encode() for k in names])
        cv = (C.c_void_p * len(names))(*[a.ctypes.data for a in keep])
        wzc, mbc = np.ascontiguousarray(wz), np.ascontiguousarray(mb)
        mags, spec, mass = np.zeros((n, 12)), np.zeros((n, g.nspec)), np.zeros(n)
        pb, wzb = np.zeros((n, api.NBGRAD)), np.zeros((n, g.nz, g.nt))
        st = C.c_int()
        shim = C.CDLL(build.SHIM_SO)
        shim.shim_batch_ext(ctx.handle, n, len(names), C.cast(cn, C.c_void_p), C.cast(cv, C.c_void_p), None,
                            C.c_void_p(wzc.ctypes.data), C.c_void_p(mbc.ctypes.data), None, None, None,
                            C.c_void_p(mags.ctypes.data), C.c_void_p(spec.ctypes.data), C.c_void_p(mass.ctypes.data),
                            None, C.c_void_p(pb.ctypes.data), 0, None, C.c_void_p(wzb.ctypes.data), 0, C.byref(st))
        assert st.value == 0, api.last_error()
    np.testing.assert_array_equal(mags, ref["mags"])
    np.testing.assert_array_equal(spec, ref["spec"])
    np.testing.assert_array_equal(mass, ref["mass"])
    np.testing.assert_array_equal(pb, ref["params_bar"])
    np.testing.assert_array_equal(wzb, ref["weights_z_bar"])
    assert np.any(pb != 0.0)
