"""Dust emission and the AGN torus on the GPU (k_emission and the chain's emission stages) against the oracle
(orc_csp_batch with an Emission, pinned by energy balance in tests/test_oracle_emission.py): spectra, magnitudes
and dust masses with the Draine & Li templates between their nodes, birth-cloud and old-star screens, the IGM,
redshifts and smoothing; the AGN torus at and between its optical depths; the IMF-per-model operator path (its L_bol
from the operator's sums); a model that needs templates the context lacks is status 2."""
import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, operator, spec_rel_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

pytestmark = pytest.mark.gpu

# FP64: the SFH weights' conditioning (1e-9) as the batch without emission; mixed: the FP32 chain
TOL = {0: (1.0e-7, 1.0e-7, 1.0e-9), 1: (1.0e-4, 1.0e-4, 2.0e-5)}


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _prm(n, seed, nz):
    prm = synth.parametric_params(n, seed=seed, logzsol=(-2.0, 0.3) if nz > 1 else (0.0, 0.0))
    rng = np.random.default_rng(seed)
    prm[SLOT["dust1"]] = np.where(rng.random(n) < 0.6, rng.uniform(0.1, 1.0, n), 0.0)
    prm[SLOT["dust3"]] = np.where(rng.random(n) < 0.3, rng.uniform(0.0, 0.5, n), 0.0)
    prm[SLOT["frac_obrun"]] = rng.uniform(0.0, 0.3, n)
    prm[SLOT["dust2"]][:2] = 0.0           # no dust: no emission
    prm[SLOT["dust1"]][:2] = 0.0
    prm[SLOT["duste_qpah"]] = rng.uniform(0.2, 5.0, n)       # beyond the grid's ends too (weights not clamped)
    prm[SLOT["duste_umin"]] = rng.uniform(0.1, 25.0, n)
    prm[SLOT["duste_gamma"]] = rng.uniform(-0.1, 1.1, n)      # (clamped)
    prm[SLOT["fagn"]] = np.where(rng.random(n) < 0.5, rng.uniform(0.0, 0.5, n), 0.0)
    prm[SLOT["agn_tau"]] = rng.uniform(3.0, 160.0, n)
    return prm


def _ctx(g, b, agn=True, duste=True):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    ctx.set_cosmology(*cosmo())
    ctx.set_options(1, 1)
    ctx.set_emission(1 if duste else 0, 1 if agn else 0)
    ctx.set_dust_emission(synth.DL07_QPAH, synth.DL07_UMIN, synth.dust_emission_templates(g.lam))
    ctx.set_agn(synth.AGN_TAU, synth.agn_templates(g.lam))
    return ctx


def _em(g, agn=True, duste=True):
    return O.Emission(synth.DL07_QPAH, synth.DL07_UMIN, synth.dust_emission_templates(g.lam), synth.AGN_TAU,
                      synth.agn_templates(g.lam), add_dust_emission=1 if duste else 0, add_agn_dust=1 if agn else 0)


def _check(ref, got, precision, what):
    ts, tm, tmd = TOL[precision]
    es = spec_rel_err(got["spec"], ref["spec"])
    if precision == 0:
        em = mag_err(got["mags"], ref["mags"])
    else:   # FP32 spectra: the bands within 30 mag of each model's brightest (fainter far-UV ones reach FP32's
        # subnormals or underflow to 0)
        q = ref["mags"]
        bright = np.nanmin(np.where(np.isfinite(q), q, np.nan), axis=1, keepdims=True)
        ok = np.isfinite(q) & (q - bright <= 30.0)
        assert np.all(np.isfinite(got["mags"][ok]))
        em = float(np.max(np.abs(got["mags"][ok] - q[ok])))
    ed = float(np.max(np.abs(got["mdust"] - ref["mdust"]) / np.abs(ref["mdust"])))
    print(f"{what}: spectra {es:.2e}, mags {em:.2e}, mdust {ed:.2e}")
    assert es < ts and em < tm and ed < tmd


@pytest.mark.parametrize("precision", [0, 1])
def test_emission_parametric(precision):
    nz = 3
    g, b = grid(nz), bands(30)
    n = 96
    prm = _prm(n, 201, nz)
    want = ("mags", "spec", "mdust")
    ref = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=want, emission=_em(g))
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: emission parametric n={n} precision={precision}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, split_parts=1, want=want)
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ctx.close()
    _check(ref, got, precision, "emission")
    # the emitting models' spectra differ from the same batch without emission
    assert np.all(ref["mdust"][2:][prm[SLOT["dust2"]][2:] > 0] > 1e-200)


@pytest.mark.parametrize("precision", [0, 1])
def test_emission_operator(precision):
    """The IMF per model with dust emission and the AGN (the AGN's L_bol from the operator's per-age sums)."""
    nz = 3
    g, b = grid(nz), bands(20)
    op = operator(nz, 40, 300)
    n = 48
    prm = synth.imf_params(n, seed=207, p=_prm(n, 205, nz))
    want = ("mags", "spec", "mdust", "lbol")
    ref = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=want, emission=_em(g), ops=op, imf_type=2,
                      remnants=1)
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    ctx.set_ssp_ops(op)
    ctx.set_imf(2, 1, 2.35, 0.5, 40.0, 8.5)
    print(f"case: emission operator n={n} precision={precision}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, split_parts=1, use_ops=1, want=want)
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ctx.close()
    _check(ref, got, precision, "emission operator")


def test_emission_statuses():
    """Templates absent: the models that need them are status 2, the others run; emission off: no status."""
    from libopensps_amd import hip as H
    g, b = grid(1), bands(10)
    n = 8
    prm = _prm(n, 211, 1)
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    ctx.set_cosmology(*cosmo())
    ctx.set_options(1, 1)
    ctx.set_emission(1, 1)
    print("case: emission statuses", flush=True)
    with pytest.raises(H.OpenSPSHipError) as e:
        ctx.batch(_dev(prm), precision=0, split_parts=1, want=("mags",))
    assert e.value.status == 2
    ctx.set_emission(0, 0)
    out = ctx.batch(_dev(prm), precision=0, split_parts=1, want=("mags", "mdust"))
    assert np.all(out["mdust"].cpu().numpy() == np.finfo(np.float64).tiny)
    ctx.close()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_api_emission_library(precision):
    """Through the public C API with the test release's libraries (dust_emission/DL07, agn_torus/clumpy, read by
    datalib and given by opensps_context_set_dust_emission_templates / _set_agn_templates): the host places them on
    the grid, the batch adds the emission; against the oracle with the same libraries placed by the oracle's
    restatement of the reference's loaders."""
    import os
    from libopensps_amd import api, datalib
    root = os.path.join(os.path.dirname(__file__), "golden", "opensps-test-data")
    dust, agn = datalib.dust_emission(root), datalib.agn_torus(root)
    nz = 3
    g = grid(nz)
    fs = synth.gaussian_filters(20)
    n = 64
    prm = _prm(n, 221, nz)
    names = synth.varied_names() + ["dust1", "dust3", "frac_obrun", "duste_qpah", "duste_umin", "duste_gamma",
                                    "fagn", "agn_tau"]
    print(f"case: api emission library {precision}", flush=True)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(fs)
        ctx.set_int("add_igm_absorption", 1)
        ctx.set_str("kernel_precision", precision)
        ctx.set_dust_emission_templates(*dust)
        ctx.set_agn_templates(*agn)
        r = ctx.compute_csp_batch_params(prm, names, want=("mags", "spec", "mdust"))
        bands = ctx.bands()
    ref = O.csp_batch(g, bands, prm, add_igm=1, cosmo=O.cosmology_table(), want=("mags", "spec", "mdust"),
                      emission=O.emission_on_grid(g.lam, dust, agn))
    assert ref["status"] == 0
    _check(ref, r, 0 if precision == "double" else 1, "api emission library")
