"""Nebular emission with the CLOUDY grids on the GPU (k_neb_q0, k_neb; the library placed on the grid by
opensps_hip_set_nebular) against the oracle (pinned in test_oracle_nebular.py): young models over the metallicity
blend with gas log Z and log U across (and beyond) the tables, escaping fractions, both dust screens, redshift and IGM,
the lines in the spectrum at one width; the lines outside the spectrum; the lines' attenuated luminosities (emlines);
and the statuses of what is not built."""
import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, spec_rel_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

pytestmark = pytest.mark.gpu

# FP64: the SFH weights' conditioning (1e-8 of a model's largest weight, as the basis path) sets all three; mixed: the
# FP32 age sums and chain, the lines' luminosities in FP64 with the transmissions rounded to FP32 (as the reference's
# line_attenuated rounds them)
TOL = {0: dict(spec=1e-7, mag=1e-7, lines=1e-7), 1: dict(spec=1e-4, mag=1e-3, lines=1e-5)}


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _params(n, seed=71):
    rng = np.random.default_rng(seed)
    prm = synth.parametric_params(n, seed=seed, logzsol=(-2.1, 0.45), zred=(0.0, 2.0), sigma=(120.0, 120.0))
    prm[SLOT["tage"]] = np.geomspace(0.004, 3.0, n)
    prm[SLOT["tau"]] = rng.uniform(0.1, 3.0, n)
    prm[SLOT["gas_logz"]] = rng.uniform(-2.3, 0.4, n)
    prm[SLOT["gas_logu"]] = rng.uniform(-4.3, -0.8, n)
    prm[SLOT["frac_obrun"]] = rng.uniform(0.0, 0.4, n)
    prm[SLOT["dust1"]] = rng.uniform(0.0, 1.0, n)
    prm[SLOT["dust2"]] = rng.uniform(0.0, 1.0, n)
    return prm


def _ctx(g, b, lib, cont_on=1, lines_in_spec=1):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    ctx.set_nebular(lib)
    ctx.set_neb_options(1, cont_on, lines_in_spec)
    return ctx


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("cont_on,lines_in_spec", [(1, 1), (0, 0)])
def test_nebular_against_oracle(precision, cont_on, lines_in_spec):
    g, b = grid(3), bands(20)
    lib = synth.nebular_library(g)
    n = 48
    prm = _params(n)
    neb = O.nebular_on_grid(g.lam, *lib, cont_on=cont_on, lines_in_spec=lines_in_spec)
    ref = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=("mags", "spec", "mass", "lbol", "emlines"),
                      nebular=neb)
    assert ref["status"] == 0
    plain = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=("mags",))
    ctx = _ctx(g, b, lib, cont_on, lines_in_spec)
    print(f"case: nebular n={n} precision={precision} cont={cont_on} lines_in_spec={lines_in_spec}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, split_parts=1, want=("mags", "spec", "mass", "lbol", "emlines"))
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    t = TOL[precision]
    es, em = spec_rel_err(got["spec"], ref["spec"]), mag_err(got["mags"], ref["mags"])
    big = ref["emlines"] > 1e-6 * ref["emlines"].max(axis=1, keepdims=True)
    el = float(np.max(np.abs(got["emlines"] / np.where(big, ref["emlines"], 1.0) - 1.0) * big))
    shift = float(np.nanmax(np.abs(ref["mags"] - plain["mags"])))
    print(f"  spectra {es:.3e}, mags {em:.3e}, emlines {el:.3e} (the nebular light moves mags by up to {shift:.3f})")
    assert shift > 0.01   # the stage acts
    assert es < t["spec"] and em < t["mag"] and el < t["lines"]
    np.testing.assert_allclose(got["mass"], ref["mass"], rtol=1e-12)


def test_nebular_statuses():
    from libopensps_amd import hip as H
    g, b = grid(3), bands(20)
    lib = synth.nebular_library(g)
    prm = _params(8)
    ctx = _ctx(g, b, lib)
    try:
        bad = prm.copy()
        bad[SLOT["sigma_smooth"], 3] = 80.0   # lines in the spectrum: one width a batch
        print("case: nebular statuses", flush=True)
        with pytest.raises(H.OpenSPSHipError) as e:
            ctx.batch(_dev(bad), precision=1, split_parts=1, want=("mags",))
        assert e.value.status == 2
        ctx.set_neb_options(1, 1, 0)
        ctx.batch(_dev(bad), precision=1, split_parts=1, want=("mags",))   # lines outside the spectrum: any width
        with pytest.raises(H.OpenSPSHipError) as e:
            ctx.batch_grad(_dev(prm), {"mags": _dev(np.ones((8, 20)))}, precision=1, split_parts=1,
                           grads=("params_bar",))
        assert e.value.status == 2
        ctx.set_emission(1, 0)
        with pytest.raises(H.OpenSPSHipError) as e:   # lines outside the spectrum with dust emission
            ctx.batch(_dev(prm), precision=1, split_parts=1, want=("mags",))
        assert e.value.status == 2
        ctx.set_emission(0, 0)
        ctx.set_nebular(None)
        with pytest.raises(H.OpenSPSHipError) as e:   # no tables
            ctx.batch(_dev(prm), precision=1, split_parts=1, want=("mags",))
        assert e.value.status == 2
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("library", ["release", "synthetic"])
def test_api_nebular_library(precision, library):
    """Through the public C API with a CLOUDY library given by opensps_context_set_nebular_tables: the host hands it
    to the device, which places it on the grid; the batch adds the continuum and the lines and returns emlines
    (get_nemline rows); against the oracle with the same library placed by the oracle's restatement of the
    reference's loader. "release": the test release's libraries/nebular as datalib reads it (its line table holds
    only FP32 round-off, so it exercises the reader, the placement and the continuum); "synthetic": a library with
    lines of the CLOUDY grid's strengths on the same path."""
    import os
    from libopensps_amd import api, datalib
    g = grid(3)
    if library == "release":
        lib = datalib.nebular(os.path.join(os.path.dirname(__file__), "golden", "opensps-test-data"))
    else:
        lib = synth.nebular_library(g)
    fs = synth.gaussian_filters(20)
    n = 40
    prm = _params(n, seed=83)
    names = synth.varied_names() + ["dust1", "frac_obrun", "gas_logz", "gas_logu"]
    print(f"case: api nebular library {library} {precision}", flush=True)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(fs)
        ctx.set_int("add_igm_absorption", 1)
        ctx.set_int("add_dust_emission", 0)
        ctx.set_int("add_neb_emission", 1)
        ctx.set_str("kernel_precision", precision)
        ctx.set_nebular_tables(*lib)
        assert ctx.nemline == lib[5].size
        r = ctx.compute_csp_batch_params(prm, names, want=("mags", "spec", "emlines"))
        b = ctx.bands()
    ref = O.csp_batch(g, b, prm, add_igm=1, cosmo=O.cosmology_table(), want=("mags", "spec", "emlines"),
                      nebular=O.nebular_on_grid(g.lam, *lib))
    plain = O.csp_batch(g, b, prm, add_igm=1, cosmo=O.cosmology_table(), want=("mags",))
    assert ref["status"] == 0
    t = TOL[0 if precision == "double" else 1]
    es, em = spec_rel_err(r["spec"], ref["spec"]), mag_err(r["mags"], ref["mags"])
    big = ref["emlines"] > 1e-6 * ref["emlines"].max(axis=1, keepdims=True)
    el = float(np.max(np.abs(r["emlines"] / np.where(big, ref["emlines"], 1.0) - 1.0) * big))
    shift = float(np.nanmax(np.abs(ref["mags"] - plain["mags"])))
    print(f"  spectra {es:.3e}, mags {em:.3e}, emlines {el:.3e} (the nebular light moves mags by up to {shift:.3f})")
    assert r["emlines"].shape == (n, lib[5].size) and shift > 0.01
    if library == "synthetic":
        assert big.sum() > n * 50
    assert es < t["spec"] and em < t["mag"] and el < t["lines"]
