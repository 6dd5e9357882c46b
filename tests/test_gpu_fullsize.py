"""The IMF-per-model and batch-adjoint workloads at their full dimensions against the oracle.

The whole batch runs on the GPU, so that every large-index path is exercised: the operator's CSR segments over
160 500 points per metallicity and its 10 000-template product, 64 column tiles of 64 models per pass; the reverse
pass's transposed product over all 2354 (age, Z) columns and the 8192-model tiles. Models are independent, so the
oracle checks a bounded set of them (the first few of each batch), at the tolerances of the smaller parity tests:
the points each IMF selects bit-exact; spectra, magnitudes, masses, every params_bar row and weights_z_bar within
the FP64 and mixed bounds.
"""
import functools

import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, spec_rel_err
from libopensps_amd import synth

import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


@functools.lru_cache(maxsize=None)
def _operator():
    g = grid(22)
    return synth.ssp_operator(g.lam, g.log_age, 22, npts=1500, ntpl=10000, seed=11)


def _ctx(g, b):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    return ctx


@pytest.mark.parametrize("precision", [0, 1])
def test_operator_full_size(precision):
    """4096 models with per-model Kroupa slopes U(0.3, 3.3) and upper limits U(60, 120) on 107 ages x 22 Z x 1500
    points and 10 000 templates (the IMF-per-model workload); the first 16 against the oracle."""
    g, b, op = grid(22), bands(100), _operator()
    n, m = 4096, 16
    prm = synth.imf_params(n, seed=13, p=synth.parametric_params(n, seed=5))
    want = ("mags", "spec", "mass", "lbol", "idx", "op_count")
    ref = O.csp_batch(g, b, np.ascontiguousarray(prm[:, :m]), add_igm=1, cosmo=cosmo(), want=want, ops=op,
                      imf_type=2, remnants=1)
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    ctx.set_ssp_ops(op)
    ctx.set_imf(2, 1)
    print(f"case: operator full size n={n} precision={precision}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, use_ops=1, want=want)
    got = {k: v.detach().cpu().numpy()[:m] for k, v in out.items()}
    ctx.close()
    assert np.array_equal(got["idx"][:, :2], ref["idx"][:, :2])
    assert np.array_equal(got["op_count"], ref["op_count"]), "the IMF limits select different isochrone points"
    es, em = spec_rel_err(got["spec"], ref["spec"]), mag_err(got["mags"], ref["mags"])
    emass = float(np.max(np.abs(got["mass"] / ref["mass"] - 1.0)))
    elbol = float(np.max(np.abs(got["lbol"] - ref["lbol"])))
    print(f"operator full size: spectra {es:.2e}, mags {em:.2e}, mass {emass:.2e}, lbol {elbol:.2e}")
    assert es < (1e-7 if precision == 0 else 1e-4)
    assert em < (1e-7 if precision == 0 else 1e-3)
    assert emass < 1e-7 and elbol < 1e-7


ROWS = ["dust2", "dust_index", "zred", "sigma_smooth", "igm_factor", "dust1", "dust3", "frac_obrun", "frac_nodust"]


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_full_size(precision):
    """8192 models with weights over 107 ages x 22 Z (a log-normal SFH of 1 dex times an enrichment history), dust,
    IGM, z U(0, 4), smoothing, the cotangents of 100 magnitudes (the batch-adjoint workload); the first 8 against the
    oracle's reverse pass: every params_bar row this path computes, the rows it leaves 0, weights_z_bar."""
    g, b = grid(22), bands(100)
    n, m = 8192, 8
    prm = synth.adjoint_params(n, seed=19)
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=17)
    mb = np.random.default_rng(23).standard_normal((n, 100))
    ref = O.csp_batch_grad(g, b, np.ascontiguousarray(prm[:, :m]), ssp_weights_z=wz[:m], mags_bar=mb[:m],
                           cosmo=cosmo())
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: adjoint full size n={n} precision={precision}", flush=True)
    out = ctx.batch_grad(_dev(prm), {"mags": _dev(mb)}, precision=precision, ssp_weights_z=_dev(wz),
                         want=("mags",), grads=("params_bar", "weights_z_bar"))
    got = {k: v.detach().cpu().numpy()[:m] for k, v in out.items()}
    ctx.close()
    tol = {0: (1e-10, 1e-11, 1e-10), 1: (2e-4, 1e-4, 1e-4)}[precision]
    em = mag_err(got["mags"], ref["mags"])
    assert em < tol[2], em
    pr, pg = ref["params_bar"], got["params_bar"]
    for name in ROWS:
        r = O.GRAD_ROWS[name]
        e = float(np.max(np.abs(pg[:, r] - pr[:, r]))) / (np.max(np.abs(pr[:, r])) + 1e-300)
        print(f"  {name:12s} {e:.2e}")
        assert e < tol[0], name
    other = [r for r in range(O.NBGRAD) if r not in [O.GRAD_ROWS[k] for k in ROWS]]
    assert np.all(pg[:, other] == 0.0)
    wr, wg = ref["weights_z_bar"].reshape(m, -1), got["weights_z_bar"].reshape(m, -1)
    ew = float(np.max(np.abs(wg - wr) / (np.max(np.abs(wr), axis=1, keepdims=True) + 1e-300)))
    print(f"  weights_z_bar {ew:.2e}")
    assert ew < tol[1]
