"""The SSP operator with the IMF per model on the GPU (libopensps_hip directly) against the oracle: per-model Kroupa
slopes and upper limits, the metallicity blend over two passes, parametric SFHs, dust, IGM, redshift and smoothing.
The isochrone points each IMF selects (the counts per pass) are bit-exact; spectra, magnitudes, masses and L_bol
within the FP64 and mixed tolerances of the basis path."""
import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, operator, spec_rel_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

# FP64: the SFH weights' conditioning (1e-8 of a model's largest weight, as the basis path); mixed: FP32 point
# weights and fold, split-BF16 template product, FP32 chain
FP64_SPEC, FP64_MAG, FP64_SUMS = 1.0e-7, 1.0e-7, 1.0e-7
MIX_SPEC, MIX_MAG, MIX_SUMS = 1.0e-4, 1.0e-3, 1.0e-12


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _run(op, g, b, prm, precision, imf_type=2, remnants=1, weights=None, split=0,
         want=("mags", "spec", "mass", "lbol", "idx", "op_count")):
    from libopensps_amd import hip as H
    ref = O.csp_batch(g, b, prm, ssp_weights=weights, add_igm=1, cosmo=cosmo(), want=want, ops=op, imf_type=imf_type,
                      remnants=remnants)
    assert ref["status"] == 0
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    ctx.set_ssp_ops(op)
    ctx.set_imf(imf_type, remnants)
    print(f"case: operator n={prm.shape[1]} nz={op.nz} ntpl={op.ntpl} precision={precision} imf={imf_type}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, split_parts=split, use_ops=1,
                    ssp_weights=None if weights is None else _dev(weights), want=want)
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    return ref, got


def _check(ref, got, precision):
    fp64 = precision == 0
    assert np.array_equal(got["idx"][:, :2], ref["idx"][:, :2])
    assert np.array_equal(got["op_count"], ref["op_count"]), "the IMF limits select different isochrone points"
    es = spec_rel_err(got["spec"], ref["spec"])
    em = mag_err(got["mags"], ref["mags"])
    emass = float(np.max(np.abs(got["mass"] / ref["mass"] - 1.0)))
    elbol = float(np.max(np.abs(got["lbol"] - ref["lbol"])))
    print(f"operator: spectra {es:.3e}, mags {em:.3e}, mass {emass:.3e}, lbol {elbol:.3e}")
    assert es < (FP64_SPEC if fp64 else MIX_SPEC)
    assert em < (FP64_MAG if fp64 else MIX_MAG)
    assert emass < FP64_SUMS and elbol < FP64_SUMS


def _imf_case(n, nz=3, seed=41):
    op = operator(nz, 40, 300)
    g = grid(nz)
    b = bands(20)
    prm = synth.parametric_params(n, seed=seed, logzsol=(-2.1, 0.45) if nz > 1 else (0.0, 0.0))
    prm = synth.imf_params(n, seed=seed + 1, p=prm)
    return op, g, b, prm


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_operator_parametric(precision):
    op, g, b, prm = _imf_case(96)
    ref, got = _run(op, g, b, prm, precision)
    _check(ref, got, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_operator_one_metallicity_salpeter_no_remnants(precision):
    op, g, b, prm = _imf_case(40, nz=1, seed=43)
    prm[SLOT["imf_lower_limit"]] = np.linspace(0.08, 0.5, prm.shape[1])
    ref, got = _run(op, g, b, prm, precision, imf_type=0, remnants=0)
    _check(ref, got, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_operator_weights_and_parts(precision):
    """Given SSP weights, a birth-cloud screen (young and old apart) and the limits across the breaks."""
    op, g, b, prm = _imf_case(48, seed=47)
    prm[SLOT["dust1"]] = 0.7
    prm[SLOT["imf_lower_limit"]] = np.linspace(0.1, 1.2, prm.shape[1])
    prm[SLOT["imf_upper_limit"]] = np.linspace(5.0, 100.0, prm.shape[1])
    w = synth.dirichlet_weights(prm.shape[1], g.nt, seed=49)
    ref, got = _run(op, g, b, prm, precision, weights=w, split=1)
    _check(ref, got, precision)
