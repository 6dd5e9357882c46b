"""opensps_hip_csp_batch_noalloc (the batch with its scratch in the fixed per-device array, no device allocation)
against opensps_hip_csp_batch on the same inputs: the same kernels, so identical outputs, on two calls in a row with
the second smaller, in both precisions; a batch whose scratch exceeds the fixed array is status 3 before any
launch."""
import numpy as np
import pytest

from helpers import bands, cosmo, grid
from libopensps_amd import synth
from libopensps_amd.params import SLOT


def _ctx(g, b):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    return ctx


def _models(n, seed, screens):
    prm = synth.parametric_params(n, seed=seed)
    if screens:
        prm[SLOT["dust1"]] = np.linspace(0.0, 1.0, n)
    return prm


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_batch_equals_batch(precision):
    import torch
    from libopensps_amd import hip as H
    g = grid(22)
    b = bands(100)
    split = 1 if precision == 0 else 0
    want = ("mags", "spec", "mass", "lbol", "idx")
    ctx = _ctx(g, b)
    try:
        for n, seed in ((480, 71), (131, 72)):
            prm = torch.as_tensor(_models(n, seed, split == 1), device="cuda")
            print(f"case: noalloc n={n} precision={precision} split={split}", flush=True)
            ref = ctx.batch(prm, precision=precision, split_parts=split, want=want)
            got = ctx.batch(prm, precision=precision, split_parts=split, want=want, noalloc=True)
            for k in want:
                a, r = got[k].cpu().numpy(), ref[k].cpu().numpy()
                assert np.array_equal(np.isnan(a), np.isnan(r)), k
                np.testing.assert_array_equal(a[~np.isnan(a)], r[~np.isnan(r)], err_msg=k)
        # beyond the fixed scratch (FP64, young and old apart: ~100 KB a model): status 3, nothing launched
        n = 1000
        prm = torch.as_tensor(_models(n, 73, True), device="cuda")
        with pytest.raises(H.OpenSPSHipError) as e:
            ctx.batch(prm, precision=0, split_parts=1, want=want, noalloc=True)
        assert e.value.status == 3
    finally:
        ctx.close()
