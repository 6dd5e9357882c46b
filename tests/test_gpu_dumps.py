"""A GPU batch written in the reference's cross-build formats (the forward dump of its batch benchmark and the OPSF
fixture container) and read back: the files hold exactly the device's outputs, and those match the oracle on the
same inputs, so a file from this build can be set against one from the reference's with its comparison tools."""
import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, spec_rel_err
from libopensps_amd import dumps, synth

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_dumps(tmp_path, precision):
    import torch
    from libopensps_amd import hip as H
    g, b = grid(3), bands(20)
    n = 64
    prm = synth.parametric_params(n, seed=61, logzsol=(-2.1, 0.45))
    ctx = H.DeviceContext(0)
    try:
        ctx.set_grid(g)
        ctx.set_bands(b)
        cz, cdl = cosmo()
        ctx.set_cosmology(cz, cdl)
        ctx.set_options(1, 1)
        print(f"case: dumps n={n} precision={precision}", flush=True)
        out = ctx.batch(torch.as_tensor(prm, device="cuda"), precision=precision,
                        want=("mags", "spec", "mass", "lbol"))
        host = {k: v.detach().cpu().numpy() for k, v in out.items()}
    finally:
        ctx.close()
    fwd = tmp_path / "fwd.bin"
    dumps.write_forward(str(fwd), host["mags"], host["spec"])
    m2, s2 = dumps.read_forward(str(fwd))
    assert m2.tobytes() == host["mags"].tobytes()
    assert np.array_equal(s2, host["spec"].astype(np.float64))
    ops = tmp_path / "case.opsf"
    from libopensps_amd.params import SLOT
    dumps.write_batch_opsf(str(ops), g.lam, host, tage=prm[SLOT["tage"]])
    r = dumps.read_opsf(str(ops))
    assert np.array_equal(r["wavelengths"], g.lam)
    assert np.array_equal(r["csp/spec"], host["spec"].astype(np.float64))
    assert r["csp/mags"].tobytes() == host["mags"].tobytes()
    assert np.array_equal(r["csp/mass"], host["mass"]) and np.array_equal(r["csp/lbol"], host["lbol"])
    # the files' contents against the oracle
    ref = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=("mags", "spec", "mass", "lbol"))
    es, em = spec_rel_err(r["csp/spec"], ref["spec"]), mag_err(r["csp/mags"], ref["mags"])
    print(f"  dumped spectra {es:.3e}, magnitudes {em:.3e}")
    assert es < (1e-7 if precision == 0 else 1e-4) and em < (1e-7 if precision == 0 else 1e-3)
