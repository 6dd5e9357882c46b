"""The emulator trained on spectra the GPU batch generates, with the batch as its judge and fallback: a short
training run (200k samples) reaches a median magnitude error well below that of the untrained network and of the
best constant predictor, infers faster than the batch on the same inputs, and the serving path sends inputs outside
the training box, and whole batches whose spot-check fails, to the batch."""
import time

import numpy as np
import pytest

from helpers import cosmo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    import torch
    from libopensps_amd import emu_train as T, hip as H, synth
    import oracle as O
    g = synth.ssp_grid(seed=1, nspec=5994, nt=107, nz=22)
    fs = synth.gaussian_filters(20, seed=7)
    dctx = H.DeviceContext(0)
    dctx.set_grid(g)
    dctx.set_bands(O.build_bands(fs, g.lam))
    cz, cdl = cosmo()
    dctx.set_cosmology(cz, cdl)
    dctx.set_options(1, 1)
    print("case: emulator training 200k samples", flush=True)
    emu = T.Emulator.train(dctx, (cz, cdl), n_train=200_000, n_pca_fit=20_000, n_pca=32, epochs=20, batch=2048,
                           lr=2e-3, hidden=256, layers=3, log=print)
    yield T, H, dctx, emu
    emu.mlp.close()
    dctx.close()


def test_training_beats_the_untrained_network(trained):
    T, H, dctx, emu = trained
    from libopensps_amd.emulator import MLP
    v = T.validate(emu, dctx, n=4096)
    print(f"  trained: {v}")
    # the network before training (same architecture and scalings)
    untrained = MLP(T.N_FEAT, emu.n_pca + emu.nbands, 256, 3, seed=5)
    p = emu.mlp.get_params()
    untrained.set_scaling(*emu.mlp.scale)
    emu0 = T.Emulator(dctx, untrained, emu.pca, emu.n_pca, emu.nbands, emu.box, emu.cosmo)
    v0 = T.validate(emu0, dctx, n=4096)
    print(f"  untrained: {v0}")
    untrained.close()
    assert p.size > 0
    assert v["median_dmag"] < 0.08 and v["median_dmag"] < 0.2 * v0["median_dmag"]
    assert emu.train_info["loss_history"][-1] < 0.5 * emu.train_info["loss_history"][0]


def test_inference_faster_than_the_batch(trained):
    import torch
    T, H, dctx, emu = trained
    n = 65536
    x = T.sample_inputs(n, 5)
    fd, _ = emu.inputs(x)
    emu.predict_raw(fd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        emu.predict_raw(fd)
    te = (time.perf_counter() - t0) / 5
    prm = T.to_prm_device(torch.as_tensor(x, device="cuda"))
    dctx.batch(prm, precision=H.MIXED, want=("mags",))
    t0 = time.perf_counter()
    for _ in range(2):
        dctx.batch(prm, precision=H.MIXED, want=("mags",))
    tb = (time.perf_counter() - t0) / 2
    print(f"  emulator {n / te:.3e} models/s, batch {n / tb:.3e} models/s, x{tb / te:.1f}")
    assert te * 5 < tb


def test_fallback_and_spot_check(trained):
    T, H, dctx, emu = trained
    x = T.sample_inputs(2000, 11)
    x[:50, T.INPUTS.index("zred")] = 5.0          # outside the box: the batch takes them
    mags, info = emu.predict(x, tol=1.0)
    assert info["fallback"] == 50 and not info["spot_failed"]
    ref = emu._batch(x[:50])
    assert float((mags[:50] - ref).abs().max()) == 0.0
    mags, info = emu.predict(x, tol=0.0)          # every spot-check fails: the whole batch falls back
    assert info["spot_failed"] and info["fallback"] == 2000
    assert 0.0 < emu.fallback_rate() <= 1.0


def test_full_input_space():
    """The emulator on the full input set (the parametric parameters, the three Kroupa slopes through the SSP
    operator, and the dust emission and AGN parameters with their templates on the context): a short training run
    learns the slopes' and the emission's effect (its median error well below the best constant predictor's), the
    batch it falls back to is the operator batch with the emission stages, and inputs outside the box go to it."""
    import torch
    from libopensps_amd import emu_train as T, hip as H, synth
    from helpers import bands, grid, operator
    sp = T.FULL
    g = grid(3)
    dctx = H.DeviceContext(0)
    dctx.set_grid(g)
    dctx.set_bands(bands(20))
    cz, cdl = cosmo()
    dctx.set_cosmology(cz, cdl)
    dctx.set_options(1, 1)
    dctx.set_ssp_ops(operator(3, 40, 300))
    dctx.set_imf(2, 1)
    dctx.set_emission(1, 1)
    dctx.set_dust_emission(synth.DL07_QPAH, synth.DL07_UMIN, synth.dust_emission_templates(g.lam))
    dctx.set_agn(synth.AGN_TAU, synth.agn_templates(g.lam))
    print("case: emulator training, full input set, 200k samples", flush=True)
    emu = T.Emulator.train(dctx, (cz, cdl), n_train=200_000, n_pca_fit=20_000, n_pca=32, epochs=20, batch=2048,
                           lr=2e-3, hidden=256, layers=3, space=sp)
    try:
        assert emu.train_info["inputs"] == sp.names and emu.mlp.dims[0] == sp.n_feat == 15
        v = T.validate(emu, dctx, n=4096)
        print(f"  trained: {v}")
        # the best constant predictor of the magnitudes less the distance term (the per-band median)
        x = sp.sample(4096, 999)
        o = dctx.batch(sp.to_prm_device(torch.as_tensor(x, device="cuda")), precision=H.FP64, want=("mags",),
                       use_ops=1)
        ref = o["mags"]
        dist = torch.as_tensor(T.distance_term(x[:, sp.iz], (cz, cdl)), device="cuda")[:, None]
        fin = torch.isfinite(ref)
        bright = torch.where(fin, ref, torch.full_like(ref, float("inf"))).amin(dim=1, keepdim=True)
        ok = fin & (ref - bright <= 20.0)
        r = ref - dist
        const = torch.nanmedian(torch.where(ok, r, torch.full_like(r, float("nan"))), dim=0).values
        d0 = (r - const[None, :]).abs()[ok]
        print(f"  constant predictor median {float(d0.median()):.4f}")
        assert v["median_dmag"] < 0.2 * float(d0.median())
        # the slopes and the emission inputs move the magnitudes the network predicts as the batch moves them
        xa = sp.sample(256, 5)
        xb = xa.copy()
        xb[:, sp.names.index("imf1")] = 0.5
        xb[:, sp.names.index("imf3")] = 3.0
        ma = emu._batch(xa)
        mb = emu._batch(xb)
        _, pa = emu.predict_raw(*emu.inputs(xa))
        _, pb = emu.predict_raw(*emu.inputs(xb))
        dm_ref = (mb - ma)[:, :5].double()
        dm_emu = (pb - pa)[:, :5].double()
        assert float(dm_ref.abs().median()) > 0.05
        assert float((dm_emu - dm_ref).abs().median()) < 0.5 * float(dm_ref.abs().median())
        # outside the box (imf1 beyond its range): the operator batch takes them
        xo = sp.sample(300, 12)
        xo[:20, sp.names.index("imf1")] = 3.5
        mags, info = emu.predict(xo, tol=1.0)
        assert info["fallback"] == 20
        assert float((mags[:20] - emu._batch(xo[:20])).abs().max()) == 0.0
    finally:
        emu.mlp.close()
        dctx.close()
