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
