"""The emulator's HIP kernels (include/opensps_emu.h) against the FP64 oracle with the kernels' bf16 roundings
emulated (oracle/emu_oracle.py): inference, the loss and its gradient with respect to every parameter, one Adam
step, and a short training run that lowers the loss."""
import numpy as np
import pytest

import emu_oracle as E  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def _setup(dims, seed=0, n=300):
    from libopensps_amd import emulator as EM
    rng = np.random.default_rng(seed)
    n_in, n_out, hidden, layers = dims
    net = EM.MLP(n_in, n_out, hidden, layers, seed=seed + 1)
    p = net.get_params().astype(np.float64)
    p = p + 0.05 * rng.standard_normal(p.size) * (p != 0)   # biases move off zero too below
    sizes = [n_in] + [hidden] * layers + [n_out]
    at = 0
    for l in range(layers + 1):
        at += sizes[l + 1] * sizes[l]
        p[at:at + sizes[l + 1]] = 0.1 * rng.standard_normal(sizes[l + 1])
        at += sizes[l + 1]
    p = p.astype(np.float32).astype(np.float64)
    net.set_params(p)
    scale = (rng.standard_normal(n_in), 0.5 + rng.random(n_in), rng.standard_normal(n_out), 0.5 + rng.random(n_out))
    net.set_scaling(*scale)
    scale = tuple(np.asarray(v, np.float32).astype(np.float64) for v in scale)
    x = (rng.standard_normal((n, n_in)) * scale[1] + scale[0]).astype(np.float32)
    t = (rng.standard_normal((n, n_out)) * scale[3] + scale[2]).astype(np.float32)
    return net, p, scale, x.astype(np.float64), t.astype(np.float64)


DIMS = [(7, 228, 512, 4), (13, 100, 256, 2)]


@pytest.mark.parametrize("per_layer", [False, True])
@pytest.mark.parametrize("dims", DIMS + [(8, 64, 384, 3), (5, 200, 128, 1)])
def test_predict_matches_oracle(dims, per_layer, monkeypatch):
    """The fused inference kernel (activations in LDS) and the layer-by-layer one."""
    if per_layer:
        monkeypatch.setenv("OPENSPS_EMU_PER_LAYER", "1")
    net, p, scale, x, _ = _setup(dims)
    print(f"case: emu predict dims={dims} per_layer={per_layer}", flush=True)
    y = net.predict(_dev(x)).cpu().numpy()
    net.sync()
    ref, _ = E.forward(p, x, dims, scale, round_bf16=True)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f"  predict rel err {err:.2e}")
    assert err < 2e-3
    net.close()


@pytest.mark.parametrize("dims", DIMS)
def test_gradients_match_oracle(dims):
    net, p, scale, x, t = _setup(dims, seed=3)
    print(f"case: emu gradients dims={dims}", flush=True)
    w = np.linspace(0.5, 1.5, dims[1])
    net.set_loss_weights(w)
    loss, g = net.gradients(_dev(x), _dev(t))
    lref, gref = E.loss_and_grad(p, x, t, dims, scale, w, round_bf16=True)
    print(f"  loss {loss:.6e} vs {lref:.6e}")
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    # per layer, relative to the layer's largest gradient
    n_in, n_out, hidden, layers = dims
    sizes = [n_in] + [hidden] * layers + [n_out]
    at = 0
    for l in range(layers + 1):
        for part, m in (("W", sizes[l + 1] * sizes[l]), ("b", sizes[l + 1])):
            a, b = g[at:at + m], gref[at:at + m]
            err = np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-30)
            print(f"  layer {l} {part} {err:.2e}")
            assert err < 2e-3, (l, part)
            at += m
    net.close()


def test_adam_step_matches_numpy():
    dims = (7, 64, 128, 2)
    net, p, scale, x, t = _setup(dims, seed=5)
    print("case: emu adam step", flush=True)
    _, g = net.gradients(_dev(x), _dev(t))
    lr = 1e-3
    net.train_step(_dev(x), _dev(t), lr)
    net.sync()
    got = net.get_params().astype(np.float64)
    g = g.astype(np.float64)
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    err = np.max(np.abs(got - ref))
    print(f"  adam max abs err {err:.2e}")
    assert err < 1e-6


def test_training_lowers_the_loss():
    dims = (4, 16, 128, 2)
    from libopensps_amd import emulator as EM
    import torch
    rng = np.random.default_rng(7)
    net = EM.MLP(*dims, seed=3)
    net.set_scaling(np.zeros(4), np.ones(4), np.zeros(16), np.ones(16))
    x = rng.uniform(-1, 1, (4096, 4))
    A = rng.standard_normal((4, 16))
    t = np.sin(x @ A)
    xd, td = _dev(x), _dev(t)
    print("case: emu training", flush=True)
    l0 = net.train_step(xd, td, 1e-3, want_loss=True)
    for _ in range(300):
        net.train_step(xd, td, 1e-3)
    l1 = net.train_step(xd, td, 1e-3, want_loss=True)
    print(f"  loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.05 * l0
    net.close()
