"""The learned emulator: a ctypes front end of libopensps_hip's MLP (include/opensps_emu.h) and its harness:
training data from the GPU batch itself, the PCA basis of log-spectra, training with the project's own forward /
backward / Adam kernels, and inference with the plain batch as fallback and judge.

The network maps a model's parameters to [the PCA coefficients of its log10 spectrum, its band magnitudes]. The
plain batch (opensps_hip_csp_batch) stays the fallback for models outside the training box and for batches whose
spot-check (a random subset recomputed by the batch) exceeds the tolerance, and it is the judge of the emulator's
error on a validation batch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import build as _build
from . import hip as H

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_build.HIP_SO)
        vp, i, f, u = C.c_void_p, C.c_int, C.c_float, C.c_uint
        L.opensps_emu_create.argtypes = [i, i, i, i, i, u, C.POINTER(vp)]
        L.opensps_emu_destroy.argtypes = [vp]
        L.opensps_emu_nparams.argtypes = [vp, C.POINTER(C.c_long)]
        L.opensps_emu_set_params.argtypes = [vp, vp]
        L.opensps_emu_get_params.argtypes = [vp, vp]
        L.opensps_emu_set_scaling.argtypes = [vp, vp, vp, vp, vp]
        L.opensps_emu_set_loss_weights.argtypes = [vp, vp]
        L.opensps_emu_set_adam.argtypes = [vp, f, f, f]
        L.opensps_emu_predict.argtypes = [vp, i, vp, vp, vp]
        L.opensps_emu_gradients.argtypes = [vp, i, vp, vp, vp, vp]
        L.opensps_emu_train_step.argtypes = [vp, i, vp, vp, f, vp]
        L.opensps_emu_sync.argtypes = [vp]
        L.opensps_emu_set_stream.argtypes = [vp, vp]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """One emulator network on the GPU (n_in -> layers x hidden GELU -> n_out)."""

    def __init__(self, n_in, n_out, hidden=512, layers=4, seed=1, device=0):
        import torch
        self.dims = (n_in, n_out, hidden, layers)
        self._e = C.c_void_p()
        H.check(lib().opensps_emu_create(device, n_in, n_out, hidden, layers, seed, C.byref(self._e)))
        # the kernels run on torch's current stream: ordered with the torch work that writes their inputs, and the
        # caching allocator's reuse of those tensors is safe (it is stream-ordered)
        self._stream = torch.cuda.current_stream(device).cuda_stream
        H.check(lib().opensps_emu_set_stream(self._e, C.c_void_p(self._stream)))

    def close(self):
        if self._e:
            lib().opensps_emu_destroy(self._e)
            self._e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nparams(self) -> int:
        n = C.c_long()
        H.check(lib().opensps_emu_nparams(self._e, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.nparams(), np.float32)
        H.check(lib().opensps_emu_get_params(self._e, p.ctypes.data))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams()
        H.check(lib().opensps_emu_set_params(self._e, p.ctypes.data))

    def set_scaling(self, in_mu, in_sd, out_mu, out_sd):
        a = [_f32(v) for v in (in_mu, in_sd, out_mu, out_sd)]
        H.check(lib().opensps_emu_set_scaling(self._e, *[v.ctypes.data for v in a]))
        self.scale = tuple(np.asarray(v, np.float64) for v in a)

    def set_loss_weights(self, w):
        w = _f32(w)
        H.check(lib().opensps_emu_set_loss_weights(self._e, w.ctypes.data))

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8):
        H.check(lib().opensps_emu_set_adam(self._e, beta1, beta2, eps))

    def predict(self, x, out=None, stream=None, sync=True):
        """x: device float32 (n, n_in) tensor -> device (n, n_out), on torch's current stream (or stream); with
        sync the call waits for the result."""
        import torch
        n = x.shape[0]
        assert x.dtype == torch.float32 and x.is_contiguous() and x.shape[1] == self.dims[0]
        if out is None:
            out = torch.empty((n, self.dims[1]), dtype=torch.float32, device=x.device)
        H.check(lib().opensps_emu_predict(self._e, n, x.data_ptr(), out.data_ptr(), stream))
        if sync:
            self.sync()
        return out

    def gradients(self, x, t):
        n = x.shape[0]
        g = np.zeros(self.nparams(), np.float32)
        loss = C.c_float()
        H.check(lib().opensps_emu_gradients(self._e, n, x.data_ptr(), t.data_ptr(), g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def train_step(self, x, t, lr, want_loss=False):
        """One Adam step on device x, t (stream-ordered after the torch work that wrote them)."""
        loss = C.c_float()
        H.check(lib().opensps_emu_train_step(self._e, x.shape[0], x.data_ptr(), t.data_ptr(), float(lr),
                                             C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def sync(self):
        H.check(lib().opensps_emu_sync(self._e))


def _infer_main(argv=None):
    """The inference kernel alone on a randomly initialised network (the serving shape: n_in features -> layers x
    hidden -> n_out outputs), timed with the device synchronised around `reps` calls; for profiling runs."""
    import argparse
    import time

    import torch
    ap = argparse.ArgumentParser(description="emulator inference micro-benchmark")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--n-in", type=int, default=8)
    ap.add_argument("--n-out", type=int, default=228)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--layers", type=int, default=4)
    a = ap.parse_args(argv)
    mlp = MLP(a.n_in, a.n_out, a.hidden, a.layers, seed=3)
    x = torch.rand((a.n, a.n_in), dtype=torch.float32, device="cuda")
    out = mlp.predict(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        mlp.predict(x, out=out, sync=False)
    mlp.sync()
    dt = (time.perf_counter() - t0) / a.reps
    flops = 2.0 * (a.n_in * a.hidden + (a.layers - 1) * a.hidden * a.hidden + a.hidden * a.n_out) * a.n
    print(f"emulator inference: {a.n} models, {1e3 * dt:.3f} ms, {a.n / dt:.4g} models/s, {flops / dt / 1e12:.1f} TFLOP/s")
    mlp.close()


if __name__ == "__main__":
    _infer_main()
