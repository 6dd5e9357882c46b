"""The reference's dump files of the batch, for comparing builds bit for bit (its tools/compare_forward_dump.py and
tools/compare_builds/compare_grad_dump.py read them):

* forward (benchmark_csp_batch with OPENSPS_BENCH_DUMP_FWD, tests/benchmark_csp_batch.f90:426-439): the int32
  header nbands, n, nspec, then the magnitudes (nbands x n) and the spectra (nspec x n) as FP64, each column-major
  (band or pixel fastest), little-endian, no padding;
* gradients (OPENSPS_BENCH_DUMP, :441-452): the int32 header NBGRAD, n, then params_bar (NBGRAD x n) as FP64.

The arrays here are in this package's layout, (n, nbands), (n, nspec) and (n, NBGRAD) C-order, which is the
column-major (quantity, model) layout of the files."""
from __future__ import annotations

import struct

import numpy as np


class DumpError(ValueError):
    pass


def write_forward(path: str, mags, spec) -> None:
    m = np.ascontiguousarray(mags, dtype="<f8")
    s = np.ascontiguousarray(spec, dtype="<f8")
    if m.ndim != 2 or s.ndim != 2 or m.shape[0] != s.shape[0]:
        raise DumpError(f"mags {m.shape} and spec {s.shape} are not (n, nbands) and (n, nspec)")
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", m.shape[1], m.shape[0], s.shape[1]))
        f.write(m.tobytes())
        f.write(s.tobytes())


def read_forward(path: str):
    """(mags (n, nbands), spec (n, nspec)) of a forward dump."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 12:
        raise DumpError(f"{path}: shorter than its header")
    nb, n, ns = struct.unpack("<3i", data[:12])
    if min(nb, n, ns) < 0 or len(data) != 12 + 8 * (nb * n + ns * n):
        raise DumpError(f"{path}: size does not match its header ({nb} x {n}, {ns})")
    body = np.frombuffer(data, dtype="<f8", offset=12)
    return body[:nb * n].reshape(n, nb).copy(), body[nb * n:].reshape(n, ns).copy()


def write_gradients(path: str, params_bar) -> None:
    p = np.ascontiguousarray(params_bar, dtype="<f8")
    if p.ndim != 2:
        raise DumpError(f"params_bar {p.shape} is not (n, NBGRAD)")
    with open(path, "wb") as f:
        f.write(struct.pack("<2i", p.shape[1], p.shape[0]))
        f.write(p.tobytes())


def read_gradients(path: str):
    """params_bar (n, NBGRAD) of a gradient dump."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 8:
        raise DumpError(f"{path}: shorter than its header")
    nr, n = struct.unpack("<2i", data[:8])
    if min(nr, n) < 0 or len(data) != 8 + 8 * nr * n:
        raise DumpError(f"{path}: size does not match its header ({nr} x {n})")
    return np.frombuffer(data, dtype="<f8", offset=8).reshape(n, nr).copy()
