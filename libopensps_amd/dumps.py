"""The reference's dump files of the batch, for comparing builds bit for bit (its tools/compare_forward_dump.py and
tools/compare_builds/compare_grad_dump.py read them):

* forward (benchmark_csp_batch with OPENSPS_BENCH_DUMP_FWD, tests/benchmark_csp_batch.f90:426-439): the int32
  header nbands, n, nspec, then the magnitudes (nbands x n) and the spectra (nspec x n) as FP64, each column-major
  (band or pixel fastest), little-endian, no padding;
* gradients (OPENSPS_BENCH_DUMP, :441-452): the int32 header NBGRAD, n, then params_bar (NBGRAD x n) as FP64.

* OPSF (the cross-build fixtures of tools/compare_builds, written by tools/legacy_parity/csp_dump.f90:76-110 and
  tests/generate_opensps_fixtures.f90, read by compare_legacy.read_opsf): the 4 bytes "OPSF", an int32 version (1),
  then records until the end of the file, each an int32 name length, the name (ASCII), an int32 rank, the int32
  Fortran dimensions, and the values as FP64 in Fortran (column-major) order. The batch's records are named as the
  reference names them: wavelengths, csp/spec (nspec x n), csp/mags (nbands x n), csp/mass, csp/lbol, csp/mdust,
  csp/age (n).

The arrays here are in this package's layout, (n, nbands), (n, nspec) and (n, NBGRAD) C-order, which is the
column-major (quantity, model) layout of the files; an OPSF record of Fortran dimensions (d1, .., dk) is the C-order
array of shape (dk, .., d1)."""
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


OPSF_MAGIC = b"OPSF"
OPSF_VERSION = 1


def write_opsf(path: str, records) -> None:
    """Writes {name: array} (C-order arrays: shape (dk, .., d1) for Fortran dimensions (d1, .., dk)) in order."""
    with open(path, "wb") as f:
        f.write(OPSF_MAGIC)
        f.write(struct.pack("<i", OPSF_VERSION))
        for name, a in records.items():
            b = name.encode("ascii")
            v = np.ascontiguousarray(np.asarray(a, dtype="<f8"))
            dims = tuple(reversed(v.shape)) if v.ndim else (1,)
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(dims)))
            f.write(struct.pack(f"<{len(dims)}i", *dims))
            f.write(v.tobytes())


def read_opsf(path: str):
    """{name: array} of an OPSF file, each array C-order of shape (dk, .., d1)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != OPSF_MAGIC or len(data) < 8:
        raise DumpError(f"{path}: not an OPSF file")
    (version,) = struct.unpack("<i", data[4:8])
    if version != OPSF_VERSION:
        raise DumpError(f"{path}: OPSF version {version}, not {OPSF_VERSION}")
    out, p = {}, 8
    while p < len(data):
        try:
            (ln,) = struct.unpack_from("<i", data, p)
            name = data[p + 4:p + 4 + ln].decode("ascii")
            p += 4 + ln
            (nd,) = struct.unpack_from("<i", data, p)
            dims = struct.unpack_from(f"<{nd}i", data, p + 4)
            p += 4 + 4 * nd
        except (struct.error, UnicodeDecodeError) as e:
            raise DumpError(f"{path}: truncated record header at byte {p}") from e
        cnt = int(np.prod(dims)) if nd else 1
        if ln < 0 or nd < 0 or min(dims, default=0) < 0 or p + 8 * cnt > len(data):
            raise DumpError(f"{path}: record {name!r} runs past the end of the file")
        out[name] = np.frombuffer(data, dtype="<f8", count=cnt, offset=p).reshape(tuple(reversed(dims))).copy()
        p += 8 * cnt
    return out


def write_batch_opsf(path: str, lam, out, tage=None) -> None:
    """The batch's outputs (host or device arrays: mags (n, nbands), and where present spec (n, nspec), mass, lbol,
    mdust (n)) as an OPSF fixture with the reference's record names; tage (n) becomes csp/age (in yr as the
    reference writes ocompsp%age: 10^9 tage)."""
    def host(a):
        return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    rec = {"wavelengths": np.asarray(lam, dtype=np.float64)}
    if tage is not None:
        rec["csp/age"] = 1.0e9 * np.asarray(tage, dtype=np.float64)
    for k in ("mass", "lbol", "mdust"):
        if k in out:
            rec["csp/" + k] = host(out[k]).astype(np.float64)
    if "spec" in out:
        rec["csp/spec"] = host(out["spec"]).astype(np.float64)
    rec["csp/mags"] = host(out["mags"]).astype(np.float64)
    write_opsf(path, rec)
