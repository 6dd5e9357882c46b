"""The dump files of the reference's batch benchmark (tests/benchmark_csp_batch.f90:426-452 of the reference): the
bytes a Fortran stream write of mags(nbands, n), spec(nspec, n) and params_bar(NBGRAD, n) gives, read back, and
refused when the size does not match the header (as its tools/compare_forward_dump.py refuses them)."""
import struct

import numpy as np
import pytest

from libopensps_amd import dumps


def test_forward_bytes_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n, nb, ns = 5, 3, 7
    mags, spec = rng.standard_normal((n, nb)), rng.standard_normal((n, ns))
    mags[1, 2] = np.nan
    p = tmp_path / "fwd.bin"
    dumps.write_forward(str(p), mags, spec)
    raw = p.read_bytes()
    # the Fortran layout: header, then mags(nbands, n) and spec(nspec, n) column-major
    want = struct.pack("<3i", nb, n, ns) + np.asfortranarray(mags.T).tobytes(order="F") + \
        np.asfortranarray(spec.T).tobytes(order="F")
    assert raw == want
    m2, s2 = dumps.read_forward(str(p))
    assert m2.tobytes() == mags.tobytes() and s2.tobytes() == spec.tobytes()
    p.write_bytes(raw[:-8])
    with pytest.raises(dumps.DumpError, match="size"):
        dumps.read_forward(str(p))


def test_gradient_bytes_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    pb = rng.standard_normal((4, 44))
    p = tmp_path / "grad.bin"
    dumps.write_gradients(str(p), pb)
    raw = p.read_bytes()
    assert raw == struct.pack("<2i", 44, 4) + np.asfortranarray(pb.T).tobytes(order="F")
    assert dumps.read_gradients(str(p)).tobytes() == pb.tobytes()
    p.write_bytes(raw + b"\0")
    with pytest.raises(dumps.DumpError, match="size"):
        dumps.read_gradients(str(p))


def _fortran_record(name, arr_f):
    """One OPSF record as tools/legacy_parity/csp_dump.f90's put() writes it (a Fortran stream write of the name
    length, the name, the rank, the dimensions, then the values in Fortran order): arr_f is the Fortran-shaped array."""
    a = np.asfortranarray(arr_f, dtype="<f8")
    dims = a.shape
    return (struct.pack("<i", len(name)) + name.encode() + struct.pack("<i", len(dims)) +
            struct.pack(f"<{len(dims)}i", *dims) + a.tobytes(order="F"))


def test_opsf_bytes_and_round_trip(tmp_path):
    """The OPSF container: the bytes the reference's writer gives for spec(nspec, n), mags(nbands, n) and mass(n),
    byte for byte, and each record read back as the C-order (n, quantity) array."""
    rng = np.random.default_rng(3)
    n, nb, ns = 4, 3, 6
    lam = np.linspace(100.0, 1e4, ns)
    spec, mags, mass = rng.random((n, ns)), rng.standard_normal((n, nb)), rng.random(n)
    mags[2, 1] = np.nan
    p = tmp_path / "case.opsf"
    dumps.write_opsf(str(p), {"wavelengths": lam, "csp/spec": spec, "csp/mags": mags, "csp/mass": mass})
    want = (b"OPSF" + struct.pack("<i", 1) + _fortran_record("wavelengths", lam) +
            _fortran_record("csp/spec", spec.T) + _fortran_record("csp/mags", mags.T) +
            _fortran_record("csp/mass", mass))
    assert p.read_bytes() == want
    r = dumps.read_opsf(str(p))
    assert list(r) == ["wavelengths", "csp/spec", "csp/mags", "csp/mass"]
    assert r["csp/spec"].shape == (n, ns) and r["csp/spec"].tobytes() == spec.tobytes()
    assert r["csp/mags"].tobytes() == mags.tobytes() and r["csp/mass"].tobytes() == mass.tobytes()
    # a file the reference wrote: a rank-2 record with Fortran dims (2, 3) is the C array of shape (3, 2)
    f = np.arange(6.0).reshape(2, 3, order="F")
    q = tmp_path / "ref.opsf"
    q.write_bytes(b"OPSF" + struct.pack("<i", 1) + _fortran_record("ssp/spec", f))
    assert np.array_equal(dumps.read_opsf(str(q))["ssp/spec"], f.T)


def test_opsf_refuses_bad_files(tmp_path):
    p = tmp_path / "bad.opsf"
    p.write_bytes(b"OPSX" + struct.pack("<i", 1))
    with pytest.raises(dumps.DumpError, match="not an OPSF"):
        dumps.read_opsf(str(p))
    p.write_bytes(b"OPSF" + struct.pack("<i", 2))
    with pytest.raises(dumps.DumpError, match="version"):
        dumps.read_opsf(str(p))
    good = b"OPSF" + struct.pack("<i", 1) + _fortran_record("csp/mass", np.ones(5))
    p.write_bytes(good[:-8])
    with pytest.raises(dumps.DumpError, match="past the end"):
        dumps.read_opsf(str(p))


def test_batch_opsf_names(tmp_path):
    """The batch's outputs under the reference's record names, ages in yr."""
    rng = np.random.default_rng(4)
    out = {"mags": rng.standard_normal((3, 2)), "spec": rng.random((3, 5)).astype(np.float32),
           "mass": rng.random(3), "lbol": rng.standard_normal(3)}
    p = tmp_path / "b.opsf"
    dumps.write_batch_opsf(str(p), np.arange(5.0), out, tage=np.array([1.0, 2.0, 13.7]))
    r = dumps.read_opsf(str(p))
    assert set(r) == {"wavelengths", "csp/age", "csp/mass", "csp/lbol", "csp/spec", "csp/mags"}
    assert np.allclose(r["csp/age"], [1e9, 2e9, 1.37e10]) and r["csp/spec"].dtype == np.float64
    assert np.array_equal(r["csp/spec"], out["spec"].astype(np.float64))
