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
