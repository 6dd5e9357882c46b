"""Parity of the GPU batch (libopensps_hip.so) with the FP64 oracle on the same seeded inputs.

Tolerances: the FP64 path agrees with the oracle to round-off amplified by the SFH moment method's cancellation
(below): spectra 1e-7 relative, magnitudes 1e-7 mag, SFH weights 1e-7 relative to the model's largest weight. The
mixed path (split-BF16 age sums, FP32 chain) is held to the targets of the build: spectra 1e-4 relative per
pixel (pixels at or above 1e-6 of the spectrum's peak), magnitudes 1e-3 mag. Index selection (it_tesc, zlo)
is bit-exact in both.
"""
import numpy as np
import pytest

from helpers import bands, cosmo, grid, mag_err, spec_rel_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

# FP64: the moment method of the SFH weights differences large indefinite integrals (the reference's algorithm,
# src/physics/opensps_sfh.f90:97-257), so a 1-ulp difference between the device's and the host's exp / pow / log10
# moves a weight by up to ~1e-8 of the model's largest; spectra and magnitudes follow linearly.
FP64_SPEC, FP64_MAG, FP64_W = 1.0e-7, 1.0e-7, 1.0e-7
MIX_SPEC, MIX_MAG = 1.0e-4, 1.0e-3


def _ctx(g, b, add_igm=1, smooth_fast=1):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(add_igm, smooth_fast)
    return ctx


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def _host(t):
    return t.detach().cpu().numpy()


def _run_both(g, b, prm, precision, weights=None, split=0, add_igm=1, want=("mags", "spec", "mass", "lbol", "idx")):
    from libopensps_amd import hip as H
    ref = O.csp_batch(g, b, prm, ssp_weights=weights, add_igm=add_igm, cosmo=cosmo(),
                      want=tuple(want) + (("w",) if weights is None else ()))
    assert ref["status"] == 0
    ctx = _ctx(g, b, add_igm)
    print(f"case: n={prm.shape[1]} nz={g.nz} precision={precision} split={split}", flush=True)
    out = ctx.batch(_dev(prm), precision=precision, split_parts=split,
                    ssp_weights=None if weights is None else _dev(weights),
                    want=tuple(want) + (("weights",) if weights is None else ()))
    got = {k: _host(v) for k, v in out.items()}
    ctx.close()
    return ref, got


def _check(ref, got, precision, check_w=True):
    from libopensps_amd import hip as H
    fp64 = precision == H.FP64
    ew = 0.0
    if check_w and "w" in ref:
        wmax = np.max(np.abs(ref["w"]), axis=1, keepdims=True)
        ew = float(np.max(np.abs(got["weights"] - ref["w"]) / wmax))
    es = spec_rel_err(got["spec"], ref["spec"])
    em = mag_err(got["mags"], ref["mags"])
    emass = float(np.max(np.abs(got["mass"] / ref["mass"] - 1.0)))
    print(f"weights err {ew:.3e}, mass rel err {emass:.3e}, max spectral rel err {es:.3e}, max mag err {em:.3e}")
    assert np.array_equal(got["idx"][:, :2], ref["idx"][:, :2]), "it_tesc / zlo differ"
    assert np.all(got["idx"][:, 2] == 0)
    assert ew < FP64_W
    assert emass < FP64_W
    np.testing.assert_allclose(got["lbol"], ref["lbol"], rtol=0, atol=FP64_W)
    assert es < (FP64_SPEC if fp64 else MIX_SPEC)
    assert em < (FP64_MAG if fp64 else MIX_MAG)


@pytest.mark.gpu
def test_selftest_mfma_f64_layout():
    from libopensps_amd import hip as H
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(64), rng.standard_normal(64)
    A, B = np.zeros((16, 4)), np.zeros((4, 16))
    for l in range(64):
        A[l % 16, l // 16] = a[l]
        B[l // 16, l % 16] = b[l]
    D = A @ B
    d = H.selftest_mfma(0, a, b)
    # the result layout the FP64 product's epilogue assumes: register r of lane l holds D[l // 16 + 4 r][l % 16]
    for l in range(64):
        for r in range(4):
            assert abs(d[l, r] - D[l // 16 + 4 * r, l % 16]) < 1e-12


@pytest.mark.gpu
def test_selftest_mfma_bf16_layout():
    from libopensps_amd import hip as H
    rng = np.random.default_rng(1)
    av = rng.integers(-4, 5, size=(64, 8)).astype(np.float32)   # exact in bf16
    bv = rng.integers(-4, 5, size=(64, 8)).astype(np.float32)
    to_bf16 = lambda x: (x.view(np.uint32) >> 16).astype(np.uint16)
    A, B = np.zeros((16, 32)), np.zeros((32, 16))
    for l in range(64):
        for e in range(8):
            A[l % 16, 8 * (l // 16) + e] = av[l, e]
            B[8 * (l // 16) + e, l % 16] = bv[l, e]
    D = A @ B
    d = H.selftest_mfma(1, to_bf16(av), to_bf16(bv))
    for l in range(64):
        for r in range(4):
            assert d[l, r] == D[4 * (l // 16) + r, l % 16]


def _weight_case(n=64):
    g = grid(1)
    b = bands(20)
    prm = synth.default_params(n)
    w = synth.dirichlet_weights(n, g.nt, seed=3)
    return g, b, prm, w


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_weight_matrix_batch(precision):
    """A given SSP-weight matrix, one metallicity, no dust, rest frame (the 256-model workload's shape)."""
    g, b, prm, w = _weight_case(64)
    ref, got = _run_both(g, b, prm, precision, weights=w)
    _check(ref, got, precision, check_w=False)


def _parametric_case(n, seed=5):
    g = grid(22)
    b = bands(100)
    prm = synth.parametric_params(n, seed=seed)
    return g, b, prm


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_parametric_batch(precision):
    """Delayed-tau SFH, logzsol blend over 22 metallicities, power-law dust, IGM, redshift, velocity smoothing."""
    g, b, prm = _parametric_case(192)
    ref, got = _run_both(g, b, prm, precision)
    _check(ref, got, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_uniform_redshift_batch(precision):
    """Every model at one redshift (the reference's uniform-z IGM form) and one sigma."""
    g, b, prm = _parametric_case(48, seed=11)
    prm[SLOT["zred"]] = 2.7
    prm[SLOT["sigma_smooth"]] = 150.0
    ref, got = _run_both(g, b, prm, precision)
    _check(ref, got, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_birth_cloud_and_parts(precision):
    """dust1, dust3, frac_obrun, frac_nodust: young and old sums apart; the parts equal the oracle's age sums."""
    g, b, prm = _parametric_case(64, seed=13)
    rng = np.random.default_rng(2)
    prm[SLOT["dust1"]] = rng.uniform(0.0, 1.5, prm.shape[1])
    prm[SLOT["dust3"], ::3] = 0.4
    prm[SLOT["frac_obrun"]] = rng.uniform(0.0, 0.5, prm.shape[1])
    prm[SLOT["frac_nodust"], ::4] = 0.2
    prm[SLOT["dust_tesc"]] = rng.uniform(6.5, 8.0, prm.shape[1])
    ref, got = _run_both(g, b, prm, precision, split=1,
                         want=("mags", "spec", "mass", "lbol", "idx", "young", "old"))
    _check(ref, got, precision)
    tol = FP64_SPEC if precision == 0 else MIX_SPEC
    assert spec_rel_err(got["young"], ref["young"]) < tol
    assert spec_rel_err(got["old"], ref["old"]) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("sfh", [0, 1, 5])
def test_other_sfh_types(sfh):
    """SSP (sfh 0), tau model (sfh 1) with constant and burst components, Simha (sfh 5: truncation, slopes of both
    signs, a start time), FP64 path."""
    g, b, prm = _parametric_case(64, seed=17)
    prm[SLOT["sfh"]] = sfh
    if sfh == 5:
        rng = np.random.default_rng(6)
        n = prm.shape[1]
        start = np.where(rng.random(n) < 0.3, 0.1 * prm[SLOT["tage"]], 0.0)
        prm[SLOT["sf_start"]] = start
        prm[SLOT["sf_trunc"]] = np.where(rng.random(n) < 0.8, start + (prm[SLOT["tage"]] - start) *
                                         rng.uniform(0.2, 0.9, n), 0.0)
        prm[SLOT["sf_slope"]] = rng.uniform(-1.5, 1.0, n)
    if sfh == 1:
        rng = np.random.default_rng(4)
        prm[SLOT["const"]] = rng.uniform(0.0, 0.3, prm.shape[1])
        prm[SLOT["fburst"]] = rng.uniform(0.0, 0.3, prm.shape[1])
        prm[SLOT["tburst"]] = rng.uniform(0.5, 12.0, prm.shape[1])
    ref, got = _run_both(g, b, prm, 0)
    _check(ref, got, 0)


@pytest.mark.gpu
def test_unsupported_sfh_status():
    from libopensps_amd import hip as H
    g, b, prm = _parametric_case(8)
    prm[SLOT["sfh"]] = 6   # no such SFH type
    ctx = _ctx(g, b)
    with pytest.raises(H.OpenSPSHipError) as e:
        ctx.batch(_dev(prm), precision=H.FP64)
    assert e.value.status == 2
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_tabular_sfh(precision):
    """sfh 3 (and 2): the context's SFH table as linear segments up to each model's age, one metallicity with the
    per-model blend, against the oracle; ages before, inside and after the table's span."""
    from libopensps_amd import hip as H
    g, b, prm = _parametric_case(48, seed=31)
    n = prm.shape[1]
    rng = np.random.default_rng(8)
    age = np.sort(np.concatenate([[0.0], rng.uniform(0.02, 11.0, 40)]))
    sfr = rng.uniform(0.1, 30.0, age.size)
    prm[SLOT["sfh"]] = np.where(np.arange(n) % 2, 3.0, 2.0)
    prm[SLOT["tage"]] = np.linspace(0.05, 13.5, n)
    O.set_sfh_tab(age, sfr)
    try:
        print(f"case: tabular sfh n={n} precision={precision}", flush=True)
        want = ("mags", "spec", "mass", "lbol", "idx")
        ref = O.csp_batch(g, b, prm, cosmo=cosmo(), want=want + ("w",))
        assert ref["status"] == 0
        ctx = _ctx(g, b)
        ctx.set_sfh_tab(age, sfr)
        out = ctx.batch(_dev(prm), precision=precision, want=want + ("weights",))
        got = {k: v.detach().cpu().numpy() for k, v in out.items()}
        ctx.close()
    finally:
        O.set_sfh_tab([], [])
    _check(ref, got, precision)


def _lcg(seed):
    while True:
        seed = (1103515245 * seed + 12345) & 2147483647
        yield seed / 2147483648.0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_band_fluxes_reference_fixture(precision):
    """The reference's band_fluxes unit test (tests/spectra/test_opensps_batch_photometry.f90:21-126): its seeded
    LCG data, supports and band mask; FP64 sums against the oracle, bit-identical repeats, exact 2x scaling, a
    zero model, skipped bands exactly 0."""
    import torch
    from libopensps_amd import hip as H
    NSPEC, NBANDS, NMAX = 700, 12, 100
    J_LO = np.array([1, 1, 699, 100, 250, 1, 1, 300, 10, 100, 2, 600], np.int32)
    J_HI = np.array([699, 1, 699, 100, 260, 256, 257, 655, 400, 611, 40, 699], np.int32)
    DO = np.array([1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 0], np.int32)
    u = _lcg(20261107)
    wt = np.zeros(NSPEC)
    bt = np.zeros((NSPEC, NBANDS))
    fo = np.zeros((NSPEC, NMAX))
    for j in range(NSPEC):
        wt[j] = 0.5 + next(u)
        for bb in range(NBANDS):
            bt[j, bb] = next(u)
        for m in range(NMAX):
            fo[j, m] = 2.0 * next(u)
        fo[j, 1] = 2.0 * fo[j, 0]
        fo[j, 2] = 0.0
    dt = np.float64 if precision == 0 else np.float32
    bt, wt, fo = bt.astype(dt), wt.astype(dt), fo.astype(dt)
    for n in (1, 2, 7, 17, 41, NMAX):
        f = _host(H.band_fluxes(precision, _dev(bt), _dev(J_LO), _dev(J_HI), _dev(DO), _dev(wt),
                                _dev(np.ascontiguousarray(fo[:, :n]))))
        f2 = _host(H.band_fluxes(precision, _dev(bt), _dev(J_LO), _dev(J_HI), _dev(DO), _dev(wt),
                                 _dev(np.ascontiguousarray(fo[:, :n]))))
        assert np.array_equal(f, f2)
        for m in range(n):
            ref = O.band_fluxes_col(bt.astype(np.float64), J_LO, J_HI, DO, wt.astype(np.float64),
                                    fo[:, m].astype(np.float64))
            np.testing.assert_allclose(f[m], ref, rtol=1e-11, atol=0)
        assert np.all(f[:, DO == 0] == 0.0)
        if n >= 3:
            assert np.array_equal(f[1], 2.0 * f[0])
            assert np.all(f[2] == 0.0)
    torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_observed_frame_reference_fixture(precision):
    """The reference's observed_frame_z test (tests/spectra/test_opensps_batch_photometry.f90:139-208): a
    non-uniform grid, redshifts below, at and above its range, a rest-frame column, NaN and negative values."""
    from libopensps_amd import hip as H
    NS, NC = 70, 9
    j = np.arange(1, NS + 1)
    lam = 900.0 * np.exp(0.045 * (j - 1)) * (1.0 + 0.01 * np.sin(j))
    u = _lcg(12345)
    total = np.array([[1.0 + next(u) for _ in range(NS)] for _ in range(NC)])
    total[2, 4] = np.nan
    total[3, 8] = -1.0
    zc = np.array([0.0, 0.05, 0.4, 1.0, 2.5, 6.0, -0.3, 1.0e-12, 0.7])
    dt = np.float64 if precision == 0 else np.float32
    tot = total.astype(dt)
    f = _host(H.observed_frame(precision, _dev(zc), _dev(lam), _dev(tot)))
    for m in range(NC):
        ref = O.observed_frame(lam, zc[m], tot[m].astype(np.float64))
        tol = 1e-15 if precision == 0 else 4.0 * np.finfo(np.float32).eps
        np.testing.assert_allclose(f[:, m], ref, rtol=tol, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("with_spec", [True, False])
def test_wide_smoothing_window(precision, with_spec):
    """Smoothing over the whole grid (min_wave_smooth 50 A, max_wave_smooth 2e8 A): more smoothed pixels than the
    single-buffer chain holds in registers, so they pass through the spectrum out, or through the staging rows
    when no spectrum is asked for."""
    g, b, prm = _parametric_case(48, seed=31)
    prm[SLOT["min_wave_smooth"]] = 50.0
    prm[SLOT["max_wave_smooth"]] = 2.0e8
    want = ("mags", "spec", "mass", "lbol", "idx") if with_spec else ("mags", "mass", "lbol", "idx")
    ref, got = _run_both(g, b, prm, precision, want=want)
    fp64 = precision == 0
    em = mag_err(got["mags"], ref["mags"])
    print(f"wide window: max mag err {em:.3e}")
    assert em < (FP64_MAG if fp64 else MIX_MAG)
    if with_spec:
        es = spec_rel_err(got["spec"], ref["spec"])
        print(f"wide window: max spectral rel err {es:.3e}")
        assert es < (FP64_SPEC if fp64 else MIX_SPEC)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_broad_velocity_kernels(precision):
    """Velocity dispersions of 50 to 3000 km/s on the 5994-pixel grid (~700 km/s a pixel): smoothing windows from one
    to about eighteen pixels a side, so the FP32 chain takes both its five-tap form (within two pixels a side) and the
    general sum, and its batched window lookup (seven pixels) both settles windows and walks on past them."""
    g, b, prm = _parametric_case(64, seed=41)
    prm[SLOT["sigma_smooth"]] = np.random.default_rng(41).uniform(50.0, 3000.0, prm.shape[1])
    ref, got = _run_both(g, b, prm, precision)
    _check(ref, got, precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_two_buffer_chain(precision, monkeypatch):
    """The two-buffer chain (taken where the band pixels outgrow the single-buffer chain's registers), forced on the
    parametric batch."""
    monkeypatch.setenv("OPENSPS_CHAIN", "two")
    g, b, prm = _parametric_case(64, seed=37)
    ref, got = _run_both(g, b, prm, precision)
    _check(ref, got, precision)
