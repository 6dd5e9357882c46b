"""The reverse pass on the GPU (opensps_hip_csp_batch_grad: k_grad_chain, the transposed product, k_grad_out,
k_sfh_grad) against the oracle's reverse pass (orc_csp_batch_grad, itself pinned by central differences of the
oracle's forward): params_bar rows (the SFH parameters' rows with SFH parameters), weights_z_bar (weights over ages
and metallicities) and weights_bar with the metallicity blend, the mass and L_bol cotangents, with and without the
birth-cloud and old-star screens. FP64 path: relative
to each output's largest magnitude per model; mixed path (FP32 chain, split-BF16 transposed product): looser."""
import numpy as np
import pytest

from helpers import cosmo, mag_err
from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

pytestmark = pytest.mark.gpu

NSPEC, NT, NZ, NB = 5994, 40, 5, 30
# FP64: summation order and the six-term IGM against the oracle; mixed: FP32 spectra and transmissions in the chain,
# bf16 hi/lo rows of y_bar in the transposed product
TOL = {0: {"params": 1.0e-10, "weights": 1.0e-11, "mags": 1.0e-10},
       1: {"params": 2.0e-4, "weights": 1.0e-4, "mags": 1.0e-4}}


def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _setup():
    g = synth.ssp_grid(seed=1, nspec=NSPEC, nt=NT, nz=NZ)
    fs = synth.gaussian_filters(NB, seed=7)
    return g, O.build_bands(fs, g.lam)


def _ctx(g, b):
    from libopensps_amd import hip as H
    ctx = H.DeviceContext(0)
    ctx.set_grid(g)
    ctx.set_bands(b)
    cz, cdl = cosmo()
    ctx.set_cosmology(cz, cdl)
    ctx.set_options(1, 1)
    return ctx


def _cot(n, seed=5):
    rng = np.random.default_rng(seed)
    return {"mags": rng.standard_normal((n, NB)), "spec": rng.standard_normal((n, NSPEC)) * 1e2,
            "mass": rng.standard_normal(n), "lbol": rng.standard_normal(n)}


def _rel(got, ref):
    """max |got - ref| / max |ref| per model (rows), over models."""
    got = np.asarray(got).reshape(got.shape[0], -1)
    ref = np.asarray(ref).reshape(ref.shape[0], -1)
    scale = np.max(np.abs(ref), axis=1, keepdims=True) + 1e-300
    return float(np.max(np.abs(got - ref) / scale))


ROWS = ["dust1", "dust2", "dust3", "dust_index", "dust1_index", "frac_obrun", "frac_nodust", "logzsol", "zred",
        "sigma_smooth", "igm_factor"]


def _compare(ref, got, precision, rows, tols=TOL):
    tol = tols[precision]
    em = mag_err(got["mags"], ref["mags"])
    print(f"  mags {em:.2e}")
    assert em < tol["mags"]
    pr, pg = ref["params_bar"], got["params_bar"]
    errs = {}
    for name in rows:
        r = O.GRAD_ROWS[name]
        scale = np.max(np.abs(pr[:, r])) + 1e-300
        err = np.abs(pg[:, r] - pr[:, r]) / scale
        errs[name] = float(np.max(err))
        print(f"  {name:12s} {errs[name]:.2e} (model {int(np.argmax(err))})")
    for name in rows:
        assert errs[name] < tol["params"], name
    other = [r for r in range(O.NBGRAD) if r not in [O.GRAD_ROWS[k] for k in rows]]
    assert np.all(pg[:, other] == 0.0)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("birth", [False, True])
def test_grad_weights_z(precision, birth):
    g, b = _setup()
    n = 24
    prm = synth.adjoint_params(n, seed=61, birth_cloud=birth)
    prm[SLOT["zred"], :3] = [0.0, 0.005, 6.5]   # the rest frame, a small redshift, beyond the IGM table's usual range
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=67)
    cot = _cot(n)
    ref = O.csp_batch_grad(g, b, prm, ssp_weights_z=wz, mags_bar=cot["mags"], spec_bar=cot["spec"],
                           mass_bar=cot["mass"], lbol_bar=cot["lbol"], cosmo=cosmo())
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: grad weights_z n={n} birth={birth} precision={precision}", flush=True)
    out = ctx.batch_grad(_dev(prm), {k: _dev(v) for k, v in cot.items()}, precision=precision,
                         split_parts=1 if birth else 0, ssp_weights_z=_dev(wz),
                         grads=("params_bar", "weights_bar", "weights_z_bar"))
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    rows = ["dust2", "dust_index", "zred", "sigma_smooth", "igm_factor", "dust1"]
    if birth:
        rows += ["dust3", "dust1_index", "frac_obrun", "frac_nodust"]
    else:
        rows += ["dust3", "frac_obrun", "frac_nodust"]
    _compare(ref, got, precision, rows)
    ew = _rel(got["weights_z_bar"], ref["weights_z_bar"])
    print(f"  weights_z_bar {ew:.2e}")
    assert ew < TOL[precision]["weights"]
    assert np.all(got["weights_bar"] == 0.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_grad_blend_weights(precision):
    g, b = _setup()
    n = 20
    prm = synth.adjoint_params(n, seed=71, birth_cloud=True)
    prm[SLOT["logzsol"]] = np.linspace(-2.0, 0.3, n)
    w = synth.dirichlet_weights(n, NT, seed=73, alpha=0.7)
    cot = _cot(n, seed=11)
    ref = O.csp_batch_grad(g, b, prm, ssp_weights=w, mags_bar=cot["mags"], spec_bar=cot["spec"],
                           mass_bar=cot["mass"], lbol_bar=cot["lbol"], cosmo=cosmo())
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: grad blend n={n} precision={precision}", flush=True)
    out = ctx.batch_grad(_dev(prm), {k: _dev(v) for k, v in cot.items()}, precision=precision, split_parts=1,
                         ssp_weights=_dev(w), grads=("params_bar", "weights_bar"))
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    _compare(ref, got, precision, ROWS)
    ew = _rel(got["weights_bar"], ref["weights_bar"])
    print(f"  weights_bar {ew:.2e}")
    assert ew < TOL[precision]["weights"]


def test_grad_statuses():
    """weights_bar alone is computed with SFH parameters."""
    from libopensps_amd import hip as H
    import torch
    g, b = _setup()
    n = 4
    prm = synth.parametric_params(n, seed=5)
    ctx = _ctx(g, b)
    print("case: grad statuses", flush=True)
    out = ctx.alloc_outputs(n, 0)
    wb = {"weights_bar": torch.zeros((n, NT), dtype=torch.float64, device="cuda")}
    ctx.launch_grad(_dev(prm), out, {"mags": _dev(np.ones((n, NB)))}, wb, precision=0)
    ctx.sync()
    assert torch.isfinite(wb["weights_bar"]).all()
    ctx.close()


SFH_ROWS = ["tage", "tau", "const", "fburst", "tburst", "sf_start", "sf_trunc", "sf_slope"]


def _sfh_prm(n, seed):
    """SFH types 0, 1, 4, 5 mixed, with const, bursts (before and after tage), start times, truncation and (Simha)
    slopes of both signs."""
    rng = np.random.default_rng(seed)
    prm = synth.adjoint_params(n, seed=seed, birth_cloud=True)
    prm[SLOT["sfh"]] = rng.choice([0, 1, 4, 5], n)
    prm[SLOT["sf_slope"]] = np.where(prm[SLOT["sfh"]] == 5, rng.uniform(-1.5, 1.0, n), 0.0)
    tage = rng.uniform(0.2, 13.0, n)
    prm[SLOT["tage"]] = tage
    prm[SLOT["tau"]] = rng.uniform(0.2, 9.0, n)
    prm[SLOT["const"]] = np.where(rng.random(n) < 0.5, rng.uniform(0.05, 0.4, n), 0.0)
    prm[SLOT["fburst"]] = np.where(rng.random(n) < 0.5, rng.uniform(0.05, 0.4, n), 0.0)
    prm[SLOT["tburst"]] = tage * rng.uniform(0.1, 1.3, n)
    start = np.where(rng.random(n) < 0.3, tage * rng.uniform(0.05, 0.3, n), 0.0)
    prm[SLOT["sf_start"]] = start
    prm[SLOT["sf_trunc"]] = np.where(rng.random(n) < 0.4, start + (tage - start) * rng.uniform(0.2, 0.9, n), 0.0)
    return prm


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("with_wbar", [False, True])
def test_grad_sfh_rows(precision, with_wbar):
    """params_bar with SFH parameters: the SFH tangents' rows (k_sfh_grad) against the oracle's batch_sfh_gradients,
    with weights_bar requested or not (then the tangents read an internal d L / d w)."""
    g, b = _setup()
    n = 45
    prm = _sfh_prm(n, 101)
    cot = _cot(n, seed=13)
    ref = O.csp_batch_grad(g, b, prm, mags_bar=cot["mags"], spec_bar=cot["spec"], mass_bar=cot["mass"],
                           lbol_bar=cot["lbol"], cosmo=cosmo())
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: grad sfh rows n={n} precision={precision} with_wbar={with_wbar}", flush=True)
    grads = ("params_bar", "weights_bar") if with_wbar else ("params_bar",)
    out = ctx.batch_grad(_dev(prm), {k: _dev(v) for k, v in cot.items()}, precision=precision, split_parts=1,
                         grads=grads)
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    _compare(ref, got, precision, ROWS + SFH_ROWS)
    if with_wbar:
        ew = _rel(got["weights_bar"], ref["weights_bar"])
        print(f"  weights_bar {ew:.2e}")
        assert ew < TOL[precision]["weights"]


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_public_api_grad(precision):
    """opensps_context_compute_csp_batch_grad through the Fortran host against the oracle."""
    from libopensps_amd import api
    g, _ = _setup()
    fs = synth.gaussian_filters(NB, seed=7)
    n = 12
    prm = synth.adjoint_params(n, seed=83, birth_cloud=True)
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=89)
    cot = _cot(n, seed=17)
    names = ["dust1", "dust2", "dust3", "dust_index", "dust1_index", "frac_obrun", "frac_nodust", "zred",
             "sigma_smooth"]
    print(f"case: public grad n={n} precision={precision}", flush=True)
    with api.Context() as ctx:
        ctx.set_ssp_basis(g)
        ctx.set_filters(fs)
        ctx.set_int("add_igm_absorption", 1)
        ctx.set_int("add_dust_emission", 0)
        ctx.set_str("kernel_precision", precision)
        r = ctx.compute_csp_batch_grad({k: prm[SLOT[k]] for k in names}, ssp_weights_z=wz, mags_bar=cot["mags"],
                                       spec_bar=cot["spec"], mass_bar=cot["mass"], lbol_bar=cot["lbol"],
                                       want=("mags", "spec", "mass", "lbol"), params_bar_rows=50)
        bands = ctx.bands()
    ref = O.csp_batch_grad(g, bands, prm, ssp_weights_z=wz, mags_bar=cot["mags"], spec_bar=cot["spec"],
                           mass_bar=cot["mass"], lbol_bar=cot["lbol"], cosmo=cosmo())
    assert ref["status"] == 0
    p = 0 if precision == "double" else 1
    assert np.all(r["params_bar"][:, O.NBGRAD:] == 0.0)   # rows past NBGRAD left as they were
    got = {"mags": r["mags"], "params_bar": r["params_bar"][:, :O.NBGRAD]}
    _compare(ref, got, p, ["dust1", "dust2", "dust3", "dust_index", "dust1_index", "frac_obrun", "frac_nodust",
                           "zred", "sigma_smooth", "igm_factor"])
    ew = _rel(r["weights_z_bar"], ref["weights_z_bar"])
    print(f"  weights_z_bar {ew:.2e}")
    assert ew < TOL[p]["weights"]


def test_autograd_chi2():
    """The torch.autograd.Function: a chi^2 loss on the magnitudes, loss.backward() -> prm.grad and the weights'
    grad against the oracle's reverse pass with the same cotangent."""
    import torch
    from libopensps_amd import autograd as AG
    g, b = _setup()
    n = 16
    prm = synth.adjoint_params(n, seed=97, birth_cloud=False)
    wz = synth.lognormal_weights_z(n, g.log_age, g.logz, seed=101)
    ctx = _ctx(g, b)
    print(f"case: autograd chi2 n={n}", flush=True)
    p_t = torch.tensor(prm, device="cuda", requires_grad=True)
    w_t = torch.tensor(wz, device="cuda", requires_grad=True)
    mags = AG.CSPMags.apply(ctx, p_t, w_t, 0)
    m_np = mags.detach().cpu().numpy()
    _, dmag = synth.chi2_cotangent(m_np)
    ok = torch.isfinite(mags)
    sig = 2.5 / np.log(10.0) * 0.05
    obs = torch.tensor(m_np - 0.5 * sig * sig * dmag, device="cuda")   # so that d loss / d mags = dmag
    r = torch.where(ok, (mags - obs) / sig, torch.zeros_like(mags))
    loss = (r * r).sum()
    loss.backward()
    ctx.close()
    ref = O.csp_batch_grad(g, b, prm, ssp_weights_z=wz, mags_bar=dmag, cosmo=cosmo())
    for name in ("dust2", "dust_index", "zred", "sigma_smooth", "igm_factor"):
        a = p_t.grad[SLOT[name]].cpu().numpy()
        e = ref["params_bar"][:, O.GRAD_ROWS[name]]
        err = float(np.max(np.abs(a - e)) / (np.max(np.abs(e)) + 1e-300))
        print(f"  {name:12s} {err:.2e}")
        assert err < 1e-9, name
    ew = _rel(w_t.grad.cpu().numpy(), ref["weights_z_bar"])
    print(f"  weights_z {ew:.2e}")
    assert ew < 1e-10


def test_autograd_sfh_parameters():
    """The torch.autograd.Function without weights: prm.grad in the SFH parameters' slots (the SFH tangents) against
    the oracle's reverse pass."""
    import torch
    from libopensps_amd import autograd as AG
    g, b = _setup()
    n = 20
    prm = _sfh_prm(n, 103)
    ctx = _ctx(g, b)
    print(f"case: autograd sfh n={n}", flush=True)
    p_t = torch.tensor(prm, device="cuda", requires_grad=True)
    mags = AG.CSPMags.apply(ctx, p_t, None, 0)
    dmag = np.random.default_rng(3).standard_normal((n, NB))
    (mags * torch.tensor(dmag, device="cuda")).sum().backward()
    ctx.close()
    ref = O.csp_batch_grad(g, b, prm, mags_bar=dmag, cosmo=cosmo())
    for name in SFH_ROWS + ["dust2", "logzsol"]:
        a = p_t.grad[SLOT[name]].cpu().numpy()
        e = ref["params_bar"][:, O.GRAD_ROWS[name]]
        err = float(np.max(np.abs(a - e)) / (np.max(np.abs(e)) + 1e-300))
        print(f"  {name:12s} {err:.2e}")
        assert err < 1e-9 or np.all(a == e), name


IMF_ROWS = {2: ["imf1", "imf2", "imf3"], 4: ["imf1", "imf2", "mdave"], 0: []}
# the operator's forward in FP64 sums 10^4-template products and the fold in another order than the oracle (its
# magnitudes agree to 1e-7 there, the faintest bands ~40 mag below the brightest setting the figure)
TOL_OP = {0: {"params": 1.0e-8, "weights": 1.0e-8, "mags": 1.0e-7},
          1: {"params": 1.0e-4, "weights": 5.0e-5, "mags": 1.0e-2}}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("imf_type,nz,remnants", [(2, 3, 1), (4, 3, 1), (0, 1, 0)])
def test_grad_operator(precision, imf_type, nz, remnants):
    """The reverse pass with the IMF per model (the operator: k_op_adj_t, k_op_imf_grad, k_op_adj_points,
    k_op_adj_out after the transposed template product) against the oracle's (orc_ops_model_adj, pinned by central
    differences): the IMF rows (Kroupa's three slopes, Dave's two slopes and break, none for Salpeter), logzsol
    through both metallicities' operators, the SFH rows from w_bar, the chain's rows, and weights_bar."""
    from helpers import bands, grid, operator
    op, g, b = operator(nz, 40, 300), grid(nz), bands(NB)
    n = 40
    prm = synth.parametric_params(n, seed=81, logzsol=(-2.1, 0.45) if nz > 1 else (0.0, 0.0))
    prm = synth.imf_params(n, seed=82, p=prm)
    prm[SLOT["dust1"]] = np.linspace(0.0, 1.0, n)
    prm[SLOT["imf_lower_limit"]] = np.linspace(0.08, 0.6, n)
    rng = np.random.default_rng(83)
    cot = {"mags": rng.standard_normal((n, NB)), "spec": rng.standard_normal((n, g.nspec)) * 1e2,
           "mass": rng.standard_normal(n), "lbol": rng.standard_normal(n)}
    mdave = 0.7
    # magnitude cotangents on the bands within 30 mag of each model's brightest: a band 40 mag fainter carries FP32
    # subnormal fluxes in the mixed chain, and its d mag / d F = -2.5 / (F ln 10) would set the mixed figures
    fwd = O.csp_batch(g, b, prm, add_igm=1, cosmo=cosmo(), want=("mags",), ops=op, imf_type=imf_type,
                      remnants=remnants, mdave=mdave)["mags"]
    cot["mags"] = np.where(fwd <= np.nanmin(fwd, axis=1, keepdims=True) + 30.0, cot["mags"], 0.0)
    ref = O.csp_batch_grad(g, b, prm, mags_bar=cot["mags"], spec_bar=cot["spec"], mass_bar=cot["mass"],
                           lbol_bar=cot["lbol"], cosmo=cosmo(), ops=op, imf_type=imf_type, remnants=remnants,
                           mdave=mdave)
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    ctx.set_ssp_ops(op)
    ctx.set_imf(imf_type, remnants, 2.35, mdave, 40.0, 8.5)
    print(f"case: grad operator n={n} nz={nz} imf={imf_type} precision={precision}", flush=True)
    out = ctx.batch_grad(_dev(prm), {k: _dev(v) for k, v in cot.items()}, precision=precision, split_parts=1,
                         grads=("params_bar", "weights_bar"), use_ops=1)
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    rows = ["dust1", "dust2", "dust3", "dust_index", "dust1_index", "frac_obrun", "frac_nodust", "zred",
            "sigma_smooth", "igm_factor"] + SFH_ROWS + (["logzsol"] if nz > 1 else []) + IMF_ROWS[imf_type]
    # the magnitudes of the bands within 30 mag of each model's brightest (the fainter ones to TOL_OP's figure). The
    # metallicities below and above the grid extrapolate (dz < 0 or > 1): one operator's sums are then subtracted
    # from the other's, which magnifies the summation-order differences (FP64: ~1e-9 mag there, 1e-12 inside)
    rm = ref["mags"]
    near = rm <= np.nanmin(rm, axis=1, keepdims=True) + 30.0
    en = float(np.nanmax(np.abs(got["mags"] - rm)[near]))
    print(f"  mags within 30 mag {en:.2e}")
    assert en < (1.0e-8 if precision == 0 else 1.0e-3)
    _compare(ref, got, precision, rows, TOL_OP)
    ew = _rel(got["weights_bar"], ref["weights_bar"])
    print(f"  weights_bar {ew:.2e}")
    assert ew < TOL_OP[precision]["weights"]


@pytest.mark.parametrize("precision", [0, 1])
def test_grad_zred_mags(precision):
    """Magnitudes at a second redshift (zred_mags): the forward's observed frame and distance term at it, the IGM
    at zred, and zred's row the IGM's part alone, against the oracle's; z = 0 and zred_mags = 0 included."""
    g, b = _setup()
    n = 24
    prm = synth.adjoint_params(n, seed=121, birth_cloud=True)
    rng = np.random.default_rng(123)
    prm[SLOT["zred"]] = rng.uniform(0.0, 4.0, n)
    zm = rng.uniform(0.0, 4.0, n)
    prm[SLOT["zred"], 0] = 0.0
    zm[1] = 0.0
    w = synth.dirichlet_weights(n, NT, seed=125, alpha=0.7)
    cot = _cot(n, seed=127)
    ref = O.csp_batch_grad(g, b, prm, ssp_weights=w, mags_bar=cot["mags"], spec_bar=cot["spec"],
                           mass_bar=cot["mass"], lbol_bar=cot["lbol"], cosmo=cosmo(), zred_mags=zm)
    assert ref["status"] == 0
    ctx = _ctx(g, b)
    print(f"case: grad zred_mags n={n} precision={precision}", flush=True)
    out = ctx.batch_grad(_dev(prm), {k: _dev(v) for k, v in cot.items()}, precision=precision, split_parts=1,
                         ssp_weights=_dev(w), grads=("params_bar", "weights_bar"), zred_mags=_dev(zm))
    got = {k: v.detach().cpu().numpy() for k, v in out.items()}
    ctx.close()
    _compare(ref, got, precision, ROWS)
    ew = _rel(got["weights_bar"], ref["weights_bar"])
    print(f"  weights_bar {ew:.2e}")
    assert ew < TOL[precision]["weights"]
