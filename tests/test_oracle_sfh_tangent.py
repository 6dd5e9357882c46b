"""Pins of the oracle's SFH tangents (orc_sfh_weights_tangent, the restatement of sfh_weights_tangent,
src/physics/opensps_sfh_tangent.f90:66-251) by differences of the oracle's SFH weights (orc_sfh_weights, itself
pinned in test_oracle.py): for sfh 0, 1, 4 with const, fburst, the burst's age, a start time and truncation, and 5
(Simha, with its slope and the linear part's mass fraction, simha_frac_t), every
parameter's derivative against central differences (h = 1e-4 relative: the weights are moment differences of
~1e10-year quantities, so smaller steps see their rounding), away from the kinks where a window end or the burst
crosses a grid age; const and fburst at zero against the one-sided (upward) difference; the weights the tangent
returns equal the forward's bit for bit. A dropped normalisation term, a wrong sign on a window end or the burst,
or a lost d Ei / dx fails one of them."""
import numpy as np
import pytest

from libopensps_amd import synth
from libopensps_amd.params import SLOT

import oracle as O

TF = synth.age_grid(107)
KINK = 2.0e-3   # dex: sampled window ends and burst ages stay this far from a grid age


def _weights(p):
    w, st = O.sfh_weights(TF, p)
    assert st == 0
    return w


def _far(v):
    lv = np.log10(v)
    return np.min(np.abs(TF - lv)) > KINK


def _sample(rng, sfh):
    """One model of SFH type sfh whose window ends, burst age and age bracket lie away from grid ages."""
    while True:
        p = synth.default_params(1)[:, 0].copy()
        p[SLOT["sfh"]] = sfh
        tage = rng.uniform(0.05, 13.5)
        p[SLOT["tage"]] = tage
        p[SLOT["tau"]] = rng.uniform(0.1, 10.0)
        if rng.random() < 0.6:
            p[SLOT["const"]] = rng.uniform(0.05, 0.4)
        if rng.random() < 0.6:
            p[SLOT["fburst"]] = rng.uniform(0.05, 0.4)
            p[SLOT["tburst"]] = tage * (rng.uniform(0.05, 0.9) if rng.random() < 0.8 else rng.uniform(1.1, 1.5))
        start = 0.0
        if sfh != 0 and rng.random() < 0.4:
            start = tage * rng.uniform(0.05, 0.4)
            p[SLOT["sf_start"]] = start
        if sfh != 0 and rng.random() < (0.9 if sfh == 5 else 0.5):
            p[SLOT["sf_trunc"]] = start + (tage - start) * rng.uniform(0.1, 0.9)
        if sfh == 5:
            p[SLOT["sf_slope"]] = rng.choice([-1.0, 1.0]) * rng.uniform(0.05, 1.0)
        ti = (tage - start) * 1e9
        ends = [ti, tage * 1e9, (tage - p[SLOT["tburst"]]) * 1e9]
        if p[SLOT["sf_trunc"]] > 0.0:
            tq = ti - (p[SLOT["sf_trunc"]] - start) * 1e9
            ends.append(tq)
            if p[SLOT["sf_slope"]] < 0.0:   # Simha's linear SFR reaches zero at lookback t0 = tq + 1e9 / sf_slope
                ends.append(tq + 1e9 / p[SLOT["sf_slope"]])   # (ends <= 0 are not checked)
        if all(e <= 0.0 or _far(e) for e in ends):
            return p


def _fd(p, name):
    x = p[SLOT[name]]
    if x == 0.0:   # const, fburst at 0: the one-sided derivative, second order (-3 f0 + 4 f1 - f2) / 2h
        h = 1e-5
        p1, p2 = p.copy(), p.copy()
        p1[SLOT[name]] = h
        p2[SLOT[name]] = 2 * h
        return (4 * _weights(p1) - _weights(p2) - 3 * _weights(p)) / (2 * h)
    h = abs(x) * 1e-4
    pp, pm = p.copy(), p.copy()
    pp[SLOT[name]] = x + h
    pm[SLOT[name]] = x - h
    return (_weights(pp) - _weights(pm)) / (2 * h)


@pytest.mark.parametrize("sfh", [0, 1, 4, 5])
def test_tangent_against_differences(sfh):
    rng = np.random.default_rng(100 + sfh)
    checked = 0
    for _ in range(40):
        p = _sample(rng, sfh)
        for q, name in enumerate(O.SFHT_PARAMS, start=1):
            w, dw, st = O.sfh_weights_tangent(TF, p, q)
            assert st == 0
            assert np.array_equal(w, _weights(p))
            if (name == "sf_slope" and sfh != 5) or (p[SLOT[name]] == 0.0 and name not in ("const", "fburst")):
                continue
            fd = _fd(p, name)
            scale = max(np.max(np.abs(fd)), np.max(np.abs(dw)))
            # 1e-4 of the largest entry, plus the forward's rounding over the step (~1e-10 of a weight per 1e-4)
            err = np.max(np.abs(fd - dw))
            assert err <= 1e-4 * scale + 2e-9 / (abs(p[SLOT[name]]) * 1e-4 + 1e-5), (sfh, name, err, scale, p)
            checked += 1
    assert checked > 100


def test_zero_rows():
    """sf_slope acts on Simha's linear part only: zero for sfh 0, 1, 4; without truncation (sf_trunc = 0), sf_trunc's
    derivative is zero; for an SSP (sfh = 0) only tage moves the weights; Simha has no const, burst or tburst."""
    rng = np.random.default_rng(7)
    for sfh in (0, 1, 4, 5):
        p = _sample(rng, sfh)
        p[SLOT["sf_trunc"]] = 0.0
        for q, name in enumerate(O.SFHT_PARAMS, start=1):
            _, dw, _ = O.sfh_weights_tangent(TF, p, q)
            if name in ("sf_slope", "sf_trunc") or (sfh == 0 and name != "tage") or \
                    (sfh == 5 and name in ("const", "fburst", "tburst")):
                assert np.all(dw == 0.0), (sfh, name)


def test_normalised_derivative_sums_to_zero():
    """The weights sum to one solar mass formed (over the ages up to imax) whatever the parameters: the tangents of
    tau, tage and sf_trunc along a pure tau / delayed-tau history with const sum to zero."""
    rng = np.random.default_rng(11)
    for sfh in (1, 4, 5):
        for _ in range(10):
            p = _sample(rng, sfh)
            p[SLOT["fburst"]] = 0.0
            for name in ("tage", "tau", "sf_trunc", "const", "sf_start") + (("sf_slope",) if sfh == 5 else ()):
                q = O.SFHT_PARAMS.index(name) + 1
                w, dw, _ = O.sfh_weights_tangent(TF, p, q)
                assert abs(np.sum(w) - 1.0) < 1e-12
                assert abs(np.sum(dw)) < 1e-9 * max(1.0, np.max(np.abs(dw))), (sfh, name, np.sum(dw))


def test_unsupported_type():
    p = synth.default_params(1)[:, 0].copy()
    p[SLOT["sfh"]] = 2   # tabular
    p[SLOT["tage"]] = 5.0
    _, _, st = O.sfh_weights_tangent(TF, p, 1)
    assert st == 2
