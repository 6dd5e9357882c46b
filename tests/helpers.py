"""Shared helpers of the tests: synthetic set-ups and the error measures."""
import functools

import numpy as np

from libopensps_amd import synth
import oracle as O


def spec_rel_err(test, ref, floor=1.0e-6):
    """Max relative error per pixel over the pixels at or above floor x the spectrum's peak (the measure of the
    reference's mixed-precision evidence, which excludes pixels below 1e-6 of the peak)."""
    test = np.asarray(test, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    peak = np.max(np.abs(ref), axis=-1, keepdims=True)
    mask = np.abs(ref) >= floor * peak
    err = np.where(mask, np.abs(test - ref) / np.where(mask, np.abs(ref), 1.0), 0.0)
    return float(err.max())


def mag_err(test, ref):
    test = np.asarray(test)
    ref = np.asarray(ref)
    both = np.isfinite(ref) & np.isfinite(test)
    assert np.array_equal(np.isfinite(ref), np.isfinite(test)), "NaN pattern differs"
    return float(np.max(np.abs(test[both] - ref[both]))) if both.any() else 0.0


@functools.lru_cache(maxsize=None)
def grid(nz=1, nspec=5994, nt=107, seed=1):
    return synth.ssp_grid(seed=seed, nspec=nspec, nt=nt, nz=nz)


@functools.lru_cache(maxsize=None)
def bands(nbands=20, nspec=5994, seed=7, mode=0):
    g = grid(1, nspec)
    fs = synth.gaussian_filters(nbands, seed=seed)
    return O.build_bands(fs, g.lam, mode)


@functools.lru_cache(maxsize=None)
def cosmo():
    return O.cosmology_table(72.0, 0.27, 0.73)


@functools.lru_cache(maxsize=None)
def operator(nz=3, npts=40, ntpl=300, seed=11):
    g = grid(nz)
    return synth.ssp_operator(g.lam, g.log_age, nz, npts=npts, ntpl=ntpl, seed=seed)
