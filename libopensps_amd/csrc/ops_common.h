// Device helpers shared by the kernels of the batched composite-population path (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/opensps_params.h"

// LDS written by some lanes of a wavefront and read by others: a compiler barrier plus wavefront-scope fences
#define WAVE_SYNC()                                           \
    do {                                                      \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                      \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
    } while (0)

#define OPS_SAFE_FLOOR 2.2250738585072014e-308   // tiny(1.0_wp)
#define OPS_C_LIGHT 2.99792458e18                // Angstrom / s
#define OPS_LY_ALPHA 1215.67

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4_t;

// Number of entries of the ascending array a[0..n) that are <= v (binary search). For a grid of
// distinct ascending values, max(1, min(n - 1, count_le)) is the reference's clamped find_interval
// (src/math/opensps_interpolation.f90:149-198 with the callers' clamp to [1, n - 1]), 1-based.
__device__ __forceinline__ int count_le(const double* __restrict__ a, int n, double v) {
    int lo = 0, hi = n;   // invariant: a[0..lo) <= v, a[hi..n) > v
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The reference's find_interval for either order with its end rules, 1-based (0 below, n above).
__host__ __device__ __forceinline__ int find_interval_ref(const double* __restrict__ a, int n, double v) {
    if (n < 2) return 1;
    bool asc = a[n - 1] >= a[0];
    int lower = 0, upper = n + 1;
    while (upper - lower > 1) {
        int mid = (upper + lower) >> 1;
        if (asc == (v >= a[mid - 1])) lower = mid; else upper = mid;
    }
    if (v == a[0]) return 1;
    if (v == a[n - 1]) return n - 1;
    return lower;
}

// #{j : lam[j] <= x} (STRICT: lam[j] < x) on the ascending grid of the lookup: the bucket of lnx (ln x to
// within a few ulp) starts a pixel or two away, and a walk with exact FP64 compares on lam finishes it.
// Always within [0, nspec]; a NaN x counts 0.
template <bool STRICT>
__device__ __forceinline__ int lam_count(int nspec, int nbk, double lnl0, double bkinv, const double* __restrict__ lam,
                                         const int* __restrict__ bkt, double x, double lnx) {
    const double u = (lnx - lnl0) * bkinv;
    int b = 0;
    if (u > 0.0) b = u >= (double)(nbk - 1) ? nbk - 1 : (int)u;
    int c = bkt[b];
    // the count is the length of the prefix of entries below x (<= x); entries before 0 count, entries past the grid
    // do not. One batch of four loads around the bucket's start settles the usual case (the count within a pixel of
    // it); otherwise the walks.
    auto in = [&](double v) { return STRICT ? v < x : v <= x; };
    const bool i0 = c - 2 < 0 || in(lam[c - 2]), i1 = c - 1 < 0 || in(lam[c - 1]);
    const bool i2 = c < nspec && in(lam[c]), i3 = c + 1 < nspec && in(lam[c + 1]);
    if (i0 && !i3) return c - 1 + (int)i1 + (int)i2;
    if (STRICT) {
        while (c < nspec && lam[c] < x) c++;
        while (c > 0 && !(lam[c - 1] < x)) c--;
    } else {
        while (c < nspec && lam[c] <= x) c++;
        while (c > 0 && lam[c - 1] > x) c--;
    }
    return c;
}

// exp(x) in FP32 for an FP64 argument: x = k ln 2 + r reduced in FP64, exp(r) in FP32, scaled by
// 2^k. Keeps ~1 ulp of FP32 where rounding x itself to FP32 would cost |x| eps (the reference's
// exp_kp, src/physics/opensps_csp_batch.f90:3886-3919). Flushes to 0 below the FP32 range.
__device__ __forceinline__ float exp_kp(double x) {
    const double LOG2E = 1.4426950408889634, LN2 = 0.6931471805599453;
    x = fmin(1.0e4, fmax(-1.0e4, x));
    double k = rint(x * LOG2E);
    if (k < -149.0) return 0.0f;
    if (k > 127.0) return 3.402823466e38f;
    float r = (float)(x - k * LN2);
    return ldexpf(__expf(r), (int)k);
}

// FP32 -> bf16, round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t bf16_rn(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Split an FP64 value into bf16 hi + lo (x ~ hi + lo to ~2^-16 relative).
__device__ __forceinline__ void split_bf16(double x, uint16_t& hi, uint16_t& lo) {
    hi = bf16_rn((float)x);
    lo = bf16_rn((float)(x - (double)bf16_to_f32(hi)));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Maximum with the first index among equal maxima (the reference's strict '>' scan).
__device__ __forceinline__ void argmax_merge(double& v, int& j, double v2, int j2) {
    if (v2 > v || (v2 == v && j2 < j)) { v = v2; j = j2; }
}

// The six-term IGM optical depth of a pixel: its terms a (igm_tau_pixel_terms) times the (1 + z) powers b6
// (igm_tau_z_terms), src/physics/opensps_cosmology.f90:458-507
__device__ __forceinline__ double igm_tau6(const double* __restrict__ aj, const double* b6) {
    return aj[0] * b6[0] + aj[1] * b6[1] + aj[2] * b6[2] + aj[3] * b6[3] + aj[4] * b6[4] + aj[5] * b6[5];
}
