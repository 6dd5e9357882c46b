// Kernels of the batched composite-population path for gfx950 (MI355X).
//
//   k_prep      one wavefront per model: index selection (it_tesc, zlo), the FP64 SFH weights by the
//               moment method, mass and L_bol, and the model's columns of the age-sum product
//   k_group     one workgroup: the models sorted by metallicity pair, cut into column tiles
//   k_gemm_f64  the FP64 age sums yo = [basis(zlo) basis(zlo+1)] . W per pair on the matrix cores (16x16x4);
//               the split-BF16 ones are k_mm_bf16x3 (ops_mm.hip)
//   k_chain     one workgroup per model: dust, IGM, velocity smoothing, the spectrum out, the
//               observed frame, the band integrals and the magnitudes, the spectrum staged in LDS
#include <limits>
#include <type_traits>

#include "ops_common.h"
#include "ops_launch.h"

// ------------------------------------------------------------------------------------------------
// SFH weights (src/physics/opensps_sfh.f90). FP64 with contraction off, so that each operation
// rounds as the reference's CPU build rounds it.
// ------------------------------------------------------------------------------------------------

#pragma clang fp contract(off)

struct SfhP {
    double tau, tage, tburst, sf_trunc, sf_slope, tq, t0, tb;
    int type;
    int simha;   // use_simha_limits: the window is [t0, tq] (Simha's linear part), else [tq, tage]
};

// The series' term ratios k / (k + 1)^2, k = 0 .. 1000 (correctly rounded at compile time): the series below
// multiplies by them where the reference divides by (k + 1)^2, the same terms to an ulp and no FP64 division
// in the loop (its index is uniform, so the ratio is a scalar operand)
#define OPS_EXPI_KMAX 1000
struct ExpiRatio {
    double v[OPS_EXPI_KMAX + 1];
};
constexpr ExpiRatio make_expi_ratio() {
    ExpiRatio t{};
    for (int k = 0; k <= OPS_EXPI_KMAX; k++) t.v[k] = (double)k / (((double)k + 1.0) * ((double)k + 1.0));
    return t;
}
__constant__ ExpiRatio c_expi = make_expi_ratio();

__device__ double d_expi(double x) {   // opensps_special_functions.f90:50-122
    const double gam = 0.57721566490153286060651209008240243104215933593992;
    if (x == 0.0) return -INFINITY;
    if (x < 0.0) return NAN;
    double res = 1.0, r = 1.0;
    if (fabs(x) <= 40.0) {
        // four terms a block: their ratios come in one scalar load, so the loop waits once per four terms
        bool done = false;
        for (int k0 = 1; k0 <= OPS_EXPI_KMAX && !done; k0 += 4) {
            double c[4];
#pragma unroll
            for (int q = 0; q < 4; q++) c[q] = c_expi.v[k0 + q];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                r = (r * x) * c[q];
                res = res + r;
                if (fabs(r) <= 1.0e-20 * fabs(res)) { done = true; break; }
                if (k0 + q == OPS_EXPI_KMAX) return NAN;
            }
        }
        return gam + log(x) + x * res;
    }
    double inv_x = 1.0 / x;
    for (int k = 1; k <= 20; k++) { r = r * (double)k * inv_x; res = res + r; }
    return exp(x) * inv_x * res;
}

// find_interval_ref on a strictly ascending array by the whole wavefront (uniform control flow, every lane active):
// the number of entries <= v, counted with ballots, is the binary search's result for an ascending array without
// repeats (0 below and for NaN, n above), with the reference's end rules
__device__ __forceinline__ int wave_find_interval(const double* __restrict__ a, int n, double v, int lane) {
    if (n < 2) return 1;
    int cnt = 0;
    for (int i0 = 0; i0 < n; i0 += 64) cnt += __popcll(__ballot(i0 + lane < n && a[i0 + lane] <= v));
    if (v == a[0]) return 1;
    if (v == a[n - 1]) return n - 1;
    return cnt;
}

__device__ void d_convert_sfh(const double* p, SfhP& s) {   // convert_sfhparams_pc (opensps_sfh.f90:1305-1355)
    int sfh = (int)p[OPS_P_SFH];
    double start = (sfh == 1 || sfh == 4 || sfh == 5) ? p[OPS_P_SF_START] * 1.0e9 : 0.0;
    s.tage = p[OPS_P_TAGE] * 1.0e9 - start;
    s.tburst = p[OPS_P_TBURST] * 1.0e9 - start;
    s.sf_trunc = p[OPS_P_SF_TRUNC] * 1.0e9 - start;
    s.tau = p[OPS_P_TAU] * 1.0e9;
    s.sf_slope = -p[OPS_P_SF_SLOPE] / 1.0e9;
    s.tb = s.tage - s.tburst;
    s.tq = (s.sf_trunc <= 0.0 || s.sf_trunc > s.tage) ? 0.0 : s.tage - s.sf_trunc;
    s.t0 = (s.sf_slope > OPS_SAFE_FLOOR) ? s.tq - 1.0 / s.sf_slope : 0.0;
    if (s.t0 > s.tq || s.t0 <= 0.0) s.t0 = 0.0;
    s.type = 0;
    s.simha = 0;
}

// eval_indefinite_moments_log (opensps_sfh.f90:852-910), types 0, 1, 4, 5
__device__ void d_moments(double log_t, const SfhP& s, double& m0, double& m1) {
    const double LOG_E = 0.4342944819032518;   // log10(e)
    double t_lin = pow(10.0, log_t);
    if (s.type == 0) {
        m0 = t_lin;
        m1 = t_lin * (log_t - LOG_E);
    } else if (s.type == 5) {   // Simha's linear part: SFR ~ 1 - slope (tq - t) in lookback time
        const double tp = fmax(0.0, s.tage - s.sf_trunc), sf = 1.0 - s.sf_slope * tp;
        const double q = s.sf_slope * (t_lin * t_lin) / 2.0;
        m0 = sf * t_lin + q;
        m1 = sf * t_lin * (log_t - LOG_E) + q * (log_t - 0.5 * LOG_E);
    } else if (s.type == 1 || s.type == 4) {
        if (s.tau <= OPS_SAFE_FLOOR) { m0 = 0.0; m1 = 0.0; return; }
        double tt = fmax(t_lin / s.tau, OPS_SAFE_FLOOR);
        double ei = d_expi(tt);
        if (!isfinite(ei)) ei = 0.0;
        double et = exp(tt);
        if (s.type == 1) {
            m0 = et;
            m1 = log_t * m0 - LOG_E * ei;
        } else {
            m0 = -(t_lin - s.tage - s.tau) * et;
            double tc = s.tau * LOG_E;
            m1 = -((t_lin - s.tage - s.tau) * log_t - tc) * et - (s.tage + s.tau) * LOG_E * ei;
        }
    } else {
        m0 = 0.0; m1 = 0.0;
    }
}

// compute_ssp_weights (opensps_sfh.f90:97-257) by one wavefront: the moments at the clamped nodes are
// independent, so each lane takes nodes; each node then adds the right part of the segment below and
// the left part of the segment above, in the reference's order. w, m0s, m1s: LDS (nt + 1).
__device__ void wave_ssp_weights(const double* __restrict__ tf, int nt, double tiny_logt, const SfhP& s, int idx_min,
                                 int idx_max, double* w, double* m0s, double* m1s, int lane) {
    for (int i = lane; i < nt; i += 64) w[i] = 0.0;
    WAVE_SYNC();
    if (s.type == -1) {
        if (lane == 0 && s.tb >= 0.0) {
            double log_tb = log10(fmax(s.tb, pow(10.0, tiny_logt)));
            if (log_tb <= tf[0]) {
                w[0] = 1.0;
            } else {
                int j = find_interval_ref(tf, nt, log_tb);
                j = max(1, min(j, nt - 1));
                double bdt = tf[j] - tf[j - 1];
                w[j - 1] = (tf[j] - log_tb) / bdt;
                w[j] = (log_tb - tf[j - 1]) / bdt;
            }
        }
        WAVE_SYNC();
        return;
    }
    double lmin = pow(10.0, tiny_logt);
    double lts = log10(fmax(s.simha ? s.t0 : s.tq, lmin)), lte = log10(fmax(s.simha ? s.tq : s.tage, lmin));
    int js = wave_find_interval(tf, nt, lts, lane);
    int je = wave_find_interval(tf, nt, lte, lane) + 1;
    js = max(js, max(idx_min, 1));
    je = min(je, min(idx_max, nt - 1));
    // node moments M(L(j)), j = js .. je + 1 (1-based), L(j) = clamp(tf(j))
    for (int j = js + lane; j <= je + 1; j += 64) {
        double L = fmin(fmax(tf[j - 1], lts), lte);
        d_moments(L, s, m0s[j], m1s[j]);
    }
    WAVE_SYNC();
    for (int j = js + lane; j <= je + 1; j += 64) {
        double Lj = fmin(fmax(tf[j - 1], lts), lte);
        double acc = 0.0;
        if (j - 1 >= js) {   // segment j - 1: node j is its upper node
            double Lp = fmin(fmax(tf[j - 2], lts), lte);
            if (Lj > Lp) {
                double m0 = m0s[j] - m0s[j - 1], m1 = m1s[j] - m1s[j - 1];
                double nlo = tf[j - 2], nhi = tf[j - 1];
                double inv = 1.0 / (nhi - nlo);
                acc = acc + (m1 - nlo * m0) * inv;
            }
        }
        if (j <= je) {       // segment j: node j is its lower node
            double Ln = fmin(fmax(tf[j], lts), lte);
            if (Ln > Lj) {
                double m0 = m0s[j + 1] - m0s[j], m1 = m1s[j + 1] - m1s[j];
                double nlo = tf[j - 1], nhi = tf[j];
                double inv = 1.0 / (nhi - nlo);
                acc = acc + (nhi * m0 - m1) * inv;
            }
        }
        w[j - 1] = acc;
    }
    WAVE_SYNC();
    if (idx_min == 0 && lane == 0) {
        double l1 = fmin(fmax(tiny_logt, lts), lte), l2 = fmin(fmax(tf[0], lts), lte);
        if (l1 < l2) {
            double a0, a1, b0, b1;
            d_moments(l2, s, a0, a1);
            d_moments(l1, s, b0, b1);
            w[0] = w[0] + (a0 - b0);
        }
    }
    WAVE_SYNC();
}

// expm1 as the reference forms it (opensps_special_functions.f90:219-233) and gammainc(2, x) (:142-167)
__device__ double d_expm1_ref(double x) {
    if (fabs(x) < 1.0e-5) return x * (1.0 + x * (0.5 + x * (1.0 / 6.0)));
    return exp(x) - 1.0;
}
__device__ double d_gammainc2(double x) { return x < 0.0 ? 0.0 : -d_expm1_ref(-x) - x * exp(-x); }

// The Simha SFH's linear mass fraction (get_sfh_properties_at_age_pc's frac_linear for sfh = 5,
// opensps_sfh.f90:327-455), Gyr; log_tmax: the oldest grid age (log10 yr)
__device__ double d_simha_frac(const double* p, double log_tmax) {
    const double sf_start = p[OPS_P_SF_START], sf_trunc = p[OPS_P_SF_TRUNC];
    const double t_max = pow(10.0, log_tmax - 9.0) - sf_start, t_prime = p[OPS_P_TAGE] - sf_start;
    const double t_trunc = (sf_trunc < OPS_SAFE_FLOOR || sf_trunc < sf_start) ? t_max : sf_trunc - sf_start;
    const double tau = fmax(p[OPS_P_TAU], OPS_SAFE_FLOOR);
    const double mass_tau = t_prime >= 0.0 ? tau * d_gammainc2(fmin(t_prime, t_trunc) / tau) : 0.0;
    const double slope_m = -p[OPS_P_SF_SLOPE];
    const double t_zero = slope_m > 0.0 ? t_trunc + 1.0 / slope_m : t_max;
    double mass_linear = 0.0;
    if (!(t_trunc <= 0.0 || t_trunc > t_max) && t_prime > t_trunc) {
        const double sfr_trunc = (t_trunc / p[OPS_P_TAU]) * exp(-t_trunc / p[OPS_P_TAU]);
        const double t_hi = fmin(t_zero, t_prime);
        mass_linear = sfr_trunc * ((t_hi - t_trunc) - (slope_m / 2.0) * (t_hi * t_hi + t_trunc * t_trunc) +
                                   slope_m * t_hi * t_trunc);
    }
    return (t_prime >= 0.0 && mass_linear + mass_tau > OPS_SAFE_FLOOR) ? mass_linear / (mass_linear + mass_tau) : 0.0;
}

// The tabular SFH (sfh_weights_core_pc, opensps_sfh.f90:1219-1292, one metallicity as the batch calls it) by one
// wavefront: each bin of the context's table (forward ages tab[0..ntab), years; SFRs tab[ntab..2 ntab)) a linear SFR
// segment over its lookback interval (the type-5 moments with the bin's slope), its hat weights normalised and
// scaled by the segment's mass. The bins go one after another (the weights of neighbouring bins share nodes).
__device__ void wave_tab_weights(const double* __restrict__ tf, int nt, double tiny_logt, const double* __restrict__ tab,
                                 int ntab, double tage, double* w, double* t1, double* m0s, double* m1s, int lane) {
    for (int i = lane; i < nt; i += 64) w[i] = 0.0;
    WAVE_SYNC();
    const double lmin = pow(10.0, tiny_logt), lmax = pow(10.0, tf[nt - 1]);
    const double* age = tab;
    const double* sfr = tab + ntab;
    for (int j = 0; j < ntab - 1; j++) {
        const double ta = tage * 1.0e9 - age[j + 1], tb = tage * 1.0e9 - age[j];
        if (tb < 0.0) continue;
        const double dt = fmax(tb - ta, OPS_SAFE_FLOOR);
        SfhP s;
        s.tau = 0.0; s.tburst = 0.0; s.tb = 0.0; s.t0 = 0.0; s.simha = 0;
        s.type = 5;
        s.sf_slope = -(sfr[j + 1] - sfr[j]) / dt / fmax(sfr[j + 1], OPS_SAFE_FLOOR);
        s.tq = fmin(fmax(ta, lmin), lmax);
        s.tage = fmin(fmax(tb, lmin), lmax);
        s.sf_trunc = s.tage - s.tq;
        const double mass2 = sfr[j + 1] * (1.0 + s.sf_slope / 2.0 * (s.tage + s.tq - 2.0 * ta)) * (s.tage - s.tq);
        const int imin = min(max(wave_find_interval(tf, nt, log10(fmax(ta, OPS_SAFE_FLOOR)), lane) - 1, 0), nt);
        const int imax = min(max(wave_find_interval(tf, nt, log10(fmax(tb, OPS_SAFE_FLOOR)), lane) + 2, 0), nt);
        wave_ssp_weights(tf, nt, tiny_logt, s, imin, imax, t1, m0s, m1s, lane);
        double mass1 = 0.0;
        for (int i = lane; i < nt; i += 64) mass1 += t1[i];
        mass1 = wave_sum(mass1);
        if (mass1 < OPS_SAFE_FLOOR) mass1 = 1.0;
        for (int i = lane; i < nt; i += 64) w[i] = w[i] + t1[i] * (mass2 / mass1);
        WAVE_SYNC();
    }
}

// sfh_weights_core_pc (opensps_sfh.f90:1077-1294) for sfh 0, 1, 4, 5 and 2, 3 (the context's table). Returns 0, or 2
// (unsupported).
__device__ int wave_sfh_weights(const double* __restrict__ tf, int nt, double tiny_logt, const double* p, double* w,
                                double* t1, double* t2, double* m0s, double* m1s, const double* tab, int ntab,
                                int lane) {
    SfhP s;
    int sfh = (int)p[OPS_P_SFH];
    if (sfh == 2 || sfh == 3) {
        wave_tab_weights(tf, nt, tiny_logt, tab, ntab, p[OPS_P_TAGE], w, t1, m0s, m1s, lane);
        return 0;
    }
    d_convert_sfh(p, s);
    int imax = wave_find_interval(tf, nt, log10(s.tage), lane) + 2;
    imax = min(max(imax, 1), nt);
    if (sfh == 0) {
        s.type = -1;
        s.tb = s.tage;
        wave_ssp_weights(tf, nt, tiny_logt, s, max(imax - 2, 1), imax, w, m0s, m1s, lane);
        return 0;
    }
    if (sfh == 5) {   // Simha: the delayed-tau part over [tq, tage], the linear part over [t0, tq], mixed
        s.type = 4;
        wave_ssp_weights(tf, nt, tiny_logt, s, 0, imax, t1, m0s, m1s, lane);
        double m1 = 0.0, m2 = 0.0;
        for (int i = 0; i < imax; i++) m1 = m1 + t1[i];
        s.type = 5;
        s.simha = 1;
        wave_ssp_weights(tf, nt, tiny_logt, s, 0, imax, t2, m0s, m1s, lane);
        for (int i = 0; i < imax; i++) m2 = m2 + t2[i];
        if (m1 < OPS_SAFE_FLOOR) m1 = 1.0;
        if (m2 < OPS_SAFE_FLOOR) m2 = 1.0;
        const double fl = d_simha_frac(p, tf[nt - 1]);
        for (int i = lane; i < nt; i += 64) w[i] = (t1[i] / m1) * (1.0 - fl) + (t2[i] / m2) * fl;
        WAVE_SYNC();
        return 0;
    }
    if (sfh != 1 && sfh != 4) return 2;
    s.type = sfh;
    wave_ssp_weights(tf, nt, tiny_logt, s, 0, imax, w, m0s, m1s, lane);
    double mass1 = 0.0;
    for (int i = 0; i < imax; i++) mass1 = mass1 + w[i];   // (every lane: the reference's sequential sum)
    if (mass1 < OPS_SAFE_FLOOR) mass1 = 1.0;
    WAVE_SYNC();
    for (int i = lane; i < nt; i += 64) w[i] = w[i] / mass1;
    WAVE_SYNC();
    double cst = p[OPS_P_CONST], fb = p[OPS_P_FBURST];
    if (cst > 0.0 || fb > OPS_SAFE_FLOOR) {
        s.type = 0;
        wave_ssp_weights(tf, nt, tiny_logt, s, 0, imax, t1, m0s, m1s, lane);
        double mc = 0.0;
        for (int i = 0; i < imax; i++) mc = mc + t1[i];
        if (mc < OPS_SAFE_FLOOR) mc = 1.0;
        double fbv = 0.0;
        if (s.tb >= 0.0) {
            s.type = -1;
            wave_ssp_weights(tf, nt, tiny_logt, s, 0, imax, t2, m0s, m1s, lane);
            fbv = fb;
        } else {
            for (int i = lane; i < nt; i += 64) t2[i] = 0.0;
            WAVE_SYNC();
        }
        for (int i = lane; i < nt; i += 64) w[i] = (1.0 - cst - fbv) * w[i] + cst * (t1[i] / mc) + fbv * t2[i];
        WAVE_SYNC();
    }
    return 0;
}

// ------------------------------------------------------------------------------------------------
// SFH tangents (src/physics/opensps_sfh_tangent.f90) reduced to params_bar's SFH rows (batch_sfh_gradients,
// src/physics/opensps_csp_batch.f90:5198-5217: row = sum over ages of d L / d w_k  d w_k / d theta).
// The weights' tangent is linear in the tangents of the internal times the parameters move (tage, tau,
// tq = tage - sf_trunc and tb = tage - tburst, in years), so one pass takes the derivative along each of them
// (per year) and the eight rows are combinations of their sums S_e = sum_k wbar_k d w_k / d e (G = 1e9 yr / Gyr):
//   tage = G (S_tage + S_tb + [tq > 0] S_tq)   tau = G S_tau       tburst = -G S_tb      sf_start = -G S_tage
//   sf_trunc = -G [tq > 0] S_tq               const, fburst: the mixture's own terms    sf_slope = 0
// (the slope acts on Simha's linear part only). sfh = 0 is a burst at tb = tage: tage's row only. Simha (sfh = 5)
// mixes a delayed-tau part (along tage, tau, tq) and a linear part over [t0, tq] (along tq and its slope) by the
// linear part's mass fraction f, whose derivatives (simha_frac_t) add (S_lin - S_tau) df / d theta.
// ------------------------------------------------------------------------------------------------

// moments_t (opensps_sfh_tangent.f90:426-467), logarithmic branch, types 0, 1, 4, 5, along ND directions: the
// moments at log time L and their derivatives for the time's tangents dL[e] and the parameters' dtg[e], dtau[e]
// (tage, tau), dtp[e] (Simha's t' = tage - sf_trunc) and dsl[e] (its slope)
template <int ND>
__device__ void d_moments_t(double L, const double* dL, const double* dtg, const double* dtau, const double* dtp,
                            const double* dsl, const SfhP& s, double& m0, double& m1, double* dm0, double* dm1) {
    const double LOG_E = 0.4342944819032518, LN10 = 2.302585092994046;
    const double t = pow(10.0, L);
    m0 = 0.0;
    m1 = 0.0;
#pragma unroll
    for (int e = 0; e < ND; e++) { dm0[e] = 0.0; dm1[e] = 0.0; }
    if (s.type == 0) {
        m0 = t;
        m1 = t * (L - LOG_E);
#pragma unroll
        for (int e = 0; e < ND; e++) {
            const double dt = t * LN10 * dL[e];
            dm0[e] = dt;
            dm1[e] = dt * (L - LOG_E) + t * dL[e];
        }
        return;
    }
    if (s.type == 5) {   // Simha's linear part: slope factor 1 - s t', t' = max(0, tage - sf_trunc)
        const double sl = s.sf_slope, tp = fmax(0.0, s.tage - s.sf_trunc), sf = 1.0 - sl * tp;
        const double h = sl * (t * t) / 2.0;
        m0 = sf * t + h;
        m1 = sf * t * (L - LOG_E) + h * (L - 0.5 * LOG_E);
#pragma unroll
        for (int e = 0; e < ND; e++) {
            const double dt = t * LN10 * dL[e];
            const double dsf = -dsl[e] * tp - sl * dtp[e], dh = dsl[e] * (t * t) / 2.0 + sl * t * dt;
            dm0[e] = dsf * t + sf * dt + dh;
            dm1[e] = dsf * t * (L - LOG_E) + sf * (dt * (L - LOG_E) + t * dL[e]) + dh * (L - 0.5 * LOG_E) + h * dL[e];
        }
        return;
    }
    if ((s.type != 1 && s.type != 4) || s.tau <= OPS_SAFE_FLOOR) return;
    const double tau = s.tau, tg = s.tage;
    double x = t / tau;
    const bool floored = x < OPS_SAFE_FLOOR;
    if (floored) x = OPS_SAFE_FLOOR;
    const double ex = exp(x);
    double ei = d_expi(x);
    const bool fin = isfinite(ei);
    if (!fin) ei = 0.0;
    const double a = t - tg - tau, b = a * L - tau * LOG_E;
    if (s.type == 1) {
        m0 = ex;
        m1 = L * ex - LOG_E * ei;
    } else {
        m0 = -a * ex;
        m1 = -b * ex - (tg + tau) * LOG_E * ei;
    }
#pragma unroll
    for (int e = 0; e < ND; e++) {
        const double dt = t * LN10 * dL[e];
        const double dx = floored ? 0.0 : dt / tau - t * dtau[e] / (tau * tau);
        const double dei = fin ? ex / x * dx : 0.0;
        if (s.type == 1) {
            dm0[e] = ex * dx;
            dm1[e] = dL[e] * ex + L * dm0[e] - LOG_E * dei;
        } else {
            const double da = dt - dtg[e] - dtau[e];
            dm0[e] = -(da * ex + a * ex * dx);
            dm1[e] = -((da * L + a * dL[e] - dtau[e] * LOG_E) * ex + b * ex * dx) - (dtg[e] + dtau[e]) * LOG_E * ei -
                     (tg + tau) * LOG_E * dei;
        }
    }
}

// One SFH component's weights and tangents over its window reduced against wbar.
struct SfhTanSums {
    double mass, B;           // sum_{k < imax} w_k; sum_k wbar_k w_k
    double dmass[3], A[3];    // per direction: sum_{k < imax} dw_k; sum_k wbar_k dw_k
};

// ssp_weights_t (opensps_sfh_tangent.f90:255-358) with idx_min = 0, by one wavefront. The window [tq, tage] with the
// directions tage, then tau (ND = 3 only), then tq (last); Simha's linear part (s.simha, ND = 2): the window [t0, tq]
// with the directions tq (t' = tage - sf_trunc and t0 = tq - 1 / slope move with it) and the internal slope. The
// moments at the clamped nodes are independent (one node a lane); node j then takes the upper end of segment j - 1
// and the lower end of segment j. Zero-width segments stay in (no weight; the tangent is the one-sided one where a
// window end sits on a grid age). mm: LDS, (2 + 2 ND) (nt + 2) doubles: node j (1-based) at slot j, the young
// edge's limits at slots 0, nt + 1.
template <int ND>
__device__ void wave_weights_t(const double* __restrict__ tf, int nt, double tiny_logt, const SfhP& s, int imax,
                               const double* __restrict__ wbar, double* mm, int lane, SfhTanSums& r) {
    const double LN10 = 2.302585092994046;
    const int S = nt + 2;
    double *M0 = mm, *M1 = mm + S, *D0 = mm + 2 * S, *D1 = mm + (2 + ND) * S;
    const double lmin = pow(10.0, tiny_logt);
    const double ts = s.simha ? s.t0 : s.tq, te = s.simha ? s.tq : s.tage;
    const double lts = log10(fmax(ts, lmin)), lte = log10(fmax(te, lmin));
    const double uts = ts > lmin ? 1.0 / (ts * LN10) : 0.0;   // d lts / d ts
    const double ute = te > lmin ? 1.0 / (te * LN10) : 0.0;   // d lte / d te
    const int js = max(wave_find_interval(tf, nt, lts, lane), 1);
    const int je = min(wave_find_interval(tf, nt, lte, lane) + 1, min(imax, nt - 1));
    const int nn = max(je + 2 - js, 0);   // nodes js .. je + 1
    // per direction: the window ends' tangents and the parameters' (tage, tau, t', slope)
    double dts[ND], dte[ND], dtg[ND], dtau[ND], dtp[ND], dsl[ND];
#pragma unroll
    for (int e = 0; e < ND; e++) { dts[e] = 0.0; dte[e] = 0.0; dtg[e] = 0.0; dtau[e] = 0.0; dtp[e] = 0.0; dsl[e] = 0.0; }
    if (s.simha) {
        dte[0] = 1.0;                                    // tq: the window's young end, t' and t0 with it
        dts[0] = s.t0 > 0.0 ? 1.0 : 0.0;
        dtp[0] = s.tage - s.sf_trunc > 0.0 ? 1.0 : 0.0;
        if (ND > 1) {                                    // the slope: t0 = tq - 1 / slope
            dts[ND - 1] = s.t0 > 0.0 ? 1.0 / (s.sf_slope * s.sf_slope) : 0.0;
            dsl[ND - 1] = 1.0;
        }
    } else {
        dte[0] = 1.0;   // tage
        dtg[0] = 1.0;
        if (ND == 3) dtau[1] = 1.0;
        dts[ND - 1] = 1.0;   // tq
    }
    auto clampL = [&](double node, int& fl) {   // clamp_t (opensps_sfh_tangent.f90:407-422)
        double v = node;
        fl = 0;
        if (lts > node) { v = lts; fl = 1; }
        if (lte < v) { v = lte; fl = 2; }
        return v;
    };
    WAVE_SYNC();
    for (int i = lane; i < nn + 2; i += 64) {
        int slot;
        double node;
        if (i < nn) { slot = js + i; node = tf[slot - 1]; }
        else if (i == nn) { slot = 0; node = tiny_logt; }
        else { slot = nt + 1; node = tf[0]; }
        int fl;
        const double L = clampL(node, fl);
        double dL[ND];
#pragma unroll
        for (int e = 0; e < ND; e++) dL[e] = fl == 1 ? dts[e] * uts : (fl == 2 ? dte[e] * ute : 0.0);
        double m0, m1, dm0[ND], dm1[ND];
        d_moments_t<ND>(L, dL, dtg, dtau, dtp, dsl, s, m0, m1, dm0, dm1);
        M0[slot] = m0;
        M1[slot] = m1;
#pragma unroll
        for (int e = 0; e < ND; e++) {
            D0[e * S + slot] = dm0[e];
            D1[e * S + slot] = dm1[e];
        }
    }
    WAVE_SYNC();
    double mass = 0.0, B = 0.0, dmass[ND], A[ND];
#pragma unroll
    for (int e = 0; e < ND; e++) { dmass[e] = 0.0; A[e] = 0.0; }
    for (int j = js + lane; j <= je + 1; j += 64) {
        double w = 0.0, dw[ND];
#pragma unroll
        for (int e = 0; e < ND; e++) dw[e] = 0.0;
        if (j - 1 >= js) {   // segment j - 1 (ages j - 1, j): node j is its upper end
            const double nlo = tf[j - 2], inv = 1.0 / (tf[j - 1] - nlo);
            w += ((M1[j] - M1[j - 1]) - nlo * (M0[j] - M0[j - 1])) * inv;
#pragma unroll
            for (int e = 0; e < ND; e++)
                dw[e] += ((D1[e * S + j] - D1[e * S + j - 1]) - nlo * (D0[e * S + j] - D0[e * S + j - 1])) * inv;
        }
        if (j <= je) {       // segment j (ages j, j + 1): node j is its lower end
            const double nhi = tf[j], inv = 1.0 / (nhi - tf[j - 1]);
            w += (nhi * (M0[j + 1] - M0[j]) - (M1[j + 1] - M1[j])) * inv;
#pragma unroll
            for (int e = 0; e < ND; e++)
                dw[e] += (nhi * (D0[e * S + j + 1] - D0[e * S + j]) - (D1[e * S + j + 1] - D1[e * S + j])) * inv;
        }
        const double wb = wbar[j - 1];
        if (j - 1 < imax) {
            mass += w;
#pragma unroll
            for (int e = 0; e < ND; e++) dmass[e] += dw[e];
        }
        B += wb * w;
#pragma unroll
        for (int e = 0; e < ND; e++) A[e] += wb * dw[e];
    }
    if (lane == 0) {   // the youngest edge: the mass formed before the first grid age goes to it
        int f1, f2;
        if (clampL(tiny_logt, f1) < clampL(tf[0], f2)) {
            const double w = M0[nt + 1] - M0[0], wb = wbar[0];
            mass += w;
            B += wb * w;
#pragma unroll
            for (int e = 0; e < ND; e++) {
                const double dw = D0[e * S + nt + 1] - D0[e * S];
                dmass[e] += dw;
                A[e] += wb * dw;
            }
        }
    }
    r.mass = wave_sum(mass);
    r.B = wave_sum(B);
#pragma unroll
    for (int e = 0; e < 3; e++) { r.dmass[e] = 0.0; r.A[e] = 0.0; }
#pragma unroll
    for (int e = 0; e < ND; e++) {
        r.dmass[e] = wave_sum(dmass[e]);
        r.A[e] = wave_sum(A[e]);
    }
    WAVE_SYNC();
}

// simha_frac_t (opensps_sfh_tangent.f90:362-403): Simha's linear mass fraction and its derivatives (Gyr; per unit of
// sf_slope) with respect to tage, tau, sf_start, sf_trunc, sf_slope (df[0..4])
__device__ double d_simha_frac_t(const double* p, double log_tmax, double* df) {
    const double sf_start = p[OPS_P_SF_START], sf_trunc = p[OPS_P_SF_TRUNC], tau_p = p[OPS_P_TAU];
    const double tmax = pow(10.0, log_tmax - 9.0) - sf_start, tpr = p[OPS_P_TAGE] - sf_start;
    const bool trunc_on = !(sf_trunc < OPS_SAFE_FLOOR || sf_trunc < sf_start);
    const double ttr = trunc_on ? sf_trunc - sf_start : tmax, tau = fmax(tau_p, OPS_SAFE_FLOOR);
    const double sm = -p[OPS_P_SF_SLOPE];
#pragma unroll
    for (int q = 0; q < 5; q++) df[q] = 0.0;
    if (tpr < 0.0 || ttr <= 0.0 || ttr > tmax || tpr <= ttr) return 0.0;
    const double x = ttr / tau, gx = d_gammainc2(x), mt = tau * gx;
    const double tz = sm > 0.0 ? ttr + 1.0 / sm : tmax;
    const bool zhi = tz <= tpr;
    const double th = zhi ? tz : tpr;
    const double r = ttr / tau_p, er = exp(-r), sfr = r * er;
    const double bb = (th - ttr) - (sm / 2.0) * (th * th + ttr * ttr) + sm * th * ttr;
    const double ml = sfr * bb;
    if (!(ml + mt > OPS_SAFE_FLOOR)) return 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) {   // seeds: tage, tau, sf_start, sf_trunc, sf_slope
        double dtmax = 0.0, dtpr = 0.0, dttr = 0.0, dtau = 0.0, dsm = 0.0;
        if (q == 0) dtpr = 1.0;
        if (q == 1 && tau_p > OPS_SAFE_FLOOR) dtau = 1.0;
        if (q == 2) { dtmax = -1.0; dtpr = -1.0; dttr = -1.0; }
        if (q == 3 && trunc_on) dttr = 1.0;
        if (q == 4) dsm = -1.0;
        if (!trunc_on) dttr = dtmax;
        const double dx = (dttr - x * dtau) / tau;   // (tpr > ttr: the tau part's mass ends at the truncation)
        const double dmt = dtau * gx + tau * x * exp(-x) * dx;
        const double dtz = sm > 0.0 ? dttr - dsm / (sm * sm) : dtmax;
        const double dth = zhi ? dtz : dtpr;
        const double dr = (dttr - r * dtau) / tau_p, dsfr = er * (1.0 - r) * dr;
        const double dbb = (dth - dttr) - (dsm / 2.0) * (th * th + ttr * ttr) - sm * (th * dth + ttr * dttr) +
                           dsm * th * ttr + sm * (dth * ttr + th * dttr);
        const double dml = dsfr * bb + sfr * dbb;
        df[q] = (dml * mt - ml * dmt) / ((ml + mt) * (ml + mt));
    }
    return ml / (ml + mt);
}

// The burst (ssp_weights_t, type -1): two weights linear in log tb between the bracketing ages. B = sum wbar w,
// A = sum wbar dw / d tb (per year).
__device__ void d_burst_t(const double* __restrict__ tf, int nt, double tiny_logt, double tb,
                          const double* __restrict__ wbar, double& B, double& A) {
    B = 0.0;
    A = 0.0;
    if (tb < 0.0) return;
    const double lmin = pow(10.0, tiny_logt), ltb = log10(fmax(tb, lmin));
    const double u = tb > lmin ? 1.0 / (tb * 2.302585092994046) : 0.0;
    if (ltb <= tf[0]) { B = wbar[0]; return; }
    const int j = max(1, min(find_interval_ref(tf, nt, ltb), nt - 1));
    const double bdt = tf[j] - tf[j - 1];
    B = wbar[j - 1] * ((tf[j] - ltb) / bdt) + wbar[j] * ((ltb - tf[j - 1]) / bdt);
    A = (wbar[j] - wbar[j - 1]) * (u / bdt);
}

// k_sfh_grad: one wavefront per model (4 per workgroup); sfh_weights_tangent (opensps_sfh_tangent.f90:66-251) for
// sfh 0, 1, 4, 5 along the internal times at once, written to params_bar's rows 16 .. 21, 24, 25 (0-based). Other
// SFH types (the forward has flagged them) get zeros.
__global__ __launch_bounds__(256) void k_sfh_grad(SfhGradArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = blockIdx.x * 4 + wv;
    if (m >= a.n) return;
    const int nt = a.nt;
    double* mm = lds + (size_t)wv * 8 * (nt + 2);
    double p[OPS_NPAR];
#pragma unroll
    for (int q = 0; q < OPS_NPAR; q++) p[q] = a.prm[(size_t)q * a.n + m];
    const double* wb = a.wbar + (size_t)m * nt;
    const double G = 1.0e9;
    double rows[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // tage tau const fburst tburst sf_start sf_trunc slope
    const int sfh = (int)p[OPS_P_SFH];
    if (sfh == 0 || sfh == 1 || sfh == 4 || sfh == 5) {
        SfhP s;
        d_convert_sfh(p, s);
        const int imax = min(max(wave_find_interval(a.time, nt, log10(s.tage), lane) + 2, 1), nt);
        if (sfh == 5) {
            // Simha: (1 - f) wt + f cn, the delayed-tau part wt along tage, tau, tq, the linear part cn along tq and
            // the slope, f the linear mass fraction (its derivatives from simha_frac_t)
            SfhTanSums T, Ln;
            s.type = 4;
            wave_weights_t<3>(a.time, nt, a.tiny_logt, s, imax, wb, mm, lane, T);
            s.type = 5;
            s.simha = 1;
            wave_weights_t<2>(a.time, nt, a.tiny_logt, s, imax, wb, mm, lane, Ln);
            double m1 = T.mass, dm1[3] = {T.dmass[0], T.dmass[1], T.dmass[2]};
            if (m1 < OPS_SAFE_FLOOR) { m1 = 1.0; dm1[0] = dm1[1] = dm1[2] = 0.0; }
            double m2 = Ln.mass, dm2[2] = {Ln.dmass[0], Ln.dmass[1]};
            if (m2 < OPS_SAFE_FLOOR) { m2 = 1.0; dm2[0] = dm2[1] = 0.0; }
            double dfl[5];
            const double fl = d_simha_frac_t(p, a.time[nt - 1], dfl);
            const double bt = T.B / m1, bl = Ln.B / m2, mix = bl - bt;   // sum wbar (cn - wt)
            const double t_tg = (T.A[0] - bt * dm1[0]) / m1, t_tau = (T.A[1] - bt * dm1[1]) / m1;
            const double t_tq = (T.A[2] - bt * dm1[2]) / m1;
            const double l_tq = (Ln.A[0] - bl * dm2[0]) / m2, l_sl = (Ln.A[1] - bl * dm2[1]) / m2;
            const bool tq_on = s.tq > 0.0;
            const double s_tq = (1.0 - fl) * t_tq + fl * l_tq;
            rows[0] = G * ((1.0 - fl) * t_tg + (tq_on ? s_tq : 0.0)) + mix * dfl[0];
            rows[1] = G * (1.0 - fl) * t_tau + mix * dfl[1];
            rows[5] = -G * (1.0 - fl) * t_tg + mix * dfl[2];
            rows[6] = (tq_on ? -G * s_tq : 0.0) + mix * dfl[3];
            rows[7] = -(1.0 / G) * fl * l_sl + mix * dfl[4];
        } else if (sfh == 0) {
            double B, A;
            d_burst_t(a.time, nt, a.tiny_logt, s.tage, wb, B, A);
            rows[0] = G * A;
        } else {
            SfhTanSums T, Cn;
            s.type = sfh;
            wave_weights_t<3>(a.time, nt, a.tiny_logt, s, imax, wb, mm, lane, T);
            s.type = 0;
            wave_weights_t<2>(a.time, nt, a.tiny_logt, s, imax, wb, mm, lane, Cn);
            double mass1 = T.mass, dm1[3] = {T.dmass[0], T.dmass[1], T.dmass[2]};
            if (mass1 < OPS_SAFE_FLOOR) { mass1 = 1.0; dm1[0] = dm1[1] = dm1[2] = 0.0; }
            double massc = Cn.mass, dmc[2] = {Cn.dmass[0], Cn.dmass[1]};
            if (massc < OPS_SAFE_FLOOR) { massc = 1.0; dmc[0] = dmc[1] = 0.0; }
            const bool burst = s.tb >= 0.0;
            double BB = 0.0, AB = 0.0;
            if (burst) d_burst_t(a.time, nt, a.tiny_logt, s.tb, wb, BB, AB);
            double c = p[OPS_P_CONST], fb = burst ? p[OPS_P_FBURST] : 0.0;
            if (!(p[OPS_P_CONST] > 0.0 || p[OPS_P_FBURST] > OPS_SAFE_FLOOR)) { c = 0.0; fb = 0.0; }
            const double bt = T.B / mass1, bc = Cn.B / massc;   // sum wbar wt, sum wbar cn (normalized parts)
            // sum wbar d(w / mass) = (A - (B / mass) d mass) / mass
            const double k = 1.0 - c - fb;
            const double s_tg = k * ((T.A[0] - bt * dm1[0]) / mass1) + c * ((Cn.A[0] - bc * dmc[0]) / massc);
            const double s_tau = k * ((T.A[1] - bt * dm1[1]) / mass1);
            const double s_tq = k * ((T.A[2] - bt * dm1[2]) / mass1) + c * ((Cn.A[1] - bc * dmc[1]) / massc);
            const double s_tb = fb * AB;
            const bool tq_on = s.tq > 0.0;
            rows[0] = G * (s_tg + s_tb + (tq_on ? s_tq : 0.0));
            rows[1] = G * s_tau;
            rows[2] = bc - bt;
            rows[3] = burst ? BB - bt : 0.0;
            rows[4] = -G * s_tb;
            rows[5] = -G * s_tg;
            rows[6] = tq_on ? -G * s_tq : 0.0;
        }
    }
    if (lane == 0) {
        double* pb = a.params_bar + (size_t)m * OPENSPS_HIP_NBGRAD_ROWS;
        const int rr[8] = {16, 17, 18, 19, 20, 21, 24, 25};
#pragma unroll
        for (int q = 0; q < 8; q++) pb[rr[q]] = rows[q];
    }
}

// The Vega system (magnitudes_from_fluxes, src/spectra/opensps_batch_photometry.f90:185-232), a thread per
// (model, band) of mags (nbands, n): vref = 0: AB less Vega's AB magnitude in every band; vref > 0: band vref stays
// AB and the others are AB - magvega(b) + magvega(vref), NaN where the model's vref magnitude is NaN. Band vref is
// only read, so no thread reads a value another writes.
__global__ void k_vega_mags(int n, int nb, double* mags, const double* magvega, int vref, double anchor) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * nb) return;
    const int m = (int)(i / nb), b = (int)(i - (size_t)m * nb);
    if (b == vref - 1) return;
    double v = mags[i];
    if (vref > 0) {
        const double vv = mags[(size_t)m * nb + vref - 1];
        v = vv == vv ? v - magvega[b] + anchor : (double)NAN;
    } else {
        v = v - magvega[b];
    }
    mags[i] = v;
}

// ------------------------------------------------------------------------------------------------
// k_prep: one wavefront per model (4 per workgroup)
// ------------------------------------------------------------------------------------------------
__device__ void wave_chain_header(const PrepArgs& a, int m, const double* p, double dz, double ls, int lane);

// four wavefronts a SIMD (at most 128 VGPRs: a few spill, and the occupancy more than repays them; 0.199 -> 0.189 ms
// on the 16 384-model batch)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_prep(PrepArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = blockIdx.x * 4 + wv;
    if (m >= a.n) return;
    const int nt = a.nt, n = a.n;
    double* w = lds + (size_t)wv * 5 * (nt + 1);
    double* t1 = w + (nt + 1);
    double* t2 = t1 + (nt + 1);
    double* m0s = t2 + (nt + 1);
    double* m1s = m0s + (nt + 1);
    double p[OPS_NPAR];
#pragma unroll
    for (int q = 0; q < OPS_NPAR; q++) p[q] = a.prm[(size_t)q * n + m];

    int st = 0;
    if (!a.weights_in && !a.wz && p[OPS_P_TAGE] <= 0.0) st = 2;                // batch_check_models :1312
    if (!a.smooth_fast && p[OPS_P_SIGMA_SMOOTH] > OPS_SAFE_FLOOR) st = 2;     // host smoothing modes: not built
    // split_parts = 0 folds young and old into one sum: only without birth-cloud and old-star screens
    bool needs_parts = (p[OPS_P_DUST1] > OPS_SAFE_FLOOR && 1.0 - p[OPS_P_FRAC_OBRUN] > OPS_SAFE_FLOOR) ||
                       fabs(p[OPS_P_DUST3]) > OPS_SAFE_FLOOR;
    if (needs_parts && a.parts == 1) st = 4;
    if (st == 0 && a.wz) {
        for (int k = lane; k < nt; k += 64) w[k] = 0.0;   // (no SFH weights: the weights run over metallicities)
        WAVE_SYNC();
    } else if (st == 0) {
        if (a.weights_in) {
            for (int k = lane; k < nt; k += 64) w[k] = a.weights_in[(size_t)m * nt + k];
            WAVE_SYNC();
        } else {
            st = wave_sfh_weights(a.time, nt, a.tiny_logt, p, w, t1, t2, m0s, m1s, a.tab, a.ntab, lane);
        }
    }
    if (st != 0) {
        if (lane == 0) {
            atomicMax(a.flag, st);
            if (a.status) a.status[m] = st;
        }
        for (int k = lane; k < nt; k += 64) w[k] = 0.0;
        WAVE_SYNC();
    } else if (lane == 0 && a.status) {
        a.status[m] = 0;
    }

    // stage 1 indices: batch_model_indices (opensps_csp_batch.f90:1541-1576), metallicity_interval (:2073-2084)
    double dal = p[OPS_P_DUST_TESC] > OPS_SAFE_FLOOR ? p[OPS_P_DUST_TESC] : 7.0;
    int it = max(1, min(wave_find_interval(a.time, nt, dal, lane), nt));
    int zlo = 1;
    double dz = 0.0;
    if (a.nz > 1 && !a.wz) {
        zlo = max(1, min(wave_find_interval(a.logz, a.nz, p[OPS_P_LOGZSOL], lane), a.nz - 1));
        dz = (p[OPS_P_LOGZSOL] - a.logz[zlo - 1]) / (a.logz[zlo] - a.logz[zlo - 1]);
    }

    // model_sums (:2129-2176): mass and L_bol over the ages (terms per lane, then a wave sum); with the operator
    // they come from the per-age sums of the IMF weights (k_op_sums)
    double ms = 0.0, ls = 0.0;
    for (int k = a.ops ? nt : lane; k < nt; k += 64) {
        double wk = w[k];
        if (wk > OPS_SAFE_FLOOR) {
            double mk, lk;
            if (a.nz > 1) {
                mk = (1.0 - dz) * a.bmass[(size_t)(zlo - 1) * nt + k] + dz * a.bmass[(size_t)zlo * nt + k];
                lk = (1.0 - dz) * a.blbol[(size_t)(zlo - 1) * nt + k] + dz * a.blbol[(size_t)zlo * nt + k];
                mk = wk * mk;
                lk = wk * pow(10.0, lk);
            } else {
                mk = wk * a.bmass[k];
                lk = wk * pow(10.0, a.blbol[k]);
            }
            ms += mk;
            ls += lk;
        }
    }
    // weights over ages and metallicities: model_sums_zw (opensps_csp_batch.f90:2086-2127), every (age, Z)
    const double* wzm = a.wz ? a.wz + (size_t)m * nt * a.nz : nullptr;
    if (wzm && st == 0)
        for (int k = lane; k < nt * a.nz; k += 64) {
            const double wk = wzm[k];
            if (wk > OPS_SAFE_FLOOR) {
                ms += wk * a.bmass[k];
                ls += wk * pow(10.0, a.blbol[k]);
            }
        }
    ms = wave_sum(ms);
    ls = wave_sum(ls);
    if (lane == 0 && a.lsum) a.lsum[m] = ls;
    if (lane == 0) {
        if (a.mass && !a.ops) a.mass[m] = ms;
        if (a.lbol && !a.ops) a.lbol[m] = ls > OPS_SAFE_FLOOR ? log10(ls) : 0.0;
        a.zlo[m] = zlo;
        a.it[m] = it;
        if (a.idx) { a.idx[3 * m] = it; a.idx[3 * m + 1] = zlo; a.idx[3 * m + 2] = st; }
    }
    if (a.w_out)
        for (int k = lane; k < nt; k += 64) a.w_out[(size_t)m * nt + k] = w[k];

    // The model's columns of the age-sum product: K = [slab zlo ages | slab zlo+1 ages] (blend) or the
    // ages (one basis); the weight below SAFE_FLOOR is 0 (model_sums); young columns keep the ages
    // <= it_tesc, old ones the others; with one part both.
    const double c2 = dz, c1 = 1.0 - dz;
    for (int part = 0; part < (a.ops ? 0 : a.parts); part++) {
        size_t col = (size_t)m * a.parts + part;
        for (int k = lane; k < a.kpad; k += 64) {
            double v = 0.0;
            if (wzm) {
                // K = every (age, Z), metallicities outer: age_sums_w (:2269-2324)
                if (k < a.kpair && st == 0) {
                    const int age = k % nt;
                    double wk = wzm[k];
                    if (!(wk > OPS_SAFE_FLOOR)) wk = 0.0;
                    const bool young = age + 1 <= it;
                    if ((a.parts == 1) || (part == 0 ? young : !young)) v = wk;
                }
            } else if (k < a.kpair) {
                int age = (a.nz > 1) ? (k % nt) : k;   // 0-based age
                double wk = w[age];
                if (!(wk > OPS_SAFE_FLOOR)) wk = 0.0;
                bool young = age + 1 <= it;
                bool keep = (a.parts == 1) || (part == 0 ? young : !young);
                if (keep) v = (a.nz > 1) ? ((k < nt) ? c1 * wk : c2 * wk) : wk;
            }
            if (a.mixed) {
                uint16_t hi, lo;
                split_bf16(v, hi, lo);
                a.whi[col * a.kpad + k] = hi;
                a.wlo[col * a.kpad + k] = lo;
            } else {
                a.w64[col * a.kpad + k] = v;
            }
        }
    }
    wave_chain_header(a, m, p, dz, ls, lane);
    // the model's rank in its pair and slot (k_group): last, after every load of the wavefront (the vector memory
    // counter completes in order, so an atomic issued earlier would hold up every later wait); OPS_GCNT_SLOTS
    // counters a pair keep the atomics on one address few
    if (lane == 0) a.rank[m] = atomicAdd(a.gcnt + (zlo - 1) * OPS_GCNT_SLOTS + (m % OPS_GCNT_SLOTS), 1);
}

#pragma clang fp contract(on)

// ------------------------------------------------------------------------------------------------
// k_group: the models grouped by metallicity pair (k_prep counted each pair's models in OPS_GCNT_SLOTS slots, model m
// in slot m % OPS_GCNT_SLOTS, and ranked each model in its slot with an atomic, so the order within a pair follows
// the slots and the order of arrival: every column of the products is formed alone, so the results do not depend
// on it), and the column tiles of each pair (OPS_TILE_COL columns, and with mtiles the OPS_MM_COLS columns of the
// split-BF16 products). A workgroup per 1024 models: each forms the slots' and pairs' prefixes (a few hundred
// counters) and scatters its models; the first writes the tile lists.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_group(GroupArgs a) {
    __shared__ int gstart[256], gtot[256], tbase[257], mbase[257];
    __shared__ int sstart[256 * OPS_GCNT_SLOTS];   // per (pair, slot) its first position within the pair
    const int t = threadIdx.x, G = a.npair, n = a.n;
    for (int g = t; g < G; g += blockDim.x) {
        int run = 0;
        for (int q = 0; q < OPS_GCNT_SLOTS; q++) {
            sstart[g * OPS_GCNT_SLOTS + q] = run;
            run += a.gcnt[g * OPS_GCNT_SLOTS + q];
        }
        gtot[g] = run;
    }
    __syncthreads();
    if (t == 0) {
        int s = 0, tb = 0, mb = 0;
        for (int g = 0; g < G; g++) {
            gstart[g] = s;
            tbase[g] = tb;
            mbase[g] = mb;
            s += gtot[g];
            tb += (gtot[g] * a.parts + OPS_TILE_COL - 1) / OPS_TILE_COL;
            mb += (gtot[g] * a.parts + OPS_MM_COLS - 1) / OPS_MM_COLS;
        }
        tbase[G] = tb;
        mbase[G] = mb;
        if (blockIdx.x == 0) {
            *a.ntiles = tb;
            if (a.mtiles) *a.nmtiles = mb;
        }
    }
    __syncthreads();
    // each workgroup (all of them form the same prefix) places its slice of the models; the first writes the tiles
    for (int m = blockIdx.x * blockDim.x + t; m < n; m += gridDim.x * blockDim.x) {
        const int g = a.zlo[m] - 1;
        a.order[gstart[g] + sstart[g * OPS_GCNT_SLOTS + m % OPS_GCNT_SLOTS] + a.rank[m]] = m;
    }
    if (blockIdx.x != 0) return;
    for (int g = t; g < G; g += blockDim.x) {
        const int ncol = gtot[g] * a.parts;
        for (int i = 0; tbase[g] + i < tbase[g + 1]; i++) {
            int4 d;
            d.x = g;
            d.y = gstart[g] * a.parts + i * OPS_TILE_COL;
            d.z = min(OPS_TILE_COL, ncol - i * OPS_TILE_COL);
            d.w = 0;
            a.tiles[tbase[g] + i] = d;
        }
        for (int i = 0; a.mtiles && mbase[g] + i < mbase[g + 1]; i++)
            a.mtiles[mbase[g] + i] = make_int4(g, gstart[g] * a.parts + i * OPS_MM_COLS,
                                               min(OPS_MM_COLS, ncol - i * OPS_MM_COLS), 0);
    }
}

// ------------------------------------------------------------------------------------------------
// FP64 age-sum products on the matrix cores. Workgroup tile: 64 pixels x 64 columns, four wavefronts of
// 32 x 32 (2 x 2 blocks of 16 x 16). A is packed in instruction order at set-up (the pair's two slabs side by
// side in K), B per model by k_prep. The blockIdx -> tile map keeps each XCD on one eighth of the pixel
// tiles, so a pair's A rows stay in that XCD's L2 across column tiles.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool gemm_tile(const GemmArgs& a, int& pt, int& ct) {
    const int b = blockIdx.x;
    const int xcd = b & 7, r = b >> 3;
    const int ptx = (a.npxt + 7) >> 3;   // pixel tiles per XCD
    pt = (r % ptx) * 8 + xcd;
    ct = r / ptx;
    return pt < a.npxt && ct < *a.ntiles;
}

__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs a) {
    int pt, ct;
    if (!gemm_tile(a, pt, ct)) return;
    const int4 td = a.tiles[ct];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int blk0 = pt * 4 + wr * 2;   // 16-pixel blocks of this wave
    const int ksteps = a.kpad / OPS_KSTEP_F64;
    const double* Wc[2];
    int model[2], part[2];
    bool valid[2];
    for (int cb = 0; cb < 2; cb++) {
        int c = wc * 32 + cb * 16 + (lane & 15);
        valid[cb] = c < td.z;
        int cc = td.y + (valid[cb] ? c : 0);
        int pos = cc / a.parts;
        part[cb] = cc - pos * a.parts;
        model[cb] = a.order[pos];
        Wc[cb] = a.w64 + ((size_t)model[cb] * a.parts + part[cb]) * a.kpad;
    }
    f64x4_t acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (f64x4_t){0.0, 0.0, 0.0, 0.0};
    const double* Ab = a.A64 + ((size_t)td.x * a.npxblk + blk0) * ksteps * 64 + lane;
    const int kl = lane >> 4;
    for (int ks = 0; ks < ksteps; ks++) {
        double av[2], bv[2];
        av[0] = Ab[(size_t)ks * 64];
        av[1] = Ab[(size_t)(ksteps + ks) * 64];
        bv[0] = valid[0] ? Wc[0][ks * 4 + kl] : 0.0;
        bv[1] = valid[1] ? Wc[1][ks * 4 + kl] : 0.0;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    double* yo = (double*)a.yo;
    for (int j = 0; j < 2; j++) {
        if (!valid[j]) continue;
        double* col = yo + ((size_t)part[j] * a.n + model[j]) * a.nsp;
        for (int i = 0; i < 2; i++) {
            // f64 16x16x4 result layout: register r of lane l holds row (l / 16) + 4 r, column l % 16
            int px = (blk0 + i) * 16 + (lane >> 4);
            for (int r = 0; r < 4; r++) col[px + 4 * r] = acc[i][j][r];
        }
    }
}

// FP64 age sums with vector FMAs: each thread 4 pixels x 4 columns of the 64 x 64 tile, K staged through LDS
// in chunks of 16 (the A operand read from its packed layout).
__global__ __launch_bounds__(256) void k_gemm_f64_vec(GemmArgs a) {
    int pt, ct;
    if (!gemm_tile(a, pt, ct)) return;
    const int4 td = a.tiles[ct];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int ksteps = a.kpad / OPS_KSTEP_F64;
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    __shared__ const double* wcol[64];
    __shared__ int ocol[64];
    if (tid < 64) {
        int c = tid;
        bool v = c < td.z;
        int cc = td.y + (v ? c : 0);
        int pos = cc / a.parts;
        int part = cc - pos * a.parts;
        int model = a.order[pos];
        wcol[c] = v ? a.w64 + ((size_t)model * a.parts + part) * a.kpad : nullptr;
        ocol[c] = v ? part * a.n + model : -1;
    }
    __syncthreads();
    double acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    const size_t pair_base = (size_t)td.x * a.npxblk;
    for (int k0 = 0; k0 < a.kpad; k0 += 16) {
        // A tile: 64 rows x 16 k (1024 values, 4 per thread)
        for (int e = tid; e < 1024; e += 256) {
            int r = e & 63, kk = e >> 6;
            int row = pt * 64 + r, k = k0 + kk;
            double v = 0.0;
            if (k < a.kpad) {
                size_t blk = pair_base + row / 16;
                v = a.A64[(blk * ksteps + k / 4) * 64 + (row & 15) + 16 * (k & 3)];
            }
            As[kk][r] = v;
        }
        for (int e = tid; e < 1024; e += 256) {
            int c = e & 63, kk = e >> 6;
            int k = k0 + kk;
            Bs[kk][c] = (wcol[c] && k < a.kpad) ? wcol[c][k] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            double av[4], bv[4];
            for (int i = 0; i < 4; i++) av[i] = As[kk][ty * 4 + i];
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx * 4 + j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
    double* yo = (double*)a.yo;
    for (int j = 0; j < 4; j++) {
        int oc = ocol[tx * 4 + j];
        if (oc < 0) continue;
        double* col = yo + (size_t)oc * a.nsp + pt * 64 + ty * 4;
        for (int i = 0; i < 4; i++) col[i] = acc[i][j];
    }
}

// Packing of the resident bases into the instruction order of the products (once per grid)
__global__ void k_pack_f64(PackArgs a) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ksteps = a.kpad / OPS_KSTEP_F64;
    size_t total = (size_t)a.npair * a.npxblk * ksteps * 64;
    if (e >= total) return;
    int lane = e % 64;
    size_t r = e / 64;
    int ks = r % ksteps;
    r /= ksteps;
    int blk = r % a.npxblk;
    int pair = r / a.npxblk;
    int row = blk * 16 + (lane & 15);
    int k = ks * 4 + (lane >> 4);
    double v = 0.0;
    if (row < a.nspec && k < a.kpair) {
        int slab = (a.nz > 1) ? pair + (k >= a.nt ? 1 : 0) : 0;
        int age = (a.nz > 1) ? k % a.nt : k;
        v = a.basis[((size_t)slab * a.nt + age) * a.nspec + row];
    }
    a.A64[e] = v;
}

__global__ void k_pack_bf16(PackArgs a) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ksteps = a.kpad / OPS_KSTEP_BF16;
    size_t total = (size_t)a.npair * a.npxblk * ksteps * 64;
    if (e >= total) return;
    int lane = e % 64;
    size_t r = e / 64;
    int ks = r % ksteps;
    r /= ksteps;
    int blk = r % a.npxblk;
    int pair = r / a.npxblk;
    int row = blk * 16 + (lane & 15);
    for (int q = 0; q < 8; q++) {
        int k = ks * 32 + 8 * (lane >> 4) + q;
        double v = 0.0;
        if (row < a.nspec && k < a.kpair) {
            int slab = (a.nz > 1) ? pair + (k >= a.nt ? 1 : 0) : 0;
            int age = (a.nz > 1) ? k % a.nt : k;
            v = a.basis[((size_t)slab * a.nt + age) * a.nspec + row];
        }
        uint16_t hi, lo;
        split_bf16(v, hi, lo);
        a.Ahi[e * 8 + q] = hi;
        a.Alo[e * 8 + q] = lo;
    }
}

// ------------------------------------------------------------------------------------------------
// The chain's per-model constants, one wave per model at the end of k_prep: the uniform FP64 math of
// the chain formed once per model rather than once per thread of the chain. The six (1 + z) powers of
// the IGM optical depth and its cap, the maximum over the pixels below Lyman-alpha (apply_igm_z,
// opensps_csp_batch.f90:4205-4319); the velocity kernel's constants and smoothed pixel range
// (build_velocity_kernel, src/spectra/opensps_smoothing.f90:553-633); the distance term (distance_term,
// src/spectra/opensps_batch_photometry.f90:784-794, with the clamped find_interval of z in the cosmology
// table, descending in z as the reference builds it, or ascending).
// ------------------------------------------------------------------------------------------------

__device__ void wave_chain_header(const PrepArgs& a, int m, const double* p, double dz, double ls, int lane) {
    // the IGM at zred; the observed frame and the distance term at the magnitudes' redshift zm (zred_mags or zred)
    const double z = p[OPS_P_ZRED], opz = 1.0 + z;
    const double zm = a.zmags ? a.zmags[m] : z;
    double pw = 0.0;
    if (lane < 6) {
        const double e = lane == 0 ? 3.46 : lane == 1 ? 1.68 : lane == 2 ? 3.0 + 0.46
                       : lane == 3 ? 1.5 + 0.18 : lane == 4 ? 3.0 - 1.32 : 1.68;
        pw = pow(opz, e);
    }
    double b6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) b6[k] = __shfl(pw, k, 64);
    // the IGM depth's maximum below Lyman-alpha: two pixels a lane at a time (their loads in flight together)
    double tmax = 0.0;
    int jmax = 1;
    if (a.add_igm && z > OPS_SAFE_FLOOR && a.jtop > 0) {
        const int jt = a.jtop;
        for (int j0 = lane; j0 < jt; j0 += 128) {
            double t4[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int j = j0 + 64 * u;
                t4[u] = j < jt ? igm_tau6(a.igm_a + (size_t)j * 6, b6) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 2; u++)
                if (j0 + 64 * u < jt) argmax_merge(tmax, jmax, t4[u], j0 + 64 * u + 1);
        }
        for (int o = 32; o > 0; o >>= 1) {
            double v2 = __shfl_xor(tmax, o, 64);
            int j2 = __shfl_xor(jmax, o, 64);
            argmax_merge(tmax, jmax, v2, j2);
        }
    }
    // the cosmology table's interval of z: eight entries a lane at a time
    const bool asc = a.cz[a.ncosmo - 1] >= a.cz[0];
    int cnt = 0;
    for (int q0 = lane; q0 < a.ncosmo; q0 += 512) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int q = q0 + 64 * u;
            if (q < a.ncosmo) cnt += (asc ? a.cz[q] <= zm : zm < a.cz[q]) ? 1 : 0;
        }
    }
    cnt = wave_sum(cnt);
    const double sig = p[OPS_P_SIGMA_SMOOTH];
    const double minl = p[OPS_P_MIN_WAVE_SMOOTH], maxl = p[OPS_P_MAX_WAVE_SMOOTH];
    const bool smooth = a.smooth_fast && sig > OPS_SAFE_FLOOR;
    const bool window = smooth && minl <= maxl;
    const double sb = sig / (OPS_C_LIGHT * 1.0e-13);
    double dm = 0.0;
    const bool dist_on = fabs(zm) > OPS_SAFE_FLOOR;
    if (dist_on) {
        const int i = max(1, min(cnt, a.ncosmo - 1));
        const double slope = (a.cdl[i] - a.cdl[i - 1]) / (a.cz[i] - a.cz[i - 1]);
        dm = a.cdl[i - 1] + slope * (zm - a.cz[i - 1]);
    }
    // the logarithms and exponentials of the header in one call each over the lanes (the lanes' cases would
    // otherwise run one after another): ln(1 + z), ln(dm / 10), ln(minl), ln(maxl) as log1p; exp(-+4 sigma)
    const double larg = lane == OPS_HDR_LNOPZ ? zm : lane == OPS_HDR_DIST ? dm / 10.0 - 1.0
                      : lane == OPS_HDR_W0 ? minl - 1.0 : lane == OPS_HDR_W1 ? maxl - 1.0 : 0.0;
    const double lg = log1p(larg);
    const double ln_opz = __shfl(lg, OPS_HDR_LNOPZ, 64);
    const double ex = exp(lane == OPS_HDR_FBLUE ? -4.0 * sb : lane == OPS_HDR_FRED ? 4.0 * sb : 0.0);
    const LamLookup& L = a.lk;
    double v = 0.0;
    switch (lane) {
    case OPS_HDR_RZ: v = 1.0 / (1.0 + zm); break;
    case OPS_HDR_LNOPZ: v = lg; break;
    case OPS_HDR_INV2: v = smooth ? 0.5 * (1.0 / sb) * (1.0 / sb) : 0.0; break;
    case OPS_HDR_FBLUE: v = smooth ? ex : 1.0; break;
    case OPS_HDR_FRED: v = smooth ? ex : 1.0; break;
    case OPS_HDR_DIST:   // 5 log10(dm / 10) - 2.5 log10(1 + z)
        if (dist_on && dm == dm && dm > OPS_SAFE_FLOOR)
            v = 5.0 * (lg * 0.43429448190325176) + (-2.5 * (ln_opz * 0.43429448190325176));
        break;
    case OPS_HDR_TMAX: v = tmax; break;
    case OPS_HDR_JMAX: v = (double)jmax; break;
    case OPS_HDR_DZ: v = dz; break;
    case OPS_HDR_LSUM: v = ls; break;   // (the operator path's comes from k_op_sums)
    default: break;
    }
    // #{lam < minl} and #{lam <= maxl}: the bucket lookup from the logarithm, then the exact walk
    if (window && (lane == OPS_HDR_W0 || lane == OPS_HDR_W1)) {
        const double x = lane == OPS_HDR_W0 ? minl : maxl;
        v = lane == OPS_HDR_W0 ? (double)lam_count<true>(L.nspec, L.nbk, L.lnl0, L.bkinv, L.lam, L.bkt, x, lg)
                               : (double)lam_count<false>(L.nspec, L.nbk, L.lnl0, L.bkinv, L.lam, L.bkt, x, lg);
    }
#pragma unroll
    for (int k = 0; k < 6; k++)
        if (lane == OPS_HDR_B6 + k) v = b6[k];
    if (lane < OPS_HDR_N) a.hdr[(size_t)m * OPS_HDR_N + lane] = v;
}

// ------------------------------------------------------------------------------------------------
// k_chain: one workgroup per model. Stages 5-8 of compute_csp_batch in the reference's order:
// attenuation (attenuate_batch, pixel_transmissions: opensps_csp_batch.f90:3576-3818), IGM
// (apply_igm_z :4205-4319), velocity smoothing (build_velocity_kernel, src/spectra/opensps_smoothing.f90:553-633,
// formed per pixel and model on the fly), the observed frame (observed_frame_z,
// src/spectra/opensps_batch_photometry.f90:819-903), the band integrals (band_fluxes :1049-1171) and the
// magnitudes (magnitudes_from_fluxes :185-232).
// T = double: FP64 throughout. T = float: the FP32 chain. Its transmissions, smoothing weights, interpolation
// and band products are FP32 (from FP32 tables); the IGM optical depths, the interval selection, the window
// edges, the band sums across lanes and the magnitudes stay FP64, so every index the chain selects is the FP64
// path's. The per-model constants come from k_prep (hdr).
// Layout: a thread takes quads of 4 consecutive pixels (16- or 32-byte loads and LDS accesses); the band
// coefficients are quad-aligned (zeros pad each band to whole quads), and the observed frame covers the padded
// band pixels, with 0 past the grid.
// Phases (one barrier each): attenuation x IGM -> s0 | smoothing s0 -> s1 and the spectrum out | observed
// frame s1 -> s0 over the band quads | band sums from s0, 16 lanes per band -> LDS | magnitudes, a thread per band.
// ------------------------------------------------------------------------------------------------
template <typename T> struct Quad;
template <> struct Quad<float> { typedef f32x4_t type; };
template <> struct Quad<double> { typedef f64x4_t type; };
template <typename T> struct Pair;
template <> struct Pair<float> { typedef float __attribute__((ext_vector_type(2))) type; };
template <> struct Pair<double> { typedef double __attribute__((ext_vector_type(2))) type; };

// The chain reads its arguments through the kernarg segment's address, made opaque again at each phase (chain_args),
// so that each phase loads the pointers and sizes it uses (scalar loads) instead of holding all of them from the
// kernel's start (with ~40 pointers live at once the compiler spilled them through VGPR lanes). The per-model
// parameters and header, which the chain only reads, are read through the constant address space (scalar loads).
typedef const __attribute__((address_space(4))) ChainArgs CArgs;
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ CArgs* chain_args(CArgs* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// The velocity kernel's window [il, ih] of pixel i (build_velocity_kernel's edges): from i down to the last pixel
// at or below lambda_i exp(-4 sigma), up to the first at or above lambda_i exp(4 sigma)
__device__ __forceinline__ int2 chain_window(CArgs& a, int i, double fblue, double fred) {
    const double li = a.lam[i];
    const double vl = li * fblue, vh = li * fred;
    int il = i, ih = i;
    while (il > 0 && a.lam[il] > vl) il--;
    while (ih < a.nspec - 1 && a.lam[ih] < vh) ih++;
    return make_int2(il, ih);
}

// The same window from one batch of loads, lambda[i - R .. i + R]: the walks' tests on those pixels are made at once
// (lambda increases, so the tests that pass are the first ones); a window wider than R a side is walked on from there
template <int R>
__device__ __forceinline__ int2 chain_window_batch(CArgs& a, int i, double fblue, double fred) {
    const double* __restrict__ lam = a.lam;
    const int nspec = a.nspec;
    double v[2 * R + 1];
#pragma unroll
    for (int o = -R; o <= R; o++) v[o + R] = lam[min(max(i + o, 0), nspec - 1)];
    const double vl = v[R] * fblue, vh = v[R] * fred;
    int nl = 0, nh = 0;
    bool rl = true, rh = true;
#pragma unroll
    for (int o = 0; o <= R; o++) {
        rl = rl && i - o > 0 && v[R - o] > vl;
        rh = rh && i + o < nspec - 1 && v[R + o] < vh;
        nl += (int)rl;
        nh += (int)rh;
    }
    int il = i - nl, ih = i + nh;
    if (nl > R)
        while (il > 0 && lam[il] > vl) il--;
    if (nh > R)
        while (ih < nspec - 1 && lam[ih] < vh) ih++;
    return make_int2(il, ih);
}

// FP32 velocity smoothing of pixel i over a window within [i - 2, i + 2]: the five taps at once, those outside the
// window weighted 0 (the sums in the window's order, as smooth_px forms them)
__device__ __forceinline__ float smooth_px5(CArgs& a, const float* __restrict__ s0, int i, int2 w, float c2, float out) {
    const int il = w.x, ih = w.y, nspec = a.nspec;
    if (ih - il + 1 < 2) return out;
    const double lli = a.loglam[i];
    float4 f[5];
    double ll[5];
#pragma unroll
    for (int o = 0; o < 5; o++) {
        const int k = min(max(i + o - 2, 0), nspec - 1);
        f[o] = a.sm32[k];
        ll[o] = a.loglam[k];
    }
    float tw = 0.0f, acc = 0.0f;
#pragma unroll
    for (int o = 0; o < 5; o++) {
        const int k = i + o - 2;
        if (k >= il && k <= ih) {
            const float d = (float)(lli - ll[o]);
            float wk = __expf(-(c2 * (d * d))) * f[o].x;
            wk *= (k > il ? f[o].y : 0.0f) + (k < ih ? f[o].z : 0.0f);
            tw += wk;
            acc += wk * s0[k];
        }
    }
    if (tw > 0.0f) out = acc * __builtin_amdgcn_rcpf(tw);
    return out;
}

// A wave-uniform FP32 value held in a scalar register (its FP64 source converts in a vector register)
__device__ __forceinline__ float sfloat(double x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)x)));
}

// The sum over each row of 16 lanes, in every lane of the row (DPP: the quad's pairs, the quad, the half-row, the row)
__device__ __forceinline__ float row16_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));   // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));   // row_mirror
    return v;
}

// The velocity-smoothed value of pixel i over its window w (the weights of build_velocity_kernel)
template <typename T>
__device__ __forceinline__ T smooth_px(CArgs& a, const T* __restrict__ s0, int i, int2 w, double inv2, T out) {
    const int il = w.x, ih = w.y;
    if (ih - il + 1 < 2) return out;
    const double lli = a.loglam[i];
    if constexpr (std::is_same<T, float>::value) {
        // four window pixels a step, their table loads issued together; the same sums in the same order
        const float c2 = (float)inv2;
        float tw = 0.0f, acc = 0.0f;
        for (int k0 = il; k0 <= ih; k0 += 4) {
            double ll[4];
            float4 f[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = min(k0 + q, ih);
                ll[q] = a.loglam[k];
                f[q] = a.sm32[k];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = k0 + q;
                if (k <= ih) {
                    const float d = (float)(lli - ll[q]);
                    float wk = __expf(-(c2 * (d * d))) * f[q].x;
                    wk *= (k > il ? f[q].y : 0.0f) + (k < ih ? f[q].z : 0.0f);
                    tw += wk;
                    acc += wk * s0[k];
                }
            }
        }
        if (tw > 0.0f) out = acc * __builtin_amdgcn_rcpf(tw);
    } else {
        double tw = 0.0, acc = 0.0;
        for (int k = il; k <= ih; k++) {
            const double d = lli - a.loglam[k];
            double wk = exp(-(inv2 * (d * d))) * a.wl[k];
            const double dl = k > il ? a.dlam[k - 1] : 0.0;
            const double dr = k < ih ? a.dlam[k] : 0.0;
            wk = 0.5 * wk * (dl + dr);
            tw += wk;
            acc += wk * s0[k];
        }
        if (tw > OPS_SAFE_FLOOR) out = acc / tw;
    }
    return out;
}

// The observed frame's source interval of the band quad q (pixels 4q .. 4q + 3 at lambda / (1 + z)): per pixel the
// interval il (-1 past the grid) and the fraction t. The first pixel finds its interval by the bucket lookup, the
// next three walk on from it. Grid and redshift only: formed before the spectrum is.
template <typename T>
struct ObsIdx {
    int il[4];
    T t[4];
};

template <typename T>
__device__ __forceinline__ ObsIdx<T> obs_index(CArgs& a, int q, double rz, double lnopz) {
    const int nspec = a.nspec;
    const __attribute__((address_space(4))) LamLookup& L = a.lk;
    ObsIdx<T> r;
    int c = -1;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int j = 4 * q + e;
        r.il[e] = -1;
        r.t[e] = (T)0;
        if (j < nspec) {
            const double tg = a.lam[j] * rz;
            if (c < 0) {
                c = lam_count<false>(L.nspec, L.nbk, L.lnl0, L.bkinv, L.lam, L.bkt, tg, a.loglam[j] - lnopz);
            } else {
                while (c < nspec && a.lam[c] <= tg) c++;
            }
            const int il = max(1, min(nspec - 1, c)) - 1;
            r.il[e] = il;
            r.t[e] = (T)((tg - a.lam[il]) * a.rdlam[il]);
        }
    }
    return r;
}

// The same intervals with the quad's four pixels looked up side by side (the FP32 chain): each pixel's bucket start,
// then the three entries around it (the count from them), then 1 / dlam of its interval.
// Three rounds of independent loads instead of the walks from pixel to pixel; the counts and fractions are
// obs_index's. A pixel whose count the four entries do not settle walks as lam_count does.
__device__ __forceinline__ ObsIdx<float> obs_index_quad(CArgs& a, int q, double rz, double lnopz) {
    const int nspec = a.nspec, nbk = a.lk.nbk;
    const double lnl0 = a.lk.lnl0, bkinv = a.lk.bkinv;
    const double* __restrict__ lam = a.lam;
    ObsIdx<float> r;
    double tg[4];
    int c[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int j = min(4 * q + e, nspec - 1);
        tg[e] = lam[j] * rz;
        const double u = (a.loglam[j] - lnopz - lnl0) * bkinv;
        int b = 0;
        if (u > 0.0) b = u >= (double)(nbk - 1) ? nbk - 1 : (int)u;
        c[e] = a.lk.bkt[b];
    }
    // the entries c - 1, c, c + 1 settle the count as c or c + 1 (the bucket holds at most one pixel boundary);
    // otherwise (a rounding of the logarithm across a bucket edge) the walk
    double v[4][3];
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int o = 0; o < 3; o++) v[e][o] = lam[min(max(c[e] - 1 + o, 0), nspec - 1)];
    int il[4];
    double lo[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const double x = tg[e];
        const bool i1 = c[e] - 1 < 0 || v[e][0] <= x, i2 = c[e] < nspec && v[e][1] <= x;
        const bool i3 = c[e] + 1 < nspec && v[e][2] <= x;
        int cnt;
        if (i1 && !i3) {
            cnt = c[e] + (int)i2;
        } else {
            cnt = c[e];
            while (cnt < nspec && lam[cnt] <= x) cnt++;
            while (cnt > 0 && lam[cnt - 1] > x) cnt--;
        }
        il[e] = max(1, min(nspec - 1, cnt)) - 1;
        const int d = il[e] - (c[e] - 1);
        lo[e] = d == 0 ? v[e][0] : d == 1 ? v[e][1] : lam[il[e]];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const bool in = 4 * q + e < nspec;
        r.il[e] = in ? il[e] : -1;
        r.t[e] = in ? (float)((tg[e] - lo[e]) * a.rdlam[il[e]]) : 0.0f;
    }
    return r;
}

// The observed-frame quad q (times 1 / lambda; 0 past the grid) from the rest-frame spectrum src and its intervals
template <typename T>
__device__ __forceinline__ typename Quad<T>::type obs_quad(CArgs& a, const T* __restrict__ src, int q,
                                                           bool rest, const ObsIdx<T>& x) {
    typename Quad<T>::type out;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int j = 4 * q + e;
        T v = (T)0;
        if (j < a.nspec) {
            if (rest) {
                v = src[j];
            } else {
                const int il = x.il[e];
                const T lo = src[il], hi = src[il + 1];
                v = lo + x.t[e] * (hi - lo);
            }
            if (v != v) v = (T)0;
            if constexpr (std::is_same<T, float>::value)
                v = fmaxf(v, 0.0f) * a.wl32[j];
            else
                v = fmax(v, 0.0) * a.wl[j];
        }
        out[e] = v;
    }
    return out;
}

// ------------------------------------------------------------------------------------------------
// Dust emission (the Draine & Li templates with the energy balance) and the AGN torus
// ------------------------------------------------------------------------------------------------
// The diffuse transmission of pixel i as the chain's attenuation forms it (pixel_transmissions,
// opensps_csp_batch.f90:3765-3818): the curve (a fixed one or the power law), then exp(-dust2 curve), in T
template <typename T>
__device__ __forceinline__ T pixel_td(const double* ltab, const float* ltab32, const double* dcurve,
                                      const float* dcurve32, double ed, double dust2, int i) {
    if constexpr (std::is_same<T, float>::value) {
        const float curve = dcurve32 ? dcurve32[i] : ((ed != 0.0) ? __expf((float)ed * ltab32[i]) : 1.0f);
        return __expf(-((float)dust2 * curve));
    } else {
        const double curve = dcurve ? dcurve[i] : ((ed != 0.0) ? exp(ed * ltab[i]) : 1.0);
        return exp(-(dust2 * curve));
    }
}

// The Draine & Li shape at pixel i (dl07_shape, opensps_csp_batch.f90:4103-4122): bilinear in q_PAH and U_min in
// FP64 on the table (nsp, nqpah, 2 numin) (column q + nqpah k; k = 2 u for U_max = U_min, 2 u + 1 for the power-law
// U distribution), the two mixed by gamma, then rounded to T and floored at T's smallest normal
template <typename T>
__device__ __forceinline__ T dl07_shape(const double* tb, int nsp, int nqpah, int iq, double wq, int iu, double wu,
                                        double gam, int i) {
    const size_t c00 = (size_t)iq + (size_t)nqpah * (2 * iu), c10 = (size_t)(iq + 1) + (size_t)nqpah * (2 * iu);
    const size_t c11 = (size_t)(iq + 1) + (size_t)nqpah * (2 * iu + 2), c01 = (size_t)iq + (size_t)nqpah * (2 * iu + 2);
    const size_t nq = (size_t)nqpah;
    const double lo = (1.0 - wq) * (1.0 - wu) * tb[c00 * nsp + i] + wq * (1.0 - wu) * tb[c10 * nsp + i] +
                      wq * wu * tb[c11 * nsp + i] + (1.0 - wq) * wu * tb[c01 * nsp + i];
    const double hi = (1.0 - wq) * (1.0 - wu) * tb[(c00 + nq) * nsp + i] + wq * (1.0 - wu) * tb[(c10 + nq) * nsp + i] +
                      wq * wu * tb[(c11 + nq) * nsp + i] + (1.0 - wq) * wu * tb[(c01 + nq) * nsp + i];
    const T sh = (T)((1.0 - gam) * lo + gam * hi);
    return sh > std::numeric_limits<T>::min() ? sh : std::numeric_limits<T>::min();
}

// 1 - t for t = exp(-x) without cancellation (one_minus_trans_kp, opensps_csp_batch.f90:3925-3939)
template <typename T>
__device__ __forceinline__ T one_minus_trans(T x, T t) {
    if (t == (T)1) return x;
    if (fabs(x) < (T)1) {
        if constexpr (std::is_same<T, float>::value)
            return (1.0f - t) * x / (-logf(t));
        else
            return (1.0 - t) * x / (-log(t));
    }
    return (T)1 - t;
}

// ------------------------------------------------------------------------------------------------
// k_emission: one workgroup (256 threads) per model, after the age sums. Into the chain's header: the dust emission's
// template indices (dl07_indices, opensps_csp_batch.f90:1611-1626: the bracketing intervals of q_PAH and U_min,
// clamped to the grid, their weights not; gamma clamped) and its energy balance (dust_emission_reductions :3963-4042:
// the absorbed light of the attenuation, young through the birth cloud and the diffuse dust, old through its screen
// and the diffuse dust, attenuate_batch :3640-3656; the template; the template times the diffuse transmission; each
// summed over frequency with prepare_batch_grid's trapezoid) as the normalisation s_abs / s_esc that
// add_dust_emission (:4045-4099) applies, for a model with add_dust_emission and dust1 or dust2 > 0
// (batch_model_indices :1569-1570); the AGN torus's interval of agn_tau with its weight clamped (agn_indices
// :1591-1606) and its luminosity fagn L_bol through log10 (add_agn_batch :4342-4433), for add_agn_dust and fagn > 0.
// mdust = 3.21e-3 / (4 pi) s_abs / s_norm (:1206-1211). A model that needs templates the context lacks: status 2.
// Transmissions in T as the chain forms them, the sums in FP64.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_emission(DustArgs a) {
    __shared__ double red[3][4];
    const int m = blockIdx.x, tid = threadIdx.x, n = a.n;
    double* h = a.hdr + (size_t)m * OPS_HDR_N;
    const double dust1 = a.prm[(size_t)OPS_P_DUST1 * n + m], dust2 = a.prm[(size_t)OPS_P_DUST2 * n + m];
    const double fagn = a.prm[(size_t)OPS_P_FAGN * n + m];
    const bool dusty = dust1 > OPS_SAFE_FLOOR || dust2 > OPS_SAFE_FLOOR;
    const bool want_e = a.duste_on && dusty, want_a = a.agn_on && fagn > OPS_SAFE_FLOOR;
    const bool emis = want_e && a.nqpah > 0, agn = want_a && a.ntau > 0;
    if (tid == 0) {
        if ((want_e && !emis) || (want_a && !agn)) {
            atomicMax(a.flag, 2);
            if (a.status) a.status[m] = 2;
        }
        double it = 0.0, wt = 0.0, lagn = 0.0;
        if (agn) {
            const double tau = a.prm[(size_t)OPS_P_AGN_TAU * n + m];
            const int ia = max(1, min(find_interval_ref(a.agn_tau, a.ntau, tau), a.ntau - 1));
            it = (double)(ia - 1);
            wt = fmax(0.0, fmin((tau - a.agn_tau[ia - 1]) / (a.agn_tau[ia] - a.agn_tau[ia - 1]), 1.0));
            const double ls = h[OPS_HDR_LSUM];
            lagn = (ls > OPS_SAFE_FLOOR ? pow(10.0, log10(ls)) : 1.0) * fagn;
        }
        h[OPS_HDR_AGN_IT] = it;
        h[OPS_HDR_AGN_WT] = wt;
        h[OPS_HDR_AGN_L] = lagn;
        h[OPS_HDR_DL_ON] = emis ? 1.0 : 0.0;
    }
    if (!emis) {
        if (tid == 0) {
            h[OPS_HDR_DL_RN] = 0.0;
            if (a.mdust) a.mdust[m] = OPS_SAFE_FLOOR;
        }
        return;
    }
    const double qp = a.prm[(size_t)OPS_P_DUSTE_QPAH * n + m], um = a.prm[(size_t)OPS_P_DUSTE_UMIN * n + m];
    const int iq = max(1, min(find_interval_ref(a.qpah, a.nqpah, qp), a.nqpah - 1)) - 1;
    const int iu = max(1, min(find_interval_ref(a.umin, a.numin, um), a.numin - 1)) - 1;
    const double wq = (qp - a.qpah[iq]) / (a.qpah[iq + 1] - a.qpah[iq]);
    const double wu = (um - a.umin[iu]) / (a.umin[iu + 1] - a.umin[iu]);
    const double gam = fmax(0.0, fmin(a.prm[(size_t)OPS_P_DUSTE_GAMMA * n + m], 1.0));
    const double dust3 = a.prm[(size_t)OPS_P_DUST3 * n + m];
    const double ed = a.prm[(size_t)OPS_P_DUST_INDEX * n + m], ed1 = a.prm[(size_t)OPS_P_DUST1_INDEX * n + m];
    const double fob = a.prm[(size_t)OPS_P_FRAC_OBRUN * n + m], fnd = a.prm[(size_t)OPS_P_FRAC_NODUST * n + m];
    const bool no_birth = dust1 <= OPS_SAFE_FLOOR || 1.0 - fob <= OPS_SAFE_FLOOR;
    const bool no_old = fabs(dust3) <= OPS_SAFE_FLOOR;
    const bool no_diffuse_mix = 1.0 - fnd <= OPS_SAFE_FLOOR;
    const T* y = (const T*)a.yo + (size_t)m * a.nsp;
    const T* o = a.parts == 2 ? (const T*)a.yo + ((size_t)n + m) * a.nsp : nullptr;
    const T cf = (T)fob, cn = (T)fnd;
    double sa = 0.0, sn = 0.0, se = 0.0;
    for (int i = tid; i < a.nspec; i += blockDim.x) {
        T curve, tr_d;
        if constexpr (std::is_same<T, float>::value) {
            curve = a.dcurve32 ? a.dcurve32[i] : ((ed != 0.0) ? __expf((float)ed * a.ltab32[i]) : 1.0f);
            tr_d = __expf(-((float)dust2 * curve));
        } else {
            curve = a.dcurve ? a.dcurve[i] : ((ed != 0.0) ? exp(ed * a.ltab[i]) : 1.0);
            tr_d = exp(-(dust2 * curve));
        }
        const T tau_d = (T)dust2 * curve;
        T tau_b = (T)0, tr_b = (T)1, tau_o = (T)0, tr_o = (T)1;
        if (!no_birth) {
            if constexpr (std::is_same<T, float>::value) {
                tau_b = (float)dust1 * __expf((float)ed1 * a.ltab32[i]);
                tr_b = __expf(-tau_b);
            } else {
                tau_b = dust1 * exp(ed1 * a.ltab[i]);
                tr_b = exp(-tau_b);
            }
        }
        if (!no_old) {
            tau_o = (T)dust3 * curve;
            if constexpr (std::is_same<T, float>::value)
                tr_o = __expf(-tau_o);
            else
                tr_o = exp(-tau_o);
        }
        const T omtb = one_minus_trans<T>(tau_b, tr_b), omto = one_minus_trans<T>(tau_o, tr_o);
        const T omtd = no_diffuse_mix ? (T)0 : ((T)1 - cn) * one_minus_trans<T>(tau_d, tr_d);
        const T ay = tr_b * ((T)1 - cf) + cf;
        const T yv = y[i], ov = o ? o[i] : (T)0;
        const T absorbed = yv * (((T)1 - cf) * omtb + ay * omtd) + ov * (omto + tr_o * omtd);
        const T sh = dl07_shape<T>(a.dustem, a.nsp, a.nqpah, iq, wq, iu, wu, gam, i);
        const double W = a.wnu[i];
        sa += W * (double)absorbed;
        sn += W * (double)sh;
        se += W * (double)(sh * tr_d);
    }
    sa = wave_sum(sa);
    sn = wave_sum(sn);
    se = wave_sum(se);
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = sa;
        red[1][tid >> 6] = sn;
        red[2][tid >> 6] = se;
    }
    __syncthreads();
    if (tid == 0) {
        sa = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        sn = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        se = red[2][0] + red[2][1] + red[2][2] + red[2][3];
        h[OPS_HDR_DL_IQ] = (double)iq;
        h[OPS_HDR_DL_WQ] = wq;
        h[OPS_HDR_DL_IU] = (double)iu;
        h[OPS_HDR_DL_WU] = wu;
        h[OPS_HDR_DL_GAM] = gam;
        h[OPS_HDR_DL_RN] = se > OPS_SAFE_FLOOR ? (double)(T)(sa / se) : 0.0;
        if (a.mdust)
            a.mdust[m] = sn > OPS_SAFE_FLOOR ? 3.21e-3 / (4.0 * 3.14159265358979323846) * (sa / sn) : OPS_SAFE_FLOOR;
    }
}
// ------------------------------------------------------------------------------------------------
// Nebular emission with the CLOUDY grids
// ------------------------------------------------------------------------------------------------
#define OPS_H_PLANCK 6.62607015e-27   // erg s (opensps_constants.f90)
#define OPS_L_SOL 3.828e33            // erg / s

// k_neb_q0: grid (nt_neb, nz), a workgroup per (nebular age, metallicity): the ionizing photons of the basis spectrum,
// the photon-number integral of L_nu / (h nu) below the Lyman limit by the frequency trapezoid over pixels
// 1 .. whlylim (nebular_kernels step 1, opensps_csp_batch.f90:2680-2723), in FP64
__global__ __launch_bounds__(256) void k_neb_q0(int nspec, int nt, int nt_neb, int whlylim, const double* lam,
                                                const double* basis, double* q0) {
    __shared__ double red[4];
    const int t = blockIdx.x, iz = blockIdx.y, tid = threadIdx.x;
    const double* b = basis + ((size_t)iz * nt + t) * nspec;
    double s = 0.0;
    for (int i = tid; i < whlylim - 1; i += blockDim.x) {
        const double nu0 = OPS_C_LIGHT / lam[i], nu1 = OPS_C_LIGHT / lam[i + 1];
        s += 0.5 * fabs(nu1 - nu0) * (b[i] / nu0 + b[i + 1] / nu1);
    }
    s = wave_sum(s);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        const double tot = (red[0] + red[1]) + (red[2] + red[3]);
        q0[(size_t)iz * nt_neb + t] = whlylim >= 2 ? tot / OPS_H_PLANCK * OPS_L_SOL : 0.0;
    }
}

// get_grid_indices_weights (opensps_gas.f90:394-412): the clamped interval (1-based) and weight on an ascending axis
__device__ __forceinline__ void neb_index(const double* g, int ng, double v, int& idx, double& w) {
    int i = find_interval_ref(g, ng, v);
    i = max(1, min(i, ng - 1));
    w = fmax(0.0, fmin((v - g[i - 1]) / (g[i] - g[i - 1]), 1.0));
    idx = i;
}

// A table (rows x nnz x nna x nnu, column-major) at row r (0-based), CLOUDY age a (1-based), interpolated bilinearly in
// log Z and log U (zu_point, :3942-3955)
struct NebZU {
    size_t sz, sa, su;
    double c00, c01, c10, c11;
    int iz, iu;
    __device__ double at(const double* tab, int r, int a) const {
        const double* p = tab + (size_t)r + (size_t)(iz - 1) * sz + (size_t)(a - 1) * sa + (size_t)(iu - 1) * su;
        return c00 * p[0] + c01 * p[su] + c10 * p[sz] + c11 * p[sz + su];
    }
};

// k_neb: a workgroup (256 threads) per model, after the age sums and before the emission stages and the chain.
// (1) the model's gas intervals and, per nebular age, w Q (Q the blend of the bracketing metallicities' photons
// times 1 - frac_obrun); (2) the EUV rows of the age sums redone with the escaping fraction clamp(frac_obrun) on the
// nebular ages (they replace the product's rows); (3) the continuum: per pixel the sum over the nebular ages of
// w Q 10^lg, lg interpolated in the gas parameters and linearly in log age, and within one CLOUDY age interval a
// geometric progression (the reference's per-interval exponential and step ratio); (4) the lines' luminosities, the
// same sums on the line tables; (5) the lines over their profiles (CSR by pixel, in line order); (6) emlines: the
// luminosities through the birth cloud and the diffuse dust at the line's wavelength. FP64 sums; the age sums in
// the path's precision. A thread takes pixels j = tid + 256 k in (2), (3) and (5), so each pixel's updates stay in
// one thread's order.
template <typename T>
__global__ __launch_bounds__(256) void k_neb(NebArgs a) {
    extern __shared__ double nsh[];
    __shared__ int s_iz, s_iu;
    __shared__ double s_wz, s_wu;
    const int m = blockIdx.x, tid = threadIdx.x, n = a.n, ntn = a.nt_neb, nl = a.nline, nt = a.nt;
    double* cq = nsh;
    double* wkv = nsh + ntn;
    double* qv = nsh + 2 * ntn;
    double* lum = nsh + 3 * ntn;
    const double fob = a.prm[(size_t)OPS_P_FRAC_OBRUN * n + m];
    if (tid == 0) {
        int i;
        double w;
        neb_index(a.lz, a.nnz, a.prm[(size_t)OPS_P_GAS_LOGZ * n + m], i, w);
        s_iz = i;
        s_wz = w;
        neb_index(a.lu, a.nnu, a.prm[(size_t)OPS_P_GAS_LOGU * n + m], i, w);
        s_iu = i;
        s_wu = w;
        if (a.lines_in_spec && a.prm[(size_t)OPS_P_SIGMA_SMOOTH * n + m] != a.sigma0) {
            atomicMax(a.flag, 2);   // one profile width a batch (opensps_csp_batch.f90:749-752)
            if (a.status) a.status[m] = 2;
        }
    }
    const int zlo = a.zlo[m], it = a.it[m];
    const bool blend = a.nz > 1;
    const double dz = blend ? a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ] : 0.0;
    const double fesc = fmin(fmax(fob, 0.0), 1.0), fabs_ = 1.0 - fob;
    const double* wm = a.w + (size_t)m * nt;
    for (int t = tid; t < ntn; t += blockDim.x) {
        double q = blend ? (1.0 - dz) * a.q0[(size_t)(zlo - 1) * ntn + t] + dz * a.q0[(size_t)zlo * ntn + t] : a.q0[t];
        q = fabs(q * fabs_);
        const double wk = wm[t] > OPS_SAFE_FLOOR ? wm[t] : 0.0;
        cq[t] = (wk != 0.0 && q > OPS_SAFE_FLOOR) ? wk * q : 0.0;
        wkv[t] = wk;
        qv[t] = q;
    }
    __syncthreads();
    NebZU zu;
    zu.iz = s_iz;
    zu.iu = s_iu;
    const double wz = s_wz, wu = s_wu;
    zu.c00 = (1.0 - wz) * (1.0 - wu);
    zu.c01 = (1.0 - wz) * wu;
    zu.c10 = wz * (1.0 - wu);
    zu.c11 = wz * wu;
    T* y = (T*)a.yo + (size_t)m * a.nsp;
    T* o = a.parts == 2 ? (T*)a.yo + ((size_t)n + m) * a.nsp : nullptr;
    // (2) the EUV rows
    const double* blo = a.basis + (size_t)(zlo - 1) * a.nspec * nt;
    const double* bhi = blend ? a.basis + (size_t)zlo * a.nspec * nt : blo;
    for (int i = tid; i < a.whlylim; i += blockDim.x) {
        double sy = 0.0, so = 0.0;
        for (int k = 0; k < nt; k++) {
            const double wk = wm[k] > OPS_SAFE_FLOOR ? wm[k] : 0.0;
            double bk = blend ? (1.0 - dz) * blo[(size_t)k * a.nspec + i] + dz * bhi[(size_t)k * a.nspec + i]
                              : (1.0 - dz) * blo[(size_t)k * a.nspec + i];
            if (k < ntn) bk = bk * fesc;
            if (k + 1 <= it) sy = sy + wk * bk;
            else so = so + wk * bk;
        }
        if (o) {
            y[i] = (T)sy;
            o[i] = (T)so;
        } else {
            y[i] = (T)(sy + so);
        }
    }
    // (3) the continuum
    if (a.cont_on) {
        zu.sz = (size_t)a.nspec;
        zu.sa = zu.sz * a.nnz;
        zu.su = zu.sa * a.nna;
        const double ln10 = 2.302585092994046;
        for (int j = tid; j < a.nspec; j += blockDim.x) {
            double ys = 0.0, os = 0.0, sy = 0.0, so = 0.0, ra = 0.0, rb = 0.0, dl = 0.0, f = 0.0, r = 1.0;
            int aprev = -1;
            for (int t = 0; t < ntn; t++) {
                const int aa = a.ia[t];
                if (aa != aprev) {
                    ys = ys + sy;
                    os = os + so;
                    sy = 0.0;
                    so = 0.0;
                    ra = aa == aprev + 1 ? rb : zu.at(a.cont, j, aa);
                    rb = zu.at(a.cont, j, aa + 1);
                    aprev = aa;
                    dl = (rb - ra) * ln10;
                    f = exp(ra * ln10 + a.wa[t] * dl);
                } else {
                    if (a.newr[t]) r = exp(a.dwa[t] * dl);
                    f = f * r;
                }
                if (t + 1 <= it) sy = sy + cq[t] * f;
                else so = so + cq[t] * f;
            }
            ys = ys + sy;
            os = os + so;
            if (o) {
                y[j] = (T)((double)y[j] + ys);
                o[j] = (T)((double)o[j] + os);
            } else {
                y[j] = (T)((double)y[j] + (ys + os));
            }
        }
    }
    // (4) the lines' luminosities
    zu.sz = (size_t)nl;
    zu.sa = zu.sz * a.nnz;
    zu.su = zu.sa * a.nna;
    for (int i = tid; i < nl; i += blockDim.x) {
        double sy = 0.0, so = 0.0;
        for (int t = 0; t < ntn; t++) {
            const double wk = wkv[t], q = qv[t];
            if (wk == 0.0 || q <= OPS_SAFE_FLOOR) continue;
            const int aa = a.ia[t];
            const double lg = (1.0 - a.wa[t]) * zu.at(a.line, i, aa) + a.wa[t] * zu.at(a.line, i, aa + 1);
            const double v = wk * (pow(10.0, lg) * q);
            if (t + 1 <= it) sy = sy + v;
            else so = so + v;
        }
        lum[i] = sy;
        lum[nl + i] = so;
    }
    __syncthreads();
    // (5) the profiles
    if (a.lines_in_spec)
        for (int j = tid; j < a.nspec; j += blockDim.x) {
            const int p0 = a.lptr[j], p1 = a.lptr[j + 1];
            if (p0 == p1) continue;
            double sy = 0.0, so = 0.0;
            for (int p = p0; p < p1; p++) {
                const double g = a.gval[p];
                const int i = a.lid[p];
                sy = sy + g * lum[i];
                so = so + g * lum[nl + i];
            }
            if (o) {
                y[j] = (T)((double)y[j] + sy);
                o[j] = (T)((double)o[j] + so);
            } else {
                y[j] = (T)((double)y[j] + (sy + so));
            }
        }
    // (6) the lines through the dust
    if (a.emlines) {
        const double fnd = a.prm[(size_t)OPS_P_FRAC_NODUST * n + m], dust2 = a.prm[(size_t)OPS_P_DUST2 * n + m];
        const double ed = a.prm[(size_t)OPS_P_DUST_INDEX * n + m], dust1 = a.prm[(size_t)OPS_P_DUST1 * n + m];
        const double ed1 = a.prm[(size_t)OPS_P_DUST1_INDEX * n + m];
        for (int i = tid; i < nl; i += blockDim.x) {
            const double pos = a.lpos[i];
            const int lo = a.llo[i];   // 1-based, 1 .. nspec - 1
            double td = 1.0;
            if (!(1.0 - fnd <= OPS_SAFE_FLOOR)) {
                const double cv0 = a.dcurve ? a.dcurve[lo - 1] : (ed != 0.0 ? exp(ed * a.ltab[lo - 1]) : 1.0);
                const double cv1 = a.dcurve ? a.dcurve[lo] : (ed != 0.0 ? exp(ed * a.ltab[lo]) : 1.0);
                const double t0 = (double)(T)exp(-dust2 * cv0), t1 = (double)(T)exp(-dust2 * cv1);
                td = t0 + (t1 - t0) / (a.lam[lo] - a.lam[lo - 1]) * (pos - a.lam[lo - 1]);
            }
            if (td != td) td = 1.0;
            double birth = 1.0;
            if (!(dust1 <= OPS_SAFE_FLOOR || 1.0 - fob <= OPS_SAFE_FLOOR)) birth = exp(-dust1 * pow(pos / 5500.0, ed1));
            double ay = lum[i] * (birth * (1.0 - fob) + fob), ao = lum[nl + i];
            if (1.0 - fnd > OPS_SAFE_FLOOR) {
                ay = ay * (td * (1.0 - fnd) + fnd);
                ao = ao * (td * (1.0 - fnd) + fnd);
            }
            a.emlines[(size_t)m * nl + i] = ay + ao;
        }
    }
}

template __global__ void k_emission<float>(DustArgs);
template __global__ void k_emission<double>(DustArgs);
template __global__ void k_neb<float>(NebArgs);
template __global__ void k_neb<double>(NebArgs);

// ONE: one spectrum buffer in LDS (nsp values): the smoothed values and the observed-frame quads wait in
// registers for a barrier and then overwrite it (windows wider than OPS_CHAIN_KS pixels a thread go through
// the spectrum out, or a staging row, in global memory). Otherwise two buffers. The host takes ONE when the
// band quads fit OPS_CHAIN_KQ a thread.
// Latency: the chain's table lookups do not depend on the spectrum, so each thread starts them early: the windows
// of its first OPS_CHAIN_KW smoothed pixels and the IGM transmissions of its pixels below Lyman-alpha
// (pixel-parallel, applied after the attenuation's barrier) before the attenuation's loads, the source intervals of
// its first observed-frame quad before the smoothing; the attenuation fetches KA quads a thread before it computes,
// and the band sums fetch four bands a group at a time.
// EM: the emission stages (dust emission before the IGM, the AGN torus after it), an instance of its own so that a
// batch without them runs the chain without their code.
template <typename T, bool ONE, bool EM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value ? 8 : 6)))
void k_chain(ChainArgs) {
    CArgs* pa = chain_args((CArgs*)__builtin_amdgcn_kernarg_segment_ptr());
    constexpr bool F32 = std::is_same<T, float>::value;
    constexpr int KA = F32 ? 2 : 1;
    typedef typename Quad<T>::type Q;
    extern __shared__ __align__(32) unsigned char lds_raw[];
    T* s0 = (T*)lds_raw;
    T* s1 = ONE ? s0 : s0 + pa->nsp;
    Q* s0q = (Q*)s0;
    Q* s1q = (Q*)s1;
    const int m = blockIdx.x, n = pa->n, nspec = pa->nspec, nq = pa->nsp >> 2;
    const int tid = threadIdx.x, nth = blockDim.x;
    cdouble* h = (cdouble*)(pa->hdr + (size_t)m * OPS_HDR_N);
    cdouble* prm = (cdouble*)pa->prm;
    const double z = prm[(size_t)OPS_P_ZRED * n + m];
    const double dust1 = prm[(size_t)OPS_P_DUST1 * n + m], dust2 = prm[(size_t)OPS_P_DUST2 * n + m];
    const double dust3 = prm[(size_t)OPS_P_DUST3 * n + m];
    const double ed = prm[(size_t)OPS_P_DUST_INDEX * n + m], ed1 = prm[(size_t)OPS_P_DUST1_INDEX * n + m];
    const double fob = prm[(size_t)OPS_P_FRAC_OBRUN * n + m], fnd = prm[(size_t)OPS_P_FRAC_NODUST * n + m];
    const double igmf = prm[(size_t)OPS_P_IGM_FACTOR * n + m];
    const bool no_birth = dust1 <= OPS_SAFE_FLOOR || 1.0 - fob <= OPS_SAFE_FLOOR;
    const bool no_old = fabs(dust3) <= OPS_SAFE_FLOOR;
    const bool no_diffuse_mix = 1.0 - fnd <= OPS_SAFE_FLOOR;
    const bool igm_on = pa->add_igm && z > OPS_SAFE_FLOOR && pa->jtop > 0;
    const double inv2 = h[OPS_HDR_INV2], fblue = h[OPS_HDR_FBLUE], fred = h[OPS_HDR_FRED];
    const int w0 = (int)h[OPS_HDR_W0], w1 = (int)h[OPS_HDR_W1];
    const bool rest = fabs(pa->zmags ? ((cdouble*)pa->zmags)[m] : z) <= OPS_SAFE_FLOOR;   // the magnitudes' redshift
    const double rz = h[OPS_HDR_RZ], lnopz = h[OPS_HDR_LNOPZ];
    const int qb0 = pa->band_j0 >> 2, qb1 = (pa->band_j1 + 1) >> 2;   // band quads [qb0, qb1)

    // ---- the lookups that need no spectrum (the FP64 chain, short of registers, does its lookups where it uses
    // their results) ----
    constexpr bool EARLY = F32;
    int2 win[OPS_CHAIN_KW];
    // IGM transmissions of pixels tid + k nth below jt (pixels i < jcap take the capped depth)
    const int jt = igm_on ? pa->jtop : 0;
    const double tmax = h[OPS_HDR_TMAX];
    const int jmax = (int)h[OPS_HDR_JMAX];
    const int jcap = jmax > 1 ? jmax : 0;
    T tig[OPS_CHAIN_KI];
#pragma unroll
    for (int k = 0; k < OPS_CHAIN_KI; k++) tig[k] = (T)1;
    if (EARLY && jt > 0) {
        T b6[6];
#pragma unroll
        for (int k = 0; k < 6; k++) b6[k] = (T)h[OPS_HDR_B6 + k];
        const size_t jn = (size_t)pa->jq;
#pragma unroll
        for (int k = 0; k < OPS_CHAIN_KI; k++) {
            const int i = tid + k * nth;
            if (i < jt) {
                if constexpr (F32) {
                    float tau = (float)tmax;
                    if (i >= jcap) {
                        const float* g = pa->igm32s + i;
                        tau = g[0] * b6[0] + g[jn] * b6[1] + g[2 * jn] * b6[2] + g[3 * jn] * b6[3] + g[4 * jn] * b6[4] +
                              g[5 * jn] * b6[5];
                    }
                    tig[k] = __expf(-(tau * (float)igmf));
                } else {
                    double tau = tmax;
                    if (i >= jcap) {
                        const double* g = pa->igm64s + i;
                        tau = g[0] * b6[0] + g[jn] * b6[1] + g[2 * jn] * b6[2] + g[3 * jn] * b6[3] + g[4 * jn] * b6[4] +
                              g[5 * jn] * b6[5];
                    }
                    tig[k] = exp(-tau * igmf);
                }
            }
        }
    }

    pa = chain_args(pa);
    // ---- phase 1: attenuation -> s0 (padded pixels: the zero age sums), KA quads a thread fetched first ----
    const Q* y4 = (const Q*)((const T*)pa->yo + (size_t)m * pa->nsp);
    const Q* o4 = pa->parts == 2 ? (const Q*)((const T*)pa->yo + ((size_t)n + m) * pa->nsp) : nullptr;
    for (int q0 = tid; q0 < nq; q0 += KA * nth) {
        Q y[KA], o[KA], L[KA];
#pragma unroll
        for (int u = 0; u < KA; u++) {
            const int q = q0 + u * nth;
            y[u] = (Q){};
            o[u] = (Q){};
            L[u] = (Q){};
            if (q < nq) {
                y[u] = y4[q];
                if (o4) o[u] = o4[q];
                if constexpr (F32)
                    L[u] = ((const f32x4_t*)pa->ltab32)[q];
                else
                    L[u] = ((const f64x4_t*)pa->ltab)[q];
            }
        }
#pragma unroll
        for (int u = 0; u < KA; u++) {
            const int q = q0 + u * nth;
            if (q >= nq) break;
            Q tot;
            if constexpr (F32) {
                const float cf = (float)fob, cn = (float)fnd;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float curve = pa->dcurve32 ? pa->dcurve32[4 * q + e]
                                                   : ((ed != 0.0) ? __expf((float)ed * L[u][e]) : 1.0f);
                    const float tr_d = __expf(-((float)dust2 * curve));
                    float tr_b = 1.0f, tr_o = 1.0f;
                    if (!no_birth) tr_b = __expf(-((float)dust1 * __expf((float)ed1 * L[u][e])));
                    if (!no_old) tr_o = __expf(-((float)dust3 * curve));
                    const float ssum = (y[u][e] * tr_b * (1.0f - cf) + y[u][e] * cf) + o[u][e] * tr_o;
                    tot[e] = no_diffuse_mix ? ssum : ssum * (tr_d * (1.0f - cn) + cn);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const double curve = pa->dcurve ? pa->dcurve[4 * q + e] : ((ed != 0.0) ? exp(ed * L[u][e]) : 1.0);
                    const double tr_d = exp(-(dust2 * curve));
                    double tr_b = 1.0, tr_o = 1.0;
                    if (!no_birth) tr_b = exp(-(dust1 * exp(ed1 * L[u][e])));
                    if (!no_old) tr_o = exp(-(dust3 * curve));
                    const double ssum = (y[u][e] * tr_b * (1.0 - fob) + y[u][e] * fob) + o[u][e] * tr_o;
                    tot[e] = no_diffuse_mix ? ssum : ssum * (tr_d * (1.0 - fnd) + fnd);
                }
            }
            s0q[q] = tot;
            if (pa->young || pa->old) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int i = 4 * q + e;
                    if (i < nspec) {
                        if (pa->young) ((T*)pa->young)[(size_t)m * nspec + i] = y[u][e];
                        if (pa->old) ((T*)pa->old)[(size_t)m * nspec + i] = o[u][e];
                    }
                }
            }
        }
    }
    // the smoothing windows of the thread's first OPS_CHAIN_KW smoothed pixels (their loads wait with the barrier)
#pragma unroll
    for (int k = 0; k < OPS_CHAIN_KW; k++) {
        const int i = w0 + tid + k * nth;
        win[k] = EARLY && i < w1 ? chain_window_batch<3>(*pa, i, fblue, fred) : make_int2(0, 0);
    }
    __syncthreads();
    pa = chain_args(pa);
    // dust emission (add_dust_emission, opensps_csp_batch.f90:4045-4099): the template times the diffuse
    // transmission times the energy balance's normalisation (k_dust_sums), before the IGM as there
    const double rn = EM ? h[OPS_HDR_DL_RN] : 0.0;
    if (EM && rn != 0.0) {
        const int iq = (int)h[OPS_HDR_DL_IQ], iu = (int)h[OPS_HDR_DL_IU];
        const double wq = h[OPS_HDR_DL_WQ], wu = h[OPS_HDR_DL_WU], gam = h[OPS_HDR_DL_GAM];
        for (int i = tid; i < nspec; i += nth) {
            const T td = pixel_td<T>(pa->ltab, pa->ltab32, pa->dcurve, pa->dcurve32, ed, dust2, i);
            const T sh = dl07_shape<T>(pa->dustem, pa->nsp, pa->nqpah, iq, wq, iu, wu, gam, i);
            s0[i] = s0[i] + sh * td * (T)rn;
        }
        __syncthreads();
    }
    pa = chain_args(pa);
    // the IGM (apply_igm_z): pixels below jt times their transmissions; pixels past KI a thread (grids with more
    // than KI x 512 pixels below Lyman-alpha), and every pixel in the FP64 chain, form theirs here
    if (jt > 0) {
        if constexpr (EARLY) {
#pragma unroll
            for (int k = 0; k < OPS_CHAIN_KI; k++) {
                const int i = tid + k * nth;
                if (i < jt) s0[i] *= tig[k];
            }
        }
        for (int i = tid + (EARLY ? OPS_CHAIN_KI * nth : 0); i < jt; i += nth) {
            double tau = tmax;
            if (i >= jcap) {
                const double* g = pa->igm64s + i;
                const size_t jn = (size_t)pa->jq;
                tau = g[0] * h[OPS_HDR_B6] + g[jn] * h[OPS_HDR_B6 + 1] + g[2 * jn] * h[OPS_HDR_B6 + 2] +
                      g[3 * jn] * h[OPS_HDR_B6 + 3] + g[4 * jn] * h[OPS_HDR_B6 + 4] + g[5 * jn] * h[OPS_HDR_B6 + 5];
            }
            s0[i] *= (T)exp(-tau * igmf);
        }
        __syncthreads();
    }
    pa = chain_args(pa);
    // the AGN torus (add_agn_batch, opensps_csp_batch.f90:4342-4433): fagn L_bol times the agn_tau template (linear
    // between its two optical depths) times the diffuse transmission, after the IGM as there
    const double lagn = EM ? h[OPS_HDR_AGN_L] : 0.0;
    if (EM && lagn != 0.0) {
        const int ia = (int)h[OPS_HDR_AGN_IT];
        const double wta = h[OPS_HDR_AGN_WT];
        for (int i = tid; i < nspec; i += nth) {
            const T td = pixel_td<T>(pa->ltab, pa->ltab32, pa->dcurve, pa->dcurve32, ed, dust2, i);
            double tmpl = (1.0 - wta) * pa->agn[(size_t)ia * pa->nsp + i] + wta * pa->agn[(size_t)(ia + 1) * pa->nsp + i];
            tmpl = tmpl * (double)td;
            s0[i] = (T)((double)s0[i] + lagn * tmpl);
        }
        __syncthreads();
    }

    pa = chain_args(pa);
    // the source intervals of the first observed-frame quad (their lookups overlap the smoothing's)
    ObsIdx<T> ox0;
    constexpr bool OXE = false;   // (the FP32 chain looks its intervals up in phase 3, three rounds of loads)
    if (OXE && !rest && qb0 + tid < qb1) ox0 = obs_index<T>(*pa, qb0 + tid, rz, lnopz);

    pa = chain_args(pa);
    // ---- phase 2: velocity smoothing (-> s1, or in place with ONE), and the spectrum out. Pixels outside the
    // smoothed range [w0, w1) keep their values; the range's pixels are spread over the threads one by one ----
    T* spec_out = pa->spec ? (T*)pa->spec + (size_t)m * nspec : nullptr;
    for (int q = tid; q < nq; q += nth) {
        const int i0 = 4 * q;
        const bool outside = i0 + 3 < w0 || i0 >= w1;
        if (!ONE || spec_out) {
            const Q v = s0q[q];
            if constexpr (!ONE) {
                if (outside) {
                    s1q[q] = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (i0 + e < w0 || i0 + e >= w1) s1[i0 + e] = v[e];
                }
            }
            if (spec_out) {
                // a whole quad outside the smoothed range leaves as one 16-byte store where the row allows it (FP32,
                // row start a multiple of 4 values), else as two 2-value stores (row start even), else value by value;
                // the vector stores non-temporal (the spectrum is written once, for the caller)
                const size_t r0 = (size_t)m * nspec;
                if (outside && i0 + 3 < nspec && (r0 & 1) == 0) {
                    typedef typename Pair<T>::type P2;
                    bool whole = false;
                    if constexpr (F32) {
                        whole = (r0 & 3) == 0;
                        if (whole) __builtin_nontemporal_store(v, (Q*)(spec_out + i0));
                    }
                    if (!whole) {
                        __builtin_nontemporal_store((P2){v[0], v[1]}, (P2*)(spec_out + i0));
                        __builtin_nontemporal_store((P2){v[2], v[3]}, (P2*)(spec_out + i0 + 2));
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (i0 + e < nspec && (i0 + e < w0 || i0 + e >= w1)) spec_out[i0 + e] = v[e];
                }
            }
        }
    }
    static_assert(OPS_CHAIN_KW == 2, "the windows looked up early: win[0], win[1]");
    auto window = [&](int k, int i) -> int2 {
        if (EARLY && k == 0) return win[0];
        if (EARLY && k == 1) return win[1];
        if constexpr (F32) return chain_window_batch<3>(*pa, i, fblue, fred);
        return chain_window(*pa, i, fblue, fred);
    };
    // the FP32 chain takes windows within two pixels a side (the usual case) five taps at once
    const float c2f = sfloat(inv2);
    auto smooth = [&](int i, int2 w) -> T {
        if constexpr (F32) {
            if (w.x >= i - 2 && w.y <= i + 2) return smooth_px5(*pa, s0, i, w, c2f, s0[i]);
        }
        return smooth_px<T>(*pa, s0, i, w, inv2, s0[i]);
    };
    if constexpr (ONE) {
        if (w1 - w0 <= OPS_CHAIN_KS * nth) {
            T sv[OPS_CHAIN_KS];
#pragma unroll
            for (int k = 0; k < OPS_CHAIN_KS; k++) {
                const int i = w0 + tid + k * nth;
                sv[k] = i < w1 ? smooth(i, window(k, i)) : (T)0;
                if (i < w1 && spec_out) spec_out[i] = sv[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < OPS_CHAIN_KS; k++) {
                const int i = w0 + tid + k * nth;
                if (i < w1) s0[i] = sv[k];
            }
        } else {
            T* G = spec_out ? spec_out : (T*)pa->stage + (size_t)m * nspec;
            int k = 0;
            for (int i = w0 + tid; i < w1; i += nth, k++) G[i] = smooth(i, window(k, i));
            __syncthreads();
            for (int i = w0 + tid; i < w1; i += nth) s0[i] = G[i];
        }
    } else {
        int k = 0;
        for (int i = w0 + tid; i < w1; i += nth, k++) {
            const T out = smooth(i, window(k, i));
            s1[i] = out;
            if (spec_out) spec_out[i] = out;
        }
    }
    __syncthreads();

    pa = chain_args(pa);
    // ---- phase 3: the observed frame over the band quads (from s1) -> s0 ----
    auto oidx = [&](int k, int q) -> ObsIdx<T> {
        if (rest) return ObsIdx<T>{};
        if (OXE && k == 0) return ox0;
        if constexpr (F32) return obs_index_quad(*pa, q, rz, lnopz);
        return obs_index<T>(*pa, q, rz, lnopz);
    };
    if constexpr (ONE) {
        Q fo[OPS_CHAIN_KQ];
#pragma unroll
        for (int k = 0; k < OPS_CHAIN_KQ; k++) {
            const int q = qb0 + tid + k * nth;
            if (q < qb1) fo[k] = obs_quad<T>(*pa, s1, q, rest, oidx(k, q));
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < OPS_CHAIN_KQ; k++) {
            const int q = qb0 + tid + k * nth;
            if (q < qb1) s0q[q] = fo[k];
        }
    } else {
        int k = 0;
        for (int q = qb0 + tid; q < qb1; q += nth, k++) s0q[q] = obs_quad<T>(*pa, s1, q, rest, oidx(k, q));
    }
    __syncthreads();

    pa = chain_args(pa);
    // ---- phase 4: band sums (16 lanes per band, 4 pixels a lane, FP64 across the lanes, four bands a group at a
    // time) -> LDS; magnitudes, a thread per band ----
    // the sums wait in LDS after the spectrum (sband_staged: the host sized the LDS for them)
    double* sband = (double*)(lds_raw + (ONE ? 1 : 2) * (size_t)pa->nsp * sizeof(T));
    const bool staged = pa->sband_staged != 0;
    const double dist = h[OPS_HDR_DIST];
    const int g = tid >> 4, gl = tid & 15, ng = nth >> 4;
    constexpr int KB = F32 ? 4 : 2;
    for (int b0 = g; b0 < pa->nbands; b0 += KB * ng) {
        int4 bm[KB];
        bool on[KB];
        int nit = 0;
#pragma unroll
        for (int u = 0; u < KB; u++) {
            const int b = b0 + u * ng;
            bm[u] = make_int4(0, 0, 0, 0);
            on[u] = false;
            if (b < pa->nbands) {
                bm[u] = pa->bmeta[b];   // {coefficient offset, first pixel, pixels}: multiples of 4
                on[u] = pa->mask[b] != 0;
            }
            if (on[u]) nit = max(nit, (bm[u].z / 4 + 15) / 16);
        }
        double s[KB];
        if constexpr (F32) {
            float acc[KB];
#pragma unroll
            for (int u = 0; u < KB; u++) acc[u] = 0.0f;
            for (int it = 0; it < nit; it++) {
                const int q = gl + 16 * it;
#pragma unroll
                for (int u = 0; u < KB; u++) {
                    if (on[u] && 4 * q < bm[u].z) {
                        const f32x4_t cc = *(const f32x4_t*)(pa->bc32 + (unsigned)(bm[u].x + 4 * q));
                        const f32x4_t ff = *(const f32x4_t*)(s0 + (unsigned)(bm[u].y + 4 * q));
                        acc[u] = fmaf(cc[0], ff[0], acc[u]);
                        acc[u] = fmaf(cc[1], ff[1], acc[u]);
                        acc[u] = fmaf(cc[2], ff[2], acc[u]);
                        acc[u] = fmaf(cc[3], ff[3], acc[u]);
                    }
                }
            }
            // the 16 lanes' partial sums added in FP32 (a tree across the row)
#pragma unroll
            for (int u = 0; u < KB; u++) s[u] = (double)row16_sum(acc[u]);
        } else {
#pragma unroll
            for (int u = 0; u < KB; u++) s[u] = 0.0;
            for (int it = 0; it < nit; it++) {
                const int q = gl + 16 * it;
#pragma unroll
                for (int u = 0; u < KB; u++) {
                    if (on[u] && 4 * q < bm[u].z) {
                        const f64x4_t cc = *(const f64x4_t*)(pa->bc64 + (unsigned)(bm[u].x + 4 * q));
                        const f64x4_t ff = *(const f64x4_t*)(s0 + (unsigned)(bm[u].y + 4 * q));
                        s[u] += cc[0] * ff[0] + cc[1] * ff[1] + cc[2] * ff[2] + cc[3] * ff[3];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < KB; u++) {
            if constexpr (!F32)
                for (int o = 8; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o, 16);
            const int b = b0 + u * ng;
            if (gl == 0 && b < pa->nbands) {
                if (staged) {
                    sband[b] = s[u];
                } else {
                    double mag = NAN;
                    if (on[u]) mag = (s[u] > 0.0 ? -2.5 * log10(s[u]) : NAN) - 48.60 - 2.5 * pa->amzl + dist;
                    pa->mags[(size_t)m * pa->nbands + b] = mag;
                    if (pa->fband) pa->fband[(size_t)m * pa->nbands + b] = s[u];
                }
            }
        }
    }
    if (!staged) return;
    __syncthreads();
    pa = chain_args(pa);
    for (int b = tid; b < pa->nbands; b += nth) {
        const double s = sband[b];
        double mag = NAN;
        if (pa->mask[b]) mag = (s > 0.0 ? -2.5 * log10(s) : NAN) - 48.60 - 2.5 * pa->amzl + dist;
        pa->mags[(size_t)m * pa->nbands + b] = mag;
        if (pa->fband) pa->fband[(size_t)m * pa->nbands + b] = s;
    }
}

template __global__ void k_chain<double, false, false>(ChainArgs);
template __global__ void k_chain<float, false, false>(ChainArgs);
template __global__ void k_chain<double, true, false>(ChainArgs);
template __global__ void k_chain<float, true, false>(ChainArgs);
template __global__ void k_chain<double, false, true>(ChainArgs);
template __global__ void k_chain<float, false, true>(ChainArgs);
template __global__ void k_chain<double, true, true>(ChainArgs);
template __global__ void k_chain<float, true, true>(ChainArgs);

// ------------------------------------------------------------------------------------------------
// Kernel-level entry points: band_fluxes and observed_frame_z in the reference's own layouts
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_band_fluxes(int n, int nspec, int nbands, const T* bt, const int* jlo, const int* jhi, const int* do_band,
                              const T* wt, const T* fobs_t, double* fband) {
    const int b = blockIdx.y, m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (m >= n) return;
    double s = 0.0;
    if (do_band[b])
        for (int j = jlo[b] + lane; j <= jhi[b]; j += 64)
            s += (double)wt[j - 1] * ((double)fobs_t[(size_t)(j - 1) * n + m] * (double)bt[(size_t)(j - 1) * nbands + b] +
                                      (double)fobs_t[(size_t)j * n + m] * (double)bt[(size_t)j * nbands + b]);
    s = wave_sum(s);
    if (lane == 0) fband[(size_t)m * nbands + b] = s;
}
template __global__ void k_band_fluxes<double>(int, int, int, const double*, const int*, const int*, const int*,
                                               const double*, const double*, double*);
template __global__ void k_band_fluxes<float>(int, int, int, const float*, const int*, const int*, const int*,
                                              const float*, const float*, double*);

template <typename T>
__global__ void k_observed_frame(int n, int nspec, const double* zc, const double* lam, const T* total, T* fobs_t) {
    const int m = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nspec) return;
    const double z = zc[m];
    const T* tc = total + (size_t)m * nspec;
    double v;
    if (fabs(z) > OPS_SAFE_FLOOR) {
        double tg = lam[j] / (1.0 + z);
        int il = max(1, min(nspec - 1, count_le(lam, nspec, tg))) - 1;
        double t = (tg - lam[il]) / (lam[il + 1] - lam[il]);
        v = (double)tc[il] + t * ((double)tc[il + 1] - (double)tc[il]);
    } else {
        v = (double)tc[j];
    }
    if (v != v) v = 0.0;
    double wl = lam[j] > 2.2250738585072014e-308 ? 1.0 / lam[j] : 0.0;
    fobs_t[(size_t)j * n + m] = (T)fmax(v, 0.0) * (T)wl;
}
template __global__ void k_observed_frame<double>(int, int, const double*, const double*, const double*, double*);
template __global__ void k_observed_frame<float>(int, int, const double*, const double*, const float*, float*);

// Self-test of the matrix-instruction lane layouts the products assume (one wavefront): which = 0 runs
// f64 16x16x4 on a[64], b[64] (doubles); which = 1 runs bf16 16x16x32 on a[64][8], b[64][8] (uint16).
// d[64][4] receives each lane's accumulator.
__global__ void k_selftest_mfma(int which, const void* a, const void* b, void* d) {
    const int lane = threadIdx.x;
    if (which == 0) {
        f64x4_t acc = (f64x4_t){0.0, 0.0, 0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(((const double*)a)[lane], ((const double*)b)[lane], acc, 0, 0, 0);
        for (int r = 0; r < 4; r++) ((double*)d)[lane * 4 + r] = acc[r];
    } else {
        f32x4_t acc = (f32x4_t){0.f, 0.f, 0.f, 0.f};
        bf16x8_t av = ((const bf16x8_t*)a)[lane], bv = ((const bf16x8_t*)b)[lane];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        for (int r = 0; r < 4; r++) ((float*)d)[lane * 4 + r] = acc[r];
    }
}
