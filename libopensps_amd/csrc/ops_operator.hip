// The SSP operator with the IMF per model on gfx950 (MI355X): the isochrone points' IMF weights in FP64,
// their per-age sums, the fold of the weighted points into template weights (a segmented reduction in the
// kernel precision), and the template-product operand.
//
//   k_op_weights  one wavefront per (tile of <= 64 models of one metallicity pair, age, pass): the lanes are
//                 the models, the points of the age stream through the wavefront (uniform loads); FP64 bin
//                 integrals piece by piece (imf_pieces_multi, src/physics/opensps_ssp_batch.f90:2913-3091,
//                 imf_bin_weights_tab, src/physics/opensps_imf.f90:439-559); per-age mass, L_bol and
//                 photon sums with the remnants (add_remnant_mass_p,
//                 src/physics/opensps_stellar_modifications.f90:729-816)
//   k_op_segments one wavefront per (segment of <= 64 entries of one template's list, tile, pass): the
//                 entries (CSR by template, sorted by age) times the points' weights and the SFH weights,
//                 split young / old at it_tesc (fold_multi's segments, opensps_ssp_batch.f90:3095-3163)
//   k_op_fold     one wavefront per (tile, block of 8 templates): the segments' sums per template over both
//                 metallicity passes (:3164-3186), written as the B operand of the template product (split
//                 BF16 [column][K] or FP64)
//   k_op_sums     per model: the surviving mass and L_bol from the per-age sums (model_sums MSUMS_OWN,
//                 opensps_csp_batch.f90:2129-2176)
//   k_op_tiles    the uniform column tiles of the template product (one shared template library)
#include <type_traits>

#include "ops_common.h"
#include "ops_context.h"
#include "ops_launch.h"

// ------------------------------------------------------------------------------------------------
// IMF integrals (FP64, contraction off: each operation rounds as the reference's CPU build rounds it)
// ------------------------------------------------------------------------------------------------
#pragma clang fp contract(off)

struct ImfP {
    double salp, a1, a2, a3, mdave, lower, upper, lower_bound, mlim_bh, mlim_ns;
};

// powerlaw_bin (opensps_imf.f90:754-792): int_lo^hi m^(e-1) dm from plo = lo^e, dl = ln(hi/lo) by Kahan's log1p
__device__ double d_powerlaw_bin(double plo, double lo, double hi, double e) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    double x = (hi - lo) / lo, u = 1.0 + x, dl;
    if (u == 1.0) dl = x;
    else dl = log(u) * x / (u - 1.0);
    if (e == 0.0) return plo * dl;
    double y = e * dl;
    if (fabs(y) < 0.25) {
        double r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return plo * dl * r;
    }
    return plo * (exp(y) - 1.0) / e;
}

// expm1_series (opensps_imf.f90:641-667)
__device__ double d_expm1_series(double y) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    if (fabs(y) < 0.25) {
        double r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return y * r;
    }
    return exp(y) - 1.0;
}

// integrate_powerlaw (opensps_imf.f90:725-748)
__device__ double d_integrate_powerlaw(double lo, double hi, double alpha, bool mw, double coeff) {
    if (hi <= lo) return 0.0;
    const double denom = mw ? 2.0 - alpha : 1.0 - alpha;
    return coeff * d_powerlaw_bin(pow(lo, denom), lo, hi, denom);
}

// imf_integral (opensps_imf.f90:206-315) for the power-law types 0 (Salpeter), 2 (Kroupa), 4 (Dave). Not inlined:
// the weights' hot loop calls it only for bins across a break, and keeps its registers for the common path.
__device__ __attribute__((noinline)) double d_imf_integral(int type, const ImfP& p, double m1, double m2, bool mw) {
    const double lo = fmin(m1, m2), hi = fmax(m1, m2);
    double v = 0.0;
    if (hi <= lo) return 0.0;
    const int t = type % 10;
    if (t == 0) {
        v = d_integrate_powerlaw(lo, hi, p.salp, mw, 1.0);
    } else if (t == 2) {
        const double c2 = pow(0.5, -p.a1 + p.a2);
        double sl = fmax(lo, 0.08), sh = fmin(hi, 0.5);
        if (sh > sl) v = v + d_integrate_powerlaw(sl, sh, p.a1, mw, 1.0);
        sl = fmax(lo, 0.5); sh = fmin(hi, 1.0);
        if (sh > sl) v = v + d_integrate_powerlaw(sl, sh, p.a2, mw, c2);
        sl = fmax(lo, 1.0); sh = hi;
        if (sh > sl) v = v + d_integrate_powerlaw(sl, sh, p.a3, mw, c2);
    } else if (t == 4) {
        double sl = fmax(lo, 0.08), sh = fmin(hi, p.mdave);
        if (sh > sl) v = v + d_integrate_powerlaw(sl, sh, p.a1, mw, 1.0);
        sl = fmax(lo, p.mdave); sh = hi;
        if (sh > sl) v = v + d_integrate_powerlaw(sl, sh, p.a2, mw, pow(p.mdave, -p.a1 + p.a2));
    }
    return v;
}

// add_remnant_mass_p (opensps_stellar_modifications.f90:729-816): black holes, neutron stars, white dwarfs
__device__ __attribute__((noinline)) double d_remnant_mass(int type, const ImfP& p, double mmax) {
    const double hi = p.upper, lbh = p.mlim_bh, lns = p.mlim_ns;
    const double norm = d_imf_integral(type, p, p.lower, hi, true);
    double add = 0.0;
    if (norm <= 0.0) return 0.0;
    double a = fmin(fmax(lbh, mmax), hi);
    if (a < hi) add = add + 0.5 * d_imf_integral(type, p, a, hi, true) / norm;
    if (mmax <= lbh) {
        a = fmin(fmax(lns, mmax), hi);
        const double b = fmin(lbh, hi);
        if (a < b) add = add + 1.4 * d_imf_integral(type, p, a, b, false) / norm;
    }
    if (mmax <= lns) {
        a = fmin(mmax, hi);
        const double b = fmin(lns, hi);
        if (a < b)
            add = add + (0.48 * d_imf_integral(type, p, a, b, false)) / norm +
                  (0.077 * d_imf_integral(type, p, a, b, true)) / norm;
    }
    return add;
}

// The state of imf_bin_weights_tab over one piece: the current power-law segment and the chain of shared edges
struct BinState {
    int cur;
    double e, c, d, ce, prev_hi, prev_phi;
};

__device__ __forceinline__ void bin_reset(BinState& s) {
    s.cur = 0; s.e = 0.0; s.c = 0.0; s.d = 1.0; s.ce = 0.0; s.prev_hi = -1.0; s.prev_phi = 0.0;
}

// A new power-law segment's constants (imf_bin_weights_tab, opensps_imf.f90:500-525), once per segment change
__device__ __attribute__((noinline)) void bin_segment(int t, const ImfP& p, int seg, BinState& s) {
    double alpha, coeff = 1.0;
    if (t == 0) {
        alpha = p.salp;
    } else if (t == 2) {
        alpha = seg == 1 ? p.a1 : (seg == 2 ? p.a2 : p.a3);
        if (seg > 1) coeff = pow(0.5, -p.a1 + p.a2);
    } else {
        alpha = seg == 1 ? p.a1 : p.a2;
        if (seg > 1) coeff = pow(p.mdave, -p.a1 + p.a2);
    }
    s.e = 1.0 - alpha; s.c = coeff; s.d = 1.0 - alpha;
    s.ce = 0.0;
    if (s.e != 0.0) s.ce = s.c / s.e;
    s.cur = seg;
    s.prev_hi = -1.0;
}

// The first bin of an age (lower edge at the IMF's lower bound): the difference of powers (:531-535)
__device__ __attribute__((noinline)) double bin_first(const BinState& s, double lo, double hi) {
    if (fabs(s.d) <= 1.0e-12) return s.c * log(hi / lo);
    return s.c * (pow(hi, s.e) - pow(lo, s.e)) / s.d;
}

// One bin of imf_bin_weights_tab (opensps_imf.f90:466-557), FP64 build: stars per unit initial mass. The
// segment classification and the limits decide which bins weigh nothing, exactly as the reference.
__device__ __attribute__((noinline)) double d_bin_weight(int type, const ImfP& p, double inv_total, double mini, double m1,
                                                      double m2,
                               double llo, double dl, int first, BinState& s) {
    if (mini < p.lower || mini > p.upper) return 0.0;
    const double lo = first == 1 ? p.lower_bound : m1, hi = m2;
    if (hi <= lo) return 0.0;
    const int t = type % 10;
    int seg = 0;
    if (t == 0) {
        seg = 1;
    } else if (t == 2) {
        if (lo >= 0.08 && hi <= 0.5) seg = 1;
        else if (lo >= 0.5 && hi <= 1.0) seg = 2;
        else if (lo >= 1.0) seg = 3;
    } else if (t == 4) {
        if (lo >= 0.08 && hi <= p.mdave) seg = 1;
        else if (lo >= p.mdave) seg = 2;
    }
    if (seg > 0 && seg != s.cur) bin_segment(t, p, seg, s);
    double val;
    if (seg > 0 && s.cur == seg) {
        if (first == 1) {
            val = bin_first(s, lo, hi);
        } else if (s.e == 0.0) {
            val = s.c * dl;
        } else {
            const double plo = (lo == s.prev_hi) ? s.prev_phi : exp(s.e * llo);
            const double em1 = d_expm1_series(s.e * dl);
            val = s.ce * plo * em1;
            s.prev_hi = hi;
            s.prev_phi = plo + plo * em1;
        }
    } else {
        val = d_imf_integral(type, p, lo, hi, false);
        s.prev_hi = -1.0;
        if (seg == 0) s.cur = 0;
    }
    return val * inv_total;
}

// The model's IMF from its parameter slots and the context's settings
__device__ __forceinline__ ImfP model_imf(const OpWArgs& a, int m) {
    const int n = a.n;
    ImfP p;
    p.salp = a.salp_ind;
    p.a1 = a.prm[(size_t)OPS_P_IMF1 * n + m];
    p.a2 = a.prm[(size_t)OPS_P_IMF2 * n + m];
    p.a3 = a.prm[(size_t)OPS_P_IMF3 * n + m];
    p.mdave = a.mdave;
    p.lower = a.prm[(size_t)OPS_P_IMF_LOWER * n + m];
    p.upper = a.prm[(size_t)OPS_P_IMF_UPPER * n + m];
    p.lower_bound = p.lower;
    p.mlim_bh = a.mlim_bh;
    p.mlim_ns = a.mlim_ns;
    return p;
}

// ------------------------------------------------------------------------------------------------
// k_op_imf: one thread per model: the total mass integral over [lower, upper] (the weights' normalisation) and
// each power-law segment's constants e = 1 - alpha, c and c / total (bin_segment), formed once per model.
// ------------------------------------------------------------------------------------------------
__global__ void k_op_imf(OpWArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const ImfP p = model_imf(a, m);
    const int t = a.imf_type % 10;
    const double tot = d_imf_integral(a.imf_type, p, p.lower, p.upper, true);
    double* o = a.impc + (size_t)m * OPS_IMFC_N;
    o[OPS_IMFC_INV_TOTAL] = tot > 0.0 ? 1.0 / tot : 1.0;
    for (int seg = 1; seg <= 3; seg++) {
        BinState s;
        bin_reset(s);
        if (t == 0 && seg > 1) {
            s.e = 0.0; s.c = 0.0; s.ce = 0.0;
        } else if (t == 4 && seg > 2) {
            s.e = 0.0; s.c = 0.0; s.ce = 0.0;
        } else {
            bin_segment(t, p, seg, s);
        }
        o[OPS_IMFC_E + seg - 1] = s.e;
        o[OPS_IMFC_C + seg - 1] = s.c;
        o[OPS_IMFC_CT + seg - 1] = s.c * o[OPS_IMFC_INV_TOTAL];
    }
}

// ------------------------------------------------------------------------------------------------
// k_op_slow: grid (ceil(nt / 4), max tiles, passes), one wavefront per (tile, age, pass) as k_op_weights. The
// age's bins that leave the common path (its first bin, whose lower edge is each model's lower limit; bins
// across a break; empty bins; listed per age by the host) by the full imf_bin_weights_tab per lane, and the
// age's remnants: their weights go to the points' rows, their sums to the age's partial sums, which k_op_weights
// adds.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_slow(OpWArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + wv, tile = a.tile0 + blockIdx.y, pass = blockIdx.z;
    if (t >= a.nt || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const int iz = a.npass == 1 ? 0 : td.x + pass;
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int nt = a.nt;
    const ImfP p = model_imf(a, m);
    const double inv_total = a.impc[(size_t)m * OPS_IMFC_N + OPS_IMFC_INV_TOTAL];
    const int zo = a.zoff[iz];
    T* u = (T*)a.u + ((size_t)pass * a.utiles + blockIdx.y) * a.npz * 64 + lane;
    double sm = 0.0, sl = 0.0, sq = 0.0;
    int cnt = 0;
    const int j0 = a.slow_ptr[iz * nt + t], j1 = a.slow_ptr[iz * nt + t + 1];
    for (int j = j0; j < j1; j++) {
        const int k = a.slow_pt[j];
        const double* pt = a.pt + (size_t)k * OPS_PT_N;
        BinState s;
        bin_reset(s);
        const double w = d_bin_weight(a.imf_type, p, inv_total, pt[0], pt[1], pt[2], pt[3], pt[4], a.first[k], s);
        u[(size_t)(k - zo) * 64] = (T)w;
        cnt += w > 0.0 ? 1 : 0;
        sm = sm + pt[5] * w;
        sl = sl + pt[6] * w;
        sq = sq + pt[7] * w;
    }
    const int k0 = a.age_ptr[iz * nt + t], k1 = a.age_ptr[iz * nt + t + 1];
    if (a.remnants && k1 > k0) sm = sm + d_remnant_mass(a.imf_type, p, a.age_mmax[iz * nt + t]);
    double* sp = a.spart + (((size_t)pass * a.maxtiles + tile) * nt + t) * 3 * 64 + lane;
    sp[0] = sm;
    sp[64] = sl;
    sp[128] = sq;
    a.scnt[(((size_t)pass * a.maxtiles + tile) * nt + t) * 64 + lane] = cnt;
}

#pragma clang fp contract(on)
// expm1(y) / y = sum_j y^j / (j + 1)!, j = 0 .. 12, by Horner with fused multiply-adds: expm1_series's
// (opensps_imf.f90:641-667) polynomial, the same terms, in 12 FMAs; |y| < 1/4 as there. The coefficients
// 1 / (j + 1)! come from a device table (scalar registers, the addends of the FMAs).
__device__ __forceinline__ double d_expm1_over_y(double y, const double* pc) {
    double s = pc[12];
#pragma unroll
    for (int j = 11; j >= 0; j--) s = fma(s, y, pc[j]);
    return s;
}

typedef __attribute__((address_space(4))) const double* OpsDoubleC;
typedef __attribute__((address_space(4))) const OpsWPoint* OpsWPointC;

// A point record through scalar loads, field by field
__device__ __forceinline__ OpsWPoint wp_load(OpsWPointC p) {
    OpsWPoint r;
    r.mini = p->mini;
    r.llo = p->llo;
    r.dl = p->dl;
    r.cm = p->cm;
    r.cl = p->cl;
    r.code = p->code;
    return r;
}

// ------------------------------------------------------------------------------------------------
// k_op_weights: grid (ceil(nt ngroup / 4), the round's tiles, passes), 256 threads = 4 (age, point group) pairs:
// a wavefront takes OPS_OP_GPTS consecutive points of one age (whole pieces of 32, where the chain of shared edges
// restarts anyway), so the weights of one round of tiles still fill the GPU. Pass 0 takes the operator of
// metallicity zlo with weight 1 - dz, pass 1 that of zlo + 1 with weight dz (one pass with one metallicity).
// The common bins of an age (not its first, inside one power-law segment, not empty): a point's record is the
// same for every lane (the lanes are models), so the records come through scalar loads and the segment and the
// chain of shared edges are uniform, as the host coded them for the context's IMF type (OpsWPoint::code). Every
// lane runs the chain hi^e = lo^e (1 + expm1(e dl)) over every common point (exp only where the host broke the
// chain: a new piece of 32 as imf_bin_weights_tab's, a new segment, a bin k_op_slow took, or edges not shared),
// and the limits select the weights. Per bin: w = c lo^e dl (expm1(e dl) / (e dl)) / total, which is c dl / total
// for e = 0 (imf_bin_weights_tab :537-551 with :502-525's constants, c and 1 / total from k_op_imf).
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_weights(OpWArgs a) {
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wi = blockIdx.x * 4 + wv, tile = a.tile0 + blockIdx.y, pass = blockIdx.z;
    const int t = wi / a.ngroup, grp = wi - t * a.ngroup;
    if (t >= a.nt || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];   // {metallicity pair (zlo - 1), first column, columns}
    const int iz = a.npass == 1 ? 0 : __builtin_amdgcn_readfirstlane(td.x) + pass;
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int n = a.n, nt = a.nt;
    const double lower = a.prm[(size_t)OPS_P_IMF_LOWER * n + m], upper = a.prm[(size_t)OPS_P_IMF_UPPER * n + m];
    const double* ic = a.impc + (size_t)m * OPS_IMFC_N;
    const double e1 = ic[OPS_IMFC_E], e2 = ic[OPS_IMFC_E + 1], e3 = ic[OPS_IMFC_E + 2];
    const double c1 = ic[OPS_IMFC_CT], c2 = ic[OPS_IMFC_CT + 1], c3 = ic[OPS_IMFC_CT + 2];
    const int ka = __builtin_amdgcn_readfirstlane(a.age_ptr[iz * nt + t]);
    const int kb = __builtin_amdgcn_readfirstlane(a.age_ptr[iz * nt + t + 1]);
    const int k0 = min(kb, ka + grp * OPS_OP_GPTS), k1 = min(kb, k0 + OPS_OP_GPTS);
    const int zo = __builtin_amdgcn_readfirstlane(a.zoff[iz]);
    T* u = (T*)a.u + ((size_t)pass * a.utiles + blockIdx.y) * a.npz * 64 + lane;
    const OpsWPointC wp = (OpsWPointC)a.wp;
    const OpsDoubleC pcs = (OpsDoubleC)a.poly;
    double pc[13];
#pragma unroll
    for (int j = 0; j < 13; j++) pc[j] = pcs[j];
    double sm = 0.0, sl = 0.0, phi = 1.0;
    int cnt = 0;
    // one point a step; the next point's record is fetched (scalar loads) before this point's arithmetic, so its
    // latency hides behind it
    OpsWPoint nxt;
    if (k0 < k1) nxt = wp_load(wp + k0);
    for (int k = k0; k < k1; k++) {
        const OpsWPoint r = nxt;
        if (k + 1 < k1) nxt = wp_load(wp + k + 1);
        const int seg = r.code & 3;
        if (seg == 0) continue;   // k_op_slow's bin
        const double e = seg == 1 ? e1 : (seg == 2 ? e2 : e3);
        const double c = seg == 1 ? c1 : (seg == 2 ? c2 : c3);
        const double dl = r.dl;
        const double plo = (r.code & 4) ? phi : exp(e * r.llo);
        const double y = e * dl;
        double sx = d_expm1_over_y(y, pc);
        if (!(fabs(y) < 0.25)) sx = (exp(y) - 1.0) / y;
        phi = fma(plo, y * sx, plo);
        const double w = (r.mini < lower || r.mini > upper) ? 0.0 : c * dl * plo * sx;
        u[(size_t)(k - zo) * 64] = (T)w;
        cnt += w > 0.0 ? 1 : 0;
        sm = fma(r.cm, w, sm);
        sl = fma(r.cl, w, sl);
    }
    // the group's sums (k_op_agesums adds the groups in order, then k_op_slow's)
    const size_t rg = (((size_t)pass * a.utiles + blockIdx.y) * nt + t) * a.ngroup + grp;
    a.gsum[rg * 2 * 64 + lane] = sm;
    a.gsum[rg * 2 * 64 + 64 + lane] = sl;
    a.gcnt[rg * 64 + lane] = cnt;
}

// ------------------------------------------------------------------------------------------------
// k_op_agesums: grid (ceil(nt / 4), the round's tiles, passes), a wavefront per (age, tile, pass): the age's mass and
// L_bol sums from its point groups (in order) and the slow bins with the remnants (k_op_slow), times the blend
// factor, the count of points with a nonzero weight, and the age's SFH x blend factor for the segments.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_op_agesums(OpWArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + wv, tile = a.tile0 + blockIdx.y, pass = blockIdx.z;
    if (t >= a.nt || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int n = a.n, nt = a.nt;
    const double dz = a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ];
    const double cz = a.npass == 1 ? 1.0 : (pass == 0 ? 1.0 - dz : dz);
    double sm = 0.0, sl = 0.0;
    int cnt = 0;
    const size_t rg0 = (((size_t)pass * a.utiles + blockIdx.y) * nt + t) * a.ngroup;
    for (int g = 0; g < a.ngroup; g++) {
        sm = sm + a.gsum[(rg0 + g) * 2 * 64 + lane];
        sl = sl + a.gsum[(rg0 + g) * 2 * 64 + 64 + lane];
        cnt += a.gcnt[(rg0 + g) * 64 + lane];
    }
    const size_t ra = ((size_t)pass * a.maxtiles + tile) * nt + t;
    const double* sp = a.spart + ra * 3 * 64 + lane;
    sm = sm + sp[0];
    sl = sl + sp[64];
    cnt += a.scnt[ra * 64 + lane];
    if (valid) {
        const size_t r = ((size_t)pass * n + m) * nt + t;
        a.smass_p[r] = cz * sm;
        a.slum_p[r] = cz * log10(fmax(sl, OPS_SAFE_FLOOR));
        if (a.count && cnt && cz != 0.0) atomicAdd(a.count + 2 * m + pass, cnt);
    }
    const double ws = a.wsfh[(size_t)m * nt + t];
    double* wc64 = (double*)a.wc;
    float* wc32 = (float*)a.wc;
    if (a.u_f32)
        wc32[ra * 64 + lane] = (float)((ws > OPS_SAFE_FLOOR ? ws : 0.0) * cz);
    else
        wc64[ra * 64 + lane] = (ws > OPS_SAFE_FLOOR ? ws : 0.0) * cz;
}
#pragma clang fp contract(off)

// k_op_sums: a wavefront per model (4 a workgroup): each lane the passes' sums of its ages, then model_sums
// (MSUMS_OWN) over the lanes
__global__ __launch_bounds__(256) void k_op_sums(OpSArgs a) {
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= a.n) return;
    const int nt = a.nt, n = a.n;
    double ms = 0.0, ls = 0.0;
    for (int t = lane; t < nt; t += 64) {
        double smass = 0.0, slog = 0.0;
        for (int pass = 0; pass < a.npass; pass++) {
            const size_t r = ((size_t)pass * n + m) * nt + t;
            smass = smass + a.smass_p[r];
            slog = slog + a.slum_p[r];
        }
        const double w = a.wsfh[(size_t)m * nt + t];
        if (w > OPS_SAFE_FLOOR) {
            ms = ms + w * smass;
            ls = ls + w * pow(10.0, slog);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        ms += __shfl_xor(ms, o, 64);
        ls += __shfl_xor(ls, o, 64);
    }
    if (lane == 0) {
        if (a.mass) a.mass[m] = ms;
        if (a.lbol) a.lbol[m] = ls > OPS_SAFE_FLOOR ? log10(ls) : 0.0;
    }
}

#pragma clang fp contract(on)

template <typename T> struct Pair;
template <> struct Pair<float> { typedef float2 type; };
template <> struct Pair<double> { typedef double2 type; };

// ------------------------------------------------------------------------------------------------
// k_op_segments: grid (ceil(nseg_max / 4), max tiles, passes), 256 threads = 4 segments. A segment is at most
// OPS_SEG consecutive entries of one template's list (sorted by age) in one metallicity, so every wavefront has
// about the same work however unevenly the templates are used. The entries (point, age, value) are uniform across
// the wavefront; they are taken eight at a time, then each lane gathers its model's point weights and SFH x blend
// factors for the eight (256-byte rows) before summing young (age <= it_tesc) and old in the kernel precision.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_segments(OpGArgs a) {
    typedef typename Pair<T>::type T2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int sg = blockIdx.x * 4 + wv, tile = a.tile0 + blockIdx.y, pass = blockIdx.z;
    if (tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const int iz = a.npass == 1 ? 0 : td.x + pass;
    if (sg >= a.segoff[iz + 1] - a.segoff[iz]) return;
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int it = a.it[m];
    const int2 rng = a.seg[a.segoff[iz] + sg];
    const T* u = (const T*)a.u + ((size_t)pass * a.utiles + blockIdx.y) * a.npz * 64 + lane;
    const T* wc = (const T*)a.wc + ((size_t)pass * a.maxtiles + tile) * a.nt * 64 + lane;
    // lane l holds entry rng.x + l (a segment has at most 64), then the entries go round by broadcast
    int2 me = make_int2(0, 0);
    T mv = (T)0;
    if (rng.x + lane < rng.y) {
        me = a.te[rng.x + lane];
        if constexpr (std::is_same<T, float>::value)
            mv = a.te_val32[rng.x + lane];
        else
            mv = a.te_val64[rng.x + lane];
    }
    const int ne = rng.y - rng.x;
    T y = (T)0, o = (T)0;
    for (int j = 0; j < ne; j += 8) {
        int px[8], ag[8];
        T val[8], uu[8], ww[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int src = min(j + q, 63);
            px[q] = __shfl(me.x, src, 64);
            ag[q] = __shfl(me.y, src, 64);
            val[q] = j + q < ne ? __shfl(mv, src, 64) : (T)0;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            uu[q] = u[(size_t)px[q] * 64];
            ww[q] = wc[(size_t)ag[q] * 64];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const T x = val[q] * (uu[q] * ww[q]);
            if (ag[q] + 1 <= it) y += x;
            else o += x;
        }
    }
    T2 r;
    r.x = y;
    r.y = o;
    ((T2*)a.part)[(((size_t)pass * a.utiles + blockIdx.y) * a.nseg_max + sg) * 64 + lane] = r;
}

// ------------------------------------------------------------------------------------------------
// k_op_fold: grid (ceil(kpad / 32), max tiles), 256 threads = 4 blocks of 8 templates: per template the sums of
// its segments over both metallicity passes (fold_multi's per-template sum, opensps_ssp_batch.f90:3164-3186),
// written as the B operand of the template product (split BF16 [column][K], or FP64), young and old apart or
// summed.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_fold(OpFArgs a) {
    typedef typename Pair<T>::type T2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e0 = (blockIdx.x * 4 + wv) * OPS_TPL_BLOCK, tile = a.tile0 + blockIdx.y;
    if (e0 >= a.kpad || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    T vy[OPS_TPL_BLOCK], vo[OPS_TPL_BLOCK];
#pragma unroll
    for (int q = 0; q < OPS_TPL_BLOCK; q++) { vy[q] = (T)0; vo[q] = (T)0; }
    for (int pass = 0; pass < a.npass; pass++) {
        const int iz = a.npass == 1 ? 0 : td.x + pass;
        const int* ts = a.tseg + (size_t)iz * (a.ntpl + 1);
        const T2* P = (const T2*)a.part + ((size_t)pass * a.utiles + blockIdx.y) * a.nseg_max * 64 + lane;
#pragma unroll
        for (int q = 0; q < OPS_TPL_BLOCK; q++) {
            const int e = e0 + q;
            if (e >= a.ntpl) break;
            for (int sg = ts[e]; sg < ts[e + 1]; sg++) {
                const T2 r = P[(size_t)sg * 64];
                vy[q] += r.x;
                vo[q] += r.y;
            }
        }
    }
    if (!valid) return;
    for (int part = 0; part < a.parts; part++) {
        const size_t col = (size_t)m * a.parts + part;
        T v[OPS_TPL_BLOCK];
#pragma unroll
        for (int q = 0; q < OPS_TPL_BLOCK; q++) v[q] = a.parts == 1 ? vy[q] + vo[q] : (part == 0 ? vy[q] : vo[q]);
        if constexpr (std::is_same<T, float>::value) {
            bf16x8_t hi, lo;
            uint16_t h[OPS_TPL_BLOCK], l[OPS_TPL_BLOCK];
#pragma unroll
            for (int q = 0; q < OPS_TPL_BLOCK; q++) {
                h[q] = bf16_rn(v[q]);
                l[q] = bf16_rn(v[q] - bf16_to_f32(h[q]));
            }
            __builtin_memcpy(&hi, h, sizeof(hi));
            __builtin_memcpy(&lo, l, sizeof(lo));
            *(bf16x8_t*)(a.whi + col * a.kpad + e0) = hi;
            *(bf16x8_t*)(a.wlo + col * a.kpad + e0) = lo;
        } else {
#pragma unroll
            for (int q = 0; q < OPS_TPL_BLOCK; q++) a.w64[col * a.kpad + e0 + q] = v[q];
        }
    }
}

// The uniform column tiles of one A for every column (the template library, the full basis): width consecutive
// columns each (position order)
__global__ void k_op_tiles(int ncols, int width, int* ntiles, int4* tiles) {
    const int nt = (ncols + width - 1) / width;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) tiles[i] = make_int4(0, i * width, min(width, ncols - i * width), 0);
    if (threadIdx.x == 0) *ntiles = nt;
}

template __global__ void k_op_weights<float>(OpWArgs);
template __global__ void k_op_weights<double>(OpWArgs);
template __global__ void k_op_slow<float>(OpWArgs);
template __global__ void k_op_slow<double>(OpWArgs);
template __global__ void k_op_segments<float>(OpGArgs);
template __global__ void k_op_segments<double>(OpGArgs);
template __global__ void k_op_fold<float>(OpFArgs);
template __global__ void k_op_fold<double>(OpFArgs);
