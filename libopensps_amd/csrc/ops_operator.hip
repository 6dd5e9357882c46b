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

// A value with its derivatives along three directions (forward mode): the reverse pass's IMF rows take the bins off
// the common path, the total mass and the remnants through the same functions as the forward, with this in place of
// double. Comparisons and branches follow the value.
struct D3 {
    double v, d[3];
    __device__ D3() : v(0.0), d{0.0, 0.0, 0.0} {}
    __device__ D3(double x) : v(x), d{0.0, 0.0, 0.0} {}
};
__device__ inline D3 operator+(const D3& a, const D3& b) {
    D3 r(a.v + b.v);
    for (int i = 0; i < 3; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__device__ inline D3 operator-(const D3& a, const D3& b) {
    D3 r(a.v - b.v);
    for (int i = 0; i < 3; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__device__ inline D3 operator-(const D3& a) {
    D3 r(-a.v);
    for (int i = 0; i < 3; i++) r.d[i] = -a.d[i];
    return r;
}
__device__ inline D3 operator*(const D3& a, const D3& b) {
    D3 r(a.v * b.v);
    for (int i = 0; i < 3; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
__device__ inline D3 operator/(const D3& a, const D3& b) {
    D3 r(a.v / b.v);
    for (int i = 0; i < 3; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
__device__ inline bool operator<(const D3& a, const D3& b) { return a.v < b.v; }
__device__ inline bool operator<=(const D3& a, const D3& b) { return a.v <= b.v; }
__device__ inline bool operator>(const D3& a, const D3& b) { return a.v > b.v; }
__device__ inline bool operator>=(const D3& a, const D3& b) { return a.v >= b.v; }
__device__ inline bool operator==(const D3& a, const D3& b) { return a.v == b.v; }
__device__ inline bool operator!=(const D3& a, const D3& b) { return a.v != b.v; }
__device__ inline D3 exp(const D3& a) {
    D3 r(::exp(a.v));
    for (int i = 0; i < 3; i++) r.d[i] = r.v * a.d[i];
    return r;
}
__device__ inline D3 log(const D3& a) {
    D3 r(::log(a.v));
    for (int i = 0; i < 3; i++) r.d[i] = a.d[i] / a.v;
    return r;
}
__device__ inline D3 pow(const D3& a, const D3& b) {
    D3 r(::pow(a.v, b.v));
    const double la = a.v > 0.0 ? ::log(a.v) : 0.0;
    for (int i = 0; i < 3; i++) r.d[i] = r.v * (b.d[i] * la + (a.d[i] != 0.0 ? b.v * a.d[i] / a.v : 0.0));
    return r;
}
__device__ inline D3 fabs(const D3& a) { return a.v < 0.0 ? -a : a; }
__device__ inline D3 fmin(const D3& a, const D3& b) { return b.v < a.v ? b : a; }
__device__ inline D3 fmax(const D3& a, const D3& b) { return b.v > a.v ? b : a; }
__device__ inline double val(double x) { return x; }
__device__ inline double val(const D3& x) { return x.v; }

template <typename S>
struct ImfPT {
    S salp, a1, a2, a3, mdave, lower, upper, lower_bound, mlim_bh, mlim_ns;
};
typedef ImfPT<double> ImfP;

// powerlaw_bin (opensps_imf.f90:754-792): int_lo^hi m^(e-1) dm from plo = lo^e, dl = ln(hi/lo) by Kahan's log1p
template <typename S>
__device__ S d_powerlaw_bin(S plo, S lo, S hi, S e) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    S x = (hi - lo) / lo, u = 1.0 + x, dl;
    if (u == 1.0) dl = x;
    else dl = log(u) * x / (u - 1.0);
    if (e == 0.0) return plo * dl;
    S y = e * dl;
    if (fabs(y) < 0.25) {
        S r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return plo * dl * r;
    }
    return plo * (exp(y) - 1.0) / e;
}

// expm1_series (opensps_imf.f90:641-667)
template <typename S>
__device__ S d_expm1_series(S y) {
    const double I2 = 1.0 / 2, I3 = 1.0 / 3, I4 = 1.0 / 4, I5 = 1.0 / 5, I6 = 1.0 / 6, I7 = 1.0 / 7, I8 = 1.0 / 8,
                 I9 = 1.0 / 9, I10 = 1.0 / 10, I11 = 1.0 / 11, I12 = 1.0 / 12, I13 = 1.0 / 13;
    if (fabs(y) < 0.25) {
        S r = 1.0 + y * I13;
        r = 1.0 + y * I12 * r; r = 1.0 + y * I11 * r; r = 1.0 + y * I10 * r; r = 1.0 + y * I9 * r;
        r = 1.0 + y * I8 * r; r = 1.0 + y * I7 * r; r = 1.0 + y * I6 * r; r = 1.0 + y * I5 * r;
        r = 1.0 + y * I4 * r; r = 1.0 + y * I3 * r; r = 1.0 + y * I2 * r;
        return y * r;
    }
    return exp(y) - 1.0;
}

// integrate_powerlaw (opensps_imf.f90:725-748)
template <typename S>
__device__ S d_integrate_powerlaw(S lo, S hi, S alpha, bool mw, S coeff) {
    if (hi <= lo) return S(0.0);
    const S denom = mw ? 2.0 - alpha : 1.0 - alpha;
    return coeff * d_powerlaw_bin<S>(pow(lo, denom), lo, hi, denom);
}

// imf_integral (opensps_imf.f90:206-315) for the power-law types 0 (Salpeter), 2 (Kroupa), 4 (Dave). Not inlined:
// the weights' hot loop calls it only for bins across a break, and keeps its registers for the common path.
template <typename S>
__device__ __attribute__((noinline)) S d_imf_integral(int type, const ImfPT<S>& p, S m1, S m2, bool mw) {
    const S lo = fmin(m1, m2), hi = fmax(m1, m2);
    S v = 0.0;
    if (hi <= lo) return S(0.0);
    const int t = type % 10;
    if (t == 0) {
        v = d_integrate_powerlaw<S>(lo, hi, p.salp, mw, S(1.0));
    } else if (t == 2) {
        const S c2 = pow(S(0.5), -p.a1 + p.a2);
        S sl = fmax(lo, S(0.08)), sh = fmin(hi, S(0.5));
        if (sh > sl) v = v + d_integrate_powerlaw<S>(sl, sh, p.a1, mw, S(1.0));
        sl = fmax(lo, S(0.5)); sh = fmin(hi, S(1.0));
        if (sh > sl) v = v + d_integrate_powerlaw<S>(sl, sh, p.a2, mw, c2);
        sl = fmax(lo, S(1.0)); sh = hi;
        if (sh > sl) v = v + d_integrate_powerlaw<S>(sl, sh, p.a3, mw, c2);
    } else if (t == 4) {
        S sl = fmax(lo, S(0.08)), sh = fmin(hi, p.mdave);
        if (sh > sl) v = v + d_integrate_powerlaw<S>(sl, sh, p.a1, mw, S(1.0));
        sl = fmax(lo, p.mdave); sh = hi;
        if (sh > sl) v = v + d_integrate_powerlaw<S>(sl, sh, p.a2, mw, pow(p.mdave, -p.a1 + p.a2));
    }
    return v;
}

// add_remnant_mass_p (opensps_stellar_modifications.f90:729-816): black holes, neutron stars, white dwarfs
template <typename S>
__device__ __attribute__((noinline)) S d_remnant_mass(int type, const ImfPT<S>& p, double mmax) {
    const S hi = p.upper, lbh = p.mlim_bh, lns = p.mlim_ns;
    const S norm = d_imf_integral<S>(type, p, p.lower, hi, true);
    S add = 0.0;
    if (norm <= 0.0) return S(0.0);
    S a = fmin(fmax(lbh, S(mmax)), hi);
    if (a < hi) add = add + 0.5 * d_imf_integral<S>(type, p, a, hi, true) / norm;
    if (mmax <= val(lbh)) {
        a = fmin(fmax(lns, S(mmax)), hi);
        const S b = fmin(lbh, hi);
        if (a < b) add = add + 1.4 * d_imf_integral<S>(type, p, a, b, false) / norm;
    }
    if (mmax <= val(lns)) {
        a = fmin(S(mmax), hi);
        const S b = fmin(lns, hi);
        if (a < b)
            add = add + (0.48 * d_imf_integral<S>(type, p, a, b, false)) / norm +
                  (0.077 * d_imf_integral<S>(type, p, a, b, true)) / norm;
    }
    return add;
}

// The state of imf_bin_weights_tab over one piece: the current power-law segment and the chain of shared edges
template <typename S>
struct BinStateT {
    int cur;
    S e, c, d, ce, prev_hi, prev_phi;
};
typedef BinStateT<double> BinState;

template <typename S>
__device__ __forceinline__ void bin_reset(BinStateT<S>& s) {
    s.cur = 0; s.e = 0.0; s.c = 0.0; s.d = 1.0; s.ce = 0.0; s.prev_hi = -1.0; s.prev_phi = 0.0;
}

// A new power-law segment's constants (imf_bin_weights_tab, opensps_imf.f90:500-525), once per segment change
template <typename S>
__device__ __attribute__((noinline)) void bin_segment(int t, const ImfPT<S>& p, int seg, BinStateT<S>& s) {
    S alpha, coeff = 1.0;
    if (t == 0) {
        alpha = p.salp;
    } else if (t == 2) {
        alpha = seg == 1 ? p.a1 : (seg == 2 ? p.a2 : p.a3);
        if (seg > 1) coeff = pow(S(0.5), -p.a1 + p.a2);
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
template <typename S>
__device__ __attribute__((noinline)) S bin_first(const BinStateT<S>& s, S lo, S hi) {
    if (fabs(s.d) <= 1.0e-12) return s.c * log(hi / lo);
    return s.c * (pow(hi, s.e) - pow(lo, s.e)) / s.d;
}

// One bin of imf_bin_weights_tab (opensps_imf.f90:466-557), FP64 build: stars per unit initial mass. The
// segment classification and the limits decide which bins weigh nothing, exactly as the reference.
template <typename S>
__device__ __attribute__((noinline)) S d_bin_weight(int type, const ImfPT<S>& p, S inv_total, double mini, double m1,
                                                    double m2, double llo, double dl, int first, BinStateT<S>& s) {
    if (mini < val(p.lower) || mini > val(p.upper)) return S(0.0);
    const S lo = first == 1 ? p.lower_bound : S(m1), hi = m2;
    if (hi <= lo) return S(0.0);
    const int t = type % 10;
    const double lov = val(lo), hiv = val(hi);
    int seg = 0;
    if (t == 0) {
        seg = 1;
    } else if (t == 2) {
        if (lov >= 0.08 && hiv <= 0.5) seg = 1;
        else if (lov >= 0.5 && hiv <= 1.0) seg = 2;
        else if (lov >= 1.0) seg = 3;
    } else if (t == 4) {
        if (lov >= 0.08 && hiv <= val(p.mdave)) seg = 1;
        else if (lov >= val(p.mdave)) seg = 2;
    }
    if (seg > 0 && seg != s.cur) bin_segment<S>(t, p, seg, s);
    S v;
    if (seg > 0 && s.cur == seg) {
        if (first == 1) {
            v = bin_first<S>(s, lo, hi);
        } else if (s.e == 0.0) {
            v = s.c * dl;
        } else {
            const S plo = (lo == s.prev_hi) ? s.prev_phi : exp(s.e * llo);
            const S em1 = d_expm1_series<S>(s.e * dl);
            v = s.ce * plo * em1;
            s.prev_hi = hi;
            s.prev_phi = plo + plo * em1;
        }
    } else {
        v = d_imf_integral<S>(type, p, lo, hi, false);
        s.prev_hi = -1.0;
        if (seg == 0) s.cur = 0;
    }
    return v * inv_total;
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
    // the segment's e and c, selected when the (uniform) segment changes: the points are sorted by mass, so it
    // changes at most at the IMF's breaks (a branch the wavefront takes rarely, not selects on every point)
    int cur = 0;
    double e = 0.0, c = 0.0;
    // one point a step; the next point's record is fetched (scalar loads) before this point's arithmetic, so its
    // latency hides behind it
    OpsWPoint nxt;
    if (k0 < k1) nxt = wp_load(wp + k0);
    for (int k = k0; k < k1; k++) {
        const OpsWPoint r = nxt;
        if (k + 1 < k1) nxt = wp_load(wp + k + 1);
        const int seg = r.code & 3;
        if (seg == 0) continue;   // k_op_slow's bin
        if (__builtin_expect(seg != cur, 0)) {
            cur = seg;
            e = seg == 1 ? e1 : (seg == 2 ? e2 : e3);
            c = seg == 1 ? c1 : (seg == 2 ? c2 : c3);
        }
        const double dl = r.dl;
        const double plo = (r.code & 4) ? phi : exp(e * r.llo);
        const double y = e * dl;
        double sx = d_expm1_over_y(y, pc);
        if (!(fabs(y) < 0.25)) sx = (exp(y) - 1.0) / y;
        phi = fma(plo, y * sx, plo);
        const double w = (r.mini < lower || r.mini > upper) ? 0.0 : c * dl * plo * sx;
        __builtin_nontemporal_store((T)w, u + (size_t)(k - zo) * 64);   // streamed: read back by the segments
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
        if (a.sraw) {
            a.sraw[2 * r] = sm;
            a.sraw[2 * r + 1] = sl;
        }
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
        if (a.lsum) a.lsum[m] = ls;
        a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_LSUM] = ls;   // (the AGN's luminosity: k_emission)
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

// ================================================================================================
// The reverse pass of the IMF per model (ssp_batch_age_sums_adjoint, src/physics/opensps_ssp_batch.f90:991-1248,
// with adj_template_bar, adj_fold_gather, adj_lq, adj_mass_dot and adj_imf_analytic). The template cotangents
// v_b = etpl y_b (the transposed product, ops_api.hip) arrive transposed to (part, template, position), so a
// wavefront's lanes (the models of a column tile) read each template's cotangents as one contiguous run. Per point
// k of an age: g_k = sum_j val(k, j) v_b(tpl(k, j)); per (pass, age, model) the fold's sum F = sum_k w_k g_k and the
// IMF sums sum_k dw_k (x g_k + A cm_k + B cl_k) (+ A drem), x the SFH weight times the blend factor, A and B the
// mass and L_bol cotangents per point; k_op_adj_out turns them into w_bar, the metallicity's and the IMF's rows.
// Directions of the IMF rows: Kroupa (2) imf1, imf2, imf3; Dave (4) imf1, imf2, mdave; Salpeter none.
// ================================================================================================
#pragma clang fp contract(on)

// The model's IMF with the directions seeded (value types as model_imf)
__device__ ImfPT<D3> model_imf_d3(const OpAdjArgs& a, int m) {
    const int n = a.n, t = a.imf_type % 10;
    ImfPT<D3> p;
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
    if (t == 2 || t == 4) {
        p.a1.d[0] = 1.0;
        p.a2.d[1] = 1.0;
        if (t == 2) p.a3.d[2] = 1.0;
        else p.mdave.d[2] = 1.0;
    }
    return p;
}

// k_op_imf_grad: one thread per model: d ln(c_seg / total) along the directions for each power-law segment (the
// common bins' weights are c_seg / total x a function of e_seg = 1 - alpha_seg and the bin alone)
__global__ void k_op_imf_grad(OpAdjArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const ImfPT<D3> p = model_imf_d3(a, m);
    const int t = a.imf_type % 10;
    const D3 tot = d_imf_integral<D3>(a.imf_type, p, p.lower, p.upper, true);
    double* o = a.impd + (size_t)m * OPS_IMFD_N;
    for (int seg = 1; seg <= 3; seg++) {
        const bool on = !(t == 0 && seg > 1) && !(t == 4 && seg > 2);
        BinStateT<D3> s;
        bin_reset(s);
        if (on) bin_segment<D3>(t, p, seg, s);
        for (int i = 0; i < 3; i++) {
            double v = 0.0;
            if (on && s.c.v != 0.0) v = s.c.d[i] / s.c.v - (tot.v > 0.0 ? tot.d[i] / tot.v : 0.0);
            o[3 * (seg - 1) + i] = v;
        }
    }
}

// k_op_adj_t: grid (ceil(ntpl / 64), ceil(n / 64), 2 parts), 256 threads: C (column m * 2 + part, template) with
// row stride ldc -> CT (part, template, position) through a 64 x 64 LDS tile, the positions' models from order
template <typename T>
__global__ __launch_bounds__(256) void k_op_adj_t(int n, int ntpl, int ldc, const int* order, const T* C, T* CT) {
    __shared__ T tile[64][65];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e0 = blockIdx.x * 64, p0 = blockIdx.y * 64, part = blockIdx.z;
    for (int r = wv; r < 64; r += 4) {
        const int pos = p0 + r, e = e0 + lane;
        T v = (T)0;
        if (pos < n && e < ntpl) v = C[((size_t)order[pos] * 2 + part) * ldc + e];
        tile[r][lane] = v;
    }
    __syncthreads();
    for (int r = wv; r < 64; r += 4) {
        const int e = e0 + r, pos = p0 + lane;
        if (e < ntpl && pos < n) CT[((size_t)part * ntpl + e) * n + pos] = tile[lane][r];
    }
}

// sx(y) = expm1(y) / y and h(y) = sx'(y) = (e^y (y - 1) + 1) / y^2 = int_0^1 s e^(y s) ds. Below |y| = 1/4 both from
// the series of d_expm1_over_y (the same FMAs for sx) by Horner's scheme with its derivative alongside (h's terms
// (j + 1) / (j + 2)! = 1 / (j! (j + 2))); above it the closed forms.
__device__ __forceinline__ void d_sx_h(double y, const double* pc, double& sx, double& h) {
    double p = pc[12], dp = 0.0;
#pragma unroll
    for (int j = 11; j >= 0; j--) {
        dp = fma(dp, y, p);
        p = fma(p, y, pc[j]);
    }
    sx = p;
    h = dp;
    if (!(fabs(y) < 0.25)) {
        const double ey = exp(y);
        sx = (ey - 1.0) / y;
        h = (ey * (y - 1.0) + 1.0) / (y * y);
    }
}

// A lane's model in the reverse pass's point kernels: its position in the tile, the blend factor of the pass, the
// template cotangents' row of its part, and the factors of the points' cotangent ge = x0 gk + A cm + B cl (the
// fold through the SFH weight x0, the mass and the log-blended L_bol sums)
template <typename T>
struct AdjLane {
    int m, pos;
    bool valid;
    double cz, x0, A, B, lower, upper;
    const T* ct;
};

template <typename T>
__device__ __forceinline__ AdjLane<T> adj_lane(const OpAdjArgs& a, int4 td, int lane, int t, int pass) {
    AdjLane<T> L;
    const int n = a.n, nt = a.nt;
    L.valid = lane < td.z;
    L.pos = td.y + (L.valid ? lane : 0);
    L.m = a.order[L.pos];
    const int m = L.m;
    const double dz = a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ];
    L.cz = a.npass == 1 ? 1.0 : (pass == 0 ? 1.0 - dz : dz);
    const int part = t + 1 <= a.it[m] ? 0 : 1;
    L.ct = (const T*)a.CT + (size_t)part * a.ntpl * n + L.pos;
    const double* go = a.gout + (size_t)m * OPS_GOUT_N;
    const double cs = go[OPS_GOUT_C];
    const double mbs = a.mass_bar ? a.mass_bar[m] / cs : 0.0, lsbs = go[OPS_GOUT_LSB];
    const double ws = a.wsfh[(size_t)m * nt + t];
    L.x0 = (ws > OPS_SAFE_FLOOR ? ws : 0.0) * L.cz;
    double slog = 0.0;
    for (int q = 0; q < a.npass; q++) slog += a.slum_p[((size_t)q * n + m) * nt + t];
    const double sl = a.sraw[(((size_t)pass * n + m) * nt + t) * 2 + 1];
    L.A = mbs * L.x0;
    L.B = sl > OPS_SAFE_FLOOR ? lsbs * L.x0 * pow(10.0, slog) / sl : 0.0;
    L.lower = a.prm[(size_t)OPS_P_IMF_LOWER * n + m];
    L.upper = a.prm[(size_t)OPS_P_IMF_UPPER * n + m];
    return L;
}

// k_op_adj_points: grid (nt x ceil(ngroup / 4), tiles, passes), 256 threads = 4 point groups of one age (OPS_OP_GPTS
// points each, whole pieces as k_op_weights takes them, so each group starts its chain of shared edges afresh); the
// lanes are a column tile's models. The common bins as k_op_weights forms them (records through scalar loads, the
// next point's record and entries fetched ahead of this point's arithmetic, the point's template cotangents
// gathered before it), each with its derivative along e: dw/de = c / total lo^e dl (ln lo sx(e dl) + dl h(e dl)),
// sx = expm1(y) / y; along a direction dw = w d ln(c / total) - [the segment's slope is the direction] dw/de. The
// four groups' sums are added in order through LDS and written per (pass, age, block of groups, quantity, position).
#define OPS_ADJ_NENT 2   // entries per point fetched ahead (points with more take the rest in place)
template <typename T>
__global__ __launch_bounds__(256) void k_op_adj_points(OpAdjArgs a) {
    __shared__ double red[4][4][64];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ngb = (a.ngroup + 3) / 4;
    const int t = blockIdx.x / ngb, gb = blockIdx.x - t * ngb, grp = gb * 4 + wv;
    const int tile = blockIdx.y, pass = blockIdx.z;
    if (t >= a.nt || tile >= *a.ntiles) return;   // (uniform over the workgroup)
    const int4 td = a.tiles[tile];
    const int iz = a.npass == 1 ? 0 : __builtin_amdgcn_readfirstlane(td.x) + pass;
    const int nt = a.nt, n = a.n, nent = a.nent, ty = a.imf_type % 10;
    const AdjLane<T> L = adj_lane<T>(a, td, lane, t, pass);
    // the model's per-segment e, c and d ln(c / total)
    const double* ic = a.impc + (size_t)L.m * OPS_IMFC_N;
    const double e1 = ic[OPS_IMFC_E], e2 = ic[OPS_IMFC_E + 1], e3 = ic[OPS_IMFC_E + 2];
    const double c1 = ic[OPS_IMFC_CT], c2 = ic[OPS_IMFC_CT + 1], c3 = ic[OPS_IMFC_CT + 2];
    double dc[OPS_IMFD_N];
#pragma unroll
    for (int q = 0; q < OPS_IMFD_N; q++) dc[q] = a.impd[(size_t)L.m * OPS_IMFD_N + q];
    // the direction of each segment's slope (Kroupa: segment s has alpha_s; Dave: alpha_1, alpha_2; Salpeter: none)
    const int sd1 = (ty == 2 || ty == 4) ? 0 : -1, sd2 = (ty == 2 || ty == 4) ? 1 : -1, sd3 = ty == 2 ? 2 : -1;
    const int ka = __builtin_amdgcn_readfirstlane(a.age_ptr[iz * nt + t]);
    const int kb = __builtin_amdgcn_readfirstlane(a.age_ptr[iz * nt + t + 1]);
    const int k0 = min(kb, ka + grp * OPS_OP_GPTS), k1 = min(kb, k0 + OPS_OP_GPTS);
    typedef __attribute__((address_space(4))) const double* DC;
    typedef __attribute__((address_space(4))) const int* IC;
    const OpsWPointC wp = (OpsWPointC)a.wp;
    const DC pcs = (DC)a.poly, pv = (DC)a.pval;
    const IC pe = (IC)a.pent;
    double pc[13];
#pragma unroll
    for (int j = 0; j < 13; j++) pc[j] = pcs[j];
    const int nf = min(nent, OPS_ADJ_NENT);
    double F = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, phi = 1.0;
    // per run of one segment (uniform; the points are sorted by mass, so a run ends at most at the IMF's breaks):
    // its e and c, and S = sum ge w, D = sum ge dw/de; at the run's end g_i += d ln(c / total)_i S - [the segment's
    // slope is direction i] D
    int cur = 0;
    double e = 0.0, c = 0.0, S = 0.0, D = 0.0;
#define OPS_ADJ_FLUSH                                                                                                  \
    if (cur != 0) {                                                                                                    \
        const double d0 = cur == 1 ? dc[0] : (cur == 2 ? dc[3] : dc[6]);                                               \
        const double d1 = cur == 1 ? dc[1] : (cur == 2 ? dc[4] : dc[7]);                                               \
        const double d2 = cur == 1 ? dc[2] : (cur == 2 ? dc[5] : dc[8]);                                               \
        const int sd = cur == 1 ? sd1 : (cur == 2 ? sd2 : sd3);                                                        \
        g0 = fma(d0, S, g0) - (sd == 0 ? D : 0.0);                                                                     \
        g1 = fma(d1, S, g1) - (sd == 1 ? D : 0.0);                                                                     \
        g2 = fma(d2, S, g2) - (sd == 2 ? D : 0.0);                                                                     \
        S = 0.0;                                                                                                       \
        D = 0.0;                                                                                                       \
    }
    OpsWPoint nxt;
    int ne[OPS_ADJ_NENT];
    double nv[OPS_ADJ_NENT];
    if (k0 < k1) {
        nxt = wp_load(wp + k0);
#pragma unroll
        for (int j = 0; j < OPS_ADJ_NENT; j++) {
            ne[j] = j < nf ? pe[(size_t)k0 * nent + j] : -1;
            nv[j] = j < nf ? pv[(size_t)k0 * nent + j] : 0.0;
        }
    }
    for (int k = k0; k < k1; k++) {
        const OpsWPoint r = nxt;
        // the point's template cotangents (a contiguous run of the tile's models each), gathered first and without a
        // branch (an absent entry reads template 0's row with the value 0), so their waits fall at their use
        T cv[OPS_ADJ_NENT];
        double cw[OPS_ADJ_NENT];
#pragma unroll
        for (int j = 0; j < OPS_ADJ_NENT; j++) {
            cv[j] = L.ct[(size_t)max(ne[j], 0) * n];
            cw[j] = ne[j] >= 0 ? nv[j] : 0.0;
        }
        if (k + 1 < k1) {
            nxt = wp_load(wp + k + 1);
#pragma unroll
            for (int j = 0; j < OPS_ADJ_NENT; j++) {
                ne[j] = j < nf ? pe[(size_t)(k + 1) * nent + j] : -1;
                nv[j] = j < nf ? pv[(size_t)(k + 1) * nent + j] : 0.0;
            }
        }
        const int seg = r.code & 3;
        if (seg == 0) continue;   // off the common path: k_op_adj_slow
        if (__builtin_expect(seg != cur, 0)) {
            OPS_ADJ_FLUSH
            cur = seg;
            e = seg == 1 ? e1 : (seg == 2 ? e2 : e3);
            c = seg == 1 ? c1 : (seg == 2 ? c2 : c3);
        }
        const double dl = r.dl;
        const double plo = (r.code & 4) ? phi : exp(e * r.llo);
        const double y = e * dl;
        double sx, h;
        d_sx_h(y, pc, sx, h);
        phi = fma(plo, y * sx, plo);
        if (r.mini < L.lower || r.mini > L.upper) continue;
        const double w = c * dl * plo * sx;
        const double dwde = c * plo * dl * fma(dl, h, r.llo * sx);
        double gk = 0.0;
#pragma unroll
        for (int j = 0; j < OPS_ADJ_NENT; j++) gk = fma(cw[j], (double)cv[j], gk);
        for (int j = OPS_ADJ_NENT; j < nent; j++) {
            const int ej = pe[(size_t)k * nent + j];
            if (ej >= 0) gk = fma(pv[(size_t)k * nent + j], (double)L.ct[(size_t)ej * n], gk);
        }
        const double ge = L.x0 * gk + L.A * r.cm + L.B * r.cl;
        F = fma(w, gk, F);
        S = fma(ge, w, S);
        D = fma(ge, dwde, D);
    }
    OPS_ADJ_FLUSH
#undef OPS_ADJ_FLUSH
    red[wv][0][lane] = F;
    red[wv][1][lane] = g0;
    red[wv][2][lane] = g1;
    red[wv][3][lane] = g2;
    __syncthreads();
    if (wv == 0 && L.valid) {
        double* o = a.gpart + (((size_t)pass * nt + t) * ngb + gb) * 4 * n + L.pos;
#pragma unroll
        for (int q = 0; q < 4; q++) o[(size_t)q * n] = ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) +
                                                       red[3][q][lane];
    }
}

// k_op_adj_slow: grid (ceil(nt / 4), tiles, passes), a wavefront per (age, column tile, pass) as k_op_slow: the
// bins off the common path and the remnants by the forward's functions on D3, added to k_op_adj_points's blocks
// of groups in order; the (pass, age, position) sums for k_op_adj_out (zero for a pass of zero weight, as the
// forward's grouping leaves it out).
template <typename T>
__global__ __launch_bounds__(256) void k_op_adj_slow(OpAdjArgs a) {
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = blockIdx.x * 4 + wv, tile = blockIdx.y, pass = blockIdx.z;
    if (t >= a.nt || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const int iz = a.npass == 1 ? 0 : __builtin_amdgcn_readfirstlane(td.x) + pass;
    const int n = a.n, nt = a.nt, nent = a.nent;
    const int ngb = (a.ngroup + 3) / 4;
    const AdjLane<T> L = adj_lane<T>(a, td, lane, t, pass);
    double F = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const double* gp = a.gpart + ((size_t)pass * nt + t) * ngb * 4 * n + L.pos;
    for (int b = 0; b < ngb; b++) {
        F = F + gp[((size_t)b * 4 + 0) * n];
        g0 = g0 + gp[((size_t)b * 4 + 1) * n];
        g1 = g1 + gp[((size_t)b * 4 + 2) * n];
        g2 = g2 + gp[((size_t)b * 4 + 3) * n];
    }
    const int ka = a.age_ptr[iz * nt + t], kb = a.age_ptr[iz * nt + t + 1];
    const int j0 = a.slow_ptr[iz * nt + t], j1 = a.slow_ptr[iz * nt + t + 1];
    const bool rem = a.remnants && kb > ka;
    if (j1 > j0 || rem) {
        const ImfPT<D3> p = model_imf_d3(a, L.m);
        const D3 tot = d_imf_integral<D3>(a.imf_type, p, p.lower, p.upper, true);
        const D3 inv = tot.v > 0.0 ? D3(1.0) / tot : D3(1.0);
        for (int j = j0; j < j1; j++) {
            const int k = a.slow_pt[j];
            const double* q = a.pt + (size_t)k * OPS_PT_N;
            BinStateT<D3> s;
            bin_reset(s);
            const D3 w = d_bin_weight<D3>(a.imf_type, p, inv, q[0], q[1], q[2], q[3], q[4], a.first[k], s);
            double gk = 0.0;
            for (int jj = 0; jj < nent; jj++) {
                const int e = a.pent[(size_t)k * nent + jj];
                if (e >= 0) gk = fma(a.pval[(size_t)k * nent + jj], (double)L.ct[(size_t)e * n], gk);
            }
            const double ge = L.x0 * gk + L.A * q[5] + L.B * q[6];
            F = fma(w.v, gk, F);
            g0 = fma(ge, w.d[0], g0);
            g1 = fma(ge, w.d[1], g1);
            g2 = fma(ge, w.d[2], g2);
        }
        if (rem) {
            const D3 r = d_remnant_mass<D3>(a.imf_type, p, a.age_mmax[iz * nt + t]);
            g0 = fma(L.A, r.d[0], g0);
            g1 = fma(L.A, r.d[1], g1);
            g2 = fma(L.A, r.d[2], g2);
        }
    }
    if (!L.valid) return;
    if (L.cz == 0.0) {
        F = 0.0; g0 = 0.0; g1 = 0.0; g2 = 0.0;
    }
    double* o = a.part + (((size_t)pass * nt + t) * n + L.pos) * 4;
    o[0] = F;
    o[1] = g0;
    o[2] = g1;
    o[3] = g2;
}

// k_op_adj_out: a wavefront per position (4 a workgroup), lanes over the ages: w_bar (the fold's part plus the
// per-age L_bol and mass terms), the metallicity weight's cotangent (the fold, the mass and the log-blended L_bol of
// both passes), the IMF rows, and params_bar as k_grad_out writes it.
__global__ __launch_bounds__(256) void k_op_adj_out(OpAdjArgs a) {
    const int lane = threadIdx.x & 63;
    const int pos = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pos >= a.n) return;
    const int n = a.n, nt = a.nt, m = a.order[pos];
    const double* g = a.gout + (size_t)m * OPS_GOUT_N;
    const double c = g[OPS_GOUT_C], lsb = g[OPS_GOUT_LSB] * c;
    const double mb = a.mass_bar ? a.mass_bar[m] : 0.0;
    const double dz = a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ];
    double dzb = 0.0, i0 = 0.0, i1 = 0.0, i2 = 0.0;
    for (int t = lane; t < nt; t += 64) {
        const double ws = a.wsfh[(size_t)m * nt + t];
        const double x = ws > OPS_SAFE_FLOOR ? ws : 0.0;
        double smass = 0.0, slog = 0.0;
        for (int q = 0; q < a.npass; q++) {
            smass += a.smass_p[((size_t)q * n + m) * nt + t];
            slog += a.slum_p[((size_t)q * n + m) * nt + t];
        }
        const double slum = pow(10.0, slog);
        double wf = 0.0;
        for (int q = 0; q < a.npass; q++) {
            const double cz = a.npass == 1 ? 1.0 : (q == 0 ? 1.0 - dz : dz);
            const double* P = a.part + (((size_t)q * nt + t) * n + pos) * 4;
            const double F = P[0] * c;
            wf += cz * F;
            i0 += P[1];
            i1 += P[2];
            i2 += P[3];
            if (a.npass == 2 && cz != 0.0) {
                const double* sr = a.sraw + (((size_t)q * n + m) * nt + t) * 2;
                dzb += (q == 0 ? -1.0 : 1.0) * x * (F + mb * sr[0] + lsb * slum * log(fmax(sr[1], OPS_SAFE_FLOOR)));
            }
        }
        if (a.weights_bar) a.weights_bar[(size_t)m * nt + t] = wf + lsb * slum + mb * smass;
    }
    for (int o = 32; o > 0; o >>= 1) {
        dzb += __shfl_xor(dzb, o, 64);
        i0 += __shfl_xor(i0, o, 64);
        i1 += __shfl_xor(i1, o, 64);
        i2 += __shfl_xor(i2, o, 64);
    }
    if (lane == 0 && a.params_bar) {
        double* pb = a.params_bar + (size_t)m * OPENSPS_HIP_NBGRAD_ROWS;
        for (int r = 0; r < OPENSPS_HIP_NBGRAD_ROWS; r++) pb[r] = 0.0;
        for (int q = 0; q < 7; q++) pb[1 + q] = g[OPS_GOUT_D + q] * c;   // dust1 .. frac_nodust (BGRAD 2..8)
        const int zl = a.zlo[m];
        if (a.nz > 1) pb[15] = dzb / (a.logz[zl] - a.logz[zl - 1]);    // logzsol (BGRAD 16)
        pb[22] = (g[OPS_GOUT_DZO] + g[OPS_GOUT_DZI]) * c + g[OPS_GOUT_DZD];  // zred (BGRAD 23)
        pb[23] = g[OPS_GOUT_DSG] * c;                                        // sigma_smooth (BGRAD 24)
        pb[32] = g[OPS_GOUT_DIGM] * c;                                       // igm_factor (BGRAD 33)
        const int ty = a.imf_type % 10;
        if (ty == 2 || ty == 4) {                                           // imf1, imf2 (BGRAD 27, 28)
            pb[26] = i0 * c;
            pb[27] = i1 * c;
            pb[ty == 2 ? 28 : 30] = i2 * c;                                  // imf3 (29) or mdave (31)
        }
    }
}
#pragma clang fp contract(off)

template __global__ void k_op_adj_t<float>(int, int, int, const int*, const float*, float*);
template __global__ void k_op_adj_t<double>(int, int, int, const int*, const double*, double*);
template __global__ void k_op_adj_points<float>(OpAdjArgs);
template __global__ void k_op_adj_points<double>(OpAdjArgs);
template __global__ void k_op_adj_slow<float>(OpAdjArgs);
template __global__ void k_op_adj_slow<double>(OpAdjArgs);
template __global__ void k_op_weights<float>(OpWArgs);
template __global__ void k_op_weights<double>(OpWArgs);
template __global__ void k_op_slow<float>(OpWArgs);
template __global__ void k_op_slow<double>(OpWArgs);
template __global__ void k_op_segments<float>(OpGArgs);
template __global__ void k_op_segments<double>(OpGArgs);
template __global__ void k_op_fold<float>(OpFArgs);
template __global__ void k_op_fold<double>(OpFArgs);
