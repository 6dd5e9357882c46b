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
//   k_op_fold     one wavefront per (tile, block of 8 templates): the entries of each template (CSR by
//                 template, sorted by age) times the points' weights and the SFH weights, split young / old
//                 at it_tesc (fold_multi, opensps_ssp_batch.f90:3095-3188), both metallicity passes summed;
//                 written as the B operand of the template product (split BF16 [column][K] or FP64)
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

// imf_integral (opensps_imf.f90:206-315) for the power-law types 0 (Salpeter), 2 (Kroupa), 4 (Dave)
__device__ double d_imf_integral(int type, const ImfP& p, double m1, double m2, bool mw) {
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
__device__ double d_remnant_mass(int type, const ImfP& p, double mmax) {
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

// One bin of imf_bin_weights_tab (opensps_imf.f90:466-557), FP64 build: stars per unit initial mass. The
// segment classification and the limits decide which bins weigh nothing, exactly as the reference.
__device__ double d_bin_weight(int type, const ImfP& p, double inv_total, double mini, double m1, double m2,
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
    if (seg > 0 && seg != s.cur) {
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
    double val;
    if (seg > 0 && s.cur == seg) {
        if (first == 1) {
            if (fabs(s.d) <= 1.0e-12) val = s.c * log(hi / lo);
            else val = s.c * (pow(hi, s.e) - pow(lo, s.e)) / s.d;
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

// ------------------------------------------------------------------------------------------------
// k_op_weights: grid (ceil(nt / 4), max tiles, passes), 256 threads = 4 ages. Pass 0 takes the operator of
// metallicity zlo with weight 1 - dz, pass 1 that of zlo + 1 with weight dz (one pass with one metallicity).
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_weights(OpWArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + wv, tile = blockIdx.y, pass = blockIdx.z;
    if (t >= a.nt || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];   // {metallicity pair (zlo - 1), first column, columns}
    const int iz = a.npass == 1 ? 0 : td.x + pass;
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int n = a.n, nt = a.nt;
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
    const double dz = a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ];
    const double cz = a.npass == 1 ? 1.0 : (pass == 0 ? 1.0 - dz : dz);
    const double tot = d_imf_integral(a.imf_type, p, p.lower, p.upper, true);
    const double inv_total = tot > 0.0 ? 1.0 / tot : 1.0;
    const int k0 = a.age_ptr[iz * nt + t], k1 = a.age_ptr[iz * nt + t + 1], zo = a.zoff[iz];
    T* u = (T*)a.u + ((size_t)pass * a.maxtiles + tile) * a.npz * 64 + lane;
    double sm = 0.0, sl = 0.0, sq = 0.0;
    int cnt = 0;
    BinState s;
    for (int q0 = k0; q0 < k1; q0 += OPS_PIECE) {
        const int q1 = min(k1, q0 + OPS_PIECE);
        bin_reset(s);
        double pm = 0.0, pl = 0.0, pq = 0.0;
        for (int k = q0; k < q1; k++) {
            const double* pt = a.pt + (size_t)k * OPS_PT_N;
            const double w = d_bin_weight(a.imf_type, p, inv_total, pt[0], pt[1], pt[2], pt[3], pt[4], a.first[k], s);
            u[(size_t)(k - zo) * 64] = (T)w;
            cnt += w > 0.0 ? 1 : 0;
            pm = pm + pt[5] * w;
            pl = pl + pt[6] * w;
            pq = pq + pt[7] * w;
        }
        sm = sm + pm; sl = sl + pl; sq = sq + pq;
    }
    if (a.remnants && k1 > k0) sm = sm + d_remnant_mass(a.imf_type, p, a.age_mmax[iz * nt + t]);
    if (valid) {
        const size_t r = ((size_t)pass * n + m) * nt + t;
        a.smass_p[r] = cz * sm;
        a.slum_p[r] = cz * log10(fmax(sl, OPS_SAFE_FLOOR));
        if (a.count && cnt && cz != 0.0) atomicAdd(a.count + 2 * m + pass, cnt);
    }
    const double ws = a.wsfh[(size_t)m * nt + t];
    T* wc = (T*)a.wc + (((size_t)pass * a.maxtiles + tile) * nt + t) * 64 + lane;
    *wc = (T)((ws > OPS_SAFE_FLOOR ? ws : 0.0) * cz);
    (void)sq;   // the ionizing photons serve nebular emission, which this build does not add
}

// k_op_sums: one thread per model; the passes' per-age sums, then model_sums (MSUMS_OWN)
__global__ void k_op_sums(OpSArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n) return;
    const int nt = a.nt, n = a.n;
    double ms = 0.0, ls = 0.0;
    for (int t = 0; t < nt; t++) {
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
    if (a.mass) a.mass[m] = ms;
    if (a.lbol) a.lbol[m] = ls > OPS_SAFE_FLOOR ? log10(ls) : 0.0;
}

#pragma clang fp contract(on)

// ------------------------------------------------------------------------------------------------
// k_op_fold: grid (ceil(kpad / 32), max tiles), 256 threads = 4 blocks of 8 templates. Per template and pass
// the entries (point, age, value) are uniform across the wavefront; each lane reads its model's point weight
// and SFH x blend factor (256-byte rows) and sums young (age <= it_tesc) and old in the kernel precision.
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_op_fold(OpFArgs a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e0 = (blockIdx.x * 4 + wv) * OPS_TPL_BLOCK, tile = blockIdx.y;
    if (e0 >= a.kpad || tile >= *a.ntiles) return;
    const int4 td = a.tiles[tile];
    const bool valid = lane < td.z;
    const int m = a.order[td.y + (valid ? lane : 0)];
    const int it = a.it[m], nt = a.nt;
    T vy[OPS_TPL_BLOCK], vo[OPS_TPL_BLOCK];
#pragma unroll
    for (int q = 0; q < OPS_TPL_BLOCK; q++) { vy[q] = (T)0; vo[q] = (T)0; }
    for (int pass = 0; pass < a.npass; pass++) {
        const int iz = a.npass == 1 ? 0 : td.x + pass;
        const T* u = (const T*)a.u + ((size_t)pass * a.maxtiles + tile) * a.npz * 64 + lane;
        const T* wc = (const T*)a.wc + ((size_t)pass * a.maxtiles + tile) * nt * 64 + lane;
        const int* ptr = a.tz_ptr + (size_t)iz * (a.ntpl + 1);
#pragma unroll
        for (int q = 0; q < OPS_TPL_BLOCK; q++) {
            const int e = e0 + q;
            if (e >= a.ntpl) break;
            const int j0 = ptr[e], j1 = ptr[e + 1];
            T y = (T)0, o = (T)0;
            for (int j = j0; j < j1; j++) {
                const int2 pe = a.te[j];
                T val;
                if constexpr (std::is_same<T, float>::value)
                    val = a.te_val32[j];
                else
                    val = a.te_val64[j];
                const T x = val * (u[(size_t)pe.x * 64] * wc[(size_t)pe.y * 64]);
                if (pe.y + 1 <= it) y += x;
                else o += x;
            }
            vy[q] += y;
            vo[q] += o;
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

// The template product's column tiles: 64 consecutive columns each (position order), one library for all
__global__ void k_op_tiles(int ncols, int* ntiles, int4* tiles) {
    const int nt = (ncols + OPS_TILE_COL - 1) / OPS_TILE_COL;
    for (int i = threadIdx.x; i < nt; i += blockDim.x)
        tiles[i] = make_int4(0, i * OPS_TILE_COL, min(OPS_TILE_COL, ncols - i * OPS_TILE_COL), 0);
    if (threadIdx.x == 0) *ntiles = nt;
}

template __global__ void k_op_weights<float>(OpWArgs);
template __global__ void k_op_weights<double>(OpWArgs);
template __global__ void k_op_fold<float>(OpFArgs);
template __global__ void k_op_fold<double>(OpFArgs);
