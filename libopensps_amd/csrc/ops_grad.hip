// The reverse pass of the batch (batch_backward, src/physics/opensps_csp_batch.f90:4713-5136) on gfx950:
//   k_grad_chain  one workgroup per model: the magnitudes' and the spectrum's cotangents back through the
//                 band integrals, the observed frame, the velocity smoothing, the IGM and the attenuation, with
//                 the forward's spectra and transmissions recomputed (not stored), to the cotangents of the age
//                 sums (the rows of the transposed product) and the per-model scalars of the parameters
//   k_gemmt_f64   w_bar = basis^T y_bar on the matrix cores in FP64 (the split-BF16 product is k_mm_bf16x3,
//                 ops_mm.hip, with the basis packed by k_pack_t)
//   k_grad_out    params_bar, weights_bar / weights_z_bar with the L_bol and mass terms and the metallicity blend
#include "ops_common.h"
#include "ops_launch.h"

#define DMAG_DF (-1.0857362047581294)   // -2.5 / ln 10

// Block reductions over the 512 threads (8 wavefronts) of k_grad_chain
template <int K>
__device__ void block_sum(double (&v)[K], double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < K; q++) v[q] = wave_sum(v[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < K; q++) red[wv * K + q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; q++) {
        double s = 0.0;
        for (int w = 0; w < nw; w++) s += red[w * K + q];
        v[q] = s;
    }
    __syncthreads();
}

__device__ double block_max(double v, double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < nw; w++) r = fmax(r, red[w]);
    __syncthreads();
    return r;
}

// The velocity-smoothing window of pixel i (build_velocity_kernel's, as the forward chain walks it)
__device__ __forceinline__ void smooth_window(const GradArgs& a, int i, double fblue, double fred, int& il, int& ih) {
    const double li = a.lam[i], vl = li * fblue, vh = li * fred;
    il = i;
    ih = i;
    while (il > 0 && a.lam[il] > vl) il--;
    while (ih < a.nspec - 1 && a.lam[ih] < vh) ih++;
}

// The raw weight of pixel k in the window [il, ih] of pixel i (before the normalisation by the window's sum), in
// the forward chain's precision and form
template <typename T>
__device__ __forceinline__ double smooth_weight(const GradArgs& a, int i, int k, int il, int ih, double inv2) {
    const double lli = a.loglam[i];
    if constexpr (std::is_same<T, float>::value) {
        const float d = (float)(lli - a.loglam[k]);
        const float4 f = a.sm32[k];
        float wk = __expf(-((float)inv2 * (d * d))) * f.x;
        wk *= (k > il ? f.y : 0.0f) + (k < ih ? f.z : 0.0f);
        return (double)wk;
    } else {
        const double d = lli - a.loglam[k];
        const double wk = exp(-(inv2 * (d * d))) * a.wl[k];
        const double dl = k > il ? a.dlam[k - 1] : 0.0, dr = k < ih ? a.dlam[k] : 0.0;
        return 0.5 * wk * (dl + dr);
    }
}

// The pixel's transmissions (pixel_transmissions, opensps_csp_batch.f90:3765-3818, dust_type 0) in FP64
struct Trans {
    double curve, td, pb, tb, to, dfac, ay;
};

__device__ __forceinline__ Trans transmissions(double L, double dust1, double dust2, double dust3, double ed, double ed1,
                                               double fob, double fnd, bool no_birth, bool no_old, bool nodust) {
    Trans r;
    r.curve = ed != 0.0 ? exp(ed * L) : 1.0;
    r.td = exp(-(dust2 * r.curve));
    r.pb = exp(ed1 * L);
    r.tb = no_birth ? 1.0 : exp(-(dust1 * r.pb));
    r.to = no_old ? 1.0 : exp(-(dust3 * r.curve));
    r.dfac = nodust ? r.td * (1.0 - fnd) + fnd : 1.0;
    r.ay = r.tb * (1.0 - fob) + fob;
    return r;
}

// ------------------------------------------------------------------------------------------------
// k_grad_chain: one workgroup (512 threads) per model. LDS: S0 (the spectrum after attenuation and IGM), S1
// (after smoothing; then u = G1 / the window's sum), G1 (the cotangent of S1), in the forward's precision T.
//   1. the band cotangents F_bar = mags_bar dmag/dF (finite magnitudes of computed bands, F > 0) and the model's
//      scale c = max(|F_bar|, |spec_bar|, |lbol_bar| L_bol) (batch_backward :4855-4880); G1 = spec_bar / c
//   2. S0 and S1 recomputed as the forward chain forms them
//   3. observed frame (backward_observed_frame :6958-7136): per band pixel j the cotangent of fobs(j), zero where
//      the forward's clamp max(v, 0) acts; dz through the interpolation fraction; scattered to G1 by LDS atomics
//   4. sigma_smooth (smooth_sigma_adjoint: dC/dsigma = C (x^2 - <x^2>) / (sigma_beta^3 c 1e-13)); u = G1 / tw
//   5. per pixel k: G0(k) = the transposed smoothing (the windows that hold k), then the IGM and the attenuation
//      (atten_adjoint_chunk :6476-6665): y_bar, o_bar to the product's rows, the parameters' partial sums in FP64
//   6. the block sums of the partials -> gout
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(512) void k_grad_chain(GradArgs a) {
    constexpr bool F32 = std::is_same<T, float>::value;
    extern __shared__ __align__(32) unsigned char lds_raw[];
    T* S0 = (T*)lds_raw;
    T* S1 = S0 + a.nsp;
    T* G1 = S1 + a.nsp;
    double* fbb = (double*)(G1 + a.nsp);
    double* red = fbb + (a.nbands > 0 ? a.nbands : 1);   // 8 wavefronts x 16
    const int m = blockIdx.x, n = a.n, nspec = a.nspec, nsp = a.nsp, nb = a.nbands;
    const int tid = threadIdx.x, nth = blockDim.x;
    const double* h = a.hdr + (size_t)m * OPS_HDR_N;
    const double z = a.prm[(size_t)OPS_P_ZRED * n + m];
    const double dust1 = a.prm[(size_t)OPS_P_DUST1 * n + m], dust2 = a.prm[(size_t)OPS_P_DUST2 * n + m];
    const double dust3 = a.prm[(size_t)OPS_P_DUST3 * n + m];
    const double ed = a.prm[(size_t)OPS_P_DUST_INDEX * n + m], ed1 = a.prm[(size_t)OPS_P_DUST1_INDEX * n + m];
    const double fob = a.prm[(size_t)OPS_P_FRAC_OBRUN * n + m], fnd = a.prm[(size_t)OPS_P_FRAC_NODUST * n + m];
    const double igmf = a.prm[(size_t)OPS_P_IGM_FACTOR * n + m];
    const double sig = a.prm[(size_t)OPS_P_SIGMA_SMOOTH * n + m];
    const bool no_birth = dust1 <= OPS_SAFE_FLOOR || 1.0 - fob <= OPS_SAFE_FLOOR;
    const bool no_old = fabs(dust3) <= OPS_SAFE_FLOOR;
    const bool nodust = 1.0 - fnd > OPS_SAFE_FLOOR;
    const bool igm_on = a.add_igm && z > OPS_SAFE_FLOOR && a.jtop > 0;
    const T* ym = (const T*)a.yo + (size_t)m * nsp;
    const T* om = a.parts == 2 ? (const T*)a.yo + ((size_t)n + m) * nsp : nullptr;

    // ---- 1. band cotangents and the model's scale ----
    double cmax = 0.0, msum = 0.0;
    for (int b = tid; b < nb; b += nth) {
        double f = 0.0;
        if (a.mags_bar && a.mask[b]) {
            const double mg = a.mags[(size_t)m * nb + b], F = a.fband[(size_t)m * nb + b];
            if (mg == mg && F > 0.0) {
                const double mb = a.mags_bar[(size_t)m * nb + b];
                f = mb * DMAG_DF / F;
                msum += mb;
            }
        }
        fbb[b] = f;
        cmax = fmax(cmax, fabs(f));
    }
    if (a.spec_bar)
        for (int i = tid; i < nspec; i += nth) cmax = fmax(cmax, fabs(a.spec_bar[(size_t)m * nspec + i]));
    const double lsum = a.lsum[m];
    const double lbx = a.lbol_bar ? a.lbol_bar[m] : 0.0;
    if (tid == 0) cmax = fmax(cmax, fabs(lbx) * fmax(lsum, 0.0));
    double c = block_max(cmax, red);
    if (!(c > 0.0) || c > 1.7976931348623157e308) c = 1.0;
    const double rc = 1.0 / c;
    for (int b = tid; b < nb; b += nth) fbb[b] *= rc;
    for (int i = tid; i < nsp; i += nth) G1[i] = (a.spec_bar && i < nspec) ? (T)(a.spec_bar[(size_t)m * nspec + i] * rc) : (T)0;

    // ---- 2. S0: attenuation and IGM (the forward's phase 1, per pixel) ----
    double b6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) b6[k] = h[OPS_HDR_B6 + k];
    const double tmax = h[OPS_HDR_TMAX];
    const int jmax = (int)h[OPS_HDR_JMAX];
    const int jcap = jmax > 1 ? jmax : 0;
    const int jt = igm_on ? a.jtop : 0;
    for (int i = tid; i < nsp; i += nth) {
        T tot = (T)0;
        if (i < nspec) {
            const T y = ym[i], o = om ? om[i] : (T)0;
            if constexpr (F32) {
                const float L = (float)a.ltab[i];
                const float cf = (float)fob, cn = (float)fnd;
                const float curve = (ed != 0.0) ? __expf((float)ed * L) : 1.0f;
                const float tr_d = __expf(-((float)dust2 * curve));
                float tr_b = 1.0f, tr_o = 1.0f;
                if (!no_birth) tr_b = __expf(-((float)dust1 * __expf((float)ed1 * L)));
                if (!no_old) tr_o = __expf(-((float)dust3 * curve));
                const float ssum = (y * tr_b * (1.0f - cf) + y * cf) + o * tr_o;
                tot = nodust ? ssum * (tr_d * (1.0f - cn) + cn) : ssum;
            } else {
                const double L = a.ltab[i];
                const double curve = (ed != 0.0) ? exp(ed * L) : 1.0;
                const double tr_d = exp(-(dust2 * curve));
                double tr_b = 1.0, tr_o = 1.0;
                if (!no_birth) tr_b = exp(-(dust1 * exp(ed1 * L)));
                if (!no_old) tr_o = exp(-(dust3 * curve));
                const double ssum = (y * tr_b * (1.0 - fob) + y * fob) + o * tr_o;
                tot = nodust ? ssum * (tr_d * (1.0 - fnd) + fnd) : ssum;
            }
            if (i < jt) {
                const double tau = i < jcap ? tmax : igm_tau6(a.igm_a + (size_t)i * 6, b6);
                if constexpr (F32)
                    tot = tot * exp_kp(-tau * igmf);
                else
                    tot = tot * exp(-tau * igmf);
            }
        }
        S0[i] = tot;
    }
    __syncthreads();
    // S1: the smoothed spectrum (pixels outside [w0, w1) and windows of fewer than 2 pixels pass through)
    const double inv2 = h[OPS_HDR_INV2], fblue = h[OPS_HDR_FBLUE], fred = h[OPS_HDR_FRED];
    const int w0 = (int)h[OPS_HDR_W0], w1 = (int)h[OPS_HDR_W1];
    for (int i = tid; i < nsp; i += nth) {
        T out = S0[i];
        if (i >= w0 && i < w1) {
            int il, ih;
            smooth_window(a, i, fblue, fred, il, ih);
            if (ih - il + 1 >= 2) {
                if constexpr (F32) {
                    float tw = 0.0f, acc = 0.0f;
                    for (int k = il; k <= ih; k++) {
                        const float wk = (float)smooth_weight<T>(a, i, k, il, ih, inv2);
                        tw += wk;
                        acc += wk * S0[k];
                    }
                    if (tw > 0.0f) out = acc / tw;
                } else {
                    double tw = 0.0, acc = 0.0;
                    for (int k = il; k <= ih; k++) {
                        const double wk = smooth_weight<T>(a, i, k, il, ih, inv2);
                        tw += wk;
                        acc += wk * S0[k];
                    }
                    if (tw > OPS_SAFE_FLOOR) out = acc / tw;
                }
            }
        }
        S1[i] = out;
    }
    __syncthreads();

    // ---- 3. the observed frame: fobs(j) = max(v(j), 0) / lambda_j, v the interpolation of S1 at lambda_j / (1 + z) ----
    const double zm = a.zmags ? a.zmags[m] : z;   // the magnitudes' redshift
    const bool rest = fabs(zm) <= OPS_SAFE_FLOOR;
    const double rz = h[OPS_HDR_RZ], lnopz = h[OPS_HDR_LNOPZ];
    double dzo = 0.0;
    const int jb1 = min(a.band_j1, nspec - 1);
    for (int j = a.band_j0 + tid; j <= jb1; j += nth) {
        const int pj = j - a.band_j0;
        double gk = 0.0;
        for (int e = a.bp_ptr[pj]; e < a.bp_ptr[pj + 1]; e++) gk += fbb[a.bp_band[e]] * a.bp_c64[e];
        if (gk == 0.0) continue;
        T v;
        int il = j;
        double t = 0.0;
        if (rest) {
            v = S1[j];
        } else {
            const double tg = a.lam[j] * rz;
            const LamLookup& L = a.lk;
            const int cc = lam_count<false>(L.nspec, L.nbk, L.lnl0, L.bkinv, L.lam, L.bkt, tg, a.loglam[j] - lnopz);
            il = max(1, min(nspec - 1, cc)) - 1;
            t = (tg - a.lam[il]) * a.rdlam[il];
            const T lo = S1[il], hi = S1[il + 1];
            v = lo + (T)t * (hi - lo);
        }
        if (!(v > (T)0)) continue;   // the clamp (and NaN -> 0): no gradient
        const double gv = gk * a.wl[j];
        if (rest) {
            atomicAdd(&G1[j], (T)gv);
        } else {
            if (!a.zmags) dzo += gv * ((double)S1[il + 1] - (double)S1[il]) * (-a.lam[j] * rz * rz) * a.rdlam[il];
            atomicAdd(&G1[il], (T)((1.0 - t) * gv));
            if (t != 0.0) atomicAdd(&G1[il + 1], (T)(t * gv));
        }
    }
    __syncthreads();

    // ---- 4. sigma_smooth's derivative, and u = G1 / tw in S1 (NaN: the pixel passes through) ----
    double dsg = 0.0;
    int rmax = 0;
    for (int i = w0 + tid; i < w1; i += nth) {
        int il, ih;
        smooth_window(a, i, fblue, fred, il, ih);
        T u = (T)NAN;
        if (ih - il + 1 >= 2) {
            const double lli = a.loglam[i];
            double tw = 0.0, sx2 = 0.0, ss = 0.0, sx2s = 0.0;
            for (int k = il; k <= ih; k++) {
                const double wk = smooth_weight<T>(a, i, k, il, ih, inv2);
                const double x = lli - a.loglam[k], s0 = (double)S0[k];
                tw += wk;
                sx2 += wk * x * x;
                ss += wk * s0;
                sx2s += wk * x * x * s0;
            }
            if (F32 ? tw > 0.0 : tw > OPS_SAFE_FLOOR) {
                const double g = (double)G1[i];
                dsg += g * (sx2s / tw - (sx2 / tw) * (ss / tw));
                u = (T)(g / tw);
                rmax = max(rmax, max(ih - i, i - il));
            }
        }
        S1[i] = u;
    }
    rmax = (int)block_max((double)rmax, red);
    // (block_max ends with a barrier: every u is in place)

    // ---- 5. per pixel: the transposed smoothing, then the IGM and the attenuation ----
    double d[7] = {0, 0, 0, 0, 0, 0, 0}, d13 = 0.0, d14 = 0.0;
    const double opz = 1.0 + z;
    double bd6[6];
    const double IE[6] = {3.46, 1.68, 3.0 + 0.46, 1.5 + 0.18, 3.0 - 1.32, 1.68};
#pragma unroll
    for (int k = 0; k < 6; k++) bd6[k] = IE[k] * b6[k] / opz;
    for (int k = tid; k < nsp; k += nth) {
        double yb = 0.0, ob = 0.0;
        if (k < nspec) {
            const bool pass = k < w0 || k >= w1 || S1[k] != S1[k];
            double g0 = pass ? (double)G1[k] : 0.0;
            const int i0 = max(w0, k - rmax), i1 = min(w1 - 1, k + rmax);
            for (int i = i0; i <= i1; i++) {
                const T u = S1[i];
                if (u != u || u == (T)0) continue;
                int il, ih;
                smooth_window(a, i, fblue, fred, il, ih);
                if (k < il || k > ih) continue;
                g0 += smooth_weight<T>(a, i, k, il, ih, inv2) * (double)u;
            }
            const double L = a.ltab[k];
            const Trans r = transmissions(L, dust1, dust2, dust3, ed, ed1, fob, fnd, no_birth, no_old, nodust);
            double tig = 1.0, tauc = 0.0;
            if (k < jt) {
                tauc = k < jcap ? tmax : igm_tau6(a.igm_a + (size_t)k * 6, b6);
                tig = exp(-tauc * igmf);
            }
            const double ga = g0 * tig;
            const double y = (double)ym[k], o = om ? (double)om[k] : 0.0;
            yb = ga * r.ay * r.dfac;
            ob = ga * r.to * r.dfac;
            const double gd = ga * (y * r.ay + o * r.to);
            const double gtd = nodust ? gd * (1.0 - fnd) : 0.0;
            const double gtb = ga * y * (1.0 - fob) * r.dfac;
            const double gto = ga * o * r.dfac;
            if (k < jt) {
                // capped pixels (below the cap pixel) take the cap pixel's d tau / dz
                const int pp = max(k, jmax - 1);
                const double* a6 = a.igm_a + (size_t)pp * 6;
                const double dtz = a6[0] * bd6[0] + a6[1] * bd6[1] + a6[2] * bd6[2] + a6[3] * bd6[3] + a6[4] * bd6[4] +
                                   a6[5] * bd6[5];
                const double apre = (y * r.ay + o * r.to) * r.dfac;
                d13 += -ga * apre * igmf * dtz;
                d14 += -ga * apre * tauc;
            }
            if (no_birth) {
                d[0] += -gtb * r.pb;
            } else {
                d[0] += -gtb * r.pb * r.tb;
                d[4] += -gtb * dust1 * r.pb * L * r.tb;
            }
            d[1] += -gtd * r.td * r.curve;
            d[2] += -gto * r.to * r.curve;
            d[3] += -(gtd * r.td * dust2 + gto * r.to * dust3) * r.curve * L;
            d[5] += ga * y * (1.0 - r.tb) * r.dfac;
            if (nodust) d[6] += gd * (1.0 - r.td);
        }
        // the product's rows (model, part): the young cotangent, then the old one
        const size_t r0 = (size_t)m * a.parts * nsp + k;
        if constexpr (F32) {
            uint16_t hi, lo;
            split_bf16(yb, hi, lo);
            ((uint16_t*)a.ybh)[r0] = hi;
            ((uint16_t*)a.ybl)[r0] = lo;
            if (a.parts == 2) {
                split_bf16(ob, hi, lo);
                ((uint16_t*)a.ybh)[r0 + nsp] = hi;
                ((uint16_t*)a.ybl)[r0 + nsp] = lo;
            }
        } else {
            ((double*)a.ybh)[r0] = yb;
            if (a.parts == 2) ((double*)a.ybh)[r0 + nsp] = ob;
        }
    }

    // ---- 6. the model's scalars ----
    double v[12] = {d[0], d[1], d[2], d[3], d[4], d[5], d[6], d13, d14, dzo, dsg, msum};
    block_sum<12>(v, red);
    if (tid == 0) {
        double* g = a.gout + (size_t)m * OPS_GOUT_N;
        g[OPS_GOUT_C] = c;
        g[OPS_GOUT_LSB] = lsum > OPS_SAFE_FLOOR ? lbx / (lsum * log(10.0)) * rc : 0.0;
        for (int q = 0; q < 7; q++) g[OPS_GOUT_D + q] = v[q];
        g[OPS_GOUT_DZI] = v[7];
        g[OPS_GOUT_DIGM] = v[8];
        g[OPS_GOUT_DZO] = v[9];
        double ds = 0.0;
        if (sig > OPS_SAFE_FLOOR && w1 > w0) {
            const double sb = sig / (OPS_C_LIGHT * 1.0e-13);
            ds = v[10] / (sb * sb * sb * OPS_C_LIGHT * 1.0e-13);
        }
        g[OPS_GOUT_DSG] = ds;
        // the distance term (distance_term_dz, opensps_batch_photometry.f90:798-813): d_L linear in the table
        double dzd = 0.0;
        if (v[11] != 0.0 && fabs(z) > OPS_SAFE_FLOOR && !a.zmags) {   // (with zred_mags: not zred's)
            const bool asc = a.cz[a.ncosmo - 1] >= a.cz[0];
            int cnt = 0;
            for (int q = 0; q < a.ncosmo; q++) cnt += (asc ? a.cz[q] <= z : z < a.cz[q]) ? 1 : 0;
            const int i = max(1, min(cnt, a.ncosmo - 1));
            const double slope = (a.cdl[i] - a.cdl[i - 1]) / (a.cz[i] - a.cz[i - 1]);
            const double dm = a.cdl[i - 1] + slope * (z - a.cz[i - 1]);
            if (dm == dm && dm > OPS_SAFE_FLOOR)
                dzd = v[11] * (5.0 / log(10.0) * slope / dm - 2.5 / (log(10.0) * (1.0 + z)));
        }
        g[OPS_GOUT_DZD] = dzd;
    }
}

template __global__ void k_grad_chain<float>(GradArgs);
template __global__ void k_grad_chain<double>(GradArgs);

// ------------------------------------------------------------------------------------------------
// The transposed product C(col, k) = sum_j At(k, j) Yb(col, j): workgroup tile 64 rows (k) x 64 columns, four
// wavefronts of 32 x 32. Both operands are pixel-contiguous rows, so each lane reads its K slice with one vector
// load (8 bf16; 4 doubles for f64, whose four K steps take one component each). The blockIdx -> tile map keeps
// each XCD on one eighth of the row tiles (At stays in that XCD's L2).
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool gemmt_tile(const GemmTArgs& a, int& rt, int& ct) {
    const int b = blockIdx.x, xcd = b & 7, r = b >> 3;
    const int rtx = (a.nrt + 7) >> 3;
    rt = (r % rtx) * 8 + xcd;
    ct = r / rtx;
    return rt < a.nrt && ct < a.nct;
}

__global__ __launch_bounds__(256) void k_gemmt_f64(GemmTArgs a) {
    int rt, ct;
    if (!gemmt_tile(a, rt, ct)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, wr = wv >> 1, wc = wv & 1;
    const int row0 = rt * 64 + wr * 32, col0 = ct * 64 + wc * 32;
    const int kq = 4 * (lane >> 4);
    const f64x4_t *A[2], *B[2];
    bool valid[2];
    for (int i = 0; i < 2; i++) {
        A[i] = (const f64x4_t*)(a.A64 + (size_t)(row0 + i * 16 + (lane & 15)) * a.nsp + kq);
        const int col = col0 + i * 16 + (lane & 15);
        valid[i] = col < a.ncols;
        B[i] = (const f64x4_t*)(a.B64 + (size_t)(valid[i] ? col : 0) * a.nsp + kq);
    }
    f64x4_t acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (f64x4_t){0.0, 0.0, 0.0, 0.0};
    const f64x4_t zero4 = (f64x4_t){0.0, 0.0, 0.0, 0.0};
    const int kchunks = a.nsp / 16;
    for (int kc = 0; kc < kchunks; kc++) {
        f64x4_t av[2], bv[2];
        for (int i = 0; i < 2; i++) {
            av[i] = A[i][kc * 4];
            bv[i] = valid[i] ? B[i][kc * 4] : zero4;
        }
        // pixel 16 kc + 4 q + s enters step s from lanes with q = lane / 16, in A and B alike
#pragma unroll
        for (int s = 0; s < 4; s++)
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][s], bv[j][s], acc[i][j], 0, 0, 0);
    }
    double* C = (double*)a.C;
    for (int j = 0; j < 2; j++) {
        if (!valid[j]) continue;
        double* col = C + (size_t)(col0 + j * 16 + (lane & 15)) * a.ldc;
        // f64 16x16x4 result layout: register r of lane l holds row (l / 16) + 4 r, column l % 16
        for (int i = 0; i < 2; i++)
            for (int r = 0; r < 4; r++) col[row0 + i * 16 + (lane >> 4) + 4 * r] = acc[i][j][r];
    }
}

// The basis as the transposed product's A: rows k = z nt + t (nrows, zero past kfull), nsp pixels (zero past nspec).
// FP64 row-major; split BF16 in k_mm_bf16x3's instruction order: 16-row block k / 16, K step j / 32, lane
// k % 16 + 16 ((j % 32) / 8), element j % 8.
__global__ void k_pack_t(int nspec, int nsp, int kfull, int nrows, const double* basis, double* A64, uint16_t* Ah,
                         uint16_t* Al) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nrows * nsp) return;
    const int k = (int)(e / nsp), j = (int)(e % nsp);
    const double v = (k < kfull && j < nspec) ? basis[(size_t)k * nspec + j] : 0.0;
    if (A64) A64[e] = v;
    if (Ah) {
        uint16_t hi, lo;
        split_bf16(v, hi, lo);
        const int ksteps = nsp / 32;
        const size_t at = (((size_t)(k >> 4) * ksteps + (j >> 5)) * 64 + (k & 15) + 16 * ((j & 31) >> 3)) * 8 + (j & 7);
        Ah[at] = hi;
        Al[at] = lo;
    }
}

// ------------------------------------------------------------------------------------------------
// k_grad_out: one workgroup per model. The weights' cotangents from the transposed product C (scaled by the
// model's c), with the L_bol and mass terms (backward_age_sums_w :5288-5330 and the weights over metallicities'
// mass term :1143-1151; backward_age_sums with the metallicity blend; batch_mass_cotangent :5142), the
// metallicity weight's cotangent into logzsol's row, and params_bar.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grad_out(GradOutArgs a) {
    __shared__ double red[4];
    const int m = blockIdx.x, tid = threadIdx.x, nt = a.nt;
    const double* g = a.gout + (size_t)m * OPS_GOUT_N;
    const double c = g[OPS_GOUT_C], lsb = g[OPS_GOUT_LSB] * c;
    const double mb = a.mass_bar ? a.mass_bar[m] : 0.0;
    const int it = a.it[m], zlo = a.zlo[m];
    auto Cv = [&](int part, int k) -> double {
        const size_t at = ((size_t)m * a.parts + part) * a.ldc + k;
        return a.mixed ? (double)((const float*)a.C)[at] : ((const double*)a.C)[at];
    };
    double dzb = 0.0;
    if (a.wz) {
        for (int k = tid; k < a.kfull; k += blockDim.x) {
            const int t = k % nt;
            const int part = (a.parts == 2 && t + 1 > it) ? 1 : 0;
            const double v = Cv(part, k) * c + lsb * pow(10.0, a.blbol[k]) + mb * a.bmass[k];
            if (a.weights_z_bar) a.weights_z_bar[(size_t)m * a.kfull + k] = v;
        }
        if (a.weights_bar)
            for (int t = tid; t < nt; t += blockDim.x) a.weights_bar[(size_t)m * nt + t] = 0.0;
    } else {
        const bool blend = a.nz > 1;
        const double dz = blend ? a.hdr[(size_t)m * OPS_HDR_N + OPS_HDR_DZ] : 0.0;
        const int zl = zlo - 1, zh = blend ? zlo : zlo - 1;
        for (int t = tid; t < nt; t += blockDim.x) {
            const int part = (a.parts == 2 && t + 1 > it) ? 1 : 0;
            const double gl = Cv(part, zl * nt + t), gh = Cv(part, zh * nt + t);
            double mk, lk, wb;
            if (blend) {
                mk = (1.0 - dz) * a.bmass[(size_t)zl * nt + t] + dz * a.bmass[(size_t)zh * nt + t];
                lk = pow(10.0, (1.0 - dz) * a.blbol[(size_t)zl * nt + t] + dz * a.blbol[(size_t)zh * nt + t]);
                wb = ((1.0 - dz) * gl + dz * gh) * c;
            } else {
                mk = a.bmass[t];
                lk = pow(10.0, a.blbol[t]);
                wb = gl * c;
            }
            if (a.weights_bar) a.weights_bar[(size_t)m * nt + t] = wb + lsb * lk + mb * mk;
            if (blend) {
                const double w = a.w[(size_t)m * nt + t];
                const double wk = w > OPS_SAFE_FLOOR ? w : 0.0;
                dzb += wk * (gh - gl) * c;
                if (wk > 0.0)
                    dzb += lsb * wk * lk * log(10.0) * (a.blbol[(size_t)zh * nt + t] - a.blbol[(size_t)zl * nt + t]) +
                           mb * wk * (a.bmass[(size_t)zh * nt + t] - a.bmass[(size_t)zl * nt + t]);
            }
        }
        dzb = wave_sum(dzb);
        if ((tid & 63) == 0) red[tid >> 6] = dzb;
        __syncthreads();
        dzb = red[0] + red[1] + red[2] + red[3];
    }
    if (tid == 0 && a.params_bar) {
        double* pb = a.params_bar + (size_t)m * OPENSPS_HIP_NBGRAD_ROWS;
        for (int r = 0; r < OPENSPS_HIP_NBGRAD_ROWS; r++) pb[r] = 0.0;
        for (int q = 0; q < 7; q++) pb[1 + q] = g[OPS_GOUT_D + q] * c;   // dust1 .. frac_nodust (BGRAD 2..8)
        if (!a.wz && a.nz > 1) pb[15] = dzb / (a.logz[zlo] - a.logz[zlo - 1]);   // logzsol (BGRAD 16)
        pb[22] = (g[OPS_GOUT_DZO] + g[OPS_GOUT_DZI]) * c + g[OPS_GOUT_DZD];  // zred (BGRAD 23)
        pb[23] = g[OPS_GOUT_DSG] * c;                                        // sigma_smooth (BGRAD 24)
        pb[32] = g[OPS_GOUT_DIGM] * c;                                       // igm_factor (BGRAD 33)
    }
}
