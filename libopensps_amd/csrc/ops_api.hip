// The C ABI of libopensps_hip.so (include/opensps_hip.h): context set-up on the host (the tables the
// kernels read, formed in FP64 once per grid) and the launch sequence of a batch.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/opensps_hip.h"
#include "ops_common.h"
#include "ops_launch.h"

static std::mutex g_err_mu;

__global__ void k_fill_i32(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ p, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- the fixed scratch of the no-allocation entry point (opensps_hip_csp_batch_noalloc) ----
// One array per device, held while the process lives; its device address is taken once per device at context
// creation (hipGetSymbolAddress with that device current) and used wherever host code needs the memory.
#define OPS_NOALLOC_BYTES ((size_t)64 << 20)
static __device__ __align__(256) unsigned char scratch[OPS_NOALLOC_BYTES];
#define OPS_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[OPS_MAX_DEVICES] = {};
static std::string g_err;

int ops_fail(OpsCtx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) {
        c->last_error = buf;
    } else {
        std::lock_guard<std::mutex> g(g_err_mu);
        g_err = buf;
    }
    return code;
}

#define HIPCHK(c, call)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return ops_fail((c), 10, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int ops_reserve(OpsCtx* c, OpsBuffer& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (c->arena) {   // the no-allocation entry point: a 256-byte aligned piece of the fixed scratch
        OpsArena& A = *c->arena;
        const size_t off = (A.used + 255) & ~(size_t)255;
        if (off > A.cap || bytes > A.cap - off)
            return ops_fail(c, 3, "batch_noalloc: the batch needs more than the %zu bytes of fixed scratch (%zu used, "
                                  "%zu more asked): fewer models per call", A.cap, off, bytes);
        b.p = A.base + off;
        b.bytes = bytes;
        A.used = off + bytes;
        return 0;
    }
    if (b.bytes >= bytes) return 0;
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    HIPCHK(c, hipMalloc(&b.p, want));
    b.bytes = want;
    return 0;
}

template <typename T>
static int upload(OpsCtx* c, T** dst, const T* src, size_t count) {
    if (*dst) {
        HIPCHK(c, hipFree(*dst));
        *dst = nullptr;
    }
    HIPCHK(c, hipMalloc((void**)dst, sizeof(T) * (count ? count : 1)));
    if (count) HIPCHK(c, hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return 0;
}

#define UPLOAD(c, dst, src, count)                 \
    do {                                           \
        int st_ = upload((c), &(dst), (src), (count)); \
        if (st_) return st_;                       \
    } while (0)

// ABS_MAG_ZEROPOINT_LOG = log10(L_sun / (4 pi (10 pc)^2)) (src/core/opensps_constants.f90:87-88)
static double abs_mag_zeropoint_log() {
    const double pi = acos(-1.0);
    const double pc = (648000.0 / pi) * 1.495978707e13;
    return log10(3.828e33 / (4.0 * pi * (10.0 * pc) * (10.0 * pc)));
}

// Madau (1995) wavelength factors of the IGM optical depth (igm_tau_pixel_terms,
// src/physics/opensps_cosmology.f90:458-492)
static void igm_pixel_terms(double w, double* a) {
    static const double LYW[17] = {1215.67, 1025.72, 972.537, 949.743, 937.803, 930.748, 926.226, 923.150, 920.963,
                                   919.352, 918.129, 917.181, 916.429, 915.824, 915.329, 914.919, 914.576};
    static const double LYC[17] = {0.0036,    0.0017,    0.0011846, 0.0009410, 0.0007960, 0.0006967,
                                   0.0006236, 0.0005665, 0.0005200, 0.0004817, 0.0004487, 0.0004200,
                                   0.0003947, 0.000372,  0.000352,  0.0003334, 0.00031644};
    const double LL = 911.75;
    for (int k = 0; k < 6; k++) a[k] = 0.0;
    if (w < LYW[0]) {
        double ser = 0.0;
        for (int i = 0; i < 17; i++)
            if (w < LYW[i]) ser = ser + LYC[i] * pow(LYW[i], -3.46);
        a[0] = ser * pow(w, 3.46);
        a[1] = 0.0017 * pow(w / LYW[0], 1.68);
    }
    if (w < LL) {
        double s = w / LL;
        a[2] = 0.25 * pow(s, 3.0) * (1.0 - pow(s, 0.46));
        a[3] = 9.4 * pow(s, 1.5) * (1.0 - pow(s, 0.18));
        a[4] = -0.7 * pow(s, 3.0) * (pow(s, -1.32) - 1.0);
        a[5] = -0.023 * (1.0 - pow(s, 1.68));
    }
}

extern "C" {

int opensps_hip_create(int device, opensps_hip_ctx_t** out) {
    if (!out) return ops_fail(nullptr, 3, "opensps_hip_create: NULL output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return ops_fail(nullptr, 10, "opensps_hip_create: no HIP device");
    if (device < 0 || device >= ndev) return ops_fail(nullptr, 3, "opensps_hip_create: device %d of %d", device, ndev);
    if (hipSetDevice(device) != hipSuccess) return ops_fail(nullptr, 10, "opensps_hip_create: hipSetDevice failed");
    if (device >= OPS_MAX_DEVICES) return ops_fail(nullptr, 3, "opensps_hip_create: device %d beyond %d", device,
                                                     OPS_MAX_DEVICES);
    {   // the fixed scratch's address on this device, once
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return ops_fail(nullptr, 10, "opensps_hip_create: the fixed scratch's address: %s", hipGetErrorString(e));
            g_scratch_ptr[device] = p;
        }
    }
    OpsCtx* c = new OpsCtx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ops_fail(nullptr, 10, "opensps_hip_create: stream");
    }
    c->own_stream = true;
    // the batch's status flag comes back through pinned host memory, copied in stream order (NULL: a plain copy at sync)
    if (hipHostMalloc((void**)&c->h_flag, sizeof(int), hipHostMallocDefault) != hipSuccess) c->h_flag = nullptr;
    // OPENSPS_F64_GEMM=vector: the FP64 age sums with vector FMAs instead of the f64 matrix instruction
    const char* e = getenv("OPENSPS_F64_GEMM");
    c->f64_vector = e && strcmp(e, "vector") == 0;
    // OPENSPS_CHAIN=two: the two-buffer chain for every batch (it otherwise runs only where the band pixels
    // outgrow the single-buffer chain's registers)
    const char* ch = getenv("OPENSPS_CHAIN");
    c->chain_two = ch && strcmp(ch, "two") == 0;
    *out = c;
    return 0;
}

static void free_operator(OpsCtx* c);
static void free_full(OpsCtx* c);

static void free_ptr(void* p) {
    if (p) (void)hipFree(p);
}

int opensps_hip_destroy(opensps_hip_ctx_t* c) {
    if (!c) return 1;
    (void)hipStreamSynchronize(c->stream);
    void* ps[] = {c->d_lam, c->d_loglam, c->d_ltab, c->d_wl, c->d_dlam, c->d_time, c->d_logz, c->d_bmass, c->d_blbol,
                  c->d_igm_a, c->d_igm64s, c->d_igm32s, c->d_basis64, c->d_A64, c->d_Ahi, c->d_Alo, c->d_bmeta, c->d_bc64,
                  c->d_bc32, c->d_mask, c->d_cz, c->d_cdl, c->yo.p, c->wcol.p, c->w64.p, c->order.p, c->tiles.p,
                  c->misc.p, c->spec.p, c->flag.p, c->d_rdlam, c->d_ltab32, c->d_wl32, c->d_sm32, c->d_bkt, c->hdr.p, c->wsfh.p,
                  c->opu.p, c->opwc.p, c->opsum.p, c->opseg.p, c->opimf.p, c->opspart.p, c->opgp.p, c->d_bp_ptr, c->d_bp_band,
                  c->d_bp_c64, c->glsum.p, c->gfband.p, c->gw.p, c->gout.p, c->gyb.p, c->gC.p, c->gwb.p, c->d_magvega, c->d_tab, c->d_dcurve, c->d_dcurve32,
                  c->opsraw.p, c->opct.p, c->oppart.p, c->opimfd.p,
                  c->d_qpah, c->d_umin, c->d_dustem, c->d_agn_tau, c->d_agn, c->d_wnu, c->nebw.p,
                  c->d_neb_lz, c->d_neb_lu, c->d_neb_cont, c->d_neb_line, c->d_neb_lpos, c->d_neb_ia, c->d_neb_newr,
                  c->d_neb_llo, c->d_neb_wa, c->d_neb_dwa, c->d_neb_q0, c->d_neb_lptr, c->d_neb_lid, c->d_neb_gval};
    for (void* p : ps) free_ptr(p);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    free_operator(c);
    free_full(c);
    for (int k = 0; k <= OpsCtx::NPROF; k++)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    for (int k = 0; k <= OpsCtx::NGPROF; k++)
        if (c->gev[k]) (void)hipEventDestroy(c->gev[k]);
    for (int k = 0; k <= OpsCtx::NOPROF; k++)
        if (c->opev[k]) (void)hipEventDestroy(c->opev[k]);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int opensps_hip_set_stream(opensps_hip_ctx_t* c, void* stream) {
    if (!c) return 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (stream) {
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else if (!c->own_stream) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return 0;
}

int opensps_hip_set_grid(opensps_hip_ctx_t* c, int nspec, const double* lam, int nt, const double* log_age, int nz,
                         const double* logz, const double* spec, const double* mass, const double* lbol) {
    if (!c) return 1;
    if (nspec < 2 || nt < 2 || nz < 1 || !lam || !log_age || !logz || !spec || !mass || !lbol)
        return ops_fail(c, 3, "set_grid: bad sizes or NULL arrays");
    for (int i = 1; i < nspec; i++)
        if (!(lam[i] > lam[i - 1])) return ops_fail(c, 3, "set_grid: wavelengths not strictly ascending at %d", i);
    for (int i = 1; i < nt; i++)
        if (!(log_age[i] > log_age[i - 1])) return ops_fail(c, 3, "set_grid: ages not strictly ascending at %d", i);
    for (int i = 1; i < nz; i++)
        if (!(logz[i] > logz[i - 1])) return ops_fail(c, 3, "set_grid: metallicities not strictly ascending at %d", i);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nspec = nspec;
    c->nt = nt;
    c->nz = nz;
    c->nsp = ((nspec + OPS_TILE_PX - 1) / OPS_TILE_PX) * OPS_TILE_PX;
    c->dcurve_on = 0;   // (a curve is per pixel of the grid)
    c->nqpah = 0;       // (so are the dust emission and AGN templates: set them again)
    c->numin = 0;
    c->ntau = 0;
    c->nnz = 0;         // and the nebular tables
    c->nbands = 0;   // the band tables index the old grid: set_bands again
    c->band_j0 = 0;
    c->band_j1 = -1;
    c->npair = nz > 1 ? nz - 1 : 1;
    c->kpair = nz > 1 ? 2 * nt : nt;
    c->k64 = ((c->kpair + OPS_KSTEP_F64 - 1) / OPS_KSTEP_F64) * OPS_KSTEP_F64;
    c->kbf = ((c->kpair + OPS_KSTEP_BF16 - 1) / OPS_KSTEP_BF16) * OPS_KSTEP_BF16;
    c->h_lam.assign(lam, lam + nspec);
    c->h_time.assign(log_age, log_age + nt);
    c->h_logz.assign(logz, logz + nz);
    // Grid-only tables in FP64 (the reference forms them once per context: prepare_batch_grid,
    // attenuation_tables_cached; src/spectra/opensps_batch_photometry.f90:237-394, opensps_csp_batch.f90:3494-3566)
    // ltab and ltab32 are padded to nsp (the chain reads them by whole quads)
    std::vector<double> loglam(nspec), ltab(c->nsp, 0.0), wl(nspec), dlam(nspec);
    for (int j = 0; j < nspec; j++) {
        loglam[j] = log(lam[j]);
        ltab[j] = log(lam[j] / 5500.0);
        wl[j] = lam[j] > 2.2250738585072014e-308 ? 1.0 / lam[j] : 0.0;
        dlam[j] = j + 1 < nspec ? lam[j + 1] - lam[j] : 0.0;
    }
    // FP32 copies and the per-pixel smoothing factors of the FP32 chain
    std::vector<double> rdlam(nspec);
    std::vector<float> ltab32(c->nsp, 0.0f), wl32(nspec);
    std::vector<float4> sm32(nspec);
    for (int j = 0; j < nspec; j++) {
        rdlam[j] = j + 1 < nspec ? 1.0 / dlam[j] : 0.0;
        ltab32[j] = (float)ltab[j];
        wl32[j] = (float)wl[j];
        sm32[j] = make_float4((float)wl[j], j > 0 ? (float)(0.5 * dlam[j - 1]) : 0.0f, (float)(0.5 * dlam[j]), 0.0f);
    }
    // Buckets over ln(lambda) no wider than the finest pixel: a lookup lands within a pixel or two of
    // the interval, and an exact walk on lambda finishes it
    double minsp = 1e300;
    for (int j = 0; j + 1 < nspec; j++) minsp = fmin(minsp, loglam[j + 1] - loglam[j]);
    const double range = loglam[nspec - 1] - loglam[0];
    double nb = range / minsp + 2.0;
    int nbk = nb > (double)(1 << 20) ? (1 << 20) : (int)nb + 1;
    double bw = nb > (double)(1 << 20) ? range / (double)(nbk - 2) : minsp;
    std::vector<int> bkt(nbk);
    for (int b = 0, j = 0; b < nbk; b++) {
        const double x = loglam[0] + (double)b * bw;
        while (j < nspec && loglam[j] < x) j++;
        bkt[b] = j;
    }
    c->nbk = nbk;
    c->lnl0 = loglam[0];
    c->bkinv = 1.0 / bw;
    // the frequency trapezoid of the dust emission's sums (dust_emission_reductions with prepare_batch_grid's wt,
    // src/spectra/opensps_batch_photometry.f90:262): sum_i wt(i) (f(i) + f(i+1)) = sum_i f(i) (wt(i-1) + wt(i))
    std::vector<double> wnu(c->nsp, 0.0);
    for (int j = 0; j + 1 < nspec; j++) {
        const double wt = 0.5 * fabs(OPS_C_LIGHT / lam[j + 1] - OPS_C_LIGHT / lam[j]);
        wnu[j] += wt;
        wnu[j + 1] += wt;
    }
    UPLOAD(c, c->d_wnu, wnu.data(), wnu.size());
    UPLOAD(c, c->d_rdlam, rdlam.data(), (size_t)nspec);
    UPLOAD(c, c->d_ltab32, ltab32.data(), ltab32.size());
    UPLOAD(c, c->d_wl32, wl32.data(), (size_t)nspec);
    UPLOAD(c, c->d_sm32, sm32.data(), (size_t)nspec);
    UPLOAD(c, c->d_bkt, bkt.data(), (size_t)nbk);
    int jtop = 0;
    for (int j = 1; j <= nspec; j++)
        if (lam[j - 1] < 1215.67) jtop = j;
    std::vector<double> igm((size_t)6 * (jtop ? jtop : 1));
    for (int j = 0; j < jtop; j++) igm_pixel_terms(lam[j], &igm[(size_t)6 * j]);
    c->jtop = jtop;
    UPLOAD(c, c->d_lam, lam, (size_t)nspec);
    UPLOAD(c, c->d_loglam, loglam.data(), (size_t)nspec);
    UPLOAD(c, c->d_ltab, ltab.data(), ltab.size());
    UPLOAD(c, c->d_wl, wl.data(), (size_t)nspec);
    UPLOAD(c, c->d_dlam, dlam.data(), (size_t)nspec);
    UPLOAD(c, c->d_time, log_age, (size_t)nt);
    UPLOAD(c, c->d_logz, logz, (size_t)nz);
    UPLOAD(c, c->d_bmass, mass, (size_t)nt * nz);
    UPLOAD(c, c->d_blbol, lbol, (size_t)nt * nz);
    UPLOAD(c, c->d_igm_a, igm.data(), igm.size());
    {
        // term-major with the stride padded to whole quads (the chain reads a quad's terms as one 16- or 32-byte load)
        const size_t jn = jtop ? (size_t)((jtop + 3) & ~3) : 4;
        c->jq = (int)jn;
        std::vector<double> i64(6 * jn, 0.0);
        std::vector<float> i32(6 * jn, 0.0f);
        for (int j = 0; j < jtop; j++)
            for (int k = 0; k < 6; k++) {
                i64[k * jn + j] = igm[(size_t)6 * j + k];
                i32[k * jn + j] = (float)igm[(size_t)6 * j + k];
            }
        UPLOAD(c, c->d_igm64s, i64.data(), i64.size());
        UPLOAD(c, c->d_igm32s, i32.data(), i32.size());
    }
    UPLOAD(c, c->d_basis64, spec, (size_t)nspec * nt * nz);
    free_full(c);   // (the operands over every (age, Z) are formed again at their first use)
    // Both packed operands now (the first batch of either precision then starts at once)
    const int npxblk = c->nsp / 16;
    PackArgs pa;
    pa.nspec = nspec;
    pa.nt = nt;
    pa.nz = nz;
    pa.npair = c->npair;
    pa.kpair = c->kpair;
    pa.npxblk = npxblk;
    pa.basis = c->d_basis64;
    {
        size_t n64 = (size_t)c->npair * npxblk * (c->k64 / OPS_KSTEP_F64) * 64;
        if (c->d_A64) HIPCHK(c, hipFree(c->d_A64));
        HIPCHK(c, hipMalloc(&c->d_A64, n64 * sizeof(double)));
        pa.kpad = c->k64;
        pa.A64 = c->d_A64;
        k_pack_f64<<<(unsigned)((n64 + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    {
        size_t nbf = (size_t)c->npair * npxblk * (c->kbf / OPS_KSTEP_BF16) * 64;
        if (c->d_Ahi) HIPCHK(c, hipFree(c->d_Ahi));
        if (c->d_Alo) HIPCHK(c, hipFree(c->d_Alo));
        HIPCHK(c, hipMalloc(&c->d_Ahi, nbf * 8 * sizeof(uint16_t)));
        HIPCHK(c, hipMalloc(&c->d_Alo, nbf * 8 * sizeof(uint16_t)));
        pa.kpad = c->kbf;
        pa.Ahi = c->d_Ahi;
        pa.Alo = c->d_Alo;
        k_pack_bf16<<<(unsigned)((nbf + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int opensps_hip_set_bands(opensps_hip_ctx_t* c, int nbands, const double* bands) {
    if (!c) return 1;
    if (c->nspec < 2) return ops_fail(c, 1, "set_bands: no grid");
    if (nbands < 0 || (nbands > 0 && !bands)) return ops_fail(c, 3, "set_bands: bad arguments");
    const int nspec = c->nspec;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> wtl(nspec, 0.0);
    for (int j = 0; j + 1 < nspec; j++) wtl[j] = 0.5 * fabs(c->h_lam[j + 1] - c->h_lam[j]);
    // Per band a quad-aligned run of pixels [q0, q0 + nq) holding the support (zeros around it), and its
    // coefficients at a quad-aligned offset: the chain reads both 4 pixels at a time
    std::vector<int4> meta(nbands ? nbands : 1, make_int4(0, 0, 0, 0));
    std::vector<double> c64;
    int j0 = nspec, j1 = -1;
    for (int b = 0; b < nbands; b++) {
        // supports: prepare_batch_grid (src/spectra/opensps_batch_photometry.f90:293-307), 1-based
        const double* B = bands + (size_t)b * nspec;
        int ilo = nspec, ihi = 1;
        for (int j = 1; j <= nspec; j++)
            if (B[j - 1] != 0.0) {
                ilo = j < ilo ? j : ilo;
                ihi = j > ihi ? j : ihi;
            }
        int jlo = ilo - 1 > 1 ? ilo - 1 : 1, jhi = ihi < nspec - 1 ? ihi : nspec - 1;
        if (ilo > ihi) { jlo = 1; jhi = 0; }
        const int nj = jhi >= jlo ? jhi - jlo + 2 : 0;   // pixels jlo .. jhi + 1 (the trapezoid reads j + 1)
        if (nj == 0) continue;
        const int p0 = jlo - 1;                           // 0-based first pixel
        const int q0 = p0 & ~3, nq = ((p0 + nj - q0) + 3) & ~3;
        meta[b] = make_int4((int)c64.size(), q0, nq, 0);
        // per pixel the trapezoid weights of the two intervals it closes, times the band (the reference's
        // sparse band matrix bk, :323-344), formed in FP64
        for (int p = q0; p < q0 + nq; p++) {
            const int j = p + 1;   // 1-based
            double cw = 0.0;
            if (j >= jlo && j < jlo + nj) {
                if (j >= jlo && j <= jhi) cw += wtl[j - 1];
                if (j - 1 >= jlo && j - 1 <= jhi) cw += wtl[j - 2];
                cw *= B[j - 1];
            }
            c64.push_back(cw);
        }
        j0 = q0 < j0 ? q0 : j0;
        j1 = q0 + nq - 1 > j1 ? q0 + nq - 1 : j1;
    }
    // the same coefficients by pixel over [j0, j1] for the reverse pass (each pixel's bands, in band order)
    {
        const int npix = j1 >= j0 ? j1 - j0 + 1 : 0;
        std::vector<int> bptr(npix + 1, 0), bband;
        std::vector<double> bc;
        std::vector<std::vector<std::pair<int, double>>> per(npix);
        for (int b = 0; b < nbands; b++) {
            const int4 mt = meta[b];
            for (int q = 0; q < mt.z; q++) {
                const double cw = c64[(size_t)mt.x + q];
                if (cw != 0.0) per[mt.y + q - j0].push_back(std::make_pair(b, cw));
            }
        }
        for (int p = 0; p < npix; p++) {
            for (auto& e : per[p]) {
                bband.push_back(e.first);
                bc.push_back(e.second);
            }
            bptr[p + 1] = (int)bband.size();
        }
        if (bband.empty()) {
            bband.push_back(0);
            bc.push_back(0.0);
        }
        UPLOAD(c, c->d_bp_ptr, bptr.data(), bptr.size());
        UPLOAD(c, c->d_bp_band, bband.data(), bband.size());
        UPLOAD(c, c->d_bp_c64, bc.data(), bc.size());
    }
    if (c64.empty()) c64.assign(4, 0.0);
    std::vector<float> c32(c64.size());
    for (size_t i = 0; i < c64.size(); i++) c32[i] = (float)c64[i];
    c->nbands = nbands;
    c->band_j0 = j1 >= 0 ? j0 : 0;
    c->band_j1 = j1;   // < nsp: the quads end at or before the padded grid's end
    UPLOAD(c, c->d_bmeta, meta.data(), meta.size());
    UPLOAD(c, c->d_bc64, c64.data(), c64.size());
    UPLOAD(c, c->d_bc32, c32.data(), c32.size());
    c->h_mask.assign(nbands ? nbands : 1, 1);
    UPLOAD(c, c->d_mask, c->h_mask.data(), c->h_mask.size());
    c->vega_on = 0;   // (the Vega system's zero points are per band: set again after new bands)
    c->vega_ref = 0;
    return 0;
}

// The device mask: the caller's, with the Vega system's reference band always computed (compute_magnitudes_columns,
// src/spectra/opensps_batch_photometry.f90:105-110)
static int upload_mask(OpsCtx* c) {
    if (!c->nbands) return 0;
    std::vector<int> m(c->h_mask.begin(), c->h_mask.begin() + c->nbands);
    if (c->vega_on && c->vega_ref > 0) m[c->vega_ref - 1] = 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_mask, m.data(), sizeof(int) * c->nbands, hipMemcpyHostToDevice));
    return 0;
}

int opensps_hip_set_mag_mask(opensps_hip_ctx_t* c, int nbands, const int* mask) {
    if (!c) return 1;
    if (nbands != c->nbands) return ops_fail(c, 3, "set_mag_mask: %d bands, context has %d", nbands, c->nbands);
    for (int b = 0; b < nbands; b++) c->h_mask[b] = mask ? (mask[b] != 0) : 1;
    return upload_mask(c);
}

int opensps_hip_set_dust_curve(opensps_hip_ctx_t* c, int nspec, const double* curve) {
    if (!c) return 1;
    if (!curve) {
        c->dcurve_on = 0;
        return 0;
    }
    if (nspec != c->nspec || c->nsp < nspec)
        return ops_fail(c, 3, "set_dust_curve: %d pixels, the grid has %d (set the grid first)", nspec, c->nspec);
    std::vector<double> d(c->nsp, 0.0);
    std::vector<float> f(c->nsp, 0.0f);
    for (int i = 0; i < nspec; i++) {
        d[i] = curve[i];
        f[i] = (float)curve[i];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    UPLOAD(c, c->d_dcurve, d.data(), d.size());
    UPLOAD(c, c->d_dcurve32, f.data(), f.size());
    c->dcurve_on = 1;
    return 0;
}

int opensps_hip_set_sfh_tab(opensps_hip_ctx_t* c, int ntab, const double* age_yr, const double* sfr) {
    if (!c) return 1;
    if (ntab < 0 || ntab > OPS_NTABMAX || (ntab > 0 && (!age_yr || !sfr)))
        return ops_fail(c, 3, "set_sfh_tab: ntab %d out of range (0 to %d) or NULL arrays", ntab, OPS_NTABMAX);
    std::vector<double> t(2 * (size_t)ntab);
    for (int i = 0; i < ntab; i++) {
        t[i] = age_yr[i];
        t[ntab + i] = sfr[i];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));   // (a batch in flight may still read the old table)
    UPLOAD(c, c->d_tab, t.data(), t.size());
    c->ntab = ntab;
    return 0;
}

int opensps_hip_set_vega(opensps_hip_ctx_t* c, int nbands, const double* magvega, int vref) {
    if (!c) return 1;
    if (!magvega || nbands == 0) {
        const bool was = c->vega_on && c->vega_ref > 0;
        c->vega_on = 0;
        c->vega_ref = 0;
        return was ? upload_mask(c) : 0;
    }
    if (nbands != c->nbands) return ops_fail(c, 3, "set_vega: %d bands, context has %d", nbands, c->nbands);
    if (vref < 0 || vref > nbands) return ops_fail(c, 3, "set_vega: reference band %d of %d", vref, nbands);
    int st;
    HIPCHK(c, hipStreamSynchronize(c->stream));   // (a batch in flight may still read the old zero points)
    UPLOAD(c, c->d_magvega, magvega, (size_t)nbands);
    c->vega_on = 1;
    c->vega_ref = vref;
    c->vega_anchor = vref > 0 ? magvega[vref - 1] : 0.0;
    if ((st = upload_mask(c))) return st;
    return 0;
}

static void free_operator(OpsCtx* c) {
    void* ps[] = {c->op.d_age_ptr, c->op.d_zoff, c->op.d_first, c->op.d_tz_ptr, c->op.d_pt, c->op.d_age_mmax,
                  c->op.d_te, c->op.d_te_val64, c->op.d_te_val32, c->op.d_A64, c->op.d_Ahi, c->op.d_Alo,
                  c->op.d_seg, c->op.d_segoff, c->op.d_tseg, c->op.d_slow_ptr, c->op.d_slow_pt, c->op.d_wp,
                  c->op.d_poly, c->op.d_pent, c->op.d_pval, c->op.d_At64, c->op.d_Ath, c->op.d_Atl};
    for (void* p : ps) free_ptr(p);
    c->op = OpsCtx::OpsOperator();
}

// k_mm_bf16x3 over every row tile and at most maxct column tiles (grouped: the kernel reads the count on the device)
static int launch_mm(OpsCtx* c, const MMArgs& mm, int maxct) {
    // the operand buffers (64 KiB), overlaid in the epilogue by the C tile (128 columns of 132 floats), then the
    // column tables
    const size_t lds = (size_t)OPS_MM_COLS * OPS_MM_TS * sizeof(float) + 2 * OPS_MM_COLS * sizeof(int);
    const unsigned blocks = 8u * (unsigned)((mm.nrt + 7) / 8) * (unsigned)maxct;
    if (blocks == 0) return 0;
    HIPCHK(c, hipFuncSetAttribute((const void*)k_mm_bf16x3, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_mm_bf16x3<<<blocks, 256, lds, c->stream>>>(mm);
    HIPCHK(c, hipGetLastError());
    return 0;
}

static void free_full(OpsCtx* c) {
    void* ps[] = {c->full.A64, c->full.Ahi, c->full.Alo, c->full.At64, c->full.Ath, c->full.Atl};
    for (void* p : ps) free_ptr(p);
    c->full = OpsCtx::Full();
}

// The basis over every (age, Z) as the operands of the weights over metallicities: the forward product's packed A
// (K = nt nz, metallicities outer, as ssp_weights_z) and, with transposed, the reverse product's rows; formed at
// their first use after set_grid, for the precision asked
static int ensure_full(OpsCtx* c, bool mixed, bool transposed) {
    OpsCtx::Full& f = c->full;
    const int kfull = c->nt * c->nz, npxblk = c->nsp / 16;
    if (f.kfull == 0) {
        f.kfull = kfull;
        f.k64 = ((kfull + OPS_KSTEP_F64 - 1) / OPS_KSTEP_F64) * OPS_KSTEP_F64;
        f.kbf = ((kfull + OPS_KSTEP_BF16 - 1) / OPS_KSTEP_BF16) * OPS_KSTEP_BF16;
        f.nrt = (kfull + 63) / 64;
    }
    PackArgs pa;
    pa.nspec = c->nspec;
    pa.nt = kfull;
    pa.nz = 1;
    pa.npair = 1;
    pa.kpair = kfull;
    pa.npxblk = npxblk;
    pa.basis = c->d_basis64;
    if (!mixed && !f.A64) {
        const size_t n64 = (size_t)npxblk * (f.k64 / OPS_KSTEP_F64) * 64;
        HIPCHK(c, hipMalloc((void**)&f.A64, n64 * sizeof(double)));
        pa.kpad = f.k64;
        pa.A64 = f.A64;
        k_pack_f64<<<(unsigned)((n64 + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    if (mixed && !f.Ahi) {
        const size_t nbf = (size_t)npxblk * (f.kbf / OPS_KSTEP_BF16) * 64;
        HIPCHK(c, hipMalloc((void**)&f.Ahi, nbf * 8 * sizeof(uint16_t)));
        HIPCHK(c, hipMalloc((void**)&f.Alo, nbf * 8 * sizeof(uint16_t)));
        pa.kpad = f.kbf;
        pa.Ahi = f.Ahi;
        pa.Alo = f.Alo;
        k_pack_bf16<<<(unsigned)((nbf + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    if (transposed) {
        const size_t nel = (size_t)f.nrt * 64 * c->nsp;
        double* a64 = nullptr;
        uint16_t *ah = nullptr, *al = nullptr;
        if (!mixed && !f.At64) {
            HIPCHK(c, hipMalloc((void**)&f.At64, nel * sizeof(double)));
            a64 = f.At64;
        }
        if (mixed && !f.Ath) {
            HIPCHK(c, hipMalloc((void**)&f.Ath, nel * sizeof(uint16_t)));
            HIPCHK(c, hipMalloc((void**)&f.Atl, nel * sizeof(uint16_t)));
            ah = f.Ath;
            al = f.Atl;
        }
        if (a64 || ah) {
            k_pack_t<<<(unsigned)((nel + 255) / 256), 256, 0, c->stream>>>(c->nspec, c->nsp, kfull, f.nrt * 64,
                                                                          c->d_basis64, a64, ah, al);
            HIPCHK(c, hipGetLastError());
        }
    }
    return 0;
}

// imf_bin_logs (src/physics/opensps_imf.f90:564-579): ln m1 and ln(m2 / m1) in Kahan's log1p form
static void bin_logs(double m1, double m2, double* llo, double* dl) {
    *llo = 0.0;
    *dl = 0.0;
    if (m1 <= 0.0 || m2 <= m1) return;
    *llo = log(m1);
    const double x = (m2 - m1) / m1, u = 1.0 + x;
    *dl = (u == 1.0) ? x : log(u) * x / (u - 1.0);
}

// The tables that depend on the context's IMF type. The bins k_op_slow takes: per (metallicity, age) the first
// bin, the bins across a power-law break and the empty bins (bin_class 0 or m2 <= m1), in point order. The records
// of k_op_weights, with each common bin's segment and whether lo^e comes down the chain: the previous common bin
// of the same piece of OPS_PIECE (imf_bin_weights_tab's pieces) is in the same segment and ends where this one
// begins, and no non-empty bin k_op_slow takes lies between them (empty bins leave the chain as they are).
static int build_imf_tables(OpsCtx* c) {
    const int nzt = c->op.nz * c->op.nt, ty = c->imf_type % 10, np = c->op.npts;
    std::vector<int> ptr(nzt + 1, 0), pts;
    std::vector<OpsWPoint> wp(np);
    for (int a = 0; a < nzt; a++) {
        const int k0 = c->op.h_age_ptr[a], k1 = c->op.h_age_ptr[a + 1];
        bool chain = false;
        int pseg = 0;
        double pm2 = 0.0;
        for (int k = k0; k < k1; k++) {
            const double* q = &c->op.h_pt[(size_t)k * OPS_PT_N];
            const double m1 = q[1], m2 = q[2];
            OpsWPoint& r = wp[k];
            r = OpsWPoint();
            r.mini = q[0];
            r.llo = q[3];
            r.dl = q[4];
            r.cm = q[5];
            r.cl = q[6];
            if ((k - k0) % OPS_PIECE == 0) chain = false;
            const int first = c->op.h_first[k] == 1;
            const int seg = (first || !(m2 > m1)) ? 0 : bin_class(ty, m1, m2, c->mdave);
            if (seg == 0) {
                pts.push_back(k);
                r.code = 0;
                if (first || m2 > m1) chain = false;
                continue;
            }
            r.code = seg | ((chain && seg == pseg && m1 == pm2) ? 4 : 0);
            chain = true;
            pseg = seg;
            pm2 = m2;
        }
        ptr[a + 1] = (int)pts.size();
    }
    c->op.nslow = (int)pts.size();
    if (pts.empty()) pts.push_back(0);
    UPLOAD(c, c->op.d_slow_ptr, ptr.data(), ptr.size());
    UPLOAD(c, c->op.d_slow_pt, pts.data(), pts.size());
    UPLOAD(c, c->op.d_wp, wp.data(), wp.size());
    double poly[13], f = 1.0;
    for (int j = 0; j < 13; j++) {
        f *= (double)(j + 1);
        poly[j] = 1.0 / f;
    }
    UPLOAD(c, c->op.d_poly, poly, 13);
    return 0;
}

int opensps_hip_set_ssp_ops(opensps_hip_ctx_t* c, const opensps_hip_ssp_ops_t* o) {
    if (!c) return 1;
    if (c->nspec < 2) return ops_fail(c, 1, "set_ssp_ops: no grid installed");
    if (!o || !o->age_ptr || !o->mini || !o->m1 || !o->m2 || !o->first || !o->cm || !o->cl || !o->q || !o->age_mmax ||
        !o->ent_tpl || !o->ent_val || !o->etpl)
        return ops_fail(c, 3, "set_ssp_ops: NULL arrays");
    if (o->nz != c->nz || o->nt != c->nt)
        return ops_fail(c, 3, "set_ssp_ops: %d metallicities x %d ages, the grid has %d x %d", o->nz, o->nt, c->nz, c->nt);
    if (o->npts < 1 || o->ntpl < 1 || o->nent < 1) return ops_fail(c, 3, "set_ssp_ops: empty operator");
    const int nz = o->nz, nt = o->nt, np = o->npts, ntpl = o->ntpl, ne = o->nent, nspec = c->nspec;
    if (o->age_ptr[0] != 0 || o->age_ptr[nz * nt] != np) return ops_fail(c, 3, "set_ssp_ops: age_ptr must span 0 .. npts");
    for (int a = 0; a < nz * nt; a++)
        if (o->age_ptr[a + 1] < o->age_ptr[a]) return ops_fail(c, 3, "set_ssp_ops: age_ptr decreases at %d", a);
    for (size_t k = 0; k < (size_t)np * ne; k++)
        if (o->ent_tpl[k] < -1 || o->ent_tpl[k] >= ntpl) return ops_fail(c, 3, "set_ssp_ops: template index out of range");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_operator(c);
    // points: the bins' logarithms formed here once (model-independent), OPS_PT_N doubles each
    std::vector<double> pt((size_t)np * OPS_PT_N);
    for (int k = 0; k < np; k++) {
        double* q = &pt[(size_t)k * OPS_PT_N];
        q[0] = o->mini[k];
        q[1] = o->m1[k];
        q[2] = o->m2[k];
        bin_logs(o->m1[k], o->m2[k], &q[3], &q[4]);
        q[5] = o->cm[k];
        q[6] = o->cl[k];
        q[7] = o->q[k];
    }
    std::vector<int> zoff(nz + 1);
    int npz = 1;
    for (int iz = 0; iz <= nz; iz++) zoff[iz] = o->age_ptr[iz * nt];
    for (int iz = 0; iz < nz; iz++) npz = std::max(npz, zoff[iz + 1] - zoff[iz]);
    int npa_max = 1;
    for (int a = 0; a < nz * nt; a++) npa_max = std::max(npa_max, o->age_ptr[a + 1] - o->age_ptr[a]);
    // entries by template per metallicity (CSR), in point order within a template, so sorted by age
    std::vector<int> tz_ptr((size_t)nz * (ntpl + 1), 0);
    for (int iz = 0; iz < nz; iz++)
        for (int k = zoff[iz]; k < zoff[iz + 1]; k++)
            for (int j = 0; j < ne; j++) {
                const int e = o->ent_tpl[(size_t)k * ne + j];
                if (e >= 0) tz_ptr[(size_t)iz * (ntpl + 1) + e + 1]++;
            }
    int tot = 0;
    for (int iz = 0; iz < nz; iz++) {
        int* tp = &tz_ptr[(size_t)iz * (ntpl + 1)];
        tp[0] = tot;
        for (int e = 0; e < ntpl; e++) tp[e + 1] += tp[e];
        tot = tp[ntpl];
    }
    const int nentries = tot;
    std::vector<int2> te(nentries > 0 ? nentries : 1);
    std::vector<double> tv64(nentries > 0 ? nentries : 1);
    std::vector<float> tv32(nentries > 0 ? nentries : 1);
    std::vector<int> fill(tz_ptr);
    for (int iz = 0; iz < nz; iz++)
        for (int t = 0; t < nt; t++)
            for (int k = o->age_ptr[iz * nt + t]; k < o->age_ptr[iz * nt + t + 1]; k++)
                for (int j = 0; j < ne; j++) {
                    const int e = o->ent_tpl[(size_t)k * ne + j];
                    if (e < 0) continue;
                    const int at = fill[(size_t)iz * (ntpl + 1) + e]++;
                    te[at] = make_int2(k - zoff[iz], t);
                    tv64[at] = o->ent_val[(size_t)k * ne + j];
                    tv32[at] = (float)tv64[at];
                }
    // segments: each template's entries in a metallicity cut into runs of at most OPS_SEG
    std::vector<int2> seg;
    std::vector<int> segoff(nz + 1), tseg((size_t)nz * (ntpl + 1));
    int nseg_max = 1;
    for (int iz = 0; iz < nz; iz++) {
        segoff[iz] = (int)seg.size();
        const int* tp = &tz_ptr[(size_t)iz * (ntpl + 1)];
        for (int e = 0; e < ntpl; e++) {
            tseg[(size_t)iz * (ntpl + 1) + e] = (int)seg.size() - segoff[iz];
            for (int j = tp[e]; j < tp[e + 1]; j += OPS_SEG) seg.push_back(make_int2(j, std::min(tp[e + 1], j + OPS_SEG)));
        }
        tseg[(size_t)iz * (ntpl + 1) + ntpl] = (int)seg.size() - segoff[iz];
        nseg_max = std::max(nseg_max, (int)seg.size() - segoff[iz]);
    }
    segoff[nz] = (int)seg.size();
    if (seg.empty()) seg.push_back(make_int2(0, 0));
    c->op.nseg_max = nseg_max;
    c->op.nz = nz;
    c->op.nt = nt;
    c->op.npts = np;
    c->op.ntpl = ntpl;
    c->op.npz = npz;
    c->op.npa_max = npa_max;
    // K padded to whole blocks of the fold (8 templates; a multiple of the f64 instruction's 4 and of 32 for bf16)
    c->op.k64 = ((ntpl + OPS_TPL_BLOCK - 1) / OPS_TPL_BLOCK) * OPS_TPL_BLOCK;
    c->op.kbf = ((ntpl + OPS_KSTEP_BF16 - 1) / OPS_KSTEP_BF16) * OPS_KSTEP_BF16;
    UPLOAD(c, c->op.d_age_ptr, o->age_ptr, (size_t)nz * nt + 1);
    UPLOAD(c, c->op.d_zoff, zoff.data(), zoff.size());
    UPLOAD(c, c->op.d_first, o->first, (size_t)np);
    UPLOAD(c, c->op.d_pt, pt.data(), pt.size());
    UPLOAD(c, c->op.d_age_mmax, o->age_mmax, (size_t)nz * nt);
    UPLOAD(c, c->op.d_seg, seg.data(), seg.size());
    UPLOAD(c, c->op.d_segoff, segoff.data(), segoff.size());
    UPLOAD(c, c->op.d_tseg, tseg.data(), tseg.size());
    UPLOAD(c, c->op.d_te, te.data(), te.size());
    UPLOAD(c, c->op.d_te_val64, tv64.data(), tv64.size());
    UPLOAD(c, c->op.d_te_val32, tv32.data(), tv32.size());
    // the library as the A operand of the template product, packed like a basis of one slab and ntpl "ages"
    double* lib = nullptr;
    HIPCHK(c, hipMalloc((void**)&lib, sizeof(double) * (size_t)ntpl * nspec));
    HIPCHK(c, hipMemcpy(lib, o->etpl, sizeof(double) * (size_t)ntpl * nspec, hipMemcpyHostToDevice));
    const int npxblk = c->nsp / 16;
    PackArgs pa;
    pa.nspec = nspec;
    pa.nt = ntpl;
    pa.nz = 1;
    pa.npair = 1;
    pa.kpair = ntpl;
    pa.npxblk = npxblk;
    pa.basis = lib;
    {
        size_t n64 = (size_t)npxblk * (c->op.k64 / OPS_KSTEP_F64) * 64;
        HIPCHK(c, hipMalloc((void**)&c->op.d_A64, n64 * sizeof(double)));
        pa.kpad = c->op.k64;
        pa.A64 = c->op.d_A64;
        k_pack_f64<<<(unsigned)((n64 + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    {
        size_t nbf = (size_t)npxblk * (c->op.kbf / OPS_KSTEP_BF16) * 64;
        HIPCHK(c, hipMalloc((void**)&c->op.d_Ahi, nbf * 8 * sizeof(uint16_t)));
        HIPCHK(c, hipMalloc((void**)&c->op.d_Alo, nbf * 8 * sizeof(uint16_t)));
        pa.kpad = c->op.kbf;
        pa.Ahi = c->op.d_Ahi;
        pa.Alo = c->op.d_Alo;
        k_pack_bf16<<<(unsigned)((nbf + 255) / 256), 256, 0, c->stream>>>(pa);
        HIPCHK(c, hipGetLastError());
    }
    // the reverse pass's operands: the library as rows (template, pixel) in FP64, each point's entries, h's series
    c->op.nent = ne;
    c->op.nrt = (ntpl + 63) / 64;
    {
        const size_t nel = (size_t)c->op.nrt * 64 * c->nsp;
        HIPCHK(c, hipMalloc((void**)&c->op.d_At64, nel * sizeof(double)));
        k_pack_t<<<(unsigned)((nel + 255) / 256), 256, 0, c->stream>>>(nspec, c->nsp, ntpl, c->op.nrt * 64, lib,
                                                                      c->op.d_At64, nullptr, nullptr);
        HIPCHK(c, hipGetLastError());
    }
    UPLOAD(c, c->op.d_pent, o->ent_tpl, (size_t)np * ne);
    UPLOAD(c, c->op.d_pval, o->ent_val, (size_t)np * ne);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(lib));
    c->op.h_age_ptr.assign(o->age_ptr, o->age_ptr + (size_t)nz * nt + 1);
    c->op.h_first.assign(o->first, o->first + np);
    c->op.h_pt.swap(pt);
    int st = build_imf_tables(c);
    if (st) return st;
    c->op.on = true;
    return 0;
}

int opensps_hip_free_ssp_ops(opensps_hip_ctx_t* c) {
    if (!c) return 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_operator(c);
    return 0;
}

int opensps_hip_set_imf(opensps_hip_ctx_t* c, int imf_type, int remnants, double salp_ind, double mdave,
                        double mlim_bh, double mlim_ns) {
    if (!c) return 1;
    const int t = imf_type % 10;
    if (imf_type < 0 || (t != 0 && t != 2 && t != 4))
        return ops_fail(c, 2, "set_imf: IMF type %d is not a power law this build integrates (0, 2, 4)", imf_type);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool relist = imf_type % 10 != c->imf_type % 10 || mdave != c->mdave;
    c->imf_type = imf_type;
    c->remnants = remnants != 0;
    c->salp_ind = salp_ind;
    c->mdave = mdave;
    c->mlim_bh = mlim_bh;
    c->mlim_ns = mlim_ns;
    if (c->op.on && relist) return build_imf_tables(c);
    return 0;
}

int opensps_hip_set_cosmology(opensps_hip_ctx_t* c, int n, const double* z, const double* dl) {
    if (!c) return 1;
    if (n < 2 || !z || !dl) return ops_fail(c, 3, "set_cosmology: bad table");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ncosmo = n;
    UPLOAD(c, c->d_cz, z, (size_t)n);
    UPLOAD(c, c->d_cdl, dl, (size_t)n);
    return 0;
}

int opensps_hip_set_options(opensps_hip_ctx_t* c, int add_igm, int smooth_fast) {
    if (!c) return 1;
    c->add_igm = add_igm != 0;
    c->smooth_fast = smooth_fast != 0;
    return 0;
}

int opensps_hip_set_emission(opensps_hip_ctx_t* c, int add_dust_emission, int add_agn_dust) {
    if (!c) return 1;
    c->add_duste = add_dust_emission == 1 ? 1 : 0;
    c->add_agn = add_agn_dust == 1 ? 1 : 0;
    return 0;
}

int opensps_hip_set_dust_emission(opensps_hip_ctx_t* c, int nqpah, const double* qpah, int numin, const double* umin,
                                  const double* table) {
    if (!c) return 1;
    if (c->nspec < 2) return ops_fail(c, 1, "set_dust_emission: no grid installed");
    if (nqpah == 0) {
        c->nqpah = 0;
        c->numin = 0;
        return 0;
    }
    if (nqpah < 2 || numin < 2 || !qpah || !umin || !table)
        return ops_fail(c, 3, "set_dust_emission: at least 2 q_PAH and 2 U_min values and the table are required");
    for (int k = 1; k < nqpah; k++)
        if (!(qpah[k] > qpah[k - 1])) return ops_fail(c, 3, "set_dust_emission: q_PAH not strictly ascending");
    for (int k = 1; k < numin; k++)
        if (!(umin[k] > umin[k - 1])) return ops_fail(c, 3, "set_dust_emission: U_min not strictly ascending");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // padded to nsp pixels per column (zero past nspec)
    const int ncol = nqpah * 2 * numin;
    std::vector<double> t((size_t)c->nsp * ncol, 0.0);
    for (int k = 0; k < ncol; k++)
        for (int i = 0; i < c->nspec; i++) t[(size_t)k * c->nsp + i] = table[(size_t)k * c->nspec + i];
    UPLOAD(c, c->d_dustem, t.data(), t.size());
    UPLOAD(c, c->d_qpah, qpah, (size_t)nqpah);
    UPLOAD(c, c->d_umin, umin, (size_t)numin);
    c->nqpah = nqpah;
    c->numin = numin;
    return 0;
}

int opensps_hip_set_agn(opensps_hip_ctx_t* c, int ntau, const double* tau, const double* table) {
    if (!c) return 1;
    if (c->nspec < 2) return ops_fail(c, 1, "set_agn: no grid installed");
    if (ntau == 0) {
        c->ntau = 0;
        return 0;
    }
    if (ntau < 2 || !tau || !table) return ops_fail(c, 3, "set_agn: at least 2 optical depths and the table are required");
    for (int k = 1; k < ntau; k++)
        if (!(tau[k] > tau[k - 1])) return ops_fail(c, 3, "set_agn: optical depths not strictly ascending");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> t((size_t)c->nsp * ntau, 0.0);
    for (int k = 0; k < ntau; k++)
        for (int i = 0; i < c->nspec; i++) t[(size_t)k * c->nsp + i] = table[(size_t)k * c->nspec + i];
    UPLOAD(c, c->d_agn, t.data(), t.size());
    UPLOAD(c, c->d_agn_tau, tau, (size_t)ntau);
    c->ntau = ntau;
    return 0;
}

int opensps_hip_set_neb_options(opensps_hip_ctx_t* c, int add_neb_emission, int add_neb_continuum, int nebemlineinspec) {
    if (!c) return 1;
    c->add_neb = add_neb_emission == 1 ? 1 : 0;
    c->neb_cont_on = add_neb_continuum == 1 ? 1 : 0;
    c->neb_lines_in_spec = nebemlineinspec == 1 ? 1 : 0;
    return 0;
}

// The CLOUDY library placed on the grid as load_nebular_grid places it (src/io/opensps_io.f90:749-823): log10(x +
// 1e-95), the continuum linear in wavelength onto the grid (the library's end intervals extrapolated, as
// interpolate_linear_scalar does), the age axis to log10; then what nebular_setup derives from the grid (:1630-1770):
// the nebular ages (those up to the library's last age), each one's CLOUDY interval and weight, the weight's steps
// within an interval, the pixels below the Lyman limit, each line's pixel interval; and step 1 of nebular_kernels,
// the ionizing photons of every nebular basis spectrum (k_neb_q0)
int opensps_hip_set_nebular(opensps_hip_ctx_t* c, int nnz, const double* logz, int nna, const double* age, int nnu,
                            const double* logu, int nlam, const double* lam, const double* cont, int nline,
                            const double* line_lam, const double* lines) {
    if (!c) return 1;
    if (c->nspec < 2) return ops_fail(c, 1, "set_nebular: no grid installed");
    if (nnz == 0) {
        c->nnz = 0;
        return 0;
    }
    if (nnz < 2 || nna < 2 || nnu < 2 || nlam < 2 || nline < 1 || !logz || !age || !logu || !lam || !cont ||
        !line_lam || !lines)
        return ops_fail(c, 3, "set_nebular: at least 2 values on every axis, one line and every array are required");
    const double* axes[4] = {logz, age, logu, lam};
    const int na[4] = {nnz, nna, nnu, nlam};
    for (int k = 0; k < 4; k++)
        for (int i = 1; i < na[k]; i++)
            if (!(axes[k][i] > axes[k][i - 1])) return ops_fail(c, 3, "set_nebular: axis %d not strictly ascending", k);
    for (int i = 0; i < nna; i++)
        if (!(age[i] > 0.0)) return ops_fail(c, 3, "set_nebular: ages must be positive (yr)");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nspec = c->nspec, nt = c->nt;
    const double floor_ = 1.0e-95;
    // the tables: cont (nnu, nna, nnz, nspec) and lines (nnu, nna, nnz, nline) in C order, from the library's
    // (nnz, nna, nnu, nlam) and (nnz, nna, nnu, nline)
    std::vector<double> tc((size_t)nnu * nna * nnz * nspec), tl((size_t)nnu * nna * nnz * nline), lc(nlam);
    for (int i = 0; i < nnz; i++)
        for (int j = 0; j < nna; j++)
            for (int k = 0; k < nnu; k++) {
                const double* src = cont + (((size_t)i * nna + j) * nnu + k) * nlam;
                for (int l = 0; l < nlam; l++) lc[l] = log10(src[l] + floor_);
                double* dst = tc.data() + (((size_t)k * nna + j) * nnz + i) * nspec;
                for (int p = 0; p < nspec; p++) {
                    const double x = c->h_lam[p];
                    int q = find_interval_ref(lam, nlam, x);
                    q = std::max(1, std::min(q, nlam - 1));
                    const double slope = (lc[q] - lc[q - 1]) / (lam[q] - lam[q - 1]);
                    dst[p] = lc[q - 1] + slope * (x - lam[q - 1]);
                }
                const double* ls = lines + (((size_t)i * nna + j) * nnu + k) * nline;
                double* ld = tl.data() + (((size_t)k * nna + j) * nnz + i) * nline;
                for (int l = 0; l < nline; l++) ld[l] = log10(ls[l] + floor_);
            }
    std::vector<double> lage(nna);
    for (int j = 0; j < nna; j++) lage[j] = log10(age[j]);
    // the nebular ages and their CLOUDY intervals
    const int ntn = std::max(0, std::min(find_interval_ref(c->h_time.data(), nt, lage[nna - 1]), nt));
    const int n1 = std::max(1, ntn);
    std::vector<int> ia(n1, 1), newr(n1, 0);
    std::vector<double> wa(n1, 0.0), dwa(n1, 0.0);
    for (int t = 0; t < ntn; t++) {
        int i = std::max(1, std::min(find_interval_ref(lage.data(), nna, c->h_time[t]), nna - 1));
        ia[t] = i;
        wa[t] = std::max(0.0, std::min((c->h_time[t] - lage[i - 1]) / (lage[i] - lage[i - 1]), 1.0));
    }
    for (int t = 1; t < ntn; t++) {
        if (ia[t] != ia[t - 1]) continue;
        dwa[t] = wa[t] - wa[t - 1];
        newr[t] = (ia[t - 1] != ia[std::max(0, t - 2)] || t == 1 || dwa[t] != dwa[t - 1]) ? 1 : 0;
    }
    // each line's pixel interval (1-based, 1 .. nspec - 1)
    std::vector<int> llo(nline);
    for (int i = 0; i < nline; i++) {
        int jc = 1, j = nspec + 1;
        while (j - jc > 1) {
            const int p = (jc + j) / 2;
            if (line_lam[i] >= c->h_lam[p - 1]) jc = p;
            else j = p;
        }
        llo[i] = std::max(1, std::min(jc, nspec - 1));
    }
    UPLOAD(c, c->d_neb_cont, tc.data(), tc.size());
    UPLOAD(c, c->d_neb_line, tl.data(), tl.size());
    UPLOAD(c, c->d_neb_lz, logz, (size_t)nnz);
    UPLOAD(c, c->d_neb_lu, logu, (size_t)nnu);
    UPLOAD(c, c->d_neb_lpos, line_lam, (size_t)nline);
    UPLOAD(c, c->d_neb_ia, ia.data(), ia.size());
    UPLOAD(c, c->d_neb_newr, newr.data(), newr.size());
    UPLOAD(c, c->d_neb_wa, wa.data(), wa.size());
    UPLOAD(c, c->d_neb_dwa, dwa.data(), dwa.size());
    UPLOAD(c, c->d_neb_llo, llo.data(), llo.size());
    c->h_neb_lpos.assign(line_lam, line_lam + nline);
    c->nnz = nnz;
    c->nna = nna;
    c->nnu = nnu;
    c->nline = nline;
    c->nt_neb = ntn;
    c->whlylim = find_interval_ref(c->h_lam.data(), nspec, 912.0);
    c->neb_sigma = -1.0;   // the profiles: built again at the next batch
    std::vector<double> zero((size_t)c->nz * n1, 0.0);
    UPLOAD(c, c->d_neb_q0, zero.data(), zero.size());
    if (ntn > 0) {
        k_neb_q0<<<dim3((unsigned)ntn, (unsigned)c->nz), 256, 0, c->stream>>>(nspec, nt, ntn, c->whlylim, c->d_lam,
                                                                                c->d_basis64, c->d_neb_q0);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

// The lines' profiles for one width (compute_line_gaussians, opensps_gas.f90:298-340, velocity smoothing, at least
// twice the pixel spacing at the line, compute_nebular_kernels, opensps_initialization.f90:1377-1385), each from the
// pixel nearest its centre until it falls to 1e-14 of its peak (nebular_setup :1725-1751), as CSR by pixel in line
// order
static int neb_profiles(OpsCtx* c, double sigma) {
    if (c->neb_sigma == sigma && c->d_neb_lptr) return 0;
    const int nspec = c->nspec, nl = c->nline;
    const double* lam = c->h_lam.data();
    const double pi = acos(-1.0), sq2pi = sqrt(2.0 * pi), kms = 1.0e13 / OPS_C_LIGHT;
    std::vector<int> jlo(nl, 1), jhi(nl, 0);
    std::vector<std::vector<double>> prof(nl);
    std::vector<double> g(nspec);
    for (int i = 0; i < nl; i++) {
        const double pos = c->h_neb_lpos[i];
        int jr = std::max(1, std::min(find_interval_ref(lam, nspec, pos), nspec - 1));
        const double res = lam[jr] - lam[jr - 1];
        double sig = pos * sigma * kms;
        sig = std::max(sig, res * 2.0);
        const double norm = (1.0 / (sq2pi * sig)) * (pos * pos / OPS_C_LIGHT);
        for (int k = 0; k < nspec; k++) {
            const double x = (lam[k] - pos) / sig;
            g[k] = norm * exp(-0.5 * x * x);
        }
        const int jc = std::max(1, std::min(find_interval_ref(lam, nspec, pos), nspec));
        const int jh = std::min(jc + 1, nspec);
        const double cut = 1.0e-14 * std::max(g[jc - 1], g[jh - 1]);
        if (cut <= 0.0) continue;
        int lo = jc, hi = jh;
        while (lo > 1 && g[lo - 2] > cut) lo--;
        while (hi < nspec && g[hi] > cut) hi++;
        jlo[i] = lo;
        jhi[i] = hi;
        prof[i].assign(g.begin() + (lo - 1), g.begin() + hi);
    }
    std::vector<int> lptr(nspec + 1, 0);
    for (int i = 0; i < nl; i++)
        for (int j = jlo[i]; j <= jhi[i]; j++) lptr[j]++;   // (pixel j - 1 counted at lptr[j])
    for (int j = 0; j < nspec; j++) lptr[j + 1] += lptr[j];
    const int tot = lptr[nspec];
    std::vector<int> lid(std::max(1, tot)), fill(lptr.begin(), lptr.end() - 1);
    std::vector<double> gval(std::max(1, tot));
    for (int i = 0; i < nl; i++)
        for (int j = jlo[i]; j <= jhi[i]; j++) {
            const int p = fill[j - 1]++;
            lid[p] = i;
            gval[p] = prof[i][j - jlo[i]];
        }
    UPLOAD(c, c->d_neb_lptr, lptr.data(), lptr.size());
    UPLOAD(c, c->d_neb_lid, lid.data(), lid.size());
    UPLOAD(c, c->d_neb_gval, gval.data(), gval.size());
    c->neb_nprof = tot;
    c->neb_sigma = sigma;
    return 0;
}

int opensps_hip_csp_batch(opensps_hip_ctx_t* c, const opensps_hip_batch_t* a) {
    if (!c) return 1;
    if (!a) return ops_fail(c, 3, "batch: NULL arguments");
    if (c->nspec < 2) return ops_fail(c, 1, "batch: no SSP grid installed");
    if (c->ncosmo < 2) return ops_fail(c, 1, "batch: no cosmology table installed");
    const int n = a->n;
    if (n < 0) return ops_fail(c, 3, "batch: n < 0");
    if (n == 0) return 0;
    if (!a->prm || (!a->mags && c->nbands > 0)) return ops_fail(c, 3, "batch: prm and mags are required");
    if (a->precision != OPENSPS_HIP_FP64 && a->precision != OPENSPS_HIP_MIXED)
        return ops_fail(c, 3, "batch: unknown precision %d", a->precision);
    const int parts = a->split_parts ? 2 : 1;
    if ((a->young || a->old) && parts != 2) return ops_fail(c, 3, "batch: young / old need split_parts = 1");
    const bool mixed = a->precision == OPENSPS_HIP_MIXED;
    const bool ops = a->use_ops != 0;
    if (ops && !c->op.on) return ops_fail(c, 1, "batch: no SSP operator installed (opensps_hip_set_ssp_ops)");
    const bool wz = a->ssp_weights_z != nullptr;
    if (wz && ops) return ops_fail(c, 2, "batch: ssp_weights_z with the SSP operator (the reference refuses it too)");
    if (wz && a->ssp_weights) return ops_fail(c, 3, "batch: ssp_weights and ssp_weights_z both given");
    // nebular emission (the CLOUDY grids): with the resident bases and their SFH or given weights; not with the IMF
    // per model or weights over metallicities (their nebular paths are not built), nor with lines outside the spectrum
    // while dust emission is on (the lines' absorbed light in its energy balance is not built), nor from the fixed
    // scratch (the profiles' tables are allocated)
    const bool neb = c->add_neb != 0;
    if (neb) {
        if (c->nnz == 0) return ops_fail(c, 2, "batch: nebular emission without its tables (opensps_hip_set_nebular)");
        if (ops || wz) return ops_fail(c, 2, "batch: nebular emission with the SSP operator or ssp_weights_z is not built");
        if (!c->neb_lines_in_spec && c->add_duste)
            return ops_fail(c, 2, "batch: nebular lines outside the spectrum with dust emission are not built");
        if (c->arena) return ops_fail(c, 2, "batch_noalloc: nebular emission allocates its profiles (use opensps_hip_csp_batch)");
    }
    int st;
    if (wz && (st = ensure_full(c, mixed, false))) return st;
    const bool uni = ops || wz;   // one A operand for every column (uniform column tiles)
    const int kpad = ops ? (mixed ? c->op.kbf : c->op.k64)
                         : wz ? (mixed ? c->full.kbf : c->full.k64) : (mixed ? c->kbf : c->k64);
    const int npxblk = c->nsp / 16, npxt = c->nsp / OPS_TILE_PX;
    const size_t tsz = mixed ? sizeof(float) : sizeof(double);
    // column tiles: by metallicity pair over model columns (the operator groups models, parts = 1), then with the
    // operator the template product's uniform tiles after them
    const int gparts = ops ? 1 : parts;
    const int maxtiles = (n * gparts + OPS_TILE_COL - 1) / OPS_TILE_COL + c->npair;
    const int optiles = (n * parts + OPS_TILE_COL - 1) / OPS_TILE_COL;
    // the split-BF16 products' wider tiles (k_mm_bf16x3)
    const int maxm = mixed ? (n * gparts + OPS_MM_COLS - 1) / OPS_MM_COLS + c->npair : 0;
    const int optm = mixed && uni ? (n * parts + OPS_MM_COLS - 1) / OPS_MM_COLS : 0;
    HIPCHK(c, hipSetDevice(c->device));

    // scratch: misc = zlo, it, status, order (n ints each) | ntiles, flag, tile counts (64) | pair slot counts
    // (256 x OPS_GCNT_SLOTS) | rank (n)
    if (c->npair > 256) return ops_fail(c, 3, "batch: more than 257 metallicities");
    const size_t ncnt = 64 + (size_t)256 * OPS_GCNT_SLOTS;
    if ((st = ops_reserve(c, c->misc, sizeof(int) * ((size_t)5 * n + ncnt)))) return st;
    if ((st = ops_reserve(c, c->tiles, sizeof(int4) * (size_t)(maxtiles + (uni ? optiles : 0) + maxm + optm)))) return st;
    if ((st = ops_reserve(c, c->hdr, sizeof(double) * OPS_HDR_N * (size_t)n))) return st;
    if ((st = ops_reserve(c, c->yo, tsz * (size_t)parts * n * c->nsp))) return st;
    if (mixed) {
        if ((st = ops_reserve(c, c->wcol, sizeof(uint16_t) * 2 * (size_t)n * parts * kpad))) return st;
    } else {
        if ((st = ops_reserve(c, c->w64, sizeof(double) * (size_t)n * parts * kpad))) return st;
    }
    int* misc = (int*)c->misc.p;
    int *d_zlo = misc, *d_it = misc + n, *d_status = misc + 2 * n, *d_order = misc + 3 * n;
    int* d_ntiles = misc + 4 * n;
    int* d_flag = misc + 4 * n + 16;
    int* d_optiles = misc + 4 * n + 32;
    int* d_nmtiles = misc + 4 * n + 40;
    int* d_opmtiles = misc + 4 * n + 48;
    int* d_gcnt = misc + 4 * n + 64;
    int* d_rank = misc + 4 * n + ncnt;
    int4* mtiles = (int4*)c->tiles.p + maxtiles + (uni ? optiles : 0);   // then the uniform ones
    const int npass = (ops && c->nz > 1) ? 2 : 1;
    // the operator's weights, segments and fold take every column tile in one round (rounds of fewer tiles, their
    // scratch reused while still in the last-level cache, measured slower: too few wavefronts a round)
    const int chunk = ops ? std::max(1, std::min(c->op_chunk, maxtiles)) : 1;
    const int ngroup = ops ? (c->op.npa_max + OPS_OP_GPTS - 1) / OPS_OP_GPTS : 1;
    const size_t ngp = (size_t)npass * chunk * c->nt * ngroup * 64;
    if (ops) {
        if ((st = ops_reserve(c, c->wsfh, sizeof(double) * (size_t)n * c->nt))) return st;
        if ((st = ops_reserve(c, c->opu, tsz * (size_t)npass * chunk * c->op.npz * 64))) return st;
        if ((st = ops_reserve(c, c->opwc, tsz * (size_t)npass * maxtiles * c->nt * 64))) return st;
        if ((st = ops_reserve(c, c->opsum, sizeof(double) * 2 * (size_t)npass * n * c->nt))) return st;
        if ((st = ops_reserve(c, c->opseg, 2 * tsz * (size_t)npass * chunk * c->op.nseg_max * 64))) return st;
        if ((st = ops_reserve(c, c->opgp, (2 * sizeof(double) + sizeof(int)) * ngp))) return st;
        if ((st = ops_reserve(c, c->opimf, sizeof(double) * OPS_IMFC_N * (size_t)n))) return st;
        if ((st = ops_reserve(c, c->opspart, (3 * sizeof(double) + sizeof(int)) * (size_t)npass * maxtiles * c->nt * 64)))
            return st;
    }
    // The chain's layout (one LDS spectrum buffer when the band quads fit the registers of the single-buffer chain,
    // else two) and its staging rows for smoothing windows wider than the registers hold: reserved before the first
    // launch, so that every scratch failure returns before any work is queued
    const int chain_threads = 512;
    const int qbands = c->band_j1 >= c->band_j0 ? (c->band_j1 + 1 - (c->band_j0 & ~3)) / 4 : 0;
    const bool one = qbands <= OPS_CHAIN_KQ * chain_threads && !c->chain_two;
    if (one && !a->spec)
        if ((st = ops_reserve(c, c->spec, tsz * (size_t)n * c->nspec))) return st;
    // the counters and the pair slot counts start at zero (a fill kernel: no memset node where a caller captures)
    k_fill_i32<<<(unsigned)((ncnt + 255) / 256), 256, 0, c->stream>>>(d_ntiles, (int)ncnt, 0);
    HIPCHK(c, hipGetLastError());
    c->d_last_flag = d_flag;
    c->flag_async = false;

    PrepArgs pa;
    pa.n = n;
    pa.nt = c->nt;
    pa.nz = c->nz;
    pa.parts = parts;
    pa.kpair = wz ? c->full.kfull : c->kpair;
    pa.kpad = kpad;
    pa.mixed = mixed;
    pa.smooth_fast = c->smooth_fast;
    pa.ops = ops;
    pa.wz = a->ssp_weights_z;
    pa.zmags = a->zred_mags;
    pa.lsum = c->lsum_out;
    pa.tiny_logt = c->tiny_logt;
    pa.prm = a->prm;
    pa.weights_in = a->ssp_weights;
    pa.time = c->d_time;
    pa.logz = c->d_logz;
    pa.bmass = c->d_bmass;
    pa.blbol = c->d_blbol;
    double* nebw = nullptr;
    if (neb && !a->weights) {
        if ((st = ops_reserve(c, c->nebw, sizeof(double) * (size_t)n * c->nt))) return st;
        nebw = (double*)c->nebw.p;
    }
    pa.w_out = ops ? (double*)c->wsfh.p : (a->weights ? a->weights : nebw);
    pa.mass = a->mass;
    pa.lbol = a->lbol;
    pa.w64 = mixed ? nullptr : (double*)c->w64.p;
    pa.whi = mixed ? (uint16_t*)c->wcol.p : nullptr;
    pa.wlo = mixed ? (uint16_t*)c->wcol.p + (size_t)n * parts * kpad : nullptr;
    pa.zlo = d_zlo;
    pa.gcnt = d_gcnt;
    pa.rank = d_rank;
    pa.tab = c->d_tab;
    pa.ntab = c->ntab;
    pa.it = d_it;
    pa.idx = a->idx;
    pa.status = d_status;
    pa.flag = d_flag;
    LamLookup lk;
    lk.nspec = c->nspec;
    lk.nbk = c->nbk;
    lk.lnl0 = c->lnl0;
    lk.bkinv = c->bkinv;
    lk.lam = c->d_lam;
    lk.bkt = c->d_bkt;
    pa.add_igm = c->add_igm;
    pa.jtop = c->jtop;
    pa.ncosmo = c->ncosmo;
    pa.lk = lk;
    pa.igm_a = c->d_igm_a;
    pa.cz = c->d_cz;
    pa.cdl = c->d_cdl;
    pa.hdr = (double*)c->hdr.p;
    size_t lds_prep = sizeof(double) * 4 * 5 * (size_t)(c->nt + 1);
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    k_prep<<<(n + 3) / 4, 256, lds_prep, c->stream>>>(pa);
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipGetLastError());

    GroupArgs ga;
    ga.n = n;
    ga.npair = c->npair;
    ga.parts = gparts;
    ga.zlo = d_zlo;
    ga.gcnt = d_gcnt;
    ga.rank = d_rank;
    ga.order = d_order;
    ga.ntiles = d_ntiles;
    ga.tiles = (int4*)c->tiles.p;
    ga.nmtiles = mixed && !uni ? d_nmtiles : nullptr;
    ga.mtiles = mixed && !uni ? mtiles : nullptr;
    k_group<<<(unsigned)max(1, (n + 1023) / 1024), 1024, 0, c->stream>>>(ga);
    HIPCHK(c, hipGetLastError());
    if (uni && mixed) {
        k_op_tiles<<<1, 256, 0, c->stream>>>(n * parts, OPS_MM_COLS, d_opmtiles, mtiles + maxm);
        HIPCHK(c, hipGetLastError());
    }
    if (wz) {
        k_op_tiles<<<1, 256, 0, c->stream>>>(n * parts, OPS_TILE_COL, d_optiles, (int4*)c->tiles.p + maxtiles);
        HIPCHK(c, hipGetLastError());
    }
    if (ops) {
        if (a->weights)
            HIPCHK(c, hipMemcpyAsync(a->weights, c->wsfh.p, sizeof(double) * (size_t)n * c->nt, hipMemcpyDeviceToDevice,
                                     c->stream));
        double* smass_p = (double*)c->opsum.p;
        double* slum_p = smass_p + (size_t)npass * n * c->nt;
        OpWArgs w;
        w.n = n;
        w.nt = c->nt;
        w.npass = npass;
        w.npz = c->op.npz;
        w.maxtiles = maxtiles;
        w.imf_type = c->imf_type;
        w.remnants = c->remnants;
        w.salp_ind = c->salp_ind;
        w.mdave = c->mdave;
        w.mlim_bh = c->mlim_bh;
        w.mlim_ns = c->mlim_ns;
        w.ntiles = d_ntiles;
        w.tiles = (const int4*)c->tiles.p;
        w.order = d_order;
        w.prm = a->prm;
        w.hdr = (const double*)c->hdr.p;
        w.wsfh = (const double*)c->wsfh.p;
        w.age_ptr = c->op.d_age_ptr;
        w.zoff = c->op.d_zoff;
        w.first = c->op.d_first;
        w.pt = c->op.d_pt;
        w.age_mmax = c->op.d_age_mmax;
        w.u = c->opu.p;
        w.wc = c->opwc.p;
        w.smass_p = smass_p;
        w.slum_p = slum_p;
        w.count = a->op_count;
        w.impc = (double*)c->opimf.p;
        w.slow_ptr = c->op.d_slow_ptr;
        w.slow_pt = c->op.d_slow_pt;
        w.wp = c->op.d_wp;
        w.poly = c->op.d_poly;
        w.spart = (double*)c->opspart.p;
        w.scnt = (int*)(w.spart + 3 * (size_t)npass * maxtiles * c->nt * 64);
        w.tile0 = 0;
        w.utiles = chunk;
        w.ngroup = ngroup;
        w.u_f32 = mixed ? 1 : 0;
        w.gsum = (double*)c->opgp.p;
        w.gcnt = (int*)(w.gsum + 2 * ngp);
        w.sraw = c->op_sraw_out;
        if (a->op_count) {
            k_fill_i32<<<(unsigned)((2 * (size_t)n + 255) / 256), 256, 0, c->stream>>>(a->op_count, 2 * n, 0);
            HIPCHK(c, hipGetLastError());
        }
        const bool optime = c->profiling && chunk >= maxtiles;   // (one round: its phases are timed)
        if (optime) HIPCHK(c, hipEventRecord(c->opev[0], c->stream));
        k_op_imf<<<(n + 255) / 256, 256, 0, c->stream>>>(w);
        HIPCHK(c, hipGetLastError());
        OpGArgs g;
        g.nt = c->nt;
        g.npass = npass;
        g.npz = c->op.npz;
        g.maxtiles = maxtiles;
        g.nseg_max = c->op.nseg_max;
        g.tile0 = 0;
        g.utiles = chunk;
        g.ntiles = d_ntiles;
        g.tiles = (const int4*)c->tiles.p;
        g.order = d_order;
        g.it = d_it;
        g.segoff = c->op.d_segoff;
        g.seg = c->op.d_seg;
        g.te = c->op.d_te;
        g.te_val64 = c->op.d_te_val64;
        g.te_val32 = c->op.d_te_val32;
        g.u = c->opu.p;
        g.wc = c->opwc.p;
        g.part = c->opseg.p;
        OpFArgs f;
        f.ntpl = c->op.ntpl;
        f.npass = npass;
        f.maxtiles = maxtiles;
        f.nseg_max = c->op.nseg_max;
        f.kpad = kpad;
        f.parts = parts;
        f.tile0 = 0;
        f.utiles = chunk;
        f.ntiles = d_ntiles;
        f.tiles = (const int4*)c->tiles.p;
        f.order = d_order;
        f.tseg = c->op.d_tseg;
        f.part = c->opseg.p;
        f.whi = mixed ? (uint16_t*)c->wcol.p : nullptr;
        f.wlo = mixed ? (uint16_t*)c->wcol.p + (size_t)n * parts * kpad : nullptr;
        f.w64 = mixed ? nullptr : (double*)c->w64.p;
        const dim3 gw((c->nt + 3) / 4, chunk, npass), gwg((c->nt * ngroup + 3) / 4, chunk, npass);
        const dim3 gg((c->op.nseg_max + 3) / 4, chunk, npass);
        const dim3 gf((kpad + 4 * OPS_TPL_BLOCK - 1) / (4 * OPS_TPL_BLOCK), chunk);
        // one round a chunk of column tiles: the off-path bins and remnants, the common bins by point groups, the
        // ages' sums, the segments, the fold (each round's tiles past the batch's exit at once)
        for (int t0 = 0; t0 < maxtiles; t0 += chunk) {
            w.tile0 = g.tile0 = f.tile0 = t0;
            if (mixed) {
                k_op_slow<float><<<gw, 256, 0, c->stream>>>(w);
                k_op_weights<float><<<gwg, 256, 0, c->stream>>>(w);
            } else {
                k_op_slow<double><<<gw, 256, 0, c->stream>>>(w);
                k_op_weights<double><<<gwg, 256, 0, c->stream>>>(w);
            }
            if (optime) HIPCHK(c, hipEventRecord(c->opev[1], c->stream));
            k_op_agesums<<<gw, 256, 0, c->stream>>>(w);
            if (mixed) {
                k_op_segments<float><<<gg, 256, 0, c->stream>>>(g);
                k_op_fold<float><<<gf, 256, 0, c->stream>>>(f);
            } else {
                k_op_segments<double><<<gg, 256, 0, c->stream>>>(g);
                k_op_fold<double><<<gf, 256, 0, c->stream>>>(f);
            }
            HIPCHK(c, hipGetLastError());
        }
        OpSArgs sa;
        sa.n = n;
        sa.nt = c->nt;
        sa.npass = npass;
        sa.wsfh = (const double*)c->wsfh.p;
        sa.smass_p = smass_p;
        sa.slum_p = slum_p;
        sa.mass = a->mass;
        sa.lbol = a->lbol;
        sa.lsum = c->lsum_out;
        sa.hdr = (double*)c->hdr.p;
        k_op_sums<<<(n + 3) / 4, 256, 0, c->stream>>>(sa);
        HIPCHK(c, hipGetLastError());
        if (optime) {
            HIPCHK(c, hipEventRecord(c->opev[2], c->stream));
            c->opev_pending = true;
        }
        k_op_tiles<<<1, 256, 0, c->stream>>>(n * parts, OPS_TILE_COL, d_optiles, (int4*)c->tiles.p + maxtiles);
        HIPCHK(c, hipGetLastError());
    }
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));

    GemmArgs gm;
    gm.n = n;
    gm.parts = parts;
    gm.nsp = c->nsp;
    gm.npxt = npxt;
    gm.npxblk = npxblk;
    gm.kpad = kpad;
    gm.order = d_order;
    gm.ntiles = uni ? d_optiles : d_ntiles;
    gm.tiles = (const int4*)c->tiles.p + (uni ? maxtiles : 0);
    gm.A64 = ops ? c->op.d_A64 : wz ? c->full.A64 : c->d_A64;
    gm.w64 = (const double*)c->w64.p;
    gm.yo = c->yo.p;
    const unsigned ptx = (unsigned)((npxt + 7) / 8);
    const unsigned gblocks = 8u * ptx * (unsigned)(uni ? optiles : maxtiles);
    if (mixed) {
        MMArgs mm;
        mm.mblk = npxblk;
        mm.nrt = (npxblk + 7) / 8;
        mm.ksteps = kpad / OPS_KSTEP_BF16;
        mm.mode = OPS_MM_GROUPED;
        mm.n = n;
        mm.parts = parts;
        mm.order = d_order;
        mm.ntiles = uni ? d_opmtiles : d_nmtiles;
        mm.tiles = uni ? mtiles + maxm : mtiles;
        mm.ncols = 0;
        mm.nct = 0;
        mm.Ah = ops ? c->op.d_Ahi : wz ? c->full.Ahi : c->d_Ahi;
        mm.Al = ops ? c->op.d_Alo : wz ? c->full.Alo : c->d_Alo;
        mm.slab = (size_t)npxblk * mm.ksteps * 64 * 8;
        mm.Bh = pa.whi;
        mm.Bl = pa.wlo;
        mm.ldb = kpad;
        mm.C = (float*)c->yo.p;
        mm.ldc = c->nsp;
        if ((st = launch_mm(c, mm, uni ? optm : maxm))) return st;
    } else if (c->f64_vector)
        k_gemm_f64_vec<<<gblocks, 256, 0, c->stream>>>(gm);
    else
        k_gemm_f64<<<gblocks, 256, 0, c->stream>>>(gm);
    HIPCHK(c, hipGetLastError());
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));

    // nebular emission into the age sums (before the emission stages and the chain, which read them)
    if (neb) {
        double sigma0 = 0.0;
        if (c->neb_lines_in_spec) {
            HIPCHK(c, hipMemcpyAsync(&sigma0, a->prm + (size_t)OPS_P_SIGMA_SMOOTH * n, sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if ((st = neb_profiles(c, sigma0))) return st;
        }
        NebArgs na;
        na.n = n;
        na.nspec = c->nspec;
        na.nsp = c->nsp;
        na.nt = c->nt;
        na.nz = c->nz;
        na.parts = parts;
        na.nt_neb = c->nt_neb;
        na.whlylim = c->whlylim;
        na.nline = c->nline;
        na.nnz = c->nnz;
        na.nna = c->nna;
        na.nnu = c->nnu;
        na.cont_on = c->neb_cont_on;
        na.lines_in_spec = c->neb_lines_in_spec;
        na.sigma0 = sigma0;
        na.prm = a->prm;
        na.hdr = (const double*)c->hdr.p;
        na.w = a->weights ? a->weights : nebw;
        na.zlo = d_zlo;
        na.it = d_it;
        na.basis = c->d_basis64;
        na.q0 = c->d_neb_q0;
        na.lz = c->d_neb_lz;
        na.lu = c->d_neb_lu;
        na.cont = c->d_neb_cont;
        na.line = c->d_neb_line;
        na.lpos = c->d_neb_lpos;
        na.ia = c->d_neb_ia;
        na.newr = c->d_neb_newr;
        na.llo = c->d_neb_llo;
        na.wa = c->d_neb_wa;
        na.dwa = c->d_neb_dwa;
        na.lptr = c->d_neb_lptr;
        na.lid = c->d_neb_lid;
        na.gval = c->d_neb_gval;
        na.lam = c->d_lam;
        na.ltab = c->d_ltab;
        na.dcurve = c->dcurve_on ? c->d_dcurve : nullptr;
        na.yo = c->yo.p;
        na.emlines = a->emlines;
        na.status = d_status;
        na.flag = d_flag;
        const size_t lds = sizeof(double) * (3 * (size_t)std::max(1, c->nt_neb) + 2 * (size_t)c->nline);
        if (mixed)
            k_neb<float><<<n, 256, lds, c->stream>>>(na);
        else
            k_neb<double><<<n, 256, lds, c->stream>>>(na);
        HIPCHK(c, hipGetLastError());
    } else if (a->emlines) {
        k_fill_f64<<<(unsigned)(((size_t)n * std::max(1, c->nline) + 255) / 256), 256, 0, c->stream>>>(
            a->emlines, (size_t)n * std::max(1, c->nline), 0.0);
        HIPCHK(c, hipGetLastError());
    }

    // dust emission and the AGN torus: their indices and the energy balance into the chain's header, the dust masses;
    // the chain with its emission stages (EM) when either can act
    const bool em = c->add_duste || c->add_agn;
    if (em || a->mdust) {
        DustArgs da;
        da.n = n;
        da.nspec = c->nspec;
        da.nsp = c->nsp;
        da.parts = parts;
        da.nqpah = c->nqpah;
        da.numin = c->numin;
        da.ntau = c->ntau;
        da.duste_on = c->add_duste;
        da.agn_on = c->add_agn;
        da.qpah = c->d_qpah;
        da.umin = c->d_umin;
        da.agn_tau = c->d_agn_tau;
        da.status = d_status;
        da.flag = d_flag;
        da.prm = a->prm;
        da.ltab = c->d_ltab;
        da.ltab32 = c->d_ltab32;
        da.dcurve = c->dcurve_on ? c->d_dcurve : nullptr;
        da.dcurve32 = c->dcurve_on ? c->d_dcurve32 : nullptr;
        da.dustem = c->d_dustem;
        da.wnu = c->d_wnu;
        da.yo = c->yo.p;
        da.hdr = (double*)c->hdr.p;
        da.mdust = a->mdust;
        if (mixed)
            k_emission<float><<<n, 256, 0, c->stream>>>(da);
        else
            k_emission<double><<<n, 256, 0, c->stream>>>(da);
        HIPCHK(c, hipGetLastError());
    }

    ChainArgs ca;
    ca.n = n;
    ca.nspec = c->nspec;
    ca.nsp = c->nsp;
    ca.parts = parts;
    ca.nbands = c->nbands;
    ca.jtop = c->jtop;
    ca.add_igm = c->add_igm;
    ca.smooth_fast = c->smooth_fast;
    ca.band_j0 = c->band_j0;
    ca.band_j1 = c->nbands > 0 ? c->band_j1 : -1;
    ca.amzl = abs_mag_zeropoint_log();
    ca.lk = lk;
    ca.hdr = (const double*)c->hdr.p;
    ca.rdlam = c->d_rdlam;
    ca.ltab32 = c->d_ltab32;
    ca.wl32 = c->d_wl32;
    ca.sm32 = c->d_sm32;
    ca.prm = a->prm;
    ca.zmags = a->zred_mags;
    ca.dustem = c->d_dustem;
    ca.agn = c->d_agn;
    ca.nqpah = c->nqpah;
    ca.numin = c->numin;
    ca.ntau = c->ntau;
    ca.yo = c->yo.p;
    ca.lam = c->d_lam;
    ca.loglam = c->d_loglam;
    ca.ltab = c->d_ltab;
    ca.wl = c->d_wl;
    ca.dlam = c->d_dlam;
    ca.igm64s = c->d_igm64s;
    ca.igm32s = c->d_igm32s;
    ca.jq = c->jq;
    ca.bmeta = c->d_bmeta;
    ca.mask = c->d_mask;
    ca.bc64 = c->d_bc64;
    ca.bc32 = c->d_bc32;
    ca.spec = a->spec;
    ca.young = a->young;
    ca.old = a->old;
    ca.mags = a->mags;
    ca.fband = a->fband;
    ca.dcurve = c->dcurve_on ? c->d_dcurve : nullptr;
    ca.dcurve32 = c->dcurve_on ? c->d_dcurve32 : nullptr;
    // the band sums (up to 2048 bands) for the magnitudes a thread per band
    ca.sband_staged = (size_t)c->nbands * sizeof(double) <= 16384;
    ca.stage = one && !a->spec ? c->spec.p : nullptr;
    const size_t lds_chain = (one ? 1 : 2) * tsz * (size_t)c->nsp + (ca.sband_staged ? sizeof(double) * c->nbands : 0);
    // the chain's instance: precision, one LDS buffer or two, the emission stages or none
    const void* kfn;
    if (em)
        kfn = mixed ? (one ? (const void*)k_chain<float, true, true> : (const void*)k_chain<float, false, true>)
                    : (one ? (const void*)k_chain<double, true, true> : (const void*)k_chain<double, false, true>);
    else
        kfn = mixed ? (one ? (const void*)k_chain<float, true, false> : (const void*)k_chain<float, false, false>)
                    : (one ? (const void*)k_chain<double, true, false> : (const void*)k_chain<double, false, false>);
    HIPCHK(c, hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_chain));
    {
        void* kargs[] = {(void*)&ca};
        HIPCHK(c, hipLaunchKernel(kfn, dim3(n), dim3(chain_threads), kargs, lds_chain, c->stream));
    }
    if (c->profiling) {
        HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
        c->ev_pending = true;
    }
    if (c->vega_on && c->nbands > 0) {   // the Vega system on the magnitudes (after the chain, on its stream)
        const size_t cnt = (size_t)n * c->nbands;
        k_vega_mags<<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(n, c->nbands, a->mags, c->d_magvega,
                                                                          c->vega_ref, c->vega_anchor);
        HIPCHK(c, hipGetLastError());
    }
    // the flag (final: only k_prep, k_neb and k_emission write it) to pinned host memory behind the batch, so that the
    // sync after it reads the status without a second round trip
    c->flag_async = false;
    if (c->h_flag) {
        HIPCHK(c, hipMemcpyAsync(c->h_flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        c->flag_async = true;
    }
    return 0;
}

// ------------------------------------------------------------------------------------------------
// The batch without allocation: the same launches as opensps_hip_csp_batch, its scratch carved from the device's
// fixed array (OPS_NOALLOC_BYTES) instead of the context's grow-only buffers. The context's own buffers are set
// aside for the call and restored after it, so both entry points can be mixed on one context.
// ------------------------------------------------------------------------------------------------
int opensps_hip_csp_batch_noalloc(opensps_hip_ctx_t* c, const opensps_hip_batch_t* a) {
    if (!c) return 1;
    if (!a) return ops_fail(c, 3, "batch_noalloc: NULL arguments");
    if (a->use_ops || a->ssp_weights_z)
        return ops_fail(c, 2, "batch_noalloc: the SSP operator and ssp_weights_z need more scratch than the fixed "
                              "%zu bytes (use opensps_hip_csp_batch)", OPS_NOALLOC_BYTES);
    void* base = g_scratch_ptr[c->device];
    if (!base) return ops_fail(c, 1, "batch_noalloc: no fixed scratch on device %d", c->device);
    OpsBuffer* mine[] = {&c->yo, &c->wcol, &c->w64, &c->tiles, &c->misc, &c->spec, &c->hdr};
    const int nb = (int)(sizeof(mine) / sizeof(mine[0]));
    OpsBuffer saved[sizeof(mine) / sizeof(mine[0])];
    for (int i = 0; i < nb; i++) {
        saved[i] = *mine[i];
        *mine[i] = OpsBuffer();
    }
    OpsArena arena;
    arena.base = (char*)base;
    arena.cap = OPS_NOALLOC_BYTES;
    c->arena = &arena;
    const int st = opensps_hip_csp_batch(c, a);
    c->arena = nullptr;
    for (int i = 0; i < nb; i++) *mine[i] = saved[i];
    return st;
}

// ------------------------------------------------------------------------------------------------
// The reverse pass: the forward batch (keeping its age sums, band fluxes, SFH weights and L_bol sums), then
// k_grad_chain -> the transposed product -> k_grad_out, on the context's stream
// ------------------------------------------------------------------------------------------------
int opensps_hip_csp_batch_grad(opensps_hip_ctx_t* c, const opensps_hip_grad_t* g) {
    if (!c) return 1;
    if (!g) return ops_fail(c, 3, "batch_grad: NULL arguments");
    const opensps_hip_batch_t& f = g->fwd;
    const int n = f.n;
    const bool ops = f.use_ops != 0;
    if (ops && !c->op.on) return ops_fail(c, 1, "batch_grad: no SSP operator installed (opensps_hip_set_ssp_ops)");
    if (ops && f.ssp_weights_z) return ops_fail(c, 2, "batch_grad: ssp_weights_z with the SSP operator");
    if (c->dcurve_on) return ops_fail(c, 2, "batch_grad: the reverse pass with a tabulated dust curve is not built");
    if (c->add_neb) return ops_fail(c, 2, "batch_grad: the reverse pass with nebular emission is not built");
    if (g->weights_z_bar && !f.ssp_weights_z) return ops_fail(c, 3, "batch_grad: weights_z_bar needs ssp_weights_z");
    if (n < 0) return ops_fail(c, 3, "batch_grad: n < 0");
    if (n == 0) return 0;
    if (c->nsp % 32 != 0) return ops_fail(c, 3, "batch_grad: padded grid not a multiple of 32");
    const bool mixed = f.precision == OPENSPS_HIP_MIXED;
    // young and old sums always apart: the derivatives of the screens' parameters (dust1, dust3, frac_obrun) need
    // the split where the forward does not (at dust1 = 0 or dust3 = 0, one-sided, as atten_adjoint_chunk)
    const int parts = 2, nsp = c->nsp, nt = c->nt, nz = c->nz;
    const bool wz = f.ssp_weights_z != nullptr;
    int st;
    if (!ops && (st = ensure_full(c, mixed, true))) return st;
    if (ops && mixed && !c->op.d_Ath) {
        // the library's rows in split BF16 (from the FP64 rows packed at set_ssp_ops), once
        const size_t nel = (size_t)c->op.nrt * 64 * nsp;
        HIPCHK(c, hipMalloc((void**)&c->op.d_Ath, nel * sizeof(uint16_t)));
        HIPCHK(c, hipMalloc((void**)&c->op.d_Atl, nel * sizeof(uint16_t)));
        k_pack_t<<<(unsigned)((nel + 255) / 256), 256, 0, c->stream>>>(nsp, nsp, c->op.ntpl, c->op.nrt * 64,
                                                                      c->op.d_At64, nullptr, c->op.d_Ath, c->op.d_Atl);
        HIPCHK(c, hipGetLastError());
    }
    // the transposed product's rows: every (age, metallicity) of the basis, or the operator's templates
    const int kfull = ops ? c->op.ntpl : c->full.kfull, nrt = ops ? c->op.nrt : c->full.nrt, ldc = nrt * 64;
    const int ncols = n * parts;
    const int npass = (ops && nz > 1) ? 2 : 1;
    const size_t tsz = mixed ? sizeof(float) : sizeof(double);
    if (ops) {
        if ((st = ops_reserve(c, c->opsraw, sizeof(double) * 2 * (size_t)npass * n * nt))) return st;
        if ((st = ops_reserve(c, c->opct, tsz * 2 * (size_t)c->op.ntpl * n))) return st;
        // the (pass, age, position) sums, then k_op_adj_points's per block of 4 point groups
        const size_t ngb = (size_t)((c->op.npa_max + OPS_OP_GPTS - 1) / OPS_OP_GPTS + 3) / 4;
        if ((st = ops_reserve(c, c->oppart, sizeof(double) * 4 * (size_t)npass * nt * n * (1 + ngb)))) return st;
        if ((st = ops_reserve(c, c->opimfd, sizeof(double) * OPS_IMFD_N * (size_t)n))) return st;
    }
    if ((st = ops_reserve(c, c->glsum, sizeof(double) * (size_t)n))) return st;
    if ((st = ops_reserve(c, c->gfband, sizeof(double) * (size_t)n * (c->nbands > 0 ? c->nbands : 1)))) return st;
    if ((st = ops_reserve(c, c->gw, sizeof(double) * (size_t)n * nt))) return st;
    if ((st = ops_reserve(c, c->gout, sizeof(double) * OPS_GOUT_N * (size_t)n))) return st;
    if ((st = ops_reserve(c, c->gyb, (mixed ? 2 * sizeof(uint16_t) : sizeof(double)) * (size_t)ncols * nsp))) return st;
    if ((st = ops_reserve(c, c->gC, (mixed ? sizeof(float) : sizeof(double)) * (size_t)ncols * ldc))) return st;
    // the SFH rows (batch_sfh_gradients, opensps_csp_batch.f90:1163-1166: only where the weights are not given) need
    // d L / d w per age
    const bool sfh_rows = g->params_bar && !f.ssp_weights && !wz;
    if (sfh_rows && !g->weights_bar && (st = ops_reserve(c, c->gwb, sizeof(double) * (size_t)n * nt))) return st;
    // the forward batch
    opensps_hip_batch_t fw = f;
    fw.split_parts = 1;
    if (!fw.fband) fw.fband = (double*)c->gfband.p;
    if (!fw.weights) fw.weights = (double*)c->gw.p;
    c->lsum_out = (double*)c->glsum.p;
    c->op_sraw_out = ops ? (double*)c->opsraw.p : nullptr;
    st = opensps_hip_csp_batch(c, &fw);
    c->lsum_out = nullptr;
    c->op_sraw_out = nullptr;
    if (st) return st;
    int* misc = (int*)c->misc.p;
    const int *d_zlo = misc, *d_it = misc + n;

    GradArgs ga;
    ga.n = n;
    ga.nspec = c->nspec;
    ga.nsp = nsp;
    ga.parts = parts;
    ga.nbands = c->nbands;
    ga.jtop = c->jtop;
    ga.add_igm = c->add_igm;
    ga.band_j0 = c->band_j0;
    ga.band_j1 = c->nbands > 0 ? c->band_j1 : -1;
    ga.ncosmo = c->ncosmo;
    LamLookup lk;
    lk.nspec = c->nspec;
    lk.nbk = c->nbk;
    lk.lnl0 = c->lnl0;
    lk.bkinv = c->bkinv;
    lk.lam = c->d_lam;
    lk.bkt = c->d_bkt;
    ga.lk = lk;
    ga.prm = f.prm;
    ga.hdr = (const double*)c->hdr.p;
    ga.yo = c->yo.p;
    ga.lam = c->d_lam;
    ga.loglam = c->d_loglam;
    ga.ltab = c->d_ltab;
    ga.wl = c->d_wl;
    ga.dlam = c->d_dlam;
    ga.rdlam = c->d_rdlam;
    ga.igm_a = c->d_igm_a;
    ga.cz = c->d_cz;
    ga.cdl = c->d_cdl;
    ga.sm32 = c->d_sm32;
    ga.mask = c->d_mask;
    ga.bp_ptr = c->d_bp_ptr;
    ga.bp_band = c->d_bp_band;
    ga.bp_c64 = c->d_bp_c64;
    ga.mags = fw.mags;
    ga.fband = fw.fband;
    ga.lsum = (const double*)c->glsum.p;
    ga.mags_bar = g->mags_bar;
    ga.spec_bar = g->spec_bar;
    ga.lbol_bar = g->lbol_bar;
    ga.ybh = c->gyb.p;
    ga.ybl = mixed ? (void*)((uint16_t*)c->gyb.p + (size_t)ncols * nsp) : nullptr;
    ga.gout = (double*)c->gout.p;
    ga.zmags = f.zred_mags;
    const size_t lds = 3 * tsz * (size_t)nsp + sizeof(double) * ((c->nbands > 0 ? c->nbands : 1) + 8 * 16);
    const void* kfn = mixed ? (const void*)k_grad_chain<float> : (const void*)k_grad_chain<double>;
    HIPCHK(c, hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (c->profiling) HIPCHK(c, hipEventRecord(c->gev[0], c->stream));
    if (mixed)
        k_grad_chain<float><<<n, 512, lds, c->stream>>>(ga);
    else
        k_grad_chain<double><<<n, 512, lds, c->stream>>>(ga);
    HIPCHK(c, hipGetLastError());
    if (c->profiling) HIPCHK(c, hipEventRecord(c->gev[1], c->stream));

    GemmTArgs gt;
    gt.krows = kfull;
    gt.nrt = nrt;
    gt.ncols = ncols;
    gt.nct = (ncols + 63) / 64;
    gt.nsp = nsp;
    gt.ldc = ldc;
    gt.A64 = ops ? c->op.d_At64 : c->full.At64;
    gt.B64 = (const double*)c->gyb.p;
    gt.C = c->gC.p;
    if (mixed) {
        MMArgs mm;
        mm.mblk = nrt * 4;
        mm.nrt = (mm.mblk + 7) / 8;
        mm.ksteps = nsp / OPS_KSTEP_BF16;
        mm.mode = OPS_MM_DIRECT;
        mm.n = n;
        mm.parts = parts;
        mm.order = nullptr;
        mm.ntiles = nullptr;
        mm.tiles = nullptr;
        mm.ncols = ncols;
        mm.nct = (ncols + OPS_MM_COLS - 1) / OPS_MM_COLS;
        mm.Ah = ops ? c->op.d_Ath : c->full.Ath;
        mm.Al = ops ? c->op.d_Atl : c->full.Atl;
        mm.slab = 0;
        mm.Bh = (const uint16_t*)ga.ybh;
        mm.Bl = (const uint16_t*)ga.ybl;
        mm.ldb = nsp;
        mm.C = (float*)c->gC.p;
        mm.ldc = ldc;
        if ((st = launch_mm(c, mm, mm.nct))) return st;
    } else {
        const unsigned gblocks = 8u * (unsigned)((nrt + 7) / 8) * (unsigned)gt.nct;
        k_gemmt_f64<<<gblocks, 256, 0, c->stream>>>(gt);
        HIPCHK(c, hipGetLastError());
    }
    if (c->profiling) HIPCHK(c, hipEventRecord(c->gev[2], c->stream));

    double* wbar_out = g->weights_bar ? g->weights_bar : (sfh_rows ? (double*)c->gwb.p : nullptr);
    if (ops) {
        // the operator's reverse pass: the template cotangents by position, the IMF constants' derivatives, the
        // points' sums per (pass, age, model), then w_bar, the metallicity's and the IMF's rows and params_bar
        OpAdjArgs q;
        q.n = n;
        q.nt = nt;
        q.nz = nz;
        q.npass = npass;
        q.ntpl = c->op.ntpl;
        q.nent = c->op.nent;
        q.imf_type = c->imf_type;
        q.remnants = c->remnants;
        q.salp_ind = c->salp_ind;
        q.mdave = c->mdave;
        q.mlim_bh = c->mlim_bh;
        q.mlim_ns = c->mlim_ns;
        q.ntiles = misc + 4 * n;
        q.tiles = (const int4*)c->tiles.p;
        q.order = misc + 3 * n;
        q.it = d_it;
        q.zlo = d_zlo;
        q.prm = f.prm;
        q.hdr = (const double*)c->hdr.p;
        q.wsfh = (const double*)c->wsfh.p;
        q.gout = (const double*)c->gout.p;
        q.mass_bar = g->mass_bar;
        q.logz = c->d_logz;
        q.impc = (const double*)c->opimf.p;
        q.impd = (double*)c->opimfd.p;
        q.age_ptr = c->op.d_age_ptr;
        q.first = c->op.d_first;
        q.slow_ptr = c->op.d_slow_ptr;
        q.slow_pt = c->op.d_slow_pt;
        q.pt = c->op.d_pt;
        q.age_mmax = c->op.d_age_mmax;
        q.wp = c->op.d_wp;
        q.poly = c->op.d_poly;
        q.pent = c->op.d_pent;
        q.pval = c->op.d_pval;
        q.smass_p = (const double*)c->opsum.p;
        q.slum_p = q.smass_p + (size_t)npass * n * nt;
        q.sraw = (const double*)c->opsraw.p;
        q.CT = c->opct.p;
        q.part = (double*)c->oppart.p;
        q.gpart = q.part + 4 * (size_t)npass * nt * n;
        q.ngroup = (c->op.npa_max + OPS_OP_GPTS - 1) / OPS_OP_GPTS;
        q.params_bar = g->params_bar;
        q.weights_bar = wbar_out;
        const dim3 gt3((unsigned)((c->op.ntpl + 63) / 64), (unsigned)((n + 63) / 64), 2);
        if (mixed)
            k_op_adj_t<float><<<gt3, 256, 0, c->stream>>>(n, c->op.ntpl, ldc, q.order, (const float*)c->gC.p,
                                                           (float*)c->opct.p);
        else
            k_op_adj_t<double><<<gt3, 256, 0, c->stream>>>(n, c->op.ntpl, ldc, q.order, (const double*)c->gC.p,
                                                            (double*)c->opct.p);
        HIPCHK(c, hipGetLastError());
        k_op_imf_grad<<<(n + 255) / 256, 256, 0, c->stream>>>(q);
        HIPCHK(c, hipGetLastError());
        const int maxtiles = (n + OPS_TILE_COL - 1) / OPS_TILE_COL + c->npair;
        const dim3 gp((unsigned)(nt * ((q.ngroup + 3) / 4)), (unsigned)maxtiles, (unsigned)npass);
        const dim3 gs((unsigned)((nt + 3) / 4), (unsigned)maxtiles, (unsigned)npass);
        if (mixed) {
            k_op_adj_points<float><<<gp, 256, 0, c->stream>>>(q);
            k_op_adj_slow<float><<<gs, 256, 0, c->stream>>>(q);
        } else {
            k_op_adj_points<double><<<gp, 256, 0, c->stream>>>(q);
            k_op_adj_slow<double><<<gs, 256, 0, c->stream>>>(q);
        }
        HIPCHK(c, hipGetLastError());
        k_op_adj_out<<<(n + 3) / 4, 256, 0, c->stream>>>(q);
        HIPCHK(c, hipGetLastError());
    }

    GradOutArgs go;
    go.n = n;
    go.nt = nt;
    go.nz = nz;
    go.parts = parts;
    go.kfull = kfull;
    go.ldc = ldc;
    go.wz = wz;
    go.zlo = d_zlo;
    go.it = d_it;
    go.hdr = (const double*)c->hdr.p;
    go.logz = c->d_logz;
    go.bmass = c->d_bmass;
    go.blbol = c->d_blbol;
    go.w = fw.weights;
    go.mass_bar = g->mass_bar;
    go.gout = (const double*)c->gout.p;
    go.C = c->gC.p;
    go.mixed = mixed;
    go.params_bar = g->params_bar;
    go.weights_bar = wbar_out;
    go.weights_z_bar = g->weights_z_bar;
    if (!ops) {
        k_grad_out<<<n, 256, 0, c->stream>>>(go);
        HIPCHK(c, hipGetLastError());
    }
    if (sfh_rows) {
        SfhGradArgs sa;
        sa.n = n;
        sa.nt = nt;
        sa.tiny_logt = c->tiny_logt;
        sa.time = c->d_time;
        sa.prm = f.prm;
        sa.wbar = go.weights_bar;
        sa.params_bar = g->params_bar;
        const size_t lds_sfh = sizeof(double) * 4 * 8 * (size_t)(nt + 2);
        if (lds_sfh > 65536)
            HIPCHK(c, hipFuncSetAttribute((const void*)k_sfh_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_sfh));
        k_sfh_grad<<<(n + 3) / 4, 256, lds_sfh, c->stream>>>(sa);
        HIPCHK(c, hipGetLastError());
    }
    if (c->profiling) {
        HIPCHK(c, hipEventRecord(c->gev[3], c->stream));
        c->gev_pending = true;
    }
    return 0;
}

int opensps_hip_grad_kernel_times(opensps_hip_ctx_t* c, double* ms, int* launches, int reset) {
    if (!c) return 1;
    for (int k = 0; k < OpsCtx::NGPROF; k++) {
        if (ms) ms[k] = c->gprof_ms[k];
        if (launches) launches[k] = (int)c->gprof_n[k];
        if (reset) {
            c->gprof_ms[k] = 0.0;
            c->gprof_n[k] = 0;
        }
    }
    return 0;
}

int opensps_hip_sync(opensps_hip_ctx_t* c) {
    if (!c) return 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->ev_pending) {
        for (int k = 0; k < OpsCtx::NPROF; k++) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
        }
        c->ev_pending = false;
    }
    if (c->opev_pending) {
        for (int k = 0; k < OpsCtx::NOPROF; k++) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->opev[k], c->opev[k + 1]));
            c->opprof_ms[k] += ms;
            c->opprof_n[k] += 1;
        }
        c->opev_pending = false;
    }
    if (c->gev_pending) {
        for (int k = 0; k < OpsCtx::NGPROF; k++) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->gev[k], c->gev[k + 1]));
            c->gprof_ms[k] += ms;
            c->gprof_n[k] += 1;
        }
        c->gev_pending = false;
    }
    if (!c->d_last_flag) return 0;
    int flag = 0;
    if (c->flag_async)
        flag = *c->h_flag;   // (the stream has drained: the batch's copy has landed)
    else
        HIPCHK(c, hipMemcpy(&flag, c->d_last_flag, sizeof(int), hipMemcpyDeviceToHost));
    c->flag_async = false;
    c->d_last_flag = nullptr;
    if (flag == 2) return ops_fail(c, 2, "batch: a model needs an option this build does not compute (SFH type, "
                                         "tage <= 0 without weights, smoothing other than the fast velocity operator, "
                                         "or dust emission / the AGN torus without their templates)");
    if (flag == 4) return ops_fail(c, 3, "batch: split_parts = 0 with a birth-cloud or old-star screen");
    return 0;
}

int opensps_hip_set_profiling(opensps_hip_ctx_t* c, int on) {
    if (!c) return 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (on && !c->ev[0])
        for (int k = 0; k <= OpsCtx::NPROF; k++) HIPCHK(c, hipEventCreate(&c->ev[k]));
    if (on && !c->gev[0])
        for (int k = 0; k <= OpsCtx::NGPROF; k++) HIPCHK(c, hipEventCreate(&c->gev[k]));
    if (on && !c->opev[0])
        for (int k = 0; k <= OpsCtx::NOPROF; k++) HIPCHK(c, hipEventCreate(&c->opev[k]));
    c->opev_pending = false;
    c->profiling = on != 0;
    c->ev_pending = false;
    c->gev_pending = false;
    return 0;
}

int opensps_hip_op_kernel_times(opensps_hip_ctx_t* c, double* ms, int* launches, int reset) {
    if (!c) return 1;
    for (int k = 0; k < OpsCtx::NOPROF; k++) {
        if (ms) ms[k] = c->opprof_ms[k];
        if (launches) launches[k] = (int)c->opprof_n[k];
        if (reset) {
            c->opprof_ms[k] = 0.0;
            c->opprof_n[k] = 0;
        }
    }
    return 0;
}

int opensps_hip_kernel_times(opensps_hip_ctx_t* c, double* ms, int* launches, int reset) {
    if (!c) return 1;
    for (int k = 0; k < OpsCtx::NPROF; k++) {
        if (ms) ms[k] = c->prof_ms[k];
        if (launches) launches[k] = (int)c->prof_n[k];
        if (reset) {
            c->prof_ms[k] = 0.0;
            c->prof_n[k] = 0;
        }
    }
    return 0;
}

int opensps_hip_band_fluxes(int precision, int n, int nspec, int nbands, const void* bands_t, const int* j_lo,
                            const int* j_hi, const int* do_band, const void* wt, const void* fobs_t, double* fband,
                            void* stream) {
    if (n <= 0 || nbands <= 0) return 0;
    if (nspec < 2 || !bands_t || !j_lo || !j_hi || !do_band || !wt || !fobs_t || !fband)
        return ops_fail(nullptr, 3, "band_fluxes: bad arguments");
    dim3 grid((n + 3) / 4, nbands);
    if (precision == OPENSPS_HIP_FP64)
        k_band_fluxes<double><<<grid, 256, 0, (hipStream_t)stream>>>(n, nspec, nbands, (const double*)bands_t, j_lo, j_hi,
                                                                     do_band, (const double*)wt, (const double*)fobs_t, fband);
    else
        k_band_fluxes<float><<<grid, 256, 0, (hipStream_t)stream>>>(n, nspec, nbands, (const float*)bands_t, j_lo, j_hi,
                                                                    do_band, (const float*)wt, (const float*)fobs_t, fband);
    HIPCHK(nullptr, hipGetLastError());
    HIPCHK(nullptr, hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int opensps_hip_observed_frame(int precision, int n, int nspec, const double* zc, const double* lambda, const void* total,
                               void* fobs_t, void* stream) {
    if (n <= 0) return 0;
    if (nspec < 2 || !zc || !lambda || !total || !fobs_t) return ops_fail(nullptr, 3, "observed_frame: bad arguments");
    dim3 grid((nspec + 255) / 256, n);
    if (precision == OPENSPS_HIP_FP64)
        k_observed_frame<double><<<grid, 256, 0, (hipStream_t)stream>>>(n, nspec, zc, lambda, (const double*)total,
                                                                        (double*)fobs_t);
    else
        k_observed_frame<float><<<grid, 256, 0, (hipStream_t)stream>>>(n, nspec, zc, lambda, (const float*)total,
                                                                       (float*)fobs_t);
    HIPCHK(nullptr, hipGetLastError());
    HIPCHK(nullptr, hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int opensps_hip_selftest_mfma(int which, const void* a, const void* b, void* d) {
    if (!a || !b || !d || (which != 0 && which != 1)) return ops_fail(nullptr, 3, "selftest_mfma: bad arguments");
    size_t ab = which == 0 ? 64 * sizeof(double) : 64 * 8 * sizeof(uint16_t);
    size_t db = which == 0 ? 64 * 4 * sizeof(double) : 64 * 4 * sizeof(float);
    void *da = nullptr, *dbp = nullptr, *dd = nullptr;
    HIPCHK(nullptr, hipMalloc(&da, ab));
    HIPCHK(nullptr, hipMalloc(&dbp, ab));
    HIPCHK(nullptr, hipMalloc(&dd, db));
    HIPCHK(nullptr, hipMemcpy(da, a, ab, hipMemcpyHostToDevice));
    HIPCHK(nullptr, hipMemcpy(dbp, b, ab, hipMemcpyHostToDevice));
    k_selftest_mfma<<<1, 64>>>(which, da, dbp, dd);
    HIPCHK(nullptr, hipGetLastError());
    HIPCHK(nullptr, hipMemcpy(d, dd, db, hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(dbp);
    (void)hipFree(dd);
    return 0;
}

int opensps_hip_malloc(void** ptr, size_t bytes) {
    if (!ptr) return 3;
    *ptr = nullptr;
    HIPCHK(nullptr, hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int opensps_hip_free(void* ptr) {
    if (ptr) HIPCHK(nullptr, hipFree(ptr));
    return 0;
}

int opensps_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes) HIPCHK(nullptr, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int opensps_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes) HIPCHK(nullptr, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int opensps_hip_device_count(int* n) {
    if (!n) return 3;
    *n = 0;
    if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
    return 0;
}

int opensps_hip_last_error(opensps_hip_ctx_t* c, char* msg, int len) {
    if (!msg || len <= 0) return 3;
    std::string s;
    if (c) {
        s = c->last_error;
    } else {
        std::lock_guard<std::mutex> g(g_err_mu);
        s = g_err;
    }
    snprintf(msg, (size_t)len, "%s", s.c_str());
    return 0;
}

}   // extern "C"
