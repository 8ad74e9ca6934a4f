// The C-ABI of libbbgpu.so: argument checks, the per-fit model object (GalmodSetup + beam kernel on one
// device), error reporting and the launches of the kernels of bb_galmod.hip, bb_conv.hip, bb_resid.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/bbgpu.h"
#include "bb_internal.h"

#define fail bb_fail
#define hip_check bb_hip_check

extern "C" int bb_launch_galmod_analytic(const bb_grid* g, bb_box box, const double* sr, int nsub, int B,
                                         const int* naz, const int4* chunks, int nchunks, const double2* az,
                                         int TX, int TY, double* sub, double* wtab, short4* bbox, short4* sbox,
                                         const double* cd2i, const double* vedges, int lin_axis, float* out,
                                         int accumulate, int* bin, int bin_cap,
                                         const int4* vbox, int variant, int boxes_ready, hipStream_t st);
extern "C" void bb_set_tile_events(void* ev_begin, void* ev_end);
extern "C" size_t bb_analytic_sub_bytes();
extern "C" size_t bb_analytic_wt_bytes();
extern "C" int bb_analytic_chunk();
extern "C" size_t bb_analytic_lds_bytes(int nz, int nlines, int TX, int TY);
extern "C" size_t bb_analytic_lds_static();
extern "C" int bb_launch_conv2d_same(const float* in, float* out, int B, int by, int bx, int nz, const float* kern,
                                     int ky, int kx, float scale, hipStream_t st);
extern "C" int bb_conv_max_kx();
extern "C" int bb_residual_blocks(int64_t nvox);
extern "C" int bb_launch_ring_maps(const bb_grid* g, bb_box box, const double* sr, int nsub, int B, const int* naz,
                                   const int4* chunks, int nchunks, const double2* az, int TX, int TY, double* sub,
                                   double* wtab, short4* bbox, short4* sbox, const int* widx, const double* wval,
                                   const double* ew, int nmap, int map0, int nw, double qunit,
                                   unsigned long long* counts, int* bin, int bin_cap, hipStream_t st);
extern "C" size_t bb_rm_tile_lds(int nw, int TX, int TY);
extern "C" int bb_launch_zero_planes(unsigned long long* counts, int B, int nmap, int map0, int nw, int64_t plane,
                                     hipStream_t st);
extern "C" int bb_launch_norm(int B, int by, int bx, int nz, int ny, int nx, int64_t blo0, int64_t blo1, double wpow,
                              int side, int ftype, double data_noise, double bweight, int normtype, int foot,
                              double hmax, const float* model, const float* data, const uint8_t* mask,
                              const double* ring, const double* chan_noise, const float* pre, const int* ilo,
                              const int* ihi, int ky, const double* thr, double* work, double* val, int64_t* nb,
                              int64_t* nt, hipStream_t st);
extern "C" size_t bb_mc_sub_bytes();
extern "C" int bb_launch_zero_u32(void* p, int64_t n, hipStream_t st);
extern "C" size_t bb_mc_rec_bytes();
extern "C" size_t bb_mc_lds_bytes(int nz, int TX, int TY);
extern "C" int bb_launch_galmod_mc(const bb_grid* g, bb_box box, const double* sr, int nsub, int B, int nvpar,
                                   uint32_t seed, int TX, int TY, int64_t cap, void* subs, double* qv, float* invq,
                                   int64_t* total, int* tcount, int64_t* toff, int* tcursor, void* recs, int* overflow,
                                   const double* cd2i, float* out, unsigned long long* counts, int accumulate,
                                   const int4* vbox, hipStream_t st);
extern "C" size_t bb_rm_tile_static();
extern "C" size_t bb_rm_finish_lds(int ky, int kx);
extern "C" int bb_launch_rm_finish(const unsigned long long* counts, double* maps, int P, int by, int bx,
                                   const double* kern, int ky, int kx, double unit, double scale, int convolve,
                                   hipStream_t st);
extern "C" int bb_gram_blocks(int npix);
extern "C" size_t bb_gram_entries(int n);
extern "C" size_t bb_nnls_lds(int n);
extern "C" int bb_nnls_max();
extern "C" int bb_launch_profile_solve(int n, int nr, int nok, double smooth, double corr, double floor_, double norm,
                                       double dmed, const double* maps, const double* wpix, const double* ypix,
                                       int npix, const double* lcoef, const int* lcol, const int* near,
                                       const double* normv, int B, double* partial, double* gram, double* Mall,
                                       double* x, hipStream_t st);
extern "C" int bb_launch_set_dens(const double* x, int B, int n, int nr, const int* widx, const double* wval,
                                  const double* ew, int nsub, double* sr, float* dens, hipStream_t st);
extern "C" int bb_launch_dilate(const uint8_t* mask, int nz, int ny, int nx, double radius, int nchan, uint8_t* tmp,
                                uint8_t* out, hipStream_t st);
extern "C" int bb_launch_residual(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2,
                                  double eps2, double* partial, double* out_ss, double* out_lg, const int4* vbox,
                                  const float2* sv, int frac, int bx, int nz, hipStream_t st);

static thread_local char g_err[512] = "";

int bb_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int bb_hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) return bb_fail(BB_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

extern "C" const char* bb_last_error(void) { return g_err; }

extern "C" int bb_device_count(int* n) {
    if (!n) return fail(BB_ERR_ARG, "bb_device_count: n is NULL");
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        *n = 0;
        return fail(BB_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return 0;
}

static int check_grid(const bb_grid* g) {
    if (!g) return fail(BB_ERR_ARG, "grid is NULL");
    if (g->nz <= 0) return fail(BB_ERR_ARG, "grid: nz = %d", g->nz);
    if (g->nlines < 1 || g->nlines > BB_MAX_LINES) return fail(BB_ERR_ARG, "grid: nlines = %d", g->nlines);
    if (g->nzq < 1 || g->nzq > 1024) return fail(BB_ERR_ARG, "grid: nzq = %d (1..1024)", g->nzq);
    if (!(g->pixarea > 0) || !(g->radsep > 0) || g->cdelt0 == 0 || g->cdelt1 == 0)
        return fail(BB_ERR_ARG, "grid: pixel sizes must be non-zero");
    if (!(g->chmin > 0)) return fail(BB_ERR_ARG, "grid: chmin must be > 0");
    return 0;
}


extern "C" int bb_model_destroy(bb_model* m) {
    if (!m) return 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(m->device);
    (void)hipFree(m->d_cd2i);
    (void)hipFree(m->d_vedges);
    (void)hipFree(m->d_kern);
    (void)hipFree(m->d_kern64);
    (void)hipSetDevice(dev);
    delete m;
    return 0;
}

extern "C" int bb_model_create(const bb_grid* g, const double* cd2i, const double* vedges, const double* kern, int ky,
                               int kx, double scale, bb_model** out) {
    if (!out) return fail(BB_ERR_ARG, "bb_model_create: out is NULL");
    *out = nullptr;
    int rc = check_grid(g);
    if (rc) return rc;
    if (!cd2i || !vedges) return fail(BB_ERR_ARG, "bb_model_create: missing setup arrays");
    if (kern && (ky < 1 || kx < 1 || ky % 2 == 0 || kx % 2 == 0 || ky > 511 || kx > bb_conv_max_kx()))
        return fail(BB_ERR_ARG, "bb_model_create: kernel %dx%d must be odd-sized with kx <= %d", ky, kx,
                    bb_conv_max_kx());
    bb_model* m = new bb_model();
    memset(m, 0, sizeof *m);
    m->g = *g;
    m->scale = scale;
    m->bin_per = BB_BIN_PER_CHUNK;
    (void)hipGetDevice(&m->device);
    rc = upload(&m->d_cd2i, cd2i, (size_t)g->nz, "upload cd2i");
    if (!rc) rc = upload(&m->d_vedges, vedges, (size_t)g->nlines * (g->nz + 1), "upload vedges");
    if (!rc && g->axtyp == 4) {
        bool lin = true;
        for (int k = 1; k < g->nz && lin; k++) lin = cd2i[k] == cd2i[0] && (float)cd2i[k] == (float)cd2i[0];
        for (int l = 0; l < g->nlines && lin; l++)
            for (int e = 0; e <= g->nz && lin; e++) {
                volatile double t = (double)e + 0.5 - g->crpix3;       // (the operations of the device, unfused)
                volatile double u = t * g->cdelt3;
                lin = u + g->crval3 == vedges[(size_t)l * (g->nz + 1) + e];
            }
        m->lin = lin ? 1 : 0;
    }
    if (!rc && kern) {
        std::vector<float> kf((size_t)ky * kx);
        for (size_t i = 0; i < kf.size(); i++) kf[i] = (float)kern[i];
        m->ky = ky;
        m->kx = kx;
        rc = upload(&m->d_kern, kf.data(), kf.size(), "upload kernel");
        if (!rc) rc = upload(&m->d_kern64, kern, kf.size(), "upload kernel");
    }
    if (!rc) {
        // the pixel tile: the most pixels (at most 256, sides at most 3:2) whose spectra, with the channel tables,
        // fit the LDS of a CU, the squarest of those, taller on a tie (fewer sample columns cross a square tile per
        // pixel: 12x13 instead of 16x8 for 96 channels, 53.3 against 57.3 ms for the 256-vector batch of the
        // 160x160x96 cube)
        m->TX = 0;
        int best = 0, bd = 1 << 30;
        for (int ty = 1; ty <= 256; ty++)
            for (int tx = (2 * ty + 2) / 3; 2 * tx <= 3 * ty && tx * ty <= 256; tx++) {
                if (bb_analytic_lds_bytes(g->nz, g->nlines, tx, ty) + bb_analytic_lds_static() > BB_LDS_MAX) break;
                const int a = tx * ty, d = tx > ty ? tx - ty : ty - tx;
                if (a > best || (a == best && (d < bd || (d == bd && ty > tx)))) {
                    best = a;
                    bd = d;
                    m->TX = tx;
                    m->TY = ty;
                }
            }
        if (!m->TX) rc = fail(BB_ERR_ARG, "bb_model_create: nz = %d channels do not fit the LDS tile", g->nz);
    }
    if (rc) {
        bb_model_destroy(m);
        return rc;
    }
    *out = m;
    return 0;
}

extern "C" int bb_model_set_tile(bb_model* m, int TX, int TY) {
    if (!m || TX < 1 || TY < 1 || TX > 64 || TY > 64) return fail(BB_ERR_ARG, "bb_model_set_tile: bad tile %dx%d", TX, TY);
    if (bb_analytic_lds_bytes(m->g.nz, m->g.nlines, TX, TY) + bb_analytic_lds_static() > BB_LDS_MAX)
        return fail(BB_ERR_ARG, "bb_model_set_tile: a %dx%d tile of %d channels exceeds the LDS", TX, TY, m->g.nz);
    m->TX = TX;
    m->TY = TY;
    return 0;
}

extern "C" int bb_model_set_bin(bb_model* m, int per_chunk) {
    if (!m || per_chunk < 0 || per_chunk > BB_BIN_PER_CHUNK)
        return fail(BB_ERR_ARG, "bb_model_set_bin: %d entries per chunk (0..%d)", per_chunk, BB_BIN_PER_CHUNK);
    m->bin_per = per_chunk;
    return 0;
}

extern "C" int bb_model_set_deposit(bb_model* m, int variant) {
    if (!m || variant < 0 || variant > 1) return fail(BB_ERR_ARG, "bb_model_set_deposit: variant %d (0 or 1)", variant);
    m->dep_variant = variant;
    return 0;
}

extern "C" int bb_model_get_tile(const bb_model* m, int* TX, int* TY) {
    if (!m || !TX || !TY) return fail(BB_ERR_ARG, "bb_model_get_tile: NULL argument");
    *TX = m->TX;
    *TY = m->TY;
    return 0;
}

extern "C" int bb_azplan_destroy(bb_azplan* p) {
    if (!p) return 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(p->device);
    (void)hipFree(p->d_naz);
    (void)hipFree(p->d_chunks);
    (void)hipFree(p->d_az);
    (void)hipSetDevice(dev);
    delete p;
    return 0;
}

extern "C" int bb_azplan_create(const bb_model* m, const double* R0, int nsub, bb_azplan** out) {
    if (!out) return fail(BB_ERR_ARG, "bb_azplan_create: out is NULL");
    *out = nullptr;
    if (!m || (nsub > 0 && !R0) || nsub < 0) return fail(BB_ERR_ARG, "bb_azplan_create: bad arguments");
    const int CS = bb_analytic_chunk();
    std::vector<int> naz(nsub);
    std::vector<int4> chunks;
    std::vector<double2> az;
    for (int s = 0; s < nsub; s++) {
        if (!(R0[s] >= 0) || !(R0[s] < 1e6)) return fail(BB_ERR_ARG, "bb_azplan_create: radius %d = %g", s, R0[s]);
        double circ_pix = 2 * M_PI * R0[s] / sqrt(m->g.pixarea);
        long long n = (long long)(3 * circ_pix + 0.5);
        if (n < 16) n = 16;
        if (n > (1 << 24)) return fail(BB_ERR_ARG, "bb_azplan_create: %lld azimuths in sub-ring %d", n, s);
        naz[s] = (int)n;
        const double daz = 2 * M_PI / (double)n;
        const double cd = cos(daz), sd = sin(daz);
        for (int kr = 0; kr < 2; kr++) {
            const int off = (int)az.size();
            const double az0 = 2 * M_PI * (0.5 * ((kr + 1) % 2) + 0.25) / (double)n;
            double caz = cos(az0), saz = sin(az0);
            for (long long ka = 0; ka < n; ka++) {
                if (ka > 0) {
                    double c2 = caz * cd - saz * sd, s2 = saz * cd + caz * sd;
                    caz = c2;
                    saz = s2;
                }
                az.push_back(make_double2(caz, saz));
            }
            for (int ka0 = 0; ka0 < n; ka0 += CS)
                chunks.push_back(make_int4(2 * s + kr, ka0, (int)((n - ka0) < CS ? (n - ka0) : CS), off + ka0));
        }
    }
    if (az.size() > (size_t)INT32_MAX) return fail(BB_ERR_ARG, "bb_azplan_create: too many azimuth samples");
    bb_azplan* p = new bb_azplan();
    p->nsub = nsub;
    p->nchunks = (int)chunks.size();
    p->R0.assign(R0, R0 + nsub);
    p->d_naz = nullptr;
    p->d_chunks = nullptr;
    p->d_az = nullptr;
    (void)hipGetDevice(&p->device);
    int rc = upload(&p->d_naz, naz.data(), naz.size(), "upload naz");
    if (!rc) rc = upload(&p->d_chunks, chunks.data(), chunks.size(), "upload chunks");
    if (!rc) rc = upload(&p->d_az, az.data(), az.size(), "upload azimuths");
    if (rc) {
        bb_azplan_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

// The chunk lists of the pixel tiles (bin_chunks): up to BB_BIN_PER_CHUNK tiles per chunk on average (the work
// buffers are sized for that; a model may use fewer, bb_model_set_bin), and at most
// nchunks tiles (beyond: the tiles test every chunk box).
static int64_t bin_cap(const bb_azplan* p) { return (int64_t)p->nchunks * BB_BIN_PER_CHUNK; }
static size_t bin_ints(const bb_azplan* p) { return (size_t)3 * (p->nchunks + 1) + 1 + (size_t)bin_cap(p); }

extern "C" int64_t bb_analytic_work_size(const bb_azplan* p, int B) {
    if (!p || B < 0) return -1;
    size_t per = (size_t)p->nsub * (bb_analytic_sub_bytes() + bb_analytic_wt_bytes()) +
                 (size_t)p->nchunks * (1 + bb_analytic_chunk()) * sizeof(short4) + bin_ints(p) * sizeof(int);
    return (int64_t)(per * B + 256);
}

// the bin region of the work buffer after the sample boxes, or NULL when the box has more tiles than chunks
static int* work_bin(const bb_model* m, const bb_azplan* p, int B, short4* sbox, long long ntiles) {
    if (m->bin_per <= 0 || ntiles > p->nchunks || bin_cap(p) > INT32_MAX) return nullptr;
    return (int*)(sbox + (size_t)B * p->nchunks * bb_analytic_chunk());
}

static int check_box(const bb_model* m, bb_box box) {
    if (box.bhi0 <= box.blo0 || box.bhi1 <= box.blo1)
        return fail(BB_ERR_ARG, "box [%d,%d)x[%d,%d) is empty", box.blo0, box.bhi0, box.blo1, box.bhi1);
    (void)m;
    return 0;
}

extern "C" int bb_galmod_analytic(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub,
                                  int B, void* work, int64_t work_bytes, float* out, int accumulate, void* stream) {
    return bb_galmod_analytic_vb(m, p, box, nullptr, sr, nsub, B, work, work_bytes, out, accumulate, stream);
}

extern "C" int bb_galmod_analytic_boxes(const bb_model* m, const bb_azplan* p, bb_box box, const int32_t* vbox,
                                        const double* sr, int nsub, int B, void* work, int64_t work_bytes, float* out,
                                        int accumulate, void* stream) {
    if (!vbox) return fail(BB_ERR_ARG, "bb_galmod_analytic_boxes: vbox is NULL");
    return bb_galmod_analytic_vb(m, p, box, (const int4*)vbox, sr, nsub, B, work, work_bytes, out, accumulate, stream);
}

int bb_galmod_analytic_vb(const bb_model* m, const bb_azplan* p, bb_box box, const int4* vbox, const double* sr,
                          int nsub, int B, void* work, int64_t work_bytes, float* out, int accumulate, void* stream,
                          int boxes_ready) {
    if (!m || !p) return fail(BB_ERR_ARG, "bb_galmod_analytic: model or plan is NULL");
    int rc = check_box(m, box);
    if (rc) return rc;
    if (nsub != p->nsub) return fail(BB_ERR_ARG, "bb_galmod_analytic: nsub %d != plan's %d", nsub, p->nsub);
    if (B < 0 || B > 65535) return fail(BB_ERR_ARG, "bb_galmod_analytic: B=%d out of range", B);
    if (box.bhi0 - box.blo0 > 32767 || box.bhi1 - box.blo1 > 32767)
        return fail(BB_ERR_ARG, "bb_galmod_analytic: box larger than 32767 pixels");
    if (nsub == 0 || B == 0) return 0;
    if (!sr || !out || !work) return fail(BB_ERR_ARG, "bb_galmod_analytic: NULL array");
    if (work_bytes < bb_analytic_work_size(p, B))
        return fail(BB_ERR_ARG, "bb_galmod_analytic: work holds %lld bytes, needs %lld", (long long)work_bytes,
                    (long long)bb_analytic_work_size(p, B));
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    const long long ntiles = (long long)((bx + m->TX - 1) / m->TX) * ((by + m->TY - 1) / m->TY);
    if (ntiles > INT32_MAX) return fail(BB_ERR_ARG, "bb_galmod_analytic: too many tiles");
    char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    double* sub = (double*)w;
    w += (size_t)B * nsub * bb_analytic_sub_bytes();
    double* wtab = (double*)w;
    w += (size_t)B * nsub * bb_analytic_wt_bytes();
    short4* bbox = (short4*)w;
    w += (size_t)B * p->nchunks * sizeof(short4);
    short4* sbox = (short4*)w;
    rc = bb_launch_galmod_analytic(&m->g, box, sr, nsub, B, p->d_naz, p->d_chunks, p->nchunks, p->d_az, m->TX, m->TY,
                                   sub, wtab, bbox, sbox, m->d_cd2i, m->d_vedges, m->lin, out,
                                   accumulate, work_bin(m, p, B, sbox, ntiles), (int)(p->nchunks * (int64_t)m->bin_per), vbox,
                                   m->dep_variant, boxes_ready && !vbox, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_galmod_analytic launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_trace_deposit(void* ev_begin, void* ev_end) {
    if ((ev_begin == nullptr) != (ev_end == nullptr))
        return fail(BB_ERR_ARG, "bb_trace_deposit: give both events or neither");
    bb_set_tile_events(ev_begin, ev_end);
    return 0;
}

extern "C" int bb_conv2d_same(const bb_model* m, const float* in, float* out, int B, int by, int bx,
                              void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_conv2d_same: model is NULL");
    if (!m->d_kern) return fail(BB_ERR_ARG, "bb_conv2d_same: the model has no beam kernel");
    if (B < 0 || by < 1 || bx < 1 || B > 65535) return fail(BB_ERR_ARG, "bb_conv2d_same: bad sizes");
    if (B == 0) return 0;
    if (!in || !out || in == out) return fail(BB_ERR_ARG, "bb_conv2d_same: in/out must be distinct, non-NULL");
    int rc = bb_launch_conv2d_same(in, out, B, by, bx, m->g.nz, m->d_kern, m->ky, m->kx, (float)m->scale,
                                   (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_conv2d_same launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int64_t bb_residual_work_size(int B, int64_t nvox) {
    if (B < 0 || nvox < 0) return -1;
    return (int64_t)B * bb_residual_blocks(nvox) * 2;
}

extern "C" int bb_residual_gauss(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2,
                                 double eps2, double* work, int64_t work_len, double* out_ss, double* out_lg,
                                 void* stream) {
    if (B < 0 || nvox < 0 || B > 65535) return fail(BB_ERR_ARG, "bb_residual_gauss: bad sizes");
    if (B == 0) return 0;
    if (!model || !dvalid || !work || !out_ss) return fail(BB_ERR_ARG, "bb_residual_gauss: NULL array");
    if (eps2 > 0 && (!out_lg || !(sigma2 > 0))) return fail(BB_ERR_ARG, "bb_residual_gauss: fracerr needs sigma2 > 0 and out_lg");
    if (work_len < bb_residual_work_size(B, nvox))
        return fail(BB_ERR_ARG, "bb_residual_gauss: work holds %lld doubles, needs %lld", (long long)work_len,
                    (long long)bb_residual_work_size(B, nvox));
    int rc = bb_launch_residual(model, dvalid, B, nvox, sigma2, eps2, work, out_ss, out_lg, nullptr, nullptr, 0, 1, 1,
                                (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_residual_gauss launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int64_t bb_conv_resid_blocks(int by, int bx, int nz);
extern "C" int bb_launch_conv_resid(const float* in, const float* dvalid, int B, int by, int bx, int nz,
                                    const float* kern, int ky, int kx, float scale, double sigma2, double eps2,
                                    double* partial, const int4* vbox, const float2* sv, int frac, hipStream_t st);
extern "C" int bb_launch_residual_final(const double* partial, int B, int nblk, double* out_ss, double* out_lg,
                                        hipStream_t st);

extern "C" int64_t bb_conv_residual_work_size(const bb_model* m, int B, int by, int bx) {
    if (!m || B < 0 || by < 1 || bx < 1) return -1;
    return (int64_t)B * bb_conv_resid_blocks(by, bx, m->g.nz) * 2;
}

extern "C" int bb_conv_residual_gauss(const bb_model* m, const float* dep, const float* dvalid, int B, int by, int bx,
                                      double sigma2, double eps2, double* work, int64_t work_len, double* out_ss,
                                      double* out_lg, void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_conv_residual_gauss: model is NULL");
    if (!m->d_kern) return fail(BB_ERR_ARG, "bb_conv_residual_gauss: the model has no beam kernel");
    if (B < 0 || by < 1 || bx < 1 || B > 65535) return fail(BB_ERR_ARG, "bb_conv_residual_gauss: bad sizes");
    if (B == 0) return 0;
    if (!dep || !dvalid || !work || !out_ss) return fail(BB_ERR_ARG, "bb_conv_residual_gauss: NULL array");
    if (eps2 > 0 && (!out_lg || !(sigma2 > 0)))
        return fail(BB_ERR_ARG, "bb_conv_residual_gauss: fracerr needs sigma2 > 0 and out_lg");
    const int64_t need = bb_conv_residual_work_size(m, B, by, bx);
    if (work_len < need)
        return fail(BB_ERR_ARG, "bb_conv_residual_gauss: work holds %lld doubles, needs %lld", (long long)work_len,
                    (long long)need);
    hipStream_t st = (hipStream_t)stream;
    int rc = bb_launch_conv_resid(dep, dvalid, B, by, bx, m->g.nz, m->d_kern, m->ky, m->kx, (float)m->scale, sigma2,
                                  eps2, work, nullptr, nullptr, 0, st);
    if (!rc) rc = bb_launch_residual_final(work, B, (int)(need / (2 * B)), out_ss, out_lg, st);
    return rc ? fail(BB_ERR_HIP, "bb_conv_residual_gauss launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// ---------------------------------------------------------------------------------------------------------
// Ring maps and the deconvolved density profile.
extern "C" int64_t bb_ring_maps_work_size(const bb_azplan* p, int B) { return bb_analytic_work_size(p, B); }

extern "C" int bb_ring_maps_deposit(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub,
                                    int B, const int32_t* widx, const double* wval, int nmap, int map0, int nw,
                                    double qunit, void* work, int64_t work_bytes, uint64_t* counts, void* stream) {
    return bb_ring_maps_deposit_ew(m, p, box, sr, nsub, B, widx, wval, nullptr, nmap, map0, nw, qunit, work, work_bytes,
                                   counts, stream);
}

int bb_ring_maps_deposit_ew(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub, int B,
                            const int32_t* widx, const double* wval, const double* ew, int nmap, int map0, int nw,
                            double qunit, void* work, int64_t work_bytes, uint64_t* counts, void* stream) {
    if (!m || !p) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: model or plan is NULL");
    int rc = check_box(m, box);
    if (rc) return rc;
    if (nsub != p->nsub) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: nsub %d != plan's %d", nsub, p->nsub);
    if (B < 0 || B > 65535) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: B=%d out of range", B);
    if (box.bhi0 - box.blo0 > 32767 || box.bhi1 - box.blo1 > 32767)
        return fail(BB_ERR_ARG, "bb_ring_maps_deposit: box larger than 32767 pixels");
    if (nw < 1 || map0 < 0 || map0 + nw > nmap)
        return fail(BB_ERR_ARG, "bb_ring_maps_deposit: maps %d..%d outside nmap %d", map0, map0 + nw - 1, nmap);
    if (!(qunit > 0) || !isfinite(qunit)) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: qunit must be > 0");
    if (B == 0) return 0;
    if (!counts) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: NULL array");
    if (nsub > 0 && (!sr || !widx || !wval || !work)) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: NULL array");
    if (nsub > 0 && work_bytes < bb_ring_maps_work_size(p, B))
        return fail(BB_ERR_ARG, "bb_ring_maps_deposit: work holds %lld bytes, needs %lld", (long long)work_bytes,
                    (long long)bb_ring_maps_work_size(p, B));
    // the largest pixel tile whose nw maps fit the LDS of a CU (a larger tile is crossed by fewer vertical columns)
    static const int cand[][2] = {{20, 20}, {16, 16}, {16, 8}, {8, 8}, {8, 4}, {4, 4}, {4, 2}, {2, 2}, {2, 1}, {1, 1}};
    int TX = 0, TY = 0;
    for (auto& c : cand)
        if (bb_rm_tile_lds(nw, c[0], c[1]) + bb_rm_tile_static() <= BB_LDS_MAX) {
            TX = c[0];
            TY = c[1];
            break;
        }
    if (!TX) return fail(BB_ERR_ARG, "bb_ring_maps_deposit: %d maps do not fit the LDS tile", nw);
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    if (nsub == 0) {
        rc = bb_launch_zero_planes((unsigned long long*)counts, B, nmap, map0, nw, (int64_t)by * bx, (hipStream_t)stream);
        return rc ? fail(BB_ERR_HIP, "bb_ring_maps_deposit: %s", hipGetErrorString((hipError_t)rc)) : 0;
    }
    char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    double* sub = (double*)w;
    w += (size_t)B * nsub * bb_analytic_sub_bytes();
    double* wtab = (double*)w;
    w += (size_t)B * nsub * bb_analytic_wt_bytes();
    short4* bbox = (short4*)w;
    w += (size_t)B * p->nchunks * sizeof(short4);
    short4* sbox = (short4*)w;
    const long long ntiles = (long long)((bx + TX - 1) / TX) * ((by + TY - 1) / TY);
    rc = bb_launch_ring_maps(&m->g, box, sr, nsub, B, p->d_naz, p->d_chunks, p->nchunks, p->d_az, TX, TY, sub, wtab,
                             bbox, sbox, (const int*)widx, wval, ew, nmap, map0, nw, qunit, (unsigned long long*)counts,
                             work_bin(m, p, B, sbox, ntiles), (int)(p->nchunks * (int64_t)m->bin_per), (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_ring_maps_deposit launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_ring_maps_finish(const bb_model* m, const uint64_t* counts, int P, int by, int bx, double unit,
                                   int convolve, double* maps, void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_ring_maps_finish: model is NULL");
    if (P < 0 || P > 65535 || by < 1 || bx < 1) return fail(BB_ERR_ARG, "bb_ring_maps_finish: bad sizes");
    if (convolve && !m->d_kern64) return fail(BB_ERR_ARG, "bb_ring_maps_finish: the model has no beam kernel");
    if (convolve && bb_rm_finish_lds(m->ky, m->kx) > BB_LDS_MAX)
        return fail(BB_ERR_ARG, "bb_ring_maps_finish: a %dx%d kernel is too large", m->ky, m->kx);
    if (P == 0) return 0;
    if (!counts || !maps) return fail(BB_ERR_ARG, "bb_ring_maps_finish: NULL array");
    int rc = bb_launch_rm_finish((const unsigned long long*)counts, maps, P, by, bx, m->d_kern64, m->ky, m->kx, unit,
                                 m->scale, convolve, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_ring_maps_finish launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_profile_destroy(bb_profile* p) {
    if (!p) return 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(p->device);
    (void)hipFree(p->d_wpix);
    (void)hipFree(p->d_ypix);
    (void)hipFree(p->d_lcoef);
    (void)hipFree(p->d_lcol);
    (void)hipFree(p->d_near);
    (void)hipSetDevice(dev);
    delete p;
    return 0;
}

extern "C" int bb_profile_create(const bb_profile_desc* d, const double* wpix, const double* ypix, const double* lcoef,
                                 const int32_t* lcol, const int32_t* near, bb_profile** out) {
    if (!out) return fail(BB_ERR_ARG, "bb_profile_create: out is NULL");
    *out = nullptr;
    if (!d || !wpix || !ypix) return fail(BB_ERR_ARG, "bb_profile_create: NULL argument");
    const int n = d->nr + 1;
    if (d->nr < 1 || n > bb_nnls_max()) return fail(BB_ERR_ARG, "bb_profile_create: nr = %d (1..%d)", d->nr, bb_nnls_max() - 1);
    if (d->npix < 1 || d->nok < 1 || d->nok > d->npix) return fail(BB_ERR_ARG, "bb_profile_create: bad pixel counts");
    const bool prior = d->smooth > 0 && n >= 3;
    if (prior && (!lcoef || !lcol || !near)) return fail(BB_ERR_ARG, "bb_profile_create: the prior needs its rows");
    if (prior && (!(d->norm > 0) || !(d->dmed > 0) || !(d->floor >= 0)))
        return fail(BB_ERR_ARG, "bb_profile_create: norm and dmed must be > 0, floor >= 0");
    if (prior)
        for (int i = 0; i < 3 * (n - 2); i++)
            if (lcol[i] < 0 || lcol[i] >= n) return fail(BB_ERR_ARG, "bb_profile_create: column %d out of range", lcol[i]);
    bb_profile* p = new bb_profile();
    memset(p, 0, sizeof *p);
    p->d = *d;
    (void)hipGetDevice(&p->device);
    int rc = upload(&p->d_wpix, wpix, (size_t)d->npix, "upload pixel weights");
    if (!rc) rc = upload(&p->d_ypix, ypix, (size_t)d->npix, "upload map");
    if (!rc && prior) rc = upload(&p->d_lcoef, lcoef, (size_t)3 * (n - 2), "upload curvature rows");
    if (!rc && prior) rc = upload(&p->d_lcol, (const int*)lcol, (size_t)3 * (n - 2), "upload curvature rows");
    if (!rc && prior) rc = upload(&p->d_near, (const int*)near, (size_t)(n - 2), "upload curvature rows");
    if (rc) {
        bb_profile_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

extern "C" int64_t bb_profile_work_size(const bb_profile* p, int B) {
    if (!p || B < 0) return -1;
    const int n = p->d.nr + 1, m = n + 1;
    const size_t dbl = (size_t)B * ((size_t)bb_gram_blocks(p->d.npix) * bb_gram_entries(n) + (size_t)m * m + (size_t)n * n);
    return (int64_t)(dbl * sizeof(double) + 256);
}

extern "C" int bb_profile_solve(const bb_profile* p, const double* maps, int B, void* work, int64_t work_bytes,
                                double* x, void* stream) {
    return bb_profile_solve_pv(p, maps, B, work, work_bytes, x, nullptr, nullptr, nullptr, nullptr, stream);
}

int bb_profile_solve_pv(const bb_profile* p, const double* maps, int B, void* work, int64_t work_bytes, double* x,
                        const double* lcoef, const int32_t* lcol, const int32_t* near, const double* normv,
                        void* stream) {
    if (!p) return fail(BB_ERR_ARG, "bb_profile_solve: profile is NULL");
    if (B < 0 || B > 65535) return fail(BB_ERR_ARG, "bb_profile_solve: B=%d out of range", B);
    if (B == 0) return 0;
    if (!maps || !work || !x) return fail(BB_ERR_ARG, "bb_profile_solve: NULL array");
    if (work_bytes < bb_profile_work_size(p, B))
        return fail(BB_ERR_ARG, "bb_profile_solve: work holds %lld bytes, needs %lld", (long long)work_bytes,
                    (long long)bb_profile_work_size(p, B));
    const int n = p->d.nr + 1, m = n + 1;
    double* partial = (double*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    double* gram = partial + (size_t)B * bb_gram_blocks(p->d.npix) * bb_gram_entries(n);
    double* Mall = gram + (size_t)B * m * m;
    const bool prior = p->d.smooth > 0 && n >= 3;
    // (the prior's rows of every vector when normv is given, else the profile's)
    const bool pv = normv != nullptr;
    if (pv && prior && (!lcoef || !lcol || !near)) return fail(BB_ERR_ARG, "bb_profile_solve: NULL prior rows");
    int rc = bb_launch_profile_solve(n, p->d.nr, p->d.nok, prior ? p->d.smooth : 0.0, p->d.corr, p->d.floor, p->d.norm,
                                     p->d.dmed, maps, p->d_wpix, p->d_ypix, p->d.npix, pv ? lcoef : p->d_lcoef,
                                     pv ? (const int*)lcol : p->d_lcol, pv ? (const int*)near : p->d_near,
                                     pv ? normv : nullptr, B, partial, gram, Mall, x, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_profile_solve launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_profile_set_dens(const double* x, int B, int n, int nr, const int32_t* widx, const double* wval,
                                   int nsub, double* sr, float* dens, void* stream) {
    return bb_profile_set_dens_ew(x, B, n, nr, widx, wval, nullptr, nsub, sr, dens, stream);
}

int bb_profile_set_dens_ew(const double* x, int B, int n, int nr, const int32_t* widx, const double* wval,
                           const double* ew, int nsub, double* sr, float* dens, void* stream) {
    if (B < 0 || B > 65535 || nr < 1 || n < nr || nsub < 0) return fail(BB_ERR_ARG, "bb_profile_set_dens: bad sizes");
    if (B == 0 || nsub == 0) return 0;
    if (!x || !widx || !wval || !sr) return fail(BB_ERR_ARG, "bb_profile_set_dens: NULL array");
    int rc = bb_launch_set_dens(x, B, n, nr, (const int*)widx, wval, ew, nsub, sr, dens, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_profile_set_dens launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// ---------------------------------------------------------------------------------------------------------
// Monte-Carlo GALMOD.
static size_t mc_align(size_t x) { return (x + 255) & ~(size_t)255; }

static void mc_tiles(const bb_model* m, bb_box box, int* ntiles) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    *ntiles = ((bx + m->TX - 1) / m->TX) * ((by + m->TY - 1) / m->TY);
}

extern "C" int64_t bb_mc_work_size(const bb_model* m, bb_box box, int nsub, int B, int64_t max_clouds) {
    if (!m || nsub < 0 || B < 0 || max_clouds < 0 || box.bhi0 <= box.blo0 || box.bhi1 <= box.blo1) return -1;
    int nt;
    mc_tiles(m, box, &nt);
    size_t w = mc_align((size_t)B * nsub * bb_mc_sub_bytes()) + mc_align((size_t)B * (sizeof(double) + sizeof(float) +
                                                                                      sizeof(int64_t))) +
               mc_align((size_t)B * nt * sizeof(int)) * 2 + mc_align((size_t)B * (nt + 1) * sizeof(int64_t)) +
               mc_align(sizeof(int)) + mc_align((size_t)B * max_clouds * bb_mc_rec_bytes()) + 256;
    return (int64_t)w;
}

extern "C" int bb_galmod_mc(const bb_model* m, bb_box box, const double* sr, int nsub, int B, int nvpar, uint32_t seed,
                            int64_t max_clouds, void* work, int64_t work_bytes, float* out, uint64_t* counts,
                            int* overflow, int accumulate, void* stream) {
    return bb_galmod_mc_vb(m, box, nullptr, sr, nsub, B, nvpar, seed, max_clouds, work, work_bytes, out, counts,
                           overflow, accumulate, stream);
}

extern "C" int bb_galmod_mc_boxes(const bb_model* m, bb_box box, const int32_t* vbox, const double* sr, int nsub,
                                  int B, int nvpar, uint32_t seed, int64_t max_clouds, void* work, int64_t work_bytes,
                                  float* out, uint64_t* counts, int* overflow, int accumulate, void* stream) {
    if (!vbox) return fail(BB_ERR_ARG, "bb_galmod_mc_boxes: vbox is NULL");
    return bb_galmod_mc_vb(m, box, (const int4*)vbox, sr, nsub, B, nvpar, seed, max_clouds, work, work_bytes, out,
                           counts, overflow, accumulate, stream);
}

int bb_galmod_mc_vb(const bb_model* m, bb_box box, const int4* vbox, const double* sr, int nsub, int B, int nvpar,
                    uint32_t seed, int64_t max_clouds, void* work, int64_t work_bytes, float* out, uint64_t* counts,
                    int* overflow, int accumulate, void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_galmod_mc: model is NULL");
    int rc = check_box(m, box);
    if (rc) return rc;
    if (B < 0 || B > 65535 || nsub < 0 || nsub > 65535) return fail(BB_ERR_ARG, "bb_galmod_mc: bad sizes");
    if (nvpar < 1 || nvpar > 65535) return fail(BB_ERR_ARG, "bb_galmod_mc: nv = %d", nvpar);
    if (max_clouds < 0) return fail(BB_ERR_ARG, "bb_galmod_mc: max_clouds < 0");
    if (box.bhi0 - box.blo0 > 32767 || box.bhi1 - box.blo1 > 32767)
        return fail(BB_ERR_ARG, "bb_galmod_mc: box larger than 32767 pixels");
    if (bb_mc_lds_bytes(m->g.nz, m->TX, m->TY) > BB_LDS_MAX)
        return fail(BB_ERR_ARG, "bb_galmod_mc: the %dx%d tile of %d channels exceeds the LDS", m->TX, m->TY, m->g.nz);
    if (B == 0) return 0;
    if (!out || !work || !overflow || (nsub > 0 && !sr)) return fail(BB_ERR_ARG, "bb_galmod_mc: NULL array");
    if (work_bytes < bb_mc_work_size(m, box, nsub, B, max_clouds))
        return fail(BB_ERR_ARG, "bb_galmod_mc: work holds %lld bytes, needs %lld", (long long)work_bytes,
                    (long long)bb_mc_work_size(m, box, nsub, B, max_clouds));
    int nt;
    mc_tiles(m, box, &nt);
    char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    void* subs = w;
    w += mc_align((size_t)B * nsub * bb_mc_sub_bytes());
    double* qv = (double*)w;
    float* invq = (float*)(qv + B);
    int64_t* total = (int64_t*)(((uintptr_t)(invq + B) + 7) & ~(uintptr_t)7);
    w += mc_align((size_t)B * (sizeof(double) + sizeof(float) + sizeof(int64_t)));
    int* tcount = (int*)w;
    w += mc_align((size_t)B * nt * sizeof(int));
    int* tcursor = (int*)w;
    w += mc_align((size_t)B * nt * sizeof(int));
    int64_t* toff = (int64_t*)w;
    w += mc_align((size_t)B * (nt + 1) * sizeof(int64_t));
    w += mc_align(sizeof(int));
    void* recs = w;
    hipStream_t st = (hipStream_t)stream;
    rc = hip_check((hipError_t)bb_launch_zero_u32(tcount, (int64_t)B * nt, st), "bb_galmod_mc");
    if (!rc) rc = hip_check((hipError_t)bb_launch_zero_u32(overflow, 1, st), "bb_galmod_mc");
    if (rc) return rc;
    if (nsub == 0) {
        const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
        const int64_t nv = (int64_t)B * by * bx * m->g.nz;
        if (!accumulate) rc = hip_check((hipError_t)bb_launch_zero_u32(out, nv, st), "bb_galmod_mc");
        if (!rc && counts) rc = hip_check((hipError_t)bb_launch_zero_u32(counts, 2 * nv, st), "bb_galmod_mc");
        return rc;
    }
    rc = bb_launch_galmod_mc(&m->g, box, sr, nsub, B, nvpar, seed, m->TX, m->TY, max_clouds, subs, qv, invq, total,
                             tcount, toff, tcursor, recs, overflow, m->d_cd2i, out, (unsigned long long*)counts,
                             accumulate, vbox, st);
    return rc ? fail(BB_ERR_HIP, "bb_galmod_mc launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// ---------------------------------------------------------------------------------------------------------
// The 3DFIT ring residual.
extern "C" int64_t bb_norm_work_size(int B, int by, int bx, int nz, int with_support) {
    if (B < 0 || by < 1 || bx < 1 || nz < 1) return -1;
    int64_t n = (int64_t)B * by * bx * 3 + B;
    if (with_support) n += (int64_t)B * by * (bx + 1) * nz;
    return n;
}

extern "C" int bb_norm3dfit(const float* model, int B, int by, int bx, int nz, const float* data, const uint8_t* mask,
                            int ny, int nx, int blo0, int blo1, const double* ring, const double* chan_noise,
                            const bb_norm_par* par, const float* pre, const int32_t* ilo, const int32_t* ihi, int ky,
                            const double* thr, double* work, int64_t work_len, double* value, int64_t* nblank,
                            int64_t* ntot, void* stream) {
    if (!par) return fail(BB_ERR_ARG, "bb_norm3dfit: par is NULL");
    if (B < 0 || B > 65535 || by < 1 || bx < 1 || nz < 1 || ny < 1 || nx < 1)
        return fail(BB_ERR_ARG, "bb_norm3dfit: bad sizes");
    if (blo0 < 0 || blo1 < 0 || blo0 + bx > nx || blo1 + by > ny)
        return fail(BB_ERR_ARG, "bb_norm3dfit: the box [%d,%d)x[%d,%d) is outside the cube", blo0, blo0 + bx, blo1,
                    blo1 + by);
    if (par->ftype < 1 || par->ftype > 4) return fail(BB_ERR_ARG, "bb_norm3dfit: ftype = %d (1..4)", par->ftype);
    if (par->normtype < 0 || par->normtype > 2) return fail(BB_ERR_ARG, "bb_norm3dfit: normtype = %d", par->normtype);
    if (pre && (ky < 1 || ky > 4096 || !ilo || !ihi || !thr)) return fail(BB_ERR_ARG, "bb_norm3dfit: bad support");
    if (B == 0) return 0;
    if (!model || !data || !mask || !ring || !chan_noise || !work || !value)
        return fail(BB_ERR_ARG, "bb_norm3dfit: NULL array");
    if (work_len < bb_norm_work_size(B, by, bx, nz, pre != nullptr))
        return fail(BB_ERR_ARG, "bb_norm3dfit: work holds %lld doubles, needs %lld", (long long)work_len,
                    (long long)bb_norm_work_size(B, by, bx, nz, pre != nullptr));
    int rc = bb_launch_norm(B, by, bx, nz, ny, nx, blo0, blo1, par->wpow, par->side, par->ftype, par->data_noise,
                            par->bweight, par->normtype, par->foot, par->hmax, model, data, mask, ring, chan_noise, pre,
                            (const int*)ilo, (const int*)ihi, ky, thr, work, value, nblank, ntot, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_norm3dfit launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// ---------------------------------------------------------------------------------------------------------
// The likelihood surrogate (bb_mlp.hip): a din -> H (SiLU) x nh -> 1 network with FP32 master parameters, their
// BF16 copies, Adam moments and the activations of up to nmax rows, all on the device that created it.
extern "C" int bb_launch_mlp_gemm(int epi, const unsigned short* A, const unsigned short* B, int M, int N, int K,
                                  const float* bias, float* Z, unsigned short* Ab, unsigned short* AbT, float* G,
                                  const float* Zp, int nvalid, float* part, hipStream_t st);
extern "C" int bb_mlp_splits(int K);
extern "C" int bb_launch_mlp_in(const float* x, int n, int din, int np, int dp, unsigned short* X, unsigned short* XT,
                                hipStream_t st);
extern "C" int bb_launch_mlp_out(const unsigned short* a, const unsigned short* w, const float* b, int np, int H,
                                 float* y, hipStream_t st);
extern "C" int bb_launch_mlp_out_bwd(const float* y, const float* t, int n, int np, int H, const unsigned short* w,
                                     const float* Z, unsigned short* D, unsigned short* DT, hipStream_t st);
extern "C" int bb_launch_mlp_out_grad(const float* y, const float* t, int n, const unsigned short* a, int H, float* gw,
                                      float* gb, float* loss, hipStream_t st);
extern "C" int bb_launch_mlp_bias_grad(const unsigned short* DT, int np, int H, float* g, hipStream_t st);
extern "C" int bb_launch_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                                  float eps, const float* cc, hipStream_t st);
extern "C" int bb_launch_mlp_tick(int* count, float* cc, const float* loss_slot, float* hist, int hist_cap, float b1,
                                  float b2, hipStream_t st);
extern "C" int bb_launch_mlp_cast(const float* W, int rows, int cols, unsigned short* Wb, unsigned short* WbT,
                                  hipStream_t st);

extern "C" int bb_launch_walk_init(int* cnt, int steps, unsigned char* moved, int n, hipStream_t st);
extern "C" int bb_launch_walk_prop(const float* x, const float* zl, int n, int din, int s, float* sc, const int* cnt,
                                   const float* xc, const float* xs, float* prop, float* pn, hipStream_t st);
extern "C" int bb_launch_walk_accept(const float* prop, const float* y, int n, int din, const float* thr, float* x,
                                     unsigned char* moved, int* cnt_s, hipStream_t st);
extern "C" int bb_launch_walk_final(float* sc, const int* cnt, int steps, int n, hipStream_t st);

#define BB_MLP_DP 32          // padded input width (one MFMA K step)
#define BB_WALK_MAX_STEPS 4096

struct bb_mlp {
    int device;
    int din, H, nh, nmax;     // nmax: rows, a multiple of 32
    int64_t np_;              // parameters
    float* P;                 // FP32 master parameters: per hidden layer W (H x K) then b (H); then w (H), b (1)
    float* Gr;                // gradient
    float* Mo;                // Adam moments
    float* Ve;
    unsigned short* Wb;       // BF16 weights: per layer (H x K) and its transpose (K x H); then w (H)
    unsigned short *X, *XT;   // inputs (nmax x DP), (DP x nmax)
    float* Z;                 // pre-activations, per layer (nmax x H)
    unsigned short *A, *AT;   // activations, per layer (nmax x H), (H x nmax)
    unsigned short *D, *DT;   // layer gradients, two (nmax x H) + (H x nmax) sets
    float *y, *t;             // outputs and targets (nmax)
    int* cnt;                 // device: Adam steps taken, losses recorded by the current bb_mlp_train call
    float* cc;                // device: Adam bias corrections of the current step
    float* loss_slot;         // device: loss of the current step
    float* hist;              // device: losses of the steps of the current bb_mlp_train call
    float* part;              // device: split-K partial weight gradients (bb_mlp_splits(nmax) x H x max K)
    hipStream_t cs;           // private stream of the graph capture
    hipGraphExec_t gx;        // one captured training step (forward, backward, Adam, BF16 copies) for gx_rows rows
    float gx_lr;
    int gx_rows;
    float *wprop, *wpn;       // device: proposals and standardised inputs of a walk step (nmax x din)
    int* wcnt;                // device: acceptances per walk step (BB_WALK_MAX_STEPS)
    hipGraphExec_t wx;        // the last captured walk, for the arguments in wkey
    const void* wkey[8];
    int wrows, wsteps;
};

#define BB_MLP_HIST 4096      // most steps of one bb_mlp_train call

static int64_t mlp_K(const bb_mlp* n, int l) { return l == 0 ? BB_MLP_DP : n->H; }
static int64_t mlp_woff(const bb_mlp* n, int l) {       // offset of W_l in P (b_l follows it)
    int64_t o = 0;
    for (int i = 0; i < l; i++) o += (int64_t)n->H * mlp_K(n, i) + n->H;
    return o;
}
static int64_t mlp_bwoff(const bb_mlp* n, int l) {      // offset of W_l (then its transpose) in Wb
    int64_t o = 0;
    for (int i = 0; i < l; i++) o += 2 * (int64_t)n->H * mlp_K(n, i);
    return o;
}

extern "C" int bb_mlp_destroy(bb_mlp* n) {
    if (!n) return 0;
    if (n->gx) (void)hipGraphExecDestroy(n->gx);
    if (n->cs) (void)hipStreamDestroy(n->cs);
    if (n->wx) (void)hipGraphExecDestroy(n->wx);
    void* ptrs[] = {n->P, n->Gr, n->Mo, n->Ve, n->Wb, n->X, n->XT, n->Z, n->A, n->AT, n->D, n->DT, n->y, n->t,
                    n->cnt, n->cc, n->loss_slot, n->hist, n->part, n->wprop, n->wpn, n->wcnt};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete n;
    return 0;
}

extern "C" int bb_mlp_create(int din, int hidden, int nhidden, int nmax, bb_mlp** out) {
    if (!out) return fail(BB_ERR_ARG, "bb_mlp_create: out is NULL");
    *out = nullptr;
    if (din < 1 || din > BB_MLP_DP) return fail(BB_ERR_ARG, "bb_mlp_create: din = %d (1..%d)", din, BB_MLP_DP);
    if (hidden < 32 || hidden > 512 || hidden % 32) return fail(BB_ERR_ARG, "bb_mlp_create: hidden = %d", hidden);
    if (nhidden < 1 || nhidden > 16) return fail(BB_ERR_ARG, "bb_mlp_create: nhidden = %d", nhidden);
    if (nmax < 1 || nmax > (1 << 22)) return fail(BB_ERR_ARG, "bb_mlp_create: nmax = %d", nmax);
    bb_mlp* n = new bb_mlp();
    memset(n, 0, sizeof *n);
    (void)hipGetDevice(&n->device);
    n->din = din;
    n->H = hidden;
    n->nh = nhidden;
    n->nmax = (nmax + 31) / 32 * 32;
    n->np_ = mlp_woff(n, nhidden) + hidden + 1;
    const int64_t nw = mlp_bwoff(n, nhidden) + hidden;
    const int64_t R = n->nmax, H = hidden;
    int rc = 0;
    auto alloc = [&](void** p, size_t bytes, const char* what) {
        if (!rc) rc = hip_check(hipMalloc(p, bytes), what);
        if (!rc) rc = hip_check(hipMemset(*p, 0, bytes), what);
    };
    alloc((void**)&n->P, n->np_ * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->Gr, n->np_ * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->Mo, n->np_ * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->Ve, n->np_ * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->Wb, nw * 2, "bb_mlp_create");
    alloc((void**)&n->X, R * BB_MLP_DP * 2, "bb_mlp_create");
    alloc((void**)&n->XT, R * BB_MLP_DP * 2, "bb_mlp_create");
    alloc((void**)&n->Z, R * H * nhidden * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->A, R * H * nhidden * 2, "bb_mlp_create");
    alloc((void**)&n->AT, R * H * nhidden * 2, "bb_mlp_create");
    alloc((void**)&n->D, R * H * 2 * 2, "bb_mlp_create");
    alloc((void**)&n->DT, R * H * 2 * 2, "bb_mlp_create");
    alloc((void**)&n->y, R * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->t, R * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->cnt, 2 * sizeof(int), "bb_mlp_create");
    alloc((void**)&n->cc, 2 * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->loss_slot, sizeof(float), "bb_mlp_create");
    alloc((void**)&n->hist, BB_MLP_HIST * sizeof(float), "bb_mlp_create");
    alloc((void**)&n->part, (size_t)bb_mlp_splits((int)R) * H * (H > BB_MLP_DP ? H : BB_MLP_DP) * sizeof(float),
          "bb_mlp_create");
    if (!rc) rc = hip_check(hipStreamCreateWithFlags(&n->cs, hipStreamNonBlocking), "bb_mlp_create");
    if (rc) {
        bb_mlp_destroy(n);
        return rc;
    }
    *out = n;
    return 0;
}

extern "C" int64_t bb_mlp_param_count(const bb_mlp* n) { return n ? n->np_ : -1; }

static int mlp_cast(bb_mlp* n, hipStream_t st) {
    int rc = 0;
    for (int l = 0; l < n->nh && !rc; l++) {
        const int K = (int)mlp_K(n, l);
        unsigned short* wb = n->Wb + mlp_bwoff(n, l);
        rc = bb_launch_mlp_cast(n->P + mlp_woff(n, l), n->H, K, wb, wb + (int64_t)n->H * K, st);
    }
    if (!rc) rc = bb_launch_mlp_cast(n->P + mlp_woff(n, n->nh), 1, n->H, n->Wb + mlp_bwoff(n, n->nh), nullptr, st);
    return rc;
}

extern "C" int bb_mlp_set_params(bb_mlp* n, const float* host) {
    if (!n || !host) return fail(BB_ERR_ARG, "bb_mlp_set_params: NULL argument");
    int rc = hip_check(hipMemcpy(n->P, host, n->np_ * sizeof(float), hipMemcpyHostToDevice), "bb_mlp_set_params");
    if (!rc) rc = hip_check(hipMemset(n->Mo, 0, n->np_ * sizeof(float)), "bb_mlp_set_params");
    if (!rc) rc = hip_check(hipMemset(n->Ve, 0, n->np_ * sizeof(float)), "bb_mlp_set_params");
    if (!rc) rc = hip_check(hipMemset(n->cnt, 0, 2 * sizeof(int)), "bb_mlp_set_params");
    if (!rc) rc = hip_check((hipError_t)mlp_cast(n, nullptr), "bb_mlp_set_params");
    if (!rc) rc = hip_check(hipDeviceSynchronize(), "bb_mlp_set_params");
    return rc;
}

extern "C" int bb_mlp_get_params(const bb_mlp* n, float* host) {
    if (!n || !host) return fail(BB_ERR_ARG, "bb_mlp_get_params: NULL argument");
    return hip_check(hipMemcpy(host, n->P, n->np_ * sizeof(float), hipMemcpyDeviceToHost), "bb_mlp_get_params");
}

// the layers of the forward pass of rows [0, rows) of n->X (activations kept for a backward pass; outputs in n->y)
static int mlp_layers(bb_mlp* n, int rows, hipStream_t st) {
    const int R = (rows + 31) / 32 * 32, H = n->H;
    int rc = 0;
    for (int l = 0; l < n->nh && !rc; l++) {
        const unsigned short* in = l == 0 ? n->X : n->A + (int64_t)(l - 1) * n->nmax * H;
        const int K = (int)mlp_K(n, l);
        rc = bb_launch_mlp_gemm(0, in, n->Wb + mlp_bwoff(n, l), R, H, K, n->P + mlp_woff(n, l) + (int64_t)H * K,
                                n->Z + (int64_t)l * n->nmax * H, n->A + (int64_t)l * n->nmax * H,
                                n->AT + (int64_t)l * n->nmax * H, nullptr, nullptr, rows, nullptr, st);
    }
    if (!rc)
        rc = bb_launch_mlp_out(n->A + (int64_t)(n->nh - 1) * n->nmax * H, n->Wb + mlp_bwoff(n, n->nh),
                               n->P + mlp_woff(n, n->nh) + H, R, H, n->y, st);
    return rc;
}

// forward of rows [0, rows) of x (device, rows x din)
static int mlp_forward(bb_mlp* n, const float* x, int rows, hipStream_t st) {
    const int R = (rows + 31) / 32 * 32;
    int rc = bb_launch_mlp_in(x, rows, n->din, R, BB_MLP_DP, n->X, n->XT, st);
    if (!rc) rc = mlp_layers(n, rows, st);
    return rc;
}

// gradient of 0.5 mean (y - t)^2 over rows [0, n) into n->Gr (after mlp_forward of the same rows); loss (device)
static int mlp_backward(bb_mlp* n, const float* t, int rows, float* loss, hipStream_t st) {
    const int R = (rows + 31) / 32 * 32, H = n->H, L = n->nh;
    const int64_t lay = (int64_t)n->nmax * H;
    float* gout = n->Gr + mlp_woff(n, L);
    int rc = bb_launch_mlp_out_grad(n->y, t, rows, n->A + (L - 1) * lay, H, gout, gout + H, loss, st);
    if (!rc)
        rc = bb_launch_mlp_out_bwd(n->y, t, rows, R, H, n->Wb + mlp_bwoff(n, L), n->Z + (L - 1) * lay, n->D, n->DT,
                                   st);
    int cur = 0;
    for (int l = L - 1; l >= 0 && !rc; l--) {
        const int K = (int)mlp_K(n, l);
        unsigned short* D = n->D + cur * lay;
        unsigned short* DT = n->DT + cur * lay;
        const unsigned short* inT = l == 0 ? n->XT : n->AT + (l - 1) * lay;
        float* gw = n->Gr + mlp_woff(n, l);
        // dW (H x K) = D^T (H x R) . in (R x K): rows of D^T and of in^T, contiguous along the R samples
        rc = bb_launch_mlp_gemm(1, DT, inT, H, K, R, nullptr, nullptr, nullptr, nullptr, gw, nullptr, rows, n->part,
                                st);
        if (!rc) rc = bb_launch_mlp_bias_grad(DT, R, H, gw + (int64_t)H * K, st);
        if (!rc && l > 0) {
            // the gradient of layer l - 1: (D (R x H) . W (H x K)) * silu'(Z_{l-1}), W^T rows contiguous along H
            const unsigned short* WT = n->Wb + mlp_bwoff(n, l) + (int64_t)H * K;
            rc = bb_launch_mlp_gemm(2, D, WT, R, K, H, nullptr, nullptr, n->D + (1 - cur) * lay,
                                    n->DT + (1 - cur) * lay, nullptr, n->Z + (l - 1) * lay, rows, nullptr, st);
            cur = 1 - cur;
        }
    }
    return rc;
}

static int mlp_check_rows(const bb_mlp* n, const void* x, int rows, const char* what) {
    if (!n || !x) return fail(BB_ERR_ARG, "%s: NULL argument", what);
    if (rows < 1 || rows > n->nmax) return fail(BB_ERR_ARG, "%s: %d rows (1..%d)", what, rows, n->nmax);
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != n->device) return fail(BB_ERR_ARG, "%s: the network lives on device %d, not %d", what, n->device, dev);
    return 0;
}

extern "C" int bb_mlp_forward(bb_mlp* n, const float* x, int rows, float* y, void* stream) {
    int rc = mlp_check_rows(n, x, rows, "bb_mlp_forward");
    if (rc) return rc;
    if (!y) return fail(BB_ERR_ARG, "bb_mlp_forward: y is NULL");
    hipStream_t st = (hipStream_t)stream;
    rc = mlp_forward(n, x, rows, st);
    if (!rc) rc = (int)hipMemcpyAsync(y, n->y, (size_t)rows * sizeof(float), hipMemcpyDeviceToDevice, st);
    return rc ? fail(BB_ERR_HIP, "bb_mlp_forward: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_mlp_grad(bb_mlp* n, const float* x, const float* t, int rows, float* grad, float* loss, void* stream) {
    int rc = mlp_check_rows(n, x, rows, "bb_mlp_grad");
    if (rc) return rc;
    if (!t || !grad) return fail(BB_ERR_ARG, "bb_mlp_grad: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    rc = mlp_forward(n, x, rows, st);
    if (!rc) rc = mlp_backward(n, t, rows, loss, st);
    if (!rc) rc = (int)hipMemcpyAsync(grad, n->Gr, n->np_ * sizeof(float), hipMemcpyDeviceToDevice, st);
    return rc ? fail(BB_ERR_HIP, "bb_mlp_grad: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// One training step on the rows of n->X and n->t: forward, backward, Adam (constants from the device counter),
// BF16 copies. The same launches every step, so it is captured once into a graph and replayed.
static int mlp_train_step(bb_mlp* n, int rows, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int rc = mlp_layers(n, rows, st);
    if (!rc) rc = mlp_backward(n, n->t, rows, n->loss_slot, st);
    if (!rc) rc = bb_launch_mlp_tick(n->cnt, n->cc, n->loss_slot, n->hist, BB_MLP_HIST, b1, b2, st);
    if (!rc) rc = bb_launch_mlp_adam(n->P, n->Gr, n->Mo, n->Ve, n->np_, lr, b1, b2, eps, n->cc, st);
    if (!rc) rc = mlp_cast(n, st);
    return rc;
}

extern "C" int bb_mlp_train(bb_mlp* n, const float* x, const float* t, int rows, int steps, float lr, float* loss,
                            void* stream) {
    int rc = mlp_check_rows(n, x, rows, "bb_mlp_train");
    if (rc) return rc;
    if (!t) return fail(BB_ERR_ARG, "bb_mlp_train: t is NULL");
    if (steps < 0 || steps > BB_MLP_HIST || !(lr > 0.f) || !isfinite(lr))
        return fail(BB_ERR_ARG, "bb_mlp_train: steps %d (0..%d), lr %g", steps, BB_MLP_HIST, lr);
    if (steps == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int R = (rows + 31) / 32 * 32;
    // the inputs and targets into the network's own buffers (the captured step reads only those)
    rc = bb_launch_mlp_in(x, rows, n->din, R, BB_MLP_DP, n->X, n->XT, st);
    if (!rc) rc = (int)hipMemcpyAsync(n->t, t, (size_t)rows * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (!rc) rc = (int)hipMemsetAsync(n->cnt + 1, 0, sizeof(int), st);
    if (rc) return fail(BB_ERR_HIP, "bb_mlp_train: %s", hipGetErrorString((hipError_t)rc));
    if (!n->gx || n->gx_rows != rows || n->gx_lr != lr) {
        // capture one step on the private stream
        if (n->gx) {
            (void)hipGraphExecDestroy(n->gx);
            n->gx = nullptr;
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(n->cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const int r2 = mlp_train_step(n, rows, lr, n->cs);
            e = hipStreamEndCapture(n->cs, &g);
            if (e == hipSuccess && r2) e = (hipError_t)r2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&n->gx, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            n->gx = nullptr;
            return fail(BB_ERR_HIP, "bb_mlp_train: graph capture: %s", hipGetErrorString(e));
        }
        n->gx_rows = rows;
        n->gx_lr = lr;
    }
    for (int s = 0; s < steps && !rc; s++) rc = (int)hipGraphLaunch(n->gx, st);
    if (!rc && loss) rc = (int)hipMemcpyAsync(loss, n->hist, (size_t)steps * sizeof(float), hipMemcpyDeviceToDevice, st);
    return rc ? fail(BB_ERR_HIP, "bb_mlp_train: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// The walk's launches: counters cleared, then per step the proposals, the forward pass and the acceptance.
static int mlp_walk_body(bb_mlp* n, float* x, const float* zl, int rows, int steps, const float* thr, const float* xc,
                         const float* xs, float* sc, unsigned char* moved, hipStream_t st) {
    int rc = bb_launch_walk_init(n->wcnt, steps, moved, rows, st);
    for (int s = 0; s < steps && !rc; s++) {
        rc = bb_launch_walk_prop(x, zl, rows, n->din, s, sc, n->wcnt, xc, xs, n->wprop, n->wpn, st);
        if (!rc) rc = mlp_forward(n, n->wpn, rows, st);
        if (!rc) rc = bb_launch_walk_accept(n->wprop, n->y, rows, n->din, thr, x, moved, n->wcnt + s, st);
    }
    if (!rc) rc = bb_launch_walk_final(sc, n->wcnt, steps, rows, st);
    return rc;
}

extern "C" int bb_mlp_walk(bb_mlp* n, float* x, const float* zl, int rows, int steps, const float* thr,
                           const float* xc, const float* xs, float* sc, unsigned char* moved, void* stream) {
    int rc = mlp_check_rows(n, x, rows, "bb_mlp_walk");
    if (rc) return rc;
    if (steps < 1 || steps > BB_WALK_MAX_STEPS)
        return fail(BB_ERR_ARG, "bb_mlp_walk: steps = %d (1..%d)", steps, BB_WALK_MAX_STEPS);
    if (!zl || !thr || !xc || !xs || !sc || !moved) return fail(BB_ERR_ARG, "bb_mlp_walk: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    if (!n->wprop) {
        rc = hip_check(hipMalloc((void**)&n->wprop, (size_t)n->nmax * n->din * sizeof(float)), "bb_mlp_walk");
        if (!rc) rc = hip_check(hipMalloc((void**)&n->wpn, (size_t)n->nmax * n->din * sizeof(float)), "bb_mlp_walk");
        if (!rc) rc = hip_check(hipMalloc((void**)&n->wcnt, BB_WALK_MAX_STEPS * sizeof(int)), "bb_mlp_walk");
        if (rc) return rc;
    }
    const void* key[8] = {x, zl, thr, xc, xs, sc, moved, nullptr};
    bool same = n->wx && n->wrows == rows && n->wsteps == steps;
    for (int i = 0; i < 8 && same; i++) same = key[i] == n->wkey[i];
    if (!same) {
        // capture the whole walk on the private stream (counters are cleared by a kernel: no memset node)
        if (n->wx) {
            (void)hipGraphExecDestroy(n->wx);
            n->wx = nullptr;
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(n->cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const int r2 = mlp_walk_body(n, x, zl, rows, steps, thr, xc, xs, sc, moved, n->cs);
            e = hipStreamEndCapture(n->cs, &g);
            if (e == hipSuccess && r2) e = (hipError_t)r2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&n->wx, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            n->wx = nullptr;
            return fail(BB_ERR_HIP, "bb_mlp_walk: graph capture: %s", hipGetErrorString(e));
        }
        memcpy(n->wkey, key, sizeof key);
        n->wrows = rows;
        n->wsteps = steps;
    }
    rc = (int)hipGraphLaunch(n->wx, st);
    return rc ? fail(BB_ERR_HIP, "bb_mlp_walk: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

// ---------------------------------------------------------------------------------------------------------
// Per-fit set-up: the likelihood mask (dilate_mask, bbarolo/bayesian/noise.py:18-30).
extern "C" int bb_dilate_mask(const uint8_t* mask, int nz, int ny, int nx, double radius, int nchan, uint8_t* tmp,
                              uint8_t* out, void* stream) {
    if (nz < 0 || ny < 0 || nx < 0 || nz > 65535 || (ny + 31) / 32 > 65535 || (nx + 31) / 32 > 65535)
        return fail(BB_ERR_ARG, "bb_dilate_mask: bad sizes %d x %d x %d", nz, ny, nx);
    if (!(radius == radius) || radius > 1e9 || nchan < 0)
        return fail(BB_ERR_ARG, "bb_dilate_mask: bad radius %g or nchan %d", radius, nchan);
    if (radius >= 1.0 && (int)floor(radius) > bb_dilate_rmax())
        return fail(BB_ERR_ARG, "bb_dilate_mask: radius %g pixels exceeds %d", radius, bb_dilate_rmax());
    if ((int64_t)nz * ny * nx == 0) return 0;
    if (!mask || !out) return fail(BB_ERR_ARG, "bb_dilate_mask: NULL array");
    if (radius >= 1.0 && nchan > 0 && !tmp) return fail(BB_ERR_ARG, "bb_dilate_mask: tmp is needed");
    if (out == mask || (tmp && (out == tmp || tmp == mask))) return fail(BB_ERR_ARG, "bb_dilate_mask: arrays alias");
    int rc = bb_launch_dilate(mask, nz, ny, nx, radius, nchan, tmp, out, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_dilate_mask launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}
