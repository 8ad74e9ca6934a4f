// The C-ABI of libbbgpu.so: argument checks, the per-fit model object (GalmodSetup + beam kernel on one
// device), error reporting and the launches of the kernels of bb_galmod.hip, bb_conv.hip, bb_resid.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/bbgpu.h"

extern "C" int bb_launch_galmod_analytic(const bb_grid* g, bb_box box, const double* sr, int nsub, int B,
                                         int chunks, const double* cd2i, const double* vedges, const double* ptab,
                                         int ntab, float* out, hipStream_t st);
extern "C" int bb_launch_conv2d_same(const float* in, float* out, int B, int by, int bx, int nz, const float* kern,
                                     const int* rowspan, int ky, int kx, float scale, hipStream_t st);
extern "C" int bb_residual_blocks(int64_t nvox);
extern "C" int bb_launch_residual(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2,
                                  double eps2, double* partial, double* out_ss, double* out_lg, hipStream_t st);

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define BB_ERR_ARG 1
#define BB_ERR_HIP 2

static int hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(BB_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

struct bb_model {
    bb_grid g;
    int device;
    int ntab;
    double* d_cd2i;
    double* d_vedges;
    double* d_ptab;
    int ky, kx;
    float* d_kern;
    int* d_span;
    double scale;
};

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

template <class T>
static int upload(T** dst, const T* src, size_t n, const char* what) {
    *dst = nullptr;
    if (n == 0) return 0;
    int rc = hip_check(hipMalloc((void**)dst, n * sizeof(T)), what);
    if (rc) return rc;
    return hip_check(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), what);
}

extern "C" int bb_model_destroy(bb_model* m) {
    if (!m) return 0;
    int dev = 0;
    hipGetDevice(&dev);
    hipSetDevice(m->device);
    hipFree(m->d_cd2i);
    hipFree(m->d_vedges);
    hipFree(m->d_ptab);
    hipFree(m->d_kern);
    hipFree(m->d_span);
    hipSetDevice(dev);
    delete m;
    return 0;
}

extern "C" int bb_model_create(const bb_grid* g, const double* cd2i, const double* vedges, const double* ptab,
                               int ntab, const double* kern, int ky, int kx, double scale, bb_model** out) {
    if (!out) return fail(BB_ERR_ARG, "bb_model_create: out is NULL");
    *out = nullptr;
    int rc = check_grid(g);
    if (rc) return rc;
    if (!cd2i || !vedges || !ptab || ntab < 2) return fail(BB_ERR_ARG, "bb_model_create: missing setup arrays");
    if (kern && (ky < 1 || kx < 1 || ky % 2 == 0 || kx % 2 == 0 || ky > 511 || ky * kx > 4096))
        return fail(BB_ERR_ARG, "bb_model_create: kernel %dx%d must be odd-sized, ky <= 511, <= 4096 taps", ky, kx);
    bb_model* m = new bb_model();
    memset(m, 0, sizeof *m);
    m->g = *g;
    m->ntab = ntab;
    m->scale = scale;
    hipGetDevice(&m->device);
    rc = upload(&m->d_cd2i, cd2i, (size_t)g->nz, "upload cd2i");
    if (!rc) rc = upload(&m->d_vedges, vedges, (size_t)g->nlines * (g->nz + 1), "upload vedges");
    if (!rc) rc = upload(&m->d_ptab, ptab, (size_t)ntab, "upload Phi table");
    if (!rc && kern) {
        // float32 taps and the non-zero span [ilo, ihi] of every kernel row (ihi < ilo: empty row)
        std::vector<float> kf((size_t)ky * kx);
        std::vector<int> span(2 * (size_t)ky);
        for (int j = 0; j < ky; j++) {
            int lo = kx, hi = -1;
            for (int i = 0; i < kx; i++) {
                double v = kern[(size_t)j * kx + i];
                kf[(size_t)j * kx + i] = (float)v;
                if (v != 0) {
                    if (lo == kx) lo = i;
                    hi = i;
                }
            }
            span[2 * j] = lo;
            span[2 * j + 1] = hi;
        }
        m->ky = ky;
        m->kx = kx;
        rc = upload(&m->d_kern, kf.data(), kf.size(), "upload kernel");
        if (!rc) rc = upload(&m->d_span, span.data(), span.size(), "upload kernel spans");
    }
    if (rc) {
        bb_model_destroy(m);
        return rc;
    }
    *out = m;
    return 0;
}

static int check_box(const bb_model* m, bb_box box) {
    if (box.bhi0 <= box.blo0 || box.bhi1 <= box.blo1)
        return fail(BB_ERR_ARG, "box [%d,%d)x[%d,%d) is empty", box.blo0, box.bhi0, box.blo1, box.bhi1);
    (void)m;
    return 0;
}

extern "C" int bb_galmod_analytic(const bb_model* m, bb_box box, const double* sr, int nsub, int B, int chunks,
                                  float* out, void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_galmod_analytic: model is NULL");
    int rc = check_box(m, box);
    if (rc) return rc;
    if (nsub < 0 || B < 0 || chunks < 1 || chunks > 65535 || 2 * nsub > 65535 || B > 65535)
        return fail(BB_ERR_ARG, "bb_galmod_analytic: nsub=%d B=%d chunks=%d out of range", nsub, B, chunks);
    if (nsub == 0 || B == 0) return 0;
    if (!sr || !out) return fail(BB_ERR_ARG, "bb_galmod_analytic: NULL array");
    rc = bb_launch_galmod_analytic(&m->g, box, sr, nsub, B, chunks, m->d_cd2i, m->d_vedges, m->d_ptab, m->ntab, out,
                                   (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_galmod_analytic launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}

extern "C" int bb_conv2d_same(const bb_model* m, const float* in, float* out, int B, int by, int bx,
                              void* stream) {
    if (!m) return fail(BB_ERR_ARG, "bb_conv2d_same: model is NULL");
    if (!m->d_kern) return fail(BB_ERR_ARG, "bb_conv2d_same: the model has no beam kernel");
    if (B < 0 || by < 1 || bx < 1 || B > 65535) return fail(BB_ERR_ARG, "bb_conv2d_same: bad sizes");
    if (B == 0) return 0;
    if (!in || !out || in == out) return fail(BB_ERR_ARG, "bb_conv2d_same: in/out must be distinct, non-NULL");
    int rc = bb_launch_conv2d_same(in, out, B, by, bx, m->g.nz, m->d_kern, m->d_span, m->ky, m->kx,
                                   (float)m->scale, (hipStream_t)stream);
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
    int rc = bb_launch_residual(model, dvalid, B, nvox, sigma2, eps2, work, out_ss, out_lg, (hipStream_t)stream);
    return rc ? fail(BB_ERR_HIP, "bb_residual_gauss launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}
