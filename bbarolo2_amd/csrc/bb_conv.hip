// Beam convolution of model cubes on the GPU: 'same' 2D convolution of every channel with the beam
// kernel, then values below 1e-12 set to zero and the flux scale applied (FFTConvolver and
// DiskModel.convolve, bbarolo/smooth.py:145-165 and galmod.py:1273-1283; _conv_sparse galmod.py:974-1037).
//
// Layout yxz (spectra contiguous): the lanes of a wavefront run along the channels, so every load is a
// contiguous segment and all lanes follow the same control flow; the kernel taps are wave-uniform (scalar
// loads). Each thread computes a block of RY x RX output pixels of one channel: it walks the RY + KY - 1
// input rows that feed the block, keeps one input row segment (RX + KX - 1 values) in registers and adds it,
// through every kernel row that links it to an output row, to the RY x RX accumulators. The kernel width KX
// is a template parameter (odd, <= 25) so that the row loops unroll into registers; taps outside the beam's
// 1e-4 cut are zero and simply add nothing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/bbgpu.h"
#include "bb_internal.h"

#define BB_CV_RY 4
#define BB_CV_RX 4
#define BB_CV_THREADS 256

template <int KX>
__global__ void __launch_bounds__(BB_CV_THREADS)
k_conv2d_same(const float* __restrict__ inall, float* __restrict__ outall, int by, int bx, int nz, int ZL,
              const float* __restrict__ kern, int ky, float scale) {
    constexpr int W = BB_CV_RX + KX - 1;
    constexpr int hx = KX / 2;
    const int hy = ky / 2;
    const int b = blockIdx.z;
    const int zl = threadIdx.x % ZL, grp = threadIdx.x / ZL, ngrp = blockDim.x / ZL;
    const int z = blockIdx.y * ZL + zl;
    const int nbx = (bx + BB_CV_RX - 1) / BB_CV_RX, nby = (by + BB_CV_RY - 1) / BB_CV_RY;
    const int blk = blockIdx.x * ngrp + grp;
    if (z >= nz || blk >= nbx * nby) return;
    const int y0 = (blk / nbx) * BB_CV_RY, x0 = (blk % nbx) * BB_CV_RX;
    const float* in = inall + (size_t)b * by * bx * nz + z;
    float acc[BB_CV_RY][BB_CV_RX];
#pragma unroll
    for (int r = 0; r < BB_CV_RY; r++)
#pragma unroll
        for (int c = 0; c < BB_CV_RX; c++) acc[r][c] = 0.f;
    // output (y, x) = sum_{j,i} K[j][i] in[y + hy - j][x + hx - i]
    const int iylo = y0 - hy, iyhi = y0 + BB_CV_RY - 1 + hy;
    for (int iy = iylo; iy <= iyhi; iy++) {
        if (iy < 0 || iy >= by) continue;
        float row[W];
        const float* ir = in + (size_t)iy * bx * nz;
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int ix = x0 - hx + w;
            row[w] = (ix >= 0 && ix < bx) ? ir[(size_t)ix * nz] : 0.f;
        }
#pragma unroll
        for (int r = 0; r < BB_CV_RY; r++) {
            const int j = y0 + r + hy - iy;
            if (j < 0 || j >= ky) continue;
            const float* kr = kern + j * KX;
#pragma unroll
            for (int i = 0; i < KX; i++) {
                const float kv = kr[i];
#pragma unroll
                for (int c = 0; c < BB_CV_RX; c++) acc[r][c] = fmaf(kv, row[c + KX - 1 - i], acc[r][c]);
            }
        }
    }
    float* out = outall + (size_t)b * by * bx * nz + z;
#pragma unroll
    for (int r = 0; r < BB_CV_RY; r++) {
        const int y = y0 + r;
        if (y >= by) break;
#pragma unroll
        for (int c = 0; c < BB_CV_RX; c++) {
            const int x = x0 + c;
            if (x < bx) {
                float v = acc[r][c];
                out[((size_t)y * bx + x) * nz] = v < 1e-12f ? 0.f : v * scale;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// The convolution fused with the masked Gaussian residual (FFTConvolver then _sumsq_residuals /
// _residuals_fracerr, bbarolo/smooth.py:145-165, bbarolo/bayesian/likelihood.py:22-77): every convolved voxel is
// compared at once with the data (dvalid, yxz, NaN where not summed) and never written. The terms are summed in
// FP64 per thread, then per block in a fixed order into partial (2 doubles per block: the residual sum and, for
// FRAC, the log-variance sum); k_residual_final reduces the blocks of every vector in a fixed order.
__device__ __forceinline__ double cv_wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <int KX, bool FRAC>
__global__ void __launch_bounds__(BB_CV_THREADS)
k_conv_resid(const float* __restrict__ inall, const float* __restrict__ dvalid, int by, int bx, int nz, int ZL,
             const float* __restrict__ kern, int ky, float scale, float sigma2f, float eps2f,
             double* __restrict__ partial, const int4* __restrict__ vbox, const float2* __restrict__ sv) {
    constexpr int W = BB_CV_RX + KX - 1;
    constexpr int hx = KX / 2;
    __shared__ double red[2][BB_CV_THREADS / 64];
    const int hy = ky / 2;
    const int b = blockIdx.z;
    // the voxels summed: the vector's own model box (box-relative [x, z) x [y, w)), where the reference's model
    // of that box lives; the rest of the region enters through the outside sum of the caller
    int rx0 = 0, ry0 = 0, rx1 = bx, ry1 = by;
    if (vbox) {
        const int4 v = vbox[b];
        rx0 = v.x;
        ry0 = v.y;
        rx1 = v.z;
        ry1 = v.w;
    }
    if (sv) {
        sigma2f = sv[b].x;
        eps2f = sv[b].y;
    }
    const int zl = threadIdx.x % ZL, grp = threadIdx.x / ZL, ngrp = blockDim.x / ZL;
    const int z = blockIdx.y * ZL + zl;
    const int nbx = (bx + BB_CV_RX - 1) / BB_CV_RX, nby = (by + BB_CV_RY - 1) / BB_CV_RY;
    const int blk = blockIdx.x * ngrp + grp;
    double s = 0., lg = 0.;
    if (z < nz && blk < nbx * nby) {
        const int y0 = (blk / nbx) * BB_CV_RY, x0 = (blk % nbx) * BB_CV_RX;
        const float* in = inall + (size_t)b * by * bx * nz + z;
        float acc[BB_CV_RY][BB_CV_RX];
#pragma unroll
        for (int r = 0; r < BB_CV_RY; r++)
#pragma unroll
            for (int c = 0; c < BB_CV_RX; c++) acc[r][c] = 0.f;
        const int iylo = y0 - hy, iyhi = y0 + BB_CV_RY - 1 + hy;
        for (int iy = iylo; iy <= iyhi; iy++) {
            if (iy < 0 || iy >= by) continue;
            float row[W];
            const float* ir = in + (size_t)iy * bx * nz;
#pragma unroll
            for (int w = 0; w < W; w++) {
                const int ix = x0 - hx + w;
                row[w] = (ix >= 0 && ix < bx) ? ir[(size_t)ix * nz] : 0.f;
            }
#pragma unroll
            for (int r = 0; r < BB_CV_RY; r++) {
                const int j = y0 + r + hy - iy;
                if (j < 0 || j >= ky) continue;
                const float* kr = kern + j * KX;
#pragma unroll
                for (int i = 0; i < KX; i++) {
                    const float kv = kr[i];
#pragma unroll
                    for (int c = 0; c < BB_CV_RX; c++) acc[r][c] = fmaf(kv, row[c + KX - 1 - i], acc[r][c]);
                }
            }
        }
        const float* dv = dvalid + z;
#pragma unroll
        for (int r = 0; r < BB_CV_RY; r++) {
            const int y = y0 + r;
            if (y >= by) break;
            if (y < ry0 || y >= ry1) continue;
#pragma unroll
            for (int c = 0; c < BB_CV_RX; c++) {
                const int x = x0 + c;
                if (x >= rx0 && x < rx1) {
                    const float d = dv[((size_t)y * bx + x) * nz];
                    if (d == d) {
                        const float v = acc[r][c];
                        const float mm = v < 1e-12f ? 0.f : v * scale;
                        const float rr = d - mm;
                        if (!FRAC) {
                            s += (double)(rr * rr);
                        } else if (mm != 0.f) {
                            const float var = fmaf(eps2f * mm, mm, sigma2f);
                            s += (double)(rr * rr / var);
                            lg += (double)logf(var / sigma2f);
                        } else {
                            s += (double)(rr * rr / sigma2f);
                        }
                    }
                }
            }
        }
    }
    s = cv_wave_sum(s);
    if (FRAC) lg = cv_wave_sum(lg);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = s;
        red[1][w] = lg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0., c = 0.;
        for (int k = 0; k < BB_CV_THREADS / 64; k++) {
            a += red[0][k];
            c += red[1][k];
        }
        const size_t o = ((size_t)b * gridDim.y * gridDim.x + (size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
        partial[o] = a;
        partial[o + 1] = c;
    }
}

// Channel lanes per group: the width in {64, 32, 16, 8} that leaves the fewest idle lanes.
static int pick_zl(int nz) {
    int best = 64, waste = 1 << 30;
    for (int zl : {64, 32, 16, 8}) {
        int w = ((nz + zl - 1) / zl) * zl - nz;
        if (w < waste) {
            waste = w;
            best = zl;
        }
    }
    return best;
}

template <int KX>
static void launch(const float* in, float* out, int B, int by, int bx, int nz, const float* kern, int ky, float scale,
                   hipStream_t st) {
    const int ZL = pick_zl(nz);
    const int ngrp = BB_CV_THREADS / ZL;
    const int nblk = ((bx + BB_CV_RX - 1) / BB_CV_RX) * ((by + BB_CV_RY - 1) / BB_CV_RY);
    dim3 grid((nblk + ngrp - 1) / ngrp, (nz + ZL - 1) / ZL, B);
    hipLaunchKernelGGL(k_conv2d_same<KX>, grid, dim3(BB_CV_THREADS), 0, st, in, out, by, bx, nz, ZL, kern, ky, scale);
}

extern "C" int bb_conv_max_kx() { return 25; }

// Blocks per vector of k_conv_resid (2 partial doubles each).
extern "C" int64_t bb_conv_resid_blocks(int by, int bx, int nz) {
    const int ZL = pick_zl(nz);
    const int ngrp = BB_CV_THREADS / ZL;
    const int64_t nblk = (int64_t)((bx + BB_CV_RX - 1) / BB_CV_RX) * ((by + BB_CV_RY - 1) / BB_CV_RY);
    return ((nblk + ngrp - 1) / ngrp) * ((nz + ZL - 1) / ZL);
}

template <int KX>
static void launch_cr(const float* in, const float* dvalid, int B, int by, int bx, int nz, const float* kern, int ky,
                      float scale, float sigma2, float eps2, double* partial, const int4* vbox, const float2* sv,
                      bool frac, hipStream_t st) {
    const int ZL = pick_zl(nz);
    const int ngrp = BB_CV_THREADS / ZL;
    const int nblk = ((bx + BB_CV_RX - 1) / BB_CV_RX) * ((by + BB_CV_RY - 1) / BB_CV_RY);
    dim3 grid((nblk + ngrp - 1) / ngrp, (nz + ZL - 1) / ZL, B);
    if (frac)
        hipLaunchKernelGGL((k_conv_resid<KX, true>), grid, dim3(BB_CV_THREADS), 0, st, in, dvalid, by, bx, nz, ZL, kern,
                           ky, scale, sigma2, eps2, partial, vbox, sv);
    else
        hipLaunchKernelGGL((k_conv_resid<KX, false>), grid, dim3(BB_CV_THREADS), 0, st, in, dvalid, by, bx, nz, ZL,
                           kern, ky, scale, 1.f, 0.f, partial, vbox, (const float2*)nullptr);
}

// The fused convolution and residual: partial (B x bb_conv_resid_blocks x 2 doubles). vbox (B, box-relative [x, z) x
// [y, w)) restricts the sums to each vector's own model box (NULL: the whole box); sv (B: sigma2, eps2) replaces the
// scalars per vector (NULL: the scalars); frac selects the fracerr form (also when eps2 is 0 for every vector).
extern "C" int bb_launch_conv_resid(const float* in, const float* dvalid, int B, int by, int bx, int nz,
                                    const float* kern, int ky, int kx, float scale, double sigma2, double eps2,
                                    double* partial, const int4* vbox, const float2* sv, int frac, hipStream_t st) {
    const bool fr = frac != 0 || eps2 > 0;
    switch (kx) {
#define BB_CASE(K) \
    case K: launch_cr<K>(in, dvalid, B, by, bx, nz, kern, ky, scale, (float)sigma2, (float)eps2, partial, vbox, sv, fr, \
                         st);                                                                                          \
        break;
        BB_CASE(1) BB_CASE(3) BB_CASE(5) BB_CASE(7) BB_CASE(9) BB_CASE(11) BB_CASE(13) BB_CASE(15) BB_CASE(17)
        BB_CASE(19) BB_CASE(21) BB_CASE(23) BB_CASE(25)
#undef BB_CASE
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

// kern: (ky, kx) float32 taps on the device, kx odd <= 25.
extern "C" int bb_launch_conv2d_same(const float* in, float* out, int B, int by, int bx, int nz, const float* kern,
                                     int ky, int kx, float scale, hipStream_t st) {
    switch (kx) {
#define BB_CASE(K) \
    case K: launch<K>(in, out, B, by, bx, nz, kern, ky, scale, st); break;
        BB_CASE(1) BB_CASE(3) BB_CASE(5) BB_CASE(7) BB_CASE(9) BB_CASE(11) BB_CASE(13) BB_CASE(15) BB_CASE(17)
        BB_CASE(19) BB_CASE(21) BB_CASE(23) BB_CASE(25)
#undef BB_CASE
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------
// The fused convolution-residual without a caller's work buffer: the per-block partial sums live in a file-scope
// device array of BB_CR_SCRATCH doubles (8 MiB; each device that runs it holds its own copy until the process
// ends). Its address on a device comes from hipGetSymbolAddress on the first call with that device current; host
// code passes that address, never the array's name, to the kernels. Every block writes its own partial pair, so
// nothing needs clearing; calls that overlap on one device (two streams, two threads) share the array and race.
#define BB_CR_SCRATCH (1 << 20)
#define BB_CR_MAX_DEVICES 64

static __device__ __align__(256) double scratch[BB_CR_SCRATCH];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[BB_CR_MAX_DEVICES] = {nullptr};

static int scratch_ptr(double** out) {
    int dev = -1;
    int rc = bb_hip_check(hipGetDevice(&dev), "bb_conv_residual_gauss_noalloc: hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= BB_CR_MAX_DEVICES)
        return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: device %d beyond %d", dev, BB_CR_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        rc = bb_hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                          "bb_conv_residual_gauss_noalloc: hipGetSymbolAddress");
        if (rc) return rc;
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int64_t bb_conv_residual_noalloc_max(void) { return BB_CR_SCRATCH; }

extern "C" int bb_conv_residual_gauss_noalloc(const bb_model* m, const float* dep, const float* dvalid, int B, int by,
                                              int bx, double sigma2, double eps2, double* out_ss, double* out_lg,
                                              void* stream) {
    if (!m) return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: model is NULL");
    if (!m->d_kern) return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: the model has no beam kernel");
    if (B < 0 || by < 1 || bx < 1 || B > 65535) return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: bad sizes");
    if (B == 0) return 0;
    if (!dep || !dvalid || !out_ss) return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: NULL array");
    if (eps2 > 0 && (!out_lg || !(sigma2 > 0)))
        return bb_fail(BB_ERR_ARG, "bb_conv_residual_gauss_noalloc: fracerr needs sigma2 > 0 and out_lg");
    const int64_t nblk = bb_conv_resid_blocks(by, bx, m->g.nz);
    const int64_t need = (int64_t)B * nblk * 2;
    if (need > BB_CR_SCRATCH)
        return bb_fail(BB_ERR_ARG,
                       "bb_conv_residual_gauss_noalloc: %d models of %d x %d x %d need %lld partial sums, the scratch "
                       "holds %d (use bb_conv_residual_gauss with a work buffer, or smaller batches)",
                       B, by, bx, m->g.nz, (long long)need, BB_CR_SCRATCH);
    double* work = nullptr;
    int rc = scratch_ptr(&work);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    rc = bb_launch_conv_resid(dep, dvalid, B, by, bx, m->g.nz, m->d_kern, m->ky, m->kx, (float)m->scale, sigma2, eps2,
                              work, nullptr, nullptr, 0, st);
    if (!rc) rc = bb_launch_residual_final(work, B, (int)nblk, out_ss, out_lg, st);
    return rc ? bb_fail(BB_ERR_HIP, "bb_conv_residual_gauss_noalloc launch: %s", hipGetErrorString((hipError_t)rc)) : 0;
}
