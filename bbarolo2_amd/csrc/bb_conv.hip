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
