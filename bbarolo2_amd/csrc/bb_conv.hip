// Beam convolution of model cubes on the GPU: 'same' 2D convolution of every channel with the beam
// kernel, then values below 1e-12 set to zero and the flux scale applied (FFTConvolver and
// DiskModel.convolve, bbarolo/smooth.py:145-165 and galmod.py:1273-1283; _conv_sparse galmod.py:974-1037).
//
// Layout yxz (spectra contiguous), so a wavefront's 64 lanes read 64 consecutive channels of one pixel
// (256-byte coalesced loads). Each thread computes BB_CV_RX consecutive output pixels of one row for one
// channel, sliding along the kernel row, so every input value loaded serves up to RX outputs. The kernel
// taps live in LDS as float32; rows are restricted to their non-zero span (the beam is cut below 1e-4 of
// its peak, an ellipse inside the (ky, kx) box).
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BB_CV_ZL 64      // channels per block (lanes)
#define BB_CV_PY 4       // output pixel groups per block (waves)
#define BB_CV_RX 8       // consecutive output pixels per thread
#define BB_CV_KMAX 4096  // kernel taps held in LDS

__global__ void __launch_bounds__(BB_CV_ZL * BB_CV_PY)
k_conv2d_same(const float* __restrict__ inall, float* __restrict__ outall, int by, int bx, int nz,
              const float* __restrict__ kern, const int* __restrict__ rowspan, int ky, int kx, float scale) {
    __shared__ float sk[BB_CV_KMAX];
    __shared__ int sspan[2 * 512];
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int i = tid; i < ky * kx; i += nt) sk[i] = kern[i];
    for (int i = tid; i < 2 * ky; i += nt) sspan[i] = rowspan[i];
    __syncthreads();
    const int b = blockIdx.z;
    const int z = blockIdx.y * BB_CV_ZL + threadIdx.x;
    const int nxg = (bx + BB_CV_RX - 1) / BB_CV_RX;
    const int grp = blockIdx.x * BB_CV_PY + threadIdx.y;
    if (z >= nz || grp >= by * nxg) return;
    const int y = grp / nxg, x0 = (grp % nxg) * BB_CV_RX;
    const float* in = inall + (size_t)b * by * bx * nz;
    float* out = outall + (size_t)b * by * bx * nz;
    const int hy = ky / 2, hx = kx / 2;
    float acc[BB_CV_RX];
#pragma unroll
    for (int r = 0; r < BB_CV_RX; r++) acc[r] = 0.f;
    for (int j = 0; j < ky; j++) {
        const int yy = y + hy - j;         // input row feeding output row y through kernel row j
        if (yy < 0 || yy >= by) continue;
        const int ilo = sspan[2 * j], ihi = sspan[2 * j + 1];
        if (ihi < ilo) continue;
        const float* row = in + (size_t)yy * bx * nz + z;
        const float* kr = sk + j * kx;
        // output x0 + r uses input column x0 + r + hx - i, i in [ilo, ihi]
        const int clo = x0 + hx - ihi, chi = x0 + BB_CV_RX - 1 + hx - ilo;
        for (int c = clo; c <= chi; c++) {
            if (c < 0 || c >= bx) continue;
            const float v = row[(size_t)c * nz];
#pragma unroll
            for (int r = 0; r < BB_CV_RX; r++) {
                const int i = x0 + r + hx - c;
                if (i >= ilo && i <= ihi) acc[r] = fmaf(kr[i], v, acc[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < BB_CV_RX; r++) {
        const int x = x0 + r;
        if (x >= bx) break;
        float v = acc[r];
        v = v < 1e-12f ? 0.f : v * scale;
        out[((size_t)y * bx + x) * nz + z] = v;
    }
}

extern "C" int bb_launch_conv2d_same(const float* in, float* out, int B, int by, int bx, int nz, const float* kern,
                                     const int* rowspan, int ky, int kx, float scale, hipStream_t st) {
    const int nxg = (bx + BB_CV_RX - 1) / BB_CV_RX;
    dim3 grid((by * nxg + BB_CV_PY - 1) / BB_CV_PY, (nz + BB_CV_ZL - 1) / BB_CV_ZL, B);
    hipLaunchKernelGGL(k_conv2d_same, grid, dim3(BB_CV_ZL, BB_CV_PY), 0, st, in, out, by, bx, nz, kern, rowspan, ky,
                       kx, scale);
    return (int)hipGetLastError();
}
