// Masked Gaussian residual of model cubes against the data on the GPU (_sumsq_residuals and
// _residuals_fracerr, bbarolo/bayesian/likelihood.py:22-77).
//
// The data enter as one float32 array "dvalid" in the layout of the models (yxz): the data of the voxels
// that the reference sums (in the region, in the likelihood mask, not blank), NaN for every other voxel,
// prepared once per fit. A residual pass therefore streams 8 bytes per voxel (model + data). Each lane
// forms its terms in FP32 and accumulates them in FP64; blocks write FP64 partial sums, reduced by a second
// kernel in a fixed order (deterministic results).
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BB_RS_THREADS 256
#define BB_RS_PER_THREAD 8

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <bool FRAC>
__global__ void __launch_bounds__(BB_RS_THREADS)
k_residual(const float* __restrict__ model, const float* __restrict__ dvalid, int64_t nvox, float sigma2f,
           float eps2f, double* __restrict__ partial, const int4* __restrict__ vbox, const float2* __restrict__ sv,
           int bx, int nz) {
    __shared__ double red[2][BB_RS_THREADS / 64];
    const int b = blockIdx.y;
    const float* m = model + (size_t)b * nvox;
    if (sv) {
        sigma2f = sv[b].x;
        eps2f = sv[b].y;
    }
    int4 vb = make_int4(0, 0, 0, 0);
    if (vbox) vb = vbox[b];
    double s = 0., lg = 0.;
    const int64_t base = (int64_t)blockIdx.x * BB_RS_THREADS * BB_RS_PER_THREAD;
#pragma unroll
    for (int k = 0; k < BB_RS_PER_THREAD; k++) {
        int64_t i = base + (int64_t)k * BB_RS_THREADS + threadIdx.x;
        if (i < nvox) {
            float d = dvalid[i];
            if (vbox) {
                // only the voxels of the vector's own model box (box-relative [x, z) x [y, w), yxz layout)
                const int64_t p = i / nz;
                const int x = (int)(p % bx), y = (int)(p / bx);
                if (x < vb.x || x >= vb.z || y < vb.y || y >= vb.w) d = __builtin_nanf("");
            }
            if (d == d) {
                float mm = m[i];
                float r = d - mm;
                if (!FRAC) {
                    s += (double)(r * r);
                } else if (mm != 0.f) {
                    float v = fmaf(eps2f * mm, mm, sigma2f);
                    s += (double)(r * r / v);
                    lg += (double)logf(v / sigma2f);
                } else {
                    s += (double)(r * r / sigma2f);
                }
            }
        }
    }
    s = wave_sum(s);
    if (FRAC) lg = wave_sum(lg);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = s;
        red[1][w] = lg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0., c = 0.;
        for (int k = 0; k < BB_RS_THREADS / 64; k++) {
            a += red[0][k];
            c += red[1][k];
        }
        size_t o = ((size_t)b * gridDim.x + blockIdx.x) * 2;
        partial[o] = a;
        partial[o + 1] = c;
    }
}

// One block per parameter vector: sum its nblk partial pairs in a fixed order.
__global__ void __launch_bounds__(BB_RS_THREADS)
k_residual_final(const double* __restrict__ partial, int nblk, double* __restrict__ out_ss,
                 double* __restrict__ out_lg) {
    __shared__ double red[2][BB_RS_THREADS / 64];
    const int b = blockIdx.x;
    double s = 0., lg = 0.;
    for (int i = threadIdx.x; i < nblk; i += BB_RS_THREADS) {
        s += partial[((size_t)b * nblk + i) * 2];
        lg += partial[((size_t)b * nblk + i) * 2 + 1];
    }
    s = wave_sum(s);
    lg = wave_sum(lg);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][w] = s;
        red[1][w] = lg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0., c = 0.;
        for (int k = 0; k < BB_RS_THREADS / 64; k++) {
            a += red[0][k];
            c += red[1][k];
        }
        out_ss[b] = a;
        if (out_lg) out_lg[b] = c;
    }
}

extern "C" int bb_launch_residual_final(const double* partial, int B, int nblk, double* out_ss, double* out_lg,
                                        hipStream_t st) {
    hipLaunchKernelGGL(k_residual_final, dim3(B), dim3(BB_RS_THREADS), 0, st, partial, nblk, out_ss, out_lg);
    return (int)hipGetLastError();
}

extern "C" int bb_residual_blocks(int64_t nvox) {
    const int64_t per = (int64_t)BB_RS_THREADS * BB_RS_PER_THREAD;
    return (int)((nvox + per - 1) / per);
}

// vbox / sv / frac as bb_launch_conv_resid (bx, nz: the box width and the channels of the yxz layout, for vbox).
extern "C" int bb_launch_residual(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2,
                                  double eps2, double* partial, double* out_ss, double* out_lg, const int4* vbox,
                                  const float2* sv, int frac, int bx, int nz, hipStream_t st) {
    const int nblk = bb_residual_blocks(nvox);
    dim3 grid(nblk, B);
    if (eps2 > 0 || frac)
        hipLaunchKernelGGL(k_residual<true>, grid, dim3(BB_RS_THREADS), 0, st, model, dvalid, nvox, (float)sigma2,
                           (float)eps2, partial, vbox, sv, bx, nz);
    else
        hipLaunchKernelGGL(k_residual<false>, grid, dim3(BB_RS_THREADS), 0, st, model, dvalid, nvox, 1.f, 0.f,
                           partial, vbox, (const float2*)nullptr, bx, nz);
    hipLaunchKernelGGL(k_residual_final, dim3(B), dim3(BB_RS_THREADS), 0, st, partial, nblk, out_ss, out_lg);
    return (int)hipGetLastError();
}
