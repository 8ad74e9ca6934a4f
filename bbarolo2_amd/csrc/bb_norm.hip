// The 3DFIT residual of ring models on the GPU (_norm, bbarolo/galfit.py:39-188): the pixels of the ring and
// their weights, the normalisation of the model (LOCAL per pixel, AZIM per ring, NONE), the masked residual
// (ftype 1-4) with the blanks (model voxels outside the mask; with the ANALYTIC support, only those where the
// expected number of sub-clouds over the kernel footprint reaches thr, galfit.py:685-723) in one fused pass.
//
// B trial models (B, by, bx, nz) float32 yxz share their box; each trial has its own ring geometry. One
// wavefront per pixel of the box: the lanes take the channels, the LOCAL factor is a wavefront reduction, and
// each pixel's sums (FP64) and counts are stored; one workgroup per trial then adds the pixels in an order fixed
// by their coordinates in the cube (anchored_sum), so a trial gives the same value bit for bit in any box that
// holds its ring (the CLASSIC stage models the trials of all its rings in the union of their boxes). AZIM takes
// a first pass for the ring's factor. The support's footprint sums use per-row prefix sums of the unconvolved
// model along x (FP64), as the reference does.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/bbgpu.h"
#include "bb_internal.h"

#define BB_NM_THREADS 256
#define BB_NM_WAVES (BB_NM_THREADS / 64)

struct NormGeo {
    int by, bx, nz, ny, nx;
    int64_t blo0, blo1;
    double wpow, data_noise, bweight, hmax;
    int side, ftype, normtype, foot, ky;
};

// inside the ring? weight (galfit.py:58-86); rp: x0, y0, phi, inc, r1, r2 of the trial (box-relative pixels)
__device__ __forceinline__ bool nm_inside(const NormGeo& G, const double* rp, int x, int y, double& w) {
    const double F = 3.141592653589793 / 180.;
    const double x0 = rp[0], y0 = rp[1], phi = rp[2], inc = rp[3], r1 = rp[4], r2 = rp[5];
    const double sphi = sin(F * phi), cphi = cos(F * phi), cinc = cos(F * inc);
    const double xr = -(x - x0) * sphi + (y - y0) * cphi;
    const double ys = -(x - x0) * cphi - (y - y0) * sphi;
    const double yr = ys / cinc;
    const double r = sqrt(xr * xr + yr * yr);
    double th = r < 0.1 ? 0.0 : atan2(yr, xr) / F;
    if (G.foot == 1) {
        if (!(fabs(xr) >= r1 && fabs(xr) <= r2 && fabs(ys) <= G.hmax)) return false;
        th = xr >= 0 ? 0.0 : 180.0;
    } else if (!(r >= r1 && r <= r2)) {
        return false;
    }
    if (G.side == 1 && !(fabs(th) <= 90.0)) return false;
    if (G.side == 2 && !(fabs(th) >= 90.0)) return false;
    if (G.foot == 1) {
        w = 1.0;
    } else {
        const double wf = G.wpow < 0 ? fabs(sin(th * F)) : fabs(cos(th * F));
        w = pow(wf, fabs(G.wpow));
    }
    return true;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// the sum over the pixels of one trial of K per-pixel values pix (npix, K), by the workgroup (BB_NM_THREADS) of
// the trial, in an order fixed by the pixels' coordinates in the cube: lane l of a wavefront adds the columns
// with (x + blo0) % 64 == l of a row, wavefront w the rows with (y + blo1) % 4 == w, the row sums going to the
// lane (y + blo1) / 4 % 64 in increasing y; then wavefront sums and the 4 wavefronts in order. A larger box
// around the same non-zero pixels only adds exact zeros (every value is >= 0), so the sum is unchanged.
template <int K>
__device__ void anchored_sum(const NormGeo& G, const double* __restrict__ pix, double* out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int xr = (int)(G.blo0 & 63), yr = (int)(G.blo1 & 3);
    double acc[K];
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    for (int y = (wv - yr + 4) & 3; y < G.by; y += BB_NM_WAVES) {
        double rs[K];
        for (int k = 0; k < K; k++) rs[k] = 0.0;
        const double* row = pix + (size_t)y * G.bx * K;
        for (int x = (lane - xr + 64) & 63; x < G.bx; x += 64)
            for (int k = 0; k < K; k++) rs[k] += row[(size_t)x * K + k];
        for (int k = 0; k < K; k++) rs[k] = wave_sum(rs[k]);
        if (lane == (int)(((y + G.blo1) >> 2) & 63))
            for (int k = 0; k < K; k++) acc[k] += rs[k];
    }
    __shared__ double red[BB_NM_WAVES][K];
    for (int k = 0; k < K; k++) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
        for (int k = 0; k < K; k++) red[wv][k] = acc[k];
    __syncthreads();
    for (int k = 0; k < K; k++) {
        double t = 0.0;
        for (int i = 0; i < BB_NM_WAVES; i++) t += red[i][k];
        out[k] = t;
    }
}

// AZIM: per pixel of the ring, the sums of the model and of the masked data over the channels (0 elsewhere)
__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_azim(NormGeo G, const float* __restrict__ model, const float* __restrict__ data,
            const uint8_t* __restrict__ mask, const double* __restrict__ ring, double* __restrict__ pix) {
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npix = G.by * G.bx;
    const double* rp = ring + (size_t)b * 6;
    const float* mb = model + (size_t)b * npix * G.nz;
    for (int p = blockIdx.x * BB_NM_WAVES + wv; p < npix; p += gridDim.x * BB_NM_WAVES) {
        const int y = p / G.bx, x = p - y * G.bx;
        double w, ms = 0.0, os = 0.0;
        if (nm_inside(G, rp, x, y, w)) {
            const size_t d0 = ((size_t)(y + G.blo1) * G.nx + (x + G.blo0)) * G.nz;
            for (int z = lane; z < G.nz; z += 64) {
                ms += (double)mb[(size_t)p * G.nz + z];
                if (mask[d0 + z]) os += (double)data[d0 + z];
            }
            ms = wave_sum(ms);
            os = wave_sum(os);
        }
        if (lane == 0) {
            pix[((size_t)b * npix + p) * 2 + 0] = ms;
            pix[((size_t)b * npix + p) * 2 + 1] = os;
        }
    }
}

__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_azim_final(NormGeo G, const double* __restrict__ pix, double* __restrict__ factor) {
    const int b = blockIdx.x;
    double s[2];
    anchored_sum<2>(G, pix + (size_t)b * G.by * G.bx * 2, s);
    if (threadIdx.x == 0) factor[b] = s[0] != 0 ? s[1] / s[0] : 0.0;
}

// prefix sums of the unconvolved model along x, per (trial, row, channel): P (B, by, bx+1, nz) FP64
__global__ void k_norm_prefix(NormGeo G, int B, const float* __restrict__ pre, double* __restrict__ P) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nl = (int64_t)B * G.by * G.nz;
    if (t >= nl) return;
    const int z = (int)(t % G.nz);
    const int64_t by_ = t / G.nz;               // b * by + y
    const float* src = pre + (size_t)by_ * G.bx * G.nz + z;
    double* dst = P + (size_t)by_ * (G.bx + 1) * G.nz + z;
    double acc = 0.0;
    dst[0] = 0.0;
    for (int x = 0; x < G.bx; x++) {
        acc += (double)src[(size_t)x * G.nz];
        dst[(size_t)(x + 1) * G.nz] = acc;
    }
}

// the residual: per pixel of the box (sum, blanks, voxels), zeros outside the ring
__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_res(NormGeo G, const float* __restrict__ model, const float* __restrict__ data, const uint8_t* __restrict__ mask,
           const double* __restrict__ ring, const double* __restrict__ chan_noise, const double* __restrict__ azf,
           const double* __restrict__ P, const int* __restrict__ ilo, const int* __restrict__ ihi,
           const double* __restrict__ thr, double* __restrict__ pix) {
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npix = G.by * G.bx, nz = G.nz;
    const double* rp = ring + (size_t)b * 6;
    const float* mb = model + (size_t)b * npix * nz;
    const double* Pb = P ? P + (size_t)b * G.by * (G.bx + 1) * nz : nullptr;
    const int hy = G.ky / 2;
    for (int p = blockIdx.x * BB_NM_WAVES + wv; p < npix; p += gridDim.x * BB_NM_WAVES) {
        const int y = p / G.bx, x = p - y * G.bx;
        double* o = pix + ((size_t)b * npix + p) * 3;
        double wi, sum = 0.0, nb = 0.0, nt = 0.0;
        if (!nm_inside(G, rp, x, y, wi)) {
            if (lane == 0) o[0] = o[1] = o[2] = 0.0;
            continue;
        }
        const size_t d0 = ((size_t)(y + G.blo1) * G.nx + (x + G.blo0)) * nz;
        double factor = 1.0;
        if (G.normtype == 0) {
            double ms = 0.0, os = 0.0;
            for (int z = lane; z < nz; z += 64) {
                ms += (double)mb[(size_t)p * nz + z];
                const double o = (double)data[d0 + z];
                if (o > 0 && mask[d0 + z]) os += o;
            }
            ms = wave_sum(ms);
            os = wave_sum(os);
            factor = ms != 0 ? os / ms : 0.0;
        } else if (G.normtype == 1) {
            factor = azf[b];
        }
        for (int z = lane; z < nz; z += 64) {
            double mod = (double)mb[(size_t)p * nz + z];
            if (G.normtype != 2) mod *= factor;
            const double o = (double)data[d0 + z];
            const double obs = o > 0 ? o : G.data_noise;
            if (!mask[d0 + z]) {
                if (mod == 0) continue;
                if (Pb) {
                    double lam = 0.0;
                    for (int j = 0; j < G.ky; j++) {
                        const int yy = y + hy - j;
                        if (yy < 0 || yy >= G.by) continue;
                        int xa = x + ilo[j], xb = x + ihi[j] + 1;
                        xa = xa < 0 ? 0 : xa;
                        xb = xb > G.bx ? G.bx : xb;
                        if (xb <= xa) continue;
                        const double* Pr = Pb + (size_t)yy * (G.bx + 1) * nz;
                        lam += Pr[(size_t)xb * nz + z] - Pr[(size_t)xa * nz + z];
                    }
                    if (lam < thr[z]) continue;
                }
                nb += 1.0;
            }
            nt += 1.0;
            const double nw = chan_noise[z];
            double t;
            if (G.ftype == 1) t = wi * (mod - obs) * (mod - obs) / sqrt(obs) / nw;
            else if (G.ftype == 2) t = wi * fabs(mod - obs) / nw;
            else if (G.ftype == 3) t = wi * fabs(mod - obs) / (mod + obs) / nw;
            else if (G.ftype == 4) t = wi * (mod - obs) * (mod - obs) / nw;
            else t = 0.0;
            sum += t;
        }
        sum = wave_sum(sum);
        nb = wave_sum(nb);
        nt = wave_sum(nt);
        if (lane == 0) {
            o[0] = sum;
            o[1] = nb;
            o[2] = nt;
        }
    }
}

__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_final(NormGeo G, const double* __restrict__ pix, double* __restrict__ val, int64_t* __restrict__ nbo,
             int64_t* __restrict__ nto) {
    const int b = blockIdx.x;
    double s[3];
    anchored_sum<3>(G, pix + (size_t)b * G.by * G.bx * 3, s);
    if (threadIdx.x != 0) return;
    const double nb = s[1], nt = s[2];
    if (nbo) nbo[b] = (int64_t)nb;
    if (nto) nto[b] = (int64_t)nt;
    val[b] = nt - nb == 0 ? __builtin_inf() : pow(1.0 + nb / nt, G.bweight) * s[0] / (nt - nb);
}

// ---------------------------------------------------------------------------------------------------------
extern "C" int bb_norm_blocks(int by, int bx) {
    const int npix = by * bx;
    int n = (npix + 8 * BB_NM_WAVES - 1) / (8 * BB_NM_WAVES);     // (about 8 pixels per wavefront)
    return n < 1 ? 1 : (n > 1024 ? 1024 : n);
}

extern "C" int bb_launch_norm(int B, int by, int bx, int nz, int ny, int nx, int64_t blo0, int64_t blo1, double wpow,
                              int side, int ftype, double data_noise, double bweight, int normtype, int foot,
                              double hmax, const float* model, const float* data, const uint8_t* mask,
                              const double* ring, const double* chan_noise, const float* pre, const int* ilo,
                              const int* ihi, int ky, const double* thr, double* work, double* val, int64_t* nb,
                              int64_t* nt, hipStream_t st) {
    NormGeo G;
    G.by = by;
    G.bx = bx;
    G.nz = nz;
    G.ny = ny;
    G.nx = nx;
    G.blo0 = blo0;
    G.blo1 = blo1;
    G.wpow = wpow;
    G.data_noise = data_noise;
    G.bweight = bweight;
    G.hmax = hmax;
    G.side = side;
    G.ftype = ftype;
    G.normtype = normtype;
    G.foot = foot;
    G.ky = pre ? ky : 0;
    const int nblk = bb_norm_blocks(by, bx);
    double* pix = work;                                          // B * by * bx * 3
    double* azf = pix + (size_t)B * by * bx * 3;                 // B
    double* P = azf + B;                                         // B * by * (bx+1) * nz (with a support)
    if (normtype == 1) {
        hipLaunchKernelGGL(k_norm_azim, dim3(nblk, B), dim3(BB_NM_THREADS), 0, st, G, model, data, mask, ring, pix);
        hipLaunchKernelGGL(k_norm_azim_final, dim3(B), dim3(BB_NM_THREADS), 0, st, G, (const double*)pix, azf);
    }
    if (pre) {
        const int64_t nl = (int64_t)B * by * nz;
        hipLaunchKernelGGL(k_norm_prefix, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, G, B, pre, P);
    }
    hipLaunchKernelGGL(k_norm_res, dim3(nblk, B), dim3(BB_NM_THREADS), 0, st, G, model, data, mask, ring, chan_noise,
                       azf, pre ? (const double*)P : (const double*)nullptr, ilo, ihi, thr, pix);
    hipLaunchKernelGGL(k_norm_final, dim3(B), dim3(BB_NM_THREADS), 0, st, G, (const double*)pix, val, nb, nt);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------
// LOCAL normalisation of a final model (galfit_out.py _local, bbarolo/galfit_out.py:128-135): every spectrum of the
// model scaled by (sum over the channels of the data in the mask) / (sum of the model), 0 where the model sums to 0.
// One wavefront per pixel, lanes along the channels (yxz layout), FP64 sums.
__global__ void __launch_bounds__(256)
k_local_norm(const float* __restrict__ model, const float* __restrict__ data, const uint8_t* __restrict__ mask,
             int64_t npix, int nz, float* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= npix) return;
    const float* m = model + p * nz;
    const float* d = data + p * nz;
    const uint8_t* k = mask + p * nz;
    double ms = 0., os = 0.;
    for (int z = lane; z < nz; z += 64) {
        ms += (double)m[z];
        if (k[z]) os += (double)d[z];
    }
    for (int off = 32; off > 0; off >>= 1) {
        ms += __shfl_xor(ms, off, 64);
        os += __shfl_xor(os, off, 64);
    }
    const double f = ms != 0. ? os / ms : 0.;
    for (int z = lane; z < nz; z += 64) out[p * nz + z] = (float)((double)m[z] * f);
}

extern "C" int bb_local_norm(const float* model, const float* data, const uint8_t* mask, int by, int bx, int nz,
                             float* out, void* stream) {
    if (by < 0 || bx < 0 || nz < 1) return bb_fail(BB_ERR_ARG, "bb_local_norm: bad sizes");
    const int64_t npix = (int64_t)by * bx;
    if (npix == 0) return 0;
    if (!model || !data || !mask || !out) return bb_fail(BB_ERR_ARG, "bb_local_norm: NULL array");
    const int64_t blocks = (npix + 3) / 4;
    if (blocks > INT32_MAX) return bb_fail(BB_ERR_ARG, "bb_local_norm: too many pixels");
    hipLaunchKernelGGL(k_local_norm, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, model, data, mask, npix,
                       nz, out);
    return bb_hip_check(hipGetLastError(), "bb_local_norm");
}
