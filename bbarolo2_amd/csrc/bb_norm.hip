// The 3DFIT residual of ring models on the GPU (_norm, bbarolo/galfit.py:39-188): the pixels of the ring and
// their weights, the normalisation of the model (LOCAL per pixel, AZIM per ring, NONE), the masked residual
// (ftype 1-4) with the blanks (model voxels outside the mask; with the ANALYTIC support, only those where the
// expected number of sub-clouds over the kernel footprint reaches thr, galfit.py:685-723) in one fused pass.
//
// B trial models (B, by, bx, nz) float32 yxz share their box; each trial has its own ring geometry. One
// wavefront per pixel of the box: the lanes take the channels, the LOCAL factor is a wavefront reduction, and
// the sums (FP64) and counts go to per-block partials reduced in a fixed order (deterministic). AZIM takes a
// first pass for the ring's factor. The support's footprint sums use per-row prefix sums of the unconvolved
// model along x (FP64), as the reference does.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

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

// AZIM: per-block partial sums of the model and of the masked data over the ring's pixels
__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_azim(NormGeo G, const float* __restrict__ model, const float* __restrict__ data,
            const uint8_t* __restrict__ mask, const double* __restrict__ ring, double* __restrict__ part) {
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npix = G.by * G.bx;
    const double* rp = ring + (size_t)b * 6;
    const float* mb = model + (size_t)b * npix * G.nz;
    double ms = 0.0, os = 0.0;
    for (int p = blockIdx.x * BB_NM_WAVES + wv; p < npix; p += gridDim.x * BB_NM_WAVES) {
        const int y = p / G.bx, x = p - y * G.bx;
        double w;
        if (!nm_inside(G, rp, x, y, w)) continue;
        const size_t d0 = ((size_t)(y + G.blo1) * G.nx + (x + G.blo0)) * G.nz;
        for (int z = lane; z < G.nz; z += 64) {
            ms += (double)mb[(size_t)p * G.nz + z];
            if (mask[d0 + z]) os += (double)data[d0 + z];
        }
    }
    __shared__ double red[2][BB_NM_WAVES];
    ms = wave_sum(ms);
    os = wave_sum(os);
    if (lane == 0) {
        red[0][wv] = ms;
        red[1][wv] = os;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, c = 0;
        for (int i = 0; i < BB_NM_WAVES; i++) {
            a += red[0][i];
            c += red[1][i];
        }
        part[((size_t)b * gridDim.x + blockIdx.x) * 2 + 0] = a;
        part[((size_t)b * gridDim.x + blockIdx.x) * 2 + 1] = c;
    }
}

__global__ void k_norm_azim_final(const double* __restrict__ part, int nblk, double* __restrict__ factor) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (int)gridDim.x * (int)blockDim.x) return;
    double a = 0, c = 0;
    for (int k = 0; k < nblk; k++) {
        a += part[((size_t)b * nblk + k) * 2 + 0];
        c += part[((size_t)b * nblk + k) * 2 + 1];
    }
    factor[b] = a != 0 ? c / a : 0.0;
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

// the residual: per-block partials (sum, blanks, voxels)
__global__ void __launch_bounds__(BB_NM_THREADS)
k_norm_res(NormGeo G, const float* __restrict__ model, const float* __restrict__ data, const uint8_t* __restrict__ mask,
           const double* __restrict__ ring, const double* __restrict__ chan_noise, const double* __restrict__ azf,
           const double* __restrict__ P, const int* __restrict__ ilo, const int* __restrict__ ihi,
           const double* __restrict__ thr, double* __restrict__ part) {
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npix = G.by * G.bx, nz = G.nz;
    const double* rp = ring + (size_t)b * 6;
    const float* mb = model + (size_t)b * npix * nz;
    const double* Pb = P ? P + (size_t)b * G.by * (G.bx + 1) * nz : nullptr;
    const int hy = G.ky / 2;
    double sum = 0.0, nb = 0.0, nt = 0.0;
    for (int p = blockIdx.x * BB_NM_WAVES + wv; p < npix; p += gridDim.x * BB_NM_WAVES) {
        const int y = p / G.bx, x = p - y * G.bx;
        double wi;
        if (!nm_inside(G, rp, x, y, wi)) continue;
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
    }
    __shared__ double red[3][BB_NM_WAVES];
    sum = wave_sum(sum);
    nb = wave_sum(nb);
    nt = wave_sum(nt);
    if (lane == 0) {
        red[0][wv] = sum;
        red[1][wv] = nb;
        red[2][wv] = nt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, c = 0, d = 0;
        for (int i = 0; i < BB_NM_WAVES; i++) {
            a += red[0][i];
            c += red[1][i];
            d += red[2][i];
        }
        double* o = part + ((size_t)b * gridDim.x + blockIdx.x) * 3;
        o[0] = a;
        o[1] = c;
        o[2] = d;
    }
}

__global__ void k_norm_final(const double* __restrict__ part, int nblk, int B, double bweight, double* __restrict__ val,
                             int64_t* __restrict__ nbo, int64_t* __restrict__ nto) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0, nb = 0, nt = 0;
    for (int k = 0; k < nblk; k++) {
        const double* p = part + ((size_t)b * nblk + k) * 3;
        s += p[0];
        nb += p[1];
        nt += p[2];
    }
    if (nbo) nbo[b] = (int64_t)nb;
    if (nto) nto[b] = (int64_t)nt;
    val[b] = nt - nb == 0 ? __builtin_inf() : pow(1.0 + nb / nt, bweight) * s / (nt - nb);
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
    double* part = work;                                   // B * nblk * 3
    double* azf = part + (size_t)B * nblk * 3;             // B
    double* P = azf + B;                                   // B * by * (bx+1) * nz (with a support)
    if (normtype == 1) {
        hipLaunchKernelGGL(k_norm_azim, dim3(nblk, B), dim3(BB_NM_THREADS), 0, st, G, model, data, mask, ring, part);
        hipLaunchKernelGGL(k_norm_azim_final, dim3(B), dim3(1), 0, st, part, nblk, azf);
    }
    if (pre) {
        const int64_t nl = (int64_t)B * by * nz;
        hipLaunchKernelGGL(k_norm_prefix, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, G, B, pre, P);
    }
    hipLaunchKernelGGL(k_norm_res, dim3(nblk, B), dim3(BB_NM_THREADS), 0, st, G, model, data, mask, ring, chan_noise,
                       azf, pre ? (const double*)P : (const double*)nullptr, ilo, ihi, thr, part);
    hipLaunchKernelGGL(k_norm_final, dim3((B + 63) / 64), dim3(64), 0, st, part, nblk, B, bweight, val, nb, nt);
    return (int)hipGetLastError();
}
