// The likelihood mask of Bayesian3DFIT on the GPU: the emission mask dilated by a disk of a given radius in each
// channel, then by +-nchan channels (dilate_mask, bbarolo/bayesian/noise.py:18-30, called by _setup_likemask,
// bbarolo/bayesian/likelihood.py:130-200).
//
// The reference takes, in every channel that holds mask pixels, the Euclidean distance transform of the
// complement and keeps the pixels within the radius: a pixel is in the dilated channel when some mask pixel of
// that channel is at a distance sqrt(dx^2 + dy^2) <= radius. k_dilate_disk forms the same set directly: one
// workgroup per (channel, 32 x 32 pixel tile) stages the tile's mask with a halo of R = floor(radius) pixels in
// LDS as per-row prefix counts, and a pixel is set when, for some row offset dy, the row y + dy holds a mask pixel
// within +-W(dy) columns (W(dy) the largest dx with sqrt(dx^2 + dy^2) <= radius, evaluated in FP64 as the distance
// transform's distances are compared): O(2R + 1) LDS reads per pixel. k_dilate_z is the maximum filter of
// 2 nchan + 1 channels with zeros outside the cube (maximum_filter1d, mode 'constant').
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define BB_DL_TILE 32
#define BB_DL_THREADS 256

extern "C" int bb_dilate_rmax() { return 56; }

__global__ void __launch_bounds__(BB_DL_THREADS)
k_dilate_disk(const uint8_t* __restrict__ m, int ny, int nx, double radius, int R, uint8_t* __restrict__ out) {
    extern __shared__ int dl_lds[];
    const int k = blockIdx.z;
    const int x0 = blockIdx.x * BB_DL_TILE, y0 = blockIdx.y * BB_DL_TILE;
    const int W = BB_DL_TILE + 2 * R;              // staged rows and columns
    const int P = W + 1;                           // prefix entries per row
    int* pre = dl_lds;                             // (W, P): pre[r][c] = mask pixels in columns [0, c) of row r
    int* wdy = pre + W * P;                        // (2R + 1): W(dy), -1 where no dx qualifies
    uint8_t* stage = (uint8_t*)(wdy + 2 * R + 1);  // (W, W) the staged mask
    const uint8_t* mk = m + (size_t)k * ny * nx;
    for (int i = threadIdx.x; i < W * W; i += blockDim.x) {
        const int r = i / W, c = i - r * W;
        const int y = y0 - R + r, x = x0 - R + c;
        stage[i] = (y >= 0 && y < ny && x >= 0 && x < nx) ? (mk[(size_t)y * nx + x] != 0) : 0;
    }
    for (int d = threadIdx.x; d <= 2 * R; d += blockDim.x) {
        const int dy = d - R;
        int w = -1;
        for (int dx = 0; dx <= R; dx++)
            if (sqrt((double)(dx * dx + dy * dy)) <= radius) w = dx;
        wdy[d] = w;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < W; r += blockDim.x) {
        int s = 0;
        pre[r * P] = 0;
        for (int c = 0; c < W; c++) {
            s += stage[r * W + c];
            pre[r * P + c + 1] = s;
        }
    }
    __syncthreads();
    uint8_t* ok = out + (size_t)k * ny * nx;
    for (int i = threadIdx.x; i < BB_DL_TILE * BB_DL_TILE; i += blockDim.x) {
        const int ty = i / BB_DL_TILE, tx = i - ty * BB_DL_TILE;
        const int y = y0 + ty, x = x0 + tx;
        if (y >= ny || x >= nx) continue;
        uint8_t hit = 0;
        for (int d = 0; d <= 2 * R && !hit; d++) {
            const int w = wdy[d];
            if (w < 0) continue;
            const int* row = pre + (ty + d) * P;       // staged row of y + (d - R)
            const int c = tx + R;                      // staged column of x
            hit = row[c + w + 1] - row[c - w] > 0;
        }
        ok[(size_t)y * nx + x] = hit;
    }
}

__global__ void k_dilate_z(const uint8_t* __restrict__ in, int nz, int64_t npix, int nchan, uint8_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (p >= npix) return;
    uint8_t v = 0;
    const int a = k - nchan < 0 ? 0 : k - nchan, b = k + nchan > nz - 1 ? nz - 1 : k + nchan;
    for (int j = a; j <= b && !v; j++) v = in[(size_t)j * npix + p] != 0;
    out[(size_t)k * npix + p] = v;
}

__global__ void k_mask_copy(const uint8_t* __restrict__ in, int64_t n, uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] != 0;
}

extern "C" size_t bb_dilate_lds(int R) {
    const size_t W = BB_DL_TILE + 2 * (size_t)R;
    return W * (W + 1) * sizeof(int) + (2 * (size_t)R + 1) * sizeof(int) + W * W;
}

// mask (nz, ny, nx) u8 (nonzero: in the mask) -> out (0 / 1); tmp (nz, ny, nx) u8 holds the disk stage when both
// stages run. The caller checks floor(radius) <= bb_dilate_rmax() and nz <= 65535.
extern "C" int bb_launch_dilate(const uint8_t* mask, int nz, int ny, int nx, double radius, int nchan, uint8_t* tmp,
                                uint8_t* out, hipStream_t st) {
    const int64_t npix = (int64_t)ny * nx, n = npix * nz;
    if (n == 0) return 0;
    const bool disk = radius >= 1.0, zed = nchan > 0;      // (a radius below one pixel adds no pixel)
    const uint8_t* zin = mask;
    if (disk) {
        uint8_t* d1 = zed ? tmp : out;
        const int R = (int)floor(radius);
        const size_t lds = bb_dilate_lds(R);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void*)k_dilate_disk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
            if (e != hipSuccess) return (int)e;
        }
        hipLaunchKernelGGL(k_dilate_disk, dim3((nx + BB_DL_TILE - 1) / BB_DL_TILE, (ny + BB_DL_TILE - 1) / BB_DL_TILE, nz),
                           dim3(BB_DL_THREADS), lds, st, mask, ny, nx, radius, R, d1);
        zin = d1;
    } else if (!zed) {
        hipLaunchKernelGGL(k_mask_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, mask, n, out);
    }
    if (zed)
        hipLaunchKernelGGL(k_dilate_z, dim3((unsigned)((npix + 255) / 256), nz), dim3(256), 0, st, zin, nz, npix,
                           nchan, out);
    return (int)hipGetLastError();
}
