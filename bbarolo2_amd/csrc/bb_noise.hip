// The noise of the cube on the GPU, once per fit: the statistics of Bayesian3DFIT._setup_noise and noise_correlation
// (bbarolo/bayesian/likelihood.py:92-128, bbarolo/bayesian/noise.py:8-186). Cubes are (nz, ny, nx) float32, zyx.
//
// * bb_noise_flags: the voxels used (finite data, and non-zero when zeros are blanks, outside the given emission
//   mask); the mask and used voxels per channel; per pixel whether a channel holds emission and how many used voxels
//   it has. One thread per pixel walks its spectrum (coalesced across the pixels of a row); the per-channel counts
//   are wave ballots added once per wave.
// * bb_masked_median: the median of the selected float32 values (or of |x - c|) as numpy.median gives it: the two
//   middle order statistics by a radix selection on the order-preserving uint32 keys of the floats (four passes of
//   8 bits, two 256-bin histograms each, the bins picked on the device), their float32 mean. Exact: no sampling, no
//   tolerance.
// * bb_masked_std: the standard deviation of the selected values (FP64 two-pass), the reference's fallback when the
//   median absolute deviation is 0.
// * bb_flag_subsample: the voxels of a random subsample given by ranks among the selected voxels (the reference
//   draws 2e6 of them with numpy when there are more): flags of the voxels whose rank is chosen.
// * bb_box_dilate: the binary dilation by a box of half-sizes (rz, ry, rx) (_dilate, noise.py:8-15: maximum filters
//   along each axis, zero outside the cube).
// * bb_like_maps: the maps of the likelihood voxels (likemask, not blank) of _setup_likemask (likelihood.py:176-185):
//   per pixel the sum of the squared data, their number and the sum of the data, in FP64, summed as the reference
//   sums them (sequentially within chunks of 16 channels, the chunks added in order: the same bits), and the data of
//   the likelihood in a box as the context keeps it ((by, bx, nz), the voxels not used NaN).
// * bb_lag_xy, bb_lag_z: the autocorrelation sums of the centred noise that the reference forms with zero-padded
//   FFTs, as direct FP64 sums over the lags it keeps: sum over pairs of used voxels of c c' and the number of such
//   pairs (c = data - median). The padding of the FFTs exceeds the largest lag, so these are the same sums.
//   Spatial lags: one workgroup per 32x32 pixel tile and group of channels stages the tile and a window of its
//   partners in LDS, and every thread sums its lags over the tile; lags (dy, dx) and (-dy, -dx) have the same sum,
//   so only the half-plane dy > 0 or (dy = 0, dx >= 0) is summed. Spectral lags: one thread per pixel keeps the
//   last K + 1 values of its spectrum in registers. Partial sums go to the work buffer per workgroup and are added
//   in a fixed order: the results do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/bbgpu.h"
#include "bb_internal.h"

#define BB_NS_THREADS 256
#define BB_LAG_T 32          // pixel tile of the spatial lag sums
#define BB_LAG_NPT 4         // lags per thread of a band of the spatial sums
#define BB_LAG_ZMAX 7        // largest spectral lag
#define BB_SUB_BLK 4096      // voxels per block of the rank scan (16 per thread)

namespace {

__device__ __forceinline__ uint32_t f2key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key2f(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__device__ __forceinline__ bool used_value(float d, int zero_blank) {
    return d == d && fabsf(d) != __builtin_inff() && !(zero_blank && d == 0.0f);
}

__global__ void k_zero_u64(unsigned long long* a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0ull;
}

__global__ void k_zero_u8(uint8_t* a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

__global__ void __launch_bounds__(BB_NS_THREADS)
k_noise_flags(const float* __restrict__ data, const uint8_t* __restrict__ mask, const uint8_t* __restrict__ emis,
              int nz, int64_t plane, int zero_blank, uint8_t* __restrict__ used,
              unsigned long long* __restrict__ chan_mask, unsigned long long* __restrict__ chan_used,
              uint8_t* __restrict__ pix_emis, uint32_t* __restrict__ pix_used) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < plane;
    const int lane = threadIdx.x & 63;
    uint8_t pe = 0;
    uint32_t pn = 0;
    for (int z = 0; z < nz; z++) {
        bool f = false, m = false;
        if (in) {
            const int64_t i = (int64_t)z * plane + p;
            const bool e = emis && emis[i];
            f = used_value(data[i], zero_blank) && !e;
            m = mask && mask[i];
            used[i] = f;
            pe |= e;
            pn += f;
        }
        const uint64_t bf = __ballot(f), bm = __ballot(m);
        if (lane == 0) {
            if (bf) atomicAdd(&chan_used[z], (unsigned long long)__popcll(bf));
            if (bm) atomicAdd(&chan_mask[z], (unsigned long long)__popcll(bm));
        }
    }
    if (in) {
        if (pix_emis) pix_emis[p] = pe;
        if (pix_used) pix_used[p] = pn;
    }
}

// the state of one radix selection: the key bits decided so far and the rank sought among the keys that share them
struct SelState {
    uint32_t prefix, mask;
    int64_t k, n;
};

__global__ void k_sel_init(SelState* st, uint32_t* hist) {
    const int t = threadIdx.x;
    if (t < 2) {
        st[t].prefix = 0;
        st[t].mask = 0;
        st[t].k = 0;
        st[t].n = 0;
    }
    for (int i = t; i < 2 * 256; i += blockDim.x) hist[i] = 0;
}

// histograms of the 8 key bits at shift of the values that agree with each selection's decided bits
__global__ void __launch_bounds__(BB_NS_THREADS)
k_sel_hist(const float* __restrict__ x, const uint8_t* __restrict__ sel, int64_t n, const float* __restrict__ center,
           const SelState* __restrict__ st, uint32_t* __restrict__ hist, int shift) {
    __shared__ uint32_t h[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const SelState a = st[0], b = st[1];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!sel[i]) continue;
        float v = x[i];
        if (center) v = fabsf(v - center[0]);      // (float32, as numpy's |out - median(out)|)
        if (v != v) continue;
        const uint32_t u = f2key(v);
        const uint32_t bin = (u >> shift) & 255u;
        if ((u & a.mask) == a.prefix) atomicAdd(&h[bin], 1u);
        if ((u & b.mask) == b.prefix) atomicAdd(&h[256 + bin], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 512; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// the bin of each selection's rank; the first pass also counts the values and sets the ranks (n - 1) / 2 and n / 2
__global__ void k_sel_pick(SelState* st, uint32_t* hist, int shift, int first) {
    if (threadIdx.x != 0) return;
    for (int s = 0; s < 2; s++) {
        uint32_t* hh = hist + 256 * s;
        if (first) {
            int64_t n = 0;
            for (int b = 0; b < 256; b++) n += hh[b];
            st[s].n = n;
            st[s].k = s == 0 ? (n - 1) / 2 : n / 2;
        }
        int64_t cum = 0;
        for (int b = 0; b < 256; b++) {
            if (cum + (int64_t)hh[b] > st[s].k) {
                st[s].prefix |= (uint32_t)b << shift;
                st[s].mask |= 255u << shift;
                st[s].k -= cum;
                break;
            }
            cum += hh[b];
        }
        for (int b = 0; b < 256; b++) hh[b] = 0;
    }
}

__global__ void k_sel_final(const SelState* st, float* out, int64_t* count) {
    if (threadIdx.x != 0) return;
    const int64_t n = st[0].n;
    if (count) count[0] = n;
    if (n == 0) {
        out[0] = __int_as_float(0x7fc00000);
        return;
    }
    const float a = key2f(st[0].prefix), b = key2f(st[1].prefix);
    // numpy: the middle value itself for an odd count, the float32 mean of the two middle values for an even one
    out[0] = (n & 1) ? a : (a + b) * 0.5f;
}

// FP64 sums of x (c NULL) or of (x - c)^2 over the selected finite values, one per workgroup
__global__ void __launch_bounds__(BB_NS_THREADS)
k_masked_sums(const float* __restrict__ x, const uint8_t* __restrict__ sel, int64_t n, const double* __restrict__ c,
              double* __restrict__ part) {
    double s = 0.0;
    const double cc = c ? c[0] : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float xf = x[i];
        if (!sel[i] || !(xf == xf)) continue;
        const double v = (double)xf;
        s += c ? (v - cc) * (v - cc) : v;
    }
    __shared__ double red[BB_NS_THREADS];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = BB_NS_THREADS / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_mean_final(const double* __restrict__ part, int nb, const int64_t* __restrict__ count,
                             double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[b];
    out[0] = count[0] > 0 ? s / (double)count[0] : 0.0;
}

__global__ void k_std_final(const double* __restrict__ part, int nb, const int64_t* __restrict__ count,
                            float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[b];
    out[0] = count[0] > 0 ? (float)sqrt(s / (double)count[0]) : __int_as_float(0x7fc00000);
}

// the rank scan of the subsample: flagged voxels per block of BB_SUB_BLK
__global__ void __launch_bounds__(BB_NS_THREADS)
k_sub_count(const uint8_t* __restrict__ flags, int64_t n, uint32_t* __restrict__ bc) {
    const int64_t b0 = (int64_t)blockIdx.x * BB_SUB_BLK;
    uint32_t c = 0;
    for (int k = threadIdx.x; k < BB_SUB_BLK; k += BB_NS_THREADS)
        if (b0 + k < n) c += flags[b0 + k] != 0;
    __shared__ uint32_t red[BB_NS_THREADS];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = BB_NS_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bc[blockIdx.x] = red[0];
}

// exclusive prefix sums of the block counts (one workgroup, 256 at a time) into off
__global__ void __launch_bounds__(BB_NS_THREADS)
k_sub_scan(const uint32_t* __restrict__ bc, int64_t nb, int64_t* __restrict__ off) {
    __shared__ int64_t s[BB_NS_THREADS];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < nb; b0 += BB_NS_THREADS) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nb ? (int64_t)bc[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < BB_NS_THREADS; d <<= 1) {
            const int64_t t = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) off[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == BB_NS_THREADS - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
}

__global__ void k_sub_mark(const int64_t* __restrict__ ranks, int64_t nr, int64_t nflag, uint8_t* __restrict__ chosen) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nr; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = ranks[j];
        if (r >= 0 && r < nflag) chosen[r] = 1;
    }
}

// out = flags && chosen[rank among the flagged voxels]; each thread walks 16 consecutive voxels of its block
__global__ void __launch_bounds__(BB_NS_THREADS)
k_sub_select(const uint8_t* __restrict__ flags, int64_t n, int64_t nflag, const int64_t* __restrict__ off,
             const uint8_t* __restrict__ chosen, uint8_t* __restrict__ out) {
    constexpr int PER = BB_SUB_BLK / BB_NS_THREADS;
    const int64_t v0 = (int64_t)blockIdx.x * BB_SUB_BLK + (int64_t)threadIdx.x * PER;
    uint32_t c = 0;
    for (int k = 0; k < PER; k++)
        if (v0 + k < n) c += flags[v0 + k] != 0;
    __shared__ uint32_t s[BB_NS_THREADS];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < BB_NS_THREADS; d <<= 1) {
        const uint32_t t = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t r = off[blockIdx.x] + (int64_t)(s[threadIdx.x] - c);
    for (int k = 0; k < PER; k++) {
        if (v0 + k >= n) break;
        const bool f = flags[v0 + k] != 0;
        out[v0 + k] = f && r < nflag && chosen[r];
        r += f;
    }
}

// one pass of the box dilation along an axis: out = max over |j - i| <= r of in along that axis, 0 outside
__global__ void k_box_max(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int nz, int ny, int nx, int axis,
                          int r) {
    const int64_t n = (int64_t)nz * ny * nx;
    const int len = axis == 0 ? nz : (axis == 1 ? ny : nx);
    const int64_t stride = axis == 0 ? (int64_t)ny * nx : (axis == 1 ? nx : 1);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int pos = axis == 0 ? (int)(i / ((int64_t)ny * nx)) : (axis == 1 ? (int)((i / nx) % ny) : (int)(i % nx));
        const int lo = pos - r < 0 ? 0 : pos - r, hi = pos + r >= len ? len - 1 : pos + r;
        uint8_t m = 0;
        for (int j = lo; j <= hi && !m; j++) m = in[i + (int64_t)(j - pos) * stride] != 0;
        out[i] = m;
    }
}

// ---- spatial lag sums ------------------------------------------------------------------------------------------------
// The half-plane of lags in index order: (0, 0 .. L), then (dy, -L .. L) for dy = 1 .. L.
__device__ __host__ __forceinline__ int64_t lag_first(int dy, int L) {
    return dy == 0 ? 0 : (int64_t)(L + 1) + (int64_t)(dy - 1) * (2 * L + 1);
}

struct LagPlan {
    int ntx, nty, nzg, zper;
    int64_t nh;
};

LagPlan lag_plan(int nz, int ny, int nx, int L) {
    LagPlan p;
    p.ntx = (nx + BB_LAG_T - 1) / BB_LAG_T;
    p.nty = (ny + BB_LAG_T - 1) / BB_LAG_T;
    p.nh = lag_first(L + 1, L);
    const int64_t nt = (int64_t)p.ntx * p.nty;
    // about 2048 workgroups, and at most 4M partial sums per quantity (a few tens of MB of work buffer)
    int64_t g = (2048 + nt - 1) / nt;
    const int64_t gcap = (4 << 20) / (nt * p.nh);
    if (g > gcap) g = gcap;
    if (g > nz) g = nz;
    if (g < 1) g = 1;
    p.zper = (int)((nz + g - 1) / g);
    p.nzg = (nz + p.zper - 1) / p.zper;
    return p;
}

// LDS of a band of lags dy in [dy0, dy1]: the tile (T x T) and the window of its partners ((T + dy1 - dy0) rows of
// T + 2L pixels), c as f64 and the flags as u8
size_t lag_lds(int L, int dy0, int dy1) {
    const size_t cells = (size_t)(BB_LAG_T + dy1 - dy0) * (BB_LAG_T + 2 * L) + BB_LAG_T * BB_LAG_T;
    return cells * 9;
}

__global__ void __launch_bounds__(BB_NS_THREADS)
k_lag_xy(const float* __restrict__ data, const uint8_t* __restrict__ w, const uint8_t* __restrict__ chan_use, int nz,
         int ny, int nx, int L, int dy0, int dy1, double med, int zper, int ntx, int64_t nh,
         double* __restrict__ part) {
    extern __shared__ double lds_lag[];
    constexpr int T = BB_LAG_T;
    const int W = T + 2 * L, H = T + dy1 - dy0;
    double* cs = lds_lag;                                   // window (H x W)
    double* ca = cs + (size_t)H * W;                        // tile (T x T)
    uint8_t* ws = (uint8_t*)(ca + T * T);
    uint8_t* wa = ws + (size_t)H * W;
    const int tile = blockIdx.x, zg = blockIdx.y;
    const int y0 = (tile / ntx) * T, x0 = (tile % ntx) * T;
    const int64_t plane = (int64_t)ny * nx;
    const int64_t l0 = lag_first(dy0, L), l1 = lag_first(dy1 + 1, L);
    // the partner of the tile pixel (ay, ax) at lag (dy, dx) is the window cell (ay + dy - dy0, ax + L + dx)
    int off[BB_LAG_NPT];
    bool act[BB_LAG_NPT];
    double sc[BB_LAG_NPT];
    uint32_t sw[BB_LAG_NPT];
#pragma unroll
    for (int k = 0; k < BB_LAG_NPT; k++) {
        const int64_t l = l0 + threadIdx.x + (int64_t)k * BB_NS_THREADS;
        act[k] = l < l1;
        int dy = dy0, dx = 0;
        if (act[k]) {
            if (l <= L) {
                dy = 0;
                dx = (int)l;
            } else {
                const int64_t q = l - (L + 1);
                dy = 1 + (int)(q / (2 * L + 1));
                dx = (int)(q % (2 * L + 1)) - L;
            }
        }
        off[k] = (dy - dy0) * W + L + dx;
        sc[k] = 0.0;
        sw[k] = 0;
    }
    const int z1 = min(nz, (zg + 1) * zper);
    for (int z = zg * zper; z < z1; z++) {
        if (!chan_use[z]) continue;
        const float* dz = data + z * plane;
        const uint8_t* wz = w + z * plane;
        __syncthreads();
        for (int i = threadIdx.x; i < H * W; i += BB_NS_THREADS) {
            const int yy = y0 + dy0 + i / W, xx = x0 - L + i % W;
            double c = 0.0;
            uint8_t f = 0;
            if (yy < ny && xx >= 0 && xx < nx) {
                const int64_t q = (int64_t)yy * nx + xx;
                f = wz[q] != 0;
                if (f) c = (double)dz[q] - med;
            }
            cs[i] = c;
            ws[i] = f;
        }
        for (int i = threadIdx.x; i < T * T; i += BB_NS_THREADS) {
            const int yy = y0 + i / T, xx = x0 + i % T;
            double c = 0.0;
            uint8_t f = 0;
            if (yy < ny && xx < nx) {
                const int64_t q = (int64_t)yy * nx + xx;
                f = wz[q] != 0;
                if (f) c = (double)dz[q] - med;
            }
            ca[i] = c;
            wa[i] = f;
        }
        __syncthreads();
        for (int ay = 0; ay < T; ay++) {
            for (int ax = 0; ax < T; ax++) {
                if (!wa[ay * T + ax]) continue;          // (uniform: every thread reads the same tile pixel)
                const double a = ca[ay * T + ax];
                const int base = ay * W + ax;
#pragma unroll
                for (int k = 0; k < BB_LAG_NPT; k++) {
                    if (!act[k]) continue;
                    const int o = base + off[k];
                    sc[k] += a * cs[o];
                    sw[k] += ws[o];
                }
            }
        }
    }
    const int64_t wg = (int64_t)zg * gridDim.x + tile;
#pragma unroll
    for (int k = 0; k < BB_LAG_NPT; k++) {
        if (!act[k]) continue;
        const int64_t l = l0 + threadIdx.x + (int64_t)k * BB_NS_THREADS;
        part[(wg * nh + l) * 2 + 0] = sc[k];
        part[(wg * nh + l) * 2 + 1] = (double)sw[k];
    }
}

// ---- spectral lag sums -----------------------------------------------------------------------------------------------
// Pairs (z - t, z) of channels of the same run (runid >= 0), both voxels used and at a used pixel (pix_use NULL:
// every pixel); part[(block * (K + 1) + t) * 2 + {0, 1}].
__global__ void __launch_bounds__(BB_NS_THREADS)
k_lag_z(const float* __restrict__ data, const uint8_t* __restrict__ w, const uint8_t* __restrict__ pix_use,
        const int* __restrict__ runid, int nz, int64_t plane, int K, double med, double* __restrict__ part) {
    constexpr int KM = BB_LAG_ZMAX + 1;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < plane && (!pix_use || pix_use[p]);
    double hc[KM], sc[KM];
    int hr[KM];
    bool hf[KM];
    uint32_t sw[KM];
#pragma unroll
    for (int t = 0; t < KM; t++) {
        hc[t] = 0.0;
        hr[t] = -1;
        hf[t] = false;
        sc[t] = 0.0;
        sw[t] = 0;
    }
    if (in) {
        for (int z = 0; z < nz; z++) {
#pragma unroll
            for (int t = KM - 1; t > 0; t--) {
                hc[t] = hc[t - 1];
                hr[t] = hr[t - 1];
                hf[t] = hf[t - 1];
            }
            const int64_t i = (int64_t)z * plane + p;
            hr[0] = runid[z];
            hf[0] = hr[0] >= 0 && w[i];
            hc[0] = hf[0] ? (double)data[i] - med : 0.0;
            if (!hf[0]) continue;
#pragma unroll
            for (int t = 0; t < KM; t++) {
                if (t <= K && hf[t] && hr[t] == hr[0]) {
                    sc[t] += hc[0] * hc[t];
                    sw[t] += 1;
                }
            }
        }
    }
    __shared__ double red[BB_NS_THREADS];
#pragma unroll
    for (int t = 0; t < KM; t++) {
        if (t > K) break;
        for (int q = 0; q < 2; q++) {
            red[threadIdx.x] = q ? (double)sw[t] : sc[t];
            __syncthreads();
            for (int s = BB_NS_THREADS / 2; s > 0; s >>= 1) {
                if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) part[((int64_t)blockIdx.x * (K + 1) + t) * 2 + q] = red[0];
            __syncthreads();
        }
    }
}

// the partial sums of every lag added in workgroup order; part[(blk * nlag + l) * 2 + {0, 1}]
__global__ void k_lag_final(const double* __restrict__ part, int64_t nlag, int64_t nblk, double* __restrict__ c,
                            double* __restrict__ wsum) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlag) return;
    double a = 0.0, b = 0.0;
    for (int64_t k = 0; k < nblk; k++) {
        a += part[(k * nlag + l) * 2];
        b += part[(k * nlag + l) * 2 + 1];
    }
    c[l] = a;
    wsum[l] = b;
}

// one thread per pixel: its spectrum in chunks of 16 channels
#define BB_MAP_CHUNK 16
__global__ void __launch_bounds__(BB_NS_THREADS)
k_like_maps(const float* __restrict__ data, const uint8_t* __restrict__ likemask, const uint8_t* __restrict__ region,
            int nz, int ny, int nx, int bx0, int by0, int bx1, int by1, float* __restrict__ dvalid,
            double* __restrict__ dm2, int32_t* __restrict__ nvox, double* __restrict__ mom0) {
    const int64_t plane = (int64_t)ny * nx;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) return;
    const int x = (int)(p % nx), y = (int)(p / nx);
    const bool inbox = dvalid && x >= bx0 && x < bx1 && y >= by0 && y < by1;
    const bool inreg = !region || region[p];
    float* dv = inbox ? dvalid + ((int64_t)(y - by0) * (bx1 - bx0) + (x - bx0)) * nz : nullptr;
    double a2 = 0.0, a1 = 0.0;
    int32_t n = 0;
    for (int z0 = 0; z0 < nz; z0 += BB_MAP_CHUNK) {
        double c2 = 0.0, c1 = 0.0;
        const int z1 = min(nz, z0 + BB_MAP_CHUNK);
        for (int z = z0; z < z1; z++) {
            const int64_t i = (int64_t)z * plane + p;
            const float d = data[i];
            const bool v = (!likemask || likemask[i]) && used_value(d, 1);
            const double dd = v ? (double)d : 0.0;
            c2 += dd * dd;
            c1 += dd;
            n += v;
            if (dv) dv[z] = (v && inreg) ? d : __int_as_float(0x7fc00000);
        }
        a2 += c2;
        a1 += c1;
    }
    if (dm2) dm2[p] = a2;
    if (nvox) nvox[p] = n;
    if (mom0) mom0[p] = a1;
}

int grid_of(int64_t n) {
    int64_t g = (n + BB_NS_THREADS - 1) / BB_NS_THREADS;
    return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------------
extern "C" int bb_noise_flags(const float* data, const uint8_t* mask, const uint8_t* emis, int nz, int ny, int nx,
                              int zero_blank, uint8_t* used, uint64_t* chan_mask, uint64_t* chan_used,
                              uint8_t* pix_emis, uint32_t* pix_used, void* stream) {
    if (nz < 0 || ny < 0 || nx < 0) return bb_fail(BB_ERR_ARG, "bb_noise_flags: bad sizes %d %d %d", nz, ny, nx);
    const int64_t plane = (int64_t)ny * nx;
    if (plane * nz == 0) return 0;
    if (!data || !used || !chan_mask || !chan_used) return bb_fail(BB_ERR_ARG, "bb_noise_flags: NULL array");
    if (plane > (int64_t)0x7fffffff * BB_NS_THREADS) return bb_fail(BB_ERR_ARG, "bb_noise_flags: plane too large");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(BB_NS_THREADS), 0, st, (unsigned long long*)chan_mask, (int64_t)nz);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(BB_NS_THREADS), 0, st, (unsigned long long*)chan_used, (int64_t)nz);
    const int64_t nb = (plane + BB_NS_THREADS - 1) / BB_NS_THREADS;
    hipLaunchKernelGGL(k_noise_flags, dim3((unsigned)nb), dim3(BB_NS_THREADS), 0, st, data, mask, emis, nz, plane,
                       zero_blank, used, (unsigned long long*)chan_mask, (unsigned long long*)chan_used, pix_emis,
                       pix_used);
    return bb_hip_check(hipGetLastError(), "bb_noise_flags");
}

extern "C" int64_t bb_masked_median_work_size(void) {
    return 2 * (int64_t)sizeof(SelState) + 512 * (int64_t)sizeof(uint32_t);
}

extern "C" int bb_masked_median(const float* x, const uint8_t* sel, int64_t n, const float* center, void* work,
                                float* out, int64_t* count, void* stream) {
    if (n < 0) return bb_fail(BB_ERR_ARG, "bb_masked_median: n = %lld", (long long)n);
    if (!x || !sel || !work || !out) return bb_fail(BB_ERR_ARG, "bb_masked_median: NULL array");
    hipStream_t st = (hipStream_t)stream;
    SelState* s = (SelState*)work;
    uint32_t* hist = (uint32_t*)(s + 2);
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(256), 0, st, s, hist);
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        hipLaunchKernelGGL(k_sel_hist, dim3(grid_of(n)), dim3(BB_NS_THREADS), 0, st, x, sel, n, center,
                           (const SelState*)s, hist, shift);
        hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(64), 0, st, s, hist, shift, pass == 0 ? 1 : 0);
    }
    hipLaunchKernelGGL(k_sel_final, dim3(1), dim3(64), 0, st, (const SelState*)s, out, count);
    return bb_hip_check(hipGetLastError(), "bb_masked_median");
}

extern "C" int64_t bb_masked_std_work_size(void) { return (65535 + 1) * (int64_t)sizeof(double); }

extern "C" int bb_masked_std(const float* x, const uint8_t* sel, int64_t n, const int64_t* count, void* work,
                             float* out, void* stream) {
    if (n < 0) return bb_fail(BB_ERR_ARG, "bb_masked_std: n = %lld", (long long)n);
    if (!x || !sel || !count || !work || !out) return bb_fail(BB_ERR_ARG, "bb_masked_std: NULL array");
    hipStream_t st = (hipStream_t)stream;
    const int nb = grid_of(n);
    double* mean = (double*)work;
    double* part = mean + 1;
    hipLaunchKernelGGL(k_masked_sums, dim3(nb), dim3(BB_NS_THREADS), 0, st, x, sel, n, (const double*)nullptr, part);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(64), 0, st, (const double*)part, nb, count, mean);
    hipLaunchKernelGGL(k_masked_sums, dim3(nb), dim3(BB_NS_THREADS), 0, st, x, sel, n, (const double*)mean, part);
    hipLaunchKernelGGL(k_std_final, dim3(1), dim3(64), 0, st, (const double*)part, nb, count, out);
    return bb_hip_check(hipGetLastError(), "bb_masked_std");
}

extern "C" int64_t bb_flag_subsample_work_size(int64_t n, int64_t nflag) {
    if (n < 0 || nflag < 0) return -1;
    const int64_t nb = (n + BB_SUB_BLK - 1) / BB_SUB_BLK;
    return nb * (int64_t)(sizeof(uint32_t) + sizeof(int64_t)) + nflag;
}

extern "C" int bb_flag_subsample(const uint8_t* flags, int64_t n, int64_t nflag, const int64_t* ranks, int64_t nr,
                                 void* work, int64_t work_bytes, uint8_t* out, void* stream) {
    if (n < 0 || nflag < 0 || nr < 0) return bb_fail(BB_ERR_ARG, "bb_flag_subsample: negative size");
    if (n == 0) return 0;
    if (!flags || !work || !out || (nr > 0 && !ranks)) return bb_fail(BB_ERR_ARG, "bb_flag_subsample: NULL array");
    if (work_bytes < bb_flag_subsample_work_size(n, nflag))
        return bb_fail(BB_ERR_ARG, "bb_flag_subsample: work of %lld bytes, %lld needed", (long long)work_bytes,
                       (long long)bb_flag_subsample_work_size(n, nflag));
    const int64_t nb = (n + BB_SUB_BLK - 1) / BB_SUB_BLK;
    if (nb > 0x7fffffff) return bb_fail(BB_ERR_ARG, "bb_flag_subsample: cube too large");
    hipStream_t st = (hipStream_t)stream;
    int64_t* off = (int64_t*)work;
    uint32_t* bc = (uint32_t*)(off + nb);
    uint8_t* chosen = (uint8_t*)(bc + nb);
    if (nflag > 0) hipLaunchKernelGGL(k_zero_u8, dim3(grid_of(nflag)), dim3(BB_NS_THREADS), 0, st, chosen, nflag);
    if (nr > 0) hipLaunchKernelGGL(k_sub_mark, dim3(grid_of(nr)), dim3(BB_NS_THREADS), 0, st, ranks, nr, nflag, chosen);
    hipLaunchKernelGGL(k_sub_count, dim3((unsigned)nb), dim3(BB_NS_THREADS), 0, st, flags, n, bc);
    hipLaunchKernelGGL(k_sub_scan, dim3(1), dim3(BB_NS_THREADS), 0, st, (const uint32_t*)bc, nb, off);
    hipLaunchKernelGGL(k_sub_select, dim3((unsigned)nb), dim3(BB_NS_THREADS), 0, st, flags, n, nflag,
                       (const int64_t*)off, (const uint8_t*)chosen, out);
    return bb_hip_check(hipGetLastError(), "bb_flag_subsample");
}

extern "C" int bb_box_dilate(const uint8_t* m, int nz, int ny, int nx, int rz, int ry, int rx, uint8_t* tmp,
                             uint8_t* out, void* stream) {
    if (nz < 0 || ny < 0 || nx < 0 || rz < 0 || ry < 0 || rx < 0) return bb_fail(BB_ERR_ARG, "bb_box_dilate: bad sizes");
    const int64_t n = (int64_t)nz * ny * nx;
    if (n == 0) return 0;
    const int rs[3] = {rz, ry, rx};
    const int npass = (rz > 0) + (ry > 0) + (rx > 0);
    if (!m || !out || m == out || (npass > 1 && (!tmp || tmp == out || tmp == m)))
        return bb_fail(BB_ERR_ARG, "bb_box_dilate: arrays missing or aliased");
    hipStream_t st = (hipStream_t)stream;
    if (npass == 0) return bb_hip_check(hipMemcpyAsync(out, m, n, hipMemcpyDeviceToDevice, st), "bb_box_dilate");
    // the passes alternate between tmp and out so that the last lands in out
    const uint8_t* src = m;
    uint8_t* dst = (npass % 2) ? out : tmp;
    for (int a = 0; a < 3; a++) {
        if (rs[a] <= 0) continue;
        hipLaunchKernelGGL(k_box_max, dim3(grid_of(n)), dim3(BB_NS_THREADS), 0, st, src, dst, nz, ny, nx, a, rs[a]);
        src = dst;
        dst = dst == out ? tmp : out;
    }
    return bb_hip_check(hipGetLastError(), "bb_box_dilate");
}

extern "C" int bb_lag_xy_max(void) {
    // the largest L whose one-row band still fits the LDS
    int L = 1;
    while (lag_lds(L + 1, 0, 0) <= BB_LDS_MAX) L++;
    return L;
}

extern "C" int64_t bb_lag_xy_work_size(int nz, int ny, int nx, int L) {
    if (nz < 1 || ny < 1 || nx < 1 || L < 1) return -1;
    const LagPlan p = lag_plan(nz, ny, nx, L);
    return (int64_t)p.ntx * p.nty * p.nzg * p.nh * 2;
}

extern "C" int bb_lag_xy(const float* data, const uint8_t* w, const uint8_t* chan_use, int nz, int ny, int nx, int L,
                         double med, double* work, int64_t work_len, double* c, double* wsum, void* stream) {
    if (nz < 1 || ny < 1 || nx < 1 || L < 1) return bb_fail(BB_ERR_ARG, "bb_lag_xy: bad sizes (L >= 1)");
    if (!data || !w || !chan_use || !work || !c || !wsum) return bb_fail(BB_ERR_ARG, "bb_lag_xy: NULL array");
    if (L > bb_lag_xy_max()) return bb_fail(BB_ERR_ARG, "bb_lag_xy: largest lag %d > %d", L, bb_lag_xy_max());
    if (work_len < bb_lag_xy_work_size(nz, ny, nx, L)) return bb_fail(BB_ERR_ARG, "bb_lag_xy: work too small");
    hipStream_t st = (hipStream_t)stream;
    const LagPlan p = lag_plan(nz, ny, nx, L);
    const dim3 grid((unsigned)(p.ntx * p.nty), (unsigned)p.nzg);
    // bands of dy whose lags fit the threads (BB_LAG_NPT each) and whose window fits the LDS; a single row of lags
    // (2L + 1 <= 1024) always does
    int dy0 = 0;
    while (dy0 <= L) {
        int dy1 = dy0;
        while (dy1 + 1 <= L && lag_first(dy1 + 2, L) - lag_first(dy0, L) <= BB_LAG_NPT * BB_NS_THREADS &&
               lag_lds(L, dy0, dy1 + 1) <= BB_LDS_MAX)
            dy1++;
        if (lag_lds(L, dy0, dy1) > 65536) {
            const int rc = bb_hip_check(hipFuncSetAttribute((const void*)k_lag_xy,
                                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                                            (int)lag_lds(L, dy0, dy1)), "bb_lag_xy");
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_lag_xy, grid, dim3(BB_NS_THREADS), lag_lds(L, dy0, dy1), st, data, w, chan_use, nz, ny,
                           nx, L, dy0, dy1, med, p.zper, p.ntx, p.nh, work);
        dy0 = dy1 + 1;
    }
    const int64_t nblk = (int64_t)p.ntx * p.nty * p.nzg;
    hipLaunchKernelGGL(k_lag_final, dim3((unsigned)((p.nh + 63) / 64)), dim3(64), 0, st, (const double*)work, p.nh,
                       nblk, c, wsum);
    return bb_hip_check(hipGetLastError(), "bb_lag_xy");
}

extern "C" int64_t bb_lag_z_work_size(int ny, int nx, int K) {
    if (ny < 1 || nx < 1 || K < 0 || K > BB_LAG_ZMAX) return -1;
    const int64_t nb = ((int64_t)ny * nx + BB_NS_THREADS - 1) / BB_NS_THREADS;
    return nb * (K + 1) * 2;
}

extern "C" int bb_lag_z(const float* data, const uint8_t* w, const uint8_t* pix_use, const int32_t* runid, int nz,
                        int ny, int nx, int K, double med, double* work, int64_t work_len, double* c, double* wsum,
                        void* stream) {
    if (nz < 1 || ny < 1 || nx < 1 || K < 0 || K > BB_LAG_ZMAX)
        return bb_fail(BB_ERR_ARG, "bb_lag_z: bad sizes (0 <= K <= %d)", BB_LAG_ZMAX);
    if (!data || !w || !runid || !work || !c || !wsum) return bb_fail(BB_ERR_ARG, "bb_lag_z: NULL array");
    if (work_len < bb_lag_z_work_size(ny, nx, K)) return bb_fail(BB_ERR_ARG, "bb_lag_z: work too small");
    hipStream_t st = (hipStream_t)stream;
    const int64_t plane = (int64_t)ny * nx;
    const int64_t nb = (plane + BB_NS_THREADS - 1) / BB_NS_THREADS;
    hipLaunchKernelGGL(k_lag_z, dim3((unsigned)nb), dim3(BB_NS_THREADS), 0, st, data, w, pix_use, (const int*)runid, nz,
                       plane, K, med, work);
    hipLaunchKernelGGL(k_lag_final, dim3(1), dim3(64), 0, st, (const double*)work, (int64_t)(K + 1), nb, c, wsum);
    return bb_hip_check(hipGetLastError(), "bb_lag_z");
}

extern "C" int bb_like_maps(const float* data, const uint8_t* likemask, const uint8_t* region, int nz, int ny, int nx,
                            bb_box box, float* dvalid, double* dm2, int32_t* nvox, double* mom0, void* stream) {
    if (nz < 1 || ny < 1 || nx < 1) return bb_fail(BB_ERR_ARG, "bb_like_maps: bad sizes %d %d %d", nz, ny, nx);
    if (!data) return bb_fail(BB_ERR_ARG, "bb_like_maps: data is NULL");
    if (dvalid && (box.blo0 < 0 || box.blo1 < 0 || box.bhi0 > nx || box.bhi1 > ny || box.bhi0 <= box.blo0 ||
                   box.bhi1 <= box.blo1))
        return bb_fail(BB_ERR_ARG, "bb_like_maps: box outside the cube or empty");
    const int64_t plane = (int64_t)ny * nx;
    const int64_t nb = (plane + BB_NS_THREADS - 1) / BB_NS_THREADS;
    hipLaunchKernelGGL(k_like_maps, dim3((unsigned)nb), dim3(BB_NS_THREADS), 0, (hipStream_t)stream, data, likemask,
                       region, nz, ny, nx, box.blo0, box.blo1, box.bhi0, box.bhi1, dvalid, dm2, nvox, mom0);
    return bb_hip_check(hipGetLastError(), "bb_like_maps");
}
