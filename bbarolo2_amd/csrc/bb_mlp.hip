// The likelihood surrogate: a small multilayer perceptron din -> H (SiLU) x nh -> 1 with BF16 weights and
// activations and FP32 accumulation on the matrix cores, trained on the device by full-batch Adam on the
// (parameters, log-likelihood) pairs a sampler has evaluated. The reference's learned component is the
// regression network of nautilus's neural bounds (bbarolo/bayesian/nautilus/mlp.py:17-222, neural.py:60-127: a
// ReLU network fitted to the likelihood rank of the live points and used to decide where to draw); this one
// regresses the log-likelihood itself and screens proposals.
//
// Every matrix product is C (M x N) = A (M x K) . B (N x K)^T with both operands contiguous along K (so the three
// products of a layer -- forward, weight gradient, input gradient -- read BF16 rows), on the gfx950 instruction
// v_mfma_f32_16x16x32_bf16: one wavefront per 16 x 16 tile of C. Lane l holds A[l % 16][8 (l / 16) .. +7],
// B[l % 16][8 (l / 16) .. +7] (rows of B, i.e. columns of B^T) and D[4 (l / 16) + j][l % 16], j = 0..3.
// Reductions (bias gradients, the output layer, the loss) are fixed-order tree sums: the training is
// deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

typedef __bf16 bb_bf16x8 __attribute__((ext_vector_type(8)));
typedef float bb_f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short mlp_bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float mlp_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ float mlp_sig(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sig(z); }
__device__ __forceinline__ float mlp_dsilu(float z) {
    const float s = mlp_sig(z);
    return s * (1.0f + z * (1.0f - s));
}

enum { MLP_FWD = 0, MLP_GRAD = 1, MLP_BWD = 2 };

// C = A . B^T (M, N multiples of 16, K of 32) with the epilogue EPI:
//   MLP_FWD  z = C + bias[col]: Z (fp32, M x N), a = silu(z) as BF16 in Ab (M x N) and AbT (N x M);
//   MLP_GRAD the partial product over K slice blockIdx.y ([y kchunk, (y + 1) kchunk)) into G + y M N (fp32; the
//            weight gradients have K = the training rows: split so that the small outputs fill the chip);
//   MLP_BWD  d = C * silu'(Zp) for rows < nvalid (0 beyond) as BF16 in Ab (M x N) and AbT (N x M).
template <int EPI>
__global__ void __launch_bounds__(256)
k_mlp_gemm(const unsigned short* __restrict__ A, const unsigned short* __restrict__ B, int M, int N, int K,
           const float* __restrict__ bias, float* __restrict__ Z, unsigned short* __restrict__ Ab,
           unsigned short* __restrict__ AbT, float* __restrict__ G, const float* __restrict__ Zp, int nvalid,
           int kchunk) {
    const int lane = threadIdx.x & 63;
    const int tiles_n = N >> 4;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= (M >> 4) * tiles_n) return;       // (whole wavefronts)
    const int m0 = (tile / tiles_n) << 4, n0 = (tile % tiles_n) << 4;
    const unsigned short* a = A + (size_t)(m0 + (lane & 15)) * K + 8 * (lane >> 4);
    const unsigned short* b = B + (size_t)(n0 + (lane & 15)) * K + 8 * (lane >> 4);
    const int kb = blockIdx.y * kchunk, ke = kb + kchunk < K ? kb + kchunk : K;
    bb_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k = kb; k < ke; k += 32) {
        const bb_bf16x8 av = *(const bb_bf16x8*)(a + k);
        const bb_bf16x8 bv = *(const bb_bf16x8*)(b + k);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    const int col = n0 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int row = m0 + 4 * (lane >> 4) + j;
        const size_t o = (size_t)row * N + col;
        if (EPI == MLP_FWD) {
            const float z = acc[j] + bias[col];
            Z[o] = z;
            const unsigned short h = mlp_bf(mlp_silu(z));
            Ab[o] = h;
            AbT[(size_t)col * M + row] = h;
        } else if (EPI == MLP_GRAD) {
            G[(size_t)blockIdx.y * M * N + o] = acc[j];
        } else {
            const float d = row < nvalid ? acc[j] * mlp_dsilu(Zp[o]) : 0.f;
            const unsigned short h = mlp_bf(d);
            Ab[o] = h;
            AbT[(size_t)col * M + row] = h;
        }
    }
}

// G[e] = sum over the splits s (in order) of part[s][e].
__global__ void k_mlp_sum_parts(const float* __restrict__ part, int nsplit, int64_t n, float* __restrict__ G) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += part[(size_t)k * n + e];
    G[e] = s;
}

// Inputs (n, din) fp32 -> BF16 rows X (np, dp) and columns XT (dp, np), zero beyond n and din.
__global__ void k_mlp_in(const float* __restrict__ x, int n, int din, int np, int dp, unsigned short* __restrict__ X,
                         unsigned short* __restrict__ XT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np * dp) return;
    const int r = i / dp, c = i % dp;
    const unsigned short h = (r < n && c < din) ? mlp_bf(x[(size_t)r * din + c]) : (unsigned short)0;
    X[i] = h;
    XT[(size_t)c * np + r] = h;
}

// Output layer: y[r] = sum_i a[r][i] w[i] + b (FP32 sums of BF16 products), rows < np.
__global__ void k_mlp_out(const unsigned short* __restrict__ a, const unsigned short* __restrict__ w, const float* b,
                          int np, int H, float* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    const unsigned short* ar = a + (size_t)r * H;
    float s = 0.f;
    for (int i = 0; i < H; i++) s = fmaf(mlp_f(ar[i]), mlp_f(w[i]), s);
    y[r] = s + b[0];
}

// Gradient into the last hidden layer: dz[r][i] = (y[r] - t[r]) / n * w[i] * silu'(z[r][i]) (rows < n; 0 beyond),
// BF16 in D (np x H) and DT (H x np).
__global__ void k_mlp_out_bwd(const float* __restrict__ y, const float* __restrict__ t, int n, int np, int H,
                              const unsigned short* __restrict__ w, const float* __restrict__ Z,
                              unsigned short* __restrict__ D, unsigned short* __restrict__ DT) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np * H) return;
    const int r = e / H, i = e % H;
    const float dy = r < n ? (y[r] - t[r]) / (float)n : 0.f;
    const unsigned short h = mlp_bf(dy * mlp_f(w[i]) * mlp_dsilu(Z[e]));
    D[e] = h;
    DT[(size_t)i * np + r] = h;
}

// Fixed-order block sum (256 threads).
__device__ float mlp_block_sum(float v, float* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const float r = red[0];
    __syncthreads();
    return r;
}

// Block i < H: gw[i] = sum_r dy[r] a[r][i]; block H: gb = sum_r dy[r] and loss = sum_r (y - t)^2 / (2 n).
__global__ void __launch_bounds__(256)
k_mlp_out_grad(const float* __restrict__ y, const float* __restrict__ t, int n, const unsigned short* __restrict__ a,
               int H, float* __restrict__ gw, float* __restrict__ gb, float* __restrict__ loss) {
    __shared__ float red[256];
    const int i = blockIdx.x;
    float s = 0.f, l = 0.f;
    for (int r = threadIdx.x; r < n; r += 256) {
        const float d = y[r] - t[r];
        if (i < H) s += d / (float)n * mlp_f(a[(size_t)r * H + i]);
        else {
            s += d / (float)n;
            l += d * d;
        }
    }
    s = mlp_block_sum(s, red);
    if (i < H) {
        if (threadIdx.x == 0) gw[i] = s;
        return;
    }
    l = mlp_block_sum(l, red);
    if (threadIdx.x == 0) {
        gb[0] = s;
        if (loss) loss[0] = 0.5f * l / (float)n;
    }
}

// Bias gradients: g[o] = sum_r DT[o][r] (the BF16 layer gradients), one block per output.
__global__ void __launch_bounds__(256)
k_mlp_bias_grad(const unsigned short* __restrict__ DT, int np, float* __restrict__ g) {
    __shared__ float red[256];
    const int o = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < np; r += 256) s += mlp_f(DT[(size_t)o * np + r]);
    s = mlp_block_sum(s, red);
    if (threadIdx.x == 0) g[o] = s;
}

// The step counter of Adam (count[0]) and its bias corrections cc = (1 - b1^t, 1 - b2^t); the loss of the step
// (loss_slot) appended to hist at count[1]. One thread; runs between the backward pass and k_mlp_adam, so that a
// captured training step replays with the right constants.
__global__ void k_mlp_tick(int* __restrict__ count, float* __restrict__ cc, const float* __restrict__ loss_slot,
                           float* __restrict__ hist, int hist_cap, float b1, float b2) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int t = count[0] + 1;
    count[0] = t;
    cc[0] = 1.f - powf(b1, (float)t);
    cc[1] = 1.f - powf(b2, (float)t);
    const int k = count[1];
    if (k < hist_cap) hist[k] = loss_slot[0];
    count[1] = k + 1;
}

// Adam (bias-corrected by cc) on the FP32 master parameters.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int64_t np_, float lr, float b1, float b2, float eps,
                           const float* __restrict__ cc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np_) return;
    const float c1 = cc[0], c2 = cc[1];
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// The BF16 copies of a weight matrix W (rows x cols, fp32): Wb (rows x cols) and WbT (cols x rows).
__global__ void k_mlp_cast(const float* __restrict__ W, int rows, int cols, unsigned short* __restrict__ Wb,
                           unsigned short* __restrict__ WbT) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int r = e / cols, c = e % cols;
    const unsigned short h = mlp_bf(W[e]);
    Wb[e] = h;
    if (WbT) WbT[(size_t)c * rows + r] = h;
}

// ---------------------------------------------------------------------------------------------------------
// Launchers (host). Shapes are checked by the C-ABI before these run.
static inline int mlp_blocks(int64_t n, int t) { return (int)((n + t - 1) / t); }

// Splits of K for the weight gradients: slices of 512 rows (at least), at most BB_MLP_SPLITS of them.
#define BB_MLP_SPLITS 64
extern "C" int bb_mlp_splits(int K) {
    int ns = (K + 511) / 512;
    return ns < 1 ? 1 : (ns > BB_MLP_SPLITS ? BB_MLP_SPLITS : ns);
}

// part: for MLP_GRAD, bb_mlp_splits(K) x M x N floats of partial sums (G receives their sum).
extern "C" int bb_launch_mlp_gemm(int epi, const unsigned short* A, const unsigned short* B, int M, int N, int K,
                                  const float* bias, float* Z, unsigned short* Ab, unsigned short* AbT, float* G,
                                  const float* Zp, int nvalid, float* part, hipStream_t st) {
    const int tiles = (M / 16) * (N / 16);
    const dim3 block(256);
    if (epi == MLP_FWD) {
        hipLaunchKernelGGL(k_mlp_gemm<MLP_FWD>, dim3(mlp_blocks(tiles, 4)), block, 0, st, A, B, M, N, K, bias, Z, Ab,
                           AbT, G, Zp, nvalid, K);
    } else if (epi == MLP_GRAD) {
        const int ns = bb_mlp_splits(K);
        const int kchunk = ((K + ns - 1) / ns + 31) / 32 * 32;
        hipLaunchKernelGGL(k_mlp_gemm<MLP_GRAD>, dim3(mlp_blocks(tiles, 4), ns), block, 0, st, A, B, M, N, K, bias, Z,
                           Ab, AbT, part, Zp, nvalid, kchunk);
        hipLaunchKernelGGL(k_mlp_sum_parts, dim3(mlp_blocks((int64_t)M * N, 256)), block, 0, st, part, ns,
                           (int64_t)M * N, G);
    } else {
        hipLaunchKernelGGL(k_mlp_gemm<MLP_BWD>, dim3(mlp_blocks(tiles, 4)), block, 0, st, A, B, M, N, K, bias, Z, Ab,
                           AbT, G, Zp, nvalid, K);
    }
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_in(const float* x, int n, int din, int np, int dp, unsigned short* X, unsigned short* XT,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_in, dim3(mlp_blocks((int64_t)np * dp, 256)), dim3(256), 0, st, x, n, din, np, dp, X, XT);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_out(const unsigned short* a, const unsigned short* w, const float* b, int np, int H,
                                 float* y, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_out, dim3(mlp_blocks(np, 256)), dim3(256), 0, st, a, w, b, np, H, y);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_out_bwd(const float* y, const float* t, int n, int np, int H, const unsigned short* w,
                                     const float* Z, unsigned short* D, unsigned short* DT, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_out_bwd, dim3(mlp_blocks((int64_t)np * H, 256)), dim3(256), 0, st, y, t, n, np, H, w, Z,
                       D, DT);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_out_grad(const float* y, const float* t, int n, const unsigned short* a, int H, float* gw,
                                      float* gb, float* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_out_grad, dim3(H + 1), dim3(256), 0, st, y, t, n, a, H, gw, gb, loss);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_bias_grad(const unsigned short* DT, int np, int H, float* g, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_bias_grad, dim3(H), dim3(256), 0, st, DT, np, g);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                                  float eps, const float* cc, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(mlp_blocks(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, cc);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_tick(int* count, float* cc, const float* loss_slot, float* hist, int hist_cap, float b1,
                                  float b2, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, st, count, cc, loss_slot, hist, hist_cap, b1, b2);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_mlp_cast(const float* W, int rows, int cols, unsigned short* Wb, unsigned short* WbT,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_cast, dim3(mlp_blocks((int64_t)rows * cols, 256)), dim3(256), 0, st, W, rows, cols, Wb,
                       WbT);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------
// Random walks on the region {network(x) >= thr} (bb_mlp_walk): per step the proposals x + scale zl[s] and their
// standardised inputs, the network's forward pass, then the acceptance. The scale of step s follows from that of
// step s - 1 and its acceptance (every thread computes it; thread 0 records it for the next launches), so the whole
// walk runs on the device.
__device__ __forceinline__ float walk_scale(float prev, int acc, int n) {
    return fminf(fmaxf(prev * expf((float)acc / (float)n - 0.3f), 1e-4f), 3.0f);
}

__global__ void k_walk_init(int* cnt, int steps, unsigned char* moved, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < steps) cnt[i] = 0;
    if (i < n) moved[i] = 0;
}

__global__ void k_walk_prop(const float* __restrict__ x, const float* __restrict__ zl, int n, int din, int s,
                            float* __restrict__ sc, const int* __restrict__ cnt, const float* __restrict__ xc,
                            const float* __restrict__ xs, float* __restrict__ prop, float* __restrict__ pn) {
    const float scale = s > 0 ? walk_scale(sc[s - 1], cnt[s - 1], n) : sc[0];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > 0 && i == 0) sc[s] = scale;
    if (i >= (int64_t)n * din) return;
    const int j = (int)(i % din);
    const float p = x[i] + zl[(size_t)s * n * din + i] * scale;
    prop[i] = p;
    pn[i] = (p - xc[j]) / xs[j];
}

__global__ void k_walk_accept(const float* __restrict__ prop, const float* __restrict__ y, int n, int din,
                              const float* __restrict__ thr, float* __restrict__ x, unsigned char* __restrict__ moved,
                              int* __restrict__ cnt_s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool acc = false;
    if (i < n) {
        bool inside = true;
        for (int j = 0; j < din; j++) {
            const float p = prop[(size_t)i * din + j];
            inside = inside && p >= 0.0f && p <= 1.0f;
        }
        acc = inside && y[i] >= thr[0];
        if (acc) {
            for (int j = 0; j < din; j++) x[(size_t)i * din + j] = prop[(size_t)i * din + j];
            moved[i] = 1;
        }
    }
    const unsigned long long m = __ballot(acc);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(cnt_s, (int)__popcll(m));
}

__global__ void k_walk_final(float* sc, const int* cnt, int steps, int n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[steps] = walk_scale(sc[steps - 1], cnt[steps - 1], n);
}

extern "C" int bb_launch_walk_init(int* cnt, int steps, unsigned char* moved, int n, hipStream_t st) {
    const int m = steps > n ? steps : n;
    hipLaunchKernelGGL(k_walk_init, dim3((m + 255) / 256), dim3(256), 0, st, cnt, steps, moved, n);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_walk_prop(const float* x, const float* zl, int n, int din, int s, float* sc, const int* cnt,
                                   const float* xc, const float* xs, float* prop, float* pn, hipStream_t st) {
    const int64_t m = (int64_t)n * din;
    hipLaunchKernelGGL(k_walk_prop, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, x, zl, n, din, s, sc, cnt, xc,
                       xs, prop, pn);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_walk_accept(const float* prop, const float* y, int n, int din, const float* thr, float* x,
                                     unsigned char* moved, int* cnt_s, hipStream_t st) {
    hipLaunchKernelGGL(k_walk_accept, dim3((n + 255) / 256), dim3(256), 0, st, prop, y, n, din, thr, x, moved, cnt_s);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_walk_final(float* sc, const int* cnt, int steps, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_walk_final, dim3(1), dim3(64), 0, st, sc, cnt, steps, n);
    return (int)hipGetLastError();
}
