// The deconvolved density profile on the GPU (bbarolo/profile.py:33-138, bayesian/density.py:13-88): the
// per-evaluation densities of the rings when the geometry is free and DENS is not.
//
//   k_rm_finish   fixed-point ring-map counts -> FP64 maps, convolved with the beam ('same', no 1e-12 cut) and
//                 scaled (DiskModel.ring_maps, galmod.py:1340-1345). One workgroup per 16 x 16 output tile of
//                 one map; the input tile with its halo and the taps live in LDS.
//   k_gram        the Gram matrix of the weighted system [A | b], A = K / s and b = y / s on the usable pixels
//                 (weight 1/s, 0 elsewhere): G = A^T A, h = A^T b and b^T b in one (n+1) x (n+1) matrix. Each
//                 workgroup sums a range of pixels staged through LDS; k_gram_reduce adds the partial matrices
//                 in a fixed order (deterministic).
//   k_nnls        one wavefront per evaluation: solve_profile on the normal equations. The QR of the
//                 reference reduces the least squares to the n x n triangle R with R^T R = A^T A and
//                 R^T c = A^T b, so min |R x - c|^2 + |w L x|^2 and min x^T (G + L^T W^2 L) x - 2 h^T x have the
//                 same minimiser over x >= 0; it is found with the Lawson-Hanson active set, the passive
//                 systems by Cholesky in LDS. First solution with the relative prior, misfit from the Gram
//                 matrix (b^T b - 2 h^T x + x^T G x), curvature weights from it, final solution.
//   k_set_dens    the float32 densities (the ring table of Bayesian3DFIT stores float32, bayesian/rings.py:54-61)
//                 interpolated into the density row of the model's sub-ring tables.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define BB_RM_T 16           // output tile side of k_rm_finish
#define BB_GR_PIX 128        // pixels staged per pass of k_gram
#define BB_GR_THREADS 256
#define BB_NN_MAX 128        // largest system (rings + annulus) of k_nnls

// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(BB_RM_T * BB_RM_T)
k_rm_finish(const unsigned long long* __restrict__ counts, double* __restrict__ maps, int by, int bx,
            const double* __restrict__ kern, int ky, int kx, double unit, double scale, int convolve) {
    extern __shared__ double rm_lds[];
    const int p = blockIdx.z;
    const int tx = threadIdx.x % BB_RM_T, ty = threadIdx.x / BB_RM_T;
    const int x0 = blockIdx.x * BB_RM_T, y0 = blockIdx.y * BB_RM_T;
    const unsigned long long* in = counts + (size_t)p * by * bx;
    double* out = maps + (size_t)p * by * bx;
    if (!convolve) {
        const int x = x0 + tx, y = y0 + ty;
        if (x < bx && y < by) out[(size_t)y * bx + x] = (double)in[(size_t)y * bx + x] * unit;
        return;
    }
    const int hy = ky / 2, hx = kx / 2;
    const int W = BB_RM_T + kx - 1, H = BB_RM_T + ky - 1;
    double* tile = rm_lds;
    double* tk = rm_lds + W * H;
    // input rows y0 - (ky-1-hy) .. y0 + T-1 + hy, columns likewise: out(y, x) = sum K[j][i] in[y + hy - j][x + hx - i]
    const int iy0 = y0 + hy - (ky - 1), ix0 = x0 + hx - (kx - 1);
    int any = 0;
    for (int e = threadIdx.x; e < W * H; e += blockDim.x) {
        const int r = e / W, c = e % W;
        const int iy = iy0 + r, ix = ix0 + c;
        const unsigned long long v = (iy >= 0 && iy < by && ix >= 0 && ix < bx) ? in[(size_t)iy * bx + ix] : 0ull;
        any |= v != 0ull;
        tile[e] = (double)v;
    }
    for (int e = threadIdx.x; e < ky * kx; e += blockDim.x) tk[e] = kern[e];
    // a map is non-zero only near its ring: tiles whose input is all zero are written as zeros
    const int nonzero = __syncthreads_or(any);
    const int x = x0 + tx, y = y0 + ty;
    if (x >= bx || y >= by) return;
    if (!nonzero) {
        out[(size_t)y * bx + x] = 0.0;
        return;
    }
    double acc = 0.0;
    for (int j = 0; j < ky; j++) {
        const double* row = tile + (ty + ky - 1 - j) * W + tx + kx - 1;
        const double* kr = tk + j * kx;
        for (int i = 0; i < kx; i++) acc = fma(kr[i], row[-i], acc);
    }
    out[(size_t)y * bx + x] = acc * unit * scale;
}

// The same convolution with the kernel width a template parameter (odd, <= 25): one wavefront per 16 x 16 output
// tile, every lane computes BB_RMF_RX consecutive outputs of one row; it holds each input row segment that feeds
// them in registers and adds it through every kernel row, so an input value is read from LDS once per kernel row
// instead of once per tap (the taps are wave-uniform LDS reads).
#define BB_RMF_RX 4
template <int KX>
__global__ void __launch_bounds__(64)
k_rm_finish_t(const unsigned long long* __restrict__ counts, double* __restrict__ maps, int by, int bx,
              const double* __restrict__ kern, int ky, double unit, double scale) {
    extern __shared__ double rm_lds[];
    constexpr int hx = KX / 2, WR = BB_RMF_RX + KX - 1, W = BB_RM_T + KX - 1;
    const int p = blockIdx.z;
    const int hy = ky / 2, H = BB_RM_T + ky - 1;
    const int x0 = blockIdx.x * BB_RM_T, y0 = blockIdx.y * BB_RM_T;
    const unsigned long long* in = counts + (size_t)p * by * bx;
    double* out = maps + (size_t)p * by * bx;
    double* tile = rm_lds;
    double* tk = rm_lds + W * H;
    const int iy0 = y0 + hy - (ky - 1), ix0 = x0 + hx - (KX - 1);
    int any = 0;
    for (int e = threadIdx.x; e < W * H; e += 64) {
        const int r = e / W, c = e % W;
        const int iy = iy0 + r, ix = ix0 + c;
        const unsigned long long v = (iy >= 0 && iy < by && ix >= 0 && ix < bx) ? in[(size_t)iy * bx + ix] : 0ull;
        any |= v != 0ull;
        tile[e] = (double)v;
    }
    for (int e = threadIdx.x; e < ky * KX; e += 64) tk[e] = kern[e];
    const int nonzero = __syncthreads_or(any);
    const int ty = threadIdx.x / (BB_RM_T / BB_RMF_RX), tx = (threadIdx.x % (BB_RM_T / BB_RMF_RX)) * BB_RMF_RX;
    const int y = y0 + ty;
    if (y >= by) return;
    double acc[BB_RMF_RX];
#pragma unroll
    for (int c = 0; c < BB_RMF_RX; c++) acc[c] = 0.0;
    if (nonzero) {
        // out(y, x) = sum_{j,i} K[j][i] tile[ty + ky - 1 - j][tx + c + KX - 1 - i]
        for (int j = 0; j < ky; j++) {
            const double* row = tile + (ty + ky - 1 - j) * W + tx;
            double rv[WR];
#pragma unroll
            for (int w = 0; w < WR; w++) rv[w] = row[w];
            const double* kr = tk + j * KX;
#pragma unroll
            for (int i = 0; i < KX; i++) {
                const double kv = kr[i];
#pragma unroll
                for (int c = 0; c < BB_RMF_RX; c++) acc[c] = fma(kv, rv[c + KX - 1 - i], acc[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < BB_RMF_RX; c++) {
        const int x = x0 + tx + c;
        if (x < bx) out[(size_t)y * bx + x] = acc[c] * unit * scale;
    }
}

// ---------------------------------------------------------------------------------------------------------
// Entry e of the upper triangle (i <= j) of an m x m matrix, row by row.
__device__ __forceinline__ void tri_ij(int e, int m, int& i, int& j) {
    int r = 0, rem = e;
    while (rem >= m - r) {
        rem -= m - r;
        r++;
    }
    i = r;
    j = r + rem;
}

// partial[b][blk][e]: sum over the pixels of block blk of A~_pi A~_pj, A~ = [K_0 w, ..., K_{n-1} w, y w].
__global__ void __launch_bounds__(BB_GR_THREADS)
k_gram(const double* __restrict__ maps, const double* __restrict__ wpix, const double* __restrict__ ypix, int n,
       int npix, int ppb, double* __restrict__ partial) {
    extern __shared__ double gr_lds[];
    __shared__ int act[BB_NN_MAX + 1];          // columns with a non-zero value among the staged pixels
    const int m = n + 1, ne = m * (m + 1) / 2;
    const int b = blockIdx.y, blk = blockIdx.x, nblk = gridDim.x;
    double* st = gr_lds;                        // [BB_GR_PIX][m + 1] (odd stride)
    const int sm = m | 1;
    constexpr int EPT = 8;                      // entries per thread per sweep
    double acc[EPT];
    int ei[EPT], ej[EPT];
    const int p0 = blk * ppb, p1 = min(npix, p0 + ppb);
    for (int sweep = 0; sweep * BB_GR_THREADS * EPT < ne; sweep++) {
#pragma unroll
        for (int k = 0; k < EPT; k++) {
            acc[k] = 0.0;
            const int e = (sweep * EPT + k) * BB_GR_THREADS + threadIdx.x;
            if (e < ne) tri_ij(e, m, ei[k], ej[k]);
            else ei[k] = -1;
        }
        for (int pa = p0; pa < p1; pa += BB_GR_PIX) {
            const int np = min(BB_GR_PIX, p1 - pa);
            __syncthreads();
            for (int c = threadIdx.x; c < m; c += BB_GR_THREADS) act[c] = 0;
            __syncthreads();
            for (int e = threadIdx.x; e < np * m; e += BB_GR_THREADS) {
                const int c = e / np, q = e % np;     // column c, pixel q: coalesced reads of one map
                const int pix = pa + q;
                const double w = wpix[pix];
                double v;
                if (c < n) v = w != 0.0 ? maps[((size_t)b * n + c) * npix + pix] * w : 0.0;
                else v = ypix[pix];
                st[q * sm + c] = v;
                if (v != 0.0) act[c] = 1;
            }
            __syncthreads();
            // (a ring's map is non-zero only near the ring: the products of columns that are zero on all the
            // staged pixels add nothing and are skipped)
#pragma unroll
            for (int k = 0; k < EPT; k++) {
                if (ei[k] < 0 || !act[ei[k]] || !act[ej[k]]) continue;
                const double* a = st + ei[k];
                const double* c = st + ej[k];
                double s = acc[k];
                for (int q = 0; q < np; q++) s = fma(a[q * sm], c[q * sm], s);
                acc[k] = s;
            }
        }
#pragma unroll
        for (int k = 0; k < EPT; k++) {
            const int e = (sweep * EPT + k) * BB_GR_THREADS + threadIdx.x;
            if (e < ne) partial[((size_t)b * nblk + blk) * ne + e] = acc[k];
        }
    }
}

// gram[b] (m x m, symmetric) = sum over the blocks of partial[b], in block order.
__global__ void k_gram_reduce(const double* __restrict__ partial, int n, int nblk, double* __restrict__ gram) {
    const int m = n + 1, ne = m * (m + 1) / 2;
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double s = 0.0;
    for (int k = 0; k < nblk; k++) s += partial[((size_t)b * nblk + k) * ne + e];
    int i, j;
    tri_ij(e, m, i, j);
    gram[(size_t)b * m * m + (size_t)i * m + j] = s;
    gram[(size_t)b * m * m + (size_t)j * m + i] = s;
}

// ---------------------------------------------------------------------------------------------------------
// Non-negative least squares on the normal equations, one wavefront (64 lanes) per system. M (n x n, row
// major, symmetric positive definite) and h (n) in LDS; the result in x (LDS). Lane l owns the entries l and
// l + 64 of every vector; the choices of the active set (the entering index, the step to the boundary) are
// wavefront reductions that pick the same index as a scan in index order (ties to the lowest index).
//
// The passive system M_PP is kept as a Cholesky factor L (L L^T = M_PP) in the order the indices entered: an
// entering index appends one row (a forward substitution), so an iteration of Lawson-Hanson costs O(k^2) instead of
// a factorisation; an index that leaves rebuilds the factor (rare). The substitutions run column by column: the
// pivot of step i is read from its lane (readlane) and every lane updates its own entries with one FMA, so a step
// is one short dependency chain, not a dot product per lane.
struct NnLds {
    double* M;      // n x n system
    double* L;      // n x n: row r of the Cholesky factor of the passive system, in entering order
    double* invd;   // n: 1 / L[r][r]
    double* z;      // n: passive solution
    double* x;      // n: current solution
    double* g;      // n: gradient h - M x
    double* h;      // n
    int* P;         // n: passive flags
    int* ban;       // n: excluded from the next choice
    int* ord;       // n: the index of factor row r
};

// (value, index) of the lane pair with the larger value, ties to the lower index; index -1 loses to any index.
__device__ __forceinline__ void nn_argmax(double& v, int& i) {
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) {
            v = ov;
            i = oi;
        }
    }
}

__device__ __forceinline__ void nn_argmin(double& v, int& i) {
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        if (oi >= 0 && (i < 0 || ov < v || (ov == v && oi < i))) {
            v = ov;
            i = oi;
        }
    }
}

__device__ __forceinline__ double nn_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double nn_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// the value of entry i of a vector held as (v0: entries lane, v1: entries lane + 64), from its lane
__device__ __forceinline__ double nn_get(double v0, double v1, int i) {
    const double v = i < 64 ? v0 : v1;
    const int l = i & 63;
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l), hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Forward substitution L y = b over rows [0, k): b in (b0, b1) (entry r in lane r & 63), y returned in place.
__device__ __forceinline__ void nn_forward(const NnLds& S, int n, int k, int lane, double& b0, double& b1) {
    for (int i = 0; i < k; i++) {
        const double yi = nn_get(b0, b1, i) * S.invd[i];
        if (lane > i && lane < k) b0 -= S.L[(size_t)lane * n + i] * yi;
        if (lane + 64 > i && lane + 64 < k) b1 -= S.L[(size_t)(lane + 64) * n + i] * yi;
        if (lane == i) b0 = yi;
        if (lane + 64 == i) b1 = yi;
    }
}

// Backward substitution L^T z = c over rows [0, k), in place.
__device__ __forceinline__ void nn_backward(const NnLds& S, int n, int k, int lane, double& c0, double& c1) {
    for (int i = k - 1; i >= 0; i--) {
        const double zi = nn_get(c0, c1, i) * S.invd[i];
        const double* Li = S.L + (size_t)i * n;
        if (lane < i) c0 -= Li[lane] * zi;
        if (lane + 64 < i) c1 -= Li[lane + 64] * zi;
        if (lane == i) c0 = zi;
        if (lane + 64 == i) c1 = zi;
    }
}

// Append index j as row k of the factor; false (nothing changed) if its pivot is at the round-off of its own
// computation (16 eps of the diagonal) or below: the column is, to working precision, in the span of the factor's.
__device__ bool nn_append(const NnLds& S, int n, int& k, int lane, int j) {
    double b0 = 0.0, b1 = 0.0;
    if (lane < k) b0 = S.M[(size_t)S.ord[lane] * n + j];
    if (lane + 64 < k) b1 = S.M[(size_t)S.ord[lane + 64] * n + j];
    nn_forward(S, n, k, lane, b0, b1);
    const double d2 = S.M[(size_t)j * n + j] - nn_sum((lane < k ? b0 * b0 : 0.0) + (lane + 64 < k ? b1 * b1 : 0.0));
    if (!(d2 > 16.0 * 2.220446049250313e-16 * S.M[(size_t)j * n + j])) return false;
    const double d = sqrt(d2);
    double* Lk = S.L + (size_t)k * n;
    if (lane < k) Lk[lane] = b0;
    if (lane + 64 < k) Lk[lane + 64] = b1;
    if (lane == 0) {
        Lk[k] = d;
        S.invd[k] = 1.0 / d;
        S.ord[k] = j;
    }
    __syncthreads();
    k++;
    return true;
}

// The factor of the passive set from scratch, in index order; a column dependent on the earlier ones leaves the
// passive set (its value 0: the others span its contribution).
__device__ bool nn_rebuild(const NnLds& S, int n, int& k, int lane) {
    k = 0;
    for (int j = 0; j < n; j++) {
        if (!S.P[j] || nn_append(S, n, k, lane, j)) continue;
        __syncthreads();
        if (lane == (j & 63)) {
            S.P[j] = 0;
            S.x[j] = 0.0;
        }
        __syncthreads();
    }
    return true;
}

// z = M_PP^-1 h_P (0 outside P) from the factor of its k rows.
__device__ void nn_factor_solve(const NnLds& S, int n, int k, int lane) {
    double c0 = lane < k ? S.h[S.ord[lane]] : 0.0, c1 = lane + 64 < k ? S.h[S.ord[lane + 64]] : 0.0;
    nn_forward(S, n, k, lane, c0, c1);
    nn_backward(S, n, k, lane, c0, c1);
    for (int i = lane; i < n; i += 64) S.z[i] = 0.0;
    __syncthreads();
    if (lane < k) S.z[S.ord[lane]] = c0;
    if (lane + 64 < k) S.z[S.ord[lane + 64]] = c1;
    __syncthreads();
}

// Lawson-Hanson active set: returns false if the iteration limits were reached or a pivot failed.
// warm (optional, n): a guess of the positive set. If the passive solution on it is feasible the active set
// starts there (a few iterations when the guess is the solution of a nearby problem), else from x = 0.
__device__ bool nn_solve(const NnLds& S, int n, int lane, const double* warm = nullptr) {
    for (int i = lane; i < n; i += 64) {
        S.x[i] = 0.0;
        S.P[i] = warm ? (warm[i] > 0.0) : 0;
        S.ban[i] = 0;
    }
    __syncthreads();
    int k = 0;
    if (warm) {
        bool ok = nn_rebuild(S, n, k, lane);
        if (ok) nn_factor_solve(S, n, k, lane);
        int bad = 0;
        for (int i = lane; i < n; i += 64)
            if (S.P[i] && !(S.z[i] > 0.0)) bad = 1;
        const bool feas = ok && __ballot(bad) == 0;
        for (int i = lane; i < n; i += 64) {
            if (feas) S.x[i] = S.z[i];
            else S.P[i] = 0;
        }
        if (!feas) k = 0;
        __syncthreads();
    }
    const int maxit = 3 * n + 10;
    int it = 0;
    for (; it < maxit; it++) {
        // gradient h - M x and the size of its round-off
        double amax = 0.0;
        for (int i = lane; i < n; i += 64) {
            double s = S.h[i], a = fabs(S.h[i]);
            const double* Mi = S.M + (size_t)i * n;
            for (int j = 0; j < n; j++) {
                const double t = Mi[j] * S.x[j];
                s -= t;
                a += fabs(t);
            }
            S.g[i] = s;
            amax = fmax(amax, a);
        }
        amax = nn_max(amax);
        // the entering index: the largest gradient above its round-off among the free, unbanned indices
        double gv = 0.0;
        int best = -1;
        const double gb = 64.0 * (double)n * 2.220446049250313e-16 * amax;
        for (int i = lane; i < n; i += 64)
            if (!S.P[i] && !S.ban[i] && S.g[i] > gb && (best < 0 || S.g[i] > gv)) {
                gv = S.g[i];
                best = i;
            }
        nn_argmax(gv, best);
        if (best < 0) break;
        const int j = best;
        if (!nn_append(S, n, k, lane, j)) {
            // dependent on the passive columns: it cannot lower the residual; barred until the solution moves
            if (lane == (j & 63)) S.ban[j] = 1;
            __syncthreads();
            continue;
        }
        if (lane == (j & 63)) S.P[j] = 1;
        __syncthreads();
        for (int inner = 0;; inner++) {
            if (inner >= maxit) return false;
            nn_factor_solve(S, n, k, lane);
            if (inner == 0 && !(S.z[j] > 0.0)) {
                // the new index cannot enter (it is the last row of the factor)
                k--;
                if (lane == (j & 63)) {
                    S.P[j] = 0;
                    S.ban[j] = 1;
                }
                __syncthreads();
                break;
            }
            // the step to the boundary: the first passive index with z <= 0 and the smallest x / (x - z)
            double av = 2.0;
            int imin = -1;
            for (int i = lane; i < n; i += 64)
                if (S.P[i] && !(S.z[i] > 0.0)) {
                    const double a = S.x[i] / (S.x[i] - S.z[i]);
                    if (imin < 0 || a < av) {
                        av = a;
                        imin = i;
                    }
                }
            nn_argmin(av, imin);
            if (imin < 0) {
                for (int i = lane; i < n; i += 64) {
                    S.x[i] = S.z[i];
                    S.ban[i] = 0;
                }
                __syncthreads();
                break;
            }
            const double alpha = av;
            for (int i = lane; i < n; i += 64) {
                S.ban[i] = 0;
                if (!S.P[i]) continue;
                const double xi = S.x[i] + alpha * (S.z[i] - S.x[i]);
                if (i == imin || !(xi > 0.0)) {
                    S.x[i] = 0.0;
                    S.P[i] = 0;
                } else {
                    S.x[i] = xi;
                }
            }
            __syncthreads();
            // indices left the passive set: the factor of the rest
            if (!nn_rebuild(S, n, k, lane)) return false;
        }
    }
    return it < maxit;
}

struct ProfConst {
    int n, nr, nok;
    double smooth, corr, floor, norm, dmed;
};

// M = G + sum_r w2[r] l_r^T l_r over the nrows curvature rows (lcoef / lcol: 3 per row; w2 NULL: wconst2).
__device__ void nn_build(double* __restrict__ M, const double* __restrict__ G, int n, const double* __restrict__ lcoef,
                         const int* __restrict__ lcol, const double* w2, double wconst2, int nrows, int lane) {
    __syncthreads();
    for (int e = lane; e < n * n; e += 64) {
        const int i = e / n, j = e % n;
        double s = G[i * (n + 1) + j];
        for (int r = 0; r < nrows; r++) {
            const int* c = lcol + 3 * r;
            const int a = c[0] == i ? 0 : c[1] == i ? 1 : c[2] == i ? 2 : -1;
            const int d = c[0] == j ? 0 : c[1] == j ? 1 : c[2] == j ? 2 : -1;
            if (a >= 0 && d >= 0) s += (w2 ? w2[r] : wconst2) * lcoef[3 * r + a] * lcoef[3 * r + d];
        }
        M[e] = s;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64)
k_nnls(ProfConst pc, const double* __restrict__ gram, const double* __restrict__ lcoef, const int* __restrict__ lcol,
       const int* __restrict__ near, const double* __restrict__ normv, double* __restrict__ Mall,
       double* __restrict__ xout) {
    extern __shared__ double nn_lds[];
    const int n = pc.n, m = n + 1, lane = threadIdx.x, b = blockIdx.x;
    (void)Mall;
    if (normv) {
        // the prior's rows of this vector (its exterior annulus follows its edge)
        const size_t nrow = n > 2 ? (size_t)(n - 2) : 0;
        lcoef += (size_t)b * 3 * nrow;
        lcol += (size_t)b * 3 * nrow;
        near += (size_t)b * nrow;
        pc.norm = normv[b];
    }
    NnLds L;
    L.M = nn_lds;
    L.L = L.M + n * n;
    L.invd = L.L + n * n;
    L.z = L.invd + n;
    L.x = L.z + n;
    L.g = L.x + n;
    L.h = L.g + n;
    double* w2 = L.h + n;            // n doubles
    double* x0 = w2 + n;             // n doubles
    L.P = (int*)(x0 + n);
    L.ban = L.P + n;
    L.ord = L.ban + n;
    const double* G = gram + (size_t)b * m * m;
    double* M = L.M;
    for (int i = lane; i < n; i += 64) L.h[i] = G[i * m + n];
    __syncthreads();
    bool ok = true;
    if (!(pc.smooth > 0.0) || n < 3) {
        nn_build(M, G, n, lcoef, lcol, nullptr, 0.0, 0, lane);
        ok = nn_solve(L, n, lane);
    } else {
        double tr = 0.0;
        for (int i = 0; i < n; i++) tr += G[i * m + i];
        const double srel = sqrt(0.1 * tr / fmax(pc.norm, 1e-300));
        nn_build(M, G, n, lcoef, lcol, nullptr, srel * srel, n - 2, lane);
        ok = nn_solve(L, n, lane);
        for (int i = lane; i < n; i += 64) x0[i] = L.x[i];
        __syncthreads();
        double xm = 0.0;
        for (int i = 0; i < n - 1; i++) xm += x0[i];
        xm /= (double)(n - 1);
        if (ok && xm > 0.0) {
            // misfit of the first solution: b^T b - 2 h^T x0 + x0^T G x0
            double q = G[n * m + n];
            for (int i = 0; i < n; i++) {
                double gx = 0.0;
                for (int j = 0; j < n; j++) gx += G[i * m + j] * x0[j];
                q += x0[i] * (gx - 2.0 * L.h[i]);
            }
            const int dof = pc.nok - n > 1 ? pc.nok - n : 1;
            const double chi2 = q / (double)dof;
            const double infl = fmax(pc.corr, 1.0) * fmax(chi2, 1.0);
            const double wfl = sqrt(pc.floor) * srel;
            for (int r = lane; r < n - 2; r += 64) {
                double w;
                if (near[r]) {
                    w = srel;
                } else {
                    const double xc = fmax(x0[lcol[3 * r + 1]], 0.1 * xm);
                    w = fmax(sqrt(infl) * pc.dmed * pc.dmed / (pc.smooth * xc), wfl);
                }
                w2[r] = w * w;
            }
            __syncthreads();
            nn_build(M, G, n, lcoef, lcol, w2, 0.0, n - 2, lane);
            ok = nn_solve(L, n, lane, x0);
        }
    }
    for (int i = lane; i < n; i += 64) xout[(size_t)b * n + i] = ok ? L.x[i] : __builtin_nan("");
}

// dens (B, nr) = float32(x[:, :nr]); sr[b, 5, s] = sum_m wval[s][m] dens[b][widx[s][m]] (times ew[b][s] if given).
__global__ void k_set_dens(const double* __restrict__ x, int n, int nr, const int2* __restrict__ widx,
                           const double2* __restrict__ wval, const double* __restrict__ ew, int nsub,
                           double* __restrict__ sr, float* __restrict__ dens) {
    const int b = blockIdx.y;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const double* xb = x + (size_t)b * n;
    if (dens && s < nr) dens[(size_t)b * nr + s] = (float)xb[s];
    if (s >= nsub) return;
    const int2 wi = widx[s];
    const double2 wv = wval[s];
    double d = 0.0;
    if (wi.x >= 0 && wi.x < nr) d += wv.x * (double)(float)xb[wi.x];
    if (wi.y >= 0 && wi.y < nr) d += wv.y * (double)(float)xb[wi.y];
    if (ew) d *= ew[(size_t)b * nsub + s];      // (the vector's edge weight of the density)
    sr[((size_t)b * 16 + 5) * nsub + s] = d;
}

// ---------------------------------------------------------------------------------------------------------
extern "C" size_t bb_rm_finish_lds(int ky, int kx) {
    return (size_t)((BB_RM_T + ky - 1) * (BB_RM_T + kx - 1) + ky * kx) * sizeof(double);
}

extern "C" int bb_launch_rm_finish(const unsigned long long* counts, double* maps, int P, int by, int bx,
                                   const double* kern, int ky, int kx, double unit, double scale, int convolve,
                                   hipStream_t st) {
    dim3 grid((bx + BB_RM_T - 1) / BB_RM_T, (by + BB_RM_T - 1) / BB_RM_T, P);
    if (convolve && kx <= 25 && (kx & 1)) {
        const size_t lt = bb_rm_finish_lds(ky, kx);
        switch (kx) {
#define BB_CASE(K)                                                                                                 \
    case K:                                                                                                        \
        if (lt > 64 * 1024) {                                                                                      \
            hipError_t e = hipFuncSetAttribute((const void*)k_rm_finish_t<K>,                                     \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);            \
            if (e != hipSuccess) return (int)e;                                                                    \
        }                                                                                                          \
        hipLaunchKernelGGL(k_rm_finish_t<K>, grid, dim3(64), lt, st, counts, maps, by, bx, kern, ky, unit, scale); \
        break;
            BB_CASE(1) BB_CASE(3) BB_CASE(5) BB_CASE(7) BB_CASE(9) BB_CASE(11) BB_CASE(13) BB_CASE(15) BB_CASE(17)
            BB_CASE(19) BB_CASE(21) BB_CASE(23) BB_CASE(25)
#undef BB_CASE
        }
        return (int)hipGetLastError();
    }
    const size_t lds = convolve ? bb_rm_finish_lds(ky, kx) : 0;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_rm_finish, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(k_rm_finish, grid, dim3(BB_RM_T * BB_RM_T), lds, st, counts, maps, by, bx, kern, ky, kx, unit,
                       scale, convolve);
    return (int)hipGetLastError();
}

extern "C" int bb_gram_blocks(int npix) { return (npix + 2047) / 2048; }
extern "C" size_t bb_gram_entries(int n) { return (size_t)(n + 1) * (n + 2) / 2; }
extern "C" size_t bb_nnls_lds(int n) { return (size_t)(2 * n * n + 7 * n) * sizeof(double) + 3 * n * sizeof(int) + 16; }
extern "C" int bb_nnls_max() { return BB_NN_MAX; }

extern "C" int bb_launch_profile_solve(int n, int nr, int nok, double smooth, double corr, double floor_, double norm,
                                       double dmed, const double* maps, const double* wpix, const double* ypix,
                                       int npix, const double* lcoef, const int* lcol, const int* near,
                                       const double* normv, int B, double* partial, double* gram, double* Mall,
                                       double* x, hipStream_t st) {
    const int nblk = bb_gram_blocks(npix);
    const int ppb = 2048;
    const int m = n + 1;
    const size_t lds = (size_t)BB_GR_PIX * (m | 1) * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_gram, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(k_gram, dim3(nblk, B), dim3(BB_GR_THREADS), lds, st, maps, wpix, ypix, n, npix, ppb, partial);
    const int ne = (int)bb_gram_entries(n);
    hipLaunchKernelGGL(k_gram_reduce, dim3((ne + 255) / 256, B), dim3(256), 0, st, partial, n, nblk, gram);
    ProfConst pc;
    pc.n = n;
    pc.nr = nr;
    pc.nok = nok;
    pc.smooth = smooth;
    pc.corr = corr;
    pc.floor = floor_;
    pc.norm = norm;
    pc.dmed = dmed;
    const size_t nl = bb_nnls_lds(n);
    if (nl > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_nnls, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(k_nnls, dim3(B), dim3(64), nl, st, pc, gram, lcoef, lcol, near, normv, Mall, x);
    return (int)hipGetLastError();
}

extern "C" int bb_launch_set_dens(const double* x, int B, int n, int nr, const int* widx, const double* wval,
                                  const double* ew, int nsub, double* sr, float* dens, hipStream_t st) {
    const int nt = nsub > nr ? nsub : nr;
    hipLaunchKernelGGL(k_set_dens, dim3((nt + 255) / 256, B), dim3(256), 0, st, x, n, nr, (const int2*)widx,
                       (const double2*)wval, ew, nsub, sr, dens);
    return (int)hipGetLastError();
}
