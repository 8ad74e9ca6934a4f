// GALMOD ANALYTIC on the GPU: clouds on a regular grid in radius / azimuth / height with channel-integrated
// Gaussian line profiles, deposited into a float32 yxz model cube. Follows bbarolo/galmod.py:275-465
// (_zgrid, _galmod_analytic) sample by sample: the same samples, pixel rounding, merging of equal-pixel
// vertical runs and line profiles. Geometry is FP64 (pixel and channel indices round as in the reference);
// the deposited values are summed in FP32.
//
// Three launches per batch of B parameter vectors that share their sub-ring radii:
//   k_an_prep  one thread per (vector, sub-ring): the per-sub-ring constants and the vertical samples
//              (their count, spacing and weights, galmod.py:275-323) into a small table;
//   k_an_bbox  one thread per (vector, chunk of 32 azimuth samples of one sub-ring and radial sample):
//              the pixel bounding box of the chunk (the pixel of a sample is monotone in its height, so the
//              two extreme vertical samples bound it);
//   k_an_tile  one workgroup per (vector, tile of TX x TY pixels): the tile's spectra live in LDS; the
//              workgroup gathers the chunks whose box meets the tile, recomputes their samples (one lane
//              per azimuth) and adds the runs that fall in the tile with LDS atomics, then writes the tile
//              once. No global atomics: every output voxel is written by exactly one workgroup.
// The azimuths come from a table built on the host with the reference's rotation recurrence
// (galmod.py:372-388), so that samples on pixel edges round as in the reference.
#include <hip/hip_runtime.h>

#include "../../include/bbgpu.h"
#include "bb_device.h"

#define BB_CS 32      // azimuth samples per chunk
#define BB_NSF 32     // doubles per sub-ring record
#define BB_ZCAP 64    // vertical weights tabulated per sub-ring (beyond: computed on the fly)
#define BB_RMAX 16    // runs of a sample inside a tile held in registers (beyond: deposited one by one)
#define BB_CB 8       // channels of a line profile computed together
#define BB_TILE_THREADS 768     // 12 wavefronts: 3 per SIMD at the LDS of one workgroup per CU
#define BB_RM_THREADS 768      // ring maps workgroup: 3 wavefronts per SIMD (maps of 256 vectors 21.1 ms against
                               // 23.4 with 512 threads; 1024 measured no faster)
#define BB_RM_WCAP 8           // vertical weights per sub-ring kept in LDS by k_rm_tile (a column of <= 16 samples)

enum {
    F_R0 = 0, F_R1, F_XPOS, F_YPOS, F_SPA, F_CPA, F_SINC, F_CINC, F_VSYS, F_VROT, F_VRAD, F_VVERT, F_SIG, F_K,
    F_FLUX, F_WREC, F_WAPP, F_Z0, F_DVDZ, F_ZCYL, F_ZDEP, F_NZZ, F_NIN, F_STEP, F_ZLIM, F_DENS, F_HALF, F_NAZ
};

// ---------------------------------------------------------------------------------------------------------
// Vertical samples: bin j (0 <= j < half) of one side of the plane spans [lo, hi] (units of z0); sample kz
// sits at -+(lo + hi)/2 of bin j = half-1-kz (kz < half) or kz-half (kz >= half), as _zgrid builds them.
__device__ __forceinline__ void zbin(int j, int nin, double step, double zlim, double& lo, double& hi) {
    lo = j < nin ? (double)j * step : (double)(nin + 4 * (j - nin)) * step;
    hi = lo + (j < nin ? step : 4.0 * step);
    if (hi > zlim) hi = zlim;
}

__device__ __forceinline__ int zbin_of(int kz, int half) { return kz >= half ? kz - half : half - 1 - kz; }

__device__ __forceinline__ double zq_at(int kz, int half, int nin, double step, double zlim) {
    double lo, hi;
    zbin(zbin_of(kz, half), nin, step, zlim, lo, hi);
    double z = 0.5 * (lo + hi);
    return kz >= half ? z : -z;
}

// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_an_prep(bb_grid g, bb_box box, const double* __restrict__ srall, int nsub, const int* __restrict__ naz_tab,
          double* __restrict__ suball, double* __restrict__ wtaball, int unit_dens, const int4* __restrict__ vbox) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (s >= nsub) return;
    if (vbox) {
        // the vector's own model box: the extent of its vertical samples depends on it (galmod.py:295-300)
        const int4 v = vbox[b];
        box.blo0 = v.x;
        box.blo1 = v.y;
        box.bhi0 = v.z;
        box.bhi1 = v.w;
    }
    const double* sr = srall + (size_t)b * 16 * nsub;
    double* q = suball + ((size_t)b * nsub + s) * BB_NSF;
    double* wt = wtaball + ((size_t)b * nsub + s) * BB_ZCAP;
    auto S = [&](int row) { return sr[(size_t)row * nsub + s]; };
    const double R0 = S(0), dens = unit_dens ? 1.0 : S(5);    // (unit_dens: the ring maps)
    // azimuths of the sub-ring: the plan's, or from its radius as the plan makes them (a table whose radii differ per
    // vector: the exterior annulus with a free edge)
    int64_t naz;
    if (naz_tab) {
        naz = naz_tab[s];
    } else {
        naz = (int64_t)(3 * (2 * BB_PI * R0 / sqrt(g.pixarea)) + 0.5);
        naz = naz < 16 ? 16 : naz > (1 << 24) ? (1 << 24) : naz;
    }
    q[F_NAZ] = (double)naz;
    const double ring_area_pix = 2 * BB_PI * R0 * g.radsep / g.pixarea;
    const double r_inner = R0 - 0.5 * g.radsep;
    q[F_R0] = sqrt(r_inner * r_inner + 2.0 * g.radsep * R0 * 0.25);
    q[F_R1] = sqrt(r_inner * r_inner + 2.0 * g.radsep * R0 * 0.75);
    double sig = sqrt(S(4) * S(4) + g.sig_instr * g.sig_instr);
    double smin = 1e-6 * fabs(g.chmin);
    sig = sig < smin ? smin : sig;
    q[F_SIG] = sig;
    q[F_K] = (double)((int64_t)(5.0 * sig / g.chmin) + 2);
    const double inc = S(9), phi = S(10);
    const double sinc = sin(inc), cinc = cos(inc);
    q[F_SINC] = sinc;
    q[F_CINC] = cinc;
    q[F_SPA] = sin(phi) * cos(g.crota2) + cos(phi) * sin(g.crota2);
    q[F_CPA] = cos(phi) * cos(g.crota2) - sin(phi) * sin(g.crota2);
    q[F_XPOS] = S(11);
    q[F_YPOS] = S(12);
    q[F_VSYS] = S(13);
    q[F_VROT] = S(1);
    q[F_VRAD] = S(2);
    q[F_VVERT] = S(3);
    q[F_WREC] = S(14);
    q[F_WAPP] = S(15);
    q[F_DENS] = dens;
    q[F_FLUX] = dens * ring_area_pix * g.pixarea / (double)(naz * 2);
    const double z0 = S(6), dvdz = S(7), zcyl = S(8);
    q[F_Z0] = z0;
    q[F_DVDZ] = dvdz;
    q[F_ZCYL] = zcyl;
    // vertical samples (galmod.py:295-323)
    const double sa = fabs(sinc);
    int nin = 1, half = 0, nzz = 1;
    double step = 0., zlim = 0.;
    if (z0 > 0 && g.nzq >= 2 && sa >= 1e-6) {
        const int nmax = g.nzq;
        double pixsize = sqrt(g.pixarea);
        double xpos = S(11), ypos = S(12);
        double dx = fmax(fabs((double)box.blo0 - 0.5 - xpos), fabs((double)box.bhi0 - 0.5 - xpos)) * fabs(g.cdelt0);
        double dy = fmax(fabs((double)box.blo1 - 0.5 - ypos), fabs((double)box.bhi1 - 0.5 - ypos)) * fabs(g.cdelt1);
        double zl2 = (R0 + g.radsep + sqrt(dx * dx + dy * dy)) / (sa * z0);
        double ext = bb_layer_extent(g.ltype);
        zlim = ext < zl2 ? ext : zl2;
        step = g.zstep * pixsize / (z0 * sa);
        if (dvdz != 0) {
            double st2 = 0.5 * sig / (fabs(dvdz) * z0 * sa);
            if (st2 < step) step = st2;
        }
        double zc = bb_layer_core(g.ltype);
        double m = zc < zlim ? zc : zlim;
        // (a few passes for finite inputs; the host rejects non-finite tables)
        if (!(step > 0.0) || !(m / step < 1e9)) step = m / 1e9;
        int64_t nin64 = 1, nout = 0;
        for (int it = 0; it < 4096; it++) {
            nin64 = (int64_t)ceil(m / step);
            if (nin64 < 1) nin64 = 1;
            double no = ceil((zlim - (double)nin64 * step) / (4.0 * step));
            nout = no > 0 ? (int64_t)fmin(no, 1e9) : 0;
            if (2 * (nin64 + nout) <= nmax) break;
            step *= 1.01 * (double)(nin64 + nout) / (double)(nmax / 2);
        }
        if (2 * (nin64 + nout) > nmax) {
            nin64 = nmax / 2;
            nout = 0;
        }
        nin = (int)nin64;
        half = (int)(nin64 + nout);
        nzz = 2 * half;
        for (int j = 0; j < half && j < BB_ZCAP; j++) {
            double lo, hi;
            zbin(j, nin, step, zlim, lo, hi);
            wt[j] = bb_layer_cdf(g.ltype, hi) - bb_layer_cdf(g.ltype, lo);
        }
    } else {
        wt[0] = 1.0;
    }
    q[F_NZZ] = nzz;
    q[F_NIN] = nin;
    q[F_HALF] = half;
    q[F_STEP] = step;
    q[F_ZLIM] = zlim;
    // does the velocity depend on height for some sample (galmod.py:369-374)? |z| is largest in the last bin
    bool zdep = false;
    if (dvdz != 0 && nzz > 1) zdep = fabs(zq_at(nzz - 1, half, nin, step, zlim) * z0) > zcyl;
    q[F_ZDEP] = zdep ? 1.0 : 0.0;
}

__device__ __forceinline__ double z_at(const double* __restrict__ q, int kz) {
    const int nzz = (int)q[F_NZZ];
    if (nzz <= 1) return 0.0;
    return zq_at(kz, (int)q[F_HALF], (int)q[F_NIN], q[F_STEP], q[F_ZLIM]) * q[F_Z0];
}

// Pixel of a sample at height z (galmod.py:393-398).
__device__ __forceinline__ void pix_at(const double* __restrict__ q, const bb_grid& g, double xs, double xc, double yb,
                                       double z, int64_t& g0, int64_t& g1) {
    double y = yb - z * q[F_SINC];
    g0 = bb_lround(q[F_XPOS] + (xs - y * q[F_CPA]) / g.cdelt0);
    g1 = bb_lround(q[F_YPOS] + (xc + y * q[F_SPA]) / g.cdelt1);
}

// ---------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_an_bbox(bb_grid g, bb_box box, const int4* __restrict__ chunks, int nchunks, const double2* __restrict__ az,
          const double* __restrict__ suball, int nsub, short4* __restrict__ bball, short4* __restrict__ sball,
          const int4* __restrict__ vbox) {
    // one thread per azimuth sample (BB_CS lanes per chunk): the box of its vertical segment, then the box of
    // the chunk by a reduction over its lanes
    const int t = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    const int c = t / BB_CS, lane = t % BB_CS;
    const bool valid = c < nchunks;
    const int4 ch = valid ? chunks[c] : make_int4(0, 0, 0, 0);
    const int s = ch.x >> 1, kr = ch.x & 1;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    // the clip: the layout box, or the vector's own model box inside it (box-relative, inclusive)
    int cx0 = 0, cx1 = bx - 1, cy0 = 0, cy1 = by - 1;
    if (vbox) {
        const int4 v = vbox[b];
        cx0 = max(cx0, v.x - box.blo0);
        cx1 = min(cx1, v.z - 1 - box.blo0);
        cy0 = max(cy0, v.y - box.blo1);
        cy1 = min(cy1, v.w - 1 - box.blo1);
    }
    int x0 = 1 << 30, x1 = -(1 << 30), y0 = 1 << 30, y1 = -(1 << 30);
    if (valid && lane < ch.z) {
        const double* q = suball + ((size_t)b * nsub + s) * BB_NSF;
        if (q[F_DENS] != 0) {
            const double R = kr ? q[F_R1] : q[F_R0];
            const int nzz = (int)q[F_NZZ];
            const double2 cs = az[ch.w + lane];
            const double caz = cs.x, saz = cs.y;
            const double wside = caz >= 0.0 ? q[F_WREC] : q[F_WAPP];
            if (wside != 0.0) {
                const double x = R * caz, yb = R * saz * q[F_CINC], xs = x * q[F_SPA], xc = x * q[F_CPA];
                int64_t a0, a1, e0, e1;
                pix_at(q, g, xs, xc, yb, z_at(q, 0), a0, a1);
                pix_at(q, g, xs, xc, yb, z_at(q, nzz - 1), e0, e1);
                // box-relative, clipped to the box (the pixel of a sample is monotone in its height)
                int64_t lx = (a0 < e0 ? a0 : e0) - box.blo0, hx = (a0 > e0 ? a0 : e0) - box.blo0;
                int64_t ly = (a1 < e1 ? a1 : e1) - box.blo1, hy = (a1 > e1 ? a1 : e1) - box.blo1;
                lx = lx < cx0 ? cx0 : lx;
                ly = ly < cy0 ? cy0 : ly;
                hx = hx > cx1 ? cx1 : hx;
                hy = hy > cy1 ? cy1 : hy;
                if (lx <= hx && ly <= hy) {
                    x0 = (int)lx;
                    x1 = (int)hx;
                    y0 = (int)ly;
                    y1 = (int)hy;
                }
            }
        }
    }
    if (valid) sball[((size_t)b * nchunks + c) * BB_CS + lane] = x0 <= x1 ? make_short4((short)x0, (short)x1, (short)y0, (short)y1)
                                                                          : make_short4(1, 0, 1, 0);
    for (int off = BB_CS / 2; off > 0; off >>= 1) {
        const int ox0 = __shfl_xor(x0, off, BB_CS), ox1 = __shfl_xor(x1, off, BB_CS);
        const int oy0 = __shfl_xor(y0, off, BB_CS), oy1 = __shfl_xor(y1, off, BB_CS);
        x0 = ox0 < x0 ? ox0 : x0;
        x1 = ox1 > x1 ? ox1 : x1;
        y0 = oy0 < y0 ? oy0 : y0;
        y1 = oy1 > y1 ? oy1 : y1;
    }
    if (valid && lane == 0)
        bball[(size_t)b * nchunks + c] = x0 <= x1 ? make_short4((short)x0, (short)x1, (short)y0, (short)y1)
                                                  : make_short4(1, 0, 1, 0);
}

// ---------------------------------------------------------------------------------------------------------
// The chunks of every pixel tile (a counting sort of the chunk boxes by the tiles they meet), so that a tile
// reads its own list instead of testing every chunk box of its vector. Per vector: cnt (ntiles) counts, then offs
// (ntiles + 1) exclusive offsets with offs[ntiles] the total, list (cap) the chunk indices. A vector whose total
// exceeds cap gets offs[ntiles] = -1: its tiles test all the chunks.
__global__ void __launch_bounds__(256)
k_tile_zero(int* __restrict__ cnt, int* __restrict__ cur, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        cnt[i] = 0;
        cur[i] = 0;
    }
}

template <bool FILL>
__global__ void __launch_bounds__(256)
k_tile_bin(const short4* __restrict__ bball, int nchunks, int TX, int TY, int ntx, int ntiles, int* __restrict__ cnt,
           const int* __restrict__ offs, int* __restrict__ cur, int* __restrict__ list, int cap) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (c >= nchunks) return;
    const short4 r = bball[(size_t)b * nchunks + c];
    if (r.x > r.y) return;                               // (empty box)
    const int ntyv = ntiles / ntx;
    const int qx0 = r.x / TX, qx1 = r.y / TX, qy0 = r.z / TY, qy1 = r.w / TY;
    if (FILL && offs[(size_t)b * (ntiles + 1) + ntiles] < 0) return;
    for (int qy = qy0; qy <= qy1 && qy < ntyv; qy++)
        for (int qx = qx0; qx <= qx1 && qx < ntx; qx++) {
            const int t = qy * ntx + qx;
            if (!FILL) {
                atomicAdd(&cnt[(size_t)b * ntiles + t], 1);
            } else {
                const int pos = offs[(size_t)b * (ntiles + 1) + t] + atomicAdd(&cur[(size_t)b * ntiles + t], 1);
                list[(size_t)b * cap + pos] = c;
            }
        }
}

// offs = exclusive scan of cnt per vector (one block per vector), offs[ntiles] = total or -1 beyond cap.
__global__ void __launch_bounds__(256)
k_tile_scan(const int* __restrict__ cnt, int ntiles, int cap, int* __restrict__ offs) {
    __shared__ int part[256];
    const int b = blockIdx.x;
    const int* cb = cnt + (size_t)b * ntiles;
    int* ob = offs + (size_t)b * (ntiles + 1);
    const int per = (ntiles + 255) / 256, t0 = threadIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    int s = 0;
    for (int t = t0; t < t1; t++) s += cb[t];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < 256; i++) {
            const int v = part[i];
            part[i] = acc;
            acc += v;
        }
        ob[ntiles] = acc <= cap ? acc : -1;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int t = t0; t < t1; t++) {
        ob[t] = acc;
        acc += cb[t];
    }
}

// the chunk lists of the tiles of TX x TY pixels of the box (bin: ntiles x 3 + 1 ints and cap list entries per
// vector, laid out as cnt | cur | offs | list)
static int bin_chunks(bb_box box, int B, const short4* bbox, int nchunks, int TX, int TY, int* bin, int cap,
                      hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    const int ntx = (bx + TX - 1) / TX, ntiles = ntx * ((by + TY - 1) / TY);
    int* cnt = bin;
    int* cur = cnt + (size_t)B * ntiles;
    int* offs = cur + (size_t)B * ntiles;
    int* list = offs + (size_t)B * (ntiles + 1);
    const int nz = B * ntiles;
    hipLaunchKernelGGL(k_tile_zero, dim3((nz + 255) / 256), dim3(256), 0, st, cnt, cur, nz);
    hipLaunchKernelGGL(k_tile_bin<false>, dim3((nchunks + 255) / 256, B), dim3(256), 0, st, bbox, nchunks, TX, TY, ntx,
                       ntiles, cnt, offs, cur, list, cap);
    hipLaunchKernelGGL(k_tile_scan, dim3(B), dim3(256), 0, st, cnt, ntiles, cap, offs);
    hipLaunchKernelGGL(k_tile_bin<true>, dim3((nchunks + 255) / 256, B), dim3(256), 0, st, bbox, nchunks, TX, TY, ntx,
                       ntiles, cnt, offs, cur, list, cap);
    return (int)hipGetLastError();
}

// The bin's layout in a kernel: the list of the tile of vector b, [lo, hi) (tlist NULL: no bin, or it overflowed).
__device__ __forceinline__ void tile_list(const int* bin, int B, int ntiles, int cap, int b, int tile,
                                          const int*& tlist, int& lo, int& hi) {
    tlist = nullptr;
    lo = hi = 0;
    if (!bin) return;
    const int* offs = bin + (size_t)2 * B * ntiles + (size_t)b * (ntiles + 1);
    if (offs[ntiles] < 0) return;
    tlist = bin + (size_t)2 * B * ntiles + (size_t)B * (ntiles + 1) + (size_t)b * cap;
    lo = offs[tile];
    hi = offs[tile + 1];
}

// ---------------------------------------------------------------------------------------------------------
// The deposit of one tile. Per-sub-ring constants are held in registers (Sub), the vertical weights of the
// chunk being processed, the channel edges and the channel factors in LDS; the line profile is built from
// the Phi table (galmod.py:183-214) and added to the LDS spectra of the runs kept.
struct Sub {
    double R, xpos, ypos, spa, cpa, sinc, cinc, vsys, vrot, vrad, vvert, sig, flux, wrec, wapp, z0, dvdz, zcyl, step,
        zlim;
    int64_t K;
    int zdep, nzz, nin, half;
};

__device__ __forceinline__ void load_sub(Sub& u, const double* __restrict__ q, int kr) {
    u.R = kr ? q[F_R1] : q[F_R0];
    u.xpos = q[F_XPOS];
    u.ypos = q[F_YPOS];
    u.spa = q[F_SPA];
    u.cpa = q[F_CPA];
    u.sinc = q[F_SINC];
    u.cinc = q[F_CINC];
    u.vsys = q[F_VSYS];
    u.vrot = q[F_VROT];
    u.vrad = q[F_VRAD];
    u.vvert = q[F_VVERT];
    u.sig = q[F_SIG];
    u.flux = q[F_FLUX];
    u.wrec = q[F_WREC];
    u.wapp = q[F_WAPP];
    u.z0 = q[F_Z0];
    u.dvdz = q[F_DVDZ];
    u.zcyl = q[F_ZCYL];
    u.step = q[F_STEP];
    u.zlim = q[F_ZLIM];
    u.K = (int64_t)q[F_K];
    u.zdep = q[F_ZDEP] != 0.0;
    u.nzz = (int)q[F_NZZ];
    u.nin = (int)q[F_NIN];
    u.half = (int)q[F_HALF];
}

__device__ __forceinline__ double uz(const Sub& u, int kz) {
    return u.nzz <= 1 ? 0.0 : zq_at(kz, u.half, u.nin, u.step, u.zlim) * u.z0;
}

__device__ __forceinline__ void upix(const Sub& u, const bb_grid& g, double xs, double xc, double yb, double z,
                                     int64_t& g0, int64_t& g1) {
    double y = yb - z * u.sinc;
    g0 = bb_lround(u.xpos + (xs - y * u.cpa) / g.cdelt0);
    g1 = bb_lround(u.ypos + (xc + y * u.spa) / g.cdelt1);
}

// The pixel track of an azimuth sample is linear in its height: X(z) = X0 + z dX, Y(z) = Y0 + z dY (the same
// sums as upix, regrouped). Its value is within ~1e-12 pixel of upix's, so it rounds to the same pixel except
// within 1e-9 of a half-integer, where upix computes the pixel exactly as the reference does.
struct Track {
    double X0, dX, Y0, dY;
};

__device__ __forceinline__ Track utrack(const Sub& u, const bb_grid& g, double xs, double xc, double yb) {
    Track t;
    const double i0 = 1.0 / g.cdelt0, i1 = 1.0 / g.cdelt1;     // (an estimate: products, not quotients)
    t.X0 = u.xpos + (xs - yb * u.cpa) * i0;
    t.dX = u.sinc * u.cpa * i0;
    t.Y0 = u.ypos + (xc + yb * u.spa) * i1;
    t.dY = -u.sinc * u.spa * i1;
    return t;
}

// The track in FP32 relative to a tile's first pixel (ox, oy), with the heights in FP32 (zq_at in FP32: within
// 4e-7 of it, relative). While the column stays within 128 pixels of the tile (ok), the FP32 track is within
// 7e-5 pixel of the FP64 one, so a pixel it rounds to with more than 1e-4 to spare is the FP64 track's (and the
// reference's); otherwise the FP64 path decides (upix_ff). Half the cost of the FP64 walk (FP32 fma and rint run at
// twice the FP64 rate), the same pixels.
struct TrackF {
    float X0, dX, Y0, dY, step, zlim, z0;
    int ox, oy;
    bool ok;
};

__device__ __forceinline__ TrackF utrack_f(const Sub& u, const Track& t, int ox, int oy) {
    TrackF f;
    const double x0 = t.X0 - (double)ox, y0 = t.Y0 - (double)oy;
    const double zm = u.nzz > 1 ? u.zlim * u.z0 : 0.0;
    f.ok = fabs(x0) + fabs(t.dX) * zm < 128.0 && fabs(y0) + fabs(t.dY) * zm < 128.0;
    f.X0 = (float)x0;
    f.dX = (float)t.dX;
    f.Y0 = (float)y0;
    f.dY = (float)t.dY;
    f.step = (float)u.step;
    f.zlim = (float)u.zlim;
    f.z0 = (float)u.z0;
    f.ox = ox;
    f.oy = oy;
    return f;
}

__device__ __forceinline__ float uz_f(const Sub& u, const TrackF& f, int kz) {
    if (u.nzz <= 1) return 0.f;
    const int j = zbin_of(kz, u.half);
    const float lo = j < u.nin ? (float)j * f.step : (float)(u.nin + 4 * (j - u.nin)) * f.step;
    float hi = lo + (j < u.nin ? f.step : 4.0f * f.step);
    if (hi > f.zlim) hi = f.zlim;
    const float z = 0.5f * (lo + hi) * f.z0;
    return kz >= u.half ? z : -z;
}

// A pixel index as int32: every box pixel fits; an index beyond +-2^30 (a sample far outside) becomes -2^30,
// outside every tile.
__device__ __forceinline__ int pix32(int64_t a) { return a > (1 << 30) || a < -(1 << 30) ? -(1 << 30) : (int)a; }

// The pixel of the track at height z: the rounded estimate unless it is within 1e-9 of a half-integer (or
// huge), where upix places the sample exactly as the reference does.
__device__ __forceinline__ void upix_fast(const Sub& u, const bb_grid& g, const Track& t, double xs, double xc,
                                          double yb, double z, int& g0, int& g1) {
    const double X = fma(z, t.dX, t.X0), Y = fma(z, t.dY, t.Y0);
    const double rx = rint(X), ry = rint(Y);
    const double lim = 0.5 - 1e-9;
    if (fabs(X) < 1e9 && fabs(Y) < 1e9 && fabs(X - rx) < lim && fabs(Y - ry) < lim) {
        g0 = (int)rx;
        g1 = (int)ry;
    } else {
        int64_t a, b;
        upix(u, g, xs, xc, yb, z, a, b);
        g0 = pix32(a);
        g1 = pix32(b);
    }
}

// The pixel of vertical sample kz from the FP32 track, or from upix_fast when that is not decisive.
__device__ __forceinline__ void upix_ff(const Sub& u, const bb_grid& g, const Track& t, const TrackF& f, double xs,
                                        double xc, double yb, int kz, int& g0, int& g1) {
    const float z = uz_f(u, f, kz);
    const float X = fmaf(z, f.dX, f.X0), Y = fmaf(z, f.dY, f.Y0);
    const float rx = rintf(X), ry = rintf(Y);
    if (f.ok && fabsf(X - rx) < 0.4999f && fabsf(Y - ry) < 0.4999f) {
        g0 = (int)rx + f.ox;
        g1 = (int)ry + f.oy;
    } else {
        upix_fast(u, g, t, xs, xc, yb, uz(u, kz), g0, g1);
    }
}

// ---------------------------------------------------------------------------------------------------------
// Accumulation in the LDS tile is fixed point: every contribution x (>= 0) is stored as round(x / q) in a
// 64-bit integer, q = amax / 2^31 with amax a bound of any single contribution of the parameter vector
// (largest relint x flux x side weight over its sub-rings, times the largest channel factor), so one
// contribution is < 2^31 units and a voxel holds 2^33 of them. Integer LDS atomics run at full rate on
// gfx950 (float ones are serialised per lane), and the sums are exact: the model does not depend on the
// order of the atomics.
struct Acc {
    unsigned long long* tile;
    float inv_q;        // 1 / q
    int nzp;
    __device__ __forceinline__ void add(int slot, int k, float x) const {
        const unsigned int u = (unsigned int)fmaf(x, inv_q, 0.5f);
        atomicAdd(tile + slot * nzp + k, (unsigned long long)u);
    }
};

// Phi(t) = erfc(-t / sqrt 2) / 2, the cumulative normal of the line profile (galmod.py:183-214 tabulates it at 1/512
// on [-6, 6] and interpolates linearly, an error of up to 1.2e-7), here in registers by Abramowitz & Stegun 7.1.26:
// erfc(x) = tau (a1 + tau (a2 + tau (a3 + tau (a4 + tau a5)))) exp(-x^2), tau = 1 / (1 + p x), x >= 0, |error| <=
// 1.5e-7. One reciprocal, one exponential and five fmas instead of two LDS reads at a random address: the deposit's
// LDS serves its atomics only (61.2 -> 56.6 ms for the 256-vector batch of the 160x160x96 cube).
__device__ __forceinline__ float phi_as(float t) {
    const float x = fabsf(t) * 0.70710678f;
    const float tau = __builtin_amdgcn_rcpf(fmaf(0.3275911f, x, 1.0f));
    float p = fmaf(tau, 1.061405429f, -1.453152027f);
    p = fmaf(tau, p, 1.421413741f);
    p = fmaf(tau, p, -0.284496736f);
    p = fmaf(tau, p, 0.254829592f);
    const float q = 0.5f * tau * p * __expf(-x * x);
    return t < 0.0f ? q : 1.0f - q;
}

// aw x the channel-integrated line of centre v (galmod.py:194-214) into pixel slot of the tile: the rare
// paths (a height-dependent velocity, more runs of a sample in the tile than BB_RMAX). Scalar arguments only:
// an argument passed by reference to a call that is not inlined would have to live in scratch memory.
__device__ __noinline__ void acc_line(unsigned long long* tile, float inv_q, int nzp, int slot, double aw, double v,
                                      double sig, const double* ve, const double* cd, int64_t kc, int64_t K, int nz) {
    int64_t k0 = kc - K;
    if (k0 < 0) k0 = 0;
    int64_t k1 = kc + K + 1;
    if (k1 > nz) k1 = nz;
    if (k1 <= k0) return;
    const double inv = 1.0 / sig;
    float pp = phi_as((float)((ve[k0] - v) * inv));
    for (int k = (int)k0; k < (int)k1; k++) {
        const float pn = phi_as((float)((ve[k + 1] - v) * inv));
        const float f = pn - pp;
        pp = pn;
        const unsigned int uq = (unsigned int)fmaf((float)aw * (fabsf(f) * (float)cd[k]), inv_q, 0.5f);
        atomicAdd(tile + slot * nzp + k, (unsigned long long)uq);
    }
}

__device__ __noinline__ double uw_cold(int j, int nin, double step, double zlim, int ltype) {
    double lo, hi;
    zbin(j, nin, step, zlim, lo, hi);
    return bb_layer_cdf(ltype, hi) - bb_layer_cdf(ltype, lo);
}

// weight of vertical sample kz: the sub-ring's table (BB_ZCAP bins, FP64), else the layer CDF
__device__ __forceinline__ double uw_(const Sub& u, const double* __restrict__ wt, int kz, int ltype) {
    if (u.nzz <= 1) return 1.0;
    const int j = zbin_of(kz, u.half);
    if (j < BB_ZCAP) return wt[j];
    return uw_cold(j, u.nin, u.step, u.zlim, ltype);
}

// The lanes of a wavefront take samples BB_SSTRIDE entries apart in a window of 64 BB_SSTRIDE (neighbouring azimuths of a
// sub-ring, about two pixels apart): far enough apart that their atomics seldom meet the same pixel, near enough that
// their columns cross the tile alike (runs per sample), 51.6 ms for the 256-vector batch against 52.5 at 4 and 52.9
// at 12 (the whole window); 64 consecutive samples: 70.4
#define BB_SSTRIDE 6
#define BB_LIST_CAP 1024     // chunks meeting the tile, gathered before their samples are tested
#define BB_AN_SAMP_CAP 8192  // samples meeting the tile, gathered before they are deposited (ANALYTIC deposit; the
                             // LDS it takes is worth more here than float copies of the vertical weights: 66.1 ms
                             // against 68.0 with 4096 samples and the weights in LDS)
#define BB_AN_WSAMP_CAP 2048 // the same for the wavefront deposit
#define BB_TILE_PAD 64       // words after the LDS tile: the wavefront deposit's lanes past a line window add 0 there
#define BB_RM_SAMP_CAP 3072  // the same for the ring maps (a longer list is slower there: 24.9 ms against 23.5)

// Odd stride of the pixel spectra in the LDS tile: consecutive pixels start in different banks.
__device__ __host__ __forceinline__ int bb_tile_stride(int nz) { return nz | 1; }

struct TileGeo {
    int ax0, ax1, ay0, ay1;     // the pixels of the tile that receive samples, absolute (inclusive)
    int ox, oy;                 // the tile's first pixel: slot (0, 0) of its LDS spectra
    int TX;
    __device__ __forceinline__ int slot(int p0, int p1) const {
        if (p0 < ax0 || p0 > ax1 || p1 < ay0 || p1 > ay1) return -1;
        return (p1 - oy) * TX + (p0 - ox);
    }
    __device__ __forceinline__ int slot(int64_t p0, int64_t p1) const { return slot(pix32(p0), pix32(p1)); }
};

// A height-dependent velocity (galmod.py:432-465): every vertical pair has its own line.
__device__ __forceinline__ void an_sample_lag(const bb_grid& g, const bb_box& box, const TileGeo& T, const Acc& A,
                                           const double* sve, const double* scd,
                                           const Sub& u, const double* __restrict__ wt, double caz,
                                           double saz) {
    const int nz = g.nz;
    const double wside = caz >= 0.0 ? u.wrec : u.wapp;
    const double x = u.R * caz, yb = u.R * saz * u.cinc, xs = x * u.spa, xc = x * u.cpa;
    const double vplane = u.vsys + (u.vrot * caz + u.vrad * saz) * u.sinc + u.vvert * u.cinc;
    const double afl = u.flux * wside;
    for (int kz = 0; kz < (u.nzz + 1) / 2; kz++) {
        const int kz2 = u.nzz - 1 - kz;
        const double za = uz(u, kz), zb = uz(u, kz2);
        int64_t ga0, ga1, gb0, gb1;
        upix(u, g, xs, xc, yb, za, ga0, ga1);
        upix(u, g, xs, xc, yb, zb, gb0, gb1);
        const bool in1 = !(ga0 < box.blo0 || ga0 >= box.bhi0 || ga1 < box.blo1 || ga1 >= box.bhi1);
        const bool in2 = kz2 != kz && !(gb0 < box.blo0 || gb0 >= box.bhi0 || gb1 < box.blo1 || gb1 >= box.bhi1);
        const int s1 = in1 ? T.slot(ga0, ga1) : -1;
        const int s2 = in2 ? T.slot(gb0, gb1) : -1;
        if (s1 < 0 && s2 < 0) continue;
        const double az_ = fabs(za);
        double v;
        if (az_ > u.zcyl)
            v = u.vsys + (bb_lagged(u.vrot, u.dvdz, az_ - u.zcyl) * caz + u.vrad * saz) * u.sinc + u.vvert * u.cinc;
        else
            v = vplane;
        const double w1 = uw_(u, wt, kz, g.ltype), w2 = uw_(u, wt, kz2, g.ltype);
        for (int nl = 0; nl < g.nlines; nl++) {
            const double* ve = sve + (size_t)nl * (nz + 1);
            const double a = g.relint[nl] * afl;
            const int64_t kc = bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
            if (in1 && in2 && ga0 == gb0 && ga1 == gb1) {
                acc_line(A.tile, A.inv_q, A.nzp, s1, a * (w1 + w2), v, u.sig, ve, scd, kc, u.K, nz);
                continue;
            }
            if (s1 >= 0) acc_line(A.tile, A.inv_q, A.nzp, s1, a * w1, v, u.sig, ve, scd, kc, u.K, nz);
            if (s2 >= 0) acc_line(A.tile, A.inv_q, A.nzp, s2, a * w2, v, u.sig, ve, scd, kc, u.K, nz);
        }
    }
}

// The vertical samples of one azimuth sample that can fall in the tile, [ka, kb]; false if none. Only they are
// placed exactly: a continuous estimate of the (monotone) pixel track picks them, widened by one sample on each
// side, and those neighbours are outside the tile, so every run of equal pixels inside it is complete.
__device__ __forceinline__ bool tile_zrange(const Sub& u, const Track& tr, const TileGeo& T, int& ka, int& kb) {
    ka = 0;
    kb = u.nzz - 1;
    if (u.nzz > 2) {
        const double X0 = tr.X0, dX = tr.dX, Y0 = tr.Y0, dY = tr.dY;
        const double m = 1e-6;
        int first = u.nzz, last = -1;
        for (int kz = 0; kz < u.nzz; kz++) {
            const double z = uz(u, kz), X = X0 + z * dX, Y = Y0 + z * dY;
            if (X >= (double)T.ax0 - 0.5 - m && X < (double)T.ax1 + 0.5 + m && Y >= (double)T.ay0 - 0.5 - m &&
                Y < (double)T.ay1 + 0.5 + m) {
                first = kz < first ? kz : first;
                last = kz;
            }
        }
        if (last < 0) return false;
        ka = first > 0 ? first - 1 : 0;
        kb = last < u.nzz - 1 ? last + 1 : u.nzz - 1;
    }
    return true;
}

// The line profile of one sample (galmod.py:411-431), shared by its runs in the tile: the runs rs[0..nr) (tile
// slots) with their vertical weights rw[0..nr); the loop over r stops at the first r that no lane of the
// wavefront has. LIN: a linear velocity axis (AXTYP 4) with one channel factor for all channels: the
// channel edges are computed as the host computes them, ((k - 1/2) + 1 - crpix3) cdelt3 + crval3 in FP64 (the
// same operations, the same values), and the channel factor is a constant, instead of being read from LDS.
template <bool LIN>
__device__ __forceinline__ void an_lines(const bb_grid& g, const Acc& A, const double* sve, const float* scf,
                                         const Sub& u, double vplane, double afl,
                                         const int* rs, const float* rw, int nr) {
    const int nz = g.nz;
    for (int nl = 0; nl < g.nlines; nl++) {
        const double* ve = sve + (size_t)nl * (nz + 1);
        const float a = (float)(g.relint[nl] * afl);
        const int64_t kc = bb_lround(bb_velgrid(vplane, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
        int64_t k0 = kc - u.K;
        if (k0 < 0) k0 = 0;
        int64_t k1 = kc + u.K + 1;
        if (k1 > nz) k1 = nz;
        const float inv = (float)(1.0 / u.sig);
        auto edge = [&](int e) -> double {
            return LIN ? ((double)e + 0.5 - g.crpix3) * g.cdelt3 + g.crval3 : ve[e];
        };
        const float cf0 = scf[0];
        // LIN: the table coordinate of edge k0 + e is x0 + e dx, one FP32 fma from the window's first edge (x0 from
        // the FP64 edge as above; the fma's error, ~1e-6 in t over a window of 64 channels, is that of the FP32
        // cast of the FP64 difference); the edge index is clamped to the window, so the channels past it take
        // two equal values and add 0
        const float x0 = k1 > k0 ? (float)(edge((int)k0) - vplane) * inv : 0.f;
        const float dx = (float)(g.cdelt3 / u.sig);
        const int ne = (int)(k1 - k0);
        float pp = k1 > k0 ? phi_as(x0) : 0.f;
        // blocks of BB_CB channels: the block's profile values first (independent table reads, one wait), then
        // every run adds the block (one address per run, the channels as immediate offsets)
        for (int kb = (int)k0; kb < (int)k1; kb += BB_CB) {
            const int nk = (int)k1 - kb;
            float pe[BB_CB], prof[BB_CB];
#pragma unroll
            for (int j = 0; j < BB_CB; j++) {            // (independent lookups: the loads issue together)
                if (LIN) {
                    const int e = min(kb - (int)k0 + j + 1, ne);
                    pe[j] = phi_as(fmaf((float)e, dx, x0));
                } else {
                    const int k = kb + j < nz ? kb + j : nz - 1;      // (past k1: computed, not added)
                    pe[j] = phi_as((float)(edge(k + 1) - vplane) * inv);
                }
            }
#pragma unroll
            for (int j = 0; j < BB_CB; j++) {
                if (LIN) {
                    prof[j] = fabsf(pe[j] - (j ? pe[j - 1] : pp)) * cf0;
                } else {
                    const int k = kb + j < nz ? kb + j : nz - 1;
                    const float v = fabsf(pe[j] - (j ? pe[j - 1] : pp)) * scf[k];
                    prof[j] = j < nk ? v : 0.f;
                }
            }
            pp = pe[BB_CB - 1];
#pragma unroll
            for (int r = 0; r < BB_RMAX; r++) {
                const bool on = r < nr;
                if (__ballot(on) == 0) break;      // (no lane has a run r: the wavefront leaves the loop)
                if (on) {
                    // (channels past the window have prof 0 and add 0: no branch per channel; the tile has BB_CB
                    // words of padding)
                    const float aw = a * rw[r] * A.inv_q;
                    unsigned long long* t = A.tile + rs[r] * A.nzp + kb;
#pragma unroll
                    for (int j = 0; j < BB_CB; j++)
                        atomicAdd(t + j, (unsigned long long)(unsigned int)fmaf(aw, prof[j], 0.5f));
                }
            }
        }
    }
}

// Deposit the runs of one azimuth sample that fall in the tile (galmod.py:389-431 for one sample).
template <bool LIN>
__device__ __forceinline__ void an_sample(const bb_grid& g, const bb_box& box, const TileGeo& T, const Acc& A,
                                          const double* sve, const double* scd, const float* scf,
                                          const Sub& u, const double* __restrict__ wt, double caz,
                                          double saz) {
    if (u.zdep) {
        an_sample_lag(g, box, T, A, sve, scd, u, wt, caz, saz);
        return;
    }
    const int nz = g.nz;
    const double wside = caz >= 0.0 ? u.wrec : u.wapp;
    const double x = u.R * caz, yb = u.R * saz * u.cinc, xs = x * u.spa, xc = x * u.cpa;
    const double vplane = u.vsys + (u.vrot * caz + u.vrad * saz) * u.sinc + u.vvert * u.cinc;
    const double afl = u.flux * wside;
    const Track tr = utrack(u, g, xs, xc, yb);
    const TrackF tf = utrack_f(u, tr, T.ox, T.oy);
    int rs[BB_RMAX];
    float rw[BB_RMAX];
    if (u.nzz <= BB_RMAX) {
        // a short column: the whole column is walked (the unrolled loop costs the same for any sub-range) and
        // each run that ends inside the tile is appended to rs[0..nr) (a select over the unrolled slots, no
        // register indexing), so the line loop runs over as many slots as the wavefront's longest list (runs
        // stored at the index of their last vertical sample instead left 3 of 4 lane slots idle: 4.7% slower)
        const int ka = 0, kb = u.nzz - 1;
        int c0 = 0, c1 = 0;
        double ws = 0.0;
        int nr = 0;
#pragma unroll
        for (int kz = 0; kz <= BB_RMAX; kz++) {
            const bool in = kz >= ka && kz <= kb;
            int h0 = 0, h1 = 0;
            if (in) upix_ff(u, g, tr, tf, xs, xc, yb, kz, h0, h1);
            const bool same = in && kz > ka && h0 == c0 && h1 == c1;
            if (kz > 0) {
                const bool close = kz - 1 >= ka && kz - 1 <= kb && !same;     // the run open at kz - 1 ends there
                const int sl = close ? T.slot(c0, c1) : -1;
                if (sl >= 0) {
#pragma unroll
                    for (int r = 0; r < BB_RMAX; r++)
                        if (r == nr) {
                            rs[r] = sl;
                            rw[r] = (float)ws;
                        }
                    nr++;
                }
            }
            if (in) {
                const double w = uw_(u, wt, kz, g.ltype);
                if (same) {
                    ws += w;
                } else {
                    c0 = h0;
                    c1 = h1;
                    ws = w;
                }
            }
        }
        if (nr > 0) an_lines<LIN>(g, A, sve, scf, u, vplane, afl, rs, rw, nr);
        return;
    }
    // a long column: the runs in the tile are compacted into rs[0..nr); beyond BB_RMAX they are deposited one by one
    int ka, kb;
    if (!tile_zrange(u, tr, T, ka, kb)) return;
    int nr = 0;
    int c0, c1;
    upix_ff(u, g, tr, tf, xs, xc, yb, ka, c0, c1);
    double ws = uw_(u, wt, ka, g.ltype);
    for (int kz = ka + 1; kz <= kb + 1; kz++) {
        int h0 = 0, h1 = 0;
        bool same = false;
        if (kz <= kb) {
            upix_ff(u, g, tr, tf, xs, xc, yb, kz, h0, h1);
            same = h0 == c0 && h1 == c1;
        }
        if (same) {
            ws += uw_(u, wt, kz, g.ltype);
            continue;
        }
        const int sl = T.slot(c0, c1);
        if (sl >= 0) {
            if (nr < BB_RMAX) {
#pragma unroll
                for (int r = 0; r < BB_RMAX; r++)
                    if (r == nr) {
                        rs[r] = sl;
                        rw[r] = (float)ws;
                    }
                nr++;
            } else {
                for (int nl = 0; nl < g.nlines; nl++) {
                    int64_t kc = bb_lround(bb_velgrid(vplane, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3,
                                                      g.crpix3));
                    acc_line(A.tile, A.inv_q, A.nzp, sl, g.relint[nl] * afl * ws, vplane, u.sig,
                             sve + (size_t)nl * (nz + 1), scd, kc, u.K, nz);
                }
            }
        }
        if (kz <= kb) {
            c0 = h0;
            c1 = h1;
            ws = uw_(u, wt, kz, g.ltype);
        }
    }
    if (nr > 0) an_lines<LIN>(g, A, sve, scf, u, vplane, afl, rs, rw, nr);
}

// The samples of a batch vector that meet the tile [tx0, tx1] x [ty0, ty1] (box-relative): the chunks whose
// box meets it are gathered in clist, then their samples whose vertical segment meets it in slist, and dep(e)
// is called for each, e = (chunk << 5) | azimuth lane, one per thread. The list holds neighbouring azimuths of
// one sub-ring next to each other; in a full window of blockDim entries (W wavefronts) lane l of wavefront w takes
// entry 64 S (w / S) + S l + w % S, S = BB_SSTRIDE (a last, partial window: l Wa + w), so the lanes of one LDS
// atomic are S azimuths apart (different pixels: same-address atomics of a wavefront serialise) on neighbouring
// sub-rings (similar work). The samples of successive batches
// of chunks accumulate in slist until it is nearly full or the tile's chunks are done (larger windows: fuller
// wavefronts). Every thread of the block calls it (it synchronises). SCAP: the entries of slist. WAVE: the
// wavefronts take 64 consecutive entries at a time instead, and every lane of a wavefront calls dep together
// (e = -1 past the end): the deposit that works across the lanes of a wavefront (an_wave).
template <int SCAP, bool WAVE = false, class Dep>
__device__ __forceinline__ void tile_samples(int nchunks, const short4* __restrict__ bb, const short4* __restrict__ sb,
                                             const int* __restrict__ tlist, int tlo, int thi, int tx0, int tx1,
                                             int ty0, int ty1, int* clist, int* slist, int& nlist, int& nsamp,
                                             bool active, Dep&& dep) {
    // the candidate chunks: entries [tlo, thi) of the tile's chunk list, or all the chunks (tlist NULL)
    const int lo = tlist ? tlo : 0, hi = tlist ? thi : nchunks;
    const int lane = threadIdx.x % BB_CS, grp = threadIdx.x / BB_CS;
    // 3. the ns samples gathered, one per thread (every thread calls it)
    auto flush = [&](int ns) {
        const int W = blockDim.x >> 6;
        if (WAVE) {
            for (int g0 = (int)(threadIdx.x >> 6) * 64; g0 < ns; g0 += blockDim.x) {
                const int idx = g0 + (int)(threadIdx.x & 63);
                dep(idx < ns ? slist[idx] : -1);
            }
        } else {
            for (int base2 = 0; base2 < ns; base2 += blockDim.x) {
                // (a last, partial window fills as few wavefronts as it can, its lanes still Wa entries apart)
                const int m = ns - base2;
                const int Wa = m >= (int)blockDim.x ? W : (m + 63) >> 6;
                const int wv = (int)(threadIdx.x >> 6), ln = (int)(threadIdx.x & 63);
                int idx;
                if (m >= (int)blockDim.x && W % BB_SSTRIDE == 0)      // (a full window: groups of BB_SSTRIDE wavefronts)
                    idx = base2 + (wv / BB_SSTRIDE) * 64 * BB_SSTRIDE + ln * BB_SSTRIDE + wv % BB_SSTRIDE;
                else
                    idx = wv < Wa ? base2 + ln * Wa + wv : ns;
                if (idx < ns) dep(slist[idx]);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) nsamp = 0;
        __syncthreads();
    };
    for (int base = lo; base < hi && active; base += blockDim.x) {
        // 1. the chunks whose bounding box meets the tile
        if (base + (int)threadIdx.x < hi) {
            const int c = tlist ? tlist[base + threadIdx.x] : base + threadIdx.x;
            const short4 r = bb[c];
            if (r.x <= r.y && r.x <= tx1 && r.y >= tx0 && r.z <= ty1 && r.w >= ty0) clist[atomicAdd(&nlist, 1)] = c;
        }
        __syncthreads();
        const int n = nlist;
        __syncthreads();                            // (every thread has read n before the list changes)
        if (!(n > BB_LIST_CAP - (int)blockDim.x || base + (int)blockDim.x >= hi)) continue;
        if (tlist) {
            // the tile's list is filled in no particular order: sort the chunks (bitonic, in LDS), so that the
            // samples come back in the order of their sub-rings and azimuths, as a scan of all the chunks gives them
            int np2 = 1;
            while (np2 < n) np2 <<= 1;
            for (int i = n + (int)threadIdx.x; i < np2; i += blockDim.x) clist[i] = 0x7fffffff;
            __syncthreads();
            for (int k = 2; k <= np2; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = threadIdx.x; i < np2; i += blockDim.x) {
                        const int ixj = i ^ j;
                        if (ixj > i) {
                            const int a = clist[i], bv = clist[ixj];
                            if ((a > bv) == ((i & k) == 0)) {
                                clist[i] = bv;
                                clist[ixj] = a;
                            }
                        }
                    }
                    __syncthreads();
                }
        }
        const int ngrp = (int)blockDim.x / BB_CS;
        for (int j0 = 0; j0 < n; j0 += ngrp) {
            // 2. their samples whose vertical segment meets the tile (one lane per azimuth)
            const int j = j0 + grp;
            if (j < n) {
                const int cj = clist[j];
                const short4 r = sb[(size_t)cj * BB_CS + lane];
                if (r.x <= r.y && r.x <= tx1 && r.y >= tx0 && r.z <= ty1 && r.w >= ty0)
                    slist[atomicAdd(&nsamp, 1)] = (cj << 5) | lane;
            }
            __syncthreads();
            const int ns = nsamp;
            __syncthreads();
            if (ns > SCAP - (int)blockDim.x) flush(ns);
        }
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
    }
    // the samples still gathered (every thread reads nsamp after the barrier that ended the loop)
    const int ns = nsamp;
    __syncthreads();
    if (ns > 0) flush(ns);
}

// ---------------------------------------------------------------------------------------------------------
// The deposit across the lanes of a wavefront (an_wave), for 64 consecutive samples of the tile's list.
//   Phase A, one lane per sample: the sample's line window (line 0) and its runs in the tile, walked BB_SEG
//   closing positions at a time (a segment of its column) into registers indexed by the position p: the run that
//   ends at vertical sample kbase + p sets bit p of m, its amplitude aw_p[p] and the LDS offset of its spectrum
//   sb_p[p]. No register array is indexed by a run count, so no select chains and no spills.
//   Phase B, the lanes along the channels: for each sample of the wavefront that has runs, lane l computes the
//   Phi table value at edge E0 + l of its window; a channel's lower edge comes from the lane below by a DPP shift
//   (lane 0's from the previous block of 64 edges), and every run of the sample adds the profile to its pixel's
//   spectrum, 64 consecutive channels per LDS atomic: no bank conflicts, and no idle lanes but past the end of
//   the window. The sample's parameters come to the wavefront by readlane.
// The values added, and so the model, are those of the lane-per-sample deposit (an_sample / an_lines) bit for
// bit: the same FP32 operations on the same operands, integer sums. Samples whose velocity depends on height take
// the lane-per-sample path (an_sample_lag) in Phase A.
#define BB_SEG 16
#define BB_WRUNS 256   // runs of a wavefront's samples held in its LDS buffer at a time

// v of the lane below (lane l receives lane l - 1's v; lane 0 receives lane0): DPP wave_shr:1
__device__ __forceinline__ float from_lane_below(float v, float lane0) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float readlane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// NL1: one spectral line (the window and amplitudes of Phase A are those of line 0); otherwise every line's window
// is computed in Phase B from the sample's velocity.
template <bool LIN, bool NL1>
__device__ __forceinline__ void an_wave(const bb_grid& g, const bb_box& box, const TileGeo& T, const Acc& A,
                                        const double* sve, const double* scd, const float* scf,
                                        const int4* __restrict__ chunks, const double2* __restrict__ az,
                                        const double* __restrict__ sub, const double* __restrict__ wtab, uint2* rbuf,
                                        int e) {
    const int lane = (int)(threadIdx.x & 63);
    const int nz = g.nz;
    // ---- Phase A: one lane per sample
    Sub u;
    bool ok = e >= 0;
    double xs = 0., xc = 0., yb = 0., vplane = 0., afl = 0.;
    const double* wt = nullptr;
    Track tr;
    tr.X0 = tr.dX = tr.Y0 = tr.dY = 0.;
    u.nzz = 0;
    u.K = 0;
    u.sig = 1.;
    if (ok) {
        const int4 ch = chunks[e >> 5];
        const int s = ch.x >> 1;
        const double* q = sub + (size_t)s * BB_NSF;
        load_sub(u, q, ch.x & 1);
        const double2 cs = az[ch.w + (e & 31)];
        wt = wtab + (size_t)s * BB_ZCAP;
        if (u.zdep) {
            an_sample_lag(g, box, T, A, sve, scd, u, wt, cs.x, cs.y);
            ok = false;
        } else {
            // (the operations of an_sample)
            const double wside = cs.x >= 0.0 ? u.wrec : u.wapp;
            const double x = u.R * cs.x;
            yb = u.R * cs.y * u.cinc;
            xs = x * u.spa;
            xc = x * u.cpa;
            vplane = u.vsys + (u.vrot * cs.x + u.vrad * cs.y) * u.sinc + u.vvert * u.cinc;
            afl = u.flux * wside;
            tr = utrack(u, g, xs, xc, yb);
        }
    }
    // the window of line 0 (an_lines): first channel k0, ne channels, the table coordinate x0 of its first edge
    int hk0 = 0, hne = 0;
    float hx0 = 0.f, ha = 0.f;
    const float hinv = (float)(1.0 / u.sig), hdx = (float)(g.cdelt3 / u.sig);
    if (NL1 && ok) {
        const int64_t kc = bb_lround(bb_velgrid(vplane, g.axtyp, g.velsys, g.lfreq0[0], g.crval3, g.cdelt3, g.crpix3));
        int64_t k0 = kc - u.K;
        if (k0 < 0) k0 = 0;
        int64_t k1 = kc + u.K + 1;
        if (k1 > nz) k1 = nz;
        if (k1 > k0) {
            hk0 = (int)k0;
            hne = (int)(k1 - k0);
        }
        ha = (float)(g.relint[0] * afl);
        if (LIN && hne > 0)
            hx0 = (float)((((double)hk0 + 0.5 - g.crpix3) * g.cdelt3 + g.crval3) - vplane) * hinv;
        if (hne == 0 || ha == 0.f) ok = false;        // (the lane-per-sample deposit adds zeros)
    }
    // the closing positions of the sample's runs: the whole column, or the part of a long one that can meet the tile
    int ka = 0, kb = -1;
    if (ok) {
        if (u.nzz <= BB_SEG) kb = u.nzz - 1;
        else if (!tile_zrange(u, tr, T, ka, kb)) ok = false;
    }
    int c0 = 0, c1 = 0;
    double ws = 0.;
    if (ok) {
        upix_fast(u, g, tr, xs, xc, yb, uz(u, ka), c0, c1);
        ws = uw_(u, wt, ka, g.ltype);
    }
    const float cf0 = scf[0];
    for (int kbase = ka; ; kbase += BB_SEG) {
        const bool more = ok && kbase <= kb;
        if (__ballot(more) == 0) break;        // (uniform: the wavefront's samples are all walked)
        unsigned m = 0;
        float aw_p[BB_SEG];
        int sb_p[BB_SEG];
#pragma unroll
        for (int j = 0; j < BB_SEG; j++) {
            const int kz = kbase + j + 1;      // the run open at kz - 1 closes unless kz is in its pixel
            if (more && kz - 1 <= kb) {
                int h0 = 0, h1 = 0;
                bool same = false;
                if (kz <= kb) {
                    upix_fast(u, g, tr, xs, xc, yb, uz(u, kz), h0, h1);
                    same = h0 == c0 && h1 == c1;
                }
                if (same) {
                    ws += uw_(u, wt, kz, g.ltype);
                } else {
                    const int sl = T.slot(c0, c1);
                    if (sl >= 0) {
                        m |= 1u << j;
                        aw_p[j] = NL1 ? ha * (float)ws * A.inv_q : (float)ws;
                        sb_p[j] = sl * A.nzp + (NL1 ? hk0 : 0);
                    }
                    if (kz <= kb) {
                        c0 = h0;
                        c1 = h1;
                        ws = uw_(u, wt, kz, g.ltype);
                    }
                }
            }
        }
        // ---- the runs of the segment, compacted in the wavefront's LDS buffer (rbuf, BB_WRUNS entries: {aw bits,
        // spectrum offset}) in lane order; more than BB_WRUNS runs go through in rounds of whole lanes
        const int nr = __popc(m);
        const unsigned long long below = (1ull << lane) - 1ull;
        int pre = 0;                             // the runs of the lanes below this one
#pragma unroll
        for (int bit = 0; bit < 5; bit++) pre += __popcll(__ballot((nr >> bit) & 1) & below) << bit;
        unsigned long long pending = __ballot(nr > 0);
        int r0 = 0;
        while (pending) {
            const bool inr = nr > 0 && pre >= r0 && pre - r0 + nr <= BB_WRUNS;
            if (inr) {
                int p = pre - r0;
#pragma unroll
                for (int j = 0; j < BB_SEG; j++)
                    if (m & (1u << j)) {
                        rbuf[p] = make_uint2(__float_as_uint(aw_p[j]), (unsigned)sb_p[j]);
                        p++;
                    }
            }
            unsigned long long lm = __ballot(inr);
            pending &= ~lm;
            const int rbn = (pre - r0) | (nr << 16);       // (the sample's first record and run count, packed)
            // ---- Phase B: the lanes along the channels, one sample of the round at a time
            while (lm) {
                const int L = __builtin_ctzll(lm);
                lm &= lm - 1;
                const int pk = __builtin_amdgcn_readlane(rbn, L), rb = pk & 0xffff, rn = pk >> 16;
                const float dx = readlane_f(hdx, L), inv = (LIN && NL1) ? 0.f : readlane_f(hinv, L);
                const double vpl = (LIN && NL1) ? 0. : readlane_d(vplane, L);
                const double aflL = NL1 ? 0. : readlane_d(afl, L);
                const int KL = NL1 ? 0 : (int)__builtin_amdgcn_readlane((int)u.K, L);
                for (int nl = 0; nl < (NL1 ? 1 : g.nlines); nl++) {
                    const double* ve = sve + (size_t)nl * (nz + 1);
                    int k0, ne;
                    float x0 = 0.f, a = 0.f;
                    if (NL1) {
                        const int kn = __builtin_amdgcn_readlane(hk0 | (hne << 16), L);
                        k0 = kn & 0xffff;
                        ne = kn >> 16;
                        if (LIN) x0 = readlane_f(hx0, L);
                    } else {
                        const int64_t kc = bb_lround(bb_velgrid(vpl, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3,
                                                                g.crpix3));
                        int64_t q0 = kc - KL;
                        if (q0 < 0) q0 = 0;
                        int64_t q1 = kc + KL + 1;
                        if (q1 > nz) q1 = nz;
                        if (q1 <= q0) continue;
                        k0 = (int)q0;
                        ne = (int)(q1 - q0);
                        a = (float)(g.relint[nl] * aflL);
                        if (LIN)
                            x0 = (float)((((double)k0 + 0.5 - g.crpix3) * g.cdelt3 + g.crval3) - vpl) * inv;
                    }
                    float carry = 0.f;
                    for (int E0 = 0; E0 <= ne; E0 += 64) {
                        // lane l: edge min(E0 + l, ne) of the window (exact in float32); the channel below it is
                        // c = E0 + l - 1. Past the window both edges are edge ne and the profile is 0, and lane 0 of
                        // the first block takes its own value as the edge below (channel -1 adds 0): no lane mask,
                        // the zeros go to distinct words of the padded tile.
                        const float Ef = (float)min(E0 + lane, ne);
                        const float q = LIN ? phi_as(fmaf(Ef, dx, x0)) : phi_as((float)(ve[k0 + (int)Ef] - vpl) * inv);
                        if (E0 == 0) carry = readlane_f(q, 0);
                        const float qlo = from_lane_below(q, carry);
                        carry = readlane_f(q, 63);
                        const int c = E0 + lane - 1;
                        const float prof = fabsf(q - qlo) * (LIN ? cf0 : scf[min(max(k0 + c, 0), nz - 1)]);
                        unsigned long long* t = A.tile + (NL1 ? 0 : k0) + c;
                        auto dep = [&](uint2 rec) {
                            const float aw = NL1 ? __uint_as_float(rec.x) : a * __uint_as_float(rec.x) * A.inv_q;
                            atomicAdd(t + (int)rec.y, (unsigned long long)(unsigned int)fmaf(aw, prof, 0.5f));
                        };
                        int r = 0;
                        for (; r + 1 < rn; r += 2) {      // (two records per LDS read)
                            const uint2 r1 = rbuf[rb + r], r2 = rbuf[rb + r + 1];
                            dep(r1);
                            dep(r2);
                        }
                        if (r < rn) dep(rbuf[rb + r]);
                    }
                }
            }
            if (pending) r0 = __builtin_amdgcn_readlane(pre, __builtin_ctzll(pending));
        }
    }
}

// MODE 0: one lane per sample (an_sample); 1: the wavefront deposit of one line (an_wave<LIN, true>); 2: the
// wavefront deposit of several lines.
template <bool LIN, int MODE>
__global__ void __launch_bounds__(BB_TILE_THREADS)
k_an_tile(bb_grid g, bb_box box, int TX, int TY, const int4* __restrict__ chunks, int nchunks,
          const double2* __restrict__ az, const double* __restrict__ suball, const double* __restrict__ wtaball,
          int nsub, const short4* __restrict__ bball, const short4* __restrict__ sball,
          const double* __restrict__ cd2i, const double* __restrict__ vedges, float* __restrict__ outall,
          int accumulate, const int* __restrict__ bin, int bin_cap,
          const int4* __restrict__ vbox) {
    // dynamic LDS: [vedges (nlines, nz+1) f64][cd2i (nz) f64][cd2i (nz) f32][pad, tile u64, pad]
    extern __shared__ double dyn[];
    // the wavefront deposit holds a smaller sample list and, in the bytes this frees, each wavefront's run buffer
    constexpr int SCAP = MODE == 0 ? BB_AN_SAMP_CAP : BB_AN_WSAMP_CAP;
    constexpr int NRB = MODE == 0 ? 1 : (BB_TILE_THREADS / 64) * BB_WRUNS;
    static_assert(SCAP * sizeof(int) + NRB * sizeof(uint2) <= (BB_AN_SAMP_CAP + 2) * sizeof(int), "LDS budget");
    __shared__ int clist[BB_LIST_CAP];
    __shared__ int slist[SCAP];
    __shared__ uint2 rbuf[NRB];
    __shared__ int nlist, nsamp;
    __shared__ double red[BB_TILE_THREADS / 64];
    int b, tid;
    xcd_tile(tid, b);
    const int nz = g.nz, nl_ = g.nlines, nzp = bb_tile_stride(nz);
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    const int ntx = (bx + TX - 1) / TX;
    const int tx0 = (tid % ntx) * TX, ty0 = (tid / ntx) * TY;
    const int tx1 = tx0 + TX - 1, ty1 = ty0 + TY - 1;
    double* sve = dyn;
    double* scd = dyn + (size_t)nl_ * (nz + 1);
    float* scf = (float*)(scd + nz);
    // (one word of padding before the tile and BB_TILE_PAD after it)
    unsigned long long* tile = (unsigned long long*)(scf + ((nz + 1) & ~1)) + 1;
    const int ntile = TX * TY * nzp;
    const double* sub = suball + (size_t)b * nsub * BB_NSF;
    const double* wtab = wtaball + (size_t)b * nsub * BB_ZCAP;
    for (int i = threadIdx.x; i < ntile; i += blockDim.x) tile[i] = 0ull;
    for (int i = threadIdx.x; i < nl_ * (nz + 1); i += blockDim.x) sve[i] = vedges[i];
    for (int i = threadIdx.x; i < nz; i += blockDim.x) {
        scd[i] = cd2i[i];
        scf[i] = (float)cd2i[i];
    }
    // the fixed-point scale of this vector: amax bounds any single contribution
    double am = 0.;
    for (int s = threadIdx.x; s < nsub; s += blockDim.x) {
        const double* q = sub + (size_t)s * BB_NSF;
        const double a = fabs(q[F_FLUX]) * fmax(fabs(q[F_WREC]), fabs(q[F_WAPP]));
        am = a > am ? a : am;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(am, off, 64);
        am = o > am ? o : am;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = am;
    if (threadIdx.x == 0) {
        nlist = 0;
        nsamp = 0;
    }
    __syncthreads();
    double amax = 0.;
    for (int w = 0; w < BB_TILE_THREADS / 64; w++) amax = red[w] > amax ? red[w] : amax;
    double rmax = 0., cmax = 0.;
    for (int nl = 0; nl < nl_; nl++) rmax = fabs(g.relint[nl]) > rmax ? fabs(g.relint[nl]) : rmax;
    for (int k = 0; k < nz; k++) cmax = fabs(scd[k]) > cmax ? fabs(scd[k]) : cmax;
    amax *= rmax * cmax;
    const double q = amax > 0. ? amax * 0x1p-31 : 1.0;
    Acc A;
    A.tile = tile;
    A.nzp = nzp;
    A.inv_q = (float)(1.0 / q);
    const short4* bb = bball + (size_t)b * nchunks;
    const short4* sb = sball + (size_t)b * nchunks * BB_CS;
    TileGeo T;
    T.ax0 = T.ox = box.blo0 + tx0;
    T.ax1 = box.blo0 + (tx1 < bx - 1 ? tx1 : bx - 1);
    T.ay0 = T.oy = box.blo1 + ty0;
    T.ay1 = box.blo1 + (ty1 < by - 1 ? ty1 : by - 1);
    T.TX = TX;
    if (vbox) {
        // only the pixels of the vector's own model box receive samples (the reference's model of that box); the
        // slots keep the tile's origin
        const int4 v = vbox[b];
        T.ax0 = max(T.ax0, v.x);
        T.ax1 = min(T.ax1, v.z - 1);
        T.ay0 = max(T.ay0, v.y);
        T.ay1 = min(T.ay1, v.w - 1);
    }
    const bool live = amax > 0. && T.ax0 <= T.ax1 && T.ay0 <= T.ay1;
    const int* tl;
    int tlo, thi;
    tile_list(bin, gridDim.y, gridDim.x, bin_cap, b, tid, tl, tlo, thi);
    tile_samples<SCAP, MODE != 0>(nchunks, bb, sb, tl, tlo, thi, tx0, tx1, ty0, ty1, clist, slist, nlist,
                                            nsamp, live, [&](int e) {
        if constexpr (MODE == 0) {
            const int4 ch = chunks[e >> 5];
            const int s = ch.x >> 1;
            Sub u;
            load_sub(u, sub + (size_t)s * BB_NSF, ch.x & 1);
            const double2 cs = az[ch.w + (e & 31)];
            an_sample<LIN>(g, box, T, A, sve, scd, scf, u, wtab + (size_t)s * BB_ZCAP, cs.x, cs.y);
        } else {
            an_wave<LIN, MODE == 1>(g, box, T, A, sve, scd, scf, chunks, az, sub, wtab,
                                    rbuf + (threadIdx.x >> 6) * BB_WRUNS, e);
        }
    });
    __syncthreads();
    float* out = outall + (size_t)b * by * bx * nz;
    for (int i = threadIdx.x; i < TX * TY * nz; i += blockDim.x) {
        const int p = i / nz, k = i - p * nz;
        const int rx = tx0 + p % TX, ry = ty0 + p / TX;
        if (rx < bx && ry < by) {
            float* o = out + ((size_t)ry * bx + rx) * nz + k;
            const float v = (float)((double)tile[(size_t)p * nzp + k] * q);
            *o = accumulate ? *o + v : v;
        }
    }
}

// LDS bytes of the dynamic part of k_an_tile: channel edges and factors (f64 and f32), the tile (u64) with its padding.
extern "C" size_t bb_analytic_lds_bytes(int nz, int nlines, int TX, int TY) {
    return ((size_t)nlines * (nz + 1) + nz) * sizeof(double) + (size_t)((nz + 1) & ~1) * sizeof(float) +
           ((size_t)TX * TY * bb_tile_stride(nz) + 1 + BB_TILE_PAD) * sizeof(unsigned long long);
}

extern "C" size_t bb_analytic_lds_static() {
    return (BB_LIST_CAP + BB_AN_SAMP_CAP) * sizeof(int) + (BB_TILE_THREADS / 64) * sizeof(double) + 64;
}

// Events recorded around every k_an_tile launch of this host thread (measurement; null: none).
static thread_local hipEvent_t t_tile_ev[2] = {nullptr, nullptr};

extern "C" void bb_set_tile_events(void* ev_begin, void* ev_end) {
    t_tile_ev[0] = (hipEvent_t)ev_begin;
    t_tile_ev[1] = (hipEvent_t)ev_end;
}

// ---------------------------------------------------------------------------------------------------------
extern "C" int bb_launch_galmod_analytic(const bb_grid* g, bb_box box, const double* sr, int nsub, int B,
                                         const int* naz, const int4* chunks, int nchunks, const double2* az,
                                         int TX, int TY, double* sub, double* wtab, short4* bbox, short4* sbox,
                                         const double* cd2i, const double* vedges, int lin_axis, float* out,
                                         int accumulate, int* bin, int bin_cap,
                                         const int4* vbox, int variant, int boxes_ready, hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    hipLaunchKernelGGL(k_an_prep, dim3((nsub + 63) / 64, B), dim3(64), 0, st, *g, box, sr, nsub, naz, sub, wtab, 0,
                       vbox);
    // (the sample boxes depend on the geometry only: boxes_ready keeps those of the ring maps of the same samples)
    if (!boxes_ready)
        hipLaunchKernelGGL(k_an_bbox, dim3((nchunks * BB_CS + 255) / 256, B), dim3(256), 0, st, *g, box, chunks,
                           nchunks, az, sub, nsub, bbox, sbox, vbox);
    const int ntiles = ((bx + TX - 1) / TX) * ((by + TY - 1) / TY);
    if (bin) {
        const int e = bin_chunks(box, B, bbox, nchunks, TX, TY, bin, bin_cap, st);
        if (e) return e;
    }
    const size_t lds = bb_analytic_lds_bytes(g->nz, g->nlines, TX, TY);
    // a linear velocity axis with one channel factor: the channel edges and factors computed, not read. variant 0:
    // one lane per sample, 1: the wavefront deposit (one line, or several)
    using Kern = void (*)(bb_grid, bb_box, int, int, const int4*, int, const double2*, const double*, const double*,
                          int, const short4*, const short4*, const double*, const double*, float*, int, const int*, int,
                          const int4*);
    static const Kern kern[2][3] = {{k_an_tile<false, 0>, k_an_tile<false, 1>, k_an_tile<false, 2>},
                                    {k_an_tile<true, 0>, k_an_tile<true, 1>, k_an_tile<true, 2>}};
    static bool attr = false;
    if (!attr) {
        const int dyn = (int)(160 * 1024 - bb_analytic_lds_static());
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 3; j++) {
                const hipError_t e = hipFuncSetAttribute((const void*)kern[i][j], hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
                if (e != hipSuccess) return (int)e;
            }
        attr = true;
    }
    if (t_tile_ev[0] && t_tile_ev[1]) {
        hipError_t e = hipEventRecord(t_tile_ev[0], st);
        if (e != hipSuccess) return (int)e;
    }
    const int mode = variant == 0 ? 0 : (g->nlines == 1 ? 1 : 2);
    hipLaunchKernelGGL(kern[lin_axis != 0][mode], dim3(ntiles, B), dim3(BB_TILE_THREADS), lds, st, *g, box, TX, TY,
                       chunks, nchunks, az, sub, wtab, nsub, bbox, sbox, cd2i, vedges, out, accumulate, bin, bin_cap,
                       vbox);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && t_tile_ev[0] && t_tile_ev[1]) e = hipEventRecord(t_tile_ev[1], st);
    return (int)e;
}

extern "C" size_t bb_analytic_sub_bytes() { return BB_NSF * sizeof(double); }
extern "C" size_t bb_analytic_wt_bytes() { return BB_ZCAP * sizeof(double); }
extern "C" int bb_analytic_chunk() { return BB_CS; }

// ---------------------------------------------------------------------------------------------------------
// Unit-density ring maps (galmod.py:468-536): the samples of the deposit above summed over the channels.
// Every run of equal-pixel vertical samples adds fluxsc x (its vertical weights) x (the side weight) x wval to
// the total-intensity map of each ring its sub-ring interpolates. The sub-ring records come from k_an_prep with
// unit densities (their flux field is fluxsc) and the sample boxes from k_an_bbox on the box of the maps. One
// workgroup per (vector, tile of TX x TY pixels) holds the tile's nw maps in LDS as 64-bit fixed point (a
// contribution x adds round(x / qunit): exact sums, independent of the order of the atomics) and writes them
// once: every count is written by exactly one workgroup.
__global__ void __launch_bounds__(BB_RM_THREADS)
k_rm_tile(bb_grid g, bb_box box, int TX, int TY, const int4* __restrict__ chunks, int nchunks,
          const double2* __restrict__ az, const double* __restrict__ suball, const double* __restrict__ wtaball,
          int nsub, const short4* __restrict__ bball, const short4* __restrict__ sball, const int2* __restrict__ widx,
          const double2* __restrict__ wval, const double* __restrict__ ewall, int nmap, int map0, int nw,
          double inv_q, unsigned long long* __restrict__ counts, const int* __restrict__ bin, int bin_cap, int wlds) {
    extern __shared__ double dyn[];
    __shared__ int clist[BB_LIST_CAP];
    __shared__ int slist[BB_RM_SAMP_CAP];
    __shared__ int nlist, nsamp;
    int b, tid;
    xcd_tile(tid, b);
    const int nwp = nw | 1;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    const int ntx = (bx + TX - 1) / TX;
    const int tx0 = (tid % ntx) * TX, ty0 = (tid / ntx) * TY;
    const int tx1 = tx0 + TX - 1, ty1 = ty0 + TY - 1;
    unsigned long long* tile = (unsigned long long*)dyn;
    // (wlds: the first BB_RM_WCAP vertical weights of every sub-ring of the vector, FP64, after the tile)
    double* swd = wlds ? (double*)(tile + (size_t)TX * TY * nwp) : nullptr;
    for (int i = threadIdx.x; i < TX * TY * nwp; i += blockDim.x) tile[i] = 0ull;
    if (threadIdx.x == 0) {
        nlist = 0;
        nsamp = 0;
    }
    __syncthreads();
    const double* sub = suball + (size_t)b * nsub * BB_NSF;
    const double* wtab = wtaball + (size_t)b * nsub * BB_ZCAP;
    const double* ew = ewall ? ewall + (size_t)b * nsub : nullptr;     // (the vector's edge weights of the density)
    if (wlds)
        for (int i = threadIdx.x; i < nsub * BB_RM_WCAP; i += blockDim.x)
            swd[i] = wtab[(size_t)(i / BB_RM_WCAP) * BB_ZCAP + i % BB_RM_WCAP];
    __syncthreads();
    TileGeo T;
    T.ax0 = T.ox = box.blo0 + tx0;
    T.ax1 = box.blo0 + (tx1 < bx - 1 ? tx1 : bx - 1);
    T.ay0 = T.oy = box.blo1 + ty0;
    T.ay1 = box.blo1 + (ty1 < by - 1 ? ty1 : by - 1);
    T.TX = TX;
    const int* tl;
    int tlo, thi;
    tile_list(bin, gridDim.y, gridDim.x, bin_cap, b, tid, tl, tlo, thi);
    tile_samples<BB_RM_SAMP_CAP>(nchunks, bball + (size_t)b * nchunks, sball + (size_t)b * nchunks * BB_CS, tl, tlo, thi, tx0, tx1, ty0,
                 ty1, clist, slist, nlist, nsamp, true, [&](int e) {
        const int4 ch = chunks[e >> 5];
        const int s = ch.x >> 1;
        int2 wi = widx[s];
        if (wi.x >= nw) wi.x = -1;      // (out-of-range ring indices are ignored)
        if (wi.y >= nw) wi.y = -1;
        if (wi.x < 0 && wi.y < 0) return;
        Sub u;
        load_sub(u, sub + (size_t)s * BB_NSF, ch.x & 1);
        const double* wt = wtab + (size_t)s * BB_ZCAP;
        const double* wd = swd ? swd + (size_t)s * BB_RM_WCAP : nullptr;
        auto wgt = [&](int kz) -> double {        // the weight of vertical sample kz, FP64
            if (u.nzz <= 1) return 1.0;
            const int j = zbin_of(kz, u.half);
            return (wd && j < BB_RM_WCAP) ? wd[j] : uw_(u, wt, kz, g.ltype);
        };
        const double2 cs = az[ch.w + (e & 31)];
        const double wside = cs.x >= 0.0 ? u.wrec : u.wapp;
        if (wside == 0.0) return;
        double2 wv = wval[s];
        if (ew) {
            wv.x *= ew[s];
            wv.y *= ew[s];
        }
        const double x = u.R * cs.x, yb = u.R * cs.y * u.cinc, xs = x * u.spa, xc = x * u.cpa;
        const Track tr = utrack(u, g, xs, xc, yb);
        const TrackF tf = utrack_f(u, tr, T.ox, T.oy);
        int ka, kb;
        ka = 0;                 // (the whole column: cheaper here than narrowing it to the tile)
        kb = u.nzz - 1;
        int c0, c1;
        upix_ff(u, g, tr, tf, xs, xc, yb, ka, c0, c1);
        double ws = wgt(ka);
        for (int kz = ka + 1; kz <= kb + 1; kz++) {
            int h0 = 0, h1 = 0;
            bool same = false;
            if (kz <= kb) {
                upix_ff(u, g, tr, tf, xs, xc, yb, kz, h0, h1);
                same = h0 == c0 && h1 == c1;
            }
            if (same) {
                ws += wgt(kz);
                continue;
            }
            const int sl = T.slot(c0, c1);
            if (sl >= 0) {
                const double base = u.flux * ws;
                if (wi.x >= 0) atomicAdd(tile + sl * nwp + wi.x, (unsigned long long)(base * wv.x * wside * inv_q + 0.5));
                if (wi.y >= 0) atomicAdd(tile + sl * nwp + wi.y, (unsigned long long)(base * wv.y * wside * inv_q + 0.5));
            }
            if (kz <= kb) {
                c0 = h0;
                c1 = h1;
                ws = wgt(kz);
            }
        }
    });
    __syncthreads();
    for (int i = threadIdx.x; i < nw * TX * TY; i += blockDim.x) {
        const int m = i / (TX * TY), p = i - m * (TX * TY);
        const int rx = tx0 + p % TX, ry = ty0 + p / TX;
        if (rx < bx && ry < by) counts[(((size_t)b * nmap + map0 + m) * by + ry) * bx + rx] = tile[(size_t)p * nwp + m];
    }
}

// The map of the exterior annulus when its radii follow each vector's edge (a table per vector, no shared azimuth
// plan): one workgroup per (sub-ring, radial sample, vector), a thread per azimuth (cos and sin of the azimuth itself
// rather than the plan's rotation recurrence), the runs of equal-pixel vertical samples of the column added to the
// vector's counts with global 64-bit atomics, as k_rm_tile adds them (the annulus is a few sub-rings beside hundreds).
// Sub-rings [0, nlay) are the thin layer (ring weight wlay.x), the rest the thick one (wlay.y).
__global__ void __launch_bounds__(256)
k_rm_annulus(bb_grid g, bb_box box, const double* __restrict__ suball, const double* __restrict__ wtaball, int nsub,
             int nlay, double2 wlay, double inv_q, int nmap, int map0, unsigned long long* __restrict__ counts) {
    const int s = blockIdx.x >> 1, kr = blockIdx.x & 1, b = blockIdx.y;
    const double* q = suball + ((size_t)b * nsub + s) * BB_NSF;
    Sub u;
    load_sub(u, q, kr);
    const double wv = s < nlay ? wlay.x : wlay.y;
    if ((u.wrec == 0.0 && u.wapp == 0.0) || wv == 0.0) return;
    const int64_t naz = (int64_t)q[F_NAZ];
    const double* wt = wtaball + ((size_t)b * nsub + s) * BB_ZCAP;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    unsigned long long* plane = counts + ((size_t)b * nmap + map0) * (size_t)by * bx;
    const double daz = 2 * BB_PI / (double)naz;
    const double az0 = 2 * BB_PI * (0.5 * ((kr + 1) % 2) + 0.25) / (double)naz;
    for (int64_t ka = threadIdx.x; ka < naz; ka += blockDim.x) {
        double saz, caz;
        sincos(az0 + (double)ka * daz, &saz, &caz);
        const double wside = caz >= 0.0 ? u.wrec : u.wapp;
        if (wside == 0.0) continue;
        const double x = u.R * caz, yb = u.R * saz * u.cinc, xs = x * u.spa, xc = x * u.cpa;
        const Track tr = utrack(u, g, xs, xc, yb);
        int c0, c1;
        upix_fast(u, g, tr, xs, xc, yb, uz(u, 0), c0, c1);
        double ws = uw_(u, wt, 0, g.ltype);
        for (int kz = 1; kz <= u.nzz; kz++) {
            int h0 = 0, h1 = 0;
            bool same = false;
            if (kz < u.nzz) {
                upix_fast(u, g, tr, xs, xc, yb, uz(u, kz), h0, h1);
                same = h0 == c0 && h1 == c1;
            }
            if (same) {
                ws += uw_(u, wt, kz, g.ltype);
                continue;
            }
            if (c0 >= box.blo0 && c0 < box.bhi0 && c1 >= box.blo1 && c1 < box.bhi1) {
                const double base = u.flux * ws;
                atomicAdd(plane + (size_t)(c1 - box.blo1) * bx + (c0 - box.blo0),
                          (unsigned long long)(base * wv * wside * inv_q + 0.5));
            }
            if (kz < u.nzz) {
                c0 = h0;
                c1 = h1;
                ws = uw_(u, wt, kz, g.ltype);
            }
        }
    }
}

// Zero planes [map0, map0 + nw) of every vector's counts (B, nmap, plane): a kernel rather than a memset, so that the
// launch can be captured in a HIP graph.
__global__ void k_zero_planes(unsigned long long* __restrict__ counts, int nmap, int map0, int nw, int64_t plane) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < plane * nw) counts[((int64_t)blockIdx.y * nmap + map0) * plane + i] = 0ull;
}

extern "C" int bb_launch_zero_planes(unsigned long long* counts, int B, int nmap, int map0, int nw, int64_t plane,
                                     hipStream_t st) {
    const int64_t n = plane * nw;
    if (n == 0 || B == 0) return 0;
    hipLaunchKernelGGL(k_zero_planes, dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, st, counts, nmap, map0, nw,
                       plane);
    return (int)hipGetLastError();
}

extern "C" size_t bb_rm_tile_lds(int nw, int TX, int TY) {
    return (size_t)TX * TY * (nw | 1) * sizeof(unsigned long long);
}

extern "C" size_t bb_rm_tile_static() { return (BB_LIST_CAP + BB_RM_SAMP_CAP) * sizeof(int) + 64; }

extern "C" int bb_launch_ring_maps(const bb_grid* g, bb_box box, const double* sr, int nsub, int B, const int* naz,
                                   const int4* chunks, int nchunks, const double2* az, int TX, int TY, double* sub,
                                   double* wtab, short4* bbox, short4* sbox, const int* widx, const double* wval,
                                   const double* ew, int nmap, int map0, int nw, double qunit,
                                   unsigned long long* counts, int* bin, int bin_cap, hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    hipLaunchKernelGGL(k_an_prep, dim3((nsub + 63) / 64, B), dim3(64), 0, st, *g, box, sr, nsub, naz, sub, wtab, 1,
                       (const int4*)nullptr);
    hipLaunchKernelGGL(k_an_bbox, dim3((nchunks * BB_CS + 255) / 256, B), dim3(256), 0, st, *g, box, chunks, nchunks,
                       az, sub, nsub, bbox, sbox, (const int4*)nullptr);
    if (bin) {
        const int e = bin_chunks(box, B, bbox, nchunks, TX, TY, bin, bin_cap, st);
        if (e) return e;
    }
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_rm_tile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(160 * 1024 - bb_rm_tile_static()));
        if (e != hipSuccess) return (int)e;
        attr = true;
    }
    const int ntiles = ((bx + TX - 1) / TX) * ((by + TY - 1) / TY);
    // the sub-rings' vertical weights in LDS beside the tile when they fit
    const size_t lt = bb_rm_tile_lds(nw, TX, TY), lw = (size_t)nsub * BB_RM_WCAP * sizeof(double);
    const int wl = lt + lw + bb_rm_tile_static() <= 160 * 1024 ? 1 : 0;
    hipLaunchKernelGGL(k_rm_tile, dim3(ntiles, B), dim3(BB_RM_THREADS), lt + (wl ? lw : 0), st, *g, box, TX, TY, chunks,
                       nchunks, az, sub, wtab, nsub, bbox, sbox, (const int2*)widx, (const double2*)wval, ew, nmap,
                       map0, nw, 1.0 / qunit, counts, bin, bin_cap, wl);
    return (int)hipGetLastError();
}

// The exterior annulus of every vector from its own table sr (B, 16, nsub): the sub-ring records (azimuths from each
// radius), the plane map0 of counts cleared, then k_rm_annulus. sub, wtab: B nsub records of work.
extern "C" int bb_launch_rm_annulus(const bb_grid* g, bb_box box, const double* sr, int nsub, int nlay, int B,
                                    double w_thin, double w_thick, double* sub, double* wtab, int nmap, int map0,
                                    double qunit, unsigned long long* counts, hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    int e = bb_launch_zero_planes(counts, B, nmap, map0, 1, (int64_t)by * bx, st);
    if (e || nsub == 0 || B == 0) return e;
    hipLaunchKernelGGL(k_an_prep, dim3((nsub + 63) / 64, B), dim3(64), 0, st, *g, box, sr, nsub, (const int*)nullptr,
                       sub, wtab, 1, (const int4*)nullptr);
    hipLaunchKernelGGL(k_rm_annulus, dim3(2 * nsub, B), dim3(256), 0, st, *g, box, sub, wtab, nsub, nlay,
                       make_double2(w_thin, w_thick), 1.0 / qunit, nmap, map0, counts);
    return (int)hipGetLastError();
}
