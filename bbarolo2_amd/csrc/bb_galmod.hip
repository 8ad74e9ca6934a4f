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
#define BB_RMAX 8     // vertical runs per sample held in registers (beyond: deposited one by one)
#define BB_TILE_THREADS 256

enum {
    F_R0 = 0, F_R1, F_XPOS, F_YPOS, F_SPA, F_CPA, F_SINC, F_CINC, F_VSYS, F_VROT, F_VRAD, F_VVERT, F_SIG, F_K,
    F_FLUX, F_WREC, F_WAPP, F_Z0, F_DVDZ, F_ZCYL, F_ZDEP, F_NZZ, F_NIN, F_STEP, F_ZLIM, F_DENS, F_HALF
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
          double* __restrict__ suball, double* __restrict__ wtaball) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (s >= nsub) return;
    const double* sr = srall + (size_t)b * 16 * nsub;
    double* q = suball + ((size_t)b * nsub + s) * BB_NSF;
    double* wt = wtaball + ((size_t)b * nsub + s) * BB_ZCAP;
    auto S = [&](int row) { return sr[(size_t)row * nsub + s]; };
    const double R0 = S(0), dens = S(5);
    const int64_t naz = naz_tab[s];
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

// Weight of vertical sample kz (the flux of the layer in its bin).
__device__ __forceinline__ double wq_at(const double* __restrict__ q, const double* __restrict__ wt, int kz,
                                        int ltype) {
    const int nzz = (int)q[F_NZZ];
    if (nzz == 1) return 1.0;
    const int half = (int)q[F_HALF];
    const int j = zbin_of(kz, half);
    if (j < BB_ZCAP) return wt[j];
    double lo, hi;
    zbin(j, (int)q[F_NIN], q[F_STEP], q[F_ZLIM], lo, hi);
    return bb_layer_cdf(ltype, hi) - bb_layer_cdf(ltype, lo);
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
          const double* __restrict__ suball, int nsub, short4* __restrict__ bball) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (c >= nchunks) return;
    const int4 ch = chunks[c];
    const int s = ch.x >> 1, kr = ch.x & 1;
    const double* q = suball + ((size_t)b * nsub + s) * BB_NSF;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    int64_t x0 = INT64_MAX, x1 = INT64_MIN, y0 = INT64_MAX, y1 = INT64_MIN;
    if (q[F_DENS] != 0) {
        const double R = kr ? q[F_R1] : q[F_R0];
        const int nzz = (int)q[F_NZZ];
        const double za = z_at(q, 0), zb = z_at(q, nzz - 1);
        for (int i = 0; i < ch.z; i++) {
            const double2 cs = az[ch.w + i];
            const double caz = cs.x, saz = cs.y;
            const double wside = caz >= 0.0 ? q[F_WREC] : q[F_WAPP];
            if (wside == 0.0) continue;
            const double x = R * caz, yb = R * saz * q[F_CINC], xs = x * q[F_SPA], xc = x * q[F_CPA];
            int64_t g0, g1;
            pix_at(q, g, xs, xc, yb, za, g0, g1);
            x0 = g0 < x0 ? g0 : x0; x1 = g0 > x1 ? g0 : x1; y0 = g1 < y0 ? g1 : y0; y1 = g1 > y1 ? g1 : y1;
            pix_at(q, g, xs, xc, yb, zb, g0, g1);
            x0 = g0 < x0 ? g0 : x0; x1 = g0 > x1 ? g0 : x1; y0 = g1 < y0 ? g1 : y0; y1 = g1 > y1 ? g1 : y1;
        }
    }
    short4 r = make_short4(1, 0, 1, 0);       // empty
    if (x0 <= x1) {
        // box-relative, widened by one pixel, clipped to the box
        int64_t a0 = x0 - box.blo0 - 1, a1 = x1 - box.blo0 + 1, c0 = y0 - box.blo1 - 1, c1 = y1 - box.blo1 + 1;
        a0 = a0 < 0 ? 0 : a0;
        c0 = c0 < 0 ? 0 : c0;
        a1 = a1 > bx - 1 ? bx - 1 : a1;
        c1 = c1 > by - 1 ? by - 1 : c1;
        if (a0 <= a1 && c0 <= c1) r = make_short4((short)a0, (short)a1, (short)c0, (short)c1);
    }
    bball[(size_t)b * nchunks + c] = r;
}

// ---------------------------------------------------------------------------------------------------------
// Add aw * the channel-integrated line (centre v) to the spectrum sp of the LDS tile (galmod.py:194-214).
__device__ __forceinline__ void lds_line(float* sp, double aw, double v, double sig, const double* __restrict__ ve,
                                         const double* __restrict__ cd2i, const double* __restrict__ ptab, int ntab,
                                         int64_t kc, int64_t K, int nz) {
    int64_t k0 = kc - K;
    if (k0 < 0) k0 = 0;
    int64_t k1 = kc + K + 1;
    if (k1 > nz) k1 = nz;
    if (k1 <= k0) return;
    const double inv = 1.0 / sig;
    double pp = bb_phi((ve[k0] - v) * inv, ptab, ntab);
    for (int64_t k = k0; k < k1; k++) {
        double pn = bb_phi((ve[k + 1] - v) * inv, ptab, ntab);
        double f = pn - pp;
        pp = pn;
        atomicAdd(sp + k, (float)(aw * ((f >= 0 ? f : -f) * cd2i[k])));
    }
}

struct TileCtx {
    int tx0, ty0, TX, TY, bx, by, nz;
    __device__ __forceinline__ int slot(int64_t g0, int64_t g1, const bb_box& box) const {
        int64_t rx = g0 - box.blo0 - tx0, ry = g1 - box.blo1 - ty0;
        if (rx < 0 || rx >= TX || ry < 0 || ry >= TY) return -1;
        if (rx + tx0 >= bx || ry + ty0 >= by) return -1;
        return (int)(ry * TX + rx);
    }
};

// One azimuth sample of a chunk: its vertical runs inside the tile, deposited into the LDS tile.
__device__ void tile_sample(const bb_grid& g, const bb_box& box, const TileCtx& t, float* tile, const double* q,
                            const double* wt, double R, double caz, double saz, const double* __restrict__ cd2i,
                            const double* __restrict__ vedges, const double* __restrict__ ptab, int ntab) {
    const double wside = caz >= 0.0 ? q[F_WREC] : q[F_WAPP];
    if (wside == 0.0) return;
    const int nz = t.nz, nzz = (int)q[F_NZZ];
    const double x = R * caz, yb = R * saz * q[F_CINC], xs = x * q[F_SPA], xc = x * q[F_CPA];
    const double sinc = q[F_SINC], cinc = q[F_CINC];
    const double vplane = q[F_VSYS] + (q[F_VROT] * caz + q[F_VRAD] * saz) * sinc + q[F_VVERT] * cinc;
    const double sig = q[F_SIG];
    const int64_t K = (int64_t)q[F_K];
    const double afl = q[F_FLUX] * wside;
    if (q[F_ZDEP] == 0.0) {
        // runs of consecutive vertical samples in the same pixel (galmod.py:414-431)
        int rslot[BB_RMAX];
        double rws[BB_RMAX];
        int nr = 0;
        int64_t c0, c1;
        pix_at(q, g, xs, xc, yb, z_at(q, 0), c0, c1);
        double ws = wq_at(q, wt, 0, g.ltype);
        for (int kz = 1; kz <= nzz; kz++) {
            int64_t h0 = 0, h1 = 0;
            bool same = false;
            if (kz < nzz) {
                pix_at(q, g, xs, xc, yb, z_at(q, kz), h0, h1);
                same = h0 == c0 && h1 == c1;
            }
            if (same) {
                ws += wq_at(q, wt, kz, g.ltype);
                continue;
            }
            // the run [.., kz) ends: keep it if its pixel is in this tile
            int sl = t.slot(c0, c1, box);
            if (sl >= 0) {
                if (nr < BB_RMAX) {
#pragma unroll
                    for (int r = 0; r < BB_RMAX; r++)
                        if (r == nr) { rslot[r] = sl; rws[r] = ws; }
                    nr++;
                } else {
                    for (int nl = 0; nl < g.nlines; nl++) {
                        double v = vplane;
                        int64_t kc = bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3,
                                                          g.crpix3));
                        lds_line(tile + (size_t)sl * nz, g.relint[nl] * afl * ws, v, sig,
                                 vedges + (size_t)nl * (nz + 1), cd2i, ptab, ntab, kc, K, nz);
                    }
                }
            }
            if (kz < nzz) {
                c0 = h0;
                c1 = h1;
                ws = wq_at(q, wt, kz, g.ltype);
            }
        }
        if (nr == 0) return;
        // one line profile per emission line, shared by the runs kept (galmod.py:411-413)
        for (int nl = 0; nl < g.nlines; nl++) {
            const double* ve = vedges + (size_t)nl * (nz + 1);
            const double a = g.relint[nl] * afl;
            const double v = vplane;
            const int64_t kc = bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
            int64_t k0 = kc - K;
            if (k0 < 0) k0 = 0;
            int64_t k1 = kc + K + 1;
            if (k1 > nz) k1 = nz;
            if (k1 <= k0) continue;
            const double inv = 1.0 / sig;
            double pp = bb_phi((ve[k0] - v) * inv, ptab, ntab);
            for (int64_t k = k0; k < k1; k++) {
                double pn = bb_phi((ve[k + 1] - v) * inv, ptab, ntab);
                double f = pn - pp;
                pp = pn;
                const double prof = (f >= 0 ? f : -f) * cd2i[k];
#pragma unroll
                for (int r = 0; r < BB_RMAX; r++)
                    if (r < nr) atomicAdd(tile + (size_t)rslot[r] * nz + k, (float)((a * rws[r]) * prof));
            }
        }
        return;
    }
    // the velocity depends on |z|: samples at +z and -z share their line (galmod.py:432-465)
    for (int kz = 0; kz < (nzz + 1) / 2; kz++) {
        const int kz2 = nzz - 1 - kz;
        const double za = z_at(q, kz), zb = z_at(q, kz2);
        int64_t ga0, ga1, gb0, gb1;
        pix_at(q, g, xs, xc, yb, za, ga0, ga1);
        pix_at(q, g, xs, xc, yb, zb, gb0, gb1);
        const bool in1 = !(ga0 < box.blo0 || ga0 >= box.bhi0 || ga1 < box.blo1 || ga1 >= box.bhi1);
        const bool in2 = kz2 != kz && !(gb0 < box.blo0 || gb0 >= box.bhi0 || gb1 < box.blo1 || gb1 >= box.bhi1);
        const int s1 = in1 ? t.slot(ga0, ga1, box) : -1;
        const int s2 = in2 ? t.slot(gb0, gb1, box) : -1;
        if (s1 < 0 && s2 < 0) continue;
        const double az_ = fabs(za);
        double v;
        if (az_ > q[F_ZCYL])
            v = q[F_VSYS] + (bb_lagged(q[F_VROT], q[F_DVDZ], az_ - q[F_ZCYL]) * caz + q[F_VRAD] * saz) * sinc +
                q[F_VVERT] * cinc;
        else
            v = vplane;
        const double w1 = wq_at(q, wt, kz, g.ltype), w2 = wq_at(q, wt, kz2, g.ltype);
        for (int nl = 0; nl < g.nlines; nl++) {
            const double* ve = vedges + (size_t)nl * (nz + 1);
            const double a = g.relint[nl] * afl;
            const int64_t kc = bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
            if (in1 && in2 && ga0 == gb0 && ga1 == gb1) {
                lds_line(tile + (size_t)s1 * nz, a * (w1 + w2), v, sig, ve, cd2i, ptab, ntab, kc, K, nz);
                continue;
            }
            if (s1 >= 0) lds_line(tile + (size_t)s1 * nz, a * w1, v, sig, ve, cd2i, ptab, ntab, kc, K, nz);
            if (s2 >= 0) lds_line(tile + (size_t)s2 * nz, a * w2, v, sig, ve, cd2i, ptab, ntab, kc, K, nz);
        }
    }
}

__global__ void __launch_bounds__(BB_TILE_THREADS)
k_an_tile(bb_grid g, bb_box box, int TX, int TY, const int4* __restrict__ chunks, int nchunks,
          const double2* __restrict__ az, const double* __restrict__ suball, const double* __restrict__ wtaball,
          int nsub, const short4* __restrict__ bball, const double* __restrict__ cd2i,
          const double* __restrict__ vedges, const double* __restrict__ ptab, int ntab, float* __restrict__ outall,
          int accumulate) {
    extern __shared__ float tile[];
    __shared__ int list[BB_TILE_THREADS];
    __shared__ int nlist;
    const int b = blockIdx.y;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    const int ntx = (bx + TX - 1) / TX;
    TileCtx t;
    t.TX = TX;
    t.TY = TY;
    t.tx0 = (blockIdx.x % ntx) * TX;
    t.ty0 = (blockIdx.x / ntx) * TY;
    t.bx = bx;
    t.by = by;
    t.nz = g.nz;
    const int nz = g.nz;
    const int ntile = TX * TY * nz;
    for (int i = threadIdx.x; i < ntile; i += blockDim.x) tile[i] = 0.f;
    const short4* bb = bball + (size_t)b * nchunks;
    const double* sub = suball + (size_t)b * nsub * BB_NSF;
    const double* wtab = wtaball + (size_t)b * nsub * BB_ZCAP;
    const int tx1 = t.tx0 + TX - 1, ty1 = t.ty0 + TY - 1;
    const int lane = threadIdx.x % BB_CS, grp = threadIdx.x / BB_CS, ngrp = blockDim.x / BB_CS;
    for (int base = 0; base < nchunks; base += blockDim.x) {
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
        const int c = base + threadIdx.x;
        if (c < nchunks) {
            const short4 r = bb[c];
            if (r.x <= r.y && r.x <= tx1 && r.y >= t.tx0 && r.z <= ty1 && r.w >= t.ty0) list[atomicAdd(&nlist, 1)] = c;
        }
        __syncthreads();
        const int n = nlist;
        for (int j = grp; j < n; j += ngrp) {
            const int4 ch = chunks[list[j]];
            if (lane < ch.z) {
                const int s = ch.x >> 1, kr = ch.x & 1;
                const double* q = sub + (size_t)s * BB_NSF;
                const double2 cs = az[ch.w + lane];
                tile_sample(g, box, t, tile, q, wtab + (size_t)s * BB_ZCAP, kr ? q[F_R1] : q[F_R0], cs.x, cs.y,
                            cd2i, vedges, ptab, ntab);
            }
        }
        __syncthreads();
    }
    float* out = outall + (size_t)b * by * bx * nz;
    for (int i = threadIdx.x; i < ntile; i += blockDim.x) {
        const int p = i / nz, k = i - p * nz;
        const int rx = t.tx0 + p % TX, ry = t.ty0 + p / TX;
        if (rx < bx && ry < by) {
            float* o = out + ((size_t)ry * bx + rx) * nz + k;
            *o = accumulate ? *o + tile[i] : tile[i];
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
extern "C" int bb_launch_galmod_analytic(const bb_grid* g, bb_box box, const double* sr, int nsub, int B,
                                         const int* naz, const int4* chunks, int nchunks, const double2* az,
                                         int TX, int TY, double* sub, double* wtab, short4* bbox, const double* cd2i,
                                         const double* vedges, const double* ptab, int ntab, float* out,
                                         int accumulate, hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    hipLaunchKernelGGL(k_an_prep, dim3((nsub + 63) / 64, B), dim3(64), 0, st, *g, box, sr, nsub, naz, sub, wtab);
    hipLaunchKernelGGL(k_an_bbox, dim3((nchunks + 255) / 256, B), dim3(256), 0, st, *g, box, chunks, nchunks, az, sub,
                       nsub, bbox);
    const int ntiles = ((bx + TX - 1) / TX) * ((by + TY - 1) / TY);
    const size_t lds = (size_t)TX * TY * g->nz * sizeof(float);
    hipLaunchKernelGGL(k_an_tile, dim3(ntiles, B), dim3(BB_TILE_THREADS), lds, st, *g, box, TX, TY, chunks, nchunks,
                       az, sub, wtab, nsub, bbox, cd2i, vedges, ptab, ntab, out, accumulate);
    return (int)hipGetLastError();
}

extern "C" size_t bb_analytic_sub_bytes() { return BB_NSF * sizeof(double); }
extern "C" size_t bb_analytic_wt_bytes() { return BB_ZCAP * sizeof(double); }
extern "C" int bb_analytic_chunk() { return BB_CS; }
