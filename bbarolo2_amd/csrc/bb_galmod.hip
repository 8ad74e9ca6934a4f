// GALMOD ANALYTIC on the GPU: clouds on a regular grid in radius / azimuth / height with
// channel-integrated Gaussian line profiles, deposited into a float32 yxz model cube.
// Follows bbarolo/galmod.py:275-465 (_zgrid, _galmod_analytic) sample by sample: the same samples,
// pixel rounding, merging of equal-pixel vertical runs and line profiles. Geometry is FP64 (pixel and
// channel indices round as in the reference); the deposited values are summed in FP32.
//
// Work decomposition: grid (CH, 2 * nsub, B). Block (cx, 2 s + kr, b) handles sub-ring s, radial sample
// kr of parameter vector b, azimuths ka = cx * 256 + t, stepping by CH * 256 (any CH >= 1 covers all
// naz azimuths; CH = ceil(naz_max / 256) gives one azimuth per thread). The vertical samples of the
// sub-ring are built once per block in LDS.
#include "bb_device.h"
#include "../../include/bbgpu.h"

#define BB_ZQ_MAX 1024
#define BB_AN_THREADS 256

struct SubRing {
    double R0, dens, vrot, vrad, vvert, sig, z0, dvdz, zcyl, sinc, cinc, spa, cpa, xpos, ypos, vsys, wrec,
        wapp, fluxsc, r_inner;
    int64_t naz, K;
};

__device__ __forceinline__ void load_subring(SubRing& q, const double* __restrict__ sr, int nsub, int ir,
                                             const bb_grid& g) {
    auto S = [&](int row) { return sr[(size_t)row * nsub + ir]; };
    q.R0 = S(0);
    q.dens = S(5);
    double ring_area_pix = 2 * BB_PI * q.R0 * g.radsep / g.pixarea;
    double circ_pix = 2 * BB_PI * q.R0 / sqrt(g.pixarea);
    int64_t naz = (int64_t)(3 * circ_pix + 0.5);
    q.naz = naz < 16 ? 16 : naz;
    q.vrot = S(1);
    q.vrad = S(2);
    q.vvert = S(3);
    double sig = sqrt(S(4) * S(4) + g.sig_instr * g.sig_instr);
    double smin = 1e-6 * fabs(g.chmin);
    q.sig = sig < smin ? smin : sig;
    q.K = (int64_t)(5.0 * q.sig / g.chmin) + 2;
    q.z0 = S(6);
    q.dvdz = S(7);
    q.zcyl = S(8);
    q.sinc = sin(S(9));
    q.cinc = cos(S(9));
    double phi = S(10);
    q.spa = sin(phi) * cos(g.crota2) + cos(phi) * sin(g.crota2);
    q.cpa = cos(phi) * cos(g.crota2) - sin(phi) * sin(g.crota2);
    q.xpos = S(11);
    q.ypos = S(12);
    q.vsys = S(13);
    q.wrec = S(14);
    q.wapp = S(15);
    q.fluxsc = q.dens * ring_area_pix * g.pixarea / (double)(q.naz * 2);
    q.r_inner = q.R0 - 0.5 * g.radsep;
}

// The vertical samples of a sub-ring (galmod.py:275-323), built cooperatively by the block into LDS.
// Returns nzz (the same value in every thread).
__device__ int block_zgrid(double* zq, double* wq, int* sh_n, double* sh_step, const bb_grid& g, const bb_box& b,
                           double z0, double sinc, double dvdz, double sig, double R0, double xpos, double ypos) {
    double s = fabs(sinc);
    int nmax = g.nzq;
    if (z0 <= 0 || nmax < 2 || s < 1e-6) {
        if (threadIdx.x == 0) {
            zq[0] = 0.0;
            wq[0] = 1.0;
        }
        __syncthreads();
        return 1;
    }
    double pixsize = sqrt(g.pixarea);
    double dx = fmax(fabs((double)b.blo0 - 0.5 - xpos), fabs((double)b.bhi0 - 0.5 - xpos)) * fabs(g.cdelt0);
    double dy = fmax(fabs((double)b.blo1 - 0.5 - ypos), fabs((double)b.bhi1 - 0.5 - ypos)) * fabs(g.cdelt1);
    double zl2 = (R0 + g.radsep + sqrt(dx * dx + dy * dy)) / (s * z0);
    double ext = bb_layer_extent(g.ltype);
    double zlim = ext < zl2 ? ext : zl2;
    if (threadIdx.x == 0) {
        double step = g.zstep * pixsize / (z0 * s);
        if (dvdz != 0) {
            double st2 = 0.5 * sig / (fabs(dvdz) * z0 * s);
            if (st2 < step) step = st2;
        }
        double zc = bb_layer_core(g.ltype);
        int64_t nin, nout;
        for (;;) {
            double m = zc < zlim ? zc : zlim;
            nin = (int64_t)ceil(m / step);
            if (nin < 1) nin = 1;
            nout = (int64_t)ceil((zlim - (double)nin * step) / (4.0 * step));
            if (nout < 0) nout = 0;
            if (2 * (nin + nout) <= nmax) break;
            step *= 1.01 * (double)(nin + nout) / (double)(nmax / 2);
        }
        sh_n[0] = (int)nin;
        sh_n[1] = (int)nout;
        *sh_step = step;
    }
    __syncthreads();
    int nin = sh_n[0], half = sh_n[0] + sh_n[1];
    double step = *sh_step;
    for (int j = threadIdx.x; j < half; j += blockDim.x) {
        double lo = j < nin ? (double)j * step : (double)(nin + 4 * (j - nin)) * step;
        double hi = lo + (j < nin ? step : 4.0 * step);
        if (hi > zlim) hi = zlim;
        double w = bb_layer_cdf(g.ltype, hi) - bb_layer_cdf(g.ltype, lo);
        zq[half + j] = 0.5 * (lo + hi);
        wq[half + j] = w;
        zq[half - 1 - j] = -0.5 * (lo + hi);
        wq[half - 1 - j] = w;
    }
    __syncthreads();
    return 2 * half;
}

// Deposit a*ws times the channel-integrated line (centre v, dispersion sig) into the spectrum o
// (galmod.py:194-214 and the deposit loops of 404-465).
__device__ __forceinline__ void deposit_line(float* __restrict__ o, double aw, double v, double sig,
                                             const double* __restrict__ ve, const double* __restrict__ cd2i,
                                             const double* __restrict__ ptab, int ntab, int64_t kc, int64_t K,
                                             int nz) {
    int64_t k0 = kc - K;
    if (k0 < 0) k0 = 0;
    int64_t k1 = kc + K + 1;
    if (k1 > nz) k1 = nz;
    if (k1 <= k0) return;
    double inv = 1.0 / sig;
    double pp = bb_phi((ve[k0] - v) * inv, ptab, ntab);
    for (int64_t k = k0; k < k1; k++) {
        double pn = bb_phi((ve[k + 1] - v) * inv, ptab, ntab);
        double f = pn - pp;
        pp = pn;
        double val = aw * ((f >= 0 ? f : -f) * cd2i[k]);
        atomicAdd(o + k, (float)val);
    }
}

__global__ void __launch_bounds__(BB_AN_THREADS)
k_galmod_analytic(bb_grid g, bb_box box, const double* __restrict__ srall, int nsub,
                  const double* __restrict__ cd2i, const double* __restrict__ vedges,
                  const double* __restrict__ ptab, int ntab, float* __restrict__ outall) {
    __shared__ double zq[BB_ZQ_MAX];
    __shared__ double wq[BB_ZQ_MAX];
    __shared__ int sh_n[2];
    __shared__ double sh_step;
    const int ir = blockIdx.y >> 1, kr = blockIdx.y & 1, b = blockIdx.z;
    const double* sr = srall + (size_t)b * 16 * nsub;
    const int nz = g.nz;
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    float* out = outall + (size_t)b * by * bx * nz;
    SubRing q;
    load_subring(q, sr, nsub, ir, g);
    if (q.dens == 0) return;
    if ((int64_t)blockIdx.x * blockDim.x >= q.naz) return;
    const int nzz = block_zgrid(zq, wq, sh_n, &sh_step, g, box, q.z0, q.sinc, q.dvdz, q.sig, q.R0, q.xpos, q.ypos);
    bool zdep = false;
    if (q.dvdz != 0 && nzz > 1)
        for (int kz = 0; kz < nzz; kz++)
            if (fabs(zq[kz] * q.z0) > q.zcyl) { zdep = true; break; }
    const double u = (kr + 0.5) / 2;
    const double R = sqrt(q.r_inner * q.r_inner + 2.0 * g.radsep * q.R0 * u);
    const double daz = 2 * BB_PI / (double)q.naz;
    const double az0 = 2 * BB_PI * (0.5 * ((kr + 1) % 2) + 0.25) / (double)q.naz;
    for (int64_t ka = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ka < q.naz;
         ka += (int64_t)gridDim.x * blockDim.x) {
        double saz, caz;
        sincos(az0 + (double)ka * daz, &saz, &caz);
        double wside = caz >= 0.0 ? q.wrec : q.wapp;
        if (wside == 0.0) continue;
        double x = R * caz, yb = R * saz * q.cinc, xs = x * q.spa, xc = x * q.cpa;
        double vplane = q.vsys + (q.vrot * caz + q.vrad * saz) * q.sinc + q.vvert * q.cinc;
        // (is some vertical sample in the box? galmod.py:402-409)
        bool inbox = false;
        for (int kz = 0; kz < nzz && !inbox; kz++) {
            double z = nzz > 1 ? zq[kz] * q.z0 : 0.0;
            double y = yb - z * q.sinc;
            int64_t g0 = bb_lround(q.xpos + (xs - y * q.cpa) / g.cdelt0);
            int64_t g1 = bb_lround(q.ypos + (xc + y * q.spa) / g.cdelt1);
            inbox = box.blo0 <= g0 && g0 < box.bhi0 && box.blo1 <= g1 && g1 < box.bhi1;
        }
        if (!inbox) continue;
        for (int nl = 0; nl < g.nlines; nl++) {
            const double* ve = vedges + (size_t)nl * (nz + 1);
            double a = g.relint[nl] * q.fluxsc * wside;
            if (!zdep) {
                double v = vplane;
                int64_t kc = bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
                int kz = 0;
                while (kz < nzz) {
                    double z = nzz > 1 ? zq[kz] * q.z0 : 0.0;
                    double y = yb - z * q.sinc;
                    int64_t g0 = bb_lround(q.xpos + (xs - y * q.cpa) / g.cdelt0);
                    int64_t g1 = bb_lround(q.ypos + (xc + y * q.spa) / g.cdelt1);
                    double ws = wq[kz];
                    kz++;
                    while (kz < nzz) {
                        double z2 = zq[kz] * q.z0;
                        double y2 = yb - z2 * q.sinc;
                        int64_t h0 = bb_lround(q.xpos + (xs - y2 * q.cpa) / g.cdelt0);
                        int64_t h1 = bb_lround(q.ypos + (xc + y2 * q.spa) / g.cdelt1);
                        if (h0 != g0 || h1 != g1) break;
                        ws += wq[kz];
                        kz++;
                    }
                    if (g0 < box.blo0 || g0 >= box.bhi0 || g1 < box.blo1 || g1 >= box.bhi1) continue;
                    float* o = out + ((size_t)(g1 - box.blo1) * bx + (size_t)(g0 - box.blo0)) * nz;
                    deposit_line(o, a * ws, v, q.sig, ve, cd2i, ptab, ntab, kc, q.K, nz);
                }
            } else {
                // the velocity depends on |z|: the samples at +z and -z share the line (galmod.py:432-465)
                for (int kz = 0; kz < (nzz + 1) / 2; kz++) {
                    int kz2 = nzz - 1 - kz;
                    double za = zq[kz] * q.z0, zb = zq[kz2] * q.z0;
                    double ya = yb - za * q.sinc, ybb = yb - zb * q.sinc;
                    int64_t ga0 = bb_lround(q.xpos + (xs - ya * q.cpa) / g.cdelt0);
                    int64_t ga1 = bb_lround(q.ypos + (xc + ya * q.spa) / g.cdelt1);
                    int64_t gb0 = bb_lround(q.xpos + (xs - ybb * q.cpa) / g.cdelt0);
                    int64_t gb1 = bb_lround(q.ypos + (xc + ybb * q.spa) / g.cdelt1);
                    bool in1 = !(ga0 < box.blo0 || ga0 >= box.bhi0 || ga1 < box.blo1 || ga1 >= box.bhi1);
                    bool in2 = kz2 != kz && !(gb0 < box.blo0 || gb0 >= box.bhi0 || gb1 < box.blo1 || gb1 >= box.bhi1);
                    if (!(in1 || in2)) continue;
                    double az_ = fabs(za), v;
                    if (az_ > q.zcyl)
                        v = q.vsys + (bb_lagged(q.vrot, q.dvdz, az_ - q.zcyl) * caz + q.vrad * saz) * q.sinc +
                            q.vvert * q.cinc;
                    else
                        v = vplane;
                    int64_t kc =
                        bb_lround(bb_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3, g.crpix3));
                    if (in1 && in2 && ga0 == gb0 && ga1 == gb1) {
                        float* o = out + ((size_t)(ga1 - box.blo1) * bx + (size_t)(ga0 - box.blo0)) * nz;
                        deposit_line(o, a * (wq[kz] + wq[kz2]), v, q.sig, ve, cd2i, ptab, ntab, kc, q.K, nz);
                        continue;
                    }
                    if (in1) {
                        float* o = out + ((size_t)(ga1 - box.blo1) * bx + (size_t)(ga0 - box.blo0)) * nz;
                        deposit_line(o, a * wq[kz], v, q.sig, ve, cd2i, ptab, ntab, kc, q.K, nz);
                    }
                    if (in2) {
                        float* o = out + ((size_t)(gb1 - box.blo1) * bx + (size_t)(gb0 - box.blo0)) * nz;
                        deposit_line(o, a * wq[kz2], v, q.sig, ve, cd2i, ptab, ntab, kc, q.K, nz);
                    }
                }
            }
        }
    }
}

extern "C" int bb_launch_galmod_analytic(const bb_grid* g, bb_box box, const double* sr, int nsub, int B,
                                         int chunks, const double* cd2i, const double* vedges, const double* ptab,
                                         int ntab, float* out, hipStream_t st) {
    dim3 grid(chunks, 2 * nsub, B);
    hipLaunchKernelGGL(k_galmod_analytic, grid, dim3(BB_AN_THREADS), 0, st, *g, box, sr, nsub, cd2i, vedges, ptab,
                       ntab, out);
    return (int)hipGetLastError();
}
