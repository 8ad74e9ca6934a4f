/*
 * ORACLE -- test infrastructure only. Never part of the product path.
 *
 * A plain-C FP64 restatement of the reference's numba kernels on the model-cube likelihood path, used
 * by tests/ (as the checker), by __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Each
 * function cites the reference lines it follows (paths relative to the reference repository).
 * Compiled with -ffp-contract=off so that every multiply and add rounds as in numba's code.
 *
 * Layouts are the reference's: models (ny, nx, nz) "yxz" float64 accumulated with +=, data (nz, ny, nx)
 * float32, the sub-ring table sr (16, nsub) float64 row-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define C_MS 2.99792458e8
#define SR(row, i) sr[(size_t)(row) * nsub + (i)]

/* galmod.py:38-46 */
static inline int64_t lround_bb(double x) {
    if (!(fabs(x) < 1e18)) return -((int64_t)1 << 62);
    if (x >= 0) return (int64_t)(x + 0.5);
    return -(int64_t)(-x + 0.5);
}

/* galmod.py:49-62 */
static inline double velgrid(double v, int axtyp, int velsys, double freq0, double crval3, double cdelt3,
                             double crpix3) {
    double w;
    if (axtyp == 4) {
        w = v;
    } else {
        if (velsys == 1) w = freq0 * (1 - v / C_MS);
        else if (velsys == 2) w = freq0 * (C_MS / (C_MS + v));
        else w = freq0 * sqrt((C_MS - v) / (C_MS + v));
        if (axtyp == 2) w = C_MS / w;
    }
    return (w - crval3) / cdelt3 + crpix3 - 1;
}

/* galmod.py:65-73 */
static inline double lagged(double vrot, double dvdz, double dz) {
    double v = vrot - dvdz * dz;
    if (vrot >= 0.0) return v > 0.0 ? v : 0.0;
    return v < 0.0 ? v : 0.0;
}

/* galmod.py:183-191 */
static inline double phi_tab(double t, const double* tab, int ntab) {
    double x = (t + 6.0) * 512.0;
    if (x <= 0.0) return 0.0;
    int64_t k = (int64_t)x;
    if (k >= ntab - 1) return 1.0;
    return tab[k] + (x - (double)k) * (tab[k + 1] - tab[k]);
}

/* galmod.py:194-214: returns nk, writes k0 */
static int line_profile(double* prof, double v, double sig, const double* vedges, const double* cd2i,
                        const double* ptab, int ntab, int64_t kc, int64_t K, int nz, int64_t* k0out) {
    int64_t k0 = kc - K;
    if (k0 < 0) k0 = 0;
    int64_t k1 = kc + K + 1;
    if (k1 > nz) k1 = nz;
    *k0out = k0;
    if (k1 <= k0) return 0;
    double inv = 1.0 / sig;
    double pp = phi_tab((vedges[k0] - v) * inv, ptab, ntab);
    for (int64_t k = k0; k < k1; k++) {
        double pn = phi_tab((vedges[k + 1] - v) * inv, ptab, ntab);
        double f = pn - pp;
        prof[k - k0] = (f >= 0 ? f : -f) * cd2i[k];
        pp = pn;
    }
    return (int)(k1 - k0);
}

/* galmod.py:217-258 */
static double layer_cdf(int ltype, double x) {
    if (ltype == 1) return 0.5 * (1.0 + erf(x / sqrt(2.0)));
    if (ltype == 2) return 0.5 * (1.0 + tanh(x));
    if (ltype == 3) return x >= 0 ? 1.0 - 0.5 * exp(-x) : 0.5 * exp(x);
    if (ltype == 4) return 0.5 + atan(x) / M_PI;
    double c = x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x);
    return 0.5 * (c + 1.0);
}
static double layer_extent(int ltype) {
    if (ltype == 1) return 5.0;
    if (ltype == 2) return 7.6;
    if (ltype == 3) return 14.0;
    if (ltype == 4) return 1e30;
    return 1.0;
}
static double layer_core(int ltype) {
    if (ltype == 1) return 3.717;
    if (ltype == 2) return 4.147;
    if (ltype == 3) return 6.908;
    if (ltype == 4) return 31.61;
    return 1e30;
}

/* galmod.py:275-323 */
int orc_zgrid(double* zq, double* wq, int nmax, int ltype, double z0, double sinc, double pixsize, double dvdz,
              double sig, double R0, double radsep, double xpos, double ypos, double cdelt0, double cdelt1,
              int64_t blo0, int64_t blo1, int64_t bhi0, int64_t bhi1, double zstep) {
    double s = fabs(sinc);
    if (z0 <= 0 || nmax < 2 || s < 1e-6) {
        zq[0] = 0.0;
        wq[0] = 1.0;
        return 1;
    }
    double dx = fmax(fabs((double)blo0 - 0.5 - xpos), fabs((double)bhi0 - 0.5 - xpos)) * fabs(cdelt0);
    double dy = fmax(fabs((double)blo1 - 0.5 - ypos), fabs((double)bhi1 - 0.5 - ypos)) * fabs(cdelt1);
    double zl2 = (R0 + radsep + sqrt(dx * dx + dy * dy)) / (s * z0);
    double zlim = layer_extent(ltype) < zl2 ? layer_extent(ltype) : zl2;
    double step = zstep * pixsize / (z0 * s);
    if (dvdz != 0) {
        double st2 = 0.5 * sig / (fabs(dvdz) * z0 * s);
        if (st2 < step) step = st2;
    }
    double zc = layer_core(ltype);
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
    int64_t half = nin + nout;
    for (int64_t j = 0; j < half; j++) {
        double lo = j < nin ? (double)j * step : (double)(nin + 4 * (j - nin)) * step;
        double hi = lo + (j < nin ? step : 4.0 * step);
        if (hi > zlim) hi = zlim;
        double w = layer_cdf(ltype, hi) - layer_cdf(ltype, lo);
        zq[half + j] = 0.5 * (lo + hi);
        wq[half + j] = w;
        zq[half - 1 - j] = -0.5 * (lo + hi);
        wq[half - 1 - j] = w;
    }
    return (int)(2 * half);
}

typedef struct {
    int nx, ny, nz;            /* box model dims: out (by, bx, nz) */
    double radsep, cdens, pixarea, cdelt0, cdelt1, crota2;
    int64_t blo0, blo1, bhi0, bhi1;
    int ltype;
    double sig_instr;
    int axtyp, velsys, nlines;
    double crval3, cdelt3, crpix3, chwidth, zstep;
} orc_grid;

/* galmod.py:326-465 (ANALYTIC GALMOD). out (by, bx, nz) float64, +=. vedges (nlines, nz+1). */
void orc_galmod_analytic(double* out, const double* sr, int nsub, int ring_lo, int ring_hi, const orc_grid* g,
                         const double* vedges, const double* relint, const double* cd2i, const double* lfreq0,
                         int nzq, const double* ptab, int ntab) {
    const int nz = g->nz, bx = (int)(g->bhi0 - g->blo0);
    const int64_t blo0 = g->blo0, blo1 = g->blo1, bhi0 = g->bhi0, bhi1 = g->bhi1;
    const int nlines = g->nlines;
    double pixsize = sqrt(g->pixarea);
    int nq = nzq > 1 ? nzq : 1;
    double* zq = (double*)malloc(sizeof(double) * nq);
    double* wq = (double*)malloc(sizeof(double) * nq);
    double* prof = (double*)malloc(sizeof(double) * (nz + 1));
    int64_t* gx = (int64_t*)malloc(sizeof(int64_t) * nq);
    int64_t* gy = (int64_t*)malloc(sizeof(int64_t) * nq);
    for (int ir = ring_lo; ir < ring_hi; ir++) {
        double R0 = SR(0, ir), dens = SR(5, ir);
        if (dens == 0) continue;
        double ring_area_pix = 2 * M_PI * R0 * g->radsep / g->pixarea;
        double circ_pix = 2 * M_PI * R0 / sqrt(g->pixarea);
        int64_t naz = (int64_t)(3 * circ_pix + 0.5);
        if (naz < 16) naz = 16;
        const int nrad = 2;
        double vrot = SR(1, ir), vrad = SR(2, ir), vvert = SR(3, ir);
        double sig = sqrt(SR(4, ir) * SR(4, ir) + g->sig_instr * g->sig_instr);
        double sgmin = 1e-6 * fabs(g->chwidth);
        if (sig < sgmin) sig = sgmin;
        int64_t K = (int64_t)(5.0 * sig / g->chwidth) + 2;
        double z0 = SR(6, ir), dvdz = SR(7, ir), zcyl = SR(8, ir);
        double sinc = sin(SR(9, ir)), cinc = cos(SR(9, ir));
        double phi = SR(10, ir);
        double spa = sin(phi) * cos(g->crota2) + cos(phi) * sin(g->crota2);
        double cpa = cos(phi) * cos(g->crota2) - sin(phi) * sin(g->crota2);
        double xpos = SR(11, ir), ypos = SR(12, ir), vsys0 = SR(13, ir);
        int nzz = orc_zgrid(zq, wq, nzq, g->ltype, z0, sinc, pixsize, dvdz, sig, R0, g->radsep, xpos, ypos,
                            g->cdelt0, g->cdelt1, blo0, blo1, bhi0, bhi1, g->zstep);
        double fluxsc = dens * ring_area_pix * g->pixarea / (double)(naz * nrad);
        int zdep = 0;
        if (dvdz != 0 && nzz > 1)
            for (int kz = 0; kz < nzz; kz++)
                if (fabs(zq[kz] * z0) > zcyl) { zdep = 1; break; }
        double r_inner = R0 - 0.5 * g->radsep;
        double daz = 2 * M_PI / (double)naz;
        double cd = cos(daz), sd = sin(daz);
        for (int kr = 0; kr < nrad; kr++) {
            double u = (kr + 0.5) / nrad;
            double R = sqrt(r_inner * r_inner + 2.0 * g->radsep * R0 * u);
            double az0 = 2 * M_PI * (0.5 * ((kr + 1) % 2) + 0.25) / (double)naz;
            double caz = cos(az0), saz = sin(az0);
            for (int64_t ka = 0; ka < naz; ka++) {
                if (ka > 0) {
                    double c2 = caz * cd - saz * sd, s2 = saz * cd + caz * sd;
                    caz = c2; saz = s2;
                }
                double wside = caz >= 0.0 ? SR(14, ir) : SR(15, ir);
                if (wside == 0.0) continue;
                double x = R * caz, yb = R * saz * cinc, xs = x * spa, xc = x * cpa;
                for (int kz = 0; kz < nzz; kz++) {
                    double z = nzz > 1 ? zq[kz] * z0 : 0.0;
                    double y = yb - z * sinc;
                    gx[kz] = lround_bb(xpos + (xs - y * cpa) / g->cdelt0);
                    gy[kz] = lround_bb(ypos + (xc + y * spa) / g->cdelt1);
                }
                double vplane = vsys0 + (vrot * caz + vrad * saz) * sinc + vvert * cinc;
                int inbox = 0;
                for (int kz = 0; kz < nzz; kz++)
                    if (blo0 <= gx[kz] && gx[kz] < bhi0 && blo1 <= gy[kz] && gy[kz] < bhi1) { inbox = 1; break; }
                if (!inbox) continue;
                for (int nl = 0; nl < nlines; nl++) {
                    const double* ve = vedges + (size_t)nl * (nz + 1);
                    double a = relint[nl] * fluxsc * wside;
                    if (!zdep) {
                        double v = vplane;
                        int64_t kc = lround_bb(velgrid(v, g->axtyp, g->velsys, lfreq0[nl], g->crval3, g->cdelt3,
                                                       g->crpix3));
                        int64_t k0;
                        int nk = line_profile(prof, v, sig, ve, cd2i, ptab, ntab, kc, K, nz, &k0);
                        if (nk == 0) continue;
                        int kz = 0;
                        while (kz < nzz) {
                            int64_t g0 = gx[kz], g1 = gy[kz];
                            double ws = wq[kz];
                            int w = 1;
                            while (kz + w < nzz && gx[kz + w] == g0 && gy[kz + w] == g1) {
                                ws += wq[kz + w];
                                w++;
                            }
                            kz += w;
                            if (g0 < blo0 || g0 >= bhi0 || g1 < blo1 || g1 >= bhi1) continue;
                            double aw = a * ws;
                            double* o = out + ((size_t)(g1 - blo1) * bx + (size_t)(g0 - blo0)) * nz;
                            for (int j = 0; j < nk; j++) o[k0 + j] += aw * prof[j];
                        }
                    } else {
                        for (int kz = 0; kz < (nzz + 1) / 2; kz++) {
                            int kz2 = nzz - 1 - kz;
                            int in1 = !(gx[kz] < blo0 || gx[kz] >= bhi0 || gy[kz] < blo1 || gy[kz] >= bhi1);
                            int in2 = kz2 != kz &&
                                      !(gx[kz2] < blo0 || gx[kz2] >= bhi0 || gy[kz2] < blo1 || gy[kz2] >= bhi1);
                            if (!(in1 || in2)) continue;
                            double az_ = fabs(zq[kz] * z0), v;
                            if (az_ > zcyl)
                                v = vsys0 + (lagged(vrot, dvdz, az_ - zcyl) * caz + vrad * saz) * sinc + vvert * cinc;
                            else
                                v = vplane;
                            int64_t kc = lround_bb(velgrid(v, g->axtyp, g->velsys, lfreq0[nl], g->crval3,
                                                           g->cdelt3, g->crpix3));
                            int64_t k0;
                            int nk = line_profile(prof, v, sig, ve, cd2i, ptab, ntab, kc, K, nz, &k0);
                            if (nk == 0) continue;
                            if (in1 && in2 && gx[kz] == gx[kz2] && gy[kz] == gy[kz2]) {
                                double* o = out + ((size_t)(gy[kz] - blo1) * bx + (size_t)(gx[kz] - blo0)) * nz;
                                double aw = a * (wq[kz] + wq[kz2]);
                                for (int j = 0; j < nk; j++) o[k0 + j] += aw * prof[j];
                                continue;
                            }
                            if (in1) {
                                double* o = out + ((size_t)(gy[kz] - blo1) * bx + (size_t)(gx[kz] - blo0)) * nz;
                                double aw = a * wq[kz];
                                for (int j = 0; j < nk; j++) o[k0 + j] += aw * prof[j];
                            }
                            if (in2) {
                                double* o = out + ((size_t)(gy[kz2] - blo1) * bx + (size_t)(gx[kz2] - blo0)) * nz;
                                double aw = a * wq[kz2];
                                for (int j = 0; j < nk; j++) o[k0 + j] += aw * prof[j];
                            }
                        }
                    }
                }
            }
        }
    }
    free(zq); free(wq); free(prof); free(gx); free(gy);
}

/* galmod.py:468-536 (unit-density total-intensity maps per ring). out (nr, by, bx), +=.
 * widx/wval (nsub, 2). */
void orc_ring_maps_analytic(double* out, int nr, const double* sr, int nsub, const int64_t* widx,
                            const double* wval, const orc_grid* g, int nzq) {
    const int64_t blo0 = g->blo0, blo1 = g->blo1, bhi0 = g->bhi0, bhi1 = g->bhi1;
    const int bx = (int)(bhi0 - blo0), by = (int)(bhi1 - blo1);
    (void)nr;
    double pixsize = sqrt(g->pixarea);
    int nq = nzq > 1 ? nzq : 1;
    double* zq = (double*)malloc(sizeof(double) * nq);
    double* wq = (double*)malloc(sizeof(double) * nq);
    int64_t* gx = (int64_t*)malloc(sizeof(int64_t) * nq);
    int64_t* gy = (int64_t*)malloc(sizeof(int64_t) * nq);
    for (int ir = 0; ir < nsub; ir++) {
        if (widx[2 * ir] < 0 && widx[2 * ir + 1] < 0) continue;
        double R0 = SR(0, ir);
        double ring_area_pix = 2 * M_PI * R0 * g->radsep / g->pixarea;
        double circ_pix = 2 * M_PI * R0 / sqrt(g->pixarea);
        int64_t naz = (int64_t)(3 * circ_pix + 0.5);
        if (naz < 16) naz = 16;
        const int nrad = 2;
        double z0 = SR(6, ir);
        double sinc = sin(SR(9, ir));
        double phi = SR(10, ir);
        double spa = sin(phi) * cos(g->crota2) + cos(phi) * sin(g->crota2);
        double cpa = cos(phi) * cos(g->crota2) - sin(phi) * sin(g->crota2);
        double xpos = SR(11, ir), ypos = SR(12, ir);
        double sig = sqrt(SR(4, ir) * SR(4, ir) + g->sig_instr * g->sig_instr);
        double sgmin = 1e-6 * fabs(g->chwidth);
        if (sig < sgmin) sig = sgmin;
        double cinc = cos(SR(9, ir));
        int nzz = orc_zgrid(zq, wq, nzq, g->ltype, z0, sinc, pixsize, SR(7, ir), sig, R0, g->radsep, xpos, ypos,
                            g->cdelt0, g->cdelt1, blo0, blo1, bhi0, bhi1, g->zstep);
        double fluxsc = ring_area_pix * g->pixarea / (double)(naz * nrad);
        double r_inner = R0 - 0.5 * g->radsep;
        double daz = 2 * M_PI / (double)naz;
        double cd = cos(daz), sd = sin(daz);
        for (int kr = 0; kr < nrad; kr++) {
            double u = (kr + 0.5) / nrad;
            double R = sqrt(r_inner * r_inner + 2.0 * g->radsep * R0 * u);
            double az0 = 2 * M_PI * (0.5 * ((kr + 1) % 2) + 0.25) / (double)naz;
            double caz = cos(az0), saz = sin(az0);
            for (int64_t ka = 0; ka < naz; ka++) {
                if (ka > 0) {
                    double c2 = caz * cd - saz * sd, s2 = saz * cd + caz * sd;
                    caz = c2; saz = s2;
                }
                double wside = caz >= 0.0 ? SR(14, ir) : SR(15, ir);
                if (wside == 0.0) continue;
                double x = R * caz, yb = R * saz * cinc, xs = x * spa, xc = x * cpa;
                for (int kz = 0; kz < nzz; kz++) {
                    double z = nzz > 1 ? zq[kz] * z0 : 0.0;
                    double y = yb - z * sinc;
                    gx[kz] = lround_bb(xpos + (xs - y * cpa) / g->cdelt0);
                    gy[kz] = lround_bb(ypos + (xc + y * spa) / g->cdelt1);
                }
                int kz = 0;
                while (kz < nzz) {
                    int64_t g0 = gx[kz], g1 = gy[kz];
                    double ws = wq[kz];
                    kz++;
                    while (kz < nzz && gx[kz] == g0 && gy[kz] == g1) { ws += wq[kz]; kz++; }
                    if (g0 < blo0 || g0 >= bhi0 || g1 < blo1 || g1 >= bhi1) continue;
                    for (int m = 0; m < 2; m++) {
                        int64_t r = widx[2 * ir + m];
                        if (r >= 0)
                            out[((size_t)r * by + (size_t)(g1 - blo1)) * bx + (size_t)(g0 - blo0)] +=
                                fluxsc * ws * wval[2 * ir + m] * wside;
                    }
                }
            }
        }
    }
    free(zq); free(wq); free(gx); free(gy);
}

/* 'Same' 2D convolution of each channel of a yxz model with kern (ky, kx): the direct-sum definition of
 * FFTConvolver (smooth.py:145-165: full linear convolution cropped at [k//2, k//2 + n)), then values below
 * 1e-12 set to zero (thresh != 0; the ring maps of DiskModel.ring_maps, galmod.py:1342-1345, are not
 * thresholded) and the scale applied (galmod.py:1273-1283, and _conv_sparse galmod.py:974-1037).
 * in/out (ny, nx, nz) float64. */
void orc_conv_same(const double* in, double* out, int ny, int nx, int nz, const double* kern, int ky, int kx,
                   double scale, int thresh) {
    int hy = ky / 2, hx = kx / 2;
    memset(out, 0, sizeof(double) * (size_t)ny * nx * nz);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const double* m = in + ((size_t)y * nx + x) * nz;
            int nonzero = 0;
            for (int z = 0; z < nz; z++) if (m[z] != 0) { nonzero = 1; break; }
            if (!nonzero) continue;
            for (int j = 0; j < ky; j++)
                for (int i = 0; i < kx; i++) {
                    double k = kern[j * kx + i];
                    if (k == 0) continue;
                    int yy = y + j - hy, xx = x + i - hx;
                    if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
                    double* o = out + ((size_t)yy * nx + xx) * nz;
                    for (int z = 0; z < nz; z++) o[z] += k * m[z];
                }
        }
    size_t n = (size_t)ny * nx * nz;
    for (size_t i = 0; i < n; i++) {
        double v = out[i];
        if (thresh && v < 1.E-12) out[i] = 0.;
        else if (scale != 1.0) out[i] = v * scale;
    }
}

/* bayesian/likelihood.py:22-45. model (nz, ny, nx) f64; data f32; mask u8 (or NULL with maskmode 0);
 * region (ny, nx) u8 or NULL. */
double orc_sumsq_residuals(const double* model, const float* data, const uint8_t* mask, int maskmode,
                           const uint8_t* region, int nz, int ny, int nx) {
    double s = 0.;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (region && !region[(size_t)j * nx + i]) continue;
                size_t v = ((size_t)k * ny + j) * nx + i;
                double d = data[v];
                if (d != d || d == 0.) continue;
                if (maskmode != 0 && !mask[v]) {
                    if (maskmode == 2) continue;
                    d = 0.;
                }
                double r = d - model[v];
                s += r * r;
            }
    return s;
}

/* bayesian/likelihood.py:48-77. Returns s; *lg = the log-variance sum. */
double orc_residuals_fracerr(const double* model, const float* data, const uint8_t* mask, int maskmode,
                             const uint8_t* region, int nz, int ny, int nx, double sigma2, double eps2,
                             double* lg) {
    double s = 0., l = 0.;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (region && !region[(size_t)j * nx + i]) continue;
                size_t v = ((size_t)k * ny + j) * nx + i;
                double d = data[v];
                if (d != d || d == 0.) continue;
                if (maskmode != 0 && !mask[v]) {
                    if (maskmode == 2) continue;
                    d = 0.;
                }
                double m = model[v], r = d - m;
                if (m != 0.) {
                    double var = sigma2 + eps2 * m * m;
                    s += r * r / var;
                    l += log(var / sigma2);
                } else {
                    s += r * r / sigma2;
                }
            }
    *lg = l;
    return s;
}
