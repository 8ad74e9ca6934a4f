/*
 * ORACLE -- test infrastructure only. FP64 C restatement of the 3DFIT residual of one ring (_norm,
 * bbarolo/galfit.py:39-188, with the support arguments of Galfit._support_args, galfit.py:685-723):
 *   - the pixels of the ring: the annulus r1 <= R <= r2 in the plane of the disk (foot 0) or the strip
 *     r1 <= |x| <= r2, |y| <= hmax along the major axis (foot 1), optionally one side (1 receding |th| <= 90,
 *     2 approaching |th| >= 90); weights |cos th|^wpow (|sin th|^|wpow| for wpow < 0), 1 in a strip;
 *   - the normalisation of the model: LOCAL (0) per pixel, sum of the positive masked data / sum of the model;
 *     AZIM (1) one factor for the ring; NONE (2);
 *   - the voxels: masked data, or model voxels outside the mask (blanks; with a support, only where the
 *     expected number of sub-clouds over the kernel footprint, lam, reaches thr[z]);
 *   - f = (1 + Nb/N)^bweight sum / (N - Nb), the sum of w x (mod-obs)^2/sqrt(obs)/n (ftype 1), |mod-obs|/n (2),
 *     |mod-obs|/(mod+obs)/n (3), (mod-obs)^2/n (4); obs = data if > 0 else data_noise; +inf if N = Nb.
 * The model (by, bx, nz) is scaled in place, as the reference does.
 *
 * Pinned by tests/test_oracle_norm.py: hand-computed values on tiny cubes for each ftype and normalisation, the
 * ring geometry against a brute-force sampler, the blank and support rules, the +inf case.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

double orc_norm(double* model, int by, int bx, int nz, const float* data, const uint8_t* mask, int ny, int nx,
                int64_t blo0, int64_t blo1, double x0, double y0, double phi, double inc, double r1, double r2,
                double wpow, int side, int ftype, const double* chan_noise, double data_noise, double bweight,
                int normtype, const double* pre, const int64_t* ilo, const int64_t* ihi, int ky, const double* thr,
                int foot, double hmax, int64_t* nblank_out, int64_t* ntot_out) {
    (void)ny;
    const double F = M_PI / 180.;
    const double sphi = sin(F * phi), cphi = cos(F * phi), cinc = cos(F * inc);
    const int usesin = wpow < 0;
    const double apow = fabs(wpow);
    const int hy = ky / 2;
    uint8_t* inside = (uint8_t*)calloc((size_t)by * bx, 1);
    double* wmap = (double*)calloc((size_t)by * bx, sizeof(double));
    for (int y = 0; y < by; y++)
        for (int x = 0; x < bx; x++) {
            const double xr = -(x - x0) * sphi + (y - y0) * cphi;
            const double ys = -(x - x0) * cphi - (y - y0) * sphi;
            const double yr = ys / cinc;
            const double r = sqrt(xr * xr + yr * yr);
            double th = r < 0.1 ? 0.0 : atan2(yr, xr) / F;
            if (foot == 1) {
                if (!(fabs(xr) >= r1 && fabs(xr) <= r2 && fabs(ys) <= hmax)) continue;
                th = xr >= 0 ? 0.0 : 180.0;
            } else if (!(r >= r1 && r <= r2)) {
                continue;
            }
            if (side == 1 && !(fabs(th) <= 90.0)) continue;
            if (side == 2 && !(fabs(th) >= 90.0)) continue;
            inside[(size_t)y * bx + x] = 1;
            if (foot == 1) {
                wmap[(size_t)y * bx + x] = 1.0;
                continue;
            }
            const double wf = usesin ? fabs(sin(th * F)) : fabs(cos(th * F));
            wmap[(size_t)y * bx + x] = pow(wf, apow);
        }
#define MOD(y, x, z) model[(((size_t)(y)) * bx + (x)) * nz + (z)]
#define DAT(y, x, z) data[(((size_t)((y) + blo1)) * nx + ((x) + blo0)) * nz + (z)]
#define MSK(y, x, z) mask[(((size_t)((y) + blo1)) * nx + ((x) + blo0)) * nz + (z)]
    double factor = 1.0;
    if (normtype == 1) {
        double ms = 0, os = 0;
        for (int y = 0; y < by; y++)
            for (int x = 0; x < bx; x++) {
                if (!inside[(size_t)y * bx + x]) continue;
                for (int z = 0; z < nz; z++) {
                    ms += MOD(y, x, z);
                    if (MSK(y, x, z)) os += DAT(y, x, z);
                }
            }
        factor = ms != 0 ? os / ms : 0.0;
    }
    double* lam = (double*)calloc((size_t)nz, sizeof(double));
    int64_t nb = 0, ntot = 0;
    double minfunc = 0.0;
    for (int y = 0; y < by; y++)
        for (int x = 0; x < bx; x++) {
            if (!inside[(size_t)y * bx + x]) continue;
            if (normtype == 0) {
                double ms = 0, os = 0;
                for (int z = 0; z < nz; z++) {
                    ms += MOD(y, x, z);
                    const double o = DAT(y, x, z);
                    if (o > 0 && MSK(y, x, z)) os += o;
                }
                factor = ms != 0 ? os / ms : 0.0;
            }
            const double wi = wmap[(size_t)y * bx + x];
            if (pre) {
                /* lam[z]: the sum of pre over the kernel footprint (rows y+hy-j, columns x+ilo[j]..x+ihi[j]) */
                for (int z = 0; z < nz; z++) lam[z] = 0;
                for (int j = 0; j < ky; j++) {
                    const int yy = y + hy - j;
                    if (yy < 0 || yy >= by) continue;
                    int64_t xa = x + ilo[j], xb = x + ihi[j] + 1;
                    xa = xa < 0 ? 0 : xa;
                    xb = xb > bx ? bx : xb;
                    for (int64_t xx = xa; xx < xb; xx++)
                        for (int z = 0; z < nz; z++) lam[z] += pre[(((size_t)yy) * bx + xx) * nz + z];
                }
            }
            for (int z = 0; z < nz; z++) {
                if (normtype != 2) MOD(y, x, z) *= factor;
                const double o = DAT(y, x, z);
                const double obs = o > 0 ? o : data_noise;
                const double mod = MOD(y, x, z);
                if (!MSK(y, x, z)) {
                    if (mod == 0) continue;
                    if (pre && lam[z] < thr[z]) continue;
                    nb++;
                }
                ntot++;
                const double nw = chan_noise[z];
                if (ftype == 1) minfunc += wi * (mod - obs) * (mod - obs) / sqrt(obs) / nw;
                else if (ftype == 2) minfunc += wi * fabs(mod - obs) / nw;
                else if (ftype == 3) minfunc += wi * fabs(mod - obs) / (mod + obs) / nw;
                else if (ftype == 4) minfunc += wi * (mod - obs) * (mod - obs) / nw;
            }
        }
#undef MOD
#undef DAT
#undef MSK
    free(inside);
    free(wmap);
    free(lam);
    if (nblank_out) *nblank_out = nb;
    if (ntot_out) *ntot_out = ntot;
    if (ntot - nb == 0) return INFINITY;
    return pow(1.0 + (double)nb / (double)ntot, bweight) * minfunc / (double)(ntot - nb);
}
