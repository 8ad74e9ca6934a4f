// Device-side helpers shared by the GALMOD kernels: rounding, channel grid, lag, the tabulated normal
// CDF and the vertical layer profiles. FP64, in the operation order of the reference
// (bbarolo/galmod.py:38-89, 174-258) so that pixel and channel indices round as there.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BB_C_MS 2.99792458e8
#define BB_PI 3.141592653589793

// Round half away from zero; NaN or |x| >= 1e18 -> a sentinel far outside any grid (galmod.py:38-46).
// Values within +-1e9 (every pixel or channel of a cube, and far beyond) round with 32-bit conversions; the
// others take the 64-bit path, so the result is the reference's for every input.
__device__ __forceinline__ int64_t bb_lround(double x) {
    const double a = fabs(x);
    if (a < 1e9) {
        const int r = (int)(a + 0.5);
        return x >= 0 ? (int64_t)r : -(int64_t)r;
    }
    if (!(a < 1e18)) return -((int64_t)1 << 62);
    if (x >= 0) return (int64_t)(x + 0.5);
    return -(int64_t)(-x + 0.5);
}

// Velocity (m/s) -> fractional channel (galmod.py:49-62).
__device__ __forceinline__ double bb_velgrid(double v, int axtyp, int velsys, double freq0, double crval3,
                                             double cdelt3, double crpix3) {
    double w;
    if (axtyp == 4) {
        w = v;
    } else {
        if (velsys == 1) w = freq0 * (1 - v / BB_C_MS);
        else if (velsys == 2) w = freq0 * (BB_C_MS / (BB_C_MS + v));
        else w = freq0 * sqrt((BB_C_MS - v) / (BB_C_MS + v));
        if (axtyp == 2) w = BB_C_MS / w;
    }
    return (w - crval3) / cdelt3 + crpix3 - 1;
}

// Rotation lagged by dvdz above zcyl, never reversing its sense (galmod.py:65-73).
__device__ __forceinline__ double bb_lagged(double vrot, double dvdz, double dz) {
    double v = vrot - dvdz * dz;
    if (vrot >= 0.0) return v > 0.0 ? v : 0.0;
    return v < 0.0 ? v : 0.0;
}

// Linear interpolation in the 1/512 table of Phi over [-6, 6] (galmod.py:183-191).
__device__ __forceinline__ double bb_phi(double t, const double* __restrict__ tab, int ntab) {
    double x = (t + 6.0) * 512.0;
    if (!(x > 0.0)) return 0.0;                     // (also NaN: never an index outside the table)
    if (!(x < (double)(ntab - 1))) return 1.0;
    int64_t k = (int64_t)x;
    double a = tab[k];
    return a + (x - (double)k) * (tab[k + 1] - a);
}

// Fraction of the layer flux below height x (units of z0), galmod.py:217-231.
__device__ __forceinline__ double bb_layer_cdf(int ltype, double x) {
    if (ltype == 1) return 0.5 * (1.0 + erf(x / sqrt(2.0)));
    if (ltype == 2) return 0.5 * (1.0 + tanh(x));
    if (ltype == 3) return x >= 0 ? 1.0 - 0.5 * exp(-x) : 0.5 * exp(x);
    if (ltype == 4) return 0.5 + atan(x) / BB_PI;
    double c = x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x);
    return 0.5 * (c + 1.0);
}
__device__ __forceinline__ double bb_layer_extent(int ltype) {
    return ltype == 1 ? 5.0 : ltype == 2 ? 7.6 : ltype == 3 ? 14.0 : ltype == 4 ? 1e30 : 1.0;
}
__device__ __forceinline__ double bb_layer_core(int ltype) {
    return ltype == 1 ? 3.717 : ltype == 2 ? 4.147 : ltype == 3 ? 6.908 : ltype == 4 ? 31.61 : 1e30;
}

// The (tile, vector) of a workgroup of a (tiles, vectors) grid. The dispatcher deals workgroups to the 8 XCDs
// in turn by linear id, and each XCD has its own L2; renumbered, XCD x takes the x-th eighth of the (vector,
// tile) pairs in order, so the tiles of one vector run on one XCD and read its chunk tables, boxes and sub-ring
// records through one L2 instead of eight. The ids past the last multiple of 8 keep their place.
#define BB_NXCD 8
__device__ __forceinline__ void xcd_tile(int& tile, int& b) {
    const int lin = blockIdx.y * gridDim.x + blockIdx.x;
    const int per = (gridDim.x * gridDim.y) / BB_NXCD;
    const int r = lin < per * BB_NXCD ? (lin % BB_NXCD) * per + lin / BB_NXCD : lin;
    b = r / gridDim.x;
    tile = r - b * gridDim.x;
}
