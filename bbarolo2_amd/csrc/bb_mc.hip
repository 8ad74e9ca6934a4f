// Monte-Carlo GALMOD on the GPU (the reference's _galmod_mc, bbarolo/galmod.py:92-171) with a counter-based
// random stream: every draw is Philox4x32-10 of (cloud, block) under the key (seed, sub-ring), so clouds and
// sub-clouds can be drawn in any order and the model of a box is the crop of the full model. The FP32
// functions below use +, -, *, fmaf and bit operations only, in the operation order of the CPU restatement, so
// the integer count cube is identical on both.
//
//   k_mc_prep    per (vector, sub-ring): cloud number, flux per sub-cloud, geometry (FP64 where the reference
//                rounds cloud numbers, FP32 for the clouds);
//   k_mc_scan    per vector: the fixed-point unit q and the clouds' offsets;
//   k_mc_clouds  per (vector, sub-ring), twice: the pixel of every cloud; pass 0 counts the clouds of every
//                pixel tile, pass 1 writes each cloud's record (tile slot, velocity, width, flux) into its tile's
//                bin (counting sort; the deposit does not depend on the order inside a bin);
//   k_mc_tile    per (vector, tile): the tile's spectra in LDS as 64-bit integers; each thread takes a
//                (cloud, block of 4 sub-clouds) item: one Philox call, two Box-Muller pairs, four channels,
//                four integer LDS atomics; then the tile is written once (count x q x cd2i).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/bbgpu.h"
#include "bb_device.h"

#define BB_MC_THREADS 512
#define BB_PI_MC 3.141592653589793

// ---- Philox4x32-10 ----------------------------------------------------------------------------------------
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// ---- FP32 functions (the same operations as the CPU restatement) -------------------------------------------
__device__ __forceinline__ float mc_u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 0x1p-23f; }

__device__ __forceinline__ float mc_rcp(float x) {
    float y = __uint_as_float(0x7EF311C3u - __float_as_uint(x));
#pragma unroll
    for (int i = 0; i < 4; i++) y = fmaf(y, fmaf(-x, y, 1.0f), y);
    return y;
}

__device__ __forceinline__ float mc_sqrt(float x) {
    if (!(x > 0.0f)) return 0.0f;
    float y = __uint_as_float(0x5F375A86u - (__float_as_uint(x) >> 1));
#pragma unroll
    for (int i = 0; i < 4; i++) y = y * fmaf(-0.5f * x * y, y, 1.5f);
    const float s = x * y;
    return fmaf(0.5f * y, fmaf(-s, s, x), s);
}

// log(x), x > 0 normal: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log(1 + y) = y P(y) with P the float32-rounded
// least-squares fit that mc_log_fit.py (beside this file) computes
__device__ __forceinline__ float mc_log(float x) {
    const uint32_t i = __float_as_uint(x);
    int e = (int)((i >> 23) & 0xFF) - 127;
    float m = __uint_as_float((i & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m *= 0.5f;
        e += 1;
    }
    const float y = m - 1.0f;
    float p = -7.638592273e-02f;
    p = fmaf(p, y, 1.291417032e-01f);
    p = fmaf(p, y, -1.324084252e-01f);
    p = fmaf(p, y, 1.418007910e-01f);
    p = fmaf(p, y, -1.660912633e-01f);
    p = fmaf(p, y, 2.000207156e-01f);
    p = fmaf(p, y, -2.500158548e-01f);
    p = fmaf(p, y, 3.333332539e-01f);
    p = fmaf(p, y, -4.999998808e-01f);
    p = fmaf(p, y, 1.0f);
    return fmaf((float)e, 0.693147182f, y * p);
}

// (cos, sin) of 2 pi t, t in [0, 1]: quadrant, then the Taylor series of the quarter turn
__device__ __forceinline__ void mc_sincos_turn(float t, float& c, float& s) {
    const float t4 = 4.0f * t;
    const float qf = floorf(t4);
    const float f = t4 - qf;
    const int q = ((int)qf) & 3;
    const float f2 = f * f;
    float ps = 5.69217278e-08f;
    ps = fmaf(ps, f2, -3.59884325e-06f);
    ps = fmaf(ps, f2, 1.60441181e-04f);
    ps = fmaf(ps, f2, -4.68175393e-03f);
    ps = fmaf(ps, f2, 7.96926245e-02f);
    ps = fmaf(ps, f2, -6.45964086e-01f);
    ps = fmaf(ps, f2, 1.57079637e+00f);
    const float sv = ps * f;
    float pc = -6.38660325e-09f;
    pc = fmaf(pc, f2, 4.71087475e-07f);
    pc = fmaf(pc, f2, -2.52020418e-05f);
    pc = fmaf(pc, f2, 9.19260259e-04f);
    pc = fmaf(pc, f2, -2.08634809e-02f);
    pc = fmaf(pc, f2, 2.53669500e-01f);
    pc = fmaf(pc, f2, -1.23370051e+00f);
    pc = fmaf(pc, f2, 1.0f);
    const float cv = pc;
    c = q == 0 ? cv : q == 1 ? -sv : q == 2 ? -cv : sv;
    s = q == 0 ? sv : q == 1 ? cv : q == 2 ? -sv : -cv;
}

__device__ __forceinline__ void mc_box_muller(uint32_t a, uint32_t b, float& n0, float& n1) {
    const float r = mc_sqrt(-2.0f * mc_log(mc_u01(a)));
    float c, s;
    mc_sincos_turn(mc_u01(b), c, s);
    n0 = r * c;
    n1 = r * s;
}

// vertical deviate of the layer (galmod.py:76-89 with these draws)
__device__ __forceinline__ float mc_layer_dev(int ltype, uint32_t a, uint32_t b) {
    if (ltype == 1) {
        float n0, n1;
        mc_box_muller(a, b, n0, n1);
        return n0;
    }
    const float x = 2.0f * mc_u01(a) - 1.0f;
    if (ltype == 2) return 0.5f * (mc_log(1.0f + x) - mc_log(1.0f - x));
    if (ltype == 3) return x >= 0.0f ? -mc_log(x) : mc_log(-x);
    if (ltype == 4) {
        float c, s;
        mc_sincos_turn(fabsf(x) * 0.25f, c, s);
        const float t = s * mc_rcp(c);
        return x >= 0.0f ? t : -t;
    }
    return x;
}

__device__ __forceinline__ int64_t mc_lround_f(float x) {
    const float a = fabsf(x);
    if (!(a < 1e9f)) return -((int64_t)1 << 62);
    const int64_t r = (int64_t)(a + 0.5f);
    return x >= 0.0f ? r : -r;
}

__device__ __forceinline__ int64_t mc_lround_d(double x) {
    if (!(fabs(x) < 1e18)) return -((int64_t)1 << 62);
    if (x >= 0) return (int64_t)(x + 0.5);
    return -(int64_t)(-x + 0.5);
}

__device__ __forceinline__ float mc_lag(float vrot, float dvdz, float dz) {
    const float v = vrot - dvdz * dz;
    if (vrot >= 0.0f) return v > 0.0f ? v : 0.0f;
    return v < 0.0f ? v : 0.0f;
}

__device__ __forceinline__ double mc_velgrid(double v, int axtyp, int velsys, double freq0, double crval3,
                                             double cdelt3, double crpix3) {
    const double C = 2.99792458e8;
    double w;
    if (axtyp == 4) {
        w = v;
    } else {
        if (velsys == 1) w = freq0 * (1 - v / C);
        else if (velsys == 2) w = freq0 * (C / (C + v));
        else w = freq0 * sqrt((C - v) / (C + v));
        if (axtyp == 2) w = C / w;
    }
    return (w - crval3) / cdelt3 + crpix3 - 1;
}

// ---- per sub-ring -----------------------------------------------------------------------------------------
struct McSub {
    int64_t nc;
    float fluxsc, r_in2, r_cross, sinc, cinc, spa, cpa, xpos, ypos, vsys, vrot, vrad, vvert, vdisp, z0, dvdz, zcyl,
        wrec, wapp;
    int nv;
};

__global__ void k_mc_prep(bb_grid g, const double* __restrict__ srall, int nsub, int nvpar, McSub* __restrict__ subs) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (s >= nsub) return;
    const double* sr = srall + (size_t)b * 16 * nsub;
    McSub q;
    q.nc = 0;
    q.fluxsc = q.r_in2 = q.r_cross = q.sinc = q.cinc = q.spa = q.cpa = q.xpos = q.ypos = 0.f;
    q.vsys = q.vrot = q.vrad = q.vvert = q.vdisp = q.z0 = q.dvdz = q.zcyl = q.wrec = q.wapp = 0.f;
    q.nv = 0;
    const double R0 = sr[s], dens = sr[5 * nsub + s];
    const int nv = sr[4 * nsub + s] == 0 ? 1 : nvpar;
    if (dens != 0) {
        const double ring_area_pix = 2 * BB_PI_MC * R0 * g.radsep / g.pixarea;
        const double dc = g.cmode == 1 ? dens : g.cmode == 2 ? dens * dens : pow(dens, (double)g.cmode);
        int64_t nc = mc_lround_d(g.cdens * dc * ring_area_pix);
        if ((double)mc_lround_d((double)nc / ring_area_pix) < 10.) nc = mc_lround_d(10 * ring_area_pix);
        if (nc > 0) {
            q.nc = nc;
            q.nv = nv;
            q.fluxsc = (float)(dens * ring_area_pix * g.pixarea / ((double)nc * (double)nv));
            const double r_inner = R0 - 0.5 * g.radsep;
            q.r_in2 = (float)(r_inner * r_inner);
            q.r_cross = (float)(2.0 * g.radsep * R0);
            const double TURN = 1.0 / (2 * BB_PI_MC);
            float c, sn;
            double ti = sr[9 * nsub + s] * TURN;
            ti -= floor(ti);
            mc_sincos_turn((float)ti, c, sn);
            q.sinc = sn;
            q.cinc = c;
            double tp = (sr[10 * nsub + s] + g.crota2) * TURN;
            tp -= floor(tp);
            mc_sincos_turn((float)tp, c, sn);
            q.spa = sn;
            q.cpa = c;
            q.xpos = (float)sr[11 * nsub + s];
            q.ypos = (float)sr[12 * nsub + s];
            q.vsys = (float)sr[13 * nsub + s];
            q.vrot = (float)sr[1 * nsub + s];
            q.vrad = (float)sr[2 * nsub + s];
            q.vvert = (float)sr[3 * nsub + s];
            q.vdisp = (float)sqrt(sr[4 * nsub + s] * sr[4 * nsub + s] + g.sig_instr * g.sig_instr);
            q.z0 = (float)sr[6 * nsub + s];
            q.dvdz = (float)sr[7 * nsub + s];
            q.zcyl = (float)sr[8 * nsub + s];
            q.wrec = (float)sr[14 * nsub + s];
            q.wapp = (float)sr[15 * nsub + s];
        }
    }
    subs[(size_t)b * nsub + s] = q;
}

// per vector: q (the unit of the counts) and 1/q, the cloud total; one block of 256 threads per vector
__global__ void k_mc_scan(bb_grid g, const McSub* __restrict__ subs, int nsub, double* __restrict__ qv,
                          float* __restrict__ invq, int64_t* __restrict__ total) {
    const int b = blockIdx.x;
    __shared__ double red[256];
    __shared__ long long cnt[256];
    double am = 0.0;
    long long n = 0;
    for (int s = threadIdx.x; s < nsub; s += blockDim.x) {
        const McSub& u = subs[(size_t)b * nsub + s];
        const double w = fabs((double)u.wrec) > fabs((double)u.wapp) ? fabs((double)u.wrec) : fabs((double)u.wapp);
        const double a = u.nc > 0 ? fabs((double)u.fluxsc) * w : 0.0;
        am = a > am ? a : am;
        n += u.nc;
    }
    red[threadIdx.x] = am;
    cnt[threadIdx.x] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        long long t = 0;
        for (int i = 0; i < (int)blockDim.x; i++) {
            m = red[i] > m ? red[i] : m;
            t += cnt[i];
        }
        double rmax = 0.0;
        for (int nl = 0; nl < g.nlines; nl++) {
            const double r = fabs((double)(float)g.relint[nl]);
            rmax = r > rmax ? r : rmax;
        }
        m *= rmax;
        const double q = m > 0 ? m * 0x1p-30 : 1.0;
        qv[b] = q;
        invq[b] = (float)(1.0 / q);
        total[b] = t;
    }
}

struct McRec {
    int slot, s;
    uint32_t ic;
    float aw, c0, sg, vs;
    int nv;
};

struct McGeo {
    int64_t blo0, blo1, bhi0, bhi1;
    int TX, TY, ntx;
    float icd0, icd1, ia, ib;
};

// the pixel of cloud ic of sub-ring u, and its record; false if it is outside the box or on a zero-weight half
__device__ __forceinline__ bool mc_cloud(const bb_grid& g, const McGeo& G, const int4& vb, const McSub& u, int s,
                                         uint32_t seed, uint32_t ic, int& tile, McRec& rec) {
    uint32_t x[4];
    philox(ic, 0u, 0u, 0u, seed, (uint32_t)s, x);
    const float R = mc_sqrt(fmaf(u.r_cross, fabsf(2.0f * mc_u01(x[0]) - 1.0f), u.r_in2));
    float caz, saz;
    mc_sincos_turn(fabsf(2.0f * mc_u01(x[1]) - 1.0f), caz, saz);
    const float z = mc_layer_dev(g.ltype, x[2], x[3]) * u.z0;
    const float xx = R * caz;
    const float yy = R * saz * u.cinc - z * u.sinc;
    const int64_t g0 = mc_lround_f(u.xpos + (xx * u.spa - yy * u.cpa) * G.icd0);
    const int64_t g1 = mc_lround_f(u.ypos + (xx * u.cpa + yy * u.spa) * G.icd1);
    const float wside = caz >= 0.0f ? u.wrec : u.wapp;
    if (wside == 0.0f || g0 < vb.x || g0 >= vb.z || g1 < vb.y || g1 >= vb.w) return false;
    const float az_ = fabsf(z);
    float vs;
    if (az_ > u.zcyl) vs = u.vsys + (mc_lag(u.vrot, u.dvdz, az_ - u.zcyl) * caz + u.vrad * saz) * u.sinc;
    else vs = u.vsys + (u.vrot * caz + u.vrad * saz) * u.sinc;
    vs += u.vvert * u.cinc;
    const int px = (int)(g0 - G.blo0), py = (int)(g1 - G.blo1);
    tile = (py / G.TY) * G.ntx + px / G.TX;
    rec.slot = (py % G.TY) * G.TX + px % G.TX;
    rec.s = s;
    rec.ic = ic;
    const bool linear = g.axtyp == 4 && g.nlines == 1;
    rec.aw = u.fluxsc * wside;
    rec.c0 = vs * G.ia + G.ib;
    rec.sg = linear ? u.vdisp * G.ia : u.vdisp;
    rec.vs = vs;
    rec.nv = u.nv;
    return true;
}

// pass 0: tcount[b][tile] += clouds; pass 1: recs[b][toff[b][tile] + cursor] = record (toff: ntiles + 1 per vector)
__global__ void __launch_bounds__(256)
k_mc_clouds(bb_grid g, McGeo G, const McSub* __restrict__ subs, int nsub, uint32_t seed, int ntiles, int pass,
            int* __restrict__ tcount, const int64_t* __restrict__ toff, int* __restrict__ tcursor,
            McRec* __restrict__ recs, int64_t cap, int* __restrict__ overflow, const int4* __restrict__ vbox) {
    const int s = blockIdx.x, b = blockIdx.y;
    const McSub u = subs[(size_t)b * nsub + s];
    // the box of the clouds kept: the layout box, or the vector's own model box within it
    int4 vb = make_int4((int)G.blo0, (int)G.blo1, (int)G.bhi0, (int)G.bhi1);
    if (vbox) {
        const int4 v = vbox[b];
        vb.x = max(vb.x, v.x);
        vb.y = max(vb.y, v.y);
        vb.z = min(vb.z, v.z);
        vb.w = min(vb.w, v.w);
    }
    for (int64_t ic = threadIdx.x; ic < u.nc; ic += blockDim.x) {
        int tile;
        McRec rec;
        if (!mc_cloud(g, G, vb, u, s, seed, (uint32_t)ic, tile, rec)) continue;
        if (pass == 0) {
            atomicAdd(tcount + (size_t)b * ntiles + tile, 1);
        } else {
            const int64_t pos = toff[(size_t)b * (ntiles + 1) + tile] + atomicAdd(tcursor + (size_t)b * ntiles + tile, 1);
            if (pos < cap) recs[(size_t)b * cap + pos] = rec;
            else atomicOr(overflow, 1);
        }
    }
}

// exclusive scan of the tile counts of each vector (one block per vector), cursors reset
__global__ void k_mc_tscan(const int* __restrict__ tcount, int ntiles, int64_t* __restrict__ toff,
                           int* __restrict__ tcursor) {
    const int b = blockIdx.x;
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        for (int t = 0; t < ntiles; t++) {
            toff[(size_t)b * (ntiles + 1) + t] = acc;
            acc += tcount[(size_t)b * ntiles + t];
        }
        toff[(size_t)b * (ntiles + 1) + ntiles] = acc;
    }
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) tcursor[(size_t)b * ntiles + t] = 0;
}

__global__ void __launch_bounds__(BB_MC_THREADS)
k_mc_tile(bb_grid g, McGeo G, int bx, int by, uint32_t seed, int ntiles, const int64_t* __restrict__ toff,
          const McRec* __restrict__ recs, int64_t cap, const double* __restrict__ qv, const float* __restrict__ invq,
          const double* __restrict__ cd2i, int nvmax, float* __restrict__ outall,
          unsigned long long* __restrict__ countsall, int accumulate) {
    extern __shared__ unsigned long long mtile[];
    int b, t;
    xcd_tile(t, b);
    const int nz = g.nz, nzp = nz | 1;
    const int TX = G.TX, TY = G.TY;
    const int tx0 = (t % G.ntx) * TX, ty0 = (t / G.ntx) * TY;
    for (int i = threadIdx.x; i < TX * TY * nzp; i += blockDim.x) mtile[i] = 0ull;
    __syncthreads();
    const int64_t r0 = toff[(size_t)b * (ntiles + 1) + t], r1 = toff[(size_t)b * (ntiles + 1) + t + 1];
    const int64_t nrec = (r1 < cap ? r1 : cap) - (r0 < cap ? r0 : cap);
    const McRec* rec = recs + (size_t)b * cap + r0;
    const int nblk = (nvmax + 3) / 4;
    const float iq = invq[b];
    const bool linear = g.axtyp == 4 && g.nlines == 1;
    float relf[BB_MAX_LINES];
#pragma unroll
    for (int nl = 0; nl < BB_MAX_LINES; nl++) relf[nl] = (float)g.relint[nl];
    for (int64_t it = threadIdx.x; it < nrec * nblk; it += blockDim.x) {
        const int64_t ir = it / nblk;
        const int j = (int)(it - ir * nblk);
        const McRec r = rec[ir];
        if (4 * j >= r.nv) continue;
        uint32_t y[4];
        philox(r.ic, 1u + (uint32_t)j, 0u, 0u, seed, (uint32_t)r.s, y);
        float n[4];
        mc_box_muller(y[0], y[1], n[0], n[1]);
        mc_box_muller(y[2], y[3], n[2], n[3]);
        unsigned long long* tp = mtile + (size_t)r.slot * nzp;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (4 * j + i >= r.nv) break;
            if (linear) {
                const int64_t k = mc_lround_f(fmaf(n[i], r.sg, r.c0));
                if (k < 0 || k >= nz) continue;
                atomicAdd(tp + k, (unsigned long long)(uint32_t)fmaf(relf[0] * r.aw, iq, 0.5f));
            } else {
                const double v = (double)fmaf(n[i], r.sg, r.vs);
                for (int nl = 0; nl < g.nlines; nl++) {
                    const int64_t k = mc_lround_d(mc_velgrid(v, g.axtyp, g.velsys, g.lfreq0[nl], g.crval3, g.cdelt3,
                                                             g.crpix3));
                    if (k < 0 || k >= nz) continue;
                    atomicAdd(tp + k, (unsigned long long)(uint32_t)fmaf(relf[nl] * r.aw, iq, 0.5f));
                }
            }
        }
    }
    __syncthreads();
    const double q = qv[b];
    float* out = outall + (size_t)b * by * bx * nz;
    for (int i = threadIdx.x; i < TX * TY * nz; i += blockDim.x) {
        const int p = i / nz, k = i - p * nz;
        const int rx = tx0 + p % TX, ry = ty0 + p / TX;
        if (rx < bx && ry < by) {
            const unsigned long long c = mtile[(size_t)p * nzp + k];
            const size_t o = ((size_t)ry * bx + rx) * nz + k;
            const float v = (float)((double)c * q * cd2i[k]);
            out[o] = accumulate ? out[o] + v : v;
            if (countsall) countsall[(size_t)b * by * bx * nz + o] = c;
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// zero n 32-bit words (a kernel, not a memset: the chain stays capturable into a HIP graph)
__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0u;
}

extern "C" int bb_launch_zero_u32(void* p, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, n);
    return (int)hipGetLastError();
}

extern "C" size_t bb_mc_sub_bytes() { return sizeof(McSub); }
extern "C" size_t bb_mc_rec_bytes() { return sizeof(McRec); }
extern "C" size_t bb_mc_lds_bytes(int nz, int TX, int TY) { return (size_t)TX * TY * (nz | 1) * sizeof(unsigned long long); }

extern "C" int bb_launch_galmod_mc(const bb_grid* g, bb_box box, const double* sr, int nsub, int B, int nvpar,
                                   uint32_t seed, int TX, int TY, int64_t cap, void* subs, double* qv, float* invq,
                                   int64_t* total, int* tcount, int64_t* toff, int* tcursor, void* recs, int* overflow,
                                   const double* cd2i, float* out, unsigned long long* counts, int accumulate,
                                   const int4* vbox, hipStream_t st) {
    const int bx = box.bhi0 - box.blo0, by = box.bhi1 - box.blo1;
    McGeo G;
    G.blo0 = box.blo0;
    G.blo1 = box.blo1;
    G.bhi0 = box.bhi0;
    G.bhi1 = box.bhi1;
    G.TX = TX;
    G.TY = TY;
    G.ntx = (bx + TX - 1) / TX;
    G.icd0 = (float)(1.0 / g->cdelt0);
    G.icd1 = (float)(1.0 / g->cdelt1);
    G.ia = (float)(1.0 / g->cdelt3);
    G.ib = (float)(-g->crval3 / g->cdelt3 + g->crpix3 - 1);
    const int ntiles = G.ntx * ((by + TY - 1) / TY);
    McSub* S = (McSub*)subs;
    McRec* R = (McRec*)recs;
    hipLaunchKernelGGL(k_mc_prep, dim3((nsub + 63) / 64, B), dim3(64), 0, st, *g, sr, nsub, nvpar, S);
    hipLaunchKernelGGL(k_mc_scan, dim3(B), dim3(256), 0, st, *g, S, nsub, qv, invq, total);
    hipLaunchKernelGGL(k_mc_clouds, dim3(nsub, B), dim3(256), 0, st, *g, G, S, nsub, seed, ntiles, 0, tcount,
                       (const int64_t*)nullptr, tcursor, R, cap, overflow, vbox);
    hipLaunchKernelGGL(k_mc_tscan, dim3(B), dim3(256), 0, st, tcount, ntiles, toff, tcursor);
    hipLaunchKernelGGL(k_mc_clouds, dim3(nsub, B), dim3(256), 0, st, *g, G, S, nsub, seed, ntiles, 1, tcount, toff,
                       tcursor, R, cap, overflow, vbox);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_mc_tile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024);
        if (e != hipSuccess) return (int)e;
        attr = true;
    }
    hipLaunchKernelGGL(k_mc_tile, dim3(ntiles, B), dim3(BB_MC_THREADS), bb_mc_lds_bytes(g->nz, TX, TY), st, *g, G, bx,
                       by, seed, ntiles, toff, R, cap, qv, invq, (const double*)cd2i, nvpar, out, counts, accumulate);
    return (int)hipGetLastError();
}
