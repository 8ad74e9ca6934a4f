// The fused likelihood of a fit: Bayesian3DFIT._log_likelihood (bbarolo/bayesian/likelihood.py:567-619) for batches
// of parameter vectors, run entirely by the library once the per-fit context exists:
//
//   theta (host) -> validity and model boxes (host, a few numbers per vector)
//                -> sub-ring tables on the device (k_like_tables: _update_rings + GalmodSetup.subrings,
//                   bayesian/fitter.py:1109-1136, galmod.py:761-851, from the plan of the fixed radii)
//                -> ring maps -> non-negative profile -> densities into the tables (when DENS is the deconvolved
//                   profile of every vector, bayesian/density.py:13-88)
//                -> ANALYTIC or MC deposit in each vector's model box -> beam convolution fused with the masked
//                   residual -> the squared data outside the box (summed-area table) -> log L (host).
//
// A launch of up to `chunk` vectors is asynchronous (bb_loglike_launch); bb_loglike_collect waits for it and forms
// the log-likelihoods. The halves of a launch run on two private streams (the least squares of one half, latency
// bound, beside the ring maps of the other). Device memory is allocated once, at bb_ctx_create.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/bbgpu.h"
#include "bb_internal.h"

namespace {

constexpr double DEG2RAD = M_PI / 180.0;     // numpy.radians multiplies by this constant

struct PlanDev {
    int nmr, nsub, thick, nsub_t;
    unsigned from_fit;
    const int* ridx;
    const double* rad;
    const double* base;       // (BB_NKEY, nmr)
    const int* idx;
    const double* dr;
    const unsigned char* cst;
    const double* rads;
    const double* densw;      // may be null
    const double* side;       // (2, nsub_t)
};

struct FitDev {
    int nr, ndim;
    int off[BB_NKEY];
    int ring[BB_NKEY];
    const double* rings;      // (BB_NKEY, nr): the float32 ring table
};

__device__ __forceinline__ double fit_value(const FitDev& F, const double* th, int k, int i) {
    if (F.off[k] >= 0) return (double)(float)th[F.off[k] + (F.ring[k] ? i : 0)];
    return F.rings[(size_t)k * F.nr + i];
}

__device__ __forceinline__ double key_value(const PlanDev& P, const FitDev& F, const double* th, int k, int m) {
    if ((P.from_fit >> k) & 1u) return fit_value(F, th, k, P.ridx[m]);
    return P.base[(size_t)k * P.nmr + m];
}

// row r (1..16) of the sub-ring table at model ring m, in the table's units (m/s, arcmin, radians), as
// SubringPlan.tables and ring_values form them
__device__ double row_value(const PlanDev& P, const FitDev& F, const double* th, int r, int m) {
    auto kv = [&](int k) { return key_value(P, F, th, k, m); };
    switch (r) {
        case 1: return kv(BB_K_VROT) * 1000.;
        case 2: return kv(BB_K_VRAD) * 1000.;
        case 3: return kv(BB_K_VVERT) * 1000.;
        case 4: return kv(BB_K_VDISP) * 1000.;
        case 5: return kv(BB_K_DENS);
        case 6: return kv(BB_K_Z0) / 60.;
        case 7: return kv(BB_K_DVDZ) * 1000. * 60.;
        case 8: return kv(BB_K_ZCYL) / 60.;
        case 9: return kv(BB_K_INC) * DEG2RAD;
        case 10: return kv(BB_K_PHI) * DEG2RAD;
        case 11: return kv(BB_K_XPOS);
        case 12: return kv(BB_K_YPOS);
        case 13: return kv(BB_K_VSYS) * 1000.;
        case 14: return kv(BB_K_FTHICK);
        case 15: {
            double z = kv(BB_K_Z0THICK);
            if (!(z >= 0)) z = kv(BB_K_Z0);          // (negative: tied to Z0)
            return z / 60.;
        }
        default: {
            double v = kv(BB_K_VDISPTHICK);
            if (!(v >= 0)) v = kv(BB_K_VDISP);
            return v * 1000.;
        }
    }
}

// The edges of the disk that theta moves (_update_rings, bayesian/fitter.py:1112-1120): RADMAX the edge of the disk (a
// sharp edge, or a Gaussian one of width `width`), RADMAXA / RADMAXR those of the approaching / receding halves (sharp;
// negative: none, galmod.py:853-863). ew (B, nsub_t), when given, receives the density edge weight of every sub-ring
// (the ring maps and the profile's densities take it, as the reference's ring weights hold it, galmod.py:1285-1314).
struct EdgeDev {
    int oe, oa, orr;          // offsets in theta (-1: not free)
    double width;             // arcsec
    double dr;                // arcsec: the sub-ring width (the setup's radsep)
    double* ew;
};

// edge_weights (galmod.py:1122-1131): the fraction of the sub-ring [R - dr/2, R + dr/2] inside the edge, or the
// Gaussian edge's 1/2 erfc((R - edge) / (sqrt(2) width))
__device__ __forceinline__ double edge_weight(double R, double dr, double edge, double width) {
    if (width > 0) return 0.5 * erfc((R - edge) / (sqrt(2.) * width));
    return fmin(fmax((edge - (R - dr / 2.)) / dr, 0.), 1.);
}

// linear interpolation of row r at sub-ring s between the model rings of its segment (interp_values)
__device__ double interp_row(const PlanDev& P, const FitDev& F, const double* th, int r, int s) {
    if (P.cst[s]) return row_value(P, F, th, r, 0);
    if (P.nmr > 1) {
        const int i1 = P.idx[s] > 1 ? P.idx[s] : 1;
        const double dur = P.rad[i1] - P.rad[i1 - 1];
        const double lo = row_value(P, F, th, r, i1 - 1);
        const double slope = (row_value(P, F, th, r, i1) - lo) / dur;
        return lo + slope * P.dr[s];
    }
    return row_value(P, F, th, r, 0) * 1.0;
}

// One thread per (sub-ring, vector): the 16 rows of the table (B, 16, nsub_t) of GalmodSetup.subrings
// (galmod.py:761-851: the edge weights of the density, the thin and thick layers, the side weights, galmod.py:833-851).
__global__ void __launch_bounds__(64)
k_like_tables(PlanDev P, FitDev F, EdgeDev E, const double* __restrict__ theta, double* __restrict__ srall) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (s >= P.nsub) return;
    const double* th = theta + (size_t)b * F.ndim;
    double* sr = srall + (size_t)b * 16 * P.nsub_t;
    double v[17];
    v[0] = P.rads[s];
    for (int r = 1; r <= 13; r++) v[r] = interp_row(P, F, th, r, s);
    double dens = v[5];
    const double Ras = v[0] * 60.;
    if (E.oe >= 0) {
        const double w = edge_weight(Ras, E.dr, th[E.oe], E.width);
        dens = dens * w;
        if (E.ew) {
            E.ew[(size_t)b * P.nsub_t + s] = w;
            if (P.thick) E.ew[(size_t)b * P.nsub_t + P.nsub + s] = w;
        }
    } else if (P.densw) {
        dens = dens * P.densw[s];
    }
    // the side weights: the plan's, or those of the vector's own edges (both layers at the sub-ring's radius)
    double wrec = P.side[s], wapp = P.side[P.nsub_t + s];
    double wrec_t = P.thick ? P.side[P.nsub + s] : 0., wapp_t = P.thick ? P.side[P.nsub_t + P.nsub + s] : 0.;
    if (E.orr >= 0) {
        const double e = th[E.orr];
        wrec = wrec_t = e >= 0 ? edge_weight(Ras, E.dr, e, 0.) : 1.0;
    }
    if (E.oa >= 0) {
        const double e = th[E.oa];
        wapp = wapp_t = e >= 0 ? edge_weight(Ras, E.dr, e, 0.) : 1.0;
    }
    auto put = [&](int row, int col, double x) { sr[(size_t)row * P.nsub_t + col] = x; };
    if (!P.thick) {
        for (int r = 0; r <= 13; r++) put(r, s, r == 5 ? dens : v[r]);
        put(14, s, wrec);
        put(15, s, wapp);
        return;
    }
    const double f = interp_row(P, F, th, 14, s);
    const double z0t = interp_row(P, F, th, 15, s);
    const double vdt = interp_row(P, F, th, 16, s);
    const int t = P.nsub + s;
    for (int r = 0; r <= 13; r++) {
        double thin = v[r], thk = v[r];
        if (r == 5) {
            thin = dens * (1. - f);
            thk = dens * f;
        } else if (r == 7) {
            thin = 0.;
        } else if (r == 6) {
            thk = z0t;
        } else if (r == 4) {
            thk = vdt;
        }
        put(r, s, thin);
        put(r, t, thk);
    }
    put(14, s, wrec);
    put(15, s, wapp);
    put(14, t, wrec_t);
    put(15, t, wapp_t);
}

// The exterior annulus of the profile when an edge of the disk is free (profile.py:45-64, galmod.py:761-832): per
// vector its radius rext = min(outer edge, RADMAX, max(RADMAXA, RADMAXR) when both halves have one) + width / 2, the
// model rings of one CENTRED ring that wide (model_rings, galmod.py:1066-1119), the float32-accumulated sub-ring radii
// (capacity cap per layer; past the last, zero side weights), every sub-ring with the values of the last ring (as the
// annulus plan gives them: its model rings are equal, so the interpolation gives their values exactly), and the rows
// of the curvature prior on the radii of the rings and of the annulus (prior_rows, profile.py:81-150), which move with
// rext. One workgroup per vector; lane 0 does the sequential parts.
struct AnnDev {
    int oe, oa, orr;          // theta offsets of RADMAX, RADMAXA, RADMAXR (-1: not free)
    double edge0, ea0, er0;   // the fixed edges (edge0 NaN: none; ea0, er0 negative: none)
    double oedge, w;          // arcsec: the outer edge of the rings, the width of the annulus
    double radsep;            // arcmin: the sub-ring width
    int empty, cap;           // EMPTY, sub-rings per layer
    int nr;                   // rings (the prior has n = nr + 1 columns)
    const double* radii;      // (nr) arcsec
    double dmin, dmed;        // a quarter of the smallest ring spacing, the median spacing
    double* lcoef;            // (B, n - 2, 3)
    int* lcol;                // (B, n - 2, 3)
    int* near;                // (B, n - 2)
    double* norm;             // (B)
    int* overflow;            // set when a vector's annulus has more sub-rings than cap
};

#define BB_ANN_CAP 2048

__global__ void __launch_bounds__(64)
k_ann_tables(PlanDev P, FitDev F, AnnDev A, const double* __restrict__ theta, double* __restrict__ srall) {
#pragma clang fp contract(off)
    __shared__ double rads[BB_ANN_CAP];
    __shared__ int nsh;
    __shared__ double rows[17];
    const int b = blockIdx.x, lane = threadIdx.x;
    const double* th = theta + (size_t)b * F.ndim;
    const int nst = A.cap * (P.thick ? 2 : 1);
    double* sr = srall + (size_t)b * 16 * nst;
    if (lane == 0) {
        double edge = A.oedge;
        const double e = A.oe >= 0 ? th[A.oe] : A.edge0;
        if (e == e) edge = fmin(edge, e);
        const double ea = A.oa >= 0 ? th[A.oa] : A.ea0, er = A.orr >= 0 ? th[A.orr] : A.er0;
        if (ea >= 0 && er >= 0) edge = fmin(edge, fmax(ea, er));
        const double rext = edge + A.w / 2.;
        // model_rings of one CENTRED ring of width w
        double rad[3];
        int nm = 0;
        const double rin = fmax(rext - A.w / 2., 0.), rout = rext + A.w / 2.;
        if (rin / 60. + A.radsep < rext / 60.) rad[nm++] = rin / 60.;
        rad[nm++] = rext / 60.;
        if (rext / 60. + A.radsep < rout / 60.) rad[nm++] = rout / 60.;
        // subring_radii: float32 accumulation
        int n = 0;
        const bool inner = !(rad[0] != 0 && A.empty);
        float r = inner ? (float)(A.radsep / 2.0) : (float)(rad[0] + A.radsep / 2.0);
        if (inner)
            for (const float r0 = (float)rad[0]; r < r0; r = (float)((double)r + A.radsep)) {
                if (n < A.cap) rads[n] = (double)r;
                if (++n > A.cap) break;
            }
        for (int i = 1; i < nm && n <= A.cap; i++)
            for (const float ri = (float)rad[i]; r < ri; r = (float)((double)r + A.radsep)) {
                if (n < A.cap) rads[n] = (double)r;
                if (++n > A.cap) break;
            }
        if (n > A.cap) {
            *A.overflow = 1;
            n = A.cap;
        }
        nsh = n;
        for (int k = 1; k <= 13; k++) rows[k] = row_value(P, F, th, k, 0);
        if (P.thick)
            for (int k = 14; k <= 16; k++) rows[k] = row_value(P, F, th, k, 0);
        // prior_rows on the radii of the rings and rext, sorted (stable: rext after the radii equal to it)
        const int nr = A.nr, nc = nr + 1, nrow = nc > 2 ? nc - 2 : 0;
        int p = 0;
        while (p < nr && A.radii[p] <= rext) p++;
        auto Rs = [&](int i) { return i < p ? A.radii[i] : i == p ? rext : A.radii[i - 1]; };
        auto ord = [&](int i) { return i < p ? i : i == p ? nr : i - 1; };
        const double ptp = fmax(A.radii[nr - 1], rext) - fmin(A.radii[0], rext);
        const double tiny = 1e-6 * fmax(ptp, 1e-300);
        const double dmin = nc > 2 ? fmax(A.dmin, tiny) : tiny;
        double* lc = A.lcoef + (size_t)b * 3 * nrow;
        int* lk = A.lcol + (size_t)b * 3 * nrow;
        int* nearb = A.near + (size_t)b * nrow;
        double norm = 0.;
        for (int i = 0; i < nrow; i++) {
            const double ra = Rs(i + 1) - Rs(i), rb = Rs(i + 2) - Rs(i + 1);
            const double a = fmax(ra, tiny), bb = fmax(rb, tiny);
            const double c = 2. / (a + bb);
            lc[3 * i] = c * (1. / a);
            lc[3 * i + 1] = c * (-(1. / a + 1. / bb));
            lc[3 * i + 2] = c * (1. / bb);
            lk[3 * i] = ord(i);
            lk[3 * i + 1] = ord(i + 1);
            lk[3 * i + 2] = ord(i + 2);
            const double a2 = fmax(a, dmin), b2 = fmax(bb, dmin);
            const double c2 = 2. / (a2 + b2);
            const double m0 = c2 * (1. / a2), m1 = c2 * (-(1. / a2 + 1. / b2)), m2 = c2 * (1. / b2);
            norm += m0 * m0 + m1 * m1 + m2 * m2;
            nearb[i] = fmin(ra, rb) < 0.25 * A.dmed ? 1 : 0;
        }
        A.norm[b] = norm;
    }
    __syncthreads();
    const int n = nsh;
    auto put = [&](int row, int col, double x) { sr[(size_t)row * nst + col] = x; };
    for (int k = lane; k < A.cap; k += 64) {
        const bool on = k < n;
        const double R0 = on ? rads[k] : 0.0;
        const double dens = rows[5];
        if (!P.thick) {
            put(0, k, R0);
            for (int r = 1; r <= 13; r++) put(r, k, rows[r]);
            put(14, k, on ? 1.0 : 0.0);
            put(15, k, on ? 1.0 : 0.0);
            continue;
        }
        const int t = A.cap + k;
        const double f = rows[14];
        for (int r = 0; r <= 13; r++) {
            double thin = r == 0 ? R0 : rows[r], thk = thin;
            if (r == 5) {
                thin = dens * (1. - f);
                thk = dens * f;
            } else if (r == 7) {
                thin = 0.;
            } else if (r == 6) {
                thk = rows[15];
            } else if (r == 4) {
                thk = rows[16];
            }
            put(r, k, thin);
            put(r, t, thk);
        }
        put(14, k, on ? 1.0 : 0.0);
        put(15, k, on ? 1.0 : 0.0);
        put(14, t, on ? 1.0 : 0.0);
        put(15, t, on ? 1.0 : 0.0);
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
struct bb_ctx {
    int device;
    const bb_model* m;
    bb_like_desc d;
    int nz, nsub_t, nsub_e;
    bb_box db;                 // the deposit box (the layout of the models and of dvalid)
    int by, bx;
    int64_t vox;               // by bx nz
    FitDev fit;
    PlanDev plan, ann;
    EdgeDev edge;                // the model's edges moved by theta (offsets -1: fixed, in the plan)
    bool ann_free = false;       // an edge is free and the profile is computed: the annulus of every vector
    AnnDev annd;
    double w_thin = 0., w_thick = 0.;   // the annulus's ring weights of its thin and thick sub-rings
    std::vector<void*> dev_allocs;
    // host copies for the validity and boxes
    std::vector<double> rings;   // (BB_NKEY, nr)
    int off[BB_NKEY], ring[BB_NKEY];
    bool thick;
    std::vector<double> sat;     // (ny + 1) (nx + 1)
    double sat_total, nvox, ss_out_fix;
    // device objects
    bb_azplan* pm = nullptr;
    bb_azplan* pe = nullptr;
    bb_profile* prof = nullptr;
    int nprof;                   // columns of the profile (nr + 1), 0 without
    const int32_t *widx = nullptr, *widx_e = nullptr;
    const double *wval = nullptr, *wval_e = nullptr;
    double qunit, unit;
    float* dvalid = nullptr;
    // per-launch device buffers (chunk vectors)
    char* stage_d = nullptr;
    char* stage_h = nullptr;     // pinned
    size_t stage_bytes = 0;
    double* sr = nullptr;
    double* sr_e = nullptr;
    double* e_sub = nullptr;     // the annulus's sub-ring records (an edge free)
    double* e_wt = nullptr;
    float* dep = nullptr;
    uint64_t* counts = nullptr;
    double* maps = nullptr;
    double* x = nullptr;
    float* dens = nullptr;
    double* cwork = nullptr;
    int64_t cper = 0;            // partial doubles per vector
    double* ss = nullptr;
    double* lg = nullptr;
    int* mcflag = nullptr;
    void* awork[2] = {nullptr, nullptr};
    int64_t awork_b[2] = {0, 0};
    void* pwork[2] = {nullptr, nullptr};
    int64_t pwork_b[2] = {0, 0};
    void* mwork[2] = {nullptr, nullptr};
    int64_t mwork_b[2] = {0, 0};
    int64_t mc_cap = 0;
    char* out_h = nullptr;       // pinned: ss, lg (chunk each), dens (chunk nr floats), mc flags (2 ints)
    hipStream_t s[2] = {nullptr, nullptr};
    hipEvent_t ev_tables = nullptr, ev_maps = nullptr, ev_part[2] = {nullptr, nullptr}, ev_done = nullptr;
    // the launch in flight
    bool pending = false;
    int B = 0, nv = 0, nparts = 1;
    bool frac = false;
    std::vector<int> valid;      // indices of the valid vectors
    std::vector<double> s2, e2, ssout;
    int64_t nnls_fail = 0, mc_retries = 0;
    hipStream_t last_stream = nullptr;
    // launches as captured HIP graphs, one per batch size (and tile, list capacity, MC capacity, halves)
    struct GraphEntry {
        int nv, TX, TY, bin, halves;
        int64_t cap;
        hipGraphExec_t exec;
    };
    int use_graphs = 0;
    bool capturing = false;
    hipStream_t scap = nullptr;
    std::vector<GraphEntry> gx;
    std::vector<int> seen;
    int64_t graph_replays = 0;
    hipEvent_t trace[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

namespace {

template <class T>
int dalloc(bb_ctx* c, T** p, size_t n, const char* what) {
    *p = nullptr;
    if (n == 0) n = 1;
    int rc = bb_hip_check(hipMalloc((void**)p, n * sizeof(T)), what);
    if (!rc) c->dev_allocs.push_back((void*)*p);
    return rc;
}

template <class T>
int dupload(bb_ctx* c, const T** dst, const T* src, size_t n, const char* what) {
    T* p = nullptr;
    int rc = dalloc(c, &p, n, what);
    if (!rc && n) rc = bb_hip_check(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice), what);
    *dst = p;
    return rc;
}

bool positive_only(int k) {
    switch (k) {
        case BB_K_VROT: case BB_K_VDISP: case BB_K_INC: case BB_K_XPOS: case BB_K_YPOS: case BB_K_Z0: case BB_K_DENS:
        case BB_K_ZCYL: case BB_K_Z0THICK: case BB_K_FTHICK: case BB_K_VDISPTHICK:
            return true;
        default:
            return false;
    }
}

// the float32 value of ring key k at ring i of vector th (host)
double host_value(const bb_ctx* c, const double* th, int k, int i) {
    if (c->off[k] >= 0) return (double)(float)th[c->off[k] + (c->ring[k] ? i : 0)];
    return c->rings[(size_t)k * c->d.nr + i];
}

int copy_plan(bb_ctx* c, const bb_ring_plan* p, PlanDev* P, const char* what) {
    if (!p || p->nmr < 1 || p->nsub < 0 || !p->ridx || !p->rad || !p->base || (p->nsub > 0 && (!p->idx || !p->dr ||
        !p->cst || !p->rads || !p->side)))
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: incomplete %s plan", what);
    for (int mi = 0; mi < p->nmr; mi++)
        if (p->ridx[mi] < 0 || p->ridx[mi] >= c->d.nr)
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: %s plan ring index %d out of range", what, p->ridx[mi]);
    for (int s = 0; s < p->nsub; s++)
        if (!p->cst[s] && (p->idx[s] < 0 || p->idx[s] >= p->nmr))
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: %s plan segment %d out of range", what, p->idx[s]);
    P->nmr = p->nmr;
    P->nsub = p->nsub;
    P->thick = p->thick ? 1 : 0;
    P->nsub_t = p->nsub * (P->thick ? 2 : 1);
    P->from_fit = p->from_fit;
    int rc = dupload(c, &P->ridx, (const int*)p->ridx, p->nmr, what);
    if (!rc) rc = dupload(c, &P->rad, p->rad, p->nmr, what);
    if (!rc) rc = dupload(c, &P->base, p->base, (size_t)BB_NKEY * p->nmr, what);
    if (!rc) rc = dupload(c, &P->idx, (const int*)p->idx, p->nsub, what);
    if (!rc) rc = dupload(c, &P->dr, p->dr, p->nsub, what);
    if (!rc) rc = dupload(c, &P->cst, (const unsigned char*)p->cst, p->nsub, what);
    if (!rc) rc = dupload(c, &P->rads, p->rads, p->nsub, what);
    P->densw = nullptr;
    if (!rc && p->densw) rc = dupload(c, &P->densw, p->densw, p->nsub, what);
    if (!rc) rc = dupload(c, &P->side, p->side, (size_t)2 * P->nsub_t, what);
    return rc;
}

// the tables of the model (with its edges moved by theta) or of the fixed exterior annulus (no edges)
int launch_tables(bb_ctx* c, const PlanDev& P, bool model, const double* theta, int B, double* out, hipStream_t st) {
    if (P.nsub == 0 || B == 0) return 0;
    EdgeDev E = {-1, -1, -1, 0., 0., nullptr};
    if (model) E = c->edge;
    hipLaunchKernelGGL(k_like_tables, dim3((P.nsub + 63) / 64, B), dim3(64), 0, st, P, c->fit, E, theta, out);
    return bb_hip_check(hipGetLastError(), "k_like_tables");
}

// the exterior annulus of every vector (an edge free): its tables and the rows of the prior
int launch_ann(bb_ctx* c, const double* theta, int B, hipStream_t st) {
    if (B == 0) return 0;
    hipLaunchKernelGGL(k_ann_tables, dim3(B), dim3(64), 0, st, c->ann, c->fit, c->annd, theta, c->sr_e);
    return bb_hip_check(hipGetLastError(), "k_ann_tables");
}

// the stage layout: theta (chunk ndim doubles), vbox_abs, vbox_rel (chunk int4 each), sv (chunk float2)
struct Stage {
    double* th;
    int4* va;
    int4* vr;
    float2* sv;
};

Stage stage_of(const bb_ctx* c, char* base) {
    const size_t C = (size_t)c->d.chunk;
    Stage S;
    S.th = (double*)base;
    S.va = (int4*)(base + C * c->d.ndim * sizeof(double));
    S.vr = S.va + C;
    S.sv = (float2*)(S.vr + C);
    return S;
}

// one part (vectors [a, a + n)) of a launch on stream st
int run_part(bb_ctx* c, int p, int a, int n, const Stage& D, hipStream_t st, hipEvent_t after_maps) {
    const bb_model* m = c->m;
    const int64_t P = (int64_t)(c->d.likebox.bhi0 - c->d.likebox.blo0) * (c->d.likebox.bhi1 - c->d.likebox.blo1);
    int rc = 0;
    double* sr = c->sr + (size_t)a * 16 * c->nsub_t;
    const bool tr = p == 0 && c->trace[0] && !c->capturing;
    auto mark = [&](int k) { return tr ? bb_hip_check(hipEventRecord(c->trace[k], st), "trace") : 0; };
    rc = mark(0);
    if (rc) return rc;
    if (c->prof) {
        const int nm = c->nprof;
        uint64_t* counts = c->counts + (size_t)a * nm * P;
        double* maps = c->maps + (size_t)a * nm * P;
        const double* ew = c->edge.ew ? c->edge.ew + (size_t)a * c->nsub_t : nullptr;
        rc = bb_ring_maps_deposit_ew(m, c->pm, c->d.likebox, sr, c->nsub_t, n, c->widx, c->wval, ew, nm, 0, c->d.nr,
                                     c->qunit, c->awork[p], c->awork_b[p], counts, st);
        double* sre = c->sr_e + (size_t)a * 16 * c->nsub_e;
        if (!rc && c->ann_free) {
            const size_t nse = (size_t)a * c->nsub_e;
            rc = bb_launch_rm_annulus(&m->g, c->d.likebox, sre, c->nsub_e, c->annd.cap, n, c->w_thin, c->w_thick,
                                      c->e_sub + nse * (bb_analytic_sub_bytes() / sizeof(double)),
                                      c->e_wt + nse * (bb_analytic_wt_bytes() / sizeof(double)), nm, c->d.nr,
                                      c->qunit, (unsigned long long*)counts, st);
            if (rc) rc = bb_fail(BB_ERR_HIP, "bb_loglike_launch: annulus maps: %s", hipGetErrorString((hipError_t)rc));
        } else if (!rc) {
            rc = bb_ring_maps_deposit(m, c->pe, c->d.likebox, sre, c->nsub_e, n, c->widx_e, c->wval_e, nm, c->d.nr, 1,
                                      c->qunit, c->awork[p], c->awork_b[p], counts, st);
        }
        const int lby = c->d.likebox.bhi1 - c->d.likebox.blo1, lbx = c->d.likebox.bhi0 - c->d.likebox.blo0;
        if (!rc) rc = bb_ring_maps_finish(m, counts, n * nm, lby, lbx, c->unit, m->d_kern64 != nullptr, maps, st);
        if (!rc) rc = mark(1);
        if (!rc && after_maps) rc = bb_hip_check(hipEventRecord(after_maps, st), "bb_loglike_launch");
        if (!rc && c->ann_free) {
            const size_t nrow = nm > 2 ? (size_t)(nm - 2) : 0;
            rc = bb_profile_solve_pv(c->prof, maps, n, c->pwork[p], c->pwork_b[p], c->x + (size_t)a * nm,
                                     c->annd.lcoef + (size_t)a * 3 * nrow, c->annd.lcol + (size_t)a * 3 * nrow,
                                     c->annd.near + (size_t)a * nrow, c->annd.norm + a, st);
        } else if (!rc) {
            rc = bb_profile_solve(c->prof, maps, n, c->pwork[p], c->pwork_b[p], c->x + (size_t)a * nm, st);
        }
        if (!rc)
            rc = bb_profile_set_dens_ew(c->x + (size_t)a * nm, n, nm, c->d.nr, c->widx, c->wval, ew, c->nsub_t, sr,
                                        c->dens + (size_t)a * c->d.nr, st);
    } else {
        rc = mark(1);
        if (!rc && after_maps) rc = bb_hip_check(hipEventRecord(after_maps, st), "bb_loglike_launch");
    }
    if (!rc) rc = mark(2);
    if (rc) return rc;
    float* dep = c->dep + (size_t)a * c->vox;
    const int4* va = c->d.box_free ? D.va + a : nullptr;
    const int4* vr = c->d.box_free ? D.vr + a : nullptr;
    if (!c->d.mc) {
        // the ring maps' sample boxes serve the deposit (they depend on the geometry only): the same samples (plan
        // pm, these vectors, this work buffer) in the same box, unless the fixed annulus's maps, on the same work
        // after them, reached that part of it
        const bb_box L = c->d.likebox, D0 = c->db;
        bool reuse = c->prof && !c->d.box_free && L.blo0 == D0.blo0 && L.blo1 == D0.blo1 && L.bhi0 == D0.bhi0 &&
                     L.bhi1 == D0.bhi1;
        if (reuse && !c->ann_free)
            reuse = bb_ring_maps_work_size(c->pe, n) - 256 <=
                    (int64_t)n * c->nsub_t * (int64_t)(bb_analytic_sub_bytes() + bb_analytic_wt_bytes());
        rc = bb_galmod_analytic_vb(m, c->pm, c->db, va, sr, c->nsub_t, n, c->awork[p], c->awork_b[p], dep, 0, st,
                                   reuse ? 1 : 0);
    } else {
        rc = bb_galmod_mc_vb(m, c->db, va, sr, c->nsub_t, n, c->d.nvpar, c->d.seed, c->mc_cap, c->mwork[p],
                             c->mwork_b[p], dep, nullptr, c->mcflag + p, 0, st);
    }
    if (!rc) rc = mark(3);
    if (rc) return rc;
    const bool sigfree = c->off[BB_K_SIGMA] >= 0, epsfree = c->off[BB_K_FRACERR] >= 0;
    const float2* sv = (sigfree || epsfree) ? D.sv + a : nullptr;
    const double s2 = c->d.sigma * c->d.sigma, e2 = c->d.eps * c->d.eps;
    double* cw = c->cwork + (size_t)a * c->cper;
    if (m->d_kern) {
        rc = bb_launch_conv_resid(dep, c->dvalid, n, c->by, c->bx, c->nz, m->d_kern, m->ky, m->kx, (float)m->scale, s2,
                                  e2, cw, vr, sv, c->frac ? 1 : 0, st);
        if (!rc) rc = bb_launch_residual_final(cw, n, (int)(c->cper / 2), c->ss + a, c->lg + a, st);
    } else {
        rc = bb_launch_residual(dep, c->dvalid, n, c->vox, s2, e2, cw, c->ss + a, c->lg + a, vr, sv, c->frac ? 1 : 0,
                                c->bx, c->nz, st);
    }
    if (rc) return bb_fail(BB_ERR_HIP, "bb_loglike_launch: %s", hipGetErrorString((hipError_t)rc));
    return mark(4);
}

// the device work of the current launch (c->nv vectors staged): tables, parts, results to the pinned buffer
int enqueue_body(bb_ctx* c, hipStream_t st) {
    const int nv = c->nv;
    Stage D = stage_of(c, c->stage_d);
    int rc = bb_hip_check(hipMemcpyAsync(c->stage_d, c->stage_h, c->stage_bytes, hipMemcpyHostToDevice, st),
                          "bb_loglike_launch: upload");
    if (!rc) rc = launch_tables(c, c->plan, true, D.th, nv, c->sr, st);
    if (!rc && c->ann_free) {
        rc = bb_hip_check((hipError_t)bb_launch_zero_u32(c->mcflag + 2, 1, st), "bb_loglike_launch");
        if (!rc) rc = launch_ann(c, D.th, nv, st);
    } else if (!rc && c->prof) {
        rc = launch_tables(c, c->ann, false, D.th, nv, c->sr_e, st);
    }
    if (rc) return rc;
    const bool halves = c->d.halves && nv >= 64;
    c->nparts = halves ? 2 : 1;
    if (!halves) {
        rc = run_part(c, 0, 0, nv, D, st, nullptr);
    } else {
        const int h = nv / 2;
        rc = bb_hip_check(hipEventRecord(c->ev_tables, st), "bb_loglike_launch");
        for (int p = 0; p < 2 && !rc; p++) {
            rc = bb_hip_check(hipStreamWaitEvent(c->s[p], c->ev_tables, 0), "bb_loglike_launch");
            if (!rc && p == 1) rc = bb_hip_check(hipStreamWaitEvent(c->s[1], c->ev_maps, 0), "bb_loglike_launch");
            if (!rc) rc = run_part(c, p, p ? h : 0, p ? nv - h : h, D, c->s[p], p == 0 ? c->ev_maps : nullptr);
            if (!rc) rc = bb_hip_check(hipEventRecord(c->ev_part[p], c->s[p]), "bb_loglike_launch");
            if (!rc) rc = bb_hip_check(hipStreamWaitEvent(st, c->ev_part[p], 0), "bb_loglike_launch");
        }
    }
    if (rc) return rc;
    const size_t C = (size_t)c->d.chunk;
    double* ss_h = (double*)c->out_h;
    double* lg_h = ss_h + C;
    float* dens_h = (float*)(lg_h + C);
    int* flag_h = (int*)(dens_h + C * c->d.nr);
    rc = bb_hip_check(hipMemcpyAsync(ss_h, c->ss, nv * sizeof(double), hipMemcpyDeviceToHost, st), "download");
    if (!rc) rc = bb_hip_check(hipMemcpyAsync(lg_h, c->lg, nv * sizeof(double), hipMemcpyDeviceToHost, st), "download");
    if (!rc && c->prof)
        rc = bb_hip_check(hipMemcpyAsync(dens_h, c->dens, (size_t)nv * c->d.nr * sizeof(float), hipMemcpyDeviceToHost,
                                         st), "download");
    if (!rc && (c->d.mc || c->ann_free))
        rc = bb_hip_check(hipMemcpyAsync(flag_h, c->mcflag, 3 * sizeof(int), hipMemcpyDeviceToHost, st), "download");
    return rc;
}

void drop_graphs(bb_ctx* c) {
    for (auto& e : c->gx) (void)hipGraphExecDestroy(e.exec);
    c->gx.clear();
    c->seen.clear();
}

// the device work, eagerly or as the replay of a captured graph of this batch size: the first launch of a size runs
// eagerly (it sets the kernels' attributes), the second is captured on a private stream, later ones replay it
int enqueue(bb_ctx* c, hipStream_t st) {
    int rc = 0;
    if (!c->use_graphs) {
        rc = enqueue_body(c, st);
    } else {
        const bb_model* m = c->m;
        hipGraphExec_t exec = nullptr;
        for (auto& e : c->gx)
            if (e.nv == c->nv && e.TX == m->TX && e.TY == m->TY && e.bin == m->bin_per && e.cap == c->mc_cap &&
                e.halves == c->d.halves)
                exec = e.exec;
        const bool seen = std::find(c->seen.begin(), c->seen.end(), c->nv) != c->seen.end();
        if (!exec && !seen) {
            c->seen.push_back(c->nv);
            rc = enqueue_body(c, st);
        } else {
            if (!exec) {
                hipGraph_t g = nullptr;
                c->capturing = true;
                hipError_t e = hipStreamBeginCapture(c->scap, hipStreamCaptureModeRelaxed);
                if (e == hipSuccess) {
                    const int r2 = enqueue_body(c, c->scap);
                    e = hipStreamEndCapture(c->scap, &g);
                    if (e == hipSuccess && r2) e = hipErrorUnknown;
                }
                c->capturing = false;
                if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e != hipSuccess) return bb_fail(BB_ERR_HIP, "bb_loglike_launch: graph capture: %s", hipGetErrorString(e));
                if (c->gx.size() >= 8) {
                    (void)hipGraphExecDestroy(c->gx.front().exec);
                    c->gx.erase(c->gx.begin());
                }
                c->gx.push_back({c->nv, m->TX, m->TY, m->bin_per, c->d.halves, c->mc_cap, exec});
            }
            rc = bb_hip_check(hipGraphLaunch(exec, st), "bb_loglike_launch: graph");
            c->graph_replays++;
        }
    }
    if (!rc) rc = bb_hip_check(hipEventRecord(c->ev_done, st), "bb_loglike_launch");
    return rc;
}

int alloc_mc(bb_ctx* c, int64_t cap) {
    drop_graphs(c);        // (they hold the old buffers)
    for (int p = 0; p < 2; p++) {
        if (c->mwork[p]) {
            auto it = std::find(c->dev_allocs.begin(), c->dev_allocs.end(), c->mwork[p]);
            if (it != c->dev_allocs.end()) c->dev_allocs.erase(it);
            (void)hipFree(c->mwork[p]);
            c->mwork[p] = nullptr;
        }
        const int nmax = p == 0 ? c->d.chunk : c->d.chunk / 2 + 1;
        c->mwork_b[p] = bb_mc_work_size(c->m, c->db, c->nsub_t, nmax, cap);
        if (c->mwork_b[p] < 0) return bb_fail(BB_ERR_ARG, "bb_ctx: bad MC work size");
        char* w = nullptr;
        int rc = dalloc(c, &w, (size_t)c->mwork_b[p], "bb_ctx: MC work");
        if (rc) return rc;
        c->mwork[p] = w;
    }
    c->mc_cap = cap;
    return 0;
}

}  // namespace

extern "C" int bb_ctx_destroy(bb_ctx* c) {
    if (!c) return 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(c->device);
    if (c->pending) (void)hipEventSynchronize(c->ev_done);
    drop_graphs(c);
    if (c->scap) (void)hipStreamDestroy(c->scap);
    for (void* p : c->dev_allocs) (void)hipFree(p);
    if (c->stage_h) (void)hipHostFree(c->stage_h);
    if (c->out_h) (void)hipHostFree(c->out_h);
    for (int p = 0; p < 2; p++) {
        if (c->s[p]) (void)hipStreamDestroy(c->s[p]);
        if (c->ev_part[p]) (void)hipEventDestroy(c->ev_part[p]);
    }
    if (c->ev_tables) (void)hipEventDestroy(c->ev_tables);
    if (c->ev_maps) (void)hipEventDestroy(c->ev_maps);
    if (c->ev_done) (void)hipEventDestroy(c->ev_done);
    if (c->pm) bb_azplan_destroy(c->pm);
    if (c->pe) bb_azplan_destroy(c->pe);
    if (c->prof) bb_profile_destroy(c->prof);
    (void)hipSetDevice(dev);
    delete c;
    return 0;
}

extern "C" int bb_ctx_create(const bb_model* m, const bb_like_desc* d, const float* data, const uint8_t* likemask,
                             const uint8_t* region, const double* rings, const int32_t* free_off,
                             const int32_t* free_ring, const bb_ring_plan* plan, const bb_ring_plan* annulus,
                             const bb_like_profile* prof, bb_ctx** out) {
    if (!out) return bb_fail(BB_ERR_ARG, "bb_ctx_create: out is NULL");
    *out = nullptr;
    if (!m || !d || !data || !rings || !free_off || !free_ring || !plan)
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: NULL argument");
    if (d->nr < 1 || d->ndim < 1 || d->chunk < 1 || d->chunk > 65535 || d->nx < 1 || d->ny < 1)
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: bad sizes (nr %d, ndim %d, chunk %d)", d->nr, d->ndim, d->chunk);
    const bb_box L = d->likebox, F = d->fixbox;
    if (L.blo0 < 0 || L.blo1 < 0 || L.bhi0 > d->nx || L.bhi1 > d->ny || L.bhi0 <= L.blo0 || L.bhi1 <= L.blo1)
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: likelihood box outside the cube or empty");
    if (!d->box_free && (F.blo0 < L.blo0 || F.blo1 < L.blo1 || F.bhi0 > L.bhi0 || F.bhi1 > L.bhi1 ||
                         F.bhi0 <= F.blo0 || F.bhi1 <= F.blo1))
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: the fixed model box must be a non-empty box of the likelihood box");
    if (!(d->sigma > 0) || !(d->likefactor > 0) || !(d->eps >= 0))
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: sigma and likefactor must be > 0, eps >= 0");
    if (d->box_free && (!(d->pixas0 > 0) || !(d->pixas1 > 0) || !(d->ext >= 0)))
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: per-vector boxes need the pixel sizes and the extent");
    if (d->mc && (d->nvpar < 1 || d->mc_cap < 1)) return bb_fail(BB_ERR_ARG, "bb_ctx_create: MC needs nvpar, mc_cap");
    for (int k = 0; k < BB_NKEY; k++) {
        const int n = free_ring[k] ? d->nr : 1;
        if (free_off[k] >= 0 && free_off[k] + n > d->ndim)
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: free key %d beyond theta", k);
        if (k >= BB_K_SIGMA && free_off[k] >= 0 && free_ring[k])
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: SIGMA, FRACERR, RADMAX, RADMAXA and RADMAXR are single values");
    }
    const bool edges_free = free_off[BB_K_RADMAX] >= 0 || free_off[BB_K_RADMAXA] >= 0 || free_off[BB_K_RADMAXR] >= 0;
    if (free_off[BB_K_RADMAX] >= 0 && !(d->edgewidth >= 0))
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: the width of the edge must be >= 0");
    if (prof && edges_free) {
        if (!(d->annw > 0) || !(d->oedge >= 0) || d->ann_cap < 1 || d->ann_cap > BB_ANN_CAP)
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: a free edge needs the annulus width, the outer edge and an "
                                       "annulus capacity of 1..%d sub-rings (%d)", BB_ANN_CAP, d->ann_cap);
        if (!prof->radii || (d->nr > 1 && !(prof->dmin > 0)))
            return bb_fail(BB_ERR_ARG, "bb_ctx_create: a free edge needs the radii of the rings and their spacing");
        for (int i = 1; i < d->nr; i++)
            if (!(prof->radii[i] > prof->radii[i - 1]))
                return bb_fail(BB_ERR_ARG, "bb_ctx_create: the radii must increase");
    }
    if (prof && !annulus) return bb_fail(BB_ERR_ARG, "bb_ctx_create: the profile needs the annulus plan");
    bb_ctx* c = new bb_ctx();
    c->m = m;
    c->d = *d;
    (void)hipGetDevice(&c->device);
    c->nz = m->g.nz;
    c->rings.assign(rings, rings + (size_t)BB_NKEY * d->nr);
    for (int k = 0; k < BB_NKEY; k++) {
        c->off[k] = free_off[k];
        c->ring[k] = free_ring[k] ? 1 : 0;
        c->fit.off[k] = free_off[k];
        c->fit.ring[k] = c->ring[k];
    }
    c->fit.nr = d->nr;
    c->fit.ndim = d->ndim;
    int rc = dupload(c, &c->fit.rings, rings, (size_t)BB_NKEY * d->nr, "bb_ctx_create: rings");
    if (!rc) rc = copy_plan(c, plan, &c->plan, "model");
    c->thick = plan->thick != 0;
    c->nsub_t = c->plan.nsub_t;
    c->edge = {free_off[BB_K_RADMAX], free_off[BB_K_RADMAXA], free_off[BB_K_RADMAXR], d->edgewidth, m->g.radsep * 60.,
               nullptr};
    if (!rc && prof) {
        rc = copy_plan(c, annulus, &c->ann, "annulus");
        c->nsub_e = c->ann.nsub_t;
    }
    c->ann_free = prof && edges_free;
    if (!rc && c->ann_free) {
        // the annulus of every vector: cap sub-rings per layer; its ring weights are those of the fixed annulus's
        // first thin and thick sub-rings (the same for all of its sub-rings: equal model rings, no edge)
        c->nsub_e = d->ann_cap * (c->ann.thick ? 2 : 1);
        if (annulus->nsub < 1 || !prof->wval_e || !prof->widx_e || prof->widx_e[0] != 0)
            rc = bb_fail(BB_ERR_ARG, "bb_ctx_create: the annulus plan has no sub-ring of its ring");
        else {
            c->w_thin = prof->wval_e[0];
            c->w_thick = c->ann.thick ? prof->wval_e[2 * annulus->nsub] : 0.;
        }
        AnnDev& A = c->annd;
        A.oe = free_off[BB_K_RADMAX];
        A.oa = free_off[BB_K_RADMAXA];
        A.orr = free_off[BB_K_RADMAXR];
        A.edge0 = d->edge0;
        A.ea0 = d->edgea0;
        A.er0 = d->edger0;
        A.oedge = d->oedge;
        A.w = d->annw;
        A.radsep = m->g.radsep;
        A.empty = d->empty ? 1 : 0;
        A.cap = d->ann_cap;
        A.nr = d->nr;
        A.dmin = prof->dmin;
        A.dmed = prof->d.dmed;
        const size_t C = (size_t)d->chunk, nrow = d->nr + 1 > 2 ? (size_t)(d->nr - 1) : 0;
        if (!rc) rc = dupload(c, &A.radii, prof->radii, (size_t)d->nr, "bb_ctx_create: radii");
        if (!rc) rc = dalloc(c, &A.lcoef, C * 3 * nrow, "bb_ctx_create: prior rows");
        if (!rc) rc = dalloc(c, &A.lcol, C * 3 * nrow, "bb_ctx_create: prior rows");
        if (!rc) rc = dalloc(c, &A.near, C * nrow, "bb_ctx_create: prior rows");
        if (!rc) rc = dalloc(c, &A.norm, C, "bb_ctx_create: prior rows");
        if (!rc) rc = dalloc(c, &c->e_sub, C * c->nsub_e * (bb_analytic_sub_bytes() / sizeof(double)),
                             "bb_ctx_create: annulus records");
        if (!rc) rc = dalloc(c, &c->e_wt, C * c->nsub_e * (bb_analytic_wt_bytes() / sizeof(double)),
                             "bb_ctx_create: annulus records");
    }
    if (!rc && prof && free_off[BB_K_RADMAX] >= 0) rc = dalloc(c, &c->edge.ew, (size_t)d->chunk * c->nsub_t,
                                                               "bb_ctx_create: edge weights");
    if (rc) {
        bb_ctx_destroy(c);
        return rc;
    }
    if (c->nsub_t < 1) {
        bb_ctx_destroy(c);
        return bb_fail(BB_ERR_ARG, "bb_ctx_create: the model has no sub-rings");
    }
    // the data: the voxels of the likelihood (likemask, not blank, in the region); the summed-area table of their
    // squared data over the region (likelihood.py:166-229)
    const int nx = d->nx, ny = d->ny, nz = c->nz;
    c->db = d->box_free ? L : F;
    c->bx = c->db.bhi0 - c->db.blo0;
    c->by = c->db.bhi1 - c->db.blo1;
    c->vox = (int64_t)c->by * c->bx * nz;
    // one pass over the cube on the device (bb_like_maps): the data of the likelihood in the box and, per pixel, the
    // squared data and number of the likelihood voxels; then the table on the host, summed down the columns first
    // and along the rows next, as numpy's cumsum(cumsum(a, 0), 1) (likelihood.py:222-229)
    const size_t npix = (size_t)ny * nx, nvx = npix * nz;
    float* t_data = nullptr;
    uint8_t *t_lm = nullptr, *t_reg = nullptr;
    double* t_dm2 = nullptr;
    int32_t* t_nv = nullptr;
    float* dvp = nullptr;
    rc = dalloc(c, &dvp, (size_t)c->vox, "bb_ctx_create: data");
    c->dvalid = dvp;
    if (!rc) rc = bb_hip_check(hipMalloc((void**)&t_data, nvx * sizeof(float)), "bb_ctx_create: staging");
    if (!rc && likemask) rc = bb_hip_check(hipMalloc((void**)&t_lm, nvx), "bb_ctx_create: staging");
    if (!rc && region) rc = bb_hip_check(hipMalloc((void**)&t_reg, npix), "bb_ctx_create: staging");
    if (!rc) rc = bb_hip_check(hipMalloc((void**)&t_dm2, npix * sizeof(double)), "bb_ctx_create: staging");
    if (!rc) rc = bb_hip_check(hipMalloc((void**)&t_nv, npix * sizeof(int32_t)), "bb_ctx_create: staging");
    if (!rc) rc = bb_hip_check(hipMemcpy(t_data, data, nvx * sizeof(float), hipMemcpyHostToDevice), "bb_ctx_create");
    if (!rc && t_lm) rc = bb_hip_check(hipMemcpy(t_lm, likemask, nvx, hipMemcpyHostToDevice), "bb_ctx_create");
    if (!rc && t_reg) rc = bb_hip_check(hipMemcpy(t_reg, region, npix, hipMemcpyHostToDevice), "bb_ctx_create");
    if (!rc) {
        const bb_box db = {c->db.blo0, c->db.blo1, c->db.bhi0, c->db.bhi1};
        rc = bb_like_maps(t_data, t_lm, t_reg, nz, ny, nx, db, dvp, t_dm2, t_nv, nullptr, nullptr);
    }
    std::vector<double> dm2(npix, 0.);
    std::vector<int32_t> nvm(npix, 0);
    if (!rc) rc = bb_hip_check(hipMemcpy(dm2.data(), t_dm2, npix * sizeof(double), hipMemcpyDeviceToHost),
                               "bb_ctx_create");
    if (!rc) rc = bb_hip_check(hipMemcpy(nvm.data(), t_nv, npix * sizeof(int32_t), hipMemcpyDeviceToHost),
                               "bb_ctx_create");
    for (void* q : {(void*)t_data, (void*)t_lm, (void*)t_reg, (void*)t_dm2, (void*)t_nv})
        if (q) (void)hipFree(q);
    if (rc) {
        bb_ctx_destroy(c);
        return rc;
    }
    double nvox = 0.;
    for (size_t q = 0; q < npix; q++) {
        const bool inreg = !region || region[q];
        if (!inreg) dm2[q] = 0.;
        else nvox += (double)nvm[q];
    }
    c->sat.assign((size_t)(ny + 1) * (nx + 1), 0.);
    std::vector<double> col(nx, 0.);
    for (int y = 0; y < ny; y++) {
        double row = 0.;
        for (int x = 0; x < nx; x++) {
            col[x] += dm2[(size_t)y * nx + x];
            row += col[x];
            c->sat[(size_t)(y + 1) * (nx + 1) + x + 1] = row;
        }
    }
    c->sat_total = c->sat[(size_t)ny * (nx + 1) + nx];
    c->nvox = nvox;
    auto boxsum = [&](int x0, int y0, int x1, int y1) {
        const size_t W = nx + 1;
        return c->sat[(size_t)y1 * W + x1] - c->sat[(size_t)y0 * W + x1] - c->sat[(size_t)y1 * W + x0] +
               c->sat[(size_t)y0 * W + x0];
    };
    c->ss_out_fix = d->box_free ? 0. : std::max(c->sat_total - boxsum(F.blo0, F.blo1, F.bhi0, F.bhi1), 0.);
    // azimuth plans of the sub-ring radii (rows 0 of the tables: one layer, or thin then thick)
    std::vector<double> R0(c->nsub_t);
    for (int s = 0; s < c->nsub_t; s++) R0[s] = plan->rads[s % plan->nsub];
    if (!rc) rc = bb_azplan_create(m, R0.data(), c->nsub_t, &c->pm);
    if (!rc && prof) {
        std::vector<double> Re(c->nsub_e);
        for (int s = 0; s < c->nsub_e; s++) Re[s] = annulus->rads[s % annulus->nsub];
        rc = bb_azplan_create(m, Re.data(), c->nsub_e, &c->pe);
        bb_profile_desc pd = prof->d;
        const int lbx = L.bhi0 - L.blo0, lby = L.bhi1 - L.blo1;
        if (!rc && (pd.nr != d->nr || pd.npix != lbx * lby))
            rc = bb_fail(BB_ERR_ARG, "bb_ctx_create: the profile must have nr rings and the pixels of the box");
        if (!rc) rc = bb_profile_create(&pd, prof->wpix, prof->ypix, prof->lcoef, prof->lcol, prof->near, &c->prof);
        if (!rc && (!prof->widx || !prof->wval || !prof->widx_e || !prof->wval_e))
            rc = bb_fail(BB_ERR_ARG, "bb_ctx_create: the profile needs the ring weights");
        if (!rc) rc = dupload(c, &c->widx, prof->widx, (size_t)2 * c->nsub_t, "bb_ctx_create: weights");
        if (!rc) rc = dupload(c, &c->wval, prof->wval, (size_t)2 * c->nsub_t, "bb_ctx_create: weights");
        if (!rc) rc = dupload(c, &c->widx_e, prof->widx_e, (size_t)2 * c->nsub_e, "bb_ctx_create: weights");
        if (!rc) rc = dupload(c, &c->wval_e, prof->wval_e, (size_t)2 * c->nsub_e, "bb_ctx_create: weights");
        c->nprof = d->nr + 1;
        c->qunit = prof->qunit;
        c->unit = prof->unit;
        if (!rc && !(c->qunit > 0)) rc = bb_fail(BB_ERR_ARG, "bb_ctx_create: qunit must be > 0");
    }
    // buffers of a launch of chunk vectors
    const size_t C = (size_t)d->chunk;
    const int64_t P = (int64_t)(L.bhi0 - L.blo0) * (L.bhi1 - L.blo1);
    c->stage_bytes = C * d->ndim * sizeof(double) + C * 2 * sizeof(int4) + C * sizeof(float2);
    if (!rc) rc = dalloc(c, &c->stage_d, c->stage_bytes, "bb_ctx_create");
    if (!rc) rc = bb_hip_check(hipHostMalloc((void**)&c->stage_h, c->stage_bytes, hipHostMallocDefault), "pinned");
    if (!rc)
        rc = bb_hip_check(hipHostMalloc((void**)&c->out_h, C * 2 * sizeof(double) + C * d->nr * sizeof(float) + 16,
                                        hipHostMallocDefault), "pinned");
    if (!rc) rc = dalloc(c, &c->sr, C * 16 * c->nsub_t, "bb_ctx_create");
    if (!rc && prof) rc = dalloc(c, &c->sr_e, C * 16 * c->nsub_e, "bb_ctx_create");
    if (!rc) rc = dalloc(c, &c->dep, C * c->vox, "bb_ctx_create: models");
    if (!rc && prof) rc = dalloc(c, &c->counts, C * c->nprof * P, "bb_ctx_create: map counts");
    if (!rc && prof) rc = dalloc(c, &c->maps, C * c->nprof * P, "bb_ctx_create: maps");
    if (!rc && prof) rc = dalloc(c, &c->x, C * c->nprof, "bb_ctx_create");
    if (!rc) rc = dalloc(c, &c->dens, C * d->nr, "bb_ctx_create");
    c->cper = m->d_kern ? (int64_t)bb_conv_resid_blocks(c->by, c->bx, nz) * 2 : (int64_t)bb_residual_blocks(c->vox) * 2;
    if (!rc) rc = dalloc(c, &c->cwork, C * c->cper, "bb_ctx_create");
    if (!rc) rc = dalloc(c, &c->ss, C, "bb_ctx_create");
    if (!rc) rc = dalloc(c, &c->lg, C, "bb_ctx_create");
    if (!rc) rc = dalloc(c, &c->mcflag, 4, "bb_ctx_create");
    if (!rc) rc = bb_hip_check(hipMemset(c->mcflag, 0, 4 * sizeof(int)), "bb_ctx_create");
    c->annd.overflow = c->mcflag + 2;
    for (int p = 0; p < 2 && !rc; p++) {
        const int nmax = p == 0 ? d->chunk : d->chunk / 2 + 1;
        int64_t a = bb_analytic_work_size(c->pm, nmax);
        if (prof) a = std::max(a, bb_ring_maps_work_size(c->pe, nmax));
        c->awork_b[p] = a;
        char* w = nullptr;
        rc = dalloc(c, &w, (size_t)a, "bb_ctx_create: deposit work");
        c->awork[p] = w;
        if (!rc && prof) {
            c->pwork_b[p] = bb_profile_work_size(c->prof, nmax);
            char* q = nullptr;
            rc = dalloc(c, &q, (size_t)c->pwork_b[p], "bb_ctx_create: profile work");
            c->pwork[p] = q;
        }
    }
    if (!rc && d->mc) rc = alloc_mc(c, d->mc_cap);
    for (int p = 0; p < 2 && !rc; p++) {
        rc = bb_hip_check(hipStreamCreateWithFlags(&c->s[p], hipStreamNonBlocking), "bb_ctx_create: streams");
        if (!rc) rc = bb_hip_check(hipEventCreateWithFlags(&c->ev_part[p], hipEventDisableTiming), "bb_ctx_create");
    }
    if (!rc) rc = bb_hip_check(hipStreamCreateWithFlags(&c->scap, hipStreamNonBlocking), "bb_ctx_create: streams");
    if (!rc) rc = bb_hip_check(hipEventCreateWithFlags(&c->ev_tables, hipEventDisableTiming), "bb_ctx_create");
    if (!rc) rc = bb_hip_check(hipEventCreateWithFlags(&c->ev_maps, hipEventDisableTiming), "bb_ctx_create");
    if (!rc) rc = bb_hip_check(hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming), "bb_ctx_create");
    if (rc) {
        bb_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

extern "C" int bb_ctx_info(const bb_ctx* c, double* nvox, double* ss_total, int32_t* deposit_box, int64_t* nnls_fail,
                           int64_t* mc_retries) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_ctx_info: ctx is NULL");
    if (nvox) *nvox = c->nvox;
    if (ss_total) *ss_total = c->sat_total;
    if (deposit_box) {
        deposit_box[0] = c->db.blo0;
        deposit_box[1] = c->db.blo1;
        deposit_box[2] = c->db.bhi0;
        deposit_box[3] = c->db.bhi1;
    }
    if (nnls_fail) *nnls_fail = c->nnls_fail;
    if (mc_retries) *mc_retries = c->mc_retries;
    return 0;
}

extern "C" int64_t bb_ctx_graph_replays(const bb_ctx* c) { return c ? c->graph_replays : -1; }

extern "C" int bb_ctx_set_halves(bb_ctx* c, int halves) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_ctx_set_halves: ctx is NULL");
    if (c->pending) return bb_fail(BB_ERR_ARG, "bb_ctx_set_halves: a launch is in flight");
    c->d.halves = halves ? 1 : 0;
    return 0;
}

extern "C" int bb_ctx_set_graphs(bb_ctx* c, int on) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_ctx_set_graphs: ctx is NULL");
    if (c->pending) return bb_fail(BB_ERR_ARG, "bb_ctx_set_graphs: a launch is in flight");
    c->use_graphs = on ? 1 : 0;
    if (!on) drop_graphs(c);
    return 0;
}

extern "C" int bb_ctx_trace(bb_ctx* c, void* const* ev) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_ctx_trace: ctx is NULL");
    for (int k = 0; k < 5; k++) c->trace[k] = ev ? (hipEvent_t)ev[k] : nullptr;
    if (ev)
        for (int k = 0; k < 5; k++)
            if (!c->trace[k]) return bb_fail(BB_ERR_ARG, "bb_ctx_trace: give 5 events or none");
    return 0;
}

extern "C" int bb_loglike_launch(bb_ctx* c, const double* theta, int B, void* stream) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_loglike_launch: ctx is NULL");
    if (c->pending) return bb_fail(BB_ERR_ARG, "bb_loglike_launch: collect the previous launch first");
    if (B < 0 || B > c->d.chunk) return bb_fail(BB_ERR_ARG, "bb_loglike_launch: B = %d (0..%d)", B, c->d.chunk);
    if (B > 0 && !theta) return bb_fail(BB_ERR_ARG, "bb_loglike_launch: theta is NULL");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != c->device) return bb_fail(BB_ERR_ARG, "bb_loglike_launch: the ctx lives on device %d, not %d", c->device, dev);
    const bb_like_desc& d = c->d;
    const int nr = d.nr;
    Stage H = stage_of(c, c->stage_h);
    c->B = B;
    c->valid.clear();
    c->s2.clear();
    c->e2.clear();
    c->ssout.clear();
    c->frac = d.eps > 0;
    const int kz0 = BB_K_Z0, kzt = BB_K_Z0THICK, kft = BB_K_FTHICK;
    for (int i = 0; i < B; i++) {
        const double* th = theta + (size_t)i * d.ndim;
        bool ok = true;
        // the likelihood is zero where a free positive-only value is negative, a free FTHICK is above 1, or (Z0 or
        // Z0THICK free) the thick layer is thinner than the thin one (likelihood.py:537-559)
        for (int k = 0; k < BB_NKEY && ok; k++) {
            if (c->off[k] < 0) continue;
            const int n = c->ring[k] ? nr : 1;
            for (int j = 0; j < n && ok; j++) {
                const double v = th[c->off[k] + j];
                if (!(v == v)) ok = false;
                else if (k == BB_K_RADMAX && v < 0) ok = false;
                else if (positive_only(k) && (float)v < 0.f) ok = false;
                else if (k == kft && (float)v > 1.f) ok = false;
            }
        }
        double sig = d.sigma, eps = d.eps;
        if (ok && c->off[BB_K_SIGMA] >= 0) {
            sig = th[c->off[BB_K_SIGMA]];
            if (!(sig > 0)) ok = false;
        }
        if (ok && c->off[BB_K_FRACERR] >= 0) {
            eps = th[c->off[BB_K_FRACERR]];
            if (eps > 0) c->frac = true;
            else eps = 0.;
        }
        if (ok && c->thick && (c->off[kz0] >= 0 || c->off[kzt] >= 0))
            for (int j = 0; j < nr && ok; j++) {
                const double zt = host_value(c, th, kzt, j);
                if (zt >= 0 && zt < host_value(c, th, kz0, j)) ok = false;
            }
        if (!ok) continue;
        const int j = (int)c->valid.size();
        c->valid.push_back(i);
        c->s2.push_back(sig * sig);
        c->e2.push_back(eps * eps);
        memcpy(H.th + (size_t)j * d.ndim, th, d.ndim * sizeof(double));
        H.sv[j] = make_float2((float)(sig * sig), (float)(eps * eps));
        double out = c->ss_out_fix;
        if (d.box_free) {
            // DiskModel.box (galmod.py:1224-1240) of the vector's rings, within the likelihood box
            double lh = host_value(c, th, kz0, nr - 1);
            if (c->thick && host_value(c, th, kft, nr - 1) > 0) {
                double zt = host_value(c, th, kzt, nr - 1);
                if (!(zt >= 0)) zt = lh;
                lh = std::max(lh, zt);
            }
            const double ext = d.ext + 3 * lh;
            auto iceil = [](double v) { return (long long)std::min(std::max(ceil(v), -1e9), 1e9); };
            const long long xdis = iceil(ext / d.pixas0), ydis = iceil(ext / d.pixas1);
            const long long xp = iceil(host_value(c, th, BB_K_XPOS, nr - 1));
            const long long yp = iceil(host_value(c, th, BB_K_YPOS, nr - 1));
            long long b0 = std::max(xp - xdis, 0LL), b1 = std::max(yp - ydis, 0LL);
            long long h0 = std::min(xp + xdis + 1, (long long)d.nx), h1 = std::min(yp + ydis + 1, (long long)d.ny);
            b0 = std::max(b0, (long long)d.likebox.blo0);
            b1 = std::max(b1, (long long)d.likebox.blo1);
            h0 = std::max(std::min(h0, (long long)d.likebox.bhi0), b0);
            h1 = std::max(std::min(h1, (long long)d.likebox.bhi1), b1);
            H.va[j] = make_int4((int)b0, (int)b1, (int)h0, (int)h1);
            H.vr[j] = make_int4((int)b0 - c->db.blo0, (int)b1 - c->db.blo1, (int)h0 - c->db.blo0,
                                (int)h1 - c->db.blo1);
            const size_t W = d.nx + 1;
            const double in = c->sat[(size_t)h1 * W + h0] - c->sat[(size_t)b1 * W + h0] - c->sat[(size_t)h1 * W + b0] +
                              c->sat[(size_t)b1 * W + b0];
            out = std::max(c->sat_total - in, 0.);
        }
        c->ssout.push_back(out);
    }
    c->nv = (int)c->valid.size();
    hipStream_t st = (hipStream_t)stream;
    c->last_stream = st;
    int rc = 0;
    if (c->nv > 0) {
        rc = enqueue(c, st);          // (the MC deposit clears its overflow flag itself)
    } else {
        rc = bb_hip_check(hipEventRecord(c->ev_done, st), "bb_loglike_launch");
    }
    if (rc) return rc;
    c->pending = true;
    return 0;
}

extern "C" int bb_loglike_collect(bb_ctx* c, double* loglike, float* dens) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_loglike_collect: ctx is NULL");
    if (!c->pending) return bb_fail(BB_ERR_ARG, "bb_loglike_collect: nothing launched");
    if (!loglike) return bb_fail(BB_ERR_ARG, "bb_loglike_collect: loglike is NULL");
    int rc = bb_hip_check(hipEventSynchronize(c->ev_done), "bb_loglike_collect");
    const size_t C = (size_t)c->d.chunk;
    double* ss_h = (double*)c->out_h;
    double* lg_h = ss_h + C;
    float* dens_h = (float*)(lg_h + C);
    int* flag_h = (int*)(dens_h + C * c->d.nr);
    // MC: more clouds in a vector's box than the records hold -> larger buffers, the same launch again
    auto overflow = [&]() { return flag_h[0] != 0 || (c->nparts > 1 && flag_h[1] != 0); };
    for (int retry = 0; !rc && c->d.mc && c->nv > 0 && overflow(); retry++) {
        if (retry >= 8 || c->mc_cap > ((int64_t)1 << 30)) {
            rc = bb_fail(BB_ERR_ARG, "bb_loglike_collect: the MC clouds of a vector exceed %lld records",
                         (long long)c->mc_cap);
            break;
        }
        rc = alloc_mc(c, c->mc_cap * 2);
        c->mc_retries++;
        if (!rc) rc = enqueue(c, c->last_stream);
        if (!rc) rc = bb_hip_check(hipEventSynchronize(c->ev_done), "bb_loglike_collect");
    }
    c->pending = false;
    if (rc) return rc;
    if (c->ann_free && c->nv > 0 && flag_h[2] != 0)
        return bb_fail(BB_ERR_ARG, "bb_loglike_collect: a vector's exterior annulus has more than %d sub-rings",
                       c->annd.cap);
    const bb_like_desc& d = c->d;
    const int nr = d.nr;
    for (int i = 0; i < c->B; i++) loglike[i] = -INFINITY;
    if (dens)
        for (size_t i = 0; i < (size_t)c->B * nr; i++) dens[i] = __builtin_nanf("");
    for (int j = 0; j < c->nv; j++) {
        const int i = c->valid[j];
        bool bad = false;
        if (c->prof) {
            for (int r = 0; r < nr; r++) {
                const float v = dens_h[(size_t)j * nr + r];
                if (!(v == v)) bad = true;
                if (dens) dens[(size_t)i * nr + r] = v;
            }
        }
        if (bad) {
            c->nnls_fail++;           // (the profile of this vector did not converge: its likelihood is zero)
            continue;
        }
        const double s2 = c->s2[j];
        double nll;
        if (c->frac)
            nll = 0.5 * (ss_h[j] + c->ssout[j] / s2) + 0.5 * c->nvox * log(2 * M_PI * s2) + 0.5 * lg_h[j];
        else
            nll = 0.5 * (ss_h[j] + c->ssout[j]) / s2 + 0.5 * c->nvox * log(2 * M_PI * s2);
        loglike[i] = -d.likefactor * nll;
    }
    return 0;
}

extern "C" int bb_loglike_batch(bb_ctx* c, const double* theta, int B, double* loglike, float* dens, void* stream) {
    if (!c) return bb_fail(BB_ERR_ARG, "bb_loglike_batch: ctx is NULL");
    if (B < 0) return bb_fail(BB_ERR_ARG, "bb_loglike_batch: B = %d", B);
    for (int a = 0; a < B; a += c->d.chunk) {
        const int n = std::min(c->d.chunk, B - a);
        int rc = bb_loglike_launch(c, theta + (size_t)a * c->d.ndim, n, stream);
        if (!rc) rc = bb_loglike_collect(c, loglike + a, dens ? dens + (size_t)a * c->d.nr : nullptr);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int bb_ctx_tables(bb_ctx* c, const double* theta, int B, double* sr, double* sr_ext) {
    if (!c || (B > 0 && (!theta || !sr))) return bb_fail(BB_ERR_ARG, "bb_ctx_tables: NULL argument");
    if (c->pending) return bb_fail(BB_ERR_ARG, "bb_ctx_tables: a launch is in flight");
    if (B < 0 || B > c->d.chunk) return bb_fail(BB_ERR_ARG, "bb_ctx_tables: B = %d (0..%d)", B, c->d.chunk);
    if (B == 0) return 0;
    int rc = bb_hip_check(hipMemcpy(c->stage_d, theta, (size_t)B * c->d.ndim * sizeof(double), hipMemcpyHostToDevice),
                          "bb_ctx_tables");
    if (!rc) rc = launch_tables(c, c->plan, true, (const double*)c->stage_d, B, c->sr, nullptr);
    if (!rc && c->prof && sr_ext)
        rc = c->ann_free ? launch_ann(c, (const double*)c->stage_d, B, nullptr)
                         : launch_tables(c, c->ann, false, (const double*)c->stage_d, B, c->sr_e, nullptr);
    if (!rc) rc = bb_hip_check(hipDeviceSynchronize(), "bb_ctx_tables");
    if (!rc)
        rc = bb_hip_check(hipMemcpy(sr, c->sr, (size_t)B * 16 * c->nsub_t * sizeof(double), hipMemcpyDeviceToHost),
                          "bb_ctx_tables");
    if (!rc && c->prof && sr_ext)
        rc = bb_hip_check(hipMemcpy(sr_ext, c->sr_e, (size_t)B * 16 * c->nsub_e * sizeof(double), hipMemcpyDeviceToHost),
                          "bb_ctx_tables");
    return rc;
}
