// Internal to libbbgpu.so: the objects behind the opaque handles of bbgpu.h and the helpers shared by the
// C-ABI files (bb_capi.hip, bb_like.hip). Not installed, not part of the C-ABI.
#ifndef BB_INTERNAL_H
#define BB_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

#include "../../include/bbgpu.h"

#define BB_ERR_ARG 1
#define BB_ERR_HIP 2
#define BB_LDS_MAX (160 * 1024)   // LDS of a gfx950 CU, all usable by one workgroup
#define BB_BIN_PER_CHUNK 8        // most entries per chunk of the tiles' chunk lists (the work buffers hold them)

// the message of the calling thread's last error (bb_last_error); returns code
int bb_fail(int code, const char* fmt, ...);
int bb_hip_check(hipError_t e, const char* what);

template <class T>
static inline int bb_upload(T** dst, const T* src, size_t n, const char* what) {
    *dst = nullptr;
    if (n == 0) return 0;
    int rc = bb_hip_check(hipMalloc((void**)dst, n * sizeof(T)), what);
    if (rc) return rc;
    return bb_hip_check(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), what);
}
#define upload bb_upload

struct bb_model {
    bb_grid g;
    int device;
    double* d_cd2i;
    double* d_vedges;
    int ky, kx;
    float* d_kern;
    double* d_kern64;
    double scale;
    int TX, TY;        // pixel tile of the ANALYTIC deposit (its spectra fit the LDS budget)
    int bin_per;       // capacity of the tiles' chunk lists, in entries per chunk (0: no lists, every tile scans)
    int dep_variant;   // ANALYTIC deposit: 0 one lane per sample, 1 the wavefront deposit (bit-identical models)
    int lin;           // linear velocity axis whose edges are exactly ((k - 1/2) + 1 - crpix3) cdelt3 + crval3 and
                       // one channel factor cd2i for all channels: the deposit computes them instead of reading them
};

// The azimuth plan of a set of sub-ring radii (ANALYTIC samples): naz per sub-ring, the chunks of BB_CS
// azimuths of every (sub-ring, radial sample) and the cosine / sine of every azimuth, built on the host with
// the reference's recurrence (galmod.py:372-388) so that the device rounds pixels as the reference does.
struct bb_azplan {
    int device;
    int nsub;
    int nchunks;
    std::vector<double> R0;
    int* d_naz;
    int4* d_chunks;
    double2* d_az;
};

struct bb_profile {
    bb_profile_desc d;
    int device;
    double* d_wpix;
    double* d_ypix;
    double* d_lcoef;
    int* d_lcol;
    int* d_near;
};


// The deposits with a model box per vector (vbox: B boxes [x, z) x [y, w) in absolute pixels inside box, the layout
// of out; NULL: box for every vector), as bb_galmod_analytic / bb_galmod_mc.
// boxes_ready: the sample boxes in work are those of the same samples in the same box from a previous launch on the
// same work (the ring maps of the likelihood context): k_an_bbox is not run again.
int bb_galmod_analytic_vb(const bb_model* m, const bb_azplan* p, bb_box box, const int4* vbox, const double* sr,
                          int nsub, int B, void* work, int64_t work_bytes, float* out, int accumulate, void* stream,
                          int boxes_ready = 0);
int bb_galmod_mc_vb(const bb_model* m, bb_box box, const int4* vbox, const double* sr, int nsub, int B, int nvpar,
                    uint32_t seed, int64_t max_clouds, void* work, int64_t work_bytes, float* out, uint64_t* counts,
                    int* overflow, int accumulate, void* stream);

// The ring maps, profile and densities with per-vector parts (the likelihood context with a free edge of the disk):
// ew (B, nsub) multiplies the ring weights of each vector's sub-rings (the edge weights of the density; NULL: 1); the
// prior's rows of every vector (lcoef, lcol: B (n-2) 3; near: B (n-2); normv: B) replace the profile's when normv is
// given.
int bb_ring_maps_deposit_ew(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub, int B,
                            const int32_t* widx, const double* wval, const double* ew, int nmap, int map0, int nw,
                            double qunit, void* work, int64_t work_bytes, uint64_t* counts, void* stream);
int bb_profile_solve_pv(const bb_profile* p, const double* maps, int B, void* work, int64_t work_bytes, double* x,
                        const double* lcoef, const int32_t* lcol, const int32_t* near, const double* normv,
                        void* stream);
int bb_profile_set_dens_ew(const double* x, int B, int n, int nr, const int32_t* widx, const double* wval,
                           const double* ew, int nsub, double* sr, float* dens, void* stream);

// kernel launchers shared with bb_like.hip
extern "C" int bb_launch_rm_annulus(const bb_grid* g, bb_box box, const double* sr, int nsub, int nlay, int B,
                                    double w_thin, double w_thick, double* sub, double* wtab, int nmap, int map0,
                                    double qunit, unsigned long long* counts, hipStream_t st);
extern "C" int bb_launch_zero_u32(void* p, int64_t n, hipStream_t st);
extern "C" size_t bb_analytic_sub_bytes();
extern "C" size_t bb_analytic_wt_bytes();
extern "C" int bb_launch_conv_resid(const float* in, const float* dvalid, int B, int by, int bx, int nz,
                                    const float* kern, int ky, int kx, float scale, double sigma2, double eps2,
                                    double* partial, const int4* vbox, const float2* sv, int frac, hipStream_t st);
extern "C" int bb_launch_residual_final(const double* partial, int B, int nblk, double* out_ss, double* out_lg,
                                        hipStream_t st);
extern "C" int bb_launch_residual(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2,
                                  double eps2, double* partial, double* out_ss, double* out_lg, const int4* vbox,
                                  const float2* sv, int frac, int bx, int nz, hipStream_t st);
extern "C" int64_t bb_conv_resid_blocks(int by, int bx, int nz);
extern "C" int bb_residual_blocks(int64_t nvox);
#endif
