/*
 * bbgpu -- the C-ABI of the MI355X model-cube likelihood (GALMOD -> beam convolution -> residual).
 *
 * The reference (BBarolo 2, pure Python + numba) has no FFI: its hot path sits behind numba kernels
 * called with positional numpy arrays. Each entry point below replaces one of those kernels (cited
 * file:line, paths relative to the reference repository) or fuses a chain of them, with the same
 * argument meaning:
 *   - device pointers (allocated by the caller, e.g. torch tensors), sizes and a HIP stream (void*,
 *     NULL = the default stream); no torch types cross this boundary;
 *   - model outputs are accumulated with += into caller-zeroed buffers, as the numba kernels do;
 *   - every entry point returns 0 on success and a non-zero code on error; bb_last_error() gives the
 *     message of the last error of the calling thread. Invalid arguments are rejected before any launch.
 * Launches are asynchronous on the given stream; the caller synchronises.
 */
#ifndef BBGPU_H
#define BBGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BB_MAX_LINES 8

/* Header-derived constants of the models of one cube (GalmodSetup, bbarolo/galmod.py:613-726), in the
 * reference's internal units: arcmin, m/s, radians. chmin is the narrowest channel (galmod.py:726), which
 * the ANALYTIC kernels use as the channel width (galmod.py:909-914). */
typedef struct bb_grid {
    int32_t nz;                  /* channels */
    int32_t ltype;               /* vertical layer: 1 gauss, 2 sech2, 3 exp, 4 lorentz, other box */
    int32_t axtyp;               /* 2 wavelength, 3 frequency, 4 velocity */
    int32_t velsys;              /* 1 radio, 2 optical, 3 relativistic */
    int32_t nlines;              /* emission lines (<= BB_MAX_LINES) */
    int32_t nzq;                 /* vertical-sample capacity of the ANALYTIC kernels (1 or 1024) */
    int32_t cmode;               /* MC cloud-number exponent */
    int32_t pad0;
    double radsep, cdens, pixarea, cdelt0, cdelt1, crota2;
    double sig_instr, crval3, cdelt3, crpix3, chmin, zstep;
    double relint[BB_MAX_LINES];
    double lfreq0[BB_MAX_LINES];
} bb_grid;

/* A box [blo, bhi) of the cube in pixels (DiskModel.box, galmod.py:1224-1240). */
typedef struct bb_box {
    int32_t blo0, blo1, bhi0, bhi1;
} bb_box;

const char* bb_last_error(void);
int bb_device_count(int* n);

/* The per-fit model object: the header-derived setup of GalmodSetup (bbarolo/galmod.py:609-758) and the beam
 * kernel of DiskModel (galmod.py:1199-1214, FFTConvolver(kernel) of smooth.py:107-143), copied to the current
 * device. cd2i (nz), vedges (nlines, nz+1) and kern (ky, kx) are HOST float64 arrays; kern may be NULL (no
 * smoothing); its sizes are odd with kx <= 25. scale: the flux scaling after the convolution (1 for DENSFLUX
 * models). The line profile's cumulative normal, which the reference tabulates (galmod.py:183-191), the deposit
 * evaluates in registers (within 1.5e-7, Abramowitz & Stegun 7.1.26): no table is passed. */
typedef struct bb_model bb_model;
int bb_model_create(const bb_grid* g, const double* cd2i, const double* vedges, const double* kern, int ky, int kx,
                    double scale, bb_model** out);
int bb_model_destroy(bb_model* m);

/* The pixel tile (TX x TY) of the ANALYTIC deposit: by default the largest whose spectra fit the LDS of a CU;
 * set_tile lets a tuner choose another (rejected if it does not fit). No counterpart in the reference. */
int bb_model_set_tile(bb_model* m, int TX, int TY);
int bb_model_get_tile(const bb_model* m, int* TX, int* TY);
/* The capacity of the per-tile chunk lists of the deposits (a counting sort of the chunks' pixel boxes by tile),
 * in entries per chunk, 0..8 (default 8). A vector whose chunk boxes meet more tiles than that has no lists and its
 * tiles test every chunk box; 0 turns the lists off. The results do not depend on it (exact integer sums). No
 * counterpart in the reference. */
int bb_model_set_bin(bb_model* m, int per_chunk);
/* The ANALYTIC deposit's kernel: 0 (default) one lane per sample for the whole deposit, 1 the wavefront deposit (one
 * lane per sample walks its column, then the wavefront's lanes along the channels add each sample's line profile to
 * its runs). Both add the same values: the models are bit-identical. No counterpart in the reference. */
int bb_model_set_deposit(bb_model* m, int variant);

/* The azimuth plan of a set of sub-ring radii R0 (arcmin, HOST array, nsub): the ANALYTIC samples of every
 * sub-ring (naz = max(int(3 circumference / pixel + 0.5), 16), two radial samples, galmod.py:339-345) and the
 * cosine and sine of every azimuth from the reference's rotation recurrence (galmod.py:372-388), computed on
 * the host once per set of radii and copied to the current device. */
typedef struct bb_azplan bb_azplan;
int bb_azplan_create(const bb_model* m, const double* R0, int nsub, bb_azplan** out);
int bb_azplan_destroy(bb_azplan* p);

/* ANALYTIC GALMOD (_galmod_analytic, bbarolo/galmod.py:326-465) of a batch of B sub-ring tables whose radii are
 * those of the plan: sr (B, 16, nsub) float64 (rows as GalmodSetup.subrings, galmod.py:761-851), out
 * (B, by, bx, nz) float32 yxz over the box. accumulate = 0 writes the models, 1 adds them to out (the numba
 * kernels accumulate with +=). work: device scratch of bb_analytic_work_size(plan, B) bytes. */
int64_t bb_analytic_work_size(const bb_azplan* p, int B);
int bb_galmod_analytic(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub, int B,
                       void* work, int64_t work_bytes, float* out, int accumulate, void* stream);

/* bb_galmod_analytic with a model box per vector: vbox (B x 4 int32, DEVICE) holds [blo0, blo1, bhi0, bhi1) of vector
 * b in absolute pixels; its model is the reference's model of that box (DiskModel.array(rings, blo, bhi),
 * galmod.py:1242-1271: samples outside it are dropped, its vertical samples reach as far as that box needs) laid out
 * in out (B, by, bx, nz) over box, zero outside it. So models of different boxes share one launch (the 3DFIT ring fits
 * of all the rings). */
int bb_galmod_analytic_boxes(const bb_model* m, const bb_azplan* p, bb_box box, const int32_t* vbox, const double* sr,
                             int nsub, int B, void* work, int64_t work_bytes, float* out, int accumulate, void* stream);

/* Measurement, no counterpart in the reference: two hipEvent_t (or both NULL to stop) that the calling host
 * thread's ANALYTIC deposits record on their stream just before and just after the deposit kernel (the tile
 * pass of bb_galmod_analytic), so that a benchmark times that kernel alone. */
int bb_trace_deposit(void* ev_begin, void* ev_end);

/* 'Same' 2D beam convolution of every channel (FFTConvolver.__call__, bbarolo/smooth.py:145-165; _conv_sparse,
 * galmod.py:974-1037), values below 1e-12 set to zero, then the scale applied (DiskModel.convolve,
 * galmod.py:1273-1283). in/out (B, by, bx, nz) float32 yxz, distinct buffers. */
int bb_conv2d_same(const bb_model* m, const float* in, float* out, int B, int by, int bx, void* stream);

/* Masked Gaussian residual sums (_sumsq_residuals / _residuals_fracerr, bbarolo/bayesian/likelihood.py:22-77)
 * of B models (B, nvox) float32 against dvalid (nvox) float32 in the same voxel order: the data of the voxels
 * that enter the reference's sum (region, likelihood mask and blanks applied), NaN for the others.
 * eps2 <= 0: out_ss[b] = sum (d - m)^2 (out_lg may be NULL). eps2 > 0: out_ss[b] = sum r^2 / v and
 * out_lg[b] = sum_{m != 0} log(v / sigma2), v = sigma2 + eps2 m^2 (sigma2 where m = 0).
 * work: device float64 scratch of bb_residual_work_size(B, nvox) elements. Results are FP64, deterministic. */
int64_t bb_residual_work_size(int B, int64_t nvox);
int bb_residual_gauss(const float* model, const float* dvalid, int B, int64_t nvox, double sigma2, double eps2,
                      double* work, int64_t work_len, double* out_ss, double* out_lg, void* stream);

/* The beam convolution fused with the masked Gaussian residual: FFTConvolver (bbarolo/smooth.py:145-165, with the
 * 1e-12 cut and the flux scale of DiskModel.convolve, galmod.py:1273-1283) followed by _sumsq_residuals /
 * _residuals_fracerr (bbarolo/bayesian/likelihood.py:22-77), the convolved model compared with the data voxel by
 * voxel and never stored. dep (B, by, bx, nz) float32 yxz: the unconvolved models (device); dvalid as in
 * bb_residual_gauss. work: bb_conv_residual_work_size(m, B, by, bx) float64 elements. */
int64_t bb_conv_residual_work_size(const bb_model* m, int B, int by, int bx);
int bb_conv_residual_gauss(const bb_model* m, const float* dep, const float* dvalid, int B, int by, int bx,
                           double sigma2, double eps2, double* work, int64_t work_len, double* out_ss, double* out_lg,
                           void* stream);
/* bb_conv_residual_gauss without a work buffer and without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops). Its partial sums live in a
 * static device array of bb_conv_residual_noalloc_max() doubles (8 MiB), one per device, held until the process ends;
 * a call needing more (B x blocks x 2, see bb_conv_residual_work_size) is rejected. Calls that overlap on one device
 * share that array and race. Same results as bb_conv_residual_gauss, bit for bit. */
int64_t bb_conv_residual_noalloc_max(void);
int bb_conv_residual_gauss_noalloc(const bb_model* m, const float* dep, const float* dvalid, int B, int by, int bx,
                                   double sigma2, double eps2, double* out_ss, double* out_lg, void* stream);

/* Monte-Carlo GALMOD (_galmod_mc, bbarolo/galmod.py:92-171) of B sub-ring tables sr (B, 16, nsub) with a
 * counter-based stream: the draws of cloud ic of sub-ring s are Philox4x32-10 of counters (ic, block) under the
 * key (seed, s) (the reference's xoroshiro128+ / ziggurat stream is sequential), so the model of a box is the crop
 * of the full model and every vector of a batch has the same draws (common random numbers, as ISEED). nvpar:
 * sub-clouds per cloud (NV, or the channel count; 1 where VDISP = 0). out (B, by, bx, nz) float32 yxz, written
 * (accumulate = 0) or added to; counts (optional, (B, by, bx, nz) uint64): the integer sums of the sub-clouds in
 * units of q (the model is count x q x cd2i). max_clouds bounds the clouds of one vector inside the box; if it
 * is exceeded *overflow (device int) is set and the model is incomplete. work: bb_mc_work_size bytes. */
int64_t bb_mc_work_size(const bb_model* m, bb_box box, int nsub, int B, int64_t max_clouds);
int bb_galmod_mc(const bb_model* m, bb_box box, const double* sr, int nsub, int B, int nvpar, uint32_t seed,
                 int64_t max_clouds, void* work, int64_t work_bytes, float* out, uint64_t* counts, int* overflow,
                 int accumulate, void* stream);

/* bb_galmod_mc with a model box per vector (vbox as bb_galmod_analytic_boxes): the clouds of vector b outside its box
 * are dropped. */
int bb_galmod_mc_boxes(const bb_model* m, bb_box box, const int32_t* vbox, const double* sr, int nsub, int B, int nvpar,
                       uint32_t seed, int64_t max_clouds, void* work, int64_t work_bytes, float* out, uint64_t* counts,
                       int* overflow, int accumulate, void* stream);

/* The 3DFIT residual of ring models (_norm + Galfit._norm_value / _support_args, bbarolo/galfit.py:39-188,
 * 685-733): B trial models (B, by, bx, nz) float32 yxz (convolved) in the box [blo, blo + (bx, by)) of the data
 * (ny, nx, nz) float32 yxz and mask (ny, nx, nz) uint8 (the transposed cube and mask of Galfit, data_T/mask_T);
 * ring (B, 6) float64: x0, y0 (box-relative pixels), phi, inc (degrees), r1, r2 (pixels) of each trial;
 * chan_noise (nz) float64. Optional support of ANALYTIC models: pre (B, by, bx, nz) float32 (unconvolved model in
 * units of sub-clouds), ilo/ihi (ky) int32 footprint columns of the kernel rows, thr (nz) float64. value (B)
 * float64 (+inf without voxels), nblank / ntot (B) int64 (may be NULL). All arrays on the device; the model is not
 * scaled in place (the reference's in-place normalisation is internal to one evaluation).
 * work: bb_norm_work_size(...) float64 elements. */
typedef struct bb_norm_par {
    int32_t side;        /* 0 both, 1 receding, 2 approaching (SIDE) */
    int32_t ftype;       /* 1 chi^2/sqrt(obs), 2 |m-o|, 3 |m-o|/(m+o), 4 (m-o)^2 (FTYPE) */
    int32_t normtype;    /* 0 LOCAL, 1 AZIM, 2 NONE (NORM) */
    int32_t foot;        /* 0 annulus, 1 strip along the major axis (|y| <= hmax) */
    double wpow;         /* weight |cos th|^wpow, |sin th|^|wpow| if negative (WFUNC) */
    double data_noise;   /* obs where the data are <= 0 */
    double bweight;      /* exponent of the blank penalty (BWEIGHT) */
    double hmax;
} bb_norm_par;
int64_t bb_norm_work_size(int B, int by, int bx, int nz, int with_support);
int bb_norm3dfit(const float* model, int B, int by, int bx, int nz, const float* data, const uint8_t* mask, int ny,
                 int nx, int blo0, int blo1, const double* ring, const double* chan_noise, const bb_norm_par* par,
                 const float* pre, const int32_t* ilo, const int32_t* ihi, int ky, const double* thr, double* work,
                 int64_t work_len, double* value, int64_t* nblank, int64_t* ntot, void* stream);

/* The LOCAL normalisation of a final 3DFIT model (bbarolo/galfit_out.py:128-135): out = model x (sum over the channels
 * of the data where mask is set) / (sum over the channels of the model), per pixel (0 where the model sums to 0).
 * model, data, out (by, bx, nz) float32 yxz, mask (by, bx, nz) uint8, all on the device (out may alias model). */
int bb_local_norm(const float* model, const float* data, const uint8_t* mask, int by, int bx, int nz, float* out,
                  void* stream);

/* Unit-density ring maps (DiskModel.ring_maps / _ring_maps_analytic, bbarolo/galmod.py:1317-1345, 468-536) in
 * two steps. bb_ring_maps_deposit writes the planes map0 .. map0+nw-1 of counts (B, nmap, by, bx) uint64: the
 * total-intensity maps, summed over the channels, of the ANALYTIC samples of the sub-ring tables sr (B, 16, nsub;
 * their density row is not read) with the radii of the plan. Sub-ring s adds to map map0 + widx[s][m] (m = 0, 1;
 * -1 or >= nw: none) with weight wval[s][m] (device arrays (nsub, 2) int32 / float64: the interpolation of the
 * ring densities, DiskModel._ring_weights, galmod.py:1285-1314). Fixed point: a contribution x adds
 * round(x / qunit), so the sums are exact and deterministic; the caller picks qunit so that no sum exceeds 2^64
 * units (every contribution is <= fluxsc = 2 pi R0 radsep / (2 naz) of its sub-ring).
 * work: bb_ring_maps_work_size(plan, B) bytes. */
int64_t bb_ring_maps_work_size(const bb_azplan* p, int B);
int bb_ring_maps_deposit(const bb_model* m, const bb_azplan* p, bb_box box, const double* sr, int nsub, int B,
                         const int32_t* widx, const double* wval, int nmap, int map0, int nw, double qunit,
                         void* work, int64_t work_bytes, uint64_t* counts, void* stream);
/* maps (P, by, bx) float64 = counts x unit, convolved with the beam ('same', without the 1e-12 cut of the
 * models) and multiplied by the model's scale when convolve != 0 (galmod.py:1342-1345). unit = qunit x
 * mean(cd2i) x sum(relint). */
int bb_ring_maps_finish(const bb_model* m, const uint64_t* counts, int P, int by, int bx, double unit, int convolve,
                        double* maps, void* stream);

/* The deconvolved density profile of the rings (profile_system / solve_profile, bbarolo/profile.py:33-138; the
 * per-evaluation density of Bayesian3DFIT, bayesian/density.py:13-88, likelihood.py:595-596). The per-fit
 * constants: the pixels of the map box (wpix = 1 / map noise where usable, else 0; ypix = data map x wpix), the
 * second-difference rows of the curvature prior on the radii of the rings and of the exterior annulus (3
 * coefficients and columns per row, profile.py:81-99) and which of them cross two radii closer than a quarter
 * of the ring spacing (near). */
typedef struct bb_profile_desc {
    int32_t nr;          /* rings; the system has n = nr + 1 columns (the rings, then the exterior annulus) */
    int32_t npix;        /* pixels of the map box */
    int32_t nok;         /* usable pixels (wpix > 0) */
    int32_t pad0;
    double smooth;       /* width of the curvature prior (like_kwargs densprof_smooth; 0: none) */
    double corr;         /* pixels per independent noise value of the map (1 / likefactor2d) */
    double floor;        /* lowest prior weight as a fraction of the relative prior's (DENSPROF_FLOOR) */
    double norm;         /* sum of the squared row coefficients, spacings >= dmin (profile.py:141-150) */
    double dmed;         /* median spacing of the ring radii */
} bb_profile_desc;
typedef struct bb_profile bb_profile;
int bb_profile_create(const bb_profile_desc* d, const double* wpix, const double* ypix, const double* lcoef,
                      const int32_t* lcol, const int32_t* near, bb_profile** out);
int bb_profile_destroy(bb_profile* p);
/* x (B, n) float64 = solve_profile of the maps (B, n, by, bx) of bb_ring_maps_finish: the two non-negative
 * solutions (relative prior, then the density-scaled prior) on the normal equations of the weighted system; NaN
 * where the active set did not converge. work: bb_profile_work_size(p, B) bytes. */
int64_t bb_profile_work_size(const bb_profile* p, int B);
int bb_profile_solve(const bb_profile* p, const double* maps, int B, void* work, int64_t work_bytes, double* x,
                     void* stream);
/* The densities into the model: dens (B, nr) float32 = x[:, :nr] (the float32 ring table,
 * bayesian/rings.py:54-61; may be NULL) and row 5 of sr (B, 16, nsub) = sum_m wval[s][m] dens[widx[s][m]]. */
int bb_profile_set_dens(const double* x, int B, int n, int nr, const int32_t* widx, const double* wval, int nsub,
                        double* sr, float* dens, void* stream);

/* The likelihood surrogate: a network din -> hidden (SiLU) x nhidden -> 1 with BF16 weights and activations and
 * FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32 master parameters and full-batch Adam
 * (beta 0.9 / 0.999, eps 1e-8), trained on the device on the (parameters, log-likelihood) pairs of a sampler. It
 * stands where the regression network of nautilus's neural bounds stands in the reference
 * (bbarolo/bayesian/nautilus/mlp.py:17-222, fitted in neural.py:60-127 and used by bounds/neural.py): a learned
 * model of the likelihood that decides which proposals to evaluate. The reference's network is a scikit-learn-like
 * NumPy MLP on the CPU; this one has no counterpart at the C level. din <= 32, hidden a multiple of 32 (<= 512),
 * nmax: the most rows of any call. Parameters (bb_mlp_param_count floats): per hidden layer l the weights
 * (hidden x K_l, K_0 = 32 with the inputs zero-padded, K_l = hidden) then the biases (hidden); then the output
 * weights (hidden) and bias (1). x (rows, din), t, y (rows) and grad are DEVICE float32 arrays; the network's own
 * buffers live on the device current at bb_mlp_create, which must be current at every call. */
typedef struct bb_mlp bb_mlp;
int bb_mlp_create(int din, int hidden, int nhidden, int nmax, bb_mlp** out);
int bb_mlp_destroy(bb_mlp* net);
int64_t bb_mlp_param_count(const bb_mlp* net);
/* Host float32 parameters in / out; set also resets the Adam moments and step count. */
int bb_mlp_set_params(bb_mlp* net, const float* host_params);
int bb_mlp_get_params(const bb_mlp* net, float* host_params);
/* y = net(x) for rows rows. */
int bb_mlp_forward(bb_mlp* net, const float* x, int rows, float* y, void* stream);
/* grad = d/dparams of 0.5 mean (net(x) - t)^2 through the BF16 copies of the parameters; loss (device, 1 float)
 * may be NULL. */
int bb_mlp_grad(bb_mlp* net, const float* x, const float* t, int rows, float* grad, float* loss, void* stream);
/* steps full-batch Adam steps of learning rate lr on (x, t); loss (device, steps floats: the loss before each
 * step) may be NULL. Deterministic (fixed-order reductions, no atomics). */
int bb_mlp_train(bb_mlp* net, const float* x, const float* t, int rows, int steps, float lr, float* loss,
                 void* stream);
/* Random walks on the network's region {net(x) >= thr} for a sampler's proposals (the screened random walks of the
 * nested sampler; the reference samples with nautilus's neural bounds instead, bbarolo/bayesian/nautilus/bounds/
 * neural.py, so this call has no counterpart there). rows walkers x (rows, din) float32 take `steps` steps: the
 * proposal x + sc[s] zl[s] (zl (steps, rows, din): the Gaussian steps, already multiplied by the Cholesky factor)
 * is kept if it lies in the unit cube and net((proposal - xc) / xs) >= thr[0]; x is updated in place and moved
 * (rows, uint8) marks the walkers that moved. sc (steps + 1 floats): sc[0] the first step's scale; the call writes
 * sc[s] = clamp(sc[s-1] exp(acceptance of step s-1 - 0.3), 1e-4, 3) for s = 1..steps. All DEVICE arrays; the walk
 * is one captured HIP graph, replayed while the arguments are the same. steps <= 4096. */
int bb_mlp_walk(bb_mlp* net, float* x, const float* zl, int rows, int steps, const float* thr, const float* xc,
                const float* xs, float* sc, uint8_t* moved, void* stream);

/* ---- per-fit set-up: the likelihood mask -----------------------------------------------------------------------
 * Replaces dilate_mask (bbarolo/bayesian/noise.py:18-30; called by _setup_likemask, likelihood.py:160-175, with the
 * radius nbeam x beam FWHM in pixels): in every channel, the pixels within a Euclidean distance radius of a mask
 * pixel of that channel, then the maximum over +-nchan channels (zeros outside the cube). mask, tmp and out are
 * DEVICE (nz, ny, nx) uint8 arrays (mask: nonzero = in the mask; out: 0 / 1); tmp is needed when radius >= 1 and
 * nchan > 0 (else it may be NULL); out must not alias mask or tmp. floor(radius) <= bb_dilate_rmax(). */
int bb_dilate_mask(const uint8_t* mask, int nz, int ny, int nx, double radius, int nchan, uint8_t* tmp,
                   uint8_t* out, void* stream);
int bb_dilate_rmax(void);

/* ---- per-fit set-up: the noise ------------------------------------------------------------------------------------
 * The statistics of Bayesian3DFIT._setup_noise (the noise rms, bbarolo/bayesian/likelihood.py:92-111) and of
 * noise_correlation (bbarolo/bayesian/noise.py:33-186, called by _set_likefactor, likelihood.py:262-268). Cubes are
 * DEVICE (nz, ny, nx) arrays, data float32; every call is asynchronous on stream.
 *
 * The voxels used: finite data (and non-zero when zero_blank: the blanks of likelihood.py:496-499) outside emis
 * (may be NULL: nothing excluded; noise.py:75-90 with the dilated mask, likelihood.py:105-108 with the mask). used
 * (uint8 0 / 1) per voxel; chan_mask and chan_used (uint64, nz) the voxels of mask (may be NULL) and the used voxels
 * per channel; pix_emis (uint8, ny x nx) whether any channel of the pixel is in emis and pix_used (uint32) its used
 * voxels (both may be NULL). */
int bb_noise_flags(const float* data, const uint8_t* mask, const uint8_t* emis, int nz, int ny, int nx, int zero_blank,
                   uint8_t* used, uint64_t* chan_mask, uint64_t* chan_used, uint8_t* pix_emis, uint32_t* pix_used,
                   void* stream);
/* numpy.median of the float32 values x[i] with sel[i] != 0 (i < n), or of |x[i] - center[0]| (float32) when center
 * (DEVICE, 1 float) is not NULL (likelihood.py:109; noise.py:93-96): out (DEVICE float) the middle value or the float32
 * mean of the two middle values, NaN when nothing is selected; count (DEVICE int64, may be NULL) the number selected.
 * Exact (radix selection). work: bb_masked_median_work_size() DEVICE bytes. */
int64_t bb_masked_median_work_size(void);
int bb_masked_median(const float* x, const uint8_t* sel, int64_t n, const float* center, void* work, float* out,
                     int64_t* count, void* stream);
/* The standard deviation of the selected values (ddof 0), the fallback of likelihood.py:110-111 (numpy.nanstd): FP64
 * sums rounded to float32; count (DEVICE int64) the number selected (as bb_masked_median returns it). */
int64_t bb_masked_std_work_size(void);
int bb_masked_std(const float* x, const uint8_t* sel, int64_t n, const int64_t* count, void* work, float* out,
                  void* stream);
/* The subsample of noise.py:94-95: out[i] = 1 for the flagged voxels (flags[i] != 0, nflag of them) whose rank among
 * the flagged voxels (C order) is one of ranks (DEVICE int64, nr of them, from numpy's
 * default_rng(0).choice(nflag, 2000000, replace=False)). work: bb_flag_subsample_work_size(n, nflag) DEVICE bytes. */
int64_t bb_flag_subsample_work_size(int64_t n, int64_t nflag);
int bb_flag_subsample(const uint8_t* flags, int64_t n, int64_t nflag, const int64_t* ranks, int64_t nr, void* work,
                      int64_t work_bytes, uint8_t* out, void* stream);
/* The binary dilation of m by a box of half-sizes (rz, ry, rx) (_dilate, noise.py:8-15; zero outside the cube): out
 * (uint8 0 / 1); tmp (same size) is needed with more than one non-zero half-size. */
int bb_box_dilate(const uint8_t* m, int nz, int ny, int nx, int rz, int ry, int rx, uint8_t* tmp, uint8_t* out,
                  void* stream);
/* The maps of the likelihood voxels (likemask (may be NULL: all), not blank: finite and non-zero) of _setup_likemask
 * (likelihood.py:176-185): per pixel dm2 the sum of their squared data, nvox their number, mom0 the sum of their data
 * (FP64, summed sequentially within chunks of 16 channels and the chunks in order, as the reference's numpy sums:
 * the same bits); and dvalid, the data of the likelihood in box as the context holds it, (by, bx, nz) yxz with NaN
 * where the voxel is not used or its pixel is outside region (may be NULL). DEVICE arrays; any output may be NULL. */
int bb_like_maps(const float* data, const uint8_t* likemask, const uint8_t* region, int nz, int ny, int nx, bb_box box,
                 float* dvalid, double* dm2, int32_t* nvox, double* mom0, void* stream);
/* The spatial autocorrelation sums of noise.py:103-119 without FFTs: for the lags (dy, dx) of the half-plane dy > 0
 * or (dy = 0, 0 <= dx <= L), in the order (0, 0..L), (1, -L..L), ..., (L, -L..L) (2L^2 + 2L + 1 of them),
 *     c[l] = sum over the channels with chan_use[z] and the pixels p of w(z, p) w(z, p + lag) (data - med)(data' - med)
 *     wsum[l] = the number of such pairs,
 * FP64 (DEVICE arrays; the other half-plane has the same sums); w the used voxels. work: bb_lag_xy_work_size doubles.
 * 1 <= L <= bb_lag_xy_max(). */
int bb_lag_xy_max(void);
int64_t bb_lag_xy_work_size(int nz, int ny, int nx, int L);
int bb_lag_xy(const float* data, const uint8_t* w, const uint8_t* chan_use, int nz, int ny, int nx, int L, double med,
              double* work, int64_t work_len, double* c, double* wsum, void* stream);
/* The spectral sums of noise.py:136-173: for t = 0..K (K <= 7), c[t] = sum over the pixels with pix_use (NULL: all)
 * and the pairs of channels (z, z + t) of the same run (runid[z] >= 0, DEVICE int32, nz) of w w' (data - med)
 * (data' - med), wsum[t] the number of pairs. work: bb_lag_z_work_size doubles. */
int64_t bb_lag_z_work_size(int ny, int nx, int K);
int bb_lag_z(const float* data, const uint8_t* w, const uint8_t* pix_use, const int32_t* runid, int nz, int ny, int nx,
             int K, double med, double* work, int64_t work_len, double* c, double* wsum, void* stream);

/* ---- the fused likelihood of a fit --------------------------------------------------------------------------------
 * Bayesian3DFIT._log_likelihood (bbarolo/bayesian/likelihood.py:567-619) of batches of parameter vectors, run by the
 * library from theta to log L: the seam that nautilus's vectorized sampler calls (bayesian/sampling.py:235-241). Per
 * vector: the free values into the float32 ring table (_update_rings, bayesian/fitter.py:1109-1136), zero likelihood
 * for a negative positive-only value (likelihood.py:537-559), the sub-ring tables (GalmodSetup.subrings,
 * galmod.py:761-851) built on the device, the deconvolved density profile when DENS is not fitted and the geometry is
 * free (bayesian/density.py:13-88), the model in its own box (likelihood.py:429-435), the beam convolution fused with
 * the masked residual, the squared data of the region outside the box (likelihood.py:437-440), and
 *     log L = -likefactor [(S + S_out) / (2 sigma^2) + nvox ln(2 pi sigma^2) / 2]        (+ the fracerr terms).
 * Ring keys (the rows of the ring table, in the reference's units: km/s, arcsec, degrees). SIGMA, FRACERR and the
 * edges are single values of theta, not ring values: RADMAX the outer edge of the disk (arcsec, used as float64 as
 * _update_rings sets rings.edge, bayesian/fitter.py:1112-1116; zero likelihood when negative), RADMAXA / RADMAXR the
 * edges of the approaching / receding halves (fitter.py:1117-1120; negative: no edge, galmod.py:853-863). */
#define BB_NKEY 21
enum {
    BB_K_VROT = 0, BB_K_VRAD, BB_K_VVERT, BB_K_VDISP, BB_K_DENS, BB_K_Z0, BB_K_DVDZ, BB_K_ZCYL, BB_K_INC, BB_K_PHI,
    BB_K_XPOS, BB_K_YPOS, BB_K_VSYS, BB_K_FTHICK, BB_K_Z0THICK, BB_K_VDISPTHICK, BB_K_SIGMA, BB_K_FRACERR,
    BB_K_RADMAX, BB_K_RADMAXA, BB_K_RADMAXR
};

/* The sub-ring plan of a set of fixed radii (the part of GalmodSetup.subrings that does not depend on the ring values,
 * galmod.py:805-832, with the ring convention of model_rings, galmod.py:1066-1119): HOST arrays. Model ring m takes the
 * value of key k from fit ring ridx[m] when bit k of from_fit is set, else base[k][m]. Sub-ring s interpolates between
 * model rings idx[s]-1 and idx[s] at dr[s] arcmin from the first (cst[s]: the values of model ring 0); rads[s] is its
 * float32-accumulated radius. densw (nsub, may be NULL) multiplies the density (the disk edge); side (2, nsub_total)
 * are the weights of the receding and approaching halves; thick: a second, thick layer (nsub_total = 2 nsub). */
typedef struct bb_ring_plan {
    int32_t nmr, nsub, thick;
    uint32_t from_fit;
    const int32_t* ridx;
    const double* rad;
    const double* base;
    const int32_t* idx;
    const double* dr;
    const uint8_t* cst;
    const double* rads;
    const double* densw;
    const double* side;
} bb_ring_plan;

typedef struct bb_like_desc {
    int32_t nr;          /* rings of the fit */
    int32_t ndim;        /* length of a parameter vector */
    int32_t chunk;       /* most vectors of one launch (device buffers are sized for it) */
    int32_t mc;          /* 1: the Monte-Carlo model (GALMODTYPE MC), 0: ANALYTIC */
    int32_t nvpar;       /* MC: sub-clouds per cloud */
    uint32_t seed;       /* MC: the seed of the counter-based stream */
    int32_t box_free;    /* 1: each vector has its own model box (XPOS, YPOS or Z0 free) */
    int32_t halves;      /* 1: the two halves of a launch on two streams */
    bb_box likebox;      /* the box of the likelihood region (and of the profile's maps) */
    bb_box fixbox;       /* box_free = 0: the model box of every vector (within likebox) */
    double sigma;        /* noise per voxel (theta's SIGMA when free) */
    double likefactor;   /* 1 / correlation volume of the noise */
    double eps;          /* fixed fractional model error (theta's FRACERR when free; 0: none) */
    double ext;          /* arcsec: max(last radius + RADSEP / 2, outer edge of the disk) (galmod.py:1231-1233) */
    double pixas0, pixas1;   /* |CDELT1|, |CDELT2| in arcsec */
    int32_t nx, ny;      /* the cube */
    int64_t mc_cap;      /* MC: first bound of the clouds of one vector in its box (grown when exceeded) */
    /* The edges of the disk when RADMAX, RADMAXA or RADMAXR is free (the model's density weights, galmod.py:833-851,
     * and the profile's exterior annulus per vector, profile.py:45-64); the fixed ones are in the plans. */
    double edgewidth;    /* RADMAX: width of a Gaussian edge (arcsec; 0: a sharp edge, like_kwargs radmax 'edge') */
    double oedge;        /* arcsec: the outer edge of the disk of the rings (outer_edge, with the ring convention) */
    double edge0;        /* arcsec: the fixed edge of the disk when RADMAX is not free (NaN: none) */
    double edgea0, edger0;   /* arcsec: the fixed edges of the approaching / receding halves (negative: none) */
    double annw;         /* arcsec: the width of the exterior annulus */
    int32_t ann_cap;     /* sub-rings of one layer of a vector's exterior annulus (capacity) */
    int32_t empty;       /* EMPTY: the disk starts at its first radius */
} bb_like_desc;

/* The deconvolved profile of every vector (NULL: DENS fitted, or the profile of the initial rings in the ring table):
 * the profile object's description and HOST arrays as bb_profile_create, the ring weights of the model sub-rings and
 * of the exterior annulus (as bb_ring_maps_deposit) and the fixed point of the maps. */
typedef struct bb_like_profile {
    bb_profile_desc d;
    const double* wpix;
    const double* ypix;
    const double* lcoef;
    const int32_t* lcol;
    const int32_t* near;
    const int32_t* widx;
    const double* wval;
    const int32_t* widx_e;
    const double* wval_e;
    double qunit, unit;
    const double* radii; /* (nr) arcsec: the radii of the rings (the prior's rows per vector when an edge is free) */
    double dmin;         /* a quarter of the smallest spacing of the radii (profile.py:141-150; 0 with one ring) */
} bb_like_profile;

/* The context of a fit on the current device (Bayesian3DFIT's set-up, likelihood.py:83-440): m the model (its grid,
 * beam and vertical-sample capacity), d, data (nz, ny, nx) float32, likemask (nz, ny, nx) uint8 (NULL: all), region
 * (ny, nx) uint8 (NULL: all); rings (BB_NKEY, nr) the float32 ring table; free_off[k] the offset of key k in theta
 * (-1: not free), free_ring[k] 1 for one value per ring; plan the model's sub-ring plan, annulus that of the exterior
 * annulus (needed with prof; with an edge free only its values and ring weights are used: its radii follow each
 * vector's edge). All HOST arrays, copied. A voxel enters the likelihood if it is in the likemask and the
 * region and its data are finite and non-zero. */
typedef struct bb_ctx bb_ctx;
int bb_ctx_create(const bb_model* m, const bb_like_desc* d, const float* data, const uint8_t* likemask,
                  const uint8_t* region, const double* rings, const int32_t* free_off, const int32_t* free_ring,
                  const bb_ring_plan* plan, const bb_ring_plan* annulus, const bb_like_profile* prof, bb_ctx** out);
int bb_ctx_destroy(bb_ctx* c);
/* The log-likelihoods of B (<= chunk) parameter vectors theta (HOST, B x ndim), launched on stream (returns at once)
 * and collected: loglike (HOST, B; -inf where the likelihood is zero: a negative positive-only value, FTHICK > 1,
 * SIGMA <= 0, a profile that did not converge), dens (HOST, B x nr float32 densities of the profile, may be NULL).
 * One launch at a time per context. */
int bb_loglike_launch(bb_ctx* c, const double* theta, int B, void* stream);
int bb_loglike_collect(bb_ctx* c, double* loglike, float* dens);
/* launch + collect over chunks of any B */
int bb_loglike_batch(bb_ctx* c, const double* theta, int B, double* loglike, float* dens, void* stream);
/* The sub-ring tables of B (<= chunk) vectors as the context builds them on the device: sr (B, 16, nsub_total) and
 * sr_ext (B, 16, nsub of the annulus, or ann_cap per layer with an edge free: the sub-rings past a vector's last
 * have zero side weights; may be NULL), HOST arrays; synchronous. */
int bb_ctx_tables(bb_ctx* c, const double* theta, int B, double* sr, double* sr_ext);
/* nvox and the squared data of the region, the deposit box [blo0, blo1, bhi0, bhi1), the vectors whose profile did not
 * converge and the MC launches repeated with larger buffers (any may be NULL). */
int bb_ctx_info(const bb_ctx* c, double* nvox, double* ss_total, int32_t* deposit_box, int64_t* nnls_fail,
                int64_t* mc_retries);
/* 0: launches run on the caller's stream only (no halves on two streams). */
int bb_ctx_set_halves(bb_ctx* c, int halves);
/* 1: the device work of a launch runs as a captured HIP graph per batch size (the first launch of a size eagerly, the
 * second captured, later ones replayed with the new vectors staged in the pinned buffer the graph copies from); 0:
 * eager launches. bb_ctx_graph_replays counts the replays. */
int bb_ctx_set_graphs(bb_ctx* c, int on);
int64_t bb_ctx_graph_replays(const bb_ctx* c);
/* Measurement, no counterpart in the reference: 5 hipEvent_t that eager launches record on the stream of their first
 * part before the ring maps, after them, after the profile, after the deposit and after the residual (NULL: none). */
int bb_ctx_trace(bb_ctx* c, void* const* ev);

#ifdef __cplusplus
}
#endif
#endif
