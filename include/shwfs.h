/* shwfs.h - C ABI of the MI355X (gfx950) Shack-Hartmann WFS hot path.
 *
 * Plain C, plain pointers and sizes.  Every device pointer is owned by the
 * caller (the C++ sutra/carma layer, or any FFI binding); the library never
 * frees what it is given.  Every entry point returns SHWFS_OK (0) or a
 * negative SHWFS_ERR_* code; the message of the last error on the calling
 * host thread is available from shwfs_last_error().  Nothing here throws or
 * aborts.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Array layout is the reference's: column-major ("x fast"), e.g. a phase
 * screen of side npup holds phase(x, y) at x + y*npup; the spot cube holds
 * bincube(px, py, s) at px + py*npix + s*npix*npix.
 */
#ifndef SHWFS_H
#define SHWFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SHWFS_OK = 0,
  SHWFS_ERR_ARG = -1,         /* null pointer / inconsistent sizes */
  SHWFS_ERR_UNSUPPORTED = -2, /* shape outside what the kernels implement */
  SHWFS_ERR_HIP = -3,         /* a HIP runtime call failed */
  SHWFS_ERR_STATE = -4,       /* call order violated (e.g. image before tables) */
  SHWFS_ERR_NOMEM = -5
};

/* Computation method of the spot image (shwfs_frame_args.method). */
enum {
  SHWFS_METHOD_FFT_FP32 = 0,   /* plain FP32: in-register radix-2 FFT */
  SHWFS_METHOD_DFT_BF16X3 = 1, /* pruned DFT on MFMA, split-BF16x3, FP32 accumulate */
  SHWFS_METHOD_DFT_FP16 = 2    /* pruned DFT on MFMA, FP16 with per-subap scale */
};

/* Centroider kinds (shwfs_centroid_args.kind). */
enum {
  SHWFS_CENTROID_NONE = 0,
  SHWFS_CENTROID_COG = 1,
  SHWFS_CENTROID_TCOG = 2,
  SHWFS_CENTROID_WCOG = 3,
  SHWFS_CENTROID_MLP = 4  /* learned centroider (shwfs_mlp_*), fused after binning */
};

/* Geometry of one SH WFS: sizes + device tables (all owned by the caller).
 * Filled by the C++ layer after SutraWfsSH::load_arrays
 * (reference: libsutra/src.cpp/sutra_wfs_sh.cpp:308-331). */
typedef struct shwfs_sh_geom {
  int32_t nvalid;  /* valid sub-apertures */
  int32_t nphase;  /* phase points per sub-aperture side (pdiam) */
  int32_t npup;    /* side of the phase screen / pupil arrays */
  int32_t nfft;    /* FFT support side (8, 16, 32 or 64) */
  int32_t ntot;    /* HR grid side: nfft, or (FFT method) a power of two in (nfft, 64]
                      onto which the |E|^2 image is placed through d_hrmap */
  int32_t npix;    /* detector pixels per sub-aperture side */
  int32_t nrebin;  /* HR pixels per detector pixel side */
  int32_t nxsub;   /* sub-apertures across the pupil; binimg side = npix*nxsub */
  const int32_t* d_istart;     /* [nvalid] validpuppixx (x of the subap corner in the screen) */
  const int32_t* d_jstart;     /* [nvalid] validpuppixy */
  const int32_t* d_validsubsx; /* [nvalid] x of the spot corner in binimg (already *npix) */
  const int32_t* d_validsubsy; /* [nvalid] */
  const float* d_offsets;      /* [nphase^2] half-pixel tip-tilt phase (halfxy) */
  const float* d_fluxPerSub;   /* [nvalid] */
  const int32_t* d_binmap;     /* [nrebin^2 * npix^2] HR flat index per binned pixel */
  int32_t binmap_block;        /* 1: d_binmap is the block form found by
                                  shwfs_binmap_block_origin (kernels compute the
                                  indices instead of reading the table) */
  int32_t bin_origin;          /* origin returned by shwfs_binmap_block_origin */
  /* enlarged HR grid and kernel convolution (SutraWfsSH::comp_generic,
   * libsutra/src.cpp/sutra_wfs_sh.cpp:409-481, 514-525); FFT method only */
  const int32_t* d_hrmap;      /* [nfft^2] ntot-grid flat index of HR element u + v*nfft
                                  (indexfill); read when ntot != nfft */
  const float* d_ftkernel;     /* [2 * ntot^2] complex (re, im) FFT of the convolution
                                  kernel / ntot^2; read when kernconv */
  int32_t kernconv;            /* 1: HR image convolved with the kernel before binning */
  /* LGS elongation (SutraWfsSH::comp_generic with lgs, libsutra/src.cpp/
   * sutra_wfs_sh.cpp:420-451, 483-512); FFT method, ntot in {16, 32, 64} */
  const float* d_ftlgskern;    /* [2 * ntot^2 * nvalid] complex: per sub-aperture FFT of its
                                  elongated spot kernel (shwfs_lgs_kernels), or NULL (NGS).
                                  The HR image is convolved with it before binning; with
                                  ntot != nfft and kernconv, then with ftkernel as well (with
                                  ntot == nfft, kernconv is not applied, as the reference) */
  /* Detector-image pixels outside every spot window (index x + y * npix * nxsub),
   * or NULL.  The reference draws the readout background on the whole binimg and
   * then overwrites the spot windows (SutraWfsSH::fill_binimage, sutra_wfs_sh.cpp:
   * 599-621); with this list a noisy frame of the FFT, DFT or fused learned-
   * centroider spot kernel draws it on these pixels only, by extra workgroups of
   * the spot launch (the same draws, element = image index, Philox domain 2).
   * NULL, and the extended-grid and traced spot kernels: a launch of its own draws
   * it on every pixel ahead of the spots. */
  const int32_t* d_bgpix;      /* [nbgpix], each index < (npix * nxsub)^2 */
  int32_t nbgpix;
} shwfs_sh_geom;

/* A centroider applied to a detector image (or fused after the spot stage).
 * Reference kernels: libsutra/src.cu/sutra_centroider_{cog,wcog,tcog}.cu. */
typedef struct shwfs_centroid_args {
  int32_t kind;            /* SHWFS_CENTROID_* */
  int32_t nvalid;
  int32_t npix;
  int32_t size;            /* side of the image */
  const int32_t* d_validx; /* [nvalid] spot corner x in the image */
  const int32_t* d_validy; /* [nvalid] */
  float offset;            /* pixel offset subtracted (npix/2 - 0.5) */
  float scale;             /* arcsec per pixel */
  float threshold;         /* TCOG / WCOG */
  const float* d_weights;  /* WCOG: [npix^2 * nvalid] */
  const float* d_ref;      /* [2*nvalid] reference slopes (may be NULL = zeros) */
  float* d_slopes;         /* [2*nvalid] */
  float* d_intensities;    /* [nvalid] centroider flux */
  int32_t slope_order;     /* 0 untied (x..., y...), 1 interlaced (x, y, x, y...) */
  int32_t fp64_moments;    /* 1: flux and first moments summed in float64 */
  /* SHWFS_CENTROID_MLP only (fused into the spot kernel: the FFT method with
   * 8x8-pixel spots at nfft 16 or 32, or an MFMA method at nfft 32 with
   * 16-point windows and 8x8 pixels binned 4x4 about the grid centre): as
   * shwfs_mlp_centroids over the spot cube, in the same launch as the spots.
   * Other kinds ignore these. */
  const float* d_mlp_params; /* [SHWFS_MLP_NPARAM] */
  float mlp_x_scale;         /* input scale of the spot pixels */
  float mlp_y_scale;         /* network output (pixels) -> slope units */
  float mlp_flux_lo;         /* spots with flux outside [lo, hi] take their CoG */
  float mlp_flux_hi;
  int32_t* d_mlp_fallback;   /* [nvalid] 1 where the CoG was used, or NULL */
} shwfs_centroid_args;

/* One layer seen by a source: the phase screen of an atmospheric layer (or
 * a DM shape) and where the source's footprint lies in it.  Reference: the
 * (type, index) -> (xoff, yoff) maps of SutraSource filled by add_layer
 * (libsutra/src.cpp/sutra_source.cpp:206-213), read by SutraSource::raytrace
 * (222-267) with the layer's accumulated shift (accumx, accumy) added. */
#define SHWFS_MAX_LAYERS 16
typedef struct shwfs_layer {
  const float* d_screen; /* [dim * dim] layer phase (x fastest), micrometres */
  int32_t dim;           /* side of the layer screen (>= 2) */
  float xoff, yoff;      /* footprint offset in the layer, pixels (accum included) */
  float G;               /* magnification (1: none) */
  float thetaML;         /* rotation, radians */
  float dx, dy;          /* extra shift, pixels */
  float delta;           /* cone compression 1 - altitude / LGS height (1: NGS) */
} shwfs_layer;

/* One frame of one SH WFS. */
typedef struct shwfs_frame_args {
  const float* d_phase;      /* [npup^2] phase screen in micrometres */
  const float* d_pupil;      /* [npup^2] amplitude mask */
  float scale;               /* 2*pi / lambda[um] */
  float nphot;               /* photons per fully illuminated sub-aperture */
  float noise;               /* spot noise: < 0 none, 0 Poisson, > 0 Poisson + readout sigma */
  float noise_background;    /* > 0: binimg background = sigma*N(0,1) before the spots */
  uint64_t seed;             /* Philox key */
  uint32_t call;             /* Philox call counter (advance once per frame) */
  int32_t method;            /* SHWFS_METHOD_* */
  int32_t wg_size;           /* threads per workgroup of the spot kernel: 0 (default), 64, 128, 256 */
  float* d_bincube;          /* [npix^2 * nvalid] */
  float* d_binimg;           /* [(npix*nxsub)^2] or NULL to skip the image */
  float* d_intensities;      /* [nvalid] raw per-subap flux before normalisation */
  const shwfs_centroid_args* centroid; /* NULL or fused centroider (validx/size ignored) */
  const int32_t* d_pupil_unit; /* NULL, or [nvalid] 1 where the pupil window of the
                                  sub-aperture is all ones (shwfs_sh_unit_pupil of
                                  this d_pupil): the kernels skip reading it */
  const shwfs_layer* layers;   /* host array [nlayers], or NULL */
  int32_t nlayers;             /* > 0: the phase is the raytrace of these layers
                                  over an npup x npup source screen, computed in
                                  the field stage (d_phase is not read and may be
                                  NULL): the values shwfs_raytrace would write
                                  there, bit for bit.  FFT method only. */
} shwfs_frame_args;

/* ---- spot image (SutraWfsSH::comp_generic + fill_binimage, fused) ---------
 * Replaces the launch sequence of libsutra/src.cpp/sutra_wfs_sh.cpp:337-621:
 * fillcamplipup (src.cu/sutra_wfs.cu:55-85), CarmaFFT (libcarma/src.cpp/
 * carma_fft.cpp:164-180), abs2c (src.cu/sutra_utils.cu:157-168), fillbincube
 * (src.cu/sutra_wfs.cu:396-414), subap_reduce (696-718), subap_norm
 * (src.cu/sutra_utils.cu:181-193), prng 'P'/'N' (libcarma/src.cpp/
 * carma_obj.cpp:490-527) and fillbinimg (src.cu/sutra_wfs.cu:205-229),
 * plus an optional fused get_centroids. */
int32_t shwfs_sh_comp_image(const shwfs_sh_geom* geom, const shwfs_frame_args* frame, void* stream);

/* The same frame with no device allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  shwfs_sh_comp_image builds the MFMA DFT
 * operand table of a shape in memory it allocates on first use; this entry
 * instead rebuilds that table, on `stream`, into a static device array of
 * SHWFS_NOALLOC_SCRATCH_BYTES held by each device until the process ends
 * (every DFT shape fits: nfft <= 64, nphase <= 32), then runs the same
 * kernels.  Results equal shwfs_sh_comp_image's bit for bit.  All calls on
 * one device share that array: calls that overlap on one device (other
 * streams, other host threads) race and must be ordered by the caller.
 * The FFT method uses no scratch. */
#define SHWFS_NOALLOC_SCRATCH_BYTES 32768
int32_t shwfs_sh_comp_image_noalloc(const shwfs_sh_geom* geom, const shwfs_frame_args* frame,
                                    void* stream);

/* d_flags[s] = 1 if every pupil value of the window of sub-aperture s is
 * exactly 1, else 0 (for shwfs_frame_args.d_pupil_unit).  No counterpart in
 * the reference: fully illuminated sub-apertures then skip the pupil read. */
int32_t shwfs_sh_unit_pupil(const shwfs_sh_geom* geom, const float* d_pupil, int32_t* d_flags,
                            void* stream);
/* Validates a geometry against what the kernels implement (no launch). */
int32_t shwfs_sh_check_geom(const shwfs_sh_geom* geom, int32_t method);

/* Host-side analysis of a binmap (no GPU): returns 1 and sets *origin when,
 * for every binned pixel p = px + py*npix, the set of its nrebin^2 HR indices
 * is {((o + px*nrebin + a) mod nfft) + ((o + py*nrebin + b) mod nfft) * nfft},
 * the form geom_init.py:872-899 builds for an NGS (centred window, fftshift);
 * returns 0 otherwise, negative on bad arguments.  Block form lets the
 * kernels skip the table; the summation order inside a wrapped block may
 * differ from the table's (float rounding only). */
int32_t shwfs_binmap_block_origin(const int32_t* h_binmap, int32_t nfft, int32_t npix,
                                  int32_t nrebin, int32_t* origin);

/* ---- centroiders on a detector image ---------------------------------------
 * Replaces get_centroids of libsutra/src.cu/sutra_centroider_cog.cu:83-150,
 * sutra_centroider_wcog.cu:118-170 and sutra_centroider_tcog.cu:78-141
 * (TCOG writes the thresholded pixels back into the image, as the reference). */
int32_t shwfs_centroids(const float* d_img_in, float* d_img_out, const shwfs_centroid_args* args,
                        void* stream);

/* Broadcast a npix x npix weight map to every sub-aperture:
 * replaces fill_weights, libsutra/src.cu/sutra_centroider_wcog.cu:27-50. */
int32_t shwfs_fill_weights(float* d_out, const float* d_in, int32_t npix, int32_t n_total,
                           void* stream);

/* Brightest-pixel CoG over the image d_img (a->kind, threshold and weights are
 * not used): per sub-aperture the window value of rank nbpix in decreasing
 * order (0: the brightest; pixels outside the image count as 0) is the
 * threshold t; pixels keep v > t ? v - t : 0, written back into d_img, and the
 * CoG of those values gives the slopes and intensities.  Replaces
 * get_centroids of libsutra/src.cu/sutra_centroider_bpcog.cu:28-78 (run by
 * SutraCentroiderBpcog::get_cog, src.cpp/sutra_centroider_bpcog.cpp:77-90).
 * 0 <= nbpix < npix^2 <= 1024. */
int32_t shwfs_centroids_bpcog(float* d_img, int32_t nbpix, const shwfs_centroid_args* a,
                              void* stream);

/* Correlation centroider over the image d_img (row stride `size`, the image
 * side npix * nxsub): per sub-aperture s the spot b[x, y] = d_img[(validx[s] +
 * x) + (validy[s] + y) * size] is correlated with the reference function r
 * (d_ref_fn [npix^2], x fastest; with ref_per_subap [npix^2 * nvalid]):
 * C[kx, ky] = sum_m b[m + k] r[m] for lags -(npix-1) .. npix-1, divided by
 * d_corrnorm [(2 npix - 1)^2] (NULL: 1).  The maximum over the central nbmax^2
 * window (nbmax = npix if 2 npix - 1 > 20, else 2 npix - 1; the lowest flat
 * index on ties) centres a sizex x sizey patch (moved inside the map when the
 * maximum sits on its border); d_interpmat [sizex*sizey, 6] (column-major:
 * element (t, c) at t + c*sizex*sizey, patch sample t = col + row*sizex) fits
 * a x^2 + b y^2 + c xy + d x + e y + f, and the vertex plus the patch position
 * gives ((v + m - 0.5 - npix/2) - offset) * scale, minus d_ref (may be NULL;
 * in the order of the slopes: slope_order 0 x then y, 1 interlaced).
 * d_corr [(2 npix - 1)^2 * nvalid] and d_corrmax [nvalid] (flat index of the
 * maximum in its map) may be NULL.  Replaces SutraCentroiderCorr::get_cog
 * (libsutra/src.cpp/sutra_centroider_corr.cpp:190-262) and its kernels
 * fillcorr, correl, roll2real, corr_norm, sortmaxi and interp_parab
 * (libsutra/src.cu/sutra_centroider_corr.cu).  npix <= 16; 6 <= sizex*sizey
 * <= 49, sizex and sizey <= 2 npix - 1. */
int32_t shwfs_centroids_corr(const float* d_img, int32_t size, const int32_t* d_validx,
                             const int32_t* d_validy, int32_t nvalid, int32_t npix,
                             const float* d_ref_fn, int32_t ref_per_subap,
                             const float* d_corrnorm, const float* d_interpmat, int32_t sizex,
                             int32_t sizey, float offset, float scale, const float* d_ref,
                             int32_t slope_order, float* d_slopes, float* d_corr,
                             int32_t* d_corrmax, void* stream);

/* ---- field stop ----------------------------------------------------------------
 * The pupil field pupil * exp(i scale phase) ([npup^2] each, x fastest) zero-
 * padded to n x n, forward transform (unnormalised), product with d_submask
 * [n^2] (element kx + ky n: the stop in the focal plane, fftshifted as
 * geom_init builds it), inverse transform (unnormalised), divided by n^2:
 * the filtered field F on the npup x npup support, returned as amplitude
 * d_ampl_out = |F| and phase d_phase_out = arg(F) / scale ([npup^2] each).
 * Passed as d_pupil / d_phase of shwfs_sh_comp_image (with d_pupil_unit NULL)
 * they give the spot field F * exp(-i offset).  d_work: [2 * n * npup] floats
 * (complex n x npup rows).  Replaces the field-stop branch of
 * SutraWfsSH::comp_generic (libsutra/src.cpp/sutra_wfs_sh.cpp:366-391:
 * fsamplipup_krnl, CarmaFFT, submask_krnl, CarmaFFT, fillcamplipup from the
 * focal field, libsutra/src.cu/sutra_wfs.cu:87-181, 1868-1888).
 * n a power of two in [16, 4096], npup <= n, scale finite and non-zero. */
int32_t shwfs_field_stop(const float* d_phase, const float* d_pupil, int32_t npup, float scale,
                         const float* d_submask, int32_t n, void* d_work, float* d_phase_out,
                         float* d_ampl_out, void* stream);

/* ---- LGS elongated spot kernels ---------------------------------------------
 * Per sub-aperture s, the sodium profile seen along the elongation, sampled on
 * the ntot HR pixels x (pixel size pixsize arcsec): with norm = 206265 *
 * doffaxis[s] / hg^2 and t(x) = ((x - c) pixsize / norm + hg - h0) / deltah
 * (c = ntot/2 + 0.5 for even ntot, ntot/2 for odd), linear interpolation of
 * prof1d at t (index clamped to [0, nprof - 2]; 0 weight below 0, weight 1 at
 * the top), or, when pixsize > deltah * norm, of profcum at t(x+1) minus at
 * t(x); then convolved with the beam: unnormalised inverse DFT of DFT(p) *
 * ftbeam.  d_prof2d [2 * ntot * nvalid] complex (x fastest).  Replaces
 * SutraLGS::lgs_update (libsutra/src.cpp/sutra_lgs.cpp:160-178: iprof_krnl,
 * CarmaFFT, tftbeam_krnl, CarmaFFT, libsutra/src.cu/sutra_lgs.cu:25-104).
 * nprof >= 2; the first nprof values of prof1d and profcum are read. */
int32_t shwfs_lgs_prof2d(float* d_prof2d, const float* d_prof1d, const float* d_profcum,
                         int32_t nprof, const float* d_doffaxis, float hg, float pixsize,
                         float h0, float deltah, const float* d_ftbeam, int32_t ntot,
                         int32_t nvalid, void* stream);
/* The elongated kernels and their transforms: image[x + y ntot] =
 * |Re prof2d[(x + ntot/2) % ntot, s]| * beam[y], rotated by azimuth[s] about
 * (ntot/2, ntot/2) with bilinear interpolation (clamped at the edges), stored
 * in d_lgskern [ntot^2 * nvalid] (may be NULL), and its forward 2-D transform
 * (unnormalised) in d_ftlgskern [2 * ntot^2 * nvalid].  With d_prof2d NULL the
 * kernels are read from d_lgskern and only transformed (SutraLGS::
 * load_kernels).  Replaces SutraLGS::lgs_makespot (sutra_lgs.cpp:180-212:
 * rollbeamexp_krnl, rotate_krnl, CarmaFFT; sutra_lgs.cu:106-300).
 * ntot in {16, 32, 64}. */
int32_t shwfs_lgs_kernels(float* d_ftlgskern, float* d_lgskern, const float* d_prof2d,
                          const float* d_beam, const float* d_azimuth, int32_t ntot,
                          int32_t nvalid, void* stream);

/* ---- counter-based noise on a device array ----------------------------------
 * Replaces CarmaObj<float>::prng (libcarma/src.cpp/carma_obj.cpp:490-527,
 * src.cu/carma_rng.cu:98-141,287-303) with Philox4x32-10 streams:
 * gtype 'P': d = Poisson(d); 'N': d = beta*d + alpha*N(0,1); 'U': d = beta*d + U(0,1). */
int32_t shwfs_prng(float* d_data, int64_t n, char gtype, float alpha, float beta, uint64_t seed,
                   uint32_t call, uint32_t domain, void* stream);

/* ---- detector calibration ----------------------------------------------------
 * cal[i] = (raw[lutPix[i]] - dark[i]) * flat[i]; replaces calib_krnl,
 * libsutra/src.cu/sutra_centroider.cu:107-114 (float input). */
int32_t shwfs_calibrate(const float* d_raw, float* d_cal, const float* d_dark, const float* d_flat,
                        const int32_t* d_lutPix, int64_t n, void* stream);

/* Same for a uint16 camera image: calib_krnl<uint16_t>
 * (libsutra/src.cu/sutra_centroider.cu:107-114, instantiated at :140-160). */
int32_t shwfs_calibrate_u16(const uint16_t* d_raw, float* d_cal, const float* d_dark,
                            const float* d_flat, const int32_t* d_lutPix, int64_t n, void* stream);

/* ---- simulated camera image (fakecam) ----------------------------------------
 * cam[i] = u16(min(fma(img[i], flat[i], dark[i]) / max_flux_per_pix, 1) * max_pix_value),
 * u16() converting toward zero with saturation to [0, 65535] (NaN -> 0), as the
 * reference's float-to-integer conversion does; replaces digitalize_krnl,
 * libsutra/src.cu/sutra_wfs.cu:2243-2255, called from SutraWfsSH::comp_image
 * (libsutra/src.cpp/sutra_wfs_sh.cpp:635-640).  max_flux_per_pix > 0. */
int32_t shwfs_digitalize(uint16_t* d_cam, const float* d_img, const float* d_dark,
                         const float* d_flat, int32_t max_flux_per_pix, int32_t max_pix_value,
                         int64_t n, void* stream);

/* ---- external detector image -> spots (acquisim) -----------------------------
 * bincube[s*npix^2 + x + y*npix] = image[(x + validx[s]) + (y + validy[s]) * size]
 * (size: the image side, npix * nxsub; 0 for an element past size^2, which the
 * reference does not guard); replaces bcube_krnl, libsutra/src.cu/
 * sutra_acquisim.cu:65-87 (SutraAcquisim::comp_image, src.cpp/
 * sutra_acquisim.cpp:52-66). */
int32_t shwfs_fill_bincube(const float* d_image, float* d_bincube, const int32_t* d_validx,
                           const int32_t* d_validy, int32_t npix, int32_t nvalid, int32_t size,
                           void* stream);
/* The same from a sub-aperture number map over the nxsub x nxsub grid:
 * num_ssp[i + nxsub*j] = 0 for an invalid sub-aperture, else 1 + its index in
 * the bincube (entries above nvalid are skipped); replaces bcube_krnl_2D,
 * libsutra/src.cu/sutra_acquisim.cu:26-48 (SutraAcquisim::comp_image_2D). */
int32_t shwfs_fill_bincube_2d(const float* d_image, float* d_bincube, const int32_t* d_num_ssp,
                              int32_t npix, int32_t nxsub, int32_t nvalid, void* stream);

/* ---- geometric slopes (mean phase gradient per sub-aperture) -----------------
 * Replaces SutraWfs::slopes_geom (libsutra/src.cpp/sutra_wfs.cpp:199-245):
 * type 0: edge differences, reduce_phasex/y (libsutra/src.cu/sutra_wfs.cu:868-913);
 * type 1: pupil-weighted central differences over the window divided by
 *         2 n fluxPerSub, derive_phasex/y (916-1033; needs d_pupil, the
 *         screen-sized pupil, and d_fluxPerSub [nvalid]);
 * type 2: projection on ttprojmat (1036-1090; needs d_ttprojmat
 *         [nphase^2 * 2 nvalid]).
 * phase in micrometres; phasemap [nphase^2 * nvalid] flat screen indices;
 * slopes [2*nvalid] (x then y); alpha as the reference (0.206265 / subapd *
 * nphase / (nphase - 1)) for types 0 and 1.  Unused pointers may be NULL. */
int32_t shwfs_slopes_geom(int32_t type, const float* d_phase, const float* d_pupil,
                          const float* d_fluxPerSub, float* d_slopes, const int32_t* d_phasemap,
                          const float* d_ttprojmat, int32_t nphase, int32_t nvalid, float alpha,
                          void* stream);

/* ---- tip/tilt filter of a slope vector ---------------------------------------
 * tip = <s, ref_tip>, tilt = <s, ref_tilt> on the unfiltered slopes, then
 * s -= tip * ref_tip + tilt * ref_tilt in place; d_tt (may be NULL) receives
 * {tip, tilt}.  One launch on the device (graph-capturable); replaces
 * SutraCentroider::apply_TT_filter (libsutra/src.cpp/sutra_centroider.cpp:322-338). */
int32_t shwfs_tt_filter(float* d_slopes, const float* d_ref_tip, const float* d_ref_tilt,
                        float* d_tt, int32_t nslopes, void* stream);

/* ---- spots into the detector image (fill_binimage) ----------------------------
 * With noise_background > 0 the whole image first becomes noise_background *
 * N(0,1) (Philox stream (seed, call, domain), element index i); then every spot
 * pixel is written: binimg[(validsubsx[s] + px) + (validsubsy[s] + py) * nsub] =
 * bincube[px + py*npix + s*npix^2].  Other pixels keep their values.  Replaces
 * SutraWfsSH::fill_binimage (libsutra/src.cpp/sutra_wfs_sh.cpp:599-621):
 * d_binimg->prng('N', noise) and fillbinimg (src.cu/sutra_wfs.cu:183-229).
 * The spot kernels draw their own background with domain 2 and the frame's
 * call, so this entry with domain 2 reproduces their image. */
int32_t shwfs_fill_binimg(const float* d_bincube, float* d_binimg, const int32_t* d_validsubsx,
                          const int32_t* d_validsubsy, int32_t npix, int32_t nvalid, int32_t nsub,
                          float noise_background, uint64_t seed, uint32_t call, uint32_t domain,
                          void* stream);

/* ---- checked host-to-device copy -----------------------------------------------
 * Copies src_elems elements of elem_size bytes to a device array of dst_elems
 * elements and waits for the copy; SHWFS_ERR_ARG (nothing copied) when the
 * sizes differ.  Used by set_bincube / set_binimg, where the reference prints
 * a message and returns success (libsutra/src.cpp/sutra_wfs_sh.cpp:665-672,
 * src.cpp/sutra_wfs.cpp:253-261): here a wrong size is a defined error. */
int32_t shwfs_upload(void* d_dst, int64_t dst_elems, const void* h_src, int64_t src_elems,
                     int32_t elem_size, void* stream);

/* ---- batched frames over several SH WFS ---------------------------------------
 * One launch per frame covers every sub-aperture of every WFS of the batch
 * (uneven sizes and methods mixed), replacing the sequential per-WFS launch
 * sequences of SutraSensors (libsutra/src.cpp/sutra_sensors.cpp:42-91, one
 * cuFFT plan per WFS shape) driven by the supervisor's loop over WFS
 * (shesha/supervisor/compassSupervisor.py:256-265).  The spot noise of WFS i
 * at frame f uses the Philox call frames[i].call + f, f being a device-side
 * frame counter (starts at 0, advanced once per frame), so captured graphs
 * replay successive frames.  All pointers in geoms / frames (and their fused
 * centroiders) are captured at creation and must stay valid.  A NULL stream
 * means an internal stream of the batch that is ordered with the default
 * stream. */
typedef struct shwfs_batch shwfs_batch; /* opaque */
typedef struct shwfs_batch_tile {
  int32_t method; /* SHWFS_METHOD_* used for this WFS in the batch */
  int32_t split;  /* FFT stage: 1, or 2 (row transforms split in two, needs
                     2*nphase <= nfft); 0 = 2 when allowed.  Ignored by DFT. */
} shwfs_batch_tile;
/* frames[i].method is the initial method of WFS i (wg_size is ignored).
 * Refused (SHWFS_ERR_UNSUPPORTED): ntot != nfft, kernconv, LGS kernels, layers, and the
 * SHWFS_CENTROID_MLP kind, which run per WFS through shwfs_sh_comp_image. */
int32_t shwfs_batch_create(const shwfs_sh_geom* geoms, const shwfs_frame_args* frames,
                           int32_t nwfs, shwfs_batch** out);
int32_t shwfs_batch_destroy(shwfs_batch* batch);
/* tiles: [nwfs] or NULL (keep); grid: workgroups of the launch, 0 = one per
 * tile (otherwise each loops over tiles); remap: 1 gives each XCD a
 * contiguous range of tiles.  Drops a captured graph. */
int32_t shwfs_batch_set_tiles(shwfs_batch* batch, const shwfs_batch_tile* tiles, int32_t grid,
                              int32_t remap);
int32_t shwfs_batch_get_tiles(const shwfs_batch* batch, shwfs_batch_tile* tiles, int32_t* grid,
                              int32_t* remap);
int32_t shwfs_batch_set_frame(shwfs_batch* batch, uint32_t frame, void* stream);
/* Synchronises the device, then reads the frame counter. */
int32_t shwfs_batch_get_frame(shwfs_batch* batch, uint32_t* frame);
/* One frame of every WFS (+ the frame counter advance). */
int32_t shwfs_batch_launch(shwfs_batch* batch, void* stream);
/* Captures `frames` consecutive frames into a hipGraph (replacing any
 * previous one); nothing runs until shwfs_batch_replay. */
int32_t shwfs_batch_capture(shwfs_batch* batch, int32_t frames, void* stream);
/* Launches the captured graph `ngraphs` times (ngraphs * frames frames). */
int32_t shwfs_batch_replay(shwfs_batch* batch, int32_t ngraphs, void* stream);
/* Tile autotuner: per WFS, the methods allowed by method_mask (bit m =
 * SHWFS_METHOD_m) and FFT splits; then the launch grid and XCD mapping; each
 * candidate timed over `reps` frames (median of 3).  Keeps the fastest
 * configuration and returns its time per frame.  Overwrites the outputs and
 * drops a captured graph; the frame counter is left unchanged. */
int32_t shwfs_batch_autotune(shwfs_batch* batch, uint32_t method_mask, int32_t reps, void* stream,
                             float* best_us);

/* ---- learned centroider (MLP 64 -> 128 -> 128 -> 2, GELU) ---------------------
 * No counterpart in the reference: a centroider type next to CoG / WCoG /
 * TCoG (libsutra/src.cu/sutra_centroider_{cog,wcog,tcog}.cu), trained on spot
 * / true-gradient pairs produced by the WFS itself.  bf16 operands on the
 * matrix cores, FP32 accumulation, FP32 master parameters.
 * d_params: flat [SHWFS_MLP_NPARAM] = W1 (128x64 row-major), b1 (128),
 * W2 (128x128), b2 (128), W3 (2x128), b3 (2).  d_work: shwfs_mlp_work_bytes()
 * of device scratch for the bf16 weight copies.  Inputs x: [n][64] spots
 * (pixel px + 8*py), multiplied by x_scale; outputs / targets [n][2] in
 * pixels.  Loss: mean over samples and outputs of (y - t)^2. */
#define SHWFS_MLP_NPARAM 25090
size_t shwfs_mlp_work_bytes(void);
/* LeCun-normal weights from a Philox stream, zero biases. */
int32_t shwfs_mlp_init(float* d_params, uint64_t seed, void* stream);
int32_t shwfs_mlp_forward(const float* d_params, void* d_work, const float* d_x, int64_t n,
                          float x_scale, float* d_y, void* stream);
/* Mean loss into *d_loss and its gradient into d_grad (both overwritten);
 * targets are d_t * t_scale.  d_index: optional [n] rows of d_x / d_t to use
 * (a shuffled batch; each row must exist), or NULL. */
int32_t shwfs_mlp_loss_grad(const float* d_params, void* d_work, const float* d_x,
                            const float* d_t, const int32_t* d_index, int64_t n, float x_scale,
                            float t_scale, float* d_grad, float* d_loss, void* stream);
/* One Adam step (step >= 1, bias-corrected) on d_params; zeroes d_grad. */
int32_t shwfs_mlp_adam(float* d_params, float* d_grad, float* d_m, float* d_v, int32_t step,
                       float lr, float beta1, float beta2, float eps, void* stream);
/* Slopes of 8x8-pixel spots: the MLP output * y_scale - ref, or the CoG of
 * the same pixels, (Sx/(I+1e-6) - offset) * scale - ref, when the spot flux I
 * is outside [flux_lo, flux_hi] (d_fallback[s] = 1 then; may be NULL).  Spots
 * are the windows at (d_validx[s], d_validy[s]) of a detector image of side
 * `size`, or, with both valid arrays NULL, a spot cube [nvalid][64].
 * d_intensities (may be NULL) receives I.  slope_order as
 * shwfs_centroid_args. */
int32_t shwfs_mlp_centroids(const float* d_params, void* d_work, const float* d_img,
                            int32_t size, const int32_t* d_validx, const int32_t* d_validy,
                            int32_t nvalid, int32_t npix, float x_scale, float flux_lo,
                            float flux_hi, float offset, float scale, float y_scale,
                            const float* d_ref, int32_t slope_order, float* d_slopes,
                            float* d_intensities, int32_t* d_fallback, void* stream);

/* ---- guide-star raytrace -----------------------------------------------------------
 * d_out (nx x ny, x fastest) = (accumulate ? d_out : 0) + sum over the layers,
 * in order, of each layer's bilinear value at the traced point:
 *   x1 = (x - nx/2) G, y1 = (y - ny/2) G,
 *   xref = ((cos t x1 - sin t y1) + nx/2 + xoff/delta) delta + dx (y alike),
 *   (1-u)(1-v) L[i,j] + u(1-v) L[i+1,j] + (1-u) v L[i,j+1] + u v L[i+1,j+1]
 * with i = trunc(xref), u = xref - i; a point whose square leaves the layer
 * adds nothing (the reference adds nothing past the far edge and reads
 * before the layer for a negative i or j; here those add nothing too).
 * Replaces SutraSource::raytrace(SutraAtmos*) (libsutra/src.cpp/
 * sutra_source.cpp:222-267): cudaMemset + target_raytrace / raytrace_krnl
 * per layer (libsutra/src.cu/sutra_target.cu:25-98), in one launch. */
int32_t shwfs_raytrace(float* d_out, int32_t nx, int32_t ny, const shwfs_layer* layers,
                       int32_t nlayers, int32_t accumulate, void* stream);

/* ---- phase window (synthetic inputs) ---------------------------------------------
 * d_dst[x + y*n] = gain * d_src[(ox + x) + (oy + y)*src_side] for x, y < n:
 * a translated, rescaled window of a larger screen (frozen-flow shift and
 * r0 scaling of a von Karman screen).  No counterpart in the reference. */
int32_t shwfs_phase_window(const float* d_src, int32_t src_side, int32_t ox, int32_t oy,
                           float gain, float* d_dst, int32_t n, void* stream);

/* ---- diagnostics ------------------------------------------------------------ */
/* One 16x32 by 32x16 product on the matrix cores with the operand layouts the
 * DFT kernels use: d_A (16x32) and d_B (32x16) row-major float, rounded to the
 * method's input type (fp16, or bf16 hi/lo for BF16X3); d_D (16x16) row-major.
 * No counterpart in the reference: it checks the MFMA lane layout. */
int32_t shwfs_mfma_probe(const float* d_A, const float* d_B, float* d_D, int32_t method,
                         void* stream);
int32_t shwfs_last_error(char* buf, size_t len);
const char* shwfs_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SHWFS_H */
