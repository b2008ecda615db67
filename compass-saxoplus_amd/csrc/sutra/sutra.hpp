// sutra: the reference's C++ class API for the SH WFS path, re-implemented
// over the libshwfs C ABI.
//
// Class and method names and argument lists follow the reference
// (libsutra/include/sutra_{telescope,source,wfs,wfs_sh,sensors,centroider,
// centroider_cog,centroider_wcog,centroider_tcog,rtc}.hpp).  What changes:
//  * the FFT plan maps keyed by {nfft, nfft, nvalid} hold a project plan
//    (ShwfsPlan: method + tile choice) instead of a vendor FFT handle;
//  * complex arrays are CarmaObj<hipFloatComplex>;
//  * comp_image is one fused launch (plus a background-noise launch when the
//    WFS has readout noise); d_camplipup / d_camplifoc exist for the API but
//    the fused kernels never materialise the complex cubes;
//  * errors raise std::runtime_error instead of DEBUG-only throws.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../carma/carma.hpp"

using std::map;
using std::string;
using std::vector;

// Execution plan of one SH WFS spot computation (replaces cufftHandle*).
struct ShwfsPlan {
  int32_t nfft = 0;
  int32_t nvalid = 0;
  int32_t method = SHWFS_METHOD_FFT_FP32;
  int32_t wg_size = 0;  // threads per workgroup of the spot kernel (0 = library default)
};
using PlanMap = map<vector<int32_t>, ShwfsPlan*>;

class SutraTelescope {
 public:
  CarmaContext* current_context;
  int32_t device;
  int64_t pup_size;
  int64_t pup_size_m;
  int64_t num_eleme_pup;
  CarmaObj<float>* d_pupil;
  CarmaObj<float>* d_pupil_m;

  SutraTelescope(CarmaContext* context, int64_t n_pup, int64_t npos, float* pupil, int64_t n_pup_m,
                 float* pupil_m);
  ~SutraTelescope();
};

class SutraPhase {
 public:
  CarmaObj<float>* d_screen;
  int64_t screen_size;
  SutraPhase(CarmaContext* context, int64_t size);
  ~SutraPhase();
};

// One turbulent layer (reference: libsutra/include/sutra_tscreen.hpp:29-80):
// its phase screen d_tscreen->d_screen (µm, x fastest), altitude and
// accumulated wind shift accumx / accumy (pixels), which raytrace adds to a
// source's offsets.  The screen is written by the caller (host2device of
// d_screen); the extrusion machinery of the reference (A / B matrices,
// stencils, move_atmos) is not part of this build.
class SutraTurbuScreen {
 public:
  int32_t device;
  SutraPhase* d_tscreen;
  int64_t screen_size;
  float r0, amplitude, altitude, windspeed, winddir, deltax, deltay;
  float accumx, accumy;
  CarmaContext* current_context;

  SutraTurbuScreen(CarmaContext* context, int64_t size, int64_t size2, float amplitude,
                   float altitude, float windspeed, float winddir, float deltax, float deltay,
                   int32_t device);
  ~SutraTurbuScreen();
};

// The layers of the atmosphere (reference: libsutra/include/sutra_atmos.hpp,
// constructor libsutra/src.cpp/sutra_atmos.cpp:24-58 with the same arguments).
class SutraAtmos {
 public:
  int32_t nscreens;
  vector<SutraTurbuScreen*> d_screens;
  float r0;
  CarmaContext* current_context;

  SutraAtmos(CarmaContext* context, int32_t nscreens, float global_r0, float* r0_per_layer,
             int64_t* size1, int64_t* size2, float* altitude, float* windspeed, float* winddir,
             float* deltax, float* deltay, int32_t device);
  ~SutraAtmos();
};

// Laser guide star of an SH WFS (reference: libsutra/include/sutra_lgs.hpp,
// src.cpp/sutra_lgs.cpp): the sodium profile seen by every sub-aperture,
// convolved with the beam (d_prof2d), and the elongated spot kernels with
// their transforms (d_lgskern, d_ftlgskern), which the spot kernel of the WFS
// multiplies into every frame's HR image.  npix is the HR support (ntot).
// The kernels are built for all sub-apertures by lgs_init / lgs_update (and
// kept until the next one); the reference rebuilds them for each chunk of
// nmaxhr sub-apertures in every frame from the same profile, which gives the
// same kernels.  Kernels given by load_kernels stay in place (in the reference
// the next frame overwrites them from the profile).
class SutraLGS {
 public:
  int32_t device;
  int64_t nvalid, npix, nmaxhr;
  int32_t nprof;
  float hg, h0, deltah, pixsize;
  bool ready;  // kernels built
  CarmaObj<float>* d_doffaxis;
  CarmaObj<float>* d_azimuth;
  CarmaObj<float>* d_prof1d;
  CarmaObj<float>* d_profcum;
  CarmaObj<hipFloatComplex>* d_prof2d;  // [npix, nvalid]
  CarmaObj<float>* d_beam;              // [npix]
  CarmaObj<hipFloatComplex>* d_ftbeam;  // [npix]
  CarmaObj<float>* d_lgskern;           // [npix, npix, nvalid]
  CarmaObj<hipFloatComplex>* d_ftlgskern;
  CarmaContext* current_context;

  // d_lgskern / d_ftlgskern of the reference's shared buffers are not used:
  // each LGS owns kernels for all its sub-apertures (pass NULL)
  SutraLGS(CarmaContext* context, CarmaObj<float>* d_lgskern,
           CarmaObj<hipFloatComplex>* d_ftlgskern, PlanMap& ftlgskern_plans, int64_t nvalid,
           int64_t npix, int64_t nmaxhr);
  ~SutraLGS();
  int32_t lgs_init(int32_t nprof, float hg, float h0, float deltah, float pixsize,
                   float* doffaxis, float* prof1d, float* profcum, float* beam,
                   hipFloatComplex* ftbeam, float* azimuth);
  int32_t load_prof(float* prof1d, float* profcum, float hg, float h0, float deltah);
  // profile -> prof2d, then the kernels (lgs_makespot) for every sub-aperture
  int32_t lgs_update(CarmaDevice* device = nullptr, hipStream_t stream = nullptr);
  // kernels of the sub-apertures from nin / npix on (the reference's chunk
  // offset nin counts profile elements); here always all of them
  int32_t lgs_makespot(CarmaDevice* device = nullptr, int32_t nin = 0,
                       hipStream_t stream = nullptr);
  int32_t load_kernels(float* h_lgskern, CarmaDevice* device = nullptr);
};

// Guide-star source: holds the phase seen by the WFS (µm) and scale = 2*pi/lambda
// (reference: libsutra/src.cpp/sutra_source.cpp:25-100), and the layers it
// sees (add_layer) with their offsets.
//
// raytrace(atmos) keeps the layers as `pending` instead of rendering them:
// the FFT spot kernel of the WFS then traces its windows through them in its
// field stage (the npup^2 screen is neither written nor read).  Any other
// reader of the screen renders it first (render_layers): the phase property,
// slopes_geom, the MFMA methods, batches.  set_phase / reset_phase drop the
// pending layers.  lazy_raytrace = false renders at once, as the reference.
class SutraSource {
 public:
  CarmaContext* current_context;
  int32_t device;
  float posx, posy;
  int64_t npts;
  float mag, lambda, zp, scale;
  bool lgs;
  float G, thetaML, dx, dy;
  string type;
  SutraPhase* d_phase;
  SutraLGS* d_lgs;  // laser guide star (lgs), else NULL
  map<std::pair<string, int32_t>, float> xoff, yoff;  // (layer type, index) -> offset
  vector<shwfs_layer> pending;  // layers of the last raytrace, not rendered yet
  bool lazy_raytrace;

  SutraSource(CarmaContext* context, float xpos, float ypos, float lambda, float mag, float zerop,
              int64_t size, string type, int32_t device);
  ~SutraSource();
  int32_t set_phase(float* data);
  int32_t reset_phase();
  int32_t add_layer(string type, int32_t idx, float xoff, float yoff);
  int32_t remove_layer(string type, int32_t idx);
  int32_t raytrace(SutraAtmos* atmos, bool async = false);
  // writes the pending layers into d_phase->d_screen (nothing when none)
  int32_t render_layers(hipStream_t stream = nullptr);
};

class SutraWfs {
 public:
  int32_t device;
  string type;
  int64_t nxsub, nvalid, npix, nrebin, nfft, ntot, npup, nphase, nmaxhr, nffthr;
  float subapd, nphot, nphot4imat, noise;
  bool lgs, kernconv, roket, is_low_order;
  bool fakecam;
  int32_t max_flux_per_pix, max_pix_value;

  ShwfsPlan* campli_plan;
  ShwfsPlan* fttotim_plan;
  CarmaObj<hipFloatComplex>* d_ftkernel;
  CarmaObj<hipFloatComplex>* d_camplipup;
  CarmaObj<hipFloatComplex>* d_camplifoc;
  CarmaObj<hipFloatComplex>* d_fttotim;

  CarmaObj<float>* d_pupil;
  CarmaObj<float>* d_bincube;
  CarmaObj<float>* d_binimg;
  CarmaObj<float>* d_binimg_notnoisy;
  CarmaObj<float>* d_intensities;
  CarmaObj<float>* d_offsets;
  CarmaObj<float>* d_fluxPerSub;
  CarmaObj<float>* d_sincar;
  CarmaObj<int32_t>* d_hrmap;
  CarmaObj<uint16_t>* d_camimg;
  CarmaObj<float>* d_dark;
  CarmaObj<float>* d_flat;
  CarmaObj<float>* d_slopes;

  SutraSource* d_gs;
  CarmaObj<int32_t>* d_phasemap;
  CarmaObj<float>* d_ttprojmat;
  CarmaObj<float>* d_ttprojvec;
  CarmaObj<int32_t>* d_validsubsx;
  CarmaObj<int32_t>* d_validsubsy;
  CarmaObj<float>* d_submask;

  CarmaContext* current_context;

  int32_t offset, nvalid_tot, rank;

  // Stream the WFS launches on (nullptr = default stream, as the reference).
  hipStream_t stream;

 public:
  virtual ~SutraWfs();

  int32_t wfs_initgs(float xpos, float ypos, float lambda, float mag, float zerop, int64_t size,
                     float noise, int64_t seed, float G, float thetaML, float dx, float dy);
  int32_t set_pupil(float* pupil);
  // called after the pupil changed (derived classes refresh what depends on it)
  virtual void pupil_changed() {}
  int32_t set_binimg(float* binimg, int32_t nElem);
  int32_t set_dark(float* dark, int32_t nElem);
  int32_t set_flat(float* flat, int32_t nElem);
  int32_t set_fakecam(bool fakecam);
  void alloc_camera();
  int32_t digitalize_image();
  int32_t set_max_flux_per_pix(int32_t max_flux_per_pix);
  int32_t set_max_pix_value(int32_t max_pix_value);
  int32_t slopes_geom(float* slopes, int32_t type = 0);
  int32_t slopes_geom(int32_t type = 0);
  int32_t set_noise(float noise, int64_t seed);

  virtual int32_t fill_binimage(int32_t async) = 0;
  virtual int32_t comp_image(bool noise = true) = 0;
  virtual int32_t define_mpi_rank(int32_t rank, int32_t size) = 0;
  virtual int32_t allocate_buffers(PlanMap& campli_plans, PlanMap& fttotim_plans) = 0;

 protected:
  virtual int32_t comp_generic() = 0;
  SutraWfs(CarmaContext* context, SutraTelescope* d_tel, CarmaObj<hipFloatComplex>* d_camplipup,
           CarmaObj<hipFloatComplex>* d_camplifoc, CarmaObj<hipFloatComplex>* d_fttotim,
           string type, int64_t nxsub, int64_t nvalid, int64_t npix, int64_t nphase,
           int64_t nrebin, int64_t nfft, int64_t ntot, int64_t npup, float pdiam, float nphotons,
           float nphot4imat, int32_t lgs, bool fakecam, int32_t max_flux_per_pix,
           int32_t max_pix_value, bool is_low_order, bool roket, int32_t device);
};

class SutraWfsSH : public SutraWfs {
 public:
  CarmaObj<int32_t>* d_binmap;
  CarmaObj<int32_t>* d_validpuppixx;
  CarmaObj<int32_t>* d_validpuppixy;
  // field stop (set_field_stop): d_fsamplipup is the n x npup work array of
  // the transforms; the spot kernels read the stopped field as the amplitude
  // d_fsampl and the phase d_fsphase ([npup^2] each).  d_fsamplifoc stays
  // NULL: the focal-plane field is never stored whole.
  CarmaObj<hipFloatComplex>* d_fsamplipup;
  CarmaObj<hipFloatComplex>* d_fsamplifoc;
  CarmaObj<float>* d_fsphase;
  CarmaObj<float>* d_fsampl;
  ShwfsPlan* fsampli_plan;

  // Philox state of the spot / image noise: (key, call counter).
  uint64_t noise_seed;
  uint32_t noise_call;
  bool arrays_loaded;
  bool own_plan;  // campli_plan was cloned by set_method and belongs to this WFS
  int32_t bin_block, bin_origin;  // binmap analysis done at load_arrays
  CarmaObj<int32_t>* d_pupil_unit;  // [nvalid] 1 where the pupil window is all ones
  // detector-image pixels outside every spot window (built by load_arrays):
  // a noisy frame draws its readout background on them in the spot launch
  CarmaObj<int32_t>* d_bgpix;
  int64_t nbgpix;

  SutraWfsSH(CarmaContext* context, SutraTelescope* d_tel, CarmaObj<hipFloatComplex>* d_camplipup,
             CarmaObj<hipFloatComplex>* d_camplifoc, CarmaObj<hipFloatComplex>* d_fttotim,
             int64_t nxsub, int64_t nvalid, int64_t npix, int64_t nphase, int64_t nrebin,
             int64_t nfft, int64_t ntot, int64_t npup, float pdiam, float nphotons,
             float nphot4imat, int32_t lgs, bool fakecam, int32_t max_flux_per_pix,
             int32_t max_pix_value, bool is_low_order, bool roket, int32_t device);
  ~SutraWfsSH();

  int32_t define_mpi_rank(int32_t rank, int32_t size);
  int32_t allocate_buffers(PlanMap& campli_plans, PlanMap& fttotim_plans);
  int32_t load_arrays(int32_t* phasemap, int32_t* hrmap, int32_t* binmap, float* offsets,
                      float* fluxPerSub, int32_t* validsubsx, int32_t* validsubsy, int32_t* istart,
                      int32_t* jstart, float* ttprojmat, hipFloatComplex* kernel);
  int32_t fill_binimage(int32_t async);
  void scatter_binimg(float background, uint32_t call, uint32_t domain);
  int32_t comp_image(bool noise = true);
  int32_t comp_nphot(float ittime, float optthroughput, float diam, int32_t nxsub, float zerop = 0,
                     float gsmag = 0, float lgsreturnperwatt = 0, float laserpower = 0);
  int32_t set_bincube(float* bincube, int32_t nElem);
  int32_t set_field_stop(PlanMap& campli_plans, float* field_stop, int32_t N);

  // Spot computation method (SHWFS_METHOD_*); FP32 FFT by default.
  int32_t set_method(int32_t method);
  int32_t get_method() const { return campli_plan ? campli_plan->method : SHWFS_METHOD_FFT_FP32; }
  int32_t set_wg_size(int32_t wg_size);
  int32_t get_wg_size() const { return campli_plan ? campli_plan->wg_size : 0; }
  // the guide star's pending layers are traced in the field stage of the spot
  // kernel (FFT method, 64-lane workgroups); otherwise they are rendered first
  bool layers_in_field_stage() const;
  // with a field stop: render the guide star's layers, then the stopped
  // field into d_fsphase / d_fsampl (read by frame_args)
  void field_stop_stage(hipStream_t st);
  // an LGS WFS computes frames only once its kernels exist
  void check_lgs_ready() const;
  // Geometry descriptor handed to the C ABI.
  shwfs_sh_geom geom() const;
  // Frame arguments of the next comp_image (advances nothing).
  shwfs_frame_args frame_args(bool with_noise) const;
  void pupil_changed() override;

 private:
  int32_t comp_generic();
  bool noise_this_call;
};

class SutraSensors {
 public:
  int32_t device;
  bool roket;
  CarmaContext* current_context;
  size_t nsensors() { return d_wfs.size(); }
  vector<SutraWfs*> d_wfs;
  PlanMap campli_plans;
  PlanMap fttotim_plans;
  PlanMap ftlgskern_plans;
  PlanMap field_stop_plans;

  CarmaObj<hipFloatComplex>* d_camplipup;
  CarmaObj<hipFloatComplex>* d_camplifoc;
  CarmaObj<hipFloatComplex>* d_fttotim;
  CarmaObj<hipFloatComplex>* d_ftlgskern;
  CarmaObj<float>* d_lgskern;

  SutraSensors(CarmaContext* context, SutraTelescope* d_tel, vector<string> type, int32_t nwfs,
               int64_t* nxsub, int64_t* nvalid, int64_t* npupils, int64_t* npix, int64_t* nphase,
               int64_t* nrebin, int64_t* nfft, int64_t* ntot, int64_t* npup, float* pdiam,
               float* nphot, float* nphot4imat, int32_t* lgs, bool* fakecam,
               int32_t* max_flux_per_pix, int32_t* max_pix_value, int32_t device, bool roket);
  ~SutraSensors();

  int32_t allocate_buffers();
  int32_t define_mpi_rank(int32_t rank, int32_t size);
  int32_t set_noise(int32_t nwfs, float noise, int64_t seed);
  int32_t set_field_stop(int32_t nwfs, float* field_stop, int32_t N);
  int32_t initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop, int64_t* size,
                 float* noise, int64_t* seed, float* G, float* thetaML, float* dx, float* dy);
  int32_t initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop, int64_t* size,
                 float* noise, float* G, float* thetaML, float* dx, float* dy);
  int32_t initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop, int64_t* size,
                 float* G, float* thetaML, float* dx, float* dy);
};

// Spots of an SH WFS from an external detector image (reference:
// libsutra/include/sutra_acquisim.hpp, src.cpp/sutra_acquisim.cpp): the image
// (side npix * nxsub) is copied to the device and its npix x npix windows at
// the valid positions are gathered into the WFS's bincube.
class SutraAcquisim {
 public:
  int32_t device;
  string type;
  int64_t nxsub;
  int64_t nvalid;
  int64_t npix;
  CarmaContext* current_context;
  SutraWfsSH* wfs;
  CarmaObj<int32_t>* d_validsubsx;  // the WFS's, or owned after set_validsubs
  CarmaObj<int32_t>* d_validsubsy;
  bool own_valid;

  SutraAcquisim(SutraSensors* sensors, int32_t wfs_num);
  ~SutraAcquisim();
  int32_t set_validsubs(int64_t nvalid, int32_t* validsubsx, int32_t* validsubsy);
  int32_t comp_image(int64_t* dims, float* bimage);
  int32_t comp_image(int64_t* dims, float* bimage, CarmaObj<float>* d_bincube);
  int32_t comp_image_2D(int64_t* dims, float* bimage, int32_t* num_ssp);
};

enum class SlopeOrder { untied = 0, interlaced = 1 };

template <class Tin, class Tout>
class SutraCentroider {
 public:
  int32_t device;
  SutraWfs* wfs;
  int32_t nvalid;
  int32_t nslopes;
  int32_t npix;
  int32_t nxsub;
  bool filter_TT;
  float offset;
  float scale;
  SlopeOrder slope_order = SlopeOrder::untied;
  CarmaContext* current_context;

  CarmaObj<Tout>* d_bincube;
  CarmaObj<float>* d_intensities;
  CarmaObj<Tout>* d_centroids_ref;
  CarmaObj<float>* d_img;
  CarmaObj<Tin>* d_img_raw;
  CarmaObj<float>* d_dark;
  CarmaObj<float>* d_flat;
  CarmaObj<int32_t>* d_lutPix;
  CarmaObj<int32_t>* d_validx;
  CarmaObj<int32_t>* d_validy;
  CarmaObj<int32_t>* d_validMask;
  CarmaObj<float>* d_centro_filtered;
  CarmaObj<float>* d_ref_Tip;
  CarmaObj<float>* d_ref_Tilt;
  CarmaObj<float>* d_TT_slopes;

 protected:
  SutraCentroider(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset, float scale,
                  bool filter_TT, int32_t device);

 public:
  virtual ~SutraCentroider();
  int32_t set_scale(float scale);
  int32_t set_offset(float offset);
  int32_t set_dark(float* dark, int32_t n);
  int32_t set_flat(float* flat, int32_t n);
  int32_t set_lutPix(int32_t* lutPix, int32_t n);
  int32_t init_calib(int32_t n, int32_t m);
  int32_t init_roi(int32_t N);
  int32_t set_centroids_ref(float* centroids_ref);
  int32_t calibrate_img() { return calibrate_img(nullptr); }
  int32_t calibrate_img(hipStream_t stream);
  int32_t load_validpos(int32_t* ivalid, int32_t* jvalid, int32_t N);
  int32_t set_npix(int32_t npix);
  int32_t set_nxsub(int32_t nxsub);
  int32_t load_img(Tin* img, int32_t n);
  int32_t load_img(Tin* img, int32_t n, int32_t location);
  int32_t load_img(Tin* img, int32_t m, int32_t n, int32_t location);
  int32_t load_img(CarmaObj<Tin>* img);
  int32_t init_img_raw(int32_t m, int32_t n);
  bool is_type(string typec) { return typec == get_type(); }
  int32_t init_TT_filter();
  int32_t apply_TT_filter(Tout* centroids);

  virtual string get_type() = 0;
  virtual int32_t get_cog(float* img, float* intensities, Tout* centroids, int32_t nvalid,
                          int32_t npix, int32_t ntot, hipStream_t stream = nullptr) = 0;
  virtual int32_t get_cog(float* intensities, Tout* slopes, bool noise) = 0;
  virtual int32_t get_cog() = 0;

  // Arguments of this centroider for the C ABI (image pointers excluded).
  virtual shwfs_centroid_args centroid_args(float* intensities, Tout* slopes, int32_t nvalid,
                                            int32_t npix, int32_t size) = 0;
  // float64 flux / first-moment sums (tcog / wcog / cog).
  bool fp64_moments = false;
  int32_t set_fp64_moments(bool v) {
    fp64_moments = v;
    return EXIT_SUCCESS;
  }
};

template <class Tin, class T>
class SutraCentroiderCog : public SutraCentroider<Tin, T> {
 public:
  SutraCentroiderCog(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                     float scale, bool filter_TT, int32_t device);
  ~SutraCentroiderCog();
  string get_type() override { return "cog"; }
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;
};

// Brightest-pixel CoG (reference: libsutra/include/sutra_centroider_bpcog.hpp,
// src.cpp/sutra_centroider_bpcog.cpp): per sub-aperture, the nmax brightest
// pixels above the (nmax+1)-th.  Taken from the detector image after the
// frame (not fused into the spot kernels).
template <class Tin, class T>
class SutraCentroiderBpcog : public SutraCentroider<Tin, T> {
 public:
  int32_t nmax;
  CarmaObj<T>* d_bpix;          // [nmax, nvalid], allocated as the reference does
  CarmaObj<uint32_t>* d_bpind;  // [nmax, nvalid]
  SutraCentroiderBpcog(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                       float scale, bool filter_TT, int32_t device, int32_t nmax);
  ~SutraCentroiderBpcog();
  string get_type() override { return "bpcog"; }
  int32_t set_nmax(int32_t nmax);
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;
};

// Correlation centroider (reference: libsutra/include/sutra_centroider_corr.hpp,
// src.cpp/sutra_centroider_corr.cpp): per sub-aperture the correlation of the
// spot with a reference function, its maximum and a parabola fitted around
// it.  The correlation is summed directly on the lags -(npix-1) .. npix-1
// (the reference's zero-padded FFT gives the same lags), so d_corrfnct holds
// the reference function in pixels and there is no spot-transform buffer
// (d_corrspot stays NULL).  Taken from the detector image after the frame.
template <class Tin, class T>
class SutraCentroiderCorr : public SutraCentroider<Tin, T> {
 public:
  int32_t interp_sizex, interp_sizey;
  CarmaObj<T>* d_corrfnct;   // [npix, npix, 1 or nvalid] reference function(s)
  CarmaObj<T>* d_corrspot;   // NULL (no transform of the spots)
  CarmaObj<T>* d_corrnorm;   // [2 npix - 1, 2 npix - 1]
  CarmaObj<int32_t>* d_corrmax;  // [nvalid] flat index of the maximum in each map
  CarmaObj<T>* d_corr;       // [2 npix - 1, 2 npix - 1, nvalid] normalised maps
  CarmaObj<T>* d_interpmat;  // [sizex * sizey, 6] parabola fit matrix
  SutraCentroiderCorr(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                      float scale, bool filter_TT, int32_t device);
  ~SutraCentroiderCorr();
  string get_type() override { return "corr"; }
  int32_t init_corr(int32_t isizex, int32_t isizey, T* interpmat);
  // corr: [npix, npix] (ndim 2, every sub-aperture) or [npix, npix, nvalid]
  // (ndim 3); corr_norm: the first (2 npix - 1)^2 values are used
  int32_t load_corr(T* corr, T* corr_norm, int32_t ndim);
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;

 private:
  bool ref_per_subap = false;
};

template <class Tin, class T>
class SutraCentroiderTcog : public SutraCentroider<Tin, T> {
 public:
  float threshold;
  SutraCentroiderTcog(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                      float scale, bool filter_TT, int32_t device);
  ~SutraCentroiderTcog();
  string get_type() override { return "tcog"; }
  int32_t set_threshold(float threshold);
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;
};

template <class Tin, class T>
class SutraCentroiderWcog : public SutraCentroider<Tin, T> {
 public:
  float threshold;
  CarmaObj<float>* d_weights;
  SutraCentroiderWcog(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                      float scale, bool filter_TT, int32_t device);
  ~SutraCentroiderWcog();
  string get_type() override { return "wcog"; }
  int32_t init_weights();
  int32_t load_weights(float* weights, int32_t ndim);
  int32_t set_threshold(float threshold);
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;
};

// Learned centroider: a per-sub-aperture MLP (64 -> 128 -> 128 -> 2, GELU)
// on 8x8-pixel spots, bf16 on the matrix cores, trained on the GPU with
// Adam; CoG of the same pixels as fallback when the spot flux is outside the
// training range.  No counterpart in the reference (a centroider type next to
// the CoG family, created by SutraRtc::add_centroider with type "mlp").
template <class Tin, class T>
class SutraCentroiderMlp : public SutraCentroider<Tin, T> {
 public:
  CarmaObj<float>* d_params;  // SHWFS_MLP_NPARAM fp32 master parameters
  CarmaObj<float>* d_grad;
  CarmaObj<float>* d_m;       // Adam moments
  CarmaObj<float>* d_v;
  CarmaObj<float>* d_work;    // bf16 weight copies
  CarmaObj<float>* d_loss;
  CarmaObj<int32_t>* d_fallback;  // [nvalid] 1 where the CoG was used
  float x_scale;              // input scale (1 / photons per sub-aperture)
  float flux_lo, flux_hi;     // in-distribution spot flux
  int32_t step;               // Adam steps taken

  SutraCentroiderMlp(CarmaContext* context, SutraWfs* wfs, int64_t nvalid, float offset,
                     float scale, bool filter_TT, int32_t device);
  ~SutraCentroiderMlp();
  string get_type() override { return "mlp"; }
  int32_t init_params(uint64_t seed);
  int32_t set_input_scale(float s);
  int32_t set_flux_range(float lo, float hi);
  // mean loss of the targets t * t_scale; gradient left in d_grad
  float loss_grad(CarmaObj<float>* x, CarmaObj<float>* t, int64_t n, float t_scale);
  int32_t forward(CarmaObj<float>* x, int64_t n, CarmaObj<float>* y);
  // epochs over shuffled batches (host shuffle, seeded); returns the mean
  // loss of the last epoch
  float train(CarmaObj<float>* x, CarmaObj<float>* t, int64_t n, float t_scale, int32_t batch,
              int32_t epochs, float lr, uint64_t seed);
  int32_t get_cog(float* img, float* intensities, T* centroids, int32_t nvalid, int32_t npix,
                  int32_t ntot, hipStream_t stream = nullptr) override;
  int32_t get_cog(float* intensities, T* slopes, bool noise) override;
  int32_t get_cog() override;
  // not fusable into the spot kernel: NONE (comp_frame runs spots then get_cog)
  shwfs_centroid_args centroid_args(float* intensities, T* slopes, int32_t nvalid, int32_t npix,
                                    int32_t size) override;
};

// Training pairs generated by the WFS itself: for each of nframes frames, a
// window of `base` (side base->get_dims(1), phase in um at r0_base) at a
// random shift, rescaled to r0 ~ U(r0_min, r0_max) ((r0/r0_base)^(-5/6)),
// becomes the guide-star phase; the noisy spots (the WFS noise) go to rows
// of X [nframes*nvalid][64] and slopes_geom(geom_type) (arcsec) to T
// [nframes*nvalid][2].  Returns the number of pairs.
int64_t make_mlp_training_set(SutraWfsSH* wfs, CarmaObj<float>* base, float r0_base, float r0_min,
                              float r0_max, int32_t nframes, uint64_t seed, int32_t geom_type,
                              CarmaObj<float>* X, CarmaObj<float>* T);

// Slope sink of the RTC: the reference's controllers are outside this build;
// the generic controller keeps what do_centroids needs (its slope vector and
// the indices of the centroiders that feed it).
template <class T, class Tout>
class SutraController {
 public:
  string type;
  int32_t device;
  int32_t nslope, nactu;
  float delay, gain;
  vector<int32_t> centro_idx;
  CarmaObj<T>* d_centroids;
  CarmaObj<T>* d_com;
  CarmaContext* current_context;
  SutraController(CarmaContext* context, string type, int32_t nslope, int32_t nactu, float delay,
                  int32_t device, const int32_t* idx_centro, int32_t ncentro);
  ~SutraController();
  string get_type() { return type; }
};

// One frame of one SH WFS with its centroider fused into the spot launch:
// same results as wfs->comp_image(noise) followed by c->get_cog(intensities,
// slopes, noise) (TCoG's write-back into d_binimg included).
int32_t comp_frame(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* intensities,
                   float* slopes, bool noise);
// Issues `nframes` frames back to back from the host (fused: one launch per
// frame; otherwise comp_image + get_cog) and returns the elapsed time in ms
// between hipEvents recorded on the WFS stream around the loop.
double run_frames(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* slopes,
                  int32_t nframes, bool noise, bool fused);
// Captures `nframes` fused frames (one launch each) into a hipGraph on a
// stream of its own, replays it `reps` times after one untimed replay, and
// returns the elapsed ms of the timed replays between hipEvents on that
// stream (frames issued by the device back to back, without host launches).
double run_frames_graph(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* slopes,
                        int32_t nframes, int32_t reps,
                        bool noise = false);

// Several SH WFS (uneven sizes) computed by one launch per frame, with their
// fused centroiders; replaces the per-WFS loop over SutraSensors::d_wfs.
// Slopes go to buffers owned by the batch (d_slopes[i], 2*nvalid each).
// Spot noise of WFS i at batch frame f uses the Philox call
// wfs[i]->noise_call + f (noise_call as it was at construction).
class SutraWfsBatch {
 public:
  vector<SutraWfsSH*> wfs;
  vector<SutraCentroider<float, float>*> centro;
  vector<CarmaObj<float>*> d_slopes;
  bool noise;
  int32_t device;
  CarmaContext* current_context;

  SutraWfsBatch(CarmaContext* context, vector<SutraWfsSH*> wfs,
                vector<SutraCentroider<float, float>*> centro, bool noise);
  ~SutraWfsBatch();
  SutraWfsBatch(const SutraWfsBatch&) = delete;
  SutraWfsBatch& operator=(const SutraWfsBatch&) = delete;

  // per-WFS (method, split), launch grid (0 = one workgroup per tile), XCD remap
  void set_tiles(const vector<int32_t>& methods, const vector<int32_t>& splits, int32_t grid,
                 int32_t remap);
  void render_sources();
  vector<int32_t> methods() const;
  vector<int32_t> splits() const;
  int32_t grid() const;
  int32_t remap() const;
  float autotune(uint32_t method_mask, int32_t reps);  // us per frame of the kept tiles
  void set_frame(uint32_t frame);
  uint32_t frame();
  void launch();                         // one frame
  void capture(int32_t frames);          // hipGraph of `frames` frames
  double replay(int32_t ngraphs);        // ms between hipEvents around the replays
  double run(int32_t nframes);           // ms for nframes plain launches

 private:
  shwfs_batch* handle;
};

template <typename Tin, typename T, typename Tout>
class SutraRtc {
 public:
  vector<SutraCentroider<Tin, T>*> d_centro;
  vector<SutraController<T, Tout>*> d_control;

  SutraRtc();
  ~SutraRtc();
  int32_t add_centroider(CarmaContext* context, int64_t nvalid, float offset, float scale,
                         bool filter_TT, int64_t device, string typec);
  int32_t add_centroider(CarmaContext* context, int64_t nvalid, float offset, float scale,
                         bool filter_TT, int64_t device, string typec, SutraWfs* wfs);
  int32_t add_controller(CarmaContext* context, string typec, int64_t device, float delay,
                         int32_t nslope, int32_t nactu, const int32_t* idx_centro,
                         int32_t ncentro);
  int32_t remove_centroider(int32_t ncentro);
  int32_t do_calibrate_img(int32_t ncntrl);
  int32_t do_centroids();
  int32_t do_centroids(int32_t ncntrl);
  int32_t do_centroids(int32_t ncntrl, bool noise);
  int32_t set_centroids_ref(float* centroids_ref);
};
