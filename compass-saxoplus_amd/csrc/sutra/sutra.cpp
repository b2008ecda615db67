// sutra classes of the SH WFS path over the libshwfs C ABI.
#include "sutra.hpp"

#include <cmath>
#include <iostream>

namespace {
constexpr double kPi = 3.14159265358979323846;

int64_t* dims1(int64_t* d, int64_t n) {
  d[0] = 1;
  d[1] = n;
  return d;
}
int64_t* dims2(int64_t* d, int64_t a, int64_t b) {
  d[0] = 2;
  d[1] = a;
  d[2] = b;
  return d;
}
int64_t* dims3(int64_t* d, int64_t a, int64_t b, int64_t c) {
  d[0] = 3;
  d[1] = a;
  d[2] = b;
  d[3] = c;
  return d;
}
template <class T>
void del(T*& p) {
  delete p;
  p = nullptr;
}
}  // namespace

// ============================================================ SutraTelescope
// reference: libsutra/src.cpp/sutra_telescope.cpp:27-60
SutraTelescope::SutraTelescope(CarmaContext* context, int64_t n_pup, int64_t npos, float* pupil,
                               int64_t n_pup_m, float* pupil_m)
    : current_context(context),
      device(context->get_active_device()),
      pup_size(n_pup),
      pup_size_m(n_pup_m),
      num_eleme_pup(npos) {
  int64_t d[4];
  d_pupil = new CarmaObj<float>(context, dims2(d, n_pup, n_pup), pupil);
  d_pupil_m = new CarmaObj<float>(context, dims2(d, n_pup_m, n_pup_m), pupil_m);
}

SutraTelescope::~SutraTelescope() {
  del(d_pupil);
  del(d_pupil_m);
}

// ============================================================ SutraPhase / SutraSource
SutraPhase::SutraPhase(CarmaContext* context, int64_t size) : screen_size(size) {
  int64_t d[4];
  d_screen = new CarmaObj<float>(context, dims2(d, size, size));
}
SutraPhase::~SutraPhase() { del(d_screen); }

// reference: libsutra/src.cpp/sutra_source.cpp:62-100 (scale = 2*pi/lambda, phase in µm)
SutraSource::SutraSource(CarmaContext* context, float xpos, float ypos, float lambda_, float mag_,
                         float zerop, int64_t size, string type_, int32_t device_)
    : current_context(context),
      device(device_),
      posx(xpos),
      posy(ypos),
      npts(size),
      mag(mag_),
      lambda(lambda_),
      zp(zerop),
      scale(float(2 * kPi / lambda_)),
      lgs(false),
      G(1.f),
      thetaML(0.f),
      dx(0.f),
      dy(0.f),
      type(type_),
      lazy_raytrace(true) {
  context->set_active_device(device, 1);
  d_phase = new SutraPhase(context, size);
  d_lgs = nullptr;
}

SutraSource::~SutraSource() {
  del(d_phase);
  del(d_lgs);
}

// ============================================================ SutraLGS
// reference: libsutra/src.cpp/sutra_lgs.cpp:26-250
SutraLGS::SutraLGS(CarmaContext* context, CarmaObj<float>* lgskern,
                   CarmaObj<hipFloatComplex>* ftlgskern, PlanMap& ftlgskern_plans,
                   int64_t nvalid_, int64_t npix_, int64_t nmaxhr_)
    : device(context->get_active_device()),
      nvalid(nvalid_),
      npix(npix_),
      nmaxhr(nmaxhr_),
      nprof(0),
      hg(0.f),
      h0(0.f),
      deltah(0.f),
      pixsize(0.f),
      ready(false),
      d_doffaxis(nullptr),
      d_azimuth(nullptr),
      d_prof1d(nullptr),
      d_profcum(nullptr),
      current_context(context) {
  (void)lgskern;
  (void)ftlgskern;
  (void)ftlgskern_plans;
  int64_t d1[2] = {1, npix};
  d_beam = new CarmaObj<float>(context, d1);
  d_ftbeam = new CarmaObj<hipFloatComplex>(context, d1);
  int64_t dv[2] = {1, nvalid};
  d_azimuth = new CarmaObj<float>(context, dv);
  int64_t d2[3] = {2, npix, nvalid};
  d_prof2d = new CarmaObj<hipFloatComplex>(context, d2);
  int64_t d3[4] = {3, npix, npix, nvalid};
  d_lgskern = new CarmaObj<float>(context, d3);
  d_ftlgskern = new CarmaObj<hipFloatComplex>(context, d3);
}

SutraLGS::~SutraLGS() {
  del(d_doffaxis);
  del(d_azimuth);
  del(d_prof1d);
  del(d_profcum);
  del(d_prof2d);
  del(d_beam);
  del(d_ftbeam);
  del(d_lgskern);
  del(d_ftlgskern);
}

int32_t SutraLGS::lgs_init(int32_t nprof_, float hg_, float h0_, float deltah_, float pixsize_,
                           float* doffaxis, float* prof1d, float* profcum, float* beam,
                           hipFloatComplex* ftbeam, float* azimuth) {
  if (nprof_ < 2) throw std::runtime_error("lgs_init: the profile needs at least 2 samples");
  current_context->set_active_device(device, 1);
  nprof = nprof_;
  hg = hg_;
  h0 = h0_;
  deltah = deltah_;
  pixsize = pixsize_;
  del(d_doffaxis);
  del(d_prof1d);
  del(d_profcum);
  int64_t dv[2] = {1, nvalid};
  d_doffaxis = new CarmaObj<float>(current_context, dv, doffaxis);
  // nprof + 1 values as the reference allocates; the profile gives nprof
  // (the last one is never read), the cumulative sum nprof + 1
  int64_t dp[2] = {1, (int64_t)nprof + 1};
  std::vector<float> p1(prof1d, prof1d + nprof);
  p1.push_back(0.0f);
  d_prof1d = new CarmaObj<float>(current_context, dp, p1.data());
  d_profcum = new CarmaObj<float>(current_context, dp, profcum);
  d_beam->host2device(beam);
  d_ftbeam->host2device(ftbeam);
  d_azimuth->host2device(azimuth);
  return lgs_update();
}

int32_t SutraLGS::load_prof(float* prof1d, float* profcum, float hg_, float h0_, float deltah_) {
  if (d_prof1d == nullptr) throw std::runtime_error("load_prof: lgs_init has not been called");
  current_context->set_active_device(device, 1);
  std::vector<float> p1(prof1d, prof1d + nprof);
  p1.push_back(0.0f);
  d_prof1d->host2device(p1.data());
  d_profcum->host2device(profcum);
  hg = hg_;
  h0 = h0_;
  deltah = deltah_;
  return EXIT_SUCCESS;
}

int32_t SutraLGS::lgs_update(CarmaDevice* dev, hipStream_t stream) {
  (void)dev;
  if (d_prof1d == nullptr) throw std::runtime_error("lgs_update: lgs_init has not been called");
  current_context->set_active_device(device, 1);
  shwfs_throw_if(shwfs_lgs_prof2d(reinterpret_cast<float*>(d_prof2d->get_data()),
                                  d_prof1d->get_data(), d_profcum->get_data(), nprof,
                                  d_doffaxis->get_data(), hg, pixsize, h0, deltah,
                                  reinterpret_cast<const float*>(d_ftbeam->get_data()),
                                  (int32_t)npix, (int32_t)nvalid, stream),
                 "lgs_update");
  return lgs_makespot(nullptr, 0, stream);
}

int32_t SutraLGS::lgs_makespot(CarmaDevice* dev, int32_t nin, hipStream_t stream) {
  (void)dev;
  (void)nin;  // every sub-aperture at once
  if (d_prof1d == nullptr) throw std::runtime_error("lgs_makespot: lgs_init has not been called");
  current_context->set_active_device(device, 1);
  shwfs_throw_if(shwfs_lgs_kernels(reinterpret_cast<float*>(d_ftlgskern->get_data()),
                                   d_lgskern->get_data(),
                                   reinterpret_cast<const float*>(d_prof2d->get_data()),
                                   d_beam->get_data(), d_azimuth->get_data(), (int32_t)npix,
                                   (int32_t)nvalid, stream),
                 "lgs_makespot");
  ready = true;
  return EXIT_SUCCESS;
}

int32_t SutraLGS::load_kernels(float* h_lgskern, CarmaDevice* dev) {
  (void)dev;
  current_context->set_active_device(device, 1);
  d_lgskern->host2device(h_lgskern);
  shwfs_throw_if(shwfs_lgs_kernels(reinterpret_cast<float*>(d_ftlgskern->get_data()),
                                   d_lgskern->get_data(), nullptr, nullptr, nullptr,
                                   (int32_t)npix, (int32_t)nvalid, nullptr),
                 "load_kernels");
  ready = true;
  return EXIT_SUCCESS;
}

int32_t SutraSource::set_phase(float* data) {
  pending.clear();
  return d_phase->d_screen->host2device(data);
}
int32_t SutraSource::reset_phase() {
  pending.clear();
  return d_phase->d_screen->reset();
}

// reference: libsutra/src.cpp/sutra_source.cpp:206-220
int32_t SutraSource::add_layer(string ltype, int32_t idx, float mxoff, float myoff) {
  current_context->set_active_device(device, 1);
  xoff[make_pair(ltype, idx)] = mxoff;
  yoff[make_pair(ltype, idx)] = myoff;
  return EXIT_SUCCESS;
}

int32_t SutraSource::remove_layer(string ltype, int32_t idx) {
  current_context->set_active_device(device, 1);
  xoff.erase(make_pair(ltype, idx));
  yoff.erase(make_pair(ltype, idx));
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_source.cpp:222-267.  The layers are the
// "atmos" entries of the offset maps, in map order, each at its offset plus
// the layer's accumulated shift; G / thetaML / dx / dy are the source's.
int32_t SutraSource::raytrace(SutraAtmos* atmos, bool async) {
  (void)async;  // one launch either way
  if (atmos == nullptr) throw std::invalid_argument("raytrace: atmos is None");
  if (lgs && (d_lgs == nullptr || d_lgs->hg == 0.0f))
    throw std::runtime_error("raytrace: LGS source without lgs_init (its altitude is unknown)");
  current_context->set_active_device(device, 1);
  vector<shwfs_layer> layers;
  for (auto& kv : xoff) {
    if (kv.first.first.find("atmos") != 0) continue;
    const int32_t idx = kv.first.second;
    if (idx < 0 || idx >= (int32_t)atmos->d_screens.size())
      throw std::out_of_range("raytrace: no atmosphere layer of that index");
    SutraTurbuScreen* ps = atmos->d_screens[idx];
    CarmaObj<float>* scr = ps->d_tscreen->d_screen;
    shwfs_layer l{};
    l.d_screen = scr->get_data();
    l.dim = (int32_t)scr->get_dims(1);
    l.xoff = kv.second + ps->accumx;
    l.yoff = yoff[kv.first] + ps->accumy;
    l.G = G;
    l.thetaML = thetaML;
    l.dx = dx;
    l.dy = dy;
    // cone compression of the layer seen from the LGS (sutra_source.cpp:250-252)
    l.delta = lgs ? 1.0f - ps->altitude / d_lgs->hg : 1.0f;
    layers.push_back(l);
  }
  if ((int32_t)layers.size() > SHWFS_MAX_LAYERS)
    throw std::runtime_error("raytrace: more layers than SHWFS_MAX_LAYERS");
  pending = layers;
  if (!lazy_raytrace || pending.empty()) {
    if (pending.empty()) return d_phase->d_screen->reset();  // no layer: the screen is zero
    return render_layers();
  }
  return EXIT_SUCCESS;
}

int32_t SutraSource::render_layers(hipStream_t stream) {
  if (pending.empty()) return EXIT_SUCCESS;
  current_context->set_active_device(device, 1);
  CarmaObj<float>* scr = d_phase->d_screen;
  shwfs_throw_if(shwfs_raytrace(scr->get_data(), (int32_t)scr->get_dims(1),
                                (int32_t)scr->get_dims(2), pending.data(),
                                (int32_t)pending.size(), 0, stream),
                 "raytrace");
  pending.clear();
  return EXIT_SUCCESS;
}

// ============================================================ atmosphere layers
// reference: libsutra/src.cpp/sutra_tscreen.cpp:25-80 (the holder part)
SutraTurbuScreen::SutraTurbuScreen(CarmaContext* context, int64_t size, int64_t size2,
                                   float amplitude_, float altitude_, float windspeed_,
                                   float winddir_, float deltax_, float deltay_, int32_t device_)
    : device(device_),
      screen_size(size),
      r0(0.f),
      amplitude(amplitude_),
      altitude(altitude_),
      windspeed(windspeed_),
      winddir(winddir_),
      deltax(deltax_),
      deltay(deltay_),
      accumx(0.f),
      accumy(0.f),
      current_context(context) {
  (void)size2;  // stencil size of the extrusion, not used here
  if (size < 2) throw std::invalid_argument("TurbuScreen: size must be at least 2");
  current_context->set_active_device(device, 1);
  d_tscreen = new SutraPhase(context, size);
  d_tscreen->d_screen->reset();
}
SutraTurbuScreen::~SutraTurbuScreen() { del(d_tscreen); }

// reference: libsutra/src.cpp/sutra_atmos.cpp:24-58
SutraAtmos::SutraAtmos(CarmaContext* context, int32_t nscreens_, float global_r0,
                       float* r0_per_layer, int64_t* size1, int64_t* size2, float* altitude,
                       float* windspeed, float* winddir, float* deltax, float* deltay,
                       int32_t device)
    : nscreens(nscreens_), r0(global_r0), current_context(context) {
  for (int32_t i = 0; i < nscreens; ++i) {
    // amplitude r0^(-5/6) of the layer (r0 in pixels), as the reference
    const float amp = std::pow(r0_per_layer[i], -5.0f / 6.0f);
    d_screens.push_back(new SutraTurbuScreen(context, size1[i], size2[i], amp, altitude[i],
                                             windspeed[i], winddir[i], deltax[i], deltay[i],
                                             device));
    d_screens.back()->r0 = r0_per_layer[i];
  }
}
SutraAtmos::~SutraAtmos() {
  for (auto* s : d_screens) delete s;
}

// ============================================================ SutraWfs
// reference: libsutra/src.cpp/sutra_wfs.cpp:25-96
SutraWfs::SutraWfs(CarmaContext* context, SutraTelescope* d_tel,
                   CarmaObj<hipFloatComplex>* camplipup, CarmaObj<hipFloatComplex>* camplifoc,
                   CarmaObj<hipFloatComplex>* fttotim, string type_, int64_t nxsub_,
                   int64_t nvalid_, int64_t npix_, int64_t nphase_, int64_t nrebin_, int64_t nfft_,
                   int64_t ntot_, int64_t npup_, float pdiam, float nphotons, float nphot4imat_,
                   int32_t lgs_, bool fakecam_, int32_t max_flux_per_pix_, int32_t max_pix_value_,
                   bool is_low_order_, bool roket_, int32_t device_)
    : device(device_),
      type(type_),
      nxsub(nxsub_),
      nvalid(nvalid_),
      npix(npix_),
      nrebin(nrebin_),
      nfft(nfft_),
      ntot(ntot_),
      npup(npup_),
      nphase(nphase_),
      nmaxhr(nvalid_),
      nffthr(1),
      subapd(pdiam),
      nphot(nphotons),
      nphot4imat(nphot4imat_),
      noise(0),
      lgs(lgs_ != 0),
      kernconv(false),
      roket(roket_),
      is_low_order(is_low_order_),
      fakecam(fakecam_),
      max_flux_per_pix(max_flux_per_pix_),
      max_pix_value(max_pix_value_),
      campli_plan(nullptr),
      fttotim_plan(nullptr),
      d_ftkernel(nullptr),
      d_camplipup(nullptr),
      d_camplifoc(nullptr),
      d_fttotim(nullptr),
      d_pupil(d_tel->d_pupil_m),
      d_bincube(nullptr),
      d_binimg(nullptr),
      d_binimg_notnoisy(nullptr),
      d_intensities(nullptr),
      d_offsets(nullptr),
      d_fluxPerSub(nullptr),
      d_sincar(nullptr),
      d_hrmap(nullptr),
      d_camimg(nullptr),
      d_dark(nullptr),
      d_flat(nullptr),
      d_slopes(nullptr),
      d_gs(nullptr),
      d_phasemap(nullptr),
      d_ttprojmat(nullptr),
      d_ttprojvec(nullptr),
      d_validsubsx(nullptr),
      d_validsubsy(nullptr),
      d_submask(nullptr),
      current_context(context),
      offset(0),
      nvalid_tot((int32_t)nvalid_),
      rank(0),
      stream(nullptr) {
  if (camplipup != nullptr && !is_low_order) d_camplipup = camplipup;
  if (camplifoc != nullptr && !is_low_order) d_camplifoc = camplifoc;
  if (fttotim != nullptr && !is_low_order) d_fttotim = fttotim;
}

SutraWfs::~SutraWfs() {}

int32_t SutraWfs::wfs_initgs(float xpos, float ypos, float lambda, float mag, float zerop,
                             int64_t size, float noise_, int64_t seed, float G, float thetaML,
                             float dx, float dy) {
  current_context->set_active_device(device, 1);
  d_gs = new SutraSource(current_context, xpos, ypos, lambda, mag, zerop, size, "wfs", device);
  d_gs->G = G;
  d_gs->thetaML = thetaML;
  d_gs->dx = dx;
  d_gs->dy = dy;
  if (lgs) {  // reference: libsutra/src.cpp/sutra_wfs.cpp:114-119
    if (ntot != 16 && ntot != 32 && ntot != 64)
      throw std::runtime_error("LGS WFS: the HR support ntot must be 16, 32 or 64");
    PlanMap none;
    d_gs->d_lgs = new SutraLGS(current_context, nullptr, nullptr, none, nvalid, ntot, nmaxhr);
    d_gs->lgs = true;
  }
  return set_noise(noise_, seed);
}

// reference: libsutra/src.cpp/sutra_wfs.cpp:123-141 (init_prng(seed) of bincube and,
// for readout noise, of binimg; here one Philox key per WFS, call counter reset).
int32_t SutraWfs::set_noise(float noise_, int64_t seed) {
  current_context->set_active_device(device, 1);
  noise = noise_;
  if (noise > -1 && d_bincube) d_bincube->init_prng(seed);
  if (noise > 0 && d_binimg) d_binimg->init_prng(seed);
  auto* sh = dynamic_cast<SutraWfsSH*>(this);
  if (sh) {
    sh->noise_seed = (uint64_t)seed;
    sh->noise_call = 0;
  }
  return EXIT_SUCCESS;
}

int32_t SutraWfs::set_pupil(float* pupil) {
  current_context->set_active_device(device, 1);
  int32_t rc = d_pupil->host2device(pupil);
  pupil_changed();
  return rc;
}

// reference: libsutra/src.cpp/sutra_wfs.cpp:253-261 (there a wrong size only
// prints; here it raises with the error of shwfs_upload)
int32_t SutraWfs::set_binimg(float* binimg, int32_t nElem) {
  current_context->set_active_device(device, 1);
  shwfs_throw_if(shwfs_upload(d_binimg->get_data(), d_binimg->get_nb_elements(), binimg, nElem,
                              (int32_t)sizeof(float), nullptr),
                 "set_binimg");
  return EXIT_SUCCESS;
}

int32_t SutraWfs::set_dark(float* dark, int32_t nElem) {
  if (d_dark == nullptr) {
    std::cout << "WFS d_dark array has not been initialized yet" << std::endl;
    return EXIT_FAILURE;
  }
  if (nElem == d_dark->get_nb_elements()) d_dark->host2device(dark);
  return EXIT_SUCCESS;
}

int32_t SutraWfs::set_flat(float* flat, int32_t nElem) {
  if (d_flat == nullptr) {
    std::cout << "WFS d_flat array has not been initialized yet" << std::endl;
    return EXIT_FAILURE;
  }
  if (nElem == d_flat->get_nb_elements()) d_flat->host2device(flat);
  return EXIT_SUCCESS;
}

int32_t SutraWfs::set_fakecam(bool f) {
  fakecam = f;
  if (f) alloc_camera();
  return EXIT_SUCCESS;
}

// Camera image (uint16), dark (0) and flat (1) with the detector image's
// extents, as the reference allocates them for fakecam
// (libsutra/src.cpp/sutra_wfs_sh.cpp:96-102); no-op before the detector image
// exists (allocate_buffers calls it again) or once allocated.
void SutraWfs::alloc_camera() {
  if (d_binimg == nullptr || d_camimg != nullptr) return;
  current_context->set_active_device(device, 1);
  d_camimg = new CarmaObj<uint16_t>(current_context, d_binimg->get_dims());
  d_dark = new CarmaObj<float>(current_context, d_binimg->get_dims());
  d_dark->reset();
  d_flat = new CarmaObj<float>(current_context, d_binimg->get_dims());
  d_flat->memset(1.0f);
}

// The camera image from the detector image (fakecam; reference:
// libsutra/src.cpp/sutra_wfs_sh.cpp:635-640).
int32_t SutraWfs::digitalize_image() {
  if (!fakecam) return EXIT_SUCCESS;
  alloc_camera();
  if (d_camimg == nullptr) throw std::runtime_error("fakecam: the detector image is not allocated");
  shwfs_throw_if(shwfs_digitalize(d_camimg->get_data(), d_binimg->get_data(), d_dark->get_data(),
                                  d_flat->get_data(), max_flux_per_pix, max_pix_value,
                                  d_binimg->get_nb_elements(), nullptr),
                 "digitalize");
  return EXIT_SUCCESS;
}
int32_t SutraWfs::set_max_flux_per_pix(int32_t v) {
  max_flux_per_pix = v;
  return EXIT_SUCCESS;
}
int32_t SutraWfs::set_max_pix_value(int32_t v) {
  max_pix_value = v;
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_wfs.cpp:199-245
int32_t SutraWfs::slopes_geom(float* slopes, int32_t type_) {
  current_context->set_active_device(device, 1);
  if (!d_phasemap) throw std::runtime_error("slopes_geom: arrays not loaded");
  d_gs->render_layers(stream);
  float alpha = float(0.206265 / subapd * nphase / (nphase - 1));
  shwfs_throw_if(shwfs_slopes_geom(type_, d_gs->d_phase->d_screen->get_data(), d_pupil->get_data(),
                                   d_fluxPerSub->get_data(), slopes, d_phasemap->get_data(),
                                   d_ttprojmat->get_data(), (int32_t)nphase, (int32_t)nvalid,
                                   alpha, stream),
                 "slopes_geom");
  return EXIT_SUCCESS;
}

int32_t SutraWfs::slopes_geom(int32_t type_) { return slopes_geom(d_slopes->get_data(), type_); }

// ============================================================ SutraWfsSH
SutraWfsSH::SutraWfsSH(CarmaContext* context, SutraTelescope* d_tel,
                       CarmaObj<hipFloatComplex>* camplipup, CarmaObj<hipFloatComplex>* camplifoc,
                       CarmaObj<hipFloatComplex>* fttotim, int64_t nxsub_, int64_t nvalid_,
                       int64_t npix_, int64_t nphase_, int64_t nrebin_, int64_t nfft_,
                       int64_t ntot_, int64_t npup_, float pdiam, float nphotons,
                       float nphot4imat_, int32_t lgs_, bool fakecam_, int32_t max_flux_per_pix_,
                       int32_t max_pix_value_, bool is_low_order_, bool roket_, int32_t device_)
    : SutraWfs(context, d_tel, camplipup, camplifoc, fttotim, "sh", nxsub_, nvalid_, npix_,
               nphase_, nrebin_, nfft_, ntot_, npup_, pdiam, nphotons, nphot4imat_, lgs_,
               fakecam_, max_flux_per_pix_, max_pix_value_, is_low_order_, roket_, device_),
      d_binmap(nullptr),
      d_validpuppixx(nullptr),
      d_validpuppixy(nullptr),
      d_fsamplipup(nullptr),
      d_fsamplifoc(nullptr),
      d_fsphase(nullptr),
      d_fsampl(nullptr),
      fsampli_plan(nullptr),
      noise_seed(1234),
      noise_call(0),
      arrays_loaded(false),
      own_plan(false),
      bin_block(0),
      bin_origin(0),
      d_pupil_unit(nullptr),
      d_bgpix(nullptr),
      nbgpix(0),
      noise_this_call(true) {}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:48-75 (vestigial MPI split;
// kept for API completeness, the kernels always see the local nvalid).
int32_t SutraWfsSH::define_mpi_rank(int32_t rank_, int32_t size) {
  if (size <= 0) throw std::runtime_error("define_mpi_rank: size must be positive");
  rank = rank_;
  int32_t r = (int32_t)(nvalid % size);
  if (rank_ < r) {
    nvalid = nvalid / size + 1;
    offset = rank_ * (int32_t)nvalid;
  } else {
    nvalid = nvalid / size;
    offset = rank_ * (int32_t)nvalid + r;
  }
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:77-271
int32_t SutraWfsSH::allocate_buffers(PlanMap& campli_plans, PlanMap& fttotim_plans) {
  (void)fttotim_plans;
  current_context->set_active_device(device, 1);
  if (ntot < nfft) throw std::runtime_error("allocate_buffers: ntot must be >= nfft");
  int64_t d[5];
  if (rank == 0) {
    d_binimg = new CarmaObj<float>(current_context, dims2(d, npix * nxsub, npix * nxsub));
    if (roket) d_binimg_notnoisy = new CarmaObj<float>(current_context, dims2(d, npix * nxsub, npix * nxsub));
    if (fakecam) alloc_camera();
  }
  if (d_camplipup == nullptr)
    d_camplipup = new CarmaObj<hipFloatComplex>(current_context, dims3(d, nfft, nfft, nvalid));
  if (d_camplifoc == nullptr)
    d_camplifoc = new CarmaObj<hipFloatComplex>(current_context, dims3(d, nfft, nfft, nvalid));

  vector<int32_t> vdims{(int32_t)nfft, (int32_t)nfft, (int32_t)nvalid};
  auto it = campli_plans.find(vdims);
  if (it == campli_plans.end()) {
    ShwfsPlan* plan = new ShwfsPlan();
    plan->nfft = (int32_t)nfft;
    plan->nvalid = (int32_t)nvalid;
    plan->method = SHWFS_METHOD_FFT_FP32;
    campli_plans[vdims] = plan;
    campli_plan = plan;
  } else {
    campli_plan = it->second;
  }

  d_bincube = new CarmaObj<float>(current_context, dims3(d, npix, npix, nvalid_tot));
  d_slopes = new CarmaObj<float>(current_context, dims1(d, 2 * nvalid));
  d_ftkernel = new CarmaObj<hipFloatComplex>(current_context, dims2(d, ntot, ntot));
  // enlarged HR grid: the ntot-grid index of each nfft-grid element
  // (reference: libsutra/src.cpp/sutra_wfs_sh.cpp:196-204)
  if (ntot != nfft) d_hrmap = new CarmaObj<int32_t>(current_context, dims1(d, nfft * nfft));
  d_offsets = new CarmaObj<float>(current_context, dims2(d, nphase, nphase));
  d_binmap = new CarmaObj<int32_t>(current_context, dims2(d, nrebin * nrebin, npix * npix));
  d_intensities = new CarmaObj<float>(current_context, dims1(d, nvalid_tot));
  d_fluxPerSub = new CarmaObj<float>(current_context, dims1(d, nvalid_tot));
  d_validsubsx = new CarmaObj<int32_t>(current_context, dims1(d, nvalid_tot));
  d_validsubsy = new CarmaObj<int32_t>(current_context, dims1(d, nvalid_tot));
  d_validpuppixx = new CarmaObj<int32_t>(current_context, dims1(d, nvalid_tot));
  d_validpuppixy = new CarmaObj<int32_t>(current_context, dims1(d, nvalid_tot));
  d_phasemap = new CarmaObj<int32_t>(current_context, dims2(d, nphase * nphase, nvalid));
  d_ttprojmat = new CarmaObj<float>(current_context, dims2(d, nphase * nphase, nvalid * 2));
  d_ttprojvec = new CarmaObj<float>(current_context, dims1(d, nphase * nphase * nvalid));
  return EXIT_SUCCESS;
}

SutraWfsSH::~SutraWfsSH() {
  current_context->set_active_device(device, 1);
  del(d_pupil_unit);
  del(d_bgpix);
  del(d_ftkernel);
  del(d_bincube);
  del(d_binimg);
  del(d_binimg_notnoisy);
  del(d_intensities);
  del(d_offsets);
  del(d_fluxPerSub);
  del(d_sincar);
  del(d_hrmap);
  del(d_slopes);
  del(d_phasemap);
  del(d_ttprojmat);
  del(d_ttprojvec);
  del(d_binmap);
  del(d_validsubsx);
  del(d_validsubsy);
  del(d_validpuppixx);
  del(d_validpuppixy);
  del(d_submask);
  del(d_fsamplipup);
  del(d_fsamplifoc);
  del(d_fsphase);
  del(d_fsampl);
  del(d_camimg);
  del(d_dark);
  del(d_flat);
  del(d_gs);
  // shared scratch (d_camplipup/d_camplifoc/d_fttotim) belongs to SutraSensors
  // unless this WFS created it
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:308-331
int32_t SutraWfsSH::load_arrays(int32_t* phasemap, int32_t* hrmap, int32_t* binmap,
                                float* offsets, float* fluxPerSub, int32_t* validsubsx,
                                int32_t* validsubsy, int32_t* istart, int32_t* jstart,
                                float* ttprojmat, hipFloatComplex* kernel) {
  if (d_bincube == nullptr)
    throw std::runtime_error("ERROR : d_bincube not initialized, did you do the allocate_buffers?");
  current_context->set_active_device(device, 1);
  // host-side bounds checks of every table the kernels index with: a
  // sub-aperture window must lie in the screen, a spot in the image, a binmap
  // entry in the HR grid.
  const int64_t nsub = npix * nxsub;
  for (int64_t s = 0; s < nvalid; ++s) {
    const int64_t i0 = istart[s], j0 = jstart[s];
    if (i0 < 0 || j0 < 0 || i0 + nphase > npup || j0 + nphase > npup)
      throw std::runtime_error("load_arrays: sub-aperture window outside the phase screen");
    const int64_t vx = validsubsx[offset + s], vy = validsubsy[offset + s];
    if (vx < 0 || vy < 0 || vx + npix > nsub || vy + npix > nsub)
      throw std::runtime_error("load_arrays: spot window outside the detector image");
  }
  for (int64_t i = 0; i < nrebin * nrebin * npix * npix; ++i)
    if (binmap[i] < 0 || binmap[i] >= ntot * ntot)
      throw std::runtime_error("load_arrays: binmap entry outside the HR grid");
  if (ntot != nfft) {
    for (int64_t i = 0; i < nfft * nfft; ++i)
      if (hrmap[i] < 0 || hrmap[i] >= ntot * ntot)
        throw std::runtime_error("load_arrays: hrmap entry outside the enlarged HR grid");
    d_hrmap->host2device(hrmap);  // reference: sutra_wfs_sh.cpp:323
  }
  for (int64_t i = 0; i < nphase * nphase * nvalid; ++i)
    if (phasemap[offset * nphase * nphase + i] < 0 || phasemap[offset * nphase * nphase + i] >= npup * npup)
      throw std::runtime_error("load_arrays: phasemap entry outside the phase screen");
  d_phasemap->host2device(&phasemap[offset * nphase * nphase]);
  d_ttprojmat->host2device(ttprojmat);
  d_offsets->host2device(offsets);
  d_binmap->host2device(binmap);
  d_fluxPerSub->host2device(&fluxPerSub[offset]);
  d_validsubsx->host2device(&validsubsx[offset]);
  d_validsubsy->host2device(&validsubsy[offset]);
  d_validpuppixx->host2device(istart);
  d_validpuppixy->host2device(jstart);
  d_ftkernel->host2device(kernel);
  // the image pixels no spot window covers (reference: fill_binimage draws
  // the background on every pixel, then fillbinimg overwrites the windows)
  {
    std::vector<uint8_t> covered((size_t)(nsub * nsub), 0);
    for (int64_t s = 0; s < nvalid; ++s) {
      const int64_t vx = validsubsx[offset + s], vy = validsubsy[offset + s];
      for (int64_t y = vy; y < vy + npix; ++y)
        for (int64_t x = vx; x < vx + npix; ++x) covered[(size_t)(x + y * nsub)] = 1;
    }
    std::vector<int32_t> bg;
    for (int64_t i = 0; i < nsub * nsub; ++i)
      if (!covered[(size_t)i]) bg.push_back((int32_t)i);
    nbgpix = (int64_t)bg.size();
    if (bg.empty()) bg.push_back(0);  // a valid pointer for an empty list
    del(d_bgpix);
    int64_t d[2] = {1, (int64_t)bg.size()};
    d_bgpix = new CarmaObj<int32_t>(current_context, d, bg.data());
  }
  int32_t origin = 0, rc = 0;
  if (ntot == nfft) {  // the block form is a property of the nfft grid
    rc = shwfs_binmap_block_origin(binmap, (int32_t)nfft, (int32_t)npix, (int32_t)nrebin, &origin);
    if (rc < 0) shwfs_throw_if(rc, "load_arrays");
  }
  bin_block = rc;
  bin_origin = rc ? origin : 0;
  arrays_loaded = true;
  pupil_changed();
  return EXIT_SUCCESS;
}

// fully illuminated sub-apertures skip the pupil read in the spot kernels
void SutraWfsSH::pupil_changed() {
  if (!arrays_loaded) return;
  current_context->set_active_device(device, 1);
  if (d_pupil_unit == nullptr) {
    int64_t d[2] = {1, nvalid};
    d_pupil_unit = new CarmaObj<int32_t>(current_context, d);
  }
  shwfs_sh_geom g = geom();
  shwfs_throw_if(shwfs_sh_unit_pupil(&g, d_pupil->get_data(), d_pupil_unit->get_data(), stream),
                 "pupil analysis");
}

shwfs_sh_geom SutraWfsSH::geom() const {
  shwfs_sh_geom g{};
  g.nvalid = (int32_t)nvalid;
  g.nphase = (int32_t)nphase;
  g.npup = (int32_t)d_pupil->get_dims(1);
  g.nfft = (int32_t)nfft;
  g.ntot = (int32_t)ntot;
  g.npix = (int32_t)npix;
  g.nrebin = (int32_t)nrebin;
  g.nxsub = (int32_t)nxsub;
  if (arrays_loaded) {
    g.d_istart = d_validpuppixx->get_data();
    g.d_jstart = d_validpuppixy->get_data();
    g.d_validsubsx = d_validsubsx->get_data();
    g.d_validsubsy = d_validsubsy->get_data();
    g.d_offsets = d_offsets->get_data();
    g.d_fluxPerSub = d_fluxPerSub->get_data();
    g.d_binmap = d_binmap->get_data();
    g.binmap_block = bin_block;
    g.bin_origin = bin_origin;
    g.d_hrmap = d_hrmap ? d_hrmap->get_data() : nullptr;
    g.d_ftkernel = reinterpret_cast<const float*>(d_ftkernel->get_data());
    g.kernconv = kernconv ? 1 : 0;
    // the LGS kernels once built (frames check lgs_ready first)
    SutraLGS* l = (lgs && d_gs) ? d_gs->d_lgs : nullptr;
    g.d_ftlgskern = (l && l->ready) ? reinterpret_cast<const float*>(l->d_ftlgskern->get_data())
                                    : nullptr;
    g.d_bgpix = d_bgpix ? d_bgpix->get_data() : nullptr;
    g.nbgpix = (int32_t)nbgpix;
  }
  return g;
}

shwfs_frame_args SutraWfsSH::frame_args(bool with_noise) const {
  shwfs_frame_args f{};
  f.d_phase = d_gs ? d_gs->d_phase->d_screen->get_data() : nullptr;
  f.d_pupil = d_pupil->get_data();
  f.scale = d_gs ? d_gs->scale : 0.f;
  f.nphot = nphot;
  f.noise = with_noise ? noise : -1.0f;
  f.noise_background = noise;  // fill_binimage uses the WFS noise even for comp_image(false)
  f.seed = noise_seed;
  f.call = noise_call;
  f.method = get_method();
  f.wg_size = get_wg_size();
  f.d_bincube = d_bincube->get_data();
  f.d_binimg = d_binimg ? d_binimg->get_data() : nullptr;
  f.d_intensities = d_intensities->get_data();
  f.centroid = nullptr;
  f.d_pupil_unit = d_pupil_unit ? d_pupil_unit->get_data() : nullptr;
  if (d_submask != nullptr) {  // the stopped field (field_stop_stage) as amplitude and phase
    f.d_phase = d_fsphase->get_data();
    f.d_pupil = d_fsampl->get_data();
    f.d_pupil_unit = nullptr;
  }
  if (layers_in_field_stage()) {
    f.layers = d_gs->pending.data();
    f.nlayers = (int32_t)d_gs->pending.size();
  }
  return f;
}

bool SutraWfsSH::layers_in_field_stage() const {
  const int32_t wg = get_wg_size();
  return d_gs != nullptr && !d_gs->pending.empty() && d_submask == nullptr &&
         get_method() == SHWFS_METHOD_FFT_FP32 && (wg == 0 || wg == 64);
}

void SutraWfsSH::check_lgs_ready() const {
  if (lgs && !(d_gs && d_gs->d_lgs && d_gs->d_lgs->ready))
    throw std::runtime_error("LGS WFS: lgs_init (or load_kernels) has not been called");
}

// reference: the field-stop branch of comp_generic
// (libsutra/src.cpp/sutra_wfs_sh.cpp:366-391)
void SutraWfsSH::field_stop_stage(hipStream_t st) {
  if (d_submask == nullptr) return;
  d_gs->render_layers(st);
  shwfs_throw_if(shwfs_field_stop(d_gs->d_phase->d_screen->get_data(), d_pupil->get_data(),
                                  (int32_t)d_pupil->get_dims(1), d_gs->scale,
                                  d_submask->get_data(), (int32_t)d_submask->get_dims(1),
                                  d_fsamplipup->get_data(), d_fsphase->get_data(),
                                  d_fsampl->get_data(), st),
                 "field stop");
}

// One fused launch: fillcamplipup + FFT + abs2c + fillbincube + subap_reduce +
// subap_norm + prng 'P' / 'N' + (readout background) + fillbinimg
// (reference sequence: libsutra/src.cpp/sutra_wfs_sh.cpp:337-621).
// With roket the reference keeps the spots before the noise in
// d_binimg_notnoisy (572-578): the spot kernel then runs noise-free into that
// image, the noise is drawn on the spot cube with the same Philox draws the
// fused kernel takes (domains 0 and 1, element = spot-cube index), and the
// spots go to d_binimg with the same background (domain 2): the same bincube
// and binimg as the fused launch, bit for bit.
int32_t SutraWfsSH::comp_generic() {
  if (d_bincube == nullptr)
    throw std::runtime_error("ERROR : d_bincube not initialized, did you do the allocate_buffers?");
  if (!arrays_loaded) throw std::runtime_error("comp_image: load_arrays has not been called");
  if (d_gs == nullptr) throw std::runtime_error("comp_image: initgs has not been called");
  if (d_gs->d_phase->d_screen->get_dims(1) != d_pupil->get_dims(1))
    throw std::runtime_error("comp_image: phase screen and pupil sizes differ");
  current_context->set_active_device(device, 1);
  check_lgs_ready();
  if (!layers_in_field_stage()) d_gs->render_layers(stream);
  field_stop_stage(stream);
  shwfs_sh_geom g = geom();
  shwfs_frame_args f = frame_args(noise_this_call);
  if (roket && d_binimg_notnoisy != nullptr && d_binimg != nullptr) {
    shwfs_frame_args clean = f;
    clean.noise = -1.0f;
    clean.noise_background = 0.0f;
    clean.d_binimg = d_binimg_notnoisy->get_data();
    shwfs_throw_if(shwfs_sh_comp_image(&g, &clean, stream), "SutraWfsSH::comp_image (roket)");
    const int64_t n = d_bincube->get_nb_elements();
    if (f.noise > -1.0f)
      shwfs_throw_if(shwfs_prng(d_bincube->get_data(), n, 'P', 0.f, 1.f, f.seed, f.call, 0u, stream),
                     "SutraWfsSH::comp_image (roket, photon noise)");
    if (f.noise > 0.0f)
      shwfs_throw_if(shwfs_prng(d_bincube->get_data(), n, 'N', f.noise, 1.f, f.seed, f.call, 1u,
                                stream),
                     "SutraWfsSH::comp_image (roket, readout noise)");
    scatter_binimg(f.noise_background, f.call, 2u);
  } else {
    shwfs_throw_if(shwfs_sh_comp_image(&g, &f, stream), "SutraWfsSH::comp_image");
  }
  noise_call++;
  return EXIT_SUCCESS;
}

// Background readout noise (noise > 0) and the spots of d_bincube into
// d_binimg (fillbinimg, libsutra/src.cu/sutra_wfs.cu:183-229).
void SutraWfsSH::scatter_binimg(float background, uint32_t call, uint32_t domain) {
  shwfs_throw_if(shwfs_fill_binimg(d_bincube->get_data(), d_binimg->get_data(),
                                   d_validsubsx->get_data(), d_validsubsy->get_data(),
                                   (int32_t)npix, (int32_t)nvalid, (int32_t)(npix * nxsub),
                                   background, noise_seed, call, domain, stream),
                 "fill_binimage");
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:599-621: background readout
// noise when noise > 0, then the spots of d_bincube (e.g. after set_bincube)
// scattered into d_binimg.  The background of this standalone call is drawn
// from its own Philox domain (3) at the current frame counter, so it never
// repeats a frame's own background.  comp_image does not call it: the spot
// kernel writes the image in the same launch.
int32_t SutraWfsSH::fill_binimage(int32_t async) {
  (void)async;
  if (d_binimg == nullptr)
    throw std::runtime_error("ERROR : d_bincube not initialized, did you do the allocate_buffers?");
  if (!arrays_loaded) throw std::runtime_error("fill_binimage: load_arrays has not been called");
  current_context->set_active_device(device, 1);
  scatter_binimg(noise, noise_call, 3u);
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:623-643
int32_t SutraWfsSH::comp_image(bool noise_) {
  current_context->set_active_device(device, 1);
  noise_this_call = noise_;
  int32_t result = comp_generic();  // includes fill_binimage's image (same launch)
  noise_this_call = true;
  if (fakecam) digitalize_image();  // throws on failure
  return result;
}

int32_t SutraWfsSH::comp_nphot(float ittime, float optthroughput, float diam, int32_t nxsub_,
                               float zerop, float gsmag, float lgsreturnperwatt,
                               float laserpower) {
  d_gs->mag = gsmag;
  if (laserpower == 0) {
    if (zerop == 0) zerop = 1e11;
    nphot = zerop * pow(10., (-0.4 * gsmag)) * ittime * optthroughput * (diam / nxsub_) *
            (diam / nxsub_);
  } else {
    nphot = lgsreturnperwatt * laserpower * optthroughput * (diam / nxsub_) * (diam / nxsub_) *
            1e4 * ittime;
  }
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:665-672 (there a wrong size
// only prints; here it raises with the error of shwfs_upload)
int32_t SutraWfsSH::set_bincube(float* bincube, int32_t nElem) {
  current_context->set_active_device(device, 1);
  shwfs_throw_if(shwfs_upload(d_bincube->get_data(), d_bincube->get_nb_elements(), bincube, nElem,
                              (int32_t)sizeof(float), nullptr),
                 "set_bincube");
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_wfs_sh.cpp:674-705.  The transform
// pair needs no plan (the library's own kernels), so campli_plans is not
// touched; the stop applies to every spot method.
int32_t SutraWfsSH::set_field_stop(PlanMap& campli_plans, float* field_stop, int32_t N) {
  (void)campli_plans;
  const int64_t np = d_pupil->get_dims(1);
  if (N < 16 || N > 4096 || (N & (N - 1)) != 0)
    throw std::runtime_error("set_field_stop: N must be a power of two in [16, 4096]");
  if (N < np) throw std::runtime_error("set_field_stop: N smaller than the pupil support");
  if (field_stop == nullptr) throw std::runtime_error("set_field_stop: no mask");
  current_context->set_active_device(device, 1);
  del(d_submask);
  del(d_fsamplipup);
  del(d_fsphase);
  del(d_fsampl);
  int64_t d2[3] = {2, N, N};
  d_submask = new CarmaObj<float>(current_context, d2, field_stop);
  int64_t dw[3] = {2, N, np};  // rows y < npup of the padded field
  d_fsamplipup = new CarmaObj<hipFloatComplex>(current_context, dw);
  int64_t dp[3] = {2, np, np};
  d_fsphase = new CarmaObj<float>(current_context, dp);
  d_fsampl = new CarmaObj<float>(current_context, dp);
  return EXIT_SUCCESS;
}

int32_t SutraWfsSH::set_wg_size(int32_t wg_size) {
  if (wg_size != 0 && wg_size != 64 && wg_size != 128 && wg_size != 256)
    throw std::runtime_error("set_wg_size: 0, 64, 128 or 256");
  if (!campli_plan) throw std::runtime_error("set_wg_size: allocate_buffers has not run");
  if (!own_plan) {
    campli_plan = new ShwfsPlan(*campli_plan);
    own_plan = true;
  }
  campli_plan->wg_size = wg_size;
  return EXIT_SUCCESS;
}

int32_t SutraWfsSH::set_method(int32_t method) {
  if (method != SHWFS_METHOD_FFT_FP32 && method != SHWFS_METHOD_DFT_BF16X3 &&
      method != SHWFS_METHOD_DFT_FP16)
    throw std::runtime_error("set_method: unknown method");
  if (method != SHWFS_METHOD_FFT_FP32 && nfft < 16)
    throw std::runtime_error("set_method: the MFMA DFT path needs nfft >= 16");
  if (!campli_plan) throw std::runtime_error("set_method: allocate_buffers has not run");
  // a plan may be shared by WFS of identical shape: give this WFS its own
  if (!own_plan) {
    campli_plan = new ShwfsPlan(*campli_plan);
    own_plan = true;
  }
  campli_plan->method = method;
  return EXIT_SUCCESS;
}

// ============================================================ SutraSensors
// reference: libsutra/src.cpp/sutra_sensors.cpp:25-135
SutraSensors::SutraSensors(CarmaContext* context, SutraTelescope* d_tel, vector<string> type,
                           int32_t nwfs, int64_t* nxsub, int64_t* nvalid, int64_t* npupils,
                           int64_t* npix, int64_t* nphase, int64_t* nrebin, int64_t* nfft,
                           int64_t* ntot, int64_t* npup, float* pdiam, float* nphot,
                           float* nphot4imat, int32_t* lgs, bool* fakecam,
                           int32_t* max_flux_per_pix, int32_t* max_pix_value, int32_t device_,
                           bool roket_)
    : device(device_),
      roket(roket_),
      current_context(context),
      d_camplipup(nullptr),
      d_camplifoc(nullptr),
      d_fttotim(nullptr),
      d_ftlgskern(nullptr),
      d_lgskern(nullptr) {
  (void)npupils;
  current_context->set_active_device(device, 1);
  if ((int32_t)type.size() < nwfs) throw std::runtime_error("Sensors: fewer types than nwfs");
  if (nwfs > 0 && type[0] == "sh") {
    int64_t maxnfft = nfft[0], maxnvalid = nvalid[0];
    for (int32_t i = 1; i < nwfs; i++) {
      if (type[i] != "shlo") {
        maxnvalid = std::max(maxnvalid, nvalid[i]);
        maxnfft = std::max(maxnfft, nfft[i]);
      }
    }
    int64_t d[5];
    // Shared complex scratch of the reference's API; the fused kernels do not use it.
    d_camplifoc = new CarmaObj<hipFloatComplex>(context, dims3(d, maxnfft, maxnfft, maxnvalid));
    d_camplipup = new CarmaObj<hipFloatComplex>(context, dims3(d, maxnfft, maxnfft, maxnvalid));
  }
  for (int32_t i = 0; i < nwfs; i++) {
    SutraWfs* wfs = nullptr;
    if (type[i] == "sh" || type[i] == "shlo") {
      wfs = new SutraWfsSH(context, d_tel, d_camplipup, d_camplifoc, d_fttotim, nxsub[i],
                           nvalid[i], npix[i], nphase[i], nrebin[i], nfft[i], ntot[i], npup[i],
                           pdiam[i], nphot[i], nphot4imat[i], lgs[i], fakecam[i],
                           max_flux_per_pix[i], max_pix_value[i], type[i] == "shlo", roket,
                           device);
    } else {
      throw std::runtime_error("Sensors: only Shack-Hartmann WFS (\"sh\") are part of this build");
    }
    d_wfs.push_back(wfs);
  }
  allocate_buffers();
}

SutraSensors::~SutraSensors() {
  current_context->set_active_device(device, 1);
  std::vector<ShwfsPlan*> owned;
  for (auto& kv : campli_plans) owned.push_back(kv.second);
  while (!d_wfs.empty()) {
    SutraWfsSH* sh = dynamic_cast<SutraWfsSH*>(d_wfs.back());
    if (sh && sh->own_plan) delete sh->campli_plan;
    if (sh && sh->d_camplipup != d_camplipup) delete sh->d_camplipup;
    if (sh && sh->d_camplifoc != d_camplifoc) delete sh->d_camplifoc;
    delete d_wfs.back();
    d_wfs.pop_back();
  }
  for (auto* p : owned) delete p;
  del(d_camplifoc);
  del(d_camplipup);
  del(d_fttotim);
  del(d_ftlgskern);
  del(d_lgskern);
}

int32_t SutraSensors::allocate_buffers() {
  current_context->set_active_device(device, 1);
  for (auto* w : d_wfs) w->allocate_buffers(campli_plans, fttotim_plans);
  return EXIT_SUCCESS;
}

int32_t SutraSensors::define_mpi_rank(int32_t rank, int32_t size) {
  for (auto* w : d_wfs) w->define_mpi_rank(rank, size);
  return EXIT_SUCCESS;
}

int32_t SutraSensors::initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop,
                             int64_t* size, float* noise, int64_t* seed, float* G, float* thetaML,
                             float* dx, float* dy) {
  current_context->set_active_device(device, 1);
  for (size_t i = 0; i < d_wfs.size(); i++)
    d_wfs[i]->wfs_initgs(xpos[i], ypos[i], lambda[i], mag[i], zerop, size[i], noise[i], seed[i],
                         G[i], thetaML[i], dx[i], dy[i]);
  return EXIT_SUCCESS;
}

int32_t SutraSensors::initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop,
                             int64_t* size, float* noise, float* G, float* thetaML, float* dx,
                             float* dy) {
  for (size_t i = 0; i < d_wfs.size(); i++)
    d_wfs[i]->wfs_initgs(xpos[i], ypos[i], lambda[i], mag[i], zerop, size[i], noise[i],
                         1234 + (int64_t)i, G[i], thetaML[i], dx[i], dy[i]);
  return EXIT_SUCCESS;
}

int32_t SutraSensors::initgs(float* xpos, float* ypos, float* lambda, float* mag, float zerop,
                             int64_t* size, float* G, float* thetaML, float* dx, float* dy) {
  for (size_t i = 0; i < d_wfs.size(); i++)
    d_wfs[i]->wfs_initgs(xpos[i], ypos[i], lambda[i], mag[i], zerop, size[i], -1,
                         1234 + (int64_t)i, G[i], thetaML[i], dx[i], dy[i]);
  return EXIT_SUCCESS;
}

int32_t SutraSensors::set_noise(int32_t nwfs, float noise, int64_t seed) {
  return d_wfs.at(nwfs)->set_noise(noise, seed);
}

int32_t SutraSensors::set_field_stop(int32_t nwfs, float* field_stop, int32_t N) {
  SutraWfsSH* wfs = dynamic_cast<SutraWfsSH*>(d_wfs.at(nwfs));
  if (wfs) return wfs->set_field_stop(campli_plans, field_stop, N);
  return EXIT_SUCCESS;
}

// ============================================================ acquisim
// reference: libsutra/src.cpp/sutra_acquisim.cpp:24-83
SutraAcquisim::SutraAcquisim(SutraSensors* sensors, int32_t wfs_num)
    : device(0), nxsub(0), nvalid(0), npix(0), current_context(nullptr), wfs(nullptr),
      d_validsubsx(nullptr), d_validsubsy(nullptr), own_valid(false) {
  if (sensors == nullptr || wfs_num < 0 || wfs_num >= (int32_t)sensors->d_wfs.size())
    throw std::runtime_error("Acquisim: no such WFS");
  SutraWfs* w = sensors->d_wfs[wfs_num];
  wfs = dynamic_cast<SutraWfsSH*>(w);
  if (wfs == nullptr) throw std::runtime_error("Acquisim: the WFS is not a Shack-Hartmann WFS");
  device = w->device;
  type = w->type;
  nxsub = w->nxsub;
  nvalid = w->nvalid;
  npix = w->npix;
  current_context = w->current_context;
  d_validsubsx = wfs->d_validsubsx;
  d_validsubsy = wfs->d_validsubsy;
}

SutraAcquisim::~SutraAcquisim() {
  if (own_valid) {
    delete d_validsubsx;
    delete d_validsubsy;
  }
}

int32_t SutraAcquisim::set_validsubs(int64_t n, int32_t* validsubsx, int32_t* validsubsy) {
  current_context->set_active_device(device, 1);
  int64_t d[2] = {1, n};
  if (own_valid) {
    delete d_validsubsx;
    delete d_validsubsy;
  }
  d_validsubsx = new CarmaObj<int32_t>(current_context, d, validsubsx);
  d_validsubsy = new CarmaObj<int32_t>(current_context, d, validsubsy);
  own_valid = true;
  nvalid = n;
  return EXIT_SUCCESS;
}

int32_t SutraAcquisim::comp_image(int64_t* dims, float* bimage) {
  return comp_image(dims, bimage, wfs->d_bincube);
}

int32_t SutraAcquisim::comp_image(int64_t* dims, float* bimage, CarmaObj<float>* d_bincube) {
  current_context->set_active_device(device, 1);
  if (d_validsubsx == nullptr || d_bincube == nullptr)
    throw std::runtime_error("Acquisim: valid positions or bincube not set");
  if (d_bincube->get_nb_elements() < nvalid * npix * npix)
    throw std::runtime_error("Acquisim: bincube smaller than nvalid * npix^2");
  // the kernel reads elements below (npix * nxsub)^2 of the image
  if (dims[0] < 1 || dims[1] * (dims[0] >= 2 ? dims[2] : 1) < (npix * nxsub) * (npix * nxsub))
    throw std::runtime_error("Acquisim.comp_image: image smaller than (npix * nxsub)^2");
  CarmaObj<float> img(current_context, dims, bimage);
  shwfs_throw_if(shwfs_fill_bincube(img.get_data(), d_bincube->get_data(),
                                    d_validsubsx->get_data(), d_validsubsy->get_data(),
                                    (int32_t)npix, (int32_t)nvalid, (int32_t)(npix * nxsub),
                                    wfs->stream),
                 "Acquisim.comp_image");
  carma_safe_call(hipStreamSynchronize(wfs->stream));  // before the image copy is freed
  return EXIT_SUCCESS;
}

int32_t SutraAcquisim::comp_image_2D(int64_t* dims, float* bimage, int32_t* num_ssp) {
  current_context->set_active_device(device, 1);
  if (dims[1] * dims[2] < (npix * nxsub) * (npix * nxsub))
    throw std::runtime_error("Acquisim.comp_image_2D: image smaller than (npix * nxsub)^2");
  CarmaObj<float> img(current_context, dims, bimage);
  int64_t d1[2] = {1, nxsub * nxsub};
  CarmaObj<int32_t> num(current_context, d1, num_ssp);
  shwfs_throw_if(shwfs_fill_bincube_2d(img.get_data(), wfs->d_bincube->get_data(), num.get_data(),
                                       (int32_t)npix, (int32_t)nxsub, (int32_t)wfs->nvalid,
                                       wfs->stream),
                 "Acquisim.comp_image_2D");
  carma_safe_call(hipStreamSynchronize(wfs->stream));
  return EXIT_SUCCESS;
}
