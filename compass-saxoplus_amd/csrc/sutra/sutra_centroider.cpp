// CoG-family centroiders and the RTC dispatch of the SH WFS path.
// reference: libsutra/src.cpp/sutra_centroider{,_cog,_wcog,_tcog}.cpp, sutra_rtc.cpp
#include <type_traits>
#include <iostream>

#include "sutra.hpp"

namespace {
template <class T>
void del(T*& p) {
  delete p;
  p = nullptr;
}
}  // namespace

// ============================================================ SutraCentroider
// reference: libsutra/src.cpp/sutra_centroider.cpp:25-88
template <class Tin, class Tout>
SutraCentroider<Tin, Tout>::SutraCentroider(CarmaContext* context, SutraWfs* wfs_,
                                            int64_t nvalid_, float offset_, float scale_,
                                            bool filter_TT_, int32_t device_)
    : device(device_),
      wfs(wfs_),
      nvalid((int32_t)nvalid_),
      nslopes(0),
      npix(0),
      nxsub(0),
      filter_TT(filter_TT_),
      offset(offset_),
      scale(scale_),
      current_context(context),
      d_bincube(nullptr),
      d_intensities(nullptr),
      d_centroids_ref(nullptr),
      d_img(nullptr),
      d_img_raw(nullptr),
      d_dark(nullptr),
      d_flat(nullptr),
      d_lutPix(nullptr),
      d_validx(nullptr),
      d_validy(nullptr),
      d_validMask(nullptr),
      d_centro_filtered(nullptr),
      d_ref_Tip(nullptr),
      d_ref_Tilt(nullptr),
      d_TT_slopes(nullptr) {
  context->set_active_device(device, 1);
  int64_t d[2] = {1, nvalid};
  d_intensities = new CarmaObj<float>(context, d);
  if (wfs != nullptr) d_img = wfs->d_binimg;
  nslopes = 2 * nvalid;
  int64_t d2[2] = {1, nslopes};
  d_centroids_ref = new CarmaObj<Tout>(context, d2);
  if (filter_TT) init_TT_filter();
}

template <class Tin, class Tout>
SutraCentroider<Tin, Tout>::~SutraCentroider() {
  del(d_intensities);
  del(d_centroids_ref);
  if (wfs == nullptr) del(d_img);
  del(d_img_raw);
  del(d_validx);
  del(d_validy);
  del(d_dark);
  del(d_flat);
  del(d_lutPix);
  del(d_bincube);
  del(d_validMask);
  del(d_TT_slopes);
  del(d_centro_filtered);
  del(d_ref_Tip);
  del(d_ref_Tilt);
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_scale(float s) {
  scale = s;
  return EXIT_SUCCESS;
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_offset(float o) {
  offset = o;
  return EXIT_SUCCESS;
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_nxsub(int32_t n) {
  nxsub = n;
  return EXIT_SUCCESS;
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_npix(int32_t n) {
  npix = n;
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_centroider.cpp:90-114
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::init_calib(int32_t n, int32_t m) {
  current_context->set_active_device(device, 1);
  int64_t d[3] = {2, n, m};
  if (d_img == nullptr) d_img = new CarmaObj<float>(current_context, d);
  if (d_dark == nullptr) {
    d_dark = new CarmaObj<float>(current_context, d);
    d_dark->reset();
  }
  if (d_flat == nullptr) {
    d_flat = new CarmaObj<float>(current_context, d);
    d_flat->memset(1.f);
  }
  if (d_lutPix == nullptr) {
    int64_t d1[2] = {1, (int64_t)n * m};
    d_lutPix = new CarmaObj<int32_t>(current_context, d1);
    std::vector<int32_t> h((size_t)n * m);
    for (int32_t i = 0; i < n * m; ++i) h[i] = i;
    d_lutPix->host2device(h.data());
  }
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::init_roi(int32_t N) {
  current_context->set_active_device(device, 1);
  if (d_validx == nullptr) {
    int64_t d[2] = {1, N};
    d_validx = new CarmaObj<int32_t>(current_context, d);
    d_validy = new CarmaObj<int32_t>(current_context, d);
  }
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_dark(float* dark, int32_t n) {
  current_context->set_active_device(device, 1);
  if (d_dark == nullptr) {
    int64_t d[3] = {2, n, n};
    d_dark = new CarmaObj<float>(current_context, d);
  }
  return d_dark->host2device(dark);
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_flat(float* flat, int32_t n) {
  current_context->set_active_device(device, 1);
  if (d_flat == nullptr) {
    int64_t d[3] = {2, n, n};
    d_flat = new CarmaObj<float>(current_context, d);
  }
  return d_flat->host2device(flat);
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_lutPix(int32_t* lutPix, int32_t n) {
  current_context->set_active_device(device, 1);
  if (d_lutPix == nullptr) {
    int64_t d[2] = {1, n};
    d_lutPix = new CarmaObj<int32_t>(current_context, d);
  }
  return d_lutPix->host2device(lutPix);
}

// reference: libsutra/src.cpp/sutra_centroider.cpp:226-244 (img = (raw[lut] - dark) * flat)
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::calibrate_img(hipStream_t stream) {
  current_context->set_active_device(device, 1);
  if (d_img_raw == nullptr) {
    std::cout << "Image not initialized\n" << std::endl;
    return EXIT_FAILURE;
  }
  init_calib((int32_t)d_img_raw->get_dims(1), (int32_t)d_img_raw->get_dims(2));
  if (d_img->get_nb_elements() != d_img_raw->get_nb_elements())
    throw std::runtime_error("calibrate_img: raw and calibrated images differ in size");
  if constexpr (std::is_same<Tin, uint16_t>::value)
    shwfs_throw_if(shwfs_calibrate_u16(d_img_raw->get_data(), d_img->get_data(),
                                       d_dark->get_data(), d_flat->get_data(),
                                       d_lutPix->get_data(), d_img->get_nb_elements(), stream),
                   "calibrate_img");
  else
    shwfs_throw_if(shwfs_calibrate(d_img_raw->get_data(), d_img->get_data(), d_dark->get_data(),
                                   d_flat->get_data(), d_lutPix->get_data(),
                                   d_img->get_nb_elements(), stream),
                   "calibrate_img");
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::load_img(CarmaObj<Tin>* img) {
  return load_img(img->get_data(), (int32_t)img->get_dims(1), img->get_device());
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::load_img(Tin* img, int32_t n) {
  return load_img(img, n, -1);
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::load_img(Tin* img, int32_t n, int32_t location) {
  return load_img(img, n, n, location);
}
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::load_img(Tin* img, int32_t m, int32_t n, int32_t location) {
  init_img_raw(m, n);
  if (d_img_raw->get_dims(1) != m || d_img_raw->get_dims(2) != n)
    throw std::runtime_error("load_img: image size differs from the first loaded image");
  if (location < 0)
    d_img_raw->host2device(img);
  else
    d_img_raw->copy_from(img, d_img_raw->get_nb_elements());
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::init_img_raw(int32_t m, int32_t n) {
  current_context->set_active_device(device, 1);
  if (d_img_raw == nullptr) {
    int64_t d[3] = {2, m, n};
    d_img_raw = new CarmaObj<Tin>(current_context, d);
  }
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::load_validpos(int32_t* ivalid, int32_t* jvalid, int32_t N) {
  current_context->set_active_device(device, 1);
  if (d_validx == nullptr) init_roi(N);
  if (d_validx->get_nb_elements() < N)
    throw std::runtime_error("load_validpos: more positions than the centroider holds");
  d_validx->host2device(ivalid);
  d_validy->host2device(jvalid);
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::set_centroids_ref(float* centroids_ref) {
  return d_centroids_ref->host2device(centroids_ref);
}

template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::init_TT_filter() {
  current_context->set_active_device(device, 1);
  int64_t d[2] = {1, 2};
  d_TT_slopes = new CarmaObj<float>(current_context, d);
  d[1] = nslopes;
  d_centro_filtered = new CarmaObj<float>(current_context, d);
  d_ref_Tip = new CarmaObj<float>(current_context, d);
  d_ref_Tilt = new CarmaObj<float>(current_context, d);
  return EXIT_SUCCESS;
}

// reference: libsutra/src.cpp/sutra_centroider.cpp:322-338: tip and tilt are
// both projected from the unfiltered slopes, then both removed; one device
// launch (no host round trip), d_centro_filtered keeps a copy as the
// reference's does.
template <class Tin, class Tout>
int32_t SutraCentroider<Tin, Tout>::apply_TT_filter(Tout* centroids) {
  static_assert(std::is_same<Tout, float>::value, "TT filter on float slopes");
  current_context->set_active_device(device, 1);
  if (d_ref_Tip == nullptr) init_TT_filter();
  hipStream_t st = wfs ? wfs->stream : nullptr;
  shwfs_throw_if(shwfs_tt_filter(centroids, d_ref_Tip->get_data(), d_ref_Tilt->get_data(),
                                 d_TT_slopes->get_data(), nslopes, st),
                 "apply_TT_filter");
  carma_safe_call(hipMemcpyAsync(d_centro_filtered->get_data(), centroids, sizeof(float) * nslopes,
                                 hipMemcpyDeviceToDevice, st));
  return EXIT_SUCCESS;
}

// ---- common get_cog plumbing ------------------------------------------------
template <class Tin, class Tout>
static int32_t run_centroids(SutraCentroider<Tin, Tout>* c, float* img, float* intensities,
                             Tout* centroids, int32_t nvalid, int32_t npix, int32_t ntot,
                             hipStream_t stream) {
  c->current_context->set_active_device(c->device, 1);
  if (c->d_validx == nullptr)
    throw std::runtime_error("get_cog: load_validpos has not been called");
  if (npix <= 0) throw std::runtime_error("get_cog: set_npix has not been called");
  if (c->d_validx->get_nb_elements() < nvalid)
    throw std::runtime_error("get_cog: fewer valid positions loaded than sub-apertures");
  if (c->get_type() == "wcog") {
    auto* w = dynamic_cast<SutraCentroiderWcog<Tin, Tout>*>(c);
    if (w && w->d_weights && w->d_weights->get_nb_elements() < (int64_t)npix * npix * nvalid)
      throw std::runtime_error("get_cog: weights smaller than npix^2 * nvalid");
  }
  shwfs_centroid_args a = c->centroid_args(intensities, centroids, nvalid, npix, ntot);
  float* out = (a.kind == SHWFS_CENTROID_TCOG) ? img : nullptr;
  shwfs_throw_if(shwfs_centroids(img, out, &a, stream), "get_cog");
  if (c->filter_TT) c->apply_TT_filter(centroids);
  return EXIT_SUCCESS;
}

template <class Tin, class Tout>
static shwfs_centroid_args base_args(SutraCentroider<Tin, Tout>* c, int32_t kind,
                                     float* intensities, Tout* slopes, int32_t nvalid,
                                     int32_t npix, int32_t size) {
  shwfs_centroid_args a{};
  a.kind = kind;
  a.nvalid = nvalid;
  a.npix = npix;
  a.size = size;
  a.d_validx = c->d_validx ? c->d_validx->get_data() : nullptr;
  a.d_validy = c->d_validy ? c->d_validy->get_data() : nullptr;
  a.offset = c->offset;
  a.scale = c->scale;
  a.threshold = 0.f;
  a.d_weights = nullptr;
  a.d_ref = c->d_centroids_ref->get_data();
  a.d_slopes = slopes;
  a.d_intensities = intensities;
  a.slope_order = (int32_t)c->slope_order;
  a.fp64_moments = c->fp64_moments ? 1 : 0;
  return a;
}

// ============================================================ Cog
// reference: libsutra/src.cpp/sutra_centroider_cog.cpp:26-98
template <class Tin, class T>
SutraCentroiderCog<Tin, T>::SutraCentroiderCog(CarmaContext* context, SutraWfs* wfs,
                                               int64_t nvalid, float offset, float scale,
                                               bool filter_TT, int32_t device)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device) {}
template <class Tin, class T>
SutraCentroiderCog<Tin, T>::~SutraCentroiderCog() {}

template <class Tin, class T>
shwfs_centroid_args SutraCentroiderCog<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                              int32_t nvalid, int32_t npix,
                                                              int32_t size) {
  return base_args(this, SHWFS_CENTROID_COG, intensities, slopes, nvalid, npix, size);
}

template <class Tin, class T>
int32_t SutraCentroiderCog<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                            int32_t nvalid, int32_t npix, int32_t ntot,
                                            hipStream_t stream) {
  return run_centroids(this, img, intensities, centroids, nvalid, npix, ntot, stream);
}

template <class Tin, class T>
int32_t SutraCentroiderCog<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs != nullptr) {
    CarmaObj<float>* img = (noise || !this->wfs->roket) ? this->wfs->d_binimg
                                                        : this->wfs->d_binimg_notnoisy;
    return get_cog(img->get_data(), intensities, slopes, this->wfs->nvalid_tot,
                   (int32_t)this->wfs->npix, (int32_t)this->wfs->d_binimg->get_dims(1),
                   this->wfs->stream);
  }
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_SUCCESS;  // as the reference (sutra_centroider_cog.cpp:86-87)
}

template <class Tin, class T>
int32_t SutraCentroiderCog<Tin, T>::get_cog() {
  if (this->wfs != nullptr)
    return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_SUCCESS;
}

// ============================================================ Bpcog
// reference: libsutra/src.cpp/sutra_centroider_bpcog.cpp:26-124
template <class Tin, class T>
SutraCentroiderBpcog<Tin, T>::SutraCentroiderBpcog(CarmaContext* context, SutraWfs* wfs,
                                                   int64_t nvalid, float offset, float scale,
                                                   bool filter_TT, int32_t device, int32_t nmax_)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device),
      nmax(0),
      d_bpix(nullptr),
      d_bpind(nullptr) {
  set_nmax(nmax_);
}
template <class Tin, class T>
SutraCentroiderBpcog<Tin, T>::~SutraCentroiderBpcog() {
  del(d_bpix);
  del(d_bpind);
}

template <class Tin, class T>
int32_t SutraCentroiderBpcog<Tin, T>::set_nmax(int32_t n) {
  if (n <= 0) throw std::runtime_error("set_nmax: nmax must be positive");
  this->current_context->set_active_device(this->device, 1);
  if (n != nmax || d_bpix == nullptr) {
    nmax = n;
    del(d_bpix);
    del(d_bpind);
    int64_t d[3] = {2, n, this->nvalid};
    d_bpix = new CarmaObj<T>(this->current_context, d);
    d_bpind = new CarmaObj<uint32_t>(this->current_context, d);
  }
  return EXIT_SUCCESS;
}

template <class Tin, class T>
shwfs_centroid_args SutraCentroiderBpcog<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                                int32_t nvalid, int32_t npix,
                                                                int32_t size) {
  // the kind is not used by shwfs_centroids_bpcog; the fused paths refuse bpcog
  return base_args(this, SHWFS_CENTROID_NONE, intensities, slopes, nvalid, npix, size);
}

template <class Tin, class T>
int32_t SutraCentroiderBpcog<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                              int32_t nvalid, int32_t npix, int32_t ntot,
                                              hipStream_t stream) {
  this->current_context->set_active_device(this->device, 1);
  if (this->d_validx == nullptr)
    throw std::runtime_error("get_cog: load_validpos has not been called");
  if (npix <= 0) throw std::runtime_error("get_cog: set_npix has not been called");
  if (this->d_validx->get_nb_elements() < nvalid)
    throw std::runtime_error("get_cog: fewer valid positions loaded than sub-apertures");
  shwfs_centroid_args a = centroid_args(intensities, centroids, nvalid, npix, ntot);
  shwfs_throw_if(shwfs_centroids_bpcog(img, nmax, &a, stream), "get_cog (bpcog)");
  if (this->filter_TT) this->apply_TT_filter(centroids);
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderBpcog<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs != nullptr) {
    CarmaObj<float>* img = (noise || !this->wfs->roket) ? this->wfs->d_binimg
                                                        : this->wfs->d_binimg_notnoisy;
    return get_cog(img->get_data(), intensities, slopes, (int32_t)this->wfs->nvalid,
                   (int32_t)this->wfs->npix, (int32_t)this->wfs->d_binimg->get_dims(1),
                   this->wfs->stream);
  }
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

template <class Tin, class T>
int32_t SutraCentroiderBpcog<Tin, T>::get_cog() {
  if (this->wfs != nullptr)
    return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

// ============================================================ Corr
// reference: libsutra/src.cpp/sutra_centroider_corr.cpp:26-285
template <class Tin, class T>
SutraCentroiderCorr<Tin, T>::SutraCentroiderCorr(CarmaContext* context, SutraWfs* wfs,
                                                 int64_t nvalid, float offset, float scale,
                                                 bool filter_TT, int32_t device)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device),
      interp_sizex(0),
      interp_sizey(0),
      d_corrfnct(nullptr),
      d_corrspot(nullptr),
      d_corrnorm(nullptr),
      d_corrmax(nullptr),
      d_corr(nullptr),
      d_interpmat(nullptr) {
  if (wfs != nullptr) {
    this->npix = (int32_t)wfs->npix;
    this->nxsub = (int32_t)wfs->nxsub;
  }
}

template <class Tin, class T>
SutraCentroiderCorr<Tin, T>::~SutraCentroiderCorr() {
  del(d_corrfnct);
  del(d_corrnorm);
  del(d_corrmax);
  del(d_corr);
  del(d_interpmat);
}

template <class Tin, class T>
int32_t SutraCentroiderCorr<Tin, T>::init_corr(int32_t isizex, int32_t isizey, T* interpmat) {
  const int32_t n = this->npix;
  if (n <= 0) throw std::runtime_error("init_corr: set_npix has not been called");
  if (isizex <= 0 || isizey <= 0 || isizex > 2 * n - 1 || isizey > 2 * n - 1)
    throw std::runtime_error("init_corr: interpolation patch larger than the correlation map");
  this->current_context->set_active_device(this->device, 1);
  del(d_corrfnct);
  del(d_corrnorm);
  del(d_corrmax);
  del(d_corr);
  del(d_interpmat);
  const int64_t N = 2 * n - 1;
  int64_t d3[4] = {3, n, n, this->nvalid};
  d_corrfnct = new CarmaObj<T>(this->current_context, d3);
  int64_t dc[4] = {3, N, N, this->nvalid};
  d_corr = new CarmaObj<T>(this->current_context, dc);
  int64_t dn[3] = {2, N, N};
  d_corrnorm = new CarmaObj<T>(this->current_context, dn);
  std::vector<T> ones((size_t)(N * N), T(1));
  d_corrnorm->host2device(ones.data());
  int64_t dm[2] = {1, this->nvalid};
  d_corrmax = new CarmaObj<int32_t>(this->current_context, dm);
  interp_sizex = isizex;
  interp_sizey = isizey;
  int64_t di[3] = {2, (int64_t)isizex * isizey, 6};
  d_interpmat = new CarmaObj<T>(this->current_context, di);
  d_interpmat->host2device(interpmat);
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderCorr<Tin, T>::load_corr(T* corr, T* corr_norm, int32_t ndim) {
  if (d_corrfnct == nullptr) throw std::runtime_error("load_corr: init_corr has not been called");
  if (ndim != 2 && ndim != 3) throw std::runtime_error("load_corr: ndim must be 2 or 3");
  this->current_context->set_active_device(this->device, 1);
  d_corrnorm->host2device(corr_norm);
  const size_t n2 = (size_t)this->npix * this->npix;
  ref_per_subap = ndim == 3;
  carma_safe_call(hipMemcpy(d_corrfnct->get_data(), corr,
                            sizeof(T) * n2 * (ref_per_subap ? (size_t)this->nvalid : 1),
                            hipMemcpyHostToDevice));
  return EXIT_SUCCESS;
}

template <class Tin, class T>
shwfs_centroid_args SutraCentroiderCorr<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                               int32_t nvalid, int32_t npix,
                                                               int32_t size) {
  // the CoG kinds do not apply; the fused paths run this centroider after the frame
  return base_args(this, SHWFS_CENTROID_NONE, intensities, slopes, nvalid, npix, size);
}

// The spots are read with the row stride npix * nxsub (the detector image
// side), as fill_bincube does in the reference; ntot is not used.
template <class Tin, class T>
int32_t SutraCentroiderCorr<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                             int32_t nvalid, int32_t npix, int32_t ntot,
                                             hipStream_t stream) {
  (void)intensities;
  (void)ntot;
  this->current_context->set_active_device(this->device, 1);
  if (d_corrfnct == nullptr) throw std::runtime_error("get_cog (corr): init_corr has not been called");
  if (this->d_validx == nullptr)
    throw std::runtime_error("get_cog: load_validpos has not been called");
  if (this->nxsub <= 0) throw std::runtime_error("get_cog (corr): nxsub not set for this centroider");
  if (npix != this->npix || nvalid != this->nvalid)
    throw std::runtime_error("get_cog (corr): npix / nvalid differ from init_corr's");
  if (this->d_validx->get_nb_elements() < nvalid)
    throw std::runtime_error("get_cog: fewer valid positions loaded than sub-apertures");
  shwfs_throw_if(
      shwfs_centroids_corr(img, npix * this->nxsub, this->d_validx->get_data(),
                           this->d_validy->get_data(), nvalid, npix, d_corrfnct->get_data(),
                           ref_per_subap ? 1 : 0, d_corrnorm->get_data(),
                           d_interpmat->get_data(), interp_sizex, interp_sizey, this->offset,
                           this->scale, this->d_centroids_ref->get_data(),
                           (int32_t)this->slope_order, centroids, d_corr->get_data(),
                           d_corrmax->get_data(), stream),
      "get_cog (corr)");
  if (this->filter_TT) this->apply_TT_filter(centroids);
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderCorr<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs != nullptr) {
    CarmaObj<float>* img = (noise || !this->wfs->roket) ? this->wfs->d_binimg
                                                        : this->wfs->d_binimg_notnoisy;
    if (img->get_dims(1) != (int64_t)this->npix * this->nxsub)
      throw std::runtime_error("get_cog (corr): detector image side is not npix * nxsub");
    return get_cog(img->get_data(), intensities, slopes, (int32_t)this->wfs->nvalid,
                   (int32_t)this->wfs->npix, (int32_t)img->get_dims(1), this->wfs->stream);
  }
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

template <class Tin, class T>
int32_t SutraCentroiderCorr<Tin, T>::get_cog() {
  if (this->wfs != nullptr)
    return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

// ============================================================ Tcog
// reference: libsutra/src.cpp/sutra_centroider_tcog.cpp:26-108
template <class Tin, class T>
SutraCentroiderTcog<Tin, T>::SutraCentroiderTcog(CarmaContext* context, SutraWfs* wfs,
                                                 int64_t nvalid, float offset, float scale,
                                                 bool filter_TT, int32_t device)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device),
      threshold(0.f) {}
template <class Tin, class T>
SutraCentroiderTcog<Tin, T>::~SutraCentroiderTcog() {}

template <class Tin, class T>
int32_t SutraCentroiderTcog<Tin, T>::set_threshold(float t) {
  threshold = t;
  return EXIT_SUCCESS;
}

template <class Tin, class T>
shwfs_centroid_args SutraCentroiderTcog<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                               int32_t nvalid, int32_t npix,
                                                               int32_t size) {
  shwfs_centroid_args a = base_args(this, SHWFS_CENTROID_TCOG, intensities, slopes, nvalid, npix, size);
  a.threshold = threshold;
  return a;
}

template <class Tin, class T>
int32_t SutraCentroiderTcog<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                             int32_t nvalid, int32_t npix, int32_t ntot,
                                             hipStream_t stream) {
  return run_centroids(this, img, intensities, centroids, nvalid, npix, ntot, stream);
}

template <class Tin, class T>
int32_t SutraCentroiderTcog<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs != nullptr) {
    CarmaObj<float>* img = (noise || !this->wfs->roket) ? this->wfs->d_binimg
                                                        : this->wfs->d_binimg_notnoisy;
    return get_cog(img->get_data(), intensities, slopes, (int32_t)this->wfs->nvalid,
                   (int32_t)this->wfs->npix, (int32_t)this->wfs->d_binimg->get_dims(1),
                   this->wfs->stream);
  }
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

template <class Tin, class T>
int32_t SutraCentroiderTcog<Tin, T>::get_cog() {
  if (this->wfs != nullptr)
    return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

// ============================================================ Wcog
// reference: libsutra/src.cpp/sutra_centroider_wcog.cpp:26-159
template <class Tin, class T>
SutraCentroiderWcog<Tin, T>::SutraCentroiderWcog(CarmaContext* context, SutraWfs* wfs,
                                                 int64_t nvalid, float offset, float scale,
                                                 bool filter_TT, int32_t device)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device),
      threshold(0.f),
      d_weights(nullptr) {}
template <class Tin, class T>
SutraCentroiderWcog<Tin, T>::~SutraCentroiderWcog() {
  delete d_weights;
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::init_weights() {
  this->current_context->set_active_device(this->device, 1);
  delete d_weights;
  int64_t d[4] = {3, this->npix, this->npix, this->nvalid};
  d_weights = new CarmaObj<float>(this->current_context, d);
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::load_weights(float* weights, int32_t ndim) {
  if (d_weights == nullptr) throw std::runtime_error("load_weights: init_weights has not been called");
  if (ndim == 3) {
    d_weights->host2device(weights);
  } else {
    int64_t d[3] = {2, this->npix, this->npix};
    CarmaObj<float> tmp(this->current_context, d, weights);
    shwfs_throw_if(shwfs_fill_weights(d_weights->get_data(), tmp.get_data(), this->npix,
                                      (int32_t)d_weights->get_nb_elements(), nullptr),
                   "load_weights");
    carma_safe_call(hipDeviceSynchronize());
  }
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::set_threshold(float t) {
  threshold = t;
  return EXIT_SUCCESS;
}

template <class Tin, class T>
shwfs_centroid_args SutraCentroiderWcog<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                               int32_t nvalid, int32_t npix,
                                                               int32_t size) {
  shwfs_centroid_args a = base_args(this, SHWFS_CENTROID_WCOG, intensities, slopes, nvalid, npix, size);
  a.threshold = threshold;
  a.d_weights = d_weights ? d_weights->get_data() : nullptr;
  return a;
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                             int32_t nvalid, int32_t npix, int32_t ntot,
                                             hipStream_t stream) {
  return run_centroids(this, img, intensities, centroids, nvalid, npix, ntot, stream);
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs != nullptr) {
    CarmaObj<float>* img = (noise || !this->wfs->roket) ? this->wfs->d_binimg
                                                        : this->wfs->d_binimg_notnoisy;
    return get_cog(img->get_data(), intensities, slopes, (int32_t)this->wfs->nvalid,
                   (int32_t)this->wfs->npix, (int32_t)this->wfs->d_binimg->get_dims(1),
                   this->wfs->stream);
  }
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

template <class Tin, class T>
int32_t SutraCentroiderWcog<Tin, T>::get_cog() {
  if (this->wfs != nullptr)
    return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
  std::cerr << "this->wfs was not initialized" << std::endl;
  return EXIT_FAILURE;
}

// ============================================================ fused frame
int32_t comp_frame(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* intensities,
                   float* slopes, bool noise) {
  if (c == nullptr || c->wfs != wfs)
    throw std::runtime_error("comp_frame: the centroider must be attached to this WFS");
  if (!wfs->arrays_loaded || wfs->d_gs == nullptr)
    throw std::runtime_error("comp_frame: load_arrays / initgs have not been called");
  if (c->npix != wfs->npix || c->nvalid != wfs->nvalid)
    throw std::runtime_error("comp_frame: centroider and WFS sizes differ");
  if (wfs->d_gs->d_phase->d_screen->get_dims(1) != wfs->d_pupil->get_dims(1))
    throw std::runtime_error("comp_frame: phase screen and pupil sizes differ");
  if (c->get_type() == "wcog") {
    auto* wc = dynamic_cast<SutraCentroiderWcog<float, float>*>(c);
    if (!wc || !wc->d_weights || wc->d_weights->get_nb_elements() < wfs->npix * wfs->npix * wfs->nvalid)
      throw std::runtime_error("comp_frame: WCoG weights not initialised for this WFS");
  }
  wfs->current_context->set_active_device(wfs->device, 1);
  shwfs_centroid_args a =
      c->centroid_args(intensities, slopes, (int32_t)wfs->nvalid, (int32_t)wfs->npix,
                       (int32_t)wfs->d_binimg->get_dims(1));
  // not fused: spots, then centroids (roket keeps the image before the noise,
  // which the fused launch does not write; the learned centroider is fused
  // only where centroid_args says so)
  if (a.kind == SHWFS_CENTROID_NONE || c->get_type() == "bpcog" || wfs->roket) {
    wfs->comp_image(noise);
    return c->get_cog(intensities, slopes, noise);
  }
  // the fused learned centroider and the traced field stage are separate kernels
  wfs->check_lgs_ready();
  if (a.kind == SHWFS_CENTROID_MLP || !wfs->layers_in_field_stage())
    wfs->d_gs->render_layers(wfs->stream);
  wfs->field_stop_stage(wfs->stream);
  shwfs_sh_geom g = wfs->geom();
  shwfs_frame_args f = wfs->frame_args(noise);
  f.centroid = &a;
  shwfs_throw_if(shwfs_sh_comp_image(&g, &f, wfs->stream), "comp_frame");
  wfs->noise_call++;
  if (c->filter_TT) c->apply_TT_filter(slopes);
  return EXIT_SUCCESS;
}

double run_frames(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* slopes,
                  int32_t nframes, bool noise, bool fused) {
  hipEvent_t e0, e1;
  carma_safe_call(hipEventCreate(&e0));
  carma_safe_call(hipEventCreate(&e1));
  carma_safe_call(hipEventRecord(e0, wfs->stream));
  for (int32_t k = 0; k < nframes; ++k) {
    if (fused) {
      comp_frame(wfs, c, c->d_intensities->get_data(), slopes, noise);
    } else {
      wfs->comp_image(noise);
      c->get_cog(c->d_intensities->get_data(), slopes, noise);
    }
  }
  carma_safe_call(hipEventRecord(e1, wfs->stream));
  carma_safe_call(hipEventSynchronize(e1));
  float ms = 0.f;
  carma_safe_call(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

double run_frames_graph(SutraWfsSH* wfs, SutraCentroider<float, float>* c, float* slopes,
                        int32_t nframes, int32_t reps, bool noise) {
  if (nframes <= 0 || reps <= 0) throw std::runtime_error("run_frames_graph: counts must be positive");
  wfs->current_context->set_active_device(wfs->device, 1);
  // one frame outside the capture: first-use work (device tables) happens here
  comp_frame(wfs, c, c->d_intensities->get_data(), slopes, noise);
  carma_safe_call(hipDeviceSynchronize());
  hipStream_t st = nullptr;
  carma_safe_call(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipStream_t keep = wfs->stream;
  wfs->stream = st;
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  double ms_total = -1.0;
  try {
    carma_safe_call(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    // with noise, each captured frame keeps the Philox call counter it was
    // captured with: replays repeat the same draws (the work is the same)
    for (int32_t k = 0; k < nframes; ++k) comp_frame(wfs, c, c->d_intensities->get_data(), slopes, noise);
    carma_safe_call(hipStreamEndCapture(st, &g));
    carma_safe_call(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    carma_safe_call(hipGraphLaunch(ex, st));
    hipEvent_t e0, e1;
    carma_safe_call(hipEventCreate(&e0));
    carma_safe_call(hipEventCreate(&e1));
    carma_safe_call(hipEventRecord(e0, st));
    for (int32_t r = 0; r < reps; ++r) carma_safe_call(hipGraphLaunch(ex, st));
    carma_safe_call(hipEventRecord(e1, st));
    carma_safe_call(hipEventSynchronize(e1));
    float ms = 0.f;
    carma_safe_call(hipEventElapsedTime(&ms, e0, e1));
    ms_total = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  } catch (...) {
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(st, &dead);  // no-op unless the capture is still open
    if (dead) (void)hipGraphDestroy(dead);
    if (ex) (void)hipGraphExecDestroy(ex);
    if (g) (void)hipGraphDestroy(g);
    wfs->stream = keep;
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    throw;
  }
  (void)hipGraphExecDestroy(ex);
  (void)hipGraphDestroy(g);
  wfs->stream = keep;
  carma_safe_call(hipStreamSynchronize(st));
  carma_safe_call(hipStreamDestroy(st));
  return ms_total;
}

template class SutraCentroider<float, float>;
template class SutraCentroiderCog<float, float>;
template class SutraCentroiderTcog<float, float>;
template class SutraCentroiderWcog<float, float>;
// uint16 camera frames (Centroider_UF; calibrate_img turns them into the
// float image the centroids are taken from)
template class SutraCentroider<uint16_t, float>;
template class SutraCentroiderCog<uint16_t, float>;
template class SutraCentroiderTcog<uint16_t, float>;
template class SutraCentroiderWcog<uint16_t, float>;
template class SutraCentroiderBpcog<float, float>;
template class SutraCentroiderBpcog<uint16_t, float>;
template class SutraCentroiderCorr<float, float>;
template class SutraCentroiderCorr<uint16_t, float>;

// ============================================================ controller / rtc
template <class T, class Tout>
SutraController<T, Tout>::SutraController(CarmaContext* context, string type_, int32_t nslope_,
                                          int32_t nactu_, float delay_, int32_t device_,
                                          const int32_t* idx_centro, int32_t ncentro)
    : type(type_),
      device(device_),
      nslope(nslope_),
      nactu(nactu_),
      delay(delay_),
      gain(0.f),
      current_context(context) {
  int64_t d[2] = {1, nslope};
  d_centroids = new CarmaObj<T>(context, d);
  d[1] = nactu > 0 ? nactu : 1;
  d_com = new CarmaObj<T>(context, d);
  for (int32_t i = 0; i < ncentro; ++i) centro_idx.push_back(idx_centro ? idx_centro[i] : i);
}

template <class T, class Tout>
SutraController<T, Tout>::~SutraController() {
  delete d_centroids;
  delete d_com;
}

template class SutraController<float, float>;

template <typename Tin, typename T, typename Tout>
SutraRtc<Tin, T, Tout>::SutraRtc() {}

template <typename Tin, typename T, typename Tout>
SutraRtc<Tin, T, Tout>::~SutraRtc() {
  for (auto* c : d_centro) delete c;
  for (auto* c : d_control) delete c;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::add_centroider(CarmaContext* context, int64_t nvalid, float offset,
                                               float scale, bool filter_TT, int64_t device,
                                               string typec) {
  return add_centroider(context, nvalid, offset, scale, filter_TT, device, typec, nullptr);
}

// reference: libsutra/src.cpp/sutra_rtc.cpp:73-115 (factory by type string)
template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::add_centroider(CarmaContext* context, int64_t nvalid, float offset,
                                               float scale, bool filter_TT, int64_t device,
                                               string typec, SutraWfs* wfs) {
  if (typec == "bpcog")  // ten brightest pixels, as the reference's factory
    d_centro.push_back(new SutraCentroiderBpcog<Tin, T>(context, wfs, nvalid, offset, scale,
                                                        filter_TT, (int32_t)device, 10));
  else if (typec == "corr")
    d_centro.push_back(new SutraCentroiderCorr<Tin, T>(context, wfs, nvalid, offset, scale,
                                                       filter_TT, (int32_t)device));
  else if (typec == "cog")
    d_centro.push_back(new SutraCentroiderCog<Tin, T>(context, wfs, nvalid, offset, scale,
                                                      filter_TT, (int32_t)device));
  else if (typec == "tcog")
    d_centro.push_back(new SutraCentroiderTcog<Tin, T>(context, wfs, nvalid, offset, scale,
                                                       filter_TT, (int32_t)device));
  else if (typec == "wcog")
    d_centro.push_back(new SutraCentroiderWcog<Tin, T>(context, wfs, nvalid, offset, scale,
                                                       filter_TT, (int32_t)device));
  else if (typec == "mlp") {  // learned centroider (this build's addition), float images
    if constexpr (std::is_same<Tin, float>::value)
      d_centro.push_back(new SutraCentroiderMlp<Tin, T>(context, wfs, nvalid, offset, scale,
                                                        filter_TT, (int32_t)device));
    else
      throw std::runtime_error("add_centroider: the mlp centroider takes float images (Rtc_FFF)");
  } else {
    std::cerr << "centroider unknown" << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::add_controller(CarmaContext* context, string typec, int64_t device,
                                               float delay, int32_t nslope, int32_t nactu,
                                               const int32_t* idx_centro, int32_t ncentro) {
  if (typec != "generic")
    throw std::runtime_error("add_controller: only the slope-holding \"generic\" controller is part of this build");
  d_control.push_back(new SutraController<T, Tout>(context, typec, nslope, nactu, delay,
                                                   (int32_t)device, idx_centro, ncentro));
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::remove_centroider(int32_t ncentro) {
  delete d_centro.at(ncentro);
  d_centro.erase(d_centro.begin() + ncentro);
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::do_calibrate_img(int32_t ncntrl) {
  for (int32_t idx : d_control.at(ncntrl)->centro_idx) d_centro.at(idx)->calibrate_img();
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::do_centroids() {
  for (auto* c : d_centro) c->get_cog();
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::do_centroids(int32_t ncntrl) {
  return do_centroids(ncntrl, true);
}

// reference: libsutra/src.cpp/sutra_rtc.cpp:485-511
template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::do_centroids(int32_t ncntrl, bool noise) {
  SutraController<T, Tout>* ctl = d_control.at(ncntrl);
  int32_t indslope = 0;
  for (int32_t idx : ctl->centro_idx) {
    SutraCentroider<Tin, T>* c = d_centro.at(idx);
    if (indslope + c->nslopes > ctl->nslope)
      throw std::runtime_error("do_centroids: the controller slope vector is too short");
    if (c->wfs != nullptr) {
      c->get_cog(c->d_intensities->get_data(), ctl->d_centroids->get_data_at(indslope), noise);
    } else {
      if (c->d_img == nullptr) throw std::runtime_error("do_centroids: no image loaded");
      c->get_cog(c->d_img->get_data(), c->d_intensities->get_data(),
                 ctl->d_centroids->get_data_at(indslope), c->nvalid, c->npix,
                 (int32_t)c->d_img->get_dims(1));
    }
    indslope += c->nslopes;
  }
  return EXIT_SUCCESS;
}

template <typename Tin, typename T, typename Tout>
int32_t SutraRtc<Tin, T, Tout>::set_centroids_ref(float* centroids_ref) {
  int32_t off = 0;
  for (auto* c : d_centro) {
    c->set_centroids_ref(centroids_ref + off);
    off += c->nslopes;
  }
  return EXIT_SUCCESS;
}

template class SutraRtc<float, float, float>;
template class SutraRtc<uint16_t, float, float>;
