// SutraCentroiderMlp: learned per-sub-aperture centroider over the C ABI
// (shwfs_mlp_*), and the generator of its training pairs.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>

#include "sutra.hpp"

template <class Tin, class T>
SutraCentroiderMlp<Tin, T>::SutraCentroiderMlp(CarmaContext* context, SutraWfs* wfs,
                                               int64_t nvalid, float offset, float scale,
                                               bool filter_TT, int32_t device)
    : SutraCentroider<Tin, T>(context, wfs, nvalid, offset, scale, filter_TT, device),
      x_scale(1.0f),
      flux_lo(-INFINITY),
      flux_hi(INFINITY),
      step(0) {
  int64_t d[2] = {1, SHWFS_MLP_NPARAM};
  d_params = new CarmaObj<float>(context, d);
  d_grad = new CarmaObj<float>(context, d);
  d_m = new CarmaObj<float>(context, d);
  d_v = new CarmaObj<float>(context, d);
  d_params->reset();
  d_grad->reset();
  d_m->reset();
  d_v->reset();
  int64_t dw[2] = {1, (int64_t)((shwfs_mlp_work_bytes() + 3) / 4)};
  d_work = new CarmaObj<float>(context, dw);
  int64_t dl[2] = {1, 1};
  d_loss = new CarmaObj<float>(context, dl);
  int64_t df[2] = {1, nvalid > 0 ? nvalid : 1};
  d_fallback = new CarmaObj<int32_t>(context, df);
  d_fallback->reset();
}

template <class Tin, class T>
SutraCentroiderMlp<Tin, T>::~SutraCentroiderMlp() {
  delete d_params;
  delete d_grad;
  delete d_m;
  delete d_v;
  delete d_work;
  delete d_loss;
  delete d_fallback;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::init_params(uint64_t seed) {
  this->current_context->set_active_device(this->device, 1);
  shwfs_throw_if(shwfs_mlp_init(d_params->get_data(), seed, nullptr), "mlp init_params");
  d_m->reset();
  d_v->reset();
  d_grad->reset();
  step = 0;
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::set_input_scale(float s) {
  x_scale = s;
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::set_flux_range(float lo, float hi) {
  if (!(lo <= hi)) throw std::invalid_argument("set_flux_range: lo > hi");
  flux_lo = lo;
  flux_hi = hi;
  return EXIT_SUCCESS;
}

static void check_xt(CarmaObj<float>* x, CarmaObj<float>* t, int64_t n) {
  if (!x || n <= 0) throw std::invalid_argument("mlp: empty input");
  if (x->get_nb_elements() < n * 64) throw std::invalid_argument("mlp: x smaller than n x 64");
  if (t && t->get_nb_elements() < n * 2) throw std::invalid_argument("mlp: t smaller than n x 2");
}

template <class Tin, class T>
float SutraCentroiderMlp<Tin, T>::loss_grad(CarmaObj<float>* x, CarmaObj<float>* t, int64_t n,
                                            float t_scale) {
  check_xt(x, t, n);
  this->current_context->set_active_device(this->device, 1);
  shwfs_throw_if(shwfs_mlp_loss_grad(d_params->get_data(), d_work->get_data(), x->get_data(),
                                     t->get_data(), nullptr, n, x_scale, t_scale,
                                     d_grad->get_data(), d_loss->get_data(), nullptr),
                 "mlp loss_grad");
  float l = 0.f;
  d_loss->device2host(&l);
  return l;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::forward(CarmaObj<float>* x, int64_t n, CarmaObj<float>* y) {
  check_xt(x, y, n);
  this->current_context->set_active_device(this->device, 1);
  shwfs_throw_if(shwfs_mlp_forward(d_params->get_data(), d_work->get_data(), x->get_data(), n,
                                   x_scale, y->get_data(), nullptr),
                 "mlp forward");
  return EXIT_SUCCESS;
}

template <class Tin, class T>
float SutraCentroiderMlp<Tin, T>::train(CarmaObj<float>* x, CarmaObj<float>* t, int64_t n,
                                        float t_scale, int32_t batch, int32_t epochs, float lr,
                                        uint64_t seed) {
  check_xt(x, t, n);
  if (batch <= 0 || epochs <= 0) throw std::invalid_argument("mlp train: batch and epochs > 0");
  if (n > 0x7fffffff) throw std::invalid_argument("mlp train: at most 2^31 - 1 pairs");
  this->current_context->set_active_device(this->device, 1);
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed);
  int64_t di[2] = {1, n};
  CarmaObj<int32_t> d_idx(this->current_context, di);
  const int64_t nb = n / batch;  // full batches only
  if (nb == 0) throw std::invalid_argument("mlp train: fewer pairs than one batch");
  int64_t dl[2] = {1, nb};
  CarmaObj<float> d_losses(this->current_context, dl);
  float last = 0.f;
  for (int32_t e = 0; e < epochs; ++e) {
    std::shuffle(perm.begin(), perm.end(), rng);
    d_idx.host2device(perm.data());
    for (int64_t b = 0; b < nb; ++b) {
      shwfs_throw_if(shwfs_mlp_loss_grad(d_params->get_data(), d_work->get_data(), x->get_data(),
                                         t->get_data(), d_idx.get_data() + b * batch, batch,
                                         x_scale, t_scale, d_grad->get_data(),
                                         d_losses.get_data() + b, nullptr),
                     "mlp train");
      ++step;
      shwfs_throw_if(shwfs_mlp_adam(d_params->get_data(), d_grad->get_data(), d_m->get_data(),
                                    d_v->get_data(), step, lr, 0.9f, 0.999f, 1e-8f, nullptr),
                     "mlp train");
    }
    std::vector<float> h(nb);
    d_losses.device2host(h.data());
    last = (float)(std::accumulate(h.begin(), h.end(), 0.0) / (double)nb);
  }
  return last;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::get_cog(float* img, float* intensities, T* centroids,
                                            int32_t nvalid, int32_t npix, int32_t ntot,
                                            hipStream_t stream) {
  this->current_context->set_active_device(this->device, 1);
  if (this->d_validx == nullptr) throw std::runtime_error("get_cog: load_validpos has not been called");
  if (this->d_validx->get_nb_elements() < nvalid)
    throw std::runtime_error("get_cog: fewer valid positions loaded than sub-apertures");
  if (d_fallback->get_nb_elements() < nvalid)
    throw std::runtime_error("get_cog: more sub-apertures than the centroider was built for");
  shwfs_throw_if(shwfs_mlp_centroids(d_params->get_data(), d_work->get_data(), img, ntot,
                                     this->d_validx->get_data(), this->d_validy->get_data(),
                                     nvalid, npix, x_scale, flux_lo, flux_hi, this->offset,
                                     this->scale, this->scale, this->d_centroids_ref->get_data(),
                                     (int32_t)this->slope_order, centroids, intensities,
                                     d_fallback->get_data(), stream),
                 "mlp get_cog");
  if (this->filter_TT) this->apply_TT_filter(centroids);
  return EXIT_SUCCESS;
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::get_cog(float* intensities, T* slopes, bool noise) {
  if (this->wfs == nullptr) throw std::runtime_error("get_cog: no WFS attached");
  (void)noise;
  return get_cog(this->wfs->d_binimg->get_data(), intensities, slopes,
                 (int32_t)this->wfs->nvalid, (int32_t)this->wfs->npix,
                 (int32_t)this->wfs->d_binimg->get_dims(1), this->wfs->stream);
}

template <class Tin, class T>
int32_t SutraCentroiderMlp<Tin, T>::get_cog() {
  if (this->wfs == nullptr) throw std::runtime_error("get_cog: no WFS attached");
  return get_cog(this->wfs->d_intensities->get_data(), this->wfs->d_slopes->get_data(), true);
}

// Fused after binning (kind SHWFS_CENTROID_MLP) when the attached WFS has
// 8x8-pixel spots and either computes them with the FFT method at nfft 16 or
// 32, or with an MFMA method at nfft 32 from 16-point windows binned 4x4 in
// the half-grid form (block binmap, origin 16); otherwise NONE (comp_frame
// runs the spots, then get_cog).
template <class Tin, class T>
shwfs_centroid_args SutraCentroiderMlp<Tin, T>::centroid_args(float* intensities, T* slopes,
                                                              int32_t nvalid, int32_t npix,
                                                              int32_t size) {
  shwfs_centroid_args a{};
  a.kind = SHWFS_CENTROID_NONE;
  auto* sh = dynamic_cast<SutraWfsSH*>(this->wfs);
  if (!sh || npix != 8 || d_fallback->get_nb_elements() < nvalid) return a;
  const bool fft = sh->get_method() == SHWFS_METHOD_FFT_FP32 && (sh->nfft == 16 || sh->nfft == 32);
  const bool mfma = sh->get_method() != SHWFS_METHOD_FFT_FP32 && sh->nfft == 32 &&
                    sh->nphase == 16 && sh->nrebin == 4 && sh->bin_block && sh->bin_origin == 16;
  if (!fft && !mfma) return a;
  a.kind = SHWFS_CENTROID_MLP;
  a.nvalid = nvalid;
  a.npix = npix;
  a.size = size;
  a.d_validx = this->d_validx ? this->d_validx->get_data() : nullptr;
  a.d_validy = this->d_validy ? this->d_validy->get_data() : nullptr;
  a.offset = this->offset;
  a.scale = this->scale;
  a.d_ref = this->d_centroids_ref->get_data();
  a.d_slopes = slopes;
  a.d_intensities = intensities;
  a.slope_order = (int32_t)this->slope_order;
  a.d_mlp_params = d_params->get_data();
  a.mlp_x_scale = x_scale;
  a.mlp_y_scale = this->scale;
  a.mlp_flux_lo = flux_lo;
  a.mlp_flux_hi = flux_hi;
  a.d_mlp_fallback = d_fallback->get_data();
  return a;
}

int64_t make_mlp_training_set(SutraWfsSH* wfs, CarmaObj<float>* base, float r0_base, float r0_min,
                              float r0_max, int32_t nframes, uint64_t seed, int32_t geom_type,
                              CarmaObj<float>* X, CarmaObj<float>* T) {
  if (!wfs || !base || !X || !T) throw std::invalid_argument("make_mlp_training_set: None argument");
  if (!wfs->arrays_loaded || !wfs->d_gs) throw std::runtime_error("make_mlp_training_set: WFS not ready");
  if (wfs->npix != 8) throw std::invalid_argument("make_mlp_training_set: needs 8x8-pixel spots");
  if (!(r0_min > 0.f && r0_min <= r0_max && r0_base > 0.f))
    throw std::invalid_argument("make_mlp_training_set: bad r0 range");
  const int64_t nv = wfs->nvalid;
  const int64_t n = nv * (int64_t)nframes;
  if (X->get_nb_elements() < n * 64 || T->get_nb_elements() < n * 2)
    throw std::invalid_argument("make_mlp_training_set: X / T too small");
  CarmaObj<float>* scr = wfs->d_gs->d_phase->d_screen;
  wfs->d_gs->pending.clear();  // the screen is written here, frame by frame
  const int32_t npup = (int32_t)scr->get_dims(1);
  const int32_t side = (int32_t)base->get_dims(1);
  if (base->get_dims(2) != side || side < npup)
    throw std::invalid_argument("make_mlp_training_set: base screen must be square and >= npup");
  wfs->current_context->set_active_device(wfs->device, 1);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> ur0(r0_min, r0_max);
  std::uniform_int_distribution<int32_t> ush(0, side - npup);
  for (int32_t f = 0; f < nframes; ++f) {
    const double r0 = ur0(rng);
    const float gain = (float)std::pow(r0 / r0_base, -5.0 / 6.0);
    const int32_t ox = ush(rng), oy = ush(rng);
    shwfs_throw_if(shwfs_phase_window(base->get_data(), side, ox, oy, gain, scr->get_data(), npup,
                                      wfs->stream),
                   "make_mlp_training_set");
    wfs->comp_image(true);
    wfs->slopes_geom(geom_type);
    float* xo = X->get_data() + (int64_t)f * nv * 64;
    float* to = T->get_data() + (int64_t)f * nv * 2;
    carma_safe_call(hipMemcpyAsync(xo, wfs->d_bincube->get_data(), sizeof(float) * nv * 64,
                                   hipMemcpyDeviceToDevice, wfs->stream));
    // slopes_geom: x[0..nv), y[nv..2nv) -> interleaved (x, y) rows
    carma_safe_call(hipMemcpy2DAsync(to, 2 * sizeof(float), wfs->d_slopes->get_data(),
                                     sizeof(float), sizeof(float), nv, hipMemcpyDeviceToDevice,
                                     wfs->stream));
    carma_safe_call(hipMemcpy2DAsync(to + 1, 2 * sizeof(float), wfs->d_slopes->get_data() + nv,
                                     sizeof(float), sizeof(float), nv, hipMemcpyDeviceToDevice,
                                     wfs->stream));
  }
  carma_safe_call(hipStreamSynchronize(wfs->stream));
  return n;
}

template class SutraCentroiderMlp<float, float>;
