// SutraWfsBatch: several SH WFS per launch through the batch C ABI.
#include <functional>
#include <stdexcept>

#include "sutra.hpp"

SutraWfsBatch::SutraWfsBatch(CarmaContext* context, vector<SutraWfsSH*> wfs_,
                             vector<SutraCentroider<float, float>*> centro_, bool noise_)
    : wfs(wfs_), centro(centro_), noise(noise_), device(0), current_context(context),
      handle(nullptr) {
  if (wfs.empty()) throw std::invalid_argument("WfsBatch: no WFS given");
  if (!centro.empty() && centro.size() != wfs.size())
    throw std::invalid_argument("WfsBatch: one centroider (or None) per WFS");
  centro.resize(wfs.size(), nullptr);
  device = wfs[0]->device;
  vector<shwfs_sh_geom> geoms;
  vector<shwfs_frame_args> frames;
  vector<shwfs_centroid_args> cargs(wfs.size());
  for (size_t i = 0; i < wfs.size(); ++i) {
    SutraWfsSH* w = wfs[i];
    if (!w) throw std::invalid_argument("WfsBatch: None WFS");
    if (!w->arrays_loaded || !w->d_gs)
      throw std::runtime_error("WfsBatch: load_arrays / initgs have not been called");
    if (w->device != device) throw std::invalid_argument("WfsBatch: WFS on different devices");
    if (w->fakecam)
      throw std::invalid_argument("WfsBatch: fakecam WFS (camera image) are run through comp_image");
    if (w->roket)
      throw std::invalid_argument("WfsBatch: roket WFS (image before the noise) are run through comp_image");
    if (w->d_submask != nullptr)
      throw std::invalid_argument("WfsBatch: WFS with a field stop are run through comp_image");
    if (w->lgs) throw std::invalid_argument("WfsBatch: LGS WFS are run through comp_image");
    int64_t d[2] = {1, 2 * w->nvalid};
    d_slopes.push_back(new CarmaObj<float>(context, d));
    geoms.push_back(w->geom());
    shwfs_frame_args f = w->frame_args(noise);
    SutraCentroider<float, float>* c = centro[i];
    if (c) {
      if (c->wfs != w) throw std::invalid_argument("WfsBatch: centroider not attached to its WFS");
      if (c->get_type() == "bpcog" || c->get_type() == "corr")
        throw std::invalid_argument("WfsBatch: " + c->get_type() +
                                    " centroiders run after the frame (get_cog)");
      if (c->get_type() == "wcog") {
        auto* wc = dynamic_cast<SutraCentroiderWcog<float, float>*>(c);
        if (!wc || !wc->d_weights ||
            wc->d_weights->get_nb_elements() < w->npix * w->npix * w->nvalid)
          throw std::runtime_error("WfsBatch: WCoG weights not initialised");
      }
      cargs[i] = c->centroid_args(c->d_intensities->get_data(), d_slopes.back()->get_data(),
                                  (int32_t)w->nvalid, (int32_t)w->npix,
                                  (int32_t)w->d_binimg->get_dims(1));
      // the learned centroider runs per WFS (get_cog), not inside a batch
      if (cargs[i].kind == SHWFS_CENTROID_MLP) cargs[i].kind = SHWFS_CENTROID_NONE;
      f.centroid = &cargs[i];
    }
    frames.push_back(f);
  }
  current_context->set_active_device(device, 1);
  shwfs_throw_if(shwfs_batch_create(geoms.data(), frames.data(), (int32_t)wfs.size(), &handle),
                 "WfsBatch");
}

SutraWfsBatch::~SutraWfsBatch() {
  shwfs_batch_destroy(handle);
  for (auto* s : d_slopes) delete s;
}

void SutraWfsBatch::set_tiles(const vector<int32_t>& methods, const vector<int32_t>& splits,
                              int32_t grid_, int32_t remap_) {
  if (methods.size() != wfs.size() || splits.size() != wfs.size())
    throw std::invalid_argument("WfsBatch.set_tiles: one method and one split per WFS");
  vector<shwfs_batch_tile> t(wfs.size());
  for (size_t i = 0; i < wfs.size(); ++i) t[i] = shwfs_batch_tile{methods[i], splits[i]};
  shwfs_throw_if(shwfs_batch_set_tiles(handle, t.data(), grid_, remap_), "WfsBatch.set_tiles");
}

vector<int32_t> SutraWfsBatch::methods() const {
  vector<shwfs_batch_tile> t(wfs.size());
  shwfs_throw_if(shwfs_batch_get_tiles(handle, t.data(), nullptr, nullptr), "WfsBatch.methods");
  vector<int32_t> m;
  for (auto& x : t) m.push_back(x.method);
  return m;
}

vector<int32_t> SutraWfsBatch::splits() const {
  vector<shwfs_batch_tile> t(wfs.size());
  shwfs_throw_if(shwfs_batch_get_tiles(handle, t.data(), nullptr, nullptr), "WfsBatch.splits");
  vector<int32_t> m;
  for (auto& x : t) m.push_back(x.split);
  return m;
}

int32_t SutraWfsBatch::grid() const {
  int32_t g = 0;
  shwfs_throw_if(shwfs_batch_get_tiles(handle, nullptr, &g, nullptr), "WfsBatch.grid");
  return g;
}

int32_t SutraWfsBatch::remap() const {
  int32_t r = 0;
  shwfs_throw_if(shwfs_batch_get_tiles(handle, nullptr, nullptr, &r), "WfsBatch.remap");
  return r;
}

// Batches read rendered screens: pending raytrace layers are drawn first.
void SutraWfsBatch::render_sources() {
  for (auto* w : wfs)
    if (w->d_gs) w->d_gs->render_layers(nullptr);
}

float SutraWfsBatch::autotune(uint32_t method_mask, int32_t reps) {
  current_context->set_active_device(device, 1);
  render_sources();
  float us = 0.f;
  shwfs_throw_if(shwfs_batch_autotune(handle, method_mask, reps, nullptr, &us), "WfsBatch.autotune");
  return us;
}

void SutraWfsBatch::set_frame(uint32_t f) {
  shwfs_throw_if(shwfs_batch_set_frame(handle, f, nullptr), "WfsBatch.set_frame");
}

uint32_t SutraWfsBatch::frame() {
  uint32_t f = 0;
  shwfs_throw_if(shwfs_batch_get_frame(handle, &f), "WfsBatch.frame");
  return f;
}

void SutraWfsBatch::launch() {
  current_context->set_active_device(device, 1);
  render_sources();
  shwfs_throw_if(shwfs_batch_launch(handle, nullptr), "WfsBatch.launch");
}

void SutraWfsBatch::capture(int32_t frames) {
  current_context->set_active_device(device, 1);
  render_sources();
  shwfs_throw_if(shwfs_batch_capture(handle, frames, nullptr), "WfsBatch.capture");
}

static double timed(const std::function<void()>& body) {
  hipEvent_t e0, e1;
  carma_safe_call(hipEventCreate(&e0));
  carma_safe_call(hipEventCreate(&e1));
  // the batch's NULL-stream work runs on a blocking stream ordered with the
  // default stream: events on the default stream bracket it
  carma_safe_call(hipEventRecord(e0, nullptr));
  body();
  carma_safe_call(hipEventRecord(e1, nullptr));
  carma_safe_call(hipEventSynchronize(e1));
  float ms = 0.f;
  carma_safe_call(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

double SutraWfsBatch::replay(int32_t ngraphs) {
  current_context->set_active_device(device, 1);
  render_sources();
  return timed([&] { shwfs_throw_if(shwfs_batch_replay(handle, ngraphs, nullptr), "WfsBatch.replay"); });
}

double SutraWfsBatch::run(int32_t nframes) {
  current_context->set_active_device(device, 1);
  render_sources();
  return timed([&] {
    for (int32_t k = 0; k < nframes; ++k)
      shwfs_throw_if(shwfs_batch_launch(handle, nullptr), "WfsBatch.run");
  });
}
