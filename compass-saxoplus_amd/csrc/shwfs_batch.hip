// Batched frames over several SH WFS: task table, one launch per frame,
// hipGraph capture of consecutive frames, and the tile autotuner.
//
// Replaces the sequential per-WFS model of the reference (SutraSensors with
// shared scratch and one cuFFT plan per WFS shape, libsutra/src.cpp/
// sutra_sensors.cpp:42-91; the supervisor's loop over WFS,
// shesha/supervisor/compassSupervisor.py:256-265).
#include <algorithm>
#include <vector>

#include "shwfs_batch.hpp"

using namespace shwfs;

struct shwfs_batch {
  int device = 0;
  std::vector<shwfs_sh_geom> geoms;
  std::vector<shwfs_frame_args> frames;  // centroid pointer not kept
  std::vector<ImageParams> params;
  std::vector<shwfs_batch_tile> tiles;
  int grid = 0;   // 0: one workgroup per tile
  int remap = 0;  // XCD-contiguous tile ranges
  bool noisy = false;
  // device task table
  BatchTask* d_tasks = nullptr;
  TaskStarts starts{};  // first tiles of the first kTaskStarts tasks
  TaskArgs args{};      // parameters of the first kTaskArgs tasks
  int ntasks = 0, total_tiles = 0;
  int kind_set = 0;  // BatchKindSet covering the tasks
  // every task the same MFMA variant: launched through the specialised
  // same-kind kernel (method >= 0) instead of the per-kind dispatch kernel
  int uni_method = -1, uni_nfft = 0, uni_pdiam = 0, uni_nr = 0;
  bool uni_full = false;
  // stage groups: the tasks fall into three stage kinds for which a group
  // kernel is compiled (group_kernel >= 0, shwfs_batch_groups.hip)
  int group_kernel = -1;
  int group_wpe = 4;  // waves per SIMD of the noise-free group kernel (3 or 4, autotuned)
  GroupArgs gargs{};
  size_t lds = 0;  // bytes of LDS per wave
  int waves = 1;   // waves per workgroup of the mixed kernel (kBatchWaves when the slots fit)
  uint32_t* d_frame = nullptr;
  hipStream_t own_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_frames = 0;
};

namespace {

__global__ void frame_advance_kernel(uint32_t* frame, uint32_t n) { frame[0] += n; }
__global__ void frame_set_kernel(uint32_t* frame, uint32_t v) { frame[0] = v; }

bool fft_kind(int nfft, int split, int* kind) {
  const int base = nfft == 16 ? BK_FFT16_1 : nfft == 32 ? BK_FFT32_1 : nfft == 64 ? BK_FFT64_1 : -1;
  if (base < 0) return false;
  *kind = base + (split == 2 ? 1 : 0);
  return true;
}

bool dft_kind(int nfft, int pdiam, int method, int* kind) {
  const bool bf = method == SHWFS_METHOD_DFT_BF16X3;
  if (!bf && method != SHWFS_METHOD_DFT_FP16) return false;
  const int off = bf ? BK_BF16_16 : BK_FP16_16;
  if (pdiam <= 16) {
    if (nfft == 16) return *kind = off + 0, true;
    if (nfft == 32) return *kind = off + 1, true;
    if (nfft == 64) return *kind = off + 2, true;
  } else if (pdiam <= 32) {
    if (nfft == 32) return *kind = off + 3, true;
    if (nfft == 64) return *kind = off + 4, true;
  }
  return false;
}

// Per-tile cost estimate used to put the heaviest tiles first.
double tile_cost(const BatchTask& t, int nfft) {
  return (double)t.per_tile * nfft * nfft;
}

int32_t fill_task(shwfs_batch* b, int i, const shwfs_batch_tile& tile, hipStream_t st,
                  BatchTask* out, size_t* lds) {
  const shwfs_sh_geom& g = b->geoms[i];
  BatchTask t{};
  t.P = b->params[i];
  const int nfft = g.nfft, pdiam = g.nphase, npix = g.npix;
  int32_t rc = check_geom(&g, tile.method);
  if (rc) return rc;
  if (tile.method == SHWFS_METHOD_FFT_FP32) {
    int split = tile.split;
    if (split == 0) split = 2 * pdiam <= nfft ? 2 : 1;
    if (split != 1 && split != 2) return fail(SHWFS_ERR_ARG, "batch: FFT split must be 0, 1 or 2");
    if (split == 2 && 2 * pdiam > nfft)
      return fail(SHWFS_ERR_ARG, "batch: FFT split 2 needs 2*nphase <= nfft");
    if (!fft_kind(nfft, split, &t.kind))
      return fail(SHWFS_ERR_UNSUPPORTED, "batch: FFT stage needs nfft in {16, 32, 64}");
    t.per_tile = 64 / nfft;
    int slot = 0;
    switch (t.kind) {
      case BK_FFT16_1: slot = FftLayout<16, 1>::slot_floats(pdiam, npix); break;
      case BK_FFT16_2: slot = FftLayout<16, 2>::slot_floats(pdiam, npix); break;
      case BK_FFT32_1: slot = FftLayout<32, 1>::slot_floats(pdiam, npix); break;
      case BK_FFT32_2: slot = FftLayout<32, 2>::slot_floats(pdiam, npix); break;
      case BK_FFT64_1: slot = FftLayout<64, 1>::slot_floats(pdiam, npix); break;
      default: slot = FftLayout<64, 2>::slot_floats(pdiam, npix); break;
    }
    *lds = sizeof(float) * (size_t)t.per_tile * slot;
  } else {
    if (!dft_kind(nfft, pdiam, tile.method, &t.kind))
      return fail(SHWFS_ERR_UNSUPPORTED, "batch: DFT stage needs nfft in {16, 32, 64}, pdiam <= 32");
    t.per_tile = kDftTaskSubaps;
    t.nr = dft_register_rebin(t.P);
    const bool gather = t.nr == 0;
    int wf = 0;
    switch (t.kind - (tile.method == SHWFS_METHOD_DFT_BF16X3 ? BK_BF16_16 : BK_FP16_16)) {
      case 0: wf = subaps_floats<16, 1>(npix, gather, kDftTaskSubaps); break;
      case 1: wf = subaps_floats<32, 1>(npix, gather, kDftTaskSubaps); break;
      case 2: wf = subaps_floats<64, 1>(npix, gather, kDftTaskSubaps); break;
      case 3: wf = subaps_floats<32, 2>(npix, gather, kDftTaskSubaps); break;
      default: wf = subaps_floats<64, 2>(npix, gather, kDftTaskSubaps); break;
    }
    *lds = sizeof(float) * (size_t)wf;
    t.wtab = dft_operand_table(nfft, pdiam, tile.method, st);
    if (!t.wtab) return SHWFS_ERR_HIP;
  }
  t.ntiles = (g.nvalid + t.per_tile - 1) / t.per_tile;
  *out = t;
  return SHWFS_OK;
}

// Stage code of a task (stage_code): the FFT with its row split, or the MFMA
// stage with its window class, full-window flag and register rebin code.
int task_stage_code(const BatchTask& t, const shwfs_sh_geom& g, int method) {
  const int n = g.nfft;
  if (method == SHWFS_METHOD_FFT_FP32) {
    const int split = (t.kind == BK_FFT16_2 || t.kind == BK_FFT32_2 || t.kind == BK_FFT64_2) ? 2 : 1;
    return stage_code(-1, n, split, false, 0);
  }
  const int q = g.nphase <= 16 ? 1 : 2;
  int nr = dft_register_rebin(t.P);
  if (nr == 4 && dft_half_grid(t.P, n)) nr |= kHalfGrid;
  return stage_code(method, n, q, g.nphase == 16 * q, nr);
}

// Group mode: the tasks' stage codes are three distinct codes in the order of
// a compiled group kernel (at most kTaskArgs tasks).  Reorders `tasks` by
// group (heaviest tiles first inside a group), sets per-tile counts, the
// group starts and operand tables, and the LDS slot; returns false when no
// group kernel covers the batch.
bool setup_groups(shwfs_batch* b, std::vector<BatchTask>& tasks, std::vector<double>& cost,
                  std::vector<int>& order, size_t* lds, hipStream_t st) {
  const int n = (int)tasks.size();
  if (n > kTaskArgs) return false;
  std::vector<int> code(n);
  std::vector<int> distinct;
  for (int i = 0; i < n; ++i) {
    code[i] = task_stage_code(tasks[i], b->geoms[i], b->tiles[i].method);
    if (std::find(distinct.begin(), distinct.end(), code[i]) == distinct.end())
      distinct.push_back(code[i]);
  }
  if (distinct.size() != 3) return false;
  std::sort(distinct.begin(), distinct.end());
  int index = -1;
  do {
    index = group_kernel_index(distinct.data());
  } while (index < 0 && std::next_permutation(distinct.begin(), distinct.end()));
  if (index < 0) return false;
  const int* gc = group_kernel_codes(index);
  std::vector<int> group(n);
  size_t wlds = 0;
  for (int i = 0; i < n; ++i) {
    group[i] = (int)(std::find(gc, gc + 3, code[i]) - gc);
    const shwfs_sh_geom& g = b->geoms[i];
    BatchTask& t = tasks[i];
    size_t l = 0;
    if (b->tiles[i].method == SHWFS_METHOD_FFT_FP32) {
      t.per_tile = 64 / g.nfft;
      const int split = (t.kind == BK_FFT16_2 || t.kind == BK_FFT32_2 || t.kind == BK_FFT64_2) ? 2 : 1;
      int f = 0;
      switch (g.nfft * 4 + split) {
        case 16 * 4 + 1: f = FftLayout<16, 1>::slot_floats(g.nphase, g.npix); break;
        case 16 * 4 + 2: f = FftLayout<16, 2>::slot_floats(g.nphase, g.npix); break;
        case 32 * 4 + 1: f = FftLayout<32, 1>::slot_floats(g.nphase, g.npix); break;
        case 32 * 4 + 2: f = FftLayout<32, 2>::slot_floats(g.nphase, g.npix); break;
        case 64 * 4 + 1: f = FftLayout<64, 1>::slot_floats(g.nphase, g.npix); break;
        default: f = FftLayout<64, 2>::slot_floats(g.nphase, g.npix); break;
      }
      l = sizeof(float) * (size_t)t.per_tile * f;
    } else {
      t.per_tile = dft_stage_subaps(g.nfft);
      const bool gather = t.nr == 0;
      const int q = g.nphase <= 16 ? 1 : 2;
      int f = 0;
      switch (g.nfft * 4 + q) {
        case 16 * 4 + 1: f = subaps_floats<16, 1>(g.npix, gather, kDftTaskSubaps); break;
        case 32 * 4 + 1: f = subaps_floats<32, 1>(g.npix, gather, kDftTaskSubaps); break;
        case 32 * 4 + 2: f = subaps_floats<32, 2>(g.npix, gather, kDftTaskSubaps); break;
        // one sub-aperture per wave (dft_subap): the single-frame kernel's slot
        case 64 * 4 + 1: f = DftLayout<64, 1>::wave_floats(g.npix, gather); break;
        default: f = DftLayout<64, 2>::wave_floats(g.npix, gather); break;
      }
      l = sizeof(float) * (size_t)f;
    }
    t.ntiles = (g.nvalid + t.per_tile - 1) / t.per_tile;
    // the 64-point MFMA stage shares each sub-aperture among the waves of a
    // workgroup: one tile per wave, a quarter sub-aperture each
    if (b->tiles[i].method != SHWFS_METHOD_FFT_FP32 && g.nfft == 64) t.ntiles = g.nvalid * kDftWaves;
    cost[i] = tile_cost(t, g.nfft);
    wlds = std::max(wlds, l);
  }
  if (wlds * kDftWaves > max_lds_per_block()) return false;
  // group order, heaviest tiles first inside each group
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    return group[a] != group[c] ? group[a] < group[c] : cost[a] > cost[c];
  });
  GroupArgs ga{};
  int first = 0;
  for (int gi = 0; gi < 3; ++gi) {
    ga.first[gi] = first;
    ga.wtab[gi] = nullptr;
    for (int k = 0; k < n; ++k) {
      const int i = order[k];
      if (group[i] != gi) continue;
      first += tasks[i].ntiles;
      if (b->tiles[i].method != SHWFS_METHOD_FFT_FP32) ga.wtab[gi] = tasks[i].wtab;
    }
  }
  (void)st;
  b->group_kernel = index;
  b->gargs = ga;
  *lds = wlds;
  return true;
}

int32_t build_tasks(shwfs_batch* b, hipStream_t st) {
  const int n = (int)b->geoms.size();
  // the table may still be read by frames in flight on any stream
  SHWFS_HIP(hipDeviceSynchronize());
  std::vector<BatchTask> tasks(n);
  std::vector<double> cost(n);
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    size_t l = 0;
    int32_t rc = fill_task(b, i, b->tiles[i], st, &tasks[i], &l);
    if (rc) return rc;
    lds = std::max(lds, l);
    cost[i] = tile_cost(tasks[i], b->geoms[i].nfft);
  }
  // same-kind MFMA batch (one kind, rebin factor and window class for all, at
  // most kTaskArgs WFS: their parameters travel as kernel arguments): tiles of
  // kDftTaskSubaps sub-apertures, one wave each, kDftWaves waves' LDS slots
  // per workgroup
  const int m0 = b->tiles[0].method;
  bool uni = m0 != SHWFS_METHOD_FFT_FP32 && n <= kTaskArgs;
  for (int i = 0; i < n && uni; ++i) {
    uni = tasks[i].kind == tasks[0].kind && tasks[i].nr == tasks[0].nr &&
          b->geoms[i].nphase == b->geoms[0].nphase;
  }
  if (uni) {
    size_t wlds = 0;
    for (int i = 0; i < n; ++i) {
      const int npix = b->geoms[i].npix;
      const bool gather = tasks[i].nr == 0;
      int f = 0;
      switch (tasks[i].kind - (m0 == SHWFS_METHOD_DFT_BF16X3 ? BK_BF16_16 : BK_FP16_16)) {
        case 0: f = subaps_floats<16, 1>(npix, gather, kDftTaskSubaps); break;
        case 1: f = subaps_floats<32, 1>(npix, gather, kDftTaskSubaps); break;
        case 2: f = subaps_floats<64, 1>(npix, gather, kDftTaskSubaps); break;
        case 3: f = subaps_floats<32, 2>(npix, gather, kDftTaskSubaps); break;
        default: f = subaps_floats<64, 2>(npix, gather, kDftTaskSubaps); break;
      }
      wlds = std::max(wlds, sizeof(float) * (size_t)f);
    }
    if (wlds * kDftWaves > max_lds_per_block()) uni = false;
    if (uni) {
      lds = wlds;
      for (int i = 0; i < n; ++i) {
        tasks[i].per_tile = kDftTaskSubaps;
        tasks[i].ntiles = (b->geoms[i].nvalid + kDftTaskSubaps - 1) / kDftTaskSubaps;
        cost[i] = tile_cost(tasks[i], b->geoms[i].nfft);
      }
    }
  }
  std::vector<int> order(n);
  b->group_kernel = -1;
  bool grouped = false;
  if (!uni) {
    size_t glds = 0;
    grouped = setup_groups(b, tasks, cost, order, &glds, st);
    if (grouped) lds = glds;
  }
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "batch: LDS footprint exceeds the per-workgroup limit");
  // waves per workgroup of the mixed kernel: kBatchWaves slots if they fit
  int waves = kBatchWaves;
  while (waves > 1 && lds * waves > max_lds_per_block()) waves /= 2;
  if (!grouped) {
    // heaviest tiles first (they start first and the light ones fill in)
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  }
  std::vector<BatchTask> sorted(n);
  int first = 0;
  for (int k = 0; k < n; ++k) {
    sorted[k] = tasks[order[k]];
    sorted[k].first_tile = first;
    first += sorted[k].ntiles;
  }
  if (b->d_tasks == nullptr) SHWFS_HIP(hipMalloc((void**)&b->d_tasks, sizeof(BatchTask) * n));
  SHWFS_HIP(hipMemcpy(b->d_tasks, sorted.data(), sizeof(BatchTask) * n, hipMemcpyHostToDevice));
  for (int k = 0; k < kTaskStarts; ++k) b->starts.first[k] = k < n ? sorted[k].first_tile : kNoTask;
  b->args = TaskArgs{};
  for (int k = 0; k < n && k < kTaskArgs; ++k) b->args.P[k] = sorted[k].P;
  b->ntasks = n;
  b->total_tiles = first;
  b->lds = lds;
  b->waves = waves;
  int set = 0;
  for (const auto& t : sorted) set |= batch_kind_set(t.kind);
  b->kind_set = (set == BKS_FFT || set == BKS_BF16 || set == BKS_FP16) ? set : BKS_ALL;
  b->uni_method = -1;
  if (uni) {
    b->uni_method = m0;
    b->uni_nfft = b->geoms[0].nfft;
    b->uni_pdiam = b->geoms[0].nphase;
    b->uni_nr = tasks[0].nr;
    // the half-grid 4x4 form when every WFS of the batch allows it
    bool half = true;
    for (int i = 0; i < n; ++i) half = half && dft_half_grid(tasks[i].P, b->uni_nfft);
    if (half) b->uni_nr |= kHalfGrid;
    b->uni_full = b->uni_pdiam == 16 || b->uni_pdiam == 32;
  }
  return SHWFS_OK;
}

void drop_graph(shwfs_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_frames = 0;
}

hipStream_t pick_stream(shwfs_batch* b, void* stream) {
  return stream ? (hipStream_t)stream : b->own_stream;
}

// One frame (frame counter + frame_off) without advancing the counter.
int32_t launch_frame(shwfs_batch* b, uint32_t frame_off, hipStream_t st) {
  for (size_t i = 0; i < b->frames.size(); ++i) {
    const shwfs_frame_args& f = b->frames[i];
    if (f.d_binimg && f.noise_background > 0.0f) {
      const long n = (long)b->params[i].nsub * b->params[i].nsub;
      int32_t rc = launch_background(f.d_binimg, n, f.noise_background, f.seed, f.call + frame_off,
                                     st, b->d_frame);
      if (rc) return rc;
    }
  }
  // mixed kernel: workgroups of b->waves waves, one tile per wave
  const int wgs = (b->total_tiles + b->waves - 1) / b->waves;
  const int grid = b->grid > 0 ? std::min(b->grid, wgs) : wgs;
  if (b->uni_method >= 0)
    return launch_dft_tasks(b->uni_method, b->uni_nfft, b->uni_pdiam, b->uni_full, b->uni_nr,
                            b->noisy, b->d_tasks, b->starts, b->args, b->ntasks, b->total_tiles, grid,
                            b->lds,
                            b->d_frame, frame_off, b->remap, st);
  if (b->group_kernel >= 0)
    return launch_batch_groups(b->group_kernel, b->noisy, b->group_wpe, b->starts, b->args,
                               b->ntasks, b->total_tiles, b->gargs, b->lds, b->d_frame, frame_off,
                               st);
  if (b->noisy)
    return launch_batch_noisy(b->kind_set, b->d_tasks, b->starts, b->ntasks, b->total_tiles, grid,
                              b->waves, b->lds, b->d_frame, frame_off, b->remap, st);
  return launch_batch_quiet(b->kind_set, b->d_tasks, b->starts, b->ntasks, b->total_tiles, grid,
                            b->waves, b->lds, b->d_frame, frame_off, b->remap, st);
}

int32_t advance(shwfs_batch* b, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(frame_advance_kernel, dim3(1), dim3(1), 0, st, b->d_frame, n);
  SHWFS_LAUNCHED("frame_advance_kernel");
  return SHWFS_OK;
}

int resident_blocks(shwfs_batch* b) {
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  const int per = b->noisy ? batch_blocks_per_cu_noisy(b->kind_set, b->waves, b->lds)
                           : batch_blocks_per_cu_quiet(b->kind_set, b->waves, b->lds);
  return cus * (per > 0 ? per : 1);
}

// Time per frame (us) of the current configuration: median of 3 runs of
// `reps` back-to-back frames.
int32_t time_launches(shwfs_batch* b, int reps, hipStream_t st, float* us);

// Median time per frame of the current configuration; for a noise-free
// group kernel, of its faster register budget (kept in b->group_wpe).
int32_t time_config(shwfs_batch* b, int reps, hipStream_t st, float* us) {
  int32_t rc = build_tasks(b, st);
  if (rc) return rc;
  if (b->group_kernel < 0 || b->noisy) return time_launches(b, reps, st, us);
  float best = 0.f;
  int best_wpe = 4;
  for (int wpe : {4, 3}) {
    b->group_wpe = wpe;
    float t = 0.f;
    rc = time_launches(b, reps, st, &t);
    if (rc) return rc;
    if (wpe == 4 || t < best) {
      best = t;
      best_wpe = wpe;
    }
  }
  b->group_wpe = best_wpe;
  *us = best;
  return SHWFS_OK;
}

int32_t time_launches(shwfs_batch* b, int reps, hipStream_t st, float* us) {
  int32_t rc = SHWFS_OK;
  hipEvent_t e0, e1;
  SHWFS_HIP(hipEventCreate(&e0));
  SHWFS_HIP(hipEventCreate(&e1));
  float runs[3];
  for (int w = 0; w < 2 && rc == SHWFS_OK; ++w) rc = launch_frame(b, 0, st);
  for (int r = 0; r < 3 && rc == SHWFS_OK; ++r) {
    (void)hipEventRecord(e0, st);
    for (int k = 0; k < reps && rc == SHWFS_OK; ++k) rc = launch_frame(b, 0, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(SHWFS_ERR_HIP, "batch: timing failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    runs[r] = 1e3f * ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  std::sort(runs, runs + 3);
  *us = runs[1];
  return SHWFS_OK;
}

}  // namespace

extern "C" {

int32_t shwfs_batch_create(const shwfs_sh_geom* geoms, const shwfs_frame_args* frames,
                           int32_t nwfs, shwfs_batch** out) {
  if (!out) return fail(SHWFS_ERR_ARG, "batch_create: out is NULL");
  *out = nullptr;
  if (!geoms || !frames || nwfs <= 0) return fail(SHWFS_ERR_ARG, "batch_create: no WFS given");
  auto* b = new (std::nothrow) shwfs_batch();
  if (!b) return fail(SHWFS_ERR_NOMEM, "batch_create: out of host memory");
  int32_t rc = SHWFS_OK;
  if (hipGetDevice(&b->device) != hipSuccess) rc = fail(SHWFS_ERR_HIP, "batch_create: no device");
  for (int i = 0; i < nwfs && rc == SHWFS_OK; ++i) {
    ImageParams p;
    rc = image_params(&geoms[i], &frames[i], &p);
    if (rc) break;
    if (p.ckind == SHWFS_CENTROID_MLP) {
      rc = fail(SHWFS_ERR_UNSUPPORTED, "batch_create: the learned centroider is not fused in batches");
      break;
    }
    if (ext_needed(&geoms[i])) {
      rc = fail(SHWFS_ERR_UNSUPPORTED, "batch_create: ntot != nfft / kernconv / LGS run per WFS");
      break;
    }
    if (frames[i].nlayers > 0) {
      rc = fail(SHWFS_ERR_UNSUPPORTED, "batch_create: batches read rendered phase screens (no layers)");
      break;
    }
    b->geoms.push_back(geoms[i]);
    shwfs_frame_args f = frames[i];
    f.centroid = nullptr;
    b->frames.push_back(f);
    b->params.push_back(p);
    b->tiles.push_back(shwfs_batch_tile{frames[i].method, 0});
    if (frames[i].noise > -1.0f) b->noisy = true;
  }
  if (rc == SHWFS_OK && hipMalloc((void**)&b->d_frame, sizeof(uint32_t)) != hipSuccess)
    rc = fail(SHWFS_ERR_HIP, "batch_create: hipMalloc failed");
  if (rc == SHWFS_OK && hipStreamCreate(&b->own_stream) != hipSuccess)
    rc = fail(SHWFS_ERR_HIP, "batch_create: hipStreamCreate failed");
  if (rc == SHWFS_OK) {
    hipLaunchKernelGGL(frame_set_kernel, dim3(1), dim3(1), 0, b->own_stream, b->d_frame, 0u);
    if (hipGetLastError() != hipSuccess) rc = fail(SHWFS_ERR_HIP, "batch_create: counter init failed");
  }
  if (rc == SHWFS_OK) rc = build_tasks(b, b->own_stream);
  if (rc == SHWFS_OK && hipStreamSynchronize(b->own_stream) != hipSuccess)
    rc = fail(SHWFS_ERR_HIP, "batch_create: synchronisation failed");
  if (rc) {
    shwfs_batch_destroy(b);
    return rc;
  }
  *out = b;
  return SHWFS_OK;
}

int32_t shwfs_batch_destroy(shwfs_batch* b) {
  if (!b) return SHWFS_OK;
  drop_graph(b);
  if (b->own_stream) (void)hipStreamSynchronize(b->own_stream);
  if (b->d_tasks) (void)hipFree(b->d_tasks);
  if (b->d_frame) (void)hipFree(b->d_frame);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
  return SHWFS_OK;
}

int32_t shwfs_batch_set_tiles(shwfs_batch* b, const shwfs_batch_tile* tiles, int32_t grid,
                              int32_t remap) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_set_tiles: NULL batch");
  if (grid < 0) return fail(SHWFS_ERR_ARG, "batch_set_tiles: grid must be >= 0");
  std::vector<shwfs_batch_tile> old = b->tiles;
  if (tiles) b->tiles.assign(tiles, tiles + b->geoms.size());
  int32_t rc = build_tasks(b, b->own_stream);
  if (rc) {
    b->tiles = old;
    (void)build_tasks(b, b->own_stream);
    return rc;
  }
  b->grid = grid;
  b->remap = remap ? 1 : 0;
  drop_graph(b);
  return SHWFS_OK;
}

int32_t shwfs_batch_get_tiles(const shwfs_batch* b, shwfs_batch_tile* tiles, int32_t* grid,
                              int32_t* remap) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_get_tiles: NULL batch");
  if (tiles) std::copy(b->tiles.begin(), b->tiles.end(), tiles);
  if (grid) *grid = b->grid;
  if (remap) *remap = b->remap;
  return SHWFS_OK;
}

int32_t shwfs_batch_set_frame(shwfs_batch* b, uint32_t frame, void* stream) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_set_frame: NULL batch");
  hipLaunchKernelGGL(frame_set_kernel, dim3(1), dim3(1), 0, pick_stream(b, stream), b->d_frame,
                     frame);
  SHWFS_LAUNCHED("frame_set_kernel");
  return SHWFS_OK;
}

int32_t shwfs_batch_get_frame(shwfs_batch* b, uint32_t* frame) {
  if (!b || !frame) return fail(SHWFS_ERR_ARG, "batch_get_frame: NULL argument");
  SHWFS_HIP(hipDeviceSynchronize());
  SHWFS_HIP(hipMemcpy(frame, b->d_frame, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SHWFS_OK;
}

int32_t shwfs_batch_launch(shwfs_batch* b, void* stream) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_launch: NULL batch");
  hipStream_t st = pick_stream(b, stream);
  int32_t rc = launch_frame(b, 0, st);
  if (rc) return rc;
  return advance(b, 1, st);
}

int32_t shwfs_batch_capture(shwfs_batch* b, int32_t frames, void* stream) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_capture: NULL batch");
  if (frames <= 0) return fail(SHWFS_ERR_ARG, "batch_capture: frames must be positive");
  hipStream_t st = pick_stream(b, stream);
  drop_graph(b);
  SHWFS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int32_t rc = SHWFS_OK;
  for (int k = 0; k < frames && rc == SHWFS_OK; ++k) rc = launch_frame(b, (uint32_t)k, st);
  if (rc == SHWFS_OK) rc = advance(b, (uint32_t)frames, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g) return fail(SHWFS_ERR_HIP, "batch_capture: stream capture failed");
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph(b);
    return fail(SHWFS_ERR_HIP, "batch_capture: graph instantiation failed");
  }
  b->graph_frames = frames;
  return SHWFS_OK;
}

int32_t shwfs_batch_replay(shwfs_batch* b, int32_t ngraphs, void* stream) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_replay: NULL batch");
  if (!b->exec) return fail(SHWFS_ERR_STATE, "batch_replay: no captured graph");
  if (ngraphs < 0) return fail(SHWFS_ERR_ARG, "batch_replay: negative count");
  hipStream_t st = pick_stream(b, stream);
  for (int k = 0; k < ngraphs; ++k) SHWFS_HIP(hipGraphLaunch(b->exec, st));
  return SHWFS_OK;
}

int32_t shwfs_batch_autotune(shwfs_batch* b, uint32_t method_mask, int32_t reps, void* stream,
                             float* best_us) {
  if (!b) return fail(SHWFS_ERR_ARG, "batch_autotune: NULL batch");
  if (reps <= 0) return fail(SHWFS_ERR_ARG, "batch_autotune: reps must be positive");
  if ((method_mask & 7u) == 0) return fail(SHWFS_ERR_ARG, "batch_autotune: empty method mask");
  hipStream_t st = pick_stream(b, stream);
  drop_graph(b);
  const int n = (int)b->geoms.size();
  // start from a valid configuration: the first allowed method per WFS
  for (int i = 0; i < n; ++i) {
    if ((method_mask >> b->tiles[i].method) & 1u) continue;
    for (int m = 0; m < 3; ++m)
      if (((method_mask >> m) & 1u) && check_geom(&b->geoms[i], m) == SHWFS_OK) {
        b->tiles[i] = shwfs_batch_tile{m, 0};
        break;
      }
  }
  float best = 0.f;
  int32_t rc = time_config(b, reps, st, &best);
  if (rc) return rc;
  // one method per geometry class (WFS of equal nfft, nphase, npix, nrebin
  // move together), every combination over at most three classes, else every
  // WFS on one method: these reach the same-kind and the stage-group kernels,
  // which changing one WFS at a time from a uniform start would not
  {
    std::vector<int> cls(n, -1);
    int ncls = 0;
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < i && cls[i] < 0; ++j) {
        const shwfs_sh_geom &a = b->geoms[i], &c = b->geoms[j];
        if (a.nfft == c.nfft && a.nphase == c.nphase && a.npix == c.npix && a.nrebin == c.nrebin)
          cls[i] = cls[j];
      }
      if (cls[i] < 0) cls[i] = ncls++;
    }
    const std::vector<shwfs_batch_tile> start = b->tiles;
    std::vector<shwfs_batch_tile> keep = start;
    const int per_class = ncls <= 3;
    int combos = 1;
    for (int c = 0; c < (per_class ? ncls : 1); ++c) combos *= 3;
    for (int code = 0; code < combos; ++code) {
      b->tiles = start;
      bool ok = true, changed = false;
      for (int i = 0; i < n && ok; ++i) {
        int m = code;
        for (int c = 0; c < (per_class ? cls[i] : 0); ++c) m /= 3;
        m %= 3;
        const shwfs_sh_geom& g = b->geoms[i];
        int k;
        ok = ((method_mask >> m) & 1u) && check_geom(&g, m) == SHWFS_OK &&
             (m == SHWFS_METHOD_FFT_FP32 || dft_kind(g.nfft, g.nphase, m, &k));
        if (ok && b->tiles[i].method != m) {
          b->tiles[i] = shwfs_batch_tile{m, 0};
          changed = true;
        }
      }
      float us = 0.f;
      if (ok && changed && time_config(b, reps, st, &us) == SHWFS_OK && us < best) {
        best = us;
        keep = b->tiles;
      }
    }
    b->tiles = keep;
  }
  // coordinate descent over the WFS, heaviest first
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    return (double)b->geoms[a].nvalid * b->geoms[a].nfft * b->geoms[a].nfft >
           (double)b->geoms[c].nvalid * b->geoms[c].nfft * b->geoms[c].nfft;
  });
  for (int i : order) {
    const shwfs_sh_geom& g = b->geoms[i];
    std::vector<shwfs_batch_tile> cands;
    for (int m = 0; m < 3; ++m) {
      if (!((method_mask >> m) & 1u) || check_geom(&g, m) != SHWFS_OK) continue;
      if (m == SHWFS_METHOD_FFT_FP32) {
        if (g.nfft != 16 && g.nfft != 32 && g.nfft != 64) continue;
        cands.push_back(shwfs_batch_tile{m, 1});
        if (2 * g.nphase <= g.nfft) cands.push_back(shwfs_batch_tile{m, 2});
      } else {
        int k;
        if (dft_kind(g.nfft, g.nphase, m, &k)) cands.push_back(shwfs_batch_tile{m, 0});
      }
    }
    shwfs_batch_tile keep = b->tiles[i];
    for (const auto& c : cands) {
      if (c.method == keep.method && c.split == keep.split) continue;
      b->tiles[i] = c;
      float us = 0.f;
      if (time_config(b, reps, st, &us) == SHWFS_OK && us < best) {
        best = us;
        keep = c;
      }
    }
    b->tiles[i] = keep;
  }
  rc = build_tasks(b, st);
  if (rc) return rc;
  // launch shape: one workgroup per tile or persistent grids; XCD mapping
  const int res = resident_blocks(b);
  const int grids[3] = {0, res, 2 * res};
  int best_grid = b->grid, best_remap = b->remap;
  for (int gi = 0; gi < 3; ++gi)
    for (int rm = 0; rm < 2; ++rm) {
      if (grids[gi] == b->grid && rm == b->remap) continue;
      // the same-kind and group kernels take one tile per wave: no persistent
      // grid, and the group kernels no XCD remapping either
      if (b->uni_method >= 0 && grids[gi] != 0) continue;
      if (b->group_kernel >= 0 && (grids[gi] != 0 || rm != 0)) continue;
      const int sg = b->grid, sr = b->remap;
      b->grid = grids[gi];
      b->remap = rm;
      float us = 0.f;
      if (time_config(b, reps, st, &us) == SHWFS_OK && us < best) {
        best = us;
        best_grid = grids[gi];
        best_remap = rm;
      }
      b->grid = sg;
      b->remap = sr;
    }
  b->grid = best_grid;
  b->remap = best_remap;
  // the kept configuration, timed once more: this also settles the register
  // budget of a group kernel (time_config keeps the faster one)
  float final_us = 0.f;
  rc = time_config(b, reps, st, &final_us);
  if (rc) return rc;
  SHWFS_HIP(hipStreamSynchronize(st));
  if (best_us) *best_us = std::min(best, final_us);
  return SHWFS_OK;
}

}  // extern "C"
