// MFMA spot kernels: the zero-padded 2-D FFT of each sub-aperture field
// becomes two dense products with the pruned DFT matrix on the matrix cores.
// With W[k][n] = exp(-2 pi i k n / N) (k < N, n < P16 = 16 Q >= pdiam; the
// zero rows of the padded field are never multiplied) and the real stacking
//     A_W = [[Wr, -Wi], [Wi, Wr]]                 (2N x 2 P16)
// the two stages are
//     stage A   [Gr | Gi]  = [Er | Ei] A_W^T       rows y,         cols (re/im, kx)
//     stage B   [Fr ; Fi]  = A_W [Gr ; Gi]         rows (re/im, ky), cols kx
// on v_mfma_f32_16x16x32 tiles with FP32 accumulation, one wave per
// sub-aperture.  A_W's A-operand fragments are, register for register, the
// B-operand fragments of A_W^T, so one prebuilt table feeds both stages.
// The contraction index (re/im, x) of stage A and (re/im, y) of stage B is
// stored in the order in which an accumulator tile lists its rows (k_slot):
// stage A's accumulators are then stage B's operand as they stand, and each
// lane evaluates the field at the points its operand slots hold; no lane
// exchanges and no LDS pass.
// Operand precisions:
//   * SHWFS_METHOD_DFT_BF16X3: every operand x = hi + lo (bf16 each),
//     product = hi*hi + hi*lo + lo*hi (lo*lo dropped), 3 MFMA per tile;
//   * SHWFS_METHOD_DFT_FP16: fp16 operands, 1 MFMA per tile; the field of a
//     sub-aperture is scaled by a power of two s such that the bound
//     pdiam * max|E| >= max|G| lands in [2^14, 2^15), and the flux by 1/s^2
//     (max|E| <= 1 for a window of unit pupil, measured otherwise).
// The field uses the hardware sine and cosine on the angle in turns (about
// 1e-6 absolute error): these methods trade accuracy for speed and state their
// error against the float64 oracle.  Binning, noise, detector image and
// centroids follow the same definitions as the FFT kernel.
#pragma once

#include <map>
#include <mutex>

#include "shwfs_mfma.hpp"
#include "shwfs_tail.hpp"

namespace shwfs {

// Bit of a compile-time rebin code (rebin factor in bits 0-2) selecting the
// half-grid 4x4 binning form (dft_half_grid).
constexpr int kHalfGrid = 8;

template <int N, int Q>
struct DftLayout {
  static constexpr int P16 = 16 * Q;             // padded phase side (pdiam <= P16 <= N)
  static constexpr int MT = N / 8;               // 16-row tiles of A_W: (re/im, k) = 2N rows
  static constexpr int KS = Q;                   // 32-wide K steps of A_W: (re/im, n) = 2 P16
  static constexpr int NTB = N / 16;             // kx tiles
  static constexpr bool kAwInRegs = MT * KS <= 8;  // operand table held in registers
  static constexpr int PE = P16 + 4;             // field row stride (floats), 16-byte rows
  static constexpr int EPLANE = P16 * PE;        // Er plane, then Ei plane
  static constexpr int EPL = (P16 * P16 + 63) / 64;  // field elements per lane
  // LDS per wave: the field planes; the general-binmap tail adds the |F|^2
  // image (over the planes) and the spot.  The moment totals use 192 doubles
  // (kRedFloats): after the spot (binned in registers) or at the start (over
  // the dead |F|^2 image).
  static constexpr int kRedFloats = 384;
  static __host__ __device__ int wave_floats(int npix, bool gather) {
    const int spot = (npix * npix + 3) & ~3;
    if (!gather) return 2 * EPLANE > spot + kRedFloats ? 2 * EPLANE : spot + kRedFloats;
    const int img = 2 * EPLANE > N * N ? 2 * EPLANE : N * N;
    return (img > kRedFloats ? img : kRedFloats) + spot;
  }
};

template <int PREC>
struct Frag;
template <>
struct Frag<SHWFS_METHOD_DFT_FP16> {
  f16x8 v;
  __device__ void set(int e, float x) { v[e] = (_Float16)x; }
};
template <>
struct Frag<SHWFS_METHOD_DFT_BF16X3> {
  bf16x8 hi, lo;
  __device__ void set(int e, float x) {
    __bf16 h, l;
    split_bf16(x, h, l);
    hi[e] = h;
    lo[e] = l;
  }
};

__device__ __forceinline__ f32x4 mma(const Frag<SHWFS_METHOD_DFT_FP16>& a,
                                     const Frag<SHWFS_METHOD_DFT_FP16>& b, f32x4 c) {
  return mfma_f16(a.v, b.v, c);
}
__device__ __forceinline__ f32x4 mma(const Frag<SHWFS_METHOD_DFT_BF16X3>& a,
                                     const Frag<SHWFS_METHOD_DFT_BF16X3>& b, f32x4 c) {
  c = mfma_bf16(a.lo, b.hi, c);
  c = mfma_bf16(a.hi, b.lo, c);
  return mfma_bf16(a.hi, b.hi, c);
}

// Contraction slot e of K step ks in lane group g (= lane >> 4) of an
// operand fragment holds the (re/im, coordinate) pair
//     Q == 1:  im = e >= 4,  c = 4 g + (e & 3)
//     Q == 2:  im = ks == 1, c = 16 (e >> 2) + 4 g + (e & 3)
// (coordinate x in stage A, y in stage B).  An accumulator tile gives lane
// j + 16 g rows 4 g .. 4 g + 3 of its column j, so the Gr and Gi tiles of
// stage A (Q == 1), or its two row blocks of Gr (then of Gi; Q == 2), fill a
// lane's slots in this order.  Both operands of a product take the same
// order, which leaves the sums over K unchanged.
template <int Q>
__device__ __forceinline__ int slot_coord(int g, int e) {
  return 4 * g + (e & 3) + (Q == 1 ? 0 : 16 * (e >> 2));
}
template <int Q>
__device__ __forceinline__ bool slot_im(int ks, int e) {
  return Q == 1 ? e >= 4 : ks == 1;
}

// Element e of lane l of the A_W fragment (mt, ks).
template <int N, int Q>
__device__ __forceinline__ float aw_value(int mt, int ks, int lane, int e) {
  const int m = 16 * mt + (lane & 15);
  const bool row_im = m >= N, col_im = slot_im<Q>(ks, e);
  const int kx = row_im ? m - N : m;
  const int x = slot_coord<Q>(lane >> 4, e);
  const int idx = (kx * x) & (N - 1);
  const float c = kTw<N>.c[idx];   // cos(2 pi idx / N)
  const float s = kTw<N>.s[idx];   // -sin(2 pi idx / N)
  if (!row_im) return col_im ? -s : c;  // [Wr, -Wi] with Wi = -sin
  return col_im ? c : s;                // [Wi,  Wr]
}

// FP16: the A_W table holds 2^t A_W with t = 14 - log2(P16), the operand
// scale of a unit-pupil field over a full window (bound P16 * max |E| of
// max |G|, brought to [2^14, 2^15)); the field then enters unscaled.  BF16x3:
// t = 0.
template <int Q, int PREC>
constexpr int table_exp() {
  return PREC == SHWFS_METHOD_DFT_FP16 ? (Q == 1 ? 10 : 9) : 0;
}

// The A_W operand (times 2^table_exp) as per-lane fragments, built once per
// device and kept in a small device table.
template <int N, int Q, int PREC>
__global__ void build_wfrag_kernel(Frag<PREC>* out) {
  using L = DftLayout<N, Q>;
  constexpr float scale = (float)(1 << table_exp<Q, PREC>());
  const int lane = threadIdx.x;
  for (int mt = 0; mt < L::MT; ++mt)
    for (int ks = 0; ks < L::KS; ++ks) {
      Frag<PREC> f;
#pragma unroll
      for (int e = 0; e < 8; ++e) f.set(e, aw_value<N, Q>(mt, ks, lane, e) * scale);
      out[(mt * L::KS + ks) * 64 + lane] = f;
    }
}

template <int N, int Q, int PREC>
static const Frag<PREC>* wfrag_table(hipStream_t st) {
  static std::mutex mu;
  static std::map<int, Frag<PREC>*> tables;  // device -> table
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = tables.find(dev);
  if (it != tables.end()) return it->second;
  using L = DftLayout<N, Q>;
  Frag<PREC>* t = nullptr;
  if (hipMalloc((void**)&t, sizeof(Frag<PREC>) * L::MT * L::KS * 64) != hipSuccess) return nullptr;
  hipLaunchKernelGGL((build_wfrag_kernel<N, Q, PREC>), dim3(1), dim3(64), 0, st, t);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFree(t);
    return nullptr;
  }
  tables[dev] = t;
  return t;
}

// Source of the A_W fragments: registers (small tables) or the device table
// (reads that stay in L1/L2).
template <int N, int Q, int PREC, bool REGS = DftLayout<N, Q>::kAwInRegs>
struct AwOperand {
  using L = DftLayout<N, Q>;
  Frag<PREC> f[L::MT][L::KS];
  __device__ __forceinline__ AwOperand(const Frag<PREC>* t, int lane) {
#pragma unroll
    for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
      for (int ks = 0; ks < L::KS; ++ks) f[mt][ks] = t[(mt * L::KS + ks) * 64 + lane];
  }
  __device__ __forceinline__ Frag<PREC> get(int mt, int ks) const { return f[mt][ks]; }
};
template <int N, int Q, int PREC>
struct AwOperand<N, Q, PREC, false> {
  using L = DftLayout<N, Q>;
  const Frag<PREC>* t;
  __device__ __forceinline__ AwOperand(const Frag<PREC>* tab, int lane) : t(tab + lane) {}
  __device__ __forceinline__ Frag<PREC> get(int mt, int ks) const {
    return t[(mt * L::KS + ks) * 64];
  }
};

// Stage B's operand from two accumulator tiles of stage A: X's four rows
// in slots 0-3, Y's in slots 4-7 (slot_coord / slot_im).
template <int PREC>
__device__ __forceinline__ Frag<PREC> acc_to_operand(const f32x4& X, const f32x4& Y) {
  Frag<PREC> f;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    f.set(v, X[v]);
    f.set(4 + v, Y[v]);
  }
  return f;
}

// Screen and pupil values of one window as the lanes of stage A's field
// operand need them (field_load), and the operand itself (field_eval).  The
// operand is A_E[y][slot] with y = 16 ma + (lane & 15); lane l evaluates the
// NP points x = field_px(l, j) of its row(s), which are the coordinates of its
// operand slots: Er in the re slots, Ei in the im slots (0 outside the pdiam
// window).
//  * Q == 1 (P16 = 16): x = 4 g + j (g = l >> 4, j < 4) of row y = l & 15;
//    slots 0-3 take Er, 4-7 Ei.
//  * Q == 2 (P16 = 32): x = 4 g + j and 16 + 4 g + j at the rows y = (l & 15)
//    and 16 + (l & 15); Er is the ks = 0 operand, Ei ks = 1.
// Screen reads: four consecutive x per lane, 16 rows per instruction (16-byte
// loads for full windows).
template <int Q>
struct FieldIn {
  static constexpr int NP = Q == 1 ? 4 : 8;  // points per lane and row
  float ph[Q][NP];
};

// x of point j of the lane (the coordinate of operand slot j).
template <int Q>
__device__ __forceinline__ int field_px(int lane, int j) {
  return slot_coord<Q>(lane >> 4, j);
}

typedef float f32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

// The offsets (the same for every window) at the lane's points, as turns:
// of = -offset / (2 pi).
template <int N, int Q, bool FULL>
__device__ __forceinline__ void field_offsets(const ImageParams& P, int lane,
                                              float (&of)[Q][FieldIn<Q>::NP]) {
  constexpr int NP = FieldIn<Q>::NP;
  const int pdiam = FULL ? DftLayout<N, Q>::P16 : P.pdiam;
#pragma unroll
  for (int ma = 0; ma < Q; ++ma) {
    const int y = 16 * ma + (lane & 15);
    if constexpr (FULL) {
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        const int x = field_px<Q>(lane, j);  // x .. x + 3 are points j .. j + 3
        const f32x4a4 o = *reinterpret_cast<const f32x4a4*>(P.offsets + (x + y * pdiam));
        of[ma][j] = o.x; of[ma][j + 1] = o.y; of[ma][j + 2] = o.z; of[ma][j + 3] = o.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int x = field_px<Q>(lane, j);
        const bool in = x < pdiam && y < pdiam;
        of[ma][j] = P.offsets[(unsigned)(in ? x + y * pdiam : 0)];
      }
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) of[ma][j] = -of[ma][j] * 0.15915494309189533577f;
  }
}

// Screen values at the lane's points of window win (flat screen index of its
// corner); points outside a window smaller than the operand tile read the
// corner (their amplitude is 0).  Row offsets y * npup through the 24-bit
// multiply (y < 32, npup < 2^15).
template <int N, int Q, bool FULL>
__device__ __forceinline__ void field_load(const ImageParams& P, unsigned win, int lane,
                                           FieldIn<Q>& f) {
  constexpr int NP = FieldIn<Q>::NP;
  const int pdiam = FULL ? DftLayout<N, Q>::P16 : P.pdiam;
  const unsigned npup = (unsigned)P.npup;
#pragma unroll
  for (int ma = 0; ma < Q; ++ma) {
    const int y = 16 * ma + (lane & 15);
    if constexpr (FULL) {
      // points j .. j + 3 are consecutive and inside the window: 16-byte
      // loads (4-byte aligned; gfx950 global loads need dword alignment only)
      const unsigned g = win + __umul24((unsigned)y, npup);
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        const unsigned gj = g + (unsigned)field_px<Q>(lane, j);
        const f32x4a4 v = *reinterpret_cast<const f32x4a4*>(
            reinterpret_cast<const char*>(P.phase) + (gj << 2));
        f.ph[ma][j] = v.x; f.ph[ma][j + 1] = v.y; f.ph[ma][j + 2] = v.z; f.ph[ma][j + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int x = field_px<Q>(lane, j);
        const bool in = x < pdiam && y < pdiam;
        const unsigned g = win + (in ? (unsigned)x + __umul24((unsigned)y, npup) : 0u);  // clamped
        f.ph[ma][j] =
            *reinterpret_cast<const float*>(reinterpret_cast<const char*>(P.phase) + (g << 2));
      }
    }
  }
}

// Amplitude at the lane's points of window win: the pupil; 1 for a window
// whose pupil is all ones (unit; nothing read); 0 outside a window smaller
// than the operand tile.  Read where it is used (few windows are not unit).
template <int N, int Q, bool FULL>
__device__ __forceinline__ void field_amplitude(const ImageParams& P, unsigned win, bool unit,
                                                int lane, float (&pu)[Q][FieldIn<Q>::NP]) {
  constexpr int NP = FieldIn<Q>::NP;
  const int pdiam = FULL ? DftLayout<N, Q>::P16 : P.pdiam;
  const unsigned npup = (unsigned)P.npup;
#pragma unroll
  for (int ma = 0; ma < Q; ++ma) {
    const int y = 16 * ma + (lane & 15);
    if constexpr (FULL) {
      const unsigned g = win + __umul24((unsigned)y, npup);
#pragma unroll
      for (int j = 0; j < NP; j += 4) {
        if (unit) {
          pu[ma][j] = pu[ma][j + 1] = pu[ma][j + 2] = pu[ma][j + 3] = 1.0f;
        } else {
          const unsigned gj = g + (unsigned)field_px<Q>(lane, j);
          const f32x4a4 v = *reinterpret_cast<const f32x4a4*>(
              reinterpret_cast<const char*>(P.pupil) + (gj << 2));
          pu[ma][j] = v.x; pu[ma][j + 1] = v.y; pu[ma][j + 2] = v.z; pu[ma][j + 3] = v.w;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int x = field_px<Q>(lane, j);
        const bool in = x < pdiam && y < pdiam;
        const unsigned g = win + (in ? (unsigned)x + __umul24((unsigned)y, npup) : 0u);
        float a = 1.0f;
        if (!unit)
          a = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(P.pupil) + (g << 2));
        pu[ma][j] = in ? a : 0.0f;
      }
    }
  }
}

// Field E = amplitude * exp(i (phase scale - offset)) at the lane's points,
// as stage A's operand.  The angle is taken in turns, reduced to [0, 1) by
// v_fract, and fed to the hardware sine and cosine.
// UNIT: a unit-pupil full window (amplitude 1 everywhere: pu is not read).
// of: the offsets in turns (field_offsets); pu: the amplitude
// (field_amplitude).
template <int N, int Q, bool UNIT>
__device__ __forceinline__ void field_eval(const ImageParams& P, int lane, const FieldIn<Q>& f,
                                           const float (&of)[Q][FieldIn<Q>::NP],
                                           const float (&pu)[Q][FieldIn<Q>::NP],
                                           float (&ev)[Q][Q][8]) {
  constexpr float kInv2Pi = 0.15915494309189533577f;
  constexpr int NP = FieldIn<Q>::NP;
  const float turn_scale = P.scale * kInv2Pi;
  float er[Q][NP], ei[Q][NP];
#pragma unroll
  for (int ma = 0; ma < Q; ++ma)
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const float r = __builtin_amdgcn_fractf(fmaf(f.ph[ma][j], turn_scale, of[ma][j]));
      er[ma][j] = __builtin_amdgcn_cosf(r);
      ei[ma][j] = __builtin_amdgcn_sinf(r);
      if constexpr (!UNIT) {
        er[ma][j] *= pu[ma][j];
        ei[ma][j] *= pu[ma][j];
      }
    }
  if constexpr (Q == 1) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ev[0][0][j] = er[0][j];
      ev[0][0][4 + j] = ei[0][j];
    }
  } else {
#pragma unroll
    for (int ma = 0; ma < Q; ++ma)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        ev[ma][0][e] = er[ma][e];
        ev[ma][1][e] = ei[ma][e];
      }
  }
}

// Flat screen index of the corner of sub-aperture s's window.
__device__ __forceinline__ unsigned window_corner(const ImageParams& P, int s) {
  return (unsigned)P.istart[s] + (unsigned)P.jstart[s] * (unsigned)P.npup;
}

// Whether sub-aperture s's pupil window is all ones.
__device__ __forceinline__ bool unit_window(const ImageParams& P, int s) {
  return P.pupil_unit != nullptr && P.pupil_unit[s] != 0;
}


// |F|^2 of one kx block (stage B's accumulator tiles fb of block ntk) into
// LDS: binned in registers (nr > 0) into the spot bin, or the whole image hr
// for the general binmap gather (nr == 0).  o: the binning origin.
template <int N, int Q, int NRC>
__device__ __forceinline__ void spot_bin_block(const f32x4 (&fb)[DftLayout<N, Q>::MT],
                                               const int ntk, const int lane, const int nr,
                                               const int o, const int npix, float* hr,
                                               float* bin) {
  using L = DftLayout<N, Q>;
  constexpr bool HALF = NRC >= 0 && (NRC & kHalfGrid) != 0;
  const int kx = 16 * ntk + (lane & 15);
  const int u = (kx - o) & (N - 1);
#pragma unroll
  for (int mt = 0; mt < L::MT / 2; ++mt) {
    float iv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float fr = fb[mt][v], fi = fb[mt + L::MT / 2][v];
      iv[v] = fmaf(fi, fi, fr * fr);  // one rounding order in every instantiation
    }
    const int ky0 = 16 * mt + 4 * (lane >> 4);  // this lane's 4 consecutive ky
    if (nr == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) hr[kx + (ky0 + v) * N] = iv[v];
    } else if (nr == 4) {
      // the 4 ky of the lane x lanes kx .. kx+3 form one binned pixel
      float b = fb[mt][0] * fb[mt][0];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (v > 0) b = fmaf(fb[mt][v], fb[mt][v], b);
        b = fmaf(fb[mt + L::MT / 2][v], fb[mt + L::MT / 2][v], b);
      }
      b += __int_as_float(dpp_row_perm(__float_as_int(b), 0));  // lane ^ 1
      b += __int_as_float(dpp_row_perm(__float_as_int(b), 1));  // lane ^ 2
      // both sums with every lane active (two DPP adds): kept from being moved
      // into the store's lane mask, where the second needs a separate DPP move
      asm volatile("" : "+v"(b));
      if constexpr (HALF) {
        // the binned cell (cx, cy) = (kx / 4, ky / 4) is pixel (cx ^ H, cy ^ H):
        // a per-lane base plus a constant offset per tile pair
        constexpr int G = N / 4, H = G / 2;
        const int a = (lane & 15) >> 2, c = lane >> 4;  // < 4
        int off;
        if constexpr (G >= 8)  // H >= 4: the XOR leaves a and c alone
          off = (a + c * G) + (((4 * ntk) ^ H) + ((4 * mt) ^ H) * G);
        else
          off = ((4 * ntk + a) ^ H) + ((4 * mt + c) ^ H) * G;
        if ((lane & 3) == 0) bin[off] = b;
      } else {
        const int px = u >> 2, py = ((ky0 - o) & (N - 1)) >> 2;
        if ((lane & 3) == 0 && px < npix && py < npix) bin[px + (int)__umul24(py, npix)] = b;
      }
    } else if (nr == 2) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        float b = iv[2 * g] + iv[2 * g + 1];
        b += __int_as_float(dpp_row_perm(__float_as_int(b), 0));  // lane ^ 1
        const int px = u >> 1, py = ((ky0 + 2 * g - o) & (N - 1)) >> 1;
        if ((lane & 1) == 0 && px < npix && py < npix) bin[px + (int)__umul24(py, npix)] = b;
      }
    } else {  // nr == 1
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int px = u, py = (ky0 + v - o) & (N - 1);
        if (px < npix && py < npix) bin[px + (int)__umul24(py, npix)] = iv[v];
      }
    }
  }
}

// Binned spot of sub-aperture s (< nvalid, uniform over the wave), before
// the flux normalisation, into bin (npix^2 floats of LDS, x fastest), in units
// of 1/inten_scale (returned through inten_scale).  hr: N^2 floats of LDS for
// the |F|^2 image (nr == 0 only); bm: the binmap table (LDS or global) used
// when nr == 0; nr: register rebin factor (dft_register_rebin), fixed at
// compile time by NRC >= 0 (NRC & 7; bit kHalfGrid: the half-grid 4x4 form,
// see dft_half_grid); FULL: pdiam == P16 (the window fills the operand
// tiles; no masking).
// PARTS > 1: the sub-aperture is shared by PARTS waves, wave `part` taking
// kx blocks part, part + PARTS, ... (N / 16 blocks; register binning only):
// each wave evaluates the whole field, computes stage A for its blocks' re
// and im tiles only, and bins its blocks into the shared spot `bin`.
template <int N, int Q, int PREC, bool FULL, int NRC = -1, int PARTS = 1>
__device__ __forceinline__ void dft_spot(const ImageParams& P, const int lane, float* hr,
                                         float* bin, const int* bm, const int nr_rt,
                                         const AwOperand<N, Q, PREC>& aw, const FieldIn<Q>& fin,
                                         const float (&of)[Q][FieldIn<Q>::NP], const unsigned win,
                                         const bool unit, float& inten_scale, const int part = 0) {
  static_assert(PARTS == 1 || (NRC >= 0 && (NRC & 7) > 0 && (N / 16) % PARTS == 0),
                "dft_spot: a shared sub-aperture needs register binning and whole kx blocks");
  const int nr = NRC >= 0 ? (NRC & 7) : nr_rt;
  using L = DftLayout<N, Q>;
  using F = Frag<PREC>;
  const int pdiam = FULL ? L::P16 : P.pdiam;

  // ---- field: stage A's operand, computed in registers.  FP16: G must land
  // in [2^14, 2^15) of its bound pdiam * max |E|: the table carries 2^t
  // (table_exp), which is the whole scale for a unit-pupil window (max |E| =
  // 1); otherwise max |E| is measured over the wave and the field takes the
  // rest, 2^(e_total - t) (scalar exponent work: the bound is wave-uniform).
  constexpr int t = table_exp<Q, PREC>();
  int e_total = t;  // G carries 2^e_total
  float ev[Q][Q][8];
  if (FULL && unit) {
    field_eval<N, Q, FULL>(P, lane, fin, of, of, ev);
  } else {
    float pu[Q][FieldIn<Q>::NP];
    field_amplitude<N, Q, FULL>(P, win, unit, lane, pu);
    field_eval<N, Q, false>(P, lane, fin, of, pu, ev);
    if constexpr (PREC == SHWFS_METHOD_DFT_FP16) {
      if (!unit) {
        float m = 0.0f;
#pragma unroll
        for (int ma = 0; ma < Q; ++ma)
#pragma unroll
          for (int ks = 0; ks < Q; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(ev[ma][ks][e]));
        const float amax = wave_max(m) * (float)pdiam;
        const int ex = ((__float_as_int(amax) >> 23) & 0xff) - 127;  // ilogb for normal amax
        // zero or subnormal bound: the table's scale alone
        e_total = amax >= 1.17549435e-38f ? 14 - ex : t;
        const float fscale = __int_as_float((127 + e_total - t) << 23);  // |e_total - t| < 127
#pragma unroll
        for (int ma = 0; ma < Q; ++ma)
#pragma unroll
          for (int ks = 0; ks < Q; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) ev[ma][ks][e] *= fscale;
      }
    }
  }
  F ae[Q][Q];
#pragma unroll
  for (int ma = 0; ma < Q; ++ma)
#pragma unroll
    for (int ks = 0; ks < Q; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) ae[ma][ks].set(e, ev[ma][ks][e]);

  // ---- stage A: G (rows y, cols (re/im, kx)) = E A_W^T
  constexpr int NTW = L::NTB / PARTS;  // kx blocks of this wave
  f32x4 ga[Q][L::MT];
  if constexpr (PARTS == 1) {
#pragma unroll
    for (int ma = 0; ma < Q; ++ma)
#pragma unroll
      for (int nt = 0; nt < L::MT; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS; ++ks) acc = mma(ae[ma][ks], aw.get(nt, ks), acc);
        ga[ma][nt] = acc;
      }
  } else {
    // the re and im tiles of blocks part + PARTS j, stored at j and H + j
#pragma unroll
    for (int ma = 0; ma < Q; ++ma)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < NTW; ++j) {
          const int nt = h * (L::MT / 2) + part + PARTS * j;
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int ks = 0; ks < L::KS; ++ks) acc = mma(ae[ma][ks], aw.get(nt, ks), acc);
          ga[ma][h * (L::MT / 2) + j] = acc;
        }
  }
  // F carries 2^(e_total + t) (stage B's table scale too): |F|^2 is scaled
  // back by 2^(-2 (e_total + t)) (may be subnormal)
  inten_scale = ldexpf(1.0f, -2 * (e_total + t));

  // ---- stage B per kx tile: F = A_W [Gr ; Gi]; K step ks takes the Gr and
  // Gi tiles of the kx block (Q == 1), or the two row blocks of Gr (ks = 0)
  // or Gi (ks = 1) (Q == 2), in slot order (slot_coord).
  // |F|^2 goes to LDS: binned in registers (nr > 0) or as the whole image for
  // the general binmap gather (nr == 0).  The field planes are dead by now.
  const int o = P.bin_origin, npix = P.npix;
  wave_lds_sync();
#pragma unroll
  for (int nt = 0; nt < NTW; ++nt) {
    const int ntk = PARTS == 1 ? nt : part + PARTS * nt;  // this tile's kx block
    F bg[L::KS];
#pragma unroll
    for (int ks = 0; ks < L::KS; ++ks) {
      constexpr int H = L::MT / 2;  // im column tiles start at H
      if constexpr (Q == 1)
        bg[ks] = acc_to_operand<PREC>(ga[0][nt], ga[0][H + nt]);
      else
        bg[ks] = acc_to_operand<PREC>(ga[0][ks * H + nt], ga[Q - 1][ks * H + nt]);
    }
    f32x4 fb[L::MT];
#pragma unroll
    for (int mt = 0; mt < L::MT; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < L::KS; ++ks) acc = mma(aw.get(mt, ks), bg[ks], acc);
      fb[mt] = acc;
    }
    spot_bin_block<N, Q, NRC>(fb, ntk, lane, nr, o, npix, hr, bin);
  }
  if (nr == 0) {
    wave_lds_sync();
    (void)bin_gather<64, N>(P, true, lane, hr, bin, bm);
  }
}

// ---- two sub-apertures of an 8-point window per stage A (N = 16) ----------
// Sub-apertures a and b with pdiam <= 8 on the 16-point split-BF16x3 stage
// with register binning share one pass of stage A: the lanes of operand rows
// y = lane & 15 below 8 carry a's window rows, rows 8-15 b's (rows y >= pdiam
// are zero in the one-sub-aperture layout, as points x >= pdiam are).  Stage
// A's accumulator tiles then hold a's rows in lanes 0-31 and b's in lanes
// 32-63 (lane j + 16 g holds rows 4 g .. 4 g + 3), and one v_permlane32_swap
// per register of stage B's operand gives a's operand (lanes 32-63 zero) and
// b's (moved to lanes 0-31, lanes 32-63 zero): each that of dft_spot on the
// sub-aperture alone, so both spots are the same bit for bit, for half the
// field evaluations, stage A products and operand splits.  FP16: each half
// takes its window's field scale from the maximum over its own 32 lanes,
// which is dft_spot's wave maximum over that window.
template <int N, int Q, int PREC, bool FULL, int NRC>
__host__ __device__ constexpr bool dft_pair_stage() {
  return N == 16 && Q == 1 && !FULL && NRC >= 0 && (NRC & 7) > 0;
}

// Offsets (in turns) at the lane's points of its pair row (row = lane & 7).
__device__ __forceinline__ void field_offsets_pair(const ImageParams& P, int lane,
                                                   float (&of)[1][FieldIn<1>::NP]) {
  const int pdiam = P.pdiam, row = lane & 7;
#pragma unroll
  for (int j = 0; j < FieldIn<1>::NP; ++j) {
    const int x = field_px<1>(lane, j);
    const bool in = x < pdiam && row < pdiam;
    of[0][j] = P.offsets[(unsigned)(in ? x + row * pdiam : 0)];
  }
#pragma unroll
  for (int j = 0; j < FieldIn<1>::NP; ++j) of[0][j] = -of[0][j] * 0.15915494309189533577f;
}

// Screen values at the lane's points: window a for rows y < 8, b above.
__device__ __forceinline__ void field_load_pair(const ImageParams& P, unsigned win_a,
                                                unsigned win_b, int lane, FieldIn<1>& f) {
  const int pdiam = P.pdiam, row = lane & 7;
  const unsigned npup = (unsigned)P.npup;
  const unsigned win = (lane & 8) ? win_b : win_a;
#pragma unroll
  for (int j = 0; j < FieldIn<1>::NP; ++j) {
    const int x = field_px<1>(lane, j);
    const bool in = x < pdiam && row < pdiam;
    const unsigned g = win + (in ? (unsigned)x + __umul24((unsigned)row, npup) : 0u);  // clamped
    f.ph[0][j] =
        *reinterpret_cast<const float*>(reinterpret_cast<const char*>(P.phase) + (g << 2));
  }
}

// Amplitude at the lane's points of its pair row (field_amplitude's rule).
__device__ __forceinline__ void field_amplitude_pair(const ImageParams& P, unsigned win_a,
                                                     unsigned win_b, bool unit_a, bool unit_b,
                                                     int lane, float (&pu)[1][FieldIn<1>::NP]) {
  const int pdiam = P.pdiam, row = lane & 7;
  const unsigned npup = (unsigned)P.npup;
  const bool second = (lane & 8) != 0;
  const unsigned win = second ? win_b : win_a;
  const bool unit = second ? unit_b : unit_a;
#pragma unroll
  for (int j = 0; j < FieldIn<1>::NP; ++j) {
    const int x = field_px<1>(lane, j);
    const bool in = x < pdiam && row < pdiam;
    const unsigned g = win + (in ? (unsigned)x + __umul24((unsigned)row, npup) : 0u);
    float a = 1.0f;
    if (!unit)
      a = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(P.pupil) + (g << 2));
    pu[0][j] = in ? a : 0.0f;
  }
}

// Lanes 0-31 of f (lanes 32-63 zero) and lanes 32-63 of f moved to lanes
// 0-31 (lanes 32-63 zero).  Every lane active.
__device__ __forceinline__ void split_lane_halves(const Frag<SHWFS_METHOD_DFT_BF16X3>& f,
                                                  Frag<SHWFS_METHOD_DFT_BF16X3>& lower,
                                                  Frag<SHWFS_METHOD_DFT_BF16X3>& upper) {
  typedef unsigned u32x4v __attribute__((ext_vector_type(4)));
  const u32x4v h = __builtin_bit_cast(u32x4v, f.hi), l = __builtin_bit_cast(u32x4v, f.lo);
  u32x4v h0, h1, l0, l1;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    // swaps the old value's lanes 32-63 with the source's lanes 0-31
    const auto rh = __builtin_amdgcn_permlane32_swap(h[d], 0u, false, false);
    const auto rl = __builtin_amdgcn_permlane32_swap(l[d], 0u, false, false);
    h0[d] = rh[0];
    h1[d] = rh[1];
    l0[d] = rl[0];
    l1[d] = rl[1];
  }
  lower.hi = __builtin_bit_cast(bf16x8, h0);
  lower.lo = __builtin_bit_cast(bf16x8, l0);
  upper.hi = __builtin_bit_cast(bf16x8, h1);
  upper.lo = __builtin_bit_cast(bf16x8, l1);
}

__device__ __forceinline__ void split_lane_halves(const Frag<SHWFS_METHOD_DFT_FP16>& f,
                                                  Frag<SHWFS_METHOD_DFT_FP16>& lower,
                                                  Frag<SHWFS_METHOD_DFT_FP16>& upper) {
  typedef unsigned u32x4v __attribute__((ext_vector_type(4)));
  const u32x4v v = __builtin_bit_cast(u32x4v, f.v);
  u32x4v v0, v1;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const auto r = __builtin_amdgcn_permlane32_swap(v[d], 0u, false, false);
    v0[d] = r[0];
    v1[d] = r[1];
  }
  lower.v = __builtin_bit_cast(f16x8, v0);
  upper.v = __builtin_bit_cast(f16x8, v1);
}

// Maximum over the 32 lanes with the same lane bit 3 (one pair half).
__device__ __forceinline__ float pair_half_max(float m) {
  m = fmaxf(m, __shfl_xor(m, 1));
  m = fmaxf(m, __shfl_xor(m, 2));
  m = fmaxf(m, __shfl_xor(m, 4));
  m = fmaxf(m, __shfl_xor(m, 16));
  return fmaxf(m, __shfl_xor(m, 32));
}

// The binned spots of sub-apertures a and b (see above) into bin_a, bin_b,
// in units of 1 / scale_a and 1 / scale_b.
template <int N, int PREC, int NRC>
__device__ __forceinline__ void dft_spot_pair(const ImageParams& P, const int lane, float* bin_a,
                                              float* bin_b, const AwOperand<N, 1, PREC>& aw,
                                              const FieldIn<1>& fin,
                                              const float (&of)[1][FieldIn<1>::NP],
                                              const unsigned win_a, const unsigned win_b,
                                              const bool unit_a, const bool unit_b,
                                              float& scale_a, float& scale_b) {
  using L = DftLayout<N, 1>;
  using F = Frag<PREC>;
  static_assert(L::MT == 2 && L::KS == 1 && L::NTB == 1, "dft_spot_pair: the 16-point stage");
  constexpr int t = table_exp<1, PREC>();
  float ev[1][1][8];
  {
    float pu[1][FieldIn<1>::NP];
    field_amplitude_pair(P, win_a, win_b, unit_a, unit_b, lane, pu);
    field_eval<N, 1, false>(P, lane, fin, of, pu, ev);
  }
  // FP16 field scale of each half (dft_spot's rule on its window alone)
  int e_a = t, e_b = t;
  if constexpr (PREC == SHWFS_METHOD_DFT_FP16) {
    if (!unit_a || !unit_b) {  // uniform
      float m = 0.0f;
#pragma unroll
      for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(ev[0][0][e]));
      const float amax = pair_half_max(m) * (float)P.pdiam;
      const int ex = ((__float_as_int(amax) >> 23) & 0xff) - 127;  // ilogb for normal amax
      const bool unit = (lane & 8) ? unit_b : unit_a;
      const int e_lane = unit ? t : (amax >= 1.17549435e-38f ? 14 - ex : t);
      const float fscale = __int_as_float((127 + e_lane - t) << 23);  // |e_lane - t| < 127
#pragma unroll
      for (int e = 0; e < 8; ++e) ev[0][0][e] *= fscale;
      e_a = __builtin_amdgcn_readlane(e_lane, 0);  // lane 0: half a
      e_b = __builtin_amdgcn_readlane(e_lane, 8);  // lane 8: half b
    }
  }
  F ae;
#pragma unroll
  for (int e = 0; e < 8; ++e) ae.set(e, ev[0][0][e]);
  f32x4 ga[L::MT];
#pragma unroll
  for (int nt = 0; nt < L::MT; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    ga[nt] = mma(ae, aw.get(nt, 0), acc);
  }
  scale_a = ldexpf(1.0f, -2 * (e_a + t));
  scale_b = ldexpf(1.0f, -2 * (e_b + t));
  const int o = P.bin_origin, npix = P.npix;
  const int nr = NRC & 7;
  wave_lds_sync();
  F both = acc_to_operand<PREC>(ga[0], ga[1]);
  F op[2];
  split_lane_halves(both, op[0], op[1]);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    f32x4 fb[L::MT];
#pragma unroll
    for (int mt = 0; mt < L::MT; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      fb[mt] = mma(aw.get(mt, 0), op[k], acc);
    }
    spot_bin_block<N, 1, NRC>(fb, 0, lane, nr, o, npix, nullptr, k == 0 ? bin_a : bin_b);
  }
}

// Spot, detector pixels and centroid of sub-aperture s by one wave.  smem: the
// wave's LDS slot (wave_floats): the spot sits at its start (nr > 0) or end
// (nr == 0), the moment totals use the 192 doubles after the spot (nr > 0)
// or at the start, over the dead |F|^2 image (nr == 0).  KEEP: no centroider;
// the spot-cube values stay in the spot.
template <int N, int Q, int PREC, bool NOISE, bool FULL, int NRC = -1, bool KEEP = false>
__device__ __forceinline__ void dft_subap(const ImageParams& P, const int s, const int lane,
                                          float* smem, const int* bm, const int nr_rt,
                                          const AwOperand<N, Q, PREC>& aw,
                                          unsigned call_add = 0u) {
  using L = DftLayout<N, Q>;
  const int nr = NRC >= 0 ? (NRC & 7) : nr_rt;
  SubapScalars pre;
  pre.load(P, s);
  const int npix = P.npix;
  const int spot_f = (npix * npix + 3) & ~3;
  float* bin = nr > 0 ? smem : smem + (L::wave_floats(npix, true) - spot_f);
  const int pdiam = FULL ? L::P16 : P.pdiam;
  const bool unit = unit_window(P, s);
  float of[Q][FieldIn<Q>::NP];
  field_offsets<N, Q, FULL>(P, lane, of);
  const unsigned win = window_corner(P, s);
  FieldIn<Q> fin;
  field_load<N, Q, FULL>(P, win, lane, fin);
  float inten_scale;
  dft_spot<N, Q, PREC, FULL, NRC>(P, lane, smem, bin, bm, nr_rt, aw, fin, of, win, unit,
                                  inten_scale);
  wave_lds_sync();
  const float I = spot_sum<64>(bin, npix * npix, true, lane);
  void* red = nr > 0 ? smem + spot_f : smem;
  finish_spot<64, NOISE, KEEP>(P, s, true, lane, bin, I, inten_scale, call_add, red, &pre);
  wave_lds_sync();
}

// Sub-aperture s shared by the PARTS waves of a workgroup (part = wave
// index): each bins its kx blocks into the spot at the start of `shared`
// (the first wave's LDS slot), then the first wave runs the tail.  All PARTS
// waves must call it (workgroup barrier).
template <int N, int Q, int PREC, bool NOISE, bool FULL, int NRC, int PARTS>
__device__ __forceinline__ void dft_subap_parts(const ImageParams& P, const int s, const int lane,
                                                const int part, float* shared,
                                                const AwOperand<N, Q, PREC>& aw,
                                                unsigned call_add = 0u) {
  static_assert(NRC >= 0 && (NRC & 7) > 0, "dft_subap_parts: register binning only");
  SubapScalars pre;
  pre.load(P, s);
  const int npix = P.npix;
  const int spot_f = (npix * npix + 3) & ~3;
  const bool unit = unit_window(P, s);
  float of[Q][FieldIn<Q>::NP];
  field_offsets<N, Q, FULL>(P, lane, of);
  const unsigned win = window_corner(P, s);
  FieldIn<Q> fin;
  field_load<N, Q, FULL>(P, win, lane, fin);
  float inten_scale;
  dft_spot<N, Q, PREC, FULL, NRC, PARTS>(P, lane, shared, shared, nullptr, NRC & 7, aw, fin, of,
                                         win, unit, inten_scale, part);
  __syncthreads();  // every kx block binned
  if (part == 0) {
    const float I = spot_sum<64>(shared, npix * npix, true, lane);
    finish_spot<64, NOISE>(P, s, true, lane, shared, I, inten_scale, call_add, shared + spot_f,
                           &pre);
  }
}

// LDS floats per wave of dft_subaps: the |F|^2 image (nr == 0), then SPW
// spots (16-byte aligned), then SPW spot scales, then 192 doubles for the
// float64 moment sums of the shared tail.
template <int N, int Q>
__host__ __device__ inline int subaps_floats(int npix, bool gather, int spw) {
  const int spot = (npix * npix + 3) & ~3;
  return (gather ? N * N : 0) + spw * spot + ((spw + 3) & ~3) + 384;
}

// Sub-apertures s0 .. s0 + SPW - 1 (those < nvalid) by one wave: their spots
// one after the other, then the tails of all of them at once with 16 lanes per
// sub-aperture (SPW == 4), so the per-sub-aperture sums, divisions and stores
// are shared.  smem: the wave's LDS slot (subaps_floats).
template <int N, int Q, int PREC, bool NOISE, bool FULL, int NRC, int SPW>
__device__ __forceinline__ void dft_subaps(const ImageParams& P, const int s0, const int lane,
                                           float* smem, const int* bm, const int nr_rt,
                                           const AwOperand<N, Q, PREC>& aw,
                                           unsigned call_add = 0u) {
  static_assert(SPW == 4, "dft_subaps: four sub-apertures per wave");
  const int nr = NRC >= 0 ? (NRC & 7) : nr_rt;
  const int npix = P.npix, npix2 = npix * npix;
  const int spot_f = (npix2 + 3) & ~3;
  float* hr = smem;
  float* spots = smem + (nr == 0 ? N * N : 0);
  float* scales = spots + SPW * spot_f;
  // this lane's sub-aperture in the tail: q = lane / 16
  const int q = lane >> 4, r = lane & 15;
  const int s = s0 + q;
  const bool valid = s < P.nvalid;
  SubapScalars pre;
  pre.load(P, valid ? s : s0);  // per-lane loads, issued ahead of the spots
  // windows and unit flags of the SPW sub-apertures (scalar loads, together);
  // each window's screen values are loaded while the previous spot is computed
  using L = DftLayout<N, Q>;
  const int pdiam = FULL ? L::P16 : P.pdiam;
  const int last = P.nvalid - 1;
  unsigned win[SPW];
  bool unit[SPW];
#pragma unroll
  for (int k = 0; k < SPW; ++k) {
    const int sk = min(s0 + k, last);
    win[k] = window_corner(P, sk);
    unit[k] = unit_window(P, sk);
  }
  float of[Q][FieldIn<Q>::NP];
  FieldIn<Q> fin[2];
  if (dft_pair_stage<N, Q, PREC, FULL, NRC>() && pdiam <= 8) {
    // two sub-apertures per stage A (dft_spot_pair); each pair's screen values
    // are loaded while the previous pair is computed
    if constexpr (dft_pair_stage<N, Q, PREC, FULL, NRC>()) {
      field_offsets_pair(P, lane, of);
      field_load_pair(P, win[0], win[1], lane, fin[0]);
#pragma unroll
      for (int k = 0; k < SPW; k += 2) {
        if (s0 + k > last) break;  // uniform
        if (k + 2 < SPW) field_load_pair(P, win[k + 2], win[k + 3], lane, fin[((k >> 1) + 1) & 1]);
        float scale_a, scale_b;
        dft_spot_pair<N, PREC, NRC>(P, lane, spots + k * spot_f, spots + (k + 1) * spot_f, aw,
                                    fin[(k >> 1) & 1], of, win[k], win[k + 1], unit[k],
                                    unit[k + 1], scale_a, scale_b);
        if (lane == 0) {
          scales[k] = scale_a;
          scales[k + 1] = scale_b;
        }
      }
    }
  } else {
    field_offsets<N, Q, FULL>(P, lane, of);
    field_load<N, Q, FULL>(P, win[0], lane, fin[0]);
#pragma unroll
    for (int k = 0; k < SPW; ++k) {
      if (s0 + k > last) break;  // uniform
      if (k + 1 < SPW) field_load<N, Q, FULL>(P, win[k + 1], lane, fin[(k + 1) & 1]);
      float inten_scale;
      dft_spot<N, Q, PREC, FULL, NRC>(P, lane, hr, spots + k * spot_f, bm, nr_rt, aw, fin[k & 1],
                                      of, win[k], unit[k], inten_scale);
      if (lane == 0) scales[k] = inten_scale;
    }
  }
  wave_lds_sync();
  float* sp = spots + q * spot_f;
  const float I = spot_sum<16>(sp, npix2, valid, r);
  double* red = reinterpret_cast<double*>(scales + ((SPW + 3) & ~3));  // 8-byte aligned
  finish_spot<16, NOISE>(P, s, valid, r, sp, I, scales[q], call_add, red, &pre);
  wave_lds_sync();
}

// Register budgets (waves per SIMD), the largest at which the code still
// fits without spilling: 7 (72 VGPRs) for the noise-free FP16 full-window
// 4x4-binning variant, 5 (96 VGPRs) for the other noise-free variants, so a
// frame of ~5000 sub-apertures is resident at once; 4 for the N = 64 tiles;
// 2 (240 VGPRs) with the noise samplers, which spill at 3 or more.
template <int PREC, bool NOISE, int N, bool FULL, int NR>
constexpr int dft_waves_per_eu() {
  if (NOISE) return 2;
  if (N == 64) return 4;
  return (PREC == SHWFS_METHOD_DFT_FP16 && FULL && (NR & 7) == 4) ? 7 : 5;
}

// Waves per workgroup of the spot kernels.  The waves are independent (one
// sub-aperture and one LDS slot each, no workgroup barrier after the start):
// grouping them lifts the per-CU workgroup limit off the number of resident
// waves.
constexpr int kDftWaves = 4;

// The 64-point transform with register binning shares each sub-aperture
// among the waves of a workgroup (dft_subap_parts): its sub-apertures are few
// and long, one wave each would leave most of the chip idle.
__host__ __device__ constexpr bool dft_split_subap(int n, int nr_code) {
  return n == 64 && (nr_code & 7) > 0;
}

// One frame: wave w of workgroup g handles sub-aperture kDftWaves g' + w with
// g' = xcd_remap(g) (workgroups of one XCD hold contiguous sub-apertures).
// One sub-aperture per wave, with no loop: values derived from the lane index
// would otherwise be hoisted out of the loop and held in registers across the
// whole body.  Spare waves of the last workgroup recompute the last
// sub-aperture (same values to the same addresses) rather than exit early: an
// exit test would hold back the parameter loads behind a load of nvalid.
template <int N, int Q, int PREC, bool NOISE, bool FULL, int NR>
__global__ __launch_bounds__(64 * kDftWaves)
__attribute__((amdgpu_waves_per_eu(dft_waves_per_eu<PREC, NOISE, N, FULL, NR>())))
void sh_image_dft_kernel(ImageParams P, const Frag<PREC>* wtab, BgParams bg) {
  constexpr int nr = NR & 7;
  using L = DftLayout<N, Q>;
  // workgroups past the spot part draw the fused background (noisy frames)
  if constexpr (NOISE) {
    if ((int)blockIdx.x >= bg.spot_blocks) {
      draw_background<64 * kDftWaves>(P, bg, (int)blockIdx.x - bg.spot_blocks);
      return;
    }
  }
  extern __shared__ float smem[];
  const int lane = (int)threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int WF = L::wave_floats(P.npix, nr == 0);
  int* s_binmap = (int*)(smem + kDftWaves * WF);
  if (nr == 0 && !P.bin_block) {
    const int nbm = P.nrebin * P.nrebin * P.npix * P.npix;
    for (int i = threadIdx.x; i < nbm; i += 64 * kDftWaves) s_binmap[i] = P.binmap[i];
    __syncthreads();
  }
  const AwOperand<N, Q, PREC> aw(wtab, lane);
  if constexpr (dft_split_subap(N, NR)) {
    // the 64-point transform: one sub-aperture per workgroup, a kx quarter per wave
    const int s = min(xcd_remap(blockIdx.x, bg.spot_blocks), P.nvalid - 1);
    dft_subap_parts<N, Q, PREC, NOISE, FULL, NR, kDftWaves>(P, s, lane, wave, smem, aw);
  } else {
    const int s = min(xcd_remap(blockIdx.x, bg.spot_blocks) * kDftWaves + wave, P.nvalid - 1);
    dft_subap<N, Q, PREC, NOISE, FULL, NR>(P, s, lane, smem + wave * WF, s_binmap, nr, aw);
  }
}

// Sub-apertures per wave of the same-kind batch kernel (its tiles).
constexpr int kDftTaskSubaps = 4;

// Same-kind batches: every task runs this (N, Q, PREC, FULL, NR) variant.
// Wave w of workgroup b handles global tile kDftWaves b' + w of the flattened
// task range (b' = b, or xcd_remap(b) with remap; task i owns tiles
// [first_tile, first_tile + ntiles), kDftTaskSubaps sub-apertures each);
// spare waves of the last workgroup recompute the last tile.  slot: LDS
// floats per wave (the largest slot of the tasks).  At most kTaskArgs tasks:
// their parameters come with the kernel arguments.
template <int N, int Q, int PREC, bool NOISE, bool FULL, int NR>
__global__ __launch_bounds__(64 * kDftWaves)
__attribute__((amdgpu_waves_per_eu(dft_waves_per_eu<PREC, NOISE, N, FULL, NR>())))
void sh_dft_tasks_kernel(const BatchTask* __restrict__ tasks, const TaskStarts starts,
                         const TaskArgs args, int ntasks, int total, int slot,
                         const Frag<PREC>* wtab, const uint32_t* __restrict__ frame_base,
                         uint32_t frame_off, int remap) {
  extern __shared__ float smem[];
  const int lane = (int)threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const uint32_t call = frame_base[0] + frame_off;
  const AwOperand<N, Q, PREC> aw(wtab, lane);
  const int G = (int)gridDim.x;
  const int b = (int)blockIdx.x;
  const int bb = b + (remap != 0) * (xcd_remap(b, G) - b);  // no branch ahead of the loads
  const int t = min(bb * kDftWaves + wave, total - 1);
  const int i = find_task_args(starts, ntasks, t);
  const ImageParams& P = args.P[i];
  dft_subaps<N, Q, PREC, NOISE, FULL, NR, kDftTaskSubaps>(
      P, (t - starts.first[i]) * kDftTaskSubaps, lane, smem + wave * slot, P.binmap, NR & 7, aw,
      call);
}

template <int N, int Q, int PREC>
static int32_t launch_dft_tasks_n(bool full, int nr, bool noisy, const BatchTask* tasks,
                                  const TaskStarts& starts, const TaskArgs& args,
                                  int ntasks, int total, int grid, size_t lds,
                                  const uint32_t* frame_base, uint32_t frame_off, int remap,
                                  hipStream_t st) {
  const Frag<PREC>* wtab = wfrag_table<N, Q, PREC>(st);
  if (!wtab) return fail(SHWFS_ERR_HIP, "DFT batch: building the DFT operand table failed");
  decltype(&sh_dft_tasks_kernel<N, Q, PREC, false, true, 4>) k = nullptr;
#define SHWFS_PICK(NZ, FL)                                                             \
  switch (nr) {                                                                        \
    case 4 | kHalfGrid: k = sh_dft_tasks_kernel<N, Q, PREC, NZ, FL, 4 | kHalfGrid>; break; \
    case 4: k = sh_dft_tasks_kernel<N, Q, PREC, NZ, FL, 4>; break;                     \
    case 2: k = sh_dft_tasks_kernel<N, Q, PREC, NZ, FL, 2>; break;                     \
    case 1: k = sh_dft_tasks_kernel<N, Q, PREC, NZ, FL, 1>; break;                     \
    default: k = sh_dft_tasks_kernel<N, Q, PREC, NZ, FL, 0>; break;                    \
  }
  if (noisy) {
    if (full) { SHWFS_PICK(true, true) } else { SHWFS_PICK(true, false) }
  } else {
    if (full) { SHWFS_PICK(false, true) } else { SHWFS_PICK(false, false) }
  }
#undef SHWFS_PICK
  (void)grid;  // one wave per sub-aperture
  if (total <= 0) return SHWFS_OK;
  if (lds * kDftWaves > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT batch: LDS footprint exceeds the per-workgroup limit");
  if (ntasks > kTaskArgs) return fail(SHWFS_ERR_UNSUPPORTED, "DFT batch: more than 16 tasks");
  const int slot = (int)(lds / sizeof(float));
  hipLaunchKernelGGL(k, dim3((total + kDftWaves - 1) / kDftWaves), dim3(64 * kDftWaves),
                     lds * kDftWaves, st, tasks, starts, args, ntasks, total, slot, wtab,
                     frame_base, frame_off, remap);
  SHWFS_LAUNCHED("sh_dft_tasks_kernel");
  return SHWFS_OK;
}

// Register binning applies when the binmap is the NGS block form with a block
// origin aligned on the rebin factor (nrebin in {1, 2, 4}); otherwise 0, the
// general gather over |F|^2 in LDS.
__host__ __device__ inline int dft_register_rebin(const ImageParams& p) {
  const bool blk = p.bin_block && p.nrebin > 0 && (p.bin_origin % p.nrebin) == 0;
  return (blk && (p.nrebin == 1 || p.nrebin == 2 || p.nrebin == 4)) ? p.nrebin : 0;
}

// 4x4 register binning of a spot that spans the whole transform (npix =
// nfft / 4) with the fftshift of a natural guide star (block origin nfft / 2):
// binned cell (c_x, c_y) = (k_x / 4, k_y / 4) is pixel (c_x, c_y) XOR nfft / 8,
// so each lane's pixel per tile is fixed at compile time.
inline bool dft_half_grid(const ImageParams& p, int nfft) {
  return dft_register_rebin(p) == 4 && p.npix * 4 == nfft && p.bin_origin * 2 == nfft;
}

template <int N, int Q, int PREC, int NR>
static int32_t launch_dft_t(const ImageParams& p, hipStream_t st, const void* wtab_in,
                            BgParams bg) {
  using L = DftLayout<N, Q>;
  const int nbm = (NR == 0 && !p.bin_block) ? p.nrebin * p.nrebin * p.npix * p.npix : 0;
  const size_t lds = sizeof(float) * ((size_t)kDftWaves * L::wave_floats(p.npix, NR == 0) + nbm);
  if (lds > max_lds_per_block())
    return fail(SHWFS_ERR_UNSUPPORTED, "DFT spot kernel: LDS footprint exceeds the per-workgroup limit");
  const bool noisy = p.noise > -1.0f;
  const bool full = p.pdiam == L::P16;
  auto kern = noisy ? (full ? sh_image_dft_kernel<N, Q, PREC, true, true, NR>
                            : sh_image_dft_kernel<N, Q, PREC, true, false, NR>)
                    : (full ? sh_image_dft_kernel<N, Q, PREC, false, true, NR>
                            : sh_image_dft_kernel<N, Q, PREC, false, false, NR>);
  // the library's cached table, or one the caller has built (wtab_in)
  const Frag<PREC>* wtab = wtab_in ? static_cast<const Frag<PREC>*>(wtab_in)
                                   : wfrag_table<N, Q, PREC>(st);
  if (!wtab) return fail(SHWFS_ERR_HIP, "DFT spot kernel: building the DFT operand table failed");
  if (p.nvalid <= 0) return SHWFS_OK;
  // one wave per sub-aperture, or one workgroup per 64-point sub-aperture;
  // then the workgroups of the fused background
  bg.spot_blocks = dft_split_subap(N, NR) ? p.nvalid : (p.nvalid + kDftWaves - 1) / kDftWaves;
  const int grid = bg.spot_blocks + (noisy ? (bg.n + 64 * kDftWaves - 1) / (64 * kDftWaves) : 0);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kDftWaves), lds, st, p, wtab, bg);
  SHWFS_LAUNCHED("sh_image_dft_kernel");
  return SHWFS_OK;
}

template <int N, int Q, int PREC>
static int32_t launch_dft_n(const ImageParams& p, hipStream_t st, const void* wtab,
                            const BgParams& bg) {
  switch (dft_register_rebin(p)) {
    case 4:
      return dft_half_grid(p, N) ? launch_dft_t<N, Q, PREC, 4 | kHalfGrid>(p, st, wtab, bg)
                                 : launch_dft_t<N, Q, PREC, 4>(p, st, wtab, bg);
    case 2: return launch_dft_t<N, Q, PREC, 2>(p, st, wtab, bg);
    case 1: return launch_dft_t<N, Q, PREC, 1>(p, st, wtab, bg);
    default: return launch_dft_t<N, Q, PREC, 0>(p, st, wtab, bg);
  }
}

// Entry point per operand precision and padded phase side 16*Q; each
// (PREC, Q) pair is instantiated in a translation unit of its own.  wtab:
// NULL (the library's table of the shape, built on first use) or a table
// the caller has built with build_dft_table_pq.
template <int PREC, int Q>
int32_t launch_dft_pq(const ImageParams& p, int nfft, hipStream_t st, const void* wtab,
                      const BgParams& bg) {
  if constexpr (Q == 1) {
    switch (nfft) {
      case 16: return launch_dft_n<16, 1, PREC>(p, st, wtab, bg);
      case 32: return launch_dft_n<32, 1, PREC>(p, st, wtab, bg);
      case 64: return launch_dft_n<64, 1, PREC>(p, st, wtab, bg);
      default: break;
    }
  } else {
    switch (nfft) {
      case 32: return launch_dft_n<32, 2, PREC>(p, st, wtab, bg);
      case 64: return launch_dft_n<64, 2, PREC>(p, st, wtab, bg);
      default: break;
    }
  }
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path: needs nfft in {16, 32, 64} and pdiam <= min(32, nfft)");
}

template <int PREC, int Q>
int32_t launch_dft_tasks_pq(int nfft, bool full, int nr, bool noisy, const BatchTask* tasks,
                            const TaskStarts& starts, const TaskArgs& args,
                            int ntasks, int total, int grid, size_t lds,
                            const uint32_t* frame_base, uint32_t frame_off, int remap,
                            hipStream_t st) {
#define SHWFS_ARGS \
  full, nr, noisy, tasks, starts, args, ntasks, total, grid, lds, frame_base, frame_off, remap, st
  if constexpr (Q == 1) {
    switch (nfft) {
      case 16: return launch_dft_tasks_n<16, 1, PREC>(SHWFS_ARGS);
      case 32: return launch_dft_tasks_n<32, 1, PREC>(SHWFS_ARGS);
      case 64: return launch_dft_tasks_n<64, 1, PREC>(SHWFS_ARGS);
      default: break;
    }
  } else {
    switch (nfft) {
      case 32: return launch_dft_tasks_n<32, 2, PREC>(SHWFS_ARGS);
      case 64: return launch_dft_tasks_n<64, 2, PREC>(SHWFS_ARGS);
      default: break;
    }
  }
#undef SHWFS_ARGS
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT batch: needs nfft in {16, 32, 64} and pdiam <= min(32, nfft)");
}

template <int PREC, int Q>
const void* dft_table_pq(int nfft, hipStream_t st) {
  if constexpr (Q == 1) {
    switch (nfft) {
      case 16: return wfrag_table<16, 1, PREC>(st);
      case 32: return wfrag_table<32, 1, PREC>(st);
      case 64: return wfrag_table<64, 1, PREC>(st);
      default: return nullptr;
    }
  } else {
    switch (nfft) {
      case 32: return wfrag_table<32, 2, PREC>(st);
      case 64: return wfrag_table<64, 2, PREC>(st);
      default: return nullptr;
    }
  }
}

// The operand table of (nfft, PREC, Q) built into dst (cap bytes) on stream
// st, with no allocation and no wait; its size in bytes into *bytes.
// SHWFS_ERR_ARG when it does not fit in cap.
template <int N, int Q, int PREC>
static int32_t build_wfrag_into(void* dst, size_t cap, hipStream_t st, size_t* bytes) {
  using L = DftLayout<N, Q>;
  const size_t need = sizeof(Frag<PREC>) * L::MT * L::KS * 64;
  if (bytes) *bytes = need;
  if (need > cap) return fail(SHWFS_ERR_ARG, "DFT operand table larger than the scratch buffer");
  hipLaunchKernelGGL((build_wfrag_kernel<N, Q, PREC>), dim3(1), dim3(64), 0, st,
                     static_cast<Frag<PREC>*>(dst));
  SHWFS_LAUNCHED("build_wfrag_kernel");
  return SHWFS_OK;
}

template <int PREC, int Q>
int32_t build_dft_table_pq(int nfft, void* dst, size_t cap, hipStream_t st, size_t* bytes) {
  if constexpr (Q == 1) {
    switch (nfft) {
      case 16: return build_wfrag_into<16, 1, PREC>(dst, cap, st, bytes);
      case 32: return build_wfrag_into<32, 1, PREC>(dst, cap, st, bytes);
      case 64: return build_wfrag_into<64, 1, PREC>(dst, cap, st, bytes);
      default: break;
    }
  } else {
    switch (nfft) {
      case 32: return build_wfrag_into<32, 2, PREC>(dst, cap, st, bytes);
      case 64: return build_wfrag_into<64, 2, PREC>(dst, cap, st, bytes);
      default: break;
    }
  }
  return fail(SHWFS_ERR_UNSUPPORTED, "MFMA DFT path: needs nfft in {16, 32, 64} and pdiam <= min(32, nfft)");
}

#define SHWFS_DFT_EXTERN(PREC, Q)                                                         \
  extern template int32_t launch_dft_pq<PREC, Q>(const ImageParams&, int, hipStream_t,  \
                                                 const void*, const BgParams&);         \
  extern template const void* dft_table_pq<PREC, Q>(int, hipStream_t);                \
  extern template int32_t build_dft_table_pq<PREC, Q>(int, void*, size_t, hipStream_t,  \
                                                      size_t*);                         \
  extern template int32_t launch_dft_tasks_pq<PREC, Q>(                                 \
      int, bool, int, bool, const BatchTask*, const TaskStarts&, const TaskArgs&, int, int,   \
      int, size_t, const uint32_t*, uint32_t, int, hipStream_t);
SHWFS_DFT_EXTERN(SHWFS_METHOD_DFT_BF16X3, 1)
SHWFS_DFT_EXTERN(SHWFS_METHOD_DFT_BF16X3, 2)
SHWFS_DFT_EXTERN(SHWFS_METHOD_DFT_FP16, 1)
SHWFS_DFT_EXTERN(SHWFS_METHOD_DFT_FP16, 2)
#undef SHWFS_DFT_EXTERN

}  // namespace shwfs
