// Isotropic undecimated wavelet transform (B3-spline a trous) for gfx950.
//
// Semantics follow the reference serial IUWT (pymoresane/iuwt.py:66-191):
//   smoothing h_s  = separable [1,4,6,4,1]/16 with taps at 0, +-2^s, +-2^(s+1),
//                    half-sample symmetric extension at the borders;
//   decomposition  C = h_s*C0, w_s = C0 - h_s*C, C0 <- C   (iuwt.py:102-106)
//   recomposition  R <- h_s*R + w_s from coarse to fine    (iuwt.py:142-147)
//
// Kernel design.  A dilated filter only couples pixels whose row (and
// column) indices agree modulo d = 2^s, so a workgroup owns one row phase p
// (rows p, p+d, p+2d, ...) and a "comb" of columns: Q consecutive column
// phases (a contiguous Q-float chunk) repeated every d columns for TJ teeth.
// In that polyphase frame every scale is an undilated 5-tap filter with a
// halo of a few teeth, whatever d is.  The workgroup walks down its row
// phase: each input row of the comb (plus halo teeth) is staged in LDS, the
// horizontal taps are read from LDS, and the vertical taps come from
// per-thread register windows that slide one row per step.  The walk runs in
// groups of rows: while one group is processed, the next group's input rows
// and per-output operands are already in flight (see k_iuwt_walk).
//
// One decomposition scale is one launch: from C0 it produces C = H C0 (5-tap
// in both axes) and w = C0 - H(H C0), using the 9-tap filter B3*B3 for the
// doubled smoothing (exact: on the symmetric extension, which is periodic
// with period 2N, the two passes compose into one circular filter).  The
// detail epilogue optionally multiplies by a {0,1} byte mask and/or reduces
// sum((E - w)^2) against a reference plane instead of storing w.
//
// A recomposition handles two consecutive detail scales per launch where the
// shape allows (k_iuwt_recomp2: the coarser scale is the same filter at
// spacing 2 in the finer scale's frame), with results bit-identical to one
// launch per scale.  Planes of at most 256^2 pixels, and planes of at most
// 512^2 in launches of few of them, use direct-stencil kernels instead of the
// walks, and pairs of scales there run as one launch (k_iuwt_small2_dec /
// k_iuwt_small2_rec); every family gives the same bits.  An autotuner (end of this file) picks the
// walks' workgroup count and the recomposition pairing per call signature.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pms_internal.h"

namespace pms {
namespace {

constexpr int kThreads = 256;

__constant__ float c_h5[5] = {1.f / 16, 4.f / 16, 6.f / 16, 4.f / 16, 1.f / 16};
__constant__ float c_h9[9] = {1.f / 256, 8.f / 256, 28.f / 256, 56.f / 256, 70.f / 256,
                              56.f / 256, 28.f / 256, 8.f / 256, 1.f / 256};

struct ScaleJob {
  const float* in;        // C0 (decomp/smooth) or R (recomp)
  float* out_c;           // C (decomp/smooth) or R' (recomp); may be null in decomp
  float* out_w;           // decomp: detail output (null = not stored)
  const float* w_in;      // recomp: detail plane added after smoothing
  const uint8_t* mask;    // decomp: mask applied to w; recomp: mask applied to w_in (null = none)
  const float* ref;       // decomp: reduce sum((ref - w)^2) into partial[] (null = off)
  double* partial;        // one double per workgroup when ref != null
  const int* pred;        // skip the launch body when *pred == 0 (null = always run)
};

// One launch serves up to kMaxBatch jobs of the same shape and scale (grid y, or z for the
// small-plane kernel, selects the job).
struct ScaleParams {
  ScaleJob job[kMaxBatch];
  int h, w;               // plane shape (row-major, contiguous)
  int d;                  // dilation 2^s
  int q_log2, tj;         // comb: Q = 1<<q_log2 phases x TJ teeth (Q*TJ <= 256)
  int k_rows;             // output rows (in phase units) per workgroup walk
  int n_cg, n_tb, n_seg;  // grid decomposition
};

template <int MODE>
struct ModeTraits;
template <>
struct ModeTraits<IUWT_DECOMP> {
  static constexpr int H = 4;  // 9-tap halo
};
template <>
struct ModeTraits<IUWT_SMOOTH> {
  static constexpr int H = 2;
};
template <>
struct ModeTraits<IUWT_RECOMP> {
  static constexpr int H = 2;
};

// Small planes (whole plane L2-resident): one thread per output pixel loads its full dilated
// 9x9 (decomposition) or 5x5 neighbourhood at once, so a scale costs one memory latency instead
// of a row walk.  Same arithmetic as k_iuwt_walk: row sums first, then the vertical combination.
// Every load is issued before the first is used (at most 4 waves per SIMD, so the compiler may keep
// all 81 values in registers instead of waiting row by row to save registers for occupancy that
// these small grids never reach).
template <int MODE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_iuwt_small(ScaleParams sp) {
  const ScaleJob a = sp.job[blockIdx.z];
  constexpr int H = ModeTraits<MODE>::H;
  constexpr int T = 2 * H + 1;
  if (a.pred && *a.pred == 0) return;
  const int w = sp.w, h = sp.h, d = sp.d;
  const int col0 = blockIdx.x * 64 + (threadIdx.x & 63);
  const int row0 = blockIdx.y * 4 + (threadIdx.x >> 6);
  const bool live = col0 < w && row0 < h;
  const int col = live ? col0 : 0;
  const int row = live ? row0 : 0;
  int cidx[T];
#pragma unroll
  for (int b = 0; b < T; ++b) cidx[b] = fold_idx(col + (b - H) * d, w);
  float vv[T][T];
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const float* src = a.in + (size_t)fold_idx(row + (i - H) * d, h) * w;
#pragma unroll
    for (int b = 0; b < T; ++b) vv[i][b] = src[cidx[b]];
  }
  float ra[T], rb[T], xc = 0.f;
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const float* v = vv[i];
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int b = 0; b < 5; ++b) sa = fmaf(c_h5[b], v[b + H - 2], sa);
    if (MODE == IUWT_DECOMP) {
#pragma unroll
      for (int b = 0; b < 9; ++b) sb = fmaf(c_h9[b], v[b], sb);
      if (i == H) xc = v[H];
    }
    ra[i] = sa;
    rb[i] = sb;
  }
  float c = 0.f;
#pragma unroll
  for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], ra[k + H - 2], c);
  const size_t off = (size_t)row * w + col;
  if (MODE == IUWT_DECOMP) {
    float c1 = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) c1 = fmaf(c_h9[k], rb[k], c1);
    float wv = xc - c1;
    if (live) {
      if (a.mask && a.mask[off] == 0) wv = 0.f;
      if (a.out_c) a.out_c[off] = c;
      if (a.out_w) a.out_w[off] = wv;
    }
    if (a.ref) {
      double acc = 0.0;
      if (live) {
        const float df = a.ref[off] - wv;
        acc = (double)df * (double)df;
      }
      acc = block_sum_f64(acc);
      if (threadIdx.x == 0) a.partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
    }
  } else if (!live) {
    return;
  } else if (MODE == IUWT_SMOOTH) {
    a.out_c[off] = c;
  } else {
    float wi = a.w_in[off];
    if (a.mask && a.mask[off] == 0) wi = 0.f;
    a.out_c[off] = c + wi;
  }
}

// Which optional operands a walk launch has (compile-time, so the step body is straight-line).
enum WalkFlags { F_OUTC = 1, F_OUTW = 2, F_MASK = 4, F_REF = 8 };

// Output slots of steps that produce no pixel (the 2H lead-in rows, padding steps, dead lanes):
// their stores land here instead of being branched around, which keeps every step's sequence
// of memory operations identical.
__device__ float g_walk_sink[1024];

// Raw buffer access to one plane: loads take a per-lane byte offset (VGPR, fixed for the walk)
// and a per-row byte offset (SGPR), so a row's loads need no per-lane address arithmetic.  Plane
// sizes stay below 2^31 bytes (checked on the host); every offset used is inside the plane.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ int buf_u8(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return (int)__builtin_amdgcn_raw_buffer_load_b8(r, (int)voff, (int)soff, 0);
}

// A walk has at most this many steps: per-step row offsets are computed once, one step per
// lane of a wave, and read with readlane.
constexpr int kWalkMaxSteps = 64;

// Row walk with group pipelining over the comb/phase decomposition described at the top of this
// file; the walk is cut into groups of K rows: while a group's K steps run, the input
// rows and per-output operands (mask, reference, detail plane) of the next group are already in
// flight, so the loop body never drains the memory pipe.  A group's K input rows are staged in LDS
// together, so the workgroup meets at one barrier per group rather than one per row.  Loads never
// branch (dead slots read a valid column) and stores never branch (no-output steps write to
// g_walk_sink).  Row offsets of
// every step come from per-lane tables (readlane) and loads are buffer loads with the row offset
// in a scalar register, which keeps per-step scalar and address arithmetic to a few instructions.
template <int MODE, int FL, int K, int ML>
__global__ __launch_bounds__(kThreads) void k_iuwt_walk(ScaleParams sp) {
  constexpr bool OUTC = (MODE != IUWT_DECOMP) || (FL & F_OUTC);
  constexpr bool OUTW = (MODE == IUWT_DECOMP) && (FL & F_OUTW);
  constexpr bool MASK = (FL & F_MASK) != 0;
  constexpr bool REF = (MODE == IUWT_DECOMP) && (FL & F_REF);
  constexpr bool WIN = (MODE == IUWT_RECOMP);
  constexpr int H = ModeTraits<MODE>::H;
  constexpr int WA = (MODE == IUWT_DECOMP) ? 7 : 5;
  static_assert(K % 2 == 0, "group length must be even (LDS double buffer parity)");
  const ScaleJob a = sp.job[blockIdx.y];
  if (a.pred && *a.pred == 0) return;

  // 2K row buffers of ML*256 slots (every lane writes its slots): a group's K input rows are staged
  // together behind one barrier, consecutive groups in alternate halves
  extern __shared__ float lds[];
  const int Q = 1 << sp.q_log2;
  const int TJ = sp.tj;
  const int row_elems = (TJ + 2 * H) * Q;

  const unsigned nblk = gridDim.x;
  unsigned b = xcd_remap(blockIdx.x, nblk);
  const int seg = b % sp.n_seg;
  b /= sp.n_seg;
  const int tb = b % sp.n_tb;
  b /= sp.n_tb;
  const int cg = b % sp.n_cg;
  const int p = b / sp.n_cg;  // row phase
  const int d = sp.d;
  const int h = sp.h, w = sp.w;

  const int t = threadIdx.x;
  const int q = t & (Q - 1);
  const int j = t >> sp.q_log2;
  const bool thread_live = j < TJ;
  const int col = cg * Q + q + (tb * TJ + j) * d;
  const bool col_ok = thread_live && col < w;
  const unsigned col_off = col_ok ? 4u * (unsigned)col : 0u;
  const int tap_lds = thread_live ? (j + H) * Q + q : H * Q;  // dead lanes read valid slots

  const int mr = (h - p + d - 1) / d;
  const int o_begin = seg * sp.k_rows;
  const int o_end = min(o_begin + sp.k_rows, mr);
  const int n_out = o_end - o_begin;
  const int i_begin = o_begin - H;
  const int n_steps = n_out + 2 * H;  // <= kWalkMaxSteps (plan)
  const int n_groups = (n_steps + K - 1) / K;

  // per-step row byte offsets, step = lane: input rows clamped to the walk (the prefetch of the two
  // groups past the end re-reads the last row, so every index stays within the extension's range),
  // output rows clamped to a real row for steps that produce none
  const int lane = t & 63;
  const int tab_in = fold_idx(p + (i_begin + min(lane, n_steps - 1)) * d, h) * w * 4;
  const int tab_out = (p + min(max(i_begin + lane - H, o_begin), o_end - 1) * d) * w * 4;
  auto step_row = [](int tab, int s) { return (unsigned)__builtin_amdgcn_readlane(tab, min(s, kWalkMaxSteps - 1)); };

  unsigned ld_off[ML];
#pragma unroll
  for (int m = 0; m < ML; ++m) {
    const int e = t + m * kThreads;
    const int jj = e >> sp.q_log2;
    const int qq = e & (Q - 1);
    ld_off[m] = e < row_elems ? 4u * (unsigned)fold_idx(cg * Q + qq + (tb * TJ + jj - H) * d, w) : 0u;
  }
  const unsigned plane = (unsigned)h * (unsigned)w;
  const auto r_in = plane_rsrc(a.in, 4u * plane);
  const auto r_win = plane_rsrc(a.w_in, 4u * plane);
  const auto r_ref = plane_rsrc(a.ref, 4u * plane);
  const auto r_mask = plane_rsrc(a.mask, plane);

  struct Group {
    float row[K][ML];
    float win[K];
    float ref[K];
    int msk[K];
  };
  auto load_group = [&](Group& G, int g) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int s = g * K + u;
      const unsigned s_in = step_row(tab_in, s);
#pragma unroll
      for (int m = 0; m < ML; ++m) G.row[u][m] = buf_f32(r_in, ld_off[m], s_in);
      if constexpr (MASK || REF || WIN) {
        const unsigned s_o = step_row(tab_out, s);
        if constexpr (MASK) G.msk[u] = buf_u8(r_mask, col_off >> 2, s_o >> 2);
        if constexpr (REF) G.ref[u] = buf_f32(r_ref, col_off, s_o);
        if constexpr (WIN) G.win[u] = buf_f32(r_win, col_off, s_o);
      }
    }
  };

  float wa[WA];
  float wb[9];
  float wx[5];
#pragma unroll
  for (int k = 0; k < WA; ++k) wa[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) wb[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 5; ++k) wx[k] = 0.f;
  double acc = 0.0;

  // a wave writes group g's half only after the barrier of group g - 1, which every wave reaches only
  // once it has read group g - 2 (the same half)
  auto stage = [&](const Group& G, int g) {
    float* half = lds + (size_t)((g & 1) * K) * ML * kThreads;
#pragma unroll
    for (int u = 0; u < K; ++u) {
#pragma unroll
      for (int m = 0; m < ML; ++m) half[(u * ML + m) * kThreads + t] = G.row[u][m];
    }
    __syncthreads();
  };
  auto step = [&](const Group& G, int u, int s) {
    const float* cur = lds + (size_t)(((s / K) & 1) * K + u) * ML * kThreads;
    const float* rowp = cur + tap_lds;
    float ha = 0.f, hb = 0.f, hx = 0.f;
#pragma unroll
    for (int k = -2; k <= 2; ++k) ha = fmaf(c_h5[k + 2], rowp[k * Q], ha);
    if (MODE == IUWT_DECOMP) {
#pragma unroll
      for (int k = -4; k <= 4; ++k) hb = fmaf(c_h9[k + 4], rowp[k * Q], hb);
      hx = rowp[0];
    }
#pragma unroll
    for (int k = 0; k < WA - 1; ++k) wa[k] = wa[k + 1];
    wa[WA - 1] = ha;
    if (MODE == IUWT_DECOMP) {
#pragma unroll
      for (int k = 0; k < 8; ++k) wb[k] = wb[k + 1];
      wb[8] = hb;
#pragma unroll
      for (int k = 0; k < 4; ++k) wx[k] = wx[k + 1];
      wx[4] = hx;
    }
    const bool out = (unsigned)(s - 2 * H) < (unsigned)n_out && col_ok;
    const unsigned off = step_row(tab_out, s) + col_off;
    float c = 0.f;
#pragma unroll
    for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], wa[k], c);
    auto dst = [&](float* base) {
      return out ? reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) : g_walk_sink + t;
    };
    if constexpr (MODE == IUWT_DECOMP) {
      float c1 = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) c1 = fmaf(c_h9[k], wb[k], c1);
      float wv = wx[0] - c1;
      if constexpr (MASK) wv = G.msk[u] == 0 ? 0.f : wv;
      if constexpr (OUTC) *dst(a.out_c) = c;
      // detail planes are not read again by the walk: non-temporal stores keep the smoothed
      // planes (re-read at the next scale) in the caches
      if constexpr (OUTW) __builtin_nontemporal_store(wv, dst(a.out_w));
      if constexpr (REF) {
        const float df = out ? G.ref[u] - wv : 0.f;
        acc = fma((double)df, (double)df, acc);
      }
    } else if constexpr (MODE == IUWT_SMOOTH) {
      *dst(a.out_c) = c;
    } else {
      float wi = G.win[u];
      if constexpr (MASK) wi = G.msk[u] == 0 ? 0.f : wi;
      *dst(a.out_c) = c + wi;
    }
  };

  Group GA, GB;
  load_group(GA, 0);
  for (int g = 0; g < n_groups; g += 2) {
    load_group(GB, g + 1);
    stage(GA, g);
#pragma unroll
    for (int u = 0; u < K; ++u) step(GA, u, g * K + u);
    load_group(GA, g + 2);
    if (g + 1 >= n_groups) break;
    stage(GB, g + 1);
#pragma unroll
    for (int u = 0; u < K; ++u) step(GB, u, (g + 1) * K + u);
  }

  if constexpr (REF) {
    const double tot = block_sum_f64(acc);
    if (t == 0) a.partial[blockIdx.x] = tot;
  }
}

// ---- two scales per launch ----------------------------------------------------------------
// Operands of one job of a two-scale recomposition: R_s = h_s (h_{s+1} R + w_{s+1}) + w_s.  In
// the frame of dilation d = 2^s the coarse scale is an undilated filter of the frame subsampled by
// 2, so one walk serves both: the first stage filters at tooth/row spacing 2, the second at
// spacing 1.
struct PairJob {
  const float* in;       // R_{s+2} (or the coarsest detail plane)
  float* out;            // R_s
  const float* w_in0;    // detail plane of scale s
  const float* w_in1;    // detail plane of scale s+1
  const uint8_t* mask0;  // masks of those planes (null = none)
  const uint8_t* mask1;
  const int* pred;
};
struct PairParams {
  PairJob job[kMaxBatch];
  int h, w, d, q_log2, tj, k_rows, n_cg, n_tb, n_seg;
};

// Whether the symmetric extension maps index i (in [-2n, 4n)) through a mirror.
__device__ __forceinline__ bool fold_mirrors(int i, int n) {
  i = i < 0 ? i + 2 * n : i;
  i = i >= 2 * n ? i - 2 * n : i;
  return i >= n;
}

// Two-scale recomposition walk.  Per input row of R (staged in LDS with a halo of 6 teeth) each
// thread computes the intermediate T = h_{s+1} R + w_{s+1} for one T lane (teeth -2 .. TJ+1 of
// the comb: the plan keeps (TJ+4)*Q <= NT, the block size), staged in a second LDS row, and one step later
// R_s = h_s T + w_s for its output lane.  That lag gives one barrier per step: at step s the
// barrier publishes R(s) and T(s-1); stage 1 reads R(s) and writes T(s), stage 2 reads T(s-1), and
// R(s+1) is written last, all into double buffers.  A walk has at most 64 steps: the row offsets
// of step s (input, T and output rows, folded) are computed once, one step per lane, and read
// with readlane, and the rows whose T is mirrored form one 64-bit mask.
//
// The result is bit-identical to two one-scale launches.  There the fine scale reads the stored
// T at folded positions; here T is computed at the unfolded position.  On the symmetric extension
// T(x) = T(fold(x)) in exact arithmetic, and the rounded sums agree too when the operands are
// accumulated in the same order: through a mirror the neighbours of x appear in reverse order, so
// mirrored lanes (columns) and mirrored rows accumulate their 5 taps from the other end.

template <int FL, int K, int ML, int NT>
__global__ __launch_bounds__(NT) void k_iuwt_recomp2(PairParams sp) {
  constexpr bool MASK = (FL & F_MASK) != 0;
  constexpr int HR = 6, HT = 2;
  constexpr int LAG_T = 4, LAG_O = 6;  // T row = input row - 4, output row = input row - 6
  static_assert(K % 2 == 0, "group length must be even (LDS double buffer parity)");
  const PairJob a = sp.job[blockIdx.y];
  if (a.pred && *a.pred == 0) return;

  // LDS: two R rows of ML*NT slots and two T rows of NT slots (every lane writes its slot)
  extern __shared__ float lds[];
  float* rbuf0 = lds;
  float* rbuf1 = lds + ML * NT;
  float* tbuf0 = lds + 2 * ML * NT;
  float* tbuf1 = tbuf0 + NT;
  const int Q = 1 << sp.q_log2;
  const int TJ = sp.tj;
  const int r_elems = (TJ + 2 * HR) * Q;
  const int t_elems = (TJ + 2 * HT) * Q;

  const unsigned nblk = gridDim.x;
  unsigned b = xcd_remap(blockIdx.x, nblk);
  const int seg = b % sp.n_seg;
  b /= sp.n_seg;
  const int tb = b % sp.n_tb;
  b /= sp.n_tb;
  const int cg = b % sp.n_cg;
  const int p = b / sp.n_cg;
  const int d = sp.d;
  const int h = sp.h, w = sp.w;

  const int t = threadIdx.x;
  const int q = t & (Q - 1);
  const int j = t >> sp.q_log2;
  const bool thread_live = j < TJ;
  const int col = cg * Q + q + (tb * TJ + j) * d;
  const bool col_ok = thread_live && col < w;
  const unsigned col_off = col_ok ? 4u * (unsigned)col : 0u;
  const int t2_lds = thread_live ? (j + HT) * Q + q : 2 * Q;  // stage-2 taps (dead lanes: valid slots)

  // this thread's T lane: tooth j - HT of the comb
  const bool tl_ok = t < t_elems;
  const int tx = cg * Q + q + (tb * TJ + j - HT) * d;
  const unsigned tl_off = tl_ok ? 4u * (unsigned)fold_idx(tx, w) : 0u;
  const bool tl_mir = tl_ok && fold_mirrors(tx, w);
  const int tl_lds = tl_ok ? (j - HT + HR) * Q + q : 4 * Q;  // dead lanes read valid slots
  const bool wave_mir = __any(tl_mir);

  const int mr = (h - p + d - 1) / d;
  const int o_begin = seg * sp.k_rows;
  const int o_end = min(o_begin + sp.k_rows, mr);
  const int n_out = o_end - o_begin;
  const int i_begin = o_begin - LAG_O;
  const int n_steps = n_out + 2 * LAG_O;  // <= kWalkMaxSteps (plan)
  // one more step than rows: stage 2 of the last row runs in it
  const int n_groups = (n_steps + 1 + K - 1) / K;

  // per-step row byte offsets, step = lane: input rows clamped to the walk (the prefetch past the
  // end re-reads the last row), T rows before the first complete T window use the first real one
  const int lane = t & 63;
  const int ls = min(lane, n_steps - 1);
  const int tab_in = fold_idx(p + (i_begin + ls) * d, h) * w * 4;
  const int tab_t = fold_idx(p + (i_begin + max(ls, 2 * LAG_T) - LAG_T) * d, h) * w * 4;
  const int tab_out = (p + min(max(i_begin + lane - LAG_O, o_begin), o_end - 1) * d) * w * 4;
  const unsigned long long row_mir = __ballot(fold_mirrors(p + (i_begin + lane - LAG_T) * d, h));
  auto step_row = [](int tab, int s) { return (unsigned)__builtin_amdgcn_readlane(tab, min(s, kWalkMaxSteps - 1)); };

  unsigned ld_off[ML];
#pragma unroll
  for (int m = 0; m < ML; ++m) {
    const int e = t + m * NT;
    const int jj = e >> sp.q_log2;
    const int qq = e & (Q - 1);
    ld_off[m] = e < r_elems ? 4u * (unsigned)fold_idx(cg * Q + qq + (tb * TJ + jj - HR) * d, w) : 0u;
  }
  const unsigned plane = (unsigned)h * (unsigned)w;
  const auto r_in = plane_rsrc(a.in, 4u * plane);
  const auto r_w0 = plane_rsrc(a.w_in0, 4u * plane);
  const auto r_w1 = plane_rsrc(a.w_in1, 4u * plane);
  const auto r_m0 = plane_rsrc(a.mask0, plane);
  const auto r_m1 = plane_rsrc(a.mask1, plane);

  struct Group {
    float row[K][ML];
    float w1[K];
    int m1[K];
    float w0[K];
    int m0[K];
  };
  auto load_group = [&](Group& G, int g) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int s = g * K + u;
      const unsigned s_in = step_row(tab_in, s), s_t = step_row(tab_t, s), s_o = step_row(tab_out, s);
#pragma unroll
      for (int m = 0; m < ML; ++m) G.row[u][m] = buf_f32(r_in, ld_off[m], s_in);
      G.w1[u] = buf_f32(r_w1, tl_off, s_t);
      G.w0[u] = buf_f32(r_w0, col_off, s_o);
      if constexpr (MASK) {
        G.m1[u] = buf_u8(r_m1, tl_off >> 2, s_t >> 2);
        G.m0[u] = buf_u8(r_m0, col_off >> 2, s_o >> 2);
      }
    }
  };
  auto put_row = [&](float* buf, const float* row) {
#pragma unroll
    for (int m = 0; m < ML; ++m) buf[t + m * NT] = row[m];
  };

  float vt[9];
  float vr[5];
#pragma unroll
  for (int k = 0; k < 9; ++k) vt[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 5; ++k) vr[k] = 0.f;
  float w0_prev = 0.f;

  auto step = [&](const Group& G, const Group& Gn, int u, int s) {
    float* rcur = (u & 1) ? rbuf1 : rbuf0;
    float* tcur = (u & 1) ? tbuf1 : tbuf0;
    float* rnext = (u & 1) ? rbuf0 : rbuf1;
    const float* tprev = (u & 1) ? tbuf0 : tbuf1;
    __syncthreads();
    // stage 1, scale s+1 (taps at tooth spacing 2 and row spacing 2): T of step s
    {
      float x[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) x[k] = rcur[tl_lds + (2 * k - 4) * Q];
      float hr = 0.f;
#pragma unroll
      for (int k = 0; k < 5; ++k) hr = fmaf(c_h5[k], x[k], hr);
      if (wave_mir) {
        float hm = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) hm = fmaf(c_h5[k], x[4 - k], hm);
        hr = tl_mir ? hm : hr;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) vt[k] = vt[k + 1];
      vt[8] = hr;
      float c = 0.f;
      if ((row_mir >> (s & 63)) & 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], vt[8 - 2 * k], c);
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], vt[2 * k], c);
      }
      float wi = G.w1[u];
      if constexpr (MASK) wi = G.m1[u] == 0 ? 0.f : wi;
      tcur[t] = c + wi;
    }
    // stage 2, scale s (taps at spacing 1): output of step s-1
    {
      float ht = 0.f;
#pragma unroll
      for (int k = -2; k <= 2; ++k) ht = fmaf(c_h5[k + 2], tprev[t2_lds + k * Q], ht);
#pragma unroll
      for (int k = 0; k < 4; ++k) vr[k] = vr[k + 1];
      vr[4] = ht;
      float c = 0.f;
#pragma unroll
      for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], vr[k], c);
      const int so = s - 1;
      const bool out = (unsigned)(so - 2 * LAG_O) < (unsigned)n_out && col_ok;
      float* dst = reinterpret_cast<float*>(reinterpret_cast<char*>(a.out) + step_row(tab_out, max(so, 0)) + col_off);
      *(out ? dst : g_walk_sink + t) = c + w0_prev;
    }
    float wi = G.w0[u];
    if constexpr (MASK) wi = G.m0[u] == 0 ? 0.f : wi;
    w0_prev = wi;
    put_row(rnext, u + 1 < K ? G.row[u + 1] : Gn.row[0]);
  };

  Group GA, GB;
  load_group(GA, 0);
  put_row(rbuf0, GA.row[0]);
  for (int g = 0; g < n_groups; g += 2) {
    load_group(GB, g + 1);
#pragma unroll
    for (int u = 0; u < K; ++u) step(GA, GB, u, g * K + u);
    load_group(GA, g + 2);
    if (g + 1 >= n_groups) break;
#pragma unroll
    for (int u = 0; u < K; ++u) step(GB, GA, u, (g + 1) * K + u);
  }
}

// ---- two decomposition scales per launch ----------------------------------------------------
// Operands of one job of a two-scale decomposition: from C_s the detail planes w_s and w_{s+1} and
// the smoothed C_{s+2}; C_{s+1} = h_s C_s exists only as LDS rows.
struct Dec2Job {
  const float* in;       // C_s
  float* out_c;          // C_{s+2} (null: not kept)
  float* out_w0;         // w_s
  float* out_w1;         // w_{s+1}
  const uint8_t* mask0;  // masks of w_s and w_{s+1} (null = none)
  const uint8_t* mask1;
  const int* pred;
};
struct Dec2Params {
  Dec2Job job[kMaxBatch];
  int h, w, d, q_log2, tj, k_rows, n_cg, n_tb, n_seg;
};

// A two-scale walk has at most this many steps: its per-step tables use two registers (steps
// 0..63 and 64..127), so the 21 lead-in and lag steps are a small share of a walk.
constexpr int kDec2MaxSteps = 128;

// Two-scale decomposition walk over the comb/phase frame of scale s (dilation d = 2^s), where scale
// s+1 is the same filter at tooth and row spacing 2.  Stage 1 filters the staged C_s row (halo 10
// teeth) into C_{s+1} for one C_{s+1} lane per thread (teeth -8 .. TJ+7, staged in an LDS row) and
// into w_s = C_s - H9 C_s for the output lane; stage 2 filters the C_{s+1} row of the previous step
// at spacing 2 into C_{s+2} and w_{s+1} = C_{s+1} - H9 C_{s+1}.  Input row i gives C_{s+1} row
// i-2 and w_s row i-4 in stage 1; C_{s+1} row i-3 gives C_{s+2} row i-7 and w_{s+1} row i-11 in
// stage 2, so a walk of n output rows reads n+20 input rows in n+21 steps, with one barrier per
// step (it publishes the step's C_s row and the previous step's C_{s+1} row; both are double
// buffered).  Every sum has the one-scale walk's operands in its order, so the planes are
// bit-identical to two one-scale launches: C_{s+1} lanes and rows beyond the border are computed at
// the unfolded position, which on the symmetric extension equals the folded one when a mirrored lane
// or row accumulates its taps from the other end (as in the two-scale recomposition).
template <int FL, int K, int ML, int NT>
__global__ __launch_bounds__(NT) void k_iuwt_dec2(Dec2Params sp) {
  constexpr bool OUTC = (FL & F_OUTC) != 0;
  constexpr bool MASK = (FL & F_MASK) != 0;
  constexpr int HT = 8, HC = 10;  // C_{s+1} lanes and staged C_s teeth beyond the output teeth
  constexpr int LAG = 10;         // input rows before the first output row
  static_assert(K % 2 == 0, "group length must be even (LDS double buffer parity)");
  const Dec2Job a = sp.job[blockIdx.y];
  if (a.pred && *a.pred == 0) return;

  // LDS: two C_s rows of ML*NT slots and two C_{s+1} rows of NT slots (every lane writes its slot)
  extern __shared__ float lds[];
  float* cbuf0 = lds;
  float* cbuf1 = lds + ML * NT;
  float* tbuf0 = lds + 2 * ML * NT;
  float* tbuf1 = tbuf0 + NT;
  const int Q = 1 << sp.q_log2;
  const int TJ = sp.tj;
  const int c_elems = (TJ + 2 * HC) * Q;
  const int t_elems = (TJ + 2 * HT) * Q;

  const unsigned nblk = gridDim.x;
  unsigned b = xcd_remap(blockIdx.x, nblk);
  const int seg = b % sp.n_seg;
  b /= sp.n_seg;
  const int tb = b % sp.n_tb;
  b /= sp.n_tb;
  const int cg = b % sp.n_cg;
  const int p = b / sp.n_cg;  // row phase
  const int d = sp.d;
  const int h = sp.h, w = sp.w;

  const int t = threadIdx.x;
  const int q = t & (Q - 1);
  const int jt = (t >> sp.q_log2) - HT;  // tooth of this thread's C_{s+1} lane and output lane
  const bool t_live = t < t_elems;
  const bool o_live = jt >= 0 && jt < TJ;
  const int col = cg * Q + q + (tb * TJ + jt) * d;  // natural column of the lane (possibly beyond a border)
  const bool col_ok = o_live && col < w;
  const unsigned col_off = col_ok ? 4u * (unsigned)col : 0u;
  const bool tl_mir = t_live && fold_mirrors(col, w);
  const int c_lds = t_live ? t + 2 * Q : 2 * Q;  // C_s slot of the lane's tooth (dead lanes: valid slots)
  const int o_lds = o_live ? t : HT * Q;         // the output lane's C_{s+1} slot; its C_s slot is o_lds + 2Q

  const int mr = (h - p + d - 1) / d;
  const int o_begin = seg * sp.k_rows;
  const int o_end = min(o_begin + sp.k_rows, mr);
  const int n_out = o_end - o_begin;
  const int i_begin = o_begin - LAG;
  const int n_in = n_out + 2 * LAG;
  const int n_steps = n_in + 1;  // <= kDec2MaxSteps (plan)
  const int n_groups = (n_steps + K - 1) / K;

  // per-step row byte offsets, step = lane (and lane + 64): input rows clamped to the walk (the
  // prefetch past the end re-reads the last row), output rows clamped to a real row for steps that
  // produce none
  const int lane = t & 63;
  auto in_row = [&](int s) { return fold_idx(p + (i_begin + min(s, n_in - 1)) * d, h) * w * 4; };
  auto out_row = [&](int s, int lag) { return (p + min(max(i_begin + s - lag, o_begin), o_end - 1) * d) * w * 4; };
  const int tin0 = in_row(lane), tin1 = in_row(lane + 64);
  const int tw0a = out_row(lane, 4), tw0b = out_row(lane + 64, 4);
  const int tca = out_row(lane, 7), tcb = out_row(lane + 64, 7);
  const int tw1a = out_row(lane, 11), tw1b = out_row(lane + 64, 11);
  // the C_{s+1} row of step s (input row - 2) lies beyond a border: its vertical taps run reversed
  auto mir_of = [&](int s) { return fold_mirrors(p + (i_begin + s - 2) * d, h); };
  const unsigned long long mira = __ballot(mir_of(lane));
  const unsigned long long mirb = __ballot(mir_of(lane + 64));
  // both table registers are read and the step's one selected, without a branch
  auto tab = [](int lo, int hi, int s) {
    const unsigned a = (unsigned)__builtin_amdgcn_readlane(lo, s & 63);
    const unsigned b = (unsigned)__builtin_amdgcn_readlane(hi, s & 63);
    return s < 64 ? a : b;
  };
  auto mirrored = [&](int s) { return (((s < 64 ? mira : mirb) >> (s & 63)) & 1ull) != 0; };

  unsigned ld_off[ML];
#pragma unroll
  for (int m = 0; m < ML; ++m) {
    const int e = t + m * NT;
    const int jj = e >> sp.q_log2;
    const int qq = e & (Q - 1);
    ld_off[m] = e < c_elems ? 4u * (unsigned)fold_idx(cg * Q + qq + (tb * TJ + jj - HC) * d, w) : 0u;
  }
  const unsigned plane = (unsigned)h * (unsigned)w;
  const auto r_in = plane_rsrc(a.in, 4u * plane);
  const auto r_m0 = plane_rsrc(a.mask0, plane);
  const auto r_m1 = plane_rsrc(a.mask1, plane);

  struct Group {
    float row[K][ML];
    int m0[K];
    int m1[K];
  };
  auto load_group = [&](Group& G, int g) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int s = g * K + u;
      const unsigned s_in = tab(tin0, tin1, s);
#pragma unroll
      for (int m = 0; m < ML; ++m) G.row[u][m] = buf_f32(r_in, ld_off[m], s_in);
      if constexpr (MASK) {
        G.m0[u] = buf_u8(r_m0, col_off >> 2, tab(tw0a, tw0b, s) >> 2);
        G.m1[u] = buf_u8(r_m1, col_off >> 2, tab(tw1a, tw1b, s) >> 2);
      }
    }
  };
  auto put_row = [&](float* buf, const float* row) {
#pragma unroll
    for (int m = 0; m < ML; ++m) buf[t + m * NT] = row[m];
  };

  float va[5], vb[9], vx[5], v2a[9], v2b[17], v2x[9];
#pragma unroll
  for (int k = 0; k < 5; ++k) va[k] = vx[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) vb[k] = v2a[k] = v2x[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 17; ++k) v2b[k] = 0.f;

  auto dst = [&](float* base, unsigned row_off, bool out) {
    return out ? reinterpret_cast<float*>(reinterpret_cast<char*>(base) + row_off + col_off) : g_walk_sink + t;
  };

  auto step = [&](const Group& G, const Group& Gn, int u, int s) {
    const float* ccur = (u & 1) ? cbuf1 : cbuf0;
    float* cnext = (u & 1) ? cbuf0 : cbuf1;
    float* tcur = (u & 1) ? tbuf1 : tbuf0;
    const float* tprev = (u & 1) ? tbuf0 : tbuf1;
    __syncthreads();
    // stage 1, scale s (spacing 1)
    {
      float x[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) x[k] = ccur[c_lds + (k - 2) * Q];
      // both accumulation orders, selected without a branch, so the two stages form one block of
      // independent work for the scheduler
      float ha = 0.f, hm = 0.f;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        ha = fmaf(c_h5[k], x[k], ha);
        hm = fmaf(c_h5[k], x[4 - k], hm);
      }
      ha = tl_mir ? hm : ha;
#pragma unroll
      for (int k = 0; k < 4; ++k) va[k] = va[k + 1];
      va[4] = ha;
      float c = 0.f, cm = 0.f;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        c = fmaf(c_h5[k], va[k], c);
        cm = fmaf(c_h5[k], va[4 - k], cm);
      }
      tcur[t] = mirrored(s) ? cm : c;  // C_{s+1} of input row - 2
      float y[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) y[k] = ccur[o_lds + (k - 2) * Q];
      float hb = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) hb = fmaf(c_h9[k], y[k], hb);
#pragma unroll
      for (int k = 0; k < 8; ++k) vb[k] = vb[k + 1];
      vb[8] = hb;
#pragma unroll
      for (int k = 0; k < 4; ++k) vx[k] = vx[k + 1];
      vx[4] = y[4];
      float c1 = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) c1 = fmaf(c_h9[k], vb[k], c1);
      float wv = vx[0] - c1;  // w_s of input row - 4
      if constexpr (MASK) wv = G.m0[u] == 0 ? 0.f : wv;
      const bool out = (unsigned)(s - 2 * LAG + 6) < (unsigned)n_out && col_ok;
      __builtin_nontemporal_store(wv, dst(a.out_w0, tab(tw0a, tw0b, s), out));
    }
    // stage 2, scale s+1 (spacing 2), on the C_{s+1} row of the previous step (input row - 3)
    {
      float z[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) z[k] = tprev[o_lds + 2 * (k - 4) * Q];
      float ha = 0.f, hb = 0.f;
#pragma unroll
      for (int k = 0; k < 5; ++k) ha = fmaf(c_h5[k], z[k + 2], ha);
#pragma unroll
      for (int k = 0; k < 9; ++k) hb = fmaf(c_h9[k], z[k], hb);
#pragma unroll
      for (int k = 0; k < 8; ++k) v2a[k] = v2a[k + 1];
      v2a[8] = ha;
#pragma unroll
      for (int k = 0; k < 16; ++k) v2b[k] = v2b[k + 1];
      v2b[16] = hb;
#pragma unroll
      for (int k = 0; k < 8; ++k) v2x[k] = v2x[k + 1];
      v2x[8] = z[4];
      if constexpr (OUTC) {
        float c = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) c = fmaf(c_h5[k], v2a[2 * k], c);
        const bool out = (unsigned)(s - 2 * LAG + 3) < (unsigned)n_out && col_ok;
        *dst(a.out_c, tab(tca, tcb, s), out) = c;  // C_{s+2} of input row - 7
      }
      float c1 = 0.f;
#pragma unroll
      for (int k = 0; k < 9; ++k) c1 = fmaf(c_h9[k], v2b[2 * k], c1);
      float wv = v2x[0] - c1;  // w_{s+1} of input row - 11
      if constexpr (MASK) wv = G.m1[u] == 0 ? 0.f : wv;
      const bool out = (unsigned)(s - 2 * LAG - 1) < (unsigned)n_out && col_ok;
      __builtin_nontemporal_store(wv, dst(a.out_w1, tab(tw1a, tw1b, s), out));
    }
    put_row(cnext, u + 1 < K ? G.row[u + 1] : Gn.row[0]);
  };

  tbuf0[t] = 0.f;
  tbuf1[t] = 0.f;
  Group GA, GB;
  load_group(GA, 0);
  put_row(cbuf0, GA.row[0]);
  for (int g = 0; g < n_groups; g += 2) {
    load_group(GB, g + 1);
#pragma unroll
    for (int u = 0; u < K; ++u) step(GA, GB, u, g * K + u);
    load_group(GA, g + 2);
    if (g + 1 >= n_groups) break;
#pragma unroll
    for (int u = 0; u < K; ++u) step(GB, GA, u, (g + 1) * K + u);
  }
}

struct Plan {
  int q_log2, tj, k_rows, n_cg, n_tb, n_seg, blocks, lds_bytes;
  int nt = kThreads;  // threads per block
};

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

// tuning knobs, read once per process
struct IuwtTuning {
  int qmax, krows, target_blocks;
  long small_pixels;  // planes up to this many pixels always use the direct-stencil kernels
  int pair_min_rows;  // two-scale recomposition needs this many rows per row phase of the finer scale
  IuwtTuning()
      : qmax(env_int("PMS_IUWT_QMAX", 32)),
        krows(env_int("PMS_IUWT_KROWS", 0)),
        target_blocks(env_int("PMS_IUWT_BLOCKS", 512)),
        small_pixels(env_int("PMS_IUWT_SMALL", 256 * 256)),
        pair_min_rows(env_int("PMS_IUWT_PAIR_MIN_ROWS", 8)) {}
};
const IuwtTuning& tuning() {
  static const IuwtTuning t;
  return t;
}

// Launch parameters an autotuned call runs with (set for the duration of the call on its thread).
struct TuneParams {
  int target_blocks;  // workgroups a one-scale walk aims for
  int pair;           // recomposition pairing (0 off, 1 from the coarsest, 2 from the finest);
                      // decomposition: two-scale launches (1) or one scale per launch (0)
};

// One decomposition scale per launch unless PMS_IUWT_DEC2=1 asks for two-scale launches where they
// apply: those give the same planes and move fewer bytes, but measured slower (their two dependent
// filter stages per step leave the walk latency-bound rather than bandwidth-bound).
constexpr int kDec2Default = 0;
thread_local const TuneParams* t_active = nullptr;

int plan_target_blocks() { return t_active ? t_active->target_blocks : tuning().target_blocks; }

// Output rows per walk once the walk count is fixed: the same number of segments, of equal length
// (a length cap can leave a short last segment, e.g. 56 + 8 rows, whose walk pays the same lead-in
// while the long one sets the launch's time), unless that would go below kmin.
int balanced_rows(int mr, int k, int kmin) {
  if (k <= 0 || getenv("PMS_IUWT_KROWS")) return k;
  const int kb = ceil_div(mr, ceil_div(mr, k));
  return kb >= kmin ? kb : k;
}

Plan make_plan(int h, int w, int scale, int halo) {
  Plan pl;
  const int d = 1 << scale;
  const int teeth = ceil_div(w, d);
  const int qmax = tuning().qmax;
  int q = d < qmax ? d : qmax;
  int tj = kThreads / q;
  int tp = 1;
  while (tp < teeth) tp <<= 1;
  if (tj > tp) tj = tp;
  while (q * tj < kThreads && q < d) q <<= 1;  // widen the chunk when teeth are few
  pl.q_log2 = ilog2(q);
  pl.tj = tj;
  const int mr = ceil_div(h, d);
  // aim for enough workgroups to fill 256 CUs several times over
  const int strips = d * ceil_div(d, q) * ceil_div(teeth, tj);
  int k = tuning().krows;
  if (k <= 0) {
    const long target = plan_target_blocks();
    long segs = (target + strips - 1) / strips;
    if (segs < 1) segs = 1;
    k = ceil_div(mr, (int)segs);
    const int kmin = 4 * halo;
    if (k < kmin) k = kmin;
  }
  if (k > kWalkMaxSteps - 2 * halo) k = kWalkMaxSteps - 2 * halo;  // a walk's steps fit one wave's lanes
  if (k > mr) k = mr;
  k = balanced_rows(mr, k, 4 * halo);
  pl.k_rows = k;
  pl.n_cg = ceil_div(d, q);
  pl.n_tb = ceil_div(teeth, tj);
  pl.n_seg = ceil_div(mr, k);
  pl.blocks = d * pl.n_cg * pl.n_tb * pl.n_seg;
  const int elems = (tj + 2 * halo) * q;
  const int ml = elems <= 2 * kThreads ? 2 : (elems <= 4 * kThreads ? 4 : 8);
  pl.lds_bytes = 2 * ml * kThreads * (int)sizeof(float);
  return pl;
}

constexpr int kWalkGroup = 2;
// rows per prefetch group of a one-scale decomposition walk: two groups of four rows in flight hide
// more of the first scale's HBM latency (its input is the only plane not just written to the caches);
// the recompositions keep groups of two, which measured faster there
constexpr int kDecompGroup = 4;

template <int MODE, int FL>
void launch_walk_fl(ScaleParams& sp, int count, const Plan& pl, int halo, hipStream_t st) {
  constexpr int K = MODE == IUWT_DECOMP ? kDecompGroup : kWalkGroup;
  const int elems = (pl.tj + 2 * halo) << pl.q_log2;
  dim3 grid(pl.blocks, count), block(kThreads);
  // 2K staged rows of ML slots per lane
  const size_t lds = [&](int ml) { return (size_t)2 * K * ml * kThreads * sizeof(float); }(
      elems <= 2 * kThreads ? 2 : (elems <= 4 * kThreads ? 4 : 8));
  if (elems <= 2 * kThreads)
    hipLaunchKernelGGL((k_iuwt_walk<MODE, FL, K, 2>), grid, block, lds, st, sp);
  else if (elems <= 4 * kThreads)
    hipLaunchKernelGGL((k_iuwt_walk<MODE, FL, K, 4>), grid, block, lds, st, sp);
  else
    hipLaunchKernelGGL((k_iuwt_walk<MODE, FL, K, 8>), grid, block, lds, st, sp);
  PMS_LAUNCH_CHECK();
}

template <int MODE>
void launch_walk(ScaleParams& sp, int count, const Plan& pl, int halo, int fl, hipStream_t st) {
  if constexpr (MODE == IUWT_DECOMP) {
    switch (fl) {
#define PMS_WALK_CASE(F) \
  case F: launch_walk_fl<MODE, F>(sp, count, pl, halo, st); break;
      PMS_WALK_CASE(0) PMS_WALK_CASE(1) PMS_WALK_CASE(2) PMS_WALK_CASE(3) PMS_WALK_CASE(4) PMS_WALK_CASE(5)
      PMS_WALK_CASE(6) PMS_WALK_CASE(7) PMS_WALK_CASE(8) PMS_WALK_CASE(9) PMS_WALK_CASE(10) PMS_WALK_CASE(11)
      PMS_WALK_CASE(12) PMS_WALK_CASE(13) PMS_WALK_CASE(14) PMS_WALK_CASE(15)
#undef PMS_WALK_CASE
      default: fail(PMS_ERR_INTERNAL, "bad IUWT walk flags");
    }
  } else if constexpr (MODE == IUWT_RECOMP) {
    if (fl & F_MASK)
      launch_walk_fl<MODE, F_MASK>(sp, count, pl, halo, st);
    else
      launch_walk_fl<MODE, 0>(sp, count, pl, halo, st);
  } else {
    launch_walk_fl<MODE, 0>(sp, count, pl, halo, st);
  }
}

// Plan of a two-scale recomposition at scales (s, s+1) on the frame of scale s, or false when the
// pair is not used for this shape (small planes, too few rows per row phase).  The comb keeps the
// T row ((TJ+4) teeth of Q phases) within one lane per thread.
bool recomp2_plan(int h, int w, int s, Plan* out) {
  if ((long)h * w <= tuning().small_pixels || (long)h * w >= (1L << 29)) return false;
  if ((4L << s) > h || (4L << s) > w) return false;
  const int d = 1 << s;
  const int mr = ceil_div(h, d);
  if (mr < tuning().pair_min_rows) return false;
  const int teeth = ceil_div(w, d);
  int q = std::min(d, tuning().qmax);
  if (q > 32) q = 32;
  // wide chunks take 512-thread blocks so the 4 halo teeth of the T row cost a small share
  const int nt = q >= 16 ? 2 * kThreads : kThreads;
  auto tj_for = [&](int qq) { return std::min(nt / qq - 4, teeth); };
  // widen the chunk while that gives more live output lanes (few teeth)
  while (q < d && 2 * q <= 32 && tj_for(2 * q) * 2 * q > tj_for(q) * q) q *= 2;
  const int tj = tj_for(q);
  if (tj < 1) return false;
  Plan pl;
  pl.nt = nt;
  pl.q_log2 = ilog2(q);
  pl.tj = tj;
  const int strips = d * ceil_div(d, q) * ceil_div(teeth, tj);
  int k = tuning().krows;
  if (k <= 0) {
    const long target = plan_target_blocks();
    long segs = (target + strips - 1) / strips;
    if (segs < 1) segs = 1;
    k = ceil_div(mr, (int)segs);
    if (k < 24) k = 24;
  }
  if (k > kWalkMaxSteps - 12) k = kWalkMaxSteps - 12;  // a walk's steps fit one wave's lanes
  if (k > mr) k = mr;
  k = balanced_rows(mr, k, 24);
  pl.k_rows = k;
  pl.n_cg = ceil_div(d, q);
  pl.n_tb = ceil_div(teeth, tj);
  pl.n_seg = ceil_div(mr, k);
  pl.blocks = d * pl.n_cg * pl.n_tb * pl.n_seg;
  const int r_elems = (tj + 12) * q;
  const int ml = r_elems <= 2 * nt ? 2 : (r_elems <= 4 * nt ? 4 : 8);
  pl.lds_bytes = 2 * (ml + 1) * nt * (int)sizeof(float);
  *out = pl;
  return true;
}

template <int FL, int NT>
void launch_recomp2_nt(PairParams& sp, int count, const Plan& pl, hipStream_t st) {
  const int elems = (pl.tj + 12) << pl.q_log2;
  dim3 grid(pl.blocks, count), block(NT);
  if (elems <= 2 * NT)
    hipLaunchKernelGGL((k_iuwt_recomp2<FL, kWalkGroup, 2, NT>), grid, block, pl.lds_bytes, st, sp);
  else if (elems <= 4 * NT)
    hipLaunchKernelGGL((k_iuwt_recomp2<FL, kWalkGroup, 4, NT>), grid, block, pl.lds_bytes, st, sp);
  else
    hipLaunchKernelGGL((k_iuwt_recomp2<FL, kWalkGroup, 8, NT>), grid, block, pl.lds_bytes, st, sp);
  PMS_LAUNCH_CHECK();
}

template <int FL>
void launch_recomp2_fl(PairParams& sp, int count, const Plan& pl, hipStream_t st) {
  if (pl.nt == 2 * kThreads)
    launch_recomp2_nt<FL, 2 * kThreads>(sp, count, pl, st);
  else
    launch_recomp2_nt<FL, kThreads>(sp, count, pl, st);
}

// jobs[0..count) share shape and scale pair; masked and unmasked jobs go to separate launches
void recomp2_batch(const PairJob* jobs, int count, int h, int w, int s, const Plan& pl, hipStream_t st) {
  PairParams sp{};
  sp.h = h;
  sp.w = w;
  sp.d = 1 << s;
  sp.q_log2 = pl.q_log2;
  sp.tj = pl.tj;
  sp.k_rows = pl.k_rows;
  sp.n_cg = pl.n_cg;
  sp.n_tb = pl.n_tb;
  sp.n_seg = pl.n_seg;
  for (int masked = 0; masked < 2; ++masked) {
    int nb = 0;
    auto flush = [&]() {
      if (nb == 0) return;
      if (masked)
        launch_recomp2_fl<F_MASK>(sp, nb, pl, st);
      else
        launch_recomp2_fl<0>(sp, nb, pl, st);
      nb = 0;
    };
    for (int k = 0; k < count; ++k) {
      if ((jobs[k].mask0 != nullptr) != (masked != 0)) continue;
      sp.job[nb++] = jobs[k];
      if (nb == kMaxBatch) flush();
    }
    flush();
  }
}

// Launch shape of a two-scale decomposition at scales (s, s+1), or false where it does not apply:
// planes for the row walks, d = 2^s with 20 d within both sides (every staged or intermediate
// index folds at most once), and chunks of at most 8 phases (with 16 the 16 halo teeth of the
// C_{s+1} lanes would take half of a block's lanes).
bool dec2_plan(int h, int w, int s, Plan* out) {
  if ((long)h * w <= tuning().small_pixels || (long)h * w >= (1L << 29)) return false;
  const int d = 1 << s;
  if (d > 8 || 20L * d > h || 20L * d > w) return false;
  const int q = d;
  const int nt = q <= 2 ? kThreads : 2 * kThreads;
  const int teeth = ceil_div(w, d);
  const int tj = std::min(nt / q - 16, teeth);
  if (tj < 8) return false;
  Plan pl;
  pl.nt = nt;
  pl.q_log2 = ilog2(q);
  pl.tj = tj;
  const int mr = ceil_div(h, d);
  const int strips = d * ceil_div(d, q) * ceil_div(teeth, tj);
  int k = tuning().krows;
  if (k <= 0) {
    const long target = plan_target_blocks();
    long segs = (target + strips - 1) / strips;
    if (segs < 1) segs = 1;
    k = ceil_div(mr, (int)segs);
    if (k < 40) k = 40;
  }
  if (k > kDec2MaxSteps - 21) k = kDec2MaxSteps - 21;  // a walk's steps fit two table registers
  if (k > mr) k = mr;
  k = balanced_rows(mr, k, 40);
  pl.k_rows = k;
  pl.n_cg = ceil_div(d, q);
  pl.n_tb = ceil_div(teeth, tj);
  pl.n_seg = ceil_div(mr, k);
  pl.blocks = d * pl.n_cg * pl.n_tb * pl.n_seg;
  pl.lds_bytes = (2 * 2 * nt + 2 * nt) * (int)sizeof(float);  // (tj + 20) q <= 2 nt C_s slots per row
  *out = pl;
  return true;
}

template <int FL, int NT>
void launch_dec2_nt(Dec2Params& sp, int count, const Plan& pl, hipStream_t st) {
  PMS_REQUIRE(((pl.tj + 20) << pl.q_log2) <= 2 * NT && ((pl.tj + 16) << pl.q_log2) <= NT,
              "two-scale decomposition plan exceeds its block");
  hipLaunchKernelGGL((k_iuwt_dec2<FL, kWalkGroup, 2, NT>), dim3(pl.blocks, count), dim3(NT), pl.lds_bytes, st, sp);
  PMS_LAUNCH_CHECK();
}

template <int FL>
void launch_dec2_fl(Dec2Params& sp, int count, const Plan& pl, hipStream_t st) {
  if (pl.nt == 2 * kThreads)
    launch_dec2_nt<FL, 2 * kThreads>(sp, count, pl, st);
  else
    launch_dec2_nt<FL, kThreads>(sp, count, pl, st);
}

// jobs[0..count) share shape and scale pair; jobs with and without a mask or a kept C_{s+2} go to
// separate launches
void dec2_batch(const Dec2Job* jobs, int count, int h, int w, int s, const Plan& pl, hipStream_t st) {
  Dec2Params sp{};
  sp.h = h;
  sp.w = w;
  sp.d = 1 << s;
  sp.q_log2 = pl.q_log2;
  sp.tj = pl.tj;
  sp.k_rows = pl.k_rows;
  sp.n_cg = pl.n_cg;
  sp.n_tb = pl.n_tb;
  sp.n_seg = pl.n_seg;
  for (int kind = 0; kind < 4; ++kind) {
    const bool masked = (kind & 1) != 0, keep_c = (kind & 2) != 0;
    int nb = 0;
    auto flush = [&]() {
      if (nb == 0) return;
      switch (kind) {
        case 0: launch_dec2_fl<F_OUTW>(sp, nb, pl, st); break;
        case 1: launch_dec2_fl<F_OUTW | F_MASK>(sp, nb, pl, st); break;
        case 2: launch_dec2_fl<F_OUTW | F_OUTC>(sp, nb, pl, st); break;
        default: launch_dec2_fl<F_OUTW | F_OUTC | F_MASK>(sp, nb, pl, st); break;
      }
      nb = 0;
    };
    for (int k = 0; k < count; ++k) {
      if ((jobs[k].mask0 != nullptr) != masked || (jobs[k].out_c != nullptr) != keep_c) continue;
      sp.job[nb++] = jobs[k];
      if (nb == kMaxBatch) flush();
    }
    flush();
  }
}

// ---- two scales per launch on small planes ---------------------------------------------------
// Planes in the direct-stencil range take a pair of scales (s, s+1) in one launch.  A workgroup owns
// a tile of one phase class: rows pr + d*i and columns pc + d*j for 16 consecutive lattice indices i
// and j (d = 2^s).  On that lattice scale s is the undilated filter and scale s+1 the same filter at
// spacing 2, so the tile needs its input only within 10 lattice steps (decomposition) or 4
// (recomposition); that region is staged in LDS once, and the intermediate plane -- C_{s+1} of a
// decomposition, T = h_{s+1} R + w_{s+1} of a recomposition -- is computed from it in LDS for the
// tile and a halo of 8 (2) lattice steps and never reaches memory.  Region and halo positions keep
// their unfolded lattice coordinates: a staged input value is the symmetric extension there, and an
// intermediate value at a mirrored position (an odd number of reflections away) accumulates its
// taps in reverse order, which makes it the same sum, in the same order, as the one-scale kernel
// forms at the folded position.  Every plane is therefore bit-identical to two k_iuwt_small launches.
constexpr int kTile2 = 16;  // tile side in lattice steps (one output pixel per thread)

template <class J>
struct Small2Params {
  J job[kMaxBatch];  // Dec2Job or PairJob
  int h, w, d;
  int ntr, ntc;      // tiles per phase along rows and columns
};

// Folded plane index of lattice coordinates origin + step*(first + k), k in [0, n), for k = lane;
// the index is stored with its mirror flag as ~index when the position lies on a mirrored period.
// Any integer coordinate maps (the extension is periodic with period 2*size).
__device__ __forceinline__ void ext_indices(int* out, int n, int origin, int step, int first, int size) {
  const int k = threadIdx.x & 63;
  if (k < n) {
    const int p = 2 * size;
    int m = (origin + step * (first + k)) % p;
    if (m < 0) m += p;
    out[k] = m >= size ? ~(p - 1 - m) : m;
  }
}
__device__ __forceinline__ int ext_pos(int e) { return e < 0 ? ~e : e; }

// Stage the RS x RS region of lattice positions (ry[i], rx[j]) of plane `in` into LDS rows of RP
// floats.  Every thread issues the same number of loads (the last ones clamped to the final element,
// which is then written twice with the same value), so all of them are in flight together.
template <int RS, int RP>
__device__ __forceinline__ void stage_region(float* dst, const float* in, const int* ry, const int* rx, int w) {
  constexpr int kPer = (RS * RS + kThreads - 1) / kThreads;
  float v[kPer];
  int at[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = min((int)threadIdx.x + k * kThreads, RS * RS - 1);
    const int ri = e / RS, rj = e - ri * RS;
    at[k] = ri * RP + rj;
    v[k] = in[(size_t)ext_pos(ry[ri]) * w + ext_pos(rx[rj])];
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) dst[at[k]] = v[k];
}

// 5-tap sum of v[0], v[step], ... v[4*step] in the order of an unmirrored position, or from v[4*step]
// down to v[0] for a mirrored one: either way the chain adds the taps in the folded position's order.
__device__ __forceinline__ float tap5(const float* v, int step, bool mirrored) {
  float x[5];
#pragma unroll
  for (int b = 0; b < 5; ++b) x[b] = v[b * step];
  float s = 0.f;
#pragma unroll
  for (int b = 0; b < 5; ++b) s = fmaf(c_h5[b], mirrored ? x[4 - b] : x[b], s);
  return s;
}
__device__ __forceinline__ float tap9(const float* v, int step) {
  float s = 0.f;
#pragma unroll
  for (int b = 0; b < 9; ++b) s = fmaf(c_h9[b], v[b * step], s);
  return s;
}

// Scales s and s+1 of a decomposition: w_s, w_{s+1} and (OUTC) C_{s+2} from C_s.  The filters are
// applied separably, row sums first (the one-scale kernel's order): pass 1 forms the 5-tap row sums
// of C_s at every C_{s+1} column and the 9-tap row sums at the tile's columns, pass 2 the column sums
// (C_{s+1} and w_s), pass 3 the row sums of C_{s+1} at spacing 2, pass 4 w_{s+1} and C_{s+2}.
template <bool MASK, bool OUTC>
__global__ __launch_bounds__(kThreads) void k_iuwt_small2_dec(Small2Params<Dec2Job> sp) {
  constexpr int RS = kTile2 + 20, RP = RS + 1;  // staged C_s: lattice -10 .. 25 (rows padded)
  constexpr int IS = kTile2 + 16, IP = IS + 1;  // C_{s+1}: lattice -8 .. 23
  constexpr int N5 = RS * IS;                   // pass 1: 5-tap row sums, every staged row
  constexpr int N9 = (kTile2 + 8) * kTile2;     // pass 1: 9-tap row sums, staged rows 6 .. 29
  constexpr int M9 = IS * kTile2;               // pass 3: 9-tap row sums of C_{s+1}, every row
  constexpr int M5 = (kTile2 + 8) * kTile2;     // pass 3: 5-tap row sums of C_{s+1}, rows 4 .. 27
  static_assert((N5 + N9) % kThreads == 0 && N5 % 64 == 0, "pass 1 splits evenly, wave-uniform");
  static_assert(IS * IS % kThreads == 0 && M9 % 64 == 0, "pass 2 splits evenly; pass 3 wave-uniform");
  static_assert(M9 + M5 <= RS * IP, "pass 3 reuses the pass-1 array");
  __shared__ float cs[RS * RP];
  __shared__ float hs[RS * IP];  // pass 1 5-tap sums [row][column]; pass 3 sums afterwards
  __shared__ float h9[N9];
  __shared__ float c1[IS * IP];
  __shared__ int ry[RS], rx[RS];
  const Dec2Job a = sp.job[blockIdx.z];
  if (a.pred && *a.pred == 0) return;
  const int h = sp.h, w = sp.w, d = sp.d;
  const int pr = blockIdx.y / sp.ntr, i0 = (blockIdx.y - pr * sp.ntr) * kTile2;
  const int pc = blockIdx.x / sp.ntc, j0 = (blockIdx.x - pc * sp.ntc) * kTile2;
  if (threadIdx.x < 64)
    ext_indices(ry, RS, pr, d, i0 - 10, h);
  else if (threadIdx.x < 128)
    ext_indices(rx, RS, pc, d, j0 - 10, w);
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int row = pr + d * (i0 + ty), col = pc + d * (j0 + tx);
  const bool live = row < h && col < w;
  const size_t off = live ? (size_t)row * w + col : 0;
  const bool keep0 = !MASK || a.mask0[off] != 0;
  const bool keep1 = !MASK || a.mask1[off] != 0;
  __syncthreads();
  stage_region<RS, RP>(cs, a.in, ry, rx, w);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < (N5 + N9) / kThreads; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < N5) {
      const int ri = e / IS, v = e - ri * IS;
      hs[ri * IP + v] = tap5(cs + ri * RP + v, 1, rx[v + 2] < 0);
    } else {
      const int f = e - N5, ri = f / kTile2, x = f - ri * kTile2;
      h9[f] = tap9(cs + (ri + 6) * RP + x + 6, 1);
    }
  }
  __syncthreads();
  // C_{s+1} slot (u, v) is centred on staged (u + 2, v + 2); w_s on staged (ty + 10, tx + 10)
#pragma unroll
  for (int k = 0; k < IS * IS / kThreads; ++k) {
    const int e = threadIdx.x + k * kThreads;
    const int u = e / IS, v = e - u * IS;
    c1[u * IP + v] = tap5(hs + u * IP + v, IP, ry[u + 2] < 0);
  }
  if (live) a.out_w0[off] = keep0 ? cs[(ty + 10) * RP + tx + 10] - tap9(h9 + ty * kTile2 + tx, kTile2) : 0.f;
  __syncthreads();
  float* g9 = hs;       // [32][16]: 9-tap sums at spacing 2 of C_{s+1} row r at tile column x
  float* g5 = hs + M9;  // [24][16]: 5-tap sums of C_{s+1} row r + 4
#pragma unroll
  for (int k = 0; k < (M9 + M5 + kThreads - 1) / kThreads; ++k) {
    const int e = threadIdx.x + k * kThreads;
    if (e < M9) {
      const int r = e / kTile2, x = e - r * kTile2;
      g9[e] = tap9(c1 + r * IP + x, 2);
    } else if (OUTC && e < M9 + M5) {
      const int f = e - M9, r = f / kTile2, x = f - r * kTile2;
      g5[f] = tap5(c1 + (r + 4) * IP + x + 4, 2, false);
    }
  }
  __syncthreads();
  if (!live) return;
  // w_{s+1} centred on slot (ty + 8, tx + 8), taps on slot rows ty + 2i; C_{s+2} on rows ty + 4 + 2k
  a.out_w1[off] = keep1 ? c1[(ty + 8) * IP + tx + 8] - tap9(g9 + ty * kTile2 + tx, 2 * kTile2) : 0.f;
  if (OUTC) a.out_c[off] = tap5(g5 + ty * kTile2 + tx, 2 * kTile2, false);
}

// Scales s+1 then s of a recomposition: R_s = h_s T + w_s with T = h_{s+1} R + w_{s+1}.  Pass 1 forms
// the row sums of R at spacing 2 at every T column, pass 2 T, pass 3 the row sums of T at the tile's
// columns, pass 4 R_s.
template <bool MASK>
__global__ __launch_bounds__(kThreads) void k_iuwt_small2_rec(Small2Params<PairJob> sp) {
  constexpr int RS = kTile2 + 12, RP = RS + 1;  // staged R: lattice -6 .. 21
  constexpr int TS = kTile2 + 4, TP = TS + 1;   // T: lattice -2 .. 17
  constexpr int N1 = RS * TS;                   // pass 1 sums
  constexpr int N2 = TS * TS;                   // T slots
  constexpr int N3 = TS * kTile2;               // pass 3 sums
  constexpr int K2 = (N2 + kThreads - 1) / kThreads;
  __shared__ float rs[RS * RP];
  __shared__ float hr[RS * TP];  // pass 1 sums [row][T column]; pass 3 sums afterwards
  __shared__ float tt[TS * TP];
  __shared__ int ry[RS], rx[RS];
  const PairJob a = sp.job[blockIdx.z];
  if (a.pred && *a.pred == 0) return;
  const int h = sp.h, w = sp.w, d = sp.d;
  const int pr = blockIdx.y / sp.ntr, i0 = (blockIdx.y - pr * sp.ntr) * kTile2;
  const int pc = blockIdx.x / sp.ntc, j0 = (blockIdx.x - pc * sp.ntc) * kTile2;
  if (threadIdx.x < 64)
    ext_indices(ry, RS, pr, d, i0 - 6, h);
  else if (threadIdx.x < 128)
    ext_indices(rx, RS, pc, d, j0 - 6, w);
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int row = pr + d * (i0 + ty), col = pc + d * (j0 + tx);
  const bool live = row < h && col < w;
  const size_t off = live ? (size_t)row * w + col : 0;
  float wi0 = a.w_in0[off];
  if (MASK && a.mask0[off] == 0) wi0 = 0.f;
  __syncthreads();
  // the detail term of T slot (u, v) (staged (u + 4, v + 4)) is read at the folded position, where
  // the one-scale launch adds it; the last slots are clamped (a repeated slot writes the same value)
  float wi1[K2];
#pragma unroll
  for (int k = 0; k < K2; ++k) {
    const int e = min((int)threadIdx.x + k * kThreads, N2 - 1);
    const int u = e / TS, v = e - u * TS;
    const size_t o = (size_t)ext_pos(ry[u + 4]) * w + ext_pos(rx[v + 4]);
    wi1[k] = a.w_in1[o];
    if (MASK && a.mask1[o] == 0) wi1[k] = 0.f;
  }
  stage_region<RS, RP>(rs, a.in, ry, rx, w);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < (N1 + kThreads - 1) / kThreads; ++k) {
    const int e = min((int)threadIdx.x + k * kThreads, N1 - 1);
    const int ri = e / TS, v = e - ri * TS;
    hr[ri * TP + v] = tap5(rs + ri * RP + v, 2, rx[v + 4] < 0);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K2; ++k) {
    const int e = min((int)threadIdx.x + k * kThreads, N2 - 1);
    const int u = e / TS, v = e - u * TS;
    tt[u * TP + v] = tap5(hr + u * TP + v, 2 * TP, ry[u + 4] < 0) + wi1[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < (N3 + kThreads - 1) / kThreads; ++k) {
    const int e = min((int)threadIdx.x + k * kThreads, N3 - 1);
    const int r = e / kTile2, x = e - r * kTile2;
    hr[e] = tap5(tt + r * TP + x, 1, false);
  }
  __syncthreads();
  if (!live) return;
  a.out[off] = tap5(hr + ty * kTile2 + tx, kTile2, false) + wi0;
}

template <class J>
Small2Params<J> small2_params(int h, int w, int s) {
  Small2Params<J> sp{};
  sp.h = h;
  sp.w = w;
  sp.d = 1 << s;
  sp.ntr = ceil_div(ceil_div(h, sp.d), kTile2);
  sp.ntc = ceil_div(ceil_div(w, sp.d), kTile2);
  return sp;
}

template <class J>
dim3 small2_grid(const Small2Params<J>& sp, int nb) { return dim3(sp.d * sp.ntc, sp.d * sp.ntr, nb); }

// jobs[0..count) share shape and scale pair; masked and unmasked jobs, and jobs that keep C_{s+2}
// and those that do not, go to separate launches
void small2_dec_batch(const Dec2Job* jobs, int count, int h, int w, int s, hipStream_t st) {
  Small2Params<Dec2Job> sp = small2_params<Dec2Job>(h, w, s);
  for (int kind = 0; kind < 4; ++kind) {
    const bool masked = (kind & 1) != 0, keep_c = (kind & 2) != 0;
    int nb = 0;
    auto flush = [&]() {
      if (nb == 0) return;
      const dim3 g = small2_grid(sp, nb);
      switch (kind) {
        case 0: hipLaunchKernelGGL((k_iuwt_small2_dec<false, false>), g, dim3(kThreads), 0, st, sp); break;
        case 1: hipLaunchKernelGGL((k_iuwt_small2_dec<true, false>), g, dim3(kThreads), 0, st, sp); break;
        case 2: hipLaunchKernelGGL((k_iuwt_small2_dec<false, true>), g, dim3(kThreads), 0, st, sp); break;
        default: hipLaunchKernelGGL((k_iuwt_small2_dec<true, true>), g, dim3(kThreads), 0, st, sp); break;
      }
      PMS_LAUNCH_CHECK();
      nb = 0;
    };
    for (int k = 0; k < count; ++k) {
      if ((jobs[k].mask0 != nullptr) != masked || (jobs[k].out_c != nullptr) != keep_c) continue;
      PMS_REQUIRE(jobs[k].out_w0 && jobs[k].out_w1 && (!masked || jobs[k].mask1),
                  "two-scale decomposition needs both detail planes (and both masks or none)");
      sp.job[nb++] = jobs[k];
      if (nb == kMaxBatch) flush();
    }
    flush();
  }
}

void small2_rec_batch(const PairJob* jobs, int count, int h, int w, int s, hipStream_t st) {
  Small2Params<PairJob> sp = small2_params<PairJob>(h, w, s);
  for (int masked = 0; masked < 2; ++masked) {
    int nb = 0;
    auto flush = [&]() {
      if (nb == 0) return;
      const dim3 g = small2_grid(sp, nb);
      if (masked)
        hipLaunchKernelGGL(k_iuwt_small2_rec<true>, g, dim3(kThreads), 0, st, sp);
      else
        hipLaunchKernelGGL(k_iuwt_small2_rec<false>, g, dim3(kThreads), 0, st, sp);
      PMS_LAUNCH_CHECK();
      nb = 0;
    };
    for (int k = 0; k < count; ++k) {
      if ((jobs[k].mask0 != nullptr) != (masked != 0)) continue;
      PMS_REQUIRE(!masked || jobs[k].mask1, "two-scale recomposition needs both masks or none");
      sp.job[nb++] = jobs[k];
      if (nb == kMaxBatch) flush();
    }
    flush();
  }
}

int walk_flags(const IuwtScaleCall& c) {
  if (c.mode == IUWT_DECOMP)
    return (c.out_c ? F_OUTC : 0) | (c.out_w ? F_OUTW : 0) | (c.mask ? F_MASK : 0) | (c.ref ? F_REF : 0);
  if (c.mode == IUWT_RECOMP) return c.mask ? F_MASK : 0;
  return 0;
}


}  // namespace

// Scales with 2^(s+1) above a side are valid (the reference reads the axis reversed there); they
// run on the direct-stencil kernels of iuwt_any.hip, the row walks take the rest.
void check_iuwt_scale(int h, int w, int scale) {
  PMS_REQUIRE(h >= 2 && w >= 2, "IUWT planes must be at least 2x2");
  PMS_REQUIRE(scale >= 0 && scale < 30, "IUWT scale out of range");
  PMS_REQUIRE((long)h * w < (1L << 29), "IUWT planes are limited to 2^29 pixels (32-bit byte offsets)");
}

bool use_small(int h, int w) { return (long)h * w <= tuning().small_pixels; }

// Kernel family of a batched decomposition or recomposition of `jobs` planes: the direct-stencil
// kernels for planes of at most 256^2 pixels, and for planes of at most 512^2 while the launch holds at
// most PMS_IUWT_SMALL_LAUNCH pixels (default 2^22: 16 planes of 512^2; read per call).  Few mid-sized
// planes leave the row walks with too few workgroups to fill the chip, while a large batch of them
// streams better through the walks (whose planes are bit-identical).
bool use_small_launch(int h, int w, int jobs) {
  const long px = (long)h * w;
  return use_small(h, w) || (px <= 512L * 512 && px * jobs <= (long)env_int("PMS_IUWT_SMALL_LAUNCH", 1 << 22));
}

// Scales (s, s+1) of a plane on the direct-stencil kernels run as one two-scale launch unless
// PMS_IUWT_SMALL2=0 (read per call; the planes are identical either way); scale s+1 must be within
// the row-walk range.
bool small2_applies(bool small, int h, int w, int s) {
  return small && !iuwt_over_long(h, w, s + 1) && env_int("PMS_IUWT_SMALL2", 1) != 0;
}

dim3 small_grid(int h, int w) { return dim3(ceil_div(w, 64), ceil_div(h, 4)); }

static ScaleJob scale_job(const IuwtScaleCall& c) {
  ScaleJob j;
  j.in = c.in;
  j.out_c = c.out_c;
  j.out_w = c.out_w;
  j.w_in = c.w_in;
  j.mask = c.mask;
  j.ref = c.ref;
  j.partial = c.partial;
  j.pred = c.pred;
  return j;
}

int iuwt_scale(const IuwtScaleCall& c, hipStream_t st) { return iuwt_scale_batch(&c, 1, st, use_small(c.h, c.w)); }

// All calls share mode, shape and scale; `small` picks the direct-stencil kernel over the row walk
// (a caller passing a reference plane passes use_small(h, w), the rule iuwt_scale_blocks counts
// partials by).  Returns the blocks (partials) per job.
int iuwt_scale_batch(const IuwtScaleCall* calls, int count, hipStream_t st, bool small) {
  PMS_REQUIRE(count >= 1, "empty IUWT batch");
  const IuwtScaleCall& c0 = calls[0];
  check_iuwt_scale(c0.h, c0.w, c0.scale);
  if (iuwt_over_long(c0.h, c0.w, c0.scale)) return iuwt_gen_scale_batch<float>(calls, count, st);
  for (int k = 1; k < count; ++k)
    PMS_REQUIRE(calls[k].mode == c0.mode && calls[k].h == c0.h && calls[k].w == c0.w && calls[k].scale == c0.scale,
                "IUWT batch mixes modes, shapes or scales");
  const int halo = (c0.mode == IUWT_DECOMP) ? 4 : 2;
  Plan pl{};
  if (!small) {
    pl = make_plan(c0.h, c0.w, c0.scale, halo);
    PMS_REQUIRE((pl.tj + 2 * halo) * (1 << pl.q_log2) <= 8 * kThreads, "IUWT tile too wide");
  }
  ScaleParams sp{};
  sp.h = c0.h;
  sp.w = c0.w;
  sp.d = 1 << c0.scale;
  sp.q_log2 = pl.q_log2;
  sp.tj = pl.tj;
  sp.k_rows = pl.k_rows;
  sp.n_cg = pl.n_cg;
  sp.n_tb = pl.n_tb;
  sp.n_seg = pl.n_seg;
  int blocks = 0;
  if (!small) {
    // one launch per operand layout (flag set), up to kMaxBatch jobs each
    std::vector<int> order(count);
    for (int k = 0; k < count; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return walk_flags(calls[x]) < walk_flags(calls[y]); });
    for (int i0 = 0; i0 < count;) {
      const int fl = walk_flags(calls[order[i0]]);
      int nb = 0;
      while (i0 + nb < count && nb < kMaxBatch && walk_flags(calls[order[i0 + nb]]) == fl) {
        sp.job[nb] = scale_job(calls[order[i0 + nb]]);
        ++nb;
      }
      if (c0.mode == IUWT_DECOMP)
        launch_walk<IUWT_DECOMP>(sp, nb, pl, halo, fl, st);
      else if (c0.mode == IUWT_SMOOTH)
        launch_walk<IUWT_SMOOTH>(sp, nb, pl, halo, fl, st);
      else
        launch_walk<IUWT_RECOMP>(sp, nb, pl, halo, fl, st);
      i0 += nb;
    }
    return pl.blocks;
  }
  for (int j0 = 0; j0 < count; j0 += kMaxBatch) {
    const int nb = std::min(kMaxBatch, count - j0);
    for (int k = 0; k < nb; ++k) sp.job[k] = scale_job(calls[j0 + k]);
    const dim3 g2 = small_grid(c0.h, c0.w);
    const dim3 grid(g2.x, g2.y, nb);
    if (c0.mode == IUWT_DECOMP)
      hipLaunchKernelGGL(k_iuwt_small<IUWT_DECOMP>, grid, dim3(kThreads), 0, st, sp);
    else if (c0.mode == IUWT_SMOOTH)
      hipLaunchKernelGGL(k_iuwt_small<IUWT_SMOOTH>, grid, dim3(kThreads), 0, st, sp);
    else
      hipLaunchKernelGGL(k_iuwt_small<IUWT_RECOMP>, grid, dim3(kThreads), 0, st, sp);
    PMS_LAUNCH_CHECK();
    blocks = (int)(g2.x * g2.y);
  }
  return blocks;
}

int iuwt_scale_blocks(int h, int w, int scale, int mode) {
  if (iuwt_over_long(h, w, scale)) return iuwt_gen_blocks(h, w);
  if (use_small(h, w)) {
    const dim3 g = small_grid(h, w);
    return (int)(g.x * g.y);
  }
  return make_plan(h, w, scale, mode == IUWT_DECOMP ? 4 : 2).blocks;
}

// Full decomposition.  detail: (scale_count - scale_adjust) planes; smoothed
// may be null; work: two planes of scratch.  mask/ref (optional) are
// per-detail-plane arrays laid out like detail.  Returns the number of
// partial sums written (ref != null), in plane order.
long iuwt_decompose(const float* in, int h, int w, int scale_count, int scale_adjust, float* detail,
                    float* smoothed, float* work, const uint8_t* mask, const float* ref,
                    double* partial, const int* pred, hipStream_t st) {
  IuwtDecompJob j{in, scale_count, scale_adjust, detail, smoothed, work, mask, ref, partial, pred, 0};
  iuwt_decompose_batch(&j, 1, h, w, st);
  return j.np;
}

// Scale by scale over all jobs: at scale s a job either pre-smooths (s < scale_adjust) or
// decomposes (scale_adjust <= s < scale_count); jobs of the same kind share one launch.
static void decompose_batch_impl(IuwtDecompJob* jobs, int count, int h, int w, bool small, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  struct State {
    const float* cur;
    int slot;
    int done;  // scales below this one are decomposed
  };
  std::vector<State> state(count);
  int smax = 0;
  for (int k = 0; k < count; ++k) {
    IuwtDecompJob& j = jobs[k];
    PMS_REQUIRE(j.scale_adjust >= 0 && j.scale_count > j.scale_adjust, "need scale_count > scale_adjust >= 0");
    check_iuwt_scale(h, w, j.scale_count - 1);
    state[k] = {j.in, 0, 0};
    j.np = 0;
    smax = std::max(smax, j.scale_count);
  }
  // detail scales in pairs from the finest up where a two-scale launch applies (1), or one scale per
  // launch (0): PMS_IUWT_DEC2 when set, else the autotuner's record, else the default (the planes are
  // identical either way)
  const int fuse = getenv("PMS_IUWT_DEC2") ? env_int("PMS_IUWT_DEC2", kDec2Default)
                                           : (t_active ? t_active->pair : kDec2Default);
  std::vector<IuwtScaleCall> calls;
  std::vector<int> who;
  std::vector<Dec2Job> pairs;
  std::vector<int> pwho;
  calls.reserve(count);
  who.reserve(count);
  for (int s = 0; s < smax; ++s) {
    for (int mode : {IUWT_SMOOTH, IUWT_DECOMP}) {
      calls.clear();
      who.clear();
      pairs.clear();
      pwho.clear();
      Plan pl;
      const bool small_pair = mode == IUWT_DECOMP && small2_applies(small, h, w, s);
      const bool pair_plan = small_pair || (!small && fuse == 1 && mode == IUWT_DECOMP && dec2_plan(h, w, s, &pl));
      for (int k = 0; k < count; ++k) {
        const IuwtDecompJob& j = jobs[k];
        float* ping = j.work + (size_t)state[k].slot * plane;
        IuwtScaleCall c{};
        c.mode = mode;
        c.in = state[k].cur;
        c.h = h;
        c.w = w;
        c.scale = s;
        c.pred = j.pred;
        if (mode == IUWT_SMOOTH) {
          if (s >= j.scale_adjust) continue;
          c.out_c = ping;
        } else {
          if (s < j.scale_adjust || s >= j.scale_count || s < state[k].done) continue;
          const int kk = s - j.scale_adjust;
          if (pair_plan && s + 1 < j.scale_count && kk % 2 == 0 && j.detail && !j.ref) {
            Dec2Job pj{};
            pj.in = state[k].cur;
            pj.out_c = (s + 1 == j.scale_count - 1) ? j.smoothed : ping;
            pj.out_w0 = j.detail + (size_t)kk * plane;
            pj.out_w1 = j.detail + (size_t)(kk + 1) * plane;
            pj.mask0 = j.mask ? j.mask + (size_t)kk * plane : nullptr;
            pj.mask1 = j.mask ? j.mask + (size_t)(kk + 1) * plane : nullptr;
            pj.pred = j.pred;
            pairs.push_back(pj);
            pwho.push_back(k);
            continue;
          }
          const bool last = (s == j.scale_count - 1);
          // the direct-stencil form of an over-long scale needs C even when it is not kept
          c.out_c = (last && (j.smoothed || !iuwt_over_long(h, w, s))) ? j.smoothed : ping;
          c.out_w = j.detail ? j.detail + (size_t)kk * plane : nullptr;
          c.mask = j.mask ? j.mask + (size_t)kk * plane : nullptr;
          c.ref = j.ref ? j.ref + (size_t)kk * plane : nullptr;
          c.partial = j.ref ? j.partial + j.np : nullptr;
        }
        calls.push_back(c);
        who.push_back(k);
      }
      if (!pairs.empty()) {
        if (small_pair)
          small2_dec_batch(pairs.data(), (int)pairs.size(), h, w, s, st);
        else
          dec2_batch(pairs.data(), (int)pairs.size(), h, w, s, pl, st);
        for (int k : pwho) {
          State& sk = state[k];
          sk.done = s + 2;
          if (s + 1 != jobs[k].scale_count - 1) {
            sk.cur = jobs[k].work + (size_t)sk.slot * plane;
            sk.slot ^= 1;
          }
        }
      }
      if (calls.empty()) continue;
      const int nb = iuwt_scale_batch(calls.data(), (int)calls.size(), st, small);
      for (size_t i = 0; i < who.size(); ++i) {
        IuwtDecompJob& j = jobs[who[i]];
        State& sk = state[who[i]];
        if (mode == IUWT_DECOMP && j.ref) j.np += nb;
        if (mode == IUWT_SMOOTH || s != j.scale_count - 1) {
          sk.cur = j.work + (size_t)sk.slot * plane;
          sk.slot ^= 1;
        }
      }
    }
  }
}

long iuwt_decompose_partials(int h, int w, int scale_count, int scale_adjust) {
  long np = 0;
  for (int s = scale_adjust; s < scale_count; ++s) np += iuwt_scale_blocks(h, w, s, IUWT_DECOMP);
  return np;
}

// Full recomposition of nplanes detail planes (scales scale_adjust ..
// scale_adjust+nplanes-1) with optional smoothed base and optional mask on the
// detail planes.  work: two planes of scratch; out: one plane.
void iuwt_recompose(const float* detail, int nplanes, int h, int w, int scale_adjust,
                    const float* smoothed, float* out, float* work, const uint8_t* mask,
                    const int* pred, hipStream_t st) {
  const IuwtRecompJob j{detail, nplanes, scale_adjust, smoothed, out, work, mask, pred};
  iuwt_recompose_batch(&j, 1, h, w, st);
}

// From the coarsest scale down: at scale s a job adds its detail plane (scale_adjust <= s) or only
// smooths (s < scale_adjust).  Pairs of consecutive detail scales run as one two-scale launch where
// the shape allows it (the same planes result).  Each job's steps form a list; round r issues every
// job's r-th step, steps of the same kind and scale sharing one launch.  Without a smoothed base R
// starts at zero, so a job's coarsest step is its (masked) detail plane itself.
static void recompose_batch_impl(const IuwtRecompJob* jobs, int count, int h, int w, bool small, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  enum { OP_RECOMP, OP_SMOOTH, OP_PAIR };
  struct Op {
    int kind, s;  // OP_PAIR covers scales s+1 then s
  };
  struct State {
    const float* cur;
    int slot;
    std::vector<Op> ops;
  };
  std::vector<State> state(count);
  size_t rounds = 0;
  // PMS_IUWT_PAIR=0 keeps every scale in its own launch, 1 pairs detail scales from the coarsest
  // down, 2 from the finest up (read per call, and it overrides the autotuner's record; the results
  // are identical)
  const int pair = getenv("PMS_IUWT_PAIR") ? env_int("PMS_IUWT_PAIR", 2) : (t_active ? t_active->pair : 2);
  for (int k = 0; k < count; ++k) {
    const IuwtRecompJob& j = jobs[k];
    PMS_REQUIRE(j.nplanes >= 1 && j.scale_adjust >= 0, "recomposition needs at least one detail plane");
    const int top = j.nplanes + j.scale_adjust;
    check_iuwt_scale(h, w, top - 1);
    State& sk = state[k];
    sk.cur = j.smoothed;
    sk.slot = 0;
    int s = top - 1;
    if (!j.smoothed) {
      const size_t off = (size_t)(top - 1 - j.scale_adjust) * plane;
      if (!j.mask) {
        sk.cur = j.detail + off;
      } else {
        masked_copy(j.detail + off, j.mask + off, j.work, plane, j.pred, st);
        sk.cur = j.work;
        sk.slot = 1;
      }
      --s;
    }
    Plan pl;
    // pairing from the finest scale up leaves the odd scale out at the top
    bool skip_top = pair == 2 && s >= j.scale_adjust && (s - j.scale_adjust + 1) % 2 == 1;
    while (s >= 0) {
      if (skip_top) {
        skip_top = false;
        sk.ops.push_back({OP_RECOMP, s});
        --s;
        continue;
      }
      if (pair && s - 1 >= j.scale_adjust && (small2_applies(small, h, w, s - 1) || (!small && recomp2_plan(h, w, s - 1, &pl)))) {
        sk.ops.push_back({OP_PAIR, s - 1});
        s -= 2;
      } else {
        sk.ops.push_back({s >= j.scale_adjust ? OP_RECOMP : OP_SMOOTH, s});
        --s;
      }
    }
    if (sk.ops.empty()) copy_plane(sk.cur, j.out, plane, j.pred, st);
    rounds = std::max(rounds, sk.ops.size());
  }
  std::vector<IuwtScaleCall> calls;
  std::vector<PairJob> pairs;
  std::vector<int> who;
  std::vector<char> issued(count);
  for (size_t r = 0; r < rounds; ++r) {
    std::fill(issued.begin(), issued.end(), 0);
    for (int k0 = 0; k0 < count; ++k0) {
      if (issued[k0] || r >= state[k0].ops.size()) continue;
      const Op op0 = state[k0].ops[r];
      calls.clear();
      pairs.clear();
      who.clear();
      for (int k = k0; k < count; ++k) {
        State& sk = state[k];
        if (issued[k] || r >= sk.ops.size() || sk.ops[r].kind != op0.kind || sk.ops[r].s != op0.s) continue;
        const IuwtRecompJob& j = jobs[k];
        float* out = r + 1 == sk.ops.size() ? j.out : j.work + (size_t)sk.slot * plane;
        if (op0.kind == OP_PAIR) {
          const size_t o0 = (size_t)(op0.s - j.scale_adjust) * plane;
          PairJob pj{};
          pj.in = sk.cur;
          pj.out = out;
          pj.w_in0 = j.detail + o0;
          pj.w_in1 = j.detail + o0 + plane;
          pj.mask0 = j.mask ? j.mask + o0 : nullptr;
          pj.mask1 = j.mask ? j.mask + o0 + plane : nullptr;
          pj.pred = j.pred;
          pairs.push_back(pj);
        } else {
          IuwtScaleCall c{};
          c.mode = op0.kind == OP_RECOMP ? IUWT_RECOMP : IUWT_SMOOTH;
          c.in = sk.cur;
          c.out_c = out;
          if (op0.kind == OP_RECOMP) {
            c.w_in = j.detail + (size_t)(op0.s - j.scale_adjust) * plane;
            c.mask = j.mask ? j.mask + (size_t)(op0.s - j.scale_adjust) * plane : nullptr;
          }
          c.h = h;
          c.w = w;
          c.scale = op0.s;
          c.pred = j.pred;
          calls.push_back(c);
        }
        who.push_back(k);
        issued[k] = 1;
        sk.cur = out;
        sk.slot ^= 1;
      }
      if (op0.kind == OP_PAIR) {
        Plan pl;
        if (small2_applies(small, h, w, op0.s)) {
          small2_rec_batch(pairs.data(), (int)pairs.size(), h, w, op0.s, st);
        } else {
          PMS_REQUIRE(recomp2_plan(h, w, op0.s, &pl), "two-scale recomposition plan changed");
          recomp2_batch(pairs.data(), (int)pairs.size(), h, w, op0.s, pl, st);
        }
      } else {
        iuwt_scale_batch(calls.data(), (int)calls.size(), st, small);
      }
    }
  }
}

// ---- autotuner ------------------------------------------------------------------------------
// The row walks' launch shape (target workgroup count) and the recomposition's scale pairing change
// only how the work is cut, never a pixel's arithmetic, so every candidate gives the same planes.
// On the first call of a (shape, scales, operands, batch size) signature the candidates run on the
// call's own buffers (pure functions of their inputs) and are timed with events; the fastest is
// kept in a per-process table (mutex-guarded) and, with PMS_AUTOTUNE_CACHE=<file>, in a JSON file
// keyed by device architecture and signature that later processes load.  PMS_AUTOTUNE=0 keeps the
// defaults.  Calls that reduce against a reference plane keep the default shape (their partial
// sums are sized by it), and calls under stream capture or with predicated jobs use the table or
// the defaults without timing.
namespace {

std::mutex g_tune_mu;
std::map<std::string, TuneParams> g_tune_table;
bool g_tune_loaded = false;

const char* tune_cache_path() { return getenv("PMS_AUTOTUNE_CACHE"); }

// architecture name of the current device (looked up once per device; caller holds g_tune_mu)
std::string device_arch_locked() {
  static std::map<int, std::string> names;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return "unknown";
  auto it = names.find(dev);
  if (it != names.end()) return it->second;
  hipDeviceProp_t prop;
  std::string a = hipGetDeviceProperties(&prop, dev) == hipSuccess ? std::string(prop.gcnArchName) : "unknown";
  const size_t colon = a.find(':');
  if (colon != std::string::npos) a = a.substr(0, colon);
  names[dev] = a;
  return a;
}

// file format: a JSON object, one entry per line: "<arch>|<signature>": {"target_blocks": N, "pair": P}
void load_tune_cache_locked() {
  g_tune_loaded = true;
  const char* path = tune_cache_path();
  if (!path) return;
  FILE* f = fopen(path, "r");
  if (!f) return;
  char line[1024];
  while (fgets(line, sizeof(line), f)) {
    char key[768];
    int tb = 0, pr = 0;
    if (sscanf(line, " \"%767[^\"]\": {\"target_blocks\": %d, \"pair\": %d", key, &tb, &pr) == 3 && tb > 0)
      g_tune_table[key] = TuneParams{tb, pr};
  }
  fclose(f);
}

void save_tune_cache_locked() {
  const char* path = tune_cache_path();
  if (!path) return;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;
  fprintf(f, "{\n");
  size_t i = 0;
  for (const auto& kv : g_tune_table)
    fprintf(f, "\"%s\": {\"target_blocks\": %d, \"pair\": %d}%s\n", kv.first.c_str(), kv.second.target_blocks,
            kv.second.pair, ++i < g_tune_table.size() ? "," : "");
  fprintf(f, "}\n");
  fclose(f);
  (void)std::rename(tmp.c_str(), path);
}

int batch_bucket(int count) {
  int b = 1;
  while (b < count) b <<= 1;
  return b;
}

struct ActiveGuard {
  explicit ActiveGuard(const TuneParams* p) { t_active = p; }
  ~ActiveGuard() { t_active = nullptr; }
};

// Runs `run` with tuned parameters for signature `sig`; `can_time` says whether the call may be
// repeated for timing (no stream capture, no predicates, no input/output aliasing).
template <class Run>
void autotuned(const std::string& sig, const std::vector<TuneParams>& cands, bool can_time, hipStream_t st,
               Run run) {
  const char* en = getenv("PMS_AUTOTUNE");
  if ((en && en[0] == '0') || cands.size() < 2) {
    run();
    return;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) can_time = false;
  std::string key;
  TuneParams tp{};
  bool have = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    key = device_arch_locked() + "|" + sig;
    if (!g_tune_loaded) load_tune_cache_locked();
    auto it = g_tune_table.find(key);
    if (it != g_tune_table.end()) {
      tp = it->second;
      have = true;
    }
  }
  if (!have && can_time) {
    hipEvent_t e0, e1;
    PMS_HIP(hipEventCreate(&e0));
    PMS_HIP(hipEventCreate(&e1));
    float best = INFINITY;
    for (const TuneParams& c : cands) {
      ActiveGuard g(&c);
      run();  // warm (caches, code objects)
      for (int rep = 0; rep < 2; ++rep) {
        PMS_HIP(hipEventRecord(e0, st));
        run();
        PMS_HIP(hipEventRecord(e1, st));
        PMS_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PMS_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          tp = c;
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    have = true;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_table[key] = tp;
    save_tune_cache_locked();
  }
  if (have) {
    ActiveGuard g(&tp);
    run();
  } else {
    run();
  }
}

std::vector<TuneParams> walk_candidates(int h, int w, bool recomp) {
  std::vector<TuneParams> c;
  if ((long)h * w <= tuning().small_pixels) return c;  // direct-stencil kernel: nothing to tune
  const bool fixed_blocks = getenv("PMS_IUWT_BLOCKS") != nullptr;
  const bool fixed_pair = getenv("PMS_IUWT_PAIR") != nullptr;
  const int pair_default = env_int("PMS_IUWT_PAIR", 2);
  std::vector<int> blocks =
      fixed_blocks ? std::vector<int>{tuning().target_blocks} : std::vector<int>{512, 1024, 2048, 4096};
  std::vector<int> pairs = (recomp && !fixed_pair) ? std::vector<int>{2, 0} : std::vector<int>{pair_default};
  // for a decomposition the field selects two-scale launches (1) or one scale per launch (0); it is
  // not timed (the two-scale walk measured slower at every tested shape), PMS_IUWT_DEC2 sets it
  if (!recomp) pairs = std::vector<int>{env_int("PMS_IUWT_DEC2", kDec2Default)};
  for (int p : pairs)
    for (int b : blocks) c.push_back(TuneParams{b, p});
  return c;
}

bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  if (!a || !b) return false;
  const char* pa = static_cast<const char*>(a);
  const char* pb = static_cast<const char*>(b);
  return pa < pb + nb && pb < pa + na;
}

}  // namespace

void iuwt_decompose_batch(IuwtDecompJob* jobs, int count, int h, int w, hipStream_t st) {
  const size_t pb = (size_t)h * w * sizeof(float);
  bool can_time = true, has_ref = false;
  std::string sig = "decompose " + std::to_string(h) + "x" + std::to_string(w) + " batch " +
                    std::to_string(batch_bucket(count));
  for (int k = 0; k < count; ++k) {
    const IuwtDecompJob& j = jobs[k];
    has_ref = has_ref || j.ref;
    if (j.pred) can_time = false;
    const size_t nd = (size_t)std::max(0, j.scale_count - j.scale_adjust) * pb;
    if (overlaps(j.in, pb, j.detail, nd) || overlaps(j.in, pb, j.smoothed, pb) || overlaps(j.in, pb, j.work, 2 * pb))
      can_time = false;
    if (k == 0)
      sig += " S" + std::to_string(j.scale_count) + "/" + std::to_string(j.scale_adjust) + (j.detail ? " w" : "") +
             (j.smoothed ? " c" : "") + (j.mask ? " m" : "");
  }
  // per-block partials against a reference plane are counted by the plane-size rule alone
  const bool small = has_ref ? use_small(h, w) : use_small_launch(h, w, count);
  if (has_ref || small) {
    decompose_batch_impl(jobs, count, h, w, small, st);
    return;
  }
  autotuned(sig, walk_candidates(h, w, false), can_time, st,
            [&]() { decompose_batch_impl(jobs, count, h, w, false, st); });
}

void iuwt_recompose_batch(const IuwtRecompJob* jobs, int count, int h, int w, hipStream_t st) {
  const size_t pb = (size_t)h * w * sizeof(float);
  bool can_time = true;
  std::string sig = "recompose " + std::to_string(h) + "x" + std::to_string(w) + " batch " +
                    std::to_string(batch_bucket(count));
  for (int k = 0; k < count; ++k) {
    const IuwtRecompJob& j = jobs[k];
    if (j.pred) can_time = false;
    const size_t nd = (size_t)std::max(0, j.nplanes) * pb;
    if (overlaps(j.out, pb, j.detail, nd) || overlaps(j.out, pb, j.smoothed, pb) ||
        overlaps(j.work, 2 * pb, j.detail, nd) || overlaps(j.work, 2 * pb, j.smoothed, pb))
      can_time = false;
    if (k == 0)
      sig += " S" + std::to_string(j.nplanes) + "/" + std::to_string(j.scale_adjust) + (j.smoothed ? " c" : "") +
             (j.mask ? " m" : "");
  }
  if (use_small_launch(h, w, count)) {
    recompose_batch_impl(jobs, count, h, w, true, st);
    return;
  }
  autotuned(sig, walk_candidates(h, w, true), can_time, st,
            [&]() { recompose_batch_impl(jobs, count, h, w, false, st); });
}

}  // namespace pms
