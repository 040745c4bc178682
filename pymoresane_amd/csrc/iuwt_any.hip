// Direct-stencil IUWT for gfx950 with the reference's exact tap semantics at every scale, templated
// on the arithmetic type.
//
// Semantics (pymoresane/iuwt.py:66-191): one smoothing is the separable B3 filter
// [1,4,6,4,1]/16 with taps at offsets 0, +-2^s, +-2^(s+1) on both axes.  A tap offset o on an
// axis of length n reads
//   * the half-sample symmetric extension (... b a | a b ... y z | z y ...) when |o| <= n: the
//     reference's slices reflect once there;
//   * the whole axis reversed (input n-1-i for output i) when |o| > n: numpy clips both slices of
//     that tap (iuwt.py:166-187), which happens when a scale count chosen for the whole image is
//     applied to a smaller subregion.
// Decomposition per scale: C = h C0, w = C0 - h C (iuwt.py:102-106); recomposition R <- h R + w
// from coarse to fine (iuwt.py:142-147).
//
// Kernel: one thread per output pixel reads its 5x5 dilated neighbourhood (row sums first, then
// the vertical combination, as the row-walk kernels do).  A decomposition scale is two launches:
// SMOOTH writes C, DETAIL reads C0 at the centre and C around it and writes w (optionally masked,
// or reduced against a reference plane).  The float32 transforms (iuwt.hip) use this for the
// scales beyond the row walks' range; the float64 transforms use it at every scale.
#include <algorithm>
#include <vector>

#include "pms_internal.h"

namespace pms {
namespace {

constexpr int kGenThreads = 256;
enum GenMode { GEN_SMOOTH = 0, GEN_DETAIL = 1, GEN_RECOMP = 2 };

template <class T>
struct GenJob {
  const T* in;        // SMOOTH/RECOMP: the plane smoothed; DETAIL: C0 (centre term)
  const T* aux;       // DETAIL: C, the plane smoothed
  T* out;             // SMOOTH: C; RECOMP: R'; DETAIL: w (null = not stored)
  const T* w_in;      // RECOMP: detail plane added after smoothing
  const uint8_t* mask;
  const T* ref;       // DETAIL: reduce sum((ref - w)^2) per block (null = off)
  double* partial;
  const int* pred;
};

template <class T>
struct GenParams {
  GenJob<T> job[kMaxBatch];
  int h, w, d;
};

// input index of the tap at offset off for output i on an axis of length n
__device__ __forceinline__ int tap_idx(int i, int off, int n) {
  if (off > n || off < -n) return n - 1 - i;
  return fold_idx(i + off, n);
}

template <class T, int MODE>
__global__ __launch_bounds__(kGenThreads) void k_iuwt_gen(GenParams<T> sp) {
  const GenJob<T> a = sp.job[blockIdx.z];
  if (a.pred && *a.pred == 0) return;
  const int w = sp.w, h = sp.h, d = sp.d;
  const int col0 = blockIdx.x * 64 + (threadIdx.x & 63);
  const int row0 = blockIdx.y * 4 + (threadIdx.x >> 6);
  const bool live = col0 < w && row0 < h;
  const int col = live ? col0 : 0;
  const int row = live ? row0 : 0;
  const T h5[5] = {(T)0.0625, (T)0.25, (T)0.375, (T)0.25, (T)0.0625};
  const T* __restrict__ src = MODE == GEN_DETAIL ? a.aux : a.in;
  int ci[5];
#pragma unroll
  for (int b = 0; b < 5; ++b) ci[b] = tap_idx(col, (b - 2) * d, w);
  T acc = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const T* r = src + (size_t)tap_idx(row, (i - 2) * d, h) * w;
    T sr = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) sr = fma(h5[b], r[ci[b]], sr);
    acc = fma(h5[i], sr, acc);
  }
  const size_t off = (size_t)row * w + col;
  if (MODE == GEN_DETAIL) {
    T wv = a.in[off] - acc;
    if (a.mask && a.mask[off] == 0) wv = 0;
    if (live && a.out) a.out[off] = wv;
    if (a.ref) {
      double e = 0.0;
      if (live) {
        const double df = (double)a.ref[off] - (double)wv;
        e = df * df;
      }
      e = block_sum_f64(e);
      if (threadIdx.x == 0) a.partial[blockIdx.y * gridDim.x + blockIdx.x] = e;
    }
    return;
  }
  if (!live) return;
  if (MODE == GEN_SMOOTH) {
    a.out[off] = acc;
  } else {
    T wi = a.w_in[off];
    if (a.mask && a.mask[off] == 0) wi = 0;
    a.out[off] = acc + wi;
  }
}

template <class T>
__global__ __launch_bounds__(kGenThreads) void k_gen_masked_copy(const T* __restrict__ src,
                                                                 const uint8_t* __restrict__ m, T* __restrict__ dst,
                                                                 size_t n, const int* pred) {
  if (pred && *pred == 0) return;
  for (size_t i = blockIdx.x * (size_t)kGenThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kGenThreads)
    dst[i] = (m == nullptr || m[i]) ? src[i] : (T)0;
}

dim3 gen_grid(int h, int w) { return dim3(ceil_div(w, 64), ceil_div(h, 4)); }

template <class T, int MODE>
void launch_gen(std::vector<GenJob<T>>& jobs, int h, int w, int d, hipStream_t st) {
  for (size_t j0 = 0; j0 < jobs.size(); j0 += kMaxBatch) {
    const int nb = (int)std::min(jobs.size() - j0, (size_t)kMaxBatch);
    GenParams<T> sp{};
    for (int k = 0; k < nb; ++k) sp.job[k] = jobs[j0 + k];
    sp.h = h;
    sp.w = w;
    sp.d = d;
    const dim3 g2 = gen_grid(h, w);
    hipLaunchKernelGGL((k_iuwt_gen<T, MODE>), dim3(g2.x, g2.y, nb), dim3(kGenThreads), 0, st, sp);
    PMS_LAUNCH_CHECK();
  }
}

template <class T>
void gen_masked_copy(const T* src, const uint8_t* mask, T* dst, size_t n, const int* pred, hipStream_t st) {
  size_t b = (n + kGenThreads * 4 - 1) / (kGenThreads * 4);
  b = std::max<size_t>(1, std::min<size_t>(b, 4096));
  hipLaunchKernelGGL(k_gen_masked_copy<T>, dim3((unsigned)b), dim3(kGenThreads), 0, st, src, mask, dst, n, pred);
  PMS_LAUNCH_CHECK();
}

void check_gen_plane(int h, int w, int scale) {
  PMS_REQUIRE(h >= 2 && w >= 2, "IUWT planes must be at least 2x2");
  PMS_REQUIRE(scale >= 0 && scale < 30, "IUWT scale out of range");
  PMS_REQUIRE((long)h * w < (1L << 31), "IUWT planes are limited to 2^31 pixels");
}

}  // namespace

bool iuwt_over_long(int h, int w, int scale) { return (2L << scale) > h || (2L << scale) > w; }

int iuwt_gen_blocks(int h, int w) {
  const dim3 g = gen_grid(h, w);
  return (int)(g.x * g.y);
}

// Calls of one mode, shape and scale.  IUWT_DECOMP calls need out_c (the smoothed plane C).
template <class T>
int iuwt_gen_scale_batch(const IuwtScaleCallT<T>* calls, int count, hipStream_t st) {
  PMS_REQUIRE(count >= 1, "empty IUWT batch");
  const IuwtScaleCallT<T>& c0 = calls[0];
  check_gen_plane(c0.h, c0.w, c0.scale);
  std::vector<GenJob<T>> a, b;
  for (int k = 0; k < count; ++k) {
    const IuwtScaleCallT<T>& c = calls[k];
    PMS_REQUIRE(c.mode == c0.mode && c.h == c0.h && c.w == c0.w && c.scale == c0.scale,
                "IUWT batch mixes modes, shapes or scales");
    if (c.mode == IUWT_DECOMP) {
      PMS_REQUIRE(c.out_c != nullptr, "a direct-stencil decomposition scale needs a plane for C");
      a.push_back(GenJob<T>{c.in, nullptr, c.out_c, nullptr, nullptr, nullptr, nullptr, c.pred});
      b.push_back(GenJob<T>{c.in, c.out_c, c.out_w, nullptr, c.mask, c.ref, c.partial, c.pred});
    } else if (c.mode == IUWT_SMOOTH) {
      a.push_back(GenJob<T>{c.in, nullptr, c.out_c, nullptr, nullptr, nullptr, nullptr, c.pred});
    } else {
      a.push_back(GenJob<T>{c.in, nullptr, c.out_c, c.w_in, c.mask, nullptr, nullptr, c.pred});
    }
  }
  const int d = 1 << c0.scale;
  if (c0.mode == IUWT_RECOMP) {
    launch_gen<T, GEN_RECOMP>(a, c0.h, c0.w, d, st);
  } else {
    launch_gen<T, GEN_SMOOTH>(a, c0.h, c0.w, d, st);
    if (c0.mode == IUWT_DECOMP) launch_gen<T, GEN_DETAIL>(b, c0.h, c0.w, d, st);
  }
  return iuwt_gen_blocks(c0.h, c0.w);
}

template int iuwt_gen_scale_batch<float>(const IuwtScaleCallT<float>*, int, hipStream_t);
template int iuwt_gen_scale_batch<double>(const IuwtScaleCallT<double>*, int, hipStream_t);

// float64 decomposition: scale by scale over all jobs (pre-smoothing below scale_adjust, then
// decomposition), ping-ponging C through the job's two work planes.
void iuwt_decompose_batch(IuwtDecompJobT<double>* jobs, int count, int h, int w, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  std::vector<const double*> cur(count);
  std::vector<int> slot(count, 0);
  int smax = 0;
  for (int k = 0; k < count; ++k) {
    IuwtDecompJobT<double>& j = jobs[k];
    PMS_REQUIRE(j.scale_adjust >= 0 && j.scale_count > j.scale_adjust, "need scale_count > scale_adjust >= 0");
    check_gen_plane(h, w, j.scale_count - 1);
    cur[k] = j.in;
    j.np = 0;
    smax = std::max(smax, j.scale_count);
  }
  std::vector<IuwtScaleCallT<double>> calls;
  std::vector<int> who;
  for (int s = 0; s < smax; ++s) {
    for (int mode : {IUWT_SMOOTH, IUWT_DECOMP}) {
      calls.clear();
      who.clear();
      for (int k = 0; k < count; ++k) {
        const IuwtDecompJobT<double>& j = jobs[k];
        double* ping = j.work + (size_t)slot[k] * plane;
        IuwtScaleCallT<double> c{};
        c.mode = mode;
        c.in = cur[k];
        c.h = h;
        c.w = w;
        c.scale = s;
        c.pred = j.pred;
        if (mode == IUWT_SMOOTH) {
          if (s >= j.scale_adjust) continue;
          c.out_c = ping;
        } else {
          if (s < j.scale_adjust || s >= j.scale_count) continue;
          const int kk = s - j.scale_adjust;
          const bool last = s == j.scale_count - 1;
          c.out_c = (last && j.smoothed) ? j.smoothed : ping;
          c.out_w = j.detail ? j.detail + (size_t)kk * plane : nullptr;
          c.mask = j.mask ? j.mask + (size_t)kk * plane : nullptr;
          c.ref = j.ref ? j.ref + (size_t)kk * plane : nullptr;
          c.partial = j.ref ? j.partial + j.np : nullptr;
        }
        calls.push_back(c);
        who.push_back(k);
      }
      if (calls.empty()) continue;
      const int nb = iuwt_gen_scale_batch<double>(calls.data(), (int)calls.size(), st);
      for (size_t i = 0; i < who.size(); ++i) {
        const int k = who[i];
        IuwtDecompJobT<double>& j = jobs[k];
        if (mode == IUWT_DECOMP && j.ref) j.np += nb;
        cur[k] = calls[i].out_c;
        if (calls[i].out_c != j.smoothed) slot[k] ^= 1;
      }
    }
  }
}

long iuwt_decompose_partials_f64(int h, int w, int scale_count, int scale_adjust) {
  return (long)(scale_count - scale_adjust) * iuwt_gen_blocks(h, w);
}

// float64 recomposition from the coarsest kept scale down (one launch per scale and job group)
void iuwt_recompose_batch(const IuwtRecompJobT<double>* jobs, int count, int h, int w, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  struct State {
    const double* cur;
    int slot;
    int s;  // next scale to apply (counts down to -1)
  };
  std::vector<State> state(count);
  int smax = -1;
  for (int k = 0; k < count; ++k) {
    const IuwtRecompJobT<double>& j = jobs[k];
    PMS_REQUIRE(j.nplanes >= 1 && j.scale_adjust >= 0, "recomposition needs at least one detail plane");
    const int top = j.nplanes + j.scale_adjust;
    check_gen_plane(h, w, top - 1);
    State& sk = state[k];
    sk.cur = j.smoothed;
    sk.slot = 0;
    sk.s = top - 1;
    if (!j.smoothed) {
      // R starts at zero: the coarsest step is the (masked) detail plane itself
      const size_t off = (size_t)(top - 1 - j.scale_adjust) * plane;
      double* dst = j.work;
      if (top - 1 == 0) dst = j.out;  // one plane, no smoothing left
      gen_masked_copy<double>(j.detail + off, j.mask ? j.mask + off : nullptr, dst, plane, j.pred, st);
      sk.cur = dst;
      sk.slot = 1;
      sk.s = top - 2;
    }
    smax = std::max(smax, sk.s);
  }
  std::vector<IuwtScaleCallT<double>> calls;
  std::vector<int> who;
  for (int s = smax; s >= 0; --s) {
    for (int mode : {IUWT_RECOMP, IUWT_SMOOTH}) {
      calls.clear();
      who.clear();
      for (int k = 0; k < count; ++k) {
        const IuwtRecompJobT<double>& j = jobs[k];
        State& sk = state[k];
        if (sk.s != s) continue;
        const bool detail = s >= j.scale_adjust;
        if ((mode == IUWT_RECOMP) != detail) continue;
        IuwtScaleCallT<double> c{};
        c.mode = mode;
        c.in = sk.cur;
        c.out_c = s == 0 ? j.out : j.work + (size_t)sk.slot * plane;
        if (detail) {
          const size_t off = (size_t)(s - j.scale_adjust) * plane;
          c.w_in = j.detail + off;
          c.mask = j.mask ? j.mask + off : nullptr;
        }
        c.h = h;
        c.w = w;
        c.scale = s;
        c.pred = j.pred;
        calls.push_back(c);
        who.push_back(k);
      }
      if (calls.empty()) continue;
      iuwt_gen_scale_batch<double>(calls.data(), (int)calls.size(), st);
      for (size_t i = 0; i < who.size(); ++i) {
        State& sk = state[who[i]];
        sk.cur = calls[i].out_c;
        sk.slot ^= 1;
        sk.s -= 1;
      }
    }
  }
  // a job whose only plane was scale 0 with a smoothed base still needs its steps: covered above;
  // a job with a smoothed base and nothing to apply cannot occur (nplanes >= 1)
}

}  // namespace pms
