// Noise estimate, thresholding, source extraction and the vector/scalar
// kernels of the conjugate-gradient minor loop, for gfx950.
//
// Semantics follow pymoresane/iuwt_toolbox.py:
//   estimate_threshold (17-48) : median(|w| over the noise mask) / 0.6745,
//                                exact: radix select on the float bits of |w|
//                                (non-negative floats order like unsigned ints),
//                                np.median's mean of the two middle values for
//                                even counts;
//   apply_threshold (50-75)    : w where |w| > sigma * t (strict), else 0;
//   source_extraction (97-154) : 8-connected labels of non-zero pixels by
//                                union-find with atomicMin links (every root is
//                                its component's smallest raster index, so the
//                                partition and the raster-order label numbering
//                                are exactly scipy.ndimage.label's), then the
//                                coarse-to-fine keep rule on the device;
//   snr_ratio (292-306)        : float64 sums of squares.
// All reductions are deterministic: fixed grids, per-block partials in a
// fixed order, integer atomics only.
#include <algorithm>
#include <map>
#include <mutex>

#include "pms_internal.h"

namespace pms {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ bool pred_off(const int* pred) { return pred && *pred == 0; }

// ----------------------------------------------------------------------------- copies
// Streaming copy/fill: 16-byte accesses, four per thread in flight, for the 16-byte-aligned bulk
// (the host checks alignment); the last n % 4 elements elementwise.
__global__ __launch_bounds__(kT) void k_copy(const float* __restrict__ src, float* __restrict__ dst, size_t n,
                                             const int* pred) {
  if (pred_off(pred)) return;
  const size_t n4 = n >> 2;
  const float4* __restrict__ s4 = reinterpret_cast<const float4*>(src);
  float4* __restrict__ d4 = reinterpret_cast<float4*>(dst);
  const size_t stride = (size_t)gridDim.x * kT;
  size_t i = blockIdx.x * (size_t)kT + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 a = s4[i], b = s4[i + stride], c = s4[i + 2 * stride], d = s4[i + 3 * stride];
    d4[i] = a;
    d4[i + stride] = b;
    d4[i + 2 * stride] = c;
    d4[i + 3 * stride] = d;
  }
  for (; i < n4; i += stride) d4[i] = s4[i];
  for (size_t t = (n4 << 2) + blockIdx.x * (size_t)kT + threadIdx.x; t < n; t += stride) dst[t] = src[t];
}

template <class T>
__global__ void k_copy1(const T* __restrict__ src, T* __restrict__ dst, size_t n, const int* pred) {
  if (pred_off(pred)) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// n elements of a 4-byte type, p 16-byte aligned
template <class T>
__global__ __launch_bounds__(kT) void k_fill16(T* __restrict__ p, size_t n, T v) {
  const size_t n4 = n >> 2;
  uint4 w;
  w.x = w.y = w.z = w.w = *reinterpret_cast<const uint32_t*>(&v);
  uint4* __restrict__ p4 = reinterpret_cast<uint4*>(p);
  const size_t stride = (size_t)gridDim.x * kT;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n4; i += stride) p4[i] = w;
  for (size_t t = (n4 << 2) + blockIdx.x * (size_t)kT + threadIdx.x; t < n; t += stride) p[t] = v;
}

template <class T>
__global__ void k_masked_copy(const T* __restrict__ src, const uint8_t* __restrict__ m, T* __restrict__ dst, size_t n,
                              const int* pred) {
  if (pred_off(pred)) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = m[i] ? src[i] : (T)0;
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// dst (n x n) = src window [r0, r0+n) x [c0, c0+n) of a (ld-wide) plane, optional multiply by mask window
template <class T, class TIN>
__global__ void k_window(const TIN* __restrict__ src, int ld, int r0, int c0, int n, const T* __restrict__ mul,
                         T* __restrict__ dst) {
  const size_t tot = (size_t)n * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / n), c = (int)(i % n);
    const size_t s = (size_t)(r0 + r) * ld + (c0 + c);
    T v = (T)src[s];
    if (mul) v *= mul[s];
    dst[i] = v;
  }
}

// ----------------------------------------------------------------------------- MAD median
struct Rect {
  int r0, r1, c0, c1;  // [r0, r1) x [c0, c1)
};

__device__ __forceinline__ bool in_noise_mask(int r, int c, Rect inc, Rect exc) {
  const bool in = r >= inc.r0 && r < inc.r1 && c >= inc.c0 && c < inc.c1;
  const bool ex = r >= exc.r0 && r < exc.r1 && c >= exc.c0 && c < exc.c1;
  return in && !ex;
}

// Sort keys of |w|: the bits of a non-negative float (or double) order like an unsigned integer.
template <class T>
struct MadKey;
template <>
struct MadKey<float> {
  using K = uint32_t;
  __device__ static K key(float v) { return __float_as_uint(v) & 0x7fffffffu; }
  __device__ static float val(K k) { return __uint_as_float(k); }
};
template <>
struct MadKey<double> {
  using K = unsigned long long;
  __device__ static K key(double v) { return (K)__double_as_longlong(v) & 0x7fffffffffffffffull; }
  __device__ static double val(K k) { return __longlong_as_double((long long)k); }
};

// per-plane select state (32 bytes for either key width)
template <class K>
struct alignas(16) SelStateT {
  K prefix;          // fixed high bits of the target key
  K hi_key;          // min key above the low median (for even counts)
  uint32_t rank;     // remaining rank of the low median inside the prefix bucket
  uint32_t eq;       // count of keys equal to the final low median
  uint32_t first;    // rank of the first key equal to the low median
  uint32_t n;        // number of keys
};
constexpr size_t kSelBytes = 32;
static_assert(sizeof(SelStateT<uint32_t>) <= kSelBytes && sizeof(SelStateT<unsigned long long>) <= kSelBytes,
              "select state exceeds its slot");

constexpr int kBins = 2048;
constexpr int kMaxPlanes = 16;  // plane stacks handled per call (scale counts up to 16)

// Histogram of one digit of the keys whose higher digits equal the prefix found so far.  With
// `above` set (the last digit), keys whose higher digits exceed the prefix are all larger than the
// low median: the smallest of them goes to hi_key (atomic min), where the last select finds it if
// the low median's bucket holds no larger key (the successor of even counts, without a pass of its own).
template <class T, bool MASKED>
__global__ __launch_bounds__(kT) void k_mad_hist(const T* __restrict__ planes, size_t plane, int w, Rect inc,
                                                 Rect exc, int shift, int nbits, int prefix_shift, int above,
                                                 SelStateT<typename MadKey<T>::K>* __restrict__ st,
                                                 uint32_t* __restrict__ hist, const int* pred) {
  using K = typename MadKey<T>::K;
  if (pred_off(pred)) return;
  __shared__ uint32_t h[kBins];
  const int s = blockIdx.y;
  for (int i = threadIdx.x; i < kBins; i += kT) h[i] = 0;
  __syncthreads();
  const T* x = planes + (size_t)s * plane;
  const K pre = st[s].prefix;
  const K dmask = ((K)1 << nbits) - 1;
  const K none = ~(K)0;
  K best = none;
  auto add = [&](T v) {
    const K key = MadKey<T>::key(v);
    if (prefix_shift < (int)(8 * sizeof(K))) {
      const K kp = key >> prefix_shift, pp = pre >> prefix_shift;
      if (kp != pp) {
        if (above && kp > pp && key < best) best = key;
        return;
      }
    }
    atomicAdd(&h[(uint32_t)((key >> shift) & dmask)], 1u);
  };
  const size_t stride = (size_t)gridDim.x * kT;
  bool vec = false;
  if constexpr (!MASKED && sizeof(T) == 4) vec = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  if constexpr (sizeof(T) == 4) {
    if (vec) {
    // eight keys per thread and iteration, loaded as two float4 before any histogram update, so a
    // thread has two loads in flight instead of one per atomic (the counts do not depend on order)
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const size_t n4 = plane >> 2;
    size_t i = blockIdx.x * (size_t)kT + threadIdx.x;
    for (; i + stride < n4; i += 2 * stride) {
      const float4 a = x4[i], b = x4[i + stride];
      add(a.x), add(a.y), add(a.z), add(a.w), add(b.x), add(b.y), add(b.z), add(b.w);
    }
    if (i < n4) {
      const float4 a = x4[i];
      add(a.x), add(a.y), add(a.z), add(a.w);
    }
    for (size_t j = 4 * n4 + blockIdx.x * (size_t)kT + threadIdx.x; j < plane; j += stride) add(x[j]);
    }
  }
  if (!vec) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += stride) {
      if (MASKED) {
        const int r = (int)(i / w), c = (int)(i % w);
        if (!in_noise_mask(r, c, inc, exc)) continue;
      }
      add(x[i]);
    }
  }
  __syncthreads();
  uint32_t* gh = hist + (size_t)s * kBins;
  for (int i = threadIdx.x; i < kBins; i += kT)
    if (h[i]) atomicAdd(&gh[i], h[i]);
  if (above) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const K v = __shfl_xor(best, o, 64);
      best = v < best ? v : best;
    }
    // one atomic per block (per-wave atomics on the one address serialise at the L2)
    __shared__ K wmin[kT / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 1; k < kT / 64; ++k) best = wmin[k] < best ? wmin[k] : best;
      if (best != none) atomicMin(&st[s].hi_key, best);
    }
  }
}

// one block per plane: locate the bucket holding the target rank, extend the prefix, clear the histogram
template <class K>
__global__ __launch_bounds__(kT) void k_mad_select(SelStateT<K>* st, uint32_t* hist, int shift, int nbits, int last,
                                                   const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.x;
  uint32_t* gh = hist + (size_t)s * kBins;
  __shared__ uint32_t cnt[kBins];
  __shared__ uint32_t part[kT];
  const int nb = 1 << nbits;
  const int per = kBins / kT;  // 8 bins per thread
  const SelStateT<K> S0 = st[s];  // read by every thread before the barrier; one thread writes it after
  uint32_t local = 0;
  for (int k = 0; k < per; ++k) {
    const int b = threadIdx.x * per + k;
    const uint32_t v = (b < nb) ? gh[b] : 0u;
    cnt[b] = v;
    local += v;
    gh[b] = 0u;
  }
  // inclusive scan of the per-thread counts (wave shuffles, then the wave totals)
  uint32_t inc = local;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) part[wid] = inc;
  __syncthreads();
  uint32_t before = 0;
  for (int k = 0; k < wid; ++k) before += part[k];
  inc += before;
  const uint32_t exc = inc - local;
  // the thread whose range of ranks [exc, inc) holds the target rank finishes the search (the
  // last thread also takes a rank beyond the total, as the serial scan did)
  const bool mine = (S0.rank >= exc && S0.rank < inc) || (threadIdx.x == kT - 1 && S0.rank >= inc);
  __shared__ SelStateT<K> s_state;
  __shared__ int s_bin;
  if (mine) {
    SelStateT<K> S = S0;
    uint32_t run = exc;
    const int t = threadIdx.x;
    int b = t * per;
    for (; b < t * per + per - 1; ++b) {
      if (run + cnt[b] > S.rank) break;
      run += cnt[b];
    }
    S.rank -= run;
    S.prefix |= (K)b << shift;
    if (last) S.eq = cnt[b];
    s_state = S;
    s_bin = b;
  }
  __syncthreads();
  if (!last) {
    if (threadIdx.x == 0) st[s] = s_state;
    return;
  }
  // the smallest key above the low median: the next non-empty bucket of this last digit (same
  // higher digits), else the smallest key with larger higher digits (hi_key, from the histogram pass)
  const int b = s_bin;
  int nxt = kBins;
  for (int k = 0; k < per; ++k) {
    const int bb = threadIdx.x * per + k;
    if (bb > b && bb < nb && cnt[bb] != 0u && bb < nxt) nxt = bb;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(nxt, o, 64);
    nxt = v < nxt ? v : nxt;
  }
  __shared__ int s_nxt[kT / 64];
  if (lane == 0) s_nxt[wid] = nxt;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kT / 64; ++k) nxt = s_nxt[k] < nxt ? s_nxt[k] : nxt;
    SelStateT<K> S = s_state;
    const K dm = (((K)1 << nbits) - 1) << shift;  // this digit's bits
    if (nxt < nb) S.hi_key = (S.prefix & ~dm) | ((K)nxt << shift);
    st[s] = S;
  }
}

template <class K>
__global__ void k_mad_init(SelStateT<K>* st, int S, uint32_t n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  SelStateT<K> z{};
  z.n = n;
  z.rank = (n - 1) / 2;
  z.prefix = 0;
  z.hi_key = ~(K)0;  // lowered by the last histogram pass and the last select
  st[s] = z;
}

template <class T>
__global__ void k_mad_finish(const SelStateT<typename MadKey<T>::K>* st, int S, double* thr, const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const SelStateT<typename MadKey<T>::K> z = st[s];
  const uint32_t klo = (z.n - 1) / 2;
  const T vlo = MadKey<T>::val(z.prefix);
  double med = (double)vlo;
  if ((z.n & 1u) == 0) {
    // rank of the last key equal to vlo = first + eq - 1, first = klo - remaining rank
    const uint32_t first = klo - z.rank;
    const uint32_t lastr = first + z.eq - 1;
    const T vhi = (klo + 1 <= lastr) ? vlo : MadKey<T>::val(z.hi_key);
    med = ((double)vlo + (double)vhi) * 0.5;
  }
  thr[s] = med / 0.6745;
}

// ----------------------------------------------------------------------------- threshold + maxima
template <class T>
__global__ __launch_bounds__(kT) void k_threshold(const T* __restrict__ in, T* __restrict__ out, size_t plane, int w,
                                                  const double* __restrict__ thr, double sigma,
                                                  const int* __restrict__ border, T* __restrict__ pmax,
                                                  const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.y;
  const int h = (int)(plane / w);
  const double lim = sigma * thr[s];
  const int b = border ? border[s] : 0;
  const T* x = in + (size_t)s * plane;
  T* y = out + (size_t)s * plane;
  T m = -INFINITY;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += (size_t)gridDim.x * kT) {
    const T v = x[i];
    T o = ((double)fabs(v) > lim) ? v : (T)0;
    if (border) {
      const int r = (int)(i / w), c = (int)(i % w);
      if (r < b || r >= h - b || c < b || c >= w - b) o = 0;
    }
    y[i] = o;
    m = o > m ? o : m;
  }
  m = block_max_t<T>(m);
  if (threadIdx.x == 0) pmax[(size_t)s * gridDim.x + blockIdx.x] = m;
}

// per-plane max (or max |w|) of a plane stack: partials[s * gridDim.x + block]
template <class T>
__global__ __launch_bounds__(kT) void k_plane_max(const T* __restrict__ in, size_t plane, int use_abs,
                                                  T* __restrict__ pmax, const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.y;
  const T* x = in + (size_t)s * plane;
  T m = -INFINITY;
  auto take = [&](T v) {
    const T a = use_abs ? (T)fabs(v) : v;
    m = a > m ? a : m;
  };
  const size_t stride = (size_t)gridDim.x * kT;
  size_t tail = 0;
  if constexpr (sizeof(T) == 4) {
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {  // float4 loads; the maximum does not depend on order
      const float4* x4 = reinterpret_cast<const float4*>(x);
      const size_t n4 = plane >> 2;
      for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n4; i += stride) {
        const float4 a = x4[i];
        take(a.x), take(a.y), take(a.z), take(a.w);
      }
      tail = 4 * n4;
    }
  }
  for (size_t i = tail + blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += stride) take(x[i]);
  m = block_max_t<T>(m);
  if (threadIdx.x == 0) pmax[(size_t)s * gridDim.x + blockIdx.x] = m;
}

// per-plane reduce of partial maxima; out_max[s] (float) and optionally nsm[s] = max / energy[s]
// (one block per plane; max is exact in any order)
template <class T>
__global__ __launch_bounds__(kT) void k_max_final(const T* pmax, int nb, T* out_max, const double* energy,
                                                  double* nsm, const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.x;
  T m = -INFINITY;
  for (int b = threadIdx.x; b < nb; b += kT) {
    const T v = pmax[(size_t)s * nb + b];
    m = v > m ? v : m;
  }
  m = block_max_t<T>(m);
  if (threadIdx.x == 0) {
    if (out_max) out_max[s] = m;
    if (nsm) nsm[s] = (double)m / energy[s];
  }
}

// ----------------------------------------------------------------------------- connected components
__device__ __forceinline__ int ccl_find(const int* lab, int x) {
  int p = lab[x];
  while (p != x) {
    x = p;
    p = lab[x];
  }
  return x;
}

__device__ __forceinline__ void ccl_union(int* lab, int a, int b) {
  while (true) {
    a = ccl_find(lab, a);
    b = ccl_find(lab, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&lab[a], b);
    if (old == a) return;
    a = old;
  }
}

// Two-level labelling: each 32x32 tile is labelled in LDS (row runs linked to their first pixel,
// then union-find with LDS atomicMin links between rows, so every tile component's root is its
// smallest raster index), the tile roots are written as global
// raster indices, then only pixels on tile borders union across tiles in global memory.  The final
// forest, and so the roots, are those of a global union-find: every component's root is its
// smallest raster index.
constexpr int kCclTile = 32;

__device__ __forceinline__ int lds_find(const int* L, int x) {
  int p = L[x];
  while (p != x) {
    x = p;
    p = L[x];
  }
  return x;
}

__device__ __forceinline__ void lds_union(int* L, int a, int b) {
  while (true) {
    a = lds_find(L, a);
    b = lds_find(L, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&L[a], b);
    if (old == a) return;
    a = old;
  }
}

template <class V>
__global__ __launch_bounds__(kT) void k_ccl_tile(const V* __restrict__ w, int* __restrict__ lab, int h, int wd,
                                                 const int* pred) {
  if (pred_off(pred)) return;
  constexpr int T = kCclTile, TT = kCclTile * kCclTile;
  __shared__ int L[TT];
  const int s = blockIdx.y;
  const int tiles_x = (wd + T - 1) / T;
  const int r0 = (blockIdx.x / tiles_x) * T, c0 = (blockIdx.x % tiles_x) * T;
  const size_t plane = (size_t)h * wd;
  const V* x = w + (size_t)s * plane;
  int* l = lab + (size_t)s * plane;
  // A tile row is half a wave: one ballot gives each pixel the unset pixels left of it in its row, so
  // every run of set pixels starts as a tree one link deep under its first pixel (its smallest index)
  // and row neighbours need no union.
  static_assert(T == 32 && TT % kT == 0 && kT % 64 == 0, "one tile row per half wave, uniform loop");
  for (int i = threadIdx.x; i < TT; i += kT) {
    const int r = r0 + i / T, c = c0 + i % T;
    const bool on = r < h && c < wd && x[(size_t)r * wd + c] != (V)0;
    const unsigned long long m = __ballot(on);
    const int lo = i & 31;  // == this lane's position within its half wave
    const unsigned gaps = ~(unsigned)(m >> (threadIdx.x & 32)) & ((1u << lo) - 1u);
    L[i] = on ? i - lo + (gaps ? 32 - __clz(gaps) : 0) : -1;
  }
  __syncthreads();
  // 8-connected unions with the row above: the pixel above alone when it is set (the diagonal ones
  // are then in its run), otherwise each diagonal pixel that is set
  for (int i = threadIdx.x; i < TT; i += kT) {
    if (L[i] < 0) continue;
    const int tr = i / T, tc = i % T;
    if (tr > 0) {
      const int up = i - T;
      if (L[up] >= 0) {
        lds_union(L, i, up);
      } else {
        if (tc > 0 && L[up - 1] >= 0) lds_union(L, i, up - 1);
        if (tc + 1 < T && L[up + 1] >= 0) lds_union(L, i, up + 1);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < TT; i += kT) {
    const int r = r0 + i / T, c = c0 + i % T;
    if (r >= h || c >= wd) continue;
    int v = L[i];
    if (v >= 0) {
      v = lds_find(L, i);
      v = (r0 + v / T) * wd + (c0 + v % T);
    }
    l[(size_t)r * wd + c] = v;
  }
}

// unions across tile borders (first row, first and last column of each tile)
// Unions across tile borders.  Only border pixels take part -- every pixel of a row that starts a
// tile row (r % T == 0, r > 0), and on the other rows the first and last column of each tile -- so
// one block walks one row at a time and its threads take that row's border pixels, with 32-bit
// indices (planes hold fewer than 2^29 pixels).  Unions commute: the forest's roots, each
// component's smallest index, do not depend on their order.
__global__ __launch_bounds__(kT) void k_ccl_border(int* __restrict__ lab, int h, int wd, const int* pred) {
  if (pred_off(pred)) return;
  constexpr int T = kCclTile;
  const int s = blockIdx.y;
  int* l = lab + (size_t)s * h * wd;
  const int edge_cols = 2 * ((wd + T - 1) / T);  // columns tT and tT + T - 1 of every tile t
  for (int r = blockIdx.x; r < h; r += gridDim.x) {
    const bool top = r > 0 && r % T == 0;
    const int n = top ? wd : edge_cols;
    for (int k = threadIdx.x; k < n; k += kT) {
      const int c = top ? k : (k >> 1) * T + ((k & 1) ? T - 1 : 0);
      if (c >= wd) continue;
      const int cr = c % T;
      const bool left = c > 0 && cr == 0, right = cr == T - 1;
      const int me = r * wd + c;
      if (!(top || left || right) || l[me] < 0) continue;
      if (left && l[me - 1] >= 0) ccl_union(l, me, me - 1);
      if (r > 0) {
        const int up = me - wd;
        if (c > 0 && (top || left) && l[up - 1] >= 0) ccl_union(l, me, up - 1);
        if (top && l[up] >= 0) ccl_union(l, me, up);
        if (c + 1 < wd && (top || right) && l[up + 1] >= 0) ccl_union(l, me, up + 1);
      }
    }
  }
}

__global__ __launch_bounds__(kT) void k_ccl_compress(int* __restrict__ lab, size_t plane, const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.y;
  int* l = lab + (size_t)s * plane;
  const size_t stride = (size_t)gridDim.x * kT;
  size_t tail = 0;
  if ((reinterpret_cast<uintptr_t>(l) & 15) == 0) {  // four labels per load; roots found per element
    const int4* l4 = reinterpret_cast<const int4*>(l);
    const size_t n4 = plane >> 2;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n4; i += stride) {
      const int4 v = l4[i];
      const int b = (int)(4 * i);
      if (v.x >= 0) l[b] = ccl_find(l, b);
      if (v.y >= 0) l[b + 1] = ccl_find(l, b + 1);
      if (v.z >= 0) l[b + 2] = ccl_find(l, b + 2);
      if (v.w >= 0) l[b + 3] = ccl_find(l, b + 3);
    }
    tail = 4 * n4;
  }
  for (size_t i = tail + blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += stride)
    if (l[i] >= 0) l[i] = ccl_find(l, (int)i);
}

// One step of the per-scale keep rule, coarsest scale first, in one pass over the pixels: applies
// scale sa's marks (mask = pixel's object kept, extracted = mask * w) and marks the roots of scale
// sm = sa - 1's objects that hold a seed pixel (value >= tol * max) inside sa's kept mask.  sa < 0:
// mark only (the coarsest scale, which has no coarser mask); sm < 0: apply only.  A mark is the
// scale's number plus one, in one of two flag planes by the scale's parity, so a launch reads the
// marks of one plane while it writes the other's and no plane needs clearing between scales.
template <class T>
__global__ __launch_bounds__(kT) void k_keep_step(const T* __restrict__ w, const int* __restrict__ lab, int sa, int sm,
                                                  const T* __restrict__ smax, double tol, int neg_comp, size_t plane,
                                                  uint8_t* __restrict__ keep2, size_t kstride,
                                                  uint8_t* __restrict__ mask, T* __restrict__ extracted,
                                                  const int* pred) {
  if (pred_off(pred)) return;
  const double lim = sm >= 0 ? tol * (double)smax[sm] : 0.0;
  const uint8_t* ka = keep2 + (size_t)(sa & 1) * kstride;
  uint8_t* km = keep2 + (size_t)(sm & 1) * kstride;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += (size_t)gridDim.x * kT) {
    bool coarse_kept = true;
    if (sa >= 0) {
      const size_t o = (size_t)sa * plane + i;
      const int root = lab[o];
      const bool k = root >= 0 && ka[root] == (uint8_t)(sa + 1);
      mask[o] = k;
      if (extracted) extracted[o] = k ? w[o] : (T)0;
      coarse_kept = k;
    }
    if (sm >= 0 && coarse_kept) {
      const size_t o = (size_t)sm * plane + i;
      const int root = lab[o];
      if (root < 0) continue;
      const T v = neg_comp ? (T)fabs(w[o]) : w[o];
      if ((double)v >= lim) km[root] = (uint8_t)(sm + 1);
    }
  }
}

__global__ void k_clear_u8(uint8_t* p, size_t n, const int* pred) {
  if (pred_off(pred)) return;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// raster-order label numbering: count roots per block, scan, write ranks
__global__ __launch_bounds__(kT) void k_root_count(const int* __restrict__ lab, size_t plane, int chunk,
                                                   uint32_t* __restrict__ counts) {
  const int s = blockIdx.y;
  const int* l = lab + (size_t)s * plane;
  const size_t lo = (size_t)blockIdx.x * chunk;
  const size_t hi = min(plane, lo + (size_t)chunk);
  uint32_t c = 0;
  for (size_t i = lo + threadIdx.x; i < hi; i += kT) c += (l[i] == (int)i);
  __shared__ uint32_t red[kT];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[(size_t)s * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_root_scan(uint32_t* counts, int nb, int S, uint32_t* totals) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  uint32_t run = 0;
  for (int b = 0; b < nb; ++b) {
    const uint32_t v = counts[(size_t)s * nb + b];
    counts[(size_t)s * nb + b] = run;
    run += v;
  }
  if (totals) totals[s] = run;
}

// serial-in-chunk ranks (one thread per chunk keeps raster order), then label each pixel by its root's rank
__global__ void k_root_rank(int* __restrict__ lab, size_t plane, int chunk, const uint32_t* __restrict__ offs,
                            int nb, int* __restrict__ rank) {
  const int s = blockIdx.y;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int* l = lab + (size_t)s * plane;
  int* rk = rank + (size_t)s * plane;
  uint32_t run = offs[(size_t)s * nb + b];
  const size_t lo = (size_t)b * chunk;
  const size_t hi = min(plane, lo + (size_t)chunk);
  for (size_t i = lo; i < hi; ++i)
    if (l[i] == (int)i) rk[i] = (int)(++run);
}

__global__ __launch_bounds__(kT) void k_label_out(const int* __restrict__ lab, const int* __restrict__ rank,
                                                  size_t plane, int* __restrict__ out) {
  const int s = blockIdx.y;
  const int* l = lab + (size_t)s * plane;
  const int* rk = rank + (size_t)s * plane;
  int* o = out + (size_t)s * plane;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += (size_t)gridDim.x * kT)
    o[i] = (l[i] >= 0) ? rk[l[i]] : 0;
}

// ----------------------------------------------------------------------------- reductions & CG vectors
// partial[b] = sum over the block's grid-stride range of a[i]*b[i] (b == null: a[i]^2)
template <class T>
__global__ __launch_bounds__(kT) void k_dot(const T* __restrict__ a, const T* __restrict__ b, size_t n,
                                            double* __restrict__ partial, const int* pred) {
  if (pred_off(pred)) return;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const double x = a[i];
    acc = fma(x, b ? (double)b[i] : x, acc);
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// per-plane sums of squares of a plane stack: partial[s * gridDim.x + block]
template <class T>
__global__ __launch_bounds__(kT) void k_plane_sumsq(const T* __restrict__ a, size_t plane,
                                                    double* __restrict__ partial, const int* pred) {
  if (pred_off(pred)) return;
  const int s = blockIdx.y;
  const T* x = a + (size_t)s * plane;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < plane; i += (size_t)gridDim.x * kT) {
    const double v = x[i];
    acc = fma(v, v, acc);
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) partial[(size_t)s * gridDim.x + blockIdx.x] = acc;
}

// sum of (x - mean)^2 where mean = *mean_sum / count
template <class T>
__global__ __launch_bounds__(kT) void k_dev2(const T* __restrict__ a, size_t n, const double* __restrict__ sum,
                                             double* __restrict__ partial) {
  const double mean = *sum / (double)n;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const double d = (double)a[i] - mean;
    acc = fma(d, d, acc);
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <class T>
__global__ __launch_bounds__(kT) void k_sum(const T* __restrict__ a, size_t n, double* __restrict__ partial) {
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) acc += (double)a[i];
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// out[g] = sum of partial[g*np .. g*np + np) in a fixed order (one block per group)
__global__ __launch_bounds__(kT) void k_sum_partials(const double* __restrict__ partial, long np,
                                                     double* __restrict__ out, const int* pred) {
  if (pred_off(pred)) return;
  const double* p = partial + (size_t)blockIdx.x * np;
  double acc = 0.0;
  for (long i = threadIdx.x; i < np; i += kT) acc += p[i];
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

struct SumBatch {
  const double* partial[kMaxBatch];
  long np[kMaxBatch];
  double* out[kMaxBatch];
  const int* pred[kMaxBatch];
};

// job blockIdx.x: same summation order as k_sum_partials for one group
__global__ __launch_bounds__(kT) void k_sum_partials_batch(SumBatch b) {
  if (pred_off(b.pred[blockIdx.x])) return;
  const double* p = b.partial[blockIdx.x];
  const long np = b.np[blockIdx.x];
  double acc = 0.0;
  for (long i = threadIdx.x; i < np; i += kT) acc += p[i];
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) *b.out[blockIdx.x] = acc;
}

}  // namespace

// ============================================================================= host wrappers
static int ew_grid(size_t n) {
  size_t b = (n + (size_t)kT * 4 - 1) / ((size_t)kT * 4);
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}

int reduce_grid(size_t n) { return ew_grid(n); }

static bool aligned16(const void* a, const void* b = nullptr) {
  return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

// blocks for a 16-byte streaming kernel doing four accesses per thread
static int vec_grid(size_t n) {
  size_t b = ((n >> 2) + (size_t)kT * 4 - 1) / ((size_t)kT * 4);
  if (b < 1) b = 1;
  if (b > 16384) b = 16384;
  return (int)b;
}

void copy_plane(const float* src, float* dst, size_t n, const int* pred, hipStream_t st) {
  if (aligned16(src, dst))
    hipLaunchKernelGGL(k_copy, dim3(vec_grid(n)), dim3(kT), 0, st, src, dst, n, pred);
  else
    hipLaunchKernelGGL(k_copy1<float>, dim3(ew_grid(n)), dim3(kT), 0, st, src, dst, n, pred);
  PMS_LAUNCH_CHECK();
}

// a double plane is copied as twice as many floats (bit copy)
void copy_plane(const double* src, double* dst, size_t n, const int* pred, hipStream_t st) {
  copy_plane(reinterpret_cast<const float*>(src), reinterpret_cast<float*>(dst), 2 * n, pred, st);
}

template <class T>
static void masked_copy_t(const T* src, const uint8_t* mask, T* dst, size_t n, const int* pred, hipStream_t st) {
  hipLaunchKernelGGL(k_masked_copy<T>, dim3(ew_grid(n)), dim3(kT), 0, st, src, mask, dst, n, pred);
  PMS_LAUNCH_CHECK();
}
void masked_copy(const float* src, const uint8_t* mask, float* dst, size_t n, const int* pred, hipStream_t st) {
  masked_copy_t(src, mask, dst, n, pred, st);
}
void masked_copy(const double* src, const uint8_t* mask, double* dst, size_t n, const int* pred, hipStream_t st) {
  masked_copy_t(src, mask, dst, n, pred, st);
}

void fill_f32(float* p, size_t n, float v, hipStream_t st) {
  if (aligned16(p))
    hipLaunchKernelGGL(k_fill16<float>, dim3(vec_grid(n)), dim3(kT), 0, st, p, n, v);
  else
    hipLaunchKernelGGL(k_fill_f32, dim3(ew_grid(n)), dim3(kT), 0, st, p, n, v);
  PMS_LAUNCH_CHECK();
}

void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
  if (aligned16(p))
    hipLaunchKernelGGL(k_fill16<uint32_t>, dim3(vec_grid(n)), dim3(kT), 0, st, p, n, v);
  else
    hipLaunchKernelGGL(k_fill_u32, dim3(ew_grid(n)), dim3(kT), 0, st, p, n, v);
  PMS_LAUNCH_CHECK();
}

void fill_zero(float* p, size_t n, hipStream_t st) { fill_u32(reinterpret_cast<uint32_t*>(p), n, 0u, st); }
void fill_zero(double* p, size_t n, hipStream_t st) { fill_u32(reinterpret_cast<uint32_t*>(p), 2 * n, 0u, st); }

template <class T, class TIN>
static void window_copy_t(const TIN* src, int ld, int r0, int c0, int n, const T* mul, T* dst, hipStream_t st) {
  hipLaunchKernelGGL((k_window<T, TIN>), dim3(ew_grid((size_t)n * n)), dim3(kT), 0, st, src, ld, r0, c0, n, mul,
                     dst);
  PMS_LAUNCH_CHECK();
}
void window_copy(const float* src, int ld, int r0, int c0, int n, const float* mul, float* dst, hipStream_t st) {
  window_copy_t(src, ld, r0, c0, n, mul, dst, st);
}
void window_copy(const float* src, int ld, int r0, int c0, int n, const double* mul, double* dst, hipStream_t st) {
  window_copy_t(src, ld, r0, c0, n, mul, dst, st);
}
void window_copy(const double* src, int ld, int r0, int c0, int n, const double* mul, double* dst, hipStream_t st) {
  window_copy_t(src, ld, r0, c0, n, mul, dst, st);
}

// Python slice(start, stop).indices(n) for step 1, as [lo, hi) with hi >= lo
static void py_slice(long start, long stop, long n, int* lo, int* hi) {
  auto fix = [n](long v) {
    if (v < 0) v += n;
    if (v < 0) v = 0;
    if (v > n) v = n;
    return v;
  };
  long a = fix(start), b = fix(stop);
  if (b < a) b = a;
  *lo = (int)a;
  *hi = (int)b;
}

NoiseMask noise_mask(int h, int w, int edge_excl, int int_excl) {
  NoiseMask m{};
  m.masked = (edge_excl != 0) || (int_excl != 0);
  m.inc_r0 = 0;
  m.inc_r1 = h;
  m.inc_c0 = 0;
  m.inc_c1 = w;
  m.exc_r0 = m.exc_r1 = m.exc_c0 = m.exc_c1 = 0;
  if (edge_excl != 0) {
    py_slice(edge_excl, -edge_excl, h, &m.inc_r0, &m.inc_r1);
    py_slice(edge_excl, -edge_excl, w, &m.inc_c0, &m.inc_c1);
  }
  if (int_excl != 0) {
    // iuwt_toolbox.py:29,40: mid = in1.shape[1]/2 is the row count of the (scales, rows, cols) stack,
    // and the same mid bounds both axes
    const int mid = h / 2;
    py_slice(mid - int_excl, mid + int_excl, h, &m.exc_r0, &m.exc_r1);
    py_slice(mid - int_excl, mid + int_excl, w, &m.exc_c0, &m.exc_c1);
  }
  const long inc = (long)(m.inc_r1 - m.inc_r0) * (m.inc_c1 - m.inc_c0);
  const int ir0 = std::max(m.inc_r0, m.exc_r0), ir1 = std::min(m.inc_r1, m.exc_r1);
  const int ic0 = std::max(m.inc_c0, m.exc_c0), ic1 = std::min(m.inc_c1, m.exc_c1);
  const long ov = (ir1 > ir0 && ic1 > ic0) ? (long)(ir1 - ir0) * (ic1 - ic0) : 0;
  m.count = inc - ov;
  return m;
}

// workspace layout is independent of the plane count and element type of a call:
// [select state: 16 slots of 32 bytes][histograms 16 x 2048]
size_t mad_work_bytes(int) { return (size_t)kMaxPlanes * (kSelBytes + kBins * sizeof(uint32_t)); }

// Allocation-free median workspace: one file-scope device array per GPU, holding the select state
// and histograms of one call of up to kMaxPlanes planes (131,584 bytes).  Host code never names
// the array in a kernel argument or copy: the array's host-side symbol is a stand-in, not a
// device address, so each device's address is looked up once with hipGetSymbolAddress (with that
// device current) and kept here.
namespace {
constexpr size_t kNoAllocMadBytes = (size_t)kMaxPlanes * (kSelBytes + kBins * sizeof(uint32_t));
static __device__ __align__(256) unsigned char scratch[kNoAllocMadBytes];
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;
}  // namespace

size_t noalloc_mad_bytes() { return kNoAllocMadBytes; }

void* noalloc_mad_scratch() {
  int dev = 0;
  PMS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  PMS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  PMS_REQUIRE(ptr != nullptr, "device scratch symbol has no address");
  g_scratch_ptr.emplace(dev, ptr);
  return ptr;
}

void mad_work_init(void* work, int, hipStream_t st) {
  uint32_t* hist = (uint32_t*)((char*)work + (size_t)kMaxPlanes * kSelBytes);
  fill_u32(hist, (size_t)kMaxPlanes * kBins, 0u, st);
}

// Exact median per plane: radix select over the key bits in digits of at most 11 bits, one
// histogram pass and one select step per digit; for even counts the last pass and select also find
// the successor of the low median.
template <class T>
static void mad_thresholds_t(const T* planes, int S, int h, int w, const NoiseMask& nm, double* thr, void* work,
                             const int* pred, hipStream_t st) {
  using K = typename MadKey<T>::K;
  PMS_REQUIRE(nm.count > 0, "noise mask selects no pixels");
  PMS_REQUIRE(nm.count < (1L << 32), "plane too large for the median");
  PMS_REQUIRE(S >= 1 && S <= kMaxPlanes, "the median handles 1..16 planes per call");
  SelStateT<K>* sst = (SelStateT<K>*)work;
  uint32_t* hist = (uint32_t*)((char*)work + (size_t)kMaxPlanes * kSelBytes);
  const size_t plane = (size_t)h * w;
  Rect inc{nm.inc_r0, nm.inc_r1, nm.inc_c0, nm.inc_c1};
  Rect exc{nm.exc_r0, nm.exc_r1, nm.exc_c0, nm.exc_c1};
  hipLaunchKernelGGL(k_mad_init<K>, dim3(ceil_div(S, 64)), dim3(64), 0, st, sst, S, (uint32_t)nm.count);
  PMS_LAUNCH_CHECK();
  const int gx = std::min(ew_grid(plane), 512);
  // digits of 11 bits from the top of the key word (its sign bit is always clear); the last digit
  // takes what is left: 11/11/10 bits for float, 11 x 5 + 9 for double
  constexpr int kKeyBits = 8 * sizeof(K);
  constexpr int kPasses = (kKeyBits + 10) / 11;
  int shifts[kPasses], bits[kPasses];
  for (int pass = 0, top = kKeyBits; pass < kPasses; ++pass) {
    bits[pass] = std::min(11, top);
    top -= bits[pass];
    shifts[pass] = top;
  }
  // even counts also need the smallest key above the low median: the last histogram pass and the
  // last select find it on the way
  const int above = (nm.count & 1) == 0 ? 1 : 0;
  for (int pass = 0; pass < kPasses; ++pass) {
    const int pshift = (pass == 0) ? (int)(8 * sizeof(K)) : shifts[pass - 1];
    const int ab = pass == kPasses - 1 ? above : 0;
    if (nm.masked)
      hipLaunchKernelGGL((k_mad_hist<T, true>), dim3(gx, S), dim3(kT), 0, st, planes, plane, w, inc, exc,
                         shifts[pass], bits[pass], pshift, ab, sst, hist, pred);
    else
      hipLaunchKernelGGL((k_mad_hist<T, false>), dim3(gx, S), dim3(kT), 0, st, planes, plane, w, inc, exc,
                         shifts[pass], bits[pass], pshift, ab, sst, hist, pred);
    PMS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mad_select<K>, dim3(S), dim3(kT), 0, st, sst, hist, shifts[pass], bits[pass],
                       pass == kPasses - 1, pred);
    PMS_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_mad_finish<T>, dim3(ceil_div(S, 64)), dim3(64), 0, st, (const SelStateT<K>*)sst, S, thr, pred);
  PMS_LAUNCH_CHECK();
}
void mad_thresholds(const float* planes, int S, int h, int w, const NoiseMask& nm, double* thr, void* work,
                    const int* pred, hipStream_t st) {
  mad_thresholds_t(planes, S, h, w, nm, thr, work, pred, st);
}
void mad_thresholds(const double* planes, int S, int h, int w, const NoiseMask& nm, double* thr, void* work,
                    const int* pred, hipStream_t st) {
  mad_thresholds_t(planes, S, h, w, nm, thr, work, pred, st);
}

int threshold_grid(size_t plane) { return std::min(ew_grid(plane), 256); }

template <class T>
static void apply_threshold_t(const T* in, T* out, int S, int h, int w, const double* thr, double sigma,
                              const int* border, T* pmax, const int* pred, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  const int gx = threshold_grid(plane);
  hipLaunchKernelGGL(k_threshold<T>, dim3(gx, S), dim3(kT), 0, st, in, out, plane, w, thr, sigma, border, pmax, pred);
  PMS_LAUNCH_CHECK();
}
void apply_threshold(const float* in, float* out, int S, int h, int w, const double* thr, double sigma,
                     const int* border, float* pmax, const int* pred, hipStream_t st) {
  apply_threshold_t(in, out, S, h, w, thr, sigma, border, pmax, pred, st);
}
void apply_threshold(const double* in, double* out, int S, int h, int w, const double* thr, double sigma,
                     const int* border, double* pmax, const int* pred, hipStream_t st) {
  apply_threshold_t(in, out, S, h, w, thr, sigma, border, pmax, pred, st);
}

template <class T>
static void plane_max_t(const T* in, int S, size_t plane, int use_abs, T* pmax, const int* pred, hipStream_t st) {
  const int gx = threshold_grid(plane);
  hipLaunchKernelGGL(k_plane_max<T>, dim3(gx, S), dim3(kT), 0, st, in, plane, use_abs, pmax, pred);
  PMS_LAUNCH_CHECK();
}
void plane_max(const float* in, int S, size_t plane, int use_abs, float* pmax, const int* pred, hipStream_t st) {
  plane_max_t(in, S, plane, use_abs, pmax, pred, st);
}
void plane_max(const double* in, int S, size_t plane, int use_abs, double* pmax, const int* pred, hipStream_t st) {
  plane_max_t(in, S, plane, use_abs, pmax, pred, st);
}

template <class T>
static void max_final_t(const T* pmax, int nb, int S, T* out_max, const double* energy, double* nsm, const int* pred,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_max_final<T>, dim3(S), dim3(kT), 0, st, pmax, nb, out_max, energy, nsm, pred);
  PMS_LAUNCH_CHECK();
}
void max_final(const float* pmax, int nb, int S, float* out_max, const double* energy, double* nsm, const int* pred,
               hipStream_t st) {
  max_final_t(pmax, nb, S, out_max, energy, nsm, pred, st);
}
void max_final(const double* pmax, int nb, int S, double* out_max, const double* energy, double* nsm,
               const int* pred, hipStream_t st) {
  max_final_t(pmax, nb, S, out_max, energy, nsm, pred, st);
}

// workspace layout: [keep flags: two planes of bytes][maxima: 64 values][partial maxima: 16 x 2048
// values][labels: S planes of int]; everything before the labels sits at fixed offsets (values sized
// for double, so one layout serves both element types).
static size_t keep_bytes(size_t plane) { return ((plane + 255) / 256) * 256; }
static size_t ext_labels_offset(size_t plane) {
  return 2 * keep_bytes(plane) + 64 * sizeof(double) + (size_t)kMaxPlanes * 2048 * sizeof(double);
}

size_t extract_work_bytes(int S, int h, int w) {
  const size_t plane = (size_t)h * w;
  return ext_labels_offset(plane) + (size_t)S * plane * sizeof(int);
}

void extract_work_init(void* work, int, int h, int w, hipStream_t st) {
  const size_t plane = (size_t)h * w;
  hipLaunchKernelGGL(k_clear_u8, dim3(ew_grid(2 * plane)), dim3(kT), 0, st, (uint8_t*)work, 2 * keep_bytes(plane),
                     (const int*)nullptr);
  PMS_LAUNCH_CHECK();
}

template <class T>
static void source_extract_t(const T* w, int S, int h, int wd, double tol, int neg_comp, uint8_t* mask, T* extracted,
                             int* labels_out, void* work, const int* pred, hipStream_t st) {
  PMS_REQUIRE((size_t)h * wd < (size_t)1 << 31, "plane too large for int labels");
  PMS_REQUIRE(S >= 1 && S <= kMaxPlanes, "source extraction handles 1..16 planes per call");
  const size_t plane = (size_t)h * wd;
  uint8_t* keep = (uint8_t*)work;
  const size_t kstride = keep_bytes(plane);
  T* smax = (T*)((char*)work + 2 * kstride);
  T* pmax = (T*)((char*)work + 2 * kstride + 64 * sizeof(double));
  int* lab = (int*)((char*)work + ext_labels_offset(plane));
  const int gx = threshold_grid(plane);
  // keep flags start cleared whatever the workspace held before
  hipLaunchKernelGGL(k_clear_u8, dim3(ew_grid(2 * plane)), dim3(kT), 0, st, keep, 2 * kstride, pred);
  PMS_LAUNCH_CHECK();
  plane_max_t(w, S, plane, neg_comp, pmax, pred, st);
  max_final_t(pmax, gx, S, smax, (const double*)nullptr, (double*)nullptr, pred, st);
  const int gl = std::min(ew_grid(plane), 1024);
  const int tiles = ceil_div(h, kCclTile) * ceil_div(wd, kCclTile);
  hipLaunchKernelGGL(k_ccl_tile<T>, dim3(tiles, S), dim3(kT), 0, st, w, lab, h, wd, pred);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ccl_border, dim3(std::min(h, 4096), S), dim3(kT), 0, st, lab, h, wd, pred);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ccl_compress, dim3(gl, S), dim3(kT), 0, st, lab, plane, pred);
  PMS_LAUNCH_CHECK();
  // keep rule from the coarsest scale down: mark S-1; apply s and mark s-1; ...; apply 0
  for (int sa = S; sa >= 0; --sa) {
    hipLaunchKernelGGL(k_keep_step<T>, dim3(gl), dim3(kT), 0, st, w, (const int*)lab, sa < S ? sa : -1, sa - 1,
                       (const T*)smax, tol, neg_comp, plane, keep, kstride, mask, extracted, pred);
    PMS_LAUNCH_CHECK();
  }
  if (labels_out) {
    // rank roots in raster order (scipy.ndimage.label numbering)
    const int chunk = 4096;
    const int nb = (int)((plane + chunk - 1) / chunk);
    uint32_t* counts = nullptr;
    PMS_HIP(hipMallocAsync((void**)&counts, sizeof(uint32_t) * (size_t)S * nb, st));
    int* rank = nullptr;
    PMS_HIP(hipMallocAsync((void**)&rank, sizeof(int) * (size_t)S * plane, st));
    hipLaunchKernelGGL(k_root_count, dim3(nb, S), dim3(kT), 0, st, (const int*)lab, plane, chunk, counts);
    PMS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_root_scan, dim3(ceil_div(S, 64)), dim3(64), 0, st, counts, nb, S, (uint32_t*)nullptr);
    PMS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_root_rank, dim3(ceil_div(nb, 64), S), dim3(64), 0, st, lab, plane, chunk,
                       (const uint32_t*)counts, nb, rank);
    PMS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_label_out, dim3(gl, S), dim3(kT), 0, st, (const int*)lab, (const int*)rank, plane,
                       labels_out);
    PMS_LAUNCH_CHECK();
    PMS_HIP(hipFreeAsync(counts, st));
    PMS_HIP(hipFreeAsync(rank, st));
  }
}
void source_extract(const float* w, int S, int h, int wd, double tol, int neg_comp, uint8_t* mask, float* extracted,
                    int* labels_out, void* work, const int* pred, hipStream_t st) {
  source_extract_t(w, S, h, wd, tol, neg_comp, mask, extracted, labels_out, work, pred, st);
}
void source_extract(const double* w, int S, int h, int wd, double tol, int neg_comp, uint8_t* mask,
                    double* extracted, int* labels_out, void* work, const int* pred, hipStream_t st) {
  source_extract_t(w, S, h, wd, tol, neg_comp, mask, extracted, labels_out, work, pred, st);
}

// ---------------------------------------------------------------- reductions
template <class T>
static void dot_partials_t(const T* a, const T* b, size_t n, double* partial, const int* pred, hipStream_t st) {
  hipLaunchKernelGGL(k_dot<T>, dim3(reduce_grid(n)), dim3(kT), 0, st, a, b, n, partial, pred);
  PMS_LAUNCH_CHECK();
}
void dot_partials(const float* a, const float* b, size_t n, double* partial, const int* pred, hipStream_t st) {
  dot_partials_t(a, b, n, partial, pred, st);
}
void dot_partials(const double* a, const double* b, size_t n, double* partial, const int* pred, hipStream_t st) {
  dot_partials_t(a, b, n, partial, pred, st);
}

void sum_partials(const double* partial, long np, int groups, double* out, const int* pred, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_partials, dim3(groups), dim3(kT), 0, st, partial, np, out, pred);
  PMS_LAUNCH_CHECK();
}

void sum_partials_batch(const SumJob* jobs, int count, hipStream_t st) {
  for (int j0 = 0; j0 < count; j0 += kMaxBatch) {
    const int nb = std::min(kMaxBatch, count - j0);
    SumBatch b{};
    for (int k = 0; k < nb; ++k) {
      b.partial[k] = jobs[j0 + k].partial;
      b.np[k] = jobs[j0 + k].np;
      b.out[k] = jobs[j0 + k].out;
      b.pred[k] = jobs[j0 + k].pred;
    }
    hipLaunchKernelGGL(k_sum_partials_batch, dim3(nb), dim3(kT), 0, st, b);
    PMS_LAUNCH_CHECK();
  }
}

template <class T>
static void plane_sumsq_partials_t(const T* a, int S, size_t plane, double* partial, const int* pred, hipStream_t st) {
  hipLaunchKernelGGL(k_plane_sumsq<T>, dim3(reduce_grid(plane), S), dim3(kT), 0, st, a, plane, partial, pred);
  PMS_LAUNCH_CHECK();
}
void plane_sumsq_partials(const float* a, int S, size_t plane, double* partial, const int* pred, hipStream_t st) {
  plane_sumsq_partials_t(a, S, plane, partial, pred, st);
}
void plane_sumsq_partials(const double* a, int S, size_t plane, double* partial, const int* pred, hipStream_t st) {
  plane_sumsq_partials_t(a, S, plane, partial, pred, st);
}

template <class T>
static void mean_std_t(const T* a, size_t n, double* partial, double* sum_out, double* dev2_out, hipStream_t st) {
  const int g = reduce_grid(n);
  hipLaunchKernelGGL(k_sum<T>, dim3(g), dim3(kT), 0, st, a, n, partial);
  PMS_LAUNCH_CHECK();
  sum_partials(partial, g, 1, sum_out, nullptr, st);
  hipLaunchKernelGGL(k_dev2<T>, dim3(g), dim3(kT), 0, st, a, n, (const double*)sum_out, partial);
  PMS_LAUNCH_CHECK();
  sum_partials(partial, g, 1, dev2_out, nullptr, st);
}
void mean_std(const float* a, size_t n, double* partial, double* sum_out, double* dev2_out, hipStream_t st) {
  mean_std_t(a, n, partial, sum_out, dev2_out, st);
}
void mean_std(const double* a, size_t n, double* partial, double* sum_out, double* dev2_out, hipStream_t st) {
  mean_std_t(a, n, partial, sum_out, dev2_out, st);
}

}  // namespace pms
