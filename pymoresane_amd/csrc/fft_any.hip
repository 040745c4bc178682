// FFT convolution for any even frame side and for float64, hand-written for gfx950.
//
// Same three-pass structure and the same semantics as the power-of-two float32 kernels
// (fft.hip, reference pymoresane/iuwt_convolution.py:18-73): rows r2c -> column pass (forward,
// PSF product, inverse) -> rows c2r with scale and crop, the image at the frame origin and the
// P/2 roll folded into the PSF frame.  Differences: the transforms are mixed radix (2, 3, 4, 5,
// 7, 8, 11, 13) with the radix plan chosen at run time, and every kernel is templated on the
// arithmetic type, so this file serves
//   * float32 frames whose side is not a power of two (the reference's numpy rfft2 takes any
//     size, iuwt_convolution.py:64-73), and
//   * the float64 path of every frame side (the deconvolution's float64 precision mode).
//
// Each transform of length L runs in LDS as in-place decimation-in-frequency stages (forward,
// output in mixed-radix digit-reversed order) or their exact mirror (inverse, decimation in time,
// digit-reversed in, natural order out), so the column pass needs no reordering: the PSF
// spectrum is stored in the same digit-reversed order.  Stage twiddles and the radix roots are
// read from one table exp(-2 pi i k / P) per frame side P and type, computed on the host in
// float64.  A block holds G transforms (G = 1 for long ones), 256/G threads each.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "pms_internal.h"

namespace pms {
namespace {

constexpr int kGThreads = 256;

template <class T>
__device__ __forceinline__ Cx<T> cmk(T a, T b) {
  Cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <class T>
__device__ __forceinline__ Cx<T> gadd(Cx<T> a, Cx<T> b) {
  return cmk<T>(a.x + b.x, a.y + b.y);
}
template <class T>
__device__ __forceinline__ Cx<T> gsub(Cx<T> a, Cx<T> b) {
  return cmk<T>(a.x - b.x, a.y - b.y);
}
template <class T>
__device__ __forceinline__ Cx<T> gmul(Cx<T> a, Cx<T> b) {
  return cmk<T>(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
template <class T>
__device__ __forceinline__ Cx<T> gconj(Cx<T> a) {
  return cmk<T>(a.x, -a.y);
}
template <class T>
__device__ __forceinline__ Cx<T> gscale(Cx<T> a, T s) {
  return cmk<T>(a.x * s, a.y * s);
}
// multiply by -i (forward) or +i (inverse)
template <class T, bool FWD>
__device__ __forceinline__ Cx<T> gmul_i(Cx<T> a) {
  return FWD ? cmk<T>(a.y, -a.x) : cmk<T>(-a.y, a.x);
}

// LDS slot of element i (one padding slot per 32 elements)
__device__ __forceinline__ int gpad(int i) { return i + (i >> 5); }
__host__ __device__ __forceinline__ int gspan(int n) { return n + (n >> 5) + 1; }

constexpr int kMaxStages = 24;

// Radix plan of one transform length: stage i (decimation-in-frequency order) has radix r[i] and
// works on sub-transforms of length Ls_i = L / (r[0] ... r[i-1]).
struct GPlan {
  int L;
  int ns;
  int r[kMaxStages];
};

// Position of frequency k in the digit-reversed output of the forward transform:
// k = d0 + r0 (d1 + r1 (d2 + ...)), position = d0 L/r0 + d1 L/(r0 r1) + ...
__host__ __device__ __forceinline__ int gperm(const GPlan& p, int k) {
  int pos = 0, span = p.L;
  for (int i = 0; i < p.ns; ++i) {
    const int r = p.r[i];
    span /= r;
    const int q = k / r;
    pos += (k - q * r) * span;
    k = q;
  }
  return pos;
}

// In-register DFT of length R: x_m <- sum_n x_n exp(-+ 2 pi i m n / R) (FWD: minus).  The roots
// exp(-2 pi i j / R) are table entries j * rstep.
template <class T, int R, bool FWD>
struct GDft {
  using C = Cx<T>;
  __device__ __forceinline__ static void run(C (&x)[R], const C* __restrict__ tw, int rstep) {
    C y[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      C acc = x[0];
#pragma unroll
      for (int n = 1; n < R; ++n) {
        C w = tw[((m * n) % R) * rstep];
        if (!FWD) w = gconj(w);
        acc = gadd(acc, gmul(x[n], w));
      }
      y[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = y[m];
  }
};

template <class T, bool FWD>
struct GDft<T, 2, FWD> {
  using C = Cx<T>;
  __device__ __forceinline__ static void run(C (&x)[2], const C*, int) {
    const C a = x[0], b = x[1];
    x[0] = gadd(a, b);
    x[1] = gsub(a, b);
  }
};

template <class T, bool FWD>
struct GDft<T, 4, FWD> {
  using C = Cx<T>;
  __device__ __forceinline__ static void run(C (&x)[4], const C*, int) {
    const C s0 = gadd(x[0], x[2]), d0 = gsub(x[0], x[2]);
    const C s1 = gadd(x[1], x[3]), d1 = gmul_i<T, FWD>(gsub(x[1], x[3]));
    x[0] = gadd(s0, s1);
    x[2] = gsub(s0, s1);
    x[1] = gadd(d0, d1);
    x[3] = gsub(d0, d1);
  }
};

template <class T, bool FWD>
struct GDft<T, 8, FWD> {
  using C = Cx<T>;
  __device__ __forceinline__ static void run(C (&x)[8], const C* tw, int rstep) {
    C e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
    GDft<T, 4, FWD>::run(e, tw, rstep);
    GDft<T, 4, FWD>::run(o, tw, rstep);
    const T h = (T)0.70710678118654752440;  // cos(pi/4)
    // o_k * exp(-+ 2 pi i k / 8) for k = 0..3
    const C t1 = FWD ? cmk<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x))
                     : cmk<T>(h * (o[1].x - o[1].y), h * (o[1].y + o[1].x));
    const C t2 = gmul_i<T, FWD>(o[2]);
    const C t3 = FWD ? cmk<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y))
                     : cmk<T>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
    x[0] = gadd(e[0], o[0]);
    x[4] = gsub(e[0], o[0]);
    x[1] = gadd(e[1], t1);
    x[5] = gsub(e[1], t1);
    x[2] = gadd(e[2], t2);
    x[6] = gsub(e[2], t2);
    x[3] = gadd(e[3], t3);
    x[7] = gsub(e[3], t3);
  }
};

// One in-place stage of radix R over a transform of length L held in LDS, sub-transform length
// Ls (M = Ls/R).  Forward: DFT then twiddles (decimation in frequency); inverse: conjugate
// twiddles then the inverse DFT, the exact mirror.  TL is the table length (the frame side).
template <class T, int R, bool FWD>
__device__ __forceinline__ void gstage(Cx<T>* sm, int L, int Ls, int TL, const Cx<T>* __restrict__ tw, int lt,
                                       int nthr) {
  using C = Cx<T>;
  const int M = Ls / R;
  const int nbf = L / R;
  const int tstep = TL / Ls;
  const int rstep = TL / R;
  for (int b = lt; b < nbf; b += nthr) {
    const int blk = b / M;
    const int n = b - blk * M;
    const int base = blk * Ls + n;
    C x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = sm[gpad(base + k * M)];
    if (FWD) {
      GDft<T, R, true>::run(x, tw, rstep);
      if (M > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) x[k] = gmul(x[k], tw[n * k * tstep]);
      }
    } else {
      if (M > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) x[k] = gmul(x[k], gconj(tw[n * k * tstep]));
      }
      GDft<T, R, false>::run(x, tw, rstep);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) sm[gpad(base + k * M)] = x[k];
  }
}

template <class T, bool FWD>
__device__ __forceinline__ void gstage_any(int r, Cx<T>* sm, int L, int Ls, int TL, const Cx<T>* tw, int lt,
                                           int nthr) {
  switch (r) {
    case 2: gstage<T, 2, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 3: gstage<T, 3, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 4: gstage<T, 4, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 5: gstage<T, 5, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 7: gstage<T, 7, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 8: gstage<T, 8, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    case 11: gstage<T, 11, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
    default: gstage<T, 13, FWD>(sm, L, Ls, TL, tw, lt, nthr); break;
  }
}

// Whole transforms; callers synchronise before (inputs staged) -- each stage ends with a barrier.
template <class T>
__device__ __forceinline__ void gfft_forward(Cx<T>* sm, const GPlan& p, int TL, const Cx<T>* tw, int lt, int nthr) {
  int Ls = p.L;
  for (int i = 0; i < p.ns; ++i) {
    gstage_any<T, true>(p.r[i], sm, p.L, Ls, TL, tw, lt, nthr);
    Ls /= p.r[i];
    __syncthreads();
  }
}

template <class T>
__device__ __forceinline__ void gfft_inverse(Cx<T>* sm, const GPlan& p, int TL, const Cx<T>* tw, int lt, int nthr) {
  int Ls = 1;
  for (int i = p.ns - 1; i >= 0; --i) {
    Ls *= p.r[i];
    gstage_any<T, false>(p.r[i], sm, p.L, Ls, TL, tw, lt, nthr);
    __syncthreads();
  }
}

template <class T>
struct GR2cBatch {
  const void* src[kMaxBatch];
  Cx<T>* T1[kMaxBatch];
  const int* pred[kMaxBatch];
};
template <class T>
struct GColBatch {
  const Cx<T>* T1[kMaxBatch];
  const Cx<T>* S[kMaxBatch];
  Cx<T>* T2[kMaxBatch];
  const int* pred[kMaxBatch];
};
template <class T>
struct GC2rBatch {
  const Cx<T>* T2[kMaxBatch];
  T* out[kMaxBatch];
  const int* pred[kMaxBatch];
};

// ---- pass 1: rows r2c.  Frame row fr of the job: the image row (off == 0, q x q image at the
// frame origin) or the PSF row of the centred, -P/2-rolled frame (off = q/2).  Output column-major:
// T1[k * ldT + fr], k in [0, P/2].
template <class T, class TIN>
__global__ __launch_bounds__(kGThreads) void k_g_r2c(GR2cBatch<T> bt, GPlan pl, int P, int q, int off, int n_rows,
                                                     int ldT, int G, const Cx<T>* __restrict__ tw) {
  using C = Cx<T>;
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  const TIN* __restrict__ src = static_cast<const TIN*>(bt.src[jb]);
  C* __restrict__ T1 = bt.T1[jb];
  extern __shared__ unsigned char g_smem[];
  const int L = pl.L;
  const int nthr = kGThreads / G;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int fr = blockIdx.x * G + g;
  C* sm = reinterpret_cast<C*>(g_smem) + g * gspan(L);
  const bool live = fr < n_rows;
  const int a = off == 0 ? fr : (fr + off) % P;
  const bool row_ok = live && a < q;
  const TIN* row = src + (size_t)(row_ok ? a : 0) * q;
  for (int e = lt; e < L; e += nthr) {
    T v0 = 0, v1 = 0;
    if (row_ok) {
      const int c0 = off == 0 ? 2 * e : (2 * e + off) % P;
      const int c1 = off == 0 ? 2 * e + 1 : (2 * e + 1 + off) % P;
      if (c0 < q) v0 = (T)row[c0];
      if (c1 < q) v1 = (T)row[c1];
    }
    sm[gpad(e)] = cmk<T>(v0, v1);
  }
  __syncthreads();
  gfft_forward(sm, pl, P, tw, lt, nthr);
  if (!live) return;
  for (int k = lt; k <= L; k += nthr) {
    const C zk = sm[gpad(gperm(pl, k == L ? 0 : k))];
    const C zl = gconj(sm[gpad(gperm(pl, k == 0 ? 0 : L - k))]);
    const C e = gscale(gadd(zk, zl), (T)0.5);
    const C dd = gscale(gsub(zk, zl), (T)0.5);
    const C o = cmk<T>(dd.y, -dd.x);  // (zk - conj z_{L-k}) / (2i)
    T1[(size_t)k * ldT + fr] = gadd(e, gmul(tw[k], o));
  }
}

// ---- pass 2: per frequency column k in [0, P/2]: the n_in stored rows (zero beyond) -> forward
// P-point FFT.  SPEC: store the spectrum column (digit-reversed order).  Otherwise multiply by the
// PSF column (same order), inverse FFT, store the first n_out rows: T2[k * n_out + i].
template <class T, bool SPEC>
__global__ __launch_bounds__(kGThreads) void k_g_cols(GColBatch<T> bt, GPlan pl, int n_in, int n_out, int G,
                                                      const Cx<T>* __restrict__ tw) {
  using C = Cx<T>;
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  extern __shared__ unsigned char g_smem[];
  const int P = pl.L;
  const int ncols = P / 2 + 1;
  const int nthr = kGThreads / G;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int k = blockIdx.x * G + g;
  const bool live = k < ncols;
  C* sm = reinterpret_cast<C*>(g_smem) + g * gspan(P);
  const C* col = bt.T1[jb] + (size_t)(live ? k : 0) * n_in;
  for (int i = lt; i < P; i += nthr) sm[gpad(i)] = (live && i < n_in) ? col[i] : cmk<T>(0, 0);
  __syncthreads();
  gfft_forward(sm, pl, P, tw, lt, nthr);
  if (SPEC) {
    if (!live) return;
    C* out = bt.T2[jb] + (size_t)k * P;
    for (int i = lt; i < P; i += nthr) out[i] = sm[gpad(i)];
    return;
  }
  const C* S = bt.S[jb] + (size_t)(live ? k : 0) * P;
  for (int i = lt; i < P; i += nthr) sm[gpad(i)] = gmul(sm[gpad(i)], S[i]);
  __syncthreads();
  gfft_inverse(sm, pl, P, tw, lt, nthr);
  if (!live) return;
  C* out = bt.T2[jb] + (size_t)k * n_out;
  for (int i = lt; i < n_out; i += nthr) out[i] = sm[gpad(i)];
}

// ---- pass 3: per kept row r: Hermitian half spectrum T2[k * ldT + r], k in [0, P/2] -> the
// P/2-point complex sequence whose inverse FFT interleaves the even and odd output samples, scaled
// by 1/P^2; the first n columns are stored.
template <class T>
__global__ __launch_bounds__(kGThreads) void k_g_c2r(GC2rBatch<T> bt, GPlan pl, int P, int n, int ldT, int G, T scale,
                                                     const Cx<T>* __restrict__ tw) {
  using C = Cx<T>;
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  extern __shared__ unsigned char g_smem[];
  const int L = pl.L;
  const int nthr = kGThreads / G;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int r = blockIdx.x * G + g;
  const bool live = r < n;
  C* sm = reinterpret_cast<C*>(g_smem) + g * gspan(L);
  const C* T2 = bt.T2[jb];
  const int rr = live ? r : 0;
  for (int k = lt; k < L; k += nthr) {
    const C xk = T2[(size_t)k * ldT + rr];
    const C xl = gconj(T2[(size_t)(L - k) * ldT + rr]);
    const C d = gmul(gsub(xk, xl), gconj(tw[k]));  // (X_k - conj X_{L-k}) exp(+2 pi i k / P)
    const C z = gadd(gadd(xk, xl), cmk<T>(-d.y, d.x));
    sm[gpad(gperm(pl, k))] = z;
  }
  __syncthreads();
  gfft_inverse(sm, pl, P, tw, lt, nthr);
  if (!live) return;
  T* out = bt.out[jb] + (size_t)r * n;
  for (int e = lt; e < L; e += nthr) {
    const C z = sm[gpad(e)];
    if (2 * e < n) out[2 * e] = z.x * scale;
    if (2 * e + 1 < n) out[2 * e + 1] = z.y * scale;
  }
}

// ---- host side --------------------------------------------------------------------------------
bool make_gplan(int L, GPlan* p) {
  p->L = L;
  p->ns = 0;
  int m = L;
  // powers of two first as radix 8 (a final 4 or 2 for the remainder), then the odd primes
  while (m % 8 == 0 && p->ns < kMaxStages) {
    p->r[p->ns++] = 8;
    m /= 8;
  }
  if (m % 4 == 0) {
    p->r[p->ns++] = 4;
    m /= 4;
  }
  if (m % 2 == 0) {
    p->r[p->ns++] = 2;
    m /= 2;
  }
  for (int r : {3, 5, 7, 11, 13}) {
    while (m % r == 0 && p->ns < kMaxStages) {
      p->r[p->ns++] = r;
      m /= r;
    }
  }
  return m == 1;
}

int groups_for(int L, size_t elem) {
  // transforms per block: up to 16, keeping at least max(16, L/8) threads per transform and at most
  // 48 KB of LDS per block
  int G = 1;
  while (G < 16 && (size_t)(2 * G) * gspan(L) * elem <= 48 * 1024 && kGThreads / (2 * G) >= std::max(16, L / 8))
    G *= 2;
  return G;
}

struct TwKey {
  int dev, p, bytes;
  bool operator<(const TwKey& o) const {
    return dev != o.dev ? dev < o.dev : (p != o.p ? p < o.p : bytes < o.bytes);
  }
};
std::mutex g_gtw_mu;
std::map<TwKey, void*> g_gtw;

// exp(-2 pi i k / P), k in [0, P), in type T (computed in float64), one table per device and P
template <class T>
const Cx<T>* g_twiddles(int p) {
  using C = Cx<T>;
  int dev = 0;
  PMS_HIP(hipGetDevice(&dev));
  const TwKey key{dev, p, (int)sizeof(T)};
  std::lock_guard<std::mutex> lk(g_gtw_mu);
  auto it = g_gtw.find(key);
  if (it != g_gtw.end()) return static_cast<const C*>(it->second);
  std::vector<C> h(p);
  for (int k = 0; k < p; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)p;
    h[k].x = (T)std::cos(a);
    h[k].y = (T)std::sin(a);
  }
  void* d = nullptr;
  PMS_HIP(hipMalloc(&d, sizeof(C) * p));
  PMS_HIP(hipMemcpy(d, h.data(), sizeof(C) * p, hipMemcpyHostToDevice));
  g_gtw.emplace(key, d);
  return static_cast<const C*>(d);
}

template <class T>
void set_g_attrs() {
  static std::once_flag once;
  std::call_once(once, [] {
    const int lds_max = 160 * 1024;
    (void)hipFuncSetAttribute((const void*)k_g_r2c<T, float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void*)k_g_r2c<T, double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void*)k_g_cols<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void*)k_g_cols<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void*)k_g_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipGetLastError();
  });
}

template <class T>
void check_g_frame(int p) {
  GPlan pl;
  PMS_REQUIRE(p >= 4 && p % 2 == 0, "FFT frame side must be even and >= 4, got " + std::to_string(p));
  PMS_REQUIRE(p <= gfft_max_frame(sizeof(T)),
              "FFT frame side " + std::to_string(p) + " exceeds the " + std::to_string(gfft_max_frame(sizeof(T))) +
                  " that one LDS transform holds in " + (sizeof(T) == 8 ? "float64" : "float32"));
  PMS_REQUIRE(make_gplan(p, &pl) && make_gplan(p / 2, &pl),
              "FFT frame side " + std::to_string(p) + " has a prime factor above 13");
}

template <class T, class TIN>
void launch_r2c(const GR2cBatch<T>& bt, int nb, int p, int q, int off, int n_rows, int ldT,
                const Cx<T>* tw, hipStream_t st) {
  using C = Cx<T>;
  GPlan pl;
  make_gplan(p / 2, &pl);
  const int G = groups_for(p / 2, sizeof(C));
  const size_t lds = (size_t)G * gspan(p / 2) * sizeof(C);
  hipLaunchKernelGGL((k_g_r2c<T, TIN>), dim3(ceil_div(n_rows, G), nb), dim3(kGThreads), lds, st, bt, pl, p, q, off,
                     n_rows, ldT, G, tw);
  PMS_LAUNCH_CHECK();
}

template <class T, bool SPEC>
void launch_cols(const GColBatch<T>& bt, int nb, int p, int n_in, int n_out, const Cx<T>* tw,
                 hipStream_t st) {
  using C = Cx<T>;
  GPlan pl;
  make_gplan(p, &pl);
  const int G = groups_for(p, sizeof(C));
  const size_t lds = (size_t)G * gspan(p) * sizeof(C);
  hipLaunchKernelGGL((k_g_cols<T, SPEC>), dim3(ceil_div(p / 2 + 1, G), nb), dim3(kGThreads), lds, st, bt, pl, n_in,
                     n_out, G, tw);
  PMS_LAUNCH_CHECK();
}

template <class T>
void launch_c2r(const GC2rBatch<T>& bt, int nb, int p, int n, int ldT, const Cx<T>* tw, hipStream_t st) {
  using C = Cx<T>;
  GPlan pl;
  make_gplan(p / 2, &pl);
  const int G = groups_for(p / 2, sizeof(C));
  const size_t lds = (size_t)G * gspan(p / 2) * sizeof(C);
  const T scale = (T)(1.0 / ((double)p * (double)p));
  hipLaunchKernelGGL(k_g_c2r<T>, dim3(ceil_div(n, G), nb), dim3(kGThreads), lds, st, bt, pl, p, n, ldT, G, scale, tw);
  PMS_LAUNCH_CHECK();
}

}  // namespace

int gfft_max_frame(size_t real_bytes) { return real_bytes == 8 ? 8192 : 16384; }

bool gfft_frame_ok(int p, size_t real_bytes) {
  GPlan pl;
  return p >= 4 && p % 2 == 0 && p <= gfft_max_frame(real_bytes) && make_gplan(p, &pl) && make_gplan(p / 2, &pl);
}

int gfft_perm(int k, int p) {
  GPlan pl;
  PMS_REQUIRE(make_gplan(p, &pl), "FFT frame side " + std::to_string(p) + " has a prime factor above 13");
  return gperm(pl, k);
}

template <class T, class TIN>
void gfft_psf_spectrum(const TIN* psf, int q, int p, Cx<T>* spec, Cx<T>* work,
                       hipStream_t st) {
  check_g_frame<T>(p);
  PMS_REQUIRE(q >= 1 && q <= p, "PSF larger than the FFT frame");
  set_g_attrs<T>();
  const auto* tw = g_twiddles<T>(p);
  GR2cBatch<T> rb{};
  rb.src[0] = psf;
  rb.T1[0] = work;
  launch_r2c<T, TIN>(rb, 1, p, q, q / 2, p, p, tw, st);
  GColBatch<T> cb{};
  cb.T1[0] = work;
  cb.T2[0] = spec;
  launch_cols<T, true>(cb, 1, p, p, p, tw, st);
}

template <class T>
void gfft_conv_batch(const GConvJob<T>* jobs, int count, int n, int p, hipStream_t st) {
  check_g_frame<T>(p);
  PMS_REQUIRE(p == n || p == 2 * n, "fft_conv: frame must be N (circular) or 2N (linear)");
  set_g_attrs<T>();
  const auto* tw = g_twiddles<T>(p);
  const int L = p / 2;
  for (int j0 = 0; j0 < count; j0 += kMaxBatch) {
    const int nb = std::min(kMaxBatch, count - j0);
    GR2cBatch<T> rb{};
    GColBatch<T> cb{};
    GC2rBatch<T> ob{};
    for (int k = 0; k < nb; ++k) {
      const GConvJob<T>& j = jobs[j0 + k];
      PMS_REQUIRE(j.in && j.spec && j.out && j.work, "fft_conv: NULL buffer");
      auto* T1 = j.work;
      auto* T2 = j.work + (size_t)(L + 1) * n;
      rb.src[k] = j.in;
      rb.T1[k] = T1;
      rb.pred[k] = j.pred;
      cb.T1[k] = T1;
      cb.S[k] = j.spec;
      cb.T2[k] = T2;
      cb.pred[k] = j.pred;
      ob.T2[k] = T2;
      ob.out[k] = j.out;
      ob.pred[k] = j.pred;
    }
    launch_r2c<T, T>(rb, nb, p, n, 0, n, n, tw, st);
    launch_cols<T, false>(cb, nb, p, n, n, tw, st);
    launch_c2r<T>(ob, nb, p, n, n, tw, st);
  }
}

template void gfft_psf_spectrum<float, float>(const float*, int, int, Cx<float>*, Cx<float>*, hipStream_t);
template void gfft_psf_spectrum<double, float>(const float*, int, int, Cx<double>*, Cx<double>*, hipStream_t);
template void gfft_psf_spectrum<double, double>(const double*, int, int, Cx<double>*, Cx<double>*, hipStream_t);
template void gfft_conv_batch<float>(const GConvJob<float>*, int, int, int, hipStream_t);
template void gfft_conv_batch<double>(const GConvJob<double>*, int, int, int, hipStream_t);

}  // namespace pms
