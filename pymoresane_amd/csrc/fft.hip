// FFT convolution with a precomputed PSF spectrum, hand-written for gfx950.
//
// Semantics follow the reference fft_convolve (pymoresane/iuwt_convolution.py:18-73):
//   linear   : out = crop_N(fftshift(irfft2(S * rfft2(pad_2N(x)))))
//   circular : out = fftshift(irfft2(S * rfft2(x)))
// Both are circular convolutions on a P x P frame (P = 2N or N) followed by a
// crop.  Here the image sits at the frame origin and the P/2 spatial roll
// that the shift/crop implies is folded into the PSF frame once, when the
// spectrum is built, so the output is simply the first N rows/columns.
//
// Passes (all FFTs run in LDS, radix-16/32 butterflies in registers):
//   1. rows r2c   : each image row (zero-padded to P) -> P/2+1 frequencies
//                   via a P/2-point complex FFT; written column-major
//                   (frequency k owns a contiguous column of rows).
//   2. columns    : per frequency column: forward P-point FFT (decimation in
//                   frequency, output left in digit-reversed order), pointwise
//                   product with the PSF column stored in that same order,
//                   inverse FFT (decimation in time, natural order out); only
//                   the N kept rows are written.  No reordering pass exists.
//   3. rows c2r   : per kept row: Hermitian pre-twiddle into a P/2-point
//                   inverse FFT, scale 1/P^2, write the N kept columns.
// Twiddles come from a device table exp(-2 pi i k / 2^14) computed on the
// host in double precision; in-butterfly powers are formed by a balanced
// product tree from one table value.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <vector>

#include "pms_internal.h"

namespace pms {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// LDS index with one padding slot per 32 complex values (breaks power-of-two bank strides)
__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }

// cos / sin of 2 pi j / 32 for j in [0, 16)
struct Roots32 {
  static constexpr float c[16] = {1.000000000e+00f,  9.807852507e-01f,  9.238795042e-01f,
                                  8.314695954e-01f,  7.071067691e-01f,  5.555702448e-01f,
                                  3.826834261e-01f,  1.950903237e-01f,  0.0f,
                                  -1.950903237e-01f, -3.826834261e-01f, -5.555702448e-01f,
                                  -7.071067691e-01f, -8.314695954e-01f, -9.238795042e-01f,
                                  -9.807852507e-01f};
  static constexpr float s[16] = {0.0f,             1.950903237e-01f, 3.826834261e-01f,
                                  5.555702448e-01f, 7.071067691e-01f, 8.314695954e-01f,
                                  9.238795042e-01f, 9.807852507e-01f, 1.000000000e+00f,
                                  9.807852507e-01f, 9.238795042e-01f, 8.314695954e-01f,
                                  7.071067691e-01f, 5.555702448e-01f, 3.826834261e-01f,
                                  1.950903237e-01f};
};

// t = o * exp(DIR * 2 pi i k / R), k < R/2 known at compile time after unrolling
template <int R, int DIR>
__device__ __forceinline__ float2 rot(float2 o, int k) {
  const int j = k * (32 / R);
  if (j == 0) return o;
  if (j == 8) return DIR < 0 ? make_float2(o.y, -o.x) : make_float2(-o.y, o.x);
  const float2 w = make_float2(Roots32::c[j], DIR * Roots32::s[j]);
  return cmul(o, w);
}

// In-register DFT: x_k <- sum_m x_m exp(DIR * 2 pi i m k / R), natural order in and out.
template <int R, int DIR>
struct Dft {
  __device__ __forceinline__ static void run(float2 (&x)[R]) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      e[k] = x[2 * k];
      o[k] = x[2 * k + 1];
    }
    Dft<R / 2, DIR>::run(e);
    Dft<R / 2, DIR>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const float2 t = rot<R, DIR>(o[k], k);
      x[k] = cadd(e[k], t);
      x[k + R / 2] = csub(e[k], t);
    }
  }
};
template <int DIR>
struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2 (&)[1]) {}
};

// powers w^0 .. w^(R-1) by a balanced product tree
template <int R>
__device__ __forceinline__ void powers(float2 w, float2 (&wk)[R]) {
  wk[0] = make_float2(1.f, 0.f);
  if (R > 1) wk[1] = w;
#pragma unroll
  for (int k = 2; k < R; ++k) wk[k] = cmul(wk[k >> 1], wk[k - (k >> 1)]);
}

// One in-place radix-R stage over an LDS array of length L, sub-transform
// length Ls (M = Ls/R).  Forward = decimation in frequency; inverse runs the
// exact mirror (conjugate twiddles before the butterfly).
template <int R, bool FWD>
__device__ __forceinline__ void fft_stage(float2* sm, int L, int Ls, int lt, int nthr, const float2* tw) {
  const int M = Ls / R;
  const int mM = __ffs(M) - 1;  // M is a power of two
  const int nbf = L / R;
  const int tw_step = kFftMaxLen / Ls;
  for (int b = lt; b < nbf; b += nthr) {
    const int blk = b >> mM;
    const int n = b & (M - 1);
    const int base = blk * Ls + n;
    float2 x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = sm[pad(base + k * M)];
    if (FWD) {
      Dft<R, -1>::run(x);
      if (M > 1) {
        float2 wk[R];
        powers<R>(tw[n * tw_step], wk);
#pragma unroll
        for (int k = 1; k < R; ++k) x[k] = cmul(x[k], wk[k]);
      }
    } else {
      if (M > 1) {
        float2 wk[R];
        powers<R>(cconj(tw[n * tw_step]), wk);
#pragma unroll
        for (int k = 1; k < R; ++k) x[k] = cmul(x[k], wk[k]);
      }
      Dft<R, +1>::run(x);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) sm[pad(base + k * M)] = x[k];
  }
}

// Radix plan for L = 2^m: ns = ceil(m/4) stages, the first of radix 2^(m - 4(ns-1)) in {2,4,8,16},
// the rest radix 16.  Stage i (DIF order) works on sub-transforms of length Ls_i:
// Ls_0 = L, Ls_i = 16^(ns-i) for i >= 1.
__host__ __device__ __forceinline__ int plan_stages(int m) { return (m + 3) >> 2; }
__host__ __device__ __forceinline__ int plan_first_log2(int m) { return m - 4 * (plan_stages(m) - 1); }

template <bool FWD>
__device__ __forceinline__ void run_stage(int lr, float2* sm, int L, int Ls, int lt, int nthr, const float2* tw) {
  switch (lr) {
    case 1: fft_stage<2, FWD>(sm, L, Ls, lt, nthr, tw); break;
    case 2: fft_stage<4, FWD>(sm, L, Ls, lt, nthr, tw); break;
    case 3: fft_stage<8, FWD>(sm, L, Ls, lt, nthr, tw); break;
    default: fft_stage<16, FWD>(sm, L, Ls, lt, nthr, tw); break;
  }
}

// storage position of frequency k after the forward transform of length 2^m:
// k = d0 + R0 (d1 + 16 (d2 + ...)), position = d0 L/R0 + d1 L/(16 R0) + ...
__host__ __device__ __forceinline__ int perm_pos(int k, int m) {
  const int ns = plan_stages(m);
  const int lr0 = plan_first_log2(m);
  int mM = m - lr0;
  int pos = (k & ((1 << lr0) - 1)) << mM;
  k >>= lr0;
  for (int i = 1; i < ns; ++i) {
    mM -= 4;
    pos += (k & 15) << mM;
    k >>= 4;
  }
  return pos;
}

__device__ __forceinline__ float frame_val(const float* src, int q, int off, int pmask, int fr, int f) {
  const int a = (fr + off) & pmask;
  const int b = (f + off) & pmask;
  return (a < q && b < q) ? src[(size_t)a * q + b] : 0.f;
}

// ---- fused stage helpers --------------------------------------------------------------
// First DIF stage (radix R0, sub-transform length L, M0 = L/R0) fused with the load of the
// transform input: element e is fetched by get(e).  Batches of U butterflies are loaded before
// any is computed so that ~16 global loads per thread are in flight.
template <int R0, class Get>
__device__ __forceinline__ void stage0_load(float2* sm, int L, int lt, int nthr, const float2* tw, Get get) {
  const int M0 = L / R0;
  const int step = kFftMaxLen / L;
  constexpr int U = (R0 >= 16) ? 1 : (R0 >= 8) ? 2 : (16 / R0);
  for (int n0 = lt; n0 < M0; n0 += nthr * U) {
    float2 x[U][R0];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = n0 + u * nthr;
#pragma unroll
      for (int k = 0; k < R0; ++k) x[u][k] = (n < M0) ? get(n + k * M0) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = n0 + u * nthr;
      if (n < M0) {
        Dft<R0, -1>::run(x[u]);
        if (M0 > 1) {
          float2 wk[R0];
          powers<R0>(tw[n * step], wk);
#pragma unroll
          for (int k = 1; k < R0; ++k) x[u][k] = cmul(x[u][k], wk[k]);
        }
#pragma unroll
        for (int k = 0; k < R0; ++k) sm[pad(n + k * M0)] = x[u][k];
      }
    }
  }
}

// Mirror of stage0_load for the inverse transform: the last inverse stage fused with the
// store of the result; element e goes to put(e, value).
template <int R0, class Put>
__device__ __forceinline__ void stage0_store(const float2* sm, int L, int lt, int nthr, const float2* tw, Put put) {
  const int M0 = L / R0;
  const int step = kFftMaxLen / L;
  for (int n = lt; n < M0; n += nthr) {
    float2 x[R0];
#pragma unroll
    for (int k = 0; k < R0; ++k) x[k] = sm[pad(n + k * M0)];
    if (M0 > 1) {
      float2 wk[R0];
      powers<R0>(cconj(tw[n * step]), wk);
#pragma unroll
      for (int k = 1; k < R0; ++k) x[k] = cmul(x[k], wk[k]);
    }
    Dft<R0, +1>::run(x);
#pragma unroll
    for (int k = 0; k < R0; ++k) put(n + k * M0, x[k]);
  }
}

// middle radix-16 stages i in [i0, i1) (forward) of an ns-stage plan: Ls = 16^(ns-i)
__device__ __forceinline__ void mid_forward(float2* sm, int L, int ns, int i0, int i1, int lt, int nthr,
                                            const float2* tw) {
  for (int i = i0; i < i1; ++i) {
    fft_stage<16, true>(sm, L, 1 << (4 * (ns - i)), lt, nthr, tw);
    __syncthreads();
  }
}

__device__ __forceinline__ void mid_inverse(float2* sm, int L, int ns, int i0, int i1, int lt, int nthr,
                                            const float2* tw) {
  for (int i = i1 - 1; i >= i0; --i) {
    fft_stage<16, false>(sm, L, 1 << (4 * (ns - i)), lt, nthr, tw);
    __syncthreads();
  }
}

// Per-job pointers of a batched launch (grid y selects the job).
struct R2cBatch {
  const float* src[kMaxBatch];
  float2* T[kMaxBatch];
  const int* pred[kMaxBatch];
};
struct ColBatch {
  const float2* T1[kMaxBatch];
  const float2* S[kMaxBatch];
  float2* T2[kMaxBatch];
  const int* pred[kMaxBatch];
};
struct C2rBatch {
  const float2* T2[kMaxBatch];
  float* out[kMaxBatch];
  const int* pred[kMaxBatch];
};

// ---- pass 1: rows r2c.  Group g of the block handles frame row fr; output T[k*ldT + fr], k in [0, P/2].
// The frame row is the image row (off == 0, q == n) or the centred/rolled PSF (off = q/2).
template <int R0>
__global__ __launch_bounds__(kThreads) void k_rows_r2c(R2cBatch bt, int q, int off, int pm, int n_rows, int rb,
                                                       int ldT, const float2* __restrict__ tw) {
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  const float* __restrict__ src = bt.src[jb];
  float2* __restrict__ T = bt.T[jb];
  extern __shared__ float2 smem[];
  const int P = 1 << pm;
  const int L = P >> 1;
  const int mL = pm - 1;
  const int nthr = kThreads / rb;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int fr = xcd_remap(blockIdx.x, gridDim.x) * rb + g;
  float2* sm = smem + g * (L + (L >> 5));
  const bool live = fr < n_rows;
  const int pmask = P - 1;
  const int a = (fr + off) & pmask;
  const bool row_ok = live && a < q;
  const float* row = src + (size_t)a * q;
  if (off == 0) {
    stage0_load<R0>(sm, L, lt, nthr, tw, [&](int e) {
      const int c = 2 * e;
      if (!row_ok || c >= q) return make_float2(0.f, 0.f);
      if (c + 1 < q) return *reinterpret_cast<const float2*>(row + c);
      return make_float2(row[c], 0.f);
    });
  } else {
    stage0_load<R0>(sm, L, lt, nthr, tw, [&](int e) {
      if (!row_ok) return make_float2(0.f, 0.f);
      const int b0 = (2 * e + off) & pmask, b1 = (2 * e + 1 + off) & pmask;
      return make_float2(b0 < q ? row[b0] : 0.f, b1 < q ? row[b1] : 0.f);
    });
  }
  __syncthreads();
  mid_forward(sm, L, plan_stages(mL), 1, plan_stages(mL), lt, nthr, tw);
  if (!live) return;
  const int step = kFftMaxLen / P;
  for (int k = lt; k <= L; k += nthr) {
    const float2 zk = sm[pad(perm_pos(k & (L - 1), mL))];
    const float2 zl = cconj(sm[pad(perm_pos((L - k) & (L - 1), mL))]);
    const float2 e = cscale(cadd(zk, zl), 0.5f);
    const float2 dd = cscale(csub(zk, zl), 0.5f);
    const float2 o = make_float2(dd.y, -dd.x);  // (zk - conj zl) / (2i)
    T[(size_t)k * ldT + fr] = cadd(e, cmul(tw[k * step], o));
  }
}

// ---- pass 2b: column forward FFT only (PSF spectrum), stored in the digit-reversed order
template <int R0>
__global__ __launch_bounds__(kThreads) void k_cols_fwd(const float2* __restrict__ T1, int pm, float2* __restrict__ S,
                                                       int cpb, const float2* __restrict__ tw) {
  extern __shared__ float2 smem[];
  const int P = 1 << pm;
  const int ncols = (P >> 1) + 1;
  const int nthr = kThreads / cpb;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int k = blockIdx.x * cpb + g;
  const bool live = k < ncols;
  float2* sm = smem + g * (P + (P >> 5));
  const float2* col = T1 + (size_t)k * P;
  stage0_load<R0>(sm, P, lt, nthr, tw, [&](int e) { return live ? col[e] : make_float2(0.f, 0.f); });
  __syncthreads();
  mid_forward(sm, P, plan_stages(pm), 1, plan_stages(pm), lt, nthr, tw);
  if (!live) return;
  float2* out = S + (size_t)k * P;
  for (int i = lt; i < P; i += nthr) out[i] = sm[pad(i)];
}

// =============================================================================================
// Convolution passes with compile-time plans.  For a frame of side P = 2^PG the radix plan, every
// LDS offset inside a butterfly and each thread's share of the work are constants, so the only
// per-butterfly integer work is one base address.  Every global operand a thread needs (stage-0
// inputs, in-stage twiddles, the PSF segment it multiplies with, the real-FFT post-processing
// twiddles) is requested at kernel start, so a kernel pays one memory latency instead of one per
// stage.  LIN marks the linear-convolution frame (P = 2N): half of every transform input is zero
// padding, and only half of every inverse output is kept.  Storage order, twiddles and arithmetic
// order per butterfly are those of the generic kernels above (same digit-reversed layout).

// LDS slot of element i: one pad per 32 and one more per 256 elements.  The second term spreads the
// digit-reversed (stride-256) accesses of the real-FFT split over all banks; both terms are linear
// over every butterfly's elements (base + k*M never carries into them), so a butterfly's offsets
// stay compile-time constants.  span() is a transform's LDS length, odd so that the transforms of
// one block start on different banks.
__host__ __device__ constexpr int padc(int i) { return i + (i >> 5) + (i >> 8); }
__host__ __device__ constexpr int span(int n) { return padc(n) + 1; }

template <int LG>
struct Plan {
  static constexpr int L = 1 << LG;
  static constexpr int NS = (LG + 3) >> 2;
  static constexpr int LR0 = LG - 4 * (NS - 1);
  static constexpr int R0 = 1 << LR0;
  static constexpr int M0 = L / R0;
  static constexpr int NBF = NS > 1 ? L / 16 : 0;  // radix-16 butterflies per stage
  static constexpr int ls(int i) { return 1 << (4 * (NS - i)); }
};

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <int B, int E, class F>
__device__ __forceinline__ void static_for_down(F&& f) {  // E-1 down to B
  if constexpr (B < E) {
    f(std::integral_constant<int, E - 1>{});
    static_for_down<B, E - 1>(f);
  }
}

// This thread's in-stage twiddles for the radix-16 stages 1..NS-1 (butterflies b = lt + j*NT).
template <int LG, int NT>
struct StageTw {
  using Pl = Plan<LG>;
  static constexpr int PER = Pl::NBF > 0 ? (Pl::NBF + NT - 1) / NT : 1;
  static constexpr int NI = Pl::NS > 1 ? Pl::NS : 1;
  float2 w[NI][PER];
  __device__ __forceinline__ void load(int lt, const float2* __restrict__ gtw) {
    static_for<1, Pl::NS>([&](auto ic) {
      constexpr int I = decltype(ic)::value;
      constexpr int Ls = Pl::ls(I);
      constexpr int M = Ls / 16;
      if constexpr (M > 1) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          const int b = lt + j * NT;
          w[I][j] = (b < Pl::NBF) ? gtw[(b & (M - 1)) * (kFftMaxLen / Ls)] : make_float2(1.f, 0.f);
        }
      }
    });
  }
};

// radix-16 stage I of one transform in LDS (sm), forward (DIF, twiddles after the butterfly) or
// inverse (the mirror: conjugate twiddles before the butterfly)
template <int LG, int NT, int I, bool FWD>
__device__ __forceinline__ void stage16(float2* sm, int lt, const StageTw<LG, NT>& tw) {
  using Pl = Plan<LG>;
  constexpr int Ls = Pl::ls(I);
  constexpr int M = Ls / 16;
  constexpr int LM = 4 * (Pl::NS - I - 1);
  constexpr int PER = StageTw<LG, NT>::PER;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int b = lt + j * NT;
    if ((Pl::NBF % NT) != 0 && b >= Pl::NBF) break;
    const int base = ((b >> LM) << (LM + 4)) + (b & (M - 1));
    float2* s0 = sm + padc(base);
    float2 x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = s0[padc(k * M)];
    if constexpr (FWD) {
      Dft<16, -1>::run(x);
      if constexpr (M > 1) {
        float2 wk[16];
        powers<16>(tw.w[I][j], wk);
#pragma unroll
        for (int k = 1; k < 16; ++k) x[k] = cmul(x[k], wk[k]);
      }
    } else {
      if constexpr (M > 1) {
        float2 wk[16];
        powers<16>(cconj(tw.w[I][j]), wk);
#pragma unroll
        for (int k = 1; k < 16; ++k) x[k] = cmul(x[k], wk[k]);
      }
      Dft<16, +1>::run(x);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) s0[padc(k * M)] = x[k];
  }
}

// stage 0 (radix R0) inputs of this thread: butterflies n = lt + u*NT < M0, elements n + k*M0;
// KL of the R0 inputs are non-zero (LIN: the upper half is zero padding)
template <int LG, int NT, bool LIN>
struct Stage0 {
  using Pl = Plan<LG>;
  static constexpr int U = (Pl::M0 + NT - 1) / NT;
  static constexpr int KL = LIN ? (Pl::R0 / 2 > 0 ? Pl::R0 / 2 : 1) : Pl::R0;
  float2 x[U][Pl::R0];
  float2 w[U];
  template <class Get>
  __device__ __forceinline__ void load(int lt, const float2* __restrict__ gtw, Get get) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = lt + u * NT;
      const bool ok = n < Pl::M0;
#pragma unroll
      for (int k = 0; k < Pl::R0; ++k) x[u][k] = (ok && k < KL) ? get(n + k * Pl::M0) : make_float2(0.f, 0.f);
      if constexpr (Pl::M0 > 1) w[u] = ok ? gtw[n * (kFftMaxLen / Pl::L)] : make_float2(1.f, 0.f);
    }
  }
  // forward stage 0 from registers into LDS
  __device__ __forceinline__ void forward(float2* sm, int lt) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = lt + u * NT;
      if ((Pl::M0 % NT) != 0 && n >= Pl::M0) break;
      Dft<Pl::R0, -1>::run(x[u]);
      if constexpr (Pl::M0 > 1) {
        float2 wk[Pl::R0];
        powers<Pl::R0>(w[u], wk);
#pragma unroll
        for (int k = 1; k < Pl::R0; ++k) x[u][k] = cmul(x[u][k], wk[k]);
      }
      float2* s0 = sm + padc(n);
#pragma unroll
      for (int k = 0; k < Pl::R0; ++k) s0[padc(k * Pl::M0)] = x[u][k];
    }
  }
  // inverse last stage from LDS to put(e, value), keeping the first KL outputs of each butterfly
  template <class Put>
  __device__ __forceinline__ void inverse(const float2* sm, int lt, Put put) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = lt + u * NT;
      if ((Pl::M0 % NT) != 0 && n >= Pl::M0) break;
      const float2* s0 = sm + padc(n);
      float2 y[Pl::R0];
#pragma unroll
      for (int k = 0; k < Pl::R0; ++k) y[k] = s0[padc(k * Pl::M0)];
      if constexpr (Pl::M0 > 1) {
        float2 wk[Pl::R0];
        powers<Pl::R0>(cconj(w[u]), wk);
#pragma unroll
        for (int k = 1; k < Pl::R0; ++k) y[k] = cmul(y[k], wk[k]);
      }
      Dft<Pl::R0, +1>::run(y);
#pragma unroll
      for (int k = 0; k < KL; ++k) put(n + k * Pl::M0, y[k]);
    }
  }
};

// points per thread in the row and column transforms (measured at P = 8192: 32 points per thread
// beats 16, whose extra threads buy occupancy but not speed)
#ifndef PMS_FFT_PTS
#define PMS_FFT_PTS 32
#endif
constexpr int kFftPts = PMS_FFT_PTS;

// launch shapes, functions of the frame only
// Rows: NT threads per row (at most kFftPts points each, at least one wave), RB rows per block as many as
// 512 threads, ~140 KB of LDS and >= 256 blocks allow: more rows per block make longer contiguous
// runs in the transposed (column-major) writes and reads.
template <int PG, bool LIN>
struct RowCfg {
  static constexpr int P = 1 << PG, L = P / 2, N = LIN ? L : P;  // N = rows = image side
  static constexpr int NT = (L / kFftPts > 64) ? L / kFftPts : 64;
  static constexpr int rb() {
    int r = 16;
    while (r > 1 && (r * NT > 512 || (size_t)r * span(L) * sizeof(float2) > 140 * 1024 || N / r < 256))
      r >>= 1;
    return r;
  }
  static constexpr int RB = rb();
  static constexpr int T = RB * NT;
  static constexpr int NPP = (RB * (L / 2 + 1) + T - 1) / T;  // real-FFT pairs per thread
  static constexpr size_t LDS = (size_t)RB * span(L) * sizeof(float2);
};

template <int PG>
struct ColCfg {
  static constexpr int P = 1 << PG;
  static constexpr int cpb() {
    int c = 8192 / P;
    if (c > 16) c = 16;
    if (c < 1) c = 1;
    while (c > 1 && (P / 2 + 1) / c < 256) c >>= 1;
    return c;
  }
  static constexpr int CPB = cpb();
  static constexpr int NT = (P / (256 / CPB) > kFftPts) ? P / kFftPts : 256 / CPB;
  static constexpr int T = CPB * NT;
  static constexpr size_t LDS = (size_t)CPB * span(P) * sizeof(float2);
};

// storage position of frequency k (compile-time plan of length 2^LG)
template <int LG>
__device__ __forceinline__ int pos_of(int k) {
  using Pl = Plan<LG>;
  int mM = LG - Pl::LR0;
  int pos = (k & (Pl::R0 - 1)) << mM;
  k >>= Pl::LR0;
#pragma unroll
  for (int i = 1; i < Pl::NS; ++i) {
    mM -= 4;
    pos += (k & 15) << mM;
    k >>= 4;
  }
  return pos;
}

// ---- pass 1: rows r2c.  Block = RB frame rows; output T[k*ldT + row], k in [0, P/2].
template <int PG, bool LIN>
__global__ __launch_bounds__((RowCfg<PG, LIN>::T)) void k_conv_r2c(R2cBatch bt, int n_rows, int ldT, const float2* __restrict__ tw) {
  using C = RowCfg<PG, LIN>;
  constexpr int LG = PG - 1;
  constexpr int L = C::L, RB = C::RB, NT = C::NT, SP = span(L);
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  const float* __restrict__ src = bt.src[jb];
  float2* __restrict__ T = bt.T[jb];
  extern __shared__ float2 smem[];
  const int g = threadIdx.x / NT;
  const int lt = threadIdx.x - g * NT;
  const int row0 = xcd_remap(blockIdx.x, gridDim.x) * RB;
  const int fr = row0 + g;
  const bool live = fr < n_rows;
  float2* sm = smem + g * SP;
  // everything global this thread needs, requested up front
  const float* row = src + (size_t)(live ? fr : 0) * C::N;
  Stage0<LG, NT, LIN> s0;
  s0.load(lt, tw, [&](int e) { return live ? *reinterpret_cast<const float2*>(row + 2 * e) : make_float2(0.f, 0.f); });
  StageTw<LG, NT> stw;
  stw.load(lt, tw);
  constexpr int step = kFftMaxLen / (2 * L);
  float2 wpp[C::NPP];
#pragma unroll
  for (int i = 0; i < C::NPP; ++i) {
    const int idx = threadIdx.x + i * C::T;
    const int k = idx / RB;
    wpp[i] = (k <= L / 2) ? tw[k * step] : make_float2(1.f, 0.f);
  }
  s0.forward(sm, lt);
  __syncthreads();
  static_for<1, Plan<LG>::NS>([&](auto ic) {
    stage16<LG, NT, decltype(ic)::value, true>(sm, lt, stw);
    __syncthreads();
  });
  // Hermitian split of the packed transform: X[k], X[L-k] from Z[k], Z[L-k]; consecutive threads
  // take consecutive rows of one k so the transposed writes pair up
#pragma unroll
  for (int i = 0; i < C::NPP; ++i) {
    const int idx = threadIdx.x + i * C::T;
    const int k = idx / RB;
    const int g2 = idx - k * RB;
    const int r2 = row0 + g2;
    if (k > L / 2 || r2 >= n_rows) continue;
    const float2* sm2 = smem + g2 * SP;
    const int pk = pos_of<LG>(k & (L - 1)), pl = pos_of<LG>((L - k) & (L - 1));
    const float2 zk = sm2[padc(pk)];
    const float2 zl = sm2[padc(pl)];
    const float2 w = wpp[i];
    {
      const float2 c = cconj(zl);
      const float2 e = cscale(cadd(zk, c), 0.5f);
      const float2 dd = cscale(csub(zk, c), 0.5f);
      T[(size_t)k * ldT + r2] = cadd(e, cmul(w, make_float2(dd.y, -dd.x)));
    }
    if (k != L - k) {
      const float2 c = cconj(zk);
      const float2 e = cscale(cadd(zl, c), 0.5f);
      const float2 dd = cscale(csub(zl, c), 0.5f);
      const float2 wl = make_float2(-w.x, w.y);  // w^(L-k) = -conj(w^k)
      T[(size_t)(L - k) * ldT + r2] = cadd(e, cmul(wl, make_float2(dd.y, -dd.x)));
    }
  }
}

// ---- pass 2: column convolution: forward FFT, product with the PSF column, inverse FFT.
template <int PG, bool LIN>
__global__ __launch_bounds__(ColCfg<PG>::T) void k_conv_cols(ColBatch bt, int n_rows, const float2* __restrict__ tw) {
  using C = ColCfg<PG>;
  using Pl = Plan<PG>;
  constexpr int P = C::P, NT = C::NT, SP = span(P);
  constexpr int NCOLS = P / 2 + 1;
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  extern __shared__ float2 smem[];
  const int g = threadIdx.x / NT;
  const int lt = threadIdx.x - g * NT;
  const int k = blockIdx.x * C::CPB + g;
  const bool live = k < NCOLS;
  float2* sm = smem + g * SP;
  const float2* __restrict__ col = bt.T1[jb] + (size_t)(live ? k : 0) * n_rows;
  const float2* __restrict__ sc = bt.S[jb] + (size_t)(live ? k : 0) * P;
  float2* __restrict__ out = bt.T2[jb] + (size_t)(live ? k : 0) * n_rows;
  if constexpr (Pl::NS == 1) {
    // whole transform in one thread's registers (P <= 16)
    if (!live || lt != 0) return;
    float2 x[P];
#pragma unroll
    for (int e = 0; e < P; ++e) x[e] = (!LIN || e < P / 2) ? col[e] : make_float2(0.f, 0.f);
    Dft<P, -1>::run(x);
#pragma unroll
    for (int e = 0; e < P; ++e) x[e] = cmul(x[e], sc[e]);
    Dft<P, +1>::run(x);
#pragma unroll
    for (int e = 0; e < (LIN ? P / 2 : P); ++e) out[e] = x[e];
  } else {
    Stage0<PG, NT, LIN> s0;
    // unconditional loads (a dead block reads column 0 and stores nothing): loads under a branch
    // leave the compiler's wait counts conservative, so that stage 0 would also wait for the PSF
    // segment requested after its inputs
    s0.load(lt, tw, [&](int e) { return col[e]; });
    StageTw<PG, NT> stw;
    stw.load(lt, tw);
    constexpr int PER = StageTw<PG, NT>::PER;
    float2 sv[PER][16];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int b = lt + j * NT;
#pragma unroll
      for (int m = 0; m < 16; ++m) sv[j][m] = (b < Pl::NBF) ? sc[16 * b + m] : make_float2(0.f, 0.f);
    }
    s0.forward(sm, lt);
    __syncthreads();
    static_for<1, Pl::NS - 1>([&](auto ic) {
      stage16<PG, NT, decltype(ic)::value, true>(sm, lt, stw);
      __syncthreads();
    });
    // last forward stage (sub-length 16), PSF product, first inverse stage, in registers
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int b = lt + j * NT;
      if ((Pl::NBF % NT) != 0 && b >= Pl::NBF) break;
      const int base = 16 * b;
      float2* s1 = sm + padc(base);
      float2 x[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) x[m] = s1[m];
      Dft<16, -1>::run(x);
#pragma unroll
      for (int m = 0; m < 16; ++m) x[m] = cmul(x[m], sv[j][m]);
      Dft<16, +1>::run(x);
#pragma unroll
      for (int m = 0; m < 16; ++m) s1[m] = x[m];
    }
    __syncthreads();
    static_for_down<1, Pl::NS - 1>([&](auto ic) {
      stage16<PG, NT, decltype(ic)::value, false>(sm, lt, stw);
      __syncthreads();
    });
    if (!live) return;
    s0.inverse(sm, lt, [&](int e, float2 v) { out[e] = v; });
  }
}

// ---- pass 3: rows c2r.  Block = RB kept rows; out[r * N + c].
template <int PG, bool LIN>
__global__ __launch_bounds__((RowCfg<PG, LIN>::T)) void k_conv_c2r(C2rBatch bt, int n_rows, int ldT, float scale,
                                                  const float2* __restrict__ tw) {
  using C = RowCfg<PG, LIN>;
  constexpr int LG = PG - 1;
  constexpr int L = C::L, RB = C::RB, NT = C::NT, SP = span(L);
  const int jb = blockIdx.y;
  const int* pred = bt.pred[jb];
  if (pred && *pred == 0) return;
  const float2* __restrict__ T2 = bt.T2[jb];
  float* __restrict__ out = bt.out[jb];
  extern __shared__ float2 smem[];
  const int g = threadIdx.x / NT;
  const int lt = threadIdx.x - g * NT;
  const int row0 = xcd_remap(blockIdx.x, gridDim.x) * RB;
  const int r = row0 + g;
  const bool live = r < n_rows;
  float2* sm = smem + g * SP;
  constexpr int step = kFftMaxLen / (2 * L);
  float2 xk[C::NPP], xl[C::NPP], wpp[C::NPP];
#pragma unroll
  for (int i = 0; i < C::NPP; ++i) {
    const int idx = threadIdx.x + i * C::T;
    const int k = idx / RB;
    const int r2 = row0 + (idx - k * RB);
    const bool ok = k <= L / 2 && r2 < n_rows;
    xk[i] = ok ? T2[(size_t)k * ldT + r2] : make_float2(0.f, 0.f);
    xl[i] = ok ? T2[(size_t)(L - k) * ldT + r2] : make_float2(0.f, 0.f);
    wpp[i] = ok ? tw[k * step] : make_float2(1.f, 0.f);
  }
  StageTw<LG, NT> stw;
  stw.load(lt, tw);
  Stage0<LG, NT, LIN> s0;
  if constexpr (Plan<LG>::M0 > 1) {
#pragma unroll
    for (int u = 0; u < Stage0<LG, NT, LIN>::U; ++u) {
      const int n = lt + u * NT;
      s0.w[u] = (n < Plan<LG>::M0) ? tw[n * (kFftMaxLen / L)] : make_float2(1.f, 0.f);
    }
  }
  // Hermitian pre-twiddle: Z'[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) conj(w^k), stored
  // digit-reversed for the inverse transform
#pragma unroll
  for (int i = 0; i < C::NPP; ++i) {
    const int idx = threadIdx.x + i * C::T;
    const int k = idx / RB;
    const int g2 = idx - k * RB;
    if (k > L / 2 || row0 + g2 >= n_rows) continue;
    float2* sm2 = smem + g2 * SP;
    const float2 w = wpp[i];
    {
      const float2 e = cadd(xk[i], cconj(xl[i]));
      const float2 o = cmul(csub(xk[i], cconj(xl[i])), cconj(w));
      const int pk = pos_of<LG>(k & (L - 1));
      sm2[padc(pk)] = make_float2(e.x - o.y, e.y + o.x);
    }
    const int kl = L - k;
    if (kl != k && kl < L) {
      const float2 e = cadd(xl[i], cconj(xk[i]));
      const float2 o = cmul(csub(xl[i], cconj(xk[i])), make_float2(-w.x, -w.y));  // conj(w^(L-k)) = -w^k
      const int pl = pos_of<LG>(kl);
      sm2[padc(pl)] = make_float2(e.x - o.y, e.y + o.x);
    }
  }
  __syncthreads();
  static_for_down<1, Plan<LG>::NS>([&](auto ic) {
    stage16<LG, NT, decltype(ic)::value, false>(sm, lt, stw);
    __syncthreads();
  });
  if (!live) return;
  float* orow = out + (size_t)r * C::N;
  s0.inverse(sm, lt, [&](int e, float2 v) {
    *reinterpret_cast<float2*>(orow + 2 * e) = make_float2(v.x * scale, v.y * scale);
  });
}

template <int PG, bool LIN>
void set_conv_attrs(int lds_max) {
  (void)hipFuncSetAttribute((const void*)k_conv_r2c<PG, LIN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_conv_cols<PG, LIN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_conv_c2r<PG, LIN>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
}

template <int PG, bool LIN>
void launch_conv(const R2cBatch& rb, const ColBatch& cb, const C2rBatch& ob, int nb, int n, float scale,
                 const float2* tw, hipStream_t st) {
  using R = RowCfg<PG, LIN>;
  using C = ColCfg<PG>;
  hipLaunchKernelGGL((k_conv_r2c<PG, LIN>), dim3((n + R::RB - 1) / R::RB, nb), dim3(R::T), R::LDS, st, rb, n, n, tw);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_conv_cols<PG, LIN>), dim3((C::P / 2 + 1 + C::CPB - 1) / C::CPB, nb), dim3(C::T), C::LDS, st,
                     cb, n, tw);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_conv_c2r<PG, LIN>), dim3((n + R::RB - 1) / R::RB, nb), dim3(R::T), R::LDS, st, ob, n, n,
                     scale, tw);
  PMS_LAUNCH_CHECK();
}

template <bool LIN>
void dispatch_conv(int pg, const R2cBatch& rb, const ColBatch& cb, const C2rBatch& ob, int nb, int n, float scale,
                   const float2* tw, hipStream_t st) {
  switch (pg) {
    case 2: launch_conv<2, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 3: launch_conv<3, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 4: launch_conv<4, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 5: launch_conv<5, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 6: launch_conv<6, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 7: launch_conv<7, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 8: launch_conv<8, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 9: launch_conv<9, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 10: launch_conv<10, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 11: launch_conv<11, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 12: launch_conv<12, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 13: launch_conv<13, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    case 14: launch_conv<14, LIN>(rb, cb, ob, nb, n, scale, tw, st); break;
    default: fail(PMS_ERR_ARG, "FFT frame out of range");
  }
}

template <bool LIN>
void set_all_conv_attrs(int lds_max) {
  set_conv_attrs<2, LIN>(lds_max);
  set_conv_attrs<3, LIN>(lds_max);
  set_conv_attrs<4, LIN>(lds_max);
  set_conv_attrs<5, LIN>(lds_max);
  set_conv_attrs<6, LIN>(lds_max);
  set_conv_attrs<7, LIN>(lds_max);
  set_conv_attrs<8, LIN>(lds_max);
  set_conv_attrs<9, LIN>(lds_max);
  set_conv_attrs<10, LIN>(lds_max);
  set_conv_attrs<11, LIN>(lds_max);
  set_conv_attrs<12, LIN>(lds_max);
  set_conv_attrs<13, LIN>(lds_max);
  set_conv_attrs<14, LIN>(lds_max);
}

struct DeviceTwiddles {
  float2* ptr = nullptr;
};
std::mutex g_tw_mu;
DeviceTwiddles g_tw[64];

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

// Rows / columns per workgroup: as many as the LDS budget allows for wide coalesced row accesses,
// but never so many that the launch has fewer than ~256 workgroups (small frames stay parallel).
int rows_per_block(int L, int n_rows) {
  static const int lim = env_int("PMS_FFT_ROW_LDS", 8192);
  int rb = lim / (L > 0 ? L : 1);
  if (rb > 16) rb = 16;
  while (rb > 1 && n_rows / rb < 256) rb >>= 1;
  if (rb < 1) rb = 1;
  return rb;
}

int cols_per_block(int P) {
  const int ncols = P / 2 + 1;
  int c = 8192 / P;
  if (c > 16) c = 16;
  while (c > 1 && ncols / c < 256) c >>= 1;
  if (c < 1) c = 1;
  return c;
}

template <int R0>
void set_lds_attrs(int lds_max) {
  (void)hipFuncSetAttribute((const void*)k_rows_r2c<R0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_cols_fwd<R0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
}

// first-stage radix of the plan for length 2^m
int first_radix(int m) { return 1 << plan_first_log2(m); }

#define PMS_FFT_DISPATCH(R, KERNEL, ...)                                          \
  switch (R) {                                                                    \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                    \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                    \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                    \
    default: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;                  \
  }

void launch_rows_r2c(const R2cBatch& bt, int count, int q, int off, int pm, int n_rows, int ldT, const float2* tw,
                     hipStream_t st) {
  const int L = (1 << pm) / 2;
  const int rb = rows_per_block(L, n_rows);
  const size_t lds = (size_t)rb * (L + (L >> 5)) * sizeof(float2);
  PMS_FFT_DISPATCH(first_radix(pm - 1), k_rows_r2c, dim3(ceil_div(n_rows, rb), count), dim3(kThreads), lds, st, bt,
                   q, off, pm, n_rows, rb, ldT, tw);
  PMS_LAUNCH_CHECK();
}


}  // namespace

void fft_init_device(int device) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  PMS_REQUIRE(device >= 0 && device < 64, "device id out of range");
  if (g_tw[device].ptr) return;
  std::vector<float2> h(kFftMaxLen);
  for (int k = 0; k < kFftMaxLen; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)kFftMaxLen;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  // allow the FFT kernels the large dynamic LDS (67.6 KB at P = 8192) that CDNA4's 160 KB LDS permits
  const int lds_max = 160 * 1024;
  set_lds_attrs<2>(lds_max);
  set_lds_attrs<4>(lds_max);
  set_lds_attrs<8>(lds_max);
  set_lds_attrs<16>(lds_max);
  set_all_conv_attrs<true>(lds_max);
  set_all_conv_attrs<false>(lds_max);
  (void)hipGetLastError();
  float2* d = nullptr;
  PMS_HIP(hipMalloc(&d, sizeof(float2) * kFftMaxLen));
  PMS_HIP(hipMemcpy(d, h.data(), sizeof(float2) * kFftMaxLen, hipMemcpyHostToDevice));
  g_tw[device].ptr = d;
}

const float2* fft_twiddles() {
  int dev = 0;
  PMS_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (g_tw[dev].ptr) return g_tw[dev].ptr;
  }
  fft_init_device(dev);
  return g_tw[dev].ptr;
}

int fft_perm(int k, int p) { return is_pow2(p) ? perm_pos(k, ilog2(p)) : gfft_perm(k, p); }

size_t fft_spectrum_elems(int p) { return (size_t)(p / 2 + 1) * p; }

size_t fft_conv_work_elems(int n, int p) { return 2 * (size_t)(p / 2 + 1) * n; }

static void check_fft_len(int p) {
  PMS_REQUIRE(is_pow2(p) && p >= 4 && p <= kFftMaxLen,
              "FFT frame side must be a power of two in [4, " + std::to_string(kFftMaxLen) +
                  "], got " + std::to_string(p));
}

void fft_psf_spectrum(const float* psf, int q, int p, float2* spec, float2* work, hipStream_t st) {
  if (!is_pow2(p)) {
    // any other even side: the mixed-radix transforms (same layout, digit-reversed columns)
    gfft_psf_spectrum<float, float>(psf, q, p, reinterpret_cast<Cx<float>*>(spec), reinterpret_cast<Cx<float>*>(work),
                                    st);
    return;
  }
  check_fft_len(p);
  PMS_REQUIRE(q >= 1 && q <= p, "PSF larger than the FFT frame");
  const float2* tw = fft_twiddles();
  const int pm = ilog2(p);
  const int L = p / 2;
  // work holds the (P/2+1) x P row spectra, column-major
  R2cBatch rb{};
  rb.src[0] = psf;
  rb.T[0] = work;
  launch_rows_r2c(rb, 1, q, q / 2, pm, p, p, tw, st);
  const int cpb = cols_per_block(p);
  const size_t lds2 = (size_t)cpb * (p + (p >> 5)) * sizeof(float2);
  PMS_FFT_DISPATCH(first_radix(pm), k_cols_fwd, dim3(ceil_div(L + 1, cpb)), dim3(kThreads), lds2, st,
                   (const float2*)work, pm, spec, cpb, tw);
  PMS_LAUNCH_CHECK();
}

void fft_conv(const float* in, int n, int p, const float2* spec, float* out, float2* work, const int* pred,
              hipStream_t st) {
  const FftConvJob j{in, spec, out, work, pred};
  fft_conv_batch(&j, 1, n, p, st);
}

// Jobs are independent: job j's blocks read and write only job j's buffers, so a batch gives each
// job exactly the result of its own single-job call.
void fft_conv_batch(const FftConvJob* jobs, int count, int n, int p, hipStream_t st) {
  if (!is_pow2(p)) {
    std::vector<GConvJob<float>> g(count);
    for (int k = 0; k < count; ++k)
      g[k] = GConvJob<float>{jobs[k].in, reinterpret_cast<const Cx<float>*>(jobs[k].spec), jobs[k].out,
                             reinterpret_cast<Cx<float>*>(jobs[k].work), jobs[k].pred};
    gfft_conv_batch<float>(g.data(), count, n, p, st);
    return;
  }
  check_fft_len(p);
  PMS_REQUIRE(p == n || p == 2 * n, "fft_conv: frame must be N (circular) or 2N (linear)");
  const float2* tw = fft_twiddles();
  const int pm = ilog2(p);
  const int L = p / 2;
  const float scale = 1.0f / ((float)p * (float)p);
  for (int j0 = 0; j0 < count; j0 += kMaxBatch) {
    const int nb = std::min(kMaxBatch, count - j0);
    R2cBatch rb{};
    ColBatch cb{};
    C2rBatch ob{};
    for (int k = 0; k < nb; ++k) {
      const FftConvJob& j = jobs[j0 + k];
      PMS_REQUIRE(j.in && j.spec && j.out && j.work, "fft_conv: NULL buffer");
      float2* T1 = j.work;
      float2* T2 = j.work + (size_t)(L + 1) * n;
      rb.src[k] = j.in;
      rb.T[k] = T1;
      rb.pred[k] = j.pred;
      cb.T1[k] = T1;
      cb.S[k] = j.spec;
      cb.T2[k] = T2;
      cb.pred[k] = j.pred;
      ob.T2[k] = T2;
      ob.out[k] = j.out;
      ob.pred[k] = j.pred;
    }
    if (p == 2 * n)
      dispatch_conv<true>(pm, rb, cb, ob, nb, n, scale, tw, st);
    else
      dispatch_conv<false>(pm, rb, cb, ob, nb, n, scale, tw, st);
  }
}

// numpy rfft2 spectrum (rows = axis-0 frequency m in [0,P), cols = k in [0,P/2]) of an
// unrolled PSF frame -> library layout: column k, position perm(m), times (-1)^(m+k)
// (the P/2 roll in both axes).
void fft_spectrum_from_numpy_host(const float* re_im, int p, float2* spec_host) {
  if (is_pow2(p))
    check_fft_len(p);
  else
    PMS_REQUIRE(gfft_frame_ok(p, sizeof(float)), "FFT frame side " + std::to_string(p) + " is not supported");
  const int nk = p / 2 + 1;
  for (int m = 0; m < p; ++m) {
    const int pos = fft_perm(m, p);
    for (int k = 0; k < nk; ++k) {
      const float sgn = ((m + k) & 1) ? -1.f : 1.f;
      const float* v = re_im + 2 * ((size_t)m * nk + k);
      spec_host[(size_t)k * p + pos] = make_float2(sgn * v[0], sgn * v[1]);
    }
  }
}

}  // namespace pms
