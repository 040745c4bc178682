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
#include <cmath>
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

// forward: natural order in LDS -> digit-reversed order in LDS (all groups in lockstep)
__device__ __forceinline__ void fft_forward(float2* sm, int m, int lt, int nthr, const float2* tw) {
  if (m <= 0) return;
  const int ns = plan_stages(m);
  const int L = 1 << m;
  run_stage<true>(plan_first_log2(m), sm, L, L, lt, nthr, tw);
  __syncthreads();
  for (int i = 1; i < ns; ++i) {
    run_stage<true>(4, sm, L, 1 << (4 * (ns - i)), lt, nthr, tw);
    __syncthreads();
  }
}

// inverse (unnormalised): digit-reversed order in LDS -> natural order in LDS
__device__ __forceinline__ void fft_inverse(float2* sm, int m, int lt, int nthr, const float2* tw) {
  if (m <= 0) return;
  const int ns = plan_stages(m);
  const int L = 1 << m;
  for (int i = ns - 1; i >= 1; --i) {
    run_stage<false>(4, sm, L, 1 << (4 * (ns - i)), lt, nthr, tw);
    __syncthreads();
  }
  run_stage<false>(plan_first_log2(m), sm, L, L, lt, nthr, tw);
  __syncthreads();
}

// storage position of frequency k after fft_forward of length 2^m:
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

// ---- pass 1: rows r2c.  Group g of the block handles frame row fr; output T[k*ldT + fr], k in [0, P/2].
__global__ __launch_bounds__(kThreads) void k_rows_r2c(const float* __restrict__ src, int q, int off, int pm,
                                                       int n_rows, int rb, float2* __restrict__ T, int ldT,
                                                       const float2* __restrict__ tw, const int* pred) {
  if (pred && *pred == 0) return;
  extern __shared__ float2 smem[];
  const int P = 1 << pm;
  const int L = P >> 1;
  const int mL = pm - 1;
  const int nthr = kThreads / rb;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int fr = blockIdx.x * rb + g;
  float2* sm = smem + g * (L + (L >> 5));
  const bool live = fr < n_rows;
  for (int n = lt; n < L; n += nthr) {
    float2 z = make_float2(0.f, 0.f);
    if (live) z = make_float2(frame_val(src, q, off, P - 1, fr, 2 * n), frame_val(src, q, off, P - 1, fr, 2 * n + 1));
    sm[pad(n)] = z;
  }
  __syncthreads();
  fft_forward(sm, mL, lt, nthr, tw);
  if (!live) return;
  const int step = kFftMaxLen / P;
  for (int k = lt; k <= L; k += nthr) {
    const float2 zk = sm[pad(perm_pos(k & (L - 1), mL))];
    const float2 zl = cconj(sm[pad(perm_pos((L - k) & (L - 1), mL))]);
    const float2 e = cscale(cadd(zk, zl), 0.5f);
    const float2 dd = cscale(csub(zk, zl), 0.5f);
    const float2 o = make_float2(dd.y, -dd.x);  // (zk - conj zl) / (2i)
    const float2 w = tw[k * step];
    T[(size_t)k * ldT + fr] = cadd(e, cmul(w, o));
  }
}

// ---- pass 2a: column convolution.  Group g handles frequency column k.
__global__ __launch_bounds__(kThreads) void k_cols_conv(const float2* __restrict__ T1, int n_in, int pm,
                                                        const float2* __restrict__ S, float2* __restrict__ T2,
                                                        int n_out, int cpb, const float2* __restrict__ tw,
                                                        const int* pred) {
  if (pred && *pred == 0) return;
  extern __shared__ float2 smem[];
  const int P = 1 << pm;
  const int ncols = (P >> 1) + 1;
  const int nthr = kThreads / cpb;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int k = blockIdx.x * cpb + g;
  const bool live = k < ncols;
  float2* sm = smem + g * (P + (P >> 5));
  const float2* col = T1 + (size_t)k * n_in;
  for (int i = lt; i < P; i += nthr) sm[pad(i)] = (live && i < n_in) ? col[i] : make_float2(0.f, 0.f);
  __syncthreads();
  fft_forward(sm, pm, lt, nthr, tw);
  if (live) {
    const float2* sc = S + (size_t)k * P;
    for (int i = lt; i < P; i += nthr) sm[pad(i)] = cmul(sm[pad(i)], sc[i]);
  }
  __syncthreads();
  fft_inverse(sm, pm, lt, nthr, tw);
  if (!live) return;
  float2* out = T2 + (size_t)k * n_out;
  for (int i = lt; i < n_out; i += nthr) out[i] = sm[pad(i)];
}

// ---- pass 2b: column forward FFT only (PSF spectrum)
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
  for (int i = lt; i < P; i += nthr) sm[pad(i)] = live ? col[i] : make_float2(0.f, 0.f);
  __syncthreads();
  fft_forward(sm, pm, lt, nthr, tw);
  if (!live) return;
  float2* out = S + (size_t)k * P;
  for (int i = lt; i < P; i += nthr) out[i] = sm[pad(i)];
}

// ---- pass 3: rows c2r.  Group g handles kept row r; out[r * n_cols + c], c < n_cols.
__global__ __launch_bounds__(kThreads) void k_rows_c2r(const float2* __restrict__ T2, int ldT, int pm, int n_rows,
                                                       int rb, float* __restrict__ out, int n_cols, float scale,
                                                       const float2* __restrict__ tw, const int* pred) {
  if (pred && *pred == 0) return;
  extern __shared__ float2 smem[];
  const int P = 1 << pm;
  const int L = P >> 1;
  const int mL = pm - 1;
  const int nthr = kThreads / rb;
  const int g = threadIdx.x / nthr;
  const int lt = threadIdx.x - g * nthr;
  const int r = blockIdx.x * rb + g;
  const bool live = r < n_rows;
  float2* sm = smem + g * (L + (L >> 5));
  const int step = kFftMaxLen / P;
  for (int k = lt; k <= (L >> 1); k += nthr) {
    float2 xk = make_float2(0.f, 0.f), xl = make_float2(0.f, 0.f);
    if (live) {
      xk = T2[(size_t)k * ldT + r];
      xl = T2[(size_t)(L - k) * ldT + r];
    }
    {
      // Z'[k] = (X[k] + conj X[L-k]) + i (X[k] - conj X[L-k]) conj(w^k)
      const float2 e = cadd(xk, cconj(xl));
      const float2 o = cmul(csub(xk, cconj(xl)), cconj(tw[k * step]));
      sm[pad(perm_pos(k & (L - 1), mL))] = make_float2(e.x - o.y, e.y + o.x);
    }
    const int kl = L - k;
    if (kl != k && kl < L) {
      const float2 e = cadd(xl, cconj(xk));
      const float2 o = cmul(csub(xl, cconj(xk)), cconj(tw[kl * step]));
      sm[pad(perm_pos(kl, mL))] = make_float2(e.x - o.y, e.y + o.x);
    }
  }
  __syncthreads();
  fft_inverse(sm, mL, lt, nthr, tw);
  if (!live) return;
  float* orow = out + (size_t)r * n_cols;
  for (int n = lt; n < L; n += nthr) {
    const int c = 2 * n;
    if (c >= n_cols) break;
    const float2 z = sm[pad(n)];
    orow[c] = z.x * scale;
    if (c + 1 < n_cols) orow[c + 1] = z.y * scale;
  }
}

struct DeviceTwiddles {
  float2* ptr = nullptr;
};
std::mutex g_tw_mu;
DeviceTwiddles g_tw[64];

int rows_per_block(int L) {
  int rb = 8192 / (L > 0 ? L : 1);
  if (rb < 1) rb = 1;
  if (rb > 16) rb = 16;
  return rb;
}

int cols_per_block(int P) {
  int c = 8192 / P;
  if (c < 1) c = 1;
  if (c > 16) c = 16;
  return c;
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
  (void)hipFuncSetAttribute((const void*)k_rows_r2c, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_cols_conv, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_cols_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  (void)hipFuncSetAttribute((const void*)k_rows_c2r, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
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

int fft_perm(int k, int p) { return perm_pos(k, ilog2(p)); }

size_t fft_spectrum_elems(int p) { return (size_t)(p / 2 + 1) * p; }

size_t fft_conv_work_elems(int n, int p) { return 2 * (size_t)(p / 2 + 1) * n; }

static void check_fft_len(int p) {
  PMS_REQUIRE(is_pow2(p) && p >= 4 && p <= kFftMaxLen,
              "FFT frame side must be a power of two in [4, " + std::to_string(kFftMaxLen) +
                  "], got " + std::to_string(p));
}

void fft_psf_spectrum(const float* psf, int q, int p, float2* spec, float2* work, hipStream_t st) {
  check_fft_len(p);
  PMS_REQUIRE(q >= 1 && q <= p, "PSF larger than the FFT frame");
  const float2* tw = fft_twiddles();
  const int pm = ilog2(p);
  const int L = p / 2;
  // work holds the (P/2+1) x P row spectra, column-major
  const int rb = rows_per_block(L);
  const size_t lds = (size_t)rb * (L + (L >> 5)) * sizeof(float2);
  hipLaunchKernelGGL(k_rows_r2c, dim3(ceil_div(p, rb)), dim3(kThreads), lds, st, psf, q, q / 2, pm, p, rb, work,
                     p, tw, (const int*)nullptr);
  PMS_LAUNCH_CHECK();
  const int cpb = cols_per_block(p);
  const size_t lds2 = (size_t)cpb * (p + (p >> 5)) * sizeof(float2);
  hipLaunchKernelGGL(k_cols_fwd, dim3(ceil_div(L + 1, cpb)), dim3(kThreads), lds2, st, (const float2*)work, pm,
                     spec, cpb, tw);
  PMS_LAUNCH_CHECK();
}

void fft_conv(const float* in, int n, int p, const float2* spec, float* out, float2* work, const int* pred,
              hipStream_t st) {
  check_fft_len(p);
  PMS_REQUIRE(p == n || p == 2 * n, "fft_conv: frame must be N (circular) or 2N (linear)");
  const float2* tw = fft_twiddles();
  const int pm = ilog2(p);
  const int L = p / 2;
  float2* T1 = work;
  float2* T2 = work + (size_t)(L + 1) * n;
  const int rb = rows_per_block(L);
  const size_t lds = (size_t)rb * (L + (L >> 5)) * sizeof(float2);
  hipLaunchKernelGGL(k_rows_r2c, dim3(ceil_div(n, rb)), dim3(kThreads), lds, st, in, n, 0, pm, n, rb, T1, n, tw,
                     pred);
  PMS_LAUNCH_CHECK();
  const int cpb = cols_per_block(p);
  const size_t lds2 = (size_t)cpb * (p + (p >> 5)) * sizeof(float2);
  hipLaunchKernelGGL(k_cols_conv, dim3(ceil_div(L + 1, cpb)), dim3(kThreads), lds2, st, (const float2*)T1, n, pm,
                     spec, T2, n, cpb, tw, pred);
  PMS_LAUNCH_CHECK();
  const float scale = 1.0f / ((float)p * (float)p);
  hipLaunchKernelGGL(k_rows_c2r, dim3(ceil_div(n, rb)), dim3(kThreads), lds, st, (const float2*)T2, n, pm, n, rb,
                     out, n, scale, tw, pred);
  PMS_LAUNCH_CHECK();
}

// numpy rfft2 spectrum (rows = axis-0 frequency m in [0,P), cols = k in [0,P/2]) of an
// unrolled PSF frame -> library layout: column k, position perm(m), times (-1)^(m+k)
// (the P/2 roll in both axes).
void fft_spectrum_from_numpy_host(const float* re_im, int p, float2* spec_host) {
  check_fft_len(p);
  const int pm = ilog2(p);
  const int nk = p / 2 + 1;
  for (int m = 0; m < p; ++m) {
    const int pos = perm_pos(m, pm);
    for (int k = 0; k < nk; ++k) {
      const float sgn = ((m + k) & 1) ? -1.f : 1.f;
      const float* v = re_im + 2 * ((size_t)m * nk + k);
      spec_host[(size_t)k * p + pos] = make_float2(sgn * v[0], sgn * v[1]);
    }
  }
}

}  // namespace pms
