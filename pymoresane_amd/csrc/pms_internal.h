// Internal (C++) interfaces between the kernel files of the MORESANE HIP library.
#pragma once

#include "pms_common.h"

namespace pms {

// ----------------------------------------------------------------- reductions
// Deterministic block-wide sums/max (fixed shuffle tree, then a fixed LDS order).
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

// Block sum of a double; valid in thread 0.  Block size must be a multiple of 64, <= 1024.
__device__ __forceinline__ double block_sum_f64(double v) {
  __shared__ double red_f64[16];
  v = wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_f64[wid] = v;
  __syncthreads();
  double tot = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) tot += red_f64[k];
  }
  return tot;
}

__device__ __forceinline__ float block_max_f32(float v) {
  __shared__ float red_f32[16];
  v = wave_max_f32(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_f32[wid] = v;
  __syncthreads();
  float m = -INFINITY;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) m = fmaxf(m, red_f32[k]);
  }
  return m;
}

__device__ __forceinline__ float block_min_f32(float v) {
  __shared__ float red_f32n[16];
  v = wave_min_f32(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_f32n[wid] = v;
  __syncthreads();
  float m = INFINITY;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) m = fminf(m, red_f32n[k]);
  }
  return m;
}

// Generic block max / min (float or double); valid in thread 0.
template <class T>
__device__ __forceinline__ T block_max_t(T v) {
  __shared__ T red_max_t[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_max_t[wid] = v;
  __syncthreads();
  T m = -INFINITY;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) m = red_max_t[k] > m ? red_max_t[k] : m;
  }
  return m;
}

template <class T>
__device__ __forceinline__ T block_min_t(T v) {
  __shared__ T red_min_t[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red_min_t[wid] = v;
  __syncthreads();
  T m = INFINITY;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) m = red_min_t[k] < m ? red_min_t[k] : m;
  }
  return m;
}

// ------------------------------------------------------------------- batching
// Kernels that can serve several jobs of one geometry in a single launch take
// their per-job pointers from fixed-capacity arrays in the kernel arguments;
// one grid dimension selects the job.  A single-job call is a batch of one.
constexpr int kMaxBatch = 32;

// ----------------------------------------------------------------------- IUWT
enum IuwtMode { IUWT_DECOMP = 0, IUWT_SMOOTH = 1, IUWT_RECOMP = 2 };

template <class T>
struct IuwtScaleCallT {
  int mode;
  const T* in;
  T* out_c;
  T* out_w;
  const T* w_in;
  const uint8_t* mask;
  const T* ref;
  double* partial;
  const int* pred;
  int h, w, scale;
};
using IuwtScaleCall = IuwtScaleCallT<float>;

void check_iuwt_scale(int h, int w, int scale);
int iuwt_scale(const IuwtScaleCall& c, hipStream_t st);  // returns launched blocks
int iuwt_scale_blocks(int h, int w, int scale, int mode);
long iuwt_decompose(const float* in, int h, int w, int scale_count, int scale_adjust, float* detail,
                    float* smoothed, float* work, const uint8_t* mask, const float* ref,
                    double* partial, const int* pred, hipStream_t st);
long iuwt_decompose_partials(int h, int w, int scale_count, int scale_adjust);
void iuwt_recompose(const float* detail, int nplanes, int h, int w, int scale_adjust,
                    const float* smoothed, float* out, float* work, const uint8_t* mask,
                    const int* pred, hipStream_t st);
// calls of one mode, shape and scale for several jobs, in launches of up to kMaxBatch jobs, on the
// direct-stencil kernel (small) or the row walk
int iuwt_scale_batch(const IuwtScaleCall* calls, int count, hipStream_t st, bool small);
// iuwt_decompose / iuwt_recompose for several jobs of one plane shape; every job keeps its own
// scale range, buffers, mask and predicate.  Decomposition jobs report their partial count in np.
template <class T>
struct IuwtDecompJobT {
  const T* in;
  int scale_count, scale_adjust;
  T* detail;
  T* smoothed;
  T* work;
  const uint8_t* mask;
  const T* ref;
  double* partial;
  const int* pred;
  long np;
};
using IuwtDecompJob = IuwtDecompJobT<float>;
void iuwt_decompose_batch(IuwtDecompJob* jobs, int count, int h, int w, hipStream_t st);
template <class T>
struct IuwtRecompJobT {
  const T* detail;
  int nplanes, scale_adjust;
  const T* smoothed;
  T* out;
  T* work;
  const uint8_t* mask;
  const int* pred;
};
using IuwtRecompJob = IuwtRecompJobT<float>;
void iuwt_recompose_batch(const IuwtRecompJob* jobs, int count, int h, int w, hipStream_t st);

// Direct-stencil IUWT (iuwt_any.hip) with the reference's exact tap semantics at every scale,
// including scales whose taps are longer than the side (2^(s+1) > side: numpy's clipped slices
// read the axis reversed, iuwt.py:166-187).  The float32 transforms above route such scales here;
// the float64 transforms use it at every scale.  A decomposition scale is two launches (smooth,
// then detail = C0 - smooth(C)), so its call needs out_c (a free work plane when not kept).
bool iuwt_over_long(int h, int w, int scale);
int iuwt_gen_blocks(int h, int w);  // partial sums per plane written by a reducing detail launch
template <class T>
int iuwt_gen_scale_batch(const IuwtScaleCallT<T>* calls, int count, hipStream_t st);
void iuwt_decompose_batch(IuwtDecompJobT<double>* jobs, int count, int h, int w, hipStream_t st);
void iuwt_recompose_batch(const IuwtRecompJobT<double>* jobs, int count, int h, int w, hipStream_t st);
long iuwt_decompose_partials_f64(int h, int w, int scale_count, int scale_adjust);

// ------------------------------------------------------------------------ FFT
// Device twiddle table exp(-2 pi i k / kFftMaxLen), k in [0, kFftMaxLen).
constexpr int kFftMaxLog2 = 14;
constexpr int kFftMaxLen = 1 << kFftMaxLog2;
void fft_init_device(int device);
const float2* fft_twiddles();

// Spectrum layout: (P/2+1) frequency columns, each holding P values along the
// other axis in the library's digit-reversed FFT order, already carrying the
// P/2 spatial roll that centres the PSF at the origin.
size_t fft_spectrum_elems(int p);
size_t fft_conv_work_elems(int n, int p);  // float2 elements of scratch for fft_conv
// psf (q x q) embedded centred in a p x p frame, rolled by -p/2 -> spectrum.
void fft_psf_spectrum(const float* psf, int q, int p, float2* spec, float2* work, hipStream_t st);
// out (n x n) = image (n x n) convolved with the PSF of spectrum spec (p = n or 2n).
void fft_conv(const float* in, int n, int p, const float2* spec, float* out, float2* work,
              const int* pred, hipStream_t st);
// fft_conv for several jobs of one image side and frame (one launch per pass per kMaxBatch jobs)
struct FftConvJob {
  const float* in;
  const float2* spec;
  float* out;
  float2* work;
  const int* pred;
};
void fft_conv_batch(const FftConvJob* jobs, int count, int n, int p, hipStream_t st);
// spectrum -> digit-reversed column-major layout (host arrays, numpy rfft2 convention)
void fft_spectrum_from_numpy_host(const float* re_im, int p, float2* spec_host);
int fft_perm(int k, int p);  // storage position of frequency k in a length-p column

// Mixed-radix FFT convolution (fft_any.hip) for any even frame side with prime factors up to 13,
// in float32 or float64; spectra and scratch in the library complex type Cx<T>.  The float32
// power-of-two frames keep the kernels above; the float32 entry points above route other frames
// here.  Layout as above ((P/2+1) columns of P values in the transform's digit-reversed order).
template <class T>
struct alignas(2 * sizeof(T)) Cx {
  T x, y;
};
template <class T>
struct GConvJob {
  const T* in;
  const Cx<T>* spec;
  T* out;
  Cx<T>* work;
  const int* pred;
};
int gfft_max_frame(size_t real_bytes);          // largest frame side one LDS transform holds
bool gfft_frame_ok(int p, size_t real_bytes);   // even, within the maximum, primes <= 13
int gfft_perm(int k, int p);                    // storage position of frequency k in a column
template <class T, class TIN>
void gfft_psf_spectrum(const TIN* psf, int q, int p, Cx<T>* spec, Cx<T>* work, hipStream_t st);
template <class T>
void gfft_conv_batch(const GConvJob<T>* jobs, int count, int n, int p, hipStream_t st);

// ------------------------------------------------------------------- toolbox
// Element-typed helpers take float or double planes (the float64 precision mode of the driver).
void copy_plane(const float* src, float* dst, size_t n, const int* pred, hipStream_t st);
void copy_plane(const double* src, double* dst, size_t n, const int* pred, hipStream_t st);
void masked_copy(const float* src, const uint8_t* mask, float* dst, size_t n, const int* pred, hipStream_t st);
void masked_copy(const double* src, const uint8_t* mask, double* dst, size_t n, const int* pred, hipStream_t st);
void fill_f32(float* p, size_t n, float v, hipStream_t st);
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);
void fill_zero(float* p, size_t n, hipStream_t st);
void fill_zero(double* p, size_t n, hipStream_t st);
// dst (n x n) = src[r0.., c0..] (row stride ld), times mul (same layout as src) when non-null
void window_copy(const float* src, int ld, int r0, int c0, int n, const float* mul, float* dst, hipStream_t st);
void window_copy(const float* src, int ld, int r0, int c0, int n, const double* mul, double* dst, hipStream_t st);
void window_copy(const double* src, int ld, int r0, int c0, int n, const double* mul, double* dst, hipStream_t st);
int reduce_grid(size_t n);      // number of partials written by dot/sumsq reductions over n elements
int threshold_grid(size_t plane);  // partial maxima per plane

// Pixels entering the MAD estimate: inside the include rectangle and outside
// the exclude rectangle (iuwt_toolbox.py:31-43, Python slice semantics).
struct NoiseMask {
  bool masked;
  int inc_r0, inc_r1, inc_c0, inc_c1;
  int exc_r0, exc_r1, exc_c0, exc_c1;
  long count;
};
NoiseMask noise_mask(int h, int w, int edge_excl, int int_excl);
size_t mad_work_bytes(int S);
// the allocation-free median workspace of the current device (looked up once per device) and its size
void* noalloc_mad_scratch();
size_t noalloc_mad_bytes();
void mad_work_init(void* work, int S, hipStream_t st);  // once per workspace (histograms start at 0)
void mad_thresholds(const float* planes, int S, int h, int w, const NoiseMask& nm, double* thr, void* work,
                    const int* pred, hipStream_t st);
void mad_thresholds(const double* planes, int S, int h, int w, const NoiseMask& nm, double* thr, void* work,
                    const int* pred, hipStream_t st);
// out = in where |in| > sigma*thr[s] else 0, zeroing a per-scale border (border may be null);
// pmax receives threshold_grid(plane) partial maxima per plane.
void apply_threshold(const float* in, float* out, int S, int h, int w, const double* thr, double sigma,
                     const int* border, float* pmax, const int* pred, hipStream_t st);
void apply_threshold(const double* in, double* out, int S, int h, int w, const double* thr, double sigma,
                     const int* border, double* pmax, const int* pred, hipStream_t st);
void plane_max(const float* in, int S, size_t plane, int use_abs, float* pmax, const int* pred, hipStream_t st);
void plane_max(const double* in, int S, size_t plane, int use_abs, double* pmax, const int* pred, hipStream_t st);
void max_final(const float* pmax, int nb, int S, float* out_max, const double* energy, double* nsm, const int* pred,
               hipStream_t st);
void max_final(const double* pmax, int nb, int S, double* out_max, const double* energy, double* nsm,
               const int* pred, hipStream_t st);
size_t extract_work_bytes(int S, int h, int w);
void extract_work_init(void* work, int S, int h, int w, hipStream_t st);
void source_extract(const float* w, int S, int h, int wd, double tol, int neg_comp, uint8_t* mask, float* extracted,
                    int* labels_out, void* work, const int* pred, hipStream_t st);
void source_extract(const double* w, int S, int h, int wd, double tol, int neg_comp, uint8_t* mask,
                    double* extracted, int* labels_out, void* work, const int* pred, hipStream_t st);
void dot_partials(const float* a, const float* b, size_t n, double* partial, const int* pred, hipStream_t st);
void dot_partials(const double* a, const double* b, size_t n, double* partial, const int* pred, hipStream_t st);
void sum_partials(const double* partial, long np, int groups, double* out, const int* pred, hipStream_t st);
// one group of partials per job: out = sum(partial[0..np))
struct SumJob {
  const double* partial;
  long np;
  double* out;
  const int* pred;  // skip the job when *pred == 0 (null = always)
};
void sum_partials_batch(const SumJob* jobs, int count, hipStream_t st);
void plane_sumsq_partials(const float* a, int S, size_t plane, double* partial, const int* pred, hipStream_t st);
void plane_sumsq_partials(const double* a, int S, size_t plane, double* partial, const int* pred, hipStream_t st);
void mean_std(const float* a, size_t n, double* partial, double* sum_out, double* dev2_out, hipStream_t st);
void mean_std(const double* a, size_t n, double* partial, double* sum_out, double* dev2_out, hipStream_t st);

}  // namespace pms
