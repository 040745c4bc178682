// Device-resident MORESANE driver (major loop, conjugate-gradient minor loop,
// by-scale outer loop), written as a per-job state machine that a Runner
// steps for one or many jobs at once.
//
// Control flow follows pymoresane/main.py:53-599 step for step; every array
// stays in HBM and only the per-iteration decision scalars (normalised scale
// maxima, the SNR sums, the residual standard deviation) are read back:
//   * major iteration (main.py:278-300): decomposition of the dirty
//     subregion, exact MAD thresholds, thresholding with the edge mask, and
//     per-scale maxima divided by the PSF scale energies;
//   * scale choice (main.py:304-340) on the host from those maxima;
//   * source extraction + recomposition (main.py:347-361) on the device;
//   * CG minor loop (main.py:365-457): vectors in float32, every inner
//     product, step size and norm in float64; the positivity clip branch is
//     predicated on a device flag so no extra synchronisation happens;
//     the host reads the SNR sums once per iteration for the exit rules;
//   * model update and residual (main.py:479-506), with revert.
// The Runner gives each job a 64-double scalar slot in one block, so a step
// of any number of jobs ends with one copy of that block and one
// synchronisation, and the CG iterations of jobs that share a plane shape
// run as batched launches (one launch per kernel for up to kMaxBatch jobs).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"

namespace pms {
namespace {

constexpr int kT = 256;

// Scalar slot layout (doubles): [0,16) thresholds, [16,32) normalised maxima, [32,48) PSF scale
// energies, 48/51 rr and rnrn (swapped every iteration), 49 p.Ap, 50 alpha, 52 beta, 53 SNR
// residual sum, 54 SNR energy sum, 56/57 mean sum and squared deviations, 58/59 the minor loop's
// current and previous SNR, 60: dmax (in the job's element type), 61/62 the minor loop's iteration
// count and exit code (-1 while it iterates) as read back by the host.
// Int slot: [0] clip flag (written only for jobs that enforce positivity), [1] minor loop still
// iterating, [2] its iteration count, [3] its exit code, [16,32) suppression borders.
constexpr int kThr = 0, kNsm = 16, kEnergy = 32, kRR = 48, kPAP = 49, kAlpha = 50, kRNRN = 51, kBeta = 52,
              kD2 = 53, kE2 = 54, kMeanSum = 56, kDev2 = 57, kSnrCur = 58, kSnrLast = 59, kDmax = 60,
              kNiterD = 61, kExitD = 62;
// Int slot: [0] clip flag, [1] minor loop active, [2] minor iterations, [3] exit code, [16,32) borders.
constexpr int kClip = 0, kActive = 1, kNiter = 2, kExit = 3, kBorders = 16;
// Upper region of a job's partial buffer (which holds at least 4096 doubles, and a CG reduction uses
// at most 1024 blocks): the xn minima and the SNR partials, written by kernels that are still
// reading the lower region's partials.
constexpr int kPartialHi = 1024;

// 16 bytes of plane elements: four floats or two doubles (the side is even, so a plane holds a
// multiple of four elements and every job's vector starts 16-byte aligned).
template <class T>
struct alignas(16) V16 {
  T a[16 / sizeof(T)];
};
template <class T>
constexpr int kLanes = 16 / sizeof(T);

// Per-job pointers of one batched CG launch (grid y selects the job).
template <class T>
struct CgJob {
  T *x, *xn, *r, *rn, *p, *ap;
  double *partial, *scal, *rr, *rnrn;
  int* iflag;  // iflag[kActive] == 0: the job's minor loop has exited, its kernels do nothing
  int enforce;
};
template <class T>
struct CgBatch {
  CgJob<T> j[kMaxBatch];
  size_t n;
};

// ---- CG vector kernels, in the job's element type T (float, or double in the float64 precision
// mode); every step size and inner product is float64.  Reductions write one partial per block; a
// one-block-per-job finishing kernel adds the partials in block order (kernel boundaries order the
// two, no fences needed).
__device__ __forceinline__ double block_total(const double* __restrict__ partial, int np) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kT) acc += partial[i];
  return block_sum_f64(acc);
}

// the same totals in every thread of the block (a streaming kernel that needs the previous
// reduction's result sums its partials itself, in block_total's order, instead of waiting for a
// one-block finishing launch)
__device__ __forceinline__ double block_total_all(const double* __restrict__ partial, int np) {
  __shared__ double s_tot;
  const double t = block_total(partial, np);
  if (threadIdx.x == 0) s_tot = t;
  __syncthreads();
  return s_tot;
}

template <class T>
__device__ __forceinline__ T block_min_all(const double* __restrict__ partial, int np) {
  __shared__ T s_min;
  T m = INFINITY;
  for (int i = threadIdx.x; i < np; i += kT) {
    const T v = (T)partial[i];
    m = v < m ? v : m;
  }
  m = block_min_t<T>(m);
  if (threadIdx.x == 0) s_min = m;
  __syncthreads();
  return s_min;
}

// per-block partials of p . Ap
template <class T>
__global__ __launch_bounds__(kT) void k_cg_pap(CgBatch<T> b) {
  const CgJob<T>& J = b.j[blockIdx.y];
  if (J.iflag[kActive] == 0) return;
  const V16<T>* __restrict__ p = reinterpret_cast<const V16<T>*>(J.p);
  const V16<T>* __restrict__ ap = reinterpret_cast<const V16<T>*>(J.ap);
  const size_t nv = b.n / kLanes<T>;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)gridDim.x * kT) {
    const V16<T> u = p[i], v = ap[i];
#pragma unroll
    for (int k = 0; k < kLanes<T>; ++k) acc = fma((double)u.a[k], (double)v.a[k], acc);
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) J.partial[blockIdx.x] = acc;
}

template <class T>
__device__ __forceinline__ T xn_of(double a, T p, T x) {
  return (T)fma(a, (double)p, (double)x);
}

// alpha = rr / (p . Ap) from k_cg_pap's np partials (every block adds them in the same order; block 0
// stores p . Ap and alpha) ; xn = x + alpha p ; per-block minima of xn in the upper partials
template <class T>
__global__ __launch_bounds__(kT) void k_cg_xn(CgBatch<T> b, int np) {
  const CgJob<T>& J = b.j[blockIdx.y];
  if (J.iflag[kActive] == 0) return;
  const double tot = block_total_all(J.partial, np);
  const double a = *J.rr / tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    J.scal[kPAP] = tot;
    J.scal[kAlpha] = a;
  }
  const V16<T>* __restrict__ x = reinterpret_cast<const V16<T>*>(J.x);
  const V16<T>* __restrict__ p = reinterpret_cast<const V16<T>*>(J.p);
  V16<T>* __restrict__ xn = reinterpret_cast<V16<T>*>(J.xn);
  const size_t nv = b.n / kLanes<T>;
  T m = INFINITY;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)gridDim.x * kT) {
    const V16<T> u = p[i], v = x[i];
    V16<T> o;
#pragma unroll
    for (int k = 0; k < kLanes<T>; ++k) {
      o.a[k] = xn_of<T>(a, u.a[k], v.a[k]);
      m = o.a[k] < m ? o.a[k] : m;
    }
    xn[i] = o;
  }
  m = block_min_t<T>(m);
  if (threadIdx.x == 0) J.partial[kPartialHi + blockIdx.x] = (double)m;
}

// positivity clip when enforce && min(xn) < 0 (the minimum of k_cg_xn's np per-block minima;
// block 0 stores the flag, which predicates the clip pass's operator launches):
// xn[xn<0] = 0 ; p = (xn - x) / alpha
template <class T>
__global__ __launch_bounds__(kT) void k_cg_clip(CgBatch<T> b, int np) {
  const CgJob<T>& J = b.j[blockIdx.y];
  if (J.iflag[kActive] == 0) {  // the clip pass's operator launches are predicated on this flag
    if (blockIdx.x == 0 && threadIdx.x == 0) J.iflag[kClip] = 0;
    return;
  }
  const int flag = (J.enforce && block_min_all<T>(J.partial + kPartialHi, np) < (T)0) ? 1 : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) J.iflag[kClip] = flag;
  if (flag == 0) return;
  const double a = J.scal[kAlpha];
  V16<T>* __restrict__ xn = reinterpret_cast<V16<T>*>(J.xn);
  const V16<T>* __restrict__ x = reinterpret_cast<const V16<T>*>(J.x);
  V16<T>* __restrict__ p = reinterpret_cast<V16<T>*>(J.p);
  const size_t nv = b.n / kLanes<T>;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)gridDim.x * kT) {
    V16<T> v = xn[i];
    const V16<T> u = x[i];
    V16<T> q;
#pragma unroll
    for (int k = 0; k < kLanes<T>; ++k) {
      if (v.a[k] < (T)0) v.a[k] = 0;
      q.a[k] = (T)(((double)v.a[k] - (double)u.a[k]) / a);
    }
    xn[i] = v;
    p[i] = q;
  }
}

template <class T>
__device__ __forceinline__ T rn_of(double a, T ap, T r) {
  return (T)fma(-a, (double)ap, (double)r);
}

// Incremental SNR sums.  The SNR of main.py:413-424 compares E with M.W(H xn).  The convolution and
// the decomposition are linear and xn = x + alpha p (the positivity clip rewrites p so that this
// holds), so M.W(H xn) = M.W(H x) + alpha M.W(H p), and M.W(H p) is this iteration's apply_op
// decomposition (Dcg).  The difference planes R = E - M.W(H x) are carried across the iterations of
// a minor loop (R = E at its start, x = 0): R <- R - alpha Dcg, D2 = sum R^2.  This replaces one
// convolution and one decomposition per iteration by one streaming pass over R and Dcg, done in the
// same launch as the residual update.
template <class T>
struct RnJob {
  const T* r;
  const T* ap;
  T* rn;
  T* rs;         // SNR difference planes (null: the SNR sums come from elsewhere)
  const T* dcg;  // this iteration's masked decomposition of H p
  double* partial;
  const double* scal;  // alpha at kAlpha
  const int* iflag;
  size_t n_snr;  // SNR elements: scales x plane (a multiple of 4)
  int nblk_snr;  // blocks and partials of the SNR pass (a function of n_snr only)
};
template <class T>
struct RnBatch {
  RnJob<T> j[kMaxBatch];
  size_t n;  // CG vector length
  int np;    // blocks (and partials) of the rn pass: cg_grid(n)
};

// rn = r - alpha Ap with per-block partials of rn . rn in blocks [0, np) (lower partials), and
// R <- R - alpha Dcg with per-block partials of R . R in blocks [0, nblk_snr) (upper partials).
template <class T>
__global__ __launch_bounds__(kT) void k_cg_rn(RnBatch<T> b) {
  const RnJob<T>& J = b.j[blockIdx.y];
  if (J.iflag[kActive] == 0) return;
  const double a = J.scal[kAlpha];
  if ((int)blockIdx.x < b.np) {
    const V16<T>* __restrict__ r = reinterpret_cast<const V16<T>*>(J.r);
    const V16<T>* __restrict__ ap = reinterpret_cast<const V16<T>*>(J.ap);
    V16<T>* __restrict__ rn = reinterpret_cast<V16<T>*>(J.rn);
    const size_t nv = b.n / kLanes<T>;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)b.np * kT) {
      const V16<T> u = ap[i], v = r[i];
      V16<T> o;
#pragma unroll
      for (int k = 0; k < kLanes<T>; ++k) o.a[k] = rn_of<T>(a, u.a[k], v.a[k]);
      rn[i] = o;
#pragma unroll
      for (int k = 0; k < kLanes<T>; ++k) acc = fma((double)o.a[k], (double)o.a[k], acc);
    }
    acc = block_sum_f64(acc);
    if (threadIdx.x == 0) J.partial[blockIdx.x] = acc;
  }
  if (J.rs && (int)blockIdx.x < J.nblk_snr) {
    V16<T>* __restrict__ rs = reinterpret_cast<V16<T>*>(J.rs);
    const V16<T>* __restrict__ dcg = reinterpret_cast<const V16<T>*>(J.dcg);
    const size_t nv = J.n_snr / kLanes<T>;
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)J.nblk_snr * kT) {
      const V16<T> d = dcg[i];
      V16<T> v = rs[i];
#pragma unroll
      for (int k = 0; k < kLanes<T>; ++k) v.a[k] = rn_of<T>(a, d.a[k], v.a[k]);
      rs[i] = v;
#pragma unroll
      for (int k = 0; k < kLanes<T>; ++k) acc = fma((double)v.a[k], (double)v.a[k], acc);
    }
    acc = block_sum_f64(acc);
    if (threadIdx.x == 0) J.partial[kPartialHi + blockIdx.x] = acc;
  }
}

// rnrn = rn . rn (from k_cg_rn's np partials) ; beta = rnrn / rr (block 0 stores both) ;
// p = rn + beta p
template <class T>
__global__ __launch_bounds__(kT) void k_cg_p(CgBatch<T> b, int np) {
  const CgJob<T>& J = b.j[blockIdx.y];
  if (J.iflag[kActive] == 0) return;
  const double tot = block_total_all(J.partial, np);
  const double be = tot / *J.rr;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *J.rnrn = tot;
    J.scal[kBeta] = be;
  }
  const V16<T>* __restrict__ rn = reinterpret_cast<const V16<T>*>(J.rn);
  V16<T>* __restrict__ p = reinterpret_cast<V16<T>*>(J.p);
  const size_t nv = b.n / kLanes<T>;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < nv; i += (size_t)gridDim.x * kT) {
    const V16<T> u = rn[i];
    V16<T> v = p[i];
#pragma unroll
    for (int k = 0; k < kLanes<T>; ++k) v.a[k] = (T)fma(be, (double)v.a[k], (double)u.a[k]);
    p[i] = v;
  }
}

// Exit rules of the minor loop (main.py:427-457) on the device, one block per job, after an
// iteration's SNR partial sums: D2 (their total), the SNR, the iteration count and, when a rule
// fires, the exit code; an exit clears the job's active flag, which turns its later queued
// iterations into no-ops.  The host reads the outcome after the step and replays the iterate swaps
// of the iterations that ran.
struct CtlBatch {
  double* scal[kMaxBatch];
  int* iflag[kMaxBatch];
  const double* partial[kMaxBatch];  // the SNR partials
  long np[kMaxBatch];
  int miter[kMaxBatch];
};

__global__ __launch_bounds__(kT) void k_minor_ctl(CtlBatch b) {
  const int j = blockIdx.x;
  double* sc = b.scal[j];
  int* fl = b.iflag[j];
  if (fl[kActive] == 0) return;
  const double d2 = block_total(b.partial[j], (int)b.np[j]);
  if (threadIdx.x != 0) return;
  sc[kD2] = d2;
  const double snr = 20.0 * log10(sqrt(sc[kE2]) / sqrt(sc[kD2]));
  const double last = sc[kSnrCur];
  const int niter = fl[kNiter] + 1;
  int code = -1;
  if (niter == 1 && snr > 40)
    code = EXIT_FIRST_HIGH;
  else if (snr > 40)
    code = EXIT_ACCEPT;
  else if (niter > 2 && snr <= last)
    code = snr > 10.5 ? EXIT_FELL_OK : EXIT_FELL_LOW;
  else if (niter >= b.miter[j])
    code = EXIT_CAP;
  sc[kSnrLast] = last;
  sc[kSnrCur] = snr;
  sc[kNiterD] = (double)niter;
  sc[kExitD] = (double)code;
  fl[kNiter] = niter;
  if (code >= 0) {
    fl[kExit] = code;
    fl[kActive] = 0;
  }
}

// start of a minor loop: iterating, no iterations yet, SNRs 0 (main.py:374-377)
__global__ void k_minor_reset(double* scal, int* iflag, int iterate) {
  if (threadIdx.x != 0) return;
  iflag[kActive] = iterate;
  iflag[kNiter] = 0;
  iflag[kExit] = EXIT_CAP;
  scal[kSnrCur] = 0.0;
  scal[kSnrLast] = 0.0;
  scal[kNiterD] = 0.0;
  scal[kExitD] = -1.0;
}

// model[r0+r, c0+c] += gain * x[r, c]
template <class T>
__global__ __launch_bounds__(kT) void k_model_add(T* __restrict__ model, int ld, int r0, int c0, int n,
                                                  const T* __restrict__ x, double gain) {
  const size_t tot = (size_t)n * n;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < tot; i += (size_t)gridDim.x * kT) {
    const int r = (int)(i / n), c = (int)(i % n);
    T* m = model + (size_t)(r0 + r) * ld + (c0 + c);
    *m = *m + (T)(gain * (double)x[i]);
  }
}

// out = a - b
template <class TA, class T>
__global__ __launch_bounds__(kT) void k_sub(const TA* __restrict__ a, const T* __restrict__ b, T* __restrict__ out,
                                            size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
    out[i] = (T)a[i] - b[i];
}

// plane += other (rounded once to the plane's type)
template <class TA, class TB>
__global__ __launch_bounds__(kT) void k_add_inplace(TA* __restrict__ a, const TB* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
    a[i] = (TA)((TB)a[i] + b[i]);
}

// float32 copy of a double plane (the float64 mode's residual output)
__global__ __launch_bounds__(kT) void k_to_f32(const double* __restrict__ a, float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) out[i] = (float)a[i];
}

int grid_for(size_t n) { return reduce_grid(n); }

pms_trace_event ev(int kind, int major, int a, int b, int c, double x, double y) {
  pms_trace_event e{};
  e.kind = kind;
  e.major = major;
  e.a = a;
  e.b = b;
  e.c = c;
  e.x = x;
  e.y = y;
  return e;
}

int cg_grid(size_t n) {
  const size_t b = ((n >> 2) + kT - 1) / kT;
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

template <class T, class K>
void launch_cg(K kernel, const CgBatch<T>& b, int nb, hipStream_t st) {
  hipLaunchKernelGGL(kernel, dim3(cg_grid(b.n), nb), dim3(kT), 0, st, b);
  PMS_LAUNCH_CHECK();
}

// a streaming kernel that also finishes the previous kernel's reduction (cg_grid(n) partials)
template <class T, class K>
void launch_cg_after(K kernel, const CgBatch<T>& b, int nb, hipStream_t st) {
  hipLaunchKernelGGL(kernel, dim3(cg_grid(b.n), nb), dim3(kT), 0, st, b, cg_grid(b.n));
  PMS_LAUNCH_CHECK();
}

// Element-typed front ends of the transform and convolution libraries.
void decompose_batch(IuwtDecompJobT<float>* jobs, int count, int h, int w, hipStream_t st) {
  iuwt_decompose_batch(jobs, count, h, w, st);
}
void decompose_batch(IuwtDecompJobT<double>* jobs, int count, int h, int w, hipStream_t st) {
  iuwt_decompose_batch(jobs, count, h, w, st);
}
void recompose_batch(const IuwtRecompJobT<float>* jobs, int count, int h, int w, hipStream_t st) {
  iuwt_recompose_batch(jobs, count, h, w, st);
}
void recompose_batch(const IuwtRecompJobT<double>* jobs, int count, int h, int w, hipStream_t st) {
  iuwt_recompose_batch(jobs, count, h, w, st);
}
void conv_batch(const FftConvJob* jobs, int count, int n, int p, hipStream_t st) {
  fft_conv_batch(jobs, count, n, p, st);
}
void conv_batch(const GConvJob<double>* jobs, int count, int n, int p, hipStream_t st) {
  gfft_conv_batch<double>(jobs, count, n, p, st);
}
template <class TIN>
void psf_spectrum_of(const TIN* psf, int q, int p, float2* spec, float2* work, hipStream_t st) {
  fft_psf_spectrum(psf, q, p, spec, work, st);
}
template <class TIN>
void psf_spectrum_of(const TIN* psf, int q, int p, Cx<double>* spec, Cx<double>* work, hipStream_t st) {
  gfft_psf_spectrum<double, TIN>(psf, q, p, spec, work, st);
}
template <class T>
void decompose1(const T* in, int n, int S, T* detail, T* work, hipStream_t st) {
  IuwtDecompJobT<T> j{in, S, 0, detail, nullptr, work, nullptr, nullptr, nullptr, nullptr, 0};
  decompose_batch(&j, 1, n, n, st);
}

}  // namespace

// --------------------------------------------------------------------------- buffers
DeviceArena::~DeviceArena() { release(); }

void DeviceArena::release() {
  for (void* p : ptrs_) (void)hipFree(p);
  ptrs_.clear();
  bytes_ = 0;
}

void* DeviceArena::alloc(size_t bytes) {
  void* p = nullptr;
  const size_t b = ((bytes + 255) / 256) * 256;
  if (b == 0) return nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) fail(PMS_ERR_ALLOC, "device allocation of " + std::to_string(b) + " bytes failed");
  ptrs_.push_back(p);
  bytes_ += b;
  return p;
}

// --------------------------------------------------------------------------- geometry
Geometry make_geometry(int n, int psf_n, const MoresaneParams& prm, int scale_count_in, int* subregion_out) {
  Geometry g{};
  PMS_REQUIRE(n >= 4 && (n % 2) == 0, "Image size is uneven. Please use even dimensions.");
  PMS_REQUIRE(prm.precision == 0 || prm.precision == 1, "precision must be 0 (float32) or 1 (float64)");
  int sub = prm.subregion;
  if (sub <= 0 || sub > n) sub = n;
  PMS_REQUIRE(sub % 2 == 0 && sub >= 4, "subregion must be even and >= 4");
  const double l2 = std::log2((double)n) - 1.0;
  int S = scale_count_in;
  if (S <= 0 || S > l2) S = (int)l2;
  PMS_REQUIRE(S >= 1, "image too small for one wavelet scale");
  g.n = n;
  g.psf_n = psf_n;
  g.sub = sub;
  g.S = S;
  g.lo = n / 2 - sub / 2;
  g.linear = prm.conv_mode == 0;
  g.double_psf = psf_n == 2 * n;
  g.f64 = prm.precision == 1;
  PMS_REQUIRE(g.double_psf || psf_n == n, "PSF must be the dirty image size or twice it");
  if (g.linear) {
    g.p_sub = 2 * sub;
    g.p_full = 2 * n;
  } else {
    g.p_sub = sub;
    g.p_full = n;
  }
  for (int pf : {g.p_sub, g.p_full}) {
    if (g.f64)
      PMS_REQUIRE(gfft_frame_ok(pf, sizeof(double)),
                  "float64 FFT frame side " + std::to_string(pf) + " is not supported (even sides up to " +
                      std::to_string(gfft_max_frame(sizeof(double))) + " with prime factors up to 13)");
    else
      PMS_REQUIRE(is_pow2(pf) ? pf <= kFftMaxLen : gfft_frame_ok(pf, sizeof(float)),
                  "FFT frame side " + std::to_string(pf) + " is not supported (even sides up to " +
                      std::to_string(kFftMaxLen) + " with prime factors up to 13)");
  }
  check_iuwt_scale(sub, sub, S - 1);
  if (subregion_out) *subregion_out = sub;
  return g;
}


// Byte sizes of every per-job buffer for geometry g (S = scale capacity), in the job's element type.
static JobSizes job_sizes(const Geometry& g, int S) {
  JobSizes z{};
  const size_t e = g.f64 ? sizeof(double) : sizeof(float);
  const size_t N2 = (size_t)g.n * g.n;
  const size_t s2 = (size_t)g.sub * g.sub;
  z.f64 = g.f64;
  z.s2 = s2 * e;
  z.S_s2 = (size_t)S * s2 * e;
  z.mask = (size_t)S * s2;
  z.iuwt = 2 * s2 * e;
  z.N2 = N2 * e;
  z.spec_sub = fft_spectrum_elems(g.p_sub) * 2 * e;
  z.spec_full = fft_spectrum_elems(g.p_full) * 2 * e;
  z.fft = std::max({fft_conv_work_elems(g.sub, g.p_sub), fft_conv_work_elems(g.n, g.p_full),
                    fft_spectrum_elems(g.p_sub), fft_spectrum_elems(g.p_full)}) *
          2 * e;
  z.psf = (size_t)g.psf_n * g.psf_n * e;
  z.ext = extract_work_bytes(S, g.sub, g.sub);
  const long dec_partials = g.f64 ? iuwt_decompose_partials_f64(g.sub, g.sub, S, 0)
                                  : iuwt_decompose_partials(g.sub, g.sub, S, 0);
  z.partial = (size_t)std::max<long>({dec_partials, (long)reduce_grid(N2) * 2, (long)reduce_grid(s2) * S, 4096L}) *
              sizeof(double);
  z.pmax = (size_t)S * 2048 * e;
  return z;
}

static bool fits(const JobSizes& need, const JobSizes& have) {
  return need.f64 == have.f64 && need.s2 <= have.s2 && need.S_s2 <= have.S_s2 && need.mask <= have.mask &&
         need.iuwt <= have.iuwt && need.N2 <= have.N2 && need.spec_sub <= have.spec_sub &&
         need.spec_full <= have.spec_full && need.fft <= have.fft && need.psf <= have.psf && need.ext <= have.ext &&
         need.partial <= have.partial && need.pmax <= have.pmax;
}

static JobSizes max_sizes(const JobSizes& a, const JobSizes& b) {
  JobSizes z;
  z.f64 = a.f64;
  z.s2 = std::max(a.s2, b.s2);
  z.S_s2 = std::max(a.S_s2, b.S_s2);
  z.mask = std::max(a.mask, b.mask);
  z.iuwt = std::max(a.iuwt, b.iuwt);
  z.N2 = std::max(a.N2, b.N2);
  z.spec_sub = std::max(a.spec_sub, b.spec_sub);
  z.spec_full = std::max(a.spec_full, b.spec_full);
  z.fft = std::max(a.fft, b.fft);
  z.psf = std::max(a.psf, b.psf);
  z.ext = std::max(a.ext, b.ext);
  z.partial = std::max(a.partial, b.partial);
  z.pmax = std::max(a.pmax, b.pmax);
  return z;
}

size_t job_footprint_bytes(const Geometry& g) {
  const JobSizes z = job_sizes(g, std::max(g.S, (int)std::log2((double)g.n) - 1));
  const size_t by_scale = z.N2;  // moresane_by_scale's private dirty copy
  return z.s2 * 10 + z.S_s2 * 4 + z.mask + z.iuwt + z.N2 * 3 + z.spec_sub + z.spec_full + z.fft + z.psf + z.ext +
         z.partial + z.pmax + by_scale + 64 * 8 + 64 * 4 + 12 * 256;
}


static std::vector<int> suppression_borders(int S, int edge_suppression, int edge_offset) {
  std::vector<int> b;
  if (edge_suppression) {
    int corruption = 0;
    for (int i = 0; i < S; ++i) {
      corruption += 2 * (1 << i);
      b.push_back(edge_offset > corruption ? edge_offset : corruption);
    }
  } else if (edge_offset > 0) {
    b.assign(S, edge_offset);
  }
  return b;
}


// --------------------------------------------------------------------------- job workspace
Job::~Job() {
  arena_.release();
  by_scale_arena_.release();
}

template <class T>
static void carve(Planes<T>& P, const JobSizes& z, DeviceArena& arena) {
  using C = typename Planes<T>::C;
  auto F = [&](size_t bytes) { return (T*)arena.alloc(bytes); };
  P.dirty_sub = F(z.s2);
  P.psf_sub = F(z.s2);
  P.D = F(z.S_s2);
  P.Dcg = F(z.S_s2);
  P.E = F(z.S_s2);
  P.Rs = F(z.S_s2);
  P.rec = F(z.s2);
  P.x = F(z.s2);
  P.xn = F(z.s2);
  P.r = F(z.s2);
  P.rn = F(z.s2);
  P.p = F(z.s2);
  P.ap = F(z.s2);
  P.cbuf = F(z.s2);
  P.iuwt_work = F(z.iuwt);
  P.model_local = F(z.N2);
  P.resid = F(z.N2);
  P.conv_full = F(z.N2);
  P.spec_sub = (C*)arena.alloc(z.spec_sub);
  P.spec_full_buf = (C*)arena.alloc(z.spec_full);
  P.spec_full = nullptr;
  P.fft_work = (C*)arena.alloc(z.fft);
  P.psf_frame = F(z.psf);
  P.pmax = F(z.pmax);
}

// Buffers grow to the largest geometry seen and are reused by smaller jobs of the same precision.
bool Job::ensure(const Geometry& g, hipStream_t st) {
  const int S = std::max(g.S, (int)std::log2((double)g.n) - 1);
  const JobSizes need = job_sizes(g, S);
  geo_ = g;
  if (have_ && fits(need, cap_)) return false;
  const JobSizes z = (have_ && cap_.f64 == need.f64) ? max_sizes(need, cap_) : need;
  arena_.release();
  cap_ = z;
  f = Planes<float>{};
  d = Planes<double>{};
  if (g.f64)
    carve(d, z, arena_);
  else
    carve(f, z, arena_);
  M = (uint8_t*)arena_.alloc(z.mask);
  mad_work = arena_.alloc(mad_work_bytes(16));
  mad_work_init(mad_work, 16, st);
  ext_work = arena_.alloc(z.ext);
  partial = (double*)arena_.alloc(z.partial);
  have_ = true;
  // the by-scale copy lives in its own arena and survives a regrowth of the others
  f.cur_dirty = (float*)by_scale_ptr_;
  d.cur_dirty = (double*)by_scale_ptr_;
  return true;
}

template <class T>
T* Job::by_scale_dirty(size_t n2) {
  const size_t bytes = n2 * sizeof(T);
  if (bytes > by_scale_bytes_) {
    by_scale_arena_.release();
    by_scale_ptr_ = by_scale_arena_.alloc(bytes);
    by_scale_bytes_ = bytes;
  }
  f.cur_dirty = (float*)by_scale_ptr_;
  d.cur_dirty = (double*)by_scale_ptr_;
  return (T*)by_scale_ptr_;
}

// --------------------------------------------------------------------------- task
Task::Task(const TaskSpec& s, Job* job, double* scal, int* iflag, double* upload_host)
    : spec(s), job_(job), scal_(scal), iflag_(iflag), upload_(upload_host) {
  if (spec.by_scale) {
    trace.push_back(ev(EV_BYSCALE, 0, spec.start_scale, 0, 0, 0.0, 0.0));
    start_call(spec.start_scale);
  } else {
    start_call(spec.prm.scale_count);
  }
}

void Task::fail(int code, const std::string& msg) {
  status = code;
  error = msg;
  phase_ = DONE;
  iterate_ = false;
}

void Task::start_call(int scale_count) {
  scale_count_ = scale_count;
  phase_ = SETUP;
}

void Task::enqueue_pre(hipStream_t st) {
  if (f64())
    enqueue_pre_t<double>(st);
  else
    enqueue_pre_t<float>(st);
}

void Task::add_shared(SharedOps& ops) {
  if (f64())
    add_shared_t<double>(ops);
  else
    add_shared_t<float>(ops);
}

void Task::enqueue_post(hipStream_t st) {
  if (f64())
    enqueue_post_t<double>(st);
  else
    enqueue_post_t<float>(st);
}

// Step work by phase.  pre: everything up to the shared part; shared: the decomposition of the dirty
// subregion (DECOMP), the recomposition of the extracted sources (first MINOR step), the full-frame
// model convolution (UPDATE / REVERT); post: the rest.  Stream order keeps pre < shared < post.
template <class T>
void Task::enqueue_pre_t(hipStream_t st) {
  if (upload_energies_ > 0) {
    PMS_HIP(hipMemcpyAsync(scal_ + kEnergy, upload_, sizeof(double) * upload_energies_, hipMemcpyHostToDevice, st));
    upload_energies_ = 0;
  }
  ++steps;
  Job& J = *job_;
  Planes<T>& B = J.pl<T>();
  const int sub = g_.sub;
  const size_t s2 = (size_t)sub * sub;
  switch (phase_) {
    case SETUP: enqueue_setup<T>(st); break;
    case DECOMP:
      if (spec.mask) {
        // decomposition of dirty_sub * mask[sub] replaces the thresholded one (main.py:283-285)
        decompose1<T>(B.dirty_sub, sub, g_.S, B.D, B.iuwt_work, st);
        mad_thresholds(B.D, g_.S, sub, sub, nm_, scal_ + kThr, J.mad_work, nullptr, st);
        window_copy(spec.mask, g_.n, g_.lo, g_.lo, sub, (const T*)nullptr, B.cbuf, st);
        window_copy(B.dirty_sub, sub, 0, 0, sub, B.cbuf, B.rec, st);
        decompose1<T>(B.rec, sub, g_.S, B.D, B.iuwt_work, st);
      }
      break;
    case MINOR:
      if (minor_first_) {
        // source extraction (main.py:347-355); masks/extracted live at planes [adj, max_scale) so the
        // decomposition can index them by scale
        const int adj = scale_adjust_;
        source_extract(B.D + (size_t)adj * s2, max_scale_ - adj, sub, sub, spec.prm.tolerance, spec.prm.neg_comp,
                       J.M + (size_t)adj * s2, B.E + (size_t)adj * s2, nullptr, J.ext_work, nullptr, st);
      }
      break;
    case UPDATE:
    case REVERT: {
      // model += gain x (a negative gain reverts the update), main.py:479-497
      const double gain = phase_ == UPDATE ? spec.prm.loop_gain : -spec.prm.loop_gain;
      hipLaunchKernelGGL(k_model_add<T>, dim3(grid_for(s2)), dim3(kT), 0, st, B.model_local, g_.n, g_.lo, g_.lo, sub,
                         (const T*)x_, gain);
      PMS_LAUNCH_CHECK();
      break;
    }
    case FINISH: enqueue_finish<T>(st); break;
    case DONE: break;
  }
}

template <class T>
void Task::add_shared_t(SharedOps& all) {
  SharedOpsT<T>& ops = all.of<T>();
  Planes<T>& B = job_->pl<T>();
  const int sub = g_.sub;
  const size_t s2 = (size_t)sub * sub;
  if (phase_ == DECOMP && !spec.mask) {
    ops.decomp.push_back({sub, IuwtDecompJobT<T>{B.dirty_sub, g_.S, 0, B.D, nullptr, B.iuwt_work, nullptr, nullptr,
                                                 nullptr, nullptr, 0}});
  } else if (phase_ == MINOR && minor_first_) {
    const int adj = scale_adjust_;
    ops.recomp.push_back({sub, IuwtRecompJobT<T>{B.E + (size_t)adj * s2, max_scale_ - adj, adj, nullptr, B.rec,
                                                 B.iuwt_work, nullptr, nullptr}});
  } else if (phase_ == UPDATE || phase_ == REVERT) {
    ops.conv.push_back(
        {{g_.n, g_.p_full}, typename EngineTypes<T>::Conv{B.model_local, B.spec_full, B.conv_full, B.fft_work, nullptr}});
  }
}

template <class T>
void Task::enqueue_post_t(hipStream_t st) {
  const MoresaneParams& prm = spec.prm;
  Job& J = *job_;
  Planes<T>& B = J.pl<T>();
  const int sub = g_.sub;
  const size_t s2 = (size_t)sub * sub;
  const size_t N2 = (size_t)g_.n * g_.n;
  switch (phase_) {
    case DECOMP:
      // MAD thresholds, thresholding with the edge mask and normalised maxima (main.py:278-300)
      if (!spec.mask) mad_thresholds(B.D, g_.S, sub, sub, nm_, scal_ + kThr, J.mad_work, nullptr, st);
      apply_threshold(B.D, B.D, g_.S, sub, sub, scal_ + kThr, prm.sigma_level,
                      borders_.empty() ? nullptr : iflag_ + kBorders, B.pmax, nullptr, st);
      max_final(B.pmax, threshold_grid(s2), g_.S, (T*)nullptr, scal_ + kEnergy, scal_ + kNsm, nullptr, st);
      break;
    case MINOR:
      if (minor_first_) {
        // CG set-up (main.py:365-377)
        const int ns = max_scale_ - scale_adjust_;
        const T* Es = B.E + (size_t)scale_adjust_ * s2;
        fill_zero((T*)x_, s2, st);
        // the incremental SNR's difference planes start at E - M.W(H x) = E (x = 0)
        copy_plane(Es, B.Rs, (size_t)ns * s2, nullptr, st);
        copy_plane(B.rec, (T*)r_, s2, nullptr, st);
        copy_plane(B.rec, B.p, s2, nullptr, st);
        dot_partials((const T*)r_, (const T*)nullptr, s2, J.partial, nullptr, st);
        sum_partials(J.partial, reduce_grid(s2), 1, rr_, nullptr, st);
        plane_sumsq_partials(Es, 1, (size_t)ns * s2, J.partial, nullptr, st);
        sum_partials(J.partial, reduce_grid((size_t)ns * s2), 1, scal_ + kE2, nullptr, st);
        hipLaunchKernelGGL(k_minor_reset, dim3(1), dim3(64), 0, st, scal_, iflag_, iterate_ ? 1 : 0);
        PMS_LAUNCH_CHECK();
      }
      break;
    case UPDATE:
    case REVERT: {
      // residual = dirty - H model, std of residual[sub], max of residual[sub] (main.py:497-506)
      if (spec.by_scale)
        hipLaunchKernelGGL((k_sub<T, T>), dim3(grid_for(N2)), dim3(kT), 0, st, (const T*)B.cur_dirty,
                           (const T*)B.conv_full, B.resid, N2);
      else
        hipLaunchKernelGGL((k_sub<float, T>), dim3(grid_for(N2)), dim3(kT), 0, st, spec.dirty, (const T*)B.conv_full,
                           B.resid, N2);
      PMS_LAUNCH_CHECK();
      window_copy(B.resid, g_.n, g_.lo, g_.lo, sub, (const T*)nullptr, B.dirty_sub, st);
      mean_std(B.dirty_sub, s2, J.partial, scal_ + kMeanSum, scal_ + kDev2, st);
      plane_max(B.dirty_sub, 1, s2, 0, B.pmax, nullptr, st);
      max_final(B.pmax, threshold_grid(s2), 1, (T*)(scal_ + kDmax), nullptr, nullptr, nullptr, st);
      break;
    }
    default: break;
  }
}

// psf_sub, PSF spectra and per-scale PSF energies (main.py:126-226), the dirty subregion and its
// maximum (main.py:229-262).
template <class T>
void Task::enqueue_setup(hipStream_t st) {
  const MoresaneParams& prm = spec.prm;
  int sub = 0;
  Geometry g = make_geometry(prm.n, prm.psf_n, prm, scale_count_, &sub);
  Job& J = *job_;
  if (J.ensure(g, st) || !x_) {
    moved_ = true;
    Planes<T>& P = J.pl<T>();
    x_ = P.x;
    xn_ = P.xn;
    r_ = P.r;
    rn_ = P.rn;
    rr_ = scal_ + kRR;
    rnrn_ = scal_ + kRNRN;
  }
  Planes<T>& B = J.pl<T>();
  g_ = J.geometry();
  const int S = g_.S;
  PMS_REQUIRE(S <= 16, "scale_count above 16 is not supported");
  const size_t N2 = (size_t)g_.n * g_.n;
  const size_t s2 = (size_t)g_.sub * g_.sub;
  if (spec.by_scale && !by_scale_copied_) {
    by_scale_copied_ = true;
    // the dirty image of each call is the previous residual; a private copy keeps the caller's intact
    T* cd = J.by_scale_dirty<T>(N2);
    window_copy(spec.dirty, g_.n, 0, 0, g_.n, (const T*)nullptr, cd, st);
  }
  B.spec_full = (g_.p_full == g_.p_sub && g_.sub == g_.n) ? B.spec_sub : B.spec_full_buf;
  const float* psf_dev = spec.psf;
  const int pn = g_.psf_n;
  const int plo = g_.double_psf ? pn / 2 - g_.sub / 2 : g_.lo;
  window_copy(psf_dev, pn, plo, plo, g_.sub, (const T*)nullptr, B.psf_sub, st);
  // spectra: the frame the reference transforms, expressed as (source window, q) for the spectrum
  using C = typename Planes<T>::C;
  auto spectrum_of = [&](int lo, int q, int p, C* dst) {
    if (lo == 0 && q == pn) {
      psf_spectrum_of(psf_dev, q, p, dst, B.fft_work, st);
    } else {
      window_copy(psf_dev, pn, lo, lo, q, (const T*)nullptr, B.psf_frame, st);
      psf_spectrum_of((const T*)B.psf_frame, q, p, dst, B.fft_work, st);
    }
  };
  const bool full = g_.sub == g_.n;
  if (g_.linear) {
    // a 2N PSF is cropped to the frame; an N x N PSF is zero-padded to 2N (pad_array): q = N
    if (full) {
      spectrum_of(0, pn, g_.p_full, B.spec_sub);
    } else {
      spectrum_of(pn / 2 - g_.sub, 2 * g_.sub, g_.p_sub, B.spec_sub);
      spectrum_of(0, pn, g_.p_full, B.spec_full);
    }
  } else {
    if (g_.double_psf) {
      spectrum_of(pn / 2 - g_.sub / 2, g_.sub, g_.p_sub, B.spec_sub);
      if (B.spec_full != B.spec_sub) spectrum_of(pn / 2 - g_.n / 2, g_.n, g_.p_full, B.spec_full);
    } else {
      spectrum_of(g_.lo, g_.sub, g_.p_sub, B.spec_sub);
      if (B.spec_full != B.spec_sub) spectrum_of(0, pn, g_.p_full, B.spec_full);
    }
  }
  // squared energies of the PSF decomposition; their square roots come back as an upload
  decompose1<T>(B.psf_sub, g_.sub, S, B.D, B.iuwt_work, st);
  plane_sumsq_partials(B.D, S, s2, J.partial, nullptr, st);
  sum_partials(J.partial, reduce_grid(s2), S, scal_ + kEnergy, nullptr, st);
  nm_ = noise_mask(g_.sub, g_.sub, prm.edge_excl, prm.int_excl);
  borders_ = suppression_borders(S, prm.edge_suppression, prm.edge_offset);
  if (!borders_.empty()) {
    int* ub = (int*)(upload_ + 32);
    std::memcpy(ub, borders_.data(), sizeof(int) * S);
    PMS_HIP(hipMemcpyAsync(iflag_ + kBorders, ub, sizeof(int) * S, hipMemcpyHostToDevice, st));
  }
  // model (local to this call) starts at zero; dirty subregion window and its maximum
  fill_zero(B.model_local, N2, st);
  if (spec.by_scale)
    window_copy((const T*)B.cur_dirty, g_.n, g_.lo, g_.lo, g_.sub, (const T*)nullptr, B.dirty_sub, st);
  else
    window_copy(spec.dirty, g_.n, g_.lo, g_.lo, g_.sub, (const T*)nullptr, B.dirty_sub, st);
  plane_max(B.dirty_sub, 1, s2, 0, B.pmax, nullptr, st);
  max_final(B.pmax, threshold_grid(s2), 1, (T*)(scal_ + kDmax), nullptr, nullptr, nullptr, st);
}

template <class T>
void Task::enqueue_finish(hipStream_t st) {
  Planes<T>& B = job_->pl<T>();
  const size_t N2 = (size_t)g_.n * g_.n;
  if (major_ > 0) {
    hipLaunchKernelGGL((k_add_inplace<float, T>), dim3(grid_for(N2)), dim3(kT), 0, st, spec.model,
                       (const T*)B.model_local, N2);
    PMS_LAUNCH_CHECK();
    if constexpr (sizeof(T) == sizeof(float)) {
      copy_plane(B.resid, spec.residual, N2, nullptr, st);
    } else {
      hipLaunchKernelGGL(k_to_f32, dim3(grid_for(N2)), dim3(kT), 0, st, (const double*)B.resid, spec.residual, N2);
      PMS_LAUNCH_CHECK();
    }
  }
  // moresane_by_scale: the residual is the next call's dirty image (main.py:584), kept in the job's
  // precision; a call that did no work leaves the residual as it was
  if (spec.by_scale) {
    if (major_ > 0)
      copy_plane((const T*)B.resid, B.cur_dirty, N2, nullptr, st);
    else
      window_copy((const float*)spec.residual, g_.n, 0, 0, g_.n, (const T*)nullptr, B.cur_dirty, st);
  }
}

IterArgs Task::iter_args() const {
  const int f = f64() ? 1 : 0;
  void* p = f ? (void*)job_->d.p : (void*)job_->f.p;
  void* ap = f ? (void*)job_->d.ap : (void*)job_->f.ap;
  return IterArgs{x_, xn_, r_, rn_, p, ap, rr_, rnrn_, max_scale_, scale_adjust_, spec.prm.enforce_positivity,
                  g_.sub, g_.p_sub, f};
}

// ---- host decisions
// the subregion maximum, stored in the slot as the job's element type
double Task::read_dmax(const double* h) const {
  if (f64()) return h[kDmax];
  float v;
  std::memcpy(&v, h + kDmax, sizeof(float));
  return (double)v;
}

void Task::consume(const double* h) {
  switch (phase_) {
    case SETUP: {
      for (int i = 0; i < g_.S; ++i) upload_[i] = std::sqrt(h[kEnergy + i]);
      upload_energies_ = g_.S;
      dmax_ = read_dmax(h);
      major_ = 0;
      max_coeff_ = 1.0;
      std_current_ = 1000.0;
      std_ratio_ = 1.0;
      min_scale_ = 0;
      nsm_.assign(g_.S, 0.0);
      major_check();
      break;
    }
    case DECOMP:
      for (int i = 0; i < g_.S; ++i) nsm_[i] = h[kNsm + i];
      choose_scales();
      break;
    case MINOR: {
      if (minor_first_) {
        e2_ = h[kE2];
        minor_first_ = false;
      }
      if (!iterate_) {
        minor_exit();
        break;
      }
      // The step ran up to K iterations; the device applied the exit rules after each
      // (k_minor_ctl).  Replay the iterate swaps of the iterations that ran: every iteration that
      // did not exit swaps (x, xn), (r, rn), (rr, rnrn) (main.py:456-457), and so does the one that
      // reaches the cap; an accepted exit takes xn as x (main.py:441).
      const int niter = (int)h[kNiterD];
      const int code = (int)h[kExitD];
      const int ran = niter - niter_;
      PMS_REQUIRE(ran >= 1, "minor loop step ran no iteration");
      niter_ = niter;
      snr_last_ = h[kSnrLast];
      snr_current_ = h[kSnrCur];
      const int swaps = (code < 0 || code == EXIT_CAP) ? ran : ran - 1;
      for (int k = 0; k < swaps; ++k) {
        std::swap(r_, rn_);
        std::swap(x_, xn_);
        std::swap(rr_, rnrn_);
      }
      if (code < 0) break;  // still iterating: the next step queues more iterations
      exit_code_ = code;
      if (code == EXIT_FIRST_HIGH || code == EXIT_FELL_LOW) {
        ++min_scale_;
      } else if (code == EXIT_ACCEPT) {
        std::swap(x_, xn_);
        min_scale_ = 0;
      } else if (code == EXIT_FELL_OK) {
        min_scale_ = 0;
      }
      minor_exit();
      break;
    }
    case UPDATE: {
      const size_t s2 = (size_t)g_.sub * g_.sub;
      const double std_last = std_current_;
      std_current_ = std::sqrt(h[kDev2] / (double)s2);
      std_ratio_ = (std_last - std_current_) / std_last;
      if (std_ratio_ < 0) {
        phase_ = REVERT;
        break;
      }
      trace.push_back(ev(EV_UPDATE, major_, 0, 0, 0, std_current_, std_ratio_));
      dmax_ = read_dmax(h);
      ++major_;
      major_check();
      break;
    }
    case REVERT:
      trace.push_back(ev(EV_UPDATE, major_, 1, 0, 0, std_current_, std_ratio_));
      dmax_ = read_dmax(h);
      ++major_;
      major_check();
      break;
    case FINISH: {
      major_iters = major_;
      if (!spec.by_scale) {
        phase_ = DONE;
        break;
      }
      ++scale_count_;
      if (complete || scale_count_ > std::log2((double)spec.prm.n) - 1 || scale_count_ > spec.stop_scale) {
        phase_ = DONE;
        break;
      }
      trace.push_back(ev(EV_BYSCALE, 0, scale_count_, 0, 0, 0.0, 0.0));
      start_call(scale_count_);
      break;
    }
    case DONE: break;
  }
}

// top of the major loop (main.py:266)
void Task::major_check() {
  const MoresaneParams& prm = spec.prm;
  if (major_ < prm.major_loop_miter && max_coeff_ > 0 && std_ratio_ > prm.accuracy &&
      dmax_ > prm.flux_threshold) {
    max_scale_ = 0;
    scale_adjust_ = 0;
    inner_check();
  } else {
    phase_ = FINISH;
  }
}

// top of the scale-selection loop (main.py:274)
void Task::inner_check() {
  if (min_scale_ < g_.S) {
    if (min_scale_ == 0)
      phase_ = DECOMP;
    else
      choose_scales();
  } else {
    after_inner();
  }
}

// scale choice from the normalised maxima (main.py:304-340)
void Task::choose_scales() {
  const int S = g_.S;
  int max_index = min_scale_;
  for (int i = min_scale_ + 1; i < S; ++i)
    if (nsm_[i] > nsm_[max_index]) max_index = i;
  max_scale_ = max_index + 1;
  max_coeff_ = nsm_[max_index];
  if (max_coeff_ == 0) {
    trace.push_back(ev(EV_NOCOEFF, major_, 0, 0, 0, 0.0, 0.0));
    after_inner();
    return;
  }
  scale_adjust_ = 0;
  for (int i = max_index - 1; i >= 0; --i)
    if (nsm_[i] == 0) {
      scale_adjust_ = i + 1;
      break;
    }
  trace.push_back(ev(EV_SCALES, major_, min_scale_, max_scale_, scale_adjust_, max_coeff_, 0.0));
  minor_first_ = true;
  iterate_ = spec.prm.minor_loop_miter > 0;
  niter_ = 0;
  exit_code_ = EXIT_CAP;
  snr_last_ = 0.0;
  snr_current_ = 0.0;
  phase_ = MINOR;
}

// end of one minor loop (main.py:459-475)
void Task::minor_exit() {
  iterate_ = false;
  trace.push_back(ev(EV_MINOR, major_, niter_, exit_code_, 0, snr_current_, 0.0));
  if (niter_ == spec.prm.minor_loop_miter && snr_current_ > 10.5) {
    min_scale_ = 0;
    after_inner();
    return;
  }
  if (min_scale_ == 0) {
    after_inner();
    return;
  }
  PMS_REQUIRE(!(exit_code_ == EXIT_CAP && min_scale_ > 0),
              "minor loop hit its cap at SNR <= 10.5 with min_scale > 0: the reference loops forever here");
  inner_check();
}

// after the scale-selection loop (main.py:477-516)
void Task::after_inner() {
  if (min_scale_ == g_.S) {
    trace.push_back(ev(EV_POOR, major_, 0, 0, 0, 0.0, 0.0));
    phase_ = FINISH;
    return;
  }
  if (max_coeff_ > 0) {
    phase_ = UPDATE;
    return;
  }
  if (major_ == 0) {
    trace.push_back(ev(EV_NOWORK, 0, 0, 0, 0, 0.0, 0.0));
    complete = 1;
    phase_ = FINISH;
    return;
  }
  major_check();
}

// --------------------------------------------------------------------------- runner
Runner::Runner(int slots, hipStream_t st) : st_(st) {
  PMS_REQUIRE(slots >= 1 && slots <= 1024, "runner slots must be in 1..1024");
  if (!st_) {
    PMS_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    owns_stream_ = true;
  }
  jobs_.resize(slots);
  tasks_.resize(slots);
  for (auto& j : jobs_) j = std::make_unique<Job>();
  PMS_HIP(hipMalloc(&scal_dev_, sizeof(double) * kSlotDoubles * slots));
  PMS_HIP(hipMalloc(&iflag_dev_, sizeof(int) * kSlotInts * slots));
  fill_u32((uint32_t*)scal_dev_, (size_t)2 * kSlotDoubles * slots, 0u, st_);
  fill_u32((uint32_t*)iflag_dev_, (size_t)kSlotInts * slots, 0u, st_);  // flags and finish counters start at 0
  PMS_HIP(hipHostMalloc((void**)&scal_host_, sizeof(double) * kSlotDoubles * slots, hipHostMallocDefault));
  PMS_HIP(hipHostMalloc((void**)&upload_host_, sizeof(double) * kSlotDoubles * slots, hipHostMallocDefault));
  refresh_options();
  PMS_HIP(hipStreamSynchronize(st_));
}

// Replaying a captured graph of the iterations is bit-identical to eager launches but measured slower
// for large frames on ROCm 7, so it is opt-in (PMS_CG_GRAPH=1).
void Runner::refresh_options() {
  const char* graph_env = getenv("PMS_CG_GRAPH");
  use_graph_ = graph_env && graph_env[0] == '1';
  // PMS_SNR_DIRECT=1: SNR sums from a convolution and a decomposition of xn each iteration, as
  // main.py:413-424 writes them, instead of the incremental update (same quantity by linearity)
  const char* snr_env = getenv("PMS_SNR_DIRECT");
  snr_direct_ = snr_env && snr_env[0] == '1';
  // PMS_CG_ITERS=K: minor-loop iterations queued per host step (the device applies the exit rules)
  const char* k_env = getenv("PMS_CG_ITERS");
  cg_iters_ = k_env ? std::max(1, std::min(32, atoi(k_env))) : (jobs_.size() == 1 ? kDefaultCgIters : 1);
}

Runner::~Runner() {
  if (st_) (void)hipStreamSynchronize(st_);
  drop_graphs();
  tasks_.clear();
  jobs_.clear();
  if (scal_dev_) (void)hipFree(scal_dev_);
  if (iflag_dev_) (void)hipFree(iflag_dev_);
  if (scal_host_) (void)hipHostFree(scal_host_);
  if (upload_host_) (void)hipHostFree(upload_host_);
  if (owns_stream_) (void)hipStreamDestroy(st_);
}

int Runner::free_slots() const {
  int n = 0;
  for (auto& t : tasks_)
    if (!t) ++n;
  return n;
}

bool Runner::idle() const {
  for (auto& t : tasks_)
    if (t && !t->done()) return false;
  return true;
}

size_t Runner::device_bytes() const {
  size_t b = 0;
  for (auto& j : jobs_) b += j->device_bytes();
  return b;
}

int Runner::first_free() const {
  for (int k = 0; k < slots(); ++k)
    if (!tasks_[k]) return k;
  return -1;
}

int Runner::add(const TaskSpec& spec, int slot) {
  const int k = slot >= 0 ? slot : first_free();
  if (k < 0) return -1;
  PMS_REQUIRE(k < slots() && !tasks_[k], "runner slot is not free");
  tasks_[k] = std::make_unique<Task>(spec, jobs_[k].get(), scal_dev_ + (size_t)k * kSlotDoubles,
                                     iflag_dev_ + (size_t)k * kSlotInts, upload_host_ + (size_t)k * kSlotDoubles);
  return k;
}

void Runner::release(int slot) { tasks_[slot].reset(); }

void Runner::run() {
  while (!idle()) step();
}

void Runner::step() {
  std::vector<int> active;
  for (int k = 0; k < slots(); ++k)
    if (tasks_[k] && !tasks_[k]->done()) active.push_back(k);
  if (active.empty()) return;
  bool moved = false;
  for (int k : active) {
    Task& t = *tasks_[k];
    try {
      t.enqueue_pre(st_);
    } catch (const Error& e) {
      t.fail(e.code, e.msg);
    }
    moved = moved || t.geometry_moved();
  }
  if (moved) drop_graphs();
  run_shared(active);
  for (int k : active) {
    Task& t = *tasks_[k];
    if (t.done()) continue;
    try {
      t.enqueue_post(st_);
    } catch (const Error& e) {
      t.fail(e.code, e.msg);
    }
  }
  std::vector<Task*> its;
  for (int k : active)
    if (tasks_[k]->wants_iteration()) its.push_back(tasks_[k].get());
  if (!its.empty()) run_iterations(its);
  PMS_HIP(hipMemcpyAsync(scal_host_, scal_dev_, sizeof(double) * kSlotDoubles * slots(), hipMemcpyDeviceToHost, st_));
  const auto t0 = std::chrono::steady_clock::now();
  PMS_HIP(hipStreamSynchronize(st_));
  wait_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k : active) {
    Task& t = *tasks_[k];
    if (t.done()) continue;
    try {
      t.consume(scal_host_ + (size_t)k * kSlotDoubles);
    } catch (const Error& e) {
      t.fail(e.code, e.msg);
    }
  }
}

// The shared part of every task's step, one batched call per plane side / frame and element type.
void Runner::run_shared(const std::vector<int>& active) {
  SharedOps ops;
  for (int k : active)
    if (!tasks_[k]->done()) tasks_[k]->add_shared(ops);
  run_shared_t<float>(ops.f);
  run_shared_t<double>(ops.d);
}

template <class T>
void Runner::run_shared_t(SharedOpsT<T>& ops) {
  std::map<int, std::vector<IuwtDecompJobT<T>>> dec;
  std::map<int, std::vector<IuwtRecompJobT<T>>> rec;
  std::map<std::pair<int, int>, std::vector<typename EngineTypes<T>::Conv>> conv;
  for (auto& e : ops.decomp) dec[e.first].push_back(e.second);
  for (auto& e : ops.recomp) rec[e.first].push_back(e.second);
  for (auto& e : ops.conv) conv[e.first].push_back(e.second);
  for (auto& kv : dec) decompose_batch(kv.second.data(), (int)kv.second.size(), kv.first, kv.first, st_);
  for (auto& kv : rec) recompose_batch(kv.second.data(), (int)kv.second.size(), kv.first, kv.first, st_);
  for (auto& kv : conv) conv_batch(kv.second.data(), (int)kv.second.size(), kv.first.first, kv.first.second, st_);
}

// The iterate buffers of iteration k of a step: every iteration that does not exit swaps (x, xn),
// (r, rn) and (rr, rnrn), so odd iterations run on the swapped set.
static std::vector<IterArgs> iteration_buffers(const std::vector<IterArgs>& its, int k) {
  std::vector<IterArgs> out = its;
  if (k & 1)
    for (IterArgs& a : out) {
      std::swap(a.x, a.xn);
      std::swap(a.r, a.rn);
      std::swap(a.rr, a.rnrn);
    }
  return out;
}

// Iterations of tasks that share an element type, plane shape and FFT frame run together, kMaxBatch
// per launch, cg_iters_ of them per step: the device's exit rules turn a finished job's remaining
// iterations into no-ops, and the host reads the outcome once per step.
void Runner::run_iterations(std::vector<Task*>& ts) {
  std::map<std::tuple<int, int, int>, std::vector<Task*>> parts;
  for (Task* t : ts) {
    const IterArgs a = t->iter_args();
    parts[std::make_tuple(a.f64, a.sub, a.p_sub)].push_back(t);
  }
  const int K = cg_iters_;
  for (auto& kv : parts) {
    std::vector<Task*>& v = kv.second;
    for (size_t j0 = 0; j0 < v.size(); j0 += kMaxBatch) {
      const size_t nb = std::min(v.size() - j0, (size_t)kMaxBatch);
      std::vector<Task*> chunk(v.begin() + j0, v.begin() + j0 + nb);
      std::vector<IterArgs> its;
      for (Task* t : chunk) its.push_back(t->iter_args());
      if (!use_graph_) {
        for (int k = 0; k < K; ++k) enqueue_iterations(iteration_buffers(its, k), chunk);
        continue;
      }
      // replay a captured graph of these K iterations on this exact set of buffers and scales
      // (recorded on first use)
      const auto key = std::make_pair(K, its);
      auto it = graphs_.find(key);
      if (it == graphs_.end()) {
        PMS_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        try {
          for (int k = 0; k < K; ++k) enqueue_iterations(iteration_buffers(its, k), chunk);
        } catch (...) {
          hipGraph_t junk = nullptr;
          (void)hipStreamEndCapture(st_, &junk);
          if (junk) (void)hipGraphDestroy(junk);
          throw;
        }
        hipGraph_t graph = nullptr;
        PMS_HIP(hipStreamEndCapture(st_, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PMS_HIP(e);
        it = graphs_.emplace(key, exec).first;
      }
      PMS_HIP(hipGraphLaunch(it->second, st_));
    }
  }
}

void Runner::drop_graphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

// Ap <- R(M . W(H v)) over scales [adj, max_scale) for every task (main.py:380-390); the clip pass
// recomputes it for tasks whose positivity clip fired (predicated on the device flag).
template <class T>
void Runner::apply_op(const std::vector<Task*>& ts, const std::vector<IterArgs>& its, bool clip_pass) {
  const int sub = its[0].sub, p_sub = its[0].p_sub;
  const size_t s2 = (size_t)sub * sub;
  std::vector<typename EngineTypes<T>::Conv> fj;
  std::vector<IuwtDecompJobT<T>> dj;
  std::vector<IuwtRecompJobT<T>> rj;
  for (size_t i = 0; i < ts.size(); ++i) {
    Job& J = *ts[i]->job();
    Planes<T>& B = J.pl<T>();
    const IterArgs& a = its[i];
    const int* pred = ts[i]->iflag() + (clip_pass ? kClip : kActive);
    fj.push_back(typename EngineTypes<T>::Conv{(const T*)a.p, B.spec_sub, B.cbuf, B.fft_work, pred});
    dj.push_back(IuwtDecompJobT<T>{B.cbuf, a.max_scale, a.adj, B.Dcg, nullptr, B.iuwt_work, J.M + (size_t)a.adj * s2,
                                   nullptr, nullptr, pred, 0});
    rj.push_back(IuwtRecompJobT<T>{B.Dcg, a.max_scale - a.adj, a.adj, nullptr, (T*)a.ap, B.iuwt_work, nullptr, pred});
  }
  conv_batch(fj.data(), (int)fj.size(), sub, p_sub, st_);
  decompose_batch(dj.data(), (int)dj.size(), sub, sub, st_);
  recompose_batch(rj.data(), (int)rj.size(), sub, sub, st_);
}

void Runner::enqueue_iterations(const std::vector<IterArgs>& its, const std::vector<Task*>& ts) {
  if (its[0].f64)
    enqueue_iterations_t<double>(its, ts);
  else
    enqueue_iterations_t<float>(its, ts);
}

// One CG iteration (main.py:380-424), its SNR sums and its exit rules (main.py:427-457) for up to
// kMaxBatch tasks of one element type.
template <class T>
void Runner::enqueue_iterations_t(const std::vector<IterArgs>& its, const std::vector<Task*>& ts) {
  const int nb = (int)ts.size();
  const int sub = its[0].sub, p_sub = its[0].p_sub;
  const size_t s2 = (size_t)sub * sub;
  CgBatch<T> cb{};
  cb.n = s2;
  std::vector<Task*> clip_ts;
  std::vector<IterArgs> clip_its;
  CgBatch<T> clip_b{};
  clip_b.n = s2;
  for (int i = 0; i < nb; ++i) {
    const IterArgs& a = its[i];
    Job& J = *ts[i]->job();
    cb.j[i] = CgJob<T>{(T*)a.x,     (T*)a.xn,      (T*)a.r, (T*)a.rn,        (T*)a.p,
                       (T*)a.ap,    J.partial,     ts[i]->scal(), a.rr,       a.rnrn,
                       ts[i]->iflag(), a.enforce};
    if (a.enforce) {
      clip_b.j[clip_ts.size()] = cb.j[i];
      clip_ts.push_back(ts[i]);
      clip_its.push_back(a);
    }
  }
  apply_op<T>(ts, its, false);
  launch_cg(k_cg_pap<T>, cb, nb, st_);
  launch_cg_after(k_cg_xn<T>, cb, nb, st_);
  if (!clip_ts.empty()) {
    launch_cg_after(k_cg_clip<T>, clip_b, (int)clip_ts.size(), st_);
    apply_op<T>(clip_ts, clip_its, true);
  }
  // rn, and in the incremental form the SNR sums (main.py:413-424) in the same launch
  RnBatch<T> rb{};
  rb.n = s2;
  rb.np = cg_grid(s2);
  int gx = rb.np;
  CtlBatch ctl{};
  for (int i = 0; i < nb; ++i) {
    Job& J = *ts[i]->job();
    Planes<T>& B = J.pl<T>();
    const IterArgs& a = its[i];
    const size_t n = (size_t)(a.max_scale - a.adj) * s2;
    const int nblk = cg_grid(n);
    rb.j[i] = RnJob<T>{(const T*)a.r, (const T*)a.ap, (T*)a.rn, snr_direct_ ? nullptr : B.Rs, B.Dcg, J.partial,
                       ts[i]->scal(), ts[i]->iflag(), n, nblk};
    if (!snr_direct_) gx = std::max(gx, nblk);
    ctl.scal[i] = ts[i]->scal();
    ctl.iflag[i] = ts[i]->iflag();
    ctl.partial[i] = J.partial + kPartialHi;
    ctl.np[i] = nblk;
    ctl.miter[i] = ts[i]->spec.prm.minor_loop_miter;
  }
  hipLaunchKernelGGL(k_cg_rn<T>, dim3(gx, nb), dim3(kT), 0, st_, rb);
  PMS_LAUNCH_CHECK();
  launch_cg_after(k_cg_p<T>, cb, nb, st_);
  if (snr_direct_) {
    // main.py:413-424 as written: convolve xn, decompose it against the extracted sources
    std::vector<typename EngineTypes<T>::Conv> fj;
    std::vector<IuwtDecompJobT<T>> dj;
    for (int i = 0; i < nb; ++i) {
      Job& J = *ts[i]->job();
      Planes<T>& B = J.pl<T>();
      const IterArgs& a = its[i];
      const int* act = ts[i]->iflag() + kActive;
      fj.push_back(typename EngineTypes<T>::Conv{(const T*)a.xn, B.spec_sub, B.cbuf, B.fft_work, act});
      dj.push_back(IuwtDecompJobT<T>{B.cbuf, a.max_scale, a.adj, nullptr, nullptr, B.iuwt_work,
                                     J.M + (size_t)a.adj * s2, B.E + (size_t)a.adj * s2, J.partial, act, 0});
    }
    conv_batch(fj.data(), nb, sub, p_sub, st_);
    decompose_batch(dj.data(), nb, sub, sub, st_);
    for (int i = 0; i < nb; ++i) {
      ctl.partial[i] = ts[i]->job()->partial;
      ctl.np[i] = dj[i].np;
    }
  }
  // exit rules of every job of the chunk (main.py:427-457)
  hipLaunchKernelGGL(k_minor_ctl, dim3(nb), dim3(kT), 0, st_, ctl);
  PMS_LAUNCH_CHECK();
}

}  // namespace pms
