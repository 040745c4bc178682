// Device-resident MORESANE driver (major loop, conjugate-gradient minor loop,
// by-scale outer loop) for one job on one HIP stream.
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
//     one host sync per iteration reads the SNR sums for the exit rules;
//   * model update and residual (main.py:479-506), with revert.
// Kernels run with a null predicate here; the predicate mechanism exists so a
// caller can enqueue work whose execution depends on device state.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine.h"

namespace pms {
namespace {

constexpr int kT = 256;




// positivity clip: xn[xn<0] = 0 ; p = (xn - x) / alpha
__global__ __launch_bounds__(kT) void k_cg_clip(float* __restrict__ xn, const float* __restrict__ x,
                                                float* __restrict__ p, const double* __restrict__ alpha, size_t n,
                                                const int* flag) {
  if (*flag == 0) return;
  const double a = *alpha;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    float v = xn[i];
    if (v < 0.f) v = 0.f;
    xn[i] = v;
    p[i] = (float)(((double)v - (double)x[i]) / a);
  }
}


// p = rn + beta p
__global__ __launch_bounds__(kT) void k_cg_p(const float* __restrict__ rn, float* __restrict__ p,
                                             const double* __restrict__ beta, size_t n) {
  const double b = *beta;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
    p[i] = (float)fma(b, (double)p[i], (double)rn[i]);
}

// model[r0+r, c0+c] += gain * x[r, c]
__global__ __launch_bounds__(kT) void k_model_add(float* __restrict__ model, int ld, int r0, int c0, int n,
                                                  const float* __restrict__ x, double gain) {
  const size_t tot = (size_t)n * n;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < tot; i += (size_t)gridDim.x * kT) {
    const int r = (int)(i / n), c = (int)(i % n);
    float* m = model + (size_t)(r0 + r) * ld + (c0 + c);
    *m = *m + (float)(gain * (double)x[i]);
  }
}

// out = a - b
__global__ __launch_bounds__(kT) void k_sub(const float* __restrict__ a, const float* __restrict__ b,
                                            float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) out[i] = a[i] - b[i];
}

// plane += other
__global__ __launch_bounds__(kT) void k_add_inplace(float* __restrict__ a, const float* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) a[i] += b[i];
}

// Deterministic grid-wide finish: every block publishes its partial, the last block to arrive sums
// the partials in block order (fixed tree) and runs `fin` on the total; the counter is left at 0.
template <class Fin>
__device__ __forceinline__ void finish_sum(double part, double* partial, unsigned* counter, Fin fin) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = part;
    __threadfence();
    is_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double acc = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) acc += ((volatile const double*)partial)[i];
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) {
    fin(acc);
    *counter = 0u;
  }
}

// alpha = rr / (p . Ap)
__global__ __launch_bounds__(kT) void k_cg_alpha(const float* __restrict__ p, const float* __restrict__ ap, size_t n,
                                                 double* partial, unsigned* counter, const double* rr, double* pap,
                                                 double* alpha) {
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
    acc = fma((double)p[i], (double)ap[i], acc);
  acc = block_sum_f64(acc);
  finish_sum(acc, partial, counter, [&](double tot) {
    *pap = tot;
    *alpha = *rr / tot;
  });
}

// xn = x + alpha p ; clip flag = enforce && min(xn) < 0
__global__ __launch_bounds__(kT) void k_cg_xn_flag(const float* __restrict__ x, const float* __restrict__ p,
                                                   const double* __restrict__ alpha, float* __restrict__ xn, size_t n,
                                                   double* partial, unsigned* counter, int enforce, int* flag) {
  const double a = *alpha;
  float m = INFINITY;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const float v = (float)fma(a, (double)p[i], (double)x[i]);
    xn[i] = v;
    m = fminf(m, v);
  }
  m = block_min_f32(m);
  // min via the same last-block pattern: publish min as a double, reduce with min in block order
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = (double)m;
    __threadfence();
    is_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  if (threadIdx.x == 0) {
    double mn = INFINITY;
    for (unsigned i = 0; i < gridDim.x; ++i) mn = fmin(mn, ((volatile const double*)partial)[i]);
    *flag = (enforce && mn < 0.0) ? 1 : 0;
    *counter = 0u;
  }
}

// rn = r - alpha Ap ; rnrn = rn . rn ; beta = rnrn / rr
__global__ __launch_bounds__(kT) void k_cg_rn_beta(const float* __restrict__ r, const float* __restrict__ ap,
                                                   const double* __restrict__ alpha, float* __restrict__ rn, size_t n,
                                                   double* partial, unsigned* counter, const double* rr, double* rnrn,
                                                   double* beta) {
  const double a = *alpha;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
    const float v = (float)fma(-a, (double)ap[i], (double)r[i]);
    rn[i] = v;
    acc = fma((double)v, (double)v, acc);
  }
  acc = block_sum_f64(acc);
  finish_sum(acc, partial, counter, [&](double tot) {
    *rnrn = tot;
    *beta = tot / *rr;
  });
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
  PMS_REQUIRE(g.double_psf || psf_n == n, "PSF must be the dirty image size or twice it");
  if (g.linear) {
    g.p_sub = 2 * sub;
    g.p_full = 2 * n;
  } else {
    g.p_sub = sub;
    g.p_full = n;
  }
  PMS_REQUIRE(is_pow2(g.p_sub) && is_pow2(g.p_full), "FFT convolution needs power-of-two image sides");
  check_iuwt_scale(sub, sub, S - 1);
  if (subregion_out) *subregion_out = sub;
  return g;
}

// --------------------------------------------------------------------------- job
// Host waits on the job's stream, timed so callers can tell GPU-bound from host-bound runs.
void Job::sync() {
  const auto t0 = std::chrono::steady_clock::now();
  PMS_HIP(hipStreamSynchronize(st_));
  wait_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// A job handed the legacy NULL stream runs on a stream of its own (graph capture needs one); every
// moresane()/moresane_by_scale() call synchronises that stream before returning.
Job::Job(hipStream_t st) : st_(st) {
  if (!st_) {
    PMS_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    owns_stream_ = true;
  }
}
Job::~Job() {
  drop_graphs();
  arena_.release();
  by_scale_arena_.release();
  if (owns_stream_) (void)hipStreamDestroy(st_);
}

// Byte sizes of every per-job buffer for geometry g (S = scale capacity).
static JobSizes job_sizes(const Geometry& g, int S) {
  JobSizes z{};
  const size_t N2 = (size_t)g.n * g.n;
  const size_t s2 = (size_t)g.sub * g.sub;
  z.s2 = s2 * sizeof(float);
  z.S_s2 = (size_t)S * s2 * sizeof(float);
  z.mask = (size_t)S * s2;
  z.iuwt = 2 * s2 * sizeof(float);
  z.N2 = N2 * sizeof(float);
  z.spec_sub = fft_spectrum_elems(g.p_sub) * sizeof(float2);
  z.spec_full = fft_spectrum_elems(g.p_full) * sizeof(float2);
  z.fft = std::max({fft_conv_work_elems(g.sub, g.p_sub), fft_conv_work_elems(g.n, g.p_full),
                    fft_spectrum_elems(g.p_sub), fft_spectrum_elems(g.p_full)}) *
          sizeof(float2);
  z.psf = (size_t)g.psf_n * g.psf_n * sizeof(float);
  z.ext = extract_work_bytes(S, g.sub, g.sub);
  z.partial = (size_t)std::max<long>({(long)iuwt_decompose_partials(g.sub, g.sub, S, 0),
                                      (long)reduce_grid(N2) * 2, (long)reduce_grid(s2) * S, 4096L}) *
              sizeof(double);
  z.pmax = (size_t)S * 2048 * sizeof(float);
  return z;
}

static bool fits(const JobSizes& need, const JobSizes& have) {
  return need.s2 <= have.s2 && need.S_s2 <= have.S_s2 && need.mask <= have.mask && need.iuwt <= have.iuwt &&
         need.N2 <= have.N2 && need.spec_sub <= have.spec_sub && need.spec_full <= have.spec_full &&
         need.fft <= have.fft && need.psf <= have.psf && need.ext <= have.ext && need.partial <= have.partial &&
         need.pmax <= have.pmax;
}

static JobSizes max_sizes(const JobSizes& a, const JobSizes& b) {
  JobSizes z;
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
  return z.s2 * 10 + z.S_s2 * 3 + z.mask + z.iuwt + z.N2 * 3 + z.spec_sub + z.spec_full + z.fft + z.psf + z.ext +
         z.partial + z.pmax + 64 * 8 + 64 * 4 + 12 * 256;
}

// Buffers grow to the largest geometry seen and are reused by smaller jobs.
void Job::ensure(const Geometry& g) {
  const int S = std::max(g.S, (int)std::log2((double)g.n) - 1);
  const JobSizes need = job_sizes(g, S);
  geo_ = g;
  if (have_ && fits(need, cap_)) return;
  const JobSizes z = have_ ? max_sizes(need, cap_) : need;
  drop_graphs();
  arena_.release();
  cap_ = z;
  auto F = [&](size_t bytes) { return (float*)arena_.alloc(bytes); };
  dirty_sub = F(z.s2);
  psf_sub = F(z.s2);
  D = F(z.S_s2);
  Dcg = F(z.S_s2);
  E = F(z.S_s2);
  M = (uint8_t*)arena_.alloc(z.mask);
  rec = F(z.s2);
  x = F(z.s2);
  xn = F(z.s2);
  r = F(z.s2);
  rn = F(z.s2);
  p = F(z.s2);
  ap = F(z.s2);
  cbuf = F(z.s2);
  iuwt_work = F(z.iuwt);
  model_local = F(z.N2);
  resid = F(z.N2);
  conv_full = F(z.N2);
  spec_sub = (float2*)arena_.alloc(z.spec_sub);
  spec_full_buf = (float2*)arena_.alloc(z.spec_full);
  fft_work = (float2*)arena_.alloc(z.fft);
  psf_frame = F(z.psf);
  mad_work = arena_.alloc(mad_work_bytes(16));
  mad_work_init(mad_work, 16, st_);
  ext_work = arena_.alloc(z.ext);
  partial = (double*)arena_.alloc(z.partial);
  pmax = (float*)arena_.alloc(z.pmax);
  scal = (double*)arena_.alloc(64 * sizeof(double));
  iflag = (int*)arena_.alloc(64 * sizeof(int));
  fill_u32((uint32_t*)iflag, 64, 0u, st_);  // flags and the grid-finish counters start at zero
  border_dev = iflag + 16;
  have_ = true;
}

// psf_sub, PSF spectra and per-scale PSF energies (main.py:126-226)
void Job::setup_psf(const float* psf_dev, std::vector<double>& energies) {
  const Geometry& g = geo_;
  spec_full = (g.p_full == g.p_sub && g.sub == g.n) ? spec_sub : spec_full_buf;
  const int pn = g.psf_n;
  const int plo = g.double_psf ? pn / 2 - g.sub / 2 : g.lo;
  window_copy(psf_dev, pn, plo, plo, g.sub, nullptr, psf_sub, st_);
  // spectra: the frame the reference transforms, expressed as (source window, q) for fft_psf_spectrum
  auto spectrum_of = [&](int lo, int q, int p, float2* dst) {
    const float* src = psf_dev;
    if (!(lo == 0 && q == pn)) {
      window_copy(psf_dev, pn, lo, lo, q, nullptr, psf_frame, st_);
      src = psf_frame;
    }
    fft_psf_spectrum(src, q, p, dst, fft_work, st_);
  };
  const bool full = g.sub == g.n;
  if (g.linear) {
    if (g.double_psf) {
      if (full) {
        spectrum_of(0, pn, g.p_full, spec_sub);
      } else {
        spectrum_of(pn / 2 - g.sub, 2 * g.sub, g.p_sub, spec_sub);
        spectrum_of(0, pn, g.p_full, spec_full);
      }
    } else {
      // N x N PSF zero-padded to 2N (pad_array): q = N inside a 2N frame
      if (full) {
        spectrum_of(0, pn, g.p_full, spec_sub);
      } else {
        spectrum_of(pn / 2 - g.sub, 2 * g.sub, g.p_sub, spec_sub);
        spectrum_of(0, pn, g.p_full, spec_full);
      }
    }
  } else {
    if (g.double_psf) {
      spectrum_of(pn / 2 - g.sub / 2, g.sub, g.p_sub, spec_sub);
      if (spec_full != spec_sub) spectrum_of(pn / 2 - g.n / 2, g.n, g.p_full, spec_full);
    } else {
      spectrum_of(g.lo, g.sub, g.p_sub, spec_sub);
      if (spec_full != spec_sub) spectrum_of(0, pn, g.p_full, spec_full);
    }
  }
  // energies of the PSF decomposition
  const size_t s2 = (size_t)g.sub * g.sub;
  iuwt_decompose(psf_sub, g.sub, g.sub, g.S, 0, D, nullptr, iuwt_work, nullptr, nullptr, nullptr, nullptr, st_);
  const int gr = reduce_grid(s2);
  plane_sumsq_partials(D, g.S, s2, partial, nullptr, st_);
  sum_partials(partial, gr, g.S, scal + 32, nullptr, st_);
  std::vector<double> e2(g.S);
  PMS_HIP(hipMemcpyAsync(e2.data(), scal + 32, sizeof(double) * g.S, hipMemcpyDeviceToHost, st_));
  sync();
  energies.resize(g.S);
  for (int i = 0; i < g.S; ++i) energies[i] = std::sqrt(e2[i]);
  PMS_HIP(hipMemcpyAsync(scal + 32, energies.data(), sizeof(double) * g.S, hipMemcpyHostToDevice, st_));
}

// residual = dirty - H model (full image); returns std of residual[subregion]
double Job::residual_and_std(const float* dirty_dev, const float* model_dev) {
  const Geometry& g = geo_;
  const size_t N2 = (size_t)g.n * g.n;
  fft_conv(model_dev, g.n, g.p_full, spec_full, conv_full, fft_work, nullptr, st_);
  hipLaunchKernelGGL(k_sub, dim3(grid_for(N2)), dim3(kT), 0, st_, dirty_dev, (const float*)conv_full, resid, N2);
  PMS_LAUNCH_CHECK();
  window_copy(resid, g.n, g.lo, g.lo, g.sub, nullptr, dirty_sub, st_);
  const size_t s2 = (size_t)g.sub * g.sub;
  mean_std(dirty_sub, s2, partial, scal + 56, scal + 57, st_);
  double v[2];
  PMS_HIP(hipMemcpyAsync(v, scal + 56, sizeof(double) * 2, hipMemcpyDeviceToHost, st_));
  sync();
  return std::sqrt(v[1] / (double)s2);
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

int Job::moresane(const MoresaneParams& prm, int scale_count, const float* dirty_dev, const float* psf_dev,
                  const float* mask_dev, float* model_dev, float* residual_dev, int* complete, Trace& trace) {
  int sub = 0;
  Geometry g = make_geometry(prm.n, prm.psf_n, prm, scale_count, &sub);
  ensure(g);
  g = geo_;
  const int S = g.S;
  const size_t s2 = (size_t)g.sub * g.sub;
  const size_t N2 = (size_t)g.n * g.n;
  const int gr = reduce_grid(s2);
  std::vector<double> energies;
  setup_psf(psf_dev, energies);
  NoiseMask nm = noise_mask(g.sub, g.sub, prm.edge_excl, prm.int_excl);
  std::vector<int> borders = suppression_borders(S, prm.edge_suppression, prm.edge_offset);
  if (!borders.empty())
    PMS_HIP(hipMemcpyAsync(border_dev, borders.data(), sizeof(int) * S, hipMemcpyHostToDevice, st_));

  // model (local to this call) starts at zero; dirty subregion window
  fill_f32(model_local, N2, 0.f, st_);
  window_copy(dirty_dev, g.n, g.lo, g.lo, g.sub, nullptr, dirty_sub, st_);
  plane_max(dirty_sub, 1, s2, 0, pmax, nullptr, st_);
  max_final(pmax, threshold_grid(s2), 1, (float*)(scal + 60), nullptr, nullptr, nullptr, st_);
  float dmax = 0.f;
  PMS_HIP(hipMemcpyAsync(&dmax, scal + 60, sizeof(float), hipMemcpyDeviceToHost, st_));
  sync();

  double* thr = scal + 0;       // [0, 16)  thresholds (S <= 16)
  double* nsm_dev = scal + 16;  // [16, 32) normalised maxima
  double* energy_dev = scal + 32;
  double* d_rr = scal + 48;
  double* d_rnrn = scal + 51;
  double* d_d2 = scal + 53;
  double* d_e2 = scal + 54;
  PMS_REQUIRE(S <= 16, "scale_count above 16 is not supported");

  // Replaying the iteration from a captured HIP graph is bit-identical to eager launches but measured
  // slower on ROCm 7 for large frames, so it is opt-in (PMS_CG_GRAPH=1).
  const char* graph_env = getenv("PMS_CG_GRAPH");
  const bool use_graph = graph_env && graph_env[0] == '1';
  int major = 0;
  double max_coeff = 1.0;
  double std_current = 1000.0, std_ratio = 1.0;
  int min_scale = 0;
  std::vector<double> nsm(S, 0.0);
  const int pmx = threshold_grid(s2);

  while (major < prm.major_loop_miter && max_coeff > 0 && std_ratio > prm.accuracy &&
         (double)dmax > prm.flux_threshold) {
    int max_scale = 0, scale_adjust = 0;
    while (min_scale < S) {
      if (min_scale == 0) {
        iuwt_decompose(dirty_sub, g.sub, g.sub, S, 0, D, nullptr, iuwt_work, nullptr, nullptr, nullptr, nullptr, st_);
        mad_thresholds(D, S, g.sub, g.sub, nm, thr, mad_work, nullptr, st_);
        if (mask_dev) {
          // decomposition of dirty_sub * mask[sub] replaces the thresholded one (main.py:283-285)
          window_copy(mask_dev, g.n, g.lo, g.lo, g.sub, nullptr, cbuf, st_);
          window_copy(dirty_sub, g.sub, 0, 0, g.sub, cbuf, rec, st_);
          iuwt_decompose(rec, g.sub, g.sub, S, 0, D, nullptr, iuwt_work, nullptr, nullptr, nullptr, nullptr, st_);
        }
        apply_threshold(D, D, S, g.sub, g.sub, thr, prm.sigma_level, borders.empty() ? nullptr : border_dev, pmax,
                        nullptr, st_);
        max_final(pmax, pmx, S, nullptr, energy_dev, nsm_dev, nullptr, st_);
        PMS_HIP(hipMemcpyAsync(nsm.data(), nsm_dev, sizeof(double) * S, hipMemcpyDeviceToHost, st_));
        sync();
      }
      int max_index = min_scale;
      for (int i = min_scale + 1; i < S; ++i)
        if (nsm[i] > nsm[max_index]) max_index = i;
      max_scale = max_index + 1;
      max_coeff = nsm[max_index];
      if (max_coeff == 0) {
        trace.push_back(ev(EV_NOCOEFF, major, 0, 0, 0, 0.0, 0.0));
        break;
      }
      scale_adjust = 0;
      for (int i = max_index - 1; i >= 0; --i)
        if (nsm[i] == 0) {
          scale_adjust = i + 1;
          break;
        }
      trace.push_back(ev(EV_SCALES, major, min_scale, max_scale, scale_adjust, max_coeff, 0.0));
      const int ns = max_scale - scale_adjust;
      const float* tslice = D + (size_t)scale_adjust * s2;
      // masks/extracted live at planes [adj, max_scale) so the decomposition can index them by scale
      uint8_t* Ms = M + (size_t)scale_adjust * s2;
      float* Es = E + (size_t)scale_adjust * s2;
      source_extract(tslice, ns, g.sub, g.sub, prm.tolerance, prm.neg_comp, Ms, Es, nullptr, ext_work, nullptr, st_);
      iuwt_recompose(Es, ns, g.sub, g.sub, scale_adjust, nullptr, rec, iuwt_work, nullptr, nullptr, st_);

      // ---- CG minor loop
      fill_f32(x, s2, 0.f, st_);
      copy_plane(rec, r, s2, nullptr, st_);
      copy_plane(rec, p, s2, nullptr, st_);
      dot_partials(r, nullptr, s2, partial, nullptr, st_);
      sum_partials(partial, gr, 1, d_rr, nullptr, st_);
      plane_sumsq_partials(Es, 1, (size_t)ns * s2, partial, nullptr, st_);
      sum_partials(partial, reduce_grid((size_t)ns * s2), 1, d_e2, nullptr, st_);
      double e2 = 0.0;
      PMS_HIP(hipMemcpyAsync(&e2, d_e2, sizeof(double), hipMemcpyDeviceToHost, st_));

      int niter = 0;
      double snr_last = 0.0, snr_current = 0.0;
      int exit_code = EXIT_CAP;
      while (niter < prm.minor_loop_miter) {
        IterArgs ia{x, xn, r, rn, p, ap, d_rr, d_rnrn, max_scale, scale_adjust, prm.enforce_positivity, g.sub,
                    g.p_sub};
        run_iteration(ia, use_graph);
        double d2 = 0.0;
        PMS_HIP(hipMemcpyAsync(&d2, d_d2, sizeof(double), hipMemcpyDeviceToHost, st_));
        sync();
        snr_last = snr_current;
        snr_current = 20.0 * std::log10(std::sqrt(e2) / std::sqrt(d2));
        ++niter;
        if (niter == 1 && snr_current > 40) {
          ++min_scale;
          exit_code = EXIT_FIRST_HIGH;
          break;
        }
        if (snr_current > 40) {
          std::swap(x, xn);
          min_scale = 0;
          exit_code = EXIT_ACCEPT;
          break;
        }
        if (niter > 2 && snr_current <= snr_last) {
          if (snr_current > 10.5) {
            min_scale = 0;
            exit_code = EXIT_FELL_OK;
          } else {
            ++min_scale;
            exit_code = EXIT_FELL_LOW;
          }
          break;
        }
        std::swap(r, rn);
        std::swap(x, xn);
        std::swap(d_rr, d_rnrn);
      }
      trace.push_back(ev(EV_MINOR, major, niter, exit_code, 0, snr_current, 0.0));
      if (niter == prm.minor_loop_miter && snr_current > 10.5) {
        min_scale = 0;
        break;
      }
      if (min_scale == 0) break;
      PMS_REQUIRE(!(exit_code == EXIT_CAP && min_scale > 0),
                  "minor loop hit its cap at SNR <= 10.5 with min_scale > 0: the reference loops forever here");
    }

    if (min_scale == S) {
      trace.push_back(ev(EV_POOR, major, 0, 0, 0, 0.0, 0.0));
      break;
    }
    if (max_coeff > 0) {
      hipLaunchKernelGGL(k_model_add, dim3(grid_for(s2)), dim3(kT), 0, st_, model_local, g.n, g.lo, g.lo, g.sub,
                         (const float*)x, prm.loop_gain);
      PMS_LAUNCH_CHECK();
      const double std_last = std_current;
      std_current = residual_and_std(dirty_dev, model_local);
      std_ratio = (std_last - std_current) / std_last;
      const bool reverted = std_ratio < 0;
      if (reverted) {
        hipLaunchKernelGGL(k_model_add, dim3(grid_for(s2)), dim3(kT), 0, st_, model_local, g.n, g.lo, g.lo, g.sub,
                           (const float*)x, -prm.loop_gain);
        PMS_LAUNCH_CHECK();
        residual_and_std(dirty_dev, model_local);
      }
      trace.push_back(ev(EV_UPDATE, major, reverted ? 1 : 0, 0, 0, std_current, std_ratio));
      // dirty_sub already holds residual[sub] (written by residual_and_std)
      plane_max(dirty_sub, 1, s2, 0, pmax, nullptr, st_);
      max_final(pmax, pmx, 1, (float*)(scal + 60), nullptr, nullptr, nullptr, st_);
      PMS_HIP(hipMemcpyAsync(&dmax, scal + 60, sizeof(float), hipMemcpyDeviceToHost, st_));
      sync();
      ++major;
    }
    if (major == 0) {
      trace.push_back(ev(EV_NOWORK, 0, 0, 0, 0, 0.0, 0.0));
      *complete = 1;
      break;
    }
  }
  if (major > 0) {
    hipLaunchKernelGGL(k_add_inplace, dim3(grid_for(N2)), dim3(kT), 0, st_, model_dev, (const float*)model_local, N2);
    PMS_LAUNCH_CHECK();
    copy_plane(resid, residual_dev, N2, nullptr, st_);
  }
  sync();
  return major;
}

// One CG iteration (main.py:380-424) up to the SNR sums, enqueued on the job's stream.
void Job::enqueue_iteration(const IterArgs& a) {
  const Geometry& g = geo_;
  const size_t s2 = (size_t)g.sub * g.sub;
  const int gr = reduce_grid(s2);
  double* d_pap = scal + 49;
  double* d_alpha = scal + 50;
  double* d_beta = scal + 52;
  double* d_d2 = scal + 53;
  int* clip = iflag + 0;
  unsigned* ctr = (unsigned*)(iflag + 40);
  uint8_t* Ms = M + (size_t)a.adj * s2;
  const float* Es = E + (size_t)a.adj * s2;
  apply_op_masked(a.p, a.ap, a.max_scale, a.adj);
  hipLaunchKernelGGL(k_cg_alpha, dim3(gr), dim3(kT), 0, st_, (const float*)a.p, (const float*)a.ap, s2, partial,
                     ctr, (const double*)a.rr, d_pap, d_alpha);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cg_xn_flag, dim3(gr), dim3(kT), 0, st_, (const float*)a.x, (const float*)a.p,
                     (const double*)d_alpha, a.xn, s2, partial, ctr + 1, a.enforce, clip);
  PMS_LAUNCH_CHECK();
  if (a.enforce) {
    hipLaunchKernelGGL(k_cg_clip, dim3(gr), dim3(kT), 0, st_, a.xn, (const float*)a.x, a.p, (const double*)d_alpha,
                       s2, (const int*)clip);
    PMS_LAUNCH_CHECK();
    apply_op_masked(a.p, a.ap, a.max_scale, a.adj, clip);
  }
  hipLaunchKernelGGL(k_cg_rn_beta, dim3(gr), dim3(kT), 0, st_, (const float*)a.r, (const float*)a.ap,
                     (const double*)d_alpha, a.rn, s2, partial, ctr + 2, (const double*)a.rr, a.rnrn, d_beta);
  PMS_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cg_p, dim3(gr), dim3(kT), 0, st_, (const float*)a.rn, a.p, (const double*)d_beta, s2);
  PMS_LAUNCH_CHECK();
  // model sources and the SNR sums (main.py:413-424)
  fft_conv(a.xn, g.sub, g.p_sub, spec_sub, cbuf, fft_work, nullptr, st_);
  const long np = iuwt_decompose(cbuf, g.sub, g.sub, a.max_scale, a.adj, nullptr, nullptr, iuwt_work, Ms, Es,
                                 partial, nullptr, st_);
  sum_partials(partial, np, 1, d_d2, nullptr, st_);
}

// Run one iteration, replaying a captured HIP graph for this exact set of buffers and scales when
// enabled (the first iteration of each configuration records it).
void Job::run_iteration(const IterArgs& a, bool use_graph) {
  if (!use_graph) {
    enqueue_iteration(a);
    return;
  }
  auto it = graphs_.find(a);
  if (it == graphs_.end()) {
    PMS_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_iteration(a);
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
    it = graphs_.emplace(a, exec).first;
  }
  PMS_HIP(hipGraphLaunch(it->second, st_));
}

void Job::drop_graphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

// Ap <- R(M . W(H v)) over scales [adj, max_scale); masks are indexed by scale
void Job::apply_op_masked(const float* v, float* out, int max_scale, int adj, const int* pred) {
  const Geometry& g = geo_;
  fft_conv(v, g.sub, g.p_sub, spec_sub, cbuf, fft_work, pred, st_);
  iuwt_decompose(cbuf, g.sub, g.sub, max_scale, adj, Dcg, nullptr, iuwt_work, M + (size_t)adj * g.sub * g.sub,
                 nullptr, nullptr, pred, st_);
  iuwt_recompose(Dcg, max_scale - adj, g.sub, g.sub, adj, nullptr, out, iuwt_work + 0, nullptr, pred, st_);
}

// main.py:523-599
void Job::moresane_by_scale(const MoresaneParams& prm, int start_scale, int stop_scale, const float* dirty_dev,
                            const float* psf_dev, const float* mask_dev, float* model_dev, float* residual_dev,
                            Trace& trace) {
  const int n = prm.n;
  const size_t N2 = (size_t)n * n;
  // the dirty image of each call is the previous residual; keep a private copy so the caller's dirty stays intact
  float* cur_dirty = (float*)by_scale_arena_.alloc(N2 * sizeof(float));
  copy_plane(dirty_dev, cur_dirty, N2, nullptr, st_);
  int complete = 0;
  int scale_count = start_scale;
  while (!complete) {
    trace.push_back(ev(EV_BYSCALE, 0, scale_count, 0, 0, 0.0, 0.0));
    moresane(prm, scale_count, cur_dirty, psf_dev, mask_dev, model_dev, residual_dev, &complete, trace);
    copy_plane(residual_dev, cur_dirty, N2, nullptr, st_);
    ++scale_count;
    if (scale_count > std::log2((double)n) - 1) break;
    if (scale_count > stop_scale) break;
  }
  sync();
  by_scale_arena_.release();
}

}  // namespace pms
