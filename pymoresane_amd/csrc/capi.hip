// extern "C" entry points of libpmship.so (declared in include/pymoresane_hip.h).
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"

namespace pms {

static thread_local std::string t_err;

void set_error(const std::string& msg) { t_err = msg; }
const char* last_error() { return t_err.c_str(); }

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return PMS_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PMS_ERR_ALLOC;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PMS_ERR_INTERNAL;
  }
}

static hipStream_t S(void* s) { return (hipStream_t)s; }

// device-side partial sums for pms_snr_ratio
namespace {
__global__ __launch_bounds__(256) void k_diffsq(const float* __restrict__ a, const float* __restrict__ b, size_t n,
                                                double* __restrict__ partial) {
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double d = (double)a[i] - (double)b[i];
    acc = fma(d, d, acc);
  }
  acc = block_sum_f64(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}
}  // namespace

// one cached Job per stream (device buffers reused across calls of the same shape)
static std::mutex g_jobs_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<Job>> g_jobs;

static Job& job_for(hipStream_t st) {
  int dev = 0;
  PMS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_jobs_mu);
  auto key = std::make_pair(dev, st);
  auto it = g_jobs.find(key);
  if (it == g_jobs.end()) it = g_jobs.emplace(key, std::make_unique<Job>(st)).first;
  return *it->second;
}

static void copy_trace(const Trace& tr, pms_trace_event* out, int cap, int* len) {
  if (len) *len = (int)tr.size();
  if (!out) return;
  for (int i = 0; i < cap && i < (int)tr.size(); ++i) out[i] = tr[i];
}

static void check_params(const pms_moresane_params* prm) {
  PMS_REQUIRE(prm != nullptr, "params must not be NULL");
  PMS_REQUIRE(prm->conv_mode == 0 || prm->conv_mode == 1, "conv_mode must be 0 (linear) or 1 (circular)");
  PMS_REQUIRE(prm->major_loop_miter >= 0 && prm->minor_loop_miter >= 0, "iteration caps must be >= 0");
}

}  // namespace pms

using namespace pms;

extern "C" {

int pms_init(int device) {
  return guarded([&] {
    int n = 0;
    PMS_HIP(hipGetDeviceCount(&n));
    PMS_REQUIRE(device >= 0 && device < n, "device " + std::to_string(device) + " not present");
    PMS_HIP(hipSetDevice(device));
    fft_init_device(device);
  });
}

const char* pms_last_error(void) { return last_error(); }

int pms_get_device_count(int* count) {
  return guarded([&] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
  });
}

int pms_device_synchronize(void) { return guarded([&] { PMS_HIP(hipDeviceSynchronize()); }); }

int pms_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    PMS_REQUIRE(ptr != nullptr, "ptr must not be NULL");
    *ptr = nullptr;
    if (bytes == 0) return;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) fail(PMS_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  });
}

int pms_free(void* ptr) {
  return guarded([&] {
    if (ptr) PMS_HIP(hipFree(ptr));
  });
}

int pms_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    if (bytes == 0) return;
    PMS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    PMS_HIP(hipStreamSynchronize(S(stream)));
  });
}

int pms_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    if (bytes == 0) return;
    PMS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
    PMS_HIP(hipStreamSynchronize(S(stream)));
  });
}

int pms_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    if (bytes == 0) return;
    PMS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
  });
}

int pms_fill_f32(float* ptr, size_t count, float value, void* stream) {
  return guarded([&] {
    if (count) fill_f32(ptr, count, value, S(stream));
  });
}

int pms_copy_f32(float* dst, const float* src, size_t count, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(dst && src, "NULL buffer");
    if (count) copy_plane(src, dst, count, nullptr, S(stream));
  });
}

int pms_stream_create(void** stream) {
  return guarded([&] {
    hipStream_t s = nullptr;
    PMS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
  });
}

int pms_stream_destroy(void* stream) {
  return guarded([&] {
    {
      std::lock_guard<std::mutex> lk(g_jobs_mu);
      for (auto it = g_jobs.begin(); it != g_jobs.end();) {
        if (it->first.second == S(stream))
          it = g_jobs.erase(it);
        else
          ++it;
      }
    }
    if (stream) PMS_HIP(hipStreamDestroy(S(stream)));
  });
}

int pms_stream_synchronize(void* stream) { return guarded([&] { PMS_HIP(hipStreamSynchronize(S(stream))); }); }

int pms_event_create(void** event) {
  return guarded([&] {
    hipEvent_t e = nullptr;
    PMS_HIP(hipEventCreate(&e));
    *event = (void*)e;
  });
}

int pms_event_destroy(void* event) {
  return guarded([&] {
    if (event) PMS_HIP(hipEventDestroy((hipEvent_t)event));
  });
}

int pms_event_record(void* event, void* stream) {
  return guarded([&] { PMS_HIP(hipEventRecord((hipEvent_t)event, S(stream))); });
}

int pms_event_elapsed_ms(void* start, void* stop, float* ms) {
  return guarded([&] {
    PMS_HIP(hipEventSynchronize((hipEvent_t)stop));
    PMS_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  });
}

// ------------------------------------------------------------------------ IUWT
size_t pms_iuwt_work_bytes(int h, int w) { return 2 * (size_t)(h > 0 ? h : 0) * (size_t)(w > 0 ? w : 0) * sizeof(float); }

int pms_iuwt_decompose(const float* in, int h, int w, int scale_count, int scale_adjust, float* detail_out,
                       float* smoothed_out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(in && detail_out && work, "NULL buffer");
    iuwt_decompose(in, h, w, scale_count, scale_adjust, detail_out, smoothed_out, (float*)work, nullptr, nullptr,
                   nullptr, nullptr, S(stream));
  });
}

int pms_iuwt_recompose(const float* detail, int n_planes, int h, int w, int scale_adjust, const float* smoothed,
                       float* out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(detail && out && work, "NULL buffer");
    iuwt_recompose(detail, n_planes, h, w, scale_adjust, smoothed, out, (float*)work, nullptr, nullptr, S(stream));
  });
}

// ------------------------------------------------------------------------- FFT
size_t pms_fft_spectrum_bytes(int p) { return p > 0 ? fft_spectrum_elems(p) * sizeof(float2) : 0; }

size_t pms_fft_work_bytes(int n, int p) {
  if (n <= 0 || p <= 0) return 0;
  const size_t a = fft_conv_work_elems(n, p), b = fft_spectrum_elems(p);
  return (a > b ? a : b) * sizeof(float2);
}

int pms_psf_spectrum(const float* psf, int q, int p, void* spectrum_out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(psf && spectrum_out && work, "NULL buffer");
    fft_psf_spectrum(psf, q, p, (float2*)spectrum_out, (float2*)work, S(stream));
  });
}

int pms_spectrum_from_rfft2(const float* rfft2_host, int p, void* spectrum_host_out) {
  return guarded([&] {
    PMS_REQUIRE(rfft2_host && spectrum_host_out, "NULL buffer");
    fft_spectrum_from_numpy_host(rfft2_host, p, (float2*)spectrum_host_out);
  });
}

int pms_fft_convolve(const float* in, int n, int p, const void* spectrum, float* out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(in && spectrum && out && work, "NULL buffer");
    fft_conv(in, n, p, (const float2*)spectrum, out, (float2*)work, nullptr, S(stream));
  });
}

// ------------------------------------------------------------------- threshold
size_t pms_threshold_work_bytes(int n_planes, int h, int w) {
  if (n_planes <= 0 || h <= 0 || w <= 0) return 0;
  (void)h;
  (void)w;
  return ((mad_work_bytes(n_planes) + 255) / 256) * 256 + (size_t)n_planes * 256 * sizeof(float);
}

int pms_estimate_threshold(const float* planes, int n_planes, int h, int w, int edge_excl, int int_excl,
                           double* thresholds_out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(planes && thresholds_out && work, "NULL buffer");
    PMS_REQUIRE(n_planes >= 1 && h >= 1 && w >= 1, "bad plane stack shape");
    NoiseMask nm = noise_mask(h, w, edge_excl, int_excl);
    mad_work_init(work, n_planes, S(stream));
    mad_thresholds(planes, n_planes, h, w, nm, thresholds_out, work, nullptr, S(stream));
  });
}

int pms_apply_threshold(const float* planes, int n_planes, int h, int w, const double* thresholds,
                        double sigma_level, float* out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(planes && thresholds && out && work, "NULL buffer");
    float* pmax = (float*)((char*)work + ((mad_work_bytes(n_planes) + 255) / 256) * 256);
    apply_threshold(planes, out, n_planes, h, w, thresholds, sigma_level, nullptr, pmax, nullptr, S(stream));
  });
}

// ------------------------------------------------------------------ extraction
size_t pms_extract_work_bytes(int n_planes, int h, int w) {
  if (n_planes <= 0 || h <= 0 || w <= 0) return 0;
  return extract_work_bytes(n_planes, h, w);
}

int pms_source_extraction(const float* planes, int n_planes, int h, int w, double tolerance, int neg_comp,
                          float* extracted_out, unsigned char* mask_out, int* labels_out, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(planes && extracted_out && mask_out && work, "NULL buffer");
    PMS_REQUIRE(n_planes >= 1 && h >= 1 && w >= 1, "bad plane stack shape");
    extract_work_init(work, n_planes, h, w, S(stream));
    source_extract(planes, n_planes, h, w, tolerance, neg_comp, mask_out, extracted_out, labels_out, work, nullptr,
                   S(stream));
  });
}

size_t pms_snr_work_bytes(size_t n) { return (2 * (size_t)reduce_grid(n) + 4) * sizeof(double); }

int pms_snr_ratio(const float* a, const float* b, size_t n, double* snr_host, void* work, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(a && b && snr_host && work, "NULL buffer");
    PMS_REQUIRE(n > 0, "empty arrays");
    const int g = reduce_grid(n);
    double* part = (double*)work;
    double* out = part + 2 * g;
    dot_partials(a, nullptr, n, part, nullptr, S(stream));
    hipLaunchKernelGGL(k_diffsq, dim3(g), dim3(256), 0, S(stream), a, b, n, part + g);
    PMS_LAUNCH_CHECK();
    sum_partials(part, g, 2, out, nullptr, S(stream));
    double v[2];
    PMS_HIP(hipMemcpyAsync(v, out, sizeof(v), hipMemcpyDeviceToHost, S(stream)));
    PMS_HIP(hipStreamSynchronize(S(stream)));
    *snr_host = 20.0 * std::log10(std::sqrt(v[0]) / std::sqrt(v[1]));
  });
}

// -------------------------------------------------------------------- MORESANE
int pms_moresane(const pms_moresane_params* prm, const float* dirty, const float* psf, const float* mask,
                 float* model_inout, float* residual_inout, int* complete_out, pms_trace_event* trace, int trace_cap,
                 int* trace_len, void* stream) {
  return guarded([&] {
    check_params(prm);
    PMS_REQUIRE(dirty && psf && model_inout && residual_inout, "NULL buffer");
    Trace tr;
    int complete = 0;
    Job& job = job_for(S(stream));
    try {
      job.moresane(*prm, prm->scale_count, dirty, psf, mask, model_inout, residual_inout, &complete, tr);
    } catch (...) {
      copy_trace(tr, trace, trace_cap, trace_len);
      throw;
    }
    if (complete_out) *complete_out = complete;
    copy_trace(tr, trace, trace_cap, trace_len);
  });
}

int pms_moresane_by_scale(const pms_moresane_params* prm, int start_scale, int stop_scale, const float* dirty,
                          const float* psf, const float* mask, float* model_inout, float* residual_inout,
                          pms_trace_event* trace, int trace_cap, int* trace_len, void* stream) {
  return guarded([&] {
    check_params(prm);
    PMS_REQUIRE(dirty && psf && model_inout && residual_inout, "NULL buffer");
    Trace tr;
    Job& job = job_for(S(stream));
    try {
      job.moresane_by_scale(*prm, start_scale, stop_scale, dirty, psf, mask, model_inout, residual_inout, tr);
    } catch (...) {
      copy_trace(tr, trace, trace_cap, trace_len);
      throw;
    }
    copy_trace(tr, trace, trace_cap, trace_len);
  });
}

// ---------------------------------------------------------------------- batch
int pms_batch_run(pms_batch_job* jobs, int n_jobs, int n_workers) {
  return guarded([&] {
    PMS_REQUIRE(n_jobs >= 0 && (n_jobs == 0 || jobs), "bad job list");
    if (n_jobs == 0) return;
    if (n_workers < 1) n_workers = 1;
    if (n_workers > n_jobs) n_workers = n_jobs;
    int dev = 0;
    PMS_HIP(hipGetDevice(&dev));
    std::atomic<int> next{0};
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    std::vector<std::string> errs(n_workers);
    for (int wk = 0; wk < n_workers; ++wk) {
      pool.emplace_back([&, wk] {
        hipStream_t st = nullptr;
        if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
          errs[wk] = "worker could not create a stream";
          return;
        }
        {
          Job job(st);
          for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n_jobs) break;
            pms_batch_job& J = jobs[i];
            const auto a = std::chrono::steady_clock::now();
            J.t_start = std::chrono::duration<double>(a - t0).count();
            Trace tr;
            job.reset_wait();
            J.status = guarded([&] {
              check_params(&J.prm);
              if (J.by_scale) {
                job.moresane_by_scale(J.prm, J.start_scale, J.stop_scale, J.dirty, J.psf, J.mask, J.model, J.residual,
                                      tr);
                J.complete = 0;
              } else {
                int complete = 0;
                J.major_iters = job.moresane(J.prm, J.prm.scale_count, J.dirty, J.psf, J.mask, J.model, J.residual,
                                             &complete, tr);
                J.complete = complete;
              }
            });
            J.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
            J.wait_seconds = job.wait_seconds();
          }
        }
        hipStreamDestroy(st);
      });
    }
    for (auto& t : pool) t.join();
    for (auto& e : errs)
      if (!e.empty()) fail(PMS_ERR_HIP, e);
  });
}

}  // extern "C"
