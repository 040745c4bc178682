// extern "C" entry points of libpmship.so (declared in include/pymoresane_hip.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

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

// one cached single-slot runner per (device, stream): device buffers reused across calls of the same
// shape.  A NULL stream gets a runner-owned stream (graph capture needs a stream of its own).
static std::mutex g_runners_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<Runner>> g_runners;

static Runner& runner_for(hipStream_t st) {
  int dev = 0;
  PMS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_runners_mu);
  auto key = std::make_pair(dev, st);
  auto it = g_runners.find(key);
  if (it == g_runners.end()) it = g_runners.emplace(key, std::make_unique<Runner>(1, st)).first;
  return *it->second;
}

// Run one task to completion on a single-slot runner; the task's error (if any) is rethrown after
// its trace has been copied out.
static void run_single(Runner& rn, const TaskSpec& spec, Trace& tr, int* complete, int* major) {
  rn.refresh_options();
  const int slot = rn.add(spec);
  PMS_REQUIRE(slot >= 0, "runner busy");
  try {
    rn.run();
  } catch (...) {
    tr = rn.task(slot)->trace;
    rn.release(slot);
    throw;
  }
  Task* t = rn.task(slot);
  tr = t->trace;
  const int status = t->status;
  const std::string err = t->error;
  if (complete) *complete = t->complete;
  if (major) *major = t->major_iters;
  rn.release(slot);
  if (status != 0) fail(status, err);
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
    (void)noalloc_mad_scratch();  // this device's address of the allocation-free median workspace
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
      std::lock_guard<std::mutex> lk(g_runners_mu);
      for (auto it = g_runners.begin(); it != g_runners.end();) {
        if (it->first.second == S(stream))
          it = g_runners.erase(it);
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

int pms_estimate_threshold_noalloc(const float* planes, int n_planes, int h, int w, int edge_excl, int int_excl,
                                   double* thresholds_out, void* stream) {
  return guarded([&] {
    PMS_REQUIRE(planes && thresholds_out, "NULL buffer");
    PMS_REQUIRE(h >= 1 && w >= 1, "bad plane stack shape");
    PMS_REQUIRE(n_planes >= 1 && n_planes <= PMS_NOALLOC_MAX_PLANES,
                "pms_estimate_threshold_noalloc takes 1.." + std::to_string(PMS_NOALLOC_MAX_PLANES) +
                    " planes per call (its fixed device workspace holds that many), got " +
                    std::to_string(n_planes));
    NoiseMask nm = noise_mask(h, w, edge_excl, int_excl);
    PMS_REQUIRE(nm.count < (1L << 32), "pms_estimate_threshold_noalloc: more than 2^32 - 1 pixels per plane");
    void* work = noalloc_mad_scratch();
    PMS_REQUIRE(mad_work_bytes(n_planes) <= noalloc_mad_bytes(), "median workspace larger than the device array");
    // the array keeps the previous call's contents: clear the histograms (the select state is
    // rewritten by the first kernel of every call)
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
    const TaskSpec spec{*prm, dirty, psf, mask, model_inout, residual_inout, 0, 0, 0};
    try {
      run_single(runner_for(S(stream)), spec, tr, &complete, nullptr);
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
    const TaskSpec spec{*prm, dirty, psf, mask, model_inout, residual_inout, 1, start_scale, stop_scale};
    try {
      run_single(runner_for(S(stream)), spec, tr, nullptr, nullptr);
    } catch (...) {
      copy_trace(tr, trace, trace_cap, trace_len);
      throw;
    }
    copy_trace(tr, trace, trace_cap, trace_len);
  });
}

// ---------------------------------------------------------------------- batch
int pms_batch_run(pms_batch_job* jobs, int n_jobs, int n_workers, int max_batch) {
  return guarded([&] {
    PMS_REQUIRE(n_jobs >= 0 && (n_jobs == 0 || jobs), "bad job list");
    PMS_REQUIRE(max_batch >= 1 && max_batch <= 1024, "max_batch must be in 1..1024");
    if (n_jobs == 0) return;
    if (n_workers < 1) n_workers = 1;
    // size classes, largest total work first; jobs keep their order inside a class
    std::map<std::tuple<int, int, int, int>, std::vector<int>> by_class;
    for (int i = 0; i < n_jobs; ++i) {
      const pms_moresane_params& p = jobs[i].prm;
      const int sub = (p.subregion <= 0 || p.subregion > p.n) ? p.n : p.subregion;
      by_class[std::make_tuple(p.n, p.psf_n, sub, p.conv_mode)].push_back(i);
    }
    // With more workers than classes each class is dealt round-robin into several groups, so two or
    // more runners (streams) step the same class and the GPU has queued work while a runner's host
    // thread turns its step around.
    const int n_classes = (int)by_class.size();
    const int split = std::max(1, n_workers / std::max(1, n_classes));
    std::vector<std::vector<int>> classes;
    for (auto& kv : by_class) {
      const int parts = std::min<int>(split, (int)kv.second.size());
      for (int q = 0; q < parts; ++q) {
        std::vector<int> part;
        for (size_t i = q; i < kv.second.size(); i += parts) part.push_back(kv.second[i]);
        classes.push_back(part);
      }
    }
    std::stable_sort(classes.begin(), classes.end(), [&](const std::vector<int>& a, const std::vector<int>& b) {
      const double wa = (double)jobs[a[0]].prm.n * jobs[a[0]].prm.n * a.size();
      const double wb = (double)jobs[b[0]].prm.n * jobs[b[0]].prm.n * b.size();
      return wa > wb;
    });
    if (n_workers > (int)classes.size()) n_workers = (int)classes.size();
    int dev = 0;
    PMS_HIP(hipGetDevice(&dev));
    // the runners' streams are non-blocking: work the caller queued on other streams (the null stream
    // included) for these buffers must be complete before the first job step reads them
    PMS_HIP(hipDeviceSynchronize());
    std::atomic<int> next_class{0};
    const auto t0 = std::chrono::steady_clock::now();
    auto since = [&](std::chrono::steady_clock::time_point t) {
      return std::chrono::duration<double>(t - t0).count();
    };
    std::vector<std::thread> pool;
    std::vector<std::string> errs(n_workers);
    for (int wk = 0; wk < n_workers; ++wk) {
      pool.emplace_back([&, wk] {
        try {
          PMS_HIP(hipSetDevice(dev));
          Runner rn(max_batch);
          std::vector<int> slot_job(max_batch, -1);
          std::vector<double> slot_wait0(max_batch, 0.0);
          for (;;) {
            const int c = next_class.fetch_add(1);
            if (c >= (int)classes.size()) break;
            const std::vector<int>& q = classes[c];
            size_t qi = 0;
            for (;;) {
              // refill free slots from this class, then step
              for (int k = rn.first_free(); k >= 0 && qi < q.size(); k = rn.first_free()) {
                pms_batch_job& J = jobs[q[qi++]];
                J.t_start = since(std::chrono::steady_clock::now());
                J.status = guarded([&] {
                  check_params(&J.prm);
                  const TaskSpec spec{J.prm,   J.dirty, J.psf, J.mask, J.model, J.residual, J.by_scale,
                                      J.start_scale, J.stop_scale};
                  rn.add(spec, k);
                });
                if (J.status != 0) {
                  J.seconds = 0.0;
                  continue;
                }
                slot_job[k] = (int)(&J - jobs);
                slot_wait0[k] = rn.wait_seconds();
              }
              if (rn.idle()) break;
              rn.step();
              const double now = since(std::chrono::steady_clock::now());
              for (int k = 0; k < max_batch; ++k) {
                Task* t = rn.task(k);
                if (!t || !t->done()) continue;
                pms_batch_job& J = jobs[slot_job[k]];
                J.status = t->status;
                if (t->status != 0) set_error(t->error);
                J.major_iters = t->major_iters;
                J.complete = J.by_scale ? 0 : t->complete;
                J.seconds = now - J.t_start;
                J.wait_seconds = rn.wait_seconds() - slot_wait0[k];
                rn.release(k);
                slot_job[k] = -1;
              }
            }
          }
        } catch (const Error& e) {
          errs[wk] = e.msg;
        } catch (const std::exception& e) {
          errs[wk] = e.what();
        }
      });
    }
    for (auto& t : pool) t.join();
    for (auto& e : errs)
      if (!e.empty()) fail(PMS_ERR_HIP, e);
  });
}

}  // extern "C"
