// Continuous-batching deconvolution server.
//
// Requests (host-resident dirty/PSF, host outputs) enter a FIFO queue.  A pool
// of worker threads, each owning a HIP stream and a grow-only device job
// workspace, admits the request at the head of the queue as soon as a worker
// is free and the sum of the device footprints of running requests plus the
// new one fits the HBM budget.  Admitted requests run concurrently on their
// own streams, so work that arrives while others are on the GPU overlaps with
// them.  Every request's result is the same as running it alone: the kernels
// are deterministic and share no state across streams.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "engine.h"

namespace pms {

class Server {
 public:
  Server(int n_workers, size_t budget) : budget_(budget), t0_(std::chrono::steady_clock::now()) {
    PMS_HIP(hipGetDevice(&dev_));
    for (int i = 0; i < n_workers; ++i) workers_.emplace_back([this] { worker(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }

  void submit(pms_serve_request* r) {
    Geometry g = make_geometry(r->prm.n, r->prm.psf_n, r->prm, r->prm.scale_count, nullptr);
    const size_t io = (size_t)(3 * (size_t)r->prm.n * r->prm.n + (size_t)r->prm.psf_n * r->prm.psf_n) * sizeof(float) +
                      (r->mask_host ? (size_t)r->prm.n * r->prm.n * sizeof(float) : 0);
    r->footprint = job_footprint_bytes(g) + io;
    PMS_REQUIRE(r->footprint <= budget_, "request footprint exceeds the server's HBM budget");
    r->status = 1;  // queued
    r->t_submit = now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      queue_.push_back(r);
      ++pending_;
    }
    cv_.notify_all();
  }

  void drain() {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
  }

  bool wait(pms_serve_request* r, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    return done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), [r] { return r->status <= 0; });
  }

  size_t used() {
    std::lock_guard<std::mutex> lk(mu_);
    return used_;
  }

 private:
  struct IoBuffers {
    DeviceArena arena;
    size_t n2 = 0, p2 = 0;
    bool mask = false;
    float *dirty = nullptr, *psf = nullptr, *mask_d = nullptr, *model = nullptr, *resid = nullptr;
    void ensure(size_t need_n2, size_t need_p2, bool need_mask) {
      if (need_n2 <= n2 && need_p2 <= p2 && (!need_mask || mask)) return;
      arena.release();
      n2 = std::max(n2, need_n2);
      p2 = std::max(p2, need_p2);
      mask = mask || need_mask;
      dirty = (float*)arena.alloc(n2 * sizeof(float));
      model = (float*)arena.alloc(n2 * sizeof(float));
      resid = (float*)arena.alloc(n2 * sizeof(float));
      psf = (float*)arena.alloc(p2 * sizeof(float));
      mask_d = mask ? (float*)arena.alloc(n2 * sizeof(float)) : nullptr;
    }
  };

  void worker() {
    if (hipSetDevice(dev_) != hipSuccess) return;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
    {
      Job job(st);
      IoBuffers io;
      for (;;) {
        pms_serve_request* r = nullptr;
        {
          std::unique_lock<std::mutex> lk(mu_);
          cv_.wait(lk, [this] {
            return stop_ || (!queue_.empty() && (used_ == 0 || used_ + queue_.front()->footprint <= budget_));
          });
          if (stop_ && queue_.empty()) break;
          if (queue_.empty()) continue;
          r = queue_.front();
          queue_.pop_front();
          used_ += r->footprint;
          r->status = 2;  // running
        }
        r->t_start = now();
        const int n = r->prm.n, pn = r->prm.psf_n;
        const size_t n2 = (size_t)n * n, p2 = (size_t)pn * pn;
        Trace tr;
        int complete = 0;
        const int status = run_guarded([&] {
          io.ensure(n2, p2, r->mask_host != nullptr);
          PMS_HIP(hipMemcpyAsync(io.dirty, r->dirty_host, n2 * sizeof(float), hipMemcpyHostToDevice, st));
          PMS_HIP(hipMemcpyAsync(io.psf, r->psf_host, p2 * sizeof(float), hipMemcpyHostToDevice, st));
          if (r->mask_host)
            PMS_HIP(hipMemcpyAsync(io.mask_d, r->mask_host, n2 * sizeof(float), hipMemcpyHostToDevice, st));
          PMS_HIP(hipMemsetAsync(io.model, 0, n2 * sizeof(float), st));
          PMS_HIP(hipMemcpyAsync(io.resid, io.dirty, n2 * sizeof(float), hipMemcpyDeviceToDevice, st));
          if (r->by_scale) {
            job.moresane_by_scale(r->prm, r->start_scale, r->stop_scale, io.dirty, io.psf,
                                  r->mask_host ? io.mask_d : nullptr, io.model, io.resid, tr);
          } else {
            r->major_iters = job.moresane(r->prm, r->prm.scale_count, io.dirty, io.psf,
                                          r->mask_host ? io.mask_d : nullptr, io.model, io.resid, &complete, tr);
          }
          PMS_HIP(hipMemcpyAsync(r->model_host, io.model, n2 * sizeof(float), hipMemcpyDeviceToHost, st));
          PMS_HIP(hipMemcpyAsync(r->residual_host, io.resid, n2 * sizeof(float), hipMemcpyDeviceToHost, st));
          PMS_HIP(hipStreamSynchronize(st));
        });
        r->t_done = now();
        {
          std::lock_guard<std::mutex> lk(mu_);
          used_ -= r->footprint;
          r->status = status;  // 0 = done, < 0 = error
          --pending_;
        }
        cv_.notify_all();
        done_cv_.notify_all();
      }
    }
    (void)hipStreamDestroy(st);
  }

  template <class F>
  static int run_guarded(F&& f) {
    try {
      f();
      return PMS_OK;
    } catch (const Error& e) {
      set_error(e.msg);
      return e.code;
    } catch (const std::exception& e) {
      set_error(e.what());
      return PMS_ERR_INTERNAL;
    }
  }

  int dev_ = 0;
  size_t budget_;
  size_t used_ = 0;
  int pending_ = 0;
  bool stop_ = false;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<pms_serve_request*> queue_;
  std::vector<std::thread> workers_;
};

}  // namespace pms

using namespace pms;

template <class F>
static int serve_guarded(F&& f) {
  try {
    f();
    return PMS_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PMS_ERR_INTERNAL;
  }
}

extern "C" {

int pms_server_create(int n_workers, double budget_bytes, void** server) {
  return serve_guarded([&] {
    PMS_REQUIRE(server != nullptr, "server must not be NULL");
    PMS_REQUIRE(n_workers >= 1 && n_workers <= 64, "n_workers must be in 1..64");
    PMS_REQUIRE(budget_bytes > 0, "budget must be positive");
    *server = new Server(n_workers, (size_t)budget_bytes);
  });
}

int pms_server_submit(void* server, pms_serve_request* req) {
  return serve_guarded([&] {
    PMS_REQUIRE(server && req, "NULL argument");
    PMS_REQUIRE(req->dirty_host && req->psf_host && req->model_host && req->residual_host, "NULL buffer");
    ((Server*)server)->submit(req);
  });
}

int pms_server_wait(void* server, pms_serve_request* req, double timeout_s, int* done) {
  return serve_guarded([&] {
    PMS_REQUIRE(server && req && done, "NULL argument");
    *done = ((Server*)server)->wait(req, timeout_s) ? 1 : 0;
  });
}

int pms_server_drain(void* server) {
  return serve_guarded([&] {
    PMS_REQUIRE(server, "NULL server");
    ((Server*)server)->drain();
  });
}

int pms_server_now(void* server, double* t) {
  return serve_guarded([&] {
    PMS_REQUIRE(server && t, "NULL argument");
    *t = ((Server*)server)->now();
  });
}

int pms_server_destroy(void* server) {
  return serve_guarded([&] { delete (Server*)server; });
}

}  // extern "C"
