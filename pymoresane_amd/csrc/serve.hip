// Continuous-batching deconvolution server.
//
// Requests (host-resident dirty/PSF, host outputs) join the FIFO queue of their
// size class (image side, PSF side, subregion, convolution mode).  Each class
// has a host thread with its own HIP stream and a Runner of max_batch slots
// that steps its admitted requests in lockstep.  Between steps the thread
// admits queued requests while a slot is free and the device footprints of all
// running requests (every class) plus the new one fit the HBM budget, so a
// request that arrives while others run joins the running batch at the next
// step, and classes overlap on the device through their streams.  Every
// request's result is the same as running it alone: each job's arithmetic is
// independent of the others in its batch and no state is shared across
// streams.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

#include "engine.h"

namespace pms {

class Server {
 public:
  Server(int max_batch, size_t budget)
      : max_batch_(max_batch), budget_(budget), t0_(std::chrono::steady_clock::now()) {
    PMS_HIP(hipGetDevice(&dev_));
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& kv : lanes_) kv.second->th.join();
  }

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }

  void submit(pms_serve_request* r) {
    Geometry g = make_geometry(r->prm.n, r->prm.psf_n, r->prm, r->prm.scale_count, nullptr);
    const size_t io = (size_t)(3 * (size_t)r->prm.n * r->prm.n + (size_t)r->prm.psf_n * r->prm.psf_n) * sizeof(float) +
                      (r->mask_host ? (size_t)r->prm.n * r->prm.n * sizeof(float) : 0);
    r->footprint = job_footprint_bytes(g) + io;
    PMS_REQUIRE(r->footprint <= budget_, "request footprint exceeds the server's HBM budget");
    // max_batch 1 is the plain path: one lane, one request at a time
    const Key key = max_batch_ == 1 ? Key{0, 0, 0, 0} : std::make_tuple(g.n, g.psf_n, g.sub, r->prm.conv_mode);
    r->status = 1;  // queued
    r->t_submit = now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = lanes_.find(key);
      if (it == lanes_.end()) {
        it = lanes_.emplace(key, std::make_unique<Lane>()).first;
        Lane* lane = it->second.get();
        lane->free_slots = max_batch_;  // its runner starts with every slot free
        lane->th = std::thread([this, lane] { lane_loop(*lane); });
      }
      it->second->queue.emplace_back(next_seq_++, r);
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

 private:
  using Key = std::tuple<int, int, int, int>;

  // device copies of one request's inputs and outputs (per slot, grow-only)
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

  struct Lane {
    std::thread th;
    std::deque<std::pair<long, pms_serve_request*>> queue;  // (submission number, request)
    int free_slots = 0;
  };

  // The head of a lane is admitted when its lane has a free slot and its footprint fits the budget.
  // Budget is granted in submission order: an older request that has a slot but waits for budget
  // in another lane holds back younger ones, so large requests are not starved by small ones.
  bool admissible(const Lane& lane, int free_slots) const {
    if (free_slots <= 0 || lane.queue.empty()) return false;
    const long seq = lane.queue.front().first;
    for (auto& kv : lanes_) {
      const Lane& o = *kv.second;
      if (&o == &lane || o.queue.empty() || o.free_slots <= 0) continue;
      if (o.queue.front().first < seq) return false;
    }
    return used_ == 0 || used_ + lane.queue.front().second->footprint <= budget_;
  }

  void lane_loop(Lane& lane) {
    std::unique_ptr<Runner> rn;
    const int status = run_guarded([&] {
      PMS_HIP(hipSetDevice(dev_));
      rn = std::make_unique<Runner>(max_batch_);
    });
    if (status != 0) {
      fail_lane(lane, status);
      return;
    }
    const hipStream_t st = rn->stream();
    std::vector<IoBuffers> io(max_batch_);
    std::vector<pms_serve_request*> slot_req(max_batch_, nullptr);
    for (;;) {
      std::vector<pms_serve_request*> admitted;
      {
        std::unique_lock<std::mutex> lk(mu_);
        lane.free_slots = rn->free_slots();
        cv_.wait(lk, [&] { return stop_ || !rn->idle() || admissible(lane, lane.free_slots); });
        if (stop_ && rn->idle() && lane.queue.empty()) break;
        while (admissible(lane, lane.free_slots)) {
          pms_serve_request* r = lane.queue.front().second;
          lane.queue.pop_front();
          used_ += r->footprint;
          r->status = 2;  // running
          admitted.push_back(r);
          --lane.free_slots;
        }
      }
      if (!admitted.empty()) cv_.notify_all();  // younger requests held back in other lanes re-check
      for (pms_serve_request* r : admitted) {
        r->t_start = now();
        const int k = rn->first_free();
        const int n = r->prm.n, pn = r->prm.psf_n;
        const size_t n2 = (size_t)n * n, p2 = (size_t)pn * pn;
        const int status = run_guarded([&] {
          IoBuffers& b = io[k];
          b.ensure(n2, p2, r->mask_host != nullptr);
          PMS_HIP(hipMemcpyAsync(b.dirty, r->dirty_host, n2 * sizeof(float), hipMemcpyHostToDevice, st));
          PMS_HIP(hipMemcpyAsync(b.psf, r->psf_host, p2 * sizeof(float), hipMemcpyHostToDevice, st));
          if (r->mask_host)
            PMS_HIP(hipMemcpyAsync(b.mask_d, r->mask_host, n2 * sizeof(float), hipMemcpyHostToDevice, st));
          fill_f32(b.model, n2, 0.f, st);
          PMS_HIP(hipMemcpyAsync(b.resid, b.dirty, n2 * sizeof(float), hipMemcpyDeviceToDevice, st));
          const TaskSpec spec{r->prm,  b.dirty, b.psf, r->mask_host ? b.mask_d : nullptr, b.model, b.resid,
                              r->by_scale, r->start_scale, r->stop_scale};
          rn->add(spec, k);
          slot_req[k] = r;
        });
        if (status != 0) finish(r, status);
      }
      if (rn->idle()) continue;
      int step_status = run_guarded([&] { rn->step(); });
      for (int k = 0; k < max_batch_; ++k) {
        Task* t = rn->task(k);
        if (!t) continue;
        if (step_status == 0 && !t->done()) continue;
        pms_serve_request* r = slot_req[k];
        int status = step_status != 0 ? step_status : t->status;
        if (status == 0) {
          status = run_guarded([&] {
            const size_t n2 = (size_t)r->prm.n * r->prm.n;
            PMS_HIP(hipMemcpyAsync(r->model_host, io[k].model, n2 * sizeof(float), hipMemcpyDeviceToHost, st));
            PMS_HIP(hipMemcpyAsync(r->residual_host, io[k].resid, n2 * sizeof(float), hipMemcpyDeviceToHost, st));
            PMS_HIP(hipStreamSynchronize(st));
          });
        } else if (step_status == 0) {
          set_error(t->error);
        }
        r->major_iters = t->major_iters;
        rn->release(k);
        slot_req[k] = nullptr;
        finish(r, status);
      }
    }
  }

  void finish(pms_serve_request* r, int status, bool admitted = true) {
    r->t_done = now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (admitted) used_ -= r->footprint;
      r->status = status;  // 0 = done, < 0 = error
      --pending_;
    }
    cv_.notify_all();
    done_cv_.notify_all();
  }

  // A lane whose thread could not set up its device stream and runner fails every request queued
  // to it (now and until the server stops) with that status, instead of leaving them pending.
  void fail_lane(Lane& lane, int status) {
    for (;;) {
      std::vector<pms_serve_request*> dead;
      {
        std::unique_lock<std::mutex> lk(mu_);
        lane.free_slots = 0;  // never holds back the requests of other lanes
        cv_.wait(lk, [&] { return stop_ || !lane.queue.empty(); });
        if (lane.queue.empty()) break;  // stopping
        for (auto& q : lane.queue) dead.push_back(q.second);
        lane.queue.clear();
      }
      cv_.notify_all();
      for (pms_serve_request* r : dead) {
        r->t_start = now();
        finish(r, status, false);
      }
    }
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
  int max_batch_;
  size_t budget_;
  size_t used_ = 0;
  int pending_ = 0;
  long next_seq_ = 0;
  bool stop_ = false;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::map<Key, std::unique_ptr<Lane>> lanes_;
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

int pms_server_create(int max_batch, double budget_bytes, void** server) {
  return serve_guarded([&] {
    PMS_REQUIRE(server != nullptr, "server must not be NULL");
    PMS_REQUIRE(max_batch >= 1 && max_batch <= 1024, "max_batch must be in 1..1024");
    PMS_REQUIRE(budget_bytes > 0, "budget must be positive");
    *server = new Server(max_batch, (size_t)budget_bytes);
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
