// Device-resident MORESANE driver: per-job workspace, the major/minor loops as a
// resumable state machine, and a runner that steps many of them in lockstep.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pymoresane_hip.h"
#include "pms_internal.h"

namespace pms {

using MoresaneParams = pms_moresane_params;

enum TraceKind {
  EV_SCALES = PMS_EV_SCALES,
  EV_MINOR = PMS_EV_MINOR,
  EV_UPDATE = PMS_EV_UPDATE,
  EV_NOCOEFF = PMS_EV_NOCOEFF,
  EV_POOR = PMS_EV_POOR,
  EV_NOWORK = PMS_EV_NOWORK,
  EV_BYSCALE = PMS_EV_BYSCALE,
};

// minor-loop exit codes (same numbering as the oracle's decision trace)
enum ExitCode { EXIT_CAP = 0, EXIT_FIRST_HIGH = 1, EXIT_ACCEPT = 2, EXIT_FELL_OK = 3, EXIT_FELL_LOW = 4 };

using Trace = std::vector<pms_trace_event>;

class DeviceArena {
 public:
  ~DeviceArena();
  void* alloc(size_t bytes);
  void release();
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_ = 0;
};

struct Geometry {
  int n, psf_n, sub, S, lo;
  bool linear, double_psf;
  int p_sub, p_full;
  bool f64;  // float64 bulk arithmetic (the precision mode) instead of float32
};

// Per-element-type pieces of the driver: spectrum element, FFT convolution job.
template <class T>
struct EngineTypes;
template <>
struct EngineTypes<float> {
  using C = float2;
  using Conv = FftConvJob;
};
template <>
struct EngineTypes<double> {
  using C = Cx<double>;
  using Conv = GConvJob<double>;
};

Geometry make_geometry(int n, int psf_n, const MoresaneParams& prm, int scale_count_in, int* subregion_out);

// byte sizes of the per-job device buffers
struct JobSizes {
  size_t s2, S_s2, mask, iuwt, N2, spec_sub, spec_full, fft, psf, ext, partial, pmax;
  bool f64;
};

// device bytes one job of this geometry needs (its workspace plus nothing else)
size_t job_footprint_bytes(const Geometry& g);

// Minor-loop iterations a one-slot runner queues per host step unless PMS_CG_ITERS says otherwise.
// A runner with more slots queues one: its batched launches and the other streams' work cover the
// host round trip, while iterations queued past a job's exit still cost their (predicated) launches.
constexpr int kDefaultCgIters = 4;

// Per-job scalar slot (doubles) in the runner's scalar block: the only values read back.
constexpr int kSlotDoubles = 64;
constexpr int kSlotInts = 64;

// The element-typed device buffers of one job (sub = subregion side, S = scale capacity).
template <class T>
struct Planes {
  using C = typename EngineTypes<T>::C;
  T *dirty_sub, *psf_sub, *D, *Dcg, *E;
  T* Rs;  // E - M.W(H x) of the running minor loop (incremental SNR)
  T *rec, *x, *xn, *r, *rn, *p, *ap, *cbuf;
  T* iuwt_work;
  T *model_local, *resid, *conv_full, *psf_frame;
  T* pmax;
  T* cur_dirty;  // moresane_by_scale: the dirty image of the current call (a private copy)
  C *spec_sub, *spec_full, *spec_full_buf, *fft_work;
};

// Device buffers of one deconvolution, kept across calls of the same or smaller shape and the same
// precision.  Only the planes of the job's element type are allocated (the other set stays null).
class Job {
 public:
  Job() = default;
  ~Job();
  Job(const Job&) = delete;
  Job& operator=(const Job&) = delete;
  // grow-only allocation for geometry g; returns true when buffers moved
  bool ensure(const Geometry& g, hipStream_t st);
  template <class T>
  T* by_scale_dirty(size_t n2);  // private copy of the dirty image for moresane_by_scale
  size_t device_bytes() const { return arena_.bytes() + by_scale_arena_.bytes(); }
  const Geometry& geometry() const { return geo_; }
  template <class T>
  Planes<T>& pl();

 private:
  DeviceArena arena_;
  DeviceArena by_scale_arena_;
  size_t by_scale_bytes_ = 0;
  void* by_scale_ptr_ = nullptr;
  Geometry geo_{};
  JobSizes cap_{};
  bool have_ = false;

 public:
  Planes<float> f{};
  Planes<double> d{};
  uint8_t* M = nullptr;
  void *mad_work = nullptr, *ext_work = nullptr;
  double* partial = nullptr;
};

template <>
inline Planes<float>& Job::pl<float>() {
  return f;
}
template <>
inline Planes<double>& Job::pl<double>() {
  return d;
}

// One request: moresane (main.py:53-521) or moresane_by_scale (main.py:523-599) of one image.
struct TaskSpec {
  MoresaneParams prm;
  const float* dirty;
  const float* psf;
  const float* mask;  // may be null
  float* model;       // accumulated into
  float* residual;    // overwritten when a major iteration ran
  int by_scale, start_scale, stop_scale;
};

class Runner;

// Work of one step that the runner issues as batched launches across tasks (between every task's
// pre and post work of that step), keyed by plane side or (image side, FFT frame), per element type.
template <class T>
struct SharedOpsT {
  std::vector<std::pair<int, IuwtDecompJobT<T>>> decomp;
  std::vector<std::pair<int, IuwtRecompJobT<T>>> recomp;
  std::vector<std::pair<std::pair<int, int>, typename EngineTypes<T>::Conv>> conv;
};
struct SharedOps {
  SharedOpsT<float> f;
  SharedOpsT<double> d;
  template <class T>
  SharedOpsT<T>& of();
};
template <>
inline SharedOpsT<float>& SharedOps::of<float>() {
  return f;
}
template <>
inline SharedOpsT<double>& SharedOps::of<double>() {
  return d;
}

// Buffers and scales of one CG iteration of one job (also the key of a captured iteration graph).
// The vector pointers are of the job's element type (f64 selects double).
struct IterArgs {
  void *x, *xn, *r, *rn, *p, *ap;
  double *rr, *rnrn;
  int max_scale, adj, enforce, sub, p_sub, f64;
  bool operator<(const IterArgs& o) const {
    return std::tie(x, xn, r, rn, p, ap, rr, rnrn, max_scale, adj, enforce, sub, p_sub, f64) <
           std::tie(o.x, o.xn, o.r, o.rn, o.p, o.ap, o.rr, o.rnrn, o.max_scale, o.adj, o.enforce, o.sub, o.p_sub,
                    o.f64);
  }
};

// The MORESANE control flow of main.py:53-599 as a state machine.  Each step the task enqueues the
// GPU work of its phase on the runner's stream; after the runner's one synchronisation the task
// reads its scalar slot and runs the host decisions up to its next GPU phase.
class Task {
 public:
  enum Phase { SETUP, DECOMP, MINOR, UPDATE, REVERT, FINISH, DONE };
  Task(const TaskSpec& spec, Job* job, double* scal, int* iflag, double* upload_host);

  Phase phase() const { return phase_; }
  bool done() const { return phase_ == DONE; }
  // GPU work of the current phase, in three parts: pre, the batched middle (add_shared), post
  void enqueue_pre(hipStream_t st);
  void add_shared(SharedOps& ops);
  void enqueue_post(hipStream_t st);
  bool wants_iteration() const { return phase_ == MINOR && iterate_; }
  IterArgs iter_args() const;
  void consume(const double* slot_host);            // host decisions after the step's sync
  void fail(int code, const std::string& msg);

  Job* job() const { return job_; }
  double* scal() const { return scal_; }
  int* iflag() const { return iflag_; }
  bool geometry_moved() const { return moved_; }

  TaskSpec spec;
  Trace trace;
  int status = 0;
  std::string error;
  int major_iters = 0;  // major iterations of the last moresane call
  int complete = 0;
  int steps = 0;

 private:
  void start_call(int scale_count);
  void major_check();
  void inner_check();
  void choose_scales();
  void minor_exit();
  void after_inner();
  template <class T>
  void enqueue_pre_t(hipStream_t st);
  template <class T>
  void add_shared_t(SharedOps& ops);
  template <class T>
  void enqueue_post_t(hipStream_t st);
  template <class T>
  void enqueue_setup(hipStream_t st);
  template <class T>
  void enqueue_finish(hipStream_t st);
  bool f64() const { return spec.prm.precision != 0; }
  double read_dmax(const double* slot_host) const;

  Job* job_;
  double* scal_;          // device slot (kSlotDoubles doubles)
  int* iflag_;            // device slot (kSlotInts ints, zero between kernels)
  double* upload_;        // pinned host staging for this slot's small uploads
  Phase phase_ = SETUP;
  bool moved_ = false;
  bool by_scale_copied_ = false;

  // per moresane call
  Geometry g_{};
  int scale_count_ = 0;
  NoiseMask nm_{};
  std::vector<int> borders_;
  int upload_energies_ = 0;  // energies waiting in upload_ for the next enqueue
  int major_ = 0, min_scale_ = 0, max_scale_ = 0, scale_adjust_ = 0;
  double max_coeff_ = 1.0, std_current_ = 1000.0, std_ratio_ = 1.0;
  double dmax_ = 0.0;
  std::vector<double> nsm_;
  // minor loop
  bool minor_first_ = false, iterate_ = false;
  int niter_ = 0, exit_code_ = EXIT_CAP;
  double e2_ = 0.0, snr_last_ = 0.0, snr_current_ = 0.0;
  void *x_ = nullptr, *xn_ = nullptr, *r_ = nullptr, *rn_ = nullptr;  // the job's element type
  double *rr_ = nullptr, *rnrn_ = nullptr;
};

// Steps a set of tasks in lockstep on one stream: per step every task enqueues its phase's work,
// the CG iterations of all tasks that share a plane shape run as batched launches (one launch per
// kernel for up to kMaxBatch jobs), then one copy of the scalar block and one synchronisation.
// Each task's arithmetic is independent of the others, so a task gives the same result alone and
// in a batch.
class Runner {
 public:
  explicit Runner(int slots, hipStream_t st = nullptr);
  ~Runner();
  Runner(const Runner&) = delete;
  Runner& operator=(const Runner&) = delete;

  int slots() const { return (int)jobs_.size(); }
  int free_slots() const;
  // start a task in a free slot (the given one, or the first free); returns the slot or -1
  int add(const TaskSpec& spec, int slot = -1);
  int first_free() const;
  Task* task(int slot) { return tasks_[slot].get(); }
  void release(int slot);  // drop a finished task (its workspace is kept)
  bool idle() const;
  void refresh_options();      // re-read PMS_CG_GRAPH and PMS_SNR_DIRECT
  void step();                 // one lockstep step (no-op when idle)
  void run();                  // step until every task is done
  hipStream_t stream() const { return st_; }
  double wait_seconds() const { return wait_seconds_; }
  size_t device_bytes() const;

 private:
  void run_shared(const std::vector<int>& active);
  template <class T>
  void run_shared_t(SharedOpsT<T>& ops);
  void run_iterations(std::vector<Task*>& ts);
  void enqueue_iterations(const std::vector<IterArgs>& its, const std::vector<Task*>& ts);
  template <class T>
  void enqueue_iterations_t(const std::vector<IterArgs>& its, const std::vector<Task*>& ts);
  template <class T>
  void apply_op(const std::vector<Task*>& ts, const std::vector<IterArgs>& its, bool clip_pass);
  void drop_graphs();

  hipStream_t st_;
  bool owns_stream_ = false;
  std::vector<std::unique_ptr<Job>> jobs_;
  std::vector<std::unique_ptr<Task>> tasks_;
  double* scal_dev_ = nullptr;
  int* iflag_dev_ = nullptr;
  double* scal_host_ = nullptr;    // pinned readback mirror
  double* upload_host_ = nullptr;  // pinned upload staging, kSlotDoubles per slot
  double wait_seconds_ = 0.0;
  bool use_graph_ = false;
  bool snr_direct_ = false;
  int cg_iters_ = 1;
  std::map<std::pair<int, std::vector<IterArgs>>, hipGraphExec_t> graphs_;
};

}  // namespace pms
