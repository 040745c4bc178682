// Device-resident MORESANE driver: per-job state and the major/minor loops.
#pragma once

#include <map>
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
};

Geometry make_geometry(int n, int psf_n, const MoresaneParams& prm, int scale_count_in, int* subregion_out);

// byte sizes of the per-job device buffers
struct JobSizes {
  size_t s2, S_s2, mask, iuwt, N2, spec_sub, spec_full, fft, psf, ext, partial, pmax;
};

// device bytes one job of this geometry needs (its workspace plus nothing else)
size_t job_footprint_bytes(const Geometry& g);

// Buffers and scales of one CG iteration: the key of a captured iteration graph.
struct IterArgs {
  float *x, *xn, *r, *rn, *p, *ap;
  double *rr, *rnrn;
  int max_scale, adj, enforce, sub, p_sub;
  bool operator<(const IterArgs& o) const {
    return std::tie(x, xn, r, rn, p, ap, rr, rnrn, max_scale, adj, enforce, sub, p_sub) <
           std::tie(o.x, o.xn, o.r, o.rn, o.p, o.ap, o.rr, o.rnrn, o.max_scale, o.adj, o.enforce, o.sub, o.p_sub);
  }
};

// Device buffers needed by one moresane() call, kept across calls of the same shape.
class Job {
 public:
  explicit Job(hipStream_t st);
  ~Job();
  int moresane(const MoresaneParams& prm, int scale_count, const float* dirty_dev, const float* psf_dev,
               const float* mask_dev, float* model_dev, float* residual_dev, int* complete, Trace& trace);
  void moresane_by_scale(const MoresaneParams& prm, int start_scale, int stop_scale, const float* dirty_dev,
                         const float* psf_dev, const float* mask_dev, float* model_dev, float* residual_dev,
                         Trace& trace);
  size_t device_bytes() const { return arena_.bytes(); }
  double wait_seconds() const { return wait_seconds_; }
  void reset_wait() { wait_seconds_ = 0.0; }
  hipStream_t stream() const { return st_; }

 private:
  void ensure(const Geometry& g);
  void setup_psf(const float* psf_dev, std::vector<double>& energies);
  void apply_op_masked(const float* v, float* out, int max_scale, int adj, const int* pred = nullptr);
  double residual_and_std(const float* dirty_dev, const float* model_dev);
  void enqueue_iteration(const IterArgs& a);
  void run_iteration(const IterArgs& a, bool use_graph);
  void drop_graphs();
  void sync();
  double wait_seconds_ = 0.0;
  std::map<IterArgs, hipGraphExec_t> graphs_;

  hipStream_t st_;
  bool owns_stream_ = false;
  DeviceArena arena_;
  DeviceArena by_scale_arena_;
  Geometry geo_{};
  JobSizes cap_{};
  bool have_ = false;

 public:
  // device buffers (sub = subregion side, S = scale capacity)
  float *dirty_sub = nullptr, *psf_sub = nullptr, *D = nullptr, *Dcg = nullptr, *E = nullptr;
  uint8_t* M = nullptr;
  float *rec = nullptr, *x = nullptr, *xn = nullptr, *r = nullptr, *rn = nullptr, *p = nullptr, *ap = nullptr,
        *cbuf = nullptr;
  float* iuwt_work = nullptr;
  float *model_local = nullptr, *resid = nullptr, *conv_full = nullptr, *psf_frame = nullptr;
  float2 *spec_sub = nullptr, *spec_full = nullptr, *spec_full_buf = nullptr, *fft_work = nullptr;
  void *mad_work = nullptr, *ext_work = nullptr;
  double* partial = nullptr;
  float* pmax = nullptr;
  double* scal = nullptr;
  int* iflag = nullptr;
  int* border_dev = nullptr;
};

}  // namespace pms
