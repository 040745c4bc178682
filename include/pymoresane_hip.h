/*
 * pymoresane_hip.h -- C ABI of the MI355X (gfx950) MORESANE deconvolution library (libpmship.so).
 *
 * Every entry point returns an int status (0 = OK, < 0 = error) and leaves a
 * message for pms_last_error().  Plain pointers and sizes only; all image
 * buffers are row-major float32 planes in device memory unless a parameter is
 * marked "host".  Plane stacks are (planes, rows, cols) with the plane index
 * outermost.  A stream argument may be NULL (the default stream).  The caller
 * owns every buffer passed in; the library frees only what it allocates.
 *
 * Each operation names the reference function it replaces (ratt-ru/PyMORESANE
 * 0.3.6); the reference binds them from Python, so the ctypes stubs in
 * pymoresane_amd/_lib.py are the binding a maintainer would add.
 */
#ifndef PYMORESANE_HIP_H
#define PYMORESANE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define PMS_STATUS_OK 0
#define PMS_STATUS_ERR_ARG (-1)
#define PMS_STATUS_ERR_HIP (-2)
#define PMS_STATUS_ERR_ALLOC (-3)
#define PMS_STATUS_ERR_STATE (-4)
#define PMS_STATUS_ERR_INTERNAL (-5)

/* Bind the calling thread to a device and build its FFT twiddle table.
 * Replaces the pyCUDA context set-up done at import (pymoresane/iuwt.py:6-14). */
int pms_init(int device);
/* Message of the last failed call on this thread (thread-local). */
const char* pms_last_error(void);
int pms_get_device_count(int* count);
int pms_device_synchronize(void);

/* ------------------------------------------------------ memory and streams */
int pms_malloc(void** ptr, size_t bytes);
int pms_free(void* ptr);
int pms_memcpy_h2d(void* dst, const void* src_host, size_t bytes, void* stream);
int pms_memcpy_d2h(void* dst_host, const void* src, size_t bytes, void* stream);
int pms_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int pms_fill_f32(float* ptr, size_t count, float value, void* stream);
/* dst[i] = src[i] for count floats with a plain streaming kernel (also the
 * reference point for calibrating HBM traffic counters). */
int pms_copy_f32(float* dst, const float* src, size_t count, void* stream);
int pms_stream_create(void** stream);
int pms_stream_destroy(void* stream);
int pms_stream_synchronize(void* stream);
int pms_event_create(void** event);
int pms_event_destroy(void* event);
int pms_event_record(void* event, void* stream);
int pms_event_elapsed_ms(void* start, void* stop, float* ms);

/* -------------------------------------------------------------------- IUWT */
/* Scratch bytes for one decomposition/recomposition of an h x w plane. */
size_t pms_iuwt_work_bytes(int h, int w);

/* Detail planes for scales scale_adjust..scale_count-1 of `in` (h x w) into
 * detail_out ((scale_count-scale_adjust) planes); smoothed_out (h x w) may be
 * NULL.  Requires 2^scale_count <= min(h, w).
 * Replaces iuwt.iuwt_decomposition (pymoresane/iuwt.py:17-41, gpu path 384-499). */
int pms_iuwt_decompose(const float* in, int h, int w, int scale_count, int scale_adjust, float* detail_out,
                       float* smoothed_out, void* work, void* stream);

/* Recompose n_planes detail planes (scales scale_adjust..) with optional
 * smoothed base (NULL = zeros) into out (h x w).
 * Replaces iuwt.iuwt_recomposition (pymoresane/iuwt.py:43-64, gpu path 501-581). */
int pms_iuwt_recompose(const float* detail, int n_planes, int h, int w, int scale_adjust, const float* smoothed,
                       float* out, void* work, void* stream);

/* --------------------------------------------------------- FFT convolution */
/* Bytes of a PSF spectrum for a p x p frame ((p/2+1) x p complex64, library layout). */
size_t pms_fft_spectrum_bytes(int p);
/* Scratch bytes for pms_fft_convolve(n, p) and pms_psf_spectrum(q, p). */
size_t pms_fft_work_bytes(int n, int p);

/* Spectrum of a q x q PSF embedded centred in a p x p frame (q == p: the frame
 * itself; q == p/2: pad_array).  p must be a power of two <= 16384.
 * Replaces the PSF-spectrum precompute (pymoresane/main.py:135-213 via
 * iuwt_convolution.gpu_r2c_fft, pymoresane/iuwt_convolution.py:76-106). */
int pms_psf_spectrum(const float* psf, int q, int p, void* spectrum_out, void* work, void* stream);

/* Host-side conversion of a numpy rfft2 spectrum (p x (p/2+1) complex64,
 * interleaved re/im, host memory) into the library layout (host memory). */
int pms_spectrum_from_rfft2(const float* rfft2_host, int p, void* spectrum_host_out);

/* out (n x n) = in (n x n) convolved with the PSF of `spectrum`; p = 2n is
 * the reference's "linear" mode, p = n its "circular" mode.
 * Replaces iuwt_convolution.fft_convolve (pymoresane/iuwt_convolution.py:18-73). */
int pms_fft_convolve(const float* in, int n, int p, const void* spectrum, float* out, void* work, void* stream);

/* ------------------------------------------------------ noise / threshold */
size_t pms_threshold_work_bytes(int n_planes, int h, int w);

/* Per-plane MAD noise estimate median(|w|)/0.6745 over the noise mask
 * (edge_excl border and int_excl centre square excluded) into
 * thresholds_out (device, n_planes doubles).
 * Replaces iuwt_toolbox.estimate_threshold (pymoresane/iuwt_toolbox.py:17-48). */
int pms_estimate_threshold(const float* planes, int n_planes, int h, int w, int edge_excl, int int_excl,
                           double* thresholds_out, void* work, void* stream);

/* pms_estimate_threshold without a workspace argument and without any device
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The median's select state and histograms live in one
 * fixed device array per GPU (131,584 bytes, held until the process ends),
 * so n_planes is limited to PMS_NOALLOC_MAX_PLANES; larger stacks are
 * rejected with PMS_STATUS_ERR_ARG.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (different streams or host threads)
 * race and must be serialised by the caller.  Same results as
 * pms_estimate_threshold, bit for bit. */
#define PMS_NOALLOC_MAX_PLANES 16
int pms_estimate_threshold_noalloc(const float* planes, int n_planes, int h, int w, int edge_excl, int int_excl,
                                   double* thresholds_out, void* stream);

/* out = planes where |planes| > sigma_level * thresholds[plane], else 0.
 * Replaces iuwt_toolbox.apply_threshold (pymoresane/iuwt_toolbox.py:50-75). */
int pms_apply_threshold(const float* planes, int n_planes, int h, int w, const double* thresholds,
                        double sigma_level, float* out, void* work, void* stream);

/* ------------------------------------------------------- source extraction */
size_t pms_extract_work_bytes(int n_planes, int h, int w);

/* 8-connected objects of the non-zero pixels per plane, kept coarse to fine
 * by the tolerance rule; writes mask_out (u8 {0,1}), extracted_out
 * (mask * planes) and, when labels_out is not NULL, the per-plane object
 * labels numbered in raster order of each object's first pixel.
 * Replaces iuwt_toolbox.source_extraction (pymoresane/iuwt_toolbox.py:77-94,
 * cpu path 97-154, gpu path 156-290). */
int pms_source_extraction(const float* planes, int n_planes, int h, int w, double tolerance, int neg_comp,
                          float* extracted_out, unsigned char* mask_out, int* labels_out, void* work, void* stream);

/* 20 log10(||a|| / ||a - b||) over n elements, float64 sums, result in host memory.
 * Replaces iuwt_toolbox.snr_ratio (pymoresane/iuwt_toolbox.py:292-306). */
size_t pms_snr_work_bytes(size_t n);
int pms_snr_ratio(const float* a, const float* b, size_t n, double* snr_host, void* work, void* stream);

/* ------------------------------------------------------------ MORESANE driver */
enum {
  PMS_EV_SCALES = 1,  /* a = min_scale, b = max_scale, c = scale_adjust, x = max_coeff */
  PMS_EV_MINOR = 2,   /* a = minor iterations, b = exit code, x = final SNR */
  PMS_EV_UPDATE = 3,  /* a = reverted, x = residual std, y = std ratio */
  PMS_EV_NOCOEFF = 4, /* no significant wavelet coefficient */
  PMS_EV_POOR = 5,    /* all scales performing poorly */
  PMS_EV_NOWORK = 6,  /* call did no work (complete) */
  PMS_EV_BYSCALE = 7  /* a = scale_count of the next moresane call */
};

typedef struct pms_trace_event {
  int kind;
  int major;
  int a, b, c, pad;
  double x, y;
} pms_trace_event;

/* Arguments of DataImage.moresane (pymoresane/main.py:53-57); 0 for
 * subregion/scale_count means "None" there. conv_mode: 0 linear, 1 circular.
 * precision: 0 = float32 bulk arithmetic (wavelet planes, convolutions, CG
 * vectors, model and residual) with float64 inner products, step sizes and
 * norms; 1 = float64 throughout the device work, the arithmetic of the
 * reference's serial numpy path.  Inputs and outputs stay float32 either way. */
typedef struct pms_moresane_params {
  int n;      /* dirty image side (square) */
  int psf_n;  /* PSF side: n or 2n */
  int subregion;
  int scale_count;
  double sigma_level, loop_gain, tolerance, accuracy;
  int major_loop_miter, minor_loop_miter;
  int conv_mode;
  int enforce_positivity;
  int edge_suppression;
  int edge_offset;
  double flux_threshold;
  int neg_comp, edge_excl, int_excl, precision;
} pms_moresane_params;

/* One MORESANE run on device buffers: model_inout accumulates the model
 * (self.model += model), residual_inout receives the residual when work was
 * done; complete_out is set when the call did no work.  The decision trace is
 * written to trace[0..trace_cap) (host memory) and its full length to
 * trace_len.
 * Replaces DataImage.moresane (pymoresane/main.py:53-521). */
int pms_moresane(const pms_moresane_params* prm, const float* dirty, const float* psf, const float* mask,
                 float* model_inout, float* residual_inout, int* complete_out, pms_trace_event* trace,
                 int trace_cap, int* trace_len, void* stream);

/* Scale-by-scale MORESANE (start_scale .. stop_scale).
 * Replaces DataImage.moresane_by_scale (pymoresane/main.py:523-599). */
int pms_moresane_by_scale(const pms_moresane_params* prm, int start_scale, int stop_scale, const float* dirty,
                          const float* psf, const float* mask, float* model_inout, float* residual_inout,
                          pms_trace_event* trace, int trace_cap, int* trace_len, void* stream);

/* ------------------------------------------------------------- batched form */
/* Independent deconvolution jobs run together.  Jobs are grouped into size
 * classes (image side, PSF side, subregion, convolution mode), each class
 * split into n_workers / n_classes groups when workers outnumber classes;
 * each of up to n_workers host threads takes a group and steps up to
 * max_batch of its jobs in lockstep on its own stream, refilling slots as
 * jobs finish: per step one
 * launch per kernel serves every job's CG iteration, followed by one copy of
 * the jobs' decision scalars and one synchronisation.  max_batch = 1 with one
 * worker is the plain one-job-at-a-time path.  Each job's result is
 * bit-identical to running it alone.  No reference counterpart: the
 * reference processes one image per call. */
typedef struct pms_batch_job {
  const float* dirty;
  const float* psf;
  const float* mask;
  float* model;     /* in/out, device */
  float* residual;  /* in/out, device */
  pms_moresane_params prm;
  int by_scale;
  int start_scale;
  int stop_scale;
  int status;       /* out: 0 or an error code */
  int major_iters;  /* out: major iterations of the last moresane call */
  int complete;     /* out */
  double seconds;   /* out: wall time of the job */
  double t_start;   /* out: start offset from the batch start, seconds */
  double wait_seconds; /* out: host time its runner spent waiting on the GPU while the job ran */
} pms_batch_job;

int pms_batch_run(pms_batch_job* jobs, int n_jobs, int n_workers, int max_batch);

/* ------------------------------------------------------------- served form */
/* Continuous-batching server: requests join the FIFO queue of their size
 * class; each class has a host thread and stream stepping up to max_batch
 * admitted requests in lockstep (as the batched form does).  A queued request
 * is admitted between steps as soon as its class has a free slot and its
 * device footprint fits the HBM budget shared by all classes, so requests that
 * arrive while others run join the running batch.  Host buffers must stay
 * valid until the request is done.  max_batch = 1 is the plain path: one
 * lane for all classes, one request at a time.  Times are seconds since the
 * server was created.  No reference counterpart: the reference has no
 * serving path. */
typedef struct pms_serve_request {
  const float* dirty_host;  /* n x n */
  const float* psf_host;    /* psf_n x psf_n */
  const float* mask_host;   /* n x n or NULL */
  float* model_host;        /* out, n x n */
  float* residual_host;     /* out, n x n */
  pms_moresane_params prm;
  int by_scale, start_scale, stop_scale;
  int status;               /* 1 queued, 2 running, 0 done, < 0 error */
  int major_iters;
  int pad;
  size_t footprint;         /* out: device bytes reserved while running */
  double t_submit, t_start, t_done;
} pms_serve_request;

int pms_server_create(int max_batch, double budget_bytes, void** server);
int pms_server_submit(void* server, pms_serve_request* req);
int pms_server_wait(void* server, pms_serve_request* req, double timeout_s, int* done);
int pms_server_drain(void* server);
int pms_server_now(void* server, double* t);
int pms_server_destroy(void* server);

#ifdef __cplusplus
}
#endif

#endif /* PYMORESANE_HIP_H */
