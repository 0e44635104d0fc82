// host_internal.h - host-side classes of libqdamd: validated job setup,
// device residency of one partial-wave job, and the step engine that launches
// the HIP kernels for one or many resident jobs.
#pragma once

#include <hip/hip_runtime_api.h>

#include <limits>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "qd.h"
#include "qd_internal.h"

namespace qd {

struct Error : std::runtime_error {
  qd_status code;
  Error(qd_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define QD_HIP(x)                                                                              \
  do {                                                                                         \
    hipError_t qd_e_ = (x);                                                                    \
    if (qd_e_ != hipSuccess)                                                                   \
      throw ::qd::Error(QD_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(qd_e_));     \
  } while (0)

#define QD_REQUIRE(cond, msg)                                   \
  do {                                                          \
    if (!(cond)) throw ::qd::Error(QD_ERR_INVALID, (msg));      \
  } while (0)

// coriolis.cpp
void coriolis_tridiagonal(int J, int l, int omega_min, int n, double *diag, double *off);
void tridiag_eigen(int n, double *d, double *e, double *z);

// Host description of one job after validation (shapes + borrowed pointers).
struct JobShape {
  int n1, n2, nt, M, L, lmax, nO, omin, kmax, J, parity;
  double dt, dr1, dr2, left1, left2, m1, m2;
  bool dump, crp, acf, mixed;
  int steps0;           // initial value of time.steps
  int legendre_n_max;
};

// Throws Error(QD_ERR_INVALID) on any inconsistency; reads only host memory.
JobShape validate_inputs(const qd_inputs *in);
// Form of the Legendre transforms of a validated job: 0 separate passes,
// 1 fused step, 2 fused step with parity-split products (host only).
int legendre_step_form(const qd_inputs *in, const JobShape &sh);
// FFT passes of a validated job: 2 (the two-pass form: r pass with the first
// half step, R pass with the R kinetic factor) or 3; *cols (if given)
// receives the two-pass r pass's columns per workgroup, 0 for 3 passes.
int fft_passes(const JobShape &sh, int *cols);

// Where a job's device memory and its large host<->device copies come from.
// The plain form allocates with hipMalloc, frees with hipFree and copies
// from and to the caller's pageable arrays; the server's form recycles
// allocations and copies through pinned staging buffers (server.cpp).
struct DeviceMemory {
  virtual ~DeviceMemory() = default;
  virtual void *allocate(size_t bytes);                 // Error(QD_ERR_NOMEM) on failure
  virtual void release(void *p, size_t bytes);
  // host -> device on `s`, ordered before later work on s; src may be reused on return
  virtual void h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
  // device -> host after the work already on `s`; complete on return
  virtual void d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
  static DeviceMemory &plain();
};

// Device memory owned by one job.
class DeviceJob {
 public:
  DeviceJob(const qd_inputs *in, const JobShape &shape, hipStream_t stream, DeviceMemory *mem = nullptr);
  ~DeviceJob();
  DeviceJob(const DeviceJob &) = delete;
  DeviceJob &operator=(const DeviceJob &) = delete;

  JobShape shape;
  JobDev dev{};          // descriptor (device pointers), uploaded by the engine
  size_t footprint = 0;  // bytes of device memory held
  int records_done = 0;  // records drained to the host
  int steps = 0;         // host mirror of the device step counter: steps enqueued
  bool prephased = false;// step-0 V(-dt/2) phase applied
  // host mirror of counters[3]: the last enqueued step ran the fused
  // Legendre kernel's sequence and its backward transform, potential
  // energy, flux values and record follow with the next step or finish
  bool pending = false;
  // observable records (and flux surface values) enqueued so far
  int records_enqueued() const { return steps - shape.steps0 - (pending ? 1 : 0); }
  int flux_steps() const { return steps - (pending ? 1 : 0); }
  std::vector<double> obs;      // [records][nO][5]
  std::vector<double> acf;      // [records][2]

  // copy the ring records not yet drained (requires the stream to be idle)
  void drain(int records_total);
  void copy_psi(double *dst, hipStream_t s) const;
  // before the first step of a job whose step counter is 0: apply exp(+i dt/2 V)
  void prephase_if_first(hipStream_t s);
  // flux: after each enqueued step or finish (steps and pending already
  // updated) fold a full window into the energy accumulators; flux_to_crp
  // folds the open window and writes P(E) for every CRP energy (synchronous,
  // after the last step's finish)
  void flux_after_step(hipStream_t s);
  // flux surface writes that still fit the open window (INT_MAX without flux)
  int steps_to_flux_flush() const;
  void flux_flush(hipStream_t s);
  void flux_to_crp(double *crp, hipStream_t s);

  static size_t estimate_footprint(const JobShape &s, int n_energies);

 private:
  DeviceMemory *mem_;
  std::vector<std::pair<void *, size_t>> allocs_;
  void *alloc(size_t bytes);
  template <class T> T *upload(std::vector<T> v, hipStream_t s);
  std::vector<std::shared_ptr<void>> staging_;   // host sources of the set-up copies
  int flux_window_start_ = 0;    // step counter at the first step of the open window
  double2 *flux_A_ = nullptr, *flux_B_ = nullptr, *flux_ph_ = nullptr;
  double *flux_E_ = nullptr, *flux_eta_ = nullptr, *flux_out_ = nullptr;
};

// central first-derivative weights on n (odd) points of spacing h
std::vector<double> first_derivative_weights(int n, double h);

// Launches the step kernels for a set of resident jobs.
class Engine {
 public:
  explicit Engine(hipStream_t s);
  ~Engine();
  void set_jobs(const std::vector<DeviceJob *> &jobs);
  // enqueue one step for every job; with ev != nullptr (2 QD_KERNEL_KINDS
  // events) records ev[2k], ev[2k+1] around the launch of each kind k in
  // timed_kinds()
  void step(hipEvent_t *ev = nullptr);
  // Whether the steps run the fused Legendre kernel (every job or none).
  // Then a step leaves each job's backward transform pending, and finish
  // runs it (the separate backward pass), with the flux surface values and
  // the reduction of that last step, for every job (q < 0) or job q only.
  bool fused() const { return fused_ != 0; }
  void finish(int q = -1);
  unsigned timed_kinds() const;
  const std::vector<DeviceJob *> &jobs() const { return jobs_; }
  hipStream_t stream() const { return stream_; }
  // graph of k consecutive steps (k = 1 or kGraphSteps), captured on first
  // use after set_jobs; replays on the stream.  A multi-step graph saves the
  // launch gap between graphs (about 9 us on the J=20 workload) k - 1 times.
  static constexpr int kGraphSteps = 8;
  void step_graph(int k = 1);
  // whether the graph of k steps is captured and instantiated (launching it
  // captures nothing)
  bool graph_ready(int k) const { return graph_exec_[k == 1 ? 0 : 1] != nullptr; }
  void drop_graph();

 private:
  hipStream_t stream_;
  std::vector<DeviceJob *> jobs_;
  JobDev *d_jobs_ = nullptr;
  int *d_bases_ = nullptr;      // [KK_COUNT][n_jobs + 1], then [n_jobs][KK_COUNT][2] for single jobs
  int n_alloc_jobs_ = 0;
  // host sources of the descriptor uploads, kept until upload_done_ passes
  std::vector<JobDev> descs_;
  std::vector<int> bases_;
  hipEvent_t upload_done_ = nullptr;
  int total_[KK_COUNT] = {};
  size_t lds_fftR_ = 0, lds_fftr_ = 0;
  int ct_n1_ = 0, ct_n2_ = 0;   // lengths of the compile-time FFT kernels, 0 = runtime path
  int fft2_ = 0;                // two-pass FFT form (every job of the engine)
  int fwd_tiles_ = 0, bwd_tiles_ = 0, fwd_kp_ = 0, bwd_kp_ = 0, nb_max_ = 1;
  int fwd_waves_ = 4, bwd_waves_ = 4;   // waves per GEMM workgroup (the same for every job)
  int mixed_ = 0, any_flux_ = 0, any_acf_ = 0;
  int fused_ = 0, split_ = 0, step_waves_ = 4, step_ttiles_ = 0, step_ltiles_ = 0;
  size_t lds_step_ = 0;
  hipGraph_t graph_[2] = {};           // [0] one step, [1] kGraphSteps steps
  hipGraphExec_t graph_exec_[2] = {};
  StepArgs args(int kind, int gate = 0) const;
  StepArgs args_one(int q, int kind) const;   // job q alone
  void launch_step_kernel(const StepArgs &a, int units, int finish, hipStream_t s) const;
  void enqueue(hipEvent_t *ev);
};

void init_kernel_attributes();

// In-place unnormalised DFT of `count` contiguous pencils of length n in
// device memory, on `stream`, with the twiddles in a static device table
// (no allocation; n <= kFftMaxN, else Error(QD_ERR_CAPACITY)).
void fft_pencils_noalloc(double2 *data, int n, int count, int inverse, hipStream_t stream);

}  // namespace qd
