// server.cpp - the serving form: many resident partial-wave jobs advanced
// together, with admission by device-memory footprint and continuous
// batching (jobs join and leave between steps).
//
// The reference runs one partial wave per process (cudaOpenMPEvolution.C:15-79,
// one-shot because of the MatlabData singletons, matlabData.C:33-53); this
// file has no counterpart there.  Every step kernel already takes an array of
// job descriptors and a per-kernel unit prefix table (Engine), so the jobs of
// one shape class share each launch; the classes run on separate streams so
// their kernels overlap on the device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "host_internal.h"

using namespace qd;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Jobs share a class when every kernel instantiation of the step is the same
// for them: grid lengths (FFT kernels and tilings, including the R-pass rows,
// which depend on whether a job has more than 5 omega blocks), Legendre tile
// shapes and kernel forms (workgroup waves, paired or one-row strips, the
// fused Legendre step) and the precision.
using ClassKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int, int, int, int, int,
                            int, int, int, int, int>;

struct ServerJob {
  int id = 0;
  qd_inputs *in = nullptr;
  JobShape shape{};
  size_t estimate = 0;       // footprint estimate used for admission
  size_t charged = 0;        // device bytes actually held while running
  int state = QD_JOB_QUEUED; // guarded by qd_server::mu
  int target = 0;            // steps to run (time.total_steps)
  int enqueued = 0;          // steps enqueued so far
  int checked = 0;           // observable records norm-checked
  std::unique_ptr<DeviceJob> job;
  std::string err;
  double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
};

struct ShapeClass {
  ClassKey key;
  hipStream_t stream = nullptr;
  std::unique_ptr<Engine> engine;
  std::vector<ServerJob *> members;
  std::deque<hipEvent_t> in_flight;   // one event per enqueued step graph, oldest first
  std::vector<hipEvent_t> spare;
  bool dirty = true;
};

// Work of the server's host worker thread: building an admitted job (device
// allocation, tables, the wavepacket upload) and retiring a finished one
// (waiting for its last step, norm check, outputs, freeing its memory), so
// that the stepping thread only launches steps.
struct Task {
  enum Kind { BUILD, RETIRE } kind;
  ServerJob *sj;
  hipEvent_t after = nullptr;   // RETIRE: recorded on the class stream after the job's last step
};

constexpr int kDrainEvery = 512;   // records per job between ring drains (ring holds 1024)
// Jobs handed to the builder per server step: the builder works through them
// while the running classes keep the GPU busy.
constexpr int kAdmitPerStep = 4;
// Step graphs a class may have queued on its stream before the stepping
// thread waits for the oldest: bounds how far the host runs ahead.
constexpr int kMaxInFlight = 3;

// Device memory of the served jobs.  Allocations of retired jobs are kept
// and handed to later jobs of the same sizes (the served shapes repeat):
// hipFree waits for the device and holds the runtime meanwhile, which
// stalled the stepping thread's launches.  A failed hipMalloc returns the
// cache and retries.  The large copies (wavepackets, potential) go through
// two pinned staging buffers of 32 MB, so that the DMA engine moves them and
// the runtime is not held by a copy from pageable memory.  Used by the
// worker thread only (and by the server's destruction after it has joined).
class ServerMemory : public DeviceMemory {
 public:
  ~ServerMemory() override {
    trim();
    for (int b = 0; b < 2; b++) {
      if (ev_[b]) (void)hipEventSynchronize(ev_[b]);
      if (ev_[b]) (void)hipEventDestroy(ev_[b]);
      if (pin_[b]) (void)hipHostFree(pin_[b]);
    }
  }
  void *allocate(size_t bytes) override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = free_.find(bytes);
      if (it != free_.end()) {
        void *p = it->second;
        free_.erase(it);
        return p;
      }
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      trim();
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(QD_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
      }
    }
    return p;
  }
  void release(void *p, size_t bytes) override {
    std::lock_guard<std::mutex> lk(mu_);
    free_.emplace(bytes, p);
  }
  void trim() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &f : free_) (void)hipFree(f.second);
    free_.clear();
  }
  void h2d(void *dst, const void *src, size_t bytes, hipStream_t s) override {
    pinned();
    int b = 0;
    for (size_t off = 0; off < bytes; off += kChunk, b ^= 1) {
      const size_t n = std::min(kChunk, bytes - off);
      if (used_[b]) QD_HIP(hipEventSynchronize(ev_[b]));
      std::memcpy(pin_[b], (const char *)src + off, n);
      QD_HIP(hipMemcpyAsync((char *)dst + off, pin_[b], n, hipMemcpyHostToDevice, s));
      QD_HIP(hipEventRecord(ev_[b], s));
      used_[b] = true;
    }
  }
  void d2h(void *dst, const void *src, size_t bytes, hipStream_t s) override {
    pinned();
    int b = 0, pb = -1;
    size_t poff = 0, pn = 0;
    auto finish = [&]() {   // the previous chunk: wait, copy out of its staging buffer
      if (pb < 0) return;
      QD_HIP(hipEventSynchronize(ev_[pb]));
      std::memcpy((char *)dst + poff, pin_[pb], pn);
    };
    for (size_t off = 0; off < bytes; off += kChunk, b ^= 1) {
      const size_t n = std::min(kChunk, bytes - off);
      if (used_[b]) QD_HIP(hipEventSynchronize(ev_[b]));
      QD_HIP(hipMemcpyAsync(pin_[b], (const char *)src + off, n, hipMemcpyDeviceToHost, s));
      QD_HIP(hipEventRecord(ev_[b], s));
      used_[b] = true;
      finish();
      pb = b;
      poff = off;
      pn = n;
    }
    finish();
  }

 private:
  static constexpr size_t kChunk = 32u << 20;
  std::mutex mu_;
  std::multimap<size_t, void *> free_;
  void *pin_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
  bool used_[2] = {false, false};
  void pinned() {
    for (int b = 0; b < 2; b++) {
      if (!pin_[b]) QD_HIP(hipHostMalloc(&pin_[b], kChunk, hipHostMallocDefault));
      if (!ev_[b]) QD_HIP(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
    }
  }
};

}  // namespace

struct qd_server {
  size_t budget = 0, used = 0;   // guarded by mu
  int max_running = 0;
  int next_id = 1;
  int active = 0;                // admitted and not yet DONE / FAILED (guarded by mu)
  hipStream_t admit_stream = nullptr;   // the worker's stream
  int device = 0;
  std::deque<ServerJob *> queue;        // submitted, not admitted (stepping thread)
  ServerMemory memory;                  // device memory and staging of the served jobs (outlives jobs)
  std::map<int, std::unique_ptr<ServerJob>> jobs;
  std::vector<std::unique_ptr<ShapeClass>> classes;
  std::string err;

  std::mutex mu;
  std::condition_variable cv;
  // held by the worker for each task and by the stepping thread while it
  // captures a step graph: HIP calls of the worker (allocation, copies,
  // frees) during a capture on another thread have invalidated it
  std::mutex hip_mu;
  std::deque<Task> tasks;               // for the worker
  std::vector<ServerJob *> built;       // built by the worker, not yet in a class
  std::vector<ServerJob *> requeue;     // builds that ran out of device memory: back to the queue
  bool stop = false;
  int pending = 0;                      // tasks queued or running in the worker
  std::thread worker;

  ShapeClass &class_for(const DeviceJob &j) {
    const JobDev &d = j.dev;
    const ClassKey key{d.n1,      d.n2,     d.nt,       d.fwd_np,       d.bwd_np,           d.fwd_kp,
                       d.bwd_kp,  d.mixed,  d.fft_rows, d.fft_rows_inv, d.gemm_fwd_waves,   d.gemm_bwd_waves,
                       d.leg_fwd_paired, d.leg_bwd_paired, d.leg_step, d.step_waves, d.leg_split,
                       d.sp_th, d.sp_te, d.sp_to, d.fft2, d.mix_cols};
    for (auto &c : classes)
      if (c->key == key) return *c;
    std::unique_ptr<ShapeClass> c(new ShapeClass());
    c->key = key;
    QD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->engine.reset(new Engine(c->stream));
    classes.push_back(std::move(c));
    return *classes.back();
  }

  int running() {
    std::lock_guard<std::mutex> lk(mu);
    return active;
  }

  // caller holds mu
  void fail_job_locked(ServerJob *sj, const std::string &msg) {
    sj->state = QD_JOB_FAILED;
    sj->err = msg;
    sj->t_done = now_s();
  }

  void push_task(const Task &t) {
    {
      std::lock_guard<std::mutex> lk(mu);
      tasks.push_back(t);
      pending++;
    }
    cv.notify_all();
  }

  // ---- worker thread
  void worker_main() {
    (void)hipSetDevice(device);
    for (;;) {
      Task t;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !tasks.empty(); });
        if (tasks.empty()) return;   // stop requested and nothing left
        t = tasks.front();
        tasks.pop_front();
      }
      {
        std::lock_guard<std::mutex> hl(hip_mu);
        if (t.kind == Task::BUILD) build(t.sj);
        else retire_job(t.sj, t.after);
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        pending--;
      }
      cv.notify_all();
    }
  }

  void build(ServerJob *sj) {
    try {
      std::unique_ptr<DeviceJob> j(new DeviceJob(sj->in, sj->shape, admit_stream, &memory));
      std::lock_guard<std::mutex> lk(mu);
      sj->job = std::move(j);
      used += sj->job->footprint > sj->charged ? sj->job->footprint - sj->charged : 0;
      sj->charged = std::max(sj->charged, sj->job->footprint);
      sj->t_admit = now_s();
      built.push_back(sj);
    } catch (const Error &e) {
      std::lock_guard<std::mutex> lk(mu);
      used -= std::min(used, sj->charged);
      sj->charged = 0;
      active--;
      if (e.code == QD_ERR_NOMEM && active > 0) {   // the estimate was short: retry later
        sj->state = QD_JOB_QUEUED;
        requeue.push_back(sj);
      } else {
        fail_job_locked(sj, e.what());
      }
    } catch (const std::exception &e) {
      std::lock_guard<std::mutex> lk(mu);
      used -= std::min(used, sj->charged);
      sj->charged = 0;
      active--;
      fail_job_locked(sj, e.what());
    }
  }

  // records not yet norm-checked; returns false (job failed) on a norm past 1
  bool check_norms(ServerJob *sj, std::string &msg) {
    DeviceJob &j = *sj->job;
    const double tol = j.shape.mixed ? 1e-6 : 1e-10;
    const int nO = j.shape.nO;
    for (int r = sj->checked; r < j.records_done; r++) {
      double tn = 0.0;
      for (int b = 0; b < nO; b++) tn += j.obs[(size_t)r * nO * 5 + b * 5];
      if (!(tn <= 1.0 + tol)) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "total norm %.17g exceeds 1 at step %d (cudaOpenMP2.cu:437-440)", tn,
                      j.shape.steps0 + r);
        *sj->in->time.steps = j.shape.steps0 + r + 1;
        msg = buf;
        return false;
      }
    }
    sj->checked = j.records_done;
    return true;
  }

  void write_outputs(ServerJob *sj, hipStream_t s) {
    DeviceJob &j = *sj->job;
    qd_inputs *in = sj->in;
    *in->time.steps = j.steps;
    const size_t rec = (size_t)j.shape.nO * 5;
    if (in->observables && sj->target > 0)
      std::memcpy(in->observables, j.obs.data(), sizeof(double) * rec * sj->target);
    if (in->autocorrelation && j.shape.acf && sj->target > 0)
      std::memcpy(in->autocorrelation, j.acf.data(), sizeof(double) * 2 * sj->target);
    if (in->final_copy) j.copy_psi(in->omega_states.wave_packets, s);
    if (j.shape.crp) j.flux_to_crp(in->crp.CRP, s);
  }

  // after the job's last step (event `after` on its class stream): drain and
  // check its records, write its outputs, free its device memory
  void retire_job(ServerJob *sj, hipEvent_t after) {
    std::string msg;
    bool ok = true;
    try {
      QD_HIP(hipStreamWaitEvent(admit_stream, after, 0));
      QD_HIP(hipEventSynchronize(after));
      // every record (a job retired early for a failed norm check may leave
      // its last step unfinished: it writes no outputs)
      sj->job->drain(sj->job->records_enqueued());
      ok = check_norms(sj, msg);
      if (ok) write_outputs(sj, admit_stream);
    } catch (const std::exception &e) {
      ok = false;
      msg = e.what();
    }
    (void)hipEventDestroy(after);
    {
      // the outputs are in place: the job is done before its memory is
      // freed (hipFree may wait for the device)
      std::lock_guard<std::mutex> lk(mu);
      if (ok) {
        sj->state = QD_JOB_DONE;
        sj->t_done = now_s();
      } else {
        fail_job_locked(sj, msg);
      }
    }
    sj->job.reset();
    std::lock_guard<std::mutex> lk(mu);
    used -= std::min(used, sj->charged);
    sj->charged = 0;
    active--;
  }

  // ---- stepping thread
  void admit() {
    {
      std::lock_guard<std::mutex> lk(mu);
      for (ServerJob *sj : requeue) queue.push_front(sj);
      requeue.clear();
    }
    int admitted = 0;
    for (auto it = queue.begin(); it != queue.end();) {
      ServerJob *sj = *it;
      if (admitted >= kAdmitPerStep) break;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (max_running > 0 && active >= max_running) break;
        if (used + sj->estimate > budget) {   // backfill: a later, smaller job may fit
          ++it;
          continue;
        }
        used += sj->estimate;
        sj->charged = sj->estimate;
        sj->state = QD_JOB_RUNNING;
        active++;
      }
      admitted++;
      push_task({Task::BUILD, sj, nullptr});
      it = queue.erase(it);
    }
  }

  // drain the observable ring of every member, check the norm (class stream
  // idle); a job whose norm passed 1 fails and is handed to the worker
  void drain_class(ShapeClass &c) {
    QD_HIP(hipStreamSynchronize(c.stream));
    for (ServerJob *sj : c.members) {
      sj->job->drain(sj->job->records_enqueued());
      std::string msg;
      if (!check_norms(sj, msg)) {
        std::lock_guard<std::mutex> lk(mu);
        sj->err = msg;
        sj->target = sj->enqueued;   // retire it now; the worker reports the failure
      }
    }
  }

  hipEvent_t class_event(ShapeClass &c) {
    hipEvent_t e;
    if (!c.spare.empty()) {
      e = c.spare.back();
      c.spare.pop_back();
    } else {
      QD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return e;
  }

  // max_steps >= Engine::kGraphSteps lets a class advance by kGraphSteps
  // steps as one graph when none of its jobs reaches its target or fills its
  // flux window within them (the ring holds kDrainEvery + kGraphSteps)
  void step(int max_steps) {
    admit();
    {
      std::unique_lock<std::mutex> lk(mu);
      // with no class to step, wait for the worker to deliver a job (or to
      // finish what it holds), so that every call makes progress
      bool idle = true;
      for (auto &c : classes) idle &= c->members.empty();
      if (idle) cv.wait(lk, [&] { return !built.empty() || pending == 0; });
      for (ServerJob *sj : built) {
        ShapeClass &c = class_for(*sj->job);
        c.members.push_back(sj);
        c.dirty = true;
      }
      built.clear();
    }
    for (auto &cp : classes) {
      ShapeClass &c = *cp;
      if (c.members.empty()) continue;
      // bound the steps queued ahead on this class's stream
      while ((int)c.in_flight.size() >= kMaxInFlight) {
        QD_HIP(hipEventSynchronize(c.in_flight.front()));
        c.spare.push_back(c.in_flight.front());
        c.in_flight.pop_front();
      }
      if (c.dirty) {
        std::vector<DeviceJob *> js;
        for (ServerJob *sj : c.members) js.push_back(sj->job.get());
        c.engine->set_jobs(js);
        c.dirty = false;
      }
      for (ServerJob *sj : c.members) sj->job->prephase_if_first(c.stream);
      int k = 1;
      if (max_steps >= Engine::kGraphSteps) {
        k = Engine::kGraphSteps;
        for (ServerJob *sj : c.members)
          if (sj->target - sj->enqueued < k || sj->job->steps_to_flux_flush() < k) k = 1;
      }
      if (c.engine->graph_ready(k)) {
        c.engine->step_graph(k);
      } else {
        // capture only while the worker is between tasks; otherwise launch
        // the steps directly and capture on a later call
        std::unique_lock<std::mutex> hl(hip_mu, std::try_to_lock);
        if (hl.owns_lock()) c.engine->step_graph(k);
        else
          for (int i = 0; i < k; i++) c.engine->step();
      }
      bool drain = false;
      for (ServerJob *sj : c.members) {
        DeviceJob &j = *sj->job;
        j.steps += k;
        sj->enqueued += k;
        if (c.engine->fused()) j.pending = true;
        j.flux_after_step(c.stream);
        drain |= sj->enqueued < sj->target && j.records_enqueued() - j.records_done >= kDrainEvery;
      }
      // jobs at their target run their last step's backward transform, flux
      // values and record alone (the engine's job q is members[q])
      for (size_t q = 0; q < c.members.size(); q++) {
        ServerJob *sj = c.members[q];
        DeviceJob &j = *sj->job;
        if (sj->enqueued >= sj->target && j.pending) {
          c.engine->finish((int)q);
          j.pending = false;
          j.flux_after_step(c.stream);
        }
      }
      hipEvent_t e = class_event(c);
      QD_HIP(hipEventRecord(e, c.stream));
      c.in_flight.push_back(e);
      if (drain) drain_class(c);
      // finished jobs leave the class; the worker retires them once the
      // stream has passed their last step
      std::vector<ServerJob *> keep;
      for (ServerJob *sj : c.members) {
        if (sj->enqueued >= sj->target) {
          hipEvent_t done;
          QD_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
          QD_HIP(hipEventRecord(done, c.stream));
          push_task({Task::RETIRE, sj, done});
          c.dirty = true;
        } else {
          keep.push_back(sj);
        }
      }
      c.members.swap(keep);
    }
  }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
  }
};

namespace {

thread_local std::string g_server_error;

qd_status sfail(qd_server *s, qd_status code, const std::string &msg) {
  if (s) s->err = msg;
  g_server_error = msg;
  return code;
}

}  // namespace

#define QD_SERVER_GUARD(s, ...)                                        \
  try {                                                                \
    __VA_ARGS__                                                        \
  } catch (const Error &e) {                                           \
    return sfail((s), e.code, e.what());                               \
  } catch (const std::bad_alloc &) {                                   \
    return sfail((s), QD_ERR_NOMEM, "host allocation failed");         \
  } catch (const std::exception &e) {                                  \
    return sfail((s), QD_ERR_DEVICE, e.what());                        \
  }

extern "C" {

qd_status qd_server_create(int64_t hbm_budget_bytes, int32_t max_running, qd_server **out) {
  if (!out) return sfail(nullptr, QD_ERR_INVALID, "out pointer is NULL");
  *out = nullptr;
  std::unique_ptr<qd_server> s(new qd_server());
  QD_SERVER_GUARD(nullptr, {
    init_kernel_attributes();
    size_t free_b = 0, total_b = 0;
    QD_HIP(hipMemGetInfo(&free_b, &total_b));
    s->budget = hbm_budget_bytes > 0 ? (size_t)hbm_budget_bytes : (size_t)(0.9 * (double)free_b);
    s->max_running = max_running > 0 ? max_running : 0;
    QD_HIP(hipStreamCreateWithFlags(&s->admit_stream, hipStreamNonBlocking));
    QD_HIP(hipGetDevice(&s->device));
    qd_server *sp = s.get();
    s->worker = std::thread([sp] { sp->worker_main(); });
    *out = s.release();
    return QD_OK;
  })
}

void qd_server_destroy(qd_server *s) {
  if (!s) return;
  s->shutdown();   // the worker finishes its queued builds and retirements first
  for (auto &c : s->classes)
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &c : s->classes) {
    c->engine.reset();
    for (hipEvent_t e : c->in_flight) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->spare) (void)hipEventDestroy(e);
  }
  s->jobs.clear();   // frees the device jobs still held
  for (auto &c : s->classes)
    if (c->stream) (void)hipStreamDestroy(c->stream);
  if (s->admit_stream) (void)hipStreamDestroy(s->admit_stream);
  delete s;
}

const char *qd_server_last_error(const qd_server *s) { return s ? s->err.c_str() : g_server_error.c_str(); }

qd_status qd_server_submit(qd_server *s, qd_inputs *in, int32_t *job_id) {
  if (!s || !in || !job_id) return sfail(s, QD_ERR_INVALID, "NULL argument");
  QD_SERVER_GUARD(s, {
    QD_REQUIRE(!in->options.wave_to_matlab && !in->options.wave_callback,
               "per-step copy-back (wave_to_matlab / wave_callback) is not available in the server; use qd_run");
    std::unique_ptr<ServerJob> sj(new ServerJob());
    sj->shape = validate_inputs(in);
    sj->in = in;
    sj->id = s->next_id++;
    sj->target = in->time.total_steps;
    sj->estimate = DeviceJob::estimate_footprint(sj->shape, sj->shape.crp ? in->crp.n_energies : 0);
    QD_REQUIRE(sj->estimate <= s->budget, "job footprint exceeds the server's memory budget");
    sj->t_submit = now_s();
    if (sj->target == 0) {
      sj->state = QD_JOB_DONE;
      sj->t_admit = sj->t_done = sj->t_submit;
    } else {
      s->queue.push_back(sj.get());
    }
    *job_id = sj->id;
    s->jobs[sj->id] = std::move(sj);
    return QD_OK;
  })
}

qd_status qd_server_step(qd_server *s, int32_t *n_running, int32_t *n_queued) {
  return qd_server_advance(s, 1, n_running, n_queued);
}

qd_status qd_server_advance(qd_server *s, int32_t max_steps, int32_t *n_running, int32_t *n_queued) {
  if (!s) return sfail(nullptr, QD_ERR_INVALID, "server is NULL");
  if (max_steps < 1) return sfail(s, QD_ERR_INVALID, "max_steps must be >= 1");
  QD_SERVER_GUARD(s, {
    s->step(max_steps);
    int nq;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      nq = (int)(s->queue.size() + s->requeue.size());
    }
    if (n_running) *n_running = s->running();
    if (n_queued) *n_queued = nq;
    return QD_OK;
  })
}

qd_status qd_server_job_status(qd_server *s, int32_t job_id, int32_t *state, int32_t *steps_done, double *latency_s,
                               double *queued_s) {
  if (!s) return sfail(nullptr, QD_ERR_INVALID, "server is NULL");
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return sfail(s, QD_ERR_INVALID, "unknown job id");
  const ServerJob &sj = *it->second;
  std::lock_guard<std::mutex> lk(s->mu);
  if (state) *state = sj.state;
  if (steps_done) *steps_done = sj.enqueued;
  if (latency_s) *latency_s = (sj.t_done >= 0 ? sj.t_done : now_s()) - sj.t_submit;
  if (queued_s) *queued_s = sj.t_admit >= 0 ? sj.t_admit - sj.t_submit : -1.0;
  return QD_OK;
}

const char *qd_server_job_error(qd_server *s, int32_t job_id) {
  if (!s) return "";
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return "unknown job id";
  std::lock_guard<std::mutex> lk(s->mu);
  return it->second->err.c_str();
}

qd_status qd_server_memory(qd_server *s, int64_t *used, int64_t *budget) {
  if (!s) return sfail(nullptr, QD_ERR_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (used) *used = (int64_t)s->used;
  if (budget) *budget = (int64_t)s->budget;
  return QD_OK;
}

}  // extern "C"
