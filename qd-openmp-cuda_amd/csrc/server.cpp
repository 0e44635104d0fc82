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
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
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
// shapes and the precision.
using ClassKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int>;

struct ServerJob {
  int id = 0;
  qd_inputs *in = nullptr;
  JobShape shape{};
  size_t estimate = 0;       // footprint estimate used for admission
  size_t charged = 0;        // device bytes actually held while running
  int state = QD_JOB_QUEUED;
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
  bool dirty = true;
};

constexpr int kDrainEvery = 512;   // records per job between ring drains (ring holds 1024)
// Jobs admitted per server step.  Building a job (allocation, tables, the
// wavepacket upload: up to 0.7 GB for 9 blocks of 256 x 192 x 96) is host
// work; admitting a few per step lets the classes already running keep the
// GPU busy meanwhile, instead of a burst of arrivals holding it idle.
constexpr int kAdmitPerStep = 4;

}  // namespace

struct qd_server {
  size_t budget = 0, used = 0;
  int max_running = 0;
  int next_id = 1;
  hipStream_t admit_stream = nullptr;
  std::deque<ServerJob *> queue;
  std::map<int, std::unique_ptr<ServerJob>> jobs;
  std::vector<std::unique_ptr<ShapeClass>> classes;
  std::string err;

  ShapeClass &class_for(const DeviceJob &j) {
    const JobDev &d = j.dev;
    const ClassKey key{d.n1,      d.n2,     d.nt,       d.fwd_np,       d.bwd_np,           d.fwd_kp,
                       d.bwd_kp,  d.mixed,  d.fft_rows, d.fft_rows_inv, d.gemm_fwd_waves,   d.gemm_bwd_waves};
    for (auto &c : classes)
      if (c->key == key) return *c;
    std::unique_ptr<ShapeClass> c(new ShapeClass());
    c->key = key;
    QD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->engine.reset(new Engine(c->stream));
    classes.push_back(std::move(c));
    return *classes.back();
  }

  int running() const {
    int n = 0;
    for (auto &c : classes) n += (int)c->members.size();
    return n;
  }

  void fail_job(ServerJob *sj, const std::string &msg) {
    sj->state = QD_JOB_FAILED;
    sj->err = msg;
    sj->t_done = now_s();
  }

  void admit() {
    int admitted = 0;
    for (auto it = queue.begin(); it != queue.end();) {
      ServerJob *sj = *it;
      if (max_running > 0 && running() >= max_running) break;
      if (admitted >= kAdmitPerStep) break;
      if (used + sj->estimate > budget) {   // backfill: a later, smaller job may fit
        ++it;
        continue;
      }
      try {
        sj->job.reset(new DeviceJob(sj->in, sj->shape, admit_stream));
      } catch (const Error &e) {
        sj->job.reset();
        if (e.code == QD_ERR_NOMEM && running() > 0) {   // the estimate was short: retry later
          ++it;
          continue;
        }
        fail_job(sj, e.what());
        it = queue.erase(it);
        continue;
      }
      sj->charged = std::max(sj->estimate, sj->job->footprint);
      used += sj->charged;
      sj->state = QD_JOB_RUNNING;
      sj->t_admit = now_s();
      admitted++;
      ShapeClass &c = class_for(*sj->job);
      c.members.push_back(sj);
      c.dirty = true;
      it = queue.erase(it);
    }
  }

  // drain the observable ring of every member, check the norm; class stream idle
  void drain_class(ShapeClass &c) {
    QD_HIP(hipStreamSynchronize(c.stream));
    for (ServerJob *sj : c.members) {
      if (sj->state != QD_JOB_RUNNING) continue;
      DeviceJob &j = *sj->job;
      j.drain(sj->enqueued);
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
          fail_job(sj, buf);
          break;
        }
      }
      if (sj->state == QD_JOB_RUNNING) sj->checked = j.records_done;
    }
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

  // retire finished and failed members of a class (class stream idle)
  void retire(ShapeClass &c) {
    std::vector<ServerJob *> keep;
    for (ServerJob *sj : c.members) {
      const bool finished = sj->state == QD_JOB_RUNNING && sj->enqueued >= sj->target;
      if (finished) {
        try {
          write_outputs(sj, c.stream);
          sj->state = QD_JOB_DONE;
          sj->t_done = now_s();
        } catch (const Error &e) {
          fail_job(sj, e.what());
        }
      }
      if (sj->state == QD_JOB_RUNNING) {
        keep.push_back(sj);
      } else {
        sj->job.reset();
        used -= std::min(used, sj->charged);
        c.dirty = true;
      }
    }
    c.members.swap(keep);
  }

  // max_steps >= Engine::kGraphSteps lets a class advance by kGraphSteps
  // steps as one graph when none of its jobs reaches its target or fills its
  // flux window within them (the ring holds kDrainEvery + kGraphSteps)
  void step(int max_steps) {
    admit();
    // enqueue the steps of every class first, so the classes' streams run
    // concurrently; then drain / retire where needed
    std::vector<ShapeClass *> settle;
    for (auto &cp : classes) {
      ShapeClass &c = *cp;
      if (c.members.empty()) continue;
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
      c.engine->step_graph(k);
      bool finishing = false, drain = false;
      for (ServerJob *sj : c.members) {
        DeviceJob &j = *sj->job;
        j.steps += k;
        sj->enqueued += k;
        j.flux_after_step(c.stream);
        finishing |= sj->enqueued >= sj->target;
        drain |= sj->enqueued - j.records_done >= kDrainEvery;
      }
      if (finishing || drain) settle.push_back(&c);
    }
    for (ShapeClass *c : settle) {
      drain_class(*c);
      retire(*c);
    }
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
    *out = s.release();
    return QD_OK;
  })
}

void qd_server_destroy(qd_server *s) {
  if (!s) return;
  for (auto &c : s->classes)
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &c : s->classes) c->engine.reset();
  s->jobs.clear();   // frees the device jobs
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
    if (n_running) *n_running = s->running();
    if (n_queued) *n_queued = (int32_t)s->queue.size();
    return QD_OK;
  })
}

qd_status qd_server_job_status(qd_server *s, int32_t job_id, int32_t *state, int32_t *steps_done, double *latency_s,
                               double *queued_s) {
  if (!s) return sfail(nullptr, QD_ERR_INVALID, "server is NULL");
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return sfail(s, QD_ERR_INVALID, "unknown job id");
  const ServerJob &sj = *it->second;
  if (state) *state = sj.state;
  if (steps_done) *steps_done = sj.enqueued;
  if (latency_s) *latency_s = (sj.t_done >= 0 ? sj.t_done : now_s()) - sj.t_submit;
  if (queued_s) *queued_s = sj.t_admit >= 0 ? sj.t_admit - sj.t_submit : -1.0;
  return QD_OK;
}

const char *qd_server_job_error(qd_server *s, int32_t job_id) {
  if (!s) return "";
  auto it = s->jobs.find(job_id);
  return it == s->jobs.end() ? "unknown job id" : it->second->err.c_str();
}

qd_status qd_server_memory(qd_server *s, int64_t *used, int64_t *budget) {
  if (!s) return sfail(nullptr, QD_ERR_INVALID, "server is NULL");
  if (used) *used = (int64_t)s->used;
  if (budget) *budget = (int64_t)s->budget;
  return QD_OK;
}

}  // extern "C"
