// qd_abi.cpp - the extern "C" entry points of include/qd.h.
//
// qd_evolve is the drop-in for mexFunction (cudaOpenMPEvolution.C:15-79):
// it validates the input struct, makes every omega block resident on the
// current device, runs time.total_steps steps with the reference's step
// sequence (cudaOpenMP2.cu:338-387), writes the step counter, the wavepackets
// and the observables back into caller memory, and frees the device.
// No C++ exception or HIP error crosses this boundary: every failure becomes a
// qd_status with a message.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "host_internal.h"
#include "qd_test.h"

using namespace qd;

struct qd_handle {
  hipStream_t stream = nullptr;
  std::unique_ptr<DeviceJob> job;
  std::unique_ptr<Engine> engine;
  const qd_inputs *in = nullptr;   // borrowed, must outlive the handle
  std::string err;
  int last_checked = 0;
  bool use_graph = true;
};

static thread_local std::string g_last_error;

static qd_status fail(qd_handle *h, qd_status code, const std::string &msg) {
  if (h) h->err = msg;
  g_last_error = msg;
  return code;
}

#define QD_GUARD(h, ...)                                                 \
  try {                                                                  \
    __VA_ARGS__                                                          \
  } catch (const Error &e) {                                             \
    return fail((h), e.code, e.what());                                  \
  } catch (const std::bad_alloc &) {                                     \
    return fail((h), QD_ERR_NOMEM, "host allocation failed");            \
  } catch (const std::exception &e) {                                    \
    return fail((h), QD_ERR_DEVICE, e.what());                           \
  }

namespace {

void print_records(qd_handle *h, int from, int to) {
  const int nO = h->job->shape.nO;
  const int *om = h->in->omega_states.omegas;
  for (int r = from; r < to; r++) {
    const double *o = h->job->obs.data() + (size_t)r * nO * 5;
    double tn = 0.0, te = 0.0;
    std::printf("\n Step: %d\n", h->job->shape.steps0 + r);
    for (int b = 0; b < nO; b++) {   // wavepackets1device.cu:443-451
      const double *x = o + b * 5;
      const double e = x[1] + x[2] + x[3] + x[4];
      std::printf(" %d %.12f %.12f %.12f %.12f %.12f %.12f\n", om[b], x[0], x[1], x[2], x[3], x[4], e);
      tn += x[0];
      te += e;
    }
    std::printf(" Total: %.12f %.12f\n", tn, te);   // cudaOpenMP2.cu:442
  }
  std::fflush(stdout);
}

// With the fused Legendre step, the last enqueued step's backward transform,
// flux values and record are pending until the next step; run them now.
void finish_pending(qd_handle *h) {
  DeviceJob &j = *h->job;
  if (!j.pending) return;
  j.flux_after_step(h->stream);   // a full window is folded before the last surface write
  h->engine->finish();
  j.pending = false;
  j.flux_after_step(h->stream);
}

// drain the ring, check the norm of each new record, publish the step counter
qd_status sync_and_check(qd_handle *h) {
  QD_HIP(hipStreamSynchronize(h->stream));
  DeviceJob &j = *h->job;
  j.drain(j.records_enqueued());
  const int nO = j.shape.nO;
  // the reference aborts on norm > 1 with zero tolerance (cudaOpenMP2.cu:437-440);
  // allow the rounding of each form: FP64 sums, or FP32-accumulated Legendre products
  const double tol = j.shape.mixed ? 1e-6 : 1e-10;
  for (int r = h->last_checked; r < j.records_done; r++) {
    double tn = 0.0;
    for (int b = 0; b < nO; b++) tn += j.obs[(size_t)r * nO * 5 + b * 5];
    if (!(tn <= 1.0 + tol)) {
      if (h->in->print_energies) print_records(h, h->last_checked, r + 1);
      h->last_checked = r + 1;
      *h->in->time.steps = j.shape.steps0 + r + 1;
      char buf[160];
      std::snprintf(buf, sizeof buf, "total norm %.17g exceeds 1 at step %d (cudaOpenMP2.cu:437-440)", tn,
                    j.shape.steps0 + r);
      return fail(h, QD_ERR_NORM, buf);
    }
  }
  if (h->in->print_energies) print_records(h, h->last_checked, j.records_done);
  h->last_checked = j.records_done;
  *h->in->time.steps = j.steps;
  return QD_OK;
}

}  // namespace

extern "C" {

const char *qd_version(void) { return "qdamd 0.1 (gfx950)"; }

const char *qd_last_global_error(void) { return g_last_error.c_str(); }

const char *qd_last_error(const qd_handle *h) { return h ? h->err.c_str() : g_last_error.c_str(); }

qd_status qd_create(const qd_inputs *in, qd_handle **out) {
  if (!out) return fail(nullptr, QD_ERR_INVALID, "out pointer is NULL");
  *out = nullptr;
  std::unique_ptr<qd_handle> h(new qd_handle());
  QD_GUARD(nullptr, {
    const JobShape shape = validate_inputs(in);   // before any device work
    init_kernel_attributes();
    QD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->in = in;
    h->job.reset(new DeviceJob(in, shape, h->stream));
    h->engine.reset(new Engine(h->stream));
    h->engine->set_jobs({h->job.get()});
    const char *g = std::getenv("QD_NO_GRAPH");
    h->use_graph = !(g && g[0] == '1');
    *out = h.release();
    return QD_OK;
  })
}

void qd_destroy(qd_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->engine.reset();
  h->job.reset();
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

qd_status qd_run(qd_handle *h, int32_t n_steps) {
  if (!h) return fail(nullptr, QD_ERR_INVALID, "handle is NULL");
  if (n_steps < 0) return fail(h, QD_ERR_INVALID, "n_steps must be >= 0");
  QD_GUARD(h, {
    DeviceJob &j = *h->job;
    const qd_options &op = h->in->options;
    const bool copy = (op.wave_to_matlab || op.wave_callback) && op.steps_to_copy_psi_from_device_to_host > 0;
    const int N = op.steps_to_copy_psi_from_device_to_host;
    Engine &e = *h->engine;
    for (int s = 0; s < n_steps;) {
      j.prephase_if_first(h->stream);   // V(-dt/2) of step 0 (cudaOpenMP2.cu:344)
      // Steps in this launch: kGraphSteps as one graph when no host-side
      // event falls inside them (a copy-back follows the step whose
      // pre-step counter is a multiple of N, the flux window fills, the
      // observable ring fills), otherwise one.  With the fused Legendre
      // step, k steps from a pending state add k records and flux writes; a
      // step from a finished state runs alone (it adds none).
      int k = 1;
      if (h->use_graph && (!e.fused() || j.pending)) {
        int lim = std::min(n_steps - s, Engine::kGraphSteps);
        lim = std::min(lim, j.steps_to_flux_flush());
        lim = std::min(lim, j.dev.ring - (j.records_enqueued() - j.records_done));
        if (copy) {
          const int r = j.steps % N;
          lim = std::min(lim, r == 0 ? 1 : N - r + 1);
        }
        if (lim == Engine::kGraphSteps) k = lim;
      }
      if (h->use_graph) e.step_graph(k);
      else e.step();
      j.steps += k;
      if (e.fused()) j.pending = true;
      s += k;
      const int steps_before = j.steps - 1;   // pre-step counter of the launch's last step
      j.flux_after_step(h->stream);
      const bool copy_now = copy && steps_before % N == 0;
      if (copy_now || s == n_steps) finish_pending(h);
      const bool ring_full = j.records_enqueued() - j.records_done >= j.dev.ring;
      if (ring_full || copy_now || s == n_steps) {
        qd_status st = sync_and_check(h);
        if (st != QD_OK) return st;
      }
      if (copy_now) {   // cudaOpenMP2.cu:380-384, before steps++ (:386)
        j.copy_psi(h->in->omega_states.wave_packets, h->stream);
        *h->in->time.steps = steps_before;
        if (op.wave_callback) op.wave_callback(op.callback_user, steps_before);
        *h->in->time.steps = j.steps;
      }
    }
    return QD_OK;
  })
}

qd_status qd_synchronize(qd_handle *h) {
  if (!h) return fail(nullptr, QD_ERR_INVALID, "handle is NULL");
  QD_GUARD(h, { return sync_and_check(h); })
}

int32_t qd_steps_done(const qd_handle *h) { return h ? h->job->steps - h->job->shape.steps0 : 0; }

qd_status qd_copy_wavepackets(qd_handle *h, double *dst) {
  if (!h || !dst) return fail(h, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(h, {
    finish_pending(h);
    h->job->copy_psi(dst, h->stream);
    return QD_OK;
  })
}

qd_status qd_get_observables(qd_handle *h, int32_t first, int32_t count, double *dst) {
  if (!h || !dst) return fail(h, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(h, {
    qd_status st = sync_and_check(h);
    if (st != QD_OK) return st;
    DeviceJob &j = *h->job;
    if (first < 0 || count < 0 || first + count > j.records_done)
      return fail(h, QD_ERR_INVALID, "observable records out of range");
    const size_t rec = (size_t)j.shape.nO * 5;
    if (count > 0) std::memcpy(dst, j.obs.data() + (size_t)first * rec, sizeof(double) * rec * count);
    return QD_OK;
  })
}

qd_status qd_get_autocorrelation(qd_handle *h, int32_t first, int32_t count, double *dst) {
  if (!h || !dst) return fail(h, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(h, {
    qd_status st = sync_and_check(h);
    if (st != QD_OK) return st;
    DeviceJob &j = *h->job;
    if (!j.shape.acf) return fail(h, QD_ERR_STATE, "autocorrelation was not requested at create");
    if (first < 0 || count < 0 || first + count > j.records_done)
      return fail(h, QD_ERR_INVALID, "autocorrelation records out of range");
    if (count > 0) std::memcpy(dst, j.acf.data() + (size_t)first * 2, sizeof(double) * 2 * count);
    return QD_OK;
  })
}

qd_status qd_get_crp(qd_handle *h, double *crp) {
  if (!h || !crp) return fail(h, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(h, {
    if (!h->job->shape.crp) return fail(h, QD_ERR_STATE, "CRP was not requested (calculate_CRP == 0)");
    finish_pending(h);
    h->job->flux_to_crp(crp, h->stream);
    return QD_OK;
  })
}

qd_status qd_profile_steps(qd_handle *h, int32_t n_steps, double *us) {
  if (!h || !us || n_steps <= 0) return fail(h, QD_ERR_INVALID, "bad arguments");
  QD_GUARD(h, {
    hipEvent_t ev[2 * QD_KERNEL_KINDS];
    for (auto &e : ev) QD_HIP(hipEventCreate(&e));
    const unsigned timed = h->engine->timed_kinds();
    double acc[QD_KERNEL_KINDS] = {};
    qd_status st = QD_OK;
    DeviceJob &j = *h->job;
    int timed_steps = 0;
    for (int s = 0; s < n_steps && st == QD_OK; s++) {
      j.prephase_if_first(h->stream);
      // with the fused Legendre step, a step from a finished state runs the
      // forward transform alone: untimed unless it is the only step
      const bool time_it = !h->engine->fused() || j.pending || n_steps == 1;
      h->engine->step(time_it ? ev : nullptr);
      j.steps++;
      if (h->engine->fused()) j.pending = true;
      j.flux_after_step(h->stream);
      QD_HIP(hipStreamSynchronize(h->stream));
      if (time_it) {
        timed_steps++;
        for (int k = 0; k < QD_KERNEL_KINDS; k++) {
          if (!(timed & (1u << k))) continue;
          float ms = 0.f;
          QD_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
          acc[k] += ms * 1000.0;
        }
      }
      if (j.records_enqueued() - j.records_done >= j.dev.ring) st = sync_and_check(h);
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (st == QD_OK) {
      finish_pending(h);
      st = sync_and_check(h);
    }
    // mean per launch; the R pass launches twice per step, as two kinds
    for (int k = 0; k < QD_KERNEL_KINDS; k++) us[k] = timed_steps ? acc[k] / timed_steps : 0.0;
    return st;
  })
}

qd_status qd_evolve(qd_inputs *in) {
  qd_handle *h = nullptr;
  qd_status st = qd_create(in, &h);
  if (st != QD_OK) return st;
  const int total = in->time.total_steps;
  st = qd_run(h, total);
  if (st == QD_OK && in->observables) st = qd_get_observables(h, 0, total, in->observables);
  if (st == QD_OK && in->autocorrelation) st = qd_get_autocorrelation(h, 0, total, in->autocorrelation);
  if (st == QD_OK && in->final_copy) st = qd_copy_wavepackets(h, in->omega_states.wave_packets);
  if (st == QD_OK && h->job->shape.crp) st = qd_get_crp(h, in->crp.CRP);
  if (st != QD_OK) g_last_error = h->err;
  qd_destroy(h);
  return st;
}

// ---------------------------------------------------------------------------
// probes (include/qd_test.h)
// ---------------------------------------------------------------------------
qd_status qd_probe_legendre_form(const qd_inputs *in, int32_t *form) {
  if (!form) return fail(nullptr, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(nullptr, {
    const JobShape shape = validate_inputs(in);
    *form = legendre_step_form(in, shape);
    return QD_OK;
  })
}

qd_status qd_probe_fft_passes(const qd_inputs *in, int32_t *passes) {
  if (!passes) return fail(nullptr, QD_ERR_INVALID, "NULL argument");
  QD_GUARD(nullptr, {
    const JobShape shape = validate_inputs(in);
    *passes = fft_passes(shape, nullptr);
    return QD_OK;
  })
}

qd_status qd_probe_gemm(const double *A, int32_t rows, int32_t k, const double *B, int32_t n, double *C) {
  QD_GUARD(nullptr, {
    init_kernel_attributes();
    QD_REQUIRE(A && B && C && rows > 0 && k > 0 && n > 0 && n <= 128, "bad probe arguments");
    const int kp = (k + 3) / 4 * 4, np = (n + 15) / 16 * 16;
    std::vector<double> Bp((size_t)kp * np, 0.0);
    for (int kk = 0; kk < k; kk++)
      for (int c = 0; c < n; c++) Bp[(size_t)kk * np + c] = B[(size_t)c * k + kk];   // B is (k x n) col-major
    double *dA, *dB, *dC;
    // the kernel reads A rows k < kp (rows k >= K masked): allocate them
    QD_HIP(hipMalloc(&dA, sizeof(double) * rows * kp));
    QD_HIP(hipMalloc(&dB, sizeof(double) * Bp.size()));
    QD_HIP(hipMalloc(&dC, sizeof(double) * rows * n));
    QD_HIP(hipMemset(dA, 0, sizeof(double) * rows * kp));
    QD_HIP(hipMemcpy(dA, A, sizeof(double) * rows * k, hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(dB, Bp.data(), sizeof(double) * Bp.size(), hipMemcpyHostToDevice));
    launch_gemm_probe(dA, rows, dB, kp, np, dC, rows, rows, k, n, nullptr);
    QD_HIP(hipGetLastError());
    QD_HIP(hipMemcpy(C, dC, sizeof(double) * rows * n, hipMemcpyDeviceToHost));
    QD_HIP(hipFree(dA));
    QD_HIP(hipFree(dB));
    QD_HIP(hipFree(dC));
    return QD_OK;
  })
}

qd_status qd_probe_fft(double *data, int32_t n, int32_t count, int32_t inverse) {
  QD_GUARD(nullptr, {
    init_kernel_attributes();
    QD_REQUIRE(data && n >= 2 && n <= kFftMaxN && count > 0, "bad probe arguments");
    int m = n;
    for (int p : {2, 3, 5})
      while (m % p == 0) m /= p;
    QD_REQUIRE(m == 1, "n must factor into 2, 3, 5");
    std::vector<double2> tw(n);
    for (int e = 0; e < n; e++) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * e / n;
      tw[e] = make_double2((double)cosl(a), (double)sinl(a));
    }
    double2 *dd, *dt;
    QD_HIP(hipMalloc(&dd, sizeof(double2) * n * count));
    QD_HIP(hipMalloc(&dt, sizeof(double2) * n));
    QD_HIP(hipMemcpy(dd, data, sizeof(double2) * n * count, hipMemcpyHostToDevice));
    QD_HIP(hipMemcpy(dt, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    launch_fft_probe(dd, dt, n, count, inverse, nullptr);
    QD_HIP(hipGetLastError());
    QD_HIP(hipMemcpy(data, dd, sizeof(double2) * n * count, hipMemcpyDeviceToHost));
    QD_HIP(hipFree(dd));
    QD_HIP(hipFree(dt));
    return QD_OK;
  })
}

qd_status qd_fft_pencils_noalloc(double *data, int32_t n, int32_t count, int32_t inverse, void *stream) {
  QD_GUARD(nullptr, {
    fft_pencils_noalloc(reinterpret_cast<double2 *>(data), n, count, inverse, static_cast<hipStream_t>(stream));
    return QD_OK;
  })
}

qd_status qd_probe_copy(int64_t bytes, int32_t repeats, double *us_per_copy) {
  QD_GUARD(nullptr, {
    QD_REQUIRE(bytes >= 16 && repeats > 0 && us_per_copy, "bad probe arguments");
    const long n = bytes / 16;
    double2 *a, *b;
    QD_HIP(hipMalloc(&a, 16 * n));
    QD_HIP(hipMalloc(&b, 16 * n));
    QD_HIP(hipMemset(a, 0, 16 * n));
    hipEvent_t e0, e1;
    QD_HIP(hipEventCreate(&e0));
    QD_HIP(hipEventCreate(&e1));
    launch_copy(a, b, n, nullptr);
    QD_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; r++) launch_copy(r % 2 ? b : a, r % 2 ? a : b, n, nullptr);
    QD_HIP(hipEventRecord(e1, nullptr));
    QD_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    QD_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us_per_copy = ms * 1000.0 / repeats;
    QD_HIP(hipEventDestroy(e0));
    QD_HIP(hipEventDestroy(e1));
    QD_HIP(hipFree(a));
    QD_HIP(hipFree(b));
    return QD_OK;
  })
}

qd_status qd_probe_bandwidth(int64_t bytes, int32_t repeats, int32_t kind, double *us_per_pass) {
  QD_GUARD(nullptr, {
    QD_REQUIRE(bytes >= 16 * 2048 && repeats > 0 && us_per_pass && (kind == 1 || kind == 2),
               "bad probe arguments");
    const long n = bytes / 16;
    double2 *a, *b;
    QD_HIP(hipMalloc(&a, 16 * n));
    QD_HIP(hipMalloc(&b, kind == 1 ? 16 * n : 16 * 2048));
    QD_HIP(hipMemset(a, 0, 16 * n));
    hipEvent_t e0, e1;
    QD_HIP(hipEventCreate(&e0));
    QD_HIP(hipEventCreate(&e1));
    launch_bandwidth(kind, a, b, n, nullptr);
    QD_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; r++) {
      const bool swap = kind == 1 && (r % 2);
      launch_bandwidth(kind, swap ? b : a, swap ? a : b, n, nullptr);
    }
    QD_HIP(hipEventRecord(e1, nullptr));
    QD_HIP(hipEventSynchronize(e1));
    QD_HIP(hipGetLastError());
    float ms = 0.f;
    QD_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us_per_pass = ms * 1000.0 / repeats;
    QD_HIP(hipEventDestroy(e0));
    QD_HIP(hipEventDestroy(e1));
    QD_HIP(hipFree(a));
    QD_HIP(hipFree(b));
    return QD_OK;
  })
}

qd_status qd_probe_coriolis(int32_t J, int32_t l, int32_t omega_min, int32_t n, double *evals, double *evecs) {
  QD_GUARD(nullptr, {
    QD_REQUIRE(evals && evecs && n >= 1 && n <= 64, "bad probe arguments");
    std::vector<double> diag(n), off(std::max(1, n - 1));
    coriolis_tridiagonal(J, l, omega_min, n, diag.data(), off.data());
    tridiag_eigen(n, diag.data(), off.data(), evecs);
    for (int a = 0; a < n; a++) evals[a] = diag[a];
    return QD_OK;
  })
}

}  // extern "C"
