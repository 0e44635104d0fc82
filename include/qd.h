/*
 * qd.h - C ABI of the MI355X split-operator wavepacket propagator.
 *
 * Drop-in boundary for the reference's MEX entry point
 *   void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
 *   (cudaOpenMPEvolution.C:15-79)
 * which parses one Matlab struct (fields r1, r2, theta, pot, time, options,
 * dump1, dump2, CRP, OmegaStates) into zero-copy views of caller memory
 * (matlabStructures.C:17-153), builds CudaOpenMPQMMD and runs
 * test_multiple_cards() (cudaOpenMP2.cu:328-390).
 *
 * Here the same fields arrive as a plain C struct of pointers and sizes.
 * All arrays are column-major (first index fastest) exactly as Matlab lays
 * them out; complex arrays are interleaved (re, im) doubles, byte-identical
 * to the reference's real (2*n1, n2, ntheta, n_omegas) wave_packets array
 * (InitWavePacket.m:35-40, FH2main.m:169).
 *
 * Ownership: every pointer is borrowed; caller memory must outlive the
 * handle. The handle owns all device memory.  Errors are returned as codes
 * with a message (qd_last_error); nothing here exits or aborts the process.
 * One host thread per handle; several handles may coexist in one process
 * (unlike the reference, whose MatlabData singletons make it one-shot,
 * matlabData.C:33-53).
 */
#ifndef QD_H
#define QD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  QD_OK = 0,
  QD_ERR_INVALID = 1,   /* input validation failed (reported before any device allocation) */
  QD_ERR_DEVICE = 2,    /* a HIP runtime call failed */
  QD_ERR_NORM = 3,      /* total norm exceeded 1 (cudaOpenMP2.cu:437-440) */
  QD_ERR_STATE = 4,     /* call not valid in the handle's current state */
  QD_ERR_NOMEM = 5,     /* device allocation failed */
  QD_ERR_CAPACITY = 6   /* request exceeds a stated capacity (server admission) */
} qd_status;

enum { QD_PREC_FP64 = 0, QD_PREC_MIXED = 1 };

/* Observable record layout: one record per step and omega block. */
enum {
  QD_OBS_NORM = 0,      /* sum_k w_k sum|psi|^2 dr1 dr2            (omegawavepacket.cu:76-99)  */
  QD_OBS_POTENTIAL = 1, /* sum_k w_k <psi_k|V_k|psi_k> dr1 dr2     (omegawavepacket.cu:101-132) */
  QD_OBS_KINETIC = 2,   /* k-space <T> after the kinetic phase     (omegawavepacket.cu:364-395) */
  QD_OBS_ROTATIONAL = 3,/* sum_l l(l+1)<I^-1>                      (omegawavepacket.cu:397-428) */
  QD_OBS_CORIOLIS = 4,  /* pre-update Coriolis energy, 2nd half    (omegawavepacket.cu:503-553) */
  QD_OBS_FIELDS = 5
};

/* RadialCoordinate: matlabStructures.h:12-33, parsed at matlabStructures.C:17-23.
 * Grid points are left + i*dr, i = 0..n-1; the FFT period is n*dr. */
typedef struct {
  int32_t n;
  double left;
  double dr;
  double mass;
} qd_radial;

/* AngleCoordinate: matlabStructures.h:35-55, parsed at matlabStructures.C:25-37.
 * associated_legendre ((m+1) x n) is parsed but unused by the propagation
 * path in the reference; it is accepted and ignored here too (may be NULL). */
typedef struct {
  int32_t n;
  int32_t m;
  const double *x;                    /* [n] quadrature nodes cos(theta_k) */
  const double *w;                    /* [n] quadrature weights */
  const double *associated_legendre;  /* optional, unused */
} qd_angle;

/* EvolutionTime: matlabStructures.h:57-74, matlabStructures.C:39-44.
 * *steps is read at create and written back in place after every step
 * (cudaOpenMP2.cu:386).  steps == 0 triggers the extra V(-dt/2) half step
 * (cudaOpenMP2.cu:344), so resuming with steps != 0 continues the
 * leapfrogged potential sequence exactly. */
typedef struct {
  int32_t total_steps;
  double time_step;
  int32_t *steps;
} qd_time;

/* Replaces mexCallMATLAB(wave_to_matlab) (wave2matlab.C:17-26): called after
 * the wavepackets were copied into omega_states.wave_packets. */
typedef void (*qd_wave_callback)(void *user, int32_t steps);

/* Options: matlabStructures.h:76-97, matlabStructures.C:46-58.
 * The copy-back fires when (wave_to_matlab != NULL || wave_callback != NULL)
 * and steps % steps_to_copy_psi_from_device_to_host == 0, tested before the
 * step counter is incremented (cudaOpenMP2.cu:380-386). use_p2p_async selects
 * a multi-GPU exchange variant in the reference; all omega blocks share one
 * device here, so it is accepted and ignored. */
typedef struct {
  const char *wave_to_matlab;
  const char *test_name;
  int32_t steps_to_copy_psi_from_device_to_host;
  int32_t use_p2p_async;
  qd_wave_callback wave_callback;     /* extension: host callback, may be NULL */
  void *callback_user;
} qd_options;

/* DumpFunction: matlabStructures.h:99-110, matlabStructures.C:60-72.
 * The absorber is applied only when both dump1 and dump2 are given
 * (matlabData.h:27). dump1 has r1.n entries, dump2 has r2.n. */
typedef struct {
  const double *dump;
} qd_dump;

/* CummulativeReactionProbabilities: matlabStructures.h:112-135,
 * matlabStructures.C:74-90.  The reference parses this struct and never
 * uses it; here CRP[e] is filled when calculate_CRP != 0:
 *   P(E) = Im( sum_omega sum_k w_k sum_i conj(A) dA/dr dR ) / (2 pi mu_r eta_sq(E))
 *   A(R_i, theta_k) = sum_n exp(i E t_n) psi(R_i, r_s, theta_k, t_n) dt
 * with r_s = r2.left + n_dividing_surface*r2.dr, dA/dr from an
 * n_gradient_points central difference, t_n = (steps_n + 1) * dt. */
typedef struct {
  const double *energies;   /* [n_energies] total energies (hartree) */
  const double *eta_sq;     /* [n_energies] |<phi_E|psi_0>|^2 of the initial translational packet */
  double *CRP;              /* [n_energies] out */
  int32_t n_dividing_surface;
  int32_t n_gradient_points;
  int32_t n_energies;
  int32_t calculate_CRP;
} qd_crp;

/* OmegaStates: matlabStructures.h:137-167, matlabStructures.C:92-153.
 * omegas must be consecutive integers starting at omegas[0].
 * associated_legendres is (ntheta, legendre_n_max, n_omegas) with
 * legendre_n_max = l_max - omegas[0] + 1; block b holds P_l^{omega_b}(x_k) for
 * l = omega_b..l_max in its first l_max - omega_b + 1 columns
 * (matlabStructures.C:121-131, FH2main.m:43-50).
 * wave_packets is complex (n1, n2, ntheta, n_omegas), interleaved, in/out. */
typedef struct {
  int32_t J;
  int32_t parity;
  int32_t l_max;
  int32_t n_omegas_max;
  int32_t n_omegas;
  const int32_t *omegas;
  int32_t legendre_n_max;
  const double *associated_legendres;
  double *wave_packets;
} qd_omega_states;

/* The whole input struct: the fields of FH2Data (FH2main.m:195-206) plus
 * extensions that have no counterpart in the reference. */
typedef struct {
  qd_radial r1;
  qd_radial r2;
  qd_angle theta;
  const double *pot;              /* (n1, n2, ntheta) */
  qd_time time;
  qd_options options;
  qd_dump dump1;
  qd_dump dump2;
  qd_crp crp;
  qd_omega_states omega_states;
  /* extensions */
  int32_t k_max;          /* helicity truncation: Coriolis matrices are cut to
                             omega <= min(J, l, k_max) before exponentiation;
                             k_max < 0 keeps the reference (untruncated) form */
  int32_t precision;      /* QD_PREC_FP64 or QD_PREC_MIXED */
  int32_t print_energies; /* print per-step lines in the reference's format */
  int32_t final_copy;     /* qd_evolve: copy the final state into wave_packets */
  double *observables;    /* qd_evolve: optional out [total_steps][n_omegas][QD_OBS_FIELDS] */
  double *autocorrelation;/* optional out, complex <psi(t_0)|psi(t_n)> per step n, summed over
                             the omega blocks with the quadrature weights and dr1 dr2.
                             psi(t_n) = exp(-i dt/2 V) psi_stored(n), the stored state with
                             the leapfrogged V(dt/2) restored, at t_n = (steps_n + 1) dt;
                             psi(t_0) is the caller's wave_packets on a fresh start
                             (steps == 0) and exp(-i dt/2 V) wave_packets on a resume.
                             When NULL at create the autocorrelation is not computed
                             (saves one read of the initial coefficients per step) */
} qd_inputs;

typedef struct qd_handle qd_handle;

/* One-shot drop-in for mexFunction: create, run time.total_steps steps,
 * fill the outputs, destroy.  On failure the message is in
 * qd_last_global_error(). */
qd_status qd_evolve(qd_inputs *in);

/* Handle API (the state machine behind qd_evolve). */
qd_status qd_create(const qd_inputs *in, qd_handle **out);
qd_status qd_run(qd_handle *h, int32_t n_steps);
qd_status qd_synchronize(qd_handle *h);
/* Copy the current DVR wavepackets, complex (n1, n2, ntheta, n_omegas). */
qd_status qd_copy_wavepackets(qd_handle *h, double *dst);
/* Per-step observables recorded so far: dst[rec][omega][QD_OBS_FIELDS],
 * records first..first+count-1 (record 0 = first step run by this handle). */
qd_status qd_get_observables(qd_handle *h, int32_t first, int32_t count, double *dst);
/* Complex autocorrelation <psi(t_0)|psi(t_n)> (defined at qd_inputs.autocorrelation)
 * for records first..first+count-1. */
qd_status qd_get_autocorrelation(qd_handle *h, int32_t first, int32_t count, double *dst);
/* Reaction probabilities at the CRP energies from the flux accumulated so far. */
qd_status qd_get_crp(qd_handle *h, double *crp);
int32_t qd_steps_done(const qd_handle *h);
void qd_destroy(qd_handle *h);
const char *qd_last_error(const qd_handle *h);
const char *qd_last_global_error(void);

/* Device-side timing of the kernels of one step, for measurement tools:
 * runs n_steps steps with HIP events around each launch on the handle's
 * stream and returns the mean microseconds per launch of each kernel kind
 * in us[QD_KERNEL_KINDS]; kinds fused into another kernel read 0.  The steps
 * advance the state like qd_run. */
enum {
  QD_K_DVR = 0,      /* absorber and potential phase: fused into QD_K_LEG_FWD      */
  QD_K_LEG_FWD = 1,  /* forward Legendre transform (with the DVR phase)            */
  QD_K_FBR1 = 2,     /* first rotational-Coriolis-rotational half step: fused into QD_K_FFT_R_FWD
                        (into QD_K_FFT_r in the two-pass FFT form, qd_probe_fft_passes) */
  QD_K_FFT_R_FWD = 3,/* forward FFT along R (reads 0 in the two-pass FFT form)        */
  QD_K_FFT_r = 4,    /* FFT along r, kinetic phase, inverse FFT (two-pass form: the first
                        half step and the r factor of the kinetic phase)            */
  QD_K_FFT_R_INV = 5,/* inverse FFT along R, 1/(n1 n2) scale, second half step (two-pass
                        form: forward FFT along R and the R factor before them)     */
  QD_K_FBR2 = 6,     /* second half step: fused into QD_K_FFT_R_INV                */
  QD_K_LEG_BWD = 7,  /* backward Legendre transform (with the potential energy)    */
  QD_K_OBS = 8,      /* flux surface values (when CRP is requested; in the fused step
                        sequence carried by the first FFT pass: QD_K_FFT_R_FWD, or
                        QD_K_FFT_r in the two-pass form)                            */
  QD_K_REDUCE = 9,   /* final reduction of the per-step observables (where the fused
                        Legendre step runs, part of the first FFT pass likewise)    */
  QD_K_LEG_STEP = 10,/* fused Legendre step: backward transform of the previous step,
                        potential energy, DVR phase and forward transform of this one;
                        where it runs, QD_K_LEG_FWD and QD_K_LEG_BWD read 0           */
  QD_KERNEL_KINDS = 11
};
qd_status qd_profile_steps(qd_handle *h, int32_t n_steps, double *us);

/* ---------------------------------------------------------------------------
 * Serving form (no counterpart in the reference, which runs one partial wave
 * per process).  A server keeps many partial-wave jobs resident on the current
 * device and advances every running job by one step per qd_server_step:
 *   - qd_server_submit validates a job and queues it (nothing is allocated);
 *   - each qd_server_step first admits up to four queued jobs whose
 *     estimated device footprint fits the byte budget (backfilling past jobs
 *     that do not fit; a few per step, so that the running classes keep the
 *     GPU busy while the host builds the next jobs),
 *     then runs one step of every running job (jobs of one grid shape and
 *     precision share each kernel launch and a HIP graph; shape classes run on
 *     their own streams), then retires jobs that reached time.total_steps:
 *     their outputs (time.steps, wave_packets with final_copy, observables,
 *     autocorrelation, CRP.CRP) are written into their qd_inputs and their
 *     device memory is freed for the next admissions.
 * A job's results do not depend on the other jobs in the mix: FP64 results
 * are bitwise those of qd_evolve on the same inputs.
 * The inputs of a submitted job are borrowed until the job is DONE or FAILED.
 * Per-step copy-back (wave_to_matlab / wave_callback) is a qd_run feature and
 * is rejected here.
 * ------------------------------------------------------------------------- */
typedef struct qd_server qd_server;
enum { QD_JOB_QUEUED = 0, QD_JOB_RUNNING = 1, QD_JOB_DONE = 2, QD_JOB_FAILED = 3 };

/* hbm_budget_bytes <= 0: 90% of the device memory free at creation;
 * max_running <= 0: no limit besides the budget. */
qd_status qd_server_create(int64_t hbm_budget_bytes, int32_t max_running, qd_server **out);
qd_status qd_server_submit(qd_server *s, qd_inputs *in, int32_t *job_id);
/* admit, advance every running job by one step, retire finished jobs;
 * *n_running / *n_queued (may be NULL) report the state after the call */
qd_status qd_server_step(qd_server *s, int32_t *n_running, int32_t *n_queued);
/* qd_server_step that advances each shape class by up to max_steps steps:
 * by 8 (replayed as one graph, saving the launch gaps between steps) when
 * max_steps >= 8 and none of the class's jobs reaches time.total_steps or
 * fills its 64-step flux window within them, otherwise by one.  Results are
 * bitwise those of single steps; admission and retirement happen between
 * calls, so larger steps trade arrival latency for throughput. */
qd_status qd_server_advance(qd_server *s, int32_t max_steps, int32_t *n_running, int32_t *n_queued);
/* state (QD_JOB_*), steps done, seconds from submit to retirement (or to now),
 * seconds from submit to admission (-1 while queued) */
qd_status qd_server_job_status(qd_server *s, int32_t job_id, int32_t *state, int32_t *steps_done,
                               double *latency_s, double *queued_s);
/* message of a FAILED job */
const char *qd_server_job_error(qd_server *s, int32_t job_id);
/* bytes of the budget in use by running jobs (estimates) and the budget */
qd_status qd_server_memory(qd_server *s, int64_t *used, int64_t *budget);
void qd_server_destroy(qd_server *s);
const char *qd_server_last_error(const qd_server *s);

/* Library version string. */
const char *qd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QD_H */
