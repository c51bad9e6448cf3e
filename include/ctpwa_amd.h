/*
 * ctpwa_amd.h - C ABI of the MI355X likelihood library (libctpwa_amd.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses on the
 * device the handle was created on; `stream` is a hipStream_t (NULL = default
 * stream).  Every call returns 0 on success and a non-zero status otherwise;
 * ctpwa_last_error() gives the message of the last failure on the calling thread.
 *
 * The entry points replace the reference's pybind11 `analysis` boundary
 * (src/main.cu:90-131) and the device work behind it:
 *
 *   ctpwa_create       <- analysis::analysis(config) / initialize()
 *                         (src/Analysis.cu:1271-1281, 3849-4208): amplitude plan,
 *                         events, SL amplitudes (src/AmpGen.cu:655-908)
 *   ctpwa_dims         <- analysis::getNVector / getNFreeTheta / getNParams /
 *                         getNPolarizations (src/Analysis.cu:1312-1316)
 *   ctpwa_param_info   <- analysis::getFreeResParams (src/AmpGen.cu:2193-2197)
 *   ctpwa_nll_grad     <- NLLFunction::forward + backward (src/Analysis.cu:625-1261)
 *   ctpwa_hessian      <- analysis::getHessian (src/Analysis.cu:2689-3378)
 *   ctpwa_intensity    <- per-event |sum_a v_a A_a|^2 summed over polarisations, the
 *                         quantity behind writeResult's weights
 *                         (src/ComputeResults.cu:22-182)
 *   ctpwa_amplitudes   <- analysis::getDataTensor/getPhspTensor-style amplitude
 *                         access (src/Analysis.cu:3640-3700)
 *   ctpwa_footprint_bytes <- DeviceManager::estimate (src/DeviceManager.cu:172-221)
 *   ctpwa_hessian_workspace_bytes <- the part of that estimate getHessian's buffers take
 *   ctpwa_amplitude_gram <- the integrals of getFitFractions / writeResult's interference
 *   ctpwa_partial_intensities <- writeResult's per-event total and partial weights
 *   ctpwa_partial_amplitudes  <- writeResult's subset totals and pair interference terms
 */
#ifndef CTPWA_AMD_H
#define CTPWA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctpwa_handle ctpwa_handle;

/* event classes */
#define CTPWA_PHSP 0
#define CTPWA_DATA 1
#define CTPWA_BKG 2

/* storage precision of the per-event inputs (the arithmetic is FP64 in both):
 * CTPWA_FP64   SL tensors and invariants in FP64 (the reference's double build)
 * CTPWA_MIXED  SL tensors and invariants in FP32, squared masses in FP64; half the
 *              HBM bytes per evaluation (the reference's default float build stores its
 *              amplitudes in FP32 as well, include/ComplexType.h:27-32) */
#define CTPWA_FP64 0
#define CTPWA_MIXED 1

/* Build the likelihood for `plan_text` (written by ctpwa_amd.model.Plan.to_text).
 * Events are host arrays [n][nfinal][4] of float64 (E, px, py, pz) in Data.order;
 * weights are host arrays [n] or NULL for unit weights.  n may be 0 for bkg.
 * The library copies what it needs; the host arrays may be freed on return. */
int ctpwa_create(const char* plan_text,
                 const double* phsp, long long n_phsp, const double* w_phsp,
                 const double* data, long long n_data, const double* w_data,
                 const double* bkg, long long n_bkg, const double* w_bkg,
                 int precision, int device, ctpwa_handle** out);

void ctpwa_destroy(ctpwa_handle* h);

/* n_amp: amplitudes; n_free: free complex couplings (getNVector); n_theta: free
 * resonance parameters (getNFreeTheta); n_pol: polarisation states. */
int ctpwa_dims(const ctpwa_handle* h, int* n_amp, int* n_free, int* n_theta, int* n_pol);

/* initial value, lower and upper bound of each free resonance parameter (host [n_theta]) */
int ctpwa_param_info(const ctpwa_handle* h, double* init, double* lower, double* upper);

/* NLL and gradient.  params_dev: device [2 n_free + n_theta] float64 laid out
 * [Re p, Im p, theta]; out_dev: device [1 + 2 n_free + n_theta] = [NLL, gradient]. */
int ctpwa_nll_grad(ctpwa_handle* h, const double* params_dev, double* out_dev, void* stream);

/* ctpwa_nll_grad that allocates nothing: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  The evaluation's per-workgroup partial sums and class
 * totals live in a library-owned static device array of 2 MiB per GPU (held until the
 * process ends) instead of the handle's buffers; the result is bit-identical to
 * ctpwa_nll_grad.  A model whose partials exceed the array (n_acc x n_workgroups +
 * 3 n_acc > 262144 doubles, n_acc = 1 + 2 n_amp + theta-gradient terms, at most 1024
 * workgroups) is rejected with an error.  All calls on one GPU share the
 * array: calls that overlap on one GPU race. */
int ctpwa_nll_grad_noalloc(ctpwa_handle* h, const double* params_dev, double* out_dev, void* stream);

/* Hessian in the same layout: device [(2 n_free + n_theta)^2], row-major. */
int ctpwa_hessian(ctpwa_handle* h, const double* params_dev, double* hess_dev, void* stream);

/* Many likelihoods of one model (same plan and precision; e.g. the toy fits of an
 * ensemble) evaluated together: one batched prep launch, one grouped per-event launch
 * whose workgroups are shared out among the likelihoods, one batched reduction.  Each
 * likelihood's result is bit-identical to ctpwa_nll_grad on its own (same partition of
 * its events, same summation order).  Replaces the per-fit NLLFunction::forward/backward
 * calls (src/Analysis.cu:625-1261) when a service runs many fits at once; the reference
 * has no batched form.  A batch object owns pinned and device descriptor space for up to
 * `max_jobs` likelihoods; calls on one batch object must not overlap.  Each likelihood
 * of a call has a handle of its own (a handle holds one coefficient table and one set of
 * partial sums); a handle named twice fails the call. */
typedef struct ctpwa_batch ctpwa_batch;
int ctpwa_batch_create(int device, int max_jobs, ctpwa_batch** out);
void ctpwa_batch_destroy(ctpwa_batch* b);
int ctpwa_batch_nll_grad(ctpwa_batch* b, ctpwa_handle* const* handles, const double* const* params_dev,
                         double* const* out_dev, int n, void* stream);

/* per-event intensity of class `cls`: device [n_cls] */
int ctpwa_intensity(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream);

/* per-event amplitudes of class `cls` (without couplings): device [n_cls][n_pol][n_amp][2] */
int ctpwa_amplitudes(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream);

/* amplitude Gram of class `cls`: device [n_amp][n_amp][2] complex
 * Q[l][k] = sum_events sum_pol conj(A_l) A_k of the amplitudes without couplings (the
 * partial-wave and interference integrals of computeResults and getFitFractions,
 * src/ComputeResults.cu:57-190, src/ComputeBF.cu, src/Analysis.cu:3382-3420,3527-3638);
 * fails for models with Custom lineshapes */
int ctpwa_amplitude_gram(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream);

/* per-event partial-wave intensities of class `cls`: device [n_wave + 1][n_cls];
 * row w < n_wave: sum_pol |P_w|^2 with P_w the coherent sum of partial wave w's
 * amplitudes times couplings, row n_wave: the total intensity (writeResult's per-event
 * weights before normalisation, src/ComputeResults.cu:22-190, src/Analysis.cu:1460-1700).
 * A partial wave is one resonance set of one chain (a plan block) together with its
 * identical-particle permuted copies; n_wave is the number of distinct values of the
 * plan's `bwave` record, or the block count when the plan has none.
 * Fails for models with Custom lineshapes */
int ctpwa_partial_intensities(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream);

/* theta (the last ntheta parameters) of the handle's most recent ctpwa_nll_grad,
 * ctpwa_nll_grad_noalloc, batched evaluation or ctpwa_hessian (the per-event outputs and
 * Grams leave it), copied to `theta_host` [ntheta] after the work queued on `stream`:
 * the resonance parameters at which the reference's stored amplitudes stand after
 * getNLL / getHessian / writeResult (src/AmpGen.cu:2344-2726), for the post-fit methods
 * that use them.  Undefined before the first evaluation */
int ctpwa_last_theta(ctpwa_handle* h, double* theta_host, void* stream);

/* per-event coherent partial-wave amplitudes of class `cls` (couplings applied): device
 * [n_wave][n_cls][n_pol][2], P_w summed over the wave's blocks as above.  The per-event
 * quantities of writeResult that need phases (the total over a subset of waves and the
 * pair interference terms 2 Re(P_i conj P_j), src/ComputeResults.cu:49-185,
 * src/Analysis.cu:1355-1700) follow from them.  Fails for models with Custom lineshapes */
int ctpwa_partial_amplitudes(ctpwa_handle* h, const double* params_dev, int cls, double* out_dev, void* stream);

/* average duration in ms of the fused per-event kernel over `iters` (<= 10000) launches
 * at `params_dev`: a HIP event pair on `stream` around every launch; the prep kernel (and
 * the Custom factor kernel) run once before */
int ctpwa_time_eval_kernel(ctpwa_handle* h, const double* params_dev, int iters, void* stream, double* avg_ms);

/* device bytes a handle for this plan, these event counts and this precision holds (the
 * upload staging, 2 x 32 MiB pinned host and device buffers plus the mixed form's FP64
 * chunk, is allocated once per device for the process and shared by all handles) */
long long ctpwa_footprint_bytes(const char* plan_text, long long n_phsp, long long n_data, long long n_bkg,
                                int precision);

/* device bytes a handle allocates on its first ctpwa_hessian call (the per-chunk
 * Jacobian rows and partial sums it keeps until ctpwa_destroy).  A handle that is alone
 * in the process, with at least 2^20 phase-space and 2^17 data + background events, runs
 * data and background on a second stream with a second buffer set; the estimate counts
 * that set when no handle is live at the time of this call */
long long ctpwa_hessian_workspace_bytes(const char* plan_text, long long n_phsp, long long n_data, long long n_bkg);

/* number of events of a class, and algorithmic bytes the fused kernel must read per
 * evaluation (SL tensors + invariants + per-event weights), its workgroup count, and the
 * polarisations stored per SL tensor (1 when the two polarisations of every SL tensor are
 * mirror images, found at ctpwa_create; else the model's count).  With
 * no free resonance parameter the phase-space term of every evaluation is the quadratic
 * form of the phase-space amplitude Gram built at ctpwa_create (the reference's stored
 * phase-space matrix, src/Analysis.cu:808-826, src/ComputeNLL.cu:575-597): the fused
 * kernel then reads the data and background events only, and so do these counts */
int ctpwa_layout(const ctpwa_handle* h, long long* n_events3, long long* bytes_per_eval,
                 int* n_sl_tensors, int* n_invariants, int* n_workgroups, int* n_pol_stored);

/* generated source of the fused kernel (NUL-terminated, owned by the handle) */
const char* ctpwa_kernel_source(const ctpwa_handle* h);

/* source of the fused kernel for a plan, without a device: writes at most cap bytes
 * (NUL-terminated) into buf and returns the full size including the NUL, -1 on error */
long long ctpwa_generate_source(const char* plan_text, int precision, char* buf, long long cap);

/* source of the generated Hessian rows kernel for a plan (the per-event part of
 * analysis::getHessian, src/Analysis.cu:2689-3378), without a device; as
 * ctpwa_generate_source.  For a model with Custom lineshapes the source of the
 * second-order factor module (ctpwa_custom2) follows a line
 * "// ---- module: ctpwa_custom2 ----" */
long long ctpwa_generate_hessian_source(const char* plan_text, int precision, char* buf, long long cap);

/* compile the fused kernel and the Hessian rows kernel for a plan with hipRTC (no device
 * needed), and for a plan with Custom lineshapes also the Hessian's second-order factor
 * kernel; returns the total code-object size, or -1 with the compiler log in
 * ctpwa_last_error() */
long long ctpwa_compile_check(const char* plan_text, int precision);

const char* ctpwa_last_error(void);
const char* ctpwa_version(void);

#ifdef __cplusplus
}
#endif
#endif
