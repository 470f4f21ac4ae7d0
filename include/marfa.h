/* marfa.h -- C ABI of the MI355X line-by-line (LBL) absorption core.
 *
 * Plain C types only (pointers and sizes; no torch or HIP types), so that the Fortran
 * host binds it with ISO_C_BINDING (marfa_amd/fortran/marfa_capi.f90) and Python with
 * ctypes (marfa_amd/lbl.py).  Every function returns 0 on success or a positive error
 * code; none of them calls exit().  Error codes 1-9 keep the meaning of the reference's
 * `stop` codes where one exists (9 = line-data/window error, src/LBL.f90:50-53); codes
 * from 20 on are this library's own (see MARFA_E_* below).
 *
 * Host buffers are owned by the caller and copied in; device buffers are owned by the
 * context.  A context is thread-compatible: calls on one context must not overlap.
 *
 * Reference interfaces replaced (Razumovskyy/MARFA):
 *   marfa_lbl_scheme    <- subroutine lblScheme(lineIdxParameter, capWVParameter)
 *                          (src/LBL.f90:25-100), bracketed by resetAbsorptionGridValues
 *                          and cascadeInterpolation (src/Grids.f90:64-139) exactly as the
 *                          only caller does (app/main.f90:148-160)
 *   marfa_set_lines     <- the 36-byte direct-access line records read at src/LBL.f90:43
 *                          (layout app/processParFile.f90:176-177)
 *   marfa_set_tips      <- TIPS(74,501) filled by readTIPS (src/EnvironmentSetup.f90:230-245)
 *   job.shape_id        <- shapeFuncPtr / setShapeFunction (src/Shapes.f90:8-15), the
 *                          `shape` plugin interface (src/Interfaces.f90:6-12)
 *   job.chi_id          <- chiFactorFuncPtr / setChiFactorFunction (src/ChiFactors.f90:7-16),
 *                          the `chifactor` plugin interface (src/Interfaces.f90:15-23)
 *   marfa_pt_tables     <- ATMOSPHERIC_LEVELS_LOOP x SUBINTERVALS_LOOP (app/main.f90:81-180)
 *                          producing every level's PT-table records at once
 *   marfa_windows       <- determineStartingSpectralLine (app/main.f90:196-248) and the
 *                          lineIdx / capWV carry (app/main.f90:133-144, src/LBL.f90:57-61)
 */
#ifndef MARFA_H
#define MARFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MARFA_NT 20481           /* points per PT-table record, src/Grids.f90:16 */
#define MARFA_RECORD_BYTES 81924 /* NT * 4, app/main.f90:111 */

/* precision forms */
#define MARFA_FP64 0  /* plain drop-in: the reference's algorithm in FP64 */
#define MARFA_FP32K 1 /* FP32 profiles, compensated FP32 or FP64 sums, FP64 offsets */

/* shapes (src/Shapes.f90) and chi factors (src/ChiFactors.f90) */
#define MARFA_SHAPE_VOIGT 0    /* voigt, Shapes.f90:22 */
#define MARFA_SHAPE_LORENTZ 1  /* lorentzShape (chi-corrected Lorentz), Shapes.f90:135 */
#define MARFA_CHI_NONE 0
#define MARFA_CHI_TONKOV 1
#define MARFA_CHI_POLLACK 2
#define MARFA_CHI_PERRIN 3

/* targets, app/main.f90:163-168 */
#define MARFA_ACS 0
#define MARFA_VAC 1

/* error codes */
#define MARFA_OK 0
#define MARFA_E_SPECIES 2     /* unsupported molecule (app/processParFile.f90:128-131) */
#define MARFA_E_PARREAD 3     /* .par field Fortran's read rejects (app/processParFile.f90:154-157) */
#define MARFA_E_WINDOW 9      /* line index outside the line list (src/LBL.f90:50-53) */
#define MARFA_E_ARG 20        /* invalid argument */
#define MARFA_E_STATE 21      /* lines / TIPS / levels not set */
#define MARFA_E_HIP 22        /* HIP runtime error (message: marfa_last_error) */
#define MARFA_E_NODEVICE 23   /* no HIP device */
#define MARFA_E_TEMPERATURE 24 /* T outside the TIPS table (src/Spectroscopy.f90:110-114) */

typedef struct marfa_ctx marfa_ctx;

typedef struct {
    double start_wv;  /* startWV [cm-1] */
    double end_wv;    /* endWV [cm-1] */
    float cutoff;     /* cutOff [cm-1] */
    int32_t shape_id; /* MARFA_SHAPE_* */
    int32_t chi_id;   /* MARFA_CHI_* */
    int32_t mol_code; /* moleculeIntCode (1 H2O, 2 CO2, ...), getSpeciesCode */
    int32_t target;   /* MARFA_ACS or MARFA_VAC */
    int32_t precision;/* MARFA_FP64 or MARFA_FP32K */
} marfa_job;

/* ---- context ---- */
int marfa_ctx_create(int device, marfa_ctx **out);
int marfa_ctx_destroy(marfa_ctx *ctx);
const char *marfa_last_error(const marfa_ctx *ctx);
int marfa_version(void);

/* ---- inputs (copied to device) ---- */
int marfa_set_tips(marfa_ctx *ctx, int32_t niso, int32_t ntemp, const float *tips /* [niso][ntemp] */);
int marfa_set_lines(marfa_ctx *ctx, int64_t n, const double *nu, const float *S,
                    const float *gair, const float *gself, const float *elow, const float *nT,
                    const float *delta, const int32_t *molIso);
/* levels: total pressure [atm], temperature [K], number density [1/(cm^2 km)] */
int marfa_set_levels(marfa_ctx *ctx, int32_t nlev, const float *P, const float *T,
                     const float *rho);

/* ---- windows (host-side index logic) ----
 * Number of 10 cm-1 subintervals of the job; lo/hi (0-based line ranges [lo,hi)),
 * recnum (1-based PT-table record numbers) and origin (subinterval starts) are filled
 * when non-NULL and max_rec is large enough. */
int marfa_windows(marfa_ctx *ctx, const marfa_job *job, int32_t max_rec, int32_t *nrec,
                  int64_t *lo, int64_t *hi, int32_t *recnum, double *origin);

/* ---- drop-in for one call of lblScheme between reset and cascade ----
 * line_idx: 1-based record index of the first line to read (in/out, carried to the
 * next subinterval as src/LBL.f90:61 does); cap_wv, start_dwv and ext_end_dwv are
 * capWV, startDeltaWV and extEndDeltaWV; P, T, pSelf the level state.  rk receives the
 * 20,481-point record (ACS).  Runs synchronously on the context's stream. */
int marfa_lbl_scheme(marfa_ctx *ctx, int64_t *line_idx, double cap_wv, double start_dwv,
                     double ext_end_dwv, float cutoff, float P, float T, float pSelf,
                     int32_t shape_id, int32_t chi_id, int32_t mol_code, int32_t precision,
                     float *rk);

/* ---- all levels x subintervals of one job ----
 * Records of level l, subinterval s land at out[(l*nrec + s)*MARFA_NT].
 * out_dev: device pointer (may be NULL) -- results stay in HBM;
 * out_host: host pointer (may be NULL) -- results copied back.
 * stream: hipStream_t to run on (NULL = the context's stream).  Synchronous unless
 * out_host is NULL and async != 0; an asynchronous call records no phase timings. */
int marfa_pt_tables(marfa_ctx *ctx, const marfa_job *job, float *out_dev, float *out_host,
                    void *stream, int32_t async);

/* marfa_pt_tables without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The caller passes out_dev (device, the layout of
 * marfa_pt_tables); the inputs are the context's (set beforehand).  The kernels' scratch is
 * a fixed array of MARFA_NOALLOC_SCRATCH_BYTES in the library, one per GPU, held until the
 * process ends; levels run in batches that fit it, and a job whose single level needs more
 * is rejected with MARFA_E_ARG (message: marfa_last_error).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads) race; calls on
 * one stream are ordered.  The call enqueues its work on stream and returns (no timings). */
#define MARFA_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int marfa_pt_tables_noalloc(marfa_ctx *ctx, const marfa_job *job, float *out_dev, void *stream);

/* marfa_pt_tables with FP64 records (no counterpart in the reference, whose records are
 * FP32): out_dev (device) receives the records before their rounding to FP32, for
 * precision studies and finite-difference checks of the adjoint.  Synchronous. */
int marfa_pt_tables_f64(marfa_ctx *ctx, const marfa_job *job, double *out_dev, void *stream);

/* ---- adjoint of marfa_pt_tables (no counterpart in the reference) ----
 * grad_out: dL/d(record point) on the device, in the layout of marfa_pt_tables' output.
 * grad_lines (host, [4][n] for the context's n lines): dL/dS, dL/dgamma_air, dL/dn_T,
 * dL/ddelta; grad_levels (host, [2][nlev]): dL/dT, dL/dP per level (density held fixed,
 * pSelf = rho 10 k T / p0 follows T).  The deposit geometry (which nodes each line samples)
 * is held fixed: the derivative is that of the piecewise-smooth model between its
 * decisions.  The level-0 sums kept by the context's last forward of the same job, inputs
 * and stream are reused (and the cascade's limiter choices, when that forward kept them:
 * marfa_keep_adjoint_state); otherwise they are recomputed.  The Perrin chi factor's
 * temperature dependence enters dL/dT. */
int marfa_pt_tables_adjoint(marfa_ctx *ctx, const marfa_job *job, const float *grad_out, void *stream,
                            double *grad_lines, double *grad_levels);
/* marfa_pt_tables_adjoint with the gradients in device memory (grad_lines [4][n], grad_levels
 * [2][nlev], FP64), enqueued on stream without waiting: for callers whose tensors live on
 * the GPU (the PyTorch operators in marfa_amd/csrc/torch_ops.cpp). */
int marfa_pt_tables_adjoint_dev(marfa_ctx *ctx, const marfa_job *job, const float *grad_out, void *stream,
                                double *grad_lines, double *grad_levels);

/* on != 0: forwards of this context keep, beside the level-0 sums they always keep, the
 * cascade's limiter choices for a following adjoint of the same job (3.2 KB per record, about
 * 1 % of the forward's time), so that the adjoint does not recompute the finer grids; on == 0:
 * they do not.  It switches on by itself once an adjoint has run on the context.  No
 * counterpart in the reference (which has no adjoint); for differentiable callers
 * (marfa_amd/autograd.py switches it on). */
int marfa_keep_adjoint_state(marfa_ctx *ctx, int32_t on);

/* Replace the per-line S, gamma_air, n_T and delta of the context's line list (n values each,
 * device pointers) by device-to-device copies on stream: the centres, and with them the
 * windows, are unchanged, so nothing is re-allocated or re-checked on the host.  No
 * counterpart in the reference, which reads these fields from the line records
 * (src/LBL.f90:43); for differentiable callers that update them every step. */
int marfa_update_line_params(marfa_ctx *ctx, const float *S, const float *gair, const float *nT,
                             const float *delta, void *stream);

/* The spectral loss of the differentiable form, L = sum_i w_i (ln k_i - ln khat_i)^2 over the
 * points where k_i > 0 and khat_i > 0 (FP64 logarithms and sum, in a fixed order), and its
 * gradient dL/dk_i = 2 w_i (ln k_i - ln khat_i) / k_i (0 elsewhere).  No counterpart in the
 * reference: the loss of the gradient workloads, fused so that a backward pass reads the
 * records once.  k, khat: n FP32 values in device memory (w: FP32 weights, or NULL for unit ones);
 * every weight is further multiplied by wscale.  marfa_log_loss writes one double to
 * loss_dev, using partial_dev (MARFA_LOSS_PARTIALS doubles, device) for the block sums;
 * marfa_log_loss_grad writes n FP32 values to grad_dev, each times *gscale_dev (device, or
 * NULL for 1).  Both are enqueued on stream (NULL: the default stream) and allocate nothing. */
#define MARFA_LOSS_PARTIALS 2048
int marfa_log_loss(int64_t n, const float *k, const float *khat, const float *w, double wscale, double *partial_dev,
                   double *loss_dev, void *stream);
int marfa_log_loss_grad(int64_t n, const float *k, const float *khat, const float *w, double wscale,
                        const double *gscale_dev, float *grad_dev, void *stream);
/* the same for FP64 records (k, khat, grad_dev in FP64; weights FP32) */
int marfa_log_loss_f64(int64_t n, const double *k, const double *khat, const float *w, double wscale,
                       double *partial_dev, double *loss_dev, void *stream);
int marfa_log_loss_grad_f64(int64_t n, const double *k, const double *khat, const float *w, double wscale,
                            const double *gscale_dev, double *grad_dev, void *stream);

/* The context's line count, level count and device (any pointer may be NULL). */
int marfa_ctx_info(const marfa_ctx *ctx, int64_t *nlines, int32_t *nlev, int32_t *device);

/* ---- batches of independent jobs in one hipGraph (no counterpart in the reference, which
 * runs one job per process) ----
 * Each job has its own line list, levels, windows and records.  marfa_batch_tune assigns
 * the jobs to `lanes` parallel branches (largest first onto the least loaded lane), times
 * the fine-kernel variants (centre lines staged in LDS per tile: 512/1024/2048, XCD-aware
 * record order on/off) on each job's first level and keeps the fastest, then runs every
 * job once (eagerly, lane_budget_gb of per-level arrays per lane, 0 = 8 GB; the captured
 * form's wave arena takes at most 12 lane budgets, 96 GB and 70 % of the free memory).
 * marfa_batch_capture records all jobs into one hipGraph: the jobs' levels are cut into
 * pieces and packed into waves; a wave prepares its pieces on the lanes (spectroscopy and
 * sort, beside the previous wave's deposits), then runs the level-0 and the fine kernel
 * each once per kernel variant over a device work list of every record (record tile) of
 * every piece (one fine launch per stage capacity the jobs' tuning chose).
 * marfa_batch_launch replays the graph with one hipGraphLaunch. */
typedef struct marfa_batch marfa_batch;
int marfa_batch_create(int device, int32_t lanes, double lane_budget_gb, marfa_batch **out);
int marfa_batch_destroy(marfa_batch *b);
const char *marfa_batch_last_error(const marfa_batch *b);
int marfa_batch_set_tips(marfa_batch *b, int32_t niso, int32_t ntemp, const float *tips);
int marfa_batch_add_job(marfa_batch *b, const marfa_job *job, int64_t n, const double *nu, const float *S,
                        const float *gair, const float *gself, const float *elow, const float *nT,
                        const float *delta, const int32_t *molIso, int32_t nlev, const float *P,
                        const float *T, const float *rho, int32_t *job_id);
/* variants (optional, [jobs]): chosen variant per job; tune_ms (optional, [jobs][6]): the
 * fine kernel's time for every variant on the job's first level */
int marfa_batch_tune(marfa_batch *b, int32_t *variants, float *tune_ms);
int marfa_batch_capture(marfa_batch *b);
/* stream NULL: lane 0's stream.  sync != 0: wait and record the launch's duration */
int marfa_batch_launch(marfa_batch *b, void *stream, int32_t sync);
/* job < 0: only last_ms (duration of the last synchronous launch) */
int marfa_batch_info(marfa_batch *b, int32_t job, int32_t *nlev, int32_t *nrec, float **records_dev,
                     float *last_ms);
int marfa_batch_records(marfa_batch *b, int32_t job, float *out_host);

/* Work counts of the job, summed over levels and subintervals (computed on the device):
 * samples[0] = exact count of distinct profile evaluations (samples) of the reference's
 * grid geometry (21 + 2m per wing line, the walks and core points of a centre line);
 * samples[1] = the part the level-0 deposit kernel evaluates: its grid-0 evaluations (the 21
 * half-integer positions and RK(1)) and the samples of grids 1..5; the rest (grids 6..9 and
 * the fine points) belongs to the fine-grid kernel. */
int marfa_count_samples(marfa_ctx *ctx, const marfa_job *job, int64_t samples[2]);

/* Pre-cascade grid sums of one record (level, subinterval index), in the reference's
 * layout: for grid k = 0..9 the arrays RKkP, RKk, RKkL (src/Grids.f90:34-45) of 10*2^k
 * cells each, then RK(1..20481) holding the direct fine-point deposits (30,690 + 20,481
 * doubles); record (optional) receives the finished record.  For tests of the grid
 * selection against a CPU restatement. */
int marfa_grid_sums(marfa_ctx *ctx, const marfa_job *job, int32_t level, int32_t rec,
                    double *grids, float *record);

/* ---- line-list ingest (host only, no device needed) ----
 * marfa_species_code <- getSpeciesCode (app/processParFile.f90:189-223): molecule name
 *   (blanks trimmed) -> two-digit HITRAN code "01".."12" in code[3]; MARFA_E_SPECIES and
 *   "00" for any other name.
 * marfa_par_convert  <- the program app/processParFile.f90:32-185: HITRAN .par text ->
 *   36-byte direct-access records (f8 nu; f4 S, gamma_air, gamma_self, E'', n_T; i4
 *   jointMolIso; f4 delta) in db_path, record r = input line r (a line with an unknown
 *   isotopologue leaves its record unwritten, a hole of zeros; no record follows the
 *   last line, as in the reference's fixture test/testDB.02).  *nrec: records in the file.
 *   MARFA_E_SPECIES (2) / MARFA_E_PARREAD (3) as the reference's stop 2 / stop 3;
 *   MARFA_E_ARG when a file cannot be opened or written. */
int marfa_species_code(const char *species, char code[3]);
int marfa_par_convert(const char *species, const char *par_path, const char *db_path, int64_t *nrec);

/* average duration [ms] of the last marfa_pt_tables call's kernels, measured with HIP
 * events on the stream they ran on: [0] spectroscopy, [1] sort, [2] coarse deposit,
 * [3] fine deposit + cascade, [4] total. */
int marfa_last_timings(const marfa_ctx *ctx, float *ms5);

#ifdef __cplusplus
}
#endif
#endif /* MARFA_H */
