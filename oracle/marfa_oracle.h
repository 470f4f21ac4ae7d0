/* marfa_oracle.h -- CPU restatement of MARFA's line-by-line PT-table path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP implementation in
 * marfa_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Three precisions of the same algorithm:
 *   MO_R32  mirrors the reference's own precision map operation by operation (FP32
 *           parameters, shapes and grid sums; FP64 positions), so that it can be checked
 *           bit for bit against the reference compiled with amdflang;
 *   MO_F64  "O_alg": the reference's gridded algorithm with every arithmetic operation in
 *           FP64 (constants keep the reference's FP32 values; the pressure shift keeps
 *           the reference's FP32 product so that line selection is identical);
 *   direct  "O_exact": no grids -- every contributing line's FP64 profile evaluated at
 *           every fine point, with the reference's line selection.
 */
#ifndef MARFA_ORACLE_H
#define MARFA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MO_NT 20481
#define MO_NGRID 30690 /* 3 * (10 + 20 + ... + 5120) */

enum { MO_R32 = 0, MO_F64 = 1 };

typedef struct {
    int64_t n;
    const double *nu;       /* [cm-1] transition wavenumber, ascending */
    const float *S;         /* [cm-1/(molecule cm-2)] at 296 K */
    const float *gair;      /* [cm-1/atm] */
    const float *gself;     /* [cm-1/atm] */
    const float *elow;      /* [cm-1] */
    const float *nT;        /* temperature exponent of gair */
    const float *delta;     /* [cm-1/atm] pressure shift */
    const int32_t *molIso;  /* isotopologue*100 + molecule */
} mo_lines;

typedef struct {
    double start_wv, end_wv; /* [cm-1] */
    float cutoff;            /* [cm-1] */
    int shape_id;            /* 0 voigt (src/Shapes.f90:22), 1 lorentzShape (:135) */
    int chi_id;              /* 0 none, 1 tonkov, 2 pollack, 3 perrin */
    int mol_code;            /* moleculeIntCode: 1 H2O, 2 CO2, ... */
    int target;              /* 0 ACS, 1 VAC */
} mo_job;

typedef struct {
    int64_t lines_left, lines_center, lines_right, lines_skipped;
    int64_t samples; /* distinct (line, position) profile evaluations */
} mo_counts;

/* per-subinterval windows, app/main.f90:117-174 + src/LBL.f90:39-61.
 * lo/hi are 0-based [lo, hi) line index ranges, recnum the 1-based record number.
 * Returns the number of subintervals (or -1 if max_rec is too small). */
int mo_windows(const mo_lines *L, const mo_job *J, int max_rec, int64_t *lo, int64_t *hi,
               int32_t *recnum, double *origin);

/* One record (one level, one subinterval): reset, LBL line loop, cascade.
 * rk_out[20481] receives RK (ACS); grids_out (optional, MO_NGRID + MO_NT doubles:
 * level-major P|M|L blocks, then RK) receives the sums before the cascade. */
int mo_record(int precision, const mo_lines *L, const mo_job *J, const float *tips,
              float P, float T, float rho, double a, int64_t lo, int64_t hi,
              double *rk_out, double *grids_out, mo_counts *counts);

/* All levels x subintervals.  out[(lev*nrec + r)*20481 + i], values as written to the
 * PT-table (ACS, or density*RK for VAC). */
int mo_pt_tables(int precision, const mo_lines *L, const mo_job *J, const float *tips,
                 int nlev, const float *P, const float *T, const float *rho,
                 int nrec, double *out, mo_counts *counts);

/* Un-gridded FP64 sum over the same contributing lines at fine points
 * i = i0, i0+stride, ... < 20481 (0-based); out[k] for the k-th point. */
int mo_direct(const mo_lines *L, const mo_job *J, const float *tips, float P, float T,
              float rho, double a, int64_t lo, int64_t hi, int i0, int stride, double *out);

/* Decision trace of the F64 path (for finite-difference checks): between begin and end,
 * mo_record / mo_cascade_f64 in MO_F64 hash the grid nodes that receive samples (out[0]),
 * the Voigt regions and chi bands taken (out[1]) and the cascade limiter choices (out[2]). */
void mo_trace_begin(void);
void mo_trace_end(uint64_t out[3]);

/* Unit-level entry points for golden-vector tests. */
float mo_voigt_r32(double X, float gD, float gL, float S, int shape_id, int chi_id,
                   int mol, float T);
double mo_voigt_f64(double X, double gD, double gL, double S, int shape_id, int chi_id,
                    int mol, double T);
float mo_chi_r32(int chi_id, double X, int mol, float T);
double mo_chi_f64(int chi_id, double X, int mol, double T);
/* out: pSelf, nu*, gD, gL, Q(T), S(T) */
void mo_spec_r32(const float *tips, double nu, float S, float gair, float gself, float E,
                 float n, int32_t molIso, float delta, float P, float T, float rho,
                 double out[6]);
void mo_spec_f64(const float *tips, double nu, float S, float gair, float gself, float E,
                 float n, int32_t molIso, float delta, float P, float T, float rho,
                 double out[6]);
/* grids: MO_NGRID values (level-major P|M|L) then RK[20481]; cascade in place into rk. */
void mo_cascade_r32(const float *grids, float *rk);
void mo_cascade_f64(const double *grids, double *rk);
float mo_wiso(int iso);
void mo_constants(uint32_t out[7]);

#ifdef __cplusplus
}
#endif
#endif
