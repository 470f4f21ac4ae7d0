/* marfa_oracle.c -- CPU restatement of MARFA's line-by-line PT-table path.
 *
 * TEST INFRASTRUCTURE ONLY (see marfa_oracle.h).  Build: oracle/Makefile, with
 * -ffp-contract=off so that a*b+c is never fused and the MO_R32 precision replays the
 * reference (built with the same setting by oracle/build_ref.sh) bit for bit.
 *
 * Reference files restated here (Razumovskyy/MARFA):
 *   src/Constants.f90:7-22      constants (kept at the FP32 values the reference folds)
 *   src/MolarMasses.f90:140-146 WISO molar masses (bit patterns read from the reference)
 *   src/Spectroscopy.f90:38-261 shift, Doppler/Lorentz HWHM, TIPS(T), S(T), wing shapes
 *   src/ChiFactors.f90:18-122  chi factors
 *   src/Shapes.f90:22-141      Humlicek-Kuntz Voigt and lorentzShape
 *   src/LineGridCalc.f90        leftLBL / centerLBL / rightLBL ten-grid deposits
 *   src/Grids.f90:64-139        grid reset and cascade interpolation
 *   src/LBL.f90:25-100          LBL_LOOP window, carry and dispatch
 *   app/main.f90:81-248         level and subinterval loops, starting line, records
 *
 * Parity pins (tests/test_oracle.py): every function against vectors produced by the
 * reference compiled with amdflang (tests/golden/), the FP64 Voigt regions against the
 * Faddeeva function, the cascade against closed forms, the deposit geometry against
 * brute-force coverage invariants.
 */
#include "marfa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ constants */
static inline float f_of_bits(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}
/* bit patterns printed by the reference (oracle/ref_probe.f90 `const`) */
#define B_PI 0x40490FDAU       /* PI = 3.1415926                (Constants.f90:7)  */
#define B_SQLN2 0x3F55224DU    /* sqln2 = sqrt(log(2.))         (Constants.f90:8)  */
#define B_C2 0x3FB829CCU       /* C2 = PLANCK*SPL/BOLsgs        (Constants.f90:15) */
#define B_DOPC 0x34C04328U     /* dopplerCONST                  (Constants.f90:22) */
#define B_SQRTPI 0x3FE2DFC4U   /* sqrt(pi)                                         */
#define B_SQRT14 0x3F97739FU   /* sqrt(1.4)                     (Shapes.f90:126)   */
#define B_LOG2 0x3F317218U     /* log(2.)                       (Spectroscopy:214) */

#define PI_F f_of_bits(B_PI)
#define SQLN2_F f_of_bits(B_SQLN2)
#define C2_F f_of_bits(B_C2)
#define DOPC_F f_of_bits(B_DOPC)
#define SQRTPI_F f_of_bits(B_SQRTPI)
#define SQRT14_F f_of_bits(B_SQRT14)
#define LOG2_F f_of_bits(B_LOG2)
#define BOLSI_F 1.3806503e-23f   /* Constants.f90:14 */
#define PATM_F 101325.f          /* Constants.f90:16 */
#define TREF_F 296.f             /* Constants.f90:24 */

static const uint32_t WISO_BITS[124] = {
#include "../marfa_amd/csrc/wiso_table.inc" /* bit patterns read from the reference, pinned by tests/golden/units.npz */
};

float mo_wiso(int iso) { return (iso >= 1 && iso <= 124) ? f_of_bits(WISO_BITS[iso - 1]) : 0.f; }

void mo_constants(uint32_t out[7]) {
    out[0] = B_PI; out[1] = B_SQLN2; out[2] = B_C2; out[3] = B_DOPC;
    out[4] = B_SQRTPI; out[5] = B_SQRT14; out[6] = B_LOG2;
}

/* grid geometry, src/Grids.f90:6-58 */
static const int NTK[10] = {10, 20, 40, 80, 160, 320, 640, 1280, 2560, 5120};
static const double HK[10] = {1.0, 0.5, 0.25, 0.125, 0.0625, 0.03125, 0.015625,
                              0.0078125, 0.00390625, 0.001953125};
#define HFINE 0.00048828125 /* H = H9/4 */
static inline int gbase(int k) { return 30 * ((1 << k) - 1); } /* P block of level k */

/* ------------------------------------------------------------------ decision trace */
/* For the adjoint's finite-difference checks: while the trace is on, the F64 deposits hash
 * every grid node they add a sample to (in order: the selection), the F64 profile every
 * Voigt region and chi band it takes, and the FP64 cascade every limiter choice.  Two
 * evaluations with equal hashes took the same branches, so a difference quotient between
 * them does not straddle a discontinuity of the piecewise-smooth model.  Single-threaded. */
static struct {
    uint64_t sel, reg, lim;
    int on;
} g_tr;
static inline uint64_t tr_mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    return h * 0xBF58476D1CE4E5B9ULL;
}
static inline int tr_node(int i) {
    if (g_tr.on) g_tr.sel = tr_mix(g_tr.sel, (uint64_t)i);
    return i;
}
static inline void tr_reg(int code) {
    if (g_tr.on) g_tr.reg = tr_mix(g_tr.reg, (uint64_t)code);
}
static inline int tr_lim(int lim) {
    if (g_tr.on) g_tr.lim = tr_mix(g_tr.lim, (uint64_t)lim);
    return lim;
}
void mo_trace_begin(void) {
    g_tr.sel = 1;
    g_tr.reg = 2;
    g_tr.lim = 3;
    g_tr.on = 1;
}
void mo_trace_end(uint64_t out[3]) {
    out[0] = g_tr.sel;
    out[1] = g_tr.reg;
    out[2] = g_tr.lim;
    g_tr.on = 0;
}
/* band of chi_id at |X| (src/ChiFactors.f90:29-122): 0 where chi = 1 */
static int chi_band_of(int chi_id, double ax, int mol) {
    if (mol != 2) return 0;
    if (chi_id == 1) return ax > 3. ? (ax <= 150. ? 1 : ax <= 300. ? 2 : 3) : 0;
    if (chi_id == 2) return ax >= 3. ? (ax < 10. ? 1 : 2) : 0;
    if (chi_id == 3) return ax > 3. ? (ax < 30. ? 1 : ax < 120. ? 2 : 3) : 0;
    return 0;
}

/* ------------------------------------------------------------------ chi factors */
/* src/ChiFactors.f90:18-122.  R32 replays the reference's mixed precision:
 * abs(X) and the exponentials are FP64, the result is FP32. */
float mo_chi_r32(int chi_id, double X, int mol, float T) {
    double ax = fabs(X);
    float r = 1.f;
    if (mol != 2) return 1.f;
    switch (chi_id) {
    case 1: /* tonkov, ChiFactors.f90:29-53 */
        if (ax > 3.) {
            if (ax <= 150.) r = (float)((double)1.084f * exp(-((double)0.027f * ax)));
            else if (ax <= 300.) r = (float)((double)0.208f * exp(-((double)0.016f * ax)));
            else r = (float)((double)0.025f * exp(-((double)0.009f * ax)));
        }
        break;
    case 2: /* pollack, ChiFactors.f90:55-80 */
        if (ax >= 3.) {
            if (ax < 10.) r = (float)((double)1.35f * exp(-ax / 10.));
            else r = (float)((double)0.614f * exp(-ax / 47.));
        }
        break;
    case 3: { /* perrin, ChiFactors.f90:82-122 */
        float B1 = 0.0888f - 0.16f * expf(-(0.00410f * T));
        float B2 = 0.0526f * expf(-(0.00152f * T));
        float B3 = 0.0232f;
        if (ax > 3.) {
            if (ax < 30.) r = (float)exp(-((double)B1 * (ax - 3.)));
            else if (ax < 120.) r = (float)exp((double)(-(B1 * (30.f - 3.f))) - (double)B2 * (ax - 30.));
            else
                r = (float)exp((double)(-(B1 * (30.f - 3.f)) - B2 * (120.f - 30.f)) -
                               (double)B3 * (ax - 120.));
        }
        break;
    }
    default:
        break;
    }
    return r;
}

double mo_chi_f64(int chi_id, double X, int mol, double T) {
    double ax = fabs(X);
    if (mol != 2) return 1.0;
    switch (chi_id) {
    case 1:
        if (ax > 3.) {
            if (ax <= 150.) return (double)1.084f * exp(-((double)0.027f * ax));
            if (ax <= 300.) return (double)0.208f * exp(-((double)0.016f * ax));
            return (double)0.025f * exp(-((double)0.009f * ax));
        }
        return 1.0;
    case 2:
        if (ax >= 3.) {
            if (ax < 10.) return (double)1.35f * exp(-ax / 10.);
            return (double)0.614f * exp(-ax / 47.);
        }
        return 1.0;
    case 3: {
        double B1 = (double)0.0888f - (double)0.16f * exp(-((double)0.00410f * T));
        double B2 = (double)0.0526f * exp(-((double)0.00152f * T));
        double B3 = (double)0.0232f;
        if (ax > 3.) {
            if (ax < 30.) return exp(-B1 * (ax - 3.));
            if (ax < 120.) return exp(-B1 * 27. - B2 * (ax - 30.));
            return exp(-B1 * 27. - B2 * 90. - B3 * (ax - 120.));
        }
        return 1.0;
    }
    default:
        return 1.0;
    }
}

/* ------------------------------------------------------------------ shapes, R32 */
typedef struct {
    float gD, gL, S;
    int shape_id, chi_id, mol;
    float T;
} r32_line;

/* lorentz * chi, Spectroscopy.f90:194-229 */
static inline float r32_chilor(double X, const r32_line *p) {
    double d = X * X + (double)(p->gL * p->gL);
    float lor = (float)(((double)p->gL / ((double)PI_F * d)) * (double)p->S);
    return lor * mo_chi_r32(p->chi_id, X, p->mol, p->T);
}

static inline float r32_doppler(double X, float gD, float S) {
    /* Spectroscopy.f90:214 */
    float a = SQLN2_F / (SQRTPI_F * gD);
    double q = X / (double)gD;
    return (float)((double)a * exp(-((q * q) * (double)LOG2_F)) * (double)S);
}

static const float ASYM_U[9] = {1.f, 1.5f, 2.f, 2.5f, 3.f, 3.5f, 4.f, 4.5f, 5.f};
static const float ASYM_W[9] = {-0.688f, 0.2667f, 0.6338f, 0.4405f, 0.2529f,
                                0.1601f, 0.1131f, 0.0853f, 0.068f};

static float r32_voigt(double X, const r32_line *p) {
    /* Shapes.f90:49-133; the VY-keyed coefficient cache (:41-47) is transparent here */
    const float gD = p->gD, gL = p->gL, S = p->S;
    float VX = (float)fabs(((double)SQLN2_F * X) / (double)gD);
    float VY = SQLN2_F * gL / gD;
    if (VX >= 15.f) return r32_chilor(X, p);
    float V2 = VX * VX;
    float num, den;
    if (VX + VY >= 15.0f) {
        float Y2 = VY * VY;
        float A1 = (0.2820948f + 0.5641896f * Y2) * VY;
        float B1 = 0.5641896f * VY;
        float A2 = 0.25f + Y2 + Y2 * Y2;
        float B2 = Y2 + Y2 - 1.f;
        num = A1 + B1 * V2;
        den = A2 + B2 * V2 + V2 * V2;
    } else if (VX + VY >= 5.5f) {
        float Y2 = VY * VY;
        float A3 = VY * (((0.56419f * Y2 + 3.10304f) * Y2 + 4.65456f) * Y2 + 1.05786f);
        float B3 = VY * ((1.69257f * Y2 + 0.56419f) * Y2 + 2.962f);
        float C3 = VY * (1.69257f * Y2 - 2.53885f);
        float D3 = VY * 0.56419f;
        float A4 = (((Y2 + 6.0f) * Y2 + 10.5f) * Y2 + 4.5f) * Y2 + 0.5625f;
        float B4 = ((4.0f * Y2 + 6.0f) * Y2 + 9.0f) * Y2 - 4.5f;
        float C4 = 10.5f + 6.0f * (Y2 - 1.0f) * Y2;
        float D4 = 4.0f * Y2 - 6.0f;
        num = ((D3 * V2 + C3) * V2 + B3) * V2 + A3;
        den = (((V2 + D4) * V2 + C4) * V2 + B4) * V2 + A4;
    } else if (VX <= 1.0f || VY >= 0.02f) {
        float Y = VY;
        float A5 = ((((((((0.564224f * Y + 7.55895f) * Y + 49.5213f) * Y + 204.510f) * Y +
                       581.746f) * Y + 1174.8f) * Y + 1678.33f) * Y + 1629.76f) * Y +
                    973.778f) * Y + 272.102f;
        float B5 = ((((((2.25689f * Y + 22.6778f) * Y + 100.705f) * Y + 247.198f) * Y +
                      336.364f) * Y + 220.843f) * Y - 2.34403f) * Y - 60.5644f;
        float C5 = ((((3.38534f * Y + 22.6798f) * Y + 52.8454f) * Y + 42.5683f) * Y + 18.546f) *
                       Y + 4.58029f;
        float D5 = ((2.25689f * Y + 7.56186f) * Y + 1.66203f) * Y - 0.128922f;
        float E5 = 0.971457E-3f + 0.564224f * Y;
        float A6 = (((((((((Y + 13.3988f) * Y + 88.2674f) * Y + 369.199f) * Y + 1074.41f) * Y +
                        2256.98f) * Y + 3447.63f) * Y + 3764.97f) * Y + 2802.87f) * Y +
                    1280.83f) * Y + 272.102f;
        float B6 = (((((((5.f * Y + 53.5952f) * Y + 266.299f) * Y + 793.427f) * Y + 1549.68f) *
                        Y + 2037.31f) * Y + 1758.34f) * Y + 902.306f) * Y + 211.678f;
        float C6 = (((((10.f * Y + 80.3928f) * Y + 269.292f) * Y + 479.258f) * Y + 497.302f) *
                        Y + 308.186f) * Y + 78.866f;
        float D6 = (((10.f * Y + 53.5952f) * Y + 92.7586f) * Y + 55.0293f) * Y + 22.0353f;
        float E6 = (5.0f * Y + 13.3988f) * Y + 1.49645f;
        num = (((E5 * V2 + D5) * V2 + C5) * V2 + B5) * V2 + A5;
        den = ((((V2 + E6) * V2 + D6) * V2 + C6) * V2 + B6) * V2 + A6;
    } else {
        /* region 4, Shapes.f90:119-130 */
        if (VX > 5.f) return r32_chilor(X, p) * (1.f + 1.5f / (VX * VX));
        if (VX > SQRT14_F) {
            int I = (int)(VX / 0.5f - 1.00001f); /* 1-based index into U, W */
            float F = 2.f * (ASYM_W[I - 1] * (ASYM_U[I] - VX) + ASYM_W[I] * (VX - ASYM_U[I - 1]));
            float dop = r32_doppler(X, gD, S);
            double lor = ((double)gL / ((double)PI_F * (X * X))) * (double)(1.f + F);
            return (float)((double)dop + lor * (double)S);
        }
        return r32_doppler(X, gD, S);
    }
    return num / den / SQRTPI_F / (gD / SQLN2_F) * S;
}

static inline float r32_shape(double X, const r32_line *p) {
    return p->shape_id == 1 ? r32_chilor(X, p) : r32_voigt(X, p);
}

float mo_voigt_r32(double X, float gD, float gL, float S, int shape_id, int chi_id, int mol,
                   float T) {
    r32_line p = {gD, gL, S, shape_id, chi_id, mol, T};
    return r32_shape(X, &p);
}

/* ------------------------------------------------------------------ shapes, F64 */
typedef struct {
    double gD, gL, S;
    int shape_id, chi_id, mol;
    double T;
} f64_line;

static inline double f64_chilor(double X, const f64_line *p) {
    if (g_tr.on) tr_reg(100 + chi_band_of(p->chi_id, fabs(X), p->mol));
    double lor = p->gL / ((double)PI_F * (X * X + p->gL * p->gL)) * p->S;
    return lor * mo_chi_f64(p->chi_id, X, p->mol, p->T);
}

static inline double f64_doppler(double X, double gD, double S) {
    double a = (double)SQLN2_F / ((double)SQRTPI_F * gD);
    double q = X / gD;
    return a * exp(-((q * q) * (double)LOG2_F)) * S;
}

static double f64_voigt(double X, const f64_line *p) {
    const double gD = p->gD, gL = p->gL, S = p->S;
    const double sq = SQLN2_F;
    double VX = fabs((sq * X) / gD);
    double VY = sq * gL / gD;
    if (VX >= 15.) return f64_chilor(X, p);
    double V2 = VX * VX, num, den;
#define D(x) ((double)(x))
    if (VX + VY >= 15.) {
        tr_reg(1);
        double Y2 = VY * VY;
        double A1 = (D(0.2820948f) + D(0.5641896f) * Y2) * VY;
        double B1 = D(0.5641896f) * VY;
        double A2 = 0.25 + Y2 + Y2 * Y2;
        double B2 = Y2 + Y2 - 1.;
        num = A1 + B1 * V2;
        den = A2 + B2 * V2 + V2 * V2;
    } else if (VX + VY >= 5.5) {
        tr_reg(2);
        double Y2 = VY * VY;
        double A3 = VY * (((D(0.56419f) * Y2 + D(3.10304f)) * Y2 + D(4.65456f)) * Y2 + D(1.05786f));
        double B3 = VY * ((D(1.69257f) * Y2 + D(0.56419f)) * Y2 + D(2.962f));
        double C3 = VY * (D(1.69257f) * Y2 - D(2.53885f));
        double D3 = VY * D(0.56419f);
        double A4 = (((Y2 + 6.0) * Y2 + 10.5) * Y2 + 4.5) * Y2 + 0.5625;
        double B4 = ((4.0 * Y2 + 6.0) * Y2 + 9.0) * Y2 - 4.5;
        double C4 = 10.5 + 6.0 * (Y2 - 1.0) * Y2;
        double D4 = 4.0 * Y2 - 6.0;
        num = ((D3 * V2 + C3) * V2 + B3) * V2 + A3;
        den = (((V2 + D4) * V2 + C4) * V2 + B4) * V2 + A4;
    } else if (VX <= 1.0 || VY >= D(0.02f)) {
        tr_reg(3);
        double Y = VY;
        double A5 = ((((((((D(0.564224f) * Y + D(7.55895f)) * Y + D(49.5213f)) * Y + D(204.510f)) * Y +
                        D(581.746f)) * Y + D(1174.8f)) * Y + D(1678.33f)) * Y + D(1629.76f)) * Y +
                     D(973.778f)) * Y + D(272.102f);
        double B5 = ((((((D(2.25689f) * Y + D(22.6778f)) * Y + D(100.705f)) * Y + D(247.198f)) * Y +
                       D(336.364f)) * Y + D(220.843f)) * Y - D(2.34403f)) * Y - D(60.5644f);
        double C5 = ((((D(3.38534f) * Y + D(22.6798f)) * Y + D(52.8454f)) * Y + D(42.5683f)) * Y +
                      D(18.546f)) * Y + D(4.58029f);
        double D5 = ((D(2.25689f) * Y + D(7.56186f)) * Y + D(1.66203f)) * Y - D(0.128922f);
        double E5 = D(0.971457E-3f) + D(0.564224f) * Y;
        double A6 = (((((((((Y + D(13.3988f)) * Y + D(88.2674f)) * Y + D(369.199f)) * Y + D(1074.41f)) * Y +
                         D(2256.98f)) * Y + D(3447.63f)) * Y + D(3764.97f)) * Y + D(2802.87f)) * Y +
                     D(1280.83f)) * Y + D(272.102f);
        double B6 = (((((((5. * Y + D(53.5952f)) * Y + D(266.299f)) * Y + D(793.427f)) * Y + D(1549.68f)) * Y +
                       D(2037.31f)) * Y + D(1758.34f)) * Y + D(902.306f)) * Y + D(211.678f);
        double C6 = (((((10. * Y + D(80.3928f)) * Y + D(269.292f)) * Y + D(479.258f)) * Y + D(497.302f)) * Y +
                      D(308.186f)) * Y + D(78.866f);
        double D6 = (((10. * Y + D(53.5952f)) * Y + D(92.7586f)) * Y + D(55.0293f)) * Y + D(22.0353f);
        double E6 = (5.0 * Y + D(13.3988f)) * Y + D(1.49645f);
        num = (((E5 * V2 + D5) * V2 + C5) * V2 + B5) * V2 + A5;
        den = ((((V2 + E6) * V2 + D6) * V2 + C6) * V2 + B6) * V2 + A6;
    } else {
        tr_reg(4);
        if (VX > 5.) return f64_chilor(X, p) * (1. + 1.5 / (VX * VX));
        if (VX > D(SQRT14_F)) {
            int I = (int)(VX / 0.5 - D(1.00001f));
            tr_reg(10 + I);
            double F = 2. * (D(ASYM_W[I - 1]) * (D(ASYM_U[I]) - VX) + D(ASYM_W[I]) * (VX - D(ASYM_U[I - 1])));
            return f64_doppler(X, gD, S) + (gL / ((double)PI_F * (X * X)) * (1. + F)) * S;
        }
        return f64_doppler(X, gD, S);
    }
#undef D
    return num / den / (double)SQRTPI_F / (gD / (double)SQLN2_F) * S;
}

static inline double f64_shape(double X, const f64_line *p) {
    return p->shape_id == 1 ? f64_chilor(X, p) : f64_voigt(X, p);
}

double mo_voigt_f64(double X, double gD, double gL, double S, int shape_id, int chi_id, int mol,
                    double T) {
    f64_line p = {gD, gL, S, shape_id, chi_id, mol, T};
    return f64_shape(X, &p);
}

/* ------------------------------------------------------------------ spectroscopy */
static inline float tips_at(const float *tips, int iso, int col) { /* TIPS(iso, col), 1-based */
    return tips[(iso - 1) * 501 + (col - 1)];
}

void mo_spec_r32(const float *tips, double nu, float S, float gair, float gself, float E,
                 float n, int32_t molIso, float delta, float P, float T, float rho,
                 double out[6]) {
    /* app/main.f90:92 */
    float pSelf = rho * 10.f * BOLSI_F * T / PATM_F;
    int iso = molIso / 100;
    float M = mo_wiso(iso);
    /* Spectroscopy.f90:48, :63, :84 */
    double nus = nu + (double)(delta * P);
    float gD = (float)(((double)DOPC_F * nus) * (double)sqrtf(T / M));
    float gL = powf(TREF_F / T, n) * (gair * (P - pSelf) + gself * pSelf);
    /* TIPSofT, Spectroscopy.f90:110-114 */
    int idx = (int)((T - 20.0f) / 2.f + 1.f);
    float tlow = (float)idx * 2.0f + 18.f;
    float wu = (T - tlow) / 2.f;
    float wl = 1.f - wu;
    float Q = wl * tips_at(tips, iso, idx) + wu * tips_at(tips, iso, idx + 1);
    /* intensityOfT, Spectroscopy.f90:137-144 */
    float tf = tips_at(tips, iso, 139) / Q;
    float bf = expf(-(C2_F * E / T)) / expf(-(C2_F * E / TREF_F));
    float ef = (float)((1.0 - exp(-(((double)C2_F * nu) / (double)T))) /
                       (1.0 - exp(-(((double)C2_F * nu) / (double)TREF_F))));
    float ST = S * tf * bf * ef;
    out[0] = pSelf; out[1] = nus; out[2] = gD; out[3] = gL; out[4] = Q; out[5] = ST;
}

void mo_spec_f64(const float *tips, double nu, float S, float gair, float gself, float E,
                 float n, int32_t molIso, float delta, float P, float T, float rho,
                 double out[6]) {
    double dT = T, dP = P;
    double pSelf = (double)rho * 10. * (double)BOLSI_F * dT / (double)PATM_F;
    int iso = molIso / 100;
    double M = mo_wiso(iso);
    /* the pressure shift keeps the reference's FP32 product: line selection is identical */
    double nus = nu + (double)(delta * P);
    double gD = ((double)DOPC_F * nus) * sqrt(dT / M);
    double gL = pow((double)TREF_F / dT, (double)n) * ((double)gair * (dP - pSelf) + (double)gself * pSelf);
    int idx = (int)((dT - 20.0) / 2. + 1.);
    double tlow = (double)idx * 2.0 + 18.;
    double wu = (dT - tlow) / 2.;
    double Q = (1. - wu) * (double)tips_at(tips, iso, idx) + wu * (double)tips_at(tips, iso, idx + 1);
    double tf = (double)tips_at(tips, iso, 139) / Q;
    double c2 = C2_F;
    double bf = exp(-(c2 * E / dT)) / exp(-(c2 * E / (double)TREF_F));
    double ef = (1.0 - exp(-((c2 * nu) / dT))) / (1.0 - exp(-((c2 * nu) / (double)TREF_F)));
    double ST = (double)S * tf * bf * ef;
    out[0] = pSelf; out[1] = nus; out[2] = gD; out[3] = gL; out[4] = Q; out[5] = ST;
}

/* ------------------------------------------------------------------ cascade */
/* src/Grids.f90:80-139.  grids: level-major P|M|L blocks, rk 1..20481 stored 0-based. */
#define CASCADE_BODY(REAL, C375, C75, C125, C5, LIM)                                         \
    for (int k = 0; k < 9; k++) {                                                            \
        const int n = NTK[k];                                                                \
        const REAL *cP = g + gbase(k), *cM = cP + n, *cL = cM + n;                            \
        REAL *fP = g + gbase(k + 1), *fM = fP + 2 * n, *fL = fM + 2 * n;                     \
        for (int j = 0; j < n; j++) {                                                        \
            int i = 2 * j;                                                                   \
            fP[i] = fP[i] + cP[j];                                                           \
            REAL A = C375 * cP[j] + C75 * cM[j] - C125 * cL[j];                              \
            if (LIM(A < cP[j] || A < cM[j])) A = C5 * (cP[j] + cM[j]);                       \
            fM[i] = fM[i] + A;                                                               \
            fL[i] = fL[i] + cM[j];                                                           \
            fP[i + 1] = fP[i + 1] + cM[j];                                                   \
            A = C375 * cL[j] + C75 * cM[j] - C125 * cP[j];                                   \
            if (LIM(A < cL[j] || A < cM[j])) A = C5 * (cL[j] + cM[j]);                       \
            fM[i + 1] = fM[i + 1] + A;                                                       \
            fL[i + 1] = fL[i + 1] + cL[j];                                                   \
        }                                                                                    \
    }                                                                                        \
    {                                                                                        \
        const REAL *P9 = g + gbase(9), *M9 = P9 + 5120, *L9 = M9 + 5120;                     \
        for (int j = 0; j < 5120; j++) {                                                     \
            REAL A = P9[j] * C375 + M9[j] * C75 - L9[j] * C125;                              \
            if (LIM(A < P9[j] || A < M9[j])) A = C5 * (P9[j] + M9[j]);                       \
            rk[4 * j + 1] = rk[4 * j + 1] + A;                                               \
            rk[4 * j + 2] = rk[4 * j + 2] + M9[j];                                           \
            A = L9[j] * C375 + M9[j] * C75 - P9[j] * C125;                                   \
            if (LIM(A < L9[j] || A < M9[j])) A = C5 * (L9[j] + M9[j]);                       \
            rk[4 * j + 3] = rk[4 * j + 3] + A;                                               \
            rk[4 * j + 4] = rk[4 * j + 4] + L9[j];                                           \
        }                                                                                    \
    }

#define NO_TRACE(x) (x)
static void cascade_r32(float *g, float *rk) { CASCADE_BODY(float, 0.375f, 0.75f, 0.125f, 0.5f, NO_TRACE) }
static void cascade_f64(double *g, double *rk) { CASCADE_BODY(double, 0.375, 0.75, 0.125, 0.5, tr_lim) }

void mo_cascade_r32(const float *grids, float *rk) {
    float *g = (float *)malloc(sizeof(float) * MO_NGRID);
    memcpy(g, grids, sizeof(float) * MO_NGRID);
    cascade_r32(g, rk);
    free(g);
}

void mo_cascade_f64(const double *grids, double *rk) {
    double *g = (double *)malloc(sizeof(double) * MO_NGRID);
    memcpy(g, grids, sizeof(double) * MO_NGRID);
    cascade_f64(g, rk);
    free(g);
}

/* ------------------------------------------------------------------ deposits */
/* src/LineGridCalc.f90, restated once for both precisions (GREAL = accumulator and
 * sample type).  Positions are FP64 and formed with the reference's exact sequence of
 * operations, so every sample sits at the bit-identical X the reference uses.  Every
 * grid accumulator receives at most one sample per line, so summation order is the line
 * order, as in the reference.  EPS is 0 (src/Grids.f90:66); its `FF < EPS` exits are
 * kept (they only act on negative or NaN samples). */
#define GREAL float
#define GLINE r32_line
#define GSHAPE r32_shape
#define GNAME(x) x##_r32
#define GTRACE(i) (i)
#include "oracle_deposit.inc"
#undef GREAL
#undef GLINE
#undef GSHAPE
#undef GNAME
#undef GTRACE
#define GREAL double
#define GLINE f64_line
#define GSHAPE f64_shape
#define GNAME(x) x##_f64
#define GTRACE(i) tr_node(i)
#include "oracle_deposit.inc"
#undef GREAL
#undef GLINE
#undef GSHAPE
#undef GNAME
#undef GTRACE

/* ------------------------------------------------------------------ windows */
int mo_windows(const mo_lines *Ls, const mo_job *J, int max_rec, int64_t *lo, int64_t *hi,
               int32_t *recnum, double *origin) {
    const int64_t n = Ls->n;
    const double *nu = Ls->nu;
    const double cut = (double)J->cutoff;
    /* determineStartingSpectralLine, app/main.f90:206-240 (0-based here) */
    int64_t start = 0;
    if (J->start_wv > cut) {
        double ext = J->start_wv - cut;
        while (start < n && nu[start] <= ext) start++;
    }
    double startWV = start < n ? nu[start] : INFINITY;
    int nrec = 0;
    int64_t lineIdx = start;
    for (double a = J->start_wv; a < J->end_wv; a = a + 10.0) {
        if (nrec >= max_rec) return -1;
        double extStart = a - cut;
        double extEnd = a + 10.0 + cut;
        double cap = extStart + 10.0;
        if (cap <= startWV) cap = startWV;
        int64_t I = lineIdx, carry = lineIdx;
        while (I < n && !(nu[I] >= extEnd)) {
            if (nu[I] <= cap) carry = I;
            I++;
        }
        lo[nrec] = lineIdx;
        hi[nrec] = I;
        if (recnum) recnum[nrec] = (int32_t)((a + 1.0) / 10.0);
        if (origin) origin[nrec] = a;
        lineIdx = carry;
        nrec++;
    }
    return nrec;
}

/* ------------------------------------------------------------------ records */
int mo_record(int precision, const mo_lines *Ls, const mo_job *J, const float *tips, float P,
              float T, float rho, double a, int64_t lo, int64_t hi, double *rk_out,
              double *grids_out, mo_counts *counts) {
    mo_counts c;
    memset(&c, 0, sizeof c);
    if (precision == MO_R32) {
        float *g = (float *)calloc(MO_NGRID + MO_NT + 1, sizeof(float));
        float *rk = g + MO_NGRID;
        for (int64_t i = lo; i < hi; i++) {
            double s[6];
            mo_spec_r32(tips, Ls->nu[i], Ls->S[i], Ls->gair[i], Ls->gself[i], Ls->elow[i],
                        Ls->nT[i], Ls->molIso[i], Ls->delta[i], P, T, rho, s);
            r32_line p = {(float)s[2], (float)s[3], (float)s[5], J->shape_id, J->chi_id,
                          J->mol_code, T};
            deposit_line_r32(g, rk, a, s[1], J->cutoff, &p, &c);
        }
        if (grids_out)
            for (int i = 0; i < MO_NGRID + MO_NT; i++) grids_out[i] = g[i];
        cascade_r32(g, rk);
        for (int i = 0; i < MO_NT; i++) rk_out[i] = rk[i];
        free(g);
    } else {
        double *g = (double *)calloc(MO_NGRID + MO_NT + 1, sizeof(double));
        double *rk = g + MO_NGRID;
        for (int64_t i = lo; i < hi; i++) {
            double s[6];
            mo_spec_f64(tips, Ls->nu[i], Ls->S[i], Ls->gair[i], Ls->gself[i], Ls->elow[i],
                        Ls->nT[i], Ls->molIso[i], Ls->delta[i], P, T, rho, s);
            f64_line p = {s[2], s[3], s[5], J->shape_id, J->chi_id, J->mol_code, (double)T};
            deposit_line_f64(g, rk, a, s[1], J->cutoff, &p, &c);
        }
        if (grids_out) memcpy(grids_out, g, sizeof(double) * (MO_NGRID + MO_NT));
        cascade_f64(g, rk);
        memcpy(rk_out, rk, sizeof(double) * MO_NT);
        free(g);
    }
    if (counts) {
        counts->lines_left += c.lines_left;
        counts->lines_center += c.lines_center;
        counts->lines_right += c.lines_right;
        counts->lines_skipped += c.lines_skipped;
        counts->samples += c.samples;
    }
    return 0;
}

int mo_pt_tables(int precision, const mo_lines *Ls, const mo_job *J, const float *tips,
                 int nlev, const float *P, const float *T, const float *rho, int nrec,
                 double *out, mo_counts *counts) {
    int64_t *lo = (int64_t *)malloc(sizeof(int64_t) * (nrec + 1));
    int64_t *hi = (int64_t *)malloc(sizeof(int64_t) * (nrec + 1));
    double *org = (double *)malloc(sizeof(double) * (nrec + 1));
    int nr = mo_windows(Ls, J, nrec + 1, lo, hi, NULL, org);
    if (nr != nrec) {
        free(lo); free(hi); free(org);
        return -1;
    }
    for (int l = 0; l < nlev; l++) {
        for (int r = 0; r < nrec; r++) {
            double *o = out + ((int64_t)l * nrec + r) * MO_NT;
            mo_record(precision, Ls, J, tips, P[l], T[l], rho[l], org[r], lo[r], hi[r], o, NULL,
                      counts);
            if (J->target == 1) {
                if (precision == MO_R32)
                    for (int i = 0; i < MO_NT; i++) o[i] = (double)(rho[l] * (float)o[i]);
                else
                    for (int i = 0; i < MO_NT; i++) o[i] = (double)rho[l] * o[i];
            }
        }
    }
    free(lo); free(hi); free(org);
    return 0;
}

int mo_direct(const mo_lines *Ls, const mo_job *J, const float *tips, float P, float T,
              float rho, double a, int64_t lo, int64_t hi, int i0, int stride, double *out) {
    int npts = 0;
    for (int i = i0; i < MO_NT; i += stride) out[npts++] = 0.0;
    const double cut = (double)J->cutoff;
    for (int64_t li = lo; li < hi; li++) {
        double s[6];
        mo_spec_f64(tips, Ls->nu[li], Ls->S[li], Ls->gair[li], Ls->gself[li], Ls->elow[li],
                    Ls->nT[li], Ls->molIso[li], Ls->delta[li], P, T, rho, s);
        f64_line p = {s[2], s[3], s[5], J->shape_id, J->chi_id, J->mol_code, (double)T};
        double nus = s[1];
        double u = nus - a;
        /* the reference's line selection (src/LBL.f90:81-98, LineGridCalc.f90:22-23,192,509) */
        if (nus < a) {
            if (-u > cut) continue;
        } else if (nus > a + 10.0) {
            if ((u - 10.0) >= cut) continue;
        } else {
            if (u >= 10.0) continue;
        }
        int k = 0;
        for (int i = i0; i < MO_NT; i += stride, k++) out[k] += f64_shape(u - (double)i * HFINE, &p);
    }
    return npts;
}
