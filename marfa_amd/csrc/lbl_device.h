// lbl_device.h -- device-side constants, line shapes and grid geometry of the LBL core.
//
// Line shapes follow the reference's Humlicek-Kuntz Voigt (src/Shapes.f90:22-133), the
// chi-corrected Lorentz and wing asymptotics (src/Spectroscopy.f90:194-261) and the chi
// factors (src/ChiFactors.f90:18-122).  Constants keep the FP32 values the reference
// folds (src/Constants.f90:7-22): their bit patterns are below.
//
// Two precision forms share this code:
//   double  (MARFA_FP64)  every operation in FP64;
//   float   (MARFA_FP32K) samples in FP32.  The region of the Humlicek-Kuntz split is
//           decided on the fast path only when the FP32 estimate of VX is clearly in the
//           Lorentz region; every other sample decides its region from the FP64 value
//           VX = |X| sqrt(ln2)/gD (2 ulp from the oracle's |sqrt(ln2) X / gD|), so both
//           forms pick the same approximation for the same sample except when VX lies
//           within ~1e-16 relative of a region boundary.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace marfa {

constexpr int NT = 20481;                 // src/Grids.f90:16
constexpr double HF = 0.00048828125;      // finest step H = 2^-11
constexpr double H9 = 0.001953125;        // 2^-9

__host__ __device__ inline float bits_f(uint32_t b) { return __builtin_bit_cast(float, b); }

// bit patterns of the reference's FP32 constants
#define MARFA_PI_F bits_f(0x40490FDAU)      // PI = 3.1415926
#define MARFA_SQLN2_F bits_f(0x3F55224DU)   // sqrt(log(2.))
#define MARFA_C2_F bits_f(0x3FB829CCU)      // h c / k
#define MARFA_DOPC_F bits_f(0x34C04328U)    // dopplerCONST
#define MARFA_SQRTPI_F bits_f(0x3FE2DFC4U)  // sqrt(pi)
#define MARFA_SQRT14_F bits_f(0x3F97739FU)  // sqrt(1.4)
#define MARFA_LOG2_F bits_f(0x3F317218U)    // log(2.)
#define MARFA_BOLSI_F 1.3806503e-23f        // src/Constants.f90:14
#define MARFA_PATM_F 101325.f               // src/Constants.f90:16
#define MARFA_TREF_F 296.f                  // src/Constants.f90:24

__host__ __device__ inline double hk(int k) { return __builtin_ldexp(1.0, -k); }  // H_k = 2^-k

// ---------------------------------------------------------------- chi factors
template <typename R> __device__ inline R dexp(R x);
template <> __device__ inline double dexp<double>(double x) { return exp(x); }
template <> __device__ inline float dexp<float>(float x) { return __expf(x); }

// chi(|X|), src/ChiFactors.f90:29-122; applied only to CO2 (moleculeIntCode 2)
template <typename R>
__device__ inline R chi_factor(int chi, R ax, R B1, R B2) {
    if (chi == 1) {
        if (ax > R(3)) {
            if (ax <= R(150)) return R(1.084f) * dexp<R>(-(R(0.027f) * ax));
            if (ax <= R(300)) return R(0.208f) * dexp<R>(-(R(0.016f) * ax));
            return R(0.025f) * dexp<R>(-(R(0.009f) * ax));
        }
    } else if (chi == 2) {
        // exp(-|X|/10) and exp(-|X|/47) with the divisions taken as products by the
        // reciprocals (one rounding of the exponent's argument apart)
        if (ax >= R(3)) {
            if (ax < R(10)) return R(1.35f) * dexp<R>(-ax * R(0.1));
            return R(0.614f) * dexp<R>(-ax * R(1.0 / 47.0));
        }
    } else if (chi == 3) {
        const R B3 = R(0.0232f);
        if (ax > R(3)) {
            if (ax < R(30)) return dexp<R>(-B1 * (ax - R(3)));
            if (ax < R(120)) return dexp<R>(-B1 * R(27) - B2 * (ax - R(30)));
            return dexp<R>(-B1 * R(27) - B2 * R(90) - B3 * (ax - R(120)));
        }
    }
    return R(1);
}

// chi_factor with the band chosen by selects instead of branches for chi 1 and 2, whose
// bands all read coef * exp(-(rate |X|)): each band is evaluated with chi_factor's
// operations, so the values are the same; Perrin (chi 3) keeps chi_factor
template <typename R>
__device__ inline R chi_factor_sel(int chi, R ax, R B1, R B2) {
    if (chi == 1) {
        const bool b1 = ax <= R(150), b2 = ax <= R(300);
        const R coef = b1 ? R(1.084f) : b2 ? R(0.208f) : R(0.025f);
        const R rate = b1 ? R(0.027f) : b2 ? R(0.016f) : R(0.009f);
        const R v = coef * dexp<R>(-(rate * ax));
        return ax > R(3) ? v : R(1);
    }
    if (chi == 2) {
        const bool b1 = ax < R(10);
        const R coef = b1 ? R(1.35f) : R(0.614f);
        const R rate = b1 ? R(0.1) : R(1.0 / 47.0);
        const R v = coef * dexp<R>(-ax * rate);
        return ax >= R(3) ? v : R(1);
    }
    return chi_factor<R>(chi, ax, B1, B2);
}

// ---------------------------------------------------------------- per-line state
// Derived once per (line, level) visit; R is the sample precision.
template <typename R>
struct LineShape {
    R xs;     // sqrt(ln2)/gD    : VX = |X| xs
    R vy;     // sqrt(ln2) gL/gD : VY
    R lorc;   // gL S / PI       : Lorentz numerator
    R gl2;    // gL^2
    R normc;  // S / sqrt(pi) / (gD/sqrt(ln2)) : K(x,y) -> profile
    R dopa;   // sqrt(ln2)/(sqrt(pi) gD) S    : Doppler amplitude
    R igd;    // 1/gD
    R gL;     // Lorentz HWHM
    R S;      // line intensity (the adjoint's dF/dgL)
    // region decisions in FP64: VX = |X| xs64 (within 2 ulp of the FP64 form's
    // |sqrt(ln2) X / gD|), VY = vy64 formed as the FP64 form forms it
    double xs64, vy64;
    // in the level's sorted arrays: shifted centre and line index, so that one load gives a
    // gather everything it tests and evaluates
    double nus;
    int32_t idx;
};
// 72 B (FP32) / 104 B (FP64).  The deposit kernels' lanes load consecutive records; a
// 64-B record (without gL: a power-of-two stride) measured 1 ms slower in each of the two
// deposit kernels than this layout.
static_assert(sizeof(LineShape<float>) == 72 && sizeof(LineShape<double>) == 104, "line record layout");

// The fields the level-0 kernel's far wings read (FP32 form), 24 B per sorted line, written
// by the sort beside the line records of a level batch ([levels][lines] after the records):
// 64 lanes' far lines take 1.5 KB in two loads instead of three loads across 4.6 KB of
// records.
struct FarRec {
    double nus;
    int32_t idx;
    float xs, lorc, gl2;
};
static_assert(sizeof(FarRec) == 24, "far record layout");

// bytes per sorted line of a level batch's line records (FP32 form: with the far records)
template <typename R>
constexpr size_t line_record_bytes() {
    return sizeof(LineShape<R>) + (sizeof(R) == 4 ? sizeof(FarRec) : 0);
}

template <typename R>
__device__ inline LineShape<R> make_line(double gD, double gL, double S) {
    LineShape<R> p;
    const double sq = (double)MARFA_SQLN2_F;
    if constexpr (sizeof(R) == 8) {
        p.xs = sq / gD;
        p.vy = sq * gL / gD;
        p.lorc = gL * S / (double)MARFA_PI_F;
        p.gl2 = gL * gL;
        p.normc = S / (double)MARFA_SQRTPI_F / (gD / sq);
        p.dopa = sq / ((double)MARFA_SQRTPI_F * gD) * S;
        p.igd = 1.0 / gD;
        p.gL = gL;
        p.S = S;
    } else {
        const float g = (float)gD, l = (float)gL, s = (float)S;
        p.xs = MARFA_SQLN2_F / g;
        p.vy = (float)(sq * gL / gD);
        p.lorc = l * s / MARFA_PI_F;
        p.gl2 = l * l;
        p.normc = s / MARFA_SQRTPI_F / (g / MARFA_SQLN2_F);
        p.dopa = MARFA_SQLN2_F / (MARFA_SQRTPI_F * g) * s;
        p.igd = 1.0f / g;
        p.gL = l;
        p.S = s;
    }
    p.xs64 = sq / gD;
    p.vy64 = sq * gL / gD;
    return p;
}

struct ShapeCtl {
    int shape_id, chi_id, mol;
    float B1, B2;     // Perrin temperature coefficients of the level (FP32 as the reference)
    double B1d, B2d;  // same in FP64
    double dB1, dB2;  // their temperature derivatives (the adjoint's Perrin term)
};

template <typename R>
__device__ inline R chi_of(const ShapeCtl &c, R ax) {
    if (c.mol != 2 || c.chi_id == 0) return R(1);
    if constexpr (sizeof(R) == 8) return chi_factor<double>(c.chi_id, ax, c.B1d, c.B2d);
    else return chi_factor<float>(c.chi_id, ax, c.B1, c.B2);
}

// chi factor chosen at compile time (CHI >= 0), or at run time from the control (CHI < 0)
template <typename R, int CHI>
__device__ inline R chi_sel(const ShapeCtl &c, R ax) {
    if constexpr (CHI < 0) {
        return chi_of<R>(c, ax);
    } else if constexpr (CHI == 0) {
        return R(1);
    } else {
        if (c.mol != 2) return R(1);
        if constexpr (sizeof(R) == 8) return chi_factor_sel<double>(CHI, ax, c.B1d, c.B2d);
        else return chi_factor_sel<float>(CHI, ax, c.B1, c.B2);
    }
}

// 1/x: the FP32 form takes the hardware reciprocal (1 ulp), the FP64 form divides
template <typename R> __device__ inline R recip(R x);
template <> __device__ inline double recip<double>(double x) { return 1.0 / x; }
template <> __device__ inline float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// a / b: the FP32 form multiplies by the hardware reciprocal (1 ulp; the FP32 form's
// samples are held to a tolerance, not to the FP64 form's bits), every other type divides
template <typename R> __device__ inline R fdiv(R a, R b) { return a / b; }
template <> __device__ inline float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Forward-mode dual number with two tangents (d/dx^2, d/dy): evaluating the Kuntz rationals
// in it gives their exact partial derivatives for the adjoint, through the same
// coefficient arithmetic as the value.
template <typename R>
struct Dual2 {
    R v, a, b;
    __device__ Dual2() {}
    __device__ Dual2(R x) : v(x), a(R(0)), b(R(0)) {}
    __device__ Dual2(R x, R da, R db) : v(x), a(da), b(db) {}
};
template <typename R> __device__ inline Dual2<R> operator+(Dual2<R> x, Dual2<R> y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
template <typename R> __device__ inline Dual2<R> operator-(Dual2<R> x, Dual2<R> y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
template <typename R> __device__ inline Dual2<R> operator*(Dual2<R> x, Dual2<R> y) {
    return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b};
}
template <typename R> __device__ inline Dual2<R> operator/(Dual2<R> x, Dual2<R> y) {
    if constexpr (sizeof(R) == 4) {  // FP32: one hardware reciprocal for the three quotients
        const R r = recip<R>(y.v), q = x.v * r;
        return {q, (x.a - q * y.a) * r, (x.b - q * y.b) * r};
    } else {
        const R q = x.v / y.v;
        return {q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v};
    }
}

// Humlicek-Kuntz K(x,y) rational regions 1-3 (src/Shapes.f90:60-117), VY = y, V2 = x^2.
// R is float, double or Dual2<float/double> (the coefficients enter as R(constant)).
template <typename R>
__device__ inline R kuntz_rational(int region, R VY, R V2) {
    R num, den;
    if (region == 1) {
        R Y2 = VY * VY;
        R A1 = (R(0.2820948f) + R(0.5641896f) * Y2) * VY;
        R B1 = R(0.5641896f) * VY;
        R A2 = R(0.25f) + Y2 + Y2 * Y2;
        R B2 = Y2 + Y2 - R(1);
        num = A1 + B1 * V2;
        den = A2 + B2 * V2 + V2 * V2;
    } else if (region == 2) {
        R Y2 = VY * VY;
        R A3 = VY * (((R(0.56419f) * Y2 + R(3.10304f)) * Y2 + R(4.65456f)) * Y2 + R(1.05786f));
        R B3 = VY * ((R(1.69257f) * Y2 + R(0.56419f)) * Y2 + R(2.962f));
        R C3 = VY * (R(1.69257f) * Y2 - R(2.53885f));
        R D3 = VY * R(0.56419f);
        R A4 = (((Y2 + R(6)) * Y2 + R(10.5f)) * Y2 + R(4.5f)) * Y2 + R(0.5625f);
        R B4 = ((R(4) * Y2 + R(6)) * Y2 + R(9)) * Y2 - R(4.5f);
        R C4 = R(10.5f) + R(6) * (Y2 - R(1)) * Y2;
        R D4 = R(4) * Y2 - R(6);
        num = ((D3 * V2 + C3) * V2 + B3) * V2 + A3;
        den = (((V2 + D4) * V2 + C4) * V2 + B4) * V2 + A4;
    } else {
        R Y = VY;
        R A5 = ((((((((R(0.564224f) * Y + R(7.55895f)) * Y + R(49.5213f)) * Y + R(204.510f)) * Y +
                     R(581.746f)) * Y + R(1174.8f)) * Y + R(1678.33f)) * Y + R(1629.76f)) * Y +
                 R(973.778f)) * Y + R(272.102f);
        R B5 = ((((((R(2.25689f) * Y + R(22.6778f)) * Y + R(100.705f)) * Y + R(247.198f)) * Y +
                   R(336.364f)) * Y + R(220.843f)) * Y - R(2.34403f)) * Y - R(60.5644f);
        R C5 = ((((R(3.38534f) * Y + R(22.6798f)) * Y + R(52.8454f)) * Y + R(42.5683f)) * Y +
                 R(18.546f)) * Y + R(4.58029f);
        R D5 = ((R(2.25689f) * Y + R(7.56186f)) * Y + R(1.66203f)) * Y - R(0.128922f);
        R E5 = R(0.971457E-3f) + R(0.564224f) * Y;
        R A6 = (((((((((Y + R(13.3988f)) * Y + R(88.2674f)) * Y + R(369.199f)) * Y + R(1074.41f)) * Y +
                     R(2256.98f)) * Y + R(3447.63f)) * Y + R(3764.97f)) * Y + R(2802.87f)) * Y +
                 R(1280.83f)) * Y + R(272.102f);
        R B6 = (((((((R(5) * Y + R(53.5952f)) * Y + R(266.299f)) * Y + R(793.427f)) * Y + R(1549.68f)) * Y +
                   R(2037.31f)) * Y + R(1758.34f)) * Y + R(902.306f)) * Y + R(211.678f);
        R C6 = (((((R(10) * Y + R(80.3928f)) * Y + R(269.292f)) * Y + R(479.258f)) * Y + R(497.302f)) * Y +
                 R(308.186f)) * Y + R(78.866f);
        R D6 = (((R(10) * Y + R(53.5952f)) * Y + R(92.7586f)) * Y + R(55.0293f)) * Y + R(22.0353f);
        R E6 = (R(5) * Y + R(13.3988f)) * Y + R(1.49645f);
        num = (((E5 * V2 + D5) * V2 + C5) * V2 + B5) * V2 + A5;
        den = ((((V2 + E6) * V2 + D6) * V2 + C6) * V2 + B6) * V2 + A6;
    }
    return fdiv<R>(num, den);
}

__constant__ static const float ASYM_U[9] = {1.f, 1.5f, 2.f, 2.5f, 3.f, 3.5f, 4.f, 4.5f, 5.f};
__constant__ static const float ASYM_W[9] = {-0.688f, 0.2667f, 0.6338f, 0.4405f, 0.2529f,
                                             0.1601f, 0.1131f, 0.0853f, 0.068f};

// Lorentz x chi (src/Spectroscopy.f90:202,227).  NEAR: the caller samples within a few
// tenths of a cm-1 of the centre; every chi law is exactly 1 for |X| < 3
// (src/ChiFactors.f90), so chi is evaluated only from there on.
template <typename R, int CHI = -1, bool NEAR = false>
__device__ inline R chilor(R X, const LineShape<R> &p, const ShapeCtl &c) {
    R lor = p.lorc * recip<R>(X * X + p.gl2);
    if constexpr (CHI != 0) {
        const R ax = X < R(0) ? -X : X;
        if (!NEAR || ax >= R(3)) lor *= chi_sel<R, CHI>(c, ax);
    }
    return lor;
}

// the same from ax = |X| (X enters only as X^2 and |X|)
template <typename R, int CHI = -1>
__device__ inline R chilor_abs(R ax, const LineShape<R> &p, const ShapeCtl &c) {
    R lor = p.lorc * recip<R>(ax * ax + p.gl2);
    if constexpr (CHI != 0) lor *= chi_sel<R, CHI>(c, ax);
    return lor;
}

template <typename R>
__device__ inline R doppler(R X, const LineShape<R> &p) {
    R q = X * p.igd;
    return p.dopa * dexp<R>(-((q * q) * R(MARFA_LOG2_F)));
}

// Non-Lorentz part of voigt(X): the region is decided from FP64 VX/VY formed exactly as
// the FP64 form forms them (src/Shapes.f90:49-50), the value in precision R.
template <typename R, bool NEAR = false>
__device__ inline R voigt_regions(double X64, const LineShape<R> &p, const ShapeCtl &c) {
    const double VX = fabs(X64) * p.xs64;
    const double VY = p.vy64;
    const R X = (R)X64;
    if (VX >= 15.0) return chilor<R, -1, NEAR>(X, p, c);
    int region;
    if (VX + VY >= 15.0) region = 1;
    else if (VX + VY >= 5.5) region = 2;
    else if (VX <= 1.0 || VY >= (double)0.02f) region = 3;
    else region = 4;
    if (region < 4) {
        R vx = (R)VX;
        return kuntz_rational<R>(region, (R)VY, vx * vx) * p.normc;
    }
    if (VX > 5.0) {
        R vx = (R)VX;
        return chilor<R, -1, NEAR>(X, p, c) * (R(1) + fdiv<R>(R(1.5f), vx * vx));
    }
    if (VX > (double)MARFA_SQRT14_F) {
        // table index decided in FP64 as the FP64 form decides it
        const int I = (int)(VX / 0.5 - (double)1.00001f);
        R vx = (R)VX;
        R F = R(2) * (R(ASYM_W[I - 1]) * (R(ASYM_U[I]) - vx) + R(ASYM_W[I]) * (vx - R(ASYM_U[I - 1])));
        return doppler<R>(X, p) + fdiv<R>(p.lorc, X * X) * (R(1) + F);
    }
    return doppler<R>(X, p);
}

// out-of-line copy for kernels where the non-Lorentz regions are rare (level-0 deposits).
// The line's fields travel as scalar arguments (VGPRs): a LineShape passed by reference
// needs an address, and the caller would then store every line's parameters to scratch
// memory, whether or not the call is taken.
template <typename R>
__device__ __noinline__ R voigt_slow(double X64, double xs64, double vy64, R lorc, R gl2, R normc, R dopa, R igd,
                                     const ShapeCtl &c) {
    LineShape<R> p;
    p.xs64 = xs64;
    p.vy64 = vy64;
    p.lorc = lorc;
    p.gl2 = gl2;
    p.normc = normc;
    p.dopa = dopa;
    p.igd = igd;
    return voigt_regions<R>(X64, p, c);
}

// Profile value at signed distance X64 (FP64) from the shifted centre.  NEAR: the caller
// samples close to line centres, where the Humlicek-Kuntz regions are the common case,
// so the FP32 form inlines them (the FP64 form keeps them out of line: inlined FP64
// regions double the fine kernel's registers and cut its occupancy to one wave per SIMD).
template <typename R, bool NEAR = false, int CHI = -1>
__device__ inline R shape_eval(double X64, const LineShape<R> &p, const ShapeCtl &c) {
    const R X = (R)X64;
    if (c.shape_id == 1) return chilor<R, CHI, NEAR>(X, p, c);  // lorentzShape
    const R ax = X < R(0) ? -X : X;
    // fast path: clearly in region 0 (pure Lorentz x chi); margin covers the FP32 error of VX
    const R margin = sizeof(R) == 8 ? R(15.0 * (1.0 + 1e-12)) : R(15.0f * (1.0f + 4e-6f));
    if (ax * p.xs >= margin) return chilor<R, CHI, NEAR>(X, p, c);
    if constexpr (NEAR && sizeof(R) == 4) return voigt_regions<R, true>(X64, p, c);
    else return voigt_slow<R>(X64, p.xs64, p.vy64, p.lorc, p.gl2, p.normc, p.dopa, p.igd, c);
}

// ---------------------------------------------------------------- derivatives (adjoint)
// d chi / d|X| of chi_factor: the exponent's slope times chi inside each band, 0 outside.
template <typename R>
__device__ inline R dchi_factor(int chi, R ax, R B1, R B2, R chiv) {
    if (chi == 1) {
        if (ax > R(3)) return (ax <= R(150) ? R(-0.027f) : ax <= R(300) ? R(-0.016f) : R(-0.009f)) * chiv;
    } else if (chi == 2) {
        if (ax >= R(3)) return (ax < R(10) ? R(-0.1) : R(-1.0 / 47.0)) * chiv;
    } else if (chi == 3) {
        if (ax > R(3)) return (ax < R(30) ? -B1 : ax < R(120) ? -B2 : R(-0.0232f)) * chiv;
    }
    return R(0);
}

// value and partial derivatives of one sample with respect to the signed offset X (the
// shifted centre minus the node position), gD and gL; the S(T) derivative is v / S.
template <typename R>
struct SampleGrad {
    R v, dX, dgD, dgL;
    R dT;  // at fixed X, gD, gL: through the Perrin chi factor's temperature coefficients
};

template <typename R>
__device__ inline void chilor_grad(R X, const LineShape<R> &p, const ShapeCtl &c, SampleGrad<R> &o) {
    const R ax = X < R(0) ? -X : X, sg = X < R(0) ? R(-1) : R(1);
    const R inv = recip<R>(X * X + p.gl2);
    const R lor = p.lorc * inv;
    R chi = R(1), dchi = R(0);
    if (c.mol == 2 && c.chi_id != 0) {
        if constexpr (sizeof(R) == 8) {
            chi = chi_factor<double>(c.chi_id, ax, c.B1d, c.B2d);
            dchi = dchi_factor<double>(c.chi_id, ax, c.B1d, c.B2d, chi);
        } else {
            chi = chi_factor<float>(c.chi_id, ax, c.B1, c.B2);
            dchi = dchi_factor<float>(c.chi_id, ax, c.B1, c.B2, chi);
        }
    }
    o.v = lor * chi;
    o.dX = R(-2) * X * lor * inv * chi + lor * dchi * sg;
    o.dgD = R(0);
    o.dgL = p.S / R(MARFA_PI_F) * (X * X - p.gl2) * inv * inv * chi;
    // Perrin (src/ChiFactors.f90:82-122): chi = exp(e(|X|, B1(T), B2(T))), d chi/dT = chi de/dT
    // with e = -B1 (|X| - 3), -27 B1 - B2 (|X| - 30) or -27 B1 - 90 B2 - B3 (|X| - 120)
    o.dT = R(0);
    if (c.mol == 2 && c.chi_id == 3 && ax > R(3)) {
        const R dB1 = (R)c.dB1, dB2 = (R)c.dB2;
        const R de = ax < R(30) ? -dB1 * (ax - R(3)) : ax < R(120) ? -R(27) * dB1 - dB2 * (ax - R(30))
                                                                 : -R(27) * dB1 - R(90) * dB2;
        o.dT = o.v * de;
    }
}

// Same region decisions as shape_eval / voigt_regions, value and derivatives of the branch
// taken (src/Shapes.f90:22-133, src/Spectroscopy.f90:194-261).
template <typename R>
__device__ inline SampleGrad<R> shape_grad(double X64, const LineShape<R> &p, const ShapeCtl &c) {
    SampleGrad<R> o;
    const R X = (R)X64;
    const R ax = X < R(0) ? -X : X, sg = X < R(0) ? R(-1) : R(1);
    const R margin = sizeof(R) == 8 ? R(15.0 * (1.0 + 1e-12)) : R(15.0f * (1.0f + 4e-6f));
    if (c.shape_id == 1 || ax * p.xs >= margin) {
        chilor_grad<R>(X, p, c, o);
        return o;
    }
    const double VX = fabs(X64) * p.xs64, VY = p.vy64;
    if (VX >= 15.0) {
        chilor_grad<R>(X, p, c, o);
        return o;
    }
    const R vx = (R)VX, vy = (R)VY;
    int region;
    if (VX + VY >= 15.0) region = 1;
    else if (VX + VY >= 5.5) region = 2;
    else if (VX <= 1.0 || VY >= (double)0.02f) region = 3;
    else region = 4;
    o.dT = R(0);  // chi enters only the chi-corrected Lorentz and asymptote 1
    if (region < 4) {
        const Dual2<R> K = kuntz_rational<Dual2<R>>(region, Dual2<R>(vy, R(0), R(1)), Dual2<R>(vx * vx, R(1), R(0)));
        o.v = K.v * p.normc;
        o.dX = p.normc * K.a * R(2) * vx * p.xs * sg;
        o.dgD = -(o.v + p.normc * (R(2) * vx * vx * K.a + vy * K.b)) * p.igd;
        o.dgL = p.normc * K.b * p.xs;
        return o;
    }
    if (VX > 5.0) {  // asymptote 1: Lorentz x chi x (1 + 1.5 / VX^2)
        chilor_grad<R>(X, p, c, o);
        const R f = R(1) + fdiv<R>(R(1.5f), vx * vx), df = fdiv<R>(R(-3), vx * vx * vx);
        o.dX = o.dX * f + o.v * df * p.xs * sg;
        o.dgD = o.v * df * (-vx * p.igd);
        o.dgL = o.dgL * f;
        o.dT = o.dT * f;
        o.v = o.v * f;
        return o;
    }
    const R q = X * p.igd;
    const R D = p.dopa * dexp<R>(-((q * q) * R(MARFA_LOG2_F)));
    o.v = D;
    o.dX = D * R(-2) * R(MARFA_LOG2_F) * q * p.igd;
    o.dgD = D * (R(2) * R(MARFA_LOG2_F) * q * q - R(1)) * p.igd;
    o.dgL = R(0);
    if (VX > (double)MARFA_SQRT14_F) {  // asymptote 2: Doppler + Lorentz wing (1 + F(VX))
        const int I = (int)(VX / 0.5 - (double)1.00001f);
        const R F = R(2) * (R(ASYM_W[I - 1]) * (R(ASYM_U[I]) - vx) + R(ASYM_W[I]) * (vx - R(ASYM_U[I - 1])));
        const R dF = R(2) * (R(ASYM_W[I]) - R(ASYM_W[I - 1]));
        const R ix2 = recip<R>(X * X);
        const R T2 = p.lorc * ix2 * (R(1) + F);
        o.v += T2;
        o.dX += fdiv<R>(R(-2) * T2, X) + p.lorc * ix2 * dF * p.xs * sg;
        o.dgD += p.lorc * ix2 * dF * (-vx * p.igd);
        o.dgL = p.S / R(MARFA_PI_F) * ix2 * (R(1) + F);
    }
    return o;
}

// ---------------------------------------------------------------- lane scans by DPP
// Lane moves through the VALU's data-parallel primitives (no LDS round trip, unlike a
// shuffle): dpp_move returns in each lane the value of the lane that CTRL selects, or `old`
// where there is none (outside the 16-lane row for row shifts) or the lane's row is not in
// ROWS.  Controls used: row_shr:n (0x110 + n: lane - n within its row), row_bcast:15
// (0x142: the previous row's last lane), row_bcast:31 (0x143: lane 31), wave_shl:1 (0x130:
// lane + 1 across the wavefront).  Every lane of the wavefront must be active.
template <int CTRL, int ROWS>
__device__ inline int dpp_move(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ inline float dpp_move(float old, float v) {
    return __builtin_bit_cast(float, dpp_move<CTRL, ROWS>(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWS>
__device__ inline double dpp_move(double old, double v) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)dpp_move<CTRL, ROWS>((int)(uint32_t)o, (int)(uint32_t)x);
    const uint32_t hi = (uint32_t)dpp_move<CTRL, ROWS>((int)(uint32_t)(o >> 32), (int)(uint32_t)(x >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// one step of seg_scan: add the selected lane's partial sums when it has the same key
template <int CTRL, int ROWS, typename T, int N>
__device__ inline void seg_step(int key, T (&v)[N]) {
    const bool take = dpp_move<CTRL, ROWS>(-2, key) == key;  // keys are >= -1: -2 never matches
#pragma unroll
    for (int c = 0; c < N; c++) {
        const T t = dpp_move<CTRL, ROWS>(T(0), v[c]);
        v[c] = take ? v[c] + t : v[c];
    }
}

// Key-segmented inclusive sums across the wavefront: afterwards each lane holds, per value,
// the sum over the lanes up to and including itself that have its key.  Keys must be
// non-decreasing along the lanes (scan_key), so each key's lanes are contiguous.  Steps:
// shifts by 1, 2, 4, 8 within the 16-lane rows, then row 0's (row 2's) last lane into row 1
// (row 3), then lane 31 into rows 2 and 3 -- the same partial sums in the same order on
// every run.
template <typename T, int N>
__device__ inline void seg_scan(int key, T (&v)[N]) {
    seg_step<0x111, 0xF>(key, v);
    seg_step<0x112, 0xF>(key, v);
    seg_step<0x114, 0xF>(key, v);
    seg_step<0x118, 0xF>(key, v);
    seg_step<0x142, 0xA>(key, v);
    seg_step<0x143, 0xC>(key, v);
}

// the next lane's key (-3 in lane 63): a lane ends its key's run when this differs
__device__ inline int next_key(int key) { return dpp_move<0x130, 0xF>(-3, key); }

// Key of a lane for a key-segmented lane scan.  Lanes with a value (has) keep their key;
// a lane without one takes the key of the nearest lane before it that has one (its own
// values are zero), or -1 before the first.  Along the lanes with values the keys are
// non-decreasing, but lanes without one can sit between two lanes of one key (a line
// outside the record's window among the sorted candidates); inheriting the key keeps every
// key's lanes contiguous, so the scan's segment ends hold each key's sum exactly once.
// Formed as an inclusive max-scan of (has ? key : -1) across the lanes (as the keys with
// values do not decrease, the maximum up to a lane is the key of the last such lane up to
// it): DPP steps, no LDS round trip.  Every lane of the wavefront must be active.
__device__ inline int scan_key(bool has, int key) {
    int k = has ? key : -1;
    k = max(k, dpp_move<0x111, 0xF>(-1, k));
    k = max(k, dpp_move<0x112, 0xF>(-1, k));
    k = max(k, dpp_move<0x114, 0xF>(-1, k));
    k = max(k, dpp_move<0x118, 0xF>(-1, k));
    k = max(k, dpp_move<0x142, 0xA>(-1, k));
    k = max(k, dpp_move<0x143, 0xC>(-1, k));
    return k;
}

// ---------------------------------------------------------------- accumulators
// Node and record sums are FP64 in both precision forms: an FP32 sample enters an FP64 sum
// exactly, which is more accurate than a compensated FP32 pair held in the same two
// registers and takes fewer operations per add.
struct Acc {
    double s;
    __device__ inline void zero() { s = 0.0; }
    __device__ inline void add(double y) { s += y; }
    __device__ inline double value() const { return s; }
};

}  // namespace marfa
