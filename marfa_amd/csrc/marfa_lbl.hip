// marfa_lbl.hip -- MI355X (gfx950) line-by-line PT-table core behind include/marfa.h.
//
// Pipeline for a batch of atmospheric levels (one stream, no atomics in any sum):
//   1. spectroscopy_kernel  one thread per (line, level): shifted centre, Doppler and
//      Lorentz HWHM, S(T) in FP64 (src/Spectroscopy.f90:38-145, src/LBL.f90:64-73).
//   2. per-level counting sort of the shifted centres on 1/512 cm-1 buckets (the gather
//      needs lines ordered by position; the bucket starts are its index).
//   3. coarse_kernel  one workgroup per record (level, 10 cm-1 subinterval): every line of
//      the record's window (in sorted order) is evaluated at the 21 positions of the
//      1 cm-1 grid and the samples are summed per (cell, edge) in registers, then reduced
//      across the workgroup in a fixed order.  This is the level-0 part of leftLBL / centerLBL /
//      rightLBL (src/LineGridCalc.f90) and carries ~90% of all samples.
//   4. fine_kernel  one workgroup per (record, 1 cm-1 tile): the sums of grids 1..9 and
//      the fine points of the tile gather the lines that deposit into them -- the inverse
//      of the reference's walks, in closed form: on grid k the lines of one walk group
//      deposit into one pair of cells -- from candidates sorted by position (LDS-staged
//      near the tile); each sum has one owner and a fixed order.  The tile then runs the
//      cascade (src/Grids.f90:80-139) in FP64 from its level-0 cell down to the 2,048
//      points of the record it owns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/marfa.h"
#include "lbl_device.h"

using namespace marfa;

// WISO molar masses (src/MolarMasses.f90:140-146), bit patterns of the reference's values
__constant__ static const uint32_t WISO_BITS[124] = {
#include "wiso_table.inc"
};

namespace {

constexpr int BLOCK = 256;
constexpr int TILES = 10;      // 1 cm-1 tiles per 10 cm-1 record
constexpr int NG0 = 32;        // grid 0 of a record: RK1, P[10], M[10], L[10], pad
constexpr int NG13 = 420;      // grids 1..3 of a record: P, mid, L blocks of 20, 40, 80 cells
constexpr int NG45 = 1440;     // grids 4, 5 of a record: P, mid, L blocks of 160, 320 cells
constexpr int KNEAR = 5;       // grids 1..KNEAR are summed by the level-0 kernel
constexpr int NCOARSE = NG0 + NG13 + NG45;  // per record: the level-0 kernel's sums
constexpr int NACC = 3066;     // per tile: 3*(2+4+...+512) nodes of grids 1..9

enum : uint8_t { CLS_NONE = 0, CLS_LEFT = 1, CLS_CENTER = 2, CLS_RIGHT = 3 };

// Sortable integer image of a double (total order, also for negative values) and back.
__host__ __device__ inline uint64_t key_of(double x) {
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    return (b & 0x8000000000000000ULL) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double double_of(uint64_t k) {
    uint64_t b = (k & 0x8000000000000000ULL) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
    return __builtin_bit_cast(double, b);
}

// Kernel arguments shared by the deposit kernels.
struct KArgs {
    int nlev, nrec;            // levels in this batch, subintervals per level
    int lev0;                  // first level of the batch (output / level arrays)
    int64_t lmin, nl;          // line range [lmin, lmin+nl) of the batch's arrays
    const double *nus, *gD, *gL, *S;          // [nlev][nl]
    const uint64_t *skey;                     // [nlev][nl] sorted keys of shifted centres
    const int32_t *sidx;                      // [nlev][nl] line index of sorted entry
    const void *shp;                          // [nlev][nl] LineShape<R> of sorted entries
    const int32_t *bst;                       // [nlev][NB+1] first sorted entry of each bucket
    double bbase;                             // shifted centre of bucket 0's left edge
    int NB;                                   // buckets (1/512 cm-1) per level
    const int64_t *lo, *hi;                   // [nrec] windows (global line indices)
    const double *org;                        // [nrec] subinterval origins
    const double *sscale;                     // [lev] sample scale (power of 2)
    const float *T;                           // [lev]
    const float *outmul;                      // [lev] 1 (ACS) or density (VAC)
    float cutoff;
    int shape_id, chi_id, mol;
    double *coarse;                           // [nlev*nrec][NCOARSE]
    float *out;                               // [lev][nrec][NT] (absolute level index)
    double *out64;                            // same layout in FP64 (instead of out) or null
    double *grids;  // optional: pre-cascade sums of record 0 in the reference's layout
                    // (grids 0..9 as P|mid|L blocks, then RK(1..20481)), sample scale removed
    uint64_t *lim;  // optional: the cascade's limiter choices, [nlev*nrec][TILES][LIMW words]
};

// Buckets of the per-level index of sorted shifted centres (1/512 cm-1 wide).
constexpr double BUCKETS_PER_CM = 512.0;

// Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a contiguous
// range of work items so that neighbouring records (which share 90+% of their line
// windows) are processed under the same L2.  The grid is padded to 8*per blocks.
__device__ inline int xcd_item(int nitems) {
    const int per = (nitems + 7) / 8;
    const int b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}

__device__ inline ShapeCtl make_ctl(const KArgs &a, int lev) {
    ShapeCtl c;
    c.shape_id = a.shape_id;
    c.chi_id = a.chi_id;
    c.mol = a.mol;
    const float T = a.T[lev];
    c.B1 = 0.0888f - 0.16f * expf(-(0.00410f * T));
    c.B2 = 0.0526f * expf(-(0.00152f * T));
    c.B1d = (double)0.0888f - (double)0.16f * exp(-((double)0.00410f * (double)T));
    c.B2d = (double)0.0526f * exp(-((double)0.00152f * (double)T));
    c.dB1 = (double)0.16f * (double)0.00410f * exp(-((double)0.00410f * (double)T));
    c.dB2 = -(double)0.0526f * (double)0.00152f * exp(-((double)0.00152f * (double)T));
    return c;
}

// ------------------------------------------------------------------ 1. spectroscopy
struct SpecArgs {
    int nlev;
    int64_t lmin, nl;
    const double *nu;
    const float *S, *gair, *gself, *elow, *nT, *delta;
    const int32_t *molIso;
    const float *P, *T;          // [nlev] (batch-relative)
    const double *pself;         // [nlev]
    const float *tips;           // [74][501]
    double *nus, *gD, *gL, *ST;  // [nlev][nl]
    uint64_t *keys;              // [nlev][nl] sortable bits of nus
    int32_t *idx;                // [nlev][nl]
    int *emax;                   // [nlev] maximum exponent of S(T) (FP32 form), or null
};

__global__ __launch_bounds__(BLOCK) void spectroscopy_kernel(SpecArgs a) {
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int lev = blockIdx.y;
    int e = -2000;  // exponent of S(T) for the level maximum
    if (j < a.nl) {
        const int64_t i = a.lmin + j;
        const float P = a.P[lev], T = a.T[lev];
        const double dT = T, dP = P, ps = a.pself[lev];
        const int iso = a.molIso[i] / 100;
        const double M = (double)bits_f(WISO_BITS[iso - 1]);
        const double nu = a.nu[i];
        // shifted centre: FP32 product delta*P as src/Spectroscopy.f90:48 forms it, so that
        // every subinterval's line selection is the reference's
        const double nus = nu + (double)(a.delta[i] * P);
        const double gD = ((double)MARFA_DOPC_F * nus) * sqrt(dT / M);
        const double gL = pow((double)MARFA_TREF_F / dT, (double)a.nT[i]) *
                          ((double)a.gair[i] * (dP - ps) + (double)a.gself[i] * ps);
        // TIPS(T) by linear interpolation, src/Spectroscopy.f90:110-114
        const int ti = (int)((dT - 20.0) / 2.0 + 1.0);
        const double wu = (dT - ((double)ti * 2.0 + 18.0)) / 2.0;
        const float *row = a.tips + (iso - 1) * 501;
        const double Q = (1.0 - wu) * (double)row[ti - 1] + wu * (double)row[ti];
        const double c2 = (double)MARFA_C2_F, E = a.elow[i];
        const double tf = (double)row[138] / Q;
        const double bf = exp(-(c2 * E / dT)) / exp(-(c2 * E / (double)MARFA_TREF_F));
        const double ef = (1.0 - exp(-((c2 * nu) / dT))) / (1.0 - exp(-((c2 * nu) / (double)MARFA_TREF_F)));
        const double ST = (double)a.S[i] * tf * bf * ef;
        const int64_t o = (int64_t)lev * a.nl + j;
        a.nus[o] = nus;
        a.gD[o] = gD;
        a.gL[o] = gL;
        a.ST[o] = ST;
        a.keys[o] = key_of(nus);
        a.idx[o] = (int32_t)i;
        if (ST > 0.0) e = ilogb(ST);
    }
    if (a.emax) {
        // per-level maximum exponent of S(T) for the FP32 form's sample scale: block maximum,
        // then an atomic only when it exceeds the value already there (the maximum only
        // grows, so a stale read merely costs an atomic; max is order-independent, so the
        // result is deterministic)
        __shared__ int wmax[BLOCK / 64];
        for (int off = 32; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = e;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < BLOCK / 64; w++) e = max(e, wmax[w]);
            if (e > *(volatile const int *)(a.emax + lev)) atomicMax(a.emax + lev, e);
        }
    }
}

// ------------------------------------------------------------------ 3. coarse (level 0)
// Bit q of a mask marks position q*0.5 cm-1 of the subinterval (q = 0..20).
constexpr uint32_t EVEN = 0x155555u, ODD = 0x0AAAAAu;

// Roles of one line at the 21 positions of grid 0 (bit q: position q/2 cm-1): P, mid and
// L edges of cells 1..10.  Returns false when the line deposits nothing in the record.
__device__ inline bool level0_masks(double nus, double org, double cut, uint32_t &mP, uint32_t &mM,
                                    uint32_t &mL) {
    const double u = nus - org;
    mP = mM = mL = 0;
    if (nus < org) {                 // leftLBL, src/LineGridCalc.f90:21-33: cells 1..10 (d >= 1)
        const double d = -u;         // or 2..10 (finer grids cover [0, 1])
        if (d > cut) return false;
        if (d >= 1.0) { mP = EVEN & 0x7FFFFu; mM = ODD; mL = EVEN & ~1u; }
        else { mP = EVEN & 0x7FFFCu; mM = ODD & ~2u; mL = EVEN & ~5u; }
        return true;
    }
    if (nus > org + 10.0) {          // rightLBL, :508-518: cells 10..1 or 9..1
        const double dr = u - 10.0;
        if (dr >= cut) return false;
        if (dr >= 1.0) { mP = EVEN & 0x7FFFFu; mM = ODD; mL = EVEN & ~1u; }
        else { mP = EVEN & 0x1FFFFu; mM = ODD & 0x3FFFFu; mL = EVEN & 0x7FFFCu; }
        return true;
    }
    if (!(u < 10.0)) return false;   // centerLBL :192 (a line at the right edge adds nothing)
    // centerLBL level-0 walks (:201-214 from the left edge, :343-358 from the right edge):
    // cells 1..e0 and 11-e0'..10 with e0 = floor(u) - 1, e0' = floor(10 - u) - 1
    const int e0 = u >= 2.0 ? (int)floor(u) - 1 : 0;
    const double v = 10.0 - u;
    const int e0r = v >= 2.0 ? (int)floor(v) - 1 : 0;
    if (e0 > 0) {
        mP = EVEN & ((1u << (2 * e0 - 1)) - 1u);
        mM = ODD & ((1u << (2 * e0)) - 1u);
        mL = EVEN & ((1u << (2 * e0 + 1)) - 1u) & ~1u;
    }
    if (e0r > 0) {
        const int qs = 20 - 2 * e0r;
        mP |= EVEN & ~((1u << qs) - 1u) & 0x7FFFFu;
        mM |= ODD & ~((1u << qs) - 1u);
        mL |= EVEN & ~((1u << (qs + 1)) - 1u);
    }
    return true;
}

// chi bands of chi_factor_sel for Tonkov (CHI 1) and Pollack (CHI 2): band index of |X| and
// the band's law coef * exp(-(rate |X|)) (band 0: chi = 1)
template <typename R, int CHI>
__device__ inline int chi_band(R ax) {
    if constexpr (CHI == 1) return ax > R(3) ? (ax <= R(150) ? 1 : ax <= R(300) ? 2 : 3) : 0;
    else if constexpr (CHI == 3) return ax > R(3) ? (ax < R(30) ? 1 : ax < R(120) ? 2 : 3) : 0;
    else return ax >= R(3) ? (ax < R(10) ? 1 : 2) : 0;
}
template <typename R, int CHI>
__device__ inline void chi_law(int b, R &coef, R &rate) {
    if constexpr (CHI == 1) {
        coef = b == 0 ? R(1) : b == 1 ? R(1.084f) : b == 2 ? R(0.208f) : R(0.025f);
        rate = b == 0 ? R(0) : b == 1 ? R(0.027f) : b == 2 ? R(0.016f) : R(0.009f);
    } else {
        coef = b == 0 ? R(1) : b == 1 ? R(1.35f) : R(0.614f);
        rate = b == 0 ? R(0) : b == 1 ? R(0.1) : R(1.0 / 47.0);
    }
}

// the exponent's slope of chi in band b (Perrin, src/ChiFactors.f90:82-122: the level's
// temperature coefficients B1, B2 and the constant 0.0232)
template <typename R, int CHI>
__device__ inline R chi_rate(int b, const ShapeCtl &c) {
    if constexpr (CHI == 3) return b <= 1 ? (R)c.B1 : b == 2 ? (R)c.B2 : R(0.0232f);
    else {
        R coef, rate;
        chi_law<R, CHI>(b, coef, rate);
        return rate;
    }
}

// Grids 1..5 of a record from one line per lane (src/LineGridCalc.f90: the walks of
// centerLBL on levels 1..5, and the cells next to the record's edges that leftLBL /
// rightLBL fill on levels k <= m).  For a centre line at u = nu* - org the left walk on grid
// k covers cells 2g-1 .. n-1, n = floor(u 2^k), g = n >> 1: 3 positions (n even) or 5 (n
// odd) of the pair (2g-1, 2g) from X = u - (4g - 4 + e) H_k/2; the right walk is the mirror
// image in v = 10 - u (pair NT_k/2 + 1 - g from the left).  A left wing with d = -u < H_{k-1}
// fills pair 1 from X = u + (e - 4) H_k/2 (5 positions when d >= H_k), a right wing with
// d = u - 10 < H_{k-1} the last pair from X = d + (4 - e) H_k/2.  Along lanes holding lines in
// sorted order, the pair index of each (grid, walk) does not decrease, so a key-segmented
// lane scan sums each pair's values and the last lane of a segment adds them to the pair's
// six slots (P, mid, L blocks of 10 2^k cells per grid): grids 1..3 in g13 (FP64), grids 4, 5
// in g45 (precision R: a slot there takes the segments of a few chunks at most).  Every lane
// of the wavefront calls this (the scans); ok: the lane holds a line of the record's window
// that deposits.
template <typename R, int CHI>
__device__ inline void near_grids_lanes(bool ok, double nus, double org, const LineShape<R> &p, const ShapeCtl &ctl,
                                        double *g13, R *g45) {
    const double u = nus - org;
    // classes as the deposit dispatch takes them (src/LBL.f90:81-98 on the shifted centre)
    const bool left = ok && nus < org, right = ok && nus > org + 10.0, centre = ok && !left && !right && u < 10.0;
    // a centre line's last position on grid k of a walk, w - (n_k - 1) H_k, is the first of
    // that walk on grid k + 1 (n_{k+1} >> 1 = n_k), formed by the same FP64 operations from
    // the same w: its value is carried (per walk) instead of evaluated again
    R carry0 = R(0), carry1 = R(0);
    bool has0 = false, has1 = false;
#pragma unroll 1
    for (int k = 1; k <= KNEAR; k++) {
        const double Hk = hk(k), hq = 0.5 * Hk;
        const int ntk = 10 << k, npair = ntk >> 1, base = 30 * ((1 << k) - 2) - (k > 3 ? NG13 : 0);
#pragma unroll 1
        for (int side = 0; side < 2; side++) {  // 0: left walk (and right wings), 1: right walk (left wings)
            int key = -1;
            bool five = false, plus = false;
            double B = 0.0, C = 0.0;
            if (centre) {
                const double w = side ? 10.0 - u : u;
                const int nn = (int)floor(__builtin_ldexp(w, k));
                const int gi = nn >> 1;
                if (gi >= 1) {
                    key = side ? npair - gi : gi - 1;
                    five = nn & 1;
                    B = w;
                    C = (double)(4 * gi - 4);
                }
            } else if (side == 0 && right && u - 10.0 < 2.0 * Hk) {
                key = npair - 1;
                five = u - 10.0 >= Hk;
                B = u - 10.0;
                C = -4.0;
            } else if (side == 1 && left && -u < 2.0 * Hk) {
                key = 0;
                five = -u >= Hk;
                B = u;
                C = -4.0;
                plus = true;
            }
            R v[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};  // e = 0..4 and e = 2 once more (n odd)
            if (key >= 0) {
                // position e of the walk: X = B -+ (C + e) H_k / 2
                auto at = [&](double e) {
                    const double m = (C + e) * hq;
                    return shape_eval<R, true, CHI>(plus ? B + m : B - m, p, ctl);
                };
                if (centre && (side ? has1 : has0)) v[0] = side ? carry1 : carry0;
                else v[0] = at(0.0);
                v[1] = at(1.0);
                v[2] = at(2.0);
                if (five) {
                    v[5] = v[2];
                    v[3] = at(3.0);
                    v[4] = at(4.0);
                }
            }
            if (centre) {
                const R last = five ? v[4] : v[2];
                if (side) {
                    carry1 = last;
                    has1 = key >= 0;
                } else {
                    carry0 = last;
                    has0 = key >= 0;
                }
            }
            // slots P1 M1 L1 P2 M2 L2 of the pair (right walk: the mirror image)
            R sv[6] = {side ? v[4] : v[0], side ? v[3] : v[1], side ? v[5] : v[2],
                       side ? v[2] : v[5], side ? v[1] : v[3], side ? v[0] : v[4]};
            key = scan_key(key >= 0, key);
            seg_scan(key, sv);
            const R s0 = sv[0], s1 = sv[1], s2 = sv[2], s3 = sv[3], s4 = sv[4], s5 = sv[5];
            if (key >= 0 && next_key(key) != key) {
                const int c = 2 * key;  // 0-based cells c, c + 1 of grid k
                if (k <= 3) {
                    g13[base + c] += (double)s0;
                    g13[base + ntk + c] += (double)s1;
                    g13[base + 2 * ntk + c] += (double)s2;
                    g13[base + c + 1] += (double)s3;
                    g13[base + ntk + c + 1] += (double)s4;
                    g13[base + 2 * ntk + c + 1] += (double)s5;
                } else {
                    g45[base + c] += s0;
                    g45[base + ntk + c] += s1;
                    g45[base + 2 * ntk + c] += s2;
                    g45[base + c + 1] += s3;
                    g45[base + ntk + c + 1] += s4;
                    g45[base + 2 * ntk + c + 1] += s5;
                }
            }
        }
    }
}

// The 21 grid-0 nodes of a far wing in the FP64 form, |X_q| = base + q step, in order:
// use(q, 1 / (X_q^2 + gL^2)), with one division per pair of nodes (1/a = b / (ab): a few
// ulps instead of a half; the denominators lie in [1, ~1e5], so the product cannot
// overflow).
template <typename R, typename F>
__device__ inline void far_nodes(R base, R step, R gl2, F use) {
#pragma unroll
    for (int q = 0; q <= 20; q += 2) {
        const R ax0 = q == 0 ? base : fma(step, (R)q, base);
        const R d0 = ax0 * ax0 + gl2;
        if (q == 20) {
            use(q, recip<R>(d0));
        } else {
            const R ax1 = fma(step, (R)(q + 1), base);
            const R d1 = ax1 * ax1 + gl2;
            const R r = recip<R>(d0 * d1);
            use(q, d1 * r);
            use(q + 1, d0 * r);
        }
    }
}

// The FP32 form's far wing of one line into per-node run sums: run[q] += F / d_q, d_q =
// X_q^2 + gL^2, |X_q| = base + q step, with one hardware reciprocal per pair of nodes:
// F / d_q = (F / (d_q d_q')) d_q'.  F carries a factor 2^32 (the caller's), which keeps
// F / (d_q d_q') a normal number for every F the records hold (d_q d_q' <= ~5e9 at the
// largest cutoffs in use, d >= 1).
__device__ inline void far_run(float base, float step, float gl2, float F, float (&run)[21]) {
#pragma unroll
    for (int q = 0; q <= 20; q += 2) {
        const float ax0 = q == 0 ? base : fmaf(step, (float)q, base);
        const float d0 = ax0 * ax0 + gl2;
        if (q == 20) {
            run[q] = fmaf(F, recip<float>(d0), run[q]);
        } else {
            const float ax1 = fmaf(step, (float)(q + 1), base);
            const float d1 = ax1 * ax1 + gl2;
            const float fr = F * recip<float>(d0 * d1);
            run[q] = fmaf(fr, d1, run[q]);
            run[q + 1] = fmaf(fr, d0, run[q + 1]);
        }
    }
}

// One workgroup per record.  Threads take the level's sorted line records (shifted centre,
// line index and the precomputed profile parameters) in order, so the 64 lines of a
// wavefront lie within a fraction of a cm-1 and take the same branches.  A line at least
// 1 cm-1 outside the subinterval whose nearest node is on the Lorentz branch deposits
// into all 21 nodes at |X| = d + q/2 (left wing) or d + (20-q)/2 (right wing), taken in
// precision R from d >= 1 (relative error ~1 ulp).  When chi keeps one band over those
// nodes (Tonkov / Pollack), chi(|X| +- 1/2) = chi(|X|) exp(-+rate/2): chi is evaluated at
// the first node only, and the ratio to the other nodes is applied to the node sums (FP32
// form: per run of lines, from a table) or carried node to node (FP64 form).  The lines
// closer than 1 cm-1 take the masked per-node path with X = u - q/2.  The FP32 form runs
// at 3 waves per SIMD, the FP64 form at 2.  r: the record (level-batch index x records +
// subinterval).
template <typename R, int CHI>
__device__ __forceinline__ void coarse_item(const KArgs &a, int r) {
    constexpr bool kRuns = sizeof(R) == 4;
    // chi laws whose bands read coef exp(-rate |X|): Tonkov and Pollack (both forms), Perrin
    // (FP32 form; its rates are the level's)
    constexpr bool kBands = CHI == 1 || CHI == 2 || (kRuns && CHI == 3);
    const int lev = r / a.nrec, s = r % a.nrec;
    const ShapeCtl ctl = make_ctl(a, a.lev0 + lev);
    // FP64 form: exp(-rate_b step), step = +1/2 (even entries), -1/2.  FP32 form: the run
    // keys' node factors 2^-32 exp(-rate_b step q) (key 2b + 0 / 1: band b, left / right
    // wing; the last key: no chi)
    constexpr int NKEY = kBands ? 9 : 1;
    __shared__ double rho_tab[8];
    __shared__ double node_fac[NKEY][21];
    if constexpr (kRuns) {
        for (int t = threadIdx.x; t < NKEY * 21; t += BLOCK) {
            const int key = t / 21, q = t % 21;
            double f = 0x1p-32;
            if (key < NKEY - 1) f *= exp(-(double)chi_rate<R, CHI>(key >> 1, ctl) * ((key & 1) ? -0.5 : 0.5) * q);
            node_fac[key][q] = f;
        }
    } else if (kBands && threadIdx.x < 8) {
        R coef, rate;
        chi_law<R, CHI>(threadIdx.x >> 1, coef, rate);
        rho_tab[threadIdx.x] = exp(-(double)rate * ((threadIdx.x & 1) ? -0.5 : 0.5));
    }
    __syncthreads();
    const double org = a.org[s];
    const int64_t lo = a.lo[s], hi = a.hi[s];
    // line indices are int32: the window test as one unsigned compare
    const int32_t lo32 = (int32_t)lo;
    const uint32_t span = hi > lo ? (uint32_t)(hi - lo) : 0u;
    const double cut = (double)a.cutoff;
    const bool nochi = CHI == 0 || ctl.mol != 2;
    const LineShape<R> *shp = reinterpret_cast<const LineShape<R> *>(a.shp) + (int64_t)lev * a.nl;
    const int32_t *bst = a.bst + (int64_t)lev * (a.NB + 1);
    // sorted lines with shifted centres in [org - cut, org + 10 + cut] (bucket superset; the
    // window and deposit tests below are exact)
    int J0, J1;
    {
        const double x0 = (org - cut - a.bbase) * BUCKETS_PER_CM, x1 = (org + 10.0 + cut - a.bbase) * BUCKETS_PER_CM;
        const int b0 = (int)fmin(fmax(floor(x0) - 1.0, 0.0), (double)a.NB);
        const int b1 = (int)fmin(fmax(floor(x1) + 2.0, 0.0), (double)a.NB);
        J0 = bst[b0];
        J1 = b1 > b0 ? bst[b1] : J0;
    }

    // The lines at least 1 cm-1 outside the subinterval (every sorted line before Ja has
    // u < -1, every line from Jb on u > 11: bucket bounds with one bucket of margin) deposit
    // into every node with every role, so they are summed per position q first (21 sums);
    // the lines in between -- centre lines and wings closer than 1 cm-1 -- then add with
    // their roles into the 31 node sums.  The two phases keep 21 and 31 sums live, not both.
    int Ja, Jb;
    {
        const double xa = (org - 1.0 - a.bbase) * BUCKETS_PER_CM, xb = (org + 11.0 - a.bbase) * BUCKETS_PER_CM;
        const int ba = (int)fmin(fmax(floor(xa) - 1.0, 0.0), (double)a.NB);
        const int bb = (int)fmin(fmax(floor(xb) + 2.0, 0.0), (double)a.NB);
        Ja = min(max(bst[ba], J0), J1);
        Jb = min(max(bst[bb], Ja), J1);
    }
    const R margin = sizeof(R) == 8 ? R(15.0 * (1.0 + 1e-12)) : R(15.0f * (1.0f + 4e-6f));

    Acc pos[21];
#pragma unroll
    for (int q = 0; q <= 20; q++) pos[q].zero();
    // FP32 form: the far lines without chi, and those whose chi keeps one band over the
    // subinterval, are first summed per node in FP32 over runs of at most FAR_RUN lines of
    // the thread with one key (band and wing side): a line adds F / d_q with F = 2^32 lorc
    // chi(|X_0|), and a run enters the FP64 node sums times the key's node factor 2^-32
    // exp(-+rate q / 2) = chi(|X_q|) / chi(|X_0|) (positive terms: relative error about
    // FAR_RUN + 3 ulp) -- an FP32 multiply-add per sample instead of a conversion, an FP64
    // multiply-add and the FP64 chi carry
    constexpr int FAR_RUN = 8;
    R run[21];
    int nrun = 0, rkey = NKEY - 1;
#pragma unroll
    for (int q = 0; q <= 20; q++) run[q] = R(0);
    auto flush = [&]() {
#pragma unroll
        for (int q = 0; q <= 20; q++) {
            pos[q].add((double)run[q] * (NKEY == 1 ? 0x1p-32 : node_fac[rkey][q]));
            run[q] = R(0);
        }
        nrun = 0;
    };
    const int nleft = Ja - J0, nfar = nleft + (J1 - Jb);
    auto far_j = [&](int v) { return v < nleft ? J0 + v : Jb + (v - nleft); };
    // The far path reads five fields of a line record (centre, index and the three Lorentz
    // parameters); those of the next two lines are loaded while the current one is
    // evaluated.  The whole record is read only by the per-node fallback.
    struct Far {
        double nus;
        int32_t idx;
        R xs, lorc, gl2;
    };
    // FP32 form: the far records after the batch's line records
    const FarRec *farr = reinterpret_cast<const FarRec *>(reinterpret_cast<const LineShape<R> *>(a.shp) +
                                                          (int64_t)a.nlev * a.nl) + (int64_t)lev * a.nl;
    auto far_rec = [&](int j) {  // the fields of sorted line j
        Far f;
        if constexpr (kRuns) {
            const FarRec q = farr[j];
            f.nus = q.nus;
            f.idx = q.idx;
            f.xs = q.xs;
            f.lorc = q.lorc;
            f.gl2 = q.gl2;
        } else {
            const LineShape<R> &q = shp[j];
            f.nus = q.nus;
            f.idx = q.idx;
            f.xs = q.xs;
            f.lorc = q.lorc;
            f.gl2 = q.gl2;
        }
        return f;
    };
    auto far_at = [&](int v) { return far_rec(far_j(v)); };
    // FP32 form: does chi keep one band b over the 21 nodes?  |X| is monotone in q, so the
    // bands of nodes 0 and 20 (as the nodes are formed) decide; Pollack: band 2 from 10 cm-1
    // on, and its other bands are narrower than the 10 cm-1 the nodes span
    auto one_band = [&](bool lw, R base, R step, int &b) {
        const R ax20 = fma(step, R(20), base);
        if constexpr (CHI == 2) {
            b = 2;
            return (lw ? base : ax20) >= R(10);
        } else {
            b = chi_band<R, CHI>(base);
            return b == chi_band<R, CHI>(ax20);
        }
    };
    int vslow0 = 0x7fffffff, vslow1 = -1;  // the FP32 form's lines for the second pass
    // one far line v of the thread (its fields f, loaded two lines ahead); LW: the lines
    // before Ja (v < nleft) lie left of the subinterval, those from Jb right
    auto far_line = [&](int v, const Far &f, auto LWc) __attribute__((always_inline)) {
        constexpr bool lw = decltype(LWc)::value;
        const int j = lw ? J0 + v : Jb + (v - nleft);
        if ((uint32_t)(f.idx - lo32) >= span) return;  // the record's window (unshifted centres)
        const double u = f.nus - org;
        // leftLBL / rightLBL cutoff tests (src/LineGridCalc.f90:23, :508-509)
        const double dmin = lw ? -u : u - 10.0;
        if (lw ? dmin > cut : dmin >= cut) return;
        const R base = lw ? (R)dmin : (R)(dmin + 10.0);  // |X| at node 0
        const R step = lw ? R(0.5) : R(-0.5);
        // every node, every role: when even the nearest node, |X| >= dmin, is on the Lorentz
        // branch (|X| xs >= margin as shape_eval tests it; rounding is monotone, so every
        // node is), the nodes are summed without per-sample branch tests
        const bool lor = ctl.shape_id == 1 || (R)dmin * f.xs >= margin;
        if constexpr (kRuns) {
            // one chi law (or none) over the 21 nodes: far_run into the runs; the lines that
            // need chi or the shape per node are left to the second pass below (the thread's
            // range [vslow0, vslow1] of them)
            int key = NKEY - 1;
            R F = f.lorc * R(0x1p32f);
            bool one_law = lor && nochi;
            if constexpr (kBands) {
                int b;
                if (lor && !nochi && one_band(lw, base, step, b)) {
                    R chi0;
                    if constexpr (CHI == 3) {
                        chi0 = chi_factor<R>(3, base, ctl.B1, ctl.B2);
                    } else {
                        R coef, rate;
                        chi_law<R, CHI>(b, coef, rate);
                        chi0 = coef * dexp<R>(-(rate * base));
                    }
                    F = f.lorc * (R(0x1p32f) * chi0);
                    key = 2 * b + (lw ? 0 : 1);
                    one_law = true;
                }
            }
            if (!one_law) {
                vslow0 = min(vslow0, v);
                vslow1 = v;
                return;
            }
            if ((key != rkey && nrun > 0) || nrun == FAR_RUN) flush();
            rkey = key;
            nrun++;
            far_run(base, step, f.gl2, F, run);
            return;
        }
        if (lor) {
            // the FP64 form: the Lorentz numerator (times chi) is one FP64 factor
            if (nochi) {
                const double lorc = (double)f.lorc;
                far_nodes<R>(base, step, f.gl2, [&](int q, R inv) { pos[q].add((double)inv * lorc); });
                return;
            }
            if constexpr (kBands && !kRuns) {
                const int b = chi_band<R, CHI>((R)dmin);
                if (b == chi_band<R, CHI>((R)(dmin + 10.0))) {
                    R coef, rate;
                    chi_law<R, CHI>(b, coef, rate);
                    double lchi = (double)f.lorc * (double)(coef * dexp<R>(-(rate * base)));
                    const double rho = rho_tab[2 * b + (lw ? 0 : 1)];
                    far_nodes<R>(base, step, f.gl2, [&](int q, R inv) {
                        if (q > 0) lchi *= rho;
                        pos[q].add((double)inv * lchi);
                    });
                    return;
                }
            }
            LineShape<R> pl;
            pl.lorc = f.lorc;
            pl.gl2 = f.gl2;
#pragma unroll
            for (int q = 0; q <= 20; q++) {
                const R ax = q == 0 ? base : fma(step, (R)q, base);
                pos[q].add((double)chilor_abs<R, CHI>(ax, pl, ctl));
            }
            return;
        }
        const LineShape<R> p = shp[j];
#pragma unroll
        for (int q = 0; q <= 20; q++) pos[q].add((double)shape_eval<R, false, CHI>(u - 0.5 * q, p, ctl));
    };
    // the thread's lines v = tid, tid + BLOCK, ... of [vbeg, vend) on one side, two per trip, so
    // that each prefetched set of fields is consumed and reloaded in place
    auto far_range = [&](int vbeg, int vend, auto LWc) __attribute__((always_inline)) {
        constexpr bool lw = decltype(LWc)::value;
        auto rec = [&](int v) { return far_rec(lw ? J0 + v : Jb + (v - nleft)); };
        Far n1{}, n2{};
        if (vbeg < vend) n1 = rec(vbeg);
        if (vbeg + BLOCK < vend) n2 = rec(vbeg + BLOCK);
        for (int v = vbeg; v < vend; v += 2 * BLOCK) {
            far_line(v, n1, LWc);
            if (v + 2 * BLOCK < vend) n1 = rec(v + 2 * BLOCK);
            if (v + BLOCK >= vend) break;
            far_line(v + BLOCK, n2, LWc);
            if (v + 3 * BLOCK < vend) n2 = rec(v + 3 * BLOCK);
        }
    };
    // the same lines in the same order as one pass over [0, nfar): the left range, then the
    // right range from the thread's first v >= nleft
    {
        const int tid = (int)threadIdx.x;
        far_range(tid, nleft, std::true_type{});
        const int vr = tid >= nleft ? tid : tid + (nleft - tid + BLOCK - 1) / BLOCK * BLOCK;
        far_range(vr, nfar, std::false_type{});
    }

    if (kRuns && nrun > 0) flush();
    if constexpr (kRuns) {
        // second pass of the FP32 form: the far lines with chi over several bands (chi per
        // node) or off the Lorentz branch (the shape per node), straight into the FP64 sums
        for (int v = vslow0; v <= vslow1; v += BLOCK) {
            const Far f = far_at(v);
            if ((uint32_t)(f.idx - lo32) >= span) continue;
            const double u = f.nus - org;
            const bool lw = v < nleft;
            const double dmin = lw ? -u : u - 10.0;
            if (lw ? dmin > cut : dmin >= cut) continue;
            const R base = lw ? (R)dmin : (R)(dmin + 10.0);
            const R step = lw ? R(0.5) : R(-0.5);
            const bool lor = ctl.shape_id == 1 || (R)dmin * f.xs >= margin;
            if (lor) {
                if (nochi) continue;
                if constexpr (kBands) {
                    int b;
                    if (one_band(lw, base, step, b)) continue;
                }
                LineShape<R> pl;  // chilor_abs reads the Lorentz fields only
                pl.lorc = f.lorc;
                pl.gl2 = f.gl2;
#pragma unroll
                for (int q = 0; q <= 20; q++) {
                    const R ax = q == 0 ? base : fma(step, (R)q, base);
                    pos[q].add((double)chilor_abs<R, CHI>(ax, pl, ctl));
                }
            } else {
                const LineShape<R> p = shp[far_j(v)];
#pragma unroll
                for (int q = 0; q <= 20; q++) pos[q].add((double)shape_eval<R, false, CHI>(u - 0.5 * q, p, ctl));
            }
        }
    }
    // node q: position q/2 of the subinterval; even q is L of cell q/2 and P of cell q/2+1
    Acc aR, aP[10], aM[10], aL[10];
    aR = pos[0];
#pragma unroll
    for (int c = 0; c < 10; c++) {
        aP[c] = pos[2 * c];
        aM[c] = pos[2 * c + 1];
        aL[c] = pos[2 * c + 2];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // near lines (centre lines and the wings within 1 cm-1 of an edge), first pass: their
    // masked grid-0 samples into the 31 node sums, one line per thread
    for (int j = Ja + (int)threadIdx.x; j < Jb; j += BLOCK) {
        const LineShape<R> p = shp[j];
        if (p.idx < lo || p.idx >= hi) continue;
        const double nus = p.nus;
        const double u = nus - org;
        uint32_t mP = 0, mM = 0, mL = 0;
        if (!level0_masks(nus, org, cut, mP, mM, mL)) continue;
        const uint32_t need = mP | mM | mL | 1u;
#pragma unroll
        for (int q = 0; q <= 20; q++) {
            // positions no role needs are moved far out, onto the cheap Lorentz branch
            const double X = ((need >> q) & 1u) ? u - 0.5 * q : 1.0e4;
            const double val = (double)shape_eval<R, false, CHI>(X, p, ctl);
            if (q == 0) aR.add(val);
            if (q & 1) {
                aM[q >> 1].add(((mM >> q) & 1u) ? val : 0.0);
            } else {
                if (q <= 18) aP[q >> 1].add(((mP >> q) & 1u) ? val : 0.0);
                if (q >= 2) aL[(q >> 1) - 1].add(((mL >> q) & 1u) ? val : 0.0);
            }
        }
    }

    // fixed-order reduction of the node sums: wave butterflies, then the 4 waves in order,
    // in FP64
    __shared__ double red[BLOCK / 64][31];
    {
        double vals[31];
        vals[0] = aR.value();
#pragma unroll
        for (int c = 0; c < 10; c++) {
            vals[1 + c] = aP[c].value();
            vals[11 + c] = aM[c].value();
            vals[21 + c] = aL[c].value();
        }
#pragma unroll
        for (int k = 0; k < 31; k++) {
            double x = vals[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            if (lane == 0) red[wv][k] = x;
        }
    }

    // Second pass: grids 1..5 of the record from the same near lines, one line per lane
    // (near_grids_lanes, the Voigt regions inlined: no node sums are live here); per
    // wavefront, partial node sums in LDS, added to by the last lane of each pair's segment
    // after a lane scan; the wavefronts' partials are summed in wave order at the end.
    __shared__ double g13[BLOCK / 64][NG13];
    __shared__ R g45[BLOCK / 64][NG45];
    for (int e = lane; e < NG13; e += 64) g13[wv][e] = 0.0;
    for (int e = lane; e < NG45; e += 64) g45[wv][e] = R(0);
    LineShape<R> nxt;  // the next centre line's record, loaded while the current one is evaluated
    if (Ja + (int)threadIdx.x < Jb) nxt = shp[Ja + threadIdx.x];
    // the wavefront's lanes take consecutive lines, all lanes every pass (the scans)
    for (int c0 = Ja + 64 * wv; c0 < Jb; c0 += BLOCK) {
        const int j = c0 + lane;
        const LineShape<R> p = nxt;
        if (j + BLOCK < Jb) nxt = shp[j + BLOCK];
        bool ok = j < Jb && !(p.idx < lo || p.idx >= hi);
        const double nus = p.nus;
        uint32_t mP = 0, mM = 0, mL = 0;
        if (ok) ok = level0_masks(nus, org, cut, mP, mM, mL);
        near_grids_lanes<R, CHI>(ok, nus, org, p, ctl, g13[wv], g45[wv]);
    }

    __syncthreads();
    if (threadIdx.x < 31) {
        double x = red[0][threadIdx.x];
        for (int w = 1; w < BLOCK / 64; w++) x += red[w][threadIdx.x];
        a.coarse[(int64_t)r * NCOARSE + threadIdx.x] = x;
    }
    for (int e = threadIdx.x; e < NG13; e += BLOCK) {
        double x = g13[0][e];
        for (int w = 1; w < BLOCK / 64; w++) x += g13[w][e];
        a.coarse[(int64_t)r * NCOARSE + NG0 + e] = x;
    }
    for (int e = threadIdx.x; e < NG45; e += BLOCK) {
        double x = (double)g45[0][e];
        for (int w = 1; w < BLOCK / 64; w++) x += (double)g45[w][e];
        a.coarse[(int64_t)r * NCOARSE + NG0 + NG13 + e] = x;
    }
}

// One job's level batch: one workgroup per record, each XCD a contiguous range of records.
template <typename R, int CHI>
__global__ __launch_bounds__(BLOCK, sizeof(R) == 4 ? 4 : 2) void coarse_kernel(KArgs a) {
    const int nitems = a.nlev * a.nrec;
    const int r = xcd_item(nitems);
    if (r >= nitems) return;
    coarse_item<R, CHI>(a, r);
}

// ------------------------------------------------------------------ 4. fine + cascade
#include "lbl_fine.h"
#include "lbl_adjoint.h"
#include "lbl_sort.h"
#include "lbl_loss.h"
#include "lbl_multi.h"

// Fine-kernel variants (the batch autotuner's choices): centre lines staged in LDS per tile
// {512, 1024, 2048} x contiguous record ranges per XCD {on, off}.  Variant 1 is the default.
constexpr int NVARIANTS = 6;
constexpr int VARIANT_CAP[NVARIANTS] = {512, 1024, 2048, 512, 1024, 2048};

template <typename R>
void launch_fine(int variant, dim3 g, hipStream_t st, const KArgs &ka) {
    switch (variant) {
    case 0: hipLaunchKernelGGL((fine_kernel<R, 512, true>), g, dim3(BLOCK), 0, st, ka); break;
    case 2: hipLaunchKernelGGL((fine_kernel<R, 2048, true>), g, dim3(BLOCK), 0, st, ka); break;
    case 3: hipLaunchKernelGGL((fine_kernel<R, 512, false>), g, dim3(BLOCK), 0, st, ka); break;
    case 4: hipLaunchKernelGGL((fine_kernel<R, 1024, false>), g, dim3(BLOCK), 0, st, ka); break;
    case 5: hipLaunchKernelGGL((fine_kernel<R, 2048, false>), g, dim3(BLOCK), 0, st, ka); break;
    default: hipLaunchKernelGGL((fine_kernel<R, 1024, true>), g, dim3(BLOCK), 0, st, ka); break;
    }
}

// ------------------------------------------------------------------ sample counts
__global__ __launch_bounds__(BLOCK) void count_kernel(KArgs a, unsigned long long *counts) {
    // counts[2r] = distinct samples of the reference's geometry in record r,
    // counts[2r+1] = the part of them the level-0 kernel evaluates: its grid-0 evaluations
    // and the samples of grids 1..KNEAR
    const int nitems = a.nlev * a.nrec;
    const int r = blockIdx.x;
    if (r >= nitems) return;
    const int lev = r / a.nrec, s = r % a.nrec;
    const double org = a.org[s];
    const double cut = (double)a.cutoff;
    const int64_t base = (int64_t)lev * a.nl - a.lmin;
    unsigned long long cnt = 0, c0 = 0;
    for (int64_t i = a.lo[s] + threadIdx.x; i < a.hi[s]; i += BLOCK) {
        const double nus = a.nus[base + i], u = nus - org;
        uint32_t mP, mM, mL;
        if (level0_masks(nus, org, cut, mP, mM, mL)) c0 += __builtin_popcount(mP | mM | mL | 1u);
        if (nus < org) {
            const double d = -u;
            if (d > cut) continue;
            int m = 0;
            while (m <= 9 && d < hk(m)) m++;
            cnt += 21 + 2 * m;
            c0 += 2 * min(m, KNEAR);  // the near grids' cells next to the edge
        } else if (nus > org + 10.0) {
            const double dr = u - 10.0;
            if (dr >= cut) continue;
            int m = 0;
            while (m <= 9 && dr < hk(m)) m++;
            cnt += 21 + 2 * m;
            c0 += 2 * min(m, KNEAR);
        } else if (u < 10.0) {
            const double v = 10.0 - u;
            long long cells = 0;
            int ep = 0, epr = 0;
            for (int k = 0; k <= 9; k++) {
                const int e = u >= 2.0 * hk(k) ? (int)floor(__builtin_ldexp(u, k)) - 1 : 0;
                const int er = v >= 2.0 * hk(k) ? (int)floor(__builtin_ldexp(v, k)) - 1 : 0;
                cells += (e - 2 * ep) + (er - 2 * epr);
                if (k >= 1 && k <= KNEAR) c0 += 2 * ((e - 2 * ep) + (er - 2 * epr));
                ep = e;
                epr = er;
            }
            const long long fu = (long long)floor(u / HF);
            long long fine = 0;
            if (u >= HF) fine += fu - 4LL * ep;
            const long long rlo = fu + 2, rhi = NT - 4LL * epr;
            if (rhi >= rlo) fine += rhi - rlo + 1;
            cnt += 2 + 2 * cells + fine;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_xor(cnt, off);
        c0 += __shfl_xor(c0, off);
    }
    __shared__ unsigned long long red[2][BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = cnt;
        red[1][threadIdx.x >> 6] = c0;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long x = 0;
        for (int w = 0; w < BLOCK / 64; w++) x += red[threadIdx.x][w];
        counts[2 * r + threadIdx.x] = x;
    }
}

}  // namespace

// ====================================================================== fixed scratch
// Scratch of marfa_pt_tables_noalloc, for callers that cannot allocate device memory: one
// array per GPU the library runs on, held until the process ends.  In host code the
// array's name is not a device address; its address on each device is looked up once, with
// that device current (fixed_scratch_of), and only that pointer is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = MARFA_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on the current device (which must be `device`)
hipError_t fixed_scratch_of(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

// ====================================================================== host side
struct marfa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    size_t budget = (size_t)24 << 30;  // bytes of per-(line, level) scratch per level batch
    int variant = 4;                   // fine-kernel variant (launch_fine): stage 1024, plain order
    // inputs
    int64_t n = 0;
    std::vector<double> h_nu;
    double *d_nu = nullptr;
    float *d_S = nullptr, *d_gair = nullptr, *d_gself = nullptr, *d_elow = nullptr, *d_nT = nullptr,
          *d_delta = nullptr;
    int32_t *d_molIso = nullptr;
    float *d_tips = nullptr;
    bool tips_set = false;
    std::vector<float> hP, hT, hrho;
    // work buffers
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
    };
    // per-run scratch; a batched job borrows its lane's set (s points there)
    struct Scratch {
        Buf params[4], keys[2], idx[2], coarse, shapes, buckets, hist, spart;
        Buf gbar, gpos, gline, glev, gacc;  // adjoint
    };
    Scratch own;
    Scratch *s = &own;
    Buf lvl, win, out, counts;  // level and window arrays of the last run, records, counters
    Buf adjout;                 // device gradients of marfa_pt_tables_adjoint before their copy
    hipEvent_t ev[5] = {};
    float timings[5] = {0, 0, 0, 0, 0};
    double *dbg_grids = nullptr;
    void *fixed_scratch = nullptr;  // this device's `scratch` (marfa_pt_tables_noalloc)
    // Checkpoint of the last forward run: the level-0 kernel's sums (grid 0 and grids 1..5,
    // NCOARSE values) of every record of every level, then the fine kernel's cascade limiter
    // choices (TILES x LIMW words per record), kept so that the adjoint of the same job and
    // parameters skips the level-0 kernel and the recomputation of the finer grids.  gen
    // changes whenever lines, levels or TIPS do.
    Buf ckpt;
    uint64_t gen = 0;
    // what run() last copied into lvl and win (host images, and the stream it copied on):
    // the same values for the same stream are not copied again.  Cleared by any other
    // writer of those buffers and when they are reallocated.
    std::vector<char> up_lvl, up_win;
    hipStream_t up_stream = nullptr;
    bool up_valid = false;
    // forwards keep the cascade's limiter choices too (marfa_keep_adjoint_state, or an
    // adjoint has run on this context; the adjoint of a forward that did not keep them
    // recomputes them)
    bool adjoint_seen = false;
    struct CkptKey {
        bool valid = false;
        bool lim = false;  // the limiter choices beside the sums were written by a forward run
        uint64_t gen = 0;
        marfa_job job{};
        hipStream_t stream = nullptr;
        int nlev = 0, nrec = 0;
    } ck;
    std::string fixed_err;
};

namespace {

int set_err(marfa_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err(c, MARFA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure(marfa_ctx *c, marfa_ctx::Buf &b, size_t bytes) {
    if (b.bytes >= bytes) return 0;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    size_t want = std::max(bytes, (size_t)256);
    HIPCHK(hipMalloc(&b.p, want));
    b.bytes = want;
    return 0;
}

void free_scratch(marfa_ctx::Scratch &s) {
    marfa_ctx::Buf *bufs[] = {&s.params[0], &s.params[1], &s.params[2], &s.params[3], &s.keys[0], &s.keys[1],
                              &s.idx[0], &s.idx[1], &s.coarse, &s.shapes, &s.buckets, &s.hist, &s.spart,
                              &s.gbar, &s.gpos, &s.gline, &s.glev, &s.gacc};
    for (auto *b : bufs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
    }
}

struct Windows {
    int nrec = 0;
    std::vector<int64_t> lo, hi;
    std::vector<int32_t> rec;
    std::vector<double> org;
};

// app/main.f90:117-174 with src/LBL.f90:39-61 (window = lines read before the exit test
// fires; carry = last line at or below capWV) and the starting line of app/main.f90:223-240.
void compute_windows(const std::vector<double> &nu, const marfa_job &j, Windows &w) {
    const int64_t n = (int64_t)nu.size();
    const double cut = (double)j.cutoff;
    int64_t start = 0;
    if (j.start_wv > cut) {
        const double ext = j.start_wv - cut;
        start = std::upper_bound(nu.begin(), nu.end(), ext) - nu.begin();
    }
    const double startWV = start < n ? nu[start] : INFINITY;
    int64_t idx = start;
    w = Windows();
    for (double a = j.start_wv; a < j.end_wv; a = a + 10.0) {
        const double extStart = a - cut, extEnd = a + 10.0 + cut;
        double cap = extStart + 10.0;
        if (cap <= startWV) cap = startWV;
        int64_t hi = std::lower_bound(nu.begin() + std::min(idx, n), nu.end(), extEnd) - nu.begin();
        hi = std::max(hi, idx);
        int64_t ub = std::upper_bound(nu.begin() + std::min(idx, n), nu.begin() + hi, cap) - nu.begin();
        int64_t carry = ub - 1 >= idx ? ub - 1 : idx;
        w.lo.push_back(idx);
        w.hi.push_back(hi);
        w.rec.push_back((int32_t)((a + 1.0) / 10.0));
        w.org.push_back(a);
        idx = carry;
        w.nrec++;
    }
}

struct LevelSet {  // host-side level state for one run
    std::vector<float> P, T, outmul;
    std::vector<double> pself;
};

// Run levels x records.  Windows are level-independent; out is [nlev][nrec][NT].
// Fixed device memory from which run() carves its scratch instead of allocating it
// (for callers that must not allocate): a bump allocator of 256-byte aligned pieces.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
    void *take(size_t bytes) {
        if (used + round(bytes) > size) return nullptr;
        void *p = base + used;
        used += round(bytes);
        return p;
    }
};

// Level geometry of a run (the same for every level batch): the line range read, the
// bucket index's origin and size, scan chunks, records per level, line blocks.
struct RunGeom {
    int64_t lmin = 0, nl = 0;
    double bbase = 0.0;
    int NB = 0, nsblk = 0, nrec = 0;
    unsigned nblk = 0;
};
// Level and window arrays of a job on the device: P, T, output multiplier, pSelf, sample
// scale, maximum exponent of S(T) per level; window bounds and origin per record.
struct RunArrays {
    float *P = nullptr, *T = nullptr, *mul = nullptr;
    double *pself = nullptr, *sscale = nullptr;
    int *emax = nullptr;
    int64_t *lo = nullptr, *hi = nullptr;
    double *org = nullptr;
};
// Device arrays of one level batch: temporaries of the spectroscopy and the sort (par: shifted
// centre, gD, gL, S(T) in line order; key1/idx1: scattered keys; hist, spart) and what the
// deposit kernels read (key0/idx0: sorted keys and line indices -- first the unsorted ones --,
// shapes: the sorted line records, buckets: the bucket index, coarse: the level-0 sums).
struct BatchBufs {
    void *par[4] = {}, *key0 = nullptr, *key1 = nullptr, *idx0 = nullptr, *idx1 = nullptr;
    double *coarse = nullptr;
    void *shapes = nullptr, *buckets = nullptr, *hist = nullptr, *spart = nullptr;
};

// The run geometry of a job's windows (bucket index over every record's reach, 1/512 cm-1
// buckets from 4 cm-1 below the first window); false when every window is empty.
static bool geom_of(const marfa_job &job, const Windows &W, RunGeom &g) {
    if (W.nrec == 0) return false;
    const int64_t lmin = *std::min_element(W.lo.begin(), W.lo.end());
    const int64_t lmax = *std::max_element(W.hi.begin(), W.hi.end());
    if (lmax <= lmin) return false;
    g.lmin = lmin;
    g.nl = lmax - lmin;
    g.bbase = W.org.front() - (double)job.cutoff - 4.0;
    const double btop = W.org.back() + 10.0 + (double)job.cutoff + 4.0;
    g.NB = (int)std::ceil((btop - g.bbase) * 512.0) + 1;
    g.nsblk = (g.NB + SCAN_CHUNK - 1) / SCAN_CHUNK;
    g.nrec = W.nrec;
    g.nblk = (unsigned)((g.nl + BLOCK - 1) / BLOCK);
    return true;
}

// The level and window arrays in their device buffers (run() writes them in this layout).
static RunArrays arrays_at(void *lvl, void *win, int nlev, int nrec) {
    RunArrays d;
    d.P = (float *)lvl;
    d.T = d.P + nlev;
    d.mul = d.T + nlev;
    d.pself = (double *)(((uintptr_t)(d.mul + nlev) + 15) & ~(uintptr_t)15);
    d.sscale = d.pself + nlev;
    d.emax = (int *)(d.sscale + nlev);
    d.lo = (int64_t *)win;
    d.hi = d.lo + nrec;
    d.org = (double *)(d.hi + nrec);
    return d;
}

// Levels [l0, l0 + nb) of a job up to the deposit kernels, enqueued on st: spectroscopy (and
// the FP32 form's per-level sample scale), then the counting sort by shifted centre into the
// sorted keys, line indices, line records and bucket index.  ka receives the deposit
// kernels' arguments for these levels.  timed: c->ev[1] is recorded between the two parts.
static int enqueue_prep(marfa_ctx *c, const marfa_job &job, const RunGeom &g, const RunArrays &d, int l0, int nb,
                        const BatchBufs &B, float *out, double *out64, hipStream_t st, bool timed, KArgs &ka,
                        bool launch = true) {
    const int64_t nl = g.nl, lmin = g.lmin;
    const int NB = g.NB, nsblk = g.nsblk;
    const double bbase = g.bbase;
    SpecArgs sa;
    sa.nlev = nb;
    sa.lmin = lmin;
    sa.nl = nl;
    sa.nu = c->d_nu;
    sa.S = c->d_S;
    sa.gair = c->d_gair;
    sa.gself = c->d_gself;
    sa.elow = c->d_elow;
    sa.nT = c->d_nT;
    sa.delta = c->d_delta;
    sa.molIso = c->d_molIso;
    sa.P = d.P + l0;
    sa.T = d.T + l0;
    sa.pself = d.pself + l0;
    sa.tips = c->d_tips;
    sa.nus = (double *)B.par[0];
    sa.gD = (double *)B.par[1];
    sa.gL = (double *)B.par[2];
    sa.ST = (double *)B.par[3];
    sa.keys = (uint64_t *)B.key0;
    sa.idx = (int32_t *)B.idx0;
    sa.emax = job.precision == MARFA_FP32K ? d.emax + l0 : nullptr;
    if (launch) {
    // FP32 form: per-level power-of-two sample scale from the maximum exponent of S(T);
    // FP64 form: 1 (the exponent stays at its initial -2000)
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(BLOCK), 0, st, d.emax + l0, (int64_t)nb, -2000);
    dim3 gs((unsigned)((nl + BLOCK - 1) / BLOCK), (unsigned)nb);
    hipLaunchKernelGGL(spectroscopy_kernel, gs, dim3(BLOCK), 0, st, sa);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(scale_kernel, dim3((nb + 63) / 64), dim3(64), 0, st, (const int *)d.emax + l0, nb,
                       d.sscale + l0);
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(c->ev[1], st));
    {
        // order by shifted centre: bucket counts, their scan (the bucket index), scatter,
        // in-bucket order
        int32_t *hist = (int32_t *)B.hist, *bst = (int32_t *)B.buckets;
        int32_t *part = (int32_t *)B.spart;
        const uint64_t *k0 = (const uint64_t *)B.key0;
        const int32_t *i0 = (const int32_t *)B.idx0;
        const int64_t nh = (int64_t)nb * NB;
        hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((nh + BLOCK - 1) / BLOCK, 65536)),
                           dim3(BLOCK), 0, st, hist, nh, 0);
        const dim3 gl((unsigned)((nl + BLOCK - 1) / BLOCK), (unsigned)nb);
        hipLaunchKernelGGL(bucket_hist_kernel, gl, dim3(BLOCK), 0, st, k0, nl, bbase, NB, hist);
        hipLaunchKernelGGL(scan_partial_kernel, dim3(nsblk, nb), dim3(BLOCK), 0, st, (const int32_t *)hist, NB, nsblk,
                           part);
        hipLaunchKernelGGL(scan_top_kernel, dim3(1, nb), dim3(BLOCK), 0, st, part, nsblk);
        hipLaunchKernelGGL(scan_final_kernel, dim3(nsblk, nb), dim3(BLOCK), 0, st, (const int32_t *)hist, NB, nsblk,
                           (const int32_t *)part, bst);
        hipLaunchKernelGGL(bucket_scatter_kernel, gl, dim3(BLOCK), 0, st, k0, i0, nl, bbase, NB, (const int32_t *)bst,
                           hist, (uint64_t *)B.key1, (int32_t *)B.idx1);
        // final order into key0 / idx0 (the unsorted copies are not read again) and the line
        // records
        if (job.precision == MARFA_FP32K)
            hipLaunchKernelGGL(order_shape_kernel<float>, gl, dim3(BLOCK), 0, st, (const uint64_t *)B.key1,
                               (const int32_t *)B.idx1, nl, bbase, NB, (const int32_t *)bst, lmin,
                               (const double *)sa.nus, (const double *)sa.gD, (const double *)sa.gL,
                               (const double *)sa.ST, (const double *)d.sscale, l0, (uint64_t *)B.key0,
                               (int32_t *)B.idx0, (LineShape<float> *)B.shapes);
        else
            hipLaunchKernelGGL(order_shape_kernel<double>, gl, dim3(BLOCK), 0, st, (const uint64_t *)B.key1,
                               (const int32_t *)B.idx1, nl, bbase, NB, (const int32_t *)bst, lmin,
                               (const double *)sa.nus, (const double *)sa.gD, (const double *)sa.gL,
                               (const double *)sa.ST, (const double *)d.sscale, l0, (uint64_t *)B.key0,
                               (int32_t *)B.idx0, (LineShape<double> *)B.shapes);
        HIPCHK(hipGetLastError());
    }
    }  // launch
    ka = KArgs();
    ka.nlev = nb;
    ka.nrec = g.nrec;
    ka.lev0 = l0;
    ka.lmin = lmin;
    ka.nl = nl;
    ka.nus = sa.nus;
    ka.gD = sa.gD;
    ka.gL = sa.gL;
    ka.S = sa.ST;
    ka.skey = (const uint64_t *)B.key0;
    ka.sidx = (const int32_t *)B.idx0;
    ka.lo = d.lo;
    ka.hi = d.hi;
    ka.org = d.org;
    ka.sscale = d.sscale;
    ka.T = d.T;
    ka.outmul = d.mul;
    ka.cutoff = job.cutoff;
    ka.shape_id = job.shape_id;
    ka.chi_id = job.chi_id;
    ka.mol = job.mol_code;
    ka.coarse = B.coarse;
    ka.out = out;
    ka.out64 = out64;
    ka.grids = c->dbg_grids;
    ka.shp = B.shapes;
    ka.bst = (const int32_t *)B.buckets;
    ka.bbase = bbase;
    ka.NB = NB;
    return 0;
}

// Adjoint request of run(): gout = dL/d(record point) in the records' layout (device);
// the per-line gradients accumulate in c->s->gacc ([nl][4] from line lmin) and land in
// glines (device, [4][n]), the per-level T, P gradients in glevels (device, [2][nlev]).
struct AdjReq {
    const float *gout = nullptr;
    double *glines = nullptr, *glevels = nullptr;
    int64_t lmin = 0, nl = 0;
};

// capture: the kernels only, for recording into a hipGraph -- the level and window arrays
// must already be on the device (an earlier run of the same job on the same context) and
// every buffer already large enough, so that no allocation, copy or synchronisation occurs.
// ckpt: device array [nlev][nrec][NCOARSE] for the level-0 sums instead of per-batch scratch
// (the forward's checkpoint); reuse: it already holds them for this job, so the level-0
// kernel is not launched.  lim: device array [nlev][nrec][TILES][LIMW] beside it, where the
// forward's fine kernel records the cascade's limiter choices; lim_ok: it holds them for
// this job, and the adjoint reads them instead of recomputing the grid sums and cascade.
int run(marfa_ctx *c, const marfa_job &job, const LevelSet &L, const Windows &W, float *out,
        hipStream_t st, bool timed, AdjReq *adj = nullptr, double *out64 = nullptr, bool capture = false,
        Arena *arena = nullptr, double *ckpt = nullptr, bool reuse = false, uint64_t *lim = nullptr,
        bool lim_ok = false) {
    if (capture) timed = false;
    const int nlev = (int)L.P.size(), nrec = W.nrec;
    if (nrec == 0 || nlev == 0) return 0;
    for (float T : L.T)
        if (!(T >= 20.f && T < 1018.f)) return set_err(c, MARFA_E_TEMPERATURE, "temperature outside the TIPS table");
    const int64_t lmin = *std::min_element(W.lo.begin(), W.lo.end());
    const int64_t lmax = *std::max_element(W.hi.begin(), W.hi.end());
    if (lmin < 0 || lmax > c->n) return set_err(c, MARFA_E_WINDOW, "window outside the line list");
    if (lmax <= lmin) {
        // every window is empty (the range lies beyond the line list): the records are zero,
        // cleared by a kernel (no memset, so a capture holds it too); no line is read
        const int64_t words = (int64_t)nlev * nrec * NT * (out64 ? 2 : 1);
        int32_t *dst = out64 ? (int32_t *)out64 : (int32_t *)out;
        if (dst) {
            hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((words + BLOCK - 1) / BLOCK, 65536)),
                               dim3(BLOCK), 0, st, dst, words, 0);
            HIPCHK(hipGetLastError());
        }
        if (adj) {
            adj->lmin = lmin;
            adj->nl = 0;
            hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((8 * c->n + BLOCK - 1) / BLOCK, 65536)),
                               dim3(BLOCK), 0, st, (int32_t *)adj->glines, 8 * c->n, 0);
            hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(BLOCK), 0, st, (int32_t *)adj->glevels,
                               (int64_t)4 * nlev, 0);
            HIPCHK(hipGetLastError());
        }
        return 0;
    }
    const int64_t nl = lmax - lmin;

    // bucket index of shifted centres: 1/512 cm-1 buckets over every record's reach
    RunGeom G0;
    geom_of(job, W, G0);
    const int NB = G0.NB;
    const size_t shp_bytes = job.precision == MARFA_FP32K ? line_record_bytes<float>() : line_record_bytes<double>();

    const int nsblk = G0.nsblk;
    const unsigned nblk = G0.nblk;

    // level batches so that per-(line, level) arrays stay within a memory budget -- or, with
    // an arena (the allocation-free entry point), within what the arena holds
    size_t per_level = (size_t)nl * (4 * 8 + 2 * 8 + 2 * 4 + shp_bytes) + (size_t)(2 * NB + 1) * 4;
    if (adj) per_level += (size_t)nrec * (30690 + NT + GPOS) * 4 + (size_t)nl * 4 * 8;
    const size_t budget = c->budget;
    int lb = (int)std::max<size_t>(1, std::min<size_t>((size_t)nlev, budget / std::max<size_t>(per_level, 1)));
    const size_t lvl_bytes = (size_t)nlev * (4 * 4 + 2 * 8) + 64, win_bytes = (size_t)nrec * 24 + 64;
    if (arena) {
        // the level and window arrays, then per level of a batch every scratch array, each
        // piece rounded up to the arena's 256-byte alignment (13 pieces per batch)
        const size_t fixed = Arena::round(lvl_bytes) + Arena::round(win_bytes) + 13 * 256;
        const size_t one = (size_t)nl * (4 * 8 + 2 * 8 + 2 * 4 + shp_bytes) + (size_t)nrec * NCOARSE * 8 +
                           (size_t)(2 * NB + 1) * 4 + (size_t)nsblk * 4;
        if (adj || arena->size < fixed + one)
            return set_err(c, MARFA_E_ARG, "the job needs " + std::to_string(fixed + one) +
                                               " bytes of scratch for one level; the fixed scratch holds " +
                                               std::to_string(arena->size));
        lb = (int)std::min<size_t>((size_t)nlev, (arena->size - fixed) / one);
    }
    // scratch arrays: the context's buffers (grown on demand) or carved from the arena
    auto get = [&](marfa_ctx::Buf &b, size_t bytes) -> void * {
        if (arena) return arena->take(bytes);
        return ensure(c, b, bytes) ? nullptr : b.p;
    };
#define MARFA_GET(var, buf, bytes)                                                                   \
    void *var = get(buf, bytes);                                                                     \
    if (!var) return arena ? set_err(c, MARFA_E_ARG, "fixed scratch exhausted") : MARFA_E_HIP;

    // level arrays [nlev]: P, T, outmul (float), pself, sscale (double), max exponent (int);
    // window arrays [nrec]: lo, hi, origin
    void *p_lvl = c->lvl.p, *p_win = c->win.p;
    if (!capture) {
        // a reallocation (capacity change) may return the old address with other contents
        const size_t lvl0 = c->lvl.bytes, win0 = c->win.bytes;
        p_lvl = get(c->lvl, lvl_bytes);
        if (!p_lvl) return MARFA_E_HIP;
        p_win = get(c->win, win_bytes);
        if (!p_win) return MARFA_E_HIP;
        if (arena || c->lvl.bytes != lvl0 || c->win.bytes != win0) c->up_valid = false;
    }
    const RunArrays D = arrays_at(p_lvl, p_win, nlev, nrec);
    float *dP = D.P, *dT = D.T, *dmul = D.mul;
    double *dps = D.pself;
    int64_t *dlo = D.lo, *dhi = D.hi;
    double *dorg = D.org;
    if (!capture) {
        // host images of what the level and window arrays receive
        std::vector<char> il((size_t)nlev * 20), iw((size_t)nrec * 24);
        memcpy(il.data(), L.P.data(), nlev * 4);
        memcpy(il.data() + nlev * 4, L.T.data(), nlev * 4);
        memcpy(il.data() + nlev * 8, L.outmul.data(), nlev * 4);
        memcpy(il.data() + nlev * 12, L.pself.data(), nlev * 8);
        memcpy(iw.data(), W.lo.data(), nrec * 8);
        memcpy(iw.data() + nrec * 8, W.hi.data(), nrec * 8);
        memcpy(iw.data() + nrec * 16, W.org.data(), nrec * 8);
        const bool same = c->up_valid && c->up_stream == st;
        c->up_valid = false;  // until every copy below is enqueued
        if (!(same && il == c->up_lvl)) {
            HIPCHK(hipMemcpyAsync(dP, L.P.data(), nlev * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dT, L.T.data(), nlev * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dmul, L.outmul.data(), nlev * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dps, L.pself.data(), nlev * 8, hipMemcpyHostToDevice, st));
        }
        if (!(same && iw == c->up_win)) {
            HIPCHK(hipMemcpyAsync(dlo, W.lo.data(), nrec * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dhi, W.hi.data(), nrec * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dorg, W.org.data(), nrec * 8, hipMemcpyHostToDevice, st));
        }
        if (!arena) {
            c->up_lvl = std::move(il);
            c->up_win = std::move(iw);
            c->up_stream = st;
            c->up_valid = true;
        }
    }

    const size_t arr = (size_t)lb * nl;
    MARFA_GET(p_par0, c->s->params[0], arr * 8)
    MARFA_GET(p_par1, c->s->params[1], arr * 8)
    MARFA_GET(p_par2, c->s->params[2], arr * 8)
    MARFA_GET(p_par3, c->s->params[3], arr * 8)
    MARFA_GET(p_key0, c->s->keys[0], arr * 8)
    MARFA_GET(p_key1, c->s->keys[1], arr * 8)
    MARFA_GET(p_idx0, c->s->idx[0], arr * 4)
    MARFA_GET(p_idx1, c->s->idx[1], arr * 4)
    void *p_coarse = nullptr;
    if (!ckpt) {
        p_coarse = get(c->s->coarse, (size_t)lb * nrec * NCOARSE * 8);
        if (!p_coarse) return arena ? set_err(c, MARFA_E_ARG, "fixed scratch exhausted") : MARFA_E_HIP;
    }
    MARFA_GET(p_shapes, c->s->shapes, arr * shp_bytes)
    MARFA_GET(p_buckets, c->s->buckets, (size_t)lb * (NB + 1) * 4)
    MARFA_GET(p_hist, c->s->hist, (size_t)lb * NB * 4)
    MARFA_GET(p_spart, c->s->spart, (size_t)lb * nsblk * 4)
    void *p_gbar = nullptr, *p_gpos = nullptr, *p_gline = nullptr, *p_glev = nullptr, *p_gacc = nullptr;
    if (adj) {
        if (!(p_gbar = get(c->s->gbar, (size_t)lb * nrec * (30690 + NT) * 4)) ||
            !(p_gpos = get(c->s->gpos, (size_t)lb * nrec * GPOS * 4)) ||
            !(p_gline = get(c->s->gline, arr * 4 * 8)) || !(p_glev = get(c->s->glev, (size_t)2 * lb * nblk * 2 * 8)) ||
            !(p_gacc = get(c->s->gacc, (size_t)nl * 4 * 8)))
            return MARFA_E_HIP;
        // cleared by a kernel, not a memset, so that the sequence can be captured
        hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((8 * nl + BLOCK - 1) / BLOCK, 65536)),
                           dim3(BLOCK), 0, st, (int32_t *)p_gacc, 8 * nl, 0);
        HIPCHK(hipGetLastError());
        adj->lmin = lmin;
        adj->nl = nl;
    }

    RunGeom G;
    geom_of(job, W, G);

    float tms[5] = {0, 0, 0, 0, 0};
    for (int l0 = 0; l0 < nlev; l0 += lb) {
        const int nb = std::min(lb, nlev - l0);
        if (timed) HIPCHK(hipEventRecord(c->ev[0], st));
        BatchBufs B;
        B.par[0] = p_par0;
        B.par[1] = p_par1;
        B.par[2] = p_par2;
        B.par[3] = p_par3;
        B.key0 = p_key0;
        B.key1 = p_key1;
        B.idx0 = p_idx0;
        B.idx1 = p_idx1;
        B.coarse = ckpt ? ckpt + (size_t)l0 * nrec * NCOARSE : (double *)p_coarse;
        B.shapes = p_shapes;
        B.buckets = p_buckets;
        B.hist = p_hist;
        B.spart = p_spart;
        KArgs ka;
        int rc = enqueue_prep(c, job, G, D, l0, nb, B, out, out64, st, timed, ka);
        if (rc) return rc;
        if (lim) ka.lim = lim + (size_t)l0 * nrec * TILES * LIMW;
        if (timed) HIPCHK(hipEventRecord(c->ev[2], st));
        const int nitems = nb * nrec;
        const unsigned gcoarse = (unsigned)(8 * ((nitems + 7) / 8));
        const unsigned gfine = (unsigned)(8 * ((nitems * TILES + 7) / 8));
        // the chi factor is a compile-time choice of the level-0 kernel (as in the reference,
        // src/ChiFactors.f90:11-16), so its hot loop carries only the selected branch
        if (!reuse) {
            const bool f32 = job.precision == MARFA_FP32K;
            const int chi = job.mol_code == 2 ? job.chi_id : 0;
#define MARFA_COARSE(CH)                                                                              \
    do {                                                                                              \
        if (f32) hipLaunchKernelGGL((coarse_kernel<float, CH>), dim3(gcoarse), dim3(BLOCK), 0, st, ka);  \
        else hipLaunchKernelGGL((coarse_kernel<double, CH>), dim3(gcoarse), dim3(BLOCK), 0, st, ka);     \
    } while (0)
            switch (chi) {
            case 1: MARFA_COARSE(1); break;
            case 2: MARFA_COARSE(2); break;
            case 3: MARFA_COARSE(3); break;
            default: MARFA_COARSE(0); break;
            }
#undef MARFA_COARSE
        }
        HIPCHK(hipGetLastError());
        if (timed) HIPCHK(hipEventRecord(c->ev[3], st));
        if (adj) {
            // adjoint: dL/dnode of every record, then the line pass and its reductions
            float *gbar = (float *)p_gbar;
            const bool lk = lim && lim_ok;
            if (job.precision == MARFA_FP32K) {
                if (lk) hipLaunchKernelGGL((adjoint_tile_kernel<float, true>), dim3(gfine), dim3(BLOCK), 0, st, ka, adj->gout, gbar);
                else hipLaunchKernelGGL((adjoint_tile_kernel<float, false>), dim3(gfine), dim3(BLOCK), 0, st, ka, adj->gout, gbar);
            } else {
                if (lk) hipLaunchKernelGGL((adjoint_tile_kernel<double, true>), dim3(gfine), dim3(BLOCK), 0, st, ka, adj->gout, gbar);
                else hipLaunchKernelGGL((adjoint_tile_kernel<double, false>), dim3(gfine), dim3(BLOCK), 0, st, ka, adj->gout, gbar);
            }
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(adjoint_gpos_kernel, dim3((unsigned)((nitems * GPOS + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                               st, (const float *)gbar, nitems, (float *)p_gpos);
            HIPCHK(hipGetLastError());
            AdjArgs ab;
            ab.nu = c->d_nu;
            ab.S = c->d_S;
            ab.gair = c->d_gair;
            ab.gself = c->d_gself;
            ab.elow = c->d_elow;
            ab.nT = c->d_nT;
            ab.delta = c->d_delta;
            ab.molIso = c->d_molIso;
            ab.tips = c->d_tips;
            ab.P = dP;
            ab.pself = dps;
            ab.gbar = gbar;
            ab.gpos = (const float *)p_gpos;
            ab.gline = (double *)p_gline;
            ab.glev = (double *)p_glev;
            // pass 0 (far wings) without the chi terms when the job has no chi
            const bool chif = job.mol_code == 2 && job.chi_id != 0;
            if (job.precision == MARFA_FP32K) {
                if (chif) hipLaunchKernelGGL((adjoint_line_kernel<float, 0, true>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
                else hipLaunchKernelGGL((adjoint_line_kernel<float, 0, false>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
                hipLaunchKernelGGL((adjoint_line_kernel<float, 1>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
            } else {
                if (chif) hipLaunchKernelGGL((adjoint_line_kernel<double, 0, true>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
                else hipLaunchKernelGGL((adjoint_line_kernel<double, 0, false>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
                hipLaunchKernelGGL((adjoint_line_kernel<double, 1>), dim3(nblk, nb), dim3(BLOCK), 0, st, ka, ab);
            }
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(nblk), dim3(BLOCK), 0, st, (const double *)p_gline, nb,
                               nl, (double *)p_gacc);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(adjoint_level_kernel, dim3(2 * nb), dim3(BLOCK), 0, st,
                               (const double *)p_glev, nb, (int)nblk, nlev, l0, adj->glevels);
            HIPCHK(hipGetLastError());
        } else if (job.precision == MARFA_FP32K) {
            launch_fine<float>(c->variant, dim3(gfine), st, ka);
        } else {
            launch_fine<double>(c->variant, dim3(gfine), st, ka);
        }
        HIPCHK(hipGetLastError());
        if (timed) {
            HIPCHK(hipEventRecord(c->ev[4], st));
            HIPCHK(hipEventSynchronize(c->ev[4]));
            float x;
            for (int q = 0; q < 4; q++) {
                HIPCHK(hipEventElapsedTime(&x, c->ev[q], c->ev[q + 1]));
                tms[q] += x;
            }
        }
    }
    if (adj) {
        hipLaunchKernelGGL(adjoint_scatter_kernel, dim3((unsigned)std::min<int64_t>((c->n + BLOCK - 1) / BLOCK, 65536)),
                           dim3(BLOCK), 0, st, (const double *)p_gacc, lmin, nl, c->n, adj->glines);
        HIPCHK(hipGetLastError());
    }
    if (timed) {
        tms[4] = tms[0] + tms[1] + tms[2] + tms[3];
        memcpy(c->timings, tms, sizeof tms);
    }
    return 0;
}
#undef MARFA_GET

}  // namespace

// ====================================================================== C ABI
extern "C" {

int marfa_version(void) { return 1; }

const char *marfa_last_error(const marfa_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int marfa_ctx_create(int device, marfa_ctx **out) {
    if (!out) return MARFA_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MARFA_E_NODEVICE;
    if (device < 0 || device >= ndev) return MARFA_E_ARG;
    marfa_ctx *c = new marfa_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return MARFA_E_HIP;
    }
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    // the allocation-free entry point's scratch on this device (looked up once per device)
    const hipError_t e = fixed_scratch_of(device, &c->fixed_scratch);
    if (e != hipSuccess) c->fixed_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    *out = c;
    return 0;
}

int marfa_ctx_destroy(marfa_ctx *c) {
    if (!c) return MARFA_E_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->d_nu, c->d_S, c->d_gair, c->d_gself, c->d_elow, c->d_nT, c->d_delta, c->d_molIso, c->d_tips};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free_scratch(c->own);
    marfa_ctx::Buf *bufs[] = {&c->lvl, &c->win, &c->out, &c->counts, &c->adjout, &c->ckpt};
    for (auto *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int marfa_set_tips(marfa_ctx *c, int32_t niso, int32_t ntemp, const float *tips) {
    if (!c || !tips || niso != 74 || ntemp != 501) return set_err(c, MARFA_E_ARG, "TIPS must be 74 x 501");
    (void)hipSetDevice(c->device);
    if (!c->d_tips) HIPCHK(hipMalloc(&c->d_tips, 74 * 501 * 4));
    HIPCHK(hipMemcpy(c->d_tips, tips, 74 * 501 * 4, hipMemcpyHostToDevice));
    c->tips_set = true;
    c->gen++;
    return 0;
}

int marfa_set_lines(marfa_ctx *c, int64_t n, const double *nu, const float *S, const float *gair,
                    const float *gself, const float *elow, const float *nT, const float *delta,
                    const int32_t *molIso) {
    if (!c || n < 0 || (n > 0 && (!nu || !S || !gair || !gself || !elow || !nT || !delta || !molIso)))
        return set_err(c, MARFA_E_ARG, "bad line arrays");
    if (n >= (int64_t)1 << 31) return set_err(c, MARFA_E_ARG, "at most 2^31-1 lines per context");
    for (int64_t i = 0; i < n; i++) {
        const int iso = molIso[i] / 100;
        if (iso < 1 || iso > 74) return set_err(c, MARFA_E_ARG, "isotopologue outside the TIPS/WISO tables");
        if (i && !(nu[i] >= nu[i - 1])) return set_err(c, MARFA_E_ARG, "line centres must be ascending");
    }
    (void)hipSetDevice(c->device);
    void **ptrs[] = {(void **)&c->d_nu, (void **)&c->d_S, (void **)&c->d_gair, (void **)&c->d_gself,
                     (void **)&c->d_elow, (void **)&c->d_nT, (void **)&c->d_delta, (void **)&c->d_molIso};
    for (void **p : ptrs)
        if (*p) {
            HIPCHK(hipFree(*p));
            *p = nullptr;
        }
    const size_t m = (size_t)std::max<int64_t>(n, 1);
    HIPCHK(hipMalloc(&c->d_nu, m * 8));
    HIPCHK(hipMalloc(&c->d_S, m * 4));
    HIPCHK(hipMalloc(&c->d_gair, m * 4));
    HIPCHK(hipMalloc(&c->d_gself, m * 4));
    HIPCHK(hipMalloc(&c->d_elow, m * 4));
    HIPCHK(hipMalloc(&c->d_nT, m * 4));
    HIPCHK(hipMalloc(&c->d_delta, m * 4));
    HIPCHK(hipMalloc(&c->d_molIso, m * 4));
    if (n > 0) {
        HIPCHK(hipMemcpy(c->d_nu, nu, n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_S, S, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_gair, gair, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_gself, gself, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_elow, elow, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_nT, nT, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_delta, delta, n * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_molIso, molIso, n * 4, hipMemcpyHostToDevice));
    }
    c->h_nu.assign(nu, nu + n);
    c->n = n;
    c->gen++;
    return 0;
}

int marfa_set_levels(marfa_ctx *c, int32_t nlev, const float *P, const float *T, const float *rho) {
    if (!c || nlev < 0 || nlev > 200 || (nlev > 0 && (!P || !T || !rho)))
        return set_err(c, MARFA_E_ARG, "bad levels (at most 200, src/Atmsosphere.f90:37)");
    c->hP.assign(P, P + nlev);
    c->hT.assign(T, T + nlev);
    c->hrho.assign(rho, rho + nlev);
    c->gen++;
    return 0;
}

int marfa_windows(marfa_ctx *c, const marfa_job *job, int32_t max_rec, int32_t *nrec, int64_t *lo,
                  int64_t *hi, int32_t *recnum, double *origin) {
    if (!c || !job || !nrec) return set_err(c, MARFA_E_ARG, "null argument");
    Windows w;
    compute_windows(c->h_nu, *job, w);
    *nrec = w.nrec;
    if (w.nrec > max_rec) return 0;
    for (int i = 0; i < w.nrec; i++) {
        if (lo) lo[i] = w.lo[i];
        if (hi) hi[i] = w.hi[i];
        if (recnum) recnum[i] = w.rec[i];
        if (origin) origin[i] = w.org[i];
    }
    return 0;
}

static int check_job(marfa_ctx *c, const marfa_job *job) {
    if (!c || !job) return MARFA_E_ARG;
    if (!c->tips_set || c->n == 0) return set_err(c, MARFA_E_STATE, "lines and TIPS must be set");
    if (job->shape_id < 0 || job->shape_id > 1 || job->chi_id < 0 || job->chi_id > 3 ||
        (job->precision != MARFA_FP64 && job->precision != MARFA_FP32K) || !(job->cutoff >= 0.f) ||
        !(job->end_wv > job->start_wv))
        return set_err(c, MARFA_E_ARG, "bad job parameters");
    return 0;
}

static LevelSet level_set(const marfa_ctx *c, const marfa_job &job) {
    LevelSet L;
    L.P = c->hP;
    L.T = c->hT;
    for (size_t i = 0; i < c->hP.size(); i++) {
        // pSelf of app/main.f90:92 in FP64
        L.pself.push_back((double)c->hrho[i] * 10.0 * (double)MARFA_BOLSI_F * (double)c->hT[i] / (double)MARFA_PATM_F);
        L.outmul.push_back(job.target == MARFA_VAC ? c->hrho[i] : 1.0f);
    }
    return L;
}

// The checkpoint buffer for a run of job over every level of the context (null when the job
// has no records); the key is cleared until the run that fills it has been enqueued.
static double *ckpt_for(marfa_ctx *c, const Windows &w) {
    c->ck.valid = false;
    if (w.nrec == 0) return nullptr;
    const size_t recs = c->hP.size() * (size_t)w.nrec;
    if (ensure(c, c->ckpt, recs * NCOARSE * 8 + recs * TILES * LIMW * 8)) return nullptr;
    return (double *)c->ckpt.p;
}
// the limiter choices beside the level-0 sums in the checkpoint buffer
static uint64_t *ckpt_lim(marfa_ctx *c, const Windows &w) {
    return (uint64_t *)((double *)c->ckpt.p + c->hP.size() * (size_t)w.nrec * NCOARSE);
}
static void ckpt_set(marfa_ctx *c, const marfa_job &job, hipStream_t st, const Windows &w, bool lim) {
    c->ck.valid = w.nrec > 0;
    c->ck.lim = lim;
    c->ck.gen = c->gen;
    c->ck.job = job;
    c->ck.stream = st;
    c->ck.nlev = (int)c->hP.size();
    c->ck.nrec = w.nrec;
}
// does the checkpoint hold this job's level-0 sums for the current inputs, written on the
// same stream (so that the adjoint's kernels are ordered after the forward that wrote it)?
static bool ckpt_holds(const marfa_ctx *c, const marfa_job &job, hipStream_t st, const Windows &w) {
    const marfa_job &k = c->ck.job;
    return c->ck.valid && c->ck.gen == c->gen && c->ck.stream == st && c->ck.nlev == (int)c->hP.size() &&
           c->ck.nrec == w.nrec && k.start_wv == job.start_wv && k.end_wv == job.end_wv && k.cutoff == job.cutoff &&
           k.shape_id == job.shape_id && k.chi_id == job.chi_id && k.mol_code == job.mol_code &&
           k.precision == job.precision;
}

int marfa_pt_tables(marfa_ctx *c, const marfa_job *job, float *out_dev, float *out_host, void *stream,
                    int32_t async) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (c->hP.empty()) return set_err(c, MARFA_E_STATE, "levels must be set");
    (void)hipSetDevice(c->device);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    Windows w;
    compute_windows(c->h_nu, *job, w);
    LevelSet L = level_set(c, *job);
    const size_t total = (size_t)L.P.size() * w.nrec * NT;
    float *out = out_dev;
    if (!out) {
        if (ensure(c, c->out, total * 4)) return MARFA_E_HIP;
        out = (float *)c->out.p;
    }
    // phases are timed (and the host waits for them) unless the call is asynchronous
    double *ck = ckpt_for(c, w);
    if (w.nrec > 0 && !ck) return MARFA_E_HIP;
    const bool keep_lim = ck && c->adjoint_seen;
    rc = run(c, *job, L, w, out, st, out_host != nullptr || !async, nullptr, nullptr, false, nullptr, ck, false,
             keep_lim ? ckpt_lim(c, w) : nullptr);
    if (rc) return rc;
    ckpt_set(c, *job, st, w, keep_lim);
    if (out_host) {
        HIPCHK(hipMemcpyAsync(out_host, out, total * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    } else if (!async) {
        HIPCHK(hipStreamSynchronize(st));
    }
    return 0;
}

int marfa_pt_tables_noalloc(marfa_ctx *c, const marfa_job *job, float *out_dev, void *stream) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!out_dev) return set_err(c, MARFA_E_ARG, "out_dev is required: this entry point allocates nothing");
    if (c->hP.empty()) return set_err(c, MARFA_E_STATE, "levels must be set");
    if (!c->fixed_scratch) return set_err(c, MARFA_E_HIP, "no fixed scratch on this device: " + c->fixed_err);
    (void)hipSetDevice(c->device);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    Windows w;
    compute_windows(c->h_nu, *job, w);
    LevelSet L = level_set(c, *job);
    Arena ar;
    ar.base = (char *)c->fixed_scratch;
    ar.size = NOALLOC_SCRATCH_BYTES;
    // no phase timing (no events to wait on): the call only enqueues its copies and kernels
    return run(c, *job, L, w, out_dev, st, false, nullptr, nullptr, false, &ar);
}

int marfa_pt_tables_f64(marfa_ctx *c, const marfa_job *job, double *out_dev, void *stream) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!out_dev) return set_err(c, MARFA_E_ARG, "null argument");
    if (c->hP.empty()) return set_err(c, MARFA_E_STATE, "levels must be set");
    (void)hipSetDevice(c->device);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    Windows w;
    compute_windows(c->h_nu, *job, w);
    LevelSet L = level_set(c, *job);
    double *ck = ckpt_for(c, w);
    if (w.nrec > 0 && !ck) return MARFA_E_HIP;
    const bool keep_lim = ck && c->adjoint_seen;
    rc = run(c, *job, L, w, nullptr, st, true, nullptr, out_dev, false, nullptr, ck, false,
             keep_lim ? ckpt_lim(c, w) : nullptr);
    if (rc) return rc;
    ckpt_set(c, *job, st, w, keep_lim);
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// adjoint into device buffers: grad_lines [4][n], grad_levels [2][nlev], on stream st
static int adjoint_dev(marfa_ctx *c, const marfa_job *job, const float *grad_out, hipStream_t st, double *grad_lines,
                       double *grad_levels, bool timed) {
    Windows w;
    compute_windows(c->h_nu, *job, w);
    LevelSet L = level_set(c, *job);
    const int nlev = (int)L.P.size();
    if (w.nrec == 0) {
        hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((8 * c->n + BLOCK - 1) / BLOCK, 65536)),
                           dim3(BLOCK), 0, st, (int32_t *)grad_lines, 8 * c->n, 0);
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(BLOCK), 0, st, (int32_t *)grad_levels, (int64_t)4 * nlev, 0);
        HIPCHK(hipGetLastError());
        return 0;
    }
    AdjReq req;
    req.gout = grad_out;
    req.glines = grad_lines;
    req.glevels = grad_levels;
    // dL/dP_level and dL/dT_level: T also enters through pSelf = rho 10 k T / p0 (in the
    // line pass) -- the density is held fixed.  The level-0 sums come from the forward's
    // checkpoint when it holds this job with the current inputs, else they are recomputed.
    c->adjoint_seen = true;
    const bool reuse = ckpt_holds(c, *job, st, w);
    const bool lim_ok = reuse && c->ck.lim;
    double *ck = reuse ? (double *)c->ckpt.p : ckpt_for(c, w);
    if (!ck) return MARFA_E_HIP;
    const int rc = run(c, *job, L, w, nullptr, st, timed, &req, nullptr, false, nullptr, ck, reuse,
                       lim_ok ? ckpt_lim(c, w) : nullptr, lim_ok);
    if (rc) return rc;
    if (!reuse) ckpt_set(c, *job, st, w, false);  // the sums, without limiter choices
    return 0;
}

int marfa_pt_tables_adjoint(marfa_ctx *c, const marfa_job *job, const float *grad_out, void *stream,
                            double *grad_lines, double *grad_levels) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!grad_out || !grad_lines || !grad_levels) return set_err(c, MARFA_E_ARG, "null argument");
    if (c->hP.empty()) return set_err(c, MARFA_E_STATE, "levels must be set");
    (void)hipSetDevice(c->device);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    const size_t nlev = c->hP.size(), nline = (size_t)c->n * 4 * 8;
    if (ensure(c, c->adjout, nline + nlev * 2 * 8 + 256)) return MARFA_E_HIP;
    double *gl = (double *)c->adjout.p, *gv = (double *)((char *)c->adjout.p + ((nline + 255) & ~(size_t)255));
    rc = adjoint_dev(c, job, grad_out, st, gl, gv, true);  // synchronous anyway: phases timed
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(grad_lines, gl, nline, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(grad_levels, gv, nlev * 2 * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int marfa_pt_tables_adjoint_dev(marfa_ctx *c, const marfa_job *job, const float *grad_out, void *stream,
                                double *grad_lines, double *grad_levels) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!grad_out || !grad_lines || !grad_levels) return set_err(c, MARFA_E_ARG, "null argument");
    if (c->hP.empty()) return set_err(c, MARFA_E_STATE, "levels must be set");
    (void)hipSetDevice(c->device);
    return adjoint_dev(c, job, grad_out, stream ? (hipStream_t)stream : c->stream, grad_lines, grad_levels, false);
}

int marfa_keep_adjoint_state(marfa_ctx *c, int32_t on) {
    if (!c) return MARFA_E_ARG;
    c->adjoint_seen = on != 0;
    return 0;
}

int marfa_update_line_params(marfa_ctx *c, const float *S, const float *gair, const float *nT, const float *delta,
                             void *stream) {
    if (!c || !S || !gair || !nT || !delta) return set_err(c, MARFA_E_ARG, "null argument");
    if (c->n == 0) return set_err(c, MARFA_E_STATE, "lines must be set");
    (void)hipSetDevice(c->device);
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    const size_t b = (size_t)c->n * 4;
    HIPCHK(hipMemcpyAsync(c->d_S, S, b, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->d_gair, gair, b, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->d_nT, nT, b, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->d_delta, delta, b, hipMemcpyDeviceToDevice, st));
    c->gen++;
    return 0;
}

int marfa_ctx_info(const marfa_ctx *c, int64_t *nlines, int32_t *nlev, int32_t *device) {
    if (!c) return MARFA_E_ARG;
    if (nlines) *nlines = c->n;
    if (nlev) *nlev = (int32_t)c->hP.size();
    if (device) *device = c->device;
    return 0;
}

int marfa_lbl_scheme(marfa_ctx *c, int64_t *line_idx, double cap_wv, double start_dwv, double ext_end_dwv,
                     float cutoff, float P, float T, float pSelf, int32_t shape_id, int32_t chi_id,
                     int32_t mol_code, int32_t precision, float *rk) {
    marfa_job job = {start_dwv, start_dwv + 10.0, cutoff, shape_id, chi_id, mol_code, MARFA_ACS, precision};
    int rc = check_job(c, &job);
    if (rc) return rc;
    if (!line_idx || !rk) return set_err(c, MARFA_E_ARG, "null argument");
    const int64_t lo = *line_idx - 1;
    if (lo < 0 || lo >= c->n) return set_err(c, MARFA_E_WINDOW, "line index outside the line list");
    (void)hipSetDevice(c->device);
    const auto &nu = c->h_nu;
    // src/LBL.f90:41-61: read until the unshifted centre reaches extEndDeltaWV, carry the
    // last line at or below capWV
    int64_t hi = std::lower_bound(nu.begin() + lo, nu.end(), ext_end_dwv) - nu.begin();
    int64_t ub = std::upper_bound(nu.begin() + lo, nu.begin() + hi, cap_wv) - nu.begin();
    if (ub - 1 >= lo) *line_idx = ub;  // 1-based index of the last line <= capWV
    Windows w;
    w.nrec = 1;
    w.lo = {lo};
    w.hi = {std::max(hi, lo)};
    w.rec = {(int32_t)((start_dwv + 1.0) / 10.0)};
    w.org = {start_dwv};
    if (hi <= lo) {
        std::fill(rk, rk + NT, 0.f);
        return 0;
    }
    LevelSet L;
    L.P = {P};
    L.T = {T};
    L.pself = {(double)pSelf};
    L.outmul = {1.0f};
    if (ensure(c, c->out, (size_t)NT * 4)) return MARFA_E_HIP;
    rc = run(c, job, L, w, (float *)c->out.p, c->stream, false);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(rk, c->out.p, NT * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int marfa_count_samples(marfa_ctx *c, const marfa_job *job, int64_t *samples) {
    // samples[0]: distinct samples of the reference geometry; samples[1]: their grid-0 part
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!samples || c->hP.empty()) return set_err(c, MARFA_E_ARG, "levels and output required");
    (void)hipSetDevice(c->device);
    Windows w;
    compute_windows(c->h_nu, *job, w);
    LevelSet L = level_set(c, *job);
    const int nlev = (int)L.P.size();
    samples[0] = samples[1] = 0;
    if (w.nrec == 0) return 0;
    // shifted centres for every level, then one workgroup per record
    const int64_t lmin = *std::min_element(w.lo.begin(), w.lo.end());
    const int64_t lmax = *std::max_element(w.hi.begin(), w.hi.end());
    if (lmax <= lmin) return 0;  // every window empty: no samples (and no line is read)
    const int64_t nl = lmax - lmin;
    hipStream_t st = c->stream;
    int64_t total = 0, total0 = 0;
    for (int l = 0; l < nlev; l++) {
        LevelSet one;
        one.P = {L.P[l]};
        one.T = {L.T[l]};
        one.pself = {L.pself[l]};
        one.outmul = {1.f};
        c->up_valid = false;  // the level and window buffers get other values here
        if (ensure(c, c->lvl, 64 + 3 * 4 + 2 * 8)) return MARFA_E_HIP;
        for (int q = 0; q < 4; q++)
            if (ensure(c, c->s->params[q], (size_t)nl * 8)) return MARFA_E_HIP;
        if (ensure(c, c->s->keys[0], (size_t)nl * 8) || ensure(c, c->s->idx[0], (size_t)nl * 4)) return MARFA_E_HIP;
        if (ensure(c, c->win, (size_t)w.nrec * 24 + 64)) return MARFA_E_HIP;
        float *dP = (float *)c->lvl.p, *dT = dP + 1;
        double *dps = (double *)(((uintptr_t)(dP + 3) + 15) & ~(uintptr_t)15);
        HIPCHK(hipMemcpyAsync(dP, &one.P[0], 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dT, &one.T[0], 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dps, &one.pself[0], 8, hipMemcpyHostToDevice, st));
        int64_t *dlo = (int64_t *)c->win.p, *dhi = dlo + w.nrec;
        double *dorg = (double *)(dhi + w.nrec);
        HIPCHK(hipMemcpyAsync(dlo, w.lo.data(), w.nrec * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dhi, w.hi.data(), w.nrec * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dorg, w.org.data(), w.nrec * 8, hipMemcpyHostToDevice, st));
        SpecArgs sa = {};
        sa.nlev = 1;
        sa.lmin = lmin;
        sa.nl = nl;
        sa.nu = c->d_nu;
        sa.S = c->d_S;
        sa.gair = c->d_gair;
        sa.gself = c->d_gself;
        sa.elow = c->d_elow;
        sa.nT = c->d_nT;
        sa.delta = c->d_delta;
        sa.molIso = c->d_molIso;
        sa.P = dP;
        sa.T = dT;
        sa.pself = dps;
        sa.tips = c->d_tips;
        sa.nus = (double *)c->s->params[0].p;
        sa.gD = (double *)c->s->params[1].p;
        sa.gL = (double *)c->s->params[2].p;
        sa.ST = (double *)c->s->params[3].p;
        sa.keys = (uint64_t *)c->s->keys[0].p;
        sa.idx = (int32_t *)c->s->idx[0].p;
        hipLaunchKernelGGL(spectroscopy_kernel, dim3((unsigned)((nl + BLOCK - 1) / BLOCK), 1), dim3(BLOCK), 0, st, sa);
        HIPCHK(hipGetLastError());
        KArgs ka = {};
        ka.nlev = 1;
        ka.nrec = w.nrec;
        ka.lmin = lmin;
        ka.nl = nl;
        ka.nus = sa.nus;
        ka.lo = dlo;
        ka.hi = dhi;
        ka.org = dorg;
        ka.cutoff = job->cutoff;
        if (ensure(c, c->counts, (size_t)w.nrec * 16 + 64)) return MARFA_E_HIP;
        hipLaunchKernelGGL(count_kernel, dim3(w.nrec), dim3(BLOCK), 0, st, ka, (unsigned long long *)c->counts.p);
        HIPCHK(hipGetLastError());
        std::vector<unsigned long long> h(2 * w.nrec);
        HIPCHK(hipMemcpyAsync(h.data(), c->counts.p, w.nrec * 16, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int q = 0; q < w.nrec; q++) {
            total += (int64_t)h[2 * q];
            total0 += (int64_t)h[2 * q + 1];
        }
    }
    samples[0] = total;
    samples[1] = total0;
    return 0;
}

int marfa_grid_sums(marfa_ctx *c, const marfa_job *job, int32_t level, int32_t rec, double *grids,
                    float *record) {
    int rc = check_job(c, job);
    if (rc) return rc;
    if (!grids || level < 0 || level >= (int)c->hP.size()) return set_err(c, MARFA_E_ARG, "bad level");
    (void)hipSetDevice(c->device);
    Windows w;
    compute_windows(c->h_nu, *job, w);
    if (rec < 0 || rec >= w.nrec) return set_err(c, MARFA_E_ARG, "bad record");
    Windows one;
    one.nrec = 1;
    one.lo = {w.lo[rec]};
    one.hi = {w.hi[rec]};
    one.rec = {w.rec[rec]};
    one.org = {w.org[rec]};
    LevelSet all = level_set(c, *job), L;
    L.P = {all.P[level]};
    L.T = {all.T[level]};
    L.pself = {all.pself[level]};
    L.outmul = {all.outmul[level]};
    const size_t gbytes = (size_t)(30690 + NT) * 8;
    if (ensure(c, c->out, (size_t)NT * 4 + gbytes + 64)) return MARFA_E_HIP;
    double *dg = (double *)(((uintptr_t)((float *)c->out.p + NT) + 15) & ~(uintptr_t)15);
    HIPCHK(hipMemsetAsync(dg, 0, gbytes, c->stream));
    c->dbg_grids = dg;
    rc = run(c, *job, L, one, (float *)c->out.p, c->stream, false);
    c->dbg_grids = nullptr;
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(grids, dg, gbytes, hipMemcpyDeviceToHost, c->stream));
    if (record) HIPCHK(hipMemcpyAsync(record, c->out.p, NT * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

namespace {
template <typename T>
int log_loss(int64_t n, const T *k, const T *khat, const float *w, double wscale, double *partial_dev,
             double *loss_dev, void *stream) {
    if (n < 0 || (n > 0 && (!k || !khat)) || !partial_dev || !loss_dev) return MARFA_E_ARG;
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(loss_partial_kernel<T>, dim3(LOSS_BLOCKS), dim3(BLOCK), 0, st, n, k, khat, w, wscale,
                       partial_dev);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(BLOCK), 0, st, (const double *)partial_dev, LOSS_BLOCKS,
                       loss_dev);
    return hipGetLastError() == hipSuccess ? 0 : MARFA_E_HIP;
}

template <typename T>
int log_loss_grad(int64_t n, const T *k, const T *khat, const float *w, double wscale, const double *gscale_dev,
                  T *grad_dev, void *stream) {
    if (n < 0 || (n > 0 && (!k || !khat || !grad_dev))) return MARFA_E_ARG;
    if (n == 0) return 0;
    const unsigned g = (unsigned)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 8192);
    hipLaunchKernelGGL(loss_grad_kernel<T>, dim3(g), dim3(BLOCK), 0, (hipStream_t)stream, n, k, khat, w, wscale,
                       gscale_dev, grad_dev);
    return hipGetLastError() == hipSuccess ? 0 : MARFA_E_HIP;
}
}  // namespace

extern "C" {

int marfa_log_loss(int64_t n, const float *k, const float *khat, const float *w, double wscale, double *partial_dev,
                   double *loss_dev, void *stream) {
    return log_loss<float>(n, k, khat, w, wscale, partial_dev, loss_dev, stream);
}
int marfa_log_loss_grad(int64_t n, const float *k, const float *khat, const float *w, double wscale,
                        const double *gscale_dev, float *grad_dev, void *stream) {
    return log_loss_grad<float>(n, k, khat, w, wscale, gscale_dev, grad_dev, stream);
}
int marfa_log_loss_f64(int64_t n, const double *k, const double *khat, const float *w, double wscale,
                       double *partial_dev, double *loss_dev, void *stream) {
    return log_loss<double>(n, k, khat, w, wscale, partial_dev, loss_dev, stream);
}
int marfa_log_loss_grad_f64(int64_t n, const double *k, const double *khat, const float *w, double wscale,
                            const double *gscale_dev, double *grad_dev, void *stream) {
    return log_loss_grad<double>(n, k, khat, w, wscale, gscale_dev, grad_dev, stream);
}

int marfa_last_timings(const marfa_ctx *c, float *ms5) {
    if (!c || !ms5) return MARFA_E_ARG;
    memcpy(ms5, c->timings, sizeof c->timings);
    return 0;
}

}  // extern "C"

// ====================================================================== batches of jobs
// Many independent jobs (each its own line list, levels, windows and records) recorded into
// one hipGraph: the jobs are spread over `lanes` parallel branches (fork/join on events);
// the jobs of a lane run in order and share that lane's scratch.  A job is a context whose
// scratch pointer, stream and TIPS are borrowed from the batch.
struct marfa_batch {
    int device = 0;
    std::string err;
    std::vector<marfa_ctx *> lanes;
    struct Job {
        marfa_ctx *ctx = nullptr;
        marfa_job spec{};
        LevelSet L;
        Windows W;
        int lane = 0;
        double work = 0;  // lines x levels x records, for the lane assignment
        bool ready = false;  // run once eagerly on its lane (buffers at their final size)
        float tune_ms[NVARIANTS] = {};
    };
    std::vector<Job> jobs;
    float *d_tips = nullptr;
    size_t lane_budget = (size_t)8 << 30;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> join;
    float last_ms = 0;
    // The captured form: every job's levels cut into pieces (level batches); pieces packed
    // into waves whose deposit inputs (sorted keys, line records, bucket index, level-0 sums)
    // fit one device arena.  A wave prepares its pieces on the lanes (spectroscopy and sort
    // per piece, lane temporaries), then runs one level-0 launch and one fine launch per
    // kernel variant over the work list of all its pieces (a device table of the pieces'
    // arguments; one workgroup per record or record tile).
    struct Piece {
        int job = 0, l0 = 0, nb = 0, lane = 0;
        BatchBufs B;
        KArgs ka;
        double work = 0;  // lines x levels x records: the order of the work lists
    };
    struct Launch {       // one launch over a work list
        bool coarse = true;
        int prec = 0, variant = 0;  // chi (level-0) or LDS stage capacity (fine)
        std::vector<int> pieces;
        size_t tab_off = 0, first_off = 0;  // in `tables`
        int items = 0, grid = 0;
    };
    struct Wave {
        std::vector<int> pieces;
        std::vector<Launch> launches;
    };
    std::vector<Piece> pieces;
    std::vector<Wave> waves;
    marfa_ctx::Buf arena, tables;
    // both halves of the wave arena: at most 96 GB, 12 lane budgets and 70 % of the free memory
    size_t arena_budget = (size_t)96 << 30;
    hipEvent_t done[2] = {nullptr, nullptr};  // deposits of the last wave in each half
};

namespace {
int bset_err(marfa_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    return code;
}
#define BHIPCHK(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return bset_err(b, MARFA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void release_job_ctx(marfa_ctx *c) {
    // borrowed members are the batch's
    c->s = &c->own;
    c->d_tips = nullptr;
    c->stream = nullptr;
    marfa_ctx_destroy(c);
}
}  // namespace

namespace {
// Cut every job into pieces (level batches whose deposit inputs take at most a quarter of
// a wave's half of the arena and whose temporaries fit a lane's budget), pack them into
// waves, give each piece a preparation lane (least prepared lines first) and its buffers
// (wave k in half k % 2 of the arena, so that a wave is prepared while the one before it
// deposits), build the work lists (pieces by descending work per item) and put their tables
// on the device.  Lanes 1.. prepare, lane 0's stream deposits (one lane: it does both).
int plan_waves(marfa_batch *b) {
    b->pieces.clear();
    b->waves.clear();
    // the lanes' arrays that only the eager runs used (the captured form keeps its deposit
    // inputs in the arena) are released first
    for (auto *l : b->lanes) {
        marfa_ctx::Buf *unused[] = {&l->own.keys[0], &l->own.idx[0], &l->own.coarse, &l->own.shapes, &l->own.buckets};
        for (auto *u : unused) {
            if (u->p) BHIPCHK(hipFree(u->p));
            u->p = nullptr;
            u->bytes = 0;
        }
    }
    size_t fr = 0, tot = 0;
    size_t budget = b->arena_budget;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = std::min(budget, fr / 10 * 7);  // room for the rest
    const size_t half = budget / 2, quarter = half / 4;
    const int K = (int)b->lanes.size();
    const int p0 = K > 1 ? 1 : 0;  // first preparation lane
    struct Need {
        size_t keep = 0, temp = 0;  // per level: deposit inputs, temporaries
    };
    std::vector<RunGeom> geo(b->jobs.size());
    std::vector<Need> need(b->jobs.size());
    for (size_t j = 0; j < b->jobs.size(); j++) {
        const auto &J = b->jobs[j];
        if (!geom_of(J.spec, J.W, geo[j])) continue;
        const RunGeom &g = geo[j];
        const size_t shp = J.spec.precision == MARFA_FP32K ? line_record_bytes<float>() : line_record_bytes<double>();
        need[j].keep = (size_t)g.nl * (8 + 4 + shp) + (size_t)(g.NB + 1) * 4 + (size_t)g.nrec * NCOARSE * 8 + 5 * 256;
        need[j].temp = (size_t)g.nl * (4 * 8 + 8 + 4) + (size_t)g.NB * 4 + (size_t)g.nsblk * 4;
        const int nlev = (int)J.L.P.size();
        const size_t cap_keep = std::max<size_t>(1, quarter / need[j].keep);
        const size_t cap_temp = std::max<size_t>(1, b->lane_budget / std::max<size_t>(need[j].temp, 1));
        const int lb = (int)std::min<size_t>((size_t)nlev, std::min(cap_keep, cap_temp));
        for (int l0 = 0; l0 < nlev; l0 += lb) {
            marfa_batch::Piece P;
            P.job = (int)j;
            P.l0 = l0;
            P.nb = std::min(lb, nlev - l0);
            P.work = (double)g.nl * g.nrec * P.nb;
            b->pieces.push_back(P);
        }
    }
    // waves: pieces in order while their deposit inputs fit half the arena (the first wave a
    // quarter of that, so that the deposits start early: the preparation of every later wave
    // runs beside the deposits of the wave before it)
    size_t used = 0, top = 0;
    for (size_t i = 0; i < b->pieces.size(); i++) {
        const size_t sz = need[b->pieces[i].job].keep * b->pieces[i].nb;
        if (b->waves.empty() || used + sz > (b->waves.size() == 1 ? quarter : half)) {
            b->waves.emplace_back();
            used = 0;
        }
        b->waves.back().pieces.push_back((int)i);
        used += sz;
        top = std::max(top, used);
    }
    // device memory: the arena's two halves, and per lane the temporaries of its largest piece
    marfa_ctx *c0 = b->lanes[0];
    const size_t hsz = (std::max<size_t>(top, 256) + 255) & ~(size_t)255;
    if (ensure(c0, b->arena, (b->waves.size() > 1 ? 2 : 1) * hsz)) return bset_err(b, MARFA_E_HIP, c0->err);
    std::vector<size_t> temp_of(K, 0);
    for (size_t w = 0; w < b->waves.size(); w++) {
        auto &W = b->waves[w];
        std::vector<double> load(K, 0.0);
        std::vector<int> order = W.pieces;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->pieces[x].work > b->pieces[y].work; });
        for (int i : order) {
            const int k = p0 + (int)(std::min_element(load.begin() + p0, load.end()) - (load.begin() + p0));
            auto &P = b->pieces[i];
            P.lane = k;
            load[k] += (double)geo[P.job].nl * P.nb;
        }
        Arena ar;
        ar.base = (char *)b->arena.p + (w % 2) * hsz;
        ar.size = hsz;
        for (int i : W.pieces) {
            auto &P = b->pieces[i];
            const auto &J = b->jobs[P.job];
            const RunGeom &g = geo[P.job];
            const size_t arr = (size_t)P.nb * g.nl;
            const size_t shp = J.spec.precision == MARFA_FP32K ? line_record_bytes<float>() : line_record_bytes<double>();
            P.B.key0 = ar.take(arr * 8);
            P.B.idx0 = ar.take(arr * 4);
            P.B.shapes = ar.take(arr * shp);
            P.B.buckets = ar.take((size_t)P.nb * (g.NB + 1) * 4);
            P.B.coarse = (double *)ar.take((size_t)P.nb * g.nrec * NCOARSE * 8);
            if (!P.B.key0 || !P.B.idx0 || !P.B.shapes || !P.B.buckets || !P.B.coarse)
                return bset_err(b, MARFA_E_ARG, "a wave outgrew the batch arena");
            temp_of[P.lane] = std::max(temp_of[P.lane], need[P.job].temp * P.nb);
        }
    }
    // lane temporaries: spectroscopy outputs, scattered keys and indices, histogram, scan
    for (int k = 0; k < K; k++) {
        marfa_ctx *l = b->lanes[k];
        marfa_ctx::Scratch &S = l->own;
        size_t arr_max = 0, nb_max = 0, ns_max = 0;
        for (const auto &P : b->pieces)
            if (P.lane == k) {
                arr_max = std::max(arr_max, (size_t)P.nb * geo[P.job].nl);
                nb_max = std::max(nb_max, (size_t)P.nb * geo[P.job].NB);
                ns_max = std::max(ns_max, (size_t)P.nb * geo[P.job].nsblk);
            }
        for (int q = 0; q < 4; q++)
            if (ensure(l, S.params[q], arr_max * 8)) return bset_err(b, MARFA_E_HIP, l->err);
        if (ensure(l, S.keys[1], arr_max * 8) || ensure(l, S.idx[1], arr_max * 4) || ensure(l, S.hist, nb_max * 4) ||
            ensure(l, S.spart, ns_max * 4))
            return bset_err(b, MARFA_E_HIP, l->err);
    }
    for (auto &P : b->pieces) {
        marfa_ctx::Scratch &S = b->lanes[P.lane]->own;
        for (int q = 0; q < 4; q++) P.B.par[q] = S.params[q].p;
        P.B.key1 = S.keys[1].p;
        P.B.idx1 = S.idx[1].p;
        P.B.hist = S.hist.p;
        P.B.spart = S.spart.p;
        auto &J = b->jobs[P.job];
        const RunArrays d = arrays_at(J.ctx->lvl.p, J.ctx->win.p, (int)J.L.P.size(), J.W.nrec);
        enqueue_prep(J.ctx, J.spec, geo[P.job], d, P.l0, P.nb, P.B, (float *)J.ctx->out.p, nullptr, nullptr, false,
                     P.ka, false);
        // the deposit kernels read the sorted arrays only: the lane temporaries are not theirs
        P.ka.nus = P.ka.gD = P.ka.gL = P.ka.S = nullptr;
    }
    // work lists: per wave, one level-0 launch per (form, chi) and one fine launch per (form,
    // stage capacity); pieces by descending work per item
    std::vector<char> host;
    auto put = [&](const void *src, size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        memcpy(host.data() + off, src, bytes);
        return off;
    };
    for (auto &W : b->waves) {
        for (int coarse = 1; coarse >= 0; coarse--) {
            std::vector<marfa_batch::Launch> ls;
            for (int i : W.pieces) {
                const auto &P = b->pieces[i];
                const auto &J = b->jobs[P.job];
                const int prec = J.spec.precision;
                const int var = coarse ? (J.spec.mol_code == 2 ? J.spec.chi_id : 0) : VARIANT_CAP[J.ctx->variant];
                auto it = std::find_if(ls.begin(), ls.end(),
                                       [&](const marfa_batch::Launch &x) { return x.prec == prec && x.variant == var; });
                if (it == ls.end()) {
                    ls.emplace_back();
                    it = ls.end() - 1;
                    it->coarse = coarse;
                    it->prec = prec;
                    it->variant = var;
                }
                it->pieces.push_back(i);
            }
            for (auto &L : ls) {
                std::stable_sort(L.pieces.begin(), L.pieces.end(), [&](int x, int y) {
                    const auto &X = b->pieces[x], &Y = b->pieces[y];
                    return (double)geo[X.job].nl / geo[X.job].nrec > (double)geo[Y.job].nl / geo[Y.job].nrec;
                });
                std::vector<KArgs> tab;
                std::vector<int> first(1, 0);
                for (int i : L.pieces) {
                    const auto &P = b->pieces[i];
                    tab.push_back(P.ka);
                    first.push_back(first.back() + P.nb * geo[P.job].nrec * (coarse ? 1 : TILES));
                }
                L.items = first.back();
                L.tab_off = put(tab.data(), tab.size() * sizeof(KArgs));
                L.first_off = put(first.data(), first.size() * sizeof(int));
                L.grid = 8 * XCD_RUN * ((L.items + 8 * XCD_RUN - 1) / (8 * XCD_RUN));
                W.launches.push_back(L);
            }
        }
    }
    if (ensure(c0, b->tables, std::max<size_t>(host.size(), 256))) return bset_err(b, MARFA_E_HIP, c0->err);
    BHIPCHK(hipMemcpy(b->tables.p, host.data(), host.size(), hipMemcpyHostToDevice));
    return 0;
}

// Record the waves into the capturing stream s0: jobs without lines (zero records), then
// per wave the pieces' preparation forked over the lanes, joined, and the work-list
// launches.
int capture_waves(marfa_batch *b, hipStream_t s0) {
    for (auto &J : b->jobs) {
        RunGeom g;
        if (J.W.nrec > 0 && !geom_of(J.spec, J.W, g)) {
            const int64_t words = (int64_t)J.L.P.size() * J.W.nrec * NT;
            hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<int64_t>((words + BLOCK - 1) / BLOCK, 65536)),
                               dim3(BLOCK), 0, s0, (int32_t *)J.ctx->out.p, words, 0);
        }
    }
    char *tb = (char *)b->tables.p;
    // the preparation lanes join the capture; wave k's preparation waits for the deposits of
    // wave k - 2, which read the same half of the arena (events done[k % 2])
    if (hipEventRecord(b->fork, s0) != hipSuccess) return bset_err(b, MARFA_E_HIP, "fork");
    for (size_t k = 1; k < b->lanes.size(); k++)
        if (hipStreamWaitEvent(b->lanes[k]->stream, b->fork, 0) != hipSuccess) return bset_err(b, MARFA_E_HIP, "fork");
    for (size_t w = 0; w < b->waves.size(); w++) {
        auto &W = b->waves[w];
        if (w >= 2)
            for (size_t k = 1; k < b->lanes.size(); k++)
                if (hipStreamWaitEvent(b->lanes[k]->stream, b->done[w % 2], 0) != hipSuccess)
                    return bset_err(b, MARFA_E_HIP, "wave order");
        for (int i : W.pieces) {
            auto &P = b->pieces[i];
            auto &J = b->jobs[P.job];
            RunGeom g;
            geom_of(J.spec, J.W, g);
            const RunArrays d = arrays_at(J.ctx->lvl.p, J.ctx->win.p, (int)J.L.P.size(), J.W.nrec);
            KArgs ka;
            const int rc = enqueue_prep(J.ctx, J.spec, g, d, P.l0, P.nb, P.B, (float *)J.ctx->out.p, nullptr,
                                        b->lanes[P.lane]->stream, false, ka);
            if (rc) return bset_err(b, rc, J.ctx->err);
        }
        for (size_t k = 1; k < b->lanes.size(); k++)
            if (hipEventRecord(b->join[k], b->lanes[k]->stream) != hipSuccess ||
                hipStreamWaitEvent(s0, b->join[k], 0) != hipSuccess)
                return bset_err(b, MARFA_E_HIP, "join");
        for (const auto &L : W.launches) {
            if (L.items == 0) continue;
            const KArgs *tab = (const KArgs *)(tb + L.tab_off);
            const int *first = (const int *)(tb + L.first_off);
            const int np = (int)L.pieces.size();
            const dim3 gr((unsigned)L.grid);
            const bool f32 = L.prec == MARFA_FP32K;
#define MARFA_MULTI(KERN) hipLaunchKernelGGL((KERN), gr, dim3(BLOCK), 0, s0, tab, first, np)
            if (L.coarse) {
                switch (L.variant) {
                case 1: if (f32) MARFA_MULTI((coarse_multi<float, 1>)); else MARFA_MULTI((coarse_multi<double, 1>)); break;
                case 2: if (f32) MARFA_MULTI((coarse_multi<float, 2>)); else MARFA_MULTI((coarse_multi<double, 2>)); break;
                case 3: if (f32) MARFA_MULTI((coarse_multi<float, 3>)); else MARFA_MULTI((coarse_multi<double, 3>)); break;
                default: if (f32) MARFA_MULTI((coarse_multi<float, 0>)); else MARFA_MULTI((coarse_multi<double, 0>)); break;
                }
            } else {
                switch (L.variant) {
                case 512: if (f32) MARFA_MULTI((fine_multi<float, 512>)); else MARFA_MULTI((fine_multi<double, 512>)); break;
                case 2048: if (f32) MARFA_MULTI((fine_multi<float, 2048>)); else MARFA_MULTI((fine_multi<double, 2048>)); break;
                default: if (f32) MARFA_MULTI((fine_multi<float, 1024>)); else MARFA_MULTI((fine_multi<double, 1024>)); break;
                }
            }
#undef MARFA_MULTI
            if (hipGetLastError() != hipSuccess) return bset_err(b, MARFA_E_HIP, "work-list launch");
        }
        if (hipEventRecord(b->done[w % 2], s0) != hipSuccess) return bset_err(b, MARFA_E_HIP, "wave order");
    }
    for (size_t k = 1; k < b->lanes.size(); k++)  // every lane joined before the capture ends
        if (hipEventRecord(b->join[k], b->lanes[k]->stream) != hipSuccess ||
            hipStreamWaitEvent(s0, b->join[k], 0) != hipSuccess)
            return bset_err(b, MARFA_E_HIP, "join");
    return 0;
}

}  // namespace

extern "C" {

int marfa_batch_create(int device, int32_t lanes, double lane_budget_gb, marfa_batch **out) {
    if (!out || lanes < 1 || lanes > 16) return MARFA_E_ARG;
    *out = nullptr;
    marfa_batch *b = new marfa_batch();
    b->device = device;
    if (lane_budget_gb > 0) b->lane_budget = (size_t)(lane_budget_gb * (double)(1u << 30));
    b->arena_budget = std::min(b->arena_budget, 12 * b->lane_budget);
    for (int k = 0; k < lanes; k++) {
        marfa_ctx *c = nullptr;
        int rc = marfa_ctx_create(device, &c);
        if (rc) {
            for (auto *l : b->lanes) marfa_ctx_destroy(l);
            delete b;
            return rc;
        }
        b->lanes.push_back(c);
    }
    (void)hipSetDevice(device);
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess ||
        hipEventCreate(&b->t1) != hipSuccess) {
        for (auto *l : b->lanes) marfa_ctx_destroy(l);
        delete b;
        return MARFA_E_HIP;
    }
    b->join.resize(lanes);
    for (auto &e : b->join) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    for (auto &e : b->done) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    *out = b;
    return 0;
}

int marfa_batch_destroy(marfa_batch *b) {
    if (!b) return MARFA_E_ARG;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto &j : b->jobs) release_job_ctx(j.ctx);
    for (auto *l : b->lanes) marfa_ctx_destroy(l);
    if (b->d_tips) (void)hipFree(b->d_tips);
    if (b->arena.p) (void)hipFree(b->arena.p);
    if (b->tables.p) (void)hipFree(b->tables.p);
    for (auto e : {b->fork, b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    for (auto e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b->done)
        if (e) (void)hipEventDestroy(e);
    delete b;
    return 0;
}

const char *marfa_batch_last_error(const marfa_batch *b) { return b ? b->err.c_str() : "null batch"; }

int marfa_batch_set_tips(marfa_batch *b, int32_t niso, int32_t ntemp, const float *tips) {
    if (!b || niso != 74 || ntemp != 501 || !tips) return bset_err(b, MARFA_E_ARG, "TIPS must be 74 x 501");
    (void)hipSetDevice(b->device);
    if (!b->d_tips) BHIPCHK(hipMalloc(&b->d_tips, (size_t)niso * ntemp * 4));
    BHIPCHK(hipMemcpy(b->d_tips, tips, (size_t)niso * ntemp * 4, hipMemcpyHostToDevice));
    return 0;
}

int marfa_batch_add_job(marfa_batch *b, const marfa_job *job, int64_t n, const double *nu, const float *S,
                        const float *gair, const float *gself, const float *elow, const float *nT,
                        const float *delta, const int32_t *molIso, int32_t nlev, const float *P, const float *T,
                        const float *rho, int32_t *job_id) {
    if (!b || !job || !job_id) return bset_err(b, MARFA_E_ARG, "null argument");
    if (!b->d_tips) return bset_err(b, MARFA_E_STATE, "TIPS must be set first");
    if (b->exec) return bset_err(b, MARFA_E_STATE, "the batch is already captured");
    (void)hipSetDevice(b->device);
    marfa_ctx *c = new marfa_ctx();
    c->device = b->device;
    for (auto &e : c->ev) (void)hipEventCreate(&e);
    auto fail = [&](int rc) {
        b->err = c->err;
        release_job_ctx(c);
        return rc;
    };
    c->d_tips = b->d_tips;
    c->tips_set = true;
    c->budget = b->lane_budget;
    int rc = marfa_set_lines(c, n, nu, S, gair, gself, elow, nT, delta, molIso);
    if (!rc) rc = marfa_set_levels(c, nlev, P, T, rho);
    if (!rc) rc = check_job(c, job);
    if (rc) return fail(rc);
    marfa_batch::Job J;
    J.ctx = c;
    J.spec = *job;
    compute_windows(c->h_nu, *job, J.W);
    J.L = level_set(c, *job);
    const size_t total = (size_t)nlev * J.W.nrec * NT;
    if (ensure(c, c->out, std::max<size_t>(total, 1) * 4)) return fail(MARFA_E_HIP);
    int64_t lines = 0;
    if (J.W.nrec > 0)
        lines = *std::max_element(J.W.hi.begin(), J.W.hi.end()) - *std::min_element(J.W.lo.begin(), J.W.lo.end());
    J.work = (double)std::max<int64_t>(lines, 1) * nlev * std::max(J.W.nrec, 1);
    *job_id = (int32_t)b->jobs.size();
    b->jobs.push_back(std::move(J));
    return 0;
}

// Lane assignment (largest job first onto the least loaded lane), then per job: time every
// fine-kernel variant on the job's first level and keep the fastest, then one full eager run
// (records, and every buffer at its final size so that the capture allocates nothing).
int marfa_batch_tune(marfa_batch *b, int32_t *variants, float *tune_ms) {
    if (!b) return MARFA_E_ARG;
    if (b->exec) return bset_err(b, MARFA_E_STATE, "the batch is already captured");
    (void)hipSetDevice(b->device);
    const int K = (int)b->lanes.size();
    std::vector<int> order(b->jobs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->jobs[x].work > b->jobs[y].work; });
    std::vector<double> load(K, 0.0);
    for (int j : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->jobs[j].lane = k;
        load[k] += b->jobs[j].work;
    }
    for (size_t j = 0; j < b->jobs.size(); j++) {
        auto &J = b->jobs[j];
        marfa_ctx *c = J.ctx;
        marfa_ctx *lane = b->lanes[J.lane];
        c->s = &lane->own;
        c->stream = lane->stream;
        float *out = (float *)c->out.p;
        int best = 1;
        if (J.W.nrec > 0) {
            LevelSet L1;
            L1.P = {J.L.P[0]};
            L1.T = {J.L.T[0]};
            L1.pself = {J.L.pself[0]};
            L1.outmul = {J.L.outmul[0]};
            float bms = 0;
            for (int v = 0; v < NVARIANTS; v++) {
                c->variant = v;
                int rc = run(c, J.spec, L1, J.W, out, lane->stream, true);
                if (!rc) rc = run(c, J.spec, L1, J.W, out, lane->stream, true);  // second run timed warm
                if (rc) {
                    b->err = c->err;
                    return rc;
                }
                J.tune_ms[v] = c->timings[3];
                if (v == 0 || J.tune_ms[v] < bms) {
                    bms = J.tune_ms[v];
                    best = v;
                }
            }
        }
        c->variant = best;
        int rc = run(c, J.spec, J.L, J.W, out, lane->stream, false);
        if (rc) {
            b->err = c->err;
            return rc;
        }
        J.ready = true;
        if (variants) variants[j] = best;
        if (tune_ms)
            for (int v = 0; v < NVARIANTS; v++) tune_ms[j * NVARIANTS + v] = J.tune_ms[v];
    }
    for (auto *l : b->lanes) BHIPCHK(hipStreamSynchronize(l->stream));
    return 0;
}

int marfa_batch_capture(marfa_batch *b) {
    if (!b) return MARFA_E_ARG;
    if (b->exec) return 0;
    for (const auto &J : b->jobs)
        if (!J.ready) return bset_err(b, MARFA_E_STATE, "tune the batch before capturing it");
    (void)hipSetDevice(b->device);
    for (auto *l : b->lanes) BHIPCHK(hipStreamSynchronize(l->stream));
    const int rc = plan_waves(b);
    if (rc) return rc;
    hipStream_t s0 = b->lanes[0]->stream;
    BHIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
    int err = capture_waves(b, s0);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s0, &g);
    if (err || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!err) b->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return err ? err : MARFA_E_HIP;
    }
    b->graph = g;
    BHIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return 0;
}

int marfa_batch_launch(marfa_batch *b, void *stream, int32_t sync) {
    if (!b) return MARFA_E_ARG;
    if (!b->exec) return bset_err(b, MARFA_E_STATE, "capture the batch first");
    (void)hipSetDevice(b->device);
    hipStream_t st = stream ? (hipStream_t)stream : b->lanes[0]->stream;
    BHIPCHK(hipEventRecord(b->t0, st));
    BHIPCHK(hipGraphLaunch(b->exec, st));
    BHIPCHK(hipEventRecord(b->t1, st));
    if (sync) {
        BHIPCHK(hipEventSynchronize(b->t1));
        BHIPCHK(hipEventElapsedTime(&b->last_ms, b->t0, b->t1));
    }
    return 0;
}

int marfa_batch_info(marfa_batch *b, int32_t job, int32_t *nlev, int32_t *nrec, float **records_dev, float *last_ms) {
    if (!b) return MARFA_E_ARG;
    if (last_ms) *last_ms = b->last_ms;
    if (job < 0) return 0;
    if (job >= (int32_t)b->jobs.size()) return bset_err(b, MARFA_E_ARG, "no such job");
    const auto &J = b->jobs[job];
    if (nlev) *nlev = (int32_t)J.L.P.size();
    if (nrec) *nrec = J.W.nrec;
    if (records_dev) *records_dev = (float *)J.ctx->out.p;
    return 0;
}

int marfa_batch_records(marfa_batch *b, int32_t job, float *out_host) {
    if (!b || !out_host) return MARFA_E_ARG;
    if (job < 0 || job >= (int32_t)b->jobs.size()) return bset_err(b, MARFA_E_ARG, "no such job");
    (void)hipSetDevice(b->device);
    for (auto *l : b->lanes) BHIPCHK(hipStreamSynchronize(l->stream));
    BHIPCHK(hipDeviceSynchronize());
    const auto &J = b->jobs[job];
    const size_t total = J.L.P.size() * (size_t)J.W.nrec * NT;
    if (total) BHIPCHK(hipMemcpy(out_host, J.ctx->out.p, total * 4, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
