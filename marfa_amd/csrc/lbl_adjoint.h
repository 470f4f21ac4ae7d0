// lbl_adjoint.h -- line pass of the adjoint: dL/d(line parameters) and dL/d(level T, P)
// from dL/d(pre-cascade node) of every record (adjoint_tile_kernel).  Included by
// marfa_lbl.hip inside its kernel namespace.
//
// One thread per (line, level) owns that line's gradient, so no sum needs atomics: it
// visits every record whose window holds the line, walks the line's samples in the
// reference's deposit geometry (src/LineGridCalc.f90; the same closed forms as the
// forward gathers), and accumulates dL/dnode x dF/d(S(T), gD, gL, nu*) in FP64, with F and
// its derivatives recomputed in the form's precision (shape_grad).  The chain to the
// parameters follows src/Spectroscopy.f90:38-145 and app/main.f90:92 (pSelf is
// proportional to T at fixed density).
#pragma once

struct AdjArgs {
    const double *nu;                                // [n] unshifted centres (context)
    const float *S, *gair, *gself, *elow, *nT, *delta;
    const int32_t *molIso;
    const float *tips;                               // [74][501]
    const float *P;                                  // [nlev] absolute level index
    const double *pself;                             // [nlev]
    const float *gbar;                               // [batch records][30690 + NT]
    const float *gpos;                               // [batch records][GPOS]: adjoint_gpos_kernel
    double *gline;                                   // [batch levels][nl][4]: S, gair, nT, delta
    double *glev;                                    // [batch levels][blocks][2]: T, P
};

__device__ inline int node_p(int k, int I) { return 30 * ((1 << k) - 1) + (I - 1); }
__device__ inline int node_m(int k, int I) { return 30 * ((1 << k) - 1) + (10 << k) + (I - 1); }
__device__ inline int node_l(int k, int I) { return 30 * ((1 << k) - 1) + 2 * (10 << k) + (I - 1); }
__device__ inline int node_rk(int i) { return 30690 + i - 1; }

// Per record, the node adjoints of the 21 positions q/2 cm-1 of grid 0 with every role at a
// position summed (q even: L of cell q/2 and P of cell q/2 + 1, and RK(1) at q = 0; q odd:
// the mid of cell (q + 1)/2): a wing line 1 cm-1 or more outside the subinterval samples
// every position in every role (src/LineGridCalc.f90:21-33, :508-518), so this is its weight.
constexpr int GPOS = 24;  // 21 positions, padded

__global__ __launch_bounds__(BLOCK) void adjoint_gpos_kernel(const float *gbar, int nrecs, float *gpos) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nrecs * GPOS) return;
    const int r = i / GPOS, q = i % GPOS;
    const float *gr = gbar + (int64_t)r * (30690 + NT);
    float g = 0.f;
    if (q <= 20) {
        if (q & 1) {
            g = gr[10 + (q >> 1)];
        } else {
            if (q > 0) g += gr[20 + (q >> 1) - 1];
            if (q < 20) g += gr[q >> 1];
            if (q == 0) g += gr[30690];
        }
    }
    gpos[i] = g;
}

template <typename R>
struct LineAdj {
    const float *gr;  // dL/dnode of the record
    const float *gp;  // the record's position adjoints (adjoint_gpos_kernel)
    const LineShape<R> *p;
    const ShapeCtl *ctl;
    double u;         // nu* - origin
    double aS, aX, aD, aL, aT;
    __device__ inline void acc(double x, double gs) {
        if (gs == 0.0) return;
        const SampleGrad<R> s = shape_grad<R>(u - x, *p, *ctl);
        aS += gs * (double)s.v;
        aX += gs * (double)s.dX;
        aD += gs * (double)s.dgD;
        aL += gs * (double)s.dgL;
        aT += gs * (double)s.dT;
    }
    // positions q = qlo..qhi (steps of H_k/2) of a run of cells of grid k: even q is the L
    // edge of cell q/2 (q > qlo) and the P edge of cell q/2+1 (q < qhi); odd q a mid point.
    // A centre line's consecutive grids share one position (the last of grid k's left walk
    // is the first of grid k+1's, the first of grid k's right walk the last of grid k+1's,
    // x formed exactly either way), sampled once: cfirst / clast are node adjoints that
    // another walk hands over for the first / last position, and with dfirst / dlast that
    // position is not sampled here but its adjoint returned for the other walk.
    __device__ inline double walk(int k, int qlo, int qhi, double cfirst = 0.0, double clast = 0.0,
                                  bool dfirst = false, bool dlast = false) {
        const double h = 0.5 * hk(k);
        double ret = 0.0;
        for (int q = qlo; q <= qhi; q++) {
            float g;
            if (q & 1) {
                g = gr[node_m(k, (q + 1) >> 1)];
            } else {
                g = 0.f;
                if (q > qlo) g += gr[node_l(k, q >> 1)];
                if (q < qhi) g += gr[node_p(k, (q >> 1) + 1)];
            }
            double gs = g;
            if (q == qlo) gs += cfirst;
            if (q == qhi) gs += clast;
            if ((q == qlo && dfirst) || (q == qhi && dlast)) {
                ret = gs;
                continue;
            }
            acc((double)q * h, gs);
        }
        return ret;
    }
    // Does a wing line dmin >= 1 cm-1 outside the subinterval have its nearest node -- and
    // so every node -- on the chi-corrected Lorentz branch (the level-0 kernel's test)?
    __device__ inline bool far_branch(double dmin) const {
        const R margin = sizeof(R) == 8 ? R(15.0 * (1.0 + 1e-12)) : R(15.0f * (1.0f + 4e-6f));
        return ctl->shape_id == 1 || (R)dmin * p->xs >= margin;
    }
    // Such a line (lw: left of the subinterval): its 21 positions at |X| = dmin + q/2 (left)
    // or dmin + (20 - q)/2 (right), weighted by the record's position adjoints gp, summed
    // over the record in precision R and added to the FP64 sums once.
    // CHIF false: the job has no chi (the caller's launch), and the three sums reduce to
    // A = sum g/d, B = sum g/d^2, C = sum g |X|/d^2 (d = X^2 + gL^2): sample sum lorc A,
    // d/d|X| -2 lorc C, d/dgL S/pi (X^2 - gL^2)/d^2 summed = S/pi (A - 2 gL^2 B).
    template <bool CHIF = true>
    __device__ inline void far(double dmin, bool lw) {
        const float4 *g4 = reinterpret_cast<const float4 *>(gp);  // GPOS floats, 16-byte aligned
        const R base = lw ? (R)dmin : (R)(dmin + 10.0), step = lw ? R(0.5) : R(-0.5);
        const R spi = p->S / R(MARFA_PI_F), gl2 = p->gl2, lorc = p->lorc;
        if constexpr (!CHIF) {
            R A = R(0), B = R(0), C = R(0);
#pragma unroll 1
            for (int q0 = 0; q0 <= 20; q0 += 4) {
                const float4 x = g4[q0 >> 2];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int q = q0 + e;
                    if (q > 20) break;
                    const R gs = (R)(e == 0 ? x.x : e == 1 ? x.y : e == 2 ? x.z : x.w);
                    const R ax = fma(step, (R)q, base);
                    const R inv = recip<R>(ax * ax + gl2);
                    const R gi = gs * inv, gi2 = gi * inv;
                    A += gi;
                    B += gi2;
                    C = fma(gi2, ax, C);
                }
            }
            const double sX = -2.0 * (double)lorc * (double)C;
            aS += (double)lorc * (double)A;
            aX += lw ? -sX : sX;  // X = -|X| on the left wing
            aL += (double)spi * ((double)A - 2.0 * (double)gl2 * (double)B);
            return;
        }
        const bool chi = ctl->mol == 2 && ctl->chi_id != 0;
        R sS = R(0), sX = R(0), sL = R(0), sT = R(0);
        // positions four at a time (one 16-byte load of their adjoints)
#pragma unroll 1
        for (int q0 = 0; q0 <= 20; q0 += 4) {
            const float4 x = g4[q0 >> 2];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int q = q0 + e;
                if (q > 20) break;
                const R gs = (R)(e == 0 ? x.x : e == 1 ? x.y : e == 2 ? x.z : x.w);
                const R ax = fma(step, (R)q, base);
                const R inv = recip<R>(ax * ax + gl2);
                R c = R(1), dc = R(0);
                if (chi) {
                    if constexpr (sizeof(R) == 8) {
                        c = chi_factor<double>(ctl->chi_id, ax, ctl->B1d, ctl->B2d);
                        dc = dchi_factor<double>(ctl->chi_id, ax, ctl->B1d, ctl->B2d, c);
                    } else {
                        c = chi_factor<float>(ctl->chi_id, ax, ctl->B1, ctl->B2);
                        dc = dchi_factor<float>(ctl->chi_id, ax, ctl->B1, ctl->B2, c);
                    }
                }
                const R gl = gs * lorc * inv;  // gs x Lorentz
                const R gv = gl * c;           // gs x sample
                sS += gv;
                // d/d|X| of Lorentz x chi: -2|X| Lorentz chi / (X^2 + gL^2) + Lorentz dchi/d|X|
                sX += gl * (dc - R(2) * ax * inv * c);
                // d/dgL: S/pi (X^2 - gL^2) / (X^2 + gL^2)^2 chi
                sL += gs * spi * (ax * ax - gl2) * inv * inv * c;
                if (chi && ctl->chi_id == 3 && ax > R(3)) {
                    // Perrin: d chi / dT = chi de/dT (chilor_grad)
                    const R dB1 = (R)ctl->dB1, dB2 = (R)ctl->dB2;
                    const R de = ax < R(30) ? -dB1 * (ax - R(3))
                                 : ax < R(120) ? -R(27) * dB1 - dB2 * (ax - R(30))
                                               : -R(27) * dB1 - R(90) * dB2;
                    sT += gv * de;
                }
            }
        }
        aS += (double)sS;
        aX += lw ? -(double)sX : (double)sX;  // X = -|X| on the left wing
        aL += (double)sL;
        aT += (double)sT;
    }
    // centerLBL (src/LineGridCalc.f90:181-494): RK(1), the two walks and the core points
    __device__ inline void centre() {
        // walk ends e_k on the left (w = u) and right (w = 10 - u) sides
        auto ends = [&](double w, int k) { return w >= 2.0 * hk(k) ? (int)floor(__builtin_ldexp(w, k)) - 1 : 0; };
        // RK(1) is at x = 0, the first position of the grid-0 left walk when there is one
        int ep = 0, e = ends(u, 0);
        double c = 0.0;
        if (e > 0) c = gr[node_rk(1)];
        else acc(0.0, gr[node_rk(1)]);
        for (int k = 0; k <= 9; k++) {
            const int en = k < 9 ? ends(u, k + 1) : 0;
            // grid k + 1 walks (en > 2 e) from this walk's last position
            if (e > 2 * ep) c = walk(k, 4 * ep, 2 * e, c, 0.0, false, k < 9 && en > 2 * e);
            ep = e;
            e = en;
        }
        const int fu = (int)floor(u / HF);
        if (u >= HF)
            for (int i = 4 * ep + 2; i <= fu + 1; i++) acc((double)(i - 1) * HF, gr[node_rk(i)]);
        const double v = 10.0 - u;
        int er = 0;
        e = ends(v, 0);
        c = 0.0;
        for (int k = 0; k <= 9; k++) {
            const int en = k < 9 ? ends(v, k + 1) : 0;
            const int two_nt = 20 << k;
            // grid k + 1 walks (en > 2 e) up to this walk's first position
            if (e > 2 * er) c = walk(k, two_nt - 2 * e, two_nt - 4 * er, 0.0, c, k < 9 && en > 2 * e, false);
            er = e;
            e = en;
        }
        for (int i = fu + 2; i <= NT - 4 * er; i++) acc((double)(i - 1) * HF, gr[node_rk(i)]);
    }
    // leftLBL (:7-179): the line's right wing from the subinterval's left edge (far wings
    // on the Lorentz branch take far() instead)
    __device__ inline void left() {
        const double d = -u;
        acc(0.0, gr[node_rk(1)]);
        if (d >= 1.0) {
            walk(0, 0, 20);
            return;
        }
        int m = 1;
        while (m <= 9 && d < hk(m)) m++;
        if (m <= 9) {
            acc(0.0, gr[node_p(m, 1)]);
            acc(0.5 * hk(m), gr[node_m(m, 1)]);
            acc(hk(m), gr[node_l(m, 1)]);
        } else {
            for (int i = 2; i <= 5; i++) acc((double)(i - 1) * HF, gr[node_rk(i)]);
        }
        for (int k = m < 9 ? m : 9; k >= 1; k--) {
            acc(hk(k), gr[node_p(k, 2)]);
            acc(1.5 * hk(k), gr[node_m(k, 2)]);
            acc(2.0 * hk(k), gr[node_l(k, 2)]);
        }
        walk(0, 2, 20);
    }
    // rightLBL (:496-674): the line's left wing from the subinterval's right edge (far wings
    // on the Lorentz branch take far() instead)
    __device__ inline void right() {
        const double dr = u - 10.0;
        acc(0.0, gr[node_rk(1)]);
        if (dr >= 1.0) {
            walk(0, 0, 20);
            return;
        }
        int m = 1;
        while (m <= 9 && dr < hk(m)) m++;
        if (m <= 9) {
            const int N = 10 << m;
            acc(10.0, gr[node_l(m, N)]);
            acc(10.0 - 0.5 * hk(m), gr[node_m(m, N)]);
            acc(10.0 - hk(m), gr[node_p(m, N)]);
        } else {
            for (int i = NT; i >= NT - 3; i--) acc((double)(i - 1) * HF, gr[node_rk(i)]);
        }
        for (int k = m < 9 ? m : 9; k >= 1; k--) {
            const int N = (10 << k) - 1;
            acc(10.0 - hk(k), gr[node_l(k, N)]);
            acc(10.0 - 1.5 * hk(k), gr[node_m(k, N)]);
            acc(10.0 - 2.0 * hk(k), gr[node_p(k, N)]);
        }
        walk(0, 0, 18);
    }
};

// Two passes over the (line, level) pairs, one thread each, so that the far wings -- the
// bulk of the samples, in few registers -- run at a high occupancy apart from the
// register-hungry rest.  PASS 0: far wings on the Lorentz branch (far()); it writes the
// line's gradient slots and its own per-block level partials.  PASS 1: every other deposit
// (centre lines, wings within 1 cm-1, far wings off that branch); launched after pass 0, its
// thread adds to the slots the same line's pass-0 thread wrote and writes its own level
// partials.  The chain rule is linear in the sums, so each pass applies it to its own.
template <typename R, int PASS, bool CHIF = true>
__global__ __launch_bounds__(BLOCK, PASS == 0 ? 4 : sizeof(R) == 4 ? 3 : 2) void adjoint_line_kernel(KArgs a, AdjArgs b) {
    const int lev = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double cT = 0.0, cP = 0.0;
    if (j < a.nl) {
        const int64_t i = a.lmin + j, o = (int64_t)lev * a.nl + j;
        const int L = a.lev0 + lev;
        const double nus = a.nus[o], gD = a.gD[o], gL = a.gL[o], ST = a.S[o];
        const double sc = a.sscale[L];
        const LineShape<R> p = make_line<R>(gD, gL, ST * sc);
        const ShapeCtl ctl = make_ctl(a, L);
        LineAdj<R> w;
        w.p = &p;
        w.ctl = &ctl;
        w.aS = w.aX = w.aD = w.aL = w.aT = 0.0;
        // records whose window [lo, hi) holds line i: hi is non-decreasing in r
        int r = 0, hiR = a.nrec;
        while (r < hiR) {
            const int mid = (r + hiR) >> 1;
            if (a.hi[mid] > i) hiR = mid;
            else r = mid + 1;
        }
        const double cut = (double)a.cutoff;
        // pass 1 works only where the line is no far wing on the Lorentz branch: a wing at
        // d >= dfar is one (d >= 1 and (R)d xs >= the branch margin, with a relative 1e-5 to
        // spare over the rounding of (R)d xs), so only the records with origins in
        // (nus - 10 - dfar, nus + dfar) are visited (subinterval origins increase with r)
        double dfar = 1.0;
        if (PASS == 1) {
            const R margin = sizeof(R) == 8 ? R(15.0 * (1.0 + 1e-12)) : R(15.0f * (1.0f + 4e-6f));
            if (ctl.shape_id != 1) dfar = fmax(1.0, (double)margin / (double)p.xs * (1.0 + 1e-5));
            const double o0 = nus - 10.0 - dfar;
            int hiO = a.nrec;
            while (r < hiO) {
                const int mid = (r + hiO) >> 1;
                if (a.org[mid] > o0) hiO = mid;
                else r = mid + 1;
            }
        }
        for (; r < a.nrec && a.lo[r] <= i; r++) {
            if (PASS == 1 && a.org[r] >= nus + dfar) break;
            if (i >= a.hi[r]) continue;
            const double org = a.org[r];
            const int64_t rr = (int64_t)lev * a.nrec + r;
            w.u = nus - org;
            w.gr = b.gbar + rr * (30690 + NT);
            w.gp = b.gpos + rr * GPOS;
            if (nus < org || nus > org + 10.0) {
                const bool lw = nus < org;
                const double d = lw ? -w.u : w.u - 10.0;
                // leftLBL / rightLBL cutoff tests (src/LineGridCalc.f90:23, :508-509)
                if (lw ? d > cut : d >= cut) continue;
                const bool fw = d >= 1.0 && w.far_branch(d);
                if (PASS == 0 && fw) w.template far<CHIF>(d, lw);
                if (PASS == 1 && !fw) {
                    if (lw) w.left();
                    else w.right();
                }
            } else if (PASS == 1 && w.u < 10.0) {
                w.centre();
            }
        }
        // chain rule to the parameters (src/Spectroscopy.f90:38-145)
        const double T = (double)a.T[L], P = (double)b.P[L], ps = b.pself[L];
        const double S0 = (double)b.S[i], ga = (double)b.gair[i], gs = (double)b.gself[i];
        const double n = (double)b.nT[i], dl = (double)b.delta[i], E = (double)b.elow[i];
        const double tr = (double)MARFA_TREF_F / T, fT = pow(tr, n);
        const double dnu = w.aX + w.aD * gD / nus;     // gD is proportional to nu*
        double *gl = b.gline + o * 4;
        const double g0 = S0 != 0.0 ? w.aS / S0 : 0.0;  // S(T) is proportional to S
        const double g1 = w.aL * fT * (P - ps);
        const double g2 = w.aL * gL * log(tr);
        const double g3 = dnu * P;                       // nu* = nu + delta P
        if (PASS == 0) {
            gl[0] = g0;
            gl[1] = g1;
            gl[2] = g2;
            gl[3] = g3;
        } else {
            gl[0] += g0;
            gl[1] += g1;
            gl[2] += g2;
            gl[3] += g3;
        }
        // d ln S(T) / dT: partition sum (linear in the TIPS table), Boltzmann factor and
        // stimulated emission (unshifted nu)
        const int iso = b.molIso[i] / 100;
        const int ti = (int)((T - 20.0) / 2.0 + 1.0);
        const double wu = (T - ((double)ti * 2.0 + 18.0)) / 2.0;
        const float *row = b.tips + (iso - 1) * 501;
        const double Q = (1.0 - wu) * (double)row[ti - 1] + wu * (double)row[ti];
        const double dQ = ((double)row[ti] - (double)row[ti - 1]) / 2.0;
        const double c2 = (double)MARFA_C2_F, x = c2 * b.nu[i] / T;
        const double dlnS = -dQ / Q + c2 * E / (T * T) - (x / T) * exp(-x) / (1.0 - exp(-x));
        cT = w.aS * dlnS + w.aD * gD / (2.0 * T) + w.aL * (-n * gL / T + fT * (gs - ga) * ps / T) + w.aT;
        cP = dnu * dl + w.aL * fT * ga;
    }
    // per-block partial sums of the level gradients, fixed order; glev is [pass][level][block]
    for (int off = 32; off > 0; off >>= 1) {
        cT += __shfl_xor(cT, off);
        cP += __shfl_xor(cP, off);
    }
    __shared__ double red[2][BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = cT;
        red[1][threadIdx.x >> 6] = cP;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double x = 0.0;
        for (int w2 = 0; w2 < BLOCK / 64; w2++) x += red[threadIdx.x][w2];
        b.glev[(((int64_t)PASS * gridDim.y + lev) * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = x;
    }
}

// running per-line gradient += the batch's levels, in level order
__global__ __launch_bounds__(BLOCK) void adjoint_reduce_kernel(const double *gline, int nb, int64_t nl, double *gacc) {
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= nl) return;
    double s0 = gacc[j * 4], s1 = gacc[j * 4 + 1], s2 = gacc[j * 4 + 2], s3 = gacc[j * 4 + 3];
    for (int l = 0; l < nb; l++) {
        const double *g = gline + ((int64_t)l * nl + j) * 4;
        s0 += g[0];
        s1 += g[1];
        s2 += g[2];
        s3 += g[3];
    }
    gacc[j * 4] = s0;
    gacc[j * 4 + 1] = s1;
    gacc[j * 4 + 2] = s2;
    gacc[j * 4 + 3] = s3;
}

// per-level dL/dT and dL/dP of a level batch: the block partials of both passes of
// adjoint_line_kernel (glev [pass][level][block][2]) summed in a fixed order (strided per
// thread, then wave butterflies, then waves in order); one workgroup per (level, quantity);
// out[l] (T) and out[nlev + l] (P) for the batch's levels from l0
__global__ __launch_bounds__(BLOCK) void adjoint_level_kernel(const double *glev, int nb, int nblk, int nlev, int l0,
                                                             double *out) {
    const int q = blockIdx.x >> 1, c = blockIdx.x & 1;
    if (q >= nb) return;
    double x = 0.0;
    for (int b2 = threadIdx.x; b2 < 2 * nblk; b2 += BLOCK) {
        const int pass = b2 >= nblk, blk = b2 - pass * nblk;
        x += glev[(((int64_t)pass * nb + q) * nblk + blk) * 2 + c];
    }
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    __shared__ double red[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
        for (int w = 1; w < BLOCK / 64; w++) s += red[w];
        out[(int64_t)c * nlev + l0 + q] = s;
    }
}

// per-line gradients [nl][4] (lines lmin .. lmin + nl - 1) into the caller's layout
// out[q * n + i] (q = S, gamma_air, n_T, delta), zero for the lines outside every window
__global__ __launch_bounds__(BLOCK) void adjoint_scatter_kernel(const double *gacc, int64_t lmin, int64_t nl, int64_t n,
                                                               double *out) {
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const bool in = i >= lmin && i < lmin + nl;
#pragma unroll
        for (int q = 0; q < 4; q++) out[q * n + i] = in ? gacc[(i - lmin) * 4 + q] : 0.0;
    }
}
