// lbl_fine.h -- grids 1..9 and the fine points of one 1 cm-1 tile, by gather, then the
// tile's cascade.  Included by marfa_lbl.hip inside its kernel namespace.
//
// Every sum of the tile has one owner, which adds its lines in the order of their shifted
// centres, so no sum depends on scheduling and no atomics are used.  Grids 1..9 are summed
// by cell pairs: the lines of one walk group deposit into one pair of cells and nowhere
// else on that grid (pair_walk), so a pair's six sums are owned by a set of threads that
// split its lines and meet in a fixed butterfly; on grids 6..9 of a staged tile a wavefront
// owns a quarter of the pairs and takes one line per lane for all four grids at once
// (level_lanes), summing each pair's lines by a lane scan.  Fine points are owned by threads, four
// per grid-9 cell, which share one candidate range.  Candidate ranges come from a
// per-level index of the sorted shifted centres in 1/512 cm-1 buckets; grids 4..9 and the
// fine points read their centre lines from an LDS copy of the tile's reach (u and sorted
// index, with a local bucket table), grids 1..3 (whose groups reach up to 2 cm-1 beyond
// the tile) and the wing lines read the level's sorted line records.  The next line's
// record is loaded while the current line is evaluated.
//
// For a line at u = nu* - a (src/LineGridCalc.f90:191) the left walk of centerLBL
// (:201-331) covers, on grid k, the cells (2 e_{k-1}, e_k] with e_k = max(0, floor(u 2^k) - 1);
// the right walk (:343-484) the mirror image in v = 10 - u; the core fine points are
// [4 e_9 + 2, floor(u/H) + 1] (:332-339) and [floor(u/H) + 2, NT - 4 e'_9] (:485-490).
// leftLBL / rightLBL deposit into cells 1, 2 (NT_k, NT_k - 1) of grids k <= m, where
// m = #{k : d < H_k}, and into fine points 2..5 (NT-3..NT) when m = 10.
#pragma once


__device__ inline int bucket_of(double nus, double base, int NB) {
    const double x = (nus - base) * BUCKETS_PER_CM;
    if (!(x >= 0.0)) return 0;
    if (x >= (double)(NB - 1)) return NB - 1;
    return (int)x;
}

template <typename R>
struct FineCtx {
    const uint64_t *skey;
    const int32_t *sidx;
    const LineShape<R> *shp;
    const int32_t *bst;
    double org, top, base, cut;
    int NB;
    int64_t wlo, whi;
    ShapeCtl ctl;
};

// sorted candidate range [j0, j1) holding every line with u = nu* - org in [u0, u1)
template <typename R>
__device__ inline void cand(const FineCtx<R> &f, double u0, double u1, int &j0, int &j1) {
    const double x0 = (f.org + u0 - f.base) * BUCKETS_PER_CM;
    const double x1 = (f.org + u1 - f.base) * BUCKETS_PER_CM;
    const int b0 = (int)fmin(fmax(floor(x0) - 1.0, 0.0), (double)f.NB);
    const int b1 = (int)fmin(fmax(floor(x1) + 2.0, 0.0), (double)f.NB);
    j0 = f.bst[b0];
    j1 = b1 > b0 ? f.bst[b1] : j0;
}

// 0: no deposit beyond grid 0; 1: left wing with m >= 1; 2: centre line; 3: right wing
// with m >= 1.  u = nu* - org as src/LineGridCalc.f90 forms it.
template <typename R>
__device__ inline int classify(const FineCtx<R> &f, int j, double &u) {
    const double nus = double_of(f.skey[j]);
    u = nus - f.org;
    const int64_t li = f.sidx[j];
    if (li < f.wlo || li >= f.whi) return 0;
    if (nus < f.org) {
        const double d = -u;
        return (!(d > f.cut) && d < 1.0) ? 1 : 0;
    }
    if (nus > f.top) {
        const double dr = u - 10.0;
        return (!(dr >= f.cut) && dr < 1.0) ? 3 : 0;
    }
    return u < 10.0 ? 2 : 0;
}

// Candidate lines read straight from the level's sorted arrays in global memory.
template <typename R>
struct GlobalSrc {
    const FineCtx<R> &f;
    const ShapeCtl &ctl;
    __device__ GlobalSrc(const FineCtx<R> &f_) : f(f_), ctl(f_.ctl) {}
    __device__ inline void range(double u0, double u1, int &j0, int &j1) const { cand(f, u0, u1, j0, j1); }
    __device__ inline int cls(int j, double &u) const { return classify(f, j, u); }
    __device__ inline const LineShape<R> &shape(int j) const { return f.shp[j]; }
    __device__ inline const LineShape<R> &line(int j) const { return f.shp[j]; }
    // classify()'s class 2 from the sorted entry's own record
    __device__ inline bool centre(const LineShape<R> &p, double &u) const {
        u = p.nus - f.org;
        if (p.idx < f.wlo || p.idx >= f.whi) return false;
        if (p.nus < f.org || p.nus > f.top) return false;
        return u < 10.0;
    }
    // classify()'s classes 1 and 3 (wing lines closer than 1 cm-1 to the record's left /
    // right edge), from the sorted entry's own record
    __device__ inline bool wing_left(const LineShape<R> &p, double u) const {
        if (p.idx < f.wlo || p.idx >= f.whi || !(p.nus < f.org)) return false;
        const double d = -u;
        return !(d > f.cut) && d < 1.0;
    }
    __device__ inline bool wing_right(const LineShape<R> &p, double u) const {
        if (p.idx < f.wlo || p.idx >= f.whi || !(p.nus > f.top)) return false;
        const double dr = u - 10.0;
        return !(dr >= f.cut) && dr < 1.0;
    }
    // candidates of a cell pair's group (u in [wlo, whi) on the left walk, v = 10 - u in
    // [wlo, whi) on the right walk) and of a grid-9 cell's fine points (u in [ulo, uhi)):
    // bucket supersets, the gathers test exactly
    template <bool RIGHT>
    __device__ inline void group_range(double wlo, double whi, int &j0, int &j1) const {
        if (!RIGHT) cand(f, wlo, whi, j0, j1);
        else cand(f, 10.0 - whi - 1e-9, 10.0 - wlo + 1e-9, j0, j1);
    }
    __device__ inline void fine_range(double ulo, double uhi, int &j0, int &j1) const {
        cand(f, ulo - 1e-9, uhi, j0, j1);
    }
    // lines of the one-line-per-lane passes, u in [ulo, uhi): a bucket superset (the passes
    // test centre membership and the deposit geometry per line)
    __device__ inline void lane_range(double ulo, double uhi, int &j0, int &j1) const { cand(f, ulo, uhi, j0, j1); }
};

// Centre lines (class 2) of one tile's fine reach, staged in LDS in sorted order: u = nu* - org
// and the sorted index, plus a local bucket table s_b[b] = first staged line with
// u >= ustart + b/512.  The cell pairs of grid k >= 4 in a tile [t, t+1) cm-1 take centre
// lines with u in (t - H_{k-1}, t + 1 + H_{k-1}) (H_3 = 0.125) and the fine points lines
// within 2 H_9 + H of the point, so the staged reach [t - STAGE_REACH, t + 1 + STAGE_REACH)
// holds every line they need.
constexpr double STAGE_REACH = 0.1328125;  // H_3 + 4/512, exact in binary
constexpr int STAGE_NB = 648;              // (1 + 2 STAGE_REACH) * 512 buckets
constexpr int STAGE_CAP = 1024;            // default staged lines; denser tiles read global memory

template <typename R>
struct CentreLds {
    const double *su;
    const int32_t *sj;
    const int32_t *sb;
    double ustart, org;
    const LineShape<R> *shp;
    ShapeCtl ctl;
    // every staged line is a centre line
    __device__ inline const LineShape<R> &line(int j) const { return shp[sj[j]]; }
    __device__ inline bool centre(const LineShape<R> &p, double &u) const {
        u = p.nus - org;  // the value staged in su
        return true;
    }
    // wing lines are not staged (the record-edge pairs read global memory)
    __device__ inline bool wing_left(const LineShape<R> &, double) const { return false; }
    __device__ inline bool wing_right(const LineShape<R> &, double) const { return false; }
    // staged bucket of a position on the 1/512 cm-1 grid (exact: ustart is on that grid)
    __device__ inline int bk(double x) const { return (int)((x - ustart) * BUCKETS_PER_CM); }
    // Group boundaries of grids 4..9 (multiples of H_{k-1} from the origin) and the fine
    // cells' boundaries are on the staged bucket grid, so the bucket table gives the ranges
    // without a search.  Left walk: exactly u in [wlo, whi).  Right walk: u >= 10 - whi is a
    // superset at the start (the gather tests v); the end steps past the lines whose
    // v = fl(10 - u) still rounds to >= wlo.
    template <bool RIGHT>
    __device__ inline void group_range(double wlo, double whi, int &j0, int &j1) const {
        if (!RIGHT) {
            j0 = sb[bk(wlo)];
            j1 = sb[bk(whi)];
        } else {
            j0 = sb[bk(10.0 - whi)];
            j1 = sb[bk(10.0 - wlo)];
            const int n = sb[STAGE_NB];
            while (j1 < n && !(10.0 - su[j1] < wlo)) j1++;
        }
    }
    __device__ inline void fine_range(double ulo, double uhi, int &j0, int &j1) const {
        j0 = sb[bk(ulo)];
        j1 = sb[bk(uhi)];
    }
    // u in [ulo, uhi) for bounds on the staged bucket grid
    __device__ inline void lane_range(double ulo, double uhi, int &j0, int &j1) const {
        j0 = sb[bk(ulo)];
        j1 = sb[bk(uhi)];
    }
};

// Centre lines on grid k, by cell pairs.  The left walk of a centre line with
// n = floor(u 2^k) >= 2 covers cells 2g-1 .. n-1 of grid k, g = n >> 1 (positions
// q = 4g-4 .. 2n-2 in half cells): the lines of one group g -- u in [g H_{k-1}, (g+1) H_{k-1})
// -- deposit into the cell pair (2g-1, 2g) and no other centre line's left walk does.  The
// right walk is the mirror image in v = 10 - u: group g' = NT_k/2 + 1 - g deposits into the
// same pair.  So the six sums (P, mid, L of both cells) of a pair have one owner, which
// visits each line of its two groups once and evaluates 3 positions (n even) or 5 (n odd).
// The lines of a group are split over the owner's L threads (sub-lane sl), each in sorted
// order.  Slots: 0..2 = P, mid, L of cell 2g-1; 3..5 = the same of cell 2g.
// The record's edge pairs (cells 1, 2 and NT_k-1, NT_k) also take the wing lines closer
// than H_{k-1} to that edge (leftLBL / rightLBL on grids k <= m, src/LineGridCalc.f90): they
// lie in the u range of the edge pair's mirror group, and a wing line's positions follow
// the same pattern as a centre line's -- nodes q = 4-e from the edge, e = 0..2, plus
// e = 3, 4 and the second cell's edge when d >= H_k (m == k) -- so they are summed in the
// same loop.  X: left wing u - q hq, right wing (u - 10) + q' hq, as the reference forms them.
// rng: the group's candidate range [rng[0], rng[1]) looked up beforehand (group_range of
// the same group), or null
template <typename R, bool RIGHT, typename Src, bool WINGS_ONLY = false>
__device__ inline void pair_walk(const Src &src, int k, int g, int sl, int L, double s[6],
                                 const int *rng = nullptr) {
    const double Hk = hk(k), hq = 0.5 * Hk;
    const double wlo = (double)(2 * g) * Hk, whi = (double)(2 * g + 2) * Hk;
    const double q0 = (double)(4 * g - 4);
    const bool edge = g == (5 << k);  // walk group starting at the record's far edge
    int j0, j1;
    if (rng) {
        j0 = rng[0];
        j1 = rng[1];
    } else {
        src.template group_range<RIGHT>(wlo, whi, j0, j1);
    }
    // sums by position e = q - q0 (a2o: e = 2 of lines with n odd, the second cell's edge)
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a2o = 0.0, a3 = 0.0, a4 = 0.0;
    int j = j0 + sl;
    if (j < j1) {
        // the next line's record is loaded while the current one is evaluated
        LineShape<R> nxt = src.line(j);
        for (; j < j1; j += L) {
            const LineShape<R> p = nxt;
            if (j + L < j1) nxt = src.line(j + L);
            double u;
            // X_e = B - (C + e) hq (centre lines, wing on the right edge) or B + (C + e) hq
            double B, C;
            bool five, plus = false;
            if (src.centre(p, u)) {
                if (WINGS_ONLY) continue;
                const double w = RIGHT ? 10.0 - u : u;  // distance walked from the pair's side
                if (w < wlo || !(w < whi)) continue;
                five = (int)floor(__builtin_ldexp(w, k)) & 1;
                B = w;
                C = q0;
            } else {
                if (!edge) continue;
                if (RIGHT) {  // left wing, u in (-H_{k-1}, 0): X = u - (4 - e) hq
                    if (!src.wing_left(p, u) || !(-u < 2.0 * Hk)) continue;
                    five = -u >= Hk;
                    B = u;
                    plus = true;
                } else {      // right wing: X = (u - 10) + (4 - e) hq
                    if (!src.wing_right(p, u) || !(u - 10.0 < 2.0 * Hk)) continue;
                    five = u - 10.0 >= Hk;
                    B = u - 10.0;
                }
                C = -4.0;
            }
            const int ne = five ? 5 : 3;
#pragma unroll 1
            for (int e = 0; e < ne; e++) {
                const double m = (C + (double)e) * hq;
                const double v = (double)shape_eval<R, true>(plus ? B + m : B - m, p, src.ctl);
                if (e == 0) a0 += v;
                else if (e == 1) a1 += v;
                else if (e == 2) {
                    a2 += v;
                    if (five) a2o += v;
                } else if (e == 3) a3 += v;
                else a4 += v;
            }
        }
    }
    // left walk: e -> P1, M1, L1 (+P2 for n odd), M2, L2; right walk: the mirror image
    if (!RIGHT) {
        s[0] += a0; s[1] += a1; s[2] += a2; s[3] += a2o; s[4] += a3; s[5] += a4;
    } else {
        s[5] += a0; s[4] += a1; s[3] += a2; s[2] += a2o; s[1] += a3; s[0] += a4;
    }
}

// direct deposits of centre lines at the four fine points i0 .. i0+3 of one grid-9 cell
// (i0 = 4m + 2, positions (i-1) H).  Point i takes the line at or right of it when
// u < (floor((i-2)/4) + 2) H9 and the line left of it when v = 10 - u <
// (floor((NT-i)/4) + 2) H9 (src/LineGridCalc.f90:332-339, :485-490); both bounds are the
// same for the four points of a cell, so they share one candidate range, u in
// [(m-1) H9, (m+2) H9), and each line is loaded once.  v[e]: sums in sorted order.
template <typename R, typename Src>
__device__ inline void gather_fine_cell(const Src &src, int i0, double v[4]) {
    const double uhi = (double)(((i0 - 2) >> 2) + 2) * H9;
    const double vhi = (double)(((NT - i0) >> 2) + 2) * H9;
    int j0, j1;
    src.fine_range(10.0 - vhi, uhi, j0, j1);
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (j0 < j1) {
        LineShape<R> nxt = src.line(j0);
        for (int j = j0; j < j1; j++) {
            const LineShape<R> p = nxt;
            if (j + 1 < j1) nxt = src.line(j + 1);
            double u;
            if (!src.centre(p, u)) continue;
#pragma unroll 1
            for (int e = 0; e < 4; e++) {
                const double xi = (double)(i0 + e - 1) * HF;
                if (u < xi ? !(10.0 - u < vhi) : !(u < uhi)) continue;
                const double x = (double)shape_eval<R, true>(u - xi, p, src.ctl);
                if (e == 0) v0 += x;
                else if (e == 1) v1 += x;
                else if (e == 2) v2 += x;
                else v3 += x;
            }
        }
    }
    v[0] = v0;
    v[1] = v1;
    v[2] = v2;
    v[3] = v3;
}

// Fine points of a staged tile in the FP32 form, one line per lane.  The four points of
// grid-9 cell m (record-level index; points 4m+2 .. 4m+5) take the lines with u in
// [(m-1) H9, (m+2) H9), i.e. a line with mu = floor(u / H9) serves cells mu-1, mu, mu+1,
// under the same per-point tests as gather_fine_cell.  Wavefront wv owns the tile's cells
// [128 wv, 128 wv + 128) -- lane l the cells 128 wv + l and 128 wv + 64 + l -- and walks
// the staged lines of their windows 64 at a time: each lane evaluates its line's 12
// points, a key-segmented lane scan (key = mu, non-decreasing along the sorted lines) sums
// them per source cell, and the segment's last lane adds its three cells' sums to the
// wavefront's chunk partials in LDS (one cell offset at a time, so no two lanes update one
// slot in the same instruction); after each chunk the owning lanes add the partials to
// their FP64 sums v[h][e] and clear them.  buf: 4 x 512 floats of free LDS.  Src: the
// tile's LDS stage, or the level's sorted records for a tile too dense to stage.
template <typename R, typename Src>
__device__ inline void fine_lanes(const Src &src, int t, float *buf, double v[2][4]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int M0 = 512 * t + 128 * wv;  // the wavefront's first cell (record level)
    float *b = buf + 512 * wv;          // [128 cells][4 points]
    for (int e = lane; e < 512; e += 64) b[e] = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int e = 0; e < 4; e++) v[h][e] = 0.0;
    int j0, j1;
    src.lane_range((double)(M0 - 1) * H9, (double)(M0 + 130) * H9, j0, j1);
#pragma unroll 1
    for (int c0 = j0; c0 < j1; c0 += 64) {
        const int j = c0 + lane;
        int key = -1;  // lanes without a centre line: see scan_key
        R a[12];
#pragma unroll
        for (int q = 0; q < 12; q++) a[q] = R(0);
        LineShape<R> p;
        double u = 0.0;
        if (j < j1) p = src.line(j);
        const bool has = j < j1 && src.centre(p, u);
        if (has) {
            const int mu = (int)floor(u * 512.0);
            key = mu - M0 + 1;  // 0 .. 130
#pragma unroll
            for (int q = 0; q < 12; q++) {
                const int m = mu - 1 + (q >> 2);
                if (m < M0 || m >= M0 + 128) continue;
                const int i0 = 2 + 4 * m, e = q & 3;
                const double uhi = (double)(m + 2) * H9;
                const double vhi = (double)(((NT - i0) >> 2) + 2) * H9;
                const double xi = (double)(i0 + e - 1) * HF;
                if (u < xi ? !(10.0 - u < vhi) : !(u < uhi)) continue;
                a[q] = shape_eval<R, true>(u - xi, p, src.ctl);
            }
        }
        key = scan_key(has, key);
        seg_scan(key, a);
        const bool last = key >= 0 && next_key(key) != key;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int cl = key - 2 + d;  // local cell of offset d (mu - 1 + d - M0)
            if (last && cl >= 0 && cl < 128) {
#pragma unroll
                for (int e = 0; e < 4; e++) b[4 * cl + e] += (float)a[4 * d + e];
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int cl = lane + 64 * h;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                v[h][e] += (double)b[4 * cl + e];
                b[4 * cl + e] = 0.0f;
            }
        }
    }
}

// direct deposits of wing lines at the first / last fine points of the record
template <typename R, typename Src>
__device__ inline void gather_fine_wings(const Src &src, int i, Acc &acc) {
    if (i <= 5) {  // leftLBL :107-112 (d < H9)
        int j0, j1;
        src.range(-H9, 0.0, j0, j1);
        for (int j = j0; j < j1; j++) {
            double u;
            if (src.cls(j, u) != 1 || !(-u < H9)) continue;
            const double X = i == 2 ? u - HF : i == 3 ? (u - HF) - HF : i == 4 ? (u + HF) - H9 : u - H9;
            acc.add(shape_eval<R, true>(X, src.shape(j), src.ctl));
        }
    }
    if (i >= NT - 3) {  // rightLBL :592-595 (d < H9)
        int j0, j1;
        src.range(10.0 - 1e-9, 10.0 + H9 + 1e-9, j0, j1);
        for (int j = j0; j < j1; j++) {
            double u;
            if (src.cls(j, u) != 3) continue;
            const double dr = u - 10.0;
            if (!(dr < H9)) continue;
            const double X = i == NT ? dr : i == NT - 1 ? dr + HF : i == NT - 2 ? (dr + HF) + HF : (dr + H9) - HF;
            acc.add(shape_eval<R, true>(X, src.shape(j), src.ctl));
        }
    }
}

// offset of grid k's P block in the tile's node sums; mid at +2^k, L at +2*2^k
__device__ inline int lvl_acc_off(int k) { return 3 * ((1 << k) - 2); }

// ---- stages of one (record, tile) work item, shared by the forward and adjoint kernels

template <typename R>
__device__ inline void tile_ctx(const KArgs &a, int lev, int s, FineCtx<R> &f) {
    f.skey = a.skey + (int64_t)lev * a.nl;
    f.sidx = a.sidx + (int64_t)lev * a.nl;
    f.shp = reinterpret_cast<const LineShape<R> *>(a.shp) + (int64_t)lev * a.nl;
    f.bst = a.bst + (int64_t)lev * (a.NB + 1);
    f.org = a.org[s];
    f.top = f.org + 10.0;
    f.base = a.bbase;
    f.cut = (double)a.cutoff;
    f.NB = a.NB;
    f.wlo = a.lo[s];
    f.whi = a.hi[s];
    f.ctl = make_ctl(a, a.lev0 + lev);
}

// Stage the centre lines of tile t's fine reach: order-preserving compaction of the sorted
// candidates (one ballot per wave, wave counts through LDS), then the local bucket table.
// Returns false (uniformly) when the reach holds more than cap centre lines.
template <typename R>
__device__ inline bool stage_tile(const FineCtx<R> &f, int t, double *s_u, int32_t *s_j, int32_t *s_b,
                                  int *s_cnt, CentreLds<R> &lsrc, int cap = STAGE_CAP) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double ustart = (double)t - STAGE_REACH, uend = (double)t + 1.0 + STAGE_REACH;
    int nst = 0;
    int J0, J1;
    cand(f, ustart, uend, J0, J1);
    for (int base = J0; base < J1; base += BLOCK) {
        const int j = base + tid;
        double u = 0.0;
        const bool keep = j < J1 && classify(f, j, u) == 2 && u >= ustart && u < uend;
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_cnt[wv] = __popcll(m);
        __syncthreads();
        int off = nst, tot = 0;
        for (int w2 = 0; w2 < BLOCK / 64; w2++) {
            if (w2 < wv) off += s_cnt[w2];
            tot += s_cnt[w2];
        }
        off += __popcll(m & ((1ull << lane) - 1ull));
        if (keep && off < cap) {
            s_u[off] = u;
            s_j[off] = j;
        }
        nst += tot;
        __syncthreads();
    }
    const bool staged = nst <= cap;
    if (staged)
        for (int b = tid; b <= STAGE_NB; b += BLOCK) {
            const double x = ustart + (double)b * (1.0 / BUCKETS_PER_CM);
            int lo = 0, hi = nst;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_u[mid] < x) lo = mid + 1;
                else hi = mid;
            }
            s_b[b] = lo;
        }
    __syncthreads();
    lsrc.su = s_u;
    lsrc.sj = s_j;
    lsrc.sb = s_b;
    lsrc.ustart = ustart;
    lsrc.org = f.org;
    lsrc.shp = f.shp;
    lsrc.ctl = f.ctl;
    return staged;
}

// Grids LANES_K0..9 of a staged tile, one line per lane.  The pairs of these grids hold a
// few lines each (2 H_{k-1} cm-1 of lines), so a pair-per-thread split leaves most lanes
// idle while the fullest pair's owner works.  Instead each wavefront owns a quarter of every
// grid's pairs and walks their lines -- per walk a contiguous range of the sorted stage --
// 64 at a time, one line per lane.  A wavefront's pairs of grid k take, on the left walk,
// u in [t + wv/4, t + (wv+1)/4) + H_{k-1} (on the right walk, the mirror image), so one
// pass over the union of these ranges serves all three grids: each lane loads its line
// once and, grid after grid, evaluates its 3 or 5 positions when the line's pair (the same
// floor as the deposit geometry) is one of the wavefront's; a key-segmented inclusive scan
// across the lanes (keys = pair index, non-decreasing along the sorted lines in both walks;
// lanes with no pair on that grid inherit the key before them, scan_key) sums every pair's values, and the last
// lane of each segment adds them to the pair's slots in LDS.  Pairs belong to one
// wavefront and its chunks run in order, so every sum has a fixed order.  Centre lines
// only (wing lines are added by the edge pair's walk); the wavefront's slots are zeroed
// first.
constexpr int LANES_K0 = 6;

template <typename R, typename Src>
__device__ inline void level_lanes(const Src &src, int t, double *g) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll 1
    for (int k = LANES_K0; k <= 9; k++) {
        const int npair = 1 << (k - 1), nk = 1 << k, accb = lvl_acc_off(k);
        const int pw = npair / (BLOCK / 64), p0 = wv * pw;  // this wavefront's pairs [p0, p0 + pw)
        for (int e = lane; e < 6 * pw; e += 64) {
            const int pi = p0 + e / 6, c = e % 6;
            g[accb + (c % 3) * nk + 2 * pi + c / 3] = 0.0;
        }
    }
    const double lo = (double)t + 0.25 * wv, hi = lo + 0.25;
    // left walk: exactly u in [lo + H_8, hi + H_{K0-1}) (bounds on the bucket grid); right
    // walk: u in (lo - H_{K0-1}, hi - H_8], one bucket wider (w = fl(10 - u)).  The two
    // ranges are walked as one sequence of lanes (left walk first), so that a chunk's lanes
    // are filled across the switch; keys side * npair + pair stay non-decreasing.
    int ja, ja1, jb, jb1;
    src.lane_range(lo + hk(8), hi + hk(LANES_K0 - 1), ja, ja1);
    src.lane_range(lo - hk(LANES_K0 - 1) - 1.0 / BUCKETS_PER_CM, hi - hk(8) + 1.0 / BUCKETS_PER_CM, jb, jb1);
    const int na = ja1 - ja, ntot = na + (jb1 - jb);
#pragma unroll 1
    for (int c0 = 0; c0 < ntot; c0 += 64) {
        const int m = c0 + lane;
        bool act = m < ntot;
        const bool side = m >= na;  // this lane's walk: false left, true right
        LineShape<R> p;
        double w = 0.0;
        if (act) {
            p = src.line(side ? jb + (m - na) : ja + m);
            double u;
            act = src.centre(p, u);  // staged lines always; sorted records: centre lines only
            w = side ? 10.0 - u : u;
        }
        // the last position of a lane's walk on grid k, w - (n_k - 1) H_k, is the first of its
        // walk on grid k + 1 (n_{k+1} >> 1 = n_k), formed by the same FP64 operations from the
        // same w: its value is carried instead of evaluated again
        R carry = R(0);
        bool has_carry = false;
#pragma unroll 1
        for (int k = LANES_K0; k <= 9; k++) {
            const double hq = 0.5 * hk(k);
            const int npair = 1 << (k - 1), nk = 1 << k, accb = lvl_acc_off(k);
            const int pw = npair / (BLOCK / 64), p0 = wv * pw;
            int key = -1;  // lanes without a pair on this grid: see scan_key
            R a0 = R(0), a1 = R(0), a2 = R(0), a2o = R(0), a3 = R(0), a4 = R(0);
            if (act) {
                const int nn = (int)floor(__builtin_ldexp(w, k));
                const int gi = nn >> 1;
                const int pi = side ? (5 << k) - gi - t * npair : gi - t * npair - 1;
                if (pi >= p0 && pi < p0 + pw) {
                    key = (side ? npair : 0) + pi;
                    const bool five = nn & 1;
                    const double q0 = (double)(4 * gi - 4);
                    if (has_carry) a0 = carry;
                    else a0 = shape_eval<R, true>(w - q0 * hq, p, src.ctl);
                    a1 = shape_eval<R, true>(w - (q0 + 1.0) * hq, p, src.ctl);
                    a2 = shape_eval<R, true>(w - (q0 + 2.0) * hq, p, src.ctl);
                    if (five) {
                        a2o = a2;
                        a3 = shape_eval<R, true>(w - (q0 + 3.0) * hq, p, src.ctl);
                        a4 = shape_eval<R, true>(w - (q0 + 4.0) * hq, p, src.ctl);
                    }
                    carry = five ? a4 : a2;
                    has_carry = true;
                } else {
                    has_carry = false;
                }
            }
            // slots P1 M1 L1 P2 M2 L2 of the lane's pair (right walk: the mirror image)
            R sv[6] = {side ? a4 : a0, side ? a3 : a1, side ? a2o : a2, side ? a2 : a2o, side ? a1 : a3, side ? a0 : a4};
            key = scan_key(key >= 0, key);
            seg_scan(key, sv);
            const R v0 = sv[0], v1 = sv[1], v2 = sv[2], v3 = sv[3], v4 = sv[4], v5 = sv[5];
            const bool last = key >= 0 && next_key(key) != key;
            const bool kside = key >= npair;  // the walk of the segment (a lane may inherit it)
            // a pair may end a segment in both walks of one chunk: left walk first, then
            // right walk, so that no two lanes update one slot in the same instruction
#pragma unroll
            for (int sd = 0; sd < 2; sd++) {
                if (last && (int)kside == sd) {
                    const int cb = accb + 2 * (key - (kside ? npair : 0));
                    g[cb] += (double)v0;
                    g[cb + nk] += (double)v1;
                    g[cb + 2 * nk] += (double)v2;
                    g[cb + 1] += (double)v3;
                    g[cb + nk + 1] += (double)v4;
                    g[cb + 2 * nk + 1] += (double)v5;
                }
            }
        }
    }
}

// Pre-cascade sums of grids 1..9 at the nodes of tile t, into g (LDS).  Grids 1..5 come
// from the level-0 kernel (cz15: the record's P, mid, L blocks of grids 1..5), which sums
// them from the record's centre and edge lines in one pass.  Grids 6..9 take one staged
// line per lane (level_lanes: their centre lines lie within H_3 of the tile), the edge
// pairs' wing lines a walk over global memory.  Ends with a barrier.
template <typename R>
__device__ inline void tile_level_sums(const GlobalSrc<R> &gsrc, const CentreLds<R> &lsrc, bool staged, int t,
                                       double *g, double *red, const double *cz15) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // grids 1..5 of the tile: cells [t 2^k, (t + 1) 2^k) of grid k, 3 x (2 + 4 + ... + 32)
    // values: thread e < 186 copies value e
    for (int e = tid; e < 186; e += BLOCK) {
        int k = 1, o = e;
        while (o >= 3 << k) {
            o -= 3 << k;
            k++;
        }
        const int nk = 1 << k, role = o / nk, c = o % nk;
        g[lvl_acc_off(k) + role * nk + c] = cz15[30 * ((1 << k) - 2) + role * (10 << k) + t * nk + c];
    }
#pragma unroll 1
    for (int k = LANES_K0; k <= 9; k++) {
        if (k >= LANES_K0) {
            if (k == LANES_K0) {
                if (staged) level_lanes<R>(lsrc, t, g);
                else level_lanes<R>(gsrc, t, g);
            }
            if ((t == 0 || t == TILES - 1) && wv == (t == 0 ? 0 : BLOCK / 64 - 1)) {
                // wing lines of the record's edge pair (not staged), by the wavefront that
                // owns the pair, after its centre lines
                double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                if (t == 0) pair_walk<R, true, GlobalSrc<R>, true>(gsrc, k, 5 << k, lane, 64, s);
                else pair_walk<R, false, GlobalSrc<R>, true>(gsrc, k, 5 << k, lane, 64, s);
                for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                    for (int c = 0; c < 6; c++) s[c] += __shfl_xor(s[c], off);
                if (lane == 0) {
                    const int nk = 1 << k, accb = lvl_acc_off(k), c1 = t == 0 ? 0 : nk - 2;
#pragma unroll
                    for (int c = 0; c < 2; c++) {
                        g[accb + c1 + c] += s[3 * c];
                        g[accb + nk + c1 + c] += s[3 * c + 1];
                        g[accb + 2 * nk + c1 + c] += s[3 * c + 2];
                    }
                }
            }
        }
    }
    __syncthreads();
}

// limiter of the cascade, src/Grids.f90:80-101: A1 is the new mid of the left child cell,
// A2 of the right one; true when the limiter replaces the Lagrange value by the mean
__device__ inline double cascade_a1(double P, double M, double L, bool &lim) {
    double A = 0.375 * P + 0.75 * M - 0.125 * L;
    lim = A < P || A < M;
    return lim ? 0.5 * (P + M) : A;
}
__device__ inline double cascade_a2(double P, double M, double L, bool &lim) {
    double A = 0.375 * L + 0.75 * M - 0.125 * P;
    lim = A < L || A < M;
    return lim ? 0.5 * (L + M) : A;
}

// Limiter choices of a tile's cascade, kept for the adjoint's transposed cascade: for the
// step from grid k (k = 1..9) to the next, the A1 and A2 flags of its cells in 64-cell
// groups, one 64-bit word per (group, flag): word lim_word(k, c) + 0 / 1, bit c & 63.
// Grids 1..6 take one group each, 7 two, 8 four, 9 eight: LIMW = 40 words per tile.  The
// words are formed by wavefront ballots (a group is one wavefront's cells) and stored by
// the wavefront's first lane.
constexpr int LIMW = 40;
__device__ inline int lim_word(int k, int c) {
    const int base = k <= 6 ? 2 * (k - 1) : k == 7 ? 12 : k == 8 ? 16 : 24;
    return base + 2 * (c >> 6);
}
// the wavefront's flags of cells c (one per lane, lanes without a cell pass false); the
// cells of the active lanes share one group, the first active lane being lane 0
__device__ inline void lim_put(uint64_t *lim, int k, int c, bool l1, bool l2) {
    const uint64_t w1 = __ballot(l1), w2 = __ballot(l2);
    if ((threadIdx.x & 63) == 0) {
        const int w = lim_word(k, c);
        lim[w] = w1;
        lim[w + 1] = w2;
    }
}
__device__ inline void lim_get(const uint64_t *lim, int k, int c, bool &l1, bool &l2) {
    const int w = lim_word(k, c), b = c & 63;
    l1 = (lim[w] >> b) & 1u;
    l2 = (lim[w + 1] >> b) & 1u;
}

// Cascade of tile t in FP64 from its grid-0 cell (cz: P, M, L of the 10 cells) down to
// grid 9 (src/Grids.f90:104-139): afterwards g holds the cascaded values of every level.
// lim (the tile's LIMW words in global memory, or null): the limiter choices of grids 1..8
// are stored there.
__device__ inline void tile_cascade(double *g, const double *cz, int t, uint64_t *lim = nullptr) {
    const int tid = threadIdx.x;
    auto step = [&](double P, double M, double L, double *fP, double *fM, double *fL, int i, bool &l1, bool &l2) {
        fP[i] += P;
        fM[i] += cascade_a1(P, M, L, l1);
        fL[i] += M;
        fP[i + 1] += M;
        fM[i + 1] += cascade_a2(P, M, L, l2);
        fL[i + 1] += L;
    };
    if (tid == 0) {
        const int k1 = lvl_acc_off(1);
        bool l1, l2;
        step(cz[1 + t], cz[11 + t], cz[21 + t], g + k1, g + k1 + 2, g + k1 + 4, 0, l1, l2);
    }
    __syncthreads();
    for (int k = 1; k < 9; k++) {
        const int nk = 1 << k, cb = lvl_acc_off(k), fb = lvl_acc_off(k + 1);
        for (int c = tid; c < nk; c += BLOCK) {  // nk <= BLOCK: one cell per thread
            bool l1, l2;
            step(g[cb + c], g[cb + nk + c], g[cb + 2 * nk + c], g + fb, g + fb + 2 * nk, g + fb + 4 * nk, 2 * c, l1, l2);
            if (lim) lim_put(lim, k, c, l1, l2);
        }
        __syncthreads();
    }
}

// One (record, tile) work item w by owner gathers: grid 1..9 sums into g, cascade,
// fine points in registers, record write.  LDS: g[NACC], staging arrays of capacity cap.
template <typename R>
__device__ __forceinline__ void fine_tile(const KArgs &a, int w, double *g, double *red, double *s_u, int32_t *s_j, int32_t *s_b,
                                 int *s_cnt, int cap) {
    const int r = w / TILES, t = w % TILES;
    const int lev = r / a.nrec, s = r % a.nrec;
    const double sc = a.sscale[a.lev0 + lev];
    const int tid = threadIdx.x;
    // the cascade's limiter choices, when the caller keeps them for the adjoint
    uint64_t *lim = a.lim ? a.lim + ((int64_t)r * TILES + t) * LIMW : nullptr;
    FineCtx<R> f;
    tile_ctx<R>(a, lev, s, f);
    const GlobalSrc<R> gsrc(f);
    CentreLds<R> lsrc;
    const bool staged = stage_tile<R>(f, t, s_u, s_j, s_b, s_cnt, lsrc, cap);
    const double *cz = a.coarse + (int64_t)r * NCOARSE;
    tile_level_sums<R>(gsrc, lsrc, staged, t, g, red, cz + NG0);
    if (a.grids && r == 0) {
        for (int e = tid; e < NACC; e += BLOCK) {
            int k = 1;
            while (k < 9 && e >= lvl_acc_off(k + 1)) k++;
            const int nk = 1 << k, o = e - lvl_acc_off(k), role = o / nk, cl = o % nk;
            a.grids[30 * ((1 << k) - 1) + role * (10 << k) + t * nk + cl] = g[e] / sc;
        }
        if (tid < 3) a.grids[tid * 10 + t] = cz[1 + 10 * tid + t] / sc;
        if (t == 0 && tid == 0) a.grids[30690] = cz[0] / sc;
    }
    tile_cascade(g, cz, t, lim);

    // ---- fine points: direct deposits summed in registers (thread owns the 4 points of
    // grid-9 cells tid and tid + 256), then the last cascade step and the record write
    const int cb9 = lvl_acc_off(9);
    const double mul = (double)a.outmul[a.lev0 + lev] / sc;
    const int64_t ro = ((int64_t)(a.lev0 + lev) * a.nrec + s) * NT;
    // FP32 form: lines per lane (grids 1..8's LDS is free after the cascade); FP64 form:
    // each thread gathers the candidates of its cells
    const bool lanes = sizeof(R) == 4;
    double fv[2][4];
    if (lanes) {
        if (staged) fine_lanes<R>(lsrc, t, (float *)g, fv);
        else fine_lanes<R>(gsrc, t, (float *)g, fv);
    }
    for (int h = 0; h < 2; h++) {
        const int c = lanes ? 128 * (tid >> 6) + 64 * h + (tid & 63) : tid + h * BLOCK;
        const int i0 = 2048 * t + 2 + 4 * c;
        double v[4];
        if (lanes) {
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = fv[h][e];
        } else if (staged) gather_fine_cell<R>(lsrc, i0, v);
        else gather_fine_cell<R>(gsrc, i0, v);
        if (i0 <= 5 || i0 + 3 >= NT - 3) {
#pragma unroll 1
            for (int e = 0; e < 4; e++) {
                const int i = i0 + e;
                if (i > 5 && i < NT - 3) continue;
                Acc acc;
                acc.zero();
                gather_fine_wings<R>(gsrc, i, acc);
                const double x = acc.value();
                v[0] += e == 0 ? x : 0.0;
                v[1] += e == 1 ? x : 0.0;
                v[2] += e == 2 ? x : 0.0;
                v[3] += e == 3 ? x : 0.0;
            }
        }
        if (a.grids && r == 0)
            for (int e = 0; e < 4; e++) a.grids[30690 + 2048 * t + 1 + 4 * c + e] = v[e] / sc;
        const double P = g[cb9 + c], M = g[cb9 + 512 + c], L = g[cb9 + 1024 + c];
        bool l1, l2;
        v[0] += cascade_a1(P, M, L, l1);
        v[1] += M;
        v[2] += cascade_a2(P, M, L, l2);
        v[3] += L;
        if (lim) lim_put(lim, 9, c, l1, l2);  // c: lane + 64 x (the wavefront's group)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (a.out64) a.out64[ro + 2048 * t + 1 + 4 * c + e] = v[e] * mul;
            else a.out[ro + 2048 * t + 1 + 4 * c + e] = (float)(v[e] * mul);
        }
    }
    if (t == 0 && tid == 0) {
        if (a.out64) a.out64[ro] = cz[0] * mul;
        else a.out[ro] = (float)(cz[0] * mul);
    }
}

// CAP: centre lines staged in LDS per tile (LDS 27.5 KB + 12 B x CAP: occupancy against
// staging); XCD: contiguous record ranges per XCD (xcd_item) or plain dispatch order.
// The FP32 form is held to 128 VGPRs (4 waves per SIMD, as many as the LDS of the default
// stage allows): the gathers wait on memory, and waves hide that better than registers.
template <typename R, int CAP = STAGE_CAP, bool XCD = true>
__global__ __launch_bounds__(BLOCK, sizeof(R) == 4 ? 4 : 1) void fine_kernel(KArgs a) {
    __shared__ double g[NACC];            // grids 1..9: P, mid, L of the tile's cells
    __shared__ double s_u[CAP];           // staged centre lines
    __shared__ int32_t s_j[CAP];
    __shared__ int32_t s_b[STAGE_NB + 1];
    __shared__ int s_cnt[BLOCK / 64];
    __shared__ double red[BLOCK / 64 * 6];  // wave partials of the level sums
    const int nitems = a.nlev * a.nrec * TILES;
    const int w = XCD ? xcd_item(nitems) : (int)blockIdx.x;
    if (w >= nitems) return;
    fine_tile<R>(a, w, g, red, s_u, s_j, s_b, s_cnt, CAP);
}

// ---- adjoint of the deposit + cascade of one (record, tile)
// gout: dL/d(record point) in the records' layout; gbar: per record of the batch,
// dL/d(pre-cascade node) in the reference's grid layout (grids 0..9 P|mid|L blocks, then
// RK(1..20481)), for samples scaled by the level's power-of-two scale.  The tile recomputes
// its grid 1..9 sums and the cascade (the limiter's choices need the cascaded values), then
// runs the transposed cascade from its 2,048 points up to its grid-0 cell.  LIM: the forward
// kept the cascade's limiter choices (a.lim), which are all the transposed cascade needs of
// the forward values, so the grid sums and the forward cascade are not recomputed.
template <typename R, bool LIM = false>
__global__ __launch_bounds__(BLOCK) void adjoint_tile_kernel(KArgs a, const float *gout, float *gbar) {
    __shared__ double g[NACC];
    __shared__ float gb[NACC];            // adjoints of the cascaded values, grids 1..9
    __shared__ double s_u[STAGE_CAP];
    __shared__ int32_t s_j[STAGE_CAP];
    __shared__ int32_t s_b[STAGE_NB + 1];
    __shared__ int s_cnt[BLOCK / 64];
    __shared__ double red[BLOCK / 64 * 6];
    const int nitems = a.nlev * a.nrec * TILES;
    const int w = xcd_item(nitems);
    if (w >= nitems) return;
    const int r = w / TILES, t = w % TILES;
    const int lev = r / a.nrec, s = r % a.nrec;
    const double sc = a.sscale[a.lev0 + lev];
    const int tid = threadIdx.x;
    const double *cz = a.coarse + (int64_t)r * NCOARSE;
    __shared__ uint64_t s_lim[LIMW];
    if constexpr (LIM) {
        if (tid < LIMW) s_lim[tid] = a.lim[((int64_t)r * TILES + t) * LIMW + tid];
        __syncthreads();
    } else {
        FineCtx<R> f;
        tile_ctx<R>(a, lev, s, f);
        const GlobalSrc<R> gsrc(f);
        CentreLds<R> lsrc;
        const bool staged = stage_tile<R>(f, t, s_u, s_j, s_b, s_cnt, lsrc);
        tile_level_sums<R>(gsrc, lsrc, staged, t, g, red, cz + NG0);
        tile_cascade(g, cz, t);
    }
    // the limiter's choices of the step from grid k's cell c (P, M, L: its cascaded values)
    auto flags = [&](int k, int c, double P, double M, double L, bool &l1, bool &l2) {
        if constexpr (LIM) {
            lim_get(s_lim, k, c, l1, l2);
        } else {
            cascade_a1(P, M, L, l1);
            cascade_a2(P, M, L, l2);
        }
    };

    const double mul = (double)a.outmul[a.lev0 + lev] / sc;
    const float *go = gout + ((int64_t)(a.lev0 + lev) * a.nrec + s) * NT;
    float *gr = gbar + (int64_t)r * (30690 + NT);
    // fine points: the direct deposits' adjoint is the record's
    for (int e = tid; e < 2048; e += BLOCK) gr[30690 + 2048 * t + 1 + e] = (float)(go[2048 * t + 1 + e] * mul);
    if (t == 0 && tid == 0) gr[30690] = (float)(go[0] * mul);
    // grid 9 from the last cascade step: points 4c+2..4c+5 of the tile get A1, M, A2, L
    {
        const int cb9 = lvl_acc_off(9);
        for (int c = tid; c < 512; c += BLOCK) {
            bool l1, l2;
            flags(9, c, g[cb9 + c], g[cb9 + 512 + c], g[cb9 + 1024 + c], l1, l2);
            const int p0 = 2048 * t + 1 + 4 * c;
            const double r0 = go[p0] * mul, r1 = go[p0 + 1] * mul, r2 = go[p0 + 2] * mul, r3 = go[p0 + 3] * mul;
            gb[cb9 + c] = (float)(r0 * (l1 ? 0.5 : 0.375) + r2 * (l2 ? 0.0 : -0.125));
            gb[cb9 + 512 + c] = (float)(r0 * (l1 ? 0.5 : 0.75) + r1 + r2 * (l2 ? 0.5 : 0.75));
            gb[cb9 + 1024 + c] = (float)(r0 * (l1 ? 0.0 : -0.125) + r2 * (l2 ? 0.5 : 0.375) + r3);
        }
    }
    __syncthreads();
    // transposed steps k -> k+1 for k = 8..1, then grid 0
    auto back = [&](bool l1, bool l2, const float *cP, const float *cM, const float *cL, int i,
                    float &oP, float &oM, float &oL) {
        const double m0 = cM[i], m1 = cM[i + 1];
        oP = (float)(cP[i] + m0 * (l1 ? 0.5 : 0.375) + m1 * (l2 ? 0.0 : -0.125));
        oM = (float)(m0 * (l1 ? 0.5 : 0.75) + cL[i] + cP[i + 1] + m1 * (l2 ? 0.5 : 0.75));
        oL = (float)(m0 * (l1 ? 0.0 : -0.125) + m1 * (l2 ? 0.5 : 0.375) + cL[i + 1]);
    };
    for (int k = 8; k >= 1; k--) {
        const int nk = 1 << k, cb = lvl_acc_off(k), fb = lvl_acc_off(k + 1);
        for (int c = tid; c < nk; c += BLOCK) {
            bool l1, l2;
            flags(k, c, g[cb + c], g[cb + nk + c], g[cb + 2 * nk + c], l1, l2);
            back(l1, l2, gb + fb, gb + fb + 2 * nk, gb + fb + 4 * nk, 2 * c, gb[cb + c], gb[cb + nk + c],
                 gb[cb + 2 * nk + c]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int k1 = lvl_acc_off(1);
        float oP, oM, oL;
        bool l1, l2;  // the grid-0 cell's step, from the level-0 sums
        cascade_a1(cz[1 + t], cz[11 + t], cz[21 + t], l1);
        cascade_a2(cz[1 + t], cz[11 + t], cz[21 + t], l2);
        back(l1, l2, gb + k1, gb + k1 + 2, gb + k1 + 4, 0, oP, oM, oL);
        gr[t] = oP;
        gr[10 + t] = oM;
        gr[20 + t] = oL;
    }
    // grids 1..9 of the tile into the reference layout
    for (int e = tid; e < NACC; e += BLOCK) {
        int k = 1;
        while (k < 9 && e >= lvl_acc_off(k + 1)) k++;
        const int nk = 1 << k, o = e - lvl_acc_off(k), role = o / nk, cl = o % nk;
        gr[30 * ((1 << k) - 1) + role * (10 << k) + t * nk + cl] = gb[e];
    }
}
