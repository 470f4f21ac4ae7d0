// lbl_sort.h -- per-level ordering of the lines by shifted centre, and the bucket index.
// Included by marfa_lbl.hip inside its kernel namespace.
//
// A counting sort on the 1/512 cm-1 buckets the fine gathers use: histogram (integer
// atomics), exclusive scan (the bucket index bst itself), scatter, then each line's rank
// among its bucket's few lines by (shifted centre, line index) gives its final position,
// where its line record is written.  The result is the stable order by shifted centre
// whatever the scheduling, and every step is a plain kernel (no library call, no memset),
// so the whole pipeline can be captured in a hipGraph.
#pragma once

constexpr int SCAN_CHUNK = BLOCK * 16;  // buckets per scan block

__global__ __launch_bounds__(BLOCK) void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) p[i] = v;
}

// Consecutive lines of a level are nearly in order of shifted centre (the line list is
// sorted by centre; the pressure shift moves each by a fraction of a cm-1), so the
// BLOCK lines of one workgroup usually fall into a short run of buckets.  When that run is
// shorter than HIST_SPAN the workgroup counts in LDS and updates each non-empty bucket
// with one global atomic; otherwise every line updates its bucket (the integer counts are
// the same either way).
constexpr int HIST_SPAN = 2048;

// bucket of line j of level lev (act: j < nl) and the workgroup's [bmin, bmax] over its
// active lines (bmin > bmax when none is active)
__device__ inline int block_buckets(const uint64_t *keys, int64_t nl, double base, int NB, int lev, int64_t j,
                                    bool act, int *red, int &bmin, int &bmax) {
    const int b = act ? bucket_of(double_of(keys[(int64_t)lev * nl + j]), base, NB) : 0;
    int lo = act ? b : INT32_MAX, hi = act ? b : -1;
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[wv] = lo;
        red[BLOCK / 64 + wv] = hi;
    }
    __syncthreads();
    bmin = red[0];
    bmax = red[BLOCK / 64];
    for (int w = 1; w < BLOCK / 64; w++) {
        bmin = min(bmin, red[w]);
        bmax = max(bmax, red[BLOCK / 64 + w]);
    }
    return b;
}

// cnt[lev][b] += 1 for every line of the level (cnt zero on entry)
__global__ __launch_bounds__(BLOCK) void bucket_hist_kernel(const uint64_t *keys, int64_t nl, double base, int NB,
                                                            int32_t *cnt) {
    __shared__ int h[HIST_SPAN];
    __shared__ int red[2 * (BLOCK / 64)];
    const int lev = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool act = j < nl;
    int bmin, bmax;
    const int b = block_buckets(keys, nl, base, NB, lev, j, act, red, bmin, bmax);
    int32_t *c = cnt + (int64_t)lev * NB;
    if (bmax - bmin >= HIST_SPAN) {
        if (act) atomicAdd(c + b, 1);
        return;
    }
    const int span = bmax - bmin + 1;
    for (int i = threadIdx.x; i < span; i += BLOCK) h[i] = 0;
    __syncthreads();
    if (act) atomicAdd(h + (b - bmin), 1);
    __syncthreads();
    for (int i = threadIdx.x; i < span; i += BLOCK)
        if (h[i]) atomicAdd(c + bmin + i, h[i]);
}

__device__ inline int block_exclusive_scan(int x, int *tmp, int &total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = x;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
    }
    if (lane == 63) tmp[wv] = inc;
    __syncthreads();
    int woff = 0;
    total = 0;
    for (int w = 0; w < BLOCK / 64; w++) {
        if (w < wv) woff += tmp[w];
        total += tmp[w];
    }
    __syncthreads();
    return woff + inc - x;
}

// part[lev][blk] = sum of the block's SCAN_CHUNK counts
__global__ __launch_bounds__(BLOCK) void scan_partial_kernel(const int32_t *cnt, int NB, int nblk, int32_t *part) {
    __shared__ int tmp[BLOCK / 64];
    const int lev = blockIdx.y, blk = blockIdx.x;
    const int32_t *c = cnt + (int64_t)lev * NB;
    int s = 0;
    for (int e = 0; e < 16; e++) {
        const int b = blk * SCAN_CHUNK + e * BLOCK + threadIdx.x;
        if (b < NB) s += c[b];
    }
    int total;
    block_exclusive_scan(s, tmp, total);
    if (threadIdx.x == 0) part[(int64_t)lev * nblk + blk] = total;
}

// exclusive scan of part[lev][0..nblk) in place, one workgroup per level
__global__ __launch_bounds__(BLOCK) void scan_top_kernel(int32_t *part, int nblk) {
    __shared__ int tmp[BLOCK / 64];
    int32_t *p = part + (int64_t)blockIdx.y * nblk;
    int carry = 0;
    for (int base = 0; base < nblk; base += BLOCK) {
        const int i = base + threadIdx.x;
        const int x = i < nblk ? p[i] : 0;
        int total;
        const int ex = block_exclusive_scan(x, tmp, total);
        if (i < nblk) p[i] = carry + ex;
        carry += total;
    }
}

// bst[lev][b] = first sorted position of bucket b (b = 0..NB; bst[NB] = nl).  The chunk
// passes through LDS (padded: element i at i + i/16) so that global reads and writes are
// coalesced while each thread scans 16 consecutive buckets.
__device__ inline int scan_pad(int i) { return i + (i >> 4); }

__global__ __launch_bounds__(BLOCK) void scan_final_kernel(const int32_t *cnt, int NB, int nblk, const int32_t *part,
                                                           int32_t *bst) {
    __shared__ int tmp[BLOCK / 64];
    __shared__ int tile[SCAN_CHUNK + SCAN_CHUNK / 16];
    const int lev = blockIdx.y, blk = blockIdx.x;
    const int32_t *c = cnt + (int64_t)lev * NB;
    int32_t *o = bst + (int64_t)lev * (NB + 1);
    const int base = blk * SCAN_CHUNK;
    for (int e = 0; e < 16; e++) {
        const int i = e * BLOCK + threadIdx.x;
        tile[scan_pad(i)] = base + i < NB ? c[base + i] : 0;
    }
    __syncthreads();
    // each thread owns 16 consecutive buckets of the chunk
    const int i0 = threadIdx.x * 16;
    int v[16], s = 0;
    for (int e = 0; e < 16; e++) {
        v[e] = tile[scan_pad(i0 + e)];
        s += v[e];
    }
    int total;
    int run = part[(int64_t)lev * nblk + blk] + block_exclusive_scan(s, tmp, total);
    for (int e = 0; e < 16; e++) {
        tile[scan_pad(i0 + e)] = run;
        run += v[e];
    }
    __syncthreads();
    for (int e = 0; e < 16; e++) {
        const int i = e * BLOCK + threadIdx.x;
        if (base + i < NB) o[base + i] = tile[scan_pad(i)];
    }
    if (blk == nblk - 1 && threadIdx.x == BLOCK - 1) o[NB] = run;
}

// place every line in its bucket's range (order inside the bucket set by the atomics);
// cnt is consumed (back to zero).  A workgroup whose lines span fewer than HIST_SPAN
// buckets ranks its lines per bucket in LDS and takes each bucket's slots with one global
// atomic: its lines of bucket b get positions bst[b] + [old - n, old), old = the count
// before the atomic, n = its lines in b.
__global__ __launch_bounds__(BLOCK) void bucket_scatter_kernel(const uint64_t *keys, const int32_t *idx, int64_t nl,
                                                               double base, int NB, const int32_t *bst, int32_t *cnt,
                                                               uint64_t *skey, int32_t *sidx) {
    __shared__ int h[HIST_SPAN];
    __shared__ int red[2 * (BLOCK / 64)];
    const int lev = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool act = j < nl;
    int bmin, bmax;
    const int b = block_buckets(keys, nl, base, NB, lev, j, act, red, bmin, bmax);
    const int64_t o = (int64_t)lev * nl;
    int32_t *c = cnt + (int64_t)lev * NB;
    const int32_t *s0 = bst + (int64_t)lev * (NB + 1);
    int pos = 0;
    if (bmax - bmin >= HIST_SPAN) {
        if (act) pos = s0[b] + atomicSub(c + b, 1) - 1;
    } else {
        const int span = bmax - bmin + 1;
        for (int i = threadIdx.x; i < span; i += BLOCK) h[i] = 0;
        __syncthreads();
        const int r = act ? atomicAdd(h + (b - bmin), 1) : 0;  // rank among the workgroup's lines of b
        __syncthreads();
        for (int i = threadIdx.x; i < span; i += BLOCK) {
            const int n = h[i];
            if (n) h[i] = atomicSub(c + bmin + i, n) - n;  // first slot of the workgroup's lines of b
        }
        __syncthreads();
        if (act) pos = s0[b] + h[b - bmin] + r;
    }
    if (act) {
        skey[o + pos] = keys[o + j];
        sidx[o + pos] = idx[o + j];
    }
}

// Final order and line records: a line at position j of the scattered arrays takes
// position bst[b] + (its rank by (key, line index) among the lines of its bucket b) --
// the stable order by shifted centre, whatever order the atomics left inside the bucket
// (buckets hold a few lines each).  The thread writes the sorted key and index there and
// the line's profile parameters (LineShape: one record per sorted line, read by the
// deposit kernels in one load).  A workgroup's lines land on its own positions
// [j0, j0 + BLOCK) except where a bucket straddles its edges, so the records go through LDS
// and out as coalesced dwords (only the slots this workgroup filled); the few that land
// outside are written directly.
template <typename R>
__global__ __launch_bounds__(BLOCK) void order_shape_kernel(const uint64_t *keys, const int32_t *idx, int64_t nl,
                                                            double base, int NB, const int32_t *bst, int64_t lmin,
                                                            const double *nus, const double *gD, const double *gL,
                                                            const double *S, const double *sscale, int lev0,
                                                            uint64_t *skey, int32_t *sidx, LineShape<R> *shp) {
    constexpr int W = sizeof(LineShape<R>) / 4;  // dwords per record
    __shared__ double rec[BLOCK * W / 2];
    __shared__ int filled[BLOCK];
    const int lev = blockIdx.y;
    const int64_t j0 = (int64_t)blockIdx.x * BLOCK, j = j0 + threadIdx.x;
    const int64_t o = (int64_t)lev * nl;
    filled[threadIdx.x] = 0;
    __syncthreads();
    if (j < nl) {
        const uint64_t k = keys[o + j];
        const int32_t x = idx[o + j];
        const int b = bucket_of(double_of(k), base, NB);
        const int32_t *s0 = bst + (int64_t)lev * (NB + 1);
        const int lo = s0[b], hi = s0[b + 1];
        int rank = 0;
        for (int i = lo; i < hi; i++) {
            const uint64_t kk = keys[o + i];
            const int32_t xx = idx[o + i];
            rank += (kk < k || (kk == k && xx < x)) ? 1 : 0;
        }
        const int64_t q = lo + rank;
        skey[o + q] = k;
        sidx[o + q] = x;
        const int64_t i = o + (x - lmin);
        LineShape<R> p = make_line<R>(gD[i], gL[i], S[i] * sscale[lev0 + lev]);
        p.nus = nus[i];
        p.idx = x;
        if constexpr (sizeof(R) == 4) {
            // the far record, after the batch's gridDim.y levels of line records
            FarRec f;
            f.nus = p.nus;
            f.idx = x;
            f.xs = p.xs;
            f.lorc = p.lorc;
            f.gl2 = p.gl2;
            reinterpret_cast<FarRec *>(shp + (int64_t)gridDim.y * nl)[o + q] = f;
        }
        if (q >= j0 && q < j0 + BLOCK) {
            *(LineShape<R> *)(rec + (q - j0) * (W / 2)) = p;
            filled[q - j0] = 1;
        } else {
            shp[o + q] = p;
        }
    }
    __syncthreads();
    const int n = (int)min((int64_t)BLOCK, nl - j0);
    const uint32_t *src = (const uint32_t *)rec;
    uint32_t *dst = (uint32_t *)(shp + o + j0);
    for (int d = threadIdx.x; d < n * W; d += BLOCK)
        if (filled[d / W]) dst[d] = src[d];
}

// FP32 form: sample scale 2^-emax of every level from the per-level maximum exponent
__global__ void scale_kernel(const int *emax, int nlev, double *sscale) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nlev) sscale[l] = emax[l] > -2000 ? __builtin_ldexp(1.0, -emax[l]) : 1.0;
}
