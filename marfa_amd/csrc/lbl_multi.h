// lbl_multi.h -- launches of the deposit kernels over a device work list that holds many
// jobs' records and tiles (the batched form).  Included by marfa_lbl.hip inside its kernel
// namespace.
//
// A work list is a table of pieces -- one job's level batch each, with the deposit kernels'
// arguments (KArgs) -- and the prefix `first` of their item counts (records for the level-0
// kernel, record tiles for the fine kernel).  One workgroup per item; the item of a
// workgroup interleaves the XCDs in runs of XCD_RUN consecutive items: workgroups are
// dispatched round-robin over the 8 XCDs, so the items of a run (neighbouring records,
// which share most of their lines) meet in one L2, while the heavy and the light parts of
// the list are spread evenly over the XCDs.  Every item is computed with the single-job
// kernels' code and its own inputs, so its result is the single-job run's.
#pragma once

constexpr int XCD_RUN = 8;

// item of workgroup b (grid: a multiple of 8 XCD_RUN)
__device__ inline int run_item(int b) {
    const int slot = b >> 3;  // the workgroup's position among its XCD's workgroups
    return (slot / XCD_RUN) * (8 * XCD_RUN) + (b & 7) * XCD_RUN + slot % XCD_RUN;
}

// the piece holding item it: the last p with first[p] <= it (pieces without items skipped)
__device__ inline int piece_of(const int *first, int npieces, int it) {
    int lo = 0, hi = npieces - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// level-0 deposit of every record of the pieces (coarse_kernel's work)
template <typename R, int CHI>
__global__ __launch_bounds__(BLOCK, sizeof(R) == 4 ? 3 : 2) void coarse_multi(const KArgs *__restrict__ tab,
                                                                               const int *__restrict__ first,
                                                                               int npieces) {
    const int it = run_item(blockIdx.x);
    if (it >= first[npieces]) return;
    const int p = piece_of(first, npieces, it);
    coarse_item<R, CHI>(tab[p], it - first[p]);
}

// grids 1..9, fine points and cascade of every record tile of the pieces (fine_kernel's work)
template <typename R, int CAP>
__global__ __launch_bounds__(BLOCK, sizeof(R) == 4 ? 4 : 1) void fine_multi(const KArgs *__restrict__ tab,
                                                                             const int *__restrict__ first,
                                                                             int npieces) {
    __shared__ double g[NACC];
    __shared__ double s_u[CAP];
    __shared__ int32_t s_j[CAP];
    __shared__ int32_t s_b[STAGE_NB + 1];
    __shared__ int s_cnt[BLOCK / 64];
    __shared__ double red[BLOCK / 64 * 6];
    const int it = run_item(blockIdx.x);
    if (it >= first[npieces]) return;
    const int p = piece_of(first, npieces, it);
    fine_tile<R>(tab[p], it - first[p], g, red, s_u, s_j, s_b, s_cnt, CAP);
}
