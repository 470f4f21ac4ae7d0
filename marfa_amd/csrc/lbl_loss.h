// lbl_loss.h -- the spectral loss of the differentiable form and its gradient with respect to
// the records, in one pass over the points each.  Included by marfa_lbl.hip inside its
// kernel namespace.
//
// L = sum_i w_i (ln k_i - ln khat_i)^2 over the points where k_i > 0 and khat_i > 0, and
// dL/dk_i = 2 w_i (ln k_i - ln khat_i) / k_i there (0 elsewhere).  The logarithms and the
// sum are FP64 (records in FP32 or FP64).  The sum has a fixed order: LOSS_BLOCKS workgroups each take a fixed,
// strided share of the points, reduce it in a fixed tree and write one partial; one
// workgroup then adds the partials in a fixed tree.
#pragma once

constexpr int LOSS_BLOCKS = MARFA_LOSS_PARTIALS;

template <typename T>
__device__ inline double log_residual(T k, T kh) {
    return (k > T(0) && kh > T(0)) ? log((double)k) - log((double)kh) : 0.0;
}

// block sum in a fixed order (wave butterflies, then the waves in order): thread 0 holds it
__device__ inline double block_sum(double x) {
    __shared__ double red[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BLOCK / 64; w++) s += red[w];
    return s;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void loss_partial_kernel(int64_t n, const T *k, const T *kh, const float *w,
                                                            double wscale, double *partial) {
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
        const double d = log_residual(k[i], kh[i]);
        acc += (w ? (double)w[i] : 1.0) * d * d;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s * wscale;
}

__global__ __launch_bounds__(BLOCK) void loss_final_kernel(const double *partial, int nparts, double *loss) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nparts; b += BLOCK) acc += partial[b];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) *loss = s;
}

// grad_i = gscale * dL/dk_i (gscale: one double in device memory, e.g. the incoming gradient
// of the loss in a backward pass, or null for 1)
template <typename T>
__global__ __launch_bounds__(BLOCK) void loss_grad_kernel(int64_t n, const T *k, const T *kh, const float *w,
                                                         double wscale, const double *gscale, T *grad) {
    const double g = wscale * (gscale ? *gscale : 1.0);
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
        const T ki = k[i];
        const double d = log_residual(ki, kh[i]);
        grad[i] = d != 0.0 ? (T)(2.0 * g * (w ? (double)w[i] : 1.0) * d / (double)ki) : T(0);
    }
}
