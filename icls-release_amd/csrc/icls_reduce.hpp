// icls_reduce.hpp -- deterministic block reductions (64-wide wavefronts).
// A reduction is two launches: blocks fold their values in a fixed tree and write one
// partial each; a single block then folds the partials in a fixed order.  The result
// is bit-identical from run to run (it differs from the reference's serial Fortran
// sum() by reassociation only).
#pragma once
#include <hip/hip_runtime.h>

namespace icls {

struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return fmax(a, b); } };
struct OpMin { __device__ double operator()(double a, double b) const { return fmin(a, b); } };

// All threads of the block must call this; nthreads <= 1024, multiple of 64.
template <class Op>
__device__ __forceinline__ double block_reduce(double v, Op op, double* smem /* >= 16 doubles */) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off));
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = (nthr + 63) >> 6;
    __syncthreads();
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    double r = 0.;
    if (tid == 0) {
        r = smem[0];
        for (int w = 1; w < nw; w++) r = op(r, smem[w]);
    }
    __syncthreads();
    return r;  // valid in thread 0
}

// linear id of the block in its 3D grid (x fastest)
__device__ __forceinline__ int block_linear() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// threads of the single block that folds the partials
constexpr int RED_FINAL_THREADS = 1024;

// Thread t folds partials t, t + T, t + 2T, ... (T threads) in that order, four loads
// in flight at a time; the block then folds the threads' values.  The order is fixed,
// so the result does not change from run to run.
template <class Op>
__device__ __forceinline__ double fold_partials(const double* __restrict__ part, int n, double init) {
    Op op;
    double v = init;
    const int T = blockDim.x;
    int q = threadIdx.x;
    for (; q + 3 * T < n; q += 4 * T) {
        const double a = part[q], b = part[q + T], c = part[q + 2 * T], d = part[q + 3 * T];
        v = op(op(op(op(v, a), b), c), d);
    }
    for (; q < n; q += T) v = op(v, part[q]);
    return v;
}

// Fold n partials into out[0] with one block of RED_FINAL_THREADS threads.
template <class Op>
__global__ void __launch_bounds__(RED_FINAL_THREADS)
k_reduce_final(const double* __restrict__ part, int n, double init, double* out) {
    __shared__ double smem[16];
    double r = block_reduce(fold_partials<Op>(part, n, init), Op(), smem);
    if (threadIdx.x == 0) out[0] = r;
}

// Block b folds the n partials at part + b * n into out[b], as k_reduce_final does
// for one array (same order, same result): one launch for several sums.
template <class Op>
__global__ void __launch_bounds__(RED_FINAL_THREADS)
k_reduce_final_rows(const double* __restrict__ part, int n, double init, double* out) {
    __shared__ double smem[16];
    double r = block_reduce(fold_partials<Op>(part + (long)blockIdx.x * n, n, init), Op(), smem);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// as k_reduce_final, skipped when the device switch flag[0] is zero
template <class Op>
__global__ void __launch_bounds__(RED_FINAL_THREADS)
k_reduce_final_pred(const double* __restrict__ part, int n, double init, double* out,
                    const double* __restrict__ flag) {
    if (flag[0] == 0.) return;
    __shared__ double smem[16];
    double r = block_reduce(fold_partials<Op>(part, n, init), Op(), smem);
    if (threadIdx.x == 0) out[0] = r;
}

}  // namespace icls
