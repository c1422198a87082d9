// dcm_device.h -- device helpers shared by the kernels: index maps of the
// one-rank boundary semantics and deterministic block reductions.
#pragma once

#include "dcm_internal.h"

namespace dcm {

// Periodic source index of a lateral halo index (0 -> n, n+1 -> 1).
__device__ __forceinline__ int wrap(int j, int n) {
    return j == 0 ? n : (j == n + 1 ? 1 : j);
}
// Source level of an edge or corner on a z plane after exchange_full: the
// receive from MPI_PROC_NULL leaves the z-wrapped local copy
// (mpi_module.f90:1244-1263): 0 -> kp, kp+1 -> 1.
__device__ __forceinline__ int zwrap(int k, int kp) {
    return k == 0 ? kp : (k == kp + 1 ? 1 : k);
}
// Neumann copy of mat_ax (dynamics.f90:206-211): 0 -> 1, kp+1 -> kp.
__device__ __forceinline__ int zclamp(int k, int kp) {
    return k < 1 ? 1 : (k > kp ? kp : k);
}

// Linear thread / block ids for 3-D launches.
__device__ __forceinline__ int lin_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ unsigned lin_bid() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ unsigned lin_nblocks() {
    return gridDim.x * gridDim.y * gridDim.z;
}

// Sum NS values over a block of kBlock threads in a fixed order and store
// the block's partial sums at partials[s*stride + block].  Every thread of
// the block must call this (inactive threads contribute zero).
template <int NS>
__device__ __forceinline__ void block_sum_store(double (&v)[NS], double *partials,
                                                long long stride) {
    __shared__ double red[NS][kBlock / 64];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[s] += __shfl_down(v[s], off, 64);
    }
    const int t = lin_tid();
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) red[s][wave] = v[s];
    }
    __syncthreads();
    if (t == 0) {
        const unsigned b = lin_bid();
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double acc = red[s][0];
#pragma unroll
            for (int q = 1; q < kBlock / 64; ++q) acc += red[s][q];
            partials[s * stride + b] = acc;
        }
    }
}

// As block_sum_store, storing at partials[s*stride + idx] (idx: the block's
// tile number when blocks are renumbered).
template <int NS>
__device__ __forceinline__ void block_sum_store_at(double (&v)[NS], double *partials,
                                                   long long stride, int idx) {
    __shared__ double red[NS][kBlock / 64];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[s] += __shfl_down(v[s], off, 64);
    }
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) red[s][wave] = v[s];
    }
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double acc = red[s][0];
#pragma unroll
            for (int q = 1; q < kBlock / 64; ++q) acc += red[s][q];
            partials[s * stride + idx] = acc;
        }
    }
}

__device__ __forceinline__ double block_min_store(double v, double *partials) {
    __shared__ double red[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(v, off, 64);
        v = o < v ? o : v;
    }
    const int t = lin_tid();
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (t == 0) {
        double acc = red[0];
#pragma unroll
        for (int q = 1; q < kBlock / 64; ++q) acc = red[q] < acc ? red[q] : acc;
        partials[lin_bid()] = acc;
    }
    return v;
}

// Final reduction of one slot by a single block of 1024 threads: thread t
// sums blocks t, t+1024, ... sequentially, then a fixed tree.  Deterministic
// for a given number of blocks.
__device__ __forceinline__ double reduce_slot_1024(const double *p, int n) {
    __shared__ double red[1024];
    const int t = threadIdx.x;
    // four independent chains keep several loads in flight
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = t;
    for (; b + 3 * 1024 < n; b += 4 * 1024) {
        a0 += p[b];
        a1 += p[b + 1024];
        a2 += p[b + 2048];
        a3 += p[b + 3072];
    }
    for (; b < n; b += 1024) a0 += p[b];
    double acc = (a0 + a1) + (a2 + a3);
    red[t] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double reduce_min_1024(const double *p, int n) {
    __shared__ double red[1024];
    const int t = threadIdx.x;
    double acc = p[0];
    for (int b = t; b < n; b += 1024) acc = p[b] < acc ? p[b] : acc;
    red[t] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) red[t] = red[t + w] < red[t] ? red[t + w] : red[t];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace dcm
