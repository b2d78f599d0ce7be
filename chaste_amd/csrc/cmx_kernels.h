// gfx950 kernels of the monodomain time-step.  One thread per node/row,
// 256-thread blocks (four 64-lane wavefronts), block indices remapped so each
// XCD sweeps a contiguous slice of rows (its private L2 then holds the
// neighbouring planes the stencil gathers).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "cmx.h"

namespace cmx {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kNumXcd = 8;

// Device-side CG control block; written only by the last block of a
// reduction kernel, read by every later kernel of the step.
struct CgScalars {
    double rz;       // (r, z) of the current iterate
    double alpha, beta;
    double zz;       // ||D^-1 r||^2 of the current iterate
    double tol2;     // squared stopping threshold
    double bb;       // ||D^-1 b||^2
    int32_t it;      // iterations done in this solve
    int32_t max_its;
    int32_t done;    // 0 running, 1 converged, 2 iteration limit
    int32_t nonfinite;
    uint32_t ticket; // last-block counter (reset by the last block)
    uint32_t pad;
};

// Hardware dispatches block b to XCD b % 8; give XCD x the contiguous range
// of logical blocks [x*q + min(x, r), ...) of its q (+1) blocks.
__device__ __forceinline__ int xcd_block(int b, int G)
{
    const int x = b & (kNumXcd - 1), k = b >> 3;
    const int q = G / kNumXcd, r = G % kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

// Row accessors.  slot() indexes the value arrays, dslot() the diagonal
// inverse, col() gives the column of slot k of row i.
struct StencilRows {
    const uint16_t* __restrict__ pid;
    const int32_t* __restrict__ off;
    int width;
    __device__ __forceinline__ int base(int i) const { return (int)pid[i] * width; }
    __device__ __forceinline__ int col(int i, int b, int k) const { return i + off[b + k]; }
    __device__ __forceinline__ int slot(int, int b, int k) const { return b + k; }
    __device__ __forceinline__ int dslot(int i, int b) const { return b / width; }
};

struct EllRows {
    const int32_t* __restrict__ cols;
    int width;
    int n;
    __device__ __forceinline__ int base(int i) const { return i; }
    __device__ __forceinline__ int col(int i, int, int k) const { return cols[(size_t)k * n + i]; }
    __device__ __forceinline__ int slot(int i, int, int k) const { return k * n + i; }
    __device__ __forceinline__ int dslot(int i, int) const { return i; }
};

template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N])
{
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1)
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], s, kWave);
}

// Block-reduces N doubles; the last block to finish (ticket) sums the
// per-block partials in block order (deterministic) and returns true on all
// of its threads with the totals in tot.
template <int N>
__device__ __forceinline__ bool reduce_last_block(double (&v)[N], double* __restrict__ partials,
                                                  uint32_t* ticket, int blk, int G, double (&tot)[N])
{
    __shared__ double red[kBlock / kWave][N];
    __shared__ bool last;
    wave_sum<N>(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) red[wid][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w) s += red[w][j];
            partials[(size_t)blk * N + j] = s;
        }
        __threadfence();
        const uint32_t t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == (uint32_t)(G - 1));
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[N];
#pragma unroll
    for (int j = 0; j < N; ++j) acc[j] = 0;
    // strided per-thread sums, then a fixed-order tree: deterministic for a given G
    for (int b = threadIdx.x; b < G; b += kBlock)
#pragma unroll
        for (int j = 0; j < N; ++j) acc[j] += partials[(size_t)b * N + j];
    __syncthreads();
    wave_sum<N>(acc);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) red[wid][j] = acc[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w][j];
        tot[j] = s;
    }
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// SimpleStimulus / RegularStimulus at time t (uA/cm^3); duration already padded.
__device__ __forceinline__ double stimulus_at(const cmx_stimulus& s, double t)
{
    if (s.period > 0.0) {
        if (t < s.start) return 0.0;
        const double beat = fmod(t - s.start, s.period);
        return (beat >= 0.0 && beat <= s.duration && t <= s.stop) ? s.magnitude : 0.0;
    }
    return (s.start <= t && t <= s.duration + s.start) ? s.magnitude : 0.0;
}

}  // namespace cmx
