// gfx950 kernels of the monodomain time-step: shared device helpers.
// 256-thread blocks (four 64-lane wavefronts); each block owns a contiguous
// slice of rows and block indices are remapped so that the blocks one XCD
// runs cover a contiguous range (its private L2 then holds the neighbouring
// planes the stencil gathers).
//
// Reductions never use fences or atomics: every producing block writes its
// partial sums to its own slot, and every block of the *next* kernel sums all
// slots in the same fixed order (the kernel boundary orders the writes).  The
// totals are therefore bitwise identical in every block and on every run.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "cmx.h"

namespace cmx {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kNumXcd = 8;
constexpr int kMaxBlocks = 1024;   // partial-sum slots per reduction

// Time of the step being computed, advanced on the device by the finish
// kernel so a step's launches have fixed arguments (and can be replayed).
struct StepClock {
    double t0, dt, t_end, eps;   // TimeStepper(t0, t_end, dt) of one cmx_step call
    int32_t k;                   // steps completed
    int32_t n_steps;
    int32_t poison;              // a step's final solve missed its iteration budget:
    int32_t failed_step;         //   every later kernel is a no-op until the host resumes it
    int32_t nonfinite;
    int32_t pad;
};

// One linear solve's control block.
struct SolveCtl {
    double tol2;                 // squared stopping threshold on ||D^-1 r||
    double bb;                   // ||D^-1 b||^2 of the step
    double zz;                   // ||D^-1 r||^2 at exit
    int32_t it;                  // iterations applied
    int32_t done;                // 0 running, 1 converged, 2 iteration limit, 3 non-finite
    int32_t max_its;
    int32_t pad;
};

__device__ __forceinline__ double step_time(const StepClock& c, int k)
{
    return c.t0 + k * c.dt;
}

// TimeStepper::CalculateNextTime (global/src/TimeStepper.cpp:106-121).
__device__ __forceinline__ double step_next_time(const StepClock& c, int k)
{
    double n = c.t0 + (k + 1) * c.dt;
    if (c.t_end - n <= c.eps) n = c.t_end;
    return n;
}

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
    __device__ __forceinline__ int dslot(int, int b) const { return b / width; }
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

// Block-sum of N doubles into out[blk*N + j] (plain stores, thread 0).
template <int N>
__device__ __forceinline__ void write_partials(double (&v)[N], double* __restrict__ out, int blk)
{
    __shared__ double red[kBlock / kWave][N];
    wave_sum<N>(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) red[wid][j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w) s += red[w][j];
            out[(size_t)blk * N + j] = s;
        }
}

// Sum of the G slots written by the previous kernel, identical in every block.
template <int N>
__device__ __forceinline__ void read_partials(const double* __restrict__ in, int G, double (&tot)[N])
{
    __shared__ double red[kBlock / kWave][N];
    double acc[N];
#pragma unroll
    for (int j = 0; j < N; ++j) acc[j] = 0;
    for (int b = threadIdx.x; b < G; b += kBlock)
#pragma unroll
        for (int j = 0; j < N; ++j) acc[j] += in[(size_t)b * N + j];
    wave_sum<N>(acc);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
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
    __syncthreads();
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
