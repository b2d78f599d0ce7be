// gfx950 kernels of the monodomain time-step: shared device helpers.
// The cell kernel runs 256-thread blocks, one node per thread.  The solver
// kernels run 1024-thread blocks (sixteen 64-lane wavefronts), by default one
// block per CU; each block owns a contiguous slice of rows and block indices
// are remapped so that the blocks one XCD runs cover a contiguous range (its
// private L2 then holds the neighbouring planes the stencil gathers).
//
// Reductions never use fences or atomics: every producing block writes its
// partial sums to its own slot, and every block of the *next* kernel sums all
// slots in the same fixed order (the kernel boundary orders the writes).  The
// totals are therefore bitwise identical in every block and on every run.
// With G blocks a reduction costs G^2 slot reads, hence few, large blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "cmx.h"

namespace cmx {

constexpr int kBlock = 256;       // cell kernel
constexpr int kCgBlock = 1024;    // solver kernels
constexpr int kCgGrid = 256;      // default solver blocks (one per CU)
constexpr int kWave = 64;
constexpr int kNumXcd = 8;
constexpr int kMaxBlocks = kCgBlock;   // partial-sum slots per reduction (one per thread of a consumer)

// Time of the step being computed, advanced on the device so that a step's
// launches have fixed arguments (and can be replayed).  Step k ends in two
// parts: the kernel that starts step k+1 (or the batch's finish kernel)
// decides step k's final solve and writes V^{k+1}; the residual kernel of
// step k+1 then advances k.
struct StepClock {
    double t0, dt, t_end, eps;   // TimeStepper(t0, t_end, dt) of one cmx_step call
    int32_t k;                   // steps completed
    int32_t n_steps;
    int32_t poison;              // a step's final solve missed its iteration budget:
    int32_t failed_step;         //   every later kernel is a no-op until the host resumes it
    int32_t nonfinite;
    int32_t pending;             // step k's solves are done, its voltage update is not
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

// Rows of one block: a contiguous chunk [lo, hi) of ceil(n/G) rows walked
// with stride kCgBlock.
struct RowRange {
    int lo, hi;
    __device__ __forceinline__ RowRange(int n, int blk, int G)
    {
        const int per = (n + G - 1) / G;
        lo = blk * per;
        hi = min(n, lo + per);
    }
};

constexpr int kMaxPatterns = 64;   // stencil format: distinct row patterns
constexpr int kMaxStencil = 16;    // stencil format: offsets per row (padded width)

// Row accessors.  row(i) caches what a row needs; col(r, k) is the column of
// slot k, vidx(r, k) the index into the value tables, didx(r) / didx_of(j)
// the index into the diagonal-inverse table of row i / of any row j.
//
// Uniform-offset stencil: offsets are kernel arguments (held in SGPRs),
// values and diagonal inverses are per-pattern tables that the kernels stage
// in LDS (kStaged).  Out-of-range columns of boundary rows are clamped to a
// valid row; their values are 0.
template <int W>
struct StencilRows {
    static constexpr bool kStaged = true;
    static constexpr int kPreW = W;      // gathers a kernel can hold in registers
    static constexpr int kIterWaves = 4; // CG iteration: the 4 wavefronts per SIMD a 1024-thread block implies
    const uint8_t* __restrict__ pid;
    int n;
    int n_patterns;
    int pad;                             // guard elements each side of a gathered vector (>= max |off|)
    int32_t off[W];
    struct Row { int i; int p; };
    __device__ __forceinline__ constexpr int width() const { return W; }
    __device__ __forceinline__ Row row(int i) const { return Row{i, (int)pid[i]}; }
    __device__ __forceinline__ int col(const Row& r, int k) const
    {
        const int j = r.i + off[k];
        return j < 0 ? 0 : (j >= n ? n - 1 : j);
    }
    __device__ __forceinline__ int vidx(const Row& r, int k) const { return r.p * W + k; }
    __device__ __forceinline__ int didx(const Row& r) const { return r.p; }
    __device__ __forceinline__ int didx_of(int j) const { return (int)pid[j]; }
    __device__ __forceinline__ int value_count() const { return n_patterns * W; }
    __device__ __forceinline__ int diag_count() const { return n_patterns; }
};

// ELL: values (and column indices) in global memory in 64-row slices, each
// slice holding its w slots one after another ([n/64][w][64]): a wavefront's
// rows read one contiguous block, slot by slot.  Either column indices (cols)
// or, with offset sets (sets != null), one set id per row: slot k of row i is
// column i + set_off[sets[i] * w + k], gathered vectors then carrying guard
// zones as wide as the largest offset.
struct EllRows {
    static constexpr bool kStaged = false;
    static constexpr int kPreW = 1;      // gathers are loaded where they are used
    // CG iteration: at least 8 wavefronts per SIMD (64 VGPRs), two blocks per
    // CU, so that more rows' gathers are in flight (FP64 ELL on a 13M-row mesh:
    // 37.2 -> 40.7 steps/s)
    static constexpr int kIterWaves = 8;
    const int32_t* __restrict__ cols;
    int w;
    int n;
    const uint8_t* __restrict__ sets = nullptr;
    const int32_t* __restrict__ set_off = nullptr;
    struct Row { int i; const int32_t* off; };
    __device__ __forceinline__ int width() const { return w; }
    __device__ __forceinline__ Row row(int i) const
    {
        return Row{i, sets ? set_off + (int)sets[i] * w : nullptr};
    }
    __device__ __forceinline__ int col(const Row& r, int k) const
    {
        return r.off ? r.i + r.off[k] : cols[vidx(r, k)];
    }
    __device__ __forceinline__ int vidx(const Row& r, int k) const { return ((r.i >> 6) * w + k) * 64 + (r.i & 63); }
    __device__ __forceinline__ int didx(const Row& r) const { return r.i; }
    __device__ __forceinline__ int didx_of(int j) const { return j; }
    __device__ __forceinline__ int value_count() const { return 0; }
    __device__ __forceinline__ int diag_count() const { return 0; }
};

static_assert(kMaxPatterns * kMaxStencil <= kCgBlock, "one table entry per solver thread");

// One entry of a per-pattern table per thread, loaded into a register so the
// load is issued together with the kernel's other loads; store() then puts it
// in LDS (the caller synchronises before reading).  The ELL format reads its
// global arrays directly.
template <class Rows, class T>
struct TableLoad {
    T v;
    bool on;
    __device__ __forceinline__ TableLoad(const T* __restrict__ g, int count)
    {
        on = Rows::kStaged && (int)threadIdx.x < count;
        v = on ? g[threadIdx.x] : T(0);
    }
    __device__ __forceinline__ const T* store(T* lds, const T* g) const
    {
        if constexpr (Rows::kStaged) {
            if (on) lds[threadIdx.x] = v;
            return lds;
        } else {
            (void)lds;
            return g;
        }
    }
};

// Raw buffer load of one element: address = descriptor base + voff + soff
// (bytes), so that all the gathers of a row share one 32-bit VGPR offset and
// take their per-neighbour offsets from SGPRs.
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <class T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, T v);
template <>
__device__ __forceinline__ void buf_store<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}
template <>
__device__ __forceinline__ void buf_store<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double v)
{
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 0);
}

// Descriptor over `bytes` bytes from p (planes of one array addressed by a
// per-thread offset and a per-plane scalar offset).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}

// Descriptor of a gathered vector x[-pad .. n+pad) (data format 32-bit, the
// gfx9 raw-buffer word 3).
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const T* x, int n, int pad)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)(x - pad), (short)0, (int)((uint32_t)(n + 2 * pad) * sizeof(T)),
                                             0x00020000);
}

// The gathered neighbour values of one row: held in registers for the
// stencil format (loaded before they are needed), read in place for ELL.
// Stencil gathers read x[i + off[k]] with no clamping: gathered vectors carry
// pad >= max |off| zero guard elements on each side, and the boundary rows'
// out-of-mesh slots have zero matrix values.
template <class Rows, class T>
struct Gathered {
    T g[Rows::kPreW];
    __device__ __forceinline__ void load(const Rows& R, const typename Rows::Row& row, const T* __restrict__ x)
    {
        if constexpr (Rows::kStaged) {
            const __amdgpu_buffer_rsrc_t rs = vec_rsrc(x, R.n, R.pad);
            const uint32_t voff = (uint32_t)row.i * (uint32_t)sizeof(T);
#pragma unroll
            for (int k = 0; k < Rows::kPreW; ++k)
                g[k] = buf_load<T>(rs, voff, (uint32_t)(R.pad + R.off[k]) * (uint32_t)sizeof(T));
        }
    }
    __device__ __forceinline__ T at(const Rows& R, const typename Rows::Row& row, const T* __restrict__ x, int k) const
    {
        if constexpr (Rows::kStaged) return g[k];
        else return x[R.col(row, k)];
    }
};

template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N])
{
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1)
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], s, kWave);
}

// Block-sum of N doubles into out[blk*N + j] (BS threads; plain stores).
template <int BS, int N>
__device__ __forceinline__ void write_partials(double (&v)[N], double* __restrict__ out, int blk)
{
    constexpr int NW = BS / kWave;
    __shared__ double red[NW][N];
    wave_sum<N>(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) red[wid][j] = v[j];
    __syncthreads();
    if (threadIdx.x < N) {
        const int j = threadIdx.x;
        double s = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[w][j];
        out[(size_t)blk * N + j] = s;
    }
}

// The G <= kMaxBlocks slots of the previous kernel's reduction, loaded with
// the kernel's other loads (BS threads, at most kMaxBlocks / BS slots each);
// total() sums them in a fixed order (identical in every block) and contains
// barriers.  The constructor issues the first slot of each thread without a
// branch (index clamped, the value masked in total()), so the kernel's later
// loads are issued before anything waits for it; slots past BS (G > BS) are
// loaded in total().
template <int N, int BS = kCgBlock>
struct PartialLoad {
    static constexpr int kTrips = (kMaxBlocks + BS - 1) / BS;
    const double* __restrict__ in;
    int G;
    double v[N];
    __device__ __forceinline__ PartialLoad(const double* __restrict__ in_, int G_) : in(in_), G(G_)
    {
        const int b = (int)threadIdx.x;
        const int bc = b < G ? b : (G > 0 ? G - 1 : 0);
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = in[(size_t)bc * N + j];
    }
    __device__ __forceinline__ void total(double (&tot)[N])
    {
        constexpr int NW = BS / kWave;
        __shared__ double red[NW][N];
        if ((int)threadIdx.x >= G)
#pragma unroll
            for (int j = 0; j < N; ++j) v[j] = 0.0;
#pragma unroll
        for (int t = 1; t < kTrips; ++t) {
            const int b = (int)threadIdx.x + t * BS;
            if (b < G)
#pragma unroll
                for (int j = 0; j < N; ++j) v[j] += in[(size_t)b * N + j];
        }
        wave_sum<N>(v);
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < N; ++j) red[wid][j] = v[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += red[w][j];
            tot[j] = s;
        }
        __syncthreads();
    }
};

// Element i of an array with a wave-uniform base, addressed as base + a
// 32-bit byte offset (arrays below 4 GiB): the compiler then uses the
// scalar-base + vector-offset form, one offset VGPR shared by every plane,
// instead of a 64-bit vector address per array.
template <class T>
__device__ __forceinline__ T ldu(const T* __restrict__ base, uint32_t i)
{
    return *(const T*)((const char*)base + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stu(T* __restrict__ base, uint32_t i, T v)
{
    *(T*)((char*)base + i * (uint32_t)sizeof(T)) = v;
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
