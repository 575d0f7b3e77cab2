// Block-wide row statistics for one wire row held in registers:
// exact order statistics by radix select over LDS histograms,
// restore_baseline (sigproc/src/OmnibusSigProc.cxx:867-906), the robust
// RMS of ROI_formation::cal_RMS (sigproc/src/ROI_formation.cxx:364-389)
// and its threshold.
//
// Order statistics follow Waveform::percentile (util/src/Waveform.cxx:
// 85-106): the value at index min((size_t)(p * (float)n), n-1) of the
// sorted set, p a float.  Select works on the monotone unsigned image of
// the float bits; three passes of 11/11/10-bit digits over LDS
// histograms narrow the key down to one value, which is returned
// exactly (the same element nth_element would return).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wct {

constexpr int kHistBins = 2048;

__device__ __forceinline__ uint32_t f2key(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// Shared scratch of the row-statistics routines.
struct RowScratch {
    uint32_t hist[kHistBins];
    uint32_t wsum[32];
    uint32_t sel_bucket;
    uint32_t sel_rank;
    double dsum[32];
    uint32_t csum[32];
};

// Inclusive scan inside a wavefront (64 lanes).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) sh->csum[wid] = v;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += sh->csum[w];
    return tot;
}

template <int NT>
__device__ __forceinline__ double block_sum_f64(double v, RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) sh->dsum[wid] = v;
    __syncthreads();
    double tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += sh->dsum[w];
    return tot;
}

// rank-th smallest (0-based) key among this block's elements whose
// flag bit is set in fl (bit i <-> val[i]).  Every thread calls it.
template <int NT, int VPT>
__device__ uint32_t block_select(const float (&val)[VPT], const uint64_t fl, uint32_t rank,
                                 RowScratch* sh)
{
    constexpr int NW = NT / 64;
    constexpr int BPT = kHistBins / NT;   // bins per thread in the scan
    uint32_t prefix = 0, pmask = 0;
    const int shifts[3] = {21, 10, 0};
    const uint32_t dmasks[3] = {0x7ffu, 0x7ffu, 0x3ffu};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll 1
    for (int pass = 0; pass < 3; ++pass) {
        const int sh_ = shifts[pass];
        const uint32_t dm = dmasks[pass];
#pragma unroll
        for (int b = 0; b < BPT; ++b) sh->hist[threadIdx.x * BPT + b] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((fl >> i) & 1ull) {
                const uint32_t k = f2key(val[i]);
                if ((k & pmask) == prefix) atomicAdd(&sh->hist[(k >> sh_) & dm], 1u);
            }
        }
        __syncthreads();
        // exclusive prefix over bins: thread t owns bins [t*BPT, t*BPT+BPT)
        uint32_t local = 0;
#pragma unroll
        for (int b = 0; b < BPT; ++b) local += sh->hist[threadIdx.x * BPT + b];
        uint32_t incl = wave_incl_scan(local);
        if (lane == 63) sh->wsum[wid] = incl;
        __syncthreads();
        uint32_t base = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) base += (w < wid) ? sh->wsum[w] : 0u;
        uint32_t excl = base + incl - local;
        if (rank >= excl && rank < excl + local) {
            uint32_t c = excl;
            for (int b = 0; b < BPT; ++b) {
                const uint32_t h = sh->hist[threadIdx.x * BPT + b];
                if (rank < c + h) {
                    sh->sel_bucket = threadIdx.x * BPT + b;
                    sh->sel_rank = rank - c;
                    break;
                }
                c += h;
            }
        }
        __syncthreads();
        const uint32_t bucket = sh->sel_bucket;
        rank = sh->sel_rank;
        prefix |= bucket << sh_;
        pmask |= dm << sh_;
        __syncthreads();
    }
    return prefix;
}

__device__ __forceinline__ uint32_t pct_index(float p, uint32_t n)
{
    uint32_t mid = (uint32_t)(p * (float)n);
    return mid < n - 1 ? mid : n - 1;
}

// restore_baseline on one row (val in registers, valid bits in vmask).
// Returns the subtracted baseline (0 if the row has no non-zero sample).
template <int NT, int VPT>
__device__ float row_restore_baseline(float (&val)[VPT], uint64_t vmask, RowScratch* sh)
{
    uint64_t nz = 0;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if (((vmask >> i) & 1ull) && val[i] != 0.0f) { nz |= 1ull << i; ++c; }
    const uint32_t n1 = block_sum_u32<NT>(c, sh);
    if (n1 == 0) return 0.0f;
    const float b1 = key2f(block_select<NT, VPT>(val, nz, pct_index(0.5f, n1), sh));
    uint64_t s2 = 0;
    c = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if (((nz >> i) & 1ull) && fabsf(val[i] - b1) < 500.0f) { s2 |= 1ull << i; ++c; }
    const uint32_t n2 = block_sum_u32<NT>(c, sh);
    const float b2 = key2f(block_select<NT, VPT>(val, s2, pct_index(0.5f, n2), sh));
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if ((nz >> i) & 1ull) val[i] -= b2;
    return b2;
}

// ROI_formation::cal_RMS over the samples flagged in pmask.
template <int NT, int VPT>
__device__ float row_cal_rms(const float (&val)[VPT], uint64_t pmask, RowScratch* sh)
{
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i) c += (pmask >> i) & 1ull;
    const uint32_t n = block_sum_u32<NT>(c, sh);
    if (n == 0) return 0.0f;
    const float p0 = key2f(block_select<NT, VPT>(val, pmask, pct_index((float)(0.5 - 0.34), n), sh));
    const float p1 = key2f(block_select<NT, VPT>(val, pmask, pct_index(0.5f, n), sh));
    const float p2 = key2f(block_select<NT, VPT>(val, pmask, pct_index((float)(0.5 + 0.34), n), sh));
    const double d21 = (double)(p2 - p1), d10 = (double)(p1 - p0);
    const float r0 = (float)sqrt((d21 * d21 + d10 * d10) / 2.);
    const double lim = 5.0 * (double)r0;
    double s = 0;
    uint32_t k = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
        if (((pmask >> i) & 1ull) && (double)fabsf(val[i]) < lim) {
            s += (double)val[i] * (double)val[i];
            ++k;
        }
    }
    const double tot = block_sum_f64<NT>(s, sh);
    const uint32_t cnt = block_sum_u32<NT>(k, sh);
    if (cnt == 0) return 0.0f;
    return sqrtf((float)(tot / (double)cnt));
}

}  // namespace wct
