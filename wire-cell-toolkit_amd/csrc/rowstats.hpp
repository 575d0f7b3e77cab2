// Block-wide row statistics for one wire row held in registers:
// exact order statistics by radix select over LDS histograms,
// restore_baseline (sigproc/src/OmnibusSigProc.cxx:867-906), the robust
// RMS of ROI_formation::cal_RMS (sigproc/src/ROI_formation.cxx:364-389)
// and its threshold.
//
// Order statistics follow Waveform::percentile (util/src/Waveform.cxx:
// 85-106): the value at index min((size_t)(p * (float)n), n-1) of the
// sorted set, p a float.  Selection works on the monotone unsigned image
// of the float bits and returns the exact element nth_element would.
//
// Fast path (no excluded samples in the row).  Every statistic the row
// needs is an order statistic of one set, the non-zero samples Z:
//  - median 1 is the middle of Z;
//  - median 2 is the middle of the samples of Z within 500 of median 1,
//    a contiguous run of sorted Z (|fl(x - m1)| grows monotonically away
//    from m1), so it is the element of Z at rank n_lo + mid(n2);
//  - after subtracting b, fl(x - b) is monotone in x and zeros stay zero,
//    so the 16/50/84% points of the row are either 0 or fl(z - b) for the
//    element z of Z at a rank shifted by the zero count.
// Z is therefore counting-sorted into LDS once, by linear bins spanning
// mean +- 8 sd of Z (plus an underflow and an overflow bin; the bin of a
// sample is a monotone function of it, so bins are contiguous in sorted
// order).  Each statistic then finds its bin from the offsets and selects
// exactly inside that bin's keys (rank counting for bins up to a block,
// radix passes above).
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

constexpr int kLinBins = 2048;          // interior linear bins
constexpr int kLinHist = 2560;          // histogram length (underflow, bins, overflow, padding)

// LDS scratch of the row-statistics routines.  The bin-sorted keys of
// the row follow the struct (row_keys()).
struct RowScratch {
    uint32_t off[kLinHist + 1];  // linear-bin counts, then exclusive offsets
    uint32_t cur[kLinHist];      // scatter cursors; reused as the radix histogram
    uint32_t wsum[32];
    uint32_t csum[4 * 32];
    double dsum[32];
    uint32_t sel_bucket;
    uint32_t sel_rank;
    uint32_t list_n;
    uint32_t pad_;
};

__device__ __forceinline__ uint32_t* row_keys(RowScratch* sh)
{
    return reinterpret_cast<uint32_t*>(sh + 1);
}

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

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums of up to four counters over the block.
template <int NT, int NC>
__device__ __forceinline__ void block_sums(uint32_t (&c)[NC], RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NC; ++q) c[q] = wave_sum(c[q]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NC; ++q) sh->csum[q * 32 + wid] = c[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += sh->csum[q * 32 + w];
        c[q] = t;
    }
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, RowScratch* sh)
{
    uint32_t c[1] = {v};
    block_sums<NT, 1>(c, sh);
    return c[0];
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

// Finds the bin of hist (nb bins) holding the rank-th element and the
// rank inside it; every thread gets them.
template <int NT>
__device__ __forceinline__ void hist_locate(const uint32_t* hist, int nb, uint32_t rank, uint32_t& bucket,
                            uint32_t& sub, RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int bpt = nb / NT;
    uint32_t local = 0;
    for (int b = 0; b < bpt; ++b) local += hist[threadIdx.x * bpt + b];
    const uint32_t incl = wave_incl_scan(local);
    if (lane == 63) sh->wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) base += (w < wid) ? sh->wsum[w] : 0u;
    const uint32_t excl = base + incl - local;
    if (rank >= excl && rank < excl + local) {
        uint32_t c = excl;
        for (int b = 0; b < bpt; ++b) {
            const uint32_t h = hist[threadIdx.x * bpt + b];
            if (rank < c + h) {
                sh->sel_bucket = threadIdx.x * bpt + b;
                sh->sel_rank = rank - c;
                break;
            }
            c += h;
        }
    }
    __syncthreads();
    bucket = sh->sel_bucket;
    sub = sh->sel_rank;
    __syncthreads();
}

// Linear binning of a row: bin 0 below lo, 1..kLinBins inside, kLinBins+1
// above.  Float subtraction and scaling are monotone, so is the bin.
struct LinBins {
    float lo;
    float inv_w;
};

__device__ __forceinline__ uint32_t lin_bin(float x, LinBins lb)
{
    const float t = (x - lb.lo) * lb.inv_w;
    return t < 0.f ? 0u : (t >= (float)kLinBins ? (uint32_t)kLinBins + 1u : 1u + (uint32_t)t);
}

// Exclusive scan of a[0..n) in place (n a multiple of NT), total in a[n].
template <int NT>
__device__ __forceinline__ void block_excl_scan(uint32_t* a, int n, RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int per = n / NT;
    uint32_t* my = a + threadIdx.x * per;
    uint32_t local = 0;
    for (int b = 0; b < per; ++b) local += my[b];
    const uint32_t incl = wave_incl_scan(local);
    if (lane == 63) sh->wsum[wid] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        base += (w < wid) ? sh->wsum[w] : 0u;
        tot += sh->wsum[w];
    }
    uint32_t run = base + incl - local;
    for (int b = 0; b < per; ++b) {
        const uint32_t c = my[b];
        my[b] = run;
        run += c;
    }
    if (threadIdx.x == 0) a[n] = tot;
    __syncthreads();
}

// Exact rank-th smallest of seg[0..cnt) (unordered keys in LDS).
template <int NT>
__device__ __forceinline__ uint32_t seg_select(const uint32_t* seg, uint32_t cnt, uint32_t rank, RowScratch* sh)
{
    if (cnt <= (uint32_t)NT) {
        // position in sorted order = smaller keys + equal keys before it
        if (threadIdx.x < cnt) {
            const uint32_t k = seg[threadIdx.x];
            uint32_t pos = 0;
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t o = seg[j];
                pos += (o < k) || (o == k && j < threadIdx.x);
            }
            if (pos == rank) sh->sel_bucket = k;
        }
        __syncthreads();
        const uint32_t key = sh->sel_bucket;
        __syncthreads();
        return key;
    }
    // three radix passes over the segment (11/11/10 bits)
    uint32_t* h = sh->cur;
    uint32_t prefix = 0, pmask = 0;
    const int shifts[3] = {21, 10, 0};
    const uint32_t dmasks[3] = {0x7ffu, 0x7ffu, 0x3ffu};
#pragma unroll 1
    for (int pass = 0; pass < 3; ++pass) {
        const int sh_ = shifts[pass];
        const uint32_t dm = dmasks[pass];
        for (int b = threadIdx.x; b < kHistBins; b += NT) h[b] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cnt; i += NT) {
            const uint32_t k = seg[i];
            if ((k & pmask) == prefix) atomicAdd(&h[(k >> sh_) & dm], 1u);
        }
        __syncthreads();
        uint32_t bucket, sub;
        hist_locate<NT>(h, pass == 2 ? 1024 : kHistBins, rank, bucket, sub, sh);
        rank = sub;
        prefix |= bucket << sh_;
        pmask |= dm << sh_;
    }
    return prefix;
}

// rank-th smallest of the bin-sorted keys (offsets in sh->off).
template <int NT>
__device__ __forceinline__ uint32_t sorted_select(uint32_t rank, RowScratch* sh)
{
    // the bin b with off[b] <= rank < off[b+1]
    const int per = kLinHist / NT;
    for (int q = 0; q < per; ++q) {
        const int b = threadIdx.x * per + q;
        if (sh->off[b] <= rank && rank < sh->off[b + 1]) {
            sh->sel_rank = b;
        }
    }
    __syncthreads();
    const uint32_t b = sh->sel_rank;
    const uint32_t o = sh->off[b], cnt = sh->off[b + 1] - o;
    __syncthreads();
    return seg_select<NT>(row_keys(sh) + o, cnt, rank - o, sh);
}

__device__ __forceinline__ uint32_t pct_index(float p, uint32_t n)
{
    uint32_t mid = (uint32_t)(p * (float)n);
    return mid < n - 1 ? mid : n - 1;
}

// Fast path: restore_baseline (if restore) and cal_RMS over all valid
// samples of the row.  The non-zero samples are counting-sorted into LDS
// by linear bin once; every order statistic is then a short exact select
// inside one bin.  Returns the RMS (0 when want_rms is false) and leaves
// the baselined samples in val.
template <int NT, int VPT>
__device__ __forceinline__ float row_stats_fast(float (&val)[VPT], uint64_t valid, bool restore, bool want_rms,
                                RowScratch* sh)
{
    uint64_t nz = 0;
    double s1 = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if (((valid >> i) & 1ull) && val[i] != 0.0f) {
            nz |= 1ull << i;
            s1 += val[i];
            s2 += (double)val[i] * val[i];
        }
    uint32_t c2[2] = {(uint32_t)__popcll(nz), (uint32_t)__popcll(valid)};
    block_sums<NT, 2>(c2, sh);
    const uint32_t n_nz = c2[0], n_valid = c2[1];
    if (n_nz > 0 && (restore || want_rms)) {
        const double S1 = block_sum_f64<NT>(s1, sh);
        const double S2 = block_sum_f64<NT>(s2, sh);
        const double mean = S1 / n_nz;
        const double sd = sqrt(fmax(S2 / n_nz - mean * mean, 0.0));
        LinBins lb;
        lb.lo = (float)(mean - 8.0 * sd);
        const double iw = sd > 0 ? (double)kLinBins / (16.0 * sd) : 0.0;
        lb.inv_w = (isfinite(lb.lo) && isfinite(iw)) ? (float)iw : 0.f;
        if (!isfinite(lb.lo)) lb.lo = 0.f;
        // counting sort of the non-zero keys by bin (bins recomputed in the
        // scatter rather than kept in registers)
        for (int b = threadIdx.x; b < kLinHist; b += NT) sh->off[b] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < VPT; ++i)
            if ((nz >> i) & 1ull) atomicAdd(&sh->off[lin_bin(val[i], lb)], 1u);
        __syncthreads();
        block_excl_scan<NT>(sh->off, kLinHist, sh);
        for (int b = threadIdx.x; b < kLinHist; b += NT) sh->cur[b] = sh->off[b];
        __syncthreads();
        uint32_t* keys = row_keys(sh);
#pragma unroll
        for (int i = 0; i < VPT; ++i)
            if ((nz >> i) & 1ull) keys[atomicAdd(&sh->cur[lin_bin(val[i], lb)], 1u)] = f2key(val[i]);
        __syncthreads();
    }
    float b = 0.0f;
    if (restore && n_nz > 0) {
        const float m1 = key2f(sorted_select<NT>(pct_index(0.5f, n_nz), sh));
        uint32_t c[2] = {0, 0};
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((nz >> i) & 1ull) {
                if (!(fabsf(val[i] - m1) < 500.0f)) {
                    if (val[i] < m1) ++c[0]; else ++c[1];
                }
            }
        }
        block_sums<NT, 2>(c, sh);
        const uint32_t n2 = n_nz - c[0] - c[1];
        b = key2f(sorted_select<NT>(c[0] + pct_index(0.5f, n2), sh));
    }
    float rms = 0.0f;
    if (want_rms && n_valid > 0) {
        // ranks of the 16/50/84% points among the baselined row
        uint32_t c[2] = {0, 0};
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((nz >> i) & 1ull) {
                const float y = val[i] - b;
                if (y < 0.0f) ++c[0];
                else if (y == 0.0f) ++c[1];
            }
        }
        block_sums<NT, 2>(c, sh);
        const uint32_t c_neg = c[0], c_eq = c[1], n_zero = n_valid - n_nz;
        const float ps[3] = {(float)(0.5 - 0.34), 0.5f, (float)(0.5 + 0.34)};
        float par[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const uint32_t r = pct_index(ps[q], n_valid);
            if (r < c_neg) {
                par[q] = key2f(sorted_select<NT>(r, sh)) - b;
            } else if (r < c_neg + n_zero + c_eq) {
                par[q] = 0.0f;
            } else {
                par[q] = key2f(sorted_select<NT>(r - n_zero, sh)) - b;
            }
        }
        const double d21 = (double)(par[2] - par[1]), d10 = (double)(par[1] - par[0]);
        const float r0 = (float)sqrt((d21 * d21 + d10 * d10) / 2.);
        const double lim = 5.0 * (double)r0;
        double s = 0;
        uint32_t k = 0;
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((valid >> i) & 1ull) {
                const float y = ((nz >> i) & 1ull) ? val[i] - b : val[i];
                if ((double)fabsf(y) < lim) {
                    s += (double)y * (double)y;
                    ++k;
                }
            }
        }
        const double tot = block_sum_f64<NT>(s, sh);
        const uint32_t cnt = block_sum_u32<NT>(k, sh);
        rms = cnt ? sqrtf((float)(tot / (double)cnt)) : 0.0f;
    }
    if (restore) {
#pragma unroll
        for (int i = 0; i < VPT; ++i)
            if ((nz >> i) & 1ull) val[i] -= b;
    }
    return rms;
}

// ---- general path: filtered radix select (rows with excluded samples) --

// rank-th smallest key among the flagged samples, three full passes.
template <int NT, int VPT>
__device__ __forceinline__ uint32_t block_select(const float (&val)[VPT], const uint64_t fl, uint32_t rank,
                                 RowScratch* sh)
{
    uint32_t prefix = 0, pmask = 0;
    const int shifts[3] = {21, 10, 0};
    const uint32_t dmasks[3] = {0x7ffu, 0x7ffu, 0x3ffu};
#pragma unroll 1
    for (int pass = 0; pass < 3; ++pass) {
        const int sh_ = shifts[pass];
        const uint32_t dm = dmasks[pass];
        for (int b = threadIdx.x; b < kHistBins; b += NT) sh->cur[b] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((fl >> i) & 1ull) {
                const uint32_t k = f2key(val[i]);
                if ((k & pmask) == prefix) atomicAdd(&sh->cur[(k >> sh_) & dm], 1u);
            }
        }
        __syncthreads();
        uint32_t bucket, sub;
        hist_locate<NT>(sh->cur, pass == 2 ? 1024 : kHistBins, rank, bucket, sub, sh);
        rank = sub;
        prefix |= bucket << sh_;
        pmask |= dm << sh_;
    }
    return prefix;
}

// restore_baseline on one row (val in registers, valid bits in vmask).
template <int NT, int VPT>
__device__ __forceinline__ float row_restore_baseline(float (&val)[VPT], uint64_t vmask, RowScratch* sh)
{
    uint64_t nz = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if (((vmask >> i) & 1ull) && val[i] != 0.0f) nz |= 1ull << i;
    const uint32_t n1 = block_sum_u32<NT>((uint32_t)__popcll(nz), sh);
    if (n1 == 0) return 0.0f;
    const float b1 = key2f(block_select<NT, VPT>(val, nz, pct_index(0.5f, n1), sh));
    uint64_t s2 = 0;
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if (((nz >> i) & 1ull) && fabsf(val[i] - b1) < 500.0f) s2 |= 1ull << i;
    const uint32_t n2 = block_sum_u32<NT>((uint32_t)__popcll(s2), sh);
    const float b2 = key2f(block_select<NT, VPT>(val, s2, pct_index(0.5f, n2), sh));
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        if ((nz >> i) & 1ull) val[i] -= b2;
    return b2;
}

// ROI_formation::cal_RMS over the samples flagged in pmask.
template <int NT, int VPT>
__device__ __forceinline__ float row_cal_rms(const float (&val)[VPT], uint64_t pmask, RowScratch* sh)
{
    const uint32_t n = block_sum_u32<NT>((uint32_t)__popcll(pmask), sh);
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
