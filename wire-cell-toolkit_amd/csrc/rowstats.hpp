// Block-wide row statistics for one wire row held in registers:
// exact order statistics, restore_baseline (sigproc/src/
// OmnibusSigProc.cxx:867-906), the robust RMS of ROI_formation::cal_RMS
// (sigproc/src/ROI_formation.cxx:364-389) and its threshold.
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
// +-8 clipped standard deviations around the clipped mean of Z (moments
// of the samples within 3 sd of the mean, so signal on the row does not
// widen the bins of its noise), plus an underflow and an overflow bin.
// The bin of a sample is a monotone function of it, so bins are
// contiguous in sorted order.  Each statistic then finds its bin from the
// offsets and selects exactly inside that bin's keys (rank counting for
// bins up to a block, 8-bit radix passes above).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sync.hpp"

namespace wct {

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
constexpr int kLinHist = 2304;          // histogram length (underflow, bins, overflow, padding)
constexpr int kRadixBins = 256;         // 8-bit digits of the radix selections
constexpr int kSortSeg = 256;           // largest bin kept in sorted order

// LDS scratch of the row-statistics routines.  The bin-sorted keys of
// the row follow the struct (row_keys()).
struct RowScratch {
    // Linear-bin counts, then exclusive offsets; once the keys are
    // scattered, off[b] holds the end of bin b (the start of bin b + 1).
    uint32_t off[kLinHist + 1];
    uint32_t hist[kRadixBins];   // radix histogram
    uint32_t wsum[8];
    uint32_t csum[4][8];
    double dsum[8];
    uint32_t sel_bucket;
    uint32_t sel_rank;
    // Fast path: every phase of a row reduces into its own slots, so no
    // phase has to wait for the reads of the one before (phases of
    // consecutive rows are separated by the row's other barriers).
    float fpart[4][8];           // per-wave float sums
    uint32_t upart[5][2][8];     // [phase][counter][wave]
    double dpart[8];
    uint32_t sbin[5];            // located bins of the selections
    uint32_t skey[5];            // selected keys
    uint32_t sless[5];           // keys of the bin below the selected one
    uint32_t seq[5];             // keys equal to the selected one
    // The bin of the first median in sorted order: each key with the
    // number of keys of the bin below it and equal to it.
    uint32_t mkey[kSortSeg];
    uint32_t mlt[kSortSeg];
    uint32_t met[kSortSeg];
};

__device__ __forceinline__ uint32_t* row_keys(RowScratch* sh)
{
    return reinterpret_cast<uint32_t*>(sh + 1);
}

// Cross-lane steps through DPP (a VALU operand modifier: no LDS round
// trip).  DPP controls of the GFX9 encoding: row_shr:n = 0x110 + n
// (lane i of a row of 16 reads lane i - n of the same row), row_bcast:15
// = 0x142 (lane 15 of each row feeds the next row), row_bcast:31 = 0x143
// (lane 31 feeds rows 2 and 3).  update_dpp(old = 0, ...) leaves 0 in a
// lane whose source is outside its row or whose row is masked, the
// identity of the sums below.  Callers run these with the whole
// wavefront active.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}

// Inclusive scan over the 64 lanes: prefix sums inside each row of 16
// (shifts 1, 2, 4, 8), rows 1 and 3 then add the last lane of the row
// before them, rows 2 and 3 add lane 31.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)dpp_i<0x111, 0xf>((int)v);
    v += (uint32_t)dpp_i<0x112, 0xf>((int)v);
    v += (uint32_t)dpp_i<0x114, 0xf>((int)v);
    v += (uint32_t)dpp_i<0x118, 0xf>((int)v);
    v += (uint32_t)dpp_i<0x142, 0xa>((int)v);
    v += (uint32_t)dpp_i<0x143, 0xc>((int)v);
    return v;
}

// Sums over the wavefront; the total (lane 63 of the scan) is returned
// to every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v)));
}

__device__ __forceinline__ float wave_sum_f(float v)
{
    v += dpp_f<0x111, 0xf>(v);
    v += dpp_f<0x112, 0xf>(v);
    v += dpp_f<0x114, 0xf>(v);
    v += dpp_f<0x118, 0xf>(v);
    v += dpp_f<0x142, 0xa>(v);
    v += dpp_f<0x143, 0xc>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v)
{
    return __hiloint2double(dpp_i<CTRL, ROWS>(__double2hiint(v)), dpp_i<CTRL, ROWS>(__double2loint(v)));
}

__device__ __forceinline__ double wave_sum_d(double v)
{
    v += dpp_d<0x111, 0xf>(v);
    v += dpp_d<0x112, 0xf>(v);
    v += dpp_d<0x114, 0xf>(v);
    v += dpp_d<0x118, 0xf>(v);
    v += dpp_d<0x142, 0xa>(v);
    v += dpp_d<0x143, 0xc>(v);
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Sums of up to four counters over the block (two barriers; the slots
// are reused).
template <int NT, int NC>
__device__ __forceinline__ void block_sums(uint32_t (&c)[NC], RowScratch* sh)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NC; ++q) c[q] = wave_sum(c[q]);
    lds_barrier();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NC; ++q) sh->csum[q][wid] = c[q];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += sh->csum[q][w];
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
    v = wave_sum_d(v);
    lds_barrier();
    if (lane == 0) sh->dsum[wid] = v;
    lds_barrier();
    double tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += sh->dsum[w];
    return tot;
}

// Sums of NC counters over the block through the slots of one phase:
// one barrier.
template <int NT, int NC>
__device__ __forceinline__ void block_sums_slot(uint32_t (&c)[NC], uint32_t (*slot)[8])
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NC; ++q) c[q] = wave_sum(c[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NC; ++q) slot[q][wid] = c[q];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += slot[q][w];
        c[q] = t;
    }
}

// Finds the bin of hist (nb bins, a multiple of NT) holding the rank-th
// element and the rank inside it; every thread gets them.
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
    lds_barrier();
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
    lds_barrier();
    bucket = sh->sel_bucket;
    sub = sh->sel_rank;
    lds_barrier();
}

// Linear binning of a row: bin 0 below lo, 1..kLinBins inside, kLinBins+1
// above.  Float subtraction and scaling are monotone, so is the bin.
struct LinBins {
    float lo;
    float inv_w;
};

// A NaN position (an +inf zero-slot sentinel times inv_w == 0, the bins
// of a constant row) goes to the overflow bin with the other sentinels;
// no NaN reaches the float-to-integer conversion.
__device__ __forceinline__ uint32_t lin_bin(float x, LinBins lb)
{
    const float t = (x - lb.lo) * lb.inv_w;
    if (!(t < (float)kLinBins)) return (uint32_t)kLinBins + 1u;
    return t < 0.f ? 0u : 1u + (uint32_t)t;
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
        lds_barrier();
        const uint32_t key = sh->sel_bucket;
        lds_barrier();
        return key;
    }
    // four radix passes over the segment, 8 bits each
    uint32_t* h = sh->hist;
    uint32_t prefix = 0, pmask = 0;
#pragma unroll 1
    for (int sh_ = 24; sh_ >= 0; sh_ -= 8) {
        for (int b = threadIdx.x; b < kRadixBins; b += NT) h[b] = 0;
        lds_barrier();
        for (uint32_t i = threadIdx.x; i < cnt; i += NT) {
            const uint32_t k = seg[i];
            if ((k & pmask) == prefix) atomicAdd(&h[(k >> sh_) & 0xffu], 1u);
        }
        lds_barrier();
        uint32_t bucket, sub;
        hist_locate<NT>(h, kRadixBins, rank, bucket, sub, sh);
        rank = sub;
        prefix |= bucket << sh_;
        pmask |= 0xffu << sh_;
    }
    return prefix;
}

__device__ __forceinline__ uint32_t pct_index(float p, uint32_t n)
{
    uint32_t mid = (uint32_t)(p * (float)n);
    return mid < n - 1 ? mid : n - 1;
}

// Bin b of the scattered keys: off[b] is its end, off[b-1] its start.
__device__ __forceinline__ uint32_t bin_start(const RowScratch* sh, uint32_t b)
{
    return b ? sh->off[b - 1] : 0u;
}

// The ends of the bins a thread owns (its share of the offsets), read
// once; write() lets the thread owning the bin that holds sorted rank r
// write that bin.  The ends are non-decreasing, so the bin is the first
// one whose end exceeds r: b0 + the number of ends <= r (no early exit,
// no per-bin branch).
template <int NT>
struct OwnBins {
    static constexpr int per = kLinHist / NT;
    int b0;
    uint32_t lo;
    uint32_t e[per];
    __device__ __forceinline__ explicit OwnBins(const RowScratch* sh)
    {
        b0 = threadIdx.x * per;
        lo = bin_start(sh, b0);
#pragma unroll
        for (int q = 0; q < per; ++q) e[q] = sh->off[b0 + q];
    }
    __device__ __forceinline__ void write(uint32_t r, uint32_t* dst) const
    {
        if (r >= lo && r < e[per - 1]) {
            uint32_t c = 0;
#pragma unroll
            for (int q = 0; q < per - 1; ++q) c += e[q] <= r;
            *dst = b0 + c;
        }
    }
};

template <int NT>
__device__ __forceinline__ void locate_write(uint32_t r, uint32_t* dst, const RowScratch* sh)
{
    OwnBins<NT>(sh).write(r, dst);
}

// Rank counting inside a segment of keys: thread t (< cnt) takes key
// seg[t] and counts the keys below it and equal to it; the thread at
// sorted position `rank` writes its key and, when asked, those counts.
__device__ __forceinline__ void seg_rank_write(const uint32_t* seg, uint32_t cnt, uint32_t rank, uint32_t t,
                                               uint32_t* dst, uint32_t* less = nullptr, uint32_t* eq = nullptr)
{
    if (t < cnt) {
        const uint32_t k = seg[t];
        uint32_t lt = 0, et = 0, eb = 0;   // below; equal; equal and before t
#pragma unroll 4
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t o = seg[j];
            lt += o < k;
            et += o == k;
            eb += (o == k) & (j < t);
        }
        if (lt + eb == rank) {
            *dst = k;
            if (less) {
                *less = lt;
                *eq = et;
            }
        }
    }
}

// Exact rank-th key of a bin larger than a block (radix passes) and the
// numbers of keys below and equal to it (o: keys of the bins before).
template <int NT>
__device__ __forceinline__ uint32_t big_select(const uint32_t* seg, uint32_t cnt, uint32_t rank, uint32_t o,
                                               RowScratch* sh, uint32_t& less, uint32_t& eq)
{
    const uint32_t key = seg_select<NT>(seg, cnt, rank, sh);
    uint32_t c[2] = {0, 0};
    for (uint32_t j = threadIdx.x; j < cnt; j += NT) {
        c[0] += seg[j] < key;
        c[1] += seg[j] == key;
    }
    block_sums<NT, 2>(c, sh);
    less = o + c[0];
    eq = c[1];
    return key;
}

// Sorts a bin of at most kSortSeg keys into sh->mkey (with the counts of
// keys below and equal to each in mlt/met): thread t places seg[t] at
// its sorted position (ties in index order).
__device__ __forceinline__ void seg_sort_write(const uint32_t* seg, uint32_t cnt, RowScratch* sh)
{
    const uint32_t t = threadIdx.x;
    if (t < cnt) {
        const uint32_t k = seg[t];
        uint32_t lt = 0, et = 0, eb = 0;
#pragma unroll 4
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t o = seg[j];
            lt += o < k;
            et += o == k;
            eb += (o == k) & (j < t);
        }
        sh->mkey[lt + eb] = k;
        sh->mlt[lt + eb] = lt;
        sh->met[lt + eb] = et;
    }
}

// rank-th smallest of the bin-sorted keys through selection slot q:
// two barriers (more only for a bin holding more than a block of keys).
// less/eq return how many keys lie below the selected one and equal it.
template <int NT>
__device__ __forceinline__ uint32_t select_slot(uint32_t rank, int q, RowScratch* sh, uint32_t& less, uint32_t& eq)
{
    locate_write<NT>(rank, &sh->sbin[q], sh);
    lds_barrier();
    const uint32_t b = sh->sbin[q];
    const uint32_t o = bin_start(sh, b), cnt = sh->off[b] - o;
    const uint32_t* seg = row_keys(sh) + o;
    if (cnt > (uint32_t)NT) return big_select<NT>(seg, cnt, rank - o, o, sh, less, eq);
    seg_rank_write(seg, cnt, rank - o, threadIdx.x, &sh->skey[q], &sh->sless[q], &sh->seq[q]);
    lds_barrier();
    less = o + sh->sless[q];
    eq = sh->seq[q];
    return sh->skey[q];
}

// restore_baseline's window (OmnibusSigProc.cxx:867-906): the numbers of
// non-zero samples x with !(|fl(x - m1)| < 500) below and above m1.  When
// the underflow and overflow bins are empty, every non-zero sample lies
// in the linear bins [lo, lo + kLinBins / inv_w) up to float rounding; if
// that span sits inside the window with a margin far above the rounding,
// both counts are 0 without a pass over the row.
template <int NT, int VPT>
__device__ __forceinline__ void far_counts(const float (&val)[VPT], float m1, LinBins lb, uint32_t nzero,
                                           RowScratch* sh, uint32_t (&c)[2])
{
    c[0] = c[1] = 0;
    // the nzero zero slots hold +inf and sit in the overflow bin
    if (lb.inv_w > 0.f && sh->off[0] == 0 && sh->off[kLinBins + 1] - sh->off[kLinBins] == nzero) {
        const double lo = lb.lo, hi = (double)lb.lo + (double)kLinBins / (double)lb.inv_w;
        const double m = m1, margin = 1e-3 * (fabs(lo) + fabs(hi) + fabs(m) + 500.0);
        if (lo - m > -500.0 + margin && hi - m < 500.0 - margin) return;
    }
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
        if (!(fabsf(val[i] - m1) < 500.0f)) {
            if (val[i] < m1) ++c[0]; else ++c[1];
        }
    }
    block_sums_slot<NT, 2>(c, sh->upart[2]);
    c[1] -= nzero;
}

// Fast path: restore_baseline (if restore) and cal_RMS over the n_valid
// samples of the row; the register slots past them hold zeros.  After the
// moments every zero slot (padding or a zero sample) holds +inf: it sorts
// above every non-zero sample, so no rank below n_nz moves, it falls out
// of the RMS window and of the clipped moments by itself, and the counts
// that see it (the +-500 window, the RMS count of zero samples) are
// corrected by their known number; no pass needs a per-sample validity
// test.  The zeros are restored at the end.  The non-zero samples are
// counting-sorted into LDS by linear bin once; every order statistic is
// then a short exact select inside one bin.  Returns the RMS (0 when
// want_rms is false) and leaves the baselined samples in val.
template <int NT, int VPT>
__device__ __forceinline__ float row_stats_fast(float (&val)[VPT], uint32_t n_valid, bool restore, bool want_rms,
                                                RowScratch* sh)
{
    constexpr int NW = NT / 64;
    constexpr int per = kLinHist / NT;
    static_assert(kLinHist % NT == 0 && kRadixBins % NT == 0, "bins per thread");
    static_assert(NT >= 192, "the three percentile ranks run on three wavefronts");
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool stats = restore || want_rms;
    // moments of the non-zero samples (float: they only place the bins)
    uint32_t nn = 0;
    float s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
        nn += val[i] != 0.0f;
        s1 += val[i];
        s2 = fmaf(val[i], val[i], s2);
    }
    if (stats) {
        for (int q = 0; q < per; ++q) sh->off[threadIdx.x * per + q] = 0;
    }
    nn = wave_sum(nn);
    s1 = wave_sum_f(s1);
    s2 = wave_sum_f(s2);
    if (lane == 0) {
        sh->upart[0][0][wid] = nn;
        sh->fpart[0][wid] = s1;
        sh->fpart[1][wid] = s2;
    }
    lds_barrier();
    uint32_t n_nz = 0;
    float S1 = 0.f, S2 = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        n_nz += sh->upart[0][0][w];
        S1 += sh->fpart[0][w];
        S2 += sh->fpart[1][w];
    }
    if (n_nz == 0 || !stats) return 0.0f;   // an all-zero row: RMS 0, nothing to subtract

    float mean = S1 / (float)n_nz;
    float sd = sqrtf(fmaxf(S2 / (float)n_nz - mean * mean, 0.0f));
    const float kInf = __int_as_float(0x7f800000);
    const uint32_t nzero = (uint32_t)(NT * VPT) - n_nz;   // zero slots, padding included
#pragma unroll
    for (int i = 0; i < VPT; ++i) val[i] = val[i] != 0.0f ? val[i] : kInf;
    {   // clipped moments: the samples within 3 sd of the mean
        const float lim = 3.0f * sd;
        float c1 = 0.f, c2 = 0.f;
        uint32_t cn = 0;
#pragma unroll
        for (int i = 0; i < VPT; ++i)
            if (fabsf(val[i] - mean) < lim) {
                c1 += val[i];
                c2 = fmaf(val[i], val[i], c2);
                ++cn;
            }
        cn = wave_sum(cn);
        c1 = wave_sum_f(c1);
        c2 = wave_sum_f(c2);
        if (lane == 0) {
            sh->upart[1][0][wid] = cn;
            sh->fpart[2][wid] = c1;
            sh->fpart[3][wid] = c2;
        }
        lds_barrier();
        uint32_t CN = 0;
        float C1 = 0.f, C2 = 0.f;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            CN += sh->upart[1][0][w];
            C1 += sh->fpart[2][w];
            C2 += sh->fpart[3][w];
        }
        if (CN >= 2) {
            const float cm = C1 / (float)CN;
            const float cs = sqrtf(fmaxf(C2 / (float)CN - cm * cm, 0.0f));
            if (cs > 0.f) {
                mean = cm;
                sd = cs;
            }
        }
    }
    LinBins lb;
    lb.lo = mean - 8.0f * sd;
    lb.inv_w = sd > 0.f ? (float)kLinBins / (16.0f * sd) : 0.f;
    if (!isfinite(lb.lo) || !isfinite(lb.inv_w)) {
        lb.lo = 0.f;
        lb.inv_w = 0.f;
    }
    // counting sort of the non-zero keys by bin
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        atomicAdd(&sh->off[lin_bin(val[i], lb)], 1u);
    lds_barrier();
    {
        uint32_t* my = sh->off + threadIdx.x * per;
        uint32_t local = 0;
        for (int q = 0; q < per; ++q) local += my[q];
        const uint32_t incl = wave_incl_scan(local);
        if (lane == 63) sh->wsum[wid] = incl;
        lds_barrier();
        uint32_t base = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) base += (w < wid) ? sh->wsum[w] : 0u;
        uint32_t run = base + incl - local;
        for (int q = 0; q < per; ++q) {
            const uint32_t c = my[q];
            my[q] = run;
            run += c;
        }
        lds_barrier();
    }
    // scatter: the offsets advance to the ends of their bins
    uint32_t* keys = row_keys(sh);
#pragma unroll
    for (int i = 0; i < VPT; ++i)
        keys[atomicAdd(&sh->off[lin_bin(val[i], lb)], 1u)] = f2key(val[i]);
    lds_barrier();

    float b = 0.0f;
    uint32_t b_less = 0, b_eq = 0;   // non-zero samples below b and equal to b
    // the first median's bin, kept sorted when it holds at most a block
    uint32_t o1 = 0, n1 = 0;
    bool sorted1 = false;
    const float ps[3] = {(float)(0.5 - 0.34), 0.5f, (float)(0.5 + 0.34)};
    float par[3] = {0.0f, 0.0f, 0.0f};   // the 16/50/84% points of the baselined row
    if (restore && want_rms && n_nz == n_valid) {
        // No zero in the row: the percentile ranks do not depend on the
        // baseline (the r-th smallest of fl(z - b) is fl(z_(r) - b)), so
        // the median and the 16/84% points are selected in one round.
        const uint32_t r1 = pct_index(0.5f, n_nz);
        const uint32_t rq[2] = {pct_index(ps[0], n_nz), pct_index(ps[2], n_nz)};
        {
            const OwnBins<NT> own(sh);
            own.write(r1, &sh->sbin[0]);
            own.write(rq[0], &sh->sbin[2]);
            own.write(rq[1], &sh->sbin[3]);
        }
        lds_barrier();
        const uint32_t B1 = sh->sbin[0];
        o1 = bin_start(sh, B1);
        n1 = sh->off[B1] - o1;
        uint32_t oq[2], nq[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t bq = sh->sbin[2 + q];
            oq[q] = bin_start(sh, bq);
            nq[q] = sh->off[bq] - oq[q];
        }
        uint32_t kq[2];
        if (n1 <= 64u && nq[0] <= 64u && nq[1] <= 64u) {
            // the median's bin sorted on wavefront 0, the 16/84% ranks on 1 and 2
            seg_sort_write(keys + o1, n1, sh);
#pragma unroll
            for (int q = 0; q < 2; ++q)
                seg_rank_write(keys + oq[q], nq[q], rq[q] - oq[q], threadIdx.x - 64u * (q + 1), &sh->skey[2 + q]);
            lds_barrier();
            sorted1 = true;
#pragma unroll
            for (int q = 0; q < 2; ++q) kq[q] = sh->skey[2 + q];
        } else {   // wider bins: one selection at a time over the block
            sorted1 = n1 <= (uint32_t)NT;
            if (sorted1) {
                seg_sort_write(keys + o1, n1, sh);
                lds_barrier();
            }
#pragma unroll 1
            for (int q = 0; q < 2; ++q) {
                if (nq[q] > (uint32_t)NT) {
                    kq[q] = seg_select<NT>(keys + oq[q], nq[q], rq[q] - oq[q], sh);
                } else {
                    seg_rank_write(keys + oq[q], nq[q], rq[q] - oq[q], threadIdx.x, &sh->skey[2 + q]);
                    lds_barrier();
                    kq[q] = sh->skey[2 + q];
                }
            }
        }
        float m1;
        uint32_t l1, e1;
        if (sorted1) {
            m1 = key2f(sh->mkey[r1 - o1]);
            l1 = o1 + sh->mlt[r1 - o1];
            e1 = sh->met[r1 - o1];
        } else {
            m1 = key2f(big_select<NT>(keys + o1, n1, r1 - o1, o1, sh, l1, e1));
        }
        uint32_t c[2];
        far_counts<NT, VPT>(val, m1, lb, nzero, sh, c);
        const uint32_t r2 = c[0] + pct_index(0.5f, n_nz - c[0] - c[1]);
        if (c[0] + c[1] == 0 || r2 == r1) {
            b = m1;   // the second median is the first
            b_less = l1;
            b_eq = e1;
        } else if (sorted1 && r2 - o1 < n1) {
            b = key2f(sh->mkey[r2 - o1]);   // inside the sorted bin
            b_less = o1 + sh->mlt[r2 - o1];
            b_eq = sh->met[r2 - o1];
        } else {
            b = key2f(select_slot<NT>(r2, 1, sh, b_less, b_eq));
        }
        par[0] = key2f(kq[0]) - b;
        par[1] = m1 - b;
        par[2] = key2f(kq[1]) - b;
    } else {
        if (restore) {
            const uint32_t r1 = pct_index(0.5f, n_nz);
            locate_write<NT>(r1, &sh->sbin[0], sh);
            lds_barrier();
            const uint32_t B1 = sh->sbin[0];
            o1 = bin_start(sh, B1);
            n1 = sh->off[B1] - o1;
            sorted1 = n1 <= (uint32_t)NT;
            float m1;
            uint32_t l1, e1;
            if (sorted1) {
                seg_sort_write(keys + o1, n1, sh);
                lds_barrier();
                m1 = key2f(sh->mkey[r1 - o1]);
                l1 = o1 + sh->mlt[r1 - o1];
                e1 = sh->met[r1 - o1];
            } else {
                m1 = key2f(big_select<NT>(keys + o1, n1, r1 - o1, o1, sh, l1, e1));
            }
            uint32_t c[2];
            far_counts<NT, VPT>(val, m1, lb, nzero, sh, c);
            const uint32_t r2 = c[0] + pct_index(0.5f, n_nz - c[0] - c[1]);
            if (c[0] + c[1] == 0 || r2 == r1) {
                b = m1;   // the second median is the first
                b_less = l1;
                b_eq = e1;
            } else if (sorted1 && r2 - o1 < n1) {
                b = key2f(sh->mkey[r2 - o1]);   // inside the sorted bin
                b_less = o1 + sh->mlt[r2 - o1];
                b_eq = sh->met[r2 - o1];
            } else {
                b = key2f(select_slot<NT>(r2, 1, sh, b_less, b_eq));
            }
        }
        if (want_rms) {
            // ranks of the 16/50/84% points among the baselined row.  For a
            // non-zero sample x, fl(x - b) < 0 exactly when x < b and == 0
            // exactly when x == b, so with b a sample of Z the counts are
            // those of its selection.
            uint32_t c[2] = {b_less, b_eq};
            if (!restore) {
                uint32_t neg[1] = {0};
#pragma unroll
                for (int i = 0; i < VPT; ++i) neg[0] += val[i] < 0.0f;
                block_sums_slot<NT, 1>(neg, sh->upart[3]);
                c[0] = neg[0];
            }
            const uint32_t c_neg = c[0], c_eq = c[1], n_zero = n_valid - n_nz;
            uint32_t sr[3];     // rank among the non-zero keys, or ~0u for a zero
            bool need[3], any = false;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const uint32_t r = pct_index(ps[q], n_valid);
                sr[q] = r < c_neg ? r : (r < c_neg + n_zero + c_eq ? ~0u : r - n_zero);
                par[q] = 0.0f;
                need[q] = false;
                if (sr[q] == ~0u) continue;
                if (sorted1 && sr[q] - o1 < n1) {
                    par[q] = key2f(sh->mkey[sr[q] - o1]) - b;   // inside the median's sorted bin
                } else {
                    need[q] = any = true;
                    locate_write<NT>(sr[q], &sh->sbin[2 + q], sh);
                }
            }
            if (any) {
                lds_barrier();
                uint32_t o[3] = {0, 0, 0}, cnt[3] = {0, 0, 0};
                bool big = false, wide = false;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (need[q]) {
                        const uint32_t bq = sh->sbin[2 + q];
                        o[q] = bin_start(sh, bq);
                        cnt[q] = sh->off[bq] - o[q];
                        if (cnt[q] > (uint32_t)NT) big = true;
                        if (cnt[q] > 64u) wide = true;
                    }
                }
                // bins of at most a wavefront of keys: statistic q on wavefront q
                const uint32_t t0 = threadIdx.x;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (need[q] && cnt[q] <= (uint32_t)NT)
                        seg_rank_write(keys + o[q], cnt[q], sr[q] - o[q], wide ? t0 : t0 - 64u * q, &sh->skey[2 + q]);
                }
                lds_barrier();
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    if (need[q] && cnt[q] <= (uint32_t)NT) par[q] = key2f(sh->skey[2 + q]) - b;
                if (big) {   // a bin larger than a block: radix passes, one statistic at a time
#pragma unroll 1
                    for (int q = 0; q < 3; ++q)
                        if (need[q] && cnt[q] > (uint32_t)NT)
                            par[q] = key2f(seg_select<NT>(keys + o[q], cnt[q], sr[q] - o[q], sh)) - b;
                }
            }
        }
    }
    float rms = 0.0f;
    if (want_rms) {
        const double d21 = (double)(par[2] - par[1]), d10 = (double)(par[1] - par[0]);
        const float r0 = (float)sqrt((d21 * d21 + d10 * d10) / 2.);
        const double lim = 5.0 * (double)r0;
        // For a float y, (double)|y| < lim exactly when |y| < L, L the
        // float nearest lim if that is not below it, else the next one up.
        float L = (float)lim;
        if ((double)L < lim) L = __uint_as_float(__float_as_uint(L) + 1u);
        // per-thread partial sums in FP32 (at most VPT terms), the block
        // sum in FP64
        float sf = 0.0f;
        uint32_t k = 0;
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const float y = val[i] - b;   // +inf for a zero slot: outside the window
            const bool in = fabsf(y) < L;
            const float yi = in ? y : 0.0f;
            sf = fmaf(yi, yi, sf);
            k += in;
        }
        double s = wave_sum_d((double)sf);
        k = wave_sum(k);
        if (lane == 0) {
            sh->dpart[wid] = s;
            sh->upart[4][0][wid] = k;
        }
        lds_barrier();
        double tot = 0;
        uint32_t cntk = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            tot += sh->dpart[w];
            cntk += sh->upart[4][0][w];
        }
        // the zero samples of the row (not the padding) are in the
        // window whenever it is not empty
        if (lim > 0.0) cntk += n_valid - n_nz;
        rms = cntk ? sqrtf((float)(tot / (double)cntk)) : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < VPT; ++i) val[i] = val[i] != kInf ? val[i] - b : 0.0f;   // b = 0 without restore
    return rms;
}

// ---- general path: filtered radix select (rows with excluded samples) --

// rank-th smallest key among the flagged samples, four 8-bit passes.
template <int NT, int VPT>
__device__ __forceinline__ uint32_t block_select(const float (&val)[VPT], const uint64_t fl, uint32_t rank,
                                                 RowScratch* sh)
{
    uint32_t prefix = 0, pmask = 0;
#pragma unroll 1
    for (int sh_ = 24; sh_ >= 0; sh_ -= 8) {
        for (int b = threadIdx.x; b < kRadixBins; b += NT) sh->hist[b] = 0;
        lds_barrier();
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if ((fl >> i) & 1ull) {
                const uint32_t k = f2key(val[i]);
                if ((k & pmask) == prefix) atomicAdd(&sh->hist[(k >> sh_) & 0xffu], 1u);
            }
        }
        lds_barrier();
        uint32_t bucket, sub;
        hist_locate<NT>(sh->hist, kRadixBins, rank, bucket, sub, sh);
        rank = sub;
        prefix |= bucket << sh_;
        pmask |= 0xffu << sh_;
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
