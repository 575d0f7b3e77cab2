// Kernels and launch templates of the fused deconvolution, shared by
// the translation units that instantiate them (decon_rows_*.hip,
// decon_wire_*.hip) and the host side of the plans (decon.hip).  The
// instantiations are spread over several files so the library builds
// in parallel.
#pragma once
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "gmem.hpp"
#include "rowstats.hpp"

struct wct_decon_kernel_s {
    int device;
    int nw, nt, kc;
    float2* K;          // kc x nw, k-major
};

namespace wct { struct DJob; }

// Jobs of equal length, launched together with that length's kernels.
struct wct_decon_group {
    int len;          // M (row passes) or nw (wire pass)
    long begin, end;  // rows or tiles
    int twcount;      // twiddle entries of the length's plan
    bool smooth5;
    bool bad;         // some job of the group has bad-channel ranges
    bool var;         // some job of the group has extra outputs (K4)
};

struct wct_decon_plan_s {
    wct::DJob* jobs = nullptr;   // device job table, grouped by row length M (row passes)
    wct::DJob* jobs_t = nullptr; // the same jobs grouped by wire count (wire pass)
    std::vector<wct_decon_group> rgroups, tgroups;
    int njobs = 0;
    long rows = 0, tiles = 0;
    int Mmax = 0, nwmax = 0, tk = 8;
    int twM = 0, twW = 0;        // largest twiddle tables (complex entries)
    int ncu = 0;                 // compute units of the device
    int row_pct = 100;           // persistent row grids: percent of the resident maximum
    int row_plan = 0;            // 1: compile-time row plans where compiled (kStaticRowLen, jobs carry
                                 // twL); 0: runtime plans (transient plans built outside plan_create)
    int r2c_blocks = 0;          // workgroups per CU of a compile-time-plan r2c pass (3, 4, 5)
    int c2r_blocks = 0;          // workgroups per CU of a compile-time-plan c2r pass (2, 3, 4)
    bool smooth5 = true;         // every length is 2^a 3^b 5^c
    bool spectrum = false;       // the third stage writes time spectra
};

// A captured sequence of plan runs, replayed with one launch.
struct wct_decon_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace wct {

constexpr int kNT = 256;

// An extra output of a job (K4): the same spectrum row through another
// per-bin time filter (rows flagged in alt_rows take g_alt), then
// optionally restore_baseline, cal_RMS and save_data's zeroing.
constexpr int kMaxVar = 6;
struct DVar {
    const float* g;
    const float* g_alt;
    const unsigned char* alt_rows;
    float* out;
    float* rms;
    int ld_out;
    int restore, save_mode;
};

// Device-side description of one plane job of a batch.
struct DJob {
    const float* adc;
    float* out;
    unsigned char* mask;
    float* rms;
    const int* bad_ptr;
    const int* bad_lo;
    const int* bad_hi;
    const float* lf1;
    const float* lf2;
    const float2* K;
    float2* S;
    const float2* twM;
    const float2* twL;    // twiddles of the odd-first schedule (compile-time row plans)
    const float2* wN;
    const float2* twW;
    long row0;
    long tile0;
    int ld_adc, ld_out, ld_mask;
    int nw, nt, M, kc, ldk;
    float th_factor;
    int restore, save_mode, adjoint;
    float2* spec;     // spectrum mode: full time spectrum rows (else null)
    int ld_spec;
    int nvar;         // extra outputs (K4)
    DVar var[kMaxVar];
    const float2* cf;     // per-channel correction table (or null)
    const int* cf_row;    // its row per wire, -1 for none
    int ld_cf;
    FftPlan planM;
    FftPlan planW;
};

__device__ __forceinline__ int find_job_rows(const DJob* jobs, int njobs, long r)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].row0 <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_job_tiles(const DJob* jobs, int njobs, long t)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool in_bad(const DJob& J, int w, int t, bool inclusive)
{
    if (!J.bad_ptr) return false;
    const auto bp = G(J.bad_ptr), blo = G(J.bad_lo), bhi = G(J.bad_hi);
    const int q0 = bp[w], q1 = bp[w + 1];
    for (int q = q0; q < q1; ++q) {
        const int lo = blo[q], hi = bhi[q];
        if (t >= lo && (inclusive ? t <= hi : t < hi)) return true;
    }
    return false;
}

// XCD-aware block order: consecutive logical tiles land on the same XCD
// (blocks are dispatched round-robin over the 8 XCDs), so tiles sharing
// cache lines share an L2.
__device__ __forceinline__ long xcd_remap(long b, long nblocks)
{
    constexpr int kXcd = 8;
    const long per = nblocks / kXcd;
    if (b >= per * kXcd) return b;
    return (b % kXcd) * per + b / kXcd;
}

// LDS layout of the row kernels: the row (padded_len(Mmax) complex), the
// twiddle table of the current length, then (K3) the statistics scratch.
// K1 stages the twiddle table in LDS after the row; K3 reads it from the
// (L1-resident) global table and instead needs room for the statistics
// scratch and one row of sorted keys, which alias the row buffer once the
// row is in registers.
struct RowLds {
    int buf_len;      // complex entries of the row buffer
    int tw_len;       // complex entries of the twiddle region
    size_t bytes() const { return (size_t)(buf_len + tw_len) * sizeof(float2); }
};

// nkeys: the row statistics' key capacity (every register slot of the
// row, padding included).
inline RowLds row_lds(int Mmax, int twmax, bool stats, int nkeys = 0)
{
    RowLds l;
    if (stats) {
        const size_t scratch = sizeof(RowScratch) + (size_t)std::max(2 * Mmax, nkeys) * sizeof(uint32_t);
        l.buf_len = std::max(padded_len(Mmax), (int)((scratch + sizeof(float2) - 1) / sizeof(float2)));
        l.tw_len = 0;
    } else {
        l.buf_len = padded_len(Mmax);
        l.tw_len = twmax;
    }
    return l;
}

// Row buffer layouts: the runtime plans keep a padded row (pidx); the
// compile-time row plans (MS > 0) an unpadded one (fft_lin).  Slot i of
// this thread is sample threadIdx.x + i NT.
template <int NT, int MS>
__device__ __forceinline__ int row_slot(int i)
{
    if constexpr (MS > 0)
        return (int)threadIdx.x + i * NT;
    else
        return pidx_slot<NT>(pidx(threadIdx.x), i);
}
template <int MS>
__device__ __forceinline__ int row_at(int k)
{
    if constexpr (MS > 0)
        return k;
    else
        return pidx(k);
}

// ---- K1: real FFT along ticks --------------------------------------
// Persistent: each block walks rows blockIdx.x, +gridDim.x, ...; the
// twiddles are read through the caches (LDS holds only the row, which
// leaves room for four workgroups per CU).
// Five workgroups per CU for rows of up to 12 slots per thread with the
// smooth radix set (a 96-register cap: a few spills for a quarter more
// resident rows than four); the larger instances keep every register.
constexpr int r2c_min_blocks(int VH, bool S5) { return S5 && VH <= 12 ? 5 : 1; }

// MS > 0: a compile-time plan for rows of exactly MS complex samples
// (fft_static: constant spans and twiddle offsets, no per-pass dispatch),
// at BLK workgroups per CU.
template <int NT, int VH, bool S5, int MS = 0, int BLK = r2c_min_blocks(VH, S5)>
__global__ __launch_bounds__(NT, BLK) void k_row_r2c(const DJob* __restrict__ jobs, int njobs,
                                                long rbeg, long rend, RowLds lay)
{
    constexpr int MMAX = NT * VH;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    (void)lay;
    // the next row's samples are fetched into registers while the current
    // row is transformed, so the loads of a row are all in flight at once
    float2 z[VH];
    auto fetch = [&](long rr) {
        const DJob& Jn = jobs[find_job_rows(jobs, njobs, rr)];
        const int wn = (int)(rr - Jn.row0);
        const auto src = G(reinterpret_cast<const float2*>(Jn.adc + (long)wn * Jn.ld_adc));
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            if (n < Jn.M) z[i] = src[n];
        }
    };
    if (rbeg + blockIdx.x < rend) fetch(rbeg + blockIdx.x);
    for (long r = rbeg + blockIdx.x; r < rend; r += gridDim.x) {
        const DJob& J = jobs[find_job_rows(jobs, njobs, r)];
        const int w = (int)(r - J.row0);
        const int M = J.M;

        const bool has_bad = J.bad_ptr && G(J.bad_ptr)[w + 1] > G(J.bad_ptr)[w];
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            if (n < M) {
                float2 v = z[i];
                if (has_bad) {   // load_data zeroes [lo, hi) of bad channels (:337-346)
                    if (in_bad(J, w, 2 * n, false)) v.x = 0.f;
                    if (in_bad(J, w, 2 * n + 1, false)) v.y = 0.f;
                }
                buf[row_slot<NT, MS>(i)] = v;
            }
        }
        if (r + gridDim.x < rend) fetch(r + gridDim.x);
        lds_barrier();
        if constexpr (MS > 0)
            fft_lin<float2, NT, MS>(buf, G(J.twL), -1, threadIdx.x);
        else
            fft_lds<float2, NT, MMAX, S5>(buf, J.planM, G(J.twM), -1, threadIdx.x);
        const auto dst = G(J.S + (long)w * J.ldk);
        for (int k = threadIdx.x; k < J.ldk; k += NT) {
            float2 X;
            X.x = 0.f;
            X.y = 0.f;
            if (k < J.kc) {
                const float2 Zk = buf[row_at<MS>(k == M ? 0 : k)];
                const float2 Zm = buf[row_at<MS>(k == 0 ? 0 : M - k)];
                // X = (Zk + conj Zm)/2 - i W^k (Zk - conj Zm)/2
                const float sr = Zk.x + Zm.x, si = Zk.y - Zm.y;
                const float dr = Zk.x - Zm.x, di = Zk.y + Zm.y;
                const float2 W = G(J.wN)[k];
                const float wr = W.x * dr - W.y * di, wi = W.x * di + W.y * dr;
                X.x = 0.5f * (sr + wi);
                X.y = 0.5f * (si - wr);
            }
            dst[k] = X;
        }
        lds_barrier();
    }
}

// K1 on a compile-time row plan of MS complex samples (6000-tick rows).
// Vector-memory counters retire in issue order and count stores too, so
// a global load issued after the next row's prefetch cannot be waited
// for without waiting for the prefetch.  Nothing in the row's transform
// therefore reads global memory: the plan's twiddle table (one table per
// length, shared by every job of the launch) is staged into LDS once per
// workgroup, and the r2c post-process twiddles W^k of the first kWReg NT
// bins are held in registers for the whole launch (bins past them, only
// when more than kWReg NT are kept, read the table in global memory).
// The row's job is found once, when its samples are fetched.
constexpr int kWReg = 4;

template <int MS>
constexpr int lin_tw_entries()
{
    return sf_twoff(MS, sf_npass(MS, true), true) + kGenericTabSize;
}

template <int NT, int MS, int BLK>
__global__ __launch_bounds__(NT, BLK) void k_row_r2c_lin(const DJob* __restrict__ jobs, int njobs,
                                                         long rbeg, long rend)
{
    constexpr int VH = (MS + NT - 1) / NT;
    constexpr int NTW = lin_tw_entries<MS>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    float2* tws = buf + MS;
    long r = rbeg + blockIdx.x;
    if (r >= rend) return;
    int jn = find_job_rows(jobs, njobs, r);
    {
        const auto src = G(jobs[jn].twL);
        for (int i = threadIdx.x; i < NTW; i += NT) tws[i] = src[i];
    }
    float2 wreg[kWReg];
    {
        const auto wN = G(jobs[jn].wN);   // M + 1 entries
#pragma unroll
        for (int i = 0; i < kWReg; ++i) {
            const int k = threadIdx.x + i * NT;
            wreg[i] = wN[k <= MS ? k : MS];
        }
    }
    float2 z[VH];
    auto fetch = [&](long rr, int j) {
        const DJob& Jn = jobs[j];
        const auto src = G(reinterpret_cast<const float2*>(Jn.adc + (rr - Jn.row0) * (long)Jn.ld_adc));
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            z[i] = src[n < MS ? n : MS - 1];   // in-row address; slots past the row are not stored
        }
    };
    fetch(r, jn);
    for (; r < rend; r += gridDim.x) {
        const DJob& J = jobs[jn];
        const int w = (int)(r - J.row0);
        const bool has_bad = J.bad_ptr && G(J.bad_ptr)[w + 1] > G(J.bad_ptr)[w];
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            if (n < MS) {
                float2 v = z[i];
                if (has_bad) {   // load_data zeroes [lo, hi) of bad channels (:337-346)
                    if (in_bad(J, w, 2 * n, false)) v.x = 0.f;
                    if (in_bad(J, w, 2 * n + 1, false)) v.y = 0.f;
                }
                buf[n] = v;
            }
        }
        const auto dst = G(J.S + (long)w * J.ldk);
        const int kc = J.kc, ldk = J.ldk;
        const auto wN = G(J.wN);
        const long rn = r + gridDim.x;
        if (rn < rend) {
            jn = find_job_rows(jobs, njobs, rn);
            fetch(rn, jn);
        }
        lds_barrier();
        fft_lin<float2, NT, MS>(buf, tws, -1, threadIdx.x);
        auto bin = [&](int k, float2 W) {
            float2 X;
            X.x = 0.f;
            X.y = 0.f;
            if (k < kc) {
                const float2 Zk = buf[k == MS ? 0 : k];
                const float2 Zm = buf[k == 0 ? 0 : MS - k];
                // X = (Zk + conj Zm)/2 - i W^k (Zk - conj Zm)/2
                const float sr = Zk.x + Zm.x, si = Zk.y - Zm.y;
                const float dr = Zk.x - Zm.x, di = Zk.y + Zm.y;
                const float wr = W.x * dr - W.y * di, wi = W.x * di + W.y * dr;
                X.x = 0.5f * (sr + wi);
                X.y = 0.5f * (si - wr);
            }
            dst[k] = X;
        };
#pragma unroll
        for (int i = 0; i < kWReg; ++i) {
            const int k = threadIdx.x + i * NT;
            if (k < ldk) bin(k, wreg[i]);
        }
        for (int k = threadIdx.x + kWReg * NT; k < ldk; k += NT) bin(k, wN[k < kc ? k : 0]);
        lds_barrier();
    }
}

// ---- K2: wire FFT, multiply by K, inverse wire FFT ------------------
template <int NT, int TK, int NWMAX, bool S5, int NWS>
__global__ __launch_bounds__(NT) void k_wire_pass(const DJob* __restrict__ jobs, int njobs,
                                                  long tbeg, long ntiles, int ldc)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    float2* twl = buf + (size_t)TK * ldc;
    const long t = tbeg + xcd_remap(blockIdx.x, ntiles);
    const DJob& J = jobs[find_job_tiles(jobs, njobs, t)];
    const int k0 = (int)(t - J.tile0) * TK;
    const int nw = NWS ? NWS : J.nw;
    const auto S = G(J.S);
    stage_twiddles(twl, G(J.twW), J.planW.twcount, threadIdx.x, NT);
    constexpr int NLD = ((NWS ? NWS : NWMAX) * TK + NT - 1) / NT;   // elements per thread, all loads in flight
    static_assert(NT % TK == 0 && (NT / TK) % 16 == 0, "wire tile layout");
    {
        float2 t[NLD];
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int idx = threadIdx.x + q * NT;
            if (idx < nw * TK) t[q] = S[(long)(idx / TK) * J.ldk + k0 + idx % TK];
        }
        if (J.cf_row) {   // channel-by-channel electronics correction, before the wire transform
            const int k = k0 + threadIdx.x % TK;
#pragma unroll
            for (int q = 0; q < NLD; ++q) {
                const int idx = threadIdx.x + q * NT;
                if (idx < nw * TK && k < J.kc) {
                    const int ri = G(J.cf_row)[idx / TK];
                    if (ri >= 0) t[q] = cmul(t[q], G(J.cf)[(long)ri * J.ld_cf + k]);
                }
            }
        }
        // idx / TK = threadIdx.x / TK + q NT/TK and idx % TK = threadIdx.x % TK
        const int lbase = (threadIdx.x % TK) * ldc + pidx(threadIdx.x / TK);
#pragma unroll
        for (int q = 0; q < NLD; ++q) {
            const int idx = threadIdx.x + q * NT;
            if (idx < nw * TK) buf[lbase + q * ((NT / TK) + (NT / TK) / 16)] = t[q];
        }
    }
    lds_barrier();
    constexpr int SL = NT / TK;
    const int slice = threadIdx.x / SL, st = threadIdx.x % SL;
    float2* col = buf + slice * ldc;
    if constexpr (NWS > 0)
        fft_static<float2, SL, NWS>(col, twl, -1, st);
    else
        fft_lds<float2, SL, NWMAX, S5>(col, J.planW, twl, -1, st);
    const int k = k0 + slice;
    const auto Kk = G(J.K + (long)k * nw);
    for (int m = st; m < nw; m += SL) {
        float2 o;
        o.x = 0.f;
        o.y = 0.f;
        if (k < J.kc) {
            float2 kk = Kk[m];
            if (J.adjoint) kk.y = -kk.y;   // adjoint of a real circular filter: conj(K)
            o = cmul(col[pidx(m)], kk);
        }
        col[pidx(m)] = o;
    }
    lds_barrier();
    if constexpr (NWS > 0)
        fft_static<float2, SL, NWS>(col, twl, +1, st);
    else
        fft_lds<float2, SL, NWMAX, S5>(col, J.planW, twl, +1, st);
    for (int idx = threadIdx.x; idx < nw * TK; idx += NT) {
        const int cc = idx % TK, w = idx / TK;
        S[(long)w * J.ldk + k0 + cc] = buf[cc * ldc + pidx(w)];
    }
}

// All threads load their Y[k], Y[M-k] pairs first (every load in
// flight), then build Z' in LDS.
template <int NT, int VH, int MS = 0, typename YP, typename GP>
__device__ __forceinline__ void c2r_stage(const DJob& J, YP Y, GP g, float2* buf)
{
    const int M = J.M, kc = J.kc;
    float2 a[VH], b[VH];
    float ga[VH], gb[VH];
#pragma unroll
    for (int i = 0; i < VH; ++i) {
        const int k = threadIdx.x + i * NT;
        const int km = M - k;
        a[i].x = a[i].y = b[i].x = b[i].y = 0.f;
        ga[i] = gb[i] = 1.f;
        if (k < M) {
            if (k < kc) { a[i] = Y[k]; if (g) ga[i] = g[k]; }
            if (km < kc) { b[i] = Y[km]; if (g) gb[i] = g[km]; }
        }
    }
#pragma unroll
    for (int i = 0; i < VH; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < M) {
            float2 av = a[i], bv = b[i];
            av.x *= ga[i]; av.y *= ga[i];
            bv.x *= gb[i]; bv.y *= gb[i];
            if (k == 0) av.y = 0.f;
            if (M - k == M) bv.y = 0.f;
            const float sr = av.x + bv.x, si = av.y - bv.y;
            const float dr = av.x - bv.x, di = av.y + bv.y;
            const float2 W = J.wN[k];   // generic, as the c2r pass's other pointers
            const float pr = W.x * dr + W.y * di, pi = W.x * di - W.y * dr;
            float2 zz;
            zz.x = sr - pi;
            zz.y = si + pr;
            buf[row_slot<NT, MS>(i)] = zz;
        }
    }
}

// ---- K3: inverse real FFT, baseline, RMS, threshold, outputs ---------
// BAD: the group has bad-channel ranges (rows with excluded samples take
// the general statistics path); without, that path is not compiled in,
// which keeps the kernel within the registers of three workgroups per CU.
// Unlike the other stages this kernel keeps its job-table pointers
// generic: typed as global memory (G()) their 64-bit bases move into
// scalar registers and the kernel spills more at the three-workgroup
// register cap (16 against 9 VGPRs, 44 against 40 SGPRs), which measured
// 12.7 against 12.2 ms per 150-APA step.
template <int NT, int VH, bool S5, bool BAD, int MS = 0, int BLK = 3>
__global__ __launch_bounds__(NT, BLK) void k_row_c2r_roi(const DJob* __restrict__ jobs, int njobs,
                                                    long rbeg, long rend, RowLds lay)
{
    constexpr int MMAX = NT * VH;
    constexpr int VPT = 2 * VH;            // real samples per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    RowScratch* sh = reinterpret_cast<RowScratch*>(buf);   // reused once the row is in registers
    for (long r = rbeg + blockIdx.x; r < rend; r += gridDim.x) {
        const DJob& J = jobs[find_job_rows(jobs, njobs, r)];
        const int w = (int)(r - J.row0);
        const int M = J.M;
        const float2* twl = J.twM;
        const float2* Y = J.S + (long)w * J.ldk;
        const bool has_bad = BAD && J.bad_ptr && J.bad_ptr[w + 1] > J.bad_ptr[w];

        c2r_stage<NT, VH, MS>(J, Y, J.lf1, buf);
        lds_barrier();
        if constexpr (MS > 0)
            fft_lin<float2, NT, MS>(buf, J.twL, +1, threadIdx.x);
        else
            fft_lds<float2, NT, MMAX, S5>(buf, J.planM, twl, +1, threadIdx.x);

        float val[VPT];
        uint64_t valid = 0;
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            float2 z;
            z.x = z.y = 0.f;
            if (n < M) {
                z = buf[row_slot<NT, MS>(i)];
                valid |= 3ull << (2 * i);
            }
            val[2 * i] = z.x;
            val[2 * i + 1] = z.y;
        }
        uint64_t pm = valid;
        const bool stats = J.restore || J.mask || J.rms;
        if (stats && has_bad) {
#pragma unroll
            for (int i = 0; i < VPT; ++i) {
                const int pos = 2 * (threadIdx.x + (i / 2) * NT) + (i & 1);
                if (((valid >> i) & 1ull) && in_bad(J, w, pos, true)) pm &= ~(1ull << i);
            }
        }
        float rms = 0.0f;
        if (stats) {
            // the statistics scratch overlays the row buffer: every
            // thread holds its samples before any scratch write
            lds_barrier();
            // the fast path needs every valid sample in the RMS set
            const bool row_bad = has_bad && (__syncthreads_or(pm != valid) != 0);
            if (!row_bad) {
                rms = row_stats_fast<NT, VPT>(val, 2u * (uint32_t)M, J.restore != 0, true, sh);
            } else if constexpr (BAD) {
                if (J.restore) row_restore_baseline<NT, VPT>(val, valid, sh);
                rms = row_cal_rms<NT, VPT>(val, pm, sh);
            }
        }
        const double th = (double)J.th_factor * (double)rms + 1.0;

        uint64_t m1 = 0, m2 = 0;   // bit i: sample i above threshold (tight, tighter)
#pragma unroll
        for (int i = 0; i < VPT; ++i)   // slots past the row hold 0 < th
            if ((!has_bad || ((pm >> i) & 1ull)) && (double)val[i] > th) m1 |= 1ull << i;

        float* o = J.out + (long)w * J.ld_out;
#pragma unroll
        for (int i = 0; i < VH; ++i) {
            const int n = threadIdx.x + i * NT;
            if (n < M) {
                float a = val[2 * i], b = val[2 * i + 1];
                if (J.save_mode) {   // save_data (:383-418): negatives and [lo, hi) bad ranges to 0
                    a = a > 0.f ? a : 0.f;
                    b = b > 0.f ? b : 0.f;
                    if (has_bad) {
                        if (in_bad(J, w, 2 * n, false)) a = 0.f;
                        if (in_bad(J, w, 2 * n + 1, false)) b = 0.f;
                    }
                }
                float2 ab;
                ab.x = a;
                ab.y = b;
                *reinterpret_cast<float2*>(o + 2 * n) = ab;
            }
        }

        if (J.lf2) {   // the "tighter" variant for the mask (decon_2D_tighterROI, :1120-1160)
            lds_barrier();
            c2r_stage<NT, VH, MS>(J, Y, J.lf2, buf);
            lds_barrier();
            if constexpr (MS > 0)
                fft_lin<float2, NT, MS>(buf, J.twL, +1, threadIdx.x);
            else
                fft_lds<float2, NT, MMAX, S5>(buf, J.planM, twl, +1, threadIdx.x);
#pragma unroll
            for (int i = 0; i < VH; ++i) {
                const int n = threadIdx.x + i * NT;
                float2 z;
                z.x = z.y = 0.f;
                if (n < M) z = buf[row_slot<NT, MS>(i)];
                val[2 * i] = z.x;
                val[2 * i + 1] = z.y;
            }
            if (J.restore) {
                lds_barrier();   // samples in registers before the scratch is written
                row_stats_fast<NT, VPT>(val, 2u * (uint32_t)M, true, false, sh);
            }
#pragma unroll
            for (int i = 0; i < VPT; ++i)
                if ((!has_bad || ((pm >> i) & 1ull)) && (double)val[i] > th) m2 |= 1ull << i;
        }

        if (J.mask) {
            unsigned char* mk = J.mask + (long)w * J.ld_mask;
#pragma unroll
            for (int i = 0; i < VH; ++i) {
                const int n = threadIdx.x + i * NT;
                if (n < M) {
                    const unsigned a0 = (unsigned)((m1 >> (2 * i)) & 1ull) | ((unsigned)((m2 >> (2 * i)) & 1ull) << 1);
                    const unsigned a1 = (unsigned)((m1 >> (2 * i + 1)) & 1ull) | ((unsigned)((m2 >> (2 * i + 1)) & 1ull) << 1);
                    *reinterpret_cast<unsigned short*>(mk + 2 * n) = (unsigned short)(a0 | (a1 << 8));
                }
            }
        }
        if (J.rms && threadIdx.x == 0) J.rms[w] = rms;
        lds_barrier();
    }
}

// ---- K4: the extra outputs of a job (DVar) ----------------------------
// Per row and per extra output: the kept bins through the output's
// filter, the inverse real FFT, then restore_baseline / cal_RMS (with the
// job's bad ranges excluded) / save_data's zeroing as the output asks.
// A launch of its own after K3 (the spectrum rows are still in the
// workspace), so the registers of the tight pass are not shared with it;
// two workgroups per CU leave it every register it asks for.
template <int NT, int VH, bool S5, bool BAD>
__global__ __launch_bounds__(NT, 2) void k_row_c2r_var(const DJob* __restrict__ jobs, int njobs,
                                                    long rbeg, long rend, RowLds lay)
{
    constexpr int MMAX = NT * VH;
    constexpr int VPT = 2 * VH;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    RowScratch* sh = reinterpret_cast<RowScratch*>(buf);
    for (long r = rbeg + blockIdx.x; r < rend; r += gridDim.x) {
        const DJob& J = jobs[find_job_rows(jobs, njobs, r)];
        if (J.nvar == 0) continue;   // uniform over the block
        const int w = (int)(r - J.row0);
        const int M = J.M;
        const auto Y = G(J.S + (long)w * J.ldk);
        const bool has_bad = BAD && J.bad_ptr && G(J.bad_ptr)[w + 1] > G(J.bad_ptr)[w];
        for (int v = 0; v < J.nvar; ++v) {
            const DVar& V = J.var[v];
            const auto g = G((V.alt_rows && G(V.alt_rows)[w]) ? V.g_alt : V.g);
            c2r_stage<NT, VH>(J, Y, g, buf);
            lds_barrier();
            fft_lds<float2, NT, MMAX, S5>(buf, J.planM, G(J.twM), +1, threadIdx.x);
            float val[VPT];
            uint64_t valid = 0;
#pragma unroll
            for (int i = 0; i < VH; ++i) {
                const int n = threadIdx.x + i * NT;
                float2 z;
                z.x = z.y = 0.f;
                if (n < M) {
                    z = buf[pidx_slot<NT>(pidx(threadIdx.x), i)];
                    valid |= 3ull << (2 * i);
                }
                val[2 * i] = z.x;
                val[2 * i + 1] = z.y;
            }
            const bool want_rms = V.rms != nullptr;
            uint64_t pm = valid;
            if (want_rms && has_bad) {
#pragma unroll
                for (int i = 0; i < VPT; ++i) {
                    const int pos = 2 * (threadIdx.x + (i / 2) * NT) + (i & 1);
                    if (((valid >> i) & 1ull) && in_bad(J, w, pos, true)) pm &= ~(1ull << i);
                }
            }
            float rms = 0.0f;
            // the statistics scratch overlays the row buffer
            lds_barrier();
            if (V.restore || want_rms) {
                const bool row_bad = want_rms && has_bad && (__syncthreads_or(pm != valid) != 0);
                if (!row_bad) {
                    rms = row_stats_fast<NT, VPT>(val, 2u * (uint32_t)M, V.restore != 0, want_rms, sh);
                } else if constexpr (BAD) {
                    if (V.restore) row_restore_baseline<NT, VPT>(val, valid, sh);
                    rms = row_cal_rms<NT, VPT>(val, pm, sh);
                }
            }
            if (V.out) {
                const auto o = G(V.out + (long)w * V.ld_out);
#pragma unroll
                for (int i = 0; i < VH; ++i) {
                    const int n = threadIdx.x + i * NT;
                    if (n < M) {
                        float a = val[2 * i], b = val[2 * i + 1];
                        if (V.save_mode) {   // save_data (:383-418)
                            a = a > 0.f ? a : 0.f;
                            b = b > 0.f ? b : 0.f;
                            if (has_bad) {
                                if (in_bad(J, w, 2 * n, false)) a = 0.f;
                                if (in_bad(J, w, 2 * n + 1, false)) b = 0.f;
                            }
                        }
                        float2 ab;
                        ab.x = a;
                        ab.y = b;
                        *reinterpret_cast<gptr<float2>>(o + 2 * n) = ab;
                    }
                }
            }
            if (V.rms && threadIdx.x == 0) G(V.rms)[w] = rms;
            lds_barrier();
        }
    }
}

// Persistent grid: as many blocks as the device keeps resident at once
// (a larger grid leaves a tail of blocks that start only as others end).
template <typename K>
static int resident_grid(K kernel, int nt, size_t lds, int ncu, long work, int pct, int* grid)
{
    int per_cu = 0;
    WCT_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, lds));
    if (per_cu < 1) per_cu = 1;
    const long full = std::max<long>(1, (long)per_cu * ncu * pct / 100);
    *grid = (int)std::max<long>(1, std::min<long>(work, full));
    return WCT_OK;
}

inline bool is_smooth5(int n)
{
    for (int q : {2, 3, 5})
        while (n % q == 0) n /= q;
    return n == 1;
}

// Wire counts with wire-pass kernels compiled for that exact count (the
// compile-time FFT plan saves the pass bookkeeping of the column
// transforms); other counts take the runtime-plan kernel.
template <typename F>
static int dispatch_static_nw(int nw, int tk, F&& f)
{
    if (tk == 8 || tk == 4 || tk == 2) {
        if (nw == 800) return f(std::integral_constant<int, 800>());
        if (nw == 960) return f(std::integral_constant<int, 960>());
    }
    return f(std::integral_constant<int, 0>());
}

template <int MS, int BLK>
static int launch_r2c_lin(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    const size_t lds = (size_t)(MS + lin_tw_entries<MS>()) * sizeof(float2);
    int rc = allow_lds(k_row_r2c_lin<kNT, MS, BLK>, lds);
    if (rc) return rc;
    int grid = 0;
    rc = resident_grid(k_row_r2c_lin<kNT, MS, BLK>, kNT, lds, p->ncu, g.end - g.begin, p->row_pct, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_row_r2c_lin<kNT, MS, BLK>), dim3(grid), dim3(kNT), lds, s, p->jobs, p->njobs, g.begin,
                       g.end);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <int VH, bool S5, int MS, int BLK>
static int launch_r2c_inst(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    const RowLds lay = row_lds(g.len, 0, false);   // twiddles read through the caches
    const size_t lds = lay.bytes();
    int rc = allow_lds(k_row_r2c<kNT, VH, S5, MS, BLK>, lds);
    if (rc) return rc;
    int grid = 0;
    rc = resident_grid(k_row_r2c<kNT, VH, S5, MS, BLK>, kNT, lds, p->ncu, g.end - g.begin, p->row_pct, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_row_r2c<kNT, VH, S5, MS, BLK>), dim3(grid), dim3(kNT), lds, s, p->jobs, p->njobs,
                       g.begin, g.end, lay);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

// Lengths with compile-time row plans (the 6000-tick rows), and whether
// a group runs one under the plan's tuning.
constexpr int kStaticRowLen = 3000;
inline bool use_static_row(wct_decon_plan p, const wct_decon_group& g, int vh, bool s5)
{
    return p->row_plan == 1 && vh == 12 && s5 && g.len == kStaticRowLen;
}

template <int VH, bool S5>
static int launch_rows_r2c(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    if constexpr (VH == 12 && S5) {
        if (use_static_row(p, g, VH, S5)) {
            switch (p->r2c_blocks) {
            case 3: return launch_r2c_lin<kStaticRowLen, 3>(p, g, s);
            case 5: return launch_r2c_inst<VH, S5, kStaticRowLen, 5>(p, g, s);
            default: return launch_r2c_inst<VH, S5, kStaticRowLen, 4>(p, g, s);
            }
        }
    }
    return launch_r2c_inst<VH, S5, 0, r2c_min_blocks(VH, S5)>(p, g, s);
}

template <int VH, bool S5, bool BAD, int MS, int BLK>
static int launch_c2r_inst(wct_decon_plan p, const wct_decon_group& g, hipStream_t s, const RowLds& lay)
{
    const size_t lds = lay.bytes();
    int rc = allow_lds(k_row_c2r_roi<kNT, VH, S5, BAD, MS, BLK>, lds);
    if (rc) return rc;
    int grid = 0;
    rc = resident_grid(k_row_c2r_roi<kNT, VH, S5, BAD, MS, BLK>, kNT, lds, p->ncu, g.end - g.begin, p->row_pct,
                       &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_row_c2r_roi<kNT, VH, S5, BAD, MS, BLK>), dim3(grid), dim3(kNT), lds, s, p->jobs,
                       p->njobs, g.begin, g.end, lay);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <int VH, bool S5, bool BAD>
static int launch_rows_c2r(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    const RowLds lay = row_lds(g.len, g.twcount, true, 2 * kNT * VH);
    const size_t lds = lay.bytes();
    int rc = WCT_OK;
    bool done = false;
    if constexpr (VH == 12 && S5) {
        if (use_static_row(p, g, VH, S5)) {
            switch (p->c2r_blocks) {
            case 2: rc = launch_c2r_inst<VH, S5, BAD, kStaticRowLen, 2>(p, g, s, lay); break;
            case 4: rc = launch_c2r_inst<VH, S5, BAD, kStaticRowLen, 4>(p, g, s, lay); break;
            default: rc = launch_c2r_inst<VH, S5, BAD, kStaticRowLen, 3>(p, g, s, lay); break;
            }
            done = true;
        }
    }
    if (!done) rc = launch_c2r_inst<VH, S5, BAD, 0, 3>(p, g, s, lay);
    if (rc) return rc;
    if (!g.var) return WCT_OK;
    // the jobs' extra outputs, after the tight pass
    rc = allow_lds(k_row_c2r_var<kNT, VH, S5, BAD>, lds);
    if (rc) return rc;
    int grid = 0;
    rc = resident_grid(k_row_c2r_var<kNT, VH, S5, BAD>, kNT, lds, p->ncu, g.end - g.begin, p->row_pct, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_row_c2r_var<kNT, VH, S5, BAD>), dim3(grid), dim3(kNT), lds, s, p->jobs, p->njobs,
                       g.begin, g.end, lay);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <int TK, int NWMAX, bool S5, int NWS>
static int launch_wire(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    const int ldc = padded_len(g.len) | 1;
    const size_t lds = ((size_t)TK * ldc + g.twcount) * sizeof(float2);
    int rc = allow_lds(k_wire_pass<kNT, TK, NWMAX, S5, NWS>, lds);
    if (rc) return rc;
    const long nt = g.end - g.begin;
    hipLaunchKernelGGL((k_wire_pass<kNT, TK, NWMAX, S5, NWS>), dim3(nt), dim3(kNT), lds, s, p->jobs_t, p->njobs,
                       g.begin, nt, ldc);
    WCT_HIP_CHECK(hipGetLastError());
    return WCT_OK;
}

template <bool S5, int VH>
static int row_pass(wct_decon_plan p, const wct_decon_group& g, hipStream_t s, bool c2r)
{
    if (!c2r) return launch_rows_r2c<VH, S5>(p, g, s);
    return g.bad ? launch_rows_c2r<VH, S5, true>(p, g, s) : launch_rows_c2r<VH, S5, false>(p, g, s);
}

// One row-pass group (r2c, or c2r with the statistics) at VH complex
// slots per thread.  Instantiated in decon_rows_*.hip.
template <int VH>
int row_pass_vh(wct_decon_plan p, const wct_decon_group& g, hipStream_t s, bool c2r)
{
    return g.smooth5 ? row_pass<true, VH>(p, g, s, c2r) : row_pass<false, VH>(p, g, s, c2r);
}

// One wire-pass group at TK bins per tile.  Instantiated in
// decon_wire_*.hip.
template <int TK>
int wire_pass_tk(wct_decon_plan p, const wct_decon_group& g, hipStream_t s)
{
    if constexpr (TK == 16) {
        return g.smooth5 ? launch_wire<16, 1024, true, 0>(p, g, s) : launch_wire<16, 1024, false, 0>(p, g, s);
    } else {
        constexpr int NWMAX = TK == 8 ? 1024 : 2048;
        return dispatch_static_nw(g.len, TK, [&](auto N) -> int {
            constexpr int NWS = decltype(N)::value;
            if (NWS > 0 || g.smooth5) return launch_wire<TK, NWMAX, true, NWS>(p, g, s);
            return launch_wire<TK, NWMAX, false, 0>(p, g, s);
        });
    }
}

#define WCT_ROW_VH(X) X(4) X(8) X(12) X(16) X(20) X(24) X(32)
#define WCT_WIRE_TK(X) X(2) X(4) X(8) X(16)
#define WCT_EXTERN_ROW(VH) extern template int row_pass_vh<VH>(wct_decon_plan, const wct_decon_group&, hipStream_t, bool);
#define WCT_EXTERN_WIRE(TK) extern template int wire_pass_tk<TK>(wct_decon_plan, const wct_decon_group&, hipStream_t);
WCT_ROW_VH(WCT_EXTERN_ROW)
WCT_WIRE_TK(WCT_EXTERN_WIRE)
#define WCT_INSTANTIATE_ROW(VH) template int row_pass_vh<VH>(wct_decon_plan, const wct_decon_group&, hipStream_t, bool);
#define WCT_INSTANTIATE_WIRE(TK) template int wire_pass_tk<TK>(wct_decon_plan, const wct_decon_group&, hipStream_t);

}  // namespace wct
