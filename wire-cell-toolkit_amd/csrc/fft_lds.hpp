// LDS-resident mixed-radix Stockham FFT for gfx950, shared by the IDFT
// kernels and the fused deconvolution kernels.
//
// One transform of length n lives in LDS (T2 = float2 or double2).  A
// group of NT threads (a whole workgroup or a slice of it) walks the
// passes; each pass loads every butterfly's inputs into registers,
// synchronises, then twiddles, runs the radix-R DFT and stores in
// autosorted (Stockham) order, so no bit reversal is ever needed.
//
// Pass p with radix R and span Ns (product of the earlier radices):
//   for j < n/R:  k = j mod Ns
//     v[r]  = x[j + r n/R] * w^{r k},  w = exp(sign 2 pi i / (Ns R))
//     v     = DFT_R(v)
//     y[(j-k) R + k + r Ns] = v[r]
// Twiddles come from a per-length table built on the host in double and
// rounded to T once (forward sign; conjugated for the inverse).
#pragma once
#include <hip/hip_runtime.h>

#include "sync.hpp"

namespace wct {

constexpr int kMaxPasses = 24;

// Radix schedule of one length.  Passed by value as a kernel argument.
struct FftPlan {
    int n;
    int npass;
    int rad[kMaxPasses];
    int twoff[kMaxPasses];   // offset (in complex entries) of each pass's twiddles
    int dftoff;              // offset of the generic small-DFT tables (7, 11, 13)
    int twcount;             // entries in the whole table (passes + generic tables)
};

// LDS index with one pad slot every 16 complex entries: the strided
// stores of the early passes (stride R entries) then spread over the
// banks instead of piling onto a few.
__host__ __device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }
__host__ __device__ __forceinline__ int padded_len(int n) { return n + (n >> 4) + 1; }

// pidx(t + i NT) for NT a multiple of 16 equals pidx(t) + i (NT + NT/16):
// a row walked with stride NT is one base address plus constant offsets
// (immediate offsets of the LDS instructions, no per-slot address register).
template <int NT>
__host__ __device__ __forceinline__ int pidx_slot(int pt, int i)
{
    static_assert(NT % 16 == 0, "stride must be a multiple of the pad period");
    return pt + i * (NT + NT / 16);
}

template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// i * sign * a
template <typename T2>
__device__ __forceinline__ T2 cisgn(T2 a, int sign)
{
    T2 r;
    if (sign > 0) { r.x = -a.y; r.y = a.x; }
    else { r.x = a.y; r.y = -a.x; }
    return r;
}

// ---- radix-R DFT kernels on registers ---------------------------------

template <typename T2, int R> struct Dft;

template <typename T2> struct Dft<T2, 2> {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int, TB)
    {
        T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T2> struct Dft<T2, 3> {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB)
    {
        using T = decltype(v[0].x);
        const T h = T(0.86602540378443864676);
        T2 s = cadd(v[1], v[2]);
        T2 d = csub(v[1], v[2]);
        T2 m;
        m.x = v[0].x - T(0.5) * s.x;
        m.y = v[0].y - T(0.5) * s.y;
        T2 rot = cisgn(d, sign);      // i*sign*d
        v[0] = cadd(v[0], s);
        v[1].x = m.x + h * rot.x; v[1].y = m.y + h * rot.y;
        v[2].x = m.x - h * rot.x; v[2].y = m.y - h * rot.y;
    }
};

template <typename T2> struct Dft<T2, 4> {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB)
    {
        T2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        T2 t2 = cadd(v[1], v[3]), t3 = cisgn(csub(v[1], v[3]), sign);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <typename T2> struct Dft<T2, 5> {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB)
    {
        using T = decltype(v[0].x);
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        T2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        T2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        T2 a1, a2, b1, b2;
        a1.x = v[0].x + c1 * t1.x + c2 * t2.x; a1.y = v[0].y + c1 * t1.y + c2 * t2.y;
        a2.x = v[0].x + c2 * t1.x + c1 * t2.x; a2.y = v[0].y + c2 * t1.y + c1 * t2.y;
        b1.x = s1 * t3.x + s2 * t4.x; b1.y = s1 * t3.y + s2 * t4.y;
        b2.x = s2 * t3.x - s1 * t4.x; b2.y = s2 * t3.y - s1 * t4.y;
        v[0] = cadd(v[0], cadd(t1, t2));
        T2 r1 = cisgn(b1, sign), r2 = cisgn(b2, sign);
        v[1] = cadd(a1, r1);
        v[4] = csub(a1, r1);
        v[2] = cadd(a2, r2);
        v[3] = csub(a2, r2);
    }
};

template <typename T2> struct Dft<T2, 8> {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB)
    {
        using T = decltype(v[0].x);
        const T r2 = T(0.70710678118654752440);
        T2 e[4] = {v[0], v[2], v[4], v[6]};
        T2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T2, 4>::run(e, sign, nullptr);
        Dft<T2, 4>::run(o, sign, nullptr);
        // w8^1 = (1 + i sign)/sqrt2, w8^2 = i sign, w8^3 = (-1 + i sign)/sqrt2
        T2 o1, o2, o3;
        {
            T2 is = cisgn(o[1], sign);
            o1.x = r2 * (o[1].x + is.x); o1.y = r2 * (o[1].y + is.y);
        }
        o2 = cisgn(o[2], sign);
        {
            T2 is = cisgn(o[3], sign);
            o3.x = r2 * (is.x - o[3].x); o3.y = r2 * (is.y - o[3].y);
        }
        v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
    }
};

// Prime-factor (Good-Thomas) butterfly of N = N1 N2 with N1, N2 coprime:
// no inner twiddles.  Input (N2 n1 + N1 n2) mod N is element n2 of the
// N2-point DFT n1; bin k1 of the N1-point DFT over n1 of those DFTs' bin
// k2 is output (N2 e1 k1 + N1 e2 k2) mod N, e1 = N2^-1 mod N1 and
// e2 = N1^-1 mod N2 (then W_N^(n k) = W_N1^(n1 k1) W_N2^(n2 k2)).
constexpr int inv_mod(int a, int m)
{
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 1;
}

template <typename T2, int N1, int N2> struct DftPfa {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB)
    {
        constexpr int N = N1 * N2;
        constexpr int o1 = N2 * inv_mod(N2 % N1, N1), o2 = N1 * inv_mod(N1 % N2, N2);
        T2 a[N1][N2];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) a[n1][n2] = v[(N2 * n1 + N1 * n2) % N];
            Dft<T2, N2>::run(a[n1], sign, nullptr);
        }
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            T2 b[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) b[n1] = a[n1][k2];
            Dft<T2, N1>::run(b, sign, nullptr);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) v[(o1 * k1 + o2 * k2) % N] = b[k1];
        }
    }
};
template <typename T2> struct Dft<T2, 10> : DftPfa<T2, 2, 5> {};
template <typename T2> struct Dft<T2, 15> : DftPfa<T2, 3, 5> {};
template <typename T2> struct Dft<T2, 20> : DftPfa<T2, 4, 5> {};

// Generic odd prime radix (7, 11, 13) from a small table of
// exp(-2 pi i q / R), q < R (forward sign).
template <typename T2, int R> struct DftGeneric {
    template <typename TB>
    static __device__ __forceinline__ void run(T2* v, int sign, TB tab)
    {
        T2 out[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            T2 acc = v[0];
#pragma unroll
            for (int j = 1; j < R; ++j) {
                T2 w = tab[(j * k) % R];
                if (sign > 0) w.y = -w.y;
                T2 p = cmul(v[j], w);
                acc = cadd(acc, p);
            }
            out[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = out[k];
    }
};
template <typename T2> struct Dft<T2, 7> : DftGeneric<T2, 7> {};
template <typename T2> struct Dft<T2, 11> : DftGeneric<T2, 11> {};
template <typename T2> struct Dft<T2, 13> : DftGeneric<T2, 13> {};

// Offset of radix R's small table inside the plan's generic-table block.
__host__ __device__ constexpr int generic_tab_offset(int R)
{
    return R == 7 ? 0 : (R == 11 ? 7 : (R == 13 ? 18 : 0));
}
constexpr int kGenericTabSize = 7 + 11 + 13;

// ---- one Stockham pass ------------------------------------------------

// LDS index of input r of butterfly j = tid + b NT: pidx(j + r m), written
// as pidx(tid + r m) plus a constant when NT is a multiple of 16.
template <int NT>
__device__ __forceinline__ int pass_read_idx(int tid, int b, int rm)
{
    if constexpr (NT % 16 == 0)
        return pidx_slot<NT>(pidx(tid + rm), b);
    else
        return pidx(tid + b * NT + rm);
}
// buf: LDS sequence of length n.  The group has NT threads, this thread
// is tid.  MAXB bounds the butterflies one thread owns in this pass.

template <typename T2, int R, int NT, int MAXB, typename TW>
__device__ __forceinline__ void stockham_pass(T2* buf, int n, int Ns, TW tw, TW gtab, int sign, int tid)
{
    const int m = n / R;
    T2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (j < m) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = buf[pass_read_idx<NT>(tid, b, r * m)];
        }
    }
    lds_barrier();
    // k = j mod Ns, stepped from the first butterfly (one division per pass)
    int k = tid % Ns;
    const int kstep = NT % Ns;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (j < m) {
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    T2 w = tw[k * (R - 1) + (r - 1)];
                    if (sign > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            Dft<T2, R>::run(v[b], sign, gtab + generic_tab_offset(R));
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pidx(base + r * Ns)] = v[b][r];
        }
        k += kstep;
        if (k >= Ns) k -= Ns;
    }
    lds_barrier();
}

// SMOOTH5 instantiates only the radix 2/3/4/5/8/10/15/20 passes (lengths whose
// prime factors are 2, 3 and 5); the generic 7/11/13 passes carry large
// constant tables that would inflate the register allocation of every
// kernel that includes them.
// Full transform of the LDS sequence buf (length plan.n <= NMAX, stored
// at padded indices pidx(i)).  tw is the plan's twiddle table, normally
// staged in LDS by the caller (stage_twiddles).  Every thread of the
// workgroup must call this (it contains barriers); when the workgroup is
// split into slices each slice passes its own buf and slice-local tid.
template <typename T2, int NT, int NMAX, bool SMOOTH5 = false, typename TW>
__device__ __forceinline__ void fft_lds(T2* buf, const FftPlan& pl, TW tw, int sign, int tid)
{
    const TW gtab = tw + pl.dftoff;
    int Ns = 1;
    for (int p = 0; p < pl.npass; ++p) {
        const int R = pl.rad[p];
        const TW t = tw + pl.twoff[p];
        switch (R) {
        case 2: stockham_pass<T2, 2, NT, (NMAX / 2 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 3: stockham_pass<T2, 3, NT, (NMAX / 3 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 4: stockham_pass<T2, 4, NT, (NMAX / 4 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 5: stockham_pass<T2, 5, NT, (NMAX / 5 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 10: stockham_pass<T2, 10, NT, (NMAX / 10 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 15: stockham_pass<T2, 15, NT, (NMAX / 15 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 20: stockham_pass<T2, 20, NT, (NMAX / 20 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 8: stockham_pass<T2, 8, NT, (NMAX / 8 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid); break;
        case 7:
            if (!SMOOTH5) stockham_pass<T2, 7, NT, (NMAX / 7 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid);
            break;
        case 11:
            if (!SMOOTH5) stockham_pass<T2, 11, NT, (NMAX / 11 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid);
            break;
        case 13:
            if (!SMOOTH5) stockham_pass<T2, 13, NT, (NMAX / 13 + NT - 1) / NT>(buf, pl.n, Ns, t, gtab, sign, tid);
            break;
        default: break;
        }
        Ns *= R;
    }
}

// Copies a plan's twiddle table to LDS (all threads of the block).
template <typename T2, typename TS>
__device__ __forceinline__ void stage_twiddles(T2* dst, TS src, int count, int tid,
                                               int nthreads)
{
    for (int i = tid; i < count; i += nthreads) dst[i] = src[i];
}

// ---- radix schedule ------------------------------------------------------
// The one radix rule, used by the host plans and the compile-time plans
// alike (so a twiddle table built for one serves the other).  The fewest
// passes come from pairing factors into the largest butterflies that
// stay cheap in registers: a 3 with a 5 (radix 15), each remaining 5 with
// one or two of the 2s (radix 10 or 20), the rest of the 2s in 8s plus a
// 4 or a 2.  The 5s take only as many 2s as save a pass, preferring
// radix 10 over 20 (fewer registers for the same count).  Order: 8s, the
// 4 or 2, 20s, 10s, 15s, 5s, 3s, 7s, 11s, 13s; with odd_first (the
// unpadded compile-time row plans) 15s, 5s and 3s come first, so that the
// first pass's stores, strided by an odd radix, hit distinct LDS banks.
// Returns 0 past the last pass (and for a length with a prime factor
// above 13).
constexpr int sf_radix(int n, int p, bool odd_first = false)
{
    if (n < 2) return 0;
    int a = 0, b = 0, c = 0, m = n;
    while (m % 2 == 0) { ++a; m /= 2; }
    while (m % 3 == 0) { ++b; m /= 3; }
    while (m % 5 == 0) { ++c; m /= 5; }
    int e7 = 0, e11 = 0, e13 = 0;
    while (m % 7 == 0) { ++e7; m /= 7; }
    while (m % 11 == 0) { ++e11; m /= 11; }
    while (m % 13 == 0) { ++e13; m /= 13; }
    if (m != 1) return 0;
    const int n15 = b < c ? b : c;
    const int c1 = c - n15;                       // 5s left after the 15s
    const int xmax = a < 2 * c1 ? a : 2 * c1;     // 2s the 5s could take
    const int lp = (a - xmax + 2) / 3;            // passes for the 2s left over
    const int x = a - 3 * lp > 0 ? a - 3 * lp : 0;   // 2s the 5s do take
    const int n20 = x > c1 ? x - c1 : 0;
    const int n10 = x - 2 * n20;
    const int n5 = c1 - n10 - n20;
    const int l2 = a - x;
    const int cnt[11] = {l2 / 3, l2 % 3 == 2 ? 1 : 0, l2 % 3 == 1 ? 1 : 0, n20, n10, n15, n5, b - n15, e7, e11, e13};
    const int rad[11] = {8, 4, 2, 20, 10, 15, 5, 3, 7, 11, 13};
    const int ord_even[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    const int ord_odd[11] = {5, 6, 7, 0, 1, 2, 3, 4, 8, 9, 10};
    int idx = p;
    for (int i = 0; i < 11; ++i) {
        const int q = odd_first ? ord_odd[i] : ord_even[i];
        if (idx < cnt[q]) return rad[q];
        idx -= cnt[q];
    }
    return 0;
}

constexpr int sf_npass(int n, bool odd_first = false)
{
    int p = 0;
    while (sf_radix(n, p, odd_first)) ++p;
    return p;
}

constexpr int sf_span(int n, int p, bool odd_first = false)
{
    int ns = 1;
    for (int q = 0; q < p; ++q) ns *= sf_radix(n, q, odd_first);
    return ns;
}

constexpr int sf_twoff(int n, int p, bool odd_first = false)
{
    int off = 0, ns = 1;
    for (int q = 0; q < p; ++q) {
        off += ns * (sf_radix(n, q, odd_first) - 1);
        ns *= sf_radix(n, q, odd_first);
    }
    return off;
}

// One Stockham pass with every size a constant.
template <typename T2, int R, int NT, int N, int NS, typename TW>
__device__ __forceinline__ void stockham_pass_c(T2* buf, TW tw, TW gtab, int sign, int tid)
{
    constexpr int m = N / R;
    constexpr int MAXB = (m + NT - 1) / NT;
    T2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (m % NT == 0 || j < m) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = buf[pass_read_idx<NT>(tid, b, r * m)];
        }
    }
    lds_barrier();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (m % NT == 0 || j < m) {
            const int k = NS == 1 ? 0 : j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    T2 w = tw[k * (R - 1) + (r - 1)];
                    if (sign > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            Dft<T2, R>::run(v[b], sign, gtab + generic_tab_offset(R));
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pidx(base + r * NS)] = v[b][r];
        }
    }
    lds_barrier();
}

template <typename T2, int NT, int N, int P, typename TW>
__device__ __forceinline__ void fft_static_from(T2* buf, TW tw, int sign, int tid)
{
    if constexpr (P < sf_npass(N)) {
        constexpr int R = sf_radix(N, P);
        stockham_pass_c<T2, R, NT, N, sf_span(N, P)>(buf, tw + sf_twoff(N, P), tw + sf_twoff(N, sf_npass(N)),
                                                     sign, tid);
        fft_static_from<T2, NT, N, P + 1>(buf, tw, sign, tid);
    }
}

// One Stockham pass of a compile-time plan on an unpadded row: input r
// of butterfly j is buf[j + r m], output r goes to buf[(j - k) R + k +
// r NS] (k = j mod NS).  Every address is one base per butterfly plus a
// constant, so the LDS instructions carry the offsets as immediates and
// no per-sample index arithmetic is left.  Reads are contiguous across
// the lanes; the stores of the first pass are strided by R, free of bank
// conflicts when R is odd (the odd-first schedule puts 15s, 5s and 3s
// first).
template <typename T2, int R, int NT, int N, int NS, typename TW>
__device__ __forceinline__ void stockham_pass_lin(T2* buf, TW tw, TW gtab, int sign, int tid)
{
    constexpr int m = N / R;
    constexpr int MAXB = (m + NT - 1) / NT;
    static_assert((MAXB - 1) * NT + (R - 1) * m < 8192 && (R - 1) * NS < 8192, "LDS immediate offsets");
    T2 v[MAXB][R];
    const T2* src = buf + tid;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (m % NT == 0 || j < m) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = src[b * NT + r * m];
        }
    }
    lds_barrier();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * NT;
        if (m % NT == 0 || j < m) {
            const int k = NS == 1 ? 0 : j % NS;
            if (NS > 1) {
                const TW t = tw + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    T2 w = t[r - 1];
                    if (sign > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            Dft<T2, R>::run(v[b], sign, gtab + generic_tab_offset(R));
            T2* dst = buf + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * NS] = v[b][r];
        }
    }
    lds_barrier();
}

template <typename T2, int NT, int N, int P, typename TW>
__device__ __forceinline__ void fft_lin_from(T2* buf, TW tw, int sign, int tid)
{
    if constexpr (P < sf_npass(N, true)) {
        constexpr int R = sf_radix(N, P, true);
        stockham_pass_lin<T2, R, NT, N, sf_span(N, P, true)>(buf, tw + sf_twoff(N, P, true),
                                                             tw + sf_twoff(N, sf_npass(N, true), true), sign, tid);
        fft_lin_from<T2, NT, N, P + 1>(buf, tw, sign, tid);
    }
}

// Full transform of length N (a compile-time constant) of an unpadded
// LDS row (entry i at buf[i], in and out), on the odd-first radix
// schedule with that schedule's twiddle table (get_plan with odd_first).
// Every thread of the group calls it.
template <typename T2, int NT, int N, typename TW>
__device__ __forceinline__ void fft_lin(T2* buf, TW tw, int sign, int tid)
{
    static_assert(sf_npass(N, true) > 0 && sf_npass(N, true) <= kMaxPasses, "length with prime factors above 13");
    fft_lin_from<T2, NT, N, 0>(buf, tw, sign, tid);
}

// Full transform of length N (a compile-time constant) in LDS; same
// contract as fft_lds, with the twiddle table of the runtime plan of N.
template <typename T2, int NT, int N, typename TW>
__device__ __forceinline__ void fft_static(T2* buf, TW tw, int sign, int tid)
{
    static_assert(sf_npass(N) > 0 && sf_npass(N) <= kMaxPasses, "length with prime factors above 13");
    fft_static_from<T2, NT, N, 0>(buf, tw, sign, tid);
}

}  // namespace wct
