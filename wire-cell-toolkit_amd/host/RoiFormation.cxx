#include "RoiFormation.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>

namespace WireCell {
namespace Hip {

RoiList row_runs(const uint8_t* m, int nt)
{
    RoiList out;
    int j = 0;
    while (j < nt - 1) {
        if (!m[j]) {
            ++j;
            continue;
        }
        int e = j;
        while (e + 1 < nt && m[e + 1]) ++e;
        if (!out.empty() && j <= out.back().second)
            out.back().second = e;
        else
            out.emplace_back(j, e);
        j = e + 2;
    }
    return out;
}

void extend_rois(RoiList& rois, int pad, int nt)
{
    RoiList out;
    out.reserve(rois.size());
    for (const Roi& r : rois) {
        const int b = r.first - pad < 0 ? 0 : r.first - pad;
        const int e = r.second + pad >= nt ? nt - 1 : r.second + pad;
        if (!out.empty() && b < out.back().second) {
            if (e > out.back().second) out.back().second = e;
        } else {
            out.emplace_back(b, e);
        }
    }
    rois.swap(out);
}

void connect_rois(std::vector<RoiList>& rows, float asy)
{
    const int nw = (int)rows.size();
    for (int i = 0; i + 2 < nw; ++i) {
        const RoiList& lo = rows[i];
        const RoiList& hi = rows[i + 2];
        RoiList& mid = rows[i + 1];
        for (const Roi& a : lo) {
            const int len1 = a.second - a.first + 1;
            for (const Roi& b : hi) {
                const int len2 = b.second - b.first + 1;
                if (!(std::abs(len2 - len1) < (len2 + len1) * asy)) continue;
                // midpoints truncated toward zero as the reference's
                // int = double assignment does
                const int s3 = (int)((a.first + b.first) / 2.);
                const int e3 = (int)((a.second + b.second) / 2.);
                if (!(s3 < e3 && s3 <= a.second && s3 <= b.second && e3 >= a.first && e3 >= b.first)) continue;
                bool overlaps = false;
                for (const Roi& c : mid) {
                    const int ms = c.first > s3 ? c.first : s3;
                    const int me = c.second < e3 ? c.second : e3;
                    if (ms < me) {
                        overlaps = true;
                        break;
                    }
                }
                if (!overlaps) mid.emplace_back(s3, e3);
            }
        }
    }
}

std::vector<RoiList> tight_rois(const uint8_t* mask, int nw, int nt, long ld, int pad, float asy)
{
    std::vector<RoiList> rows(nw);
    for (int w = 0; w < nw; ++w) {
        rows[w] = row_runs(mask + (size_t)w * ld, nt);
        extend_rois(rows[w], pad, nt);
    }
    connect_rois(rows, asy);
    return rows;
}

namespace {

// Waveform::percentile: element (size_t)(p * (float)n), at most n - 1, of
// the sorted set (p a float).
float percentile(std::vector<float>& v, float p)
{
    const size_t n = v.size();
    if (n == 1) return v[0];
    size_t k = (size_t)(p * (float)n);
    if (k > n - 1) k = n - 1;
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

// The rebinned row with the reference's wrap-around reads.
struct Groups {
    std::vector<float> v;
    int n() const { return (int)v.size(); }
    float at(int i) const { return v[i]; }
    float wrapped(int i) const
    {
        const int m = n();
        while (i < 0) i += m;
        while (i >= m) i -= m;
        return v[i];
    }
    // local_ave(signal, i, 1)
    double avg3(int i) const
    {
        double s = 0;
        for (int d = -1; d <= 1; ++d) s += wrapped(i + d);
        return s / 3.0;
    }
    // find_ROI_end: forward while above th, then downhill
    int walk_end(int i, double th, bool jump) const
    {
        const int last = n() - 1;
        int e = i;
        double c = at(e);
        while (c > th) {
            ++e;
            c = wrapped(e);
            if (e >= last) {
                e = last;
                break;
            }
        }
        while (avg3(e + 1) < avg3(e) + 25 || (jump && avg3(e + 2) < avg3(e) + 25)) {
            if (++e >= last) {
                e = last;
                break;
            }
        }
        return e;
    }
    // find_ROI_begin: backward while above th, then downhill
    int walk_begin(int i, double th, bool jump) const
    {
        int b = i;
        double c = at(b);
        while (c > th) {
            --b;
            c = wrapped(b);
            if (b <= 0) {
                b = 0;
                break;
            }
        }
        while (avg3(b - 1) < avg3(b) + 25 || (jump && avg3(b - 2) < avg3(b) + 25)) {
            if (--b <= 0) {
                b = 0;
                break;
            }
        }
        return b;
    }
    int argmax(int b, int e) const
    {
        int m = b;
        for (int k = b; k <= e; ++k)
            if (v[k] > v[m]) m = k;
        return m;
    }
};

bool in_ranges(const RoiList& bad, int t)
{
    for (const Roi& r : bad)
        if (t >= r.first && t <= r.second) return true;
    return false;
}

int x86_int(float v)
{
    if (v >= -2147483648.0f && v < 2147483648.0f) return (int)v;
    return INT_MIN;
}

}  // namespace

double cal_rms(std::vector<float> samples)
{
    if (samples.empty()) return 0;
    std::vector<float> w(samples);
    const float p0 = percentile(w, (float)(0.5 - 0.34));
    const float p1 = percentile(w, 0.5f);
    const float p2 = percentile(w, (float)(0.5 + 0.34));
    const float spread = (float)std::sqrt((std::pow((double)(p2 - p1), 2) + std::pow((double)(p1 - p0), 2)) / 2.);
    float sum = 0, count = 0;
    for (float x : samples) {
        if (std::fabs(x) < 5.0 * spread) {
            sum = (float)(sum + std::pow((double)x, 2));
            count += 1;
        }
    }
    return count > 0 ? (double)std::sqrt(sum / count) : 0.0;
}

RoiList loose_row(const float* row, int nt, const RoiList& bad, double rms, const LooseParams& lp)
{
    Groups g;
    const int ng = lp.rebin > 0 ? nt / lp.rebin : 0;
    g.v.resize(ng);
    for (int i = 0; i < ng; ++i) {
        double sum = 0;
        for (int j = 0; j < lp.rebin; ++j) {
            const int t = lp.rebin * i + j;
            if (bad.empty() || !in_ranges(bad, t)) sum += row[t];
        }
        g.v[i] = (float)sum;
    }
    float th = (float)(rms * lp.rebin * lp.factor);
    if (th > lp.max_th) th = (float)lp.max_th;
    const double th1 = th * lp.factor1;
    const bool jump = lp.jump_one_bin != 0;
    RoiList out;
    for (int j = 1; j < ng - 1; ++j) {
        const double c = g.at(j), prev = g.at(j - 1), next = g.at(j + 1);
        int b = 0, e = 0;
        bool take = false;
        if (c > th) {
            b = g.walk_begin(j, th1, jump);
            e = g.walk_end(j, th1, jump);
            take = true;
        } else if (c > prev && c > next) {
            b = g.walk_begin(j, prev, jump);
            e = g.walk_end(j, next, jump);
            const int m = g.argmax(b, e);
            const float rise = g.at(m) - g.at(b) + g.at(m) - g.at(e);
            if (rise > th * 2) take = true;
            const int tb = std::max(m - lp.short_length, b);
            const int te = std::min(m + lp.short_length, e);
            const float dl = g.at(m) - g.at(tb), dr = g.at(m) - g.at(te);
            if (dl > th1 && dr > th1) take = true;
        }
        if (!take) continue;
        if (!out.empty() && b <= out.back().second)
            out.back().second = e;
        else
            out.emplace_back(b, e);
        j = e < ng ? e : ng;
    }
    for (Roi& r : out) {
        r.first *= lp.rebin;
        r.second = r.second * lp.rebin + (lp.rebin - 1);
    }
    return out;
}

void extend_loose(RoiList& loose, const RoiList& tight)
{
    RoiList out;
    for (Roi r : loose) {
        for (const Roi& t : tight) {
            if (r.first > t.first && r.first < t.second) r.first = t.first;
            if (r.second > t.first && r.second < t.second) r.second = t.second;
        }
        if (!out.empty() && r.first < out.back().second)
            out.back().second = r.second;
        else
            out.push_back(r);
    }
    loose.swap(out);
}

std::vector<RoiList> loose_rois(const float* r, int nw, int nt, long ld, const int* bad_ptr, const int* bad_lo,
                                const int* bad_hi, const float* rms, const std::vector<RoiList>* tight,
                                const LooseParams& lp)
{
    std::vector<RoiList> out(nw);
    for (int w = 0; w < nw; ++w) {
        const float* row = r + (size_t)w * ld;
        RoiList bad;
        if (bad_ptr)
            for (int q = bad_ptr[w]; q < bad_ptr[w + 1]; ++q) bad.emplace_back(bad_lo[q], bad_hi[q]);
        double wr;
        if (rms) {
            wr = rms[w];
        } else {
            std::vector<float> kept;
            kept.reserve(nt);
            for (int t = 0; t < nt; ++t)
                if (bad.empty() || !in_ranges(bad, t)) kept.push_back(row[t]);
            wr = cal_rms(std::move(kept));
        }
        out[w] = loose_row(row, nt, bad, wr, lp);
        if (tight && w < (int)tight->size()) extend_loose(out[w], (*tight)[w]);
    }
    return out;
}

void apply_roi(float* r, int nw, int nt, long ld, const std::vector<RoiList>& rois, bool linear)
{
    std::vector<float> out(nt);
    for (int w = 0; w < nw; ++w) {
        float* row = r + (size_t)w * ld;
        std::fill(out.begin(), out.end(), 0.0f);
        if (w < (int)rois.size()) {
            for (const Roi& q : rois[w]) {
                const float a = row[q.first], z = row[q.second];
                const float span = (float)(q.second - q.first);
                for (int i = q.first; i <= q.second; ++i) {
                    float v = row[i];
                    if (linear) v = row[i] - ((z - a) * (float)(i - q.first) / span + a);
                    out[i] = (float)x86_int(v);
                }
            }
        }
        std::copy(out.begin(), out.end(), row);
    }
}

}  // namespace Hip
}  // namespace WireCell
