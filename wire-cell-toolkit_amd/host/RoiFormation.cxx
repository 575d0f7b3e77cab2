#include "RoiFormation.h"

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

}  // namespace Hip
}  // namespace WireCell
