// High-resolution peak search (sigproc/src/PeakFinding.cxx:34-474).
//
// One working array of 7 (n + 2 numberIterations) doubles holds the
// stages in fixed regions of ext = n + 2 numberIterations entries each,
// as the reference lays them out: region 0 the response and then the
// iterated solution, 1 the extended (smoothed) source and then A^T A,
// 2 |source| and then the vector p, 3 the iteration scratch, 4 p while
// it is built, 6 an untouched copy of the extended source.  p is longer
// than one region (ext + 2 (response length - 1)) and spills over into
// the region that follows it; the iterations keep, where their update is
// skipped, whatever that region holds, so the layout is part of the
// result and is kept as is.
#include "PeakFinding.h"

#include <algorithm>
#include <cmath>

namespace WireCell {
namespace Hip {

namespace {

// Returns the number of peaks; px receives their positions.
int search_high_res(const std::vector<double>& src, const PeakSearch& ps, std::vector<double>& px)
{
    const int n = (int)src.size();
    const double sigma = ps.sigma;
    const int iters = (int)(7 * sigma + 0.5);
    if (sigma < 1) return 0;
    if (ps.threshold <= 0 || ps.threshold >= 100) return 0;
    if ((int)(5.0 * sigma + 0.5) >= 1024 / 2) return 0;
    if (ps.markov && ps.aver_window <= 0) return 0;
    if (ps.background_remove && n < 2 * iters + 1) return 0;

    // slope of a straight-line fit to the first 2 sigma samples, at most 0
    double slope = 0;
    const int kfit = int(2 * sigma + 0.5);
    if (kfit >= 2) {
        double m0 = 0, m1 = 0, m2 = 0, l0 = 0, l1 = 0;
        for (int i = 0; i < kfit; i++) {
            const double a = i, b = src[i];
            m0 += 1, m1 += a, m2 += a * a, l0 += b, l1 += a * b;
        }
        const double det = m0 * m2 - m1 * m1;
        slope = det != 0 ? (-l0 * m1 + l1 * m0) / det : 0;
        if (slope > 0) slope = 0;
    }

    const int ext = n + 2 * iters, shift = iters, bw = 2;
    std::vector<double> ws(7 * (size_t)(n + 2 * iters), 0.0);
    double* R0 = ws.data();
    double* R1 = R0 + ext;
    double* R2 = R0 + 2 * ext;
    double* R3 = R0 + 3 * ext;
    double* R4 = R0 + 4 * ext;
    double* R6 = R0 + 6 * ext;

    // the source extended by iters samples each side: the fitted line
    // (clipped at 0) before it, its last sample (clipped) after it
    for (int i = 0; i < ext; i++) {
        double v;
        if (i < shift) {
            v = src[0] + slope * (double)(i - shift);
            if (v < 0) v = 0;
        } else if (i >= n + shift) {
            v = src[n - 1];
            if (v < 0) v = 0;
        } else {
            v = src[i - shift];
        }
        R1[i] = v;
    }

    if (ps.background_remove) {   // clipping filter (not used by BreakROI)
        for (int i = 1; i <= iters; i++) {
            for (int j = i; j < ext - i; j++) {
                double a = R1[j];
                if (!ps.markov) {
                    const double b = (R1[j - i] + R1[j + i]) / 2.0;
                    if (b < a) a = b;
                    R0[j] = a;
                } else {
                    double av = 0, men = 0;
                    for (int w = j - bw; w <= j + bw; w++)
                        if (w >= 0 && w < ext) av += R1[w], men += 1;
                    av = av / men;
                    double b = 0;
                    men = 0;
                    for (int w = j - i - bw; w <= j - i + bw; w++)
                        if (w >= 0 && w < ext) b += R1[w], men += 1;
                    b = b / men;
                    double c = 0;
                    men = 0;
                    for (int w = j + i - bw; w <= j + i + bw; w++)
                        if (w >= 0 && w < ext) c += R1[w], men += 1;
                    c = c / men;
                    b = (b + c) / 2;
                    if (b < a) av = b;
                    R0[j] = av;
                }
            }
            for (int j = i; j < ext - i; j++) R1[j] = R0[j];
        }
        for (int j = 0; j < ext; j++) {
            double b;
            if (j < shift) {
                b = src[0] + slope * (double)(j - shift);
                if (b < 0) b = 0;
            } else if (j >= n + shift) {
                b = src[n - 1];
                if (b < 0) b = 0;
            } else {
                b = src[j - shift];
            }
            R1[j] = b - R1[j];
        }
        for (int j = 0; j < ext; j++)
            if (R1[j] < 0) R1[j] = 0;
    }

    for (int i = 0; i < ext; i++) R6[i] = R1[i];

    if (ps.markov) {   // Markov-chain smoothing of the extended source
        for (int j = 0; j < ext; j++) R2[j] = R1[j];
        const int xmin = 0, xmax = ext - 1;
        double maxch = 0, area = 0;
        for (int i = 0; i < ext; i++) {
            R0[i] = 0;
            if (maxch < R2[i]) maxch = R2[i];
            area += R2[i];
        }
        if (maxch == 0) return 0;
        double nom = 1;
        R0[xmin] = 1;
        for (int i = xmin; i < xmax; i++) {
            const double nip = R2[i] / maxch, nim = R2[i + 1] / maxch;
            double sp = 0, sm = 0;
            for (int l = 1; l <= ps.aver_window; l++) {
                double a = (i + l > xmax) ? R2[xmax] / maxch : R2[i + l] / maxch;
                double b = a - nip;
                a = (a + nip <= 0) ? 1 : std::sqrt(a + nip);
                b = b / a;
                b = std::exp(b);
                sp = sp + b;
                a = (i - l + 1 < xmin) ? R2[xmin] / maxch : R2[i - l + 1] / maxch;
                b = a - nim;
                a = (a + nim <= 0) ? 1 : std::sqrt(a + nim);
                b = b / a;
                b = std::exp(b);
                sm = sm + b;
            }
            const double a = R0[i + 1] = R0[i] * (sp / sm);
            nom = nom + a;
        }
        for (int i = xmin; i <= xmax; i++) R0[i] = R0[i] / nom;
        for (int j = 0; j < ext; j++) R1[j] = R0[j] * area;
        for (int j = 0; j < ext; j++) R2[j] = R1[j];
        if (ps.background_remove) {
            for (int i = 1; i <= iters; i++) {
                for (int j = i; j < ext - i; j++) {
                    double a = R1[j];
                    const double b = (R1[j - i] + R1[j + i]) / 2.0;
                    if (b < a) a = b;
                    R0[j] = a;
                }
                for (int j = i; j < ext - i; j++) R1[j] = R0[j];
            }
            for (int j = 0; j < ext; j++) R1[j] = R2[j] - R1[j];
        }
    }

    // Gold deconvolution with a Gaussian response of width sigma,
    // integer-valued (1000 exp(-x^2 / 2 sigma^2) truncated)
    double area = 0, maximum = 0;
    int lh = -1, posit = 0;
    for (int i = 0; i < ext; i++) {
        double lda = (double)i - 3 * sigma;
        lda = lda * lda / (2 * sigma * sigma);
        const int q = (int)(1000 * std::exp(-lda));
        lda = q;
        if (lda != 0) lh = i + 1;
        R0[i] = lda;
        area = area + lda;
        if (lda > maximum) {
            maximum = lda;
            posit = i;
        }
    }
    for (int i = 0; i < ext; i++) R2[i] = std::fabs(R1[i]);
    {   // b = A^T A (the response's autocorrelation) into region 1
        int h = lh - 1;
        if (h > ext) h = ext;
        for (int i = -h; i <= h; i++) {
            double lda = 0;
            const int jmin = i < 0 ? -i : 0;
            int jmax = lh - 1 - i;
            if (jmax > lh - 1) jmax = lh - 1;
            for (int j = jmin; j <= jmax; j++) lda = lda + R0[j] * R0[i + j];
            R1[i + h] = lda;
        }
    }
    {   // p = A^T y into region 4 (longer than a region), then moved to 2
        const int imin = -(lh - 1), imax = ext + lh - 2;
        for (int i = imin; i <= imax; i++) {
            double lda = 0;
            for (int j = 0; j <= lh - 1; j++) {
                const int k = i + j;
                if (k >= 0 && k < ext) lda = lda + R0[j] * R2[k];
            }
            R4[i - imin] = lda;
        }
        for (int i = imin; i <= imax; i++) R2[i - imin] = R4[i - imin];
    }
    for (int i = 0; i < ext; i++) R0[i] = 1;
    for (int it = 0; it < ps.decon_iterations; it++) {
        for (int i = 0; i < ext; i++) {
            if (std::fabs(R2[i]) > 0.00001 && std::fabs(R0[i]) > 0.00001) {
                double lda = 0;
                int jmin = lh - 1;
                if (jmin > i) jmin = i;
                jmin = -jmin;
                int jmax = lh - 1;
                if (jmax > ext - 1 - i) jmax = ext - 1 - i;
                for (int j = jmin; j <= jmax; j++) lda = lda + R1[j + lh - 1] * R0[i + j];
                double ldb = R2[i];
                lda = lda != 0 ? ldb / lda : 0;
                ldb = R0[i];
                R3[i] = lda * ldb;
            }
        }
        for (int i = 0; i < ext; i++) R0[i] = R3[i];
    }
    for (int i = 0; i < ext; i++) R1[(i + posit) % ext] = R0[i];   // undo the response's offset

    double max_src = 0, max_dec = 0;
    {
        const int j = lh - 1;
        for (int i = 0; i < ext - j; i++) {
            if (i >= shift && i < n + shift) {
                R0[i] = area * R1[i + j];
                if (max_dec < R0[i]) max_dec = R0[i];
                if (max_src < R6[i]) max_src = R6[i];
            } else {
                R0[i] = 0;
            }
        }
    }
    double frac = 1;
    if (frac > ps.threshold) frac = ps.threshold;
    frac = frac / 100;

    // local maxima of the deconvolved spectrum, inserted in order of the
    // source height at their (truncated) position
    int np = 0;
    px.assign((size_t)std::max(n, ps.max_peaks + 1), 0.0);
    for (int i = 1; i < ext - 1; i++) {
        if (!(R0[i] > R0[i - 1] && R0[i] > R0[i + 1])) continue;
        if (!(i >= shift && i < n + shift)) continue;
        if (!(R0[i] > frac * max_dec && R6[i] > ps.threshold * max_src / 100.0)) continue;
        double a = 0, b = 0;
        for (int j = i - 1; j <= i + 1; j++) {
            a += (double)(j - shift) * R0[j];
            b += R0[j];
        }
        a = a / b;
        if (a < 0) a = 0;
        if (a >= n) a = n - 1;
        if (np == 0) {
            px[0] = a;
            np = 1;
            continue;
        }
        int j = 0;
        bool higher = false;
        for (; j < np && !higher; j++)
            if (R6[shift + (int)a] > R6[shift + (int)px[j]]) higher = true;
        if (!higher) {
            if (j < ps.max_peaks) px[j] = a;
        } else {
            for (int k = np; k >= j; k--)
                if (k < ps.max_peaks) px[k] = px[k - 1];
            px[j - 1] = a;
        }
        if (np < ps.max_peaks) np += 1;
    }
    return np;
}

}  // namespace

Peaks find_peaks(const std::vector<float>& signal, const PeakSearch& ps)
{
    Peaks out;
    std::vector<double> src(signal.begin(), signal.end());
    const int n = (int)src.size();
    if (!(n >= 2 * ps.sigma + 1)) return out;
    std::vector<double> px;
    const int np = search_high_res(src, ps, px);
    out.x.assign(px.begin(), px.begin() + np);
    out.y.resize(np);
    for (int i = 0; i < np; i++) out.y[i] = src[(int)std::round(out.x[i])];
    return out;
}

}  // namespace Hip
}  // namespace WireCell
