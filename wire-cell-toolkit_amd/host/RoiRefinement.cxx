// ROI refinement (sigproc/src/ROI_refinement.cxx, SignalROI.cxx).
#include "RoiRefinement.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace WireCell {
namespace Hip {

RoiRefinement::RoiRefinement(int plane, int nw, int nt, const RefineParams& p)
    : m_plane(plane), m_nw(nw), m_nt(nt), m_p(p), m_tight(nw), m_loose(nw)
{
    if (plane < 0 || plane > 2 || nw <= 0 || nt <= 0) throw std::invalid_argument("RoiRefinement: bad plane shape");
}

// SignalROI(plane, chid, start, end, signal) (SignalROI.cxx:9-22): the
// signal on [start, end] minus the line through its end samples, float.
int RoiRefinement::make_roi(int wire, int start, int end, const std::vector<float>& signal)
{
    Sroi r;
    r.wire = wire;
    r.start = start;
    r.end = end;
    r.ext_start = start;
    r.ext_end = end;
    const float sc = signal.at(start), ec = signal.at(end);
    r.contents.resize(end - start + 1);
    for (int i = start; i <= end; i++)
        r.contents[i - start] = signal.at(i) - ((ec - sc) * (i - start) / (end - start) + sc);
    m_pool.push_back(std::move(r));
    return (int)m_pool.size() - 1;
}

int RoiRefinement::copy_roi(int id)
{
    Sroi r = m_pool[id];
    r.ext_start = r.start;
    r.ext_end = r.end;
    m_pool.push_back(std::move(r));
    return (int)m_pool.size() - 1;
}

RoiRefinement::Links& RoiRefinement::grow(std::vector<Links>& m, int id)
{
    if ((int)m.size() <= id) m.resize(id + 1);
    return m[id];
}

// SignalROI::overlap (SignalROI.cxx:68-85): the common range holds more
// than one tick.
bool RoiRefinement::overlap(int a, int b) const
{
    const Sroi &A = m_pool[a], &B = m_pool[b];
    const int lo = std::max(A.start, B.start), hi = std::min(A.end, B.end);
    return hi > lo;
}

// SignalROI::overlap(roi, th, th1) (:48-66): some tick of the common
// range is above th here and above th1 there.
bool RoiRefinement::overlap_th(int a, int b, float th, float th1) const
{
    const Sroi &A = m_pool[a], &B = m_pool[b];
    const int lo = std::max(A.start, B.start), hi = std::min(A.end, B.end);
    if (!(hi > lo)) return false;
    for (int i = lo; i <= hi; i++)
        if (A.contents[i - A.start] > th && B.contents[i - B.start] > th1) return true;
    return false;
}

// SignalROI::get_above_threshold (:102-125): runs of contents above th,
// as offsets into the ROI.
std::vector<std::pair<int, int>> RoiRefinement::above(int id, float th) const
{
    const std::vector<float>& c = m_pool[id].contents;
    std::vector<std::pair<int, int>> runs;
    const int n = (int)c.size();
    for (int i = 0; i < n; i++) {
        if (!(c[i] > th)) continue;
        int e = i;
        while (e + 1 < n && c[e + 1] > th) ++e;
        runs.emplace_back(i, e);
        i = e;
    }
    return runs;
}

double RoiRefinement::average(int id) const
{
    double s = 0, n = 0;
    for (float v : m_pool[id].contents) s += v, n++;
    return n != 0 ? s / n : 0;
}

void RoiRefinement::add(std::vector<Links>& m, int key, int val)
{
    Links& l = grow(m, key);
    l.key = true;
    l.ids.push_back(val);
}

// ROI_refinement::link / unlink (:230-267): link adds each end to the
// other's list once; unlink removes the first occurrence where the key
// exists.
void RoiRefinement::link(int prev, int next)
{
    Links& f = grow(m_front, prev);
    if (!f.key || std::find(f.ids.begin(), f.ids.end(), next) == f.ids.end()) f.ids.push_back(next);
    f.key = true;
    Links& b = grow(m_back, next);
    if (!b.key || std::find(b.ids.begin(), b.ids.end(), prev) == b.ids.end()) b.ids.push_back(prev);
    b.key = true;
}

void RoiRefinement::unlink(int prev, int next)
{
    Links& f = grow(m_front, prev);
    if (f.key) {
        auto it = std::find(f.ids.begin(), f.ids.end(), next);
        if (it != f.ids.end()) f.ids.erase(it);
    }
    Links& b = grow(m_back, next);
    if (b.key) {
        auto it = std::find(b.ids.begin(), b.ids.end(), prev);
        if (it != b.ids.end()) b.ids.erase(it);
    }
}

void RoiRefinement::drop_links(int id)
{
    if (grow(m_front, id).key) {
        const std::vector<int> next = m_front[id].ids;
        for (int n : next) unlink(id, n);
        m_front[id] = Links();
    }
    if (grow(m_back, id).key) {
        const std::vector<int> prev = m_back[id].ids;
        for (int p : prev) unlink(p, id);
        m_back[id] = Links();
    }
}

// The common tail of ShrinkROI (:1783-1841) and BreakROI1 (:2381-2445):
// repl takes id's place at the end of the wire's list, inherits the
// neighbour links it overlaps (neighbour by neighbour) and the tight ROIs
// it overlaps; id's keys are erased.
void RoiRefinement::replace(int id, const std::vector<int>& repl)
{
    std::vector<int>& l = list(m_pool[id].wire);
    l.erase(std::find(l.begin(), l.end(), id));
    l.insert(l.end(), repl.begin(), repl.end());
    if (grow(m_front, id).key) {
        const std::vector<int> next = m_front[id].ids;
        for (int n : next) {
            unlink(id, n);
            for (int r : repl)
                if (overlap(r, n)) link(r, n);
        }
        m_front[id] = Links();
    }
    if (grow(m_back, id).key) {
        const std::vector<int> prev = m_back[id].ids;
        for (int p : prev) {
            unlink(p, id);
            for (int r : repl)
                if (overlap(r, p)) link(p, r);
        }
        m_back[id] = Links();
    }
    if (grow(m_contained, id).key) {
        const std::vector<int> tight = m_contained[id].ids;
        for (int t : tight)
            for (int r : repl)
                if (overlap(r, t)) add(m_contained, r, t);
        m_contained[id] = Links();
    }
}

void RoiRefinement::load(const float* r, long ld, const std::vector<RoiList>& bad, const std::vector<float>& rms,
                         const std::vector<RoiList>& tight, const std::vector<RoiList>& loose)
{
    if ((int)rms.size() != m_nw || (int)tight.size() != m_nw || (m_plane != 2 && (int)loose.size() != m_nw) ||
        (!bad.empty() && (int)bad.size() != m_nw))
        throw std::invalid_argument("RoiRefinement::load: per-wire lists do not match the plane");
    m_rms = rms;
    std::vector<float> signal(m_nt);
    for (int w = 0; w < m_nw; w++) {
        const float* row = r + (long)w * ld;
        for (int t = 0; t < m_nt; t++) {
            bool ok = true;
            if (!bad.empty())
                for (const Roi& b : bad[w])
                    if (t >= b.first && t <= b.second) {
                        ok = false;
                        break;
                    }
            signal[t] = ok ? row[t] : 0.0f;
        }
        const float th = m_rms[w] * m_p.th_factor;
        for (const Roi& q : tight[w]) {
            if (q.first < 0 || q.second >= m_nt || q.second < q.first)
                throw std::invalid_argument("RoiRefinement::load: tight ROI outside the row");
            const int id = make_roi(w, q.first, q.second, signal);
            if (above(id, th).empty()) {
                m_pool.pop_back();
                continue;
            }
            m_tight[w].push_back(id);
            // collection plane: tight ROIs overlapping the previous wire's
            // (:422-449; no link across the middle of a wrapped plane)
            if (m_plane == 2 && w > 0 && (!m_p.wrapped || (double)w != m_nw / 2.)) {
                for (int pv : m_tight[w - 1])
                    if (overlap(id, pv)) {
                        add(m_front, pv, id);
                        add(m_back, id, pv);
                    }
            }
        }
        if (m_plane == 2) continue;
        for (const Roi& q : loose[w]) {
            if (q.first < 0 || q.second >= m_nt || q.second < q.first)
                throw std::invalid_argument("RoiRefinement::load: loose ROI outside the row");
            const int id = make_roi(w, q.first, q.second, signal);
            if (above(id, th).empty()) {
                m_pool.pop_back();
                continue;
            }
            m_loose[w].push_back(id);
            if (w > 0) {
                for (int pv : m_loose[w - 1])
                    if (overlap(id, pv)) {
                        add(m_front, pv, id);
                        add(m_back, id, pv);
                    }
            }
            if (w == m_nw - 1 && m_p.wrapped) {
                for (int nx : m_loose[0])
                    if (overlap(id, nx)) {
                        add(m_back, nx, id);
                        add(m_front, id, nx);
                    }
            }
            for (int t : m_tight[w])
                if (overlap(t, id)) add(m_contained, id, t);
        }
    }
}

// Loose ROIs not connected, through the neighbour links, to one that
// contains a tight ROI are removed (induction planes).
void RoiRefinement::cleanup_rois()
{
    if (m_plane == 2) return;
    std::vector<char> saved(m_pool.size(), 0);
    for (int w = 0; w < m_nw; w++) {
        for (int id : m_loose[w]) {
            if (saved[id] || !grow(m_contained, id).key) continue;
            std::vector<int> stack{id};
            saved[id] = 1;
            while (!stack.empty()) {
                const int t = stack.back();
                stack.pop_back();
                for (std::vector<Links>* m : {&m_front, &m_back}) {
                    const Links& l = grow(*m, t);
                    if (!l.key) continue;
                    for (int n : l.ids)
                        if (!saved[n]) {
                            stack.push_back(n);
                            saved[n] = 1;
                        }
                }
            }
        }
    }
    for (int w = 0; w < m_nw; w++) {
        std::vector<int> gone;
        for (int id : m_loose[w])
            if (!saved[id]) gone.push_back(id);
        for (int id : gone) drop_links(id);
        for (int id : gone) m_loose[w].erase(std::find(m_loose[w].begin(), m_loose[w].end(), id));
    }
}

// Tight ROIs no loose ROI contains become loose ROIs of their own, linked
// to the overlapping loose ROIs of the next and previous wires.
void RoiRefinement::generate_merge_rois()
{
    if (m_plane == 2) return;
    for (int i = 0; i < m_nw; i++) {
        std::vector<char> covered(m_pool.size(), 0);
        for (int id : m_loose[i]) {
            const Links& c = grow(m_contained, id);
            if (c.key)
                for (int t : c.ids) covered[t] = 1;
        }
        std::vector<int> todo;
        for (int t : m_tight[i])
            if (!covered[t]) todo.push_back(t);
        for (int t : todo) {
            const int L = copy_roi(t);
            m_loose[i].push_back(L);
            Links& c = grow(m_contained, L);
            c.key = true;
            c.ids.assign(1, t);
            auto to_next = [&](int nw) {
                for (int nx : m_loose[nw])
                    if (overlap(L, nx)) {
                        add(m_back, nx, L);
                        add(m_front, L, nx);
                    }
            };
            auto to_prev = [&](int pw) {
                for (int pv : m_loose[pw])
                    if (overlap(L, pv)) {
                        add(m_front, pv, L);
                        add(m_back, L, pv);
                    }
            };
            if (i < m_nw - 1) to_next(i + 1);
            if (i == m_nw - 1 && m_p.wrapped) to_next(0);
            if (i > 0) to_prev(i - 1);
            if (i == 0 && m_p.wrapped) to_prev(m_nw - 1);
        }
    }
}

// Neighbour links whose ROIs are never above threshold together are cut.
void RoiRefinement::check_rois()
{
    if (m_plane == 2) return;
    for (int w = 0; w < m_nw; w++) {
        for (int id : m_loose[w]) {
            const float th = m_p.th_factor * m_rms[w];
            if (grow(m_front, id).key) {
                std::vector<int> cut;
                for (int r1 : m_front[id].ids)
                    if (!overlap_th(id, r1, th, m_p.th_factor * m_rms[m_pool[r1].wire])) cut.push_back(r1);
                for (int r1 : cut) unlink(id, r1);
            }
            if (grow(m_back, id).key) {
                std::vector<int> cut;
                for (int r1 : m_back[id].ids)
                    if (!overlap_th(id, r1, th, m_p.th_factor * m_rms[m_pool[r1].wire])) cut.push_back(r1);
                for (int r1 : cut) unlink(r1, id);
            }
        }
    }
}

// BreakROI (:1867-2328) without protected zones: the ROI's peaks are
// found; when more than one stands above th_peak rms, the ROI is split
// at the valleys between peaks separated by more than the low-peak
// separation threshold and each section gets its own straight baseline;
// tight ROIs left empty by that are copied back.  Then, twice, the
// stretches below th_factor rms are located (their sample nearest zero,
// or lowest) and each stretch between consecutive ones gets a straight
// baseline through them.  The contents are replaced in place.
void RoiRefinement::break_roi(int id, float rms)
{
    Sroi& R = m_pool[id];
    const int start = R.start, end = R.end;
    if (start < 0 || end < 0) return;
    std::vector<float> temp = R.contents;
    const int n = (int)temp.size();

    float low_sep = m_p.low_peak_sep_threshold_pre;
    if (low_sep < m_p.sep_peak * rms) low_sep = m_p.sep_peak * rms;

    PeakSearch ps;
    ps.max_peaks = m_p.max_npeaks;
    ps.sigma = m_p.sigma;
    ps.threshold = m_p.th_percent;
    const Peaks pk = find_peaks(temp, ps);
    const int nfound = (int)pk.x.size();
    if (nfound > 1) {
        int npeaks = nfound;
        std::vector<int> opp;
        int above_th = 0;
        for (int j = 0; j < npeaks; j++) {
            opp.push_back((int)(pk.x[j] + start));
            if (pk.y[j] > m_p.th_peak * rms) above_th++;
        }
        if (above_th > 1) {
            auto T = [&](double pos) -> float { return temp.at((size_t)((int)pos - start)); };
            std::sort(opp.begin(), opp.end());
            std::vector<float> vp(npeaks + 3, 0.f);
            vp[0] = start;
            float mn = 1e9;
            for (int k = 0; k < opp[0] - start; k++)
                if (temp.at(k) < mn) {
                    mn = temp[k];
                    vp[0] = k + start;
                }
            if (vp[0] != start) {   // a peak at the start, before the first valley
                opp.insert(opp.begin(), start);
                npeaks++;
                for (int j = start; j != vp[0]; j++)
                    if (temp.at(j - start) > temp.at(opp[0] - start)) opp[0] = j;
                vp[0] = start;
            }
            for (int j = 0; j != npeaks - 1; j++) {
                mn = 1e9;
                vp[j + 1] = opp[j];
                for (int k = opp[j] - start; k < opp[j + 1] - start; k++)
                    if (temp.at(k) < mn) {
                        mn = temp[k];
                        vp[j + 1] = k + start;
                    }
            }
            vp[npeaks] = end;
            mn = 1e9;
            for (int k = opp[npeaks - 1] - start; k <= end - start; k++)
                if (temp.at(k) < mn) {
                    mn = temp[k];
                    vp[npeaks] = k + start;
                }
            if (vp[npeaks] != end) {   // a peak at the end, after the last valley
                npeaks++;
                vp[npeaks] = end;
                opp.push_back(end);
                for (int j = (int)vp[npeaks - 1]; j != vp[npeaks]; j++)
                    if (temp.at(j - start) > temp.at(opp[npeaks - 1] - start)) opp[npeaks - 1] = j;
            }

            // merge peaks not separated by low_sep from their valleys
            std::vector<float> vp1(npeaks + 2, 0.f), pp1(npeaks + 2, 0.f);
            int n1 = 0;
            vp1[0] = vp[0];
            for (int j = 0; j < npeaks; j++) {
                if (n1 > 0 && T(vp[j]) < T(vp1[n1])) vp1[n1] = vp[j];
                if (T(opp[j]) - T(vp1[n1]) > low_sep) {
                    pp1[n1] = opp[j];
                    n1++;
                    bool found = false;
                    for (int k = j + 1; k != npeaks + 1; k++) {
                        if (k <= npeaks && T(opp[k - 1]) > T(pp1[n1 - 1])) pp1[n1 - 1] = opp[k - 1];
                        if (T(pp1[n1 - 1]) - T(vp[k]) > low_sep) {
                            vp1[n1] = vp[k];
                            j = k - 1;
                            found = true;
                            break;
                        }
                    }
                    if (!found) {
                        vp1[n1] = vp[npeaks];
                        j = npeaks;
                    }
                }
            }
            vp1[n1] = vp[npeaks];
            if (n1 == 0) {
                n1 = 1;
                vp1[0] = vp[0];
                vp1[1] = vp[npeaks];
            }

            // each section minus its straight baseline (double arithmetic)
            const std::vector<float> t1 = temp;
            std::fill(temp.begin(), temp.end(), 0.f);
            for (int j = 0; j != n1; j++) {
                const int sp = (int)vp1[j], ep = (int)vp1[j + 1];
                const double sc = t1.at(sp - start), ec = t1.at(ep - start);
                if (ep < sp) throw std::runtime_error("RoiRefinement: break sections out of order");
                for (int k = sp; k != ep + 1; k++) {
                    const double v = t1.at(k - start) - (sc + (ec - sc) * (k - sp) / (ep - sp));
                    temp.at(k - start) = v;
                }
            }

            // tight ROIs left without signal get their contents back
            Links& c = grow(m_contained, id);
            c.key = true;   // the reference's contained_rois[roi] creates the entry
            const std::vector<int> tights = c.ids;
            for (int t : tights) {
                const Sroi& Tr = m_pool[t];
                bool any = false;
                for (int i = Tr.start; i <= Tr.end; i++)
                    if (i - start >= 0 && i - start < n && temp[i - start] != 0) {
                        any = true;
                        break;
                    }
                if (!any)
                    for (int i = Tr.start; i <= Tr.end; i++)
                        if (i - start >= 0 && i - start < n) temp[i - start] = Tr.contents.at(i - Tr.start);
            }
        }
    }

    for (int qx = 0; qx != 2; qx++) {
        const float th = m_p.th_factor * rms;
        std::vector<int> saved;
        for (int i = 0; i < n; i++) {
            if (!(temp[i] < th)) continue;
            int e = i;
            while (e + 1 < n && temp[e + 1] < th) ++e;
            int bmin = i;
            float mn = 1e9;
            for (int j = i; j <= e; j++) {
                if (std::fabs(temp[j]) < 1e-3) {
                    bmin = j;
                    break;
                }
                if (temp[j] < mn) {
                    mn = temp[j];
                    bmin = j;
                }
            }
            saved.push_back(bmin);
            i = e;
        }
        // (no stretch below threshold: the reference indexes an empty
        // list here; every sample then ends up 0, as its sections would
        // leave them)
        const std::vector<float> t1 = temp;
        std::fill(temp.begin(), temp.end(), 0.f);
        for (int j = 0; j + 1 < (int)saved.size(); j++) {
            const int sp = saved[j], ep = saved[j + 1];
            const float sc = t1[sp], ec = t1[ep];
            for (int k = sp; k != ep + 1; k++) {
                const double v = t1[k] - (sc + (ec - sc) * (k - sp) / (ep - sp));
                temp[k] = v;
            }
        }
    }
    m_pool[id].contents = temp;
}

// BreakROI1 (:2330-2446): the ROI splits at its samples within 1e-3 of
// zero into ROIs between consecutive ones.
void RoiRefinement::break_roi1(int id)
{
    const Sroi& R = m_pool[id];
    const int start = R.start, end = R.end, wire = R.wire;
    if (start < 0 || end < 0) return;
    const std::vector<float> temp = R.contents;
    std::vector<int> bins;
    for (int i = 0; i < (int)temp.size(); i++)
        if (std::fabs(temp[i]) < 1e-3) bins.push_back(i + start);
    std::vector<int> repl;
    std::vector<float> signal(end + 1, 0.f);
    for (int i = 0; i + 1 < (int)bins.size(); i++) {
        const int s1 = bins[i], e1 = bins[i + 1];
        std::fill(signal.begin(), signal.end(), 0.f);
        for (int j = s1; j <= e1; j++) signal[j] = temp[j - start];
        if (s1 >= 0 && e1 > s1) repl.push_back(make_roi(wire, s1, e1, signal));
    }
    replace(id, repl);
}

void RoiRefinement::break_rois()
{
    if (m_plane == 2) return;
    std::vector<int> all;
    for (int w = 0; w < m_nw; w++)
        for (int id : m_loose[w]) {
            break_roi(id, m_rms[w]);
            all.push_back(id);
        }
    for (int id : all) break_roi1(id);
}

// ShrinkROI (:1603-1842): the ROI shrinks to the span of its contained
// tight ROIs and of the ticks where a neighbour is above its threshold
// while this ROI is above its own, padded by pad and kept inside the old
// range; the new ROI takes a straight baseline through its new ends.
void RoiRefinement::shrink_roi(int id)
{
    const Sroi R = m_pool[id];
    const int start = R.start, end = R.end;
    if (start < 0 || end < 0) return;
    const int n = end - start + 1;
    const float th1 = m_rms[R.wire] * m_p.th_factor;
    std::vector<float> mark(n, 0.f);
    if (grow(m_contained, id).key)
        for (int t : m_contained[id].ids)
            for (int i = m_pool[t].start; i <= m_pool[t].end; i++)
                if (i - start >= 0 && i - start < n) mark[i - start] = 1;
    for (std::vector<Links>* m : {&m_front, &m_back}) {
        if (!grow(*m, id).key) continue;
        for (int nb : (*m)[id].ids) {
            const Sroi& N = m_pool[nb];
            const float th = m_rms[N.wire] * m_p.th_factor;
            for (const auto& run : above(nb, th))
                for (int j = run.first; j <= run.second; j++) {
                    const int idx = j + N.start - start;
                    if (idx >= 0 && idx < n && R.contents[idx] > th1) mark[idx] = 1;
                }
        }
    }
    int ns = start, ne = end;
    for (int i = 0; i < n; i++)
        if (mark[i] > 0) {
            ns = i + start;
            break;
        }
    for (int i = n - 1; i >= 0; i--)
        if (mark[i] > 0) {
            ne = i + start;
            break;
        }
    ns -= m_p.pad;
    ne += m_p.pad;
    if (ns < start) ns = start;
    if (ne > end) ne = end;
    std::vector<float> signal(end + 1, 0.f);
    for (int i = ns; i <= ne; i++) signal[i] = R.contents[i - start];
    std::vector<int> repl;
    if (ns >= 0 && ne > ns) repl.push_back(make_roi(R.wire, ns, ne, signal));
    replace(id, repl);
}

void RoiRefinement::shrink_rois()
{
    if (m_plane == 2) return;
    std::vector<int> all;
    for (int w = 0; w < m_nw; w++) all.insert(all.end(), m_loose[w].begin(), m_loose[w].end());
    for (int id : all) shrink_roi(id);
}

// CleanUpCollectionROIs / CleanUpInductionROIs: an ROI neither above the
// high threshold anywhere nor (where used) above the low one on average,
// with no neighbour that is, is removed.
void RoiRefinement::remove_unsupported(bool use_mean, float high_th, float mean_th)
{
    std::vector<char> good(m_pool.size(), 0);
    for (int w = 0; w < m_nw; w++)
        for (int id : list(w))
            if (!above(id, high_th).empty() || (use_mean && average(id) > mean_th)) good[id] = 1;
    std::vector<int> bad;
    for (int w = 0; w < m_nw; w++)
        for (int id : list(w)) {
            if (good[id]) continue;
            bool near_good = false;
            for (std::vector<Links>* m : {&m_front, &m_back}) {
                const Links& l = grow(*m, id);
                if (!l.key) continue;
                for (int nb : l.ids)
                    if (good[nb]) {
                        near_good = true;
                        break;
                    }
                if (near_good) break;
            }
            if (!near_good) bad.push_back(id);
        }
    for (int id : bad) {
        drop_links(id);
        std::vector<int>& l = list(m_pool[id].wire);
        auto it = std::find(l.begin(), l.end(), id);
        if (it != l.end()) l.erase(it);
    }
}

void RoiRefinement::cleanup_collection_rois()
{
    if (m_plane != 2) return;
    remove_unsupported(true, m_p.fake_signal_high_th, m_p.fake_signal_low_th);
}

void RoiRefinement::cleanup_induction_rois()
{
    if (m_plane == 2) return;
    const float mean_th = m_p.fake_signal_low_th * m_p.fake_signal_low_th_ind_factor;
    const float high_th = m_p.fake_signal_high_th * m_p.fake_signal_high_th_ind_factor;
    // the V plane uses the high threshold only (:1483-1487)
    remove_unsupported(m_plane == 0, high_th, mean_th);
}

// ExtendROIs (:2946-3136): each wire's list is sorted by start tick; an
// ROI's extended range covers its neighbours on the adjacent wires; where
// extended ranges of consecutive ROIs overlap, both end at the midpoint of
// the gap between the ROIs.
void RoiRefinement::extend_rois()
{
    for (int w = 0; w < m_nw; w++) {
        std::vector<int>& l = list(w);
        std::stable_sort(l.begin(), l.end(), [&](int a, int b) { return m_pool[a].start < m_pool[b].start; });
        for (int id : l) {
            Sroi& R = m_pool[id];
            R.ext_start = R.start;
            R.ext_end = R.end;
            for (std::vector<Links>* m : {&m_front, &m_back})
                for (int nb : grow(*m, id).ids) {
                    if (R.ext_start > m_pool[nb].start) R.ext_start = m_pool[nb].start;
                    if (R.ext_end < m_pool[nb].end) R.ext_end = m_pool[nb].end;
                }
        }
        int prev = -1;
        for (int id : l) {
            if (prev >= 0 && m_pool[prev].ext_end > m_pool[id].ext_start) {
                const int mid = int(m_pool[prev].end * 0.5 + m_pool[id].start * 0.5);
                m_pool[prev].ext_end = mid;
                m_pool[id].ext_start = mid;
            }
            prev = id;
        }
    }
}

void RoiRefinement::refine()
{
    cleanup_rois();
    generate_merge_rois();
    for (int qx = 0; qx != m_p.break_roi_loop; qx++) {
        break_rois();
        check_rois();
        cleanup_rois();
    }
    shrink_rois();
    check_rois();
    cleanup_rois();
    if (m_plane == 2)
        cleanup_collection_rois();
    else
        cleanup_induction_rois();
    extend_rois();
}

std::vector<RoiList> RoiRefinement::rois(bool extended) const
{
    std::vector<RoiList> out(m_nw);
    for (int w = 0; w < m_nw; w++)
        for (int id : list(w)) {
            const Sroi& R = m_pool[id];
            out[w].emplace_back(extended ? R.ext_start : R.start, extended ? R.ext_end : R.end);
        }
    return out;
}

std::vector<std::vector<float>> RoiRefinement::contents(int wire) const
{
    std::vector<std::vector<float>> out;
    for (int id : list(wire)) out.push_back(m_pool[id].contents);
    return out;
}

}  // namespace Hip
}  // namespace WireCell
