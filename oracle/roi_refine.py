"""Pure-Python restatement of Wire-Cell's ROI refinement
(sigproc/src/ROI_refinement.cxx, SignalROI.cxx, PeakFinding.cxx).

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, and only as
the checker.  It follows the reference's control flow statement by
statement: SignalROI objects, wire lists, and the front/back/contained maps
keyed by ROI object (dicts, whose key presence is tested as the reference
tests its std::map keys).  Float arithmetic of the reference is done in
numpy float32 scalars, double arithmetic in Python floats (IEEE double).
The multi-plane protection (MP2ROI/MP3ROI) is not restated: BreakROI runs
with no protected zones.
"""
import math

import numpy as np

F = np.float32


class SignalROI:
    """SignalROI (SignalROI.cxx:9-32)."""

    def __init__(self, plane, chid, start_bin, end_bin, signal=None, copy_of=None):
        if copy_of is not None:   # SignalROI(SignalROI*) (:24-32)
            self.plane, self.chid = copy_of.plane, copy_of.chid
            self.start_bin, self.end_bin = copy_of.start_bin, copy_of.end_bin
            self.contents = list(copy_of.contents)
        else:                     # :9-22, float arithmetic
            self.plane, self.chid = plane, chid
            self.start_bin, self.end_bin = start_bin, end_bin
            sc = F(signal[start_bin])
            ec = F(signal[end_bin])
            with np.errstate(divide="ignore", invalid="ignore"):
                self.contents = [F(signal[i]) - ((ec - sc) * F(i - start_bin) / F(end_bin - start_bin) + sc)
                                 for i in range(start_bin, end_bin + 1)]
        self.ext_start_bin = None
        self.ext_end_bin = None

    def get_above_threshold(self, th):   # :102-125
        bins = []
        i = 0
        n = len(self.contents)
        while i < n:
            if self.contents[i] > th:
                start = end = i
                for j in range(i + 1, n):
                    if self.contents[j] > th:
                        end = j
                    else:
                        break
                bins.append((start, end))
                i = end
            i += 1
        return bins

    def get_average_heights(self):   # :34-46
        s1 = 0.0
        s2 = 0.0
        for v in self.contents:
            s1 += float(v)
            s2 += 1
        return s1 / s2 if s2 != 0 else 0.0

    def overlap(self, roi):   # :68-85
        min_start = self.start_bin
        if self.start_bin < roi.start_bin:
            min_start = roi.start_bin
        min_end = self.end_bin
        if self.end_bin > roi.end_bin:
            min_end = roi.end_bin
        return min_end > min_start

    def overlap_th(self, roi1, th, th1):   # :48-66
        min_start = self.start_bin
        if self.start_bin < roi1.start_bin:
            min_start = roi1.start_bin
        min_end = self.end_bin
        if self.end_bin > roi1.end_bin:
            min_end = roi1.end_bin
        if min_end > min_start:
            for i in range(min_start, min_end + 1):
                if self.contents[i - self.start_bin] > th and roi1.contents[i - roi1.start_bin] > th1:
                    return True
            return False
        return False


# ---- PeakFinding (PeakFinding.cxx:34-474) ---------------------------------

def find_peak(signal, fMaxPeaks=200, sigma=1.0, threshold=0.05, backgroundRemove=False,
              deconIterations=3, markov=True, averWindow=3):
    """PeakFinding::find_peak: (positions, heights)."""
    source = [float(v) for v in signal]
    ssize = len(source)
    fpx = [0.0] * max(ssize, fMaxPeaks + 1)
    if ssize >= 2 * sigma + 1:
        npeaks = _search_high_res(source, ssize, fpx, fMaxPeaks, sigma, threshold, backgroundRemove,
                                  deconIterations, markov, averWindow)
    else:
        npeaks = 0
    xs = fpx[:npeaks]
    ys = [source[int(math.floor(x + 0.5))] for x in xs]   # std::round of a non-negative position
    return xs, ys


def _search_high_res(source, ssize, fPositionX, fMaxPeaks, sigma, threshold, backgroundRemove,
                     deconIterations, markov, averWindow):
    PEAK_WINDOW = 1024
    numberIterations = int(7 * sigma + 0.5)
    size_ext = ssize + 2 * numberIterations
    shift = numberIterations
    bw = 2
    if sigma < 1:
        return 0
    if threshold <= 0 or threshold >= 100:
        return 0
    j = int(5.0 * sigma + 0.5)
    if j >= PEAK_WINDOW / 2:
        return 0
    if markov and averWindow <= 0:
        return 0
    if backgroundRemove and ssize < 2 * numberIterations + 1:
        return 0
    m0low = m1low = m2low = l0low = l1low = 0.0
    k = int(2 * sigma + 0.5)
    if k >= 2:
        for i in range(k):
            a = float(i)
            b = source[i]
            m0low += 1
            m1low += a
            m2low += a * a
            l0low += b
            l1low += a * b
        detlow = m0low * m2low - m1low * m1low
        if detlow != 0:
            l1low = (-l0low * m1low + l1low * m0low) / detlow
        else:
            l1low = 0.0
        if l1low > 0:
            l1low = 0.0
    else:
        l1low = 0.0
    i = int(7 * sigma + 0.5)
    i = 2 * i
    ws = [0.0] * (7 * (ssize + i))
    for i in range(size_ext):
        if i < shift:
            a = float(i - shift)
            ws[i + size_ext] = source[0] + l1low * a
            if ws[i + size_ext] < 0:
                ws[i + size_ext] = 0.0
        elif i >= ssize + shift:
            ws[i + size_ext] = source[ssize - 1]
            if ws[i + size_ext] < 0:
                ws[i + size_ext] = 0.0
        else:
            ws[i + size_ext] = source[i - shift]
    if backgroundRemove:
        for i in range(1, numberIterations + 1):
            for j in range(i, size_ext - i):
                if not markov:
                    a = ws[size_ext + j]
                    b = (ws[size_ext + j - i] + ws[size_ext + j + i]) / 2.0
                    if b < a:
                        a = b
                    ws[j] = a
                else:
                    a = ws[size_ext + j]
                    av = 0.0
                    men = 0.0
                    for w in range(j - bw, j + bw + 1):
                        if 0 <= w < size_ext:
                            av += ws[size_ext + w]
                            men += 1
                    av = av / men
                    b = 0.0
                    men = 0.0
                    for w in range(j - i - bw, j - i + bw + 1):
                        if 0 <= w < size_ext:
                            b += ws[size_ext + w]
                            men += 1
                    b = b / men
                    c = 0.0
                    men = 0.0
                    for w in range(j + i - bw, j + i + bw + 1):
                        if 0 <= w < size_ext:
                            c += ws[size_ext + w]
                            men += 1
                    c = c / men
                    b = (b + c) / 2
                    if b < a:
                        av = b
                    ws[j] = av
            for j in range(i, size_ext - i):
                ws[size_ext + j] = ws[j]
        for j in range(size_ext):
            if j < shift:
                b = source[0] + l1low * float(j - shift)
                if b < 0:
                    b = 0.0
                ws[size_ext + j] = b - ws[size_ext + j]
            elif j >= ssize + shift:
                b = source[ssize - 1]
                if b < 0:
                    b = 0.0
                ws[size_ext + j] = b - ws[size_ext + j]
            else:
                ws[size_ext + j] = source[j - shift] - ws[size_ext + j]
        for j in range(size_ext):
            if ws[size_ext + j] < 0:
                ws[size_ext + j] = 0.0
    for i in range(size_ext):
        ws[i + 6 * size_ext] = ws[i + size_ext]
    if markov:
        for j in range(size_ext):
            ws[2 * size_ext + j] = ws[size_ext + j]
        xmin, xmax = 0, size_ext - 1
        maxch = 0.0
        plocha = 0.0
        for i in range(size_ext):
            ws[i] = 0.0
            if maxch < ws[2 * size_ext + i]:
                maxch = ws[2 * size_ext + i]
            plocha += ws[2 * size_ext + i]
        if maxch == 0:
            return 0
        nom = 1.0
        ws[xmin] = 1.0
        for i in range(xmin, xmax):
            nip = ws[2 * size_ext + i] / maxch
            nim = ws[2 * size_ext + i + 1] / maxch
            sp = 0.0
            sm = 0.0
            for l in range(1, averWindow + 1):
                if i + l > xmax:
                    a = ws[2 * size_ext + xmax] / maxch
                else:
                    a = ws[2 * size_ext + i + l] / maxch
                b = a - nip
                a = 1.0 if a + nip <= 0 else math.sqrt(a + nip)
                b = b / a
                b = math.exp(b)
                sp = sp + b
                if i - l + 1 < xmin:
                    a = ws[2 * size_ext + xmin] / maxch
                else:
                    a = ws[2 * size_ext + i - l + 1] / maxch
                b = a - nim
                a = 1.0 if a + nim <= 0 else math.sqrt(a + nim)
                b = b / a
                b = math.exp(b)
                sm = sm + b
            a = sp / sm
            ws[i + 1] = ws[i] * a
            a = ws[i + 1]
            nom = nom + a
        for i in range(xmin, xmax + 1):
            ws[i] = ws[i] / nom
        for j in range(size_ext):
            ws[size_ext + j] = ws[j] * plocha
        for j in range(size_ext):
            ws[2 * size_ext + j] = ws[size_ext + j]
        if backgroundRemove:
            for i in range(1, numberIterations + 1):
                for j in range(i, size_ext - i):
                    a = ws[size_ext + j]
                    b = (ws[size_ext + j - i] + ws[size_ext + j + i]) / 2.0
                    if b < a:
                        a = b
                    ws[j] = a
                for j in range(i, size_ext - i):
                    ws[size_ext + j] = ws[j]
            for j in range(size_ext):
                ws[size_ext + j] = ws[2 * size_ext + j] - ws[size_ext + j]
    area = 0.0
    lh_gold = -1
    posit = 0
    maximum = 0.0
    for i in range(size_ext):
        lda = float(i) - 3 * sigma
        lda = lda * lda / (2 * sigma * sigma)
        j = int(1000 * math.exp(-lda))
        lda = float(j)
        if lda != 0:
            lh_gold = i + 1
        ws[i] = lda
        area = area + lda
        if lda > maximum:
            maximum = lda
            posit = i
    for i in range(size_ext):
        ws[2 * size_ext + i] = abs(ws[size_ext + i])
    i = lh_gold - 1
    if i > size_ext:
        i = size_ext
    imin, imax = -i, i
    for i in range(imin, imax + 1):
        lda = 0.0
        jmin = -i if i < 0 else 0
        jmax = lh_gold - 1 - i
        if jmax > lh_gold - 1:
            jmax = lh_gold - 1
        for j in range(jmin, jmax + 1):
            lda = lda + ws[j] * ws[i + j]
        ws[size_ext + i - imin] = lda
    i = lh_gold - 1
    imin, imax = -i, size_ext + i - 1
    for i in range(imin, imax + 1):
        lda = 0.0
        for j in range(lh_gold):
            k = i + j
            if 0 <= k < size_ext:
                lda = lda + ws[j] * ws[2 * size_ext + k]
        ws[4 * size_ext + i - imin] = lda
    for i in range(imin, imax + 1):
        ws[2 * size_ext + i - imin] = ws[4 * size_ext + i - imin]
    for i in range(size_ext):
        ws[i] = 1.0
    for lindex in range(deconIterations):
        for i in range(size_ext):
            if abs(ws[2 * size_ext + i]) > 0.00001 and abs(ws[i]) > 0.00001:
                lda = 0.0
                jmin = lh_gold - 1
                if jmin > i:
                    jmin = i
                jmin = -jmin
                jmax = lh_gold - 1
                if jmax > size_ext - 1 - i:
                    jmax = size_ext - 1 - i
                for j in range(jmin, jmax + 1):
                    lda = lda + ws[j + lh_gold - 1 + size_ext] * ws[i + j]
                ldb = ws[2 * size_ext + i]
                lda = ldb / lda if lda != 0 else 0.0
                ldb = ws[i]
                lda = lda * ldb
                ws[3 * size_ext + i] = lda
        for i in range(size_ext):
            ws[i] = ws[3 * size_ext + i]
    for i in range(size_ext):
        lda = ws[i]
        j = (i + posit) % size_ext
        ws[size_ext + j] = lda
    maximum = 0.0
    maximum_decon = 0.0
    j = lh_gold - 1
    for i in range(size_ext - j):
        if shift <= i < ssize + shift:
            ws[i] = area * ws[size_ext + i + j]
            if maximum_decon < ws[i]:
                maximum_decon = ws[i]
            if maximum < ws[6 * size_ext + i]:
                maximum = ws[6 * size_ext + i]
        else:
            ws[i] = 0.0
    lda = 1.0
    if lda > threshold:
        lda = threshold
    lda = lda / 100
    peak_index = 0
    for i in range(1, size_ext - 1):
        if ws[i] > ws[i - 1] and ws[i] > ws[i + 1]:
            if shift <= i < ssize + shift:
                if ws[i] > lda * maximum_decon and ws[6 * size_ext + i] > threshold * maximum / 100.0:
                    a = 0.0
                    b = 0.0
                    for j in range(i - 1, i + 2):
                        a += float(j - shift) * ws[j]
                        b += ws[j]
                    a = a / b
                    if a < 0:
                        a = 0.0
                    if a >= ssize:
                        a = float(ssize - 1)
                    if peak_index == 0:
                        fPositionX[0] = a
                        peak_index = 1
                    else:
                        j = 0
                        priz = 0
                        while j < peak_index and priz == 0:
                            if ws[6 * size_ext + shift + int(a)] > ws[6 * size_ext + shift + int(fPositionX[j])]:
                                priz = 1
                            j += 1
                        if priz == 0:
                            if j < fMaxPeaks:
                                fPositionX[j] = a
                        else:
                            for k in range(peak_index, j - 1, -1):
                                if k < fMaxPeaks:
                                    fPositionX[k] = fPositionX[k - 1]
                            fPositionX[j - 1] = a
                        if peak_index < fMaxPeaks:
                            peak_index += 1
    return peak_index


# ---- ROI_refinement -------------------------------------------------------

class RoiRefinement:
    """ROI_refinement for one plane (plane 0/1 induction, 2 collection);
    chid is the wire index within the plane."""

    def __init__(self, plane, nwire, th_factor=3.0, fake_signal_low_th=500, fake_signal_high_th=1000,
                 fake_signal_low_th_ind_factor=1.0, fake_signal_high_th_ind_factor=1.0, pad=5,
                 break_roi_loop=2, th_peak=3.0, sep_peak=6.0, low_peak_sep_threshold_pre=1200,
                 max_npeaks=200, sigma=2.0, th_percent=0.1, isWrapped=False):
        self.plane = plane
        self.nwire = nwire
        # float members of the reference class
        self.th_factor = F(th_factor)
        self.fake_signal_low_th = F(fake_signal_low_th)
        self.fake_signal_high_th = F(fake_signal_high_th)
        self.fake_signal_low_th_ind_factor = F(fake_signal_low_th_ind_factor)
        self.fake_signal_high_th_ind_factor = F(fake_signal_high_th_ind_factor)
        self.pad = int(pad)
        self.break_roi_loop = int(break_roi_loop)
        self.th_peak = F(th_peak)
        self.sep_peak = F(sep_peak)
        self.low_peak_sep_threshold_pre = F(low_peak_sep_threshold_pre)
        self.max_npeaks = int(max_npeaks)
        self.sigma = F(sigma)
        self.th_percent = F(th_percent)
        self.isWrapped = bool(isWrapped)
        self.rois_tight = [[] for _ in range(nwire)]
        self.rois_loose = [[] for _ in range(nwire)]
        self.front_rois = {}
        self.back_rois = {}
        self.contained_rois = {}
        self.rms = None

    def _lists(self):
        return self.rois_tight if self.plane == 2 else self.rois_loose

    # -- links (:230-267)
    def unlink(self, prev_roi, next_roi):
        if prev_roi in self.front_rois:
            t = self.front_rois[prev_roi]
            if next_roi in t:
                t.remove(next_roi)
        if next_roi in self.back_rois:
            t = self.back_rois[next_roi]
            if prev_roi in t:
                t.remove(prev_roi)

    def link(self, prev_roi, next_roi):
        if prev_roi in self.front_rois:
            t = self.front_rois[prev_roi]
            if next_roi not in t:
                t.append(next_roi)
        else:
            self.front_rois[prev_roi] = [next_roi]
        if next_roi in self.back_rois:
            t = self.back_rois[next_roi]
            if prev_roi not in t:
                t.append(prev_roi)
        else:
            self.back_rois[next_roi] = [prev_roi]

    @staticmethod
    def _push(m, key, val):
        if key in m:
            m[key].append(val)
        else:
            m[key] = [val]

    # -- load_data (:269-699)
    def load_data(self, r_data, rms, tight, loose=None, bad=None):
        """r_data (nw, nt) float32; rms per wire; tight/loose: per-wire
        lists of inclusive (start, end); bad: per-wire inclusive ranges."""
        self.rms = [F(v) for v in rms]
        nw, nt = r_data.shape
        for irow in range(nw):
            signal = [F(v) for v in r_data[irow]]
            if bad is not None and len(bad[irow]):
                for icol in range(nt):
                    for (lo, hi) in bad[irow]:
                        if lo <= icol <= hi:
                            signal[icol] = F(0)
                            break
            chid = irow
            for (s, e) in tight[irow]:
                tight_roi = SignalROI(self.plane, chid, s, e, signal)
                threshold = self.rms[irow] * self.th_factor
                if len(tight_roi.get_above_threshold(threshold)) == 0:
                    continue
                self.rois_tight[chid].append(tight_roi)
                if self.plane == 2 and chid > 0:
                    if chid != nw / 2. or not self.isWrapped:
                        for prev_roi in self.rois_tight[chid - 1]:
                            if tight_roi.overlap(prev_roi):
                                self._push(self.front_rois, prev_roi, tight_roi)
                                self._push(self.back_rois, tight_roi, prev_roi)
            if self.plane != 2:
                for (s, e) in loose[irow]:
                    loose_roi = SignalROI(self.plane, chid, s, e, signal)
                    threshold = self.rms[irow] * self.th_factor
                    if len(loose_roi.get_above_threshold(threshold)) == 0:
                        continue
                    self.rois_loose[chid].append(loose_roi)
                    if chid > 0:
                        for prev_roi in self.rois_loose[chid - 1]:
                            if loose_roi.overlap(prev_roi):
                                self._push(self.front_rois, prev_roi, loose_roi)
                                self._push(self.back_rois, loose_roi, prev_roi)
                    if chid == nw - 1 and self.isWrapped:
                        for next_roi in self.rois_loose[0]:
                            if loose_roi.overlap(next_roi):
                                self._push(self.back_rois, next_roi, loose_roi)
                                self._push(self.front_rois, loose_roi, next_roi)
                    for tight_roi in self.rois_tight[chid]:
                        if tight_roi.overlap(loose_roi):
                            self._push(self.contained_rois, loose_roi, tight_roi)

    # -- CleanUpROIs (:701-873)
    def CleanUpROIs(self):
        if self.plane == 2:
            return
        saved = set()
        for i in range(self.nwire):
            for roi in self.rois_loose[i]:
                if roi in saved:
                    continue
                if roi in self.contained_rois:
                    temp_rois = [roi]
                    saved.add(roi)
                    while temp_rois:
                        temp_roi = temp_rois.pop()
                        if temp_roi in self.front_rois:
                            for r1 in self.front_rois[temp_roi]:
                                if r1 not in saved:
                                    temp_rois.append(r1)
                                    saved.add(r1)
                        if temp_roi in self.back_rois:
                            for r1 in self.back_rois[temp_roi]:
                                if r1 not in saved:
                                    temp_rois.append(r1)
                                    saved.add(r1)
        for i in range(self.nwire):
            to_be_removed = []
            for roi in self.rois_loose[i]:
                if roi not in saved:
                    to_be_removed.append(roi)
                    if roi in self.front_rois:
                        for r1 in self.front_rois[roi]:
                            self.back_rois.setdefault(r1, []).remove(roi)
                        del self.front_rois[roi]
                    if roi in self.back_rois:
                        for r1 in self.back_rois[roi]:
                            self.front_rois.setdefault(r1, []).remove(roi)
                        del self.back_rois[roi]
            for roi in to_be_removed:
                self.rois_loose[i].remove(roi)

    # -- generate_merge_ROIs (:875-1159)
    def generate_merge_ROIs(self):
        if self.plane == 2:
            return
        nw = self.nwire
        for i in range(nw):
            covered = set()
            for roi in self.rois_loose[i]:
                if roi in self.contained_rois:
                    for r1 in self.contained_rois[roi]:
                        covered.add(r1)
            saved_rois = [roi for roi in self.rois_tight[i] if roi not in covered]
            for roi in saved_rois:
                loose_roi = SignalROI(0, 0, 0, 0, copy_of=roi)
                self.rois_loose[i].append(loose_roi)
                self.contained_rois[loose_roi] = [roi]
                if i < nw - 1:
                    for next_roi in self.rois_loose[i + 1]:
                        if loose_roi.overlap(next_roi):
                            self._push(self.back_rois, next_roi, loose_roi)
                            self._push(self.front_rois, loose_roi, next_roi)
                if i == nw - 1 and self.isWrapped:
                    for next_roi in self.rois_loose[0]:
                        if loose_roi.overlap(next_roi):
                            self._push(self.back_rois, next_roi, loose_roi)
                            self._push(self.front_rois, loose_roi, next_roi)
                if i > 0:
                    for prev_roi in self.rois_loose[i - 1]:
                        if loose_roi.overlap(prev_roi):
                            self._push(self.front_rois, prev_roi, loose_roi)
                            self._push(self.back_rois, loose_roi, prev_roi)
                if i == 0 and self.isWrapped:
                    for prev_roi in self.rois_loose[nw - 1]:
                        if loose_roi.overlap(prev_roi):
                            self._push(self.front_rois, prev_roi, loose_roi)
                            self._push(self.back_rois, loose_roi, prev_roi)

    # -- CheckROIs (:1161-1284)
    def CheckROIs(self):
        if self.plane == 2:
            return
        for i in range(self.nwire):
            for roi in self.rois_loose[i]:
                th = self.th_factor * self.rms[roi.chid]
                if roi in self.front_rois:
                    temp_rois = [r1 for r1 in self.front_rois[roi]
                                 if not roi.overlap_th(r1, th, self.th_factor * self.rms[r1.chid])]
                    for r1 in temp_rois:
                        self.unlink(roi, r1)
                if roi in self.back_rois:
                    temp_rois = [r1 for r1 in self.back_rois[roi]
                                 if not roi.overlap_th(r1, th, self.th_factor * self.rms[r1.chid])]
                    for r1 in temp_rois:
                        self.unlink(r1, roi)

    # -- CleanUpCollectionROIs (:1286-1366), CleanUpInductionROIs (:1368-1601)
    def _cleanup_fake(self, lists, use_mean, threshold, mean_threshold):
        good = set()
        for i in range(self.nwire):
            for roi in lists[i]:
                if len(roi.get_above_threshold(threshold)) != 0 or (
                        use_mean and roi.get_average_heights() > mean_threshold):
                    good.add(roi)
        bad = []
        for i in range(self.nwire):
            for roi in lists[i]:
                if roi in good:
                    continue
                if roi in self.front_rois and any(r1 in good for r1 in self.front_rois[roi]):
                    continue
                if roi in self.back_rois and any(r1 in good for r1 in self.back_rois[roi]):
                    continue
                bad.append(roi)
        for roi in bad:
            if roi in self.front_rois:
                for r1 in list(self.front_rois[roi]):
                    self.unlink(roi, r1)
                del self.front_rois[roi]
            if roi in self.back_rois:
                for r1 in list(self.back_rois[roi]):
                    self.unlink(r1, roi)
                del self.back_rois[roi]
            if roi in lists[roi.chid]:
                lists[roi.chid].remove(roi)

    def CleanUpCollectionROIs(self):
        self._cleanup_fake(self.rois_tight, True, self.fake_signal_high_th, self.fake_signal_low_th)

    def CleanUpInductionROIs(self):
        mean_threshold = self.fake_signal_low_th * self.fake_signal_low_th_ind_factor
        threshold = self.fake_signal_high_th * self.fake_signal_high_th_ind_factor
        self._cleanup_fake(self.rois_loose, self.plane == 0, threshold, mean_threshold)

    # -- replacing an ROI (ShrinkROI :1783-1841, BreakROI1 :2381-2445)
    def _replace(self, roi, new_rois):
        lst = self.rois_loose[roi.chid]
        lst.remove(roi)
        lst.extend(new_rois)
        if roi in self.front_rois:
            next_rois = list(self.front_rois[roi])
            for n in next_rois:
                self.unlink(roi, n)
                for nr in new_rois:
                    if nr.overlap(n):
                        self.link(nr, n)
            del self.front_rois[roi]
        if roi in self.back_rois:
            prev_rois = list(self.back_rois[roi])
            for p in prev_rois:
                self.unlink(p, roi)
                for nr in new_rois:
                    if nr.overlap(p):
                        self.link(p, nr)
            del self.back_rois[roi]
        if roi in self.contained_rois:
            tight_rois = list(self.contained_rois[roi])
            for t in tight_rois:
                for nr in new_rois:
                    if nr.overlap(t):
                        self._push(self.contained_rois, nr, t)
            del self.contained_rois[roi]

    # -- ShrinkROI (:1603-1842), ShrinkROIs (:1844-1865)
    def ShrinkROI(self, roi):
        start_bin, end_bin = roi.start_bin, roi.end_bin
        if start_bin < 0 or end_bin < 0:
            return
        chid = roi.chid
        contents = roi.contents
        threshold1 = self.rms[chid] * self.th_factor
        temp_signal = [F(0)] * (end_bin - start_bin + 1)
        if roi in self.contained_rois:
            for tight_roi in self.contained_rois[roi]:
                for i in range(tight_roi.start_bin, tight_roi.end_bin + 1):
                    if 0 <= i - start_bin < len(temp_signal):
                        temp_signal[i - start_bin] = F(1)
        for m in (self.front_rois, self.back_rois):
            if roi in m:
                for r1 in m[roi]:
                    threshold = self.rms[r1.chid] * self.th_factor
                    for (a, b) in r1.get_above_threshold(threshold):
                        for j in range(a, b + 1):
                            k = j + r1.start_bin - start_bin
                            if 0 <= k < len(temp_signal):
                                if contents[k] > threshold1:
                                    temp_signal[k] = F(1)
        new_start_bin = start_bin
        new_end_bin = end_bin
        for i in range(len(temp_signal)):
            if temp_signal[i] > 0:
                new_start_bin = i + start_bin
                break
        for i in range(len(temp_signal) - 1, -1, -1):
            if temp_signal[i] > 0:
                new_end_bin = i + start_bin
                break
        new_start_bin -= self.pad
        new_end_bin += self.pad
        if new_start_bin < start_bin:
            new_start_bin = start_bin
        if new_end_bin > end_bin:
            new_end_bin = end_bin
        signal = [F(0)] * (end_bin + 1)
        for i in range(new_start_bin, new_end_bin + 1):
            signal[i] = contents[i - start_bin]
        new_rois = []
        if new_start_bin >= 0 and new_end_bin > new_start_bin:
            new_rois.append(SignalROI(roi.plane, chid, new_start_bin, new_end_bin, signal))
        self._replace(roi, new_rois)

    def ShrinkROIs(self):
        if self.plane == 2:
            return
        all_rois = [roi for i in range(self.nwire) for roi in self.rois_loose[i]]
        for roi in all_rois:
            self.ShrinkROI(roi)

    # -- BreakROI (:1867-2328)
    def BreakROI(self, roi, rms):
        start_bin, end_bin = roi.start_bin, roi.end_bin
        if start_bin < 0 or end_bin < 0:
            return
        temp_signal = list(roi.contents)
        low_peak_sep_threshold = self.low_peak_sep_threshold_pre
        if low_peak_sep_threshold < self.sep_peak * rms:
            low_peak_sep_threshold = self.sep_peak * rms
        px, py = find_peak(temp_signal, self.max_npeaks, float(self.sigma), float(self.th_percent))
        nfound = len(px)
        if nfound > 1:
            npeaks = nfound
            order_peak_pos = []
            npeaks_threshold = 0
            for j in range(npeaks):
                order_peak_pos.append(int(px[j] + start_bin))
                if py[j] > float(self.th_peak * rms):
                    npeaks_threshold += 1
            if npeaks_threshold > 1:
                order_peak_pos.sort()
                valley_pos = [F(0)] * 205
                valley_pos[0] = F(start_bin)
                mn = F(1e9)
                for k in range(0, order_peak_pos[0] - start_bin):
                    if temp_signal[k] < mn:
                        mn = temp_signal[k]
                        valley_pos[0] = F(k + start_bin)
                if valley_pos[0] != start_bin:
                    order_peak_pos.append(0)
                    for j in range(npeaks - 1, -1, -1):
                        order_peak_pos[j + 1] = order_peak_pos[j]
                    npeaks += 1
                    order_peak_pos[0] = start_bin
                    j = start_bin
                    while j != valley_pos[0]:
                        if temp_signal[j - start_bin] > temp_signal[order_peak_pos[0] - start_bin]:
                            order_peak_pos[0] = j
                        j += 1
                    valley_pos[0] = F(start_bin)
                for j in range(npeaks - 1):
                    mn = F(1e9)
                    valley_pos[j + 1] = F(order_peak_pos[j])
                    for k in range(order_peak_pos[j] - start_bin, order_peak_pos[j + 1] - start_bin):
                        if temp_signal[k] < mn:
                            mn = temp_signal[k]
                            valley_pos[j + 1] = F(k + start_bin)
                valley_pos[npeaks] = F(end_bin)
                mn = F(1e9)
                for k in range(order_peak_pos[npeaks - 1] - start_bin, end_bin - start_bin + 1):
                    if temp_signal[k] < mn:
                        mn = temp_signal[k]
                        valley_pos[npeaks] = F(k + start_bin)
                if valley_pos[npeaks] != end_bin:
                    npeaks += 1
                    valley_pos[npeaks] = F(end_bin)
                    order_peak_pos.append(end_bin)
                    j = int(valley_pos[npeaks - 1])
                    while j != valley_pos[npeaks]:
                        if temp_signal[j - start_bin] > temp_signal[order_peak_pos[npeaks - 1] - start_bin]:
                            order_peak_pos[npeaks - 1] = j
                        j += 1

                def T(pos):
                    return temp_signal[int(pos) - start_bin]

                valley_pos1 = [F(0)] * 205
                peak_pos1 = [F(0)] * 205
                npeaks1 = 0
                valley_pos1[0] = valley_pos[0]
                j = 0
                while j < npeaks:
                    if npeaks1 > 0:
                        if T(valley_pos[j]) < T(valley_pos1[npeaks1]):
                            valley_pos1[npeaks1] = valley_pos[j]
                    if T(order_peak_pos[j]) - T(valley_pos1[npeaks1]) > low_peak_sep_threshold:
                        peak_pos1[npeaks1] = F(order_peak_pos[j])
                        npeaks1 += 1
                        flag1 = 0
                        k = j + 1
                        while k != npeaks + 1:
                            if k <= npeaks:
                                if T(order_peak_pos[k - 1]) > T(peak_pos1[npeaks1 - 1]):
                                    peak_pos1[npeaks1 - 1] = F(order_peak_pos[k - 1])
                            if T(peak_pos1[npeaks1 - 1]) - T(valley_pos[k]) > low_peak_sep_threshold:
                                valley_pos1[npeaks1] = valley_pos[k]
                                j = k - 1
                                flag1 = 1
                                break
                            k += 1
                        if flag1 == 0:
                            valley_pos1[npeaks1] = valley_pos[npeaks]
                            j = npeaks
                    j += 1
                valley_pos1[npeaks1] = valley_pos[npeaks]
                if npeaks1 == 0:
                    npeaks1 = 1
                    valley_pos1[0] = valley_pos[0]
                    valley_pos1[1] = valley_pos[npeaks]
                saved_boundaries = set()
                for j in range(npeaks1):
                    saved_boundaries.add(int(valley_pos1[j]))
                    saved_boundaries.add(int(valley_pos1[j + 1]))
                temp1_signal = list(temp_signal)
                temp_signal = [F(0)] * len(temp1_signal)
                for j in range(npeaks1):
                    start_pos = int(valley_pos1[j])
                    start_content = float(temp1_signal[int(valley_pos1[j]) - start_bin])
                    end_pos = int(valley_pos1[j + 1])
                    end_content = float(temp1_signal[int(valley_pos1[j + 1]) - start_bin])
                    if start_pos in saved_boundaries or end_pos in saved_boundaries:
                        k = start_pos
                        while k != end_pos + 1:
                            if end_pos == start_pos:   # 0/0 in double
                                temp_content = math.nan
                            else:
                                temp_content = float(temp1_signal[k - start_bin]) - (
                                    start_content + (end_content - start_content) * (k - start_pos)
                                    / (end_pos - start_pos))
                            temp_signal[k - start_bin] = F(temp_content)
                            k += 1
                if roi not in self.contained_rois:   # the reference's operator[] creates the entry
                    self.contained_rois[roi] = []
                for temp_roi in self.contained_rois[roi]:
                    temp_flag = 0
                    for i in range(temp_roi.start_bin, temp_roi.end_bin + 1):
                        if 0 <= i - roi.start_bin < len(temp_signal):
                            if temp_signal[i - roi.start_bin] != 0:
                                temp_flag = 1
                                break
                    if temp_flag == 0:
                        for i in range(temp_roi.start_bin, temp_roi.end_bin + 1):
                            if 0 <= i - roi.start_bin < len(temp_signal):
                                temp_signal[i - roi.start_bin] = temp_roi.contents[i - temp_roi.start_bin]
        for qx in range(2):
            th = self.th_factor * rms
            bins = []
            i = 0
            n = len(temp_signal)
            while i < n:
                if temp_signal[i] < th:
                    start = end = i
                    for j in range(i + 1, n):
                        if temp_signal[j] < th:
                            end = j
                        else:
                            break
                    bins.append((start, end))
                    i = end
                i += 1
            saved_b = []
            for (start, end) in bins:
                mn = F(1e9)
                bin_min = start
                for j in range(start, end + 1):
                    if abs(float(temp_signal[j])) < 1e-3:
                        bin_min = j
                        break
                    elif temp_signal[j] < mn:
                        mn = temp_signal[j]
                        bin_min = j
                saved_b.append(bin_min)
            temp1_signal = list(temp_signal)
            temp_signal = [F(0)] * len(temp1_signal)
            for j in range(len(saved_b) - 1):
                start_pos = saved_b[j]
                start_content = temp1_signal[start_pos]
                end_pos = saved_b[j + 1]
                end_content = temp1_signal[end_pos]
                for k in range(start_pos, end_pos + 1):
                    temp_content = temp1_signal[k] - (start_content + (end_content - start_content) * F(k - start_pos)
                                                      / F(end_pos - start_pos))
                    temp_signal[k] = F(temp_content)
        roi.contents = list(temp_signal)

    # -- BreakROI1 (:2330-2446)
    def BreakROI1(self, roi):
        start_bin, end_bin = roi.start_bin, roi.end_bin
        if start_bin < 0 or end_bin < 0:
            return
        temp_signal = list(roi.contents)
        bins = [i + start_bin for i in range(len(temp_signal)) if abs(float(temp_signal[i])) < 1e-3]
        new_rois = []
        for i in range(len(bins) - 1):
            start_bin1 = bins[i]
            end_bin1 = bins[i + 1]
            signal = [F(0)] * (end_bin + 1)
            for j in range(start_bin1, end_bin1 + 1):
                signal[j] = temp_signal[j - start_bin]
            if start_bin1 >= 0 and end_bin1 > start_bin1:
                new_rois.append(SignalROI(roi.plane, roi.chid, start_bin1, end_bin1, signal))
        self._replace(roi, new_rois)

    def BreakROIs(self):
        if self.plane == 2:
            return
        all_rois = []
        for i in range(self.nwire):
            for roi in self.rois_loose[i]:
                self.BreakROI(roi, self.rms[i])
                all_rois.append(roi)
        for roi in all_rois:
            self.BreakROI1(roi)

    # -- ExtendROIs (:2946-3136)
    def ExtendROIs(self):
        lists = self._lists()
        for chid in range(self.nwire):
            lists[chid].sort(key=lambda r: r.start_bin)   # stable, as std::list::sort
            for roi in lists[chid]:
                roi.ext_start_bin = roi.start_bin
                roi.ext_end_bin = roi.end_bin
                for m in (self.front_rois, self.back_rois):
                    for r1 in m.get(roi, []):
                        if roi.ext_start_bin > r1.start_bin:
                            roi.ext_start_bin = r1.start_bin
                        if roi.ext_end_bin < r1.end_bin:
                            roi.ext_end_bin = r1.end_bin
            prev_roi = None
            for roi in lists[chid]:
                if prev_roi is not None:
                    if prev_roi.ext_end_bin > roi.ext_start_bin:
                        mid = int(prev_roi.end_bin * 0.5 + roi.start_bin * 0.5)
                        prev_roi.ext_end_bin = mid
                        roi.ext_start_bin = mid
                prev_roi = roi

    # -- OmnibusSigProc.cxx:1506-1590 for one plane (no multi-plane protection)
    def refine(self):
        self.CleanUpROIs()
        self.generate_merge_ROIs()
        for qx in range(self.break_roi_loop):
            self.BreakROIs()
            self.CheckROIs()
            self.CleanUpROIs()
        self.ShrinkROIs()
        self.CheckROIs()
        self.CleanUpROIs()
        if self.plane == 2:
            self.CleanUpCollectionROIs()
        else:
            self.CleanUpInductionROIs()
        self.ExtendROIs()

    def rois(self, extended=True):
        out = []
        for lst in self._lists():
            out.append([(r.ext_start_bin, r.ext_end_bin) if extended else (r.start_bin, r.end_bin) for r in lst])
        return out

    def contents(self, wire):
        return [list(r.contents) for r in self._lists()[wire]]
