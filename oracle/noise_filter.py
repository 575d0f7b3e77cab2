"""Restatement of Wire-Cell's per-channel noise filters: OmnibusNoiseFilter
(sigproc/src/OmnibusNoiseFilter.cxx:95-260) with Protodune::OneChannelNoise
(sigproc/src/Protodune.cxx:753-918) and RelGainCalib (:975-1009),
Microboone::OneChannelNoise (sigproc/src/Microboone.cxx:902-1064), the
coherent subtraction (:53-462, :798-872), OneChannelStatus (:1263-1334),
ADCBitShift (:1106-1216) and the parts of OmniChannelNoiseDB they read
(sigproc/src/OmniChannelNoiseDB.cxx).

TEST INFRASTRUCTURE ONLY.  Only tests/ and the CPU-baseline leg of the
benchmark may import this module, and only as the checker.  Float
arithmetic of the reference is done in numpy float32 scalars and arrays,
double arithmetic in float64; the reference's FFTW float transforms are
done in complex128 and rounded to complex64 where the reference holds a
float spectrum.  Pure-Python loops follow the reference's sequential
algorithms, so this is for test-sized inputs (tens of channels).
"""
import math

import numpy as np

F = np.float32
FLAG = F(10000.0)


# ---- util/src/Waveform.cxx --------------------------------------------

def mean_rms(wf):
    """Waveform::mean_rms (Waveform.cxx:19-37): double sums, population RMS."""
    n = len(wf)
    if n == 0:
        return 0.0, 0.0
    if n == 1:
        return float(wf[0]), 0.0
    s = s2 = 0.0
    for x in np.asarray(wf, dtype=np.float64):
        s += x
        s2 += x * x
    mean = s / n
    v = (s2 - s * s / n) / n
    return mean, math.sqrt(v) if v >= 0 else float("nan")


def percentile_binned(wave, p):
    """Waveform::percentile_binned (Waveform.cxx:108-131): an n-bin float
    histogram between min and max, the lower edge of the bin where the
    count passes n p."""
    w = np.asarray(wave, dtype=F)
    vmin, vmax = F(w.min()), F(w.max())
    nbins = len(w)
    binsize = F((vmax - vmin) / F(nbins))
    hist = np.zeros(nbins, dtype=np.int64)
    with np.errstate(divide="ignore", invalid="ignore"):
        for val in w:
            r = F(F(val - vmin) / binsize)
            if np.isnan(r):
                b = -2 ** 31          # x86 conversion of NaN
            else:
                b = int(math.floor(abs(float(r)) + 0.5)) * (1 if r >= 0 else -1)   # round half away
            b = max(0, b)
            b = min(nbins - 1, b)
            hist[b] += 1
    imed = int(F(F(len(w)) * F(p)))   # size_t * float, in float
    count = 0
    for ind in range(nbins):
        count += hist[ind]
        if count > imed:
            return F(vmin + F(F(ind) * binsize))
    return F(vmin + (vmax - vmin) * F(p))


def median_binned(wave):
    return percentile_binned(wave, 0.5)


# ---- sigproc/src/Microboone.cxx helpers ------------------------------

def calc_rms_with_flags(sig):
    """Microboone::CalcRMSWithFlags (:545-566): samples below 4096 only."""
    temp = [x for x in np.asarray(sig, dtype=F) if x < 4096]
    if not temp:
        return F(0.0)
    # the reference passes the double 0.5 -+ 0.34 to a float parameter
    p0 = percentile_binned(temp, F(0.5 - 0.34))
    p1 = percentile_binned(temp, F(0.5))
    p2 = percentile_binned(temp, F(0.5 + 0.34))
    return F(math.sqrt((float(p2 - p1) ** 2 + float(p1 - p0) ** 2) / 2.0))


def signal_filter(sig):
    """Microboone::SignalFilter (:568-610), in place: samples within 8 ticks
    of one above 4 RMS get 20000 added (once)."""
    sig = np.asarray(sig, dtype=F)
    thr = F(4.0 * float(calc_rms_with_flags(sig)))
    n = len(sig)
    region = [(x > thr or float(x) < -1.0 * float(thr)) and x < 4096.0 for x in sig]
    for i in range(n):
        if region[i]:
            for j in range(max(i - 8, 0), min(i + 8, n)):
                if sig[j] < 4096.0:
                    sig[j] = F(float(sig[j]) + 20000.0)
    return sig


def raw_adaptive_baseline(sig):
    """Microboone::RawAdapativeBaselineAlg (:612-737), in place."""
    sig = np.asarray(sig, dtype=F)
    W, n = 20, len(sig)
    minw = W // 2
    if all(x == FLAG for x in sig):
        return sig
    base = [0.0] * n
    filled = [False] * n
    bval, wb = 0.0, 0
    for j in range(0, W // 2 + 1):
        if sig[j] < 4096.0:
            bval += float(sig[j])
            wb += 1
    base[0] = 0.0 if wb == 0 else bval / wb
    filled[0] = wb >= minw
    for j in range(1, n):
        old, new = j - W // 2 - 1, j + W // 2
        if old >= 0 and sig[old] < 4096.0:
            bval -= float(sig[old])
            wb -= 1
        if new < n and sig[new] < 4096:
            bval += float(sig[new])
            wb += 1
        base[j] = 0.0 if wb == 0 else bval / wb
        filled[j] = wb >= minw
    for j in range(n):
        x = sig[j]
        if x != FLAG:
            if not filled[j]:
                d = j
                while not filled[d] and d > 0 and sig[d] != FLAG:
                    d -= 1
                dflag = not filled[d]
                u = j
                while not filled[u] and u < n - 1 and sig[u] != FLAG:
                    u += 1
                uflag = not filled[u]
                if not dflag and not uflag:
                    base[j] = ((j - d) * base[d] + (u - j) * base[u]) / float(u - d)
                elif dflag and not uflag:
                    base[j] = base[u]
                elif not dflag and uflag:
                    base[j] = base[d]
                else:
                    base[j] = 0.0
            sig[j] = F(float(x) - base[j])
    return sig


def remove_filter_flags(sig):
    """Microboone::RemoveFilterFlags (:739-757), in place."""
    sig = np.asarray(sig, dtype=F)
    for i in range(len(sig)):
        x = float(sig[i])
        if x > 4096.0:
            sig[i] = F(x - 20000.0) if x > 10000.0 else F(0.0)
    return sig


def noisy_filter(sig, min_rms, max_rms):
    """Microboone::NoisyFilterAlg (:464-533), in place; True if noisy."""
    sig = np.asarray(sig, dtype=F)
    rms = float(calc_rms_with_flags(sig))
    if rms > float(F(max_rms)) or rms < float(F(min_rms)):
        sig[:] = FLAG
        return True
    return False


# ---- sigproc/src/Diagnostics.cxx -----------------------------------

def is_partial(spec, nfreqs=4, maxpower=6000.0):
    """Diagnostics::Partial (:15-31): bins 1..nfreqs+1 falling and their
    mean magnitude above maxpower."""
    mag0 = abs(complex(spec[1]))
    s = mag0
    for ind in range(1, nfreqs + 1):
        if ind >= len(spec):
            break
        magi = abs(complex(spec[ind + 1]))
        if mag0 <= magi:
            return False
        s += magi
    return s / (nfreqs + 1) > maxpower


# ---- OmniChannelNoiseDB spectra ----------------------------------------

def rcrc_spectrum(rcrc, tick, nsamples, layers=1):
    """parse_rcrc (OmniChannelNoiseDB.cxx:171-208): Response::SimpleRC
    sampled at t = i tick (util/src/Response.cxx:517-536, :267-276), its
    float spectrum raised to the power rc_layers."""
    t = np.arange(nsamples) * ((nsamples * tick) / nsamples)
    w = np.where(rcrc > 0, -tick / rcrc * np.exp(-t / rcrc), 0.0)
    w = w + (t < tick)
    spec = np.fft.fft(w.astype(F).astype(np.float64)).astype(np.complex64)
    out = spec.copy()
    for _ in range(layers - 1):
        out = (out * spec).astype(np.complex64)
    return out


def freqmask_spectrum(masks, nsamples):
    """parse_freqmasks (:152-169): ones, with [lobin, hibin] (inclusive)
    set to value, in list order."""
    s = np.ones(nsamples, dtype=np.complex64)
    for value, lo, hi in masks:
        lo, hi = max(lo, 0), min(hi, nsamples - 1)
        s[lo:hi + 1] = value
    return s


def flagged_bins(noise):
    """The runs of bins whose spectrum value has a real part below 0.5
    (Protodune.cxx:791-806), as inclusive (first, last) pairs."""
    bins = []
    for i, v in enumerate(noise):
        if np.real(v) < 0.5:
            if bins and i == bins[-1][1] + 1:
                bins[-1][1] = i
            else:
                bins.append([i, i])
    return [tuple(b) for b in bins]


# ---- DftTools --------------------------------------------------------

def fwd_r2c(x):
    return np.fft.fft(np.asarray(x, dtype=F).astype(np.float64)).astype(np.complex64)


def inv_c2r(spec):
    """DftTools::inv_c2r (aux/src/DftTools.cxx:39-50) with
    Spectrum::hermitian_mirror (Spectrum.h:56-89)."""
    s = np.asarray(spec, dtype=np.complex128).copy()
    n = len(s)
    s[0] = s[0].real
    half = n // 2
    if n % 2 == 0:
        s[half] = s[half].real
        s[half + 1:] = np.conj(s[1:half][::-1])
    else:
        s[half + 1:] = np.conj(s[1:half + 1][::-1])
    return np.fft.ifft(s).real.astype(F)


def fft_scaling(signal, nsamples):
    """Protodune::FftScaling (:532-552)."""
    nsig = len(signal)
    tran = list(fwd_r2c(signal)) + [0j] * max(0, nsamples - nsig)
    tran = np.asarray(tran[:nsamples], dtype=np.complex64)
    mid = nsig // 2 if nsig % 2 == 0 else (nsig + 1) // 2
    tran = np.concatenate([tran[:mid], tran[nsig:], tran[mid:nsig]])
    out = inv_c2r(tran)
    return (out * F(nsamples // nsig)).astype(F)


# ---- Protodune::OneChannelNoise::apply ---------------------------------

def one_channel_noise(signal, plane, rcrc=None, noise=None, min_rms=0.5, max_rms=10.0, resmp=None):
    """Returns (filtered signal, {mask name: True}) for one channel
    (Protodune.cxx:753-918).  rcrc, noise: the DB spectra (None: empty)."""
    sig = np.asarray(signal, dtype=F).copy()
    masks = {}
    if resmp is not None:
        smpout = len(sig)
        sig = fft_scaling(sig[:resmp], smpout)
    spectrum = fwd_r2c(sig)
    partial = is_partial(spectrum)
    if not partial and rcrc is not None and len(rcrc) == len(spectrum):
        with np.errstate(divide="ignore", invalid="ignore"):
            spectrum = (spectrum.astype(np.complex128) / np.asarray(rcrc, np.complex128)).astype(np.complex64)
    if plane == 2:
        mag = np.abs(spectrum.astype(np.complex128)).astype(F)
        mag[0] = 0
        mag = raw_adaptive_baseline(mag)
        spec_n = [] if noise is None else noise
        fb = flagged_bins(spec_n)
        n_harmonic = 0
        m = len(mag)
        for _ in range(5):
            for lo, hi in fb:
                iend = hi + 1
                mean, rms = mean_rms(mag[lo:iend])
                cut = mean + 2.7 * rms
                if lo > 1050:
                    cut = mean + 3 * rms
                for j in range(lo, iend):
                    if float(mag[j]) > cut:
                        spectrum[j] = 0
                        spectrum[m + 1 - j] = 0
                        mag[j] = 0
                        mag[m + 1 - j] = 0
                        n_harmonic += 1
        if n_harmonic > 5:
            masks["harmonic"] = True
    spectrum[0] = 0
    sig = inv_c2r(spectrum)
    mean, rms = mean_rms(sig)
    temp = sig.copy()
    for i in range(len(temp)):
        if abs(float(temp[i]) - mean) > 6 * rms:
            temp[i] = F(mean)
    baseline = median_binned(temp)
    sig = (sig + F(-baseline)).astype(F)
    if partial:
        if plane != 2:
            masks["lf_noisy"] = True
        signal_filter(sig)
        raw_adaptive_baseline(sig)
        remove_filter_flags(sig)
    signal_filter(sig)
    noisy = noisy_filter(sig, min_rms, max_rms)
    remove_filter_flags(sig)
    if noisy:
        masks["noisy"] = True
    return sig, masks


# ---- MicroBooNE's one-channel filter -------------------------------------

def chirp(sig, window=20, min_rms=0.9, max_frac=0.20):
    """Diagnostics::Chirp (sigproc/src/Diagnostics.cxx:33-137): RMS of
    consecutive 20-sample windows (double), the low-RMS run bounds; returns
    (is_chirp, beg, end) with [beg, end) the chirped samples."""
    n = len(sig)
    counter, mean, rms2 = 0, 0.0, 0.0
    num_low, first, last, lowflag = 0, -1, -1, False
    r1 = r2 = r3 = 0.0
    normal = 0
    for ibin in range(n):
        x = float(sig[ibin])
        mean += x
        rms2 += x * x
        counter += 1
        if counter == window:
            mean /= window
            rms2 /= window
            v = rms2 - mean * mean
            rms = math.sqrt(v) if v >= 0 else float("nan")
            r1, r2, r3 = r2, r3, rms
            if rms < min_rms:
                num_low += 1
            if ibin >= 3 * window - 1:
                if r2 < min_rms and (r1 > min_rms or r3 > min_rms):
                    normal += 1
                if not lowflag:
                    if r2 < min_rms and r3 < min_rms:
                        lowflag = True
                        first, last = ibin - 2 * window + 1, ibin - window + 1
                    if ibin == 3 * window - 1 and r1 < min_rms and r2 < min_rms:
                        lowflag = True
                        first, last = ibin - 3 * window + 1, ibin - 2 * window + 1
                elif r2 < min_rms and r3 < min_rms:
                    last = ibin - window + 1
            counter, mean, rms2 = 0, 0.0, 0.0
    chirp_frac = num_low / (n / window)
    normal_frac = normal / num_low if num_low else float("nan")
    if num_low > 4 and (normal_frac < max_frac or (num_low < 2.0 / max_frac and last - first == num_low * window)):
        first = max(1, first - window)
        last = min(n, last + 2 * window)
        if n - last < window:
            last = n
        if chirp_frac > 0.99:
            first, last = 1, n
        return True, first - 1, last
    return False, -1, -1


def chirp_raise_baseline(sig, bin1, bin2):
    """Microboone::Chirp_raise_baseline (:535-543), in place."""
    sig[max(bin1, 0):min(bin2, len(sig))] = FLAG
    return sig


def mb_one_channel_noise(signal, plane, baseline=0.0, gain=1.0, rcrc=None, config=None, noise=None,
                         min_rms=0.5, max_rms=10.0):
    """Microboone::OneChannelNoise::apply (sigproc/src/Microboone.cxx:902-1064);
    returns (signal, {mask name: [(first, second), ...]}).  rcrc, config
    and noise are the DB spectra (None: not applied; rcrc only when the
    waveform is not partial)."""
    sig = (np.asarray(signal, dtype=F) + F(F(baseline) * -1)).astype(F)
    n = len(sig)
    sig_gc = (sig * F(gain)).astype(F)
    masks = {}
    is_chirp, beg, end = chirp(sig_gc)
    if is_chirp:
        masks.setdefault("chirp", []).append((beg, end))
        if plane != 2 and (beg > 0 or end < n):
            masks.setdefault("lf_noisy", []).append((beg, end))
    spectrum = fwd_r2c(sig)
    partial = is_partial(spectrum)
    if not partial and rcrc is not None:
        spectrum = (spectrum.astype(np.complex128) / np.asarray(rcrc, np.complex128)).astype(np.complex64)
    if config is not None:
        spectrum = (spectrum * np.asarray(config, np.complex64)).astype(np.complex64)
    if noise is not None:
        spectrum = (spectrum * np.asarray(noise, np.complex64)).astype(np.complex64)
    spectrum[0] = 0
    sig = inv_c2r(spectrum)
    mean, rms = mean_rms(sig)
    temp = sig.copy()
    for i in range(n):
        if abs(float(temp[i]) - mean) > 6 * rms:
            temp[i] = F(mean)
    base = median_binned(temp)
    sig = (sig + F(-base)).astype(F)
    if is_chirp:
        chirp_raise_baseline(sig, beg, end)
        signal_filter(sig)
        raw_adaptive_baseline(sig)
    if partial:
        if plane != 2:
            masks.setdefault("lf_noisy", []).append((0, n))
        signal_filter(sig)
        raw_adaptive_baseline(sig)
    signal_filter(sig)
    noisy = noisy_filter(sig, min_rms, max_rms)
    remove_filter_flags(sig)
    if noisy:
        masks["noisy"] = [(0, n)]
        masks.pop("lf_noisy", None)
    return sig, masks


def rel_gain_calib(signal, gain_fac, gain_def=1.0, gain_min=0.8, gain_max=1.25):
    """Protodune::RelGainCalib::apply (Protodune.cxx:975-1009): gain_fac
    (signal - 0) + 0 in float, unless the factor is outside the cuts."""
    sig = np.asarray(signal, dtype=F).copy()
    g = F(gain_fac)
    if g > F(gain_max) or g < F(gain_min):
        return sig
    return (g * (sig - F(0))).astype(F) + F(0)


# ---- MicroBooNE's coherent noise subtraction (grouped filter) -----------

def filter_time(freq):
    """Microboone.cxx:33-38."""
    return (freq > 0) * math.exp(-0.5 * (freq / 0.143555) ** 4.95096)


def filter_low(freq, cut_off=0.08):
    """Microboone.cxx:40-49."""
    if (0.177 < freq < 0.18) or (0.2143 < freq < 0.215) or (0.106 <= freq <= 0.109) or (0.25 < freq < 0.251):
        return 0.0
    return 1 - math.exp(-((freq / cut_off) ** 8))


def filter_low_loose(freq):
    """Microboone.cxx:51."""
    return 1 - math.exp(-((freq / 0.005) ** 2))


def _freq(i, n):
    return i / (1.0 * n) * 2.0 if i < n / 2.0 else (n - i) / (1.0 * n) * 2.0


def calc_rms(signal):
    """Derivations::CalcRMS (sigproc/src/Derivations.cxx:7-21): mean and RMS
    of the samples within 4.5 RMS of the mean."""
    mean, rms = mean_rms(signal)
    keep = [x for x in np.asarray(signal, dtype=F) if abs(float(x) - mean) < 4.5 * rms]
    return mean_rms(keep)


def calc_median(signals):
    """Derivations::CalcMedian (:23-69): per tick, the binned median of the
    channels' samples within 5 (float) mean RMS and above 0.001 in size."""
    max_rms, count = F(0), F(0)
    for s in signals:
        r = mean_rms(s)[1]
        if r > 0:
            max_rms = F(float(max_rms) + r)
            count = F(count + F(1))
    if count > 0:
        max_rms = F(max_rms / count)
    lim = F(F(5) * max_rms)
    arr = np.stack([np.asarray(s, dtype=F) for s in signals])
    med = np.zeros(arr.shape[1], dtype=F)
    for t in range(arr.shape[1]):
        col = [v for v in arr[:, t] if abs(v) < lim and abs(float(v)) > 0.001]
        med[t] = median_binned(col) if col else F(0)
    return med


def _decon(sig, respec, lowfilt):
    spec = fwd_r2c(sig).astype(np.complex128) / np.asarray(respec, np.complex128)
    spec = spec.astype(np.complex64)
    n = len(spec)
    fac = np.array([F(filter_time(_freq(i, n)) * lowfilt(_freq(i, n))) for i in range(n)], dtype=np.complex64)
    return inv_c2r((spec * fac).astype(np.complex64))


def _uses_response(respec, res_offset):
    return respec is not None and len(respec) > 0 and (np.real(respec[0]) != 1 or np.imag(respec[0]) != 0) \
        and res_offset != 0


def _runs(flags):
    rois, inside = [], False
    for i, f in enumerate(flags):
        if f:
            if inside:
                rois[-1].append(i)
            else:
                rois.append([i])
                inside = True
        else:
            inside = False
    return rois


def _interp_rois(vals, rois, which):
    n = len(vals)
    for roi in rois:
        if not which(roi):
            continue
        b0, bf = max(roi[0] - 1, 0), min(roi[-1] + 1, n - 1)
        m0, mf = float(vals[b0]), float(vals[bf])
        run, rise = float(bf - b0), mf - m0
        for b in roi:
            vals[b] = F(m0 + (b - b0) / run * rise)
    return vals


def signal_protection(medians, respec, res_offset, pad_f, pad_b, upper_decon_limit, decon_lf_cutoff,
                      upper_adc_limit, protection_factor, min_adc_limit):
    """Microboone::SignalProtection (Microboone.cxx:276-462); medians
    changed in place; returns the ROIs (lists of ticks)."""
    n = len(medians)
    flags = [False] * n
    mean, rms = calc_rms(medians)

    def mark(j):
        flags[j] = True
        for k in range(pad_b):
            flags[min(j + k + 1, n - 1)] = True
        for k in range(pad_f):
            flags[max(j - k - 1, 0)] = True

    pf = float(F(protection_factor))
    limit = F(pf * rms) if pf * rms > float(F(upper_adc_limit)) else F(upper_adc_limit)
    if F(min_adc_limit) < limit:
        limit = F(min_adc_limit)
    for j in range(n):
        if abs(float(medians[j]) - mean) > float(limit):
            mark(j)
    use = _uses_response(respec, res_offset)
    if use:
        dec = _decon(medians, respec, lambda f: filter_low(f, float(F(decon_lf_cutoff))))
        mean, rms = calc_rms(dec)
        limit = F(pf * rms) if pf * rms > float(F(upper_decon_limit)) else F(upper_decon_limit)
        for j in range(n):
            if float(dec[j]) - mean > float(limit):
                tb = j + res_offset
                if tb >= n:
                    tb -= n
                mark(tb)
    rois = _runs(flags)
    _interp_rois(medians, rois, lambda roi: not use)
    return rois


def subtract_wscaling(signals, medians, respec, res_offset, rois, decon_limit1, roi_min_max_ratio,
                      rms_threshold=0.0):
    """Microboone::Subtract_WScaling (Microboone.cxx:53-274); signals (in
    channel order) changed in place."""
    n = len(medians)
    coefs = []
    ave, cnt = 0.0, 0.0
    for s in signals:
        r = calc_rms(s)[1]
        s2 = s3 = 0.0
        for j in range(n):
            if abs(float(s[j])) < 4 * r:
                s2 += float(F(s[j] * medians[j]))
                s3 += float(F(medians[j] * medians[j]))
        c = s2 / s3 if s3 > 0 else 0.0
        coefs.append(c)
        if c != 0:
            ave += c
            cnt += 1
    if cnt > 0:
        ave /= cnt
    use = _uses_response(respec, res_offset)
    lim1 = float(F(decon_limit1))
    for i, s in enumerate(signals):
        if ave != 0:
            sc = F(coefs[i] / ave)
            sc = max(sc, F(0))
            sc = min(sc, F(1.5))
        else:
            sc = F(0)
        sub = medians
        if use:
            if rms_threshold:
                roi_sig = np.asarray(s, dtype=F).copy()
            else:
                roi_sig = np.zeros(n, dtype=F)
                for roi in rois:
                    b0, bf = max(roi[0] - 1, 0), min(roi[-1] + 1, n - 1)
                    m0, mf = float(s[b0]), float(s[bf])
                    run, rise = float(bf - b0), mf - m0
                    for b in roi:
                        roi_sig[b] = F(float(s[b]) - (m0 + (b - b0) / run * rise))
            dec = _decon(roi_sig, respec, filter_low_loose)
            if rms_threshold:
                m, r = calc_rms(dec)
                dec = np.asarray([F(float(v) - m) for v in dec], dtype=F)
                rl = float(F(rms_threshold)) * r
                if 0 < rl < 0.02:
                    rl = 0.02
                if rl > lim1:
                    rl = lim1
                lim1 = float(F(rl))
            good = set()
            for roi in rois:
                b0, bf = max(roi[0] - 1, 0), min(roi[-1] + 1, n - 1)
                vals = []
                for t in range(b0, bf + 1):
                    tb = t - res_offset
                    if tb < 0:
                        tb += n
                    if tb >= n:
                        tb -= n
                    vals.append(float(dec[tb]))
                mx, mn = max(vals), min(vals)
                if mx > lim1 and abs(mn) < mx * float(F(roi_min_max_ratio)):
                    good.add(roi[0])
            sub = _interp_rois(np.asarray(medians, dtype=F).copy(), rois, lambda roi: roi[0] in good)
        for t in range(n):
            if abs(float(s[t])) > 0.001:
                s[t] = F(s[t] - F(sub[t] * sc))
    return signals


def coherent_noise_sub(signals, respec=None, res_offset=0, pad_f=0, pad_b=0, decon_limit=0.02,
                       decon_lf_cutoff=0.08, adc_limit=0.0, decon_limit1=0.08, roi_min_max_ratio=0.8,
                       protection_factor=5.0, min_adc_limit=50.0, rms_threshold=0.0):
    """Microboone::CoherentNoiseSub::apply (Microboone.cxx:798-872) on one
    group (signals in channel order, float32 arrays changed in place); the
    DB values are those of the group's first channel."""
    signals = [np.asarray(s, dtype=F).copy() for s in signals]
    med = calc_median(signals)
    rois = signal_protection(med, respec, res_offset, pad_f, pad_b, decon_limit, decon_lf_cutoff, adc_limit,
                             protection_factor, min_adc_limit)
    subtract_wscaling(signals, med, respec, res_offset, rois, decon_limit1, roi_min_max_ratio, rms_threshold)
    return signals


# ---- Microboone::OneChannelStatus (Microboone.cxx:1263-1334) -----------

def lf_noisy_masked(sig, threshold=3.5, window=5):
    """ID_lf_noisy's first part (:1287-1318): the binned median and 16/84 %
    points, rms from them (double), the non-zero count, and a copy of the
    samples in which every sample within [i - window, i + window) of a
    sample i beyond threshold x rms from the median is set to the median
    (marks from the original samples).  Returns (masked, median, rms,
    valid)."""
    s = np.asarray(sig, dtype=F)
    mean = float(percentile_binned(s, 0.5))
    val1 = float(percentile_binned(s, 0.5 - 0.34))
    val2 = float(percentile_binned(s, 0.5 + 0.34))
    rms = math.sqrt(((val1 - mean) ** 2 + (val2 - mean) ** 2) / 2.0)
    valid = int(np.count_nonzero(s))
    out = s.copy()
    n = len(s)
    for i in range(n):
        if abs(float(s[i]) - mean) > threshold * rms:
            for k in range(-window, window):
                if 0 <= k + i < n:
                    out[k + i] = F(mean)
    return out, mean, rms, valid


def lf_noisy(sig, threshold=3.5, window=5, nbins=250, cut=14.0):
    """OneChannelStatus::ID_lf_noisy (:1281-1334): the summed magnitude of
    bins 1..nbins of the masked samples' spectrum (float transform) over
    the count of non-zero samples, against cut.  Returns (flag, ratio);
    no non-zero sample: (False, 0)."""
    masked, _, _, valid = lf_noisy_masked(sig, threshold, window)
    if not valid:
        return False, 0.0
    spec = fwd_r2c(masked)
    content = 0.0
    for i in range(nbins):
        content += float(F(abs(spec[i + 1])))
    return content / valid > cut, content / valid


def one_channel_status(signal, plane, threshold=3.5, window=5, nbins=250, cut=14.0):
    """OneChannelStatus::apply (:1263-1279): off the collection plane, the
    whole channel is masked lf_noisy when ID_lf_noisy holds.  Returns the
    mask list of the channel ([(0, n)] or [])."""
    if plane == 2:
        return []
    flag, _ = lf_noisy(signal, threshold, window, nbins, cut)
    return [(0, len(signal))] if flag else []


# ---- Microboone::ADCBitShift (Microboone.cxx:1106-1216) -----------------

def _i32(v):
    """C int wrap-around of a Python int."""
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def bits_shift_right(value, n, filling, total_bits):
    """Bits::shift_right (util/src/Bits.cxx:4-7): (value >> n) | (filling
    << (total_bits - n)), C int arithmetic."""
    return _i32((value >> n) | _i32(filling << (total_bits - n)))


def bits_lowest(value, n):
    """Bits::lowest_bits (util/src/Bits.cxx:9): value & (2^n - 1)."""
    return value & ((1 << n) - 1)


def _trunc(x):
    """C conversion of a double to int (toward zero)."""
    return int(math.trunc(x))


def adc_bit_shift(signal, nbits=12, exam_nticks=500, threshold_sigma=7.5, threshold_fix=0.8):
    """ADCBitShift::apply (:1106-1216).  The low nbits bits of each of the
    first exam_nticks samples (as int, truncated) are compared with the
    previous sample's (zero before the first); nshift is the number of
    low bits, from bit 0 up, that flip fewer than nbin / 2 - nbin / 2
    sqrt(1 / nbin) threshold_sigma times.  For 0 < nshift < 11 every
    sample becomes shift_right(x, nshift, lowest_bits(x, nshift), 12), and
    a sample far (pow(2, nbits - nshift) threshold_fix, as int) from the
    integer mean and from a neighbour takes, in the fillings' increasing
    order, each shift_right(x, nshift, filling, 12) closer to the
    neighbours' extrapolation (sequential: the previous samples are the
    corrected ones; the first sample's next two are the mean).  Returns
    (signal as float32, nshift; 0 when unchanged).  The reference does not
    terminate on a negative sample in the examined ticks (its bit loop
    shifts a negative int forever); this takes the two's-complement bits."""
    sig = np.asarray(signal, dtype=F)
    nbin = exam_nticks
    counter = [0] * nbits
    prev = [0] * nbits
    for i in range(nbin):
        x = _trunc(float(sig[i]))
        cur = [(x >> j) & 1 for j in range(nbits)]
        for j in range(nbits):
            counter[j] += abs(cur[j] - prev[j])
        prev = cur
    lim = nbin / 2. - nbin / 2. * math.sqrt(1. / nbin) * threshold_sigma
    nshift = 0
    for j in range(nbits):
        if counter[j] < lim:
            nshift += 1
        else:
            break
    if not (nshift != 0 and nshift < 11):
        return sig.copy(), 0
    nl = len(sig)
    x_orig = [_trunc(float(v)) for v in sig]
    x = [0] * nl
    fillings = set()
    total = 0
    for i in range(nl):
        f = bits_lowest(x_orig[i], nshift)
        x[i] = bits_shift_right(x_orig[i], nshift, f, 12)
        fillings.add(f)
        total += x[i]
    mean = _i32(total)
    mean = int(mean / nl) if mean >= 0 else -int(-mean / nl)   # C int division
    exp_diff = _trunc(math.pow(2, nbits - nshift) * threshold_fix)
    prev1 = prev_c = nxt = nxt1 = mean
    for i in range(nl):
        cur = x[i]
        if abs(cur - mean) > exp_diff and (abs(cur - prev_c) > exp_diff or abs(cur - nxt) > exp_diff):
            ev = _trunc(((2 * prev_c - prev1) + (prev_c + nxt) / 2. + (prev_c * 2. / 3. + nxt1 / 3.)) / 3.)
            for f in sorted(fillings):
                y = bits_shift_right(x_orig[i], nshift, f, 12)
                if abs(y - ev) < abs(x[i] - ev):
                    x[i] = y
        prev1 = prev_c
        prev_c = x[i]
        nxt = x[i + 2] if i + 2 < nl else mean
        nxt1 = x[i + 3] if i + 3 < nl else mean
    return np.asarray(x, dtype=F), nshift
