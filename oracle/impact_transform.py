"""numpy restatement of Wire-Cell's forward simulation convolution
(gen/src/ImpactTransform.cxx:19-347).

TEST INFRASTRUCTURE ONLY.  Only tests/ and the CPU-baseline leg of the
benchmark may import this module, and only as the checker.  It follows the
reference's steps in complex128: the frame geometry (:28-29, :84-115), the
2-for-1 packing of mirrored impact groups into the real and imaginary parts
of one frame (:122-212), the central group (:216-301), the inverse time
transform and the real + row-reversed imaginary output (:303-306), and
waveform(iwire) with the optional long-range response (:312-347).

Inputs are what BinnedDiffusion_transform::get_charge_vec and the plane's
impact responses hand the reference:
  charges[g]  -- list of (channel, tick, charge) for impact group g
  resp[g][k + P] -- the time waveform of the response of group g at wire
                 offset k in [-P, P] (IImpactResponse::waveform(); the
                 reference recovers it as inv_c2r(spectrum()), :158-168)
"""
import math

import numpy as np

import oracle as O


def geometry(num_group, nwires_resp, impact_range, time_range, waveform_pad):
    """The frame of one transform (:28-29, :90-115)."""
    num_pad_wire = int(math.floor((nwires_resp - 1) / 2.0 + 0.5))   # std::round of a value >= 0
    start_ch = math.floor(impact_range[0] / (num_group - 1)) - 1
    end_ch = math.ceil(impact_range[1] / (num_group - 1)) + 2
    if (end_ch - start_ch) % 2 == 1:
        end_ch += 1
    start_tick = time_range[0] - 1
    end_tick = time_range[1] + 2
    if (end_tick - start_tick) % 2 == 1:
        end_tick += 1
    ntotal_wires = O.fft_best_length(end_ch - start_ch + 2 * num_pad_wire, True)
    npad_wire = (ntotal_wires - end_ch + start_ch) // 2
    ntotal_ticks = O.fft_best_length(end_tick - start_tick + waveform_pad)
    npad_time = ntotal_ticks - end_tick + start_tick
    return dict(num_group=num_group, num_pad_wire=num_pad_wire, start_ch=start_ch, end_ch=end_ch,
                npad_wire=npad_wire, m_start_ch=start_ch - npad_wire, m_end_ch=end_ch + npad_wire,
                m_start_tick=start_tick, m_end_tick=end_tick + npad_time,
                nw=end_ch - start_ch + 2 * npad_wire, nt=end_tick + npad_time - start_tick)


def _response_frame(resp_g, P, nw, nt):
    """resp_f_w of one group (:155-201): each offset's waveform cut or
    zero-padded to nt ticks at row k (k >= 0) or nw + k (k < 0), then the
    time and wire transforms."""
    frame = np.zeros((nw, nt), dtype=np.complex128)
    for k in range(-P, P + 1):
        w = np.asarray(resp_g[k + P], dtype=np.float64)
        row = k if k >= 0 else nw + k
        m = min(nt, len(w))
        frame[row, :m] = w[:m]
    return np.fft.fft(np.fft.fft(frame, axis=1), axis=0)


def _cols(group):
    """(channel, tick, charge) columns of one group's entries (a list of
    tuples or an (n, 3) array)."""
    a = np.asarray(group, dtype=np.float64).reshape(-1, 3)
    return a[:, 0].astype(np.int64), a[:, 1].astype(np.int64), a[:, 2]


def transform(charges, resp, geo):
    """m_decon_data (nw x nt) of the constructor (:117-306)."""
    G, P = geo["num_group"], geo["num_pad_wire"]
    nw, nt = geo["nw"], geo["nt"]
    start_ch, end_ch, npad = geo["start_ch"], geo["end_ch"], geo["npad_wire"]
    t0 = geo["m_start_tick"]
    num_double = (G - 1) // 2
    acc = np.zeros((nw, nt), dtype=np.complex128)
    for i in range(num_double):
        c = np.zeros((nw, nt), dtype=np.complex128)
        ch, tick, q = _cols(charges[i])
        np.add.at(c, (ch + npad - start_ch, tick - t0), q)
        ii = 2 * num_double - i
        ch, tick, q = _cols(charges[ii])
        np.add.at(c, (end_ch + npad - 1 - ch, tick - t0), 1j * q)
        c = np.fft.fft2(c)
        c = c * _response_frame(resp[i], P, nw, nt)
        acc += np.fft.ifft(c, axis=0)
    i = num_double
    d = np.zeros((nw, nt), dtype=np.float64)
    ch, tick, q = _cols(charges[i])
    np.add.at(d, (ch + npad - start_ch, tick - t0), q)
    c = np.fft.fft(np.fft.fft(d, axis=1), axis=0) * _response_frame(resp[i], P, nw, nt)
    acc += np.fft.ifft(c, axis=0)
    acc = np.fft.ifft(acc, axis=1)
    return acc.real + acc.imag[::-1, :]


def waveform(decon, geo, iwire, nsamples, long_aux=None, long_pad=0):
    """ImpactTransform::waveform (:312-347)."""
    if iwire < geo["m_start_ch"] or iwire >= geo["m_end_ch"]:
        return np.zeros(nsamples)
    wf = np.zeros(nsamples)
    t0, t1 = geo["m_start_tick"], geo["m_end_tick"]
    lo, hi = max(t0, 0), min(t1, nsamples)
    if hi > lo:
        wf[lo:hi] = decon[iwire - geo["m_start_ch"], lo - t0:hi - t0]
    if long_aux is not None and len(long_aux) > 0:
        n = O.fft_best_length(nsamples + long_pad)
        a = np.zeros(n)
        a[:nsamples] = wf
        b = np.zeros(n)
        m = min(n, len(long_aux))
        b[:m] = long_aux[:m]
        spec = np.fft.fft(a) * np.fft.fft(b)
        # inv_c2r mirrors the lower half of the spectrum before the inverse
        half = n // 2 + 1
        sym = spec.copy()
        sym[half:] = np.conj(spec[1:n - half + 1][::-1])
        sym[0] = sym[0].real
        if n % 2 == 0:
            sym[n // 2] = sym[n // 2].real
        wf = np.fft.ifft(sym).real[:nsamples]
    return wf
