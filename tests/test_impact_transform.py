"""The forward simulation convolution of Gen::ImpactTransform
(gen/src/ImpactTransform.cxx:19-347): the oracle pinned on the CPU.

The reference ships no test or golden vector for ImpactTransform, so the
oracle (oracle/impact_transform.py, a statement-by-statement numpy
restatement with 2D transforms and the 2-for-1 real/imaginary packing) is
pinned by what the mathematics fixes:
  - the whole transform equals the plain sum over impact groups of direct
    circular convolutions (no transforms, no packing): group g < nd
    convolved with its response rows, the mirror group 2 nd - g with the
    wire-reversed rows of group 2 nd - g's partner, the central group with
    its own; a wrong packing row, a missing row reversal or a swapped
    real/imaginary part fails it;
  - the frame geometry on hand-worked cases (floor/ceil of the impact
    range over the group count, the parity fix-ups, fft_best_length's
    tables, which are pinned against the reference's own tables
    elsewhere), and the C ABI's geometry equal to it;
  - waveform(iwire): the time window on hand-built cases and the
    long-range convolution against numpy's direct convolution.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
import impact_transform as IT


def direct(charges, resp, geo):
    """Sum over groups of circular convolutions over (wire, tick), no
    transforms.  Group g > nd uses the rows of group 2 nd - g mirrored in
    wire (h_g(dw, dt) = h_{2nd-g}(-dw, dt))."""
    G, P = geo["num_group"], geo["num_pad_wire"]
    nw, nt = geo["nw"], geo["nt"]
    nd = (G - 1) // 2
    out = np.zeros((nw, nt))
    for g in range(G):
        if g > 2 * nd:
            continue
        src, sgn = (g, 1) if g <= nd else (2 * nd - g, -1)
        ker = np.zeros((nw, nt))
        for k in range(-P, P + 1):
            w = np.asarray(resp[src][k + P], dtype=np.float64)
            m = min(nt, len(w))
            ker[(sgn * k) % nw, :m] += w[:m]
        for ch, tick, q in charges[g]:
            r = ch - geo["m_start_ch"]
            c = tick - geo["m_start_tick"]
            out += q * np.roll(np.roll(ker, r, axis=0), c, axis=1)
    return out


def make_case(rng, G, nwires_resp, imp, tim, pad, nresp, ncharge, dup=0):
    geo = IT.geometry(G, nwires_resp, imp, tim, pad)
    P = geo["num_pad_wire"]
    resp = rng.normal(size=(G, 2 * P + 1, nresp))
    charges = []
    for g in range(G):
        lst = []
        for _ in range(ncharge):
            ch = int(rng.integers(geo["start_ch"] + 1, geo["end_ch"] - 1))
            tk = int(rng.integers(tim[0], tim[1] + 1))
            lst.append((ch, tk, float(rng.normal())))
        for _ in range(dup):   # repeated cells add up
            lst.append(lst[int(rng.integers(len(lst)))][:2] + (float(rng.normal()),))
        charges.append(lst)
    return geo, resp, charges


@pytest.mark.parametrize("G,nwr,imp,tim,pad,nresp", [
    (11, 21, (37, 95), (100, 300), 50, 120),
    (11, 21, (0, 40), (0, 30), 10, 300),     # frame starts before channel 0 / tick 0
    (5, 5, (3, 17), (5, 60), 8, 40),
    (3, 3, (2, 9), (1, 20), 4, 10),
    (10, 7, (4, 40), (2, 25), 6, 30),       # even group count: group 9 is unused
])
def test_oracle_equals_direct_convolution(G, nwr, imp, tim, pad, nresp):
    rng = np.random.default_rng(G * 100 + nwr)
    geo, resp, charges = make_case(rng, G, nwr, imp, tim, pad, nresp, ncharge=6, dup=2)
    got = IT.transform(charges, resp, geo)
    want = direct(charges, resp, geo)
    assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max())


def test_oracle_groups_are_separable():
    # one charge in one group at a time: the output is that group's
    # response (mirrored for g > nd) shifted to the charge
    rng = np.random.default_rng(3)
    G, nwr = 5, 5
    geo = IT.geometry(G, nwr, (10, 30), (20, 60), 8)
    P, nd = geo["num_pad_wire"], (G - 1) // 2
    resp = rng.normal(size=(G, nwr, 30))
    ch, tk = geo["start_ch"] + 3, 25
    for g in range(G):
        charges = [[] for _ in range(G)]
        charges[g] = [(ch, tk, 1.0)]
        out = IT.transform(charges, resp, geo)
        r0, c0 = ch - geo["m_start_ch"], tk - geo["m_start_tick"]
        src = g if g <= nd else 2 * nd - g
        for k in range(-P, P + 1):
            row = (r0 + (k if g <= nd else -k)) % geo["nw"]
            np.testing.assert_allclose(out[row, c0:c0 + 30], resp[src][k + P], atol=1e-10)
        assert abs(out.sum() - resp[src].sum()) < 1e-9


def test_geometry_hand_worked():
    # impact (37, 95) over 10 impacts per pitch: channels floor(3.7)-1 = 2
    # to ceil(9.5)+2 = 12 (even span 10); 10 + 2*10 = 30 wires -> even
    # table 40, npad 15; ticks 99..302 -> 303 (span 204), 204 + 50 = 254
    # -> 256, npad_time 52
    g = IT.geometry(11, 21, (37, 95), (100, 300), 50)
    assert (g["start_ch"], g["end_ch"], g["npad_wire"]) == (2, 12, 15)
    assert (g["m_start_ch"], g["m_end_ch"], g["nw"]) == (-13, 27, 40)
    assert (g["m_start_tick"], g["m_end_tick"], g["nt"]) == (99, 355, 256)
    # a full 800-wire plane: -1 .. 802 -> 803 (span 804), 824 -> 896,
    # npad 46; 6000 ticks + 200: -1 .. 6003 (span 6004), 6204 -> 6272
    g = IT.geometry(11, 21, (0, 8000), (0, 6000), 200)
    assert (g["start_ch"], g["end_ch"], g["npad_wire"], g["nw"]) == (-1, 803, 46, 896)
    assert (g["m_start_tick"], g["m_end_tick"], g["nt"]) == (-1, 6271, 6272)


ONAME = {"frame_start_ch": "m_start_ch", "frame_end_ch": "m_end_ch", "frame_start_tick": "m_start_tick",
         "frame_end_tick": "m_end_tick"}


def test_geometry_c_abi_matches_oracle():
    from wct_amd import impact as I
    from wct_amd.capi import lib
    lib()
    rng = np.random.default_rng(11)
    for _ in range(200):
        G = int(rng.choice([3, 5, 7, 11, 21]))
        nwr = int(rng.choice([3, 5, 11, 21]))
        a = int(rng.integers(0, 30000)); b = a + int(rng.integers(0, 30000))
        t0 = int(rng.integers(0, 9000)); t1 = t0 + int(rng.integers(0, 9000))
        pad = int(rng.integers(0, 600))
        want = IT.geometry(G, nwr, (a, b), (t0, t1), pad)
        got = I.geometry(G, nwr, (a, b), (t0, t1), pad)
        assert got == {k: want[ONAME.get(k, k)] for k in I.FRAME_KEYS}
    for n in list(range(1, 2000)) + [16383, 16384, 16385, 20000]:
        for keep in (0, 1):
            assert I.fft_best_length(n, keep) == O.fft_best_length(n, bool(keep))


def test_create_rejects_bad_arguments():
    from wct_amd.capi import lib, WctError
    L = lib()
    h = C.c_void_p()
    dummy = C.c_void_p(1)
    assert L.wct_impact_create(C.byref(h), 11, 20, dummy, 100, None) == -1    # even response wire count
    assert L.wct_impact_create(C.byref(h), 0, 21, dummy, 100, None) == -1
    assert L.wct_impact_create(C.byref(h), 65, 21, dummy, 100, None) == -1
    assert L.wct_impact_create(C.byref(h), 11, 21, None, 100, None) == -1
    assert L.wct_impact_destroy(None) == 0


def test_waveform_window_and_long_response():
    rng = np.random.default_rng(5)
    geo = IT.geometry(11, 21, (0, 40), (0, 30), 10)   # frame ticks [-1, ...)
    decon = rng.normal(size=(geo["nw"], geo["nt"]))
    ns = 50
    w = IT.waveform(decon, geo, 3, ns)
    t0 = geo["m_start_tick"]
    hi = min(geo["m_end_tick"], ns)
    np.testing.assert_array_equal(w[:hi], decon[3 - geo["m_start_ch"], -t0:hi - t0])
    assert not w[hi:].any()
    assert not IT.waveform(decon, geo, geo["m_end_ch"], ns).any()
    assert not IT.waveform(decon, geo, geo["m_start_ch"] - 1, ns).any()
    # long-range response: circular convolution over fft_best_length(ns + pad)
    aux = rng.normal(size=37)
    pad = 20
    n = O.fft_best_length(ns + pad)
    got = IT.waveform(decon, geo, 3, ns, aux, pad)
    a = np.zeros(n); a[:ns] = w
    full = np.convolve(a, aux)
    circ = full[:n].copy()
    circ[:len(full) - n] += full[n:]
    np.testing.assert_allclose(got, circ[:ns], atol=1e-10)
