"""Gen::ImpactTransform on the device (wct_impact_*) against the oracle
(oracle/impact_transform.py, FP64).

Tolerance: FP32 transforms and float atomics give a relative RMS error
(||dev - ref||_2 / ||ref||_2) of about 1e-7..1e-6; the tests require
<= 1e-5, and a maximum error <= 1e-4 of max |ref|.  At the full plane
size (896 x 6272) the properties that do not need the oracle are checked
instead: linearity in the charges, an isolated charge reproducing its
group's response rows, and the frame sum.
"""
import numpy as np
import pytest

import impact_transform as IT

pytestmark = pytest.mark.gpu

REL_RMS = 1e-5
REL_MAX = 1e-4


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel_err(got, want):
    d = got - want
    return (np.sqrt((d * d).sum() / max((want * want).sum(), 1e-300)),
            np.abs(d).max() / max(np.abs(want).max(), 1e-300))


def depo_charges(rng, geo, G, ndepo, tim):
    """Charge lists of ndepo Gaussian-ish depositions: each spreads over a
    few impact groups, a few wires and ~10 ticks."""
    lists = [[] for _ in range(G)]
    for _ in range(ndepo):
        ch = int(rng.integers(geo["start_ch"] + 1, geo["end_ch"] - 2))
        t = int(rng.integers(tim[0], max(tim[0] + 1, tim[1] - 12)))
        amp = float(rng.uniform(100, 5000))
        for g in rng.choice(G, size=3, replace=False):
            for dc in (0, 1):
                for dt in range(10):
                    q = amp * np.exp(-0.5 * ((dt - 5) / 2.0) ** 2) * (0.6 if dc else 1.0)
                    lists[int(g)].append((ch + dc, t + dt, q))
    return lists


def responses(rng, G, nwr, nresp):
    """Smooth bipolar/unipolar response rows falling off with |offset|."""
    P = (nwr - 1) // 2
    t = np.arange(nresp)
    r = np.zeros((G, nwr, nresp))
    for g in range(G):
        for j in range(nwr):
            k = j - P
            t0 = 40 + 3 * abs(k) + g
            shape = np.exp(-0.5 * ((t - t0) / 6.0) ** 2)
            if k != 0:
                shape = shape - np.exp(-0.5 * ((t - t0 - 14) / 6.0) ** 2)
            r[g, j] = shape * (1.0 / (1 + abs(k))) * (1 + 0.01 * rng.normal())
    return r


def run_dev(torch, it, geo, charges):
    offs, ch, tk, q = it.pack_charges(charges)
    out = it.run(geo, offs, ch, tk, q)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("G,nwr,imp,tim,pad,nresp,ndepo", [
    (11, 21, (37, 95), (100, 300), 50, 120, 20),
    (11, 21, (0, 40), (0, 30), 10, 300, 10),
    (5, 5, (3, 17), (5, 60), 8, 40, 10),
    (3, 3, (2, 9), (1, 20), 4, 10, 5),
    (10, 7, (4, 40), (2, 25), 6, 30, 5),
    (11, 21, (0, 2400), (0, 1000), 200, 1000, 300),   # 288 x 1280 frame
    (11, 11, (150, 990), (17, 2217), 100, 2300, 200),  # 100 x 2560 frame
])
def test_transform_matches_oracle(G, nwr, imp, tim, pad, nresp, ndepo):
    torch = _torch()
    from wct_amd.impact import ImpactTransform
    rng = np.random.default_rng(nwr * 1000 + nresp)
    geo = IT.geometry(G, nwr, imp, tim, pad)
    resp = responses(rng, G, nwr, nresp)
    charges = depo_charges(rng, geo, G, ndepo, tim)
    print(f"impact case G={G} nwr={nwr} frame {geo['nw']}x{geo['nt']}", flush=True)
    it = ImpactTransform(resp.astype(np.float32))
    fr = it.geometry(imp, tim, pad)
    assert (fr["nw"], fr["nt"]) == (geo["nw"], geo["nt"])
    got = run_dev(torch, it, fr, charges)
    want = IT.transform(charges, resp.astype(np.float32).astype(np.float64), geo)
    rms, mx = rel_err(got, want)
    assert rms <= REL_RMS and mx <= REL_MAX, (rms, mx)
    it.close()


def test_waveforms_match_oracle():
    torch = _torch()
    from wct_amd.impact import ImpactTransform
    rng = np.random.default_rng(7)
    G, nwr, imp, tim, pad = 11, 21, (0, 400), (0, 500), 50
    geo = IT.geometry(G, nwr, imp, tim, pad)
    resp = responses(rng, G, nwr, 600).astype(np.float32)
    charges = depo_charges(rng, geo, G, 40, tim)
    it = ImpactTransform(resp)
    fr = it.geometry(imp, tim, pad)
    offs, ch, tk, q = it.pack_charges(charges)
    decon = it.run(fr, offs, ch, tk, q)
    dec_np = decon.cpu().numpy().astype(np.float64)
    ns = 520
    aux = (np.exp(-np.arange(300) / 80.0) * 1e-3).astype(np.float32)
    for long_aux, long_pad in ((None, 0), (aux, 100)):
        wf = it.waveforms(fr, decon, ns, long_aux, long_pad).cpu().numpy().astype(np.float64)
        want = np.stack([IT.waveform(dec_np, geo, geo["m_start_ch"] + r, ns,
                                     None if long_aux is None else long_aux.astype(np.float64), long_pad)
                         for r in range(geo["nw"])])
        rms, mx = rel_err(wf, want)
        assert rms <= REL_RMS and mx <= REL_MAX, (long_pad, rms, mx)
    it.close()


def test_empty_and_out_of_frame_charges():
    torch = _torch()
    from wct_amd.impact import ImpactTransform
    rng = np.random.default_rng(9)
    G, nwr = 11, 21
    imp, tim, pad = (37, 95), (100, 300), 50
    geo = IT.geometry(G, nwr, imp, tim, pad)
    resp = responses(rng, G, nwr, 100).astype(np.float32)
    it = ImpactTransform(resp)
    fr = it.geometry(imp, tim, pad)
    empty = [[] for _ in range(G)]
    assert not run_dev(torch, it, fr, empty).any()
    inside = [[(5, 150, 1000.0)] if g == 2 else [] for g in range(G)]
    outside = [list(x) for x in inside]
    outside[3] = [(geo["m_end_ch"] + 3, 150, 1e6), (5, geo["m_end_tick"] + 1, 1e6), (5, geo["m_start_tick"] - 1, 1e6)]
    np.testing.assert_array_equal(run_dev(torch, it, fr, outside), run_dev(torch, it, fr, inside))
    it.close()


def test_full_plane_properties():
    """896 wires x 6272 ticks (an 800-wire plane over 6000 ticks): linearity,
    an isolated charge, the frame sum; a smaller frame on the same handle
    afterwards equals a fresh handle's result."""
    torch = _torch()
    from wct_amd.impact import ImpactTransform
    rng = np.random.default_rng(21)
    G, nwr, nresp = 11, 21, 6000
    imp, tim, pad = (0, 8000), (0, 6000), 200
    geo = IT.geometry(G, nwr, imp, tim, pad)
    resp = responses(rng, G, nwr, nresp).astype(np.float32)
    it = ImpactTransform(resp)
    fr = it.geometry(imp, tim, pad)
    assert (fr["nw"], fr["nt"]) == (896, 6272)
    a = depo_charges(rng, geo, G, 2000, tim)
    b = depo_charges(rng, geo, G, 2000, tim)
    ab = [x + y for x, y in zip(a, b)]
    oa, ob, oab = (run_dev(torch, it, fr, c) for c in (a, b, ab))
    rms, mx = rel_err(oab, oa + ob)
    assert rms <= REL_RMS and mx <= REL_MAX, (rms, mx)
    # frame sum: each charge contributes q times the sum of its group's
    # rows (the mirror uses its partner's rows)
    nd = (G - 1) // 2
    rs = resp.astype(np.float64).sum(axis=(1, 2))
    want_sum = sum(q * rs[g if g <= nd else 2 * nd - g] for g in range(G) for _, _, q in ab[g])
    assert abs(oab.sum() - want_sum) <= 1e-4 * abs(want_sum)
    # an isolated charge in mirror group 8 at channel 400, tick 3000
    one = [[] for _ in range(G)]
    one[8] = [(400, 3000, 1.0)]
    o1 = run_dev(torch, it, fr, one)
    r0, c0 = 400 - fr["frame_start_ch"], 3000 - fr["frame_start_tick"]
    P = (nwr - 1) // 2
    for k in range(-P, P + 1):
        np.testing.assert_allclose(o1[r0 - k, c0:c0 + 3000], resp[2 * nd - 8][k + P][:3000], atol=5e-6)
    # smaller frame on the reused handle
    imp2, tim2 = (100, 900), (50, 700)
    fr2 = it.geometry(imp2, tim2, pad)
    c2 = depo_charges(rng, IT.geometry(G, nwr, imp2, tim2, pad), G, 100, tim2)
    reused = run_dev(torch, it, fr2, c2)
    it.close()
    it2 = ImpactTransform(resp)
    fresh = run_dev(torch, it2, fr2, c2)
    rms, mx = rel_err(reused, fresh)
    assert rms <= REL_RMS and mx <= REL_MAX
    it2.close()
