"""Synthetic planes of the benchmark workloads (WCT units: ns, mm, e+).

Field responses are Garfield-style path sets at pitch positions
-D..D pitches (the full symmetric set: Response::wire_region_average,
util/src/Response.cxx:79-196, only sizes its averages when a position
repeats).  Collection: a unipolar Gaussian current, sigma = 3 ticks,
neighbour d scaled by 1/(1+d^2), d <= 5.  Induction: the time
derivative of that Gaussian (positive lobe of unit area), d <= 10.
Electronics: cold electronics, 14 mV/fC, 2 us; ADC_mV = 4096/2000 mV;
postgain 1.  Decon filters: Wiener tight sigma 0.1 MHz with the PDSP
Wiener_tight_W power 3.35324 (cfg/pgrapher/experiment/pdsp/
sp-filters.jsonnet:95-98), induction adds ROI_tight_lf (0.014 MHz) and
the mask's ROI_tighter_lf (0.06 MHz); wire filter sigma 0.75, power 2.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import host

US = 1000.0
MHZ = 1e-3
MV = 1e-9
FC = 1.0e-15 / 1.602176487e-19

TICK = 0.5 * US
FINE_PERIOD = 0.1 * US
FR_SAMPLES = 1000
PITCH = 4.0
FR_PEAK = 30.0 * US
ORIGIN = 48.0                 # mm
SPEED = 1.6 / US              # mm/ns  -> origin/speed = 30 us
CTOFFSET = 0.0
FTOFFSET = 0.0
GAIN = 14.0 * MV / FC
SHAPING = 2.0 * US
ADC_MV = 4096 / (2000.0 * MV)
INTER_GAIN = 1.0
SIGMA_T = 3 * TICK

WIENER = dict(max_freq=1 * MHZ, sigma=0.1 * MHZ, power=3.35324, flag=True)
WIRE = dict(max_freq=1.0, sigma=0.75, power=2, flag=False)
LF_TIGHT = dict(max_freq=1 * MHZ, tau=0.014 * MHZ)
LF_TIGHTER = dict(max_freq=1 * MHZ, tau=0.06 * MHZ)
# the other decon variants' filters, PDSP values
# (cfg/pgrapher/experiment/pdsp/sp-filters.jsonnet:79-114)
LF_LOOSE = dict(max_freq=1 * MHZ, tau=0.002 * MHZ)
WIENER_WIDE = dict(max_freq=1 * MHZ, sigma=0.175722 * MHZ, power=4.37928, flag=True)
GAUS_WIDE = dict(max_freq=1 * MHZ, sigma=0.12 * MHZ, power=2, flag=True)
TH_FACTOR = 3.0
NOISE_ADC = 1.5
Q_PER_TICK = 5000.0


def fr_paths(kind):
    """(pitchpos[], current[npaths, FR_SAMPLES]) for 'col' or 'ind'."""
    dmax = 5 if kind == "col" else 10
    t = np.arange(FR_SAMPLES) * FINE_PERIOD
    g = np.exp(-0.5 * ((t - FR_PEAK) / SIGMA_T) ** 2) / (SIGMA_T * math.sqrt(2 * math.pi))
    if kind == "ind":
        g = SIGMA_T * math.sqrt(2 * math.pi) * (-(t - FR_PEAK) / SIGMA_T ** 2) * g
    pos, cur = [], []
    for d in range(-dmax, dmax + 1):
        pos.append(d * PITCH)
        cur.append(-g / (1.0 + d * d))
    return np.array(pos), np.array(cur)


@dataclass
class PlaneSpec:
    kind: str            # "col" or "ind"
    nw: int
    nt: int
    fr_rows: np.ndarray = field(repr=False, default=None)   # region-averaged rows
    elec: np.ndarray = field(repr=False, default=None)
    nfine: int = 0
    wire_shift: int = 0
    time_shift: int = 0
    wire_filter: np.ndarray = field(repr=False, default=None)
    time_filter: np.ndarray = field(repr=False, default=None)   # Wiener (x tight LF if folded)
    lf_tight: np.ndarray = field(repr=False, default=None)
    lf_tighter: np.ndarray = field(repr=False, default=None)
    lf_folded: bool = False          # time_filter holds Wiener x ROI_tight_lf


def plane_spec(kind, nw, nt, fold_lf=True):
    """All host-side inputs of one plane's decon (OmnibusSigProc.cxx:707-865)."""
    pos, cur = fr_paths(kind)
    regions, rows = host.wire_region_average(pos, cur, PITCH)
    nfine = host.fft_best_length(FR_SAMPLES)
    elec = host.coldelec_samples(nfine, 0.0, nfine * FINE_PERIOD, GAIN, SHAPING, 1.0)
    sp = PlaneSpec(kind=kind, nw=nw, nt=nt, fr_rows=rows, elec=elec, nfine=nfine)
    sp.wire_shift = (len(regions) - 1) // 2
    sp.time_shift = int((CTOFFSET + ORIGIN / SPEED) / TICK)
    sp.wire_filter = host.hf_filter(nw, **WIRE)
    wiener = host.hf_filter(nt, **WIENER)
    if kind == "ind":
        lt = host.lf_filter(nt, **LF_TIGHT)
        sp.lf_tight = lt
        sp.lf_tighter = host.lf_filter(nt, **LF_TIGHTER)
        sp.time_filter = wiener * lt if fold_lf else wiener
        sp.lf_folded = bool(fold_lf)
    else:
        sp.time_filter = wiener
    return sp


# Plane lists of the workloads.
APA_PLANES = (("ind", 800), ("ind", 800), ("col", 960))


def config_planes(name):
    if name == "col480":          # one collection plane, 480 x 6000
        return [("col", 480, 6000)]
    if name == "apa":             # one APA, U/V/W 800/800/960 x 6000
        return [(k, nw, 6000) for k, nw in APA_PLANES]
    if name == "batch24":         # 8 APAs x 3 planes, Nt cycling 6000/8000/10000
        return [(k, nw, (6000, 8000, 10000)[a % 3]) for a in range(8) for k, nw in APA_PLANES]
    if name.startswith("apa") and name[3:].isdigit():   # N APAs of U/V/W 800/800/960 x 6000
        return [(k, nw, 6000) for a in range(int(name[3:])) for k, nw in APA_PLANES]
    if name == "apa150":          # 150 APAs x 2560 channels x 6000 ticks
        return [(k, nw, 6000) for a in range(150) for k, nw in APA_PLANES]
    raise KeyError(name)


def tracks_raster(rng, ntracks, nw, nt, q_per_tick=Q_PER_TICK):
    """Straight tracks with uniform endpoints in [0,nw) x [0,nt), each
    depositing q_per_tick electrons per tick along its length (float32)."""
    q = np.zeros((nw, nt), np.float32)
    for _ in range(ntracks):
        w0, w1 = rng.uniform(0, nw, 2)
        t0, t1 = rng.uniform(0, nt, 2)
        dw, dt = abs(w1 - w0), abs(t1 - t0)
        nstep = int(math.ceil(max(dw, dt))) + 1
        f = np.linspace(0.0, 1.0, nstep) if nstep > 1 else np.zeros(1)
        w = np.clip(np.floor(w0 + f * (w1 - w0)).astype(int), 0, nw - 1)
        t = np.clip(np.floor(t0 + f * (t1 - t0)).astype(int), 0, nt - 1)
        np.add.at(q, (w, t), q_per_tick * (dt + 1.0) / nstep)
    return q


def charge_raster(rng, nw, nt, occupancy=0.02, ntracks=None, blip_mean=20000.0):
    """Sparse tracks plus exponential-energy point blips (electrons per
    pixel) until the given fraction of pixels is nonzero (float32)."""
    q = tracks_raster(rng, tracks_per_plane(nw) if ntracks is None else ntracks, nw, nt)
    want = int(occupancy * nw * nt)
    have = int(np.count_nonzero(q))
    if want > have:
        n = want - have
        w = rng.integers(0, nw, n)
        t = rng.integers(0, nt, n)
        np.add.at(q, (w, t), rng.exponential(blip_mean, n).astype(np.float32))
    return q


def tracks_per_plane(nw, per_480=20):
    """Track count at the density of 20 tracks per 480 wires."""
    return max(1, int(round(per_480 * nw / 480.0)))


# ---- forward simulation convolution (Gen::ImpactTransform) ----
# One induction plane of 800 wires read out over 6000 ticks: 11 impact
# groups (10 impact positions per pitch), responses over 21 wires, the
# frame of a transform whose charge spans the whole plane (896 x 6272).
IMPACT = dict(num_group=11, nwires_resp=21, nresp=6000, impact_range=(0, 8000), time_range=(0, 6000),
              waveform_pad=200, ndepo=20000)


def impact_responses(num_group, nwires_resp, nresp, seed=7):
    """Response rows (num_group, nwires_resp, nresp): a bipolar induction
    shape on the offset wires, unipolar on the central one, shifted with
    the impact position and falling off with the offset; zero after 400
    ticks as the reference's short-padded responses are."""
    rng = np.random.default_rng(seed)
    P = (nwires_resp - 1) // 2
    t = np.arange(min(nresp, 400), dtype=np.float64)
    r = np.zeros((num_group, nwires_resp, nresp))
    for g in range(num_group):
        for j in range(nwires_resp):
            k = j - P
            t0 = 60 + 2 * abs(k) + 0.3 * g
            s = np.exp(-0.5 * ((t - t0) / 5.0) ** 2)
            if k != 0:
                s = s - np.exp(-0.5 * ((t - t0 - 12) / 5.0) ** 2)
            r[g, j, :len(t)] = s / (1 + abs(k)) * (1 + 0.02 * rng.normal())
    return r


def impact_charges(frame, num_group, ndepo, time_range, seed=11):
    """Charge entries of ndepo depositions as get_charge_vec hands them
    over: each deposition spreads over 3 neighbouring impact groups,
    2 wires and 12 ticks with Gaussian weights.  Returns (offsets[G + 1],
    channel, tick, charge) with the entries grouped by impact group."""
    rng = np.random.default_rng(seed)
    ch0 = rng.integers(frame["start_ch"] + 1, frame["end_ch"] - 2, size=ndepo)
    t0 = rng.integers(time_range[0], max(time_range[0] + 1, time_range[1] - 13), size=ndepo)
    g0 = rng.integers(0, num_group - 2, size=ndepo)
    amp = rng.uniform(200.0, 8000.0, size=ndepo)
    dg, dc, dt = np.meshgrid(np.arange(3), np.arange(2), np.arange(12), indexing="ij")
    w = (np.exp(-0.5 * ((dt - 5.5) / 2.0) ** 2) * np.where(dc == 0, 1.0, 0.5) *
         np.exp(-0.5 * (dg - 1.0) ** 2)).ravel()
    grp = (g0[:, None] + dg.ravel()[None, :]).ravel()
    ch = (ch0[:, None] + dc.ravel()[None, :]).ravel()
    tk = (t0[:, None] + dt.ravel()[None, :]).ravel()
    q = (amp[:, None] * w[None, :]).ravel()
    order = np.argsort(grp, kind="stable")
    counts = np.bincount(grp, minlength=num_group)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return offsets, ch[order].astype(np.int32), tk[order].astype(np.int32), q[order]
