// HipOmnibusSigProc: the IFrameFilter of OmnibusSigProc
// (sigproc/src/OmnibusSigProc.cxx) with its per-plane deconvolution and
// tight-ROI stage on the gfx950 kernels of libwct_hip.
//
// Per frame (operator(), :1382-1702): EOS and empty frames are forwarded
// as in the reference; channel masks are re-indexed per plane (:1402-1416);
// every processed plane is loaded (load_data, :304-352), deconvolved
// (decon_2D_init, :944-1051) and filtered into the outputs the
// configuration asks for, all planes of a frame in one batched launch per
// kernel stage:
//  - the tight decon (decon_2D_tightROI, :1077-1117, and for the mask the
//    tighter one, :1120-1160), baseline-restored (:867-906) and
//    thresholded (find_ROI_by_decon_itself with cal_RMS,
//    sigproc/src/ROI_formation.cxx:364-531), for the collection plane and,
//    with use_roi_refinement, the induction planes (:1452-1467);
//  - with use_roi_refinement, the ROI refinement (:1467-1620): for the
//    induction planes the loose decon (decon_2D_looseROI, :1162-1238,
//    with its tight-LF rows next to bad or lf_noisy channels) and its
//    loose ROIs (ROI_formation::find_ROI_loose, extend_ROI_loose), the
//    refine decon (decon_2D_ROI_refine, :1053-1075); per plane the
//    ROI_refinement steps on the host (RoiRefinement.h, planes in
//    parallel threads); then the hits and charge decons (decon_2D_hits,
//    decon_2D_charge, :1316-1380) cut to the refined ROIs by apply_roi and
//    saved as the "wiener" traces (with the per-wire RMS as summary) and
//    the "gauss" traces;
//  - with use_roi_debug_mode, the reference's debug outputs: tight_lf
//    (:1471-1474; an induction plane without refinement saves the
//    unfiltered decon_2D_init frame), loose_lf (decon_2D_looseROI_
//    debug_mode, :1478-1497) and decon_charge (decon_2D_charge, :1349-1380,
//    :1508-1511,1590-1601).
// The output follows save_data (:371-472).  The multi-plane protection
// (MP2ROI/MP3ROI) needs the anode's wire geometry and is refused.  The
// tight ROIs (extended and connected) are added to the output channel-mask
// map under "tight_roi_mask" as half-open tick ranges.
//
// The response spectra are built once per (nticks, tick) on the device
// in FP64 and reused for later frames of the same shape (the reference
// rebuilds them per frame, :1429; the inputs do not change).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "../../include/wct_hip.h"
#include "Components.h"
#include "Factory.h"
#include "Frame.h"
#include "Interface.h"
#include "Response.h"
#include "Buffers.h"
#include "RoiFormation.h"
#include "RoiRefinement.h"

namespace WireCell {
namespace Hip {

struct SigProcError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

static void hcheck(int rc, const char* what)
{
    if (rc != WCT_OK) throw SigProcError(std::string("HipOmnibusSigProc ") + what + ": " + wct_last_error());
}

class HipOmnibusSigProc : public IFrameFilter, public IConfigurable {
  public:
    ~HipOmnibusSigProc() override { release(); }

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["anode"] = m_anode_tn;
        cfg["dft"] = m_dft_tn;
        cfg["field_response"] = m_field_response;
        cfg["elecresponse"] = m_elecresponse_tn;
        cfg["per_chan_resp"] = m_per_chan_resp;
        cfg["ftoffset"] = m_fine_time_offset;
        cfg["ctoffset"] = m_coarse_time_offset;
        cfg["fft_flag"] = 0;
        cfg["postgain"] = m_inter_gain;
        cfg["ADC_mV"] = m_ADC_mV;
        cfg["troi_ind_th_factor"] = (double)m_th_factor_ind;
        cfg["troi_col_th_factor"] = (double)m_th_factor_col;
        cfg["troi_pad"] = m_pad;
        cfg["troi_asy"] = (double)m_asy;
        Configuration pp = Configuration::makeArray();
        for (int p : m_process_planes) pp.append(p);
        cfg["process_planes"] = pp;
        cfg["wiener_tag"] = m_wiener_tag;
        cfg["gauss_tag"] = m_gauss_tag;
        cfg["frame_tag"] = m_frame_tag;
        cfg["tight_lf_tag"] = m_tight_lf_tag;
        cfg["loose_lf_tag"] = m_loose_lf_tag;
        cfg["decon_charge_tag"] = m_decon_charge_tag;
        cfg["use_roi_debug_mode"] = m_use_roi_debug_mode;
        cfg["use_multi_plane_protection"] = false;
        cfg["use_roi_refinement"] = m_use_roi_refinement;
        cfg["sparse"] = m_sparse;
        cfg["lroi_rebin"] = m_loose.rebin;
        cfg["lroi_th_factor"] = m_loose.factor;
        cfg["lroi_max_th"] = m_loose.max_th;
        cfg["lroi_th_factor1"] = m_loose.factor1;
        cfg["lroi_short_length"] = m_loose.short_length;
        cfg["lroi_jump_one_bin"] = m_loose.jump_one_bin;
        cfg["r_th_factor"] = (double)m_refine.th_factor;
        cfg["r_fake_signal_low_th"] = (double)m_refine.fake_signal_low_th;
        cfg["r_fake_signal_high_th"] = (double)m_refine.fake_signal_high_th;
        cfg["r_fake_signal_low_th_ind_factor"] = (double)m_refine.fake_signal_low_th_ind_factor;
        cfg["r_fake_signal_high_th_ind_factor"] = (double)m_refine.fake_signal_high_th_ind_factor;
        cfg["r_pad"] = m_refine.pad;
        cfg["r_break_roi_loop"] = m_refine.break_roi_loop;
        cfg["r_th_peak"] = (double)m_refine.th_peak;
        cfg["r_sep_peak"] = (double)m_refine.sep_peak;
        cfg["r_low_peak_sep_threshold_pre"] = (double)m_refine.low_peak_sep_threshold_pre;
        cfg["r_max_npeaks"] = m_refine.max_npeaks;
        cfg["r_sigma"] = (double)m_refine.sigma;
        // the reference's defaults carry these two keys misspelt
        // (OmnibusSigProc.cxx:265,297); configure reads r_th_percent and
        // isWrapped, as the reference's does (:114,160)
        cfg["r_th_precent"] = (double)m_refine.th_percent;
        cfg["isWarped"] = m_refine.wrapped;
        // Keys of this component only.
        cfg["tight_roi_mask"] = m_tight_roi_mask;
        cfg["spectral_cut"] = m_spectral_cut;
        cfg["response_floor"] = m_response_floor;
        cfg["device"] = m_device;
        return cfg;
    }

    void configure(const Configuration& config) override
    {
        m_anode_tn = get(config, "anode", m_anode_tn);
        m_dft_tn = get(config, "dft", m_dft_tn);
        m_field_response = get(config, "field_response", m_field_response);
        m_elecresponse_tn = get(config, "elecresponse", m_elecresponse_tn);
        m_per_chan_resp = get(config, "per_chan_resp", m_per_chan_resp);
        if (get(config, "fft_flag", 0) != 0)
            throw ConfigError("HipOmnibusSigProc: fft_flag != 0 (padded transforms) is not supported");
        m_fine_time_offset = get(config, "ftoffset", m_fine_time_offset);
        m_coarse_time_offset = get(config, "ctoffset", m_coarse_time_offset);
        m_inter_gain = get(config, "postgain", m_inter_gain);
        m_ADC_mV = get(config, "ADC_mV", m_ADC_mV);
        m_th_factor_ind = get(config, "troi_ind_th_factor", m_th_factor_ind);
        m_th_factor_col = get(config, "troi_col_th_factor", m_th_factor_col);
        m_pad = get(config, "troi_pad", m_pad);
        m_asy = get(config, "troi_asy", m_asy);
        if (config["process_planes"].isArray()) {
            m_process_planes.clear();
            for (const auto& p : config["process_planes"]) {
                const int ip = p.asInt();
                if (ip < 0 || ip > 2) throw ConfigError("HipOmnibusSigProc: process_planes entries must be 0, 1 or 2");
                m_process_planes.push_back(ip);
            }
        }
        m_wiener_tag = get(config, "wiener_tag", m_wiener_tag);
        m_gauss_tag = get(config, "gauss_tag", m_gauss_tag);
        m_frame_tag = get(config, "frame_tag", m_frame_tag);
        m_tight_lf_tag = get(config, "tight_lf_tag", m_tight_lf_tag);
        m_loose_lf_tag = get(config, "loose_lf_tag", m_loose_lf_tag);
        m_decon_charge_tag = get(config, "decon_charge_tag", m_decon_charge_tag);
        m_use_roi_debug_mode = get(config, "use_roi_debug_mode", m_use_roi_debug_mode);
        m_use_roi_refinement = get(config, "use_roi_refinement", m_use_roi_refinement);
        if (get(config, "use_multi_plane_protection", false))
            throw ConfigError("HipOmnibusSigProc: use_multi_plane_protection (MP2ROI/MP3ROI) needs the anode's "
                              "wire geometry, which this component does not read");
        m_loose.rebin = get(config, "lroi_rebin", m_loose.rebin);
        m_loose.factor = get(config, "lroi_th_factor", m_loose.factor);
        m_loose.max_th = get(config, "lroi_max_th", m_loose.max_th);
        m_loose.factor1 = get(config, "lroi_th_factor1", m_loose.factor1);
        m_loose.short_length = get(config, "lroi_short_length", m_loose.short_length);
        m_loose.jump_one_bin = get(config, "lroi_jump_one_bin", m_loose.jump_one_bin);
        if (m_loose.rebin < 1) throw ConfigError("HipOmnibusSigProc: lroi_rebin must be at least 1");
        auto getf = [&](const char* key, float& v) { v = (float)get(config, key, (double)v); };
        getf("r_th_factor", m_refine.th_factor);
        getf("r_fake_signal_low_th", m_refine.fake_signal_low_th);
        getf("r_fake_signal_high_th", m_refine.fake_signal_high_th);
        getf("r_fake_signal_low_th_ind_factor", m_refine.fake_signal_low_th_ind_factor);
        getf("r_fake_signal_high_th_ind_factor", m_refine.fake_signal_high_th_ind_factor);
        m_refine.pad = get(config, "r_pad", m_refine.pad);
        m_refine.break_roi_loop = get(config, "r_break_roi_loop", m_refine.break_roi_loop);
        getf("r_th_peak", m_refine.th_peak);
        getf("r_sep_peak", m_refine.sep_peak);
        getf("r_low_peak_sep_threshold_pre", m_refine.low_peak_sep_threshold_pre);
        m_refine.max_npeaks = get(config, "r_max_npeaks", m_refine.max_npeaks);
        getf("r_sigma", m_refine.sigma);
        getf("r_th_percent", m_refine.th_percent);
        m_refine.wrapped = get(config, "isWrapped", m_refine.wrapped);
        m_response_floor = get(config, "response_floor", m_response_floor);
        if (m_response_floor < 0 || m_response_floor >= 1)
            throw ConfigError("HipOmnibusSigProc: response_floor must be in [0, 1)");
        m_sparse = get(config, "sparse", m_sparse);
        m_tight_roi_mask = get(config, "tight_roi_mask", m_tight_roi_mask);
        m_spectral_cut = get(config, "spectral_cut", m_spectral_cut);
        m_device = get(config, "device", m_device);

        // The dft component must exist, as in the reference (:62-63).
        Factory::find_tn<IDFT>(m_dft_tn);

        // OSP channel map (:171-216): WCT channel -> (plane, wire).
        auto anode = Factory::find_tn<IAnodeChannels>(m_anode_tn);
        m_chmap.clear();
        for (int ip = 0; ip < 3; ++ip) {
            m_channels[ip] = anode->plane_channels(ip);
            for (size_t w = 0; w < m_channels[ip].size(); ++w) m_chmap[m_channels[ip][w]] = {ip, (int)w};
        }
        if (m_chmap.empty())
            throw ConfigError("HipOmnibusSigProc: anode " + m_anode_tn + " has no channels in any plane");
        release();
    }

    bool operator()(const input_pointer& in, output_pointer& out) override
    {
        out = nullptr;
        if (!in) return true;   // EOS
        auto traces = in->traces();
        if (!traces || traces->empty()) {
            out = std::make_shared<Aux::SimpleFrame>(in->ident(), in->time(),
                                                     std::make_shared<ITrace::vector>(), in->tick());
            return true;
        }
        hcheck(wct_set_device(m_device), "set device");

        // Frame extent as init_overall_response takes it (:709-722).
        int tbinmin = traces->front()->tbin(), tbinmax = tbinmin;
        for (const auto& t : *traces) {
            tbinmin = std::min(tbinmin, t->tbin());
            tbinmax = std::max(tbinmax, t->tbin() + (int)t->charge().size());
        }
        const int nticks = tbinmax - tbinmin;
        if (nticks <= 0) throw SigProcError("HipOmnibusSigProc: frame without samples");
        if (!wct_dft_native_length(nticks) || nticks % 2)
            throw SigProcError("HipOmnibusSigProc: nticks " + std::to_string(nticks) +
                               " is odd or has a prime factor above 13 (the fused decon's transforms)");
        prepare(nticks, in->tick());
        if (m_adc_floats == 0) {
            // without ROI refinement or the debug mode the reference saves
            // no traces (:1452-1514): an empty frame with the input masks
            auto sframe = std::make_shared<Aux::SimpleFrame>(in->ident(), in->time(),
                                                             std::make_shared<ITrace::vector>(), in->tick(),
                                                             in->masks());
            sframe->tag_frame(m_frame_tag);
            out = sframe;
            return true;
        }

        // Bad ranges per plane, indexed by wire (:1402-1416).
        std::vector<std::vector<Waveform::BinRange>> bad[3];
        for (int ip = 0; ip < 3; ++ip) bad[ip].assign(m_channels[ip].size(), {});
        const auto inmasks = in->masks();
        auto badit = inmasks.find("bad");
        if (badit != inmasks.end()) {
            for (const auto& cm : badit->second) {
                auto it = m_chmap.find(cm.first);
                if (it == m_chmap.end()) continue;   // another anode's channel
                auto& dst = bad[it->second.first][it->second.second];
                dst.insert(dst.end(), cm.second.begin(), cm.second.end());
            }
        }

        // load_data: dense (wire, tick) frames in pinned memory.
        float* hadc = (float*)m_h_adc.p;
        for (int ip : m_process_planes) {
            Plane& pl = m_planes[ip];
            std::memset(hadc + pl.off, 0, sizeof(float) * (size_t)pl.nw * nticks);
        }
        for (const auto& t : *traces) {
            auto it = m_chmap.find(t->channel());
            if (it == m_chmap.end()) continue;
            Plane& pl = m_planes[it->second.first];
            if (!pl.active) continue;
            const auto& q = t->charge();
            const int tb = t->tbin();
            const int n = std::min((int)q.size(), nticks);
            float* row = hadc + pl.off + (size_t)it->second.second * nticks;
            for (int i = 0; i < n; ++i) {
                const int k = tb + i;
                if (k >= 0 && k < nticks) row[k] = q[i];
            }
        }

        // Bad ranges as CSR for the device: [lo, hi) zeroed on input,
        // [lo, hi] left out of the RMS and the mask.
        size_t nbad_tot = 0;
        for (int ip : m_process_planes)
            for (auto& v : bad[ip]) nbad_tot += v.size();
        const size_t csr_ints = [&] {
            size_t n = 0;
            for (int ip : m_process_planes) n += m_planes[ip].nw + 1;
            return n + 2 * nbad_tot;
        }();
        std::vector<int> csr(csr_ints);
        size_t cpos = 0;
        std::vector<size_t> ptr_off(3), lo_off(3), hi_off(3);
        for (int ip : m_process_planes) {
            const int nw = m_planes[ip].nw;
            size_t nb = 0;
            for (auto& v : bad[ip]) nb += v.size();
            ptr_off[ip] = cpos;
            lo_off[ip] = cpos + nw + 1;
            hi_off[ip] = lo_off[ip] + nb;
            int k = 0;
            for (int w = 0; w < nw; ++w) {
                csr[ptr_off[ip] + w] = k;
                for (auto& r : bad[ip][w]) {
                    csr[lo_off[ip] + k] = std::max(0, r.first);
                    csr[hi_off[ip] + k] = std::min(nticks, r.second);
                    ++k;
                }
            }
            csr[ptr_off[ip] + nw] = k;
            cpos = hi_off[ip] + nb;
        }
        const bool plan_stale = csr.size() * sizeof(int) > m_d_bad.bytes || !m_plan;
        m_d_bad.ensure(std::max<size_t>(csr.size(), 1) * sizeof(int));

        // decon_2D_looseROI's rows that take the tight LF (:1215-1231): a
        // bad channel within 2 wires (1 on the V plane) or an lf_noisy one
        // within 2 wires, in the same plane
        if (m_any_extra[kLoose]) {
            auto lfn = inmasks.find("lf_noisy");
            for (int ip = 0; ip < 2; ++ip) {
                Plane& pl = m_planes[ip];
                if (!pl.active || !pl.has_extra[kLoose]) continue;
                std::vector<char> isbad(pl.nw, 0), islfn(pl.nw, 0);
                for (int w = 0; w < pl.nw; ++w) isbad[w] = !bad[ip][w].empty() || (badit != inmasks.end() &&
                                                          badit->second.count(m_channels[ip][w]));
                if (lfn != inmasks.end())
                    for (int w = 0; w < pl.nw; ++w) islfn[w] = lfn->second.count(m_channels[ip][w]) ? 1 : 0;
                const int nb = ip ? 1 : 2, nl = 2;
                unsigned char* alt = (unsigned char*)m_h_alt.p + pl.alt_off;
                for (int w = 0; w < pl.nw; ++w) {
                    bool a = false;
                    for (int q = std::max(0, w - nb); q <= std::min(pl.nw - 1, w + nb) && !a; ++q) a = isbad[q];
                    for (int q = std::max(0, w - nl); q <= std::min(pl.nw - 1, w + nl) && !a; ++q) a = islfn[q];
                    alt[w] = a ? 1 : 0;
                }
            }
            hcheck(wct_memcpy_htod_async(m_d_alt.p, m_h_alt.p, m_alt_bytes, m_stream), "upload");
        }
        hcheck(wct_memcpy_htod_async(m_d_adc.p, m_h_adc.p, m_adc_floats * sizeof(float), m_stream), "upload");
        if (!csr.empty())
            hcheck(wct_memcpy_htod_async(m_d_bad.p, csr.data(), csr.size() * sizeof(int), m_stream), "upload");
        if (plan_stale || m_bad_layout != std::vector<size_t>{ptr_off[0], ptr_off[1], ptr_off[2], lo_off[0],
                                                              lo_off[1], lo_off[2], hi_off[0], hi_off[1],
                                                              hi_off[2]}) {
            m_bad_layout = {ptr_off[0], ptr_off[1], ptr_off[2], lo_off[0], lo_off[1], lo_off[2],
                            hi_off[0], hi_off[1], hi_off[2]};
            hcheck(wct_stream_sync(m_stream), "sync");
            make_plan(ptr_off, lo_off, hi_off);
        }
        hcheck(wct_decon_plan_run(m_plan, m_stream), "decon");
        hcheck(wct_memcpy_dtoh_async(m_h_out.p, m_d_out.p, m_adc_floats * sizeof(float), m_stream), "download");
        hcheck(wct_memcpy_dtoh_async(m_h_mask.p, m_d_mask.p, m_adc_floats, m_stream), "download");
        hcheck(wct_memcpy_dtoh_async(m_h_rms.p, m_d_rms.p, m_rms_floats * sizeof(float), m_stream), "download");
        for (int k = 0; k < kNExtra; ++k)
            if (m_any_extra[k])
                hcheck(wct_memcpy_dtoh_async(m_h_extra[k].p, m_d_extra[k].p, m_adc_floats * sizeof(float),
                                             m_stream),
                       "download");
        hcheck(wct_stream_sync(m_stream), "sync");

        // save_data (:371-472) per processed plane, in plane order.
        auto itraces = std::make_shared<ITrace::vector>();
        const float* hrms = (const float*)m_h_rms.p;
        const uint8_t* hmask = (const uint8_t*)m_h_mask.p;
        // One plane of one decon output as traces (dense, or sparse
        // positive runs); the device already applied save_data's zeroing
        // in the nominal mode, the debug mode keeps negatives and zeroes
        // only the bad ranges.
        // host_zero: the frame is not zeroed on the device (the ROI-cut
        // wiener and gauss decons); save_data's nominal zeroing of
        // negatives and bad ranges is done here.
        auto save = [&](const float* frame, int ip, IFrame::trace_list_t& idx, IFrame::trace_summary_t* thr,
                        bool host_zero = false) {
            const Plane& pl = m_planes[ip];
            for (int w = 0; w < pl.nw; ++w) {
                const float* row = frame + pl.off + (size_t)w * nticks;
                ITrace::ChargeSequence charge(row, row + nticks);
                if (host_zero && !m_use_roi_debug_mode)
                    for (auto& q : charge) q = q > 0.0f ? q : 0.0f;
                if (m_use_roi_debug_mode || host_zero) {
                    for (auto& r : bad[ip][w])
                        for (int k = std::max(0, r.first); k < std::min(nticks, r.second); ++k) charge[k] = 0;
                }
                const double thresh = pl.tight ? hrms[pl.rms_off + w] : 0.0;
                const int ident = m_channels[ip][w];
                if (m_sparse) {
                    int i = 0;
                    while (i < nticks) {
                        while (i < nticks && !(charge[i] > 0)) ++i;
                        if (i >= nticks) break;
                        int e = i;
                        while (e < nticks && charge[e] != 0) ++e;
                        idx.push_back(itraces->size());
                        itraces->push_back(std::make_shared<Aux::SimpleTrace>(
                            ident, i, ITrace::ChargeSequence(charge.begin() + i, charge.begin() + e)));
                        if (thr) thr->push_back(thresh);
                        i = e;
                    }
                } else {
                    idx.push_back(itraces->size());
                    itraces->push_back(std::make_shared<Aux::SimpleTrace>(ident, 0, std::move(charge)));
                    if (thr) thr->push_back(thresh);
                }
            }
        };
        const bool debug = m_use_roi_debug_mode, refine = m_use_roi_refinement;
        IFrame::trace_list_t tight_traces, loose_traces, charge_traces, wiener_traces, gauss_traces;
        IFrame::trace_summary_t thresholds;
        Waveform::ChannelMaskMap outmasks = inmasks;
        const float* hout = (const float*)m_h_out.p;

        // ROI refinement of every plane, planes in parallel
        std::vector<RoiList> refined[3];
        if (refine) {
            std::vector<std::future<void>> work;
            for (int ip = 0; ip < 3; ++ip)
                if (m_planes[ip].active)
                    work.push_back(std::async(std::launch::async, [&, ip] {
                        refined[ip] = refine_plane(ip, nticks, bad[ip], hmask, hrms);
                    }));
            for (auto& f : work) f.get();
        }

        for (int ip = 0; ip < 3; ++ip) {
            const Plane& pl = m_planes[ip];
            if (!pl.active) continue;
            if (debug && !m_tight_lf_tag.empty()) save(hout, ip, tight_traces, nullptr);
            if (debug && !m_loose_lf_tag.empty()) {
                // the collection plane has no loose decon: the reference
                // saves its tight one under loose_lf (:1503-1507)
                const float* src = pl.has_extra[kLooseDbg] ? (const float*)m_h_extra[kLooseDbg].p : hout;
                save(src, ip, loose_traces, nullptr);
            }
            if (pl.tight && !m_tight_roi_mask.empty()) {
                auto rois = tight_rois(hmask + pl.off, pl.nw, nticks, nticks, m_pad, m_asy);
                auto& cm = outmasks[m_tight_roi_mask];
                for (int w = 0; w < pl.nw; ++w) {
                    if (rois[w].empty()) continue;
                    auto& dst = cm[m_channels[ip][w]];
                    for (auto& r : rois[w]) dst.emplace_back(r.first, r.second + 1);
                }
            }
        }
        // per plane: wiener (hits cut to the ROIs), the debug mode's
        // decon_charge (before the cut), gauss (charge cut to the ROIs)
        std::vector<float> cut;
        auto cut_plane = [&](int ip, int extra) {
            const Plane& pl = m_planes[ip];
            const float* src = (const float*)m_h_extra[extra].p + pl.off;
            cut.assign(src, src + (size_t)pl.nw * nticks);
            apply_roi(cut.data(), pl.nw, nticks, nticks, refined[ip], ip < 2);
            return cut.data() - pl.off;
        };
        for (int ip = 0; ip < 3; ++ip) {
            const Plane& pl = m_planes[ip];
            if (!pl.active) continue;
            if (refine && !m_wiener_tag.empty()) save(cut_plane(ip, kHits), ip, wiener_traces, &thresholds, true);
            if (debug && pl.has_extra[kCharge] && !m_decon_charge_tag.empty())
                save((const float*)m_h_extra[kCharge].p, ip, charge_traces, nullptr);
            if (refine && !m_gauss_tag.empty()) save(cut_plane(ip, kCharge), ip, gauss_traces, nullptr, true);
        }

        auto sframe = std::make_shared<Aux::SimpleFrame>(in->ident(), in->time(), itraces, in->tick(), outmasks);
        sframe->tag_frame(m_frame_tag);
        if (refine) {
            if (!m_wiener_tag.empty()) sframe->tag_traces(m_wiener_tag, wiener_traces, thresholds);
            if (!m_gauss_tag.empty()) sframe->tag_traces(m_gauss_tag, gauss_traces);
        }
        if (debug) {
            if (!m_tight_lf_tag.empty()) sframe->tag_traces(m_tight_lf_tag, tight_traces);
            if (!m_loose_lf_tag.empty()) sframe->tag_traces(m_loose_lf_tag, loose_traces);
            if (!m_decon_charge_tag.empty()) sframe->tag_traces(m_decon_charge_tag, charge_traces);
        }
        out = sframe;
        return true;
    }

  private:
    // Extra outputs of a plane's job (wct_decon_output), in job order.
    enum Extra { kLooseDbg = 0, kCharge = 1, kLoose = 2, kRefine = 3, kHits = 4, kNExtra = 5 };

    // The refinement of one plane: its tight ROIs from the device mask,
    // on induction planes the loose ROIs of the loose decon (RMS without
    // the bad ticks, extended by the tight ROIs) and the refine decon as
    // the ROIs' data, on the collection plane the tight decon; then the
    // ROI_refinement steps.  Returns the refined (extended) ROIs per wire.
    std::vector<RoiList> refine_plane(int ip, int nticks, const std::vector<std::vector<Waveform::BinRange>>& badw,
                                      const uint8_t* hmask, const float* hrms) const
    {
        const Plane& pl = m_planes[ip];
        const bool ind = ip < 2;
        const auto tight = tight_rois(hmask + pl.off, pl.nw, nticks, nticks, m_pad, m_asy);
        std::vector<RoiList> badl(pl.nw);
        std::vector<int> bptr(pl.nw + 1, 0), blo, bhi;
        for (int w = 0; w < pl.nw; ++w) {
            for (const auto& r : badw[w]) {
                badl[w].emplace_back(r.first, r.second);
                blo.push_back(r.first);
                bhi.push_back(r.second);
            }
            bptr[w + 1] = (int)blo.size();
        }
        const float* base = (const float*)m_h_out.p + pl.off;
        std::vector<RoiList> loose;
        if (ind) {
            loose = loose_rois((const float*)m_h_extra[kLoose].p + pl.off, pl.nw, nticks, nticks, bptr.data(),
                               blo.data(), bhi.data(), nullptr, &tight, m_loose);
            base = (const float*)m_h_extra[kRefine].p + pl.off;
        }
        RoiRefinement R(ip, pl.nw, nticks, m_refine);
        R.load(base, nticks, badl, std::vector<float>(hrms + pl.rms_off, hrms + pl.rms_off + pl.nw), tight, loose);
        R.refine();
        return R.rois(true);
    }

    struct Plane {
        bool active{false};
        int nw{0};
        size_t off{0};       // first sample of this plane in the frame buffers
        size_t rms_off{0};
        wct_decon_kernel kernel{nullptr};
        bool tight{true};    // the job's own output is the tight decon (else the unfiltered frame)
        DevBuf f_base, f_tighter, f_extra[kNExtra];
        DevBuf f_loose_alt;  // the loose decon's filter on its tight-LF rows
        size_t alt_off{0};   // this plane's rows in the loose fallback-row flags
        DevBuf cf_rows;      // per-channel correction: table row of each wire
        bool has_tighter{false};
        bool has_extra[kNExtra]{};
        int extra_restore[kNExtra]{};
        float th_factor{3};
    };

    void release()
    {
        if (m_plan) wct_decon_plan_destroy(m_plan);
        m_plan = nullptr;
        for (auto& pl : m_planes) {
            if (pl.kernel) wct_decon_kernel_destroy(pl.kernel);
            pl.kernel = nullptr;
            pl.f_base.release();
            pl.f_tighter.release();
            pl.f_loose_alt.release();
            for (auto& b : pl.f_extra) b.release();
            pl.cf_rows.release();
            pl.has_tighter = false;
            for (auto& h : pl.has_extra) h = false;
            pl.active = false;
        }
        for (DevBuf* b : {&m_d_adc, &m_d_out, &m_d_mask, &m_d_rms, &m_d_bad, &m_d_ws, &m_d_cf, &m_d_alt}) b->release();
        m_cf_ld = 0;
        m_alt_bytes = 0;
        for (HostBuf* b : {&m_h_adc, &m_h_out, &m_h_mask, &m_h_rms, &m_h_alt}) b->release();
        for (int k = 0; k < kNExtra; ++k) {
            m_d_extra[k].release();
            m_h_extra[k].release();
            m_any_extra[k] = false;
        }
        if (m_stream) wct_stream_destroy(m_stream);
        m_stream = nullptr;
        m_built_nticks = -1;
        m_bad_layout.clear();
    }

    // Response spectra and buffers for frames of nticks at this tick
    // (init_overall_response, :707-865).
    void prepare(int nticks, double period)
    {
        if (nticks == m_built_nticks && period == m_built_period) return;
        release();
        hcheck(wct_set_device(m_device), "set device");
        hcheck(wct_stream_create(&m_stream), "stream");

        auto ifr = Factory::find_tn<IFieldResponse>(m_field_response);
        const FieldResponse& fr = ifr->field_response();
        if (fr.planes.size() < 3) throw SigProcError("HipOmnibusSigProc: field response with fewer than 3 planes");
        std::vector<RegionAverage> avg(3);
        for (int ip = 0; ip < 3; ++ip) avg[ip] = wire_region_average(fr.planes[ip]);
        if (avg[0].rows.empty()) throw SigProcError("HipOmnibusSigProc: empty region-averaged field response");
        // The reference sizes every plane's response by plane 0 (:760-761).
        const int fine_nwires = (int)avg[0].rows.size();
        const int nin = (int)avg[0].rows[0].size();
        const int nfine = (int)fft_best_length((size_t)nin);
        const Binning tbins(nfine, 0, nfine * fr.period);
        std::vector<double> elec;
        auto iwf = Factory::find_tn<IWaveform>(m_elecresponse_tn);
        if (auto ce = std::dynamic_pointer_cast<ColdElecResponse>(iwf))
            elec = ce->samples_f64(tbins);
        else {
            auto f = iwf->waveform_samples(tbins);
            elec.assign(f.begin(), f.end());
        }
        const double intrinsic = fr.origin / fr.speed;
        const int time_shift = (int)((m_coarse_time_offset + intrinsic) / period);
        const int wire_shift = (fine_nwires - 1) / 2;

        static const char* wiener_names[3] = {"Wiener_tight_U", "Wiener_tight_V", "Wiener_tight_W"};
        const bool refine = m_use_roi_refinement, debug = m_use_roi_debug_mode;
        size_t off = 0, rms_off = 0;
        for (int ip : m_process_planes) {
            Plane& pl = m_planes[ip];
            if (pl.active) continue;   // listed twice
            pl.nw = (int)m_channels[ip].size();
            if (pl.nw == 0) continue;
            const bool ind = ip < 2;
            // the reference saves nothing unless it refines ROIs or runs
            // the debug outputs (:1452-1620)
            pl.tight = !ind || refine;
            if (!refine && !debug) continue;
            const auto& rows = avg[ip].rows;
            if ((int)rows.size() > fine_nwires)
                throw SigProcError("HipOmnibusSigProc: plane " + std::to_string(ip) +
                                   " has more response regions than plane 0");
            std::vector<double> fr_rows((size_t)fine_nwires * nin, 0.0);
            for (size_t r = 0; r < rows.size(); ++r) {
                if ((int)rows[r].size() != nin) throw SigProcError("HipOmnibusSigProc: ragged field response paths");
                std::copy(rows[r].begin(), rows[r].end(), fr_rows.begin() + r * nin);
            }
            std::vector<double> resp((size_t)fine_nwires * nticks);
            hcheck(wct_overall_response(fr_rows.data(), fine_nwires, nin, elec.data(), nfine, fr.period, period,
                                        m_inter_gain, m_ADC_mV, m_fine_time_offset, nticks, resp.data(),
                                        m_stream),
                   "response");
            // The time filters of this plane's outputs, in FP64: the
            // products the reference forms in float (:1082-1102,
            // :1166-1213, :1355-1364).
            const auto wiener = filter_f64("HfFilter", wiener_names[ip], nticks);
            auto lf = [&](const char* name) { return filter_f64("LfFilter", name, nticks); };
            auto times = [](std::vector<double> a, const std::vector<double>& b) {
                for (size_t i = 0; i < a.size(); ++i) a[i] *= b[i];
                return a;
            };
            std::vector<double> f_base, f_tighter, f_extra[kNExtra];
            if (pl.tight) {
                f_base = ind ? times(wiener, lf("ROI_tight_lf")) : wiener;
                if (ind) {
                    f_tighter = times(wiener, lf("ROI_tighter_lf"));
                    pl.has_tighter = true;
                }
            } else {
                f_base.assign(nticks, 1.0);   // decon_2D_init's frame: no time filter
            }
            if (debug && ind) {
                f_extra[kLooseDbg] = times(wiener, lf("ROI_loose_lf"));
                pl.has_extra[kLooseDbg] = true;
                pl.extra_restore[kLooseDbg] = 1;
            }
            if (refine || (debug && !m_decon_charge_tag.empty())) {
                f_extra[kCharge] = filter_f64("HfFilter", "Gaus_wide", nticks);
                pl.has_extra[kCharge] = true;
                pl.extra_restore[kCharge] = ind ? 0 : 1;   // :1377-1379
            }
            std::vector<double> f_alt;
            if (refine) {
                static const char* wide_names[3] = {"Wiener_wide_U", "Wiener_wide_V", "Wiener_wide_W"};
                if (ind) {   // decon_2D_looseROI (:1162-1238), decon_2D_ROI_refine (:1053-1075)
                    f_extra[kLoose] = times(wiener, lf("ROI_loose_lf"));
                    f_alt = times(wiener, lf("ROI_tight_lf"));
                    pl.has_extra[kLoose] = true;
                    pl.extra_restore[kLoose] = 1;
                    f_extra[kRefine] = wiener;
                    pl.has_extra[kRefine] = true;
                    pl.extra_restore[kRefine] = 1;
                }
                f_extra[kHits] = filter_f64("HfFilter", wide_names[ip], nticks);   // decon_2D_hits (:1316-1347)
                pl.has_extra[kHits] = true;
                pl.extra_restore[kHits] = ind ? 0 : 1;
            }
            std::vector<double> env(f_base);
            for (const auto* f : {&f_tighter, &f_alt})
                for (size_t i = 0; i < f->size(); ++i) env[i] = std::max(env[i], std::fabs((*f)[i]));
            for (const auto& f : f_extra)
                for (size_t i = 0; i < f.size(); ++i) env[i] = std::max(env[i], std::fabs(f[i]));
            auto wire_f = filter_f64("HfFilter", ind ? "Wire_ind" : "Wire_col", pl.nw);
            hcheck(wct_decon_kernel_create_unfiltered(&pl.kernel, resp.data(), fine_nwires, pl.nw, nticks,
                                                      wire_f.data(), env.data(), wire_shift, time_shift,
                                                      m_spectral_cut, m_response_floor, m_stream),
                   "kernel");
            int knw = 0, knt = 0, kc = 0;
            hcheck(wct_decon_kernel_info(pl.kernel, &knw, &knt, &kc), "kernel");
            auto upload = [&](DevBuf& dst, const std::vector<double>& f) {
                std::vector<float> h(f.begin(), f.begin() + kc);
                dst.ensure(sizeof(float) * kc);
                hcheck(wct_memcpy_htod(dst.p, h.data(), sizeof(float) * kc, m_stream), "upload");
            };
            upload(pl.f_base, f_base);
            if (pl.has_tighter) upload(pl.f_tighter, f_tighter);
            for (int k = 0; k < kNExtra; ++k)
                if (pl.has_extra[k]) {
                    upload(pl.f_extra[k], f_extra[k]);
                    m_any_extra[k] = true;
                }
            if (pl.has_extra[kLoose]) {
                upload(pl.f_loose_alt, f_alt);
                pl.alt_off = m_alt_bytes;
                m_alt_bytes += pl.nw;
            }
            pl.th_factor = ind ? m_th_factor_ind : m_th_factor_col;
            pl.off = off;
            pl.rms_off = rms_off;
            off += (size_t)pl.nw * nticks;
            rms_off += pl.nw;
            pl.active = true;
        }
        if (!m_per_chan_resp.empty()) build_channel_factors(nticks, period);
        m_adc_floats = off;
        m_rms_floats = rms_off;
        m_nticks = nticks;
        m_d_adc.ensure(sizeof(float) * std::max<size_t>(off, 1));
        m_d_out.ensure(sizeof(float) * std::max<size_t>(off, 1));
        m_d_mask.ensure(std::max<size_t>(off, 1));
        m_d_rms.ensure(sizeof(float) * std::max<size_t>(rms_off, 1));
        m_h_adc.ensure(sizeof(float) * std::max<size_t>(off, 1));
        m_h_out.ensure(sizeof(float) * std::max<size_t>(off, 1));
        m_h_mask.ensure(std::max<size_t>(off, 1));
        m_h_rms.ensure(sizeof(float) * std::max<size_t>(rms_off, 1));
        if (m_alt_bytes) {
            m_d_alt.ensure(m_alt_bytes);
            m_h_alt.ensure(m_alt_bytes);
        }
        for (int k = 0; k < kNExtra; ++k)
            if (m_any_extra[k]) {
                m_d_extra[k].ensure(sizeof(float) * std::max<size_t>(off, 1));
                m_h_extra[k].ensure(sizeof(float) * std::max<size_t>(off, 1));
            }
        m_built_nticks = nticks;
        m_built_period = period;
    }

    // The channel-by-channel electronics correction of decon_2D_init
    // (:949-982): the nominal electronics response sampled over nticks
    // from the channel responses' start, one table row of factors per
    // distinct channel response (built in FP64 on the device), and per
    // plane the table row of each wire.  A channel without a response
    // throws the response's KeyError, as in the reference.
    void build_channel_factors(int nticks, double period)
    {
        auto cr = Factory::find_tn<IChannelResponse>(m_per_chan_resp);
        const Binning crb = cr->channel_response_binning();
        if (crb.binsize() != period)
            throw SigProcError("OmnibusSigProc::decon_2D_init: channel response size mismatch");
        const Binning tb(nticks, crb.minval, crb.minval + nticks * period);
        std::vector<double> ewave;
        auto iwf = Factory::find_tn<IWaveform>(m_elecresponse_tn);
        if (auto ce = std::dynamic_pointer_cast<ColdElecResponse>(iwf))
            ewave = ce->samples_f64(tb);
        else {
            auto f = iwf->waveform_samples(tb);
            ewave.assign(f.begin(), f.end());
        }
        std::map<std::vector<float>, int> index;
        std::vector<const std::vector<float>*> uniq;
        size_t nsamp = 1;
        for (int ip = 0; ip < 3; ++ip) {
            Plane& pl = m_planes[ip];
            if (!pl.active) continue;
            std::vector<int> rows(pl.nw);
            for (int w = 0; w < pl.nw; ++w) {
                const auto& r = cr->channel_response(m_channels[ip][w]);
                auto it = index.find(r);
                if (it == index.end()) {
                    it = index.emplace(r, (int)uniq.size()).first;
                    uniq.push_back(&it->first);
                    nsamp = std::max(nsamp, r.size());
                }
                rows[w] = it->second;
            }
            pl.cf_rows.ensure(sizeof(int) * rows.size());
            hcheck(wct_memcpy_htod(pl.cf_rows.p, rows.data(), sizeof(int) * rows.size(), m_stream), "upload");
        }
        if (uniq.empty()) return;
        std::vector<float> h(uniq.size() * nsamp, 0.0f);
        for (size_t i = 0; i < uniq.size(); ++i) std::copy(uniq[i]->begin(), uniq[i]->end(), h.begin() + i * nsamp);
        const int ld = nticks / 2 + 1;
        m_d_cf.ensure(sizeof(float) * 2 * (size_t)ld * uniq.size());
        hcheck(wct_channel_factors((float*)m_d_cf.p, ld, h.data(), (int)uniq.size(), (int)nsamp, ewave.data(),
                                   nticks, m_stream),
               "channel factors");
        m_cf_ld = ld;
    }

    std::vector<double> filter_f64(const char* type, const char* name, int n) const
    {
        auto f = Factory::find<IFilterWaveform>(type, name);
        if (auto hf = std::dynamic_pointer_cast<HfFilter>(f)) return hf->filter_waveform_f64(n);
        if (auto lf = std::dynamic_pointer_cast<LfFilter>(f)) return lf->filter_waveform_f64(n);
        auto v = f->filter_waveform(n);
        return std::vector<double>(v.begin(), v.end());
    }

    void make_plan(const std::vector<size_t>& ptr_off, const std::vector<size_t>& lo_off,
                   const std::vector<size_t>& hi_off)
    {
        if (m_plan) wct_decon_plan_destroy(m_plan);
        m_plan = nullptr;
        std::vector<wct_decon_job> jobs;
        const int* dbad = (const int*)m_d_bad.p;
        for (int ip = 0; ip < 3; ++ip) {
            const Plane& pl = m_planes[ip];
            if (!pl.active) continue;
            wct_decon_job j;
            std::memset(&j, 0, sizeof(j));
            j.kernel = pl.kernel;
            j.adc = (const float*)m_d_adc.p + pl.off;
            j.ld_adc = m_nticks;
            j.out = (float*)m_d_out.p + pl.off;
            j.ld_out = m_nticks;
            j.mask = (unsigned char*)m_d_mask.p + pl.off;
            j.ld_mask = m_nticks;
            j.rms = (float*)m_d_rms.p + pl.rms_off;
            j.bad_ptr = dbad + ptr_off[ip];
            j.bad_lo = dbad + lo_off[ip];
            j.bad_hi = dbad + hi_off[ip];
            j.lf_tight = (const float*)pl.f_base.p;
            j.lf_tighter = pl.has_tighter ? (const float*)pl.f_tighter.p : nullptr;
            if (m_cf_ld) {
                j.chan_factors = (const float*)m_d_cf.p;
                j.ld_chan_factors = m_cf_ld;
                j.chan_factor_row = (const int*)pl.cf_rows.p;
            }
            j.th_factor = pl.th_factor;
            j.restore_baseline = pl.tight ? 1 : 0;
            // the tight decon is saved only in the debug mode (negatives
            // kept) and is the collection plane's refinement input
            j.save_mode = 0;
            if (!pl.tight) {   // the unfiltered frame: no threshold, no ROIs
                j.mask = nullptr;
                j.rms = nullptr;
            }
            std::vector<wct_decon_output>& outs = m_outputs[ip];
            outs.clear();
            for (int k = 0; k < kNExtra; ++k) {
                if (!pl.has_extra[k]) continue;
                wct_decon_output o;
                std::memset(&o, 0, sizeof(o));
                o.filter = (const float*)pl.f_extra[k].p;
                if (k == kLoose) {
                    o.filter_alt = (const float*)pl.f_loose_alt.p;
                    o.alt_rows = (const unsigned char*)m_d_alt.p + pl.alt_off;
                }
                o.out = (float*)m_d_extra[k].p + pl.off;
                o.ld_out = m_nticks;
                o.restore_baseline = pl.extra_restore[k];
                o.save_mode = 0;   // saved (and zeroed) on the host, after apply_roi
                outs.push_back(o);
            }
            j.outputs = outs.empty() ? nullptr : outs.data();
            j.noutputs = (int)outs.size();
            jobs.push_back(j);
        }
        const size_t ws = wct_decon_workspace_size(jobs.data(), (int)jobs.size());
        m_d_ws.ensure(std::max<size_t>(ws, 1));
        hcheck(wct_decon_plan_create(&m_plan, jobs.data(), (int)jobs.size(), m_d_ws.p, m_d_ws.bytes), "plan");
    }

    // configuration (reference keys and defaults,
    // sigproc/inc/WireCellSigProc/OmnibusSigProc.h:100-234, except where
    // noted: the anode stand-in and the dft default name this build's
    // components, per_chan_resp defaults to "" as the PDSP layer sets it)
    std::string m_anode_tn{"AnodeChannels"};
    std::string m_dft_tn{"HipDFT"};
    std::string m_field_response{"FieldResponse"};
    std::string m_elecresponse_tn{"ColdElecResponse"};
    std::string m_per_chan_resp{""};
    double m_fine_time_offset{0.0 * units::microsecond};
    double m_coarse_time_offset{-8.0 * units::microsecond};
    double m_inter_gain{1.2};
    double m_ADC_mV{4096 / (2000. * units::mV)};
    float m_th_factor_ind{3};
    float m_th_factor_col{5};
    int m_pad{5};
    float m_asy{0.1f};
    std::vector<int> m_process_planes{0, 1, 2};
    std::string m_wiener_tag{"wiener"};
    std::string m_gauss_tag{"gauss"};
    std::string m_frame_tag{"sigproc"};
    std::string m_tight_lf_tag{"tight_lf"};
    std::string m_loose_lf_tag{"loose_lf"};
    std::string m_decon_charge_tag{"decon_charge"};
    bool m_use_roi_debug_mode{false};
    bool m_use_roi_refinement{true};
    bool m_sparse{false};
    std::string m_tight_roi_mask{"tight_roi"};
    LooseParams m_loose;
    RefineParams m_refine;
    double m_spectral_cut{1e-7};
    double m_response_floor{0.0};
    int m_device{0};

    std::vector<int> m_channels[3];
    std::unordered_map<int, std::pair<int, int>> m_chmap;

    Plane m_planes[3];
    void* m_stream{nullptr};
    wct_decon_plan m_plan{nullptr};
    DevBuf m_d_adc, m_d_out, m_d_mask, m_d_rms, m_d_bad, m_d_ws;
    DevBuf m_d_cf;       // per-channel correction table
    int m_cf_ld{0};      // its row stride (complex values), 0 without the correction
    HostBuf m_h_adc, m_h_out, m_h_mask, m_h_rms;
    DevBuf m_d_extra[kNExtra];
    HostBuf m_h_extra[kNExtra];
    bool m_any_extra[kNExtra]{};
    DevBuf m_d_alt;      // loose decon: 1 on rows that take the tight LF (all planes)
    HostBuf m_h_alt;
    size_t m_alt_bytes{0};
    std::vector<wct_decon_output> m_outputs[3];
    size_t m_adc_floats{0}, m_rms_floats{0};
    int m_nticks{0};
    int m_built_nticks{-1};
    double m_built_period{0};
    std::vector<size_t> m_bad_layout;
};

}  // namespace Hip
}  // namespace WireCell

WIRECELL_HIP_FACTORY(HipOmnibusSigProc, WireCell::Hip::HipOmnibusSigProc)
