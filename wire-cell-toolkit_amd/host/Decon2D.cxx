// The tensor-set deconvolution pair on the device:
//
//   Decon2DResponse  sig/src/Decon2DResponse.cxx:28-415
//     waveform tensor {nchans, nticks} f4  ->  its 2D deconvolution as a
//     complex time spectrum {nchans, nticks} c8 (no time filter), from
//     the spectrum mode of the fused plan (wct_decon_job.spectrum), with
//     the channel-by-channel electronics correction when per_chan_resp
//     names an IChannelResponse
//   Decon2DFilter    sig/src/Decon2DFilter.cxx:22-172
//     that spectrum times a product of IFilterWaveform time filters,
//     inverse real transform, Sig::restore_baseline, bad:cmm_range
//     ranges zeroed  ->  waveform tensor {nchans, nticks} f4, from
//     wct_spectrum_filter
//
// Tensors hold their elements column by column as the reference's do,
// so each call transposes on the device on the way in and out.  Both
// components keep their stream, device buffers and (Decon2DResponse) the
// response inverse of the last frame shape, which the reference rebuilds
// on every call (the response depends only on the configuration, the
// tick and the frame shape).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/wct_hip.h"
#include "Buffers.h"
#include "Components.h"
#include "Factory.h"
#include "Interface.h"
#include "Response.h"
#include "Tensor.h"

namespace WireCell {
namespace Hip {

struct ValueError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

namespace {

// The component "type:name", or null when the factory has none.
template <class IFACE>
std::shared_ptr<IFACE> find_or_null(const std::string& tn)
{
    try {
        return Factory::find_tn<IFACE>(tn);
    } catch (const FactoryException&) {
        return nullptr;
    }
}

std::vector<double> filter_f64(const IFilterWaveform::pointer& f, int n)
{
    if (auto hf = std::dynamic_pointer_cast<HfFilter>(f)) return hf->filter_waveform_f64(n);
    auto v = f->filter_waveform(n);
    return std::vector<double>(v.begin(), v.end());
}

// The 2D waveform tensor of the tag and its shape, or a ValueError as
// the reference raises (Decon2DResponse.cxx:214-218).
ITensor::pointer waveform_tensor(const ITensorSet::pointer& in, const std::string& tag, const char* who,
                                 const char* dt, size_t& nrows, size_t& ncols)
{
    auto t = Aux::get_tens(in, tag, "waveform");
    if (!t) throw ValueError(std::string(who) + ": no waveform tensor for tag \"" + tag + "\"");
    const auto shape = t->shape();
    if (shape.size() != 2) throw ValueError(std::string(who) + ": waveform tensor is not 2D");
    if (t->dtype() != dt)
        throw ValueError(std::string(who) + ": waveform tensor dtype " + t->dtype() + ", expected " + dt);
    nrows = shape[0];
    ncols = shape[1];
    if (t->size() < nrows * ncols * t->element_size())
        throw ValueError(std::string(who) + ": waveform tensor store smaller than its shape");
    return t;
}

// The output set: the new waveform, the input's channels (and summary),
// the input set's metadata with "tags" = [tag] (Decon2DResponse.cxx:
// 392-410).
ITensorSet::pointer output_set(const ITensorSet::pointer& in, const std::string& tag,
                               std::shared_ptr<Aux::SimpleTensor> owf, const ITensor::pointer& iwf,
                               std::vector<ITensor::pointer> extra)
{
    const Configuration iwf_md = iwf->metadata();
    Configuration& md = owf->metadata();
    md["tag"] = tag;
    md["pad"] = get<double>(iwf_md, "pad", 0.0);
    md["tbin"] = get<double>(iwf_md, "tbin", 0.0);
    md["type"] = "waveform";
    auto itv = std::make_shared<ITensor::vector>();
    itv->push_back(owf);
    for (auto& t : extra)
        if (t) itv->push_back(t);
    Configuration set_md = in->metadata();
    if (!set_md.isObject()) set_md = Configuration::makeObject();
    Configuration tags = Configuration::makeArray();
    tags.append(Configuration(tag));
    set_md["tags"] = tags;
    return std::make_shared<Aux::SimpleTensorSet>(in->ident(), set_md, itv);
}

}  // namespace

class Decon2DResponse : public ITensorSetFilter, public IConfigurable {
  public:
    ~Decon2DResponse() override { release(); }

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["dft"] = "FftwDFT";
        return cfg;
    }

    // Decon2DResponse.cxx:42-73 (the anode, the field response and, unless
    // per_chan_resp is "", the channel response must exist).
    void configure(const Configuration& cfg) override
    {
        m_cfg = cfg;
        const auto anode = get<std::string>(cfg, "anode", "Anode");
        if (!find_or_null<IAnodeChannels>(anode))
            throw ValueError("Decon2DResponse::configure: no anode \"" + anode + "\"");
        const auto cresp = get<std::string>(cfg, "per_chan_resp", "PerChannelResponse");
        m_cresp = nullptr;
        if (!cresp.empty()) {
            m_cresp = find_or_null<IChannelResponse>(cresp);
            if (!m_cresp) throw ValueError("Sig::Decon2DResponse::configure !m_cresp");
        }
        m_fresp_tn = get<std::string>(cfg, "field_response", "FieldResponse");
        if (!find_or_null<IFieldResponse>(m_fresp_tn))
            throw ValueError("Decon2DResponse::configure: no field response \"" + m_fresp_tn + "\"");
        m_device = get<int>(cfg, "device", 0);
        release();
    }

    bool operator()(const input_pointer& in, output_pointer& out) override
    {
        if (!in) {
            out = nullptr;
            return true;
        }
        const auto tag = get<std::string>(m_cfg, "tag", "trace_tag");
        size_t nchans = 0, nticks = 0;
        auto iwf = waveform_tensor(in, tag, "Decon2DResponse", "f4", nchans, nticks);
        auto ch = Aux::get_tens(in, tag, "channels");
        if (!ch) throw ValueError("Decon2DResponse: no channels tensor for tag \"" + tag + "\"");
        auto summary = Aux::get_tens(in, tag, "summary");
        const double period = get<double>(in->metadata(), "tick", 0.5 * units::microsecond);
        prepare((int)nchans, (int)nticks, period);
        if (m_cresp) {
            if (ch->dtype() != "i4" || ch->size() < nchans * sizeof(int))
                throw ValueError("Decon2DResponse: channels tensor is not {nchans} i4");
            channel_rows((const int*)ch->data(), (int)nchans, (int)nticks, period);
        }

        const size_t n = nchans * nticks;
        std::memcpy(m_h_io.p, iwf->data(), n * sizeof(float));
        device_check(wct_memcpy_htod_async(m_d_a.p, m_h_io.p, n * sizeof(float), m_stream), "upload");
        // column-major {nchans, nticks} is row-major {nticks, nchans}
        device_check(wct_dft_transpose_real((const float*)m_d_a.p, (float*)m_d_b.p, (int)nticks, (int)nchans,
                                            m_stream),
                     "transpose");
        device_check(wct_decon_plan_run(m_plan, m_stream), "decon");
        device_check(wct_dft_transpose_complex((const float*)m_d_spec.p, (float*)m_d_a.p, (int)nchans,
                                               (int)nticks, m_stream),
                     "transpose");
        device_check(wct_memcpy_dtoh_async(m_h_io.p, m_d_a.p, n * 2 * sizeof(float), m_stream), "download");
        device_check(wct_stream_sync(m_stream), "sync");
        auto owf = std::make_shared<Aux::SimpleTensor>("c8", ITensor::shape_t{nchans, nticks}, m_h_io.p);
        out = output_set(in, tag, owf, iwf, {ch, summary});
        return true;
    }

  private:
    // Response inverse and buffers for frames of nchans x nticks at this
    // tick (init_overall_response, Decon2DResponse.cxx:94-194).
    void prepare(int nchans, int nticks, double period)
    {
        if (nchans == m_nchans && nticks == m_nticks && period == m_period) return;
        release();
        if (nticks % 2) throw ValueError("Decon2DResponse: odd tick count " + std::to_string(nticks));
        device_check(wct_set_device(m_device), "set device");
        device_check(wct_stream_create(&m_stream), "stream");

        const int iplane = get<int>(m_cfg, "iplane", 0);
        if (iplane < 0 || iplane > 2) throw ValueError("Decon2DResponse: iplane " + std::to_string(iplane));
        const double gain = get<double>(m_cfg, "gain", 14.0 * units::mV / units::fC);
        const double shaping = get<double>(m_cfg, "shaping_time", 2.2 * units::microsecond);
        const double inter_gain = get<double>(m_cfg, "inter_gain", 1.2);
        const double adc_mv = get<double>(m_cfg, "ADC_mV", 4096 / (2000. * units::mV));
        const double ftoffset = get<double>(m_cfg, "ftoffset", 0.0);
        const double ctoffset = get<double>(m_cfg, "ctoffset", -8.0 * units::microsecond);

        const FieldResponse& fr = Factory::find_tn<IFieldResponse>(m_fresp_tn)->field_response();
        if ((int)fr.planes.size() <= iplane)
            throw ValueError("Decon2DResponse: field response has no plane " + std::to_string(iplane));
        const RegionAverage avg0 = wire_region_average(fr.planes[0]);
        const RegionAverage avg = iplane ? wire_region_average(fr.planes[iplane]) : avg0;
        if (avg0.rows.empty() || avg.rows.empty())
            throw ValueError("Decon2DResponse: empty region-averaged field response");
        // sized by plane 0, filled from plane iplane (:124-125, :156)
        const int fine_nwires = (int)avg0.rows.size();
        const int nin = (int)avg0.rows[0].size();
        if ((int)avg.rows.size() > fine_nwires)
            throw ValueError("Decon2DResponse: plane " + std::to_string(iplane) +
                             " has more response regions than plane 0");
        if (fine_nwires > nchans)
            throw ValueError("Decon2DResponse: " + std::to_string(nchans) + " channels, fewer than the " +
                             std::to_string(fine_nwires) + " response rows");
        std::vector<double> fr_rows((size_t)fine_nwires * nin, 0.0);
        for (size_t r = 0; r < avg.rows.size(); ++r) {
            if ((int)avg.rows[r].size() != nin) throw ValueError("Decon2DResponse: ragged field response paths");
            std::copy(avg.rows[r].begin(), avg.rows[r].end(), fr_rows.begin() + r * nin);
        }
        const int nfine = (int)fft_best_length((size_t)nin);
        // Response::ColdElec(gain, shaping) at bin centres (:131-134)
        const auto elec = coldelec_samples(nfine, 0.0, nfine * fr.period, gain, shaping, 1.0);
        std::vector<double> resp((size_t)fine_nwires * nticks);
        device_check(wct_overall_response(fr_rows.data(), fine_nwires, nin, elec.data(), nfine, fr.period, period,
                                          inter_gain, adc_mv, ftoffset, nticks, resp.data(), m_stream),
                     "response");
        static const char* wire_names[3] = {"Wire_ind", "Wire_ind", "Wire_col"};
        const auto wire_f = filter_f64(Factory::find<IFilterWaveform>("HfFilter", wire_names[iplane]), nchans);
        const std::vector<double> time_f(nticks, 1.0);
        const int wire_shift = (fine_nwires - 1) / 2;
        const int time_shift = (int)((ctoffset + fr.origin / fr.speed) / period);
        device_check(wct_decon_kernel_create(&m_kernel, resp.data(), fine_nwires, nchans, nticks, wire_f.data(),
                                             time_f.data(), wire_shift, time_shift, 0.0, m_stream),
                     "kernel");

        const size_t n = (size_t)nchans * nticks;
        m_d_a.ensure(n * 2 * sizeof(float));
        m_d_b.ensure(n * sizeof(float));
        m_d_spec.ensure(n * 2 * sizeof(float));
        m_h_io.ensure(n * 2 * sizeof(float));
        m_nchans = nchans;
        m_nticks = nticks;
        m_period = period;
        m_gain = gain;
        m_shaping = shaping;
        make_plan();
    }

    void make_plan()
    {
        if (m_plan) wct_decon_plan_destroy(m_plan);
        m_plan = nullptr;
        wct_decon_job j;
        std::memset(&j, 0, sizeof(j));
        j.kernel = m_kernel;
        j.adc = (const float*)m_d_b.p;
        j.ld_adc = m_nticks;
        j.spectrum = (float*)m_d_spec.p;
        j.ld_spectrum = m_nticks;
        if (m_cf_ld) {
            j.chan_factors = (const float*)m_d_cf.p;
            j.ld_chan_factors = m_cf_ld;
            j.chan_factor_row = (const int*)m_d_cfrows.p;
        }
        const size_t ws = wct_decon_workspace_size(&j, 1);
        m_d_ws.ensure(std::max<size_t>(ws, 1));
        device_check(wct_decon_plan_create(&m_plan, &j, 1, m_d_ws.p, m_d_ws.bytes), "plan");
    }

    // The channel-by-channel electronics correction (Decon2DResponse.cxx:
    // 280-308) for this frame's channels: a table row of factors per
    // distinct channel response, built in FP64 on the device the first
    // time the response is seen (the nominal response is
    // Response::ColdElec(gain, shaping_time) at the bin centres of nticks
    // bins from the channel responses' start), and the table row of each
    // channel.  A channel without a response throws the response's
    // KeyError.
    void channel_rows(const int* chans, int nchans, int nticks, double period)
    {
        const Binning crb = m_cresp->channel_response_binning();
        if (crb.binsize() != period)
            throw ValueError("Decon2DResponse::decon_2D_init: channel response size mismatch");
        std::vector<int> rows(nchans);
        bool grown = false;
        for (int i = 0; i < nchans; ++i) {
            const auto& r = m_cresp->channel_response(chans[i]);
            auto it = m_cf_index.find(r);
            if (it == m_cf_index.end()) {
                it = m_cf_index.emplace(r, (int)m_cf_list.size()).first;
                m_cf_list.push_back(&it->first);
                grown = true;
            }
            rows[i] = it->second;
        }
        if (grown || crb.minval != m_cf_tmin) {
            const Binning tb(nticks, crb.minval, crb.minval + nticks * period);
            const auto ewave = coldelec_samples(nticks, tb.minval, tb.maxval, m_gain, m_shaping, 1.0);
            size_t nsamp = 1;
            for (const auto* r : m_cf_list) nsamp = std::max(nsamp, r->size());
            std::vector<float> h(m_cf_list.size() * nsamp, 0.0f);
            for (size_t i = 0; i < m_cf_list.size(); ++i)
                std::copy(m_cf_list[i]->begin(), m_cf_list[i]->end(), h.begin() + i * nsamp);
            const int ld = nticks / 2 + 1;
            m_d_cf.ensure(sizeof(float) * 2 * (size_t)ld * m_cf_list.size());
            device_check(wct_channel_factors((float*)m_d_cf.p, ld, h.data(), (int)m_cf_list.size(), (int)nsamp,
                                             ewave.data(), nticks, m_stream),
                         "channel factors");
            m_d_cfrows.ensure(sizeof(int) * (size_t)nchans);
            m_cf_ld = ld;
            m_cf_tmin = crb.minval;
            make_plan();   // the table may have moved
        }
        device_check(wct_memcpy_htod(m_d_cfrows.p, rows.data(), sizeof(int) * (size_t)nchans, m_stream), "upload");
    }

    void release()
    {
        if (m_plan) wct_decon_plan_destroy(m_plan);
        if (m_kernel) wct_decon_kernel_destroy(m_kernel);
        if (m_stream) {
            wct_stream_sync(m_stream);
            wct_stream_destroy(m_stream);
        }
        m_plan = nullptr;
        m_kernel = nullptr;
        m_stream = nullptr;
        for (DevBuf* b : {&m_d_a, &m_d_b, &m_d_spec, &m_d_ws, &m_d_cf, &m_d_cfrows}) b->release();
        m_h_io.release();
        m_nchans = m_nticks = -1;
        m_period = 0;
        m_cf_index.clear();
        m_cf_list.clear();
        m_cf_ld = 0;
        m_cf_tmin = 0;
    }

    Configuration m_cfg;
    std::shared_ptr<IChannelResponse> m_cresp;
    std::map<std::vector<float>, int> m_cf_index;      // distinct channel responses -> table row
    std::vector<const std::vector<float>*> m_cf_list;  // in table order
    DevBuf m_d_cf, m_d_cfrows;
    int m_cf_ld{0};
    double m_cf_tmin{0};
    double m_gain{0}, m_shaping{0};
    std::string m_fresp_tn{"FieldResponse"};
    int m_device{0};
    int m_nchans{-1}, m_nticks{-1};
    double m_period{0};
    void* m_stream{nullptr};
    wct_decon_kernel m_kernel{nullptr};
    wct_decon_plan m_plan{nullptr};
    DevBuf m_d_a, m_d_b, m_d_spec, m_d_ws;
    HostBuf m_h_io;
};

class Decon2DFilter : public ITensorSetFilter, public IConfigurable {
  public:
    ~Decon2DFilter() override { release(); }

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["dft"] = "FftwDFT";
        return cfg;
    }

    void configure(const Configuration& cfg) override
    {
        m_cfg = cfg;
        m_device = get<int>(cfg, "device", 0);
        release();
    }

    bool operator()(const input_pointer& in, output_pointer& out) override
    {
        if (!in) {
            out = nullptr;
            return true;
        }
        const auto tag = get<std::string>(m_cfg, "tag", "trace_tag");
        size_t nchans = 0, nticks = 0;
        auto iwf = waveform_tensor(in, tag, "Decon2DFilter", "c8", nchans, nticks);
        auto ch = Aux::get_tens(in, tag, "channels");
        if (!ch) throw ValueError("Decon2DFilter: no channels tensor for tag \"" + tag + "\"");
        auto cmm_range = Aux::get_tens(in, tag, "bad:cmm_range");
        auto cmm_channel = Aux::get_tens(in, tag, "bad:cmm_channel");
        if (nticks % 2) throw ValueError("Decon2DFilter: odd tick count " + std::to_string(nticks));

        // the product of the configured time filters (:97-111)
        std::vector<float> filt(nticks, 1.0f);
        const Configuration flist = m_cfg["filters"];
        for (size_t i = 0; flist.isArray() && i < flist.size(); ++i) {
            const std::string tn = flist[(int)i].asString();
            auto fw = find_or_null<IFilterWaveform>(tn);
            if (!fw) throw ValueError("Decon2DFilter: no filter \"" + tn + "\"");
            const auto wave = fw->filter_waveform((int)nticks);
            for (size_t k = 0; k < wave.size() && k < nticks; ++k) filt[k] *= wave[k];
        }
        prepare((int)nchans, (int)nticks);

        const size_t n = nchans * nticks;
        std::memcpy(m_h_io.p, iwf->data(), n * 2 * sizeof(float));
        device_check(wct_memcpy_htod_async(m_d_a.p, m_h_io.p, n * 2 * sizeof(float), m_stream), "upload");
        std::memcpy(m_h_filt.p, filt.data(), (nticks / 2 + 1) * sizeof(float));
        device_check(wct_memcpy_htod_async(m_d_filt.p, m_h_filt.p, (nticks / 2 + 1) * sizeof(float), m_stream),
                     "upload");
        device_check(wct_dft_transpose_complex((const float*)m_d_a.p, (float*)m_d_b.p, (int)nticks, (int)nchans,
                                               m_stream),
                     "transpose");
        device_check(wct_spectrum_filter((const float*)m_d_b.p, (int)nticks, (int)nchans, (int)nticks,
                                         (const float*)m_d_filt.p, 1, (float*)m_d_a.p, (int)nticks, m_stream),
                     "filter");
        device_check(wct_dft_transpose_real((const float*)m_d_a.p, (float*)m_d_b.p, (int)nchans, (int)nticks,
                                            m_stream),
                     "transpose");
        auto owf = std::make_shared<Aux::SimpleTensor>("f4", ITensor::shape_t{nchans, nticks}, nullptr);
        float* y = reinterpret_cast<float*>(owf->store().data());
        device_check(wct_memcpy_dtoh_async(m_h_io.p, m_d_b.p, n * sizeof(float), m_stream), "download");
        device_check(wct_stream_sync(m_stream), "sync");
        std::memcpy(y, m_h_io.p, n * sizeof(float));

        // Sig::restore_baseline takes the median of a row's non-zero
        // samples, which Waveform::percentile refuses for an all-zero row
        // (util/src/Waveform.cxx:95-97).
        std::vector<char> nonzero(nchans, 0);
        for (size_t t = 0; t < nticks; ++t)
            for (size_t r = 0; r < nchans; ++r)
                if (y[r + t * nchans] != 0.0f) nonzero[r] = 1;
        for (size_t r = 0; r < nchans; ++r)
            if (!nonzero[r]) throw ValueError("Decon2DFilter: empty waveform (row " + std::to_string(r) + ")");

        // bad:cmm_range {nranges, 2} and bad:cmm_channel {nranges}, f8
        // (:126-145): row ch zeroed over [tmin, tmax); rows past the end
        // are skipped
        if (cmm_range && cmm_channel) {
            const auto rs = cmm_range->shape();
            const auto cs = cmm_channel->shape();
            if (cmm_range->dtype() != "f8" || cmm_channel->dtype() != "f8" || rs.size() != 2 || rs[1] != 2 ||
                cs.empty() || cs[0] != rs[0])
                throw ValueError("Decon2DFilter: bad:cmm_range must be {n, 2} f8 and bad:cmm_channel {n} f8");
            const size_t nr = rs[0];
            const double* ranges = reinterpret_cast<const double*>(cmm_range->data());
            const double* chans = reinterpret_cast<const double*>(cmm_channel->data());
            for (size_t i = 0; i < nr; ++i) {
                const double c = chans[i];
                if (!(c < (double)nchans)) continue;
                const double tmin = ranges[i], tmax = ranges[i + nr];
                if (!(c >= 0) || !(tmin >= 0) || !(tmin < tmax) || !(tmax <= (double)nticks))
                    throw ValueError("Decon2DFilter: cmm range " + std::to_string(i) + " out of order or bounds");
                const size_t r = (size_t)c;
                for (size_t t = (size_t)tmin; t < (size_t)tmax; ++t) y[r + t * nchans] = 0.0f;
            }
        }
        out = output_set(in, tag, owf, iwf, {ch});
        return true;
    }

  private:
    void prepare(int nchans, int nticks)
    {
        if (nchans == m_nchans && nticks == m_nticks) return;
        release();
        device_check(wct_set_device(m_device), "set device");
        device_check(wct_stream_create(&m_stream), "stream");
        const size_t n = (size_t)nchans * nticks;
        m_d_a.ensure(n * 2 * sizeof(float));
        m_d_b.ensure(n * 2 * sizeof(float));
        m_d_filt.ensure((nticks / 2 + 1) * sizeof(float));
        m_h_io.ensure(n * 2 * sizeof(float));
        m_h_filt.ensure((nticks / 2 + 1) * sizeof(float));
        m_nchans = nchans;
        m_nticks = nticks;
    }

    void release()
    {
        if (m_stream) {
            wct_stream_sync(m_stream);
            wct_stream_destroy(m_stream);
        }
        m_stream = nullptr;
        for (DevBuf* b : {&m_d_a, &m_d_b, &m_d_filt}) b->release();
        for (HostBuf* b : {&m_h_io, &m_h_filt}) b->release();
        m_nchans = m_nticks = -1;
    }

    Configuration m_cfg;
    int m_device{0};
    int m_nchans{-1}, m_nticks{-1};
    void* m_stream{nullptr};
    DevBuf m_d_a, m_d_b, m_d_filt;
    HostBuf m_h_io, m_h_filt;
};

}  // namespace Hip
}  // namespace WireCell

WIRECELL_HIP_FACTORY(Decon2DResponse, WireCell::Hip::Decon2DResponse)
WIRECELL_HIP_FACTORY(Decon2DFilter, WireCell::Hip::Decon2DFilter)
