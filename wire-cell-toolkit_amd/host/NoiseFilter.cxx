// The noise filter components (see NoiseFilter.h).  OmnibusNoiseFilter's
// control flow and masks are on the host, as in the reference; the
// per-channel filter of ProtoDUNE runs for all channels of a frame in one
// device batch (wct_nf_one_channel).
#include "NoiseFilter.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "../../include/wct_hip.h"
#include "Buffers.h"
#include "Factory.h"
#include "Frame.h"

namespace WireCell {

namespace Waveform {

BinRangeList merge(const BinRangeList& brl)
{
    BinRangeList tmp(brl.begin(), brl.end()), out;
    if (tmp.empty()) return out;
    std::sort(tmp.begin(), tmp.end());
    out.push_back(tmp[0]);
    for (size_t i = 1; i < tmp.size(); ++i) {
        if (out.back().second >= tmp[i].first) {
            out.back().second = tmp[i].second;
            continue;
        }
        out.push_back(tmp[i]);
    }
    return out;
}

BinRangeList merge(const BinRangeList& a, const BinRangeList& b)
{
    BinRangeList all(a);
    all.insert(all.end(), b.begin(), b.end());
    return merge(all);
}

ChannelMasks merge(const ChannelMasks& a, const ChannelMasks& b)
{
    ChannelMasks out = a;
    for (const auto& it : b) out[it.first] = merge(out[it.first], it.second);
    return out;
}

void merge(ChannelMaskMap& one, const ChannelMaskMap& two, const std::map<std::string, std::string>& name_map)
{
    for (const auto& it : two) {
        auto nm = name_map.find(it.first);
        const std::string mapped = nm == name_map.end() ? it.first : nm->second;
        auto dst = one.find(mapped);
        if (dst != one.end())
            dst->second = merge(dst->second, it.second);
        else
            one[mapped] = it.second;
    }
}

}  // namespace Waveform

namespace Hip {

namespace {

constexpr double kMillisecond = 1e6;   // ns

void ncheck(int rc, const std::string& what)
{
    if (rc != WCT_OK) throw std::runtime_error(what + ": " + wct_last_error());
}

// Plane index of a channel from the anode's plane lists (-1: not on it).
struct PlaneLookup {
    std::unordered_map<int, int> plane;
    void build(const IAnodeChannels& anode)
    {
        plane.clear();
        for (int p = 0; p < 3; ++p)
            for (int ch : anode.plane_channels(p)) plane[ch] = p;
    }
    int operator()(int ch) const
    {
        auto it = plane.find(ch);
        return it == plane.end() ? -1 : it->second;
    }
};

}  // namespace

// ---- OmniChannelNoiseDB (sigproc/src/OmniChannelNoiseDB.cxx) ----------
class OmniChannelNoiseDB : public IChannelNoiseDatabase, public IConfigurable {
  public:
    struct ChannelInfo {   // defaults of :30-50
        double nominal_baseline{0.0}, gain_correction{1.0}, response_offset{0.0};
        double min_rms_cut{0.5}, max_rms_cut{10.0};
        int pad_window_front{0}, pad_window_back{0};
        float decon_limit{0.02f}, decon_lf_cutoff{0.08f}, adc_limit{0.0f}, decon_limit1{0.08f};
        float protection_factor{5.0f}, min_adc_limit{50}, roi_min_max_ratio{0.8f};
        std::shared_ptr<filter_t> rcrc, config, noise, response;
    };

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["tick"] = m_tick;
        cfg["nsamples"] = m_nsamples;
        cfg["anode"] = m_anode_tn;
        cfg["field_response"] = "FieldResponse";
        cfg["dft"] = "HipDFT";
        cfg["groups"] = Configuration::makeArray();
        cfg["bad"] = Configuration::makeArray();
        cfg["channel_info"] = Configuration::makeArray();
        return cfg;
    }

    void configure(const Configuration& cfg) override
    {
        m_tick = get(cfg, "tick", m_tick);
        m_nsamples = get(cfg, "nsamples", m_nsamples);
        m_anode_tn = get<std::string>(cfg, "anode", m_anode_tn);
        m_anode = Factory::find_tn<IAnodeChannels>(m_anode_tn);
        m_db.clear();
        for (int p = 0; p < 3; ++p)
            for (int ch : m_anode->plane_channels(p)) m_db[ch] = ChannelInfo();
        m_groups.clear();
        for (const auto& jg : cfg["groups"]) {
            channel_group_t g;
            for (const auto& jch : jg) g.push_back(jch.asInt());
            m_groups.push_back(g);
        }
        m_bad.clear();
        for (const auto& jch : cfg["bad"]) m_bad.push_back(jch.asInt());
        std::sort(m_bad.begin(), m_bad.end());
        m_rc_layers = 2;
        m_rcrc_cache.clear();
        m_reconfig_cache.clear();
        m_response_cache.clear();
        m_unity.reset();
        for (const auto& jci : cfg["channel_info"]) update_channels(jci);
    }

    double sample_time() const override { return m_tick; }
    double nominal_baseline(int ch) const override { return dbget(ch).nominal_baseline; }
    double gain_correction(int ch) const override { return dbget(ch).gain_correction; }
    double min_rms_cut(int ch) const override { return dbget(ch).min_rms_cut; }
    double response_offset(int ch) const override { return dbget(ch).response_offset; }
    int pad_window_front(int ch) const override { return dbget(ch).pad_window_front; }
    int pad_window_back(int ch) const override { return dbget(ch).pad_window_back; }
    float coherent_nf_decon_limit(int ch) const override { return dbget(ch).decon_limit; }
    float coherent_nf_decon_lf_cutoff(int ch) const override { return dbget(ch).decon_lf_cutoff; }
    float coherent_nf_decon_limit1(int ch) const override { return dbget(ch).decon_limit1; }
    float coherent_nf_adc_limit(int ch) const override { return dbget(ch).adc_limit; }
    float coherent_nf_protection_factor(int ch) const override { return dbget(ch).protection_factor; }
    float coherent_nf_min_adc_limit(int ch) const override { return dbget(ch).min_adc_limit; }
    float coherent_nf_roi_min_max_ratio(int ch) const override { return dbget(ch).roi_min_max_ratio; }
    const filter_t& response(int ch) const override
    {
        static const filter_t dummy;
        const auto& f = dbget(ch).response;
        return f ? *f : dummy;
    }
    double max_rms_cut(int ch) const override { return dbget(ch).max_rms_cut; }
    const filter_t& rcrc(int ch) const override
    {
        static const filter_t dummy;
        const auto& f = dbget(ch).rcrc;
        return f ? *f : dummy;
    }
    const filter_t& config(int ch) const override
    {
        static const filter_t dummy;
        const auto& f = dbget(ch).config;
        return f ? *f : dummy;
    }
    const filter_t& noise(int ch) const override
    {
        static const filter_t dummy;
        const auto& f = dbget(ch).noise;
        return f ? *f : dummy;
    }
    const std::vector<channel_group_t> coherent_channels() const override { return m_groups; }
    std::vector<int> bad_channels() const override { return m_bad; }

  private:
    const ChannelInfo& dbget(int ch) const
    {
        auto it = m_db.find(ch);
        if (it == m_db.end()) throw std::out_of_range("OmniChannelNoiseDB: no db info for channel " + std::to_string(ch));
        return it->second;
    }
    ChannelInfo& get_ci(int ch)
    {
        auto it = m_db.find(ch);
        if (it == m_db.end()) throw std::out_of_range("OmniChannelNoiseDB: no db info for channel " + std::to_string(ch));
        return it->second;
    }

    // parse_channels (:85-138): a number, a list, {first, last} or {wpid}
    // (the low three bits of the plane id are the layer: 1 U, 2 V, 4 W).
    std::vector<int> parse_channels(const Configuration& j) const
    {
        std::vector<int> ret;
        if (j.isNumeric()) return {j.asInt()};
        if (j.isArray()) {
            for (const auto& c : j) ret.push_back(c.asInt());
            return ret;
        }
        if (j.isMember("first") && j.isMember("last")) {
            for (int c = j["first"].asInt(); c <= j["last"].asInt(); ++c) ret.push_back(c);
            return ret;
        }
        if (j.isMember("wpid")) {
            const int layer = j["wpid"].asInt() & 7;
            const int plane = layer == 1 ? 0 : layer == 2 ? 1 : layer == 4 ? 2 : -1;
            if (plane >= 0) ret = m_anode->plane_channels(plane);
        }
        return ret;
    }

    // parse_rcrc (:171-208): cached by round(1000 rcrc / ms) only, as the
    // reference's cache is (a later rc_layers for the same value reuses
    // the first spectrum).
    std::shared_ptr<filter_t> parse_rcrc(double val, int layers)
    {
        const int key = (int)std::round(1000 * val / kMillisecond);
        auto it = m_rcrc_cache.find(key);
        if (it != m_rcrc_cache.end()) return it->second;
        auto spec = std::make_shared<filter_t>(m_nsamples);
        DevBuf d;
        d.ensure(sizeof(float) * 2 * m_nsamples);
        void* s = nullptr;
        ncheck(wct_stream_create(&s), "OmniChannelNoiseDB stream");
        int rc = wct_nf_rcrc_spectrum(val, m_tick, m_nsamples, layers, (float*)d.p, s);
        if (rc == WCT_OK) rc = wct_memcpy_dtoh(spec->data(), d.p, sizeof(float) * 2 * m_nsamples, s);
        wct_stream_destroy(s);
        ncheck(rc, "OmniChannelNoiseDB rcrc spectrum");
        m_rcrc_cache[key] = spec;
        return spec;
    }

    // A flat unit spectrum of nsamples bins (default_filter, :141-150).
    std::shared_ptr<filter_t> unity()
    {
        if (!m_unity) m_unity = std::make_shared<filter_t>(m_nsamples, std::complex<float>(1, 0));
        return m_unity;
    }

    // The float spectrum of nsamples samples (FP64 transform on the device).
    filter_t spectrum_of(const std::vector<float>& wave) const
    {
        const int n = (int)wave.size();
        std::vector<double> buf(2 * (size_t)n, 0.0);
        for (int i = 0; i < n; ++i) buf[2 * i] = wave[i];
        DevBuf d;
        d.ensure(sizeof(double) * buf.size());
        void* s = nullptr;
        ncheck(wct_stream_create(&s), "OmniChannelNoiseDB stream");
        int rc = wct_memcpy_htod(d.p, buf.data(), sizeof(double) * buf.size(), s);
        if (rc == WCT_OK) rc = wct_dft_fwd1b_f64((const double*)d.p, (double*)d.p, 1, n, 1, s);
        if (rc == WCT_OK) rc = wct_memcpy_dtoh(buf.data(), d.p, sizeof(double) * buf.size(), s);
        wct_stream_destroy(s);
        ncheck(rc, "OmniChannelNoiseDB spectrum");
        filter_t out(n);
        for (int i = 0; i < n; ++i) out[i] = std::complex<float>((float)buf[2 * i], (float)buf[2 * i + 1]);
        return out;
    }

    // parse_reconfig / get_reconfig (:220-273): the ratio of the "to" and
    // "from" cold-electronics spectra (Response::ColdElec sampled at
    // i tick), cached by the four rounded parameters as the reference
    // caches them; an empty object gives the unit spectrum.
    std::shared_ptr<filter_t> parse_reconfig(const Configuration& j)
    {
        if (!j.isObject() || j.getMemberNames().empty()) return unity();
        const double mV_fC = units::mV / units::fC, us = units::microsecond;
        const double fg = j["from"]["gain"].asDouble(), fs = j["from"]["shaping"].asDouble();
        const double tg = j["to"]["gain"].asDouble(), ts = j["to"]["shaping"].asDouble();
        const int key = (int)std::round(10.0 * fg / mV_fC) << 24 | (int)std::round(10.0 * fs / us) << 16 |
                        (int)std::round(10.0 * tg / mV_fC) << 8 | (int)std::round(10.0 * ts / us);
        auto it = m_reconfig_cache.find(key);
        if (it != m_reconfig_cache.end()) return it->second;
        std::vector<float> to_sig(m_nsamples), from_sig(m_nsamples);
        const double dt = (m_nsamples * m_tick) / m_nsamples;
        for (int i = 0; i < m_nsamples; ++i) {
            to_sig[i] = (float)coldelec(i * dt, tg, ts);
            from_sig[i] = (float)coldelec(i * dt, fg, fs);
        }
        const filter_t to_f = spectrum_of(to_sig), from_f = spectrum_of(from_sig);
        auto filt = std::make_shared<filter_t>(m_nsamples);
        for (int i = 0; i < m_nsamples; ++i) (*filt)[i] = to_f[i] / from_f[i];
        m_reconfig_cache[key] = filt;
        return filt;
    }

    // parse_response (:302-350), the explicit-waveform form: the first
    // nsamples samples (float), zero-padded, as a spectrum, cached by
    // waveformid.  The form averaging the anode's field response ({wpid})
    // is refused (this build's anode has no field response).  An object
    // with neither gives an empty spectrum.
    std::shared_ptr<filter_t> parse_response(const Configuration& j)
    {
        if (j.isMember("wpid"))
            throw std::runtime_error("OmniChannelNoiseDB: a response from the anode's field response ({wpid}) is "
                                     "not built; give {waveform, waveformid}");
        if (!(j.isMember("waveform") && j.isMember("waveformid"))) return std::make_shared<filter_t>();
        const int id = j["waveformid"].asInt();
        auto it = m_response_cache.find(id);
        if (it != m_response_cache.end()) return it->second;
        std::vector<float> wave(m_nsamples, 0.f);
        const int nsamp = std::min(m_nsamples, (int)j["waveform"].size());
        for (int i = 0; i < nsamp; ++i) wave[i] = (float)j["waveform"][i].asDouble();
        auto spec = std::make_shared<filter_t>(spectrum_of(wave));
        m_response_cache[id] = spec;
        return spec;
    }

    // parse_freqmasks (:152-169)
    std::shared_ptr<filter_t> parse_freqmasks(const Configuration& j) const
    {
        auto spec = std::make_shared<filter_t>(m_nsamples, std::complex<float>(1, 0));
        for (const auto& one : j) {
            const double value = one["value"].asDouble();
            const int lo = std::max(one["lobin"].asInt(), 0);
            const int hi = std::min(one["hibin"].asInt(), m_nsamples - 1);
            for (int i = lo; i <= hi; ++i) (*spec)[i] = (float)value;
        }
        return spec;
    }

    // update_channels (:376-573): the keys the one-channel filters and the
    // accessors above read; the response spectra (read by the coherent
    // filter) are not kept.
    void update_channels(const Configuration& cfg)
    {
        auto chans = parse_channels(cfg["channels"]);
        auto setd = [&](const char* key, auto member) {
            if (!cfg.isMember(key)) return;
            const double v = cfg[key].asDouble();
            for (int ch : chans) get_ci(ch).*member = v;
        };
        setd("nominal_baseline", &ChannelInfo::nominal_baseline);
        setd("gain_correction", &ChannelInfo::gain_correction);
        setd("response_offset", &ChannelInfo::response_offset);
        setd("min_rms_cut", &ChannelInfo::min_rms_cut);
        setd("max_rms_cut", &ChannelInfo::max_rms_cut);
        auto seti = [&](const char* key, int ChannelInfo::*member) {
            if (!cfg.isMember(key)) return;
            const int v = cfg[key].asInt();
            for (int ch : chans) get_ci(ch).*member = v;
        };
        seti("pad_window_front", &ChannelInfo::pad_window_front);
        seti("pad_window_back", &ChannelInfo::pad_window_back);
        auto setf = [&](const char* key, float ChannelInfo::*member) {
            if (!cfg.isMember(key)) return;
            const float v = (float)cfg[key].asDouble();
            for (int ch : chans) get_ci(ch).*member = v;
        };
        setf("decon_limit", &ChannelInfo::decon_limit);
        setf("decon_lf_cutoff", &ChannelInfo::decon_lf_cutoff);
        setf("decon_limit1", &ChannelInfo::decon_limit1);
        setf("adc_limit", &ChannelInfo::adc_limit);
        setf("protection_factor", &ChannelInfo::protection_factor);
        setf("min_adc_limit", &ChannelInfo::min_adc_limit);
        setf("roi_min_max_ratio", &ChannelInfo::roi_min_max_ratio);
        if (!cfg["rcrc"].isNull()) {
            if (cfg.isMember("rc_layers")) m_rc_layers = cfg["rc_layers"].asInt();
            auto val = parse_rcrc(cfg["rcrc"].asDouble(), m_rc_layers);
            for (int ch : chans) get_ci(ch).rcrc = val;
        }
        if (!cfg["reconfig"].isNull()) {
            auto val = parse_reconfig(cfg["reconfig"]);
            for (int ch : chans) get_ci(ch).config = val;
        }
        if (!cfg["response"].isNull()) {
            auto val = parse_response(cfg["response"]);
            for (int ch : chans) get_ci(ch).response = val;
        }
        if (!cfg["freqmasks"].isNull()) {
            auto val = parse_freqmasks(cfg["freqmasks"]);
            for (int ch : chans) get_ci(ch).noise = val;
        }
    }

    double m_tick{0.5 * 1000.0};
    int m_nsamples{9600};
    int m_rc_layers{2};
    std::string m_anode_tn{"AnodeChannels"};
    std::shared_ptr<IAnodeChannels> m_anode;
    std::map<int, ChannelInfo> m_db;
    std::vector<channel_group_t> m_groups;
    std::vector<int> m_bad;
    std::map<int, std::shared_ptr<filter_t>> m_rcrc_cache, m_reconfig_cache, m_response_cache;
    std::shared_ptr<filter_t> m_unity;
};

// ---- pdOneChannelNoise (sigproc/src/Protodune.cxx:712-923) and
//      mbOneChannelNoise (sigproc/src/Microboone.cxx:894-1069) -----------
class OneChannelNoise : public IChannelFilter, public IBatchChannelFilter, public IConfigurable {
  public:
    typedef IChannelNoiseDatabase::filter_t filter_t;
    explicit OneChannelNoise(int mode) : m_mode(mode) {}
    ~OneChannelNoise() override
    {
        if (m_stream) wct_stream_destroy(m_stream);
    }
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["anode"] = m_anode_tn;
        cfg["noisedb"] = m_noisedb_tn;
        cfg["dft"] = "HipDFT";
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_anode_tn = get<std::string>(cfg, "anode", m_anode_tn);
        m_anode = Factory::find_tn<IAnodeChannels>(m_anode_tn);
        m_planes.build(*m_anode);
        m_noisedb_tn = get<std::string>(cfg, "noisedb", m_noisedb_tn);
        m_noisedb = Factory::find_tn<IChannelNoiseDatabase>(m_noisedb_tn);
        m_resmp.clear();
        if (m_mode != WCT_NF_PROTODUNE) return;   // FEMB resampling is ProtoDUNE's
        for (const auto& jone : cfg["resmp"]) {
            const int smpin = jone["sample_from"].asInt();
            for (const auto& jch : jone["channels"]) m_resmp[jch.asInt()] = smpin;
        }
    }

    Waveform::ChannelMaskMap apply(int ch, signal_t& signal) const override
    {
        std::vector<float> rows(signal.begin(), signal.end());
        std::vector<Waveform::ChannelMaskMap> masks;
        apply_batch({ch}, rows, (int)signal.size(), masks);
        signal.assign(rows.begin(), rows.end());
        return masks[0];
    }
    Waveform::ChannelMaskMap apply(channel_signals_t&) const override { return Waveform::ChannelMaskMap(); }

    void apply_batch(const std::vector<int>& channels, std::vector<float>& rows, int nt,
                     std::vector<Waveform::ChannelMaskMap>& masks) const override
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        const int nch = (int)channels.size();
        masks.assign(nch, Waveform::ChannelMaskMap());
        if (nch == 0) return;
        if (!m_stream) ncheck(wct_stream_create(&m_stream), "pdOneChannelNoise stream");
        resample(channels, rows, nt);

        // per-channel entries, spectra rows and flagged runs, shared spectra once
        std::vector<wct_nf_channel> ent(nch);
        std::vector<const filter_t*> rc_rows;
        std::map<const filter_t*, int> rc_id;
        auto row_of = [&](const filter_t& f, const char* what, int ch) {
            if (f.empty()) return -1;
            if ((int)f.size() != nt) {
                if (m_mode == WCT_NF_PROTODUNE) return -1;   // :771-779: skipped (with a warning)
                throw std::runtime_error(std::string("mbOneChannelNoise: ") + what + " spectrum of channel " +
                                         std::to_string(ch) + " has " + std::to_string(f.size()) +
                                         " bins for " + std::to_string(nt) + " ticks");
            }
            auto it = rc_id.find(&f);
            if (it == rc_id.end()) {
                it = rc_id.emplace(&f, (int)rc_rows.size()).first;
                rc_rows.push_back(&f);
            }
            return it->second;
        };
        std::vector<int> bins;
        std::map<const filter_t*, std::pair<int, int>> bin_id;
        for (int i = 0; i < nch; ++i) {
            const int ch = channels[i];
            wct_nf_channel& e = ent[i];
            e.plane = m_planes(ch);
            e.min_rms = (float)m_noisedb->min_rms_cut(ch);
            e.max_rms = (float)m_noisedb->max_rms_cut(ch);
            e.mode = m_mode;
            e.rcrc = row_of(m_noisedb->rcrc(ch), "rcrc", ch);
            e.baseline = 0.f;
            e.gain = 1.f;
            e.config = e.noise = -1;
            e.bins_begin = e.bins_end = 0;
            if (m_mode == WCT_NF_MICROBOONE) {
                e.baseline = (float)m_noisedb->nominal_baseline(ch);
                e.gain = (float)m_noisedb->gain_correction(ch);
                e.config = row_of(m_noisedb->config(ch), "config", ch);
                e.noise = row_of(m_noisedb->noise(ch), "noise", ch);
                continue;
            }
            if (e.plane == 2) {
                const filter_t& nz = m_noisedb->noise(ch);
                auto it = bin_id.find(&nz);
                if (it == bin_id.end()) {
                    const int begin = (int)bins.size() / 2;
                    for (int k = 0; k < (int)nz.size(); ++k) {   // :791-806
                        if (nz[k].real() < 0.5f) {
                            if ((int)bins.size() / 2 > begin && bins.back() == k - 1)
                                bins.back() = k;
                            else {
                                bins.push_back(k);
                                bins.push_back(k);
                            }
                        }
                    }
                    it = bin_id.emplace(&nz, std::make_pair(begin, (int)bins.size() / 2)).first;
                }
                e.bins_begin = it->second.first;
                e.bins_end = it->second.second;
            }
        }
        const size_t row_bytes = sizeof(float) * (size_t)nch * nt;
        m_d_in.ensure(row_bytes);
        m_d_out.ensure(row_bytes);
        m_d_chans.ensure(sizeof(wct_nf_channel) * nch);
        m_d_flags.ensure(sizeof(int) * 3 * nch);
        m_d_rcrc.ensure(sizeof(float) * 2 * std::max<size_t>(1, rc_rows.size()) * nt);
        m_d_bins.ensure(sizeof(int) * std::max<size_t>(2, bins.size()));
        const size_t ws = wct_nf_workspace_size(nch, nt);
        m_d_ws.ensure(ws);
        ncheck(wct_memcpy_htod_async(m_d_in.p, rows.data(), row_bytes, m_stream), "pdOneChannelNoise upload");
        ncheck(wct_memcpy_htod_async(m_d_chans.p, ent.data(), sizeof(wct_nf_channel) * nch, m_stream),
               "pdOneChannelNoise upload");
        for (size_t r = 0; r < rc_rows.size(); ++r)
            ncheck(wct_memcpy_htod_async((char*)m_d_rcrc.p + sizeof(float) * 2 * nt * r, rc_rows[r]->data(),
                                         sizeof(float) * 2 * nt, m_stream),
                   "pdOneChannelNoise upload");
        if (!bins.empty())
            ncheck(wct_memcpy_htod_async(m_d_bins.p, bins.data(), sizeof(int) * bins.size(), m_stream),
                   "pdOneChannelNoise upload");
        ncheck(wct_nf_one_channel((const float*)m_d_in.p, nt, (float*)m_d_out.p, nt, nch, nt,
                                  (const wct_nf_channel*)m_d_chans.p, (const float*)m_d_rcrc.p,
                                  (const int*)m_d_bins.p, (int*)m_d_flags.p, m_d_ws.p, ws, m_stream),
               "pdOneChannelNoise");
        std::vector<int> flags(3 * (size_t)nch);
        ncheck(wct_memcpy_dtoh_async(rows.data(), m_d_out.p, row_bytes, m_stream), "pdOneChannelNoise download");
        ncheck(wct_memcpy_dtoh_async(flags.data(), m_d_flags.p, sizeof(int) * 3 * nch, m_stream),
               "pdOneChannelNoise download");
        ncheck(wct_stream_sync(m_stream), "pdOneChannelNoise sync");
        for (int i = 0; i < nch; ++i) {
            const int bits = flags[3 * i], ch = channels[i];
            const Waveform::BinRange all(0, nt);
            auto& m = masks[i];
            if (m_mode == WCT_NF_MICROBOONE) {   // Microboone.cxx:924-935, 1007-1016, 1044-1061
                const Waveform::BinRange chirped(flags[3 * i + 1], flags[3 * i + 2]);
                if (bits & WCT_NF_CHIRP) {
                    m["chirp"][ch].push_back(chirped);
                    if (ent[i].plane != 2 && (chirped.first > 0 || chirped.second < nt))
                        m["lf_noisy"][ch].push_back(chirped);
                }
                if ((bits & WCT_NF_PARTIAL) && ent[i].plane != 2) m["lf_noisy"][ch].push_back(all);
                if (bits & WCT_NF_NOISY) {
                    m["noisy"][ch].push_back(all);
                    auto lf = m.find("lf_noisy");
                    if (lf != m.end()) {
                        lf->second.erase(ch);
                    }
                }
                continue;
            }
            if (bits & WCT_NF_HARMONIC) m["harmonic"][ch].push_back(all);
            if ((bits & WCT_NF_PARTIAL) && ent[i].plane != 2) m["lf_noisy"][ch].push_back(all);
            if (bits & WCT_NF_NOISY) m["noisy"][ch].push_back(all);
        }
    }

  private:
    // The FEMB clock correction (:760-768, FftScaling :532-552): the
    // channel's first sample_from samples resampled to nt, through the
    // device transforms.
    void resample(const std::vector<int>& channels, std::vector<float>& rows, int nt) const
    {
        std::map<int, std::vector<int>> by_len;
        for (int i = 0; i < (int)channels.size(); ++i) {
            auto it = m_resmp.find(channels[i]);
            if (it != m_resmp.end()) by_len[it->second].push_back(i);
        }
        for (const auto& bl : by_len) {
            const int nin = bl.first, k = (int)bl.second.size();
            if (nin <= 0) continue;
            std::vector<std::complex<float>> buf((size_t)k * std::max(nin, nt));
            for (int r = 0; r < k; ++r) {
                const float* src = rows.data() + (size_t)bl.second[r] * nt;
                for (int t = 0; t < nin; ++t) buf[(size_t)r * nin + t] = t < nt ? src[t] : 0.f;
            }
            DevBuf d;
            const size_t bytes = sizeof(std::complex<float>) * buf.size();
            d.ensure(bytes);
            ncheck(wct_memcpy_htod_async(d.p, buf.data(), sizeof(std::complex<float>) * k * nin, m_stream),
                   "resample upload");
            ncheck(wct_dft_fwd1b((const float*)d.p, (float*)d.p, k, nin, 1, m_stream), "resample fwd");
            std::vector<std::complex<float>> spec((size_t)k * nin);
            ncheck(wct_memcpy_dtoh_async(spec.data(), d.p, sizeof(std::complex<float>) * k * nin, m_stream),
                   "resample download");
            ncheck(wct_stream_sync(m_stream), "resample sync");
            const int mid = nin % 2 == 0 ? nin / 2 : (nin + 1) / 2;
            for (int r = 0; r < k; ++r) {
                std::vector<std::complex<float>> tran(nt, 0.f);
                const std::complex<float>* s = spec.data() + (size_t)r * nin;
                // resize to nt, then rotate [mid, nin) to the end
                std::vector<std::complex<float>> padded(s, s + std::min(nin, nt));
                padded.resize(nt, 0.f);
                std::rotate(padded.begin() + std::min(mid, nt), padded.begin() + std::min(nin, nt), padded.end());
                // inv_c2r's Hermitian mirror (Spectrum.h:56-89)
                padded[0] = padded[0].real();
                const int half = nt / 2;
                if (nt % 2 == 0) {
                    padded[half] = padded[half].real();
                    for (int j = 1; j < half; ++j) padded[nt - j] = std::conj(padded[j]);
                } else {
                    for (int j = 1; j <= half; ++j) padded[nt - j] = std::conj(padded[j]);
                }
                std::copy(padded.begin(), padded.end(), buf.begin() + (size_t)r * nt);
            }
            ncheck(wct_memcpy_htod_async(d.p, buf.data(), sizeof(std::complex<float>) * k * nt, m_stream),
                   "resample upload");
            ncheck(wct_dft_inv1b((const float*)d.p, (float*)d.p, k, nt, 1, m_stream), "resample inv");
            ncheck(wct_memcpy_dtoh_async(buf.data(), d.p, sizeof(std::complex<float>) * k * nt, m_stream),
                   "resample download");
            ncheck(wct_stream_sync(m_stream), "resample sync");
            const float scale = (float)(nt / nin);   // integer ratio, as Waveform::scale gets it
            for (int r = 0; r < k; ++r) {
                float* dst = rows.data() + (size_t)bl.second[r] * nt;
                for (int t = 0; t < nt; ++t) dst[t] = buf[(size_t)r * nt + t].real() * scale;
            }
        }
    }

    int m_mode;
    std::string m_anode_tn{"AnodeChannels"}, m_noisedb_tn{"OmniChannelNoiseDB"};
    std::shared_ptr<IAnodeChannels> m_anode;
    std::shared_ptr<IChannelNoiseDatabase> m_noisedb;
    PlaneLookup m_planes;
    std::map<int, int> m_resmp;
    mutable std::mutex m_mutex;
    mutable void* m_stream{nullptr};
    mutable DevBuf m_d_in, m_d_out, m_d_chans, m_d_flags, m_d_rcrc, m_d_bins, m_d_ws;
};

class PdOneChannelNoise : public OneChannelNoise {
  public:
    PdOneChannelNoise() : OneChannelNoise(WCT_NF_PROTODUNE) {}
};
class MbOneChannelNoise : public OneChannelNoise {
  public:
    MbOneChannelNoise() : OneChannelNoise(WCT_NF_MICROBOONE) {}
};

// ---- pdRelGainCalib (sigproc/src/Protodune.cxx:925-1014) ---------------
class RelGainCalib : public IChannelFilter, public IBatchChannelFilter, public IConfigurable {
  public:
    ~RelGainCalib() override
    {
        if (m_stream) wct_stream_destroy(m_stream);
    }
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["anode"] = m_anode_tn;
        cfg["noisedb"] = m_noisedb_tn;
        cfg["gain_def"] = (double)m_gain_def;
        cfg["gain_min_cut"] = (double)m_gain_min;
        cfg["gain_max_cut"] = (double)m_gain_max;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_anode_tn = get<std::string>(cfg, "anode", m_anode_tn);
        Factory::find_tn<IAnodeChannels>(m_anode_tn);
        m_noisedb_tn = get<std::string>(cfg, "noisedb", m_noisedb_tn);
        Factory::find_tn<IChannelNoiseDatabase>(m_noisedb_tn);
        m_gain_def = get(cfg, "gain_def", m_gain_def);
        m_gain_min = get(cfg, "gain_min_cut", m_gain_min);
        m_gain_max = get(cfg, "gain_max_cut", m_gain_max);
        m_rel_gain.clear();
        for (const auto& j : cfg["rel_gain"]) m_rel_gain.push_back((float)j.asDouble());
    }
    Waveform::ChannelMaskMap apply(int ch, signal_t& signal) const override
    {
        std::vector<float> rows(signal.begin(), signal.end());
        std::vector<Waveform::ChannelMaskMap> masks;
        apply_batch({ch}, rows, (int)signal.size(), masks);
        signal.assign(rows.begin(), rows.end());
        return Waveform::ChannelMaskMap();
    }
    Waveform::ChannelMaskMap apply(channel_signals_t&) const override { return Waveform::ChannelMaskMap(); }
    // The factor of channel ch is rel_gain[ch] (an array indexed by channel
    // number, :984); out of range throws as the reference's at() does.
    void apply_batch(const std::vector<int>& channels, std::vector<float>& rows, int nt,
                     std::vector<Waveform::ChannelMaskMap>& masks) const override
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        const int nch = (int)channels.size();
        masks.assign(nch, Waveform::ChannelMaskMap());
        if (nch == 0 || nt == 0) return;
        std::vector<float> g(nch);
        for (int i = 0; i < nch; ++i) g[i] = m_rel_gain.at(channels[i]);
        if (!m_stream) ncheck(wct_stream_create(&m_stream), "pdRelGainCalib stream");
        const size_t bytes = sizeof(float) * (size_t)nch * nt;
        m_d_rows.ensure(bytes);
        m_d_gain.ensure(sizeof(float) * nch);
        ncheck(wct_memcpy_htod_async(m_d_rows.p, rows.data(), bytes, m_stream), "pdRelGainCalib upload");
        ncheck(wct_memcpy_htod_async(m_d_gain.p, g.data(), sizeof(float) * nch, m_stream), "pdRelGainCalib upload");
        ncheck(wct_nf_rel_gain((float*)m_d_rows.p, nt, nch, nt, (const float*)m_d_gain.p, m_gain_min, m_gain_max,
                               m_stream),
               "pdRelGainCalib");
        ncheck(wct_memcpy_dtoh_async(rows.data(), m_d_rows.p, bytes, m_stream), "pdRelGainCalib download");
        ncheck(wct_stream_sync(m_stream), "pdRelGainCalib sync");
    }

  private:
    std::string m_anode_tn{"AnodeChannels"}, m_noisedb_tn{"OmniChannelNoiseDB"};
    float m_gain_def{1.0f}, m_gain_min{0.8f}, m_gain_max{1.25f};
    std::vector<float> m_rel_gain;
    mutable std::mutex m_mutex;
    mutable void* m_stream{nullptr};
    mutable DevBuf m_d_rows, m_d_gain;
};

// ---- mbADCBitShift (sigproc/src/Microboone.cxx:1071-1216) --------------
// Every channel of a frame in one device batch (wct_nf_adc_bit_shift); a
// corrected channel gets the mask ADCBitShift with the range (nshift,
// nshift), as the reference records it.
class ADCBitShift : public IChannelFilter, public IBatchChannelFilter, public IConfigurable {
  public:
    ~ADCBitShift() override
    {
        if (m_stream) wct_stream_destroy(m_stream);
    }
    Configuration default_configuration() const override   // :1088-1097
    {
        Configuration cfg;
        cfg["Number_of_ADC_bits"] = m_nbits;
        cfg["Exam_number_of_ticks_test"] = m_exam;
        cfg["Threshold_sigma_test"] = m_sigma;
        cfg["Threshold_fix"] = m_fix;
        return cfg;
    }
    void configure(const Configuration& cfg) override   // :1080-1086
    {
        m_nbits = get(cfg, "Number_of_ADC_bits", m_nbits);
        m_exam = get(cfg, "Exam_number_of_ticks_test", m_exam);
        m_sigma = get(cfg, "Threshold_sigma_test", m_sigma);
        m_fix = get(cfg, "Threshold_fix", m_fix);
        if (m_nbits < 1 || m_nbits > 31) throw ConfigError("mbADCBitShift: Number_of_ADC_bits outside 1..31");
        if (m_exam < 1) throw ConfigError("mbADCBitShift: Exam_number_of_ticks_test below 1");
    }
    Waveform::ChannelMaskMap apply(int ch, signal_t& signal) const override
    {
        std::vector<float> rows(signal.begin(), signal.end());
        std::vector<Waveform::ChannelMaskMap> masks;
        apply_batch({ch}, rows, (int)signal.size(), masks);
        signal.assign(rows.begin(), rows.end());
        return masks[0];
    }
    Waveform::ChannelMaskMap apply(channel_signals_t&) const override { return Waveform::ChannelMaskMap(); }
    void apply_batch(const std::vector<int>& channels, std::vector<float>& rows, int nt,
                     std::vector<Waveform::ChannelMaskMap>& masks) const override
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        const int nch = (int)channels.size();
        masks.assign(nch, Waveform::ChannelMaskMap());
        if (nch == 0 || nt == 0) return;
        if (m_exam > nt)   // the reference's at() would throw
            throw std::runtime_error("mbADCBitShift: Exam_number_of_ticks_test " + std::to_string(m_exam) +
                                     " above " + std::to_string(nt) + " ticks");
        if (!m_stream) ncheck(wct_stream_create(&m_stream), "mbADCBitShift stream");
        const size_t bytes = sizeof(float) * (size_t)nch * nt;
        m_d_rows.ensure(bytes);
        m_d_ns.ensure(sizeof(int) * nch);
        ncheck(wct_memcpy_htod_async(m_d_rows.p, rows.data(), bytes, m_stream), "mbADCBitShift upload");
        ncheck(wct_nf_adc_bit_shift((float*)m_d_rows.p, nt, nch, nt, m_nbits, m_exam, m_sigma, m_fix,
                                    (int*)m_d_ns.p, m_stream),
               "mbADCBitShift");
        std::vector<int> ns(nch);
        ncheck(wct_memcpy_dtoh_async(rows.data(), m_d_rows.p, bytes, m_stream), "mbADCBitShift download");
        ncheck(wct_memcpy_dtoh_async(ns.data(), m_d_ns.p, sizeof(int) * nch, m_stream), "mbADCBitShift download");
        ncheck(wct_stream_sync(m_stream), "mbADCBitShift sync");
        for (int i = 0; i < nch; ++i)
            if (ns[i]) masks[i]["ADCBitShift"][channels[i]].push_back({ns[i], ns[i]});
    }

  private:
    int m_nbits{12}, m_exam{500};
    double m_sigma{7.5}, m_fix{0.8};
    mutable std::mutex m_mutex;
    mutable void* m_stream{nullptr};
    mutable DevBuf m_d_rows, m_d_ns;
};

// ---- mbOneChannelStatus (sigproc/src/Microboone.cxx:1218-1334) ---------
// A status filter: channels off the collection plane whose low-frequency
// content (ID_lf_noisy) exceeds the cut are masked lf_noisy over the whole
// waveform; the samples are not changed.  Every tested channel of a frame
// in one device batch (wct_nf_lf_noisy).
class OneChannelStatus : public IChannelFilter, public IBatchChannelFilter, public IConfigurable {
  public:
    ~OneChannelStatus() override
    {
        if (m_stream) wct_stream_destroy(m_stream);
    }
    Configuration default_configuration() const override   // :1245-1255
    {
        Configuration cfg;
        cfg["Threshold"] = m_threshold;
        cfg["Window"] = m_window;
        cfg["Nbins"] = m_nbins;
        cfg["Cut"] = m_cut;
        cfg["anode"] = m_anode_tn;
        cfg["dft"] = "HipDFT";
        return cfg;
    }
    void configure(const Configuration& cfg) override   // :1229-1243
    {
        m_threshold = get(cfg, "Threshold", m_threshold);
        m_window = get(cfg, "Window", m_window);
        m_nbins = (int)get<double>(cfg, "Nbins", m_nbins);
        m_cut = get(cfg, "Cut", m_cut);
        if (m_window < 0) throw ConfigError("mbOneChannelStatus: Window below 0");
        if (m_nbins < 1) throw ConfigError("mbOneChannelStatus: Nbins below 1");
        m_anode_tn = get<std::string>(cfg, "anode", m_anode_tn);
        m_planes.build(*Factory::find_tn<IAnodeChannels>(m_anode_tn));
    }
    Waveform::ChannelMaskMap apply(int ch, signal_t& signal) const override
    {
        std::vector<float> rows(signal.begin(), signal.end());
        std::vector<Waveform::ChannelMaskMap> masks;
        apply_batch({ch}, rows, (int)signal.size(), masks);
        return masks[0];
    }
    Waveform::ChannelMaskMap apply(channel_signals_t&) const override { return Waveform::ChannelMaskMap(); }
    void apply_batch(const std::vector<int>& channels, std::vector<float>& rows, int nt,
                     std::vector<Waveform::ChannelMaskMap>& masks) const override
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        const int nch = (int)channels.size();
        masks.assign(nch, Waveform::ChannelMaskMap());
        std::vector<int> pick;   // rows off the collection plane (:1267)
        for (int i = 0; i < nch; ++i)
            if (m_planes(channels[i]) != 2) pick.push_back(i);
        if (pick.empty() || nt == 0) return;
        if (m_nbins >= nt)   // the reference's at(nbins) would throw
            throw std::runtime_error("mbOneChannelStatus: Nbins " + std::to_string(m_nbins) + " needs more than " +
                                     std::to_string(nt) + " ticks");
        const int np = (int)pick.size();
        std::vector<float> sel((size_t)np * nt);
        for (int k = 0; k < np; ++k)
            std::copy(rows.begin() + (size_t)pick[k] * nt, rows.begin() + (size_t)(pick[k] + 1) * nt,
                      sel.begin() + (size_t)k * nt);
        if (!m_stream) ncheck(wct_stream_create(&m_stream), "mbOneChannelStatus stream");
        const size_t bytes = sizeof(float) * sel.size();
        const size_t ws = wct_nf_status_workspace_size(np, nt);
        m_d_rows.ensure(bytes);
        m_d_flags.ensure(sizeof(int) * np);
        m_d_ws.ensure(ws);
        ncheck(wct_memcpy_htod_async(m_d_rows.p, sel.data(), bytes, m_stream), "mbOneChannelStatus upload");
        ncheck(wct_nf_lf_noisy((const float*)m_d_rows.p, nt, np, nt, m_threshold, m_window, m_nbins, m_cut,
                               (int*)m_d_flags.p, nullptr, m_d_ws.p, ws, m_stream),
               "mbOneChannelStatus");
        std::vector<int> flags(np);
        ncheck(wct_memcpy_dtoh_async(flags.data(), m_d_flags.p, sizeof(int) * np, m_stream),
               "mbOneChannelStatus download");
        ncheck(wct_stream_sync(m_stream), "mbOneChannelStatus sync");
        for (int k = 0; k < np; ++k)
            if (flags[k]) masks[pick[k]]["lf_noisy"][channels[pick[k]]].push_back({0, nt});
    }

  private:
    std::string m_anode_tn{"AnodeChannels"};
    double m_threshold{3.5}, m_cut{14.0};
    int m_window{5}, m_nbins{250};
    PlaneLookup m_planes;
    mutable std::mutex m_mutex;
    mutable void* m_stream{nullptr};
    mutable DevBuf m_d_rows, m_d_flags, m_d_ws;
};

// ---- mbCoherentNoiseSub (sigproc/src/Microboone.cxx:790-892) -----------
class CoherentNoiseSub : public IChannelFilter, public IBatchGroupFilter, public IConfigurable {
  public:
    typedef IChannelNoiseDatabase::filter_t filter_t;
    ~CoherentNoiseSub() override
    {
        if (m_stream) wct_stream_destroy(m_stream);
    }
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["anode"] = m_anode_tn;
        cfg["noisedb"] = m_noisedb_tn;
        cfg["dft"] = "HipDFT";
        cfg["rms_threshold"] = (double)m_rms_threshold;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_anode_tn = get<std::string>(cfg, "anode", m_anode_tn);
        Factory::find_tn<IAnodeChannels>(m_anode_tn);
        m_noisedb_tn = get<std::string>(cfg, "noisedb", m_noisedb_tn);
        m_noisedb = Factory::find_tn<IChannelNoiseDatabase>(m_noisedb_tn);
        m_rms_threshold = get(cfg, "rms_threshold", m_rms_threshold);
        if (m_rms_threshold != 0)
            throw ConfigError("mbCoherentNoiseSub: rms_threshold other than 0 (the per-channel limit chain of "
                              "Subtract_WScaling) is not built");
    }
    Waveform::ChannelMaskMap apply(int, signal_t&) const override { return Waveform::ChannelMaskMap(); }
    Waveform::ChannelMaskMap apply(channel_signals_t& chansig) const override
    {
        if (chansig.empty()) return Waveform::ChannelMaskMap();
        std::vector<int> channels;
        int nt = (int)chansig.begin()->second.size();
        std::vector<float> rows;
        std::vector<int> grp;
        for (auto& cs : chansig) {
            grp.push_back((int)channels.size());
            channels.push_back(cs.first);
            rows.insert(rows.end(), cs.second.begin(), cs.second.end());
        }
        apply_groups(channels, {grp}, rows, nt);
        size_t i = 0;
        for (auto& cs : chansig) {
            std::copy(rows.begin() + i * nt, rows.begin() + (i + 1) * nt, cs.second.begin());
            ++i;
        }
        return Waveform::ChannelMaskMap();   // :871
    }
    // Every group in one batch; a group's rows are taken in channel order
    // and its parameters from its first channel (:807-822).
    void apply_groups(const std::vector<int>& channels, const std::vector<std::vector<int>>& groups,
                      std::vector<float>& rows, int nt) const override
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        if (groups.empty() || nt == 0) return;
        const int nrows = (int)channels.size();
        std::vector<wct_nf_group> ent;
        std::vector<int> grows, row_group(nrows, -1);
        std::vector<const filter_t*> spec_rows;
        std::map<const filter_t*, int> spec_id;
        for (const auto& g0 : groups) {
            std::vector<int> g(g0);
            std::sort(g.begin(), g.end(), [&](int a, int b) { return channels[a] < channels[b]; });
            if (g.size() > 128) throw std::runtime_error("mbCoherentNoiseSub: groups of up to 128 channels");
            wct_nf_group e{};
            e.first = (int)grows.size();
            e.count = (int)g.size();
            const int ach = channels[g[0]];
            const filter_t& rs = m_noisedb->response(ach);
            e.res_offset = (int)m_noisedb->response_offset(ach);
            e.pad_f = m_noisedb->pad_window_front(ach);
            e.pad_b = m_noisedb->pad_window_back(ach);
            e.decon_limit = m_noisedb->coherent_nf_decon_limit(ach);
            e.decon_lf_cutoff = m_noisedb->coherent_nf_decon_lf_cutoff(ach);
            e.adc_limit = m_noisedb->coherent_nf_adc_limit(ach);
            e.decon_limit1 = m_noisedb->coherent_nf_decon_limit1(ach);
            e.roi_min_max_ratio = m_noisedb->coherent_nf_roi_min_max_ratio(ach);
            e.protection_factor = m_noisedb->coherent_nf_protection_factor(ach);
            e.min_adc_limit = m_noisedb->coherent_nf_min_adc_limit(ach);
            e.respec = -1;
            if (!rs.empty() && (rs[0].real() != 1 || rs[0].imag() != 0) && e.res_offset != 0) {
                if ((int)rs.size() != nt)
                    throw std::runtime_error("mbCoherentNoiseSub: response spectrum of " + std::to_string(rs.size()) +
                                             " bins for " + std::to_string(nt) + " ticks");
                auto it = spec_id.find(&rs);
                if (it == spec_id.end()) {
                    it = spec_id.emplace(&rs, (int)spec_rows.size()).first;
                    spec_rows.push_back(&rs);
                }
                e.respec = it->second;
            }
            for (int r : g) {
                if (row_group[r] >= 0) throw std::runtime_error("mbCoherentNoiseSub: a channel in two groups");
                row_group[r] = (int)ent.size();
                grows.push_back(r);
            }
            ent.push_back(e);
        }
        if (!m_stream) ncheck(wct_stream_create(&m_stream), "mbCoherentNoiseSub stream");
        const size_t row_bytes = sizeof(float) * (size_t)nrows * nt;
        m_d_rows.ensure(row_bytes);
        m_d_groups.ensure(sizeof(wct_nf_group) * ent.size());
        m_d_grows.ensure(sizeof(int) * grows.size());
        m_d_rowgroup.ensure(sizeof(int) * nrows);
        m_d_spectra.ensure(sizeof(float) * 2 * std::max<size_t>(1, spec_rows.size()) * nt);
        const size_t ws = wct_nf_coherent_workspace_size((int)ent.size(), nrows, nt);
        m_d_ws.ensure(ws);
        ncheck(wct_memcpy_htod_async(m_d_rows.p, rows.data(), row_bytes, m_stream), "mbCoherentNoiseSub upload");
        ncheck(wct_memcpy_htod_async(m_d_groups.p, ent.data(), sizeof(wct_nf_group) * ent.size(), m_stream),
               "mbCoherentNoiseSub upload");
        ncheck(wct_memcpy_htod_async(m_d_grows.p, grows.data(), sizeof(int) * grows.size(), m_stream),
               "mbCoherentNoiseSub upload");
        ncheck(wct_memcpy_htod_async(m_d_rowgroup.p, row_group.data(), sizeof(int) * nrows, m_stream),
               "mbCoherentNoiseSub upload");
        for (size_t r = 0; r < spec_rows.size(); ++r)
            ncheck(wct_memcpy_htod_async((char*)m_d_spectra.p + sizeof(float) * 2 * nt * r, spec_rows[r]->data(),
                                         sizeof(float) * 2 * nt, m_stream),
                   "mbCoherentNoiseSub upload");
        ncheck(wct_nf_coherent((float*)m_d_rows.p, nt, nrows, nt, (const wct_nf_group*)m_d_groups.p, (int)ent.size(),
                               (const int*)m_d_grows.p, (const int*)m_d_rowgroup.p, (const float*)m_d_spectra.p,
                               spec_rows.empty() ? 0 : 1, m_d_ws.p, ws, m_stream),
               "mbCoherentNoiseSub");
        ncheck(wct_memcpy_dtoh_async(rows.data(), m_d_rows.p, row_bytes, m_stream), "mbCoherentNoiseSub download");
        ncheck(wct_stream_sync(m_stream), "mbCoherentNoiseSub sync");
    }

  private:
    std::string m_anode_tn{"AnodeChannels"}, m_noisedb_tn{"OmniChannelNoiseDB"};
    std::shared_ptr<IChannelNoiseDatabase> m_noisedb;
    float m_rms_threshold{0.0f};
    mutable std::mutex m_mutex;
    mutable void* m_stream{nullptr};
    mutable DevBuf m_d_rows, m_d_groups, m_d_grows, m_d_rowgroup, m_d_spectra, m_d_ws;
};

// ---- OmnibusNoiseFilter (sigproc/src/OmnibusNoiseFilter.cxx) -----------
class OmnibusNoiseFilter : public IFrameFilter, public IConfigurable {
  public:
    Configuration default_configuration() const override   // :77-97
    {
        Configuration cfg;
        cfg["nticks"] = (int)m_nticks;
        cfg["maskmap"]["chirp"] = "bad";
        cfg["maskmap"]["noisy"] = "bad";
        cfg["channel_filters"][0] = "mbOneChannelNoise";
        cfg["channel_status_filters"][0] = "mbOneChannelStatus";
        cfg["grouped_filters"][0] = "mbCoherentNoiseSub";
        cfg["noisedb"] = "";
        cfg["intraces"] = m_intag;
        cfg["outtraces"] = m_outtag;
        return cfg;
    }
    void configure(const Configuration& cfg) override   // :37-75
    {
        m_nticks = (size_t)get(cfg, "nticks", (int)m_nticks);
        m_maskmap.clear();
        const auto& jmm = cfg["maskmap"];
        if (jmm.isObject())
            for (const auto& name : jmm.getMemberNames()) m_maskmap[name] = jmm[name].asString();
        m_perchan.clear();
        m_perchan_status.clear();
        m_grouped.clear();
        for (const auto& jf : cfg["channel_filters"]) m_perchan.push_back(Factory::find_tn<IChannelFilter>(jf.asString()));
        for (const auto& jf : cfg["channel_status_filters"])
            m_perchan_status.push_back(Factory::find_tn<IChannelFilter>(jf.asString()));
        for (const auto& jf : cfg["grouped_filters"]) m_grouped.push_back(Factory::find_tn<IChannelFilter>(jf.asString()));
        m_noisedb = Factory::find_tn<IChannelNoiseDatabase>(cfg["noisedb"].asString());
        m_intag = get<std::string>(cfg, "intraces", m_intag);
        m_outtag = get<std::string>(cfg, "outtraces", m_outtag);
    }

    bool operator()(const input_pointer& inframe, output_pointer& outframe) override   // :99-260
    {
        outframe = nullptr;
        if (!inframe) return true;
        const auto traces = tagged_traces(*inframe, m_intag);
        if (traces.empty()) {
            outframe = std::make_shared<Aux::SimpleFrame>(inframe->ident(), inframe->time(),
                                                          std::make_shared<ITrace::vector>(), inframe->tick());
            return true;
        }
        if (!m_nticks) m_nticks = traces.at(0)->charge().size();
        const int nt = (int)m_nticks;

        Waveform::ChannelMaskMap cmm;
        Waveform::merge(cmm, inframe->masks(), m_maskmap);
        const std::vector<int> bad = m_noisedb->bad_channels();
        {
            Waveform::ChannelMasks temp;
            for (int ch : bad) temp[ch].push_back(Waveform::BinRange(0, nt));
            Waveform::ChannelMaskMap tm;
            tm["bad"] = temp;
            Waveform::merge(cmm, tm, m_maskmap);
        }

        // working rows, one per channel (a later trace of the same channel
        // takes its row), good channels' samples from tick 0
        std::vector<int> channels;
        std::unordered_map<int, int> row_of;
        for (const auto& t : traces) {
            if (!row_of.count(t->channel())) {
                row_of[t->channel()] = (int)channels.size();
                channels.push_back(t->channel());
            }
        }
        std::vector<float> rows((size_t)channels.size() * nt, 0.f);
        for (const auto& t : traces) {
            float* row = rows.data() + (size_t)row_of[t->channel()] * nt;
            std::fill(row, row + nt, 0.f);
            if (std::binary_search(bad.begin(), bad.end(), t->channel())) continue;
            const auto& q = t->charge();
            std::copy(q.begin(), q.begin() + std::min((size_t)nt, q.size()), row);
        }

        for (const auto& filter : m_perchan) run_filter(*filter, channels, rows, nt, cmm);
        // grouped filters (:193-224): the complete groups (every channel
        // present), all at once when the filter takes batches
        std::vector<std::vector<int>> complete_groups;
        for (const auto& group : m_noisedb->coherent_channels()) {
            std::vector<int> g;
            bool ok = true;
            for (int ch : group) {
                auto it = row_of.find(ch);
                if (it == row_of.end()) {
                    ok = false;
                    break;
                }
                g.push_back(it->second);
            }
            if (ok) complete_groups.push_back(g);
        }
        std::vector<std::shared_ptr<IChannelFilter>> unbatched;
        for (const auto& filter : m_grouped) {
            if (auto* batch = dynamic_cast<const IBatchGroupFilter*>(filter.get()))
                batch->apply_groups(channels, complete_groups, rows, nt);
            else
                unbatched.push_back(filter);
        }
        for (const auto& group : unbatched.empty() ? std::vector<std::vector<int>>()
                                                   : m_noisedb->coherent_channels()) {
            bool complete = true;
            IChannelFilter::channel_signals_t chgrp;
            for (int ch : group) {
                auto it = row_of.find(ch);
                if (it == row_of.end()) {
                    complete = false;
                    continue;
                }
                const float* row = rows.data() + (size_t)it->second * nt;
                chgrp[ch].assign(row, row + nt);
            }
            if (!complete) continue;
            for (const auto& filter : unbatched) Waveform::merge(cmm, filter->apply(chgrp), m_maskmap);
            for (const auto& cs : chgrp) std::copy(cs.second.begin(), cs.second.end(), rows.data() + (size_t)row_of[cs.first] * nt);
        }
        for (const auto& filter : m_perchan_status) run_filter(*filter, channels, rows, nt, cmm);

        auto itraces = std::make_shared<ITrace::vector>();
        for (size_t i = 0; i < channels.size(); ++i)
            itraces->push_back(std::make_shared<Aux::SimpleTrace>(
                channels[i], 0, ITrace::ChargeSequence(rows.begin() + i * nt, rows.begin() + (i + 1) * nt)));
        auto sframe = std::make_shared<Aux::SimpleFrame>(inframe->ident(), inframe->time(), itraces, inframe->tick(),
                                                         cmm);
        IFrame::trace_list_t indices(itraces->size());
        for (size_t i = 0; i < indices.size(); ++i) indices[i] = i;
        if (!m_outtag.empty()) sframe->tag_traces(m_outtag, indices);
        sframe->tag_frame("noisefilter");
        outframe = sframe;
        return true;
    }

  private:
    // Aux::tagged_traces (aux/src/FrameTools.cxx:103-150).
    static ITrace::vector tagged_traces(const IFrame& frame, const std::string& tag)
    {
        ITrace::vector ret;
        const auto all = frame.traces();
        if (!all) return ret;
        if (tag.empty()) {   // untagged traces
            std::vector<size_t> tagged;
            for (const auto& t : frame.trace_tags()) {
                const auto& l = frame.tagged_traces(t);
                tagged.insert(tagged.end(), l.begin(), l.end());
            }
            std::sort(tagged.begin(), tagged.end());
            for (size_t i = 0; i < all->size(); ++i)
                if (!std::binary_search(tagged.begin(), tagged.end(), i)) ret.push_back(all->at(i));
            return ret;
        }
        for (size_t i : frame.tagged_traces(tag)) ret.push_back(all->at(i));
        if (!ret.empty()) return ret;
        const auto& ft = frame.frame_tags();
        if (std::find(ft.begin(), ft.end(), tag) == ft.end()) return ret;
        return *all;
    }

    // One per-channel filter over every row: batched when the filter
    // takes batches, else channel by channel; masks merged in channel
    // order (:176-187).
    void run_filter(const IChannelFilter& f, const std::vector<int>& channels, std::vector<float>& rows, int nt,
                    Waveform::ChannelMaskMap& cmm) const
    {
        if (auto* batch = dynamic_cast<const IBatchChannelFilter*>(&f)) {
            std::vector<Waveform::ChannelMaskMap> masks;
            batch->apply_batch(channels, rows, nt, masks);
            for (auto& m : masks) Waveform::merge(cmm, m, m_maskmap);
            return;
        }
        for (size_t i = 0; i < channels.size(); ++i) {
            IChannelFilter::signal_t sig(rows.begin() + i * nt, rows.begin() + (i + 1) * nt);
            Waveform::merge(cmm, f.apply(channels[i], sig), m_maskmap);
            std::copy(sig.begin(), sig.begin() + std::min((size_t)nt, sig.size()), rows.begin() + i * nt);
        }
    }

    size_t m_nticks{0};
    std::map<std::string, std::string> m_maskmap;
    std::vector<std::shared_ptr<IChannelFilter>> m_perchan, m_perchan_status, m_grouped;
    std::shared_ptr<IChannelNoiseDatabase> m_noisedb;
    std::string m_intag{"orig"}, m_outtag{"raw"};
};

}  // namespace Hip
}  // namespace WireCell

WIRECELL_HIP_FACTORY(OmniChannelNoiseDB, WireCell::Hip::OmniChannelNoiseDB)
WIRECELL_HIP_FACTORY(pdOneChannelNoise, WireCell::Hip::PdOneChannelNoise)
WIRECELL_HIP_FACTORY(mbOneChannelNoise, WireCell::Hip::MbOneChannelNoise)
WIRECELL_HIP_FACTORY(pdRelGainCalib, WireCell::Hip::RelGainCalib)
WIRECELL_HIP_FACTORY(mbCoherentNoiseSub, WireCell::Hip::CoherentNoiseSub)
WIRECELL_HIP_FACTORY(mbOneChannelStatus, WireCell::Hip::OneChannelStatus)
WIRECELL_HIP_FACTORY(mbADCBitShift, WireCell::Hip::ADCBitShift)
WIRECELL_HIP_FACTORY(OmnibusNoiseFilter, WireCell::Hip::OmnibusNoiseFilter)
