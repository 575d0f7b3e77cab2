// The small configurable components the deconvolution looks up by name:
// the time/wire filters, the electronics response, the field response
// file and the anode's channel planes.  Each keeps the reference type
// name, configuration keys and defaults.
#pragma once
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "Interface.h"
#include "Response.h"

namespace WireCell {
namespace Hip {

// SigProc::HfFilter (sigproc/src/HfFilter.cxx:13-52; defaults
// sigproc/inc/WireCellSigProc/HfFilter.h:13-14).
class HfFilter : public IFilterWaveform, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["sigma"] = m_sigma;
        cfg["power"] = m_power;
        cfg["flag"] = m_flag;
        cfg["max_freq"] = m_max_freq;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_sigma = get(cfg, "sigma", m_sigma);
        m_power = get(cfg, "power", m_power);
        m_flag = get(cfg, "flag", m_flag);
        m_max_freq = get(cfg, "max_freq", m_max_freq);
    }
    const Waveform::realseq_t filter_waveform(int nbins) const override
    {
        auto d = hf_filter_waveform(nbins, m_max_freq, m_sigma, m_power, m_flag);
        return Waveform::realseq_t(d.begin(), d.end());
    }
    // The FP64 values before the float cast of realseq_t; the device
    // builds the response inverse from these.
    std::vector<double> filter_waveform_f64(int nbins) const
    {
        return hf_filter_waveform(nbins, m_max_freq, m_sigma, m_power, m_flag);
    }

  private:
    double m_max_freq{1 * units::megahertz};
    double m_sigma{3.0 * units::megahertz};
    double m_power{2};
    bool m_flag{true};
};

// SigProc::LfFilter (sigproc/src/LfFilter.cxx:9-45; defaults
// sigproc/inc/WireCellSigProc/LfFilter.h:13).
class LfFilter : public IFilterWaveform, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["tau"] = m_tau;
        cfg["max_freq"] = m_max_freq;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_tau = get(cfg, "tau", m_tau);
        m_max_freq = get(cfg, "max_freq", m_max_freq);
    }
    const Waveform::realseq_t filter_waveform(int nbins) const override
    {
        auto d = lf_filter_waveform(nbins, m_max_freq, m_tau);
        return Waveform::realseq_t(d.begin(), d.end());
    }
    std::vector<double> filter_waveform_f64(int nbins) const
    {
        return lf_filter_waveform(nbins, m_max_freq, m_tau);
    }

  private:
    double m_max_freq{1 * units::megahertz};
    double m_tau{0.02 * units::megahertz};
};

// Gen::ColdElecResponse (gen/src/ColdElecResponse.cxx:11-49; defaults
// gen/inc/WireCellGen/ColdElecResponse.h:18-19).
class ColdElecResponse : public IWaveform, public IConfigurable {
  public:
    ColdElecResponse() { m_cfg = default_configuration(); }
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["gain"] = 14.0 * units::mV / units::fC;
        cfg["shaping"] = 2.0 * units::microsecond;
        cfg["postgain"] = 1.0;
        cfg["start"] = 0.0;
        cfg["tick"] = 0.5 * units::microsecond;
        cfg["nticks"] = 10000;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_cfg = update(default_configuration(), cfg);
        const int n = m_cfg["nticks"].asInt();
        const double t0 = m_cfg["start"].asDouble(), tick = m_cfg["tick"].asDouble();
        auto d = coldelec_samples(n, t0, t0 + n * tick, m_cfg["gain"].asDouble(), m_cfg["shaping"].asDouble(),
                                  m_cfg["postgain"].asDouble());
        m_wave.assign(d.begin(), d.end());
    }
    double waveform_start() const override { return m_cfg["start"].asDouble(); }
    double waveform_period() const override { return m_cfg["tick"].asDouble(); }
    const sequence_type& waveform_samples() const override { return m_wave; }
    sequence_type waveform_samples(const Binning& tb) const override
    {
        auto d = samples_f64(tb);
        return sequence_type(d.begin(), d.end());
    }
    std::vector<double> samples_f64(const Binning& tb) const
    {
        return coldelec_samples(tb.nbins, tb.minval, tb.maxval, m_cfg["gain"].asDouble(),
                                m_cfg["shaping"].asDouble(), m_cfg["postgain"].asDouble());
    }

  private:
    Configuration m_cfg;
    sequence_type m_wave;
};

// SigProc::FieldResponse (sigproc/src/FieldResponse.cxx:14-40) reading the
// reference's JSON schema (util/src/Response.cxx:32-72).  Plain .json
// only (no bzip2 in this build); relative names are searched along
// WIRECELL_PATH as the reference's Persist::resolve does.
class FieldResponseFile : public IFieldResponse, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["filename"] = m_fname;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_fname = get(cfg, "filename", m_fname);
        if (m_fname.empty()) throw ConfigError("FieldResponse: no filename given");
        m_fr = load(resolve(m_fname));
    }
    const Hip::FieldResponse& field_response() const override { return m_fr; }

    static std::string resolve(const std::string& fname)
    {
        if (fname.empty() || fname[0] == '/') return fname;
        if (std::ifstream(fname).good()) return fname;
        const char* wp = std::getenv("WIRECELL_PATH");
        if (wp) {
            std::stringstream ss(wp);
            std::string dir;
            while (std::getline(ss, dir, ':')) {
                const std::string full = dir + "/" + fname;
                if (std::ifstream(full).good()) return full;
            }
        }
        return fname;
    }
    static Hip::FieldResponse load(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) throw ConfigError("FieldResponse: cannot open " + path);
        std::stringstream buf;
        buf << in.rdbuf();
        const Configuration top = Configuration::parse(buf.str());
        const Configuration& fr = top["FieldResponse"];
        if (!fr.isObject()) throw ConfigError("FieldResponse: no FieldResponse object in " + path);
        Hip::FieldResponse out;
        for (const auto& plane : fr["planes"]) {
            const Configuration& plr = plane["PlaneResponse"];
            PlaneResponse pr;
            for (const auto& path_ : plr["paths"]) {
                const Configuration& par = path_["PathResponse"];
                PathResponse p;
                for (const auto& c : par["current"]["array"]["elements"]) p.current.push_back(c.asDouble());
                p.pitchpos = par["pitchpos"].asDouble();
                pr.paths.push_back(std::move(p));
            }
            pr.planeid = plr["planeid"].asInt();
            pr.location = plr["location"].asDouble();
            pr.pitch = plr["pitch"].asDouble();
            out.planes.push_back(std::move(pr));
        }
        out.origin = fr["origin"].asDouble();
        out.tstart = fr["tstart"].asDouble();
        out.period = fr["period"].asDouble();
        out.speed = fr["speed"].asDouble();
        return out;
    }

  private:
    std::string m_fname;
    Hip::FieldResponse m_fr;
};

struct KeyError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

// SigProc::PerChannelResponse (sigproc/src/PerChannelResponse.cxx:28-78):
// a JSON file {"tick": T, "t0": T0, "channels": [[ch, [samples]], [[ch,
// ch, ...], [samples]], ...]}.  tick, t0 and the samples are read as
// floats, and the binning is that of the first entry: nsamp bins from t0
// at tick.  Plain .json only (no bzip2 in this build), resolved along
// WIRECELL_PATH.
class PerChannelResponse : public IChannelResponse, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["filename"] = m_fname;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_fname = get(cfg, "filename", m_fname);
        if (m_fname.empty()) throw ConfigError("must supply a PerChannelResponse filename");
        const std::string path = FieldResponseFile::resolve(m_fname);
        std::ifstream in(path);
        if (!in) throw ConfigError("PerChannelResponse: cannot open " + path);
        std::stringstream buf;
        buf << in.rdbuf();
        const Configuration top = Configuration::parse(buf.str());
        const double tick = (float)top["tick"].asDouble();
        const double t0 = (float)top["t0"].asDouble();
        const Configuration& chans = top["channels"];
        if (chans.isNull()) throw ConfigError("no channels given in file " + m_fname);
        m_cr.clear();
        m_bins = Binning(0, 0, 0);
        for (const auto& entry : chans) {
            const Configuration& samples = entry[1];
            const int n = (int)samples.size();
            if (n == 0) throw ConfigError("zero length response in file " + m_fname);
            realseq_t resp(n);
            for (int i = 0; i < n; ++i) resp[i] = (float)samples[i].asDouble();
            if (entry[0].isArray()) {
                for (const auto& ch : entry[0]) m_cr[ch.asInt()] = resp;
            } else {
                m_cr[entry[0].asInt()] = resp;
            }
            if (!m_bins.nbins) m_bins = Binning(n, t0, t0 + n * tick);
        }
    }
    const realseq_t& channel_response(int channel_ident) const override
    {
        auto it = m_cr.find(channel_ident);
        if (it == m_cr.end()) throw KeyError("no response for channel " + std::to_string(channel_ident));
        return it->second;
    }
    Binning channel_response_binning() const override { return m_bins; }

  private:
    std::string m_fname;
    std::map<int, realseq_t> m_cr;
    Binning m_bins{0, 0, 0};
};

// The channel planes of one anode.  Configuration:
//   {"ident": N, "planes": [[ch, ...], [ch, ...], [ch, ...]]}
// or, for consecutive channel numbers,
//   {"ident": N, "plane_ranges": [[first, count], [first, count], [first, count]]}
class AnodeChannels : public IAnodeChannels, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["ident"] = 0;
        return cfg;
    }
    void configure(const Configuration& cfg) override
    {
        m_ident = get(cfg, "ident", 0);
        for (auto& p : m_planes) p.clear();
        if (cfg["planes"].isArray()) {
            int ip = 0;
            for (const auto& plane : cfg["planes"]) {
                if (ip >= 3) throw ConfigError("AnodeChannels: more than three planes");
                for (const auto& ch : plane) m_planes[ip].push_back(ch.asInt());
                ++ip;
            }
        } else if (cfg["plane_ranges"].isArray()) {
            int ip = 0;
            for (const auto& r : cfg["plane_ranges"]) {
                if (ip >= 3) throw ConfigError("AnodeChannels: more than three planes");
                const int first = r[0].asInt(), count = r[1].asInt();
                for (int i = 0; i < count; ++i) m_planes[ip].push_back(first + i);
                ++ip;
            }
        } else {
            throw ConfigError("AnodeChannels: give \"planes\" or \"plane_ranges\"");
        }
    }
    int ident() const override { return m_ident; }
    const std::vector<int>& plane_channels(int plane) const override
    {
        static const std::vector<int> empty;
        return (plane >= 0 && plane < 3) ? m_planes[plane] : empty;
    }

  private:
    int m_ident{0};
    std::vector<int> m_planes[3];
};

}  // namespace Hip
}  // namespace WireCell
