// The noise filter's component interfaces and components:
//   IChannelNoiseDatabase  iface/inc/WireCellIface/IChannelNoiseDatabase.h
//                          (the accessors the ProtoDUNE filter reads)
//   IChannelFilter         iface/inc/WireCellIface/IChannelFilter.h
//   OmniChannelNoiseDB     sigproc/src/OmniChannelNoiseDB.cxx
//   pdOneChannelNoise      sigproc/src/Protodune.cxx:712-923, every
//                          channel of a frame in one device batch
//   mbOneChannelNoise      sigproc/src/Microboone.cxx:894-1069, the same
//   pdRelGainCalib         sigproc/src/Protodune.cxx:925-1014
//   mbCoherentNoiseSub     sigproc/src/Microboone.cxx:790-892, every
//                          channel group of a frame in one device batch
//   mbOneChannelStatus     sigproc/src/Microboone.cxx:1218-1334, the
//                          same
//   mbADCBitShift          sigproc/src/Microboone.cxx:1071-1216, the same
//   OmnibusNoiseFilter     sigproc/src/OmnibusNoiseFilter.cxx
#pragma once
#include <complex>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Interface.h"

namespace WireCell {

class IChannelNoiseDatabase : public IComponent<IChannelNoiseDatabase> {
  public:
    typedef std::vector<std::complex<float>> filter_t;
    typedef std::vector<int> channel_group_t;
    virtual ~IChannelNoiseDatabase() = default;
    virtual double sample_time() const = 0;
    virtual double nominal_baseline(int channel) const = 0;
    virtual double gain_correction(int channel) const = 0;
    virtual double min_rms_cut(int channel) const = 0;
    virtual double max_rms_cut(int channel) const = 0;
    virtual double response_offset(int channel) const = 0;
    virtual int pad_window_front(int channel) const = 0;
    virtual int pad_window_back(int channel) const = 0;
    virtual float coherent_nf_decon_limit(int channel) const = 0;
    virtual float coherent_nf_decon_lf_cutoff(int channel) const = 0;
    virtual float coherent_nf_decon_limit1(int channel) const = 0;
    virtual float coherent_nf_adc_limit(int channel) const = 0;
    virtual float coherent_nf_protection_factor(int channel) const = 0;
    virtual float coherent_nf_min_adc_limit(int channel) const = 0;
    virtual float coherent_nf_roi_min_max_ratio(int channel) const = 0;
    virtual const filter_t& rcrc(int channel) const = 0;
    virtual const filter_t& config(int channel) const = 0;
    virtual const filter_t& noise(int channel) const = 0;
    virtual const filter_t& response(int channel) const = 0;
    virtual const std::vector<channel_group_t> coherent_channels() const = 0;
    virtual std::vector<int> bad_channels() const = 0;
};

class IChannelFilter : public IComponent<IChannelFilter> {
  public:
    typedef std::vector<float> signal_t;
    typedef std::map<int, signal_t> channel_signals_t;
    virtual ~IChannelFilter() = default;
    virtual Waveform::ChannelMaskMap apply(int channel, signal_t& sig) const = 0;
    virtual Waveform::ChannelMaskMap apply(channel_signals_t& chansig) const = 0;
};

namespace Waveform {
// Waveform::merge (util/src/Waveform.cxx:155-216): the union of bin-range
// lists (sorted; a range that starts inside the last one sets its end, as
// the reference does), per channel, and of mask maps with renaming.
BinRangeList merge(const BinRangeList& brl);
BinRangeList merge(const BinRangeList& a, const BinRangeList& b);
ChannelMasks merge(const ChannelMasks& a, const ChannelMasks& b);
void merge(ChannelMaskMap& one, const ChannelMaskMap& two, const std::map<std::string, std::string>& name_map);
}  // namespace Waveform

namespace Hip {

// A frame's channels through a batched per-channel filter: rows holds
// nch rows of nticks samples (host), filtered in place; masks receives one
// mask map per row.
class IBatchChannelFilter {
  public:
    virtual ~IBatchChannelFilter() = default;
    virtual void apply_batch(const std::vector<int>& channels, std::vector<float>& rows, int nticks,
                             std::vector<Waveform::ChannelMaskMap>& masks) const = 0;
};

// A frame's channel groups through a batched grouped filter: groups lists
// row indices of rows (nticks samples each), filtered in place.
class IBatchGroupFilter {
  public:
    virtual ~IBatchGroupFilter() = default;
    virtual void apply_groups(const std::vector<int>& channels, const std::vector<std::vector<int>>& groups,
                              std::vector<float>& rows, int nticks) const = 0;
};

}  // namespace Hip
}  // namespace WireCell
