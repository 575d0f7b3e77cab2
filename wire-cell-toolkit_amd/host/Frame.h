// Concrete trace and frame: Aux::SimpleTrace / Aux::SimpleFrame restated
// (aux/inc/WireCellAux/SimpleTrace.h, aux/inc/WireCellAux/SimpleFrame.h:20-70).
#pragma once
#include <algorithm>

#include "Interface.h"

namespace WireCell {
namespace Aux {

class SimpleTrace : public ITrace {
  public:
    SimpleTrace(int ident, int tbin, const ChargeSequence& charge)
        : m_ident(ident), m_tbin(tbin), m_charge(charge) {}
    SimpleTrace(int ident, int tbin, ChargeSequence&& charge)
        : m_ident(ident), m_tbin(tbin), m_charge(std::move(charge)) {}
    int channel() const override { return m_ident; }
    int tbin() const override { return m_tbin; }
    const ChargeSequence& charge() const override { return m_charge; }

  private:
    int m_ident, m_tbin;
    ChargeSequence m_charge;
};

class SimpleFrame : public IFrame {
  public:
    SimpleFrame(int ident, double time, ITrace::shared_vector traces, double tick = 0.5 * 1000.0,
                const Waveform::ChannelMaskMap& masks = Waveform::ChannelMaskMap())
        : m_ident(ident), m_time(time), m_tick(tick), m_traces(std::move(traces)), m_masks(masks) {}

    const tag_list_t& frame_tags() const override { return m_frame_tags; }
    const tag_list_t& trace_tags() const override { return m_trace_tags; }
    const trace_list_t& tagged_traces(const tag_t& tag) const override
    {
        static const trace_list_t empty;
        auto it = m_tagged.find(tag);
        return it == m_tagged.end() ? empty : it->second;
    }
    const trace_summary_t& trace_summary(const tag_t& tag) const override
    {
        static const trace_summary_t empty;
        auto it = m_summary.find(tag);
        return it == m_summary.end() ? empty : it->second;
    }
    ITrace::shared_vector traces() const override { return m_traces; }
    Waveform::ChannelMaskMap masks() const override { return m_masks; }
    int ident() const override { return m_ident; }
    double time() const override { return m_time; }
    double tick() const override { return m_tick; }

    void tag_frame(const tag_t& tag)
    {
        if (std::find(m_frame_tags.begin(), m_frame_tags.end(), tag) == m_frame_tags.end())
            m_frame_tags.push_back(tag);
    }
    void tag_traces(const tag_t& tag, const trace_list_t& indices,
                    const trace_summary_t& summary = trace_summary_t())
    {
        if (std::find(m_trace_tags.begin(), m_trace_tags.end(), tag) == m_trace_tags.end())
            m_trace_tags.push_back(tag);
        m_tagged[tag] = indices;
        if (!summary.empty()) m_summary[tag] = summary;
    }
    void set_masks(const Waveform::ChannelMaskMap& masks) { m_masks = masks; }

  private:
    int m_ident;
    double m_time, m_tick;
    ITrace::shared_vector m_traces;
    Waveform::ChannelMaskMap m_masks;
    tag_list_t m_frame_tags, m_trace_tags;
    std::map<tag_t, trace_list_t> m_tagged;
    std::map<tag_t, trace_summary_t> m_summary;
};

}  // namespace Aux
}  // namespace WireCell
