// The subset of Wire-Cell's component interfaces that the deconvolution
// path crosses, restated with the reference's names, virtual signatures
// and argument meanings so host code written against the reference reads
// the same here:
//   IDFT            iface/inc/WireCellIface/IDFT.h:80-127
//   IConfigurable   iface/inc/WireCellIface/IConfigurable.h:13-21
//   ITrace          iface/inc/WireCellIface/ITrace.h:17-37
//   IFrame          iface/inc/WireCellIface/IFrame.h:20-71
//   IFrameFilter    iface/inc/WireCellIface/IFrameFilter.h:20-30 over
//                   IFunctionNode.h:26-55
//   IFilterWaveform iface/inc/WireCellIface/IFilterWaveform.h:12-18
//   IWaveform       iface/inc/WireCellIface/IWaveform.h:16-29
//   IFieldResponse  iface/inc/WireCellIface/IFieldResponse.h
//   IChannelResponse iface/inc/WireCellIface/IChannelResponse.h:17-27
//   IFrameSink, IFrameSource  iface/inc/WireCellIface/IFrameSink.h:11-21,
//                   IFrameSource.h:12-20
// IAnodeChannels stands in for the part of IAnodePlane that
// OmnibusSigProc reads (its channel -> (plane, wire) map,
// sigproc/src/OmnibusSigProc.cxx:171-216); wire geometry is not part of
// the deconvolution.
#pragma once
#include <complex>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "Configuration.h"
#include "Response.h"

namespace WireCell {

// Common base so the factory can hold any component and cast to the
// interface a caller asks for.
class Interface {
  public:
    virtual ~Interface() = default;
};

template <class Type>
class IComponent : virtual public Interface {
  public:
    typedef std::shared_ptr<Type> pointer;
    virtual ~IComponent() = default;
};

class IConfigurable : virtual public IComponent<IConfigurable> {
  public:
    virtual ~IConfigurable() = default;
    virtual Configuration default_configuration() const { return Configuration(); }
    virtual void configure(const Configuration& config) = 0;
};

// All methods const and thread safe; row-major; forward unnormalised,
// inverse scaled by 1/n of the transformed extent; in == out allowed.
class IDFT : public IComponent<IDFT> {
  public:
    virtual ~IDFT() = default;
    using scalar_t = float;
    using complex_t = std::complex<scalar_t>;

    virtual void fwd1d(const complex_t* in, complex_t* out, int size) const = 0;
    virtual void inv1d(const complex_t* in, complex_t* out, int size) const = 0;
    virtual void fwd1b(const complex_t* in, complex_t* out, int nrows, int ncols, int axis) const = 0;
    virtual void inv1b(const complex_t* in, complex_t* out, int nrows, int ncols, int axis) const = 0;
    virtual void fwd2d(const complex_t* in, complex_t* out, int nrows, int ncols) const = 0;
    virtual void inv2d(const complex_t* in, complex_t* out, int nrows, int ncols) const = 0;
    virtual void transpose(const scalar_t* in, scalar_t* out, int nrows, int ncols) const = 0;
    virtual void transpose(const complex_t* in, complex_t* out, int nrows, int ncols) const = 0;
};

class ITrace : public IComponent<ITrace> {
  public:
    typedef std::vector<float> ChargeSequence;
    typedef std::shared_ptr<const ITrace> pointer;
    typedef std::vector<pointer> vector;
    typedef std::shared_ptr<const vector> shared_vector;
    virtual ~ITrace() = default;
    virtual int channel() const = 0;
    virtual int tbin() const = 0;
    virtual const ChargeSequence& charge() const = 0;
};

namespace Waveform {
typedef std::vector<float> realseq_t;
// A bin range [first, second) (util/inc/WireCellUtil/Waveform.h).
typedef std::pair<int, int> BinRange;
typedef std::vector<BinRange> BinRangeList;
typedef std::map<int, BinRangeList> ChannelMasks;
typedef std::map<std::string, ChannelMasks> ChannelMaskMap;
}  // namespace Waveform

class IFrame : public IComponent<IFrame> {
  public:
    typedef std::shared_ptr<const IFrame> pointer;
    typedef std::string tag_t;
    typedef std::vector<tag_t> tag_list_t;
    typedef std::vector<size_t> trace_list_t;
    typedef std::vector<double> trace_summary_t;
    virtual ~IFrame() = default;
    virtual const tag_list_t& frame_tags() const = 0;
    virtual const tag_list_t& trace_tags() const = 0;
    virtual const trace_list_t& tagged_traces(const tag_t& tag) const = 0;
    virtual const trace_summary_t& trace_summary(const tag_t& tag) const = 0;
    virtual ITrace::shared_vector traces() const = 0;
    virtual Waveform::ChannelMaskMap masks() const { return Waveform::ChannelMaskMap(); }
    virtual int ident() const = 0;
    virtual double time() const = 0;
    virtual double tick() const = 0;
};

template <typename InputType, typename OutputType>
class IFunctionNode : virtual public Interface {
  public:
    typedef std::shared_ptr<const InputType> input_pointer;
    typedef std::shared_ptr<const OutputType> output_pointer;
    virtual ~IFunctionNode() = default;
    // in == nullptr is end-of-stream: out = nullptr, return true.
    virtual bool operator()(const input_pointer& in, output_pointer& out) = 0;
};

class IFrameFilter : public IFunctionNode<IFrame, IFrame>, public IComponent<IFrameFilter> {
  public:
    typedef std::shared_ptr<IFrameFilter> pointer;
    virtual ~IFrameFilter() = default;
};

// Sink and source nodes of frames (iface/inc/WireCellIface/IFrameSink.h,
// IFrameSource.h over ISinkNode / ISourceNode): a sink consumes a frame
// (nullptr is end-of-stream); a source produces the next frame, nullptr
// once at end-of-stream, and returns false after that.
class IFrameSink : public IComponent<IFrameSink> {
  public:
    virtual ~IFrameSink() = default;
    virtual bool operator()(const IFrame::pointer& frame) = 0;
};

class IFrameSource : public IComponent<IFrameSource> {
  public:
    virtual ~IFrameSource() = default;
    virtual bool operator()(IFrame::pointer& frame) = 0;
};

class IFilterWaveform : public IComponent<IFilterWaveform> {
  public:
    virtual ~IFilterWaveform() = default;
    virtual const Waveform::realseq_t filter_waveform(int nfbins) const = 0;
};

// util/inc/WireCellUtil/Binning.h: nbins uniform bins over [minval, maxval).
struct Binning {
    int nbins{0};
    double minval{0}, maxval{0};
    Binning(int n, double lo, double hi) : nbins(n), minval(lo), maxval(hi) {}
    double binsize() const { return (maxval - minval) / nbins; }
    double center(int i) const { return minval + (i + 0.5) * binsize(); }
};

class IWaveform : public IComponent<IWaveform> {
  public:
    typedef std::vector<float> sequence_type;
    virtual ~IWaveform() = default;
    virtual double waveform_start() const = 0;
    virtual double waveform_period() const = 0;
    virtual const sequence_type& waveform_samples() const = 0;
    virtual sequence_type waveform_samples(const Binning& tbins) const = 0;
};

// A per-channel (electronics) response: samples on the binning that
// channel_response_binning() returns.
class IChannelResponse : public IComponent<IChannelResponse> {
  public:
    typedef std::vector<float> realseq_t;
    virtual ~IChannelResponse() = default;
    virtual const realseq_t& channel_response(int channel_ident) const = 0;
    virtual Binning channel_response_binning() const = 0;
};

class IFieldResponse : public IComponent<IFieldResponse> {
  public:
    virtual ~IFieldResponse() = default;
    virtual const Hip::FieldResponse& field_response() const = 0;
};

class IAnodeChannels : public IComponent<IAnodeChannels> {
  public:
    virtual ~IAnodeChannels() = default;
    virtual int ident() const = 0;
    // WCT channel idents of plane p (0, 1, 2) in wire order.
    virtual const std::vector<int>& plane_channels(int plane) const = 0;
};

}  // namespace WireCell
