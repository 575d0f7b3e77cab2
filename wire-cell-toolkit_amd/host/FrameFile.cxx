// Frame files: FrameFileSink and FrameFileSource restated
// (sio/src/FrameFileSink.cxx:22-279, sio/src/FrameFileSource.cxx:25-398)
// over the tar / .npy containers of Archive.h.  Per frame and trace tag
// the archive holds, in this order,
//   frame_<tag>_<ident>.npy     {nchannels, nticks} float32 (int16 when
//                               digitized), column by column as the
//                               reference's Eigen arrays are written
//   channels_<tag>_<ident>.npy  {nchannels} int32
//   tickinfo_<tag>_<ident>.npy  {3} float64: time, tick, first tbin
//   summary_<tag>_<ident>.npy   {nchannels} float64, when the tag has one
// then chanmask_<name>_<ident>.npy {n, 3} int32 (channel, first, end) per
// channel-mask-map entry.  Tag "*" names every trace, tag "" the
// untagged ones.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "Archive.h"
#include "Factory.h"
#include "Frame.h"
#include "Interface.h"

namespace WireCell {
namespace Hip {

namespace {

std::vector<std::string> split(const std::string& s, char d)
{
    std::vector<std::string> out;
    size_t b = 0;
    while (true) {
        const size_t e = s.find(d, b);
        out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
        if (e == std::string::npos) break;
        b = e + 1;
    }
    return out;
}

template <typename T>
NpyArray make_npy(const char* descr, std::vector<size_t> shape, const T* data, size_t n, bool fortran = false)
{
    NpyArray a;
    a.descr = descr;
    a.fortran_order = fortran;
    a.shape = std::move(shape);
    a.data.resize(n * sizeof(T));
    if (n) std::memcpy(a.data.data(), data, n * sizeof(T));
    return a;
}

// Aux::untagged_traces / Aux::tagged_traces (aux/src/FrameTools.cxx:
// 103-149): the traces of the tag ("" = not in any tag), or every trace
// when the tag is a frame tag and names no trace.
ITrace::vector tagged_traces(const IFrame::pointer& frame, const std::string& tag)
{
    ITrace::vector ret;
    const auto all = frame->traces();
    if (tag.empty()) {
        std::vector<char> tagged(all->size(), 0);
        for (const auto& t : frame->trace_tags())
            for (size_t i : frame->tagged_traces(t))
                if (i < tagged.size()) tagged[i] = 1;
        for (size_t i = 0; i < all->size(); ++i)
            if (!tagged[i]) ret.push_back(all->at(i));
        return ret;
    }
    for (size_t i : frame->tagged_traces(tag)) ret.push_back(all->at(i));
    if (!ret.empty()) return ret;
    const auto& ft = frame->frame_tags();
    if (std::find(ft.begin(), ft.end(), tag) == ft.end()) return ret;
    return *all;
}

}  // namespace

class FrameFileSink : public IFrameSink, public IConfigurable {
  public:
    ~FrameFileSink() override = default;

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["outname"] = m_outname;
        cfg["tags"] = Configuration::makeArray();
        cfg["digitize"] = false;
        cfg["baseline"] = 0.0;
        cfg["scale"] = 1.0;
        cfg["offset"] = 0.0;
        cfg["masks"] = true;
        return cfg;
    }

    // FrameFileSink.cxx:72-121
    void configure(const Configuration& cfg) override
    {
        m_outname = get<std::string>(cfg, "outname", m_outname);
        m_masks = get<bool>(cfg, "masks", m_masks);
        if (!TarWriter::supported(m_outname)) throw IOError("FrameFileSink: unsupported outname: " + m_outname);
        m_out.reset(new TarWriter(m_outname));
        m_tags.clear();
        const Configuration tags = cfg["tags"];
        for (size_t i = 0; tags.isArray() && i < tags.size(); ++i) m_tags.push_back(tags[(int)i].asString());
        if (m_tags.empty()) m_tags.push_back("*");
        m_digitize = get<bool>(cfg, "digitize", m_digitize);
        m_baseline = get<double>(cfg, "baseline", m_baseline);
        m_scale = get<double>(cfg, "scale", m_scale);
        m_offset = get<double>(cfg, "offset", m_offset);
        m_dense = false;
        const Configuration dense = cfg["dense"];
        if (dense.isObject() && dense.size()) {
            m_dense = true;
            m_chbeg = get<int>(dense, "chbeg", 0);
            m_chend = get<int>(dense, "chend", 0);
            m_tbbeg = get<int>(dense, "tbbeg", 0);
            m_tbend = get<int>(dense, "tbend", 0);
        }
        m_count = 0;
    }

    bool operator()(const IFrame::pointer& frame) override
    {
        if (!frame) {   // end of stream: the archive is complete
            ++m_count;
            if (m_out) m_out->close();
            return true;
        }
        if (!m_out) throw IOError("FrameFileSink: not configured");
        for (const auto& tag : m_tags) one_tag(frame, tag);
        if (m_masks) masks(frame);
        m_out->flush();
        ++m_count;
        return true;
    }

  private:
    // FrameFileSink.cxx:123-206
    void one_tag(const IFrame::pointer& frame, const std::string& tag)
    {
        ITrace::vector traces;
        IFrame::trace_summary_t summary;
        if (tag == "*") {
            const auto sv = frame->traces();
            traces.assign(sv->begin(), sv->end());
        } else {
            traces = tagged_traces(frame, tag);
            summary = frame->trace_summary(tag);
        }
        if (traces.empty()) return;

        std::vector<int> channels;
        int tb0 = 0, tb1 = 0;
        if (m_dense) {
            for (int c = m_chbeg; c < m_chend; ++c) channels.push_back(c);
            tb0 = m_tbbeg;
            tb1 = m_tbend;
        } else {
            for (const auto& t : traces) channels.push_back(t->channel());
            std::sort(channels.begin(), channels.end());
            channels.erase(std::unique(channels.begin(), channels.end()), channels.end());
            tb0 = traces[0]->tbin();
            tb1 = traces[0]->tbin() + (int)traces[0]->charge().size();
            for (const auto& t : traces) {
                tb0 = std::min(tb0, t->tbin());
                tb1 = std::max(tb1, t->tbin() + (int)t->charge().size());
            }
        }
        const size_t nrows = channels.size(), ncols = (size_t)std::max(0, tb1 - tb0);
        // Aux::fill (aux/src/FrameTools.cxx:181-223): charges summed into
        // rows by channel, clipped to [tb0, tb1); stored column by column
        // the double-valued keys enter the float array arithmetic as
        // floats, as Eigen promotes a scalar operand to the array's type
        const float fbase = (float)m_baseline, fscale = (float)m_scale, foff = (float)m_offset;
        std::vector<float> arr(nrows * ncols, fbase);
        std::unordered_map<int, size_t> row;
        for (size_t i = 0; i < nrows; ++i) row[channels[i]] = i;
        for (const auto& t : traces) {
            auto it = row.find(t->channel());
            if (it == row.end()) continue;
            const auto& q = t->charge();
            const int d = t->tbin() - tb0;
            const int c0 = std::max(0, d), k0 = std::max(0, -d);
            const int n = std::min((int)ncols - c0, (int)q.size() - k0);
            for (int k = 0; k < n; ++k) arr[it->second + (size_t)(c0 + k) * nrows] += q[k0 + k];
        }
        for (float& v : arr) {
            const float p = v * fscale;
            v = p + foff;
        }
        const std::string id = "_" + tag + "_" + std::to_string(frame->ident()) + ".npy";
        if (m_digitize) {
            std::vector<int16_t> s(arr.size());
            for (size_t i = 0; i < arr.size(); ++i) s[i] = (int16_t)arr[i];   // Eigen cast<short>: truncation
            m_out->add("frame" + id, npy_encode(make_npy("<i2", {nrows, ncols}, s.data(), s.size(), true)));
        } else {
            m_out->add("frame" + id, npy_encode(make_npy("<f4", {nrows, ncols}, arr.data(), arr.size(), true)));
        }
        m_out->add("channels" + id, npy_encode(make_npy("<i4", {nrows}, channels.data(), nrows)));
        const double tickinfo[3] = {frame->time(), frame->tick(), (double)tb0};
        m_out->add("tickinfo" + id, npy_encode(make_npy("<f8", {3}, tickinfo, 3)));
        if (!summary.empty())
            m_out->add("summary" + id, npy_encode(make_npy("<f8", {summary.size()}, summary.data(), summary.size())));
    }

    // FrameFileSink.cxx:218-258: each channel-mask-map entry as {n, 3}
    void masks(const IFrame::pointer& frame)
    {
        for (const auto& [name, cms] : frame->masks()) {
            std::vector<int> ch, lo, hi;
            for (const auto& [c, brl] : cms)
                for (const auto& br : brl) {
                    ch.push_back(c);
                    lo.push_back(br.first);
                    hi.push_back(br.second);
                }
            const size_t n = ch.size();
            std::vector<int> arr;   // column by column: channels, firsts, ends
            arr.insert(arr.end(), ch.begin(), ch.end());
            arr.insert(arr.end(), lo.begin(), lo.end());
            arr.insert(arr.end(), hi.begin(), hi.end());
            m_out->add("chanmask_" + name + "_" + std::to_string(frame->ident()) + ".npy",
                       npy_encode(make_npy("<i4", {n, 3}, arr.data(), arr.size(), true)));
        }
    }

    std::string m_outname{"wct-frame.tar.gz"};
    std::vector<std::string> m_tags;
    bool m_digitize{false}, m_masks{true}, m_dense{false};
    double m_baseline{0}, m_scale{1}, m_offset{0};
    int m_chbeg{0}, m_chend{0}, m_tbbeg{0}, m_tbend{0};
    size_t m_count{0};
    std::unique_ptr<TarWriter> m_out;
};

class FrameFileSource : public IFrameSource, public IConfigurable {
  public:
    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["inname"] = m_inname;
        cfg["tags"] = Configuration::makeArray();
        cfg["frame_tags"] = Configuration::makeArray();
        return cfg;
    }

    // FrameFileSource.cxx:48-66
    void configure(const Configuration& cfg) override
    {
        m_inname = get<std::string>(cfg, "inname", m_inname);
        if (!TarReader::supported(m_inname)) throw IOError("FrameFileSource: unsupported inname: " + m_inname);
        m_in.reset(new TarReader(m_inname));
        m_tags.clear();
        m_frame_tags.clear();
        const Configuration tags = cfg["tags"], ftags = cfg["frame_tags"];
        for (size_t i = 0; tags.isArray() && i < tags.size(); ++i) m_tags.push_back(tags[(int)i].asString());
        for (size_t i = 0; ftags.isArray() && i < ftags.size(); ++i) m_frame_tags.push_back(ftags[(int)i].asString());
        m_have = false;
        m_eos_sent = false;
    }

    // FrameFileSource.cxx:375-397
    bool operator()(IFrame::pointer& frame) override
    {
        frame = nullptr;
        if (m_eos_sent) return false;
        if (!m_in) throw IOError("FrameFileSource: not configured");
        frame = load();
        if (!frame) m_eos_sent = true;
        return true;
    }

  private:
    struct Entry {
        std::string name, type, tag;
        int ident{0};
        NpyArray arr;
    };
    struct Framelet {
        std::string tag;
        std::vector<double> tickinfo, summary;
        std::vector<int> channels;
        NpyArray frame;
        bool has_frame{false};
    };

    static bool untagged(const std::string& t) { return t.empty() || t == "*"; }
    bool excluded(const std::string& t) const
    {
        if (untagged(t)) return false;
        return std::find(m_tags.begin(), m_tags.end(), t) == m_tags.end();
    }

    // The next member, its name parsed as <type>_<tag>_<ident>.<ext>
    // (FrameFileSource.cxx:96-123); false at the end of the archive.
    bool read_entry()
    {
        std::vector<unsigned char> bytes;
        std::string name;
        if (!m_in->next(name, bytes)) return false;
        const auto parts = split(name, '_');
        if (parts.size() != 3) throw IOError("FrameFileSource: cannot parse member name " + name + " in " + m_inname);
        const auto rparts = split(parts[2], '.');
        m_cur.name = name;
        m_cur.type = parts[0];
        m_cur.tag = parts[1];
        m_cur.ident = std::atoi(rparts[0].c_str());
        m_cur.arr = npy_decode(bytes);
        m_have = true;
        return true;
    }

    static std::vector<double> to_f64(const NpyArray& a)
    {
        std::vector<double> v(a.count());
        for (size_t i = 0; i < v.size(); ++i) v[i] = a.at(i);
        return v;
    }
    static std::vector<int> to_i32(const NpyArray& a)
    {
        if (a.descr[1] != 'i' && a.descr[1] != 'u') throw IOError("FrameFileSource: integer array expected, got " + a.descr);
        std::vector<int> v(a.count());
        for (size_t i = 0; i < v.size(); ++i) v[i] = (int)a.at(i);
        return v;
    }

    // FrameFileSource.cxx:125-366: the members of one ident
    IFrame::pointer load()
    {
        int ident = -1;
        std::vector<Framelet> framelets;
        std::map<std::string, size_t> index;
        auto framelet = [&](const std::string& tag) -> Framelet& {
            auto it = index.find(tag);
            if (it != index.end()) return framelets[it->second];
            index[tag] = framelets.size();
            framelets.emplace_back();
            framelets.back().tag = tag;
            return framelets.back();
        };
        Waveform::ChannelMaskMap cmm;
        while (true) {
            if (!m_have && !read_entry()) break;
            if (ident < 0) ident = m_cur.ident;
            if (ident != m_cur.ident) break;   // the next frame's member stays for the next call
            Entry e = std::move(m_cur);
            m_have = false;
            if (e.type == "chanmask") {
                if (e.arr.shape.size() != 2 || e.arr.shape[1] != 3)
                    throw IOError("FrameFileSource: chanmask " + e.name + " is not {n, 3}");
                Waveform::ChannelMasks cms;
                for (size_t r = 0; r < e.arr.shape[0]; ++r)
                    cms[(int)e.arr.at(r, 0)].push_back({(int)e.arr.at(r, 1), (int)e.arr.at(r, 2)});
                cmm[e.tag] = cms;
                continue;
            }
            if (excluded(e.tag)) continue;
            if (e.type == "frame") {
                if (e.arr.shape.size() != 2) throw IOError("FrameFileSource: frame " + e.name + " is not 2D");
                if (e.arr.descr != "<f4" && e.arr.descr != "<i2")
                    throw IOError("FrameFileSource: frame " + e.name + " has dtype " + e.arr.descr + " (float32 or int16)");
                auto& f = framelet(e.tag);
                f.frame = std::move(e.arr);
                f.has_frame = true;
            } else if (e.type == "channels") {
                framelet(e.tag).channels = to_i32(e.arr);
            } else if (e.type == "tickinfo") {
                framelet(e.tag).tickinfo = to_f64(e.arr);
            } else if (e.type == "summary") {
                framelet(e.tag).summary = to_f64(e.arr);
            }   // other members are skipped, as the reference does
        }
        if (framelets.empty()) return nullptr;

        ITrace::vector traces;
        for (auto& f : framelets) {
            const size_t nrows = f.has_frame ? f.frame.shape[0] : 0, ncols = f.has_frame ? f.frame.shape[1] : 0;
            if (nrows != f.channels.size())
                throw IOError("FrameFileSource: frame and channel arrays differ in size for ident " +
                              std::to_string(ident) + " in " + m_inname);
            if (f.tickinfo.size() < 3)
                throw IOError("FrameFileSource: missing tickinfo for ident " + std::to_string(ident));
            const int tbin0 = (int)f.tickinfo[2];
            for (size_t r = 0; r < nrows; ++r) {
                ITrace::ChargeSequence q(ncols);
                for (size_t c = 0; c < ncols; ++c) q[c] = (float)f.frame.at(r, c);
                traces.push_back(std::make_shared<Aux::SimpleTrace>(f.channels[r], tbin0, std::move(q)));
            }
        }
        auto tv = std::make_shared<ITrace::vector>(traces);
        auto frame = std::make_shared<Aux::SimpleFrame>(ident, framelets[0].tickinfo[0], tv, framelets[0].tickinfo[1],
                                                        cmm);
        for (const auto& t : m_frame_tags) frame->tag_frame(t);
        size_t first = 0;
        for (auto& f : framelets) {
            const size_t n = f.channels.size();
            if (!untagged(f.tag)) {
                IFrame::trace_list_t idx(n);
                std::iota(idx.begin(), idx.end(), first);
                frame->tag_traces(f.tag, idx, f.summary);
            }
            first += n;
        }
        return frame;
    }

    std::string m_inname{"wct-frame.tar.gz"};
    std::vector<std::string> m_tags, m_frame_tags;
    std::unique_ptr<TarReader> m_in;
    Entry m_cur;
    bool m_have{false}, m_eos_sent{false};
};

}  // namespace Hip
}  // namespace WireCell

WIRECELL_HIP_FACTORY(FrameFileSink, WireCell::Hip::FrameFileSink)
WIRECELL_HIP_FACTORY(FrameFileSource, WireCell::Hip::FrameFileSource)
