// extern "C" front of the component layer (include/wct_host.h): lets a C
// or Python caller configure components by JSON, call an IDFT through its
// virtual interface and push frames through an IFrameFilter, the way the
// reference's application and tests drive them.
#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/wct_host.h"
#include "Factory.h"
#include "Frame.h"
#include "Interface.h"
#include "RoiFormation.h"

using namespace WireCell;

namespace {
thread_local std::string g_err;

int fail(const std::exception& e)
{
    g_err = e.what();
    return -1;
}

// A frame under construction (input side) or a filter's result.
struct FrameHandle {
    int ident{0};
    double time{0}, tick{500.0};
    std::vector<ITrace::pointer> traces;
    Waveform::ChannelMaskMap masks;
    IFrame::pointer frame;   // set once built or when it is an output
    std::vector<std::string> frame_tags, trace_tags;
    IFrame::pointer get()
    {
        if (!frame) {
            auto tv = std::make_shared<ITrace::vector>(traces);
            frame = std::make_shared<Aux::SimpleFrame>(ident, time, tv, tick, masks);
        }
        return frame;
    }
};
}  // namespace

extern "C" {

const char* wcth_last_error(void) { return g_err.c_str(); }

int wcth_configure(const char* json)
{
    try {
        Factory::configure_all(Configuration::parse(json ? json : ""));
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_reset(void)
{
    try {
        Factory::clear_instances();
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_known_types(char* buf, size_t cap)
{
    std::string s;
    for (auto& t : Factory::known_types()) s += t + "\n";
    if (buf && cap) {
        std::strncpy(buf, s.c_str(), cap - 1);
        buf[cap - 1] = 0;
    }
    return (int)s.size();
}

int wcth_dft(const char* tn, int op, const float* in, float* out, int nrows, int ncols, int axis)
{
    try {
        auto dft = Factory::find_tn<IDFT>(tn);
        auto ci = reinterpret_cast<const IDFT::complex_t*>(in);
        auto co = reinterpret_cast<IDFT::complex_t*>(out);
        switch (op) {
        case WCTH_FWD1D: dft->fwd1d(ci, co, nrows); break;
        case WCTH_INV1D: dft->inv1d(ci, co, nrows); break;
        case WCTH_FWD1B: dft->fwd1b(ci, co, nrows, ncols, axis); break;
        case WCTH_INV1B: dft->inv1b(ci, co, nrows, ncols, axis); break;
        case WCTH_FWD2D: dft->fwd2d(ci, co, nrows, ncols); break;
        case WCTH_INV2D: dft->inv2d(ci, co, nrows, ncols); break;
        case WCTH_TRANSPOSE_REAL: dft->transpose(in, out, nrows, ncols); break;
        case WCTH_TRANSPOSE_COMPLEX: dft->transpose(ci, co, nrows, ncols); break;
        default: g_err = "unknown IDFT operation"; return -1;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_filter_waveform(const char* tn, int n, float* out)
{
    try {
        auto f = Factory::find_tn<IFilterWaveform>(tn);
        auto v = f->filter_waveform(n);
        std::copy(v.begin(), v.end(), out);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void* wcth_frame_new(int ident, double time, double tick)
{
    auto h = new FrameHandle;
    h->ident = ident;
    h->time = time;
    h->tick = tick;
    return h;
}

void wcth_frame_free(void* h) { delete (FrameHandle*)h; }

int wcth_frame_add_trace(void* h, int channel, int tbin, int n, const float* charge)
{
    auto f = (FrameHandle*)h;
    if (!f || f->frame || n < 0 || (n && !charge)) {
        g_err = "bad frame handle or trace";
        return -1;
    }
    f->traces.push_back(std::make_shared<Aux::SimpleTrace>(channel, tbin, ITrace::ChargeSequence(charge, charge + n)));
    return 0;
}

int wcth_frame_add_mask(void* h, const char* name, int channel, int first, int second)
{
    auto f = (FrameHandle*)h;
    if (!f || f->frame || !name) {
        g_err = "bad frame handle";
        return -1;
    }
    f->masks[name][channel].emplace_back(first, second);
    return 0;
}

int wcth_frame_filter(const char* tn, void* in, void** out)
{
    try {
        if (!out) {
            g_err = "null output handle";
            return -1;
        }
        *out = nullptr;
        auto filt = Factory::find_tn<IFrameFilter>(tn);
        IFrame::pointer fin = in ? ((FrameHandle*)in)->get() : nullptr;
        IFrame::pointer fout;
        if (!(*filt)(fin, fout)) {
            g_err = "frame filter returned false";
            return -1;
        }
        if (fout) {
            auto h = new FrameHandle;
            h->frame = fout;
            *out = h;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_frame_info(void* h, int* ident, double* time, double* tick, int* ntraces)
{
    auto f = (FrameHandle*)h;
    if (!f) return -1;
    auto fr = f->get();
    if (ident) *ident = fr->ident();
    if (time) *time = fr->time();
    if (tick) *tick = fr->tick();
    if (ntraces) *ntraces = (int)fr->traces()->size();
    return 0;
}

int wcth_frame_trace(void* h, int i, int* channel, int* tbin, int* n, const float** charge)
{
    auto f = (FrameHandle*)h;
    if (!f) return -1;
    auto tv = f->get()->traces();
    if (i < 0 || i >= (int)tv->size()) {
        g_err = "trace index out of range";
        return -1;
    }
    const auto& t = (*tv)[i];
    if (channel) *channel = t->channel();
    if (tbin) *tbin = t->tbin();
    if (n) *n = (int)t->charge().size();
    if (charge) *charge = t->charge().data();
    return 0;
}

int wcth_frame_tags(void* h, int which, int i, const char** tag)
{
    auto f = (FrameHandle*)h;
    if (!f) return -1;
    const auto& tags = which == 0 ? f->get()->frame_tags() : f->get()->trace_tags();
    if (i < 0) return (int)tags.size();
    if (i >= (int)tags.size()) return -1;
    *tag = tags[i].c_str();
    return (int)tags.size();
}

int wcth_frame_tagged(void* h, const char* tag, const size_t** idx, int* n)
{
    auto f = (FrameHandle*)h;
    if (!f || !tag) return -1;
    const auto& v = f->get()->tagged_traces(tag);
    *idx = v.data();
    *n = (int)v.size();
    return 0;
}

int wcth_frame_summary(void* h, const char* tag, const double** val, int* n)
{
    auto f = (FrameHandle*)h;
    if (!f || !tag) return -1;
    const auto& v = f->get()->trace_summary(tag);
    *val = v.data();
    *n = (int)v.size();
    return 0;
}

int wcth_frame_masks(void* h, const char* name, int* channels, int* firsts, int* seconds, int cap)
{
    auto f = (FrameHandle*)h;
    if (!f || !name) return -1;
    const auto masks = f->get()->masks();
    auto it = masks.find(name);
    if (it == masks.end()) return 0;
    int k = 0;
    for (const auto& cm : it->second)
        for (const auto& r : cm.second) {
            if (k < cap) {
                channels[k] = cm.first;
                firsts[k] = r.first;
                seconds[k] = r.second;
            }
            ++k;
        }
    return k;
}

int wcth_roi_lists(const unsigned char* mask, int nw, int nt, long ld, int pad, float asy, int* roi_count,
                   int* roi_pairs, int cap)
{
    if (!mask || nw < 0 || nt < 0 || ld < nt || !roi_count) return -1;
    auto rows = Hip::tight_rois(mask, nw, nt, ld, pad, asy);
    int tot = 0;
    bool over = false;
    for (int w = 0; w < nw; ++w) {
        roi_count[w] = (int)rows[w].size();
        for (const auto& r : rows[w]) {
            if (tot < cap) {
                roi_pairs[2 * tot] = r.first;
                roi_pairs[2 * tot + 1] = r.second;
            } else {
                over = true;
            }
            ++tot;
        }
    }
    return over ? -2 : tot;
}

}  // extern "C"
