// extern "C" front of the component layer (include/wct_host.h): lets a C
// or Python caller configure components by JSON, call an IDFT through its
// virtual interface and push frames through an IFrameFilter, the way the
// reference's application and tests drive them.
#include <algorithm>
#include <cstring>
#include <string>
#include <tuple>

#include "../../include/wct_host.h"
#include "Factory.h"
#include "Frame.h"
#include "Interface.h"
#include "PeakFinding.h"
#include "RoiFormation.h"
#include "RoiRefinement.h"
#include "Tensor.h"

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
    // input side: tags to apply when the frame is built
    std::vector<std::string> in_frame_tags;
    std::vector<std::tuple<std::string, IFrame::trace_list_t, IFrame::trace_summary_t>> in_trace_tags;
    IFrame::pointer get()
    {
        if (!frame) {
            auto tv = std::make_shared<ITrace::vector>(traces);
            auto sf = std::make_shared<Aux::SimpleFrame>(ident, time, tv, tick, masks);
            for (const auto& t : in_frame_tags) sf->tag_frame(t);
            for (const auto& [t, idx, sum] : in_trace_tags) sf->tag_traces(t, idx, sum);
            frame = sf;
        }
        return frame;
    }
};
// A tensor set under construction (input side) or a filter's result.
struct TensorSetHandle {
    int ident{0};
    Configuration md{Configuration::makeObject()};
    std::vector<ITensor::pointer> tensors;
    ITensorSet::pointer set;   // set once built or when it is an output
    ITensorSet::pointer get()
    {
        if (!set) set = std::make_shared<Aux::SimpleTensorSet>(ident, md, std::make_shared<ITensor::vector>(tensors));
        return set;
    }
};

int copy_json(const Configuration& c, char* buf, size_t cap)
{
    const std::string s = c.dump();
    if (buf && cap) {
        std::strncpy(buf, s.c_str(), cap - 1);
        buf[cap - 1] = 0;
    }
    return (int)s.size();
}

Configuration parse_md(const char* json)
{
    if (!json || !*json) return Configuration::makeObject();
    return Configuration::parse(json);
}
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

int wcth_channel_response(const char* tn, int channel, float* out, int cap, int* n, double* tmin,
                          double* binsize)
{
    try {
        auto cr = Factory::find_tn<IChannelResponse>(tn);
        const auto& v = cr->channel_response(channel);
        const Binning b = cr->channel_response_binning();
        if (n) *n = (int)v.size();
        if (tmin) *tmin = b.minval;
        if (binsize) *binsize = b.nbins ? b.binsize() : 0.0;
        if (out) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), cap > 0 ? cap : 0), out);
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


static std::vector<Hip::RoiList> unflatten(const int* count, const int* pairs, int nw)
{
    std::vector<Hip::RoiList> rows(nw);
    int k = 0;
    for (int w = 0; w < nw; ++w)
        for (int i = 0; i < count[w]; ++i, ++k) rows[w].emplace_back(pairs[2 * k], pairs[2 * k + 1]);
    return rows;
}

int wcth_loose_rois(const float* r, int nw, int nt, long ld, const int* bad_ptr, const int* bad_lo,
                    const int* bad_hi, const float* rms, const int* tight_count, const int* tight_pairs,
                    int rebin, double l_factor, double l_max_th, double l_factor1, int l_short_length,
                    int l_jump_one_bin, int* roi_count, int* roi_pairs, int cap)
{
    if (!r || nw < 0 || nt < 0 || ld < nt || !roi_count || rebin < 1 || (bad_ptr && (!bad_lo || !bad_hi)) ||
        (tight_count && !tight_pairs)) {
        g_err = "loose_rois: bad arguments";
        return -1;
    }
    try {
        Hip::LooseParams lp;
        lp.rebin = rebin;
        lp.factor = l_factor;
        lp.max_th = l_max_th;
        lp.factor1 = l_factor1;
        lp.short_length = l_short_length;
        lp.jump_one_bin = l_jump_one_bin;
        std::vector<Hip::RoiList> tight;
        if (tight_count) tight = unflatten(tight_count, tight_pairs, nw);
        auto rows = Hip::loose_rois(r, nw, nt, ld, bad_ptr, bad_lo, bad_hi, rms, tight_count ? &tight : nullptr, lp);
        int tot = 0;
        bool over = false;
        for (int w = 0; w < nw; ++w) {
            roi_count[w] = (int)rows[w].size();
            for (const auto& q : rows[w]) {
                if (tot < cap) {
                    roi_pairs[2 * tot] = q.first;
                    roi_pairs[2 * tot + 1] = q.second;
                } else {
                    over = true;
                }
                ++tot;
            }
        }
        return over ? -2 : tot;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_apply_roi(float* r, int nw, int nt, long ld, const int* roi_count, const int* roi_pairs, int linear)
{
    if (!r || nw < 0 || nt < 0 || ld < nt || !roi_count || !roi_pairs) {
        g_err = "apply_roi: bad arguments";
        return -1;
    }
    auto rows = unflatten(roi_count, roi_pairs, nw);
    for (int w = 0; w < nw; ++w)
        for (const auto& q : rows[w])
            if (q.first < 0 || q.second >= nt || q.first > q.second) {
                g_err = "apply_roi: ROI [" + std::to_string(q.first) + ", " + std::to_string(q.second) +
                        "] of row " + std::to_string(w) + " outside [0, " + std::to_string(nt) + ")";
                return -1;
            }
    Hip::apply_roi(r, nw, nt, ld, rows, linear != 0);
    return 0;
}

int wcth_find_peaks(const float* signal, int n, int max_peaks, double sigma, double threshold, double* x,
                    double* y, int cap)
{
    if (!signal || n < 0 || cap < 0 || (cap > 0 && (!x || !y))) {
        g_err = "find_peaks: bad arguments";
        return -1;
    }
    try {
        Hip::PeakSearch ps;
        ps.max_peaks = max_peaks;
        ps.sigma = sigma;
        ps.threshold = threshold;
        const auto pk = Hip::find_peaks(std::vector<float>(signal, signal + n), ps);
        const int np = (int)pk.x.size();
        for (int i = 0; i < np && i < cap; ++i) {
            x[i] = pk.x[i];
            y[i] = pk.y[i];
        }
        return np > cap ? -2 : np;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void* wcth_refine_new(int plane, int nw, int nt, const double* params, int wrapped)
{
    try {
        Hip::RefineParams p;
        if (params) {
            p.th_factor = (float)params[0];
            p.fake_signal_low_th = (float)params[1];
            p.fake_signal_high_th = (float)params[2];
            p.fake_signal_low_th_ind_factor = (float)params[3];
            p.fake_signal_high_th_ind_factor = (float)params[4];
            p.pad = (int)params[5];
            p.break_roi_loop = (int)params[6];
            p.th_peak = (float)params[7];
            p.sep_peak = (float)params[8];
            p.low_peak_sep_threshold_pre = (float)params[9];
            p.max_npeaks = (int)params[10];
            p.sigma = (float)params[11];
            p.th_percent = (float)params[12];
        }
        p.wrapped = wrapped != 0;
        return new Hip::RoiRefinement(plane, nw, nt, p);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void wcth_refine_free(void* h) { delete static_cast<Hip::RoiRefinement*>(h); }

int wcth_refine_load(void* h, const float* r, long ld, const int* bad_ptr, const int* bad_lo, const int* bad_hi,
                     const float* rms, const int* tight_count, const int* tight_pairs, const int* loose_count,
                     const int* loose_pairs)
{
    auto* R = static_cast<Hip::RoiRefinement*>(h);
    if (!R || !r || !rms || !tight_count || (bad_ptr && (!bad_lo || !bad_hi))) {
        g_err = "refine_load: bad arguments";
        return -1;
    }
    try {
        const int nw = R->nwires(), nt = R->nticks();
        if (ld < nt) throw std::invalid_argument("refine_load: row stride below the tick count");
        std::vector<Hip::RoiList> bad;
        if (bad_ptr) {
            bad.resize(nw);
            for (int w = 0; w < nw; ++w)
                for (int q = bad_ptr[w]; q < bad_ptr[w + 1]; ++q) bad[w].emplace_back(bad_lo[q], bad_hi[q]);
        }
        std::vector<Hip::RoiList> loose;
        if (loose_count) loose = unflatten(loose_count, loose_pairs, nw);
        R->load(r, ld, bad, std::vector<float>(rms, rms + nw), unflatten(tight_count, tight_pairs, nw), loose);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_refine_step(void* h, const char* step)
{
    auto* R = static_cast<Hip::RoiRefinement*>(h);
    if (!R || !step) {
        g_err = "refine_step: bad arguments";
        return -1;
    }
    try {
        const std::string s(step);
        if (s == "cleanup") R->cleanup_rois();
        else if (s == "merge") R->generate_merge_rois();
        else if (s == "check") R->check_rois();
        else if (s == "break") R->break_rois();
        else if (s == "shrink") R->shrink_rois();
        else if (s == "cleanup_collection") R->cleanup_collection_rois();
        else if (s == "cleanup_induction") R->cleanup_induction_rois();
        else if (s == "extend") R->extend_rois();
        else if (s == "all") R->refine();
        else throw std::invalid_argument("refine_step: unknown step '" + s + "'");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_refine_rois(void* h, int extended, int* roi_count, int* roi_pairs, int cap)
{
    auto* R = static_cast<Hip::RoiRefinement*>(h);
    if (!R || !roi_count || cap < 0 || (cap > 0 && !roi_pairs)) {
        g_err = "refine_rois: bad arguments";
        return -1;
    }
    const auto rows = R->rois(extended != 0);
    int tot = 0;
    for (size_t w = 0; w < rows.size(); ++w) {
        roi_count[w] = (int)rows[w].size();
        for (const auto& q : rows[w]) {
            if (tot < cap) {
                roi_pairs[2 * tot] = q.first;
                roi_pairs[2 * tot + 1] = q.second;
            }
            ++tot;
        }
    }
    return tot > cap ? -2 : tot;
}

int wcth_refine_contents(void* h, int w, float* out, int cap)
{
    auto* R = static_cast<Hip::RoiRefinement*>(h);
    if (!R || w < 0 || w >= R->nwires() || cap < 0 || (cap > 0 && !out)) {
        g_err = "refine_contents: bad arguments";
        return -1;
    }
    int tot = 0;
    for (const auto& c : R->contents(w))
        for (float v : c) {
            if (tot < cap) out[tot] = v;
            ++tot;
        }
    return tot > cap ? -2 : tot;
}

void* wcth_tensorset_new(int ident, const char* metadata_json)
{
    try {
        auto h = new TensorSetHandle;
        h->ident = ident;
        h->md = parse_md(metadata_json);
        return h;
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}

void wcth_tensorset_free(void* set) { delete (TensorSetHandle*)set; }

int wcth_tensorset_add(void* set, const char* dtype, int ndim, const size_t* shape, const void* data,
                       const char* metadata_json)
{
    try {
        auto h = (TensorSetHandle*)set;
        if (!h || h->set || !dtype || ndim < 0 || (ndim && !shape)) {
            g_err = "tensorset_add: bad arguments (or the set was already used)";
            return -1;
        }
        ITensor::shape_t sh(shape, shape + ndim);
        h->tensors.push_back(std::make_shared<Aux::SimpleTensor>(dtype, sh, data, parse_md(metadata_json)));
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_tensorset_filter(const char* tn, void* in, void** out)
{
    try {
        if (!out) {
            g_err = "null output handle";
            return -1;
        }
        *out = nullptr;
        auto filt = Factory::find_tn<ITensorSetFilter>(tn);
        ITensorSet::pointer sin = in ? ((TensorSetHandle*)in)->get() : nullptr;
        ITensorSet::pointer sout;
        if (!(*filt)(sin, sout)) {
            g_err = "tensor set filter returned false";
            return -1;
        }
        if (sout) {
            auto h = new TensorSetHandle;
            h->set = sout;
            *out = h;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_tensorset_info(void* set, int* ident, int* ntensors, char* metadata_json, size_t cap)
{
    auto h = (TensorSetHandle*)set;
    if (!h) return -1;
    auto s = h->get();
    if (ident) *ident = s->ident();
    auto tv = s->tensors();
    if (ntensors) *ntensors = tv ? (int)tv->size() : 0;
    return copy_json(s->metadata(), metadata_json, cap);
}

int wcth_tensorset_tensor(void* set, int i, char* dtype, int* ndim, size_t* shape, const void** data,
                          size_t* nbytes)
{
    auto h = (TensorSetHandle*)set;
    if (!h) return -1;
    auto tv = h->get()->tensors();
    if (!tv || i < 0 || i >= (int)tv->size()) return -1;
    const auto& t = tv->at(i);
    const auto sh = t->shape();
    if (dtype) {
        std::strncpy(dtype, t->dtype().c_str(), 7);
        dtype[7] = 0;
    }
    if (ndim) *ndim = (int)sh.size();
    if (shape)
        for (size_t k = 0; k < sh.size() && k < 8; ++k) shape[k] = sh[k];
    if (data) *data = t->data();
    if (nbytes) *nbytes = t->size();
    return 0;
}

int wcth_tensorset_tensor_md(void* set, int i, char* metadata_json, size_t cap)
{
    auto h = (TensorSetHandle*)set;
    if (!h) return -1;
    auto tv = h->get()->tensors();
    if (!tv || i < 0 || i >= (int)tv->size()) return -1;
    return copy_json(tv->at(i)->metadata(), metadata_json, cap);
}

int wcth_frame_sink(const char* tn, void* frame)
{
    try {
        auto sink = Factory::find_tn<IFrameSink>(tn);
        IFrame::pointer f = frame ? ((FrameHandle*)frame)->get() : nullptr;
        if (!(*sink)(f)) {
            g_err = "frame sink returned false";
            return -1;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_frame_source(const char* tn, void** out)
{
    try {
        if (!out) {
            g_err = "null output handle";
            return -1;
        }
        *out = nullptr;
        auto src = Factory::find_tn<IFrameSource>(tn);
        IFrame::pointer f;
        if (!(*src)(f)) return 1;
        if (f) {
            auto h = new FrameHandle;
            h->frame = f;
            *out = h;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int wcth_frame_tag_frame(void* frame, const char* tag)
{
    auto f = (FrameHandle*)frame;
    if (!f || f->frame || !tag) {
        g_err = "frame_tag_frame: bad handle (or the frame was already used)";
        return -1;
    }
    f->in_frame_tags.push_back(tag);
    return 0;
}

int wcth_frame_tag_traces(void* frame, const char* tag, const size_t* idx, int n, const double* summary,
                          int nsummary)
{
    auto f = (FrameHandle*)frame;
    if (!f || f->frame || !tag || n < 0 || (n && !idx) || nsummary < 0 || (nsummary && !summary)) {
        g_err = "frame_tag_traces: bad arguments (or the frame was already used)";
        return -1;
    }
    f->in_trace_tags.emplace_back(tag, IFrame::trace_list_t(idx, idx + n),
                                  IFrame::trace_summary_t(summary, summary + nsummary));
    return 0;
}
}  // extern "C"
