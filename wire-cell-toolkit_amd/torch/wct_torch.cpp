// libtorch front of the differentiable deconvolution: the forward
// response convolution and the filtered deconvolution as
// torch::autograd::Function subclasses whose backward passes are the
// hand-written adjoints of libwct_hip (the fused row/wire/row kernels run
// with the conjugate kernel, wct_decon_job.adjoint).
//
// Both operators are real 2D circular filters:
//   forward  D = F q,    kernel R exp(+i phase) / (nw nt)
//   decon    y = G d,    kernel Fw(s_w) Ft(s_t) exp(-i phase) / (R nw nt)
// so dL/dd = G^T dL/dy and dL/dq = F^T dL/dD, and the filter-width
// gradients are dL/ds = < G_{dF/ds} d, dL/dy > with the closed-form
// derivative of SigProc::HfFilter (sigproc/src/HfFilter.cxx:38-52),
//   g(f) = exp(-(|f|/s)^p / 2),  dg/ds = g (p/2) (|f|/s)^p / s.
// Kernels are rebuilt from the response in FP64 on the device for each
// width (a small cache keyed by the widths), never stored with the graph.
#include <torch/extension.h>

#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "../../include/wct_hip.h"

namespace {

void check(int rc, const char* what)
{
    if (rc != WCT_OK) throw std::runtime_error(std::string("wct_torch ") + what + ": " + wct_last_error());
}

// |f| of the HfFilter frequency grid (HfFilter.cxx:45-49).
std::vector<double> freq_grid(int n, double max_freq)
{
    std::vector<double> f(n);
    for (int i = 0; i < n; ++i) {
        double v = i * 1.0 / n * 2 * max_freq;
        if (v > max_freq) v -= 2 * max_freq;
        f[i] = std::fabs(v);
    }
    return f;
}

std::vector<double> hf_filter(int n, double max_freq, double sigma, double power, bool flag, bool deriv)
{
    const auto a = freq_grid(n, max_freq);
    std::vector<double> g(n);
    for (int i = 0; i < n; ++i) {
        const double x = std::pow(a[i] / sigma, power);
        const double v = std::exp(-0.5 * x);
        g[i] = deriv ? v * 0.5 * power * x / sigma : v;
    }
    if (flag) g[0] = 0.0;
    return g;
}

// One plane type: response rows, geometry, filter shapes, the forward
// kernel and the decon kernels built so far.
class Model {
  public:
    Model(std::vector<double> resp, int nr, int nw, int nt, int wire_shift, int time_shift, double time_power,
          double time_max_freq, bool time_flag, double wire_power, double wire_max_freq, double cut_tol,
          int64_t stream, int cache)
        : m_resp(std::move(resp)), m_nr(nr), m_nw(nw), m_nt(nt), m_ws(wire_shift), m_ts(time_shift),
          m_tp(time_power), m_tf(time_max_freq), m_tflag(time_flag), m_wp(wire_power), m_wf(wire_max_freq),
          m_cut(cut_tol), m_cap(cache)
    {
        if ((int)m_resp.size() != nr * nt) throw std::invalid_argument("wct_torch: response size != nr * nt");
        // a backward pass holds three kernels at once (value, d/ds_t,
        // d/ds_w); the eviction below never drops the newest entry
        if (cache < 3) throw std::invalid_argument("wct_torch: the kernel cache must hold at least 3 kernels");
        check(wct_forward_kernel_create(&m_fwd, m_resp.data(), nr, nw, nt, wire_shift, time_shift,
                                        reinterpret_cast<void*>(stream)),
              "forward kernel");
    }
    ~Model()
    {
        for (auto& kv : m_cache) wct_decon_kernel_destroy(kv.second);
        if (m_fwd) wct_decon_kernel_destroy(m_fwd);
    }
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;

    int nw() const { return m_nw; }
    int nt() const { return m_nt; }

    // which: 0 value, 1 d/ds_t, 2 d/ds_w
    wct_decon_kernel kernel(double st, double sw, int which, int64_t stream)
    {
        std::lock_guard<std::mutex> lk(m_mutex);
        const auto key = std::make_tuple(st, sw, which);
        auto it = m_cache.find(key);
        if (it != m_cache.end()) return it->second;
        const auto tf = hf_filter(m_nt, m_tf, st, m_tp, m_tflag, which == 1);
        const auto wf = hf_filter(m_nw, m_wf, sw, m_wp, false, which == 2);
        wct_decon_kernel k = nullptr;
        check(wct_decon_kernel_create(&k, m_resp.data(), m_nr, m_nw, m_nt, wf.data(), tf.data(), m_ws, m_ts,
                                      m_cut, reinterpret_cast<void*>(stream)),
              "decon kernel");
        m_cache[key] = k;
        m_order.push_back(key);
        while ((int)m_order.size() > m_cap) {
            // the oldest kernel may still be read by queued work
            check(wct_stream_sync(reinterpret_cast<void*>(stream)), "sync");
            wct_decon_kernel_destroy(m_cache[m_order.front()]);
            m_cache.erase(m_order.front());
            m_order.pop_front();
        }
        return k;
    }

    wct_decon_kernel forward_kernel() const { return m_fwd; }

  private:
    std::vector<double> m_resp;
    int m_nr, m_nw, m_nt, m_ws, m_ts;
    double m_tp, m_tf;
    bool m_tflag;
    double m_wp, m_wf, m_cut;
    int m_cap;
    wct_decon_kernel m_fwd{nullptr};
    std::mutex m_mutex;
    std::map<std::tuple<double, double, int>, wct_decon_kernel> m_cache;
    std::list<std::tuple<double, double, int>> m_order;
};

// Applies kernel k (or its adjoint) to every frame of x (B, nw, nt).
torch::Tensor run_op(Model& m, wct_decon_kernel k, const torch::Tensor& x_, bool adjoint, int64_t stream)
{
    TORCH_CHECK(x_.is_cuda() && x_.scalar_type() == torch::kFloat32 && x_.dim() == 3, "wct_torch: x must be a "
                "float32 device tensor (B, nw, nt)");
    TORCH_CHECK(x_.size(1) == m.nw() && x_.size(2) == m.nt(), "wct_torch: frame shape does not match the model");
    const torch::Tensor x = x_.contiguous();
    torch::Tensor out = torch::empty_like(x);
    const int B = (int)x.size(0);
    if (B == 0) return out;
    std::vector<wct_decon_job> jobs(B);
    for (int b = 0; b < B; ++b) {
        wct_decon_job& j = jobs[b];
        std::memset(&j, 0, sizeof(j));
        j.kernel = k;
        j.adc = x[b].data_ptr<float>();
        j.ld_adc = m.nt();
        j.out = out[b].data_ptr<float>();
        j.ld_out = m.nt();
        j.adjoint = adjoint ? 1 : 0;
    }
    const size_t ws = wct_decon_workspace_size(jobs.data(), B);
    torch::Tensor work = torch::empty({(int64_t)ws}, x.options().dtype(torch::kUInt8));
    check(wct_decon_run(jobs.data(), B, work.data_ptr(), ws, reinterpret_cast<void*>(stream)), "run");
    return out;
}

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

struct ForwardConv : public torch::autograd::Function<ForwardConv> {
    static torch::Tensor forward(AutogradContext* ctx, const torch::Tensor& q, int64_t model, int64_t stream)
    {
        ctx->saved_data["model"] = model;
        ctx->saved_data["stream"] = stream;
        Model& m = *reinterpret_cast<Model*>(model);
        return run_op(m, m.forward_kernel(), q, false, stream);
    }
    static variable_list backward(AutogradContext* ctx, variable_list g)
    {
        Model& m = *reinterpret_cast<Model*>(ctx->saved_data["model"].toInt());
        const int64_t stream = ctx->saved_data["stream"].toInt();
        return {run_op(m, m.forward_kernel(), g[0], true, stream), torch::Tensor(), torch::Tensor()};
    }
};

struct Decon : public torch::autograd::Function<Decon> {
    static torch::Tensor forward(AutogradContext* ctx, const torch::Tensor& d, const torch::Tensor& sigma_t,
                                 const torch::Tensor& sigma_w, int64_t model, int64_t stream)
    {
        const double st = sigma_t.item<double>(), sw = sigma_w.item<double>();
        ctx->saved_data["model"] = model;
        ctx->saved_data["stream"] = stream;
        ctx->saved_data["st"] = st;
        ctx->saved_data["sw"] = sw;
        ctx->save_for_backward({d, sigma_t, sigma_w});
        Model& m = *reinterpret_cast<Model*>(model);
        return run_op(m, m.kernel(st, sw, 0, stream), d, false, stream);
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        Model& m = *reinterpret_cast<Model*>(ctx->saved_data["model"].toInt());
        const int64_t stream = ctx->saved_data["stream"].toInt();
        const double st = ctx->saved_data["st"].toDouble(), sw = ctx->saved_data["sw"].toDouble();
        const auto saved = ctx->get_saved_variables();
        const torch::Tensor& d = saved[0];
        const torch::Tensor g = grads[0].contiguous();
        torch::Tensor gd, gst, gsw;
        if (ctx->needs_input_grad(0)) gd = run_op(m, m.kernel(st, sw, 0, stream), g, true, stream);
        if (ctx->needs_input_grad(1)) {
            const torch::Tensor yt = run_op(m, m.kernel(st, sw, 1, stream), d, false, stream);
            gst = (yt.to(torch::kFloat64) * g.to(torch::kFloat64)).sum().to(saved[1].options()).reshape({});
        }
        if (ctx->needs_input_grad(2)) {
            const torch::Tensor yw = run_op(m, m.kernel(st, sw, 2, stream), d, false, stream);
            gsw = (yw.to(torch::kFloat64) * g.to(torch::kFloat64)).sum().to(saved[2].options()).reshape({});
        }
        return {gd, gst, gsw, torch::Tensor(), torch::Tensor()};
    }
};

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, mod)
{
    mod.doc() = "Differentiable forward response convolution and filtered deconvolution (libtorch autograd "
                "Functions over libwct_hip)";
    py::class_<Model>(mod, "Model")
        .def(py::init([](std::vector<double> resp, int nr, int nw, int nt, int wire_shift, int time_shift,
                         double time_power, double time_max_freq, bool time_flag, double wire_power,
                         double wire_max_freq, double cut_tol, int64_t stream, int cache) {
                 return new Model(std::move(resp), nr, nw, nt, wire_shift, time_shift, time_power, time_max_freq,
                                  time_flag, wire_power, wire_max_freq, cut_tol, stream, cache);
             }),
             py::arg("resp"), py::arg("nr"), py::arg("nw"), py::arg("nt"), py::arg("wire_shift"),
             py::arg("time_shift"), py::arg("time_power"), py::arg("time_max_freq"), py::arg("time_flag"),
             py::arg("wire_power"), py::arg("wire_max_freq"), py::arg("cut_tol"), py::arg("stream"),
             py::arg("cache") = 8)
        .def("ptr", [](Model& m) { return reinterpret_cast<int64_t>(&m); });
    mod.def("forward_conv",
            [](const torch::Tensor& q, int64_t model, int64_t stream) { return ForwardConv::apply(q, model, stream); },
            "D = F q (autograd: backward applies F^T)");
    mod.def("decon",
            [](const torch::Tensor& d, const torch::Tensor& st, const torch::Tensor& sw, int64_t model,
               int64_t stream) { return Decon::apply(d, st, sw, model, stream); },
            "y = G(s_t, s_w) d (autograd: G^T for d, derivative-filter passes for the widths)");
    mod.def("apply",
            [](int64_t model, const torch::Tensor& x, double st, double sw, int which, bool adjoint, int64_t stream) {
                Model& m = *reinterpret_cast<Model*>(model);
                wct_decon_kernel k = which < 0 ? m.forward_kernel() : m.kernel(st, sw, which, stream);
                return run_op(m, k, x, adjoint, stream);
            },
            "one operator without autograd: which = -1 forward, 0 decon, 1 d/ds_t, 2 d/ds_w");
}
