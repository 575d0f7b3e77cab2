// HipDFT: the IDFT component on the gfx950 FFT kernels.  Host arrays in
// and out, as IDFT requires (iface/inc/WireCellIface/IDFT.h:40-66); each
// call stages its array through device memory on a stream of the calling
// thread, so concurrent calls from different threads are independent
// (the reference's CudaDFT/TorchDFT slot, pytorch/src/DFT.cxx:42-123,
// which serialises on a semaphore instead).  Errors throw, carrying the
// C ABI's message.
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "../../include/wct_hip.h"
#include "Factory.h"
#include "Interface.h"

namespace WireCell {
namespace Hip {

struct DftError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

static void check(int rc, const char* what)
{
    if (rc != WCT_OK) throw DftError(std::string("HipDFT ") + what + ": " + wct_last_error());
}

class HipDFT : public IDFT, public IConfigurable {
  public:
    ~HipDFT() override
    {
        for (auto& kv : m_ctx) {
            wct_free(kv.second.dbuf);
            wct_stream_destroy(kv.second.stream);
        }
    }

    Configuration default_configuration() const override
    {
        Configuration cfg;
        cfg["device"] = m_device;
        return cfg;
    }
    void configure(const Configuration& cfg) override { m_device = get(cfg, "device", m_device); }

    void fwd1d(const complex_t* in, complex_t* out, int size) const override
    {
        run(in, out, size, 1, [&](float* d, void* s) { return wct_dft_fwd1d(d, d, size, s); }, "fwd1d");
    }
    void inv1d(const complex_t* in, complex_t* out, int size) const override
    {
        run(in, out, size, 1, [&](float* d, void* s) { return wct_dft_inv1d(d, d, size, s); }, "inv1d");
    }
    void fwd1b(const complex_t* in, complex_t* out, int nrows, int ncols, int axis) const override
    {
        run(in, out, nrows, ncols,
            [&](float* d, void* s) { return wct_dft_fwd1b(d, d, nrows, ncols, axis, s); }, "fwd1b");
    }
    void inv1b(const complex_t* in, complex_t* out, int nrows, int ncols, int axis) const override
    {
        run(in, out, nrows, ncols,
            [&](float* d, void* s) { return wct_dft_inv1b(d, d, nrows, ncols, axis, s); }, "inv1b");
    }
    void fwd2d(const complex_t* in, complex_t* out, int nrows, int ncols) const override
    {
        run(in, out, nrows, ncols, [&](float* d, void* s) { return wct_dft_fwd2d(d, d, nrows, ncols, s); },
            "fwd2d");
    }
    void inv2d(const complex_t* in, complex_t* out, int nrows, int ncols) const override
    {
        run(in, out, nrows, ncols, [&](float* d, void* s) { return wct_dft_inv2d(d, d, nrows, ncols, s); },
            "inv2d");
    }
    void transpose(const scalar_t* in, scalar_t* out, int nrows, int ncols) const override
    {
        stage(in, out, (size_t)nrows * ncols * sizeof(scalar_t), nrows, ncols,
              [&](float* d, void* s) { return wct_dft_transpose_real(d, d, nrows, ncols, s); }, "transpose");
    }
    void transpose(const complex_t* in, complex_t* out, int nrows, int ncols) const override
    {
        stage(in, out, (size_t)nrows * ncols * sizeof(complex_t), nrows, ncols,
              [&](float* d, void* s) { return wct_dft_transpose_complex(d, d, nrows, ncols, s); },
              "transpose");
    }

  private:
    struct Ctx {
        void* stream{nullptr};
        void* dbuf{nullptr};
        size_t cap{0};
    };

    Ctx& context(size_t bytes) const
    {
        std::lock_guard<std::mutex> lk(m_mutex);
        Ctx& c = m_ctx[std::this_thread::get_id()];
        check(wct_set_device(m_device), "set device");
        if (!c.stream) check(wct_stream_create(&c.stream), "stream");
        if (c.cap < bytes) {
            check(wct_stream_sync(c.stream), "sync");
            check(wct_free(c.dbuf), "free");
            c.dbuf = nullptr;
            c.cap = 0;
            check(wct_malloc(&c.dbuf, bytes), "malloc");
            c.cap = bytes;
        }
        return c;
    }

    template <typename T, typename F>
    void stage(const T* in, T* out, size_t bytes, int nrows, int ncols, F&& op, const char* what) const
    {
        if (nrows < 0 || ncols < 0) throw DftError(std::string("HipDFT ") + what + ": negative size");
        if (bytes == 0) return;
        if (!in || !out) throw DftError(std::string("HipDFT ") + what + ": null array");
        Ctx& c = context(bytes);
        check(wct_memcpy_htod_async(c.dbuf, in, bytes, c.stream), what);
        check(op((float*)c.dbuf, c.stream), what);
        check(wct_memcpy_dtoh(out, c.dbuf, bytes, c.stream), what);
    }

    template <typename F>
    void run(const complex_t* in, complex_t* out, int nrows, int ncols, F&& op, const char* what) const
    {
        stage(in, out, (size_t)nrows * ncols * sizeof(complex_t), nrows, ncols, op, what);
    }

    int m_device{0};
    mutable std::mutex m_mutex;
    mutable std::unordered_map<std::thread::id, Ctx> m_ctx;
};

}  // namespace Hip
}  // namespace WireCell

WIRECELL_HIP_FACTORY(HipDFT, WireCell::Hip::HipDFT)
