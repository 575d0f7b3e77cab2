// The containers of Wire-Cell's frame files: NumPy .npy arrays
// (format 1.0: magic, little-endian header length, a Python-literal
// dict with descr / fortran_order / shape, then the data) inside a tar
// stream (512-byte ustar headers, data padded to 512, two zero blocks at
// the end), optionally gzip-compressed, as util/inc/WireCellUtil/
// Stream.h writes and reads them through its custard and pigenc
// helpers (.tar and .tar.gz here; .tar.bz2 needs a bzip2 library).
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace WireCell {
namespace Hip {

struct IOError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

// One .npy array.  data holds the raw element bytes in the stored order
// (column by column when fortran_order).
struct NpyArray {
    std::string descr;               // e.g. "<f4", "<i2", "<i4", "<f8"
    bool fortran_order{false};
    std::vector<size_t> shape;
    std::vector<unsigned char> data;

    size_t count() const;
    size_t item_size() const;
    // Element (r, c) of a 2D array (or element r of a 1D one, c = 0)
    // converted to double.
    double at(size_t r, size_t c = 0) const;
};

std::vector<unsigned char> npy_encode(const NpyArray& a);
NpyArray npy_decode(const std::vector<unsigned char>& bytes);   // throws IOError

// Sequential writer and reader of a tar (or gzip-compressed tar) file.
class TarWriter {
  public:
    explicit TarWriter(const std::string& path);   // ".tar" or ".tar.gz"/".tgz"; throws IOError
    ~TarWriter();
    void add(const std::string& name, const std::vector<unsigned char>& bytes);
    void flush();
    void close();   // writes the end-of-archive blocks
    static bool supported(const std::string& path);

  private:
    void put(const void* p, size_t n);
    struct Impl;
    std::unique_ptr<Impl> m;
};

class TarReader {
  public:
    explicit TarReader(const std::string& path);   // throws IOError
    ~TarReader();
    // The next regular file of the archive; false at the end.
    bool next(std::string& name, std::vector<unsigned char>& bytes);
    static bool supported(const std::string& path);

  private:
    bool get(void* p, size_t n);
    struct Impl;
    std::unique_ptr<Impl> m;
};

}  // namespace Hip
}  // namespace WireCell
