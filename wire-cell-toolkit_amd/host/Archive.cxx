#include "Archive.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <sstream>

namespace WireCell {
namespace Hip {

namespace {

bool ends_with(const std::string& s, const std::string& e)
{
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

bool gzip_name(const std::string& path) { return ends_with(path, ".tar.gz") || ends_with(path, ".tgz"); }

size_t descr_size(const std::string& d)
{
    if (d.size() < 3) throw IOError("npy: bad dtype \"" + d + "\"");
    if (d[0] == '>') throw IOError("npy: big-endian dtype \"" + d + "\" is not supported");
    return (size_t)std::stoul(d.substr(2));
}

// The value of key k in the header dict (text up to the next top-level
// ',' or '}'; a tuple value is kept whole).
std::string dict_value(const std::string& h, const std::string& k)
{
    const auto p = h.find("'" + k + "'");
    if (p == std::string::npos) throw IOError("npy: header without '" + k + "'");
    size_t i = h.find(':', p);
    if (i == std::string::npos) throw IOError("npy: bad header");
    ++i;
    while (i < h.size() && h[i] == ' ') ++i;
    size_t j = i;
    if (h[i] == '(') {
        j = h.find(')', i);
        if (j == std::string::npos) throw IOError("npy: bad shape");
        return h.substr(i, j - i + 1);
    }
    while (j < h.size() && h[j] != ',' && h[j] != '}') ++j;
    std::string v = h.substr(i, j - i);
    while (!v.empty() && v.back() == ' ') v.pop_back();
    return v;
}

}  // namespace

size_t NpyArray::count() const
{
    size_t n = 1;
    for (size_t s : shape) n *= s;
    return n;
}

size_t NpyArray::item_size() const { return descr_size(descr); }

double NpyArray::at(size_t r, size_t c) const
{
    size_t idx = r;
    if (shape.size() == 2) idx = fortran_order ? r + c * shape[0] : r * shape[1] + c;
    const size_t sz = item_size();
    const unsigned char* p = data.data() + idx * sz;
    const char kind = descr[1];
    if (kind == 'f' && sz == 4) { float v; std::memcpy(&v, p, 4); return v; }
    if (kind == 'f' && sz == 8) { double v; std::memcpy(&v, p, 8); return v; }
    if (kind == 'i' && sz == 1) { int8_t v; std::memcpy(&v, p, 1); return v; }
    if (kind == 'i' && sz == 2) { int16_t v; std::memcpy(&v, p, 2); return v; }
    if (kind == 'i' && sz == 4) { int32_t v; std::memcpy(&v, p, 4); return v; }
    if (kind == 'i' && sz == 8) { int64_t v; std::memcpy(&v, p, 8); return (double)v; }
    if (kind == 'u' && sz == 1) { uint8_t v; std::memcpy(&v, p, 1); return v; }
    if (kind == 'u' && sz == 2) { uint16_t v; std::memcpy(&v, p, 2); return v; }
    if (kind == 'u' && sz == 4) { uint32_t v; std::memcpy(&v, p, 4); return v; }
    if (kind == 'u' && sz == 8) { uint64_t v; std::memcpy(&v, p, 8); return (double)v; }
    throw IOError("npy: unsupported dtype \"" + descr + "\"");
}

std::vector<unsigned char> npy_encode(const NpyArray& a)
{
    std::ostringstream h;
    h << "{'descr': '" << a.descr << "', 'fortran_order': " << (a.fortran_order ? "True" : "False")
      << ", 'shape': (";
    for (size_t i = 0; i < a.shape.size(); ++i) h << a.shape[i] << (a.shape.size() == 1 || i + 1 < a.shape.size() ? ", " : "");
    h << "), }";
    std::string hs = h.str();
    // magic (6) + version (2) + length (2) + header, padded with spaces
    // and a newline to a multiple of 64 bytes
    const size_t total = ((10 + hs.size() + 1 + 63) / 64) * 64;
    hs.append(total - 10 - hs.size() - 1, ' ');
    hs.push_back('\n');
    if (a.data.size() != a.count() * a.item_size()) throw IOError("npy: data size does not match the shape");
    std::vector<unsigned char> out(total + a.data.size());
    const unsigned char lead[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(hs.size() & 0xff),
                                    (unsigned char)(hs.size() >> 8)};
    std::memcpy(out.data(), lead, 10);
    std::memcpy(out.data() + 10, hs.data(), hs.size());
    if (!a.data.empty()) std::memcpy(out.data() + total, a.data.data(), a.data.size());
    return out;
}

NpyArray npy_decode(const std::vector<unsigned char>& b)
{
    if (b.size() < 10 || b[0] != 0x93 || std::memcmp(&b[1], "NUMPY", 5) != 0) throw IOError("npy: bad magic");
    size_t hlen = 0, hoff = 0;
    if (b[6] == 1) {
        hlen = b[8] | (b[9] << 8);
        hoff = 10;
    } else if (b[6] == 2 || b[6] == 3) {
        if (b.size() < 12) throw IOError("npy: truncated header");
        hlen = (size_t)b[8] | ((size_t)b[9] << 8) | ((size_t)b[10] << 16) | ((size_t)b[11] << 24);
        hoff = 12;
    } else {
        throw IOError("npy: unsupported format version");
    }
    if (b.size() < hoff + hlen) throw IOError("npy: truncated header");
    const std::string h(b.begin() + hoff, b.begin() + hoff + hlen);
    NpyArray a;
    std::string d = dict_value(h, "descr");
    if (d.size() < 2 || (d.front() != '\'' && d.front() != '"')) throw IOError("npy: bad descr");
    a.descr = d.substr(1, d.size() - 2);
    if (a.descr.size() == 2) a.descr = "|" + a.descr;   // e.g. 'u1' written without a byte order
    const std::string fo = dict_value(h, "fortran_order");
    if (fo != "True" && fo != "False") throw IOError("npy: bad fortran_order");
    a.fortran_order = fo == "True";
    const std::string sh = dict_value(h, "shape");
    for (size_t i = 1; i < sh.size();) {
        while (i < sh.size() && (sh[i] == ' ' || sh[i] == ',')) ++i;
        if (i >= sh.size() || sh[i] == ')') break;
        size_t j = i;
        while (j < sh.size() && std::isdigit((unsigned char)sh[j])) ++j;
        if (j == i) throw IOError("npy: bad shape " + sh);
        a.shape.push_back((size_t)std::stoull(sh.substr(i, j - i)));
        i = j;
    }
    const size_t nbytes = a.count() * a.item_size();
    if (b.size() < hoff + hlen + nbytes) throw IOError("npy: truncated data");
    a.data.assign(b.begin() + hoff + hlen, b.begin() + hoff + hlen + nbytes);
    return a;
}

// ---- tar ----------------------------------------------------------------

struct TarWriter::Impl {
    FILE* f{nullptr};
    gzFile gz{nullptr};
    bool closed{false};
};

TarWriter::TarWriter(const std::string& path) : m(new Impl)
{
    if (!supported(path)) throw IOError("tar: unsupported output name " + path + " (.tar, .tar.gz or .tgz)");
    if (gzip_name(path))
        m->gz = gzopen(path.c_str(), "wb");
    else
        m->f = std::fopen(path.c_str(), "wb");
    if (!m->gz && !m->f) throw IOError("tar: cannot open " + path + " for writing");
}

TarWriter::~TarWriter()
{
    try {
        close();
    } catch (...) {
    }
}

bool TarWriter::supported(const std::string& path)
{
    return ends_with(path, ".tar") || gzip_name(path);
}

void TarWriter::put(const void* p, size_t n)
{
    if (!n) return;
    const bool ok = m->gz ? gzwrite(m->gz, p, (unsigned)n) == (int)n : std::fwrite(p, 1, n, m->f) == n;
    if (!ok) throw IOError("tar: write failed");
}

void TarWriter::add(const std::string& name, const std::vector<unsigned char>& bytes)
{
    if (m->closed) throw IOError("tar: archive already closed");
    if (name.size() > 99) throw IOError("tar: member name longer than 99 characters: " + name);
    unsigned char h[512];
    std::memset(h, 0, sizeof(h));
    std::memcpy(h, name.data(), name.size());
    std::snprintf((char*)h + 100, 8, "%07o", 0644);
    std::snprintf((char*)h + 108, 8, "%07o", 0);
    std::snprintf((char*)h + 116, 8, "%07o", 0);
    std::snprintf((char*)h + 124, 12, "%011llo", (unsigned long long)bytes.size());
    std::snprintf((char*)h + 136, 12, "%011llo", 0ull);
    h[156] = '0';
    std::memcpy(h + 257, "ustar", 6);   // POSIX magic "ustar\0", version "00"
    std::memcpy(h + 263, "00", 2);
    std::memset(h + 148, ' ', 8);
    unsigned sum = 0;
    for (unsigned char c : h) sum += c;
    std::snprintf((char*)h + 148, 8, "%06o", sum);
    h[155] = ' ';
    put(h, 512);
    put(bytes.data(), bytes.size());
    static const unsigned char zero[512] = {0};
    put(zero, (512 - bytes.size() % 512) % 512);
}

void TarWriter::flush()
{
    if (m->gz) gzflush(m->gz, Z_SYNC_FLUSH);
    if (m->f) std::fflush(m->f);
}

void TarWriter::close()
{
    if (m->closed) return;
    m->closed = true;
    static const unsigned char zero[1024] = {0};
    put(zero, 1024);
    if (m->gz) gzclose(m->gz);
    if (m->f) std::fclose(m->f);
    m->gz = nullptr;
    m->f = nullptr;
}

struct TarReader::Impl {
    gzFile gz{nullptr};   // reads plain files as they are
};

TarReader::TarReader(const std::string& path) : m(new Impl)
{
    if (!supported(path)) throw IOError("tar: unsupported input name " + path + " (.tar, .tar.gz or .tgz)");
    m->gz = gzopen(path.c_str(), "rb");
    if (!m->gz) throw IOError("tar: cannot open " + path);
}

TarReader::~TarReader()
{
    if (m->gz) gzclose(m->gz);
}

bool TarReader::supported(const std::string& path) { return TarWriter::supported(path); }

bool TarReader::get(void* p, size_t n)
{
    if (!n) return true;
    return gzread(m->gz, p, (unsigned)n) == (int)n;
}

bool TarReader::next(std::string& name, std::vector<unsigned char>& bytes)
{
    std::string longname;
    while (true) {
        unsigned char h[512];
        if (!get(h, 512)) return false;   // end of file without the zero blocks
        bool zero = true;
        for (unsigned char c : h) zero = zero && c == 0;
        if (zero) return false;
        unsigned sum = 0;
        for (int i = 0; i < 512; ++i) sum += (i >= 148 && i < 156) ? ' ' : h[i];
        const unsigned stored = (unsigned)std::strtoul(std::string((char*)h + 148, 8).c_str(), nullptr, 8);
        if (sum != stored) throw IOError("tar: header checksum mismatch");
        size_t size = 0;
        if (h[124] & 0x80) {   // base-256 size
            for (int i = 125; i < 136; ++i) size = (size << 8) | h[i];
        } else {
            size = (size_t)std::strtoull(std::string((char*)h + 124, 12).c_str(), nullptr, 8);
        }
        std::vector<unsigned char> data(size);
        if (!get(data.data(), size)) throw IOError("tar: truncated member");
        unsigned char pad[512];
        if (!get(pad, (512 - size % 512) % 512)) throw IOError("tar: truncated member");
        const char type = (char)h[156];
        if (type == 'L') {   // GNU long name of the next member
            longname.assign(data.begin(), data.end());
            while (!longname.empty() && longname.back() == 0) longname.pop_back();
            continue;
        }
        if (type != '0' && type != 0) continue;   // directories, links, pax headers
        if (!longname.empty()) {
            name = longname;
        } else {
            name.assign((char*)h, strnlen((char*)h, 100));
            if (std::memcmp(h + 257, "ustar", 5) == 0 && h[345]) {
                const std::string prefix((char*)h + 345, strnlen((char*)h + 345, 155));
                name = prefix + "/" + name;
            }
        }
        bytes.swap(data);
        return true;
    }
}

}  // namespace Hip
}  // namespace WireCell
