#include "Configuration.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <sstream>

namespace WireCell {

namespace {

// Recursive-descent JSON reader.  Accepts // line comments, which the
// reference's Jsonnet-produced configuration never holds but hand-written
// test configurations sometimes do.
struct Parser {
    const std::string& s;
    size_t i{0};

    [[noreturn]] void fail(const char* what)
    {
        throw ConfigError(std::string("JSON parse error: ") + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size()) {
            if (std::isspace((unsigned char)s[i])) {
                ++i;
                continue;
            }
            if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') ++i;
                continue;
            }
            break;
        }
    }
    Configuration value()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        const char c = s[i];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Configuration(str());
        if (s.compare(i, 4, "true") == 0) { i += 4; return Configuration(true); }
        if (s.compare(i, 5, "false") == 0) { i += 5; return Configuration(false); }
        if (s.compare(i, 4, "null") == 0) { i += 4; return Configuration(); }
        return number();
    }
    std::string str()
    {
        if (i >= s.size() || s[i] != '"') fail("expected string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            const char c = s[i++];
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("bad escape");
            const char e = s[i++];
            switch (e) {
            case 'n': out += '\n'; break;
            case 't': out += '\t'; break;
            case 'r': out += '\r'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'u': {
                if (i + 4 > s.size()) fail("bad unicode escape");
                const unsigned cp = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
                i += 4;
                if (cp < 0x80) {
                    out += (char)cp;
                } else if (cp < 0x800) {
                    out += (char)(0xc0 | (cp >> 6));
                    out += (char)(0x80 | (cp & 0x3f));
                } else {
                    out += (char)(0xe0 | (cp >> 12));
                    out += (char)(0x80 | ((cp >> 6) & 0x3f));
                    out += (char)(0x80 | (cp & 0x3f));
                }
                break;
            }
            default: out += e;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    Configuration number()
    {
        const size_t b = i;
        while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' || s[i] == 'e' ||
                                s[i] == 'E' || s[i] == '-' || s[i] == '+'))
            ++i;
        if (b == i) fail("unexpected character");
        // strtod, not stod: subnormal values (field responses hold them)
        // are returned as parsed rather than rejected as out of range
        const std::string tok = s.substr(b, i - b);
        char* end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size()) fail("bad number");
        return Configuration(v);
    }
    Configuration array()
    {
        ++i;
        Configuration res = Configuration::makeArray();
        ws();
        if (i < s.size() && s[i] == ']') {
            ++i;
            return res;
        }
        while (true) {
            res.append(value());
            ws();
            if (i >= s.size()) fail("unterminated array");
            if (s[i] == ',') { ++i; continue; }
            if (s[i] == ']') { ++i; break; }
            fail("expected , or ]");
        }
        return res;
    }
    Configuration object()
    {
        ++i;
        Configuration res = Configuration::makeObject();
        ws();
        if (i < s.size() && s[i] == '}') {
            ++i;
            return res;
        }
        while (true) {
            ws();
            const std::string k = str();
            ws();
            if (i >= s.size() || s[i] != ':') fail("expected :");
            ++i;
            res[k] = value();
            ws();
            if (i >= s.size()) fail("unterminated object");
            if (s[i] == ',') { ++i; continue; }
            if (s[i] == '}') { ++i; break; }
            fail("expected , or }");
        }
        return res;
    }
};

}  // namespace

Configuration Configuration::makeArray()
{
    Configuration c;
    c.m_type = Array;
    return c;
}

Configuration Configuration::makeObject()
{
    Configuration c;
    c.m_type = Object;
    return c;
}

Configuration Configuration::parse(const std::string& text)
{
    Parser p{text};
    Configuration v = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

bool Configuration::isMember(const std::string& key) const
{
    return m_type == Object && m_obj.count(key) != 0;
}

double Configuration::asDouble() const
{
    if (m_type == Number || m_type == Bool) return m_num;
    if (m_type == Null) return 0;
    throw ConfigError("configuration value is not numeric");
}

bool Configuration::asBool() const
{
    if (m_type == Bool || m_type == Number) return m_num != 0;
    if (m_type == Null) return false;
    throw ConfigError("configuration value is not boolean");
}

std::string Configuration::asString() const
{
    if (m_type == String) return m_str;
    if (m_type == Null) return "";
    throw ConfigError("configuration value is not a string");
}

size_t Configuration::size() const
{
    if (m_type == Array) return m_arr.size();
    if (m_type == Object) return m_obj.size();
    return 0;
}

Configuration& Configuration::operator[](const std::string& key)
{
    if (m_type == Null) m_type = Object;
    if (m_type != Object) throw ConfigError("configuration value is not an object");
    return m_obj[key];
}

const Configuration& Configuration::operator[](const std::string& key) const
{
    static const Configuration null;
    if (m_type != Object) return null;
    auto it = m_obj.find(key);
    return it == m_obj.end() ? null : it->second;
}

Configuration& Configuration::operator[](int idx)
{
    if (m_type == Null) m_type = Array;
    if (m_type != Array) throw ConfigError("configuration value is not an array");
    if (idx < 0) throw ConfigError("negative index");
    if ((size_t)idx >= m_arr.size()) m_arr.resize(idx + 1);
    return m_arr[idx];
}

const Configuration& Configuration::operator[](int idx) const
{
    static const Configuration null;
    if (m_type != Array || idx < 0 || (size_t)idx >= m_arr.size()) return null;
    return m_arr[idx];
}

void Configuration::append(const Configuration& v)
{
    if (m_type == Null) m_type = Array;
    if (m_type != Array) throw ConfigError("configuration value is not an array");
    m_arr.push_back(v);
}

std::vector<std::string> Configuration::getMemberNames() const
{
    std::vector<std::string> out;
    if (m_type == Object)
        for (auto& kv : m_obj) out.push_back(kv.first);
    return out;
}

static void dump_str(std::ostringstream& os, const std::string& s)
{
    os << '"';
    for (char c : s) {
        if (c == '"' || c == '\\') os << '\\' << c;
        else if (c == '\n') os << "\\n";
        else if (c == '\t') os << "\\t";
        else os << c;
    }
    os << '"';
}

static void dump_to(std::ostringstream& os, const Configuration& v)
{
    switch (v.type()) {
    case Configuration::Null: os << "null"; break;
    case Configuration::Bool: os << (v.asBool() ? "true" : "false"); break;
    case Configuration::Number: {
        const double d = v.asDouble();
        if (std::isfinite(d)) {
            char buf[40];
            snprintf(buf, sizeof(buf), "%.17g", d);
            os << buf;
        } else {
            os << "null";
        }
        break;
    }
    case Configuration::String: dump_str(os, v.asString()); break;
    case Configuration::Array: {
        os << '[';
        bool first = true;
        for (auto& e : v) {
            if (!first) os << ',';
            first = false;
            dump_to(os, e);
        }
        os << ']';
        break;
    }
    case Configuration::Object: {
        os << '{';
        bool first = true;
        for (auto& k : v.getMemberNames()) {
            if (!first) os << ',';
            first = false;
            dump_str(os, k);
            os << ':';
            dump_to(os, v[k]);
        }
        os << '}';
        break;
    }
    }
}

std::string Configuration::dump() const
{
    std::ostringstream os;
    dump_to(os, *this);
    return os.str();
}

Configuration update(Configuration a, const Configuration& b)
{
    if (b.isNull()) return a;
    if (!b.isObject() || !a.isObject()) return b;
    for (auto& k : b.getMemberNames()) {
        if (a[k].isObject() && b[k].isObject()) a[k] = update(a[k], b[k]);
        else a[k] = b[k];
    }
    return a;
}

}  // namespace WireCell
