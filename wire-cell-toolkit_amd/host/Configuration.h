// A small JSON value for component configuration.  The reference uses
// jsoncpp's Json::Value as WireCell::Configuration
// (util/inc/WireCellUtil/Configuration.h:145-169); this restates the part
// of that API the deconvolution components use: parsing, object/array
// access, isMember/isNull, as*() conversions and get(cfg, key, default).
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace WireCell {

class Configuration {
  public:
    enum Type { Null, Bool, Number, String, Array, Object };

    Configuration() = default;
    Configuration(bool b) : m_type(Bool), m_num(b ? 1 : 0) {}
    Configuration(int v) : m_type(Number), m_num(v) {}
    Configuration(double v) : m_type(Number), m_num(v) {}
    Configuration(const char* s) : m_type(String), m_str(s) {}
    Configuration(const std::string& s) : m_type(String), m_str(s) {}

    static Configuration parse(const std::string& text);
    static Configuration makeArray();
    static Configuration makeObject();
    std::string dump() const;

    Type type() const { return m_type; }
    bool isNull() const { return m_type == Null; }
    bool isArray() const { return m_type == Array; }
    bool isObject() const { return m_type == Object; }
    bool isString() const { return m_type == String; }
    bool isNumeric() const { return m_type == Number || m_type == Bool; }
    bool isMember(const std::string& key) const;

    double asDouble() const;
    int asInt() const { return (int)asDouble(); }
    bool asBool() const;
    std::string asString() const;

    size_t size() const;
    // Object access (creates on write, returns null on missing read).
    Configuration& operator[](const std::string& key);
    const Configuration& operator[](const std::string& key) const;
    Configuration& operator[](const char* key) { return (*this)[std::string(key)]; }
    const Configuration& operator[](const char* key) const { return (*this)[std::string(key)]; }
    // Array access.
    Configuration& operator[](int idx);
    const Configuration& operator[](int idx) const;
    void append(const Configuration& v);
    std::vector<std::string> getMemberNames() const;

    // Iteration over array elements.
    std::vector<Configuration>::const_iterator begin() const { return m_arr.begin(); }
    std::vector<Configuration>::const_iterator end() const { return m_arr.end(); }

  private:
    Type m_type{Null};
    double m_num{0};
    std::string m_str;
    std::vector<Configuration> m_arr;
    std::map<std::string, Configuration> m_obj;
};

// Merge: keys of b override those of a (objects merged recursively).
Configuration update(Configuration a, const Configuration& b);

template <typename T>
T convert(const Configuration& v);
template <> inline double convert<double>(const Configuration& v) { return v.asDouble(); }
template <> inline float convert<float>(const Configuration& v) { return (float)v.asDouble(); }
template <> inline int convert<int>(const Configuration& v) { return v.asInt(); }
template <> inline bool convert<bool>(const Configuration& v) { return v.asBool(); }
template <> inline std::string convert<std::string>(const Configuration& v) { return v.asString(); }

// WireCell::get(cfg, key, default) as in the reference.
template <typename T>
T get(const Configuration& cfg, const std::string& key, const T& def = T())
{
    if (!cfg.isObject() || !cfg.isMember(key)) return def;
    const Configuration& v = cfg[key];
    if (v.isNull()) return def;
    return convert<T>(v);
}
inline std::string get(const Configuration& cfg, const std::string& key, const char* def)
{
    return get<std::string>(cfg, key, std::string(def));
}

struct ConfigError : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace WireCell
