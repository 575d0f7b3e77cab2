// Component factory: the named-instance registry of
// util/inc/WireCellUtil/NamedFactory.h restated.  A component type
// registers a maker under its type name; instances are created on first
// lookup and are unique per "type:name" (an empty name is the default
// instance), as Factory::lookup / find_tn do in the reference.  Each type
// also exports the reference's plugin entry point
//   extern "C" void* make_<TYPE>_factory()
// (NamedFactory.h:326-349), which PluginManager-style loaders resolve
// with dlsym after dlopen("lib<Plugin>.so").
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "Configuration.h"
#include "Interface.h"

namespace WireCell {

struct FactoryException : public std::runtime_error {
    using std::runtime_error::runtime_error;
};

class NamedFactoryBase {
  public:
    explicit NamedFactoryBase(std::string type) : m_type(std::move(type)) {}
    virtual ~NamedFactoryBase() = default;
    const std::string& type_name() const { return m_type; }
    virtual std::shared_ptr<Interface> create() const = 0;

  private:
    std::string m_type;
};

template <class Concrete>
class NamedFactory : public NamedFactoryBase {
  public:
    using NamedFactoryBase::NamedFactoryBase;
    std::shared_ptr<Interface> create() const override { return std::make_shared<Concrete>(); }
};

namespace Factory {

void register_type(NamedFactoryBase* factory);
std::vector<std::string> known_types();

// The instance "type:name", created on first use.  Throws
// FactoryException for an unknown type.
std::shared_ptr<Interface> lookup_instance(const std::string& type, const std::string& name,
                                           bool create = true);

// Split "type:name" (name may be empty).
std::pair<std::string, std::string> parse_tn(const std::string& tn);

template <class IFACE>
std::shared_ptr<IFACE> find(const std::string& type, const std::string& name = "")
{
    auto base = lookup_instance(type, name, true);
    auto ret = std::dynamic_pointer_cast<IFACE>(base);
    if (!ret)
        throw FactoryException("component " + type + ":" + name + " does not provide the requested interface");
    return ret;
}

template <class IFACE>
std::shared_ptr<IFACE> find_tn(const std::string& tn)
{
    auto p = parse_tn(tn);
    return find<IFACE>(p.first, p.second);
}

// The application's configuration step (apps/src/Main.cxx:330-350):
// for each {"type", "name", "data"} entry, in order, create the instance
// and configure it with default_configuration() updated by "data".
void configure_all(const Configuration& cfg_list);

// Drop every instance (types stay registered).
void clear_instances();

}  // namespace Factory

template <class Concrete>
struct FactoryRegistrar {
    explicit FactoryRegistrar(NamedFactoryBase* f) { Factory::register_type(f); }
};

}  // namespace WireCell

// Defines the plugin entry point and registers the type when the library
// is loaded.
#define WIRECELL_HIP_FACTORY(NAME, CONCRETE)                                               \
    extern "C" void* make_##NAME##_factory()                                               \
    {                                                                                      \
        static WireCell::NamedFactory<CONCRETE> nf(#NAME);                                 \
        static WireCell::FactoryRegistrar<CONCRETE> reg(&nf);                              \
        return &nf;                                                                        \
    }                                                                                      \
    namespace {                                                                            \
    [[maybe_unused]] void* const wct_hip_factory_init_##NAME = make_##NAME##_factory();   \
    }
