#include "Factory.h"

namespace WireCell {
namespace Factory {

namespace {
struct Registry {
    std::mutex mutex;
    std::map<std::string, NamedFactoryBase*> types;
    std::map<std::string, std::shared_ptr<Interface>> instances;   // "type:name"
};
Registry& registry()
{
    static Registry* r = new Registry();   // outlives static destructors
    return *r;
}
}  // namespace

void register_type(NamedFactoryBase* factory)
{
    auto& r = registry();
    std::lock_guard<std::mutex> lk(r.mutex);
    r.types[factory->type_name()] = factory;
}

std::vector<std::string> known_types()
{
    auto& r = registry();
    std::lock_guard<std::mutex> lk(r.mutex);
    std::vector<std::string> out;
    for (auto& kv : r.types) out.push_back(kv.first);
    return out;
}

std::pair<std::string, std::string> parse_tn(const std::string& tn)
{
    const auto pos = tn.find(':');
    if (pos == std::string::npos) return {tn, ""};
    return {tn.substr(0, pos), tn.substr(pos + 1)};
}

std::shared_ptr<Interface> lookup_instance(const std::string& type, const std::string& name, bool create)
{
    auto& r = registry();
    std::lock_guard<std::mutex> lk(r.mutex);
    const std::string key = type + ":" + name;
    auto it = r.instances.find(key);
    if (it != r.instances.end()) return it->second;
    if (!create) throw FactoryException("no instance " + key);
    auto ft = r.types.find(type);
    if (ft == r.types.end()) throw FactoryException("unknown component type " + type);
    auto inst = ft->second->create();
    r.instances[key] = inst;
    return inst;
}

void configure_all(const Configuration& cfg_list)
{
    if (!cfg_list.isArray()) throw ConfigError("configuration must be a list of components");
    for (const auto& one : cfg_list) {
        const std::string type = get<std::string>(one, "type", "");
        if (type.empty()) throw ConfigError("configuration entry without a type");
        const std::string name = get<std::string>(one, "name", "");
        auto inst = lookup_instance(type, name, true);
        auto cfgable = std::dynamic_pointer_cast<IConfigurable>(inst);
        if (!cfgable) continue;
        Configuration cfg = update(cfgable->default_configuration(), one["data"]);
        cfgable->configure(cfg);
    }
}

void clear_instances()
{
    auto& r = registry();
    std::lock_guard<std::mutex> lk(r.mutex);
    r.instances.clear();
}

}  // namespace Factory
}  // namespace WireCell
