#include "Tensor.h"

#include <cstring>
#include <stdexcept>

namespace WireCell {

namespace {
struct DtypeEntry {
    const char* name;
    const std::type_info* type;
    size_t size;
};
const DtypeEntry kDtypes[] = {
    {"i1", &typeid(int8_t), 1},   {"i2", &typeid(int16_t), 2},  {"i4", &typeid(int32_t), 4},
    {"i8", &typeid(int64_t), 8},  {"u1", &typeid(uint8_t), 1},  {"u2", &typeid(uint16_t), 2},
    {"u4", &typeid(uint32_t), 4}, {"u8", &typeid(uint64_t), 8}, {"f4", &typeid(float), 4},
    {"f8", &typeid(double), 8},   {"c8", &typeid(std::complex<float>), 8},
    {"c16", &typeid(std::complex<double>), 16},
};
}  // namespace

std::string dtype(const std::type_info& ti)
{
    for (const auto& e : kDtypes)
        if (*e.type == ti) return e.name;
    return "";
}

const std::type_info& dtype_info(const std::string& dt)
{
    for (const auto& e : kDtypes)
        if (dt == e.name) return *e.type;
    return typeid(void);
}

size_t dtype_size(const std::string& dt)
{
    for (const auto& e : kDtypes)
        if (dt == e.name) return e.size;
    return 0;
}

namespace Aux {

SimpleTensor::SimpleTensor(const std::string& dt, const shape_t& shape, const void* data, const Configuration& md)
    : m_type(&dtype_info(dt)), m_size(dtype_size(dt)), m_shape(shape), m_md(md)
{
    if (!m_size) throw std::invalid_argument("SimpleTensor: unknown dtype \"" + dt + "\"");
    size_t nbytes = m_size;
    for (size_t s : m_shape) nbytes *= s;
    m_store.resize(nbytes);
    if (data && nbytes) std::memcpy(m_store.data(), data, nbytes);
}

ITensor::pointer get_tens(const ITensorSet::pointer& in, const std::string& tag, const std::string& type)
{
    if (!in) return nullptr;
    auto tens = in->tensors();
    if (!tens) return nullptr;
    for (const auto& t : *tens) {
        if (!t) continue;
        const Configuration md = t->metadata();
        if (get<std::string>(md, "tag", "") == tag && get<std::string>(md, "type", "") == type) return t;
    }
    return nullptr;
}

}  // namespace Aux
}  // namespace WireCell
