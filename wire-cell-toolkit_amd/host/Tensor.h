// Tensor data interfaces restated with the reference's names and
// signatures (iface/inc/WireCellIface/ITensor.h:14-65, ITensorSet.h:9-20,
// ITensorSetFilter.h:12-20), the concrete Aux::SimpleTensor /
// Aux::SimpleTensorSet (aux/inc/WireCellAux/SimpleTensor.h,
// SimpleTensorSet.h) and Aux::get_tens (aux/src/TensUtil.cxx:4-18).
//
// As in the reference's tensor-set signal processing, a 2D waveform
// tensor of shape {nrows, ncols} holds its elements column by column
// (element (r, c) at r + c * nrows), the storage of the Eigen arrays it
// is mapped to (aux/inc/WireCellAux/Util.h:119-151).
#pragma once
#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <typeinfo>
#include <vector>

#include "Interface.h"

namespace WireCell {

// util/src/Dtype.cxx:17-37: "f4", "f8", "c8", "c16", "i1".."i8", "u1".."u8".
std::string dtype(const std::type_info& ti);
const std::type_info& dtype_info(const std::string& dt);
size_t dtype_size(const std::string& dt);

class ITensor {
  public:
    typedef std::shared_ptr<const ITensor> pointer;
    typedef std::vector<pointer> vector;
    typedef std::shared_ptr<const vector> shared_vector;
    typedef std::vector<size_t> shape_t;
    typedef std::vector<size_t> order_t;

    virtual ~ITensor() = default;
    virtual const std::type_info& element_type() const = 0;
    virtual size_t element_size() const = 0;
    virtual std::string dtype() const { return WireCell::dtype(element_type()); }
    virtual shape_t shape() const = 0;
    virtual order_t order() const { return order_t(); }
    virtual const std::byte* data() const = 0;
    virtual size_t size() const = 0;
    virtual Configuration metadata() const { return Configuration(); }
};

class ITensorSet {
  public:
    typedef std::shared_ptr<const ITensorSet> pointer;
    virtual ~ITensorSet() = default;
    virtual int ident() const = 0;
    virtual Configuration metadata() const { return Configuration(); }
    virtual ITensor::shared_vector tensors() const = 0;
};

class ITensorSetFilter : public IFunctionNode<ITensorSet, ITensorSet>, public IComponent<ITensorSetFilter> {
  public:
    typedef std::shared_ptr<ITensorSetFilter> pointer;
    virtual ~ITensorSetFilter() = default;
};

namespace Aux {

class SimpleTensor : public ITensor {
  public:
    // A tensor of element type dt ("f4", "c8", ...) and shape; data, when
    // given, holds at least the product of the shape elements and is
    // copied, else the store is zeroed.
    SimpleTensor(const std::string& dt, const shape_t& shape, const void* data,
                 const Configuration& md = Configuration::makeObject());
    template <typename ElementType>
    SimpleTensor(const shape_t& shape, const ElementType* data,
                 const Configuration& md = Configuration::makeObject())
        : SimpleTensor(WireCell::dtype(typeid(ElementType)), shape, data, md)
    {
    }

    std::vector<std::byte>& store() { return m_store; }
    Configuration& metadata() { return m_md; }

    const std::type_info& element_type() const override { return *m_type; }
    size_t element_size() const override { return m_size; }
    shape_t shape() const override { return m_shape; }
    const std::byte* data() const override { return m_store.data(); }
    size_t size() const override { return m_store.size(); }
    Configuration metadata() const override { return m_md; }

  private:
    const std::type_info* m_type;
    size_t m_size;
    shape_t m_shape;
    std::vector<std::byte> m_store;
    Configuration m_md;
};

class SimpleTensorSet : public ITensorSet {
  public:
    SimpleTensorSet(int ident, const Configuration& md, ITensor::shared_vector tensors)
        : m_ident(ident), m_md(md), m_tensors(std::move(tensors))
    {
    }
    int ident() const override { return m_ident; }
    Configuration metadata() const override { return m_md; }
    ITensor::shared_vector tensors() const override { return m_tensors; }

  private:
    int m_ident;
    Configuration m_md;
    ITensor::shared_vector m_tensors;
};

// The first tensor of the set whose metadata "tag" and "type" match, or
// null.
ITensor::pointer get_tens(const ITensorSet::pointer& in, const std::string& tag, const std::string& type);

}  // namespace Aux
}  // namespace WireCell
