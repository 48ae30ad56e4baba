// Host-side mirror of the reference types the operator path touches.
//
// These re-state, in a self-contained form, the reference interfaces a caller of the
// matrix-free path uses (paths relative to the reference root):
//   Array<OneD, T>          LibUtilities/BasicUtils/SharedArray.hpp (shared array)
//   NekError / NEKERROR     LibUtilities/BasicUtils/ErrorUtil.hpp:59, 150-171
//   NekFactory              LibUtilities/BasicUtils/NekFactory.hpp:101-240
//   ShapeType               LibUtilities/BasicUtils/ShapeType.hpp
//   ConstFactorMap          StdRegions/StdRegions.hpp (eFactorLambda)
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace Nektar {

using NekDouble = double;

struct OneD {};

// Shared 1D array with value semantics of the handle (copies share storage).
template <typename Dim, typename T>
class Array;

template <typename T>
class Array<OneD, T> {
public:
    Array() = default;
    explicit Array(size_t n) : m_data(std::make_shared<std::vector<T>>(n)), m_size(n) {}
    Array(size_t n, const T &v) : m_data(std::make_shared<std::vector<T>>(n, v)), m_size(n) {}
    Array(size_t n, const T *src)
        : m_data(std::make_shared<std::vector<T>>(src, src + n)), m_size(n) {}
    // Offset view sharing storage (Array(offset, other) in the reference is
    // `other + offset`).
    Array(const Array &a, size_t offset)
        : m_data(a.m_data), m_offset(a.m_offset + offset), m_size(a.m_size - offset) {}
    size_t size() const { return m_size; }
    size_t num_elements() const { return m_size; }
    T *data() { return m_data ? m_data->data() + m_offset : nullptr; }
    const T *data() const { return m_data ? m_data->data() + m_offset : nullptr; }
    T &operator[](size_t i) { return (*m_data)[m_offset + i]; }
    const T &operator[](size_t i) const { return (*m_data)[m_offset + i]; }
    operator Array<OneD, const T>() const
    {
        Array<OneD, const T> r;
        r.m_data = std::const_pointer_cast<const std::vector<T>>(m_data);
        r.m_offset = m_offset;
        r.m_size = m_size;
        return r;
    }
    template <typename D, typename U>
    friend class Array;

private:
    std::shared_ptr<std::vector<T>> m_data;
    size_t m_offset = 0, m_size = 0;
};

template <typename T>
class Array<OneD, const T> {
public:
    Array() = default;
    size_t size() const { return m_size; }
    size_t num_elements() const { return m_size; }
    const T *data() const { return m_data ? m_data->data() + m_offset : nullptr; }
    const T &operator[](size_t i) const { return (*m_data)[m_offset + i]; }
    template <typename D, typename U>
    friend class Array;

private:
    std::shared_ptr<const std::vector<T>> m_data;
    size_t m_offset = 0, m_size = 0;
};

static Array<OneD, NekDouble> NullNekDouble1DArray;

// Error handling: fatal errors throw, as NEKERROR(efatal) does in the reference.
class NekError : public std::runtime_error {
public:
    explicit NekError(const std::string &m) : std::runtime_error(m) {}
};

namespace ErrorUtil {
enum ErrType { efatal, ewarning };
}

#define NEKERROR(type, msg)                                                        \
    do {                                                                           \
        if ((type) == ::Nektar::ErrorUtil::efatal) throw ::Nektar::NekError(msg);  \
    } while (0)
#define ASSERTL0(cond, msg)                                                        \
    do {                                                                           \
        if (!(cond)) throw ::Nektar::NekError(std::string("Level 0 assertion violation\n") + (msg)); \
    } while (0)

namespace LibUtilities {

enum ShapeType {
    eNoShapeType,
    ePoint,
    eSegment,
    eTriangle,
    eQuadrilateral,
    eTetrahedron,
    ePyramid,
    ePrism,
    eHexahedron,
    SIZE_ShapeType
};

const char *const ShapeTypeMap[] = {"NoGeomShapeType", "Point", "Segment", "Triangle",
                                    "Quadrilateral", "Tetrahedron", "Pyramid", "Prism",
                                    "Hexahedron"};

// Factory keyed on tKey creating shared pointers to tBase from tArgs
// (NekFactory.hpp:101-240: RegisterCreatorFunction / CreateInstance / ModuleExists).
template <typename tKey, typename tBase, typename... tArgs>
class NekFactory {
public:
    using tBaseSharedPtr = std::shared_ptr<tBase>;
    using CreatorFunction = std::function<tBaseSharedPtr(tArgs...)>;

    tKey RegisterCreatorFunction(tKey key, CreatorFunction fn, std::string desc = "")
    {
        m_map[key] = {fn, desc};
        return key;
    }
    bool ModuleExists(tKey key) const { return m_map.count(key) > 0; }
    tBaseSharedPtr CreateInstance(tKey key, tArgs... args) const
    {
        auto it = m_map.find(key);
        if (it == m_map.end()) {
            std::stringstream ss;
            ss << "No such module: " << key;
            throw NekError(ss.str());
        }
        return it->second.first(args...);
    }
    std::vector<tKey> Keys() const
    {
        std::vector<tKey> k;
        for (auto &p : m_map) k.push_back(p.first);
        return k;
    }

private:
    std::map<tKey, std::pair<CreatorFunction, std::string>> m_map;
};

}  // namespace LibUtilities

namespace StdRegions {
enum ConstFactorType { eFactorLambda, eFactorTau, eFactorTime, SIZE_ConstFactorType };
using ConstFactorMap = std::map<ConstFactorType, NekDouble>;
static ConstFactorMap NullConstFactorMap;
}  // namespace StdRegions

}  // namespace Nektar
