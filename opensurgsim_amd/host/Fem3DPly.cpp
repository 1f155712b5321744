// Fem3D PLY reading (ascii, binary_little_endian, binary_big_endian) and Fem3DRepresentation::loadFem / setFem.
//
// Semantics follow Physics/FemPlyReaderDelegate.cpp:31-145 and Physics/Fem3DPlyReaderDelegate.cpp:52-96:
//   - required: element "vertex" with x, y, z; element "3d_element" with list property vertex_indices;
//   - optional: vertex thetaX/thetaY/thetaZ (ignored for Fem3D), "boundary_condition" with vertex_index,
//     a single "material" row (mass_density, poisson_ratio, young_modulus) applied to every element, or the
//     same three properties per 3d_element; the "material" element wins when both are present; no material
//     leaves zeros;
//   - other elements (e.g. "face") are read and skipped.
// The file formats and property types are those of the PLY library the reference reads through
// (DataStructures/ply.c:61-72, 760-762): char, short, int, uchar, ushort, uint, float, double (and their sized
// names int8 ... float64); binary data in either byte order.
#include "SurgSimGpu.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

namespace SurgSimGpu {
namespace Physics {

namespace {

[[noreturn]] void fail(const std::string& m) { throw Framework::AssertionFailure(m); }

enum PlyType { kNoType = 0, kChar, kShort, kInt, kUChar, kUShort, kUInt, kFloat, kDouble };

PlyType plyType(const std::string& t) {
    if (t == "char" || t == "int8") return kChar;
    if (t == "short" || t == "int16") return kShort;
    if (t == "int" || t == "int32") return kInt;
    if (t == "uchar" || t == "uint8") return kUChar;
    if (t == "ushort" || t == "uint16") return kUShort;
    if (t == "uint" || t == "uint32") return kUInt;
    if (t == "float" || t == "float32") return kFloat;
    if (t == "double" || t == "float64") return kDouble;
    return kNoType;
}

size_t typeSize(PlyType t) {
    switch (t) {
        case kChar: case kUChar: return 1;
        case kShort: case kUShort: return 2;
        case kInt: case kUInt: case kFloat: return 4;
        case kDouble: return 8;
        default: return 0;
    }
}

struct PlyProperty {
    std::string name;
    bool list = false;
    PlyType type = kNoType;       // scalar type, or the list items' type
    PlyType countType = kNoType;  // list length type
};
struct PlyElement {
    std::string name;
    size_t count = 0;
    std::vector<PlyProperty> props;
    int find(const std::string& n) const {
        for (size_t i = 0; i < props.size(); ++i)
            if (props[i].name == n) return (int)i;
        return -1;
    }
};

// Values of the data section, row by row: ascii rows are lines of whitespace-separated numbers; binary rows are
// the properties' values back to back in the file's byte order.
class PlyValues {
public:
    PlyValues(std::ifstream& in, int format, const std::string& file) : m_in(in), m_format(format), m_file(file) {}
    void beginRow(const std::string& element) {
        if (m_format != 0) return;
        std::string line;
        if (!std::getline(m_in, line)) fail("'" + m_file + "': unexpected end of data in element " + element);
        m_line.clear();
        m_line.str(line);
    }
    double read(PlyType t, const std::string& element) {
        if (m_format == 0) {
            double v;
            if (!(m_line >> v)) fail("'" + m_file + "': bad value in element " + element);
            return v;
        }
        unsigned char b[8];
        const size_t n = typeSize(t);
        if (!m_in.read(reinterpret_cast<char*>(b), (std::streamsize)n))
            fail("'" + m_file + "': unexpected end of data in element " + element);
        const bool big = m_format == 2;
        const uint16_t probe = 1;
        const bool hostBig = *reinterpret_cast<const unsigned char*>(&probe) == 0;
        if (big != hostBig)
            for (size_t i = 0; i < n / 2; ++i) std::swap(b[i], b[n - 1 - i]);
        switch (t) {
            case kChar: { int8_t v; std::memcpy(&v, b, 1); return v; }
            case kUChar: { uint8_t v; std::memcpy(&v, b, 1); return v; }
            case kShort: { int16_t v; std::memcpy(&v, b, 2); return v; }
            case kUShort: { uint16_t v; std::memcpy(&v, b, 2); return v; }
            case kInt: { int32_t v; std::memcpy(&v, b, 4); return v; }
            case kUInt: { uint32_t v; std::memcpy(&v, b, 4); return v; }
            case kFloat: { float v; std::memcpy(&v, b, 4); return v; }
            case kDouble: { double v; std::memcpy(&v, b, 8); return v; }
            default: fail("'" + m_file + "': property of unknown type in element " + element);
        }
    }

private:
    std::ifstream& m_in;
    int m_format;  // 0 ascii, 1 binary little endian, 2 binary big endian
    std::string m_file;
    std::istringstream m_line;
};

}  // namespace

void Fem3D::load(const std::string& fileName) {
    std::ifstream in(fileName, std::ios::binary);
    if (!in) fail("Could not open '" + fileName + "'");
    std::string line;
    auto next_line = [&]() {
        if (!std::getline(in, line)) return false;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return true;
    };
    next_line();
    if (line.rfind("ply", 0) != 0) fail("'" + fileName + "' is not a ply file");
    std::vector<PlyElement> elems;
    int format = -1;
    while (next_line()) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "format") {
            std::string f;
            ls >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) fail("'" + fileName + "': unknown ply format " + f);
        } else if (key == "element") {
            PlyElement e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (key == "property") {
            if (elems.empty()) fail("property before any element in '" + fileName + "'");
            PlyProperty p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it;
                p.list = true;
                p.countType = plyType(ct);
                p.type = plyType(it);
                if (p.countType == kNoType || p.type == kNoType)
                    fail("'" + fileName + "': unknown list property type " + ct + " " + it);
            } else {
                p.type = plyType(t);
                if (p.type == kNoType) fail("'" + fileName + "': unknown property type " + t);
            }
            ls >> p.name;
            elems.back().props.push_back(p);
        } else if (key == "end_header") {
            break;
        }
    }
    if (format < 0) fail("'" + fileName + "': no format line");
    const PlyElement* ev = nullptr;
    const PlyElement* ee = nullptr;
    const PlyElement* em = nullptr;
    const PlyElement* eb = nullptr;
    for (const auto& e : elems) {
        if (e.name == "vertex") ev = &e;
        if (e.name == "3d_element") ee = &e;
        if (e.name == "material") em = &e;
        if (e.name == "boundary_condition") eb = &e;
    }
    // FemPlyReaderDelegate::fileIsAcceptable
    if (!ev || !ee || ev->find("x") < 0 || ev->find("y") < 0 || ev->find("z") < 0)
        fail("'" + fileName + "' is not an acceptable Fem3D ply file (vertex x, y, z and 3d_element needed)");
    const int vi = ee->find("vertex_indices");
    if (vi < 0 || !ee->props[vi].list) fail("'" + fileName + "': 3d_element needs a vertex_indices list");
    const bool globalMaterial = em && em->find("mass_density") >= 0 && em->find("poisson_ratio") >= 0 &&
                                em->find("young_modulus") >= 0;
    const bool elementMaterial = ee->find("mass_density") >= 0 && ee->find("poisson_ratio") >= 0 &&
                                 ee->find("young_modulus") >= 0;
    m_vertices.clear();
    m_elements.clear();
    m_boundaryConditions.clear();
    Element material;  // zeros unless a material is given
    PlyValues values(in, format, fileName);
    for (const auto& e : elems) {
        for (size_t r = 0; r < e.count; ++r) {
            values.beginRow(e.name);
            std::vector<double> scal(e.props.size(), 0.0);
            std::vector<size_t> list;
            for (size_t p = 0; p < e.props.size(); ++p) {
                if (e.props[p].list) {
                    const double nd = values.read(e.props[p].countType, e.name);
                    if (!(nd >= 0.0)) fail("'" + fileName + "': bad list in element " + e.name);
                    const size_t n = (size_t)nd;
                    list.resize(n);
                    for (size_t k = 0; k < n; ++k) {
                        const double v = values.read(e.props[p].type, e.name);
                        if (!(v >= 0.0)) fail("'" + fileName + "': bad list in element " + e.name);
                        list[k] = (size_t)v;
                    }
                } else {
                    scal[p] = values.read(e.props[p].type, e.name);
                }
            }
            if (&e == ev) {
                addVertex({{scal[ev->find("x")], scal[ev->find("y")], scal[ev->find("z")]}});
            } else if (&e == ee) {
                Element el;
                if (list.size() != 4 && list.size() != 8)
                    fail("'" + fileName + "': 3d_element with " + std::to_string(list.size()) +
                         " nodes (4 or 8 expected)");
                el.nodeIds = list;
                if (elementMaterial && !globalMaterial) {
                    el.massDensity = scal[ee->find("mass_density")];
                    el.poissonRatio = scal[ee->find("poisson_ratio")];
                    el.youngModulus = scal[ee->find("young_modulus")];
                }
                m_elements.push_back(el);
            } else if (&e == em && globalMaterial && r == 0) {
                material.massDensity = scal[em->find("mass_density")];
                material.poissonRatio = scal[em->find("poisson_ratio")];
                material.youngModulus = scal[em->find("young_modulus")];
            } else if (&e == eb && eb->find("vertex_index") >= 0) {
                addBoundaryCondition((size_t)scal[eb->find("vertex_index")]);
            }
        }
    }
    if (globalMaterial)  // FemPlyReaderDelegate::endMaterials applies the one material to every element
        for (auto& el : m_elements) {
            el.massDensity = material.massDensity;
            el.poissonRatio = material.poissonRatio;
            el.youngModulus = material.youngModulus;
        }
    for (const auto& el : m_elements)
        for (size_t id : el.nodeIds)
            if (id >= getNumVertices()) fail("'" + fileName + "': element node id out of range");
    for (size_t b : m_boundaryConditions)
        if (b >= getNumVertices()) fail("'" + fileName + "': boundary condition vertex out of range");
}

void Fem3DRepresentation::loadFem(const std::string& fileName) {
    auto mesh = std::make_shared<Fem3D>();
    mesh->load(fileName);
    setFem(mesh);
}

void Fem3DRepresentation::setFem(std::shared_ptr<Fem3D> mesh) {
    if (isInitialized()) fail("The Fem cannot be set after initialization");
    if (!mesh) fail("Mesh for Fem3DRepresentation cannot be a nullptr");
    m_fem = mesh;
    auto state = std::make_shared<Math::OdeState>();
    state->setNumDof(3, mesh->getNumVertices());
    state->getPositions() = mesh->getVertexPositions();
    for (size_t b : mesh->getBoundaryConditions()) state->addBoundaryCondition(b);
    if (mesh->getNumElements() > 0) {
        const size_t nn = mesh->getElement(0).nodeIds.size();
        for (const auto& e : mesh->getElements())
            if (e.nodeIds.size() != nn)
                fail("Cannot mix and match elements of different nature. Found an element with " +
                     std::to_string(e.nodeIds.size()) + " nodes but was expecting " + std::to_string(nn));
        if (m_elementType.empty())
            m_elementType = nn == 4 ? std::string("SurgSim::Physics::Fem3DElementTetrahedron")
                                    : std::string("SurgSim::Physics::Fem3DElementCube");
    }
    setInitialState(state);
    m_elements.clear();
    for (const auto& e : mesh->getElements()) {
        std::shared_ptr<FemElement> el;
        if (m_elementType == "SurgSim::Physics::Fem3DElementCube") {
            if (e.nodeIds.size() != 8) fail("Fem3DElementCube needs 8 nodes");
            std::array<size_t, 8> ids;
            std::copy(e.nodeIds.begin(), e.nodeIds.end(), ids.begin());
            el = std::make_shared<Fem3DElementCube>(ids);
        } else {
            if (e.nodeIds.size() != 4) fail(m_elementType + " needs 4 nodes");
            std::array<size_t, 4> ids;
            std::copy(e.nodeIds.begin(), e.nodeIds.end(), ids.begin());
            if (m_elementType == "SurgSim::Physics::Fem3DElementCorotationalTetrahedron")
                el = std::make_shared<Fem3DElementCorotationalTetrahedron>(ids);
            else
                el = std::make_shared<Fem3DElementTetrahedron>(ids);
        }
        el->setMassDensity(e.massDensity);
        el->setPoissonRatio(e.poissonRatio);
        el->setYoungModulus(e.youngModulus);
        m_elements.push_back(el);
    }
}

}  // namespace Physics
}  // namespace SurgSimGpu
