// Fem3D PLY reading (ASCII) and Fem3DRepresentation::loadFem / setFem.
//
// Semantics follow Physics/FemPlyReaderDelegate.cpp:31-145 and Physics/Fem3DPlyReaderDelegate.cpp:52-96:
//   - required: element "vertex" with x, y, z; element "3d_element" with list property vertex_indices;
//   - optional: vertex thetaX/thetaY/thetaZ (ignored for Fem3D), "boundary_condition" with vertex_index,
//     a single "material" row (mass_density, poisson_ratio, young_modulus) applied to every element, or the
//     same three properties per 3d_element; the "material" element wins when both are present; no material
//     leaves zeros;
//   - other elements (e.g. "face") are read and skipped.
#include "SurgSimGpu.h"

#include <fstream>
#include <sstream>

namespace SurgSimGpu {
namespace Physics {

namespace {

[[noreturn]] void fail(const std::string& m) { throw Framework::AssertionFailure(m); }

struct PlyProperty {
    std::string name;
    bool list = false;
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

}  // namespace

void Fem3D::load(const std::string& fileName) {
    std::ifstream in(fileName);
    if (!in) fail("Could not open '" + fileName + "'");
    std::string line;
    std::getline(in, line);
    if (line.rfind("ply", 0) != 0) fail("'" + fileName + "' is not a ply file");
    std::vector<PlyElement> elems;
    bool ascii = false;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "format") {
            std::string f;
            ls >> f;
            ascii = f == "ascii";
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
            }
            ls >> p.name;
            elems.back().props.push_back(p);
        } else if (key == "end_header") {
            break;
        }
    }
    if (!ascii) fail("'" + fileName + "': only ASCII ply files are read");
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
    for (const auto& e : elems) {
        for (size_t r = 0; r < e.count; ++r) {
            if (!std::getline(in, line)) fail("'" + fileName + "': unexpected end of data in element " + e.name);
            std::istringstream ls(line);
            std::vector<double> scal(e.props.size(), 0.0);
            std::vector<size_t> list;
            for (size_t p = 0; p < e.props.size(); ++p) {
                if (e.props[p].list) {
                    size_t n = 0;
                    if (!(ls >> n)) fail("'" + fileName + "': bad list in element " + e.name);
                    list.resize(n);
                    for (size_t k = 0; k < n; ++k)
                        if (!(ls >> list[k])) fail("'" + fileName + "': bad list in element " + e.name);
                } else if (!(ls >> scal[p])) {
                    fail("'" + fileName + "': bad value in element " + e.name);
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
