// Tests of the C++ adapters (SurgSimGpu::Physics::Fem3DRepresentation and friends), written after the
// reference's unit tests (Math/UnitTests/OdeStateTests.cpp, Physics/UnitTests/Fem3DRepresentationTests.cpp,
// Physics/UnitTests/FemRepresentationTests.cpp). GPU cases compare the GPU step with the FP64 oracle
// (oracle/fem_oracle.cpp, C API) on the same mesh.
//
// usage: fem3d_representation_test --cpu | --gpu [ply data directory]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iterator>
#include <string>
#include <vector>

#include "../../opensurgsim_amd/host/SurgSimGpu.h"

using namespace SurgSimGpu;
using SurgSimGpu::Framework::AssertionFailure;

// ---- oracle C API (test infrastructure) --------------------------------------------------------
extern "C" {
void* orc_sim_create(int kind, int num_nodes, const double* x0, int num_elems, const int* conn, const double* E,
                     const double* nu, const double* rho);
void orc_sim_destroy(void* h);
void orc_sim_add_bc_dof(void* h, int dof);
void orc_sim_set_gravity(void* h, double gx, double gy, double gz, int enabled);
void orc_sim_set_external_force(void* h, const double* fext);
void orc_sim_set_solver(void* h, int scheme, int linsolver, double tol, long max_iter, long newton_max_iter,
                        int faithful_bc);
void orc_sim_set_state(void* h, const double* x, const double* v);
void orc_sim_get_state(void* h, double* x, double* v);
int orc_sim_step(void* h, double dt, int nsteps);
long orc_sim_nnz(void* h);
void orc_sim_pattern(void* h, int* outer, int* inner);
void orc_sim_set_damping(void* h, double rayleigh_mass, double rayleigh_stiffness, int lumping);
int orc_sim_add_localized_force(void* h, int element, const double* coord, const double* force, const double* K,
                                const double* D);
int orc_sim_apply_compliance(void* h, const double* b, int num_columns, double* out);
int orc_sim_compliance_matrix(void* h, double* out);
}

// ---- minimal test harness ------------------------------------------------------------------------
namespace {
int g_failures = 0;
std::string g_current;

#define EXPECT_TRUE(c)                                                                           \
    do {                                                                                         \
        if (!(c)) {                                                                              \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                          \
            ++g_failures;                                                                        \
        }                                                                                        \
    } while (0)
#define EXPECT_THROW_ASSERT(stmt)                                                                \
    do {                                                                                         \
        bool _thrown = false;                                                                    \
        try {                                                                                    \
            stmt;                                                                                \
        } catch (const AssertionFailure&) {                                                      \
            _thrown = true;                                                                      \
        }                                                                                        \
        if (!_thrown) {                                                                          \
            std::printf("  FAILED %s:%d: no AssertionFailure from %s\n", __FILE__, __LINE__, #stmt); \
            ++g_failures;                                                                        \
        }                                                                                        \
    } while (0)

void run(const char* name, const std::function<void()>& f) {
    g_current = name;
    std::printf("[ RUN  ] %s\n", name);
    std::fflush(stdout);
    const int before = g_failures;
    try {
        f();
    } catch (const std::exception& e) {
        std::printf("  FAILED: unexpected exception: %s\n", e.what());
        ++g_failures;
    }
    std::printf("[ %s ] %s\n", g_failures == before ? " OK " : "FAIL", name);
    std::fflush(stdout);
}

double relErr(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1.0));
}

// Block mesh of nx x ny x nz cubes (edge h): node id(i,j,k) = i (ny+1)(nz+1) + j (nz+1) + k, y vertical.
struct Block {
    int nx, ny, nz;
    std::vector<double> x0;
    std::vector<int> tets, hexes;
    int id(int i, int j, int k) const { return i * (ny + 1) * (nz + 1) + j * (nz + 1) + k; }
    int numNodes() const { return (nx + 1) * (ny + 1) * (nz + 1); }
};

Block makeBlock(int nx, int ny, int nz, double h = 0.01) {
    Block b{nx, ny, nz, {}, {}, {}};
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                b.x0.push_back(i * h);
                b.x0.push_back(j * h);
                b.x0.push_back(k * h);
            }
    // 5-tet split of RenderTestFem3DCorotational.cpp:53-73 on even cubes; mirrored (corner L -> L xor 1,
    // last two ids swapped) on odd cubes so that face diagonals conform.
    const int split[5][4] = {{4, 7, 1, 2}, {4, 1, 7, 5}, {4, 2, 1, 0}, {4, 7, 2, 6}, {1, 2, 7, 3}};
    const int hexOrder[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int corner[8];
                for (int L = 0; L < 8; ++L) corner[L] = b.id(i + (L & 1), j + ((L >> 1) & 1), k + ((L >> 2) & 1));
                const bool odd = (i + j + k) & 1;
                for (int t = 0; t < 5; ++t) {
                    int ids[4];
                    for (int a = 0; a < 4; ++a) ids[a] = corner[odd ? (split[t][a] ^ 1) : split[t][a]];
                    if (odd) std::swap(ids[2], ids[3]);
                    for (int a = 0; a < 4; ++a) b.tets.push_back(ids[a]);
                }
                for (int a = 0; a < 8; ++a) b.hexes.push_back(corner[hexOrder[a]]);
            }
    return b;
}

struct Material {
    double E, nu, rho;
};

// Builds a representation the way the reference's tests do: state, boundary conditions, elements.
std::shared_ptr<Physics::Fem3DRepresentation> makeRepresentation(const Block& b, const std::string& type,
                                                                 const Material& m,
                                                                 const std::vector<double>& v0) {
    auto fem = std::make_shared<Physics::Fem3DRepresentation>("fem");
    fem->setFemElementType(type);
    fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
    auto state = std::make_shared<Math::OdeState>();
    state->setNumDof(3, b.numNodes());
    state->getPositions() = b.x0;
    if (!v0.empty()) state->getVelocities() = v0;
    for (int i = 0; i <= b.nx; ++i)
        for (int k = 0; k <= b.nz; ++k) state->addBoundaryCondition(b.id(i, 0, k));  // y = 0 face fixed
    fem->setInitialState(state);
    const bool hex = type == "SurgSim::Physics::Fem3DElementCube";
    const size_t ne = hex ? b.hexes.size() / 8 : b.tets.size() / 4;
    for (size_t e = 0; e < ne; ++e) {
        std::shared_ptr<Physics::FemElement> el;
        if (hex) {
            std::array<size_t, 8> ids;
            for (int a = 0; a < 8; ++a) ids[a] = b.hexes[8 * e + a];
            el = std::make_shared<Physics::Fem3DElementCube>(ids);
        } else {
            std::array<size_t, 4> ids;
            for (int a = 0; a < 4; ++a) ids[a] = b.tets[4 * e + a];
            if (type == "SurgSim::Physics::Fem3DElementTetrahedron")
                el = std::make_shared<Physics::Fem3DElementTetrahedron>(ids);
            else
                el = std::make_shared<Physics::Fem3DElementCorotationalTetrahedron>(ids);
        }
        el->setYoungModulus(m.E);
        el->setPoissonRatio(m.nu);
        el->setMassDensity(m.rho);
        fem->addFemElement(el);
    }
    return fem;
}

void* makeOracle(const Block& b, int kind, const Material& m, const std::vector<double>& v0, int scheme,
                 int linsolver, double tol) {
    const bool hex = kind == 2;
    const std::vector<int>& conn = hex ? b.hexes : b.tets;
    const int ne = (int)conn.size() / (hex ? 8 : 4);
    std::vector<double> E(ne, m.E), nu(ne, m.nu), rho(ne, m.rho);
    void* o = orc_sim_create(kind, b.numNodes(), b.x0.data(), ne, conn.data(), E.data(), nu.data(), rho.data());
    for (int i = 0; i <= b.nx; ++i)
        for (int k = 0; k <= b.nz; ++k)
            for (int c = 0; c < 3; ++c) orc_sim_add_bc_dof(o, 3 * b.id(i, 0, k) + c);
    orc_sim_set_solver(o, scheme, linsolver, tol, -1, 1, 0);
    if (!v0.empty()) orc_sim_set_state(o, nullptr, v0.data());
    return o;
}

std::vector<double> topVelocities(const Block& b, double amplitude) {
    std::vector<double> v(3 * b.numNodes(), 0.0);
    unsigned s = 12345u;
    for (int i = 0; i <= b.nx; ++i)
        for (int k = 0; k <= b.nz; ++k) {
            const int n = b.id(i, b.ny, k);
            s = s * 1664525u + 1013904223u;
            v[3 * n] = amplitude * ((s >> 8) * (1.0 / 16777216.0) - 0.5);
            s = s * 1664525u + 1013904223u;
            v[3 * n + 2] = amplitude * ((s >> 8) * (1.0 / 16777216.0) - 0.5);
        }
    return v;
}

const std::string kTet = "SurgSim::Physics::Fem3DElementTetrahedron";
const std::string kCoro = "SurgSim::Physics::Fem3DElementCorotationalTetrahedron";
const std::string kCube = "SurgSim::Physics::Fem3DElementCube";

#ifndef FEM_TEST_DATA_DIR
#define FEM_TEST_DATA_DIR "tests/golden/ply"
#endif
std::string g_data_dir = FEM_TEST_DATA_DIR;  // the reference's PLY fixtures (second argument overrides)

std::string dataPath(const std::string& name) { return g_data_dir + "/" + name; }

bool near(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::max(1.0, std::fabs(b)); }

// A Fem3D written as a binary PLY (ply.c's binary formats, DataStructures/ply.c:760-762): vertex x, y, z as
// `coord` ("double" or "float"), 3d_element vertex_indices as a list of `count` / `index` types, a material row in
// double, boundary conditions as uint. bigEndian swaps every value's bytes.
template <typename Coord, typename Count, typename Index>
void writeBinaryPly(const Physics::Fem3D& fem, const std::string& path, bool bigEndian, const char* coord,
                    const char* count, const char* index) {
    std::ofstream out(path, std::ios::binary);
    out << "ply\nformat " << (bigEndian ? "binary_big_endian" : "binary_little_endian") << " 1.0\n"
        << "comment written by the adapter tests\n"
        << "element vertex " << fem.getNumVertices() << "\n"
        << "property " << coord << " x\nproperty " << coord << " y\nproperty " << coord << " z\n"
        << "element 3d_element " << fem.getNumElements() << "\n"
        << "property list " << count << " " << index << " vertex_indices\n"
        << "element material 1\nproperty double mass_density\nproperty double poisson_ratio\n"
        << "property double young_modulus\n"
        << "element boundary_condition " << fem.getBoundaryConditions().size() << "\n"
        << "property uint vertex_index\nend_header\n";
    auto put = [&](auto v) {
        unsigned char b[sizeof(v)];
        std::memcpy(b, &v, sizeof(v));
        if (bigEndian) std::reverse(b, b + sizeof(v));
        out.write(reinterpret_cast<const char*>(b), sizeof(v));
    };
    for (size_t i = 0; i < fem.getNumVertices(); ++i)
        for (double c : fem.getVertexPosition(i)) put((Coord)c);
    for (const auto& e : fem.getElements()) {
        put((Count)e.nodeIds.size());
        for (size_t id : e.nodeIds) put((Index)id);
    }
    const auto& e0 = fem.getElement(0);
    put(e0.massDensity);
    put(e0.poissonRatio);
    put(e0.youngModulus);
    for (size_t b : fem.getBoundaryConditions()) put((uint32_t)b);
}

// ---- CPU-only tests (no GPU context is created) ----------------------------------------------------
void plyTests() {
    run("Fem3DRepresentationReaderTests.BinaryFormats", [] {
        // the reference's wound mesh written back in ply.c's two binary formats reads as the ascii original
        Physics::Fem3D ascii;
        ascii.load(dataPath("wound_deformable.ply"));
        const std::string dir = "/tmp";
        const std::string le = dir + "/fem3d_wound_le.ply", be = dir + "/fem3d_wound_be.ply";
        writeBinaryPly<double, uint8_t, int32_t>(ascii, le, false, "double", "uchar", "int");
        writeBinaryPly<float, uint16_t, uint32_t>(ascii, be, true, "float", "ushort", "uint");
        Physics::Fem3D a, b;
        a.load(le);
        b.load(be);
        EXPECT_TRUE(a.getVertexPositions() == ascii.getVertexPositions());
        bool floatsMatch = b.getNumVertices() == ascii.getNumVertices();
        for (size_t i = 0; floatsMatch && i < 3 * ascii.getNumVertices(); ++i)
            floatsMatch = b.getVertexPositions()[i] == (double)(float)ascii.getVertexPositions()[i];
        EXPECT_TRUE(floatsMatch);
        for (const Physics::Fem3D* f : {&a, &b}) {
            EXPECT_TRUE(f->getNumElements() == ascii.getNumElements());
            bool same = true;
            for (size_t e = 0; same && e < ascii.getNumElements(); ++e)
                same = f->getElement(e).nodeIds == ascii.getElement(e).nodeIds &&
                       f->getElement(e).youngModulus == ascii.getElement(e).youngModulus &&
                       f->getElement(e).poissonRatio == ascii.getElement(e).poissonRatio &&
                       f->getElement(e).massDensity == ascii.getElement(e).massDensity;
            EXPECT_TRUE(same);
            EXPECT_TRUE(f->getBoundaryConditions() == ascii.getBoundaryConditions());
        }
        // a truncated binary file is rejected
        {
            std::ifstream src(le, std::ios::binary);
            std::string all((std::istreambuf_iterator<char>(src)), std::istreambuf_iterator<char>());
            std::ofstream cut(dir + "/fem3d_wound_cut.ply", std::ios::binary);
            cut.write(all.data(), (std::streamsize)(all.size() - 40));
        }
        Physics::Fem3D c;
        EXPECT_THROW_ASSERT(c.load(dir + "/fem3d_wound_cut.ply"));
        std::remove(le.c_str());
        std::remove(be.c_str());
        std::remove((dir + "/fem3d_wound_cut.ply").c_str());
    });
    // Physics/UnitTests/Fem3DPlyReaderDelegateTests.cpp, on the reference's own data files
    run("Fem3DRepresentationReaderTests.TetrahedronMeshDelegateTest", [] {
        Physics::Fem3D fem;
        fem.load(dataPath("Tetrahedron.ply"));
        const auto v0 = fem.getVertexPosition(0), v25 = fem.getVertexPosition(25);
        EXPECT_TRUE(near(v0[0], -1.0) && near(v0[1], 0.0) && near(v0[2], 0.0));
        EXPECT_TRUE(near(v25[0], 1.0) && near(v25[1], 1.0) && near(v25[2], 1.0));
        EXPECT_TRUE(fem.getNumElements() == 44u);
        const std::vector<size_t> t0 = {16, 0, 7, 5}, t11 = {13, 14, 16, 2};
        EXPECT_TRUE(fem.getElement(0).nodeIds == t0);
        EXPECT_TRUE(fem.getElement(11).nodeIds == t11);
        EXPECT_TRUE(fem.getBoundaryConditions().size() == 4u);
        EXPECT_TRUE(fem.getBoundaryCondition(0) == 19u && fem.getBoundaryCondition(1) == 21u &&
                    fem.getBoundaryCondition(2) == 23u && fem.getBoundaryCondition(3) == 25u);
        for (size_t e : {2u, 8u}) {
            EXPECT_TRUE(fem.getElement(e).massDensity == 900.0);
            EXPECT_TRUE(fem.getElement(e).poissonRatio == 0.45);
            EXPECT_TRUE(fem.getElement(e).youngModulus == 1750000000.0);
        }
    });
    run("Fem3DRepresentationReaderTests.CubeMeshDelegateTest", [] {
        Physics::Fem3D fem;
        fem.load(dataPath("Fem3DCube.ply"));
        const auto v0 = fem.getVertexPosition(0), v5 = fem.getVertexPosition(5);
        EXPECT_TRUE(near(v0[0], 1.0) && near(v0[1], 1.0) && near(v0[2], 1.0));
        EXPECT_TRUE(near(v5[0], 2.0) && near(v5[1], 2.0) && near(v5[2], 2.0));
        EXPECT_TRUE(fem.getNumElements() == 3u);
        const std::vector<size_t> c0 = {0, 1, 2, 3, 4, 5, 6, 7}, c2 = {3, 4, 5, 0, 2, 6, 8, 7};
        EXPECT_TRUE(fem.getElement(0).nodeIds == c0);
        EXPECT_TRUE(fem.getElement(2).nodeIds == c2);
        EXPECT_TRUE(fem.getBoundaryConditions().size() == 2u);
        EXPECT_TRUE(fem.getBoundaryCondition(0) == 9u && fem.getBoundaryCondition(1) == 5u);
        for (size_t e : {1u, 2u}) {
            EXPECT_TRUE(fem.getElement(e).massDensity == 0.2);
            EXPECT_TRUE(fem.getElement(e).poissonRatio == 0.3);
            EXPECT_TRUE(fem.getElement(e).youngModulus == 0.4);
        }
    });
    run("Fem3DRepresentationReaderTests.WrongPlyWithRotationDof", [] {
        Physics::Fem3D fem;
        fem.load(dataPath("Wrong3DFileWithRotationData.ply"));  // must not throw
        EXPECT_TRUE(fem.getNumElements() == 1u);
    });
    run("Fem3DRepresentationReaderTests.PerElementMaterial", [] {
        Physics::Fem3D fem;
        fem.load(dataPath("Fem3DCubeMaterial.ply"));
        double value = 1.0;
        for (size_t i = 0; i < fem.getNumElements(); ++i) {
            EXPECT_TRUE(fem.getElement(i).massDensity == value++);
            EXPECT_TRUE(fem.getElement(i).poissonRatio == value++);
            EXPECT_TRUE(fem.getElement(i).youngModulus == value++);
        }
    });
    run("Fem3DRepresentationReaderTests.NoMaterials", [] {
        Physics::Fem3D fem;
        fem.load(dataPath("Fem3DCubeNoMaterial.ply"));
        for (const auto& e : fem.getElements())
            EXPECT_TRUE(e.massDensity == 0.0 && e.poissonRatio == 0.0 && e.youngModulus == 0.0);
    });
    run("Fem3DRepresentationReaderTests.MissingFileAndLoadFem", [] {
        Physics::Fem3D fem;
        EXPECT_THROW_ASSERT(fem.load(dataPath("does_not_exist.ply")));
        Physics::Fem3DRepresentation rep("wound");
        rep.loadFem(dataPath("wound_deformable.ply"));
        EXPECT_TRUE(rep.getNumFemElements() == 1476u);
        EXPECT_TRUE(rep.getNumDof() == 3u * 391u);
        EXPECT_TRUE(rep.getInitialState()->getNumBoundaryConditions() == 3u * 79u);
        EXPECT_TRUE(rep.getFemElementType() == "SurgSim::Physics::Fem3DElementTetrahedron");
        EXPECT_TRUE(rep.getFemElement(0)->getYoungModulus() == 7.5e4 && rep.getFemElement(0)->getMassDensity() == 1000.0);
    });
}

void cpuTests() {
    plyTests();
    run("OdeStateTest.SetNumDof", [] {
        Math::OdeState s;
        s.setNumDof(3, 4);
        EXPECT_TRUE(s.getNumDof() == 12 && s.getNumNodes() == 4);
        EXPECT_TRUE(s.getPositions().size() == 12 && s.getVelocities().size() == 12);
        for (double a : s.getPositions()) EXPECT_TRUE(a == 0.0);
        EXPECT_TRUE(s.getNumBoundaryConditions() == 0);
        EXPECT_THROW_ASSERT(s.setNumDof(0, 4));
    });
    run("OdeStateTest.BoundaryConditions", [] {
        Math::OdeState s;
        EXPECT_THROW_ASSERT(s.addBoundaryCondition(0));  // setNumDof first
        s.setNumDof(3, 4);
        s.addBoundaryCondition(1);
        s.addBoundaryCondition(1);  // duplicates are ignored
        s.addBoundaryCondition(3, 2);
        EXPECT_TRUE(s.getNumBoundaryConditions() == 4);
        EXPECT_TRUE(s.isBoundaryCondition(3) && s.isBoundaryCondition(4) && s.isBoundaryCondition(5));
        EXPECT_TRUE(s.isBoundaryCondition(11) && !s.isBoundaryCondition(9) && !s.isBoundaryCondition(0));
        EXPECT_THROW_ASSERT(s.addBoundaryCondition(4));
        EXPECT_THROW_ASSERT(s.addBoundaryCondition(0, 3));
        Math::Vector v(12, 1.0);
        s.applyBoundaryConditionsToVector(&v);
        EXPECT_TRUE(v[3] == 0.0 && v[4] == 0.0 && v[5] == 0.0 && v[11] == 0.0 && v[0] == 1.0 && v[10] == 1.0);
        s.reset();
        EXPECT_TRUE(s.getNumBoundaryConditions() == 0 && !s.isBoundaryCondition(3));
    });
    run("OdeStateTest.IsValid", [] {
        Math::OdeState s;
        s.setNumDof(3, 2);
        EXPECT_TRUE(s.isValid());
        s.getVelocities()[4] = std::nan("");
        EXPECT_TRUE(!s.isValid());
        s.getVelocities()[4] = 0.0;
        s.getPositions()[1] = INFINITY;
        EXPECT_TRUE(!s.isValid());
    });
    run("LinearSparseSolveAndInverseTest.CGSettings", [] {
        Math::LinearSparseSolveAndInverseCG cg;
        EXPECT_TRUE(cg.getTolerance() == 2.220446049250313e-16 && cg.getMaxIterations() < 0);
        cg.setTolerance(1e-8);
        cg.setMaxIterations(100);
        EXPECT_TRUE(cg.getTolerance() == 1e-8 && cg.getMaxIterations() == 100);
        EXPECT_THROW_ASSERT(cg.setTolerance(0.0));
        EXPECT_TRUE(cg.getType() == Math::LINEARSOLVER_CONJUGATEGRADIENT);
    });
    run("OdeSolverTest.SupportedSchemes", [] {
        Math::OdeSolver implicitSolver(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        EXPECT_TRUE(implicitSolver.getName() == "Ode Solver Euler Implicit");
        Math::OdeSolver linear(Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT);
        EXPECT_TRUE(linear.getNewtonRaphsonMaximumIteration() == 1);
        // every scheme of Math/OdeSolver.h:36-49, with the reference's solver names (Math/OdeSolver*.cpp)
        EXPECT_TRUE(Math::OdeSolver(Math::INTEGRATIONSCHEME_STATIC).getName() == "Ode Solver Static");
        EXPECT_TRUE(Math::OdeSolver(Math::INTEGRATIONSCHEME_EULER_EXPLICIT).getName() == "Ode Solver Euler Explicit");
        EXPECT_TRUE(Math::OdeSolver(Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED).getName() ==
                    "Ode Solver Linear Euler Explicit Modified");
        EXPECT_TRUE(Math::OdeSolver(Math::INTEGRATIONSCHEME_RUNGE_KUTTA_4).getName() == "Ode Solver Runge Kutta 4");
        EXPECT_THROW_ASSERT(Math::OdeSolver(Math::MAX_INTEGRATIONSCHEMES));
        EXPECT_THROW_ASSERT(linear.setNewtonRaphsonMaximumIteration(0));
    });
    run("Fem3DRepresentationTest.ElementTypeAndSettings", [] {
        Physics::Fem3DRepresentation fem("fem");
        EXPECT_TRUE(fem.getFemElementType().empty());
        EXPECT_THROW_ASSERT(fem.setFemElementType("SurgSim::Physics::Fem3DElementUnknown"));
        fem.setFemElementType(kCoro);
        EXPECT_TRUE(fem.getFemElementType() == kCoro);
        EXPECT_TRUE(fem.getIntegrationScheme() == Math::INTEGRATIONSCHEME_EULER_EXPLICIT);  // reference default
        EXPECT_TRUE(fem.getLinearSolver() == Math::LINEARSOLVER_LU);                        // reference default
        EXPECT_TRUE(fem.getNumDofPerNode() == 3 && fem.getNumDof() == 0 && !fem.isInitialized());
        EXPECT_TRUE(fem.isGravityEnabled());
    });
    run("Fem3DRepresentationTest.InitializePreconditions", [] {
        const Block b = makeBlock(1, 1, 1);
        const Material m{1e5, 0.45, 1000.0};
        {
            Physics::Fem3DRepresentation fem("fem");
            EXPECT_THROW_ASSERT(fem.initialize());  // no initial state
        }
        {
            auto fem = makeRepresentation(b, kCoro, m, {});
            auto bad = std::make_shared<Physics::Fem3DElementCorotationalTetrahedron>(std::array<size_t, 4>{0, 1, 2, 99});
            bad->setYoungModulus(1e5);
            bad->setPoissonRatio(0.45);
            bad->setMassDensity(1000);
            fem->addFemElement(bad);
            EXPECT_THROW_ASSERT(fem->initialize());  // node id out of range
        }
        {
            auto fem = makeRepresentation(b, kCoro, m, {});
            fem->getFemElement(0)->setPoissonRatio(0.5);
            EXPECT_THROW_ASSERT(fem->initialize());  // Physics/FemElement.cpp:33-43
        }
        {
            auto fem = makeRepresentation(b, kCoro, m, {});
            fem->addFemElement(std::make_shared<Physics::Fem3DElementCube>(std::array<size_t, 8>{0, 1, 2, 3, 4, 5, 6, 7}));
            EXPECT_THROW_ASSERT(fem->initialize());  // mixed element classes
        }
    });
}

// ---- GPU tests -----------------------------------------------------------------------------------------
void compareWithOracle(const std::string& type, int kind, Math::IntegrationScheme scheme, Math::LinearSolver ls,
                       double tol, int steps, double bound) {
    const Block b = kind == 2 ? makeBlock(3, 2, 2) : makeBlock(3, 3, 2);
    const Material m{kind == 2 ? 2e4 : 5e4, 0.45, 1000.0};
    const auto v0 = topVelocities(b, 0.02);
    auto fem = makeRepresentation(b, type, m, v0);
    fem->setIntegrationScheme(scheme);
    fem->setLinearSolver(ls);
    fem->initialize();
    if (auto cg = std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(fem->getOdeSolver()->getLinearSolver()))
        cg->setTolerance(tol);
    const int oscheme = scheme == Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT ? 1 : 0;
    void* o = makeOracle(b, kind, m, v0, oscheme, ls == Math::LINEARSOLVER_LU ? 0 : 1, tol);
    const double dt = 1e-3;
    std::vector<double> xo(b.x0.size()), vo(b.x0.size()), xprev;
    for (int s = 0; s < steps; ++s) {
        fem->update(dt);
        fem->afterUpdate(dt);
        if (s == steps - 2) xprev = fem->getCurrentState()->getPositions();
    }
    EXPECT_TRUE(orc_sim_step(o, dt, steps) == 0);
    orc_sim_get_state(o, xo.data(), vo.data());
    const auto& x = fem->getCurrentState()->getPositions();
    const auto& v = fem->getCurrentState()->getVelocities();
    std::vector<double> u(x.size()), uo(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        u[i] = x[i] - b.x0[i];
        uo[i] = xo[i] - b.x0[i];
    }
    const double ex = relErr(x, xo), eu = relErr(u, uo), ev = relErr(v, vo);
    std::printf("  %s: rel err x %.3e u %.3e v %.3e\n", g_current.c_str(), ex, eu, ev);
    EXPECT_TRUE(ex < bound && eu < bound && ev < bound);
    EXPECT_TRUE(fem->getPreviousState()->getPositions() == xprev);
    EXPECT_TRUE(*fem->getFinalState() == *fem->getCurrentState());
    EXPECT_TRUE(fem->isActive());
    EXPECT_TRUE(fem->getStatistics().steps == steps);
    orc_sim_destroy(o);
}

// The oracle's (and the library's) number of a Math::IntegrationScheme (Math/OdeSolver.h:36-49).
int oracleScheme(Math::IntegrationScheme s) {
    switch (s) {
        case Math::INTEGRATIONSCHEME_EULER_IMPLICIT: return 0;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT: return 1;
        case Math::INTEGRATIONSCHEME_STATIC: return 2;
        case Math::INTEGRATIONSCHEME_LINEAR_STATIC: return 3;
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT: return 4;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT: return 5;
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED: return 6;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED: return 7;
        case Math::INTEGRATIONSCHEME_RUNGE_KUTTA_4: return 8;
        default: return 9;
    }
}

// A representation with Rayleigh damping (and mass lumping) under scheme, element-localized external terms
// before the first update, against the oracle.
void compareDampedScheme(Math::IntegrationScheme scheme, bool lumping, bool localized, double dt, int steps) {
    const Block b = makeBlock(3, 3, 2);
    const Material m{5e3, 0.45, 1000.0};
    const auto v0 = topVelocities(b, 0.02);
    auto fem = makeRepresentation(b, kCoro, m, v0);
    fem->setIntegrationScheme(scheme);
    fem->setLinearSolver(Math::LINEARSOLVER_CONJUGATEGRADIENT);
    fem->setRayleighDampingMass(2.0);
    fem->setRayleighDampingStiffness(1e-4);
    fem->setMassLumping(lumping);
    EXPECT_TRUE(fem->getRayleighDampingMass() == 2.0 && fem->getRayleighDampingStiffness() == 1e-4);
    EXPECT_TRUE(fem->getMassLumping() == lumping);
    fem->initialize();
    std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(fem->getOdeSolver()->getLinearSolver())
        ->setTolerance(1e-12);
    void* o = makeOracle(b, 1, m, v0, oracleScheme(scheme), 1, 1e-12);
    orc_sim_set_damping(o, 2.0, 1e-4, lumping ? 1 : 0);
    const Math::Vector coord{0.1, 0.2, 0.3, 0.4};
    const Math::Vector3d force{{0.02, -0.05, 0.01}};
    const Math::Matrix33d K{{40.0, 2.0, 0.0, 1.0, 50.0, 3.0, 0.0, -2.0, 45.0}};
    const Math::Matrix33d D{{0.1, 0.0, 0.01, 0.0, 0.2, 0.0, 0.02, 0.0, 0.15}};
    for (int s = 0; s < steps; ++s) {
        if (localized && s == 0) {
            fem->addExternalGeneralizedForce(7, coord, force, &K, &D);
            EXPECT_TRUE(orc_sim_add_localized_force(o, 7, coord.data(), force.data(), K.data(), D.data()) == 0);
        }
        fem->update(dt);
        fem->afterUpdate(dt);
        EXPECT_TRUE(orc_sim_step(o, dt, 1) == 0);
    }
    std::vector<double> xo(b.x0.size()), vo(b.x0.size());
    orc_sim_get_state(o, xo.data(), vo.data());
    const auto& x = fem->getCurrentState()->getPositions();
    std::vector<double> u(x.size()), uo(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        u[i] = x[i] - b.x0[i];
        uo[i] = xo[i] - b.x0[i];
    }
    const double eu = relErr(u, uo);
    double vn = 0.0;
    for (double q : vo) vn += q * q;
    const double ev = vn > 0 ? relErr(fem->getCurrentState()->getVelocities(), vo) : 0.0;
    std::printf("  %s: rel err u %.3e v %.3e\n", g_current.c_str(), eu, ev);
    EXPECT_TRUE(eu < 1e-9 && ev < 1e-9);
    orc_sim_destroy(o);
}

void gpuTests() {
    run("Fem3DRepresentationGpuTest.RayleighLumpingLocalizedImplicit", [] {
        compareDampedScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT, true, true, 1e-3, 6);
    });
    run("Fem3DRepresentationGpuTest.RayleighRungeKutta4", [] {
        compareDampedScheme(Math::INTEGRATIONSCHEME_RUNGE_KUTTA_4, false, true, 2e-4, 6);
    });
    run("Fem3DRepresentationGpuTest.RayleighModifiedExplicitEuler", [] {
        compareDampedScheme(Math::INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED, false, false, 2e-4, 6);
    });
    run("Fem3DRepresentationGpuTest.StaticWithLocalizedStiffness", [] {
        compareDampedScheme(Math::INTEGRATIONSCHEME_STATIC, false, true, 1e-3, 2);
    });
    run("Fem3DRepresentationGpuTest.CorotationalImplicitEulerCG", [] {
        compareWithOracle(kCoro, 1, Math::INTEGRATIONSCHEME_EULER_IMPLICIT, Math::LINEARSOLVER_CONJUGATEGRADIENT,
                          1e-12, 20, 1e-9);
    });
    run("Fem3DRepresentationGpuTest.LinearTetLinearEulerImplicitLU", [] {
        compareWithOracle(kTet, 0, Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT, Math::LINEARSOLVER_LU, 1e-13, 10,
                          1e-9);
    });
    run("Fem3DRepresentationGpuTest.ComplianceMatrixMatchesOracle", [] {
        // getComplianceMatrix after a linear implicit Euler step with the direct solver: the oracle's dense
        // inverse with boundary rows and columns zeroed (Math/OdeSolver.cpp:69-76)
        const Block b = makeBlock(3, 2, 3);
        const Material m{5e4, 0.45, 1000.0};
        auto fem = makeRepresentation(b, kTet, m, {});
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT);
        fem->setLinearSolver(Math::LINEARSOLVER_LU);
        fem->initialize();
        fem->update(1e-3);
        fem->afterUpdate(1e-3);
        const Math::Vector C = fem->getComplianceMatrix();
        void* o = makeOracle(b, 0, m, {}, 1, 0, 1e-13);
        EXPECT_TRUE(orc_sim_step(o, 1e-3, 1) == 0);
        std::vector<double> Co(C.size());
        EXPECT_TRUE(orc_sim_compliance_matrix(o, Co.data()) == 0);
        double err = 0.0, mx = 0.0;
        for (size_t i = 0; i < C.size(); ++i) {
            err = std::max(err, std::fabs(C[i] - Co[i]));
            mx = std::max(mx, std::fabs(Co[i]));
        }
        std::printf("  compliance matrix max abs err / max %.3e\n", err / mx);
        EXPECT_TRUE(err <= 1e-12 * mx);
        orc_sim_destroy(o);
    });
    run("Fem3DRepresentationGpuTest.CubeImplicitEulerCG", [] {
        compareWithOracle(kCube, 2, Math::INTEGRATIONSCHEME_EULER_IMPLICIT, Math::LINEARSOLVER_CONJUGATEGRADIENT,
                          1e-12, 10, 1e-9);
    });
    run("Fem3DRepresentationGpuTest.PatternMatchesReference", [] {
        const Block b = makeBlock(2, 3, 2);
        const Material m{5e4, 0.45, 1000.0};
        auto fem = makeRepresentation(b, kCoro, m, {});
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        fem->initialize();
        std::vector<int32_t> rowptr, colidx;
        fem->getPattern(&rowptr, &colidx);
        void* o = makeOracle(b, 1, m, {}, 0, 1, 1e-10);
        std::vector<int> outer(rowptr.size()), inner((size_t)orc_sim_nnz(o));
        orc_sim_pattern(o, outer.data(), inner.data());
        EXPECT_TRUE(inner.size() == colidx.size());
        EXPECT_TRUE(std::equal(outer.begin(), outer.end(), rowptr.begin()));
        EXPECT_TRUE(std::equal(inner.begin(), inner.end(), colidx.begin()));
        // total mass = rho V of the block (Physics/FemRepresentation.cpp:192-200)
        EXPECT_TRUE(std::fabs(fem->getTotalMass() - 1000.0 * 2 * 3 * 2 * 1e-6) < 1e-12);
        orc_sim_destroy(o);
    });
    run("Fem3DRepresentationGpuTest.ConstraintComplianceMatchesOracle", [] {
        // C H^T and H C H^T of a fixed-point constraint (FemConstraintFixedPoint) and a frictionless contact
        // (FemConstraintFrictionlessContact) in element 57 (cube (1, 2, 1), clear of the fixed y = 0 face) after
        // one update: C H^T equals applyCompliance of
        // the dense H^T bit for bit and the oracle's solve to the CG tolerance; H C H^T equals H (C H^T).
        const Block b = makeBlock(3, 3, 2);
        const Material m{5e3, 0.45, 1000.0};
        const auto v0 = topVelocities(b, 0.02);
        auto fem = makeRepresentation(b, kCoro, m, v0);
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        fem->setLinearSolver(Math::LINEARSOLVER_CONJUGATEGRADIENT);
        EXPECT_THROW_ASSERT(fem->applyCompliance(Math::Vector(b.x0.size(), 0.0), 1));
        fem->initialize();
        std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(fem->getOdeSolver()->getLinearSolver())
            ->setTolerance(1e-12);
        EXPECT_THROW_ASSERT(fem->applyCompliance(Math::Vector(b.x0.size(), 0.0), 1));  // no system matrix yet
        const double dt = 1e-3;
        fem->update(dt);
        fem->afterUpdate(dt);
        void* o = makeOracle(b, 1, m, v0, 0, 1, 1e-12);
        EXPECT_TRUE(orc_sim_step(o, dt, 1) == 0);
        const size_t n = b.x0.size();
        const Math::Vector coord{0.1, 0.0, 0.45, 0.45};
        const auto& ids = fem->getFemElement(57)->getNodeIds();
        const double nrm = std::sqrt(0.09 + 0.64 + 0.25);
        for (const auto& dirs : {std::vector<Math::Vector3d>{{{1, 0, 0}}, {{0, 1, 0}}, {{0, 0, 1}}},
                                 std::vector<Math::Vector3d>{{{0.3 / nrm, -0.8 / nrm, 0.5 / nrm}}}}) {
            const size_t k = dirs.size();
            const double scale = dt * 0.7;
            Math::Vector Ht(n * k, 0.0);
            for (size_t j = 0; j < k; ++j)
                for (size_t a = 0; a < 4; ++a)
                    if (coord[a] != 0.0)
                        for (size_t c = 0; c < 3; ++c) Ht[j * n + 3 * ids[a] + c] = coord[a] * dirs[j][c] * scale;
            Math::Vector cht, hcht;
            fem->applyComplianceToConstraint(57, coord, dirs, scale, &cht, &hcht);
            EXPECT_TRUE(cht == fem->applyCompliance(Ht, k));
            Math::Vector ref(n * k);
            EXPECT_TRUE(orc_sim_apply_compliance(o, Ht.data(), (int)k, ref.data()) == 0);
            const double err = relErr(cht, ref);
            double hmax = 0.0, hdiff = 0.0;
            for (size_t i = 0; i < k; ++i)
                for (size_t j = 0; j < k; ++j) {
                    double s = 0.0;
                    for (size_t q = 0; q < n; ++q) s += Ht[i * n + q] * ref[j * n + q];
                    hmax = std::max(hmax, std::fabs(s));
                    hdiff = std::max(hdiff, std::fabs(s - hcht[i * k + j]));
                }
            std::printf("  %zu direction(s): C H^T rel err %.3e, H C H^T rel err %.3e\n", k, err, hdiff / hmax);
            EXPECT_TRUE(hmax > 0.0 && err < 1e-9 && hdiff < 1e-9 * hmax);
        }
        Math::Vector cht;
        EXPECT_THROW_ASSERT(fem->applyComplianceToConstraint(57, {0.5, 0.5}, {{{1, 0, 0}}}, 1.0, &cht, nullptr));
        EXPECT_THROW_ASSERT(fem->applyComplianceToConstraint(1000, coord, {{{1, 0, 0}}}, 1.0, &cht, nullptr));
        orc_sim_destroy(o);
    });
    run("Fem3DRepresentationGpuTest.ExternalForceLastsOneUpdate", [] {
        const Block b = makeBlock(2, 2, 2);
        const Material m{5e4, 0.45, 1000.0};
        auto fem = makeRepresentation(b, kCoro, m, {});
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        fem->setLinearSolver(Math::LINEARSOLVER_CONJUGATEGRADIENT);
        fem->initialize();
        std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(fem->getOdeSolver()->getLinearSolver())
            ->setTolerance(1e-12);
        const int top = b.id(1, 2, 1);
        fem->addExternalGeneralizedForce(top, {{0.5, -1.0, 0.25}});
        const double dt = 1e-3;
        fem->update(dt);
        fem->afterUpdate(dt);
        for (double f : fem->getExternalGeneralizedForce()) EXPECT_TRUE(f == 0.0);
        fem->update(dt);
        fem->afterUpdate(dt);
        void* o = makeOracle(b, 1, m, {}, 0, 1, 1e-12);
        std::vector<double> f(b.x0.size(), 0.0);
        f[3 * top] = 0.5;
        f[3 * top + 1] = -1.0;
        f[3 * top + 2] = 0.25;
        orc_sim_set_external_force(o, f.data());
        orc_sim_step(o, dt, 1);
        orc_sim_set_external_force(o, nullptr);
        orc_sim_step(o, dt, 1);
        std::vector<double> xo(b.x0.size()), vo(b.x0.size());
        orc_sim_get_state(o, xo.data(), vo.data());
        const double ev = relErr(fem->getCurrentState()->getVelocities(), vo);
        std::printf("  velocity rel err %.3e\n", ev);
        EXPECT_TRUE(ev < 1e-9);
        orc_sim_destroy(o);
    });
    run("Fem3DRepresentationGpuTest.NonFiniteStateDeactivates", [] {
        const Block b = makeBlock(1, 1, 1);
        const Material m{5e4, 0.45, 1000.0};
        std::vector<double> v0(3 * b.numNodes(), 0.0);
        v0[3 * b.id(1, 1, 1) + 1] = std::nan("");
        auto fem = makeRepresentation(b, kCoro, m, v0);
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        fem->setLinearSolver(Math::LINEARSOLVER_CONJUGATEGRADIENT);
        fem->initialize();
        EXPECT_TRUE(fem->isActive());
        fem->update(1e-3);
        EXPECT_TRUE(!fem->isActive());
        EXPECT_TRUE(!fem->getCurrentState()->isValid());
        fem->update(1e-3);  // inactive: nothing happens
        EXPECT_TRUE(fem->getStatistics().steps == 1);
    });
    run("Fem3DRepresentationGpuTest.WoundMeshFromPlyMatchesOracle", [] {
        // the reference's wound_deformable.ply (391 nodes, 1476 tets, 79 fixed nodes, E 7.5e4, nu 0.45,
        // rho 1000) as corotational tets, 30 implicit steps with a downward push on one free node
        Physics::Fem3DRepresentation fem("wound");
        fem.setFemElementType(kCoro);
        fem.loadFem(dataPath("wound_deformable.ply"));
        fem.setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        fem.setLinearSolver(Math::LINEARSOLVER_CONJUGATEGRADIENT);
        fem.initialize();
        std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(fem.getOdeSolver()->getLinearSolver())
            ->setTolerance(1e-12);
        const auto mesh = fem.getFem();
        std::vector<int> conn;
        for (const auto& e : mesh->getElements())
            for (size_t id : e.nodeIds) conn.push_back((int)id);
        const int ne = (int)mesh->getNumElements(), nn = (int)mesh->getNumVertices();
        std::vector<double> E(ne, 7.5e4), nu(ne, 0.45), rho(ne, 1000.0);
        void* o = orc_sim_create(1, nn, mesh->getVertexPositions().data(), ne, conn.data(), E.data(), nu.data(),
                                 rho.data());
        for (size_t b : mesh->getBoundaryConditions())
            for (int c = 0; c < 3; ++c) orc_sim_add_bc_dof(o, 3 * (int)b + c);
        orc_sim_set_solver(o, 0, 1, 1e-12, -1, 1, 0);
        size_t push = 0;
        while (fem.getInitialState()->isBoundaryCondition(3 * push)) ++push;
        std::vector<double> f(3 * nn, 0.0);
        f[3 * push + 1] = -0.5;
        const double dt = 1e-3;
        for (int s = 0; s < 30; ++s) {
            fem.addExternalGeneralizedForce(push, {{0.0, -0.5, 0.0}});
            fem.update(dt);
            fem.afterUpdate(dt);
        }
        orc_sim_set_external_force(o, f.data());
        EXPECT_TRUE(orc_sim_step(o, dt, 30) == 0);
        std::vector<double> xo(3 * nn), vo(3 * nn), u(3 * nn), uo(3 * nn);
        orc_sim_get_state(o, xo.data(), vo.data());
        const auto& x = fem.getCurrentState()->getPositions();
        for (int i = 0; i < 3 * nn; ++i) {
            u[i] = x[i] - mesh->getVertexPositions()[i];
            uo[i] = xo[i] - mesh->getVertexPositions()[i];
        }
        const double eu = relErr(u, uo), ev = relErr(fem.getCurrentState()->getVelocities(), vo);
        std::printf("  wound mesh: rel err u %.3e v %.3e\n", eu, ev);
        EXPECT_TRUE(eu < 1e-9 && ev < 1e-9);
        std::vector<int32_t> rowptr, colidx;
        fem.getPattern(&rowptr, &colidx);
        std::vector<int> outer(rowptr.size()), inner((size_t)orc_sim_nnz(o));
        orc_sim_pattern(o, outer.data(), inner.data());
        EXPECT_TRUE(std::equal(outer.begin(), outer.end(), rowptr.begin()) && inner.size() == colidx.size() &&
                    std::equal(inner.begin(), inner.end(), colidx.begin()));
        orc_sim_destroy(o);
    });
    run("Fem3DRepresentationGpuTest.DegenerateElementRejected", [] {
        Block b = makeBlock(1, 1, 1);
        b.x0[3 * b.id(1, 1, 1)] = b.x0[3 * b.id(0, 1, 1)];  // collapse the cube: tets lose volume
        b.x0[3 * b.id(1, 1, 0)] = b.x0[3 * b.id(0, 1, 0)];
        b.x0[3 * b.id(1, 0, 1)] = b.x0[3 * b.id(0, 0, 1)];
        b.x0[3 * b.id(1, 0, 0)] = b.x0[3 * b.id(0, 0, 0)];
        const Material m{5e4, 0.45, 1000.0};
        auto fem = makeRepresentation(b, kCoro, m, {});
        fem->setIntegrationScheme(Math::INTEGRATIONSCHEME_EULER_IMPLICIT);
        EXPECT_THROW_ASSERT(fem->initialize());
        EXPECT_TRUE(!fem->isInitialized());
    });
}

}  // namespace

int main(int argc, char** argv) {
    const bool cpu = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    if (!cpu && !gpu) {
        std::printf("usage: %s --cpu | --gpu [ply data directory]\n", argv[0]);
        return 2;
    }
    if (argc > 2) g_data_dir = argv[2];
    if (cpu) cpuTests();
    if (gpu) gpuTests();
    std::printf("%s: %d failure(s)\n", cpu ? "cpu" : "gpu", g_failures);
    return g_failures == 0 ? 0 : 1;
}
