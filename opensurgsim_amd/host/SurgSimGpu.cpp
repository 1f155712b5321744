// C++ host adapters over libfem_gpu.so (see SurgSimGpu.h).
#include "SurgSimGpu.h"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace SurgSimGpu {

namespace {

const std::string kTetName = "SurgSim::Physics::Fem3DElementTetrahedron";
const std::string kCoroName = "SurgSim::Physics::Fem3DElementCorotationalTetrahedron";
const std::string kCubeName = "SurgSim::Physics::Fem3DElementCube";

[[noreturn]] void fail(const std::string& message) { throw Framework::AssertionFailure(message); }

int elementKind(const std::string& type) {
    if (type == kTetName) return FEM_ELEMENT_TET_LINEAR;
    if (type == kCoroName) return FEM_ELEMENT_TET_COROTATIONAL;
    if (type == kCubeName) return FEM_ELEMENT_HEX;
    return -1;
}

const char* schemeName(Math::IntegrationScheme s) {
    switch (s) {
        case Math::INTEGRATIONSCHEME_STATIC: return "Ode Solver Static";
        case Math::INTEGRATIONSCHEME_LINEAR_STATIC: return "Ode Solver Linear Static";
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT: return "Ode Solver Euler Explicit";
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT: return "Ode Solver Linear Euler Explicit";
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED: return "Ode Solver Euler Explicit Modified";
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED: return "Ode Solver Linear Euler Explicit Modified";
        case Math::INTEGRATIONSCHEME_EULER_IMPLICIT: return "Ode Solver Euler Implicit";
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT: return "Ode Solver Linear Euler Implicit";
        case Math::INTEGRATIONSCHEME_RUNGE_KUTTA_4: return "Ode Solver Runge Kutta 4";
        case Math::INTEGRATIONSCHEME_LINEAR_RUNGE_KUTTA_4: return "Ode Solver Linear Runge Kutta 4";
        default: return "";
    }
}

// The library's numbering of Math::IntegrationScheme (Math/OdeSolver.h:36-49).
int32_t schemeCode(Math::IntegrationScheme s) {
    switch (s) {
        case Math::INTEGRATIONSCHEME_STATIC: return FEM_SCHEME_STATIC;
        case Math::INTEGRATIONSCHEME_LINEAR_STATIC: return FEM_SCHEME_LINEAR_STATIC;
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT: return FEM_SCHEME_EULER_EXPLICIT;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT: return FEM_SCHEME_LINEAR_EULER_EXPLICIT;
        case Math::INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED: return FEM_SCHEME_EULER_EXPLICIT_MODIFIED;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED: return FEM_SCHEME_LINEAR_EULER_EXPLICIT_MODIFIED;
        case Math::INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT: return FEM_SCHEME_LINEAR_EULER_IMPLICIT;
        case Math::INTEGRATIONSCHEME_RUNGE_KUTTA_4: return FEM_SCHEME_RUNGE_KUTTA_4;
        case Math::INTEGRATIONSCHEME_LINEAR_RUNGE_KUTTA_4: return FEM_SCHEME_LINEAR_RUNGE_KUTTA_4;
        default: return FEM_SCHEME_EULER_IMPLICIT;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Math::OdeState
// ---------------------------------------------------------------------------------------------
namespace Math {

bool OdeState::operator==(const OdeState& o) const {
    return m_x == o.m_x && m_v == o.m_v && m_bcMask == o.m_bcMask;
}

void OdeState::reset() {
    std::fill(m_x.begin(), m_x.end(), 0.0);
    std::fill(m_v.begin(), m_v.end(), 0.0);
    std::fill(m_bcMask.begin(), m_bcMask.end(), 0);
    m_bcDofs.clear();
}

void OdeState::setNumDof(size_t numDofPerNode, size_t numNodes) {
    if (numDofPerNode == 0 || numNodes == 0) fail("OdeState::setNumDof: number of DOF per node and nodes must be > 0");
    m_numDofPerNode = numDofPerNode;
    m_numNodes = numNodes;
    const size_t n = numDofPerNode * numNodes;
    m_x.assign(n, 0.0);
    m_v.assign(n, 0.0);
    m_bcMask.assign(n, 0);
    m_bcDofs.clear();
}

Vector3d OdeState::getPosition(size_t nodeId) const {
    if (m_numDofPerNode != 3 || nodeId >= m_numNodes) fail("OdeState::getPosition: invalid node id");
    return {{m_x[3 * nodeId], m_x[3 * nodeId + 1], m_x[3 * nodeId + 2]}};
}

Vector3d OdeState::getVelocity(size_t nodeId) const {
    if (m_numDofPerNode != 3 || nodeId >= m_numNodes) fail("OdeState::getVelocity: invalid node id");
    return {{m_v[3 * nodeId], m_v[3 * nodeId + 1], m_v[3 * nodeId + 2]}};
}

void OdeState::addBoundaryCondition(size_t nodeId) {
    if (m_numDofPerNode == 0) fail("Number of dof per node = 0. Make sure to call setNumDof() prior to adding boundary conditions.");
    if (nodeId >= m_numNodes) fail("Invalid nodeId " + std::to_string(nodeId) + " number of nodes is " +
                                   std::to_string(m_numNodes));
    for (size_t d = 0; d < m_numDofPerNode; ++d) addBoundaryCondition(nodeId, d);
}

void OdeState::addBoundaryCondition(size_t nodeId, size_t nodeDofId) {
    if (m_numDofPerNode == 0) fail("Number of dof per node = 0. Make sure to call setNumDof() prior to adding boundary conditions.");
    if (nodeId >= m_numNodes) fail("Invalid nodeId " + std::to_string(nodeId) + " number of nodes is " +
                                   std::to_string(m_numNodes));
    if (nodeDofId >= m_numDofPerNode) fail("Invalid nodeDofId " + std::to_string(nodeDofId));
    const size_t dof = m_numDofPerNode * nodeId + nodeDofId;
    if (!m_bcMask[dof]) {
        m_bcMask[dof] = 1;
        m_bcDofs.push_back(dof);
    }
}

bool OdeState::isBoundaryCondition(size_t dof) const {
    if (dof >= m_bcMask.size()) fail("Invalid dof " + std::to_string(dof));
    return m_bcMask[dof] != 0;
}

Vector* OdeState::applyBoundaryConditionsToVector(Vector* vector) const {
    if (!vector || vector->size() != m_x.size()) fail("Invalid vector to apply boundary conditions on");
    for (size_t d : m_bcDofs) (*vector)[d] = 0.0;
    return vector;
}

bool OdeState::isValid() const {
    for (double a : m_x)
        if (!std::isfinite(a)) return false;
    for (double a : m_v)
        if (!std::isfinite(a)) return false;
    return true;
}

void LinearSparseSolveAndInverseCG::setTolerance(double tolerance) {
    if (!(tolerance > 0.0)) fail("LinearSparseSolveAndInverseCG: tolerance must be positive");
    m_tolerance = tolerance;
}

OdeSolver::OdeSolver(IntegrationScheme scheme) : m_scheme(scheme), m_name(schemeName(scheme)) {
    if (m_name.empty())
        fail("Integration scheme " + std::to_string((int)scheme) + " is unknown");
}

void OdeSolver::setLinearSolver(std::shared_ptr<LinearSparseSolveAndInverse> linearSolver) {
    if (!linearSolver) fail("OdeSolver::setLinearSolver: null solver");
    m_linearSolver = std::move(linearSolver);
}

void OdeSolver::setNewtonRaphsonMaximumIteration(size_t maxIterations) {
    if (maxIterations < 1) fail("Newton-Raphson needs at least one iteration");
    m_newtonMaxIterations = maxIterations;
}

void OdeSolver::setNewtonRaphsonEpsilonConvergence(double epsilon) {
    if (!(epsilon >= 0.0)) fail("The epsilon convergence cannot be negative");
    m_newtonEpsilon = epsilon;
}

}  // namespace Math

// ---------------------------------------------------------------------------------------------
// Physics
// ---------------------------------------------------------------------------------------------
namespace Physics {

void FemElement::checkMaterial() const {
    // Physics/FemElement.cpp:33-43
    if (!(m_rho > 0.0)) fail("Fem element mass density must be positive, " + std::to_string(m_rho));
    if (!(m_E > 0.0)) fail("Fem element Young's modulus must be positive, " + std::to_string(m_E));
    if (!(m_nu >= 0.0 && m_nu < 0.5)) fail("Fem element Poisson ratio must be in [0, 0.5[, " + std::to_string(m_nu));
}

Fem3DElementTetrahedron::Fem3DElementTetrahedron(std::array<size_t, 4> ids) : FemElement({ids.begin(), ids.end()}) {}
const std::string& Fem3DElementTetrahedron::getClassName() const { return kTetName; }

Fem3DElementCorotationalTetrahedron::Fem3DElementCorotationalTetrahedron(std::array<size_t, 4> ids)
    : FemElement({ids.begin(), ids.end()}) {}
const std::string& Fem3DElementCorotationalTetrahedron::getClassName() const { return kCoroName; }

Fem3DElementCube::Fem3DElementCube(std::array<size_t, 8> ids) : FemElement({ids.begin(), ids.end()}) {}
const std::string& Fem3DElementCube::getClassName() const { return kCubeName; }

Fem3DRepresentation::Fem3DRepresentation(const std::string& name, int device)
    : m_name(name), m_device(device) {}

Fem3DRepresentation::~Fem3DRepresentation() {
    if (m_ctx) fem_destroy(m_ctx);
}

void Fem3DRepresentation::check(int status, const char* what) const {
    if (status < 0) {
        std::ostringstream os;
        os << m_name << ": " << what << " failed (" << status << "): " << (m_ctx ? fem_last_error(m_ctx) : "");
        fail(os.str());
    }
}

void Fem3DRepresentation::setFemElementType(const std::string& type) {
    if (isInitialized()) fail("You must set the FemElement type before initializing");
    if (elementKind(type) < 0) fail("FemElement type '" + type + "' is not provided by the GPU path");
    m_elementType = type;
}

void Fem3DRepresentation::addFemElement(std::shared_ptr<FemElement> element) {
    if (isInitialized()) fail("Elements must be added before initializing");
    if (!element) fail("Null FemElement");
    m_elements.push_back(std::move(element));
}

void Fem3DRepresentation::setInitialState(std::shared_ptr<Math::OdeState> initialState) {
    if (!initialState || initialState->getNumDof() == 0) fail("Invalid initial state");
    if (initialState->getNumDofPerNode() != 3) fail("Fem3D states have 3 DOF per node");
    m_initialState = initialState;
    m_currentState = std::make_shared<Math::OdeState>(*initialState);
    m_previousState = std::make_shared<Math::OdeState>(*initialState);
    m_finalState = std::make_shared<Math::OdeState>(*initialState);
    m_externalForce.assign(initialState->getNumDof(), 0.0);
    m_hasExternalForce = false;
    m_currentStale = false;
    if (m_ctx) check(fem_set_state(m_ctx, initialState->getPositions().data(), initialState->getVelocities().data()),
                     "setInitialState");
}

std::shared_ptr<Math::OdeState> Fem3DRepresentation::getCurrentState() const {
    if (m_currentStale) fetchCurrentState();
    return m_currentState;
}

void Fem3DRepresentation::fetchCurrentState() const {
    check(fem_get_state(m_ctx, m_currentState->getPositions().data(), m_currentState->getVelocities().data()),
          "getCurrentState");
    m_currentStale = false;
}

void Fem3DRepresentation::setIntegrationScheme(Math::IntegrationScheme scheme) {
    if (isInitialized()) fail("You must set the integration scheme before initializing");
    m_scheme = scheme;
}

void Fem3DRepresentation::setLinearSolver(Math::LinearSolver linearSolver) {
    if (isInitialized()) fail("You must set the linear solver before initializing");
    m_linearSolverType = linearSolver;
}

void Fem3DRepresentation::addExternalGeneralizedForce(size_t nodeId, const Math::Vector3d& force) {
    if (!m_initialState) fail("Set the initial state before adding external forces");
    if (nodeId >= m_initialState->getNumNodes()) fail("Invalid node id " + std::to_string(nodeId));
    for (int c = 0; c < 3; ++c) m_externalForce[3 * nodeId + c] += force[c];
    m_hasExternalForce = true;
}

void Fem3DRepresentation::addExternalGeneralizedForce(const Math::Vector& force) {
    if (!m_initialState) fail("Set the initial state before adding external forces");
    if (force.size() != m_externalForce.size()) fail("External force of the wrong size");
    for (size_t i = 0; i < force.size(); ++i) m_externalForce[i] += force[i];
    m_hasExternalForce = true;
}

void Fem3DRepresentation::addExternalGeneralizedForce(size_t elementId, const Math::Vector& naturalCoordinate,
                                                      const Math::Vector3d& force, const Math::Matrix33d* K,
                                                      const Math::Matrix33d* D) {
    if (!m_ctx) fail("Initialize the representation before adding localized external forces");
    if (elementId >= m_elements.size()) fail("Invalid localization (element id out of range)");
    if (naturalCoordinate.size() != m_elements[elementId]->getNodeIds().size())
        fail("Invalid localization (natural coordinate size)");
    check(fem_add_external_generalized_force(m_ctx, (int64_t)elementId, naturalCoordinate.data(), force.data(),
                                             K ? K->data() : nullptr, D ? D->data() : nullptr),
          "external generalized force");
}

bool Fem3DRepresentation::initialize() {
    if (isInitialized()) fail("Fem3DRepresentation already initialized");
    if (!m_initialState) fail("Initial state has not been set yet. Did you call setInitialState() ?");
    if (m_elements.empty()) fail("No FemElement");
    // no type set: the type of the elements (Physics/Fem3DRepresentation.cpp:109-121 defaults by node count)
    if (m_elementType.empty()) m_elementType = m_elements[0]->getClassName();
    const int kind = elementKind(m_elementType);
    const size_t nn = kind == FEM_ELEMENT_HEX ? 8 : 4;
    const size_t numNodes = m_initialState->getNumNodes();
    std::vector<int32_t> conn(m_elements.size() * nn);
    std::vector<double> E(m_elements.size()), nu(m_elements.size()), rho(m_elements.size());
    for (size_t e = 0; e < m_elements.size(); ++e) {
        const FemElement& el = *m_elements[e];
        if (el.getClassName() != m_elementType)
            fail("Element " + std::to_string(e) + " is a " + el.getClassName() + ", the representation uses " +
                 m_elementType);
        if (el.getNumNodes() != nn) fail("Element " + std::to_string(e) + " has the wrong number of nodes");
        el.checkMaterial();
        for (size_t a = 0; a < nn; ++a) {
            const size_t id = el.getNodeId(a);
            if (id >= numNodes) fail("Element " + std::to_string(e) + " references node " + std::to_string(id));
            conn[e * nn + a] = (int32_t)id;
        }
        E[e] = el.getYoungModulus();
        nu[e] = el.getPoissonRatio();
        rho[e] = el.getMassDensity();
    }
    // ODE solver and linear solver as DeformableRepresentation::doInitialize builds them
    m_odeSolver = std::make_shared<Math::OdeSolver>(m_scheme);
    if (m_linearSolverType == Math::LINEARSOLVER_LU)
        m_odeSolver->setLinearSolver(std::make_shared<Math::LinearSparseSolveAndInverseLU>());
    else if (m_linearSolverType == Math::LINEARSOLVER_CONJUGATEGRADIENT)
        m_odeSolver->setLinearSolver(std::make_shared<Math::LinearSparseSolveAndInverseCG>());
    else
        fail("Linear solver not initialized, the linear solver is invalid");

    fem_ctx* ctx = nullptr;
    if (fem_create(m_device, &ctx) != FEM_OK || !ctx) {
        if (ctx) fem_destroy(ctx);
        fail(m_name + ": could not create a GPU context on device " + std::to_string(m_device));
    }
    m_ctx = ctx;
    try {
        check(fem_upload_mesh(m_ctx, kind, (int64_t)numNodes, m_initialState->getPositions().data(),
                              (int64_t)m_elements.size(), conn.data(), E.data(), nu.data(), rho.data()),
              "initialize");
        for (size_t dof : m_initialState->getBoundaryConditions())
            check(fem_add_boundary_condition_dof(m_ctx, (int64_t)dof), "boundary condition");
        check(fem_set_state(m_ctx, m_initialState->getPositions().data(), m_initialState->getVelocities().data()),
              "initial state");
    } catch (...) {
        fem_destroy(m_ctx);
        m_ctx = nullptr;
        throw;
    }
    std::vector<double> mass(numNodes);
    check(fem_get_node_masses(m_ctx, mass.data()), "node masses");
    m_totalMass = 0.0;
    for (double m : mass) m_totalMass += m;
    return true;
}

void Fem3DRepresentation::applyOptions(double dt) {
    fem_options o;
    fem_default_options(&o);
    o.dt = dt;
    o.gravity[0] = m_gravity[0];
    o.gravity[1] = m_gravity[1];
    o.gravity[2] = m_gravity[2];
    o.gravity_enabled = m_gravityEnabled ? 1 : 0;
    o.scheme = schemeCode(m_scheme);
    o.rayleigh_mass = m_rayleighMass;
    o.rayleigh_stiffness = m_rayleighStiffness;
    o.mass_lumping = m_massLumping ? 1 : 0;
    o.precision = (int32_t)m_precision;
    o.newton_max_iterations = (int32_t)m_odeSolver->getNewtonRaphsonMaximumIteration();
    o.newton_epsilon = m_odeSolver->getNewtonRaphsonEpsilonConvergence();
    auto solver = m_odeSolver->getLinearSolver();
    if (auto cg = std::dynamic_pointer_cast<Math::LinearSparseSolveAndInverseCG>(solver)) {
        o.cg_tolerance = cg->getTolerance();
        o.cg_max_iterations = cg->getMaxIterations();
    } else if (elementKind(m_elementType) != FEM_ELEMENT_HEX) {
        o.linear_solver = FEM_LINEAR_SOLVER_LU;  // the banded LU of the assembled operator
    } else {
        // hexahedra keep no assembled operator (element-by-element products): CG to 1e-13 stands in for the LU
        o.cg_tolerance = 1e-13;
        o.cg_max_iterations = -1;
    }
    check(fem_set_options(m_ctx, &o), "options");
    m_lastDt = dt;
}

void Fem3DRepresentation::update(double dt) {
    if (!isActive()) return;
    if (!m_odeSolver) fail("Ode solver has not been set yet. Did you call initialize() ?");
    if (!m_initialState) fail("Initial state has not been set yet. Did you call setInitialState() ?");
    applyOptions(dt);
    if (m_hasExternalForce) {
        check(fem_set_external_force(m_ctx, m_externalForce.data()), "external force");
        m_uploadedExternalForce = true;
    } else if (m_uploadedExternalForce) {
        check(fem_set_external_force(m_ctx, nullptr), "external force");
        m_uploadedExternalForce = false;
    }
    const int rc = fem_step(m_ctx, 1);
    check(rc, "update");
    if (m_syncStates) {
        if (m_currentStale) fetchCurrentState();
        m_currentState.swap(m_previousState);
        fetchCurrentState();
    } else {
        m_currentStale = true;
    }
    if (rc == FEM_INACTIVE) setLocalActive(false);  // the new state is not finite
}

void Fem3DRepresentation::afterUpdate(double) {
    if (!isActive()) return;
    *m_finalState = *getCurrentState();
    if (m_hasExternalForce) {
        std::fill(m_externalForce.begin(), m_externalForce.end(), 0.0);
        m_hasExternalForce = false;
    }
}

Math::Vector Fem3DRepresentation::applyCompliance(const Math::Vector& b, size_t numColumns) const {
    if (!m_ctx) fail("Ode solver not initialized, it should have been initialized on wake-up");
    const size_t n = m_initialState->getNumDof();
    if (b.size() != n * numColumns) fail("applyCompliance: b needs getNumDof() values per column");
    Math::Vector out(b.size(), 0.0);
    check(fem_apply_compliance(m_ctx, b.data(), (int64_t)numColumns, out.data()), "applyCompliance");
    return out;
}

Math::Vector Fem3DRepresentation::getComplianceMatrix() const {
    if (!m_ctx) fail("Ode solver not initialized, it should have been initialized on wake-up");
    const size_t n = m_initialState->getNumDof();
    Math::Vector out(n * n, 0.0);
    check(fem_get_compliance_matrix(m_ctx, out.data()), "getComplianceMatrix");
    return out;
}

void Fem3DRepresentation::applyComplianceToConstraint(size_t elementId, const Math::Vector& naturalCoordinate,
                                                      const std::vector<Math::Vector3d>& directions, double scale,
                                                      Math::Vector* cht, Math::Vector* hcht) const {
    if (!m_ctx) fail("Ode solver not initialized, it should have been initialized on wake-up");
    if (elementId >= m_elements.size()) fail("Invalid localization (element id out of range)");
    if (naturalCoordinate.size() != m_elements[elementId]->getNodeIds().size())
        fail("Invalid localization (natural coordinate size)");
    const size_t m = directions.size();
    std::vector<double> dirs;
    for (const auto& d : directions) dirs.insert(dirs.end(), d.begin(), d.end());
    if (cht) cht->assign(m_initialState->getNumDof() * m, 0.0);
    if (hcht) hcht->assign(m * m, 0.0);
    check(fem_constraint_compliance(m_ctx, (int64_t)elementId, naturalCoordinate.data(), (int)m, dirs.data(), scale,
                                    cht ? cht->data() : nullptr, hcht ? hcht->data() : nullptr),
          "constraint compliance");
}

fem_stats Fem3DRepresentation::getStatistics() const {
    if (!m_ctx) fail("Not initialized");
    fem_stats s;
    check(fem_query(m_ctx, &s), "statistics");
    return s;
}

void Fem3DRepresentation::getPattern(std::vector<int32_t>* rowptr, std::vector<int32_t>* colidx) const {
    if (!m_ctx) fail("Not initialized");
    rowptr->assign(getNumDof() + 1, 0);
    colidx->assign((size_t)fem_num_nonzeros(m_ctx), 0);
    check(fem_get_pattern(m_ctx, rowptr->data(), colidx->data()), "pattern");
}

}  // namespace Physics
}  // namespace SurgSimGpu
