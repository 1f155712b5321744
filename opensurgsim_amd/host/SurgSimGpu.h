// C++ host adapters with the reference's class names, over the C ABI of libfem_gpu.so.
//
// A Fem3DRepresentation here is configured like the reference's (Physics/Fem3DRepresentation.h,
// Physics/FemRepresentation.h, Physics/DeformableRepresentation.h): element type by factory name,
// elements with node ids and material, an initial OdeState with boundary conditions, an integration
// scheme and a linear solver, then update(dt) / afterUpdate(dt) per time step. The state lives in
// GPU memory; OdeState objects on the host mirror it after each update (configurable).
//
// Error behaviour follows the reference: precondition violations throw
// SurgSimGpu::Framework::AssertionFailure (a std::runtime_error, as SurgSim::Framework::AssertionFailure,
// Framework/Assert.h:77-95); a state that becomes non-finite deactivates the representation
// (Physics/FemRepresentation.cpp:285-293).
//
// Only what the stepping path and its compliance consumers need is mirrored. The dense compliance matrix is
// formed on request (getComplianceMatrix), not on every update; compliance warping and Localization objects are
// not provided (node and element localizations are passed as ids and natural coordinates).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fem_gpu.h"

namespace SurgSimGpu {

namespace Framework {

/// Thrown where the reference's SURGSIM_ASSERT fires (Framework/Assert.h:77-95).
class AssertionFailure : public std::runtime_error {
public:
    explicit AssertionFailure(const std::string& message) : std::runtime_error(message) {}
};

}  // namespace Framework

namespace Math {

/// Math/OdeSolver.h:36-49 (same names and order). Only the two implicit Euler schemes run here.
enum IntegrationScheme {
    INTEGRATIONSCHEME_STATIC = 0,
    INTEGRATIONSCHEME_LINEAR_STATIC,
    INTEGRATIONSCHEME_EULER_EXPLICIT,
    INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT,
    INTEGRATIONSCHEME_EULER_EXPLICIT_MODIFIED,
    INTEGRATIONSCHEME_LINEAR_EULER_EXPLICIT_MODIFIED,
    INTEGRATIONSCHEME_EULER_IMPLICIT,
    INTEGRATIONSCHEME_LINEAR_EULER_IMPLICIT,
    INTEGRATIONSCHEME_RUNGE_KUTTA_4,
    INTEGRATIONSCHEME_LINEAR_RUNGE_KUTTA_4,
    MAX_INTEGRATIONSCHEMES
};

/// Math/LinearSparseSolveAndInverse.h:41-46 (same names and order).
enum LinearSolver { LINEARSOLVER_LU = 0, LINEARSOLVER_CONJUGATEGRADIENT, MAX_LINEARSOLVER };

using Vector = std::vector<double>;
using Vector3d = std::array<double, 3>;
using Matrix33d = std::array<double, 9>;  // row-major

/// Math/OdeState.h:38-207: positions, velocities (node-major interleaved DOFs) and boundary conditions.
class OdeState {
public:
    OdeState() = default;
    bool operator==(const OdeState& other) const;
    bool operator!=(const OdeState& other) const { return !(*this == other); }
    /// Zeroes x, v and removes the boundary conditions (Math/OdeState.cpp:40-53).
    void reset();
    /// Allocates 0-initialised x, v and the BC mask (Math/OdeState.cpp:55-68). Fem3D uses 3 DOF per node.
    void setNumDof(size_t numDofPerNode, size_t numNodes);
    size_t getNumDof() const { return m_x.size(); }
    size_t getNumNodes() const { return m_numNodes; }
    size_t getNumDofPerNode() const { return m_numDofPerNode; }
    Vector& getPositions() { return m_x; }
    const Vector& getPositions() const { return m_x; }
    Vector3d getPosition(size_t nodeId) const;
    Vector& getVelocities() { return m_v; }
    const Vector& getVelocities() const { return m_v; }
    Vector3d getVelocity(size_t nodeId) const;
    /// Fixes every DOF of a node (Math/OdeState.cpp:115-127); duplicates are ignored.
    void addBoundaryCondition(size_t nodeId);
    /// Fixes one DOF of a node (Math/OdeState.cpp:129-142).
    void addBoundaryCondition(size_t nodeId, size_t nodeDofId);
    size_t getNumBoundaryConditions() const { return m_bcDofs.size(); }
    const std::vector<size_t>& getBoundaryConditions() const { return m_bcDofs; }
    bool isBoundaryCondition(size_t dof) const;
    /// Zeroes the BC entries of a vector (Math/OdeState.cpp:159-172).
    Vector* applyBoundaryConditionsToVector(Vector* vector) const;
    /// All positions and velocities finite (Math/OdeState.cpp:268-278).
    bool isValid() const;

private:
    size_t m_numDofPerNode = 0, m_numNodes = 0;
    Vector m_x, m_v;
    std::vector<size_t> m_bcDofs;
    std::vector<uint8_t> m_bcMask;
};

/// Math/LinearSparseSolveAndInverse.h:57-80. The GPU context performs the solve; these objects carry the
/// solver choice and its settings into it.
class LinearSparseSolveAndInverse {
public:
    virtual ~LinearSparseSolveAndInverse() = default;
    virtual LinearSolver getType() const = 0;
};

/// LINEARSOLVER_LU (Math/LinearSparseSolveAndInverse.h:83-95). Served by the GPU CG at a tolerance of
/// 1e-13 (no sparse LU on the device).
class LinearSparseSolveAndInverseLU : public LinearSparseSolveAndInverse {
public:
    LinearSolver getType() const override { return LINEARSOLVER_LU; }
};

/// Math/LinearSparseSolveAndInverse.h:98-123: Jacobi-preconditioned CG with Eigen's defaults
/// (tolerance = machine epsilon, max iterations = 2 x DOF; a negative value keeps the default).
class LinearSparseSolveAndInverseCG : public LinearSparseSolveAndInverse {
public:
    LinearSolver getType() const override { return LINEARSOLVER_CONJUGATEGRADIENT; }
    void setTolerance(double tolerance);
    double getTolerance() const { return m_tolerance; }
    void setMaxIterations(int64_t maxIterations) { m_maxIterations = maxIterations; }
    int64_t getMaxIterations() const { return m_maxIterations; }

private:
    double m_tolerance = 2.220446049250313e-16;
    int64_t m_maxIterations = -1;
};

/// Math/OdeSolver.h:78-159 as far as a caller configures it: scheme, linear solver and the
/// Newton-Raphson settings of OdeSolverEulerImplicit (Math/OdeSolverEulerImplicit.h).
class OdeSolver {
public:
    explicit OdeSolver(IntegrationScheme scheme);
    IntegrationScheme getIntegrationScheme() const { return m_scheme; }
    const std::string& getName() const { return m_name; }
    std::shared_ptr<LinearSparseSolveAndInverse> getLinearSolver() const { return m_linearSolver; }
    void setLinearSolver(std::shared_ptr<LinearSparseSolveAndInverse> linearSolver);
    void setNewtonRaphsonMaximumIteration(size_t maxIterations);
    size_t getNewtonRaphsonMaximumIteration() const { return m_newtonMaxIterations; }
    /// Math/OdeSolverEulerImplicit.cpp:43-48: the iterations stop once ||dv - dv_prev||_inf < epsilon.
    void setNewtonRaphsonEpsilonConvergence(double epsilon);
    double getNewtonRaphsonEpsilonConvergence() const { return m_newtonEpsilon; }

private:
    IntegrationScheme m_scheme;
    std::string m_name;
    std::shared_ptr<LinearSparseSolveAndInverse> m_linearSolver;
    size_t m_newtonMaxIterations = 1;
    double m_newtonEpsilon = 1e-5;
};

}  // namespace Math

namespace Physics {

/// Physics/FemElement.h:45-270, as element data: node ids and material. Computation happens on the GPU.
class FemElement {
public:
    virtual ~FemElement() = default;
    size_t getNumNodes() const { return m_nodeIds.size(); }
    size_t getNodeId(size_t elementNodeId) const { return m_nodeIds.at(elementNodeId); }
    const std::vector<size_t>& getNodeIds() const { return m_nodeIds; }
    void setYoungModulus(double E) { m_E = E; }
    double getYoungModulus() const { return m_E; }
    void setPoissonRatio(double nu) { m_nu = nu; }
    double getPoissonRatio() const { return m_nu; }
    void setMassDensity(double rho) { m_rho = rho; }
    double getMassDensity() const { return m_rho; }
    /// The reference's factory name of the element class.
    virtual const std::string& getClassName() const = 0;
    /// Material preconditions of Physics/FemElement.cpp:33-43.
    void checkMaterial() const;

protected:
    explicit FemElement(std::vector<size_t> nodeIds) : m_nodeIds(std::move(nodeIds)) {}

private:
    std::vector<size_t> m_nodeIds;
    double m_E = 0.0, m_nu = 0.0, m_rho = 0.0;
};

/// Physics/Fem3DElementTetrahedron.h (nodes ABC counter-clockwise seen from D).
class Fem3DElementTetrahedron : public FemElement {
public:
    explicit Fem3DElementTetrahedron(std::array<size_t, 4> nodeIds);
    const std::string& getClassName() const override;
};

/// Physics/Fem3DElementCorotationalTetrahedron.h.
class Fem3DElementCorotationalTetrahedron : public FemElement {
public:
    explicit Fem3DElementCorotationalTetrahedron(std::array<size_t, 4> nodeIds);
    const std::string& getClassName() const override;
};

/// Physics/Fem3DElementCube.h (node order of Physics/Fem3DElementCube.cpp:72-77).
class Fem3DElementCube : public FemElement {
public:
    explicit Fem3DElementCube(std::array<size_t, 8> nodeIds);
    const std::string& getClassName() const override;
};

/// Fem3D mesh asset (Physics/Fem3D.h, Physics/Fem-inl.h:90-111): vertices, elements with node ids and
/// material, boundary-condition vertices. load() reads the Fem3D PLY format (Physics/FemPlyFormat.dox,
/// Physics/FemPlyReaderDelegate.cpp:31-145, Physics/Fem3DPlyReaderDelegate.cpp:52-96), ASCII only.
class Fem3D {
public:
    struct Element {
        std::vector<size_t> nodeIds;
        double massDensity = 0.0, poissonRatio = 0.0, youngModulus = 0.0;
    };
    /// Throws AssertionFailure for a file that is not an acceptable Fem3D PLY file.
    void load(const std::string& fileName);
    size_t getNumVertices() const { return m_vertices.size() / 3; }
    Math::Vector3d getVertexPosition(size_t i) const {
        return {{m_vertices.at(3 * i), m_vertices.at(3 * i + 1), m_vertices.at(3 * i + 2)}};
    }
    const std::vector<double>& getVertexPositions() const { return m_vertices; }
    size_t getNumElements() const { return m_elements.size(); }
    const Element& getElement(size_t i) const { return m_elements.at(i); }
    const std::vector<Element>& getElements() const { return m_elements; }
    const std::vector<size_t>& getBoundaryConditions() const { return m_boundaryConditions; }
    size_t getBoundaryCondition(size_t i) const { return m_boundaryConditions.at(i); }

    void addVertex(const Math::Vector3d& p) { m_vertices.insert(m_vertices.end(), p.begin(), p.end()); }
    void addElement(const Element& e) { m_elements.push_back(e); }
    void addBoundaryCondition(size_t vertex) { m_boundaryConditions.push_back(vertex); }

private:
    std::vector<double> m_vertices;
    std::vector<Element> m_elements;
    std::vector<size_t> m_boundaryConditions;
};

/// Arithmetic of the GPU step (no reference counterpart: the reference is FP64 throughout).
enum class Precision { FP64 = FEM_PRECISION_FP64, FP32_REFINED = FEM_PRECISION_FP32_REFINED };

/// Physics/Fem3DRepresentation.h over FemRepresentation / DeformableRepresentation, stepping on a GPU.
class Fem3DRepresentation {
public:
    explicit Fem3DRepresentation(const std::string& name, int device = 0);
    ~Fem3DRepresentation();
    Fem3DRepresentation(const Fem3DRepresentation&) = delete;
    Fem3DRepresentation& operator=(const Fem3DRepresentation&) = delete;

    const std::string& getName() const { return m_name; }

    /// Physics/Fem3DRepresentation.cpp:70-126: loads a Fem3D PLY file, builds the initial state (positions and
    /// boundary conditions) and one element per mesh element of the representation's element type (default:
    /// Fem3DElementTetrahedron for 4 nodes, Fem3DElementCube for 8) with the mesh's material.
    void loadFem(const std::string& fileName);
    void setFem(std::shared_ptr<Fem3D> mesh);
    std::shared_ptr<Fem3D> getFem() const { return m_fem; }

    /// Physics/FemRepresentation.cpp:73-77: one of the three factory names; locked after initialize().
    /// Left empty, initialize() takes the class of the elements.
    void setFemElementType(const std::string& type);
    const std::string& getFemElementType() const { return m_elementType; }
    /// Physics/FemRepresentation.cpp:79-83.
    void addFemElement(std::shared_ptr<FemElement> element);
    size_t getNumFemElements() const { return m_elements.size(); }
    std::shared_ptr<FemElement> getFemElement(size_t femElementId) const { return m_elements.at(femElementId); }

    /// Physics/DeformableRepresentation.cpp:86-119: copies the state into initial, current, previous,
    /// new and final states.
    void setInitialState(std::shared_ptr<Math::OdeState> initialState);
    std::shared_ptr<Math::OdeState> getInitialState() const { return m_initialState; }
    std::shared_ptr<Math::OdeState> getCurrentState() const;
    std::shared_ptr<Math::OdeState> getPreviousState() const { return m_previousState; }
    std::shared_ptr<Math::OdeState> getFinalState() const { return m_finalState; }
    size_t getNumDof() const { return m_initialState ? m_initialState->getNumDof() : 0; }
    size_t getNumDofPerNode() const { return 3; }

    /// Physics/DeformableRepresentation.cpp:121-139; locked after initialize().
    void setIntegrationScheme(Math::IntegrationScheme scheme);
    Math::IntegrationScheme getIntegrationScheme() const { return m_scheme; }
    void setLinearSolver(Math::LinearSolver linearSolver);
    Math::LinearSolver getLinearSolver() const { return m_linearSolverType; }
    /// Valid after initialize() (Physics/DeformableRepresentation.cpp:333-385).
    std::shared_ptr<Math::OdeSolver> getOdeSolver() const { return m_odeSolver; }

    /// Physics/Representation.cpp:36-38: gravity (0, -9.81, 0), enabled by default.
    void setIsGravityEnabled(bool enabled) { m_gravityEnabled = enabled; }
    bool isGravityEnabled() const { return m_gravityEnabled; }
    void setGravity(const Math::Vector3d& g) { m_gravity = g; }

    /// Node localization of DeformableRepresentation::addExternalGeneralizedForce
    /// (Physics/Fem3DRepresentation.cpp:133-198): f is added to the node's three DOFs until afterUpdate.
    void addExternalGeneralizedForce(size_t nodeId, const Math::Vector3d& force);
    /// A full generalized force vector (getNumDof() values), added until afterUpdate.
    void addExternalGeneralizedForce(const Math::Vector& force);
    /// Element localization (Physics/Fem3DRepresentation.cpp:133-198): force spread over the element's nodes
    /// with its natural coordinates, optional stiffness K and damping D on the node blocks; applied by the
    /// next update() only. K and D need tetrahedra and a scheme other than the linear variants.
    void addExternalGeneralizedForce(size_t elementId, const Math::Vector& naturalCoordinate,
                                     const Math::Vector3d& force, const Math::Matrix33d* K = nullptr,
                                     const Math::Matrix33d* D = nullptr);

    /// DeformableRepresentation::applyCompliance (Physics/DeformableRepresentation.cpp:175-190) on numColumns
    /// columns of b (getNumDof() values each, column-major): boundary rows zeroed, one CG solve per column
    /// against the last update's system matrix, boundary rows of the result zeroed.
    Math::Vector applyCompliance(const Math::Vector& b, size_t numColumns) const;
    /// OdeSolver::getComplianceMatrix after computeComplianceMatrixFromSystemMatrix (Math/OdeSolver.cpp:69-76):
    /// the inverse of the last update's system matrix with boundary rows and columns zero, getNumDof()^2
    /// values, column-major (tetrahedral meshes).
    Math::Vector getComplianceMatrix() const;
    /// The compliance terms a constraint localized in one element hands to the MLCP
    /// (Physics/FemConstraintFixedPoint.cpp:84-106, Physics/FemConstraintFrictionlessContact.cpp:84-112,
    /// Physics/MlcpPhysicsProblem.cpp:46-76): row j of H is scale * naturalCoordinate[a] * directions[j] at
    /// node a's DOFs (fixed point: the three axes; frictionless contact: the normal; scale: dt times the
    /// builder's scale). cht (or null) receives C H^T, getNumDof() x directions.size() column-major; hcht (or
    /// null) receives H C H^T, row-major.
    void applyComplianceToConstraint(size_t elementId, const Math::Vector& naturalCoordinate,
                                     const std::vector<Math::Vector3d>& directions, double scale, Math::Vector* cht,
                                     Math::Vector* hcht) const;

    /// Physics/FemRepresentation.cpp:202-220: Rayleigh damping D = massCoef M + stiffnessCoef K.
    void setRayleighDampingMass(double massCoef) { m_rayleighMass = massCoef; }
    double getRayleighDampingMass() const { return m_rayleighMass; }
    void setRayleighDampingStiffness(double stiffnessCoef) { m_rayleighStiffness = stiffnessCoef; }
    double getRayleighDampingStiffness() const { return m_rayleighStiffness; }
    /// Physics/FemRepresentation.cpp:320-327.
    void setMassLumping(bool useMassLumping) { m_massLumping = useMassLumping; }
    bool getMassLumping() const { return m_massLumping; }
    const Math::Vector& getExternalGeneralizedForce() const { return m_externalForce; }

    /// Element initialisation, node masses and the sparsity pattern on the device
    /// (Physics/FemRepresentation.cpp:85-168; the reference's initialize(runtime) + wakeUp()).
    bool initialize();
    bool isInitialized() const { return m_ctx != nullptr; }
    double getTotalMass() const { return m_totalMass; }

    bool isActive() const { return m_active; }
    void setLocalActive(bool active) { m_active = active; }

    /// Physics/FemRepresentation.cpp:232-294 (without compliance computation): one implicit step.
    void update(double dt);
    /// Physics/DeformableRepresentation.cpp:229-253: final state = current state; external force cleared.
    void afterUpdate(double dt);

    // GPU extensions (no reference counterpart).
    void setPrecision(Precision precision) { m_precision = precision; }
    Precision getPrecision() const { return m_precision; }
    /// When false, update() leaves the state on the device: getCurrentState() fetches it on demand and the
    /// previous state is not tracked. Default true (the reference keeps all states on the host).
    void setStateSynchronization(bool synchronize) { m_syncStates = synchronize; }
    fem_stats getStatistics() const;
    /// Block-CSR-expanded scalar pattern (bit-exact with the reference's CSC pattern).
    void getPattern(std::vector<int32_t>* rowptr, std::vector<int32_t>* colidx) const;
    fem_ctx* getContext() const { return m_ctx; }

private:
    void check(int status, const char* what) const;
    void applyOptions(double dt);
    void fetchCurrentState() const;

    std::string m_name;
    int m_device;
    fem_ctx* m_ctx = nullptr;
    std::shared_ptr<Fem3D> m_fem;
    std::string m_elementType;
    std::vector<std::shared_ptr<FemElement>> m_elements;
    std::shared_ptr<Math::OdeState> m_initialState, m_currentState, m_previousState, m_finalState;
    mutable bool m_currentStale = false;
    Math::IntegrationScheme m_scheme = Math::INTEGRATIONSCHEME_EULER_EXPLICIT;  // reference default
    Math::LinearSolver m_linearSolverType = Math::LINEARSOLVER_LU;               // reference default
    std::shared_ptr<Math::OdeSolver> m_odeSolver;
    bool m_gravityEnabled = true;
    Math::Vector3d m_gravity{{0.0, -9.81, 0.0}};
    double m_rayleighMass = 0.0, m_rayleighStiffness = 0.0;
    bool m_massLumping = false;
    Math::Vector m_externalForce;
    bool m_hasExternalForce = false, m_uploadedExternalForce = false;
    bool m_active = true;
    bool m_syncStates = true;
    double m_totalMass = 0.0;
    double m_lastDt = -1.0;
    Precision m_precision = Precision::FP64;
};

}  // namespace Physics
}  // namespace SurgSimGpu
