// model3d.hpp — FP64 construction of a 3D curvilinear session (host side).
// Restates Seis3D::Curvilinear::CpuSimulation3D::initialize (src/three_d/CurvilinearSolver.cpp:258-442)
// with the reference's evaluation order, so coordinates, inverse metrics, Jacobians, normals, the
// CFL geometry, dt, step counts, receiver index maps and every model array agree bit for bit with
// the reference's CPU path.
#pragma once
#include "swl.h"
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace swl {

struct Model3D {
    swl_params3d p{};
    std::vector<double> dem; // owned DEM copy (p.dem_elevations points here)
    int nx = 0, ny = 0, nz = 0; // padded (physical + 2*3)
    size_t count = 0;
    double h = 0, transition = 0;
    bool terrain = false;
    std::vector<double> coordinates, metrics, jacobian, normals; // 3N, 9N, N, 3N
    double lambda = 0, mu = 0;                                    // homogeneous moduli
    std::vector<double> lambdaNode, muNode, densityNode;          // per node (filled also when homogeneous)
    std::vector<double> damping, source;
    std::array<std::vector<double>, 3> gradient; // isotropic moment only
    std::array<double, 3> direction{};
    double amplitude = 0;
    double dt = 0;
    int total = 0;
    double minJacobian = 0, maxMetricNorm = 0, minLength = 0, maxEdge = 0, waveDt = 0, dampingDt = 0, sourceDt = 0;
    double maxDamping = 0, maxVp = 0, minVs = 0, maxVs = 0, maxDensity = 0;
    std::vector<int64_t> receiverIndex;
    std::vector<double> receiverPosition;
    std::array<int, 3> sourceNode{};
    // Separable forms the engine uses instead of streaming full arrays (each is checked against
    // the full array when the engine is built):
    int flatFrom = 0;                                        // first padded k with identity metrics, J = 1 for k >= it
    std::array<std::vector<double>, 3> dampingAxis;          // rate = (X(i) + Y(j)) + Z(k), per padded index
    bool materialByDepth = false;                            // lambda, mu, rho depend on k only
    size_t index(int i, int j, int k) const { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k); }
};

// Throws std::runtime_error with the reference's messages for invalid input.  device >= 0: the volume
// geometry and the CFL reductions are computed on that GPU in FP64 (bit-identical to the host loops).
Model3D buildModel3D(const swl_params3d &requested, int device = -1, int precision = 0);
// geometry3d.hip: coordinates, metrics, Jacobian, normals and the CFL figures from the per-column surface
// displacement and the per-plane fade.
void buildGeometry3DOnDevice(Model3D &m, const std::vector<double> &surface, const std::vector<double> &fade, int device);
swl_mesh3d_desc meshDescriptor3D(const Model3D &m, int precision);
int layerOfNode3D(const Model3D &m, size_t q); // layered models: the node's layer index

// estimateModelMemory3D (include/Seis3D/ModelPresets3D.hpp:38-49): padded points, CPU MiB, CUDA device MiB,
// CUDA host MiB, and whether the grid is within the padded-node budget.
struct ModelMemory3D {
    size_t points = 0;
    double cpuMiB = 0, deviceMiB = 0, cudaHostMiB = 0;
    bool withinBudget = false;
};
ModelMemory3D estimateModelMemory3D(const swl_params3d &p);
// Device bytes of a 3D HIP engine for the session p describes, with `rows` per-step result rows reserved (the
// session steps 256 at a time; fewer when max_steps is smaller).  An upper bound: metrics, materials and damping are counted on every node.
double hipDeviceBytes3D(const swl_params3d &p, int precision, int rows = 256);
// fitsMemory3D (ModelPresets3D.hpp:50-56) for a HIP session on `device`, with the reference's message
// (CurvilinearSolver.cpp:309); implemented in geometry3d.hip.
void preflightMemory3D(const swl_params3d &p, int precision, int device);

} // namespace swl
