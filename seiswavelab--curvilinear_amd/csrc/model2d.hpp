// model2d.hpp — FP64 construction of a 2D curvilinear session (host side).
// Restates Seis2D::Curvilinear::Simulation::initialize (src/two_d/CurvilinearSolver2D.cpp:151-540)
// with the reference's evaluation order, so dt, step counts, index maps and every model array
// agree bit for bit with the reference's CPU path.
#pragma once
#include "swl.h"
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace swl {

struct GslsFit { // SeisWave::Rheology fit restricted to one (Qp, Qs) pair, src/common/Gsls.cpp:11-78
    int mechanisms = 0;
    std::array<double, 3> rates{}, strengthP{}, strengthS{};
    // Re(1/sqrt(shape)) of the P and S strengths at `scaleFref` (Gsls.cpp:45): a property of the fit alone, so
    // per-node materials (relaxed modulus = rho v^2 inv^2) reuse it instead of repeating the complex algebra
    double scaleFref = -1, invP = 0, invS = 0;
};
struct GslsMaterial {
    double pRelaxed = 0, sRelaxed = 0, pUnrelaxed = 0, sUnrelaxed = 0, maxRelativeQError = 0;
    std::array<double, 3> deltaP{}, deltaS{};
};
GslsFit gslsFit(int mechanisms, double fmin, double fmax, double qp, double qs, double fref);
GslsMaterial gslsMaterial(const GslsFit &fit, double fref, double fmin, double fmax, double vp, double vs, double rho,
                          double qp, double qs, bool qError);

// PCHIP terrain with linear continuation (src/two_d/SurfaceProfile.cpp:8-65).
class Pchip {
  public:
    explicit Pchip(const std::vector<std::array<double, 2>> &points);
    std::array<double, 2> sample(double x) const;

  private:
    std::vector<std::array<double, 2>> p_;
    std::vector<double> m_;
};

struct Model2D {
    swl_params2d logical{}, p{};   // p: expanded computational parameters
    int nx = 0, nz = 0, offsetX = 0, widthX = 0, widthZ = 0, fields = 5, mechanisms = 0;
    size_t count = 0;
    int64_t cpmlCount = 0;
    bool explosive = false;
    double h = 0, flatten = 0, maxMetric = 0, minJacobian = 0, maxEdge = 0, amplitude = 0;
    std::vector<double> coordinates, normals, a, b, jac, vp, vs, rho, lambda, mu, source, damping, gradX, gradZ;
    std::array<std::vector<double>, 3> deltaP, deltaS;
    std::vector<double> cpmlX, cpmlZ;
    std::vector<double> surfaceDepth, surfaceSlope, fade, fadeDerivative; // separable terrain profiles
    std::array<double, 3> rates{}, direction{};
    double dt = 0;
    int total = 0;
    double waveDt = 0, dampingDt = 0, sourceDt = 0, cpmlDt = 0, relaxationDt = 0, ppw = 0, maxQError = 0;
    double sourceWeightSum = 0, sourceRmsRadius = 0;
    std::vector<int32_t> receiverX, receiverZ;
    std::vector<double> receiverPosition;
    std::vector<int64_t> probeIndex;
    size_t recordReserve = 0;
    std::vector<double> surfaceSamples; // keeps imported samples alive for `logical`
    std::vector<uint8_t> material;      // layer index per node (attenuating models: selects the node's Qp/Qs)
    double qp(size_t q) const { return logical.layered ? logical.layer_qp[material[q]] : logical.qp; }
    double qs(size_t q) const { return logical.layered ? logical.layer_qs[material[q]] : logical.qs; }
};

// Throws std::runtime_error with the reference's messages for invalid input.
// device >= 0: coordinates, J, a, b and min J / max metric are built on that GPU (geometry2d.hip), bit for bit
// the host's values; device < 0: all on the host.
Model2D buildModel2D(const swl_params2d &requested, int device = -1);
long long buildGeometry2DOnDevice(Model2D &m, int device);
void validateGsls(int mechanisms, double fmin, double fmax, double fref, double qp, double qs);
swl_mesh2d_desc meshDescriptor(const Model2D &m);

// Detail::memoryBudget (src/two_d/MemoryBudget2D.hpp:12-38): the additional host and device bytes of a session,
// for the reference's backends (0 CPU, 1 CUDA).  backend 2 (HIP): host bytes as for CUDA (the FP64 model on the
// host, the RK state on the device) and device bytes bounding what this library's 2D engine allocates.
struct MemoryBudget2D {
    double host = 0, device = 0;
};
MemoryBudget2D memoryBudget2D(const swl_params2d &p, int backend);
// Device bytes of a 2D HIP engine for the session p describes, with `rows` per-step result rows reserved (the
// session steps 512 at a time; fewer when max_steps is smaller).  An upper bound: the border-node list is counted at every computational node.
double hipDeviceBytes2D(const swl_params2d &p, int rows = 512);
// The reference's admission check before a session is built (CurvilinearSolver2D.cpp:247-266): throws unless the
// budget is below 80 % of the free host memory and 85 % of the free memory of `device` (geometry2d.hip).
void preflightMemory2D(const swl_params2d &p, int device);
// setUniformSurfaceReceivers / synchronizeSurfaceReceivers (include/Seis2D/SurfaceSurvey2D.hpp:8-44) for the
// elastic session (surface depth 0); false where the reference's returns false.
bool synchronizeSurfaceReceivers(swl_params2d &p, int limit, int margin);
double lerpExact(double a, double b, double t); // C++20 std::lerp contract (P0811R3)

} // namespace swl
