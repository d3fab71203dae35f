// engine2d.hip — gfx950 HIP engine for the 2D P-SV elastic / Qp-Qs GSLS curvilinear solver.
//
// Replaces the reference's CudaEngine (src/two_d/CudaEngine2D.cu, .hpp:5-24).  The per-node
// numerics follow src/two_d/PointKernels2D.hpp (rhs :121-172, project :70-78, closure :79-89,
// traction-image momentum :94-120, CFS-PML ADE :50-64) in the same floating-point order, so FP64
// runs reproduce the reference's CPU path to rounding.  The execution model is the device's own:
//   * one fused kernel pass per RK stage: RHS + stage combine + surface projection + PML / GSLS
//     memory update; the stage-4 pass also produces the energy / finiteness inspection;
//   * four rotating state buffers plus one accumulator (acc = k1 + 2k2 + 2k3, rounded exactly as the
//     reference's a+2b+2c+d) instead of seven RK field sets; rollback is a buffer-index change;
//   * an aligned padded layout (node (i,k) at off + i + ld*k, column 3 on a 256-byte boundary, ld a
//     multiple of 64) so the bulk of the grid streams with 16-byte lane accesses;
//   * an interior kernel for the rectangle below the surface band and inside the absorber strips:
//     each lane owns 4 (FP32) or 2 (FP64) x-adjacent nodes, the direction schedule is a template
//     parameter, the separable curvilinear metrics a, b, J are recomputed in FP64 from the 1D terrain
//     and fade profiles (bit-identical to the stored arrays, checked when the engine is built), and
//     results leave with nontemporal 16-byte stores;
//   * a general scalar kernel for every node outside that rectangle (surface closure in FP64,
//     reduced stencils, traction image, CFS-PML ADE memories, sponge, source support);
//   * tiles mapped XCD-contiguously (each of the 8 XCDs sweeps its own band of rows so the stencil's
//     z-neighbours are reused in that XCD's L2);
//   * device-side fault rule, peak tracking and receiver gather (swl_engine2d_run): no per-step
//     host synchronisation; each step is ended by one block of the next step's first stage launch, so a
//     step is four launches.
#include "common.hpp"
#include "swl.h"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace swl {
namespace {

constexpr int TX = 64, TZ = 4, THREADS = TX * TZ; // general tiles: 64 x 4 nodes
constexpr int ORIGIN = 61;                        // off: column 3 lands on a 64-element boundary

typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
template <class R> struct Lane;
template <> struct Lane<float> {
    using V = vf4;
    static constexpr int N = 4;
};
template <> struct Lane<double> {
    using V = vd2;
    static constexpr int N = 2;
};

template <class R> struct DMesh {
    int nx, nz, ld, off, wx, wz;
    long long cpml;
    int fields, mech, has_damp, has_grad, metric;
    R h, hinv, rates[3], dir0, dir2; // hinv = 1/h (FP64: only when h is a power of two, so exact), else 0
    const R *a, *b, *jac, *lam, *mu, *rho, *src, *damp, *grad0, *grad1, *nrm, *cx, *cz;
    const R *cxf, *czf; // FP32: per PML position 1/kappa, d/kappa + alpha, d/kappa^2 (no divisions per node)
    const R *dP[3], *dS[3];
    const double *sa, *sb, *slam, *smu, *snrm;   // FP64 surface row k = 3, indexed by i
    const double *tz_, *tslope, *fade, *fadeder; // separable metric profiles (by i / by k)
    int ri0, ri1, rk0, rk1;                      // interior rectangle [ri0,ri1) x [rk0,rk1)
    int si0, si1, sk0, sk1;                      // source support box (inclusive), si1 < 0 if none
    int by_row;                                  // lambda, mu, rho, dP, dS depend on k only (per-row tables)
    const R *lam_k, *mu_k, *rho_k, *dP_k[3], *dS_k[3];
};
template <class R> __device__ __host__ inline ptrdiff_t at(const DMesh<R> &m, int i, int k) {
    return (ptrdiff_t)m.off + i + (ptrdiff_t)m.ld * k;
}
template <class R> __device__ inline bool in_rect(const DMesh<R> &m, int i, int k) {
    return i >= m.ri0 && i < m.ri1 && k >= m.rk0 && k < m.rk1;
}

struct Partial {
    double energy, velocity, traction;
    int bad, pad;
};

// End of the previous step in a device-resident run, done by block 0 of the next step's first stage launch
// while the other blocks work: the reduction of the previous step's inspection partials (written by the
// previous launch), the fault rule and, if the step is accepted, its receiver row read from the stage input
// (which is that step's state).  The engine keeps four rotating state buffers, so the first stage writes
// neither the accepted state nor the one before it: a step rejected here still leaves the state it started
// from intact, and every later launch sees the fault and does nothing.
struct StepEnd {
    const Partial *partials; // the previous step's partials
    int n;
    int *status, *accepted;
    double *peak, *diag, *samples;
    const long long *probes;
    int nrec;
    int row, step; // row and step number of the previous step
    double dt, frequency;
};

template <class R> struct StageArgs {
    DMesh<R> m;
    const R *in, *base;
    R *out, *acc;
    const R *pin, *pbase;
    R *pout, *pacc;
    size_t count;
    R forcing, w, wy; // wy: weight of this stage's k in Y = U^n + dt(k1 + 2k2 + 2k3 + k4)/6 (FP32 interior)
    R hdt, dt, dt6;   // dt / 2, dt, dt / 6 (memory variables rebuilt from U^n, FP32 interior)
    int stage;        // 0..3
    int kind;         // 0 first stage, 1 middle stages, 2 final stage (+ inspection)
    int ystage;       // FP32 interior, stage 2: Y = U^n + dt/6 k1 is rebuilt from the input and the base
    const int *status;
    Partial *partials;       // border blocks first, then interior tiles
    const int2 *border;      // (i, k) of every node outside the interior rectangle, row by row, 256 per block
    int nborder;             // border nodes
    int ngeneral, gblocks, ninterior, rtx;
    const int2 *pml;  // FP32 elastic: (i0, k) of every CFS-PML vector lane below the surface band, 256 per block
    int npml, pml0, npmlb; // lanes, first border-class block index and number of their blocks
    int ipart0;       // partial slot of interior tile 0 (after the general and PML blocks' slots)
    int nt;          // interior: nontemporal stores
    int spread;      // border blocks spread evenly over the launch (else all before the interior tiles)
    StepEnd end;     // first stage of a device-resident run: end.partials non-null ends the previous step
};

// ---- tile mapping: contiguous band of tiles per XCD (blocks are dispatched round-robin over 8 XCDs)
__device__ inline int tile_of_block(int b, int n) {
    const int x = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + j;
}

// FP64 keeps the reference's operation order exactly (separate multiply and add, true divisions), so
// FP64 runs reproduce the reference to the bit.  FP32 is held to a tolerance, not to the bit, so it fuses
// multiply-adds and replaces repeated divisions by reciprocals.
template <class R> constexpr bool kFast = sizeof(R) == 4;
template <class R> __device__ inline R mad(R a, R b, R c) {
    if constexpr (kFast<R>) return __builtin_fmaf(a, b, c);
    else return a * b + c;
}

// ---- MacCormack-DRP operators (FiniteDifference.hpp:14-25), direction as a template parameter
template <class R, int DIR> __device__ inline R coef(int l) {
    const R c = l == 0 ? (R)-.30874 : l == 1 ? (R)-.6326 : l == 2 ? (R)1.233 : l == 3 ? (R)-.3334 : (R).04168;
    return DIR ? -c : c;
}
template <int DIR> __device__ constexpr int offs(int l) { return DIR ? -(l - 1) : (l - 1); }
// r / h, computed as r * (1/h) when h is a power of two (the two are then the same IEEE result)
// FP64 multiplies by 1/h only where that is exact (h a power of two) and divides otherwise; FP32 always has hinv.
// FLAT (FP32 bodies with two or three GSLS mechanisms, and the small-grid elastic kernel): multiply
// unconditionally, which keeps the stencil code one basic block, so that the compiler overlaps several fields' loads
// (a small grid's launch is bounded by that latency chain); they stay at 3 waves per SIMD.  The elastic bodies of
// the large-grid kernel keep the branch (at their 5-wave register budget the overlapped loads spill), and so does
// GSLS-1 (it would drop from 3 to 2 waves).
template <int MECH, bool LAT = false> constexpr bool kFlatLoads = LAT || MECH >= 2;
template <class R, bool FLAT = false> __device__ inline R over_h(R r, R h, R hinv) {
    if constexpr (kFast<R> && FLAT) return r * hinv;
    else return hinv != (R)0 ? r * hinv : r / h;
}

template <class R, int DIR, bool FLAT = false>
__device__ inline R d5(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h, R hinv) {
    R r = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) r += coef<R, DIR>(l) * v[q + offs<DIR>(l) * s];
    return over_h<R, FLAT>(r, h, hinv);
}
template <class R, int DIR, bool FLAT = false>
__device__ inline R d3(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h, R hinv) {
    const R l0 = ((R)-7) / ((R)6), l1 = ((R)8) / ((R)6), l2 = ((R)-1) / ((R)6);
    const R c0 = DIR ? -l0 : l0, c1 = DIR ? -l1 : l1, c2 = DIR ? -l2 : l2;
    constexpr int o1 = DIR ? -1 : 1, o2 = DIR ? -2 : 2;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    r += c2 * v[q + o2 * s];
    return over_h<R, FLAT>(r, h, hinv);
}
template <class R, int DIR, bool FLAT = false>
__device__ inline R d2(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h, R hinv) {
    const R c0 = DIR ? (R)1 : (R)-1, c1 = DIR ? (R)-1 : (R)1;
    constexpr int o1 = DIR ? -1 : 1;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    return over_h<R, FLAT>(r, h, hinv);
}

// ---- compact CFS-PML slots (PointKernels2D.hpp:50-56)
__device__ inline long long pml_slot(int nx, int nz, int wx, int wz, long long cpml, int i, int k, int axis) {
    if (!cpml) return -1;
    const int ii = i - 3, kk = k - 3, nxc = nx - 6, nzc = nz - 6;
    if (axis == 0) {
        const int j = ii < wx ? ii : ii >= nxc - wx ? wx + ii - (nxc - wx) : -1;
        return j < 0 ? -1 : 4 * ((long long)kk * 2 * wx + j);
    }
    return kk < nzc - wz ? -1 : 4 * ((long long)(2 * wx) * nzc + (long long)(kk - (nzc - wz)) * nxc + ii);
}

// traction flux J*(grad a)*sigma (PointKernels2D.hpp:90-93)
template <class R>
__device__ inline R flux(const DMesh<R> &m, const R *__restrict__ w, size_t count, ptrdiff_t p, int axis, int comp) {
    const R x = axis ? m.a[p] : (R)1, z = axis ? m.b[p] : (R)0;
    return m.jac[p] * (x * w[(size_t)(comp ? 4 : 2) * count + p] + z * w[(size_t)(comp ? 3 : 4) * count + p]);
}

// traction-image momentum (PointKernels2D.hpp:94-120); only = -1 both axes, 0 xi only, 1 eta only
template <class R, int DX, int DZ>
__device__ R momentum(const DMesh<R> &m, const R *__restrict__ w, size_t count, ptrdiff_t q, int comp, int depth,
                      int only) {
    R result = 0;
    if (only != 1) {
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            R sign = 1;
            result += sign * coef<R, DX>(l) * flux(m, w, count, q + offs<DX>(l), 0, comp);
        }
    }
    if (only != 0) {
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            int o = offs<DZ>(l);
            R sign = 1;
            const int kk = depth + o;
            if (kk == 0) continue;
            if (kk < 0) {
                o = -kk - depth;
                sign = -1;
            }
            result += sign * coef<R, DZ>(l) * flux(m, w, count, q + (ptrdiff_t)o * m.ld, 1, comp);
        }
    }
    return result / (m.h * m.jac[q] * m.rho[q]);
}

// ADE memory rate and stretched derivative (PointKernels2D.hpp:57-64)
template <class R>
__device__ inline R stretched(const DMesh<R> &m, const R *__restrict__ pin, long long slot, int at_, int axis, int field,
                              R d, R *rate) {
    if (slot < 0) return d;
    if constexpr (kFast<R>) { // the same ADE with the profile's quotients precomputed per position
        const R *pf = (axis ? m.czf : m.cxf) + 3 * at_;
        const R psi = pin[slot + field];
        rate[4 * axis + field] = -mad(pf[1], psi, pf[2] * d);
        return mad(d, pf[0], psi);
    }
    const R *pr = (axis ? m.cz : m.cx) + 3 * at_;
    const R dd = pr[0], kappa = pr[1], alpha = pr[2], psi = pin[slot + field];
    rate[4 * axis + field] = -(dd / kappa + alpha) * psi - dd / (kappa * kappa) * d;
    return d / kappa + psi;
}

// FP32 with GSLS mechanisms: a memory variable obeys the pointwise ODE dm/dt = r (d - m)
// (- damp m in a sponge), whose drive d is a combination of three strains (exx, ezz, exz) with the
// mechanism's dP, dS.  Its stage inputs are therefore rebuilt from U^n and the strains of the step's earlier
// stages instead of being streamed: each of stages 1-3 stores the node's three strains in otherwise unused
// memory slots (strain_dst / strain_src below), and stage 4 forms
// k1..k4 and writes U^{n+1} = U^n + dt (k1 + 2 k2 + 2 k3 + k4) / 6.  Syy is pointwise too and no other
// rate reads it (dSyy/dt = lambda div - sum of the yy memories - damp Syy), so it and the yy memories are
// formed at stage 4 only.  The same RK4, rounded differently; per interior node and step the 13 x 14
// accesses of the memories and Syy become 80 (three mechanisms).  Every node of the grid (interior and
// border) follows this scheme, so the scratch slots are per node.
template <class R, int MECH> constexpr bool kRebuildMemory = kFast<R> && MECH >= 1;
// drive of memory component f (xx, zz, xz, yy) for a mechanism with moduli dP, dS
template <class R> __device__ __forceinline__ R memory_drive(int f, R dp, R ds, R exx, R ezz, R exz) {
    const R dl = dp - 2 * ds, div = exx + ezz;
    return f == 0 ? mad(dl, div, 2 * ds * exx) : f == 1 ? mad(dl, div, 2 * ds * ezz) : f == 2 ? ds * exz : dl * div;
}
// rate of a memory variable (DAMP: with the sponge term)
template <class R, bool DAMP> __device__ __forceinline__ R memory_rate(R r, R dmp, R d, R m) {
    if constexpr (DAMP) return mad(-dmp, m, r * (d - m));
    else return r * (d - m);
}
// the stage input m_s of a memory variable after s earlier stages with drives d[0..s-1]; with s = 3 also
// the weighted sums k1 + 2 k2 + 2 k3 of its rates and m_0 + 2 m_1 + 2 m_2 of its stage inputs
template <class R, bool DAMP = false>
__device__ __forceinline__ R memory_stage(int s, R r, R dmp, R m0, const R (&d)[3], R hdt, R dt, R &ksum, R &msum) {
    R m = m0;
    ksum = 0, msum = m0;
    if (s >= 1) {
        const R k = memory_rate<R, DAMP>(r, dmp, d[0], m);
        ksum = k, m = mad(hdt, k, m0), msum = mad((R)2, m, msum);
    }
    if (s >= 2) {
        const R k = memory_rate<R, DAMP>(r, dmp, d[1], m);
        ksum = mad((R)2, k, ksum), m = mad(hdt, k, m0), msum = mad((R)2, m, msum);
    }
    if (s >= 3) {
        const R k = memory_rate<R, DAMP>(r, dmp, d[2], m);
        ksum = mad((R)2, k, ksum), m = mad(dt, k, m0);
    }
    return m;
}
// Memory slots that hold the strains of stage s (0..2) of the step.  With two or more mechanisms: stages 1 and 2
// in the accumulator's slots 6 + 3 s, stage 3 in its output's slots 6..8 (stage 4's input).  With one mechanism the
// accumulator has four memory slots: stage 1 there, stage 2 in its output's (stage 3's input, stage 4's output,
// read there before stage 4 writes the memories), stage 3 as above.  Stage buffers: in = U^n, b1, b2, b1 and
// out = b1, b2, b1, b2 for stages 1..4.
template <class R, int MECH> __device__ __forceinline__ R *strain_dst(const StageArgs<R> &A) {
    const int s = A.stage;
    if (s == 0 || (s == 1 && MECH >= 2)) return A.acc + (size_t)(6 + 3 * s) * A.count;
    return A.out + (size_t)6 * A.count;
}
template <class R, int MECH> __device__ __forceinline__ const R *strain_src(const StageArgs<R> &A, int s) {
    if (s == 0 || (s == 1 && MECH >= 2)) return A.acc + (size_t)(6 + 3 * s) * A.count;
    if (s == 1) return (A.stage == 2 ? A.in : A.out) + (size_t)6 * A.count;
    return A.in + (size_t)6 * A.count;
}

// traction-free closure for the surface z-derivatives, evaluated in FP64 (PointKernels2D.hpp:79-89)
// memIn: the memory variables' stage inputs (xx, zz, xz, yy per mechanism) where they are rebuilt, else null
template <class R>
__device__ inline void closure(const DMesh<R> &m, const R *__restrict__ w, size_t count, ptrdiff_t q, int i, R xi0, R xi1,
                               double *result, const R (*memIn)[4] = nullptr) {
    const double c0 = m.sa[i], c1 = m.sb[i], c2 = c0 * c0 + c1 * c1, length = sqrt(c2);
    const double n0 = c0 / length, n1 = c1 / length;
    const double dx = xi0, dz = xi1, lambda = m.slam[i], mu = m.smu[i];
    double t0 = (lambda + 2 * mu) * dx * c0 + mu * dz * c1, t1 = mu * dz * c0 + lambda * dx * c1;
    for (int l = 0; l < m.mech; ++l) {
        const int k = 6 + 4 * l;
        const double sxx = memIn ? memIn[l][0] : w[(size_t)k * count + q], szz = memIn ? memIn[l][1] : w[(size_t)(k + 1) * count + q],
                     sxz = memIn ? memIn[l][2] : w[(size_t)(k + 2) * count + q];
        t0 -= sxx * c0 + sxz * c1;
        t1 -= sxz * c0 + szz * c1;
    }
    const double normal = t0 * n0 + t1 * n1;
    result[0] = -(t0 - n0 * normal) / (mu * c2) - n0 * normal / ((lambda + 2 * mu) * c2);
    result[1] = -(t1 - n1 * normal) / (mu * c2) - n1 * normal / ((lambda + 2 * mu) * c2);
}

// Full RHS of one node outside the interior rectangle (PointKernels2D.hpp:121-172).  REBUILD: the memory
// variables' stage inputs come from memIn (the rates of the memories, Syy and their sponge terms are then
// left to the caller) and the node's strains exx, ezz, exz go to strain.
template <class R, int DX, int DZ, int MECH, bool REBUILD = false, bool LAT = false>
__device__ inline void rhs_general(const DMesh<R> &m, const R *__restrict__ w, const R *__restrict__ pin, size_t count,
                                   int i, int k, R forcing, R *kout, R *prate, long long &sx, long long &sz,
                                   const R (*memIn)[4] = nullptr, R *strain = nullptr) {
    constexpr int F = MECH ? 6 + 4 * MECH : 5;
    constexpr bool FLAT = kFlatLoads<MECH, LAT>;
    const int depth = k - 3;
    const ptrdiff_t q = at(m, i, k), s = m.ld;
    const R h = m.h;
    sx = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 0);
    sz = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 1);
    R xi[5], eta[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) xi[f] = d5<R, DX, FLAT>(w + (size_t)f * count, q, 1, h, m.hinv);
#pragma unroll
    for (int f = 2; f < 5; ++f) eta[f] = d5<R, DZ, FLAT>(w + (size_t)f * count, q, s, h, m.hinv);
    R vx[2], vz[2];
    if (depth == 0) {
        R xs[2];
        for (int v = 0; v < 2; ++v) xs[v] = stretched(m, pin, sx, i - 3, 0, v, xi[v], prate);
        double sur[2];
        closure(m, w, count, q, i, xs[0], xs[1], sur, REBUILD ? memIn : nullptr);
        eta[0] = (R)sur[0];
        eta[1] = (R)sur[1];
    } else {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const R *fv = w + (size_t)v * count;
            eta[v] = depth == 1   ? d2<R, DZ, FLAT>(fv, q, s, h, m.hinv)
                     : depth == 2 ? d3<R, DZ, FLAT>(fv, q, s, h, m.hinv)
                                  : d5<R, DZ, FLAT>(fv, q, s, h, m.hinv);
        }
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        vx[v] = stretched(m, pin, sx, i - 3, 0, v, xi[v], prate);
        vz[v] = stretched(m, pin, sz, k - 3, 1, v, eta[v], prate);
    }
    const R a = m.a[q], b = m.b[q], rho = m.rho[q];
    R dx[5], dz[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const R ux = f < 2 ? vx[f] : xi[f], uz = f < 2 ? vz[f] : eta[f];
        dx[f] = ux + a * uz;
        dz[f] = b * uz;
    }
    const bool image = depth <= (DZ ? 3 : 1);
    const R rinv = kFast<R> ? (R)1 / rho : (R)0; // FP32: one reciprocal instead of a division per term
    auto per_rho = [&](R x) -> R {
        if constexpr (kFast<R>) return x * rinv;
        else return x / rho;
    };
    R o0 = per_rho(dx[2] + dz[4]), o1 = per_rho(dx[4] + dz[3]);
    if (image && !m.cpml) {
        o0 = momentum<R, DX, DZ>(m, w, count, q, 0, depth, -1);
        o1 = momentum<R, DX, DZ>(m, w, count, q, 1, depth, -1);
    }
    if (m.cpml) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int f = v ? 4 : 2, other = v ? 3 : 4;
            R x = per_rho(xi[f]), z = per_rho(a * eta[f] + b * eta[other]);
            if (image) {
                x = momentum<R, DX, DZ>(m, w, count, q, v, depth, 0);
                z = momentum<R, DX, DZ>(m, w, count, q, v, depth, 1);
            }
            const R r = stretched(m, pin, sx, i - 3, 0, 2 + v, x, prate) + stretched(m, pin, sz, k - 3, 1, 2 + v, z, prate);
            if (v) o1 = r;
            else o0 = r;
        }
    }
    if (m.has_grad) {
        o0 += per_rho(forcing * m.grad0[q]);
        o1 += per_rho(forcing * m.grad1[q]);
    } else {
        const R sw = m.src[q];
        o0 += per_rho(forcing * (sw * m.dir0));
        o1 += per_rho(forcing * (sw * m.dir2));
    }
    const R lambda = m.lam[q], mu = m.mu[q], div = dx[0] + dz[1];
    kout[0] = o0;
    kout[1] = o1;
    kout[2] = lambda * div + 2 * mu * dx[0];
    kout[3] = lambda * div + 2 * mu * dz[1];
    kout[4] = mu * (dz[0] + dx[1]);
    if constexpr (MECH > 0) {
        kout[5] = lambda * div;
#pragma unroll
        for (int l = 0; l < MECH; ++l) {
            const int kk = 6 + 4 * l;
            const R dp = m.dP[l][q], ds = m.dS[l][q], dl = dp - 2 * ds;
            const R drive[4] = {dl * div + 2 * ds * dx[0], dl * div + 2 * ds * dz[1], ds * (dz[0] + dx[1]), dl * div};
            R mem[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) mem[f] = REBUILD ? memIn[l][f] : w[(size_t)(kk + f) * count + q];
            kout[2] -= mem[0];
            kout[3] -= mem[1];
            kout[4] -= mem[2];
            kout[5] -= mem[3];
#pragma unroll
            for (int f = 0; f < 4; ++f) kout[kk + f] = m.rates[l] * (drive[f] - mem[f]);
        }
    }
    if (m.has_damp) {
        const R d = m.damp[q];
#pragma unroll
        for (int f = 0; f < (REBUILD ? 5 : F); ++f) kout[f] -= d * w[(size_t)f * count + q];
    }
    if constexpr (REBUILD) strain[0] = dx[0], strain[1] = dz[1], strain[2] = dz[0] + dx[1];
}

// surface stress projection sigma <- P sigma P on row k = 3 (PointKernels2D.hpp:70-78), FP64
__device__ inline void project(const double *nrm, double &sxx, double &szz, double &sxz) {
    const double nx = nrm[0], nz = nrm[1];
    const double tx = sxx * nx + sxz * nz, tz = sxz * nx + szz * nz, nn = nx * tx + nz * tz;
    sxx += -2 * nx * tx + nn * nx * nx;
    szz += -2 * nz * tz + nn * nz * nz;
    sxz += -nx * tz - tx * nz + nn * nx * nz;
}

// RK stage combine (CurvilinearSolver2D.cpp:80-97) with the accumulator form of a+2b+2c+d.
template <class R> __device__ inline R combine(int kind, R base, R k, R w, R &acc) {
    if (kind == 0) {
        acc = k;
        return base + w * k;
    }
    if (kind == 1) {
        acc = acc + 2 * k;
        return base + w * k;
    }
    if constexpr (kFast<R>) return mad(w * (acc + k), (R)(1.0 / 6.0), base);
    else return base + w * (acc + k) / (R)6;
}

__device__ inline void merge(Partial &a, const Partial &b) {
    a.energy += b.energy;
    a.velocity = fmax(a.velocity, b.velocity);
    a.traction = fmax(a.traction, b.traction);
    a.bad |= b.bad;
}
// Block reduction of the inspection partials: a butterfly over each wave's 64 lanes, then thread 0 merges the
// waves' results (fixed order, so the same inputs give the same sum); sh[0] holds the block's partial afterwards.
constexpr int kShSlots = THREADS / 64;
__device__ void block_reduce(Partial &p, Partial *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Partial q;
        q.energy = __shfl_xor(p.energy, o);
        q.velocity = __shfl_xor(p.velocity, o);
        q.traction = __shfl_xor(p.traction, o);
        q.bad = __shfl_xor(p.bad, o);
        merge(p, q);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        Partial r = sh[0];
        for (int w = 1; w < kShSlots; ++w) merge(r, sh[w]);
        sh[0] = r;
    }
    __syncthreads();
}
// energy (PointKernels2D.hpp:181-184) and inspection terms (CurvilinearSolver2D.cpp:102-111) of one node
// FAST: one FP64 division (xz^2/(2mu) = xz^2 * 4(lambda+mu) / (8 mu (lambda+mu))), same value to rounding
template <class R, bool FAST = false>
__device__ inline void node_stats(R rho, R lam, R mu, R jac, R h, const R *u, bool surface, const double *nrm, Partial &p) {
    const double l = lam, mm = mu, xx = u[2], zz = u[3], xz = u[4], vx = u[0], vz = u[1];
    double e;
    if constexpr (FAST)
        e = (.5 * (double)rho * (vx * vx + vz * vz) +
             ((l + 2 * mm) * (xx * xx + zz * zz) - 2 * l * xx * zz + xz * xz * (4 * (l + mm))) / (8 * mm * (l + mm))) *
            (double)jac * (double)h * (double)h;
    else
        e = (.5 * (double)rho * (vx * vx + vz * vz) +
             ((l + 2 * mm) * (xx * xx + zz * zz) - 2 * l * xx * zz) / (8 * mm * (l + mm)) + xz * xz / (2 * mm)) *
            (double)jac * (double)h * (double)h;
    const R v = sqrt(u[0] * u[0] + u[1] * u[1]);
    p.energy += e;
    p.velocity = fmax(p.velocity, (double)v);
    if (!isfinite(e) || !isfinite(v)) p.bad = 1;
    if (surface) {
        const double n0 = nrm[0], n1 = nrm[1];
        const double tx = xx * n0 + xz * n1, tz = xz * n0 + zz * n1, norm = sqrt(xx * xx + zz * zz + 2 * xz * xz);
        p.traction = fmax(p.traction, sqrt(tx * tx + tz * tz) / fmax(norm, 1e-25));
    }
}

// ------------------------------------------------------------------------------------------------
// General kernel: block li takes border nodes [256 li, 256 li + 256) of the compacted list (every lane
// holds a node; the lists follow the rows of 64x4 tiles, so a wave still reads row segments).
template <class R, int DX, int DZ, int MECH, bool LAT = false>
__device__ __forceinline__ void general_body(const StageArgs<R> &A, int li, Partial *sh) {
    constexpr int F = MECH ? 6 + 4 * MECH : 5;
    const int n = li * THREADS + (int)threadIdx.x;
    const bool live = n < A.nborder;
    const int2 ik = live ? A.border[n] : int2{3, 3};
    const int i = ik.x, k = ik.y;
    const DMesh<R> &m = A.m;
    Partial p{0, 0, 0, 0, 0};
    if (live) {
        R kv[F], pr[8];
        long long sx, sz;
        const ptrdiff_t q = at(m, i, k);
        R nxt[F];
        constexpr bool REB = kRebuildMemory<R, MECH>;
        constexpr int FW = REB ? 5 : F; // fields combined here (the wavefields, where the rest is rebuilt)
        R ep[3][3], strain[3];          // strains of the step's earlier stages and of this one (rebuild)
        const size_t cnt = A.count;
        const R dmp = REB && m.has_damp ? m.damp[q] : (R)0;
        if constexpr (REB) {
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    ep[s][c] = s < A.stage ? strain_src<R, MECH>(A, s)[(size_t)c * cnt + q] : (R)0;
            R memIn[MECH][4];
#pragma unroll
            for (int l = 0; l < MECH; ++l) {
                const R dp = m.dP[l][q], ds = m.dS[l][q];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    R d[3], ks, ms;
#pragma unroll
                    for (int s = 0; s < 3; ++s) d[s] = memory_drive(f, dp, ds, ep[s][0], ep[s][1], ep[s][2]);
                    memIn[l][f] = memory_stage<R, true>(A.stage, m.rates[l], dmp, A.base[(size_t)(6 + 4 * l + f) * cnt + q], d,
                                                        A.hdt, A.dt, ks, ms);
                }
                memIn[l][3] = 0; // the yy memories enter Syy only, which is formed at the last stage
            }
            rhs_general<R, DX, DZ, MECH, true, LAT>(m, A.in, A.pin, cnt, i, k, A.forcing, kv, pr, sx, sz, memIn, strain);
        } else {
            rhs_general<R, DX, DZ, MECH, false, LAT>(m, A.in, A.pin, cnt, i, k, A.forcing, kv, pr, sx, sz);
        }
#pragma unroll
        for (int f = 0; f < FW; ++f) {
            const size_t o = (size_t)f * cnt + q;
            R acc = A.kind == 0 ? (R)0 : A.acc[o];
            nxt[f] = combine(A.kind, A.base[o], kv[f], A.w, acc);
            if (A.kind != 2) A.acc[o] = acc;
        }
        if (k == 3) {
            double sxx = nxt[2], szz = nxt[3], sxz = nxt[4];
            project(m.snrm + 2 * i, sxx, szz, sxz);
            nxt[2] = (R)sxx, nxt[3] = (R)szz, nxt[4] = (R)sxz;
        }
#pragma unroll
        for (int f = 0; f < FW; ++f) A.out[(size_t)f * cnt + q] = nxt[f];
        if constexpr (REB) {
            if (A.stage < 3) { // this stage's strains, for the later stages of the step
                R *dst = strain_dst<R, MECH>(A);
#pragma unroll
                for (int c = 0; c < 3; ++c) dst[(size_t)c * cnt + q] = strain[c];
            } else { // U^{n+1} of the memories and of Syy, with the sponge terms
                R yy[4] = {0, 0, 0, 0}; // sum over mechanisms of the yy memories' stage inputs
#pragma unroll
                for (int l = 0; l < MECH; ++l) {
                    const R dp = m.dP[l][q], ds = m.dS[l][q], r = m.rates[l];
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        const size_t o = (size_t)(6 + 4 * l + f) * cnt + q;
                        R d[4];
#pragma unroll
                        for (int s = 0; s < 3; ++s) d[s] = memory_drive(f, dp, ds, ep[s][0], ep[s][1], ep[s][2]);
                        d[3] = memory_drive(f, dp, ds, strain[0], strain[1], strain[2]);
                        const R m0 = A.base[o];
                        const R k1 = memory_rate<R, true>(r, dmp, d[0], m0), m1 = mad(A.hdt, k1, m0);
                        const R k2 = memory_rate<R, true>(r, dmp, d[1], m1), m2 = mad(A.hdt, k2, m0);
                        const R k3 = memory_rate<R, true>(r, dmp, d[2], m2), m3 = mad(A.dt, k3, m0);
                        const R k4 = memory_rate<R, true>(r, dmp, d[3], m3);
                        nxt[6 + 4 * l + f] = mad(A.dt6, k1 + 2 * (k2 + k3) + k4, m0);
                        A.out[o] = nxt[6 + 4 * l + f];
                        if (f == 3) yy[0] += m0, yy[1] += m1, yy[2] += m2, yy[3] += m3;
                    }
                }
                const size_t o = (size_t)5 * cnt + q;
                const R lambda = m.lam[q], s0 = A.base[o];
                R div[4];
#pragma unroll
                for (int s = 0; s < 3; ++s) div[s] = ep[s][0] + ep[s][1];
                div[3] = strain[0] + strain[1];
                const R k1 = mad(lambda, div[0], -yy[0]) - dmp * s0, s1 = mad(A.hdt, k1, s0);
                const R k2 = mad(lambda, div[1], -yy[1]) - dmp * s1, s2 = mad(A.hdt, k2, s0);
                const R k3 = mad(lambda, div[2], -yy[2]) - dmp * s2, s3 = mad(A.dt, k3, s0);
                const R k4 = mad(lambda, div[3], -yy[3]) - dmp * s3;
                nxt[5] = mad(A.dt6, k1 + 2 * (k2 + k3) + k4, s0);
                A.out[o] = nxt[5];
            }
        }
        R pn[8];
        const long long slots[2] = {sx, sz};
#pragma unroll
        for (int axis = 0; axis < 2; ++axis) {
            const long long sl = slots[axis];
            if (sl < 0) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                R acc = A.kind == 0 ? (R)0 : A.pacc[sl + c];
                pn[4 * axis + c] = combine(A.kind, A.pbase[sl + c], pr[4 * axis + c], A.w, acc);
                if (A.kind != 2) A.pacc[sl + c] = acc;
                A.pout[sl + c] = pn[4 * axis + c];
            }
        }
        if (A.kind == 2) {
#pragma unroll
            for (int f = 0; f < F; ++f)
                if (!isfinite(nxt[f])) p.bad = 1;
            const R big = sizeof(R) == 4 ? (R)3.4028234663852886e38 : (R)1.7976931348623157e308;
#pragma unroll
            for (int axis = 0; axis < 2; ++axis)
                if (slots[axis] >= 0)
                    for (int c = 0; c < 4; ++c)
                        if (!(pn[4 * axis + c] <= big && pn[4 * axis + c] >= -big)) p.bad = 1;
            const double nrm2[2] = {(double)m.nrm[2 * q], (double)m.nrm[2 * q + 1]};
            node_stats(m.rho[q], m.lam[q], m.mu[q], m.jac[q], m.h, nxt, k == 3, nrm2, p);
        }
    }
    if (A.kind == 2) {
        block_reduce(p, sh);
        if (threadIdx.x == 0) A.partials[li] = sh[0];
    }
}

// ------------------------------------------------------------------------------------------------
// Interior kernel: each lane owns NV x-adjacent nodes of one row; a wave covers 64*NV nodes of a row
// and a block 4 rows.  Depth >= 4, no absorber, so the z-derivatives of every field use the 5-point
// DRP and the RHS has none of the boundary terms (PointKernels2D.hpp:133-171 restricted to them).
template <class R> __device__ inline typename Lane<R>::V vload(const R *p) {
    return *reinterpret_cast<const typename Lane<R>::V *>(p);
}
template <class R> __device__ inline void vstore(R *p, typename Lane<R>::V v, int nt) {
    auto *d = reinterpret_cast<typename Lane<R>::V *>(p);
    if (nt) __builtin_nontemporal_store(v, d);
    else *d = v;
}

// Interior arithmetic.  FP64 keeps the reference's operation order exactly (separate multiply and add,
// true divisions), so FP64 runs reproduce the reference to the bit.  FP32 is held to a tolerance, not
// to the bit, so its interior fuses multiply-adds and replaces repeated divisions by one reciprocal.

// Stage combine of one interior field at offset o.  FP32: the accumulator holds
// Y = U^n + dt/6 k1 + dt/3 k2 + dt/3 k3, so the last stage reads Y and the stage input only (the compulsory 3S
// traffic), not U^n as well; stage 1 does not store Y1 = U^n + dt/6 k1: stage 2 rebuilds it from what it reads
// anyway, Y1 = U^n + (U2 - U^n) / 3 (U2 = U^n + dt/2 k1), which saves one write and one read of S.  FP64: the
// reference's accumulator form.  vin is the field's stage input (which is the base in stage 1) when
// `haveInput`; k holds the field's rates; the new values go to nxt.
template <class R>
__device__ __forceinline__ void combine_field(const StageArgs<R> &A, size_t o, bool haveInput, const typename Lane<R>::V &vin,
                                              const R (&k)[Lane<R>::N], R (&nxt)[Lane<R>::N]) {
    using V = typename Lane<R>::V;
    constexpr int NV = Lane<R>::N;
    V vo, vacc;
    if constexpr (kFast<R>) {
        const V vb = A.kind == 2 ? V{} : A.kind == 0 && haveInput ? vin : vload(A.base + o);
        V vy;
        if (A.kind == 0) {
            vy = vb;
        } else if (A.ystage) {
#pragma unroll
            for (int j = 0; j < NV; ++j) vy[j] = mad(vin[j] - vb[j], (R)(1.0 / 3.0), vb[j]);
        } else {
            vy = vload(A.acc + o);
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if (A.kind == 2) {
                nxt[j] = mad(A.wy, k[j], vy[j]);
            } else {
                nxt[j] = mad(A.w, k[j], vb[j]);
                vacc[j] = mad(A.wy, k[j], vy[j]);
            }
            vo[j] = nxt[j];
        }
    } else {
        const V vb = A.kind == 0 && haveInput ? vin : vload(A.base + o); // stage 1: base is the input
        V va;
        if (A.kind != 0) va = vload(A.acc + o);
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            R acc = A.kind == 0 ? (R)0 : va[j];
            nxt[j] = combine(A.kind, vb[j], k[j], A.w, acc);
            vo[j] = nxt[j];
            vacc[j] = acc;
        }
    }
    vstore(A.out + o, vo, A.nt);
    if (A.kind == 1 || (A.kind == 0 && !kFast<R>)) vstore(A.acc + o, vacc, A.nt);
}
// LDS staging of the stencil input (the `tile` variant): per field, the block's 4 rows plus the 3 + 1
// halo rows of the z-stencil (8 rows) and HV halo vectors on each side of the 64 lane vectors are loaded
// once into LDS (double-buffered over the fields); the lanes then read their x- and z-taps from LDS.
template <class R> struct Tile {
    static constexpr int NV = Lane<R>::N, HV = (3 + NV - 1) / NV, W = 64 + 2 * HV, ROWS = 8, BUFS = 2;
    using V = typename Lane<R>::V;
    V v[BUFS][ROWS][W];
};
// One field's stencil taps from an LDS tile: the 2 HV + 1 vectors of the lane's row (x) and the four other rows of
// its column (z), read by one block of ds_read_b128 that waits for them itself.  The reads are written out so that
// the compiler, which cannot tell the tile's two buffers apart, does not wait for the other buffer's loads (still
// in flight by design) before them.  base: LDS address of the wave's first staged row at the lane's first vector;
// the immediates are byte offsets from it (x taps in row X0, z taps in rows Z[0..3], column HV).  The outputs are
// early-clobber: each read is issued while base is still needed by the next.
template <class R, int X0, int Z0, int Z1, int Z2, int Z3>
__device__ __forceinline__ void lds_taps(unsigned base, typename Lane<R>::V (&x)[2 * Tile<R>::HV + 1],
                                         typename Lane<R>::V (&z)[4]) {
    using T = Tile<R>;
    constexpr int S = 16, RW = T::W * S, C = T::HV * S;
    if constexpr (T::HV == 1) {
        asm volatile("ds_read_b128 %0, %7 offset:%8\n\tds_read_b128 %1, %7 offset:%9\n\tds_read_b128 %2, %7 offset:%10\n\t"
                     "ds_read_b128 %3, %7 offset:%11\n\tds_read_b128 %4, %7 offset:%12\n\tds_read_b128 %5, %7 offset:%13\n\t"
                     "ds_read_b128 %6, %7 offset:%14\n\ts_waitcnt lgkmcnt(0)"
                     : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(z[0]), "=&v"(z[1]), "=&v"(z[2]), "=&v"(z[3])
                     : "v"(base), "i"(X0 * RW), "i"(X0 * RW + S), "i"(X0 * RW + 2 * S), "i"(Z0 * RW + C),
                       "i"(Z1 * RW + C), "i"(Z2 * RW + C), "i"(Z3 * RW + C)
                     : "memory");
    } else {
        static_assert(T::HV == 2, "two or four nodes per lane");
        asm volatile("ds_read_b128 %0, %9 offset:%10\n\tds_read_b128 %1, %9 offset:%11\n\tds_read_b128 %2, %9 offset:%12\n\t"
                     "ds_read_b128 %3, %9 offset:%13\n\tds_read_b128 %4, %9 offset:%14\n\tds_read_b128 %5, %9 offset:%15\n\t"
                     "ds_read_b128 %6, %9 offset:%16\n\tds_read_b128 %7, %9 offset:%17\n\tds_read_b128 %8, %9 offset:%18\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(x[0]), "=&v"(x[1]), "=&v"(x[2]), "=&v"(x[3]), "=&v"(x[4]), "=&v"(z[0]), "=&v"(z[1]),
                       "=&v"(z[2]), "=&v"(z[3])
                     : "v"(base), "i"(X0 * RW), "i"(X0 * RW + S), "i"(X0 * RW + 2 * S), "i"(X0 * RW + 3 * S),
                       "i"(X0 * RW + 4 * S), "i"(Z0 * RW + C), "i"(Z1 * RW + C), "i"(Z2 * RW + C), "i"(Z3 * RW + C)
                     : "memory");
    }
}
template <class R, int DX, int DZ, int MECH, bool TILE = false, bool LAT = false>
__device__ __forceinline__ void interior_body(const StageArgs<R> &A, int tl, Partial *sh, Tile<R> *tile = nullptr) {
    using V = typename Lane<R>::V;
    constexpr int NV = Lane<R>::N;
    constexpr int F = MECH ? 6 + 4 * MECH : 5;
    const DMesh<R> &m = A.m;
    const int bx = m.ri0 + (tl % A.rtx) * 64 * NV, bk = m.rk0 + (tl / A.rtx) * 4; // block's first node, first row
    const int lane = threadIdx.x % 64, wrow = threadIdx.x / 64;
    const int i0 = bx + lane * NV, k = bk + wrow;
    const bool live = i0 < m.ri1 && k < m.rk1;
    Partial p{0, 0, 0, 0, 0};
    R xi[5][NV], eta[5][NV];
    V centre[5];
    if constexpr (TILE) {
        using T = Tile<R>;
        constexpr int rlo = DZ ? -3 : -1; // first staged row relative to bk
        const size_t cnt = A.count;
        // Field f's tile (rows bk + rlo .. bk + rlo + 7, vectors bx - HV NV .. bx + (64 + HV) NV) goes straight from
        // global memory into LDS buffer f % 2 (global_load_lds, 16 bytes per lane; lane L of an instruction writes
        // slot L after the wave-uniform LDS address).  Wave w loads rows 2w and 2w + 1, which are 2 W contiguous
        // slots (halo vectors included): lane L of instruction c fills slot 64 c + L of them, i.e. row 2w + s / W,
        // vector s % W (s = 64 c + L), so three instructions cover both rows (the third with the remaining
        // 2 W - 128 lanes).  Every wave issues exactly DMA_PER_FIELD instructions per field, whatever the grid
        // edge: addresses of slots that only dead lanes read are clamped into the row instead of skipped, so the
        // counter arithmetic below is exact.
        constexpr int SPAN = 2 * T::W, DMA_PER_FIELD = (SPAN + 63) / 64;
        static_assert(DMA_PER_FIELD == 3, "two rows of a tile take three 64-lane loads");
        const auto dma = [&](const R *g, V *l) {
            __builtin_amdgcn_global_load_lds(static_cast<const void *>(g), (__attribute__((address_space(3))) void *)l, 16, 0, 0);
        };
        const auto clamp_row = [&](int kr) { return kr < m.nz ? kr : m.nz - 1; };
        const auto clamp_x = [&](int ix) { return ix < -m.off ? -m.off : ix + NV > m.ld - m.off ? m.ld - m.off - NV : ix; };
        const auto issue = [&](int f) {
            const R *fp = A.in + (size_t)f * cnt;
            auto &buf = tile->v[f % T::BUFS];
            V *first = &buf[2 * wrow][0]; // the wave's 2 W slots, contiguous
#pragma unroll
            for (int c = 0; c < DMA_PER_FIELD; ++c) {
                const int sl = 64 * c + lane, r = 2 * wrow + sl / T::W, v = sl % T::W;
                const R *g = fp + at(m, clamp_x(bx + (v - T::HV) * NV), clamp_row(bk + rlo + r));
                if (64 * c + 63 < SPAN || lane < SPAN - 64 * c) dma(g, first + 64 * c);
            }
        };
        // vmcnt / expcnt / lgkmcnt fields of s_waitcnt (gfx9 encoding); the unused counters at their maximum
        constexpr int kWaitNextField = (DMA_PER_FIELD & 15) | ((DMA_PER_FIELD >> 4) << 14) | (7 << 4) | (15 << 8);
        constexpr int kWaitAllLoads = (7 << 4) | (15 << 8);
        issue(0);
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            auto &buf = tile->v[f % T::BUFS];
            if (f + 1 < 5) {
                issue(f + 1); // into the other buffer, released by the barrier that ended field f - 1
                __builtin_amdgcn_s_waitcnt(kWaitNextField); // this wave's field-f loads have landed
            } else {
                __builtin_amdgcn_s_waitcnt(kWaitAllLoads);
            }
            __builtin_amdgcn_s_barrier(); // ... and every wave's
            if (live) {
                // rows relative to the wave's first staged row (wrow): the lane's own row is -rlo, tap l's row
                // offs(l) - rlo (l = 1 is the centre, in the x taps)
                V xt[2 * T::HV + 1], zt[4];
                const unsigned base = (unsigned)(uintptr_t)(__attribute__((address_space(3))) V *)&buf[wrow][lane];
                lds_taps<R, -rlo, offs<DZ>(0) - rlo, offs<DZ>(2) - rlo, offs<DZ>(3) - rlo, offs<DZ>(4) - rlo>(base, xt, zt);
                R row[(2 * T::HV + 1) * NV];
#pragma unroll
                for (int v = 0; v <= 2 * T::HV; ++v) {
                    const V x = xt[v];
                    if (v == T::HV) centre[f] = x;
#pragma unroll
                    for (int j = 0; j < NV; ++j) row[v * NV + j] = x[j];
                }
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    R r = 0;
#pragma unroll
                    for (int l = 0; l < 5; ++l) r = mad(coef<R, DX>(l), row[T::HV * NV + j + offs<DX>(l)], r);
                    xi[f][j] = over_h<R, kFlatLoads<MECH, LAT>>(r, m.h, m.hinv);
                }
                const V col[5] = {zt[0], centre[f], zt[1], zt[2], zt[3]};
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    R r = 0;
#pragma unroll
                    for (int l = 0; l < 5; ++l) r = mad(coef<R, DZ>(l), col[l][j], r);
                    eta[f][j] = over_h<R, kFlatLoads<MECH, LAT>>(r, m.h, m.hinv);
                }
            }
            __builtin_amdgcn_s_barrier(); // every wave's reads of this buffer have returned before it is refilled
        }
    }
    if (live) {
        const ptrdiff_t q = at(m, i0, k), s = m.ld;
        const size_t cnt = A.count;
        const R h = m.h;
#pragma unroll
        for (int f = 0; f < 5 && !TILE; ++f) {
            const R *fp = A.in + (size_t)f * cnt;
            // x: the lane's NV nodes plus HV vectors on each side (the biased stencil reaches 3 nodes)
            constexpr int HV = (3 + NV - 1) / NV;
            R row[(2 * HV + 1) * NV];
            V mid;
#pragma unroll
            for (int v = 0; v <= 2 * HV; ++v) {
                const V x = vload(fp + q + (v - HV) * NV);
                if (v == HV) mid = x, centre[f] = x;
#pragma unroll
                for (int j = 0; j < NV; ++j) row[v * NV + j] = x[j];
            }
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                R r = 0;
#pragma unroll
                for (int l = 0; l < 5; ++l) r = mad(coef<R, DX>(l), row[HV * NV + j + offs<DX>(l)], r);
                xi[f][j] = over_h<R, kFlatLoads<MECH, LAT>>(r, h, m.hinv);
            }
            // z: five rows of the same columns
            V col[5];
#pragma unroll
            for (int l = 0; l < 5; ++l) col[l] = offs<DZ>(l) == 0 ? mid : vload(fp + q + offs<DZ>(l) * s);
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                R r = 0;
#pragma unroll
                for (int l = 0; l < 5; ++l) r = mad(coef<R, DZ>(l), col[l][j], r);
                eta[f][j] = over_h<R, kFlatLoads<MECH, LAT>>(r, h, m.hinv);
            }
        }
        // separable metrics in FP64 (bit-identical to the host arrays; the FP32 path takes one FP64
        // reciprocal, within 1 FP32 ulp of the stored values) or the stored arrays
        R a[NV], b[NV], jac[NV];
        if (m.metric) {
            const double fk = m.fade[k], fdk = m.fadeder[k];
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                const double zx = m.tslope[i0 + j] * fk, jj = 1 + m.tz_[i0 + j] * fdk;
                if constexpr (kFast<R>) {
                    const double inv = 1 / jj;
                    a[j] = (R)(-zx * inv);
                    b[j] = (R)inv;
                } else {
                    a[j] = (R)(-zx / jj);
                    b[j] = (R)(1 / jj);
                }
                jac[j] = (R)jj;
            }
        } else {
            const V va = vload(m.a + q), vb = vload(m.b + q), vj = vload(m.jac + q);
#pragma unroll
            for (int j = 0; j < NV; ++j) a[j] = va[j], b[j] = vb[j], jac[j] = vj[j];
        }
        // layered models: materials depend on the row k only (grid-conformal layers), read per row
        const V vrho = m.by_row ? V(m.rho_k[k]) : vload(m.rho + q), vlam = m.by_row ? V(m.lam_k[k]) : vload(m.lam + q),
                vmu = m.by_row ? V(m.mu_k[k]) : vload(m.mu + q);
        R kv[MECH ? 6 : 5][NV];
        R ecur[3][NV]; // this stage's strains exx, ezz, exz (memory rebuild only)
        const bool src = m.si1 >= 0 && k >= m.sk0 && k <= m.sk1 && i0 + NV - 1 >= m.si0 && i0 <= m.si1;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const R rho = vrho[j];
            R dx[5], dz[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                dx[f] = mad(a[j], eta[f][j], xi[f][j]);
                dz[f] = b[j] * eta[f][j];
            }
            R o0, o1;
            if constexpr (kFast<R>) {
                // 1/rho once; the split (CFS-PML) and plain momentum forms agree up to rounding here
                const R rinv = (R)1 / rho;
                o0 = (xi[2][j] + mad(a[j], eta[2][j], b[j] * eta[4][j])) * rinv;
                o1 = (xi[4][j] + mad(a[j], eta[4][j], b[j] * eta[3][j])) * rinv;
                if (src) {
                    const ptrdiff_t qq = q + j;
                    const R g0 = m.has_grad ? m.grad0[qq] : m.src[qq] * m.dir0;
                    const R g1 = m.has_grad ? m.grad1[qq] : m.src[qq] * m.dir2;
                    o0 = mad(A.forcing * g0, rinv, o0);
                    o1 = mad(A.forcing * g1, rinv, o1);
                }
            } else {
                if (m.cpml) { // outside the strips the ADE stretch is the identity, but the momentum keeps its split form
                    o0 = xi[2][j] / rho + (a[j] * eta[2][j] + b[j] * eta[4][j]) / rho;
                    o1 = xi[4][j] / rho + (a[j] * eta[4][j] + b[j] * eta[3][j]) / rho;
                } else {
                    o0 = (dx[2] + dz[4]) / rho;
                    o1 = (dx[4] + dz[3]) / rho;
                }
                if (src) {
                    const ptrdiff_t qq = q + j;
                    if (m.has_grad) {
                        o0 += A.forcing * m.grad0[qq] / rho;
                        o1 += A.forcing * m.grad1[qq] / rho;
                    } else {
                        const R sw = m.src[qq];
                        o0 += A.forcing * (sw * m.dir0) / rho;
                        o1 += A.forcing * (sw * m.dir2) / rho;
                    }
                }
            }
            const R lambda = vlam[j], mu = vmu[j], div = dx[0] + dz[1];
            kv[0][j] = o0;
            kv[1][j] = o1;
            kv[2][j] = mad(lambda, div, 2 * mu * dx[0]);
            kv[3][j] = mad(lambda, div, 2 * mu * dz[1]);
            kv[4][j] = mu * (dz[0] + dx[1]);
            if constexpr (MECH > 0) kv[5][j] = lambda * div;
            ecur[0][j] = dx[0], ecur[1][j] = dz[1], ecur[2][j] = dz[0] + dx[1];
        }
        // GSLS, in three passes so that few rates are live at once: (1) the stress rates take each mechanism's
        // memories in turn (the reference's order); (2) velocities and stresses are combined, stored and
        // inspected; (3) mechanism by mechanism, the memories' rates are formed and their four fields combined.
        // Pass 3 reads the memories again, from the caches, instead of holding them in registers.
        // strains of the step's earlier stages (memory rebuild only); loaded again where needed after the
        // stores of this stage, so that they are not held in registers across them
        const auto load_strains = [&](R (&e3)[3][3][NV]) {
#pragma unroll
            for (int s = 0; s < 3; ++s)
                if (s < A.stage) {
                    const R *src3 = strain_src<R, MECH>(A, s);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const V e = vload(src3 + (size_t)c * cnt + q);
#pragma unroll
                        for (int j = 0; j < NV; ++j) e3[s][c][j] = e[j];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int j = 0; j < NV; ++j) e3[s][c][j] = 0;
                }
        };
        if constexpr (kRebuildMemory<R, MECH>) {
            R eprev[3][3][NV];
            load_strains(eprev);
#pragma unroll 1
            for (int l = 0; l < MECH; ++l) {
                const V vdp = m.by_row ? V(m.dP_k[l][k]) : vload(m.dP[l] + q), vds = m.by_row ? V(m.dS_k[l][k]) : vload(m.dS[l] + q);
#pragma unroll
                for (int f = 0; f < 3; ++f) { // xx, zz, xz (the yy memories enter Syy only)
                    const V m0 = vload(A.base + (size_t)(6 + 4 * l + f) * cnt + q);
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        R d[3], ks, ms;
#pragma unroll
                        for (int s = 0; s < 3; ++s)
                            d[s] = memory_drive(f, vdp[j], vds[j], eprev[s][0][j], eprev[s][1][j], eprev[s][2][j]);
                        kv[2 + f][j] -= memory_stage(A.stage, m.rates[l], (R)0, m0[j], d, A.hdt, A.dt, ks, ms);
                    }
                }
            }
        } else if constexpr (MECH > 0) {
#pragma unroll
            for (int l = 0; l < MECH; ++l)
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const V mv = vload(A.in + (size_t)(6 + 4 * l + f) * cnt + q);
#pragma unroll
                    for (int j = 0; j < NV; ++j) kv[2 + f][j] -= mv[j];
                }
        }
        {
            R nxt[5][NV];
#pragma unroll
            for (int f = 0; f < (MECH && !kRebuildMemory<R, MECH> ? 6 : 5); ++f) {
                const size_t o = (size_t)f * cnt + q;
                // stage input: the stencil centre; Syy's only where FP32 stage 2 rebuilds Y1 from it
                const V vin = f < 5 ? centre[f < 5 ? f : 0] : (kFast<R> && A.ystage ? vload(A.in + o) : V{});
                R nx[NV];
                combine_field<R>(A, o, f < 5, vin, kv[f], nx);
                if (A.kind == 2)
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        if (!isfinite(nx[j])) p.bad = 1;
                        if (f < 5) nxt[f < 5 ? f : 0][j] = nx[j];
                    }
            }
            if (A.kind == 2) {
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    R u[5];
#pragma unroll
                    for (int f = 0; f < 5; ++f) u[f] = nxt[f][j];
                    node_stats<R, kFast<R>>(vrho[j], vlam[j], vmu[j], jac[j], h, u, false, nullptr, p);
                }
            }
        }
        if constexpr (kRebuildMemory<R, MECH>) {
            if (A.stage < 3) { // this stage's strains, for the later stages of the step
                R *dst = strain_dst<R, MECH>(A);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    V e;
#pragma unroll
                    for (int j = 0; j < NV; ++j) e[j] = ecur[c][j];
                    vstore(dst + (size_t)c * cnt + q, e, A.nt);
                }
            } else { // U^{n+1} of every memory variable and of Syy
                R eprev[3][3][NV], yy[NV];
                load_strains(eprev);
#pragma unroll
                for (int j = 0; j < NV; ++j) yy[j] = 0;
#pragma unroll 1
                for (int l = 0; l < MECH; ++l) {
                    const V vdp = m.by_row ? V(m.dP_k[l][k]) : vload(m.dP[l] + q), vds = m.by_row ? V(m.dS_k[l][k]) : vload(m.dS[l] + q);
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        const size_t o = (size_t)(6 + 4 * l + f) * cnt + q;
                        const V m0 = vload(A.base + o);
                        V mn;
#pragma unroll
                        for (int j = 0; j < NV; ++j) {
                            R d[3], ks, ms;
#pragma unroll
                            for (int s = 0; s < 3; ++s)
                                d[s] = memory_drive(f, vdp[j], vds[j], eprev[s][0][j], eprev[s][1][j], eprev[s][2][j]);
                            const R m3 = memory_stage(3, m.rates[l], (R)0, m0[j], d, A.hdt, A.dt, ks, ms);
                            const R k4 = m.rates[l] * (memory_drive(f, vdp[j], vds[j], ecur[0][j], ecur[1][j], ecur[2][j]) - m3);
                            mn[j] = mad(A.dt6, ks + k4, m0[j]);
                            if (!isfinite(mn[j])) p.bad = 1;
                            if (f == 3) yy[j] += ms + m3; // m_0 + 2 m_1 + 2 m_2 + m_3 of the yy memory
                        }
                        vstore(A.out + o, mn, A.nt);
                    }
                }
                // Syy: k_s = lambda div_s - (sum of the yy memories' stage inputs), so that
                // U^{n+1} = U^n + dt / 6 (lambda (div_1 + 2 div_2 + 2 div_3 + div_4) - sum of their weighted sums)
                const size_t o = (size_t)5 * cnt + q;
                const V s0 = vload(A.base + o);
                V sn;
#pragma unroll
                for (int j = 0; j < NV; ++j) {
                    R dsum = ecur[0][j] + ecur[1][j];
#pragma unroll
                    for (int s = 0; s < 3; ++s) dsum = mad(s ? (R)2 : (R)1, eprev[s][0][j] + eprev[s][1][j], dsum);
                    sn[j] = mad(A.dt6, mad(vlam[j], dsum, -yy[j]), s0[j]);
                    if (!isfinite(sn[j])) p.bad = 1;
                }
                vstore(A.out + o, sn, A.nt);
            }
        } else if constexpr (MECH > 0) {
#pragma unroll
            for (int l = 0; l < MECH; ++l) {
                const int kk = 6 + 4 * l;
                const V vdp = m.by_row ? V(m.dP_k[l][k]) : vload(m.dP[l] + q), vds = m.by_row ? V(m.dS_k[l][k]) : vload(m.dS[l] + q);
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const size_t o = (size_t)(kk + f) * cnt + q;
                    const V mem = vload(A.in + o);
                    R rate[NV];
#pragma unroll
                    for (int j = 0; j < NV; ++j) {
                        const R dp = vdp[j], ds = vds[j], dl = dp - 2 * ds;
                        const R dx0 = xi[0][j] + a[j] * eta[0][j], dx1 = xi[1][j] + a[j] * eta[1][j];
                        const R dz0 = b[j] * eta[0][j], dz1 = b[j] * eta[1][j], div = dx0 + dz1;
                        const R drive = f == 0 ? dl * div + 2 * ds * dx0
                                        : f == 1 ? dl * div + 2 * ds * dz1
                                        : f == 2 ? ds * (dz0 + dx1)
                                                 : dl * div;
                        rate[j] = m.rates[l] * (drive - mem[j]);
                    }
                    R nx[NV];
                    combine_field<R>(A, o, true, mem, rate, nx);
                    if (A.kind == 2)
#pragma unroll
                        for (int j = 0; j < NV; ++j)
                            if (!isfinite(nx[j])) p.bad = 1;
                }
            }
        }
    }
    if (A.kind == 2) {
        block_reduce(p, sh);
        if (threadIdx.x == 0) A.partials[A.ipart0 + tl] = sh[0];
    }
}

// One launch per RK stage: blocks [0, gblocks) take the border tiles (gblocks is ngeneral rounded up to
// a multiple of the 8 XCDs, so the interior blocks keep their XCD-contiguous band mapping), the rest
// take the interior tiles.  A single in-order launch avoids cross-stream joins between stages.
// The FP32 elastic stage fits 96 VGPRs without spilling, which gives 5 waves per SIMD instead of 4.
// LAT: the small-grid kernel (a few waves per SIMD at most, so the register budget of 3 waves costs nothing).
template <class R, int MECH, bool LAT = false> constexpr int kWavesPerEU = LAT ? 3 : sizeof(R) == 4 && MECH == 0 ? 5 : 1;

// The reference's fault rule and peak update for step `step` on the reduced inspection r
// (CurvilinearSolver2D.cpp:576-590): not finite, or after 3/f an energy above 1e4 x the accepted peak, faults the
// step; otherwise the peak, the diagnostics row and the accepted count advance.  One thread.  Returns the fault.
__device__ inline bool apply_rule(const Partial &r, int *status, int *accepted, double *peak, double *diag, int row,
                                  int step, double dt, double frequency) {
    const bool finite = !r.bad && isfinite(r.energy);
    const double before = *peak;
    const bool fault = !finite || ((step + 1) * dt > 3 / frequency && before > 0 && r.energy > before * 10000);
    if (fault) {
        *status = 1;
    } else {
        *peak = fmax(before, r.energy);
        double *d = diag + 4 * (size_t)row;
        d[0] = r.energy, d[1] = r.velocity, d[2] = r.traction, d[3] = 1;
        *accepted = row + 1;
    }
    return fault;
}
// receiver row `row`: Vx and Vz at every probe of the state (whole block)
template <class R>
__device__ inline void sample_state(const R *state, const long long *probes, int nrec, double *samples, int row) {
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        double *o = samples + (size_t)row * 2 * nrec + 2 * r;
        o[0] = state[probes[2 * r]];
        o[1] = state[probes[2 * r + 1]];
    }
}
// The previous step's end (see StepEnd), by one block of a first stage launch.
template <class R> __device__ void end_previous_step(const StageArgs<R> &A, Partial *sh) {
    __shared__ int verdict;
    const StepEnd &E = A.end;
    Partial p{0, 0, 0, 0, 0};
    for (int t = threadIdx.x; t < E.n; t += THREADS) merge(p, E.partials[t]);
    block_reduce(p, sh);
    if (threadIdx.x == 0)
        verdict = apply_rule(sh[0], E.status, E.accepted, E.peak, E.diag, E.row, E.step, E.dt, E.frequency);
    __syncthreads();
    if (!verdict) sample_state(A.in, E.probes, E.nrec, E.samples, E.row);
}

// FP32 elastic CFS-PML strips below the surface band (depth >= 4: no closure, traction image or reduced stencil;
// no sponge, checked on the host).  Each lane owns NV x-adjacent nodes as in the interior (vector
// stencil loads, metrics recomputed from the 1D profiles), and each node's four ADE memories per axis, contiguous in
// the compact slot map (PointKernels2D.hpp:50-56), are read and written as one vector.  The arithmetic is
// general_body's for these nodes (PointKernels2D.hpp:121-172 with the stretched derivatives of :57-64) in FP32 with
// fused multiply-adds; the wavefields use the interior's accumulator form, the memories general_body's.
constexpr int PML_NV = 2; // nodes per PML lane (8-byte vectors: half the interior's live stencil values)
template <class R, int DX, int DZ> __device__ __forceinline__ void pml_body(const StageArgs<R> &A, int li, Partial *sh) {
    static_assert(kFast<R>, "FP32 only");
    constexpr int NV = PML_NV, HV = (3 + NV - 1) / NV;
    typedef R V __attribute__((ext_vector_type(NV)));
    using V4 = typename Lane<R>::V; // one node's four ADE memories
    const auto vl = [](const R *p) { return *reinterpret_cast<const V *>(p); };
    const DMesh<R> &m = A.m;
    const int n = li * THREADS + (int)threadIdx.x;
    Partial p{0, 0, 0, 0, 0};
    if (n < A.npml) {
        const int2 ik = A.pml[n];
        const int i0 = ik.x, k = ik.y, valid = min(NV, m.nx - 3 - i0);
        const ptrdiff_t q = at(m, i0, k), s = m.ld;
        const size_t cnt = A.count;
        const R h = m.h;
        R xi[5][NV], eta[5][NV];
        V centre[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const R *fp = A.in + (size_t)f * cnt;
            R row[(2 * HV + 1) * NV];
#pragma unroll
            for (int v = 0; v <= 2 * HV; ++v) {
                const V x = vl(fp + q + (v - HV) * NV);
                if (v == HV) centre[f] = x;
#pragma unroll
                for (int j = 0; j < NV; ++j) row[v * NV + j] = x[j];
            }
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                R r = 0;
#pragma unroll
                for (int l = 0; l < 5; ++l) r = mad(coef<R, DX>(l), row[HV * NV + j + offs<DX>(l)], r);
                xi[f][j] = over_h(r, h, m.hinv);
            }
            V col[5];
#pragma unroll
            for (int l = 0; l < 5; ++l) col[l] = offs<DZ>(l) == 0 ? centre[f] : vl(fp + q + offs<DZ>(l) * s);
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                R r = 0;
#pragma unroll
                for (int l = 0; l < 5; ++l) r = mad(coef<R, DZ>(l), col[l][j], r);
                eta[f][j] = over_h(r, h, m.hinv);
            }
        }
        R a[NV], b[NV], jac[NV];
        if (m.metric) {
            const double fk = m.fade[k], fdk = m.fadeder[k];
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                const double zx = m.tslope[i0 + j] * fk, jj = 1 + m.tz_[i0 + j] * fdk, inv = 1 / jj;
                a[j] = (R)(-zx * inv), b[j] = (R)inv, jac[j] = (R)jj;
            }
        } else {
            const V va = vl(m.a + q), vb = vl(m.b + q), vj = vl(m.jac + q);
#pragma unroll
            for (int j = 0; j < NV; ++j) a[j] = va[j], b[j] = vb[j], jac[j] = vj[j];
        }
        const V vrho = m.by_row ? V(m.rho_k[k]) : vl(m.rho + q), vlam = m.by_row ? V(m.lam_k[k]) : vl(m.lam + q),
                vmu = m.by_row ? V(m.mu_k[k]) : vl(m.mu + q);
        const R *fz = m.czf + 3 * (k - 3);
        const bool src = m.si1 >= 0 && k >= m.sk0 && k <= m.sk1 && i0 + NV - 1 >= m.si0 && i0 <= m.si1;
        R kv[5][NV];
        bool bad = false;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int i = i0 + j;
            const long long sx = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 0),
                            sz = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 1);
            const bool ox = j < valid && sx >= 0, oz = j < valid && sz >= 0;
            const V4 px = ox ? vload(A.pin + sx) : V4{}, pz = oz ? vload(A.pin + sz) : V4{};
            const R *fx = m.cxf + 3 * (i - 3);
            R rx[4], rz[4]; // memory rates: fields 0, 1 (velocity derivatives) and 2, 3 (momentum parts)
            // stretched derivative (PointKernels2D.hpp:57-64) with the profile's quotients per position
            const auto stretch = [](bool on, const R *f, R psi, R d, R &rate) -> R {
                if (!on) return rate = 0, d;
                rate = -mad(f[1], psi, f[2] * d);
                return mad(d, f[0], psi);
            };
            const R vx0 = stretch(ox, fx, px[0], xi[0][j], rx[0]), vx1 = stretch(ox, fx, px[1], xi[1][j], rx[1]);
            const R vz0 = stretch(oz, fz, pz[0], eta[0][j], rz[0]), vz1 = stretch(oz, fz, pz[1], eta[1][j], rz[1]);
            const R dx0 = mad(a[j], vz0, vx0), dz0 = b[j] * vz0, dx1 = mad(a[j], vz1, vx1), dz1 = b[j] * vz1;
            const R rinv = (R)1 / vrho[j];
            const R x2 = xi[2][j] * rinv, z2 = mad(a[j], eta[2][j], b[j] * eta[4][j]) * rinv;
            const R x3 = xi[4][j] * rinv, z3 = mad(a[j], eta[4][j], b[j] * eta[3][j]) * rinv;
            kv[0][j] = stretch(ox, fx, px[2], x2, rx[2]) + stretch(oz, fz, pz[2], z2, rz[2]);
            kv[1][j] = stretch(ox, fx, px[3], x3, rx[3]) + stretch(oz, fz, pz[3], z3, rz[3]);
            if (src) { // the body force where the lane meets the source's support (as in the interior)
                const ptrdiff_t qq = q + j;
                const R g0 = m.has_grad ? m.grad0[qq] : m.src[qq] * m.dir0;
                const R g1 = m.has_grad ? m.grad1[qq] : m.src[qq] * m.dir2;
                kv[0][j] = mad(A.forcing * g0, rinv, kv[0][j]);
                kv[1][j] = mad(A.forcing * g1, rinv, kv[1][j]);
            }
            const R lambda = vlam[j], mu = vmu[j], div = dx0 + dz1;
            kv[2][j] = mad(lambda, div, 2 * mu * dx0);
            kv[3][j] = mad(lambda, div, 2 * mu * dz1);
            kv[4][j] = mu * (dz0 + dx1);
            // the node's memories: combine (CurvilinearSolver2D.cpp:80-97) with general_body's accumulator form
#pragma unroll
            for (int axis = 0; axis < 2; ++axis) {
                if (!(axis ? oz : ox)) continue;
                const long long sl = axis ? sz : sx;
                const R *rate = axis ? rz : rx;
                const V4 pb = A.kind == 0 ? (axis ? pz : px) : vload(A.pbase + sl);
                const V4 pa = A.kind == 0 ? V4{} : vload(A.pacc + sl);
                V4 pn, pacc;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R acc = pa[c];
                    pn[c] = combine(A.kind, pb[c], rate[c], A.w, acc);
                    pacc[c] = acc;
                    if (A.kind == 2 && !(pn[c] <= (R)3.4028234663852886e38 && pn[c] >= (R)-3.4028234663852886e38)) bad = true;
                }
                vstore(A.pout + sl, pn, 0);
                if (A.kind != 2) vstore(A.pacc + sl, pacc, 0);
            }
        }
        // the wavefields (interior_body's FP32 accumulator form); a lane past the grid's last column stores its nodes
        // one by one
        R nxt[5][NV];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const size_t o = (size_t)f * cnt + q;
            const V vin = centre[f], vb = A.kind == 0 ? vin : A.kind == 2 ? V{} : vl(A.base + o);
            const V vy = A.kind == 0 ? vb : A.ystage ? V{} : vl(A.acc + o);
            V vo, va;
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                const R y = A.ystage ? mad(vin[j] - vb[j], (R)(1.0 / 3.0), vb[j]) : vy[j];
                if (A.kind == 2) {
                    nxt[f][j] = mad(A.wy, kv[f][j], y);
                } else {
                    nxt[f][j] = mad(A.w, kv[f][j], vb[j]);
                    va[j] = mad(A.wy, kv[f][j], y);
                }
                vo[j] = nxt[f][j];
            }
            if (valid == NV) {
                *reinterpret_cast<V *>(A.out + o) = vo;
                if (A.kind == 1) *reinterpret_cast<V *>(A.acc + o) = va;
            } else {
                for (int j = 0; j < valid; ++j) {
                    A.out[o + j] = vo[j];
                    if (A.kind == 1) A.acc[o + j] = va[j];
                }
            }
        }
        if (A.kind == 2) {
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                if (j >= valid) continue;
                R u[5];
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    u[f] = nxt[f][j];
                    if (!isfinite(u[f])) bad = true;
                }
                node_stats<R, true>(vrho[j], vlam[j], vmu[j], jac[j], h, u, false, nullptr, p);
            }
            if (bad) p.bad = 1;
        }
    }
    if (A.kind == 2) {
        block_reduce(p, sh);
        if (threadIdx.x == 0) A.partials[A.ngeneral + li] = sh[0];
    }
}

// Work of block b of a whole-grid launch: a border block index (border >= 0) or an interior block index
// (interior >= 0).  Blocks go round-robin over the 8 XCDs, so the launch is handled in groups of 8 blocks.
// Not spread: the border groups come first.  Spread: the gblocks / 8 border groups sit at groups floor(j T / Gb)
// of the T groups (j < Gb), evenly between the interior groups, so the latency-bound border bodies overlap the
// bandwidth-bound interior tiles instead of holding most of the device at the start of the launch.  Interior
// block i keeps i % 8 == b % 8 either way, so tile_of_block places every interior tile on the same XCD.
template <class R> __device__ inline void block_role(const StageArgs<R> &A, int b, int &border, int &interior) {
    if (!A.spread) {
        border = b < A.gblocks ? b : -1;
        interior = b < A.gblocks ? -1 : b - A.gblocks;
        return;
    }
    const long long g = b >> 3, Gb = A.gblocks >> 3, T = (A.gblocks + A.ninterior + 7) >> 3;
    const int before = (int)((g * Gb + T - 1) / T), upto = (int)(((g + 1) * Gb + T - 1) / T); // border groups < g, <= g
    if (upto > before) {
        border = before * 8 + (b & 7), interior = -1;
    } else {
        border = -1, interior = (int)(g - before) * 8 + (b & 7);
    }
}

// PART: 0 one launch for the whole grid (border blocks and interior tiles, see block_role); 1 the border blocks
// only; 2 the interior tiles only.  Parts 1 and 2 run as two concurrent launches, each with its own register budget.
// LAT: the small-grid variant (flat stencil loads at a 3-wave register budget; same arithmetic, same results).
template <class R, int DX, int DZ, int MECH, bool TILE = false, int PART = 0, bool LAT = false>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<R, MECH, LAT>))) void stage_kernel(
    StageArgs<R> A) {
    __shared__ Partial sh[kShSlots];
    if (*A.status) return;
    const int b = PART == 2 ? (int)blockIdx.x + A.gblocks : (int)blockIdx.x;
    // block 0 of the launch is dispatched first: the previous step's end overlaps the rest of it
    if (A.end.partials && (PART == 2 ? A.gblocks == 0 && blockIdx.x == 0 : b == 0)) {
        end_previous_step(A, sh);
        __syncthreads(); // sh is used again below
    }
    int border, interior;
    if constexpr (PART == 0) {
        block_role(A, b, border, interior);
    } else {
        border = PART == 1 ? b : -1;
        interior = PART == 2 ? b - A.gblocks : -1;
    }
    if (border >= 0) {
        if (border < A.ngeneral) {
            general_body<R, DX, DZ, MECH, LAT>(A, border, sh);
        } else if constexpr (kFast<R> && MECH == 0) {
            if (border >= A.pml0 && border < A.pml0 + A.npmlb) pml_body<R, DX, DZ>(A, border - A.pml0, sh);
        }
    } else if constexpr (PART == 1) {
        return;
    } else if (interior < A.ninterior) {
        if constexpr (TILE) {
            __shared__ Tile<R> tile;
            interior_body<R, DX, DZ, MECH, true>(A, tile_of_block(interior, A.ninterior), sh, &tile);
        } else {
            interior_body<R, DX, DZ, MECH, false, LAT>(A, tile_of_block(interior, A.ninterior), sh);
        }
    }
}

// Fixed-partition reduction of the per-tile partials (deterministic: same order every run).
__global__ __launch_bounds__(THREADS) void reduce_kernel(const Partial *partials, int n, Partial *level2,
                                                         const int *status) {
    __shared__ Partial sh[kShSlots];
    if (*status) return;
    const int per = (n + gridDim.x - 1) / gridDim.x, begin = blockIdx.x * per, end = min(n, begin + per);
    Partial p{0, 0, 0, 0, 0};
    for (int t = begin + threadIdx.x; t < end; t += THREADS) merge(p, partials[t]);
    block_reduce(p, sh);
    if (threadIdx.x == 0) level2[blockIdx.x] = sh[0];
}

// The reference's fault rule, peak update and receiver gather for step s
// (CurvilinearSolver2D.cpp:576-593), on the reduced diagnostics.
template <class R> struct FinalArgs {
    const Partial *partials; // level-2 partials
    int n;
    int *status;     // 0 running, 1 faulted
    int *accepted;   // steps accepted in this run
    double *peak;    // accepted peak energy
    double *diag;    // 4 per step row
    double *samples; // 2*nrec per step row
    const R *state;  // new state (Vx at probes[2r], Vz at probes[2r+1])
    const long long *probes;
    int nrec, row, step;
    double dt, frequency;
};
template <class R> __global__ __launch_bounds__(THREADS) void finalize_kernel(FinalArgs<R> A);
template <class R> __device__ void finalize_body(const FinalArgs<R> &A);
template <class R> __global__ __launch_bounds__(THREADS) void finalize_kernel(FinalArgs<R> A) { finalize_body(A); }

// ---- multi-shot batch: one launch per RK stage walks the tiles of every shot in the batch.
// work[item] = (shot, tile): tile >= 0 a border tile index, tile | BATCH_INTERIOR an interior tile
// (already XCD-band mapped), -1 padding.  args[shot] are the shot's stage arguments for this launch;
// kind < 0 marks a shot that has no work in it (finished, or another direction pair).  Grid-stride
// loop, so the grid may be smaller than the work list (persistent blocks).
constexpr int BATCH_INTERIOR = 1 << 30, BATCH_PML = 1 << 29;
template <class R, int DX, int DZ, int MECH>
__device__ __forceinline__ void batch_item(const StageArgs<R> *__restrict__ args, const int2 w, Partial *sh) {
    if (w.y < 0) return;
    const StageArgs<R> &A = args[w.x];
    if (A.kind < 0 || *A.status) return;
    if (w.y & BATCH_INTERIOR) {
        interior_body<R, DX, DZ, MECH>(A, w.y & ~BATCH_INTERIOR, sh);
    } else if (w.y & BATCH_PML) {
        if constexpr (kFast<R> && MECH == 0) pml_body<R, DX, DZ>(A, w.y & ~BATCH_PML, sh);
    } else {
        general_body<R, DX, DZ, MECH>(A, w.y, sh);
    }
}
// PERSIST = 0: one block per work item (grid = nwork); 1: grid-stride loop over the items
template <class R, int DX, int DZ, int MECH, int PERSIST>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<R, MECH>))) void batch_stage_kernel(
    const StageArgs<R> *__restrict__ args, const int2 *__restrict__ work, int nwork) {
    __shared__ Partial sh[kShSlots];
    if constexpr (!PERSIST) {
        batch_item<R, DX, DZ, MECH>(args, work[blockIdx.x], sh);
    } else {
        for (int item = blockIdx.x; item < nwork; item += gridDim.x) {
            batch_item<R, DX, DZ, MECH>(args, work[item], sh);
            __syncthreads(); // the next item reuses sh
        }
    }
}
// one block per shot: all partials of the shot's last stage -> fault rule, diagnostics row, receivers
template <class R> __global__ __launch_bounds__(THREADS) void batch_final_kernel(const FinalArgs<R> *__restrict__ fa) {
    const FinalArgs<R> &A = fa[blockIdx.x];
    if (A.n <= 0) return;
    finalize_body(A);
}
// zeroes the per-shot run counters at the start of a captured chunk (no memset nodes in graphs)
// start of a device-resident run: counters cleared and the peak seeded from a kernel argument (no host
// memory is read after the launch returns)
__global__ void run_begin_kernel(int *status, int *accepted, double *peak, double peak0) {
    if (threadIdx.x == 0) *status = 0, *accepted = 0, *peak = peak0;
}
// start of a continuation run: its rows count from 0; the fault latch and the accepted peak carry over
__global__ void run_continue_kernel(int *accepted) {
    if (threadIdx.x == 0) *accepted = 0;
}
__global__ void batch_begin_kernel(int *const *status, int *const *accepted, double *const *peak, const double *peak0,
                                   int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) *status[s] = 0, *accepted[s] = 0, *peak[s] = peak0[s];
}

template <class R> __device__ void finalize_body(const FinalArgs<R> &A) {
    __shared__ Partial sh[kShSlots];
    __shared__ int verdict;
    if (*A.status) return;
    Partial p{0, 0, 0, 0, 0};
    for (int t = threadIdx.x; t < A.n; t += THREADS) merge(p, A.partials[t]);
    block_reduce(p, sh);
    if (threadIdx.x == 0)
        verdict = apply_rule(sh[0], A.status, A.accepted, A.peak, A.diag, A.row, A.step, A.dt, A.frequency);
    __syncthreads();
    if (verdict) return;
    sample_state(A.state, A.probes, A.nrec, A.samples, A.row);
}

// Inspection of the accepted state (CudaEngine::diagnostics path), 64x4 tiles over all nodes.
template <class R>
__global__ __launch_bounds__(THREADS) void stats_kernel(DMesh<R> m, const R *w, const R *pml, size_t count, Partial *partials,
                                                        int ntx, int ntiles) {
    __shared__ Partial sh[kShSlots];
    const int t = tile_of_block(blockIdx.x, ntiles);
    const int i = 3 + (t % ntx) * TX + (threadIdx.x % TX), k = 3 + (t / ntx) * TZ + (threadIdx.x / TX);
    Partial p{0, 0, 0, 0, 0};
    if (i < m.nx - 3 && k < m.nz - 3) {
        const ptrdiff_t q = at(m, i, k);
        R u[5];
        for (int f = 0; f < m.fields; ++f) {
            const R x = w[(size_t)f * count + q];
            if (f < 5) u[f] = x;
            if (!isfinite(x)) p.bad = 1;
        }
        const R big = sizeof(R) == 4 ? (R)3.4028234663852886e38 : (R)1.7976931348623157e308;
        for (int axis = 0; axis < 2; ++axis) {
            const long long sl = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, axis);
            if (sl >= 0)
                for (int c = 0; c < 4; ++c)
                    if (!(pml[sl + c] <= big && pml[sl + c] >= -big)) p.bad = 1;
        }
        const double nrm2[2] = {(double)m.nrm[2 * q], (double)m.nrm[2 * q + 1]};
        node_stats(m.rho[q], m.lam[q], m.mu[q], m.jac[q], m.h, u, k == 3, nrm2, p);
    }
    block_reduce(p, sh);
    if (threadIdx.x == 0) partials[t] = sh[0];
}

template <class R>
__global__ void gather_kernel(const R *w, size_t count, const long long *idx, int n, int fields, double *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        for (int f = 0; f < fields; ++f) out[(size_t)fields * r + f] = w[(size_t)f * count + idx[r]];
}

template <class R>
__global__ void frame_kernel(DMesh<R> m, const R *w, size_t count, int component, float *out, int nx, int nz, int off) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nx * nz) return;
    const ptrdiff_t p = at(m, q % nx + 3 + off, q / nx + 3);
    const R vx = w[p], vz = w[count + p], sxx = w[2 * count + p], szz = w[3 * count + p];
    R v;
    // Point::value (PointKernels2D.hpp:173-180)
    if (component == 0) v = vz;
    else if (component == 1) v = vx;
    else if (component == 2) v = sqrt(vx * vx + vz * vz);
    else if (m.mech) v = -(sxx + szz + w[5 * count + p]) / 3;
    else v = -(1 + m.lam[p] / (2 * (m.lam[p] + m.mu[p]))) * (sxx + szz) / 3;
    out[q] = (float)v;
}

__global__ __launch_bounds__(256) void stream_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

} // namespace
} // namespace swl

#include "engine2d_host.inl"
#include "batch2d_host.inl"
