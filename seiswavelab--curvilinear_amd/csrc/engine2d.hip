// engine2d.hip — gfx950 HIP engine for the 2D P-SV elastic / Qp-Qs GSLS curvilinear solver.
//
// Replaces the reference's CudaEngine (src/two_d/CudaEngine2D.cu, .hpp:5-24).  The per-node
// numerics follow src/two_d/PointKernels2D.hpp (rhs :121-172, project :70-78, closure :79-89,
// traction-image momentum :94-120, CFS-PML ADE :50-64) in the same floating-point order, but the
// execution model is the device's own:
//   * one fused kernel per RK stage: RHS + stage combine + surface projection + PML/GSLS memory
//     update, with the stage-4 kernel also producing the energy / finiteness inspection;
//   * three rotating state buffers plus one accumulator (acc = k1 + 2k2 + 2k3, rounded exactly as the
//     reference's a+2b+2c+d) instead of seven RK field sets; rollback is a buffer-index change;
//   * direction schedule as template parameters (no runtime stencil tables);
//   * 64x4-node tiles mapped XCD-contiguously (each of the 8 XCDs sweeps its own band of rows so
//     the stencil's z-neighbours are reused in that XCD's L2);
//   * device-side fault rule and receiver gather (swl_engine2d_run), no per-step host sync;
//   * FP32 runs keep the free-surface row (closure and projection) in FP64.
#include "common.hpp"
#include "swl.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace swl {
namespace {

constexpr int TX = 64, TZ = 4, THREADS = TX * TZ;

template <class R> struct DMesh {
    int nx, nz, wx, wz;
    long long cpml;
    int fields, mech, has_damp, has_grad;
    R h, rates[3], dir0, dir2;
    const R *a, *b, *jac, *lam, *mu, *rho, *src, *damp, *grad0, *grad1, *nrm, *cx, *cz;
    const R *dP[3], *dS[3];
    const double *sa, *sb, *slam, *smu, *snrm; // FP64 surface row k = 3, indexed by i
};

struct Partial {
    double energy, velocity, traction;
    int bad, pad;
};

template <class R> struct StageArgs {
    DMesh<R> m;
    const R *in, *base;
    R *out, *acc;
    const R *pin, *pbase;
    R *pout, *pacc;
    size_t count;
    R forcing, w;
    int kind; // 0 first stage, 1 middle stages, 2 final stage (+ inspection)
    const int *status;
    Partial *partials;
    int ntx, ntiles;
    const int *tiles; // tile ids handled by this launch (interior or general list)
    int nlist;
};

// ---- tile mapping: contiguous band of tiles per XCD (blocks are dispatched round-robin over 8 XCDs)
__device__ inline int tile_of_block(int b, int n) {
    const int x = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + j;
}

// ---- MacCormack-DRP operators (FiniteDifference.hpp:14-25), direction as a template parameter
template <class R, int DIR> __device__ inline R d5(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h) {
    const R c0 = DIR ? -(R)-.30874 : (R)-.30874, c1 = DIR ? -(R)-.6326 : (R)-.6326, c2 = DIR ? -(R)1.233 : (R)1.233,
            c3 = DIR ? -(R)-.3334 : (R)-.3334, c4 = DIR ? -(R).04168 : (R).04168;
    constexpr int o0 = DIR ? 1 : -1, o2 = DIR ? -1 : 1, o3 = DIR ? -2 : 2, o4 = DIR ? -3 : 3;
    R r = 0;
    r += c0 * v[q + o0 * s];
    r += c1 * v[q];
    r += c2 * v[q + o2 * s];
    r += c3 * v[q + o3 * s];
    r += c4 * v[q + o4 * s];
    return r / h;
}
template <class R, int DIR> __device__ inline R d3(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h) {
    const R l0 = ((R)-7) / ((R)6), l1 = ((R)8) / ((R)6), l2 = ((R)-1) / ((R)6);
    const R c0 = DIR ? -l0 : l0, c1 = DIR ? -l1 : l1, c2 = DIR ? -l2 : l2;
    constexpr int o1 = DIR ? -1 : 1, o2 = DIR ? -2 : 2;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    r += c2 * v[q + o2 * s];
    return r / h;
}
template <class R, int DIR> __device__ inline R d2(const R *__restrict__ v, ptrdiff_t q, ptrdiff_t s, R h) {
    const R c0 = DIR ? (R)1 : (R)-1, c1 = DIR ? (R)-1 : (R)1;
    constexpr int o1 = DIR ? -1 : 1;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    return r / h;
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

template <class R> struct Node {
    const DMesh<R> &m;
    const R *__restrict__ w;
    size_t count;
    ptrdiff_t q;
    int i, k;
    __device__ const R *f(int n) const { return w + (size_t)n * count; }
};

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
    const R c[5] = {(R)-.30874, (R)-.6326, (R)1.233, (R)-.3334, (R).04168};
    if (only != 1) {
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            const int o = DX ? -(l - 1) : (l - 1);
            const R cc = DX ? -c[l] : c[l];
            R sign = 1;
            result += sign * cc * flux(m, w, count, q + o, 0, comp);
        }
    }
    if (only != 0) {
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            int o = DZ ? -(l - 1) : (l - 1);
            const R cc = DZ ? -c[l] : c[l];
            R sign = 1;
            const int kk = depth + o;
            if (kk == 0) continue;
            if (kk < 0) {
                o = -kk - depth;
                sign = -1;
            }
            result += sign * cc * flux(m, w, count, q + (ptrdiff_t)o * m.nx, 1, comp);
        }
    }
    return result / (m.h * m.jac[q] * m.rho[q]);
}

// ADE memory rate and stretched derivative (PointKernels2D.hpp:57-64)
template <class R>
__device__ inline R stretched(const DMesh<R> &m, const R *__restrict__ pin, long long slot, int at, int axis, int field,
                              R d, R *rate) {
    if (slot < 0) return d;
    const R *pr = (axis ? m.cz : m.cx) + 3 * at;
    const R dd = pr[0], kappa = pr[1], alpha = pr[2], psi = pin[slot + field];
    rate[4 * axis + field] = -(dd / kappa + alpha) * psi - dd / (kappa * kappa) * d;
    return d / kappa + psi;
}

// traction-free closure for the surface z-derivatives, evaluated in FP64 (PointKernels2D.hpp:79-89)
template <class R>
__device__ inline void closure(const DMesh<R> &m, const R *__restrict__ w, size_t count, ptrdiff_t q, int i, R xi0, R xi1,
                               double *result) {
    const double c0 = m.sa[i], c1 = m.sb[i], c2 = c0 * c0 + c1 * c1, length = sqrt(c2);
    const double n0 = c0 / length, n1 = c1 / length;
    const double dx = xi0, dz = xi1, lambda = m.slam[i], mu = m.smu[i];
    double t0 = (lambda + 2 * mu) * dx * c0 + mu * dz * c1, t1 = mu * dz * c0 + lambda * dx * c1;
    for (int l = 0; l < m.mech; ++l) {
        const int k = 6 + 4 * l;
        const double sxx = w[(size_t)k * count + q], szz = w[(size_t)(k + 1) * count + q],
                     sxz = w[(size_t)(k + 2) * count + q];
        t0 -= sxx * c0 + sxz * c1;
        t1 -= sxz * c0 + szz * c1;
    }
    const double normal = t0 * n0 + t1 * n1;
    result[0] = -(t0 - n0 * normal) / (mu * c2) - n0 * normal / ((lambda + 2 * mu) * c2);
    result[1] = -(t1 - n1 * normal) / (mu * c2) - n1 * normal / ((lambda + 2 * mu) * c2);
}

// Full RHS of one physical node (PointKernels2D.hpp:121-172).  kout[F], prate[8] (X then Z memories).
// GENERAL = false is the interior specialisation used for tiles below the surface band (depth >= 4),
// outside the absorber strips and outside the source support: the surface closure, reduced
// stencils, traction image, ADE memories, damping and source terms are compiled out there, exactly
// as they are absent from the reference's arithmetic at those nodes.
template <class R, int DX, int DZ, int MECH, bool GENERAL>
__device__ inline void rhs_node(const DMesh<R> &m, const R *__restrict__ w, const R *__restrict__ pin, size_t count, int i,
                                int k, R forcing, R *kout, R *prate, long long &sx, long long &sz) {
    constexpr int F = MECH ? 6 + 4 * MECH : 5;
    const int depth = k - 3;
    const ptrdiff_t q = i + (ptrdiff_t)m.nx * k, s = m.nx;
    const R h = m.h;
    sx = sz = -1;
    if constexpr (GENERAL) {
        sx = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 0);
        sz = pml_slot(m.nx, m.nz, m.wx, m.wz, m.cpml, i, k, 1);
    }
    R xi[5], eta[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) xi[f] = d5<R, DX>(w + (size_t)f * count, q, 1, h);
#pragma unroll
    for (int f = 2; f < 5; ++f) eta[f] = d5<R, DZ>(w + (size_t)f * count, q, s, h);
    R vx[2], vz[2];
    if constexpr (GENERAL) {
        if (depth == 0) {
            R xs[2];
            for (int v = 0; v < 2; ++v) xs[v] = stretched(m, pin, sx, i - 3, 0, v, xi[v], prate);
            double sur[2];
            closure(m, w, count, q, i, xs[0], xs[1], sur);
            eta[0] = (R)sur[0];
            eta[1] = (R)sur[1];
        } else {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const R *fv = w + (size_t)v * count;
                eta[v] = depth == 1 ? d2<R, DZ>(fv, q, s, h) : depth == 2 ? d3<R, DZ>(fv, q, s, h) : d5<R, DZ>(fv, q, s, h);
            }
        }
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            vx[v] = stretched(m, pin, sx, i - 3, 0, v, xi[v], prate);
            vz[v] = stretched(m, pin, sz, k - 3, 1, v, eta[v], prate);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            eta[v] = d5<R, DZ>(w + (size_t)v * count, q, s, h);
            vx[v] = xi[v];
            vz[v] = eta[v];
        }
    }
    const R a = m.a[q], b = m.b[q], rho = m.rho[q];
    R dx[5], dz[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const R ux = f < 2 ? vx[f] : xi[f], uz = f < 2 ? vz[f] : eta[f];
        dx[f] = ux + a * uz;
        dz[f] = b * uz;
    }
    R o0 = (dx[2] + dz[4]) / rho, o1 = (dx[4] + dz[3]) / rho;
    if constexpr (GENERAL) {
        const bool image = depth <= (DZ ? 3 : 1);
        if (image && !m.cpml) {
            o0 = momentum<R, DX, DZ>(m, w, count, q, 0, depth, -1);
            o1 = momentum<R, DX, DZ>(m, w, count, q, 1, depth, -1);
        }
        if (m.cpml) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int f = v ? 4 : 2, other = v ? 3 : 4;
                R x = xi[f] / rho, z = (a * eta[f] + b * eta[other]) / rho;
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
            o0 += forcing * m.grad0[q] / rho;
            o1 += forcing * m.grad1[q] / rho;
        } else {
            const R sw = m.src[q];
            o0 += forcing * (sw * m.dir0) / rho;
            o1 += forcing * (sw * m.dir2) / rho;
        }
    } else if (m.cpml) {
        // outside the strips the ADE stretch is the identity, but the reference still splits the
        // momentum into its xi and eta parts (PointKernels2D.hpp:148-153)
        o0 = xi[2] / rho + (a * eta[2] + b * eta[4]) / rho;
        o1 = xi[4] / rho + (a * eta[4] + b * eta[3]) / rho;
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
            for (int f = 0; f < 4; ++f) mem[f] = w[(size_t)(kk + f) * count + q];
            kout[2] -= mem[0];
            kout[3] -= mem[1];
            kout[4] -= mem[2];
            kout[5] -= mem[3];
#pragma unroll
            for (int f = 0; f < 4; ++f) kout[kk + f] = m.rates[l] * (drive[f] - mem[f]);
        }
    }
    if constexpr (GENERAL) {
        if (m.has_damp) {
            const R d = m.damp[q];
#pragma unroll
            for (int f = 0; f < F; ++f) kout[f] -= d * w[(size_t)f * count + q];
        }
    }
}

// surface stress projection sigma <- P sigma P on row k = 3 (PointKernels2D.hpp:70-78), FP64
__device__ inline void project(const double *nrm, double &sxx, double &szz, double &sxz) {
    const double nx = nrm[0], nz = nrm[1];
    const double tx = sxx * nx + sxz * nz, tz = sxz * nx + szz * nz, nn = nx * tx + nz * tz;
    sxx += -2 * nx * tx + nn * nx * nx;
    szz += -2 * nz * tz + nn * nz * nz;
    sxz += -nx * tz - tx * nz + nn * nx * nz;
}

template <class R> __device__ inline R combine(int kind, R base, R k, R w, R &acc) {
    if (kind == 0) {
        acc = k;
        return base + w * k;
    }
    if (kind == 1) {
        acc = acc + 2 * k;
        return base + w * k;
    }
    return base + w * (acc + k) / (R)6;
}

template <class R> __device__ inline bool finite_r(R x) { return isfinite(x); }

template <class R> __device__ void block_reduce(Partial &p, Partial *sh) {
    const int t = threadIdx.x;
    sh[t] = p;
    __syncthreads();
    for (int s = THREADS / 2; s > 0; s >>= 1) {
        if (t < s) {
            Partial &a = sh[t];
            const Partial b = sh[t + s];
            a.energy += b.energy;
            a.velocity = fmax(a.velocity, b.velocity);
            a.traction = fmax(a.traction, b.traction);
            a.bad |= b.bad;
        }
        __syncthreads();
    }
}

template <class R> __device__ inline void node_stats(const DMesh<R> &m, int k, ptrdiff_t q, const R *u, Partial &p) {
    // PointKernels2D.hpp:181-184 energy; CurvilinearSolver2D.cpp:102-111 inspection
    const double l = m.lam[q], mm = m.mu[q], xx = u[2], zz = u[3], xz = u[4], vx = u[0], vz = u[1];
    p.energy = (.5 * (double)m.rho[q] * (vx * vx + vz * vz) +
                ((l + 2 * mm) * (xx * xx + zz * zz) - 2 * l * xx * zz) / (8 * mm * (l + mm)) + xz * xz / (2 * mm)) *
               (double)m.jac[q] * (double)m.h * (double)m.h;
    const R v = sqrt(u[0] * u[0] + u[1] * u[1]);
    p.velocity = v;
    if (!isfinite(p.energy) || !isfinite(p.velocity)) p.bad = 1;
    if (k == 3) {
        const double n0 = m.nrm[2 * q], n1 = m.nrm[2 * q + 1];
        const double tx = xx * n0 + xz * n1, tz = xz * n0 + zz * n1, norm = sqrt(xx * xx + zz * zz + 2 * xz * xz);
        p.traction = sqrt(tx * tx + tz * tz) / fmax(norm, 1e-25);
    }
}

template <class R, int DX, int DZ, int MECH, bool GENERAL>
__global__ __launch_bounds__(THREADS) void stage_kernel(StageArgs<R> A) {
    constexpr int F = MECH ? 6 + 4 * MECH : 5;
    __shared__ Partial sh[THREADS];
    if (*A.status) return;
    const int t = A.tiles[tile_of_block(blockIdx.x, A.nlist)];
    const int tx = t % A.ntx, tz = t / A.ntx;
    const int i = 3 + tx * TX + (threadIdx.x % TX), k = 3 + tz * TZ + (threadIdx.x / TX);
    const DMesh<R> &m = A.m;
    const bool live = i < m.nx - 3 && k < m.nz - 3;
    Partial p{0, 0, 0, 0, 0};
    if (live) {
        R kv[F], pr[8];
        long long sx, sz;
        rhs_node<R, DX, DZ, MECH, GENERAL>(m, A.in, A.pin, A.count, i, k, A.forcing, kv, pr, sx, sz);
        const ptrdiff_t q = i + (ptrdiff_t)m.nx * k;
        R nxt[F];
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const size_t at = (size_t)f * A.count + q;
            R acc = A.kind == 0 ? (R)0 : A.acc[at];
            nxt[f] = combine(A.kind, A.base[at], kv[f], A.w, acc);
            if (A.kind != 2) A.acc[at] = acc;
        }
        R pn[8];
        const long long slots[2] = {sx, sz};
        if constexpr (GENERAL) {
            if (k == 3) {
                double sxx = nxt[2], szz = nxt[3], sxz = nxt[4];
                project(m.snrm + 2 * i, sxx, szz, sxz);
                nxt[2] = (R)sxx, nxt[3] = (R)szz, nxt[4] = (R)sxz;
            }
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
        }
#pragma unroll
        for (int f = 0; f < F; ++f) A.out[(size_t)f * A.count + q] = nxt[f];
        if (A.kind == 2) {
#pragma unroll
            for (int f = 0; f < F; ++f)
                if (!finite_r(nxt[f])) p.bad = 1;
            if constexpr (GENERAL) {
                const R big = sizeof(R) == 4 ? (R)3.4028234663852886e38 : (R)1.7976931348623157e308;
#pragma unroll
                for (int axis = 0; axis < 2; ++axis)
                    if (slots[axis] >= 0)
                        for (int c = 0; c < 4; ++c)
                            if (!(pn[4 * axis + c] <= big && pn[4 * axis + c] >= -big)) p.bad = 1;
            }
            node_stats(m, GENERAL ? k : 4, q, nxt, p);
        }
    }
    if (A.kind == 2) {
        block_reduce<R>(p, sh);
        if (threadIdx.x == 0) A.partials[t] = sh[0];
    }
}

// Fixed-partition reduction of the per-tile partials (deterministic: same order every run).
__global__ __launch_bounds__(THREADS) void reduce_kernel(const Partial *partials, int ntiles, Partial *level2,
                                                         const int *status) {
    __shared__ Partial sh[THREADS];
    if (*status) return;
    const int per = (ntiles + gridDim.x - 1) / gridDim.x, begin = blockIdx.x * per, end = min(ntiles, begin + per);
    Partial p{0, 0, 0, 0, 0};
    for (int t = begin + threadIdx.x; t < end; t += THREADS) {
        const Partial b = partials[t];
        p.energy += b.energy;
        p.velocity = fmax(p.velocity, b.velocity);
        p.traction = fmax(p.traction, b.traction);
        p.bad |= b.bad;
    }
    block_reduce<double>(p, sh);
    if (threadIdx.x == 0) level2[blockIdx.x] = sh[0];
}

// The reference's fault rule, peak update and receiver gather for step s
// (CurvilinearSolver2D.cpp:576-593), on the reduced diagnostics.
template <class R> struct FinalArgs {
    const Partial *partials; // level-2 partials
    int ntiles;
    int *status;      // 0 running, 1 faulted
    int *accepted;    // steps accepted in this run
    double *peak;     // accepted peak energy
    double *diag;     // 4 per step row
    double *samples;  // 2*nrec per step row
    const R *state;   // new state (fields 0, 1 = Vx, Vz)
    const long long *probes;
    int nrec, row, step;
    double dt, frequency;
};
template <class R> __global__ __launch_bounds__(THREADS) void finalize_kernel(FinalArgs<R> A) {
    __shared__ Partial sh[THREADS];
    __shared__ int verdict;
    if (*A.status) return;
    Partial p{0, 0, 0, 0, 0};
    for (int t = threadIdx.x; t < A.ntiles; t += THREADS) {
        const Partial b = A.partials[t];
        p.energy += b.energy;
        p.velocity = fmax(p.velocity, b.velocity);
        p.traction = fmax(p.traction, b.traction);
        p.bad |= b.bad;
    }
    block_reduce<R>(p, sh);
    if (threadIdx.x == 0) {
        const Partial r = sh[0];
        const bool finite = !r.bad && isfinite(r.energy);
        const double before = *A.peak;
        const bool fault = !finite || ((A.step + 1) * A.dt > 3 / A.frequency && before > 0 && r.energy > before * 10000);
        if (fault) {
            *A.status = 1;
        } else {
            *A.peak = fmax(before, r.energy);
            double *d = A.diag + 4 * (size_t)A.row;
            d[0] = r.energy, d[1] = r.velocity, d[2] = r.traction, d[3] = 1;
            *A.accepted = A.row + 1;
        }
        verdict = fault;
    }
    __syncthreads();
    if (verdict) return;
    for (int r = threadIdx.x; r < A.nrec; r += THREADS) {
        double *o = A.samples + (size_t)A.row * 2 * A.nrec + 2 * r;
        o[0] = A.state[A.probes[2 * r]];
        o[1] = A.state[A.probes[2 * r + 1]];
    }
}

template <class R> __global__ void stats_kernel(DMesh<R> m, const R *w, const R *pml, size_t count, Partial *partials,
                                                int ntx, int ntiles) {
    __shared__ Partial sh[THREADS];
    const int t = tile_of_block(blockIdx.x, ntiles);
    const int i = 3 + (t % ntx) * TX + (threadIdx.x % TX), k = 3 + (t / ntx) * TZ + (threadIdx.x / TX);
    Partial p{0, 0, 0, 0, 0};
    if (i < m.nx - 3 && k < m.nz - 3) {
        const ptrdiff_t q = i + (ptrdiff_t)m.nx * k;
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
        node_stats(m, k, q, u, p);
    }
    block_reduce<R>(p, sh);
    if (threadIdx.x == 0) partials[t] = sh[0];
}

template <class R> __global__ void gather_kernel(const R *w, size_t count, const long long *idx, int n, int fields,
                                                 double *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        for (int f = 0; f < fields; ++f) out[(size_t)fields * r + f] = w[(size_t)f * count + idx[r]];
}

template <class R> __global__ void frame_kernel(DMesh<R> m, const R *w, size_t count, int component, float *out, int nx,
                                                int nz, int off) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nx * nz) return;
    const size_t p = (size_t)(q % nx + 3 + off) + (size_t)m.nx * (q / nx + 3);
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

// ---------------------------------------------------------------------------------------------
struct EngineBase {
    virtual ~EngineBase() = default;
};

template <class R> struct Engine2D : EngineBase {
    int device = 0, nrec = 0, ntx = 0, ntiles = 0, fields = 5, mech = 0, precision = 0;
    size_t count = 0;
    long long cpml = 0;
    hipStream_t stream{}, side{};
    DMesh<R> m{};
    DeviceBuffer mesh_arrays[24];
    // tiles below the surface band, outside absorber strips and source support take the
    // interior kernel; the rest take the general kernel on a second stream, concurrently
    DeviceBuffer tileLists, level2;
    int nInterior = 0, nGeneral = 0, nLevel2 = 0;
    hipEvent_t fork[4]{}, join[4]{};
    DeviceBuffer state[3], accbuf, pstate[3], paccbuf, partials, probes, status, accepted, peak, diag, samples, gather,
        display;
    int cur = 0, prev = 0;
    int diag_rows = 0;
    int registers = 0;
    size_t local_bytes = 0;
    std::vector<Partial> host_partials;
    // optional per-launch timing with HIP events on the engine stream
    bool profiling = false;
    std::vector<hipEvent_t> events;
    std::vector<int> eventKind; // 0 stage, 1 finalize (one entry per event pair)
    size_t eventsUsed = 0;
    double stageMs = 0, finalMs = 0;
    long long stageLaunches = 0, finalLaunches = 0;

    void mark(int kind, bool begin) {
        if (!profiling) return;
        if (eventsUsed == events.size()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "event");
            events.push_back(e);
        }
        check(hipEventRecord(events[eventsUsed++], stream), "event record");
        if (begin) eventKind.push_back(kind);
    }
    void collect() {
        if (!profiling) return;
        check(hipStreamSynchronize(stream), "timing");
        for (size_t j = 0; j + 1 < eventsUsed; j += 2) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, events[j], events[j + 1]), "event elapsed");
            if (eventKind[j / 2] == 0) stageMs += ms, ++stageLaunches;
            else finalMs += ms, ++finalLaunches;
        }
        eventsUsed = 0;
        eventKind.clear();
    }

    R *S(int b) const { return state[b].as<R>(); }
    R *P(int b) const { return pstate[b].as<R>(); }

    const R *put(int slot, const double *host, size_t n) {
        if (!host) return nullptr;
        std::vector<R> v(n);
        for (size_t i = 0; i < n; ++i) v[i] = (R)host[i];
        mesh_arrays[slot].upload(v.data(), n * sizeof(R));
        return mesh_arrays[slot].as<R>();
    }
    const double *put64(int slot, const double *host, size_t n) {
        mesh_arrays[slot].upload(host, n * sizeof(double));
        return mesh_arrays[slot].as<double>();
    }

    Engine2D(const swl_mesh2d_desc &d, const int64_t *rec, int n, int dev) : device(dev), nrec(n) {
        DeviceScope scope(device);
        precision = d.precision;
        count = (size_t)d.nx * d.nz;
        fields = d.field_count;
        mech = d.mechanisms;
        cpml = d.cpml_count;
        require(d.nx >= 14 && d.nz >= 14, "2D mesh needs at least 8 computational nodes per axis");
        require(mech >= 0 && mech <= 3 && fields == (mech ? 6 + 4 * mech : 5), "inconsistent GSLS field count");
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        m.nx = d.nx, m.nz = d.nz, m.wx = d.cpml_width_x, m.wz = d.cpml_width_z, m.cpml = cpml;
        m.fields = fields, m.mech = mech;
        m.h = (R)d.h;
        for (int l = 0; l < 3; ++l) m.rates[l] = (R)d.rates[l];
        m.dir0 = (R)d.source_direction[0], m.dir2 = (R)d.source_direction[2];
        m.a = put(0, d.a, count), m.b = put(1, d.b, count), m.jac = put(2, d.jac, count);
        m.lam = put(3, d.lambda, count), m.mu = put(4, d.mu, count), m.rho = put(5, d.rho, count);
        m.src = put(6, d.source, count), m.nrm = put(8, d.normals, 2 * count);
        m.has_damp = 0;
        if (d.damping)
            for (size_t q = 0; q < count && !m.has_damp; ++q) m.has_damp = d.damping[q] != 0;
        m.damp = m.has_damp ? put(7, d.damping, count) : nullptr;
        m.has_grad = d.source_gradient[0] != nullptr;
        if (m.has_grad) m.grad0 = put(9, d.source_gradient[0], count), m.grad1 = put(10, d.source_gradient[1], count);
        if (cpml) {
            m.cx = put(11, d.cpml_x, 3 * (size_t)(d.nx - 6));
            m.cz = put(12, d.cpml_z, 3 * (size_t)(d.nz - 6));
        }
        for (int l = 0; l < mech; ++l) m.dP[l] = put(13 + l, d.delta_p[l], count), m.dS[l] = put(16 + l, d.delta_s[l], count);
        // FP64 surface row (k = 3)
        std::vector<double> sa(d.nx), sb(d.nx), sl(d.nx), sm(d.nx), sn(2 * (size_t)d.nx);
        for (int i = 0; i < d.nx; ++i) {
            const size_t q = (size_t)i + (size_t)d.nx * 3;
            sa[i] = d.a[q], sb[i] = d.b[q], sl[i] = d.lambda[q], sm[i] = d.mu[q];
            sn[2 * i] = d.normals[2 * q], sn[2 * i + 1] = d.normals[2 * q + 1];
        }
        m.sa = put64(19, sa.data(), sa.size()), m.sb = put64(20, sb.data(), sb.size());
        m.slam = put64(21, sl.data(), sl.size()), m.smu = put64(22, sm.data(), sm.size());
        m.snrm = put64(23, sn.data(), sn.size());
        for (auto &b : state) b.allocate((size_t)fields * count * sizeof(R));
        accbuf.allocate((size_t)fields * count * sizeof(R));
        if (cpml) {
            for (auto &b : pstate) b.allocate((size_t)cpml * sizeof(R));
            paccbuf.allocate((size_t)cpml * sizeof(R));
        }
        ntx = (d.nx - 6 + TX - 1) / TX;
        ntiles = ntx * ((d.nz - 6 + TZ - 1) / TZ);
        partials.allocate((size_t)ntiles * sizeof(Partial));
        classify_tiles(d);
        nLevel2 = std::min(ntiles, 128);
        level2.allocate((size_t)nLevel2 * sizeof(Partial));
        check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "stream");
        for (int e = 0; e < 4; ++e) {
            check(hipEventCreateWithFlags(&fork[e], hipEventDisableTiming), "event");
            check(hipEventCreateWithFlags(&join[e], hipEventDisableTiming), "event");
        }
        host_partials.resize(ntiles);
        std::vector<long long> pi(2 * (size_t)std::max(n, 1));
        for (int r = 0; r < n; ++r) pi[2 * r] = rec[r], pi[2 * r + 1] = (long long)count + rec[r];
        probes.upload(pi.data(), pi.size() * sizeof(long long));
        status.allocate(sizeof(int));
        accepted.allocate(sizeof(int));
        peak.allocate(sizeof(double));
        gather.allocate((size_t)std::max(n, 1) * 5 * sizeof(double));
        display.allocate((size_t)(d.nx - 6) * (d.nz - 6) * sizeof(float));
        hipFuncAttributes attr{};
        if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&stage_kernel<R, 0, 0, 0, false>)) == hipSuccess) {
            registers = attr.numRegs;
            local_bytes = attr.localSizeBytes;
        }
    }
    ~Engine2D() override {
        for (auto e : events) (void)hipEventDestroy(e);
        for (int e = 0; e < 4; ++e) {
            if (fork[e]) (void)hipEventDestroy(fork[e]);
            if (join[e]) (void)hipEventDestroy(join[e]);
        }
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Tile classification: general = surface band (first tile row: depth 0..3), any overlap with the
    // absorber strips (where CFS-PML memories or sponge damping live) or with the source support.
    void classify_tiles(const swl_mesh2d_desc &d) {
        const int nxc = d.nx - 6, nzc = d.nz - 6;
        int si0 = d.nx, si1 = -1, sk0 = d.nz, sk1 = -1;
        const double *srcs[3] = {d.source, d.source_gradient[0], d.source_gradient[1]};
        for (const double *s : srcs) {
            if (!s) continue;
            for (int k = 0; k < d.nz; ++k)
                for (int i = 0; i < d.nx; ++i)
                    if (s[(size_t)i + (size_t)d.nx * k] != 0) {
                        si0 = std::min(si0, i), si1 = std::max(si1, i), sk0 = std::min(sk0, k), sk1 = std::max(sk1, k);
                    }
        }
        bool damp = false;
        int di0 = d.nx, di1 = -1, dk0 = d.nz; // damping support: columns < di0 or > di1, rows >= dk0
        if (d.damping) {
            for (int k = 3; k < d.nz - 3; ++k)
                for (int i = 3; i < d.nx - 3; ++i)
                    if (d.damping[(size_t)i + (size_t)d.nx * k] != 0) damp = true;
        }
        const int wx = cpml || damp ? std::max(d.cpml_width_x, 0) : 0, wz = cpml || damp ? std::max(d.cpml_width_z, 0) : 0;
        int swx = wx, swz = wz;
        if (damp && !cpml) { // sponge: nonzero where min(i, nx-1-i) < w or k >= nz - w (computational indices)
            swx = std::max(1, std::min(d.cpml_width_x, nxc / 3));
            swz = std::max(1, std::min(d.cpml_width_z, nzc / 3));
        }
        (void)di0, (void)di1, (void)dk0;
        std::vector<int> interior, general;
        for (int t = 0; t < ntiles; ++t) {
            const int tx = t % ntx, tz = t / ntx;
            const int i0 = tx * TX, i1 = std::min(nxc, i0 + TX) - 1, k0 = tz * TZ, k1 = std::min(nzc, k0 + TZ) - 1;
            bool g = tz == 0;
            if ((cpml || damp) && (i0 < swx || i1 >= nxc - swx || k1 >= nzc - swz)) g = true;
            if (si1 >= 0 && !(i1 + 3 < si0 || i0 + 3 > si1 || k1 + 3 < sk0 || k0 + 3 > sk1)) g = true;
            (g ? general : interior).push_back(t);
        }
        nInterior = (int)interior.size(), nGeneral = (int)general.size();
        std::vector<int> all(interior);
        all.insert(all.end(), general.begin(), general.end());
        tileLists.upload(all.data(), all.size() * sizeof(int));
    }
    const int *interiorList() const { return tileLists.as<int>(); }
    const int *generalList() const { return tileLists.as<int>() + nInterior; }

    template <int DX, int DZ, bool G> void launch_dirs(StageArgs<R> a, hipStream_t s) {
        a.tiles = G ? generalList() : interiorList();
        a.nlist = G ? nGeneral : nInterior;
        const dim3 grid(a.nlist), block(THREADS);
        switch (mech) {
        case 0: stage_kernel<R, DX, DZ, 0, G><<<grid, block, 0, s>>>(a); break;
        case 1: stage_kernel<R, DX, DZ, 1, G><<<grid, block, 0, s>>>(a); break;
        case 2: stage_kernel<R, DX, DZ, 2, G><<<grid, block, 0, s>>>(a); break;
        default: stage_kernel<R, DX, DZ, 3, G><<<grid, block, 0, s>>>(a); break;
        }
    }
    template <bool G> void launch_kind(int dx, int dz, const StageArgs<R> &a, hipStream_t s) {
        if (dx == 0 && dz == 0) launch_dirs<0, 0, G>(a, s);
        else if (dx == 0) launch_dirs<0, 1, G>(a, s);
        else if (dz == 0) launch_dirs<1, 0, G>(a, s);
        else launch_dirs<1, 1, G>(a, s);
        check(hipGetLastError(), "2D stage launch");
    }
    // General tiles on the side stream, interior tiles on the engine stream, joined before the
    // next stage (both read the same input buffer and write disjoint nodes).
    void launch_stage(int dx, int dz, const StageArgs<R> &a, int slot) {
        if (nGeneral) {
            check(hipEventRecord(fork[slot], stream), "fork");
            check(hipStreamWaitEvent(side, fork[slot], 0), "fork wait");
            launch_kind<true>(dx, dz, a, side);
            check(hipEventRecord(join[slot], side), "join");
        }
        if (nInterior) {
            mark(0, true);
            launch_kind<false>(dx, dz, a, stream);
            mark(0, false);
        }
        if (nGeneral) check(hipStreamWaitEvent(stream, join[slot], 0), "join wait");
    }

    static int direction(int step, int stage, int axis) {
        static const int flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                        {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
        return (flags[step % 8][axis] + stage) % 2;
    }
    static double pulse(double t, double f) {
        const double a = 3.14159265358979323846 * f * (t - 1.5 / f);
        return (1 - 2 * a * a) * std::exp(-a * a);
    }

    // Enqueue one RK4 step from buffer c; returns the buffer holding the new state.
    int enqueue_step(int c, double dt, double frequency, double scale, int step, bool inspect) {
        const int b1 = (c + 1) % 3, b2 = (c + 2) % 3;
        const int in[4] = {c, b1, b2, b1}, out[4] = {b1, b2, b1, b2};
        const double times[4] = {step * dt, (step + .5) * dt, (step + .5) * dt, (step + 1) * dt};
        const R delta = (R)dt;
        const R weights[4] = {(R)(dt / 2), (R)(dt / 2), (R)dt, delta};
        for (int s = 0; s < 4; ++s) {
            StageArgs<R> a{};
            a.m = m;
            a.in = S(in[s]), a.base = S(c), a.out = S(out[s]), a.acc = accbuf.as<R>();
            if (cpml) a.pin = P(in[s]), a.pbase = P(c), a.pout = P(out[s]), a.pacc = paccbuf.as<R>();
            a.count = count;
            a.forcing = (R)(scale * pulse(times[s], frequency));
            a.w = weights[s];
            a.kind = s == 0 ? 0 : s == 3 ? 2 : 1;
            a.status = status.as<int>();
            a.partials = partials.as<Partial>();
            a.ntx = ntx, a.ntiles = ntiles;
            (void)inspect;
            launch_stage(direction(step, s, 0), direction(step, s, 2), a, s);
        }
        return b2;
    }

    void step(double time, double dt, double frequency, double scale, int number) {
        (void)time;
        DeviceScope scope(device);
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "2D status clear");
        prev = cur;
        cur = enqueue_step(cur, dt, frequency, scale, number, false);
        check(hipStreamSynchronize(stream), "2D step");
    }
    void rollback() { cur = prev; }
    void reset() {
        DeviceScope scope(device);
        for (auto &b : state) check(hipMemsetAsync(b.ptr, 0, b.bytes, stream), "2D reset");
        check(hipMemsetAsync(accbuf.ptr, 0, accbuf.bytes, stream), "2D reset");
        if (cpml) {
            for (auto &b : pstate) check(hipMemsetAsync(b.ptr, 0, b.bytes, stream), "2D PML reset");
            check(hipMemsetAsync(paccbuf.ptr, 0, paccbuf.bytes, stream), "2D PML reset");
        }
        check(hipStreamSynchronize(stream), "2D clear completion");
        cur = prev = 0;
    }
    void diagnostics(swl_diag2d &d) {
        DeviceScope scope(device);
        stats_kernel<R><<<ntiles, THREADS, 0, stream>>>(m, S(cur), cpml ? P(cur) : nullptr, count,
                                                          partials.as<Partial>(), ntx, ntiles);
        check(hipGetLastError(), "2D diagnostics");
        check(hipMemcpyAsync(host_partials.data(), partials.ptr, ntiles * sizeof(Partial), hipMemcpyDeviceToHost, stream),
              "2D diagnostics readback");
        check(hipStreamSynchronize(stream), "2D diagnostics");
        d.energy = d.max_velocity = d.traction_relative = 0;
        d.finite = 1;
        for (const auto &p : host_partials) {
            d.energy += p.energy;
            d.max_velocity = std::max(d.max_velocity, p.velocity);
            d.traction_relative = std::max(d.traction_relative, p.traction);
            if (p.bad) d.finite = 0;
        }
        if (!std::isfinite(d.energy)) d.finite = 0;
        d.rhs_registers = registers;
        d.rhs_local_bytes = local_bytes;
    }
    void read_indices(const long long *dev_idx, int n, int nf, double *out) {
        gather.allocate((size_t)std::max(n, 1) * nf * sizeof(double));
        gather_kernel<R><<<(n + 127) / 128, 128, 0, stream>>>(S(cur), count, dev_idx, n, nf, gather.as<double>());
        check(hipGetLastError(), "2D gather");
        check(hipMemcpyAsync(out, gather.ptr, (size_t)n * nf * sizeof(double), hipMemcpyDeviceToHost, stream), "2D gather");
        check(hipStreamSynchronize(stream), "2D gather");
    }
    void sample(double *out) {
        DeviceScope scope(device);
        std::vector<long long> idx(nrec);
        std::vector<long long> pi(2 * (size_t)std::max(nrec, 1));
        check(hipMemcpy(pi.data(), probes.ptr, pi.size() * sizeof(long long), hipMemcpyDeviceToHost), "probe indices");
        for (int r = 0; r < nrec; ++r) idx[r] = pi[2 * r];
        DeviceBuffer di;
        di.upload(idx.data(), idx.size() * sizeof(long long));
        read_indices(di.as<long long>(), nrec, 2, out);
    }
    void probes_at(const int64_t *indices, int n, double *out) {
        DeviceScope scope(device);
        for (int r = 0; r < n; ++r) require(indices[r] >= 0 && (size_t)indices[r] < count, "probe index outside storage");
        DeviceBuffer di;
        std::vector<long long> v(indices, indices + n);
        di.upload(v.data(), v.size() * sizeof(long long));
        read_indices(di.as<long long>(), n, 5, out);
    }
    void snapshot(int f, double *out) {
        DeviceScope scope(device);
        require(f >= 0 && f < fields, "2D snapshot field out of range");
        std::vector<R> v(count);
        check(hipMemcpy(v.data(), S(cur) + (size_t)f * count, count * sizeof(R), hipMemcpyDeviceToHost), "2D snapshot");
        for (size_t q = 0; q < count; ++q) out[q] = v[q];
    }
    void pml(double *out) {
        DeviceScope scope(device);
        if (!cpml) return;
        std::vector<R> v(cpml);
        check(hipMemcpy(v.data(), P(cur), cpml * sizeof(R), hipMemcpyDeviceToHost), "2D PML snapshot");
        for (long long q = 0; q < cpml; ++q) out[q] = v[q];
    }
    void frame(int component, float *out, int nx, int nz, int off) {
        DeviceScope scope(device);
        require(component >= 0 && component < 4, "Invalid 2D display component.");
        require(nx >= 1 && nz >= 1 && off >= 0 && off + nx <= m.nx - 6 && nz <= m.nz - 6, "frame outside the domain");
        const int n = nx * nz;
        frame_kernel<R><<<(n + 127) / 128, 128, 0, stream>>>(m, S(cur), count, component, display.as<float>(), nx, nz, off);
        check(hipGetLastError(), "2D display launch");
        check(hipMemcpyAsync(out, display.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream), "2D display");
        check(hipStreamSynchronize(stream), "2D display");
    }
    uint64_t resident() const {
        uint64_t n = 0;
        for (const auto &b : mesh_arrays) n += b.bytes;
        for (const auto &b : state) n += b.bytes;
        for (const auto &b : pstate) n += b.bytes;
        for (const DeviceBuffer *b : {&accbuf, &paccbuf, &partials, &probes, &status, &accepted, &peak, &diag, &samples,
                                      &gather, &display})
            n += b->bytes;
        return n;
    }

    // Device-resident stepping with the reference's inspection / fault rule / sampling per step.
    void run(int first, int count_steps, double dt, double frequency, double scale, double peak0, double *diag_out,
             double *samples_out, swl_run2d &res) {
        DeviceScope scope(device);
        require(count_steps >= 0, "negative step count");
        res.accepted = 0, res.faulted = 0, res.peak_energy = peak0;
        if (!count_steps) return;
        if (diag_rows < count_steps) {
            diag.allocate((size_t)count_steps * 4 * sizeof(double));
            samples.allocate((size_t)count_steps * 2 * std::max(nrec, 1) * sizeof(double));
            diag_rows = count_steps;
        }
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "2D run status");
        check(hipMemsetAsync(accepted.ptr, 0, sizeof(int), stream), "2D run counter");
        check(hipMemcpyAsync(peak.ptr, &peak0, sizeof(double), hipMemcpyHostToDevice, stream), "2D run peak");
        std::vector<int> states(count_steps + 1);
        int c = cur;
        states[0] = c;
        for (int j = 0; j < count_steps; ++j) {
            const int s = first + j;
            const int n = enqueue_step(c, dt, frequency, scale, s, true);
            mark(1, true);
            reduce_kernel<<<nLevel2, THREADS, 0, stream>>>(partials.as<Partial>(), ntiles, level2.as<Partial>(),
                                                          status.as<int>());
            FinalArgs<R> f{};
            f.partials = level2.as<Partial>(), f.ntiles = nLevel2;
            f.status = status.as<int>(), f.accepted = accepted.as<int>(), f.peak = peak.as<double>();
            f.diag = diag.as<double>(), f.samples = samples.as<double>();
            f.state = S(n), f.probes = probes.as<long long>(), f.nrec = nrec, f.row = j, f.step = s;
            f.dt = dt, f.frequency = frequency;
            finalize_kernel<R><<<1, THREADS, 0, stream>>>(f);
            mark(1, false);
            check(hipGetLastError(), "2D finalize launch");
            c = n;
            states[j + 1] = c;
        }
        int acc_steps = 0, st = 0;
        double pk = peak0;
        check(hipMemcpyAsync(&acc_steps, accepted.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "2D run readback");
        check(hipMemcpyAsync(&st, status.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "2D run readback");
        check(hipMemcpyAsync(&pk, peak.ptr, sizeof(double), hipMemcpyDeviceToHost, stream), "2D run readback");
        check(hipStreamSynchronize(stream), "2D run");
        collect();
        if (acc_steps) {
            if (diag_out)
                check(hipMemcpy(diag_out, diag.ptr, (size_t)acc_steps * 4 * sizeof(double), hipMemcpyDeviceToHost),
                      "2D run diagnostics");
            if (samples_out && nrec)
                check(hipMemcpy(samples_out, samples.ptr, (size_t)acc_steps * 2 * nrec * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "2D run samples");
        }
        res.accepted = acc_steps;
        res.faulted = st != 0;
        res.peak_energy = pk;
        prev = acc_steps ? states[acc_steps - 1] : cur;
        cur = states[acc_steps];
    }
};

} // namespace
} // namespace swl

struct swl_engine2d {
    int precision;
    std::unique_ptr<swl::Engine2D<double>> e64;
    std::unique_ptr<swl::Engine2D<float>> e32;
    template <class F> void with(F &&f) {
        if (e64) f(*e64);
        else f(*e32);
    }
};

extern "C" {
int swl_engine2d_create(const swl_mesh2d_desc *mesh, const int64_t *receiver_index, int receivers, int device,
                        swl_engine2d **out) {
    return swl::guarded([&] {
        swl::require(mesh && out && receivers >= 0 && (receivers == 0 || receiver_index), "invalid 2D engine arguments");
        auto e = std::make_unique<swl_engine2d>();
        e->precision = mesh->precision;
        if (mesh->precision == 1) e->e32 = std::make_unique<swl::Engine2D<float>>(*mesh, receiver_index, receivers, device);
        else e->e64 = std::make_unique<swl::Engine2D<double>>(*mesh, receiver_index, receivers, device);
        *out = e.release();
    });
}
void swl_engine2d_destroy(swl_engine2d *e) { delete e; }
int swl_engine2d_step(swl_engine2d *e, double time, double dt, double frequency, double scale, int step) {
    return swl::guarded([&] { e->with([&](auto &x) { x.step(time, dt, frequency, scale, step); }); });
}
int swl_engine2d_rollback(swl_engine2d *e) {
    return swl::guarded([&] { e->with([&](auto &x) { x.rollback(); }); });
}
int swl_engine2d_reset(swl_engine2d *e) {
    return swl::guarded([&] { e->with([&](auto &x) { x.reset(); }); });
}
int swl_engine2d_diagnostics(swl_engine2d *e, swl_diag2d *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.diagnostics(*out); }); });
}
int swl_engine2d_sample(swl_engine2d *e, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.sample(out); }); });
}
int swl_engine2d_frame(swl_engine2d *e, int component, float *out, int nx, int nz, int offset_x) {
    return swl::guarded([&] { e->with([&](auto &x) { x.frame(component, out, nx, nz, offset_x); }); });
}
int swl_engine2d_snapshot(swl_engine2d *e, int field, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.snapshot(field, out); }); });
}
int swl_engine2d_pml(swl_engine2d *e, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.pml(out); }); });
}
int swl_engine2d_probes(swl_engine2d *e, const int64_t *indices, int count, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.probes_at(indices, count, out); }); });
}
uint64_t swl_engine2d_resident_bytes(swl_engine2d *e) {
    uint64_t n = 0;
    e->with([&](auto &x) { n = x.resident(); });
    return n;
}
int swl_engine2d_profile(swl_engine2d *e, int enable) {
    return swl::guarded([&] {
        e->with([&](auto &x) {
            x.profiling = enable != 0;
            x.stageMs = x.finalMs = 0;
            x.stageLaunches = x.finalLaunches = 0;
        });
    });
}
int swl_engine2d_kernel_timing(swl_engine2d *e, double *out) {
    return swl::guarded([&] {
        e->with([&](auto &x) {
            out[0] = x.stageMs, out[1] = (double)x.stageLaunches, out[2] = x.finalMs, out[3] = (double)x.finalLaunches;
            out[4] = (double)x.ntiles, out[5] = (double)x.registers;
        });
    });
}
int swl_engine2d_run(swl_engine2d *e, int first_step, int count, double dt, double frequency, double scale,
                     double peak_energy, double *diag, double *samples, swl_run2d *result) {
    return swl::guarded([&] {
        e->with([&](auto &x) { x.run(first_step, count, dt, frequency, scale, peak_energy, diag, samples, *result); });
    });
}
}

// ---------------------------------------------------------------------------------------------
// Plain streaming copy of a known size: calibrates hardware byte counters and measures the copy
// bandwidth this device actually sustains (a practical ceiling beside the 8 TB/s nominal peak).
namespace {
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
} // namespace
extern "C" int swl_stream_probe(int device, uint64_t bytes, int repeats, double *ms_per_copy) {
    return swl::guarded([&] {
        swl::require(bytes >= 16 && repeats > 0, "invalid probe size");
        swl::DeviceScope scope(device);
        const size_t n = bytes / sizeof(float4);
        swl::DeviceBuffer a, b;
        a.allocate(n * sizeof(float4));
        b.allocate(n * sizeof(float4));
        hipEvent_t e0, e1;
        swl::check(hipEventCreate(&e0), "event");
        swl::check(hipEventCreate(&e1), "event");
        const int grid = 256 * 8 * 4;
        stream_copy_kernel<<<grid, 256>>>(a.as<float4>(), b.as<float4>(), n);
        swl::check(hipEventRecord(e0, nullptr), "event");
        for (int r = 0; r < repeats; ++r) stream_copy_kernel<<<grid, 256>>>(a.as<float4>(), b.as<float4>(), n);
        swl::check(hipEventRecord(e1, nullptr), "event");
        swl::check(hipEventSynchronize(e1), "probe");
        float ms = 0;
        swl::check(hipEventElapsedTime(&ms, e0, e1), "event");
        *ms_per_copy = ms / repeats;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}
