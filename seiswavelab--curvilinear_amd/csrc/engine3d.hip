// engine3d.hip — gfx950 HIP engine for the 3D nine-field elastic curvilinear solver.
//
// Replaces the reference's CudaEngine3D (src/three_d/CudaEngine3D.hpp:7-24, .cu:68-150).  The per-node
// numerics follow src/three_d/PointKernels3D.hpp (rhs :92-112, surfaceDerivative :52-66,
// surfaceMomentum :67-91, project :36-51, energyDensity :113-120) in the same floating-point order,
// so FP64 runs reproduce the reference's CPU path (CpuSimulation3D::advance,
// src/three_d/CurvilinearSolver.cpp:445-513) to rounding.  Execution model:
//   * one fused kernel per RK stage: RHS + stage combine + surface projection; the last stage also
//     evaluates the energy / peaks / finiteness and the traction-rate residual from k4;
//   * three rotating state buffers and one accumulator.  The accumulator holds the partial sums of
//     base + dt/6 k1 + dt/3 k2 + dt/3 k3 in the reference's left-to-right order (CurvilinearSolver.cpp:472),
//     so the last stage reads the accumulator and its input only, and rollback is a buffer index;
//   * aligned layout (node (i,j,k) at off + i + ld*(j + ny*k), column 3 on a 64-element boundary);
//   * metrics and Jacobians are stored only for the terrain band k < flat_from (identity below, checked
//     when the engine is built); layered materials depending on k only and the separable sponge
//     rate (X(i) + Y(j)) + Z(k) are read from 1D tables;
//   * the surface closure and projection run in FP64 also in the FP32 variant;
//   * device-side fault rule with the source-window peak and receiver gather (swl_engine3d_run).
#include "common.hpp"
#include "swl.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

namespace swl {
namespace {

// A tile (one block, one k plane) is one 128-byte line of a row in x (16 FP64 / 32 FP32 nodes) by THREADS / that
// rows in y: a row segment is still one whole cache line, and the taller tile re-reads fewer y-halo rows.
constexpr int THREADS = 256;
template <class R> constexpr int TX = 128 / (int)sizeof(R);
template <class R> constexpr int TY = THREADS / TX<R>;
constexpr int ORIGIN = 61; // off: column 3 lands on a 64-element boundary

template <class R> struct DMesh3 {
    int nx, ny, nz, ld, off, flat;
    long long plane; // ld * ny
    R h, hinv;       // hinv = 1/h: FP32 always, FP64 when h is a power of two (exact), else 0 (divide)
    const R *met;    // 9 arrays of mcount elements (k < flat)
    const R *jac;    // mcount (k < flat)
    long long mcount;
    int by_depth;
    const R *lam, *mu, *rho; // per node, or per k when by_depth
    const R *dax, *day, *daz, *damp; // separable sponge tables, or full array
    R rho_global, dir[3];
    const R *src, *grad[3]; // full arrays, read inside the source box only
    int si0, si1, sj0, sj1, sk0, sk1;
    // FP64 surface plane (k = 3), indexed by i + ld*j
    const double *smet, *slam, *smu, *snrm;
};
template <class R> __device__ __host__ inline long long at3(const DMesh3<R> &m, int i, int j, int k) {
    return (long long)m.off + i + (long long)m.ld * j + m.plane * k;
}

struct Partial3 {
    double energy, velocity, peak[5], traction, norm;
    int bad, pad;
};
__device__ inline void merge3(Partial3 &a, const Partial3 &b) {
    a.energy += b.energy;
    a.velocity = fmax(a.velocity, b.velocity);
    for (int c = 0; c < 5; ++c) a.peak[c] = fmax(a.peak[c], b.peak[c]);
    a.traction = fmax(a.traction, b.traction);
    a.norm = fmax(a.norm, b.norm);
    a.bad |= b.bad;
}
__device__ void block_reduce3(Partial3 &p, Partial3 *sh) {
    const int t = threadIdx.x;
    sh[t] = p;
    __syncthreads();
    for (int s = THREADS / 2; s > 0; s >>= 1) {
        if (t < s) merge3(sh[t], sh[t + s]);
        __syncthreads();
    }
}

// Block reduction through wave shuffles and one partial per wave in LDS (4 x 88 bytes instead of 256 x 88).
__device__ inline double shfl_down_d(double v, int off) { return __shfl_down(v, off, 64); }
__device__ void block_reduce3_waves(Partial3 &p, Partial3 *sh4) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Partial3 o;
        o.energy = shfl_down_d(p.energy, off), o.velocity = shfl_down_d(p.velocity, off);
#pragma unroll
        for (int c = 0; c < 5; ++c) o.peak[c] = shfl_down_d(p.peak[c], off);
        o.traction = shfl_down_d(p.traction, off), o.norm = shfl_down_d(p.norm, off);
        o.bad = __shfl_down(p.bad, off, 64), o.pad = 0;
        merge3(p, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh4[wave] = p;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < THREADS / 64; ++w) merge3(p, sh4[w]);
}

template <class R> struct Stage3 {
    DMesh3<R> m;
    const R *in, *base;
    R *out, *acc;
    size_t count;
    R forcing, w, wy; // stage weight and accumulator weight (dt/6, dt/3, dt/3, dt/6)
    int kind;         // 0 first, 1 middle, 2 final (+ inspection)
    int ystage;       // stage 2 (FP32 deep bands rebuild Y1 = U^n + dt/6 k1 instead of reading it)
    const int *status;
    Partial3 *partials;
    int ntx, nty, ntiles; // tiles of TX x TY nodes per plane row / column, tiles of all planes
    int t0, nband;        // this launch covers tiles [t0, t0 + nband) (a band of k planes)
    int lds;              // deep bands: stage each field's tile and x / y halo in LDS (else direct loads)
    int bx, by, bz;       // tile bricks per XCD run (bx x by tiles by bz planes); bx = 0: whole tile columns
    int k0;               // the band's first plane
};

// Position of the u-th block of an XCD-contiguous run within a band of `planes` planes of ntx x nty tiles.
// bx = 0: tile columns, planes fastest.  Otherwise bricks of bx x by tiles by bz planes, in patch rows, patches,
// slabs; planes fastest within a brick, then x tiles, then y tiles.  Returns the tile's plane-major index.
__device__ inline int tile_in_band(int u, int planes, int ntx, int nty, int bx, int by, int bz) {
    const int perPlane = ntx * nty;
    if (bx <= 0) return (u % planes) * perPlane + u / planes;
    const int rowBlocks = by * ntx * planes, npy = (nty + by - 1) / by;
    const int py = min(u / rowBlocks, npy - 1), hy = min(by, nty - py * by);
    int r = u - py * rowBlocks;
    const int patchBlocks = bx * hy * planes, npx = (ntx + bx - 1) / bx;
    const int px = min(r / patchBlocks, npx - 1), hx = min(bx, ntx - px * bx);
    r -= px * patchBlocks;
    const int slabBlocks = hx * hy * bz, ns = (planes + bz - 1) / bz;
    const int s = min(r / slabBlocks, ns - 1), hz = min(bz, planes - s * bz);
    r -= s * slabBlocks;
    const int plane = s * bz + r % hz, cell = r / hz;
    const int tx = px * bx + cell % hx, ty = py * by + cell / hx;
    return plane * perPlane + ty * ntx + tx;
}

__device__ inline int tile_of_block(int b, int n) { // XCD-contiguous band of tiles
    const int x = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + j;
}

// ---- MacCormack-DRP operators (FiniteDifference.hpp:14-25)
template <class R, int DIR> __device__ inline R cf(int l) {
    const R c = l == 0 ? (R)-.30874 : l == 1 ? (R)-.6326 : l == 2 ? (R)1.233 : l == 3 ? (R)-.3334 : (R).04168;
    return DIR ? -c : c;
}
template <int DIR> __device__ constexpr int of(int l) { return DIR ? -(l - 1) : (l - 1); }
template <class R> __device__ inline R byh(R r, R h, R hinv) { return hinv != (R)0 ? r * hinv : r / h; }
template <class R, int DIR> __device__ inline R drp(const R *__restrict__ v, long long q, long long s, R h, R hinv) {
    R r = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) r += cf<R, DIR>(l) * v[q + of<DIR>(l) * s];
    return byh(r, h, hinv);
}
// reduced one-sided stencils for velocity z-derivatives at depth 1 (2-pt) and 2 (3-pt)
template <class R, int DIR> __device__ inline R low3(const R *__restrict__ v, long long q, long long s, R h, R hinv) {
    const R l0 = ((R)-7) / ((R)6), l1 = ((R)8) / ((R)6), l2 = ((R)-1) / ((R)6);
    const R c0 = DIR ? -l0 : l0, c1 = DIR ? -l1 : l1, c2 = DIR ? -l2 : l2;
    constexpr int o1 = DIR ? -1 : 1, o2 = DIR ? -2 : 2;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    r += c2 * v[q + o2 * s];
    return byh(r, h, hinv);
}
template <class R, int DIR> __device__ inline R low2(const R *__restrict__ v, long long q, long long s, R h, R hinv) {
    const R c0 = DIR ? (R)1 : (R)-1, c1 = DIR ? (R)-1 : (R)1;
    constexpr int o1 = DIR ? -1 : 1;
    R r = 0;
    r += c0 * v[q];
    r += c1 * v[q + o1 * s];
    return byh(r, h, hinv);
}

template <class R> __device__ inline R metric(const DMesh3<R> &m, long long q, int k, int a) {
    return k < m.flat ? m.met[(size_t)a * m.mcount + q] : (a % 4 == 0 ? (R)1 : (R)0);
}
// ZR: metrics 0..5 are known to be (1, 0, 0, 0, 1, 0) at every stored plane (TERRAIN_Z below): constants instead of
// loads, the same values in the same operations
template <class R, bool ZR> __device__ inline R metric_z(const DMesh3<R> &m, long long q, int k, int a) {
    if (ZR && a < 6) return a % 4 == 0 ? (R)1 : (R)0;
    return metric(m, q, k, a);
}
template <class R> __device__ inline R jacobian(const DMesh3<R> &m, long long q, int k) {
    return k < m.flat ? m.jac[q] : (R)1;
}

// traction flux J*(grad axis).sigma (PointKernels3D.hpp:67-72)
template <class R, bool ZR = false>
__device__ inline R tflux(const DMesh3<R> &m, const R *__restrict__ w, size_t count, long long q, int k, int axis, int comp) {
    const int f0 = comp == 0 ? 3 : comp == 1 ? 6 : 7, f1 = comp == 0 ? 6 : comp == 1 ? 4 : 8, f2 = comp == 0 ? 7 : comp == 1 ? 8 : 5;
    R v = 0;
    v += metric_z<R, ZR>(m, q, k, 3 * axis + 0) * w[(size_t)f0 * count + q];
    v += metric_z<R, ZR>(m, q, k, 3 * axis + 1) * w[(size_t)f1 * count + q];
    v += metric_z<R, ZR>(m, q, k, 3 * axis + 2) * w[(size_t)f2 * count + q];
    return jacobian(m, q, k) * v;
}
// surfaceMomentum (PointKernels3D.hpp:73-91)
template <class R, int DX, int DY, int DZ, bool ZR = false>
__device__ R smomentum(const DMesh3<R> &m, const R *__restrict__ w, size_t count, long long q, int k, int comp, int depth,
                       R rho) {
    R sum = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) sum += cf<R, DX>(l) * tflux<R, ZR>(m, w, count, q + of<DX>(l), k, 0, comp);
#pragma unroll
    for (int l = 0; l < 5; ++l) sum += cf<R, DY>(l) * tflux<R, ZR>(m, w, count, q + of<DY>(l) * (long long)m.ld, k, 1, comp);
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        int o = of<DZ>(l);
        R sign = 1;
        const int z = depth + o;
        if (z == 0) continue;
        if (z < 0) o = -z - depth, sign = -1;
        sum += sign * cf<R, DZ>(l) * tflux<R, ZR>(m, w, count, q + o * m.plane, k + o, 2, comp);
    }
    return sum / (m.h * jacobian(m, q, k) * rho);
}

// surfaceDerivative (PointKernels3D.hpp:52-66), FP64
template <class R, int DX, int DY>
__device__ void closure(const DMesh3<R> &m, const R *__restrict__ w, size_t count, long long q, int i, int j, double *res) {
    const long long sq = i + (long long)m.ld * j;
    const double *mt = m.smet + 9 * sq;
    const double lambda = m.slam[sq], mu = m.smu[sq];
    const double c[3] = {mt[6], mt[7], mt[8]};
    const double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double length = sqrt(c2), n[3] = {c[0] / length, c[1] / length, c[2] / length};
    const double h = (double)m.h;
    double g[9], tr[3] = {0, 0, 0};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const R *f = w + (size_t)v * count;
        double dx = 0, dy = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) dx += cf<double, DX>(l) * (double)f[q + of<DX>(l)];
#pragma unroll
        for (int l = 0; l < 5; ++l) dy += cf<double, DY>(l) * (double)f[q + of<DY>(l) * (long long)m.ld];
        dx = dx / h, dy = dy / h;
        for (int a = 0; a < 3; ++a) g[3 * v + a] = mt[a] * dx + mt[3 + a] * dy;
    }
    const double trace = g[0] + g[4] + g[8];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) tr[a] += ((a == b ? lambda * trace : 0) + mu * (g[3 * a + b] + g[3 * b + a])) * c[b];
    const double normal = tr[0] * n[0] + tr[1] * n[1] + tr[2] * n[2];
    for (int a = 0; a < 3; ++a) res[a] = -(tr[a] - n[a] * normal) / (mu * c2) - n[a] * normal / ((lambda + 2 * mu) * c2);
}

// project (PointKernels3D.hpp:36-51), FP64, on the six stresses of one surface node
__device__ inline void project9(const double *n, double *s6 /* xx yy zz xy xz yz */) {
    const double sg[9] = {s6[0], s6[3], s6[4], s6[3], s6[1], s6[5], s6[4], s6[5], s6[2]};
    double t[3] = {0, 0, 0}, o[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) t[a] += sg[3 * a + b] * n[b];
    const double nn = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o[3 * a + b] = sg[3 * a + b] - n[a] * t[b] - t[a] * n[b] + nn * n[a] * n[b];
    s6[0] = o[0], s6[1] = o[4], s6[2] = o[8];
    s6[3] = 0.5 * (o[1] + o[3]), s6[4] = 0.5 * (o[2] + o[6]), s6[5] = 0.5 * (o[5] + o[7]);
}

template <class R> __device__ inline R mat(const DMesh3<R> &m, const R *a, long long q, int k) {
    return m.by_depth ? a[k] : a[q];
}

// One RK stage over every physical node (one node per lane, TX x TY nodes per block, one k plane).
// Depth bands of one stage: SURFACE = planes k = 3..6 (closure, reduced stencils, traction image; the
// general per-node path below), TERRAIN = deeper planes with stored metrics, FLAT = planes with identity
// metrics and J = 1 (below the terrain transition).  The deep bands evaluate the same expressions of
// PointKernels3D.hpp:92-112 field by field, so far fewer values are live at once.
// TERRAIN_Z: the terrain band of a grid whose coordinates are x = i h, y = j h exactly and z terrain-following, so that
// the fourth-order coordinate derivatives are (1, 0, z_xi) and (0, 1, z_eta) and six of the nine inverse metrics are
// exactly (1, 0, 0, 0, 1, 0) (checked on the host): only the z row (metrics 6, 7, 8) is read, and each Cartesian
// component drops its zero terms, which changes no value (at most the sign of a zero).
constexpr int SURFACE = 0, TERRAIN = 1, FLAT = 2, TERRAIN_Z = 3;

template <class R, int DX, int DY, int DZ, int BAND>
__device__ __forceinline__ void deep_node(const Stage3<R> &A, long long q, int i, int j, int k, Partial3 &p) {
    const DMesh3<R> &m = A.m;
    const size_t cnt = A.count;
    const long long sy = m.ld, sz = m.plane;
    const R *w = A.in;
    const R h = m.h, hinv = m.hinv;
    const R lambda = mat(m, m.lam, q, k), mu = mat(m, m.mu, q, k), density = mat(m, m.rho, q, k);
    R mt[9];
    if constexpr (BAND == TERRAIN) {
#pragma unroll
        for (int a = 0; a < 9; ++a) mt[a] = __builtin_nontemporal_load(m.met + (size_t)a * m.mcount + q); // read once
    }
    // derivatives of field f along xi, eta, zeta (DRP) and the requested Cartesian components
    auto deriv = [&](int f, R d[3]) {
        const R *v = w + (size_t)f * cnt;
        d[0] = drp<R, DX>(v, q, 1, h, hinv);
        d[1] = drp<R, DY>(v, q, sy, h, hinv);
        d[2] = drp<R, DZ>(v, q, sz, h, hinv);
    };
    auto comp = [&](const R d[3], int a) -> R {
        if constexpr (BAND == FLAT) return d[a]; // identity metrics: m[a]*d0 + m[3+a]*d1 + m[6+a]*d2 = d[a]
        else return mt[a] * d[0] + mt[3 + a] * d[1] + mt[6 + a] * d[2];
    };
    R kv[9], d[3];
    // velocity gradient -> stress rates
    R g[3][3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        deriv(f, d);
#pragma unroll
        for (int a = 0; a < 3; ++a) g[f][a] = comp(d, a);
    }
    const R div = g[0][0] + g[1][1] + g[2][2];
    kv[3] = lambda * div + 2 * mu * g[0][0];
    kv[4] = lambda * div + 2 * mu * g[1][1];
    kv[5] = lambda * div + 2 * mu * g[2][2];
    kv[6] = mu * (g[0][1] + g[1][0]);
    kv[7] = mu * (g[0][2] + g[2][0]);
    kv[8] = mu * (g[1][2] + g[2][1]);
    // momentum: ((g30 + g61) + g72), ((g60 + g41) + g82), ((g70 + g81) + g52), each / density
    deriv(6, d);
    const R g60 = comp(d, 0), g61 = comp(d, 1);
    deriv(3, d);
    const R m0 = comp(d, 0) + g61;
    deriv(4, d);
    const R m1 = g60 + comp(d, 1);
    deriv(7, d);
    const R g70 = comp(d, 0);
    kv[0] = (m0 + comp(d, 2)) / density;
    deriv(8, d);
    const R m2 = g70 + comp(d, 1);
    kv[1] = (m1 + comp(d, 2)) / density;
    deriv(5, d);
    kv[2] = (m2 + comp(d, 2)) / density;
    if (m.si1 >= 0 && i >= m.si0 && i <= m.si1 && j >= m.sj0 && j <= m.sj1 && k >= m.sk0 && k <= m.sk1) {
        for (int a = 0; a < 3; ++a)
            kv[a] += A.forcing * (m.grad[0] ? m.grad[a][q] : m.src[q] * m.dir[a]) * (m.rho_global / density);
    }
    const R damp = m.damp ? m.damp[q] : (m.dax[i] + m.day[j]) + m.daz[k];
    // the stage output and the accumulator are not read again within the stage: nontemporal stores keep them
    // from displacing the y- and z-halo lines of the input in L2 and the Infinity Cache
    R nxt[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        const size_t o = (size_t)f * cnt + q;
        const R kf = kv[f] - damp * w[o];
        if (A.kind == 2) {
            nxt[f] = __builtin_nontemporal_load(A.acc + o) + A.wy * kf;
        } else {
            const R b = __builtin_nontemporal_load(A.base + o);
            nxt[f] = b + A.w * kf;
            if constexpr (sizeof(R) == 4) {
                // FP32: stage 1 does not store Y1 = U^n + dt/6 k1; stage 2 rebuilds it from its input
                // U2 = U^n + dt/2 k1 (unprojected in these bands) as U^n + (U2 - U^n) / 3
                if (A.kind == 1) {
                    const R y = A.ystage ? b + (w[o] - b) * (R)(1.0 / 3.0) : __builtin_nontemporal_load(A.acc + o);
                    __builtin_nontemporal_store(y + A.wy * kf, A.acc + o);
                }
            } else {
                __builtin_nontemporal_store((A.kind == 0 ? b : __builtin_nontemporal_load(A.acc + o)) + A.wy * kf,
                                            A.acc + o);
            }
        }
        __builtin_nontemporal_store(nxt[f], A.out + o);
    }
    if (A.kind == 2) {
        bool bad = false;
        for (int f = 0; f < 9; ++f) bad = bad || !isfinite(nxt[f]);
        const double vx = nxt[0], vy = nxt[1], vz = nxt[2];
        const double v2 = vx * vx + vy * vy + vz * vz, sv = sqrt(v2);
        p.velocity = sv;
        p.peak[0] = fabs(vz), p.peak[1] = fabs(vx), p.peak[2] = fabs(vy), p.peak[3] = sv;
        const double sxx = nxt[3], syy = nxt[4], szz = nxt[5], sxy = nxt[6], sxz = nxt[7], syz = nxt[8];
        p.peak[4] = fabs((sxx + syy + szz) / 3.0);
        const double dl = lambda, dm = mu, dr = density, mean = (sxx + syy + szz) / 3;
        double dev = 0;
        const double tt[3] = {sxx - mean, syy - mean, szz - mean};
        for (int c = 0; c < 3; ++c) dev += tt[c] * tt[c];
        dev += 2 * sxy * sxy, dev += 2 * sxz * sxz, dev += 2 * syz * syz;
        const double bulk = dl + 2 * dm / 3, hh = (double)h;
        const double jj = BAND == FLAT ? 1.0 : (double)m.jac[q];
        p.energy = (0.5 * dr * v2 + dev / (4 * dm) + mean * mean / (2 * bulk)) * jj * hh * hh * hh;
        if (!isfinite(v2) || !isfinite(p.energy)) bad = true;
        p.bad = bad;
    }
}

// The deep-band body with each field's stencil input staged in LDS: the block's tile plus 4 columns on each side and
// the y halo of the stencil direction (1 and 3 rows), loaded once with contiguous row reads, so the x and y taps
// come from LDS and only the z taps are global loads (aligned whole rows).  Two buffers: the next field is loaded
// while the current one is used, one barrier per field.  The arithmetic is deep_node's, term for term.
template <class R, int DY> struct LdsTile {
    static constexpr int W = TX<R> + 8, YLO = DY ? 3 : 1, H = TY<R> + 4, N = W * H;
};
template <class R, int DX, int DY, int DZ, int BAND>
__device__ __forceinline__ void deep_node_lds(const Stage3<R> &A, long long q0, int tx, int ty, bool active, int i, int j,
                                              int k, R *lds, Partial3 &p) {
    using T = LdsTile<R, DY>;
    const DMesh3<R> &m = A.m;
    const size_t cnt = A.count;
    const long long sy = m.ld, sz = m.plane, q = q0 + tx + (long long)sy * ty;
    const R *w = A.in;
    const R h = m.h, hinv = m.hinv;
    constexpr int ord[9] = {0, 1, 2, 6, 3, 4, 7, 8, 5};
    const int c0 = (ty + T::YLO) * T::W + tx + 4; // this node in the tile
    auto stage = [&](int f, R *buf) {
        const R *v = w + (size_t)f * cnt + q0 - 4 - (long long)T::YLO * sy;
#pragma unroll
        for (int e = threadIdx.x; e < T::N; e += THREADS) {
            const int ry = e / T::W, rx = e - ry * T::W;
            buf[e] = v[rx + (long long)ry * sy];
        }
    };
    R mt[9];
    R lambda = 0, mu = 0, density = 1;
    if (active) {
        lambda = mat(m, m.lam, q, k), mu = mat(m, m.mu, q, k), density = mat(m, m.rho, q, k);
        if constexpr (BAND == TERRAIN) {
#pragma unroll
            for (int a = 0; a < 9; ++a) mt[a] = __builtin_nontemporal_load(m.met + (size_t)a * m.mcount + q);
        }
    }
    auto comp = [&](const R d[3], int a) -> R {
        if constexpr (BAND == FLAT) return d[a];
        else return mt[a] * d[0] + mt[3 + a] * d[1] + mt[6 + a] * d[2];
    };
    R g[3][3], ctr[9], kv[9], d[3];
    R g60 = 0, g61 = 0, m0 = 0, m1 = 0, g70 = 0, m2 = 0;
    stage(ord[0], lds);
    __syncthreads();
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        const int f = ord[n];
        const R *cur = lds + (n & 1) * T::N;
        if (n + 1 < 9) stage(ord[n + 1], lds + ((n + 1) & 1) * T::N);
        if (active) {
            ctr[f] = cur[c0];
            R r = 0;
#pragma unroll
            for (int l = 0; l < 5; ++l) r += cf<R, DX>(l) * cur[c0 + of<DX>(l)];
            d[0] = byh(r, h, hinv);
            r = 0;
#pragma unroll
            for (int l = 0; l < 5; ++l) r += cf<R, DY>(l) * cur[c0 + of<DY>(l) * T::W];
            d[1] = byh(r, h, hinv);
            const R *v = w + (size_t)f * cnt;
            r = 0;
#pragma unroll
            for (int l = 0; l < 5; ++l) r += cf<R, DZ>(l) * (of<DZ>(l) == 0 ? ctr[f] : v[q + of<DZ>(l) * sz]);
            d[2] = byh(r, h, hinv);
            if (n < 3) {
#pragma unroll
                for (int a = 0; a < 3; ++a) g[f][a] = comp(d, a);
            } else if (f == 6) {
                g60 = comp(d, 0), g61 = comp(d, 1);
            } else if (f == 3) {
                m0 = comp(d, 0) + g61;
            } else if (f == 4) {
                m1 = g60 + comp(d, 1);
            } else if (f == 7) {
                g70 = comp(d, 0);
                kv[0] = (m0 + comp(d, 2)) / density;
            } else if (f == 8) {
                m2 = g70 + comp(d, 1);
                kv[1] = (m1 + comp(d, 2)) / density;
            } else { // f == 5
                kv[2] = (m2 + comp(d, 2)) / density;
            }
        }
        __syncthreads();
    }
    if (!active) return;
    const R div = g[0][0] + g[1][1] + g[2][2];
    kv[3] = lambda * div + 2 * mu * g[0][0];
    kv[4] = lambda * div + 2 * mu * g[1][1];
    kv[5] = lambda * div + 2 * mu * g[2][2];
    kv[6] = mu * (g[0][1] + g[1][0]);
    kv[7] = mu * (g[0][2] + g[2][0]);
    kv[8] = mu * (g[1][2] + g[2][1]);
    if (m.si1 >= 0 && i >= m.si0 && i <= m.si1 && j >= m.sj0 && j <= m.sj1 && k >= m.sk0 && k <= m.sk1) {
        for (int a = 0; a < 3; ++a)
            kv[a] += A.forcing * (m.grad[0] ? m.grad[a][q] : m.src[q] * m.dir[a]) * (m.rho_global / density);
    }
    const R damp = m.damp ? m.damp[q] : (m.dax[i] + m.day[j]) + m.daz[k];
    R nxt[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        const size_t o = (size_t)f * cnt + q;
        const R kf = kv[f] - damp * ctr[f];
        if (A.kind == 2) {
            nxt[f] = __builtin_nontemporal_load(A.acc + o) + A.wy * kf;
        } else {
            const R b = __builtin_nontemporal_load(A.base + o);
            nxt[f] = b + A.w * kf;
            if constexpr (sizeof(R) == 4) {
                if (A.kind == 1) {
                    const R y = A.ystage ? b + (ctr[f] - b) * (R)(1.0 / 3.0) : __builtin_nontemporal_load(A.acc + o);
                    __builtin_nontemporal_store(y + A.wy * kf, A.acc + o);
                }
            } else {
                __builtin_nontemporal_store((A.kind == 0 ? b : __builtin_nontemporal_load(A.acc + o)) + A.wy * kf,
                                            A.acc + o);
            }
        }
        __builtin_nontemporal_store(nxt[f], A.out + o);
    }
    if (A.kind == 2) {
        bool bad = false;
        for (int f = 0; f < 9; ++f) bad = bad || !isfinite(nxt[f]);
        const double vx = nxt[0], vy = nxt[1], vz = nxt[2];
        const double v2 = vx * vx + vy * vy + vz * vz, sv = sqrt(v2);
        p.velocity = sv;
        p.peak[0] = fabs(vz), p.peak[1] = fabs(vx), p.peak[2] = fabs(vy), p.peak[3] = sv;
        const double sxx = nxt[3], syy = nxt[4], szz = nxt[5], sxy = nxt[6], sxz = nxt[7], syz = nxt[8];
        p.peak[4] = fabs((sxx + syy + szz) / 3.0);
        const double dl = lambda, dm = mu, dr = density, mean = (sxx + syy + szz) / 3;
        double dev = 0;
        const double tt[3] = {sxx - mean, syy - mean, szz - mean};
        for (int c = 0; c < 3; ++c) dev += tt[c] * tt[c];
        dev += 2 * sxy * sxy, dev += 2 * sxz * sxz, dev += 2 * syz * syz;
        const double bulk = dl + 2 * dm / 3, hh = (double)h;
        const double jj = BAND == FLAT ? 1.0 : (double)m.jac[q];
        p.energy = (0.5 * dr * v2 + dev / (4 * dm) + mean * mean / (2 * bulk)) * jj * hh * hh * hh;
        if (!isfinite(v2) || !isfinite(p.energy)) bad = true;
        p.bad = bad;
    }
}

template <class R, int DX, int DY, int DZ, int BAND, bool LDS = false, int SD = -1, bool ZR = false>
__global__ __launch_bounds__(THREADS) void stage3_kernel(Stage3<R> A) {
    if (*A.status) return;
    const DMesh3<R> &m = A.m;
    // Each XCD takes a contiguous run of tile columns, planes varying fastest and then the y tile: the z- and
    // y-neighbours of a block's tile (the 3 + 1 halo planes and rows) are tiles its XCD's other resident blocks work
    // on at the same time (L2 hits), instead of tiles a plane or a row of tiles later.  tl is the tile's
    // plane-major index within the band (planes from A.k0), t its partial slot, so partials keep their order.
    const int tilesPerPlane = A.ntx * A.nty, u = tile_of_block(blockIdx.x, A.nband), planes = A.nband / tilesPerPlane;
    const int tl = tile_in_band(u, planes, A.ntx, A.nty, A.bx, A.by, A.bz), t = A.t0 + tl;
    if constexpr (BAND != SURFACE && LDS) {
        __shared__ R lds[2 * LdsTile<R, DY>::N];
        __shared__ Partial3 sh4[THREADS / 64];
        const int per = A.ntx * A.nty, kk = tl / per, rem = tl % per;
        const int tx = threadIdx.x % TX<R>, ty = threadIdx.x / TX<R>, i0 = 3 + (rem % A.ntx) * TX<R>, j0 = 3 + (rem / A.ntx) * TY<R>;
        const int i = i0 + tx, j = j0 + ty, k = A.k0 + kk;
        const bool active = i < m.nx - 3 && j < m.ny - 3;
        Partial3 p{};
        deep_node_lds<R, DX, DY, DZ, BAND>(A, at3(m, i0, j0, k), tx, ty, active, i, j, k, lds, p);
        if (A.kind == 2) {
            block_reduce3_waves(p, sh4);
            if (threadIdx.x == 0) A.partials[t] = p;
        }
        return;
    }
    if constexpr (BAND != SURFACE && !LDS) {
        __shared__ Partial3 sh[THREADS];
        const int per = A.ntx * A.nty, kk = tl / per, rem = tl % per;
        const int i = 3 + (rem % A.ntx) * TX<R> + (threadIdx.x % TX<R>), j = 3 + (rem / A.ntx) * TY<R> + threadIdx.x / TX<R>, k = A.k0 + kk;
        Partial3 p{};
        if (i < m.nx - 3 && j < m.ny - 3) deep_node<R, DX, DY, DZ, BAND>(A, at3(m, i, j, k), i, j, k, p);
        if (A.kind == 2) {
            block_reduce3(p, sh);
            if (threadIdx.x == 0) A.partials[t] = sh[0];
        }
        return;
    }
    __shared__ Partial3 sh[THREADS];
    const int per = A.ntx * A.nty, kk = tl / per, rem = tl % per;
    const int i = 3 + (rem % A.ntx) * TX<R> + (threadIdx.x % TX<R>), j = 3 + (rem / A.ntx) * TY<R> + threadIdx.x / TX<R>, k = A.k0 + kk;
    Partial3 p{};
    if (i < m.nx - 3 && j < m.ny - 3) {
        const size_t cnt = A.count;
        const long long q = at3(m, i, j, k), sy = m.ld, sz = m.plane;
        const R *w = A.in;
        const R h = m.h, hinv = m.hinv;
        const int depth = SD >= 0 ? SD : k - 3; // SD >= 0: a launch over the one surface plane at that depth
        const R lambda = mat(m, m.lam, q, k), mu = mat(m, m.mu, q, k), density = mat(m, m.rho, q, k);
        double dzs[3] = {0, 0, 0};
        if (depth == 0) closure<R, DX, DY>(m, w, cnt, q, i, j, dzs);
        R grad[9][3];
        R mt[9];
#pragma unroll
        for (int a = 0; a < 9; ++a) mt[a] = metric_z<R, ZR>(m, q, k, a);
#pragma unroll
        for (int f = 0; f < 9; ++f) {
            const R *v = w + (size_t)f * cnt;
            R d[3];
            d[0] = drp<R, DX>(v, q, 1, h, hinv);
            d[1] = drp<R, DY>(v, q, sy, h, hinv);
            if (f < 3 && depth == 0) d[2] = (R)dzs[f];
            else if (f < 3 && depth == 1) d[2] = low2<R, DZ>(v, q, sz, h, hinv);
            else if (f < 3 && depth == 2) d[2] = low3<R, DZ>(v, q, sz, h, hinv);
            else d[2] = drp<R, DZ>(v, q, sz, h, hinv);
#pragma unroll
            for (int a = 0; a < 3; ++a) grad[f][a] = mt[a] * d[0] + mt[3 + a] * d[1] + mt[6 + a] * d[2];
        }
        R kv[9];
        kv[0] = (grad[3][0] + grad[6][1] + grad[7][2]) / density;
        kv[1] = (grad[6][0] + grad[4][1] + grad[8][2]) / density;
        kv[2] = (grad[7][0] + grad[8][1] + grad[5][2]) / density;
        if (depth <= (DZ ? 3 : 1))
            for (int a = 0; a < 3; ++a) kv[a] = smomentum<R, DX, DY, DZ, ZR>(m, w, cnt, q, k, a, depth, density);
        if (m.si1 >= 0 && i >= m.si0 && i <= m.si1 && j >= m.sj0 && j <= m.sj1 && k >= m.sk0 && k <= m.sk1) {
            for (int a = 0; a < 3; ++a)
                kv[a] += A.forcing * (m.grad[0] ? m.grad[a][q] : m.src[q] * m.dir[a]) * (m.rho_global / density);
        }
        const R div = grad[0][0] + grad[1][1] + grad[2][2];
        kv[3] = lambda * div + 2 * mu * grad[0][0];
        kv[4] = lambda * div + 2 * mu * grad[1][1];
        kv[5] = lambda * div + 2 * mu * grad[2][2];
        kv[6] = mu * (grad[0][1] + grad[1][0]);
        kv[7] = mu * (grad[0][2] + grad[2][0]);
        kv[8] = mu * (grad[1][2] + grad[2][1]);
        const R damp = m.damp ? m.damp[q] : (m.dax[i] + m.day[j]) + m.daz[k];
#pragma unroll
        for (int f = 0; f < 9; ++f) kv[f] -= damp * w[(size_t)f * cnt + q];
        // combine (CurvilinearSolver.cpp:466-472) with the accumulator partial sums
        R nxt[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) {
            const size_t o = (size_t)f * cnt + q;
            if (A.kind == 2) {
                nxt[f] = A.acc[o] + A.wy * kv[f];
            } else {
                const R b = A.base[o];
                nxt[f] = b + A.w * kv[f];
                A.acc[o] = (A.kind == 0 ? b : A.acc[o]) + A.wy * kv[f];
            }
        }
        if (depth == 0) { // projection on the surface plane, FP64
            const double *n = m.snrm + 3 * (i + (long long)m.ld * j);
            double s6[6];
            for (int f = 0; f < 6; ++f) s6[f] = (double)nxt[3 + f];
            project9(n, s6);
            for (int f = 0; f < 6; ++f) nxt[3 + f] = (R)s6[f];
        }
#pragma unroll
        for (int f = 0; f < 9; ++f) A.out[(size_t)f * cnt + q] = nxt[f];
        if (A.kind == 2) {
            // Impl::updateDiagnostics (CurvilinearSolver.cpp:226-245) on the new state, rate residual from k4
            bool bad = false;
            for (int f = 0; f < 9; ++f) bad = bad || !isfinite(nxt[f]);
            const double vx = nxt[0], vy = nxt[1], vz = nxt[2];
            const double v2 = vx * vx + vy * vy + vz * vz, sv = sqrt(v2);
            p.velocity = sv;
            p.peak[0] = fabs(vz), p.peak[1] = fabs(vx), p.peak[2] = fabs(vy), p.peak[3] = sv;
            const double sxx = nxt[3], syy = nxt[4], szz = nxt[5], sxy = nxt[6], sxz = nxt[7], syz = nxt[8];
            p.peak[4] = fabs((sxx + syy + szz) / 3.0);
            const double dl = lambda, dm = mu, dr = density;
            const double mean = (sxx + syy + szz) / 3;
            double dev = 0;
            const double tt[3] = {sxx - mean, syy - mean, szz - mean};
            for (int c = 0; c < 3; ++c) dev += tt[c] * tt[c];
            dev += 2 * sxy * sxy, dev += 2 * sxz * sxz, dev += 2 * syz * syz;
            const double bulk = dl + 2 * dm / 3;
            const double hh = (double)h;
            p.energy = (0.5 * dr * v2 + dev / (4 * dm) + mean * mean / (2 * bulk)) * (double)jacobian(m, q, k) * hh * hh * hh;
            if (!isfinite(v2) || !isfinite(p.energy)) bad = true;
            if (depth == 0) {
                const double *n = m.snrm + 3 * (i + (long long)m.ld * j);
                const double t3[3] = {(double)kv[3] * n[0] + (double)kv[6] * n[1] + (double)kv[7] * n[2],
                                      (double)kv[6] * n[0] + (double)kv[4] * n[1] + (double)kv[8] * n[2],
                                      (double)kv[7] * n[0] + (double)kv[8] * n[1] + (double)kv[5] * n[2]};
                p.traction = sqrt(t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2]);
                double norm2 = 0.0;
                for (int f = 3; f < 9; ++f) norm2 += (f >= 6 ? 2.0 : 1.0) * (double)kv[f] * (double)kv[f];
                p.norm = sqrt(norm2);
            }
            p.bad = bad;
        }
    }
    if (A.kind == 2) {
        block_reduce3(p, sh);
        if (threadIdx.x == 0) A.partials[t] = sh[0];
    }
}

// ---- deep bands with vector lanes: each lane owns NV x-adjacent nodes of one row (one vector of NV values), the
// LPR lanes of a 128-byte row segment and a block THREADS / LPR rows.  The x taps come from the aligned vectors
// around the lane's own, the y and z taps are whole aligned vectors: a field costs 11-12 vector loads per NV nodes
// instead of 13 scalar loads per node.  TERRAIN reads the lane's 9 metrics once, as vectors; FLAT has identity
// metrics.  Every node's arithmetic is deep_node's, term for term, so the results are the same to the bit as the
// one-node-per-lane body.
template <class R, int NV> struct VLane3 {
    typedef R V __attribute__((ext_vector_type(NV)));
};
template <class R, int NV> __device__ inline typename VLane3<R, NV>::V vld3(const R *p) {
    return *reinterpret_cast<const typename VLane3<R, NV>::V *>(p);
}
template <class R, int NV> __device__ inline typename VLane3<R, NV>::V vld3_nt(const R *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const typename VLane3<R, NV>::V *>(p));
}
template <class R, int NV> __device__ inline void vst3_nt(R *p, typename VLane3<R, NV>::V v) {
    __builtin_nontemporal_store(v, reinterpret_cast<typename VLane3<R, NV>::V *>(p));
}

template <class R, int DX, int DY, int DZ, int NV, int BAND>
__global__ __launch_bounds__(THREADS) void band_vec_kernel(Stage3<R> A) {
    using V = typename VLane3<R, NV>::V;
    constexpr int LPR = 128 / (NV * (int)sizeof(R)), VTX = LPR * NV, VROWS = THREADS / LPR; // 128-byte row segments
    // x window: the aligned vectors holding offsets -1..3 (DX = 0) or -3..1 (DX = 1) of every node of the lane
    constexpr int XLO = DX ? -3 : -1, XHI = DX ? 1 : 3;
    constexpr int WLO = (XLO - (NV - 1)) / NV, WHI = (NV - 1 + XHI) / NV; // floor division (XLO < 0)
    __shared__ Partial3 sh4[THREADS / 64];
    if (*A.status) return;
    const DMesh3<R> &m = A.m;
    const int per = A.ntx * A.nty, u = tile_of_block(blockIdx.x, A.nband), planes = A.nband / per;
    const int tv = tile_in_band(u, planes, A.ntx, A.nty, A.bx, A.by, A.bz), kk = tv / per, rem = tv % per;
    const int i0 = 3 + (rem % A.ntx) * VTX + (threadIdx.x % LPR) * NV, j = 3 + (rem / A.ntx) * VROWS + threadIdx.x / LPR;
    const int k = A.k0 + kk;
    Partial3 p{};
    if (i0 < m.nx - 3 && j < m.ny - 3) {
        const int valid = min(NV, m.nx - 3 - i0); // nodes of the lane inside the physical grid
        const size_t cnt = A.count;
        const long long q = at3(m, i0, j, k), sy = m.ld, sz = m.plane;
        const R h = m.h, hinv = m.hinv;
        R lambda[NV], mu[NV], density[NV];
        if (m.by_depth) {
#pragma unroll
            for (int n = 0; n < NV; ++n) lambda[n] = m.lam[k], mu[n] = m.mu[k], density[n] = m.rho[k];
        } else {
            const V l = vld3<R, NV>(m.lam + q), u2 = vld3<R, NV>(m.mu + q), r = vld3<R, NV>(m.rho + q);
#pragma unroll
            for (int n = 0; n < NV; ++n) lambda[n] = l[n], mu[n] = u2[n], density[n] = r[n];
        }
        R mt[BAND == FLAT ? 1 : 9][NV];
        if constexpr (BAND == TERRAIN || BAND == TERRAIN_Z) {
#pragma unroll
            for (int a = BAND == TERRAIN_Z ? 6 : 0; a < 9; ++a) { // read once
                const V x = vld3_nt<R, NV>(m.met + (size_t)a * m.mcount + q);
#pragma unroll
                for (int n = 0; n < NV; ++n) mt[a][n] = x[n];
            }
        }
        // Cartesian component a of a derivative triple at node n: identity metrics, or deep_node's comp()
        auto cmp = [&](const R (&d)[3][NV], int a, int n) -> R {
            if constexpr (BAND == FLAT) return d[a][n];
            else if constexpr (BAND == TERRAIN_Z) return a == 2 ? mt[8][n] * d[2][n] : d[a][n] + mt[6 + a][n] * d[2][n];
            else return mt[a][n] * d[0][n] + mt[3 + a][n] * d[1][n] + mt[6 + a][n] * d[2][n];
        };
        // DRP derivatives of field f along x, y, z at the lane's NV nodes (drp's sum order for each node)
        auto deriv = [&](int f, R d[3][NV]) {
            const R *v = A.in + (size_t)f * cnt;
            R win[(WHI - WLO + 1) * NV];
#pragma unroll
            for (int b = WLO; b <= WHI; ++b) {
                const V x = vld3<R, NV>(v + q + b * NV);
#pragma unroll
                for (int n = 0; n < NV; ++n) win[(b - WLO) * NV + n] = x[n];
            }
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                R r = 0;
#pragma unroll
                for (int l = 0; l < 5; ++l) r += cf<R, DX>(l) * win[-WLO * NV + n + of<DX>(l)];
                d[0][n] = byh(r, h, hinv);
            }
            R ry[NV], rz[NV];
#pragma unroll
            for (int n = 0; n < NV; ++n) ry[n] = 0, rz[n] = 0;
#pragma unroll
            for (int l = 0; l < 5; ++l) {
                const V a = of<DY>(l) == 0 ? V{} : vld3<R, NV>(v + q + of<DY>(l) * sy);
                const V b = of<DZ>(l) == 0 ? V{} : vld3<R, NV>(v + q + of<DZ>(l) * sz);
#pragma unroll
                for (int n = 0; n < NV; ++n) {
                    ry[n] += cf<R, DY>(l) * (of<DY>(l) == 0 ? win[-WLO * NV + n] : a[n]);
                    rz[n] += cf<R, DZ>(l) * (of<DZ>(l) == 0 ? win[-WLO * NV + n] : b[n]);
                }
            }
#pragma unroll
            for (int n = 0; n < NV; ++n) d[1][n] = byh(ry[n], h, hinv), d[2][n] = byh(rz[n], h, hinv);
        };
        R d[3][NV], kv[9][NV];
        {
            R g[3][3][NV];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                deriv(f, d);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int n = 0; n < NV; ++n) g[f][a][n] = cmp(d, a, n);
            }
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                const R div = g[0][0][n] + g[1][1][n] + g[2][2][n];
                kv[3][n] = lambda[n] * div + 2 * mu[n] * g[0][0][n];
                kv[4][n] = lambda[n] * div + 2 * mu[n] * g[1][1][n];
                kv[5][n] = lambda[n] * div + 2 * mu[n] * g[2][2][n];
                kv[6][n] = mu[n] * (g[0][1][n] + g[1][0][n]);
                kv[7][n] = mu[n] * (g[0][2][n] + g[2][0][n]);
                kv[8][n] = mu[n] * (g[1][2][n] + g[2][1][n]);
            }
        }
        // momentum, in deep_node's order
        R g60[NV], m0[NV], m1[NV], g70[NV], m2[NV];
        deriv(6, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) g60[n] = cmp(d, 0, n), m0[n] = cmp(d, 1, n); // m0 holds g61 until field 3
        deriv(3, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) m0[n] = cmp(d, 0, n) + m0[n];
        deriv(4, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) m1[n] = g60[n] + cmp(d, 1, n);
        deriv(7, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) g70[n] = cmp(d, 0, n), kv[0][n] = (m0[n] + cmp(d, 2, n)) / density[n];
        deriv(8, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) m2[n] = g70[n] + cmp(d, 1, n), kv[1][n] = (m1[n] + cmp(d, 2, n)) / density[n];
        deriv(5, d);
#pragma unroll
        for (int n = 0; n < NV; ++n) kv[2][n] = (m2[n] + cmp(d, 2, n)) / density[n];
        if (m.si1 >= 0 && j >= m.sj0 && j <= m.sj1 && k >= m.sk0 && k <= m.sk1) {
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                const int i = i0 + n;
                if (i < m.si0 || i > m.si1) continue;
                for (int a = 0; a < 3; ++a)
                    kv[a][n] += A.forcing * (m.grad[0] ? m.grad[a][q + n] : m.src[q + n] * m.dir[a]) * (m.rho_global / density[n]);
            }
        }
        R damp[NV];
        if (m.damp) {
            const V dv = vld3<R, NV>(m.damp + q);
#pragma unroll
            for (int n = 0; n < NV; ++n) damp[n] = dv[n];
        } else {
#pragma unroll
            for (int n = 0; n < NV; ++n) damp[n] = (m.dax[i0 + n] + m.day[j]) + m.daz[k];
        }
        R nxt[9][NV];
#pragma unroll
        for (int f = 0; f < 9; ++f) {
            const size_t o = (size_t)f * cnt + q;
            const V c = vld3<R, NV>(A.in + o);
            V kf;
#pragma unroll
            for (int n = 0; n < NV; ++n) kf[n] = kv[f][n] - damp[n] * c[n];
            V out, accv;
            if (A.kind == 2) {
                const V y = vld3_nt<R, NV>(A.acc + o);
#pragma unroll
                for (int n = 0; n < NV; ++n) out[n] = y[n] + A.wy * kf[n];
            } else {
                const V b = vld3_nt<R, NV>(A.base + o);
#pragma unroll
                for (int n = 0; n < NV; ++n) out[n] = b[n] + A.w * kf[n];
                bool storeAcc = true;
                if constexpr (sizeof(R) == 4) {
                    if (A.kind == 1) {
                        const V y = A.ystage ? V{} : vld3_nt<R, NV>(A.acc + o);
#pragma unroll
                        for (int n = 0; n < NV; ++n)
                            accv[n] = (A.ystage ? b[n] + (c[n] - b[n]) * (R)(1.0 / 3.0) : y[n]) + A.wy * kf[n];
                    } else {
                        storeAcc = false; // stage 1: Y1 is rebuilt at stage 2
                    }
                } else {
                    const V y = A.kind == 0 ? b : vld3_nt<R, NV>(A.acc + o);
#pragma unroll
                    for (int n = 0; n < NV; ++n) accv[n] = y[n] + A.wy * kf[n];
                }
                if (storeAcc) {
                    if (valid == NV) vst3_nt<R, NV>(A.acc + o, accv);
                    else
                        for (int n = 0; n < valid; ++n) A.acc[o + n] = accv[n];
                }
            }
            if (valid == NV) vst3_nt<R, NV>(A.out + o, out);
            else
                for (int n = 0; n < valid; ++n) A.out[o + n] = out[n];
#pragma unroll
            for (int n = 0; n < NV; ++n) nxt[f][n] = out[n];
        }
        if (A.kind == 2) {
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                if (n >= valid) break;
                Partial3 e{};
                bool bad = false;
                for (int f = 0; f < 9; ++f) bad = bad || !isfinite(nxt[f][n]);
                const double vx = nxt[0][n], vy = nxt[1][n], vz = nxt[2][n];
                const double v2 = vx * vx + vy * vy + vz * vz, sv = sqrt(v2);
                e.velocity = sv;
                e.peak[0] = fabs(vz), e.peak[1] = fabs(vx), e.peak[2] = fabs(vy), e.peak[3] = sv;
                const double sxx = nxt[3][n], syy = nxt[4][n], szz = nxt[5][n], sxy = nxt[6][n], sxz = nxt[7][n],
                             syz = nxt[8][n];
                e.peak[4] = fabs((sxx + syy + szz) / 3.0);
                const double dl = lambda[n], dm = mu[n], dr = density[n], mean = (sxx + syy + szz) / 3;
                double dev = 0;
                const double tt[3] = {sxx - mean, syy - mean, szz - mean};
                for (int c = 0; c < 3; ++c) dev += tt[c] * tt[c];
                dev += 2 * sxy * sxy, dev += 2 * sxz * sxz, dev += 2 * syz * syz;
                const double bulk = dl + 2 * dm / 3, hh = (double)h;
                const double jj = BAND == FLAT ? 1.0 : (double)m.jac[q + n];
                e.energy = (0.5 * dr * v2 + dev / (4 * dm) + mean * mean / (2 * bulk)) * jj * hh * hh * hh;
                if (!isfinite(v2) || !isfinite(e.energy)) bad = true;
                e.bad = bad;
                merge3(p, e);
            }
        }
    }
    if (A.kind == 2) {
        block_reduce3_waves(p, sh4);
        if (threadIdx.x == 0) A.partials[A.t0 + tv] = p;
    }
}

__global__ __launch_bounds__(THREADS) void reduce3_kernel(const Partial3 *partials, int n, Partial3 *level2,
                                                          const int *status) {
    __shared__ Partial3 sh[THREADS];
    if (status && *status) return;
    const int per = (n + gridDim.x - 1) / gridDim.x, begin = blockIdx.x * per, end = min(n, begin + per);
    Partial3 p{};
    for (int t = begin + threadIdx.x; t < end; t += THREADS) merge3(p, partials[t]);
    block_reduce3(p, sh);
    if (threadIdx.x == 0) level2[blockIdx.x] = sh[0];
}

template <class R> struct Final3 {
    const Partial3 *partials;
    int n;
    int *status, *accepted;
    double *peak, *source_peak;
    double *diag, *samples, *last; // last: diagnostics of the latest evaluated step (10 values)
    const R *state;
    const long long *probes;
    int nrec, row, step, rule; // rule 0: always accept (single-step API)
    double dt, frequency;
};
__device__ inline void diag_row(const Partial3 &r, double *d) {
    const double ratio = r.traction / fmax(r.norm, 1.0e-30);
    const bool finite = !r.bad && isfinite(r.energy) && isfinite(ratio);
    d[0] = r.energy, d[1] = r.velocity, d[2] = ratio, d[3] = finite;
    for (int c = 0; c < 5; ++c) d[4 + c] = r.peak[c];
    d[9] = 0;
}
template <class R> __global__ __launch_bounds__(THREADS) void final3_kernel(Final3<R> A) {
    __shared__ Partial3 sh[THREADS];
    __shared__ int verdict;
    if (*A.status) return;
    Partial3 p{};
    for (int t = threadIdx.x; t < A.n; t += THREADS) merge3(p, A.partials[t]);
    block_reduce3(p, sh);
    if (threadIdx.x == 0) {
        double d[10];
        diag_row(sh[0], d);
        for (int c = 0; c < 10; ++c) A.last[c] = d[c];
        const double E = d[0];
        bool fault = false;
        if (A.rule) {
            const double t = (A.step + 1) * A.dt, window = 3.0 / A.frequency;
            if (t <= window) *A.source_peak = fmax(*A.source_peak, E);
            const double sp = *A.source_peak;
            const bool growth = t > window && sp > 0 && E > 1e4 * sp;
            fault = d[3] == 0 || growth;
        }
        if (fault) {
            *A.status = 1;
        } else {
            *A.peak = fmax(*A.peak, E);
            if (A.diag)
                for (int c = 0; c < 10; ++c) A.diag[10 * (size_t)A.row + c] = d[c];
            *A.accepted = A.row + 1;
        }
        verdict = fault;
    }
    __syncthreads();
    if (verdict || !A.samples) return;
    for (int r = threadIdx.x; r < A.nrec; r += THREADS) {
        double *o = A.samples + (size_t)A.row * 3 * A.nrec + 3 * r;
        for (int c = 0; c < 3; ++c) o[c] = A.state[A.probes[3 * r + c]];
    }
}

template <class R> __global__ void gather3_kernel(const R *state, const long long *probes, int n, double *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = state[probes[3 * r + c]];
}
// readDisplayFrame / slices (CudaEngine3D.cu:59-66): plane 0 XZ at y, 1 YZ at x, 2 XY at z
template <class R>
__global__ void slice3_kernel(const R *state, size_t count, DMesh3<R> m, int x, int y, int z, int comp, int plane,
                              int width, int height, float *out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= width * height) return;
    const int a = q % width, b = q / width;
    const int i = (plane == 1 ? x : a) + 3, j = (plane == 0 ? y : plane == 1 ? a : b) + 3, k = (plane == 2 ? z : b) + 3;
    const long long p = at3(m, i, j, k);
    const double vx = state[p], vy = state[count + p], vz = state[2 * count + p];
    const double value = comp == 0   ? vz
                         : comp == 1 ? vx
                         : comp == 2 ? vy
                         : comp == 4 ? -((double)state[3 * count + p] + state[4 * count + p] + state[5 * count + p]) / 3
                                     : sqrt(vx * vx + vy * vy + vz * vz);
    out[q] = (float)value;
}
template <class R> __global__ void plain3_kernel(const R *field, DMesh3<R> m, double *out) {
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)m.nx * m.ny * m.nz;
    if (q >= n) return;
    const int i = (int)(q % m.nx), j = (int)((q / m.nx) % m.ny), k = (int)(q / ((long long)m.nx * m.ny));
    out[q] = (double)field[at3(m, i, j, k)];
}

} // namespace
} // namespace swl

#include "engine3d_host.inl"
