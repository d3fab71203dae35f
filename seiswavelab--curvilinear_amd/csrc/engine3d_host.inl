// engine3d_host.inl — host side of the 3D HIP engine (included by engine3d.hip).
// Owns every device allocation, converts the reference's padded storage q = i + nx*(j + ny*k) into the
// aligned layout, checks the separable forms (metric band, sponge tables, materials by depth)
// against the full arrays, and enqueues the stage / inspection kernels on the engine's stream.
namespace swl {
namespace {

template <class R> struct Engine3D {
    static constexpr int real_bytes = sizeof(R);
    int device = 0, nrec = 0;
    int nx = 0, ny = 0, nz = 0, ld = 0, off = ORIGIN;
    size_t count = 0;
    hipStream_t stream{};
    hipStream_t side{};                // the surface band, concurrent with the deep bands of the same stage
    hipEvent_t fork{}, join{};
    DMesh3<R> m{};
    DeviceBuffer met, jac, lam, mu, rho, dax, day, daz, damp, src, grad[3], smet, slam, smu, snrm;
    DeviceBuffer state[3], accbuf, partials[2], level2, probes, status, accepted, peak, speak, last, diag, samples,
        gather, slice, plain;
    int ntx = 0, nty = 0, ntiles = 0, nLevel2 = 0, cur = 0, prev = 0, curPart = 0, prevPart = 0;
    int diagRows = 0;
    bool profiling = false;
    int ldsTiles = 0; // deep bands: LDS-staged stencil input (option "lds_tiles")
    int brick[3] = {0, 0, 0}; // tile order: bricks of x by y tiles by z planes per XCD run (0: tile columns)
    // deep bands with this many x-adjacent nodes per lane, as one vector (options "vector_terrain", "vector_flat":
    // 2, or 4 in FP32; 0 / 1: one node per lane)
    int vectorTerrain = 2, vectorFlat = 16 / (int)sizeof(R);
    int splitSurface = 0; // surface band: one launch per plane (option "split_surface"; 0: one launch for 4 planes)
    int overlapBands = 1; // surface band on the side stream beside the deep bands (option "overlap_bands")
    int zRowOk = 0;       // the terrain band's metrics 0..5 are (1, 0, 0, 0, 1, 0) (checked at construction)
    int zRowMetrics = 1;  // use that: read only metrics 6..8 in the terrain band (option "z_row_metrics")
    int zRowSurface = 1;  // and take metrics 0..5 as constants in the surface band (option "z_row_surface")
    int written[2] = {0, 0}; // partial slots the last step wrote into partials[0], partials[1]
    std::vector<hipEvent_t> events;
    size_t eventsUsed = 0;
    double stageMs = 0;
    long long stageLaunches = 0;

    long long idx(int i, int j, int k) const { return (long long)off + i + (long long)ld * j + (long long)ld * ny * k; }
    R *S(int b) const { return state[b].as<R>(); }

    template <class T> const T *put(DeviceBuffer &b, const std::vector<T> &v) {
        b.upload(v.data(), v.size() * sizeof(T));
        return b.as<T>();
    }
    // reference layout -> engine layout for planes k0..k1-1, `comps` interleaved values per node -> planar
    template <class T>
    const T *put_layout(DeviceBuffer &b, const double *host, int comps, int k1, size_t stride_elems) {
        std::vector<T> v((size_t)comps * stride_elems, T(0));
        for (int k = 0; k < k1; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t r = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    for (int c = 0; c < comps; ++c) v[(size_t)c * stride_elems + idx(i, j, k)] = (T)host[(size_t)comps * r + c];
                }
        return put(b, v);
    }

    void mark() {
        if (!profiling) return;
        if (eventsUsed == events.size()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "event");
            events.push_back(e);
        }
        check(hipEventRecord(events[eventsUsed++], stream), "event record");
    }
    void collect() {
        if (!profiling) return;
        check(hipStreamSynchronize(stream), "timing");
        for (size_t j = 0; j + 1 < eventsUsed; j += 2) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, events[j], events[j + 1]), "event elapsed");
            stageMs += ms, stageLaunches += 4;
        }
        eventsUsed = 0;
    }

    Engine3D(const swl_mesh3d_desc &d, const int64_t *rec, int n, int dev) : device(dev), nrec(n) {
        DeviceScope scope(device);
        nx = d.nx, ny = d.ny, nz = d.nz;
        require(nx >= 14 && ny >= 14 && nz >= 14, "3D mesh needs at least 8 physical nodes per axis");
        ld = (nx + 63) / 64 * 64;
        count = ((size_t)off + (size_t)ld * ny * nz + 63) / 64 * 64;
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "stream");
        check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "event");
        check(hipEventCreateWithFlags(&join, hipEventDisableTiming), "event");
        const size_t N = (size_t)nx * ny * nz;
        m.nx = nx, m.ny = ny, m.nz = nz, m.ld = ld, m.off = off, m.plane = (long long)ld * ny;
        m.h = (R)d.h;
        int e2 = 0;
        m.hinv = sizeof(R) == 4 || std::frexp(d.h, &e2) == 0.5 ? (R)(1 / d.h) : (R)0;
        // metric band: identity and J = 1 for k >= flat (checked)
        int flat = std::clamp(d.flat_from, 3, nz);
        for (int k = flat; k < nz - 3 && flat < nz; ++k)
            for (int j = 3; j < ny - 3; ++j)
                for (int i = 3; i < nx - 3; ++i) {
                    const size_t q = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    bool ok = d.jacobian[q] == 1.0;
                    for (int a = 0; a < 9; ++a) ok = ok && d.metrics[9 * q + a] == (a % 4 == 0 ? 1.0 : 0.0);
                    if (!ok) flat = nz, k = nz, j = ny, i = nx;
                }
        m.flat = flat;
        // z-row metric band: metrics 0..5 are (1, 0, 0, 0, 1, 0) exactly at every node of every plane with stored
        // metrics (the terrain band then reads only metrics 6..8, the surface band takes 0..5 as constants)
        zRowOk = 1;
        for (int k = 3; k < flat && k < nz - 3 && zRowOk; ++k)
            for (int j = 3; j < ny - 3 && zRowOk; ++j)
                for (int i = 3; i < nx - 3; ++i) {
                    const double *mt = d.metrics + 9 * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k));
                    if (mt[0] != 1.0 || mt[1] != 0.0 || mt[2] != 0.0 || mt[3] != 0.0 || mt[4] != 1.0 || mt[5] != 0.0) {
                        zRowOk = 0;
                        break;
                    }
                }
        m.mcount = ((long long)off + (long long)ld * ny * std::max(flat, 1) + 63) / 64 * 64;
        m.met = put_layout<R>(met, d.metrics, 9, flat, (size_t)m.mcount);
        m.jac = put_layout<R>(jac, d.jacobian, 1, flat, (size_t)m.mcount);
        // materials
        bool byDepth = d.material_by_depth != 0;
        if (byDepth)
            for (int k = 0; k < nz && byDepth; ++k) {
                const size_t q0 = (size_t)nx * ny * k;
                for (size_t q = q0; q < q0 + (size_t)nx * ny; ++q)
                    if (d.lambda[q] != d.lambda[q0] || d.mu[q] != d.mu[q0] || d.density[q] != d.density[q0]) {
                        byDepth = false;
                        break;
                    }
            }
        m.by_depth = byDepth;
        if (byDepth) {
            std::vector<R> l(nz), u(nz), r(nz);
            for (int k = 0; k < nz; ++k) {
                const size_t q0 = (size_t)nx * ny * k;
                l[k] = (R)d.lambda[q0], u[k] = (R)d.mu[q0], r[k] = (R)d.density[q0];
            }
            m.lam = put(lam, l), m.mu = put(mu, u), m.rho = put(rho, r);
        } else {
            m.lam = put_layout<R>(lam, d.lambda, 1, nz, count);
            m.mu = put_layout<R>(mu, d.mu, 1, nz, count);
            m.rho = put_layout<R>(rho, d.density, 1, nz, count);
        }
        // sponge: separable (X + Y) + Z tables when they reproduce every rate, else the full array
        bool separable = d.damping_axis[0] && d.damping_axis[1] && d.damping_axis[2];
        if (separable)
            for (int k = 3; k < nz - 3 && separable; ++k)
                for (int j = 3; j < ny - 3 && separable; ++j)
                    for (int i = 3; i < nx - 3; ++i) {
                        const double v = (d.damping_axis[0][i] + d.damping_axis[1][j]) + d.damping_axis[2][k];
                        if (std::memcmp(&v, &d.damping[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)], 8)) {
                            separable = false;
                            break;
                        }
                    }
        if (separable) {
            std::vector<R> x(nx), y(ny), z(nz);
            for (int i = 0; i < nx; ++i) x[i] = (R)d.damping_axis[0][i];
            for (int j = 0; j < ny; ++j) y[j] = (R)d.damping_axis[1][j];
            for (int k = 0; k < nz; ++k) z[k] = (R)d.damping_axis[2][k];
            if (sizeof(R) == 4) separable = false; // FP32: the rounded sum of rounded terms may differ; stream it
            else m.dax = put(dax, x), m.day = put(day, y), m.daz = put(daz, z), m.damp = nullptr;
        }
        if (!separable) m.damp = put_layout<R>(damp, d.damping, 1, nz, count);
        // source support box and arrays
        m.si1 = -1;
        int box[6] = {nx, -1, ny, -1, nz, -1};
        for (size_t q = 0; q < N; ++q) {
            bool any = d.source[q] != 0;
            for (int a = 0; a < 3; ++a) any = any || (d.source_gradient[a] && d.source_gradient[a][q] != 0);
            if (!any) continue;
            const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
            box[0] = std::min(box[0], i), box[1] = std::max(box[1], i), box[2] = std::min(box[2], j);
            box[3] = std::max(box[3], j), box[4] = std::min(box[4], k), box[5] = std::max(box[5], k);
        }
        if (box[1] >= 0) m.si0 = box[0], m.si1 = box[1], m.sj0 = box[2], m.sj1 = box[3], m.sk0 = box[4], m.sk1 = box[5];
        m.src = put_layout<R>(src, d.source, 1, nz, count);
        for (int a = 0; a < 3; ++a) m.grad[a] = d.source_gradient[a] ? put_layout<R>(grad[a], d.source_gradient[a], 1, nz, count) : nullptr;
        for (int a = 0; a < 3; ++a) m.dir[a] = (R)d.source_direction[a];
        m.rho_global = (R)d.density_global;
        // FP64 surface plane
        {
            const size_t ps = (size_t)ld * ny;
            std::vector<double> sm(9 * ps, 0.0), sl(ps, 0.0), su(ps, 0.0), sn(3 * ps, 0.0);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const size_t q = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * 3), s = (size_t)i + (size_t)ld * j;
                    for (int a = 0; a < 9; ++a) sm[9 * s + a] = d.metrics[9 * q + a];
                    for (int a = 0; a < 3; ++a) sn[3 * s + a] = d.normals[3 * q + a];
                    sl[s] = d.lambda[q], su[s] = d.mu[q];
                }
            m.smet = put(smet, sm), m.slam = put(slam, sl), m.smu = put(smu, su), m.snrm = put(snrm, sn);
        }
        for (auto &b : state) b.allocate((size_t)9 * count * sizeof(R));
        accbuf.allocate((size_t)9 * count * sizeof(R));
        ntx = (nx - 6 + TX<R> - 1) / TX<R>, nty = (ny - 6 + TY<R> - 1) / TY<R>;
        ntiles = ntx * nty * (nz - 6);
        // vector bands write one partial per (larger) tile: at most as many slots as one node per lane
        for (auto &b : partials) b.allocate((size_t)std::max(ntiles, partialSlots(1, 1)) * sizeof(Partial3));
        written[0] = written[1] = ntiles;
        nLevel2 = std::min(128, std::max(1, ntiles / 64));
        level2.allocate((size_t)nLevel2 * sizeof(Partial3));
        std::vector<long long> pi(3 * (size_t)std::max(n, 1));
        for (int r = 0; r < n; ++r) {
            require(rec[r] >= 0 && (size_t)rec[r] < N, "receiver index outside storage");
            const int i = (int)(rec[r] % nx), j = (int)((rec[r] / nx) % ny), k = (int)(rec[r] / ((int64_t)nx * ny));
            for (int c = 0; c < 3; ++c) pi[3 * r + c] = (long long)c * (long long)count + idx(i, j, k);
        }
        probes.upload(pi.data(), pi.size() * sizeof(long long));
        status.allocate(sizeof(int)), accepted.allocate(sizeof(int));
        peak.allocate(sizeof(double)), speak.allocate(sizeof(double)), last.allocate(10 * sizeof(double));
        gather.allocate((size_t)std::max(n, 1) * 3 * sizeof(double));
        const int pl = std::max({(nx - 6) * (nz - 6), (ny - 6) * (nz - 6), (nx - 6) * (ny - 6)});
        slice.allocate((size_t)pl * sizeof(float));
        check(hipStreamSynchronize(stream), "3D engine setup");
    }
    ~Engine3D() {
        for (auto e : events) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    static int direction(int step, int stage, int axis) { // FiniteDifference.hpp:9-13
        static const int flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
        return (flags[step % 8][axis] + stage) % 2;
    }
    // planes [3, 7) surface band, [7, flat) terrain band, [max(7, flat), nz - 3) flat band
    int flatFrom() const { return std::clamp(m.flat, std::min(7, nz - 3), nz - 3); }
    // VGPRs of the flat band's kernel as configured
    int flatRegisters() const {
        const void *k = vectorFlat == 4 ? reinterpret_cast<const void *>(&band_vec_kernel<R, 0, 0, 0, 16 / sizeof(R), FLAT>)
                        : vectorFlat == 2 ? reinterpret_cast<const void *>(&band_vec_kernel<R, 0, 0, 0, 2, FLAT>)
                                          : reinterpret_cast<const void *>(&stage3_kernel<R, 0, 0, 0, FLAT>);
        hipFuncAttributes attr{};
        return hipFuncGetAttributes(&attr, k) == hipSuccess ? attr.numRegs : 0;
    }
    // tiles per plane of a deep band: TX x TY nodes, or with nv nodes per lane 128-byte row segments by
    // THREADS / (128 / (nv * sizeof(R))) rows
    void bandTiles(int nv, int &tx, int &ty) const {
        if (nv <= 1) {
            tx = ntx, ty = nty;
            return;
        }
        const int lanes = 128 / (nv * (int)sizeof(R)), w = lanes * nv, h = THREADS / lanes;
        tx = (nx - 6 + w - 1) / w, ty = (ny - 6 + h - 1) / h;
    }
    // band b (0 surface, 1 terrain, 2 flat): planes [k0, k1), first partial slot t0, tiles per plane
    struct Band {
        int k0, k1, t0, tx, ty;
    };
    Band band(int b, int vterrain, int vflat) const {
        const int kEnd = nz - 3, kSurf = std::min(7, kEnd), kFlat = flatFrom();
        const int k[4] = {3, kSurf, kFlat, kEnd};
        const int vec[3] = {0, vterrain, vflat};
        Band r{k[b], k[b + 1], 0, 0, 0};
        for (int c = 0; c <= b; ++c) {
            int tx, ty;
            bandTiles(vec[c], tx, ty);
            if (c < b) r.t0 += tx * ty * (k[c + 1] - k[c]);
            else r.tx = tx, r.ty = ty;
        }
        return r;
    }
    // partial slots of a step: one per tile of the three bands
    int partialSlots(int vterrain, int vflat) const {
        const Band f = band(2, vterrain, vflat);
        return f.t0 + f.tx * f.ty * (f.k1 - f.k0);
    }
    // The bands of a stage write disjoint planes and partial slots and only read the stage input, so with
    // overlapBands the surface band (few, long-latency blocks) runs on the side stream beside the deep bands;
    // the next launch on the engine stream waits for both.
    template <int DX, int DY, int DZ, int BAND> void launch_vec(int nv, const Stage3<R> &a, hipStream_t s) {
        if (nv == 4) band_vec_kernel<R, DX, DY, DZ, 16 / sizeof(R), BAND><<<a.nband, THREADS, 0, s>>>(a);
        else band_vec_kernel<R, DX, DY, DZ, 2, BAND><<<a.nband, THREADS, 0, s>>>(a);
    }
    template <int DX, int DY, int DZ> void launch(Stage3<R> a) {
        bool forked = false;
        for (int b = 0; b < 3; ++b) {
            const Band d = band(b, vectorTerrain, vectorFlat);
            if (d.k1 <= d.k0) continue;
            a.t0 = d.t0, a.ntx = d.tx, a.nty = d.ty, a.nband = d.tx * d.ty * (d.k1 - d.k0), a.k0 = d.k0;
            hipStream_t stream = this->stream;
            if (b == 0 && overlapBands && d.k1 < nz - 3) {
                check(hipEventRecord(fork, stream), "3D band fork");
                check(hipStreamWaitEvent(side, fork, 0), "3D band fork");
                stream = side, forked = true;
            }
            if (b == 0 && splitSurface) {
                // one launch per surface plane, each compiled for its depth (fewer live registers than one body)
                const int per = d.tx * d.ty;
                Stage3<R> s = a;
                s.nband = per;
                for (int depth = 0; 3 + depth < d.k1; ++depth) {
                    s.k0 = 3 + depth, s.t0 = d.t0 + per * depth;
                    if (depth == 0) stage3_kernel<R, DX, DY, DZ, SURFACE, false, 0><<<per, THREADS, 0, stream>>>(s);
                    else if (depth == 1) stage3_kernel<R, DX, DY, DZ, SURFACE, false, 1><<<per, THREADS, 0, stream>>>(s);
                    else if (depth == 2) stage3_kernel<R, DX, DY, DZ, SURFACE, false, 2><<<per, THREADS, 0, stream>>>(s);
                    else stage3_kernel<R, DX, DY, DZ, SURFACE, false, 3><<<per, THREADS, 0, stream>>>(s);
                }
            } else if (b == 0 && zRowOk && zRowSurface) {
                stage3_kernel<R, DX, DY, DZ, SURFACE, false, -1, true><<<a.nband, THREADS, 0, stream>>>(a);
            } else if (b == 0) stage3_kernel<R, DX, DY, DZ, SURFACE><<<a.nband, THREADS, 0, stream>>>(a);
            else if (b == 1 && vectorTerrain && zRowOk && zRowMetrics) launch_vec<DX, DY, DZ, TERRAIN_Z>(vectorTerrain, a, stream);
            else if (b == 1 && vectorTerrain) launch_vec<DX, DY, DZ, TERRAIN>(vectorTerrain, a, stream);
            else if (b == 1 && a.lds) stage3_kernel<R, DX, DY, DZ, TERRAIN, true><<<a.nband, THREADS, 0, stream>>>(a);
            else if (b == 1) stage3_kernel<R, DX, DY, DZ, TERRAIN><<<a.nband, THREADS, 0, stream>>>(a);
            else if (vectorFlat == 4) band_vec_kernel<R, DX, DY, DZ, 16 / sizeof(R), FLAT><<<a.nband, THREADS, 0, stream>>>(a);
            else if (vectorFlat == 2) band_vec_kernel<R, DX, DY, DZ, 2, FLAT><<<a.nband, THREADS, 0, stream>>>(a);
            else if (a.lds) stage3_kernel<R, DX, DY, DZ, FLAT, true><<<a.nband, THREADS, 0, stream>>>(a);
            else stage3_kernel<R, DX, DY, DZ, FLAT><<<a.nband, THREADS, 0, stream>>>(a);
            if (b == 0 && forked) check(hipEventRecord(join, side), "3D band join");
        }
        if (forked) check(hipStreamWaitEvent(stream, join, 0), "3D band join");
    }
    void launch_stage(int dx, int dy, int dz, const Stage3<R> &a) {
        switch (dx * 4 + dy * 2 + dz) {
        case 0: launch<0, 0, 0>(a); break;
        case 1: launch<0, 0, 1>(a); break;
        case 2: launch<0, 1, 0>(a); break;
        case 3: launch<0, 1, 1>(a); break;
        case 4: launch<1, 0, 0>(a); break;
        case 5: launch<1, 0, 1>(a); break;
        case 6: launch<1, 1, 0>(a); break;
        default: launch<1, 1, 1>(a); break;
        }
        check(hipGetLastError(), "3D stage launch");
    }
    // one RK4 step from buffer c; the stage-4 partials go to partials[part]; returns the new buffer
    int enqueue_step(int c, double dt, double frequency, double scale, int step, int part) {
        const int b1 = (c + 1) % 3, b2 = (c + 2) % 3;
        const int in[4] = {c, b1, b2, b1}, out[4] = {b1, b2, b1, b2};
        const double times[4] = {step * dt, step * dt + 0.5 * dt, step * dt + 0.5 * dt, step * dt + dt};
        const double w[4] = {0.5 * dt, 0.5 * dt, dt, dt}, wy[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
        mark();
        for (int s = 0; s < 4; ++s) {
            Stage3<R> a{};
            a.m = m;
            a.in = S(in[s]), a.base = S(c), a.out = S(out[s]), a.acc = accbuf.as<R>();
            a.count = count;
            const double tau = 3.14159265358979323846 * frequency * (times[s] - 1.5 / frequency);
            a.forcing = (R)(scale * (1.0 - 2.0 * tau * tau) * std::exp(-tau * tau));
            a.w = (R)w[s], a.wy = (R)wy[s];
            a.kind = s == 0 ? 0 : s == 3 ? 2 : 1;
            a.ystage = s == 1;
            a.status = status.as<int>();
            a.partials = partials[part].as<Partial3>();
            a.ntx = ntx, a.nty = nty, a.ntiles = ntiles;
            a.lds = ldsTiles;
            a.bx = brick[0], a.by = brick[1], a.bz = brick[2];
            launch_stage(direction(step, s, 0), direction(step, s, 1), direction(step, s, 2), a);
        }
        written[part] = partialSlots(vectorTerrain, vectorFlat);
        mark();
        return b2;
    }
    Final3<R> final_args(int part, int row, int stepNumber, int rule, double dt, double frequency, double *diagp,
                         double *samplesp, int newState) {
        Final3<R> f{};
        f.partials = level2.as<Partial3>(), f.n = nLevel2;
        f.status = status.as<int>(), f.accepted = accepted.as<int>();
        f.peak = peak.as<double>(), f.source_peak = speak.as<double>();
        f.diag = diagp, f.samples = samplesp, f.last = last.as<double>();
        f.state = S(newState), f.probes = probes.as<long long>(), f.nrec = nrec;
        f.row = row, f.step = stepNumber, f.rule = rule, f.dt = dt, f.frequency = frequency;
        (void)part;
        return f;
    }
    void reduce(int part) {
        reduce3_kernel<<<nLevel2, THREADS, 0, stream>>>(partials[part].as<Partial3>(), written[part], level2.as<Partial3>(),
                                                         status.as<int>());
        check(hipGetLastError(), "3D reduce launch");
    }

    // CudaEngine3D::step: no fault rule (the caller inspects diagnostics and may roll back)
    void step(double time, double dt, double frequency, double scale, int number) {
        DeviceScope scope(device);
        (void)time;
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "3D status");
        const int part = 1 - curPart;
        const int n = enqueue_step(cur, dt, frequency, scale, number, part);
        reduce(part);
        final3_kernel<R><<<1, THREADS, 0, stream>>>(final_args(part, 0, number, 0, dt, frequency, nullptr, nullptr, n));
        check(hipGetLastError(), "3D finalize launch");
        check(hipStreamSynchronize(stream), "3D step");
        prev = cur, cur = n, prevPart = curPart, curPart = part;
        collect();
    }
    void rollback() { cur = prev, curPart = prevPart; }
    void reset() {
        DeviceScope scope(device);
        for (auto &b : state) check(hipMemsetAsync(b.ptr, 0, b.bytes, stream), "3D clear");
        check(hipMemsetAsync(accbuf.ptr, 0, accbuf.bytes, stream), "3D clear");
        for (auto &b : partials) check(hipMemsetAsync(b.ptr, 0, b.bytes, stream), "3D clear");
        check(hipStreamSynchronize(stream), "3D clear completion");
        cur = prev = 0;
    }
    // diagnostics of the current state (the partials written by the stage that produced it)
    void diagnostics(double *out) {
        DeviceScope scope(device);
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "3D status");
        reduce(curPart);
        final3_kernel<R><<<1, THREADS, 0, stream>>>(final_args(curPart, 0, 0, 0, 1, 1, nullptr, nullptr, cur));
        check(hipGetLastError(), "3D diagnostics launch");
        check(hipMemcpyAsync(out, last.ptr, 10 * sizeof(double), hipMemcpyDeviceToHost, stream), "3D diagnostics");
        check(hipStreamSynchronize(stream), "3D diagnostics");
    }
    void sample(double *out) {
        DeviceScope scope(device);
        if (!nrec) return;
        gather3_kernel<R><<<(nrec + 63) / 64, 64, 0, stream>>>(S(cur), probes.as<long long>(), nrec, gather.as<double>());
        check(hipGetLastError(), "3D receiver launch");
        check(hipMemcpyAsync(out, gather.ptr, (size_t)nrec * 3 * sizeof(double), hipMemcpyDeviceToHost, stream), "3D receivers");
        check(hipStreamSynchronize(stream), "3D receivers");
    }
    void slices(int x, int y, int z, int comp, float *out) {
        DeviceScope scope(device);
        const int px = nx - 6, py = ny - 6, pz = nz - 6;
        require(x >= 0 && x < px && y >= 0 && y < py && z >= 0 && z < pz && comp >= 0 && comp <= 4,
                "Display request is outside the active physical grid.");
        const int widths[3] = {px, py, px}, heights[3] = {pz, pz, py};
        size_t offset = 0;
        for (int p = 0; p < 3; ++p) {
            const int nq = widths[p] * heights[p];
            slice3_kernel<R><<<(nq + 255) / 256, 256, 0, stream>>>(S(cur), count, m, x, y, z, comp, p, widths[p], heights[p],
                                                                   slice.as<float>());
            check(hipGetLastError(), "3D slice launch");
            check(hipMemcpyAsync(out + offset, slice.ptr, (size_t)nq * sizeof(float), hipMemcpyDeviceToHost, stream),
                  "3D slice download");
            check(hipStreamSynchronize(stream), "3D slice");
            offset += (size_t)nq;
        }
    }
    void download(int f, double *out) {
        DeviceScope scope(device);
        require(f >= 0 && f < 9, "3D field index out of range");
        const size_t N = (size_t)nx * ny * nz;
        if (plain.bytes < N * sizeof(double)) plain.allocate(N * sizeof(double));
        plain3_kernel<R><<<(unsigned)((N + 255) / 256), 256, 0, stream>>>(S(cur) + (size_t)f * count, m, plain.as<double>());
        check(hipGetLastError(), "3D download launch");
        check(hipMemcpyAsync(out, plain.ptr, N * sizeof(double), hipMemcpyDeviceToHost, stream), "3D download");
        check(hipStreamSynchronize(stream), "3D download");
    }
    uint64_t resident() const {
        uint64_t b = 0;
        for (const DeviceBuffer *x : {&met, &jac, &lam, &mu, &rho, &dax, &day, &daz, &damp, &src, &smet, &slam, &smu, &snrm,
                                      &accbuf, &level2, &probes, &status, &accepted, &peak, &speak, &last, &diag,
                                      &samples, &gather, &slice, &plain})
            b += x->bytes;
        for (const auto &x : grad) b += x.bytes;
        for (const auto &x : state) b += x.bytes;
        for (const auto &x : partials) b += x.bytes;
        return b;
    }
    void run(int first, int count_steps, double dt, double frequency, double scale, double source_peak, double *diagOut,
             double *samplesOut, swl_run3d &res) {
        DeviceScope scope(device);
        if (count_steps <= 0) {
            res.accepted = 0, res.faulted = 0, res.source_peak = source_peak;
            return;
        }
        if (diagRows < count_steps) {
            diag.allocate((size_t)count_steps * 10 * sizeof(double));
            samples.allocate((size_t)count_steps * 3 * std::max(nrec, 1) * sizeof(double));
            diagRows = count_steps;
        }
        const double zero = 0;
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "3D run status");
        check(hipMemsetAsync(accepted.ptr, 0, sizeof(int), stream), "3D run counter");
        check(hipMemcpyAsync(speak.ptr, &source_peak, sizeof(double), hipMemcpyHostToDevice, stream), "3D run peak");
        check(hipMemcpyAsync(peak.ptr, &zero, sizeof(double), hipMemcpyHostToDevice, stream), "3D run peak");
        std::vector<int> states(count_steps + 1), parts(count_steps + 1);
        int c = cur, part = curPart;
        states[0] = c, parts[0] = part;
        for (int j = 0; j < count_steps; ++j) {
            const int np = 1 - part;
            const int n = enqueue_step(c, dt, frequency, scale, first + j, np);
            reduce(np);
            final3_kernel<R><<<1, THREADS, 0, stream>>>(final_args(np, j, first + j, 1, dt, frequency, diag.as<double>(),
                                                                  samples.as<double>(), n));
            check(hipGetLastError(), "3D finalize launch");
            c = n, part = np;
            states[j + 1] = c, parts[j + 1] = part;
        }
        int acc = 0, st = 0;
        double pk = 0, sp = 0;
        check(hipMemcpyAsync(&acc, accepted.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "3D run result");
        check(hipMemcpyAsync(&st, status.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "3D run result");
        check(hipMemcpyAsync(&pk, peak.ptr, sizeof(double), hipMemcpyDeviceToHost, stream), "3D run result");
        check(hipMemcpyAsync(&sp, speak.ptr, sizeof(double), hipMemcpyDeviceToHost, stream), "3D run result");
        check(hipStreamSynchronize(stream), "3D run");
        collect();
        if (acc > 0) {
            check(hipMemcpy(diagOut, diag.ptr, (size_t)acc * 10 * sizeof(double), hipMemcpyDeviceToHost), "3D run diagnostics");
            if (nrec)
                check(hipMemcpy(samplesOut, samples.ptr, (size_t)acc * 3 * nrec * sizeof(double), hipMemcpyDeviceToHost),
                      "3D run samples");
        }
        prev = states[acc > 0 ? acc - 1 : 0], prevPart = parts[acc > 0 ? acc - 1 : 0];
        cur = states[acc], curPart = parts[acc];
        res.accepted = acc, res.faulted = st != 0, res.peak_energy = pk, res.source_peak = sp, res.nonfinite = 0;
        if (st) {
            double lastRow[10];
            check(hipMemcpy(lastRow, last.ptr, sizeof lastRow, hipMemcpyDeviceToHost), "3D run verdict");
            res.nonfinite = lastRow[3] == 0;
        }
    }
};

} // namespace
} // namespace swl

struct swl_engine3d {
    int precision = 0;
    std::unique_ptr<swl::Engine3D<double>> e64;
    std::unique_ptr<swl::Engine3D<float>> e32;
    template <class F> void with(F &&f) {
        if (e64) f(*e64);
        else f(*e32);
    }
};

extern "C" {
int swl_engine3d_create(const swl_mesh3d_desc *d, const int64_t *rec, int n, int device, swl_engine3d **out) {
    return swl::guarded([&] {
        std::lock_guard<std::mutex> setup(swl::device_setup_mutex());
        swl::require(d && out && (n == 0 || rec), "invalid 3D engine arguments");
        auto e = std::make_unique<swl_engine3d>();
        e->precision = d->precision;
        if (d->precision == 1) e->e32 = std::make_unique<swl::Engine3D<float>>(*d, rec, n, device);
        else e->e64 = std::make_unique<swl::Engine3D<double>>(*d, rec, n, device);
        *out = e.release();
    });
}
void swl_engine3d_destroy(swl_engine3d *e) {
    std::lock_guard<std::mutex> setup(swl::device_setup_mutex());
    delete e;
}
int swl_engine3d_step(swl_engine3d *e, double time, double dt, double frequency, double scale, int step) {
    return swl::guarded([&] { e->with([&](auto &x) { x.step(time, dt, frequency, scale, step); }); });
}
int swl_engine3d_rollback(swl_engine3d *e) {
    return swl::guarded([&] { e->with([&](auto &x) { x.rollback(); }); });
}
int swl_engine3d_reset(swl_engine3d *e) {
    return swl::guarded([&] { e->with([&](auto &x) { x.reset(); }); });
}
int swl_engine3d_diagnostics(swl_engine3d *e, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.diagnostics(out); }); });
}
int swl_engine3d_sample(swl_engine3d *e, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.sample(out); }); });
}
int swl_engine3d_slices(swl_engine3d *e, int x, int y, int z, int component, float *out) {
    return swl::guarded([&] { e->with([&](auto &s) { s.slices(x, y, z, component, out); }); });
}
int swl_engine3d_download(swl_engine3d *e, int field, double *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.download(field, out); }); });
}
uint64_t swl_engine3d_resident_bytes(swl_engine3d *e) {
    uint64_t b = 0;
    e->with([&](auto &x) { b = x.resident(); });
    return b;
}
int swl_engine3d_run(swl_engine3d *e, int first_step, int count, double dt, double frequency, double scale,
                     double source_peak, double *diag, double *samples, swl_run3d *result) {
    return swl::guarded([&] {
        e->with([&](auto &x) { x.run(first_step, count, dt, frequency, scale, source_peak, diag, samples, *result); });
    });
}
int swl_engine3d_profile(swl_engine3d *e, int enable) {
    return swl::guarded([&] {
        e->with([&](auto &x) {
            x.profiling = enable != 0;
            x.stageMs = 0, x.stageLaunches = 0, x.eventsUsed = 0;
        });
    });
}
int swl_engine3d_set_option(swl_engine3d *e, const char *name, int value) {
    return swl::guarded([&] {
        const std::string key = name ? name : "";
        e->with([&](auto &x) {
            if (key == "lds_tiles") x.ldsTiles = value != 0;
            else if (key == "brick_x") x.brick[0] = std::max(0, value);
            else if (key == "brick_y") x.brick[1] = std::max(1, value);
            else if (key == "brick_z") x.brick[2] = std::max(1, value);
            else if (key == "split_surface") x.splitSurface = value != 0;
            else if (key == "overlap_bands") x.overlapBands = value != 0;
            else if (key == "z_row_metrics") x.zRowMetrics = value != 0;
            else if (key == "z_row_surface") x.zRowSurface = value != 0;
            else if (key == "vector_flat" || key == "vector_terrain") {
                using T = std::remove_reference_t<decltype(x)>;
                swl::require(value == 0 || value == 1 || value == 2 || (value == 4 && T::real_bytes == 4),
                             "vector_flat / vector_terrain take 0 or 1 (one node per lane), 2, or 4 (FP32)");
                (key == "vector_flat" ? x.vectorFlat : x.vectorTerrain) = value;
            }
            else throw std::runtime_error("unknown engine option: " + key);
        });
    });
}
int swl_engine3d_kernel_timing(swl_engine3d *e, double *out) {
    return swl::guarded([&] {
        e->with([&](auto &x) {
            out[0] = x.stageMs, out[1] = (double)x.stageLaunches;
            out[2] = (double)(x.nx - 6) * (double)(x.ny - 6) * (double)(x.nz - 6);
            out[3] = x.flatRegisters(), out[4] = x.m.flat, out[5] = x.m.by_depth;
            out[6] = x.zRowOk && x.zRowMetrics && x.vectorTerrain;
        });
    });
}
} // extern "C"
