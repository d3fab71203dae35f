// engine2d_host.inl — host side of the 2D HIP engine (included by engine2d.hip).
// Owns every device allocation, converts between the reference's padded storage (i + nx*k) and
// the engine's aligned layout (off + i + ld*k), classifies the grid into the interior rectangle and
// the general border, and enqueues the stage / inspection kernels on the engine's streams.

// Fixed device scratch of swl_engine2d_diagnostics_noalloc: the per-tile partials of stats_kernel
// (32 B each), so that entry point calls no allocator.  Each GPU holds its own copy until the process
// ends.  In host code this name is not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (swl::scratch_slot).
static __device__ __align__(256) unsigned char scratch[size_t(4) << 20];

namespace swl {
namespace {

constexpr size_t kScratchBytes = sizeof(scratch); // 131072 partials: grids of up to 131072 tiles of 64x4 nodes

// Device address of `scratch` on one GPU, looked up once per device with that device current, and the
// mutex that makes calls on that GPU take turns (they share its one copy of the array).
struct ScratchSlot {
    std::mutex use;
    void *ptr = nullptr;
};
ScratchSlot &scratch_slot(int device) {
    static std::mutex lookup;
    static std::vector<std::unique_ptr<ScratchSlot>> slots;
    std::lock_guard<std::mutex> guard(lookup);
    require(device >= 0, "invalid device");
    if ((size_t)device >= slots.size()) slots.resize((size_t)device + 1);
    if (!slots[device]) {
        auto slot = std::make_unique<ScratchSlot>();
        DeviceScope scope(device);
        check(hipGetSymbolAddress(&slot->ptr, HIP_SYMBOL(scratch)), "scratch address");
        require(slot->ptr != nullptr, "scratch address");
        slots[device] = std::move(slot);
    }
    return *slots[device];
}

struct PinnedBuffer { // page-locked host staging (hipHostMalloc; never registers existing memory)
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
    ~PinnedBuffer() {
        if (ptr) (void)hipHostFree(ptr);
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (ptr) (void)hipHostFree(ptr), ptr = nullptr;
        check(hipHostMalloc(&ptr, n, hipHostMallocDefault), "pinned staging allocation");
        bytes = n;
    }
};

inline uint64_t next_engine_uid() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}

template <class R> struct Engine2D {
    static constexpr int NV = Lane<R>::N;
    const uint64_t uid = next_engine_uid(); // identity for caches keyed on engines (addresses can be reused)
    int device = 0, nrec = 0, fields = 5, mech = 0, precision = 0;
    int nx = 0, nz = 0, ld = 0, off = ORIGIN;
    size_t count = 0; // field stride (elements)
    long long cpml = 0;
    hipStream_t stream{}, side{};
    DMesh<R> m{};
    DeviceBuffer mesh_arrays[32];
    // NB rotating state buffers (and PML memories): a step from buffer c writes c + 1 and c + 2 (mod NB); with four,
    // the next step's first stage writes neither c + 2 nor c, so that stage can end the step before it (StepEnd)
    static constexpr int NB = 4;
    DeviceBuffer state[NB], accbuf, pstate[NB], paccbuf, partials, level2, probes, status, accepted, peak, diag, samples,
        gather, display, borderNodes, pmlLanes;
    int ntx = 0, ntiles = 0;      // 64x4 tiles over the computational grid
    int nBorder = 0, nGeneral = 0; // nodes outside the interior rectangle (and the PML lanes), blocks of 256 of them
    int nPml = 0, nPmlBlocks = 0;  // FP32 elastic: CFS-PML vector lanes below the surface band, blocks of 256 of them
    int gBlocks = 0;               // border-class blocks of a stage launch (general, then PML), a multiple of 8
    // PML vector lanes (option "pml_vector"): -1 automatic, 0 off, 1 on; the lists are rebuilt when the choice changes.
    // Automatic: grids of at least kPmlTiles interior tiles, where they measured 2-5 % faster per step (fewer,
    // longer border threads lengthen the tail of a launch that fits the device in one round, as small grids' do)
    int pmlVector = -1;
    static constexpr int kPmlTiles = 1024;
    // small-grid FP32 elastic kernel (option "latency_kernel"): -1 automatic (launches of fewer than kLatBlocks
    // blocks, a few waves per SIMD at most), 0 off, 1 on; same arithmetic, same results
    int latencyKernel = -1;
    static constexpr int kLatBlocks = 1536;
    bool latency_kernel() const {
        if (!kFast<R> || mech != 0) return false;
        return latencyKernel < 0 ? gBlocks + nInterior < kLatBlocks : latencyKernel != 0;
    }
    bool tile_lds() const {
        if (mech != 0) return false;
        return tileLds < 0 ? gBlocks + nInterior >= kLatBlocks && !latency_kernel() : tileLds != 0;
    }
    bool pml_vector() const { return pmlVector < 0 ? nInterior >= kPmlTiles : pmlVector != 0; }
    int rtx = 0, nInterior = 0;   // interior tiles (64*NV x 4 nodes)
    int nPartials = 0, nLevel2 = 0;
    static constexpr int kDirectPartials = 2048;
    int endInStage = 1; // device-resident runs end each step in the next step's first stage launch
    // border blocks as their own launch on the side stream, beside the interior tiles: -1 automatic (FP64 grids of
    // at least kSplitTiles interior tiles, where it was measured faster; FP32 and small grids are slower split)
    int splitBorder = -1;
    static constexpr int kSplitTiles = 8192;
    bool split_border() const { return splitBorder < 0 ? !kFast<R> && nInterior >= kSplitTiles : splitBorder != 0; }
    int spreadBorder = 0; // border blocks spread evenly over the whole-grid launch (measured: no gain)
    hipEvent_t fork[4]{}, join[4]{};
    int cur = 0, prev = 0, diag_rows = 0, registers = 0, nt = 1, metricOk = 0;
    // elastic interior: stencil input staged through LDS tiles, loaded straight into LDS one field ahead (option
    // "lds_tiles": -1 automatic, 0 off, 1 on).  Automatic: elastic grids of at least kLatBlocks blocks per launch
    // (smaller grids take the small-grid kernel), where the tiles measured 2-3 % faster per step than direct loads.
    int tileLds = -1;
    size_t local_bytes = 0;
    std::vector<Partial> host_partials;
    // optional per-launch timing with HIP events (interior stage kernel and step finalisation)
    bool profiling = false;
    std::vector<hipEvent_t> events;
    std::vector<int> eventKind;
    size_t eventsUsed = 0;
    double stageMs = 0, finalMs = 0;
    long long stageLaunches = 0, finalLaunches = 0;

    R *S(int b) const { return state[b].as<R>(); }
    R *P(int b) const { return pstate[b].as<R>(); }
    size_t idx(int i, int k) const { return (size_t)off + i + (size_t)ld * k; }
    size_t internal_of(int64_t ref) const { return idx((int)(ref % nx), (int)(ref / nx)); }

    // reference layout (i + nx*k, `comps` values per node) -> engine layout, converted to R; written straight
    // into page-locked staging and uploaded on the set-up stream (completed by setup_finish at construction end)
    template <class T> const T *put_layout(int slot, const double *host, int comps) {
        if (!host) return nullptr;
        const size_t n = (size_t)comps * count;
        T *v = static_cast<T *>(pinned_arena().take(n * sizeof(T)));
        std::fill(v, v + n, T(0));
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i)
                for (int c = 0; c < comps; ++c)
                    v[(size_t)comps * idx(i, k) + c] = (T)host[(size_t)comps * ((size_t)i + (size_t)nx * k) + c];
        mesh_arrays[slot].upload_staged(v, n * sizeof(T));
        return mesh_arrays[slot].as<T>();
    }
    template <class T> const T *put_plain(int slot, const double *host, size_t n) {
        if (!host) return nullptr;
        T *v = static_cast<T *>(pinned_arena().take(std::max<size_t>(n, 1) * sizeof(T)));
        for (size_t i = 0; i < n; ++i) v[i] = (T)host[i];
        mesh_arrays[slot].upload_staged(v, n * sizeof(T));
        return mesh_arrays[slot].as<T>();
    }

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
            if (eventKind[j / 2] == 0) stageMs += ms, stageLaunches += 4; // one pair spans a step's 4 stages
            else finalMs += ms, ++finalLaunches;
        }
        eventsUsed = 0;
        eventKind.clear();
    }

    Engine2D(const swl_mesh2d_desc &d, const int64_t *rec, int n, int dev) : device(dev), nrec(n) {
        DeviceScope scope(device);
        precision = d.precision;
        nx = d.nx, nz = d.nz;
        fields = d.field_count;
        mech = d.mechanisms;
        cpml = d.cpml_count;
        require(d.nx >= 14 && d.nz >= 14, "2D mesh needs at least 8 computational nodes per axis");
        require(mech >= 0 && mech <= 3 && fields == (mech ? 6 + 4 * mech : 5), "inconsistent GSLS field count");
        ld = (nx + off + 63) / 64 * 64;
        count = ((size_t)off + (size_t)ld * nz + 63) / 64 * 64;
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "stream");
        for (int e = 0; e < 4; ++e) {
            check(hipEventCreateWithFlags(&fork[e], hipEventDisableTiming), "event");
            check(hipEventCreateWithFlags(&join[e], hipEventDisableTiming), "event");
        }
        m.nx = nx, m.nz = nz, m.ld = ld, m.off = off, m.wx = d.cpml_width_x, m.wz = d.cpml_width_z, m.cpml = cpml;
        m.fields = fields, m.mech = mech;
        m.h = (R)d.h;
        {
            int e = 0;
            const double frac = std::frexp(d.h, &e); // h = frac * 2^e, power of two iff frac == 0.5
            // FP32 (held to a tolerance) always multiplies by 1/h; FP64 only where that is exact
            m.hinv = frac == 0.5 || kFast<R> ? (R)(1 / d.h) : (R)0;
        }
        for (int l = 0; l < 3; ++l) m.rates[l] = (R)d.rates[l];
        m.dir0 = (R)d.source_direction[0], m.dir2 = (R)d.source_direction[2];
        m.a = put_layout<R>(0, d.a, 1), m.b = put_layout<R>(1, d.b, 1), m.jac = put_layout<R>(2, d.jac, 1);
        m.lam = put_layout<R>(3, d.lambda, 1), m.mu = put_layout<R>(4, d.mu, 1), m.rho = put_layout<R>(5, d.rho, 1);
        m.src = put_layout<R>(6, d.source, 1), m.nrm = put_layout<R>(8, d.normals, 2);
        m.has_damp = 0;
        if (d.damping)
            for (size_t q = 0; q < (size_t)nx * nz && !m.has_damp; ++q) m.has_damp = d.damping[q] != 0;
        m.damp = m.has_damp ? put_layout<R>(7, d.damping, 1) : nullptr;
        m.has_grad = d.source_gradient[0] != nullptr;
        if (m.has_grad) m.grad0 = put_layout<R>(9, d.source_gradient[0], 1), m.grad1 = put_layout<R>(10, d.source_gradient[1], 1);
        if (cpml) {
            m.cx = put_plain<R>(11, d.cpml_x, 3 * (size_t)(nx - 6));
            m.cz = put_plain<R>(12, d.cpml_z, 3 * (size_t)(nz - 6));
            if (kFast<R>) { // 1/kappa, d/kappa + alpha, d/kappa^2 per position, evaluated in FP64
                for (int axis = 0; axis < 2; ++axis) {
                    const double *pr = axis ? d.cpml_z : d.cpml_x;
                    const size_t n = (size_t)(axis ? nz : nx) - 6;
                    std::vector<double> f(3 * n);
                    for (size_t j = 0; j < n; ++j) {
                        const double dd = pr[3 * j], kappa = pr[3 * j + 1], alpha = pr[3 * j + 2];
                        f[3 * j] = 1 / kappa, f[3 * j + 1] = dd / kappa + alpha, f[3 * j + 2] = dd / (kappa * kappa);
                    }
                    (axis ? m.czf : m.cxf) = put_plain<R>(axis ? 29 : 28, f.data(), f.size());
                }
            }
        }
        for (int l = 0; l < mech; ++l)
            m.dP[l] = put_layout<R>(13 + l, d.delta_p[l], 1), m.dS[l] = put_layout<R>(16 + l, d.delta_s[l], 1);
        // FP64 surface row (k = 3)
        std::vector<double> sa(nx), sb(nx), sl(nx), sm(nx), sn(2 * (size_t)nx);
        for (int i = 0; i < nx; ++i) {
            const size_t q = (size_t)i + (size_t)nx * 3;
            sa[i] = d.a[q], sb[i] = d.b[q], sl[i] = d.lambda[q], sm[i] = d.mu[q];
            sn[2 * i] = d.normals[2 * q], sn[2 * i + 1] = d.normals[2 * q + 1];
        }
        m.sa = put_plain<double>(19, sa.data(), sa.size()), m.sb = put_plain<double>(20, sb.data(), sb.size());
        m.slam = put_plain<double>(21, sl.data(), sl.size()), m.smu = put_plain<double>(22, sm.data(), sm.size());
        m.snrm = put_plain<double>(23, sn.data(), sn.size());
        setup_metric(d);
        setup_rows(d);
        setup_regions(d);
        for (auto &b : state) b.allocate_async((size_t)fields * count * sizeof(R));
        accbuf.allocate_async((size_t)fields * count * sizeof(R));
        if (cpml) {
            for (auto &b : pstate) b.allocate_async((size_t)cpml * sizeof(R));
            paccbuf.allocate_async((size_t)cpml * sizeof(R));
        }
        host_partials.resize(ntiles);
        std::vector<long long> pi(2 * (size_t)std::max(n, 1));
        for (int r = 0; r < n; ++r) {
            require(rec[r] >= 0 && rec[r] < (int64_t)nx * nz, "receiver index outside storage");
            pi[2 * r] = (long long)internal_of(rec[r]), pi[2 * r + 1] = (long long)count + pi[2 * r];
        }
        probes.upload_async(pi.data(), pi.size() * sizeof(long long));
        status.allocate_async(sizeof(int));
        accepted.allocate_async(sizeof(int));
        peak.allocate_async(sizeof(double));
        gather.allocate_async((size_t)std::max(n, 1) * 5 * sizeof(double));
        display.allocate_async((size_t)(nx - 6) * (nz - 6) * sizeof(float));
        // one wait for every upload and clear above, on this thread's set-up stream only: a session built
        // while other sessions run does not queue behind their work
        setup_finish();
        hipFuncAttributes attr{};
        if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&stage_kernel<R, 0, 0, 0>)) == hipSuccess) {
            registers = attr.numRegs;
            local_bytes = attr.localSizeBytes;
        }
    }
    ~Engine2D() {
        if (stream) (void)hipStreamSynchronize(stream); // a pipelined run may still be in flight
        for (auto &P : pending)
            if (P.done) (void)hipEventDestroy(P.done);
        for (auto e : events) (void)hipEventDestroy(e);
        for (int e = 0; e < 4; ++e) {
            if (fork[e]) (void)hipEventDestroy(fork[e]);
            if (join[e]) (void)hipEventDestroy(join[e]);
        }
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Separable terrain metrics (CurvilinearSolver2D.cpp:349-354): a = -(H' F)/(1 + H F'),
    // b = 1/(1 + H F'), J = 1 + H F'.  Used on the device only if they reproduce every stored value.
    void setup_metric(const swl_mesh2d_desc &d) {
        m.metric = 0;
        if (!d.surface_depth || !d.surface_slope || !d.fade || !d.fade_derivative) return;
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                const size_t q = (size_t)i + (size_t)nx * k;
                const double zx = d.surface_slope[i] * d.fade[k], j = 1 + d.surface_depth[i] * d.fade_derivative[k];
                const double a = -zx / j, b = 1 / j;
                if (std::memcmp(&a, &d.a[q], 8) || std::memcmp(&b, &d.b[q], 8) || std::memcmp(&j, &d.jac[q], 8)) return;
            }
        m.tz_ = put_plain<double>(24, d.surface_depth, nx);
        m.tslope = put_plain<double>(25, d.surface_slope, nx);
        m.fade = put_plain<double>(26, d.fade, nz);
        m.fadeder = put_plain<double>(27, d.fade_derivative, nz);
        m.metric = metricOk = 1;
    }

    // Per-row material tables when every material array is constant along each row (layered models,
    // grid-conformal by k); the interior then reads one value per row instead of streaming the arrays.
    DeviceBuffer rowTables;
    int rowsOk = 0; // the tables exist (the per-node arrays are always resident as well)
    void setup_rows(const swl_mesh2d_desc &d) {
        m.by_row = rowsOk = 0;
        std::vector<const double *> arrays = {d.lambda, d.mu, d.rho};
        for (int l = 0; l < mech; ++l) arrays.push_back(d.delta_p[l]), arrays.push_back(d.delta_s[l]);
        for (const double *a : arrays)
            for (int k = 0; k < nz; ++k)
                for (int i = 1; i < nx; ++i)
                    if (std::memcmp(&a[(size_t)i + (size_t)nx * k], &a[(size_t)nx * k], 8)) return;
        std::vector<R> table(arrays.size() * (size_t)nz);
        for (size_t a = 0; a < arrays.size(); ++a)
            for (int k = 0; k < nz; ++k) table[a * nz + k] = (R)arrays[a][(size_t)nx * k];
        rowTables.upload_async(table.data(), table.size() * sizeof(R));
        const R *base = rowTables.as<R>();
        m.lam_k = base, m.mu_k = base + nz, m.rho_k = base + 2 * nz;
        for (int l = 0; l < mech; ++l) m.dP_k[l] = base + (3 + 2 * l) * (size_t)nz, m.dS_k[l] = base + (4 + 2 * l) * (size_t)nz;
        m.by_row = rowsOk = 1;
    }

    // Interior rectangle: depth >= 4, outside every CFS-PML slot, zero sponge damping; x-extent
    // aligned to the lane width.  The source support box stays inside either kernel.
    void setup_regions(const swl_mesh2d_desc &d) {
        const int nxc = nx - 6, nzc = nz - 6;
        int i0 = 3, i1 = nx - 3, k0 = 7, k1 = nz - 3;
        if (cpml) {
            i0 = std::max(i0, 3 + d.cpml_width_x), i1 = std::min(i1, nx - 3 - d.cpml_width_x);
            k1 = std::min(k1, nz - 3 - d.cpml_width_z);
        }
        if (m.has_damp) {
            // columns damped on the top physical row and rows damped in an undamped column
            int left = 3, right = nx - 3;
            while (left < nx - 3 && d.damping[(size_t)left + (size_t)nx * 3] != 0) ++left;
            while (right > left && d.damping[(size_t)(right - 1) + (size_t)nx * 3] != 0) --right;
            int bottom = nz - 3;
            const int probe = (left + right) / 2;
            while (bottom > k0 && d.damping[(size_t)probe + (size_t)nx * (bottom - 1)] != 0) --bottom;
            i0 = std::max(i0, left), i1 = std::min(i1, right), k1 = std::min(k1, bottom);
        }
        i0 = 3 + (i0 - 3 + NV - 1) / NV * NV;
        i1 = 3 + (std::max(i1, i0) - 3) / NV * NV;
        if (i1 <= i0 || k1 <= k0) i0 = i1 = k0 = k1 = 0;
        // verify: no damping and no PML slot inside the rectangle
        for (int k = k0; k < k1; ++k)
            for (int i = i0; i < i1; ++i) {
                bool bad = m.has_damp && d.damping[(size_t)i + (size_t)nx * k] != 0;
                if (cpml) {
                    const int ii = i - 3, kk = k - 3, wx = d.cpml_width_x, wz = d.cpml_width_z;
                    bad = bad || ii < wx || ii >= nxc - wx || kk >= nzc - wz;
                }
                if (bad) {
                    i0 = i1 = k0 = k1 = 0;
                    k = k1;
                    break;
                }
            }
        m.ri0 = i0, m.ri1 = i1, m.rk0 = k0, m.rk1 = k1;
        // source support box
        m.si0 = nx, m.si1 = -1, m.sk0 = nz, m.sk1 = -1;
        const double *srcs[3] = {d.source, d.source_gradient[0], d.source_gradient[1]};
        for (const double *sarr : srcs) {
            if (!sarr) continue;
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; ++i)
                    if (sarr[(size_t)i + (size_t)nx * k] != 0)
                        m.si0 = std::min(m.si0, i), m.si1 = std::max(m.si1, i), m.sk0 = std::min(m.sk0, k),
                        m.sk1 = std::max(m.sk1, k);
        }
        ntx = (nxc + TX - 1) / TX;
        ntiles = ntx * ((nzc + TZ - 1) / TZ);
        rtx = i1 > i0 ? (i1 - i0 + 64 * NV - 1) / (64 * NV) : 0;
        nInterior = rtx ? rtx * ((k1 - k0 + 3) / 4) : 0;
        build_border_lists();
    }
    // The nodes outside the interior rectangle: FP32 elastic CFS-PML models take the ones below the surface band
    // (depth >= 4) as vector lanes of PML_NV x-adjacent nodes (pml_body; the rectangle's x bounds are multiples of
    // NV, hence of PML_NV, from column 3, so a lane is wholly inside or outside it), the rest
    // go to the general list, compacted: visited tile by tile over 64x4 tiles, row by row inside each, so
    // consecutive lanes read consecutive nodes of a row.
    void build_border_lists() {
        const int i0 = m.ri0, i1 = m.ri1, k0 = m.rk0, k1 = m.rk1;
        const auto inRect = [&](int i, int k) { return i >= i0 && i < i1 && k >= k0 && k < k1; };
        std::vector<int2> lanes;
        std::vector<unsigned char> covered((size_t)nx * nz, 0);
        if (kFast<R> && mech == 0 && cpml && !m.has_damp && pml_vector())
            for (int k = 7; k < nz - 3; ++k)
                for (int iv = 3; iv < nx - 3; iv += PML_NV) {
                    if (inRect(iv, k)) continue;
                    lanes.push_back(int2{iv, k});
                    for (int j = 0; j < PML_NV && iv + j < nx - 3; ++j) covered[(size_t)(iv + j) + (size_t)nx * k] = 1;
                }
        std::vector<int2> border;
        for (int t = 0; t < ntiles; ++t) {
            const int ti0 = 3 + (t % ntx) * TX, ti1 = std::min(nx - 3, ti0 + TX), tk0 = 3 + (t / ntx) * TZ,
                      tk1 = std::min(nz - 3, tk0 + TZ);
            if (ti0 >= i0 && ti1 <= i1 && tk0 >= k0 && tk1 <= k1) continue;
            for (int k = tk0; k < tk1; ++k)
                for (int i = ti0; i < ti1; ++i)
                    if (!inRect(i, k) && !covered[(size_t)i + (size_t)nx * k]) border.push_back(int2{i, k});
        }
        nBorder = (int)border.size();
        nGeneral = (nBorder + THREADS - 1) / THREADS;
        borderNodes.upload_async(border.empty() ? nullptr : border.data(), std::max<size_t>(border.size(), 1) * sizeof(int2));
        nPml = (int)lanes.size();
        nPmlBlocks = (nPml + THREADS - 1) / THREADS;
        pmlLanes.upload_async(lanes.empty() ? nullptr : lanes.data(), std::max<size_t>(lanes.size(), 1) * sizeof(int2));
        gBlocks = ((nGeneral + 7) / 8 * 8 + nPmlBlocks + 7) / 8 * 8;
        nPartials = nGeneral + nPmlBlocks + nInterior;
        partials.allocate_async((size_t)std::max(nPartials, 1) * sizeof(Partial));
        nLevel2 = std::max(1, std::min(nPartials, 128));
        level2.allocate_async((size_t)nLevel2 * sizeof(Partial));
    }
    void set_pml_vector(int v) {
        const bool before = pml_vector();
        const int old = pmlVector;
        pmlVector = v < 0 ? -1 : v != 0;
        if (pml_vector() == before) return;
        if (pending[0].active || pending[1].active) {
            pmlVector = old;
            require(false, "pml_vector: a pipelined run is in flight");
        }
        DeviceScope scope(device);
        check(hipStreamSynchronize(stream), "2D border lists"); // the lists and partials are replaced
        build_border_lists();
        setup_finish();
    }

    // slot = RK stage 0..3; with profiling on, one event pair brackets the four back-to-back stage launches
    // (an event between launches would itself add a dispatch gap)
    template <int DX, int DZ, int PART> void launch_part(const StageArgs<R> &a, int blocks, hipStream_t s) {
        switch (mech) {
        case 0:
            // the interior-only launch takes the LDS tiles where the one-launch layout would
            if (PART == 2 && tile_lds()) stage_kernel<R, DX, DZ, 0, true, PART><<<blocks, THREADS, 0, s>>>(a);
            else stage_kernel<R, DX, DZ, 0, false, PART><<<blocks, THREADS, 0, s>>>(a);
            break;
        case 1: stage_kernel<R, DX, DZ, 1, false, PART><<<blocks, THREADS, 0, s>>>(a); break;
        case 2: stage_kernel<R, DX, DZ, 2, false, PART><<<blocks, THREADS, 0, s>>>(a); break;
        default: stage_kernel<R, DX, DZ, 3, false, PART><<<blocks, THREADS, 0, s>>>(a); break;
        }
    }
    template <int DX, int DZ> void launch_both(const StageArgs<R> &a, int slot) {
        const int blocks = a.gblocks + nInterior;
        if (slot == 0) mark(0, true);
        if (split_border() && a.gblocks > 0 && nInterior > 0) {
            // border blocks on the side stream beside the interior tiles; the next launch waits for both
            check(hipEventRecord(fork[slot], stream), "2D fork");
            check(hipStreamWaitEvent(side, fork[slot], 0), "2D fork");
            launch_part<DX, DZ, 1>(a, a.gblocks, side);
            check(hipGetLastError(), "2D border launch");
            check(hipEventRecord(join[slot], side), "2D join");
            launch_part<DX, DZ, 2>(a, nInterior, stream);
            check(hipGetLastError(), "2D interior launch");
            check(hipStreamWaitEvent(stream, join[slot], 0), "2D join");
            if (slot == 3) mark(0, false);
            return;
        }
        switch (mech) {
        case 0:
            if (tile_lds()) stage_kernel<R, DX, DZ, 0, true><<<blocks, THREADS, 0, stream>>>(a);
            else if (kFast<R> && latency_kernel()) {
                if constexpr (kFast<R>) stage_kernel<R, DX, DZ, 0, false, 0, true><<<blocks, THREADS, 0, stream>>>(a);
            } else stage_kernel<R, DX, DZ, 0><<<blocks, THREADS, 0, stream>>>(a);
            break;
        case 1: stage_kernel<R, DX, DZ, 1><<<blocks, THREADS, 0, stream>>>(a); break;
        case 2: stage_kernel<R, DX, DZ, 2><<<blocks, THREADS, 0, stream>>>(a); break;
        default: stage_kernel<R, DX, DZ, 3><<<blocks, THREADS, 0, stream>>>(a); break;
        }
        check(hipGetLastError(), "2D stage launch");
        if (slot == 3) mark(0, false);
    }
    void launch_stage(int dx, int dz, const StageArgs<R> &a, int slot) {
        if (dx == 0 && dz == 0) launch_both<0, 0>(a, slot);
        else if (dx == 0) launch_both<0, 1>(a, slot);
        else if (dz == 0) launch_both<1, 0>(a, slot);
        else launch_both<1, 1>(a, slot);
    }

    static int direction(int step, int stage, int axis) { // FiniteDifference.hpp:9-13
        static const int flags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                        {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
        return (flags[step % 8][axis] + stage) % 2;
    }
    static double pulse(double t, double f) { // CurvilinearSolver2D.cpp:31-34
        const double a = 3.14159265358979323846 * f * (t - 1.5 / f);
        return (1 - 2 * a * a) * std::exp(-a * a);
    }

    // Arguments of RK stage s (0..3) of step `step` taken from buffer c (stages read c, b1, b2, b1 and write
    // b1, b2, b1, b2; the step ends in b2 = (c + 2) % NB).
    StageArgs<R> stage_args(int c, double dt, double frequency, double scale, int step, int s) const {
        const int b1 = (c + 1) % NB, b2 = (c + 2) % NB;
        const int in[4] = {c, b1, b2, b1}, out[4] = {b1, b2, b1, b2};
        const double times[4] = {step * dt, (step + .5) * dt, (step + .5) * dt, (step + 1) * dt};
        const R weights[4] = {(R)(dt / 2), (R)(dt / 2), (R)dt, (R)dt};
        StageArgs<R> a{};
        a.m = m;
        a.in = S(in[s]), a.base = S(c), a.out = S(out[s]), a.acc = accbuf.as<R>();
        if (cpml) a.pin = P(in[s]), a.pbase = P(c), a.pout = P(out[s]), a.pacc = paccbuf.as<R>();
        a.count = count;
        a.forcing = (R)(scale * pulse(times[s], frequency));
        a.w = weights[s];
        a.wy = (R)(s == 0 || s == 3 ? dt / 6 : dt / 3);
        a.hdt = (R)(dt / 2), a.dt = (R)dt, a.dt6 = (R)(dt / 6), a.stage = s;
        a.kind = s == 0 ? 0 : s == 3 ? 2 : 1;
        a.ystage = s == 1;
        a.status = status.as<int>();
        a.partials = partials.as<Partial>();
        a.border = borderNodes.as<int2>(), a.nborder = nBorder, a.ngeneral = nGeneral, a.gblocks = gBlocks;
        a.pml = pmlLanes.as<int2>(), a.npml = nPml, a.pml0 = (nGeneral + 7) / 8 * 8, a.npmlb = nPmlBlocks;
        a.ipart0 = nGeneral + nPmlBlocks;
        a.ninterior = nInterior, a.rtx = rtx, a.nt = nt, a.spread = spreadBorder;
        return a;
    }
    // Enqueue one RK4 step from buffer c; returns the buffer holding the new state.
    int enqueue_step(int c, double dt, double frequency, double scale, int step) {
        for (int s = 0; s < 4; ++s)
            launch_stage(direction(step, s, 0), direction(step, s, 2), stage_args(c, dt, frequency, scale, step, s), s);
        return (c + 2) % NB;
    }

    void step(double, double dt, double frequency, double scale, int number) {
        DeviceScope scope(device);
        check(hipMemsetAsync(status.ptr, 0, sizeof(int), stream), "2D status clear");
        prev = cur;
        cur = enqueue_step(cur, dt, frequency, scale, number);
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
    void diagnostics(swl_diag2d &out) {
        DeviceScope scope(device);
        DeviceBuffer tmp;
        tmp.allocate((size_t)ntiles * sizeof(Partial));
        diagnostics_with(tmp.as<Partial>(), out);
    }
    // The same with the partials in the GPU's fixed scratch array: no allocation.  Every tile's block writes
    // its partial, so nothing needs clearing, and only this engine's ntiles partials are read back.
    void diagnostics_noalloc(swl_diag2d &out) {
        DeviceScope scope(device);
        const size_t need = (size_t)ntiles * sizeof(Partial);
        if (need > kScratchBytes)
            throw std::runtime_error("no-allocation diagnostics: this 2D grid has " + std::to_string(ntiles) +
                                     " tiles, whose partials need " + std::to_string(need) + " bytes; the fixed scratch holds " +
                                     std::to_string(kScratchBytes) + " bytes (" + std::to_string(kScratchBytes / sizeof(Partial)) +
                                     " tiles of 64x4 nodes)");
        ScratchSlot &slot = scratch_slot(device);
        std::lock_guard<std::mutex> use(slot.use);
        diagnostics_with(static_cast<Partial *>(slot.ptr), out);
    }
    // stats_kernel over the accepted state into `partials` (at least ntiles entries), reduced on the host
    void diagnostics_with(Partial *partials, swl_diag2d &out) {
        stats_kernel<R><<<ntiles, THREADS, 0, stream>>>(m, S(cur), cpml ? P(cur) : nullptr, count, partials, ntx,
                                                          ntiles);
        check(hipGetLastError(), "2D diagnostics");
        check(hipMemcpyAsync(host_partials.data(), partials, ntiles * sizeof(Partial), hipMemcpyDeviceToHost, stream),
              "2D diagnostics readback");
        check(hipStreamSynchronize(stream), "2D diagnostics");
        out.energy = out.max_velocity = out.traction_relative = 0;
        out.finite = 1;
        for (const auto &p : host_partials) {
            out.energy += p.energy;
            out.max_velocity = std::max(out.max_velocity, p.velocity);
            out.traction_relative = std::max(out.traction_relative, p.traction);
            if (p.bad) out.finite = 0;
        }
        if (!std::isfinite(out.energy)) out.finite = 0;
        out.rhs_registers = registers;
        out.rhs_local_bytes = local_bytes;
    }
    void read_indices(const std::vector<long long> &internal, int nf, double *out) {
        const int n = (int)internal.size();
        DeviceBuffer di;
        di.upload(internal.data(), internal.size() * sizeof(long long));
        gather.allocate((size_t)std::max(n, 1) * nf * sizeof(double));
        gather_kernel<R><<<(n + 127) / 128, 128, 0, stream>>>(S(cur), count, di.as<long long>(), n, nf, gather.as<double>());
        check(hipGetLastError(), "2D gather");
        check(hipMemcpyAsync(out, gather.ptr, (size_t)n * nf * sizeof(double), hipMemcpyDeviceToHost, stream), "2D gather");
        check(hipStreamSynchronize(stream), "2D gather");
    }
    void sample(double *out) {
        DeviceScope scope(device);
        std::vector<long long> pi(2 * (size_t)std::max(nrec, 1));
        check(hipMemcpy(pi.data(), probes.ptr, pi.size() * sizeof(long long), hipMemcpyDeviceToHost), "probe indices");
        std::vector<long long> idx_(nrec);
        for (int r = 0; r < nrec; ++r) idx_[r] = pi[2 * r];
        if (nrec) read_indices(idx_, 2, out);
    }
    void probes_at(const int64_t *indices, int n, double *out) {
        DeviceScope scope(device);
        std::vector<long long> v(n);
        for (int r = 0; r < n; ++r) {
            require(indices[r] >= 0 && indices[r] < (int64_t)nx * nz, "probe index outside storage");
            v[r] = (long long)internal_of(indices[r]);
        }
        if (n) read_indices(v, 5, out);
    }
    void snapshot(int f, double *out) {
        DeviceScope scope(device);
        require(f >= 0 && f < fields, "2D snapshot field out of range");
        std::vector<R> v(count);
        check(hipStreamSynchronize(stream), "2D snapshot"); // the null-stream copy does not wait for it
        check(hipMemcpy(v.data(), S(cur) + (size_t)f * count, count * sizeof(R), hipMemcpyDeviceToHost), "2D snapshot");
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) out[(size_t)i + (size_t)nx * k] = v[idx(i, k)];
    }
    void pml(double *out) {
        DeviceScope scope(device);
        if (!cpml) return;
        std::vector<R> v(cpml);
        check(hipStreamSynchronize(stream), "2D PML snapshot");
        check(hipMemcpy(v.data(), P(cur), cpml * sizeof(R), hipMemcpyDeviceToHost), "2D PML snapshot");
        for (long long q = 0; q < cpml; ++q) out[q] = v[q];
    }
    void frame(int component, float *out, int fnx, int fnz, int foff) {
        DeviceScope scope(device);
        require(component >= 0 && component < 4, "Invalid 2D display component.");
        require(fnx >= 1 && fnz >= 1 && foff >= 0 && foff + fnx <= nx - 6 && fnz <= nz - 6, "frame outside the domain");
        const int n = fnx * fnz;
        frame_kernel<R><<<(n + 127) / 128, 128, 0, stream>>>(m, S(cur), count, component, display.as<float>(), fnx, fnz, foff);
        check(hipGetLastError(), "2D display launch");
        check(hipMemcpyAsync(out, display.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream), "2D display");
        check(hipStreamSynchronize(stream), "2D display");
    }
    uint64_t resident() const {
        uint64_t b = 0;
        for (const auto &x : mesh_arrays) b += x.bytes;
        for (const auto &x : state) b += x.bytes;
        for (const auto &x : pstate) b += x.bytes;
        for (const DeviceBuffer *x : {&accbuf, &paccbuf, &partials, &level2, &probes, &status, &accepted, &peak, &diag,
                                      &samples, &gather, &display, &borderNodes, &pmlLanes, &rowTables})
            b += x->bytes;
        return b;
    }

    // Device-resident stepping with the reference's inspection / fault rule / sampling per step.
    // Device-resident run in three parts, so a runner can interleave many engines over a few streams:
    // run_begin (counters), run_steps (steps j0..j1-1 of the run, rows j0..j1-1), run_end (verdict copy);
    // after the stream has finished, run_finish reads the verdict and the rows.  Every part enqueues on
    // the stream it is given; the engine's own stream is used by run().
    struct Verdict {
        int accepted, status;
        double peak;
    };
    struct OnStream { // launches of the engine go to `s` for the scope
        hipStream_t &ref, old;
        OnStream(hipStream_t &r, hipStream_t s) : ref(r), old(r) { ref = s; }
        ~OnStream() { ref = old; }
    };
    double step_bytes() const { return (double)count * fields * sizeof(R); } // state bytes, a load-balance weight
    void reserve_rows(int count_steps) {
        if (diag_rows >= count_steps) return;
        diag.allocate((size_t)count_steps * 4 * sizeof(double));
        samples.allocate((size_t)count_steps * 2 * std::max(nrec, 1) * sizeof(double));
        diag_rows = count_steps;
    }
    // rows the enqueued kernels write: slot 0 diag / samples, slot 1 diagAlt / samplesAlt (pipelined runs)
    DeviceBuffer diagAlt, samplesAlt;
    int rowSlot = 0, altRows = 0;
    double *rows_diag() const { return (rowSlot ? diagAlt : diag).as<double>(); }
    double *rows_samples() const { return (rowSlot ? samplesAlt : samples).as<double>(); }
    void run_begin(hipStream_t s, double peak0) {
        run_begin_kernel<<<1, 64, 0, s>>>(status.as<int>(), accepted.as<int>(), peak.as<double>(), peak0);
        check(hipGetLastError(), "2D run start");
    }
    void run_steps(hipStream_t s, int first, int j0, int j1, double dt, double frequency, double scale) {
        require(j1 <= (rowSlot ? altRows : diag_rows), "run rows not reserved");
        OnStream on(stream, s);
        int c = (cur + 2 * j0) % NB;
        if (endInStage) {
            // step j ends in the first stage launch of step j + 1; the run's last step by the kernels below
            for (int j = j0; j < j1; ++j) {
                const int st = first + j;
                for (int q = 0; q < 4; ++q) {
                    StageArgs<R> a = stage_args(c, dt, frequency, scale, st, q);
                    if (q == 0 && j > j0) {
                        StepEnd &E = a.end;
                        E.partials = partials.as<Partial>(), E.n = nPartials;
                        E.status = status.as<int>(), E.accepted = accepted.as<int>(), E.peak = peak.as<double>();
                        E.diag = rows_diag(), E.samples = rows_samples(), E.probes = probes.as<long long>();
                        E.nrec = nrec, E.row = j - 1, E.step = st - 1, E.dt = dt, E.frequency = frequency;
                    }
                    launch_stage(direction(st, q, 0), direction(st, q, 2), a, q);
                }
                c = (c + 2) % NB;
            }
            if (j1 > j0) finish_step(j1 - 1, first + j1 - 1, c, dt, frequency);
            return;
        }
        for (int j = j0; j < j1; ++j) {
            const int st = first + j;
            const int n = enqueue_step(c, dt, frequency, scale, st);
            const bool timed = j % 32 == 0; // finalisation is timed on every 32nd step only (events add gaps)
            if (timed) mark(1, true);
            finish_step(j, st, n, dt, frequency);
            if (timed) mark(1, false);
            c = n;
        }
    }
    // Separate kernels end step `st` (diagnostics row j, new state in buffer n): up to kDirectPartials tile
    // partials the finalisation block reduces them itself; beyond, a first level spreads the reduction over up to
    // 128 blocks.
    void finish_step(int j, int st, int n, double dt, double frequency) {
        const bool direct = nPartials <= kDirectPartials;
        if (!direct)
            reduce_kernel<<<nLevel2, THREADS, 0, stream>>>(partials.as<Partial>(), nPartials, level2.as<Partial>(),
                                                          status.as<int>());
        FinalArgs<R> f{};
        f.partials = direct ? partials.as<Partial>() : level2.as<Partial>(), f.n = direct ? nPartials : nLevel2;
        f.status = status.as<int>(), f.accepted = accepted.as<int>(), f.peak = peak.as<double>();
        f.diag = rows_diag(), f.samples = rows_samples();
        f.state = S(n), f.probes = probes.as<long long>(), f.nrec = nrec, f.row = j, f.step = st;
        f.dt = dt, f.frequency = frequency;
        finalize_kernel<R><<<1, THREADS, 0, stream>>>(f);
        check(hipGetLastError(), "2D finalize launch");
    }
    void run_end(hipStream_t s, Verdict *v) {
        check(hipMemcpyAsync(&v->accepted, accepted.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "2D run readback");
        check(hipMemcpyAsync(&v->status, status.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "2D run readback");
        check(hipMemcpyAsync(&v->peak, peak.ptr, sizeof(double), hipMemcpyDeviceToHost, s), "2D run readback");
    }
    void run_finish(const Verdict &v, double *diag_out, double *samples_out, swl_run2d &res) {
        collect();
        const int a = v.accepted;
        if (a) {
            if (diag_out)
                check(hipMemcpy(diag_out, diag.ptr, (size_t)a * 4 * sizeof(double), hipMemcpyDeviceToHost),
                      "2D run diagnostics");
            if (samples_out && nrec)
                check(hipMemcpy(samples_out, samples.ptr, (size_t)a * 2 * nrec * sizeof(double), hipMemcpyDeviceToHost),
                      "2D run samples");
        }
        res.accepted = a;
        res.faulted = v.status != 0;
        res.peak_energy = v.peak;
        // the step from buffer c ends in (c + 2) % NB
        prev = a ? (cur + 2 * (a - 1)) % NB : cur;
        cur = (cur + 2 * a) % NB;
    }
    void run(int first, int count_steps, double dt, double frequency, double scale, double peak0, double *diag_out,
             double *samples_out, swl_run2d &res) {
        DeviceScope scope(device);
        require(count_steps >= 0, "negative step count");
        res.accepted = 0, res.faulted = 0, res.peak_energy = peak0;
        if (!count_steps) return;
        reserve_rows(count_steps);
        run_begin(stream, peak0);
        run_steps(stream, first, 0, count_steps, dt, frequency, scale);
        Verdict v{};
        run_end(stream, &v);
        check(hipStreamSynchronize(stream), "2D run");
        run_finish(v, diag_out, samples_out, res);
    }

    // ---- pipelined runs: a caller enqueues the next run before it reads this one's rows, so the host's work
    // on them overlaps the device's.  Two row slots; a continuation run keeps the device's fault latch and
    // accepted peak (the device has already applied the previous run's steps to them).  Each slot remembers
    // the buffer its run starts from: the state buffer advances by the steps actually accepted, and a run
    // enqueued after a faulted one accepts none.
    struct Pending {
        int start = 0, count = 0, active = 0;
        hipEvent_t done{};
    };
    Pending pending[2];
    PinnedBuffer verdictHost; // one Verdict per slot
    int plannedCur = 0;       // buffer after every enqueued run, if all their steps are accepted
    void run_enqueue(int slot, int first, int count_steps, double dt, double frequency, double scale, double peak0,
                     bool fresh) {
        DeviceScope scope(device);
        require(slot == 0 || slot == 1, "run slot must be 0 or 1");
        require(count_steps > 0, "a pipelined run needs at least one step");
        Pending &P = pending[slot];
        require(!P.active, "run slot already in use");
        require(fresh || pending[slot ^ 1].active || plannedCur == cur, "continuation without a run before it");
        if (slot) {
            if (altRows < count_steps) {
                diagAlt.allocate((size_t)count_steps * 4 * sizeof(double));
                samplesAlt.allocate((size_t)count_steps * 2 * std::max(nrec, 1) * sizeof(double));
                altRows = count_steps;
            }
        } else {
            reserve_rows(count_steps);
        }
        verdictHost.reserve(2 * sizeof(Verdict));
        if (!P.done) check(hipEventCreateWithFlags(&P.done, hipEventDisableTiming), "event");
        if (fresh) {
            plannedCur = cur;
            run_begin(stream, peak0);
        } else {
            run_continue_kernel<<<1, 64, 0, stream>>>(accepted.as<int>());
            check(hipGetLastError(), "2D run continuation");
        }
        const int keep = cur;
        cur = plannedCur, rowSlot = slot;
        run_steps(stream, first, 0, count_steps, dt, frequency, scale);
        cur = keep, rowSlot = 0;
        run_end(stream, static_cast<Verdict *>(verdictHost.ptr) + slot);
        check(hipEventRecord(P.done, stream), "2D run event");
        P.start = plannedCur, P.count = count_steps, P.active = 1;
        plannedCur = (plannedCur + 2 * count_steps) % NB;
    }
    void run_collect(int slot, double *diag_out, double *samples_out, swl_run2d &res) {
        DeviceScope scope(device);
        require(slot == 0 || slot == 1, "run slot must be 0 or 1");
        Pending &P = pending[slot];
        require(P.active, "no run enqueued in this slot");
        check(hipEventSynchronize(P.done), "2D run");
        P.active = 0;
        const Verdict v = static_cast<const Verdict *>(verdictHost.ptr)[slot];
        const int a = v.accepted;
        if (a) {
            if (diag_out)
                check(hipMemcpy(diag_out, (slot ? diagAlt : diag).ptr, (size_t)a * 4 * sizeof(double), hipMemcpyDeviceToHost),
                      "2D run diagnostics");
            if (samples_out && nrec)
                check(hipMemcpy(samples_out, (slot ? samplesAlt : samples).ptr, (size_t)a * 2 * nrec * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "2D run samples");
            prev = (P.start + 2 * (a - 1)) % NB;
            cur = (P.start + 2 * a) % NB;
        }
        res.accepted = a;
        res.faulted = v.status != 0;
        res.peak_energy = v.peak;
        if (!pending[slot ^ 1].active) {
            plannedCur = cur;
            collect(); // profiling events, once nothing is in flight
        }
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
        std::lock_guard<std::mutex> setup(swl::device_setup_mutex());
        swl::require(mesh && out && receivers >= 0 && (receivers == 0 || receiver_index), "invalid 2D engine arguments");
        auto e = std::make_unique<swl_engine2d>();
        e->precision = mesh->precision;
        if (mesh->precision == 1) e->e32 = std::make_unique<swl::Engine2D<float>>(*mesh, receiver_index, receivers, device);
        else e->e64 = std::make_unique<swl::Engine2D<double>>(*mesh, receiver_index, receivers, device);
        *out = e.release();
    });
}
void swl_engine2d_destroy(swl_engine2d *e) {
    std::lock_guard<std::mutex> setup(swl::device_setup_mutex()); // frees synchronise the device
    delete e;
}
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
int swl_engine2d_diagnostics_noalloc(swl_engine2d *e, swl_diag2d *out) {
    return swl::guarded([&] { e->with([&](auto &x) { x.diagnostics_noalloc(*out); }); });
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
            out[4] = (double)x.nInterior, out[5] = (double)x.registers;
            out[6] = (double)((x.nx - 6) * (long long)(x.nz - 6)); // nodes updated per stage launch
            out[7] = (double)x.nGeneral, out[8] = (double)x.m.metric, out[9] = (double)x.m.by_row;
            out[10] = (double)x.nPml, out[11] = (double)x.nPmlBlocks, out[12] = (double)x.latency_kernel();
            out[13] = (double)x.tile_lds();
        });
    });
}
int swl_engine2d_set_option(swl_engine2d *e, const char *name, int value) {
    return swl::guarded([&] {
        const std::string key = name ? name : "";
        e->with([&](auto &x) {
            if (key == "nontemporal_stores") x.nt = value != 0;
            else if (key == "stored_metrics") x.m.metric = value == 0 && x.metricOk; // read a, b, J from memory
            else if (key == "lds_tiles") x.tileLds = value < 0 ? -1 : value != 0;
            else if (key == "material_rows") x.m.by_row = value != 0 && x.rowsOk; // 0: read the per-node arrays
            else if (key == "end_in_stage") x.endInStage = value != 0; // 0: reduce + finalize kernels after every step
            else if (key == "split_border") x.splitBorder = value < 0 ? -1 : value != 0; // -1: automatic
            else if (key == "spread_border") x.spreadBorder = value != 0;
            else if (key == "pml_vector") x.set_pml_vector(value);
            else if (key == "latency_kernel") x.latencyKernel = value < 0 ? -1 : value != 0;
            else throw std::runtime_error("unknown engine option: " + key);
        });
    });
}
int swl_engine2d_run_enqueue(swl_engine2d *e, int slot, int first_step, int count, double dt, double frequency,
                             double scale, double peak_energy, int fresh) {
    return swl::guarded([&] {
        e->with([&](auto &x) { x.run_enqueue(slot, first_step, count, dt, frequency, scale, peak_energy, fresh != 0); });
    });
}
int swl_engine2d_run_collect(swl_engine2d *e, int slot, double *diag, double *samples, swl_run2d *result) {
    return swl::guarded([&] { e->with([&](auto &x) { x.run_collect(slot, diag, samples, *result); }); });
}
int swl_engine2d_run(swl_engine2d *e, int first_step, int count, double dt, double frequency, double scale,
                     double peak_energy, double *diag, double *samples, swl_run2d *result) {
    return swl::guarded([&] {
        e->with([&](auto &x) { x.run(first_step, count, dt, frequency, scale, peak_energy, diag, samples, *result); });
    });
}
int swl_stream_probe(int device, uint64_t bytes, int repeats, double *ms_per_copy) {
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
        const unsigned grid = (unsigned)((n + 255) / 256);
        swl::stream_copy_kernel<<<grid, 256>>>(a.as<float4>(), b.as<float4>(), n);
        swl::check(hipEventRecord(e0, nullptr), "event");
        for (int r = 0; r < repeats; ++r) swl::stream_copy_kernel<<<grid, 256>>>(a.as<float4>(), b.as<float4>(), n);
        swl::check(hipEventRecord(e1, nullptr), "event");
        swl::check(hipEventSynchronize(e1), "probe");
        float ms = 0;
        swl::check(hipEventElapsedTime(&ms, e0, e1), "event");
        *ms_per_copy = ms / repeats;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}
}
