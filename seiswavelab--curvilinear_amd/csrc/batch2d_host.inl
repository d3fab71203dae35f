// batch2d_host.inl — multi-shot runner for 2D engines (included by engine2d.hip after the engine).
//
// Steps many independent shots together: per RK stage, ONE launch of batch_stage_kernel walks the
// tiles of every shot of a (precision, GSLS mechanism count, stencil direction pair) group, so small
// shots share the device instead of each under-filling it with its own launches.  A chunk of K steps
// (stage launches + one per-shot finalisation launch per step) is captured into a hipGraph and
// replayed while the launch structure repeats; the per-chunk arguments (rotating buffers, Ricker
// forcing, counters) live in device buffers refreshed before each replay.  A small autotuner times
// the first chunks with different persistent-grid sizes and keeps the fastest.
// No counterpart in the reference: each of its sessions steps one shot (CudaEngine2D.cu:143-152).
namespace swl {
namespace {

struct BatchLaunch {
    int kind; // 0 stage, 1 final, 2 begin
    int group, dx, dz, mech, grid;
    size_t offset; // element offset into the args / finals array
    int count;     // shots in the launch (finals / begin)
};

template <class R> struct BatchTyped {
    struct Group {
        int mech = 0;
        std::vector<int> members; // indices into the shot list
        std::vector<int2> work;
        DeviceBuffer workBuf;
    };
    std::vector<Group> groups;
    std::vector<uint64_t> key; // engines (by uid) the groups were built for
    DeviceBuffer args, finals, begin;
    PinnedBuffer hargs, hfinals, hbegin;
    std::vector<BatchLaunch> launches;
    std::string signature, cachedSignature;
    hipGraphExec_t graph = nullptr;

    ~BatchTyped() {
        if (graph) (void)hipGraphExecDestroy(graph);
    }

    void build_groups(const std::vector<Engine2D<R> *> &engines) {
        std::vector<uint64_t> ids;
        for (auto *e : engines) ids.push_back(e->uid), ids.push_back((uint64_t)e->nGeneral << 32 | (uint32_t)e->nPmlBlocks);
        if (ids == key) return;
        key = ids;
        groups.clear();
        for (int mech = 0; mech <= 3; ++mech) {
            Group g;
            g.mech = mech;
            for (int i = 0; i < (int)engines.size(); ++i)
                if (engines[i]->mech == mech) g.members.push_back(i);
            if (g.members.empty()) continue;
            // each shot's segment starts on a multiple of 8 (XCD round-robin), border tiles first, then the
            // interior tiles in the XCD-band order of the single-shot kernel
            for (int local = 0; local < (int)g.members.size(); ++local) {
                const Engine2D<R> &e = *engines[g.members[local]];
                const int gblocks = (e.nGeneral + 7) / 8 * 8, pblocks = (e.nPmlBlocks + 7) / 8 * 8;
                for (int b = 0; b < gblocks; ++b) g.work.push_back(int2{local, b < e.nGeneral ? b : -1});
                for (int b = 0; b < pblocks; ++b) g.work.push_back(int2{local, b < e.nPmlBlocks ? (b | BATCH_PML) : -1});
                const int n = e.nInterior, padded = (n + 7) / 8 * 8;
                for (int b = 0; b < padded; ++b) {
                    // tile_of_block(b, n): position b (XCD b % 8) takes tile x*q + min(x, r) + j, a bijection of [0, n)
                    const int x = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
                    g.work.push_back(int2{local, b < n ? ((x * q + std::min(x, r) + j) | BATCH_INTERIOR) : -1});
                }
            }
            g.workBuf.upload(g.work.data(), g.work.size() * sizeof(int2));
            groups.push_back(std::move(g));
        }
        cachedSignature.clear();
    }
};

struct Runner2D {
    int device = 0;
    hipStream_t stream{};
    // fused = 1: the batch kernels above (one launch per stage over every shot's tiles);
    // fused = 0: each shot keeps its own specialised stage kernels and the shots are spread over `lanes`
    // streams (longest-first onto the least-loaded lane), so small shots' launches overlap on the device.
    // Three lanes measured best for shots of ~0.1-1 M nodes (GSLS-3 FP32 with the memory variables rebuilt
    // per node: ~70 MB of state per 0.33 M-node shot); four lanes' working sets overflow the 256 MB
    // Infinity Cache.
    int fused = 0, lanes = 3, laneSteps = 32;
    std::vector<hipStream_t> laneStreams;
    std::vector<hipEvent_t> laneEvents; // laneDepth per lane: ends of that lane's last turns (a ring)
    int laneDepth = 2;                  // turns a lane may have enqueued ahead of the device
    PinnedBuffer verdicts;
    long long laneRuns = 0;
    int chunk = 32, useGraphs = 1, autotune = 1, fixedGrid = 0;
    int cus = 256;
    // autotuner: candidate persistent grids (0 = one block per work item)
    std::vector<int> candidates;
    std::vector<double> candidateMs;
    int tuneNext = 0, chosen = -1;
    long long captures = 0, replays = 0, directChunks = 0, chunks = 0, captureFailures = 0;
    hipEvent_t t0{}, t1{};
    BatchTyped<double> b64;
    BatchTyped<float> b32;

    explicit Runner2D(int dev) : device(dev) {
        DeviceScope scope(device);
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "batch stream");
        check(hipEventCreate(&t0), "event");
        check(hipEventCreate(&t1), "event");
        hipDeviceProp_t prop{};
        check(hipGetDeviceProperties(&prop, device), "device properties");
        cus = prop.multiProcessorCount;
        candidates = {0, 8 * cus, 4 * cus, 2 * cus};
    }
    ~Runner2D() {
        for (auto e : laneEvents) (void)hipEventDestroy(e);
        for (auto s : laneStreams) (void)hipStreamDestroy(s);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int grid_for(int nwork) const {
        int g = fixedGrid > 0 ? fixedGrid : candidates[chosen >= 0 ? chosen : (autotune ? tuneNext : 0)];
        return g <= 0 || g > nwork ? nwork : g;
    }

    template <class R> void launch_stage(const BatchLaunch &L, BatchTyped<R> &B) {
        const auto &g = B.groups[L.group];
        const StageArgs<R> *a = B.args.template as<StageArgs<R>>() + L.offset;
        const int2 *w = g.workBuf.template as<int2>();
        const int n = (int)g.work.size();
#define SWL_BATCH_LAUNCH(DX, DZ, MECH)                                                                               \
    if (L.grid >= n) batch_stage_kernel<R, DX, DZ, MECH, 0><<<n, THREADS, 0, stream>>>(a, w, n);                     \
    else batch_stage_kernel<R, DX, DZ, MECH, 1><<<L.grid, THREADS, 0, stream>>>(a, w, n)
#define SWL_BATCH_MECH(DX, DZ)                                                                                       \
    switch (L.mech) {                                                                                                \
    case 0: SWL_BATCH_LAUNCH(DX, DZ, 0); break;                                                                      \
    case 1: SWL_BATCH_LAUNCH(DX, DZ, 1); break;                                                                      \
    case 2: SWL_BATCH_LAUNCH(DX, DZ, 2); break;                                                                      \
    default: SWL_BATCH_LAUNCH(DX, DZ, 3); break;                                                                     \
    }
        if (L.dx == 0 && L.dz == 0) { SWL_BATCH_MECH(0, 0) }
        else if (L.dx == 0) { SWL_BATCH_MECH(0, 1) }
        else if (L.dz == 0) { SWL_BATCH_MECH(1, 0) }
        else { SWL_BATCH_MECH(1, 1) }
#undef SWL_BATCH_MECH
#undef SWL_BATCH_LAUNCH
        check(hipGetLastError(), "batch stage launch");
    }
    template <class R> void enqueue(BatchTyped<R> &B, int nshots) {
        for (const auto &L : B.launches) {
            if (L.kind == 0) launch_stage(L, B);
            else if (L.kind == 1)
                batch_final_kernel<R><<<L.count, THREADS, 0, stream>>>(B.finals.template as<FinalArgs<R>>() + L.offset);
            else {
                const char *base = B.begin.template as<char>();
                batch_begin_kernel<<<(nshots + 63) / 64, 64, 0, stream>>>(
                    reinterpret_cast<int *const *>(base), reinterpret_cast<int *const *>(base + 8 * (size_t)nshots),
                    reinterpret_cast<double *const *>(base + 16 * (size_t)nshots),
                    reinterpret_cast<const double *>(base + 24 * (size_t)nshots), nshots);
            }
            check(hipGetLastError(), "batch launch");
        }
    }

    // Runs every shot of one precision for up to its `count` steps.
    template <class R> void run_typed(BatchTyped<R> &B, std::vector<Engine2D<R> *> &E, std::vector<swl_batch_shot *> &S) {
        const int n = (int)E.size();
        if (!n) return;
        B.build_groups(E);
        std::vector<int> done(n, 0), left(n), faulted(n, 0);
        std::vector<double> peak(n);
        for (int i = 0; i < n; ++i) {
            left[i] = std::max(0, S[i]->count), peak[i] = S[i]->peak_energy;
            S[i]->result.accepted = 0, S[i]->result.faulted = 0, S[i]->result.peak_energy = peak[i];
            if (left[i] > 0 && E[i]->diag_rows < chunk) {
                E[i]->diag.allocate((size_t)chunk * 4 * sizeof(double));
                E[i]->samples.allocate((size_t)chunk * 2 * std::max(E[i]->nrec, 1) * sizeof(double));
                E[i]->diag_rows = chunk;
            }
        }
        for (;;) {
            int K = 0;
            for (int i = 0; i < n; ++i) K = std::max(K, std::min(left[i], chunk));
            if (!K) break;
            // ---- arguments of this chunk
            B.launches.clear();
            std::vector<StageArgs<R>> args;
            std::vector<FinalArgs<R>> fin;
            std::ostringstream sig;
            sig << K << ':' << (void *)B.args.ptr << ':' << (void *)B.finals.ptr;
            B.launches.push_back(BatchLaunch{2, -1, 0, 0, 0, 0, 0, n});
            for (int j = 0; j < K; ++j) {
                for (int s = 0; s < 4; ++s)
                    for (int gi = 0; gi < (int)B.groups.size(); ++gi) {
                        const auto &g = B.groups[gi];
                        for (int combo = 0; combo < 4; ++combo) {
                            const int dx = combo >> 1, dz = combo & 1;
                            bool any = false;
                            const size_t offset = args.size();
                            for (int idx : g.members) {
                                Engine2D<R> &e = *E[idx];
                                const int step = S[idx]->first_step + done[idx] + j;
                                StageArgs<R> a{};
                                a.kind = -1;
                                if (j < std::min(left[idx], chunk) &&
                                    Engine2D<R>::direction(step, s, 0) == dx && Engine2D<R>::direction(step, s, 2) == dz) {
                                    const int c = (e.cur + 2 * j) % Engine2D<R>::NB;
                                    a = e.stage_args(c, S[idx]->dt, S[idx]->frequency, S[idx]->scale, step, s);
                                    any = true;
                                }
                                args.push_back(a);
                            }
                            if (!any) {
                                args.resize(offset);
                                continue;
                            }
                            const int grid = grid_for((int)g.work.size());
                            B.launches.push_back(BatchLaunch{0, gi, dx, dz, g.mech, grid, offset, (int)g.members.size()});
                            sig << "|s" << gi << dx << dz << ',' << grid << ',' << offset;
                        }
                    }
                for (int gi = 0; gi < (int)B.groups.size(); ++gi) {
                    const auto &g = B.groups[gi];
                    const size_t offset = fin.size();
                    for (int idx : g.members) {
                        Engine2D<R> &e = *E[idx];
                        FinalArgs<R> f{};
                        if (j < std::min(left[idx], chunk)) {
                            f.partials = e.partials.template as<Partial>(), f.n = e.nPartials;
                            f.status = e.status.template as<int>(), f.accepted = e.accepted.template as<int>();
                            f.peak = e.peak.template as<double>();
                            f.diag = e.diag.template as<double>(), f.samples = e.samples.template as<double>();
                            f.state = e.S((e.cur + 2 * (j + 1)) % Engine2D<R>::NB), f.probes = e.probes.template as<long long>();
                            f.nrec = e.nrec, f.row = j, f.step = S[idx]->first_step + done[idx] + j;
                            f.dt = S[idx]->dt, f.frequency = S[idx]->frequency;
                        }
                        fin.push_back(f);
                    }
                    B.launches.push_back(BatchLaunch{1, gi, 0, 0, g.mech, (int)g.members.size(), offset, (int)g.members.size()});
                    sig << "|f" << gi << ',' << offset;
                }
            }
            // per-shot counters (status, accepted, peak pointers, initial peak) for the begin kernel
            const size_t beginBytes = 32 * (size_t)n;
            B.hbegin.reserve(beginBytes);
            {
                char *h = static_cast<char *>(B.hbegin.ptr);
                for (int i = 0; i < n; ++i) {
                    int *st = E[i]->status.template as<int>(), *ac = E[i]->accepted.template as<int>();
                    double *pk = E[i]->peak.template as<double>();
                    std::memcpy(h + 8 * (size_t)i, &st, 8);
                    std::memcpy(h + 8 * ((size_t)n + i), &ac, 8);
                    std::memcpy(h + 8 * (2 * (size_t)n + i), &pk, 8);
                    std::memcpy(h + 8 * (3 * (size_t)n + i), &peak[i], 8);
                }
            }
            // device buffers keep their addresses while they are large enough (graph replays depend on it)
            const size_t argBytes = args.size() * sizeof(StageArgs<R>), finBytes = fin.size() * sizeof(FinalArgs<R>);
            if (B.args.bytes < argBytes) B.args.allocate(argBytes + argBytes / 4);
            if (B.finals.bytes < finBytes) B.finals.allocate(finBytes + finBytes / 4);
            if (B.begin.bytes < beginBytes) B.begin.allocate(beginBytes);
            B.hargs.reserve(argBytes), B.hfinals.reserve(finBytes);
            std::memcpy(B.hargs.ptr, args.data(), argBytes);
            std::memcpy(B.hfinals.ptr, fin.data(), finBytes);
            check(hipMemcpyAsync(B.args.ptr, B.hargs.ptr, argBytes, hipMemcpyHostToDevice, stream), "batch arguments");
            check(hipMemcpyAsync(B.finals.ptr, B.hfinals.ptr, finBytes, hipMemcpyHostToDevice, stream), "batch arguments");
            check(hipMemcpyAsync(B.begin.ptr, B.hbegin.ptr, beginBytes, hipMemcpyHostToDevice, stream), "batch counters");
            sig << '#' << (void *)B.args.ptr << (void *)B.finals.ptr << (void *)B.begin.ptr;
            B.signature = sig.str();
            // ---- launch (graph replay when the structure repeats)
            const bool timing = autotune && chosen < 0 && fixedGrid == 0 && K == chunk;
            if (timing) check(hipEventRecord(t0, stream), "batch timing");
            bool launched = false;
            if (useGraphs) {
                if (B.signature != B.cachedSignature || !B.graph) {
                    if (B.graph) (void)hipGraphExecDestroy(B.graph), B.graph = nullptr;
                    B.cachedSignature.clear();
                    // relaxed mode: sessions being built on other threads (allocations, uploads on their own
                    // streams) do not invalidate this capture
                    hipGraph_t g{};
                    std::unique_lock<std::mutex> setup(device_setup_mutex());
                    hipError_t ce = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed);
                    if (ce == hipSuccess) {
                        try {
                            enqueue(B, n);
                        } catch (const std::exception &) {
                            ce = hipErrorStreamCaptureInvalidated;
                        }
                        const hipError_t ee = hipStreamEndCapture(stream, &g);
                        if (ce == hipSuccess) ce = ee;
                    }
                    setup.unlock();
                    if (ce == hipSuccess) ce = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
                    if (g) (void)hipGraphDestroy(g);
                    if (ce == hipSuccess) {
                        B.cachedSignature = B.signature;
                        ++captures;
                    } else { // capture unavailable: launch directly from now on (same kernels, same results)
                        (void)hipGetLastError();
                        if (B.graph) (void)hipGraphExecDestroy(B.graph), B.graph = nullptr;
                        useGraphs = 0, ++captureFailures;
                        // continue on a fresh stream: the failed capture may leave state on the old one
                        hipStream_t fresh{};
                        check(hipStreamCreateWithFlags(&fresh, hipStreamNonBlocking), "batch stream");
                        (void)hipStreamDestroy(stream);
                        stream = fresh;
                        check(hipMemcpyAsync(B.args.ptr, B.hargs.ptr, argBytes, hipMemcpyHostToDevice, stream),
                              "batch arguments");
                        check(hipMemcpyAsync(B.finals.ptr, B.hfinals.ptr, finBytes, hipMemcpyHostToDevice, stream),
                              "batch arguments");
                        check(hipMemcpyAsync(B.begin.ptr, B.hbegin.ptr, beginBytes, hipMemcpyHostToDevice, stream),
                              "batch counters");
                    }
                } else {
                    ++replays;
                }
                if (B.graph) {
                    check(hipGraphLaunch(B.graph, stream), "batch graph launch");
                    launched = true;
                }
            }
            if (!launched) {
                enqueue(B, n);
                ++directChunks;
            }
            ++chunks;
            if (timing) {
                check(hipEventRecord(t1, stream), "batch timing");
                check(hipEventSynchronize(t1), "batch timing");
                float ms = 0;
                check(hipEventElapsedTime(&ms, t0, t1), "batch timing");
                candidateMs.push_back(ms);
                if (++tuneNext >= (int)candidates.size()) {
                    chosen = (int)(std::min_element(candidateMs.begin(), candidateMs.end()) - candidateMs.begin());
                }
            }
            // ---- results: per shot accepted / status / peak, rows
            std::vector<int> acc(n), st(n);
            std::vector<double> pk(n);
            for (int i = 0; i < n; ++i) {
                check(hipMemcpyAsync(&acc[i], E[i]->accepted.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "batch result");
                check(hipMemcpyAsync(&st[i], E[i]->status.ptr, sizeof(int), hipMemcpyDeviceToHost, stream), "batch result");
                check(hipMemcpyAsync(&pk[i], E[i]->peak.ptr, sizeof(double), hipMemcpyDeviceToHost, stream), "batch result");
            }
            check(hipStreamSynchronize(stream), "batch chunk");
            for (int i = 0; i < n; ++i) {
                const int ran = std::min(left[i], chunk);
                if (ran <= 0) continue;
                Engine2D<R> &e = *E[i];
                const int a = acc[i];
                if (a > 0) {
                    check(hipMemcpy(S[i]->diag + 4 * (size_t)done[i], e.diag.ptr, (size_t)a * 4 * sizeof(double),
                                    hipMemcpyDeviceToHost),
                          "batch diagnostics");
                    if (e.nrec)
                        check(hipMemcpy(S[i]->samples + 2 * (size_t)e.nrec * done[i], e.samples.ptr,
                                        (size_t)a * 2 * e.nrec * sizeof(double), hipMemcpyDeviceToHost),
                              "batch samples");
                    e.prev = (e.cur + 2 * (a - 1)) % Engine2D<R>::NB;
                    e.cur = (e.cur + 2 * a) % Engine2D<R>::NB;
                }
                done[i] += a, peak[i] = pk[i];
                S[i]->result.accepted = done[i], S[i]->result.peak_energy = pk[i];
                if (st[i]) S[i]->result.faulted = 1, left[i] = 0;
                else left[i] -= a;
            }
        }
    }

    // Lanes: every shot runs its own engine's kernels (as swl_engine2d_run would), enqueued `laneSteps`
    // steps at a time round-robin over the lanes so that every lane's hardware queue is fed from the start.
    void run_lanes(swl_batch_shot *shots, int n) {
        while ((int)laneStreams.size() < lanes) {
            hipStream_t s{};
            check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch lane stream");
            laneStreams.push_back(s);
        }
        while (laneEvents.size() < (size_t)laneDepth * laneStreams.size()) {
            hipEvent_t e{};
            check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch lane event");
            laneEvents.push_back(e);
        }
        struct Item {
            swl_batch_shot *shot;
            double cost;
        };
        std::vector<Item> items;
        for (int i = 0; i < n; ++i) {
            auto &sh = shots[i];
            sh.result.accepted = 0, sh.result.faulted = 0, sh.result.peak_energy = sh.peak_energy;
            if (sh.count <= 0) continue;
            double cost = 0;
            sh.engine->with([&](auto &e) {
                require(e.device == device, "batch shots must live on the runner's device");
                e.reserve_rows(sh.count); // allocations synchronize: all of them before anything is enqueued
                cost = e.step_bytes() * sh.count;
            });
            items.push_back({&sh, cost});
        }
        if (items.empty()) return;
        std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
        std::vector<std::vector<swl_batch_shot *>> plan(lanes);
        std::vector<double> load(lanes, 0.0);
        for (const auto &it : items) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            plan[l].push_back(it.shot), load[l] += it.cost;
        }
        verdicts.reserve(items.size() * 16);
        using V = Engine2D<double>::Verdict;
        static_assert(sizeof(V) == sizeof(Engine2D<float>::Verdict), "verdict layout");
        V *vs = static_cast<V *>(verdicts.ptr);
        std::vector<std::pair<swl_batch_shot *, V *>> order;
        for (int l = 0; l < lanes; ++l)
            for (auto *sh : plan[l]) order.push_back({sh, vs + order.size()});
        // per lane: index of the current shot in plan[l] and the next step of that shot
        std::vector<size_t> at(lanes, 0);
        std::vector<int> next(lanes, 0);
        std::vector<size_t> base(lanes, 0);
        std::vector<long long> turns(lanes, 0);
        for (int l = 1; l < lanes; ++l) base[l] = base[l - 1] + plan[l - 1].size();
        for (bool any = true; any;) {
            any = false;
            for (int l = 0; l < lanes; ++l) {
                if (at[l] >= plan[l].size()) continue;
                any = true;
                swl_batch_shot *sh = plan[l][at[l]];
                V *v = vs + base[l] + at[l];
                const hipStream_t s = laneStreams[l];
                // at most `laneDepth` turns of this lane in flight: the queues stay short, so set-up work that
                // shares a hardware queue with a lane (sessions being built on other threads) is not held
                // behind a whole chunk of stages
                hipEvent_t &done = laneEvents[(size_t)laneDepth * l + turns[l] % laneDepth];
                if (turns[l] >= laneDepth) check(hipEventSynchronize(done), "batch lane turn");
                const int j0 = next[l], j1 = std::min(sh->count, j0 + laneSteps);
                sh->engine->with([&](auto &e) {
                    if (j0 == 0) e.run_begin(s, sh->peak_energy);
                    e.run_steps(s, sh->first_step, j0, j1, sh->dt, sh->frequency, sh->scale);
                    if (j1 == sh->count) e.run_end(s, reinterpret_cast<typename std::decay_t<decltype(e)>::Verdict *>(v));
                });
                check(hipEventRecord(done, s), "batch lane turn");
                ++turns[l];
                next[l] = j1;
                if (j1 == sh->count) ++at[l], next[l] = 0;
            }
        }
        for (auto s : laneStreams) check(hipStreamSynchronize(s), "batch lanes");
        for (auto &[sh, v] : order)
            sh->engine->with([&](auto &e) {
                e.run_finish(*reinterpret_cast<const typename std::decay_t<decltype(e)>::Verdict *>(v), sh->diag,
                             sh->samples, sh->result);
            });
        ++laneRuns, ++chunks;
    }

    void run(swl_batch_shot *shots, int n) {
        DeviceScope scope(device);
        if (!fused) {
            run_lanes(shots, n);
            return;
        }
        std::vector<Engine2D<double> *> e64;
        std::vector<Engine2D<float> *> e32;
        std::vector<swl_batch_shot *> s64, s32;
        for (int i = 0; i < n; ++i) {
            require(shots[i].engine != nullptr, "batch shot without an engine");
            if (shots[i].engine->e64) e64.push_back(shots[i].engine->e64.get()), s64.push_back(&shots[i]);
            else e32.push_back(shots[i].engine->e32.get()), s32.push_back(&shots[i]);
        }
        for (auto *e : e64) require(e->device == device, "batch shots must live on the runner's device");
        for (auto *e : e32) require(e->device == device, "batch shots must live on the runner's device");
        run_typed(b64, e64, s64);
        run_typed(b32, e32, s32);
    }
};

} // namespace
} // namespace swl

struct swl_batch2d_runner {
    std::unique_ptr<swl::Runner2D> r;
};

extern "C" {
int swl_batch2d_runner_create(int device, swl_batch2d_runner **out) {
    return swl::guarded([&] {
        swl::require(out != nullptr, "invalid runner arguments");
        auto r = std::make_unique<swl_batch2d_runner>();
        r->r = std::make_unique<swl::Runner2D>(device);
        *out = r.release();
    });
}
void swl_batch2d_runner_destroy(swl_batch2d_runner *r) { delete r; }
int swl_batch2d_runner_set_option(swl_batch2d_runner *r, const char *name, int value) {
    return swl::guarded([&] {
        const std::string key = name ? name : "";
        auto &x = *r->r;
        if (key == "graphs") x.useGraphs = value != 0;
        else if (key == "autotune") x.autotune = value != 0, x.chosen = -1, x.tuneNext = 0, x.candidateMs.clear();
        else if (key == "grid") x.fixedGrid = std::max(0, value);
        else if (key == "chunk") x.chunk = std::clamp(value, 1, 4096);
        else if (key == "fused") x.fused = value != 0;
        else if (key == "lanes") {
            swl::require(value >= 1 && value <= 16, "lanes must be in 1..16");
            if (value != x.lanes) {
                for (auto e : x.laneEvents) (void)hipEventDestroy(e);
                x.laneEvents.clear();
                for (auto s : x.laneStreams) (void)hipStreamDestroy(s);
                x.laneStreams.clear();
                x.lanes = value;
            }
        } else if (key == "lane_steps") x.laneSteps = std::clamp(value, 1, 4096);
        else if (key == "lane_depth") {
            const int depth = std::clamp(value, 1, 64);
            if (depth != x.laneDepth) {
                for (auto e : x.laneEvents) (void)hipEventDestroy(e);
                x.laneEvents.clear();
                x.laneDepth = depth;
            }
        }
        else throw std::runtime_error("unknown batch option: " + key);
    });
}
int swl_batch2d_run(swl_batch2d_runner *r, swl_batch_shot *shots, int n) {
    return swl::guarded([&] { r->r->run(shots, n); });
}
int swl_batch2d_runner_stats(const swl_batch2d_runner *r, double *out) {
    const auto &x = *r->r;
    out[0] = x.chosen >= 0 ? x.candidates[x.chosen] : -1;
    out[1] = (double)x.captures, out[2] = (double)x.replays, out[3] = (double)x.directChunks, out[4] = (double)x.chunks;
    for (int c = 0; c < 4; ++c) out[5 + c] = c < (int)x.candidateMs.size() ? x.candidateMs[c] : -1;
    out[9] = x.chunk;
    out[10] = (double)x.captureFailures;
    out[11] = x.fused, out[12] = x.lanes, out[13] = (double)x.laneRuns;
    return 0;
}
} // extern "C"
