// async2d.cpp — background stepping of a 2D session: Seis2D::Curvilinear::AsyncSimulation
// (include/Seis2D/AsyncSimulation2D.hpp, src/two_d/AsyncSimulation2D.cpp) over swl_sim2d.
//
// One persistent worker thread owns the session while a batch runs.  queue(steps, budget) submits at
// most 8 single steps that stop early when the time budget is spent, the session completes, or the
// caller pauses.  The batch publishes the step, the diagnostics, the receiver samples recorded since the
// batch was queued and, when the caller has asked for one, a display frame.  poll() / synchronize() fold
// a finished publication into the caller's copy, so readers on the caller's thread only ever see
// published state.  synchronize() (and every call that touches the session itself) cancels the batch
// between steps and waits for it, as SeisWave::BackgroundRunner::wait does.
#include "background.hpp"
#include "session2d.hpp"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <optional>
#include <thread>

namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

constexpr int kDiag = 22; // swl_sim2d_diagnostics values

struct Publication {
    int step = 0, advanced = 0;
    double diag[kDiag]{};
    std::vector<std::vector<double>> vx, vz; // samples recorded during the batch, per receiver
    std::vector<float> frame;
    int component = -1, includeAbsorber = 0;
    double computeMs = 0, transferMs = 0;
    std::string error;
};

} // namespace

struct swl_async2d {
    swl_sim2d *core = nullptr; // owned
    swl::BackgroundWorker<Publication> worker;
    // the caller's published copy
    bool initialized = false;
    int step = 0, total = 0;
    double dt = 0, recordDt = 0;
    double diag[kDiag]{};
    std::vector<std::vector<double>> vx, vz;
    std::vector<float> frame;
    int requested = -1, requestedAbsorber = 0, component = -1, frameAbsorber = 0, frameStep = -1;
    double computeMs = 0, transferMs = 0;
    uint64_t batches = 0, steps = 0;
    std::string workerError;

    ~swl_async2d() {
        try {
            synchronize();
        } catch (...) {
        }
        if (core) swl_sim2d_destroy(core);
    }
    void refresh() { // the session's state becomes the published state
        initialized = core != nullptr && core->engine != nullptr;
        if (!initialized) return;
        step = core->step, total = core->model.total, dt = core->model.dt, recordDt = core->recordDt();
        swl_sim2d_diagnostics(core, diag);
        vx = core->vx, vz = core->vz;
        frameStep = -1;
    }
    void apply(Publication &&p) {
        std::memcpy(diag, p.diag, sizeof diag);
        step = p.step;
        for (size_t r = 0; r < vx.size() && r < p.vx.size(); ++r) {
            vx[r].insert(vx[r].end(), p.vx[r].begin(), p.vx[r].end());
            vz[r].insert(vz[r].end(), p.vz[r].begin(), p.vz[r].end());
        }
        if (p.component >= 0) frame = std::move(p.frame), component = p.component, frameAbsorber = p.includeAbsorber, frameStep = step;
        computeMs += p.computeMs, transferMs += p.transferMs, ++batches, steps += p.advanced;
        workerError = std::move(p.error);
    }
    void synchronize() {
        if (auto p = worker.wait()) apply(std::move(*p));
    }
    bool completed() const { return initialized && step >= total; }
    bool queue(int count, double budgetMs) {
        if (!initialized || completed() || worker.busy()) return false;
        std::vector<size_t> counts;
        for (const auto &r : vx) counts.push_back(r.size());
        const int want = requested, absorber = requestedAbsorber, n = std::clamp(count, 1, 8);
        swl_sim2d *s = core;
        return worker.submit([s, counts, want, absorber, n, budgetMs](const std::atomic_bool &cancel) {
            Publication p;
            auto begin = Clock::now();
            const int before = s->step;
            for (int i = 0; i < n && !cancel && s->step < s->model.total; ++i) {
                if (swl_sim2d_advance(s, 1)) {
                    p.error = swl::last_error();
                    break;
                }
                if (since(begin) >= budgetMs) break;
            }
            p.computeMs = since(begin);
            p.advanced = s->step - before;
            begin = Clock::now();
            p.step = s->step;
            swl_sim2d_diagnostics(s, p.diag);
            p.vx.resize(s->vx.size()), p.vz.resize(s->vz.size());
            for (size_t r = 0; r < s->vx.size(); ++r) {
                p.vx[r].assign(s->vx[r].begin() + static_cast<std::ptrdiff_t>(counts[r]), s->vx[r].end());
                p.vz[r].assign(s->vz[r].begin() + static_cast<std::ptrdiff_t>(counts[r]), s->vz[r].end());
            }
            if (want >= 0) {
                const auto &m = s->model;
                const size_t len = static_cast<size_t>(absorber ? m.p.nx : m.logical.nx) * (absorber ? m.p.nz : m.logical.nz);
                p.frame.resize(len);
                if (swl_sim2d_read_frame(s, want, p.frame.data(), absorber) == 0) p.component = want, p.includeAbsorber = absorber;
                else if (p.error.empty()) p.error = swl::last_error();
            }
            p.transferMs = since(begin);
            return p;
        });
    }
    size_t frameSize(int absorber) const {
        const auto &m = core->model;
        return static_cast<size_t>(absorber ? m.p.nx : m.logical.nx) * (absorber ? m.p.nz : m.logical.nz);
    }
};

extern "C" {
int swl_async2d_create(swl_async2d **out) {
    return swl::guarded([&] {
        swl::require(out != nullptr, "invalid async session arguments");
        *out = new swl_async2d();
    });
}
void swl_async2d_destroy(swl_async2d *a) { delete a; }
int swl_async2d_initialize(swl_async2d *a, const swl_params2d *p, int backend, int device) {
    return swl::guarded([&] {
        a->synchronize();
        swl_sim2d *s = nullptr;
        const int rc = swl_sim2d_create(p, backend, device, &s);
        if (rc == 0) {
            if (a->core) swl_sim2d_destroy(a->core);
            a->core = s;
        }
        a->refresh();
        a->requested = -1, a->computeMs = a->transferMs = 0, a->batches = a->steps = 0, a->workerError.clear();
        if (rc) throw std::runtime_error(swl::last_error());
    });
}
int swl_async2d_queue(swl_async2d *a, int steps, double budget_ms) { return a->queue(steps, budget_ms) ? 1 : 0; }
int swl_async2d_poll(swl_async2d *a) {
    return swl::guarded([&] {
        if (auto p = a->worker.poll()) a->apply(std::move(*p));
        if (!a->workerError.empty()) throw std::runtime_error(std::exchange(a->workerError, {}));
    });
}
void swl_async2d_pause(swl_async2d *a) { a->worker.cancel(); }
int swl_async2d_synchronize(swl_async2d *a) {
    return swl::guarded([&] { a->synchronize(); });
}
int swl_async2d_busy(swl_async2d *a) { return a->worker.busy() ? 1 : 0; }
int swl_async2d_timing(const swl_async2d *a, double *out) {
    out[0] = a->computeMs, out[1] = a->transferMs, out[2] = static_cast<double>(a->batches), out[3] = static_cast<double>(a->steps);
    return 0;
}
int swl_async2d_advance(swl_async2d *a, int count) {
    return swl::guarded([&] {
        a->synchronize();
        swl::require(a->core != nullptr, "The 2D session is not initialized.");
        const int rc = swl_sim2d_advance(a->core, count);
        const std::string error = rc ? swl::last_error() : std::string();
        a->refresh();
        if (rc) throw std::runtime_error(error);
    });
}
int swl_async2d_reset(swl_async2d *a) {
    return swl::guarded([&] {
        a->synchronize();
        swl::require(a->core != nullptr, "The 2D session is not initialized.");
        if (swl_sim2d_reset(a->core)) throw std::runtime_error(swl::last_error());
        a->refresh();
        a->computeMs = a->transferMs = 0, a->batches = a->steps = 0, a->workerError.clear();
    });
}
// published state: initialized, completed, step, total, receivers, records per receiver
int swl_async2d_state(const swl_async2d *a, int64_t *out) {
    out[0] = a->initialized, out[1] = a->completed(), out[2] = a->step, out[3] = a->total;
    out[4] = static_cast<int64_t>(a->vx.size()), out[5] = a->vx.empty() ? 0 : static_cast<int64_t>(a->vx.front().size());
    return 0;
}
int swl_async2d_times(const swl_async2d *a, double *out) {
    out[0] = a->dt, out[1] = a->recordDt;
    return 0;
}
int swl_async2d_diagnostics(const swl_async2d *a, double *out) {
    std::memcpy(out, a->diag, sizeof a->diag);
    return 0;
}
int swl_async2d_receivers(const swl_async2d *a, double *vx, double *vz) {
    const size_t len = a->vx.empty() ? 0 : a->vx.front().size();
    for (size_t r = 0; r < a->vx.size(); ++r) {
        std::copy(a->vx[r].begin(), a->vx[r].end(), vx + r * len);
        std::copy(a->vz[r].begin(), a->vz[r].end(), vz + r * len);
    }
    return 0;
}
// readFrame (AsyncSimulation2D.cpp:60): the published frame of the current step, else (worker idle) read
// now; *ready = 0 while a batch runs and no current frame exists (the request is remembered for the next
// batch).  `out` holds the logical (or, with include_absorber, the computational) grid.
int swl_async2d_read_frame(swl_async2d *a, int component, float *out, int include_absorber, int *ready) {
    return swl::guarded([&] {
        swl::require(a->core != nullptr, "The 2D session is not initialized.");
        a->requested = component, a->requestedAbsorber = include_absorber;
        *ready = 0;
        if (a->frameStep == a->step && a->component == component && a->frameAbsorber == include_absorber) {
            std::copy(a->frame.begin(), a->frame.end(), out);
            *ready = 1;
            return;
        }
        if (a->worker.busy()) return;
        a->frame.resize(a->frameSize(include_absorber));
        if (swl_sim2d_read_frame(a->core, component, a->frame.data(), include_absorber)) throw std::runtime_error(swl::last_error());
        a->component = component, a->frameAbsorber = include_absorber, a->frameStep = a->step;
        std::copy(a->frame.begin(), a->frame.end(), out);
        *ready = 1;
    });
}
// operator const Simulation& (AsyncSimulation2D.cpp:62): finishes the batch, then hands out the session
// (owned by the async object; for snapshots, probes and model queries).
swl_sim2d *swl_async2d_session(swl_async2d *a) {
    try {
        a->synchronize();
    } catch (...) {
    }
    return a->core;
}
} // extern "C"
