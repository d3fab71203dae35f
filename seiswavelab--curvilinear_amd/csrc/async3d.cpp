// async3d.cpp — background stepping of a 3D session: Seis3D::Curvilinear::AsyncSimulation3D
// (include/Seis3D/AsyncSimulation3D.hpp, src/three_d/AsyncSimulation3D.cpp) over swl_sim3d.
//
// As the 2D façade (async2d.cpp): one persistent worker thread owns the session while a batch runs.  queue(steps,
// budget) submits at most 8 single steps that stop early when the time budget is spent, the session completes,
// or the caller pauses.  The batch publishes the step, the diagnostics, the Vx / Vy / Vz samples recorded since the
// batch was queued and, when the caller has asked for one, the display slices.  poll() / synchronize() fold a
// finished publication into the caller's copy; every call that touches the session itself first cancels the
// batch between steps and waits for it.
#include "background.hpp"
#include "session3d.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>

namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

constexpr int kDiag3 = 21; // swl_sim3d_diagnostics values

struct Request3D {
    int x = 0, y = 0, z = 0, component = -1;
    bool operator==(const Request3D &) const = default;
};

struct Publication3D {
    int step = 0, advanced = 0;
    double diag[kDiag3]{};
    std::vector<std::vector<double>> vx, vy, vz; // samples recorded during the batch, per receiver
    Request3D request;                           // component -1: no slices published
    std::vector<float> slices;
    double computeMs = 0, transferMs = 0;
    std::string error;
};
} // namespace

struct swl_async3d {
    swl_sim3d *core = nullptr; // owned
    swl::BackgroundWorker<Publication3D> worker;
    bool initialized = false;
    int step = 0, total = 0;
    double dt = 0;
    double diag[kDiag3]{};
    std::vector<std::vector<double>> vx, vy, vz;
    Request3D requested, shown;
    std::vector<float> slices;
    int slicesStep = -1;
    double computeMs = 0, transferMs = 0;
    uint64_t batches = 0, steps = 0;
    std::string workerError;

    ~swl_async3d() {
        try {
            synchronize();
        } catch (...) {
        }
        if (core) swl_sim3d_destroy(core);
    }
    size_t sliceSize() const {
        const int px = core->model.nx - 6, py = core->model.ny - 6, pz = core->model.nz - 6;
        return (size_t)px * pz + (size_t)py * pz + (size_t)px * py;
    }
    void refresh() { // the session's state becomes the published state
        initialized = core != nullptr && core->engine != nullptr;
        if (!initialized) return;
        step = core->step, total = core->model.total, dt = core->model.dt;
        swl_sim3d_diagnostics(core, diag);
        vx = core->vx, vy = core->vy, vz = core->vz;
        slicesStep = -1;
    }
    void apply(Publication3D &&p) {
        std::memcpy(diag, p.diag, sizeof diag);
        step = p.step;
        for (size_t r = 0; r < vx.size() && r < p.vx.size(); ++r) {
            vx[r].insert(vx[r].end(), p.vx[r].begin(), p.vx[r].end());
            vy[r].insert(vy[r].end(), p.vy[r].begin(), p.vy[r].end());
            vz[r].insert(vz[r].end(), p.vz[r].begin(), p.vz[r].end());
        }
        if (p.request.component >= 0) slices = std::move(p.slices), shown = p.request, slicesStep = step;
        computeMs += p.computeMs, transferMs += p.transferMs, ++batches, steps += p.advanced;
        workerError = std::move(p.error);
    }
    void synchronize() {
        if (auto p = worker.wait()) apply(std::move(*p));
    }
    bool completed() const { return initialized && step >= total; }
    // AsyncSimulation3D::queue (AsyncSimulation3D.cpp:49-66)
    bool queue(int count, double budgetMs) {
        if (!initialized || completed() || worker.busy()) return false;
        std::vector<size_t> counts;
        for (const auto &r : vz) counts.push_back(r.size());
        const Request3D want = requested;
        const int n = std::clamp(count, 1, 8);
        const size_t planes = sliceSize();
        swl_sim3d *s = core;
        return worker.submit([s, counts, want, n, budgetMs, planes](const std::atomic_bool &cancel) {
            Publication3D p;
            auto begin = Clock::now();
            const int before = s->step;
            for (int i = 0; i < n && !cancel && s->step < s->model.total; ++i) {
                if (swl_sim3d_advance(s, 1)) {
                    p.error = swl::last_error();
                    break;
                }
                if (since(begin) >= budgetMs) break;
            }
            p.computeMs = since(begin);
            p.advanced = s->step - before;
            begin = Clock::now();
            p.step = s->step;
            swl_sim3d_diagnostics(s, p.diag);
            p.vx.resize(s->vx.size()), p.vy.resize(s->vy.size()), p.vz.resize(s->vz.size());
            for (size_t r = 0; r < s->vz.size(); ++r) {
                const auto at = static_cast<std::ptrdiff_t>(counts[r]);
                p.vx[r].assign(s->vx[r].begin() + at, s->vx[r].end());
                p.vy[r].assign(s->vy[r].begin() + at, s->vy[r].end());
                p.vz[r].assign(s->vz[r].begin() + at, s->vz[r].end());
            }
            if (want.component >= 0) {
                p.slices.resize(planes);
                if (swl_sim3d_read_display_frame(s, want.x, want.y, want.z, want.component, p.slices.data()) == 0)
                    p.request = want;
                else if (p.error.empty())
                    p.error = swl::last_error();
            }
            p.transferMs = since(begin);
            return p;
        });
    }
};

extern "C" {
int swl_async3d_create(swl_async3d **out) {
    return swl::guarded([&] {
        swl::require(out != nullptr, "invalid async session arguments");
        *out = new swl_async3d();
    });
}
void swl_async3d_destroy(swl_async3d *a) { delete a; }
int swl_async3d_initialize(swl_async3d *a, const swl_params3d *p, int backend, int device, int precision) {
    return swl::guarded([&] {
        a->synchronize();
        swl_sim3d *s = nullptr;
        const int rc = swl_sim3d_create(p, backend, device, precision, &s);
        if (rc == 0) {
            if (a->core) swl_sim3d_destroy(a->core);
            a->core = s;
        }
        a->refresh();
        a->requested = {}, a->computeMs = a->transferMs = 0, a->batches = a->steps = 0, a->workerError.clear();
        if (rc) throw std::runtime_error(swl::last_error());
    });
}
int swl_async3d_queue(swl_async3d *a, int steps, double budget_ms) { return a->queue(steps, budget_ms) ? 1 : 0; }
int swl_async3d_poll(swl_async3d *a) {
    return swl::guarded([&] {
        if (auto p = a->worker.poll()) a->apply(std::move(*p));
        if (!a->workerError.empty()) throw std::runtime_error(std::exchange(a->workerError, {}));
    });
}
void swl_async3d_pause(swl_async3d *a) { a->worker.cancel(); }
int swl_async3d_synchronize(swl_async3d *a) {
    return swl::guarded([&] { a->synchronize(); });
}
int swl_async3d_busy(swl_async3d *a) { return a->worker.busy() ? 1 : 0; }
int swl_async3d_timing(const swl_async3d *a, double *out) {
    out[0] = a->computeMs, out[1] = a->transferMs, out[2] = static_cast<double>(a->batches), out[3] = static_cast<double>(a->steps);
    return 0;
}
int swl_async3d_advance(swl_async3d *a, int count) {
    return swl::guarded([&] {
        a->synchronize();
        swl::require(a->core != nullptr, "Initialize a model first.");
        const int rc = swl_sim3d_advance(a->core, count);
        const std::string error = rc ? swl::last_error() : std::string();
        a->refresh();
        if (rc) throw std::runtime_error(error);
    });
}
int swl_async3d_reset(swl_async3d *a) {
    return swl::guarded([&] {
        a->synchronize();
        swl::require(a->core != nullptr, "Initialize a model first.");
        if (swl_sim3d_reset(a->core)) throw std::runtime_error(swl::last_error());
        a->refresh();
        a->computeMs = a->transferMs = 0, a->batches = a->steps = 0, a->workerError.clear();
    });
}
int swl_async3d_state(const swl_async3d *a, int64_t *out) {
    out[0] = a->initialized, out[1] = a->completed(), out[2] = a->step, out[3] = a->total;
    out[4] = static_cast<int64_t>(a->vz.size()), out[5] = a->vz.empty() ? 0 : static_cast<int64_t>(a->vz.front().size());
    return 0;
}
double swl_async3d_dt(const swl_async3d *a) { return a->dt; }
int swl_async3d_diagnostics(const swl_async3d *a, double *out) {
    std::memcpy(out, a->diag, sizeof a->diag);
    return 0;
}
int swl_async3d_receivers(const swl_async3d *a, double *vx, double *vy, double *vz) {
    const size_t len = a->vz.empty() ? 0 : a->vz.front().size();
    for (size_t r = 0; r < a->vz.size(); ++r) {
        std::copy(a->vx[r].begin(), a->vx[r].end(), vx + r * len);
        std::copy(a->vy[r].begin(), a->vy[r].end(), vy + r * len);
        std::copy(a->vz[r].begin(), a->vz[r].end(), vz + r * len);
    }
    return 0;
}
int swl_async3d_read_display_frame(swl_async3d *a, int x, int y, int z, int component, float *out, int *ready) {
    return swl::guarded([&] {
        swl::require(a->core != nullptr, "Initialize a model before requesting slices.");
        const Request3D r{x, y, z, component};
        a->requested = r;
        *ready = 0;
        if (a->slicesStep == a->step && a->shown == r) {
            std::copy(a->slices.begin(), a->slices.end(), out);
            *ready = 1;
            return;
        }
        if (a->worker.busy()) return; // a pending view, never old data under the new label
        a->slices.resize(a->sliceSize());
        if (swl_sim3d_read_display_frame(a->core, x, y, z, component, a->slices.data())) throw std::runtime_error(swl::last_error());
        a->shown = r, a->slicesStep = a->step;
        std::copy(a->slices.begin(), a->slices.end(), out);
        *ready = 1;
    });
}
swl_sim3d *swl_async3d_session(swl_async3d *a) {
    try {
        a->synchronize();
    } catch (...) {
    }
    return a->core;
}
} // extern "C"
