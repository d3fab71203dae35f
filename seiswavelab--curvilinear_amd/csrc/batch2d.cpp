// batch2d.cpp — session-level multi-shot batch: steps many 2D sessions (Seis2D::Curvilinear::Simulation
// semantics, session2d.hpp) through one runner (swl_batch2d_run) so their work shares each launch.
// Every member keeps its own step counter, receivers, lerp resampling and latched fault; the result of
// a shot is the same whether it runs alone (swl_sim2d_advance) or in a batch.
#include "session2d.hpp"
#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>

struct swl_batch2d {
    int device = 0;
    swl_batch2d_runner *runner = nullptr;
    std::vector<swl_sim2d *> members;
    ~swl_batch2d() {
        if (runner) swl_batch2d_runner_destroy(runner);
    }
    void advance(int count) {
        constexpr int chunk = 512; // host rows per round (the runner splits further into graph chunks)
        while (count > 0) {
            std::vector<swl_sim2d *> live;
            std::vector<swl_batch_shot> shots;
            int n = 0;
            for (auto *s : members) {
                if (!s->fault.empty() || s->step >= s->model.total || !s->engine) continue;
                const int k = std::min({count, chunk, s->model.total - s->step});
                s->diagRows.resize(4 * static_cast<size_t>(k));
                s->sampleRows.resize(2 * std::max<size_t>(s->model.receiverX.size(), 1) * static_cast<size_t>(k));
                swl_batch_shot b{};
                b.engine = s->engine, b.first_step = s->step, b.count = k;
                b.dt = s->model.dt, b.frequency = s->model.p.frequency, b.scale = s->scale(), b.peak_energy = s->peak;
                b.diag = s->diagRows.data(), b.samples = s->sampleRows.data();
                live.push_back(s), shots.push_back(b);
                n = std::max(n, k);
            }
            if (live.empty()) return;
            const auto start = std::chrono::steady_clock::now();
            if (swl_batch2d_run(runner, shots.data(), static_cast<int>(shots.size())))
                throw std::runtime_error("batch run failed: " + swl::last_error());
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
            for (size_t i = 0; i < live.size(); ++i) {
                try {
                    live[i]->absorb(shots[i].result, ms);
                } catch (const std::exception &) {
                    // the session latched its fault; the batch carries on with the others
                }
            }
            count -= n;
        }
    }
};

extern "C" {
int swl_batch2d_create(int device, swl_batch2d **out) {
    return swl::guarded([&] {
        swl::require(out != nullptr, "invalid batch arguments");
        auto b = std::make_unique<swl_batch2d>();
        b->device = device;
        if (swl_batch2d_runner_create(device, &b->runner)) throw std::runtime_error(swl::last_error());
        *out = b.release();
    });
}
void swl_batch2d_destroy(swl_batch2d *b) { delete b; }
int swl_batch2d_add(swl_batch2d *b, swl_sim2d *s) {
    return swl::guarded([&] {
        swl::require(b && s, "invalid batch arguments");
        swl::require(s->engine != nullptr, "Only sessions created on the HIP backend can join a batch.");
        swl::require(s->device == b->device, "Sessions of a batch must live on the batch's device.");
        swl::require(std::find(b->members.begin(), b->members.end(), s) == b->members.end(), "Session is already in the batch.");
        b->members.push_back(s);
    });
}
int swl_batch2d_remove(swl_batch2d *b, swl_sim2d *s) {
    return swl::guarded([&] {
        auto it = std::find(b->members.begin(), b->members.end(), s);
        swl::require(it != b->members.end(), "Session is not in the batch.");
        b->members.erase(it);
    });
}
int swl_batch2d_size(const swl_batch2d *b) { return static_cast<int>(b->members.size()); }
int swl_batch2d_advance(swl_batch2d *b, int count) {
    return swl::guarded([&] { b->advance(count); });
}
int swl_batch2d_set_option(swl_batch2d *b, const char *name, int value) {
    return swl_batch2d_runner_set_option(b->runner, name, value);
}
int swl_batch2d_stats(const swl_batch2d *b, double *out) { return swl_batch2d_runner_stats(b->runner, out); }
const char *swl_sim2d_fault(const swl_sim2d *s) { return s->fault.c_str(); }
} // extern "C"
