// session2d.hpp — the 2D session object shared by the single-session C-ABI (sim2d.cpp) and the
// multi-shot batch (batch2d.cpp).  Behaviour of Seis2D::Curvilinear::Simulation
// (src/two_d/CurvilinearSolver2D.cpp:120-147, :541-604): per-step inspection, fault rule with a
// latched fault, receiver lerp resampling; the engine runs chunks of steps on the device.
#pragma once
#include "model2d.hpp"
#include "common.hpp"
#include "swl.h"
#include <algorithm>
#include <array>
#include <chrono>
#include <string>
#include <vector>

struct swl_sim2d {
    swl::Model2D model;
    swl_engine2d *engine = nullptr;
    int device = 0;
    int step = 0;
    std::string fault;
    double energy = 0, peak = 0, maxVelocity = 0, traction = 0;
    bool finite = true;
    double initializationMs = 0, totalStepMs = 0, lastStepMs = 0;
    std::vector<std::vector<double>> vx, vz; // per receiver, recorded samples
    std::vector<std::array<double, 2>> previous;
    std::vector<double> diagRows, sampleRows;
    ~swl_sim2d() {
        if (engine) swl_engine2d_destroy(engine);
    }
    double recordDt() const {
        return model.p.sample_interval_ms > 0 ? model.p.sample_interval_ms * .001 : model.dt;
    }
    void clearRecords() {
        const size_t n = model.receiverX.size();
        vx.assign(n, {}), vz.assign(n, {});
        for (size_t r = 0; r < n; ++r) {
            vx[r].reserve(model.recordReserve), vz[r].reserve(model.recordReserve);
            vx[r].push_back(0), vz[r].push_back(0);
        }
        previous.assign(n, {0, 0});
    }
    // CurvilinearSolver2D.cpp:120-147, with the raw probe values of the step just accepted.
    void sample(const double *raw) {
        const double interval = recordDt(), dt = model.dt, end = (step + 1) * dt;
        for (size_t n = vx.front().size(); n * interval <= end + 1e-9 * dt; ++n) {
            const double a = std::clamp((n * interval - step * dt) / dt, 0.0, 1.0);
            for (size_t r = 0; r < vx.size(); ++r) {
                vx[r].push_back(swl::lerpExact(previous[r][0], raw[2 * r], a));
                vz[r].push_back(swl::lerpExact(previous[r][1], raw[2 * r + 1], a));
            }
        }
        for (size_t r = 0; r < vx.size(); ++r) previous[r] = {raw[2 * r], raw[2 * r + 1]};
    }
    double scale() const { return model.amplitude / (model.p.spacing * model.p.spacing); }
    // folds the rows of an engine run (diagRows / sampleRows) into the session, as advance() does per step
    void absorb(const swl_run2d &res, double ms) {
        const size_t nrec = model.receiverX.size();
        for (int j = 0; j < res.accepted; ++j) {
            const double *d = &diagRows[4 * static_cast<size_t>(j)];
            energy = d[0], maxVelocity = d[1], traction = d[2], finite = d[3] != 0;
            sample(&sampleRows[2 * nrec * static_cast<size_t>(j)]);
            ++step;
            peak = std::max(peak, energy);
        }
        totalStepMs += ms;
        lastStepMs = res.accepted ? ms / res.accepted : ms;
        if (res.faulted) {
            fault = "2D step rejected: nonfinite fields or excessive late energy growth. Rebuild the session.";
            throw std::runtime_error(fault);
        }
    }
    // Runs of up to `chunk` steps, pipelined: run k + 1 is enqueued (continuing on the device from run k) before
    // run k's rows are read and folded, so the host's receiver resampling overlaps the device's steps.
    std::vector<double> diagRowsAlt, sampleRowsAlt;
    void advance(int count) {
        if (!fault.empty()) throw std::runtime_error(fault);
        constexpr int chunk = 512;
        const size_t nrec = model.receiverX.size();
        std::vector<int> runs;
        for (int left = std::min(count, model.total - step); left > 0; left -= chunk) runs.push_back(std::min(left, chunk));
        if (runs.empty()) return;
        const auto failed = [&] {
            fault = "Compute backend failed: " + swl::last_error() + ". Rebuild the session.";
            return std::runtime_error(fault);
        };
        const auto enqueue = [&](size_t k, int first) {
            const int slot = static_cast<int>(k & 1);
            if (swl_engine2d_run_enqueue(engine, slot, first, runs[k], model.dt, model.p.frequency, scale(), peak, k == 0))
                throw failed();
        };
        for (auto *rows : {&diagRows, &diagRowsAlt}) rows->resize(4 * static_cast<size_t>(runs[0]));
        for (auto *rows : {&sampleRows, &sampleRowsAlt}) rows->resize(2 * nrec * static_cast<size_t>(runs[0]));
        enqueue(0, step);
        auto start = std::chrono::steady_clock::now();
        for (size_t k = 0; k < runs.size(); ++k) {
            const int slot = static_cast<int>(k & 1);
            if (k + 1 < runs.size()) enqueue(k + 1, step + runs[k]);
            swl_run2d res{};
            double *d = slot ? diagRowsAlt.data() : diagRows.data(), *s = slot ? sampleRowsAlt.data() : sampleRows.data();
            if (swl_engine2d_run_collect(engine, slot, d, s, &res)) throw failed();
            if (res.accepted < runs[k] && k + 1 < runs.size()) { // the run after a faulted one accepts nothing
                swl_run2d rest{};
                if (swl_engine2d_run_collect(engine, slot ^ 1, nullptr, nullptr, &rest)) throw failed();
            }
            const auto now = std::chrono::steady_clock::now();
            const double ms = std::chrono::duration<double, std::milli>(now - start).count();
            start = now;
            if (slot) diagRows.swap(diagRowsAlt), sampleRows.swap(sampleRowsAlt); // absorb() reads diagRows / sampleRows
            absorb(res, ms);
            if (slot) diagRows.swap(diagRowsAlt), sampleRows.swap(sampleRowsAlt);
            if (res.accepted < runs[k]) break;
        }
    }
};

