// session3d.hpp — the 3D session object behind the swl_sim3d_* C-ABI (sim3d.cpp) and the 3D export
// routines (vts3d.cpp).  Behaviour of CpuSimulation3D (src/three_d/CurvilinearSolver.cpp:301-592): per-step
// inspection, source-window peak, fault rule with rollback and a latched fault, receiver samples every step
// with t = 0 included, while the engine runs chunks of steps on the device.
#pragma once
#include "model3d.hpp"
#include "common.hpp"
#include "swl.h"
#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

struct swl_sim3d {
    swl::Model3D model;
    swl_engine3d *engine = nullptr;
    int device = 0, precision = 0, step = 0;
    std::string fault;
    double energy = 0, peak = 0, sourcePeak = 0, maxVelocity = 0, traction = 0, display[5] = {0, 0, 0, 0, 0};
    bool finite = true;
    double initializationMs = 0, totalStepMs = 0;
    std::vector<std::vector<double>> vx, vy, vz;
    std::vector<double> diagRows, sampleRows;
    ~swl_sim3d() {
        if (engine) swl_engine3d_destroy(engine);
    }
    void clearRecords() {
        const size_t n = model.receiverIndex.size();
        vx.assign(n, {0.0}), vy.assign(n, {0.0}), vz.assign(n, {0.0});
        for (size_t r = 0; r < n; ++r) {
            vx[r].reserve(model.total + 1), vy[r].reserve(model.total + 1), vz[r].reserve(model.total + 1);
        }
    }
    void apply(const double *d) {
        energy = d[0], maxVelocity = d[1], traction = d[2], finite = d[3] != 0;
        for (int c = 0; c < 5; ++c) display[c] = d[4 + c];
        peak = std::max(peak, energy);
    }
    // CpuSimulation3D::advance (CurvilinearSolver.cpp:445-513)
    void advance(int count) {
        if (count < 0 || count > 20000) throw std::runtime_error("Invalid step batch.");
        if (!fault.empty()) throw std::runtime_error(fault);
        if (!finite) throw std::runtime_error("Fields are non-finite. Rebuild the session.");
        constexpr int chunk = 256;
        const size_t nrec = model.receiverIndex.size();
        const double scale = model.amplitude / (model.p.density * model.h * model.h * model.h);
        int remaining = std::min(count, model.total - step);
        while (remaining > 0) {
            const int n = std::min(remaining, chunk);
            diagRows.resize(10 * static_cast<size_t>(n));
            sampleRows.resize(3 * std::max<size_t>(nrec, 1) * static_cast<size_t>(n));
            swl_run3d res{};
            const auto start = std::chrono::steady_clock::now();
            if (swl_engine3d_run(engine, step, n, model.dt, model.p.dominant_frequency, scale, sourcePeak, diagRows.data(),
                                 sampleRows.data(), &res)) {
                fault = "Compute backend failed: " + swl::last_error() + ". Rebuild the session.";
                throw std::runtime_error(fault);
            }
            totalStepMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
            for (int j = 0; j < res.accepted; ++j) {
                apply(&diagRows[10 * static_cast<size_t>(j)]);
                const double *smp = &sampleRows[3 * nrec * static_cast<size_t>(j)];
                for (size_t r = 0; r < nrec; ++r) {
                    vx[r].push_back(smp[3 * r]), vy[r].push_back(smp[3 * r + 1]), vz[r].push_back(smp[3 * r + 2]);
                }
                ++step;
            }
            sourcePeak = res.source_peak;
            if (res.faulted) {
                fault = res.nonfinite ? "Non-finite RK4 step rejected. Last accepted fields retained; rebuild with smaller CFL."
                                      : "Energy grew above 10,000 times the source-window peak. Last accepted step retained; "
                                        "rebuild with smaller CFL or gentler terrain.";
                throw std::runtime_error(fault);
            }
            remaining -= n;
        }
        if (!finite) throw std::runtime_error("Simulation produced non-finite fields; stop and reduce the time/grid distortion.");
    }
};
