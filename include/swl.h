/*
 * swl.h — C-ABI of the MI355X-native curvilinear velocity-stress solver (libswl.so).
 *
 * Every entry point returns an int status (0 = ok, nonzero = error) unless it
 * returns a value; on error the message is available from swl_last_error()
 * (thread-local).  No C++ types or exceptions cross this boundary; all arrays are
 * plain host pointers with explicit sizes.  The engine owns every device
 * allocation; host arrays passed in are copied during the call.
 *
 * Three layers are exported:
 *  1. Engines (swl_engine*): the HIP backends that replace the reference's
 *     CudaEngine (src/two_d/CudaEngine2D.hpp:5-24) and CudaEngine3D
 *     (src/three_d/CudaEngine3D.hpp:7-24).  They take an already-built FP64 mesh.
 *  2. Sessions (swl_sim2d* / swl_sim3d*): the 2D Simulation
 *     (include/Seis2D/CurvilinearSolver2D.hpp:93-123) and 3D ISimulationBackend3D
 *     (include/Seis3D/CurvilinearSolver.hpp:165-185) façades in C++20, which build
 *     geometry, materials, absorbers, sources, the dt plan and receivers exactly as
 *     the reference does and drive an engine.
 *  3. Batches (swl_batch2d*): many independent 2D shots stepped by one launch per
 *     RK stage (no counterpart in the reference; used by the serving loop).
 */
#ifndef SWL_H
#define SWL_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SWL_MAX_LAYERS_2D 16
#define SWL_MAX_LAYERS_3D 8

/* Mirrors Seis2D::Curvilinear::Parameters (include/Seis2D/CurvilinearSolver2D.hpp:29-57).
   node_vp/node_vs/node_density: optional heterogeneous model on the logical nx*nz grid
   (X fastest), an extension the reference 2D API cannot express. */
typedef struct swl_params2d {
    int precision; /* 0 = FP64, 1 = FP32 (Seis2D::Curvilinear::Precision) */
    int nx, nz;
    double spacing, vp, vs, density;
    double relief, wavelength, flatten_depth, frequency, duration, cfl, force_per_length;
    int source_x, source_z, max_steps;
    int source_kind; /* SeisWave::SourceKind: 0 directional force, 1 isotropic moment */
    double source_direction[3];
    double moment_per_length;
    double source_sigma_m, source_radius_m;
    int absorbing;
    int sponge_x, sponge_z;
    double sponge_strength;
    int absorber; /* Seis2D::Curvilinear::Absorber: 0 sponge, 1 CFS-PML */
    double pml_reflection, pml_kappa_max, pml_frequency;
    int external_absorption;
    int receiver_start, receiver_count, receiver_stride, receiver_depth;
    double sample_interval_ms;
    int layered, n_layers;
    double layer_thickness[SWL_MAX_LAYERS_2D], layer_vp[SWL_MAX_LAYERS_2D], layer_vs[SWL_MAX_LAYERS_2D],
        layer_density[SWL_MAX_LAYERS_2D], layer_qp[SWL_MAX_LAYERS_2D], layer_qs[SWL_MAX_LAYERS_2D];
    int attenuation, mechanisms;
    double q_min_frequency, q_max_frequency, q_reference_frequency, qp, qs;
    int imported_surface;
    double surface_datum;
    int n_surface;
    const double *surface_xz;
    const double *node_vp, *node_vs, *node_density;
    int threads;
    /* receiver policy of the interactive survey (Parameters::autoSurfaceReceivers / receiverTargetCount,
       CurvilinearSolver2D.hpp:45-46): with auto_surface_receivers the session lays the line out itself
       (synchronizeSurfaceReceivers, include/Seis2D/SurfaceSurvey2D.hpp:27-44) and reports it back through
       swl_sim2d_parameters */
    int auto_surface_receivers, receiver_target_count;
} swl_params2d;

/* Mirrors Seis3D::Curvilinear::SimulationParameters (include/Seis3D/CurvilinearSolver.hpp:70-95). */
typedef struct swl_params3d {
    int nx, ny, nz;
    double spacing, vp, vs, density;
    int layered, layer_depth_mode, n_layers;
    double layer_thickness[SWL_MAX_LAYERS_3D], layer_vp[SWL_MAX_LAYERS_3D], layer_vs[SWL_MAX_LAYERS_3D],
        layer_density[SWL_MAX_LAYERS_3D];
    double dominant_frequency;
    int source_x, source_y, source_z;
    int source_kind;
    double source_direction[3];
    double source_force, source_moment;
    double duration;
    int max_steps;
    double cfl;
    double topography_amplitude, topography_wavelength, terrain_transition_depth;
    int dem_nx, dem_ny;
    double dem_spacing, dem_origin_x, dem_origin_y, dem_datum;
    const double *dem_elevations;
    int receiver_kind, receiver_start_x, receiver_start_y, receiver_z;
    int receiver_count_x, receiver_count_y, receiver_stride_x, receiver_stride_y;
    int absorbing, absorbing_cells[3];
    double absorbing_strength;
    int threads;
} swl_params3d;

/* ---------------- errors and devices ---------------- */
/* Message of the last failed call on this thread ("" if none).  Replaces the
   std::runtime_error("CUDA <action>: <msg>") convention of src/common/CudaSupport.cuh:6-9. */
const char *swl_last_error(void);

typedef struct swl_device_info {
    int index;
    char name[128];
    char arch[64];
    uint64_t total_bytes, free_bytes;
    int compute_units;
    int available;
} swl_device_info;
/* queryCudaDevices3D / queryCudaMemory3D (src/common/CudaDevices.cu:10-61). */
int swl_device_count(void);
int swl_device_query(int device, swl_device_info *out);

/* ---------------- 2D engine (replaces CudaEngine, src/two_d/CudaEngine2D.hpp:5-24) ---------------- */
/* FP64 host mesh in padded storage (Point::MeshT, src/two_d/PointKernels2D.hpp:14-27).
   The engine converts to the run precision; in FP32 it keeps FP64 copies of the
   surface row for the traction closure and projection. */
typedef struct swl_mesh2d_desc {
    int precision;   /* 0 FP64, 1 FP32 */
    int nx, nz;      /* padded node counts (computational + 2*3 ghosts) */
    double h;
    const double *a, *b, *jac, *normals, *lambda, *mu, *rho, *source, *damping; /* normals: 2 per node */
    double source_direction[3];
    const double *source_gradient[2]; /* both null for a directional force */
    int mechanisms, field_count;
    const double *cpml_x, *cpml_z; /* 3 per computational node along x / z, or null */
    int cpml_width_x, cpml_width_z;
    int64_t cpml_count;
    double rates[3];
    const double *delta_p[3], *delta_s[3];
    /* Optional separable terrain description (CurvilinearSolver2D.cpp:344-354): surface depth
       H(x_i) and slope H'(x_i) per padded column, fade F(eta_k) and F'(eta_k) per padded row.
       When a = -(H' F)/(1 + H F'), b = 1/(1 + H F') and J = 1 + H F' reproduce every stored
       value bit for bit, the engine recomputes them in FP64 instead of streaming the arrays. */
    const double *surface_depth, *surface_slope, *fade, *fade_derivative;
} swl_mesh2d_desc;

typedef struct swl_diag2d {
    double energy, max_velocity, traction_relative;
    int finite;
    int rhs_registers;
    uint64_t rhs_local_bytes;
} swl_diag2d;

typedef struct swl_engine2d swl_engine2d;

/* CudaEngine(const MeshT&, const std::vector<Receiver>&, int device), CudaEngine2D.cu:127-142.
   receiver_index: padded storage index of each probe (Point::index(r.x+3, r.z+3)). */
int swl_engine2d_create(const swl_mesh2d_desc *mesh, const int64_t *receiver_index, int receivers, int device,
                        swl_engine2d **out);
void swl_engine2d_destroy(swl_engine2d *e);
/* CudaEngine::step(time, dt, frequency, scale, stepNumber), CudaEngine2D.cu:143-152. */
int swl_engine2d_step(swl_engine2d *e, double time, double dt, double frequency, double scale, int step);
/* CudaEngine::rollback / reset, CudaEngine2D.cu:153-154. */
int swl_engine2d_rollback(swl_engine2d *e);
int swl_engine2d_reset(swl_engine2d *e);
/* CudaEngine::diagnostics, CudaEngine2D.cu:155-158. */
int swl_engine2d_diagnostics(swl_engine2d *e, swl_diag2d *out);
/* No counterpart in the reference.  swl_engine2d_diagnostics for callers that cannot allocate device memory
   (code between HIP graph captures, tight serving loops, hosts that forbid hipMalloc): the same values, with
   the per-tile partials of the inspection kernel in a fixed 4 MiB device array (one per GPU, held until the
   process ends) instead of a per-call allocation.  Grids of more than 131072 tiles of 64x4 computational
   nodes (about 33.5 M nodes) are rejected with an error.  Calls on one GPU share that array and take turns
   (a per-device lock in this process); the call waits for the engine's stream. */
int swl_engine2d_diagnostics_noalloc(swl_engine2d *e, swl_diag2d *out);
/* CudaEngine::sample, CudaEngine2D.cu:159-162: out[2r] = Vx, out[2r+1] = Vz. */
int swl_engine2d_sample(swl_engine2d *e, double *out);
/* CudaEngine::frame, CudaEngine2D.cu:163-165: component 0 Vz, 1 Vx, 2 |V|, 3 pressure. */
int swl_engine2d_frame(swl_engine2d *e, int component, float *out, int nx, int nz, int offset_x);
/* CudaEngine::snapshot, CudaEngine2D.cu:174-178: field f (0..field_count-1) of the accepted
   state in padded storage promoted to double; swl_engine2d_pml: compact CFS-PML memories. */
int swl_engine2d_snapshot(swl_engine2d *e, int field, double *out);
int swl_engine2d_pml(swl_engine2d *e, double *out);
/* CudaEngine::probes, CudaEngine2D.cu:166-173: out[5n+f] for f = Vx,Vz,Sxx,Szz,Sxz. */
int swl_engine2d_probes(swl_engine2d *e, const int64_t *indices, int count, double *out);
/* CudaEngine::residentBytes, CudaEngine2D.cu:179-183. */
uint64_t swl_engine2d_resident_bytes(swl_engine2d *e);

/* Device-resident stepping (no reference counterpart; same semantics as the reference's
   advance() loop, CurvilinearSolver2D.cpp:553-597, with no per-step host synchronisation):
   runs up to `count` steps from `first_step`, each followed on the device by the energy /
   finiteness inspection, the fault rule
       !finite || ((s+1)*dt > 3/frequency && peak > 0 && E > 1e4*peak)
   and the receiver gather.  `peak` is the accepted peak energy before the call.
   On a rejected step the state stays at the last accepted step and later steps do nothing.
   Outputs: accepted (steps accepted), faulted (0/1), per accepted step s:
   diag[4s..4s+3] = energy, max_velocity, traction_relative, finite and
   samples[(2*receivers)*s + 2r + {0,1}] = Vx, Vz of receiver r.  Arrays hold `count` rows. */
typedef struct swl_run2d {
    int accepted, faulted;
    double peak_energy;
} swl_run2d;
int swl_engine2d_run(swl_engine2d *e, int first_step, int count, double dt, double frequency, double scale,
                     double peak_energy, double *diag, double *samples, swl_run2d *result);
/* Pipelined form of swl_engine2d_run (no reference counterpart; the same steps as the reference's advance()
   loop, CurvilinearSolver2D.cpp:553-597): swl_engine2d_run_enqueue returns once `count` steps are enqueued
   with their rows in `slot` (0 or 1); swl_engine2d_run_collect waits for that run and copies its rows, as
   swl_engine2d_run returns them.  A caller enqueues the next run before it collects this one, so its host work
   on the rows overlaps the device's steps.  fresh = 1 starts from `peak_energy` with the fault latch cleared;
   fresh = 0 continues the run before it (same slot pairing): the device's peak and fault latch carry over, so a
   run enqueued after a faulted one accepts no step.  A slot is collected once per enqueue. */
int swl_engine2d_run_enqueue(swl_engine2d *e, int slot, int first_step, int count, double dt, double frequency,
                             double scale, double peak_energy, int fresh);
int swl_engine2d_run_collect(swl_engine2d *e, int slot, double *diag, double *samples, swl_run2d *result);
/* Timing with HIP events on the engine's stream during swl_engine2d_run (no reference
   counterpart; the reference times whole steps on the host).  One event pair brackets the four
   back-to-back stage launches of a step; the finalisation is timed on every 32nd step.
   enable resets the counters. */
int swl_engine2d_profile(swl_engine2d *e, int enable);
/* out[14]: stage-kernel ms, stage launches, finalisation ms, finalisation launches, interior tiles,
   stage-kernel VGPRs, computational nodes per stage launch, border tiles (general blocks), on-the-fly metrics
   (0/1), per-row material tables (1: lambda, mu, rho, dP, dS depend on the row only; 0: per-node arrays),
   PML vector lanes (two x-adjacent nodes each), PML blocks, small-grid elastic kernel in use (0/1), LDS tiles in
   use (0/1). */
int swl_engine2d_kernel_timing(swl_engine2d *e, double *out);
/* Engine tuning switches: "nontemporal_stores" (default 1) for the interior outputs;
   "stored_metrics" (default 0) reads a, b, J from memory instead of recomputing them;
   "lds_tiles" (-1 automatic: elastic launches of at least 1536 blocks; 0 off; 1 on) loads each field's tile of
   the elastic interior's stencil input straight into LDS (global_load_lds), one field ahead of the arithmetic
   (identical results);
   "material_rows" (default 1 where the materials depend on the row only) reads the per-row tables,
   0 reads the per-node arrays;
   "pml_vector" (-1 automatic: grids of at least 1024 interior tiles; 0 off; 1 on): FP32 elastic CFS-PML models
   update the absorbing strips below the surface band (depth >= 4, no sponge) as lanes of two x-adjacent nodes
   instead of one node per lane (same equations, rounded differently: FP32 tolerance, not bit-identical).
   "latency_kernel" (-1 automatic: FP32 elastic launches of fewer than 1536 blocks; 0 off; 1 on): a variant of the
   stage kernel that keeps several fields' stencil loads in flight at a 3-wave register budget, for grids too small
   to fill the device (identical results).
   "end_in_stage" (default 1): step j is ended (reduction, fault rule, diagnostics and receiver rows) by block 0
   of step j + 1's first stage launch; 0 launches a reduction and a finalisation kernel after every step.
   "split_border" (-1 automatic: large FP64 grids; 0 off; 1 on): the border blocks run as their own launch on the
   engine's second stream beside the interior tiles.  "spread_border" (default 0): the border blocks are spread
   evenly over the one launch instead of dispatched first.  These three change the schedule only. */
int swl_engine2d_set_option(swl_engine2d *e, const char *name, int value);
/* Measurement utility (no reference counterpart): `repeats` plain streaming copies of `bytes`
   (read `bytes`, write `bytes`) on `device`; ms_per_copy = mean HIP-event time per copy.  Used to
   calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE and as the sustained-copy bandwidth reference. */
int swl_stream_probe(int device, uint64_t bytes, int repeats, double *ms_per_copy);

/* ---------------- 2D session (Seis2D::Curvilinear::Simulation) ---------------- */
typedef struct swl_sim2d swl_sim2d;
/* backend: 0 CPU is not provided by this library (the CPU path is the reference's own);
   2 = HIP.  Returns an error for unknown backends instead of falling back. */
int swl_sim2d_create(const swl_params2d *p, int backend, int device, swl_sim2d **out);
/* Builds the FP64 session model (validation, geometry, materials, absorbers, sources, dt plan,
   receivers) without a device: model/diagnostic/layout queries work, stepping is refused. */
int swl_sim2d_plan(const swl_params2d *p, swl_sim2d **out);
void swl_sim2d_destroy(swl_sim2d *s);
int swl_sim2d_advance(swl_sim2d *s, int count);
int swl_sim2d_reset(swl_sim2d *s);
int swl_sim2d_step(const swl_sim2d *s);
/* Simulation::parameters() (CurvilinearSolver2D.hpp:106): the session's logical parameters, with the automatic
   surface receiver layout applied (CurvilinearSolver2D.cpp:164-168); surface_xz points into the session, the
   per-node arrays are not echoed (null). */
int swl_sim2d_parameters(const swl_sim2d *s, swl_params2d *out);
/* Replaces the accepted peak energy that the late-growth fault rule compares against
   (!finite || (t > 3/f && peak > 0 && E > 1e4*peak), CurvilinearSolver2D.cpp:576-578, where the
   reference keeps it private).  Lets a caller restore a session's energy history, and tests drive the
   growth branch of the rule from the host; the next advance (single session or batch) uses it. */
int swl_sim2d_override_peak_energy(swl_sim2d *s, double peak);
int swl_sim2d_total_steps(const swl_sim2d *s);
double swl_sim2d_dt(const swl_sim2d *s);
double swl_sim2d_record_dt(const swl_sim2d *s);
/* layout: nx_pad, nz_pad, offset_x, cpml_width_x, cpml_width_z, field_count, cpml_count,
   receivers, record_length, mechanisms */
int swl_sim2d_layout(const swl_sim2d *s, int64_t *out);
/* diagnostics in the order energy, peak, max_velocity, traction, finite, wave_dt, damping_dt,
   source_dt, cpml_dt, relaxation_dt, ppw, max_q_error, max_metric, min_jacobian, max_edge,
   flatten_depth, source_weight_sum, source_rms_radius, initialization_ms, total_step_ms,
   last_step_ms, resident_device_bytes */
int swl_sim2d_diagnostics(const swl_sim2d *s, double *out);
/* receiver records (record_length samples each, t = 0 included) */
int swl_sim2d_receivers(const swl_sim2d *s, double *vx, double *vz, double *position, int32_t *xz);
int swl_sim2d_probe_indices(const swl_sim2d *s, int64_t *out);
int swl_sim2d_snapshot(const swl_sim2d *s, int field, double *out);
int swl_sim2d_pml(const swl_sim2d *s, double *out);
int swl_sim2d_profile(swl_sim2d *s, int enable);
int swl_sim2d_kernel_timing(const swl_sim2d *s, double *out);
int swl_sim2d_set_option(swl_sim2d *s, const char *name, int value);
/* FP64 construction arrays, same numbering as the oracle (0 a .. 21 deltaS[2]). */
int swl_sim2d_model(const swl_sim2d *s, int which, double *out);
int swl_sim2d_read_frame(const swl_sim2d *s, int component, float *out, int include_absorber);
int swl_sim2d_read_probes(const swl_sim2d *s, const int32_t *xz, int count, double *out);
/* writeReceivers (CurvilinearSolver2D.cpp:811-854): receiver CSV with provenance header. */
int swl_sim2d_write_receivers(const swl_sim2d *s, const char *path);
int64_t swl_pml_slot2d(int nx_pad, int nz_pad, int width_x, int width_z, int64_t q, int axis);
/* Detail::memoryBudget (src/two_d/MemoryBudget2D.hpp:12-38): out[0] = host bytes, out[1] = device bytes a session
   adds, for backend 0 (the reference's CPU path) and 1 (its CUDA engine) as the reference computes them, and for
   2 (this library's HIP engine: host bytes as for CUDA, device bytes an upper bound of the engine's allocations).
   swl_sim2d_create refuses a session whose HIP budget exceeds 80 % of the free host memory or 85 % of the
   device's free memory, with the reference's message (CurvilinearSolver2D.cpp:247-266). */
int swl_memory_budget2d(const swl_params2d *p, int backend, double *out);
size_t swl_params2d_size(void);
size_t swl_params3d_size(void);

/* ---------------- 2D multi-shot batch (no counterpart in the reference) ----------------
   The reference steps one shot per session (CudaEngine2D.cu:143-152).  A runner steps many 2D engines
   together in one of two modes.  Lanes (default, option "fused" 0): every shot keeps its own specialised
   stage kernels; shots are placed longest-first on the least-loaded of `lanes` streams and enqueued
   `lane_steps` steps at a time round-robin, so small shots' launches overlap on the device.  Fused
   ("fused" 1): per RK stage one launch covers the tiles of every shot that shares a precision, a GSLS
   mechanism count and the stencil direction pair; chunks of steps are captured into hipGraphs and
   replayed; an autotuner picks the persistent grid size.  Either way each shot keeps the semantics of
   swl_engine2d_run (device-side fault rule, receiver gather, its own dt / frequency / step number) and
   its results are bit-identical to running it alone. */
typedef struct swl_batch_shot {
    swl_engine2d *engine;
    int first_step, count;                        /* steps first_step .. first_step + count - 1 */
    double dt, frequency, scale, peak_energy;     /* as swl_engine2d_run */
    double *diag, *samples;                       /* host rows, as swl_engine2d_run (count rows) */
    swl_run2d result;                             /* out */
} swl_batch_shot;
typedef struct swl_batch2d_runner swl_batch2d_runner;
int swl_batch2d_runner_create(int device, swl_batch2d_runner **out);
void swl_batch2d_runner_destroy(swl_batch2d_runner *r);
/* options: "fused" (default 0 = lanes), "lanes" (3 streams), "lane_steps" (32), "lane_depth" (2 turns in flight
   per lane); fused mode: "graphs" (1),
   "autotune" (1), "grid" (0 = tuned / one block per tile), "chunk" (32 steps) */
int swl_batch2d_runner_set_option(swl_batch2d_runner *r, const char *name, int value);
int swl_batch2d_run(swl_batch2d_runner *r, swl_batch_shot *shots, int n);
/* out[14]: chosen grid (-1 untuned, 0 = one block per tile), graph captures, graph replays, direct chunks,
   chunks, autotuner candidate ms x4, chunk length, failed captures (then launched directly), fused (0/1),
   lanes, lane runs */
int swl_batch2d_runner_stats(const swl_batch2d_runner *r, double *out);

/* Session-level batch over swl_sim2d sessions created on the runner's device (not owned): advance steps
   every member up to `count` steps (each stops at its own total).  A member's fault latches on that
   session only (swl_sim2d_fault) and the others continue. */
typedef struct swl_batch2d swl_batch2d;
int swl_batch2d_create(int device, swl_batch2d **out);
void swl_batch2d_destroy(swl_batch2d *b);
int swl_batch2d_add(swl_batch2d *b, swl_sim2d *session);
int swl_batch2d_remove(swl_batch2d *b, swl_sim2d *session);
int swl_batch2d_size(const swl_batch2d *b);
int swl_batch2d_advance(swl_batch2d *b, int count);
int swl_batch2d_set_option(swl_batch2d *b, const char *name, int value);
int swl_batch2d_stats(const swl_batch2d *b, double *out);
/* "" when the session has no latched fault */
const char *swl_sim2d_fault(const swl_sim2d *s);

/* ---------------- Background stepping (Seis2D::Curvilinear::AsyncSimulation, include/Seis2D/AsyncSimulation2D.hpp,
   src/two_d/AsyncSimulation2D.cpp) ----------------
   One persistent worker per object.  queue(steps, budget_ms) runs at most 8 single steps (stopping at the time
   budget, at completion or on pause) and publishes step, diagnostics, the receiver samples recorded since the
   batch was queued and, if requested, a display frame; poll / synchronize fold the publication into the state
   that the accessors below return.  synchronize, advance, reset, initialize and session() cancel a running
   batch between steps and wait for it. */
typedef struct swl_async2d swl_async2d;
int swl_async2d_create(swl_async2d **out);
void swl_async2d_destroy(swl_async2d *a);
int swl_async2d_initialize(swl_async2d *a, const swl_params2d *p, int backend, int device);
/* returns 1 when a batch was queued, 0 when not initialized, completed or busy (AsyncSimulation::queue) */
int swl_async2d_queue(swl_async2d *a, int steps, double budget_ms);
int swl_async2d_poll(swl_async2d *a);          /* non-zero: the last batch's error (swl_last_error) */
void swl_async2d_pause(swl_async2d *a);
int swl_async2d_synchronize(swl_async2d *a);
int swl_async2d_busy(swl_async2d *a);
/* BackgroundTiming: compute ms, transfer ms, batches, steps */
int swl_async2d_timing(const swl_async2d *a, double *out);
int swl_async2d_advance(swl_async2d *a, int count);
int swl_async2d_reset(swl_async2d *a);
/* published state: initialized, completed, step, total steps, receivers, samples per receiver */
int swl_async2d_state(const swl_async2d *a, int64_t *out);
int swl_async2d_times(const swl_async2d *a, double *out); /* dt, recordDt */
int swl_async2d_diagnostics(const swl_async2d *a, double *out); /* as swl_sim2d_diagnostics */
int swl_async2d_receivers(const swl_async2d *a, double *vx, double *vz);
int swl_async2d_read_frame(swl_async2d *a, int component, float *out, int include_absorber, int *ready);
swl_sim2d *swl_async2d_session(swl_async2d *a);

/* ---------------- Receiver records as SEG-Y (SeisWave::writeSegy / readSegy, include/SeisCommon/SegyIo.hpp:5-19,
   src/common/SegyIo.cpp) ----------------
   Byte-compatible with the reference: SEG-Y rev 2 (rev 1.0 for the compatible profiles), big endian, one
   component per file, 3200-byte textual header (CP037 for the IBM profile), 400-byte binary header with the
   exact IEEE64 interval, 240-byte trace headers, coordinates in millimetres (scalar -1000). */
typedef struct swl_segy_trace {
    double position[3], source[3]; /* metres; z positive down */
    double start_time;             /* seconds, a whole number of milliseconds */
    const double *samples;         /* `count` samples of the selected component */
} swl_segy_trace;
/* profile: 0 ExactIeee32, 1 ExactIeee64, 2 LegacyIbm32, 3 CompatibleIeee32 (SegyExportProfile);
   gain: SegyExportOptions::sampleGain; sample_interval: 0 keeps dt, > 0 resamples linearly (seconds).
   Written to `path`.partial and renamed, so a reader sees the old file or the complete new one. */
int swl_segy_write(const char *path, const swl_segy_trace *traces, int ntraces, int64_t count, double dt,
                   const char *component, const char *units, int profile, double gain, double sample_interval);
/* readSegy: rev 0/1/2, formats 1 (IBM), 2, 3, 5, 6, 8, either byte order, with the reference's limits. */
typedef struct swl_segy_records swl_segy_records;
int swl_segy_read(const char *path, swl_segy_records **out);
void swl_segy_records_destroy(swl_segy_records *r);
int swl_segy_records_info(const swl_segy_records *r, int *ntraces, double *dt, char *component, char *units,
                          char *description, int text_capacity);
int swl_segy_records_trace(const swl_segy_records *r, int index, double *position, double *source,
                           double *start_time, int64_t *count);
int swl_segy_records_samples(const swl_segy_records *r, int index, double *out);
/* A 2D session's receivers, component 0 Vx / 1 Vz, as src/two_d/Experiment2D.cpp:44-71 assembles them: trace
   positions (x, 0, z), the source at the coordinates of its node, dt = recordDt. */
int swl_sim2d_write_segy(const swl_sim2d *s, const char *path, int component, int profile, double gain,
                         double sample_interval);

/* ---------------- 2D model documents and wavefield snapshots as ASCII VTS ----------------
   Model documents are byte-compatible with the reference's (versioned FieldData parameter record, six point
   arrays, coordinates, FNV-1a 64-bit checksum, include/SeisCommon/VtsModelIo.hpp:18-86).  A document is a
   model, not a wavefield checkpoint: a session built from it starts at t = 0. */
/* writeModelVts (include/Seis2D/ModelIo2D.hpp:8-10, src/two_d/ModelIo2D.cpp:75-144).  cpu_threads and
   cuda_device are the ComputeSettings3D recorded in the document; BackendKind is 1 for a HIP session (the
   accelerator slot of BackendKind3D) and 0 for a device-free plan.  Written to `path`.tmp-<ticks> and renamed,
   so an existing model survives a failed write. */
int swl_sim2d_write_model_vts(const swl_sim2d *s, const char *path, int cpu_threads, int cuda_device);
/* readModelVts (ModelIo2D.hpp:11-13, ModelIo2D.cpp:146-254): validates shape, finite data, schema 1..7 and
   the checksum; third-party VTS and wavefield snapshots are rejected.  On failure *out is untouched. */
typedef struct swl_model_doc2d swl_model_doc2d;
int swl_model_vts_read(const char *path, swl_model_doc2d **out);
void swl_model_doc2d_destroy(swl_model_doc2d *d);
/* ModelDocument::parameters; surface_xz points into the document (valid until it is destroyed). */
int swl_model_doc2d_params(const swl_model_doc2d *d, swl_params2d *out);
/* out[5]: backend kind, compute.cpuThreads, compute.cudaDevice, referenceDt, schema version. */
int swl_model_doc2d_info(const swl_model_doc2d *d, double *out);
/* writeSnapshot (CurvilinearSolver2D.hpp:124-139, CurvilinearSolver2D.cpp:725-810).  components[7]: Vx, Vz,
   velocity magnitude, pressure, Sxx, Szz, Sxz (SnapshotExportOptions::components); include_material adds Vp,
   Vs, Density (and Qp, Qs when attenuating).  An empty selection fails before the file is touched. */
int swl_sim2d_write_snapshot(const swl_sim2d *s, const char *path, const int32_t *components, int include_material);

/* ---------------- 3D engine (replaces CudaEngine3D, src/three_d/CudaEngine3D.hpp:7-24) ---------------- */
/* FP64 host mesh in padded storage q = i + nx*(j + ny*k) (Point::Mesh, src/three_d/PointKernels3D.hpp:13-21).
   Material arrays are per node (the homogeneous case repeats the constants).  The optional
   separable forms let the engine stream less; each is checked against the full arrays when the
   engine is built and ignored if any value differs. */
typedef struct swl_mesh3d_desc {
    int precision;                   /* 0 FP64, 1 FP32 */
    int nx, ny, nz;                  /* padded node counts (physical + 2*3) */
    double h;
    const double *metrics;           /* 9 per node: grad(xi), grad(eta), grad(zeta) */
    const double *jacobian, *normals; /* normals: 3 per node, set on the surface plane k = 3 */
    const double *lambda, *mu, *density;
    double density_global;           /* SimulationParameters::density (source scaling, PointKernels3D.hpp:107) */
    const double *damping, *source;
    const double *source_gradient[3]; /* all null for a directional force */
    double source_direction[3];
    int flat_from;                   /* metrics are the identity and J = 1 for k >= flat_from */
    const double *damping_axis[3];   /* damping = (X(i) + Y(j)) + Z(k), or null */
    int material_by_depth;           /* lambda, mu, density depend on k only */
} swl_mesh3d_desc;
typedef struct swl_engine3d swl_engine3d;
/* CudaEngine3D(const Point::Mesh&, receivers, device), CudaEngine3D.cu:80-103 */
int swl_engine3d_create(const swl_mesh3d_desc *mesh, const int64_t *receiver_index, int receivers, int device,
                        swl_engine3d **out);
void swl_engine3d_destroy(swl_engine3d *e);
/* CudaEngine3D::step(time, dt, frequency, step), CudaEngine3D.cu:113-124; scale = amplitude/(rho h^3) */
int swl_engine3d_step(swl_engine3d *e, double time, double dt, double frequency, double scale, int step);
/* rollback / resetFields, CudaEngine3D.cu:125-126 */
int swl_engine3d_rollback(swl_engine3d *e);
int swl_engine3d_reset(swl_engine3d *e);
/* CudaEngine3D::diagnostics, CudaEngine3D.cu:127-133: out[10] = energy, max |V|, traction-rate ratio,
   finite, display peaks |Vz|, |Vx|, |Vy|, |V|, |mean stress|, 0 */
int swl_engine3d_diagnostics(swl_engine3d *e, double *out);
/* CudaEngine3D::sample, CudaEngine3D.cu:134-138: out[3r..3r+2] = Vx, Vy, Vz */
int swl_engine3d_sample(swl_engine3d *e, double *out);
/* CudaEngine3D::slices, CudaEngine3D.cu:139-147: planes XZ at y, YZ at x, XY at z of `component`
   (0 Vz, 1 Vx, 2 Vy, 3 |V|, 4 pressure); out holds nx*nz + ny*nz + nx*ny floats (physical sizes). */
int swl_engine3d_slices(swl_engine3d *e, int x, int y, int z, int component, float *out);
/* CudaEngine3D::download, CudaEngine3D.cu:148-150: field f (0..8) of the accepted state, padded storage. */
int swl_engine3d_download(swl_engine3d *e, int field, double *out);
uint64_t swl_engine3d_resident_bytes(swl_engine3d *e);
/* Device-resident stepping (no reference counterpart; same semantics as CpuSimulation3D::advance,
   CurvilinearSolver.cpp:445-513): up to `count` steps from `first_step`; after each the inspection,
   the source-window peak update for (s+1)*dt <= 3/f, the fault rule
       !finite || ((s+1)*dt > 3/f && source_peak > 0 && E > 1e4*source_peak)
   and the receiver gather.  Per accepted step s: diag[10s..10s+9] as swl_engine3d_diagnostics,
   samples[3*receivers*s + 3r + {0,1,2}] = Vx, Vy, Vz. */
typedef struct swl_run3d {
    int accepted, faulted, nonfinite; /* nonfinite: the rejected step failed the finiteness check */
    double peak_energy, source_peak;
} swl_run3d;
int swl_engine3d_run(swl_engine3d *e, int first_step, int count, double dt, double frequency, double scale,
                     double source_peak, double *diag, double *samples, swl_run3d *result);
/* HIP-event timing of the stage kernels (one event pair per step), as for 2D. out[7]: stage ms,
   stage launches, nodes per launch, VGPRs of the flat band's kernel, metric planes stored, material by depth
   (0/1), terrain band reading the z-row metrics only (0/1). */
int swl_engine3d_profile(swl_engine3d *e, int enable);
int swl_engine3d_kernel_timing(swl_engine3d *e, double *out);
/* Engine tuning switches, all with identical fields (the energy sum may differ in its last bits, summation order):
   "vector_terrain", "vector_flat": x-adjacent nodes per lane of the terrain and flat bands, read and written as
   one vector (2, or 4 in FP32; 0 or 1: one node per lane).  Defaults: FP64 2 / 2, FP32 2 / 4.
   "overlap_bands" (default 1): the surface band runs on a second stream beside the deep bands of the same stage.
   "split_surface" (default 0): one launch per surface plane, each compiled for its depth.
   "lds_tiles" (default 0): the one-node-per-lane deep bands stage each field's tile and x / y halo in LDS.
   "brick_x", "brick_y", "brick_z": the deep bands' tile order within each XCD's run, bricks of x by y tiles by
   z planes (brick_x 0 = tile columns, the default).
   "z_row_metrics" (default 1): where the terrain band's inverse metrics 0..5 are exactly (1, 0, 0, 0, 1, 0), as the
   fourth-order derivatives of x = i h, y = j h give them, the vector terrain band reads only metrics 6..8.
   "z_row_surface" (default 1): there, the surface band takes metrics 0..5 as the constants they are. */
int swl_engine3d_set_option(swl_engine3d *e, const char *name, int value);

/* ---------------- 3D session (ISimulationBackend3D, include/Seis3D/CurvilinearSolver.hpp:165-185) ---------------- */
typedef struct swl_sim3d swl_sim3d;
/* backend 2 = HIP; precision 0 FP64 (the reference's), 1 FP32 state with FP64 closure/projection. */
int swl_sim3d_create(const swl_params3d *p, int backend, int device, int precision, swl_sim3d **out);
/* FP64 model only, no device (for checking the construction). */
int swl_sim3d_plan(const swl_params3d *p, swl_sim3d **out);
/* estimateModelMemory3D (include/Seis3D/ModelPresets3D.hpp:38-49): out[0] padded points, out[1] CPU MiB,
   out[2] CUDA device MiB, out[3] CUDA host MiB, out[4] within the padded-node budget (0/1); out[5] this library's
   HIP engine device bytes at `precision` (an upper bound).  swl_sim3d_create applies fitsMemory3D
   (ModelPresets3D.hpp:50-56) to the host estimate and the HIP device bytes (90 % of the device's free memory). */
int swl_memory_estimate3d(const swl_params3d *p, int precision, double *out);
void swl_sim3d_destroy(swl_sim3d *s);
int swl_sim3d_advance(swl_sim3d *s, int count);
int swl_sim3d_reset(swl_sim3d *s);
int swl_sim3d_step(const swl_sim3d *s);
/* Replaces the source-window peak energy of the late-growth fault rule (CurvilinearSolver.cpp:477-488:
   E > 1e4*source_peak after 3/f), as swl_sim2d_override_peak_energy does for 2D. */
int swl_sim3d_override_source_peak(swl_sim3d *s, double peak);
int swl_sim3d_total_steps(const swl_sim3d *s);
double swl_sim3d_dt(const swl_sim3d *s);
/* layout: nx_pad, ny_pad, nz_pad, receivers, samples recorded so far (t = 0 included) */
int swl_sim3d_layout(const swl_sim3d *s, int64_t *out);
/* diagnostics: the 18 values of the oracle order (energy, peak, source peak, max |V|, traction-rate
   ratio, finite, min J, max metric norm, min cell length, max edge, wave/damping/source dt,
   display peaks x5) followed by initialization ms, total step ms, resident device bytes. */
int swl_sim3d_diagnostics(const swl_sim3d *s, double *out);
int swl_sim3d_receivers(const swl_sim3d *s, double *vx, double *vy, double *vz, double *position, int64_t *index);
int swl_sim3d_snapshot(const swl_sim3d *s, int field, double *out);
/* 0 metrics(9N), 1 J, 2 normals(3N), 3 coordinates(3N), 4 lambda, 5 mu, 6 density, 7 damping, 8 source,
   9..11 moment gradient (FP64 construction values, same numbering as the oracle) */
int swl_sim3d_model(const swl_sim3d *s, int which, double *out);
/* readDisplayFrame (CurvilinearSolver.cpp:545-576): component 0..4 */
int swl_sim3d_read_display_frame(const swl_sim3d *s, int x, int y, int z, int component, float *out);
int swl_sim3d_profile(swl_sim3d *s, int enable);
int swl_sim3d_kernel_timing(const swl_sim3d *s, double *out);
int swl_sim3d_set_option(swl_sim3d *s, const char *name, int value); /* swl_engine3d_set_option */
/* writeReceiversCsv(simulation, path, components, error) (include/Seis3D/ReceiverIo.hpp:5-7,
   src/three_d/ReceiverIo.cpp:9-51): components[3] = Vx, Vy, Vz; one row per accepted step, t = 0 included. */
int swl_sim3d_write_receivers_csv(const swl_sim3d *s, const char *path, const int32_t *components);
/* writeReceiversCsv(simulation, path, error, allVelocityComponents) (ReceiverIo.hpp:8-9): Vz only with the legacy
   column names, or all three components. */
int swl_sim3d_write_receivers_vz_csv(const swl_sim3d *s, const char *path, int all_velocity_components);
/* writeVts(snapshot(), path, VtsExportOptions3D, error) (include/Seis3D/VtsWriter.hpp:11-24,
   src/three_d/VtsWriter.cpp:17-117): fields[11] = Vx, Vy, Vz, |V|, mean normal stress, Sxx, Syy, Szz, Sxy, Sxz,
   Syz; materials (per node with layer ids for layered models); the velocity vector; written to `path`.tmp-<ticks>
   and renamed. */
int swl_sim3d_write_vts(const swl_sim3d *s, const char *path, const int32_t *fields, int include_material,
                        int include_velocity_vector);

/* ---------------- 3D background stepping (Seis3D::Curvilinear::AsyncSimulation3D) ----------------
   include/Seis3D/AsyncSimulation3D.hpp:5-33, src/three_d/AsyncSimulation3D.cpp: queue(maxSteps, budgetMs) runs at
   most 8 steps on a worker thread; poll / synchronize publish its step, diagnostics, the Vx, Vy, Vz samples it
   recorded and the requested display slices; every other call first cancels the batch between steps and waits. */
typedef struct swl_async3d swl_async3d;
int swl_async3d_create(swl_async3d **out);
void swl_async3d_destroy(swl_async3d *a);
int swl_async3d_initialize(swl_async3d *a, const swl_params3d *p, int backend, int device, int precision);
int swl_async3d_queue(swl_async3d *a, int steps, double budget_ms); /* 1 queued, 0 busy / completed / uninitialized */
int swl_async3d_poll(swl_async3d *a);          /* non-zero: the last batch's error (swl_last_error) */
void swl_async3d_pause(swl_async3d *a);
int swl_async3d_synchronize(swl_async3d *a);
int swl_async3d_busy(swl_async3d *a);
/* backgroundTiming: compute ms, transfer ms, batches, steps */
int swl_async3d_timing(const swl_async3d *a, double *out);
int swl_async3d_advance(swl_async3d *a, int count);
int swl_async3d_reset(swl_async3d *a); /* resetFields */
/* published state: initialized, completed, step, total, receivers, samples per receiver */
int swl_async3d_state(const swl_async3d *a, int64_t *out);
double swl_async3d_dt(const swl_async3d *a);
int swl_async3d_diagnostics(const swl_async3d *a, double *out); /* as swl_sim3d_diagnostics */
int swl_async3d_receivers(const swl_async3d *a, double *vx, double *vy, double *vz);
/* readDisplayFrame (AsyncSimulation3D.cpp:86-92): the published slices of the current step, else (worker idle)
   read now; *ready = 0 while a batch runs and no current slices exist (the request is kept for the next batch). */
int swl_async3d_read_display_frame(swl_async3d *a, int x, int y, int z, int component, float *out, int *ready);
/* the session (owned by the async object), after the running batch finishes */
swl_sim3d *swl_async3d_session(swl_async3d *a);

#ifdef __cplusplus
}
#endif
#endif
