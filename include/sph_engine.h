/*
 * sph_engine.h - C ABI of the MI355X (gfx950) SPH step engine.
 *
 * The engine replaces the CPU solver that ttnghia/SimpleFluid drives through
 * the QtSPHSolver adapter.  Every entry point below names the reference call
 * site whose behaviour it takes over:
 *
 *   reference                                         here
 *   Include/QtSPHSolver.h:30-31  QtSPHSolver(data, params)   sph_create
 *   Source/Controller.cpp:52-64  updateSimParams/updateParams sph_set_params
 *   Source/Simulator.cpp:94-96   setupScene -> particle arrays sph_set_particles
 *   Source/Simulator.cpp:42      m_SPHSolver->makeReady()     sph_make_ready
 *   Source/Simulator.cpp:49      m_SPHSolver->advanceFrame()  sph_advance_frame
 *   Source/Simulator.cpp:46-51   inner frame loop (< 1/30 s)  sph_run_frame
 *   Include/QtSPHSolver.h:33     getNumParticles()            sph_num_particles
 *   Include/QtSPHSolver.h:34     getParticles()               sph_get_particles
 *   Include/QtSPHSolver.h:35     getVelocity()                sph_get_velocity
 *   Source/SceneManager.cpp:21-173 setupScene*                sph_scene_generate
 *
 * Conventions
 *   - Particle arrays cross the ABI as caller-owned host buffers of packed
 *     float triples (x, y, z) in ORIGINAL particle order (the order passed to
 *     sph_set_particles), the layout of the reference's Vec_Vec3<float>.
 *   - Device memory belongs to the handle.  A handle is used by one thread at
 *     a time (the reference's single worker thread, Source/Simulator.cpp:29).
 *   - Every int-returning call returns SPH_OK (0) or a negative SPH_E_* code;
 *     sph_last_error(h) gives the message (sph_last_error(NULL): the last
 *     failure on this thread without a handle, e.g. in sph_create).  No
 *     exception crosses the ABI.
 *   - Only sph_get_*, sph_synchronize and the debug fetches wait for the GPU.
 */
#ifndef SPH_ENGINE_H
#define SPH_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK            0
#define SPH_E_INVALID    -1  /* bad argument or parameter                       */
#define SPH_E_OOM        -2  /* device or host allocation failed                */
#define SPH_E_HIP        -3  /* a HIP runtime call failed                       */
#define SPH_E_NOT_READY  -4  /* no particles, or make_ready not called          */
#define SPH_E_DEVICE     -5  /* a kernel raised its error flag (bounded wait)    */
#define SPH_E_CAPACITY   -6  /* output buffer too small                         */

/* Reference scenes, Include/Common.h:49-60 (same enumerator values). */
#define SPH_SCENE_SPHERE_DROP     0
#define SPH_SCENE_CUBE_DROP       1
#define SPH_SCENE_DAMBREAK        2
#define SPH_SCENE_DOUBLE_DAMBREAK 3

/*
 * Solver parameters.  Field meaning follows SPHParameters<float> as set in
 * Source/Controller.cpp:52-64 (kernelRadius, pressureStiffness, viscosity,
 * boundaryRestitution, bUseAttractivePressure, stopTime, numThreads) and
 * Source/SceneManager.cpp:45 (particleRadius).  The remaining fields are the
 * engine's own formulation constants.
 */
typedef struct sph_params {
    float particle_radius;           /* r; lattice spacing is 2r                  */
    float kernel_radius;             /* h (support radius), reference: 2/res      */
    float rest_density;              /* rho0                                      */
    float pressure_stiffness;        /* B in p = B((rho/rho0)^7 - 1)              */
    float viscosity;                 /* mu (dynamic)                              */
    float boundary_restitution;      /* normal-velocity factor at the box walls   */
    int   use_attractive_pressure;   /* 0: p < 0 clamped to 0; 1: p *= ratio      */
    float attractive_pressure_ratio;
    float gravity[3];
    float dt;                        /* fixed step returned by advance_frame      */
    float box_min[3];                /* box faces: each bounds a half-space of    */
    float box_max[3];                /* rest fluid (its poly6 density and spiky
                                        pressure reach particles within h of the
                                        face); positions are also clamped to
                                        [box_min + r, box_max - r]               */
    float stop_time;                 /* driver only (Source/Simulator.cpp:44)     */
    int   num_threads;               /* accepted for parity, unused on the GPU    */
    float cfl_number;                /* 0: every step is dt.  > 0: a step is
                                        min(dt, cfl_number h / (c0 + max|v|)),
                                        c0 = sqrt(7 B / rho0), max over the scene's
                                        velocities at the start of the step
                                        (single-scene handles only)              */
} sph_params;

typedef struct sph_handle sph_handle;

/* Reference GUI defaults (panels in Captured/1.png, 2.png, 3.png): resolution 24, stiffness
 * 50000, viscosity 0.05, restitution 0.1, no attractive pressure, stop 5.0 s,
 * box [-1,1]^3, plus rho0 1000, gravity -9.81 y, dt 1e-4.                   */
int sph_params_default(sph_params* out);
/* kernel_radius = 2/res (Source/Controller.cpp:55), particle_radius = h/4.  */
int sph_params_set_resolution(sph_params* p, int resolution);

int  sph_create(const sph_params* p, sph_handle** out);
void sph_destroy(sph_handle* h);
int  sph_set_params(sph_handle* h, const sph_params* p);
int  sph_get_params(const sph_handle* h, sph_params* out);

/* Copies n particles in; vel may be NULL (zero velocity, as the reference's
 * scenes set, Source/SceneManager.cpp:63).                                  */
int  sph_set_particles(sph_handle* h, const float* xyz, const float* vel, uint32_t n);
int  sph_make_ready(sph_handle* h);
/* One step; *dt_out receives its length (the value advanceFrame() returns):
 * dt, or the CFL step when cfl_number > 0 (then the call waits for the GPU). */
int  sph_advance_frame(sph_handle* h, float* dt_out);
/* k steps, replayed from a captured HIP graph when graphs are enabled.     */
int  sph_advance_steps(sph_handle* h, uint32_t k);
/* Steps until the accumulated float frame time reaches 0.0333333333
 * (Source/Simulator.cpp:46-51); returns that time and the step count.
 * With cfl_number > 0 the frame runs on the device: a step graph that stops
 * itself at the end of the frame is replayed in batches, and the host reads
 * the frame state once per batch rather than every step's length.        */
int  sph_run_frame(sph_handle* h, float* frame_time_out, uint32_t* steps_out);

uint32_t sph_num_particles(const sph_handle* h);
int  sph_get_particles(sph_handle* h, float* xyz_out);
int  sph_get_velocity(sph_handle* h, float* vel_out);
int  sph_get_density(sph_handle* h, float* rho_out);   /* density of the last step */
int  sph_synchronize(sph_handle* h);
/* Frame-cadence position snapshot that overlaps the PCIe copy with further
 * steps (no reference counterpart: the reference's renderer reads the
 * solver's array while the worker writes it, Source/Simulator.cpp:50 ->
 * Source/FluidRenderWidget.cpp:211).  begin() captures the current positions
 * (original order) in stream order and returns at once; steps enqueued after
 * it run while a second stream copies the capture to page-locked memory;
 * wait() blocks until that copy is done and writes 3n floats.  One snapshot
 * is in flight at a time; a begin() waits on the GPU for the previous copy.
 * If the particle count changed in between (sph_set_particles), wait()
 * returns SPH_E_INVALID and drops the snapshot.                             */
int  sph_snapshot_begin(sph_handle* h);
int  sph_snapshot_wait(sph_handle* h, float* xyz_out);
/* State dump of one frame (no reference counterpart: the reference exports
 * only screenshots, Source/MainWindow.cpp:108-113): a NumPy .npy file of
 * float32, shape (n, 6) = x, y, z, vx, vy, vz per particle in original
 * order.  Waits for the GPU.                                                 */
int  sph_export_frame(sph_handle* h, const char* path);
const char* sph_last_error(const sph_handle* h);
/* Graph capture of the step on (default) or off.                           */
int  sph_set_use_graph(sph_handle* h, int on);

/* Reference scene generators (Source/SceneManager.cpp:21-173) in float, with
 * particle_radius = (2/res)/4.  Pass xyz_out = NULL to query the count.     */
int  sph_scene_generate(int scene, int resolution, float* xyz_out, uint32_t capacity,
                        uint32_t* n_out);
/* Seeded uniform sample of a ball: `count` points inside the ball of
 * `radius` around centre[3], no two closer than min_spacing (random
 * sequential addition: uniform darts, each kept when no kept point lies
 * within min_spacing).  The random counterpart of the SphereDrop lattice
 * (Source/SceneManager.cpp:67-91) for the sphere-drop workload.  Writes
 * 3 count floats; *n_out = points placed.  SPH_E_CAPACITY when 64 count
 * darts leave fewer than count points (the ball is too full at that
 * spacing); the points placed are written.                                  */
int  sph_scene_sample_ball(uint64_t seed, const float* centre, float radius, float min_spacing,
                           uint32_t count, float* xyz_out, uint32_t* n_out);

/* ---- one step that allocates nothing ---- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): one step of n particles whose state the caller holds in device
 * memory, enqueued on `stream` (a hipStream_t; NULL = the default stream),
 * with every intermediate in a fixed scratch array of the library (about
 * 17 MB per GPU, held until the process ends).  xyz_in, vel_in (may be NULL:
 * zero velocity), xyz_out, vel_out: device arrays of 3n floats in the
 * caller's particle order.  The step and its results equal the first
 * sph_advance_steps step of a handle given the same particles and params
 * (default launch shape).  Limits: n <= 65536, fixed dt (cfl_number 0), a
 * grid of at most 2^20 cells; past them SPH_E_INVALID, with the reason in
 * sph_last_error(NULL).  Returns without waiting for the GPU, so the
 * device-side checks a handle reports at its next synchronising call (a
 * cell of more than 1024 particles, keys out of range) are not reported
 * here.  All calls on one GPU share its scratch: calls that overlap on one
 * GPU race.                                                                 */
int  sph_step_noalloc(const sph_params* p, const float* xyz_in, const float* vel_in, float* xyz_out,
                      float* vel_out, uint32_t n, void* stream);

/* ---- batched ensemble (no reference counterpart: many scenes per launch) ---- */
typedef struct sph_scene_desc {
    sph_params params;               /* per-scene physics, box and dt             */
    uint32_t   num_particles;
} sph_scene_desc;

/* All scenes must share dt.  Particles are concatenated in scene order.     */
int  sph_batch_create(const sph_scene_desc* scenes, uint32_t nscenes, sph_handle** out);
uint32_t sph_num_scenes(const sph_handle* h);
/* offsets[s] .. offsets[s+1] is scene s's range in the concatenated arrays. */
int  sph_scene_offsets(const sph_handle* h, uint32_t* offsets);
/* Replaces scene s's stiffness, viscosity, restitution and attractive
 * settings (box, radii and dt must match the scene's create-time values).  */
int  sph_set_scene_params(sph_handle* h, uint32_t scene, const sph_params* p);

/* ---- differentiable rollout (no reference counterpart) ---- */
/* Runs `steps` steps from the current state, keeping a checkpoint of the
 * state every `checkpoint_every` steps for the reverse pass.                */
int  sph_rollout_forward(sph_handle* h, uint32_t steps, uint32_t checkpoint_every);
/* Reverse pass of the last sph_rollout_forward.  dL_dx_final: 3n floats in
 * original order.  Outputs (any may be NULL): dL/dstiffness and dL/dviscosity
 * per scene, dL/d(initial velocity) as 3n floats in original order.  The
 * handle's state is left at the rollout's final state.                       */
int  sph_rollout_backward(sph_handle* h, const float* dL_dx_final, float* dL_dstiffness,
                          float* dL_dviscosity, float* dL_dv0);

/* ---- device memory (no reference counterpart) ---- */
/* Caps the device memory this handle's buffers may hold (bytes; 0 = no cap,
 * the default).  Every call on the handle charges the buffers it allocates to
 * it; an allocation that would pass the cap is not made, the call returns
 * SPH_E_OOM with the limit in sph_last_error, and the handle keeps working
 * once the cap is raised.  The HIP runtime's own allocations (graphs, events,
 * page-locked download buffers) are not counted.  For callers that share a
 * GPU between several handles or with other work.                           */
int  sph_set_device_memory_limit(sph_handle* h, uint64_t bytes);
/* Device bytes the handle's buffers hold now.                               */
int  sph_device_memory_used(const sph_handle* h, uint64_t* bytes);

/* ---- measurement ---- */
#define SPH_PASS_HASH      0   /* standalone cell hash (first step only)        */
#define SPH_PASS_SORT      1   /* cell-start scan + sort of keys                */
#define SPH_PASS_REORDER   2   /* gather of the state into sorted order         */
#define SPH_PASS_DENSITY   3   /* density, EOS, neighbour masks                 */
#define SPH_PASS_FORCE     4   /* pressure + viscosity + integrate + clamp      */
#define SPH_NUM_PASSES     5
/* When on, every step records HIP events around each pass on the handle's
 * stream (graphs are bypassed).  read returns summed milliseconds per pass
 * and the number of steps measured, then resets the sums.                   */
int  sph_profile_enable(sph_handle* h, int on);
int  sph_profile_read(sph_handle* h, double* ms_per_pass, uint32_t* steps);
/* Sweeps block sizes / sort tile per pass on the current state and keeps the
 * fastest (state is saved and restored).  Results are cached per process
 * keyed by (device arch, particle-count bucket, scene count).               */
int  sph_autotune(sph_handle* h);
/* Current choice (5 ints): block sizes of density, force and reorder, the
 * sort items per thread, and the launch form: bit 0 set when density and
 * force run as persistent blocks taking tiles from per-XCD queues (clear:
 * one block per tile), bit 1 set when the force pass stages each row's
 * candidates in LDS (clear: it gathers them from global memory), bit 2 the
 * same for the density pass.  Every choice gives the same results bit for
 * bit.                                                                      */
int  sph_get_tuning(const sph_handle* h, int* cfg5);
int  sph_set_tuning(sph_handle* h, const int* cfg5);
/* Sort used to order particles by cell (no reference counterpart; both give
 * the same stable permutation and cell-start table):
 *   SPH_SORT_CELL_COUNT  one-digit (counting) radix sort over the whole cell
 *                        key; its histogram scan is the cell-start table (default)
 *   SPH_SORT_RADIX       multi-pass 8-bit LSD radix sort with decoupled look-back */
#define SPH_SORT_CELL_COUNT 0
#define SPH_SORT_RADIX      1
int  sph_set_sort_mode(sph_handle* h, int mode);
int  sph_get_sort_mode(const sph_handle* h);
/* Cells are cs = h(1 + 2^-10) in y and z and cs/xsub in x (no reference
 * counterpart).  xsub = 0 picks automatically (the largest of 4, 2, 1 with
 * at most 12 cells per particle, or 2^20 cells);
 * 1, 2, 4 force it.  Finer x cells let each neighbour row skip the x cells
 * outside the support sphere's chord.  Limits: 1024 cells per axis (times
 * xsub in x), 2^30 - 2 cells per handle, fewer than 2^28 particles; larger
 * setups return SPH_E_INVALID.                                              */
int  sph_set_cell_subdivision(sph_handle* h, int xsub);

/* ---- test-only views (parity against the CPU oracle) ---- */
/* Derived constants of scene s: h, h2, inv_cs, mass, cD, cP, cV, inv_rho0,
 * lo[3], hi[3] (14 floats) and the grid G[3] plus the x subdivision (4 ints). */
int  sph_debug_constants(const sph_handle* h, uint32_t scene, float* out14, int* G4);
/* State in the engine's internal order (the sorted order of the last step):
 * positions and velocities as 3n floats, original ids as n uint32.          */
int  sph_debug_get_state(sph_handle* h, float* xyz, float* vel, uint32_t* ids);
int  sph_debug_set_state(sph_handle* h, const float* xyz, const float* vel, const uint32_t* ids);
/* Intermediates of the last step: input keys (input state order), the sort
 * permutation and sorted keys, density and acceleration (sorted order).
 * acc3 is only filled when sph_debug_keep_acc(h, 1) was set before the step. */
int  sph_debug_keep_acc(sph_handle* h, int on);
int  sph_debug_step_outputs(sph_handle* h, uint32_t* keys_in, uint32_t* perm,
                            uint32_t* sorted_keys, float* rho, float* acc3);
uint32_t sph_debug_num_cells(const sph_handle* h);
int  sph_debug_cells(sph_handle* h, uint32_t* start, uint32_t* end);
/* Neighbour sets the force pass used in the last step (mask bits plus the
 * re-tested tail of long rows), self included, as CSR in sorted index space. */
int  sph_debug_neighbors(sph_handle* h, uint32_t* offsets, uint32_t* list, uint64_t capacity,
                         uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* SPH_ENGINE_H */
