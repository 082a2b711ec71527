/*
 * dem.h - C ABI of the MI355X (gfx950) DEM time step.
 *
 * This library stands in for the compute half of the reference's wgpu host
 * (WGPUComputeProg, src/wgpu_prog.rs:677-696): it owns the device buffers of one
 * scene, runs the two-pass time step (2D_LOM.wgsl + 2D_Simulation.wgsl) as HIP
 * kernels, and converts between the reference's host layouts (src/state.rs:69-93)
 * and its own device layout on upload and download.
 *
 * Conventions
 *   - Every entry point returns DEM_OK (0) or a negative DEM_E* code; none throws
 *     or aborts across the ABI.  dem_last_error() returns a thread-local message
 *     describing the last failure on the calling thread.
 *   - One dem_ctx is bound to one HIP stream.  Calls on one ctx must be
 *     serialised by the caller; distinct ctxs may be used from different threads.
 *   - Arrays handed in through dem_state_view are caller-owned and are copied;
 *     the library keeps no pointer to them after the call returns.
 */
#ifndef PSIM_DEM_H
#define PSIM_DEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_OK          0
#define DEM_EINVAL     -1   /* bad argument */
#define DEM_ENOMEM     -2   /* device or host allocation failed */
#define DEM_EDEVICE    -3   /* HIP runtime error */
#define DEM_ECAPACITY  -4   /* a fixed capacity (contacts, bonds, slots) was exceeded */
#define DEM_ESTATE     -5   /* call out of order (e.g. step before upload) */
#define DEM_EBUSY      -6   /* serving: no room for the scene right now */

/* Physics uniform: 29 words in the reference's order.
 * Packed by Settings::collision_settings (src/settings.rs:2086-2119) and read as
 * `struct Settings` by 2D_LOM.wgsl:26-56 and 2D_Simulation.wgsl:34-64. */
typedef struct dem_uniform {
    int32_t walls;                      /*  0 (the step applies walls regardless, 2D_LOM.wgsl:102) */
    float   hor_bound;                  /*  1 full box width */
    float   vert_bound;                 /*  2 full box height */
    int32_t round_bounds;               /*  3 circular wall instead of the box */
    float   bound_radius;               /*  4 */
    float   wall_friction;              /*  5 not read by the step */
    int32_t gravity;                    /*  6 */
    int32_t planet_mode;                /*  7 gravity toward a point */
    int32_t bonds;                      /*  8 bond law: 0 none, 1 normal, 2 linear contact, 3 parallel */
    int32_t collisions;                 /*  9 narrow phase on/off */
    float   friction_coefficient;       /* 10 */
    float   gravity_acc;                /* 11 multiplier on 9.81 */
    float   bond_normal_stiffness;      /* 12 */
    int32_t bonds_tear;                 /* 13 */
    float   bond_normal_strength;       /* 14 */
    float   contact_damping;            /* 15 force factor of bond law 2 */
    float   bond_damping;               /* 16 not read by the step */
    float   drag;                       /* 17 not read by the step */
    float   bond_shear_strength;        /* 18 */
    float   dT;                         /* 19 time step */
    float   bond_shear_stiffness;       /* 20 */
    float   gravity_x;                  /* 21 planet-mode centre (mouse) */
    float   gravity_y;                  /* 22 */
    int32_t mouse_gravity;              /* 23 */
    float   moment_contribution_factor; /* 24 */
    int32_t collision_interval;         /* 25 narrow phase every k-th step */
    int32_t local_damping;              /* 26 not read by the step */
    float   local_damping_alpha;        /* 27 not read by the step */
    int32_t particles;                  /* 28 active particle count N */
} dem_uniform;

/* Material record, 6 floats (2D_Simulation.wgsl:66-73, settings.rs:266). */
typedef struct dem_material {
    float red, green, blue, density, normal_stiffness, shear_stiffness;
} dem_material;

/* Broad-phase grid in the reference's GridInfo convention (2D_LOM.wgsl:19-24,
 * wgpu_prog.rs:698-717): w x h cells of side cell_size centred on the origin,
 * column 0 at x = -cell_size*w/2, row 0 at y = +cell_size*h/2 (rows top-down).
 * Positions outside the grid clamp to the edge cells.  cell_size must be at
 * least twice the largest radius plus a few ulp of the grid extent; what it exceeds
 * that by is the skin of the neighbour lists (lists are rebuilt when a particle has
 * moved 0.45 skin).  Large bodies do not count: with the radii in decreasing order
 * r_0 >= r_1 >= ..., the first k <= 64 with r_{k-1} > 2 r_k makes the top k discs
 * large bodies (a striker among small discs), which pair with every disc through a
 * per-scene list, and "the largest radius" is r_k.  cell_cap is kept for layout
 * compatibility; the counting sort has no per-cell capacity. */
typedef struct dem_grid {
    float   cell_size;
    int32_t cell_cap;
    int32_t w;
    int32_t h;
} dem_grid;

typedef struct dem_params {
    dem_uniform         u;
    const dem_material* materials;      /* n_materials records, copied */
    int32_t             n_materials;
    dem_grid            grid;
    int32_t use_f64;           /* reference "$ F64" variant: distances in f64 (2D_Simulation.wgsl:264-284) */
    int32_t record_data;       /* maintain data[7] per particle (Σfn, Σft, ΣM, 0, v*x, v*y, ω*) */
    int32_t reorder_interval;  /* re-sort particle storage into cell order (with a neighbour-list
                                  rebuild) every k steps; 0: only after an upload; < 0: never */
    int32_t block_size;        /* threads per block for particle kernels; 0 = 256 */
    int32_t graph;             /* 1: capture each dem_step call's launches in a HIP graph and replay
                                  it when a later call repeats the same launches (same step count,
                                  buffers, rebuild and re-sort pattern); 0: launch directly */
    float   restitution;       /* extension (the reference's contact law has no dashpot): e in (0, 1)
                                  adds a normal dashpot 2 zeta sqrt(m_eff kn) v_n to every contact,
                                  zeta = -ln e / sqrt(pi^2 + ln^2 e), so that a binary collision
                                  rebounds with restitution e; 0 (or >= 1): no dashpot, the
                                  reference's law */
} dem_params;

/* Caller-owned arrays in the reference host layouts (state.rs:69-93).
 *   pos, vel, acc      2N floats (x, y interleaved)
 *   rot, rot_vel, rot_acc, radii   N floats
 *   forces             6N floats {x, y, rot, delX, delY, delRot}; only x, y, rot are read
 *   fixity             6N int32 {x_vel, y_vel, rot_vel, x_vel_2, y_vel_2, rot_vel_2}
 *   material_pointers  N int32
 *   bonds              3*n_bonds int32 {index (1 intact, -9/-10/-11 broken), angle bits, length bits}
 *   bond_strengths     optional 2*n_bonds floats {normal, shear} per bond; NULL = the uniform's
 *   contacts           contact_slots*6*N words {a, b, tangent_force, bond_tangent_force, theta_b, bonded}
 *                      per particle; b == -1 marks an empty slot
 *   data               optional 7N floats
 * Download writes each particle's occupied slots sorted by partner id b, then
 * fills the rest with -1 words. */
typedef struct dem_state_view {
    int32_t  n;
    float*   pos;
    float*   vel;
    float*   acc;
    float*   rot;
    float*   rot_vel;
    float*   rot_acc;
    float*   forces;
    float*   radii;
    int32_t* fixity;
    int32_t* material_pointers;
    int32_t* bonds;
    int32_t  n_bonds;
    float*   bond_strengths;
    float*   contacts;
    int32_t  contact_slots;
    float*   data;
} dem_state_view;

typedef struct dem_ctx dem_ctx;

/* Thread-local description of the last error on this thread ("" if none). */
const char* dem_last_error(void);

/* Library version string. */
const char* dem_version(void);

/* Number of visible HIP devices (0 if none).  Never fails. */
int dem_device_count(void);

/* Stands in for WGPUComputeProg::new (wgpu_prog.rs:720-1125) minus scene setup:
 * allocates device buffers for up to `capacity` particles and `bond_capacity`
 * bonds on the current HIP device. */
int dem_create(const dem_params* params, int32_t capacity, int32_t bond_capacity, dem_ctx** out);

/* Frees every device buffer of the ctx. */
int dem_destroy(dem_ctx* ctx);

/* Uniform / material / grid update: stands in for the collision_settings and
 * material_buffer updateUniform calls (client.rs:1129-1135, 1189-1195). */
int dem_set_params(dem_ctx* ctx, const dem_params* params);

/* Host -> device copy of a whole scene: stands in for WGPUComputeProg::restore
 * (wgpu_prog.rs:1235-1262) and resize_buffers (wgpu_prog.rs:1214-1233). */
int dem_upload_state(dem_ctx* ctx, const dem_state_view* state);

/* n time steps: stands in for WGPUComputeProg::compute with gen_per_frame = n
 * (wgpu_prog.rs:1420-1473).  Asynchronous on the ctx's stream. */
int dem_step(dem_ctx* ctx, int32_t n);

/* Device -> host copy: stands in for WGPUComputeProg::update_state
 * (wgpu_prog.rs:1166, state.rs:264-289).  Synchronous.  Any pointer in the view
 * may be NULL to skip that array. */
int dem_download_state(dem_ctx* ctx, dem_state_view* state);

/* Waits for all work queued on the ctx's stream. */
int dem_sync(dem_ctx* ctx);

/* Observables computed on the device (replace full readback + State::get_datum,
 * state.rs:708-737).  kinetic energy = Σ ½ m |v|² + ½ I ω² over active particles,
 * accumulated in f64 in particle-id order. */
int dem_kinetic_energy(dem_ctx* ctx, double* out);
/* Largest scene, in particle slots (the count rounded up to the block size), that
 * dem_kinetic_energy_noalloc accepts. */
#define DEM_KE_NOALLOC_MAX 4194304
/* dem_kinetic_energy without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Partial results live in a static device array of
 * the library (one per GPU, 33.6 MB, held until the process ends), reduced on the
 * device in a fixed order: the result equals dem_kinetic_energy's to f64 rounding
 * (about 1e-15 relative).  Scenes above DEM_KE_NOALLOC_MAX slots get DEM_ECAPACITY.
 * Every call on one GPU shares its array: calls that overlap on one GPU race and must
 * be serialised by the caller.  Synchronous. */
int dem_kinetic_energy_noalloc(dem_ctx* ctx, double* out);
/* Number of bonds whose index word is negative (broken). */
int dem_broken_bond_count(dem_ctx* ctx, int64_t* out);
/* Counters: [0] steps since upload, [1] neighbour-list rebuilds (broad phases),
 * [2] contact-list overflows, [3] bond-list overflows, [4] active contacts after the
 * last step, [5] active bond endpoints after the last step, [6] storage re-sorts,
 * [7] scan timeouts (always 0), [8] scatter errors, [9] particles that took the
 * stencil fallback because their neighbour list overflowed (summed over rebuilds),
 * [10] HIP graphs captured, [11] HIP graph replays.
 * [1], [2], [3], [8], [9], [10] and [11] count over every scene of the context. */
int dem_counters(dem_ctx* ctx, int64_t* out, int32_t n_out);

/* Sum over the particles flagged in `select` (length N, nonzero = selected) of
 * the reference's get_datum quantities (state.rs:708-737): out[10] =
 * {mean x, mean y, mean v*x, mean v*y, mean rot, mean ω*, Σ|Σfn|, mean Σft,
 *  mean ΣM, mean data3}; returns DEM_ESTATE if nothing is selected.
 * Requires record_data. */
int dem_get_datum(dem_ctx* ctx, const int32_t* select, double* out);

/* Sums over a group of particles given by id (ids[n_ids], each in [0, N)), computed on
 * the device from the resident state: the grip force and grip displacement of a
 * tensile test without reading the state back.  The reference gets the same
 * quantities by reading everything back and summing on the host (State::get_datum,
 * state.rs:708-737, over the selected grip particles).  With m = density * pi * r * r
 * in f32 (the mass the step divides by) and a the stored acceleration:
 *   out[0] Σ m a_x      out[1] Σ m a_y      (net force on the group: the reaction of
 *                                            a grip whose velocity is prescribed)
 *   out[2] Σ x          out[3] Σ y          (positions)
 *   out[4] Σ v_x        out[5] Σ v_y
 *   out[6] Σ |data[0]|  (normal-force datum, as get_datum; 0 without record_data)
 *   out[7] n_ids
 * Terms are f64 products of the f32 state, summed in f64 in a fixed order (equal to
 * any other summation order to f64 rounding).  Synchronous. */
int dem_group_sums(dem_ctx* ctx, const int32_t* ids, int32_t n_ids, double* out);

/* Stage-level views for parity checks.  Each copies device arrays to the host
 * after running the broad phase (stages 1-3) on the current positions.
 *   cell_keys[N]     centre-cell key cy*w+cx of particle i (by particle id),
 *                    w*h for inactive particles (r == 0)
 *   sorted_ids[N]    particle ids in (cell, id) order
 *   cell_start[w*h+1] first sorted slot of each cell
 * Any output may be NULL. */
int dem_debug_broad_phase(dem_ctx* ctx, int32_t* cell_keys, int32_t* sorted_ids, int32_t* cell_start);

/* Contact lists of the last narrow phase, by particle id: counts[N] and
 * partners[N*max_per] (sorted ascending, -1 padded).  Bond partners are not
 * included. */
int dem_debug_contacts(dem_ctx* ctx, int32_t max_per, int32_t* counts, int32_t* partners);

/* Bond builder on the cell list: the rule of State::regen_bonds
 * (state.rs:291-367) - pair (i, j), j > i, bonded iff
 * sqrt((xj-xi)^2 + (yj-yi)^2) < (ri + rj) * factor + gap in f32 (the reference:
 * factor 1.05, gap 0; a bonded-assembly packing: factor 1, gap 0.1 r_min, i.e. every
 * pair closer than gap), at most max_bonds per lower index i taken in increasing
 * j - evaluated on the device over the current positions.  Writes up to bond_cap
 * pairs (i, j) in (i, j) order and their count; returns DEM_ECAPACITY if more
 * exist.  gap >= 0. */
int dem_build_bond_pairs(dem_ctx* ctx, int32_t max_bonds, float factor, float gap,
                         int32_t* pairs, int32_t bond_cap, int32_t* n_out);

/* Timing of the last dem_step call measured with HIP events on the ctx's
 * stream: total milliseconds and per-kernel milliseconds
 * (out[0] total, out[1] force+drift kernel, out[2] cell scan, out[3] scatter,
 *  out[4] re-sort, out[5] bond pass), and the number of launches of each: both
 * arrays hold 6 entries. */
int dem_last_step_timing(dem_ctx* ctx, int32_t enable, double* ms_out, int64_t* launches_out);

/* ---- Serving loop: continuous batching of independent scenes ----
 * No counterpart in the reference (it runs one scene in one window).  One server owns
 * one set of device pools on one HIP stream.  dem_server_submit queues a scene (its
 * parameters, initial state and step count; everything is copied).  Each
 * dem_server_tick admits queued scenes in arrival order while their footprint fits
 * the pools (up to `bypass` later, smaller requests may pass a head that does not
 * fit), advances every resident scene by up to `substeps` steps in shared launches,
 * and retires the finished ones into host staging.  A scene's result is the same as
 * its solo run through dem_ctx, bit for bit. */
typedef struct dem_server dem_server;

typedef struct dem_server_params {
    int64_t slot_capacity;      /* particle slots of the pool (each scene rounds up to block_size) */
    int64_t cell_capacity;      /* grid cells of the pool (each scene takes w*h + 1) */
    int64_t bond_capacity;      /* bonds of the pool */
    int64_t record_capacity;    /* bond endpoint records; 0 = 2*bond_capacity + slot_capacity */
    int32_t scene_capacity;     /* scenes resident at once, at most */
    int32_t material_capacity;  /* material records; 0 = 8 per scene */
    int32_t block_size;         /* 0 = 256 */
    int32_t record_data;        /* maintain data[7] for the scenes */
    int32_t substeps;           /* steps per tick for each resident scene (> 0) */
    int32_t max_resident;       /* <= 0: no limit beyond the pools; 1 = one scene at a time */
    int32_t bypass;             /* later requests tried when the head does not fit */
    int32_t contact_slots;      /* contact table width of results whose request gave none; 0 = none */
    int32_t graph;              /* 1: a tick's launches run as a HIP graph, replayed when a later
                                   tick repeats them (dem_params.graph) */
} dem_server_params;

#define DEM_REQ_QUEUED   0
#define DEM_REQ_RESIDENT 1
#define DEM_REQ_DONE     2
#define DEM_REQ_FAILED   3   /* refused at admission (its parameters or state); dem_server_take
                                returns the error and its message */

int dem_server_create(const dem_server_params* params, dem_server** out);
int dem_server_destroy(dem_server* server);
/* Queues a scene: `params` as for dem_create, `state` as for dem_upload_state, run for
 * `nsteps` steps.  Writes a ticket.  Parameters, cell size and material count are checked
 * here (DEM_EINVAL / DEM_ECAPACITY: nothing is queued); a scene whose state the engine
 * refuses at admission (e.g. a contact table beyond 14 slots) becomes DEM_REQ_FAILED and
 * the queue moves on. */
int dem_server_submit(dem_server* server, const dem_params* params, const dem_state_view* state, int32_t nsteps,
                      int64_t* ticket);
/* One tick (admit, advance, retire); writes how many scenes finished in it. Synchronous. */
int dem_server_tick(dem_server* server, int32_t* completed);
/* DEM_REQ_QUEUED / DEM_REQ_RESIDENT / DEM_REQ_DONE / DEM_REQ_FAILED */
int dem_server_status(dem_server* server, int64_t ticket, int32_t* status);
/* Copies a finished scene's state into `out` (n must match; NULL arrays skipped),
 * writes its submit-to-done latency in seconds, and forgets the ticket.  For a failed
 * scene it returns the admission's error code (message in dem_last_error) and forgets
 * the ticket. */
int dem_server_take(dem_server* server, int64_t ticket, dem_state_view* out, double* latency_s);
/* out[0] completed, [1] resident, [2] queued, [3] p50 latency s, [4] p99 latency s,
 * [5] scenes/s from the first submit to the last completion, [6] particle-steps done,
 * [7] ticks, [8] step launches (max substeps per tick, summed), [9] scenes failed at admission. */
int dem_server_stats(dem_server* server, double* out, int32_t n_out);

#ifdef __cplusplus
}
#endif

#endif /* PSIM_DEM_H */
