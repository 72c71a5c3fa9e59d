/*
 * fsi_pbf.h — C-ABI of the MI355X (gfx950) position-based-fluids substep.
 *
 * This is the drop-in boundary for the fluid hot path of Realtime-FSI-Engine.
 * The reference drives the substep from the C++ class `PBFluids`
 * (include/fluid/PBFluids.h:88-268), which dispatches GLSL compute shaders over
 * OpenGL SSBOs (src/fluid/PBFluids.cpp:323-517).  Every entry point below
 * replaces one member of that class (cited per function); the C++ mirror class
 * `PBFluids` in realtime-fsi-engine_amd/host/PBFluids.h is written on top of it.
 *
 * Conventions
 *   - Plain pointers and sizes only; host pointers unless the name says `dev`.
 *   - Every call returns an int status: FSI_OK (0) or a negative FSI_ERR_*.
 *     fsi_last_error() returns the text of the last failure on this thread.
 *   - One handle owns one HIP stream on the device that was current when it was
 *     created.  A handle is not thread-safe; distinct handles are independent.
 *   - `step` is asynchronous (enqueued on the handle's stream); every `read_*`
 *     call synchronises that stream before copying out.
 *   - Particle records are the reference's 48-byte std430 layout
 *     (include/fluid/Particle.h:8-12): pos.xyz, pad, predPos.xyz, pad, vel.xyz, pad.
 */
#ifndef FSI_PBF_H
#define FSI_PBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSI_OK 0
#define FSI_ERR_INVALID_ARG (-1) /* bad pointer, size or parameter value          */
#define FSI_ERR_HIP (-2)         /* a HIP runtime call failed                      */
#define FSI_ERR_STATE (-3)       /* call not valid in the handle's current state   */
#define FSI_ERR_CAPACITY (-4)    /* caller buffer too small                        */
#define FSI_ERR_NO_DEVICE (-5)   /* no usable gfx950 device                        */

#define FSI_PBF_ABI_VERSION 1

/* ------------------------------------------------------------------------- */
/* Types                                                                      */
/* ------------------------------------------------------------------------- */

/* One particle, 48 bytes (include/fluid/Particle.h:8-12). */
typedef struct fsi_particle {
    float pos[3];
    float pad0;
    float pred_pos[3]; /* never written by the solver: kept only for the layout */
    float pad1;
    float vel[3];
    float pad2;
} fsi_particle;

/* Contact candidate, 32 bytes (include/fluid/PBFluids.h:67-70). */
typedef struct fsi_contact_candidate {
    float pos[4]; /* xyz world position, w = 0              */
    float vel[4]; /* xyz velocity, w = the particle's pad2  */
} fsi_contact_candidate;

/* Solver parameters, field for field the reference's FluidConfig
 * (include/core/config.h:9-63), same order; bools are uint32 (0/1). */
typedef struct fsi_fluid_config {
    float dt;
    float gravity[3];
    float h;
    float rho0;
    uint32_t solver_iterations;
    uint32_t substep_iterations;
    float eps;
    uint32_t particle_count;
    float spawn_min[3];
    float spawn_max[3];
    uint32_t spawn_random;
    float spacing;
    float initial_velocity[3];
    float bounds_min[3];
    float bounds_max[3];
    float bound_damping;
    uint32_t hash_size;
    uint32_t enable_scorr;
    float k_corr;
    float n_corr;
    float delta_q;
    float cohesion_strength;
    uint32_t enable_viscosity;
    float viscosity;
    float interaction_radius;
    float interaction_strength;
    uint32_t enable_vorticity;
    float vorticity_epsilon;
    uint32_t enable_apbf;
    uint32_t min_lod;
    uint32_t max_lod;
    float lod_max_dist;
} fsi_fluid_config;

/* The packed per-step constant block, byte for byte the reference's std140
 * FluidConfigUBO (include/fluid/PBFluids.h:21-61), 144 bytes.  Exposed so a
 * caller can check the coefficient precompute of PBFluids::setParams
 * (src/fluid/PBFluids.cpp:39-109). */
typedef struct fsi_pbf_constants {
    float bounds_min[4];
    float bounds_max[4];
    float gravity_dt[4]; /* xyz gravity, w = dt / substeps */
    float h, rho0, eps, wq;
    float k_corr, n_corr, viscosity, bound_damping;
    uint32_t hash_size, particle_count, enable_scorr, enable_viscosity;
    float cohesion_strength, interaction_radius, interaction_strength, w0_self;
    uint32_t hash_mask;
    float poly6_coeff, spiky_coeff, inv_rho0;
    uint32_t min_lod, max_lod;
    float lod_max_dist;
    uint32_t enable_apbf;
} fsi_pbf_constants;

/* Per-pass device timing (HIP events on the handle's stream), accumulated
 * since the last fsi_pbf_timing_reset. */
enum fsi_pbf_pass {
    FSI_PASS_PREDICT = 0,   /* predict + hash                               */
    FSI_PASS_SORT = 1,      /* all radix passes of one substep              */
    FSI_PASS_TABLES = 2,    /* cell start/end tables + reorder              */
    FSI_PASS_LAMBDA = 3,    /* density / lambda                             */
    FSI_PASS_DELTAP = 4,    /* position correction + apply + SDF projection */
    FSI_PASS_INTEGRATE = 5, /* XSPH, curl, collisions, velocity             */
    FSI_PASS_VORTICITY = 6, /* vorticity confinement, once per frame        */
    FSI_PASS_LOD = 7,       /* APBF level of detail                         */
    FSI_PASS_COUNT = 8
};

typedef struct fsi_pbf_timing {
    double ms[FSI_PASS_COUNT];       /* summed kernel time per pass           */
    uint64_t launches[FSI_PASS_COUNT]; /* number of timed launches per pass   */
    uint64_t substeps;
    uint64_t frames;
} fsi_pbf_timing;

typedef struct fsi_pbf fsi_pbf; /* opaque handle */

/* ------------------------------------------------------------------------- */
/* Library                                                                    */
/* ------------------------------------------------------------------------- */

/* Text of the last error raised on the calling thread ("" if none). */
const char* fsi_last_error(void);

/* ABI version of the loaded library (FSI_PBF_ABI_VERSION). */
int fsi_abi_version(void);

/* The reference's FluidConfig defaults (include/core/config.h:9-63). */
int fsi_fluid_config_default(fsi_fluid_config* out);

/* Number of visible HIP devices, without touching any of them. */
int fsi_device_count(int* count);

/* ------------------------------------------------------------------------- */
/* Solver handle — replaces class PBFluids (include/fluid/PBFluids.h:88-268)  */
/* ------------------------------------------------------------------------- */

/* PBFluids::PBFluids(const FluidConfig&)  (src/fluid/PBFluids.cpp:12-37).
 * Binds the current HIP device, creates the stream, applies setParams. */
int fsi_pbf_create(const fsi_fluid_config* cfg, fsi_pbf** out);

/* ~PBFluids (RAII release of every buffer, src/opengl/SSBO.cpp:18-21). */
int fsi_pbf_destroy(fsi_pbf* h);

/* PBFluids::setParams  (src/fluid/PBFluids.cpp:39-109).  Resets the bounds
 * from cfg (setBounds with the current collision padding), recomputes the
 * packed constants.  hash_size must be a power of two >= 2; a changed
 * hash_size reallocates the cell table.  The particle count stays the one
 * given to fsi_pbf_set_particles (cfg->particle_count is ignored after that). */
int fsi_pbf_set_params(fsi_pbf* h, const fsi_fluid_config* cfg);

/* PBFluids::params()  (include/fluid/PBFluids.h:97). */
int fsi_pbf_get_params(fsi_pbf* h, fsi_fluid_config* out);

/* The packed constants of the last setParams/setBounds/setParticles. */
int fsi_pbf_get_constants(fsi_pbf* h, fsi_pbf_constants* out);

/* PBFluids::setParticles  (src/fluid/PBFluids.cpp:111-152): copies n records
 * to the device and (re)allocates every per-particle buffer. n may be 0. */
int fsi_pbf_set_particles(fsi_pbf* h, const fsi_particle* particles, uint32_t n);

/* Number of particles currently held. */
int fsi_pbf_particle_count(fsi_pbf* h, uint32_t* n);

/* PBFluids::readbackParticles  (src/fluid/PBFluids.cpp:154-191).  Copies the
 * current state (all 48 bytes per record) into out[0..n); cap must be >= n. */
int fsi_pbf_read_particles(fsi_pbf* h, fsi_particle* out, uint32_t cap, uint32_t* n);

/* PBFluids::bindParticlesForRendering  (include/fluid/PBFluids.h:143): the
 * particle buffer the renderer reads, the std430 `Particle` records
 * {vec4 pos, vec4 predPos, vec4 vel} at binding 0 of
 * resources/shaders/graphics/VS_Particle.glsl:3-12.  Writes the current state
 * as n 48-byte records (pads included, predPos as last set) into device_out
 * (device memory of n * 48 bytes, 16-byte aligned), on the device, enqueued
 * after every step already enqueued; if stream is not NULL it waits for the
 * write. */
int fsi_pbf_write_render_buffer(fsi_pbf* h, void* device_out, void* stream);

/* PBFluids::setCollisionPadding  (include/fluid/PBFluids.h:108).  Like the
 * reference it takes effect at the next set_bounds / set_params. */
int fsi_pbf_set_collision_padding(fsi_pbf* h, float padding);

/* PBFluids::setBounds  (src/fluid/PBFluids.cpp:299-306): bounds shrunk by the
 * collision padding, damping clamped to [0,1]. */
int fsi_pbf_set_bounds(fsi_pbf* h, const float bounds_min[3], const float bounds_max[3],
                       float damping);

/* PBFluids::setSDFBoundary + SDFBoundary::uploadToGPU
 * (include/fluid/PBFluids.h:120; src/fluid/SDFBoundary.cpp:580-598).
 * grid holds dims[0]*dims[1]*dims[2] signed distances, x fastest (the texture
 * order of the reference), voxel (i,j,k) centred at origin + (i,j,k)*cell.
 * grid == NULL detaches the boundary.  The data is copied. */
int fsi_pbf_set_sdf(fsi_pbf* h, const float* grid, const int32_t dims[3],
                    const float origin[3], float cell);

/* PBFluids::setSDFPadding / setSDFContainment  (include/fluid/PBFluids.h:124,130). */
int fsi_pbf_set_sdf_padding(fsi_pbf* h, float padding);
int fsi_pbf_set_sdf_containment(fsi_pbf* h, int inside);

/* PBFluids::setInteraction  (src/fluid/PBFluids.cpp:312-317). */
int fsi_pbf_set_interaction(fsi_pbf* h, int active, const float pos[3]);

/* PBFluids::setCameraPos  (include/fluid/PBFluids.h:100). */
int fsi_pbf_set_camera(fsi_pbf* h, const float pos[3]);

/* PBFluids::step  (src/fluid/PBFluids.cpp:323-517): substep_iterations
 * substeps and one vorticity pass, enqueued on the handle's stream. */
int fsi_pbf_step(fsi_pbf* h);

/* Block until all work enqueued on the handle has finished. */
int fsi_pbf_sync(fsi_pbf* h);

/* PBFluids::filterContactCandidates  (src/fluid/PBFluids.cpp:193-237):
 * particles with |sdf(pos)| < band are compacted, in particle-id order. */
int fsi_pbf_filter_contacts(fsi_pbf* h, float band);

/* PBFluids::readbackContactCandidates  (src/fluid/PBFluids.cpp:239-284).
 * *count receives the number of candidates (clamped to the particle count);
 * at most cap records are copied. */
int fsi_pbf_read_contacts(fsi_pbf* h, fsi_contact_candidate* out, uint32_t cap,
                          uint32_t* count);

/* ------------------------------------------------------------------------- */
/* Inspection taps (parity checks; the reference exposes these buffers only   */
/* as SSBOs: hashGrid, offsets, SolverBuffer — PBFluids.h:216-233)            */
/* ------------------------------------------------------------------------- */

/* Sorted (key, id) pairs and the cell table of the last substep.
 * keys/ids: n entries each; cells: 2*hash_size int32 (start, end), -1 = empty.
 * Any pointer may be NULL to skip it. */
int fsi_pbf_read_grid(fsi_pbf* h, uint32_t* keys, uint32_t* ids, int32_t* cells);

/* SolverBuffer in particle-id order: predPos_lambda and deltaP_rho as 4 floats
 * each (CS_ComputeLambdas.glsl:50-53).  Either pointer may be NULL. */
int fsi_pbf_read_solver(fsi_pbf* h, float* pred_pos_lambda, float* delta_p_rho);

/* Enable the diagnostic taps (neighbour counts of every density pass).
 * Off by default; the taps cost one extra store per particle per pass. */
int fsi_pbf_set_taps(fsi_pbf* h, int enable);

/* Per-particle counts from the last density pass, in particle-id order:
 * candidates = bucket entries visited other than the particle itself,
 * neighbours = those with |x_i - x_j|^2 <= h^2 (aliased buckets count twice,
 * as in CS_ComputeLambdas.glsl:126-150).  Requires taps. */
int fsi_pbf_read_counts(fsi_pbf* h, uint32_t* candidates, uint32_t* neighbours);

/* Per-pass device timing. */
int fsi_pbf_set_timing(fsi_pbf* h, int enable);
/* Timing of the passes whose bit (1 << FSI_PASS_*) is set in mask only: each
 * timed launch is bracketed by two events, which cost time of their own, so
 * a run that times one pass is closer to an untimed one. */
int fsi_pbf_set_timing_passes(fsi_pbf* h, uint32_t mask);
int fsi_pbf_timing_reset(fsi_pbf* h);
int fsi_pbf_read_timing(fsi_pbf* h, fsi_pbf_timing* out);

/* Launch-shape knobs picked by the autotuner (no counterpart in the
 * reference, whose shaders hard-code local_size_x = 256 and four 8-bit radix
 * passes, src/fluid/PBFluids.cpp:330,355-361).  Every setting computes the
 * same results; only speed differs. */
enum fsi_pbf_knob {
    FSI_TUNE_GATHER = 0,          /* neighbour-gather form: 0 single loop, 1 two-phase list, */
                                  /* 2 LDS-staged cell tiles                                 */
    FSI_TUNE_SORT_DIGIT_BITS = 1, /* widest radix digit, 1..10                                */
    FSI_TUNE_ORDER = 2,           /* thread -> particle order of the gather passes: 0 sorted  */
                                  /* slots as the sort leaves them, 1 Morton order of the      */
                                  /* predicted cells in per-XCD ranges over sorted-slot        */
                                  /* storage, 2 the same with the particle arrays stored in    */
                                  /* that order, whole buckets in Morton order of their cells  */
                                  /* (default); form 2 needs 1 or 2 and runs as form 1 under 0 */
    FSI_TUNE_BLOCK = 3,           /* workgroup size of the form-0/1 gathers: 128 (default) or */
                                  /* 256 (the reference's local_size_x)                       */
    FSI_TUNE_TILE = 4,            /* cell-tile shape of form 2: 0 4x4x4 cells (default),      */
                                  /* 1 8x4x4, 2 8x8x4                                          */
    FSI_TUNE_COUNT = 5
};
int fsi_pbf_set_tuning(fsi_pbf* h, int knob, int value);
int fsi_pbf_get_tuning(fsi_pbf* h, int knob, int* value);

/* Autotuner: times `frames` frames of the current state (after one
 * untimed frame) under every combination of processing order 1 or 2 and
 * digit width 6, 8 or 10 with gather form 0 or 1 at workgroup size 128 or
 * 256, and with gather form 2 at each tile shape (when the scene's box is
 * binned by 4x4x4-cell tiles),
 * restoring the particle state after each trial; keeps the fastest and
 * returns its gather form and digit width (the workgroup size and tile shape
 * through fsi_pbf_get_tuning).  The reference hard-codes its launch shape
 * (src/fluid/PBFluids.cpp:330-332,353); no counterpart. */
int fsi_pbf_autotune(fsi_pbf* h, uint32_t frames, int* gather, int* digit_bits);

/* One timed setting of the last fsi_pbf_autotune / fsi_batch_autotune call:
 * the knobs (FSI_TUNE_* values) and the milliseconds its frames took.
 * No counterpart in the reference. */
typedef struct fsi_tune_trial {
    int32_t gather, sort_digit_bits, order, block, tile;
    float ms;
} fsi_tune_trial;
/* The settings the last autotune timed, in the order it timed them (a
 * form-2 setting appears only where the scene's box could be binned by
 * tiles): up to cap records into out, their total count in *n. */
int fsi_pbf_read_autotune(fsi_pbf* h, fsi_tune_trial* out, uint32_t cap, uint32_t* n);

/* Replay each frame from a HIP graph captured on first use (re-captured when
 * parameters, buffers or knobs change).  Per-pass timing is not recorded in
 * this mode.  No counterpart in the reference. */
int fsi_pbf_set_graph(fsi_pbf* h, int enable);

/* Device pointer of the HIP stream the handle enqueues on (hipStream_t). */
int fsi_pbf_get_stream(fsi_pbf* h, void** stream);

/* Enqueue every later launch and copy of the handle on the caller's HIP
 * stream (hipStream_t of the handle's device; NULL: back to the handle's own
 * stream).  The handle's current stream is synchronised first.  The caller
 * keeps the stream alive while the handle uses it.  Several handles on one
 * stream run their frames in submission order, as the reference's dispatches
 * do on its one GL context (src/fluid/PBFluids.cpp:323-517); that is the only
 * counterpart. */
int fsi_pbf_set_stream(fsi_pbf* h, void* stream);

/* ------------------------------------------------------------------------- */
/* Multi-scene solver: many independent scenes, one launch per pass            */
/* (no counterpart in the reference, which runs one scene per PBFluids; every */
/* scene follows PBFluids::step, src/fluid/PBFluids.cpp:323-517)               */
/* ------------------------------------------------------------------------- */

typedef struct fsi_scene_desc {
    uint32_t particle_offset; /* scenes tile the particle array in order      */
    uint32_t particle_count;
    uint32_t hash_size;       /* power of two: this scene's cell table        */
    float bounds_min[3];      /* the box particles are kept in, i.e. the      */
    float bounds_max[3];      /* bounds after PBFluids::setBounds' padding;   */
                              /* damping is the shared config's               */
    const float* sdf_grid;    /* NULL: no boundary; else x fastest, copied    */
    int32_t sdf_dims[3];
    float sdf_origin[3];
    float sdf_cell;
    float sdf_padding;
    uint32_t sdf_containment;
} fsi_scene_desc;

typedef struct fsi_batch fsi_batch; /* opaque handle */

/* Shared solver parameters of every scene (cfg->hash_size, bounds and
 * particle_count are taken per scene). */
int fsi_batch_create(const fsi_fluid_config* cfg, fsi_batch** out);
int fsi_batch_destroy(fsi_batch* b);
/* particles: total records, scene s owning [offset_s, offset_s + count_s). */
int fsi_batch_set_scenes(fsi_batch* b, const fsi_particle* particles, uint32_t total, const fsi_scene_desc* scenes,
                         uint32_t nscenes);
/* One frame of every scene, enqueued on the handle's stream. */
int fsi_batch_step(fsi_batch* b);
int fsi_batch_sync(fsi_batch* b);
int fsi_batch_read_particles(fsi_batch* b, fsi_particle* out, uint32_t cap, uint32_t* n);
/* Scene s of the last substep, numbered as a single-scene handle numbers it:
 * keys inside the scene's table, ids from the scene's first particle. */
int fsi_batch_read_grid(fsi_batch* b, uint32_t scene, uint32_t* keys, uint32_t* ids, int32_t* cells);
int fsi_batch_set_graph(fsi_batch* b, int enable);
int fsi_batch_set_tuning(fsi_batch* b, int knob, int value);
int fsi_batch_get_tuning(fsi_batch* b, int knob, int* value);
int fsi_batch_autotune(fsi_batch* b, uint32_t frames, int* gather, int* digit_bits);
int fsi_batch_read_autotune(fsi_batch* b, fsi_tune_trial* out, uint32_t cap, uint32_t* n);
int fsi_batch_set_timing(fsi_batch* b, int enable);
int fsi_batch_set_timing_passes(fsi_batch* b, uint32_t mask);
int fsi_batch_timing_reset(fsi_batch* b);
int fsi_batch_read_timing(fsi_batch* b, fsi_pbf_timing* out);

/* ------------------------------------------------------------------------- */
/* Host-side helpers for scenes and boundaries                                */
/* ------------------------------------------------------------------------- */

/* Lattice block spawn (the grid branch of spawnParticles, src/main.cpp:160-188,
 * with a seeded jitter): nx*ny*nz particles at
 * origin + (i + 0.5)*spacing + U(-jitter, +jitter) per axis, ids x-fastest,
 * then y, then z; velocity v0; all pads 0.  The jitter draws come from an
 * MT19937 generator seeded with `seed`, three draws per particle (x, y, z),
 * u = (draw >> 8) * 2^-24, offset = (2u - 1) * jitter. */
int fsi_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float origin[3],
                      float spacing, float jitter, uint32_t seed, const float v0[3],
                      fsi_particle* out);

/* spawnParticles  (src/main.cpp:160-188): cfg->particle_count particles in
 * [spawn_min, spawn_max].  Grid mode: x, y, z stepped by max(1e-6, spacing)
 * from spawn_min in FP32 (z outermost, x fastest) while inside spawn_max and
 * below the count.  Random mode: std::mt19937 seeded 1337 and
 * std::uniform_real_distribution<float>(0, 1), three draws (x, y, z) per
 * particle, inside [spawn_min + spacing, spawn_max - spacing].  predPos = pos,
 * vel = initial_velocity, pads 0.  *count gets the number spawned (the grid
 * can hold fewer than particle_count); out may be NULL to ask for it. */
int fsi_spawn_particles(const fsi_fluid_config* cfg, fsi_particle* out, uint32_t cap, uint32_t* count);

/* loadOFF  (src/main.cpp:364-415): an OFF file, polygons fan-triangulated.
 * Comment (#) and blank lines are skipped; vertex lines may carry extra
 * values (normals, colours), which are ignored.  Two calls: with verts and
 * faces NULL only *nverts and *nfaces (triangles) are returned; then
 * verts (3 * nverts doubles) and faces (3 * nfaces ints) are filled.
 * FSI_ERR_INVALID_ARG if the file cannot be opened or is not an OFF file. */
int fsi_load_off(const char* path, double* verts, uint32_t vcap, int32_t* faces, uint32_t fcap,
                 uint32_t* nverts, uint32_t* nfaces);

/* fitMeshToDomain  (src/main.cpp:419-432): centre the mesh's bounding box on
 * `center` and scale its largest extent (at least 1e-9) to target_size, in
 * place. */
int fsi_mesh_fit_to_domain(double* verts, uint32_t nverts, const float center[3], double target_size);

/* Icosphere: subdivisions of a unit icosahedron projected on the sphere
 * (centre c, radius r).  Sizes: V = 10*4^s + 2, F = 20*4^s. */
int fsi_mesh_icosphere(uint32_t subdivisions, const double centre[3], double radius,
                       double* verts /* 3V */, int32_t* faces /* 3F */);

/* Signed-distance grid of a closed triangle mesh, as
 * SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-436): FP64 BVH
 * closest point, pseudo-normal sign, origin = bbox - 5*cell, dims =
 * ceil(span / cell).  Two calls: with grid == NULL only origin/dims are
 * returned; then grid (dims product floats) is filled in x-fastest order,
 * ready for fsi_pbf_set_sdf. */
int fsi_sdf_build(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces,
                  double cell, double origin[3], int32_t dims[3], float* grid);

/* ------------------------------------------------------------------------- */
/* FSI coupling on the device: the obstacle mesh, its signed-distance grid    */
/* and the contact forces.  The reference keeps these on the host            */
/* (SDFBoundary, src/fluid/SDFBoundary.cpp; computeContactForces,             */
/* src/main.cpp:604-710) and reads the contact candidates back every frame.   */
/* ------------------------------------------------------------------------- */

typedef struct fsi_mesh fsi_mesh; /* opaque handle */

/* SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-398) with the
 * grid on the device: triangles, pseudo-normals and the BVH are prepared on
 * the host (FP64), uploaded, and every voxel is computed on the device with
 * the same arithmetic as fsi_sdf_build (which gives the identical grid). */
int fsi_mesh_create(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces,
                    double cell, fsi_mesh** out);
int fsi_mesh_destroy(fsi_mesh* m);
/* Grid placement: origin = bbox - 5*cell, dims = ceil(span / cell). */
int fsi_mesh_grid_info(fsi_mesh* m, double origin[3], double* cell, int32_t dims[3]);
/* The grid (dims product floats, x fastest) to the host. */
int fsi_mesh_read_sdf(fsi_mesh* m, float* grid);
/* SDFBoundary::refitFromMesh (src/fluid/SDFBoundary.cpp:454-467): moved
 * vertices (same topology, 3 doubles each): corners and face normals
 * refreshed, BVH boxes refit, every voxel recomputed.  Edge and vertex
 * pseudo-normals keep their build values, as in the reference. */
int fsi_mesh_refit(fsi_mesh* m, const double* verts);
/* SDFBoundary::refitDirty (src/fluid/SDFBoundary.cpp:469-548): only voxels
 * within band_radius of a vertex that moved more than move_eps since the
 * last refit (around its old and its new position) are recomputed.  Out:
 * the number of recomputed voxels and their bounding box in voxel (x, y, z)
 * coordinates (min > max when none). */
int fsi_mesh_refit_dirty(fsi_mesh* m, const double* verts, double band_radius, double move_eps,
                         int32_t dirty_min[3], int32_t dirty_max[3], uint32_t* dirty_voxels);
/* SDFBoundary::closestTriangle (src/fluid/SDFBoundary.cpp:600-610) of n
 * points (3 floats each, host): per point the index of the nearest triangle
 * in the face list the mesh was created with (-1 for none) and the closest
 * point on it (3 floats), from the BVH of the mesh's current vertices, in
 * FP64 as the reference's query. */
int fsi_mesh_closest(fsi_mesh* m, const float* points, uint32_t n, int32_t* faces, float* closest);
/* The per-vertex forces of the last fsi_pbf_contact_forces (3 floats each). */
int fsi_mesh_read_forces(fsi_mesh* m, float* forces);

/* The mesh's current grid as the solver's boundary (PBFluids::setSDFBoundary
 * + SDFBoundary::uploadToGPU, include/fluid/PBFluids.h:120): a device copy,
 * ordered after the mesh's last refit; call again after a refit. */
int fsi_pbf_set_sdf_from_mesh(fsi_pbf* h, fsi_mesh* m);

/* computeContactForces (src/main.cpp:604-710) on the device over the
 * candidates of the last fsi_pbf_filter_contacts (no readback): closest
 * triangle, push direction (particle to surface, or the inward face normal
 * within 1e-6), magnitude max(0, v . dir) or 1, barycentric split onto the
 * three vertices.  Per-vertex sums run in particle-id order of the
 * candidates.  forces (3 floats per vertex, may be NULL) and counts (may be
 * NULL: [0] contacts, [1] vertices with |F| >= max(1e-5, control_threshold))
 * are copied to the host. */
int fsi_pbf_contact_forces(fsi_pbf* h, fsi_mesh* m, float band, int use_vel_magnitude,
                           float control_threshold, float* forces, uint32_t counts[2]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the contact
 * forces of fsi_pbf_contact_forces, same kernels and the same results, on
 * `count` candidates given in device memory (32-byte fsi_contact_candidate
 * records, e.g. those fsi_pbf_read_contacts returns), with every scratch
 * buffer in one fixed device array (6.6 MiB per device, kept until the process
 * ends) instead of buffers allocated on demand.  It calls no allocator.
 * Limits: count <= 65536 candidates and a mesh of <= 65535 vertices, else
 * FSI_ERR_INVALID_ARG.  Out, in device memory: forces, 4 floats per vertex
 * (x, y, z, 0; 16-byte aligned); counts[2] as fsi_pbf_contact_forces.
 * Enqueued on `stream` (NULL: the mesh's own) after the mesh's last update;
 * the results are ready when the stream reaches this point.  Calls on one
 * device share the scratch: calls that overlap in time on one device race
 * (order them on one stream). */
int fsi_mesh_contact_forces_noalloc(fsi_mesh* m, const fsi_contact_candidate* candidates, uint32_t count,
                                    float band, int use_vel_magnitude, float control_threshold, float* forces,
                                    uint32_t* counts, void* stream);

/* ------------------------------------------------------------------------- */
/* Differentiable frames.  No counterpart in the reference (its PBFluids has  */
/* no gradient): the reverse-mode derivative of `frames` whole frames of      */
/* PBFluids::step (src/fluid/PBFluids.cpp:323-517) with respect to the        */
/* initial positions and velocities, the XSPH coefficient (viscosity) and the */
/* relaxation epsilon, for fitting them to a target state.  APBF levels of    */
/* detail, the SDF boundary and the interaction force are not part of it.     */
/* States are device arrays of n float4 in particle-id order (xyz + the       */
/* record's pad, carried through).  `stream` is the caller's hipStream_t     */
/* (NULL: the null stream): a call starts after the work already queued on it */
/* and work queued on it afterwards starts after the call's work.             */
/* ------------------------------------------------------------------------- */

typedef struct fsi_adjoint fsi_adjoint; /* opaque handle */

/* cfg: everything but viscosity and eps, which each forward call sets. */
int fsi_adj_create(const fsi_fluid_config* cfg, uint32_t n, uint32_t frames, fsi_adjoint** out);
int fsi_adj_destroy(fsi_adjoint* a);
/* Precision of the differentiated trajectory: 32 (default) runs the frames
 * on the solver's own FP32 kernels, and the gradient is that of the FP32
 * trajectory, whose rounding (about 1e-7 of a position per operation) moves
 * a derivative of a many-frame run by up to about 1e-2 relative; 64 runs
 * the same passes in FP64 (coefficients of PBFluids::setParams,
 * src/fluid/PBFluids.cpp:44-101, evaluated in FP64) over explicit neighbour
 * lists, for derivatives that central differences of an FP64 simulation
 * confirm to 1e-3 and better.  The reverse passes compute in FP64 either
 * way.  Drops the checkpoints of the last forward. */
int fsi_adj_set_precision(fsi_adjoint* a, uint32_t bits);
/* Runs the frames from (pos0, vel0), keeping a checkpoint of every frame's
 * starting state and every substep's sorted permutation (n u32, which the
 * backward's replay uses instead of sorting again); pos_out / vel_out (may be
 * NULL) receive the final state (rounded to FP32 in the FP64 mode). */
int fsi_adj_forward(fsi_adjoint* a, const float* pos0, const float* vel0, float viscosity, float eps, float* pos_out,
                    float* vel_out, void* stream);
/* Given dL/d(final pos) and dL/d(final vel) (either may be NULL: zero),
 * writes dL/d(pos0) and dL/d(vel0) (may be NULL) and, on the host,
 * grad_params[0] = dL/d(viscosity), grad_params[1] = dL/d(eps).  Replays
 * each frame from its checkpoint; fails with FSI_ERR_STATE if a replay does
 * not reproduce the forward bit for bit, FSI_ERR_CAPACITY if a particle has
 * more than 96 neighbours or is the neighbour of more than 128 particles in a
 * pass (aliased buckets count once per visit).  Synchronises with the device. */
int fsi_adj_backward(fsi_adjoint* a, const float* grad_pos, const float* grad_vel, float* grad_pos0, float* grad_vel0,
                     double* grad_params, void* stream);
/* The same two calls with FP64 states (n double4 each) and parameters, for
 * the FP64 trajectory (FSI_ERR_STATE otherwise): a loss evaluated on FP32
 * positions carries their rounding into its gradient. */
int fsi_adj_forward64(fsi_adjoint* a, const double* pos0, const double* vel0, double viscosity, double eps,
                      double* pos_out, double* vel_out, void* stream);
int fsi_adj_backward64(fsi_adjoint* a, const double* grad_pos, const double* grad_vel, double* grad_pos0,
                       double* grad_vel0, double* grad_params, void* stream);
/* Of the last backward: out[0] capacity bits (1: neighbours, 2: pairs into one
 * particle), out[1] replayed records that differed from the forward, out[2]
 * the largest neighbour count and out[3] the largest in-degree of any pass. */
int fsi_adj_stats(fsi_adjoint* a, uint32_t out[4]);
/* Of the last backward: out[2f], out[2f+1] = the largest |component| of the
 * adjoint of the positions and of the velocities at the start of frame f
 * (NaN if any was NaN); `frames` pairs. */
int fsi_adj_frame_norms(fsi_adjoint* a, float* out);

/* ------------------------------------------------------------------------- */
/* Host mesh deformation of the two-way coupling: LaplacianDeformation        */
/* (src/deformation/LaplacianDeformation.cpp, include/deformation/            */
/* LaplacianDeformation.h) and its per-frame driver deformStep                */
/* (src/main.cpp:862-944).  Host code, FP64 arithmetic on vertex positions  */
/* kept as floats (the reference's mesh holds Vec3f points), with a          */
/* project-written sparse                                                    */
/* LDL^T (reverse Cuthill-McKee order) where the reference uses Eigen's        */
/* SimplicialLDLT, CHOLMOD and its IncrementalCholesky.                       */
/* ------------------------------------------------------------------------- */

typedef struct fsi_deform fsi_deform; /* opaque handle */

/* The deformStep tuning (src/main.cpp:251, 300-304, 343). */
typedef struct fsi_deform_params {
    float stiffness;             /* force -> displacement gain (1.0)             */
    float max_displacement;      /* per-frame push clamp (0.02)                  */
    float max_total_disp;        /* clamp of |target - rest| (0.40)              */
    uint32_t restore_enabled;    /* elastic pull back toward rest (1)            */
    float restore_strength;      /* fraction of the offset removed per frame (0.05) */
    float control_threshold;     /* |force| to enter the handle set (0.50)       */
    float control_threshold_low; /* |force| to stay in it (0.30)                 */
} fsi_deform_params;
int fsi_deform_params_default(fsi_deform_params* p);

/* LaplacianDeformation(mesh) + initialize() (LaplacianDeformation.cpp:7-23,
 * 77-195): cotangent weights max(0.5 * sum cot, 1e-4), L_ii = 1,
 * L_ij = -w_ij / sum w, delta = L V.  The given vertices are the rest pose. */
int fsi_deform_create(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces,
                      fsi_deform** out);
int fsi_deform_destroy(fsi_deform* d);
/* setUseLocal / setLocalKRing / setUseIncremental (LaplacianDeformation.h:
 * 40-66).  Defaults: local patch on, k = 3, incremental off. */
int fsi_deform_set_mode(fsi_deform* d, int use_local, int kring, int use_incremental);
/* addAnchor (LaplacianDeformation.cpp:407-417): pinned at its current
 * position; duplicates ignored. */
int fsi_deform_add_anchor(fsi_deform* d, int32_t vertex);
/* setControlPoints (LaplacianDeformation.cpp:424-428): count handles and
 * their targets (3 doubles each); replaces the previous set. */
int fsi_deform_set_control_points(fsi_deform* d, const int32_t* idx, const double* targets, uint32_t count);
/* precomputeSystem / solve (LaplacianDeformation.cpp:434-623): factor for
 * the current constrained set, then the positions for the current targets. */
int fsi_deform_precompute(fsi_deform* d);
int fsi_deform_solve(fsi_deform* d);
/* The current vertices (3 doubles each, every one a float value: the
 * rest pose and each solve's result are rounded to float, as the
 * reference's OpenMesh points are, LaplacianDeformation.cpp:388-389,568,617). */
int fsi_deform_read_vertices(fsi_deform* d, double* verts);
/* deformStep (src/main.cpp:862-944) given the per-vertex contact forces
 * (3 floats each, as fsi_pbf_contact_forces returns them) and the frame dt:
 * handles = non-anchor vertices pushed (|F| >= threshold, or >= the low
 * threshold if a handle last frame) or relaxing (a handle last frame, more
 * than 1e-3 from rest, restoring on); target = current + clamped push, eased
 * toward rest, clamped to max_total_disp from rest.  Refactors when the handle
 * set changed, solves when there are handles or the set changed (*solved). */
int fsi_deform_step(fsi_deform* d, const float* forces, float dt, const fsi_deform_params* p,
                    uint32_t* handle_count, int* solved);
/* out[0] local patch size, out[1] factor steps of the last (re)factorization,
 * out[2] handles of the last step, out[3] factorizations so far. */
int fsi_deform_info(fsi_deform* d, uint32_t out[4]);
/* verifyIncremental (LaplacianDeformation.cpp:671-716): max difference of a
 * solve through the incrementally updated factor and through a fresh one;
 * -1 when the incremental path is not active. */
int fsi_deform_verify_incremental(fsi_deform* d, double* max_diff);

#ifdef __cplusplus
}
#endif

#endif /* FSI_PBF_H */
