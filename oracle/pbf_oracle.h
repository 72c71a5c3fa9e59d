/*
 * pbf_oracle.h — CPU oracle for the PBF substep (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference's position-based-fluids substep,
 * dispatch by dispatch, in particle-id order, as the GLSL compute shaders of
 * Realtime-FSI-Engine compute it (resources/shaders/compute/CS_*.glsl,
 * orchestrated by src/fluid/PBFluids.cpp:323-517).  It is the checker of the
 * HIP implementation and the timed CPU baseline; nothing in the product links
 * it.  Each function names the reference file:line it follows.
 *
 * Two instantiations of one body: *_f32 (the reference's FP32 arithmetic,
 * no contraction, IEEE division and square root) and *_f64 (the same
 * semantics in FP64, used for the density-constraint error report).
 *
 * Deliberate semantic choices, shared with the HIP path:
 *   - CS_Integrate reads neighbour velocities while other invocations write
 *     them (CS_Integrate.glsl:171 vs :277), a data race in the reference.  The
 *     oracle reads the post-predict velocities v*_j (a Jacobi snapshot).
 *   - pow(x, n) of the s_corr term (CS_ComputeDeltaP.glsl:103) is x multiplied
 *     by itself by binary powering when n is an integer in [0, 64], powf/pow
 *     otherwise.  GLSL leaves pow's precision to the driver.
 *   - texture() under GL_LINEAR / CLAMP_TO_EDGE (src/opengl/Texture3D.cpp:25-32)
 *     is the GL-spec trilinear filter evaluated in the working precision with
 *     nested (1-a)*t0 + a*t1 blends along x, then y, then z.
 *   - GLSL round() in CS_ComputeLOD.glsl:76 rounds half away from zero.
 *
 * Parity pins (tests/test_oracle_pins.py): closed forms of the kernels and the
 * coefficient block, integer-hash brute force, stable-sort invariants,
 * O(N^2) brute-force neighbour visits including aliased duplicates, isolated
 * particle / particle pair closed forms, ballistic flight, wall damping, an
 * analytic sphere SDF, the XSPH pair closed form, the curl of a rigid
 * rotation, vorticity confinement on a linear |w| field, the LOD and
 * interaction-force closed forms.  The reference holds no golden vectors or tests for
 * this path, and its shaders cannot run here (no GL context), so any result
 * those pins leave open is "parity unpinned".
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stdint.h>
#include "../include/fsi_pbf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Packed constants (FluidConfigUBO, include/fluid/PBFluids.h:21-61).  Values
 * are held as double; the f32 instantiation stores FP32 results (exactly
 * representable) and reads them back as float. */
typedef struct oracle_ubo {
    double bounds_min[3], bounds_max[3], gravity[3], sub_dt;
    double h, rho0, eps, wq;
    double k_corr, n_corr, viscosity, bound_damping;
    uint32_t hash_size, particle_count, enable_scorr, enable_viscosity;
    double cohesion_strength, interaction_radius, interaction_strength, w0_self;
    uint32_t hash_mask, min_lod, max_lod, enable_apbf;
    double poly6_coeff, spiky_coeff, inv_rho0, lod_max_dist;
} oracle_ubo;

/* Per-dispatch uniforms set by PBFluids::step (src/fluid/PBFluids.cpp:432-504)
 * and setInteraction (:312-317).  Float-valued inputs are given as doubles
 * holding FP32 values for the f32 instantiation. */
typedef struct oracle_uniforms {
    uint32_t substeps, iterations, enable_vorticity, is_interacting;
    double vorticity_epsilon;
    double interaction_pos[3];
    double camera_pos[3];
    uint32_t enable_sdf, sdf_mode;
    int32_t sdf_dims[3];
    double sdf_origin[3];
    double sdf_cell;
    double sdf_padding;
    const float* sdf_grid; /* dims product values, x fastest */
} oracle_uniforms;

/* Solver state in particle-id order.  REAL = float (f32) or double (f64). */
typedef struct oracle_state {
    uint32_t n, hash_size;
    void* pos;    /* REAL[4n]: pos.xyz, pad0            (Particle.pos)        */
    void* vel;    /* REAL[4n]: vel.xyz, pad2            (Particle.vel)        */
    void* solver; /* REAL[8n]: predPos_lambda, deltaP_rho (SolverBuffer)      */
    uint32_t* keys; /* n sorted hash keys   (hashGrid[k].x)                  */
    uint32_t* ids;  /* n sorted particle ids (hashGrid[k].y)                 */
    int32_t* cells; /* 2*hash_size (start, end), -1 = empty  (offsets)        */
    uint32_t* lod;  /* n                                                      */
    void* vsnap;    /* REAL[4n] scratch: post-predict velocities             */
    uint32_t* cand_count; /* optional n: candidates of the last density pass */
    uint32_t* nbr_count;  /* optional n: neighbours (r^2 <= h^2)             */
} oracle_state;

/* src/fluid/PBFluids.cpp:39-109 (+ Kernels.cpp:9-27 for wq).  bounds are the
 * padded bounds of setBounds (:299-306); damping already clamped. */
void oracle_pack_ubo_f32(const fsi_fluid_config* cfg, const float bmin[3], const float bmax[3],
                         float damping, uint32_t n, oracle_ubo* out);
void oracle_pack_ubo_f64(const fsi_fluid_config* cfg, const float bmin[3], const float bmax[3],
                         float damping, uint32_t n, oracle_ubo* out);

/* CS_PredictAndHash.glsl:77-83 */
uint32_t oracle_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t mask);

#define ORACLE_DECLARE(SFX)                                                                   \
    void oracle_predict_hash_##SFX(const oracle_ubo* u, const oracle_uniforms* s,             \
                                   oracle_state* st);                                         \
    void oracle_sort_##SFX(oracle_state* st);                                                 \
    void oracle_build_cells_##SFX(oracle_state* st);                                          \
    void oracle_lod_##SFX(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st);   \
    void oracle_lambda_##SFX(const oracle_ubo* u, oracle_state* st, uint32_t iter);           \
    void oracle_deltap_##SFX(const oracle_ubo* u, oracle_state* st, uint32_t iter);           \
    void oracle_apply_##SFX(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st,  \
                            uint32_t iter);                                                   \
    void oracle_integrate_##SFX(const oracle_ubo* u, const oracle_uniforms* s,                \
                                oracle_state* st);                                            \
    void oracle_vorticity_##SFX(const oracle_ubo* u, const oracle_uniforms* s,                \
                                oracle_state* st);                                            \
    void oracle_substep_##SFX(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st); \
    void oracle_step_##SFX(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st);  \
    uint32_t oracle_contact_filter_##SFX(const oracle_ubo* u, const oracle_uniforms* s,       \
                                         const oracle_state* st, double band,                 \
                                         fsi_contact_candidate* out, uint32_t cap);           \
    double oracle_poly6_##SFX(const oracle_ubo* u, const double r[3]);                        \
    void oracle_grad_spiky_##SFX(const oracle_ubo* u, const double r[3], double out[3]);      \
    double oracle_sdf_query_##SFX(const oracle_uniforms* s, const double p[3]);               \
    void oracle_sdf_gradient_##SFX(const oracle_uniforms* s, const double p[3], double out[3]);

ORACLE_DECLARE(f32)
ORACLE_DECLARE(f64)

/* Density-constraint error of a set of positions, evaluated in FP64 with the
 * density-pass semantics of CS_ComputeLambdas.glsl:108-162 (27 buckets,
 * aliased buckets visited twice).  x4: n float4 positions (w ignored).
 * Outputs: mean |C_i| and max C_i with C_i = max(rho_i/rho0 - 1, -cohesion),
 * and the mean density. */
void oracle_density_error_f64(const float* x4, uint32_t n, const oracle_ubo* u64,
                              double* mean_abs_c, double* max_c, double* mean_rho);

#ifdef __cplusplus
}
#endif
#endif /* PBF_ORACLE_H */
