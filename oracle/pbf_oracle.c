/*
 * pbf_oracle.c — CPU oracle for the PBF substep (TEST INFRASTRUCTURE ONLY:
 * linked by tests/ and by the CPU-baseline leg of bench.py, never by the
 * product library).  See pbf_oracle.h for the semantics and the pins.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, no fast-math).
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* CS_PredictAndHash.glsl:77-83: wrapping int32 products of the cell
 * coordinates with three primes, XOR, reinterpret as uint, mask. */
uint32_t oracle_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t mask) {
    const uint32_t n = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return n & mask;
}

/* Kernels.cpp:9-27 (calcPoly6Kernel, through calcSCorrKernel :216-219) in
 * FP32: the wq reference value of the s_corr term.  Note the coefficient is
 * (315/(64 pi)) / h^9 with h^9 = h4*h4*h, unlike the packed poly6Coeff. */
static float kernels_cpp_poly6_f32(float rx, float ry, float rz, float h) {
    const float pi = 3.14159265358979323846f;
    const float coeff = 315.0f / (64.0f * pi);
    const float h2 = h * h;
    const float r2 = rx * rx + ry * ry + rz * rz;
    if (r2 > h2) return 0.0f;
    const float h4 = h2 * h2;
    const float h9 = h4 * h4 * h;
    const float diff = h2 - r2;
    const float diff3 = diff * diff * diff;
    return (coeff / h9) * diff3;
}

static double kernels_cpp_poly6_f64(double rx, double ry, double rz, double h) {
    const double pi = 3.14159265358979323846;
    const double coeff = 315.0 / (64.0 * pi);
    const double h2 = h * h;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 > h2) return 0.0;
    const double h4 = h2 * h2;
    const double h9 = h4 * h4 * h;
    const double diff = h2 - r2;
    return (coeff / h9) * (diff * diff * diff);
}

static void pack_common(const fsi_fluid_config* p, const float bmin[3], const float bmax[3], float damping,
                        uint32_t n, oracle_ubo* u) {
    memset(u, 0, sizeof(*u));
    for (int k = 0; k < 3; ++k) {
        u->bounds_min[k] = bmin[k];
        u->bounds_max[k] = bmax[k];
        u->gravity[k] = p->gravity[k];
    }
    u->h = p->h;
    u->rho0 = p->rho0;
    u->eps = p->eps;
    u->k_corr = p->k_corr;
    u->n_corr = p->n_corr;
    u->viscosity = p->viscosity;
    u->bound_damping = damping;
    u->hash_size = p->hash_size;
    u->particle_count = n;
    u->enable_scorr = p->enable_scorr ? 1u : 0u;
    u->enable_viscosity = p->enable_viscosity ? 1u : 0u;
    u->cohesion_strength = p->cohesion_strength;
    u->interaction_radius = p->interaction_radius;
    u->interaction_strength = p->interaction_strength;
    u->hash_mask = p->hash_size - 1u;
    u->min_lod = p->enable_apbf ? p->min_lod : p->solver_iterations;
    u->max_lod = p->enable_apbf ? p->max_lod : p->solver_iterations;
    u->lod_max_dist = p->lod_max_dist;
    u->enable_apbf = p->enable_apbf ? 1u : 0u;
}

/* src/fluid/PBFluids.cpp:39-109 in FP32. */
void oracle_pack_ubo_f32(const fsi_fluid_config* p, const float bmin[3], const float bmax[3], float damping,
                         uint32_t n, oracle_ubo* u) {
    pack_common(p, bmin, bmax, damping, n, u);
    const float dq = p->delta_q * p->h;
    u->wq = kernels_cpp_poly6_f32(dq, 0.0f, 0.0f, p->h);
    u->sub_dt = p->dt / (float)p->substep_iterations;
    const float PI = 3.14159265358979323846f;
    const float h = p->h;
    const float h2 = h * h;
    const float h3 = h2 * h;
    const float h6 = h3 * h3;
    const float h9 = h6 * h3;
    u->w0_self = (315.0f / (64.0f * PI)) / h3;
    u->poly6_coeff = 315.0f / (64.0f * PI * h9);
    u->spiky_coeff = 45.0f / (PI * h6);
    u->inv_rho0 = 1.0f / p->rho0;
}

/* The same coefficients in FP64 (the twin). */
void oracle_pack_ubo_f64(const fsi_fluid_config* p, const float bmin[3], const float bmax[3], float damping,
                         uint32_t n, oracle_ubo* u) {
    pack_common(p, bmin, bmax, damping, n, u);
    const double h = p->h;
    const double dq = (double)p->delta_q * h;
    u->wq = kernels_cpp_poly6_f64(dq, 0.0, 0.0, h);
    u->sub_dt = (double)p->dt / (double)p->substep_iterations;
    const double PI = 3.14159265358979323846;
    const double h3 = h * h * h;
    const double h6 = h3 * h3;
    const double h9 = h6 * h3;
    u->w0_self = (315.0 / (64.0 * PI)) / h3;
    u->poly6_coeff = 315.0 / (64.0 * PI * h9);
    u->spiky_coeff = 45.0 / (PI * h6);
    u->inv_rho0 = 1.0 / (double)p->rho0;
}

/* ---- FP32 instance ---- */
#define REAL float
#define SFX f32
#define RC(x) x##f
#define SQRT sqrtf
#define FLOOR floorf
#define POW powf
#define ROUND roundf
#include "pbf_oracle_body.h"
#undef REAL
#undef SFX
#undef RC
#undef SQRT
#undef FLOOR
#undef POW
#undef ROUND

/* ---- FP64 instance ---- */
#define REAL double
#define SFX f64
#define RC(x) x
#define SQRT sqrt
#define FLOOR floor
#define POW pow
#define ROUND round
#include "pbf_oracle_body.h"
#undef REAL
#undef SFX
#undef RC
#undef SQRT
#undef FLOOR
#undef POW
#undef ROUND

/* Density-constraint error in FP64 (see pbf_oracle.h). */
void oracle_density_error_f64(const float* x4, uint32_t n, const oracle_ubo* u, double* mean_abs_c,
                              double* max_c, double* mean_rho) {
    *mean_abs_c = 0.0;
    *max_c = 0.0;
    *mean_rho = 0.0;
    if (n == 0) return;
    const uint32_t H = u->hash_size;
    oracle_state st;
    memset(&st, 0, sizeof(st));
    st.n = n;
    st.hash_size = H;
    double* sol = (double*)calloc((size_t)8 * n, sizeof(double));
    st.solver = sol;
    st.keys = (uint32_t*)malloc(sizeof(uint32_t) * n);
    st.ids = (uint32_t*)malloc(sizeof(uint32_t) * n);
    st.cells = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)H);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) sol[8 * (size_t)i + k] = x4[4 * (size_t)i + k];
        const int32_t cx = (int32_t)floor(sol[8 * (size_t)i] / u->h);
        const int32_t cy = (int32_t)floor(sol[8 * (size_t)i + 1] / u->h);
        const int32_t cz = (int32_t)floor(sol[8 * (size_t)i + 2] / u->h);
        st.keys[i] = oracle_cell_hash(cx, cy, cz, u->hash_mask);
        st.ids[i] = i;
    }
    oracle_sort_f64(&st);
    oracle_build_cells_f64(&st);
    oracle_ubo uu = *u;
    uu.enable_apbf = 0;
    oracle_lambda_f64(&uu, &st, 1);
    double sum_abs = 0.0, mx = -1e300, sum_rho = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double rho = sol[8 * (size_t)i + 7];
        double c = rho / u->rho0 - 1.0;
        if (c < -u->cohesion_strength) c = -u->cohesion_strength;
        sum_abs += fabs(c);
        if (c > mx) mx = c;
        sum_rho += rho;
    }
    *mean_abs_c = sum_abs / n;
    *max_c = mx;
    *mean_rho = sum_rho / n;
    free(sol);
    free(st.keys);
    free(st.ids);
    free(st.cells);
}
