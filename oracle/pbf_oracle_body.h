/*
 * pbf_oracle_body.h — body of the CPU oracle, included twice by pbf_oracle.c
 * with REAL / SFX / RC / SQRT / FLOOR / POW defined (TEST INFRASTRUCTURE ONLY).
 *
 * Every expression keeps the operand order of the GLSL it restates, so that
 * with contraction off and IEEE division / square root the FP32 instance
 * computes the same bits as a conforming FP32 evaluation of the shader.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)
#define V3 FN(ov3)

typedef struct {
    REAL x, y, z;
} V3;

static inline V3 FN(mk)(REAL x, REAL y, REAL z) {
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
static inline V3 FN(add)(V3 a, V3 b) { return FN(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(sub)(V3 a, V3 b) { return FN(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(scl)(V3 a, REAL s) { return FN(mk)(a.x * s, a.y * s, a.z * s); }
static inline V3 FN(dvs)(V3 a, REAL s) { return FN(mk)(a.x / s, a.y / s, a.z / s); }
static inline REAL FN(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline REAL FN(len)(V3 a) { return SQRT(FN(dot)(a, a)); }
static inline V3 FN(cross)(V3 a, V3 b) {
    return FN(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

static inline V3 FN(ld3)(const REAL* p, uint32_t i) { return FN(mk)(p[4 * i], p[4 * i + 1], p[4 * i + 2]); }
static inline void FN(st3)(REAL* p, uint32_t i, V3 v) {
    p[4 * i] = v.x;
    p[4 * i + 1] = v.y;
    p[4 * i + 2] = v.z;
}

/* ---- kernels (CS_ComputeLambdas.glsl:74-92, identical in the other shaders) ---- */

static inline REAL FN(poly6)(const oracle_ubo* u, V3 r) {
    const REAL h = (REAL)u->h;
    const REAL r2 = FN(dot)(r, r);
    const REAL h2 = h * h;
    if (r2 > h2) return (REAL)0;
    const REAL diff = h2 - r2;
    return (REAL)u->poly6_coeff * diff * diff * diff;
}

static inline V3 FN(spiky)(const oracle_ubo* u, V3 r) {
    const REAL h = (REAL)u->h;
    const REAL r2 = FN(dot)(r, r);
    const REAL h2 = h * h;
    if (r2 > h2 || r2 < RC(1e-10)) return FN(mk)(0, 0, 0);
    const REAL dist = SQRT(r2);
    const REAL diff = h - dist;
    const REAL scalar = -(REAL)u->spiky_coeff * diff * diff / dist;
    return FN(scl)(r, scalar);
}

/* x^n for the s_corr term (CS_ComputeDeltaP.glsl:103). */
static inline REAL FN(powr)(REAL x, REAL n) {
    if (n >= 0 && n <= 64 && n == FLOOR(n)) {
        uint32_t e = (uint32_t)n;
        REAL result = (REAL)1;
        REAL base = x;
        while (e) {
            if (e & 1u) result = result * base;
            e >>= 1u;
            if (e) base = base * base;
        }
        return result;
    }
    return POW(x, n);
}

/* CS_ComputeDeltaP.glsl:94-104 */
static inline REAL FN(scorr)(const oracle_ubo* u, V3 r) {
    if (u->enable_scorr == 0u) return (REAL)0;
    const REAL w = FN(poly6)(u, r);
    const REAL wq = (REAL)u->wq;
    if (wq <= RC(1e-12) || w <= (REAL)0) return (REAL)0;
    const REAL ratio = w / wq;
    if (ratio <= (REAL)0) return (REAL)0;
    return -(REAL)u->k_corr * FN(powr)(ratio, (REAL)u->n_corr);
}

/* ivec3(floor(x / h)) (CS_PredictAndHash.glsl:125) */
static inline void FN(cell_of)(const oracle_ubo* u, V3 x, int32_t c[3]) {
    const REAL h = (REAL)u->h;
    c[0] = (int32_t)FLOOR(x.x / h);
    c[1] = (int32_t)FLOOR(x.y / h);
    c[2] = (int32_t)FLOOR(x.z / h);
}

/* ---- SDF texture emulation (CS_Integrate.glsl:86-99, Texture3D.cpp:20-35) ---- */

static inline REAL FN(texel)(const oracle_uniforms* s, int32_t i, int32_t j, int32_t k) {
    const int32_t W = s->sdf_dims[0], H = s->sdf_dims[1];
    return (REAL)s->sdf_grid[(size_t)i + (size_t)W * ((size_t)j + (size_t)H * (size_t)k)];
}

/* One axis of GL_LINEAR + CLAMP_TO_EDGE: u = tc * size - 0.5, i0 = floor(u),
 * alpha = frac(u), both taps clamped to [0, size-1]. */
static inline void FN(taps)(REAL tc, int32_t size, int32_t* i0, int32_t* i1, REAL* a) {
    const REAL uu = tc * (REAL)size - RC(0.5);
    REAL f = FLOOR(uu);
    *a = uu - f;
    if (!(f >= (REAL)-1)) f = (REAL)-1; /* also catches NaN */
    if (f > (REAL)size) f = (REAL)size;
    int32_t lo = (int32_t)f, hi = lo + 1;
    lo = lo < 0 ? 0 : (lo > size - 1 ? size - 1 : lo);
    hi = hi < 0 ? 0 : (hi > size - 1 ? size - 1 : hi);
    *i0 = lo;
    *i1 = hi;
}

static inline REAL FN(lerp)(REAL t0, REAL t1, REAL a) { return ((REAL)1 - a) * t0 + a * t1; }

static REAL FN(query)(const oracle_uniforms* s, V3 p) {
    const REAL cell = (REAL)s->sdf_cell;
    const REAL ox = (REAL)s->sdf_origin[0], oy = (REAL)s->sdf_origin[1], oz = (REAL)s->sdf_origin[2];
    const REAL rx = (REAL)s->sdf_dims[0], ry = (REAL)s->sdf_dims[1], rz = (REAL)s->sdf_dims[2];
    const REAL tx = ((p.x - ox) / cell + RC(0.5)) / rx;
    const REAL ty = ((p.y - oy) / cell + RC(0.5)) / ry;
    const REAL tz = ((p.z - oz) / cell + RC(0.5)) / rz;
    int32_t x0, x1, y0, y1, z0, z1;
    REAL ax, ay, az;
    FN(taps)(tx, s->sdf_dims[0], &x0, &x1, &ax);
    FN(taps)(ty, s->sdf_dims[1], &y0, &y1, &ay);
    FN(taps)(tz, s->sdf_dims[2], &z0, &z1, &az);
    const REAL c00 = FN(lerp)(FN(texel)(s, x0, y0, z0), FN(texel)(s, x1, y0, z0), ax);
    const REAL c10 = FN(lerp)(FN(texel)(s, x0, y1, z0), FN(texel)(s, x1, y1, z0), ax);
    const REAL c01 = FN(lerp)(FN(texel)(s, x0, y0, z1), FN(texel)(s, x1, y0, z1), ax);
    const REAL c11 = FN(lerp)(FN(texel)(s, x0, y1, z1), FN(texel)(s, x1, y1, z1), ax);
    const REAL c0 = FN(lerp)(c00, c10, ay);
    const REAL c1 = FN(lerp)(c01, c11, ay);
    return FN(lerp)(c0, c1, az);
}

/* sdfGradient: central differences with step cellSize (CS_ApplyDeltaP.glsl:58-65) */
static V3 FN(grad)(const oracle_uniforms* s, V3 p) {
    const REAL e = (REAL)s->sdf_cell;
    const REAL z = (REAL)0;
    const REAL gx = FN(query)(s, FN(mk)(p.x + e, p.y + z, p.z + z)) - FN(query)(s, FN(mk)(p.x - e, p.y - z, p.z - z));
    const REAL gy = FN(query)(s, FN(mk)(p.x + z, p.y + e, p.z + z)) - FN(query)(s, FN(mk)(p.x - z, p.y - e, p.z - z));
    const REAL gz = FN(query)(s, FN(mk)(p.x + z, p.y + z, p.z + e)) - FN(query)(s, FN(mk)(p.x - z, p.y - z, p.z - e));
    return FN(dvs)(FN(mk)(gx, gy, gz), (REAL)2 * e);
}

double FN(oracle_sdf_query)(const oracle_uniforms* s, const double p[3]) {
    return (double)FN(query)(s, FN(mk)((REAL)p[0], (REAL)p[1], (REAL)p[2]));
}
void FN(oracle_sdf_gradient)(const oracle_uniforms* s, const double p[3], double out[3]) {
    const V3 g = FN(grad)(s, FN(mk)((REAL)p[0], (REAL)p[1], (REAL)p[2]));
    out[0] = g.x;
    out[1] = g.y;
    out[2] = g.z;
}
double FN(oracle_poly6)(const oracle_ubo* u, const double r[3]) {
    return (double)FN(poly6)(u, FN(mk)((REAL)r[0], (REAL)r[1], (REAL)r[2]));
}
void FN(oracle_grad_spiky)(const oracle_ubo* u, const double r[3], double out[3]) {
    const V3 g = FN(spiky)(u, FN(mk)((REAL)r[0], (REAL)r[1], (REAL)r[2]));
    out[0] = g.x;
    out[1] = g.y;
    out[2] = g.z;
}

/* ---- passes ---- */

/* CS_PredictAndHash.glsl:88-130 */
void FN(oracle_predict_hash)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    REAL* pos = (REAL*)st->pos;
    REAL* vel = (REAL*)st->vel;
    REAL* sol = (REAL*)st->solver;
    const REAL dt = (REAL)u->sub_dt;
    const V3 g = FN(mk)((REAL)u->gravity[0], (REAL)u->gravity[1], (REAL)u->gravity[2]);
    const V3 ip = FN(mk)((REAL)s->interaction_pos[0], (REAL)s->interaction_pos[1], (REAL)s->interaction_pos[2]);
    const REAL irad = (REAL)u->interaction_radius, istr = (REAL)u->interaction_strength;
    const int64_t n = st->n;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        const V3 p = FN(ld3)(pos, id);
        V3 v = FN(ld3)(vel, id);
        v = FN(add)(v, FN(scl)(g, dt));
        if (s->is_interacting == 1u) {
            const V3 to = FN(sub)(ip, p);
            const REAL dist = FN(len)(to);
            if (dist < irad && dist > RC(1e-5)) {
                const REAL falloff = (REAL)1 - (dist / irad);
                const V3 nrm = FN(dvs)(to, FN(len)(to));
                v = FN(add)(v, FN(scl)(FN(scl)(FN(scl)(nrm, istr), falloff), dt));
            }
        }
        const V3 xp = FN(add)(p, FN(scl)(v, dt));
        FN(st3)(vel, id, v);
        sol[8 * (size_t)id + 0] = xp.x;
        sol[8 * (size_t)id + 1] = xp.y;
        sol[8 * (size_t)id + 2] = xp.z;
        sol[8 * (size_t)id + 3] = (REAL)0;
        int32_t c[3];
        FN(cell_of)(u, xp, c);
        st->keys[id] = oracle_cell_hash(c[0], c[1], c[2], u->hash_mask);
        st->ids[id] = id;
    }
}

/* CS_RadixHistogram/PrefixSum/Scatter.glsl with PBFluids.cpp:355-404: an LSD
 * radix sort of (key, id) pairs emitted in id order, i.e. a stable sort by key.
 * Restated as a stable counting sort (keys < hash_size). */
void FN(oracle_sort)(oracle_state* st) {
    const uint32_t n = st->n, H = st->hash_size;
    uint32_t* count = (uint32_t*)calloc((size_t)H + 1, sizeof(uint32_t));
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* i2 = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) count[st->keys[i] + 1]++;
    for (uint32_t b = 0; b < H; ++b) count[b + 1] += count[b];
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t d = count[st->keys[i]]++;
        k2[d] = st->keys[i];
        i2[d] = st->ids[i];
    }
    memcpy(st->keys, k2, sizeof(uint32_t) * n);
    memcpy(st->ids, i2, sizeof(uint32_t) * n);
    free(count);
    free(k2);
    free(i2);
}

/* CS_BuildOffsets.glsl:55-78 (clear, then first/last of every key run) */
void FN(oracle_build_cells)(oracle_state* st) {
    const uint32_t n = st->n;
    for (size_t b = 0; b < 2 * (size_t)st->hash_size; ++b) st->cells[b] = -1;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t k = st->keys[i];
        if (i == 0 || st->keys[i - 1] != k) st->cells[2 * (size_t)k] = (int32_t)i;
        if (i == n - 1 || st->keys[i + 1] != k) st->cells[2 * (size_t)k + 1] = (int32_t)(i + 1);
    }
}

/* CS_ComputeLOD.glsl:66-79 */
void FN(oracle_lod)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    const REAL* sol = (const REAL*)st->solver;
    const V3 cam = FN(mk)((REAL)s->camera_pos[0], (REAL)s->camera_pos[1], (REAL)s->camera_pos[2]);
    const REAL lo = (REAL)u->min_lod, hi = (REAL)u->max_lod;
    for (uint32_t id = 0; id < st->n; ++id) {
        const V3 x = FN(mk)(sol[8 * (size_t)id], sol[8 * (size_t)id + 1], sol[8 * (size_t)id + 2]);
        const REAL dist = FN(len)(FN(sub)(cam, x));
        REAL t = dist / (REAL)u->lod_max_dist;
        t = t < (REAL)0 ? (REAL)0 : (t > (REAL)1 ? (REAL)1 : t);
        const REAL m = hi * ((REAL)1 - t) + lo * t;
        uint32_t l = (uint32_t)ROUND(m);
        l = l < u->min_lod ? u->min_lod : (l > u->max_lod ? u->max_lod : l);
        st->lod[id] = l;
    }
}

static inline int FN(inactive)(const oracle_ubo* u, const oracle_state* st, uint32_t id, uint32_t iter) {
    return u->enable_apbf != 0u && st->lod[id] < iter;
}

/* CS_ComputeLambdas.glsl:108-162 */
void FN(oracle_lambda)(const oracle_ubo* u, oracle_state* st, uint32_t iter) {
    REAL* sol = (REAL*)st->solver;
    const REAL inv = (REAL)u->inv_rho0, rho0 = (REAL)u->rho0, h = (REAL)u->h;
    const int64_t n = st->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        if (FN(inactive)(u, st, id, iter)) continue;
        const V3 xi = FN(mk)(sol[8 * (size_t)id], sol[8 * (size_t)id + 1], sol[8 * (size_t)id + 2]);
        REAL rho = (REAL)u->w0_self;
        REAL sum_g2 = 0;
        V3 gi = FN(mk)(0, 0, 0);
        uint32_t ncand = 0, nnbr = 0;
        const REAL h2 = h * h;
        int32_t c[3];
        FN(cell_of)(u, xi, c);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint32_t b = oracle_cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, u->hash_mask);
                    const int32_t s0 = st->cells[2 * (size_t)b], s1 = st->cells[2 * (size_t)b + 1];
                    if (s0 == -1) continue;
                    for (int32_t k = s0; k < s1; ++k) {
                        const uint32_t j = st->ids[k];
                        if (j == id) continue;
                        const V3 xj = FN(mk)(sol[8 * (size_t)j], sol[8 * (size_t)j + 1], sol[8 * (size_t)j + 2]);
                        const V3 r = FN(sub)(xi, xj);
                        ++ncand;
                        if (FN(dot)(r, r) <= h2) ++nnbr;
                        rho = rho + FN(poly6)(u, r);
                        const V3 gw = FN(spiky)(u, r);
                        const REAL gj2 = FN(dot)(gw, gw);
                        sum_g2 = sum_g2 + gj2 * (inv * inv);
                        gi = FN(add)(gi, gw);
                    }
                }
        gi = FN(scl)(gi, inv);
        REAL C = rho / rho0 - (REAL)1;
        const REAL floorc = -(REAL)u->cohesion_strength;
        C = C > floorc ? C : floorc; /* GLSL max(C, -cohesion) */
        sum_g2 = sum_g2 + FN(dot)(gi, gi);
        const REAL lambda = -C / (sum_g2 + (REAL)u->eps);
        sol[8 * (size_t)id + 3] = lambda;
        sol[8 * (size_t)id + 7] = rho;
        if (st->cand_count) st->cand_count[id] = ncand;
        if (st->nbr_count) st->nbr_count[id] = nnbr;
    }
}

/* CS_ComputeDeltaP.glsl:120-166 */
void FN(oracle_deltap)(const oracle_ubo* u, oracle_state* st, uint32_t iter) {
    REAL* sol = (REAL*)st->solver;
    const REAL inv = (REAL)u->inv_rho0;
    const int64_t n = st->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        if (FN(inactive)(u, st, id, iter)) continue;
        const V3 xi = FN(mk)(sol[8 * (size_t)id], sol[8 * (size_t)id + 1], sol[8 * (size_t)id + 2]);
        const REAL li = sol[8 * (size_t)id + 3];
        V3 dp = FN(mk)(0, 0, 0);
        int32_t c[3];
        FN(cell_of)(u, xi, c);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint32_t b = oracle_cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, u->hash_mask);
                    const int32_t s0 = st->cells[2 * (size_t)b], s1 = st->cells[2 * (size_t)b + 1];
                    if (s0 == -1) continue;
                    for (int32_t k = s0; k < s1; ++k) {
                        const uint32_t j = st->ids[k];
                        if (j == id) continue;
                        const V3 xj = FN(mk)(sol[8 * (size_t)j], sol[8 * (size_t)j + 1], sol[8 * (size_t)j + 2]);
                        const REAL lj = sol[8 * (size_t)j + 3];
                        const V3 r = FN(sub)(xi, xj);
                        const V3 gw = FN(spiky)(u, r);
                        const REAL sc = FN(scorr)(u, r);
                        dp = FN(add)(dp, FN(scl)(gw, li + lj + sc));
                    }
                }
        dp = FN(scl)(dp, inv);
        sol[8 * (size_t)id + 4] = dp.x;
        sol[8 * (size_t)id + 5] = dp.y;
        sol[8 * (size_t)id + 6] = dp.z;
    }
}

/* SDF projection shared by CS_ApplyDeltaP.glsl:108-121 and CS_Integrate.glsl:216-240.
 * Returns 1 and the push direction when the particle was moved. */
static inline int FN(sdf_project)(const oracle_uniforms* s, V3* p, V3* nrm_out) {
    const REAL pad = (REAL)s->sdf_padding;
    const REAL d = FN(query)(s, *p);
    const V3 g = FN(grad)(s, *p);
    const REAL len = FN(len)(g);
    if (len > RC(1e-6)) {
        if (s->sdf_mode == 1u) {
            if (d > -pad) {
                const V3 nrm = FN(dvs)(FN(mk)(-g.x, -g.y, -g.z), len);
                *p = FN(add)(*p, FN(scl)(nrm, d + pad));
                *nrm_out = nrm;
                return 1;
            }
        } else {
            if (d < pad) {
                const V3 nrm = FN(dvs)(g, len);
                *p = FN(add)(*p, FN(scl)(nrm, pad - d));
                *nrm_out = nrm;
                return 1;
            }
        }
    }
    return 0;
}

/* CS_ApplyDeltaP.glsl:80-124 */
void FN(oracle_apply)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st, uint32_t iter) {
    REAL* sol = (REAL*)st->solver;
    const REAL h = (REAL)u->h;
    const int64_t n = st->n;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        if (FN(inactive)(u, st, id, iter)) continue;
        V3 dp = FN(mk)(sol[8 * (size_t)id + 4], sol[8 * (size_t)id + 5], sol[8 * (size_t)id + 6]);
        const REAL maxdp = RC(0.5) * h;
        const REAL l2 = FN(dot)(dp, dp);
        const REAL max2 = maxdp * maxdp;
        if (l2 > max2) {
            const REAL l = SQRT(l2);
            dp = FN(scl)(dp, maxdp / l);
        }
        V3 p = FN(add)(FN(mk)(sol[8 * (size_t)id], sol[8 * (size_t)id + 1], sol[8 * (size_t)id + 2]), dp);
        if (s->enable_sdf == 1u) {
            V3 nrm;
            (void)FN(sdf_project)(s, &p, &nrm);
        }
        sol[8 * (size_t)id + 0] = p.x;
        sol[8 * (size_t)id + 1] = p.y;
        sol[8 * (size_t)id + 2] = p.z;
    }
}

/* CS_Integrate.glsl:135-280, neighbour velocities read from the post-predict
 * snapshot (Jacobi form of the racy read at :171). */
void FN(oracle_integrate)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    REAL* pos = (REAL*)st->pos;
    REAL* vel = (REAL*)st->vel;
    REAL* sol = (REAL*)st->solver;
    REAL* vs = (REAL*)st->vsnap;
    const int64_t n = st->n;
    memcpy(vs, vel, sizeof(REAL) * 4 * (size_t)n);
    const REAL inv = (REAL)u->inv_rho0;
    const int do_visc = u->enable_viscosity == 1u;
    const int do_vort = s->enable_vorticity == 1u;
    const V3 bmin = FN(mk)((REAL)u->bounds_min[0], (REAL)u->bounds_min[1], (REAL)u->bounds_min[2]);
    const V3 bmax = FN(mk)((REAL)u->bounds_max[0], (REAL)u->bounds_max[1], (REAL)u->bounds_max[2]);
    const REAL damp = (REAL)u->bound_damping;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        const V3 old_pos = FN(ld3)(pos, id);
        const V3 old_vel = FN(ld3)(vs, id);
        V3 xp = FN(mk)(sol[8 * (size_t)id], sol[8 * (size_t)id + 1], sol[8 * (size_t)id + 2]);
        V3 visc_delta = FN(mk)(0, 0, 0);
        if (do_visc || do_vort) {
            V3 visc = FN(mk)(0, 0, 0);
            V3 curl = FN(mk)(0, 0, 0);
            int32_t c[3];
            FN(cell_of)(u, xp, c);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const uint32_t b = oracle_cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, u->hash_mask);
                        const int32_t s0 = st->cells[2 * (size_t)b], s1 = st->cells[2 * (size_t)b + 1];
                        if (s0 == -1) continue;
                        for (int32_t k = s0; k < s1; ++k) {
                            const uint32_t j = st->ids[k];
                            if (j == id) continue;
                            const V3 xj = FN(mk)(sol[8 * (size_t)j], sol[8 * (size_t)j + 1], sol[8 * (size_t)j + 2]);
                            const V3 vj = FN(ld3)(vs, j);
                            const V3 r = FN(sub)(xp, xj);
                            const V3 vij = FN(sub)(vj, old_vel);
                            if (do_visc) {
                                const REAL w = FN(poly6)(u, r);
                                visc = FN(add)(visc, FN(scl)(FN(scl)(vij, w), inv));
                            }
                            if (do_vort) {
                                const V3 gw = FN(spiky)(u, r);
                                curl = FN(add)(curl, FN(cross)(vij, gw));
                            }
                        }
                    }
            if (do_visc) visc_delta = FN(scl)(visc, (REAL)u->viscosity);
            if (do_vort) {
                const V3 cw = FN(scl)(curl, inv);
                sol[8 * (size_t)id + 4] = cw.x;
                sol[8 * (size_t)id + 5] = cw.y;
                sol[8 * (size_t)id + 6] = cw.z;
            }
        }
        int cx = 0, cy = 0, cz = 0;
        if (xp.x <= bmin.x) { xp.x = bmin.x; cx = 1; }
        if (xp.x >= bmax.x) { xp.x = bmax.x; cx = 1; }
        if (xp.y <= bmin.y) { xp.y = bmin.y; cy = 1; }
        if (xp.y >= bmax.y) { xp.y = bmax.y; cy = 1; }
        if (xp.z <= bmin.z) { xp.z = bmin.z; cz = 1; }
        if (xp.z >= bmax.z) { xp.z = bmax.z; cz = 1; }
        int hit = 0;
        V3 nrm = FN(mk)(0, 0, 0);
        if (s->enable_sdf == 1u) hit = FN(sdf_project)(s, &xp, &nrm);
        const REAL sub_dt = (REAL)u->sub_dt;
        const REAL inv_dt = (sub_dt > (REAL)0) ? ((REAL)1 / sub_dt) : (REAL)0;
        V3 v = FN(scl)(FN(sub)(xp, old_pos), inv_dt);
        v = FN(add)(v, visc_delta);
        const REAL sp2 = FN(dot)(v, v);
        if (sp2 > RC(2500.0)) v = FN(scl)(v, RC(50.0) / SQRT(sp2));
        if (cx) v.x = v.x * -damp;
        if (cy) v.y = v.y * -damp;
        if (cz) v.z = v.z * -damp;
        if (hit) {
            const REAL vn = FN(dot)(v, nrm);
            v = FN(sub)(v, FN(scl)(nrm, vn));
        }
        FN(st3)(pos, id, xp);
        FN(st3)(vel, id, v);
        sol[8 * (size_t)id + 3] = (REAL)0;
    }
}

/* CS_VorticityConfinement.glsl:133-167 (pass 1, the only one dispatched,
 * PBFluids.cpp:499-505).  Reads committed positions and the curl left in
 * deltaP_rho.xyz by the last Integrate; writes only its own velocity. */
void FN(oracle_vorticity)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    REAL* pos = (REAL*)st->pos;
    REAL* vel = (REAL*)st->vel;
    const REAL* sol = (const REAL*)st->solver;
    const REAL inv = (REAL)u->inv_rho0;
    const REAL veps = (REAL)s->vorticity_epsilon;
    const REAL dt = (REAL)u->sub_dt;
    const int64_t n = st->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < n; ++ii) {
        const uint32_t id = (uint32_t)ii;
        const V3 pi = FN(ld3)(pos, id);
        int32_t c[3];
        FN(cell_of)(u, pi, c);
        V3 eta = FN(mk)(0, 0, 0);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint32_t b = oracle_cell_hash(c[0] + dx, c[1] + dy, c[2] + dz, u->hash_mask);
                    const int32_t s0 = st->cells[2 * (size_t)b], s1 = st->cells[2 * (size_t)b + 1];
                    if (s0 == -1) continue;
                    for (int32_t k = s0; k < s1; ++k) {
                        const uint32_t j = st->ids[k];
                        if (j == id) continue;
                        const V3 r = FN(sub)(pi, FN(ld3)(pos, j));
                        const V3 gw = FN(spiky)(u, r);
                        const V3 wj = FN(mk)(sol[8 * (size_t)j + 4], sol[8 * (size_t)j + 5], sol[8 * (size_t)j + 6]);
                        const REAL om = FN(len)(wj);
                        eta = FN(add)(eta, FN(scl)(gw, om));
                    }
                }
        eta = FN(scl)(eta, inv);
        const REAL el = FN(len)(eta);
        if (el < RC(1e-6)) continue;
        const V3 nn = FN(dvs)(eta, el);
        const V3 om = FN(mk)(sol[8 * (size_t)id + 4], sol[8 * (size_t)id + 5], sol[8 * (size_t)id + 6]);
        const V3 dv = FN(scl)(FN(scl)(FN(cross)(nn, om), veps), dt);
        FN(st3)(vel, id, FN(add)(FN(ld3)(vel, id), dv));
    }
}

/* One substep of PBFluids::step (src/fluid/PBFluids.cpp:344-496). */
void FN(oracle_substep)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    FN(oracle_predict_hash)(u, s, st);
    FN(oracle_sort)(st);
    FN(oracle_build_cells)(st);
    if (u->enable_apbf) FN(oracle_lod)(u, s, st);
    const uint32_t max_iter = u->enable_apbf ? u->max_lod : s->iterations;
    for (uint32_t it = 1; it <= max_iter; ++it) {
        FN(oracle_lambda)(u, st, it);
        FN(oracle_deltap)(u, st, it);
        FN(oracle_apply)(u, s, st, it);
    }
    FN(oracle_integrate)(u, s, st);
}

/* PBFluids::step (src/fluid/PBFluids.cpp:323-517). */
void FN(oracle_step)(const oracle_ubo* u, const oracle_uniforms* s, oracle_state* st) {
    if (st->n == 0) return;
    for (uint32_t sub = 0; sub < s->substeps; ++sub) FN(oracle_substep)(u, s, st);
    if (s->enable_vorticity) FN(oracle_vorticity)(u, s, st);
}

/* CS_ContactFilter.glsl:68-83, compacted in id order. */
uint32_t FN(oracle_contact_filter)(const oracle_ubo* u, const oracle_uniforms* s, const oracle_state* st,
                                   double band, fsi_contact_candidate* out, uint32_t cap) {
    (void)u;
    const REAL* pos = (const REAL*)st->pos;
    const REAL* vel = (const REAL*)st->vel;
    const REAL bd = (REAL)band;
    uint32_t count = 0;
    for (uint32_t id = 0; id < st->n; ++id) {
        const V3 p = FN(ld3)(pos, id);
        const REAL d = FN(query)(s, p);
        const REAL ad = d < (REAL)0 ? -d : d;
        if (ad < bd) {
            if (count < st->n && count < cap) {
                fsi_contact_candidate* c = &out[count];
                c->pos[0] = (float)p.x;
                c->pos[1] = (float)p.y;
                c->pos[2] = (float)p.z;
                c->pos[3] = 0.0f;
                c->vel[0] = (float)vel[4 * (size_t)id];
                c->vel[1] = (float)vel[4 * (size_t)id + 1];
                c->vel[2] = (float)vel[4 * (size_t)id + 2];
                c->vel[3] = (float)vel[4 * (size_t)id + 3];
            }
            ++count;
        }
    }
    return count;
}

#undef V3
#undef FN
#undef CAT
#undef CAT_
