// scene.cpp — host helpers of the C-ABI: lattice spawn, the reference's
// spawnParticles, the OFF loader and mesh fitting, icosphere mesh and the FP64
// signed-distance grid of a triangle mesh.
//
// The SDF build follows SDFBoundary::buildFromMesh (src/fluid/SDFBoundary.cpp:278-436):
//   - grid origin = bbox - 5 cells, dims = ceil(span / cell)             (:296-308)
//   - triangle list with face normals, angle-weighted vertex normals and
//     edge normals (normalised sum of the two incident face normals)     (:318-383)
//   - a median-split BVH over triangle AABBs, leaves of at most 4        (:78-265)
//   - per voxel: nearest-first branch-and-bound closest point (Ericson,
//     Real-Time Collision Detection 5.1.5) and the sign of (P - closest)
//     against the pseudo-normal of the closest feature (Baerentzen-Aanaes) (:400-419)
// The grid is written x-fastest, the order the reference uploads to its 3D
// texture (:580-598), so it can be handed to fsi_pbf_set_sdf directly.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fsi_pbf.h"
#include "mesh_bvh.h"

namespace fsi {
int fail(int code, const std::string& msg);  // pbf_core.hip: records fsi_last_error()
namespace {

// ---------------- MT19937 (Matsumoto & Nishimura, 1998) ----------------
struct Mt19937 {
    uint32_t mt[624];
    int idx = 624;
    explicit Mt19937(uint32_t seed) {
        mt[0] = seed;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    }
    uint32_t next() {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
};

using mesh::D3;
using mesh::norm;

// Particle records of spawnParticles (src/main.cpp:160-188), appended to out
// while out->size() < limit.
void spawn(const fsi_fluid_config& c, std::vector<fsi_particle>& out) {
    auto record = [&](float x, float y, float z) {
        fsi_particle pt;
        std::memset(&pt, 0, sizeof(pt));
        pt.pos[0] = x;
        pt.pos[1] = y;
        pt.pos[2] = z;
        for (int a = 0; a < 3; ++a) {
            pt.pred_pos[a] = pt.pos[a];
            pt.vel[a] = c.initial_velocity[a];
        }
        out.push_back(pt);
    };
    std::mt19937 rng(1337);
    if (c.spawn_random) {
        // frand01: uniform_real_distribution<float>(0, 1) on the same engine
        std::uniform_real_distribution<float> d(0.f, 1.f);
        const float lo[3] = {c.spawn_min[0] + c.spacing, c.spawn_min[1] + c.spacing, c.spawn_min[2] + c.spacing};
        const float hi[3] = {c.spawn_max[0] - c.spacing, c.spawn_max[1] - c.spacing, c.spawn_max[2] - c.spacing};
        for (uint32_t i = 0; i < c.particle_count; ++i) {
            const float x = lo[0] + (hi[0] - lo[0]) * d(rng);
            const float y = lo[1] + (hi[1] - lo[1]) * d(rng);
            const float z = lo[2] + (hi[2] - lo[2]) * d(rng);
            record(x, y, z);
        }
    } else {
        const float dx = std::max(1e-6f, c.spacing);
        const size_t limit = c.particle_count;
        for (float z = c.spawn_min[2]; z <= c.spawn_max[2] && out.size() < limit; z += dx)
            for (float y = c.spawn_min[1]; y <= c.spawn_max[1] && out.size() < limit; y += dx)
                for (float x = c.spawn_min[0]; x <= c.spawn_max[0] && out.size() < limit; x += dx) record(x, y, z);
    }
}

// loadOFF (src/main.cpp:364-415): comment (#) and blank lines skipped before
// the header, which must contain "OFF", and before the counts line; each
// vertex line gives its first three numbers (the rest of the line, normals or
// colours, is ignored); each face "n i0 .. i(n-1)" is fan-triangulated
// (i0, ij, ij+1).  Lines missing at the end of the file read as empty: their
// vertices are 0 and their faces add nothing.
int load_off(const char* path, std::vector<std::array<double, 3>>& verts, std::vector<std::array<int32_t, 3>>& faces) {
    std::ifstream f(path);
    if (!f) return fail(FSI_ERR_INVALID_ARG, std::string("cannot open: ") + path);
    std::string header;
    while (std::getline(f, header)) {
        if (header.empty() || header[0] == '#') continue;
        while (!header.empty() && (header.back() == '\r' || header.back() == ' ')) header.pop_back();
        break;
    }
    if (header.find("OFF") == std::string::npos) return fail(FSI_ERR_INVALID_ARG, std::string("not an OFF file: ") + path);
    int nv = 0, nf = 0, ne = 0;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        if (ss >> nv >> nf >> ne) break;
    }
    if (nv < 0 || nf < 0) return fail(FSI_ERR_INVALID_ARG, std::string("negative counts in ") + path);
    auto next_line = [&]() {
        while (std::getline(f, line))
            if (!line.empty() && line[0] != '#') return;
    };
    verts.assign((size_t)nv, std::array<double, 3>{0.0, 0.0, 0.0});
    for (int i = 0; i < nv; ++i) {
        next_line();
        std::istringstream ss(line);
        ss >> verts[(size_t)i][0] >> verts[(size_t)i][1] >> verts[(size_t)i][2];
    }
    faces.clear();
    faces.reserve((size_t)nf);
    for (int i = 0; i < nf; ++i) {
        next_line();
        std::istringstream ss(line);
        int n = 0;
        ss >> n;
        std::vector<int> poly((size_t)std::max(n, 0), 0);
        for (int& v : poly) ss >> v;
        for (int j = 1; j + 1 < n; ++j) faces.push_back({poly[0], poly[(size_t)j], poly[(size_t)j + 1]});
    }
    return FSI_OK;
}

}  // namespace
}  // namespace fsi

using namespace fsi;

extern "C" {

int fsi_spawn_particles(const fsi_fluid_config* cfg, fsi_particle* out, uint32_t cap, uint32_t* count) {
    if (!cfg || !count) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    std::vector<fsi_particle> ps;
    ps.reserve(std::min<uint32_t>(cfg->particle_count, 1u << 24));
    spawn(*cfg, ps);
    *count = (uint32_t)ps.size();
    if (!out) return FSI_OK;
    if (cap < ps.size()) return fail(FSI_ERR_CAPACITY, "cap smaller than the spawned count");
    std::memcpy(out, ps.data(), ps.size() * sizeof(fsi_particle));
    return FSI_OK;
}

int fsi_load_off(const char* path, double* verts, uint32_t vcap, int32_t* faces, uint32_t fcap, uint32_t* nverts,
                 uint32_t* nfaces) {
    if (!path || !nverts || !nfaces) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<int32_t, 3>> fc;
    const int rc = load_off(path, v, fc);
    if (rc) return rc;
    *nverts = (uint32_t)v.size();
    *nfaces = (uint32_t)fc.size();
    if (!verts && !faces) return FSI_OK;
    if (!verts || !faces) return fail(FSI_ERR_INVALID_ARG, "verts and faces must both be given");
    if (vcap < v.size() || fcap < fc.size()) return fail(FSI_ERR_CAPACITY, "verts or faces too small");
    for (size_t i = 0; i < v.size(); ++i)
        for (int k = 0; k < 3; ++k) verts[3 * i + k] = v[i][k];
    for (size_t i = 0; i < fc.size(); ++i)
        for (int k = 0; k < 3; ++k) faces[3 * i + k] = fc[i][k];
    return FSI_OK;
}

int fsi_mesh_fit_to_domain(double* verts, uint32_t nverts, const float center[3], double target_size) {
    if (!center || (nverts > 0 && !verts)) return fail(FSI_ERR_INVALID_ARG, "NULL argument");
    if (nverts == 0) return FSI_OK;
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (uint32_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], verts[3 * (size_t)i + k]);
            hi[k] = std::max(hi[k], verts[3 * (size_t)i + k]);
        }
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    const double extent = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-9});
    const double s = target_size / extent;
    for (uint32_t i = 0; i < nverts; ++i) {
        double* v = verts + 3 * (size_t)i;
        v[0] = (v[0] - cx) * s + center[0];
        v[1] = (v[1] - cy) * s + center[1];
        v[2] = (v[2] - cz) * s + center[2];
    }
    return FSI_OK;
}

int fsi_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float origin[3], float spacing, float jitter,
                      uint32_t seed, const float v0[3], fsi_particle* out) {
    if (!origin || !out) return FSI_ERR_INVALID_ARG;
    Mt19937 rng(seed);
    const float vx = v0 ? v0[0] : 0.0f, vy = v0 ? v0[1] : 0.0f, vz = v0 ? v0[2] : 0.0f;
    size_t id = 0;
    auto draw = [&]() {
        const float u = (float)(rng.next() >> 8) * (1.0f / 16777216.0f);
        return (2.0f * u - 1.0f) * jitter;
    };
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i, ++id) {
                fsi_particle& p = out[id];
                std::memset(&p, 0, sizeof(p));
                const float jx = draw(), jy = draw(), jz = draw();
                p.pos[0] = origin[0] + ((float)i + 0.5f) * spacing + jx;
                p.pos[1] = origin[1] + ((float)j + 0.5f) * spacing + jy;
                p.pos[2] = origin[2] + ((float)k + 0.5f) * spacing + jz;
                for (int a = 0; a < 3; ++a) p.pred_pos[a] = p.pos[a];
                p.vel[0] = vx;
                p.vel[1] = vy;
                p.vel[2] = vz;
            }
    return FSI_OK;
}

int fsi_mesh_icosphere(uint32_t subdivisions, const double centre[3], double radius, double* verts,
                       int32_t* faces) {
    if (!centre || !verts || !faces || subdivisions > 9) return FSI_ERR_INVALID_ARG;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (D3& p : v) p = p * (1.0 / norm(p));
    std::vector<std::array<int, 3>> f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                         {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (uint32_t s = 0; s < subdivisions; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const std::pair<int, int> key(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            D3 m = (v[a] + v[b]) * 0.5;
            m = m * (1.0 / norm(m));
            v.push_back(m);
            const int idx = (int)v.size() - 1;
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto& tri : f) {
            const int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); ++i) {
        verts[3 * i + 0] = centre[0] + radius * v[i].x;
        verts[3 * i + 1] = centre[1] + radius * v[i].y;
        verts[3 * i + 2] = centre[2] + radius * v[i].z;
    }
    for (size_t i = 0; i < f.size(); ++i)
        for (int k = 0; k < 3; ++k) faces[3 * i + k] = f[i][k];
    return FSI_OK;
}

int fsi_sdf_build(const double* verts, uint32_t nverts, const int32_t* faces, uint32_t nfaces, double cell,
                  double origin[3], int32_t dims[3], float* grid) {
    if (!verts || !faces || !origin || !dims || nverts == 0 || nfaces == 0 || !(cell > 0.0))
        return FSI_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < 3 * nfaces; ++i)
        if (faces[i] < 0 || (uint32_t)faces[i] >= nverts) return FSI_ERR_INVALID_ARG;
    D3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30};
    for (uint32_t i = 0; i < nverts; ++i) {
        const D3 p{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    const double buffer = cell * 5.0;
    const D3 o{lo.x - buffer, lo.y - buffer, lo.z - buffer};
    const D3 span{(hi.x + buffer) - o.x, (hi.y + buffer) - o.y, (hi.z + buffer) - o.z};
    int32_t dm[3] = {(int32_t)std::ceil(span.x / cell), (int32_t)std::ceil(span.y / cell),
                     (int32_t)std::ceil(span.z / cell)};
    origin[0] = o.x;
    origin[1] = o.y;
    origin[2] = o.z;
    for (int k = 0; k < 3; ++k) dims[k] = dm[k];
    if (!grid) return FSI_OK;

    mesh::Mesh m;
    m.build(verts, nverts, faces, nfaces);
    const int W = dm[0], H = dm[1], D = dm[2];
    unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w) {
        pool.emplace_back([&, w]() {
            for (int z = (int)w; z < D; z += (int)nthreads)
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x < W; ++x) {
                        const D3 p{o.x + x * cell, o.y + y * cell, o.z + z * cell};
                        grid[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * (size_t)z)] = m.signed_distance(p);
                    }
        });
    }
    for (std::thread& t : pool) t.join();
    return FSI_OK;
}

}  // extern "C"
