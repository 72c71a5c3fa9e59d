// Drives the C++ PBFluids / SDFBoundary mirror the way the reference's
// src/main.cpp drives its PBFluids (construct :1116, setCollisionPadding
// :1117, setBounds :1199, setSDFBoundary / setSDFPadding :449-451,
// setParticles :950, step :1702, filterContactCandidates :1707,
// readbackContactCandidates :615), on a small dam break beside an icosphere
// obstacle, and writes the particles and the contact candidates as raw
// records for tests/test_gpu_cpp_mirror.py to compare with the C-ABI run.
//
// usage: pbfluids_driver <particles.bin> <contacts.bin> <frames>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../realtime-fsi-engine_amd/host/PBFluids.h"
#include "fsi_pbf.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s particles.bin contacts.bin frames\n", argv[0]);
        return 2;
    }
    const int frames = std::atoi(argv[3]);

    FluidConfig cfg;
    cfg.dt = 1.0f / 120.0f;
    cfg.gravity = make_pvec3(0.0f, -9.81f, 0.0f);
    cfg.h = 0.1f;
    cfg.rho0 = 6378.0f;
    cfg.solverIterations = 3;
    cfg.substepIterations = 2;
    cfg.eps = 600.0f;
    cfg.boundsMin = make_pvec3(0.0f, 0.0f, 0.0f);
    cfg.boundsMax = make_pvec3(2.4f, 1.6f, 1.6f);
    cfg.hashSize = 1u << 14;

    // a 16 x 12 x 12 lattice block at spacing 0.05, ids x fastest
    std::vector<Particle> ps;
    for (int z = 0; z < 12; ++z)
        for (int y = 0; y < 12; ++y)
            for (int x = 0; x < 16; ++x) {
                Particle p;
                p.pos = make_pvec3(0.025f + 0.05f * (float)x, 0.025f + 0.05f * (float)y, 0.025f + 0.05f * (float)z);
                p.predPos = p.pos;
                p.vel = make_pvec3(0.0f, 0.0f, 0.0f);
                ps.push_back(p);
            }

    // obstacle: subdivision-2 icosphere of radius 0.25 at (0.9, 0.2, 0.3), at the
    // block's front face
    const int sub = 2, nv = 10 * 16 + 2, nf = 20 * 16;
    std::vector<double> v(3 * (size_t)nv);
    std::vector<int32_t> f(3 * (size_t)nf);
    const double centre[3] = {0.9, 0.2, 0.3};
    if (fsi_mesh_icosphere(sub, centre, 0.25, v.data(), f.data()) != FSI_OK) return 1;
    std::vector<std::array<double, 3>> verts((size_t)nv);
    std::vector<std::array<int, 3>> faces((size_t)nf);
    for (int i = 0; i < nv; ++i) verts[(size_t)i] = {v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]};
    for (int i = 0; i < nf; ++i) faces[(size_t)i] = {f[3 * (size_t)i], f[3 * (size_t)i + 1], f[3 * (size_t)i + 2]};
    SDFBoundary sdf;
    sdf.buildFromMesh(verts, faces, 0.05, false);
    if (!sdf.valid()) return 1;

    PBFluids fluid(cfg);
    fluid.setCollisionPadding(0.0f);
    fluid.setBounds(cfg.boundsMin, cfg.boundsMax, 0.5f);
    fluid.setSDFBoundary(&sdf);
    fluid.setSDFPadding(0.025f);
    fluid.setSDFContainment(false);
    fluid.setParticles(ps);
    for (int i = 0; i < frames; ++i) fluid.step();
    fluid.filterContactCandidates(0.1f);
    std::vector<Particle> out;
    fluid.readbackParticles(out);
    std::vector<ContactCandidate> contacts;
    fluid.readbackContactCandidates(contacts);
    if (fluid.lastStatus() != 0) return 1;

    FILE* fp = std::fopen(argv[1], "wb");
    if (!fp || std::fwrite(out.data(), sizeof(Particle), out.size(), fp) != out.size()) return 1;
    std::fclose(fp);
    FILE* fc = std::fopen(argv[2], "wb");
    if (!fc) return 1;
    if (!contacts.empty() && std::fwrite(contacts.data(), sizeof(ContactCandidate), contacts.size(), fc) != contacts.size())
        return 1;
    std::fclose(fc);
    std::printf("particles %zu contacts %zu\n", out.size(), contacts.size());
    return 0;
}
