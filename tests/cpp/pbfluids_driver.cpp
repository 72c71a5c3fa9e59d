// Drives the C++ PBFluids / SDFBoundary mirror the way the reference's
// src/main.cpp drives its PBFluids (construct :1116, setCollisionPadding
// :1117, setBounds :1199, setSDFBoundary / setSDFPadding :449-451,
// setParticles :950, step :1702, filterContactCandidates :1707,
// readbackContactCandidates :615), on a small dam break beside an icosphere
// obstacle, and writes the particles and the contact candidates as raw
// records for tests/test_gpu_cpp_mirror.py to compare with the C-ABI run.
// Then `couple` frames of the deforming-obstacle loop of syncDeformedMesh
// (src/main.cpp:803-860) and computeContactForces (:604-710): a prescribed
// dent of the sphere's upper half, SDFBoundary::refitDirty +
// uploadDirtyRegion (every third frame refitFromMesh + uploadToGPU), step,
// contact filter, and at the end SDFBoundary::closestTriangle of every
// candidate.  Written to <dir>: particles.bin, contacts.bin (the first
// part), coupled_particles.bin, coupled_contacts.bin, closest.bin (per
// candidate int32 face + 3 floats), grid.bin (distances(), the reference's
// x*Ny*Nz + y*Nz + z order) and dirty.txt (per coupling frame: dirty flag and
// box).
//
// usage: pbfluids_driver <dir> <frames> <couple>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../realtime-fsi-engine_amd/host/PBFluids.h"
#include "fsi_pbf.h"

template <class T>
static bool write_all(const std::string& path, const std::vector<T>& v) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), fp) == v.size();
    return std::fclose(fp) == 0 && ok;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s dir frames couple\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int frames = std::atoi(argv[2]);
    const int couple = std::atoi(argv[3]);

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
    if (!write_all(dir + "/particles.bin", out) || !write_all(dir + "/contacts.bin", contacts)) return 1;
    std::printf("particles %zu contacts %zu\n", out.size(), contacts.size());

    // the deforming obstacle: vertices above the centre pulled down by
    // 1% of their height above it per frame
    FILE* fd = std::fopen((dir + "/dirty.txt").c_str(), "w");
    if (!fd) return 1;
    std::vector<std::array<double, 3>> cur = verts;
    for (int c = 0; c < couple; ++c) {
        for (size_t i = 0; i < verts.size(); ++i)
            if (verts[i][1] > centre[1]) cur[i][1] = verts[i][1] - 0.01 * (c + 1) * (verts[i][1] - centre[1]);
        int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
        bool dirty = false;
        if (c % 3 == 2) {
            sdf.refitFromMesh(cur);
            sdf.uploadToGPU();
        } else {
            dirty = sdf.refitDirty(cur, 0.2, 1e-4);
            if (dirty) {
                sdf.uploadDirtyRegion();
                sdf.dirtyRegion(lo, hi);
            }
        }
        std::fprintf(fd, "%d %d %d %d %d %d %d\n", dirty ? 1 : 0, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
        fluid.step();
        fluid.filterContactCandidates(0.1f);
    }
    std::fclose(fd);
    fluid.readbackParticles(out);
    fluid.readbackContactCandidates(contacts);
    if (fluid.lastStatus() != 0) return 1;
    std::vector<PVec3> pts;
    for (const ContactCandidate& cc : contacts) pts.push_back(make_pvec3(cc.pos.x, cc.pos.y, cc.pos.z));
    const std::vector<SDFBoundary::ClosestSurface> cs = sdf.closestTriangles(pts);
    struct Rec {
        int32_t face;
        float p[3];
    };
    std::vector<Rec> closest;
    for (size_t i = 0; i < cs.size(); ++i) {
        closest.push_back(Rec{cs[i].face, {cs[i].point.x, cs[i].point.y, cs[i].point.z}});
        if (i < 4) {  // the one-point query gives the same answer
            const SDFBoundary::ClosestSurface one = sdf.closestTriangle(pts[i]);
            if (one.face != cs[i].face || one.point.x != cs[i].point.x || one.point.y != cs[i].point.y ||
                one.point.z != cs[i].point.z)
                return 3;
        }
    }
    if (!write_all(dir + "/coupled_particles.bin", out) || !write_all(dir + "/coupled_contacts.bin", contacts) ||
        !write_all(dir + "/closest.bin", closest) || !write_all(dir + "/grid.bin", sdf.distances()))
        return 1;
    std::printf("coupled frames %d contacts %zu\n", couple, contacts.size());
    return 0;
}
