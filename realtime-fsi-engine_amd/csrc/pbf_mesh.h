// pbf_mesh.h — the solver-side entry of the device mesh (pbf_mesh.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fsi_pbf.h"

namespace fsi {

// The contact forces of computeContactForces (src/main.cpp:604-710) over the
// candidate records (pos, vel float4 pairs, *count of them on the device, at
// most cap) that the solver's contact filter wrote on stream st.  Forces land
// in the mesh (device, per vertex); counts[0] = contacts, counts[1] = control
// points, copied to the host when counts is not NULL (synchronising).
int mesh_contact_forces(fsi_mesh* m, const float4* contacts, const uint32_t* count, uint32_t cap, hipStream_t st,
                        float band, int use_vel_magnitude, float control_threshold, float* forces_host,
                        uint32_t counts[2]);

// The mesh's current grid for a solver: device pointer (valid until the mesh
// is destroyed), dims, origin and cell; `ready` is recorded on the mesh's
// stream after the last grid write, for the solver's stream to wait on.
int mesh_grid_for_solver(fsi_mesh* m, const float** grid, int32_t dims[3], double origin[3], double* cell,
                         hipEvent_t* ready);

// The mesh's stream waits for the work queued so far on st (a reader of the
// mesh's grid or forces), so that a later refit does not overwrite them early.
int mesh_wait_for(fsi_mesh* m, hipStream_t st);

// The device the mesh lives on (-1 for NULL).
int mesh_device(const fsi_mesh* m);

}  // namespace fsi
