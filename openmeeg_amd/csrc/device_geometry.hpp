// Device-resident geometry tables and the unknown layout used on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "bem_device.hpp"
#include "geometry.hpp"

namespace omh {

// Unknown ordering of a device matrix.  The reference orders vertices first
// and triangles after (geometry.cpp:256-295).  OMH_ORDER_TRIANGLES_FIRST
// rotates that order so that all triangle (P0) unknowns come first; in that
// order the HeadMat admits a pivot-free signed Cholesky factorisation
// (S block SPD, vertex Schur complement negative definite).
struct Layout {
    int64_t ld;     // leading dimension (column-major)
    int64_t nv;     // vertex unknowns
    int64_t nt;     // non-barrier triangle unknowns
    int tri_first;  // 1: triangles-first order, 0: reference order
};

__host__ __device__ __forceinline__ int64_t dev_index(const Layout& L, int64_t u) {
    if (!L.tri_first) return u;
    return (u >= L.nv) ? u - L.nv : u + L.nt;
}

struct MeshDev {
    int64_t tri_begin, tri_end;      // global triangle ids
    int64_t slot_begin, slot_end;    // vertex slots
    int64_t tri_unk0;                // reference unknown index of the first triangle
    int barrier;
};

struct DeviceGeometry {
    int device = -1;
    TriRec* tri = nullptr;           // nt records
    int32_t* tri_vid = nullptr;      // nt x 3 global vertex ids
    int64_t* vert_unk = nullptr;     // nv: reference unknown index (-1 invalid)
    int32_t* slot_vid = nullptr;     // vertex slot -> global vertex id
    int64_t* adj_off = nullptr;      // slot -> [adj_off[s], adj_off[s+1]) in adj
    AdjRec* adj = nullptr;
    double* gp = nullptr;            // far-field Gauss points, GP_STRIDE per triangle
    std::vector<MeshDev> meshes;
    int64_t n_tris = 0, n_verts = 0, n_slots = 0;
    ~DeviceGeometry();
};

// Builds (once) the device tables of a geometry on the current device.
DeviceGeometry& device_geometry(Geometry& g);

}  // namespace omh
