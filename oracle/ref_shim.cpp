// Test infrastructure only: a C-ABI shim over the reference's own CPU kernels.
//
// This file is compiled together with the reference's src/cpu/*.cpp (read in place
// from the reference tree by oracle/Makefile) into oracle/_ref/libxrecon_ref.so.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load that
// library, as a checker and as the timed CPU baseline.  It is never part of the
// product path.
//
// Every entry point forwards to the reference function named in its comment
// (include/functions.h:276-458 of the reference) with the POD descriptors built from
// plain scalars, exactly as the reference binding's unpack2d/unpack3d do
// (src/api/torch_binding.cpp:62-124).

#include <cstdint>
#include <cstring>
#include <vector>

#include "functions.h"
#include "types.h"

using xrecon::GeometryType;

namespace {

xrecon::Volume3d vol3(const int* n, const float* f)
{
    // n = {Nx, Ny, Nz}; f = {dx, dy, dz, shx, shy, shz}
    return xrecon::Volume3d(n[0], n[1], n[2], f[0], f[1], f[2], f[3], f[4], f[5]);
}

xrecon::Geometry3d geo3(const int* n, const float* f)
{
    // n = {type, Nu, Nv}; f = {du, dv, shu, shv, SOD, SDD}
    return xrecon::Geometry3d(static_cast<GeometryType>(n[0]), n[1], n[2],
                              f[0], f[1], f[2], f[3], f[4], f[5]);
}

xrecon::Volume2d vol2(const int* n, const float* f)
{
    // n = {Nx, Ny}; f = {dx, dy, shx, shy}
    return xrecon::Volume2d(n[0], n[1], f[0], f[1], f[2], f[3]);
}

xrecon::Geometry2d geo2(const int* n, const float* f)
{
    // n = {type, Nu}; f = {du, shu, SOD, SDD}
    return xrecon::Geometry2d(static_cast<GeometryType>(n[0]), n[1], f[0], f[1], f[2], f[3]);
}

struct Triplets {
    std::vector<int64_t> rows, cols;
    std::vector<float> vals;
};

Triplets g_last;

} // namespace

extern "C" {

// op: 0 forward3d_cpu, 1 forward3d_t_cpu, 2 backward3d_cpu, 3 weighted_backward3d_cpu,
//     4 backward3d_t_cpu, 5 weighted_backward3d_t_cpu   (reference functions.h:453-458)
int ref_op3d(int op, float* image, float* proj, float* ang,
             const int* vn, const float* vf, const int* gn, const float* gf,
             int num_batch, int num_view)
{
    const xrecon::Volume3d vol = vol3(vn, vf);
    const xrecon::Geometry3d geo = geo3(gn, gf);
    switch (op) {
    case 0: xrecon::forward3d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 1: xrecon::forward3d_t_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 2: xrecon::backward3d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 3: xrecon::weighted_backward3d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 4: xrecon::backward3d_t_cpu(proj, image, ang, vol, geo, num_batch, num_view); return 0;
    case 5: xrecon::weighted_backward3d_t_cpu(proj, image, ang, vol, geo, num_batch, num_view); return 0;
    default: return -1;
    }
}

// op: 0 forward2d_cpu, 1 forward2d_t_cpu, 2 backward2d_cpu, 3 weighted_backward2d_cpu,
//     4 backward2d_t_cpu, 5 weighted_backward2d_t_cpu   (reference functions.h:276-372)
int ref_op2d(int op, float* image, float* proj, float* ang,
             const int* vn, const float* vf, const int* gn, const float* gf,
             int num_batch, int num_view)
{
    const xrecon::Volume2d vol = vol2(vn, vf);
    const xrecon::Geometry2d geo = geo2(gn, gf);
    switch (op) {
    case 0: xrecon::forward2d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 1: xrecon::forward2d_t_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 2: xrecon::backward2d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 3: xrecon::weighted_backward2d_cpu(image, proj, ang, vol, geo, num_batch, num_view); return 0;
    case 4: xrecon::backward2d_t_cpu(proj, image, ang, vol, geo, num_batch, num_view); return 0;
    case 5: xrecon::weighted_backward2d_t_cpu(proj, image, ang, vol, geo, num_batch, num_view); return 0;
    default: return -1;
    }
}

// System-matrix triplets (reference system_matrix3d.cpp:39-62 / system_matrix2d_cpu).
// Returns the triplet count; ref_sm_fetch copies them out and releases the buffers.
int64_t ref_sm3d(float* ang, const int* vn, const float* vf, const int* gn, const float* gf,
                 int num_view)
{
    g_last = Triplets();
    xrecon::system_matrix3d_cpu(g_last.rows, g_last.cols, g_last.vals, ang,
                                vol3(vn, vf), geo3(gn, gf), num_view);
    return static_cast<int64_t>(g_last.vals.size());
}

int64_t ref_sm2d(float* ang, const int* vn, const float* vf, const int* gn, const float* gf,
                 int num_view)
{
    g_last = Triplets();
    xrecon::system_matrix2d_cpu(g_last.rows, g_last.cols, g_last.vals, ang,
                                vol2(vn, vf), geo2(gn, gf), num_view);
    return static_cast<int64_t>(g_last.vals.size());
}

void ref_sm_fetch(int64_t* rows, int64_t* cols, float* vals)
{
    const size_t n = g_last.vals.size();
    if (n) {
        std::memcpy(rows, g_last.rows.data(), n * sizeof(int64_t));
        std::memcpy(cols, g_last.cols.data(), n * sizeof(int64_t));
        std::memcpy(vals, g_last.vals.data(), n * sizeof(float));
    }
    g_last = Triplets();
}

} // extern "C"
