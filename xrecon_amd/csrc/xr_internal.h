// Internal definitions shared by the host plan builder (plan.cpp), the kernels
// (dd_kernels.hip) and the C-ABI (capi.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xrecon_amd.h"

namespace xr {

// Per-view header, one per view, FP32 values computed in FP64 on the host.
//
// Every geometry is reduced to the same picture: a driving axis (the x axis on y = 0
// for x-driven views, the y axis on x = 0 otherwise, include/cpu3d_common.h:65-78), on
// which the detector column boundaries d_c (c = 0..Nu, in increasing "internal" order,
// include/cpu3d_common.h:87) and the voxel footprints are compared.  Voxels are
// organised as "lines" (image rows for x-driven views, image columns otherwise) whose
// pixel edges land on the axis as an arithmetic sequence P0 + m * Delta.
struct alignas(16) ViewHdr {
    int xdrv;    // 1: x-driven (cos^2 > 0.5 in FP32), 0: y-driven
    int keep;    // 1: internal column c is physical column c; 0: physical Nu-1-c
    int nline;   // Ny (x-driven) or Nx (y-driven)
    int nalong;  // Nx (x-driven) or Ny (y-driven)
    float P, S;      // cos(theta) = |S| / sqrt((dbar - P)^2 + hbar^2 + S^2)
    float delta;     // forward-projection numerator: dy (x-driven) or dx (y-driven)
    float colw;      // backprojection column width w
    float tp1, tp0;  // T(a) = (tp1 * a + tp0) / (tq1 * a + tq0)
    float tq1, tq0;
    float tk0, tk1;  // fractional internal column of axis point a: tk0 + tk1 * g(T(a))
    float sx, sy;    // source position (cone) for the FDK weights
    float dlo, dhi;  // d_0 and d_Nu
    int span;        // 2 / 3: every column overlaps at most 2 / 3 pixels of any line; 0: no bound
                     // (the widest class over the view's column groups; see CellGroup)
    float pad1;
};
static_assert(sizeof(ViewHdr) == 80, "ViewHdr layout");

// Per (view, 16 x 16 tile of voxel columns): the detector columns c0 .. c0+nc-1 the tile's
// footprint reaches on that view (with a one-column margin), and the range of the depth
// scale zeta * kappa over the tile's lines and all of those columns, from which a
// voxel-gather workgroup derives the rows its slice chunk reads.  Both ranges bound every
// voxel of the tile, so a workgroup whose chunk's rows lie off the detector passes over
// the view; what the workgroup stages in LDS from them decides speed only.  nc = -1: the
// tile misses the detector; nc = 0: no range (nothing staged, nothing skipped).
struct VoxTile {
    int c0, nc;
    float zk_lo, zk_hi;
};
static_assert(sizeof(VoxTile) == 16, "VoxTile layout");

// Per (view, 64-column group of the cell gather): the range of the depth scale sigma over
// the group's columns (the plan's staging bounds use it; no kernel reads it), the group's span class -- 2 / 3: each of its
// columns overlaps at most 2 / 3 consecutive pixels of any line, 0: no bound -- and its
// widest column on the axis, wmax.  On a diagonal view the edge columns of a flat panel
// are wider on the axis than the central ones, so most groups of a three-pixel view
// still take the two-pixel path; within a group, a line whose pixel footprint Delta_L
// exceeds wmax / 0.99 (wmax / 1.99) takes the two- (three-) pixel path.
struct CellGroup {
    float sig_lo, sig_hi;
    int span;
    float wmax;
};
static_assert(sizeof(CellGroup) == 16, "CellGroup layout");

// Per (cell-gather workgroup, line): what the workgroup's 4 groups can look up in the
// z-prefix table on that line, for staging it in LDS.  Pixels m_lo .. m_lo+nm-1 of the
// line: the union over the groups of the pixels their columns overlap, one pixel of
// margin on each side, plus the one (two-pixel class) or two (three-pixel class)
// entries a fixed-span read takes past its first pixel.  Slices: z = s h + zoff for a
// row boundary at height h on the axis, with s = sigma * lc.z in [sg_lo, sg_hi] over the
// groups' columns (sigma: the column's depth scale; lc.z: the line's z factor).
// nm = 0: no column of the workgroup overlaps the line.
struct QuadLine {
    int m_lo, nm;
    float sg_lo, sg_hi;
};
static_assert(sizeof(QuadLine) == 16, "QuadLine layout");

// Geometry family of a plan; selects the column inverse g() and the weights.
enum GeoKind : int { GEO_PAR = 0, GEO_FLAT = 1, GEO_ARC = 2 };

struct Plan3d {
    xr_volume3d vol{};
    xr_geometry3d geo{};
    int kind = GEO_FLAT;     // geometry family
    int is2d = 0;            // one-slice/one-row embedding of a 2D operator
    int device = 0;
    int nview = 0;
    int maxline = 0;         // max(Nx, Ny)
    int any_x = 0, any_y = 0;
    int big = 0;             // z-prefix table of one batch item >= 2^32 bytes or plane >= 2^24 bytes
    int vsplit = 1;          // voxel gathers: views split over this many workgroup sets (summed after)
    // Scalars shared by all views.
    float zb = 0, dz = 1;    // volume bottom z and slice thickness
    float zc = 0;            // volume centre z (slice boundary k sits at (k - Nz/2) dz + zc)
    float v0 = 0, dvv = 1;   // first row boundary and row pitch on the iso plane
    float rv0 = 0;           // v0 / dvv
    float zoff = 0;          // -zb / dz
    float sod = 0;
    // Host copies of the per-view flags.
    std::vector<int> xdrv, keep;
    // Device tables.
    ViewHdr *d_hdr = nullptr;   // [nview]
    float4 *d_lineC = nullptr;  // [nview][maxline] {toff, invd, iz, lam}: cell kernels
    float4 *d_lineV = nullptr;  // [nview][maxline] {P0, Delta, zeta, 1/(Delta zeta)}: voxel kernels
    float4 *d_col = nullptr;    // [nview][Nu+1]    {d, sigma, kappa, W}
    VoxTile *d_vtile = nullptr; // [nview][ceil(Ny/16)][ceil(Nx/16)]: voxel-gather staging estimates
    CellGroup *d_cgrp = nullptr; // [nview][ceil(Nu/64)]: cell-gather column groups
    int *d_citems = nullptr;     // the groups v * ceil(Nu/64) + g listed by span class (2, then 3, then 0),
                                 // within a class by g, then v
    int ncls[3] = {0, 0, 0};     // number of groups of span class 2, 3 and 0
    QuadLine *d_cquad = nullptr; // [workgroups of classes 2 and 3][maxline]: the cell gather's LDS staging
                                 // bounds; workgroup q of class k is entry cquad_off[k] + q
    int *d_cqstage = nullptr;    // [workgroups of classes 2 and 3][cell_tile_count(Nv)]: 1 for the
                                 // (group quad, row tile) pairs the cell gather stages (the 4 views share
                                 // one orientation, every block fits a slot), per tile size 4, 8, 16
    int cquad_off[2] = {0, 0};
};

// Host-side tables before upload (kept separate so tests can inspect the builder).
struct PlanTables {
    std::vector<ViewHdr> hdr;
    std::vector<float4> lineC, lineV, col;
    std::vector<VoxTile> vtile;
    std::vector<CellGroup> cgrp;
    std::vector<int> citems;  // as Plan3d::d_citems
    int ncls[3] = {0, 0, 0};
    std::vector<QuadLine> cquad;  // as Plan3d::d_cquad
    std::vector<int> cqstage;       // as Plan3d::d_cqstage
};

// capi.cpp: the geometry checks of xr_plan3d_create (sizes, spacings, index limits).
int validate_plan_geometry(const xr_volume3d *vol, const xr_geometry3d *geo, int num_view);

// plan.cpp
int build_tables(Plan3d &p, const float *angles_host, PlanTables &t, std::string &err);
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// dd_kernels.hip
size_t workspace_bytes(const Plan3d &p, int op, int nbatch);
int execute(const Plan3d &p, int op, const float *in, float *out, int nbatch, void *ws,
            size_t ws_bytes, int tile_rows, int voxel_chunk, hipStream_t stream);
// XR_KERNEL_CELL / XR_KERNEL_VOXEL for op on this plan, XR_EINVAL if unknown.
int kernel_kind(const Plan3d &p, int op);
// phase 0: prefix + gather; 1: prefix pass only (fills the workspace from `in`);
// 2: gather pass only (reads the workspace, writes `out`).
int execute_fdk(const Plan3d &p, const float *in, float *out, int nbatch, const float *filter, const float *weight,
                void *ws, size_t ws_bytes, int voxel_chunk, hipStream_t s);
int execute_phase(const Plan3d &p, int op, int phase, const float *in, float *out, int nbatch,
                  void *ws, size_t ws_bytes, int tile_rows, int voxel_chunk, hipStream_t stream);


}  // namespace xr

// The opaque C-ABI handle (include/xrecon_amd.h).
struct xr_plan3d {
    xr::Plan3d p;
};
