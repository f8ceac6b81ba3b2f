// Host execution of the operators for CPU tensors.
//
// The reference dispatches on the input's device: GPU tensors run its CUDA kernels, CPU
// tensors its std::thread kernels (src/api/torch_binding.cpp:421-425 ->
// src/cpu/forward3d.cpp:599-603, src/cpu/backward3d.cpp, src/cpu/*2d.cpp).  Here a CPU
// tensor runs the same per-thread bodies as the gfx950 kernels (dd_body.h) -- the same
// FP64-built geometry tables (plan.cpp), the same prefix tables, the same FP32 overlap
// arithmetic -- on the host's cores with std::thread.  It never runs for GPU tensors.
//
// Work split: the cell gathers (A, B^T, B_W^T) take one (batch item, view, 64-column
// group) per task, each column of the group in 8-row tiles on the group's span class;
// the voxel gathers (B, B_W, A^T) one (batch item, slice chunk, image row) per task, each
// voxel column reading its lookups straight from the row-prefix table; the prefix passes
// one (batch item, view) or one batch item's plane range per task.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dd_body.h"

struct xr_host_plan3d {
    xr::Plan3d p;
    xr::PlanTables t;
};

namespace xr {
namespace {

int thread_count(int requested)
{
    if (requested > 0) return requested;
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// f(i) for i in [0, n), tasks handed out dynamically over nt threads.
template <class F>
void parallel_for(int64_t n, int nt, F &&f)
{
    if (n <= 0) return;
    nt = (int)std::min<int64_t>(nt, n);
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (int64_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
        });
    for (auto &x : th) x.join();
}

constexpr int kHostRows = 8;  // rows per cell-gather "thread" (tilings give bitwise-equal results)

template <int OPK, int GEO>
void host_cell(const Plan3d &p, const float2 *WX, const float2 *WY, float *out, int nb, int nt)
{
    const CellArgs a = make_cell_args(p, nb);
    const int ntile = ceil_div(a.nv, kHostRows);
    const int ncg = ceil_div(a.nu, kWave);
    const float inf = std::numeric_limits<float>::infinity();
    parallel_for((int64_t)nb * p.nview * ncg, nt, [&](int64_t it) {
        const int cg = (int)(it % ncg);
        const int v = (int)((it / ncg) % p.nview);
        const int b = (int)(it / ((int64_t)ncg * p.nview));
        const CellGroup grp = p.d_cgrp[(int64_t)v * ncg + cg];
        const int cend = std::min(a.nu, (cg + 1) * kWave);
        for (int g = 0; g < ntile; ++g)
            for (int c = cg * kWave; c < cend; ++c) {
                const int r0 = g * kHostRows;
                if (p.big)
                    cell_body<OPK, GEO, kHostRows, 0, true>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v, b,
                                                            inf);
                else if (grp.span == 2)
                    cell_body<OPK, GEO, kHostRows, 2, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v, b,
                                                             grp.wmax);
                else if (grp.span == 3)
                    cell_body<OPK, GEO, kHostRows, 3, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v, b,
                                                             grp.wmax);
                else
                    cell_body<OPK, GEO, kHostRows, 0, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v, b,
                                                             grp.wmax);
            }
    });
}

template <int OPK, int GEO, int KC>
void host_voxel_KC(const Plan3d &p, const float2 *Q, float *out, int nb, int nt)
{
    const VoxArgs a = make_vox_args(p, nb, KC);
    parallel_for((int64_t)nb * a.nchunk * a.ny, nt, [&](int64_t it) {
        const int j = (int)(it % a.ny);
        const int chunk = (int)((it / a.ny) % a.nchunk);
        const int b = (int)(it / ((int64_t)a.ny * a.nchunk));
        for (int i = 0; i < a.nx; ++i)
            voxel_body<OPK, GEO, KC>(p.d_hdr, p.d_lineV, p.d_col, Q, out, a, i, j, chunk * KC, b, 0, a.nview);
    });
}

template <int OPK, int GEO>
void host_voxel(const Plan3d &p, const float2 *Q, float *out, int nb, int nt)
{
    const int nz = p.vol.nz;
    if (nz >= 16) host_voxel_KC<OPK, GEO, 16>(p, Q, out, nb, nt);
    else if (nz >= 8) host_voxel_KC<OPK, GEO, 8>(p, Q, out, nb, nt);
    else if (nz >= 4) host_voxel_KC<OPK, GEO, 4>(p, Q, out, nb, nt);
    else host_voxel_KC<OPK, GEO, 1>(p, Q, out, nb, nt);
}

template <int GEO>
int host_execute(const Plan3d &p, const Route &rt, const float *in, float *out, int nb, int nt)
{
    if (rt.cell) {
        const int64_t plane = (int64_t)p.vol.nx * p.vol.ny;
        const size_t n = (size_t)nb * (p.vol.nz + 1) * plane;
        std::vector<float2> WX(n + kVolPad, float2{0.0f, 0.0f}), WY;
        const int64_t per = std::max<int64_t>(1, plane / 64);
        const int64_t ntask = (plane + per - 1) / per;
        parallel_for((int64_t)nb * ntask, nt, [&](int64_t it) {
            const int b = (int)(it / ntask);
            const int64_t e0 = (it % ntask) * per, e1 = std::min(plane, e0 + per);
            for (int64_t e = e0; e < e1; ++e) vol_prefix_body(in, WX.data(), plane, p.vol.nz, b, e);
        });
        if (p.any_y) {
            // y-driven views read the planes transposed: (i, j) -> i * Ny + (Ny - 1 - j)
            WY.assign(n + kVolPad, float2{0.0f, 0.0f});
            const int nx = p.vol.nx, ny = p.vol.ny;
            parallel_for((int64_t)nb * (p.vol.nz + 1), nt, [&](int64_t q) {
                const float2 *src = WX.data() + q * plane;
                float2 *dst = WY.data() + q * plane;
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) dst[(int64_t)i * ny + (ny - 1 - j)] = src[(int64_t)j * nx + i];
            });
        }
        const float2 *wy = p.any_y ? WY.data() : WX.data();
        if (rt.kop == CELL_A) host_cell<CELL_A, GEO>(p, WX.data(), wy, out, nb, nt);
        else if (rt.kop == CELL_BT) host_cell<CELL_BT, GEO>(p, WX.data(), wy, out, nb, nt);
        else host_cell<CELL_BWT, GEO>(p, WX.data(), wy, out, nb, nt);
        return XR_OK;
    }
    const SinoArgs sa = make_sino_args(p, nb);
    std::vector<float2> Q((size_t)nb * p.nview * (p.geo.nv + 1) * p.geo.nu + kSinoPad, float2{0.0f, 0.0f});
    parallel_for((int64_t)nb * p.nview, nt, [&](int64_t it) {
        const int b = (int)(it / p.nview), v = (int)(it % p.nview);
        for (int c = 0; c < p.geo.nu; ++c) {
            if (rt.kop == VOX_AT) sino_prefix_body<SINO_AT, GEO>(p.d_hdr, p.d_col, in, Q.data(), sa, b, v, c);
            else sino_prefix_body<SINO_PLAIN, GEO>(p.d_hdr, p.d_col, in, Q.data(), sa, b, v, c);
        }
    });
    if (rt.kop == VOX_B) host_voxel<VOX_B, GEO>(p, Q.data(), out, nb, nt);
    else if (rt.kop == VOX_BW) host_voxel<VOX_BW, GEO>(p, Q.data(), out, nb, nt);
    else host_voxel<VOX_AT, GEO>(p, Q.data(), out, nb, nt);
    return XR_OK;
}

int host_one_shot(int op, const float *in, float *out, const float *angles, const xr_volume3d *vol,
                  const xr_geometry3d *geo, int num_batch, int num_view)
{
    xr_host_plan3d *plan = nullptr;
    int rc = xr_host_plan3d_create(vol, geo, angles, num_view, &plan);
    if (rc) return rc;
    rc = xr_host_plan3d_execute(plan, op, in, out, num_batch, 0);
    xr_host_plan3d_destroy(plan);
    return rc;
}

}  // namespace
}  // namespace xr

using xr::fail;

extern "C" {

int xr_host_plan3d_create(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host, int num_view,
                          xr_host_plan3d **out)
{
    xr::set_error("");
    if (!out) return fail(XR_EINVAL, "null output plan pointer");
    *out = nullptr;
    int rc = xr::validate_plan_geometry(vol, geo, num_view);
    if (rc) return rc;
    if (num_view > 0 && !angles_host) return fail(XR_EINVAL, "null angles");
    std::unique_ptr<xr_host_plan3d> plan(new xr_host_plan3d());
    xr::Plan3d &p = plan->p;
    p.vol = *vol;
    p.geo = *geo;
    p.is2d = (geo->type & XR_GEOMETRY_2D) ? 1 : 0;
    p.geo.type = geo->type & ~XR_GEOMETRY_2D;
    p.kind = p.geo.type;
    p.device = -1;
    p.nview = num_view;
    std::string err;
    rc = xr::build_tables(p, angles_host, plan->t, err);
    if (rc) return fail(rc, err);
    xr::PlanTables &t = plan->t;
    p.d_hdr = t.hdr.data();
    p.d_lineC = t.lineC.data();
    p.d_lineV = t.lineV.data();
    p.d_col = t.col.data();
    p.d_vtile = t.vtile.data();
    p.d_cgrp = t.cgrp.data();
    p.d_citems = t.citems.data();
    p.d_cquad = t.cquad.data();
    p.d_cqstage = t.cqstage.data();
    *out = plan.release();
    return XR_OK;
}

void xr_host_plan3d_destroy(xr_host_plan3d *plan) { delete plan; }

int xr_host_plan3d_execute(const xr_host_plan3d *plan, int op, const float *in, float *out, int num_batch,
                           int num_threads)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    const xr::Plan3d &p = plan->p;
    const xr::Route rt = xr::route(p, op);
    if (rt.kop < 0) return fail(XR_EINVAL, "unknown operator " + std::to_string(op));
    if (num_batch < 0) return fail(XR_EINVAL, "negative batch count");
    if (num_batch == 0) return XR_OK;
    if (!in || !out) return fail(XR_EINVAL, "null input or output pointer");
    const int nt = xr::thread_count(num_threads);
    if (p.nview == 0) {
        if (!rt.cell) std::memset(out, 0, sizeof(float) * (size_t)num_batch * p.vol.nz * p.vol.ny * p.vol.nx);
        return XR_OK;
    }
    if (p.kind == xr::GEO_FLAT) return xr::host_execute<xr::GEO_FLAT>(p, rt, in, out, num_batch, nt);
    if (p.kind == xr::GEO_ARC) return xr::host_execute<xr::GEO_ARC>(p, rt, in, out, num_batch, nt);
    return xr::host_execute<xr::GEO_PAR>(p, rt, in, out, num_batch, nt);
}

int xr_forward3d_cpu(const float *image, float *proj, const float *angles, const xr_volume3d *vol,
                     const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_PROJECTION, image, proj, angles, vol, geo, num_batch, num_view);
}

int xr_forward3d_t_cpu(float *image, const float *proj, const float *angles, const xr_volume3d *vol,
                       const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_PROJECTION_T, proj, image, angles, vol, geo, num_batch, num_view);
}

int xr_backward3d_cpu(float *image, const float *proj, const float *angles, const xr_volume3d *vol,
                      const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_BACKPROJECTION, proj, image, angles, vol, geo, num_batch, num_view);
}

int xr_weighted_backward3d_cpu(float *image, const float *proj, const float *angles, const xr_volume3d *vol,
                               const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_BACKPROJECTION_W, proj, image, angles, vol, geo, num_batch, num_view);
}

int xr_backward3d_t_cpu(float *proj, const float *image, const float *angles, const xr_volume3d *vol,
                        const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_BACKPROJECTION_T, image, proj, angles, vol, geo, num_batch, num_view);
}

int xr_weighted_backward3d_t_cpu(float *proj, const float *image, const float *angles, const xr_volume3d *vol,
                                 const xr_geometry3d *geo, int num_batch, int num_view)
{
    return xr::host_one_shot(XR_OP_BACKPROJECTION_W_T, image, proj, angles, vol, geo, num_batch, num_view);
}

int xr_op2d_cpu(int op, const float *in, float *out, const float *angles, const xr_volume2d *vol,
                const xr_geometry2d *geo, int num_batch, int num_view)
{
    xr_volume3d v3;
    xr_geometry3d g3;
    const int rc = xr_geometry2d_as_3d(vol, geo, &v3, &g3);
    if (rc) return rc;
    return xr::host_one_shot(op, in, out, angles, &v3, &g3, num_batch, num_view);
}

}  // extern "C"
