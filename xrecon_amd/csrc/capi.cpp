// C-ABI of libxrecon_amd.so: plan lifetime, workspace queries, execution and the
// one-shot entry points that mirror the reference's C++ operator API
// (include/functions.h:216-458).  All argument validation happens here and reports
// through xr_last_error(); the reference's TORCH_CHECK / CUDA_CHECK paths
// (src/api/torch_binding.cpp:14-31, include/utils.h:7-11) become status codes.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "xr_internal.h"

namespace xr {
const char *last_error();
}

using xr::fail;
using xr::Plan3d;

namespace {

const int kType2d = 0x100;  // flag OR-ed into xr_geometry3d.type by xr_geometry2d_as_3d

bool finite_pos(float v) { return std::isfinite(v) && v > 0.0f; }

int validate(const xr_volume3d *vol, const xr_geometry3d *geo, int num_view)
{
    if (!vol || !geo) return fail(XR_EINVAL, "null volume or geometry descriptor");
    if (vol->nx <= 0 || vol->ny <= 0 || vol->nz <= 0) return fail(XR_EINVAL, "volume sizes must be positive");
    if (geo->nu <= 0 || geo->nv <= 0) return fail(XR_EINVAL, "detector sizes must be positive");
    if (num_view < 0) return fail(XR_EINVAL, "negative number of views");
    if ((int64_t)num_view * ((geo->nu + 63) / 64) >= ((int64_t)1 << 31))
        return fail(XR_EINVAL, "too many views for this detector width (views x ceil(Nu/64) must stay below 2^31)");
    if (!finite_pos(vol->dx) || !finite_pos(vol->dy) || !finite_pos(vol->dz))
        return fail(XR_EINVAL, "voxel spacings must be positive and finite");
    if (!finite_pos(geo->du) || !finite_pos(geo->dv))
        return fail(XR_EINVAL, "detector spacings must be positive and finite");
    if (!std::isfinite(vol->shift_x) || !std::isfinite(vol->shift_y) || !std::isfinite(vol->shift_z) ||
        !std::isfinite(geo->shift_u) || !std::isfinite(geo->shift_v))
        return fail(XR_EINVAL, "shifts must be finite");
    const int type = geo->type & ~kType2d;
    if (type != XR_PARALLEL && type != XR_FLAT && type != XR_ARC)
        return fail(XR_EINVAL, "scan type must be parallel (0), flat (1) or arc (2)");
    if (type != XR_PARALLEL && (!finite_pos(geo->sod) || !finite_pos(geo->sdd)))
        return fail(XR_EINVAL, "cone/fan geometries need positive iso_source and source_detector distances");
    // 32-bit indexing inside one batch item of a kernel (the reference is int32 throughout,
    // src/kernels/forward3d_flat.cu:140): (Nz+1)*Ny*Nx and (Nv+1)*Nu must fit.
    const int64_t volp = (int64_t)(vol->nz + 1) * vol->ny * vol->nx;
    if (volp >= ((int64_t)1 << 31)) return fail(XR_EINVAL, "volume too large for 32-bit slice indexing");
    // One view's row-prefix table is addressed with 32-bit byte offsets.
    if ((int64_t)(geo->nv + 1) * geo->nu * 8 >= ((int64_t)1 << 32))
        return fail(XR_EINVAL, "detector too large for 32-bit row indexing");
    if ((int64_t)vol->nx * vol->ny >= ((int64_t)1 << 24))
        return fail(XR_EINVAL, "volume plane too large (Nx*Ny must stay below 2^24)");
    // 24-bit index arithmetic in the gathers: rows, slices and byte strides of a row.
    if (geo->nu >= (1 << 21) || geo->nv >= (1 << 24) || vol->nz >= (1 << 24))
        return fail(XR_EINVAL, "detector or volume dimension too large for 24-bit indexing");
    return XR_OK;
}

}  // namespace

namespace xr {
int validate_plan_geometry(const xr_volume3d *vol, const xr_geometry3d *geo, int num_view)
{
    return validate(vol, geo, num_view);
}
}  // namespace xr

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T>
int upload(T **dst, const std::vector<T> &src)
{
    *dst = nullptr;
    if (src.empty()) return XR_OK;
    if (hipMalloc(reinterpret_cast<void **>(dst), src.size() * sizeof(T)) != hipSuccess)
        return fail(XR_ENOMEM, "hipMalloc failed for plan tables");
    if (hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return fail(XR_EHIP, "hipMemcpy failed for plan tables");
    return XR_OK;
}

void free_tables(Plan3d &p)
{
    if (p.d_hdr) (void)hipFree(p.d_hdr);
    if (p.d_lineC) (void)hipFree(p.d_lineC);
    if (p.d_lineV) (void)hipFree(p.d_lineV);
    if (p.d_col) (void)hipFree(p.d_col);
    if (p.d_vtile) (void)hipFree(p.d_vtile);
    if (p.d_cgrp) (void)hipFree(p.d_cgrp);
    if (p.d_citems) (void)hipFree(p.d_citems);
    if (p.d_cquad) (void)hipFree(p.d_cquad);
    if (p.d_cqstage) (void)hipFree(p.d_cqstage);
    p.d_hdr = nullptr;
    p.d_lineC = p.d_lineV = p.d_col = nullptr;
    p.d_vtile = nullptr;
    p.d_cgrp = nullptr;
    p.d_citems = nullptr;
    p.d_cquad = nullptr;
    p.d_cqstage = nullptr;
}

}  // namespace

extern "C" {

const char *xr_last_error(void) { return xr::last_error(); }

const char *xr_version(void) { return "xrecon_amd 0.1.0 (gfx950)"; }

int xr_plan3d_create(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host, int num_view,
                     int device, xr_plan3d **out)
{
    xr::set_error("");
    if (!out) return fail(XR_EINVAL, "null output plan pointer");
    *out = nullptr;
    int rc = validate(vol, geo, num_view);
    if (rc) return rc;
    if (num_view > 0 && !angles_host) return fail(XR_EINVAL, "null angles");
    std::unique_ptr<xr_plan3d> plan(new xr_plan3d());
    Plan3d &p = plan->p;
    p.vol = *vol;
    p.geo = *geo;
    p.is2d = (geo->type & kType2d) ? 1 : 0;
    p.geo.type = geo->type & ~kType2d;
    p.kind = p.geo.type;
    p.device = device;
    p.nview = num_view;
    xr::PlanTables t;
    std::string err;
    rc = xr::build_tables(p, angles_host, t, err);
    if (rc) return fail(rc, err);
    DeviceGuard g(device);
    if ((rc = upload(&p.d_hdr, t.hdr)) || (rc = upload(&p.d_lineC, t.lineC)) || (rc = upload(&p.d_lineV, t.lineV)) ||
        (rc = upload(&p.d_col, t.col)) || (rc = upload(&p.d_vtile, t.vtile)) ||
        (rc = upload(&p.d_cgrp, t.cgrp)) || (rc = upload(&p.d_citems, t.citems)) ||
        (rc = upload(&p.d_cquad, t.cquad)) || (rc = upload(&p.d_cqstage, t.cqstage))) {
        free_tables(p);
        return rc;
    }
    *out = plan.release();
    return XR_OK;
}

void xr_plan3d_destroy(xr_plan3d *plan)
{
    if (!plan) return;
    {
        DeviceGuard g(plan->p.device);
        free_tables(plan->p);
    }
    delete plan;
}

size_t xr_plan3d_workspace_bytes(const xr_plan3d *plan, int op, int num_batch)
{
    if (!plan) return 0;
    return xr::workspace_bytes(plan->p, op, num_batch);
}

int xr_plan3d_execute_tuned(const xr_plan3d *plan, int op, const float *in, float *out, int num_batch,
                            void *workspace, size_t workspace_bytes, int tile_rows, int voxel_chunk, void *stream)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    return xr::execute(plan->p, op, in, out, num_batch, workspace, workspace_bytes, tile_rows, voxel_chunk,
                       static_cast<hipStream_t>(stream));
}

int xr_plan3d_execute_phase(const xr_plan3d *plan, int op, int phase, const float *in, float *out, int num_batch,
                            void *workspace, size_t workspace_bytes, int tile_rows, int voxel_chunk, void *stream)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    return xr::execute_phase(plan->p, op, phase, in, out, num_batch, workspace, workspace_bytes, tile_rows,
                             voxel_chunk, static_cast<hipStream_t>(stream));
}

int xr_plan3d_execute_fdk(const xr_plan3d *plan, const float *in, float *out, int num_batch, const float *filter,
                          const float *weight, void *workspace, size_t workspace_bytes, int voxel_chunk, void *stream)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    return xr::execute_fdk(plan->p, in, out, num_batch, filter, weight, workspace, workspace_bytes, voxel_chunk,
                           static_cast<hipStream_t>(stream));
}

int xr_plan3d_execute(const xr_plan3d *plan, int op, const float *in, float *out, int num_batch, void *workspace,
                      size_t workspace_bytes, void *stream)
{
    return xr_plan3d_execute_tuned(plan, op, in, out, num_batch, workspace, workspace_bytes, 0, 0, stream);
}

int xr_plan3d_num_view(const xr_plan3d *plan) { return plan ? plan->p.nview : XR_EINVAL; }

int xr_plan3d_kernel_kind(const xr_plan3d *plan, int op)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    return xr::kernel_kind(plan->p, op);
}

int xr_plan3d_cell_classes(const xr_plan3d *plan, int counts[3])
{
    if (!plan || !counts) return fail(XR_EINVAL, "null plan or counts");
    for (int k = 0; k < 3; ++k) counts[k] = plan->p.ncls[k];
    return XR_OK;
}

int xr_plan3d_view_flags(const xr_plan3d *plan, int *xdriven, int *keep_order)
{
    if (!plan) return fail(XR_EINVAL, "null plan");
    for (int v = 0; v < plan->p.nview; ++v) {
        if (xdriven) xdriven[v] = plan->p.xdrv[v];
        if (keep_order) keep_order[v] = plan->p.keep[v];
    }
    return XR_OK;
}

static int one_shot(int op, const float *in, float *out, const float *angles_host, const xr_volume3d *vol,
                    const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(XR_EHIP, "no HIP device");
    xr_plan3d *plan = nullptr;
    int rc = xr_plan3d_create(vol, geo, angles_host, num_view, dev, &plan);
    if (rc) return rc;
    const size_t ws_bytes = xr_plan3d_workspace_bytes(plan, op, num_batch);
    void *ws = nullptr;
    if (ws_bytes && hipMalloc(&ws, ws_bytes) != hipSuccess) {
        xr_plan3d_destroy(plan);
        return fail(XR_ENOMEM, "hipMalloc failed for workspace");
    }
    rc = xr_plan3d_execute(plan, op, in, out, num_batch, ws, ws_bytes, stream);
    if (!rc && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
        rc = fail(XR_EHIP, "stream synchronisation failed");
    if (ws) (void)hipFree(ws);
    xr_plan3d_destroy(plan);
    return rc;
}

int xr_forward3d(const float *image, float *proj, const float *angles_host, const xr_volume3d *vol,
                 const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_PROJECTION, image, proj, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_forward3d_t(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                   const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_PROJECTION_T, proj, image, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_backward3d(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                  const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_BACKPROJECTION, proj, image, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_weighted_backward3d(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                           const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_BACKPROJECTION_W, proj, image, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_backward3d_t(float *proj, const float *image, const float *angles_host, const xr_volume3d *vol,
                    const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_BACKPROJECTION_T, image, proj, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_weighted_backward3d_t(float *proj, const float *image, const float *angles_host, const xr_volume3d *vol,
                             const xr_geometry3d *geo, int num_batch, int num_view, void *stream)
{
    return one_shot(XR_OP_BACKPROJECTION_W_T, image, proj, angles_host, vol, geo, num_batch, num_view, stream);
}

int xr_geometry2d_as_3d(const xr_volume2d *vol2, const xr_geometry2d *geo2, xr_volume3d *vol3,
                        xr_geometry3d *geo3)
{
    if (!vol2 || !geo2 || !vol3 || !geo3) return fail(XR_EINVAL, "null 2D/3D descriptor");
    *vol3 = xr_volume3d{vol2->nx, vol2->ny, 1, vol2->dx, vol2->dy, 1.0f, vol2->shift_x, vol2->shift_y, 0.0f};
    *geo3 = xr_geometry3d{geo2->type | kType2d, geo2->nu, 1, geo2->du, 1.0f, geo2->shift_u, 0.0f, geo2->sod, geo2->sdd};
    return XR_OK;
}

int xr_op2d(int op, const float *in, float *out, const float *angles_host, const xr_volume2d *vol,
            const xr_geometry2d *geo, int num_batch, int num_view, void *stream)
{
    xr_volume3d v3;
    xr_geometry3d g3;
    int rc = xr_geometry2d_as_3d(vol, geo, &v3, &g3);
    if (rc) return rc;
    return one_shot(op, in, out, angles_host, &v3, &g3, num_batch, num_view, stream);
}

}  // extern "C"
