/*
 * xrecon_amd — C-ABI of the MI355X (gfx950) distance-driven projector library
 * (libxrecon_amd.so).  Plain C types only: host descriptors, raw device pointers,
 * byte counts and an opaque hipStream_t passed as void*.  No torch types.
 *
 * Layouts are the reference's: volumes are [num_batch, Nz, Ny, Nx] float32 with x fastest
 * (the reference's [B, C, D, H, W] folded to num_batch = B*C), sinograms are
 * [num_batch, num_view, Nv, Nu] float32 (the reference's [B, C, V, Nv, Nu]).
 *
 * Every call returns 0 on success or a negative XR_E* code; xr_last_error() then
 * returns a message (thread-local).  Nothing calls exit(), nothing synchronises the
 * device, and the plan API never allocates on the execute path, so a sequence of
 * xr_plan*_execute calls can be captured into a hipGraph.
 */
#ifndef XRECON_AMD_H
#define XRECON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define XR_OK 0
#define XR_EINVAL -1   /* bad argument (shape, geometry, null pointer, unknown op) */
#define XR_EHIP -2     /* a HIP runtime call failed */
#define XR_ENOMEM -3   /* allocation failed (plan creation / one-shot entry points only) */
#define XR_EWORKSPACE -4 /* workspace smaller than xr_plan3d_workspace_bytes() */

/* Geometry types: the reference's GeometryType (include/types.h:7-11). */
#define XR_PARALLEL 0
#define XR_FLAT 1
#define XR_ARC 2

/* Operators.  The reference binds one function per operator
 * (src/api/torch_binding.cpp:407-523); here they are one enum. */
#define XR_OP_PROJECTION 0            /* A     : volume -> sinogram  (forward3d_cuda)            */
#define XR_OP_PROJECTION_T 1          /* A^T   : sinogram -> volume  (forward3d_t_cuda)          */
#define XR_OP_BACKPROJECTION 2        /* B     : sinogram -> volume  (backward3d_cuda)           */
#define XR_OP_BACKPROJECTION_W 3      /* B_W   : sinogram -> volume  (weighted_backward3d_cuda)  */
#define XR_OP_BACKPROJECTION_T 4      /* B^T   : volume -> sinogram  (backward3d_t_cuda)         */
#define XR_OP_BACKPROJECTION_W_T 5    /* B_W^T : volume -> sinogram  (weighted_backward3d_t_cuda)*/

/* Volume descriptor: the reference's Volume3d (include/types.h:62-80).  Shifts in mm. */
typedef struct xr_volume3d {
    int nx, ny, nz;
    float dx, dy, dz;
    float shift_x, shift_y, shift_z;
} xr_volume3d;

/* Detector/scan descriptor: the reference's Geometry3d (include/types.h:82-119) without
 * the helical fields.  du/shift_u follow the kernel convention of options[8]/[15]
 * (xrecon/projector/projector3d.py:81-82): flat = physical mm, arc = radians,
 * parallel = mm.  dv/shift_v are physical mm. */
typedef struct xr_geometry3d {
    int type;
    int nu, nv;
    float du, dv;
    float shift_u, shift_v;
    float sod, sdd;
} xr_geometry3d;

/* 2D descriptors: the reference's Volume2d / Geometry2d (include/types.h:35-60).  In 2D,
 * du is already the effective pitch (xrecon/projector/geometry.py:106-109). */
typedef struct xr_volume2d {
    int nx, ny;
    float dx, dy;
    float shift_x, shift_y;
} xr_volume2d;

typedef struct xr_geometry2d {
    int type;
    int nu;
    float du, shift_u;
    float sod, sdd;
} xr_geometry2d;

typedef struct xr_plan3d xr_plan3d;

/* Last error message of the calling thread ("" if none). */
const char *xr_last_error(void);

/* Library version string. */
const char *xr_version(void);

/* ---------------------------------------------------------------- plan API (3D) */

/* Builds the per-view geometry tables in FP64 on the host from float32 host angles
 * (the reference computes sinf/cosf per voxel per view on the device instead,
 * include/xrmath.h:116-123), stores them as FP32 device tables on `device`, and
 * returns an opaque plan.  Allocates device memory; not capturable. */
int xr_plan3d_create(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host,
                     int num_view, int device, xr_plan3d **out);

/* Releases a plan and its device tables. */
void xr_plan3d_destroy(xr_plan3d *plan);

/* Device workspace, in bytes, that xr_plan3d_execute needs for `op` on `num_batch`
 * volumes (0 is possible).  The workspace holds prefix-summed copies of the input. */
size_t xr_plan3d_workspace_bytes(const xr_plan3d *plan, int op, int num_batch);

/* Runs `op` on `num_batch` instances on `stream` (a hipStream_t; NULL = the legacy
 * default stream).  `in`/`out` are device pointers in the layouts above; `out` is fully
 * overwritten.  No allocation, no synchronisation: capturable into a hipGraph. */
int xr_plan3d_execute(const xr_plan3d *plan, int op, const float *in, float *out, int num_batch,
                      void *workspace, size_t workspace_bytes, void *stream);

/* The two passes of xr_plan3d_execute, callable separately (both capturable):
 * phase 1 builds the prefix tables of `in` in the workspace (out unused), phase 2 runs
 * the gather from the workspace into `out` (in unused); phase 0 does both.  Used to
 * time the gather kernels on their own, and to reuse one prefix pass for several
 * gathers over the same input. */
#define XR_PHASE_ALL 0
#define XR_PHASE_PREFIX 1
#define XR_PHASE_GATHER 2
/* voxel_chunk: slices per voxel-gather thread, 1, 4, 8 or 16 (0 = the plan's default);
 * 16 | XR_VOXEL_HALVES runs the 16 slices as two 8-slice halves, 32 | XR_VOXEL_HALVES 32
 * slices as two 16-slice halves, each half skipped by a wave whose lanes have no row of
 * it on the detector (pays on tall volumes). */
#define XR_VOXEL_HALVES 0x100
/* tile_rows: detector rows per cell-gather thread, 1, 2, 4, 8 or 16 (0 = the plan's
 * default).  With 4 rows and more a cell gather may stage, per line, the block of the
 * z-prefix table each workgroup looks up in LDS (two lines ahead of its use) instead of
 * reading every lookup from device memory; the results are the same bit for bit.  By
 * default it stages where the z-prefix table of one batch item exceeds 128 MiB (its
 * lookups then miss the last-level caches most); tile_rows | XR_CELL_DIRECT never
 * stages, tile_rows | XR_CELL_LDS stages wherever it applies (the autotuner tries
 * both).  Staging applies to the kernel launches of at least 4096 workgroups. */
#define XR_CELL_DIRECT 0x100
#define XR_CELL_LDS 0x200
int xr_plan3d_execute_phase(const xr_plan3d *plan, int op, int phase, const float *in, float *out,
                            int num_batch, void *workspace, size_t workspace_bytes, int tile_rows,
                            int voxel_chunk, void *stream);

/* FDK / FBP in one pass over the sinogram: replaces, for inference, the Python-level
 * chain of Projector3d.filtered_backprojection (xrecon/projector/projector3d.py:183-195;
 * 2D: projector.py:195-206) -- cosine weighting, the zero-padded row correlation with the
 * 2*Nu-1 tap filter, the FDK-weighted backprojection -- for detectors of at most 1024
 * columns.  `filter` (2*Nu-1 taps, already multiplied by the mean angular step and by 1/2
 * for a full scan) and `weight` (Nu column weights) are device pointers in physical
 * column order.  The filter stage accumulates in FP64 and writes the row-prefix table of
 * the filtered sinogram directly; the workspace is that of XR_OP_BACKPROJECTION_W.
 * Not a reference entry point: the reference composes these steps with torch ops. */
int xr_plan3d_execute_fdk(const xr_plan3d *plan, const float *in, float *out, int num_batch, const float *filter,
                          const float *weight, void *workspace, size_t workspace_bytes, int voxel_chunk,
                          void *stream);

/* Which gather family runs `op` on this plan: XR_KERNEL_CELL (detector-cell gather,
 * tuned by tile_rows) or XR_KERNEL_VOXEL (voxel-column gather, tuned by voxel_chunk);
 * negative for an unknown operator. */
#define XR_KERNEL_CELL 1
#define XR_KERNEL_VOXEL 2
int xr_plan3d_kernel_kind(const xr_plan3d *plan, int op);

/* Number of views / a copy of the plan's per-view flags (x-driven, keep-order), as
 * decided on the host with the reference's FP32 rule (include/cpu3d_common.h:65, :87). */
int xr_plan3d_num_view(const xr_plan3d *plan);
int xr_plan3d_view_flags(const xr_plan3d *plan, int *xdriven, int *keep_order);

/* Number of (view, 64-column group) pairs of the cell gathers (A, B^T, B_W^T) in each span
 * class: counts[0] two-pixel, counts[1] three-pixel, counts[2] general.  The cell gather
 * launches one kernel per non-empty class.  Introspection only; no reference counterpart. */
int xr_plan3d_cell_classes(const xr_plan3d *plan, int counts[3]);

/* ------------------------------------------ batched uneven instances (one launch) */

/* A batch groups up to XR_BATCH_MAX independent plans of one geometry family (uneven
 * volumes, detectors and view counts), each with its own batch count.  Executing an
 * operator on a batch launches each pass (prefix, transpose, gather) ONCE for all
 * instances: every block of the launch maps its index to (instance, tile) from the
 * per-instance descriptors uploaded at creation.  The per-call input/output pointers
 * travel as kernel arguments, so execution copies nothing and allocates nothing and
 * can be captured into a hipGraph.  The plans must outlive the batch. */
#define XR_BATCH_MAX 32
typedef struct xr_batch3d xr_batch3d;

int xr_batch3d_create(xr_plan3d *const *plans, int num_instances, const int *num_batch,
                      xr_batch3d **out);
void xr_batch3d_destroy(xr_batch3d *batch);
/* Workspace (device bytes) that xr_batch3d_execute needs for `op`. */
size_t xr_batch3d_workspace_bytes(const xr_batch3d *batch, int op);
/* in[i] / out[i]: instance i's device buffers in the layouts above.  `tile_rows`
 * and `voxel_chunk` (values as for xr_plan3d_execute_phase; 0 = default) apply to all
 * instances of the launch; the batched cell gather reads its lookups from device memory
 * (XR_CELL_DIRECT and XR_CELL_LDS are accepted and change nothing). */
int xr_batch3d_execute(const xr_batch3d *batch, int op, const float *const *in, float *const *out,
                       void *workspace, size_t workspace_bytes, int tile_rows, int voxel_chunk,
                       void *stream);

/* ------------------------------------------------- execution without allocation */

/* No counterpart in the reference.  For callers that cannot allocate device memory on the
 * call path (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): xr_plan3d_execute with the workspace taken from a fixed device array of
 * XR_NOALLOC_SCRATCH_BYTES inside the library instead of from the caller.  The caller
 * still passes the plan (created beforehand), the input and the output.  Operators whose
 * xr_plan3d_workspace_bytes() exceeds the limit fail with XR_EWORKSPACE and launch
 * nothing.  Each GPU holds its own array from the first call on it until the process
 * ends.  Calls that overlap in time on one GPU (several streams or threads) race on that
 * GPU's array: the caller serialises them. */
#define XR_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int xr_plan3d_execute_noalloc(const xr_plan3d *plan, int op, const float *in, float *out, int num_batch,
                              void *stream);

/* Same as xr_plan3d_execute with explicit tiling parameters (see above); used by the
 * autotuner. */
int xr_plan3d_execute_tuned(const xr_plan3d *plan, int op, const float *in, float *out,
                            int num_batch, void *workspace, size_t workspace_bytes, int tile_rows,
                            int voxel_chunk, void *stream);

/* ----------------------------------------------- one-shot entry points (3D, 2D) */
/* Drop-in counterparts of the reference's C++ operator API (include/functions.h).
 * They build a plan, allocate the workspace, run on `stream`, synchronise `stream`,
 * and free everything: convenient, but they allocate and synchronise, so they are not
 * capturable.  Unlike the reference, `angles_host` is a HOST pointer (the geometry
 * tables are computed on the host in FP64). */

/* include/functions.h:374-382  forward3d_cuda(image, proj, ang, vol, geo, num_batch, num_view) */
int xr_forward3d(const float *image, float *proj, const float *angles_host, const xr_volume3d *vol,
                 const xr_geometry3d *geo, int num_batch, int num_view, void *stream);
/* include/functions.h:384-392  forward3d_t_cuda */
int xr_forward3d_t(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                   const xr_geometry3d *geo, int num_batch, int num_view, void *stream);
/* include/functions.h:394-402  backward3d_cuda */
int xr_backward3d(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                  const xr_geometry3d *geo, int num_batch, int num_view, void *stream);
/* include/functions.h:404-412  weighted_backward3d_cuda */
int xr_weighted_backward3d(float *image, const float *proj, const float *angles_host,
                           const xr_volume3d *vol, const xr_geometry3d *geo, int num_batch,
                           int num_view, void *stream);
/* include/functions.h:414-422  backward3d_t_cuda(proj, image, ...) */
int xr_backward3d_t(float *proj, const float *image, const float *angles_host, const xr_volume3d *vol,
                    const xr_geometry3d *geo, int num_batch, int num_view, void *stream);
/* include/functions.h:424-432  weighted_backward3d_t_cuda(proj, image, ...) */
int xr_weighted_backward3d_t(float *proj, const float *image, const float *angles_host,
                             const xr_volume3d *vol, const xr_geometry3d *geo, int num_batch,
                             int num_view, void *stream);

/* 2D operators run as the 3D kernels on a one-slice volume and a one-row detector.
 * include/functions.h:216-274  forward2d_cuda, backward2d_cuda, weighted_backward2d_cuda,
 * forward2d_t_cuda, backward2d_t_cuda, weighted_backward2d_t_cuda.  Image [nb, Ny, Nx],
 * sinogram [nb, num_view, Nu]. */
int xr_op2d(int op, const float *in, float *out, const float *angles_host, const xr_volume2d *vol,
            const xr_geometry2d *geo, int num_batch, int num_view, void *stream);

/* Maps a 2D descriptor pair onto the equivalent one-slice 3D descriptors used by the
 * plan API (so 2D operators can use plans, workspaces and graphs too).  The returned
 * geometry type carries XR_GEOMETRY_2D, which tells xr_plan3d_create to apply the 2D
 * conventions (effective pitch, no depth magnification). */
#define XR_GEOMETRY_2D 0x100
int xr_geometry2d_as_3d(const xr_volume2d *vol2, const xr_geometry2d *geo2, xr_volume3d *vol3,
                        xr_geometry3d *geo3);

/* ------------------------------------------------- host execution (CPU tensors) */

/* The operators on host memory, for inputs that live on the CPU: the reference runs its
 * std::thread CPU kernels for CPU tensors (src/api/torch_binding.cpp:421-425).  Here the
 * host runs the same per-thread bodies as the GPU kernels on the same FP64-built tables,
 * on `num_threads` threads (0 = all hardware threads).  Pointers are host pointers. */
typedef struct xr_host_plan3d xr_host_plan3d;
int xr_host_plan3d_create(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host,
                          int num_view, xr_host_plan3d **out);
void xr_host_plan3d_destroy(xr_host_plan3d *plan);
int xr_host_plan3d_execute(const xr_host_plan3d *plan, int op, const float *in, float *out, int num_batch,
                           int num_threads);

/* include/functions.h:249  forward3d_cpu(image, proj, ang, vol, geo, num_batch, num_view) */
int xr_forward3d_cpu(const float *image, float *proj, const float *angles_host, const xr_volume3d *vol,
                     const xr_geometry3d *geo, int num_batch, int num_view);
/* include/functions.h:250  forward3d_t_cpu */
int xr_forward3d_t_cpu(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                       const xr_geometry3d *geo, int num_batch, int num_view);
/* include/functions.h:251  backward3d_cpu */
int xr_backward3d_cpu(float *image, const float *proj, const float *angles_host, const xr_volume3d *vol,
                      const xr_geometry3d *geo, int num_batch, int num_view);
/* include/functions.h:252  weighted_backward3d_cpu */
int xr_weighted_backward3d_cpu(float *image, const float *proj, const float *angles_host,
                               const xr_volume3d *vol, const xr_geometry3d *geo, int num_batch,
                               int num_view);
/* include/functions.h:253  backward3d_t_cpu(proj, image, ...) */
int xr_backward3d_t_cpu(float *proj, const float *image, const float *angles_host, const xr_volume3d *vol,
                        const xr_geometry3d *geo, int num_batch, int num_view);
/* include/functions.h:254  weighted_backward3d_t_cpu(proj, image, ...) */
int xr_weighted_backward3d_t_cpu(float *proj, const float *image, const float *angles_host,
                                 const xr_volume3d *vol, const xr_geometry3d *geo, int num_batch,
                                 int num_view);
/* include/functions.h:72-165  forward2d_cpu ... weighted_backward2d_t_cpu: image
 * [nb, Ny, Nx] <-> sinogram [nb, num_view, Nu] on the host. */
int xr_op2d_cpu(int op, const float *in, float *out, const float *angles_host, const xr_volume2d *vol,
                const xr_geometry2d *geo, int num_batch, int num_view);

/* ------------------------------------------------------- system matrix (host) */

/* COO triplets of the forward projector A: rows index the sinogram
 * ((view * Nv + row) * Nu + col), columns the volume ((z * Ny + y) * Nx + x); values that
 * are 0 or not finite are dropped.  Built on the host by the reference's own FP32
 * detector-driven traversal so that the index and sparsity pattern are identical to the
 * reference's (include/functions.h:296-323 system_matrix2d_cpu / system_matrix3d_cpu,
 * src/cpu/system_matrix3d.cpp:39-62, src/cpu/forward3d.cpp:611-658,
 * src/cpu/system_matrix2d.cpp:359-389).  Triplets come in traversal order, not
 * coalesced. */
typedef struct xr_coo xr_coo;
int xr_system_matrix3d(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host,
                       int num_view, xr_coo **out);
int xr_system_matrix2d(const xr_volume2d *vol, const xr_geometry2d *geo, const float *angles_host,
                       int num_view, xr_coo **out);
int64_t xr_coo_nnz(const xr_coo *coo);
int xr_coo_copy(const xr_coo *coo, int64_t *rows, int64_t *cols, float *vals);
void xr_coo_destroy(xr_coo *coo);

/* --------------------------------------------------- system matrix (GPU emission) */

/* The same triplets as xr_system_matrix3d / xr_system_matrix2d -- same order, same
 * values bit for bit -- emitted into device memory by `device`.  No reference
 * counterpart: the reference builds the matrix on the CPU only
 * (src/api/torch_binding.cpp:380-405 rejects GPU tensors).  Creation builds the per-view
 * tables on the host (every transcendental of the traversal), runs a counting pass and a
 * scan on `stream` and synchronises it to learn the number of triplets; emission then
 * writes them into caller-provided device buffers of xr_sysmat_gpu_nnz() entries on
 * `stream` without synchronising. */
typedef struct xr_sysmat_gpu xr_sysmat_gpu;
int xr_system_matrix3d_gpu(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host,
                           int num_view, int device, void *stream, xr_sysmat_gpu **out);
int xr_system_matrix2d_gpu(const xr_volume2d *vol, const xr_geometry2d *geo, const float *angles_host,
                           int num_view, int device, void *stream, xr_sysmat_gpu **out);
int64_t xr_sysmat_gpu_nnz(const xr_sysmat_gpu *sm);
int xr_sysmat_gpu_emit(const xr_sysmat_gpu *sm, int64_t *rows, int64_t *cols, float *vals, void *stream);
void xr_sysmat_gpu_destroy(xr_sysmat_gpu *sm);

#ifdef __cplusplus
}
#endif

#endif /* XRECON_AMD_H */
