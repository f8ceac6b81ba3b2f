// Batched execution of one operator over many uneven instances in single launches.
//
// The reference runs one instance per call (B*C folded into the grid,
// src/api/torch_binding.cpp:422).  Here a batch of up to XR_BATCH_MAX plans of one
// geometry family (different volumes, detectors, view counts, batch sizes) runs each pass
// -- volume prefix, transpose, cell gather; or sinogram prefix, voxel gather -- as ONE
// launch: block b of a launch walks the per-instance block counts (computed from the
// static descriptors uploaded at batch creation and the launch's tile shape) to find its
// instance and tile, then runs the same per-thread body as the single-instance kernels.
// Per-call pointers travel by value in the kernel arguments, so executing a batch copies
// nothing, allocates nothing and is capturable into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "dd_body.h"

namespace xr {

namespace {

struct BatchDesc {
    const ViewHdr *hdr;
    const float4 *lineC, *lineV, *col;
    const VoxTile *vtile;
    const CellGroup *cgrp;
    const int *citems;  // the plan's column groups by span class (Plan3d::d_citems)
    int ncls[3];
    CellArgs ca;
    VoxArgs va;  // nchunk is filled per launch from the launch's voxel chunk
    SinoArgs sa;
    int nb, any_y;
    int big;  // Plan3d::big
    unsigned long long ws_volx, ws_voly, ws_sino;  // byte offsets into the workspace
    unsigned long long ws_part;                    // view-set partial volumes (va.vsplit > 1)
};

struct IOPtrs {
    const float *in[XR_BATCH_MAX];
    float *out[XR_BATCH_MAX];
};

// Block counts per instance and pass (identical on host and device).
__host__ __device__ inline int blocks_vol_prefix(const BatchDesc &d)
{
    return ceil_div((int64_t)d.nb * d.ca.nx * d.ca.ny, kBlock);
}
__host__ __device__ inline int blocks_transpose(const BatchDesc &d)
{
    return d.any_y ? ceil_div(d.ca.nx, 32) * ceil_div(d.ca.ny, 32) * d.nb * (d.ca.nz + 1) : 0;
}
// Column groups of span-class index K (0: two-pixel, 1: three-pixel, 2: general, 3: all).
__host__ __device__ inline int cell_class_items(const BatchDesc &d, int K)
{
    return K == 3 ? d.ncls[0] + d.ncls[1] + d.ncls[2] : d.ncls[K];
}
__host__ __device__ inline int cell_class_offset(const BatchDesc &d, int K)
{
    return K == 1 ? d.ncls[0] : K == 2 ? d.ncls[0] + d.ncls[1] : 0;
}
// Workgroups of class index K at R rows: quads of groups x row tiles x batch items.
__host__ __device__ inline int blocks_cell(const BatchDesc &d, int R, int K)
{
    return ceil_div(cell_class_items(d, K), kWavesPerBlock) * ceil_div(d.ca.nv, R) * d.nb;
}
// Rows per thread of span-class index K: R for the two-pixel class, at most 8 otherwise.
__host__ __device__ inline int cell_class_rows(int R, int K) { return K == 0 || R <= 8 ? R : 8; }
__host__ __device__ inline int blocks_sino_prefix(const BatchDesc &d)
{
    return ceil_div(d.sa.nu, kBlock) * d.sa.nview * d.nb;
}
__host__ __device__ inline int blocks_voxel(const BatchDesc &d, int KC)
{
    return ceil_div(d.va.nx, kVoxTile) * ceil_div(d.va.ny, kVoxTile) * d.nb * ceil_div(d.va.nz, chunk_slices(KC)) *
           d.va.vsplit;
}
__host__ __device__ inline int blocks_split_sum(const BatchDesc &d)
{
    return d.va.vsplit > 1 ? ceil_div((int64_t)d.nb * d.va.nz * d.va.ny * d.va.nx, kBlock) : 0;
}

// Host check that every per-instance block count of a launch, computed in 64 bits, stays
// below 2^31, so that the 32-bit counts above (and the kernels' walk over them) are exact.
inline bool block_counts_fit(const BatchDesc &d, bool cell, int R, int KC)
{
    const int64_t lim = INT32_MAX;
    auto cd = [](int64_t a, int64_t b) { return (a + b - 1) / b; };
    const int64_t nb = d.nb;
    if (cell) {
        if (cd(nb * d.ca.nx * d.ca.ny, kBlock) >= lim) return false;
        if (d.any_y && cd(d.ca.nx, 32) * cd(d.ca.ny, 32) * nb * (d.ca.nz + 1) >= lim) return false;
        for (int k = 0; k < 4; ++k) {
            const int64_t items = k == 3 ? (int64_t)d.ncls[0] + d.ncls[1] + d.ncls[2] : d.ncls[k];
            const int rows = k == 3 ? R : cell_class_rows(R, k);
            if (cd(items, kWavesPerBlock) * cd(d.ca.nv, rows) * nb >= lim) return false;
        }
        return true;
    }
    if (cd(d.sa.nu, kBlock) * d.sa.nview * nb >= lim) return false;
    if (cd(d.va.nx, kVoxTile) * cd(d.va.ny, kVoxTile) * nb * cd(d.va.nz, chunk_slices(KC)) * d.va.vsplit >= lim)
        return false;
    return !(d.va.vsplit > 1 && cd(nb * d.va.nz * d.va.ny * d.va.nx, kBlock) >= lim);
}

// First block of each instance in one launch (start[n] = the launch's block count),
// computed on the host and passed by value, so that a workgroup finds its instance with
// a scan over kernel arguments: no divisions and no memory loads (walking the descriptor
// table instead costs one dependent scalar load per instance before the work starts).
struct BlockStarts {
    int start[XR_BATCH_MAX + 1];
};

template <typename F>
inline BlockStarts block_starts(const std::vector<BatchDesc> &desc, F count)
{
    BlockStarts st{};
    int64_t acc = 0;
    for (size_t i = 0; i < desc.size(); ++i) {
        st.start[i] = (int)acc;
        acc += count(desc[i]);
    }
    st.start[desc.size()] = (int)acc;  // the caller has checked that every total stays below 2^31
    return st;
}

// Maps the flat block index to (instance, local block); returns the instance or -1.
__device__ inline int find_instance(const BlockStarts &st, int n, int &bid)
{
    if (bid >= st.start[n]) return -1;
    int i = 0;
    while (i + 1 < n && bid >= st.start[i + 1]) ++i;
    bid -= st.start[i];
    return i;
}

__global__ __launch_bounds__(kBlock) void k_vol_prefix_batch(const BatchDesc *__restrict__ D, int n, BlockStarts st,
                                                             IOPtrs io, char *__restrict__ ws)
{
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    const int64_t plane = (int64_t)d.ca.nx * d.ca.ny;
    const int64_t idx = (int64_t)bid * kBlock + threadIdx.x;
    if (idx < kVolPad)
        reinterpret_cast<float2 *>(ws + d.ws_volx)[plane * (d.ca.nz + 1) * d.nb + idx] = make_float2(0.0f, 0.0f);
    if (idx >= plane * d.nb) return;
    const int64_t b = idx / plane;
    if (d.ca.nz <= kPrefixRegRows)
        vol_prefix_reg(io.in[i], reinterpret_cast<float2 *>(ws + d.ws_volx), plane, d.ca.nz, b, idx - b * plane);
    else
        vol_prefix_body(io.in[i], reinterpret_cast<float2 *>(ws + d.ws_volx), plane, d.ca.nz, b, idx - b * plane);
}

__global__ __launch_bounds__(kBlock) void k_transpose_batch(const BatchDesc *__restrict__ D, int n, BlockStarts st,
                                                            char *__restrict__ ws)
{
    __shared__ float2 tile[32][33];
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    const int nx = d.ca.nx, ny = d.ca.ny;
    const int tx_n = ceil_div(nx, 32), ty_n = ceil_div(ny, 32);
    const int bx = bid % tx_n, by = (bid / tx_n) % ty_n, p = bid / (tx_n * ty_n);
    const int64_t plane = (int64_t)nx * ny;
    if (bid == 0 && threadIdx.x < kVolPad)
        reinterpret_cast<float2 *>(ws + d.ws_voly)[plane * (d.ca.nz + 1) * d.nb + threadIdx.x] = make_float2(0.0f, 0.0f);
    const float2 *src = reinterpret_cast<const float2 *>(ws + d.ws_volx) + (int64_t)p * plane;
    float2 *dst = reinterpret_cast<float2 *>(ws + d.ws_voly) + (int64_t)p * plane;
    const int i0 = bx * 32, j0 = by * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + r, ii = i0 + tx;
        if (j < ny && ii < nx) tile[r][tx] = src[(int64_t)j * nx + ii];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int ii = i0 + r, j = j0 + tx;
        if (ii < nx && j < ny) dst[(int64_t)ii * ny + (ny - 1 - j)] = tile[tx][r];
    }
}

// Cell gather over the column groups of span-class index K (as cell_class_items) of
// every instance: the instance's groups in class order, then row groups, then batch
// items.  One launch per class, as for a single plan.
template <int OPK, int GEO, int R, int K, bool BIG>
__global__ __launch_bounds__(kBlock) void k_cell_batch(
    const BatchDesc *__restrict__ D, int n, BlockStarts st, IOPtrs io, char *__restrict__ ws)
{
    constexpr int SPAN = K == 0 ? 2 : K == 1 ? 3 : 0;
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    const CellArgs a = d.ca;
    // as k_cell: 4 waves on 4 consecutive groups of the class, one row tile, one batch item
    const int ncg = ceil_div(a.nu, kWave), nrt = ceil_div(a.nv, R);
    const int nk = cell_class_items(d, K), nquad = ceil_div(nk, kWavesPerBlock);
    const int lane = threadIdx.x & (kWave - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int rt = bid % nrt, rest = bid / nrt;
    const int quad = rest % nquad, b = rest / nquad;
    const int idx = quad * kWavesPerBlock + wave;
    if (idx >= nk) return;  // the whole wave
    const int item = ro_load(d.citems + cell_class_offset(d, K) + idx);
    const int v = item / ncg, cg = item - v * ncg;
    const int c = cg * kWave + lane;
    const int r0 = rt * R;
    const float2 *WX = reinterpret_cast<const float2 *>(ws + d.ws_volx);
    const float2 *WY = reinterpret_cast<const float2 *>(ws + d.ws_voly);
    const float wgrp = BIG ? INFINITY : ro_load(&d.cgrp[item].wmax);
    if (c < a.nu) cell_body<OPK, GEO, R, SPAN, BIG>(d.hdr, d.lineC, d.col, WX, WY, io.out[i], a, c, r0, v, b, wgrp);
}

template <int MODE, int GEO>
__global__ __launch_bounds__(kBlock) void k_sino_prefix_batch(const BatchDesc *__restrict__ D, int n, BlockStarts st,
                                                              IOPtrs io, char *__restrict__ ws)
{
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    const int gx = ceil_div(d.sa.nu, kBlock);
    const int c = (bid % gx) * kBlock + threadIdx.x;
    const int v = (bid / gx) % d.sa.nview, b = bid / (gx * d.sa.nview);
    if (c >= d.sa.nu) return;
    if (MODE == SINO_PLAIN && d.sa.nv <= kPrefixRegRows)
        sino_prefix_reg<MODE, GEO>(d.hdr, d.col, io.in[i], reinterpret_cast<float2 *>(ws + d.ws_sino), d.sa, b, v, c);
    else
        sino_prefix_body<MODE, GEO>(d.hdr, d.col, io.in[i], reinterpret_cast<float2 *>(ws + d.ws_sino), d.sa, b, v, c);
}

template <int OPK, int GEO, int KC, bool HALVES>
__global__ __launch_bounds__(kBlock, KC >= 32 ? 4 : 1) void k_voxel_batch(const BatchDesc *__restrict__ D, int n, BlockStarts st,
                                                        IOPtrs io, char *__restrict__ ws)
{
    __shared__ __attribute__((aligned(16))) float2 tile[vox_lds(KC)];
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    VoxArgs a = d.va;
    a.nchunk = ceil_div(a.nz, KC);
    const int gx = ceil_div(a.nx, kVoxTile), gy = ceil_div(a.ny, kVoxTile);
    const int bx = bid % gx, by = (bid / gx) % gy, bz = bid / (gx * gy);
    const int set = bz % a.vsplit, rest = bz / a.vsplit;
    const int v0 = set * a.vper, v1 = imin(a.nview, v0 + a.vper);
    float *dst = a.vsplit > 1
                     ? reinterpret_cast<float *>(ws + d.ws_part) + (int64_t)set * d.nb * a.nz * a.ny * a.nx
                     : io.out[i];
    int ii, j;
    voxel_coords(bx, by, threadIdx.x, ii, j);
    voxel_wg<OPK, GEO, KC, HALVES>(d.hdr, d.lineV, d.col, d.vtile, reinterpret_cast<const float2 *>(ws + d.ws_sino),
                                   dst, a, ii, j, ii < a.nx && j < a.ny, by * gx + bx, (rest % a.nchunk) * KC,
                                   rest / a.nchunk, v0, v1, tile);
}

// Sum of the view-set partial volumes of every split instance, in set order.
__global__ __launch_bounds__(kBlock) void k_split_sum_batch(const BatchDesc *__restrict__ D, int n, BlockStarts st,
                                                            IOPtrs io, const char *__restrict__ ws)
{
    int bid = blockIdx.x;
    const int i = find_instance(st, n, bid);
    if (i < 0) return;
    const BatchDesc &d = D[i];
    const int64_t total = (int64_t)d.nb * d.va.nz * d.va.ny * d.va.nx;
    const int64_t e = (int64_t)bid * kBlock + threadIdx.x;
    if (e >= total) return;
    const float *part = reinterpret_cast<const float *>(ws + d.ws_part);
    float acc = part[e];
    for (int s = 1; s < d.va.vsplit; ++s) acc += part[(int64_t)s * total + e];
    io.out[i][e] = acc;
}

template <int OPK, int GEO, int K>
void launch_cell_batch_class(int R, const BlockStarts &st, const BatchDesc *D, int n, const IOPtrs &io, char *ws,
                             hipStream_t s)
{
    const int grid = st.start[n];
    if (grid <= 0) return;
#define XR_CELLB(RR) \
    hipLaunchKernelGGL((k_cell_batch<OPK, GEO, RR, K, false>), dim3(grid), dim3(kBlock), 0, s, D, n, st, io, ws)
    switch (R) {
    case 1: XR_CELLB(1); break;
    case 2: XR_CELLB(2); break;
    case 4: XR_CELLB(4); break;
    case 16:
        if constexpr (K == 0) XR_CELLB(16);  // the wider classes run at most 8 rows
        else XR_CELLB(8);
        break;
    default: XR_CELLB(8); break;
    }
#undef XR_CELLB
}

// gs[k]: block starts of span-class index k (k = 3: every group, big batches).
template <int OPK, int GEO>
void launch_cell_batch(int R, bool big, const BlockStarts *gs, const BatchDesc *D, int n, const IOPtrs &io, char *ws,
                       hipStream_t s)
{
    if (big) {
        hipLaunchKernelGGL((k_cell_batch<OPK, GEO, 8, 3, true>), dim3(gs[3].start[n]), dim3(kBlock), 0, s, D, n, gs[3],
                           io, ws);
        return;
    }
    launch_cell_batch_class<OPK, GEO, 0>(R, gs[0], D, n, io, ws, s);
    launch_cell_batch_class<OPK, GEO, 1>(cell_class_rows(R, 1), gs[1], D, n, io, ws, s);
    launch_cell_batch_class<OPK, GEO, 2>(cell_class_rows(R, 2), gs[2], D, n, io, ws, s);
}

template <int OPK, int GEO>
void launch_voxel_batch(int KC, const BlockStarts &st, const BatchDesc *D, int n, const IOPtrs &io, char *ws,
                        hipStream_t s)
{
    const dim3 grid(st.start[n]);
    switch (KC) {
    case 1: hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 1, false>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    case 4: hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 4, false>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    case 8: hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 8, false>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    case 16: hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 16, false>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    case 16 | XR_VOXEL_HALVES: hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 16, true>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    default:  // 32 | XR_VOXEL_HALVES: valid_chunk admits nothing else
        hipLaunchKernelGGL((k_voxel_batch<OPK, GEO, 32, true>), grid, dim3(kBlock), 0, s, D, n, st, io, ws); break;
    }
}

template <int GEO>
void dispatch_gather(const Route &rt, int R, bool big, int KC, const BlockStarts *gs, const BatchDesc *D, int n,
                     const IOPtrs &io, char *ws, hipStream_t s)
{
    if (rt.cell) {
        if (rt.kop == CELL_A) launch_cell_batch<CELL_A, GEO>(R, big, gs, D, n, io, ws, s);
        else if (rt.kop == CELL_BT) launch_cell_batch<CELL_BT, GEO>(R, big, gs, D, n, io, ws, s);
        else launch_cell_batch<CELL_BWT, GEO>(R, big, gs, D, n, io, ws, s);
    } else {
        if (rt.kop == VOX_B) launch_voxel_batch<VOX_B, GEO>(KC, gs[0], D, n, io, ws, s);
        else if (rt.kop == VOX_BW) launch_voxel_batch<VOX_BW, GEO>(KC, gs[0], D, n, io, ws, s);
        else launch_voxel_batch<VOX_AT, GEO>(KC, gs[0], D, n, io, ws, s);
    }
}

template <int GEO>
void dispatch_sino_prefix(bool at, const BlockStarts &st, const BatchDesc *D, int n, const IOPtrs &io, char *ws,
                          hipStream_t s)
{
    const dim3 grid(st.start[n]);
    if (at) hipLaunchKernelGGL((k_sino_prefix_batch<SINO_AT, GEO>), grid, dim3(kBlock), 0, s, D, n, st, io, ws);
    else hipLaunchKernelGGL((k_sino_prefix_batch<SINO_PLAIN, GEO>), grid, dim3(kBlock), 0, s, D, n, st, io, ws);
}

}  // namespace

}  // namespace xr

struct xr_batch3d {
    std::vector<const xr::Plan3d *> plans;
    std::vector<xr::BatchDesc> desc;
    xr::BatchDesc *d_desc = nullptr;
    int device = 0;
    size_t ws_cell = 0, ws_vox = 0;
};

using xr::fail;

extern "C" {

int xr_batch3d_create(xr_plan3d *const *plans, int n, const int *num_batch, xr_batch3d **out)
{
    xr::set_error("");
    if (!out) return fail(XR_EINVAL, "null output batch pointer");
    *out = nullptr;
    if (!plans || !num_batch || n <= 0 || n > XR_BATCH_MAX)
        return fail(XR_EINVAL, "a batch holds 1.." + std::to_string(XR_BATCH_MAX) + " plans");
    std::unique_ptr<xr_batch3d> b(new xr_batch3d());
    size_t off_cell = 0, off_vox = 0;
    for (int i = 0; i < n; ++i) {
        if (!plans[i]) return fail(XR_EINVAL, "null plan in batch");
        const xr::Plan3d &p = plans[i]->p;
        if (num_batch[i] <= 0) return fail(XR_EINVAL, "batch counts must be positive");
        if (i > 0) {
            const xr::Plan3d &p0 = plans[0]->p;
            if (p.kind != p0.kind || p.is2d != p0.is2d)
                return fail(XR_EINVAL, "a batch must use one geometry family (all flat, all arc or all parallel)");
            if (p.device != p0.device) return fail(XR_EINVAL, "batched plans must live on one device");
        }
        if (p.nview <= 0) return fail(XR_EINVAL, "batched plans need at least one view");
        xr::BatchDesc d{};
        d.hdr = p.d_hdr;
        d.lineC = p.d_lineC;
        d.lineV = p.d_lineV;
        d.col = p.d_col;
        d.vtile = p.d_vtile;
        d.cgrp = p.d_cgrp;
        d.citems = p.d_citems;
        for (int k = 0; k < 3; ++k) d.ncls[k] = p.ncls[k];
        d.ca = xr::make_cell_args(p, num_batch[i]);
        d.va = xr::make_vox_args(p, num_batch[i], 1);
        d.sa = xr::make_sino_args(p, num_batch[i]);
        d.nb = num_batch[i];
        d.any_y = p.any_y;
        d.big = p.big;
        const size_t vb = xr::align256(xr::vol_prefix_bytes(p, num_batch[i]));
        d.ws_volx = off_cell;
        d.ws_voly = off_cell + vb;
        off_cell += vb * (p.any_y ? 2 : 1);
        d.ws_sino = off_vox;
        off_vox += xr::align256(xr::sino_prefix_bytes(p, num_batch[i]));
        d.ws_part = off_vox;
        off_vox += xr::align256(xr::split_partial_bytes(p, num_batch[i]));
        b->desc.push_back(d);
        b->plans.push_back(&p);
    }
    b->ws_cell = off_cell;
    b->ws_vox = off_vox;
    b->device = plans[0]->p.device;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != b->device) (void)hipSetDevice(b->device);
    hipError_t e = hipMalloc(&b->d_desc, sizeof(xr::BatchDesc) * n);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_desc, b->desc.data(), sizeof(xr::BatchDesc) * n, hipMemcpyHostToDevice);
    if (prev >= 0 && prev != b->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        if (b->d_desc) (void)hipFree(b->d_desc);
        return fail(XR_ENOMEM, std::string("batch descriptor upload failed: ") + hipGetErrorString(e));
    }
    *out = b.release();
    return XR_OK;
}

void xr_batch3d_destroy(xr_batch3d *b)
{
    if (!b) return;
    if (b->d_desc) (void)hipFree(b->d_desc);
    delete b;
}

size_t xr_batch3d_workspace_bytes(const xr_batch3d *b, int op)
{
    if (!b) return 0;
    const xr::Route rt = xr::route(*b->plans[0], op);
    if (rt.kop < 0) return 0;
    return rt.cell ? b->ws_cell : b->ws_vox;
}

int xr_batch3d_execute(const xr_batch3d *b, int op, const float *const *in, float *const *out, void *workspace,
                       size_t workspace_bytes, int tile_rows, int voxel_chunk, void *stream)
{
    if (!b) return fail(XR_EINVAL, "null batch");
    if (!in || !out) return fail(XR_EINVAL, "null input/output arrays");
    const int n = (int)b->desc.size();
    const xr::Plan3d &p0 = *b->plans[0];
    const xr::Route rt = xr::route(p0, op);
    if (rt.kop < 0) return fail(XR_EINVAL, "unknown operator " + std::to_string(op));
    const size_t need = rt.cell ? b->ws_cell : b->ws_vox;
    if (workspace_bytes < need || !workspace)
        return fail(XR_EWORKSPACE, "workspace of " + std::to_string(workspace_bytes) + " bytes, need " +
                                       std::to_string(need));
    xr::IOPtrs io{};
    for (int i = 0; i < n; ++i) {
        if (!in[i] || !out[i]) return fail(XR_EINVAL, "null buffer for batch instance " + std::to_string(i));
        io.in[i] = in[i];
        io.out[i] = out[i];
    }
    // the batched cell gather reads the tables directly: the staging flags change nothing
    if ((tile_rows & (XR_CELL_DIRECT | XR_CELL_LDS)) == (XR_CELL_DIRECT | XR_CELL_LDS))
        return fail(XR_EINVAL, "tile_rows must be 1, 2, 4, 8 or 16, with at most one of XR_CELL_DIRECT, XR_CELL_LDS");
    int R = tile_rows & ~(XR_CELL_DIRECT | XR_CELL_LDS), KC = voxel_chunk;
    if (R == 0) {
        R = 1;
        for (const xr::Plan3d *p : b->plans) R = std::max(R, xr::default_rows(*p));
    }
    if (KC == 0) {
        KC = 1;
        for (const xr::Plan3d *p : b->plans) KC = std::max(KC, xr::default_chunk(*p));
    }
    if (!xr::valid_rows(R))
        return fail(XR_EINVAL, "tile_rows must be 1, 2, 4, 8 or 16, with at most one of XR_CELL_DIRECT, XR_CELL_LDS");
    if (!xr::valid_chunk(KC)) return fail(XR_EINVAL, "voxel_chunk must be 1, 4, 8, 16, 16 | XR_VOXEL_HALVES or 32 | XR_VOXEL_HALVES");
    bool big = false;
    for (const xr::BatchDesc &d : b->desc) big = big || d.big;
    if (big) R = 8;  // tables of 4 GiB and more: the 64-bit general variant with 8-row tiles
    hipStream_t s = static_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    const xr::BatchDesc *D = b->d_desc;
    int64_t g_pre = 0, g_tr = 0, g_sum = 0;
    int64_t g_gather[4] = {0, 0, 0, 0};  // voxel gathers: [0]; cell gathers: per span class, [3] = all (big)
    for (int i = 0; i < n; ++i) {
        const xr::BatchDesc &d = b->desc[i];
        if (!xr::block_counts_fit(d, rt.cell, R, KC))
            return fail(XR_EINVAL, "batch instance " + std::to_string(i) + " is too large for one launch");
        if (rt.cell) {
            g_pre += xr::blocks_vol_prefix(d);
            g_tr += xr::blocks_transpose(d);
            for (int k = 0; k < 3; ++k) g_gather[k] += xr::blocks_cell(d, xr::cell_class_rows(R, k), k);
            g_gather[3] += xr::blocks_cell(d, R, 3);
        } else {
            g_pre += xr::blocks_sino_prefix(d);
            g_gather[0] += xr::blocks_voxel(d, KC);
            g_sum += xr::blocks_split_sum(d);
        }
    }
    if (g_pre >= INT32_MAX || g_tr >= INT32_MAX || g_sum >= INT32_MAX || g_gather[0] >= INT32_MAX ||
        g_gather[1] >= INT32_MAX || g_gather[2] >= INT32_MAX || g_gather[3] >= INT32_MAX)
        return fail(XR_EINVAL, "batch too large for one launch");
    // per-launch first block of every instance (kernel arguments)
    using xr::BatchDesc;
    xr::BlockStarts gs[4];
    const int kind = p0.kind;
    if (rt.cell) {
        const xr::BlockStarts pre = xr::block_starts(b->desc, [](const BatchDesc &d) { return xr::blocks_vol_prefix(d); });
        const xr::BlockStarts tr = xr::block_starts(b->desc, [](const BatchDesc &d) { return xr::blocks_transpose(d); });
        for (int k = 0; k < 4; ++k) {
            const int rows = k == 3 ? R : xr::cell_class_rows(R, k);
            gs[k] = xr::block_starts(b->desc, [rows, k](const BatchDesc &d) { return xr::blocks_cell(d, rows, k); });
        }
        hipLaunchKernelGGL(xr::k_vol_prefix_batch, dim3((int)g_pre), dim3(xr::kBlock), 0, s, D, n, pre, io, ws);
        if (g_tr > 0)
            hipLaunchKernelGGL(xr::k_transpose_batch, dim3((int)g_tr), dim3(xr::kBlock), 0, s, D, n, tr, ws);
    } else {
        const bool at = rt.kop == xr::VOX_AT;
        const xr::BlockStarts pre =
            xr::block_starts(b->desc, [](const BatchDesc &d) { return xr::blocks_sino_prefix(d); });
        gs[0] = xr::block_starts(b->desc, [KC](const BatchDesc &d) { return xr::blocks_voxel(d, KC); });
        if (kind == xr::GEO_FLAT) xr::dispatch_sino_prefix<xr::GEO_FLAT>(at, pre, D, n, io, ws, s);
        else if (kind == xr::GEO_ARC) xr::dispatch_sino_prefix<xr::GEO_ARC>(at, pre, D, n, io, ws, s);
        else xr::dispatch_sino_prefix<xr::GEO_PAR>(at, pre, D, n, io, ws, s);
    }
    if (kind == xr::GEO_FLAT) xr::dispatch_gather<xr::GEO_FLAT>(rt, R, big, KC, gs, D, n, io, ws, s);
    else if (kind == xr::GEO_ARC) xr::dispatch_gather<xr::GEO_ARC>(rt, R, big, KC, gs, D, n, io, ws, s);
    else xr::dispatch_gather<xr::GEO_PAR>(rt, R, big, KC, gs, D, n, io, ws, s);
    if (g_sum > 0) {
        const xr::BlockStarts sum =
            xr::block_starts(b->desc, [](const BatchDesc &d) { return xr::blocks_split_sum(d); });
        hipLaunchKernelGGL(xr::k_split_sum_batch, dim3((int)g_sum), dim3(xr::kBlock), 0, s, D, n, sum, io, ws);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(XR_EHIP, std::string("batched launch failed: ") + hipGetErrorString(e));
    return XR_OK;
}

}  // extern "C"
