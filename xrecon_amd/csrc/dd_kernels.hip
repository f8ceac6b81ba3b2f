// Distance-driven projector kernels for gfx950 (MI355X).
//
// Two kernel families cover every operator of the reference's projector core
// (src/kernels/forward3d_*.cu, backward3d_*.cu, forward2d*.cu, backward2d*.cu), and both
// are gathers: every output element is written once by the thread that owns it, with no
// atomics, so results are deterministic run to run.
//
//  * Cell gathers (A, B^T, B_W^T -> sinogram).  A thread owns one detector column and R
//    consecutive detector rows of one view; a wave is 64 adjacent columns, so the
//    volume reads of a wave are contiguous along the line.  The thread walks the
//    volume's lines (image rows for x-driven views, image columns otherwise, read from a
//    y/x-transposed copy so both orientations are coalesced); on each line the column's
//    axis interval [d_c, d_c+1] covers a few pixels m, and each row's height interval
//    covers a fraction of each pixel's z column.  That z overlap is read from a z-prefix
//    table of the volume, J(q) = J[k] + (q - k) f[k] (stored as J[k] - k f[k] beside f[k],
//    read as one fmaf), so one row costs one table lookup
//    per boundary instead of a loop over the voxels it overlaps, and the R rows share
//    the line's column overlaps and their common row boundaries.  The 64-column groups
//    of every view are classed by how many pixels their widest column can cover (two,
//    three, more); each class is one launch over its listed groups, and inside a wider
//    group each line takes the narrowest fixed path its pixel footprint allows.
//
//  * Voxel gathers (B, B_W, A^T -> volume).  A thread owns one voxel column (i, j) and a
//    chunk of KC slices; a workgroup is a 16 x 16 tile of voxel columns.  For each view
//    it finds the detector columns its footprint overlaps (closed-form inverse of the
//    column map, corrected against the FP32 table), and for each such column reads the
//    row overlaps of all KC voxels from a row-prefix table of the sinogram,
//    I(rho) = B[r] + rho p[r] with B[r] = Q[r] - r p[r]; consecutive voxels share their
//    boundary lookups.
//
// The per-thread bodies live in dd_body.h.  Geometry comes from per-view FP32 tables
// computed in FP64 on the host (plan.cpp); all overlap arithmetic is FP32, the prefix
// sums are accumulated in FP64 and stored FP32.  The path is bound by the vector ALU
// and the vector-memory address path (no dense contraction, so no MFMA).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dd_body.h"

namespace xr {

namespace {

// Zero fill (a kernel rather than hipMemsetAsync so that it is safe inside captured graphs).
__global__ __launch_bounds__(kBlock) void k_zero(float *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_vol_prefix(const float *__restrict__ f, float2 *__restrict__ W, int nx,
                                                       int ny, int nz, int nbatch)
{
    const int64_t plane = (int64_t)nx * ny;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < kVolPad) W[plane * (nz + 1) * nbatch + idx] = make_float2(0.0f, 0.0f);  // pad entries
    if (idx >= plane * nbatch) return;
    const int64_t b = idx / plane, e = idx - b * plane;
    if (nz <= kPrefixRegRows) vol_prefix_reg(f, W, plane, nz, b, e);
    else vol_prefix_body(f, W, plane, nz, b, e);
}

// y-driven line orientation: WY[p][i][ny-1-j] = WX[p][j][i] for every plane p of every
// batch item, an LDS-tiled transpose (32x32 tiles padded against bank conflicts).
__global__ __launch_bounds__(kBlock) void k_plane_transpose(const float2 *__restrict__ WX, float2 *__restrict__ WY,
                                                            int nx, int ny)
{
    __shared__ float2 tile[32][33];
    const int64_t plane = (int64_t)nx * ny;
    if ((blockIdx.x | blockIdx.y | blockIdx.z) == 0 && threadIdx.x < kVolPad)
        WY[plane * gridDim.z + threadIdx.x] = make_float2(0.0f, 0.0f);  // pad entries after the last plane
    const float2 *src = WX + (int64_t)blockIdx.z * plane;
    float2 *dst = WY + (int64_t)blockIdx.z * plane;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + r, i = i0 + tx;
        if (j < ny && i < nx) tile[r][tx] = src[(int64_t)j * nx + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < nx && j < ny) dst[(int64_t)i * ny + (ny - 1 - j)] = tile[tx][r];
    }
}

template <int MODE, int GEO>
__global__ __launch_bounds__(kBlock) void k_sino_prefix(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                        const float *__restrict__ p, float2 *__restrict__ Q, SinoArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nu) return;
    // the A^T scaling (a square root and divisions per row) makes the register path
    // slower than the generic one: plain mode only
    if (MODE == SINO_PLAIN && a.nv <= kPrefixRegRows)
        sino_prefix_reg<MODE, GEO>(hdr, colt, p, Q, a, blockIdx.z, blockIdx.y, c);
    else sino_prefix_body<MODE, GEO>(hdr, colt, p, Q, a, blockIdx.z, blockIdx.y, c);
}

// FDK filter stage.  Every detector row of every view and batch item is filtered with the
// same taps, so rows are the parallel axis: a workgroup takes kFdkRows consecutive rows
// (one per lane) and kFdkCols output columns (kFdkOut consecutive columns per wave).
// The weighted samples p[s] w[s] of a chunk of kFdkChunk columns are staged in LDS in
// FP64, transposed so that each lane reads its own row conflict-free, while the next
// chunk's samples load into registers; the taps are wave-uniform (they depend on s and
// the output column only) and come from the FP64 table T through the scalar cache.  Each
// output accumulates fma(p[s] w[s], h[.], acc) over s ascending, as fdk_filter_output
// states it.
//
// FUSED (Nv divides kFdkRows, so a row group holds whole views): the workgroup then
// builds the row-prefix table of its views' columns itself (prefix_column over the
// filtered samples kept in LDS).  Otherwise it writes the filtered samples to Q[.].y and
// k_fdk_prefix builds the table.
constexpr int kFdkRows = kWave;
constexpr int kFdkOut = 16;
constexpr int kFdkCols = kFdkOut * kWavesPerBlock;
constexpr int kFdkChunk = 64;
constexpr int kFdkLoads = kFdkRows / kWavesPerBlock;  // samples per thread per chunk
static_assert(kFdkTapPad >= kFdkCols && kFdkTapPad >= kFdkChunk, "tap table padding must cover the last blocks");
static_assert(kFdkChunk == kWave, "one staged sample column per lane");

// T[i] = h[i - kFdkTapPad] inside the filter, 0 in the padding on either side.
__global__ __launch_bounds__(kBlock) void k_fdk_taps(const float *__restrict__ h, double *__restrict__ T, int nu)
{
    const int n = 2 * nu - 1 + 2 * kFdkTapPad;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int k = i - kFdkTapPad;
        T[i] = (k >= 0 && k < 2 * nu - 1) ? (double)h[k] : 0.0;
    }
}

template <int MODE, int GEO, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_fdk_filter(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                       const float *__restrict__ p, const float *__restrict__ wgt,
                                                       const double *__restrict__ T, float2 *__restrict__ Q,
                                                       SinoArgs a, int64_t nrows)
{
    __shared__ double tile[kFdkChunk][kFdkRows + 1];  // [sample][row], +1 against bank conflicts
    const int nu = a.nu;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t g0 = (int64_t)blockIdx.x * kFdkRows;
    const int ub = blockIdx.y * kFdkCols + wave * kFdkOut;
    // Tb[s - j] is the tap of sample s for output column ub + j: h[s - (ub + j) + nu - 1]
    const double *Tb = T + kFdkTapPad + nu - 1 - ub;
    double acc[kFdkOut];
#pragma unroll
    for (int j = 0; j < kFdkOut; ++j) acc[j] = 0.0;
    float next[kFdkLoads];  // thread's samples of the next chunk: rows wave + 4 i, column lane
    auto load = [&](int s0) {
        const int s = s0 + lane;
#pragma unroll
        for (int i = 0; i < kFdkLoads; ++i) {
            const int64_t g = g0 + wave + kWavesPerBlock * i;
            next[i] = (g < nrows && s < nu) ? p[g * nu + s] : 0.0f;
        }
    };
    load(0);
    for (int s0 = 0; s0 < nu; s0 += kFdkChunk) {
        const int s = s0 + lane;
        const double ws = s < nu ? (double)wgt[s] : 0.0;
        __syncthreads();  // the previous chunk has been read
#pragma unroll
        for (int i = 0; i < kFdkLoads; ++i) tile[lane][wave + kWavesPerBlock * i] = (double)next[i] * ws;
        __syncthreads();
        if (s0 + kFdkChunk < nu) load(s0 + kFdkChunk);
        for (int k0 = 0; k0 < kFdkChunk; k0 += 8) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double x = tile[k0 + k][lane];
                const int t = s0 + k0 + k;
#pragma unroll
                for (int j = 0; j < kFdkOut; ++j) acc[j] = fma(x, Tb[t - j], acc[j]);
            }
        }
    }
    if (!FUSED) {
        const int64_t g = g0 + lane;
        if (g >= nrows) return;
        const int64_t sv = g / a.nv;
        const int r = (int)(g - sv * a.nv);
        const bool keep = hdr[sv % a.nview].keep;
        float2 *dst = Q + (sv * (a.nv + 1) + r) * nu;
#pragma unroll
        for (int j = 0; j < kFdkOut; ++j) {
            const int u = ub + j;
            if (u < nu) dst[keep ? u : nu - 1 - u] = make_float2(0.0f, (float)acc[j]);
        }
        return;
    }
    // filtered samples of the row group, [row][column], in the staging buffer
    constexpr int kQPitch = kFdkCols + 1;
    static_assert(kFdkRows * kQPitch * sizeof(float) <= sizeof(tile), "filtered block must fit the staging buffer");
    float *qbuf = reinterpret_cast<float *>(&tile[0][0]);
    __syncthreads();  // every wave has read its last chunk
#pragma unroll
    for (int j = 0; j < kFdkOut; ++j) qbuf[lane * kQPitch + wave * kFdkOut + j] = (float)acc[j];
    __syncthreads();
    const int nvg = kFdkRows / a.nv;  // whole views in the row group
    const int64_t sv0 = g0 / a.nv;
    for (int task = threadIdx.x; task < nvg * kFdkCols; task += kBlock) {
        const int vg = task / kFdkCols, cl = task - vg * kFdkCols;
        const int u = blockIdx.y * kFdkCols + cl;
        const int64_t sv = sv0 + vg;
        if (u >= nu || sv * a.nv >= nrows) continue;
        const int v = (int)(sv % a.nview);
        const ViewHdr h = hdr[v];
        const int c = h.keep ? u : nu - 1 - u;
        const float *qcol = qbuf + vg * a.nv * kQPitch + cl;
        prefix_column<MODE, GEO>(h, colt, a, v, c, [qcol](int r) { return qcol[r * kQPitch]; },
                                     Q + sv * (int64_t)(a.nv + 1) * nu + c);
    }
}

template <int MODE, int GEO>
__global__ __launch_bounds__(kBlock) void k_fdk_prefix(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                       float2 *__restrict__ Q, SinoArgs a)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nu) return;
    fdk_prefix_body<MODE, GEO>(hdr, colt, Q, a, blockIdx.z, blockIdx.y, c);
}

// Cell gather over the column groups of one span class: a workgroup is 4 waves on 4
// consecutive listed groups (64 columns each; one column group on adjacent views, as the
// lists run), rows r0 .. r0+R-1 of one row tile,
// one batch item; blockIdx.x = (batch item, group quad, row tile), row tile fastest
// (a.band: (batch item, row tile, group quad), see kCellBandBytes).
// Every wave of a launch but the last quad's has work whatever Nv is.  SPAN = the class (2, 3 or 0), so each
// class runs with the registers of its own paths; SPAN = -1: every class in one launch,
// each workgroup reading its group's class (small problems, where launches cost more
// than registers).  STAGE (classes 2 and 3, 4 to 16 rows): a workgroup whose (group
// quad, row tile) the plan marks (cqstage, cell_tile_count entries per quad) serves its
// lookups from LDS (cell_wg), with cquad the class's QuadLine table (maxline entries per
// quad); the others read the table directly.
// Staged kernels run at least 4 waves per SIMD at 8 rows and fewer (128 registers), 3 at
// 16 rows (168): uncapped they take 130 and 177 (the 256^2 x 64 cone-flat A at 16 rows:
// 3.31 ms uncapped, 2.72 ms capped; the 512^2 x 256 cone-arc A at 8 rows: 34.7 against
// 30.4 ms per item).
template <int OPK, int GEO, int R, int SPAN, bool BIG, bool STAGE>
__global__ __launch_bounds__(kBlock, STAGE ? (R <= 8 ? 4 : 3) : 1) void k_cell(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineC,
                                                 const float4 *__restrict__ colt, const CellGroup *__restrict__ cgrp,
                                                 const int *__restrict__ items, int nitems,
                                                 const QuadLine *__restrict__ cquad, const int *__restrict__ cqstage,
                                                 const float2 *__restrict__ WX, const float2 *__restrict__ WY,
                                                 float *__restrict__ out, CellArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int ncg = ceil_div(a.nu, kWave);
    const int nrt = ceil_div(a.nv, R), nquad = ceil_div(nitems, kWavesPerBlock);
    int rt, quad, b;
    if (a.band) {  // row tiles outermost (large tables)
        quad = blockIdx.x % nquad;
        const int rest = blockIdx.x / nquad;
        rt = rest % nrt;
        b = rest / nrt;
    } else {
        rt = blockIdx.x % nrt;
        const int rest = blockIdx.x / nrt;
        quad = rest % nquad;
        b = rest / nquad;
    }
    const int idx = quad * kWavesPerBlock + wave;
    if constexpr (STAGE) {
        static_assert((SPAN == 2 || SPAN == 3) && !BIG && (R == 4 || R == 8 || R == 16),
                      "staged cell gathers: span classes 2 and 3, 4 to 16 rows");
        if (ro_load(cqstage + (int64_t)quad * cell_tile_count(a.nv) + cell_tile_offset(a.nv, R) + rt)) {
            // (the same for the whole workgroup)
            const QuadLine *ql = cquad + (int64_t)quad * a.maxline;
            __shared__ __attribute__((aligned(16))) float2 s0[kCellLds];
            __shared__ __attribute__((aligned(16))) float2 s1[kCellLds];
            __shared__ __attribute__((aligned(16))) float2 s2[kCellLds];
            const bool on = idx < nitems;  // a wave past the list stages and waits with the others
            const int item = items[on ? idx : quad * kWavesPerBlock];
            const int v = item / ncg, cg = item - v * ncg;
            const int c = cg * kWave + lane;
            cell_wg<OPK, GEO, R, SPAN>(hdr, lineC, colt, ql, WX, WY, out, a, c, rt * R, v, b, ro_load(cgrp + item).wmax,
                                       on && c < a.nu, s0, s1, s2);
            return;
        }
    }
    if (idx >= nitems) return;  // the whole wave
    const int item = items[idx];
    const int v = item / ncg, cg = item - v * ncg;
    const int c = cg * kWave + lane;
    const int r0 = rt * R;
    const float wgrp = BIG ? INFINITY : cgrp[item].wmax;
    if (c >= a.nu) return;
    if (SPAN >= 0) {
        cell_body<OPK, GEO, R, (SPAN >= 0 ? SPAN : 0), BIG>(hdr, lineC, colt, WX, WY, out, a, c, r0, v, b, wgrp);
        return;
    }
    const int span = cgrp[item].span;
    if (span == 2) cell_body<OPK, GEO, R, 2, false>(hdr, lineC, colt, WX, WY, out, a, c, r0, v, b, wgrp);
    else if (span == 3) cell_body<OPK, GEO, R, 3, false>(hdr, lineC, colt, WX, WY, out, a, c, r0, v, b, wgrp);
    else cell_body<OPK, GEO, R, 0, false>(hdr, lineC, colt, WX, WY, out, a, c, r0, v, b, wgrp);
}

// 32-slice chunks: at most 128 registers (4 waves per SIMD); uncapped they take 133 and
// run at 3 waves, 24 % slower on the 512^2 x 256 cone-arc workload.
template <int OPK, int GEO, int KC, bool HALVES>
__global__ __launch_bounds__(kBlock, KC >= 32 ? 4 : 1) void k_voxel(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineV,
                                                  const float4 *__restrict__ colt, const VoxTile *__restrict__ vtile,
                                                  const float2 *__restrict__ Q, float *__restrict__ out,
                                                  float *__restrict__ part, VoxArgs a)
{
    __shared__ __attribute__((aligned(16))) float2 tile[vox_lds(KC)];
    int i, j;
    voxel_coords(blockIdx.x, blockIdx.y, threadIdx.x, i, j);
    const int set = blockIdx.z % a.vsplit, rest = blockIdx.z / a.vsplit;
    const int chunk = rest % a.nchunk;
    const int b = rest / a.nchunk;
    const int v0 = set * a.vper, v1 = imin(a.nview, v0 + a.vper);
    // one view set: its own partial volume; no split: the output itself
    float *dst = a.vsplit > 1 ? part + (int64_t)set * a.nbatch * a.nz * a.ny * a.nx : out;
    voxel_wg<OPK, GEO, KC, HALVES>(hdr, lineV, colt, vtile, Q, dst, a, i, j, i < a.nx && j < a.ny,
                                   blockIdx.y * gridDim.x + blockIdx.x, chunk * KC, b, v0, v1, tile);
}

// out[e] = sum over sets s = 0..nset-1 of part[s * n + e], in set order.
__global__ __launch_bounds__(kBlock) void k_split_sum(const float *__restrict__ part, float *__restrict__ out,
                                                      int64_t n, int nset)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float acc = part[e];
        for (int s = 1; s < nset; ++s) acc += part[(int64_t)s * n + e];
        out[e] = acc;
    }
}

// Workgroups of a cell-gather launch over `nitems` groups (see k_cell).
inline int64_t cell_blocks(const CellArgs &a, int R, int nitems, int nbatch)
{
    return (int64_t)ceil_div(nitems, kWavesPerBlock) * ceil_div(a.nv, R) * nbatch;
}

// cquad / cqstage: the class's QuadLine table and staging flags, or null for the direct path.
// Staging applies to tiles of 4 rows and more (fewer rows reuse a staged line too little).
template <int OPK, int GEO, int SPAN, bool BIG>
void launch_cell_class(const Plan3d &p, const CellArgs &a, int R, const int *items, int nitems,
                       const QuadLine *cquad, const int *cqstage, const float2 *WX, const float2 *WY, float *out,
                       int nbatch, hipStream_t s)
{
    if (nitems == 0) return;
    const dim3 grid((unsigned)cell_blocks(a, R, nitems, nbatch));
#define XR_CELL(RR, ST)                                                                                              \
    hipLaunchKernelGGL((k_cell<OPK, GEO, RR, SPAN, BIG, ST>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineC,         \
                       p.d_col, p.d_cgrp, items, nitems, cquad, cqstage, WX, WY, out, a)
    constexpr bool can_stage = (SPAN == 2 || SPAN == 3) && !BIG;
    const bool st = can_stage && cquad != nullptr;
    switch (R) {
    case 1: XR_CELL(1, false); break;
    case 2: XR_CELL(2, false); break;
    case 4:
        if (st) XR_CELL(4, can_stage);
        else XR_CELL(4, false);
        break;
    case 16:
        if constexpr (SPAN == 2 || SPAN == -1) {  // the wider classes run at most 8 rows
            if (st) XR_CELL(16, can_stage);
            else XR_CELL(16, false);
            break;
        }
        [[fallthrough]];
    default:
        if (st) XR_CELL(8, can_stage);
        else XR_CELL(8, false);
        break;
    }
#undef XR_CELL
}

// One launch per span class.  tile_rows R applies to the two-pixel class; the wider
// classes carry more registers per row and run at most 8 rows per thread.  Grids of
// fewer than kCellSplitMin workgroups (2D images, small volumes) take one launch for
// every class instead.
constexpr int64_t kCellSplitMin = 4096;
constexpr int64_t kCellStageMin = 4096;

template <int OPK, int GEO>
void launch_cell(const Plan3d &p, int R, bool stage, const float2 *WX, const float2 *WY, float *out, int nbatch,
                 hipStream_t s)
{
    const CellArgs a = make_cell_args(p, nbatch);
    const int ntot = p.ncls[0] + p.ncls[1] + p.ncls[2];
    if (p.big) {
        // tables of 4 GiB and more: one 64-bit general variant over every group
        launch_cell_class<OPK, GEO, 0, true>(p, a, 8, p.d_citems, ntot, nullptr, nullptr, WX, WY, out, nbatch, s);
        return;
    }
    if (cell_blocks(a, R, ntot, nbatch) < kCellSplitMin) {
        launch_cell_class<OPK, GEO, -1, false>(p, a, R, p.d_citems, ntot, nullptr, nullptr, WX, WY, out, nbatch, s);
        return;
    }
    const int Rw = R > 8 ? 8 : R;
    // A class launch of fewer than kCellStageMin workgroups reads directly: too few
    // workgroups per CU to hide the staging's per-line barrier (the 256^2 x 64 cone-flat
    // three-pixel class, 1344 workgroups: 0.40 ms staged against 0.31 ms direct).
    const bool st2 = stage && p.d_cquad && cell_blocks(a, R, p.ncls[0], nbatch) >= kCellStageMin;
    const bool st3 = stage && p.d_cquad && cell_blocks(a, Rw, p.ncls[1], nbatch) >= kCellStageMin;
    const QuadLine *q2 = st2 ? p.d_cquad + (int64_t)p.cquad_off[0] * p.maxline : nullptr;
    const QuadLine *q3 = st3 ? p.d_cquad + (int64_t)p.cquad_off[1] * p.maxline : nullptr;
    const int nt = cell_tile_count(p.geo.nv);
    const int *f2 = st2 ? p.d_cqstage + (int64_t)p.cquad_off[0] * nt : nullptr;
    const int *f3 = st3 ? p.d_cqstage + (int64_t)p.cquad_off[1] * nt : nullptr;
    launch_cell_class<OPK, GEO, 2, false>(p, a, R, p.d_citems, p.ncls[0], q2, f2, WX, WY, out, nbatch, s);
    launch_cell_class<OPK, GEO, 3, false>(p, a, Rw, p.d_citems + p.ncls[0], p.ncls[1], q3, f3, WX, WY, out, nbatch,
                                          s);
    launch_cell_class<OPK, GEO, 0, false>(p, a, Rw, p.d_citems + p.ncls[0] + p.ncls[1], p.ncls[2], nullptr, nullptr,
                                          WX, WY, out, nbatch, s);
}

template <int OPK, int GEO>
void launch_voxel(const Plan3d &p, int KC, const float2 *Q, float *out, float *part, int nbatch, hipStream_t s)
{
    const VoxArgs a = make_vox_args(p, nbatch, KC);
    dim3 grid(ceil_div(a.nx, kVoxTile), ceil_div(a.ny, kVoxTile), nbatch * a.nchunk * a.vsplit);
    switch (KC) {
    case 1: hipLaunchKernelGGL((k_voxel<OPK, GEO, 1, false>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    case 4: hipLaunchKernelGGL((k_voxel<OPK, GEO, 4, false>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    case 8: hipLaunchKernelGGL((k_voxel<OPK, GEO, 8, false>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    case 16: hipLaunchKernelGGL((k_voxel<OPK, GEO, 16, false>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    case 16 | XR_VOXEL_HALVES: hipLaunchKernelGGL((k_voxel<OPK, GEO, 16, true>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    default:  // 32 | XR_VOXEL_HALVES: valid_chunk admits nothing else
        hipLaunchKernelGGL((k_voxel<OPK, GEO, 32, true>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_lineV, p.d_col, p.d_vtile, Q, out, part, a); break;
    }
}

template <int GEO>
void launch_cell_geo(const Plan3d &p, int kop, int R, bool stage, const float2 *WX, const float2 *WY, float *out,
                     int nb, hipStream_t s)
{
    if (kop == CELL_A) launch_cell<CELL_A, GEO>(p, R, stage, WX, WY, out, nb, s);
    else if (kop == CELL_BT) launch_cell<CELL_BT, GEO>(p, R, stage, WX, WY, out, nb, s);
    else launch_cell<CELL_BWT, GEO>(p, R, stage, WX, WY, out, nb, s);
}

template <int GEO>
void launch_voxel_geo(const Plan3d &p, int kop, int KC, const float2 *Q, float *out, float *part, int nb,
                      hipStream_t s)
{
    if (kop == VOX_B) launch_voxel<VOX_B, GEO>(p, KC, Q, out, part, nb, s);
    else if (kop == VOX_BW) launch_voxel<VOX_BW, GEO>(p, KC, Q, out, part, nb, s);
    else launch_voxel<VOX_AT, GEO>(p, KC, Q, out, part, nb, s);
}

template <int GEO>
void launch_sino_prefix(const Plan3d &p, bool at, const float *in, float2 *Q, int nbatch, hipStream_t s)
{
    const SinoArgs a = make_sino_args(p, nbatch);
    dim3 grid(ceil_div(a.nu, kBlock), p.nview, nbatch);
    if (at) hipLaunchKernelGGL((k_sino_prefix<SINO_AT, GEO>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_col, in, Q, a);
    else hipLaunchKernelGGL((k_sino_prefix<SINO_PLAIN, GEO>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_col, in, Q, a);
}

// The voxel gather over the row-prefix table at the start of ws, with the view-set sum
// pass when the plan splits its views.
void voxel_gather(const Plan3d &p, int kop, int KC, void *ws, float *out, int nbatch, hipStream_t s)
{
    const float2 *Q = static_cast<const float2 *>(ws);
    float *part = p.vsplit > 1 ? reinterpret_cast<float *>(static_cast<char *>(ws) + align256(sino_prefix_bytes(p, nbatch)))
                               : nullptr;
    if (p.kind == GEO_FLAT) launch_voxel_geo<GEO_FLAT>(p, kop, KC, Q, out, part, nbatch, s);
    else if (p.kind == GEO_ARC) launch_voxel_geo<GEO_ARC>(p, kop, KC, Q, out, part, nbatch, s);
    else launch_voxel_geo<GEO_PAR>(p, kop, KC, Q, out, part, nbatch, s);
    if (p.vsplit > 1) {
        const int64_t n = (int64_t)nbatch * p.vol.nz * p.vol.ny * p.vol.nx;
        hipLaunchKernelGGL(k_split_sum, dim3((unsigned)std::min<int64_t>(ceil_div(n, kBlock), 8192)), dim3(kBlock), 0,
                           s, part, out, n, p.vsplit);
    }
}

// Filter stage + row-prefix table of the filtered sinogram (fused when Nv divides the
// row group, two kernels otherwise).
template <int MODE, int GEO>
void launch_fdk_table(const Plan3d &p, const float *in, const float *weight, const double *T, float2 *Q, int nbatch,
                      hipStream_t s)
{
    const SinoArgs a = make_sino_args(p, nbatch);
    const int64_t nrows = (int64_t)nbatch * p.nview * p.geo.nv;
    const dim3 grid((unsigned)ceil_div(nrows, kFdkRows), (unsigned)ceil_div(a.nu, kFdkCols));
    if (kFdkRows % a.nv == 0) {
        hipLaunchKernelGGL((k_fdk_filter<MODE, GEO, true>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_col, in, weight, T,
                           Q, a, nrows);
        return;
    }
    // the unfused filter does not depend on the mode or geometry: one instance
    hipLaunchKernelGGL((k_fdk_filter<SINO_PLAIN, GEO_FLAT, false>), grid, dim3(kBlock), 0, s, p.d_hdr, p.d_col, in,
                       weight, T, Q, a, nrows);
    hipLaunchKernelGGL((k_fdk_prefix<MODE, GEO>), dim3(ceil_div(a.nu, kBlock), p.nview, nbatch), dim3(kBlock), 0, s,
                       p.d_hdr, p.d_col, Q, a);
}

}  // namespace

int execute_fdk(const Plan3d &p, const float *in, float *out, int nbatch, const float *filter, const float *weight,
                void *ws, size_t ws_bytes, int voxel_chunk, hipStream_t s)
{
    const Route rt = route(p, XR_OP_BACKPROJECTION_W);
    if (nbatch < 0) return fail(XR_EINVAL, "negative batch count");
    if (nbatch == 0) return XR_OK;
    if (!in || !out || !filter || !weight) return fail(XR_EINVAL, "null input, output, filter or weight pointer");
    if (p.geo.nu > kFdkMaxNu)
        return fail(XR_EINVAL, "the fused FDK filter stage takes at most " + std::to_string(kFdkMaxNu) +
                                   " detector columns");
    if (p.nview == 0) {
        const int64_t n = (int64_t)nbatch * p.vol.nz * p.vol.ny * p.vol.nx;
        hipLaunchKernelGGL(k_zero, dim3(std::min<int64_t>(ceil_div(n, kBlock), 4096)), dim3(kBlock), 0, s, out, n);
        if (hipGetLastError() != hipSuccess) return fail(XR_EHIP, "zero-fill launch failed");
        return XR_OK;
    }
    const size_t need = workspace_bytes(p, XR_OP_BACKPROJECTION_W, nbatch);
    if (ws_bytes < need || !ws)
        return fail(XR_EWORKSPACE, "workspace of " + std::to_string(ws_bytes) + " bytes, need " + std::to_string(need));
    if (voxel_chunk == 0) voxel_chunk = default_chunk(p);
    if (!valid_chunk(voxel_chunk)) return fail(XR_EINVAL, "voxel_chunk must be 1, 4, 8, 16, 16 | XR_VOXEL_HALVES or 32 | XR_VOXEL_HALVES");
    float2 *Q = static_cast<float2 *>(ws);
    // the taps live where the view-split partial volumes go later (workspace_bytes
    // reserves room for both); the gather runs after the filter stage on the same stream
    double *T = reinterpret_cast<double *>(static_cast<char *>(ws) + align256(sino_prefix_bytes(p, nbatch)));
    const int nu = p.geo.nu;
    hipLaunchKernelGGL(k_fdk_taps, dim3((unsigned)ceil_div(2 * nu - 1 + 2 * kFdkTapPad, kBlock)), dim3(kBlock), 0, s,
                       filter, T, nu);
    const bool at = rt.kop == VOX_AT;
    if (p.kind == GEO_FLAT) at ? launch_fdk_table<SINO_AT, GEO_FLAT>(p, in, weight, T, Q, nbatch, s)
                               : launch_fdk_table<SINO_PLAIN, GEO_FLAT>(p, in, weight, T, Q, nbatch, s);
    else if (p.kind == GEO_ARC) at ? launch_fdk_table<SINO_AT, GEO_ARC>(p, in, weight, T, Q, nbatch, s)
                                   : launch_fdk_table<SINO_PLAIN, GEO_ARC>(p, in, weight, T, Q, nbatch, s);
    else at ? launch_fdk_table<SINO_AT, GEO_PAR>(p, in, weight, T, Q, nbatch, s)
            : launch_fdk_table<SINO_PLAIN, GEO_PAR>(p, in, weight, T, Q, nbatch, s);
    voxel_gather(p, rt.kop, voxel_chunk, ws, out, nbatch, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(XR_EHIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return XR_OK;
}

int kernel_kind(const Plan3d &p, int op)
{
    const Route rt = route(p, op);
    if (rt.kop < 0) return fail(XR_EINVAL, "unknown operator " + std::to_string(op));
    return rt.cell ? XR_KERNEL_CELL : XR_KERNEL_VOXEL;
}

size_t workspace_bytes(const Plan3d &p, int op, int nbatch)
{
    if (nbatch <= 0) return 0;
    const Route rt = route(p, op);
    if (rt.kop < 0) return 0;
    if (rt.cell) return align256(vol_prefix_bytes(p, nbatch)) * (p.any_y ? 2 : 1);
    size_t tail = split_partial_bytes(p, nbatch);
    if (op == XR_OP_BACKPROJECTION_W) tail = std::max(tail, fdk_taps_bytes(p));  // FDK filter taps
    return align256(sino_prefix_bytes(p, nbatch)) + align256(tail);
}

int execute(const Plan3d &p, int op, const float *in, float *out, int nbatch, void *ws, size_t ws_bytes,
            int tile_rows, int voxel_chunk, hipStream_t s)
{
    return execute_phase(p, op, 0, in, out, nbatch, ws, ws_bytes, tile_rows, voxel_chunk, s);
}

int execute_phase(const Plan3d &p, int op, int phase, const float *in, float *out, int nbatch, void *ws,
                  size_t ws_bytes, int tile_rows, int voxel_chunk, hipStream_t s)
{
    if (phase < 0 || phase > 2) return fail(XR_EINVAL, "phase must be 0 (all), 1 (prefix) or 2 (gather)");
    const bool do_prefix = phase != 2, do_gather = phase != 1;
    const Route rt = route(p, op);
    if (rt.kop < 0) return fail(XR_EINVAL, "unknown operator " + std::to_string(op));
    if (nbatch < 0) return fail(XR_EINVAL, "negative batch count");
    if (nbatch == 0) return XR_OK;
    if ((do_prefix && !in) || (do_gather && !out)) return fail(XR_EINVAL, "null input or output pointer");
    if (p.nview == 0) {
        // No views: A and B^T produce an empty sinogram; B and A^T an all-zero volume.
        if (!rt.cell && do_gather) {
            const int64_t n = (int64_t)nbatch * p.vol.nz * p.vol.ny * p.vol.nx;
            hipLaunchKernelGGL(k_zero, dim3(std::min<int64_t>(ceil_div(n, kBlock), 4096)), dim3(kBlock), 0, s, out, n);
            if (hipGetLastError() != hipSuccess) return fail(XR_EHIP, "zero-fill launch failed");
        }
        return XR_OK;
    }
    const size_t need = workspace_bytes(p, op, nbatch);
    if (ws_bytes < need || (need && !ws))
        return fail(XR_EWORKSPACE, "workspace of " + std::to_string(ws_bytes) + " bytes, need " + std::to_string(need));
    const int cell_mode = tile_rows & (XR_CELL_DIRECT | XR_CELL_LDS);
    tile_rows &= ~(XR_CELL_DIRECT | XR_CELL_LDS);
    if (tile_rows == 0) tile_rows = default_rows(p);
    if (voxel_chunk == 0) voxel_chunk = default_chunk(p);
    if (!valid_rows(tile_rows) || cell_mode == (XR_CELL_DIRECT | XR_CELL_LDS))
        return fail(XR_EINVAL, "tile_rows must be 1, 2, 4, 8 or 16, with at most one of XR_CELL_DIRECT, XR_CELL_LDS");
    const bool stage = cell_mode == XR_CELL_LDS || (cell_mode == 0 && make_cell_args(p, 1).band);
    if (!valid_chunk(voxel_chunk)) return fail(XR_EINVAL, "voxel_chunk must be 1, 4, 8, 16, 16 | XR_VOXEL_HALVES or 32 | XR_VOXEL_HALVES");

    if (rt.cell && cell_blocks(make_cell_args(p, nbatch), 1, p.ncls[0] + p.ncls[1] + p.ncls[2], nbatch) >=
                       ((int64_t)1 << 31))
        return fail(XR_EINVAL, "batch too large for one launch of the cell gather (column groups / 4 x rows x batch "
                               "must stay below 2^31)");
    if (rt.cell) {
        float2 *WX = static_cast<float2 *>(ws);
        float2 *WY = p.any_y ? reinterpret_cast<float2 *>(static_cast<char *>(ws) + align256(vol_prefix_bytes(p, nbatch)))
                             : nullptr;
        if (do_prefix) {
            const int64_t cols = (int64_t)p.vol.nx * p.vol.ny * nbatch;
            hipLaunchKernelGGL(k_vol_prefix, dim3(ceil_div(cols, kBlock)), dim3(kBlock), 0, s, in, WX, p.vol.nx,
                               p.vol.ny, p.vol.nz, nbatch);
            if (p.any_y) {
                dim3 grid(ceil_div(p.vol.nx, 32), ceil_div(p.vol.ny, 32), nbatch * (p.vol.nz + 1));
                hipLaunchKernelGGL(k_plane_transpose, grid, dim3(kBlock), 0, s, WX, WY, p.vol.nx, p.vol.ny);
            }
        }
        if (do_gather) {
            if (p.kind == GEO_FLAT) launch_cell_geo<GEO_FLAT>(p, rt.kop, tile_rows, stage, WX, WY, out, nbatch, s);
            else if (p.kind == GEO_ARC) launch_cell_geo<GEO_ARC>(p, rt.kop, tile_rows, stage, WX, WY, out, nbatch, s);
            else launch_cell_geo<GEO_PAR>(p, rt.kop, tile_rows, stage, WX, WY, out, nbatch, s);
        }
    } else {
        float2 *Q = static_cast<float2 *>(ws);
        const bool at = rt.kop == VOX_AT;
        if (do_prefix) {
            if (p.kind == GEO_FLAT) launch_sino_prefix<GEO_FLAT>(p, at, in, Q, nbatch, s);
            else if (p.kind == GEO_ARC) launch_sino_prefix<GEO_ARC>(p, at, in, Q, nbatch, s);
            else launch_sino_prefix<GEO_PAR>(p, at, in, Q, nbatch, s);
        }
        if (do_gather) voxel_gather(p, rt.kop, voxel_chunk, ws, out, nbatch, s);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(XR_EHIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return XR_OK;
}

}  // namespace xr
