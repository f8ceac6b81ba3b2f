// Per-thread bodies of the distance-driven kernels, written as __host__ __device__
// functions of the thread's coordinates so that the exact same code runs in the gfx950
// kernels (dd_kernels.hip) and, one "thread" at a time, in the CPU emulator that the
// test suite uses to check kernel logic against the oracle without a GPU.
//
// See dd_kernels.hip for the design; the formulas cite the reference's files.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>

#include "xr_internal.h"

namespace xr {

enum CellOp : int { CELL_A = 0, CELL_BT = 1, CELL_BWT = 2 };
enum VoxOp : int { VOX_B = 0, VOX_BW = 1, VOX_AT = 2 };
enum SinoMode : int { SINO_PLAIN = 0, SINO_AT = 1 };

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// Voxel-gather workgroup: a kVoxTile x kVoxTile block of voxel columns, each wave an
// 8 x 8 tile of it.  The 64 lanes of a wave then project onto a compact detector patch
// (a few cache lines per lookup) for views of either driving axis, where a 64 x 1 strip
// would spread one lookup over ~128 columns.
constexpr int kVoxTile = 16;

__host__ __device__ inline void voxel_coords(int bx, int by, int tid, int &i, int &j)
{
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    i = bx * kVoxTile + (wave & 1) * 8 + (lane & 7);
    j = by * kVoxTile + (wave >> 1) * 8 + (lane >> 3);
}

// A wave-uniform read of a plan table (per-view headers, per-line entries, column-group
// lists): read-only for the whole launch, so on the device it is read through the
// constant address space, which the compiler serves with scalar loads.  The per-instance
// kernels get those from their __restrict__ kernel arguments anyway; the batched kernels
// take their table pointers from a descriptor in memory, where a plain dereference would
// be a generic vector ("flat") load that every wave waits on.
// The copy goes word by word: a struct with its own copy constructor (float4) would
// otherwise bind the source to a generic reference and lose the address space.
template <class T>
__host__ __device__ inline T ro_load(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(sizeof(T) % 4 == 0, "tables of 32-bit words");
    const __attribute__((address_space(4))) unsigned *q = (const __attribute__((address_space(4))) unsigned *)p;
    unsigned w[sizeof(T) / 4];
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; ++k) w[k] = q[k];
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
#else
    return *p;
#endif
}

// A per-lane read of a plan table in global memory, through the global address space:
// vector loads, never generic ones, in the batched kernels.  16-byte entries (aligned in
// their tables) are read as one 16-byte vector, so that the load stays one instruction
// whichever of its words the caller uses.
template <class T>
__host__ __device__ inline T g_load(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    T r;
    if constexpr (sizeof(T) % 16 == 0) {
        const __attribute__((address_space(1))) u32x4 *q = (const __attribute__((address_space(1))) u32x4 *)p;
        u32x4 w[sizeof(T) / 16];
#pragma unroll
        for (unsigned k = 0; k < sizeof(T) / 16; ++k) w[k] = q[k];
        __builtin_memcpy(&r, w, sizeof(T));
    } else {
        static_assert(sizeof(T) % 4 == 0, "tables of 32-bit words");
        const __attribute__((address_space(1))) unsigned *q = (const __attribute__((address_space(1))) unsigned *)p;
        unsigned w[sizeof(T) / 4];
#pragma unroll
        for (unsigned k = 0; k < sizeof(T) / 4; ++k) w[k] = q[k];
        __builtin_memcpy(&r, w, sizeof(T));
    }
    return r;
#else
    return *p;
#endif
}

// The first three words {d, sigma, kappa} of a column-table entry (w = 0): what the voxel
// gathers use of it, as one 12-byte load.
__host__ __device__ inline float4 col_load3(const float4 *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef float f32x3 __attribute__((ext_vector_type(3)));
    const f32x3 v = *(const __attribute__((address_space(1))) f32x3 *)p;
    return make_float4(v.x, v.y, v.z, 0.0f);
#else
    return make_float4(p->x, p->y, p->z, 0.0f);
#endif
}

struct CellArgs {
    int nx, ny, nz, nu, nv, nview, maxline, nbatch;
    float v0, dvv, zoff, sod;
    float dx, dy, shx, shy;
    unsigned plane8;  // Nx*Ny*8: byte stride of a z-plane (below 2^24 unless the plan is big)
    int band;         // cell launch order: 1 = row tiles outermost (tables larger than kCellBandBytes)
};

// z-prefix tables (one orientation, one batch item) above this size are walked with the
// row tiles outermost in the cell gathers' launch order: the workgroups in flight then
// share a band of z-planes, which stays in the 256 MiB last-level cache; smaller tables
// keep each group quad's row tiles together.
constexpr uint64_t kCellBandBytes = 128ull << 20;

struct VoxArgs {
    int nx, ny, nz, nu, nv, nview, maxline, nbatch, nchunk;
    float zb, dz, rv0, sod;
    float zc, kmid;  // volume centre z and Nz / 2: slice boundary k at (k - kmid) dz + zc
    float dx, dy, shx, shy;
    float dvv;
    float hscale;  // dvv / dz
    int vsplit;    // view sets: set s sums views [s * vper, min(V, (s + 1) * vper))
    int vper;
};

struct SinoArgs {
    int nu, nv, nview, nbatch;
    float v0, dvv;
};

// Operator -> kernel route.  Parallel-beam 3D backprojection is the transpose of the
// projection and its weighted form equals the plain one (src/cpu/backward3d.cpp:101-135,
// src/kernels/forward3d_parallel.cu:201-219); 2D parallel has its own backprojection
// (src/cpu/backward2d.cpp:165-205) whose weighted form is again the plain one.
struct Route {
    bool cell;  // cell gather (writes sinogram) vs voxel gather (writes volume)
    int kop;    // CellOp or VoxOp, -1 for an unknown operator
};

inline Route route(const Plan3d &p, int op)
{
    const bool par3d = p.kind == GEO_PAR && !p.is2d;
    switch (op) {
    case XR_OP_PROJECTION: return {true, CELL_A};
    case XR_OP_PROJECTION_T: return {false, VOX_AT};
    case XR_OP_BACKPROJECTION: return par3d ? Route{false, VOX_AT} : Route{false, VOX_B};
    case XR_OP_BACKPROJECTION_W:
        if (par3d) return {false, VOX_AT};
        return p.kind == GEO_PAR ? Route{false, VOX_B} : Route{false, VOX_BW};
    case XR_OP_BACKPROJECTION_T: return par3d ? Route{true, CELL_A} : Route{true, CELL_BT};
    case XR_OP_BACKPROJECTION_W_T:
        if (par3d) return {true, CELL_A};
        return p.kind == GEO_PAR ? Route{true, CELL_BT} : Route{true, CELL_BWT};
    }
    return {true, -1};
}

__host__ __device__ inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Workspace of the cell gathers: z-prefix of the volume in one orientation, plus
// kVolPad zeroed pad entries after the table (the fixed-span loads of the cell gather may
// read up to three entries past the last pixel, with weight zero).
constexpr int kVolPad = 4;

inline size_t vol_prefix_entries(const Plan3d &p, int nbatch)
{
    return (size_t)nbatch * (p.vol.nz + 1) * (size_t)p.vol.nx * p.vol.ny;
}

inline size_t vol_prefix_bytes(const Plan3d &p, int nbatch)
{
    return (vol_prefix_entries(p, nbatch) + kVolPad) * sizeof(float2);
}

// Workspace of the voxel gathers: row-prefix of the sinogram, plus kSinoPad entries after
// it (the paired staging loads may read one entry past a view's last column; it is
// never used).
constexpr int kSinoPad = 2;

inline size_t sino_prefix_bytes(const Plan3d &p, int nbatch)
{
    return ((size_t)nbatch * p.nview * (size_t)(p.geo.nv + 1) * p.geo.nu + kSinoPad) * sizeof(float2);
}

// View splitting of the voxel gathers.  A plan whose volume offers fewer than
// kVoxMinGroups workgroups (16 x 16 voxel columns x default slice chunk) splits its views
// into up to 16 sets, each summed by its own workgroups into a partial volume; a second
// pass adds the partials in set order.  Deterministic, and identical in the batched path.
// The tiles' work is uneven (tiles outside the field of view pass over most views), so
// a grid of a few rounds of workgroups balances better than one: the 256^2 x 64 cone-flat
// B (1024 tiles x chunks) takes 2.66 ms unsplit, 2.65 / 2.56 / 2.51 / 2.54 ms in 2 / 4 /
// 8 / 16 sets.
constexpr int kVoxMinGroups = 8192;

inline size_t split_partial_bytes(const Plan3d &p, int nbatch)
{
    if (p.vsplit <= 1) return 0;
    return (size_t)p.vsplit * nbatch * (size_t)p.vol.nz * p.vol.ny * p.vol.nx * sizeof(float);
}

// Default tile shapes (the autotuner may choose others).
inline int default_rows(const Plan3d &p)
{
    const int nv = p.geo.nv;
    if (nv >= 32) return 8;
    if (nv >= 16) return 4;
    if (nv >= 4) return 2;
    return 1;
}

// Voxel-gather chunk: slices per thread in the low bits, XR_VOXEL_HALVES for chunks run
// as two wave-culled halves (16 slices: halves of 8; 32 slices: halves of 16).  Halves
// pay where many chunks straddle the detector's edge rows, and 32-slice chunks halve the
// per-view work per slice (column search, weights, staging) on volumes tall enough to
// keep the GPU busy: B per item of the 512^2 x 256 cone-arc workload 32.1 ms (16), 28.4
// (16 in halves), 23.4 (32 in halves); on 256^2 x 64 plain 16-slice chunks win (2.71 /
// 2.85 / 2.99 ms).  The autotuner tries every form.
inline int chunk_slices(int k) { return k & 0xff; }
inline bool chunk_halves(int k) { return (k & XR_VOXEL_HALVES) != 0; }

inline int default_chunk(const Plan3d &p)
{
    const int nz = p.vol.nz;
    if (nz >= 128) return 32 | XR_VOXEL_HALVES;
    if (nz >= 64) return 16;
    if (nz >= 8) return 8;
    if (nz >= 4) return 4;
    return 1;
}

inline bool valid_rows(int r) { return r == 1 || r == 2 || r == 4 || r == 8 || r == 16; }
inline bool valid_chunk(int k)
{
    return k == 1 || k == 4 || k == 8 || k == 16 || k == (16 | XR_VOXEL_HALVES) || k == (32 | XR_VOXEL_HALVES);
}

inline CellArgs make_cell_args(const Plan3d &p, int nbatch)
{
    CellArgs a;
    a.nx = p.vol.nx; a.ny = p.vol.ny; a.nz = p.vol.nz; a.nu = p.geo.nu; a.nv = p.geo.nv;
    a.nview = p.nview; a.maxline = p.maxline; a.nbatch = nbatch;
    a.v0 = p.v0; a.dvv = p.dvv; a.zoff = p.zoff; a.sod = p.sod;
    a.dx = p.vol.dx; a.dy = p.vol.dy; a.shx = p.vol.shift_x; a.shy = p.vol.shift_y;
    a.plane8 = (unsigned)((uint64_t)p.vol.nx * p.vol.ny * 8u);
    a.band = (uint64_t)(p.vol.nz + 1) * p.vol.nx * p.vol.ny * 8u > kCellBandBytes ? 1 : 0;
    return a;
}

inline VoxArgs make_vox_args(const Plan3d &p, int nbatch, int KC)
{
    VoxArgs a;
    a.nx = p.vol.nx; a.ny = p.vol.ny; a.nz = p.vol.nz; a.nu = p.geo.nu; a.nv = p.geo.nv;
    a.nview = p.nview; a.maxline = p.maxline; a.nbatch = nbatch;
    KC = chunk_slices(KC);
    a.nchunk = (a.nz + KC - 1) / KC;
    a.zb = p.zb; a.dz = p.dz; a.rv0 = p.rv0; a.sod = p.sod;
    a.zc = p.zc; a.kmid = 0.5f * (float)a.nz;
    a.dx = p.vol.dx; a.dy = p.vol.dy; a.shx = p.vol.shift_x; a.shy = p.vol.shift_y;
    a.dvv = p.dvv;
    a.hscale = p.dvv / p.dz;
    a.vsplit = p.vsplit;
    a.vper = (p.nview + p.vsplit - 1) / p.vsplit;
    return a;
}

inline SinoArgs make_sino_args(const Plan3d &p, int nbatch)
{
    SinoArgs a;
    a.nu = p.geo.nu; a.nv = p.geo.nv; a.nview = p.nview; a.nbatch = nbatch;
    a.v0 = p.v0; a.dvv = p.dvv;
    return a;
}

// clamp(x, lo, hi) for finite x: one v_med3_f32 on the device.
// min / max of values that are never NaN, as one instruction on the device (fminf and
// fmaxf of values read from memory cost an extra quieting instruction per operand there).
__host__ __device__ inline float minf_finite(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
#else
    return fminf(a, b);
#endif
}
__host__ __device__ inline float maxf_finite(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
#else
    return fmaxf(a, b);
#endif
}

__host__ __device__ inline float clampf(float x, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
    return fminf(fmaxf(x, lo), hi);
#endif
}

// a * b for operands below 2^24: one full-rate v_mul_u32_u24 on the device (a 32-bit
// v_mul_lo_u32 issues at quarter rate).  Table offsets here stay far below 2^24 per
// factor: rows/slices <= 2^16, plane strides < 2^24 (checked at plan creation).
__host__ __device__ inline unsigned mul24(unsigned a, unsigned b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// a * b + c for a, b below 2^24: one full-rate v_mad_u32_u24 on the device.  Written as
// an instruction because the compiler otherwise turns some of these (runtime operands it
// cannot bound) into a mask and a quarter-rate v_mul_lo_u32.
// A value that is the same on every lane of the wave, marked so for the compiler (kept in
// a scalar register even when it was loaded through a pointer the compiler cannot prove
// uniform, as in the batched kernels' descriptor tables).
__host__ __device__ inline int wave_uniform(int x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(x);
#else
    return x;
#endif
}

__host__ __device__ inline unsigned mad24(unsigned a, unsigned b, unsigned c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
#else
    return a * b + c;
#endif
}

// Two consecutive float2 table entries (pixels m and m+1 of a line) in one 16-byte load.
// The entries are only 8-byte aligned, which gfx950 vector memory accesses allow.
struct __attribute__((aligned(8))) Pair2 {
    float x0, y0, x1, y1;
};

// Block of a prefix table staged in LDS: columns c0 .. c0+nc-1 and rows r0 .. r0+nr-1,
// stored row-major with an even pitch >= nc (whole 16-byte pairs per row).  nc = 0:
// nothing staged.  Voxel gather: a view's row-prefix table (columns = detector columns,
// rows = detector rows); skip = 1: no voxel of the workgroup's tile and slice chunk
// reaches the detector on this view (bounds, not estimates: see VoxTile), so the
// workgroup passes over the view.  Cell gather: a line of the z-prefix table (columns =
// pixels along the line, rows = slices); skip unused.
struct VoxBox {
    int c0, nc, r0, nr, pitch, skip;
};

#if !defined(__HIP_DEVICE_COMPILE__)
// Host builds: staged cell-gather lookups found outside their LDS block (cell_line).
inline std::atomic<long long> cell_box_misses{0};
#endif

// A float2 at byte offset off of an LDS block / of a table in global memory.  On the
// device the loads name their address space, so that the compiler keeps them apart
// (ds_read vs global_load) instead of merging the two into generic "flat" loads through
// a selected pointer.
__host__ __device__ inline float2 lds_float2(const char *base, unsigned off)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(3))) float2 *)((const __attribute__((address_space(3))) char *)base + off);
#else
    return *reinterpret_cast<const float2 *>(base + off);
#endif
}
__host__ __device__ inline float2 global_float2(const char *base, unsigned off)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) float2 *)((const __attribute__((address_space(1))) char *)base + off);
#else
    return *reinterpret_cast<const float2 *>(base + off);
#endif
}
__host__ __device__ inline Pair2 lds_pair(const char *base, unsigned off)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(3))) Pair2 *)((const __attribute__((address_space(3))) char *)base + off);
#else
    return *reinterpret_cast<const Pair2 *>(base + off);
#endif
}

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }

// x - floor(x) for x >= 0: one v_fract_f32 on the device.
__host__ __device__ inline float fractf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fractf(x);
#else
    return x - floorf(x);
#endif
}

// 1 / y where an error of about one ulp is harmless: v_rcp_f32 on the device (a true
// division compiles to a ten-instruction correctly rounded sequence), a division on the
// host.
__host__ __device__ inline float approx_rcp(float y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(y);
#else
    return 1.0f / y;
#endif
}

// x / y where the quotient only seeds a search that is corrected against exact tables:
// x times v_rcp_f32(y) on the device (about 2 ulp), a true division on the host.
__host__ __device__ inline float seed_div(float x, float y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return x * __builtin_amdgcn_rcpf(y);
#else
    return x / y;
#endif
}

// FDK distance weight U of a voxel centre p for a view with source s
// (include/xrmath.h:76-86): flat SOD^2/(SOD - s.p/SOD)^2, arc 1/|s - p|^2.
template <int GEO>
__host__ __device__ inline float fdk_weight(float sx, float sy, float px, float py, float sod)
{
    if (GEO == GEO_FLAT) {
        // sod^2 / (sod - (s.p)/sod)^2 written as 1 / (1 - (s.p)/sod^2)^2: one division per
        // call (1/sod^2 is loop-invariant and hoisted)
        const float t = fmaf(-(sx * px + sy * py), 1.0f / (sod * sod), 1.0f);
        return 1.0f / (t * t);
    } else if (GEO == GEO_ARC) {
        const float ex = sx - px, ey = sy - py;
        return 1.0f / (ex * ex + ey * ey);
    }
    return 1.0f;
}

// ---------------------------------------------------------------- prefix tables

// The prefix passes read a column kPrefixChunk entries at a time into registers, so that
// a chunk's loads are in flight together; the sums still run over the entries in order.
constexpr int kPrefixChunk = 32;

// z-prefix of one volume column (b, e = j*nx + i) in the x-driven line orientation:
// W[b][k][j][i] = (J_k - k f_k, f_k), J_k = sum_{k'<k} f[k'] - total/2 (FP64 accumulation,
// the entry rounded once), so that the depth integral up to z in [k, k+1] is
// J_k + (z - k) f_k = fmaf(z, f_k, W.x): one fmaf, no fract.  Entry Nz is (J_Nz, 0).
__host__ __device__ inline void vol_prefix_body(const float *__restrict__ f, float2 *__restrict__ W, int64_t plane,
                                                int nz, int64_t b, int64_t e)
{
    const float *src = f + b * plane * nz + e;
    float2 *dst = W + b * plane * (nz + 1) + e;
    double total = 0.0;
    for (int k0 = 0; k0 < nz; k0 += kPrefixChunk) {
        float v[kPrefixChunk];
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k) v[k] = k0 + k < nz ? src[(int64_t)(k0 + k) * plane] : 0.0f;
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k)
            if (k0 + k < nz) total += (double)v[k];
    }
    double run = -0.5 * total;
    for (int k0 = 0; k0 < nz; k0 += kPrefixChunk) {
        float v[kPrefixChunk];
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k) v[k] = k0 + k < nz ? src[(int64_t)(k0 + k) * plane] : 0.0f;
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k)
            if (k0 + k < nz) {
                dst[(int64_t)(k0 + k) * plane] = make_float2((float)(run - (double)(k0 + k) * (double)v[k]), v[k]);
                run += (double)v[k];
            }
    }
    dst[(int64_t)nz * plane] = make_float2((float)run, 0.0f);
}

// Forward-projection cell normalisation applied to a sinogram sample before the A^T
// gather: Delta / (W_c cos theta_cr) (src/cpu/forward3d.cpp:330-344; the row height
// cancels against the depth overlap, which the voxel gather measures in row units).
template <int GEO>
__host__ __device__ inline float at_scale(const ViewHdr &h, float scale0, float dbp, float sg, float v0, float dvv,
                                          int r)
{
    if (GEO == GEO_PAR) return scale0;
    const float hb = sg * (v0 + ((float)r + 0.5f) * dvv);
    const float cosw = fabsf(h.S) / sqrtf(dbp * dbp + hb * hb + h.S * h.S);
    return scale0 / cosw;
}

// Row-prefix table entries of internal column c of view v (header h) from the samples
// q(r), r = 0 .. Nv-1, written to dst[r * Nu], r = 0 .. Nv: entry r = (B_r, q_r) with
//   Q_r = sum_{r' < r} q_r' - total / 2   (FP64, centred),   B_r = Q_r - r q_r,
// so that the row integral up to a fractional row rho in [r, r+1] is
//   I(rho) = Q_r + (rho - r) q_r = B_r + rho q_r    (one fmaf, no fract),
// and entry Nv = (Q_Nv, 0).  A^T mode scales each sample by the forward normalisation
// first.
template <int MODE, int GEO, class QOf>
__host__ __device__ inline void prefix_column(const ViewHdr &h, const float4 *__restrict__ colt, const SinoArgs &a,
                                              int v, int c, QOf qof, float2 *__restrict__ dst)
{
    float scale0 = 1.0f, dbp = 0.0f, sg = 1.0f;
    if (MODE == SINO_AT) {
        const float4 cc = colt[(int64_t)v * (a.nu + 1) + c];
        const float dn = colt[(int64_t)v * (a.nu + 1) + c + 1].x;
        scale0 = h.delta / cc.w;
        dbp = 0.5f * (cc.x + dn) - h.P;
        sg = cc.y;
    }
    double total = 0.0;
    for (int r0 = 0; r0 < a.nv; r0 += kPrefixChunk) {
        float q[kPrefixChunk];
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k) q[k] = r0 + k < a.nv ? qof(r0 + k) : 0.0f;
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k)
            if (r0 + k < a.nv) {
                if (MODE == SINO_AT) q[k] *= at_scale<GEO>(h, scale0, dbp, sg, a.v0, a.dvv, r0 + k);
                total += (double)q[k];
            }
    }
    double run = -0.5 * total;
    for (int r0 = 0; r0 < a.nv; r0 += kPrefixChunk) {
        float q[kPrefixChunk];
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k) q[k] = r0 + k < a.nv ? qof(r0 + k) : 0.0f;
#pragma unroll
        for (int k = 0; k < kPrefixChunk; ++k)
            if (r0 + k < a.nv) {
                const int r = r0 + k;
                if (MODE == SINO_AT) q[k] *= at_scale<GEO>(h, scale0, dbp, sg, a.v0, a.dvv, r);
                dst[(int64_t)r * a.nu] = make_float2((float)(run - (double)r * (double)q[k]), q[k]);
                run += (double)q[k];
            }
    }
    dst[(int64_t)a.nv * a.nu] = make_float2((float)run, 0.0f);
}

// Row-prefix table of one sinogram column (b, v, internal column c), in internal column
// order, from the sinogram p: q_r = p[b][v][r][col(c)] (times the A^T scale).
template <int MODE, int GEO>
__host__ __device__ inline void sino_prefix_body(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                 const float *__restrict__ p, float2 *__restrict__ Q,
                                                 const SinoArgs &a, int b, int v, int c)
{
    const ViewHdr h = hdr[v];
    const int col = h.keep ? c : (a.nu - 1 - c);
    const int64_t sv = (int64_t)b * a.nview + v;
    const float *src = p + sv * a.nv * a.nu + col;
    prefix_column<MODE, GEO>(h, colt, a, v, c, [src, &a](int r) { return src[(int64_t)r * a.nu]; },
                             Q + sv * (int64_t)(a.nv + 1) * a.nu + c);
}

// Prefix columns of at most kPrefixRegRows entries keep their samples in registers on the
// device: one read of the input instead of two, all loads in flight at once.  The
// operations and their order are those of vol_prefix_body / prefix_column.
constexpr int kPrefixRegRows = 64;

__host__ __device__ inline void vol_prefix_reg(const float *__restrict__ f, float2 *__restrict__ W, int64_t plane,
                                               int nz, int64_t b, int64_t e)
{
    const float *src = f + b * plane * nz + e;
    float2 *dst = W + b * plane * (nz + 1) + e;
    float v[kPrefixRegRows];
#pragma unroll
    for (int k = 0; k < kPrefixRegRows; ++k) v[k] = k < nz ? src[(int64_t)k * plane] : 0.0f;
    double total = 0.0;
#pragma unroll
    for (int k = 0; k < kPrefixRegRows; ++k)
        if (k < nz) total += (double)v[k];
    double run = -0.5 * total;
#pragma unroll
    for (int k = 0; k < kPrefixRegRows; ++k)
        if (k < nz) {
            dst[(int64_t)k * plane] = make_float2((float)(run - (double)k * (double)v[k]), v[k]);
            run += (double)v[k];
        }
    dst[(int64_t)nz * plane] = make_float2((float)run, 0.0f);
}

template <int MODE, int GEO>
__host__ __device__ inline void sino_prefix_reg(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                const float *__restrict__ p, float2 *__restrict__ Q, const SinoArgs &a,
                                                int b, int v, int c)
{
    const ViewHdr h = hdr[v];
    const int col = h.keep ? c : (a.nu - 1 - c);
    const int64_t sv = (int64_t)b * a.nview + v;
    const float *src = p + sv * a.nv * a.nu + col;
    float2 *dst = Q + sv * (int64_t)(a.nv + 1) * a.nu + c;
    float scale0 = 1.0f, dbp = 0.0f, sg = 1.0f;
    if (MODE == SINO_AT) {
        const float4 cc = colt[(int64_t)v * (a.nu + 1) + c];
        const float dn = colt[(int64_t)v * (a.nu + 1) + c + 1].x;
        scale0 = h.delta / cc.w;
        dbp = 0.5f * (cc.x + dn) - h.P;
        sg = cc.y;
    }
    float q[kPrefixRegRows];
#pragma unroll
    for (int r = 0; r < kPrefixRegRows; ++r) {
        q[r] = r < a.nv ? src[(int64_t)r * a.nu] : 0.0f;
        if (MODE == SINO_AT && r < a.nv) q[r] *= at_scale<GEO>(h, scale0, dbp, sg, a.v0, a.dvv, r);
    }
    double total = 0.0;
#pragma unroll
    for (int r = 0; r < kPrefixRegRows; ++r)
        if (r < a.nv) total += (double)q[r];
    double run = -0.5 * total;
#pragma unroll
    for (int r = 0; r < kPrefixRegRows; ++r)
        if (r < a.nv) {
            dst[(int64_t)r * a.nu] = make_float2((float)(run - (double)r * (double)q[r]), q[r]);
            run += (double)q[r];
        }
    dst[(int64_t)a.nv * a.nu] = make_float2((float)run, 0.0f);
}

// ---------------------------------------------------------------- FDK filter stage

// Largest detector width the fused FDK filter stage handles (per-thread work grows with
// the width; wider detectors take the composite path).
constexpr int kFdkMaxNu = 1024;

// The filter stage reads its taps from an FP64 table with kFdkTapPad zeros on each side
// (the last row and column blocks index up to that far outside the 2 Nu - 1 taps).
constexpr int kFdkTapPad = 64;

inline size_t fdk_taps_bytes(const Plan3d &p) { return (size_t)(2 * p.geo.nu - 1 + 2 * kFdkTapPad) * sizeof(double); }

// Filtered, weighted sample of detector column u (physical order) of one row p:
//   f[u] = sum_s p[s] w[s] h[s - u + Nu - 1],  s ascending, products and sum in FP64,
// which is the reference's zero-padded correlation of the cosine-weighted row with the
// 2 Nu - 1 tap filter (projector3d.py:189-190, projector.py:200-201).  k_fdk_filter
// computes the same sums in the same order.
__host__ __device__ inline double fdk_filter_output(const float *__restrict__ p, const float *__restrict__ w,
                                                    const float *__restrict__ h, int nu, int u)
{
    double acc = 0.0;
    for (int s = 0; s < nu; ++s) acc = fma((double)p[s] * (double)w[s], (double)h[s - u + nu - 1], acc);
    return acc;
}

// In-place row-prefix of filtered values already stored in Q[.].y, for internal column
// c of view v, batch item b.
template <int MODE, int GEO>
__host__ __device__ inline void fdk_prefix_body(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ colt,
                                                float2 *__restrict__ Q, const SinoArgs &a, int b, int v, int c)
{
    const int64_t sv = (int64_t)b * a.nview + v;
    float2 *dst = Q + sv * (int64_t)(a.nv + 1) * a.nu + c;
    prefix_column<MODE, GEO>(hdr[v], colt, a, v, c, [dst, &a](int r) { return dst[(int64_t)r * a.nu].y; }, dst);
}

// ---------------------------------------------------------------- cell gather

// Per-thread state of the cell gather: internal column c (boundaries d0, d1, table entry
// cc), rows r0..r0+R-1 (row boundary heights hv on the axis), accumulators.
template <int R>
struct CellLane {
    float4 cc;
    float d0, d1;
    float hv[R + 1];
    float acc[R];
};

template <int R>
__host__ __device__ inline void cell_init(CellLane<R> &st, const float4 *__restrict__ ct, const CellArgs &a, int c,
                                          int r0)
{
    st.cc = g_load(ct + c);
    st.d0 = st.cc.x;
    st.d1 = g_load(&ct[c + 1].x);
    for (int q = 0; q <= R; ++q) st.hv[q] = st.cc.y * (a.v0 + (float)(r0 + q) * a.dvv);
    for (int q = 0; q < R; ++q) st.acc[q] = 0.0f;
}

// One line L (table coefficients lc) of the cell gather for this thread.  W: the batch
// item's z-prefix table in the view's orientation.  SPAN = 2, 3 or 4: the thread's column
// overlaps at most SPAN consecutive pixels of this line (the caller's span test) and the
// line takes a fixed SPAN-pixel path; SPAN = 0: the general pixel loop.  BIG: 64-bit
// table offsets (entries), otherwise 32-bit byte offsets over the wave-uniform line base.
//
// LDS (SPAN 2 and 3 only): the lookups come from the block bx of this line's table staged
// at lds (bx.r0 .. bx.r0+bx.nr-1 slices of bx.pitch entries from pixel bx.c0, an exact
// copy of the table's entries, so the results are those of the table).  The block holds
// every lookup of the thread by construction (cell_box); the host build counts lookups
// outside it in cell_box_misses, which the tests check stays zero.
template <int OPK, int GEO, int R, int SPAN, bool BIG, bool LDS = false>
__host__ __device__ inline void cell_line(CellLane<R> &st, const ViewHdr &h, const CellArgs &a, const float4 &lc,
                                          int L, const float2 *__restrict__ W, const char *lds = nullptr,
                                          const VoxBox &bx = VoxBox{0, 0, 0, 0, 0, 0})
{
    static_assert(!LDS || ((SPAN == 2 || SPAN == 3) && !BIG), "LDS lookups: fixed two- or three-pixel paths");
    const int nalong = wave_uniform(h.nalong);
    const float fal = (float)nalong;
    // Column boundaries in pixel units along the line.
    const float t0 = fmaf(st.d0, lc.y, lc.x);
    const float t1 = fmaf(st.d1, lc.y, lc.x);
    if (!(t1 > 0.0f && t0 < fal)) return;

    // Weight of pixel m: its column overlap ou (pixel units) times the operator's factor.
    auto pixel_weight = [&](float ou, int m) -> float {
        if (OPK == CELL_A) return ou * lc.w;
        if (OPK == CELL_BWT) {
            float px, py;
            if (h.xdrv) {
                px = ((float)m - 0.5f * a.nx + 0.5f) * a.dx + a.shx;
                py = (0.5f * a.ny - (float)L - 0.5f) * a.dy + a.shy;
            } else {
                px = ((float)L - 0.5f * a.nx + 0.5f) * a.dx + a.shx;
                py = ((float)m - 0.5f * a.ny + 0.5f) * a.dy + a.shy;
            }
            return ou * fdk_weight<GEO>(h.sx, h.sy, px, py, a.sod);
        }
        return ou;
    };

    // Row boundaries in voxel units along z at this line's magnification: entry ki of the
    // z-prefix table, interpolated at z (J = fmaf(z, f, W.x), see vol_prefix_body).
    const float fnz = (float)a.nz;
    int ki[R + 1];
    float fr[R + 1];
#pragma unroll
    for (int q = 0; q <= R; ++q) {
        fr[q] = clampf(fmaf(st.hv[q], lc.z, a.zoff), 0.0f, fnz);
        ki[q] = (int)fr[q];
    }
    const int plane = h.nline * nalong;
    const float2 *Wl = W + (int64_t)L * nalong;
    const char *Wb = reinterpret_cast<const char *>(Wl);

    if (SPAN == 2 || SPAN == 3 || SPAN == 4) {
        const int m0 = t0 > 0.0f ? (int)t0 : 0;
#if !defined(__HIP_DEVICE_COMPILE__)
        if (LDS) {
            const int klo = imin(ki[0], ki[R]), khi = imax(ki[0], ki[R]);  // ki[] is monotone in q
            if (!(m0 >= bx.c0 && m0 + (SPAN - 1) < bx.c0 + bx.pitch && klo >= bx.r0 && khi < bx.r0 + bx.nr))
                cell_box_misses.fetch_add(1);
        }
#endif
        // Pixel weights of m0 .. m0+SPAN-1 (the column spans less than SPAN-1 pixels, so
        // it ends before m0+SPAN); entries past the line's end are weighted by zero.
        float w0, w1, w2 = 0.0f, w3 = 0.0f;
        w0 = pixel_weight(fminf(t1, (float)(m0 + 1)) - fmaxf(t0, (float)m0), m0);
        if (SPAN == 2) {
            w1 = m0 + 1 < nalong ? pixel_weight(fmaxf(t1 - (float)(m0 + 1), 0.0f), m0 + 1) : 0.0f;
        } else {
            w1 = m0 + 1 < nalong ? pixel_weight(fmaxf(fminf(t1, (float)(m0 + 2)) - (float)(m0 + 1), 0.0f), m0 + 1)
                                 : 0.0f;
            if (SPAN == 3) {
                w2 = m0 + 2 < nalong ? pixel_weight(fmaxf(t1 - (float)(m0 + 2), 0.0f), m0 + 2) : 0.0f;
            } else {
                w2 = m0 + 2 < nalong
                         ? pixel_weight(fmaxf(fminf(t1, (float)(m0 + 3)) - (float)(m0 + 2), 0.0f), m0 + 2)
                         : 0.0f;
                w3 = m0 + 3 < nalong ? pixel_weight(fmaxf(t1 - (float)(m0 + 3), 0.0f), m0 + 3) : 0.0f;
            }
        }
        // Lookups per row boundary: entries m0, m0+1 (one 16-byte pair), and m0+2 (three
        // pixels) or m0+2, m0+3 (a second pair, four pixels).
        Pair2 e[R + 1];
        float2 g[R + 1];
        Pair2 e2[R + 1];
#pragma unroll
        for (int q = 0; q <= R; ++q) {
            if (LDS) {
                // (ki - r0) pitch 8 + (m0 - c0) 8, the row base folded into the addend
                // (modulo 2^32; the offset itself lies in the block)
                const unsigned off = mad24((unsigned)ki[q], (unsigned)bx.pitch * 8u,
                                           (unsigned)(m0 - bx.c0) * 8u - (unsigned)(bx.r0 * bx.pitch) * 8u);
                e[q] = lds_pair(lds, off);
                if (SPAN == 3) g[q] = lds_float2(lds, off + 16u);
                continue;
            }
            const float2 *pq;
            if (BIG) pq = Wl + ((int64_t)mul24((unsigned)ki[q], (unsigned)plane) + m0);
            else pq = reinterpret_cast<const float2 *>(Wb + mad24((unsigned)ki[q], a.plane8, (unsigned)m0 * 8u));
            e[q] = *reinterpret_cast<const Pair2 *>(pq);
            if (SPAN == 3) g[q] = pq[2];
            if (SPAN == 4) e2[q] = *reinterpret_cast<const Pair2 *>(pq + 2);
        }
        float J0 = fmaf(fr[0], e[0].y0, e[0].x0);
        float J1 = fmaf(fr[0], e[0].y1, e[0].x1);
        float J2 = SPAN == 3 ? fmaf(fr[0], g[0].y, g[0].x) : SPAN == 4 ? fmaf(fr[0], e2[0].y0, e2[0].x0) : 0.0f;
        float J3 = SPAN == 4 ? fmaf(fr[0], e2[0].y1, e2[0].x1) : 0.0f;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const float K0 = fmaf(fr[q + 1], e[q + 1].y0, e[q + 1].x0);
            const float K1 = fmaf(fr[q + 1], e[q + 1].y1, e[q + 1].x1);
            float t = fmaf(w1, K1 - J1, st.acc[q]);
            if (SPAN == 3) {
                const float K2 = fmaf(fr[q + 1], g[q + 1].y, g[q + 1].x);
                t = fmaf(w2, K2 - J2, t);
                J2 = K2;
            }
            if (SPAN == 4) {
                const float K2 = fmaf(fr[q + 1], e2[q + 1].y0, e2[q + 1].x0);
                const float K3 = fmaf(fr[q + 1], e2[q + 1].y1, e2[q + 1].x1);
                t = fmaf(w3, K3 - J3, t);
                t = fmaf(w2, K2 - J2, t);
                J2 = K2;
                J3 = K3;
            }
            st.acc[q] = fmaf(w0, K0 - J0, t);
            J0 = K0;
            J1 = K1;
        }
    } else {
        const int m0 = t0 > 0.0f ? (int)floorf(t0) : 0;
        const int m1 = t1 < fal ? (int)ceilf(t1) : nalong;
        for (int m = m0; m < m1; ++m) {
            const float w = pixel_weight(fminf(t1, (float)(m + 1)) - fmaxf(t0, (float)m), m);
            float2 e[R + 1];
#pragma unroll
            for (int q = 0; q <= R; ++q) {
                if (BIG) e[q] = Wl[(int64_t)mul24((unsigned)ki[q], (unsigned)plane) + m];
                else e[q] = *reinterpret_cast<const float2 *>(Wb + mad24((unsigned)ki[q], a.plane8, (unsigned)m * 8u));
            }
            float Jp = fmaf(fr[0], e[0].y, e[0].x);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const float Jn = fmaf(fr[q + 1], e[q + 1].y, e[q + 1].x);
                st.acc[q] = fmaf(w, Jn - Jp, st.acc[q]);
                Jp = Jn;
            }
        }
    }
}

// Writes the thread's R detector cells (output column order restored).
template <int OPK, int GEO, int R>
__host__ __device__ inline void cell_store(const CellLane<R> &st, const ViewHdr &h, const CellArgs &a,
                                           float *__restrict__ out, int c, int r0, int v, int b)
{
    const int col = h.keep ? c : (a.nu - 1 - c);
    float *o = out + ((int64_t)b * a.nview + v) * a.nv * a.nu + col;
    for (int q = 0; q < R; ++q) {
        const int r = r0 + q;
        if (r >= a.nv) break;
        float s;
        if (OPK == CELL_A) {
            // Delta / (W_c H_cr cos theta_cr) with H = 1/kappa (src/cpu/forward3d.cpp:330-344)
            s = h.delta * st.cc.z / st.cc.w;
            if (GEO != GEO_PAR) {
                const float dbp = 0.5f * (st.d0 + st.d1) - h.P;
                const float hb = st.cc.y * (a.v0 + ((float)r + 0.5f) * a.dvv);
                const float cosw = fabsf(h.S) / sqrtf(dbp * dbp + hb * hb + h.S * h.S);
                s /= cosw;
            }
        } else {
            s = h.colw;
        }
        o[(int64_t)r * a.nu] = st.acc[q] * s;
    }
}

// The z-prefix table of batch item b in the view's orientation (the same for every lane
// of a wave: the header fields are marked wave-uniform so that the table base stays in
// scalar registers also where the header was read through a descriptor table).
__host__ __device__ inline const float2 *cell_table(const float2 *WX, const float2 *WY, const ViewHdr &h,
                                                    const CellArgs &a, int b)
{
    const int64_t plane = (int64_t)wave_uniform(h.nline) * wave_uniform(h.nalong);
    return (wave_uniform(h.xdrv) ? WX : WY) + (int64_t)b * (int64_t)(a.nz + 1) * plane;
}

// One thread, reading the tables in HBM only: internal column c, rows r0..r0+R-1 of view
// v, batch b.  SPAN: the class of the thread's column group (CellGroup::span), whose
// widest column spans wgrp on the axis; in a group of class 3 or 0 each line takes the
// narrowest fixed path (two, three or four pixels) its pixel footprint allows, the
// general loop beyond (wgrp = +inf: the general loop on every line).  BIG as for
// cell_line.
template <int OPK, int GEO, int R, int SPAN, bool BIG>
__host__ __device__ inline void cell_body(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineC,
                                          const float4 *__restrict__ colt, const float2 *__restrict__ WX,
                                          const float2 *__restrict__ WY, float *__restrict__ out, const CellArgs &a,
                                          int c, int r0, int v, int b, float wgrp)
{
    const ViewHdr h = ro_load(hdr + v);
    CellLane<R> st;
    cell_init<R>(st, colt + (int64_t)v * (a.nu + 1), a, c, r0);
    const float2 *W = cell_table(WX, WY, h, a, b);
    const float4 *lt = lineC + (int64_t)v * a.maxline;
    for (int L = 0; L < h.nline; ++L) {
        const float4 lc = ro_load(lt + L);
        if (SPAN == 2) {
            cell_line<OPK, GEO, R, 2, BIG>(st, h, a, lc, L, W);
            continue;
        }
        // wgrp / Delta_L: the group's widest column in pixels of this line (wave-uniform)
        const float s = wgrp * lc.y;
        if (s < 0.99f) cell_line<OPK, GEO, R, 2, BIG>(st, h, a, lc, L, W);
        else if (SPAN == 3 || s < 1.99f) cell_line<OPK, GEO, R, 3, BIG>(st, h, a, lc, L, W);
        else if (s < 2.99f) cell_line<OPK, GEO, R, 4, BIG>(st, h, a, lc, L, W);
        else cell_line<OPK, GEO, R, 0, BIG>(st, h, a, lc, L, W);
    }
    cell_store<OPK, GEO, R>(st, h, a, out, c, r0, v, b);
}

// The cell gather's LDS staging: kCellSlots slots of kCellLds float2 entries (12 KiB
// each, 36 KiB per workgroup), one line per slot, lines staged kCellSlots - 1 ahead of
// the line being gathered.  The 4 groups of a workgroup stage blocks of 230 entries on
// average (616 at most) per line of the 256^2 x 64 cone-flat workload for 8 rows; 555
// (1860) of the 512^2 x 256 cone-arc one, where 0.06 % of the blocks exceed a slot (their
// (workgroup, row tile) pairs, 2 % with those mixing orientations, read the table).
constexpr int kCellLds = 1536;
constexpr int kCellSlots = 3;
static_assert(kCellLds / 2 % kBlock == 0, "staging copies whole rounds of pairs");
constexpr int kCellChunk = 64;  // lines per staging-pitch chunk (one per lane of a wave)

// Copy rounds (global-to-LDS loads per thread) of a staged block (voxel_stage).
__host__ __device__ inline int stage_rounds(const VoxBox &bx)
{
    return (bx.nr * (bx.pitch >> 1) + kBlock - 1) / kBlock;
}

// Staging pitch of line L's block: its QuadLine's pixel count in whole pairs (0: nothing
// to stage).  A workgroup stages every line of a 64-line chunk at the chunk's largest
// pitch, so that each thread's share of the copies sits at the same place in every block
// of the chunk.
__host__ __device__ inline int cell_line_pitch(const QuadLine &q) { return q.nm > 0 ? (q.nm + 1) & ~1 : 0; }

// The block of line L's z-prefix table a cell-gather workgroup stages for its row
// boundaries r0 .. r0+R at pitch (>= the line's own pitch): pixels from the QuadLine's
// first, moved left where the block would reach more than three entries past the line's
// end (the table's last line is followed by kVolPad pad entries; a QuadLine covers at most
// Nalong + 2 pixels, so the block stays at or after the line's start); slices over the
// QuadLine's range of s = sigma * lc.z at the two extreme heights (z = s h + zoff is
// linear in h for each column, so its extremes over the rows lie at the end rows).  A
// thread evaluates z as fmaf(sigma * (v0 + r dvv), lc.z, zoff) and this bound as s h +
// zoff, both within a few units in the last place of |s h| + |zoff| of the exact value;
// the bound is widened by 1e-4 (|s h| + |zoff|) + 0.01 slices, far beyond that.  The
// block so holds every lookup of the workgroup's threads.  nc = 0: nothing staged (no
// column reaches the line, or the block exceeds a slot).
template <int R>
__host__ __device__ inline VoxBox cell_box(const QuadLine &q, const CellArgs &a, int r0, int pitch, int nalong)
{
    VoxBox bx{0, 0, 0, 0, 0, 0};
    if (q.nm <= 0) return bx;
    const float h0 = fmaf((float)r0, a.dvv, a.v0), h1 = fmaf((float)(r0 + R), a.dvv, a.v0);
    const float z0 = q.sg_lo * h0, z1 = q.sg_hi * h0, z2 = q.sg_lo * h1, z3 = q.sg_hi * h1;
    const float plo = fminf(fminf(z0, z1), fminf(z2, z3)), phi = fmaxf(fmaxf(z0, z1), fmaxf(z2, z3));
    if (!(plo <= phi)) return bx;
    const float tol = fmaf(1e-4f, fmaxf(fabsf(plo), fabsf(phi)) + fabsf(a.zoff), 0.01f);
    const float fnz = (float)a.nz;
    const int s0 = (int)clampf(floorf(plo + a.zoff - tol), 0.0f, fnz);
    const int s1 = (int)clampf(floorf(phi + a.zoff + tol), 0.0f, fnz);
    const int ns = s1 - s0 + 1;
    if (ns * pitch > kCellLds) return bx;
    bx.c0 = imin(q.m_lo, nalong + 3 - pitch);
    if (bx.c0 < 0) return VoxBox{0, 0, 0, 0, 0, 0};  // (pitch above Nalong + 3: not a QuadLine's)
    bx.nc = pitch;
    bx.r0 = s0;
    bx.nr = ns;
    bx.pitch = pitch;
    return bx;
}

// Chunk pitches of a workgroup's lines (host): pitch[L / 64] for lines 0 .. nline-1.
inline void cell_chunk_pitches(const QuadLine *ql, int nline, int *pitch)
{
    for (int c = 0; c * kCellChunk < nline; ++c) {
        int p = 0;
        for (int L = c * kCellChunk; L < imin(nline, (c + 1) * kCellChunk); ++L) p = imax(p, cell_line_pitch(ql[L]));
        pitch[c] = p;
    }
}

// Row tiles of R rows (4, 8, 16) of all three sizes, as the plan lists the cell gather's
// staging flags per workgroup: tiles of 4 rows first, then 8, then 16.
__host__ __device__ inline int cell_tile_count(int nv) { return ceil_div(nv, 4) + ceil_div(nv, 8) + ceil_div(nv, 16); }
__host__ __device__ inline int cell_tile_offset(int nv, int R)
{
    return R == 4 ? 0 : R == 8 ? ceil_div(nv, 4) : ceil_div(nv, 4) + ceil_div(nv, 8);
}

// Line L of a cell-gather thread in a group of span class SPAN (2 or 3), each line on
// the path cell_body gives it, with the lookups from the staged block bx at lds (in a
// staged workgroup every line with a column on it has its block: the plan's flags).
template <int OPK, int GEO, int R, int SPAN>
__host__ __device__ inline void cell_line_staged(CellLane<R> &st, const ViewHdr &h, const CellArgs &a,
                                                 const float4 &lc, int L, const float2 *__restrict__ W, float wgrp,
                                                 const char *lds, const VoxBox &bx)
{
    static_assert(SPAN == 2 || SPAN == 3, "staged lines: span classes 2 and 3");
    if (SPAN == 2 || wgrp * lc.y < 0.99f) cell_line<OPK, GEO, R, 2, false, true>(st, h, a, lc, L, W, lds, bx);
    else cell_line<OPK, GEO, R, 3, false, true>(st, h, a, lc, L, W, lds, bx);
}

// ---------------------------------------------------------------- voxel gather

// First column index to visit for axis point a: the closed-form inverse of the column map
// (projective for flat and parallel, atan of a projective for arc).  A seed only: callers
// correct it against the exact FP32 table.
template <int GEO>
__host__ __device__ inline float column_seed(const ViewHdr &h, float a)
{
    const float T = seed_div(fmaf(h.tp1, a, h.tp0), fmaf(h.tq1, a, h.tq0));
    return h.tk0 + h.tk1 * (GEO == GEO_ARC ? atanf(T) : T);
}


// Capacity of the staged block, in float2 entries (20 KiB of LDS per workgroup; 40 KiB
// for chunks of 32 slices).  The blocks of 16 x 16 voxel columns x 16 slices of a 1 mm
// volume seen by a 0.5 mm (at the isocentre) detector stay below 2400 entries.
constexpr int kVoxLds = 2560;
__host__ __device__ constexpr int vox_lds(int KC) { return KC >= 32 ? 2 * kVoxLds : kVoxLds; }

// VoxTile for voxel columns [i0, i1] x [j0, j1] on view v, from the tile's corner pixels:
// the axis map of a pixel edge is monotone in the pixel and in the line (a projective
// map away from its pole), and the depth magnification zeta monotone in the line.  The
// column range carries a one-column margin over the closed-form inverse's rounding, and
// the depth-scale range is taken over every column of it, so both bound every voxel of
// the tile (not only the corners): the voxel gather skips a view on them.  nc = -1: the
// tile's footprint misses the detector (by more than a thousandth of a column).
// Computed once per plan.
template <int GEO>
inline VoxTile voxel_tile(const ViewHdr &h, const float4 *lineV, const float4 *ct, int ny, int nu, int maxline, int v,
                          int i0, int i1, int j0, int j1)
{
    VoxTile t{0, 0, 0.0f, 0.0f};
    int La, Lb, mlo, mhi;
    if (h.xdrv) { La = j0; Lb = j1; mlo = i0; mhi = i1; }
    else { La = i0; Lb = i1; mlo = ny - 1 - j1; mhi = ny - 1 - j0; }
    const float4 la = lineV[(int64_t)v * maxline + La], lb = lineV[(int64_t)v * maxline + Lb];
    const float alo = fminf(fmaf((float)mlo, la.y, la.x), fmaf((float)mlo, lb.y, lb.x));
    const float ahi = fmaxf(fmaf((float)(mhi + 1), la.y, la.x), fmaf((float)(mhi + 1), lb.y, lb.x));
    const float tol = 1e-3f * (h.dhi - h.dlo) / (float)nu;
    if (!(ahi > h.dlo - tol && alo < h.dhi + tol)) {
        t.nc = -1;
        return t;
    }
    const int c0 = imax((int)clampf(column_seed<GEO>(h, alo), 0.0f, (float)(nu - 1)) - 1, 0);
    const int c1 = imin((int)clampf(column_seed<GEO>(h, ahi), 0.0f, (float)(nu - 1)) + 1, nu - 1);
    if (c1 < c0) return t;
    float klo = ct[c0].z, khi = klo;
    for (int c = c0 + 1; c <= c1; ++c) {
        klo = fminf(klo, ct[c].z);
        khi = fmaxf(khi, ct[c].z);
    }
    const float zs[4] = {la.z * klo, la.z * khi, lb.z * klo, lb.z * khi};
    t.c0 = c0;
    t.nc = c1 - c0 + 1;
    t.zk_lo = fminf(fminf(zs[0], zs[1]), fminf(zs[2], zs[3]));
    t.zk_hi = fmaxf(fmaxf(zs[0], zs[1]), fmaxf(zs[2], zs[3]));
    return t;
}

// The block a voxel-gather workgroup stages for slices [k0, k0 + KC): the tile's columns
// and the rows rho = zeta kappa z - v0/dvv of its chunk's bottom and top boundaries over
// the tile's depth-scale range, with a margin of one row, clipped to the LDS capacity.
//
// skip is set when the tile misses the detector, or when the chunk's rows lie more than
// one row outside [0, Nv] over the tile's whole depth-scale range (a voxel lane then fails
// its own test rhi > 0 && rlo < Nv on every column: rounding moves its rows by far less
// than a row).
template <int KC>
__host__ __device__ inline VoxBox voxel_box(const VoxTile &t, const VoxArgs &a, int k0)
{
    VoxBox bx{0, 0, 0, 0, 0, 0};
    if (t.nc < 0) bx.skip = 1;
    if (t.nc <= 0) return bx;
    const float zlo = fmaf((float)k0, a.dz, a.zb), zhi = fmaf((float)(k0 + KC), a.dz, a.zb);
    float lo = fminf(zlo * t.zk_lo, zlo * t.zk_hi) - a.rv0;
    float hi = fmaxf(zhi * t.zk_lo, zhi * t.zk_hi) - a.rv0;
    if (!(hi > -1.0f && lo < (float)a.nv + 1.0f)) {
        bx.skip = 1;
        return bx;
    }
    lo = fmaxf(lo, -1.0f);
    hi = fminf(hi, (float)a.nv + 1.0f);
    const int r0 = imax((int)floorf(lo) - 1, 0);
    const int r1 = imin((int)floorf(hi) + 1, a.nv);
    const int pitch = (t.nc + 1) & ~1;
    int nr = r1 - r0 + 1;
    constexpr int cap = vox_lds(KC);
    if (nr <= 0 || pitch > cap) return bx;
    if (pitch * nr > cap) nr = cap / pitch;
    bx.c0 = t.c0;
    bx.nc = t.nc;
    bx.r0 = r0;
    bx.nr = nr;
    bx.pitch = pitch;
    return bx;
}

// Thread tid's share of copying the block from the table Qv (rows nu entries apart:
// a view's row-prefix table, or a line of a z-prefix table with nu = its plane) into
// LDS, in 16-byte pairs of adjacent columns: pair p = tid + s * nthreads goes to lds[2p, 2p+1]
// (at most capacity / 2 pairs, a multiple of nthreads for every buffer here: vox_lds(KC)
// and kCellLds, so s < capacity / 2 / nthreads).  A voxel-gather pair may end one entry
// past the view's last column (the next row's first entry, or the pad after the table),
// an entry never looked up; a cell-gather block may reach three entries past the line's
// end (the next line's first entries, or the table's kVolPad pad entries).
//
// On the device the copies are direct global-to-LDS loads (no registers held): lane l of
// the wave writes 16 bytes at the wave's LDS base + 16 l, which is where pair p lands
// for p = s * nthreads + 64 * wave + l.  Lanes past the block load the table's first
// pair into unused LDS.  The caller waits for the loads (s_waitcnt) before its barrier.
__host__ __device__ inline void voxel_stage(const float2 *__restrict__ Qv, const VoxBox &bx, float2 *__restrict__ lds,
                                            int nu, int tid, int nthreads)
{
    const int half = bx.pitch >> 1;
    const int E2 = bx.nr * half;
    if (E2 <= 0) return;
    const char *Qb = reinterpret_cast<const char *>(Qv);
    const unsigned nu8 = (unsigned)nu * 8u;
    // (row, pair) of p = tid, advanced by nthreads = dq rows + dc pairs per step.  The
    // first division: (tid + 1/2) / half stays >= 1/(2 half) >= 1/5120 from an integer,
    // far above the error of the approximate reciprocal and product (tid < 2^10:
    // below 2^10 * 2^-22).
    int r = (int)(((float)tid + 0.5f) * approx_rcp((float)half));
    int c = tid - r * half;
    const int dq = nthreads / half, dc = nthreads - dq * half;
    const unsigned cb = (unsigned)bx.c0 * 8u;
    for (int s = 0; s * nthreads < E2; ++s) {
        const int p = tid + s * nthreads;
        const unsigned off = p < E2 ? mad24((unsigned)(bx.r0 + r), nu8, cb + (unsigned)c * 16u) : 0u;
#if defined(__HIP_DEVICE_COMPILE__)
        float2 *wave_base = lds + 2 * (s * nthreads + (tid & ~(kWave - 1)));
        __builtin_amdgcn_global_load_lds(Qb + off, (__attribute__((address_space(3))) void *)wave_base, 16, 0, 0);
#else
        if (p < E2) *reinterpret_cast<Pair2 *>(lds + 2 * p) = *reinterpret_cast<const Pair2 *>(Qb + off);
#endif
        r += dq;
        c += dc;
        if (c >= half) {
            c -= half;
            ++r;
        }
    }
}

// True on every lane if the predicate holds on any active lane of the wave (device); the
// predicate itself on the host, where each thread runs alone.
__host__ __device__ inline bool wave_any(bool p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(p) != 0;
#else
    return p;
#endif
}


// Where one column's row-prefix entries are read from: the staged LDS block (when the
// column lies in it) or the view's table in HBM.  Lb is the LDS block itself (never
// replaced by a null pointer: a pointer that may be either would compile to generic
// "flat" loads for the LDS reads); in_lds says whether this column is staged.
struct VoxCol {
    const char *Qb;   // view's table
    unsigned nu8;     // byte stride of a table row
    unsigned c8;      // byte offset of the column in a table row
    const char *Lb;   // LDS block (nullptr only where no block exists at all)
    bool in_lds;      // the column lies in the staged block
    unsigned lcb;     // (c - c0) * 8 - r0 * pitch * 8: LDS byte offset of row 0 (mod 2^32)
    unsigned nc8;     // LDS row pitch in bytes
    int r0, r1;       // staged rows [r0, r1)
};

// Slice boundaries q = Q0..Q1 of one column: rows rho_q = clamp(rlo + q step, 0, Nv), one
// row-prefix lookup each, I(rho) = B_r + rho p_r; accumulates slices Q0..Q1-1 with weight
// w and returns I at boundary Q1.  With have_first the caller already holds I at Q0
// (Ifirst) and that lookup is not repeated.  All lookups are independent: the rows are
// computed, the loads issued back to back, then consumed.
template <int Q0, int Q1>
__host__ __device__ inline float voxel_slices(const VoxCol &cs, float rlo, float step, float fnv, float w,
                                              bool have_first, float Ifirst, float *acc)
{
    constexpr int N = Q1 - Q0;
    int ri[N + 1];
    float rh[N + 1];
#pragma unroll
    for (int q = 0; q <= N; ++q) {
        rh[q] = clampf(fmaf((float)(Q0 + q), step, rlo), 0.0f, fnv);
        ri[q] = (int)rh[q];
    }
    float2 e[N + 1];
    // rows increase with q, so the first and last lookups bound the span's rows
    const bool staged = cs.in_lds && ri[0] >= cs.r0 && ri[N] < cs.r1;
    if (staged) {
        // byte offset (ri - r0) * pitch * 8 + (c - c0) * 8 as one 24-bit multiply-add
        // (unsigned wrap-around of the constant part is exact)
        if (!have_first) e[0] = lds_float2(cs.Lb, mad24((unsigned)ri[0], cs.nc8, cs.lcb));
#pragma unroll
        for (int q = 1; q <= N; ++q) e[q] = lds_float2(cs.Lb, mad24((unsigned)ri[q], cs.nc8, cs.lcb));
    } else {
        if (!have_first) e[0] = global_float2(cs.Qb, mul24((unsigned)ri[0], cs.nu8) + cs.c8);
#pragma unroll
        for (int q = 1; q <= N; ++q) e[q] = global_float2(cs.Qb, mul24((unsigned)ri[q], cs.nu8) + cs.c8);
    }
    // I(rho) = B_r + rho q_r (see prefix_column)
    float Ip = have_first ? Ifirst : fmaf(rh[0], e[0].y, e[0].x);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const float In = fmaf(rh[q + 1], e[q + 1].y, e[q + 1].x);
        acc[Q0 + q] = fmaf(w, In - Ip, acc[Q0 + q]);
        Ip = In;
    }
    return Ip;
}

// One view's contribution to the KC slice accumulators of voxel column (i, j).  With a
// staged block (lds != nullptr), lookups inside it read LDS, the rest the table in HBM.
//
// With HALVES, chunks of 16 slices run as two halves of 8, each skipped when no lane of
// the wave has a row of that half on the detector: every boundary of a skipped half then
// clamps to the same row (0 below, Nv above) for every lane, so its differences are
// exactly zero and skipping it changes no result.
template <int OPK, int GEO, int KC, bool HALVES = false>
__host__ __device__ inline void voxel_view(const ViewHdr &h, const float4 *__restrict__ lineV,
                                           const float4 *__restrict__ ct, const float2 *__restrict__ Qv,
                                           const VoxArgs &a, int v, int i, int j, int k0, float px, float py,
                                           float *acc, const VoxBox &bx, const float2 *__restrict__ lds)
{
    static_assert(KC <= 16 || HALVES, "chunks beyond 16 slices run in halves");
    const float fnv = (float)a.nv;
    const int nu = a.nu;
    int L, m;
    if (h.xdrv) { L = j; m = i; } else { L = i; m = a.ny - 1 - j; }
    const float4 lv = g_load(lineV + (int64_t)v * a.maxline + L);  // {P0, Delta, zeta, 1/(Delta zeta)}
    const float a0 = fmaf((float)m, lv.y, lv.x);
    const float a1 = a0 + lv.y;
    if (!(a1 > h.dlo && a0 < h.dhi)) return;

    float hB = 1.0f;
    if (OPK != VOX_AT) {
        // U * w / A_vox with A_vox = Delta * zeta * dz (include/cpu3d_common.h:80-85),
        // times the row pitch dvv of the depth overlap measured in row units.
        const float U = (OPK == VOX_BW) ? fdk_weight<GEO>(h.sx, h.sy, px, py, a.sod) : 1.0f;
        hB = U * h.colw * a.hscale * lv.w;
    }

    // A column at or before the first one whose right boundary exceeds a0: closed-form
    // seed, corrected downwards against the table.  Columns between it and the footprint
    // have no overlap and are skipped by the walk below, so the result is exact w.r.t. the
    // FP32 boundaries whatever the seed's rounding.
    int c = (int)clampf(column_seed<GEO>(h, a0), 0.0f, (float)(nu - 1));
    float4 cc = col_load3(ct + c), cn = col_load3(ct + c + 1);
    while (c > 0 && cc.x > a0) { --c; cn = cc; cc = col_load3(ct + c); }

    // Slice boundaries are measured from the volume centre, (k - Nz/2) dz + zc: the rows
    // of the slices that reach the detector then come out of small products instead of
    // the difference of two large ones (from the volume bottom, a tall volume lost about
    // Nz/2 x 2^-24 rows to cancellation).
    const float zcz = lv.z * a.zc, zsz = lv.z * a.dz;
    // Byte offsets below 2^32 within one view's table ((Nv+1)*Nu*8, checked at plan
    // creation) over a wave-uniform base.
    const char *Qb = reinterpret_cast<const char *>(Qv);
    const unsigned nu8 = (unsigned)nu * 8u;
    const unsigned nc8 = (unsigned)bx.pitch * 8u;
    const char *ctb = reinterpret_cast<const char *>(ct);
    while (c < nu && cc.x < a1) {
        // prefetch the next column's table entry one iteration ahead (32-bit byte offset
        // over the view's table: (Nu+1) * 16 bytes)
        const float4 cnn = col_load3(reinterpret_cast<const float4 *>(ctb + (unsigned)imin(c + 2, nu) * 16u));
        const float ou = minf_finite(a1, cn.x) - maxf_finite(a0, cc.x);
        if (ou > 0.0f) {
            // Row coordinate of slice boundary k: rho_k = (zeta z_k / sigma - v0) / dvv.
            const float kap = cc.z;
            const float step = zsz * kap;
            const float base = fmaf(zcz, kap, -a.rv0);
            const float rlo = fmaf((float)k0 - a.kmid, step, base);
            const float rhi = fmaf((float)(k0 + KC) - a.kmid, step, base);
            if (rhi > 0.0f && rlo < fnv) {
                // B: ou * U w sigma dvv / (Delta zeta dz); A^T: ou (normalisation folded
                // into the prefix table).
                const float w = (OPK == VOX_AT) ? ou : ou * hB * cc.y;
                const bool in_lds = lds != nullptr && (unsigned)(c - bx.c0) < (unsigned)bx.nc;
                const VoxCol cs{Qb, nu8, (unsigned)c * 8u, reinterpret_cast<const char *>(lds), in_lds,
                                (unsigned)(c - bx.c0) * 8u - (unsigned)bx.r0 * nc8, nc8, bx.r0, bx.r0 + bx.nr};
                if (!HALVES || KC < 16) {
                    voxel_slices<0, KC>(cs, rlo, step, fnv, w, false, 0.0f, acc);
                } else {
                    constexpr int H = KC / 2;
                    const float rmid = fmaf((float)H, step, rlo);
                    // this lane's lower half reaches row 0 iff rmid > 0 (its chunk starts
                    // below Nv), the upper half iff rmid < Nv (its chunk ends above 0)
                    const bool lo = wave_any(rmid > 0.0f), hi = wave_any(rmid < fnv);
                    float Imid = 0.0f;
                    if (lo) Imid = voxel_slices<0, H>(cs, rlo, step, fnv, w, false, 0.0f, acc);
                    if (hi) voxel_slices<H, KC>(cs, rlo, step, fnv, w, lo, Imid, acc);
                }
            }
        }
        ++c;
        cc = cn;
        cn = cnn;
    }
}

__host__ __device__ inline float voxel_px(const VoxArgs &a, int i) { return ((float)i - 0.5f * a.nx + 0.5f) * a.dx + a.shx; }
__host__ __device__ inline float voxel_py(const VoxArgs &a, int j) { return (0.5f * a.ny - (float)j - 0.5f) * a.dy + a.shy; }

__host__ __device__ inline const float2 *view_table(const float2 *Q, const VoxArgs &a, int b, int v)
{
    return Q + ((int64_t)b * a.nview + v) * (int64_t)(a.nv + 1) * a.nu;
}

template <int KC>
__host__ __device__ inline void voxel_store(float *__restrict__ out, const VoxArgs &a, int i, int j, int k0, int b,
                                            const float *acc)
{
    float *o = out + (int64_t)b * a.nz * a.ny * a.nx + (int64_t)j * a.nx + i;
    for (int q = 0; q < KC; ++q) {
        const int k = k0 + q;
        if (k >= a.nz) break;
        o[(int64_t)k * a.ny * a.nx] = acc[q];
    }
}

// One thread, reading the tables in HBM only: voxel column (i, j), slices k0..k0+KC-1,
// batch b, views [v0, v1).
template <int OPK, int GEO, int KC, bool HALVES = false>
__host__ __device__ inline void voxel_body(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineV,
                                           const float4 *__restrict__ colt, const float2 *__restrict__ Q,
                                           float *__restrict__ out, const VoxArgs &a, int i, int j, int k0, int b,
                                           int v0, int v1)
{
    const float px = voxel_px(a, i), py = voxel_py(a, j);
    float acc[KC];
    for (int q = 0; q < KC; ++q) acc[q] = 0.0f;
    const VoxBox none{0, 0, 0, 0, 0, 0};
    for (int v = v0; v < v1; ++v)
        voxel_view<OPK, GEO, KC, HALVES>(hdr[v], lineV, colt + (int64_t)v * (a.nu + 1), view_table(Q, a, b, v), a, v,
                                         i, j, k0, px, py, acc, none, nullptr);
    voxel_store<KC>(out, a, i, j, k0, b, acc);
}

// A workgroup of the voxel gather (device only): the 16 x 16 tile tile_id of voxel
// columns, slices k0..k0+KC-1, batch b, views [v0, v1); this thread owns (i, j) when
// `active`.  Per view the
// workgroup stages the block of the row-prefix table its footprint reads into LDS, then
// every thread gathers from it.  Inactive threads take part in staging and barriers.
// (A second, alternating buffer that overlaps the next view's staging with this view's
// gather measured slower: the gather's own table loads must wait for the earlier-issued
// staging loads anyway, and the doubled code and registers cost more than they save.)
// Chunks of fewer than kVoxStageMin slices read the table directly: a staged block then
// serves too few lookups to repay its copy and the two barriers per view.
constexpr int kVoxStageMin = 8;

template <int OPK, int GEO, int KC, bool HALVES = false>
__device__ inline void voxel_wg(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineV,
                                const float4 *__restrict__ colt, const VoxTile *__restrict__ vtile,
                                const float2 *__restrict__ Q, float *__restrict__ out, const VoxArgs &a, int i, int j,
                                bool active, int tile_id, int k0, int b, int v0, int v1,
                                float2 *__restrict__ tile)
{
    if (KC < kVoxStageMin) {
        if (active) voxel_body<OPK, GEO, KC, HALVES>(hdr, lineV, colt, Q, out, a, i, j, k0, b, v0, v1);
        return;
    }
    const float px = voxel_px(a, i), py = voxel_py(a, j);
    float acc[KC];
    for (int q = 0; q < KC; ++q) acc[q] = 0.0f;
    const int ntile = ceil_div(a.nx, kVoxTile) * ceil_div(a.ny, kVoxTile);
    for (int v = v0; v < v1; ++v) {
        const ViewHdr h = ro_load(hdr + v);
        const float4 *ct = colt + (int64_t)v * (a.nu + 1);
        const float2 *Qv = view_table(Q, a, b, v);
        const VoxBox bx = voxel_box<KC>(ro_load(vtile + (int64_t)v * ntile + tile_id), a, k0);
        if (bx.skip) continue;  // the same for every thread of the workgroup: no barrier is missed
        voxel_stage(Qv, bx, tile, a.nu, threadIdx.x, kBlock);
        __builtin_amdgcn_s_waitcnt(0);  // this wave's direct-to-LDS loads have landed
        __syncthreads();
        if (active) voxel_view<OPK, GEO, KC, HALVES>(h, lineV, ct, Qv, a, v, i, j, k0, px, py, acc, bx, tile);
        __syncthreads();
    }
    if (active) voxel_store<KC>(out, a, i, j, k0, b, acc);
}

// Waits until at most n (0..4) of this wave's vector memory loads are outstanding (they
// complete in order); n above 4 waits as for 4 (more than needed, never less).
__device__ inline void wait_vmem(int n)
{
    // s_waitcnt immediate: vmcnt in bits 3:0, expcnt (7: no wait) 6:4, lgkmcnt (15) 11:8
    switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 1: __builtin_amdgcn_s_waitcnt(0xF71); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF72); break;
    case 3: __builtin_amdgcn_s_waitcnt(0xF73); break;
    default: __builtin_amdgcn_s_waitcnt(0xF74); break;
    }
}

// Workgroup barrier that leaves this wave's outstanding global loads in flight (a
// __syncthreads() fence waits for all of them).  The caller has waited for the loads
// whose LDS data it shares across the barrier (wait_vmem).  This wave's LDS reads have
// completed (lgkmcnt 0) before other waves may overwrite what they read, and the empty
// asm statements keep the compiler from moving memory accesses across the barrier.
__device__ inline void wg_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0) only
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Cell gather of one workgroup with its z-prefix lookups served from LDS.  The 4 waves
// take 4 groups of span class SPAN (2 or 3) on rows r0..r0+R-1 (see k_cell); ql: the
// workgroup's QuadLine per line.  Line L's block (cell_box) is staged into slot L mod 3
// two lines ahead of its gather: at line L each wave waits for its own copies of L (the
// copies of L+1, issued later, may stay in flight), the workgroup meets at a barrier
// (everyone's copies of L have landed; everyone is done with line L-1, whose slot L+2
// then takes), and gathers L.  The plan stages a workgroup only where its 4 views share
// one orientation (so the line count and the table are the same for the 4 waves) and
// every block of its row tile fits a slot.  Every thread of the workgroup takes part in
// staging and barriers; active: the thread has a column (its wave a group, its lane a
// column below Nu).  The lines go three at a time so that each slot is its own array at
// every access.
//
// Per 64-line chunk each wave evaluates the chunk's blocks over its lanes (line base + l
// on lane l, read back per line with readlane) at the chunk's pitch, and each thread the
// places of its copies in a block of that pitch: pair p = tid + 256 s at row p / (pitch /
// 2), pair p mod (pitch / 2), as voxel_stage lays them out.
template <int OPK, int GEO, int R, int SPAN>
__device__ inline void cell_wg(const ViewHdr *__restrict__ hdr, const float4 *__restrict__ lineC,
                               const float4 *__restrict__ colt, const QuadLine *__restrict__ ql,
                               const float2 *__restrict__ WX, const float2 *__restrict__ WY, float *__restrict__ out,
                               const CellArgs &a, int c, int r0, int v, int b, float wgrp, bool active,
                               float2 *__restrict__ s0, float2 *__restrict__ s1, float2 *__restrict__ s2)
{
    static_assert(kCellSlots == 3 && kCellChunk == kWave, "three slots, lines staged two ahead");
    constexpr int kRounds = kCellLds / 2 / kBlock;  // copy rounds of a full slot
    const ViewHdr h = ro_load(hdr + v);
    CellLane<R> st;
    cell_init<R>(st, colt + (int64_t)v * (a.nu + 1), a, active ? c : 0, r0);
    const float2 *W = cell_table(WX, WY, h, a, b);
    const float4 *lt = lineC + (int64_t)v * a.maxline;
    const int nline = wave_uniform(h.nline), nalong = wave_uniform(h.nalong);
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const VoxBox none{0, 0, 0, 0, 0, 0};
    int kc0 = 0, kr0 = 0, knr = 0, kpitch = 0;  // lane l: line base + l's block; the chunk's pitch
    unsigned cplace[kRounds];  // this thread's copies: row << 16 | pair offset in bytes (< 2^14 each)
    auto chunk = [&](int base) {
        const QuadLine q = g_load(ql + imin(base + lane, nline - 1));
        int pm = base + lane < nline ? cell_line_pitch(q) : 0;
#pragma unroll
        for (int m = kWave / 2; m > 0; m >>= 1) pm = imax(pm, __shfl_xor(pm, m));
        kpitch = wave_uniform(pm);
        const VoxBox bx = cell_box<R>(q, a, r0, kpitch, nalong);
        kc0 = bx.c0;
        kr0 = bx.r0;
        knr = bx.nr;
        const int half = imax(kpitch >> 1, 1);
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            const int p = tid + k * kBlock, r = p / half;
            cplace[k] = (unsigned)r << 16 | (unsigned)(p - r * half) * 16u;
        }
    };
    auto box_of = [&](int j) {
        const int l = j & (kWave - 1);
        VoxBox bx;
        bx.c0 = __builtin_amdgcn_readlane(kc0, l);
        bx.r0 = __builtin_amdgcn_readlane(kr0, l);
        bx.nr = __builtin_amdgcn_readlane(knr, l);
        bx.pitch = bx.nr > 0 ? kpitch : 0;
        bx.nc = bx.pitch;
        bx.skip = 0;
        return bx;
    };
    // the copies of line L's block: as voxel_stage, with this thread's places from chunk()
    auto stage = [&](int L, const VoxBox &bx, float2 *slot) {
        const int E2 = bx.nr * (bx.pitch >> 1);
        const char *Wl = reinterpret_cast<const char *>(W + (int64_t)L * nalong);
        const unsigned base = (unsigned)bx.r0 * a.plane8 + (unsigned)bx.c0 * 8u;
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            if (k * kBlock >= E2) break;
            const unsigned off =
                tid + k * kBlock < E2 ? mad24(cplace[k] >> 16, a.plane8, base + (cplace[k] & 0xFFFFu)) : 0u;
            float2 *wave_base = slot + 2 * (k * kBlock + (tid & ~(kWave - 1)));
            __builtin_amdgcn_global_load_lds(Wl + off, (__attribute__((address_space(3))) void *)wave_base, 16, 0, 0);
        }
    };
    chunk(0);
    VoxBox b0 = box_of(0);
    VoxBox b1 = nline > 1 ? box_of(1) : none;
    VoxBox b2 = none;
    stage(0, b0, s0);
    stage(1, b1, s1);
    // The line coefficients of L, read one step ahead: a step issues the next step's read
    // after its barrier, and it lands during its gather.
    float4 lc = ro_load(lt);
    // line L from slot rd (block cur); n1: the block of L+1; n2 (slot wr) is set to L+2's
    auto step = [&](int L, const VoxBox &cur, const VoxBox &n1, VoxBox &n2, const float2 *rd, float2 *wr) {
        wait_vmem(stage_rounds(n1));
        wg_barrier();
        const float4 lc_next = ro_load(lt + imin(L + 1, nline - 1));
        if (L + 2 < nline) {
            if (((L + 2) & (kCellChunk - 1)) == 0) chunk(L + 2);
            n2 = box_of(L + 2);
        } else {
            n2 = none;
        }
        stage(L + 2, n2, wr);
        __builtin_amdgcn_sched_barrier(0);  // issue the copies before the gather
        if (active)
            cell_line_staged<OPK, GEO, R, SPAN>(st, h, a, lc, L, W, wgrp, reinterpret_cast<const char *>(rd), cur);
        lc = lc_next;
    };
    for (int L = 0; L < nline; L += 3) {
        step(L, b0, b1, b2, s0, s2);
        if (L + 1 >= nline) break;
        step(L + 1, b1, b2, b0, s1, s0);
        if (L + 2 >= nline) break;
        step(L + 2, b2, b0, b1, s2, s1);
    }
    if (active) cell_store<OPK, GEO, R>(st, h, a, out, c, r0, v, b);
}

}  // namespace xr
