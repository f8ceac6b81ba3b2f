// GPU emission of the system-matrix triplets: the same triplets as the host builder
// (sysmat.cpp), in the same order, bit for bit.
//
// Both run the traversals of sysmat_walk.h over the same per-view tables (built on the
// host, where every transcendental of the traversal is evaluated); on the GPU one thread
// walks one traversal item.  Three passes:
//   1. count:  each item counts the coefficients it keeps (non-zero and finite);
//   2. scan:   per-256-item block sums, then one workgroup turns them into exclusive block
//              offsets in item order (and the total, nnz);
//   3. emit:   each workgroup scans its 256 counts in LDS and every item walks again,
//              writing its triplets from its offset on.
// Item order is the host builder's loop order, so the concatenated output is the host's
// triplet list.  The count pass and the walks of the emit pass evaluate identical FP32
// expressions, so they agree on which coefficients are kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "sysmat_internal.h"
#include "xr_internal.h"

namespace xr {
namespace sm {
namespace {

constexpr int kSmBlock = 256;

enum SmKind : int { SM_CONE = 0, SM_PAR3 = 1, SM_2D = 2 };

// Everything a thread needs to decode its item and walk it.
struct SmArgs {
    xr_volume3d vol;
    xr_geometry3d geo;
    xr_volume2d vol2;
    xr_geometry2d geo2;
    int nview, maxline, nblk;
    long long nitems;
    const SmView *views;
    const SmView2 *views2;
    const float *cax;     // cone: [V][Nu + 1]; 2D fan: dax [V][Nu + 1] (after the pad)
    const float *center;  // cone: [V][Nu]
    const float *delta;   // arc: [V][maxline]
};

// Walks item `it` and calls out(row, col, value) for every kept coefficient, in order.
template <int KIND, class Out>
__device__ inline void walk_item(const SmArgs &a, long long it, Out &out)
{
    if (KIND == SM_PAR3) {
        const xr_volume3d &vol = a.vol;
        const xr_geometry3d &g = a.geo;
        const int ix = (int)(it % vol.nx);
        long long rest = it / vol.nx;
        const int iy = (int)(rest % vol.ny);
        rest /= vol.ny;
        const int iz = (int)(rest % vol.nz);
        const int v = (int)(rest / vol.nz);
        const long long icol = ((long long)iz * vol.ny + iy) * vol.nx + ix;
        auto emit = [&](int row, int col, float c) {
            if (keep_value(c)) out(((long long)v * g.nv + row) * g.nu + col, icol, c);
        };
        parallel_voxel(vol, g, a.views[v].sinv, a.views[v].cosv, ix, iy, iz, emit);
    } else if (KIND == SM_CONE) {
        const xr_volume3d &vol = a.vol;
        const xr_geometry3d &g = a.geo;
        const int L = (int)(it % a.maxline);
        long long rest = it / a.maxline;
        const int bk = (int)(rest % a.nblk);
        rest /= a.nblk;
        const int r = (int)(rest % g.nv);
        const int v = (int)(rest / g.nv);
        const SmView vw = a.views[v];
        if (L >= (vw.xdrv ? vol.ny : vol.nx)) return;
        const int c0 = bk * kBlockCols;
        const int nc = smin(kBlockCols, g.nu - c0);
        const float delta = g.type == XR_ARC ? a.delta[(long long)v * a.maxline + L] : 0.0f;
        auto emit = [&](int lc, int ix, int iy, int iz, float c) {
            if (!keep_value(c)) return;
            const int col = c0 + lc;
            const int o = vw.positive ? col : (g.nu - 1 - col);
            out(((long long)v * g.nv + r) * g.nu + o, ((long long)iz * vol.ny + iy) * vol.nx + ix, c);
        };
        cone_line(vol, g, vw, a.cax + (long long)v * (g.nu + 1), a.center + (long long)v * g.nu, delta, r, c0, nc, L,
                  emit);
    } else {
        const xr_volume2d &vol = a.vol2;
        const xr_geometry2d &g = a.geo2;
        const int L = (int)(it % a.maxline);
        const int v = (int)(it / a.maxline);
        const SmView2 w = a.views2[v];
        if (L >= (w.xdrv ? vol.ny : vol.nx)) return;
        const float delta = g.type == XR_ARC ? a.delta[(long long)v * a.maxline + L] : 0.0f;
        auto emit = [&](int det, int row, int col, float c) {
            if (keep_value(c)) out((long long)v * g.nu + det, (long long)row * vol.nx + col, c);
        };
        line2d(vol, g, w, a.cax + (long long)v * (g.nu + 1), delta, L, emit);
    }
}

template <int KIND>
__global__ __launch_bounds__(kSmBlock) void k_sm_count(SmArgs a, int *__restrict__ counts)
{
    const long long it = (long long)blockIdx.x * kSmBlock + threadIdx.x;
    if (it >= a.nitems) return;
    int n = 0;
    auto out = [&](long long, long long, float) { ++n; };
    walk_item<KIND>(a, it, out);
    counts[it] = n;
}

// Inclusive scan of one value per thread over the workgroup (LDS, Hillis-Steele).
__device__ inline long long block_scan_inclusive(long long x, long long *lds)
{
    const int t = threadIdx.x;
    lds[t] = x;
    __syncthreads();
    for (int d = 1; d < kSmBlock; d <<= 1) {
        const long long y = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += y;
        __syncthreads();
    }
    const long long r = lds[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kSmBlock) void k_sm_block_sums(const int *__restrict__ counts, long long nitems,
                                                            long long *__restrict__ bsum)
{
    __shared__ long long lds[kSmBlock];
    const long long it = (long long)blockIdx.x * kSmBlock + threadIdx.x;
    const long long s = block_scan_inclusive(it < nitems ? counts[it] : 0, lds);
    if (threadIdx.x == kSmBlock - 1) bsum[blockIdx.x] = s;
}

// One workgroup: bsum[0..nb) -> exclusive prefix, total -> *total.
__global__ __launch_bounds__(kSmBlock) void k_sm_scan_sums(long long *__restrict__ bsum, long long nb,
                                                           long long *__restrict__ total)
{
    __shared__ long long lds[kSmBlock];
    long long carry = 0;
    for (long long base = 0; base < nb; base += kSmBlock) {
        const long long i = base + threadIdx.x;
        const long long x = i < nb ? bsum[i] : 0;
        const long long inc = block_scan_inclusive(x, lds);
        if (i < nb) bsum[i] = carry + inc - x;
        // the last thread's inclusive value is the chunk's sum; every thread reads it
        if (threadIdx.x == kSmBlock - 1) lds[0] = inc;
        __syncthreads();
        carry += lds[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <int KIND>
__global__ __launch_bounds__(kSmBlock) void k_sm_emit(SmArgs a, const int *__restrict__ counts,
                                                      const long long *__restrict__ boff,
                                                      long long *__restrict__ rows, long long *__restrict__ cols,
                                                      float *__restrict__ vals)
{
    __shared__ long long lds[kSmBlock];
    const long long it = (long long)blockIdx.x * kSmBlock + threadIdx.x;
    const long long n = it < a.nitems ? counts[it] : 0;
    long long pos = boff[blockIdx.x] + block_scan_inclusive(n, lds) - n;
    if (it >= a.nitems || n == 0) return;
    auto out = [&](long long r, long long c, float v) {
        rows[pos] = r;
        cols[pos] = c;
        vals[pos] = v;
        ++pos;
    };
    walk_item<KIND>(a, it, out);
}

template <typename T>
int upload(T **dst, const std::vector<T> &src)
{
    *dst = nullptr;
    if (src.empty()) return XR_OK;
    if (hipMalloc(reinterpret_cast<void **>(dst), src.size() * sizeof(T)) != hipSuccess)
        return fail(XR_ENOMEM, "hipMalloc failed for system-matrix tables");
    if (hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return fail(XR_EHIP, "hipMemcpy failed for system-matrix tables");
    return XR_OK;
}

struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceScope()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace
}  // namespace sm
}  // namespace xr

struct xr_sysmat_gpu {
    int kind = 0, device = 0;
    xr::sm::SmArgs args{};
    xr::sm::SmView *d_views = nullptr;
    xr::sm::SmView2 *d_views2 = nullptr;
    float *d_cax = nullptr, *d_center = nullptr, *d_delta = nullptr;
    int *d_counts = nullptr;
    long long *d_boff = nullptr, *d_total = nullptr;
    long long nblocks = 0;
    int64_t nnz = 0;

    ~xr_sysmat_gpu()
    {
        xr::sm::DeviceScope g(device);
        for (void *p : {(void *)d_views, (void *)d_views2, (void *)d_cax, (void *)d_center, (void *)d_delta,
                        (void *)d_counts, (void *)d_boff, (void *)d_total})
            if (p) (void)hipFree(p);
    }
};

namespace {

using xr::fail;

// Count pass, block sums and their scan on `s`, then the total to the host.
int count_and_scan(xr_sysmat_gpu &m, hipStream_t s)
{
    using namespace xr::sm;
    const long long n = m.args.nitems;
    m.nblocks = (n + kSmBlock - 1) / kSmBlock;
    if (m.nblocks >= ((long long)1 << 31)) return fail(XR_EINVAL, "system matrix traversal too large for one launch");
    if (hipMalloc(reinterpret_cast<void **>(&m.d_counts), std::max<long long>(n, 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&m.d_boff), std::max<long long>(m.nblocks, 1) * sizeof(long long)) !=
            hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&m.d_total), sizeof(long long)) != hipSuccess)
        return fail(XR_ENOMEM, "hipMalloc failed for the system-matrix counts");
    long long total = 0;
    if (n > 0) {
        const dim3 grid((unsigned)m.nblocks), block(kSmBlock);
        if (m.kind == SM_CONE) hipLaunchKernelGGL(k_sm_count<SM_CONE>, grid, block, 0, s, m.args, m.d_counts);
        else if (m.kind == SM_PAR3) hipLaunchKernelGGL(k_sm_count<SM_PAR3>, grid, block, 0, s, m.args, m.d_counts);
        else hipLaunchKernelGGL(k_sm_count<SM_2D>, grid, block, 0, s, m.args, m.d_counts);
        hipLaunchKernelGGL(k_sm_block_sums, grid, block, 0, s, m.d_counts, n, m.d_boff);
        hipLaunchKernelGGL(k_sm_scan_sums, dim3(1), block, 0, s, m.d_boff, m.nblocks, m.d_total);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&total, m.d_total, sizeof(long long), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail(XR_EHIP, std::string("system-matrix count pass failed: ") + hipGetErrorString(e));
    }
    m.nnz = total;
    return XR_OK;
}

}  // namespace

extern "C" {

int xr_system_matrix3d_gpu(const xr_volume3d *vol, const xr_geometry3d *geo, const float *angles_host, int num_view,
                           int device, void *stream, xr_sysmat_gpu **out)
{
    using namespace xr::sm;
    xr::set_error("");
    if (!vol || !geo || !out || (num_view > 0 && !angles_host)) return fail(XR_EINVAL, "null argument");
    *out = nullptr;
    if (vol->nx <= 0 || vol->ny <= 0 || vol->nz <= 0 || geo->nu <= 0 || geo->nv <= 0 || num_view < 0)
        return fail(XR_EINVAL, "sizes must be positive");
    if (geo->type != XR_PARALLEL && geo->type != XR_FLAT && geo->type != XR_ARC)
        return fail(XR_EINVAL, "scan type must be parallel (0), flat (1) or arc (2)");
    if ((int64_t)num_view * geo->nv * geo->nu >= ((int64_t)1 << 31) ||
        (int64_t)vol->nx * vol->ny * vol->nz >= ((int64_t)1 << 31))
        return fail(XR_EINVAL, "system matrix too large for 32-bit indices");
    std::unique_ptr<xr_sysmat_gpu> m(new xr_sysmat_gpu());
    m->device = device;
    DeviceScope g(device);
    Tables3d t;
    build_tables3d(*vol, *geo, angles_host, num_view, t);
    SmArgs &a = m->args;
    a.vol = *vol;
    a.geo = *geo;
    a.nview = num_view;
    a.maxline = t.maxline;
    a.nblk = (geo->nu - 1) / kBlockCols + 1;
    m->kind = geo->type == XR_PARALLEL ? SM_PAR3 : SM_CONE;
    a.nitems = m->kind == SM_PAR3 ? (long long)num_view * vol->nz * vol->ny * vol->nx
                                  : (long long)num_view * geo->nv * a.nblk * a.maxline;
    int rc;
    if ((rc = upload(&m->d_views, t.views)) || (rc = upload(&m->d_cax, t.cax)) ||
        (rc = upload(&m->d_center, t.center)) || (rc = upload(&m->d_delta, t.delta)))
        return rc;
    a.views = m->d_views;
    a.cax = m->d_cax ? m->d_cax + kTablePad : nullptr;
    a.center = m->d_center ? m->d_center + kTablePad : nullptr;
    a.delta = m->d_delta;
    if ((rc = count_and_scan(*m, static_cast<hipStream_t>(stream)))) return rc;
    *out = m.release();
    return XR_OK;
}

int xr_system_matrix2d_gpu(const xr_volume2d *vol, const xr_geometry2d *geo, const float *angles_host, int num_view,
                           int device, void *stream, xr_sysmat_gpu **out)
{
    using namespace xr::sm;
    xr::set_error("");
    if (!vol || !geo || !out || (num_view > 0 && !angles_host)) return fail(XR_EINVAL, "null argument");
    *out = nullptr;
    if (vol->nx <= 0 || vol->ny <= 0 || geo->nu <= 0 || num_view < 0) return fail(XR_EINVAL, "sizes must be positive");
    if (geo->type != XR_PARALLEL && geo->type != XR_FLAT && geo->type != XR_ARC)
        return fail(XR_EINVAL, "scan type must be parallel (0), flat (1) or arc (2)");
    std::unique_ptr<xr_sysmat_gpu> m(new xr_sysmat_gpu());
    m->device = device;
    DeviceScope g(device);
    Tables2d t;
    build_tables2d(*vol, *geo, angles_host, num_view, t);
    SmArgs &a = m->args;
    a.vol2 = *vol;
    a.geo2 = *geo;
    a.nview = num_view;
    a.maxline = t.maxline;
    m->kind = SM_2D;
    a.nitems = (long long)num_view * a.maxline;
    int rc;
    if ((rc = upload(&m->d_views2, t.views)) || (rc = upload(&m->d_cax, t.dax)) || (rc = upload(&m->d_delta, t.delta)))
        return rc;
    a.views2 = m->d_views2;
    a.cax = m->d_cax ? m->d_cax + kTablePad : nullptr;
    a.delta = m->d_delta;
    if ((rc = count_and_scan(*m, static_cast<hipStream_t>(stream)))) return rc;
    *out = m.release();
    return XR_OK;
}

int64_t xr_sysmat_gpu_nnz(const xr_sysmat_gpu *m) { return m ? m->nnz : XR_EINVAL; }

int xr_sysmat_gpu_emit(const xr_sysmat_gpu *m, int64_t *rows, int64_t *cols, float *vals, void *stream)
{
    using namespace xr::sm;
    if (!m) return fail(XR_EINVAL, "null system matrix");
    if (m->nnz == 0) return XR_OK;
    if (!rows || !cols || !vals) return fail(XR_EINVAL, "null output buffer");
    DeviceScope g(m->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)m->nblocks), block(kSmBlock);
    auto *r = reinterpret_cast<long long *>(rows);
    auto *c = reinterpret_cast<long long *>(cols);
    if (m->kind == SM_CONE) hipLaunchKernelGGL(k_sm_emit<SM_CONE>, grid, block, 0, s, m->args, m->d_counts, m->d_boff, r, c, vals);
    else if (m->kind == SM_PAR3) hipLaunchKernelGGL(k_sm_emit<SM_PAR3>, grid, block, 0, s, m->args, m->d_counts, m->d_boff, r, c, vals);
    else hipLaunchKernelGGL(k_sm_emit<SM_2D>, grid, block, 0, s, m->args, m->d_counts, m->d_boff, r, c, vals);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(XR_EHIP, std::string("system-matrix emission failed: ") + hipGetErrorString(e));
    return XR_OK;
}

void xr_sysmat_gpu_destroy(xr_sysmat_gpu *m) { delete m; }

}  // extern "C"
