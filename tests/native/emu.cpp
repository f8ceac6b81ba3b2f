// Test-only CPU emulator of the product kernels.
//
// Runs the exact per-thread bodies of the gfx950 kernels (xrecon_amd/csrc/dd_body.h),
// with the product's own host plan builder (plan.cpp) and operator routing, one
// "thread" at a time on the CPU.  The CPU test suite uses it to check the kernel logic
// against the oracle without a GPU; the GPU suite checks the compiled kernels.  It is
// not part of the product library and nothing in xrecon_amd loads it.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <limits>
#include <thread>
#include <vector>

#include "../../xrecon_amd/csrc/dd_body.h"

using namespace xr;

namespace {

// Cell-gather variant (see run_cell_RV): 0 = each thread from the table, 1 = general
// 32-bit, 2 = general 64-bit, 3 = the staged workgroups as launched (run_cell_staged).
int g_cell_variant = 0;

// Staged cell gather: (thread, line) pairs served from an LDS block; workgroups
// staged and not staged (the plan's flags).
std::atomic<long long> g_cell_lds_lines{0}, g_cell_wg_staged{0}, g_cell_wg_direct{0};
// Staged blocks whose copies would start before their line or end more than three
// entries past it.
std::atomic<long long> g_cell_block_out{0};

// Number of (workgroup, view) pairs the voxel gather passed over (VoxBox::skip).
std::atomic<long long> g_voxel_skips{0};

// MODE 0: what the library launches (per column group: the group's class, each line of
// a wider group on the narrowest path it allows; big plans 64-bit general); 1: general
// loop, 32-bit offsets, on every line; 2: general loop, 64-bit offsets.
template <int OPK, int GEO, int R, int MODE>
void run_cell_RV(const Plan3d &p, const float2 *WX, const float2 *WY, float *out, int nb, int nthreads)
{
    const CellArgs a = make_cell_args(p, nb);
    const int ngroups = (a.nv + R - 1) / R;
    const int ncg = ceil_div(a.nu, kWave);
    const int items = nb * p.nview;
    const float inf = std::numeric_limits<float>::infinity();
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&, t]() {
            for (int it = t; it < items; it += nthreads) {
                const int v = it % p.nview, b = it / p.nview;
                for (int g = 0; g < ngroups; ++g)
                    for (int cg = 0; cg < ncg; ++cg) {
                        const CellGroup grp = p.d_cgrp[(int64_t)v * ncg + cg];
                        const int span = MODE == 0 && !p.big ? grp.span : 0;
                        const float wg = MODE == 0 && !p.big ? grp.wmax : inf;
                        for (int c = cg * kWave; c < std::min(a.nu, (cg + 1) * kWave); ++c) {
                            if (span == 2)
                                cell_body<OPK, GEO, R, 2, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, g * R,
                                                                 v, b, wg);
                            else if (span == 3)
                                cell_body<OPK, GEO, R, 3, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, g * R,
                                                                 v, b, wg);
                            else if (MODE == 2 || p.big)
                                cell_body<OPK, GEO, R, 0, true>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, g * R,
                                                                v, b, inf);
                            else
                                cell_body<OPK, GEO, R, 0, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, g * R,
                                                                 v, b, wg);
                        }
                    }
            }
        });
    }
    for (auto &x : th) x.join();
}

// One staged workgroup of k_cell (class SPAN 2 or 3, quad q of the class's list `items`
// of n entries, row tile r0, batch item b), line by line as cell_wg runs it: the line's
// block (at its 64-line chunk's pitch) copied into a one-slot "LDS" by every thread's
// share of voxel_stage, the rest of the slot NaN (a lookup outside the staged entries
// would show in the output), then each thread's line.
template <int OPK, int GEO, int R, int SPAN>
void run_cell_quad(const Plan3d &p, const CellArgs &a, const int *items, int n, const QuadLine *ql, int q, int r0,
                   int b, const float2 *WX, const float2 *WY, float *out)
{
    const int ncg = ceil_div(a.nu, kWave);
    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float2> buf(kCellLds);
    std::vector<CellLane<R>> st(kBlock);
    int v[kWavesPerBlock];
    bool on[kBlock];
    for (int w = 0; w < kWavesPerBlock; ++w) {
        const int idx = q * kWavesPerBlock + w;
        const int item = items[idx < n ? idx : q * kWavesPerBlock];
        v[w] = item / ncg;
        for (int l = 0; l < kWave; ++l) {
            const int c = (item - v[w] * ncg) * kWave + l;
            on[w * kWave + l] = idx < n && c < a.nu;
            if (on[w * kWave + l]) cell_init<R>(st[w * kWave + l], p.d_col + (int64_t)v[w] * (a.nu + 1), a, c, r0);
        }
    }
    const ViewHdr &h0 = p.d_hdr[v[0]];
    const float2 *W = cell_table(WX, WY, h0, a, b);
    const int plane = h0.nline * h0.nalong;
    std::vector<int> pitch((size_t)ceil_div(h0.nline, kCellChunk));
    cell_chunk_pitches(ql, h0.nline, pitch.data());
    for (int L = 0; L < h0.nline; ++L) {
        const VoxBox bx = cell_box<R>(ql[L], a, r0, pitch[L / kCellChunk], h0.nalong);
        // the copies stay within the line and three entries past it (the pad after the table)
        if (bx.nc > 0 && (bx.c0 < 0 || bx.c0 + bx.pitch > h0.nalong + 3)) g_cell_block_out.fetch_add(1);
        std::fill(buf.begin(), buf.end(), float2{nan, nan});
        for (int tid = 0; tid < kBlock; ++tid) voxel_stage(W + (int64_t)L * h0.nalong, bx, buf.data(), plane, tid, kBlock);
        for (int t = 0; t < kBlock; ++t) {
            if (!on[t]) continue;
            const int w = t / kWave;
            const int item = items[q * kWavesPerBlock + w];
            const ViewHdr &h = p.d_hdr[v[w]];
            const float4 lc = p.d_lineC[(int64_t)v[w] * a.maxline + L];
            cell_line_staged<OPK, GEO, R, SPAN>(st[t], h, a, lc, L, W, p.d_cgrp[item].wmax,
                                                reinterpret_cast<const char *>(buf.data()), bx);
            if (bx.nc > 0) g_cell_lds_lines.fetch_add(1);
        }
    }
    for (int t = 0; t < kBlock; ++t)
        if (on[t]) {
            const int w = t / kWave;
            const int item = items[q * kWavesPerBlock + w];
            cell_store<OPK, GEO, R>(st[t], p.d_hdr[v[w]], a, out, (item - v[w] * ncg) * kWave + t % kWave, r0, v[w], b);
        }
}

// The cell gather as launched with staging (tiles of 4, 8 or 16 rows): the (group quad,
// row tile) pairs of classes 2 and 3 that the plan marks staged, the others and class 0
// directly (k_cell).
template <int OPK, int GEO, int R>
void run_cell_staged(const Plan3d &p, const float2 *WX, const float2 *WY, float *out, int nb, int nthreads)
{
    const CellArgs a = make_cell_args(p, nb);
    const int ncg = ceil_div(a.nu, kWave);
    const int ntile = cell_tile_count(a.nv);
    struct Job {
        int k, q, b;
    };
    std::vector<Job> jobs;
    for (int k = 0; k < 3; ++k)
        for (int q = 0; q * kWavesPerBlock < p.ncls[k]; ++q)
            for (int b = 0; b < nb; ++b) jobs.push_back(Job{k, q, b});
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&, t]() {
            for (size_t j = t; j < jobs.size(); j += nthreads) {
                const Job jb = jobs[j];
                const int start = jb.k == 0 ? 0 : jb.k == 1 ? p.ncls[0] : p.ncls[0] + p.ncls[1];
                const int *items = p.d_citems + start;
                const int n = p.ncls[jb.k];
                const bool can = jb.k < 2 && (R == 4 || R == 8 || R == 16);
                const QuadLine *ql = can ? p.d_cquad + (int64_t)(p.cquad_off[jb.k] + jb.q) * a.maxline : nullptr;
                const int *flag = can ? p.d_cqstage + (int64_t)(p.cquad_off[jb.k] + jb.q) * ntile + cell_tile_offset(a.nv, R)
                                      : nullptr;
                for (int r0 = 0; r0 < a.nv; r0 += R) {
                    if (flag && flag[r0 / R]) {
                        if (jb.b == 0) g_cell_wg_staged.fetch_add(1);
                        if (jb.k == 0) run_cell_quad<OPK, GEO, R, 2>(p, a, items, n, ql, jb.q, r0, jb.b, WX, WY, out);
                        else run_cell_quad<OPK, GEO, R, 3>(p, a, items, n, ql, jb.q, r0, jb.b, WX, WY, out);
                        continue;
                    }
                    if (flag && jb.b == 0) g_cell_wg_direct.fetch_add(1);
                    for (int idx = jb.q * kWavesPerBlock; idx < std::min(n, (jb.q + 1) * kWavesPerBlock); ++idx) {
                        const int v = items[idx] / ncg, cg = items[idx] - v * ncg;
                        const float wg = p.d_cgrp[items[idx]].wmax;
                        for (int c = cg * kWave; c < std::min(a.nu, (cg + 1) * kWave); ++c) {
                            if (jb.k == 0)
                                cell_body<OPK, GEO, R, 2, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v,
                                                                 jb.b, wg);
                            else if (jb.k == 1)
                                cell_body<OPK, GEO, R, 3, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v,
                                                                 jb.b, wg);
                            else
                                cell_body<OPK, GEO, R, 0, false>(p.d_hdr, p.d_lineC, p.d_col, WX, WY, out, a, c, r0, v,
                                                                 jb.b, wg);
                        }
                    }
                }
            }
        });
    }
    for (auto &x : th) x.join();
}

template <int OPK, int GEO, int R>
void run_cell_R(const Plan3d &p, const float2 *WX, const float2 *WY, float *out, int nb, int nt)
{
    if (g_cell_variant == 3 && !p.big) run_cell_staged<OPK, GEO, R>(p, WX, WY, out, nb, nt);
    else if (g_cell_variant == 1) run_cell_RV<OPK, GEO, R, 1>(p, WX, WY, out, nb, nt);
    else if (g_cell_variant == 2) run_cell_RV<OPK, GEO, R, 2>(p, WX, WY, out, nb, nt);
    else run_cell_RV<OPK, GEO, R, 0>(p, WX, WY, out, nb, nt);
}

template <int OPK, int GEO>
void run_cell(const Plan3d &p, int R, const float2 *WX, const float2 *WY, float *out, int nb, int nt)
{
    if (R == 1) run_cell_R<OPK, GEO, 1>(p, WX, WY, out, nb, nt);
    else if (R == 2) run_cell_R<OPK, GEO, 2>(p, WX, WY, out, nb, nt);
    else if (R == 4) run_cell_R<OPK, GEO, 4>(p, WX, WY, out, nb, nt);
    else if (R == 16) run_cell_R<OPK, GEO, 16>(p, WX, WY, out, nb, nt);
    else run_cell_R<OPK, GEO, 8>(p, WX, WY, out, nb, nt);
}

// Voxel-gather variant: 0 = the workgroup path the library launches (per-view block
// staged in "LDS", lookups outside it from the table), 1 = every lookup from the table.
int g_voxel_variant = 0;

// Workgroup emulation of the staged voxel gather: per view, the block estimate, all
// threads' staging shares, then every thread's gather -- the order the barriers impose.
template <int OPK, int GEO, int KC, bool HALVES>
void run_voxel_wg_KC(const Plan3d &p, const float2 *Q, float *out, int nb, int nthreads)
{
    const VoxArgs a = make_vox_args(p, nb, KC);
    const int gx = (a.nx + kVoxTile - 1) / kVoxTile, gy = (a.ny + kVoxTile - 1) / kVoxTile;
    const int items = gx * gy * a.nchunk * nb;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&, t]() {
            std::vector<float2> tile(vox_lds(KC));
            std::vector<float> acc((size_t)kBlock * KC);
            for (int it = t; it < items; it += nthreads) {
                const int bx = it % gx, by = (it / gx) % gy, bz = it / (gx * gy);
                const int k0 = (bz % a.nchunk) * KC, b = bz / a.nchunk;
                std::fill(acc.begin(), acc.end(), 0.0f);
                for (int v = 0; v < a.nview; ++v) {
                    const ViewHdr &h = p.d_hdr[v];
                    const float4 *ct = p.d_col + (int64_t)v * (a.nu + 1);
                    const float2 *Qv = view_table(Q, a, b, v);
                    const VoxBox box = voxel_box<KC>(p.d_vtile[((int64_t)v * gy + by) * gx + bx], a, k0);
                    if (box.skip) {
                        g_voxel_skips.fetch_add(1);
                        continue;
                    }
                    for (int tid = 0; tid < kBlock; ++tid) voxel_stage(Qv, box, tile.data(), a.nu, tid, kBlock);
                    for (int tid = 0; tid < kBlock; ++tid) {
                        int i, j;
                        voxel_coords(bx, by, tid, i, j);
                        if (i >= a.nx || j >= a.ny) continue;
                        voxel_view<OPK, GEO, KC, HALVES>(h, p.d_lineV, ct, Qv, a, v, i, j, k0, voxel_px(a, i),
                                                         voxel_py(a, j), &acc[(size_t)tid * KC], box, tile.data());
                    }
                }
                for (int tid = 0; tid < kBlock; ++tid) {
                    int i, j;
                    voxel_coords(bx, by, tid, i, j);
                    if (i < a.nx && j < a.ny) voxel_store<KC>(out, a, i, j, k0, b, &acc[(size_t)tid * KC]);
                }
            }
        });
    }
    for (auto &x : th) x.join();
}

template <int OPK, int GEO, int KC, bool HALVES>
void run_voxel_KC(const Plan3d &p, const float2 *Q, float *out, int nb, int nthreads)
{
    if (g_voxel_variant == 0 && KC >= kVoxStageMin) {
        run_voxel_wg_KC<OPK, GEO, KC, HALVES>(p, Q, out, nb, nthreads);
        return;
    }
    const VoxArgs a = make_vox_args(p, nb, KC);
    const int items = nb * a.nchunk * a.ny;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&, t]() {
            for (int it = t; it < items; it += nthreads) {
                const int j = it % a.ny, rest = it / a.ny;
                const int chunk = rest % a.nchunk, b = rest / a.nchunk;
                for (int i = 0; i < a.nx; ++i)
                    voxel_body<OPK, GEO, KC, HALVES>(p.d_hdr, p.d_lineV, p.d_col, Q, out, a, i, j, chunk * KC, b, 0,
                                                     a.nview);
            }
        });
    }
    for (auto &x : th) x.join();
}

template <int OPK, int GEO>
void run_voxel(const Plan3d &p, int KC, const float2 *Q, float *out, int nb, int nt)
{
    if (KC == 1) run_voxel_KC<OPK, GEO, 1, false>(p, Q, out, nb, nt);
    else if (KC == 4) run_voxel_KC<OPK, GEO, 4, false>(p, Q, out, nb, nt);
    else if (KC == 8) run_voxel_KC<OPK, GEO, 8, false>(p, Q, out, nb, nt);
    else if (KC == 16) run_voxel_KC<OPK, GEO, 16, false>(p, Q, out, nb, nt);
    else if (KC == (16 | XR_VOXEL_HALVES)) run_voxel_KC<OPK, GEO, 16, true>(p, Q, out, nb, nt);
    else run_voxel_KC<OPK, GEO, 32, true>(p, Q, out, nb, nt);
}

template <int GEO>
void dispatch(const Plan3d &p, const Route &rt, int R, int KC, const float *in, float *out, int nb, int nt)
{
    if (rt.cell) {
        const int64_t plane = (int64_t)p.vol.nx * p.vol.ny;
        std::vector<float2> WX((size_t)nb * (p.vol.nz + 1) * plane + kVolPad), WY(WX.size());  // + zeroed pad
        for (int b = 0; b < nb; ++b)
            for (int64_t e = 0; e < plane; ++e) vol_prefix_body(in, WX.data(), plane, p.vol.nz, b, e);
        const int nx = p.vol.nx, ny = p.vol.ny;
        for (int64_t q = 0; q < (int64_t)nb * (p.vol.nz + 1); ++q)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) WY[q * plane + (int64_t)i * ny + (ny - 1 - j)] = WX[q * plane + (int64_t)j * nx + i];
        if (rt.kop == CELL_A) run_cell<CELL_A, GEO>(p, R, WX.data(), WY.data(), out, nb, nt);
        else if (rt.kop == CELL_BT) run_cell<CELL_BT, GEO>(p, R, WX.data(), WY.data(), out, nb, nt);
        else run_cell<CELL_BWT, GEO>(p, R, WX.data(), WY.data(), out, nb, nt);
    } else {
        const SinoArgs sa = make_sino_args(p, nb);
        std::vector<float2> Q((size_t)nb * p.nview * (p.geo.nv + 1) * p.geo.nu + kSinoPad);
        for (int b = 0; b < nb; ++b)
            for (int v = 0; v < p.nview; ++v)
                for (int c = 0; c < p.geo.nu; ++c) {
                    if (rt.kop == VOX_AT) sino_prefix_body<SINO_AT, GEO>(p.d_hdr, p.d_col, in, Q.data(), sa, b, v, c);
                    else sino_prefix_body<SINO_PLAIN, GEO>(p.d_hdr, p.d_col, in, Q.data(), sa, b, v, c);
                }
        if (rt.kop == VOX_B) run_voxel<VOX_B, GEO>(p, KC, Q.data(), out, nb, nt);
        else if (rt.kop == VOX_BW) run_voxel<VOX_BW, GEO>(p, KC, Q.data(), out, nb, nt);
        else run_voxel<VOX_AT, GEO>(p, KC, Q.data(), out, nb, nt);
    }
}

std::string g_err;

// The fused FDK pass on the CPU: the filter stage as fdk_filter_output (the device
// kernel's order), the in-place prefix, then the voxel gather of the weighted route.
template <int GEO>
void fdk_dispatch(const Plan3d &p, int KC, const float *in, const float *filt, const float *wgt, float *out, int nb,
                  int nt)
{
    const Route rt = route(p, XR_OP_BACKPROJECTION_W);
    const SinoArgs sa = make_sino_args(p, nb);
    const int nu = p.geo.nu, nv = p.geo.nv;
    std::vector<float2> Q((size_t)nb * p.nview * (nv + 1) * nu + kSinoPad);
    for (int b = 0; b < nb; ++b)
        for (int v = 0; v < p.nview; ++v) {
            const int64_t sv = (int64_t)b * p.nview + v;
            for (int r = 0; r < nv; ++r) {
                const float *row = in + (sv * nv + r) * nu;
                for (int u = 0; u < nu; ++u) {
                    const int c = p.d_hdr[v].keep ? u : nu - 1 - u;
                    Q[(sv * (nv + 1) + r) * nu + c] = float2{0.0f, (float)fdk_filter_output(row, wgt, filt, nu, u)};
                }
            }
            for (int c = 0; c < nu; ++c) {
                if (rt.kop == VOX_AT) fdk_prefix_body<SINO_AT, GEO>(p.d_hdr, p.d_col, Q.data(), sa, b, v, c);
                else fdk_prefix_body<SINO_PLAIN, GEO>(p.d_hdr, p.d_col, Q.data(), sa, b, v, c);
            }
        }
    if (rt.kop == VOX_B) run_voxel<VOX_B, GEO>(p, KC, Q.data(), out, nb, nt);
    else if (rt.kop == VOX_BW) run_voxel<VOX_BW, GEO>(p, KC, Q.data(), out, nb, nt);
    else run_voxel<VOX_AT, GEO>(p, KC, Q.data(), out, nb, nt);
}

int make_plan(Plan3d &p, PlanTables &t, const int *vn, const float *vf, const int *gn, const float *gf,
              const float *angles, int nview)
{
    p.vol = xr_volume3d{vn[0], vn[1], vn[2], vf[0], vf[1], vf[2], vf[3], vf[4], vf[5]};
    p.geo = xr_geometry3d{gn[0] & 0xff, gn[1], gn[2], gf[0], gf[1], gf[2], gf[3], gf[4], gf[5]};
    p.is2d = (gn[0] & 0x100) ? 1 : 0;
    p.kind = p.geo.type;
    p.nview = nview;
    std::string err;
    const int rc = build_tables(p, angles, t, err);
    if (rc) {
        g_err = err;
        return rc;
    }
    p.d_hdr = t.hdr.data();
    p.d_lineC = t.lineC.data();
    p.d_lineV = t.lineV.data();
    p.d_col = t.col.data();
    p.d_vtile = t.vtile.data();
    p.d_cgrp = t.cgrp.data();
    p.d_citems = t.citems.data();
    p.d_cquad = t.cquad.data();
    p.d_cqstage = t.cqstage.data();
    return XR_OK;
}

}  // namespace

extern "C" {

const char *emu_error(void) { return g_err.c_str(); }

void emu_set_cell_variant(int v) { g_cell_variant = v; }

void emu_set_voxel_variant(int v) { g_voxel_variant = v; }

// The staged cell gather's counts since the last reset: (thread, line) pairs served from
// an LDS block [0], lookups outside their block and blocks reaching outside their line
// and its three-entry margin [1] (must stay 0), (workgroup, row tile) pairs staged [2]
// and direct [3] (per batch item 0).
void emu_cell_lines(long long *counts, int reset)
{
#if !defined(__HIP_DEVICE_COMPILE__)  // (this file's device pass compiles nothing it uses)
    counts[0] = g_cell_lds_lines.load();
    counts[1] = cell_box_misses.load() + g_cell_block_out.load();
    counts[2] = g_cell_wg_staged.load();
    counts[3] = g_cell_wg_direct.load();
    if (reset) {
        g_cell_lds_lines = 0;
        cell_box_misses = 0;
        g_cell_block_out = 0;
        g_cell_wg_staged = 0;
        g_cell_wg_direct = 0;
    }
#endif
}

long long emu_voxel_skips(int reset)
{
    const long long n = g_voxel_skips.load();
    if (reset) g_voxel_skips = 0;
    return n;
}

// The fused FDK pass (xr_plan3d_execute_fdk) on the CPU: `in` [nb, V, Nv, Nu], filter taps
// (scaled, 2 Nu - 1), column weights (Nu), `out` [nb, Nz, Ny, Nx].
int emu_fdk3d(const int *vn, const float *vf, const int *gn, const float *gf, const float *angles, int nview,
              const float *in, const float *filt, const float *wgt, float *out, int nbatch, int voxel_chunk,
              int nthreads)
{
    Plan3d p;
    PlanTables t;
    const int rc = make_plan(p, t, vn, vf, gn, gf, angles, nview);
    if (rc) return rc;
    int KC = voxel_chunk;
    if (KC == 0) KC = default_chunk(p);
    if (nthreads < 1) nthreads = 1;
    if (p.kind == GEO_FLAT) fdk_dispatch<GEO_FLAT>(p, KC, in, filt, wgt, out, nbatch, nthreads);
    else if (p.kind == GEO_ARC) fdk_dispatch<GEO_ARC>(p, KC, in, filt, wgt, out, nbatch, nthreads);
    else fdk_dispatch<GEO_PAR>(p, KC, in, filt, wgt, out, nbatch, nthreads);
    return XR_OK;
}

// The register-held prefix bodies (columns of at most kPrefixRegRows entries) against the
// generic ones, on `in` (a volume and a sinogram of the plan's shapes, nb items): 1 when
// both tables agree bitwise, 0 otherwise.
int emu_prefix_bodies_agree(const int *vn, const float *vf, const int *gn, const float *gf, const float *angles,
                            int nview, const float *vol, const float *sino, int nb)
{
    Plan3d p;
    PlanTables t;
    const int rc = make_plan(p, t, vn, vf, gn, gf, angles, nview);
    if (rc) return rc;
    const int64_t plane = (int64_t)p.vol.nx * p.vol.ny;
    const size_t nvol = (size_t)nb * (p.vol.nz + 1) * plane;
    std::vector<float2> W1(nvol), W2(nvol);
    if (p.vol.nz <= kPrefixRegRows)
        for (int b = 0; b < nb; ++b)
            for (int64_t e = 0; e < plane; ++e) {
                vol_prefix_body(vol, W1.data(), plane, p.vol.nz, b, e);
                vol_prefix_reg(vol, W2.data(), plane, p.vol.nz, b, e);
            }
    const SinoArgs sa = make_sino_args(p, nb);
    const size_t nq = (size_t)nb * p.nview * (p.geo.nv + 1) * p.geo.nu;
    std::vector<float2> Q1(nq), Q2(nq), Q3(nq), Q4(nq);
    if (p.geo.nv <= kPrefixRegRows)
        for (int b = 0; b < nb; ++b)
            for (int v = 0; v < p.nview; ++v)
                for (int c = 0; c < p.geo.nu; ++c) {
                    sino_prefix_body<SINO_PLAIN, GEO_FLAT>(p.d_hdr, p.d_col, sino, Q1.data(), sa, b, v, c);
                    sino_prefix_reg<SINO_PLAIN, GEO_FLAT>(p.d_hdr, p.d_col, sino, Q2.data(), sa, b, v, c);
                    sino_prefix_body<SINO_AT, GEO_FLAT>(p.d_hdr, p.d_col, sino, Q3.data(), sa, b, v, c);
                    sino_prefix_reg<SINO_AT, GEO_FLAT>(p.d_hdr, p.d_col, sino, Q4.data(), sa, b, v, c);
                }
    const bool same = std::memcmp(W1.data(), W2.data(), nvol * sizeof(float2)) == 0 &&
                      std::memcmp(Q1.data(), Q2.data(), nq * sizeof(float2)) == 0 &&
                      std::memcmp(Q3.data(), Q4.data(), nq * sizeof(float2)) == 0;
    return same ? 1 : 0;
}

// Span classes of the cell gather for a geometry: per view (ViewHdr::span) into
// `spans` [nview], and per (view, 64-column group) (CellGroup::span) into `group_spans`
// [nview][ceil(Nu/64)] when it is not null.
int emu_view_spans(const int *vn, const float *vf, const int *gn, const float *gf, const float *angles, int nview,
                   int *spans, int *group_spans)
{
    Plan3d p;
    PlanTables t;
    const int rc = make_plan(p, t, vn, vf, gn, gf, angles, nview);
    if (rc) return rc;
    for (int v = 0; v < nview; ++v) spans[v] = t.hdr[v].span;
    if (group_spans)
        for (size_t k = 0; k < t.cgrp.size(); ++k) group_spans[k] = t.cgrp[k].span;
    return XR_OK;
}

// The staged cell gather's blocks for tiles of R rows (R2 for class 2, R3 for class 3),
// over every (workgroup, row tile, line) with a column on the line, at the chunk pitches:
// stats[0] blocks, stats[1] over a slot, stats[2] (workgroup, row tile) pairs staged,
// stats[3] pairs direct (the plan's flags), stats[4] entries staged (pitch x slices) in the
// staged pairs, stats[5] the largest block (entries), stats[6] blocks over 1024 entries.
int emu_cell_box_stats(const int *vn, const float *vf, const int *gn, const float *gf, const float *angles, int nview,
                       int R2, int R3, long long *stats)
{
    Plan3d p;
    PlanTables t;
    const int rc = make_plan(p, t, vn, vf, gn, gf, angles, nview);
    if (rc) return rc;
    const CellArgs a = make_cell_args(p, 1);
    const int ntile = cell_tile_count(a.nv);
    for (int k = 0; k < 7; ++k) stats[k] = 0;
    std::vector<int> pitch((size_t)ceil_div(p.maxline, kCellChunk));
    for (int k = 0; k < 2; ++k) {
        const int R = k == 0 ? R2 : R3;
        for (int q = 0; q * kWavesPerBlock < p.ncls[k]; ++q) {
            const QuadLine *ql = t.cquad.data() + (size_t)(p.cquad_off[k] + q) * p.maxline;
            const int *flag = t.cqstage.data() + (size_t)(p.cquad_off[k] + q) * ntile + cell_tile_offset(a.nv, R);
            const int v0 = t.citems[(k == 0 ? 0 : p.ncls[0]) + q * kWavesPerBlock] / ceil_div(a.nu, kWave);
            const int nline = t.hdr[v0].nline, nalong = t.hdr[v0].nalong;
            cell_chunk_pitches(ql, nline, pitch.data());
            for (int r0 = 0; r0 < a.nv; r0 += R) {
                const bool staged = flag[r0 / R] != 0;
                ++stats[staged ? 2 : 3];
                for (int L = 0; L < nline; ++L) {
                    if (ql[L].nm <= 0) continue;
                    ++stats[0];
                    const int pt = pitch[L / kCellChunk];
                    VoxBox bx;
                    if (R == 4) bx = cell_box<4>(ql[L], a, r0, pt, nalong);
                    else if (R == 8) bx = cell_box<8>(ql[L], a, r0, pt, nalong);
                    else bx = cell_box<16>(ql[L], a, r0, pt, nalong);
                    long long n = (long long)bx.pitch * bx.nr;
                    if (bx.nc == 0) {  // over a slot: its size without the rounding margin
                        const float h0 = fmaf((float)r0, a.dvv, a.v0), h1 = fmaf((float)(r0 + R), a.dvv, a.v0);
                        const QuadLine &qq = ql[L];
                        const float lo = std::min(std::min(qq.sg_lo * h0, qq.sg_hi * h0), std::min(qq.sg_lo * h1, qq.sg_hi * h1));
                        const float hi = std::max(std::max(qq.sg_lo * h0, qq.sg_hi * h0), std::max(qq.sg_lo * h1, qq.sg_hi * h1));
                        const int ns = (int)std::min(std::floor(hi + a.zoff), (float)a.nz) -
                                       (int)std::max(std::floor(lo + a.zoff), 0.0f) + 1;
                        n = (long long)ns * pt;
                        ++stats[1];
                    } else if (staged) {
                        stats[4] += n;
                    }
                    stats[5] = std::max(stats[5], n);
                    if (n > 1024) ++stats[6];
                }
            }
        }
    }
    return XR_OK;
}

// vn = {nx, ny, nz}, vf = {dx, dy, dz, shx, shy, shz}, gn = {type(|0x100 for 2D), nu, nv},
// gf = {du, dv, shu, shv, sod, sdd}; tile_rows / voxel_chunk 0 = the library defaults
// (8/4/2/1 by Nv, 16/8/4/1 by Nz).
int emu_op3d(int op, const int *vn, const float *vf, const int *gn, const float *gf, const float *angles, int nview,
             const float *in, float *out, int nbatch, int tile_rows, int voxel_chunk, int nthreads)
{
    Plan3d p;
    PlanTables t;
    const int rc = make_plan(p, t, vn, vf, gn, gf, angles, nview);
    if (rc) return rc;
    const Route rt = route(p, op);
    if (rt.kop < 0) {
        g_err = "unknown op";
        return XR_EINVAL;
    }
    int R = tile_rows, KC = voxel_chunk;
    if (R == 0) R = p.geo.nv >= 32 ? 8 : p.geo.nv >= 16 ? 4 : p.geo.nv >= 4 ? 2 : 1;
    if (KC == 0) KC = default_chunk(p);
    if (nthreads < 1) nthreads = 1;
    if (p.kind == GEO_FLAT) dispatch<GEO_FLAT>(p, rt, R, KC, in, out, nbatch, nthreads);
    else if (p.kind == GEO_ARC) dispatch<GEO_ARC>(p, rt, R, KC, in, out, nbatch, nthreads);
    else dispatch<GEO_PAR>(p, rt, R, KC, in, out, nbatch, nthreads);
    return XR_OK;
}

}  // extern "C"
