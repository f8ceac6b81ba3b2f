/*
 * ORACLE — test infrastructure, not product code.
 *
 * FP64 CPU restatement of the distance-driven (DD) cone-beam discretisation that the
 * reference's CPU kernels implement in FP32.  Only tests/, __graft_entry__.smoke() and
 * bench.py (its cpu_baseline leg and its nnz count) may load this library, and only as
 * the checker.  The product path never calls it.
 *
 * Conventions restated here (reference file:line):
 *   voxel centres            include/cpu3d_common.h:60-62 (row 0 at +y)
 *   source position          include/xrmath.h:116-123   s = (-SOD sin b, SOD cos b)
 *   x-driven / keep flags    include/cpu3d_common.h:65, :87 ; src/cpu/forward3d.cpp:309-310
 *                            decided in FP32 exactly as the reference does (sinf/cosf of
 *                            the float angle, squared in FP32, compared with 0.5f)
 *   axis maps                include/xrmath.h:8-38 (map2x / map2y)
 *   flat detector columns    include/cpu3d_common.h:94-149 ; forward3d.cpp:229-269
 *   arc detector columns     include/cpu3d_common.h:150-201 ; forward3d.cpp:271-297
 *   row boundaries, z-scale  cpu3d_common.h:127-131, :178-182 ; forward3d.cpp:318-365
 *   forward scale            forward3d.cpp:330-344 (dy or dx / (area * cos))
 *   backprojection coeff     cpu3d_common.h:80-85, :141, :192 (col_width*ov_u*ov_z/area * U)
 *   FDK voxel weight         xrmath.h:76-86 (flat SOD^2/(SOD - s.p/SOD)^2, arc 1/|s-p|^2)
 *   parallel-beam visitor    forward3d.cpp:37-109 (A; B == A^T and weighted == plain,
 *                            backward3d.cpp:101-135)
 *
 * For the cone geometries every operator is one coefficient per (voxel, view, cell):
 *   coefA = ov_u * ov_z * Delta / (W_c * H_cr * cos theta_cr)       (projection A)
 *   coefB = U * w * ov_u * ov_z / A_vox                             (backprojection B)
 * and A^T, B^T are their exact transposes.  This file enumerates those pairs voxel by
 * voxel (a different traversal from the reference's detector-driven walk for A), in
 * double precision, so that agreement with the reference pins both.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, shx, shy, shz;
    int type; /* 0 parallel, 1 flat, 2 arc */
    int nu, nv;
    double du, dv, shu, shv, sod, sdd;
} geo3_t;

typedef struct {
    int xdrv, keep;
    double ang, sb, cb, sx, sy;
    double *d;   /* nu+1 column boundaries on the driving axis, increasing order */
    double *sig; /* nu per-column z scale */
} view3_t;

static void geo3_from(geo3_t *g, const int *vn, const float *vf, const int *gn, const float *gf)
{
    g->nx = vn[0]; g->ny = vn[1]; g->nz = vn[2];
    g->dx = vf[0]; g->dy = vf[1]; g->dz = vf[2];
    g->shx = vf[3]; g->shy = vf[4]; g->shz = vf[5];
    g->type = gn[0]; g->nu = gn[1]; g->nv = gn[2];
    g->du = gf[0]; g->dv = gf[1]; g->shu = gf[2]; g->shv = gf[3];
    g->sod = gf[4]; g->sdd = gf[5];
}

static double axis_map(int xdrv, double sx, double sy, double px, double py)
{
    return xdrv ? (sx * py - sy * px) / (py - sy) : (sy * px - sx * py) / (px - sx);
}

static void view3_build(const geo3_t *g, float angf, view3_t *w)
{
    const float sf = sinf(angf), cf = cosf(angf);
    w->xdrv = (cf * cf > 0.5f);
    w->keep = w->xdrv ? (cf >= 0.0f) : (sf >= 0.0f);
    w->ang = (double)angf;
    w->sb = sin(w->ang);
    w->cb = cos(w->ang);
    w->sx = -g->sod * w->sb;
    w->sy = g->sod * w->cb;
    const int nu = g->nu;
    w->d = (double *)malloc(sizeof(double) * (nu + 1));
    w->sig = (double *)malloc(sizeof(double) * nu);
    const double mag = g->sod / g->sdd;
    for (int c = 0; c <= nu; ++c) {
        const double b = w->keep ? c : (nu - c);
        if (g->type == 1) {
            const double u = ((b - nu * 0.5) * g->du + g->shu) * mag;
            w->d[c] = axis_map(w->xdrv, w->sx, w->sy, u * w->cb, u * w->sb);
        } else {
            const double gam = (b - nu * 0.5) * g->du + g->shu + w->ang;
            w->d[c] = w->xdrv ? w->sx + w->sy * tan(gam) : w->sy + w->sx / tan(gam);
        }
    }
    for (int c = 0; c < nu; ++c) {
        const double bm = w->keep ? (c + 0.5) : (nu - c - 0.5);
        double mx, my;
        if (g->type == 1) {
            const double u = ((bm - nu * 0.5) * g->du + g->shu) * mag;
            mx = u * w->cb;
            my = u * w->sb;
        } else {
            const double gam = (bm - nu * 0.5) * g->du + g->shu + w->ang;
            mx = sin(gam) * g->sdd + w->sx;
            my = -cos(gam) * g->sdd + w->sy;
        }
        w->sig[c] = w->xdrv ? w->sy / (w->sy - my) : w->sx / (w->sx - mx);
    }
}

static void view3_free(view3_t *w)
{
    free(w->d);
    free(w->sig);
}

static double row_bound(const geo3_t *g, int r)
{
    const double scale = (g->type == 1) ? g->sod / g->sdd : 1.0;
    return ((r - g->nv * 0.5) * g->dv + g->shv) * scale;
}

static double ovl(double a0, double a1, double b0, double b1)
{
    const double lo = fmax(fmin(a0, a1), fmin(b0, b1));
    const double hi = fmin(fmax(a0, a1), fmax(b0, b1));
    return hi - lo > 0.0 ? hi - lo : 0.0;
}

/* ---------------------------------------------------------------- cone operators */

enum { OP_A = 0, OP_AT = 1, OP_B = 2, OP_BW = 3, OP_BT = 4, OP_BWT = 5, OP_COUNT = 6 };

typedef struct {
    const geo3_t *g;
    const float *ang;
    int nview, nbatch, op;
    const double *in;
    double *out;
    int64_t nnz;
    int lo, hi; /* work range: views for A/B^T/count, image rows j for A^T/B */
} job3_t;

/* Visit every (voxel, cell) pair of view w for image row j: calls the accumulation for
 * the operator in job. */
static void cone_row(job3_t *jb, const view3_t *w, int v, int j)
{
    const geo3_t *g = jb->g;
    const int nx = g->nx, ny = g->ny, nz = g->nz, nu = g->nu, nv = g->nv;
    const int64_t nvox = (int64_t)nx * ny * nz;
    const int64_t nsino = (int64_t)jb->nview * nv * nu;
    const double delta = w->xdrv ? g->dy : g->dx;
    const double colw = (g->type == 1) ? g->du * g->sod / g->sdd : 2.0 * g->sod * tan(0.5 * g->du);
    const double P = w->xdrv ? w->sx : w->sy;
    const double S = w->xdrv ? w->sy : w->sx;
    const double py = (ny * 0.5 - j - 0.5) * g->dy + g->shy;
    for (int i = 0; i < nx; ++i) {
        const double px = (i - nx * 0.5 + 0.5) * g->dx + g->shx;
        double a0, a1, zeta;
        if (w->xdrv) {
            a0 = axis_map(1, w->sx, w->sy, px - 0.5 * g->dx, py);
            a1 = axis_map(1, w->sx, w->sy, px + 0.5 * g->dx, py);
            zeta = w->sy / (w->sy - py);
        } else {
            a0 = axis_map(0, w->sx, w->sy, px, py - 0.5 * g->dy);
            a1 = axis_map(0, w->sx, w->sy, px, py + 0.5 * g->dy);
            zeta = w->sx / (w->sx - px);
        }
        if (!(fmax(a0, w->d[0]) < fmin(a1, w->d[nu]))) continue;
        double U = 1.0;
        if (jb->op == OP_BW || jb->op == OP_BWT) {
            if (g->type == 1) {
                const double dd = (w->sx * px + w->sy * py) / g->sod;
                U = g->sod * g->sod / ((g->sod - dd) * (g->sod - dd));
            } else {
                U = 1.0 / ((w->sx - px) * (w->sx - px) + (w->sy - py) * (w->sy - py));
            }
        }
        /* first column whose right boundary exceeds a0 */
        int lo = 0, hi = nu - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (w->d[mid + 1] > a0) hi = mid; else lo = mid + 1;
        }
        for (int c = lo; c < nu && w->d[c] < a1; ++c) {
            const double ou = fmin(a1, w->d[c + 1]) - fmax(a0, w->d[c]);
            if (!(ou > 0.0)) continue;
            const double W = w->d[c + 1] - w->d[c];
            const double sg = w->sig[c];
            const double dbar = 0.5 * (w->d[c] + w->d[c + 1]);
            const int col = w->keep ? c : (nu - 1 - c);
            const double h0 = sg * row_bound(g, 0);
            const double hstep = sg * (row_bound(g, 1) - row_bound(g, 0));
            for (int k = 0; k < nz; ++k) {
                const double pz = (k - nz * 0.5 + 0.5) * g->dz + g->shz;
                const double z0 = zeta * (pz - 0.5 * g->dz);
                const double z1 = zeta * (pz + 0.5 * g->dz);
                const double avox = fabs((a1 - a0) * (z1 - z0));
                if (hstep == 0.0) break;
                double t0 = (fmin(z0, z1) - h0) / hstep, t1 = (fmax(z0, z1) - h0) / hstep;
                if (t0 > t1) { const double t = t0; t0 = t1; t1 = t; }
                int rb = (int)floor(t0) - 1, re = (int)floor(t1) + 1;
                if (rb < 0) rb = 0;
                if (re > nv - 1) re = nv - 1;
                for (int r = rb; r <= re; ++r) {
                    const double hr0 = sg * row_bound(g, r), hr1 = sg * row_bound(g, r + 1);
                    const double oz = ovl(z0, z1, hr0, hr1);
                    if (!(oz > 0.0)) continue;
                    if (jb->op == OP_COUNT) { jb->nnz++; continue; }
                    double coef;
                    if (jb->op == OP_A || jb->op == OP_AT) {
                        const double H = hr1 - hr0;
                        const double hbar = 0.5 * (hr0 + hr1);
                        const double cosw = fabs(S) / sqrt((dbar - P) * (dbar - P) + hbar * hbar + S * S);
                        coef = ou * oz * delta / (W * H * cosw);
                    } else {
                        if (!(avox > 1e-12)) continue;
                        coef = U * colw * ou * oz / avox;
                    }
                    const int64_t vi = ((int64_t)k * ny + j) * nx + i;
                    const int64_t si = ((int64_t)v * nv + r) * nu + col;
                    for (int b = 0; b < jb->nbatch; ++b) {
                        if (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT)
                            jb->out[b * nsino + si] += coef * jb->in[b * nvox + vi];
                        else
                            jb->out[b * nvox + vi] += coef * jb->in[b * nsino + si];
                    }
                }
            }
        }
    }
}

static void *cone_worker(void *arg)
{
    job3_t *jb = (job3_t *)arg;
    const geo3_t *g = jb->g;
    const int scatter_to_sino = (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT || jb->op == OP_COUNT);
    if (scatter_to_sino) {
        for (int v = jb->lo; v < jb->hi; ++v) {
            view3_t w;
            view3_build(g, jb->ang[v], &w);
            for (int j = 0; j < g->ny; ++j) cone_row(jb, &w, v, j);
            view3_free(&w);
        }
    } else {
        for (int v = 0; v < jb->nview; ++v) {
            view3_t w;
            view3_build(g, jb->ang[v], &w);
            for (int j = jb->lo; j < jb->hi; ++j) cone_row(jb, &w, v, j);
            view3_free(&w);
        }
    }
    return NULL;
}

/* ------------------------------------------------------------ parallel-beam 3D */

/* Restates visit_parallel3d_coefficients (forward3d.cpp:37-109) in FP64: the pixel
 * strip on the detector axis, divided by |cos| or |sin|, against uniform columns;
 * depth against uniform rows; coefficient = ov_axis * (dy or dx)/du * ov_z / dv. */
static void par_row(job3_t *jb, float angf, int v, int j)
{
    const geo3_t *g = jb->g;
    const int nx = g->nx, ny = g->ny, nz = g->nz, nu = g->nu, nv = g->nv;
    const int64_t nvox = (int64_t)nx * ny * nz;
    const int64_t nsino = (int64_t)jb->nview * nv * nu;
    const float sf = sinf(angf), cf = cosf(angf);
    const int xdrv = cf * cf > 0.5f;
    const int keep = xdrv ? (cf >= 0.0f) : (sf >= 0.0f);
    const double s = sin((double)angf), c = cos((double)angf);
    const double idx0 = keep ? -0.5 * nu : 0.5 * nu;
    const double det0 = xdrv ? (idx0 * g->du + g->shu) / c : (idx0 * g->du + g->shu) / s;
    const double dint = xdrv ? g->du / fabs(c) : g->du / fabs(s);
    const double pint = xdrv ? g->dx : g->dy;
    const double xys = xdrv ? g->dy / g->du : g->dx / g->du;
    const double dz0 = -0.5 * nv * g->dv + g->shv;
    for (int i = 0; i < nx; ++i) {
        double ax0;
        if (xdrv) {
            const double px0 = (i - 0.5 * nx) * g->dx + g->shx;
            const double py = (0.5 * ny - j - 0.5) * g->dy + g->shy;
            ax0 = (s / c) * py + px0;
        } else {
            const double px = (i - 0.5 * nx + 0.5) * g->dx + g->shx;
            const double py0 = (0.5 * ny - j - 1.0) * g->dy + g->shy;
            ax0 = (c / s) * px + py0;
        }
        const double ax1 = ax0 + pint;
        for (int u = 0; u < nu; ++u) {
            const double ou = ovl(ax0, ax1, det0 + u * dint, det0 + (u + 1) * dint);
            if (!(ou > 0.0)) continue;
            const int col = keep ? u : (nu - 1 - u);
            for (int k = 0; k < nz; ++k) {
                const double z0 = (k - 0.5 * nz) * g->dz + g->shz;
                for (int r = 0; r < nv; ++r) {
                    const double oz = ovl(z0, z0 + g->dz, dz0 + r * g->dv, dz0 + (r + 1) * g->dv);
                    if (!(oz > 0.0)) continue;
                    if (jb->op == OP_COUNT) { jb->nnz++; continue; }
                    const double coef = ou * xys * oz / g->dv;
                    const int64_t vi = ((int64_t)k * ny + j) * nx + i;
                    const int64_t si = ((int64_t)v * nv + r) * nu + col;
                    for (int b = 0; b < jb->nbatch; ++b) {
                        if (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT)
                            jb->out[b * nsino + si] += coef * jb->in[b * nvox + vi];
                        else
                            jb->out[b * nvox + vi] += coef * jb->in[b * nsino + si];
                    }
                }
            }
        }
    }
}

static void *par_worker(void *arg)
{
    job3_t *jb = (job3_t *)arg;
    const int scatter_to_sino = (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT || jb->op == OP_COUNT);
    if (scatter_to_sino) {
        for (int v = jb->lo; v < jb->hi; ++v)
            for (int j = 0; j < jb->g->ny; ++j) par_row(jb, jb->ang[v], v, j);
    } else {
        for (int v = 0; v < jb->nview; ++v)
            for (int j = jb->lo; j < jb->hi; ++j) par_row(jb, jb->ang[v], v, j);
    }
    return NULL;
}

/* op: 0 A, 1 A^T, 2 B, 3 B_W, 4 B^T, 5 B_W^T, 6 count (returns nnz, ignores in/out).
 * For A, B^T, B_W^T `in` is the volume [nbatch, nz, ny, nx] and `out` the sinogram
 * [nbatch, nview, nv, nu]; for A^T, B, B_W the reverse.  `out` is overwritten. */
int64_t oracle_op3d(int op, const double *in, double *out, const float *ang,
                    const int *vn, const float *vf, const int *gn, const float *gf,
                    int nbatch, int nview, int nthreads)
{
    geo3_t g;
    geo3_from(&g, vn, vf, gn, gf);
    if (op < 0 || op > OP_COUNT) return -1;
    const int scatter_to_sino = (op == OP_A || op == OP_BT || op == OP_BWT || op == OP_COUNT);
    if (op != OP_COUNT) {
        const int64_t n = scatter_to_sino ? (int64_t)nbatch * nview * g.nv * g.nu
                                          : (int64_t)nbatch * g.nz * g.ny * g.nx;
        memset(out, 0, sizeof(double) * (size_t)n);
    }
    if (nthreads < 1) nthreads = 1;
    const int items = scatter_to_sino ? nview : g.ny;
    if (nthreads > items) nthreads = items > 0 ? items : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    job3_t *jobs = (job3_t *)calloc(nthreads, sizeof(job3_t));
    for (int t = 0; t < nthreads; ++t) {
        job3_t *jb = &jobs[t];
        jb->g = &g; jb->ang = ang; jb->nview = nview; jb->nbatch = nbatch; jb->op = op;
        jb->in = in; jb->out = out; jb->nnz = 0;
        jb->lo = (int)((int64_t)items * t / nthreads);
        jb->hi = (int)((int64_t)items * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, g.type == 0 ? par_worker : cone_worker, jb);
    }
    int64_t nnz = 0;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        nnz += jobs[t].nnz;
    }
    free(th);
    free(jobs);
    return op == OP_COUNT ? nnz : 0;
}

/* ------------------------------------------------------------------- 2D operators */

/* The reference's 2D path, restated separately from the 3D one:
 *   A    src/cpu/system_matrix2d.cpp:87-174 (parallel), :176-267 (fan-flat), :269-357 (fan-arc)
 *        coefficient = ov * (dy or dx) / (W_c * cos2d), cos2d = |S|/sqrt((dbar-P)^2 + S^2)
 *        (cos_weight_x/y, :58-68); parallel: ov * (dy or dx) / du.
 *   B    src/cpu/backward2d.cpp:30-98 (fan-flat), :100-163 (fan-arc), :165-205 (parallel)
 *        coefficient = ov * w * U / (p1 - p0) with w = du (flat), 2 SOD tan(du/2) (arc);
 *        parallel: ov * du / (dx or dy), no weighting.
 *   A^T, B^T are the transposes (forward2d.cpp:80-111, backward2d.cpp:433-475).
 * Here du is the effective pitch (geometry.py:106-109); geometry passed as
 * vn = {Nx, Ny}, vf = {dx, dy, shx, shy}, gn = {type, Nu}, gf = {du, shu, SOD, SDD}. */
typedef struct {
    int nx, ny, type, nu, nview, nbatch, op;
    double dx, dy, shx, shy, du, shu, sod, sdd;
    const float *ang;
    const double *in;
    double *out;
    int lo, hi;
} job2_t;

static void view2_pixel(job2_t *jb, int v, int j, double *dcol, int xdrv, int keep, double sb, double cb)
{
    const int nx = jb->nx, ny = jb->ny, nu = jb->nu;
    const int64_t nimg = (int64_t)nx * ny, nsino = (int64_t)jb->nview * nu;
    const double sx = -jb->sod * sb, sy = jb->sod * cb;
    const double P = xdrv ? sx : sy, S = xdrv ? sy : sx;
    const double delta = xdrv ? jb->dy : jb->dx;
    const double py = (ny * 0.5 - j - 0.5) * jb->dy + jb->shy;
    for (int i = 0; i < nx; ++i) {
        const double px = (i - nx * 0.5 + 0.5) * jb->dx + jb->shx;
        double a0, a1;
        if (jb->type == 0) {
            if (xdrv) { a0 = (sb / cb) * py + px - 0.5 * jb->dx; a1 = a0 + jb->dx; }
            else { a0 = (cb / sb) * px + py - 0.5 * jb->dy; a1 = a0 + jb->dy; }
        } else if (xdrv) {
            a0 = axis_map(1, sx, sy, px - 0.5 * jb->dx, py);
            a1 = axis_map(1, sx, sy, px + 0.5 * jb->dx, py);
        } else {
            a0 = axis_map(0, sx, sy, px, py - 0.5 * jb->dy);
            a1 = axis_map(0, sx, sy, px, py + 0.5 * jb->dy);
        }
        double U = 1.0;
        if (jb->op == OP_BW || jb->op == OP_BWT) {
            if (jb->type == 1) {
                const double dd = (sx * px + sy * py) / jb->sod;
                U = jb->sod * jb->sod / ((jb->sod - dd) * (jb->sod - dd));
            } else if (jb->type == 2) {
                U = 1.0 / ((sx - px) * (sx - px) + (sy - py) * (sy - py));
            }
        }
        const double colw = jb->type == 2 ? 2.0 * jb->sod * tan(0.5 * jb->du) : jb->du;
        for (int c = 0; c < nu; ++c) {
            const double ou = ovl(a0, a1, dcol[c], dcol[c + 1]);
            if (!(ou > 0.0)) continue;
            double coef;
            if (jb->op == OP_A || jb->op == OP_AT) {
                if (jb->type == 0) {
                    coef = ou * delta / jb->du;
                } else {
                    const double W = dcol[c + 1] - dcol[c], dbar = 0.5 * (dcol[c] + dcol[c + 1]);
                    const double cosw = fabs(S) / sqrt((dbar - P) * (dbar - P) + S * S);
                    coef = ou * delta / (W * cosw);
                }
            } else {
                coef = ou * colw * U / (a1 - a0);
            }
            const int col = keep ? c : (nu - 1 - c);
            const int64_t ii = (int64_t)j * nx + i, si = (int64_t)v * nu + col;
            for (int b = 0; b < jb->nbatch; ++b) {
                if (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT)
                    jb->out[b * nsino + si] += coef * jb->in[b * nimg + ii];
                else
                    jb->out[b * nimg + ii] += coef * jb->in[b * nsino + si];
            }
        }
    }
}

static void view2_cols(const job2_t *jb, float angf, double *dcol, int *xdrv, int *keep, double *sb, double *cb)
{
    const float sf = sinf(angf), cf = cosf(angf);
    *xdrv = cf * cf > 0.5f;
    *keep = *xdrv ? (cf >= 0.0f) : (sf >= 0.0f);
    const double a = (double)angf;
    *sb = sin(a);
    *cb = cos(a);
    const double sx = -jb->sod * *sb, sy = jb->sod * *cb;
    for (int c = 0; c <= jb->nu; ++c) {
        const double idx = *keep ? (c - jb->nu * 0.5) : (jb->nu * 0.5 - c);
        if (jb->type == 0) {
            dcol[c] = (idx * jb->du + jb->shu) / (*xdrv ? *cb : *sb);
        } else if (jb->type == 1) {
            const double u = idx * jb->du + jb->shu;
            dcol[c] = axis_map(*xdrv, sx, sy, u * *cb, u * *sb);
        } else {
            const double g = idx * jb->du + jb->shu + a;
            dcol[c] = *xdrv ? sx + sy * tan(g) : sy + sx / tan(g);
        }
    }
}

static void *worker2d(void *arg)
{
    job2_t *jb = (job2_t *)arg;
    double *dcol = (double *)malloc(sizeof(double) * (jb->nu + 1));
    const int to_sino = (jb->op == OP_A || jb->op == OP_BT || jb->op == OP_BWT);
    int xdrv, keep;
    double sb, cb;
    if (to_sino) {
        for (int v = jb->lo; v < jb->hi; ++v) {
            view2_cols(jb, jb->ang[v], dcol, &xdrv, &keep, &sb, &cb);
            for (int j = 0; j < jb->ny; ++j) view2_pixel(jb, v, j, dcol, xdrv, keep, sb, cb);
        }
    } else {
        for (int v = 0; v < jb->nview; ++v) {
            view2_cols(jb, jb->ang[v], dcol, &xdrv, &keep, &sb, &cb);
            for (int j = jb->lo; j < jb->hi; ++j) view2_pixel(jb, v, j, dcol, xdrv, keep, sb, cb);
        }
    }
    free(dcol);
    return NULL;
}

/* op as oracle_op3d (0..5).  Image [nbatch, Ny, Nx], sinogram [nbatch, nview, Nu]. */
int oracle_op2d(int op, const double *in, double *out, const float *ang, const int *vn, const float *vf,
                const int *gn, const float *gf, int nbatch, int nview, int nthreads)
{
    if (op < 0 || op > OP_BWT) return -1;
    job2_t base;
    memset(&base, 0, sizeof(base));
    base.nx = vn[0]; base.ny = vn[1];
    base.dx = vf[0]; base.dy = vf[1]; base.shx = vf[2]; base.shy = vf[3];
    base.type = gn[0]; base.nu = gn[1];
    base.du = gf[0]; base.shu = gf[1]; base.sod = gf[2]; base.sdd = gf[3];
    base.nview = nview; base.nbatch = nbatch; base.op = op; base.ang = ang; base.in = in; base.out = out;
    const int to_sino = (op == OP_A || op == OP_BT || op == OP_BWT);
    const int64_t n = to_sino ? (int64_t)nbatch * nview * base.nu : (int64_t)nbatch * base.nx * base.ny;
    memset(out, 0, sizeof(double) * (size_t)n);
    const int items = to_sino ? nview : base.ny;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > items) nthreads = items > 0 ? items : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    job2_t *jobs = (job2_t *)malloc(sizeof(job2_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = base;
        jobs[t].lo = (int)((int64_t)items * t / nthreads);
        jobs[t].hi = (int)((int64_t)items * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, worker2d, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    return 0;
}

/* Per-view FP64 geometry of the cone oracle, exported so tests can check the tables
 * the product builds: flags[2] = {xdrv, keep}, src[2] = {sx, sy}, d[nu+1], sig[nu]. */
int oracle_view3d(float ang, const int *vn, const float *vf, const int *gn, const float *gf,
                  int *flags, double *src, double *d, double *sig)
{
    geo3_t g;
    geo3_from(&g, vn, vf, gn, gf);
    view3_t w;
    view3_build(&g, ang, &w);
    flags[0] = w.xdrv;
    flags[1] = w.keep;
    src[0] = w.sx;
    src[1] = w.sy;
    memcpy(d, w.d, sizeof(double) * (g.nu + 1));
    memcpy(sig, w.sig, sizeof(double) * g.nu);
    view3_free(&w);
    return 0;
}
