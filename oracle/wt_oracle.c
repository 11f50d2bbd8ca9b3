/* wt_oracle.c -- CPU restatement of the reference solver and voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the HIP library is checked
 * against: tests/, the smoke check and the CPU-baseline leg of the benchmark
 * load it; the product path never does.
 *
 * Follows js/lbm.js (solver), js/lbmWorker.js:43-79 and js/pool.js:97-125
 * (slab split, per-step force totals in worker order) and js/voxel.js
 * (three-axis ray-parity voxelizer). Arithmetic is IEEE binary64 in the
 * reference's expression order; the file must be compiled with
 * -ffp-contract=off and without fast-math so that, with float storage, every
 * stored value equals the reference's Float32Array contents bit for bit.
 * The "double" storage variant is the FP64 oracle of the precision study.
 *
 * Pinning: golden fixtures written from the reference under Node
 * (tests/golden/) pin solver trajectories, voxel grids and statistics
 * bitwise; closed-form checks (uniform-flow fixed point, conservation,
 * brute-force point-in-polyhedron) cover the rest in tests/.
 */
#include <math.h>
#include <pthread.h>

/* Mistakes a test plants on purpose to show that the gray bounce-back
 * checks catch them (1: force sign, 2: one-sided link weight, 3: mixing
 * weights swapped); 0, the only value the Makefile builds, plants none. */
#ifndef WTO_PLANT_GRAY
#define WTO_PLANT_GRAY 0
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* velocity set, opposite table and weights (js/lbm.js:28-40) */
static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static double W[19];
#define W0 (1.0 / 3)
#define WF (1.0 / 18)
#define WE (1.0 / 36)
static double SQ18;
enum { FLUID = 0, SOLID_OBJ = 1, SOLID_WALL = 2 };

typedef struct {
    int nx, ny, nz, prec; /* prec: 4 = float storage, 8 = double storage */
    long n;
    void *buf[2];         /* bufA, bufB; f = buf[parity], g = buf[parity^1] */
    uint8_t *solid, *near;
    void *ux, *uy, *uz, *rho;
    float *uwx, *uwy, *uwz;
    float *wgt;           /* gray bounce-back weights in (0,1) per cell, or NULL */
    long off[19];
    double u0, tau, csm, lambda;
    int les, trt, movingWalls, ground, parity;
    int channel;          /* 1: channel boundaries (free-slip sides, equilibrium outflow) */
    long steps, frontalArea, planformArea, solidCells, rideHeight;
    double force[3];
} wto_lbm;

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* ECMAScript ToInt32 (the `|0` operator). */
static int js_toint32(double d)
{
    if (!(d == d) || isinf(d)) return 0;
    double t = trunc(d);
    double m = fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    uint32_t u = (uint32_t)m;
    return (int32_t)u;
}

/* Direction with the same e_x and e_y / e_z flipped where cy / cz is set. */
static int reflect_q(int q, int cy, int cz)
{
    const int ey = cy ? -EY[q] : EY[q], ez = cz ? -EZ[q] : EZ[q];
    for (int r = 0; r < 19; r++)
        if (EX[r] == EX[q] && EY[r] == ey && EZ[r] == ez) return r;
    return q;
}

static void init_tables(void)
{
    W[0] = 1.0 / 3;
    for (int q = 1; q <= 6; q++) W[q] = 1.0 / 18;
    for (int q = 7; q < 19; q++) W[q] = 1.0 / 36;
    SQ18 = 18 * sqrt(2.0);
}

/* _rebuildNear (js/lbm.js:268-282). */
static void rebuild_near(wto_lbm *L)
{
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    memset(L->near, 0, (size_t)L->n);
    for (int z = 1; z < nz - 1; z++) {
        for (int y = 1; y < ny - 1; y++) {
            long i = 1 + (long)nx * (y + (long)ny * z);
            for (int x = 1; x < nx - 1; x++, i++) {
                if (L->solid[i]) continue;
                if (L->wgt && L->wgt[i] > 0) { L->near[i] = 1; continue; }  /* gray cell: every link is gray */
                for (int q = 1; q < 19; q++) {
                    const long j = i - L->off[q];
                    if (L->solid[j] || (L->wgt && L->wgt[j] > 0)) { L->near[i] = 1; break; }
                }
            }
        }
    }
}

#define REAL float
#define SFX _f32
#include "lbm_core.inc"
#undef REAL
#undef SFX
#define REAL double
#define SFX _f64
#include "lbm_core.inc"
#undef REAL
#undef SFX

#define DISPATCH(L, fn, ...) ((L)->prec == 8 ? fn##_f64(__VA_ARGS__) : fn##_f32(__VA_ARGS__))

/* ------------------------------------------------------------------ API */

/* new LBM(nx,ny,nz) with the constructor defaults, then reset()
 * (js/lbm.js:58-128). */
wto_lbm *wto_create(int nx, int ny, int nz, int prec)
{
    init_tables();
    if (nx < 3 || ny < 3 || nz < 3 || (prec != 4 && prec != 8)) return NULL;
    wto_lbm *L = calloc(1, sizeof(*L));
    L->nx = nx; L->ny = ny; L->nz = nz; L->prec = prec;
    L->n = (long)nx * ny * nz;
    const size_t n = (size_t)L->n;
    L->buf[0] = calloc(19 * n, (size_t)prec);
    L->buf[1] = calloc(19 * n, (size_t)prec);
    L->solid = calloc(n, 1);
    L->near = calloc(n, 1);
    L->ux = calloc(n, (size_t)prec); L->uy = calloc(n, (size_t)prec);
    L->uz = calloc(n, (size_t)prec); L->rho = calloc(n, (size_t)prec);
    L->uwx = calloc(n, 4); L->uwy = calloc(n, 4); L->uwz = calloc(n, 4);
    for (int q = 0; q < 19; q++) L->off[q] = EX[q] + (long)nx * (EY[q] + (long)ny * EZ[q]);
    L->u0 = 0.06; L->tau = 0.5035; L->les = 1; L->csm = 0.14; L->trt = 1; L->lambda = 0.25;
    L->frontalArea = 1; L->planformArea = 1;
    DISPATCH(L, reset, L);
    return L;
}

void wto_destroy(wto_lbm *L)
{
    if (!L) return;
    free(L->buf[0]); free(L->buf[1]); free(L->solid); free(L->near);
    free(L->ux); free(L->uy); free(L->uz); free(L->rho);
    free(L->uwx); free(L->uwy); free(L->uwz); free(L->wgt);
    free(L);
}

void wto_set_params(wto_lbm *L, double tau, int les, double csm, int trt, double lambda, double u0,
                    int movingWalls)
{
    L->tau = tau; L->les = les; L->csm = csm; L->trt = trt; L->lambda = lambda; L->u0 = u0;
    L->movingWalls = movingWalls;
}

void wto_set_channel(wto_lbm *L, int on) { L->channel = !!on; }

void wto_get_params(const wto_lbm *L, double out[4], int iout[3])
{
    out[0] = L->tau; out[1] = L->csm; out[2] = L->lambda; out[3] = L->u0;
    iout[0] = L->les; iout[1] = L->trt; iout[2] = L->movingWalls;
}

void wto_reset(wto_lbm *L) { DISPATCH(L, reset, L); }

void wto_voxelize(wto_lbm *L, const int box[6], const uint8_t *mask, const float *wv, int ground,
                  double roadVel, long stats[5])
{
    DISPATCH(L, voxelize, L, box, mask, wv, ground, roadVel);
    stats[0] = L->frontalArea; stats[1] = L->planformArea; stats[2] = L->solidCells;
    stats[3] = L->rideHeight; stats[4] = L->movingWalls;
}

/* step() n times (js/lbm.js:285-291); forces[3*s+k] = force after step s. */
void wto_step(wto_lbm *L, int nsteps, double *forces)
{
    for (int s = 0; s < nsteps; s++) {
        double sf[3];
        DISPATCH(L, step_slab, L, 1, L->nz - 1, sf);
        DISPATCH(L, boundaries, L, 1, 0, 0, 0, 0);
        if (L->channel) DISPATCH(L, channel, L);
        L->parity ^= 1;
        L->force[0] = sf[0]; L->force[1] = sf[1]; L->force[2] = sf[2];
        L->steps++;
        if (forces) { forces[3 * s] = sf[0]; forces[3 * s + 1] = sf[1]; forces[3 * s + 2] = sf[2]; }
    }
}

/* Worker-pool step (js/pool.js:97-125 driving js/lbmWorker.js:59-79): the
 * interior is split into z-slabs as in the worker 'init' handler, each worker
 * runs its slab and its share of the boundaries, one barrier per step, and
 * the per-step force is the sum over workers in worker order. nthreads > 1
 * runs the workers on POSIX threads; 1 runs them one after another, which
 * double buffering makes equivalent. */
typedef struct {
    wto_lbm *L;
    int id, nw, nsteps, z0, z1, bz0, bz1;
    double *per; /* [nsteps][nw][3] */
    pthread_barrier_t *bar;
} pool_arg;

static void pool_one(pool_arg *a, int s)
{
    double sf[3];
    DISPATCH(a->L, step_slab, a->L, a->z0, a->z1, sf);
    DISPATCH(a->L, boundaries, a->L, 0, a->bz0, a->bz1, a->id == 0, a->id == a->nw - 1);
    double *o = a->per + ((long)s * a->nw + a->id) * 3;
    o[0] = sf[0]; o[1] = sf[1]; o[2] = sf[2];
}

static void *pool_thread(void *p)
{
    pool_arg *a = p;
    for (int s = 0; s < a->nsteps; s++) {
        pool_one(a, s);
        int r = pthread_barrier_wait(a->bar);
        if (r == PTHREAD_BARRIER_SERIAL_THREAD) a->L->parity ^= 1;
        pthread_barrier_wait(a->bar);
    }
    return NULL;
}

int wto_step_pool(wto_lbm *L, int nsteps, int nworkers, int threaded, double *forces)
{
    if (nworkers < 1 || nworkers > L->nz - 2) return -1;
    if (L->channel) return -2; /* the pool form keeps the reference boundaries */
    pool_arg *A = calloc((size_t)nworkers, sizeof(pool_arg));
    double *per = calloc((size_t)nsteps * nworkers * 3, sizeof(double));
    const int lo = 1, hi = L->nz - 1, span = hi - lo;
    for (int id = 0; id < nworkers; id++) {
        pool_arg *a = &A[id];
        a->L = L; a->id = id; a->nw = nworkers; a->nsteps = nsteps; a->per = per;
        a->z0 = lo + (int)(((long)span * id) / nworkers);
        a->z1 = lo + (int)(((long)span * (id + 1)) / nworkers);
        a->bz0 = id == 0 ? 0 : a->z0;
        a->bz1 = id == nworkers - 1 ? L->nz : a->z1;
    }
    if (threaded && nworkers > 1) {
        pthread_barrier_t bar;
        pthread_barrier_init(&bar, NULL, (unsigned)nworkers);
        pthread_t *th = calloc((size_t)nworkers, sizeof(pthread_t));
        for (int id = 0; id < nworkers; id++) { A[id].bar = &bar; pthread_create(&th[id], NULL, pool_thread, &A[id]); }
        for (int id = 0; id < nworkers; id++) pthread_join(th[id], NULL);
        pthread_barrier_destroy(&bar);
        free(th);
    } else {
        for (int s = 0; s < nsteps; s++) {
            for (int id = 0; id < nworkers; id++) pool_one(&A[id], s);
            L->parity ^= 1;
        }
    }
    for (int s = 0; s < nsteps; s++) {
        double fx = 0, fy = 0, fz = 0;
        for (int w = 0; w < nworkers; w++) {
            const double *o = per + ((long)s * nworkers + w) * 3;
            fx += o[0]; fy += o[1]; fz += o[2];
        }
        if (forces) { forces[3 * s] = fx; forces[3 * s + 1] = fy; forces[3 * s + 2] = fz; }
        L->force[0] = fx; L->force[1] = fy; L->force[2] = fz;
    }
    L->steps += nsteps;
    free(per); free(A);
    return 0;
}

/* Raw stepSlab / _boundaries / parity access for slab-level tests. */
void wto_step_slab(wto_lbm *L, int z0, int z1, double sf[3]) { DISPATCH(L, step_slab, L, z0, z1, sf); }
void wto_boundaries(wto_lbm *L, int full, int z0, int z1, int doLo, int doHi)
{
    DISPATCH(L, boundaries, L, full, z0, z1, doLo, doHi);
}
void wto_flip(wto_lbm *L) { L->parity ^= 1; L->steps++; }

/* Array views: 0 f, 1 g, 2 solid, 3 near, 4 ux, 5 uy, 6 uz, 7 rho, 8 uwx,
 * 9 uwy, 10 uwz, 11 bufA, 12 bufB. */
void *wto_array(wto_lbm *L, int which)
{
    switch (which) {
    case 0: return L->buf[L->parity];
    case 1: return L->buf[L->parity ^ 1];
    case 2: return L->solid;
    case 3: return L->near;
    case 4: return L->ux;
    case 5: return L->uy;
    case 6: return L->uz;
    case 7: return L->rho;
    case 8: return L->uwx;
    case 9: return L->uwy;
    case 10: return L->uwz;
    case 11: return L->buf[0];
    case 12: return L->buf[1];
    }
    return NULL;
}

void wto_state(const wto_lbm *L, long out[8], double force[3])
{
    out[0] = L->parity; out[1] = L->steps; out[2] = L->frontalArea; out[3] = L->planformArea;
    out[4] = L->solidCells; out[5] = L->rideHeight; out[6] = L->movingWalls; out[7] = L->ground;
    force[0] = L->force[0]; force[1] = L->force[1]; force[2] = L->force[2];
}

void wto_sample(const wto_lbm *L, double x, double y, double z, double out[3])
{
    DISPATCH(L, sample, L, x, y, z, out);
}

/* isSolid (js/lbm.js:541-545). */
int wto_is_solid(const wto_lbm *L, double x, double y, double z)
{
    const int xi = js_toint32(x), yi = js_toint32(y), zi = js_toint32(z);
    if (xi < 0 || yi < 0 || zi < 0 || xi >= L->nx || yi >= L->ny || zi >= L->nz) return 0;
    return L->solid[xi + (long)L->nx * (yi + (long)L->ny * zi)] != 0;
}

/* ============================================================ voxelizer */

#define BUCKETS 48
#define MAX_SAMPLES 4096
#define MAX_HITS 512

typedef struct {
    int nx, ny, nz, ntri;
    double cell, inv, min[3], ext[3], centre[3], bbox[6];
    long solid, disagree;
    double openness;
} wto_voxmeta;

/* bounds over every vertex coordinate present (js/voxel.js:32-42). */
void wto_voxel_bounds(const float *P, long len, double b[6])
{
    double x0 = INFINITY, y0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, z1 = -INFINITY;
    for (long i = 0; i < len; i += 3) {
        const double x = P[i];
        if (x < x0) x0 = x;
        if (x > x1) x1 = x;
        if (i + 1 < len) { const double y = P[i + 1]; if (y < y0) y0 = y; if (y > y1) y1 = y; }
        if (i + 2 < len) { const double z = P[i + 2]; if (z < z0) z0 = z; if (z > z1) z1 = z; }
    }
    b[0] = x0; b[1] = y0; b[2] = z0; b[3] = x1; b[4] = y1; b[5] = z1;
}

static double max3(double a, double b, double c) { double m = a > b ? a : b; return m > c ? m : c; }
static double min3(double a, double b, double c) { double m = a < b ? a : b; return m < c ? m : c; }

/* Grid geometry of build(P,res) (js/voxel.js:48-71). Returns 0, or -1 no
 * triangles, -2 degenerate, -3 grid too large. */
int wto_voxel_plan(const float *P, long len, int res, wto_voxmeta *m)
{
    if (res <= 0) res = 96;
    const long ntri = len / 9;
    memset(m, 0, sizeof(*m));
    if (!ntri) return -1;
    double b[6];
    wto_voxel_bounds(P, len, b);
    const double sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    const double span = max3(sx, sy, sz);
    if (!(span > 0)) return -2;
    const double cell = span / res;
    const int PAD = 2;
    const double dx = ceil(sx / cell) + 2 * PAD, dy = ceil(sy / cell) + 2 * PAD, dz = ceil(sz / cell) + 2 * PAD;
    const double nxd = dx > 3 ? dx : 3, nyd = dy > 3 ? dy : 3, nzd = dz > 3 ? dz : 3;
    if (nxd * nyd * nzd > 40e6) return -3;
    m->nx = (int)nxd; m->ny = (int)nyd; m->nz = (int)nzd; m->ntri = (int)ntri;
    m->cell = cell;
    m->min[0] = b[0] - (nxd * cell - sx) * 0.5;
    m->min[1] = b[1] - (nyd * cell - sy) * 0.5;
    m->min[2] = b[2] - (nzd * cell - sz) * 0.5;
    m->inv = 1 / cell;
    m->ext[0] = sx * 0.5; m->ext[1] = sy * 0.5; m->ext[2] = sz * 0.5;
    m->centre[0] = (b[0] + b[3]) * 0.5; m->centre[1] = (b[1] + b[4]) * 0.5; m->centre[2] = (b[2] + b[5]) * 0.5;
    memcpy(m->bbox, b, sizeof(b));
    return 0;
}

/* rasteriseSurface (js/voxel.js:105-137). */
void wto_voxel_shell(const float *P, const wto_voxmeta *m, uint8_t *shell)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const double inv = 1 / m->cell;
    for (long t = 0; t < m->ntri; t++) {
        const float *o = P + t * 9;
        const double ax = o[0], ay = o[1], az = o[2], bx = o[3], by = o[4], bz = o[5];
        const double cx = o[6], cy = o[7], cz = o[8];
        const double e1 = max3(fabs(bx - ax), fabs(by - ay), fabs(bz - az));
        const double e2 = max3(fabs(cx - ax), fabs(cy - ay), fabs(cz - az));
        double n1 = ceil(e1 * inv * 2) + 1, n2 = ceil(e2 * inv * 2) + 1;
        if (n1 * n2 > MAX_SAMPLES) {
            const double k = sqrt(MAX_SAMPLES / (n1 * n2));
            const double a1 = js_toint32(n1 * k), a2 = js_toint32(n2 * k);
            n1 = a1 > 2 ? a1 : 2; n2 = a2 > 2 ? a2 : 2;
        }
        for (double i = 0; i <= n1; i++) {
            const double u = i / n1;
            const double jmax = ceil((1 - u) * n2);
            for (double j = 0; j <= jmax; j++) {
                const double jv = j / n2, ru = 1 - u;
                const double v = jv < ru ? jv : ru;
                const double x = ax + (bx - ax) * u + (cx - ax) * v;
                const double y = ay + (by - ay) * u + (cy - ay) * v;
                const double z = az + (bz - az) * u + (cz - az) * v;
                const int gx = js_toint32((x - m->min[0]) * inv);
                const int gy = js_toint32((y - m->min[1]) * inv);
                const int gz = js_toint32((z - m->min[2]) * inv);
                if (gx < 0 || gy < 0 || gz < 0 || gx >= nx || gy >= ny || gz >= nz) continue;
                shell[gx + (long)nx * (gy + (long)ny * gz)] = 1;
            }
        }
    }
}

/* Numeric ascending order with NaN last, as TypedArray.prototype.sort. */
static int cmp_hit(const void *pa, const void *pb)
{
    const double a = *(const double *)pa, b = *(const double *)pb;
    const int na = a != a, nb = b != b;
    if (na || nb) return na - nb;
    if (a < b) return -1;
    if (a > b) return 1;
    if (a == 0 && b == 0) return (signbit(b) != 0) - (signbit(a) != 0);
    return 0;
}

/* parityFill(axis) (js/voxel.js:144-232): bucketed rays along `axis`,
 * sorted crossings, votes between alternate pairs. */
void wto_voxel_parity(const float *P, const wto_voxmeta *m, int axis, uint8_t *votes)
{
    const int A = axis, B = (axis + 1) % 3, C = (axis + 2) % 3;
    const int dim[3] = {m->nx, m->ny, m->nz};
    const int nA = dim[A], nB = dim[B], nC = dim[C];
    const double cell = m->cell, minA = m->min[A], minB = m->min[B], minC = m->min[C];
    const double spanB = nB * cell, spanC = nC * cell;
    const int nbB = nB < BUCKETS ? nB : BUCKETS, nbC = nC < BUCKETS ? nC : BUCKETS;
    const double kB = nbB / spanB, kC = nbC / spanC;
    const long nb = (long)nbB * nbC, ntri = m->ntri;
    int32_t *count = calloc((size_t)nb + 1, sizeof(int32_t));
    int32_t *lo = malloc(sizeof(int32_t) * 2 * (size_t)ntri), *hi = malloc(sizeof(int32_t) * 2 * (size_t)ntri);
    for (long t = 0; t < ntri; t++) {
        const float *o = P + t * 9;
        const double b0 = o[B], b1 = o[3 + B], b2 = o[6 + B], c0 = o[C], c1 = o[3 + C], c2 = o[6 + C];
        double bl = floor((min3(b0, b1, b2) - minB) * kB), bh = floor((max3(b0, b1, b2) - minB) * kB);
        double cl = floor((min3(c0, c1, c2) - minC) * kC), ch = floor((max3(c0, c1, c2) - minC) * kC);
        bl = bl < 0 ? 0 : bl; bh = bh >= nbB ? nbB - 1 : bh;
        cl = cl < 0 ? 0 : cl; ch = ch >= nbC ? nbC - 1 : ch;
        /* Int32Array stores apply ToInt32 */
        lo[2 * t] = js_toint32(bl); hi[2 * t] = js_toint32(bh);
        lo[2 * t + 1] = js_toint32(cl); hi[2 * t + 1] = js_toint32(ch);
        if (bl > bh || cl > ch) continue;
        if (bl != bl || bh != bh || cl != cl || ch != ch) continue; /* loops below would not run */
        for (int cc = (int)cl; cc <= (int)ch; cc++)
            for (int bb = (int)bl; bb <= (int)bh; bb++) count[bb + (long)nbB * cc + 1]++;
    }
    for (long i = 0; i < nb; i++) count[i + 1] += count[i];
    int32_t *items = malloc(sizeof(int32_t) * ((size_t)count[nb] + 1));
    int32_t *cursor = malloc(sizeof(int32_t) * (size_t)nb);
    memcpy(cursor, count, sizeof(int32_t) * (size_t)nb);
    for (long t = 0; t < ntri; t++) {
        const int bl = lo[2 * t], bh = hi[2 * t], cl = lo[2 * t + 1], ch = hi[2 * t + 1];
        if (bl > bh || cl > ch) continue;
        for (int cc = cl; cc <= ch; cc++)
            for (int bb = bl; bb <= bh; bb++) items[cursor[bb + (long)nbB * cc]++] = (int32_t)t;
    }
    const long sA = A == 0 ? 1 : (A == 1 ? dim[0] : (long)dim[0] * dim[1]);
    const long sB = B == 0 ? 1 : (B == 1 ? dim[0] : (long)dim[0] * dim[1]);
    const long sC = C == 0 ? 1 : (C == 1 ? dim[0] : (long)dim[0] * dim[1]);
    double hits[MAX_HITS];
    for (int ic = 0; ic < nC; ic++) {
        const double pc = minC + (ic + 0.5003719) * cell;
        double bcd = floor((pc - minC) * kC);
        bcd = bcd > 0 ? bcd : 0; bcd = bcd < nbC - 1 ? bcd : nbC - 1;
        const int bc = (int)bcd;
        for (int ib = 0; ib < nB; ib++) {
            const double pb = minB + (ib + 0.4996281) * cell;
            double bbd = floor((pb - minB) * kB);
            bbd = bbd > 0 ? bbd : 0; bbd = bbd < nbB - 1 ? bbd : nbB - 1;
            const int bb = (int)bbd;
            const int s = count[bb + (long)nbB * bc], e = count[bb + (long)nbB * bc + 1];
            if (s == e) continue;
            int nh = 0;
            for (int k = s; k < e && nh < MAX_HITS; k++) {
                const float *o = P + (long)items[k] * 9;
                const double b0 = o[B], c0 = o[C], b1 = o[3 + B], c1 = o[3 + C], b2 = o[6 + B], c2 = o[6 + C];
                const double d = (c1 - c2) * (b0 - b2) + (b2 - b1) * (c0 - c2);
                if (d == 0) continue;
                const double id = 1 / d;
                const double l0 = ((c1 - c2) * (pb - b2) + (b2 - b1) * (pc - c2)) * id;
                if (l0 < 0 || l0 > 1) continue;
                const double l1 = ((c2 - c0) * (pb - b2) + (b0 - b2) * (pc - c2)) * id;
                if (l1 < 0 || l1 > 1) continue;
                const double l2 = 1 - l0 - l1;
                if (l2 < 0 || l2 > 1) continue;
                hits[nh++] = l0 * (double)o[A] + l1 * (double)o[3 + A] + l2 * (double)o[6 + A];
            }
            if (nh < 2) continue;
            qsort(hits, (size_t)nh, sizeof(double), cmp_hit);
            const long colBase = ib * sB + ic * sC;
            for (int k = 0; k + 1 < nh; k += 2) {
                double i0 = ceil((hits[k] - minA) / cell - 0.5);
                double i1 = floor((hits[k + 1] - minA) / cell - 0.5);
                if (i0 < 0) i0 = 0;
                if (i1 >= nA) i1 = nA - 1;
                if (!(i0 <= i1)) continue;
                for (long ia = (long)i0; ia <= (long)i1; ia++) votes[colBase + ia * sA]++;
            }
        }
    }
    free(count); free(lo); free(hi); free(items); free(cursor);
}

/* build(P,res) (js/voxel.js:48-102): grid must hold nx*ny*nz bytes from
 * wto_voxel_plan. Returns the plan's status. */
int wto_voxel_build(const float *P, long len, int res, wto_voxmeta *m, uint8_t *grid)
{
    int rc = wto_voxel_plan(P, len, res, m);
    if (rc) return rc;
    const long n = (long)m->nx * m->ny * m->nz;
    uint8_t *votes = calloc((size_t)n, 1), *shell = calloc((size_t)n, 1);
    wto_voxel_shell(P, m, shell);
    for (int axis = 0; axis < 3; axis++) wto_voxel_parity(P, m, axis, votes);
    long solid = 0, disagree = 0;
    for (long i = 0; i < n; i++) {
        const int v = votes[i];
        const int inside = v >= 2 || shell[i] == 1;
        if (inside) solid++;
        if (v == 1 || v == 2) disagree++;
        grid[i] = inside ? 1 : 0;
    }
    m->solid = solid; m->disagree = disagree;
    m->openness = solid ? (double)disagree / (double)solid : 0;
    free(votes); free(shell);
    return 0;
}

/* Grid-backed Body.makeTest evaluated over a clamped lattice box
 * (js/shapes.js:369-388): world -> body by M^T and 1/scale, then an
 * integer-truncated voxel lookup. mask receives 1 where inside. */
void wto_grid_mask(const uint8_t *grid, const int gdim[3], const double gmin[3], double ginv,
                   const double M[9], const double pos[3], double scale, const int box[6], uint8_t *mask)
{
    const double inv = 1 / scale;
    long k = 0;
    for (int z = box[2]; z <= box[5]; z++)
        for (int y = box[1]; y <= box[4]; y++)
            for (int x = box[0]; x <= box[3]; x++, k++) {
                const double X = x + 0.5, Y = y + 0.5, Z = z + 0.5;
                const double dx = X - pos[0], dy = Y - pos[1], dz = Z - pos[2];
                const double lx = (M[0] * dx + M[3] * dy + M[6] * dz) * inv;
                const double ly = (M[1] * dx + M[4] * dy + M[7] * dz) * inv;
                const double lz = (M[2] * dx + M[5] * dy + M[8] * dz) * inv;
                mask[k] = 0;
                const int a = js_toint32((lx - gmin[0]) * ginv);
                if (a < 0 || a >= gdim[0]) continue;
                const int b = js_toint32((ly - gmin[1]) * ginv);
                if (b < 0 || b >= gdim[1]) continue;
                const int c = js_toint32((lz - gmin[2]) * ginv);
                if (c < 0 || c >= gdim[2]) continue;
                mask[k] = grid[a + (long)gdim[0] * (b + (long)gdim[1] * c)] != 0;
            }
}

/* ===================================================== gray bounce-back
 *
 * No reference counterpart (an addition of this library). A cell with weight
 * w in (0,1) stays a fluid cell and every link (i, j = i - c_q) between two
 * fluid cells transmits the fraction t = (1 - w_i)(1 - w_j) of what streams
 * along it; the rest is bounced back halfway, as at a wall:
 *   f_q(i) = t g_q(j) + (1 - t) g_qbar(i),   F -= (1 - t) 2 g_qbar(i) c_q.
 * Cell j takes the same t for the opposite link, so each pair of links
 * conserves mass exactly, and F is exactly the momentum the fluid loses
 * (partial bounce-back on links; the transmission of a link is the product
 * of what its two end cells let through). w = 0 on both ends is plain
 * streaming and w = 1 is a SOLID_OBJ cell (selected, not blended), so 0/1
 * weights reduce exactly to the reference. In a uniform gray medium the
 * momentum of a uniform flow decays by the factor 2t - 1 per step. */

/* Weights per cell; 0 or NaN = fluid, (0,1) = gray. Cells that are solid keep
 * their solid type; a weight >= 1 on a fluid cell is rejected (voxelize it
 * as body first). NULL clears the gray field. */
int wto_set_gray(wto_lbm *L, const float *w)
{
    if (!w) { free(L->wgt); L->wgt = NULL; rebuild_near(L); return 0; }
    if (!L->wgt) L->wgt = malloc(sizeof(float) * (size_t)L->n);
    for (long i = 0; i < L->n; i++) {
        float v = w[i];
        const int x = (int)(i % L->nx), y = (int)((i / L->nx) % L->ny), z = (int)(i / ((long)L->nx * L->ny));
        const int bnd = x == 0 || y == 0 || z == 0 || x == L->nx - 1 || y == L->ny - 1 || z == L->nz - 1;
        if (L->solid[i] || bnd || !(v > 0)) v = 0;   /* boundary planes carry no weight */
        if (v >= 1) return -1;
        L->wgt[i] = v;
    }
    rebuild_near(L);
    return 0;
}

/* ===================================================== discrete adjoint
 *
 * FP64 reference of the adjoint the GPU runs: reverse mode through N steps
 * of the gray step (no moving walls), all intermediate states stored. The
 * objective is J = sum_k Fbar[k] . F_k with F_k the force of step k. */

static int computed(const wto_lbm *L, long i)
{
    const int x = (int)(i % L->nx), y = (int)((i / L->nx) % L->ny), z = (int)(i / ((long)L->nx * L->ny));
    return x >= 1 && x <= L->nx - 2 && y >= 1 && y <= L->ny - 2 && z >= 1 && z <= L->nz - 2 && !L->solid[i];
}

/* gray weight of cell m (0 off the gray cells) */
static double wcell(const wto_lbm *L, long m)
{
    return L->wgt ? (double)L->wgt[m] : 0;
}

/* the link (i <- j) is gray: neither end solid, one of them gray (the rest
 * population, j = i, is no link) */
static int gray_link(const wto_lbm *L, long i, long j)
{
    return i != j && !L->solid[j] && L->wgt && (L->wgt[i] > 0 || L->wgt[j] > 0);
}

/* transmission of the link (i <- j) */
static double trans(const wto_lbm *L, long i, long j)
{
    return (1 - wcell(L, i)) * (1 - wcell(L, j));
}

/* pre-collision populations of computed cell i from state S (double storage) */
static void gather_f64(const wto_lbm *L, const double *S, long i, double f[19])
{
    const long n = L->n;
    for (int q = 0; q < 19; q++) {
        const long j = i - L->off[q];
        if (L->solid[j]) f[q] = S[OPP[q] * n + i];
        else if (gray_link(L, i, j)) {
            const double t = trans(L, i, j);
            f[q] = t * S[q * n + j] + (1 - t) * S[OPP[q] * n + i];
        } else f[q] = S[q * n + j];
    }
}

/* mu = (d collide / d f)^T lam at f, for the collision of js/lbm.js:377-460
 * written in its generic form e_q = W_q (A + 3 r cu + 4.5 r cu^2). */
void wto_collide_adj(const double f[19], const double lam[19], double tau0, double csm, double lamb, int les,
                     int trt, double mu[19])
{
    double dr = 0, jx = 0, jy = 0, jz = 0;
    for (int q = 0; q < 19; q++) { dr += f[q]; jx += EX[q] * f[q]; jy += EY[q] * f[q]; jz += EZ[q] * f[q]; }
    const double r = 1 + dr, inv = 1 / r;
    const double vx = jx * inv, vy = jy * inv, vz = jz * inv;
    const double usq = vx * vx + vy * vy + vz * vz;
    const double A = dr - 1.5 * r * usq;
    double cu[19], e[19], nn[19];
    for (int q = 0; q < 19; q++) {
        cu[q] = EX[q] * vx + EY[q] * vy + EZ[q] * vz;
        e[q] = W[q] * (A + 3 * r * cu[q] + 4.5 * r * cu[q] * cu[q]);
        nn[q] = f[q] - e[q];
    }
    double pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
    for (int q = 0; q < 19; q++) {
        pxx += EX[q] * EX[q] * nn[q]; pyy += EY[q] * EY[q] * nn[q]; pzz += EZ[q] * EZ[q] * nn[q];
        pxy += EX[q] * EY[q] * nn[q]; pxz += EX[q] * EZ[q] * nn[q]; pyz += EY[q] * EZ[q] * nn[q];
    }
    const double X = pxx * pxx + pyy * pyy + pzz * pzz + 2 * (pxy * pxy + pxz * pxz + pyz * pyz);
    const double mag = sqrt(2 * X);
    const double K = SQ18 * csm * csm;
    const double Sq = sqrt(tau0 * tau0 + K * mag * inv);
    const double te = les ? 0.5 * (tau0 + Sq) : tau0;
    const double om = 1 / te;
    const double omm = trt ? 1 / (0.5 + lamb / (te - 0.5)) : om;

    /* g = f - om P+ n - omm P- n */
    double nb[19], omb = 0, ommb = 0;
    for (int q = 0; q < 19; q++) mu[q] = lam[q];
    nb[0] = -om * lam[0];
    omb -= nn[0] * lam[0];
    for (int a = 1; a < 19; a += 2) {
        const int b = a + 1;
        const double sp = 0.5 * (nn[a] + nn[b]), sm = 0.5 * (nn[a] - nn[b]);
        const double spb = -om * (lam[a] + lam[b]), smb = -omm * (lam[a] - lam[b]);
        nb[a] = 0.5 * (spb + smb);
        nb[b] = 0.5 * (spb - smb);
        omb -= sp * (lam[a] + lam[b]);
        ommb -= sm * (lam[a] - lam[b]);
    }
    double teb = 0;
    if (trt) teb += ommb * omm * omm * lamb / ((te - 0.5) * (te - 0.5));
    else omb += ommb;
    double invb = 0;
    if (les) {
        teb += -omb / (te * te);
        const double magb = teb * K * inv / (4 * Sq);
        invb += teb * K * mag / (4 * Sq);
        const double Xb = mag > 0 ? magb / mag : 0;
        const double pxxb = 2 * Xb * pxx, pyyb = 2 * Xb * pyy, pzzb = 2 * Xb * pzz;
        const double pxyb = 4 * Xb * pxy, pxzb = 4 * Xb * pxz, pyzb = 4 * Xb * pyz;
        for (int q = 0; q < 19; q++)
            nb[q] += pxxb * EX[q] * EX[q] + pyyb * EY[q] * EY[q] + pzzb * EZ[q] * EZ[q] +
                     pxyb * EX[q] * EY[q] + pxzb * EX[q] * EZ[q] + pyzb * EY[q] * EZ[q];
    }
    /* n = f - e(rho, u) */
    double Ab = 0, rb = 0, vxb = 0, vyb = 0, vzb = 0;
    for (int q = 0; q < 19; q++) {
        mu[q] += nb[q];
        const double eb = -nb[q];
        Ab += W[q] * eb;
        rb += W[q] * eb * (3 * cu[q] + 4.5 * cu[q] * cu[q]);
        const double cub = W[q] * eb * (3 * r + 9 * r * cu[q]);
        vxb += EX[q] * cub; vyb += EY[q] * cub; vzb += EZ[q] * cub;
    }
    double drb = Ab;
    rb += -1.5 * usq * Ab;
    const double usqb = -1.5 * r * Ab;
    vxb += 2 * vx * usqb; vyb += 2 * vy * usqb; vzb += 2 * vz * usqb;
    const double jxb = vxb * inv, jyb = vyb * inv, jzb = vzb * inv;
    invb += vxb * jx + vyb * jy + vzb * jz;
    rb += -inv * inv * invb;
    drb += rb;
    for (int q = 0; q < 19; q++) mu[q] += drb + EX[q] * jxb + EY[q] * jyb + EZ[q] * jzb;
}

/* Run N steps from the current state and return dJ/dw for every gray cell
 * (0 elsewhere). Fbar[3k..3k+2] = dJ/dF_k. forces (may be NULL) gets F_k.
 * FP64 storage only; moving walls unsupported.
 *
 * mu = dJ/d(pre-collision populations of step k), lam = dJ/d(state s_k that
 * step k pulls from). A gray link (i <- j) along q contributes to dJ/dW of
 * its bounced fraction W = 1 - t
 *   G = mu_q(i) (s_qbar(i) - s_q(j)) - 2 s_qbar(i) c_q . Fbar_k,
 * and dW/dw_i = 1 - w_j, dW/dw_j = 1 - w_i: a gray cell collects G of the 18
 * links it receives on and of the 18 links it sends on. */
int wto_adjoint(wto_lbm *L, int N, const double *Fbar, double *dJdw, double *forces)
{
    if (L->prec != 8 || L->movingWalls || N < 1) return -1;
    const long n = L->n, sz = 19 * n;
    double *S = malloc(sizeof(double) * (size_t)sz * (size_t)N);
    if (!S) return -2;
    for (int k = 0; k < N; k++) {
        memcpy(S + (long)k * sz, L->buf[L->parity], sizeof(double) * (size_t)sz);
        wto_step(L, 1, forces ? forces + 3 * k : NULL);
    }
    double *mu = calloc((size_t)sz, sizeof(double)), *lam = calloc((size_t)sz, sizeof(double));
    double *mun = calloc((size_t)sz, sizeof(double));
    memset(dJdw, 0, sizeof(double) * (size_t)n);
    for (int k = N - 1; k >= 0; k--) {
        const double *Sk = S + (long)k * sz, *Fb = Fbar + 3 * k;
        for (long m = 0; m < n; m++) {
            const int cm = computed(L, m);
            for (int q = 0; q < 19; q++) {
                double v = 0;
                const long n1 = m + L->off[q];
                const int in = n1 >= 0 && n1 < n;
                /* s_q(m) streams into n1 */
                if (in && computed(L, n1) && !L->solid[m])
                    v += (gray_link(L, n1, m) ? trans(L, n1, m) : 1) * mu[q * n + n1];
                /* s_q(m) is bounced back at m: m pulls along qbar from j = n1 */
                if (cm) {
                    const long j = n1;
                    const int qb = OPP[q];
                    const double cF = EX[qb] * Fb[0] + EY[qb] * Fb[1] + EZ[qb] * Fb[2];
                    if (L->solid[j]) v += mu[qb * n + m] - (L->solid[j] == SOLID_OBJ ? 2 * cF : 0);
                    else if (gray_link(L, m, j)) v += (1 - trans(L, m, j)) * (mu[qb * n + m] - 2 * cF);
                }
                lam[q * n + m] = v;
            }
        }
        if (L->wgt) {
            for (long g = 0; g < n; g++) {
                if (!(L->wgt[g] > 0) || L->solid[g]) continue;
                double d = 0;
                for (int q = 0; q < 19; q++) {
                    const double cF = EX[q] * Fb[0] + EY[q] * Fb[1] + EZ[q] * Fb[2];
                    if (q == 0) continue;
                    /* g receives along q from j */
                    const long j = g - L->off[q];
                    if (!L->solid[j]) {
                        const double fd = Sk[OPP[q] * n + g];
                        d += (1 - wcell(L, j)) * (mu[q * n + g] * (fd - Sk[q * n + j]) - 2 * fd * cF);
                    }
                    /* n1 receives along q from g */
                    const long n1 = g + L->off[q];
                    if (n1 >= 0 && n1 < n && computed(L, n1)) {
                        const double fd = Sk[OPP[q] * n + n1];
                        d += (1 - wcell(L, n1)) * (mu[q * n + n1] * (fd - Sk[q * n + g]) - 2 * fd * cF);
                    }
                }
                dJdw[g] += d;
            }
        }
        if (k == 0) break;
        const double *Sp = S + (long)(k - 1) * sz;
        for (long i = 0; i < n; i++) {
            if (!computed(L, i)) { for (int q = 0; q < 19; q++) mun[q * n + i] = 0; continue; }
            double f[19], lv[19], mv[19];
            gather_f64(L, Sp, i, f);
            const int x = (int)(i % L->nx);
            for (int q = 0; q < 19; q++) lv[q] = lam[q * n + i] + (x == L->nx - 2 ? lam[q * n + i + 1] : 0);
            wto_collide_adj(f, lv, L->tau, L->csm, L->lambda, L->les, L->trt, mv);
            for (int q = 0; q < 19; q++) mun[q * n + i] = mv[q];
        }
        double *t = mu; mu = mun; mun = t;
    }
    free(S); free(mu); free(lam); free(mun);
    return 0;
}
