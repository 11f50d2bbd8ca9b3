/* wt_napi.c -- Node.js N-API addon over the libwindtunnel C ABI.
 *
 * A thin 1:1 binding: every export forwards to one wt_* entry point of
 * include/windtunnel.h, passing typed-array storage as plain pointers. The
 * JavaScript side (windtunnel_amd/js/lbm.js) rebuilds the reference's LBM
 * object API, the lbmWorker.js message protocol and Voxel.build on top.
 * Failures throw a JavaScript Error carrying wt_last_error(); voxelizer
 * errors therefore carry the reference's own messages (js/voxel.js:51,56,71).
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/windtunnel.h"

#define CALL(env, x)                                                      \
    do {                                                                  \
        if ((x) != napi_ok) {                                             \
            napi_throw_error((env), NULL, "N-API call failed: " #x);      \
            return NULL;                                                  \
        }                                                                 \
    } while (0)

#define WT(env, x)                                                        \
    do {                                                                  \
        int rc_ = (x);                                                    \
        if (rc_ != 0) {                                                   \
            char code_[16];                                               \
            snprintf(code_, sizeof code_, "WT%d", rc_);                   \
            napi_throw_error((env), code_, wt_last_error());              \
            return NULL;                                                  \
        }                                                                 \
    } while (0)

static napi_value args(napi_env env, napi_callback_info info, size_t want, napi_value *argv)
{
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return NULL;
    for (size_t k = argc; k < want; k++) napi_get_undefined(env, &argv[k]);
    return argv[0];
}

static int is_nullish(napi_env env, napi_value v)
{
    napi_valuetype t;
    napi_typeof(env, v, &t);
    return t == napi_undefined || t == napi_null;
}

static wt_lattice *handle(napi_env env, napi_value v)
{
    void *p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, NULL, "expected a lattice handle");
        return NULL;
    }
    return (wt_lattice *)p;
}

/* typed-array data pointer and byte length (NULL for null/undefined) */
static void *tdata(napi_env env, napi_value v, size_t *bytes, int *ok)
{
    *ok = 1;
    if (bytes) *bytes = 0;
    if (is_nullish(env, v)) return NULL;
    napi_typedarray_type type;
    size_t len = 0, off = 0;
    void *data = NULL;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &type, &len, &data, &ab, &off) != napi_ok) {
        napi_throw_type_error(env, NULL, "expected a typed array");
        *ok = 0;
        return NULL;
    }
    size_t esz = 1;
    switch (type) {
    case napi_int16_array: case napi_uint16_array: esz = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: esz = 4; break;
    case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: esz = 8; break;
    default: esz = 1;
    }
    if (bytes) *bytes = len * esz;
    return data;
}

static double num(napi_env env, napi_value v, double dflt)
{
    double d;
    if (napi_get_value_double(env, v, &d) != napi_ok) return dflt;
    return d;
}

static int boolish(napi_env env, napi_value v)
{
    napi_value b;
    bool r = false;
    if (napi_coerce_to_bool(env, v, &b) == napi_ok) napi_get_value_bool(env, b, &r);
    return r ? 1 : 0;
}

static double prop_num(napi_env env, napi_value o, const char *k, double dflt)
{
    napi_value v;
    bool has = false;
    if (napi_has_named_property(env, o, k, &has) != napi_ok || !has) return dflt;
    napi_get_named_property(env, o, k, &v);
    return num(env, v, dflt);
}

static void set_num(napi_env env, napi_value o, const char *k, double v)
{
    napi_value x;
    napi_create_double(env, v, &x);
    napi_set_named_property(env, o, k, x);
}

static void set_bool(napi_env env, napi_value o, const char *k, int v)
{
    napi_value x;
    napi_get_boolean(env, v != 0, &x);
    napi_set_named_property(env, o, k, x);
}

static void finalize(napi_env env, void *data, void *hint)
{
    (void)env; (void)hint;
    wt_destroy((wt_lattice *)data);
}

/* Lattice geometry for argument checks: cells, and bytes per field value
 * and per population value as wt_get_fields / wt_get_pops write them. */
static int lattice_info(napi_env env, wt_lattice *h, size_t *cells, size_t *fbytes)
{
    wt_state_t st;
    if (wt_get_state(h, &st) != 0) {
        napi_throw_error(env, NULL, wt_last_error());
        return 0;
    }
    *cells = (size_t)st.nx * (size_t)st.ny * (size_t)st.nz;
    *fbytes = st.storage == WT_STORE_F64 ? 8 : 4;
    return 1;
}

/* true when a typed array of `have` bytes holds at least `need` bytes; throws otherwise */
static int enough(napi_env env, size_t have, size_t need, const char *what)
{
    if (have >= need) return 1;
    char msg[160];
    snprintf(msg, sizeof msg, "%s: typed array of %zu bytes, %zu needed", what, have, need);
    napi_throw_range_error(env, NULL, msg);
    return 0;
}

/* create(nx, ny, nz, storage, arith, layout) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info)
{
    napi_value a[6], out;
    args(env, info, 6, a);
    wt_lattice *h = NULL;
    WT(env, wt_create((int)num(env, a[0], 0), (int)num(env, a[1], 0), (int)num(env, a[2], 0),
                      (int)num(env, a[3], 0), (int)num(env, a[4], 0), (int)num(env, a[5], 0), &h));
    CALL(env, napi_create_external(env, h, finalize, NULL, &out));
    return out;
}

static napi_value js_set_params(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    wt_params p;
    WT(env, wt_get_params(h, &p));
    p.tau = prop_num(env, a[1], "tau", p.tau);
    p.csm = prop_num(env, a[1], "csm", p.csm);
    p.lambda = prop_num(env, a[1], "lambda", p.lambda);
    p.u0 = prop_num(env, a[1], "u0", p.u0);
    p.les = (int)prop_num(env, a[1], "les", p.les) != 0;
    p.trt = (int)prop_num(env, a[1], "trt", p.trt) != 0;
    p.movingWalls = (int)prop_num(env, a[1], "movingWalls", p.movingWalls) != 0;
    WT(env, wt_set_params(h, &p));
    return NULL;
}

static napi_value js_get_params(napi_env env, napi_callback_info info)
{
    napi_value a[1], o;
    args(env, info, 1, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    wt_params p;
    WT(env, wt_get_params(h, &p));
    CALL(env, napi_create_object(env, &o));
    set_num(env, o, "tau", p.tau); set_num(env, o, "csm", p.csm); set_num(env, o, "lambda", p.lambda);
    set_num(env, o, "u0", p.u0); set_bool(env, o, "les", p.les); set_bool(env, o, "trt", p.trt);
    set_bool(env, o, "movingWalls", p.movingWalls);
    return o;
}

static napi_value js_reset(napi_env env, napi_callback_info info)
{
    napi_value a[1];
    args(env, info, 1, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    WT(env, wt_reset(h));
    return NULL;
}

static napi_value stats_object(napi_env env, const wt_body_stats *s)
{
    napi_value o;
    if (napi_create_object(env, &o) != napi_ok) return NULL;
    set_num(env, o, "frontalArea", (double)s->frontalArea);
    set_num(env, o, "planformArea", (double)s->planformArea);
    set_num(env, o, "solidCells", (double)s->solidCells);
    set_num(env, o, "rideHeight", (double)s->rideHeight);
    set_bool(env, o, "movingWalls", s->movingWalls);
    set_bool(env, o, "ground", s->ground);
    return o;
}

/* voxelize(h, box Int32Array(6), mask Uint8Array, wv Float32Array|null, ground, roadVel) */
static napi_value js_voxelize(napi_env env, napi_callback_info info)
{
    napi_value a[6];
    args(env, info, 6, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const int *box = tdata(env, a[1], &nb, &ok);
    if (!ok) return NULL;
    if (!box || nb < 24) { napi_throw_type_error(env, NULL, "box must be an Int32Array(6)"); return NULL; }
    size_t mb = 0, wb = 0;
    const uint8_t *mask = tdata(env, a[2], &mb, &ok);
    if (!ok) return NULL;
    const float *wv = tdata(env, a[3], &wb, &ok);
    if (!ok) return NULL;
    if (box[3] >= box[0] && box[4] >= box[1] && box[5] >= box[2]) {
        const size_t cnt = (size_t)(box[3] - box[0] + 1) * (size_t)(box[4] - box[1] + 1) * (size_t)(box[5] - box[2] + 1);
        if (mask && !enough(env, mb, cnt, "voxelize mask")) return NULL;
        size_t moving = 0;
        if (mask)
            for (size_t k = 0; k < cnt; k++) moving += mask[k] == 2;
        if (moving && !enough(env, wv ? wb : 0, moving * 12, "voxelize wall velocities")) return NULL;
    }
    wt_body_stats s;
    WT(env, wt_voxelize(h, box, mask, wv, boolish(env, a[4]), num(env, a[5], 0), &s));
    return stats_object(env, &s);
}

/* voxelizeGrid(h, grid Uint8Array, meta {nx,ny,nz,min,inv,cell}, M Float64Array(9),
 *              pos Float64Array(3), scale, box Int32Array(6), ground, roadVel) */
static napi_value js_voxelize_grid(napi_env env, napi_callback_info info)
{
    napi_value a[9];
    args(env, info, 9, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t gb = 0;
    const uint8_t *grid = tdata(env, a[1], &gb, &ok);
    if (!ok) return NULL;
    if (!grid) gb = 0;
    wt_voxmeta vm;
    memset(&vm, 0, sizeof vm);
    vm.nx = (int)prop_num(env, a[2], "nx", 0);
    vm.ny = (int)prop_num(env, a[2], "ny", 0);
    vm.nz = (int)prop_num(env, a[2], "nz", 0);
    vm.inv = prop_num(env, a[2], "inv", 0);
    vm.cell = prop_num(env, a[2], "cell", 0);
    napi_value mn;
    if (napi_get_named_property(env, a[2], "min", &mn) == napi_ok) {
        for (uint32_t k = 0; k < 3; k++) {
            napi_value e;
            if (napi_get_element(env, mn, k, &e) == napi_ok) vm.min[k] = num(env, e, 0);
        }
    }
    size_t Mb = 0, pb = 0, bb = 0;
    const double *M = tdata(env, a[3], &Mb, &ok);
    if (!ok) return NULL;
    const double *pos = tdata(env, a[4], &pb, &ok);
    if (!ok) return NULL;
    const int *box = tdata(env, a[6], &bb, &ok);
    if (!ok) return NULL;
    if (!enough(env, gb, (size_t)(vm.nx > 0 ? vm.nx : 0) * (size_t)(vm.ny > 0 ? vm.ny : 0) *
                             (size_t)(vm.nz > 0 ? vm.nz : 0), "voxel grid") ||
        !enough(env, M ? Mb : 0, 72, "rotation M") || !enough(env, pos ? pb : 0, 24, "position") ||
        !enough(env, box ? bb : 0, 24, "box"))
        return NULL;
    wt_body_stats s;
    WT(env, wt_voxelize_grid(h, grid, &vm, M, pos, num(env, a[5], 1), box, boolish(env, a[7]), num(env, a[8], 0), &s));
    return stats_object(env, &s);
}

/* step(h, n, forces Float64Array|null) */
static napi_value js_step(napi_env env, napi_callback_info info)
{
    napi_value a[3];
    args(env, info, 3, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const int n = (int)num(env, a[1], 1);
    double *f = tdata(env, a[2], &nb, &ok);
    if (!ok) return NULL;
    if (f && nb < (size_t)n * 24) { napi_throw_range_error(env, NULL, "forces array too short"); return NULL; }
    WT(env, wt_step(h, n, f));
    return NULL;
}

/* stepSlabs(h, n, nslabs, forces Float64Array|null) */
static napi_value js_step_slabs(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const int n = (int)num(env, a[1], 1), ns = (int)num(env, a[2], 1);
    double *f = tdata(env, a[3], &nb, &ok);
    if (!ok) return NULL;
    if (f && nb < (size_t)n * ns * 24) { napi_throw_range_error(env, NULL, "forces array too short"); return NULL; }
    WT(env, wt_step_slabs(h, n, ns, f));
    return NULL;
}

/* stepSlabsAsync(h, n, nslabs, forces) -> Promise: wt_step_slabs on a libuv
 * worker thread, so the event loop keeps running while the device steps
 * (as the reference's worker threads leave the page free, js/pool.js:97-125).
 * The handle and the forces array are held until the batch ends; the
 * lattice must not be used otherwise meanwhile (the pool serialises its
 * batches). The promise rejects with the library's message on failure. */
typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    napi_ref href, fref;
    wt_lattice *h;
    int n, ns, rc;
    double *f;
    char msg[512];
} step_job;

static void step_job_run(napi_env env, void *data)
{
    (void)env;
    step_job *j = (step_job *)data;
    j->rc = wt_step_slabs(j->h, j->n, j->ns, j->f);
    if (j->rc) snprintf(j->msg, sizeof j->msg, "%s", wt_last_error());
}

static void step_job_done(napi_env env, napi_status status, void *data)
{
    step_job *j = (step_job *)data;
    if (status == napi_ok && j->rc == 0) {
        napi_value u;
        napi_get_undefined(env, &u);
        napi_resolve_deferred(env, j->deferred, u);
    } else {
        char code[16];
        snprintf(code, sizeof code, "WT%d", j->rc);
        napi_value c, m, e;
        napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &c);
        napi_create_string_utf8(env, status == napi_ok ? j->msg : "stepSlabsAsync was cancelled", NAPI_AUTO_LENGTH, &m);
        napi_create_error(env, c, m, &e);
        napi_reject_deferred(env, j->deferred, e);
    }
    napi_delete_reference(env, j->href);
    if (j->fref) napi_delete_reference(env, j->fref);
    napi_delete_async_work(env, j->work);
    free(j);
}

static napi_value js_step_slabs_async(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const int n = (int)num(env, a[1], 1), ns = (int)num(env, a[2], 1);
    double *f = tdata(env, a[3], &nb, &ok);
    if (!ok) return NULL;
    if (f && nb < (size_t)n * ns * 24) { napi_throw_range_error(env, NULL, "forces array too short"); return NULL; }
    step_job *j = (step_job *)calloc(1, sizeof *j);
    if (!j) { napi_throw_error(env, NULL, "out of host memory"); return NULL; }
    j->h = h; j->n = n; j->ns = ns; j->f = f;
    napi_value promise, name;
    if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
        napi_create_reference(env, a[0], 1, &j->href) != napi_ok ||
        (f && napi_create_reference(env, a[3], 1, &j->fref) != napi_ok) ||
        napi_create_string_utf8(env, "wt_step_slabs", NAPI_AUTO_LENGTH, &name) != napi_ok ||
        napi_create_async_work(env, NULL, name, step_job_run, step_job_done, j, &j->work) != napi_ok ||
        napi_queue_async_work(env, j->work) != napi_ok) {
        free(j);
        napi_throw_error(env, NULL, "stepSlabsAsync: N-API call failed");
        return NULL;
    }
    return promise;
}

/* getWalls(h, uwx, uwy, uwz) -- Float32Arrays of n cells or null */
static napi_value js_get_walls(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    size_t cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    float *p[3];
    for (int k = 0; k < 3; k++) {
        int ok;
        size_t nb = 0;
        p[k] = tdata(env, a[1 + k], &nb, &ok);
        if (!ok) return NULL;
        if (p[k] && !enough(env, nb, cells * 4, "getWalls output")) return NULL;
    }
    WT(env, wt_get_walls(h, p[0], p[1], p[2]));
    return NULL;
}

static int prim_type(napi_env env, napi_value part)
{
    static const char *names[] = {"box", "ell", "cyl", "cone", "foil"};
    napi_value v;
    char buf[16] = "";
    size_t len = 0;
    if (napi_get_named_property(env, part, "t", &v) != napi_ok ||
        napi_get_value_string_utf8(env, v, buf, sizeof buf, &len) != napi_ok)
        return -1;
    for (int i = 0; i < 5; i++)
        if (strcmp(buf, names[i]) == 0) return i;
    return -1;
}

/* doubles of array property `k` of object o into out[0..n-1] (missing: keep) */
static int prop_doubles(napi_env env, napi_value o, const char *k, double *out, uint32_t n)
{
    napi_value v;
    bool has = false;
    if (napi_has_named_property(env, o, k, &has) != napi_ok || !has) return 0;
    napi_get_named_property(env, o, k, &v);
    if (is_nullish(env, v)) return 0;
    for (uint32_t i = 0; i < n; i++) {
        napi_value e;
        if (napi_get_element(env, v, i, &e) != napi_ok) return 0;
        napi_valuetype t;
        napi_typeof(env, e, &t);
        if (t == napi_number) out[i] = num(env, e, 0);
    }
    return 1;
}

/* voxelizeParts(h, parts [{t, c, s, m, naca?, wheel?}], M Float64Array(9),
 *               pos Float64Array(3), scale, box Int32Array(6), ground, roadVel, wheelU0) */
static napi_value js_voxelize_parts(napi_env env, napi_callback_info info)
{
    napi_value a[9];
    args(env, info, 9, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    uint32_t np = 0;
    if (napi_get_array_length(env, a[1], &np) != napi_ok || np > 64) {
        napi_throw_type_error(env, NULL, "parts must be an array of at most 64 primitives");
        return NULL;
    }
    wt_part parts[64];
    memset(parts, 0, sizeof parts);
    for (uint32_t i = 0; i < np; i++) {
        napi_value p;
        napi_get_element(env, a[1], i, &p);
        const int t = prim_type(env, p);
        if (t < 0) { napi_throw_type_error(env, NULL, "part type must be box, ell, cyl, cone or foil"); return NULL; }
        parts[i].type = t;
        napi_value w;
        bool has = false;
        if (napi_has_named_property(env, p, "wheel", &has) == napi_ok && has) {
            napi_get_named_property(env, p, "wheel", &w);
            parts[i].wheel = boolish(env, w);
        }
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        memcpy(parts[i].m, I, sizeof I);
        prop_doubles(env, p, "c", parts[i].c, 3);
        prop_doubles(env, p, "s", parts[i].s, t == WT_PRIM_CYL || t == WT_PRIM_CONE ? 2 : 3);
        prop_doubles(env, p, "m", parts[i].m, 9);
        parts[i].has_naca = prop_doubles(env, p, "naca", parts[i].naca, 3);
    }
    int ok;
    size_t Mb = 0, pb = 0, bb = 0;
    const double *M = tdata(env, a[2], &Mb, &ok);
    if (!ok) return NULL;
    const double *pos = tdata(env, a[3], &pb, &ok);
    if (!ok) return NULL;
    const int *box = tdata(env, a[5], &bb, &ok);
    if (!ok) return NULL;
    if (!enough(env, M ? Mb : 0, 72, "rotation M") || !enough(env, pos ? pb : 0, 24, "position") ||
        !enough(env, box ? bb : 0, 24, "box"))
        return NULL;
    wt_body_stats s;
    WT(env, wt_voxelize_parts(h, parts, (int)np, M, pos, num(env, a[4], 1), box, boolish(env, a[6]),
                              num(env, a[7], 0), num(env, a[8], 0), &s));
    return stats_object(env, &s);
}

/* stepSlab(h, z0, z1, force Float64Array(3) | null) */
static napi_value js_step_slab(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb = 0;
    double *f = tdata(env, a[3], &nb, &ok);
    if (!ok) return NULL;
    if (f && !enough(env, nb, 24, "stepSlab force")) return NULL;
    WT(env, wt_step_slab(h, (int)num(env, a[1], 1), (int)num(env, a[2], 1), f));
    return NULL;
}

/* boundaries(h, z0, z1, doLo, doHi) */
static napi_value js_boundaries(napi_env env, napi_callback_info info)
{
    napi_value a[5];
    args(env, info, 5, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    WT(env, wt_boundaries(h, (int)num(env, a[1], 0), (int)num(env, a[2], 0), boolish(env, a[3]), boolish(env, a[4])));
    return NULL;
}

/* setGray(h, w Float32Array(n) | null) */
static napi_value js_set_gray(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const float *w = tdata(env, a[1], &nb, &ok);
    if (!ok) return NULL;
    size_t cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    if (w && !enough(env, nb, cells * 4, "setGray weights")) return NULL;
    WT(env, wt_set_gray(h, w));
    return NULL;
}

/* adjointForward(h, n, interval, forces Float64Array(3n) | null) */
static napi_value js_adjoint_forward(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb;
    const int n = (int)num(env, a[1], 1);
    double *f = tdata(env, a[3], &nb, &ok);
    if (!ok) return NULL;
    if (f && nb < (size_t)n * 24) { napi_throw_range_error(env, NULL, "forces array too short"); return NULL; }
    WT(env, wt_adjoint_forward(h, n, (int)num(env, a[2], 0), f));
    return NULL;
}

/* adjointBackward(h, nsteps, Fbar Float64Array(3*nsteps), dJdw Float64Array(cells)) -> reverse-sweep ms */
static napi_value js_adjoint_backward(napi_env env, napi_callback_info info)
{
    napi_value a[4];
    args(env, info, 4, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok1, ok2;
    size_t nb1, nb2;
    const int n = (int)num(env, a[1], 0);
    const double *Fb = tdata(env, a[2], &nb1, &ok1);
    double *g = tdata(env, a[3], &nb2, &ok2);
    if (!ok1 || !ok2) return NULL;
    size_t cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    if (!Fb || !g) { napi_throw_type_error(env, NULL, "adjointBackward needs Fbar and dJdw arrays"); return NULL; }
    if (n < 1 || !enough(env, nb1, (size_t)n * 24, "adjointBackward Fbar") ||
        !enough(env, nb2, cells * 8, "adjointBackward dJdw"))
        return NULL;
    double ms = 0;
    WT(env, wt_adjoint_backward(h, n, Fb, g, &ms));
    napi_value v;
    napi_create_double(env, ms, &v);
    return v;
}

static napi_value js_adjoint_release(napi_env env, napi_callback_info info)
{
    napi_value a[1];
    args(env, info, 1, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    WT(env, wt_adjoint_release(h));
    return NULL;
}

/* adjointRecord(h) -> id of the recorded forward pass */
static napi_value js_adjoint_record(napi_env env, napi_callback_info info)
{
    napi_value a[1];
    args(env, info, 1, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int64_t id = 0;
    WT(env, wt_adjoint_record(h, &id));
    napi_value v;
    napi_create_double(env, (double)id, &v);
    return v;
}

static napi_value js_set_parity(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    WT(env, wt_set_parity(h, (int)num(env, a[1], 0)));
    return NULL;
}

static napi_value js_set_boundary_mode(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    WT(env, wt_set_boundary_mode(h, (int)num(env, a[1], 0)));
    return NULL;
}

static napi_value js_get_state(napi_env env, napi_callback_info info)
{
    napi_value a[1], o, f;
    args(env, info, 1, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    wt_state_t s;
    WT(env, wt_get_state(h, &s));
    CALL(env, napi_create_object(env, &o));
    set_num(env, o, "parity", (double)s.parity);
    set_num(env, o, "steps", (double)s.steps);
    set_num(env, o, "blockCells", (double)s.block_cells);
    set_num(env, o, "nx", s.nx); set_num(env, o, "ny", s.ny); set_num(env, o, "nz", s.nz);
    CALL(env, napi_create_array_with_length(env, 3, &f));
    for (uint32_t k = 0; k < 3; k++) {
        napi_value e;
        napi_create_double(env, s.force[k], &e);
        napi_set_element(env, f, k, e);
    }
    napi_set_named_property(env, o, "force", f);
    return o;
}

/* getFields(h, rho, ux, uy, uz) -- typed arrays or null */
static napi_value js_get_fields(napi_env env, napi_callback_info info)
{
    napi_value a[5];
    args(env, info, 5, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    void *p[4];
    size_t cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    for (int k = 0; k < 4; k++) {
        size_t nb = 0;
        p[k] = tdata(env, a[1 + k], &nb, &ok);
        if (!ok) return NULL;
        if (p[k] && !enough(env, nb, cells * fb, "getFields output")) return NULL;
    }
    WT(env, wt_get_fields(h, p[0], p[1], p[2], p[3]));
    return NULL;
}

/* slice(h, axis, k, field, out Float32Array, solid Uint8Array | null) */
static napi_value js_slice(napi_env env, napi_callback_info info)
{
    napi_value a[6];
    args(env, info, 6, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    wt_state_t st;
    if (wt_get_state(h, &st) != 0) {
        napi_throw_error(env, NULL, wt_last_error());
        return NULL;
    }
    const int axis = (int)num(env, a[1], 0);
    const size_t cells = (size_t)st.nx * (size_t)(axis == 0 ? st.ny : st.nz);
    int ok;
    size_t ob = 0, sb = 0;
    void *out = tdata(env, a[4], &ob, &ok);
    if (!ok) return NULL;
    void *sol = tdata(env, a[5], &sb, &ok);
    if (!ok) return NULL;
    if (!out) {
        napi_throw_type_error(env, NULL, "slice needs an output Float32Array");
        return NULL;
    }
    if (!enough(env, ob, cells * 4, "slice output") || (sol && !enough(env, sb, cells, "slice solid"))) return NULL;
    WT(env, wt_slice(h, axis, (int)num(env, a[2], 0), (int)num(env, a[3], 0), out, sol));
    return NULL;
}

/* sampleMany(h, xyz Float64Array, out Float64Array) */
static napi_value js_sample_many(napi_env env, napi_callback_info info)
{
    napi_value a[3];
    args(env, info, 3, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t xb = 0, ob = 0;
    void *xyz = tdata(env, a[1], &xb, &ok);
    if (!ok) return NULL;
    void *out = tdata(env, a[2], &ob, &ok);
    if (!ok) return NULL;
    if (!xyz || !out) {
        napi_throw_type_error(env, NULL, "sampleMany needs Float64Array points and output");
        return NULL;
    }
    const size_t count = xb / 24;
    if (!enough(env, ob, count * 24, "sampleMany output")) return NULL;
    WT(env, wt_sample(h, xyz, (int64_t)count, out));
    return NULL;
}

static napi_value js_get_pops(napi_env env, napi_callback_info info)
{
    napi_value a[3];
    args(env, info, 3, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb = 0, cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    void *out = tdata(env, a[2], &nb, &ok);
    if (!ok) return NULL;
    if (!enough(env, out ? nb : 0, 19 * cells * fb, "getPops output")) return NULL;
    WT(env, wt_get_pops(h, (int)num(env, a[1], 0), out));
    return NULL;
}

static napi_value js_set_pops(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t nb = 0, cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    void *in = tdata(env, a[1], &nb, &ok);
    if (!ok) return NULL;
    if (!enough(env, in ? nb : 0, 19 * cells * fb, "setPops input")) return NULL;
    WT(env, wt_set_pops(h, in));
    return NULL;
}

static napi_value js_get_solid(napi_env env, napi_callback_info info)
{
    napi_value a[3];
    args(env, info, 3, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    int ok;
    size_t sb = 0, rb = 0, cells, fb;
    if (!lattice_info(env, h, &cells, &fb)) return NULL;
    uint8_t *s = tdata(env, a[1], &sb, &ok);
    if (!ok) return NULL;
    uint8_t *nr = tdata(env, a[2], &rb, &ok);
    if (!ok) return NULL;
    if ((s && !enough(env, sb, cells, "getSolid solid")) || (nr && !enough(env, rb, cells, "getSolid near")))
        return NULL;
    WT(env, wt_get_solid(h, s, nr));
    return NULL;
}

static napi_value js_time_steps(napi_env env, napi_callback_info info)
{
    napi_value a[2], out;
    args(env, info, 2, a);
    wt_lattice *h = handle(env, a[0]);
    if (!h) return NULL;
    double ms = 0;
    WT(env, wt_time_steps(h, (int)num(env, a[1], 1), &ms));
    CALL(env, napi_create_double(env, ms, &out));
    return out;
}

static napi_value js_device_memory(napi_env env, napi_callback_info info)
{
    (void)info;
    uint64_t f = 0, t = 0;
    WT(env, wt_device_memory(&f, &t));
    napi_value o;
    CALL(env, napi_create_object(env, &o));
    set_num(env, o, "free", (double)f);
    set_num(env, o, "total", (double)t);
    return o;
}

static napi_value meta_object(napi_env env, const wt_voxmeta *m)
{
    napi_value o, arr;
    if (napi_create_object(env, &o) != napi_ok) return NULL;
    set_num(env, o, "nx", m->nx); set_num(env, o, "ny", m->ny); set_num(env, o, "nz", m->nz);
    set_num(env, o, "cell", m->cell); set_num(env, o, "inv", m->inv);
    const struct { const char *k; const double *v; uint32_t n; } vec[] = {
        {"min", m->min, 3}, {"ext", m->ext, 3}, {"centre", m->centre, 3}, {"bbox", m->bbox, 6}};
    for (int i = 0; i < 4; i++) {
        napi_create_array_with_length(env, vec[i].n, &arr);
        for (uint32_t k = 0; k < vec[i].n; k++) {
            napi_value e;
            napi_create_double(env, vec[i].v[k], &e);
            napi_set_element(env, arr, k, e);
        }
        napi_set_named_property(env, o, vec[i].k, arr);
    }
    set_num(env, o, "tris", m->ntri);
    set_num(env, o, "solid", (double)m->solid);
    set_num(env, o, "disagree", (double)m->disagree);
    set_num(env, o, "openness", m->openness);
    return o;
}

/* voxelBounds(P Float32Array) -> [x0,y0,z0,x1,y1,z1] */
static napi_value js_voxel_bounds(napi_env env, napi_callback_info info)
{
    napi_value a[1], arr;
    args(env, info, 1, a);
    int ok;
    size_t nb;
    const float *P = tdata(env, a[0], &nb, &ok);
    if (!ok) return NULL;
    double b[6];
    WT(env, wt_voxel_bounds(P, (int64_t)(nb / 4), b));
    CALL(env, napi_create_array_with_length(env, 6, &arr));
    for (uint32_t k = 0; k < 6; k++) {
        napi_value e;
        napi_create_double(env, b[k], &e);
        napi_set_element(env, arr, k, e);
    }
    return arr;
}

/* voxelPlan(P, res) -> meta ; voxelBuild(P, res, grid Uint8Array) -> meta */
static napi_value js_voxel_plan(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    int ok;
    size_t nb;
    const float *P = tdata(env, a[0], &nb, &ok);
    if (!ok) return NULL;
    wt_voxmeta m;
    WT(env, wt_voxel_plan(P, (int64_t)(nb / 4), (int)num(env, a[1], 96), &m));
    return meta_object(env, &m);
}

/* meshParse(name, ArrayBuffer | typed array) -> Float32Array (MeshIO.parse) */
static napi_value js_mesh_parse(napi_env env, napi_callback_info info)
{
    napi_value a[2];
    args(env, info, 2, a);
    char name[1024];
    size_t nl = 0;
    if (napi_get_value_string_utf8(env, a[0], name, sizeof name, &nl) != napi_ok) {
        napi_throw_type_error(env, NULL, "meshParse needs a file name");
        return NULL;
    }
    void *data = NULL;
    size_t bytes = 0;
    bool isab = false;
    napi_is_arraybuffer(env, a[1], &isab);
    if (isab) {
        CALL(env, napi_get_arraybuffer_info(env, a[1], &data, &bytes));
    } else {
        int ok;
        data = tdata(env, a[1], &bytes, &ok);
        if (!ok) return NULL;
    }
    int64_t count = 0;
    WT(env, wt_mesh_parse(name, data, (int64_t)bytes, NULL, &count));
    void *out = NULL;
    napi_value ab, arr;
    CALL(env, napi_create_arraybuffer(env, (size_t)count * 4, &out, &ab));
    WT(env, wt_mesh_parse(name, data, (int64_t)bytes, (float *)out, &count));
    CALL(env, napi_create_typedarray(env, napi_float32_array, (size_t)count, ab, 0, &arr));
    return arr;
}

static napi_value js_voxel_build(napi_env env, napi_callback_info info)
{
    napi_value a[3];
    args(env, info, 3, a);
    int ok;
    size_t nb, gb;
    const float *P = tdata(env, a[0], &nb, &ok);
    if (!ok) return NULL;
    uint8_t *grid = tdata(env, a[2], &gb, &ok);
    if (!ok) return NULL;
    wt_voxmeta m;
    WT(env, wt_voxel_plan(P, (int64_t)(nb / 4), (int)num(env, a[1], 96), &m));
    if (!grid || gb < (size_t)m.nx * m.ny * m.nz) {
        napi_throw_range_error(env, NULL, "grid array too short");
        return NULL;
    }
    WT(env, wt_voxel_build(P, (int64_t)(nb / 4), (int)num(env, a[1], 96), &m, grid));
    return meta_object(env, &m);
}

static napi_value js_abi(napi_env env, napi_callback_info info)
{
    (void)info;
    napi_value v;
    napi_create_int32(env, wt_abi_version(), &v);
    return v;
}

static napi_value init(napi_env env, napi_value exports)
{
    const struct { const char *name; napi_callback cb; } fns[] = {
        {"abiVersion", js_abi}, {"create", js_create}, {"setParams", js_set_params},
        {"getParams", js_get_params}, {"reset", js_reset}, {"voxelize", js_voxelize},
        {"voxelizeGrid", js_voxelize_grid}, {"step", js_step}, {"stepSlabs", js_step_slabs}, {"stepSlabsAsync", js_step_slabs_async},
        {"setParity", js_set_parity}, {"getState", js_get_state}, {"getFields", js_get_fields},
        {"getPops", js_get_pops}, {"setPops", js_set_pops}, {"getSolid", js_get_solid}, {"timeSteps", js_time_steps}, {"deviceMemory", js_device_memory}, {"setBoundaryMode", js_set_boundary_mode}, {"slice", js_slice}, {"sampleMany", js_sample_many}, {"meshParse", js_mesh_parse},
        {"voxelizeParts", js_voxelize_parts}, {"getWalls", js_get_walls}, {"stepSlab", js_step_slab}, {"boundaries", js_boundaries}, {"setGray", js_set_gray}, {"adjointForward", js_adjoint_forward}, {"adjointBackward", js_adjoint_backward},
        {"adjointRelease", js_adjoint_release}, {"adjointRecord", js_adjoint_record},
        {"voxelBounds", js_voxel_bounds}, {"voxelPlan", js_voxel_plan}, {"voxelBuild", js_voxel_build}};
    for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
        napi_value f;
        if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].cb, NULL, &f) != napi_ok) return NULL;
        if (napi_set_named_property(env, exports, fns[i].name, f) != napi_ok) return NULL;
    }
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
