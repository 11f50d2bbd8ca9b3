// extern "C" entry points of libwindtunnel (declared in include/windtunnel.h).
//
// Host-side bookkeeping that the reference does on its typed arrays (the
// solid map, wall velocities, body statistics of voxelize) runs here on host
// mirrors; everything per cell and per step runs in the kernels of lbm.hip
// and voxel.hip on the lattice's own HIP stream.

#include "wt_internal.h"

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

static thread_local char g_err[512] = "";

int wt_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define WT_CHECK_H(h)                                                   \
    do {                                                                \
        if (!(h)) return wt_fail(WT_E_ARG, "null lattice handle");      \
    } while (0)

namespace {

// Host-side fault injection for failure-handling tests:
// WT_FAULT=step:N makes the call that would run step N fail with
// WT_E_FAULT before launching anything.
int64_t fault_step()
{
    static int64_t v = -2;
    if (v == -2) {
        v = -1;
        const char *e = std::getenv("WT_FAULT");
        if (e && std::strncmp(e, "step:", 5) == 0) v = std::atoll(e + 5);
    }
    return v;
}

int check_device()
{
    int dev = 0, count = 0;
    WT_HIP(hipGetDeviceCount(&count));
    if (count < 1) return wt_fail(WT_E_DEVICE, "no HIP device");
    WT_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    WT_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return wt_fail(WT_E_DEVICE, "device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    return 0;
}

int alloc(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return wt_fail(WT_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    return 0;
}

void release(wt_lattice *h)
{
    adj_release(h);
    for (void *p : {(void *)h->d_buf[0], (void *)h->d_solid, (void *)h->d_sbits,
                    (void *)h->d_nbits, (void *)h->d_uw, (void *)h->d_mac, (void *)h->d_partial,
                    (void *)h->d_forces, (void *)h->d_near, (void *)h->d_ring,
                    (void *)h->d_prod, (void *)h->d_ops, h->d_stage, (void *)h->d_vgrid,
                    (void *)h->d_near2, (void *)h->d_pcnt, h->d_rvwork, (void *)h->d_wgt})
        if (p) (void)hipFree(p);
    if (h->h_stage) (void)hipHostFree(h->h_stage);
    for (hipEvent_t e : h->tev) (void)hipEventDestroy(e);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

int upload_body(wt_lattice *h)
{
    WT_HIP(hipMemcpyAsync(h->d_solid, h->h_solid.data(), (size_t)h->n, hipMemcpyHostToDevice, h->stream));
    WT_HIP(hipMemcpyAsync(h->d_uw, h->h_uw.data(), sizeof(float) * 3 * (size_t)h->n, hipMemcpyHostToDevice,
                          h->stream));
    return lbm_rebuild_masks(h);
}

// Pinned host staging and its device twin of at least `bytes` each.
static int stage_reserve(wt_lattice *h, size_t bytes)
{
    if (bytes <= h->stage_cap) return 0;
    if (h->h_stage) (void)hipHostFree(h->h_stage);
    if (h->d_stage) (void)hipFree(h->d_stage);
    h->h_stage = nullptr;
    h->d_stage = nullptr;
    h->stage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, (size_t)1 << 20);
    if (hipHostMalloc((void **)&h->h_stage, cap, hipHostMallocDefault) != hipSuccess) {
        h->h_stage = nullptr;
        return wt_fail(WT_E_NOMEM, "hipHostMalloc of %zu bytes of voxelize staging failed", cap);
    }
    if (int rc = alloc(&h->d_stage, cap)) return rc;
    h->stage_cap = cap;
    return 0;
}

// The cells of box b (x0, y0, z0, x1, y1, z1, inclusive) of a lattice-shaped
// host array (elem bytes per cell), packed in x, y, z order into out.
static void gather_box(const wt_lattice *h, const void *src, int elem, const int *b, uint8_t *out)
{
    const size_t row = (size_t)(b[3] - b[0] + 1) * (size_t)elem;
    for (int z = b[2]; z <= b[5]; z++)
        for (int y = b[1]; y <= b[4]; y++) {
            const int64_t i = b[0] + h->nx * ((int64_t)y + (int64_t)h->ny * z);
            std::memcpy(out, (const uint8_t *)src + (size_t)i * (size_t)elem, row);
            out += row;
        }
}

// Worker slab split of js/lbmWorker.js:43-48.
void slab_range(int nz, int id, int nw, int *bz0, int *bz1)
{
    const int lo = 1, hi = nz - 1, span = hi - lo;
    const int z0 = lo + (int)(((int64_t)span * id) / nw);
    const int z1 = lo + (int)(((int64_t)span * (id + 1)) / nw);
    *bz0 = id == 0 ? 0 : z0;
    *bz1 = id == nw - 1 ? nz : z1;
}

}  // namespace

extern "C" {

const char *wt_last_error(void) { return g_err; }
int wt_abi_version(void) { return WT_ABI_VERSION; }

int wt_device_memory(uint64_t *free_bytes, uint64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return wt_fail(WT_E_ARG, "null output pointer");
    if (int rc = check_device()) return rc;
    size_t f = 0, t = 0;
    WT_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return 0;
}

int wt_create(int nx, int ny, int nz, int storage, int arith, int layout, wt_lattice **out)
{
    if (!out) return wt_fail(WT_E_ARG, "null output pointer");
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3) return wt_fail(WT_E_ARG, "lattice must be at least 3x3x3, got %dx%dx%d", nx, ny, nz);
    // Cell indices in the near-wall and cold-cell lists are 32-bit, a lattice
    // plane's cell count is an int, and the streaming kernels take y and z
    // from the grid's y and z (at most 65535) and x offsets within a row from
    // 24-bit products; population offsets are 64-bit everywhere. Beyond that
    // the limit is device memory (about 1.5 G cells for float32 with two
    // buffers, 2.7 G for binary16 on 288 GB), which the allocation reports.
    if ((int64_t)nx * ny * nz > (int64_t)UINT32_MAX)
        return wt_fail(WT_E_ARG, "lattice of %lld cells exceeds the 32-bit cell index range",
                       (long long)nx * ny * nz);
    if (ny > 65535 || nz > 65535) return wt_fail(WT_E_ARG, "ny and nz must be below 65536");
    if (nx >= (1 << 24) || (int64_t)nx * ny > (int64_t)INT32_MAX)
        return wt_fail(WT_E_ARG, "lattice plane of %dx%d cells is too large", nx, ny);
    const bool ok = (storage == WT_STORE_F32 && (arith == WT_ARITH_F32 || arith == WT_ARITH_F64)) ||
                    (storage == WT_STORE_F16 && arith == WT_ARITH_F32) ||
                    (storage == WT_STORE_F64 && arith == WT_ARITH_F64);
    if (!ok) return wt_fail(WT_E_ARG, "unsupported storage/arith combination %d/%d", storage, arith);
    if (layout != WT_LAYOUT_AB && layout != WT_LAYOUT_ESOTERIC) return wt_fail(WT_E_ARG, "unknown layout %d", layout);
    if (int rc = check_device()) return rc;

    wt_lattice *h = new wt_lattice();
    h->nx = nx; h->ny = ny; h->nz = nz; h->n = (int64_t)nx * ny * nz;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return wt_fail(WT_E_DEVICE, "hipGetDevice failed");
    }
    h->storage = storage; h->arith = arith; h->layout = layout;
    h->pop_elem = storage == WT_STORE_F16 ? 2 : (storage == WT_STORE_F64 ? 8 : 4);
    h->mac_elem = storage == WT_STORE_F64 ? 8 : 4;
    h->nwx = (nx + 31) / 32;
    const size_t n = (size_t)h->n;
    // Population planes are padded so that consecutive planes do not start
    // at the same offset modulo large powers of two: with 2^k-sized planes
    // all 38 streams of a step would hit the same HBM channel at once.
    // WT_BUFFER_PAD (elements) offsets the second AB buffer: a tuning knob.
    int64_t bpad = 0;
    if (const char *e = std::getenv("WT_BUFFER_PAD")) bpad = std::atoll(e);
    if (bpad < 0) bpad = 0;
    // Population layout, measured on MI355X at 512x256x256: 64-cell blocks
    // for 4- and 8-byte values, 128-cell blocks (256 B per direction) for
    // binary16 (WT_BLOCK_SHIFT = log2 of the block overrides it, 0 = planar:
    // a tuning and test knob).
    h->bsh = storage == WT_STORE_F16 ? 7 : 6;
    if (const char *e = std::getenv("WT_BLOCK_SHIFT")) {
        const int v = std::atoi(e);
        h->bsh = v <= 0 ? 0 : std::max(2, std::min(10, v));
    }
    // Row order of the step (measured on MI355X at 512x256x256): alternating
    // z sweeps for the in-place layout (5% faster), forward sweeps for two
    // buffers (alternation costs them 3%). WT_SWEEP (0..3) overrides it: a
    // tuning knob.
    h->sweep = layout == WT_LAYOUT_ESOTERIC ? 1 : 0;
    if (const char *e = std::getenv("WT_SWEEP")) h->sweep = std::atoi(e) & 3;
    // binary16 storage with two buffers and rows of 128-cell multiples steps
    // two cells per thread (half the memory instructions). WT_PAIRS=0 keeps
    // one cell per thread, WT_PAIRS=2 takes two for float32 storage as well
    // (the same populations and fields bit for bit: a test and tuning knob)
    if (const char *e = std::getenv("WT_PAIRS")) h->pairs = std::atoi(e);
    // The step addresses a block-laid row as whole blocks, so the blocked
    // layout needs rows of a whole number of blocks; other widths use planes.
    if (h->bsh && nx % (1 << h->bsh)) h->bsh = 0;
    // planar planes are padded so that they do not start at the same offset
    // modulo large powers of two (WT_PLANE_PAD, elements, overrides it)
    int64_t ppad = 512;
    if (const char *e = std::getenv("WT_PLANE_PAD")) ppad = std::max<int64_t>(0, std::atoll(e));
    if (h->bsh) h->ns = (h->n + (int64_t(1) << h->bsh) - 1) >> h->bsh << h->bsh;
    else h->ns = (h->n + 63) / 64 * 64 + (ppad + 63) / 64 * 64;
    const size_t ns = (size_t)h->ns;
    // both AB buffers live in one allocation, the second starting bpad
    // elements (rounded to 64) after the end of the first
    const size_t bstride = 19 * ns + (size_t)(bpad + 63) / 64 * 64;
    int rc = 0;
    if (!rc) rc = alloc(&h->d_buf[0], (layout == WT_LAYOUT_AB ? bstride + 19 * ns : 19 * ns) * h->pop_elem);
    if (!rc && layout == WT_LAYOUT_AB) h->d_buf[1] = (char *)h->d_buf[0] + bstride * h->pop_elem;
    // in place: slots no cell ever writes (outside-pointing slots of the
    // boundary planes) are never read either; zero them for reproducibility
    if (!rc && layout == WT_LAYOUT_ESOTERIC && hipMemset(h->d_buf[0], 0, 19 * ns * h->pop_elem) != hipSuccess)
        rc = wt_fail(WT_E_DEVICE, "hipMemset failed");
    if (!rc) rc = alloc((void **)&h->d_solid, n);
    if (!rc) rc = alloc((void **)&h->d_sbits, sizeof(uint32_t) * h->nwx * (size_t)ny * nz);
    if (!rc) rc = alloc((void **)&h->d_nbits, sizeof(uint32_t) * h->nwx * (size_t)ny * nz);
    if (!rc) rc = alloc((void **)&h->d_uw, sizeof(float) * 3 * n);
    if (!rc) rc = alloc(&h->d_mac, 4 * n * h->mac_elem);
    // force partials: one per near-wall block, at most (nz + 1) planes of
    // row blocks (every listed cell is a lattice cell)
    const int bx = wt::row_block(nx);
    h->partial_cap = (int64_t)((nx + bx - 1) / bx) * ny * ((int64_t)nz + 1);
    if (!rc) rc = alloc((void **)&h->d_partial, sizeof(double) * 6 * (size_t)h->partial_cap);
    h->forces_cap = 4096;
    if (!rc) rc = alloc((void **)&h->d_forces, sizeof(double) * 3 * (size_t)h->forces_cap);
    if (!rc && hipStreamCreate(&h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "hipStreamCreate failed");

    if (!rc && (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess))
        rc = wt_fail(WT_E_DEVICE, "hipEventCreate failed");
    if (rc) { release(h); delete h; return rc; }
    h->h_solid.assign(n, 0);
    h->h_uw.assign(3 * n, 0.0f);
    rc = upload_body(h);
    if (!rc) rc = lbm_reset(h);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "initial reset failed");
    if (rc) { release(h); delete h; return rc; }
    *out = h;
    return 0;
}

int wt_destroy(wt_lattice *h)
{
    if (!h) return 0;
    (void)hipStreamSynchronize(h->stream);
    release(h);
    delete h;
    return 0;
}

int wt_set_params(wt_lattice *h, const wt_params *p)
{
    WT_CHECK_H(h);
    if (!p) return wt_fail(WT_E_ARG, "null params");
    h->p = *p;
    h->epoch++;
    return 0;
}

int wt_get_params(const wt_lattice *h, wt_params *p)
{
    WT_CHECK_H(h);
    if (!p) return wt_fail(WT_E_ARG, "null params");
    *p = h->p;
    return 0;
}

int wt_reset(wt_lattice *h)
{
    WT_CHECK_H(h);
    if (int rc = lbm_reset(h)) return rc;
    WT_HIP(hipStreamSynchronize(h->stream));
    h->steps = 0;
    h->force[0] = h->force[1] = h->force[2] = 0;
    return 0;
}

int wt_voxelize(wt_lattice *h, const int box[6], const uint8_t *mask, const float *wv, int ground,
                double roadVel, wt_body_stats *out)
{
    WT_CHECK_H(h);
    if (!box) return wt_fail(WT_E_ARG, "null box");
    h->epoch++;
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
    const bool empty = x1 < x0 || y1 < y0 || z1 < z0;
    if (!empty && (x0 < 1 || y0 < 1 || z0 < 1 || x1 > nx - 2 || y1 > ny - 2 || z1 > nz - 2))
        return wt_fail(WT_E_ARG, "box [%d,%d,%d]-[%d,%d,%d] is not within the interior", x0, y0, z0, x1, y1, z1);
    if (!empty && !mask) return wt_fail(WT_E_ARG, "null mask for a non-empty box");
    if (!empty && !wv) {
        const int64_t cnt = (int64_t)(x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
        if (std::memchr(mask, 2, (size_t)cnt)) return wt_fail(WT_E_ARG, "mask holds wall-velocity cells but wv is null");
        if (voxelize_device_applies()) {
            // the mask goes up and the whole update runs on the device
            if ((size_t)cnt > h->vgrid_cap) {
                if (h->d_vgrid) (void)hipFree(h->d_vgrid);
                h->d_vgrid = nullptr;
                h->vgrid_cap = 0;
                if (int rc = alloc((void **)&h->d_vgrid, (size_t)cnt)) return rc;
                h->vgrid_cap = (size_t)cnt;
            }
            WT_HIP(hipMemcpyAsync(h->d_vgrid, mask, (size_t)cnt, hipMemcpyHostToDevice, h->stream));
            return voxelize_device(h, box, h->d_vgrid, nullptr, ground, roadVel, out);
        }
    }
    const int64_t n = h->n, plane = (int64_t)nx * ny;
    if (h->h_ops.empty()) h->h_ops.assign((size_t)n, 0);
    uint8_t *ops = h->h_ops.data();
    uint8_t *solid = h->h_solid.data();
    float *uwx = h->h_uw.data(), *uwy = uwx + n, *uwz = uwx + 2 * n;
    // boxes (x0, y0, z0, x1, y1, z1) of the cells whose solid state or ops
    // change, and of the cells whose wall velocities change
    int cb[6] = {nx, ny, nz, -1, -1, -1}, ub[6] = {nx, ny, nz, -1, -1, -1};
    auto touch = [](int *b, int x, int y, int z) {
        b[0] = std::min(b[0], x); b[1] = std::min(b[1], y); b[2] = std::min(b[2], z);
        b[3] = std::max(b[3], x); b[4] = std::max(b[4], y); b[5] = std::max(b[5], z);
    };
    // h->uw_cells counts the cells with a nonzero wall velocity, so a zero
    // velocity needs no look at the mirror while there are none
    auto set_uw = [&](int64_t i, int x, int y, int z, float vx, float vy, float vz) {
        const bool nz_new = vx != 0 || vy != 0 || vz != 0;
        if (!nz_new && h->uw_cells == 0) return;
        if (uwx[i] != vx || uwy[i] != vy || uwz[i] != vz) {
            const bool nz_old = uwx[i] != 0 || uwy[i] != 0 || uwz[i] != 0;
            h->uw_cells += (int64_t)nz_new - (int64_t)nz_old;
            uwx[i] = vx; uwy[i] = vy; uwz[i] = vz;
            touch(ub, x, y, z);
        }
    };
    // js/lbm.js:194-202: the old body becomes fluid at u0 (all of it lies in
    // the previous body's box)
    h->ground = !!ground;
    int mov = 0;
    {
        const int *ob = h->obj_box;
        for (int z = ob[2]; z <= ob[5]; z++)
            for (int y = ob[1]; y <= ob[4]; y++) {
                const int64_t r = nx * ((int64_t)y + (int64_t)ny * z);
                uint8_t *srow = solid + r, *orow = ops + r;
                if (h->uw_cells == 0) {
                    // no wall velocity to reset: a branch-free pass over the row,
                    // the changed box taken as the box's extent along x
                    uint8_t any = 0;
                    for (int x = ob[0]; x <= ob[3]; x++) {
                        const uint8_t m = srow[x] == wt::SOLID_OBJ;
                        srow[x] = m ? (uint8_t)wt::FLUID : srow[x];
                        orow[x] |= m;
                        any |= m;
                    }
                    if (any) { touch(cb, ob[0], y, z); touch(cb, ob[3], y, z); }
                    continue;
                }
                int xa = nx, xb = -1;  // first and last cleared cell of the row
                for (int x = ob[0]; x <= ob[3]; x++) {
                    if (srow[x] != wt::SOLID_OBJ) continue;
                    srow[x] = wt::FLUID; orow[x] |= 1;
                    set_uw(r + x, x, y, z, 0, 0, 0);
                    xa = std::min(xa, x); xb = x;
                }
                if (xb >= 0) { touch(cb, xa, y, z); touch(cb, xb, y, z); }
            }
    }
    // js/lbm.js:204-214: the road row
    const int gy = ground ? wt::SOLID_WALL : wt::FLUID;
    const double rv = (roadVel != roadVel || roadVel == 0) ? 0 : roadVel;
    if (rv != 0) mov = 1;
    for (int z = 0; z < nz; z++) {
        for (int x = 0; x < nx; x++) {
            const int64_t i = x + plane * z;
            if (solid[i] != gy) {
                solid[i] = (uint8_t)gy;
                if (gy == wt::FLUID) ops[i] |= 1;
                touch(cb, x, 0, z);
            }
            set_uw(i, x, 0, z, gy == wt::SOLID_WALL ? (float)rv : 0.0f, 0, 0);
        }
    }
    // js/lbm.js:216-246: the new body, its projections and lowest cells
    std::vector<uint8_t> col((size_t)ny * nz, 0), plan((size_t)nx * nz, 0);
    std::vector<int16_t> low((size_t)nx * nz, 32767);
    int64_t cells = 0, k = 0, kw = 0;
    int nb[6] = {nx, ny, nz, -1, -1, -1};  // box of the new body's cells
    if (!empty) {
        const int w = x1 - x0 + 1;
        for (int z = z0; z <= z1; z++)
            for (int y = y0; y <= y1; y++, k += w) {
                const uint8_t *mrow = mask + k;
                int xa = w, xb = -1;  // first and last body cell of the row (box coordinates)
                for (int t = 0; t < w; t++)
                    if (mrow[t]) { xa = t; break; }
                if (xa == w) continue;
                for (int t = w - 1; t >= xa; t--)
                    if (mrow[t]) { xb = t; break; }
                const int64_t r = (int64_t)nx * (y + (int64_t)ny * z);
                if (h->uw_cells == 0 && !std::memchr(mrow + xa, 2, (size_t)(xb - xa + 1))) {
                    // no wall velocities anywhere: a branch-free pass over the row
                    uint8_t *srow = solid + r + x0, *orow = ops + r + x0, *prow = plan.data() + (size_t)nx * z + x0;
                    int16_t *lrow = low.data() + (size_t)nx * z + x0;
                    const int16_t yy = (int16_t)y;
                    int64_t c = 0;
                    for (int t = xa; t <= xb; t++) {
                        const uint8_t v = mrow[t] != 0;
                        srow[t] = v ? (uint8_t)wt::SOLID_OBJ : srow[t];
                        orow[t] |= (uint8_t)(v << 1);
                        prow[t] |= v;
                        lrow[t] = (v && yy < lrow[t]) ? yy : lrow[t];
                        c += v;
                    }
                    cells += c;
                } else {
                    for (int t = xa; t <= xb; t++) {
                        if (!mrow[t]) continue;
                        const int x = x0 + t;
                        const int64_t i = r + x;
                        solid[i] = wt::SOLID_OBJ;
                        ops[i] |= 2;
                        float wx = 0, wy = 0, wz = 0;
                        if (mrow[t] == 2) {
                            wx = wv[3 * kw]; wy = wv[3 * kw + 1]; wz = wv[3 * kw + 2];
                            kw++;
                            mov = 1;
                        }
                        set_uw(i, x, y, z, wx, wy, wz);
                        const size_t pk = x + (size_t)nx * z;
                        plan[pk] = 1;
                        if (y < low[pk]) low[pk] = (int16_t)y;
                        cells++;
                    }
                }
                col[y + (size_t)ny * z] = 1;
                touch(cb, x0 + xa, y, z); touch(cb, x0 + xb, y, z);
                touch(nb, x0 + xa, y, z); touch(nb, x0 + xb, y, z);
            }
    }
    std::copy(nb, nb + 6, h->obj_box);
    int64_t a = 0, p = 0;
    for (uint8_t c : col) a += c;
    for (uint8_t c : plan) p += c;
    h->body.frontalArea = std::max<int64_t>(1, a);
    h->body.planformArea = std::max<int64_t>(1, p);
    h->body.solidCells = cells;
    std::vector<int> gaps;
    for (size_t c = 0; c < plan.size(); c++)
        if (plan[c] && low[c] < 32767) { const int g = low[c] - 1; if (g > 0) gaps.push_back(g); }
    std::sort(gaps.begin(), gaps.end());
    h->body.rideHeight = gaps.empty() ? 0 : gaps[gaps.size() >> 1];
    h->p.movingWalls = mov;

    // device: the changed boxes only, gathered into pinned staging, uploaded
    // at once and scattered (the ops kernel then runs over their planes)
    int rc = 0;
    if (!h->d_ops) {
        rc = alloc((void **)&h->d_ops, (size_t)n);
        if (!rc && hipMemset(h->d_ops, 0, (size_t)n) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "hipMemset failed");
    }
    const bool cdirty = cb[5] >= cb[2], udirty = ub[5] >= ub[2];
    auto box_cells = [](const int *b) {
        return (size_t)(b[3] - b[0] + 1) * (size_t)(b[4] - b[1] + 1) * (size_t)(b[5] - b[2] + 1);
    };
    const size_t cn = cdirty ? box_cells(cb) : 0, un = udirty ? box_cells(ub) : 0;
    const size_t ubase = (2 * cn + 15) / 16 * 16;  // wall velocities after solid and ops
    if (!rc) rc = stage_reserve(h, ubase + 3 * sizeof(float) * un);
    if (!rc && cdirty) {
        gather_box(h, solid, 1, cb, h->h_stage);
        gather_box(h, ops, 1, cb, h->h_stage + cn);
        if (hipMemcpyAsync(h->d_stage, h->h_stage, 2 * cn, hipMemcpyHostToDevice, h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "voxelize upload failed");
        if (!rc) rc = lbm_scatter_box(h, h->d_stage, h->d_solid, 1, cb);
        if (!rc) rc = lbm_scatter_box(h, (uint8_t *)h->d_stage + cn, h->d_ops, 1, cb);
        if (!rc) rc = lbm_apply_voxelize(h, h->d_ops, plane * cb[2], plane * (cb[5] + 1));
    }
    if (!rc && udirty) {
        for (int c = 0; c < 3; c++) gather_box(h, h->h_uw.data() + c * n, 4, ub, h->h_stage + ubase + c * 4 * un);
        if (hipMemcpyAsync((uint8_t *)h->d_stage + ubase, h->h_stage + ubase, 12 * un, hipMemcpyHostToDevice,
                           h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "voxelize upload failed");
        for (int c = 0; c < 3 && !rc; c++)
            rc = lbm_scatter_box(h, (uint8_t *)h->d_stage + ubase + c * 4 * un, h->d_uw + c * n, 4, ub);
    }
    // the near-wall mask of a plane depends on the planes beside it
    if (!rc && cdirty) rc = lbm_rebuild_masks_range(h, std::max(0, cb[2] - 1), std::min(nz - 1, cb[5] + 1));
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "voxelize failed");
    if (cdirty)  // the host ops are kept zero between calls
        for (int z = cb[2]; z <= cb[5]; z++)
            for (int y = cb[1]; y <= cb[4]; y++)
                std::fill(ops + cb[0] + nx * ((int64_t)y + (int64_t)ny * z),
                          ops + cb[3] + 1 + nx * ((int64_t)y + (int64_t)ny * z), 0);
    if (rc) return rc;
    if (out) {
        out->frontalArea = h->body.frontalArea; out->planformArea = h->body.planformArea;
        out->solidCells = h->body.solidCells; out->rideHeight = h->body.rideHeight;
        out->movingWalls = mov; out->ground = h->ground;
    }
    return 0;
}

}  // extern "C"

// wt_voxelize for a body already rasterised into a device mask over box
// (d_wv: the wall velocities of its mask-2 cells, 3 per box cell, or null):
// the diff against the previous body, the road row, the ops, the wall
// velocities, the projections and the lowest cells are computed on the
// device, and only a few scalars, a histogram and the changed boxes of the
// solid map and wall velocities (for the host mirrors) come back. The result
// equals wt_voxelize's on the same mask.
bool voxelize_device_applies()
{
    return !std::getenv("WT_VOXELIZE_HOST");
}

int voxelize_device(wt_lattice *h, const int box[6], const uint8_t *d_mask, const float *d_wv, int ground,
                    double roadVel, wt_body_stats *out)
{
    h->epoch++;
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const int x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
    const bool empty = x1 < x0 || y1 < y0 || z1 < z0;
    if (!empty && (x0 < 1 || y0 < 1 || z0 < 1 || x1 > nx - 2 || y1 > ny - 2 || z1 > nz - 2))
        return wt_fail(WT_E_ARG, "box [%d,%d,%d]-[%d,%d,%d] is not within the interior", x0, y0, z0, x1, y1, z1);
    const int64_t n = h->n, plane = (int64_t)nx * ny;
    const double rv = (roadVel != roadVel || roadVel == 0) ? 0 : roadVel;
    const int *ob = h->obj_box;
    const bool oempty = ob[3] < ob[0] || ob[4] < ob[1] || ob[5] < ob[2];
    int mbox[6] = {1, 1, 1, 0, 0, 0}, ub[6] = {1, 1, 1, 0, 0, 0};
    if (!empty) std::copy(box, box + 6, mbox);
    if (!empty && !oempty)
        for (int k = 0; k < 3; k++) { ub[k] = std::min(box[k], ob[k]); ub[k + 3] = std::max(box[k + 3], ob[k + 3]); }
    else if (!empty) std::copy(box, box + 6, ub);
    else if (!oempty) std::copy(ob, ob + 6, ub);
    if (!h->d_ops) {
        if (int rc = alloc((void **)&h->d_ops, (size_t)n)) return rc;
        WT_HIP(hipMemset(h->d_ops, 0, (size_t)n));
    }
    const size_t work = ((size_t)ny * nz + 255) / 256 * 256 + ((size_t)nx * nz + 255) / 256 * 256 +
                        sizeof(int) * ((size_t)nx * nz + RV_N + (size_t)ny);
    if (work > h->rvwork_cap) {
        if (h->d_rvwork) (void)hipFree(h->d_rvwork);
        h->d_rvwork = nullptr;
        h->rvwork_cap = 0;
        if (int rc = alloc(&h->d_rvwork, work)) return rc;
        h->rvwork_cap = work;
    }
    const int big = 0x7fffffff;
    int res[RV_N] = {0};
    for (int k = 0; k < 6; k++) {
        const int init = k < 3 ? big : -1;
        res[RV_CB + k] = init; res[RV_NB + k] = init; res[RV_UB + k] = init;
    }
    std::vector<int> hist((size_t)ny, 0);
    const float road = ground ? (float)rv : 0.0f;
    h->ground = !!ground;
    if (int rc = lbm_revoxelize_device(h, ub, mbox, d_mask, d_wv, h->ground, road, h->d_rvwork, res, hist.data()))
        return rc;
    // js/lbm.js:240-246: projections and the median gap under the body
    const int64_t cells = res[RV_CELLS];
    h->body.frontalArea = std::max<int64_t>(1, res[RV_FRONTAL]);
    h->body.planformArea = std::max<int64_t>(1, res[RV_PLAN]);
    h->body.solidCells = cells;
    int64_t total = 0;
    for (int c : hist) total += c;
    int ride = 0;
    if (total > 0) {
        const int64_t mid = total >> 1;
        int64_t acc = 0;
        for (int g = 0; g < ny; g++) {
            acc += hist[(size_t)g];
            if (acc > mid) { ride = g; break; }
        }
    }
    h->body.rideHeight = ride;
    const int mov = (rv != 0 || res[RV_WHEELS] > 0) ? 1 : 0;
    h->p.movingWalls = mov;
    h->uw_cells += res[RV_NZ];
    if (cells > 0) std::copy(res + RV_NB, res + RV_NB + 6, h->obj_box);
    else { const int e[6] = {nx, ny, nz, -1, -1, -1}; std::copy(e, e + 6, h->obj_box); }
    const int *cb = res + RV_CB, *wb = res + RV_UB;
    const bool cdirty = cb[5] >= cb[2], wdirty = wb[5] >= wb[2];
    if (cdirty) {
        if (int rc = lbm_apply_voxelize(h, h->d_ops, plane * cb[2], plane * (cb[5] + 1))) return rc;
        if (int rc = lbm_rebuild_masks_range(h, std::max(0, cb[2] - 1), std::min(nz - 1, cb[5] + 1))) return rc;
    }
    // the host mirrors over the changed boxes
    auto cells_of = [](const int *b) {
        return (size_t)(b[3] - b[0] + 1) * (size_t)(b[4] - b[1] + 1) * (size_t)(b[5] - b[2] + 1);
    };
    const size_t cn = cdirty ? cells_of(cb) : 0, wn = wdirty ? cells_of(wb) : 0;
    const size_t wbase = (cn + 15) / 16 * 16;
    if (cn + wn > 0) {
        if (int rc = stage_reserve(h, wbase + 12 * wn)) return rc;
        if (cdirty) {
            if (int rc = lbm_gather_box(h, h->d_solid, h->d_stage, 1, cb)) return rc;
        }
        for (int c = 0; wdirty && c < 3; c++)
            if (int rc = lbm_gather_box(h, h->d_uw + c * n, (uint8_t *)h->d_stage + wbase + c * 4 * wn, 4, wb)) return rc;
        WT_HIP(hipMemcpyAsync(h->h_stage, h->d_stage, wbase + 12 * wn, hipMemcpyDeviceToHost, h->stream));
        WT_HIP(hipStreamSynchronize(h->stream));
        auto scatter = [&](const uint8_t *src, uint8_t *dst, const int *b, int elem) {
            const size_t row = (size_t)(b[3] - b[0] + 1) * (size_t)elem;
            for (int z = b[2]; z <= b[5]; z++)
                for (int y = b[1]; y <= b[4]; y++, src += row)
                    std::memcpy(dst + (size_t)(b[0] + nx * ((int64_t)y + (int64_t)ny * z)) * (size_t)elem, src, row);
        };
        if (cdirty) scatter(h->h_stage, h->h_solid.data(), cb, 1);
        for (int c = 0; wdirty && c < 3; c++)
            scatter(h->h_stage + wbase + c * 4 * wn, (uint8_t *)(h->h_uw.data() + c * n), wb, 4);
    }
    WT_HIP(hipStreamSynchronize(h->stream));
    if (out) {
        out->frontalArea = h->body.frontalArea; out->planformArea = h->body.planformArea;
        out->solidCells = h->body.solidCells; out->rideHeight = h->body.rideHeight;
        out->movingWalls = mov; out->ground = h->ground;
    }
    return 0;
}

extern "C" {

int wt_voxelize_grid(wt_lattice *h, const uint8_t *grid, const wt_voxmeta *vm, const double M[9],
                     const double pos[3], double scale, const int box[6], int ground, double roadVel,
                     wt_body_stats *out)
{
    WT_CHECK_H(h);
    if (!grid || !vm || !M || !pos || !box) return wt_fail(WT_E_ARG, "null argument");
    if (vm->nx < 1 || vm->ny < 1 || vm->nz < 1) return wt_fail(WT_E_ARG, "empty voxel grid");
    h->epoch++;
    const int64_t cnt = std::max(0, box[3] - box[0] + 1) * (int64_t)std::max(0, box[4] - box[1] + 1) *
                        std::max(0, box[5] - box[2] + 1);
    std::vector<uint8_t> mask((size_t)cnt);
    if (cnt > 0) {
        // grid and mask share one device buffer, kept with the lattice
        const size_t gbytes = (size_t)vm->nx * vm->ny * vm->nz, moff = (gbytes + 255) / 256 * 256;
        if (moff + (size_t)cnt > h->vgrid_cap) {
            if (h->d_vgrid) (void)hipFree(h->d_vgrid);
            h->d_vgrid = nullptr;
            h->vgrid_cap = 0;
            if (int rc = alloc((void **)&h->d_vgrid, moff + (size_t)cnt)) return rc;
            h->vgrid_cap = moff + (size_t)cnt;
        }
        uint8_t *d_grid = h->d_vgrid, *d_mask = h->d_vgrid + moff;
        int rc = 0;
        if (hipMemcpyAsync(d_grid, grid, gbytes, hipMemcpyHostToDevice, h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "grid upload failed");
        if (!rc) rc = lbm_grid_mask(h, d_grid, vm, M, pos, scale, box, d_mask);
        // the whole update on the device (env WT_VOXELIZE_HOST=1 keeps the
        // host bookkeeping)
        if (!rc && voxelize_device_applies()) return voxelize_device(h, box, d_mask, nullptr, ground, roadVel, out);
        if (!rc && hipMemcpyAsync(mask.data(), d_mask, (size_t)cnt, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "mask download failed");
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "grid lookup failed");
        if (rc) return rc;
    }
    return wt_voxelize(h, box, mask.data(), nullptr, ground, roadVel, out);
}

static int run_steps(wt_lattice *h, int nsteps, int nslabs, double *forces)
{
    WT_CHECK_H(h);
    if (nsteps < 0) return wt_fail(WT_E_ARG, "negative step count");
    if (nslabs < 1 || nslabs > h->nz - 2) return wt_fail(WT_E_ARG, "slab count %d out of range", nslabs);
    // one step's slab forces must fit the device force buffer
    if (nslabs > h->forces_cap)
        return wt_fail(WT_E_ARG, "slab count %d above the limit of %d slabs", nslabs, h->forces_cap);
    const int64_t fs = fault_step();
    if (fs >= 0 && h->steps <= fs && fs < h->steps + nsteps)
        return wt_fail(WT_E_FAULT, "injected fault at step %lld", (long long)fs);
    // the lattice's device, also when the call comes from another host
    // thread (the N-API addon steps on a worker thread); the caller's
    // current device is restored on return
    struct DeviceGuard {
        int prev = -1;
        ~DeviceGuard() { if (prev >= 0) (void)hipSetDevice(prev); }
    } guard;
    int cur = -1;
    WT_HIP(hipGetDevice(&cur));
    if (cur != h->device) {
        WT_HIP(hipSetDevice(h->device));
        guard.prev = cur;
    }
    const int per = 3 * nslabs;
    const int chunk = std::max(1, h->forces_cap / nslabs);
    for (int s0 = 0; s0 < nsteps; s0 += chunk) {
        const int cnt = std::min(chunk, nsteps - s0);
        for (int s = 0; s < cnt; s++) {
            const bool macro = s0 + s == nsteps - 1;
            if (h->timing) WT_HIP(hipEventRecord(h->tev[2 * s], h->stream));
            for (int w = 0; w < nslabs; w++) {
                int bz0 = 0, bz1 = h->nz;
                if (nslabs > 1) slab_range(h->nz, w, nslabs, &bz0, &bz1);
                if (int rc = lbm_launch_step(h, bz0, bz1, macro, h->d_forces + (int64_t)s * per + 3 * w)) return rc;
            }
            if (h->bc == WT_BC_CHANNEL)
                if (int rc = lbm_channel(h, macro)) return rc;
            if (h->timing) WT_HIP(hipEventRecord(h->tev[2 * s + 1], h->stream));
            h->parity ^= 1;
            h->steps++;
        }
        if (int rc = lbm_flush_force(h)) return rc;
        std::vector<double> tmp((size_t)cnt * per);
        WT_HIP(hipMemcpyAsync(tmp.data(), h->d_forces, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, h->stream));
        WT_HIP(hipStreamSynchronize(h->stream));
        if (h->timing) {
            for (int s = 0; s < cnt; s++) {
                float t = 0;
                WT_HIP(hipEventElapsedTime(&t, h->tev[2 * s], h->tev[2 * s + 1]));
                h->kernel_ms += t;
                h->kernel_launches += nslabs;
            }
        }
        if (forces) std::memcpy(forces + (int64_t)s0 * per, tmp.data(), sizeof(double) * tmp.size());
        for (int k = 0; k < 3; k++) {
            double t = 0;
            for (int w = 0; w < nslabs; w++) t += tmp[(size_t)(cnt - 1) * per + 3 * w + k];
            h->force[k] = t;
        }
    }
    return 0;
}

int wt_step(wt_lattice *h, int nsteps, double *forces) { return run_steps(h, nsteps, 1, forces); }

int wt_set_boundary_mode(wt_lattice *h, int mode)
{
    WT_CHECK_H(h);
    if (mode != WT_BC_REFERENCE && mode != WT_BC_CHANNEL) return wt_fail(WT_E_ARG, "unknown boundary mode %d", mode);
    if (mode == WT_BC_CHANNEL && h->layout != WT_LAYOUT_AB)
        return wt_fail(WT_E_ARG, "channel boundaries need the AB layout");
    if (mode == WT_BC_CHANNEL && !h->d_ring) {
        // interior (y, z) rows next to a side wall, each once
        std::vector<int> ring;
        for (int z = 1; z < h->nz - 1; z++)
            for (int y = 1; y < h->ny - 1; y++)
                if (y == 1 || y == h->ny - 2 || z == 1 || z == h->nz - 2) { ring.push_back(y); ring.push_back(z); }
        if (int rc = alloc((void **)&h->d_ring, sizeof(int) * ring.size())) return rc;
        WT_HIP(hipMemcpy(h->d_ring, ring.data(), sizeof(int) * ring.size(), hipMemcpyHostToDevice));
        h->nring = (int64_t)ring.size() / 2;
    }
    h->bc = mode;
    h->epoch++;
    return 0;
}

int wt_step_slab(wt_lattice *h, int z0, int z1, double force[3])
{
    WT_CHECK_H(h);
    if (h->bc != WT_BC_REFERENCE) return wt_fail(WT_E_STATE, "the slab calls take the reference boundaries only");
    if (z0 < 1 || z1 > h->nz - 1 || z0 > z1)
        return wt_fail(WT_E_ARG, "slab [%d,%d) is not within the interior rows [1,%d)", z0, z1, h->nz - 1);
    double f[3] = {0, 0, 0};
    if (z1 > z0) {
        if (int rc = lbm_launch_step(h, z0, z1, true, h->d_forces)) return rc;
        if (int rc = lbm_flush_force(h)) return rc;
        WT_HIP(hipMemcpyAsync(f, h->d_forces, sizeof f, hipMemcpyDeviceToHost, h->stream));
        WT_HIP(hipStreamSynchronize(h->stream));
    }
    if (force) std::memcpy(force, f, sizeof f);
    return 0;
}

int wt_boundaries(wt_lattice *h, int z0, int z1, int doLo, int doHi)
{
    WT_CHECK_H(h);
    if (h->bc != WT_BC_REFERENCE) return wt_fail(WT_E_STATE, "the slab calls take the reference boundaries only");
    if (z0 < 0 || z1 > h->nz || z0 > z1) return wt_fail(WT_E_ARG, "rows [%d,%d) out of range", z0, z1);
    // the planes of interior rows were written by wt_step_slab with those
    // rows; the end planes belong to whoever is told to take them
    if (doLo)
        if (int rc = lbm_launch_step(h, 0, 1, true, nullptr)) return rc;
    if (doHi)
        if (int rc = lbm_launch_step(h, h->nz - 1, h->nz, true, nullptr)) return rc;
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int wt_step_slabs(wt_lattice *h, int nsteps, int nslabs, double *forces)
{
    return run_steps(h, nsteps, nslabs, forces);
}

int wt_set_timing(wt_lattice *h, int on)
{
    WT_CHECK_H(h);
    if (on && h->tev.empty()) {
        h->tev.resize(2 * (size_t)h->forces_cap, nullptr);
        for (hipEvent_t &e : h->tev) WT_HIP(hipEventCreate(&e));
    }
    h->timing = !!on;
    h->kernel_ms = 0;
    h->kernel_launches = 0;
    return 0;
}

int wt_get_timing(const wt_lattice *h, double *kernel_ms, int64_t *launches)
{
    WT_CHECK_H(h);
    if (kernel_ms) *kernel_ms = h->kernel_ms;
    if (launches) *launches = h->kernel_launches;
    return 0;
}

int wt_set_parity(wt_lattice *h, int parity)
{
    WT_CHECK_H(h);
    if (parity != 0 && parity != 1) return wt_fail(WT_E_ARG, "parity must be 0 or 1");
    h->parity = parity;
    return 0;
}

int wt_get_state(const wt_lattice *h, wt_state_t *out)
{
    WT_CHECK_H(h);
    if (!out) return wt_fail(WT_E_ARG, "null output");
    out->parity = h->parity; out->steps = h->steps;
    for (int k = 0; k < 3; k++) out->force[k] = h->force[k];
    out->nx = h->nx; out->ny = h->ny; out->nz = h->nz;
    out->storage = h->storage; out->arith = h->arith; out->layout = h->layout;
    out->block_cells = h->bsh ? 1 << h->bsh : 0;
    out->reserved = 0;
    return 0;
}

int wt_get_fields(wt_lattice *h, void *rho, void *ux, void *uy, void *uz)
{
    WT_CHECK_H(h);
    void *dst[4] = {rho, ux, uy, uz};
    const size_t bytes = (size_t)h->n * h->mac_elem;
    for (int k = 0; k < 4; k++)
        if (dst[k])
            WT_HIP(hipMemcpyAsync(dst[k], (char *)h->d_mac + k * bytes, bytes, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int wt_set_pops(wt_lattice *h, const void *in)
{
    WT_CHECK_H(h);
    if (!in) return wt_fail(WT_E_ARG, "null populations");
    const size_t cnt = 19 * (size_t)h->n;
    const size_t eb = h->storage == WT_STORE_F64 ? 8 : 4;
    void *d_in = nullptr;
    if (int rc = alloc(&d_in, cnt * eb)) return rc;
    int rc = 0;
    if (hipMemcpyAsync(d_in, in, cnt * eb, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = wt_fail(WT_E_DEVICE, "population upload failed");
    if (!rc) rc = lbm_set_pops(h, d_in);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "population write failed");
    (void)hipFree(d_in);
    return rc;
}

int wt_get_pops(wt_lattice *h, int which, void *out)
{
    WT_CHECK_H(h);
    if (!out || (which != 0 && which != 1)) return wt_fail(WT_E_ARG, "bad population request");
    const size_t cnt = 19 * (size_t)h->n;
    if (h->layout == WT_LAYOUT_ESOTERIC && which != 0)
        return wt_fail(WT_E_ARG, "an in-place lattice holds only the current populations");
    const size_t eb = h->storage == WT_STORE_F64 ? 8 : 4;
    void *d_out = nullptr;
    if (int rc = alloc(&d_out, cnt * eb)) return rc;
    int rc = lbm_get_pops(h, which, d_out);
    if (!rc && hipMemcpyAsync(out, d_out, cnt * eb, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
        rc = wt_fail(WT_E_DEVICE, "population download failed");
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "population read failed");
    (void)hipFree(d_out);
    return rc;
}

int wt_get_solid(wt_lattice *h, uint8_t *solid, uint8_t *near)
{
    WT_CHECK_H(h);
    if (solid) std::memcpy(solid, h->h_solid.data(), (size_t)h->n);
    if (near) {
        const size_t words = (size_t)h->nwx * h->ny * h->nz;
        std::vector<uint32_t> nb(words);
        WT_HIP(hipMemcpyAsync(nb.data(), h->d_nbits, words * 4, hipMemcpyDeviceToHost, h->stream));
        WT_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < h->n; i++) {
            const int64_t x = i % h->nx, row = i / h->nx;
            near[i] = (nb[(x >> 5) + h->nwx * row] >> (x & 31)) & 1u;
        }
    }
    return 0;
}

int wt_get_walls(const wt_lattice *h, float *uwx, float *uwy, float *uwz)
{
    WT_CHECK_H(h);
    float *dst[3] = {uwx, uwy, uwz};
    for (int k = 0; k < 3; k++)
        if (dst[k]) std::memcpy(dst[k], h->h_uw.data() + (size_t)k * h->n, sizeof(float) * (size_t)h->n);
    return 0;
}

int wt_time_steps(wt_lattice *h, int nsteps, double *ms)
{
    WT_CHECK_H(h);
    if (nsteps < 1 || !ms) return wt_fail(WT_E_ARG, "bad timing request");
    WT_HIP(hipEventRecord(h->ev0, h->stream));
    for (int s = 0; s < nsteps; s++) {
        if (int rc = lbm_launch_step(h, 0, h->nz, s == nsteps - 1, h->d_forces + 3 * (s % h->forces_cap))) return rc;
        if (h->bc == WT_BC_CHANNEL)
            if (int rc = lbm_channel(h, s == nsteps - 1)) return rc;
        h->parity ^= 1;
        h->steps++;
    }
    WT_HIP(hipEventRecord(h->ev1, h->stream));
    if (int rc = lbm_flush_force(h)) return rc;
    WT_HIP(hipEventSynchronize(h->ev1));
    float t = 0;
    WT_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    const int last = (nsteps - 1) % h->forces_cap;
    WT_HIP(hipMemcpyAsync(h->force, h->d_forces + 3 * last, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

}  // extern "C"
