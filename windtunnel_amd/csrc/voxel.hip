// Voxelizer: Voxel.bounds / Voxel.build (js/voxel.js:32-232) for gfx950.
//
// Grid geometry is a handful of scalars computed on the host with the
// reference's binary64 formulas. The per-triangle and per-ray work runs in
// the kernels below, in binary64 with contraction off, so the grid and the
// openness statistic come out bit-identical to the reference:
//
//   k_shell        rasteriseSurface: one thread per triangle (js/voxel.js:105-137)
//   k_bucket_*     48x48 triangle buckets per axis, filled with atomics
//                  (js/voxel.js:150-182)
//   k_ray_count    crossings per ray                       (js/voxel.js:190-217)
//   k_ray_hits     the crossings, with their triangle index
//   k_ray_fill     first 512 crossings in triangle order, numeric sort,
//                  votes between alternate pairs           (js/voxel.js:218-230)
//   k_combine      majority vote OR shell, solid/disagree counts (js/voxel.js:79-87)
//
// The reference keeps the first 512 crossings of a ray in bucket order, and
// its buckets list triangles in ascending index; selecting the 512 smallest
// triangle indices reproduces that from buckets filled in any order.

#include "wt_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int BUCKETS = 48;
constexpr double MAX_SAMPLES = 4096;
constexpr int MAX_HITS = 512;

__device__ __forceinline__ int js_toint32(double d)
{
    if (!(d == d) || isinf(d)) return 0;
    const double t = trunc(d);
    if (t > -2147483649.0 && t < 2147483648.0) return (int)t;
    double m = fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    return (int)(uint32_t)m;
}

__device__ __forceinline__ double dmax3(double a, double b, double c)
{
    const double m = a > b ? a : b;
    return m > c ? m : c;
}
__device__ __forceinline__ double dmin3(double a, double b, double c)
{
    const double m = a < b ? a : b;
    return m < c ? m : c;
}

struct Geom {
    int nx, ny, nz;
    double cell, inv, min[3];
};

__global__ void __launch_bounds__(256) k_shell(const float *P, int64_t ntri, Geom g, uint8_t *shell)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const float *o = P + t * 9;
    const double ax = o[0], ay = o[1], az = o[2], bx = o[3], by = o[4], bz = o[5];
    const double cx = o[6], cy = o[7], cz = o[8];
    const double inv = 1 / g.cell;
    const double e1 = dmax3(fabs(bx - ax), fabs(by - ay), fabs(bz - az));
    const double e2 = dmax3(fabs(cx - ax), fabs(cy - ay), fabs(cz - az));
    double n1 = ceil(e1 * inv * 2) + 1, n2 = ceil(e2 * inv * 2) + 1;
    if (n1 * n2 > MAX_SAMPLES) {
        const double k = sqrt(MAX_SAMPLES / (n1 * n2));
        const double a1 = js_toint32(n1 * k), a2 = js_toint32(n2 * k);
        n1 = a1 > 2 ? a1 : 2;
        n2 = a2 > 2 ? a2 : 2;
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
            const int gx = js_toint32((x - g.min[0]) * inv);
            const int gy = js_toint32((y - g.min[1]) * inv);
            const int gz = js_toint32((z - g.min[2]) * inv);
            if (gx < 0 || gy < 0 || gz < 0 || gx >= g.nx || gy >= g.ny || gz >= g.nz) continue;
            shell[gx + (int64_t)g.nx * (gy + (int64_t)g.ny * gz)] = 1;
        }
    }
}

struct Axis {
    int A, B, C, nA, nB, nC, nbB, nbC;
    double cell, minA, minB, minC, kB, kC;
    int64_t sA, sB, sC;
};

// Bucket rectangle of every triangle and the per-bucket counts.
__global__ void __launch_bounds__(256) k_bucket_range(const float *P, int64_t ntri, Axis a, int4 *rng,
                                                      uint32_t *count)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const float *o = P + t * 9;
    const double b0 = o[a.B], b1 = o[3 + a.B], b2 = o[6 + a.B], c0 = o[a.C], c1 = o[3 + a.C], c2 = o[6 + a.C];
    double bl = floor((dmin3(b0, b1, b2) - a.minB) * a.kB), bh = floor((dmax3(b0, b1, b2) - a.minB) * a.kB);
    double cl = floor((dmin3(c0, c1, c2) - a.minC) * a.kC), ch = floor((dmax3(c0, c1, c2) - a.minC) * a.kC);
    bl = bl < 0 ? 0 : bl; bh = bh >= a.nbB ? a.nbB - 1 : bh;
    cl = cl < 0 ? 0 : cl; ch = ch >= a.nbC ? a.nbC - 1 : ch;
    int4 r = make_int4(0, -1, 0, -1);
    if (!(bl > bh || cl > ch)) r = make_int4((int)bl, (int)bh, (int)cl, (int)ch);
    rng[t] = r;
    for (int cc = r.z; cc <= r.w; cc++)
        for (int bb = r.x; bb <= r.y; bb++) atomicAdd(&count[bb + a.nbB * cc], 1u);
}

__global__ void __launch_bounds__(256) k_bucket_fill(const int4 *rng, int64_t ntri, int nbB, uint32_t *cursor,
                                                     uint32_t *items)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const int4 r = rng[t];
    for (int cc = r.z; cc <= r.w; cc++)
        for (int bb = r.x; bb <= r.y; bb++) items[atomicAdd(&cursor[bb + nbB * cc], 1u)] = (uint32_t)t;
}

struct Ray {
    double pb, pc;
    uint32_t s, e;
};

__device__ __forceinline__ Ray ray_of(const Axis &a, const uint32_t *start, int ib, int ic)
{
    Ray r;
    r.pc = a.minC + (ic + 0.5003719) * a.cell;
    double bc = floor((r.pc - a.minC) * a.kC);
    bc = bc > 0 ? bc : 0;
    bc = bc < a.nbC - 1 ? bc : a.nbC - 1;
    r.pb = a.minB + (ib + 0.4996281) * a.cell;
    double bb = floor((r.pb - a.minB) * a.kB);
    bb = bb > 0 ? bb : 0;
    bb = bb < a.nbB - 1 ? bb : a.nbB - 1;
    const int b = (int)bb + a.nbB * (int)bc;
    r.s = start[b];
    r.e = start[b + 1];
    return r;
}

// Barycentric crossing of the ray with triangle t in the (B,C) projection;
// false when it misses (js/voxel.js:206-216).
__device__ __forceinline__ bool cross(const float *P, const Axis &a, uint32_t t, double pb, double pc, double *h)
{
    const float *o = P + (int64_t)t * 9;
    const double b0 = o[a.B], c0 = o[a.C], b1 = o[3 + a.B], c1 = o[3 + a.C], b2 = o[6 + a.B], c2 = o[6 + a.C];
    const double d = (c1 - c2) * (b0 - b2) + (b2 - b1) * (c0 - c2);
    if (d == 0) return false;
    const double id = 1 / d;
    const double l0 = ((c1 - c2) * (pb - b2) + (b2 - b1) * (pc - c2)) * id;
    if (l0 < 0 || l0 > 1) return false;
    const double l1 = ((c2 - c0) * (pb - b2) + (b0 - b2) * (pc - c2)) * id;
    if (l1 < 0 || l1 > 1) return false;
    const double l2 = 1 - l0 - l1;
    if (l2 < 0 || l2 > 1) return false;
    *h = l0 * (double)o[a.A] + l1 * (double)o[3 + a.A] + l2 * (double)o[6 + a.A];
    return true;
}

__global__ void __launch_bounds__(256) k_ray_count(const float *P, Axis a, const uint32_t *start,
                                                   const uint32_t *items, uint32_t *rcount)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= (int64_t)a.nB * a.nC) return;
    const Ray ray = ray_of(a, start, (int)(r % a.nB), (int)(r / a.nB));
    uint32_t n = 0;
    double h;
    for (uint32_t k = ray.s; k < ray.e; k++) n += cross(P, a, items[k], ray.pb, ray.pc, &h);
    rcount[r] = n;
}

__global__ void __launch_bounds__(256) k_ray_hits(const float *P, Axis a, const uint32_t *start,
                                                  const uint32_t *items, const uint64_t *roff, double *hv,
                                                  uint32_t *ht)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= (int64_t)a.nB * a.nC) return;
    const Ray ray = ray_of(a, start, (int)(r % a.nB), (int)(r / a.nB));
    uint64_t o = roff[r];
    double h;
    for (uint32_t k = ray.s; k < ray.e; k++) {
        const uint32_t t = items[k];
        if (cross(P, a, t, ray.pb, ray.pc, &h)) { hv[o] = h; ht[o] = t; o++; }
    }
}

// In-place heap sorts for the rare long rays.
__device__ void heap_sort_by_tri(double *v, uint32_t *t, int64_t n)
{
    auto sift = [&](int64_t root, int64_t end) {
        while (2 * root + 1 < end) {
            int64_t c = 2 * root + 1;
            if (c + 1 < end && t[c + 1] > t[c]) c++;
            if (t[root] >= t[c]) return;
            const uint32_t tt = t[root]; t[root] = t[c]; t[c] = tt;
            const double vv = v[root]; v[root] = v[c]; v[c] = vv;
            root = c;
        }
    };
    for (int64_t s = n / 2 - 1; s >= 0; s--) sift(s, n);
    for (int64_t e = n - 1; e > 0; e--) {
        const uint32_t tt = t[0]; t[0] = t[e]; t[e] = tt;
        const double vv = v[0]; v[0] = v[e]; v[e] = vv;
        sift(0, e);
    }
}

__device__ void sort_values(double *v, int n)
{
    if (n <= 32) {
        for (int i = 1; i < n; i++) {
            const double x = v[i];
            int j = i - 1;
            while (j >= 0 && v[j] > x) { v[j + 1] = v[j]; j--; }
            v[j + 1] = x;
        }
        return;
    }
    auto sift = [&](int root, int end) {
        while (2 * root + 1 < end) {
            int c = 2 * root + 1;
            if (c + 1 < end && v[c + 1] > v[c]) c++;
            if (!(v[root] < v[c])) return;
            const double x = v[root]; v[root] = v[c]; v[c] = x;
            root = c;
        }
    };
    for (int s = n / 2 - 1; s >= 0; s--) sift(s, n);
    for (int e = n - 1; e > 0; e--) {
        const double x = v[0]; v[0] = v[e]; v[e] = x;
        sift(0, e);
    }
}

__global__ void __launch_bounds__(256) k_ray_fill(Axis a, const uint64_t *roff, const uint32_t *rcount, double *hv,
                                                  uint32_t *ht, uint8_t *votes)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= (int64_t)a.nB * a.nC) return;
    int64_t n = rcount[r];
    if (n < 2) return;
    double *v = hv + roff[r];
    if (n > MAX_HITS) {
        heap_sort_by_tri(v, ht + roff[r], n);
        n = MAX_HITS;
    }
    sort_values(v, (int)n);
    const int ib = (int)(r % a.nB), ic = (int)(r / a.nB);
    const int64_t colBase = ib * a.sB + ic * a.sC;
    for (int k = 0; k + 1 < n; k += 2) {
        double i0 = ceil((v[k] - a.minA) / a.cell - 0.5);
        double i1 = floor((v[k + 1] - a.minA) / a.cell - 0.5);
        if (i0 < 0) i0 = 0;
        if (i1 >= a.nA) i1 = a.nA - 1;
        for (int64_t ia = (int64_t)i0; ia <= (int64_t)i1; ia++) votes[colBase + ia * a.sA]++;
    }
}

__global__ void __launch_bounds__(256) k_combine(const uint8_t *votes, const uint8_t *shell, uint8_t *grid, int64_t n,
                                                 unsigned long long *cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    unsigned long long s = 0, d = 0;
    if (i < n) {
        const int v = votes[i];
        const bool in = v >= 2 || shell[i] == 1;
        s = in;
        d = v == 1 || v == 2;
        grid[i] = in ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { s += __shfl_down(s, o, 64); d += __shfl_down(d, o, 64); }
    if ((threadIdx.x & 63) == 0 && (s || d)) { atomicAdd(&cnt[0], s); atomicAdd(&cnt[1], d); }
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem() { for (void *p : ptrs) (void)hipFree(p); }
    template <class T> int get(T **p, size_t count)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 1);
        if (e != hipSuccess) return wt_fail(WT_E_NOMEM, "hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
        ptrs.push_back(q);
        *p = (T *)q;
        return 0;
    }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

int parity_axis(const float *dP, const wt_voxmeta *m, int axis, hipStream_t st, uint8_t *d_votes)
{
    DevMem mem;
    Axis a{};
    a.A = axis; a.B = (axis + 1) % 3; a.C = (axis + 2) % 3;
    const int dim[3] = {m->nx, m->ny, m->nz};
    a.nA = dim[a.A]; a.nB = dim[a.B]; a.nC = dim[a.C];
    a.cell = m->cell; a.minA = m->min[a.A]; a.minB = m->min[a.B]; a.minC = m->min[a.C];
    const double spanB = a.nB * m->cell, spanC = a.nC * m->cell;
    a.nbB = std::min(BUCKETS, a.nB); a.nbC = std::min(BUCKETS, a.nC);
    a.kB = a.nbB / spanB; a.kC = a.nbC / spanC;
    auto stride = [&](int ax) -> int64_t { return ax == 0 ? 1 : (ax == 1 ? dim[0] : (int64_t)dim[0] * dim[1]); };
    a.sA = stride(a.A); a.sB = stride(a.B); a.sC = stride(a.C);
    const int64_t ntri = m->ntri, nb = (int64_t)a.nbB * a.nbC, nrays = (int64_t)a.nB * a.nC;

    uint32_t *d_count, *d_start, *d_items, *d_rcount, *d_ht;
    int4 *d_rng;
    uint64_t *d_roff;
    double *d_hv;
    if (int rc = mem.get(&d_count, nb)) return rc;
    if (int rc = mem.get(&d_rng, ntri)) return rc;
    WT_HIP(hipMemsetAsync(d_count, 0, nb * 4, st));
    hipLaunchKernelGGL(k_bucket_range, dim3(blocks(ntri)), dim3(256), 0, st, dP, ntri, a, d_rng, d_count);
    WT_HIP(hipGetLastError());
    std::vector<uint32_t> cnt(nb), start(nb + 1);
    WT_HIP(hipMemcpyAsync(cnt.data(), d_count, nb * 4, hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    uint64_t acc = 0;
    for (int64_t b = 0; b < nb; b++) { start[b] = (uint32_t)acc; acc += cnt[b]; }
    if (acc >= 0xFFFFFFFFull) return wt_fail(WT_E_ARG, "too many bucket entries for one axis");
    start[nb] = (uint32_t)acc;
    if (int rc = mem.get(&d_start, nb + 1)) return rc;
    if (int rc = mem.get(&d_items, acc)) return rc;
    WT_HIP(hipMemcpyAsync(d_start, start.data(), (nb + 1) * 4, hipMemcpyHostToDevice, st));
    WT_HIP(hipMemcpyAsync(d_count, start.data(), nb * 4, hipMemcpyHostToDevice, st));  // cursors
    hipLaunchKernelGGL(k_bucket_fill, dim3(blocks(ntri)), dim3(256), 0, st, d_rng, ntri, a.nbB, d_count, d_items);
    WT_HIP(hipGetLastError());

    if (int rc = mem.get(&d_rcount, nrays)) return rc;
    hipLaunchKernelGGL(k_ray_count, dim3(blocks(nrays)), dim3(256), 0, st, dP, a, d_start, d_items, d_rcount);
    WT_HIP(hipGetLastError());
    std::vector<uint32_t> rcount(nrays);
    std::vector<uint64_t> roff(nrays);
    WT_HIP(hipMemcpyAsync(rcount.data(), d_rcount, nrays * 4, hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    uint64_t tot = 0;
    for (int64_t r = 0; r < nrays; r++) { roff[r] = tot; tot += rcount[r]; }
    if (int rc = mem.get(&d_roff, nrays)) return rc;
    if (int rc = mem.get(&d_hv, tot)) return rc;
    if (int rc = mem.get(&d_ht, tot)) return rc;
    WT_HIP(hipMemcpyAsync(d_roff, roff.data(), nrays * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ray_hits, dim3(blocks(nrays)), dim3(256), 0, st, dP, a, d_start, d_items, d_roff, d_hv, d_ht);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ray_fill, dim3(blocks(nrays)), dim3(256), 0, st, a, d_roff, d_rcount, d_hv, d_ht, d_votes);
    WT_HIP(hipGetLastError());
    WT_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

extern "C" {

int wt_voxel_bounds(const float *P, int64_t len, double b[6])
{
    if ((!P && len > 0) || !b || len < 0) return wt_fail(WT_E_ARG, "bad bounds request");
    double x0 = INFINITY, y0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, z1 = -INFINITY;
    for (int64_t i = 0; i < len; i += 3) {
        const double x = P[i];
        if (x < x0) x0 = x;
        if (x > x1) x1 = x;
        if (i + 1 < len) { const double y = P[i + 1]; if (y < y0) y0 = y; if (y > y1) y1 = y; }
        if (i + 2 < len) { const double z = P[i + 2]; if (z < z0) z0 = z; if (z > z1) z1 = z; }
    }
    b[0] = x0; b[1] = y0; b[2] = z0; b[3] = x1; b[4] = y1; b[5] = z1;
    return 0;
}

int wt_voxel_plan(const float *P, int64_t len, int res, wt_voxmeta *m)
{
    if (!m || len < 0 || (!P && len > 0)) return wt_fail(WT_E_ARG, "bad voxel plan request");
    std::memset(m, 0, sizeof(*m));
    if (res <= 0) res = 96;
    const int64_t ntri = len / 9;
    if (!ntri) return wt_fail(WT_E_VOX_EMPTY, "Mesh has no triangles.");
    if (ntri > 0x7FFFFFFF) return wt_fail(WT_E_ARG, "too many triangles");
    for (int64_t i = 0; i < ntri * 9; i++)
        if (!std::isfinite(P[i])) return wt_fail(WT_E_VOX_NONFINITE, "Mesh has a non-finite vertex coordinate.");
    double b[6];
    wt_voxel_bounds(P, len, b);
    const double sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    const double span = std::max(std::max(sx, sy), sz);
    if (!(span > 0)) return wt_fail(WT_E_VOX_DEGENERATE, "Mesh is degenerate (zero size).");
    const double cell = span / res;
    const double nxd = std::max(3.0, std::ceil(sx / cell) + 4), nyd = std::max(3.0, std::ceil(sy / cell) + 4),
                 nzd = std::max(3.0, std::ceil(sz / cell) + 4);
    if (nxd * nyd * nzd > 40e6) return wt_fail(WT_E_VOX_TOO_LARGE, "Voxel grid too large; lower the resolution.");
    m->nx = (int)nxd; m->ny = (int)nyd; m->nz = (int)nzd; m->ntri = (int)ntri;
    m->cell = cell;
    m->min[0] = b[0] - (nxd * cell - sx) * 0.5;
    m->min[1] = b[1] - (nyd * cell - sy) * 0.5;
    m->min[2] = b[2] - (nzd * cell - sz) * 0.5;
    m->inv = 1 / cell;
    m->ext[0] = sx * 0.5; m->ext[1] = sy * 0.5; m->ext[2] = sz * 0.5;
    m->centre[0] = (b[0] + b[3]) * 0.5; m->centre[1] = (b[1] + b[4]) * 0.5; m->centre[2] = (b[2] + b[5]) * 0.5;
    std::memcpy(m->bbox, b, sizeof b);
    return 0;
}

int wt_voxel_build(const float *P, int64_t len, int res, wt_voxmeta *m, uint8_t *grid)
{
    if (int rc = wt_voxel_plan(P, len, res, m)) return rc;
    if (!grid) return wt_fail(WT_E_ARG, "null grid");
    const int64_t n = (int64_t)m->nx * m->ny * m->nz, ntri = m->ntri;
    Stream st;
    WT_HIP(hipStreamCreate(&st.s));
    DevMem mem;
    float *dP;
    uint8_t *d_shell, *d_votes, *d_grid;
    unsigned long long *d_cnt;
    if (int rc = mem.get(&dP, (size_t)ntri * 9)) return rc;
    if (int rc = mem.get(&d_shell, n)) return rc;
    if (int rc = mem.get(&d_votes, n)) return rc;
    if (int rc = mem.get(&d_grid, n)) return rc;
    if (int rc = mem.get(&d_cnt, 2)) return rc;
    WT_HIP(hipMemcpyAsync(dP, P, (size_t)ntri * 9 * 4, hipMemcpyHostToDevice, st.s));
    WT_HIP(hipMemsetAsync(d_shell, 0, n, st.s));
    WT_HIP(hipMemsetAsync(d_votes, 0, n, st.s));
    WT_HIP(hipMemsetAsync(d_cnt, 0, 16, st.s));
    Geom g{m->nx, m->ny, m->nz, m->cell, m->inv, {m->min[0], m->min[1], m->min[2]}};
    hipLaunchKernelGGL(k_shell, dim3(blocks(ntri)), dim3(256), 0, st.s, dP, ntri, g, d_shell);
    WT_HIP(hipGetLastError());
    for (int axis = 0; axis < 3; axis++)
        if (int rc = parity_axis(dP, m, axis, st.s, d_votes)) return rc;
    hipLaunchKernelGGL(k_combine, dim3(blocks(n)), dim3(256), 0, st.s, d_votes, d_shell, d_grid, n, d_cnt);
    WT_HIP(hipGetLastError());
    unsigned long long cnt[2];
    WT_HIP(hipMemcpyAsync(grid, d_grid, n, hipMemcpyDeviceToHost, st.s));
    WT_HIP(hipMemcpyAsync(cnt, d_cnt, 16, hipMemcpyDeviceToHost, st.s));
    WT_HIP(hipStreamSynchronize(st.s));
    m->solid = (int64_t)cnt[0];
    m->disagree = (int64_t)cnt[1];
    m->openness = m->solid ? (double)m->disagree / (double)m->solid : 0;
    return 0;
}

}  // extern "C"
