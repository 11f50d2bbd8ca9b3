// Visual data products read on the device from the lattice's macroscopic
// fields, so a viewer downloads one plane or a list of probe velocities
// instead of the whole rho/u fields:
//  - the app's slice view (js/main.js:925-965): speed, C_p of the
//    neighbour-averaged density, or the in-plane vorticity on a z or y plane;
//  - sample() (js/lbm.js:515-539) at many points at once (particles, ribbons).
// Binary64 arithmetic in the reference's expression order (the library is
// compiled with contraction off).

#include "wt_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace {

template <class MAC> struct SliceArgs {
    const MAC *rho, *ux, *uy, *uz;
    const uint8_t *solid;
    float *out;
    uint8_t *solid_out;
    int nx, ny, nz, axis, k, field, h;
    double u0;
};

// Cell index of in-plane position (a, b): plane z = k (axis 0) or y = k (axis 1).
template <class MAC> __device__ __forceinline__ int64_t at(const SliceArgs<MAC> &s, int a, int b)
{
    return s.axis == 0 ? a + (int64_t)s.nx * (b + (int64_t)s.ny * s.k) : a + (int64_t)s.nx * (s.k + (int64_t)s.ny * b);
}

template <class MAC> __global__ void __launch_bounds__(256) k_slice(SliceArgs<MAC> s)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)s.nx * s.h) return;
    const int a = (int)(t % s.nx), b = (int)(t / s.nx);
    const int64_t i = at(s, a, b);
    const bool sol = s.solid[i] != 0;
    if (s.solid_out) s.solid_out[t] = s.solid[i];
    if (sol) {
        s.out[t] = 0.0f;
        return;
    }
    // in-plane neighbours, the cell itself at the plane's edges
    const int64_t ip = a < s.nx - 1 ? at(s, a + 1, b) : i, im = a > 0 ? at(s, a - 1, b) : i;
    const int64_t jp = b < s.h - 1 ? at(s, a, b + 1) : i, jm = b > 0 ? at(s, a, b - 1) : i;
    double v;
    if (s.field == 0) {
        const double x = s.ux[i], y = s.uy[i], z = s.uz[i];
        v = sqrt(x * x + y * y + z * z) / s.u0;
    } else if (s.field == 1) {
        const double r = ((double)s.rho[i] * 2 + (double)s.rho[ip] + (double)s.rho[im] + (double)s.rho[jp] +
                          (double)s.rho[jm]) / 6;
        v = ((r - 1) / 3) / (0.5 * s.u0 * s.u0);
    } else if (s.axis == 0) {
        v = ((double)s.uy[ip] - (double)s.uy[im]) * 0.5 - ((double)s.ux[jp] - (double)s.ux[jm]) * 0.5;
    } else {
        v = ((double)s.ux[jp] - (double)s.ux[jm]) * 0.5 - ((double)s.uz[ip] - (double)s.uz[im]) * 0.5;
    }
    s.out[t] = (float)v;
}

// ECMAScript ToInt32 (the `|0` operator).
__device__ __forceinline__ int toint32(double d)
{
    if (!(d == d) || isinf(d)) return 0;
    const double t = trunc(d);
    if (t > -2147483649.0 && t < 2147483648.0) return (int)t;
    double m = fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    return (int)(uint32_t)m;
}

template <class MAC>
__global__ void __launch_bounds__(256) k_sample(const MAC *ux, const MAC *uy, const MAC *uz, const double *xyz,
                                                double *uvw, int64_t count, int nx, int ny, int nz, double u0)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= count) return;
    const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
    double *o = uvw + 3 * p;
    if (x < 0.5 || y < 0.5 || z < 0.5 || x > nx - 1.5 || y > ny - 1.5 || z > nz - 1.5) {
        o[0] = u0; o[1] = 0; o[2] = 0;
        return;
    }
    const double xf = x - 0.5, yf = y - 0.5, zf = z - 0.5;
    const int x0 = toint32(xf), y0 = toint32(yf), z0 = toint32(zf);
    const double tx = xf - x0, ty = yf - y0, tz = zf - z0;
    double vx = 0, vy = 0, vz = 0;
    for (int k = 0; k < 2; k++) {
        const double wz = k ? tz : 1 - tz;
        for (int j = 0; j < 2; j++) {
            const double wy = (j ? ty : 1 - ty) * wz;
            const int64_t ia = x0 + (int64_t)nx * ((y0 + j) + (int64_t)ny * (z0 + k)), ib = ia + 1;
            const double wa = (1 - tx) * wy, wb = tx * wy;
            vx += (double)ux[ia] * wa + (double)ux[ib] * wb;
            vy += (double)uy[ia] * wa + (double)uy[ib] * wb;
            vz += (double)uz[ia] * wa + (double)uz[ib] * wb;
        }
    }
    o[0] = vx; o[1] = vy; o[2] = vz;
}

// Device staging of the products, grown on demand and kept with the lattice.
int staging(wt_lattice *h, size_t bytes, void **p)
{
    if (bytes > h->prod_cap) {
        if (h->d_prod) WT_HIP(hipFree(h->d_prod));
        h->d_prod = nullptr;
        h->prod_cap = 0;
        if (hipMalloc(&h->d_prod, bytes) != hipSuccess) {
            h->d_prod = nullptr;
            return wt_fail(WT_E_NOMEM, "hipMalloc of %zu bytes of product staging failed", bytes);
        }
        h->prod_cap = bytes;
    }
    *p = h->d_prod;
    return 0;
}

template <class MAC> int slice_t(wt_lattice *h, int axis, int k, int field, float *out, uint8_t *solid)
{
    SliceArgs<MAC> s{};
    const MAC *m = (const MAC *)h->d_mac;
    s.rho = m; s.ux = m + h->n; s.uy = m + 2 * h->n; s.uz = m + 3 * h->n;
    s.solid = h->d_solid;
    s.nx = h->nx; s.ny = h->ny; s.nz = h->nz; s.axis = axis; s.k = k; s.field = field;
    s.h = axis == 0 ? h->ny : h->nz;
    s.u0 = h->p.u0;
    const int64_t cnt = (int64_t)s.nx * s.h;
    void *buf = nullptr;
    const size_t fb = sizeof(float) * (size_t)cnt, sb = solid ? (size_t)cnt : 0;
    if (int rc = staging(h, fb + sb, &buf)) return rc;
    s.out = (float *)buf;
    s.solid_out = solid ? (uint8_t *)buf + fb : nullptr;
    hipLaunchKernelGGL(k_slice<MAC>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, s);
    WT_HIP(hipGetLastError());
    WT_HIP(hipMemcpyAsync(out, s.out, fb, hipMemcpyDeviceToHost, h->stream));
    if (solid) WT_HIP(hipMemcpyAsync(solid, s.solid_out, sb, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

template <class MAC> int sample_t(wt_lattice *h, const double *xyz, int64_t count, double *uvw)
{
    const MAC *m = (const MAC *)h->d_mac;
    void *buf = nullptr;
    const size_t b = sizeof(double) * 3 * (size_t)count;
    if (int rc = staging(h, 2 * b, &buf)) return rc;
    double *dx = (double *)buf, *du = dx + 3 * count;
    WT_HIP(hipMemcpyAsync(dx, xyz, b, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_sample<MAC>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, m + h->n,
                       m + 2 * h->n, m + 3 * h->n, (const double *)dx, du, count, h->nx, h->ny, h->nz, h->p.u0);
    WT_HIP(hipGetLastError());
    WT_HIP(hipMemcpyAsync(uvw, du, b, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

// ------------------------------------------------- no-allocation staging
// The device staging of wt_sample_noalloc: WT_SAMPLE_NOALLOC_MAX points in
// and as many velocities out (12 MiB), held by every GPU the library runs
// on until the process ends. Host code never uses this name as an address:
// each device's copy lives at its own address, looked up once per device
// with hipGetSymbolAddress and kept in scratch_dev.
static __device__ __align__(256) double scratch[6 * (size_t)WT_SAMPLE_NOALLOC_MAX];

constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_DEVICES];  // address of `scratch` on each device, null until looked up

// The device address of `scratch` on device dev, which must be current.
int scratch_of(int dev, double **out)
{
    if (dev < 0 || dev >= SCRATCH_DEVICES) return wt_fail(WT_E_DEVICE, "device %d beyond the %d supported", dev, SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return wt_fail(WT_E_DEVICE, "hipGetSymbolAddress of the sample staging failed on device %d: %s", dev,
                           hipGetErrorString(e));
        scratch_dev[dev] = (double *)p;
    }
    *out = scratch_dev[dev];
    return 0;
}

template <class MAC> int sample_noalloc_t(wt_lattice *h, double *stage, const double *xyz, int64_t count, double *uvw)
{
    const MAC *m = (const MAC *)h->d_mac;
    const size_t b = sizeof(double) * 3 * (size_t)count;
    double *dx = stage, *du = stage + 3 * (size_t)WT_SAMPLE_NOALLOC_MAX;
    WT_HIP(hipMemcpyAsync(dx, xyz, b, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_sample<MAC>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, m + h->n,
                       m + 2 * h->n, m + 3 * h->n, (const double *)dx, du, count, h->nx, h->ny, h->nz, h->p.u0);
    WT_HIP(hipGetLastError());
    WT_HIP(hipMemcpyAsync(uvw, du, b, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

}  // namespace

extern "C" {

int wt_slice(wt_lattice *h, int axis, int k, int field, float *out, uint8_t *solid)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    if (!out) return wt_fail(WT_E_ARG, "null output");
    if (axis != 0 && axis != 1) return wt_fail(WT_E_ARG, "slice axis must be 0 (plane z = k) or 1 (plane y = k)");
    if (field < 0 || field > 2) return wt_fail(WT_E_ARG, "slice field must be 0 (speed), 1 (C_p) or 2 (vorticity)");
    const int lim = axis == 0 ? h->nz : h->ny;
    if (k < 0 || k >= lim) return wt_fail(WT_E_ARG, "plane %d outside [0, %d)", k, lim);
    if (h->storage == WT_STORE_F64) return slice_t<double>(h, axis, k, field, out, solid);
    return slice_t<float>(h, axis, k, field, out, solid);
}

int wt_sample(wt_lattice *h, const double *xyz, int64_t count, double *uvw)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    if (count < 0 || (count > 0 && (!xyz || !uvw))) return wt_fail(WT_E_ARG, "bad sample request");
    if (count == 0) return 0;
    if (h->storage == WT_STORE_F64) return sample_t<double>(h, xyz, count, uvw);
    return sample_t<float>(h, xyz, count, uvw);
}

int wt_sample_noalloc(wt_lattice *h, const double *xyz, int64_t count, double *uvw)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    if (count < 0 || (count > 0 && (!xyz || !uvw))) return wt_fail(WT_E_ARG, "bad sample request");
    if (count > WT_SAMPLE_NOALLOC_MAX)
        return wt_fail(WT_E_ARG, "%lld points exceed the %d of the no-allocation sample", (long long)count,
                       WT_SAMPLE_NOALLOC_MAX);
    if (count == 0) return 0;
    int dev = -1;
    WT_HIP(hipGetDevice(&dev));
    if (dev != h->device)
        return wt_fail(WT_E_STATE, "the lattice lives on device %d, device %d is current", h->device, dev);
    double *stage = nullptr;
    if (int rc = scratch_of(dev, &stage)) return rc;
    if (h->storage == WT_STORE_F64) return sample_noalloc_t<double>(h, stage, xyz, count, uvw);
    return sample_noalloc_t<float>(h, stage, xyz, count, uvw);
}

}  // extern "C"
