// Analytic bodies on the GPU: Body.makeTest / makeWheelVel of the body
// catalogue (js/shapes.js:42-75, 365-470) evaluated at every cell centre of
// the body box, feeding voxelize (js/lbm.js:191-266).
//
// The point tests run in binary64 with contraction off, in the reference's
// operation order, so a cell is inside exactly when the reference's closure
// says so. The body and wheel constants (rotations, wheel axes, centres and
// spin rates) are computed on the host, also in the reference's order.

#include "wt_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr int MAX_PARTS = 64;

// a wheel of makeWheelVel: its part, world axis, world centre, spin rate
struct Wheel {
    int part;
    double axis[3], centre[3], omega;
};

struct PartsArgs {
    const wt_part *part;  // device copies of the parts and wheels
    const Wheel *wheel;
    int nparts, nwheels;
    double M[9], pos[3], inv;
    int x0, y0, z0, bx, by, bz;
};

// NACA 4-digit half thickness and camber line (js/shapes.js:17-27).
__device__ double naca_thickness(double xc, double t)
{
    return 5 * t * (0.2969 * sqrt(xc) - 0.1260 * xc - 0.3516 * xc * xc + 0.2843 * xc * xc * xc -
                    0.1036 * xc * xc * xc * xc);
}
__device__ double naca_camber(double xc, double m, double p)
{
    if (m == 0) return 0;
    return xc < p ? (m / (p * p)) * (2 * p * xc - xc * xc)
                  : (m / ((1 - p) * (1 - p))) * ((1 - 2 * p) + 2 * p * xc - xc * xc);
}

// insidePrim (js/shapes.js:42-75): q in primitive space.
__device__ bool inside_prim(const wt_part &p, const double q[3])
{
    const double *s = p.s;
    switch (p.type) {
    case WT_PRIM_BOX:
        return fabs(q[0]) <= s[0] && fabs(q[1]) <= s[1] && fabs(q[2]) <= s[2];
    case WT_PRIM_ELL: {
        const double a = q[0] / s[0], b = q[1] / s[1], c = q[2] / s[2];
        return a * a + b * b + c * c <= 1;
    }
    case WT_PRIM_CYL:
        return fabs(q[1]) <= s[1] && (q[0] * q[0] + q[2] * q[2]) <= s[0] * s[0];
    case WT_PRIM_CONE: {
        if (fabs(q[1]) > s[1]) return false;
        const double r = s[0] * (s[1] - q[1]) / (2 * s[1]);
        return (q[0] * q[0] + q[2] * q[2]) <= r * r;
    }
    case WT_PRIM_FOIL: {
        const double ch = s[0], span = s[2];
        if (fabs(q[2]) > span * 0.5) return false;
        const double xc = (q[0] + ch * 0.5) / ch;
        if (xc < 0 || xc > 1) return false;
        const double t = p.has_naca ? p.naca[2] : 0.12;
        const double m = p.has_naca ? p.naca[0] : 0;
        const double pp = p.has_naca ? p.naca[1] : 0.4;
        const double yt = naca_thickness(xc, t) * ch;
        const double yc = naca_camber(xc, m, pp) * ch;
        return fabs(q[1] - yc) <= yt;
    }
    }
    return false;
}

// part space of body-space point l: m^T (l - c)
__device__ void to_part(const wt_part &p, const double l[3], double q[3])
{
    const double *m = p.m, *c = p.c;
    const double ax = l[0] - c[0], ay = l[1] - c[1], az = l[2] - c[2];
    q[0] = m[0] * ax + m[3] * ay + m[6] * az;
    q[1] = m[1] * ax + m[4] * ay + m[7] * az;
    q[2] = m[2] * ax + m[5] * ay + m[8] * az;
}

// mask: 0 outside, 1 inside, 2 inside a wheel (wv = its velocity)
__global__ void __launch_bounds__(256) k_parts_mask(PartsArgs a, uint8_t *mask, float *wv)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t cnt = (int64_t)a.bx * a.by * a.bz;
    if (k >= cnt) return;
    const int x = a.x0 + (int)(k % a.bx);
    const int y = a.y0 + (int)((k / a.bx) % a.by);
    const int z = a.z0 + (int)(k / ((int64_t)a.bx * a.by));
    const double X = x + 0.5, Y = y + 0.5, Z = z + 0.5;
    const double dx = X - a.pos[0], dy = Y - a.pos[1], dz = Z - a.pos[2];
    const double *M = a.M;
    const double l[3] = {(M[0] * dx + M[3] * dy + M[6] * dz) * a.inv, (M[1] * dx + M[4] * dy + M[7] * dz) * a.inv,
                         (M[2] * dx + M[5] * dy + M[8] * dz) * a.inv};
    uint8_t in = 0;
    double q[3];
    for (int i = 0; i < a.nparts && !in; i++) {
        to_part(a.part[i], l, q);
        if (inside_prim(a.part[i], q)) in = 1;
    }
    if (in) {
        for (int w = 0; w < a.nwheels; w++) {
            const Wheel &wh = a.wheel[w];
            const wt_part &p = a.part[wh.part];
            to_part(p, l, q);
            if (!inside_prim(p, q)) continue;
            // u = omega (axis x r), r from the wheel centre (js/shapes.js:458-466)
            const double rx = X - wh.centre[0], ry = Y - wh.centre[1], rz = Z - wh.centre[2];
            const double *ax = wh.axis, o = wh.omega;
            wv[3 * k] = (float)(o * (ax[1] * rz - ax[2] * ry));
            wv[3 * k + 1] = (float)(o * (ax[2] * rx - ax[0] * rz));
            wv[3 * k + 2] = (float)(o * (ax[0] * ry - ax[1] * rx));
            in = 2;
            break;
        }
    }
    mask[k] = in;
}

}  // namespace

extern "C" int wt_voxelize_parts(wt_lattice *h, const wt_part *parts, int nparts, const double M[9],
                                 const double pos[3], double scale, const int box[6], int ground, double roadVel,
                                 double wheel_u0, wt_body_stats *out)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    if (!M || !pos || !box || (nparts > 0 && !parts)) return wt_fail(WT_E_ARG, "null argument");
    if (nparts < 0 || nparts > MAX_PARTS) return wt_fail(WT_E_ARG, "part count %d outside [0, %d]", nparts, MAX_PARTS);
    if (!(scale > 0)) return wt_fail(WT_E_ARG, "scale must be positive");
    for (int i = 0; i < nparts; i++)
        if (parts[i].type < WT_PRIM_BOX || parts[i].type > WT_PRIM_FOIL)
            return wt_fail(WT_E_ARG, "part %d has unknown type %d", i, parts[i].type);
    PartsArgs a{};
    a.nparts = nparts;
    std::vector<Wheel> wheels;
    std::memcpy(a.M, M, sizeof a.M);
    std::memcpy(a.pos, pos, sizeof a.pos);
    a.inv = 1 / scale;
    // makeWheelVel (js/shapes.js:426-447): world axis (sense along +z),
    // world centre and spin rate of every wheel part, in part order
    if (wheel_u0 > 0) {
        for (int i = 0; i < nparts; i++) {
            if (!parts[i].wheel) continue;
            const double *m = parts[i].m, *c = parts[i].c;
            double ab[3] = {m[1], m[4], m[7]};
            if (ab[2] < 0) { ab[0] = -ab[0]; ab[1] = -ab[1]; ab[2] = -ab[2]; }
            Wheel w{};
            w.part = i;
            for (int k = 0; k < 3; k++) {
                w.axis[k] = M[k * 3] * ab[0] + M[k * 3 + 1] * ab[1] + M[k * 3 + 2] * ab[2];
                w.centre[k] = pos[k] + scale * (M[k * 3] * c[0] + M[k * 3 + 1] * c[1] + M[k * 3 + 2] * c[2]);
            }
            const double R = parts[i].s[0] * scale;
            w.omega = R > 0 ? wheel_u0 / R : 0;
            wheels.push_back(w);
        }
    }
    a.x0 = box[0]; a.y0 = box[1]; a.z0 = box[2];
    a.bx = std::max(0, box[3] - box[0] + 1); a.by = std::max(0, box[4] - box[1] + 1); a.bz = std::max(0, box[5] - box[2] + 1);
    const int64_t cnt = (int64_t)a.bx * a.by * a.bz;
    std::vector<uint8_t> mask((size_t)cnt);
    std::vector<float> wvc;
    a.nwheels = (int)wheels.size();
    if (cnt > 0) {
        uint8_t *d_mask = nullptr;
        float *d_wv = nullptr;
        void *d_consts = nullptr;
        const size_t pbytes = sizeof(wt_part) * (size_t)std::max(1, nparts);
        const size_t cbytes = pbytes + sizeof(Wheel) * std::max<size_t>(1, wheels.size());
        if (hipMalloc((void **)&d_mask, (size_t)cnt) != hipSuccess) return wt_fail(WT_E_NOMEM, "mask allocation failed");
        if (hipMalloc((void **)&d_wv, sizeof(float) * 3 * (size_t)cnt) != hipSuccess ||
            hipMalloc(&d_consts, cbytes) != hipSuccess) {
            (void)hipFree(d_mask);
            if (d_wv) (void)hipFree(d_wv);
            return wt_fail(WT_E_NOMEM, "body buffer allocation failed");
        }
        a.part = (const wt_part *)d_consts;
        a.wheel = (const Wheel *)((char *)d_consts + pbytes);
        std::vector<float> wv;
        int rc = 0;
        if (nparts && hipMemcpyAsync(d_consts, parts, sizeof(wt_part) * nparts, hipMemcpyHostToDevice, h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "part upload failed");
        if (!rc && !wheels.empty() &&
            hipMemcpyAsync((char *)d_consts + pbytes, wheels.data(), sizeof(Wheel) * wheels.size(), hipMemcpyHostToDevice,
                           h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "wheel upload failed");
        if (!rc) hipLaunchKernelGGL(k_parts_mask, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, a, d_mask, d_wv);
        if (!rc && hipGetLastError() != hipSuccess) rc = wt_fail(WT_E_DEVICE, "body kernel launch failed");
        if (!rc && voxelize_device_applies()) {
            // the whole update on the device
            rc = voxelize_device(h, box, d_mask, a.nwheels ? d_wv : nullptr, ground, roadVel, out);
            (void)hipFree(d_mask);
            (void)hipFree(d_wv);
            (void)hipFree(d_consts);
            return rc;
        }
        if (!rc && hipMemcpyAsync(mask.data(), d_mask, (size_t)cnt, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
            rc = wt_fail(WT_E_DEVICE, "mask download failed");
        if (!rc && a.nwheels) {
            wv.resize(3 * (size_t)cnt);
            if (hipMemcpyAsync(wv.data(), d_wv, sizeof(float) * wv.size(), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
                rc = wt_fail(WT_E_DEVICE, "wall-velocity download failed");
        }
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = wt_fail(WT_E_DEVICE, "body kernel failed");
        (void)hipFree(d_mask);
        (void)hipFree(d_wv);
        (void)hipFree(d_consts);
        if (rc) return rc;
        // voxelize takes the wall velocities of mask-2 cells in box order
        for (int64_t k = 0; k < cnt; k++)
            if (mask[k] == 2) wvc.insert(wvc.end(), {wv[3 * k], wv[3 * k + 1], wv[3 * k + 2]});
    }
    return wt_voxelize(h, box, mask.data(), wvc.empty() ? nullptr : wvc.data(), ground, roadVel, out);
}
