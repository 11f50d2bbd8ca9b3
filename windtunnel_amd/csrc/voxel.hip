// Voxelizer: Voxel.bounds / Voxel.build (js/voxel.js:32-102) for gfx950.
//
// Grid geometry is a handful of scalars and is computed on the host with
// the reference's binary64 formulas; the per-triangle and per-ray work runs
// in the kernels below.

#include "wt_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

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
    return wt_fail(WT_E_STATE, "GPU voxel build not available in this build");
}

}  // extern "C"
