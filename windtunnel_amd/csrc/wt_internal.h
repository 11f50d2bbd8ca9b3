// Internal declarations shared by the HIP translation units of libwindtunnel.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/windtunnel.h"

namespace wt {

// D3Q19 velocity set in the reference's slot order (js/lbm.js:28-37). The
// slot index q is part of the storage format: f[q*n + cell].
constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
enum { FLUID = 0, SOLID_OBJ = 1, SOLID_WALL = 2 };

// Populations are stored in blocks of 2^bsh cells (64 = one wavefront): the values
// of one direction for 64 consecutive cells are contiguous (coalesced per
// direction), and a block's 19 directions follow each other, so a wave's 19
// loads come from a few kilobytes of neighbouring rows and its 19 stores go
// to one contiguous 4.75 KiB run (19 planes a lattice apart cost ~11% of
// HBM bandwidth on MI355X). Any nx works; readback converts to the
// reference's f[q*n + cell] order.
// sh = 0 selects the planar layout instead: plane q starts at q*ns.
struct Lay {
    int64_t ns;  // cells per plane
    int sh;      // log2 of the cells per block, or 0 for planar
};
// Storage position of direction q within a block (or plane order). With
// WT_SLOT_ORDER 1 the directions a cell pulls from the same neighbouring
// row (same e_y, e_z) sit next to each other, so a wave's pulls from one
// row form one contiguous run: (0,0) 0 1 2 | (1,0) 3 7 10 | (-1,0) 4 8 9 |
// (0,1) 5 11 14 | (0,-1) 6 12 13 | 15 | 16 | 17 | 18. Order 0 is q itself.
#ifndef WT_SLOT_ORDER
#define WT_SLOT_ORDER 1
#endif
#if WT_SLOT_ORDER
constexpr int SLOT[Q] = {0, 1, 2, 3, 6, 9, 12, 4, 7, 8, 5, 10, 13, 14, 11, 15, 16, 17, 18};
#else
constexpr int SLOT[Q] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18};
#endif
__host__ __device__ __forceinline__ int64_t pix(int q, int64_t cell, Lay L)
{
    const int s = SLOT[q];
    if (L.sh == 0) return s * L.ns + cell;
    return (((cell >> L.sh) * Q + s) << L.sh) + (cell & ((int64_t(1) << L.sh) - 1));
}

// binary16 storage keeps (f_q - w_q) * F16_SCALE: a power of two, so the
// scaling itself is exact, sized so |f_q - w_q| up to 1.99 stays below the
// binary16 maximum while resolvable deviations stay far above its subnormals.
constexpr float F16_SCALE = 32768.0f;
constexpr float F16_INV_SCALE = 1.0f / 32768.0f;

// Threads per block along x: one block spans a whole lattice row (up to
// 1024 cells), so the +-1 x-shifted pull loads of neighbouring waves hit
// the same CU's cache instead of re-fetching the straddled line from HBM.
#ifndef WT_ROW_BLOCK_MAX
#define WT_ROW_BLOCK_MAX 1024
#endif
inline int row_block(int nx) { int b = (nx + 63) / 64 * 64; return b > WT_ROW_BLOCK_MAX ? WT_ROW_BLOCK_MAX : b; }

struct Body {
    int64_t frontalArea = 1, planformArea = 1, solidCells = 0, rideHeight = 0;
};

}  // namespace wt

struct wt_lattice {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;
    int64_t ns = 0;     // cells per population plane, n rounded up to a whole block
    int bsh = 6;        // log2 of the cells per storage block, 0 = planar planes
    int sweep = 0;      // row order of the step: bit 0 alternates z, bit 1 alternates y
    int storage = WT_STORE_F32, arith = WT_ARITH_F32, layout = WT_LAYOUT_AB;
    int pop_elem = 4;   // bytes per stored population
    int mac_elem = 4;   // bytes per macroscopic field value
    void *d_buf[2] = {nullptr, nullptr};
    uint8_t *d_solid = nullptr;   // 0 fluid, 1 body, 2 road
    uint32_t *d_sbits = nullptr;  // bit-packed solid mask, 32 cells along x per word
    uint32_t *d_nbits = nullptr;  // bit-packed near-wall mask
    int nwx = 0;                  // words per row of a bit mask
    float *d_uw = nullptr;        // wall velocity, 3 planes of n floats
    float *d_wgt = nullptr;       // gray bounce-back weights, or null
    void *ovF = nullptr, *ovG = nullptr;  // explicit step buffers (adjoint replay), AB only
    int replay = 0;                       // replayed steps leave the macroscopic fields alone
    int64_t seed_step = -1;               // >= 0: the step whose rounding seed a replay uses
    void *adj = nullptr;                  // recorded adjoint forward pass (adjoint.hip), or null
    int64_t adj_seq = 0;                  // id of the last recorded forward pass
    int device = 0;                       // HIP device the lattice was created on
    int pairs = 1;                        // two cells per thread: 0 never, 1 binary16, 2 binary16 and float32
    uint64_t epoch = 0;                   // bumped by every change of parameters, body or weights
    void *d_mac = nullptr;        // rho, ux, uy, uz planes of n values
    double *d_partial = nullptr;  // per-block force partials, two buffers of partial_cap triples
    int64_t partial_cap = 0;
    int pslot = 0;                // buffer the last step launch wrote
    // force reduction still owed for the last step launch (done by the next
    // launch, or by lbm_flush_force): partials, their count, destination
    const double *pend_partial = nullptr;
    double *pend_out = nullptr;
    int64_t pend_count = 0;
    double *d_forces = nullptr;   // per-step force totals of the current batch
    int forces_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-launch timing of the step kernel (wt_set_timing)
    int timing = 0;
    std::vector<hipEvent_t> tev;  // 2 per step of a chunk
    double kernel_ms = 0;
    int64_t kernel_launches = 0;

    wt_params p{0.5035, 0.14, 0.25, 0.06, 1, 1, 0};
    int parity = 0;
    int64_t steps = 0;
    double force[3] = {0, 0, 0};
    int ground = 0;
    wt::Body body;
    // near-wall cells (interior fluid cells with a solid or gray upstream
    // neighbour) in cell order; near_z[z] is the index of the first one in
    // plane z (nz + 1 entries), so a z range is one slice of the list
    uint32_t *d_near = nullptr;
    // boundary variant (WT_BC_*); channel: the (y, z) rows next to a side wall
    int bc = 0;
    int *d_ring = nullptr;
    int64_t nring = 0;
    // staging of the visual products (wt_slice, wt_sample), grown on demand
    void *d_prod = nullptr;
    size_t prod_cap = 0;
    int64_t near_cap = 0;
    std::vector<int64_t> near_z;

    std::vector<uint8_t> h_solid;  // host mirror of the solid array
    std::vector<float> h_uw;       // host mirror of the wall velocities (3n)
    int64_t uw_cells = 0;          // cells of h_uw with a nonzero velocity
    // voxelize: the last body's box (every SOLID_OBJ cell lies in it; x0 > x1:
    // none), and the per-cell ops on the host and device, kept zero between calls
    int obj_box[6] = {1, 1, 1, 0, 0, 0};
    std::vector<uint8_t> h_ops;
    uint8_t *d_ops = nullptr;
    // voxelize uploads: the changed box of the mirrors gathered into pinned
    // host memory, uploaded at once and scattered on the device; both grown
    // on demand
    uint8_t *h_stage = nullptr;
    void *d_stage = nullptr;
    size_t stage_cap = 0;
    // device copy of a voxel grid for the grid-backed body lookup, grown on demand
    uint8_t *d_vgrid = nullptr;
    size_t vgrid_cap = 0;
    // device re-voxelize work space (projections, lowest cells, results)
    void *d_rvwork = nullptr;
    size_t rvwork_cap = 0;
    // a plane range of the near-wall list is rebuilt on the device into the
    // second buffer (d_near2, same capacity), around the unchanged planes
    // copied from d_near; the two are then swapped. d_pcnt: per-plane counts
    // (nz). near_full: the next rebuild covers every plane
    uint32_t *d_near2 = nullptr;
    int *d_pcnt = nullptr;
    bool near_full = true;
};

// error reporting (capi.hip)
int wt_fail(int code, const char *fmt, ...);
#define WT_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return wt_fail(WT_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// lattice kernels (lbm.hip)
int lbm_reset(wt_lattice *h);
// apply the voxelize ops of cells [c0, c1) (and clear them in d_ops)
int lbm_apply_voxelize(wt_lattice *h, uint8_t *d_ops, int64_t c0, int64_t c1);
int lbm_rebuild_masks(wt_lattice *h);
int lbm_set_pops(wt_lattice *h, const void *d_in);
// masks of planes z0..z1 only, and the near-wall list of those planes
// spliced into the rest of it
int lbm_rebuild_masks_range(wt_lattice *h, int z0, int z1);
// box[0..5] = x0, y0, z0, x1, y1, z1 (inclusive): the box's cells packed in
// x, y, z order at src (elem bytes each, elem 1 or 4) written into the
// lattice-shaped array dst
int lbm_scatter_box(wt_lattice *h, const void *src, void *dst, int elem, const int box[6]);
// wt_voxelize from a body mask already on the device (capi.hip), with
// wt_voxelize's result on the same mask. d_wv: 3 floats per mask cell (its
// mask-2 cells' wall velocities), or null. voxelize_device_applies: false
// only when the env WT_VOXELIZE_HOST asks for the host bookkeeping.
bool voxelize_device_applies();
int voxelize_device(wt_lattice *h, const int box[6], const uint8_t *d_mask, const float *d_wv, int ground,
                    double roadVel, wt_body_stats *out);
// the reverse: box cells of the lattice-shaped byte array src packed into dst
int lbm_gather_box(wt_lattice *h, const void *src, void *dst, int elem, const int box[6]);
// result slots of lbm_revoxelize_device: body cells, changed box (6), new
// body box (6), frontal and planform areas, wheel cells, change of the count
// of nonzero wall velocities, changed wall-velocity box (6)
enum { RV_CELLS = 0, RV_CB = 1, RV_NB = 7, RV_FRONTAL = 13, RV_PLAN = 14, RV_WHEELS = 15, RV_NZ = 16, RV_UB = 17,
       RV_N = 24 };
// Re-voxelize on the device from a device mask over mbox: ubox is the union
// of the previous body's box and mbox; d_wv as for voxelize_device; ground:
// the road row is wall; road: its x wall velocity. d_work
// holds ny*nz and nx*nz bytes (each rounded up to 256) and then nx*nz + 24 +
// ny ints. res (24 ints, in/out; the RV_* slots of lbm.hip: body cells,
// changed box, new body box, frontal and planform areas, wheel cells, change
// of the nonzero wall-velocity count, changed wall-velocity box) and hist
// (ny ints: (x, z) columns whose lowest body cell is at y = g + 1) come back
// on the host.
int lbm_revoxelize_device(wt_lattice *h, const int ubox[6], const int mbox[6], const uint8_t *d_mask, const float *d_wv,
                          int ground, float road, void *d_work, int *res, int *hist);
int lbm_channel(wt_lattice *h, bool macro);
// Launches one step over z rows [bz0,bz1). Its force (when force_out is
// given) is reduced into force_out by the NEXT step launch or by
// lbm_flush_force, which must run before force_out is read.
int lbm_launch_step(wt_lattice *h, int bz0, int bz1, bool macro, double *force_out);
int lbm_flush_force(wt_lattice *h);
int lbm_get_pops(wt_lattice *h, int which, void *out);
int lbm_grid_mask(wt_lattice *h, const uint8_t *grid, const wt_voxmeta *vm, const double M[9],
                  const double pos[3], double scale, const int box[6], uint8_t *mask);

// adjoint record (adjoint.hip): frees the recorded forward pass, if any
void adj_release(wt_lattice *h);

// host reference formulas shared by host and kernels
void eq_deviation(double r, double vx, double vy, double vz, double out[19]);
