/* windtunnel.h -- C ABI of the MI355X-native D3Q19 lattice-Boltzmann and
 * voxelizer library (libwindtunnel.so, HIP for gfx950).
 *
 * The ABI is the boundary a host binding (the Node N-API addon in
 * windtunnel_amd/addon, the Python ctypes binding in windtunnel_amd/lbm.py)
 * calls to replace the reference's solver object and voxelizer:
 *
 *   js/lbm.js:51-546     class LBM            -> wt_create ... wt_destroy
 *   js/lbmWorker.js:35-80 worker protocol     -> wt_step_slabs, wt_set_parity
 *   js/pool.js:97-125    Pool.step totals     -> wt_step_slabs (per-slab forces)
 *   js/voxel.js:32-102   Voxel.bounds/build   -> wt_voxel_bounds, wt_voxel_build
 *   js/shapes.js:369-388 grid-backed makeTest -> wt_voxelize_grid
 *
 * Conventions: every function returns 0 on success and a negative WT_E*
 * code on failure; wt_last_error() then names the failure (thread-local
 * text, valid until the next failing call on the same thread). Handles are
 * not thread-safe: one host thread drives one lattice. Arrays are plain
 * host pointers; the library copies to and from its own device memory.
 */
#ifndef WINDTUNNEL_H
#define WINDTUNNEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WT_ABI_VERSION 1

enum {
    WT_OK = 0,
    WT_E_ARG = -1,      /* bad argument (dimension, size, enum, null pointer) */
    WT_E_DEVICE = -2,   /* HIP runtime error, or no usable gfx950 device */
    WT_E_NOMEM = -3,    /* device or host allocation failed */
    WT_E_STATE = -4,    /* call not valid in the handle's current state */
    WT_E_FAULT = -5,    /* injected fault (WT_FAULT environment variable) */
    WT_E_VOX_EMPTY = -11,      /* "Mesh has no triangles." (js/voxel.js:51) */
    WT_E_VOX_DEGENERATE = -12, /* "Mesh is degenerate (zero size)." (js/voxel.js:56) */
    WT_E_VOX_TOO_LARGE = -13,  /* "Voxel grid too large; lower the resolution." (js/voxel.js:71) */
    WT_E_VOX_NONFINITE = -14   /* a vertex coordinate is NaN or infinite */
};

/* Population storage. F32 is the reference's Float32Array of deviations
 * f_q - w_q (js/lbm.js:74-75,145-162); F16 stores the same deviations times
 * 2^15 as binary16 (|f_q - w_q| up to 65504 * 2^-15, about 2.0: collision
 * results round stochastically and saturate at that bound); F64 stores
 * binary64 (the FP64 oracle run). */
enum { WT_STORE_F32 = 0, WT_STORE_F16 = 1, WT_STORE_F64 = 2 };

/* Collision arithmetic. F64 evaluates the reference's binary64 expression
 * tree with contraction off (with F32 storage: the reference's own numbers);
 * F32 is the fast path. */
enum { WT_ARITH_F32 = 0, WT_ARITH_F64 = 1 };

/* Population layout in HBM. AB = two buffers swapped by parity, the
 * reference's scheme (js/lbm.js:71-80,130-132); ESOTERIC = one buffer
 * streamed in place (esoteric pull). */
enum { WT_LAYOUT_AB = 0, WT_LAYOUT_ESOTERIC = 1 };

typedef struct wt_lattice wt_lattice;

/* Solver knobs of js/lbm.js:106-111 (settable properties of LBM). */
typedef struct {
    double tau;      /* BGK / symmetric relaxation time tau0 */
    double csm;      /* Smagorinsky constant */
    double lambda;   /* TRT magic parameter */
    double u0;       /* inflow speed, lattice units */
    int les;         /* Smagorinsky eddy viscosity on/off */
    int trt;         /* two-relaxation-time collision on/off (off = BGK) */
    int movingWalls; /* apply the moving-wall bounce-back correction */
} wt_params;

/* Results of LBM.voxelize (js/lbm.js:247-265). */
typedef struct {
    int64_t frontalArea, planformArea, solidCells, rideHeight;
    int32_t movingWalls, ground;
} wt_body_stats;

/* Observable scalar state of an LBM object (js/lbm.js:80,113,117), plus
 * its geometry and types as created, and the device storage layout:
 * block_cells = cells per storage block of the population buffers (0 = one
 * plane per direction). */
typedef struct {
    int64_t parity, steps;
    double force[3];
    int32_t nx, ny, nz;
    int32_t storage, arith, layout;
    int32_t block_cells, reserved;
} wt_state_t;

/* Voxel grid geometry and statistics: the object Voxel.build returns
 * (js/voxel.js:89-101), minus the grid itself. */
typedef struct {
    int32_t nx, ny, nz, ntri;
    double cell, inv, min[3], ext[3], centre[3], bbox[6];
    int64_t solid, disagree;
    double openness;
} wt_voxmeta;

const char *wt_last_error(void);
int wt_abi_version(void);

/* new LBM(nx,ny,nz) including its reset() (js/lbm.js:58-128). */
int wt_create(int nx, int ny, int nz, int storage, int arith, int layout, wt_lattice **out);
int wt_destroy(wt_lattice *h);

/* Property setters/getters of js/lbm.js:106-111 (and the worker 'params'
 * message, js/lbmWorker.js:53-56). */
int wt_set_params(wt_lattice *h, const wt_params *p);
int wt_get_params(const wt_lattice *h, wt_params *p);

/* reset(): equilibrium at u0 in fluid, at rest in solid, both buffers
 * (js/lbm.js:164-174). */
int wt_reset(wt_lattice *h);

/* voxelize(inside, box, ground, wallVel, roadVel) (js/lbm.js:191-266), with
 * the host closures already evaluated: box = the clamped box
 * {x0,y0,z0,x1,y1,z1} (inclusive, within [1,n-2] per axis), mask one byte
 * per box cell in x-fastest order (0 outside, 1 inside, 2 inside with a wall
 * velocity), wv three floats per mask==2 cell in the same order. */
int wt_voxelize(wt_lattice *h, const int box[6], const uint8_t *mask, const float *wv, int ground,
                double roadVel, wt_body_stats *out);

/* Grid-backed body (js/shapes.js:369-388): the voxel lookup is evaluated on
 * the GPU for every cell centre of the clamped box, then voxelize proceeds
 * as above without wall velocities. grid is nx*ny*nz bytes of a voxel grid
 * (wt_voxel_build), M the 3x3 body->world rotation (row-major), pos the body
 * origin in lattice cells, scale cells per body unit. */
int wt_voxelize_grid(wt_lattice *h, const uint8_t *grid, const wt_voxmeta *vm, const double M[9],
                     const double pos[3], double scale, const int box[6], int ground,
                     double roadVel, wt_body_stats *out);

/* One analytic primitive of a body (js/shapes.js:42-75 insidePrim, the
 * catalogue at :77-169). c: centre in body units; s: half-extents (box),
 * radii (ellipsoid), [radius, half-height] (cylinder and cone, axis local
 * +y, cone apex at +y), [chord, -, span] (foil); m: primitive -> body
 * rotation, row-major as eulerMat builds it (:29-40); naca: 4-digit
 * [camber m, position p, thickness t] of a foil when has_naca (else the
 * defaults 0, 0.4, 0.12); wheel: one of makeWheelVel's wheels. */
enum { WT_PRIM_BOX = 0, WT_PRIM_ELL = 1, WT_PRIM_CYL = 2, WT_PRIM_CONE = 3, WT_PRIM_FOIL = 4 };
typedef struct {
    int32_t type, wheel, has_naca, reserved;
    double c[3], s[3], m[9], naca[3];
} wt_part;

/* voxelize(body.makeTest(), box, ground, body.makeWheelVel(u0), roadVel)
 * for an analytic body (js/shapes.js:365-470 with js/lbm.js:191-266): the
 * union of the parts is tested on the GPU at every cell centre of the
 * clamped box (world -> body by M^T (X - pos) / scale, then each part's
 * m^T), in binary64 in the reference's operation order. wheel_u0 > 0 gives
 * the cells of wheel parts the rolling velocity of makeWheelVel(wheel_u0);
 * 0 gives no wall velocities. Then voxelize proceeds as wt_voxelize. */
int wt_voxelize_parts(wt_lattice *h, const wt_part *parts, int nparts, const double M[9], const double pos[3],
                      double scale, const int box[6], int ground, double roadVel, double wheel_u0,
                      wt_body_stats *out);

/* step() nsteps times (js/lbm.js:285-291). forces (may be NULL) receives the
 * 3 force components after each step, FP64, nsteps*3 doubles. */
int wt_step(wt_lattice *h, int nsteps, double *forces);

/* stepSlab(z0,z1) (js/lbm.js:293-465): cells of the z rows [z0,z1),
 * 1 <= z0 <= z1 <= nz-1, pull from the current populations and write the
 * next ones; the x/y boundary planes and the outlet copy of those rows are
 * written in the same launch (the step is fused), so _boundaries only has
 * the z = 0 and z = nz-1 planes left. No parity flip and no step count, as
 * in the reference; force receives the slab's force. */
int wt_step_slab(wt_lattice *h, int z0, int z1, double force[3]);

/* _boundaries(g, z0, z1, doLo, doHi) (js/lbm.js:467-512) after wt_step_slab:
 * writes the z = 0 plane when doLo and the z = nz-1 plane when doHi, with
 * their outlet copies; the planes of the interior rows of [z0,z1) came with
 * wt_step_slab. */
int wt_boundaries(wt_lattice *h, int z0, int z1, int doLo, int doHi);

/* Worker-pool form (js/pool.js:97-125, js/lbmWorker.js:43-79): the
 * interior is split into nslabs z-slabs as the worker 'init' handler does
 * and each step runs slab by slab; forces (may be NULL) receives
 * nsteps*nslabs*3 doubles, [step][slab][xyz], the per-worker totals the
 * pool sums in worker order. */
int wt_step_slabs(wt_lattice *h, int nsteps, int nslabs, double *forces);

/* The worker 'step' message carries the page's parity (js/lbmWorker.js:60). */
int wt_set_parity(wt_lattice *h, int parity);
int wt_get_state(const wt_lattice *h, wt_state_t *out);

/* Macroscopic fields rho, ux, uy, uz (js/lbm.js:85-88,385): element type is
 * float, or double for WT_STORE_F64; any pointer may be NULL. The values are
 * those of the last completed step (they are written on the last step of
 * every wt_step / wt_step_slabs call). */
int wt_get_fields(wt_lattice *h, void *rho, void *ux, void *uy, void *uz);

/* Populations f (which=0, current) or g (which=1) in the reference's AB
 * order f[q*n + cell] as deviations f_q - w_q: float, or double for
 * WT_STORE_F64 (F16 storage is decoded to float). With the esoteric layout
 * only which=0 is available; solid cells read as 0 there. */
int wt_get_pops(wt_lattice *h, int which, void *out);

/* The inverse, for restoring a saved state (no reference counterpart): the
 * current populations from f[q*n + cell] deviations in the same types
 * (binary16 storage rounds the floats to nearest). In place, values of
 * solid cells and of slots pointing out of the lattice are not stored. The
 * macroscopic fields keep their values until the next step computes them.
 * Binary16 storage rounds collision results with bits drawn from the step
 * count, so a restored binary16 run repeats the original bit for bit only
 * at the same step count. */
int wt_set_pops(wt_lattice *h, const void *in);

/* solid (0 fluid, 1 body, 2 road) and near-wall flags (js/lbm.js:82-83). */
int wt_get_solid(wt_lattice *h, uint8_t *solid, uint8_t *near);

/* Wall velocities uwx, uwy, uwz (js/lbm.js:90-93, set by voxelize): n
 * floats each; any pointer may be NULL. */
int wt_get_walls(const wt_lattice *h, float *uwx, float *uwy, float *uwz);

/* The app's slice view (js/main.js:925-965) computed on the device from the
 * fields of the last step, for plane z = k (axis 0, the 'xy' view) or
 * y = k (axis 1, the 'xz' view): nx * h floats, x fastest, h = ny or nz.
 * field 0: speed |u| / u0; 1: C_p = ((r - 1) / 3) / (u0^2 / 2) of the density
 * averaged with its four in-plane neighbours (weights 2,1,1,1,1 over 6);
 * 2: the in-plane vorticity component, central differences (the cell
 * itself stands in for a missing neighbour at the plane's edge). Solid
 * cells read 0; solid (may be NULL) receives the plane's solid map. */
int wt_slice(wt_lattice *h, int axis, int k, int field, float *out, uint8_t *solid);

/* sample() (js/lbm.js:515-539) at count points: trilinear velocity from
 * the fields of the last step at xyz[3p..3p+2], u0 outside the cell-centre
 * range, into uvw[3p..3p+2]. */
int wt_sample(wt_lattice *h, const double *xyz, int64_t count, double *uvw);

/* wt_sample without any device allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture, tight serving loops (particles and probes sampled every
 * frame) and harnesses that forbid hipMalloc. Same inputs, outputs and
 * results bit for bit as wt_sample, but the device staging of the points
 * and velocities is a fixed array of the library on each GPU, so count may
 * be at most WT_SAMPLE_NOALLOC_MAX (WT_E_ARG beyond it). Calls that overlap
 * on one GPU (from several threads) race on that array: serialise them.
 * The lattice's device must be the current device (WT_E_STATE otherwise). */
#define WT_SAMPLE_NOALLOC_MAX 262144
int wt_sample_noalloc(wt_lattice *h, const double *xyz, int64_t count, double *uvw);

/* MeshIO.parse (js/meshio.js:322-340 with its readers, :19-214): a model
 * file's bytes into a float32 triangle soup, 9 values per triangle in the
 * file's coordinates, for Voxel.build. The extension of `name` (or, when it
 * is none of stl/obj/glb/gltf, the content) picks binary or ASCII STL, OBJ
 * (polygons fan-triangulated) or binary glTF 2.0 (node transforms applied).
 * Host code; no device needed. With out == NULL, *count receives the number
 * of floats; otherwise *count is the capacity of out on entry and the number
 * written on return. Unreadable files fail with WT_E_ARG and the reference's
 * message in wt_last_error(). */
int wt_mesh_parse(const char *name, const void *buf, int64_t len, float *out, int64_t *count);

/* Boundary variants (set per lattice; the default is WT_BC_REFERENCE).
 * WT_BC_REFERENCE: the reference's _boundaries (js/lbm.js:473-512):
 *   equilibrium inlet, ceiling, floor and spanwise end planes, then the
 *   zero-gradient outlet copy.
 * WT_BC_CHANNEL (no reference counterpart: the "equilibrium outflow,
 *   free-slip side walls" of the 96x48x48 configuration): applied after
 *   the reference's boundaries. The outlet x = nx-1 becomes the equilibrium
 *   at rho = 1 and the velocity of x = nx-2. The y and z boundary planes
 *   become free-slip walls half a cell outside the first fluid rows:
 *   each population leaving a fluid cell towards a wall returns mirrored
 *   (specular reflection). The inlet stays the equilibrium at u0. AB layout; wt_step,
 *   wt_step_slabs and wt_time_steps only (the slab calls and the adjoint
 *   refuse it with WT_E_STATE / WT_E_ARG). */
enum { WT_BC_REFERENCE = 0, WT_BC_CHANNEL = 1 };
int wt_set_boundary_mode(wt_lattice *h, int mode);

/* Device time of nsteps steps, measured with HIP events on the lattice's
 * stream (no host sync inside the timed region). */
int wt_time_steps(wt_lattice *h, int nsteps, double *ms);

/* Free and total bytes of the current device's memory (hipMemGetInfo), for
 * sizing long adjoint runs (no reference counterpart): the checkpoints of
 * wt_adjoint_forward take at most half of what is free. */
int wt_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

/* Per-launch timing of the step kernel: while on, every step kernel is
 * bracketed by HIP events on the lattice's stream and its duration added to
 * a running total (the force-sum kernel is excluded). Turning it on clears
 * the totals. */
int wt_set_timing(wt_lattice *h, int on);
int wt_get_timing(const wt_lattice *h, double *kernel_ms, int64_t *launches);

/* Gray bounce-back (no reference counterpart): w holds one weight per cell;
 * a fluid cell j with w_j in (0,1) bounces the fraction w_j of every link
 * leaving it back into the downstream cell n = j + c_q:
 *   f_q(n) = (1 - w_j) g_q(j) + w_j g_qbar(n),   F -= 2 w_j g_qbar(n) c_q.
 * Weights 0 keep plain streaming and body cells (voxelized as solid) keep
 * the reference's bounce-back, so 0/1 fields reproduce the reference
 * exactly. A weight >= 1 on a fluid cell is an error; NULL clears the field.
 * AB layout only. */
int wt_set_gray(wt_lattice *h, const float *w);

/* Discrete adjoint (no reference counterpart): runs nsteps forward steps
 * from the current state, as wt_step does (forces may be NULL), then the
 * reverse sweep with sqrt(N) checkpointing and returns dJ/dw per cell for
 * the objective J = sum_k Fbar[3k..3k+2] . F_k; cells without a gray weight
 * get 0. interval <= 0 picks ceil(sqrt(nsteps)) steps per checkpoint, or
 * more when the checkpoints would take over half the free device memory.
 * The reverse sweep keeps a replay store of up to interval + 1 states; when
 * that does not fit, it replays each segment in shorter windows (more
 * recomputation, the same result bit for bit). reverse_ms (may be NULL)
 * receives the device time of the reverse sweep (recomputation plus adjoint
 * steps). AB layout, no moving walls. Any storage: binary16 states keep the
 * forward's own offsets, and the adjoint fields are float. */
int wt_adjoint(wt_lattice *h, int nsteps, const double *Fbar, double *dJdw, double *forces, int interval,
               double *reverse_ms);

/* The same as a forward/backward pair (torch.autograd.Function semantics;
 * no reference counterpart). wt_adjoint_forward runs nsteps from the current
 * state like wt_step and records sqrt(N) checkpoints on the lattice
 * (ceil(nsteps/interval) states in device memory); forces may be NULL.
 * wt_adjoint_backward takes the cotangent Fbar of the forward's forces
 * [nsteps][3] (nsteps must be the recorded count) and returns dJ/dw; it
 * leaves the lattice's state, parity and step count alone and may be called
 * any number of times. It fails with WT_E_STATE when no forward was
 * recorded or when parameters, body or weights changed since. wt_adjoint_release frees the record (so do the
 * next forward and wt_destroy). */
int wt_adjoint_forward(wt_lattice *h, int nsteps, int interval, double *forces);
int wt_adjoint_backward(wt_lattice *h, int nsteps, const double *Fbar, double *dJdw, double *reverse_ms);
int wt_adjoint_release(wt_lattice *h);

/* Identity of the recorded forward pass (no reference counterpart): every
 * wt_adjoint_forward on a lattice gets the next id (1, 2, ...), so a caller
 * that holds several graphs can tell that its record was replaced by a later
 * forward before it calls wt_adjoint_backward. WT_E_STATE when none. */
int wt_adjoint_record(const wt_lattice *h, int64_t *id);

/* Voxel.bounds (js/voxel.js:32-42) over len floats. */
int wt_voxel_bounds(const float *P, int64_t len, double out[6]);

/* Voxel.build(P,res) geometry only (js/voxel.js:48-71): fills vm. */
int wt_voxel_plan(const float *P, int64_t len, int res, wt_voxmeta *vm);

/* Voxel.build(P,res) on the GPU (js/voxel.js:48-102): grid receives
 * nx*ny*nz bytes (size them with wt_voxel_plan), vm the geometry and stats. */
int wt_voxel_build(const float *P, int64_t len, int res, wt_voxmeta *vm, uint8_t *grid);

#ifdef __cplusplus
}
#endif
#endif
