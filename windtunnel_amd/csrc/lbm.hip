// D3Q19 lattice kernels for gfx950: fused stream/collide step, reset,
// voxelize updates, mask builds, grid-backed body lookup.
//
// The step restates js/lbm.js:303-512 as ONE kernel per time step: every
// thread owns one cell of the full domain (x fastest, one block per lattice
// row), pulls its 19 populations, applies halfway bounce-back, the
// moving-wall correction and momentum exchange on near-wall cells, collides
// (BGK/TRT, optional Smagorinsky LES) and writes its 19 post-collision
// populations. Domain-boundary cells are handled by the same launch: the
// equilibrium planes write their equilibrium, and the x = nx-2 threads also
// write the zero-gradient outlet copy into x = nx-1 (js/lbm.js:504-511), so
// no second pass and no barrier are needed.
//
// Two population layouts share the code:
//  * AB: two buffers, f read and g written, the reference's own scheme
//    (js/lbm.js:71-80); g[q][cell] holds the post-collision value at home.
//  * esoteric pull: one buffer updated in place. In state `rep` the
//    post-collision value g_q(m) sits in slot_of(q, m, rep); a step loads
//    in representation phase and stores in 1-phase, and every cell reads
//    and writes only slots no other cell touches in that step (see
//    slot_of). This halves the population memory.
//
// Arithmetic type T is float (fast path) or double. With T = double and
// float storage the expression tree is the reference's, evaluated with
// contraction off, so the stored values are the reference's own.

#include "wt_internal.h"

#include <cmath>
#include <type_traits>

namespace wt {
namespace {

// ---------------------------------------------------------------- storage

template <class S> struct Pop;
// Build switches: WT_NT_LOAD / WT_NT_STORE mark the population streams
// non-temporal (each value is read once and written once per step).
template <class E> __device__ __forceinline__ E ld_pop(const E *p, int64_t k)
{
#ifdef WT_NT_LOAD
    return __builtin_nontemporal_load(p + k);
#else
    return p[k];
#endif
}
template <class E> __device__ __forceinline__ void st_pop(E *p, int64_t k, E v)
{
#ifdef WT_NT_STORE
    __builtin_nontemporal_store(v, p + k);
#else
    p[k] = v;
#endif
}

// binary16 through its bit pattern (the nontemporal builtins take integer
// and floating types only)
__device__ __forceinline__ __half ld_pop(const __half *p, int64_t k)
{
    return __ushort_as_half(ld_pop((const unsigned short *)p, k));
}
__device__ __forceinline__ void st_pop(__half *p, int64_t k, __half v)
{
    st_pop((unsigned short *)p, k, __half_as_ushort(v));
}

template <> struct Pop<float> {
    template <class T> __device__ static T ld(const float *p, int64_t k) { return (T)ld_pop(p, k); }
    template <class T> __device__ static void st(float *p, int64_t k, T v) { st_pop(p, k, (float)v); }
};
template <> struct Pop<double> {
    template <class T> __device__ static T ld(const double *p, int64_t k) { return (T)ld_pop(p, k); }
    template <class T> __device__ static void st(double *p, int64_t k, T v) { st_pop(p, k, (double)v); }
};
template <> struct Pop<__half> {
    template <class T> __device__ static T ld(const __half *p, int64_t k)
    {
        return (T)(__half2float(ld_pop(p, k)) * F16_INV_SCALE);
    }
    template <class T> __device__ static void st(__half *p, int64_t k, T v)
    {
        st_pop(p, k, __float2half_rn((float)v * F16_SCALE));
    }
};

// Population at byte address base + off: base is uniform across the wave,
// off a 32-bit per-lane offset (one scalar base register pair + one vector
// register per access).
template <class S, class T> __device__ __forceinline__ T ld_at(const S *base, uint32_t off)
{
    return Pop<S>::template ld<T>((const S *)((const char *)base + off), 0);
}
template <class S, class T> __device__ __forceinline__ void st_at(S *base, uint32_t off, T v)
{
    Pop<S>::template st<T>((S *)((char *)base + off), 0, v);
}

// Binary16 stores of collision results round stochastically: the scaled
// value's float32 bit pattern gets 13 random low bits added to its
// magnitude (the bits binary16 drops), and the conversion then rounds
// toward zero, so a value rounds up with probability equal to its distance
// from the value below.
// Round-to-nearest loses every update smaller than half a binary16 step
// and biases slowly relaxing steady flows (C_d 1.3 % off binary64 in a
// steady LES wake at 128x64x64, against 1e-5 with this rounding). The bits
// come from a hash of the cell index and the step's seed, so runs are
// reproducible and independent of the layout and of the order of cells.
// Build switch WT_F16_NEAREST=1 restores round-to-nearest. Constant
// writes (equilibrium planes, reset, voxelize) keep round-to-nearest.
#ifndef WT_F16_NEAREST
#define WT_F16_NEAREST 0
#endif
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t z)
{
    z = (z ^ (z >> 16)) * 0x85EBCA6Bu;
    z = (z ^ (z >> 13)) * 0xC2B2AE35u;
    return z ^ (z >> 16);
}
template <class S> constexpr bool rounds_stochastically() { return std::is_same<S, __half>::value && !WT_F16_NEAREST; }
// 32 random bits of cell i in this step (0 when the storage does not need them)
template <class S> __device__ __forceinline__ uint32_t cell_bits(int64_t i, uint32_t seed)
{
    if constexpr (rounds_stochastically<S>()) return mix32((uint32_t)i ^ seed);
    return 0;
}
// The 13 rounding bits of direction q: the cell's top 13 random bits XOR a
// constant of the direction (the top 13 bits of q times the golden ratio
// 2^32/phi, well spread over [0, 8192)). XOR with a constant is a bijection,
// so each direction's bits are uniform; adding them to a float's bit
// pattern compiles to one v_xad_u32 per direction.
__host__ __device__ constexpr uint32_t dir_key(int q) { return ((uint32_t)q * 0x9E3779B9u) >> 19; }
__device__ __forceinline__ uint32_t dir_bits(uint32_t r, int q) { return (r >> 19) ^ dir_key(q); }
// Two values already in binary16 units (x F16_SCALE) to binary16 with one
// conversion instruction: stochastically rounded with bits ra, rb (round
// toward zero after the random bits are added), or to nearest.
typedef __fp16 wt_half2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ wt_half2 f16_pair(float a, float b, uint32_t ra, uint32_t rb)
{
    if constexpr (!WT_F16_NEAREST) {
        a = __uint_as_float(__float_as_uint(a) + ra);
        b = __uint_as_float(__float_as_uint(b) + rb);
        return __builtin_amdgcn_cvt_pkrtz(a, b);
    } else {
        return wt_half2{(__fp16)a, (__fp16)b};
    }
}
// Two values in binary16 units to binary16 with stochastic rounding, the
// bits of directions qa and qb taken from the cell's top 13 random bits r13
// (r >> 19): dir_bits(r, q) added to the bit pattern as one v_xad_u32
// ((r13 ^ key) + bits) per value.
template <int QA, int QB> __device__ __forceinline__ wt_half2 f16_pair_dir(float a, float b, uint32_t r13)
{
    if constexpr (!WT_F16_NEAREST) {
        uint32_t ua, ub;
        asm("v_xad_u32 %0, %1, %2, %3" : "=v"(ua) : "v"(r13), "s"(dir_key(QA)), "v"(__float_as_uint(a)));
        asm("v_xad_u32 %0, %1, %2, %3" : "=v"(ub) : "v"(r13), "s"(dir_key(QB)), "v"(__float_as_uint(b)));
        return __builtin_amdgcn_cvt_pkrtz(__uint_as_float(ua), __uint_as_float(ub));
    } else {
        return wt_half2{(__fp16)a, (__fp16)b};
    }
}
// binary16 store of one half of a pair (the high half is stored straight
// from the pair's register)
__device__ __forceinline__ void st_h16(__half *p, __fp16 v)
{
#ifdef WT_NT_STORE
    __builtin_nontemporal_store(v, (__fp16 *)p);
#else
    *(__fp16 *)p = v;
#endif
}

// store of a collision result at element k (ptr + k), rounding bits rq
template <class S, class T> __device__ __forceinline__ void st_col(S *p, int64_t k, T v, uint32_t rq)
{
    if constexpr (rounds_stochastically<S>()) {
        st_h16(p + k, f16_pair((float)v * F16_SCALE, 0.0f, rq, 0).x);
    } else {
        Pop<S>::template st<T>(p, k, v);
    }
}
template <class S, class T> __device__ __forceinline__ void st_col_at(S *base, uint32_t off, T v, uint32_t rq)
{
    st_col<S, T>((S *)((char *)base + off), 0, v, rq);
}

__host__ __device__ constexpr int64_t cell_off(int q, int nx, int ny)
{
    return EX[q] + (int64_t)nx * (EY[q] + (int64_t)ny * EZ[q]);
}

// Esoteric-pull slot of g_q(m) in representation rep. Pairs (i, i+1) with
// i odd are opposite directions. rep 0: g_i(m) in slot i+1 of cell m+c_i,
// g_{i+1}(m) in slot i of m. rep 1: g_i(m) in slot i of m+c_i, g_{i+1}(m)
// in slot i+1 of m. The rest population stays in slot 0 of m.
__device__ __forceinline__ int64_t slot_of(int q, int64_t m, int rep, Lay sh, int nx, int ny)
{
    if (q == 0) return pix(0, m, sh);
    if (q & 1) return pix(rep ? q : q + 1, m + cell_off(q, nx, ny), sh);
    return pix(rep ? q : q - 1, m, sh);
}

// Is the slot of g_q at cell (x,y,z) inside the lattice? Only odd q place
// the value at a neighbour.
__device__ __forceinline__ bool slot_inside(int q, int x, int y, int z, int nx, int ny, int nz)
{
    if (!(q & 1)) return true;
    const int X = x + EX[q], Y = y + EY[q], Z = z + EZ[q];
    return X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz;
}

template <class S, class MAC> struct StepArgs {
    const S *F;  // AB: f (read); esoteric: the single buffer
    S *G;        // AB: g (written); esoteric: the single buffer
    const uint8_t *solid;
    const uint32_t *sbits, *nbits;
    const float *uwx, *uwy, *uwz;
    const float *wgt;  // gray bounce-back weights (AB layout), or null
    MAC *rho, *ux, *uy, *uz;
    double *partial;
    // force reduction of the previous launch, done by workgroup (0,0,0):
    // prev_count partials at prev_partial summed into prev_out (null: none)
    const double *prev_partial;
    double *prev_out;
    int64_t prev_count;
    int64_t n, ns;              // ns: cells per population plane (whole blocks)
    Lay lay;                    // population layout
    // near-wall cells of the launch: near[near0 .. near0 + near_cnt), taken
    // by the blocks of the first near_planes z planes of the grid
    const uint32_t *near;
    int64_t near0, near_cnt;
    int near_planes;
    int nx, ny, nz, nwx, z0, mov, phase, rho_near;
    int revy, revz;  // sweep the rows with y / z descending
    uint32_t seed;   // this step's seed of the binary16 rounding bits
    double tau, sq18cs2, lambda, u0;
    double om;  // 1/tau, divided on the host (the same IEEE quotient as on the device)
    int tau_le_half;  // tau <= 1/2: the straight path takes the guarded collision
    float tauf, omf, sq18f, lamf;  // the same in float: tau, 1/tau, 18 sqrt(2) Cs^2, Lambda
    double eqv[Q];  // boundary equilibrium deviation at (rho=1, u0, 0, 0)
    // Interior-cell addressing. Rows are whole storage blocks, so the
    // address of every population a cell reads or writes is
    //   lp[q] (or sp[q]) + row term (uniform per lattice row)
    //                    + lane term of x - shift (3 distinct values per cell)
    // lp/sp: address of direction q's load/store slot for the cell at the
    // start of row 0, including the y/z part of the streaming shift and, in
    // place, the slot of the current phase (it may lie before the buffer:
    // only base + row term + lane term is dereferenced). Row term:
    // (row >> lsh) * lstride elements; lane term of xs:
    // (xs >> lsh) * lstride + (xs & lmask).
    const S *lp[Q];
    S *sp[Q];
    int lsh, lstride, lmask;
    // pair launch (k_step_pair): byte offset of direction q's pull slot
    // from the row's load base F + row term + lmin (lmin <= 0 is the lowest
    // of those slots, so every offset is non-negative), and each
    // direction's equilibrium-plane value as a binary16 bit pattern in both
    // halves of a word
    int32_t loff[Q];
    int64_t lmin;
    uint32_t eqw[Q];
    // one-cell straight path through buffer resources (two buffers, when
    // every offset fits 31 bits): store offset of direction q from G + row
    // term, and whether loff/soff are valid
    int32_t soff[Q];
    int bufok;
};

// x shift of the cell whose slot holds the value a cell pulls along q (AB:
// the upstream cell; in place, odd q sit at home) and of the cell whose
// slot receives its post-collision g_q (in place, odd q go downstream).
template <bool ESO> __device__ constexpr int ld_shift(int q) { return ESO && (q & 1) ? 0 : -EX[q]; }
template <bool ESO> __device__ constexpr int st_shift(int q) { return ESO && (q & 1) ? EX[q] : 0; }

__device__ __forceinline__ bool bit(const uint32_t *bits, int nwx, int ny, int x, int y, int z)
{
    return (bits[(x >> 5) + (int64_t)nwx * (y + (int64_t)ny * z)] >> (x & 31)) & 1u;
}

// Build switch: WT_FAST_COLLIDE=0 runs float arithmetic through the
// reference-order collision below instead of collide_fast.
#ifndef WT_FAST_COLLIDE
#define WT_FAST_COLLIDE 1
#endif

// IEEE quotient and square root of binary64 without the range scaling and
// special-case fix-ups of the general sequences. The refinement is the one
// a binary64 '/' and sqrt() compile to on gfx950 (hardware reciprocal or
// reciprocal square root, two Newton-Raphson steps, a remainder-corrected
// final FMA), so the results are the correctly rounded ones, bit for bit,
// whenever no scaling would apply: divisors and quotients far inside the
// normal range, and square-root arguments of at least 2^-767. Every
// operand of the collision satisfies that (densities and relaxation times
// near 1, sums of squares of deviations that are 0 or above 1e-200);
// NaN stays non-finite, and sqrt_nz maps 0 to 0.
template <class T> __device__ __forceinline__ T rdiv(T a, T b) { return a / b; }
template <class T> __device__ __forceinline__ T rrcp(T b) { return T(1) / b; }
template <class T> __device__ __forceinline__ T sqrt_pos(T x) { return sqrt(x); }
template <class T> __device__ __forceinline__ T sqrt_nz(T x) { return sqrt(x); }
__device__ __forceinline__ double rcp_refined(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    return __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
}
template <> __device__ __forceinline__ double rdiv<double>(double a, double b)
{
    const double y = rcp_refined(b), q = a * y;
    return __builtin_fma(__builtin_fma(-b, q, a), y, q);
}
// a = 1: q = y exactly, so the final step needs no product
template <> __device__ __forceinline__ double rrcp<double>(double b)
{
    const double y = rcp_refined(b);
    return __builtin_fma(__builtin_fma(-b, y, 1.0), y, y);
}
template <> __device__ __forceinline__ double sqrt_pos<double>(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = y * 0.5;
    const double e = __builtin_fma(-h, s, 0.5);
    s = __builtin_fma(s, e, s);
    h = __builtin_fma(h, e, h);
    s = __builtin_fma(__builtin_fma(-s, s, x), h, s);
    return __builtin_fma(__builtin_fma(-s, s, x), h, s);
}
template <> __device__ __forceinline__ double sqrt_nz<double>(double x) { return x > 0.0 ? sqrt_pos(x) : x; }

template <class T, class A> __device__ __forceinline__ T omega0_of(const A &a)
{
    if constexpr (std::is_same<T, double>::value) return a.om;
    else return T(1) / (T)a.tau;
}

// Moments, deviation equilibrium, LES and TRT/BGK relaxation of one cell
// (js/lbm.js:377-460) in the reference's expression order (up to exact
// factors of two, below).
// GUARD: tau_eff = 1/2 exactly takes the IEEE operators (needed only when
// tau <= 1/2: for tau > 1/2, tau_eff >= tau > 1/2 in every cell).
template <class T, bool LES, bool TRT, bool GUARD = true>
__device__ __forceinline__ void collide(const T (&f)[Q], T tau0, T omega0, T sq18cs2, T lam, T (&g)[Q],
                                        T &r, T &vx, T &vy, T &vz)
{
    const T f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4], f5 = f[5], f6 = f[6],
            f7 = f[7], f8 = f[8], f9 = f[9], f10 = f[10], f11 = f[11], f12 = f[12], f13 = f[13],
            f14 = f[14], f15 = f[15], f16 = f[16], f17 = f[17], f18 = f[18];
    const T dr = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9 + f10 + f11 + f12 + f13 + f14 +
                 f15 + f16 + f17 + f18;
    r = T(1) + dr;
    const T inv = rrcp(r);
    vx = (f1 - f2 + f7 - f8 + f9 - f10 + f11 - f12 + f13 - f14) * inv;
    vy = (f3 - f4 + f7 - f8 - f9 + f10 + f15 - f16 + f17 - f18) * inv;
    vz = (f5 - f6 + f11 - f12 - f13 + f14 + f15 - f16 - f17 + f18) * inv;

    const T W0 = T(1.0 / 3), WF = T(1.0 / 18), WE = T(1.0 / 36);
    const T A = dr - r * T(1.5) * (vx * vx + vy * vy + vz * vz);
    const T r3 = T(3) * r, r45 = T(4.5) * r;
    const T bxy = vx + vy, bxY = vx - vy, bxz = vx + vz, bxZ = vx - vz, byz = vy + vz, byZ = vy - vz;
    const T e0 = W0 * A;
    const T hx = r45 * vx * vx, hy = r45 * vy * vy, hz = r45 * vz * vz;
    const T e1 = WF * (A + r3 * vx + hx), e2 = WF * (A - r3 * vx + hx);
    const T e3 = WF * (A + r3 * vy + hy), e4 = WF * (A - r3 * vy + hy);
    const T e5 = WF * (A + r3 * vz + hz), e6 = WF * (A - r3 * vz + hz);
    const T h7 = r45 * bxy * bxy, h9 = r45 * bxY * bxY, h11 = r45 * bxz * bxz;
    const T h13 = r45 * bxZ * bxZ, h15 = r45 * byz * byz, h17 = r45 * byZ * byZ;
    const T e7 = WE * (A + r3 * bxy + h7), e8 = WE * (A - r3 * bxy + h7);
    const T e9 = WE * (A + r3 * bxY + h9), e10 = WE * (A - r3 * bxY + h9);
    const T e11 = WE * (A + r3 * bxz + h11), e12 = WE * (A - r3 * bxz + h11);
    const T e13 = WE * (A + r3 * bxZ + h13), e14 = WE * (A - r3 * bxZ + h13);
    const T e15 = WE * (A + r3 * byz + h15), e16 = WE * (A - r3 * byz + h15);
    const T e17 = WE * (A + r3 * byZ + h17), e18 = WE * (A - r3 * byZ + h17);
    const T n1 = f1 - e1, n2 = f2 - e2, n3 = f3 - e3, n4 = f4 - e4, n5 = f5 - e5, n6 = f6 - e6;
    const T n7 = f7 - e7, n8 = f8 - e8, n9 = f9 - e9, n10 = f10 - e10, n11 = f11 - e11;
    const T n12 = f12 - e12, n13 = f13 - e13, n14 = f14 - e14, n15 = f15 - e15, n16 = f16 - e16;
    const T n17 = f17 - e17, n18 = f18 - e18;

    T om = omega0, tauEff = tau0;
    if (LES) {
        const T pxx = n1 + n2 + n7 + n8 + n9 + n10 + n11 + n12 + n13 + n14;
        const T pyy = n3 + n4 + n7 + n8 + n9 + n10 + n15 + n16 + n17 + n18;
        const T pzz = n5 + n6 + n11 + n12 + n13 + n14 + n15 + n16 + n17 + n18;
        const T pxy = n7 + n8 - n9 - n10;
        const T pxz = n11 + n12 - n13 - n14;
        const T pyz = n15 + n16 - n17 - n18;
        const T mag = sqrt_nz(T(2) * (pxx * pxx + pyy * pyy + pzz * pzz + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz)));
        tauEff = T(0.5) * (tau0 + sqrt_pos(tau0 * tau0 + sq18cs2 * mag * inv));
        om = rrcp(tauEff);
    }
    // tauEff = 0.5 exactly (tau = 0.5 in a cell at rest) divides by zero:
    // IEEE gives lam/0 = +-inf and 1/(0.5 +- inf) = +-0 (NaN for lam = 0)
    const T dt = tauEff - T(0.5);
    T omm = om;
    if (TRT) {
        if constexpr (GUARD) omm = dt != T(0) ? rrcp(T(0.5) + rdiv(lam, dt)) : T(1) / (T(0.5) + lam / dt);
        else omm = rrcp(T(0.5) + rdiv(lam, dt));
    }

    g[0] = f0 - om * (f0 - e0);
    // om * (0.5 (n_a + n_b)) rounds to exactly (0.5 om) * (n_a + n_b): a
    // factor of two commutes with rounding (no value here is near the
    // subnormal range), so the halves move into the rates, bit for bit
    const T omh = T(0.5) * om, ommh = T(0.5) * omm;
    T sp, sm;
#define WT_PAIR(a, b)                                \
    sp = omh * (n##a + n##b);                        \
    sm = ommh * (n##a - n##b);                       \
    g[a] = f##a - sp - sm;                           \
    g[b] = f##b - sp + sm;
    WT_PAIR(1, 2) WT_PAIR(3, 4) WT_PAIR(5, 6) WT_PAIR(7, 8) WT_PAIR(9, 10)
    WT_PAIR(11, 12) WT_PAIR(13, 14) WT_PAIR(15, 16) WT_PAIR(17, 18)
#undef WT_PAIR
}

// The same collision for float arithmetic, rearranged for instruction count:
// opposite pairs (a, a+1) enter through s = f_a + f_b and d = f_a - f_b, the
// equilibrium pair shares its even part W(A + 4.5 r cu^2) and odd part
// 3 W r cu, products are fused, and 1/x and sqrt use the hardware's
// single-precision approximations (about 1 ulp). Same model, same results to
// float rounding; the binary64 path above keeps the reference's order.
// The pair terms are carried doubled (2 sp = s - 2 even, 2 sm = d - 6 W r cu,
// an exact power-of-two scaling), which saves a multiply per pair.
// SCL: f and g are in binary16 units (deviations x F16_SCALE, a power of
// two), so binary16 loads and stores need no scaling; every product and sum
// is the unscaled one times the scale, exactly. r and u are unscaled.
// IN: float, or the raw binary16 loads (H16, the bit patterns zero-extended
// to 32 bits): every use of a loaded value is then an operand of a
// mixed-precision FMA (v_fma_mix_f32: f16 operands are widened exactly and
// the result is rounded once to f32, as a conversion and an add would give),
// so the loads need no conversion instructions.
// HI selects the half of the register a value sits in (0: low, 1: high),
// so the two cells of a packed pair of binary16 loads read the same register.
template <int HI> struct H16t {
    uint32_t bits;
};
using H16 = H16t<0>;
// a + b, a - b, a + x, a - x and a * b + h for binary16 a, b, h and float x
template <int HI> __device__ __forceinline__ float mix_hh(H16t<HI> a, H16t<HI> b)
{
    float d;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(d) : "v"(a.bits), "v"(b.bits));
    else asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(d) : "v"(a.bits), "v"(b.bits));
    return d;
}
template <int HI> __device__ __forceinline__ float mix_hh_sub(H16t<HI> a, H16t<HI> b)
{
    float d;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(d) : "v"(a.bits), "v"(b.bits));
    else asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(d) : "v"(a.bits), "v"(b.bits));
    return d;
}
template <int HI> __device__ __forceinline__ float mix_hf(H16t<HI> a, float x)
{
    float d;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(a.bits), "v"(x));
    else asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(a.bits), "v"(x));
    return d;
}
template <int HI> __device__ __forceinline__ float mix_hf_sub(H16t<HI> a, float x)
{
    float d;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(a.bits), "v"(x));
    else asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(a.bits), "v"(x));
    return d;
}
template <int HI> __device__ __forceinline__ float mix_ffh(float a, float b, H16t<HI> h)
{
    float d;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(b), "v"(h.bits));
    else asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(b), "v"(h.bits));
    return d;
}
__device__ __forceinline__ float add_in(float a, float b) { return a + b; }
template <int HI> __device__ __forceinline__ float add_in(H16t<HI> a, H16t<HI> b) { return mix_hh(a, b); }
__device__ __forceinline__ float sub_in(float a, float b) { return a - b; }
template <int HI> __device__ __forceinline__ float sub_in(H16t<HI> a, H16t<HI> b) { return mix_hh_sub(a, b); }
__device__ __forceinline__ float add_in(float a, float x, int) { return a + x; }
template <int HI> __device__ __forceinline__ float add_in(H16t<HI> a, float x, int) { return mix_hf(a, x); }
__device__ __forceinline__ float sub_in(float a, float x, int) { return a - x; }
template <int HI> __device__ __forceinline__ float sub_in(H16t<HI> a, float x, int) { return mix_hf_sub(a, x); }
__device__ __forceinline__ float fma_in(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <int HI> __device__ __forceinline__ float fma_in(float a, float b, H16t<HI> c) { return mix_ffh(a, b, c); }

template <bool LES, bool TRT, bool SCL = false, class IN = float>
__device__ __forceinline__ void collide_fast(const IN (&f)[Q], float tau0, float omega0, float sq18cs2, float lam,
                                             float (&g)[Q], float &r, float &vx, float &vy, float &vz)
{
    constexpr float SC = SCL ? F16_SCALE : 1.0f;
    float s[9], d[9];
#pragma unroll
    for (int p = 0; p < 9; p++) {
        s[p] = add_in(f[2 * p + 1], f[2 * p + 2]);
        d[p] = sub_in(f[2 * p + 1], f[2 * p + 2]);
    }
    const float dr = add_in(f[0], ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + s[8], 0);
    if constexpr (SCL) r = __builtin_fmaf(dr, F16_INV_SCALE, 1.0f);
    else r = 1.0f + dr;
    const float inv = __builtin_amdgcn_rcpf(r);
    const float invs = SCL ? inv * F16_INV_SCALE : inv;  // 1/r for scaled sums
    // pair directions: x, y, z, x+y, x-y, x+z, x-z, y+z, y-z
    vx = (d[0] + d[3] + d[4] + d[5] + d[6]) * invs;
    vy = (d[1] + d[3] - d[4] + d[7] + d[8]) * invs;
    vz = (d[2] + d[5] - d[6] + d[7] - d[8]) * invs;
    const float usq = __builtin_fmaf(vx, vx, __builtin_fmaf(vy, vy, vz * vz));
    const float A = __builtin_fmaf((-1.5f * SC) * r, usq, dr);
    const float cu[9] = {vx, vy, vz, vx + vy, vx - vy, vx + vz, vx - vz, vy + vz, vy - vz};
    const float WF = 1.0f / 18, WE = 1.0f / 36;
    // doubled equilibrium pair terms: 2 even = c1 cu^2 + c0, 2 odd = o2 cu
    const float c1F = ((9.0f * WF) * SC) * r, c1E = ((9.0f * WE) * SC) * r;
    const float c0F = (2.0f * WF) * A, c0E = (2.0f * WE) * A;
    const float o2F = ((6.0f * WF) * SC) * r, o2E = ((6.0f * WE) * SC) * r;
    float sp[9], sm[9];  // n_a + n_b and n_a - n_b of each pair (twice the half sums)
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const float c1 = p < 3 ? c1F : c1E, c0 = p < 3 ? c0F : c0E, o2 = p < 3 ? o2F : o2E;
        sp[p] = s[p] - __builtin_fmaf(c1 * cu[p], cu[p], c0);
        sm[p] = __builtin_fmaf(-o2, cu[p], d[p]);
    }
    const float n0 = fma_in(-1.0f / 3, A, f[0]);
    float om = omega0, tauEff = tau0;
    if (LES) {
        const float pxx = sp[0] + sp[3] + sp[4] + sp[5] + sp[6];
        const float pyy = sp[1] + sp[3] + sp[4] + sp[7] + sp[8];
        const float pzz = sp[2] + sp[5] + sp[6] + sp[7] + sp[8];
        const float pxy = sp[3] - sp[4], pxz = sp[5] - sp[6], pyz = sp[7] - sp[8];
        // the sums above are the tensor components: mag = sqrt(2 (...))
        const float q2 = __builtin_fmaf(pxx, pxx, __builtin_fmaf(pyy, pyy, pzz * pzz)) +
                         2.0f * __builtin_fmaf(pxy, pxy, __builtin_fmaf(pxz, pxz, pyz * pyz));
        const float mag = __builtin_amdgcn_sqrtf(2.0f * q2);
        tauEff = 0.5f * (tau0 + __builtin_amdgcn_sqrtf(__builtin_fmaf(tau0, tau0, sq18cs2 * mag * invs)));
        om = __builtin_amdgcn_rcpf(tauEff);
    }
    const float omm = TRT ? __builtin_amdgcn_rcpf(__builtin_fmaf(lam, __builtin_amdgcn_rcpf(tauEff - 0.5f), 0.5f)) : om;
    const float omh = 0.5f * om, ommh = 0.5f * omm;
    g[0] = fma_in(-om, n0, f[0]);
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const float x = omh * sp[p];
        g[2 * p + 1] = sub_in(f[2 * p + 1], __builtin_fmaf(ommh, sm[p], x), 0);
        g[2 * p + 2] = sub_in(f[2 * p + 2], __builtin_fmaf(-ommh, sm[p], x), 0);
    }
}

#ifndef WT_DEFER_FORCE
#define WT_DEFER_FORCE 1
#endif

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Fixed-order sum of `count` force partials into out[0..2] by one whole
// wavefront: lane l adds partials l, l+64, ...; then a shuffle tree. Every
// path that produces a step's force uses it, so forces are deterministic.
__device__ __forceinline__ void reduce_partials(const double *p, int64_t count, double *out)
{
    const int lane = threadIdx.x & 63;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int64_t b = lane; b < count; b += 64) { s0 += p[3 * b]; s1 += p[3 * b + 1]; s2 += p[3 * b + 2]; }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2);
    if (lane == 0) { out[0] = s0; out[1] = s1; out[2] = s2; }
}

// Resident waves per SIMD of the step kernel, by storage type (measured on
// MI355X at 512x256x256): binary16 moves half the bytes per load, so it
// needs 8 waves (at most 64 VGPRs) in flight to cover HBM latency; 4- and
// 8-byte storage run best with 4 waves per SIMD (twice as many cost ~7% at
// fp32). Build switches WT_F16_WAVES / WT_WAVES.
#ifndef WT_F16_WAVES
#define WT_F16_WAVES 8
#endif
#ifndef WT_WAVES
#define WT_WAVES 4
#endif
// Float32 storage with binary64 arithmetic (the reference's own numbers):
// build switch WT_EXACT_WAVES.
#ifndef WT_EXACT_WAVES
#define WT_EXACT_WAVES WT_WAVES
#endif
template <class S, class T> constexpr int step_waves()
{
    return std::is_same<S, __half>::value ? WT_F16_WAVES
                                          : (std::is_same<S, float>::value && std::is_same<T, double>::value ? WT_EXACT_WAVES
                                                                                                             : WT_WAVES);
}

// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle, no index stride).
constexpr int RAW_BUFFER_FLAGS = 0x00020000;
typedef unsigned int wt_u32x2 __attribute__((ext_vector_type(2)));
// One stored population through a buffer resource (voff: the lane's byte
// offset, soff: the uniform per-direction offset).
template <class S> __device__ __forceinline__ S buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <> __device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    const wt_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return __hiloint2double((int)v.y, (int)v.x);
}
template <class S> __device__ __forceinline__ void buf_st(S v, __amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ void buf_st<float>(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}
template <> __device__ __forceinline__ void buf_st<double>(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    wt_u32x2 u;
    u.x = (unsigned)__double2loint(v);
    u.y = (unsigned)__double2hiint(v);
    __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, 0);
}
// Binary64 arithmetic (the reference's own numbers) takes the one-cell
// straight path through buffer resources: fewer scalar instructions and
// registers beside its longer collision (0.930 against 0.945 ms at
// 512x256x256 on MI355X). Float arithmetic keeps the 64-bit scalar bases,
// which measured 3 % faster there than buffer resources. Build switch
// WT_BUF_STEP=0: 64-bit scalar bases for binary64 arithmetic too.
// Wave priority (0-3) while a binary64-arithmetic cell issues its 19 pulls:
// waves about to load go ahead of waves in their long binary64 collision,
// so more loads are in flight (+0.7 % at 512x256x256; on the float path the
// same raise costs 1.2 %, and it is not used there).
#ifndef WT_LOAD_PRIO
#define WT_LOAD_PRIO 2
#endif
#ifndef WT_BUF_STEP
#define WT_BUF_STEP 1
#endif

// One near-wall cell (an interior fluid cell with a solid upstream
// neighbour, a gray cell or a neighbour of one): halfway bounce-back, moving
// wall, momentum exchange and gray links, then the collision of the
// straight path.
template <class S, class T, class MAC, bool MACRO, bool LES, bool TRT, bool ESO, bool WIDE>
__device__ __forceinline__ void near_cell(const StepArgs<S, MAC> &a, int64_t i, double &fx, double &fy, double &fz)
{
    using P = Pop<S>;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    const int ph = a.phase, nph = a.phase ^ 1;
    const bool out = x == nx - 2;  // also writes the outlet copy at x = nx-1
    const bool out_sol = out && bit(a.sbits, a.nwx, ny, x + 1, y, z);
    const uint32_t rc = cell_bits<S>(i, a.seed);
    auto put = [&](int q, int64_t m, int cx, int cy, int cz, T v) {
        if (ESO) {
            if (slot_inside(q, cx, cy, cz, nx, ny, nz))
                st_col<S, T>(a.G, slot_of(q, m, nph, a.lay, nx, ny), v, dir_bits(rc, q));
        } else {
            st_col<S, T>(a.G, pix(q, m, a.lay), v, dir_bits(rc, q));
        }
    };

    // near-wall cell: halfway bounce-back, moving wall, momentum
    // exchange (js/lbm.js:334-368). Gray links (no reference counterpart):
    // the link from a fluid cell j transmits t = (1 - w_i)(1 - w_j) of the
    // pulled value and bounces the rest back, F -= (1 - t) 2 g_qbar(i) c_q;
    // j applies the same t to the opposite link, so the pair conserves mass.
    const float wi = (!ESO && a.wgt) ? a.wgt[i] : 0.0f;
    T f[Q];
    if constexpr (!WIDE) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int64_t j = i - cell_off(q, nx, ny);
            const int sj = a.solid[j];
            if (sj) {
                const int64_t k = ESO ? slot_of(OPP[q], i, ph, a.lay, nx, ny) : pix(OPP[q], i, a.lay);
                const T fd = P::template ld<T>(a.F, k);
                T corr = 0;
                if (a.mov) {
                    const T cu = (T)EX[q] * (T)a.uwx[j] + (T)EY[q] * (T)a.uwy[j] + (T)EZ[q] * (T)a.uwz[j];
                    if (cu != T(0)) {
                        const T wq = q == 0 ? T(1.0 / 3) : (q <= 6 ? T(1.0 / 18) : T(1.0 / 36));
                        corr = T(6) * wq * (T)a.rho[i] * cu;
                    }
                }
                f[q] = fd + corr;
                if (sj == SOLID_OBJ) {
                    const double d = (double)(T(2) * fd + corr);
                    fx -= d * EX[q]; fy -= d * EY[q]; fz -= d * EZ[q];
                }
            } else if (q && !ESO && a.wgt && (wi > 0.0f || a.wgt[j] > 0.0f)) {
                const T t = (T(1) - (T)wi) * (T(1) - (T)a.wgt[j]);
                const T fd = P::template ld<T>(a.F, pix(OPP[q], i, a.lay));
                f[q] = t * P::template ld<T>(a.F, pix(q, j, a.lay)) + (T(1) - t) * fd;
                const double d = (double)((T(1) - t) * (T(2) * fd));
                fx -= d * EX[q]; fy -= d * EY[q]; fz -= d * EZ[q];
            } else {
                const int64_t k = ESO ? slot_of(q, j, ph, a.lay, nx, ny) : pix(q, j, a.lay);
                f[q] = P::template ld<T>(a.F, k);
            }
        }
    }
    else {
        // WIDE: both candidates of every link (the pulled value and the
        // bounce-back value) are loaded up front with the upstream solid
        // states, so the cell costs one memory round trip (more registers)
        int sj[Q];
        T fp[Q], fb[Q];
        float wj[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int64_t j = i - cell_off(q, nx, ny);
            sj[q] = a.solid[j];
            fp[q] = P::template ld<T>(a.F, ESO ? slot_of(q, j, ph, a.lay, nx, ny) : pix(q, j, a.lay));
            fb[q] = P::template ld<T>(a.F, ESO ? slot_of(OPP[q], i, ph, a.lay, nx, ny) : pix(OPP[q], i, a.lay));
            wj[q] = (!ESO && a.wgt) ? a.wgt[j] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            T corr = 0;
            if (a.mov && sj[q]) {
                const int64_t j = i - cell_off(q, nx, ny);
                const T cu = (T)EX[q] * (T)a.uwx[j] + (T)EY[q] * (T)a.uwy[j] + (T)EZ[q] * (T)a.uwz[j];
                if (cu != T(0)) {
                    const T wq = q == 0 ? T(1.0 / 3) : (q <= 6 ? T(1.0 / 18) : T(1.0 / 36));
                    corr = T(6) * wq * (T)a.rho[i] * cu;
                }
            }
            const bool gray = q && !sj[q] && (wi > 0.0f || wj[q] > 0.0f);
            const T t = (T(1) - (T)wi) * (T(1) - (T)wj[q]);
            const T mix = t * fp[q] + (T(1) - t) * fb[q];
            f[q] = sj[q] ? fb[q] + corr : (gray ? mix : fp[q]);
            const double d = sj[q] == SOLID_OBJ ? (double)(T(2) * fb[q] + corr)
                                                : (gray ? (double)((T(1) - t) * (T(2) * fb[q])) : 0.0);
            fx -= d * EX[q]; fy -= d * EY[q]; fz -= d * EZ[q];
        }
    }
    T g[Q], r, vx, vy, vz;
    if constexpr (std::is_same<T, float>::value && WT_FAST_COLLIDE)
        collide_fast<LES, TRT>(f, a.tauf, a.omf, a.sq18f, a.lamf, g, r,
                               vx, vy, vz);
    else
        collide<T, LES, TRT>(f, (T)a.tau, omega0_of<T>(a), (T)a.sq18cs2, (T)a.lambda, g, r, vx, vy, vz);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        put(q, i, x, y, z, g[q]);
        if (out && !(ESO && out_sol)) put(q, i + 1, x + 1, y, z, g[q]);
    }
    if (MACRO) {
        a.rho[i] = (MAC)r; a.ux[i] = (MAC)vx; a.uy[i] = (MAC)vy; a.uz[i] = (MAC)vz;
        if (out) { a.rho[i + 1] = (MAC)r; a.ux[i + 1] = (MAC)vx; a.uy[i + 1] = (MAC)vy; a.uz[i + 1] = (MAC)vz; }
    } else if (a.rho_near) {
        a.rho[i] = (MAC)r;  // next step's moving-wall correction reads it
    }
}

// Where the near-wall cells are stepped by the one-cell kernel: binary16
// storage keeps it within 64 VGPRs (8 waves per SIMD), where the near-wall
// path would spill, so it steps them in a launch of its own (k_near) after
// the streaming kernel; wider storage runs them as the first blocks of the
// streaming launch (no second launch, no gap between launches). The
// binary16 pair kernel, the default binary16 step, runs them as its own
// first blocks as well.
// Build switch WT_EXACT_FUSED_NEAR=0: float32 storage with binary64
// arithmetic steps its near-wall cells in a launch of their own.
#ifndef WT_EXACT_FUSED_NEAR
#define WT_EXACT_FUSED_NEAR 1
#endif
template <class S, class T> constexpr bool fused_near()
{
    return !std::is_same<S, __half>::value &&
           (WT_EXACT_FUSED_NEAR || !(std::is_same<S, float>::value && std::is_same<T, double>::value));
}

// Block sum of the per-thread force: wave shuffles, then the waves' sums in
// wave order, written by thread 0 to out[0..2].
__device__ __forceinline__ void block_force(double fx, double fy, double fz, double *out)
{
    __shared__ double red[16][3];
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[wv][0] = fx; red[wv][1] = fy; red[wv][2] = fz; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sx = 0, sy = 0, sz = 0;
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; w++) { sx += red[w][0]; sy += red[w][1]; sz += red[w][2]; }
        out[0] = sx; out[1] = sy; out[2] = sz;
    }
}

// One cell of the streaming launch (x < nx - 1; x = nx - 1 is the outlet,
// written by x = nx - 2): the straight path of an interior fluid cell, or
// the equilibrium planes and the outlet copy of boundary and solid cells.
// Near-wall cells are left to the near-wall blocks or launch.
template <class S, class T, class MAC, bool MACRO, bool LES, bool TRT, bool ESO>
__device__ __forceinline__ void step_cell(const StepArgs<S, MAC> &a, int x, int y, int z)
{
    using P = Pop<S>;
    // binary16 storage with the float collision: the straight path works in
    // binary16 units (no scaling at loads and stores)
    constexpr bool HALF_UNITS = std::is_same<S, __half>::value && std::is_same<T, float>::value && WT_FAST_COLLIDE;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int nph = a.phase ^ 1;

    // store of one value of cell m (AB: home slot in g; esoteric: slot in the new representation)
    auto put = [&](int q, int64_t m, int cx, int cy, int cz, T v) {
        if (ESO) {
            if (slot_inside(q, cx, cy, cz, nx, ny, nz)) P::template st<T>(a.G, slot_of(q, m, nph, a.lay, nx, ny), v);
        } else {
            P::template st<T>(a.G, pix(q, m, a.lay), v);
        }
    };

    if (x < nx - 1) {
        const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
        const bool out = x == nx - 2;  // also writes the outlet copy at x = nx-1
        const bool bnd = x == 0 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
        const bool sol = bit(a.sbits, a.nwx, ny, x, y, z);
        const bool out_sol = out && bit(a.sbits, a.nwx, ny, x + 1, y, z);
        const bool nr = bit(a.nbits, a.nwx, ny, x, y, z);  // only interior fluid cells are near
        if (!(bnd || sol || nr)) {
            // Interior fluid cell with no wall link: pull, collide, store.
            // Rows are whole blocks, so every address is a per-direction
            // uniform base plus the lane term of x-1, x or x+1.
            const int64_t row = (int64_t)nx * (y + (int64_t)ny * z);
            const int64_t rb = (row >> a.lsh) * a.lstride;
            // lane term; a row spans fewer than 2^24 cells and a block fewer
            // than 2^24 elements, so the block offset is a 24-bit product
            auto lane = [&](int xs) {
                return (uint32_t)((__umul24((unsigned)(xs >> a.lsh), (unsigned)a.lstride) + (xs & a.lmask)) * (int)sizeof(S));
            };
            const uint32_t o[3] = {lane(x - 1), lane(x), lane(x + 1)};
            T f[Q];
            T g[Q], r, vx, vy, vz;
#if WT_BUF_STEP
            if constexpr (!ESO && !std::is_same<S, __half>::value && std::is_same<T, double>::value) {
                if (a.bufok) {
                    // the same pulls and stores through two buffer resources:
                    // a scalar base per row and a scalar offset per direction
                    const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(
                        (void *)((const char *)a.F + (rb * (int64_t)sizeof(S) + a.lmin)), 0, -1, RAW_BUFFER_FLAGS);
                    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(
                        (void *)((char *)a.G + rb * (int64_t)sizeof(S)), 0, -1, RAW_BUFFER_FLAGS);
                    if (WT_LOAD_PRIO) __builtin_amdgcn_s_setprio(WT_LOAD_PRIO);
#pragma unroll
                    for (int q = 0; q < Q; q++) f[q] = (T)buf_ld<S>(rf, (int)o[1 - EX[q]], a.loff[q]);
                    if (WT_LOAD_PRIO) __builtin_amdgcn_s_setprio(0);
                    if constexpr (std::is_same<T, float>::value && WT_FAST_COLLIDE)
                        collide_fast<LES, TRT>(f, a.tauf, a.omf, a.sq18f, a.lamf, g, r, vx, vy, vz);
                    else if (a.tau_le_half)
                        collide<T, LES, TRT, true>(f, (T)a.tau, omega0_of<T>(a), (T)a.sq18cs2, (T)a.lambda, g, r, vx, vy, vz);
                    else
                        collide<T, LES, TRT, false>(f, (T)a.tau, omega0_of<T>(a), (T)a.sq18cs2, (T)a.lambda, g, r, vx, vy, vz);
#pragma unroll
                    for (int q = 0; q < Q; q++) buf_st<S>((S)g[q], rg, (int)o[1], a.soff[q]);
                    if (out) {
                        // zero-gradient outlet: the same values again for x = nx-1
#pragma unroll
                        for (int q = 0; q < Q; q++) buf_st<S>((S)g[q], rg, (int)o[2], a.soff[q]);
                    }
                    if (MACRO) {
                        a.rho[i] = (MAC)r; a.ux[i] = (MAC)vx; a.uy[i] = (MAC)vy; a.uz[i] = (MAC)vz;
                        if (out) { a.rho[i + 1] = (MAC)r; a.ux[i + 1] = (MAC)vx; a.uy[i + 1] = (MAC)vy; a.uz[i + 1] = (MAC)vz; }
                    }
                    return;
                }
            }
#endif
            if constexpr (HALF_UNITS) {
                // binary16 storage: collide in binary16 units, no scaling,
                // straight from the raw loads
                H16 h[Q];
#pragma unroll
                for (int q = 0; q < Q; q++)
                    h[q].bits = ld_pop((const unsigned short *)((const char *)(a.lp[q] + rb) + o[1 + ld_shift<ESO>(q)]), 0);
                collide_fast<LES, TRT, true, H16>(h, a.tauf, a.omf, a.sq18f, a.lamf, g, r, vx, vy, vz);
            } else {
#pragma unroll
                for (int q = 0; q < Q; q++) f[q] = ld_at<S, T>(a.lp[q] + rb, o[1 + ld_shift<ESO>(q)]);
                if constexpr (std::is_same<T, float>::value && WT_FAST_COLLIDE)
                    collide_fast<LES, TRT>(f, a.tauf, a.omf, a.sq18f, a.lamf, g, r,
                                           vx, vy, vz);
                else
                    collide<T, LES, TRT>(f, (T)a.tau, omega0_of<T>(a), (T)a.sq18cs2, (T)a.lambda, g, r, vx, vy, vz);
            }
            const uint32_t rc = cell_bits<S>(i, a.seed);
            // the post-collision values of direction q go to lane offset os(q, xo)
            // of the row, for xo = x (home) or x + 1 (outlet copy)
            auto store_all = [&](const uint32_t (&os)[3]) {
                if constexpr (HALF_UNITS) {
                    auto at = [&](int q) { return (__half *)((char *)(a.sp[q] + rb) + os[1 + st_shift<ESO>(q)]); };
                    const uint32_t r13 = rc >> 19;
                    st_h16(at(0), f16_pair(g[0], 0.0f, dir_bits(rc, 0), 0).x);
                    auto pair = [&](auto qa) {
                        constexpr int q = decltype(qa)::value;
                        const wt_half2 h = f16_pair_dir<q, q + 1>(g[q], g[q + 1], r13);
                        st_h16(at(q), h.x);
                        st_h16(at(q + 1), h.y);
                    };
                    pair(std::integral_constant<int, 1>{}); pair(std::integral_constant<int, 3>{});
                    pair(std::integral_constant<int, 5>{}); pair(std::integral_constant<int, 7>{});
                    pair(std::integral_constant<int, 9>{}); pair(std::integral_constant<int, 11>{});
                    pair(std::integral_constant<int, 13>{}); pair(std::integral_constant<int, 15>{});
                    pair(std::integral_constant<int, 17>{});
                } else {
#pragma unroll
                    for (int q = 0; q < Q; q++) st_col_at<S, T>(a.sp[q] + rb, os[1 + st_shift<ESO>(q)], g[q], dir_bits(rc, q));
                }
            };
            store_all(o);
            if (out && !(ESO && out_sol)) {
                // zero-gradient outlet: the same values again for x = nx-1
                const uint32_t o2[3] = {o[1], o[2], lane(x + 2)};
                if constexpr (HALF_UNITS) {
                    // convert again here (one wave in eight), so the home
                    // stores above take the high halves straight from the
                    // packed registers
#pragma unroll
                    for (int q = 0; q < Q; q++) asm volatile("" : "+v"(g[q]));
                }
                store_all(o2);
            }
            if (MACRO) {
                a.rho[i] = (MAC)r; a.ux[i] = (MAC)vx; a.uy[i] = (MAC)vy; a.uz[i] = (MAC)vz;
                if (out) { a.rho[i + 1] = (MAC)r; a.ux[i + 1] = (MAC)vx; a.uy[i + 1] = (MAC)vy; a.uz[i + 1] = (MAC)vz; }
            }
        } else if (bnd || sol) {
            // Equilibrium planes (js/lbm.js:488-503): all of them write except
            // the floor where it is solid and no other plane covers the cell.
            // In place, a solid cell never writes: its slots are where its
            // fluid neighbours read their bounce-back values.
            const bool eq = bnd && (ESO ? !sol : !(sol && y == 0 && x != 0 && z != 0 && z != nz - 1 && y != ny - 1));
            if (eq) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    put(q, i, x, y, z, (T)a.eqv[q]);
                    if (out && !(ESO && out_sol)) put(q, i + 1, x + 1, y, z, (T)a.eqv[q]);
                }
                if (MACRO) {
                    a.rho[i] = (MAC)1; a.ux[i] = (MAC)a.u0; a.uy[i] = 0; a.uz[i] = 0;
                    if (out) { a.rho[i + 1] = (MAC)1; a.ux[i + 1] = (MAC)a.u0; a.uy[i + 1] = 0; a.uz[i + 1] = 0; }
                }
            } else if (out) {
                // the outlet copies whatever a solid x = nx-2 cell holds (AB
                // only: in place, a solid cell holds no values of its own)
                if (!ESO) {
#pragma unroll
                    for (int q = 0; q < Q; q++) a.G[pix(q, i + 1, a.lay)] = a.G[pix(q, i, a.lay)];
                }
                if (MACRO) {
                    a.rho[i + 1] = a.rho[i]; a.ux[i + 1] = a.ux[i];
                    a.uy[i + 1] = a.uy[i]; a.uz[i + 1] = a.uz[i];
                }
            }
        }
        // near-wall cells (interior fluid with a solid upstream neighbour,
        // gray cells and their neighbours) are stepped by the near-wall blocks
    }
}

template <class S, class T, class MAC, bool MACRO, bool LES, bool TRT, bool ESO, int WAVES>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(WAVES))) k_step(StepArgs<S, MAC> a)
{
    // Four waves per SIMD at most: claiming v127 gives every wave a
    // 128-register file, so at most four waves fit a SIMD.
    if constexpr (WAVES <= 4) asm volatile("" ::: "v127");
    if (a.prev_out && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x < 64)
        reduce_partials(a.prev_partial, a.prev_count, a.prev_out);
    if constexpr (fused_near<S, T>()) {
        if ((int)blockIdx.z < a.near_planes) {
            // The first near_planes z planes of the grid are near-wall
            // blocks: one listed cell per thread. They are dispatched first,
            // so their long dependent load chains overlap the streaming
            // blocks; each writes its FP64 force partial.
            const int64_t nb = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
            const int64_t k = nb * blockDim.x + threadIdx.x;
            double fx = 0, fy = 0, fz = 0;
            if (k < a.near_cnt)
                near_cell<S, T, MAC, MACRO, LES, TRT, ESO, false>(a, (int64_t)a.near[a.near0 + k], fx, fy, fz);
            block_force(fx, fy, fz, a.partial + nb * 3);
            return;
        }
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = a.revy ? a.ny - 1 - (int)blockIdx.y : (int)blockIdx.y;
    const int zb = (int)blockIdx.z - a.near_planes, nzr = (int)gridDim.z - a.near_planes;
    const int z = a.z0 + (a.revz ? nzr - 1 - zb : zb);
    step_cell<S, T, MAC, MACRO, LES, TRT, ESO>(a, x, y, z);
}

// The near-wall cells of one step as a launch of their own (binary16
// storage): 256-thread blocks over the list (grid-stride when it is longer
// than the partial buffer allows), the one-round-trip gather, one force
// partial per block.
constexpr int NEAR_THREADS = 256;
template <class S, class T, class MAC, bool MACRO, bool LES, bool TRT, bool ESO>
__global__ void __launch_bounds__(NEAR_THREADS) k_near(StepArgs<S, MAC> a)
{
    double fx = 0, fy = 0, fz = 0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < a.near_cnt;
         k += (int64_t)gridDim.x * blockDim.x)
        near_cell<S, T, MAC, MACRO, LES, TRT, ESO, true>(a, (int64_t)a.near[a.near0 + k], fx, fy, fz);
    block_force(fx, fy, fz, a.partial + (int64_t)blockIdx.x * 3);
}

// f(integral_constant<int, q>) for q = 0 .. 18
#define WT_FOR_Q(f)                                                                                            \
    f(std::integral_constant<int, 0>{}); f(std::integral_constant<int, 1>{}); f(std::integral_constant<int, 2>{}); \
    f(std::integral_constant<int, 3>{}); f(std::integral_constant<int, 4>{}); f(std::integral_constant<int, 5>{}); \
    f(std::integral_constant<int, 6>{}); f(std::integral_constant<int, 7>{}); f(std::integral_constant<int, 8>{}); \
    f(std::integral_constant<int, 9>{}); f(std::integral_constant<int, 10>{}); f(std::integral_constant<int, 11>{}); \
    f(std::integral_constant<int, 12>{}); f(std::integral_constant<int, 13>{}); f(std::integral_constant<int, 14>{}); \
    f(std::integral_constant<int, 15>{}); f(std::integral_constant<int, 16>{}); f(std::integral_constant<int, 17>{}); \
    f(std::integral_constant<int, 18>{});

// ------------------------------------------------------- binary16 pairs
// The binary16 step with two cells per thread (x, x + 1; x even), for two
// buffers in 128-cell blocks. A wave then covers one block, and every
// population a thread pulls or stores is one 4-byte word holding both
// cells: half the memory instructions of one cell per thread, whose 2-byte
// accesses keep the texture addresser busy (on MI355X it, not HBM, bounds
// the one-cell binary16 step). The pulls along +x and -x are the word of
// the same cells shifted across lanes (DPP wavefront shift by one lane);
// the lane at the block's edge takes the missing half from the neighbouring
// block's word, one uniform load per shifted direction and wave. Each
// cell's collision, rounding and results are
// those of the one-cell path bit for bit: the same code, reading the low or
// the high half of the shared registers.
constexpr int PAIR_DIRS[10] = {1, 7, 9, 11, 13, 2, 8, 10, 12, 14};  // e_x = +1 (5), then -1 (5)
// Four waves per SIMD (at most 128 VGPRs): two cells per thread fill them
// with as many cells as eight one-cell waves, and the near-wall path fits
// beside the streaming path without spilling, so the near-wall cells run
// as the first blocks of the same launch.
// Wave priority (0-3) while a pair of cells issues its pulls and edge loads
// (binary16: 1.8 % faster at 512x256x256).
#ifndef WT_PAIR_PRIO
#define WT_PAIR_PRIO 2
#endif
#ifndef WT_PAIR_WAVES
#define WT_PAIR_WAVES 4
#endif

// One pair word through a buffer resource: voff is the lane's byte offset,
// soff the uniform per-direction offset.
template <class W> __device__ __forceinline__ W pair_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ __forceinline__ uint32_t pair_load<uint32_t>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
template <> __device__ __forceinline__ float2 pair_load<float2>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    const wt_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ void pair_store(uint32_t v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 0);
}
__device__ __forceinline__ void pair_store(float2 v, __amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    wt_u32x2 u;
    u.x = __float_as_uint(v.x);
    u.y = __float_as_uint(v.y);
    __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, 0);
}

// Float32 storage the same way: a pair is two 4-byte values (one 8-byte
// access), the x shifts move one of them across lanes, and the edge values
// are single elements.
template <class S, bool MACRO, bool LES, bool TRT, int WAVES>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_step_pair(StepArgs<S, float> a)
{
    constexpr bool HALF = std::is_same<S, __half>::value;
    using MAC = float;
    using Word = typename std::conditional<HALF, uint32_t, float2>::type;  // the pair's two values
    if (a.prev_out && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x < 64)
        reduce_partials(a.prev_partial, a.prev_count, a.prev_out);
    if ((int)blockIdx.z < a.near_planes) {
        // near-wall blocks, dispatched first: two listed cells per thread,
        // one force partial per block
        const int64_t nb = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
        double fx = 0, fy = 0, fz = 0;
#pragma unroll 1
        for (int r = 0; r < 2; r++) {
            const int64_t k = (2 * nb + r) * blockDim.x + threadIdx.x;
            if (k < a.near_cnt)
                near_cell<S, float, MAC, MACRO, LES, TRT, false, false>(a, (int64_t)a.near[a.near0 + k], fx, fy, fz);
        }
        block_force(fx, fy, fz, a.partial + nb * 3);
        return;
    }
    const int x = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = a.revy ? a.ny - 1 - (int)blockIdx.y : (int)blockIdx.y;
    const int zb = (int)blockIdx.z - a.near_planes, nzr = (int)gridDim.z - a.near_planes;
    const int z = a.z0 + (a.revz ? nzr - 1 - zb : zb);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1) {
        // equilibrium rows (whole workgroups): the one-cell code
        step_cell<S, float, MAC, MACRO, LES, TRT, false>(a, x, y, z);
        step_cell<S, float, MAC, MACRO, LES, TRT, false>(a, x + 1, y, z);
        return;
    }
    const int lane = (int)(threadIdx.x & 63);
    const int x0 = x - 2 * lane;  // first cell of the wave's 128
    const int64_t row = (int64_t)nx * (y + (int64_t)ny * z);
    const int64_t rb = (row >> a.lsh) * a.lstride;
    // element of cell xs within its row (a row spans fewer than 2^24 cells)
    auto elem = [&](int xs) {
        return __umul24((unsigned)(xs >> a.lsh), (unsigned)a.lstride) + (unsigned)(xs & a.lmask);
    };
    const uint32_t eb = (uint32_t)sizeof(S) * elem(x);  // byte offset of the pair (x, x + 1)
    // Buffer resources over the row's slots: every access is a scalar base,
    // a scalar per-direction offset and the lane's byte offset eb, so the
    // 19 directions cost no vector address arithmetic and two 64-bit bases
    // instead of 38 (raw buffers, no range check; the offsets stay below
    // 2^31 and every address lies inside the population buffers).
    const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc(
        (void *)((const char *)a.F + (rb * (int64_t)sizeof(S) + a.lmin)), 0, -1, RAW_BUFFER_FLAGS);
    const __amdgpu_buffer_rsrc_t rg =
        __builtin_amdgcn_make_buffer_rsrc((void *)((char *)a.G + rb * (int64_t)sizeof(S)), 0, -1, RAW_BUFFER_FLAGS);
    // every lane of the row pulls (the shifted pairs need their neighbours)
    Word w[Q];
    if (WT_PAIR_PRIO) __builtin_amdgcn_s_setprio(WT_PAIR_PRIO);
#pragma unroll
    for (int q = 0; q < Q; q++) w[q] = pair_load<Word>(rf, (int)eb, a.loff[q]);
    // the value beyond the wave's edge of each direction PAIR_DIRS[k]
    // (before it for e_x = +1, after it for -1; clamped into the row at its
    // ends, where no cell uses it): one uniform load per direction, the
    // same word in every lane; binary16 loads the aligned word that holds it
    uint32_t ev[10];
    {
        const int xw = __builtin_amdgcn_readfirstlane(x0);
        const int lo = HALF ? 2 : 1;
        const int eb_lo = (int)(sizeof(S) * elem(xw >= lo ? xw - lo : 0));
        const int eb_hi = (int)(sizeof(S) * elem(xw + 128 <= nx - lo ? xw + 128 : nx - lo));
#pragma unroll
        for (int k = 0; k < 10; k++)
            ev[k] = __builtin_amdgcn_raw_buffer_load_b32(rf, 0, a.loff[PAIR_DIRS[k]] + (k < 5 ? eb_lo : eb_hi), 0);
    }
    if (WT_PAIR_PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int k = 0; k < 10; k++) {
        const int q = PAIR_DIRS[k];
        const int edge = (int)ev[k];
        if constexpr (HALF) {
            if (k < 5) {
                // pull from x - 1: (high half of the lane below, low half of this lane)
                const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(edge, (int)w[q], 0x138, 0xf, 0xf, false);
                w[q] = __builtin_amdgcn_alignbit(w[q], prev, 16);
            } else {
                // pull from x + 1: (high half of this lane, low half of the lane above)
                const uint32_t next = (uint32_t)__builtin_amdgcn_update_dpp(edge, (int)w[q], 0x130, 0xf, 0xf, false);
                w[q] = __builtin_amdgcn_alignbit(next, w[q], 16);
            }
        } else {
            if (k < 5) {
                // pull from x - 1: (the lane below's x + 1, this lane's x)
                const float prev =
                    __int_as_float(__builtin_amdgcn_update_dpp(edge, __float_as_int(w[q].y), 0x138, 0xf, 0xf, false));
                w[q] = make_float2(prev, w[q].x);
            } else {
                // pull from x + 1: (this lane's x + 1, the lane above's x)
                const float next =
                    __int_as_float(__builtin_amdgcn_update_dpp(edge, __float_as_int(w[q].x), 0x130, 0xf, 0xf, false));
                w[q] = make_float2(w[q].y, next);
            }
        }
    }
    const int64_t wi = (x >> 5) + (int64_t)a.nwx * (y + (int64_t)ny * z);
    const uint32_t sw = a.sbits[wi] >> (x & 31), nw = a.nbits[wi] >> (x & 31);
    const bool solA = sw & 1u, solB = (sw >> 1) & 1u, nrA = nw & 1u, nrB = (nw >> 1) & 1u;
    const bool bndA = x == 0, outB = x + 1 == nx - 1;
    // both cells collide (the values of boundary cells are replaced, those of
    // solid and near-wall cells masked below: the near-wall blocks write
    // near-wall cells)
    float gA[Q], gB[Q], rA, uxA, uyA, uzA, rB, uxB, uyB, uzB;
    if constexpr (HALF) {
        H16t<0> hA[Q];
        H16t<1> hB[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) { hA[q].bits = w[q]; hB[q].bits = w[q]; }
        collide_fast<LES, TRT, true>(hA, a.tauf, a.omf, a.sq18f, a.lamf, gA, rA, uxA, uyA, uzA);
        collide_fast<LES, TRT, true>(hB, a.tauf, a.omf, a.sq18f, a.lamf, gB, rB, uxB, uyB, uzB);
    } else {
        float fA[Q], fB[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) { fA[q] = w[q].x; fB[q] = w[q].y; }
        collide_fast<LES, TRT>(fA, a.tauf, a.omf, a.sq18f, a.lamf, gA, rA, uxA, uyA, uzA);
        collide_fast<LES, TRT>(fB, a.tauf, a.omf, a.sq18f, a.lamf, gB, rB, uxB, uyB, uzB);
    }
    const int64_t iA = row + x;
    // the pair to store: cell A (the equilibrium value at x = 0) and cell B
    // (at x + 1 = nx - 1, the outlet: a copy of cell A)
    // ENDS: the wave holds the row's first or last pair (equilibrium inlet
    // value, outlet copy); other waves store the collided pair as it is
    auto word = [&](auto qc, auto ends) -> Word {
        constexpr int q = decltype(qc)::value;
        constexpr bool ENDS = decltype(ends)::value;
        if constexpr (HALF) {
            const uint32_t r13A = cell_bits<S>(iA, a.seed) >> 19, r13B = cell_bits<S>(iA + 1, a.seed) >> 19;
            const uint32_t sel = bndA ? 0x07060100u : (outB ? 0x05040504u : 0x07060504u);
            uint32_t ua, ub;
            if constexpr (!WT_F16_NEAREST) {
                asm("v_xad_u32 %0, %1, %2, %3" : "=v"(ua) : "v"(r13A), "s"(dir_key(q)), "v"(__float_as_uint(gA[q])));
                asm("v_xad_u32 %0, %1, %2, %3" : "=v"(ub) : "v"(r13B), "s"(dir_key(q)), "v"(__float_as_uint(gB[q])));
            } else {
                ua = __float_as_uint(gA[q]);
                ub = __float_as_uint(gB[q]);
            }
            wt_half2 pk = WT_F16_NEAREST ? wt_half2{(__fp16)gA[q], (__fp16)gB[q]}
                                         : __builtin_amdgcn_cvt_pkrtz(__uint_as_float(ua), __uint_as_float(ub));
            uint32_t pw;
            __builtin_memcpy(&pw, &pk, 4);
            if constexpr (!ENDS) return pw;
            else return __builtin_amdgcn_perm(pw, a.eqw[q], sel);
        } else {
            if constexpr (!ENDS) return make_float2(gA[q], gB[q]);
            const float va = bndA ? (float)a.eqv[q] : gA[q];
            return make_float2(va, outB ? va : gB[q]);
        }
    };
    // halves written here: every cell but a solid or near-wall one (the
    // inlet cell x = 0 writes its equilibrium whatever it is); the outlet's
    // when cell A is written (a near-wall x = nx - 2 cell writes the outlet
    // from its near-wall block, a solid one copies itself into it in the
    // one-cell code below)
    const bool wA = bndA || (!solA && !nrA), wB = outB ? (!solA && !nrA) : (!solB && !nrB);
    // store offset of direction q (two buffers: its home slot in the block)
    auto soff = [&](int q) { return (int)((SLOT[q] << a.lsh) * (int)sizeof(S)); };
    if (__builtin_expect(!__any(solA || solB || nrA || nrB), 1)) {
        if (x0 != 0 && x0 + 128 < nx) {
            // a wave away from both row ends: no equilibrium or outlet value
            auto st = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                pair_store(word(qc, std::false_type{}), rg, (int)eb, soff(q));
            };
            WT_FOR_Q(st)
        } else {
            auto st = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                pair_store(word(qc, std::true_type{}), rg, (int)eb, soff(q));
            };
            WT_FOR_Q(st)
        }
    } else {
        auto st = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const Word v = word(qc, std::true_type{});
            if (wA && wB) {
                pair_store(v, rg, (int)eb, soff(q));
            } else if constexpr (HALF) {
                if (wA) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, rg, (int)eb, soff(q), 0);
                else if (wB) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(v >> 16), rg, (int)eb + 2, soff(q), 0);
            } else {
                if (wA) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.x), rg, (int)eb, soff(q), 0);
                else if (wB) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v.y), rg, (int)eb + 4, soff(q), 0);
            }
        };
        WT_FOR_Q(st)
        if (solA && outB) step_cell<S, float, MAC, MACRO, LES, TRT, false>(a, x, y, z);
    }
    if (MACRO) {
        // fields of the written straight and equilibrium cells (a near-wall
        // cell's are its near-wall block's)
        if (bndA || (!solA && !nrA)) {
            a.rho[iA] = bndA ? 1.0f : rA; a.ux[iA] = bndA ? (float)a.u0 : uxA;
            a.uy[iA] = bndA ? 0.0f : uyA; a.uz[iA] = bndA ? 0.0f : uzA;
        }
        if (outB ? (!solA && !nrA) : (!solB && !nrB)) {
            a.rho[iA + 1] = outB ? rA : rB; a.ux[iA + 1] = outB ? uxA : uxB;
            a.uy[iA + 1] = outB ? uyA : uyB; a.uz[iA + 1] = outB ? uzA : uzB;
        }
    }
}

// A pending force reduction on its own (one wavefront).
__global__ void __launch_bounds__(64) k_force_sum(const double *partial, int64_t count, double *out)
{
    reduce_partials(partial, count, out);
}

// ------------------------------------------------ channel boundaries
// Configuration variant with free-slip side walls and an equilibrium
// outflow (no reference function), applied after the step as a fix-up of
// what the step wrote (AB layout). Thread roles:
//  - one per interior (y, z) of the outlet plane x = nx-1: f^eq(1, u) with u
//    the velocity of the post-collision populations of x = nx-2 (0 at a solid
//    cell), its fields on the last step of a batch, then its reflections;
//  - one per interior fluid cell x in [1, nx-2] of a row next to a side wall
//    (`ring`): its reflections.
// Reflection: each population q that leaves the cell towards a side wall is
// copied into the ghost cell it faces across that wall, in the slot of q with
// e_y and/or e_z flipped, where the neighbour pulls it next step (halfway
// specular reflection). Every ghost slot has exactly one writer.
constexpr int reflect_dir(int q, bool fy, bool fz)
{
    for (int r = 0; r < Q; r++)
        if (EX[r] == EX[q] && EY[r] == (fy ? -EY[q] : EY[q]) && EZ[r] == (fz ? -EZ[q] : EZ[q])) return r;
    return q;
}

template <class S, class MAC> struct ChanArgs {
    S *G;
    const uint8_t *solid;
    MAC *rho, *ux, *uy, *uz;
    const int2 *ring;  // (y, z) of the interior rows next to a side wall
    int64_t nring;
    Lay lay;
    int nx, ny, nz, macro;
};

template <class S> __device__ __forceinline__ S to_store(double v);
template <> __device__ __forceinline__ float to_store<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ double to_store<double>(double v) { return v; }
template <> __device__ __forceinline__ __half to_store<__half>(double v) { return __float2half_rn((float)v * F16_SCALE); }

template <class S, class MAC>
__device__ __forceinline__ void reflect_out(const ChanArgs<S, MAC> &a, int x, int y, int z, const S (&v)[Q])
{
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int ty = y + EY[q], tz = z + EZ[q];
        const bool cy = ty == 0 || ty == a.ny - 1, cz = tz == 0 || tz == a.nz - 1;
        if (!(cy || cz)) continue;
        const int64_t ghost = x + (int64_t)a.nx * ((cy ? ty : y) + (int64_t)a.ny * (cz ? tz : z));
        const int64_t k = cy ? (cz ? pix(reflect_dir(q, true, true), ghost, a.lay) : pix(reflect_dir(q, true, false), ghost, a.lay))
                             : pix(reflect_dir(q, false, true), ghost, a.lay);
        a.G[k] = v[q];
    }
}

template <class S, class T, class MAC>
__global__ void __launch_bounds__(256) k_channel(ChanArgs<S, MAC> a)
{
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nout = (int64_t)(a.ny - 2) * (a.nz - 2);
    if (k < nout) {
        const int y = 1 + (int)(k % (a.ny - 2)), z = 1 + (int)(k / (a.ny - 2));
        const int64_t i = (a.nx - 1) + (int64_t)a.nx * (y + (int64_t)a.ny * z), j = i - 1;
        T vx = 0, vy = 0, vz = 0;
        if (!a.solid[j]) {
            T f[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) f[q] = Pop<S>::template ld<T>(a.G, pix(q, j, a.lay));
            const T dr = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
                         f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
            const T inv = T(1) / (T(1) + dr);
            vx = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) * inv;
            vy = (f[3] - f[4] + f[7] - f[8] - f[9] + f[10] + f[15] - f[16] + f[17] - f[18]) * inv;
            vz = (f[5] - f[6] + f[11] - f[12] - f[13] + f[14] + f[15] - f[16] - f[17] + f[18]) * inv;
        }
        // _writeEq at rho = 1 in the reference's expression order (js/lbm.js:155-162)
        const T r = T(1);
        const T usq = T(1.5) * (vx * vx + vy * vy + vz * vz);
        S v[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const T cu = T(3) * ((T)EX[q] * vx + (T)EY[q] * vy + (T)EZ[q] * vz);
            const T w = q == 0 ? T(1.0 / 3) : (q <= 6 ? T(1.0 / 18) : T(1.0 / 36));
            v[q] = to_store<S>((double)(w * ((r - T(1)) + r * (cu + T(0.5) * cu * cu - usq))));
            a.G[pix(q, i, a.lay)] = v[q];
        }
        if (a.macro) { a.rho[i] = (MAC)1; a.ux[i] = (MAC)vx; a.uy[i] = (MAC)vy; a.uz[i] = (MAC)vz; }
        if (y == 1 || y == a.ny - 2 || z == 1 || z == a.nz - 2) reflect_out(a, a.nx - 1, y, z, v);
        return;
    }
    k -= nout;
    if (k >= (int64_t)(a.nx - 2) * a.nring) return;
    const int x = 1 + (int)(k % (a.nx - 2));
    const int2 yz = a.ring[k / (a.nx - 2)];
    const int64_t i = x + (int64_t)a.nx * (yz.x + (int64_t)a.ny * yz.y);
    if (a.solid[i]) return;
    S v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) v[q] = a.G[pix(q, i, a.lay)];
    reflect_out(a, x, yz.x, yz.y, v);
}

// ------------------------------------------------------- reset / voxelize

template <class S, class MAC> struct InitArgs {
    S *f, *g;  // AB: f and g; esoteric: f = the single buffer, g unused
    MAC *rho, *ux, *uy, *uz;
    const uint8_t *solid;
    uint8_t *ops;               // voxelize: per-cell ops, cleared once applied
    int64_t n, ns;
    int64_t c0, c1;             // voxelize: cells [c0, c1) that may carry ops
    Lay lay;
    int nx, ny, nz, rep, eso;
    double u0;
    double eq_u0[Q], eq_rest[Q];
};

// Write one cell's populations as the current state (both AB buffers when
// `both`, or the current esoteric representation).
template <class S, class MAC>
__device__ void write_state(const InitArgs<S, MAC> &a, int64_t i, const double *v, bool both)
{
    if (a.eso) {
        const int x = (int)(i % a.nx), y = (int)((i / a.nx) % a.ny), z = (int)(i / ((int64_t)a.nx * a.ny));
        for (int q = 0; q < Q; q++)
            if (slot_inside(q, x, y, z, a.nx, a.ny, a.nz))
                Pop<S>::template st<double>(a.f, slot_of(q, i, a.rep, a.lay, a.nx, a.ny), v[q]);
        return;
    }
    for (int q = 0; q < Q; q++) {
        Pop<S>::template st<double>(a.f, pix(q, i, a.lay), v[q]);
        if (both) Pop<S>::template st<double>(a.g, pix(q, i, a.lay), v[q]);
    }
}

// reset(): equilibrium at u0 (fluid) or at rest (solid) in both buffers
// (js/lbm.js:164-174).
template <class S, class MAC>
__global__ void __launch_bounds__(256) k_reset(InitArgs<S, MAC> a)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const bool s = a.solid[i] != 0;
    write_state(a, i, s ? a.eq_rest : a.eq_u0, true);
    a.rho[i] = 1; a.ux[i] = s ? (MAC)0 : (MAC)a.u0; a.uy[i] = 0; a.uz[i] = 0;
}

// voxelize updates (js/lbm.js:197-233): bit 0 = _seedCell (u0 equilibrium in
// f and g), bit 1 = body cell (rest equilibrium in f, fields at rest);
// applied in that order.
template <class S, class MAC>
__global__ void __launch_bounds__(256) k_apply_ops(InitArgs<S, MAC> a)
{
    const int64_t i = a.c0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.c1) return;
    const int op = a.ops[i];
    if (!op) return;
    a.ops[i] = 0;  // the buffer stays zero between voxelize calls
    if (op & 1) {
        write_state(a, i, a.eq_u0, true);
        a.rho[i] = 1; a.ux[i] = (MAC)a.u0; a.uy[i] = 0; a.uz[i] = 0;
    }
    if (op & 2) {
        write_state(a, i, a.eq_rest, false);
        a.rho[i] = 1; a.ux[i] = 0; a.uy[i] = 0; a.uz[i] = 0;
    }
}

// Device buffer -> the reference's f[q*n + cell] order. AB: the buffer's
// value of every cell. Esoteric: the current representation; values a cell
// never stores (solid cells; out-of-lattice slots of boundary cells) read 0.
// The inverse: the reference's f[q*n + cell] order into the device buffer
// (AB: the buffer `B`; in place: the current representation, where a solid
// cell and a boundary cell's outward slots hold nothing and are skipped).
template <class S, class IN>
__global__ void __launch_bounds__(256) k_write_pops(S *B, const uint8_t *solid, const IN *in, int64_t n, int nx,
                                                    int ny, int nz, int rep, int eso, Lay lay)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    for (int q = 0; q < Q; q++) {
        const double v = (double)in[q * n + i];
        if (!eso) Pop<S>::template st<double>(B, pix(q, i, lay), v);
        else if (!solid[i] && slot_inside(q, x, y, z, nx, ny, nz))
            Pop<S>::template st<double>(B, slot_of(q, i, rep, lay, nx, ny), v);
    }
}

template <class S, class OUT>
__global__ void __launch_bounds__(256) k_read_pops(const S *B, const uint8_t *solid, OUT *out, int64_t n, int nx,
                                                   int ny, int nz, int rep, int eso, Lay lay)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    for (int q = 0; q < Q; q++) {
        OUT v = 0;
        if (!eso) v = (OUT)Pop<S>::template ld<double>(B, pix(q, i, lay));
        else if (!solid[i] && slot_inside(q, x, y, z, nx, ny, nz))
            v = (OUT)Pop<S>::template ld<double>(B, slot_of(q, i, rep, lay, nx, ny));
        out[q * n + i] = v;
    }
}

// Bit-packed solid mask and near-wall mask (js/lbm.js:268-282): near = an
// interior fluid cell with a solid cell among its 18 upstream neighbours
// (or, with gray weights, a gray cell or one next to a gray cell).
__global__ void __launch_bounds__(256) k_masks(const uint8_t *solid, const float *wgt, uint32_t *sbits,
                                               uint32_t *nbits, int nx, int ny, int nz, int nwx, int64_t w0,
                                               int64_t w1)
{
    // one thread per cell of words [w0, w1) (rows padded to whole words);
    // each half wavefront is one word, assembled by a ballot
    const int64_t t = w0 * 32 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t w = t >> 5;
    const int rw = nwx * 32;  // cells of a padded row
    const int64_t row = t / rw;
    const int x = (int)(t - row * rw);
    const int y = (int)(row % ny), z = (int)(row / ny);
    bool s = false, near = false;
    if (w < w1 && x < nx) {
        const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
        s = solid[i] != 0;
        if (!s && x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2) {
            near = wgt && wgt[i] > 0.0f;  // a gray cell: every link of it is gray
#pragma unroll
            for (int q = 1; q < Q; q++) {
                const int64_t j = i - cell_off(q, nx, ny);
                near |= solid[j] != 0 || (wgt && wgt[j] > 0.0f);
            }
        }
    }
    const uint64_t sm = __ballot(s), nm = __ballot(near);
    const int lane = threadIdx.x & 63;
    if ((lane & 31) == 0 && w < w1) {
        const int sh = lane;  // 0 or 32
        sbits[w] = (uint32_t)(sm >> sh);
        nbits[w] = (uint32_t)(nm >> sh);
    }
}

// Near-wall cells of each plane z0 + blockIdx.x (one block per plane).
__global__ void __launch_bounds__(256) k_near_count(const uint32_t *nbits, int64_t pw, int z0, int *cnt)
{
    const uint32_t *pb = nbits + (int64_t)(z0 + (int)blockIdx.x) * pw;
    int s = 0;
    for (int64_t w = threadIdx.x; w < pw; w += blockDim.x) s += __popc(pb[w]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// The near-wall list of planes z0 + blockIdx.x (one block per plane), in
// cell order: plane p starts at head + cnt[0] + ... + cnt[p-1]; each thread
// takes a run of the plane's words, placed by a block scan of their counts.
__global__ void __launch_bounds__(256) k_near_fill(const uint32_t *nbits, int64_t pw, int nwx, int nx, int ny, int z0,
                                                   const int *cnt, int64_t head, uint32_t *out)
{
    const int p = (int)blockIdx.x, z = z0 + p, t = (int)threadIdx.x;
    const uint32_t *pb = nbits + (int64_t)z * pw;
    __shared__ int64_t base;
    __shared__ int sc[256];
    if (t == 0) {
        int64_t b = head;
        for (int k = 0; k < p; k++) b += cnt[k];
        base = b;
    }
    const int64_t per = (pw + 255) / 256;
    const int64_t wa = std::min<int64_t>(pw, t * per), wb = std::min<int64_t>(pw, wa + per);
    int s = 0;
    for (int64_t w = wa; w < wb; w++) s += __popc(pb[w]);
    sc[t] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = t >= off ? sc[t - off] : 0;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    int64_t o = base + sc[t] - s;
    for (int64_t w = wa; w < wb; w++) {
        uint32_t bits = pb[w];
        const int wx = (int)(w % nwx), y = (int)(w / nwx);
        while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            out[o++] = (uint32_t)(wx * 32 + b + (int64_t)nx * (y + (int64_t)ny * z));
        }
    }
}

// ECMAScript ToInt32 of a double (the `|0` operator).
__device__ __forceinline__ int js_toint32(double d)
{
    if (!(d == d) || isinf(d)) return 0;
    const double t = trunc(d);
    if (t > -2147483649.0 && t < 2147483648.0) return (int)t;
    double m = fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    return (int)(uint32_t)m;
}

struct GridArgs {
    const uint8_t *grid;
    uint8_t *mask;
    int gnx, gny, gnz;
    double gmin[3], ginv, M[9], pos[3], inv;
    int x0, y0, z0, bx, by, bz;
};

// Grid-backed Body.makeTest at every cell centre of the box
// (js/shapes.js:376-388): world -> body by M^T / scale, truncating lookup.
__global__ void __launch_bounds__(256) k_grid_mask(GridArgs a)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t cnt = (int64_t)a.bx * a.by * a.bz;
    if (k >= cnt) return;
    const int x = a.x0 + (int)(k % a.bx);
    const int y = a.y0 + (int)((k / a.bx) % a.by);
    const int z = a.z0 + (int)(k / ((int64_t)a.bx * a.by));
    const double X = x + 0.5, Y = y + 0.5, Z = z + 0.5;
    const double dx = X - a.pos[0], dy = Y - a.pos[1], dz = Z - a.pos[2];
    const double lx = (a.M[0] * dx + a.M[3] * dy + a.M[6] * dz) * a.inv;
    const double ly = (a.M[1] * dx + a.M[4] * dy + a.M[7] * dz) * a.inv;
    const double lz = (a.M[2] * dx + a.M[5] * dy + a.M[8] * dz) * a.inv;
    uint8_t m = 0;
    const int ia = js_toint32((lx - a.gmin[0]) * a.ginv);
    if (ia >= 0 && ia < a.gnx) {
        const int ib = js_toint32((ly - a.gmin[1]) * a.ginv);
        if (ib >= 0 && ib < a.gny) {
            const int ic = js_toint32((lz - a.gmin[2]) * a.ginv);
            if (ic >= 0 && ic < a.gnz) m = a.grid[ia + (int64_t)a.gnx * (ib + (int64_t)a.gny * ic)] != 0;
        }
    }
    a.mask[k] = m;
}

template <class S, class MAC> InitArgs<S, MAC> init_args(wt_lattice *h, uint8_t *ops)
{
    InitArgs<S, MAC> a{};
    const bool eso = h->layout == WT_LAYOUT_ESOTERIC;
    a.f = (S *)h->d_buf[eso ? 0 : h->parity];
    a.g = eso ? nullptr : (S *)h->d_buf[h->parity ^ 1];
    MAC *m = (MAC *)h->d_mac;
    a.rho = m; a.ux = m + h->n; a.uy = m + 2 * h->n; a.uz = m + 3 * h->n;
    a.solid = h->d_solid; a.ops = ops; a.n = h->n; a.ns = h->ns; a.lay = Lay{h->ns, h->bsh}; a.u0 = h->p.u0;
    a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.rep = h->parity; a.eso = eso;
    eq_deviation(1, h->p.u0, 0, 0, a.eq_u0);
    eq_deviation(1, 0, 0, 0, a.eq_rest);
    return a;
}

template <class S, class MAC> int reset_t(wt_lattice *h)
{
    auto a = init_args<S, MAC>(h, nullptr);
    const int64_t nb = (h->n + 255) / 256;
    hipLaunchKernelGGL((k_reset<S, MAC>), dim3((unsigned)nb), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    return 0;
}

template <class S, class MAC> int ops_t(wt_lattice *h, uint8_t *d_ops, int64_t c0, int64_t c1)
{
    auto a = init_args<S, MAC>(h, d_ops);
    a.c0 = c0;
    a.c1 = c1;
    const int64_t nb = (c1 - c0 + 255) / 256;
    if (nb <= 0) return 0;
    hipLaunchKernelGGL((k_apply_ops<S, MAC>), dim3((unsigned)nb), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    return 0;
}

template <class S, class T, class MAC, bool MACRO, bool LES, bool TRT, bool ESO>
int step_t(wt_lattice *h, int bz0, int bz1, double *d_force)
{
    StepArgs<S, MAC> a{};
    if (ESO) {
        a.F = (const S *)h->d_buf[0];
        a.G = (S *)h->d_buf[0];
    } else {
        a.F = (const S *)(h->ovF ? h->ovF : h->d_buf[h->parity]);
        a.G = (S *)(h->ovG ? h->ovG : h->d_buf[h->parity ^ 1]);
    }
    a.phase = h->parity;
    {
        // adjoint replays take the recorded step's seed, so they reproduce it bit for bit
        const uint64_t st = (uint64_t)(h->seed_step >= 0 ? h->seed_step : h->steps);
        a.seed = mix32((uint32_t)st ^ mix32((uint32_t)(st >> 32) + 0x51ED270Bu));
    }
    a.solid = h->d_solid; a.sbits = h->d_sbits; a.nbits = h->d_nbits;
    a.uwx = h->d_uw; a.uwy = h->d_uw + h->n; a.uwz = h->d_uw + 2 * h->n;
    a.wgt = h->d_wgt;
    MAC *m = (MAC *)h->d_mac;
    a.rho = m; a.ux = m + h->n; a.uy = m + 2 * h->n; a.uz = m + 3 * h->n;
    // two partial buffers: this launch writes one while it reduces the
    // previous launch's partials from the other
    h->pslot ^= 1;
    a.partial = h->d_partial + (size_t)h->pslot * 3 * h->partial_cap;
    a.prev_partial = h->pend_partial;
    a.prev_out = h->pend_out;
    a.prev_count = h->pend_count;
    a.n = h->n; a.ns = h->ns; a.lay = Lay{h->ns, h->bsh}; a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.nwx = h->nwx; a.z0 = bz0;
    a.mov = h->p.movingWalls;
    a.rho_near = !h->replay;
    // Sweep order (h->sweep): bit 0 reverses z, bit 1 reverses y on every
    // other step, so the rows a step wrote last are the first the next step
    // reads while they are still in the 256 MB memory-side cache. The
    // result does not depend on the order.
    const int odd = (int)(h->steps & 1);
    a.revz = (h->sweep & 1) ? odd : 0;
    a.revy = (h->sweep & 2) ? odd : 0;
    a.tau = h->p.tau; a.sq18cs2 = (18 * std::sqrt(2.0)) * (h->p.csm * h->p.csm); a.lambda = h->p.lambda;
    a.om = 1 / a.tau;
    a.tau_le_half = !(a.tau > 0.5);
    a.tauf = (float)a.tau; a.omf = (float)(1 / a.tau); a.sq18f = (float)a.sq18cs2; a.lamf = (float)a.lambda;
    a.u0 = h->p.u0;
    eq_deviation(1, h->p.u0, 0, 0, a.eqv);
    // interior addressing tables (rows are whole blocks: nx % 2^bsh == 0)
    a.lsh = h->bsh;
    a.lstride = h->bsh ? Q << h->bsh : 1;
    a.lmask = h->bsh ? (1 << h->bsh) - 1 : 0;
    {
        const int ph = h->parity, nph = ph ^ 1;
        // element offset of slot `slot` of the row-0 cell shifted by dyz cells (a multiple of nx)
        auto uoff = [&](int slot, int64_t dyz) -> int64_t {
            const int64_t s = SLOT[slot];
            return (h->bsh ? s << h->bsh : s * h->ns) + (dyz >> h->bsh) * a.lstride;
        };
        for (int q = 0; q < Q; q++) {
            const int64_t dyz = (int64_t)h->nx * (EY[q] + (int64_t)h->ny * EZ[q]);
            int ls = q, ss = q;
            int64_t ld = -dyz, sd = 0;
            if (ESO && q) {
                if (q & 1) { ls = ph ? q : q + 1; ld = 0; ss = nph ? q : q + 1; sd = dyz; }
                else { ls = ph ? q : q - 1; ss = nph ? q : q - 1; }
            }
            a.lp[q] = (const S *)((uintptr_t)a.F + (uintptr_t)(uoff(ls, ld) * (int64_t)sizeof(S)));
            a.sp[q] = (S *)((uintptr_t)a.G + (uintptr_t)(uoff(ss, sd) * (int64_t)sizeof(S)));
        }
    }
    if (!ESO) {
        // buffer-resource offsets of the straight path: pulls as bytes from
        // the lowest pull slot of the row, stores from G + row term
        int64_t lmin = 0;
        for (int q = 0; q < Q; q++) lmin = std::min<int64_t>(lmin, (int64_t)a.lp[q] - (int64_t)a.F);
        bool ok = true;
        for (int q = 0; q < Q; q++) {
            const int64_t lo = (int64_t)a.lp[q] - (int64_t)a.F - lmin, so = (int64_t)a.sp[q] - (int64_t)a.G;
            ok = ok && lo <= INT32_MAX && so >= 0 && so <= INT32_MAX;
            a.loff[q] = ok ? (int32_t)lo : 0;
            a.soff[q] = ok ? (int32_t)so : 0;
        }
        a.lmin = lmin;
        a.bufok = ok;
    }
    const int bx = row_block(h->nx);
    const int gx = (h->nx + bx - 1) / bx;
    // the near-wall cells of rows z in [bz0, bz1), one per thread of the
    // near-wall blocks; each of those blocks writes one force partial
    a.near = h->d_near;
    a.near0 = h->near_z[(size_t)bz0];
    a.near_cnt = h->near_z[(size_t)bz1] - a.near0;
    const int64_t row_blocks = (int64_t)gx * h->ny;
    int64_t near_blocks = fused_near<S, T>() ? (a.near_cnt + bx - 1) / bx
                                          : std::min<int64_t>((a.near_cnt + NEAR_THREADS - 1) / NEAR_THREADS,
                                                              h->partial_cap);
    a.near_planes = fused_near<S, T>() ? (int)((near_blocks + row_blocks - 1) / row_blocks) : 0;
    bool paired = false;
    if constexpr ((std::is_same<S, __half>::value || std::is_same<S, float>::value) && std::is_same<T, float>::value &&
                  !ESO && WT_FAST_COLLIDE) {
        // two cells per thread, a wave 128 cells: blocked rows of 128-cell multiples
        const int tpr = h->bsh && h->nx % 128 == 0 ? h->nx / 2 : 0;  // threads per row: a multiple of 64
        int bxp = 1024;                                // the widest workgroup of whole waves dividing the row
        while (bxp > 64 && tpr % bxp) bxp -= 64;
        const int64_t prow = (int64_t)(tpr / bxp) * h->ny;
        // near-wall blocks of 2 bxp listed cells: at most nz planes of them
        const int64_t nblk = (a.near_cnt + 2 * bxp - 1) / (2 * bxp);
        const int np = prow > 0 ? (int)((nblk + prow - 1) / prow) : 0;
        // binary16 by default; float32 only on request (WT_PAIRS=2): with
        // 4-byte values, twice the cells in flight per wave made the float32
        // step 11 % slower on MI355X (0.989 against 0.892 ms at 512x256x256)
        const bool want = std::is_same<S, __half>::value ? h->pairs >= 1 : h->pairs >= 2;
        if (want && a.bufok && tpr > 0 && tpr % 64 == 0 && np * prow <= h->partial_cap) {
            for (int q = 0; q < Q; q++) {
                const uint32_t b = __half_as_ushort(__float2half_rn((float)a.eqv[q] * F16_SCALE));
                a.eqw[q] = b | (b << 16);
            }
            a.near_planes = np;
            // Register budget: 4 waves per SIMD, or 5 when a workgroup's
            // wave count does not divide the CU's 16 four-wave slots (a
            // 640-cell row is 5 waves: 15 of 16 slots at 4, all 20 at 5;
            // 640x320x320 steps 4.5 % faster, while 512-cell rows lose 1-3 %
            // at 5)
            const int wpw = bxp / 64;
            if (16 % wpw)
                hipLaunchKernelGGL((k_step_pair<S, MACRO, LES, TRT, WT_PAIR_WAVES + 1>),
                                   dim3(tpr / bxp, h->ny, bz1 - bz0 + a.near_planes), dim3(bxp), 0, h->stream, a);
            else
                hipLaunchKernelGGL((k_step_pair<S, MACRO, LES, TRT, WT_PAIR_WAVES>),
                                   dim3(tpr / bxp, h->ny, bz1 - bz0 + a.near_planes), dim3(bxp), 0, h->stream, a);
            paired = true;
            near_blocks = a.near_planes * prow;
        }
    }
    if (!paired)
        hipLaunchKernelGGL((k_step<S, T, MAC, MACRO, LES, TRT, ESO, step_waves<S, T>()>),
                           dim3(gx, h->ny, bz1 - bz0 + a.near_planes), dim3(bx), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    // force partials of this launch: one per near-wall block
    const int64_t count = paired ? near_blocks : fused_near<S, T>() ? a.near_planes * row_blocks : near_blocks;
    if constexpr (!fused_near<S, T>()) {
        if (near_blocks > 0 && !paired) {
            hipLaunchKernelGGL((k_near<S, T, MAC, MACRO, LES, TRT, ESO>), dim3((unsigned)near_blocks),
                               dim3(NEAR_THREADS), 0, h->stream, a);
            WT_HIP(hipGetLastError());
        }
    }
    // this launch's force is reduced by the next launch, or by lbm_flush_force
    // (build switch WT_DEFER_FORCE=0: by its own launch right away)
    h->pend_partial = a.partial;
    h->pend_out = d_force;
    h->pend_count = d_force ? count : 0;
    if (!d_force) h->pend_partial = nullptr;
#if !WT_DEFER_FORCE
    return lbm_flush_force(h);
#else
    return 0;
#endif
}

template <class S, class T, class MAC, bool ESO>
int step_dispatch(wt_lattice *h, int bz0, int bz1, bool macro, double *f)
{
    const bool les = h->p.les, trt = h->p.trt;
#define WT_CASE(M, L, R) \
    if (macro == M && les == L && trt == R) return step_t<S, T, MAC, M, L, R, ESO>(h, bz0, bz1, f);
    WT_CASE(true, true, true) WT_CASE(true, true, false) WT_CASE(true, false, true) WT_CASE(true, false, false)
    WT_CASE(false, true, true) WT_CASE(false, true, false) WT_CASE(false, false, true) WT_CASE(false, false, false)
#undef WT_CASE
    return wt_fail(WT_E_STATE, "unreachable step configuration");
}

template <class S, class T, class MAC> int step_layout(wt_lattice *h, int bz0, int bz1, bool macro, double *f)
{
    if (h->layout == WT_LAYOUT_ESOTERIC) return step_dispatch<S, T, MAC, true>(h, bz0, bz1, macro, f);
    return step_dispatch<S, T, MAC, false>(h, bz0, bz1, macro, f);
}

}  // namespace
}  // namespace wt

using namespace wt;

// _writeEq's formula for one cell (js/lbm.js:155-162), binary64 in the
// reference's order; the kernels round these values to the storage type.
void eq_deviation(double r, double vx, double vy, double vz, double out[19])
{
    const double W[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
    const double usq = 1.5 * (vx * vx + vy * vy + vz * vz);
    for (int q = 0; q < Q; q++) {
        const double cu = 3 * ((double)EX[q] * vx + (double)EY[q] * vy + (double)EZ[q] * vz);
        const double w = q == 0 ? W[0] : (q <= 6 ? W[1] : W[2]);
        out[q] = w * ((r - 1) + r * (cu + 0.5 * cu * cu - usq));
    }
}

int lbm_reset(wt_lattice *h)
{
    switch (h->storage) {
    case WT_STORE_F32: return reset_t<float, float>(h);
    case WT_STORE_F16: return reset_t<__half, float>(h);
    default: return reset_t<double, double>(h);
    }
}

int lbm_apply_voxelize(wt_lattice *h, uint8_t *d_ops, int64_t c0, int64_t c1)
{
    switch (h->storage) {
    case WT_STORE_F32: return ops_t<float, float>(h, d_ops, c0, c1);
    case WT_STORE_F16: return ops_t<__half, float>(h, d_ops, c0, c1);
    default: return ops_t<double, double>(h, d_ops, c0, c1);
    }
}

int lbm_rebuild_masks(wt_lattice *h) { return lbm_rebuild_masks_range(h, 0, h->nz - 1); }

int lbm_rebuild_masks_range(wt_lattice *h, int z0, int z1)
{
    // without a valid list, the whole lattice
    if (h->near_full || h->near_z.size() != (size_t)h->nz + 1) { z0 = 0; z1 = h->nz - 1; }
    const int64_t pw = (int64_t)h->nwx * h->ny;  // words per plane
    const int64_t w0 = pw * z0, w1 = pw * (z1 + 1);
    const int P = z1 - z0 + 1;
    if (P <= 0) return 0;
    hipLaunchKernelGGL(k_masks, dim3((unsigned)(((w1 - w0) * 32 + 255) / 256)), dim3(256), 0, h->stream,
                       h->d_solid, h->d_wgt, h->d_sbits, h->d_nbits, h->nx, h->ny, h->nz, h->nwx, w0, w1);
    WT_HIP(hipGetLastError());
    // near-wall cells per plane of the range
    if (!h->d_pcnt && hipMalloc((void **)&h->d_pcnt, sizeof(int) * (size_t)h->nz) != hipSuccess) {
        h->d_pcnt = nullptr;
        return wt_fail(WT_E_NOMEM, "hipMalloc of the near-wall plane counts failed");
    }
    hipLaunchKernelGGL(k_near_count, dim3((unsigned)P), dim3(256), 0, h->stream, (const uint32_t *)h->d_nbits, pw, z0,
                       h->d_pcnt);
    WT_HIP(hipGetLastError());
    std::vector<int> cnt((size_t)P);
    WT_HIP(hipMemcpyAsync(cnt.data(), h->d_pcnt, sizeof(int) * (size_t)P, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    // new layout: the planes before z0 and after z1 keep their cells
    std::vector<int64_t> oz = h->near_z;
    if (oz.size() != (size_t)h->nz + 1) oz.assign((size_t)h->nz + 1, 0);
    const int64_t head = oz[(size_t)z0], tail0 = oz[(size_t)z1 + 1], told = oz[(size_t)h->nz];
    int64_t mid = 0;
    for (int c : cnt) mid += c;
    const int64_t total = head + mid + (told - tail0);
    h->near_full = true;  // until the new list is in place
    uint32_t *src = h->d_near, *dst = h->d_near2;
    uint32_t *grown[2] = {nullptr, nullptr};
    if (total > h->near_cap) {
        const int64_t cap = total + total / 4 + 1024;
        for (auto &g : grown)
            if (hipMalloc((void **)&g, sizeof(uint32_t) * (size_t)cap) != hipSuccess) {
                if (grown[0]) (void)hipFree(grown[0]);
                if (grown[1]) (void)hipFree(grown[1]);
                return wt_fail(WT_E_NOMEM, "hipMalloc of the near-wall list (%lld cells) failed", (long long)cap);
            }
        dst = grown[0];
        h->near_cap = cap;
    }
    if (head > 0)
        WT_HIP(hipMemcpyAsync(dst, src, sizeof(uint32_t) * (size_t)head, hipMemcpyDeviceToDevice, h->stream));
    if (mid > 0) {
        hipLaunchKernelGGL(k_near_fill, dim3((unsigned)P), dim3(256), 0, h->stream, (const uint32_t *)h->d_nbits, pw,
                           h->nwx, h->nx, h->ny, z0, (const int *)h->d_pcnt, head, dst);
        WT_HIP(hipGetLastError());
    }
    if (told > tail0)
        WT_HIP(hipMemcpyAsync(dst + head + mid, src + tail0, sizeof(uint32_t) * (size_t)(told - tail0),
                              hipMemcpyDeviceToDevice, h->stream));
    if (grown[0]) {
        WT_HIP(hipStreamSynchronize(h->stream));
        if (h->d_near) WT_HIP(hipFree(h->d_near));
        if (h->d_near2) WT_HIP(hipFree(h->d_near2));
        h->d_near = grown[0];
        h->d_near2 = grown[1];
    } else {
        std::swap(h->d_near, h->d_near2);
    }
    h->near_z.assign((size_t)h->nz + 1, 0);
    int64_t acc = head;
    for (int z = 0; z <= h->nz; z++) {
        if (z < z0) h->near_z[(size_t)z] = oz[(size_t)z];
        else if (z <= z1) { h->near_z[(size_t)z] = acc; acc += cnt[(size_t)(z - z0)]; }
        else h->near_z[(size_t)z] = oz[(size_t)z] + (head + mid - tail0);
    }
    h->near_full = false;
    return 0;
}

template <class E>
__global__ void __launch_bounds__(256) k_scatter_box(const E *src, E *dst, int bx0, int by0, int bz0, int w, int hgt,
                                                     int64_t cnt, int nx, int ny)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int x = (int)(t % w);
    const int64_t r = t / w;
    const int y = (int)(r % hgt), z = (int)(r / hgt);
    dst[(bx0 + x) + (int64_t)nx * ((by0 + y) + (int64_t)ny * (bz0 + z))] = src[t];
}

// Device re-voxelize of a body given as a device mask (the road's solid
// state unchanged: the caller checks). One thread per cell of the union of
// the previous body's box and the new mask's box: the previous body's cells
// become fluid (op 1), the mask's cells body (op 2), and their wall
// velocities are set (mask 2: the per-cell velocity, otherwise zero), as the
// host bookkeeping does (js/lbm.js:194-246). Reduced per workgroup into res:
// body cells, the changed box, the new body's box, wheel cells, the change of
// the count of cells with a nonzero wall velocity and the box of changed wall
// velocities. The projections go to col / plan (bytes) and the lowest cell of
// each (x, z) column to low.
struct RevoxArgs {
    uint8_t *solid, *ops;
    const uint8_t *mask;
    const float *wv;                  // 3 per mask cell (mask 2 cells), or null
    float *uwx, *uwy, *uwz;
    int nx, ny, nz;
    int ux, uy, uz, uw, uh, ud;       // union box origin and extents
    int mx0, my0, mz0, mw, mh, md;    // mask box origin and extents (md 0: no mask)
    float road;                       // the road row's x velocity (ground), else 0
    int ground;                       // the road row is wall
    int *res;  // REVOX_* slots
    uint8_t *col, *plan;
    int *low, *hist;
};

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_add(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the 21 reduced values of a thread: kind 0 sum, 1 min, 2 max; slot in res
__device__ constexpr int rv_kind(int k)
{
    return (k == 0 || k == 13 || k == 14) ? 0
                                          : ((k >= 1 && k <= 3) || (k >= 7 && k <= 9) || (k >= 15 && k <= 17)) ? 1 : 2;
}
__device__ constexpr int rv_slot(int k) { return k < 13 ? k : (k == 13 ? RV_WHEELS : (k == 14 ? RV_NZ : RV_UB + (k - 15))); }

// wall velocity of cell i set to (vx, vy, vz) where it differs; records the
// change in v (nonzero-count delta at 14, changed box at 15-20)
__device__ __forceinline__ void revox_set_uw(const RevoxArgs &a, int64_t i, int x, int y, int z, float vx, float vy,
                                             float vz, int (&v)[21])
{
    const float ox = a.uwx[i], oy = a.uwy[i], oz = a.uwz[i];
    if (ox != vx || oy != vy || oz != vz) {
        a.uwx[i] = vx; a.uwy[i] = vy; a.uwz[i] = vz;
        v[14] += (int)(vx != 0 || vy != 0 || vz != 0) - (int)(ox != 0 || oy != 0 || oz != 0);
        v[15] = min(v[15], x); v[16] = min(v[16], y); v[17] = min(v[17], z);
        v[18] = max(v[18], x); v[19] = max(v[19], y); v[20] = max(v[20], z);
    }
}

__device__ __forceinline__ void revox_reduce(const RevoxArgs &a, int (&v)[21])
{
    const int big = 0x7fffffff;
    __shared__ int part[4][21];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 21; k++) {
        const int r = rv_kind(k) == 0 ? wave_add(v[k]) : (rv_kind(k) == 1 ? wave_min(v[k]) : wave_max(v[k]));
        if (lane == 0) part[wv][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 21) {
        const int k = threadIdx.x, kind = rv_kind(k);
        int w = part[0][k];
        for (int j = 1; j < 4; j++) {
            const int u = part[j][k];
            w = kind == 0 ? w + u : (kind == 1 ? min(w, u) : max(w, u));
        }
        int *dst = a.res + rv_slot(k);
        if (kind == 0) { if (w) atomicAdd(dst, w); }
        else if (kind == 1) { if (w != big) atomicMin(dst, w); }
        else if (w >= 0) atomicMax(dst, w);
    }
}

__device__ __forceinline__ void revox_init(int (&v)[21])
{
    const int big = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < 21; k++) v[k] = rv_kind(k) == 0 ? 0 : (rv_kind(k) == 1 ? big : -1);
}

__global__ void __launch_bounds__(256) k_revox(RevoxArgs a)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t cnt = (int64_t)a.uw * a.uh * a.ud;
    int v[21];
    revox_init(v);
    if (t < cnt) {
        const int x = a.ux + (int)(t % a.uw);
        const int64_t r = t / a.uw;
        const int y = a.uy + (int)(r % a.uh), z = a.uz + (int)(r / a.uh);
        const int64_t i = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
        const int mx = x - a.mx0, my = y - a.my0, mz = z - a.mz0;
        const bool in_box = a.md > 0 && mx >= 0 && mx < a.mw && my >= 0 && my < a.mh && mz >= 0 && mz < a.md;
        const int64_t k = in_box ? mx + (int64_t)a.mw * (my + (int64_t)a.mh * mz) : 0;
        const int mv = in_box ? a.mask[k] : 0;
        const bool in_new = mv != 0;
        const bool cleared = a.solid[i] == SOLID_OBJ;
        if (cleared || in_new) {
            a.ops[i] = (uint8_t)((cleared ? 1 : 0) | (in_new ? 2 : 0));
            a.solid[i] = in_new ? (uint8_t)SOLID_OBJ : (uint8_t)FLUID;
            v[1] = x; v[2] = y; v[3] = z; v[4] = x; v[5] = y; v[6] = z;
            float vx = 0, vy = 0, vz = 0;
            if (mv == 2 && a.wv) { vx = a.wv[3 * k]; vy = a.wv[3 * k + 1]; vz = a.wv[3 * k + 2]; }
            revox_set_uw(a, i, x, y, z, vx, vy, vz, v);
        }
        if (in_new) {
            v[0] = 1;
            v[13] = mv == 2;
            v[7] = x; v[8] = y; v[9] = z; v[10] = x; v[11] = y; v[12] = z;
            a.col[y + (int64_t)a.ny * z] = 1;
            a.plan[x + (int64_t)a.nx * z] = 1;
            atomicMin(&a.low[x + (int64_t)a.nx * z], y);
        }
    }
    revox_reduce(a, v);
}

// The road row (y = 0, js/lbm.js:204-214): wall (a.ground) or fluid (re-seeded:
// op 1), and its wall velocity (a.road, 0, 0), wherever they differ.
__global__ void __launch_bounds__(256) k_revox_road(RevoxArgs a)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int v[21];
    revox_init(v);
    if (t < (int64_t)a.nx * a.nz) {
        const int x = (int)(t % a.nx), z = (int)(t / a.nx);
        const int64_t i = x + (int64_t)a.nx * a.ny * z;
        const uint8_t gy = a.ground ? (uint8_t)SOLID_WALL : (uint8_t)FLUID;
        if (a.solid[i] != gy) {
            a.solid[i] = gy;
            if (gy == FLUID) a.ops[i] = 1;
            v[1] = x; v[2] = 0; v[3] = z; v[4] = x; v[5] = 0; v[6] = z;
        }
        revox_set_uw(a, i, x, 0, z, a.road, 0.0f, 0.0f, v);
    }
    revox_reduce(a, v);
}

// Projections of the new body: frontal area (y, z columns), planform area
// (x, z columns) and the histogram of its lowest cells' gaps above the floor
// (y - 1 > 0), from which the host takes the median (the ride height).
__global__ void __launch_bounds__(256) k_revox_stats(RevoxArgs a)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int c = 0, p = 0;
    if (t < (int64_t)a.ny * a.nz) c = a.col[t] != 0;
    if (t < (int64_t)a.nx * a.nz && a.plan[t]) {
        p = 1;
        const int g = a.low[t] - 1;
        if (g > 0) atomicAdd(&a.hist[g], 1);
    }
    c = wave_add(c);
    p = wave_add(p);
    if ((threadIdx.x & 63) == 0) {
        if (c) atomicAdd(&a.res[RV_FRONTAL], c);
        if (p) atomicAdd(&a.res[RV_PLAN], p);
    }
}

template <class E>
__global__ void __launch_bounds__(256) k_gather_box(const E *src, E *dst, int bx0, int by0, int bz0, int w, int hgt,
                                                    int64_t cnt, int nx, int ny)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int x = (int)(t % w);
    const int64_t r = t / w;
    const int y = (int)(r % hgt), z = (int)(r / hgt);
    dst[t] = src[(bx0 + x) + (int64_t)nx * ((by0 + y) + (int64_t)ny * (bz0 + z))];
}

int lbm_revoxelize_device(wt_lattice *h, const int ubox[6], const int mbox[6], const uint8_t *d_mask, const float *d_wv,
                          int ground, float road, void *d_work, int *res, int *hist)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    uint8_t *col = (uint8_t *)d_work, *plan = col + ((size_t)ny * nz + 255) / 256 * 256;
    int *low = (int *)(plan + ((size_t)nx * nz + 255) / 256 * 256);
    int *dres = low + (size_t)nx * nz, *dhist = dres + RV_N;
    WT_HIP(hipMemsetAsync(col, 0, (size_t)ny * nz, h->stream));
    WT_HIP(hipMemsetAsync(plan, 0, (size_t)nx * nz, h->stream));
    WT_HIP(hipMemsetAsync(low, 0x7f, sizeof(int) * (size_t)nx * nz, h->stream));
    WT_HIP(hipMemsetAsync(dhist, 0, sizeof(int) * (size_t)ny, h->stream));
    WT_HIP(hipMemcpyAsync(dres, res, sizeof(int) * RV_N, hipMemcpyHostToDevice, h->stream));
    RevoxArgs a{};
    a.solid = h->d_solid; a.ops = h->d_ops; a.mask = d_mask; a.wv = d_wv;
    a.uwx = h->d_uw; a.uwy = h->d_uw + h->n; a.uwz = h->d_uw + 2 * h->n;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.ux = ubox[0]; a.uy = ubox[1]; a.uz = ubox[2];
    a.uw = ubox[3] - ubox[0] + 1; a.uh = ubox[4] - ubox[1] + 1; a.ud = ubox[5] - ubox[2] + 1;
    a.mx0 = mbox[0]; a.my0 = mbox[1]; a.mz0 = mbox[2];
    const bool mask_empty = mbox[3] < mbox[0] || mbox[4] < mbox[1] || mbox[5] < mbox[2];
    a.mw = mask_empty ? 0 : mbox[3] - mbox[0] + 1;
    a.mh = mask_empty ? 0 : mbox[4] - mbox[1] + 1;
    a.md = mask_empty ? 0 : mbox[5] - mbox[2] + 1;
    a.road = road;
    a.ground = ground;
    a.res = dres; a.col = col; a.plan = plan; a.low = low; a.hist = dhist;
    if (a.uw > 0 && a.uh > 0 && a.ud > 0) {
        const int64_t cnt = (int64_t)a.uw * a.uh * a.ud;
        hipLaunchKernelGGL(k_revox, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, a);
        WT_HIP(hipGetLastError());
    }
    const int64_t rc = (int64_t)nx * nz;
    hipLaunchKernelGGL(k_revox_road, dim3((unsigned)((rc + 255) / 256)), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    const int64_t sc = std::max<int64_t>((int64_t)ny * nz, (int64_t)nx * nz);
    hipLaunchKernelGGL(k_revox_stats, dim3((unsigned)((sc + 255) / 256)), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    WT_HIP(hipMemcpyAsync(res, dres, sizeof(int) * RV_N, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipMemcpyAsync(hist, dhist, sizeof(int) * (size_t)ny, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int lbm_gather_box(wt_lattice *h, const void *src, void *dst, int elem, const int box[6])
{
    const int w = box[3] - box[0] + 1, hgt = box[4] - box[1] + 1, d = box[5] - box[2] + 1;
    if (w <= 0 || hgt <= 0 || d <= 0) return 0;
    if (box[0] < 0 || box[1] < 0 || box[2] < 0 || box[3] >= h->nx || box[4] >= h->ny || box[5] >= h->nz)
        return wt_fail(WT_E_ARG, "gather box outside the lattice");
    const int64_t cnt = (int64_t)w * hgt * d;
    const dim3 grid((unsigned)((cnt + 255) / 256));
    if (elem == 1)
        hipLaunchKernelGGL(k_gather_box<uint8_t>, grid, dim3(256), 0, h->stream, (const uint8_t *)src, (uint8_t *)dst,
                           box[0], box[1], box[2], w, hgt, cnt, h->nx, h->ny);
    else if (elem == 4)
        hipLaunchKernelGGL(k_gather_box<uint32_t>, grid, dim3(256), 0, h->stream, (const uint32_t *)src,
                           (uint32_t *)dst, box[0], box[1], box[2], w, hgt, cnt, h->nx, h->ny);
    else
        return wt_fail(WT_E_ARG, "gather element of %d bytes", elem);
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_scatter_box(wt_lattice *h, const void *src, void *dst, int elem, const int box[6])
{
    const int w = box[3] - box[0] + 1, hgt = box[4] - box[1] + 1, d = box[5] - box[2] + 1;
    if (w <= 0 || hgt <= 0 || d <= 0) return 0;
    if (box[0] < 0 || box[1] < 0 || box[2] < 0 || box[3] >= h->nx || box[4] >= h->ny || box[5] >= h->nz)
        return wt_fail(WT_E_ARG, "scatter box outside the lattice");
    const int64_t cnt = (int64_t)w * hgt * d;
    const dim3 grid((unsigned)((cnt + 255) / 256));
    if (elem == 1)
        hipLaunchKernelGGL(k_scatter_box<uint8_t>, grid, dim3(256), 0, h->stream, (const uint8_t *)src, (uint8_t *)dst,
                           box[0], box[1], box[2], w, hgt, cnt, h->nx, h->ny);
    else if (elem == 4)
        hipLaunchKernelGGL(k_scatter_box<uint32_t>, grid, dim3(256), 0, h->stream, (const uint32_t *)src,
                           (uint32_t *)dst, box[0], box[1], box[2], w, hgt, cnt, h->nx, h->ny);
    else
        return wt_fail(WT_E_ARG, "scatter element of %d bytes", elem);
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_launch_step(wt_lattice *h, int bz0, int bz1, bool macro, double *d_force)
{
    if (h->storage == WT_STORE_F32 && h->arith == WT_ARITH_F32)
        return step_layout<float, float, float>(h, bz0, bz1, macro, d_force);
    if (h->storage == WT_STORE_F32 && h->arith == WT_ARITH_F64)
        return step_layout<float, double, float>(h, bz0, bz1, macro, d_force);
    if (h->storage == WT_STORE_F16)
        return step_layout<__half, float, float>(h, bz0, bz1, macro, d_force);
    return step_layout<double, double, double>(h, bz0, bz1, macro, d_force);
}

template <class S, class T, class MAC> int channel_t(wt_lattice *h, bool macro)
{
    ChanArgs<S, MAC> a{};
    a.G = (S *)h->d_buf[h->parity ^ 1];
    a.solid = h->d_solid;
    MAC *m = (MAC *)h->d_mac;
    a.rho = m; a.ux = m + h->n; a.uy = m + 2 * h->n; a.uz = m + 3 * h->n;
    a.ring = (const int2 *)h->d_ring;
    a.nring = h->nring;
    a.lay = Lay{h->ns, h->bsh};
    a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.macro = macro;
    const int64_t total = (int64_t)(h->ny - 2) * (h->nz - 2) + (int64_t)(h->nx - 2) * h->nring;
    hipLaunchKernelGGL((k_channel<S, T, MAC>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_channel(wt_lattice *h, bool macro)
{
    if (h->storage == WT_STORE_F32 && h->arith == WT_ARITH_F32) return channel_t<float, float, float>(h, macro);
    if (h->storage == WT_STORE_F32) return channel_t<float, double, float>(h, macro);
    if (h->storage == WT_STORE_F16) return channel_t<__half, float, float>(h, macro);
    return channel_t<double, double, double>(h, macro);
}

int lbm_flush_force(wt_lattice *h)
{
    if (!h->pend_out) return 0;
    hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(64), 0, h->stream, h->pend_partial, h->pend_count, h->pend_out);
    h->pend_partial = nullptr;
    h->pend_out = nullptr;
    h->pend_count = 0;
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_set_pops(wt_lattice *h, const void *d_in)
{
    const bool eso = h->layout == WT_LAYOUT_ESOTERIC;
    void *B = eso ? h->d_buf[0] : h->d_buf[h->parity];
    const unsigned nb = (unsigned)((h->n + 255) / 256);
    const Lay lay{h->ns, h->bsh};
    if (h->storage == WT_STORE_F64)
        hipLaunchKernelGGL((k_write_pops<double, double>), dim3(nb), dim3(256), 0, h->stream, (double *)B, h->d_solid,
                           (const double *)d_in, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, lay);
    else if (h->storage == WT_STORE_F16)
        hipLaunchKernelGGL((k_write_pops<__half, float>), dim3(nb), dim3(256), 0, h->stream, (__half *)B, h->d_solid,
                           (const float *)d_in, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, lay);
    else
        hipLaunchKernelGGL((k_write_pops<float, float>), dim3(nb), dim3(256), 0, h->stream, (float *)B, h->d_solid,
                           (const float *)d_in, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, lay);
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_get_pops(wt_lattice *h, int which, void *d_out)
{
    const bool eso = h->layout == WT_LAYOUT_ESOTERIC;
    if (eso && which != 0) return wt_fail(WT_E_ARG, "an in-place lattice holds only the current populations");
    const void *B = eso ? h->d_buf[0] : h->d_buf[h->parity ^ which];
    const unsigned nb = (unsigned)((h->n + 255) / 256);
    if (h->storage == WT_STORE_F64)
        hipLaunchKernelGGL((k_read_pops<double, double>), dim3(nb), dim3(256), 0, h->stream, (const double *)B,
                           h->d_solid, (double *)d_out, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, Lay{h->ns, h->bsh});
    else if (h->storage == WT_STORE_F16)
        hipLaunchKernelGGL((k_read_pops<__half, float>), dim3(nb), dim3(256), 0, h->stream, (const __half *)B,
                           h->d_solid, (float *)d_out, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, Lay{h->ns, h->bsh});
    else
        hipLaunchKernelGGL((k_read_pops<float, float>), dim3(nb), dim3(256), 0, h->stream, (const float *)B,
                           h->d_solid, (float *)d_out, h->n, h->nx, h->ny, h->nz, h->parity, (int)eso, Lay{h->ns, h->bsh});
    WT_HIP(hipGetLastError());
    return 0;
}

int lbm_grid_mask(wt_lattice *h, const uint8_t *d_grid, const wt_voxmeta *vm, const double M[9],
                  const double pos[3], double scale, const int box[6], uint8_t *d_mask)
{
    GridArgs a{};
    a.grid = d_grid; a.mask = d_mask;
    a.gnx = vm->nx; a.gny = vm->ny; a.gnz = vm->nz;
    for (int k = 0; k < 3; k++) { a.gmin[k] = vm->min[k]; a.pos[k] = pos[k]; }
    for (int k = 0; k < 9; k++) a.M[k] = M[k];
    a.ginv = vm->inv; a.inv = 1 / scale;
    a.x0 = box[0]; a.y0 = box[1]; a.z0 = box[2];
    a.bx = box[3] - box[0] + 1; a.by = box[4] - box[1] + 1; a.bz = box[5] - box[2] + 1;
    const int64_t cnt = (int64_t)a.bx * a.by * a.bz;
    if (cnt <= 0) return 0;
    hipLaunchKernelGGL(k_grid_mask, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, a);
    WT_HIP(hipGetLastError());
    return 0;
}
