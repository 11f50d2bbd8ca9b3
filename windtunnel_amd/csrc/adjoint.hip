// Discrete adjoint of the gray-bounce-back step (no reference counterpart:
// an addition of this library) with sqrt(N) checkpointing.
//
// Forward step k maps the state s_k (post-collision populations, AB layout)
// to s_{k+1} and yields the force F_k. The objective is J = sum_k Fbar_k.F_k.
// A link (i <- j = i - c_q) between fluid cells transmits t_ij =
// (1 - w_i)(1 - w_j) of the pulled value and bounces W_ij = 1 - t_ij back
// (W = 1 from a solid cell; t = 1 between two cells of weight 0). With mu_k
// the adjoint of the pre-collision populations f_k of step k and lambda_k
// the adjoint of s_k, one reverse iteration k is ONE fused kernel:
//
//   lambda_k(m)[q] = t_{m+c_q,m} mu_k[q](m + c_q)                     (streaming)
//                  + W_{m,j} (mu_k[qbar](m) - 2 c_qbar . Fbar_k)       (bounce-back, j = m + c_q)
//   dJ/dw_m       += sum over the links (i <- j) with m = i or m = j of
//                    (1 - w_other) [mu_k[q](i) (s_k[qbar](i) - s_k[q](j)) - 2 s_k[qbar](i) c_q . Fbar_k]
//   mu_{k-1}(m)    = (d collide / d f)^T at f_{k-1}(m), applied to lambda_k(m)
//                    (plus lambda_k of the outlet cell for x = nx-2: the outlet
//                    copies that cell, js/lbm.js:504-511)
//
// It reads mu_k (19 gathered), s_{k-1} (19 gathered) and writes mu_{k-1}
// (19): 1.5x the bytes of a forward step when mu has the states' element
// size. States inside a checkpoint segment are recomputed by the forward
// kernel into a replay store, so the reverse sweep costs one recompute step
// plus one adjoint step per forward step.
//
// Binary16 storage: the states (checkpoints, replays) keep the forward's own
// binary16 offsets, so every replay is bit-identical to the recorded forward,
// and the rounding of a store is taken as the identity. The adjoint fields
// are binary16 too (float with WT_ADJ_FIELD=f32 in the environment): field
// mu_k is stored as mu_k * sigma_k, rounded to nearest, where sigma_k is the
// power of two that puts max|mu_{k+1}| + 2 |Fbar_{k+1}|_1 (a bound of
// lambda_{k+1} without the outlet fold) in [2^10, 2^11). One reverse step
// grows the field by far less than the 2^5 left below binary16's largest
// value; the kernel records each field's max |mu| (one atomic per wave that
// raises it) for the next launch's scale, and the host checks after the
// sweep that no stored value overflowed. When the replay store of a whole
// segment does not fit in device memory, each segment is swept in windows:
// the window's states are recomputed from the segment's checkpoint, and only
// those are kept.

#include "collide_adj.h"
#include "wt_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

using wt::EX;
using wt::EY;
using wt::EZ;
using wt::OPP;
using wt::Q;
using wt::pix;

__host__ __device__ constexpr int64_t off_of(int q, int nx, int ny)
{
    return EX[q] + (int64_t)nx * (EY[q] + (int64_t)ny * EZ[q]);
}

using wt_adj::collide_adj;
using wt_adj::collide_adj_fast;

// float arithmetic takes the rearranged form, double the generic one
template <class T, bool LES, bool TRT>
__device__ __forceinline__ void adj_collide(const T (&f)[Q], const T (&lam)[Q], double tau, double K, double lambda,
                                            T (&mu)[Q])
{
    if constexpr (std::is_same<T, float>::value)
        collide_adj_fast<LES, TRT>(f, lam, (float)tau, (float)K, (float)lambda, mu);
    else
        collide_adj<T, LES, TRT>(f, lam, (T)tau, (T)K, (T)lambda, mu);
}

// Stored state value (the population offset f - w) as T.
template <class T, class E> __device__ __forceinline__ T sval(const E *p, int64_t k) { return (T)p[k]; }
// A binary16 value times a float scale: fma(h, c, +0) is one mixed-precision
// fma (v_fma_mix_f32), the same value as the product but for the sign of a
// zero result.
__device__ __forceinline__ float scaled_half(__half h, float c) { return __builtin_fmaf(__half2float(h), c, 0.0f); }
template <class T> __device__ __forceinline__ T sval(const __half *p, int64_t k)
{
    return (T)scaled_half(p[k], wt::F16_INV_SCALE);
}

// Exponent of the scale of a binary16 adjoint field: 2^e * bound lies in
// [2^10, 2^11) (e = 0 for a zero bound). The same float computation runs on
// the host to check the stored maxima after the sweep.
__host__ __device__ inline int mu_exp(float bound)
{
    if (!(bound > 0.0f)) return 0;
    int e = 0;
    (void)frexpf(bound, &e);  // bound in [2^(e-1), 2^e)
    return std::min(std::max(11 - e, -120), 120);
}

__device__ __forceinline__ float pow2f(int e) { return __uint_as_float((uint32_t)(127 + e) << 23); }

// S: storage type of the forward states; M: type of the adjoint fields.
template <class S, class M> struct AdjArgs {
    const M *mu;     // mu_k
    const S *Sk;     // s_k (gradient at gray cells)
    const S *Sp;     // s_{k-1} (pre-collision gather of step k-1), or null at k = 0
    M *mun;          // mu_{k-1}
    const uint8_t *solid;
    const uint32_t *sbits, *nbits;
    const float *wgt;
    double *dJdw;
    int64_t ns;
    wt::Lay lay;
    int nx, ny, nz, nwx;
    double Fb[3];
    double tau, K, lambda;
    // Straight-line path of the cells set in `hot` (interior, not gray, no
    // solid or gray neighbour). mu is never written at boundary-plane or
    // solid cells, so it stays 0 there and the plain gather already drops
    // the links the general path skips:
    // element of slot q at the row-0 cell shifted by the streaming y/z part
    // (muo: mu at m + c_q; spo: s_{k-1} at m - c_q; sto: home), plus the row
    // term (row >> lsh) * lstride and the lane term of the x shift, as in
    // the forward step (rows are whole storage blocks).
    const uint32_t *hot;
    // cold cells (computed, not hot) in cell order; in the fused launch
    // they are taken by the blocks of the first cold_planes z planes
    const uint32_t *cold;
    int64_t ncold;
    int cold_planes;
    // (32 bits, which halves the scalar registers: checked on the host)
    int32_t muo[Q], spo[Q], sto[Q];
    int lsh, lstride, lmask;
    // the same offsets for buffer resources: byte offsets from the lowest
    // gathered element of the row (mu at muo + row + mumin, s_{k-1} at
    // spo + row + spmin) and from the row's home element (mu_{k-1}); bufok
    // when all of them fit 31 bits
    int64_t mumin, spmin;
    int32_t bmu[Q], bsp[Q], bst[Q];
    int bufok;
    // binary16 fields: mx[j] = max |mu_j| as float bits, j = 0 .. N (filled
    // by the launch that writes mu_j; the starting field mu_{N-1} is zero and
    // mx[N - 1] = mx[N] = 0), k the reverse step, src_in / src_out =
    // 2 |Fbar|_1 of steps k + 1 and k (0 for step N)
    uint32_t *mx;
    int k;
    float src_in, src_out;
};

// Multipliers of one launch: mu_k = stored * si, stored mu_{k-1} = mu * so.
struct MuScale {
    float si = 1.0f, so = 1.0f;
};

template <class S, class M> __device__ __forceinline__ MuScale mu_scale(const AdjArgs<S, M> &a)
{
    MuScale m;
    if constexpr (std::is_same<M, __half>::value) {
        m.si = pow2f(-mu_exp(__uint_as_float(a.mx[a.k + 1]) + a.src_in));
        m.so = pow2f(mu_exp(__uint_as_float(a.mx[a.k]) + a.src_out));
    }
    return m;
}

// Adjoint field value as T.
template <class T> __device__ __forceinline__ T mval(const float *p, int64_t k, float) { return (T)p[k]; }
template <class T> __device__ __forceinline__ T mval(const double *p, int64_t k, float) { return (T)p[k]; }
template <class T> __device__ __forceinline__ T mval(const __half *p, int64_t k, float si)
{
    return (T)scaled_half(p[k], si);
}
template <class M, class T> __device__ __forceinline__ M mstore(T v, float so)
{
    if constexpr (std::is_same<M, __half>::value) return __float2half_rn((float)v * so);
    else return (M)v;
}

// Raise the recorded max |mu| of a field to the wave's largest m (every lane
// of the wave takes part; blocks are whole waves).
__device__ __forceinline__ void mu_max(uint32_t *dst, float m)
{
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0 &&
        __float_as_uint(m) > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(dst, __float_as_uint(m));
}

__device__ __forceinline__ bool bitset(const uint32_t *bits, int nwx, int ny, int x, int y, int z)
{
    return (bits[(x >> 5) + (int64_t)nwx * (y + (int64_t)ny * z)] >> (x & 31)) & 1u;
}

template <class S, class M> __device__ __forceinline__ bool computed(const AdjArgs<S, M> &a, int x, int y, int z)
{
    return x >= 1 && x <= a.nx - 2 && y >= 1 && y <= a.ny - 2 && z >= 1 && z <= a.nz - 2 &&
           !bitset(a.sbits, a.nwx, a.ny, x, y, z);
}

// Cells the straight-line adjoint path may take (see AdjArgs::hot).
__global__ void __launch_bounds__(256) k_adj_hot(const uint32_t *sbits, const uint32_t *nbits, const float *wgt,
                                                 uint32_t *hot, int nx, int ny, int nz, int nwx)
{
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nwords = (int64_t)nwx * ny * nz;
    if (w >= nwords) return;
    const int wx = (int)(w % nwx);
    const int64_t row = w / nwx;
    const int y = (int)(row % ny), z = (int)(row / ny);
    uint32_t hb = 0;
    if (y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2) {
        const uint32_t cold = sbits[w] | nbits[w];
        for (int b = 0; b < 32; b++) {
            const int x = wx * 32 + b;
            if (x < 1 || x > nx - 2 || ((cold >> b) & 1u)) continue;
            if (wgt && wgt[x + (int64_t)nx * (y + (int64_t)ny * z)] > 0.0f) continue;
            hb |= 1u << b;
        }
    }
    hot[w] = hb;
}

template <class T, class E> __device__ __forceinline__ T ld_el(const E *base, int64_t uoff, int64_t rb, uint32_t lane)
{
    return sval<T>((const E *)((const char *)(base + (uoff + rb)) + lane), 0);
}
template <class T, class E>
__device__ __forceinline__ T ld_mu(const E *base, int64_t uoff, int64_t rb, uint32_t lane, float si)
{
    return mval<T>((const E *)((const char *)(base + (uoff + rb)) + lane), 0, si);
}

// Workgroups of at most ADJ_THREADS threads: the cold path holds several
// 19-value arrays at once, and a 256-thread workgroup lets it keep them in
// registers (1024 threads would cap every wave at 128 VGPRs and spill).
// Build switch: WT_ADJ_WAVES asks for that many resident waves per SIMD
// (a register budget) instead.
constexpr int ADJ_THREADS = 256;
inline int adj_block(int nx) { return std::min(wt::row_block(nx), ADJ_THREADS); }
#ifdef WT_ADJ_WAVES
#define WT_ADJ_ATTR __launch_bounds__(ADJ_THREADS) __attribute__((amdgpu_waves_per_eu(WT_ADJ_WAVES)))
#else
#define WT_ADJ_ATTR __launch_bounds__(ADJ_THREADS)
#endif

// One cold cell (near-wall or gray, or next to a gray cell): the general
// reverse step. Every load is issued up front; links whose far cell is not
// computed read mu = 0 there (mu is only ever written at computed cells) or
// are masked by a select.
template <class S, class M, class T, bool LES, bool TRT>
__device__ __forceinline__ float adj_cold(const AdjArgs<S, M> &a, const MuScale &sc, int x, int y, int z)
{
    const int64_t i = x + (int64_t)a.nx * (y + (int64_t)a.ny * z);
    const float wi = a.wgt ? a.wgt[i] : 0.0f;
    T mun[Q], muh[Q];  // mu_k at i + c_q, and at i itself
    int sj[Q];         // solid state of i + c_q
    float wj[Q];       // gray weight of i + c_q
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int64_t n1 = i + off_of(q, a.nx, a.ny);
        mun[q] = mval<T>(a.mu, pix(q, n1, a.lay), sc.si);
        muh[q] = mval<T>(a.mu, pix(q, i, a.lay), sc.si);
        sj[q] = a.solid[n1];
        wj[q] = a.wgt ? a.wgt[n1] : 0.0f;
    }
    // lambda_k(i): gray-aware streaming, plus the bounce-back links of i
    // (from j = i + c_q along qbar)
    T lam[Q];
    const T ti = T(1) - (T)wi;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const bool gl = q && !sj[q] && (wi > 0.0f || wj[q] > 0.0f);  // gray link either way
        const T t = gl ? ti * (T(1) - (T)wj[q]) : T(1);
        T v = sj[q] ? T(0) : t * mun[q];
        if (q && (sj[q] || gl)) {
            const int qb = OPP[q];
            const bool flink = sj[q] == wt::SOLID_OBJ || gl;
            const double cF = flink ? EX[qb] * a.Fb[0] + EY[qb] * a.Fb[1] + EZ[qb] * a.Fb[2] : 0.0;
            v += (sj[q] ? T(1) : T(1) - t) * (muh[qb] - T(2) * (T)cF);
        }
        lam[q] = v;
    }
    if (wi > 0.0f) {
        // dJ/dw_i over the 18 links i receives on (from i - c_q) and the 18
        // it sends on (into i + c_q)
        double sk[Q], fdn[Q], skj[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            sk[q] = sval<double>(a.Sk, pix(q, i, a.lay));
            fdn[q] = sval<double>(a.Sk, pix(OPP[q], i + off_of(q, a.nx, a.ny), a.lay));
            skj[q] = sval<double>(a.Sk, pix(q, i - off_of(q, a.nx, a.ny), a.lay));
        }
        double g = 0;
#pragma unroll
        for (int q = 1; q < Q; q++) {
            const int qb = OPP[q];
            const double cF = EX[q] * a.Fb[0] + EY[q] * a.Fb[1] + EZ[q] * a.Fb[2];
            // i receives along q from j = i - c_q = i + c_qbar
            if (!sj[qb]) g += (1.0 - (double)wj[qb]) * ((double)muh[q] * (sk[qb] - skj[q]) - 2 * sk[qb] * cF);
            // n1 = i + c_q receives along q from i
            const int X = x + EX[q], Y = y + EY[q], Z = z + EZ[q];
            const bool comp = X >= 1 && X <= a.nx - 2 && Y >= 1 && Y <= a.ny - 2 && Z >= 1 && Z <= a.nz - 2 && !sj[q];
            if (comp) g += (1.0 - (double)wj[q]) * ((double)mun[q] * (fdn[q] - sk[q]) - 2 * fdn[q] * cF);
        }
        a.dJdw[i] += g;
    }
    if (!a.Sp) return 0.0f;
    if (x == a.nx - 2) {
        // the outlet copies this cell: fold in lambda of x = nx-1
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (EX[q] != 1) lam[q] += mval<T>(a.mu, pix(q, i + 1 + off_of(q, a.nx, a.ny), a.lay), sc.si);
    }
    // pre-collision populations of step k-1 (gray-aware pull from j = i - c_q = i + c_qbar)
    T spj[Q], sph[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        spj[q] = sval<T>(a.Sp, pix(q, i - off_of(q, a.nx, a.ny), a.lay));
        sph[q] = sval<T>(a.Sp, pix(q, i, a.lay));
    }
    T f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int qb = OPP[q];
        if (q && sj[qb]) {
            f[q] = sph[qb];
        } else if (q && (wi > 0.0f || wj[qb] > 0.0f)) {
            const T t = ti * (T(1) - (T)wj[qb]);
            f[q] = t * spj[q] + (T(1) - t) * sph[qb];
        } else {
            f[q] = spj[q];
        }
    }
    T mu[Q];
    adj_collide<T, LES, TRT>(f, lam, a.tau, a.K, a.lambda, mu);
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        a.mun[pix(q, i, a.lay)] = mstore<M>(mu[q], sc.so);
        m = fmaxf(m, fabsf((float)mu[q]));
    }
    return m;
}

// Buffer-resource access (raw buffers: no range check, 32-bit offsets; word
// 3 is the gfx9-family 32-bit data format). One scalar base per array and
// row, one scalar byte offset per direction, the lane's byte offset in a
// VGPR: no 64-bit address arithmetic per access.
constexpr int ADJ_RAW_FLAGS = 0x00020000;
typedef unsigned int adj_u32x2 __attribute__((ext_vector_type(2)));
template <class E> __device__ __forceinline__ __amdgpu_buffer_rsrc_t adj_rsrc(const E *p)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, 0, -1, ADJ_RAW_FLAGS);
}
template <class E> __device__ __forceinline__ E buf_get(__amdgpu_buffer_rsrc_t r, uint32_t voff, int32_t soff)
{
    if constexpr (std::is_same<E, __half>::value) {
        return __ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0));
    } else if constexpr (std::is_same<E, float>::value) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
    } else {
        const adj_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0);
        return __hiloint2double((int)v.y, (int)v.x);
    }
}
template <class E> __device__ __forceinline__ void buf_put(E v, __amdgpu_buffer_rsrc_t r, uint32_t voff, int32_t soff)
{
    if constexpr (std::is_same<E, __half>::value) {
        __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(v), r, (int)voff, soff, 0);
    } else if constexpr (std::is_same<E, float>::value) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, soff, 0);
    } else {
        adj_u32x2 u;
        u.x = (unsigned)__double2loint(v);
        u.y = (unsigned)__double2hiint(v);
        __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)voff, soff, 0);
    }
}
// Adjoint field value as T from a stored element (binary16 fields scaled).
template <class T, class M> __device__ __forceinline__ T mget(M v, float si)
{
    if constexpr (std::is_same<M, __half>::value) return (T)scaled_half(v, si);
    else return (T)v;
}
// State value as T from a stored element, as sval (binary16: the stored
// offset times F16_INV_SCALE).
template <class T, class S> __device__ __forceinline__ T sget(S v)
{
    if constexpr (std::is_same<S, __half>::value) return (T)scaled_half(v, wt::F16_INV_SCALE);
    else return (T)v;
}
// Build switch WT_ADJ_BUF=0: 64-bit scalar bases for every access.
#ifndef WT_ADJ_BUF
#define WT_ADJ_BUF 1
#endif
// Wave priority (0-3) while a row-kernel cell with binary16 adjoint fields
// issues its gathers (1.2 % faster reverse step at 640x320x320; float
// fields keep priority 0).
#ifndef WT_ADJ_PRIO
#define WT_ADJ_PRIO 2
#endif

// One hot cell: no gray weight in reach and every neighbour computed, so
// lambda_k is the plain gather of mu_k and f_{k-1} the plain pull of s_{k-1}.
// Returns max |mu_{k-1}| of the cell.
// BUF: take the buffer-resource form when the offsets fit (the row kernel;
// the fused kernel, which also carries the cold path, keeps one form so its
// registers do not spill).
template <class S, class M, class T, bool LES, bool TRT, bool BUF>
__device__ __forceinline__ float adj_hot(const AdjArgs<S, M> &a, const MuScale &sc, int x, int y, int z)
{
    const int64_t row = (int64_t)a.nx * (y + (int64_t)a.ny * z);
    const int64_t rb = (row >> a.lsh) * a.lstride;
    auto lane = [&](int xs) { return (uint32_t)((xs >> a.lsh) * a.lstride + (xs & a.lmask)); };
    const uint32_t e[3] = {lane(x - 1), lane(x), lane(x + 1)};
    // byte offsets of the lanes in the adjoint (M) and state (S) arrays
    const uint32_t o[3] = {e[0] * (uint32_t)sizeof(M), e[1] * (uint32_t)sizeof(M), e[2] * (uint32_t)sizeof(M)};
    const uint32_t os[3] = {e[0] * (uint32_t)sizeof(S), e[1] * (uint32_t)sizeof(S), e[2] * (uint32_t)sizeof(S)};
    T lam[Q], f[Q], mu[Q];
#if WT_ADJ_BUF
    if (BUF && a.bufok) {
        const __amdgpu_buffer_rsrc_t rmu = adj_rsrc(a.mu + (rb + a.mumin)), rsp = adj_rsrc(a.Sp + (rb + a.spmin)),
                                     rst = adj_rsrc(a.mun + rb);
        constexpr int prio = std::is_same<M, __half>::value ? WT_ADJ_PRIO : 0;
        if (prio) __builtin_amdgcn_s_setprio(prio);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            lam[q] = mget<T>(buf_get<M>(rmu, o[1 + EX[q]], a.bmu[q]), sc.si);
            f[q] = sget<T>(buf_get<S>(rsp, os[1 - EX[q]], a.bsp[q]));
        }
        if (x == a.nx - 2) {
#pragma unroll
            for (int q = 0; q < Q; q++)
                if (EX[q] != 1) lam[q] += mget<T>(buf_get<M>(rmu, o[2 + EX[q]], a.bmu[q]), sc.si);
        }
        if (prio) __builtin_amdgcn_s_setprio(0);
        adj_collide<T, LES, TRT>(f, lam, a.tau, a.K, a.lambda, mu);
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            buf_put<M>(mstore<M>(mu[q], sc.so), rst, o[1], a.bst[q]);
            m = fmaxf(m, fabsf((float)mu[q]));
        }
        return m;
    }
#endif
#pragma unroll
    for (int q = 0; q < Q; q++) {
        lam[q] = ld_mu<T>(a.mu, a.muo[q], rb, o[1 + EX[q]], sc.si);
        f[q] = ld_el<T>(a.Sp, a.spo[q], rb, os[1 - EX[q]]);
    }
    if (x == a.nx - 2) {
        // the outlet copies this cell: fold in lambda of x = nx-1, whose
        // links leaving the lattice (e_x = +1) carry nothing
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (EX[q] != 1) lam[q] += ld_mu<T>(a.mu, a.muo[q], rb, o[2 + EX[q]], sc.si);
    }
    adj_collide<T, LES, TRT>(f, lam, a.tau, a.K, a.lambda, mu);
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        *(M *)((char *)(a.mun + (a.sto[q] + rb)) + o[1]) = mstore<M>(mu[q], sc.so);
        m = fmaxf(m, fabsf((float)mu[q]));
    }
    return m;
}

// One reverse step is two launches: the cold cells (one listed cell per
// thread; few, with long dependent load chains and many live values) and
// the hot cells (one block per lattice row, as the forward step, on the
// straight path). Two kernels let the hot one keep its own register budget,
// so it runs at several times the cold one's occupancy.
template <class S, class M, class T, bool LES, bool TRT>
__global__ void __launch_bounds__(ADJ_THREADS) k_adjoint_cold(AdjArgs<S, M> a)
{
    const MuScale sc = mu_scale(a);
    float m = 0.0f;
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < a.ncold) {
        const int64_t i = a.cold[k];
        const int x = (int)(i % a.nx), y = (int)((i / a.nx) % a.ny), z = (int)(i / ((int64_t)a.nx * a.ny));
        m = adj_cold<S, M, T, LES, TRT>(a, sc, x, y, z);
    }
    if constexpr (std::is_same<M, __half>::value)
        if (a.Sp) mu_max(a.mx + (a.k - 1), m);
}

template <class S, class M, class T, bool LES, bool TRT>
__global__ void WT_ADJ_ATTR k_adjoint(AdjArgs<S, M> a)
{
    const MuScale sc = mu_scale(a);
    float m = 0.0f;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = (int)blockIdx.y, z = (int)blockIdx.z;
    if (x < a.nx && bitset(a.hot, a.nwx, a.ny, x, y, z)) m = adj_hot<S, M, T, LES, TRT, true>(a, sc, x, y, z);
    if constexpr (std::is_same<M, __half>::value) mu_max(a.mx + (a.k - 1), m);
}

// The fused form of one reverse step: one launch whose first cold_planes z
// planes of blocks take the listed cold cells (dispatched first, so their
// long dependent load chains overlap the rows) and whose other blocks take
// one lattice row each, on the straight path. Every wave then carries the
// cold path's register budget.
template <class S, class M, class T, bool LES, bool TRT>
__global__ void WT_ADJ_ATTR k_adjoint_fused(AdjArgs<S, M> a)
{
    const MuScale sc = mu_scale(a);
    float m = 0.0f;
    if ((int)blockIdx.z < a.cold_planes) {
        const int64_t nb = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
        const int64_t k = nb * blockDim.x + threadIdx.x;
        if (k < a.ncold) {
            const int64_t i = a.cold[k];
            const int x = (int)(i % a.nx), y = (int)((i / a.nx) % a.ny), z = (int)(i / ((int64_t)a.nx * a.ny));
            m = adj_cold<S, M, T, LES, TRT>(a, sc, x, y, z);
        }
    } else {
        // row blocks exist only when there is a step k - 1 (a.Sp set)
        const int x = blockIdx.x * blockDim.x + threadIdx.x;
        const int y = (int)blockIdx.y, z = (int)blockIdx.z - a.cold_planes;
        if (x < a.nx && bitset(a.hot, a.nwx, a.ny, x, y, z)) m = adj_hot<S, M, T, LES, TRT, false>(a, sc, x, y, z);
    }
    if constexpr (std::is_same<M, __half>::value)
        if (a.Sp) mu_max(a.mx + (a.k - 1), m);
}

// Computed cells that are not hot: interior fluid cells next to a solid or
// gray cell, and gray cells (any order; each cell is independent).
__global__ void __launch_bounds__(256) k_adj_cold_list(const uint32_t *sbits, const uint32_t *hot, uint32_t *cold,
                                                       unsigned int *count, int nx, int ny, int nz, int nwx)
{
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nwords = (int64_t)nwx * ny * nz;
    if (w >= nwords) return;
    const int wx = (int)(w % nwx);
    const int64_t row = w / nwx;
    const int y = (int)(row % ny), z = (int)(row / ny);
    if (y < 1 || y > ny - 2 || z < 1 || z > nz - 2) return;
    const uint32_t skip = sbits[w] | hot[w];
    for (int b = 0; b < 32; b++) {
        const int x = wx * 32 + b;
        if (x < 1 || x > nx - 2 || ((skip >> b) & 1u)) continue;
        cold[atomicAdd(count, 1u)] = (uint32_t)(x + (int64_t)nx * (y + (int64_t)ny * z));
    }
}


template <class S, class M, class T, bool LES, bool TRT> void launch_adj_lt(wt_lattice *h, const AdjArgs<S, M> &a)
{
    if (a.cold_planes >= 0) {
        // fused: cold planes first, then (with a step k - 1) the rows
        const int bx = adj_block(h->nx), gx = (h->nx + bx - 1) / bx;
        const int gz = a.cold_planes + (a.Sp ? h->nz : 0);
        if (gz > 0)
            hipLaunchKernelGGL((k_adjoint_fused<S, M, T, LES, TRT>), dim3(gx, h->ny, gz), dim3(bx), 0, h->stream, a);
        return;
    }
    if (a.ncold > 0)
        hipLaunchKernelGGL((k_adjoint_cold<S, M, T, LES, TRT>), dim3((unsigned)((a.ncold + ADJ_THREADS - 1) / ADJ_THREADS)),
                           dim3(ADJ_THREADS), 0, h->stream, a);
    // at k = 0 only the gradient is left (no mu_{-1}): the hot cells carry none
    if (!a.Sp) return;
    const int bx = adj_block(h->nx), gx = (h->nx + bx - 1) / bx;
    hipLaunchKernelGGL((k_adjoint<S, M, T, LES, TRT>), dim3(gx, h->ny, h->nz), dim3(bx), 0, h->stream, a);
}

template <class S, class M, class T>
int launch_adj(wt_lattice *h, const AdjArgs<S, M> &a)
{
    const bool les = h->p.les, trt = h->p.trt;
    if (les && trt) launch_adj_lt<S, M, T, true, true>(h, a);
    else if (les) launch_adj_lt<S, M, T, true, false>(h, a);
    else if (trt) launch_adj_lt<S, M, T, false, true>(h, a);
    else launch_adj_lt<S, M, T, false, false>(h, a);
    WT_HIP(hipGetLastError());
    return 0;
}

struct Buf {
    void *p = nullptr;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
};

// A recorded forward pass: sqrt(N) checkpoints of the post-collision state
// (AB layout) and the forward forces, kept on the lattice until released
// or replaced, so the reverse sweep can run any number of times.
struct AdjRecord {
    int N = 0, K = 0, nseg = 0;
    int64_t step0 = 0;         // lattice step count when the forward started
    size_t st = 0;             // bytes of one stored state
    int storage = 0, arith = 0;
    uint64_t epoch = 0;        // lattice configuration epoch at record time
    int64_t id = 0;            // wt_adjoint_record
    void *ck = nullptr;        // nseg states: ck + seg*st = s_{seg*K - 1} (s_0 for seg 0)
    std::vector<double> forces;
    ~AdjRecord() { if (ck) (void)hipFree(ck); }
};

int record_forward(wt_lattice *h, int N, int K, double *forces)
{
    auto *R = new AdjRecord();
    R->N = N; R->K = K; R->nseg = (N + K - 1) / K;
    R->st = 19 * (size_t)h->ns * h->pop_elem;
    R->storage = h->storage; R->arith = h->arith; R->epoch = h->epoch;
    R->step0 = h->steps;
    R->forces.assign((size_t)N * 3, 0.0);
    if (hipMalloc(&R->ck, R->st * (size_t)R->nseg) != hipSuccess) {
        R->ck = nullptr;
        delete R;
        return wt_fail(WT_E_NOMEM, "hipMalloc of %d adjoint checkpoints (%zu bytes each) failed", (N + K - 1) / K,
                       19 * (size_t)h->ns * h->pop_elem);
    }
    auto ck = [&](int seg) { return (char *)R->ck + (size_t)seg * R->st; };
    for (int k = 0; k < N; k++) {
        const int seg = (k + 1) / K;
        if (k == 0) WT_HIP(hipMemcpyAsync(ck(0), h->d_buf[h->parity], R->st, hipMemcpyDeviceToDevice, h->stream));
        if ((k + 1) % K == 0 && seg >= 1 && seg < R->nseg)
            WT_HIP(hipMemcpyAsync(ck(seg), h->d_buf[h->parity], R->st, hipMemcpyDeviceToDevice, h->stream));
        if (int rc = lbm_launch_step(h, 0, h->nz, k == N - 1, h->d_forces + 3 * (k % h->forces_cap))) {
            delete R;
            return rc;
        }
        h->parity ^= 1;
        h->steps++;
        if (k % h->forces_cap == h->forces_cap - 1 || k == N - 1) {
            if (int rc = lbm_flush_force(h)) {
                delete R;
                return rc;
            }
            const int base = k - (k % h->forces_cap);
            WT_HIP(hipMemcpyAsync(R->forces.data() + 3 * base, h->d_forces, sizeof(double) * 3 * (k - base + 1),
                                  hipMemcpyDeviceToHost, h->stream));
            WT_HIP(hipStreamSynchronize(h->stream));
        }
    }
    if (forces) std::copy(R->forces.begin(), R->forces.end(), forces);
    for (int k = 0; k < 3; k++) h->force[k] = R->forces[3 * (N - 1) + k];
    adj_release(h);
    R->id = ++h->adj_seq;
    h->adj = R;
    return 0;
}

// Steps per replay window: the whole segment (K) when its K + 1 states fit
// in free device memory next to a reserve, else as many as fit. The
// environment variable WT_ADJ_WINDOW caps it (tests of the windowed sweep).
int replay_window(int K, size_t st, int *W)
{
    size_t freeb = 0, total = 0;
    WT_HIP(hipMemGetInfo(&freeb, &total));
    const size_t reserve = std::max<size_t>((size_t)1 << 30, total / 64);
    const size_t fit = freeb > reserve ? (freeb - reserve) / st : 0;
    int w = (int)std::min<size_t>((size_t)K, fit > 0 ? fit - 1 : 0);
    if (const char *e = std::getenv("WT_ADJ_WINDOW")) {
        const int cap = std::atoi(e);
        if (cap >= 1) w = std::min(w, cap);
    }
    if (w < 1)
        return wt_fail(WT_E_NOMEM, "no room for the adjoint replay store: %zu bytes free, %zu per state", freeb, st);
    *W = w;
    return 0;
}

template <class S, class M, class T>
int reverse_t(wt_lattice *h, const AdjRecord &R, const double *Fbar, double *dJdw, double *ms)
{
    constexpr bool H = std::is_same<M, __half>::value;
    const int N = R.N, K = R.K, nseg = R.nseg;
    const size_t st = R.st;                                  // one stored state
    const size_t sm = 19 * (size_t)h->ns * sizeof(M);        // one adjoint field
    Buf mu0, mu1, dg, mx;
    WT_HIP(hipMalloc(&mu0.p, sm));
    WT_HIP(hipMalloc(&mu1.p, sm));
    WT_HIP(hipMalloc(&dg.p, sizeof(double) * (size_t)h->n));
    // source bound 2 |Fbar_k|_1 of each step (0 for k = N)
    std::vector<float> src((size_t)N + 1, 0.0f);
    for (int k = 0; k < N; k++)
        src[(size_t)k] = (float)(2.0 * (std::fabs(Fbar[3 * k]) + std::fabs(Fbar[3 * k + 1]) + std::fabs(Fbar[3 * k + 2])));
    if (H) {
        WT_HIP(hipMalloc(&mx.p, sizeof(uint32_t) * ((size_t)N + 1)));
        WT_HIP(hipMemsetAsync(mx.p, 0, sizeof(uint32_t) * ((size_t)N + 1), h->stream));
    }
    const int64_t nwords = (int64_t)h->nwx * h->ny * h->nz;
    Buf hot;
    WT_HIP(hipMalloc(&hot.p, sizeof(uint32_t) * (size_t)nwords));
    hipLaunchKernelGGL(k_adj_hot, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, h->stream, h->d_sbits,
                       h->d_nbits, h->d_wgt, (uint32_t *)hot.p, h->nx, h->ny, h->nz, h->nwx);
    WT_HIP(hipGetLastError());
    Buf cold, ccount;
    WT_HIP(hipMalloc(&cold.p, sizeof(uint32_t) * (size_t)h->n));
    WT_HIP(hipMalloc(&ccount.p, sizeof(unsigned int)));
    WT_HIP(hipMemsetAsync(ccount.p, 0, sizeof(unsigned int), h->stream));
    hipLaunchKernelGGL(k_adj_cold_list, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, h->stream, h->d_sbits,
                       (const uint32_t *)hot.p, (uint32_t *)cold.p, (unsigned int *)ccount.p, h->nx, h->ny, h->nz,
                       h->nwx);
    WT_HIP(hipGetLastError());
    unsigned int ncold = 0;
    WT_HIP(hipMemcpyAsync(&ncold, ccount.p, sizeof ncold, hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    // cold cells in cell order (the list kernel appends them in any order)
    std::vector<uint32_t> cl(ncold);
    if (ncold) {
        WT_HIP(hipMemcpy(cl.data(), cold.p, sizeof(uint32_t) * ncold, hipMemcpyDeviceToHost));
        std::sort(cl.begin(), cl.end());
        WT_HIP(hipMemcpy(cold.p, cl.data(), sizeof(uint32_t) * ncold, hipMemcpyHostToDevice));
    }
    int W = 0;
    if (int rc = replay_window(K, st, &W)) return rc;
    std::vector<Buf> rep((size_t)W + 1);
    for (auto &b : rep) WT_HIP(hipMalloc(&b.p, st));
    std::vector<void *> rp((size_t)W + 1);
    for (int r = 0; r <= W; r++) rp[(size_t)r] = rep[(size_t)r].p;
    WT_HIP(hipMemsetAsync(mu0.p, 0, sm, h->stream));
    WT_HIP(hipMemsetAsync(mu1.p, 0, sm, h->stream));
    WT_HIP(hipMemsetAsync(dg.p, 0, sizeof(double) * (size_t)h->n, h->stream));
    // Overlap: when two replay stores of a whole segment fit, segment
    // seg - 1 is replayed on a second stream while the adjoint steps of
    // segment seg run on the lattice's stream (the two touch different
    // buffers), so the small launches of each fill the other's gaps and
    // tails. WT_ADJ_OVERLAP=0 keeps one stream; the gradient is the same bit
    // for bit either way.
    bool overlap = W >= K && nseg > 1;
    if (const char *e = std::getenv("WT_ADJ_OVERLAP")) overlap = overlap && std::atoi(e) != 0;
    if (overlap) {
        size_t freeb = 0, total = 0;
        WT_HIP(hipMemGetInfo(&freeb, &total));
        const size_t reserve = std::max<size_t>((size_t)1 << 30, total / 64);
        overlap = freeb > reserve && (freeb - reserve) / st >= (size_t)K + 1;
    }
    std::vector<Buf> rep2(overlap ? (size_t)K + 1 : 0);
    // the side stream and its events live until the sweep has drained
    struct Side {
        hipStream_t s = nullptr;
        std::vector<hipEvent_t> ev;
        ~Side()
        {
            if (s) (void)hipStreamSynchronize(s);
            for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
            if (s) (void)hipStreamDestroy(s);
        }
    } side;
    if (overlap) {
        // allocated before the timed region starts, like the first store
        for (auto &b : rep2) WT_HIP(hipMalloc(&b.p, st));
        WT_HIP(hipStreamCreate(&side.s));
        side.ev.assign(2 * (size_t)nseg, nullptr);
        for (hipEvent_t &e : side.ev) WT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }

    WT_HIP(hipEventRecord(h->ev0, h->stream));
    AdjArgs<S, M> a{};
    a.solid = h->d_solid; a.sbits = h->d_sbits; a.nbits = h->d_nbits; a.wgt = h->d_wgt; a.dJdw = (double *)dg.p;
    a.ns = h->ns; a.lay = wt::Lay{h->ns, h->bsh}; a.nx = h->nx; a.ny = h->ny; a.nz = h->nz; a.nwx = h->nwx;
    a.tau = h->p.tau; a.K = (18 * std::sqrt(2.0)) * (h->p.csm * h->p.csm); a.lambda = h->p.lambda;
    a.hot = (const uint32_t *)hot.p;
    a.cold = (const uint32_t *)cold.p;
    a.ncold = ncold;
    // One launch per reverse step (cold cells as the first planes of the
    // row grid) or two (cold cells, then the rows with a register budget of
    // their own). Measured on MI355X: one launch is faster with float
    // fields, two with binary16 fields. WT_ADJ_FUSE=0/1 picks one.
    bool fuse = !H;
    if (const char *e = std::getenv("WT_ADJ_FUSE")) fuse = std::atoi(e) != 0;
    a.cold_planes = -1;
    if (fuse) {
        const int bx = adj_block(h->nx), gx = (h->nx + bx - 1) / bx;
        const int64_t per_plane = (int64_t)gx * h->ny * bx;
        a.cold_planes = (int)((ncold + per_plane - 1) / per_plane);
    }
    a.mx = (uint32_t *)mx.p;
    a.lsh = h->bsh;
    a.lstride = h->bsh ? Q << h->bsh : 1;
    a.lmask = h->bsh ? (1 << h->bsh) - 1 : 0;
    for (int q = 0; q < Q; q++) {
        // slot q of the row-0 cell shifted by dyz cells (a multiple of nx, so of the block)
        const int64_t dyz = (int64_t)h->nx * (EY[q] + (int64_t)h->ny * EZ[q]);
        const int64_t slot = h->bsh ? (int64_t)wt::SLOT[q] << h->bsh : wt::SLOT[q] * h->ns;
        const int64_t mo = slot + (dyz >> h->bsh) * a.lstride, so = slot + ((-dyz) >> h->bsh) * a.lstride;
        if (std::max({std::llabs(mo), std::llabs(so), std::llabs(slot)}) > INT32_MAX)
            return wt_fail(WT_E_ARG, "lattice too large for the adjoint's 32-bit offsets");
        a.muo[q] = (int32_t)mo;
        a.spo[q] = (int32_t)so;
        a.sto[q] = (int32_t)slot;
    }
    a.mumin = *std::min_element(a.muo, a.muo + Q);
    a.spmin = *std::min_element(a.spo, a.spo + Q);
    a.bufok = 1;
    for (int q = 0; q < Q; q++) {
        const int64_t bm = (a.muo[q] - a.mumin) * (int64_t)sizeof(M), bs = (a.spo[q] - a.spmin) * (int64_t)sizeof(S),
                      bt = a.sto[q] * (int64_t)sizeof(M);
        const bool fit = bm <= INT32_MAX && bs <= INT32_MAX && bt >= 0 && bt <= INT32_MAX;
        a.bufok = a.bufok && fit;
        a.bmu[q] = fit ? (int32_t)bm : 0;
        a.bsp[q] = fit ? (int32_t)bs : 0;
        a.bst[q] = fit ? (int32_t)bt : 0;
    }
    void *mu_k = mu0.p, *mu_n = mu1.p;
    // the replays below drive the forward kernel on explicit buffers; the
    // lattice's own state and counters are left untouched
    const int parity = h->parity;
    // forward step k (s_k -> s_{k+1}) again, as recorded
    auto replay = [&](void *from, void *to, int k) {
        h->ovF = from;
        h->ovG = to;
        h->replay = 1;
        h->seed_step = R.step0 + k;
        const int rc = lbm_launch_step(h, 0, h->nz, false, nullptr);
        h->ovF = h->ovG = nullptr;
        h->replay = 0;
        h->seed_step = -1;
        return rc;
    };
    // one reverse step of the sweep: step k from the replayed states
    // store[k - s0] (s_k) and store[k - 1 - s0] (s_{k-1})
    auto adjoint_step = [&](const std::vector<void *> &store, int s0, int k) {
        a.mu = (const M *)mu_k;
        a.mun = (M *)mu_n;
        a.Sk = (const S *)store[(size_t)(k - s0)];
        a.Sp = k > 0 ? (const S *)store[(size_t)(k - 1 - s0)] : nullptr;
        for (int c = 0; c < 3; c++) a.Fb[c] = Fbar[3 * k + c];
        a.k = k;
        a.src_in = src[(size_t)k + 1];
        a.src_out = src[(size_t)k];
        const int rc = launch_adj<S, M, T>(h, a);
        std::swap(mu_k, mu_n);
        return rc;
    };
    if (overlap) {
        std::vector<void *> stores[2] = {rp, std::vector<void *>((size_t)K + 1)};
        for (int r = 0; r <= K; r++) stores[1][(size_t)r] = rep2[(size_t)r].p;
        auto evB = [&](int seg) { return side.ev[(size_t)seg]; };            // replays of seg done
        auto evA = [&](int seg) { return side.ev[(size_t)(nseg + seg)]; };   // adjoint of seg done
        WT_HIP(hipStreamWaitEvent(side.s, h->ev0, 0));
        // replay segment seg into a store on the side stream: s_{s0} .. s_{k1-1}
        auto prep = [&](int seg) {
            const int k0 = seg * K, k1 = std::min(N, k0 + K), s0 = std::max(k0 - 1, 0);
            const std::vector<void *> &store = stores[seg & 1];
            WT_HIP(hipMemcpyAsync(store[0], (char *)R.ck + (size_t)seg * st, st, hipMemcpyDeviceToDevice, side.s));
            hipStream_t own = h->stream;
            h->stream = side.s;
            for (int r = 0; s0 + r < k1 - 1; r++) {
                if (int rc = replay(store[(size_t)r], store[(size_t)r + 1], s0 + r)) {
                    h->stream = own;
                    return rc;
                }
            }
            h->stream = own;
            WT_HIP(hipEventRecord(evB(seg), side.s));
            return 0;
        };
        if (int rc = prep(nseg - 1)) return rc;
        for (int seg = nseg - 1; seg >= 0; seg--) {
            if (seg >= 1) {
                // the store of segment seg - 1 last served segment seg + 1
                if (seg + 1 < nseg) WT_HIP(hipStreamWaitEvent(side.s, evA(seg + 1), 0));
                if (int rc = prep(seg - 1)) return rc;
            }
            WT_HIP(hipStreamWaitEvent(h->stream, evB(seg), 0));
            const int k0 = seg * K, k1 = std::min(N, k0 + K), s0 = std::max(k0 - 1, 0);
            for (int k = k1 - 1; k >= k0; k--)
                if (int rc = adjoint_step(stores[seg & 1], s0, k)) return rc;
            WT_HIP(hipEventRecord(evA(seg), h->stream));
        }
        // every side-stream launch precedes evB(0), which the last segment waited for
    }
    for (int seg = overlap ? -1 : nseg - 1; seg >= 0; seg--) {
        // checkpoint of the segment: s_base; its reverse steps are [k0, k1)
        const int base = seg == 0 ? 0 : seg * K - 1;
        const int k0 = seg * K, k1 = std::min(N, k0 + K);
        for (int b = k1; b > k0;) {
            // window of reverse steps [a0, b): replay store rp[r] = s_{s0 + r}
            // for r = 0 .. b-1-s0 (at most W + 1 states)
            const int a0 = std::max(k0, b - W);
            const int s0 = std::max(a0 - 1, 0);
            WT_HIP(hipMemcpyAsync(rp[0], (char *)R.ck + (size_t)seg * st, st, hipMemcpyDeviceToDevice, h->stream));
            for (int c = base; c < s0; c++) {
                if (int rc = replay(rp[0], rp[1], c)) return rc;
                std::swap(rp[0], rp[1]);
            }
            for (int r = 0; s0 + r < b - 1; r++)
                if (int rc = replay(rp[(size_t)r], rp[(size_t)r + 1], s0 + r)) return rc;
            for (int k = b - 1; k >= a0; k--)
                if (int rc = adjoint_step(rp, s0, k)) return rc;
            b = a0;
        }
    }
    h->parity = parity;
    WT_HIP(hipEventRecord(h->ev1, h->stream));
    WT_HIP(hipMemcpyAsync(dJdw, dg.p, sizeof(double) * (size_t)h->n, hipMemcpyDeviceToHost, h->stream));
    std::vector<uint32_t> mxh(H ? (size_t)N + 1 : 0);
    if (H) WT_HIP(hipMemcpyAsync(mxh.data(), mx.p, sizeof(uint32_t) * ((size_t)N + 1), hipMemcpyDeviceToHost, h->stream));
    WT_HIP(hipStreamSynchronize(h->stream));
    float t = 0;
    WT_HIP(hipEventElapsedTime(&t, h->ev0, h->ev1));
    if (ms) *ms = t;
    if (H) {
        // every stored field mu_j (j = N-2 .. 0) within binary16's range:
        // values from 65520 on round to infinity
        auto fl = [](uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; };
        for (int j = N - 2; j >= 0; j--) {
            const float mj = fl(mxh[(size_t)j]);
            const int e = mu_exp(fl(mxh[(size_t)j + 1]) + src[(size_t)j + 1]);
            if (!(std::ldexp((double)mj, e) < 65520.0))
                return wt_fail(WT_E_STATE,
                               "binary16 adjoint field of step %d out of range (max |mu| %g, scale 2^%d): "
                               "set WT_ADJ_FIELD=f32 for float fields",
                               j, (double)mj, e);
        }
    }
    return 0;
}

int check_adjoint_lattice(const wt_lattice *h)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    if (h->layout != WT_LAYOUT_AB) return wt_fail(WT_E_ARG, "the adjoint needs the AB layout");
    if (h->p.movingWalls) return wt_fail(WT_E_ARG, "the adjoint does not support moving walls");
    if (h->bc != WT_BC_REFERENCE) return wt_fail(WT_E_ARG, "the adjoint takes the reference boundaries only");
    return 0;
}

}  // namespace

void adj_release(wt_lattice *h)
{
    delete (AdjRecord *)h->adj;
    h->adj = nullptr;
}

extern "C" {

int wt_set_gray(wt_lattice *h, const float *w)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    h->epoch++;
    if (!w) {
        if (h->d_wgt) (void)hipFree(h->d_wgt);
        h->d_wgt = nullptr;
        if (int rc = lbm_rebuild_masks(h)) return rc;
        WT_HIP(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (h->layout != WT_LAYOUT_AB) return wt_fail(WT_E_ARG, "gray bounce-back needs the AB layout");
    std::vector<float> v((size_t)h->n);
    for (int64_t i = 0; i < h->n; i++) {
        float x = w[i];
        const int cx = (int)(i % h->nx), cy = (int)((i / h->nx) % h->ny), cz = (int)(i / ((int64_t)h->nx * h->ny));
        const bool bnd = cx == 0 || cy == 0 || cz == 0 || cx == h->nx - 1 || cy == h->ny - 1 || cz == h->nz - 1;
        if (h->h_solid[i] || bnd || !(x > 0)) x = 0;  // boundary planes carry no weight
        if (x >= 1) return wt_fail(WT_E_ARG, "gray weight >= 1 on fluid cell %lld: voxelize it as body", (long long)i);
        v[i] = x;
    }
    if (!h->d_wgt && hipMalloc((void **)&h->d_wgt, sizeof(float) * (size_t)h->n) != hipSuccess)
        return wt_fail(WT_E_NOMEM, "hipMalloc of the weight field failed");
    WT_HIP(hipMemcpyAsync(h->d_wgt, v.data(), sizeof(float) * (size_t)h->n, hipMemcpyHostToDevice, h->stream));
    if (int rc = lbm_rebuild_masks(h)) return rc;
    WT_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int wt_adjoint_forward(wt_lattice *h, int nsteps, int interval, double *forces)
{
    if (int rc = check_adjoint_lattice(h)) return rc;
    if (nsteps < 1) return wt_fail(WT_E_ARG, "nsteps must be positive, got %d", nsteps);
    int K = interval > 0 ? interval : (int)std::ceil(std::sqrt((double)nsteps));
    if (interval <= 0) {
        // the checkpoints take at most half the free device memory; the
        // reverse sweep sizes its replay store to what is left
        size_t freeb = 0, total = 0;
        WT_HIP(hipMemGetInfo(&freeb, &total));
        const size_t st = 19 * (size_t)h->ns * h->pop_elem;
        const int64_t fit = (int64_t)(freeb / 2 / st);
        if (fit < 1) return wt_fail(WT_E_NOMEM, "no room for one adjoint checkpoint (%zu bytes free)", freeb);
        K = std::max<int64_t>(K, (nsteps + fit - 1) / fit);
    }
    K = std::max(1, std::min(K, nsteps));
    return record_forward(h, nsteps, K, forces);
}

int wt_adjoint_backward(wt_lattice *h, int nsteps, const double *Fbar, double *dJdw, double *reverse_ms)
{
    if (int rc = check_adjoint_lattice(h)) return rc;
    if (!Fbar || !dJdw) return wt_fail(WT_E_ARG, "null adjoint buffer");
    const auto *R = (const AdjRecord *)h->adj;
    if (!R) return wt_fail(WT_E_STATE, "no recorded forward pass: call wt_adjoint_forward first");
    if (nsteps != R->N)
        return wt_fail(WT_E_ARG, "Fbar covers %d steps, the recorded forward pass %d", nsteps, R->N);
    if (R->epoch != h->epoch || R->storage != h->storage || R->arith != h->arith)
        return wt_fail(WT_E_STATE, "the lattice configuration changed since the recorded forward pass");
    if (h->storage == WT_STORE_F64) return reverse_t<double, double, double>(h, *R, Fbar, dJdw, reverse_ms);
    if (h->storage == WT_STORE_F16) {
        const char *f = std::getenv("WT_ADJ_FIELD");
        if (f && std::strcmp(f, "f32") == 0) return reverse_t<__half, float, float>(h, *R, Fbar, dJdw, reverse_ms);
        return reverse_t<__half, __half, float>(h, *R, Fbar, dJdw, reverse_ms);
    }
    if (h->arith == WT_ARITH_F64) return reverse_t<float, float, double>(h, *R, Fbar, dJdw, reverse_ms);
    return reverse_t<float, float, float>(h, *R, Fbar, dJdw, reverse_ms);
}

int wt_adjoint_record(const wt_lattice *h, int64_t *id)
{
    if (!h || !id) return wt_fail(WT_E_ARG, "null argument");
    const auto *R = (const AdjRecord *)h->adj;
    if (!R) return wt_fail(WT_E_STATE, "no recorded forward pass");
    *id = R->id;
    return 0;
}

int wt_adjoint_release(wt_lattice *h)
{
    if (!h) return wt_fail(WT_E_ARG, "null lattice handle");
    adj_release(h);
    return 0;
}

int wt_adjoint(wt_lattice *h, int nsteps, const double *Fbar, double *dJdw, double *forces, int interval,
               double *reverse_ms)
{
    if (int rc = check_adjoint_lattice(h)) return rc;
    if (nsteps < 1 || !Fbar || !dJdw) return wt_fail(WT_E_ARG, "bad adjoint request");
    if (int rc = wt_adjoint_forward(h, nsteps, interval, forces)) return rc;
    const int rc = wt_adjoint_backward(h, nsteps, Fbar, dJdw, reverse_ms);
    adj_release(h);
    return rc;
}

}  // extern "C"
