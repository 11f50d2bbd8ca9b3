// Transposed Jacobian of the D3Q19 collision (js/lbm.js:377-460): the
// adjoint of one cell's moments -> deviation equilibrium -> LES -> TRT/BGK
// relaxation, mu = (d collide / d f)^T lam. Host and device (the host build
// lets a CPU harness compare the two forms).
#pragma once

#include "wt_internal.h"

#include <cmath>

namespace wt_adj {

using wt::EX;
using wt::EY;
using wt::EZ;
using wt::Q;

template <class T>
__host__ __device__ __forceinline__ T wq(int q) { return q == 0 ? T(1.0 / 3) : (q <= 6 ? T(1.0 / 18) : T(1.0 / 36)); }

// mu = (d collide / d f)^T lam for the collision of js/lbm.js:377-460 in its
// generic form e_q = W_q (A + 3 r cu + 4.5 r cu^2); LES and TRT included.
template <class T, bool LES, bool TRT>
__host__ __device__ __forceinline__ void collide_adj(const T (&f)[Q], const T (&lam)[Q], T tau0, T K, T lamb, T (&mu)[Q])
{
    T dr = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) { dr += f[q]; jx += (T)EX[q] * f[q]; jy += (T)EY[q] * f[q]; jz += (T)EZ[q] * f[q]; }
    const T r = T(1) + dr, inv = T(1) / r;
    const T vx = jx * inv, vy = jy * inv, vz = jz * inv;
    const T usq = vx * vx + vy * vy + vz * vz;
    const T A = dr - T(1.5) * r * usq;
    T cu[Q], nn[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        cu[q] = (T)EX[q] * vx + (T)EY[q] * vy + (T)EZ[q] * vz;
        nn[q] = f[q] - wq<T>(q) * (A + T(3) * r * cu[q] + T(4.5) * r * cu[q] * cu[q]);
    }
    T te = tau0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0, mag = 0, Sq = 0;
    if (LES) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            pxx += (T)(EX[q] * EX[q]) * nn[q]; pyy += (T)(EY[q] * EY[q]) * nn[q]; pzz += (T)(EZ[q] * EZ[q]) * nn[q];
            pxy += (T)(EX[q] * EY[q]) * nn[q]; pxz += (T)(EX[q] * EZ[q]) * nn[q]; pyz += (T)(EY[q] * EZ[q]) * nn[q];
        }
        mag = sqrt(T(2) * (pxx * pxx + pyy * pyy + pzz * pzz + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz)));
        Sq = sqrt(tau0 * tau0 + K * mag * inv);
        te = T(0.5) * (tau0 + Sq);
    }
    const T om = T(1) / te;
    const T omm = TRT ? T(1) / (T(0.5) + lamb / (te - T(0.5))) : om;

    T nb[Q], omb = 0, ommb = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) mu[q] = lam[q];
    nb[0] = -om * lam[0];
    omb -= nn[0] * lam[0];
#pragma unroll
    for (int a = 1; a < Q; a += 2) {
        const int b = a + 1;
        const T sp = T(0.5) * (nn[a] + nn[b]), sm = T(0.5) * (nn[a] - nn[b]);
        const T spb = -om * (lam[a] + lam[b]), smb = -omm * (lam[a] - lam[b]);
        nb[a] = T(0.5) * (spb + smb);
        nb[b] = T(0.5) * (spb - smb);
        omb -= sp * (lam[a] + lam[b]);
        ommb -= sm * (lam[a] - lam[b]);
    }
    T teb = 0;
    if (TRT) teb += ommb * omm * omm * lamb / ((te - T(0.5)) * (te - T(0.5)));
    else omb += ommb;
    T invb = 0;
    if (LES) {
        teb += -omb / (te * te);
        const T magb = teb * K * inv / (T(4) * Sq);
        invb += teb * K * mag / (T(4) * Sq);
        const T Xb = mag > T(0) ? magb / mag : T(0);
        const T pxxb = T(2) * Xb * pxx, pyyb = T(2) * Xb * pyy, pzzb = T(2) * Xb * pzz;
        const T pxyb = T(4) * Xb * pxy, pxzb = T(4) * Xb * pxz, pyzb = T(4) * Xb * pyz;
#pragma unroll
        for (int q = 0; q < Q; q++)
            nb[q] += pxxb * (T)(EX[q] * EX[q]) + pyyb * (T)(EY[q] * EY[q]) + pzzb * (T)(EZ[q] * EZ[q]) +
                     pxyb * (T)(EX[q] * EY[q]) + pxzb * (T)(EX[q] * EZ[q]) + pyzb * (T)(EY[q] * EZ[q]);
    }
    T Ab = 0, rb = 0, vxb = 0, vyb = 0, vzb = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        mu[q] += nb[q];
        const T eb = -nb[q] * wq<T>(q);
        Ab += eb;
        rb += eb * (T(3) * cu[q] + T(4.5) * cu[q] * cu[q]);
        const T cub = eb * (T(3) * r + T(9) * r * cu[q]);
        vxb += (T)EX[q] * cub; vyb += (T)EY[q] * cub; vzb += (T)EZ[q] * cub;
    }
    T drb = Ab;
    rb += -T(1.5) * usq * Ab;
    const T usqb = -T(1.5) * r * Ab;
    vxb += T(2) * vx * usqb; vyb += T(2) * vy * usqb; vzb += T(2) * vz * usqb;
    const T jxb = vxb * inv, jyb = vyb * inv, jzb = vzb * inv;
    invb += vxb * jx + vyb * jy + vzb * jz;
    rb += -inv * inv * invb;
    drb += rb;
#pragma unroll
    for (int q = 0; q < Q; q++) mu[q] += drb + (T)EX[q] * jxb + (T)EY[q] * jyb + (T)EZ[q] * jzb;
}


// Hardware single-precision reciprocal and square root (about 1 ulp) on the
// device; plain IEEE operations in a host build.
__host__ __device__ __forceinline__ float fast_rcp(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ float fast_sqrt(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}

// The same transposed Jacobian in float, rearranged for instruction count
// like the forward collision: opposite pairs (a, a+1) enter through sums and
// differences, and every per-pair quantity of the generic form is built
// from them:
//   S_p = nb_a + nb_b = -om (lam_a + lam_b) + 2 P_p (P_p: LES term of the
//   pair's direction), D_p = nb_a - nb_b = -omm (lam_a - lam_b),
//   sum over the pair of e-gradients = -w_p S_p, their difference -w_p D_p.
// Pair directions: x, y, z, x+y, x-y, x+z, x-z, y+z, y-z (a = 1, 3, ..., 17).
// Products and sums contract into fmas here (unlike the binary64 path,
// which must round like the reference): the reverse step is bound by its
// vector instruction count.
template <bool LES, bool TRT>
__host__ __device__ __forceinline__ void collide_adj_fast(const float (&f)[Q], const float (&lam)[Q], float tau0,
                                                          float K, float lamb, float (&mu)[Q])
{
#pragma clang fp contract(fast)
    float s[9], d[9];
#pragma unroll
    for (int p = 0; p < 9; p++) { s[p] = f[2 * p + 1] + f[2 * p + 2]; d[p] = f[2 * p + 1] - f[2 * p + 2]; }
    const float dr = f[0] + (((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + s[8]);
    const float jx = d[0] + d[3] + d[4] + d[5] + d[6];
    const float jy = d[1] + d[3] - d[4] + d[7] + d[8];
    const float jz = d[2] + d[5] - d[6] + d[7] - d[8];
    const float r = 1.0f + dr, inv = fast_rcp(r);
    const float vx = jx * inv, vy = jy * inv, vz = jz * inv;
    const float usq = __builtin_fmaf(vx, vx, __builtin_fmaf(vy, vy, vz * vz));
    const float A = __builtin_fmaf(-1.5f * r, usq, dr);
    const float cu[9] = {vx, vy, vz, vx + vy, vx - vy, vx + vz, vx - vz, vy + vz, vy - vz};
    const float WF = 1.0f / 18, WE = 1.0f / 36;
    const float r45 = 4.5f * r;
    float sp[9], sm[9];  // (nn_a + nn_b)/2 and (nn_a - nn_b)/2
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const float w = p < 3 ? WF : WE;
        sp[p] = __builtin_fmaf(0.5f, s[p], -w * __builtin_fmaf(r45 * cu[p], cu[p], A));
        sm[p] = __builtin_fmaf(0.5f, d[p], -3.0f * w * r * cu[p]);
    }
    const float n0 = __builtin_fmaf(-1.0f / 3, A, f[0]);
    float te = tau0, mag = 0, Sq = 0, Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
    if (LES) {
        // full non-equilibrium stress: twice the half sums
        Pxx = 2.0f * (sp[0] + sp[3] + sp[4] + sp[5] + sp[6]);
        Pyy = 2.0f * (sp[1] + sp[3] + sp[4] + sp[7] + sp[8]);
        Pzz = 2.0f * (sp[2] + sp[5] + sp[6] + sp[7] + sp[8]);
        Pxy = 2.0f * (sp[3] - sp[4]);
        Pxz = 2.0f * (sp[5] - sp[6]);
        Pyz = 2.0f * (sp[7] - sp[8]);
        const float q2 = __builtin_fmaf(Pxx, Pxx, __builtin_fmaf(Pyy, Pyy, Pzz * Pzz)) +
                         2.0f * __builtin_fmaf(Pxy, Pxy, __builtin_fmaf(Pxz, Pxz, Pyz * Pyz));
        mag = fast_sqrt(2.0f * q2);
        Sq = fast_sqrt(__builtin_fmaf(tau0, tau0, K * mag * inv));
        te = 0.5f * (tau0 + Sq);
    }
    const float om = fast_rcp(te);
    const float tm = fast_rcp(te - 0.5f);
    const float omm = TRT ? fast_rcp(__builtin_fmaf(lamb, tm, 0.5f)) : om;

    // backward through the relaxation
    float S[9], D[9];
    float omb = -n0 * lam[0], ommb = 0;
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const float Ls = lam[2 * p + 1] + lam[2 * p + 2], Ld = lam[2 * p + 1] - lam[2 * p + 2];
        S[p] = -om * Ls;
        D[p] = -omm * Ld;
        omb = __builtin_fmaf(-sp[p], Ls, omb);
        ommb = __builtin_fmaf(-sm[p], Ld, ommb);
    }
    const float nb0 = -om * lam[0];
    float teb = 0, invb = 0;
    if (TRT) teb = ommb * omm * omm * lamb * tm * tm;
    else omb += ommb;
    if (LES) {
        teb = __builtin_fmaf(-omb, om * om, teb);
        const float k4 = teb * K * fast_rcp(4.0f * Sq);
        const float magb = k4 * inv;
        invb += k4 * mag;
        const float Xb = mag > 0.0f ? magb * fast_rcp(mag) : 0.0f;
        const float bxx = 2.0f * Xb * Pxx, byy = 2.0f * Xb * Pyy, bzz = 2.0f * Xb * Pzz;
        const float bxy = 4.0f * Xb * Pxy, bxz = 4.0f * Xb * Pxz, byz = 4.0f * Xb * Pyz;
        const float P[9] = {bxx, byy, bzz, bxx + byy + bxy, bxx + byy - bxy, bxx + bzz + bxz, bxx + bzz - bxz,
                            byy + bzz + byz, byy + bzz - byz};
#pragma unroll
        for (int p = 0; p < 9; p++) S[p] = __builtin_fmaf(2.0f, P[p], S[p]);
    }
    // backward through the equilibrium
    float Ab = -nb0 * (1.0f / 3), rb = 0, cub[9];
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const float w = p < 3 ? WF : WE;
        const float es = -w * S[p], ed = -w * D[p];   // pair sum / difference of the e-gradients
        Ab += es;
        rb = __builtin_fmaf(3.0f * cu[p], ed, __builtin_fmaf(4.5f * cu[p] * cu[p], es, rb));
        cub[p] = __builtin_fmaf(3.0f * r, ed, 9.0f * r * cu[p] * es);
    }
    float vxb = cub[0] + cub[3] + cub[4] + cub[5] + cub[6];
    float vyb = cub[1] + cub[3] - cub[4] + cub[7] + cub[8];
    float vzb = cub[2] + cub[5] - cub[6] + cub[7] - cub[8];
    float drb = Ab;
    rb = __builtin_fmaf(-1.5f * usq, Ab, rb);
    const float usqb = -1.5f * r * Ab;
    vxb = __builtin_fmaf(2.0f * vx, usqb, vxb);
    vyb = __builtin_fmaf(2.0f * vy, usqb, vyb);
    vzb = __builtin_fmaf(2.0f * vz, usqb, vzb);
    const float jxb = vxb * inv, jyb = vyb * inv, jzb = vzb * inv;
    invb = __builtin_fmaf(vxb, jx, __builtin_fmaf(vyb, jy, __builtin_fmaf(vzb, jz, invb)));
    rb = __builtin_fmaf(-inv * inv, invb, rb);
    drb += rb;
    const float cj[9] = {jxb, jyb, jzb, jxb + jyb, jxb - jyb, jxb + jzb, jxb - jzb, jyb + jzb, jyb - jzb};
    mu[0] = lam[0] + nb0 + drb;
#pragma unroll
    for (int p = 0; p < 9; p++) {
        // mu_a = lam_a + (S + D)/2 + drb + cj, mu_b = lam_b + (S - D)/2 + drb - cj
        const float h = __builtin_fmaf(0.5f, S[p], drb), e = __builtin_fmaf(0.5f, D[p], cj[p]);
        mu[2 * p + 1] = lam[2 * p + 1] + (h + e);
        mu[2 * p + 2] = lam[2 * p + 2] + (h - e);
    }
}

}  // namespace wt_adj
