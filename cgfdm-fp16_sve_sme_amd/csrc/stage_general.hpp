// stage_general.hpp — one RK stage at one grid point, every boundary branch included.
//
// Per point this reproduces the reference's per-stage composite, in the reference order:
//   wave_deriv (wave_deriv.cpp:58-203)            -> h
//   free_surface_deriv on the top 3 layers        -> h overwritten (freeSurface.cpp:22-258)
//   pml_deriv x1,y1,z1,x2,y2[,z2]                 -> h += -beta*Aux*DT, hAux (pml_deriv.cpp:90-544)
//   pml_free_surface x1,y1,x2,y2 at the surface   -> hAux stress += ... (pml_freeSurface.cpp:12-235)
//   wave_rk / wave_pml_rk                         -> W, Aux (wave_rk.cpp:13-99, pml_rk.cpp:14-70)
// The stage increment h stays in FP32 registers instead of a 9-plane h_W array, and the RK
// update writes the next stage's field into a separate buffer (ping-pong), so the derivative
// and the update fuse into one pass without a read-after-write hazard.
#pragma once
#include "common.hpp"

namespace cg {

// One-sided derivative of variable v along direction d at (i, j, k):
// FB*(a_-1 f[-FB] + a0 f[0] + a1 f[FB] + a2 f[2FB] + a3 f[3FB])*rDH  (axis.h:129-133)
template <class S>
__device__ __forceinline__ float Lraw(const S* __restrict__ W, const Layout& L, int v, int i, int j, int k, int d,
                                      int FB, float rDH) {
    const int di = d == 0 ? FB : 0, dj = d == 1 ? FB : 0, dk = d == 2 ? FB : 0;
    auto f = [&](int m) { return SV<S>::ld(W[L.at(9, v, i + m * di, j + m * dj, k + m * dk)]); };
    return (float)FB * (af_1 * f(-1) + af0 * f(0) + af1 * f(1) + af2 * f(2) + af3 * f(3)) * rDH;
}

__device__ __forceinline__ float LJT(const float* J_T, int FB, float rDH) {
    return (float)FB *
           (af_1 * J_T[3 - FB] + af0 * J_T[3] + af1 * J_T[3 + FB] + af2 * J_T[3 + 2 * FB] + af3 * J_T[3 + 3 * FB]) * rDH;
}

// RK register update of the 9 wavefield variables at (i, j, k) (wave_rk0..3). The accumulator
// after stage 0 is not stored; stage 1 rebuilds it from its input (kT0).
template <class S>
__device__ __forceinline__ void rk_update(const StageParams& p, int i, int j, int k, const float* h) {
    S* __restrict__ Wo = (S*)p.Wout;
    const S* __restrict__ A = (const S*)p.A;
    const S* __restrict__ Wi = (const S*)p.Win;
    S* __restrict__ T = (S*)p.T;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        const long long q = p.lay.at(9, v, i, j, k);
        if (p.irk == 0) {
            float m = SV<S>::ld(A[q]);
            Wo[q] = SV<S>::st(m + alpha2 * h[v]);
        } else if (p.irk == 1) {
            float m = SV<S>::ld(A[q]), w = SV<S>::ld(Wi[q]);
            Wo[q] = SV<S>::st(m + alpha3 * h[v]);
            T[q] = SV<S>::st((m + (w - m) * kT0) + beta2 * h[v]);
        } else if (p.irk == 2) {
            float m = SV<S>::ld(A[q]), t = SV<S>::ld(T[q]);
            Wo[q] = SV<S>::st(m + h[v]);
            T[q] = SV<S>::st(t + beta3 * h[v]);
        } else {
            float t = SV<S>::ld(T[q]);
            Wo[q] = SV<S>::st(t + beta4 * h[v]);
        }
    }
}

template <class S>
__device__ void stage_point_general(const StageParams& p, int i, int j, int k) {
    const S* __restrict__ W = (const S*)p.Win;
    const S* __restrict__ Cj = (const S*)p.cjm;
    const Layout& L = p.lay;
    auto w = [&](int v, int ii, int jj, int kk) { return SV<S>::ld(W[L.at(9, v, ii, jj, kk)]); };
    auto cj = [&](int c, int ii, int jj, int kk) { return SV<S>::ld(Cj[L.atc(c, ii, jj, kk)]); };
    const long long rdzi = (i + (long long)j * p.nx_) * 9;
    const float rDH = p.rDH, DT = p.DT;
    const int _nz = p.nz + HALO;
    const int FBs[3] = {p.FB1, p.FB2, p.FB3};

    float M[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) M[c] = cj(c, i, j, k);
    const float mu = cj(10, i, j, k);
    const float lambda = cj(11, i, j, k);
    const float buoyancy = cj(12, i, j, k) * p.Crho;

    // raw one-sided derivatives D[dir][var] (no PML factor)
    float D[3][9];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < 9; ++v) D[d][v] = Lraw<S>(W, L, v, i, j, k, d, FBs[d], rDH);

    const float bx = p.pb[0][i], by = p.pb[1][j], bz = p.pb[2][k];
    float h[9];
    const bool fsl = p.fs && k >= _nz - HALO;
    if (!fsl) {
        // wave_deriv.cpp:132-201
        float B[3][9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            B[0][v] = D[0][v] * bx;
            B[1][v] = D[1][v] * by;
            B[2][v] = D[2][v] * bz;
        }
        float Vx[3], Vy[3], Vz[3], Txx[3], Tyy[3], Tzz[3], Txy[3], Txz[3], Tyz[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float mx = M[3 * d], my = M[3 * d + 1], mz = M[3 * d + 2];
            const float* b = B[d];
            Vx[d] = (mx * b[3] + my * b[6] + mz * b[7]) * buoyancy;
            Vy[d] = (mx * b[6] + my * b[4] + mz * b[8]) * buoyancy;
            Vz[d] = (mx * b[7] + my * b[8] + mz * b[5]) * buoyancy;
            const float dv = (mx * b[0] + my * b[1] + mz * b[2]) * lambda;
            Txx[d] = dv + 2.0f * mu * (mx * b[0]);
            Tyy[d] = dv + 2.0f * mu * (my * b[1]);
            Tzz[d] = dv + 2.0f * mu * (mz * b[2]);
            Txy[d] = (my * b[0] + mx * b[1]) * mu;
            Txz[d] = (mz * b[0] + mx * b[2]) * mu;
            Tyz[d] = (mz * b[1] + my * b[2]) * mu;
        }
        h[0] = (Vx[0] + Vx[1] + Vx[2]) * DT;
        h[1] = (Vy[0] + Vy[1] + Vy[2]) * DT;
        h[2] = (Vz[0] + Vz[1] + Vz[2]) * DT;
        h[3] = (Txx[0] + Txx[1] + Txx[2]) * DT;
        h[4] = (Tyy[0] + Tyy[1] + Tyy[2]) * DT;
        h[5] = (Tzz[0] + Tzz[1] + Tzz[2]) * DT;
        h[6] = (Txy[0] + Txy[1] + Txy[2]) * DT;
        h[7] = (Txz[0] + Txz[1] + Txz[2]) * DT;
        h[8] = (Tyz[0] + Tyz[1] + Tyz[2]) * DT;
    } else {
        // freeSurface.cpp:104-256: conservative J*(metric . T) form for the velocity with
        // antisymmetric imaging of the stress about the surface
        const int k_s = (_nz - 1) - k + HALO;
        float JT[3][3][7];  // [dir][component][l]
#pragma unroll
        for (int l = 0; l <= 2 * HALO; ++l) {
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int ii = i + (d == 0 ? l - HALO : 0), jj = j + (d == 1 ? l - HALO : 0);
                const float m0 = cj(3 * d, ii, jj, k);
                const float m1 = cj(3 * d + 1, ii, jj, k);
                const float m2 = cj(3 * d + 2, ii, jj, k);
                const float jac = cj(9, ii, jj, k);
                const float w3 = w(3, ii, jj, k), w4 = w(4, ii, jj, k);
                const float w5 = w(5, ii, jj, k), w6 = w(6, ii, jj, k);
                const float w7 = w(7, ii, jj, k), w8 = w(8, ii, jj, k);
                JT[d][0][l] = (m0 * w3 + m1 * w6 + m2 * w7) * jac;
                JT[d][1][l] = (m0 * w6 + m1 * w4 + m2 * w8) * jac;
                JT[d][2][l] = (m0 * w7 + m1 * w8 + m2 * w5) * jac;
            }
        }
        for (int l = 0; l < k_s; ++l) {
            const int kk = k + l - HALO;
            const float m0 = cj(6, i, j, kk);
            const float m1 = cj(7, i, j, kk);
            const float m2 = cj(8, i, j, kk);
            const float jac = cj(9, i, j, kk);
            const float w3 = w(3, i, j, kk), w4 = w(4, i, j, kk);
            const float w5 = w(5, i, j, kk), w6 = w(6, i, j, kk);
            const float w7 = w(7, i, j, kk), w8 = w(8, i, j, kk);
            JT[2][0][l] = (m0 * w3 + m1 * w6 + m2 * w7) * jac;
            JT[2][1][l] = (m0 * w6 + m1 * w4 + m2 * w8) * jac;
            JT[2][2][l] = (m0 * w7 + m1 * w8 + m2 * w5) * jac;
        }
        for (int c = 0; c < 3; ++c) {
            JT[2][c][k_s] = 0.f;
            for (int l = k_s + 1; l <= 2 * HALO; ++l) JT[2][c][l] = -JT[2][c][2 * k_s - l];
        }
        const float jacb = cj(9, i, j, k);
        const float Jinv = 1.0f / jacb;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            h[c] = buoyancy * Jinv *
                   (LJT(JT[0][c], p.FB1, rDH) * bx + LJT(JT[1][c], p.FB2, rDH) * by + LJT(JT[2][c], p.FB3, rDH)) * DT;

        const float Vx_xi = D[0][0] * bx, Vx_et = D[1][0] * by;
        const float Vy_xi = D[0][1] * bx, Vy_et = D[1][1] * by;
        const float Vz_xi = D[0][2] * bx, Vz_et = D[1][2] * by;
        float Vx_zt, Vy_zt, Vz_zt;
        if (k == _nz - 1) {
            const float* X = p.rdz + rdzi;
            const float* Y = p.rdz + (long long)p.nx_ * p.ny_ * 9 + rdzi;
            Vx_zt = (X[0] * Vx_xi + X[1] * Vy_xi + X[2] * Vz_xi) + (Y[0] * Vx_et + Y[1] * Vy_et + Y[2] * Vz_et);
            Vy_zt = (X[3] * Vx_xi + X[4] * Vy_xi + X[5] * Vz_xi) + (Y[3] * Vx_et + Y[4] * Vy_et + Y[5] * Vz_et);
            Vz_zt = (X[6] * Vx_xi + X[7] * Vy_xi + X[8] * Vz_xi) + (Y[6] * Vx_et + Y[7] * Vy_et + Y[8] * Vz_et);
        } else if (k == _nz - 2) {  // L2 (axis.h:92)
            const int k1 = k + p.FB3;
            Vx_zt = (float)p.FB3 * (w(0, i, j, k1) - w(0, i, j, k)) * rDH;
            Vy_zt = (float)p.FB3 * (w(1, i, j, k1) - w(1, i, j, k)) * rDH;
            Vz_zt = (float)p.FB3 * (w(2, i, j, k1) - w(2, i, j, k)) * rDH;
        } else {  // L3 (axis.h:93)
            const int k1 = k + p.FB3, k2 = k + 2 * p.FB3;
            float r3[3];
#pragma unroll
            for (int v = 0; v < 3; ++v)
                r3[v] = (float)p.FB3 * (Cf1 * w(v, i, j, k) + Cf2 * w(v, i, j, k1) + Cf3 * w(v, i, j, k2)) * rDH;
            Vx_zt = r3[0]; Vy_zt = r3[1]; Vz_zt = r3[2];
        }
        const float Vd[3][3] = {{Vx_xi, Vy_xi, Vz_xi}, {Vx_et, Vy_et, Vz_et}, {Vx_zt, Vy_zt, Vz_zt}};
        float Txx = 0.f, Tyy = 0.f, Tzz = 0.f, Txy = 0.f, Txz = 0.f, Tyz = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float mx = M[3 * d], my = M[3 * d + 1], mz = M[3 * d + 2];
            const float* b = Vd[d];
            const float dv = (mx * b[0] + my * b[1] + mz * b[2]) * lambda;
            Txx += dv + 2.0f * mu * (mx * b[0]);
            Tyy += dv + 2.0f * mu * (my * b[1]);
            Tzz += dv + 2.0f * mu * (mz * b[2]);
            Txy += (my * b[0] + mx * b[1]) * mu;
            Txz += (mz * b[0] + mx * b[2]) * mu;
            Tyz += (mz * b[1] + my * b[2]) * mu;
        }
        h[3] = Txx * DT; h[4] = Tyy * DT; h[5] = Tzz * DT;
        h[6] = Txy * DT; h[7] = Txz * DT; h[8] = Tyz * DT;
    }

    if (p.pml) {
        const int ijk[3] = {i, j, k};
#pragma unroll 1
        for (int s = 0; s < 6; ++s) {
            const Slab& X = p.slab[s];
            if (!X.on) continue;
            const int a = X.axis;
            const int c = ijk[a];
            if (c < X.lo || c >= X.lo + p.nPML) continue;
            const int i0 = (a == 0 ? c - X.lo : i - HALO);
            const int j0 = (a == 1 ? c - X.lo : j - HALO);
            const int k0 = (a == 2 ? c - X.lo : k - HALO);
            const long long pos = i0 + (long long)j0 * X.n0 + (long long)k0 * X.n0 * X.n1;
            const float beta = p.pb[a][c], dd = p.pd[a][c], alpha_d = dd + p.pa[a][c];
            // the registers hold Aux*DT in the storage type
            S* __restrict__ Xa = (S*)X.a;
            S* __restrict__ Xm = (S*)X.m;
            S* __restrict__ Xt = (S*)X.t;
            float aold[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) aold[v] = SV<S>::ld(Xa[v * X.n + pos]);
#pragma unroll
            for (int v = 0; v < 9; ++v) h[v] += -beta * aold[v];
            // single-direction right-hand side with d (pml_deriv.cpp:176-207)
            const float mx = M[3 * a], my = M[3 * a + 1], mz = M[3 * a + 2];
            float b[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) b[v] = D[a][v] * dd;
            float r[9];
            r[0] = (mx * b[3] + my * b[6] + mz * b[7]) * buoyancy;
            r[1] = (mx * b[6] + my * b[4] + mz * b[8]) * buoyancy;
            r[2] = (mx * b[7] + my * b[8] + mz * b[5]) * buoyancy;
            const float dv = (mx * b[0] + my * b[1] + mz * b[2]) * lambda;
            r[3] = dv + 2.0f * mu * (mx * b[0]);
            r[4] = dv + 2.0f * mu * (my * b[1]);
            r[5] = dv + 2.0f * mu * (mz * b[2]);
            r[6] = (my * b[0] + mx * b[1]) * mu;
            r[7] = (mz * b[0] + mx * b[2]) * mu;
            r[8] = (mz * b[1] + my * b[2]) * mu;
            // right-hand side of Aux*DT, times DT below: DT * (rhs*DT - (alpha+d)*Aux*DT)
            float ha[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) ha[v] = r[v] * DT - alpha_d * aold[v];
            if (p.fs && a < 2 && k == _nz - 1) {
                // pml_free_surface_x/y (pml_freeSurface.cpp:81-116)
                const float* Mz = p.rdz + (a == 0 ? 0 : (long long)p.nx_ * p.ny_ * 9) + rdzi;
                const float zx = M[6], zy = M[7], zz = M[8];
                const float Vz0 = Mz[0] * b[0] + Mz[1] * b[1] + Mz[2] * b[2];
                const float Vz1 = Mz[3] * b[0] + Mz[4] * b[1] + Mz[5] * b[2];
                const float Vz2 = Mz[6] * b[0] + Mz[7] * b[1] + Mz[8] * b[2];
                const float dz = (zx * Vz0 + zy * Vz1 + zz * Vz2) * lambda;
                ha[3] += (dz + 2.0f * mu * (zx * Vz0)) * DT;
                ha[4] += (dz + 2.0f * mu * (zy * Vz1)) * DT;
                ha[5] += (dz + 2.0f * mu * (zz * Vz2)) * DT;
                ha[6] += ((zy * Vz0 + zx * Vz1) * mu) * DT;
                ha[7] += ((zz * Vz0 + zx * Vz2) * mu) * DT;
                ha[8] += ((zz * Vz1 + zy * Vz2) * mu) * DT;
            }
#pragma unroll
            for (int v = 0; v < 9; ++v) ha[v] *= DT;
            // wave_pml_rk0..3 (pml_rk.cpp:14-70), in place: each point owns its aux values; the
            // accumulator after stage 0 is rebuilt in stage 1 (kT0)
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                const long long q = v * X.n + pos;
                if (p.irk == 0) {
                    const float m = aold[v];
                    Xm[q] = SV<S>::st(m);
                    Xa[q] = SV<S>::st(m + alpha2 * ha[v]);
                } else if (p.irk == 1) {
                    const float m = SV<S>::ld(Xm[q]);
                    Xt[q] = SV<S>::st((m + (aold[v] - m) * kT0) + beta2 * ha[v]);
                    Xa[q] = SV<S>::st(m + alpha3 * ha[v]);
                } else if (p.irk == 2) {
                    Xt[q] = SV<S>::st(SV<S>::ld(Xt[q]) + beta3 * ha[v]);
                    Xa[q] = SV<S>::st(SV<S>::ld(Xm[q]) + ha[v]);
                } else {
                    Xa[q] = SV<S>::st(SV<S>::ld(Xt[q]) + beta4 * ha[v]);
                }
            }
        }
    }
    rk_update<S>(p, i, j, k, h);
}

// General kernel over a list of boxes (flattened with prefix sums).
template <class S>
__global__ void __launch_bounds__(256) k_stage_general(StageParams p) {
    const long long t = xcd_order(blockIdx.x, (long long)gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= p.boxStart[p.nbox]) return;
    int b = 0;
    while (b + 1 < p.nbox && t >= p.boxStart[b + 1]) ++b;
    const long long r = t - p.boxStart[b];
    const int* B = p.box[b];
    const int nxb = B[1] - B[0], nyb = B[3] - B[2];
    const int i = B[0] + (int)(r % nxb);
    const int j = B[2] + (int)((r / nxb) % nyb);
    const int k = B[4] + (int)(r / ((long long)nxb * nyb));
    stage_point_general<S>(p, i, j, k);
}

}  // namespace cg
