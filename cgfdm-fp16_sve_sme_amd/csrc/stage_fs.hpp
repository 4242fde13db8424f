// stage_fs.hpp — one RK stage on the three free-surface layers above the bulk columns (no PML
// there), as a tile kernel: the same arithmetic as the free-surface branch of the general
// point kernel (freeSurface.cpp:104-256, wave_rk.cpp:13-99), with the in-plane neighbours read
// from LDS instead of from memory once per neighbour and variable.
//
// A workgroup owns a TX x TY column tile and does the three layers one after the other. For
// each layer every tile cell and halo cell writes nine values to LDS: the three velocities and
// the two horizontal traction fluxes J*(xi_. . T) and J*(eta_. . T) (three components each);
// the horizontal one-sided derivatives of the velocities and of the fluxes are then LDS
// stencils. The vertical flux of a point's own column (planes k-3 .. k+3, imaged
// antisymmetrically about the surface) is computed once per column from the five planes below
// the surface.
#pragma once
#include <type_traits>

#include "stage_general.hpp"

namespace cg {

constexpr int kFsTx = 64, kFsTy = 4;  // free-surface tile shape

template <int TX, int TY> struct FsLds {
    static constexpr int LX = TX + 4, LY = TY + 4, VS = LX * LY;
    float v[9 * VS];
};

// The nine LDS values of cell (i, j, k): V_x, V_y, V_z, then J*(m_d . T) row by row for the
// xi (d = 0) and eta (d = 1) metric rows, as freeSurface.cpp:138-181 forms them.
template <class S>
__device__ __forceinline__ void fs_cell_values(const StageParams& p, int i, int j, int k, float* o) {
    const S* __restrict__ W = (const S*)p.Win;
    const S* __restrict__ Cj = (const S*)p.cjm;
    const Layout& L = p.lay;
    float w[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) w[v] = SV<S>::ld(W[L.at(9, v, i, j, k)]);
    const float jac = SV<S>::ld(Cj[L.atc(9, i, j, k)]);
    o[0] = w[0];
    o[1] = w[1];
    o[2] = w[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const float m0 = SV<S>::ld(Cj[L.atc(3 * d, i, j, k)]);
        const float m1 = SV<S>::ld(Cj[L.atc(3 * d + 1, i, j, k)]);
        const float m2 = SV<S>::ld(Cj[L.atc(3 * d + 2, i, j, k)]);
        o[3 + 3 * d] = (m0 * w[3] + m1 * w[6] + m2 * w[7]) * jac;
        o[4 + 3 * d] = (m0 * w[6] + m1 * w[4] + m2 * w[8]) * jac;
        o[5 + 3 * d] = (m0 * w[7] + m1 * w[8] + m2 * w[5]) * jac;
    }
}

// Tile (bx, by) of the columns [bi0,bi1) x [bj0,bj1), planes fk0 .. fk1-1 (the three layers
// below and at the surface, fk1 - 1 = HALO + nz - 1). Every thread of the workgroup calls it;
// it starts by writing LDS, so a caller running several tiles in a row separates them by a
// barrier.
template <class S, int TX, int TY>
__device__ __forceinline__ void fs_tile(const StageParams& p, int bx, int by, FsLds<TX, TY>& lds) {
    constexpr int LX = FsLds<TX, TY>::LX, VS = FsLds<TX, TY>::VS, NT = TX * TY;
    const S* __restrict__ W = (const S*)p.Win;
    const S* __restrict__ Cj = (const S*)p.cjm;
    const Layout& L = p.lay;

    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int i0 = p.bi0 + bx * TX, j0 = p.bj0 + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool inside = i < p.bi1 && j < p.bj1;
    const int FB1 = p.FB1, FB2 = p.FB2, FB3 = p.FB3;
    const int hx = FB1 > 0 ? 1 : 3, hy = FB2 > 0 ? 1 : 3;
    const int kt = p.fk1 - 1;  // the surface plane
    const float rDH = p.rDH, DT = p.DT;

    // halo cells of this thread (same enumeration as the bulk tile: no corner cells)
    constexpr int NHX = 4 * TY, NHY = 4 * TX, NH = NHX + NHY;
    constexpr int HPT = (NH + NT - 1) / NT;
    int hi_[HPT], hj_[HPT], hl[HPT];
#pragma unroll
    for (int r = 0; r < HPT; ++r) {
        const int c = tid + r * NT;
        int hr = 0, hc = 0;
        if (c < NHX) {
            const int rr = c / 4, c4 = c % 4;
            hr = hy + rr;
            hc = c4 < hx ? c4 : TX + c4;
        } else if (c < NH) {
            const int tt = c - NHX;
            const int r4 = tt / TX;
            hc = hx + tt % TX;
            hr = r4 < hy ? r4 : TY + r4;
        }
        hi_[r] = i0 - hx + hc;
        hj_[r] = j0 - hy + hr;
        hl[r] = c < NH ? hr * LX + hc : -1;
    }
    const int lc = (ty + hy) * LX + tx + hx;

    // vertical flux of the own column on the five planes below the surface, kt-5 .. kt-1
    float Z[5][3];
    float M[9] = {}, mu = 0.f, lambda = 0.f, buoyancy = 0.f, bxv = 1.f, byv = 1.f;
    if (inside) {
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            const int kk = kt - 5 + s;
            const float m0 = SV<S>::ld(Cj[L.atc(6, i, j, kk)]);
            const float m1 = SV<S>::ld(Cj[L.atc(7, i, j, kk)]);
            const float m2 = SV<S>::ld(Cj[L.atc(8, i, j, kk)]);
            const float jac = SV<S>::ld(Cj[L.atc(9, i, j, kk)]);
            const float w3 = SV<S>::ld(W[L.at(9, 3, i, j, kk)]), w4 = SV<S>::ld(W[L.at(9, 4, i, j, kk)]);
            const float w5 = SV<S>::ld(W[L.at(9, 5, i, j, kk)]), w6 = SV<S>::ld(W[L.at(9, 6, i, j, kk)]);
            const float w7 = SV<S>::ld(W[L.at(9, 7, i, j, kk)]), w8 = SV<S>::ld(W[L.at(9, 8, i, j, kk)]);
            Z[s][0] = (m0 * w3 + m1 * w6 + m2 * w7) * jac;
            Z[s][1] = (m0 * w6 + m1 * w4 + m2 * w8) * jac;
            Z[s][2] = (m0 * w7 + m1 * w8 + m2 * w5) * jac;
        }
        bxv = p.pb[0][i];
        byv = p.pb[1][j];
    } else {
#pragma unroll
        for (int s = 0; s < 5; ++s) Z[s][0] = Z[s][1] = Z[s][2] = 0.f;
    }

    // layer s = 0, 1, 2 is plane kt-2+s (a template constant, so every window index is static)
    auto layer = [&](auto Sc) {
        constexpr int s = decltype(Sc)::value;
        const int k = kt - 2 + s;
        constexpr int ks = 5 - s;  // index of the surface in the window k-3 .. k+3
        // LDS: own cell and halo cells of plane k
        if constexpr (s > 0) __syncthreads();
        {
            float o[9];
            if (i < L.nx_ && j < L.ny_) fs_cell_values<S>(p, i, j, k, o);
            else
#pragma unroll
                for (int v = 0; v < 9; ++v) o[v] = 0.f;
#pragma unroll
            for (int v = 0; v < 9; ++v) lds.v[v * VS + lc] = o[v];
        }
#pragma unroll
        for (int r = 0; r < HPT; ++r) {
            if (hl[r] < 0) continue;
            float o[9];
            if (hi_[r] >= 0 && hi_[r] < L.nx_ && hj_[r] >= 0 && hj_[r] < L.ny_) fs_cell_values<S>(p, hi_[r], hj_[r], k, o);
            else
#pragma unroll
                for (int v = 0; v < 9; ++v) o[v] = 0.f;
#pragma unroll
            for (int v = 0; v < 9; ++v) lds.v[v * VS + hl[r]] = o[v];
        }
        __syncthreads();
        if (!inside) return;

#pragma unroll
        for (int c = 0; c < 9; ++c) M[c] = SV<S>::ld(Cj[L.atc(c, i, j, k)]);
        mu = SV<S>::ld(Cj[L.atc(10, i, j, k)]);
        lambda = SV<S>::ld(Cj[L.atc(11, i, j, k)]);
        buoyancy = SV<S>::ld(Cj[L.atc(12, i, j, k)]) * p.Crho;
        const float jacb = SV<S>::ld(Cj[L.atc(9, i, j, k)]);

        auto dx = [&](int v) {
            const float* Lv = lds.v + v * VS;
            return (float)FB1 *
                   (af_1 * Lv[lc - FB1] + af0 * Lv[lc] + af1 * Lv[lc + FB1] + af2 * Lv[lc + 2 * FB1] +
                    af3 * Lv[lc + 3 * FB1]) *
                   rDH;
        };
        auto dy = [&](int v) {
            const float* Lv = lds.v + v * VS;
            return (float)FB2 *
                   (af_1 * Lv[lc - FB2 * LX] + af0 * Lv[lc] + af1 * Lv[lc + FB2 * LX] + af2 * Lv[lc + 2 * FB2 * LX] +
                    af3 * Lv[lc + 3 * FB2 * LX]) *
                   rDH;
        };

        // velocity: h = b/J * (d/dxi J xi.T + d/deta J eta.T + d/dzeta J zeta.T) * DT, the zeta
        // flux imaged antisymmetrically about the surface (freeSurface.cpp:183-203)
        float h[9];
        const float Jinv = 1.0f / jacb;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // window entry l is plane k-3+l = kt-5+(l+s): column plane l+s below the surface,
            // zero at the surface, the negated mirror image above it
            float J_T[7];
#pragma unroll
            for (int l = 0; l < 7; ++l) {
                if (l < ks) J_T[l] = Z[l + s][c];
                else if (l == ks) J_T[l] = 0.f;
                else J_T[l] = -J_T[2 * ks - l];
            }
            const float lz = LJT(J_T, FB3, rDH);
            h[c] = buoyancy * Jinv * (dx(3 + c) * bxv + dy(6 + c) * byv + lz) * DT;
        }

        const float Vx_xi = dx(0) * bxv, Vx_et = dy(0) * byv;
        const float Vy_xi = dx(1) * bxv, Vy_et = dy(1) * byv;
        const float Vz_xi = dx(2) * bxv, Vz_et = dy(2) * byv;
        float Vx_zt, Vy_zt, Vz_zt;
        if constexpr (s == 2) {
            const long long rdzi = (i + (long long)j * p.nx_) * 9;
            const float* X = p.rdz + rdzi;
            const float* Y = p.rdz + (long long)p.nx_ * p.ny_ * 9 + rdzi;
            Vx_zt = (X[0] * Vx_xi + X[1] * Vy_xi + X[2] * Vz_xi) + (Y[0] * Vx_et + Y[1] * Vy_et + Y[2] * Vz_et);
            Vy_zt = (X[3] * Vx_xi + X[4] * Vy_xi + X[5] * Vz_xi) + (Y[3] * Vx_et + Y[4] * Vy_et + Y[5] * Vz_et);
            Vz_zt = (X[6] * Vx_xi + X[7] * Vy_xi + X[8] * Vz_xi) + (Y[6] * Vx_et + Y[7] * Vy_et + Y[8] * Vz_et);
        } else if constexpr (s == 1) {  // L2 (axis.h:92)
            const int k1 = k + FB3;
            float r2[3];
#pragma unroll
            for (int v = 0; v < 3; ++v)
                r2[v] = (float)FB3 * (SV<S>::ld(W[L.at(9, v, i, j, k1)]) - lds.v[v * VS + lc]) * rDH;
            Vx_zt = r2[0]; Vy_zt = r2[1]; Vz_zt = r2[2];
        } else {  // L3 (axis.h:93)
            const int k1 = k + FB3, k2 = k + 2 * FB3;
            float r3[3];
#pragma unroll
            for (int v = 0; v < 3; ++v)
                r3[v] = (float)FB3 *
                        (Cf1 * lds.v[v * VS + lc] + Cf2 * SV<S>::ld(W[L.at(9, v, i, j, k1)]) +
                         Cf3 * SV<S>::ld(W[L.at(9, v, i, j, k2)])) *
                        rDH;
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
        rk_update<S>(p, i, j, k, h);
    };
    layer(std::integral_constant<int, 0>{});
    layer(std::integral_constant<int, 1>{});
    layer(std::integral_constant<int, 2>{});
}

// Single-domain launch: gx * gy workgroups, dealt to the XCDs in contiguous runs.
template <class S, int TX, int TY>
__global__ void __launch_bounds__(TX* TY) k_stage_fs(StageParams p, int gx, int gy) {
    __shared__ FsLds<TX, TY> lds;
    const long long t = xcd_order(blockIdx.x, (long long)gridDim.x);
    fs_tile<S, TX, TY>(p, (int)(t % gx), (int)(t / gx), lds);
}

}  // namespace cg
