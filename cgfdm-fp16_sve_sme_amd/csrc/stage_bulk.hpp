// stage_bulk.hpp — the hot kernel: one fused RK stage over the bulk of the grid (every point that
// is neither in a PML slab nor in the top three free-surface layers).
//
// Work decomposition (CDNA4, wave64):
//  - a workgroup owns a TX x TY column tile in (x, y) and marches a chunk of planes along z in the
//    direction FB3, so the one-sided z stencil (planes k-FB3 .. k+3FB3) lives in a 5-plane
//    register queue and each plane of W is read from HBM once per chunk;
//  - the current plane of the 9 fields, with its (4-wide) x/y halo, is staged in LDS as FP32 and
//    double-buffered, so one barrier per plane suffices; x and y neighbours are LDS reads;
//  - every global load of an iteration (the plane k+4FB3 entering the queue, the halo cells, the
//    12 metric/medium values and the RK registers) is unconditional and issued at the top of the
//    iteration, before any wait: no load sits in a branch, where the compiler would have to wait
//    for it inside the branch and serialise the iteration's memory latency;
//  - global accesses are buffer loads/stores: one descriptor per array and plane (k-major
//    [k][var][j][i] layout), a per-lane byte offset that does not change along the march and a
//    wave-uniform per-variable offset, so no VALU address arithmetic is spent;
//  - the RK stage is a template parameter, so the register update has no run-time branches;
//  - the z loop is unrolled by 5 (the queue rotates by renaming) and always runs a multiple of 5
//    iterations: the ones past the chunk load clamped planes and store nothing, so there is a
//    single copy of the unrolled body.
// Arithmetic is FP32 with the stencil coefficients pre-multiplied by FB*rDH and DT folded into
// lambda, mu and the buoyancy. In the bulk every PML factor is 1, so it is not applied.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace cg {

template <int TX, int TY> struct BulkLds {
    static constexpr int LX = TX + 4, LY = TY + 4, VS = LX * LY;  // LDS row pitch and var stride
    float v[2][9 * VS];
};

// One tile: columns [bi0 + bx*TX, +TX) x [bj0 + by*TY, +TY) of the bulk box, planes [kb, ke).
// Every thread of the workgroup calls it with the same arguments.
template <class S, int TX, int TY, int IRK>
__device__ __forceinline__ void bulk_tile(const StageParams& p, int FB1, int FB2, int FB3, int bx, int by, int kb,
                                          int ke, float (&lds)[2][9 * BulkLds<TX, TY>::VS]) {
    constexpr int LX = BulkLds<TX, TY>::LX, VS = BulkLds<TX, TY>::VS;
    constexpr int NT = TX * TY;

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int i0 = p.bi0 + bx * TX, j0 = p.bj0 + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int hx = FB1 > 0 ? 1 : 3, hy = FB2 > 0 ? 1 : 3;  // halo width on the low side
    const Layout L = p.lay;
    const long long plane = L.plane;
    const bool has = i < L.nx_ && j < L.ny_;     // a grid point (possibly outside the box)
    const bool inside = i < p.bi1 && j < p.bj1;  // computed by this kernel
    const unsigned col = has ? (unsigned)L.col(i, j) * (unsigned)sizeof(S) : 0u;  // byte offset in a plane
    const int vs = (int)(plane * (long long)sizeof(S));  // bytes per variable plane

    // z chunk, traversed in the direction FB3
    const int nk = ke - kb;
    const int k0 = FB3 > 0 ? kb : ke - 1;
    const int kmax = L.nz_ - 1;

    const char* __restrict__ Wb = (const char*)p.Win;
    const char* __restrict__ Cb = (const char*)p.cjm;
    const char* __restrict__ Ab = (const char*)p.A;
    char* __restrict__ Tb = (char*)p.T;
    char* __restrict__ Ob = (char*)p.Wout;
    const long long wpl = 9 * plane * (long long)sizeof(S);   // bytes per plane of W, A, T, Wout
    const long long cpl = 13 * plane * (long long)sizeof(S);  // bytes per plane of the metric/medium
    const long long wrange = 9 * plane, crange = 13 * plane;  // elements per plane

    // stencil weights with direction and 1/DH folded in
    const float rDH = p.rDH;
    const float cx0 = FB1 * af_1 * rDH, cx1 = FB1 * af0 * rDH, cx2 = FB1 * af1 * rDH, cx3 = FB1 * af2 * rDH,
                cx4 = FB1 * af3 * rDH;
    const float cy0 = FB2 * af_1 * rDH, cy1 = FB2 * af0 * rDH, cy2 = FB2 * af1 * rDH, cy3 = FB2 * af2 * rDH,
                cy4 = FB2 * af3 * rDH;
    const float cz0 = FB3 * af_1 * rDH, cz1 = FB3 * af0 * rDH, cz2 = FB3 * af1 * rDH, cz3 = FB3 * af2 * rDH,
                cz4 = FB3 * af3 * rDH;

    // register queue: slot (m + R) % 5 holds W(k + FB3*(m-1)), m = 0..4, at unrolled step R
    float q[5][9];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const Rsrc rm = make_rsrc((const S*)(Wb + (long long)(k0 + FB3 * (m - 1)) * wpl), wrange);
#pragma unroll
        for (int v = 0; v < 9; ++v) q[m][v] = bld<S>(rm, col, v * vs);
    }

    // halo cells of this thread: cell c = tid + r*NT (r < HPT) of the list x-halo (4 columns x TY
    // rows) then y-halo (4 rows x LX columns); a thread without a cell loads element 0 in vain
    constexpr int NHX = 4 * TY, NHY = 4 * LX, NH = NHX + NHY;
    constexpr int HPT = (NH + NT - 1) / NT;
    int hl[HPT];
    unsigned hcol[HPT];
    bool hown[HPT];
#pragma unroll
    for (int r = 0; r < HPT; ++r) {
        const int c = tid + r * NT;
        hl[r] = 0;
        hcol[r] = 0;
        hown[r] = c < NH;
        if (c < NH) {
            int hr, hc;
            if (c < NHX) {
                const int rr = c / 4, c4 = c % 4;
                hr = hy + rr;
                hc = c4 < hx ? c4 : TX + c4;  // low-side columns [0,hx), high-side [hx+TX, LX)
            } else {
                const int t = c - NHX;
                const int r4 = t / LX;
                hc = t % LX;
                hr = r4 < hy ? r4 : TY + r4;
            }
            const int ig = i0 - hx + hc, jg = j0 - hy + hr;
            const bool hhas = ig >= 0 && ig < L.nx_ && jg >= 0 && jg < L.ny_;
            hcol[r] = hhas ? (unsigned)L.col(ig, jg) * (unsigned)sizeof(S) : 0u;
            hl[r] = hr * LX + hc;
        }
    }
    const int lc = (ty + hy) * LX + tx + hx;
    const float DT = p.DT;
    const float Crho = p.Crho;
    int k = k0;

    auto step = [&](auto Rc, int it) {
        constexpr int R = decltype(Rc)::value;
        float(&q0)[9] = q[(0 + R) % 5];
        float(&q1)[9] = q[(1 + R) % 5];
        float(&q2)[9] = q[(2 + R) % 5];
        float(&q3)[9] = q[(3 + R) % 5];
        float(&q4)[9] = q[(4 + R) % 5];
        const bool active = it < nk;  // iterations past the chunk store nothing
        float* Lb = lds[it & 1];
        const int kc = min(max(k, 0), kmax);  // clamped plane (only the inactive iterations need it)
        const long long pw = (long long)kc * wpl;
        const Rsrc rW = make_rsrc((const S*)(Wb + pw), wrange);
        float hv[HPT][9];
#pragma unroll
        for (int r = 0; r < HPT; ++r)
#pragma unroll
            for (int v = 0; v < 9; ++v) hv[r][v] = bld<S>(rW, hcol[r], v * vs);
        // the plane entering the queue (clamped at the grid end; unused there)
        float nq[9];
        {
            const int kn = min(max(k + 4 * FB3, 0), kmax);
            const Rsrc rN = make_rsrc((const S*)(Wb + (long long)kn * wpl), wrange);
#pragma unroll
            for (int v = 0; v < 9; ++v) nq[v] = bld<S>(rN, col, v * vs);
        }
        float M[12];
        {
            const Rsrc rC = make_rsrc((const S*)(Cb + (long long)kc * cpl), crange);
#pragma unroll
            for (int c = 0; c < 9; ++c) M[c] = bld<S>(rC, col, c * vs);
#pragma unroll
            for (int c = 10; c < 13; ++c) M[c - 1] = bld<S>(rC, col, c * vs);
        }
        const Rsrc rT = make_rsrc((const S*)(Tb + pw), wrange);
        float ra[9], rt[9];
        if constexpr (IRK == 1 || IRK == 2) {
            const Rsrc rA = make_rsrc((const S*)(Ab + pw), wrange);
#pragma unroll
            for (int v = 0; v < 9; ++v) ra[v] = bld<S>(rA, col, v * vs);
        }
        if constexpr (IRK >= 1) {
#pragma unroll
            for (int v = 0; v < 9; ++v) rt[v] = bld<S>(rT, col, v * vs);
        }
        // stage the centre plane and its halo in LDS
#pragma unroll
        for (int v = 0; v < 9; ++v) Lb[v * VS + lc] = q1[v];
#pragma unroll
        for (int r = 0; r < HPT; ++r)
            if (hown[r]) {
#pragma unroll
                for (int v = 0; v < 9; ++v) Lb[v * VS + hl[r]] = hv[r][v];
            }
        __syncthreads();
        if (inside && active) {
            const float mu = M[9] * DT, lam = M[10] * DT, bu = M[11] * (Crho * DT);
            const int ox0 = lc - FB1, ox2 = lc + FB1, ox3 = lc + 2 * FB1, ox4 = lc + 3 * FB1;
            const int oy0 = lc - FB2 * LX, oy2 = lc + FB2 * LX, oy3 = lc + 2 * FB2 * LX, oy4 = lc + 3 * FB2 * LX;
            // The 27 one-sided derivatives are folded into the physical velocity gradient
            // G[a][c] = sum_d m^d_c dV_a/dxi_d and the stress divergence dv[a] as they are formed,
            // so only the three derivatives of one variable are live at a time.
            float G[3][3];
            float dv[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                const float* Lv = Lb + v * VS;
                const float c = q1[v];
                const float d0 = cx0 * Lv[ox0] + cx1 * c + cx2 * Lv[ox2] + cx3 * Lv[ox3] + cx4 * Lv[ox4];
                const float d1 = cy0 * Lv[oy0] + cy1 * c + cy2 * Lv[oy2] + cy3 * Lv[oy3] + cy4 * Lv[oy4];
                const float d2 = cz0 * q0[v] + cz1 * c + cz2 * q2[v] + cz3 * q3[v] + cz4 * q4[v];
                if (v < 3) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) G[v][cc] = M[cc] * d0 + M[3 + cc] * d1 + M[6 + cc] * d2;
                } else {
                    // stress component (row r0, column c0) contributes d/dx_c0 to dv[r0] and, off the
                    // diagonal, d/dx_r0 to dv[c0]
                    const int r0 = v == 3 ? 0 : v == 4 ? 1 : v == 5 ? 2 : v == 6 ? 0 : v == 7 ? 0 : 1;
                    const int c0 = v == 3 ? 0 : v == 4 ? 1 : v == 5 ? 2 : v == 6 ? 1 : v == 7 ? 2 : 2;
                    dv[r0] += M[c0] * d0 + M[3 + c0] * d1 + M[6 + c0] * d2;
                    if (v >= 6) dv[c0] += M[r0] * d0 + M[3 + r0] * d1 + M[6 + r0] * d2;
                }
            }
            float h[9];
            h[0] = dv[0] * bu;
            h[1] = dv[1] * bu;
            h[2] = dv[2] * bu;
            const float div = lam * (G[0][0] + G[1][1] + G[2][2]);
            const float mu2 = 2.0f * mu;
            h[3] = div + mu2 * G[0][0];
            h[4] = div + mu2 * G[1][1];
            h[5] = div + mu2 * G[2][2];
            h[6] = mu * (G[0][1] + G[1][0]);
            h[7] = mu * (G[0][2] + G[2][0]);
            h[8] = mu * (G[1][2] + G[2][1]);
            // RK registers (wave_rk0..3)
            const Rsrc rO = make_rsrc((const S*)(Ob + pw), wrange);
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                if constexpr (IRK == 0) {
                    const float m0 = q1[v];
                    bst<S>(rO, col, v * vs, m0 + alpha2 * h[v]);
                    bst<S>(rT, col, v * vs, m0 + beta1 * h[v]);
                } else if constexpr (IRK == 1) {
                    bst<S>(rO, col, v * vs, ra[v] + alpha3 * h[v]);
                    bst<S>(rT, col, v * vs, rt[v] + beta2 * h[v]);
                } else if constexpr (IRK == 2) {
                    bst<S>(rO, col, v * vs, ra[v] + h[v]);
                    bst<S>(rT, col, v * vs, rt[v] + beta3 * h[v]);
                } else {
                    bst<S>(rO, col, v * vs, rt[v] + beta4 * h[v]);
                }
            }
        }
        // the slot of the oldest plane receives the new one
#pragma unroll
        for (int v = 0; v < 9; ++v) q0[v] = nq[v];
        k += FB3;
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    for (int it = 0; it < nk; it += 5) {
        step(I0{}, it);
        step(I1{}, it + 1);
        step(I2{}, it + 2);
        step(I3{}, it + 3);
        step(I4{}, it + 4);
    }
}

// Single-domain launch: a 1-D grid of gx*gy*nch workgroups (x tiles fastest, then y tiles, then
// z chunks of kc planes), dealt to the XCDs in contiguous runs (xcd_order). One instantiation per
// RK stage.
template <class S, int TX, int TY, int IRK>
__global__ void __launch_bounds__(TX* TY) k_stage_bulk(StageParams p, int kc, int gx, int gy) {
    __shared__ BulkLds<TX, TY> lds;
    const long long t = xcd_order(blockIdx.x, (long long)gridDim.x);
    const int bx = (int)(t % gx), by = (int)((t / gx) % gy), bz = (int)(t / ((long long)gx * gy));
    const int kb = p.bk0 + bz * kc;
    const int ke = min(kb + kc, p.bk1);
    bulk_tile<S, TX, TY, IRK>(p, p.FB1, p.FB2, p.FB3, bx, by, kb, ke, lds.v);
}

}  // namespace cg
