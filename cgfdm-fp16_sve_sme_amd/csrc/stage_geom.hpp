// stage_geom.hpp — host-side geometry shared by the launchers, the C ABI and the serving engine:
// the internal layout of a domain, the PML slab sizes, the decomposition of the interior into
// boxes (bulk, PML faces/edges/corners, free-surface and general boxes) and the launch geometry
// of the bulk and PML tile kernels. Plain C++ over the structures of common.hpp and
// stage_bulk.hpp, so it also compiles in a host-only build of the kernels.
#pragma once
#include <algorithm>

#include "../../include/cgfdm.h"
#include "common.hpp"
#include "stage_bulk.hpp"

namespace cg {

// Internal layout (Layout in common.hpp): rows padded to whole 64-column blocks and shifted so
// that the first column of the bulk box (3 + nPML) starts a block.
inline Layout make_layout(const cgfdm_desc& d) {
    Layout L{};
    L.nx_ = d.NX + 6; L.ny_ = d.NY + 6; L.nz_ = d.NZ + 6;
    const int bi0 = HALO + (d.pml ? d.nPML : 0);
    L.lead = (kBlk - bi0 % kBlk) % kBlk;
    L.px = (L.lead + L.nx_ + kBlk - 1) / kBlk * kBlk;
    L.nb = L.px / kBlk;
    L.plane = (long long)L.px * L.ny_;
    return L;
}

// Points of PML slab s (x1, y1, z1, x2, y2, z2; the top z slab is absent under a free surface).
inline long long slab_points(const cgfdm_desc& d, int s) {
    if (!d.pml) return 0;
    if (s == 5 && d.free_surface) return 0;
    const long long n = d.nPML;
    switch (s % 3) {
        case 0: return n * d.NY * d.NZ;
        case 1: return n * d.NX * d.NZ;
        default: return n * d.NX * d.NY;
    }
}

inline void add_box(StageParams& q, int i0, int i1, int j0, int j1, int k0, int k1) {
    if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
    int* B = q.box[q.nbox];
    B[0] = i0; B[1] = i1; B[2] = j0; B[3] = j1; B[4] = k0; B[5] = k1;
    q.boxStart[q.nbox + 1] = q.boxStart[q.nbox] + (long long)(i1 - i0) * (j1 - j0) * (k1 - k0);
    ++q.nbox;
}

// PML box inside the slabs sx (x normal), sy, sz (-1: none).
inline void add_pbox(StageParams& q, int i0, int i1, int j0, int j1, int k0, int k1, int sx, int sy, int sz) {
    if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
    int* B = q.pbox[q.npbox++];
    B[0] = i0; B[1] = i1; B[2] = j0; B[3] = j1; B[4] = k0; B[5] = k1;
    B[6] = (sx >= 0 ? 1 : 0) | (sy >= 0 ? 2 : 0) | (sz >= 0 ? 4 : 0);
    B[7] = (sx >= 0 ? sx : 15) | (sy >= 0 ? sy : 15) << 4 | (sz >= 0 ? sz : 15) << 8;
}

// Partition of the interior. Along each axis: the low slab [lo, lo+np), the middle and the high
// slab (z: bottom slab, middle, and either the three free-surface layers or the top slab). The
// middle^3 cell is the bulk box; every other cell below the free-surface layers is a PML box of
// the slabs it lies in (faces: one slab, edges: two, corners: three); with a free surface, its
// three layers are the free-surface box above the bulk columns and four general boxes over the
// PML columns. With use_fast == 0 the whole interior is one general box.
inline void stage_boxes(StageParams& q, int use_fast) {
    const int lo = HALO, hiI = HALO + q.nx, hiJ = HALO + q.ny, hiK = HALO + q.nz;
    const int np = q.pml ? q.nPML : 0;
    const int I0 = lo, I1 = lo + np, I2 = hiI - np;
    const int J0 = lo, J1 = lo + np, J2 = hiJ - np;
    const int K0 = lo, K1 = lo + np;
    const int K2 = q.fs ? hiK - HALO : hiK - np;  // start of the top range
    q.nbox = 0;
    q.npbox = 0;
    q.fs_on = 0;
    q.boxStart[0] = 0;
    q.bi0 = I1; q.bi1 = I2;
    q.bj0 = J1; q.bj1 = J2;
    q.bk0 = K1; q.bk1 = K2;
    q.bulk_on = use_fast && I2 > I1 && J2 > J1 && K2 > K1;
    if (!use_fast || !(I2 > I1 && J2 > J1 && K2 > K1)) {
        q.bulk_on = 0;
        add_box(q, lo, hiI, lo, hiJ, lo, hiK);
        return;
    }
    if (np > 0) {
        // cells of the 3x3x3 split (the top z cell only without a free surface); slab indices
        // x1,y1,z1,x2,y2,z2 = 0..5
        const int xr[3][3] = {{I0, I1, 0}, {I1, I2, -1}, {I2, hiI, 3}};
        const int yr[3][3] = {{J0, J1, 1}, {J1, J2, -1}, {J2, hiJ, 4}};
        const int zr[3][3] = {{K0, K1, 2}, {K1, K2, -1}, {K2, hiK, 5}};
        const int nzc = q.fs ? 2 : 3;
        // faces first, then edges, then corners (the kinds' launch order)
        for (int nslab = 1; nslab <= 3; ++nslab)
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a) {
                        if (c >= nzc) continue;
                        const int n = (xr[a][2] >= 0) + (yr[b][2] >= 0) + (zr[c][2] >= 0);
                        if (n != nslab) continue;
                        add_pbox(q, xr[a][0], xr[a][1], yr[b][0], yr[b][1], zr[c][0], zr[c][1], xr[a][2], yr[b][2],
                                 zr[c][2]);
                    }
    }
    if (q.fs) {
        // free-surface layers: above the bulk columns by the free-surface tile kernel, the ring
        // of PML columns by the general kernel
        q.fs_on = 1;
        q.fk0 = K2;
        q.fk1 = hiK;
        add_box(q, lo, hiI, lo, J1, K2, hiK);
        add_box(q, lo, hiI, J2, hiJ, K2, hiK);
        add_box(q, lo, I1, J1, J2, K2, hiK);
        add_box(q, I2, hiI, J1, J2, K2, hiK);
    }
}

// A z-chunk length for the tile kernels, whose loop runs a multiple of 5 iterations: kc rounded
// up to a multiple of 5 (at most nzb), so only the last chunk of a box has idle iterations. Used
// for the bulk box; the PML boxes keep more, shorter chunks (their idle iterations load only
// planes L2 holds, and a domain served alone needs the parallelism: rounding their chunks up
// served one request at a time 11 % slower).
inline int chunk5(int kc, int nzb) { return std::max(1, std::min(nzb, (kc + 4) / 5 * 5)); }

// Tiles of every PML box, grouped by kind (pml_mask_at order); z chunks sized so each kind has
// about `want` tiles (a single domain: enough workgroups to fill the GPU; one of many batched
// domains: its share), with chunks of at least 5 planes (the tile loop's unit): the small edge
// and corner boxes are cut into short chunks so their launches are not one long z march.
inline PmlTiles pml_tiles(const StageParams& q, int want) {
    PmlTiles t{};
    t.start[0] = 0;
    for (int g = 0; g < kNumPmlMasks; ++g) {
        const int mask = pml_mask_at(g);
        const int TX = pml_tx(mask), TY = pml_ty(mask);
        t.first[g] = t.n;
        int cols = 0;
        for (int b = 0; b < q.npbox; ++b) {
            const int* B = q.pbox[b];
            if (B[6] != mask) continue;
            const int e = t.n++;
            t.box[e] = b;
            t.gx[e] = (B[1] - pml_tile_x0(q.lay, B[0], B[1], TX) + TX - 1) / TX;
            t.gy[e] = (B[3] - B[2] + TY - 1) / TY;
            cols += t.gx[e] * t.gy[e];
        }
        for (int e = t.first[g]; e < t.n; ++e) {
            const int* B = q.pbox[t.box[e]];
            const int nzb = B[5] - B[4];
            const int nch = std::max(1, std::min(nzb / 5, (want + cols - 1) / std::max(cols, 1)));
            t.kc[e] = (nzb + nch - 1) / nch;
            t.nch[e] = (nzb + t.kc[e] - 1) / t.kc[e];
            t.start[e + 1] = t.start[e] + t.gx[e] * t.gy[e] * t.nch[e];
        }
    }
    t.first[kNumPmlMasks] = t.n;
    return t;
}

// Bulk tile shapes (x by y columns per workgroup); the choice is a launch parameter.
struct BulkCfg { int tx, ty; };
constexpr BulkCfg kBulkCfg[] = {{64, 8}, {64, 4}, {64, 16}, {128, 8}};
constexpr int kNumBulkCfg = 4;

// Launch geometry of the bulk kernel over the bulk box of q (after stage_boxes): tile shape cfg,
// gx x gy column tiles and nch chunks of kc planes.
struct BulkGeom { int cfg, tx, ty, gx, gy, nch, kc; };
inline BulkGeom bulk_geom(const StageParams& q) {
    BulkGeom g{};
    g.cfg = (q.bulk_cfg >= 0 && q.bulk_cfg < kNumBulkCfg) ? q.bulk_cfg : 0;
    g.tx = kBulkCfg[g.cfg].tx;
    g.ty = kBulkCfg[g.cfg].ty;
    const int nxb = q.bi1 - q.bi0, nyb = q.bj1 - q.bj0, nzb = q.bk1 - q.bk0;
    g.gx = (nxb + g.tx - 1) / g.tx;
    g.gy = (nyb + g.ty - 1) / g.ty;
    int kc = q.bulk_kc;
    if (kc <= 0) {
        const long long want = 2048;  // enough workgroups to fill 256 CUs several times
        const int nch = (int)std::max<long long>(1, (want + (long long)g.gx * g.gy - 1) / ((long long)g.gx * g.gy));
        kc = (nzb + nch - 1) / nch;
    }
    kc = std::max(1, std::min(kc, nzb));
    // a requested length is the longest chunk: the planes are split evenly over as many chunks
    // (a short last chunk would pay the queue's start-up planes for little work), then the length
    // is rounded up to a multiple of 5 (chunk5)
    g.nch = (nzb + kc - 1) / kc;
    g.kc = chunk5((nzb + g.nch - 1) / g.nch, nzb);
    g.nch = (nzb + g.kc - 1) / g.kc;
    return g;
}

}  // namespace cg
