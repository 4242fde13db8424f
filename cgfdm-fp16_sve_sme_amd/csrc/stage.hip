// stage.hip — launchers of the fused RK-stage kernels: the bulk kernel over the interior box, the
// PML tile kernels over the slabs around it, the free-surface tile kernel and the general kernel
// over the free-surface layers of the PML columns.
#include "stage_general.hpp"
#include "stage_bulk.hpp"
#include "stage_fs.hpp"

namespace cg {

// Bulk tile shapes (x by y columns per workgroup); the choice is a launch parameter.
struct BulkCfg { int tx, ty; };
static const BulkCfg kBulkCfg[] = {{64, 8}, {64, 4}, {64, 16}, {128, 8}};
constexpr int kNumBulkCfg = 4;

template <class S, int TX, int TY>
static void launch_bulk(const StageParams& q, unsigned grid, int kc, int gx, int gy, hipStream_t s) {
    switch (q.irk) {
        case 0: k_stage_bulk<S, TX, TY, 0><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        case 1: k_stage_bulk<S, TX, TY, 1><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        case 2: k_stage_bulk<S, TX, TY, 2><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        default: k_stage_bulk<S, TX, TY, 3><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
    }
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

static void add_box(StageParams& q, int i0, int i1, int j0, int j1, int k0, int k1) {
    if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
    int* B = q.box[q.nbox];
    B[0] = i0; B[1] = i1; B[2] = j0; B[3] = j1; B[4] = k0; B[5] = k1;
    q.boxStart[q.nbox + 1] = q.boxStart[q.nbox] + (long long)(i1 - i0) * (j1 - j0) * (k1 - k0);
    ++q.nbox;
}

// PML box inside the slabs sx (x normal), sy, sz (-1: none).
static void add_pbox(StageParams& q, int i0, int i1, int j0, int j1, int k0, int k1, int sx, int sy, int sz) {
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
void stage_boxes(StageParams& q, int use_fast) {
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

// Tiles of every PML box, grouped by kind (pml_mask_at order); z chunks sized so each kind has
// about `want` tiles (a single domain: enough workgroups to fill the GPU; one of many batched
// domains: its share), with chunks of at least 5 planes (the tile loop's unit): the small edge
// and corner boxes are cut into short chunks so their launches are not one long z march.
PmlTiles pml_tiles(const StageParams& q, int want) {
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

template <class S, int MASK>
static void launch_pml(const StageParams& q, const PmlTiles& t, hipStream_t s) {
    constexpr int g = pml_group_of(MASK), TX = pml_tx(MASK), TY = pml_ty(MASK);
    const int tiles = t.start[t.first[g + 1]] - t.start[t.first[g]];
    if (tiles <= 0) return;
    const unsigned grid = (unsigned)tiles;
    switch (q.irk) {
        case 0: k_stage_pml<S, TX, TY, 0, MASK><<<grid, TX * TY, 0, s>>>(q, t); break;
        case 1: k_stage_pml<S, TX, TY, 1, MASK><<<grid, TX * TY, 0, s>>>(q, t); break;
        case 2: k_stage_pml<S, TX, TY, 2, MASK><<<grid, TX * TY, 0, s>>>(q, t); break;
        default: k_stage_pml<S, TX, TY, 3, MASK><<<grid, TX * TY, 0, s>>>(q, t); break;
    }
}

template <class S>
static int launch_stage_t(const StageParams& p, int use_fast, hipStream_t s, hipEvent_t mid, const StageFork* fk) {
    StageParams q = p;
    stage_boxes(q, use_fast);
    // boundary kernels on the side stream, after everything enqueued on s so far
    const bool forked = fk && fk->side && use_fast != 2;
    if (forked) {
        if (hipEventRecord(fk->fork, s) != hipSuccess || hipStreamWaitEvent(fk->side, fk->fork, 0) != hipSuccess)
            return -3;
    }
    const hipStream_t sb = forked ? fk->side : s;
    if (q.bulk_on) {
        const int cfg = (p.bulk_cfg >= 0 && p.bulk_cfg < kNumBulkCfg) ? p.bulk_cfg : 0;
        const int TX = kBulkCfg[cfg].tx, TY = kBulkCfg[cfg].ty;
        const int nxb = q.bi1 - q.bi0, nyb = q.bj1 - q.bj0, nzb = q.bk1 - q.bk0;
        const int gx = (nxb + TX - 1) / TX, gy = (nyb + TY - 1) / TY;
        int kc = p.bulk_kc;
        if (kc <= 0) {
            const long long want = 2048;  // enough workgroups to fill 256 CUs several times
            const int nch = (int)std::max<long long>(1, (want + (long long)gx * gy - 1) / ((long long)gx * gy));
            kc = (nzb + nch - 1) / nch;
        }
        kc = std::max(1, std::min(kc, nzb));
        // a requested length is the longest chunk: the planes are split evenly over as many
        // chunks (a short last chunk would pay the queue's start-up planes for little work)
        const int nch = (nzb + kc - 1) / kc;
        kc = (nzb + nch - 1) / nch;
        const unsigned grid = (unsigned)((long long)gx * gy * nch);
        switch (cfg) {
            case 0: launch_bulk<S, 64, 8>(q, grid, kc, gx, gy, s); break;
            case 1: launch_bulk<S, 64, 4>(q, grid, kc, gx, gy, s); break;
            case 2: launch_bulk<S, 64, 16>(q, grid, kc, gx, gy, s); break;
            default: launch_bulk<S, 128, 8>(q, grid, kc, gx, gy, s); break;
        }
        if (hipGetLastError() != hipSuccess) return -3;
    }
    if (mid && hipEventRecord(mid, s) != hipSuccess) return -3;
    if (use_fast != 2) {
        const PmlTiles t = pml_tiles(q, 1024);
        launch_pml<S, 1>(q, t, sb);
        launch_pml<S, 2>(q, t, sb);
        launch_pml<S, 4>(q, t, sb);
        launch_pml<S, 3>(q, t, sb);
        launch_pml<S, 5>(q, t, sb);
        launch_pml<S, 6>(q, t, sb);
        launch_pml<S, 7>(q, t, sb);
        if (hipGetLastError() != hipSuccess) return -3;
    }
    if (q.fs_on && use_fast != 2) {
        const int gx = (q.bi1 - q.bi0 + kFsTx - 1) / kFsTx, gy = (q.bj1 - q.bj0 + kFsTy - 1) / kFsTy;
        k_stage_fs<S, kFsTx, kFsTy><<<(unsigned)(gx * gy), kFsTx * kFsTy, 0, sb>>>(q, gx, gy);
    }
    const long long n = q.boxStart[q.nbox];
    if (n > 0 && use_fast != 2) {
        const int bs = 256;
        k_stage_general<S><<<(unsigned)((n + bs - 1) / bs), bs, 0, sb>>>(q);
    }
    if (hipGetLastError() != hipSuccess) return -3;
    if (forked) {
        if (hipEventRecord(fk->join, fk->side) != hipSuccess || hipStreamWaitEvent(s, fk->join, 0) != hipSuccess)
            return -3;
    }
    return 0;
}

int launch_stage(int prec, const StageParams& p, int use_fast, hipStream_t s, hipEvent_t mid, const StageFork* fk) {
    switch (prec) {
        case P_F32: return launch_stage_t<float>(p, use_fast, s, mid, fk);
        case P_F16: return launch_stage_t<_Float16>(p, use_fast, s, mid, fk);
        case P_BF16: return launch_stage_t<__bf16>(p, use_fast, s, mid, fk);
    }
    return -1;
}

}  // namespace cg
