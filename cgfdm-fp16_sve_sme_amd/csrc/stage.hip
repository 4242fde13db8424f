// stage.hip — launchers of the fused RK-stage kernels: the bulk kernel over the interior box, the
// PML tile kernels over the slabs around it, the free-surface tile kernel and the general kernel
// over the free-surface layers of the PML columns.
#include "stage_general.hpp"
#include "stage_bulk.hpp"
#include "stage_fs.hpp"
#include "stage_geom.hpp"

namespace cg {

template <class S, int TX, int TY>
static void launch_bulk(const StageParams& q, unsigned grid, int kc, int gx, int gy, hipStream_t s) {
    switch (q.irk) {
        case 0: k_stage_bulk<S, TX, TY, 0><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        case 1: k_stage_bulk<S, TX, TY, 1><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        case 2: k_stage_bulk<S, TX, TY, 2><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
        default: k_stage_bulk<S, TX, TY, 3><<<grid, TX * TY, 0, s>>>(q, kc, gx, gy); break;
    }
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
    // boundary kernels on the side stream(s), after everything enqueued on s so far: with three,
    // the x faces on the first, the y and bottom faces on the second, edges, corners, the
    // free-surface layers and the general boxes on the third
    const bool forked = fk && fk->side && use_fast != 2;
    const bool three = forked && fk->side2[0] && fk->side2[1];
    if (forked) {
        if (hipEventRecord(fk->fork, s) != hipSuccess || hipStreamWaitEvent(fk->side, fk->fork, 0) != hipSuccess)
            return -3;
        if (three)
            for (int w = 0; w < 2; ++w)
                if (hipStreamWaitEvent(fk->side2[w], fk->fork, 0) != hipSuccess) return -3;
    }
    const hipStream_t sb = forked ? fk->side : s;
    const hipStream_t s1 = three ? fk->side2[0] : sb, s2 = three ? fk->side2[1] : sb;
    if (q.bulk_on) {
        const BulkGeom bg = bulk_geom(q);
        const int cfg = bg.cfg, kc = bg.kc, gx = bg.gx, gy = bg.gy;
        const unsigned grid = (unsigned)((long long)gx * gy * bg.nch);
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
        launch_pml<S, 2>(q, t, s1);
        launch_pml<S, 4>(q, t, s1);
        launch_pml<S, 3>(q, t, s2);
        launch_pml<S, 5>(q, t, s2);
        launch_pml<S, 6>(q, t, s2);
        launch_pml<S, 7>(q, t, s2);
        if (hipGetLastError() != hipSuccess) return -3;
    }
    if (q.fs_on && use_fast != 2) {
        const int gx = (q.bi1 - q.bi0 + kFsTx - 1) / kFsTx, gy = (q.bj1 - q.bj0 + kFsTy - 1) / kFsTy;
        k_stage_fs<S, kFsTx, kFsTy><<<(unsigned)(gx * gy), kFsTx * kFsTy, 0, s2>>>(q, gx, gy);
    }
    const long long n = q.boxStart[q.nbox];
    if (n > 0 && use_fast != 2) {
        const int bs = 256;
        k_stage_general<S><<<(unsigned)((n + bs - 1) / bs), bs, 0, s2>>>(q);
    }
    if (hipGetLastError() != hipSuccess) return -3;
    if (forked) {
        if (hipEventRecord(fk->join, fk->side) != hipSuccess || hipStreamWaitEvent(s, fk->join, 0) != hipSuccess)
            return -3;
        if (three)
            for (int w = 0; w < 2; ++w)
                if (hipEventRecord(fk->join2[w], fk->side2[w]) != hipSuccess ||
                    hipStreamWaitEvent(s, fk->join2[w], 0) != hipSuccess)
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
