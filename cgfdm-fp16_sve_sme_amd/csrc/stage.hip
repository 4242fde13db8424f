// stage.hip — launchers of the fused RK-stage kernels: the bulk kernel over the interior box and
// the general kernel over the PML slabs and free-surface layers around it.
#include "stage_general.hpp"
#include "stage_bulk.hpp"

namespace cg {

// Bulk tile shapes (x by y columns per workgroup); the choice is a launch parameter.
struct BulkCfg { int tx, ty; };
static const BulkCfg kBulkCfg[] = {{64, 8}, {64, 4}};
constexpr int kNumBulkCfg = 2;

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

// Bulk box and the up to six boxes around it that the general kernel covers.
void stage_boxes(StageParams& q, int use_fast) {
    const int lo = HALO, hiI = HALO + q.nx, hiJ = HALO + q.ny, hiK = HALO + q.nz;
    const int np = q.pml ? q.nPML : 0;
    q.bi0 = lo + np; q.bi1 = hiI - np;
    q.bj0 = lo + np; q.bj1 = hiJ - np;
    q.bk0 = lo + np;
    q.bk1 = q.fs ? hiK - HALO : hiK - np;
    q.bulk_on = use_fast && q.bi1 > q.bi0 && q.bj1 > q.bj0 && q.bk1 > q.bk0;
    q.nbox = 0;
    q.boxStart[0] = 0;
    if (q.bulk_on) {
        add_box(q, lo, q.bi0, lo, hiJ, lo, hiK);            // x low side, full j, k
        add_box(q, q.bi1, hiI, lo, hiJ, lo, hiK);           // x high side
        add_box(q, q.bi0, q.bi1, lo, q.bj0, lo, hiK);       // y low side
        add_box(q, q.bi0, q.bi1, q.bj1, hiJ, lo, hiK);      // y high side
        add_box(q, q.bi0, q.bi1, q.bj0, q.bj1, lo, q.bk0);  // bottom
        add_box(q, q.bi0, q.bi1, q.bj0, q.bj1, q.bk1, hiK); // top (free surface or top PML)
    } else {
        add_box(q, lo, hiI, lo, hiJ, lo, hiK);
    }
}

template <class S>
static int launch_stage_t(const StageParams& p, int use_fast, hipStream_t s, hipEvent_t mid) {
    StageParams q = p;
    stage_boxes(q, use_fast);
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
        const int nch = (nzb + kc - 1) / kc;
        const unsigned grid = (unsigned)((long long)gx * gy * nch);
        if (cfg == 0) launch_bulk<S, 64, 8>(q, grid, kc, gx, gy, s);
        else launch_bulk<S, 64, 4>(q, grid, kc, gx, gy, s);
        if (hipGetLastError() != hipSuccess) return -3;
    }
    if (mid && hipEventRecord(mid, s) != hipSuccess) return -3;
    const long long n = q.boxStart[q.nbox];
    if (n > 0 && use_fast != 2) {
        const int bs = 256;
        k_stage_general<S><<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(q);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_stage(int prec, const StageParams& p, int use_fast, hipStream_t s, hipEvent_t mid) {
    switch (prec) {
        case P_F32: return launch_stage_t<float>(p, use_fast, s, mid);
        case P_F16: return launch_stage_t<_Float16>(p, use_fast, s, mid);
        case P_BF16: return launch_stage_t<__bf16>(p, use_fast, s, mid);
    }
    return -1;
}

}  // namespace cg
