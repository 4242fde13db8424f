// batch.hip — multi-domain execution: many independent simulations advanced together, one
// launch per kernel for all of them.
//
// Every kernel here is persistent over a fixed grid and reads what it works on from device
// tables (slot table, per-slot stage parameters, tile and chunk lists). A sequence of whole time
// steps built from these launches can therefore be captured once in a HIP graph and replayed
// for any set of resident domains: admitting or retiring a domain only rewrites the tables
// (k_write_slot, k_clear_slot, k_build_tables, all outside the graph) and never changes a grid.
// Each domain keeps its own step counter on the device; k_advance moves it and the MacCormack
// signs of the next step, and a domain whose counter reached its NT is skipped by every kernel.
//
// The per-point arithmetic is the single-domain code (bulk_tile, fs_tile, stage_point_general
// and the side bodies), so a domain computes bit-identical results alone and in any mix.
#include "batch.hpp"
#include "side.hpp"
#include "stage_general.hpp"

namespace cg {

namespace {

__device__ const int kFB[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                                  {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};  // propagate.cpp:107-117

// Stage input/output buffers by RK stage: A -> B -> C -> B -> A (T accumulates).
__device__ __forceinline__ int stage_in_idx(int irk) { return irk == 0 ? 0 : (irk == 2 ? 2 : 1); }
__device__ __forceinline__ int stage_out_idx(int irk) { return irk == 0 ? 1 : (irk == 1 ? 2 : (irk == 2 ? 1 : 0)); }

__device__ __forceinline__ void set_fb(StageParams& p, int it, int irk) {
    const int sgn = (irk & 1) ? -1 : 1;
    p.FB1 = kFB[it % 8][0] * sgn;
    p.FB2 = kFB[it % 8][1] * sgn;
    p.FB3 = kFB[it % 8][2] * sgn;
}

// Uniform binary search: last entry whose start <= t (entries sorted by start, n >= 1).
template <class E, class T>
__device__ __forceinline__ int find_entry(const E* __restrict__ e, int n, T t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid].start <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// ---------------------------------------------------------------- table maintenance
__global__ void k_write_slot(BatchDev b, int s, SlotDev sd, StageParams base) {
    const int irk = threadIdx.x;
    if (irk < 4) {
        StageParams p = base;
        p.irk = irk;
        p.Win = sd.W[stage_in_idx(irk)];
        p.Wout = sd.W[stage_out_idx(irk)];
        set_fb(p, sd.it, irk);
        b.sp[s * 4 + irk] = p;
    }
    if (irk == 0) b.slots[s] = sd;
}

__global__ void k_clear_slot(BatchDev b, int s) {
    if (threadIdx.x == 0) b.slots[s].active = 0;
}

// Prefix sums of bulk tiles and general chunks over the active slots, in slot order.
__global__ void k_build_tables(BatchDev b) {
    if (threadIdx.x != 0) return;
    int nb = 0, ng = 0, tiles = 0, nf = 0, ftiles = 0;
    int np[kNumPmlMasks] = {}, ptiles[kNumPmlMasks] = {};
    long long chunks = 0;
    for (int s = 0; s < b.maxSlots; ++s) {
        const SlotDev& sd = b.slots[s];
        if (!sd.active) continue;
        if (sd.gx > 0) {
            BulkEnt& e = b.tab->bulk[nb++];
            e.slot = s; e.gx = sd.gx; e.gy = sd.gy; e.nch = sd.nch; e.kc = sd.kc; e.start = tiles;
            tiles += sd.gx * sd.gy * sd.nch;
        }
        if (sd.fgx > 0) {
            FsEnt& e = b.tab->fs[nf++];
            e.slot = s; e.gx = sd.fgx; e.gy = sd.fgy; e.start = ftiles;
            ftiles += sd.fgx * sd.fgy;
        }
        for (int x = 0; x < sd.nbox; ++x) {
            GenEnt& e = b.tab->gen[ng++];
            e.slot = s; e.box = x; e.start = chunks;
            chunks += sd.boxChunks[x];
        }
        const PmlTiles& t = sd.pt;
        for (int g = 0; g < kNumPmlMasks; ++g)
            for (int x = t.first[g]; x < t.first[g + 1]; ++x) {
                PmlEnt& e = b.tab->pml[g][np[g]++];
                e.slot = s; e.box = t.box[x]; e.gx = t.gx[x]; e.gy = t.gy[x]; e.nch = t.nch[x]; e.kc = t.kc[x];
                e.start = ptiles[g];
                ptiles[g] += t.gx[x] * t.gy[x] * t.nch[x];
            }
    }
    b.tab->nBulk = nb;
    b.tab->nTiles = tiles;
    b.tab->nGen = ng;
    b.tab->nChunks = chunks;
    b.tab->nFs = nf;
    b.tab->nFsTiles = ftiles;
    for (int g = 0; g < kNumPmlMasks; ++g) {
        b.tab->nPml[g] = np[g];
        b.tab->nPmlTiles[g] = ptiles[g];
    }
}

// ---------------------------------------------------------------- per-step kernels
template <class S>
__global__ void k_source_multi(BatchDev b) {
    const SlotDev& sd = b.slots[blockIdx.y];
    if (!sd.active || sd.it >= sd.NT) return;
    S* A = (S*)sd.W[0];
    if (sd.srcMode == 1) {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q < 343)
            point_source_at<S>(A, (S*)sd.W[1], (S*)sd.W[2], sd.lay, sd.SX, sd.SY, sd.SZ, sd.coef,
                               ricker_value(sd.fc, sd.it, sd.DT), q);
    } else if (sd.srcMode == 2) {
        int srcIt;
        float tw;
        if (!table_time(sd.it, sd.DT, sd.sdt, sd.snt, srcIt, tw)) return;
        for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < sd.npts;
             p += (long long)gridDim.x * blockDim.x)
            table_source_at<S>(A, sd.lay, sd.srcIdx, sd.mrate, sd.npts, srcIt, tw, sd.DT, p);
    }
}

template <class S, int TX, int TY, int IRK>
__global__ void __launch_bounds__(TX* TY) k_bulk_multi(BatchDev b) {
    constexpr int irk = IRK;
    __shared__ BulkLds<S, TX, TY> lds;
    const int nEnt = b.tab->nBulk, nTiles = b.tab->nTiles;
    // the persistent grid is a multiple of the XCD count, so work index w keeps its hardware
    // block's XCD (w % kXcds) and xcd_order gives every XCD a contiguous run of tiles
    for (int w = blockIdx.x; w < nTiles; w += gridDim.x) {
        const int t = (int)xcd_order(w, nTiles);
        const BulkEnt& e = b.tab->bulk[find_entry(b.tab->bulk, nEnt, t)];
        const SlotDev& sd = b.slots[e.slot];
        if (sd.it >= sd.NT) continue;
        const StageParams& p = b.sp[e.slot * 4 + irk];
        const int r = t - e.start;
        const int bx = r % e.gx, by = (r / e.gx) % e.gy, ch = r / (e.gx * e.gy);
        const int kb = p.bk0 + ch * e.kc;
        const int ke = min(kb + e.kc, p.bk1);
        bulk_tile<S, TX, TY, IRK>(p, p.FB1, p.FB2, p.FB3, p.bi0, p.bi1, p.bj0, p.bj1, 0, bx, by, kb, ke, lds.v);
        __syncthreads();  // the next tile rewrites the LDS planes
    }
}

template <class S, int IRK, int MASK>
__global__ void __launch_bounds__(pml_tx(MASK) * pml_ty(MASK)) k_pml_multi(BatchDev b) {
    constexpr int TX = pml_tx(MASK), TY = pml_ty(MASK), g = pml_group_of(MASK);
    __shared__ BulkLds<S, TX, TY> lds;
    const int nEnt = b.tab->nPml[g], nTiles = b.tab->nPmlTiles[g];
    for (int w = blockIdx.x; w < nTiles; w += gridDim.x) {
        const int t = (int)xcd_order(w, nTiles);
        const PmlEnt& e = b.tab->pml[g][find_entry(b.tab->pml[g], nEnt, t)];
        const SlotDev& sd = b.slots[e.slot];
        if (sd.it >= sd.NT) continue;
        const StageParams& p = b.sp[e.slot * 4 + IRK];
        const int r = t - e.start;
        const int bx = r % e.gx, by = (r / e.gx) % e.gy, ch = r / (e.gx * e.gy);
        const int* B = p.pbox[e.box];
        const int kb = B[4] + ch * e.kc;
        const int ke = min(kb + e.kc, B[5]);
        bulk_tile<S, TX, TY, IRK, MASK>(p, p.FB1, p.FB2, p.FB3, B[0], B[1], B[2], B[3], B[7], bx, by, kb, ke, lds.v);
        __syncthreads();
    }
}

template <class S, int IRK>
__global__ void __launch_bounds__(kFsTx* kFsTy) k_fs_multi(BatchDev b) {
    __shared__ FsLds<kFsTx, kFsTy> lds;
    const int nEnt = b.tab->nFs, nTiles = b.tab->nFsTiles;
    for (int w = blockIdx.x; w < nTiles; w += gridDim.x) {
        const int t = (int)xcd_order(w, nTiles);
        const FsEnt& e = b.tab->fs[find_entry(b.tab->fs, nEnt, t)];
        const SlotDev& sd = b.slots[e.slot];
        if (sd.it >= sd.NT) continue;
        const StageParams& p = b.sp[e.slot * 4 + IRK];
        const int r = t - e.start;
        fs_tile<S, kFsTx, kFsTy>(p, r % e.gx, r / e.gx, lds);
        __syncthreads();  // the next tile rewrites the LDS plane
    }
}

template <class S>
__global__ void __launch_bounds__(kGenChunk) k_general_multi(BatchDev b, int irk) {
    const int nEnt = b.tab->nGen;
    const long long nChunks = b.tab->nChunks;
    for (long long w = blockIdx.x; w < nChunks; w += gridDim.x) {
        const long long c = xcd_order(w, nChunks);
        const GenEnt& e = b.tab->gen[find_entry(b.tab->gen, nEnt, c)];
        const SlotDev& sd = b.slots[e.slot];
        if (sd.it >= sd.NT) continue;
        const StageParams& p = b.sp[e.slot * 4 + irk];
        const int* B = p.box[e.box];
        const long long r = (c - e.start) * kGenChunk + threadIdx.x;
        const int nxb = B[1] - B[0], nyb = B[3] - B[2];
        const long long n = (long long)nxb * nyb * (B[5] - B[4]);
        if (r < n) {
            const int i = B[0] + (int)(r % nxb);
            const int j = B[2] + (int)((r / nxb) % nyb);
            const int k = B[4] + (int)(r / ((long long)nxb * nyb));
            stage_point_general<S>(p, i, j, k);
        }
    }
}

template <class S>
__global__ void k_record_multi(BatchDev b) {
    const SlotDev& sd = b.slots[blockIdx.y];
    if (!sd.active || sd.it >= sd.NT) return;
    const S* A = (const S*)sd.W[0];
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int t = t0; t < sd.nst * 9; t += stride) station_at<S>(A, sd.lay, sd.stIdx, sd.stWave, sd.NT, sd.it, sd.rCv, sd.rCs, t);
    if (sd.fs)
        for (int t = t0; t < sd.nx * sd.ny; t += stride) pgv_at<S>(A, sd.lay, sd.nx, sd.ktop, sd.pgv, sd.rCv, t);
}

// One thread per slot: the step counter and the MacCormack signs of the next step.
__global__ void k_advance(BatchDev b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= b.maxSlots) return;
    SlotDev& sd = b.slots[s];
    if (!sd.active || sd.it >= sd.NT) return;
    const int it = sd.it + 1;
    sd.it = it;
    for (int irk = 0; irk < 4; ++irk) set_fb(b.sp[s * 4 + irk], it, irk);
}

// ---------------------------------------------------------------- launchers
int batch_write_slot(const BatchDev& b, int s, const SlotDev& sd, const StageParams& base, hipStream_t st) {
    k_write_slot<<<1, 64, 0, st>>>(b, s, sd, base);
    k_build_tables<<<1, 64, 0, st>>>(b);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int batch_clear_slot(const BatchDev& b, int s, hipStream_t st) {
    k_clear_slot<<<1, 64, 0, st>>>(b, s);
    k_build_tables<<<1, 64, 0, st>>>(b);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// One RK stage of every resident domain: the bulk tiles on st; the PML, free-surface and general
// work on the side stream(s) of fk (if given) concurrently, joined back into st. With three side
// streams: the x faces on the first, the y and bottom faces on the second, edges, corners, the
// free-surface layers and the general boxes on the third, so the boundary kernels of small
// domains (a large share of their points) overlap each other instead of running one by one.
template <class S, int TX, int TY, int IRK>
static int batch_stage(const BatchDev& b, const BatchGrid& g, hipStream_t st, const StageFork* fk) {
    hipStream_t s0 = st, s1 = st, s2 = st;
    const bool three = fk && fk->side && fk->side2[0] && fk->side2[1];
    if (fk && fk->side) {
        if (hipEventRecord(fk->fork, st) != hipSuccess || hipStreamWaitEvent(fk->side, fk->fork, 0) != hipSuccess)
            return -3;
        s0 = s1 = s2 = fk->side;
        if (three) {
            for (int q = 0; q < 2; ++q)
                if (hipStreamWaitEvent(fk->side2[q], fk->fork, 0) != hipSuccess) return -3;
            s1 = fk->side2[0];
            s2 = fk->side2[1];
        }
    }
    k_bulk_multi<S, TX, TY, IRK><<<g.bulkBlocks, TX * TY, 0, st>>>(b);
    k_pml_multi<S, IRK, 1><<<g.pmlBlocks[0], pml_tx(1) * pml_ty(1), 0, s0>>>(b);
    k_pml_multi<S, IRK, 2><<<g.pmlBlocks[1], pml_tx(2) * pml_ty(2), 0, s1>>>(b);
    k_pml_multi<S, IRK, 4><<<g.pmlBlocks[2], pml_tx(4) * pml_ty(4), 0, s1>>>(b);
    k_pml_multi<S, IRK, 3><<<g.pmlBlocks[3], pml_tx(3) * pml_ty(3), 0, s2>>>(b);
    k_pml_multi<S, IRK, 5><<<g.pmlBlocks[4], pml_tx(5) * pml_ty(5), 0, s2>>>(b);
    k_pml_multi<S, IRK, 6><<<g.pmlBlocks[5], pml_tx(6) * pml_ty(6), 0, s2>>>(b);
    k_pml_multi<S, IRK, 7><<<g.pmlBlocks[6], pml_tx(7) * pml_ty(7), 0, s2>>>(b);
    k_fs_multi<S, IRK><<<g.fsBlocks, kFsTx * kFsTy, 0, s2>>>(b);
    k_general_multi<S><<<g.genBlocks, kGenChunk, 0, s2>>>(b, IRK);
    if (s0 != st) {
        if (hipEventRecord(fk->join, s0) != hipSuccess || hipStreamWaitEvent(st, fk->join, 0) != hipSuccess)
            return -3;
    }
    if (three) {
        const hipStream_t sq[2] = {s1, s2};
        for (int q = 0; q < 2; ++q)
            if (hipEventRecord(fk->join2[q], sq[q]) != hipSuccess || hipStreamWaitEvent(st, fk->join2[q], 0) != hipSuccess)
                return -3;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <class S, int TX, int TY>
static int batch_step_t(const BatchDev& b, const BatchGrid& g, hipStream_t st, const StageFork* fk) {
    k_source_multi<S><<<dim3(2, b.maxSlots), 256, 0, st>>>(b);
    if (batch_stage<S, TX, TY, 0>(b, g, st, fk) || batch_stage<S, TX, TY, 1>(b, g, st, fk) ||
        batch_stage<S, TX, TY, 2>(b, g, st, fk) || batch_stage<S, TX, TY, 3>(b, g, st, fk))
        return -3;
    k_record_multi<S><<<dim3(16, b.maxSlots), 256, 0, st>>>(b);
    k_advance<<<(b.maxSlots + 63) / 64, 64, 0, st>>>(b);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int batch_step(int prec, int cfg, const BatchDev& b, const BatchGrid& g, hipStream_t st, const StageFork* fk) {
    switch (prec * 2 + (cfg ? 1 : 0)) {
        case P_F32 * 2: return batch_step_t<float, 64, 8>(b, g, st, fk);
        case P_F32 * 2 + 1: return batch_step_t<float, 64, 4>(b, g, st, fk);
        case P_F16 * 2: return batch_step_t<_Float16, 64, 8>(b, g, st, fk);
        case P_F16 * 2 + 1: return batch_step_t<_Float16, 64, 4>(b, g, st, fk);
        case P_BF16 * 2: return batch_step_t<__bf16, 64, 8>(b, g, st, fk);
        case P_BF16 * 2 + 1: return batch_step_t<__bf16, 64, 4>(b, g, st, fk);
    }
    return -1;
}

// Persistent grid sizes: as many workgroups as fit on the device at once.
int batch_grid(int prec, int cfg, BatchGrid& g) {
    int dev = 0, ncu = 0, nb = 0, ng = 0, nf = 0;
    int npm[kNumPmlMasks] = {};
    if (hipGetDevice(&dev) != hipSuccess) return -3;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -3;
    hipError_t e1 = hipSuccess, e2 = hipSuccess;
    const int bt = cfg ? 64 * 4 : 64 * 8;
#define CG_OCC(S)                                                                                                 \
    e1 = cfg ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bulk_multi<S, 64, 4, 1>, bt, 0)               \
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bulk_multi<S, 64, 8, 1>, bt, 0);              \
    e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&ng, k_general_multi<S>, kGenChunk, 0);                  \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[0], k_pml_multi<S, 1, 1>, pml_tx(1) * pml_ty(1), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[1], k_pml_multi<S, 1, 2>, pml_tx(2) * pml_ty(2), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[2], k_pml_multi<S, 1, 4>, pml_tx(4) * pml_ty(4), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[3], k_pml_multi<S, 1, 3>, pml_tx(3) * pml_ty(3), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[4], k_pml_multi<S, 1, 5>, pml_tx(5) * pml_ty(5), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[5], k_pml_multi<S, 1, 6>, pml_tx(6) * pml_ty(6), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&npm[6], k_pml_multi<S, 1, 7>, pml_tx(7) * pml_ty(7), 0); \
    if (e2 == hipSuccess) e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nf, k_fs_multi<S, 1>, kFsTx * kFsTy, 0);
    switch (prec) {
        case P_F32: CG_OCC(float) break;
        case P_F16: CG_OCC(_Float16) break;
        case P_BF16: CG_OCC(__bf16) break;
        default: return -1;
    }
#undef CG_OCC
    if (e1 != hipSuccess || e2 != hipSuccess) return -3;
    // multiples of the XCD count (see k_bulk_multi)
    g.bulkBlocks = (ncu * (nb > 0 ? nb : 1) + kXcds - 1) / kXcds * kXcds;
    g.genBlocks = (ncu * (ng > 0 ? ng : 1) + kXcds - 1) / kXcds * kXcds;
    for (int a = 0; a < kNumPmlMasks; ++a) g.pmlBlocks[a] = (ncu * (npm[a] > 0 ? npm[a] : 1) + kXcds - 1) / kXcds * kXcds;
    g.fsBlocks = (ncu * (nf > 0 ? nf : 1) + kXcds - 1) / kXcds * kXcds;
    return 0;
}

}  // namespace cg
