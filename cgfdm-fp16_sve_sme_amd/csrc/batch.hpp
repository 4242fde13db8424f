// batch.hpp — device-side tables of the multi-domain executor (batch.hip) and its launchers.
#pragma once
#include "common.hpp"
#include "stage_bulk.hpp"
#include "stage_fs.hpp"
#include "stage_geom.hpp"

namespace cg {

constexpr int kMaxSlots = 64;    // resident domains per batch
constexpr int kGenChunk = 256;   // points per workgroup iteration of the general kernel

// Everything the per-step kernels need to know about one resident domain.
struct SlotDev {
    int active;         // 1 while the slot holds a domain
    int it, NT;         // step counter (advanced on the device) and its end
    int fs;
    // bulk tiling: gx x gy column tiles, nch chunks of kc planes; general: chunks per box
    int gx, gy, nch, kc;
    int fgx, fgy;       // free-surface tiles over the bulk columns (0: no free surface)
    PmlTiles pt;        // PML boxes by kind
    int nbox;
    long long boxChunks[kMaxGenBoxes];
    Layout lay;
    void* W[4];         // A, B, C, T
    float DT;
    // sources
    int srcMode;        // 0 none, 1 point, 2 table
    int SX, SY, SZ;     // padded indices of the point source
    float fc;
    const float* coef;
    const int* srcIdx;
    const float* mrate;
    long long npts;
    int snt;
    float sdt;
    // stations and PGV
    const int* stIdx;
    float* stWave;
    int nst;
    float* pgv;
    int nx, ny, ktop;
    float rCv, rCs;
};

struct BulkEnt { int slot, gx, gy, nch, kc, start; };
struct PmlEnt { int slot, box, gx, gy, nch, kc, start; };
struct GenEnt { int slot, box; long long start; };
struct FsEnt { int slot, gx, gy, start; };

struct BatchTables {
    int nBulk, nTiles, nGen, pad;
    long long nChunks;
    int nPml[kNumPmlMasks], nPmlTiles[kNumPmlMasks];
    BulkEnt bulk[kMaxSlots];
    PmlEnt pml[kNumPmlMasks][kMaxSlots * 8];  // at most 8 boxes of one kind per domain
    GenEnt gen[kMaxSlots * kMaxGenBoxes];
    int nFs, nFsTiles;
    FsEnt fs[kMaxSlots];
};

// Device pointers of one batch (passed by value to every batch kernel).
struct BatchDev {
    SlotDev* slots;     // [maxSlots]
    StageParams* sp;    // [maxSlots][4], one per RK stage
    BatchTables* tab;
    int maxSlots;
};

struct BatchGrid { int bulkBlocks, genBlocks, pmlBlocks[kNumPmlMasks], fsBlocks; };

int batch_write_slot(const BatchDev& b, int s, const SlotDev& sd, const StageParams& base, hipStream_t st);
int batch_clear_slot(const BatchDev& b, int s, hipStream_t st);
int batch_step(int prec, int cfg, const BatchDev& b, const BatchGrid& g, hipStream_t st, const StageFork* fk);
int batch_grid(int prec, int cfg, BatchGrid& g);

}  // namespace cg
