// batch.hpp — device-side tables of the multi-domain executor (batch.hip) and its launchers.
#pragma once
#include "common.hpp"

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
    int nbox;
    long long boxChunks[6];
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
struct GenEnt { int slot, box; long long start; };

struct BatchTables {
    int nBulk, nTiles, nGen, pad;
    long long nChunks;
    BulkEnt bulk[kMaxSlots];
    GenEnt gen[kMaxSlots * 6];
};

// Device pointers of one batch (passed by value to every batch kernel).
struct BatchDev {
    SlotDev* slots;     // [maxSlots]
    StageParams* sp;    // [maxSlots][4], one per RK stage
    BatchTables* tab;
    int maxSlots;
};

struct BatchGrid { int bulkBlocks, genBlocks; };

void stage_boxes(StageParams& q, int use_fast);
int batch_write_slot(const BatchDev& b, int s, const SlotDev& sd, const StageParams& base, hipStream_t st);
int batch_clear_slot(const BatchDev& b, int s, hipStream_t st);
int batch_step(int prec, int cfg, const BatchDev& b, const BatchGrid& g, hipStream_t st);
int batch_grid(int prec, int cfg, BatchGrid& g);

}  // namespace cg
