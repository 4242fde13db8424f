// domain.hpp — the state behind a cgfdm_domain handle, shared by the C ABI (api.hip) and the
// serving engine (server.hip).
//
// A domain runs in one of two modes:
//  - standalone (cgfdm_domain_create): its own stream, one hipMalloc per buffer, and every
//    set-up call waits for its device work before returning;
//  - pooled (created by a server): device buffers are carved from one block of the server's
//    arena, host-to-device copies go through a block of the server's pinned memory, the stream is
//    the server's, and set-up calls only enqueue work (no host synchronisation), so admitting a
//    request never stalls the simulations already running.
#pragma once
#include <string>
#include <vector>

#include "../../include/cgfdm.h"
#include "common.hpp"

namespace cg {

void set_err(const char* fmt, ...);
extern thread_local std::string g_err;

// Bump allocator over a fixed block (device or pinned host memory); 256-byte aligned pieces.
struct Bump {
    char* base = nullptr;
    size_t size = 0, used = 0;
    void* take(size_t n) {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + n > size) return nullptr;
        used = off + n;
        return base + off;
    }
};

inline size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace cg

struct cgfdm_domain {
    cgfdm_desc d{};
    int dev = 0;
    hipStream_t stream = nullptr;
    bool ownStream = true;
    // side stream of the boundary kernels and its fork/join events (own-stream domains)
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    bool async = false;       // pooled mode: no host synchronisation in set-up calls
    cg::Bump devPool, hostPool;
    bool pooled = false;
    int prec = cg::P_F32;
    int sb = 4;  // storage bytes
    int nx = 0, ny = 0, nz = 0, nx_ = 0, ny_ = 0, nz_ = 0;
    cg::Layout lay{};
    double Cv = 1, Cs = 1, Crho = 1;
    void* W[4] = {nullptr, nullptr, nullptr, nullptr};  // A, B, C, T
    void* cjm = nullptr;
    void* m0 = nullptr;     // 16 bytes: whether (xi_x, eta_y) is uniform, and its value
    bool pair0Stream = false;  // read the first metric pair per point even when uniform
    float* prof = nullptr;  // [ax bx dx][ay by dy][az bz dz]
    float* rdz = nullptr;
    float* auxmem = nullptr;
    cg::Slab slab[6]{};
    float* pgv = nullptr;
    // stations (padded i, j, k triples)
    int* stIdx = nullptr;
    float* stWave = nullptr;
    int nst = 0;
    // sources
    int srcMode = 0;
    int SX = 0, SY = 0, SZ = 0;
    float fc = 0.f;
    float* srcCoef = nullptr;
    int* srcIdx = nullptr;  // padded i, j, k triples
    float* mrate = nullptr;
    long long npts = 0;
    int snt = 0;
    float sdt = 0.f;
    int it = 0;
    bool modelSet = false;
    int bulk_cfg = -1, bulk_kc = 0;  // launch shape of the bulk stage kernel (-1: tuned)
    long long bytes = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> evStart, evStop;
    size_t evUsed = 0;
    double tBulk = 0.0, tGen = 0.0;  // summed kernel times (ms)
    long long tCount = 0;
    std::vector<void*> allocs;

    // Zeroed device buffer of n bytes.
    int alloc(void** p, size_t n);
    // Frees a buffer from alloc (standalone mode; pooled buffers live as long as the block).
    void release(void* p);
    // Enqueues a host-to-device copy; in pooled mode the source is first copied to pinned memory.
    int h2d(void* dst, const void* src, size_t n);
    // Waits for the stream unless pooled.
    int settle();
};

namespace cg {

// Set-up entry points shared by the C ABI and the server.
int domain_init(cgfdm_domain* h, const cgfdm_desc* desc, int device);
int domain_set_model(cgfdm_domain* h, const double* terrain, const MediumSrc& host);
int domain_bulk_tiling(const cgfdm_domain* h, int cfg, int& gx, int& gy, int& nch, int& kc);
StageParams domain_stage_params(const cgfdm_domain* h);
long long domain_estimate(const cgfdm_desc& d);
int validate_desc(const cgfdm_desc* d);

}  // namespace cg
