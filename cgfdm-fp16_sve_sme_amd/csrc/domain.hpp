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
#include "stage_geom.hpp"

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
    float* auxmem = nullptr;  // PML auxiliaries of every slab (FP32)
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

// Slab geometry and auxiliary registers of the six PML slabs (x1 y1 z1 x2 y2 z2), carved in that
// order from h->auxmem (three FP32 registers of 9 variables each). h->d, h->nx..nz and h->auxmem
// must be set; the slab sizes are slab_points.
inline void assign_slabs(cgfdm_domain* h) {
    long long off = 0;
    for (int s = 0; s < 6; ++s) {
        Slab& X = h->slab[s];
        X.n = slab_points(h->d, s);
        X.on = X.n > 0;
        X.axis = s % 3;
        const int nAxis = X.axis == 0 ? h->nx : (X.axis == 1 ? h->ny : h->nz);
        X.lo = s < 3 ? HALO : HALO + nAxis - h->d.nPML;
        X.n0 = X.axis == 0 ? h->d.nPML : h->nx;
        X.n1 = X.axis == 1 ? h->d.nPML : h->ny;
        if (X.on) {
            X.a = h->auxmem + off;
            X.m = X.a + 9 * X.n;
            X.t = X.m + 9 * X.n;
            off += 27 * X.n;
        }
    }
}

// Wavefield registers of RK stage irk: input and output (A -> B -> C -> B -> A).
inline void* stage_in(const cgfdm_domain* h, int irk) { return h->W[irk == 0 ? 0 : (irk == 2 ? 2 : 1)]; }
inline void* stage_out(const cgfdm_domain* h, int irk) {
    return h->W[irk == 0 ? 1 : (irk == 1 ? 2 : (irk == 2 ? 1 : 0))];
}

// Parameters of a stage of domain h (stage registers, FB signs and irk are set per stage).
inline StageParams domain_stage_params(const cgfdm_domain* h) {
    StageParams p{};
    p.A = h->W[0];
    p.T = h->W[3];
    p.cjm = h->cjm;
    p.m0 = h->pair0Stream ? nullptr : h->m0;
    const int offs[3] = {0, 3 * h->nx_, 3 * (h->nx_ + h->ny_)};
    const int lens[3] = {h->nx_, h->ny_, h->nz_};
    for (int a = 0; a < 3; ++a) {
        p.pa[a] = h->prof + offs[a];
        p.pb[a] = h->prof + offs[a] + lens[a];
        p.pd[a] = h->prof + offs[a] + 2 * lens[a];
    }
    p.rdz = h->rdz;
    for (int s = 0; s < 6; ++s) p.slab[s] = h->slab[s];
    p.lay = h->lay;
    p.nx = h->nx; p.ny = h->ny; p.nz = h->nz;
    p.nx_ = h->nx_; p.ny_ = h->ny_;
    p.nPML = h->d.nPML;
    p.fs = h->d.free_surface;
    p.pml = h->d.pml;
    p.bulk_cfg = h->bulk_cfg < 0 ? 1 : h->bulk_cfg;
    p.bulk_kc = h->bulk_kc;
    p.rDH = (float)(1.0 / h->d.DH);
    p.DT = (float)h->d.DT;
    p.Crho = (float)h->Crho;
    return p;
}
long long domain_estimate(const cgfdm_desc& d);
int validate_desc(const cgfdm_desc* d);

}  // namespace cg
