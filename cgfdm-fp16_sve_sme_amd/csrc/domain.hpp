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
#include <algorithm>
#include <cmath>
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
    hipStream_t side2[2] = {nullptr, nullptr};  // two more side streams for the boundary kernels
    hipEvent_t evJoin2[2] = {nullptr, nullptr};
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
    float* zcol = nullptr;  // [3][ny_][nx_] per-column zeta (k_zcol)
    bool pair0Stream = false;  // read the first metric pair per point even when uniform
    float* prof = nullptr;  // [ax bx dx][ay by dy][az bz dz]
    float* rdz = nullptr;
    char* auxmem = nullptr;  // PML auxiliaries of every slab (storage type, Aux*DT)
    cg::Slab slab[6]{};
    float* pgv = nullptr;
    // non-finite guard: every checkEvery steps (0: never) the register A is scanned; *nfFlag holds
    // the step count after which the first non-finite value was found (0: none)
    int* nfFlag = nullptr;
    int checkEvery = 100;
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

// Reference scale factors of 16-bit storage (macro.h:65-73): Cv/Cs = 5e9 fixes the stored medium
// (rho*Cs/Cv), kept by every choice below.
constexpr double kRefCv = 5.0e3, kRefCs = 1.0e-6, kRefCrho = 1.0e6;
// Targets of the scaled peak particle velocity and stress that source_scales aims at. The stored
// state is not the largest stored number: next to a single-node moment source the RK stage values
// of the velocity (W = m + alpha h) overshoot the state's peak, by up to ~4.5x for a slow source
// (Ricker fc 0.5 Hz, Vs 4000 m/s, DT 0.008 s, DH 100 m: the FP16 oracle goes non-finite once the
// stored peak passes ~1.4e4), so the velocity target leaves 16x below the FP16 maximum (65504).
constexpr double kVelTarget = 4.0e3, kStressTarget = 1.0e4;
// Peak velocity one cell from a single node of moment rate M' in the discrete scheme, in units of
// the continuum's M' / (4 pi mu DH^2): 1.7 to 2.3 measured with the FP32 oracle over 19 double
// couples (fc 0.6 to 3 Hz, Vs 3400 m/s); the rule takes 2.5.
constexpr double kNodeField = 2.5;

// Scale factors (Cv, Cs) of 16-bit storage chosen from the source. Without a table (rate ==
// null): the reference's (5e3, 1e-6) times (DH/100)^2 (point_source_scales below estimates the
// point source from its wavelet and the medium instead). Moment-rate
// table (nt x npts x 6, tensor * area * slip rate, as cgfdm_set_source_table takes it, at grid
// nodes xyz): the peak particle velocity is estimated as the largest, over nodes p and radii
// R = DH, 2 DH, ..., 16 DH, of 2 * sum_{|q - p| <= R} max_t|rate_q| / (4 pi R^2), except at
// R = DH: (kNodeField * max_t|rate_p| + sum_{0 < |q - p| <= DH} max_t|rate_q|) / (4 pi DH^2).
// That is the near field of one node's moment rate M' = mu * rate one cell away,
// M'/(4 pi mu DH^2) (mu cancels), with the discrete scheme's excess for the node itself; at
// larger R it is the field of a patch of nodes slipping together, which tends to the slip
// rate / 2 of a planar fault (each node holding area * slip rate). The peak stress is
// estimated from the static near field, mu_max * max_t|sum rate dt| / (4 pi DH^3). Cv is the
// larger value that keeps both scaled estimates below their targets, Cs = Cv / 5e9.
inline void source_scales(double DH, const int* xyz, const float* rate, long long npts, int nt, double dt,
                          double mu_max, double& Cv, double& Cs) {
    const double ratio = kRefCv / kRefCs;
    if (!rate) {
        Cv = kRefCv * (DH / 100.0) * (DH / 100.0);
        Cs = Cv / ratio;
        return;
    }
    std::vector<double> peak((size_t)npts, 0.0);  // max over time and components of |rate| per node
    double mmax = 0.0;
    for (long long p = 0; p < npts; ++p)
        for (int c = 0; c < 6; ++c) {
            double acc = 0.0;
            for (int t = 0; t < nt; ++t) {
                const double r = rate[((long long)t * npts + p) * 6 + c];
                peak[p] = std::max(peak[p], std::fabs(r));
                acc += r * dt;
                mmax = std::max(mmax, std::fabs(acc));
            }
        }
    const double pi4 = 4.0 * 3.14159265358979323846;
    double v = 0.0;
    // patch sums around every node (every stride-th node as a centre above 20000 nodes)
    const long long stride = npts > 20000 ? (npts + 19999) / 20000 : 1;
    for (long long p = 0; p < npts; p += stride) {
        double sum[5] = {0, 0, 0, 0, 0};
        for (long long q = 0; q < npts; ++q) {
            const double dx = xyz[3 * q] - xyz[3 * p], dy = xyz[3 * q + 1] - xyz[3 * p + 1],
                         dz = xyz[3 * q + 2] - xyz[3 * p + 2];
            const double d2 = dx * dx + dy * dy + dz * dz;  // in cells squared
            for (int r = 0; r < 5; ++r)
                if (d2 <= (double)(1 << r) * (1 << r)) sum[r] += peak[q];
        }
        // R = DH: the node's own rate with the discrete excess, its neighbours' without (a
        // smooth distribution of nodes follows the continuum)
        v = std::max(v, (sum[0] + (kNodeField - 1.0) * peak[p]) / (pi4 * DH * DH));
        for (int r = 1; r < 5; ++r) {
            const double R = DH * (1 << r);
            v = std::max(v, 2.0 * sum[r] / (pi4 * R * R));
        }
    }
    const double T = mu_max * mmax / (pi4 * DH * DH * DH);
    Cv = kRefCv;
    if (v > 0.0 || T > 0.0) {
        Cv = v > 0.0 ? kVelTarget / v : 1e300;
        if (T > 0.0) Cv = std::min(Cv, ratio * kStressTarget / T);
    }
    Cs = Cv / ratio;
}

// Scale factors of 16-bit storage for the explosive point source (singleSource.cpp:29-127: the
// Ricker moment rate s(t) = M0 2 pi fc r (2 r^2 - 3) exp(-r^2), r = pi fc (t - 1.2/fc), M0 =
// 1e16 N m, spread over the 7^3 Gaussian weights GAUSS_FUN(d, 1.5, 0) / 0.998125703461425 into
// Txx, Tyy and Tzz): the rule of source_scales applied to the equivalent table, one node per
// stencil point holding rate = s(t) * weight / mu_src on the three normal components, sampled at
// fc/40 over the wavelet. mu_src is the shear modulus at the source (the smallest of the medium
// where that is not known: a softer medium moves faster for the same moment rate), mu_max the
// largest of the medium.
inline void point_source_scales(double DH, double fc, double mu_src, double mu_max, double& Cv, double& Cs) {
    const double pi = 3.14159265358979323846, M0 = 1.0e16;
    const double dt = 1.0 / (40.0 * fc);
    const int nt = (int)std::ceil(2.4 / fc / dt) + 1;
    double w[7];
    for (int a = 0; a < 7; ++a) w[a] = std::exp(-0.5 * (a - 3) * (a - 3) / (1.5 * 1.5));
    double wsum = 0.0;
    for (int a = 0; a < 7; ++a)
        for (int b = 0; b < 7; ++b)
            for (int c = 0; c < 7; ++c) wsum += w[a] * w[b] * w[c];
    std::vector<int> xyz(343 * 3);
    std::vector<float> rate((size_t)nt * 343 * 6, 0.0f);
    for (int q = 0; q < 343; ++q) {
        const int a = q % 7, b = (q / 7) % 7, c = q / 49;
        xyz[3 * q] = a - 3; xyz[3 * q + 1] = b - 3; xyz[3 * q + 2] = c - 3;
        const double wq = w[a] * w[b] * w[c] / wsum;
        for (int t = 0; t < nt; ++t) {
            const double r = pi * fc * (t * dt - 1.2 / fc);
            const double sv = M0 * 2.0 * pi * fc * r * (2.0 * r * r - 3.0) * std::exp(-r * r);
            for (int comp = 0; comp < 3; ++comp) rate[((size_t)t * 343 + q) * 6 + comp] = (float)(sv * wq / mu_src);
        }
    }
    source_scales(DH, xyz.data(), rate.data(), 343, nt, dt, mu_max, Cv, Cs);
}

// Set-up entry points shared by the C ABI and the server.
int domain_init(cgfdm_domain* h, const cgfdm_desc* desc, int device);
int domain_set_model(cgfdm_domain* h, const double* terrain, const MediumSrc& host);
int domain_bulk_tiling(const cgfdm_domain* h, int cfg, int& gx, int& gy, int& nch, int& kc);

// Slab geometry and auxiliary registers of the six PML slabs (x1 y1 z1 x2 y2 z2), carved in that
// order from h->auxmem (three registers of 9 variables each, h->sb bytes per value). h->d,
// h->nx..nz, h->sb and h->auxmem must be set; the slab sizes are slab_points.
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
            X.m = h->auxmem + off + 9 * X.n * h->sb;
            X.t = h->auxmem + off + 18 * X.n * h->sb;
            off += 27 * X.n * h->sb;
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
    p.zcol = h->zcol;
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
