// api.hip — host implementation of the C ABI declared in include/cgfdm.h.
//
// One domain = one simulation on one GPU. Device memory layout: every 3-D array is k-major,
// [k][var][j][i] (Layout in common.hpp):
//   W buffers A, B, C and accumulator T : 9 variables, storage type
//   CJM                                  : 13 components, storage type
//   PML slabs                            : 3 registers x 9 x slab points, FP32
//   rdz, profiles, stations, PGV         : FP32
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "domain.hpp"
#include "stage_geom.hpp"

using namespace cg;

namespace cg {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

}  // namespace cg

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            set_err("%s failed: %s", #x, hipGetErrorString(e_));                     \
            return CGFDM_EHIP;                                                      \
        }                                                                           \
    } while (0)

#define LCHK(x)                                                                     \
    do {                                                                            \
        int r_ = (x);                                                               \
        if (r_ != 0) {                                                              \
            set_err("kernel launch failed in %s: %s", #x, hipGetErrorString(hipGetLastError())); \
            return CGFDM_EHIP;                                                      \
        }                                                                           \
    } while (0)

namespace {

const int FB_TABLE[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                            {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};  // propagate.cpp:107-117

// Staging plane of cgfdm_get_slice_noalloc: one file-scope device array per GPU, 16 MiB (any
// plane of at most 4 Mi points, e.g. 2048 x 2048), held by every GPU the library runs on until the
// process ends. In host code `scratch` names a host-side stand-in, not a device address: its
// device address is looked up once per device with hipGetSymbolAddress (slice_scratch) and only
// that address is passed to kernels and copies.
constexpr long long kSliceScratchFloats = 4LL << 20;
__device__ __align__(256) float scratch[kSliceScratchFloats];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;               // guards the table and serialises the calls that use it
float* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device

// Device address of `scratch` on the current device `dev` (looked up on the first call there).
// Called with g_scratch_mu held.
hipError_t slice_scratch(int dev, float** out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[dev]) {
        float* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

// ------------------------------------------------------------------ domain memory
int cgfdm_domain::alloc(void** p, size_t n) {
    *p = nullptr;
    if (n == 0) return 0;
    if (pooled) {
        *p = devPool.take(n);
        if (!*p) {
            set_err("domain block exhausted (%zu bytes more requested, %zu of %zu used)", n, devPool.used, devPool.size);
            return CGFDM_ENOMEM;
        }
    } else {
        hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_err("hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
            return CGFDM_ENOMEM;
        }
        allocs.push_back(*p);
    }
    bytes += (long long)n;
    hipError_t e = hipMemsetAsync(*p, 0, n, stream);
    if (e != hipSuccess) { set_err("hipMemsetAsync failed: %s", hipGetErrorString(e)); return CGFDM_EHIP; }
    return 0;
}

void cgfdm_domain::release(void* p) {
    if (!p || pooled) return;
    for (size_t q = 0; q < allocs.size(); ++q)
        if (allocs[q] == p) { allocs.erase(allocs.begin() + q); break; }
    (void)hipFree(p);
}

int cgfdm_domain::h2d(void* dst, const void* src, size_t n) {
    if (n == 0) return 0;
    const void* from = src;
    if (pooled) {
        void* st = hostPool.take(n);
        if (!st) { set_err("pinned staging block exhausted (%zu bytes)", n); return CGFDM_ENOMEM; }
        std::memcpy(st, src, n);
        from = st;
    }
    HIPCHK(hipMemcpyAsync(dst, from, n, hipMemcpyHostToDevice, stream));
    return 0;
}

int cgfdm_domain::settle() {
    if (async) return 0;
    HIPCHK(hipStreamSynchronize(stream));
    return 0;
}

namespace {

int validate(const cgfdm_desc* d) {
    if (!d) { set_err("null descriptor"); return CGFDM_EINVAL; }
    if (d->NX < 1 || d->NY < 1 || d->NZ < 4) { set_err("grid too small (NX,NY >= 1, NZ >= 4)"); return CGFDM_EINVAL; }
    if (d->precision != CGFDM_FP32 && d->precision != CGFDM_FP16 && d->precision != CGFDM_BF16) {
        set_err("unknown precision %d", d->precision);
        return CGFDM_EINVAL;
    }
    if (!(d->DH > 0) || !(d->DT > 0) || d->NT < 0) { set_err("DH, DT must be > 0 and NT >= 0"); return CGFDM_EINVAL; }
    long long num = (long long)(d->NX + 6) * (d->NY + 6) * (d->NZ + 6);
    if (num >= (1LL << 31)) { set_err("padded grid exceeds 2^31 points"); return CGFDM_EINVAL; }
    if (d->pml) {
        // init_grid.cpp:108-118, pml_deriv.cpp:53-59
        if (d->nPML < 2 || d->nPML > d->NX || d->nPML > d->NY || d->nPML > d->NZ) {
            set_err("The PML layer(nPML=%d) must be >= 2 and not bigger than the grid", d->nPML);
            return CGFDM_EPML;
        }
    }
    if (d->use_terrain && !(d->depth != 0.0)) { set_err("terrain stretch needs a nonzero depth"); return CGFDM_EINVAL; }
    if (d->chi_mode != 0 && d->chi_mode != 1) { set_err("chi_mode must be 0 or 1"); return CGFDM_EINVAL; }
    return 0;
}

long long estimate(const cgfdm_desc& d) {
    const Layout L = make_layout(d);
    int sb = d.precision == CGFDM_FP32 ? 4 : 2;
    long long b = (L.size(9) * 4 + L.size(13)) * sb;
    for (int s = 0; s < 6; ++s) b += slab_points(d, s) * 9 * 3 * sb;
    b += (long long)(d.NX + 6) * (d.NY + 6) * 18 * 4;
    b += 3LL * (d.NX + d.NY + d.NZ + 18) * 4;
    b += (long long)d.NX * d.NY * 2 * 4;
    b += 256;  // uniform first metric pair
    b += (long long)(d.NX + 6) * (d.NY + 6) * 3 * 4;  // per-column zeta
    b += 256;  // non-finite flag
    return b;
}

// cpu_init_pml_parameter (init_pml_para.cpp:99-234) for one rank owning the whole domain.
void pml_profiles(const cgfdm_desc& d, std::vector<float>& out) {
    const int nn[3] = {d.NX + 6, d.NY + 6, d.NZ + 6};
    out.clear();
    const int nPML = d.nPML;
    const double DH = d.DH, L = DH * (nPML - 1);
    const double R = std::pow(10.0, -(std::log10((double)nPML) - 1.0) / std::log10(2.0) - 3.0);
    const double d0 = -(2.0 + 1.0) * 6000.0 / (2.0 * L) * std::log(R);
    const double alpha0 = (double)PIf * 0.5, beta0 = 3.0;
    for (int a = 0; a < 3; ++a) {
        const int n = nn[a], hi = n - HALO;
        std::vector<double> pa(n, 0.0), pb(n, 1.0), pd(n, 0.0), Ld(n, -1.0);
        if (d.pml) {
            for (int i = HALO; i < hi; ++i) {
                if (i < HALO + nPML) Ld[i] = (HALO + nPML - 1 - i) * DH;
                if (i >= hi - nPML && !(a == 2 && d.free_surface)) Ld[i] = (i - (hi - nPML)) * DH;
            }
            for (int i = HALO; i < hi; ++i) {
                const double l = Ld[i];
                pa[i] = l < 0.0 ? 0.0 : alpha0 * (1.0 - std::pow(l / L, 1.0));
                pb[i] = l < 0.0 ? 1.0 : 1.0 + (beta0 - 1.0) * std::pow(l / L, 2.0);
                pd[i] = l < 0.0 ? 0.0 : d0 * std::pow(l / L, 2.0);
            }
        }
        for (int i = 0; i < n; ++i) {
            pd[i] /= pb[i];
            pb[i] = 1.0 / pb[i];
        }
        for (double v : pa) out.push_back((float)v);
        for (double v : pb) out.push_back((float)v);
        for (double v : pd) out.push_back((float)v);
    }
}

SetupParams setup_params(const cgfdm_domain* h, const double* dterrain) {
    SetupParams q{};
    q.nx = h->nx; q.ny = h->ny; q.nz = h->nz;
    q.nx_ = h->nx_; q.ny_ = h->ny_; q.nz_ = h->nz_;
    q.lay = h->lay;
    q.DH = h->d.DH;
    q.use_terrain = h->d.use_terrain;
    q.depth = h->d.depth;
    q.centerX = h->d.centerX; q.centerY = h->d.centerY;
    q.terrain = dterrain;
    q.fs = h->d.free_surface;
    q.Cv = h->Cv; q.Cs = h->Cs; q.Crho = h->Crho;
    q.chi_mode = h->d.chi_mode;
    return q;
}


// ------------------------------------------------------------------ bulk launch-shape tuner
// Candidate shapes of the bulk kernel: tile (64x8 or 64x4 columns) and z-chunk length (0 picks
// enough chunks for ~2048 workgroups). The tuner times each on stage 0 of the domain's first
// step, which only reads A and the metric and only writes B and T, so repeating it is harmless;
// the choice is cached per (precision, grid, nPML, free surface) for the process.
struct TuneKey {
    int prec, nx, ny, nz, npml, fs;
    bool operator<(const TuneKey& o) const {
        return std::tie(prec, nx, ny, nz, npml, fs) < std::tie(o.prec, o.nx, o.ny, o.nz, o.npml, o.fs);
    }
};
std::mutex g_tuneMu;
std::map<TuneKey, std::pair<int, int>> g_tuned;
const int kTuneCands[][2] = {{0, 0}, {0, 64}, {1, 0}, {1, 64}, {2, 0}, {2, 64}, {3, 0}, {3, 64}};

int tune_bulk(cgfdm_domain* h, StageParams p) {
    const TuneKey key{h->prec, h->nx, h->ny, h->nz, h->d.nPML, h->d.free_surface};
    {
        std::lock_guard<std::mutex> g(g_tuneMu);
        auto f = g_tuned.find(key);
        if (f != g_tuned.end()) {
            h->bulk_cfg = f->second.first;
            h->bulk_kc = f->second.second;
            return 0;
        }
    }
    p.Win = stage_in(h, 0);
    p.Wout = stage_out(h, 0);
    p.irk = 0;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    int bc = 1, bk = 0;
    for (auto& c : kTuneCands) {
        p.bulk_cfg = c[0];
        p.bulk_kc = c[1];
        LCHK(launch_stage(h->prec, p, 2, h->stream, nullptr));  // warm-up
        HIPCHK(hipEventRecord(e0, h->stream));
        for (int r = 0; r < 2; ++r) LCHK(launch_stage(h->prec, p, 2, h->stream, nullptr));
        HIPCHK(hipEventRecord(e1, h->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; bc = c[0]; bk = c[1]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->bulk_cfg = bc;
    h->bulk_kc = bk;
    std::lock_guard<std::mutex> g(g_tuneMu);
    g_tuned[key] = {bc, bk};
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ shared set-up
namespace cg {

int validate_desc(const cgfdm_desc* d) { return validate(d); }
long long domain_estimate(const cgfdm_desc& d) { return estimate(d); }

int domain_init(cgfdm_domain* h, const cgfdm_desc* desc, int device) {
    h->d = *desc;
    h->dev = device;
    h->prec = desc->precision;
    h->sb = storage_bytes(h->prec);
    h->nx = desc->NX; h->ny = desc->NY; h->nz = desc->NZ;
    h->nx_ = h->nx + 6; h->ny_ = h->ny + 6; h->nz_ = h->nz + 6;
    h->lay = make_layout(*desc);
    const bool low = h->prec != P_F32;
    h->Cv = desc->Cv > 0 ? desc->Cv : (low ? 5.0e3 : 1.0);
    h->Cs = desc->Cs > 0 ? desc->Cs : (low ? 1.0e-6 : 1.0);
    h->Crho = desc->Crho > 0 ? desc->Crho : (low ? 1.0e6 : 1.0);
    if (const char* e = std::getenv("CGFDM_BULK_CFG")) h->bulk_cfg = std::atoi(e);
    if (const char* e = std::getenv("CGFDM_BULK_KC")) h->bulk_kc = std::atoi(e);
    if (const char* e = std::getenv("CGFDM_STREAM_PAIR0")) h->pair0Stream = std::atoi(e) != 0;
    int rc;
    for (int b = 0; b < 4; ++b)
        if ((rc = h->alloc(&h->W[b], (size_t)h->lay.size(9) * h->sb))) return rc;
    if ((rc = h->alloc(&h->cjm, (size_t)h->lay.size(13) * h->sb))) return rc;
    if ((rc = h->alloc(&h->m0, 16))) return rc;
    if ((rc = h->alloc((void**)&h->zcol, sizeof(float) * 3 * (size_t)h->nx_ * h->ny_))) return rc;
    // PML slabs: x1 y1 z1 x2 y2 z2
    long long auxTotal = 0;
    for (int s = 0; s < 6; ++s) auxTotal += slab_points(h->d, s);
    if (auxTotal > 0 && (rc = h->alloc((void**)&h->auxmem, (size_t)auxTotal * 9 * 3 * h->sb))) return rc;
    assign_slabs(h);
    if ((rc = h->alloc((void**)&h->prof, sizeof(float) * 3 * (h->nx_ + h->ny_ + h->nz_)))) return rc;
    if ((rc = h->alloc((void**)&h->rdz, sizeof(float) * 18 * (size_t)h->nx_ * h->ny_))) return rc;
    if ((rc = h->alloc((void**)&h->pgv, sizeof(float) * 2 * (size_t)h->nx * h->ny))) return rc;
    if ((rc = h->alloc((void**)&h->nfFlag, sizeof(int)))) return rc;
    std::vector<float> prof;
    pml_profiles(h->d, prof);
    if ((rc = h->h2d(h->prof, prof.data(), prof.size() * sizeof(float)))) return rc;
    return h->settle();
}

int domain_set_model(cgfdm_domain* h, const double* terrain, const MediumSrc& host) {
    if (h->d.use_terrain && !terrain) { set_err("use_terrain set but no terrain given"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long nin = (long long)h->nx * h->ny * h->nz;
    MediumSrc dev{};
    std::vector<void*> temps;
    auto cleanup = [&]() { for (void* p : temps) h->release(p); };
    auto temp = [&](void** p, size_t n) {
        int rc = h->alloc(p, n);
        if (rc == 0) { temps.push_back(*p); h->bytes -= (long long)n; }
        return rc;
    };
    int rc = 0;
    float vpmax_host = 0.f;
    if (host.vs) {
        for (long long q = 0; q < nin; ++q) {
            if (!(host.vs[q] >= 0.f) || !(host.vp[q] > 0.f) || !(host.rho[q] > 0.f)) {
                set_err("medium must be finite with vs >= 0, vp, rho > 0 (index %lld)", q);
                return CGFDM_EINVAL;
            }
            vpmax_host = std::max(vpmax_host, host.vp[q]);
        }
        const size_t mb = sizeof(float) * (size_t)nin;
        float *a = nullptr, *b = nullptr, *c = nullptr;
        if ((rc = temp((void**)&a, mb)) || (rc = temp((void**)&b, mb)) || (rc = temp((void**)&c, mb))) {
            cleanup();
            return rc;
        }
        if ((rc = h->h2d(a, host.vs, mb)) || (rc = h->h2d(b, host.vp, mb)) || (rc = h->h2d(c, host.rho, mb))) {
            cleanup();
            return rc;
        }
        dev.vs = a; dev.vp = b; dev.rho = c;
    } else {
        if (host.nl < 1 || !host.layers) { set_err("layered medium needs at least one layer"); return CGFDM_EINVAL; }
        for (int l = 0; l < host.nl; ++l) {
            const float* L = host.layers;
            if (!(L[host.nl + l] >= 0.f) || !(L[2 * host.nl + l] > 0.f) || !(L[3 * host.nl + l] > 0.f) ||
                (l > 0 && !(L[l] >= L[l - 1]))) {
                set_err("layer %d: depths must increase and vs >= 0, vp, rho > 0", l);
                return CGFDM_EINVAL;
            }
            vpmax_host = std::max(vpmax_host, L[2 * host.nl + l]);
        }
        vpmax_host = std::max(vpmax_host, 6000.0f);  // points above the first interface keep the default
        float* t = nullptr;
        const size_t n = sizeof(float) * 4 * host.nl;
        if ((rc = temp((void**)&t, n)) || (rc = h->h2d(t, host.layers, n))) { cleanup(); return rc; }
        dev.layers = t;
        dev.nl = host.nl;
    }
    double* dter = nullptr;
    if (h->d.use_terrain) {
        const size_t tb = sizeof(double) * (size_t)h->nx_ * h->ny_;
        if ((rc = temp((void**)&dter, tb)) || (rc = h->h2d(dter, terrain, tb))) { cleanup(); return rc; }
    }
    SetupParams q = setup_params(h, dter);
    LCHK(launch_setup_medium(h->prec, q, dev, h->cjm, h->stream));
    LCHK(launch_setup_metric(h->prec, q, h->cjm, h->zcol, h->m0, h->stream));
    LCHK(launch_pair0_check(h->prec, q, h->cjm, h->m0, h->stream));
    if (h->d.free_surface) LCHK(launch_setup_rdz(q, dev, h->rdz, h->stream));
    if (h->d.check_cfl) {
        if (h->async) { cleanup(); set_err("the CFL check needs a synchronous domain"); return CGFDM_EINVAL; }
        // calc_CFL.cpp:331-352: dtmax = 1.34*Hmin/Vpmax
        unsigned long long* dmin = nullptr;
        unsigned* dvp = nullptr;
        if ((rc = temp((void**)&dmin, sizeof(unsigned long long))) || (rc = temp((void**)&dvp, sizeof(unsigned)))) {
            cleanup();
            return rc;
        }
        double init = 1.0e20;
        HIPCHK(hipMemcpyAsync(dmin, &init, 8, hipMemcpyHostToDevice, h->stream));
        LCHK(launch_hmin(q, dmin, h->stream));
        LCHK(launch_vpmax(q, dev, dvp, h->stream));
        double hmin = 0;
        unsigned vbits = 0;
        HIPCHK(hipMemcpyAsync(&hmin, dmin, 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(&vbits, dvp, 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        float vpmax = 0.f;
        std::memcpy(&vpmax, &vbits, 4);
        const double dtmax = 1.34 * hmin / vpmax;
        if (h->d.DT > dtmax) {
            cleanup();
            set_err("DT = %.4e exceeds the CFL limit dtmax = %.4e (Hmin %.4e, Vpmax %.4e)", h->d.DT, dtmax, hmin, vpmax);
            return CGFDM_ECFL;
        }
    }
    if ((rc = h->settle())) return rc;
    cleanup();
    if (!h->async) {
        // the tile kernels take zeta from its per-column form (k_zcol), which every grid built
        // here has; a grid without it would be an error of this library's set-up
        int notLinear = 0;
        HIPCHK(hipMemcpy(&notLinear, (const char*)h->m0 + 4, sizeof(int), hipMemcpyDeviceToHost));
        if (notLinear) {
            set_err("the metric zeta is not (k - 3) times a per-column value on this grid");
            return CGFDM_EINVAL;
        }
    }
    h->modelSet = true;
    return CGFDM_OK;
}


}  // namespace cg

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* cgfdm_last_error(void) { return g_err.c_str(); }

long long cgfdm_estimate_bytes(const cgfdm_desc* desc) {
    if (validate(desc) != 0) return -1;
    return estimate(*desc);
}

int cgfdm_source_scales(double DH, const int* xyz, const float* rate, long long npts, int nt, float dt,
                        double mu_max, double* Cv, double* Cs) {
    if (!Cv || !Cs || !(DH > 0) || (rate && (!xyz || npts < 1 || nt < 1 || !(dt > 0) || !(mu_max > 0)))) {
        set_err("source_scales needs DH > 0 and, with a table, its nodes, npts, nt >= 1, dt > 0 and mu_max > 0");
        return CGFDM_EINVAL;
    }
    source_scales(DH, xyz, rate, npts, nt, dt, mu_max, *Cv, *Cs);
    if (!(*Cv > 0) || !std::isfinite(*Cv)) {
        set_err("source_scales: the table gives no finite scale");
        return CGFDM_EINVAL;
    }
    return CGFDM_OK;
}

int cgfdm_point_source_scales(double DH, double fc, double mu_src, double mu_max, double* Cv, double* Cs) {
    if (!Cv || !Cs || !(DH > 0) || !(fc > 0) || !(mu_src > 0) || !(mu_max > 0)) {
        set_err("point_source_scales needs DH, fc, mu_src and mu_max > 0");
        return CGFDM_EINVAL;
    }
    point_source_scales(DH, fc, mu_src, mu_max, *Cv, *Cs);
    if (!(*Cv > 0) || !std::isfinite(*Cv)) {
        set_err("point_source_scales: no finite scale");
        return CGFDM_EINVAL;
    }
    return CGFDM_OK;
}

int cgfdm_domain_create(const cgfdm_desc* desc, int device, cgfdm_domain** out) {
    if (!out) { set_err("null output handle"); return CGFDM_EINVAL; }
    *out = nullptr;
    int rc = validate(desc);
    if (rc) return rc;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) { set_err("device %d not present (%d devices)", device, ndev); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(device));
    cgfdm_domain* h = new cgfdm_domain();
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) { set_err("hipStreamCreate: %s", hipGetErrorString(e)); delete h; return CGFDM_EHIP; }
    if ((e = hipStreamCreate(&h->side)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&h->evFork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&h->evJoin, hipEventDisableTiming)) != hipSuccess) {
        set_err("side stream: %s", hipGetErrorString(e));
        cgfdm_domain_destroy(h);
        return CGFDM_EHIP;
    }
    for (int q = 0; q < 2; ++q)
        if ((e = hipStreamCreate(&h->side2[q])) != hipSuccess ||
            (e = hipEventCreateWithFlags(&h->evJoin2[q], hipEventDisableTiming)) != hipSuccess) {
            set_err("side stream: %s", hipGetErrorString(e));
            cgfdm_domain_destroy(h);
            return CGFDM_EHIP;
        }
    if ((rc = domain_init(h, desc, device))) {
        cgfdm_domain_destroy(h);
        return rc;
    }
    *out = h;
    return CGFDM_OK;
}

int cgfdm_domain_destroy(cgfdm_domain* h) {
    if (!h) return CGFDM_OK;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    for (auto e : h->evStart) (void)hipEventDestroy(e);
    for (auto e : h->evStop) (void)hipEventDestroy(e);
    if (h->side) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
    }
    for (int q = 0; q < 2; ++q) {
        if (h->side2[q]) {
            (void)hipStreamSynchronize(h->side2[q]);
            (void)hipStreamDestroy(h->side2[q]);
        }
        if (h->evJoin2[q]) (void)hipEventDestroy(h->evJoin2[q]);
    }
    if (h->evFork) (void)hipEventDestroy(h->evFork);
    if (h->evJoin) (void)hipEventDestroy(h->evJoin);
    if (h->stream && h->ownStream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CGFDM_OK;
}

long long cgfdm_domain_bytes(const cgfdm_domain* h) { return h ? h->bytes : 0; }

int cgfdm_set_model(cgfdm_domain* h, const double* terrain, const float* vs, const float* vp, const float* rho) {
    if (!h || !vs || !vp || !rho) { set_err("null argument"); return CGFDM_EINVAL; }
    MediumSrc m{vs, vp, rho, nullptr, 0};
    return domain_set_model(h, terrain, m);
}

int cgfdm_set_model_layered(cgfdm_domain* h, const double* terrain, int nlayer, const float* depth, const float* vs,
                            const float* vp, const float* rho) {
    if (!h || nlayer < 1 || !depth || !vs || !vp || !rho) { set_err("bad layered model"); return CGFDM_EINVAL; }
    std::vector<float> L(4 * (size_t)nlayer);
    for (int l = 0; l < nlayer; ++l) {
        L[l] = depth[l];
        L[nlayer + l] = vs[l];
        L[2 * nlayer + l] = vp[l];
        L[3 * nlayer + l] = rho[l];
    }
    MediumSrc m{nullptr, nullptr, nullptr, L.data(), nlayer};
    return domain_set_model(h, terrain, m);
}

int cgfdm_set_source_point(cgfdm_domain* h, int X, int Y, int Z, float fc) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    if (!h->modelSet) { set_err("set_model must precede set_source_point"); return CGFDM_ESTATE; }
    if (X < 0 || X >= h->nx || Y < 0 || Y >= h->ny || Z < 0 || Z >= h->nz || !(fc > 0)) {
        set_err("source (%d,%d,%d) outside the grid or fc <= 0", X, Y, Z);
        return CGFDM_EINVAL;
    }
    const int SX = X + HALO, SY = Y + HALO, SZ = Z + HALO;
    // the 7^3 smoothing stencil must stay inside the padded grid
    if (SX + 3 >= h->nx_ || SY + 3 >= h->ny_ || SZ + 3 >= h->nz_) {
        set_err("source too close to the grid edge");
        return CGFDM_EINVAL;
    }
    HIPCHK(hipSetDevice(h->dev));
    // normalised Gaussian weights (GAUSS_FUN, macro.h:39; singleSource.cpp:100-110), in double
    std::vector<float> amp(343);
    const double ra = 3 * 0.5;
    for (int q = 0; q < 343; ++q) {
        const int di = q % 7 - 3, dj = (q / 7) % 7 - 3, dk = q / 49 - 3;
        auto G = [&](int t) { return (float)(std::exp(-((t / ra) * (t / ra))) / (ra * 1.772453850905516)); };
        float a = G(di) * G(dj) * G(dk);
        amp[q] = (float)(a / 0.998125703461425);
    }
    int rc;
    float* damp = nullptr;
    if (!h->srcCoef && (rc = h->alloc((void**)&h->srcCoef, 343 * sizeof(float)))) return rc;
    if ((rc = h->alloc((void**)&damp, 343 * sizeof(float)))) return rc;
    h->bytes -= 343 * sizeof(float);
    if ((rc = h->h2d(damp, amp.data(), 343 * sizeof(float)))) return rc;
    LCHK(launch_source_coef(h->prec, h->cjm, h->lay, SX, SY, SZ, damp, (float)h->d.DH,
                            h->Cs * (double)(float)h->d.DT, h->srcCoef, h->stream));
    if ((rc = h->settle())) return rc;
    h->release(damp);
    h->SX = SX; h->SY = SY; h->SZ = SZ; h->fc = fc;
    h->srcMode = 1;
    return CGFDM_OK;
}

int cgfdm_set_source_table(cgfdm_domain* h, const int* xyz, const float* rate, long long npts, int nt, float dt) {
    if (!h || !xyz || !rate || npts <= 0 || nt < 2 || !(dt > 0)) { set_err("bad source table"); return CGFDM_EINVAL; }
    if (!h->modelSet) { set_err("set_model must precede set_source_table"); return CGFDM_ESTATE; }
    HIPCHK(hipSetDevice(h->dev));
    // merge duplicate grid nodes (dealDuplicateIndex, multiSource.cpp:566-624), ascending index
    std::map<long long, long long> uniq;
    std::vector<long long> idxOf(npts);
    for (long long p = 0; p < npts; ++p) {
        const int X = xyz[3 * p], Y = xyz[3 * p + 1], Z = xyz[3 * p + 2];
        if (X < 0 || X >= h->nx || Y < 0 || Y >= h->ny || Z < 0 || Z >= h->nz) {
            set_err("source point %lld (%d,%d,%d) outside the grid", p, X, Y, Z);
            return CGFDM_EINVAL;
        }
        idxOf[p] = (X + HALO) + (long long)(Y + HALO) * h->nx_ + (long long)(Z + HALO) * h->nx_ * h->ny_;
        uniq.emplace(idxOf[p], 0);
    }
    long long u = 0;
    for (auto& kv : uniq) kv.second = u++;
    std::vector<int> sidx(3 * u);
    for (auto& kv : uniq) {
        const long long key = kv.first, q = kv.second;
        sidx[3 * q] = (int)(key % h->nx_);
        sidx[3 * q + 1] = (int)((key / h->nx_) % h->ny_);
        sidx[3 * q + 2] = (int)(key / ((long long)h->nx_ * h->ny_));
    }
    std::vector<float> mr((size_t)nt * u * 6, 0.f);
    for (long long p = 0; p < npts; ++p) {
        const long long q = uniq[idxOf[p]];
        for (int t = 0; t < nt; ++t)
            for (int c = 0; c < 6; ++c) mr[((size_t)t * u + q) * 6 + c] += rate[((size_t)t * npts + p) * 6 + c];
    }
    h->release(h->srcIdx); h->srcIdx = nullptr;
    h->release(h->mrate); h->mrate = nullptr;
    int rc;
    if ((rc = h->alloc((void**)&h->srcIdx, sizeof(int) * 3 * u))) return rc;
    if ((rc = h->alloc((void**)&h->mrate, sizeof(float) * mr.size()))) return rc;
    if ((rc = h->h2d(h->srcIdx, sidx.data(), sizeof(int) * 3 * u))) return rc;
    if ((rc = h->h2d(h->mrate, mr.data(), sizeof(float) * mr.size()))) return rc;
    LCHK(launch_scale_mrate(h->prec, h->cjm, h->lay, h->srcIdx, h->mrate, u, nt, (float)h->d.DH, (float)h->Cv,
                            h->stream));
    if ((rc = h->settle())) return rc;
    h->npts = u; h->snt = nt; h->sdt = dt;
    h->srcMode = 2;
    return CGFDM_OK;
}

int cgfdm_set_stations(cgfdm_domain* h, const int* xyz, int n) {
    if (!h || (n > 0 && !xyz) || n < 0) { set_err("bad station list"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    std::vector<int> idx(3 * (size_t)n);
    for (int s = 0; s < n; ++s) {
        const int X = xyz[3 * s], Y = xyz[3 * s + 1], Z = xyz[3 * s + 2];
        if (X < 0 || X >= h->nx || Y < 0 || Y >= h->ny || Z < 0 || Z >= h->nz) {
            set_err("station %d (%d,%d,%d) outside the grid", s, X, Y, Z);
            return CGFDM_EINVAL;
        }
        idx[3 * s] = X + HALO; idx[3 * s + 1] = Y + HALO; idx[3 * s + 2] = Z + HALO;
    }
    h->release(h->stIdx); h->stIdx = nullptr;
    h->release(h->stWave); h->stWave = nullptr;
    h->nst = n;
    if (n == 0) return CGFDM_OK;
    int rc;
    if ((rc = h->alloc((void**)&h->stIdx, sizeof(int) * 3 * n))) return rc;
    if ((rc = h->alloc((void**)&h->stWave, sizeof(float) * (size_t)n * h->d.NT * 9))) return rc;
    if ((rc = h->h2d(h->stIdx, idx.data(), sizeof(int) * 3 * n))) return rc;
    return h->settle();
}

int cgfdm_step(cgfdm_domain* h, int nsteps) {
    if (!h || nsteps < 0) { set_err("bad arguments"); return CGFDM_EINVAL; }
    if (!h->modelSet) { set_err("set_model must precede step"); return CGFDM_ESTATE; }
    if (h->pooled) { set_err("a served domain is stepped by its server"); return CGFDM_ESTATE; }
    HIPCHK(hipSetDevice(h->dev));
    const float DT = (float)h->d.DT;
    StageParams p = domain_stage_params(h);
    for (int n = 0; n < nsteps; ++n) {
        int FB1 = FB_TABLE[h->it % 8][0], FB2 = FB_TABLE[h->it % 8][1], FB3 = FB_TABLE[h->it % 8][2];
        if (h->srcMode == 1)
            LCHK(launch_point_source(h->prec, h->W[0], h->W[1], h->W[2], h->lay, h->SX, h->SY, h->SZ, h->srcCoef,
                                     h->fc, h->it, DT, h->stream));
        if (h->srcMode == 2)  // addMomenteRate (addMoment.cpp:238-264) at irk = 0
            LCHK(launch_multi_source(h->prec, h->W[0], h->lay, h->srcIdx, h->mrate, h->npts, h->snt, h->sdt, h->it,
                                     DT, h->stream));
        if (h->bulk_cfg < 0) {
            p.FB1 = FB1; p.FB2 = FB2; p.FB3 = FB3;
            int rc = tune_bulk(h, p);
            if (rc) return rc;
            p.bulk_cfg = h->bulk_cfg;
            p.bulk_kc = h->bulk_kc;
        }
        for (int irk = 0; irk < 4; ++irk) {
            p.Win = stage_in(h, irk);
            p.Wout = stage_out(h, irk);
            p.irk = irk;
            p.FB1 = FB1; p.FB2 = FB2; p.FB3 = FB3;
            if (h->timing) {
                if (h->evUsed + 3 > h->evStart.size()) {
                    for (int r = 0; r < 3; ++r) {
                        hipEvent_t a;
                        HIPCHK(hipEventCreate(&a));
                        h->evStart.push_back(a);
                    }
                }
                HIPCHK(hipEventRecord(h->evStart[h->evUsed], h->stream));
                LCHK(launch_stage(h->prec, p, 1, h->stream, h->evStart[h->evUsed + 1]));
                HIPCHK(hipEventRecord(h->evStart[h->evUsed + 2], h->stream));
                h->evUsed += 3;
                if (h->evUsed >= 3 * 2048) {  // fold completed events to bound their number
                    HIPCHK(hipStreamSynchronize(h->stream));
                    for (size_t q = 0; q < h->evUsed; q += 3) {
                        float a = 0, b = 0;
                        HIPCHK(hipEventElapsedTime(&a, h->evStart[q], h->evStart[q + 1]));
                        HIPCHK(hipEventElapsedTime(&b, h->evStart[q + 1], h->evStart[q + 2]));
                        h->tBulk += a;
                        h->tGen += b;
                        ++h->tCount;
                    }
                    h->evUsed = 0;
                }
            } else {
                // untimed: the boundary kernels run beside the bulk kernel on the side streams
                StageFork fk{h->side, h->evFork, h->evJoin};
                for (int q = 0; q < 2; ++q) {
                    fk.side2[q] = h->side2[q];
                    fk.join2[q] = h->evJoin2[q];
                }
                LCHK(launch_stage(h->prec, p, 1, h->stream, nullptr, h->side ? &fk : nullptr));
            }
            FB1 = -FB1; FB2 = -FB2; FB3 = -FB3;
        }
        if (h->checkEvery > 0 && (h->it + 1) % h->checkEvery == 0)
            LCHK(launch_nonfinite(h->prec, h->W[0], h->lay, h->nfFlag, h->it + 1, h->stream));
        if (h->nst > 0 && h->it < h->d.NT)
            LCHK(launch_station(h->prec, h->W[0], h->lay, h->stIdx, h->nst, h->stWave, h->d.NT, h->it,
                                (float)(1.0 / h->Cv), (float)(1.0 / h->Cs), h->stream));
        if (h->d.free_surface)
            LCHK(launch_pgv(h->prec, h->W[0], h->lay, h->nx, h->ny, h->nz_ - HALO - 1, h->pgv, (float)(1.0 / h->Cv),
                            h->stream));
        ++h->it;
    }
    return CGFDM_OK;
}

namespace {
// After the domain's stream has drained: CGFDM_ENONFINITE once the guard has found a value that
// is not finite.
int nonfinite_status(cgfdm_domain* h) {
    int f = 0;
    HIPCHK(hipMemcpy(&f, h->nfFlag, sizeof(int), hipMemcpyDeviceToHost));
    if (f != 0) {
        set_err("non-finite wavefield value found after step %d (16-bit storage out of range? see cgfdm_source_scales)",
                f);
        return CGFDM_ENONFINITE;
    }
    return CGFDM_OK;
}
}  // namespace

int cgfdm_sync(cgfdm_domain* h) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    return nonfinite_status(h);
}

int cgfdm_set_check_interval(cgfdm_domain* h, int every) {
    if (!h || every < 0) { set_err("bad arguments"); return CGFDM_EINVAL; }
    if (h->pooled) { set_err("a served domain is checked by its server"); return CGFDM_ESTATE; }
    h->checkEvery = every;
    return CGFDM_OK;
}

int cgfdm_nonfinite_step(cgfdm_domain* h, int* step) {
    if (!h || !step) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    int f = 0;
    HIPCHK(hipMemcpy(&f, h->nfFlag, sizeof(int), hipMemcpyDeviceToHost));
    *step = f != 0 ? f : -1;
    return CGFDM_OK;
}

int cgfdm_current_step(const cgfdm_domain* h) { return h ? h->it : -1; }

int cgfdm_get_stations(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    if (h->nst == 0) return CGFDM_OK;
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipMemcpyAsync(out, h->stWave, sizeof(float) * (size_t)h->nst * h->d.NT * 9, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return nonfinite_status(h);
}

int cgfdm_get_pgv(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipMemcpyAsync(out, h->pgv, sizeof(float) * 2 * (size_t)h->nx * h->ny, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return nonfinite_status(h);
}
int cgfdm_get_slice(cgfdm_domain* h, int axis, int index, int var, float* out) {
    if (!h || !out || axis < 0 || axis > 2 || var < 0 || var > 8) { set_err("bad slice request"); return CGFDM_EINVAL; }
    const int lim = axis == 0 ? h->nx : (axis == 1 ? h->ny : h->nz);
    if (index < 0 || index >= lim) { set_err("slice index %d outside [0,%d)", index, lim); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long n = axis == 0 ? (long long)h->ny * h->nz : (axis == 1 ? (long long)h->nx * h->nz : (long long)h->nx * h->ny);
    float* tmp = nullptr;
    int rc = h->alloc((void**)&tmp, sizeof(float) * n);
    if (rc) return rc;
    h->bytes -= sizeof(float) * n;
    const float c = (float)(1.0 / (var < 3 ? h->Cv : h->Cs));
    LCHK(launch_slice(h->prec, h->W[0], h->lay, h->nx, h->ny, h->nz, axis, index, var, c, tmp, h->stream));
    HIPCHK(hipMemcpyAsync(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->release(tmp);
    return CGFDM_OK;
}

int cgfdm_get_slice_noalloc(cgfdm_domain* h, int axis, int index, int var, float* out) {
    if (!h || !out || axis < 0 || axis > 2 || var < 0 || var > 8) { set_err("bad slice request"); return CGFDM_EINVAL; }
    const int lim = axis == 0 ? h->nx : (axis == 1 ? h->ny : h->nz);
    if (index < 0 || index >= lim) { set_err("slice index %d outside [0,%d)", index, lim); return CGFDM_EINVAL; }
    const long long n = axis == 0 ? (long long)h->ny * h->nz : (axis == 1 ? (long long)h->nx * h->nz : (long long)h->nx * h->ny);
    if (n > kSliceScratchFloats) {
        set_err("slice of %lld points exceeds the %lld-point staging plane of cgfdm_get_slice_noalloc", n,
                kSliceScratchFloats);
        return CGFDM_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    HIPCHK(hipSetDevice(h->dev));
    float* tmp = nullptr;
    HIPCHK(slice_scratch(h->dev, &tmp));
    const float c = (float)(1.0 / (var < 3 ? h->Cv : h->Cs));
    // the kernel writes all n values, so nothing left by an earlier call is read
    LCHK(launch_slice(h->prec, h->W[0], h->lay, h->nx, h->ny, h->nz, axis, index, var, c, tmp, h->stream));
    HIPCHK(hipMemcpyAsync(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGFDM_OK;
}

int cgfdm_get_state(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long n = (long long)h->nx * h->ny * h->nz * 9;
    float* tmp = nullptr;
    int rc = h->alloc((void**)&tmp, sizeof(float) * n);
    if (rc) return rc;
    h->bytes -= sizeof(float) * n;
    LCHK(launch_gather(h->prec, h->W[0], h->lay, 9, h->nx, h->ny, h->nz, tmp, h->stream));
    HIPCHK(hipMemcpyAsync(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->release(tmp);
    return CGFDM_OK;
}

int cgfdm_set_state(cgfdm_domain* h, const float* in) {
    if (!h || !in) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long n = (long long)h->nx * h->ny * h->nz * 9;
    float* tmp = nullptr;
    int rc = h->alloc((void**)&tmp, sizeof(float) * n);
    if (rc) return rc;
    h->bytes -= sizeof(float) * n;
    HIPCHK(hipMemcpyAsync(tmp, in, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    LCHK(launch_scatter(h->prec, h->W[0], h->lay, 9, h->nx, h->ny, h->nz, tmp, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->release(tmp);
    return CGFDM_OK;
}

int cgfdm_get_cjm(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long n = (long long)h->nx_ * h->ny_ * h->nz_ * 13;
    float* tmp = nullptr;
    int rc = h->alloc((void**)&tmp, sizeof(float) * n);
    if (rc) return rc;
    h->bytes -= sizeof(float) * n;
    LCHK(launch_unpack(h->prec, h->cjm, h->lay, 13, tmp, h->stream));
    HIPCHK(hipMemcpyAsync(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->release(tmp);
    return CGFDM_OK;
}

int cgfdm_enable_timing(cgfdm_domain* h, int on) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    h->timing = on != 0;
    return CGFDM_OK;
}

int cgfdm_kernel_timing(cgfdm_domain* h, int reset, double* out, long long* count) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t q = 0; q < h->evUsed; q += 3) {
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, h->evStart[q], h->evStart[q + 1]));
        HIPCHK(hipEventElapsedTime(&b, h->evStart[q + 1], h->evStart[q + 2]));
        h->tBulk += a;
        h->tGen += b;
        ++h->tCount;
    }
    h->evUsed = 0;
    const double n = h->tCount ? (double)h->tCount : 1.0;
    if (out) {
        out[0] = h->tCount ? (h->tBulk + h->tGen) / n : 0.0;
        out[1] = h->tCount ? h->tBulk / n : 0.0;
        out[2] = h->tCount ? h->tGen / n : 0.0;
    }
    if (count) *count = h->tCount;
    if (reset) { h->tBulk = 0; h->tGen = 0; h->tCount = 0; }
    return CGFDM_OK;
}

int cgfdm_stage_timing(cgfdm_domain* h, int reset, double* avg_ms, long long* count) {
    double t[3];
    int rc = cgfdm_kernel_timing(h, reset, t, count);
    if (rc == 0 && avg_ms) *avg_ms = t[0];
    return rc;
}

int cgfdm_decompose(const cgfdm_desc* desc, int* boxes, int max_boxes) {
    int rc = validate(desc);
    if (rc) return rc;
    if (!boxes && max_boxes > 0) { set_err("null box array"); return CGFDM_EINVAL; }
    StageParams q{};
    q.nx = desc->NX; q.ny = desc->NY; q.nz = desc->NZ;
    q.nPML = desc->nPML; q.fs = desc->free_surface; q.pml = desc->pml;
    stage_boxes(q, 1);
    int n = 0;
    auto put = [&](int kind, int axis, const int* B) {
        if (n < max_boxes) {
            int* o = boxes + 8 * n;
            for (int c = 0; c < 6; ++c) o[c] = B[c];
            o[6] = kind;
            o[7] = axis;
        }
        ++n;
    };
    if (q.bulk_on) {
        const int B[6] = {q.bi0, q.bi1, q.bj0, q.bj1, q.bk0, q.bk1};
        put(0, -1, B);
    }
    for (int b = 0; b < q.npbox; ++b) {
        const int mask = q.pbox[b][6];
        if (mask == 1 || mask == 2 || mask == 4) put(1, mask == 1 ? 0 : mask == 2 ? 1 : 2, q.pbox[b]);
        else put(4, mask, q.pbox[b]);
    }
    if (q.fs_on) {
        const int B[6] = {q.bi0, q.bi1, q.bj0, q.bj1, q.fk0, q.fk1};
        put(3, -1, B);
    }
    for (int b = 0; b < q.nbox; ++b) put(2, -1, q.box[b]);
    return n;
}

int cgfdm_uniform_metric_pair(const cgfdm_domain* h, int* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    if (!h->modelSet) { set_err("set_model must precede uniform_metric_pair"); return CGFDM_ESTATE; }
    HIPCHK(hipSetDevice(h->dev));
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, h->m0, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = (flag != 0 && !h->pair0Stream) ? 1 : 0;
    return CGFDM_OK;
}

int cgfdm_partition(const cgfdm_domain* h, long long* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    StageParams q = domain_stage_params(h);
    stage_boxes(q, 1);
    out[0] = q.bulk_on ? (long long)(q.bi1 - q.bi0) * (q.bj1 - q.bj0) * (q.bk1 - q.bk0) : 0;
    out[1] = (long long)q.nx * q.ny * q.nz - out[0];
    out[2] = h->bulk_cfg;
    out[3] = h->bulk_kc;
    return CGFDM_OK;
}

}  // extern "C"
