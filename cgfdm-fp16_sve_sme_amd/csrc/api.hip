// api.hip — host implementation of the C ABI declared in include/cgfdm.h.
//
// One domain = one simulation on one GPU with its own HIP stream (default priority).
// Device memory layout: every 3-D array is k-major, [k][var][j][i] (see Layout in common.hpp):
//   W buffers A, B, C and accumulator T : 9 variables, storage type
//   CJM                                  : 13 components, storage type
//   PML slabs                            : 3 registers x 9 x slab points, FP32
//   rdz, profiles, stations, PGV         : FP32
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/cgfdm.h"
#include "common.hpp"

using namespace cg;

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

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

const int FB_TABLE[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                            {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};  // propagate.cpp:107-117

}  // namespace

struct cgfdm_domain {
    cgfdm_desc d{};
    int dev = 0;
    hipStream_t stream = nullptr;
    int prec = P_F32;
    int sb = 4;  // storage bytes
    int nx = 0, ny = 0, nz = 0, nx_ = 0, ny_ = 0, nz_ = 0;
    Layout lay{};
    double Cv = 1, Cs = 1, Crho = 1;
    void* W[4] = {nullptr, nullptr, nullptr, nullptr};  // A, B, C, T
    void* cjm = nullptr;
    float* prof = nullptr;  // [ax bx dx][ay by dy][az bz dz]
    float* rdz = nullptr;
    float* auxmem = nullptr;
    Slab slab[6]{};
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
    bool multiActive = false;
    int it = 0;
    bool modelSet = false;
    int bulk_cfg = 0, bulk_kc = 0;  // launch shape of the bulk stage kernel
    long long bytes = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> evStart, evStop;
    size_t evUsed = 0;
    double tAccum = 0.0;
    long long tCount = 0;
    std::vector<void*> allocs;

    int alloc(void** p, size_t n) {
        if (n == 0) { *p = nullptr; return 0; }
        hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_err("hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
            return CGFDM_ENOMEM;
        }
        allocs.push_back(*p);
        bytes += (long long)n;
        e = hipMemsetAsync(*p, 0, n, stream);
        if (e != hipSuccess) { set_err("hipMemsetAsync failed: %s", hipGetErrorString(e)); return CGFDM_EHIP; }
        return 0;
    }
    void release(void* p) {
        if (!p) return;
        for (size_t q = 0; q < allocs.size(); ++q)
            if (allocs[q] == p) { allocs.erase(allocs.begin() + q); break; }
        (void)hipFree(p);
    }
};

namespace {

long long slab_points(const cgfdm_desc& d, int s) {
    if (!d.pml) return 0;
    if (s == 5 && d.free_surface) return 0;
    const long long n = d.nPML;
    switch (s % 3) {
        case 0: return n * d.NY * d.NZ;
        case 1: return n * d.NX * d.NZ;
        default: return n * d.NX * d.NY;
    }
}

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
    return 0;
}

// Internal layout: rows padded to a multiple of 32 elements and shifted so that the first column of
// the bulk box (3 + nPML) is 64-byte aligned for both storage widths.
Layout make_layout(const cgfdm_desc& d) {
    Layout L{};
    L.nx_ = d.NX + 6; L.ny_ = d.NY + 6; L.nz_ = d.NZ + 6;
    const int align = 32;
    const int bi0 = HALO + (d.pml ? d.nPML : 0);
    L.lead = (align - bi0 % align) % align;
    L.px = (L.lead + L.nx_ + align - 1) / align * align;
    L.plane = (long long)L.px * L.ny_;
    return L;
}

long long estimate(const cgfdm_desc& d) {
    const Layout L = make_layout(d);
    int sb = d.precision == CGFDM_FP32 ? 4 : 2;
    long long b = (L.size(9) * 4 + L.size(13)) * sb;
    for (int s = 0; s < 6; ++s) b += slab_points(d, s) * 9 * 3 * 4;
    b += (long long)(d.NX + 6) * (d.NY + 6) * 18 * 4;
    b += 3LL * (d.NX + d.NY + d.NZ + 18) * 4;
    b += (long long)d.NX * d.NY * 2 * 4;
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
    return q;
}

float ricker_moment(float rickerfc, int it, float DT) {
    // sourceFunction at irk = 0 (singleSource.cpp:29-56), M0 = 1e16
    float t = ((float)it + 0.0f) * DT;
    float tdelay = 1.2f / rickerfc;
    float f0 = 2;
    float r = PIf * rickerfc * (t - tdelay);
    float rr = r * r;
    float s = (float)(r * (2.0f * rr - 3.0) * std::exp(-rr) * f0 * PIf * rickerfc);
    return s * 1e16f;
}

void* stage_in(cgfdm_domain* h, int irk) { return h->W[irk == 0 ? 0 : (irk == 2 ? 2 : 1)]; }
void* stage_out(cgfdm_domain* h, int irk) { return h->W[irk == 0 ? 1 : (irk == 1 ? 2 : (irk == 2 ? 1 : 0))]; }

}  // namespace

extern "C" {

const char* cgfdm_last_error(void) { return g_err.c_str(); }

long long cgfdm_estimate_bytes(const cgfdm_desc* desc) {
    if (validate(desc) != 0) return -1;
    return estimate(*desc);
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
    hipError_t e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) { set_err("hipStreamCreate: %s", hipGetErrorString(e)); delete h; return CGFDM_EHIP; }
    auto fail = [&](int code) { cgfdm_domain_destroy(h); return code; };
    for (int b = 0; b < 4; ++b)
        if ((rc = h->alloc(&h->W[b], (size_t)h->lay.size(9) * h->sb))) return fail(rc);
    if ((rc = h->alloc(&h->cjm, (size_t)h->lay.size(13) * h->sb))) return fail(rc);
    // PML slabs: x1 y1 z1 x2 y2 z2
    long long auxTotal = 0;
    for (int s = 0; s < 6; ++s) auxTotal += slab_points(h->d, s);
    if (auxTotal > 0) {
        if ((rc = h->alloc((void**)&h->auxmem, (size_t)auxTotal * 9 * 3 * sizeof(float)))) return fail(rc);
    }
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
    if ((rc = h->alloc((void**)&h->prof, sizeof(float) * 3 * (h->nx_ + h->ny_ + h->nz_)))) return fail(rc);
    if ((rc = h->alloc((void**)&h->rdz, sizeof(float) * 18 * (size_t)h->nx_ * h->ny_))) return fail(rc);
    if ((rc = h->alloc((void**)&h->pgv, sizeof(float) * 2 * (size_t)h->nx * h->ny))) return fail(rc);
    std::vector<float> prof;
    pml_profiles(h->d, prof);
    e = hipMemcpyAsync(h->prof, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) { set_err("profile upload: %s", hipGetErrorString(e)); return fail(CGFDM_EHIP); }
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { set_err("create sync: %s", hipGetErrorString(e)); return fail(CGFDM_EHIP); }
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
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CGFDM_OK;
}

long long cgfdm_domain_bytes(const cgfdm_domain* h) { return h ? h->bytes : 0; }

int cgfdm_set_model(cgfdm_domain* h, const double* terrain, const float* vs, const float* vp, const float* rho) {
    if (!h || !vs || !vp || !rho) { set_err("null argument"); return CGFDM_EINVAL; }
    if (h->d.use_terrain && !terrain) { set_err("use_terrain set but no terrain given"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    const long long nin = (long long)h->nx * h->ny * h->nz;
    float vpmax = 0.f;
    for (long long q = 0; q < nin; ++q) {
        if (!(vs[q] >= 0.f) || !(vp[q] > 0.f) || !(rho[q] > 0.f)) {
            set_err("medium must be finite with vp, rho > 0 (index %lld)", q);
            return CGFDM_EINVAL;
        }
        vpmax = std::max(vpmax, vp[q]);
    }
    float *dvs = nullptr, *dvp = nullptr, *drho = nullptr;
    double* dter = nullptr;
    int rc = 0;
    const size_t mb = sizeof(float) * (size_t)nin;
    if ((rc = h->alloc((void**)&dvs, mb)) || (rc = h->alloc((void**)&dvp, mb)) || (rc = h->alloc((void**)&drho, mb))) {
        h->release(dvs); h->release(dvp); h->release(drho);
        return rc;
    }
    h->bytes -= 3 * (long long)mb;
    auto cleanup = [&]() {
        h->release(dvs); h->release(dvp); h->release(drho); h->release(dter);
    };
    if (h->d.use_terrain) {
        const size_t tb = sizeof(double) * (size_t)h->nx_ * h->ny_;
        if ((rc = h->alloc((void**)&dter, tb))) { cleanup(); return rc; }
        h->bytes -= (long long)tb;
        HIPCHK(hipMemcpyAsync(dter, terrain, tb, hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(hipMemcpyAsync(dvs, vs, mb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dvp, vp, mb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(drho, rho, mb, hipMemcpyHostToDevice, h->stream));
    SetupParams q = setup_params(h, dter);
    LCHK(launch_setup_medium(h->prec, q, dvs, dvp, drho, h->cjm, h->stream));
    LCHK(launch_setup_metric(h->prec, q, h->cjm, h->stream));
    if (h->d.free_surface) LCHK(launch_setup_rdz(q, dvs, dvp, drho, h->rdz, h->stream));
    if (h->d.check_cfl) {
        // calc_CFL.cpp:331-352: dtmax = 1.34*Hmin/Vpmax
        unsigned long long* dmin = nullptr;
        if ((rc = h->alloc((void**)&dmin, sizeof(unsigned long long)))) { cleanup(); return rc; }
        h->bytes -= 8;
        double init = 1.0e20;
        HIPCHK(hipMemcpyAsync(dmin, &init, 8, hipMemcpyHostToDevice, h->stream));
        LCHK(launch_hmin(q, dmin, h->stream));
        double hmin = 0;
        HIPCHK(hipMemcpyAsync(&hmin, dmin, 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->release(dmin);
        const double dtmax = 1.34 * hmin / vpmax;
        if (h->d.DT > dtmax) {
            cleanup();
            set_err("DT = %.4e exceeds the CFL limit dtmax = %.4e (Hmin %.4e, Vpmax %.4e)", h->d.DT, dtmax, hmin, vpmax);
            return CGFDM_ECFL;
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    cleanup();
    h->modelSet = true;
    return CGFDM_OK;
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
    if (SX < 3 || SY < 3 || SZ < 3 || SX + 3 >= h->nx_ || SY + 3 >= h->ny_ || SZ + 3 >= h->nz_) {
        set_err("source too close to the grid edge");
        return CGFDM_EINVAL;
    }
    HIPCHK(hipSetDevice(h->dev));
    // J at the source from the stored CJM, as load_smooth_source reads it
    float* tmp = nullptr;
    int rc = h->alloc((void**)&tmp, sizeof(float));
    if (rc) return rc;
    h->bytes -= 4;
    const long long idx = h->lay.at(13, 9, SX, SY, SZ);
    LCHK(launch_decode(h->prec, (const char*)h->cjm + idx * h->sb, tmp, 1, h->stream));
    float jacb = 0;
    HIPCHK(hipMemcpyAsync(&jacb, tmp, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->release(tmp);
    const float DH = (float)h->d.DH, DT = (float)h->d.DT;
    std::vector<float> coef(343);
    const double ra = 3 * 0.5;
    for (int q = 0; q < 343; ++q) {
        const int di = q % 7 - 3, dj = (q / 7) % 7 - 3, dk = q / 49 - 3;
        auto G = [&](int t) { return (float)(std::exp(-((t / ra) * (t / ra))) / (ra * 1.772453850905516)); };
        float amp = G(di) * G(dj) * G(dk);
        amp = (float)(amp / 0.998125703461425);
        float value = -1.0f * amp / (jacb * (DH * DH * DH));
        coef[q] = (float)(value * (h->Cs * DT));
    }
    if (!h->srcCoef) {
        if ((rc = h->alloc((void**)&h->srcCoef, 343 * sizeof(float)))) return rc;
    }
    HIPCHK(hipMemcpyAsync(h->srcCoef, coef.data(), 343 * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
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
    HIPCHK(hipMemcpyAsync(h->srcIdx, sidx.data(), sizeof(int) * 3 * u, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->mrate, mr.data(), sizeof(float) * mr.size(), hipMemcpyHostToDevice, h->stream));
    LCHK(launch_scale_mrate(h->prec, h->cjm, h->lay, h->srcIdx, h->mrate, u, nt, (float)h->d.DH, (float)h->Cv,
                            h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->npts = u; h->snt = nt; h->sdt = dt;
    h->srcMode = 2;
    h->multiActive = true;
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
    HIPCHK(hipMemcpyAsync(h->stIdx, idx.data(), sizeof(int) * 3 * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGFDM_OK;
}

int cgfdm_step(cgfdm_domain* h, int nsteps) {
    if (!h || nsteps < 0) { set_err("bad arguments"); return CGFDM_EINVAL; }
    if (!h->modelSet) { set_err("set_model must precede step"); return CGFDM_ESTATE; }
    HIPCHK(hipSetDevice(h->dev));
    const float DT = (float)h->d.DT;
    StageParams p{};
    p.A = h->W[0];
    p.T = h->W[3];
    p.cjm = h->cjm;
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
    p.bulk_cfg = h->bulk_cfg;
    p.bulk_kc = h->bulk_kc;
    p.rDH = (float)(1.0 / h->d.DH);
    p.DT = DT;
    p.Crho = (float)h->Crho;
    for (int n = 0; n < nsteps; ++n) {
        int FB1 = FB_TABLE[h->it % 8][0], FB2 = FB_TABLE[h->it % 8][1], FB3 = FB_TABLE[h->it % 8][2];
        if (h->srcMode == 1)
            LCHK(launch_point_source(h->prec, h->W[0], h->W[1], h->W[2], h->lay, h->SX, h->SY, h->SZ, h->srcCoef,
                                     ricker_moment(h->fc, h->it, DT), h->stream));
        if (h->srcMode == 2 && h->multiActive) {
            // addMomenteRate (addMoment.cpp:238-264) at irk = 0
            const float tmpT = ((float)h->it + 0.0f) * DT;
            const int srcIt = int(tmpT / h->sdt);
            if (srcIt + 1 >= h->snt) {
                h->multiActive = false;
            } else {
                const float t1 = float(srcIt) * h->sdt;
                const float tw = (tmpT - t1) / h->sdt;
                LCHK(launch_multi_source(h->prec, h->W[0], h->lay, h->srcIdx, h->mrate, h->npts, srcIt, tw, DT,
                                         h->stream));
            }
        }
        for (int irk = 0; irk < 4; ++irk) {
            p.Win = stage_in(h, irk);
            p.Wout = stage_out(h, irk);
            p.irk = irk;
            p.FB1 = FB1; p.FB2 = FB2; p.FB3 = FB3;
            if (h->timing) {
                if (h->evUsed == h->evStart.size()) {
                    hipEvent_t a, b;
                    HIPCHK(hipEventCreate(&a));
                    HIPCHK(hipEventCreate(&b));
                    h->evStart.push_back(a);
                    h->evStop.push_back(b);
                }
                HIPCHK(hipEventRecord(h->evStart[h->evUsed], h->stream));
            }
            LCHK(launch_stage(h->prec, p, 1, h->stream));
            if (h->timing) {
                HIPCHK(hipEventRecord(h->evStop[h->evUsed], h->stream));
                ++h->evUsed;
                if (h->evUsed >= 4096) {  // fold completed events to bound their number
                    HIPCHK(hipStreamSynchronize(h->stream));
                    for (size_t q = 0; q < h->evUsed; ++q) {
                        float ms = 0;
                        HIPCHK(hipEventElapsedTime(&ms, h->evStart[q], h->evStop[q]));
                        h->tAccum += ms;
                        ++h->tCount;
                    }
                    h->evUsed = 0;
                }
            }
            FB1 = -FB1; FB2 = -FB2; FB3 = -FB3;
        }
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

int cgfdm_sync(cgfdm_domain* h) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGFDM_OK;
}

int cgfdm_current_step(const cgfdm_domain* h) { return h ? h->it : -1; }

int cgfdm_get_stations(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    if (h->nst == 0) return CGFDM_OK;
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipMemcpyAsync(out, h->stWave, sizeof(float) * (size_t)h->nst * h->d.NT * 9, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGFDM_OK;
}

int cgfdm_get_pgv(cgfdm_domain* h, float* out) {
    if (!h || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipMemcpyAsync(out, h->pgv, sizeof(float) * 2 * (size_t)h->nx * h->ny, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return CGFDM_OK;
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

int cgfdm_stage_timing(cgfdm_domain* h, int reset, double* avg_ms, long long* count) {
    if (!h) { set_err("null handle"); return CGFDM_EINVAL; }
    HIPCHK(hipSetDevice(h->dev));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t q = 0; q < h->evUsed; ++q) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->evStart[q], h->evStop[q]));
        h->tAccum += ms;
        ++h->tCount;
    }
    h->evUsed = 0;
    if (avg_ms) *avg_ms = h->tCount ? h->tAccum / h->tCount : 0.0;
    if (count) *count = h->tCount;
    if (reset) { h->tAccum = 0; h->tCount = 0; }
    return CGFDM_OK;
}

}  // extern "C"
