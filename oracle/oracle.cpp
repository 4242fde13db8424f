/*
 * oracle.cpp — CPU restatement of the reference CGFDM time loop.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP path is compared
 * against; it is never linked into, loaded by or called from the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It restates, loop for loop, the reference's scalar CPU path (the FP32-CGFDM build:
 * src-new-FP32-SVE/ sources compiled without GPU_CUDA/SVE) for one MPI rank that owns the
 * whole domain (every boundary physical). One template covers:
 *   Sim<double,double>  FP64 oracle (same discrete scheme, same float literals),
 *   Sim<float,float>    FP32 restatement (the reference plain FP32 CPU path; timed as
 *                       the CPU baseline, OpenMP over k-slabs),
 *   Sim<half16,float>   FP16-storage restatement of FLOAT16 builds (round to nearest
 *                       even at every FLOAT store), and Sim<bf16,float> likewise.
 * The setup (coordinates, metric, Jacobian, rDZ, medium conversion, PML profiles) is
 * computed in FP64 for every variant, then rounded to the storage type, as the
 * MI355X build does.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * src-new-FP16-SVE/, identical in src-new-FP32-SVE/).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>
#include <map>
#include <algorithm>
#include <type_traits>
#include <omp.h>

#include "../include/cgfdm.h"

namespace {

constexpr int HALO = 3;

/* FD coefficients and RK weights, float literals of macro.h:126-155. */
constexpr float af_1 = -0.30874f, af0 = -0.6326f, af1 = 1.2330f, af2 = -0.3334f, af3 = 0.04168f;
constexpr float alpha2 = 0.5f, alpha3 = 0.5f;
constexpr float beta1 = 0.16666667f, beta2 = 0.33333333f, beta3 = 0.33333333f, beta4 = 0.16666667f;
constexpr float Cf1 = -1.16666667f, Cf2 = 1.33333333f, Cf3 = -0.16666667f;
constexpr float PIf = 3.141592657f; /* macro.h:60 */

/* ---------------- storage types ---------------- */
struct half16 { uint16_t b; };
struct bf16 { uint16_t b; };

inline uint16_t f2h_bits(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t mant = x & 0x007fffffu;
    int32_t exp = (int32_t)((x >> 23) & 0xff);
    if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
    int32_t e = exp - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        int shift = 14 - e;
        uint32_t hm = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t hm = mant >> 13;
    uint32_t rem = mant & 0x1fffu;
    uint32_t h = sign | ((uint32_t)e << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
    return (uint16_t)h;
}
inline float h2f_bits(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float f;
    if (e == 0) {
        f = std::ldexp((float)m, -24);
        uint32_t x; std::memcpy(&x, &f, 4); x |= sign; std::memcpy(&f, &x, 4);
        return f;
    }
    uint32_t x;
    if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e - 15 + 127) << 23) | (m << 13);
    std::memcpy(&f, &x, 4);
    return f;
}
inline uint16_t f2bf_bits(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
inline float bf2f_bits(uint16_t b) {
    uint32_t x = (uint32_t)b << 16; float f; std::memcpy(&f, &x, 4); return f;
}

template <class S> struct ST;
template <> struct ST<double> {
    static double get(double v) { return v; }
    static double put(double v) { return v; }
};
template <> struct ST<float> {
    static double get(float v) { return v; }
    static float put(double v) { return (float)v; }
};
template <> struct ST<half16> {
    static double get(half16 v) { return h2f_bits(v.b); }
    static half16 put(double v) { return half16{f2h_bits((float)v)}; }
};
template <> struct ST<bf16> {
    static double get(bf16 v) { return bf2f_bits(v.b); }
    static bf16 put(double v) { return bf16{f2bf_bits((float)v)}; }
};

thread_local std::string g_err;

/* ---------------- FP64 setup (common to every variant) ---------------- */
struct Setup {
    cgfdm_desc d;
    int nx, ny, nz, nx_, ny_, nz_;
    long long num;
    double Cv, Cs, Crho;
    std::vector<double> cjm;          /* 13 x num, SoA (CJM.cpp:117-131) */
    std::vector<double> rdzdx, rdzdy; /* nx_*ny_*9 (contravariant.cpp:179-200) */
    std::vector<double> pa[3], pb[3], pd[3]; /* alpha, 1/beta, d/beta per axis */

    long long INDEX(int i, int j, int k) const { return i + (long long)j * nx_ + (long long)k * nx_ * ny_; }

    /* Coordinates: flat (coord.cpp:106-142) or terrain-following stretch
     * (coord.cpp:255-307, terrain.cpp:930-937). One rank: frontN* = 0. */
    void coord(const double* terrain, int i, int j, int k, double c[3]) const {
        c[0] = (i - HALO) * d.DH - d.centerX * d.DH;
        c[1] = (j - HALO) * d.DH - d.centerY * d.DH;
        if (d.use_terrain) {
            double h = terrain[i + (long long)j * nx_];
            double D = std::fabs(d.depth);
            double DZ = (h + D) / double(d.NZ - 1);
            c[2] = -D + DZ * (k - HALO);
        } else {
            c[2] = (k - HALO + 1) * d.DH - d.NZ * d.DH;
        }
    }

    void build(const double* terrain, const float* vs, const float* vp, const float* rho) {
        cjm.assign((size_t)13 * num, 0.0);
        const double rDH = 1.0 / d.DH;
        const int _nx = nx_ - HALO, _ny = ny_ - HALO, _nz = nz_ - HALO;
        /* coordinates of the whole padded grid */
        std::vector<double> C((size_t)num * 3);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < nz_; ++k)
            for (int j = 0; j < ny_; ++j)
                for (int i = 0; i < nx_; ++i) coord(terrain, i, j, k, &C[(size_t)INDEX(i, j, k) * 3]);
        /* solve_con_jac: metric as the adjugate of the averaged +/- MacCormack coordinate
         * derivatives (contravariant.cpp:71-171). */
        auto LC = [&](long long idx, long long stride, int var, int FB) {
            return FB * (af_1 * C[(idx - FB * stride) * 3 + var] + af0 * C[idx * 3 + var] +
                         af1 * C[(idx + FB * stride) * 3 + var] + af2 * C[(idx + 2 * FB * stride) * 3 + var] +
                         af3 * C[(idx + 3 * FB * stride) * 3 + var]) * rDH;
        };
        const long long sx = 1, sy = nx_, sz = (long long)nx_ * ny_;
#pragma omp parallel for collapse(2)
        for (int k = HALO; k < _nz; ++k)
            for (int j = HALO; j < _ny; ++j)
                for (int i = HALO; i < _nx; ++i) {
                    long long idx = INDEX(i, j, k);
                    double x_xi = 0.5 * (LC(idx, sx, 0, 1) + LC(idx, sx, 0, -1));
                    double x_et = 0.5 * (LC(idx, sy, 0, 1) + LC(idx, sy, 0, -1));
                    double x_zt = 0.5 * (LC(idx, sz, 0, 1) + LC(idx, sz, 0, -1));
                    double y_xi = 0.5 * (LC(idx, sx, 1, 1) + LC(idx, sx, 1, -1));
                    double y_et = 0.5 * (LC(idx, sy, 1, 1) + LC(idx, sy, 1, -1));
                    double y_zt = 0.5 * (LC(idx, sz, 1, 1) + LC(idx, sz, 1, -1));
                    double z_xi = 0.5 * (LC(idx, sx, 2, 1) + LC(idx, sx, 2, -1));
                    double z_et = 0.5 * (LC(idx, sy, 2, 1) + LC(idx, sy, 2, -1));
                    double z_zt = 0.5 * (LC(idx, sz, 2, 1) + LC(idx, sz, 2, -1));
                    double J = x_xi * y_et * z_zt + x_et * y_zt * z_xi + x_zt * y_xi * z_et -
                               x_zt * y_et * z_xi - x_et * y_xi * z_zt - x_xi * z_et * y_zt;
                    double Ji = 1.0 / J;
                    double m[9] = {(y_et * z_zt - y_zt * z_et) * Ji, (x_zt * z_et - x_et * z_zt) * Ji,
                                   (x_et * y_zt - x_zt * y_et) * Ji, (y_zt * z_xi - y_xi * z_zt) * Ji,
                                   (x_xi * z_zt - x_zt * z_xi) * Ji, (x_zt * y_xi - x_xi * y_zt) * Ji,
                                   (y_xi * z_et - y_et * z_xi) * Ji, (x_et * z_xi - x_xi * z_et) * Ji,
                                   (x_xi * y_et - x_et * y_xi) * Ji};
                    for (int c = 0; c < 9; ++c) cjm[(size_t)c * num + idx] = m[c];
                    cjm[(size_t)9 * num + idx] = J;
                }
        /* medium: calculate_medium (CJM.cpp:46-73) with the FP16 scaling folded in. */
        const double cfac = 1.0 * Cs / Cv;
#pragma omp parallel for collapse(2)
        for (int k = HALO; k < _nz; ++k)
            for (int j = HALO; j < _ny; ++j)
                for (int i = HALO; i < _nx; ++i) {
                    long long idx = INDEX(i, j, k);
                    long long p = (i - HALO) + (long long)(j - HALO) * d.NX + (long long)(k - HALO) * d.NX * d.NY;
                    double Vs = vs[p], Vp = vp[p], r = rho[p] * cfac;
                    cjm[(size_t)10 * num + idx] = r * (Vs * Vs);
                    cjm[(size_t)11 * num + idx] = r * (Vp * Vp - 2.0 * Vs * Vs);
                    cjm[(size_t)12 * num + idx] = 1.0 / r / Crho;
                }
        /* free surface: rDZ matrices at the surface and mirrored metrics/J above it
         * (contravariant.cpp:208-412). */
        rdzdx.assign((size_t)nx_ * ny_ * 9, 0.0);
        rdzdy.assign((size_t)nx_ * ny_ * 9, 0.0);
        if (d.free_surface) {
            const int k = _nz - 1;
#pragma omp parallel for
            for (int j = HALO; j < _ny; ++j)
                for (int i = HALO; i < _nx; ++i) {
                    long long idx = INDEX(i, j, k);
                    double xi_x = cjm[0 * num + idx], xi_y = cjm[1 * num + idx], xi_z = cjm[2 * num + idx];
                    double et_x = cjm[3 * num + idx], et_y = cjm[4 * num + idx], et_z = cjm[5 * num + idx];
                    double zt_x = cjm[6 * num + idx], zt_y = cjm[7 * num + idx], zt_z = cjm[8 * num + idx];
                    double mu = cjm[10 * num + idx], lambda = cjm[11 * num + idx];
                    /* contravariant.cpp:266; chi_mode 1 is the opt-in traction-free lambda + 2 mu */
                    double chi = d.chi_mode ? lambda + 2.0 * mu : 2.0 * lambda + mu;
                    double lam_mu = lambda + mu;
                    double DZ[9], DX[9], DY[9], rDZ[9];
                    DZ[0] = chi * zt_x * zt_x + mu * (zt_y * zt_y + zt_z * zt_z);
                    DZ[1] = lam_mu * zt_x * zt_y;
                    DZ[2] = lam_mu * zt_x * zt_z;
                    DZ[3] = DZ[1];
                    DZ[4] = chi * zt_y * zt_y + mu * (zt_x * zt_x + zt_z * zt_z);
                    DZ[5] = lam_mu * zt_y * zt_z;
                    DZ[6] = DZ[2];
                    DZ[7] = DZ[5];
                    DZ[8] = chi * zt_z * zt_z + mu * (zt_x * zt_x + zt_y * zt_y);
                    double det = DZ[0] * DZ[4] * DZ[8] + DZ[1] * DZ[5] * DZ[6] + DZ[2] * DZ[7] * DZ[3] -
                                 DZ[2] * DZ[4] * DZ[6] - DZ[1] * DZ[3] * DZ[8] - DZ[0] * DZ[7] * DZ[5];
                    DX[0] = chi * zt_x * xi_x + mu * (zt_y * xi_y + zt_z * xi_z);
                    DX[1] = lambda * zt_x * xi_y + mu * zt_y * xi_x;
                    DX[2] = lambda * zt_x * xi_z + mu * zt_z * xi_x;
                    DX[3] = lambda * zt_y * xi_x + mu * zt_x * xi_y;
                    DX[4] = chi * zt_y * xi_y + mu * (zt_x * xi_x + zt_z * xi_z);
                    DX[5] = lambda * zt_y * xi_z + mu * zt_z * xi_y;
                    DX[6] = lambda * zt_z * xi_x + mu * zt_x * xi_z;
                    DX[7] = lambda * zt_z * xi_y + mu * zt_y * xi_z;
                    DX[8] = chi * zt_z * xi_z + mu * (zt_x * xi_x + zt_y * xi_y);
                    DY[0] = chi * zt_x * et_x + mu * (zt_y * et_y + zt_z * et_z);
                    DY[1] = lambda * zt_x * et_y + mu * zt_y * et_x;
                    DY[2] = lambda * zt_x * et_z + mu * zt_z * et_x;
                    DY[3] = lambda * zt_y * et_x + mu * zt_x * et_y;
                    DY[4] = chi * zt_y * et_y + mu * (zt_x * et_x + zt_z * et_z);
                    DY[5] = lambda * zt_y * et_z + mu * zt_z * et_y;
                    DY[6] = lambda * zt_z * et_x + mu * zt_x * et_z;
                    DY[7] = lambda * zt_z * et_y + mu * zt_y * et_z;
                    DY[8] = chi * zt_z * et_z + mu * (zt_x * et_x + zt_y * et_y);
                    rDZ[0] = (DZ[4] * DZ[8] - DZ[5] * DZ[7]) / det;
                    rDZ[1] = (-DZ[3] * DZ[8] + DZ[5] * DZ[6]) / det;
                    rDZ[2] = (DZ[3] * DZ[7] - DZ[4] * DZ[6]) / det;
                    rDZ[3] = (-DZ[1] * DZ[8] + DZ[2] * DZ[7]) / det;
                    rDZ[4] = (DZ[0] * DZ[8] - DZ[2] * DZ[6]) / det;
                    rDZ[5] = (-DZ[0] * DZ[7] + DZ[1] * DZ[6]) / det;
                    rDZ[6] = (DZ[1] * DZ[5] - DZ[2] * DZ[4]) / det;
                    rDZ[7] = (-DZ[0] * DZ[5] + DZ[2] * DZ[3]) / det;
                    rDZ[8] = (DZ[0] * DZ[4] - DZ[1] * DZ[3]) / det;
                    long long s = (i + (long long)j * nx_) * 9;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            rdzdx[s + r * 3 + c] = -(rDZ[r * 3 + 0] * DX[0 + c] + rDZ[r * 3 + 1] * DX[3 + c] + rDZ[r * 3 + 2] * DX[6 + c]);
                            rdzdy[s + r * 3 + c] = -(rDZ[r * 3 + 0] * DY[0 + c] + rDZ[r * 3 + 1] * DY[3 + c] + rDZ[r * 3 + 2] * DY[6 + c]);
                        }
                }
            for (int kk = _nz; kk < nz_; ++kk)
                for (int j = HALO; j < _ny; ++j)
                    for (int i = HALO; i < _nx; ++i) {
                        long long idx = INDEX(i, j, kk), pos = INDEX(i, j, 2 * (_nz - 1) - kk);
                        for (int c = 0; c < 10; ++c) cjm[(size_t)c * num + idx] = cjm[(size_t)c * num + pos];
                    }
        }
        /* PML profiles: cpu_init_pml_parameter (init_pml_para.cpp:99-234), every face physical. */
        const int nPML = d.nPML;
        int nn[3] = {nx_, ny_, nz_};
        int lo[3] = {_nx, _ny, _nz};
        for (int a = 0; a < 3; ++a) {
            pa[a].assign(nn[a], 0.0);
            pb[a].assign(nn[a], 1.0);
            pd[a].assign(nn[a], 0.0);
        }
        if (d.pml) {
            const double DH = d.DH;
            const double Cp = 6000.0, pml_fc = 0.5;
            const double L = DH * (nPML - 1);
            const double R = std::pow(10.0, -(std::log10((double)nPML) - 1.0) / std::log10(2.0) - 3.0);
            const double d0 = -(2.0 + 1.0) * Cp / (2.0 * L) * std::log(R);
            const double alpha0 = (double)PIf * pml_fc, beta0 = 3.0;
            for (int a = 0; a < 3; ++a) {
                std::vector<double> Ld(nn[a], -1.0);
                for (int i = HALO; i < lo[a]; ++i) {
                    if (i < HALO + nPML) Ld[i] = (HALO + nPML - 1 - i) * DH;
                    bool top = (a == 2);
                    if (i >= lo[a] - nPML && !(top && d.free_surface)) Ld[i] = (i - (lo[a] - nPML)) * DH;
                }
                for (int i = HALO; i < lo[a]; ++i) {
                    double l = Ld[i];
                    pa[a][i] = l < 0.0 ? 0.0 : alpha0 * (1.0 - std::pow(l / L, 1.0));
                    pb[a][i] = l < 0.0 ? 1.0 : 1.0 + (beta0 - 1.0) * std::pow(l / L, 2.0);
                    pd[a][i] = l < 0.0 ? 0.0 : d0 * std::pow(l / L, 2.0);
                }
                for (int i = 0; i < nn[a]; ++i) {
                    pd[a][i] /= pb[a][i];
                    pb[a][i] = 1.0 / pb[a][i];
                }
            }
        }
    }
};

/* ---------------- time loop ---------------- */
template <class S, class R>
struct Sim {
    Setup su;
    int nx_, ny_, nz_, nx, ny, nz, _nx, _ny, _nz;
    long long num;
    R rDH, DT, Crho_r;
    std::vector<S> hW, W, tW, mW, cjm;
    std::vector<R> rdzdx, rdzdy, pa[3], pb[3], pd[3];
    struct Aux { bool on = false; long long n = 0; std::vector<R> h, a, t, m; };
    Aux aux[6]; /* x1, y1, z1, x2, y2, z2 (struct.h:565-574) */
    /* sources */
    int srcMode = 0; /* 0 none, 1 point, 2 table */
    int SX = 0, SY = 0, SZ = 0; float fc = 0.f;
    std::vector<long long> srcIndex; std::vector<R> momentRate; long long npts = 0; int snt = 0; float sdt = 0.f;
    bool multiActive = false;
    /* stations / pgv */
    std::vector<int> st; int nst = 0; std::vector<float> stWave; std::vector<float> pgv;
    int it = 0;
    bool modelSet = false;

    long long INDEX(int i, int j, int k) const { return i + (long long)j * nx_ + (long long)k * nx_ * ny_; }
    static R ld(const S& v) { return (R)ST<S>::get(v); }
    static S st_(R v) { return ST<S>::put((double)v); }

    explicit Sim(const cgfdm_desc& d) {
        su.d = d;
        nx = d.NX; ny = d.NY; nz = d.NZ;
        nx_ = nx + 2 * HALO; ny_ = ny + 2 * HALO; nz_ = nz + 2 * HALO;
        _nx = nx_ - HALO; _ny = ny_ - HALO; _nz = nz_ - HALO;
        su.nx = nx; su.ny = ny; su.nz = nz; su.nx_ = nx_; su.ny_ = ny_; su.nz_ = nz_;
        num = (long long)nx_ * ny_ * nz_;
        su.num = num;
        rDH = (R)(1.0 / d.DH);
        DT = (R)d.DT;
        hW.assign((size_t)9 * num, st_(0)); W = hW; tW = hW; mW = hW;
        if (d.pml) {
            long long n1[6] = {(long long)d.nPML * ny * nz, (long long)d.nPML * nx * nz, (long long)d.nPML * nx * ny,
                               (long long)d.nPML * ny * nz, (long long)d.nPML * nx * nz, (long long)d.nPML * nx * ny};
            for (int s = 0; s < 6; ++s) {
                if (s == 5 && d.free_surface) continue;
                aux[s].on = true; aux[s].n = n1[s];
                aux[s].h.assign((size_t)9 * n1[s], 0); aux[s].a = aux[s].h; aux[s].t = aux[s].h; aux[s].m = aux[s].h;
            }
        }
    }

    void setModel(const double* terrain, const float* vs, const float* vp, const float* rho, double Cv, double Cs, double Crho) {
        su.Cv = Cv; su.Cs = Cs; su.Crho = Crho;
        su.build(terrain, vs, vp, rho);
        cjm.resize((size_t)13 * num);
        for (size_t q = 0; q < cjm.size(); ++q) cjm[q] = st_((R)su.cjm[q]);
        rdzdx.resize(su.rdzdx.size()); rdzdy.resize(su.rdzdy.size());
        for (size_t q = 0; q < rdzdx.size(); ++q) { rdzdx[q] = (R)su.rdzdx[q]; rdzdy[q] = (R)su.rdzdy[q]; }
        for (int a = 0; a < 3; ++a) {
            pa[a].assign(su.pa[a].begin(), su.pa[a].end());
            pb[a].assign(su.pb[a].begin(), su.pb[a].end());
            pd[a].assign(su.pd[a].begin(), su.pd[a].end());
        }
        Crho_r = (R)Crho;
        modelSet = true;
    }

    /* operator L (axis.h:129-133) */
    inline R L(const std::vector<S>& w, int var, long long idx, long long stride, int FB) const {
        const S* p = w.data() + (size_t)var * num;
        return (R)FB * ((R)af_1 * ld(p[idx - FB * stride]) + (R)af0 * ld(p[idx]) + (R)af1 * ld(p[idx + FB * stride]) +
                        (R)af2 * ld(p[idx + 2 * FB * stride]) + (R)af3 * ld(p[idx + 3 * FB * stride])) * rDH;
    }
    inline R C(int c, long long idx) const { return ld(cjm[(size_t)c * num + idx]); }
    inline R w(const std::vector<S>& a, int v, long long idx) const { return ld(a[(size_t)v * num + idx]); }

    /* wave_deriv (wave_deriv.cpp:58-203) */
    void waveDeriv(int FB1, int FB2, int FB3) {
        const long long sx = 1, sy = nx_, sz = (long long)nx_ * ny_;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = HALO; k < _nz; ++k)
            for (int j = HALO; j < _ny; ++j)
                for (int i = HALO; i < _nx; ++i) {
                    long long index = INDEX(i, j, k);
                    R bx = pb[0][i], by = pb[1][j], bz = pb[2][k];
                    R xi_x = C(0, index), xi_y = C(1, index), xi_z = C(2, index);
                    R et_x = C(3, index), et_y = C(4, index), et_z = C(5, index);
                    R zt_x = C(6, index), zt_y = C(7, index), zt_z = C(8, index);
                    R mu = C(10, index), lambda = C(11, index), buoyancy = C(12, index);
                    buoyancy *= Crho_r;
                    R Vx_xi = L(W, 0, index, sx, FB1) * bx, Vy_xi = L(W, 1, index, sx, FB1) * bx, Vz_xi = L(W, 2, index, sx, FB1) * bx;
                    R Txx_xi = L(W, 3, index, sx, FB1) * bx, Tyy_xi = L(W, 4, index, sx, FB1) * bx, Tzz_xi = L(W, 5, index, sx, FB1) * bx;
                    R Txy_xi = L(W, 6, index, sx, FB1) * bx, Txz_xi = L(W, 7, index, sx, FB1) * bx, Tyz_xi = L(W, 8, index, sx, FB1) * bx;
                    R Vx_et = L(W, 0, index, sy, FB2) * by, Vy_et = L(W, 1, index, sy, FB2) * by, Vz_et = L(W, 2, index, sy, FB2) * by;
                    R Txx_et = L(W, 3, index, sy, FB2) * by, Tyy_et = L(W, 4, index, sy, FB2) * by, Tzz_et = L(W, 5, index, sy, FB2) * by;
                    R Txy_et = L(W, 6, index, sy, FB2) * by, Txz_et = L(W, 7, index, sy, FB2) * by, Tyz_et = L(W, 8, index, sy, FB2) * by;
                    R Vx_zt = L(W, 0, index, sz, FB3) * bz, Vy_zt = L(W, 1, index, sz, FB3) * bz, Vz_zt = L(W, 2, index, sz, FB3) * bz;
                    R Txx_zt = L(W, 3, index, sz, FB3) * bz, Tyy_zt = L(W, 4, index, sz, FB3) * bz, Tzz_zt = L(W, 5, index, sz, FB3) * bz;
                    R Txy_zt = L(W, 6, index, sz, FB3) * bz, Txz_zt = L(W, 7, index, sz, FB3) * bz, Tyz_zt = L(W, 8, index, sz, FB3) * bz;
                    R Vx1 = (xi_x * Txx_xi + xi_y * Txy_xi + xi_z * Txz_xi) * buoyancy;
                    R Vx2 = (et_x * Txx_et + et_y * Txy_et + et_z * Txz_et) * buoyancy;
                    R Vx3 = (zt_x * Txx_zt + zt_y * Txy_zt + zt_z * Txz_zt) * buoyancy;
                    R Vy1 = (xi_x * Txy_xi + xi_y * Tyy_xi + xi_z * Tyz_xi) * buoyancy;
                    R Vy2 = (et_x * Txy_et + et_y * Tyy_et + et_z * Tyz_et) * buoyancy;
                    R Vy3 = (zt_x * Txy_zt + zt_y * Tyy_zt + zt_z * Tyz_zt) * buoyancy;
                    R Vz1 = (xi_x * Txz_xi + xi_y * Tyz_xi + xi_z * Tzz_xi) * buoyancy;
                    R Vz2 = (et_x * Txz_et + et_y * Tyz_et + et_z * Tzz_et) * buoyancy;
                    R Vz3 = (zt_x * Txz_zt + zt_y * Tyz_zt + zt_z * Tzz_zt) * buoyancy;
                    R d1 = (xi_x * Vx_xi + xi_y * Vy_xi + xi_z * Vz_xi) * lambda;
                    R d2 = (et_x * Vx_et + et_y * Vy_et + et_z * Vz_et) * lambda;
                    R d3 = (zt_x * Vx_zt + zt_y * Vy_zt + zt_z * Vz_zt) * lambda;
                    R two = 2;
                    R Txx1 = d1 + two * mu * (xi_x * Vx_xi), Txx2 = d2 + two * mu * (et_x * Vx_et), Txx3 = d3 + two * mu * (zt_x * Vx_zt);
                    R Tyy1 = d1 + two * mu * (xi_y * Vy_xi), Tyy2 = d2 + two * mu * (et_y * Vy_et), Tyy3 = d3 + two * mu * (zt_y * Vy_zt);
                    R Tzz1 = d1 + two * mu * (xi_z * Vz_xi), Tzz2 = d2 + two * mu * (et_z * Vz_et), Tzz3 = d3 + two * mu * (zt_z * Vz_zt);
                    R Txy1 = (xi_y * Vx_xi + xi_x * Vy_xi) * mu, Txy2 = (et_y * Vx_et + et_x * Vy_et) * mu, Txy3 = (zt_y * Vx_zt + zt_x * Vy_zt) * mu;
                    R Txz1 = (xi_z * Vx_xi + xi_x * Vz_xi) * mu, Txz2 = (et_z * Vx_et + et_x * Vz_et) * mu, Txz3 = (zt_z * Vx_zt + zt_x * Vz_zt) * mu;
                    R Tyz1 = (xi_z * Vy_xi + xi_y * Vz_xi) * mu, Tyz2 = (et_z * Vy_et + et_y * Vz_et) * mu, Tyz3 = (zt_z * Vy_zt + zt_y * Vz_zt) * mu;
                    S* h = hW.data();
                    h[index + num * 0] = st_((Vx1 + Vx2 + Vx3) * DT);
                    h[index + num * 1] = st_((Vy1 + Vy2 + Vy3) * DT);
                    h[index + num * 2] = st_((Vz1 + Vz2 + Vz3) * DT);
                    h[index + num * 3] = st_((Txx1 + Txx2 + Txx3) * DT);
                    h[index + num * 4] = st_((Tyy1 + Tyy2 + Tyy3) * DT);
                    h[index + num * 5] = st_((Tzz1 + Tzz2 + Tzz3) * DT);
                    h[index + num * 6] = st_((Txy1 + Txy2 + Txy3) * DT);
                    h[index + num * 7] = st_((Txz1 + Txz2 + Txz3) * DT);
                    h[index + num * 8] = st_((Tyz1 + Tyz2 + Tyz3) * DT);
                }
    }

    /* free_surface_deriv (freeSurface.cpp:22-258) */
    void freeSurfaceDeriv(int FB1, int FB2, int FB3) {
        const long long sx = 1, sy = nx_, sz = (long long)nx_ * ny_;
        for (int k = _nz - HALO; k < _nz; ++k) {
#pragma omp parallel for schedule(static)
            for (int j = HALO; j < _ny; ++j)
                for (int i = HALO; i < _nx; ++i) {
                    long long index = INDEX(i, j, k);
                    R mu = C(10, index), lambda = C(11, index), buoyancy = C(12, index) * Crho_r;
                    R bx = pb[0][i], by = pb[1][j];
                    int k_s = (_nz - 1) - k + HALO;
                    R J_T1x[7], J_T2x[7], J_T3x[7], J_T1y[7], J_T2y[7], J_T3y[7], J_T1z[7], J_T2z[7], J_T3z[7];
                    for (int l = 0; l <= 2 * HALO; ++l) {
                        long long pos = INDEX(i + (l - HALO), j, k);
                        R xi_x = C(0, pos), xi_y = C(1, pos), xi_z = C(2, pos), jacb = C(9, pos);
                        J_T1x[l] = (xi_x * w(W, 3, pos) + xi_y * w(W, 6, pos) + xi_z * w(W, 7, pos)) * jacb;
                        J_T2x[l] = (xi_x * w(W, 6, pos) + xi_y * w(W, 4, pos) + xi_z * w(W, 8, pos)) * jacb;
                        J_T3x[l] = (xi_x * w(W, 7, pos) + xi_y * w(W, 8, pos) + xi_z * w(W, 5, pos)) * jacb;
                        pos = INDEX(i, j + (l - HALO), k);
                        R et_x = C(3, pos), et_y = C(4, pos), et_z = C(5, pos);
                        jacb = C(9, pos);
                        J_T1y[l] = (et_x * w(W, 3, pos) + et_y * w(W, 6, pos) + et_z * w(W, 7, pos)) * jacb;
                        J_T2y[l] = (et_x * w(W, 6, pos) + et_y * w(W, 4, pos) + et_z * w(W, 8, pos)) * jacb;
                        J_T3y[l] = (et_x * w(W, 7, pos) + et_y * w(W, 8, pos) + et_z * w(W, 5, pos)) * jacb;
                    }
                    for (int l = 0; l < k_s; ++l) {
                        long long pos = INDEX(i, j, k + (l - HALO));
                        R zt_x = C(6, pos), zt_y = C(7, pos), zt_z = C(8, pos), jacb = C(9, pos);
                        J_T1z[l] = (zt_x * w(W, 3, pos) + zt_y * w(W, 6, pos) + zt_z * w(W, 7, pos)) * jacb;
                        J_T2z[l] = (zt_x * w(W, 6, pos) + zt_y * w(W, 4, pos) + zt_z * w(W, 8, pos)) * jacb;
                        J_T3z[l] = (zt_x * w(W, 7, pos) + zt_y * w(W, 8, pos) + zt_z * w(W, 5, pos)) * jacb;
                    }
                    J_T1z[k_s] = 0; J_T2z[k_s] = 0; J_T3z[k_s] = 0;
                    for (int l = k_s + 1; l <= 2 * HALO; ++l) {
                        J_T1z[l] = -J_T1z[2 * k_s - l];
                        J_T2z[l] = -J_T2z[2 * k_s - l];
                        J_T3z[l] = -J_T3z[2 * k_s - l];
                    }
                    R jacb = C(9, index);
                    R Jinv = (R)1 / jacb;
                    auto LJT = [&](const R* J_T, int FB) {
                        return (R)FB * ((R)af_1 * J_T[3 - FB] + (R)af0 * J_T[3] + (R)af1 * J_T[3 + FB] + (R)af2 * J_T[3 + 2 * FB] +
                                        (R)af3 * J_T[3 + 3 * FB]) * rDH;
                    };
                    R hVx = buoyancy * Jinv * (LJT(J_T1x, FB1) * bx + LJT(J_T1y, FB2) * by + LJT(J_T1z, FB3));
                    R hVy = buoyancy * Jinv * (LJT(J_T2x, FB1) * bx + LJT(J_T2y, FB2) * by + LJT(J_T2z, FB3));
                    R hVz = buoyancy * Jinv * (LJT(J_T3x, FB1) * bx + LJT(J_T3y, FB2) * by + LJT(J_T3z, FB3));
                    R Vx_xi = L(W, 0, index, sx, FB1) * bx, Vx_et = L(W, 0, index, sy, FB2) * by;
                    R Vy_xi = L(W, 1, index, sx, FB1) * bx, Vy_et = L(W, 1, index, sy, FB2) * by;
                    R Vz_xi = L(W, 2, index, sx, FB1) * bx, Vz_et = L(W, 2, index, sy, FB2) * by;
                    R xi_x = C(0, index), xi_y = C(1, index), xi_z = C(2, index);
                    R et_x = C(3, index), et_y = C(4, index), et_z = C(5, index);
                    R zt_x = C(6, index), zt_y = C(7, index), zt_z = C(8, index);
                    R Vx_zt = 0, Vy_zt = 0, Vz_zt = 0;
                    if (k == _nz - 1) {
                        long long s = (i + (long long)j * nx_) * 9;
                        const R* X = &rdzdx[s];
                        const R* Y = &rdzdy[s];
                        Vx_zt = (X[0] * Vx_xi + X[1] * Vy_xi + X[2] * Vz_xi) + (Y[0] * Vx_et + Y[1] * Vy_et + Y[2] * Vz_et);
                        Vy_zt = (X[3] * Vx_xi + X[4] * Vy_xi + X[5] * Vz_xi) + (Y[3] * Vx_et + Y[4] * Vy_et + Y[5] * Vz_et);
                        Vz_zt = (X[6] * Vx_xi + X[7] * Vy_xi + X[8] * Vz_xi) + (Y[6] * Vx_et + Y[7] * Vy_et + Y[8] * Vz_et);
                    }
                    if (k == _nz - 2) { /* L2 (axis.h:92) */
                        Vx_zt = (R)FB3 * (w(W, 0, index + FB3 * sz) - w(W, 0, index)) * rDH;
                        Vy_zt = (R)FB3 * (w(W, 1, index + FB3 * sz) - w(W, 1, index)) * rDH;
                        Vz_zt = (R)FB3 * (w(W, 2, index + FB3 * sz) - w(W, 2, index)) * rDH;
                    }
                    if (k == _nz - 3) { /* L3 (axis.h:93) */
                        for (int v = 0; v < 3; ++v) {
                            R val = (R)FB3 * ((R)Cf1 * w(W, v, index) + (R)Cf2 * w(W, v, index + FB3 * sz) +
                                              (R)Cf3 * w(W, v, index + 2 * FB3 * sz)) * rDH;
                            if (v == 0) Vx_zt = val; else if (v == 1) Vy_zt = val; else Vz_zt = val;
                        }
                    }
                    R d1 = (xi_x * Vx_xi + xi_y * Vy_xi + xi_z * Vz_xi) * lambda;
                    R d2 = (et_x * Vx_et + et_y * Vy_et + et_z * Vz_et) * lambda;
                    R d3 = (zt_x * Vx_zt + zt_y * Vy_zt + zt_z * Vz_zt) * lambda;
                    R two = 2;
                    R hTxx = (d1 + two * mu * (xi_x * Vx_xi)) + (d2 + two * mu * (et_x * Vx_et)) + (d3 + two * mu * (zt_x * Vx_zt));
                    R hTyy = (d1 + two * mu * (xi_y * Vy_xi)) + (d2 + two * mu * (et_y * Vy_et)) + (d3 + two * mu * (zt_y * Vy_zt));
                    R hTzz = (d1 + two * mu * (xi_z * Vz_xi)) + (d2 + two * mu * (et_z * Vz_et)) + (d3 + two * mu * (zt_z * Vz_zt));
                    R hTxy = (xi_y * Vx_xi + xi_x * Vy_xi) * mu + (et_y * Vx_et + et_x * Vy_et) * mu + (zt_y * Vx_zt + zt_x * Vy_zt) * mu;
                    R hTxz = (xi_z * Vx_xi + xi_x * Vz_xi) * mu + (et_z * Vx_et + et_x * Vz_et) * mu + (zt_z * Vx_zt + zt_x * Vz_zt) * mu;
                    R hTyz = (xi_z * Vy_xi + xi_y * Vz_xi) * mu + (et_z * Vy_et + et_y * Vz_et) * mu + (zt_z * Vy_zt + zt_y * Vz_zt) * mu;
                    S* h = hW.data();
                    h[index + num * 0] = st_(hVx * DT);
                    h[index + num * 1] = st_(hVy * DT);
                    h[index + num * 2] = st_(hVz * DT);
                    h[index + num * 3] = st_(hTxx * DT);
                    h[index + num * 4] = st_(hTyy * DT);
                    h[index + num * 5] = st_(hTzz * DT);
                    h[index + num * 6] = st_(hTxy * DT);
                    h[index + num * 7] = st_(hTxz * DT);
                    h[index + num * 8] = st_(hTyz * DT);
                }
        }
    }

    /* pml_deriv_x/y/z (pml_deriv.cpp:90-544): axis a, FLAG 0 (low face) or 1 (high face). */
    void pmlDerivSlab(int a, int FLAG, Aux& X, int FB) {
        const long long strides[3] = {1, nx_, (long long)nx_ * ny_};
        const long long stride = strides[a];
        const int nPML = su.d.nPML;
        int n0 = a == 0 ? nPML : nx, n1 = a == 1 ? nPML : ny, n2 = a == 2 ? nPML : nz;
        int nAxis = a == 0 ? nx : (a == 1 ? ny : nz);
        int off = FLAG * (nAxis - nPML);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k0 = 0; k0 < n2; ++k0)
            for (int j0 = 0; j0 < n1; ++j0)
                for (int i0 = 0; i0 < n0; ++i0) {
                    int i = i0 + HALO + (a == 0 ? off : 0);
                    int j = j0 + HALO + (a == 1 ? off : 0);
                    int k = k0 + HALO + (a == 2 ? off : 0);
                    int ia = a == 0 ? i : (a == 1 ? j : k);
                    long long index = INDEX(i, j, k);
                    long long pos = i0 + (long long)j0 * n0 + (long long)k0 * n0 * n1;
                    R mu = C(10, index), lambda = C(11, index), buoyancy = C(12, index) * Crho_r;
                    R beta = pb[a][ia], dd = pd[a][ia], alpha_d = dd + pa[a][ia];
                    R m_x = C(3 * a + 0, index), m_y = C(3 * a + 1, index), m_z = C(3 * a + 2, index);
                    R Vx = L(W, 0, index, stride, FB) * dd, Vy = L(W, 1, index, stride, FB) * dd, Vz = L(W, 2, index, stride, FB) * dd;
                    R Txx = L(W, 3, index, stride, FB) * dd, Tyy = L(W, 4, index, stride, FB) * dd, Tzz = L(W, 5, index, stride, FB) * dd;
                    R Txy = L(W, 6, index, stride, FB) * dd, Txz = L(W, 7, index, stride, FB) * dd, Tyz = L(W, 8, index, stride, FB) * dd;
                    R rhs[9];
                    rhs[0] = (m_x * Txx + m_y * Txy + m_z * Txz) * buoyancy;
                    rhs[1] = (m_x * Txy + m_y * Tyy + m_z * Tyz) * buoyancy;
                    rhs[2] = (m_x * Txz + m_y * Tyz + m_z * Tzz) * buoyancy;
                    R dv = (m_x * Vx + m_y * Vy + m_z * Vz) * lambda;
                    R two = 2;
                    rhs[3] = dv + two * mu * (m_x * Vx);
                    rhs[4] = dv + two * mu * (m_y * Vy);
                    rhs[5] = dv + two * mu * (m_z * Vz);
                    rhs[6] = (m_y * Vx + m_x * Vy) * mu;
                    rhs[7] = (m_z * Vx + m_x * Vz) * mu;
                    rhs[8] = (m_z * Vy + m_y * Vz) * mu;
                    for (int v = 0; v < 9; ++v) X.h[pos * 9 + v] = (rhs[v] - alpha_d * X.a[pos * 9 + v]) * DT;
                    S* h = hW.data();
                    for (int v = 0; v < 9; ++v) {
                        R hv = ld(h[index + num * v]);
                        hv += -beta * X.a[pos * 9 + v] * DT;
                        h[index + num * v] = st_(hv);
                    }
                }
    }

    void pmlDeriv(int FB1, int FB2, int FB3) { /* order of pml_deriv.cpp:607-616 */
        if (aux[0].on) pmlDerivSlab(0, 0, aux[0], FB1);
        if (aux[1].on) pmlDerivSlab(1, 0, aux[1], FB2);
        if (aux[2].on) pmlDerivSlab(2, 0, aux[2], FB3);
        if (aux[3].on) pmlDerivSlab(0, 1, aux[3], FB1);
        if (aux[4].on) pmlDerivSlab(1, 1, aux[4], FB2);
        if (aux[5].on) pmlDerivSlab(2, 1, aux[5], FB3);
    }

    /* pml_free_surface_x/y (pml_freeSurface.cpp:12-235) */
    void pmlFreeSurfaceSlab(int a, int FLAG, Aux& X, int FB) {
        const long long stride = a == 0 ? 1 : nx_;
        const int nPML = su.d.nPML;
        int n0 = a == 0 ? nPML : nx, n1 = a == 1 ? nPML : ny;
        int nAxis = a == 0 ? nx : ny;
        int off = FLAG * (nAxis - nPML);
        int k0 = nz - 1;
        const std::vector<R>& M = a == 0 ? rdzdx : rdzdy;
        for (int j0 = 0; j0 < n1; ++j0)
            for (int i0 = 0; i0 < n0; ++i0) {
                int i = i0 + HALO + (a == 0 ? off : 0);
                int j = j0 + HALO + (a == 1 ? off : 0);
                int k = k0 + HALO;
                long long index = INDEX(i, j, k);
                long long s = (i + (long long)j * nx_) * 9;
                long long pos = i0 + (long long)j0 * n0 + (long long)k0 * n0 * n1;
                R mu = C(10, index), lambda = C(11, index);
                R dd = pd[a][a == 0 ? i : j];
                R zt_x = C(6, index), zt_y = C(7, index), zt_z = C(8, index);
                R Vx_ = L(W, 0, index, stride, FB) * dd, Vy_ = L(W, 1, index, stride, FB) * dd, Vz_ = L(W, 2, index, stride, FB) * dd;
                R Vx_zt = M[s + 0] * Vx_ + M[s + 1] * Vy_ + M[s + 2] * Vz_;
                R Vy_zt = M[s + 3] * Vx_ + M[s + 4] * Vy_ + M[s + 5] * Vz_;
                R Vz_zt = M[s + 6] * Vx_ + M[s + 7] * Vy_ + M[s + 8] * Vz_;
                R dv = (zt_x * Vx_zt + zt_y * Vy_zt + zt_z * Vz_zt) * lambda;
                R two = 2;
                R T[6];
                T[0] = dv + two * mu * (zt_x * Vx_zt);
                T[1] = dv + two * mu * (zt_y * Vy_zt);
                T[2] = dv + two * mu * (zt_z * Vz_zt);
                T[3] = (zt_y * Vx_zt + zt_x * Vy_zt) * mu;
                T[4] = (zt_z * Vx_zt + zt_x * Vz_zt) * mu;
                T[5] = (zt_z * Vy_zt + zt_y * Vz_zt) * mu;
                for (int v = 0; v < 6; ++v) X.h[pos * 9 + 3 + v] += T[v] * DT;
            }
    }

    void pmlFreeSurfaceDeriv(int FB1, int FB2) { /* pml_freeSurface.cpp:306-332 */
        if (aux[0].on) pmlFreeSurfaceSlab(0, 0, aux[0], FB1);
        if (aux[1].on) pmlFreeSurfaceSlab(1, 0, aux[1], FB2);
        if (aux[3].on) pmlFreeSurfaceSlab(0, 1, aux[3], FB1);
        if (aux[4].on) pmlFreeSurfaceSlab(1, 1, aux[4], FB2);
    }

    /* wave_rk0..3 (wave_rk.cpp:13-99), over the whole padded array */
    void waveRk(int irk) {
        const long long n = 9 * num;
        S* h = hW.data(); S* Wp = W.data(); S* t = tW.data(); S* m = mW.data();
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < n; ++q) {
            R hw = ld(h[q]);
            if (irk == 0) {
                R mw = ld(Wp[q]);
                R tw = mw + (R)beta1 * hw;
                R ww = mw + (R)alpha2 * hw;
                m[q] = st_(mw); t[q] = st_(tw); Wp[q] = st_(ww);
            } else if (irk == 1) {
                R tw = ld(t[q]), mw = ld(m[q]);
                tw += (R)beta2 * hw;
                R ww = mw + (R)alpha3 * hw;
                t[q] = st_(tw); Wp[q] = st_(ww);
            } else if (irk == 2) {
                R tw = ld(t[q]), mw = ld(m[q]);
                tw += (R)beta3 * hw;
                R ww = mw + hw;
                t[q] = st_(tw); Wp[q] = st_(ww);
            } else {
                R tw = ld(t[q]);
                Wp[q] = st_(tw + (R)beta4 * hw);
            }
        }
    }

    /* wave_pml_rk0..3 (pml_rk.cpp:14-70) */
    void pmlRk(int irk) {
        for (int s = 0; s < 6; ++s) {
            if (!aux[s].on) continue;
            Aux& X = aux[s];
            const long long n = 9 * X.n;
#pragma omp parallel for schedule(static)
            for (long long q = 0; q < n; ++q) {
                if (irk == 0) {
                    X.m[q] = X.a[q];
                    X.t[q] = X.m[q] + (R)beta1 * X.h[q];
                    X.a[q] = X.m[q] + (R)alpha2 * X.h[q];
                } else if (irk == 1) {
                    X.t[q] += (R)beta2 * X.h[q];
                    X.a[q] = X.m[q] + (R)alpha3 * X.h[q];
                } else if (irk == 2) {
                    X.t[q] += (R)beta3 * X.h[q];
                    X.a[q] = X.m[q] + X.h[q];
                } else {
                    X.a[q] = X.t[q] + (R)beta4 * X.h[q];
                }
            }
        }
    }

    /* sourceFunction (singleSource.cpp:29-56) at irk = 0 */
    static R sourceFunction(R rickerfc, int it, R DT) {
        R t = ((R)it + (R)0) * DT;
        R tdelay = (R)1.2f / rickerfc;
        R f0 = 2;
        R r = (R)PIf * rickerfc * (t - tdelay);
        R rr = r * r;
        R s = r * ((R)2 * rr - (R)3.0) * (R)std::exp((double)-rr) * f0 * (R)PIf * rickerfc;
        R M0 = (R)1e16;
        return s * M0;
    }

    /* load_smooth_source (singleSource.cpp:77-127) into W at irk = 0 */
    void loadPointSource() {
        const int nGauss = 3;
        if (!(SX >= HALO && SX < _nx && SY >= HALO && SY < _ny && SZ >= HALO && SZ < _nz)) return;
        R s = sourceFunction((R)fc, it, DT);
        const R DH = (R)su.d.DH;
        const double ra = nGauss * 0.5;
        R jacb = C(9, INDEX(SX, SY, SZ));
        for (int k = SZ - nGauss; k < SZ + nGauss + 1; ++k)
            for (int j = SY - nGauss; j < SY + nGauss + 1; ++j)
                for (int i = SX - nGauss; i < SX + nGauss + 1; ++i) {
                    long long index = INDEX(i, j, k);
                    auto G = [&](int t) { return (R)(std::exp(-((t / ra) * (t / ra))) / (ra * 1.772453850905516)); };
                    R amp = G(i - SX) * G(j - SY) * G(k - SZ);
                    amp = (R)(amp / 0.998125703461425);
                    R value = (R)-1 * s * amp / (jacb * (DH * DH * DH));
                    value *= (R)(su.Cs * (double)DT);
                    for (int v = 3; v <= 5; ++v) W[index + v * num] = st_(ld(W[index + v * num]) + value);
                }
    }

    /* addMomenteRate (addMoment.cpp:222-339) with NO_SOURCE_SMOOTH: interp + addSource1 */
    void addMomentRate() {
        if (!multiActive) return;
        R tmpT = ((R)it + (R)0) * DT;
        int srcIt = int(tmpT / (R)sdt);
        if (srcIt + 1 >= snt) { multiActive = false; return; }
        R t1 = (R)srcIt * (R)sdt;
        R t_weight = (tmpT - t1) / (R)sdt;
        for (long long p = 0; p < npts; ++p) {
            long long idx = srcIndex[p];
            for (int c = 0; c < 6; ++c) {
                R m0 = momentRate[((size_t)srcIt * npts + p) * 6 + c];
                R m1 = momentRate[((size_t)(srcIt + 1) * npts + p) * 6 + c];
                R slice = (m1 - m0) * t_weight + m0;
                W[idx + (3 + c) * num] = st_(ld(W[idx + (3 + c) * num]) - slice * DT);
            }
        }
    }

    /* storage_station (station.cpp:309-347) */
    void storageStation() {
        const int NT = su.d.NT;
        if (it >= NT) return;
        float C1 = (float)(1.0 / su.Cv), C2 = (float)(1.0 / su.Cs);
        for (int s = 0; s < nst; ++s) {
            long long index = INDEX(st[s * 3 + 0], st[s * 3 + 1], st[s * 3 + 2]);
            long long pos = it + (long long)s * NT;
            for (int v = 0; v < 9; ++v) stWave[pos * 9 + v] = (float)ST<S>::get(W[index + num * v]) * (v < 3 ? C1 : C2);
        }
    }

    /* compare_pgv (PGV.cpp:48-101) */
    void comparePGV() {
        const double c = 1.0 / su.Cv;
        int k = nz_ - HALO - 1;
        for (int j0 = 0; j0 < ny; ++j0)
            for (int i0 = 0; i0 < nx; ++i0) {
                long long index = INDEX(i0 + HALO, j0 + HALO, k);
                long long pos = i0 + (long long)j0 * nx;
                float Vx = (float)(ST<S>::get(W[index + num * 0]) * c);
                float Vy = (float)(ST<S>::get(W[index + num * 1]) * c);
                float Vz = (float)(ST<S>::get(W[index + num * 2]) * c);
                float Vh = std::sqrt(Vx * Vx + Vy * Vy);
                float V = std::sqrt(Vx * Vx + Vy * Vy + Vz * Vz);
                if (pgv[pos * 2 + 0] < Vh) pgv[pos * 2 + 0] = Vh;
                if (pgv[pos * 2 + 1] < V) pgv[pos * 2 + 1] = V;
            }
    }

    /* one iteration of the loop at propagate.cpp:142-198 */
    void step() {
        static const int FB[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                                     {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};
        int FB1 = FB[it % 8][0], FB2 = FB[it % 8][1], FB3 = FB[it % 8][2];
        if (srcMode == 1) loadPointSource();
        if (srcMode == 2) addMomentRate();
        for (int irk = 0; irk < 4; ++irk) {
            waveDeriv(FB1, FB2, FB3);
            if (su.d.free_surface) freeSurfaceDeriv(FB1, FB2, FB3);
            if (su.d.pml) pmlDeriv(FB1, FB2, FB3);
            if (su.d.pml && su.d.free_surface) pmlFreeSurfaceDeriv(FB1, FB2);
            waveRk(irk);
            if (su.d.pml) pmlRk(irk);
            FB1 *= -1; FB2 *= -1; FB3 *= -1;
        }
        if (nst > 0) storageStation();
        if (su.d.free_surface) comparePGV();
        ++it;
    }
};

struct Handle {
    int prec; /* 0 f64, 1 f32, 2 f16, 3 bf16 */
    void* sim;
};

template <class F>
int dispatch(Handle* h, F&& f) {
    switch (h->prec) {
        case 0: return f(*static_cast<Sim<double, double>*>(h->sim));
        case 1: return f(*static_cast<Sim<float, float>*>(h->sim));
        case 2: return f(*static_cast<Sim<half16, float>*>(h->sim));
        case 3: return f(*static_cast<Sim<bf16, float>*>(h->sim));
    }
    return CGFDM_EINVAL;
}

void scales(const cgfdm_desc& d, int prec, double& Cv, double& Cs, double& Crho) {
    bool low = prec >= 2;
    Cv = d.Cv > 0 ? d.Cv : (low ? 5.0e3 : 1.0);
    Cs = d.Cs > 0 ? d.Cs : (low ? 1.0e-6 : 1.0);
    Crho = d.Crho > 0 ? d.Crho : (low ? 1e6 : 1.0);
}

} // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

/* precision: 0 = FP64 oracle, 1 = FP32, 2 = FP16 storage, 3 = BF16 storage */
void* orc_create(const cgfdm_desc* d, int precision) {
    if (!d || d->NX < 1 || d->NY < 1 || d->NZ < 2 || precision < 0 || precision > 3) { g_err = "bad descriptor"; return nullptr; }
    if (d->pml && (d->nPML > d->NX || d->nPML > d->NY || d->nPML > d->NZ || d->nPML < 2)) { g_err = "nPML larger than the grid"; return nullptr; }
    Handle* h = new Handle();
    h->prec = precision;
    switch (precision) {
        case 0: h->sim = new Sim<double, double>(*d); break;
        case 1: h->sim = new Sim<float, float>(*d); break;
        case 2: h->sim = new Sim<half16, float>(*d); break;
        case 3: h->sim = new Sim<bf16, float>(*d); break;
    }
    return h;
}

void orc_destroy(void* hp) {
    Handle* h = static_cast<Handle*>(hp);
    if (!h) return;
    switch (h->prec) {
        case 0: delete static_cast<Sim<double, double>*>(h->sim); break;
        case 1: delete static_cast<Sim<float, float>*>(h->sim); break;
        case 2: delete static_cast<Sim<half16, float>*>(h->sim); break;
        case 3: delete static_cast<Sim<bf16, float>*>(h->sim); break;
    }
    delete h;
}

int orc_set_model(void* hp, const double* terrain, const float* vs, const float* vp, const float* rho) {
    Handle* h = static_cast<Handle*>(hp);
    return dispatch(h, [&](auto& s) {
        double Cv, Cs, Crho;
        scales(s.su.d, h->prec, Cv, Cs, Crho);
        s.setModel(terrain, vs, vp, rho, Cv, Cs, Crho);
        return 0;
    });
}

int orc_set_source_point(void* hp, int X, int Y, int Z, float fc) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        s.srcMode = 1; s.SX = X + HALO; s.SY = Y + HALO; s.SZ = Z + HALO; s.fc = fc;
        return 0;
    });
}

/* xyz global interior indices; rate [nt][npts][6]. Duplicates are merged (dealDuplicateIndex)
 * and the table is scaled by mu*Cv/(J*DH^3) with J, mu read from the stored CJM
 * (calculate_MomentRate, addMoment.cpp:11-65). */
int orc_set_source_table(void* hp, const int* xyz, const float* rate, long long npts, int nt, float dt) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        using R = std::decay_t<decltype(s.DT)>;
        std::map<long long, long long> uniq;
        std::vector<long long> idxOf(npts);
        for (long long p = 0; p < npts; ++p) {
            long long idx = s.INDEX(xyz[p * 3] + HALO, xyz[p * 3 + 1] + HALO, xyz[p * 3 + 2] + HALO);
            idxOf[p] = idx;
            uniq.emplace(idx, 0);
        }
        long long u = 0;
        for (auto& kv : uniq) kv.second = u++;
        s.npts = u; s.snt = nt; s.sdt = dt;
        s.srcIndex.assign(u, 0);
        for (auto& kv : uniq) s.srcIndex[kv.second] = kv.first;
        s.momentRate.assign((size_t)nt * u * 6, (R)0);
        for (long long p = 0; p < npts; ++p) {
            long long q = uniq[idxOf[p]];
            for (int t = 0; t < nt; ++t)
                for (int c = 0; c < 6; ++c) s.momentRate[((size_t)t * u + q) * 6 + c] += (R)rate[((size_t)t * npts + p) * 6 + c];
        }
        const R DH = (R)s.su.d.DH;
        for (long long q = 0; q < u; ++q) {
            R jacb = s.C(9, s.srcIndex[q]), mu = s.C(10, s.srcIndex[q]);
            R V = jacb * DH * DH * DH;
            V = (R)1.0 / V;
            V = V * (R)s.su.Cv;
            for (int t = 0; t < nt; ++t)
                for (int c = 0; c < 6; ++c) s.momentRate[((size_t)t * u + q) * 6 + c] *= mu * V;
        }
        s.srcMode = 2; s.multiActive = true;
        return 0;
    });
}

int orc_set_stations(void* hp, const int* xyz, int n) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        s.st.assign(xyz, xyz + 3 * n);
        for (int q = 0; q < 3 * n; ++q) s.st[q] += HALO;
        s.nst = n;
        s.stWave.assign((size_t)n * s.su.d.NT * 9, 0.f);
        return 0;
    });
}

int orc_step(void* hp, int nsteps) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        if (!s.modelSet) { g_err = "set_model first"; return (int)CGFDM_ESTATE; }
        if (s.pgv.empty()) s.pgv.assign((size_t)s.nx * s.ny * 2, 0.f);
        for (int q = 0; q < nsteps; ++q) s.step();
        return 0;
    });
}

int orc_get_stations(void* hp, float* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        std::copy(s.stWave.begin(), s.stWave.end(), out);
        return 0;
    });
}

int orc_get_pgv(void* hp, float* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        if (s.pgv.empty()) s.pgv.assign((size_t)s.nx * s.ny * 2, 0.f);
        std::copy(s.pgv.begin(), s.pgv.end(), out);
        return 0;
    });
}

/* interior wavefield, 9 x NZ x NY x NX, storage units */
int orc_get_state(void* hp, double* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        using S = std::decay_t<decltype(s.W[0])>;
        long long n = (long long)s.nx * s.ny * s.nz;
        for (int v = 0; v < 9; ++v)
            for (int k = 0; k < s.nz; ++k)
                for (int j = 0; j < s.ny; ++j)
                    for (int i = 0; i < s.nx; ++i)
                        out[v * n + i + (long long)j * s.nx + (long long)k * s.nx * s.ny] =
                            ST<S>::get(s.W[(size_t)v * s.num + s.INDEX(i + HALO, j + HALO, k + HALO)]);
        return 0;
    });
}

int orc_set_state(void* hp, const double* in) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        using S = std::decay_t<decltype(s.W[0])>;
        long long n = (long long)s.nx * s.ny * s.nz;
        for (int v = 0; v < 9; ++v)
            for (int k = 0; k < s.nz; ++k)
                for (int j = 0; j < s.ny; ++j)
                    for (int i = 0; i < s.nx; ++i)
                        s.W[(size_t)v * s.num + s.INDEX(i + HALO, j + HALO, k + HALO)] =
                            ST<S>::put(in[v * n + i + (long long)j * s.nx + (long long)k * s.nx * s.ny]);
        return 0;
    });
}

/* packed metric/medium in FP64 before rounding (13 x num) */
int orc_get_cjm64(void* hp, double* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        std::copy(s.su.cjm.begin(), s.su.cjm.end(), out);
        return 0;
    });
}

/* rdzdx then rdzdy, each nx_*ny_*9 */
int orc_get_rdz(void* hp, double* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        std::copy(s.su.rdzdx.begin(), s.su.rdzdx.end(), out);
        std::copy(s.su.rdzdy.begin(), s.su.rdzdy.end(), out + s.su.rdzdx.size());
        return 0;
    });
}

/* alpha, 1/beta, d/beta for x (nx_), y (ny_), z (nz_), concatenated as in
 * allocPMLParameter (init_pml_para.cpp:21-51): [ax bx dx ay by dy az bz dz] */
int orc_get_pml(void* hp, double* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        size_t o = 0;
        for (int a = 0; a < 3; ++a) {
            for (double v : s.su.pa[a]) out[o++] = v;
            for (double v : s.su.pb[a]) out[o++] = v;
            for (double v : s.su.pd[a]) out[o++] = v;
        }
        return 0;
    });
}

int orc_current_step(void* hp) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) { return s.it; });
}

/* One stage's increment h_W (deriv, free surface, PML with the current Aux, PML free surface)
 * for the given FB signs, interior points, 9 x NZ x NY x NX; W and Aux are left unchanged. */
int orc_debug_rhs(void* hp, int FB1, int FB2, int FB3, double* out) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        using S = std::decay_t<decltype(s.W[0])>;
        if (!s.modelSet) { g_err = "set_model first"; return (int)CGFDM_ESTATE; }
        s.waveDeriv(FB1, FB2, FB3);
        if (s.su.d.free_surface) s.freeSurfaceDeriv(FB1, FB2, FB3);
        if (s.su.d.pml) s.pmlDeriv(FB1, FB2, FB3);
        if (s.su.d.pml && s.su.d.free_surface) s.pmlFreeSurfaceDeriv(FB1, FB2);
        long long n = (long long)s.nx * s.ny * s.nz;
        for (int v = 0; v < 9; ++v)
            for (int k = 0; k < s.nz; ++k)
                for (int j = 0; j < s.ny; ++j)
                    for (int i = 0; i < s.nx; ++i)
                        out[v * n + i + (long long)j * s.nx + (long long)k * s.nx * s.ny] =
                            ST<S>::get(s.hW[(size_t)v * s.num + s.INDEX(i + HALO, j + HALO, k + HALO)]);
        return 0;
    });
}

/* The auxiliary increments h_Aux of PML slab s (0..5: x1 y1 z1 x2 y2 z2, struct.h:565-574) as the
 * last orc_debug_rhs left them (pml_deriv.cpp:176-207 plus pml_freeSurface.cpp:81-116): 9 values
 * per slab point, [pos][9] with pos = i0 + j0*n0 + k0*n0*n1 (pml_deriv.cpp:164, 313, 464).
 * Returns the number of slab points (0 if the slab is absent) or a negative status if cap is
 * smaller. */
long long orc_debug_aux_rhs(void* hp, int s, double* out, long long cap) {
    Handle* h = static_cast<Handle*>(hp);
    long long res = 0;
    int rc = dispatch(h, [&](auto& sim) {
        if (s < 0 || s > 5) { g_err = "slab index out of range"; return (int)CGFDM_EINVAL; }
        const auto& X = sim.aux[s];
        if (!X.on) { res = 0; return 0; }
        if (cap < X.n * 9) { g_err = "output too small"; return (int)CGFDM_EINVAL; }
        for (long long q = 0; q < X.n * 9; ++q) out[q] = (double)X.h[q];
        res = X.n;
        return 0;
    });
    return rc ? rc : res;
}

/* Applies only this step's source injection into W (the first action of propagate.cpp:145-150)
 * and leaves the step counter unchanged. */
int orc_debug_source(void* hp) {
    return dispatch(static_cast<Handle*>(hp), [&](auto& s) {
        if (s.srcMode == 1) s.loadPointSource();
        if (s.srcMode == 2) s.addMomentRate();
        return 0;
    });
}

/* ---- pieces exposed for known-answer tests ---- */

/* operator L applied along a uniform 1-D line (axis.h:129-133), FP64, same literals.
 * out[i] for i in [3, n-3) */
void orc_apply_L_line(const double* f, int n, int FB, double rDH, double* out) {
    for (int i = 3; i < n - 3; ++i)
        out[i] = FB * (af_1 * f[i - FB] + af0 * f[i] + af1 * f[i + FB] + af2 * f[i + 2 * FB] + af3 * f[i + 3 * FB]) * rDH;
}

/* one RK stage (wave_rk.cpp:13-99) on plain FP64 arrays of length n */
void orc_wave_rk(int irk, const double* h, double* W, double* t, double* m, long long n) {
    for (long long q = 0; q < n; ++q) {
        double hw = h[q];
        if (irk == 0) { m[q] = W[q]; t[q] = m[q] + (double)beta1 * hw; W[q] = m[q] + (double)alpha2 * hw; }
        else if (irk == 1) { t[q] += (double)beta2 * hw; W[q] = m[q] + (double)alpha3 * hw; }
        else if (irk == 2) { t[q] += (double)beta3 * hw; W[q] = m[q] + hw; }
        else W[q] = t[q] + (double)beta4 * hw;
    }
}

/* Ricker source time function at step it (singleSource.cpp:29-56), FP64 */
double orc_source_function(double fc, int it, double DT) { return Sim<double, double>::sourceFunction(fc, it, DT); }

/* 7^3 Gaussian weights of load_smooth_source, normalised (singleSource.cpp:428-434) */
void orc_gauss_weights(double* out343) {
    const double ra = 1.5;
    int q = 0;
    for (int k = -3; k <= 3; ++k)
        for (int j = -3; j <= 3; ++j)
            for (int i = -3; i <= 3; ++i) {
                auto G = [&](int t) { return std::exp(-((t / ra) * (t / ra))) / (ra * 1.772453850905516); };
                out343[q++] = G(i) * G(j) * G(k) / 0.998125703461425;
            }
}

/* moment-rate components (xx,yy,zz,xy,xz,yz) of solveMomentRate (multiSource.cpp:478-561), FP64 */
void orc_moment_tensor(double strike, double dip, double rake, double area, double rate, double* out6) {
    const double D2R = (double)PIf / 180.0;
    double s = strike * D2R, d = dip * D2R, r = rake * D2R;
    double M11 = -(std::sin(d) * std::cos(r) * std::sin(2.0 * s) + std::sin(2.0 * d) * std::sin(r) * std::sin(s) * std::sin(s));
    double M22 = std::sin(d) * std::cos(r) * std::sin(2.0 * s) - std::sin(2.0 * d) * std::sin(r) * std::cos(s) * std::cos(s);
    double M33 = -(M11 + M22);
    double M12 = std::sin(d) * std::cos(r) * std::cos(2.0 * s) + 0.5 * std::sin(2.0 * d) * std::sin(r) * std::sin(2.0 * s);
    double M13 = -(std::cos(d) * std::cos(r) * std::cos(s) + std::cos(2.0 * d) * std::sin(r) * std::sin(s));
    double M23 = -(std::cos(d) * std::cos(r) * std::sin(s) - std::cos(2.0 * d) * std::sin(r) * std::cos(s));
    out6[0] = M22 * area * rate;
    out6[1] = M11 * area * rate;
    out6[2] = M33 * area * rate;
    out6[3] = M12 * area * rate;
    out6[4] = -M23 * area * rate;
    out6[5] = -M13 * area * rate;
}

/* fp16 / bf16 rounding used by the storage emulation (round to nearest even) */
void orc_round_f16(const float* in, float* out, long long n) { for (long long q = 0; q < n; ++q) out[q] = h2f_bits(f2h_bits(in[q])); }
void orc_round_bf16(const float* in, float* out, long long n) { for (long long q = 0; q < n; ++q) out[q] = bf2f_bits(f2bf_bits(in[q])); }

int orc_num_threads(void) { return omp_get_max_threads(); }

} /* extern "C" */
