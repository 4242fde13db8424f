// setup.hip — FP64 construction of the packed metric/Jacobian/medium array on the GPU.
//
// Coordinates are never stored: the reference builds them as x = (I-3-centerX)*DH,
// y = (J-3-centerY)*DH and either a flat z (coord.cpp:106-142) or the terrain-following
// stretch z = -|D| + (h+|D|)/(NZ-1)*(K-3) (coord.cpp:255-307, terrain.cpp:930-937), so each
// thread evaluates the coordinates of its stencil neighbours on the fly from the 2-D
// height map. All arithmetic is FP64; the result is rounded once to the storage type.
#include "common.hpp"

namespace cg {

__device__ __forceinline__ void coord_at(const SetupParams& q, int i, int j, int k, double c[3]) {
#pragma clang fp contract(off)
    c[0] = (i - HALO) * q.DH - q.centerX * q.DH;
    c[1] = (j - HALO) * q.DH - q.centerY * q.DH;
    if (q.use_terrain) {
        double h = q.terrain[i + (long long)j * q.nx_];
        double D = fabs(q.depth);
        double DZ = (h + D) / double(q.nz - 1);
        c[2] = -D + DZ * (k - HALO);
    } else {
        c[2] = (k - HALO + 1) * q.DH - q.nz * q.DH;
    }
}

// Averaged forward/backward MacCormack derivative of the coordinates and the contravariant
// metric as the adjugate over J (reference solve_con_jac, contravariant.cpp:71-171). Evaluated
// without multiply-add contraction, as the reference's CPU arithmetic: the forward and backward
// derivatives of a coordinate that is constant along an axis then cancel bit for bit, so
// x_eta, x_zeta, y_xi, y_zeta and the metric terms built from them (xi_y, xi_z, eta_x, eta_z) are
// exactly zero on these terrain-following grids, which the stage tile kernels rely on.
__device__ void metric_at(const SetupParams& q, int i, int j, int k, double m[9], double& J) {
#pragma clang fp contract(off)
    const double rDH = 1.0 / q.DH;
    const double a[5] = {(double)af_1, (double)af0, (double)af1, (double)af2, (double)af3};
    double der[3][3];  // der[var][dir]
    for (int dir = 0; dir < 3; ++dir) {
        double cp[5][3], cm[5][3];
        for (int s = 0; s < 5; ++s) {
            int o = s - 1;  // offsets -1..3
            int ip = i + (dir == 0 ? o : 0), jp = j + (dir == 1 ? o : 0), kp = k + (dir == 2 ? o : 0);
            int im = i - (dir == 0 ? o : 0), jm = j - (dir == 1 ? o : 0), km = k - (dir == 2 ? o : 0);
            coord_at(q, ip, jp, kp, cp[s]);
            coord_at(q, im, jm, km, cm[s]);
        }
        for (int v = 0; v < 3; ++v) {
            double lp = (a[0] * cp[0][v] + a[1] * cp[1][v] + a[2] * cp[2][v] + a[3] * cp[3][v] + a[4] * cp[4][v]) * rDH;
            double lm = -(a[0] * cm[0][v] + a[1] * cm[1][v] + a[2] * cm[2][v] + a[3] * cm[3][v] + a[4] * cm[4][v]) * rDH;
            der[v][dir] = 0.5 * (lp + lm);
        }
    }
    const double x_xi = der[0][0], x_et = der[0][1], x_zt = der[0][2];
    const double y_xi = der[1][0], y_et = der[1][1], y_zt = der[1][2];
    const double z_xi = der[2][0], z_et = der[2][1], z_zt = der[2][2];
    J = x_xi * y_et * z_zt + x_et * y_zt * z_xi + x_zt * y_xi * z_et - x_zt * y_et * z_xi - x_et * y_xi * z_zt -
        x_xi * z_et * y_zt;
    const double Ji = 1.0 / J;
    m[0] = (y_et * z_zt - y_zt * z_et) * Ji;
    m[1] = (x_zt * z_et - x_et * z_zt) * Ji;
    m[2] = (x_et * y_zt - x_zt * y_et) * Ji;
    m[3] = (y_zt * z_xi - y_xi * z_zt) * Ji;
    m[4] = (x_xi * z_zt - x_zt * z_xi) * Ji;
    m[5] = (x_zt * y_xi - x_xi * y_zt) * Ji;
    m[6] = (y_xi * z_et - y_et * z_xi) * Ji;
    m[7] = (x_et * z_xi - x_xi * z_et) * Ji;
    m[8] = (x_xi * y_et - x_et * y_xi) * Ji;
}

// mu, lambda, buoyancy with the storage scaling folded in (reference calculate_medium,
// CJM.cpp:46-73): rho' = rho*Cs/Cv, mu = rho' Vs^2, lambda = rho'(Vp^2 - 2Vs^2),
// buoyancy = 1/(rho' Crho).
__device__ __forceinline__ void medium_at(const SetupParams& q, float vs, float vp, float rho, double& mu,
                                          double& lambda, double& buoy) {
    double r = (double)rho * (1.0 * q.Cs / q.Cv);
    double Vs = vs, Vp = vp;
    mu = r * (Vs * Vs);
    lambda = r * (Vp * Vp - 2.0 * Vs * Vs);
    buoy = 1.0 / r / q.Crho;
}

// Medium at interior point (i0, j0, k0): from the 3-D arrays when given, otherwise from the layer
// table by the reference's rule for a layered model following the terrain
// (generate_1D_medium with LayeredStructureTerrain, medium.cpp:84-216): start from the
// homogeneous default, then for every interface F the top layer is taken above -H[0] + h, layer
// F-1 between -H[F-1] + h and -H[F] + h, layer F below -H[F] + h (1 m tolerance, float compares),
// the last match winning.
__device__ __forceinline__ void medium_sample(const SetupParams& q, const MediumSrc& ms, int i0, int j0, int k0,
                                              float& vs, float& vp, float& rho) {
    if (ms.vs) {
        const long long t = i0 + (long long)j0 * q.nx + (long long)k0 * q.nx * q.ny;
        vs = ms.vs[t];
        vp = ms.vp[t];
        rho = ms.rho[t];
        return;
    }
    double c[3];
    coord_at(q, i0 + HALO, j0 + HALO, k0 + HALO, c);
    const float cz = (float)c[2];
    const double ter = q.use_terrain ? q.terrain[(i0 + HALO) + (long long)(j0 + HALO) * q.nx_] : 0.0;
    const float* H = ms.layers;
    const float* VS = H + ms.nl;
    const float* VP = VS + ms.nl;
    const float* RH = VP + ms.nl;
    vs = 3464.0f;
    vp = 6000.0f;
    rho = 2670.0f;
    const float tol = 1.0f;
    for (int F = 1; F < ms.nl; ++F) {
        const float h0 = (float)(-(double)H[0] + ter), h1 = (float)(-(double)H[F - 1] + ter),
                    h2 = (float)(-(double)H[F] + ter);
        if (cz > h0 - tol) { vs = VS[0]; vp = VP[0]; rho = RH[0]; }
        if (h1 - tol > cz && cz > h2 - tol) { vs = VS[F - 1]; vp = VP[F - 1]; rho = RH[F - 1]; }
        if (h2 - tol > cz) { vs = VS[F]; vp = VP[F]; rho = RH[F]; }
    }
}

template <class S>
__global__ void k_medium(SetupParams q, MediumSrc ms, S* __restrict__ cjm) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)q.nx * q.ny * q.nz;
    if (t >= n) return;
    int i = (int)(t % q.nx), j = (int)((t / q.nx) % q.ny), k = (int)(t / ((long long)q.nx * q.ny));
    float vs, vp, rho;
    medium_sample(q, ms, i, j, k, vs, vp, rho);
    double mu, lambda, buoy;
    medium_at(q, vs, vp, rho, mu, lambda, buoy);
    cjm[q.lay.atc(10, i + HALO, j + HALO, k + HALO)] = SV<S>::st((float)mu);
    cjm[q.lay.atc(11, i + HALO, j + HALO, k + HALO)] = SV<S>::st((float)lambda);
    cjm[q.lay.atc(12, i + HALO, j + HALO, k + HALO)] = SV<S>::st((float)buoy);
}

// Largest Vp of the medium (for the CFL check), as an atomic maximum on the float bits.
__global__ void k_vpmax(SetupParams q, MediumSrc ms, unsigned* out) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)q.nx * q.ny * q.nz;
    float vp = 0.f;
    if (t < n) {
        int i = (int)(t % q.nx), j = (int)((t / q.nx) % q.ny), k = (int)(t / ((long long)q.nx * q.ny));
        float vs, rho;
        medium_sample(q, ms, i, j, k, vs, vp, rho);
    }
    for (int off = 32; off > 0; off >>= 1) vp = fmaxf(vp, __shfl_down(vp, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(vp));
}

// Per-column zeta: along a column of these grids z = -|D| + (h + |D|)/(NZ-1)*(k-3) is linear in
// k and x, y do not depend on k, so z_zeta is constant, z_xi and z_eta are (k-3) times a
// per-column value, and zeta_x = -z_xi/(x_xi z_zeta), zeta_y = -z_eta/(y_eta z_zeta) are
// (k-3) times a per-column value while zeta_z = 1/z_zeta is constant. zcol = [3][ny_][nx_]
// (FP32) holds those values, from the FP64 metric of the top interior layer, for the interior
// columns (0 elsewhere); the tile kernels form zeta from them instead of reading it per point.
__global__ void k_zcol(SetupParams q, float* __restrict__ zcol) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long pl = (long long)q.nx_ * q.ny_;
    if (t >= pl) return;
    const int i = (int)(t % q.nx_), j = (int)(t / q.nx_);
    float a = 0.f, b = 0.f, g = 0.f;
    if (i >= HALO && i < q.nx + HALO && j >= HALO && j < q.ny + HALO) {
        const int kr = HALO + q.nz - 1;  // top interior layer, the largest k - 3
        double m[9], J;
        metric_at(q, i, j, kr, m, J);
        a = (float)(m[6] / (double)(kr - HALO));
        b = (float)(m[7] / (double)(kr - HALO));
        g = (float)m[8];
    }
    zcol[t] = a;
    zcol[pl + t] = b;
    zcol[2 * pl + t] = g;
}

// Metric and J on the interior and, with a free surface, the mirrored values above it
// (solve_coordinate_on_free_surface, contravariant.cpp:393-408: value at 2(_nz-1)-k). Each
// interior point also checks that its zeta matches the per-column form (k_zcol) to 1e-5 of its
// size and raises the int at m0 + 4 where it does not.
template <class S>
__global__ void k_metric(SetupParams q, S* __restrict__ cjm, const float* __restrict__ zcol, void* m0) {
    const int _nz = q.nz_ - HALO;
    const int kTop = q.fs ? q.nz_ : _nz;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)q.nx * q.ny * (kTop - HALO);
    if (t >= n) return;
    int i = (int)(t % q.nx) + HALO, j = (int)((t / q.nx) % q.ny) + HALO;
    int k = (int)(t / ((long long)q.nx * q.ny)) + HALO;
    int kk = k < _nz ? k : 2 * (_nz - 1) - k;
    double m[9], J;
    metric_at(q, i, j, kk, m, J);
    for (int c = 0; c < 9; ++c) cjm[q.lay.atc(c, i, j, k)] = SV<S>::st((float)m[c]);
    cjm[q.lay.atc(9, i, j, k)] = SV<S>::st((float)J);
    if (k < _nz) {
        const long long pl = (long long)q.nx_ * q.ny_, c0 = i + (long long)j * q.nx_;
        const double kf = (double)(k - HALO), gz = zcol[2 * pl + c0];
        const double ex = fabs(m[6] - zcol[c0] * kf), ey = fabs(m[7] - zcol[pl + c0] * kf), ez = fabs(m[8] - gz);
        const double tol = 1e-5 * (fabs(m[6]) + fabs(m[7]) + fabs(m[8]));
        if (!(ex <= tol && ey <= tol && ez <= tol))
            __builtin_amdgcn_raw_buffer_store_b32(1u, make_rsrc((const int*)m0, 4), 4, 0, 0);
    }
}

// Uniform first metric pair. The metric array's first variable pair is (xi_x, eta_y). On the
// terrain-following grids built here x depends on i only and y on j only, so both are constant
// up to FP64 rounding far below the storage precision, and in practice every interior point
// stores the same bits. m0 (16 bytes) = {int uniform, int unused, the stored pair of point
// (3, 3, 3)}: k_pair0_ref copies the pair and sets uniform = 1, k_pair0_check clears it at any
// interior point whose stored pair differs. The stage kernels read the pair from m0 while
// uniform is set, instead of streaming it from the metric array.
template <class S> __device__ __forceinline__ unsigned raw_bits(S x) {
    if constexpr (sizeof(S) == 4) return __builtin_bit_cast(unsigned, x);
    else return __builtin_bit_cast(unsigned short, x);
}

template <class S>
__global__ void k_pair0_ref(SetupParams q, const S* __restrict__ cjm, void* m0) {
    if (threadIdx.x != 0) return;
    const long long o = q.lay.at(13, 0, HALO, HALO, HALO);  // slots 0 and 1 are adjacent
    const Rsrc r = make_rsrc((const int*)m0, 4);
    if constexpr (sizeof(S) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(raw_bits(cjm[o]), r, 8, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(raw_bits(cjm[o + 1]), r, 12, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)raw_bits(cjm[o]), r, 8, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)raw_bits(cjm[o + 1]), r, 10, 0, 0);
    }
    __builtin_amdgcn_raw_buffer_store_b32(1u, r, 0, 0, 0);
}

template <class S>
__global__ void k_pair0_check(SetupParams q, const S* __restrict__ cjm, const void* m0) {
    const long long n = (long long)q.nx * q.ny * q.nz;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % q.nx) + HALO, j = (int)((t / q.nx) % q.ny) + HALO;
    const int k = (int)(t / ((long long)q.nx * q.ny)) + HALO;
    const long long o = q.lay.at(13, 0, i, j, k);
    const S* ref = (const S*)((const char*)m0 + 8);
    if (raw_bits(cjm[o]) != raw_bits(ref[0]) || raw_bits(cjm[o + 1]) != raw_bits(ref[1]))
        __builtin_amdgcn_raw_buffer_store_b32(0u, make_rsrc((const int*)m0, 4), 0, 0, 0);
}

// Free-surface matrices -rDZ^-1 DX and -rDZ^-1 DY at the surface layer
// (contravariant.cpp:257-351). chi is the reference's 2*lambda + mu (contravariant.cpp:266) unless
// the descriptor asks for lambda + 2*mu, the coefficient of the traction-free condition.
__global__ void k_rdz(SetupParams q, MediumSrc ms, float* __restrict__ rdz) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)q.nx * q.ny) return;
    int i0 = (int)(t % q.nx), j0 = (int)(t / q.nx);
    int i = i0 + HALO, j = j0 + HALO, k = q.nz_ - HALO - 1;
    double m[9], J, mu, lambda, buoy;
    metric_at(q, i, j, k, m, J);
    float vs, vp, rho;
    medium_sample(q, ms, i0, j0, q.nz - 1, vs, vp, rho);
    medium_at(q, vs, vp, rho, mu, lambda, buoy);
    const double xi_x = m[0], xi_y = m[1], xi_z = m[2], et_x = m[3], et_y = m[4], et_z = m[5];
    const double zt_x = m[6], zt_y = m[7], zt_z = m[8];
    const double chi = q.chi_mode ? lambda + 2.0 * mu : 2.0 * lambda + mu, lam_mu = lambda + mu;
    double DZ[9], DX[9], DY[9], rD[9];
    DZ[0] = chi * zt_x * zt_x + mu * (zt_y * zt_y + zt_z * zt_z);
    DZ[1] = lam_mu * zt_x * zt_y;
    DZ[2] = lam_mu * zt_x * zt_z;
    DZ[3] = DZ[1];
    DZ[4] = chi * zt_y * zt_y + mu * (zt_x * zt_x + zt_z * zt_z);
    DZ[5] = lam_mu * zt_y * zt_z;
    DZ[6] = DZ[2];
    DZ[7] = DZ[5];
    DZ[8] = chi * zt_z * zt_z + mu * (zt_x * zt_x + zt_y * zt_y);
    const double det = DZ[0] * DZ[4] * DZ[8] + DZ[1] * DZ[5] * DZ[6] + DZ[2] * DZ[7] * DZ[3] - DZ[2] * DZ[4] * DZ[6] -
                       DZ[1] * DZ[3] * DZ[8] - DZ[0] * DZ[7] * DZ[5];
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
    rD[0] = (DZ[4] * DZ[8] - DZ[5] * DZ[7]) / det;
    rD[1] = (-DZ[3] * DZ[8] + DZ[5] * DZ[6]) / det;
    rD[2] = (DZ[3] * DZ[7] - DZ[4] * DZ[6]) / det;
    rD[3] = (-DZ[1] * DZ[8] + DZ[2] * DZ[7]) / det;
    rD[4] = (DZ[0] * DZ[8] - DZ[2] * DZ[6]) / det;
    rD[5] = (-DZ[0] * DZ[7] + DZ[1] * DZ[6]) / det;
    rD[6] = (DZ[1] * DZ[5] - DZ[2] * DZ[4]) / det;
    rD[7] = (-DZ[0] * DZ[5] + DZ[2] * DZ[3]) / det;
    rD[8] = (DZ[0] * DZ[4] - DZ[1] * DZ[3]) / det;
    const long long nxy = (long long)q.nx_ * q.ny_;
    const long long s = (i + (long long)j * q.nx_) * 9;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            rdz[s + r * 3 + c] = (float)(-(rD[r * 3 + 0] * DX[c] + rD[r * 3 + 1] * DX[3 + c] + rD[r * 3 + 2] * DX[6 + c]));
            rdz[nxy * 9 + s + r * 3 + c] =
                (float)(-(rD[r * 3 + 0] * DY[c] + rD[r * 3 + 1] * DY[3 + c] + rD[r * 3 + 2] * DY[6 + c]));
        }
}

// Smallest point-to-plane distance among the 8 neighbour triangles of every interior point
// (reference calculate_DH_Range / distance_point2plane, calc_CFL.cpp:98-170), reduced with an
// atomic minimum on the bit pattern of a positive double.
__global__ void k_hmin(SetupParams q, unsigned long long* out) {
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long n = (long long)q.nx * q.ny * q.nz;
    double hmin = 1.0e20;
    if (t < n) {
        int i = (int)(t % q.nx) + HALO, j = (int)((t / q.nx) % q.ny) + HALO;
        int k = (int)(t / ((long long)q.nx * q.ny)) + HALO;
        double P[3], A[3], B[3], C[3];
        coord_at(q, i, j, k, P);
        for (int ii = -1; ii <= 1; ii += 2)
            for (int jj = -1; jj <= 1; jj += 2)
                for (int kk = -1; kk <= 1; kk += 2) {
                    coord_at(q, i - ii, j, k, A);
                    coord_at(q, i, j - jj, k, B);
                    coord_at(q, i, j, k - kk, C);
                    double AB[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
                    double AC[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
                    double nn[3] = {AB[1] * AC[2] - AB[2] * AC[1], AB[2] * AC[0] - AB[0] * AC[2],
                                    AB[0] * AC[1] - AB[1] * AC[0]};
                    double nd = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                    double d = fabs(((P[0] - A[0]) * nn[0] + (P[1] - A[1]) * nn[1] + (P[2] - A[2]) * nn[2]) / nd);
                    hmin = fmin(hmin, d);
                }
    }
    for (int off = 32; off > 0; off >>= 1) hmin = fmin(hmin, __shfl_down(hmin, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMin(out, (unsigned long long)__double_as_longlong(hmin));
}

static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int launch_setup_medium(int prec, const SetupParams& q, const MediumSrc& ms, void* cjm, hipStream_t s) {
    long long n = (long long)q.nx * q.ny * q.nz;
    const int bs = 256;
    switch (prec) {
        case P_F32: k_medium<float><<<blocks_for(n, bs), bs, 0, s>>>(q, ms, (float*)cjm); break;
        case P_F16: k_medium<_Float16><<<blocks_for(n, bs), bs, 0, s>>>(q, ms, (_Float16*)cjm); break;
        case P_BF16: k_medium<__bf16><<<blocks_for(n, bs), bs, 0, s>>>(q, ms, (__bf16*)cjm); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_vpmax(const SetupParams& q, const MediumSrc& ms, unsigned* out, hipStream_t s) {
    long long n = (long long)q.nx * q.ny * q.nz;
    k_vpmax<<<blocks_for(n, 256), 256, 0, s>>>(q, ms, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_setup_metric(int prec, const SetupParams& q, void* cjm, float* zcol, void* m0, hipStream_t s) {
    const int kTop = q.fs ? q.nz_ : q.nz_ - HALO;
    long long n = (long long)q.nx * q.ny * (kTop - HALO);
    const int bs = 256;
    k_zcol<<<blocks_for((long long)q.nx_ * q.ny_, bs), bs, 0, s>>>(q, zcol);
    switch (prec) {
        case P_F32: k_metric<float><<<blocks_for(n, bs), bs, 0, s>>>(q, (float*)cjm, zcol, m0); break;
        case P_F16: k_metric<_Float16><<<blocks_for(n, bs), bs, 0, s>>>(q, (_Float16*)cjm, zcol, m0); break;
        case P_BF16: k_metric<__bf16><<<blocks_for(n, bs), bs, 0, s>>>(q, (__bf16*)cjm, zcol, m0); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_pair0_check(int prec, const SetupParams& q, const void* cjm, void* m0, hipStream_t s) {
    const long long n = (long long)q.nx * q.ny * q.nz;
    const int bs = 256;
    switch (prec) {
        case P_F32:
            k_pair0_ref<float><<<1, 64, 0, s>>>(q, (const float*)cjm, m0);
            k_pair0_check<float><<<blocks_for(n, bs), bs, 0, s>>>(q, (const float*)cjm, m0);
            break;
        case P_F16:
            k_pair0_ref<_Float16><<<1, 64, 0, s>>>(q, (const _Float16*)cjm, m0);
            k_pair0_check<_Float16><<<blocks_for(n, bs), bs, 0, s>>>(q, (const _Float16*)cjm, m0);
            break;
        case P_BF16:
            k_pair0_ref<__bf16><<<1, 64, 0, s>>>(q, (const __bf16*)cjm, m0);
            k_pair0_check<__bf16><<<blocks_for(n, bs), bs, 0, s>>>(q, (const __bf16*)cjm, m0);
            break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_setup_rdz(const SetupParams& q, const MediumSrc& ms, float* rdz, hipStream_t s) {
    long long n = (long long)q.nx * q.ny;
    k_rdz<<<blocks_for(n, 256), 256, 0, s>>>(q, ms, rdz);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_hmin(const SetupParams& q, unsigned long long* out, hipStream_t s) {
    long long n = (long long)q.nx * q.ny * q.nz;
    k_hmin<<<blocks_for(n, 256), 256, 0, s>>>(q, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace cg
