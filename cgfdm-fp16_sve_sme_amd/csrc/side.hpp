// side.hpp — device bodies of the per-step side work (sources, stations, PGV), shared by the
// single-domain kernels (misc.hip) and the multi-domain kernels (batch.hip) so both paths do the
// same arithmetic on every point. The step counter `it` is an argument: the source time, the
// table interpolation weights and the station column are derived from it on the device.
#pragma once
#include "common.hpp"

namespace cg {

// Ricker moment-rate value at t = it*DT, M0 = 1e16 (sourceFunction at irk = 0,
// singleSource.cpp:29-56), with the reference's float/double mix.
__device__ __forceinline__ float ricker_value(float fc, int it, float DT) {
#pragma clang fp contract(off)
    const float t = ((float)it + 0.0f) * DT;
    const float tdelay = 1.2f / fc;
    const float f0 = 2;
    const float r = PIf * fc * (t - tdelay);
    const float rr = r * r;
    const float s = (float)(r * (2.0f * rr - 3.0) * (double)expf(-rr) * f0 * PIf * fc);
    return s * 1e16f;
}

// Point q (0..342) of the 7^3 Gaussian-smoothed explosive source: adds sval*coef[q] to Txx, Tyy,
// Tzz of W before stage 0 (load_smooth_source, singleSource.cpp:77-127). Points in the halo are
// also added to the two stage buffers, whose halos must equal W's (they are never recomputed).
template <class S>
__device__ __forceinline__ void point_source_at(S* A, S* B, S* C, const Layout& L, int SX, int SY, int SZ,
                                                const float* __restrict__ coef, float sval, int q) {
#pragma clang fp contract(off)
    const int di = q % 7 - 3, dj = (q / 7) % 7 - 3, dk = q / 49 - 3;
    const int i = SX + di, j = SY + dj, k = SZ + dk;
    const float value = sval * coef[q];
    const bool halo = i < HALO || i >= L.nx_ - HALO || j < HALO || j >= L.ny_ - HALO || k < HALO || k >= L.nz_ - HALO;
    for (int v = 3; v <= 5; ++v) {
        const long long p = L.at(9, v, i, j, k);
        A[p] = SV<S>::st(SV<S>::ld(A[p]) + value);
        if (halo) {
            B[p] = SV<S>::st(SV<S>::ld(B[p]) + value);
            C[p] = SV<S>::st(SV<S>::ld(C[p]) + value);
        }
    }
}

// Table time index and weight for step it (addMomenteRate, addMoment.cpp:238-264); false once
// srcIt + 1 reaches the table end, after which the reference stops injecting.
__device__ __forceinline__ bool table_time(int it, float DT, float sdt, int snt, int& srcIt, float& tw) {
#pragma clang fp contract(off)
    const float tmpT = ((float)it + 0.0f) * DT;
    srcIt = int(tmpT / sdt);
    if (srcIt + 1 >= snt) return false;
    const float t1 = float(srcIt) * sdt;
    tw = (tmpT - t1) / sdt;
    return true;
}

// Linear-in-time interpolation of the moment-rate table and injection into the six stresses of
// point p (interp_momentRate + addSource1, addMoment.cpp:98-144, 194-219).
template <class S>
__device__ __forceinline__ void table_source_at(S* A, const Layout& L, const int* __restrict__ ijk,
                                                const float* __restrict__ mrate, long long npts, int srcIt, float tw,
                                                float DT, long long p) {
#pragma clang fp contract(off)
    const int i = ijk[3 * p], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
    const float* m0 = mrate + ((long long)srcIt * npts + p) * 6;
    const float* m1 = mrate + ((long long)(srcIt + 1) * npts + p) * 6;
    for (int c = 0; c < 6; ++c) {
        const float slice = (m1[c] - m0[c]) * tw + m0[c];
        const long long q = L.at(9, 3 + c, i, j, k);
        A[q] = SV<S>::st(SV<S>::ld(A[q]) - slice * DT);
    }
}

// Value t = s*9 + v of the station record after step it (storage_station, station.cpp:309-347).
template <class S>
__device__ __forceinline__ void station_at(const S* __restrict__ A, const Layout& L, const int* __restrict__ ijk,
                                           float* wave, int NT, int it, float c1, float c2, int t) {
#pragma clang fp contract(off)
    const int s = t / 9, v = t % 9;
    const long long pos = it + (long long)s * NT;
    wave[pos * 9 + v] = SV<S>::ld(A[L.at(9, v, ijk[3 * s], ijk[3 * s + 1], ijk[3 * s + 2])]) * (v < 3 ? c1 : c2);
}

// Running maximum of horizontal and total surface speed at column t (compare_pgv, PGV.cpp:48-101).
template <class S>
__device__ __forceinline__ void pgv_at(const S* __restrict__ A, const Layout& L, int nx, int k, float* __restrict__ pgv,
                                       float c, int t) {
#pragma clang fp contract(off)
    const int i = t % nx + HALO, j = t / nx + HALO;
    const float Vx = SV<S>::ld(A[L.at(9, 0, i, j, k)]) * c;
    const float Vy = SV<S>::ld(A[L.at(9, 1, i, j, k)]) * c;
    const float Vz = SV<S>::ld(A[L.at(9, 2, i, j, k)]) * c;
    const float Vh = sqrtf(Vx * Vx + Vy * Vy);
    const float V = sqrtf(Vx * Vx + Vy * Vy + Vz * Vz);
    if (pgv[t * 2 + 0] < Vh) pgv[t * 2 + 0] = Vh;
    if (pgv[t * 2 + 1] < V) pgv[t * 2 + 1] = V;
}

// Source coefficient q of the 7^3 stencil: -amp_q/(J*DH^3)*Cs*DT with J read from the stored
// metric at the source (load_smooth_source, singleSource.cpp:100-121). amp is the normalised
// Gaussian weight, computed on the host in double as the reference does.
template <class S>
__device__ __forceinline__ float source_coef_at(const S* __restrict__ cjm, const Layout& L, int SX, int SY, int SZ,
                                                float amp, float DH, double CsDT) {
#pragma clang fp contract(off)
    const float jacb = SV<S>::ld(cjm[L.atc(9, SX, SY, SZ)]);
    const float value = -1.0f * amp / (jacb * (DH * DH * DH));
    return (float)(value * CsDT);
}

}  // namespace cg
