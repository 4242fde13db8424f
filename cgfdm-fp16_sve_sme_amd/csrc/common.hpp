// common.hpp — shared definitions of the MI355X CGFDM library (device and host side).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cg {

constexpr int HALO = 3;

// MacCormack coefficients (reference macro.h:126-130) and RK weights (macro.h:138-147),
// kept as the same float literals: their sum is -6e-5, not 0, and the scheme depends on it.
constexpr float af_1 = -0.30874f, af0 = -0.6326f, af1 = 1.2330f, af2 = -0.3334f, af3 = 0.04168f;
constexpr float alpha2 = 0.5f, alpha3 = 0.5f;
constexpr float beta1 = 0.16666667f, beta2 = 0.33333333f, beta3 = 0.33333333f, beta4 = 0.16666667f;
constexpr float Cf1 = -1.16666667f, Cf2 = 1.33333333f, Cf3 = -0.16666667f;
constexpr float PIf = 3.141592657f;

// Storage codecs: values live in HBM as S and are widened to FP32 for every operation.
template <class S> struct SV;
template <> struct SV<float> {
    __device__ __forceinline__ static float ld(float v) { return v; }
    __device__ __forceinline__ static float st(float v) { return v; }
};
template <> struct SV<_Float16> {
    __device__ __forceinline__ static float ld(_Float16 v) { return (float)v; }
    __device__ __forceinline__ static _Float16 st(float v) { return (_Float16)v; }
};
template <> struct SV<__bf16> {
    __device__ __forceinline__ static float ld(__bf16 v) { return (float)v; }
    __device__ __forceinline__ static __bf16 st(float v) { return (__bf16)v; }
};

// One PML slab: auxiliary variables of the 9 fields, SoA [9][n], FP32, three RK registers
// (current value, stage-0 copy, accumulator). Slab order x1,y1,z1,x2,y2,z2 as AUX6SURF
// (reference struct.h:565-574).
struct Slab {
    float* a;
    float* m;
    float* t;
    long long n;
    int on;
    int axis;   // 0 x, 1 y, 2 z
    int lo;     // first padded index of the slab along its axis
    int n0, n1; // pos = i0 + j0*n0 + k0*n0*n1 with (i0,j0,k0) relative to the slab origin
};

// Everything one RK stage needs. Passed by value to the stage kernels.
struct StageParams {
    const void* Win;  // stage input wavefield (9 planes)
    void* Wout;       // stage output wavefield
    const void* A;    // W at the start of the step (RK register m)
    void* T;          // RK accumulator t
    const void* cjm;  // 13 planes: xi_x..zt_z, J, mu, lambda, buoyancy
    const float* pa[3];  // PML alpha per axis (padded length)
    const float* pb[3];  // 1/beta
    const float* pd[3];  // d/beta
    const float* rdz;    // [2][nx_*ny_][9]: -rDZ*DX then -rDZ*DY
    Slab slab[6];
    long long num;       // padded points per variable
    int nx, ny, nz;      // interior sizes
    int nx_, ny_;        // padded sizes
    int nPML, fs, pml, irk;
    int FB1, FB2, FB3;
    float rDH, DT, Crho;
    // bulk box (handled by the fast kernel); the general kernel skips it when bulk_on
    int bulk_on;
    int bi0, bi1, bj0, bj1, bk0, bk1;
};

struct SetupParams {
    int nx, ny, nz, nx_, ny_, nz_;
    long long num;
    double DH;
    int use_terrain;
    double depth;
    int centerX, centerY;
    const double* terrain;  // padded (ny_, nx_)
    int fs;
    double Cv, Cs, Crho;
};

enum Prec { P_F32 = 1, P_F16 = 2, P_BF16 = 3 };

// ---- launchers (defined in the .hip files) ----
int launch_setup_medium(int prec, const SetupParams& q, const float* vs, const float* vp, const float* rho,
                        void* cjm, hipStream_t s);
int launch_setup_metric(int prec, const SetupParams& q, void* cjm, hipStream_t s);
int launch_setup_rdz(const SetupParams& q, const float* vs, const float* vp, const float* rho, float* rdz,
                     hipStream_t s);
int launch_hmin(const SetupParams& q, unsigned long long* out, hipStream_t s);

int launch_stage(int prec, const StageParams& p, int use_fast, hipStream_t s);

int launch_point_source(int prec, void* A, void* B, void* C, long long num, int nx_, int ny_, int SX, int SY,
                        int SZ, const float* coef, float sval, hipStream_t s);
int launch_multi_source(int prec, void* A, long long num, const long long* idx, const float* mrate,
                        long long npts, int srcIt, float tw, float DT, hipStream_t s);
int launch_scale_mrate(int prec, const void* cjm, long long num, const long long* idx, float* mrate,
                       long long npts, int nt, float DH, float Cv, hipStream_t s);
int launch_station(int prec, const void* A, long long num, const long long* idx, int nst, float* wave, int NT,
                   int it, float c1, float c2, hipStream_t s);
int launch_pgv(int prec, const void* A, long long num, int nx, int ny, int nx_, int ny_, int k, float* pgv,
               float c, hipStream_t s);
int launch_slice(int prec, const void* A, long long num, int nx, int ny, int nz, int nx_, int ny_, int axis,
                 int index, int var, float c, float* out, hipStream_t s);
int launch_gather(int prec, const void* A, long long num, int nx, int ny, int nz, int nx_, int ny_, int nvar,
                  float* out, hipStream_t s);
int launch_scatter(int prec, void* A, long long num, int nx, int ny, int nz, int nx_, int ny_, int nvar,
                   const float* in, hipStream_t s);
int launch_decode(int prec, const void* src, float* dst, long long n, hipStream_t s);

inline int storage_bytes(int prec) { return prec == P_F32 ? 4 : 2; }

}  // namespace cg
