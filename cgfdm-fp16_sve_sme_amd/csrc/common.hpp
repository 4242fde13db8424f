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
// The RK accumulator after stage 0, t = m + beta1*h0, is not stored: stage 1 rebuilds it from
// the stage-0 output w1 = m + alpha2*h0 it reads anyway, t = m + (w1 - m)*kT0 with
// kT0 = beta1/alpha2 (exact in float). 18 of the 192 values a point moves per step are saved;
// t carries the rounding of w1 instead of its own (one rounding either way).
constexpr float kT0 = beta1 / alpha2;
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

// Internal memory layout of every 3-D array with V variables: block-interleaved by variable
// pairs, [k][j][x block][...][64]. Rows are padded to px = nb*64 elements and shifted right by
// `lead` so that the first column of the bulk box starts a block. A block (V*64 elements) holds,
// for each pair of variables (2p, 2p+1), the 64 points' values interleaved point by point
// ([64][2]), followed, for odd V, by the 64 values of the last variable alone. One 32-bit (16-bit
// storage) or 64-bit (FP32 storage) load per lane thus brings both variables of a pair, a wave's
// loads of all variables of 64 consecutive points cover one contiguous run with no unused
// element, and a k-plane of all variables is one contiguous region of V*plane elements
// (plane = px*ny_). The bandwidth probe of this access mix served contiguous per-block runs
// ~25 % faster than one stream per variable plane.
//
// The metric/medium array (13 components) stores the reference components (0-8 metrics xi_x,
// xi_y, xi_z, eta_x, eta_y, eta_z, zeta_x, zeta_y, zeta_z, 9 J, 10 mu, 11 lambda, 12 buoyancy)
// in the slot order xi_x, eta_y | mu, lambda | zeta_x, zeta_y | zeta_z, J | xi_y, xi_z |
// eta_x, eta_z | buoyancy (the last, unpaired). The grids of this library are terrain-following
// (x depends on i only, y on j only, z linear in k along every column), so x_eta = x_zeta = y_xi =
// y_zeta = 0 exactly (the averaged forward/backward differences of a coordinate constant along an
// axis cancel bit for bit) and with them xi_y, xi_z, eta_x, eta_z, whose adjugate terms are
// products of those zeros; and zeta_x, zeta_y are (k - 3) times a per-column value, zeta_z is
// constant along a column. The tile kernels therefore read per point only the first two pairs
// and the buoyancy, and take zeta from three per-column values (StageParams::zcol). atc() takes
// the reference numbering.
constexpr int kBlk = 64;
__host__ __device__ __forceinline__ constexpr int cjm_slot(int c) {
    return c == 0 ? 0 : c == 1 ? 8 : c == 2 ? 9 : c == 3 ? 10 : c == 4 ? 1 : c == 5 ? 11 : c == 6 ? 4 : c == 7 ? 5
         : c == 8 ? 6 : c == 9 ? 7 : c == 10 ? 2 : c == 11 ? 3 : 12;
}
struct Layout {
    int nx_, ny_, nz_;  // padded sizes (interior + 2*HALO)
    int px, lead, nb;   // row pitch, column shift (elements), blocks per row
    long long plane;    // px * ny_
    // offset of the block holding (i, j) in a k-plane of a V-variable array
    __host__ __device__ __forceinline__ long long blk(int V, int i, int j) const {
        return ((long long)j * nb + ((i + lead) >> 6)) * V * kBlk;
    }
    __host__ __device__ __forceinline__ int lane(int i) const { return (i + lead) & (kBlk - 1); }
    // offset of variable (slot) v of the point in lane ln, relative to its block
    __host__ __device__ __forceinline__ static long long var_off(int V, int v, int ln) {
        return v < (V & ~1) ? (long long)(v >> 1) * 2 * kBlk + 2 * ln + (v & 1) : (long long)(V - 1) * kBlk + ln;
    }
    // offsets inside a k-plane of (i, j): its first pair, and the unpaired last variable (odd V)
    __host__ __device__ __forceinline__ long long col(int V, int i, int j) const { return blk(V, i, j) + 2 * lane(i); }
    __host__ __device__ __forceinline__ long long col1(int V, int i, int j) const {
        return blk(V, i, j) + (long long)(V - 1) * kBlk + lane(i);
    }
    __host__ __device__ __forceinline__ long long kplane(int V) const { return (long long)V * plane; }
    __host__ __device__ __forceinline__ long long at(int V, int v, int i, int j, int k) const {
        return (long long)k * kplane(V) + blk(V, i, j) + var_off(V, v, lane(i));
    }
    // metric/medium array, reference component numbering
    __host__ __device__ __forceinline__ long long atc(int c, int i, int j, int k) const {
        return at(13, cjm_slot(c), i, j, k);
    }
    __host__ __device__ __forceinline__ long long size(int V) const { return (long long)nz_ * kplane(V); }
};

// Global access with a wave-uniform base pointer and a 32-bit byte offset per lane, so the load
// uses the scalar-base addressing mode instead of a 64-bit address per variable in VGPRs.
// Every array here is below 4 GiB per variable plane (padded grid < 2^31 points, <= 4 bytes).
template <class S>
__device__ __forceinline__ float gld(const S* __restrict__ base, unsigned elem) {
    return SV<S>::ld(*(const S*)((const char*)base + elem * (unsigned)sizeof(S)));
}
template <class S>
__device__ __forceinline__ void gst(S* __restrict__ base, unsigned elem, float v) {
    *(S*)((char*)base + elem * (unsigned)sizeof(S)) = SV<S>::st(v);
}

// Buffer-resource access: a 128-bit descriptor (base, byte size) in SGPRs, a per-lane 32-bit byte
// offset and a wave-uniform byte offset, so a load or store needs no VALU address arithmetic.
// Accesses past the descriptor's size return 0 (loads) or are dropped (stores).
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, 32-bit data format (gfx9 family)

template <class S>
__device__ __forceinline__ Rsrc make_rsrc(const S* base, long long nelem) {
    const long long bytes = nelem * (long long)sizeof(S);
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL),
                                             kRsrcWord3);
}
template <class S> __device__ __forceinline__ float bld(Rsrc r, unsigned voff, int soff);
template <> __device__ __forceinline__ float bld<float>(Rsrc r, unsigned voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
}
template <> __device__ __forceinline__ float bld<_Float16>(Rsrc r, unsigned voff, int soff) {
    return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0));
}
template <> __device__ __forceinline__ float bld<__bf16>(Rsrc r, unsigned voff, int soff) {
    return (float)__builtin_bit_cast(__bf16, __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0));
}
template <class S> __device__ __forceinline__ void bst(Rsrc r, unsigned voff, int soff, float v);
template <> __device__ __forceinline__ void bst<float>(Rsrc r, unsigned voff, int soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, soff, 0);
}
template <> __device__ __forceinline__ void bst<_Float16>(Rsrc r, unsigned voff, int soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), r, (int)voff, soff, 0);
}
template <> __device__ __forceinline__ void bst<__bf16>(Rsrc r, unsigned voff, int soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (__bf16)v), r, (int)voff, soff, 0);
}

// One PML slab: auxiliary variables of the 9 fields, SoA [9][n], three RK registers (current
// value, stage-0 copy, accumulator) in the wavefield's storage type. A register holds Aux*DT,
// the auxiliary's contribution to one stage's increment (h += -beta*Aux*DT, pml_deriv.cpp:
// 209-237), so each auxiliary carries the scale of its own wavefield variable (Cv for the
// velocities, Cs for the stresses) and stays inside that variable's 16-bit range; the reference
// keeps Aux unscaled in FP32 (struct.h:271-277). Slab order x1,y1,z1,x2,y2,z2 as AUX6SURF
// (reference struct.h:565-574).
struct Slab {
    void* a;
    void* m;
    void* t;
    long long n;
    int on;
    int axis;   // 0 x, 1 y, 2 z
    int lo;     // first padded index of the slab along its axis
    int n0, n1; // pos = i0 + j0*n0 + k0*n0*n1 with (i0,j0,k0) relative to the slab origin
};

constexpr int kMaxGenBoxes = 4;   // the free-surface layers of the PML columns
constexpr int kMaxPBoxes = 26;    // PML boxes: 6 faces, 12 edges, 8 corners

// PML box kinds by the set of slab normals they lie in (bit 0 x, bit 1 y, bit 2 z), in launch
// order: faces x, y, z, then edges xy, xz, yz, then corners xyz.
constexpr int kNumPmlMasks = 7;
__host__ __device__ __forceinline__ constexpr int pml_mask_at(int g) {
    return g == 0 ? 1 : g == 1 ? 2 : g == 2 ? 4 : g == 3 ? 3 : g == 4 ? 5 : g == 5 ? 6 : 7;
}
__host__ __device__ __forceinline__ constexpr int pml_group_of(int mask) {
    return mask == 1 ? 0 : mask == 2 ? 1 : mask == 4 ? 2 : mask == 3 ? 3 : mask == 5 ? 4 : mask == 6 ? 5 : 6;
}

// Everything one RK stage needs. Passed by value to the stage kernels.
struct StageParams {
    const void* Win;  // stage input wavefield (9 planes)
    void* Wout;       // stage output wavefield
    const void* A;    // W at the start of the step (RK register m)
    void* T;          // RK accumulator t
    const void* cjm;  // 13 planes: xi_x..zt_z, J, mu, lambda, buoyancy
    const void* m0;   // {int uniform, int, first metric pair}: the pair's value when uniform (or null)
    const float* zcol;  // [3][ny_][nx_]: zeta_x / (k - 3), zeta_y / (k - 3), zeta_z of every column
    const float* pa[3];  // PML alpha per axis (padded length)
    const float* pb[3];  // 1/beta
    const float* pd[3];  // d/beta
    const float* rdz;    // [2][nx_*ny_][9]: -rDZ*DX then -rDZ*DY, indexed by i + j*nx_
    Slab slab[6];
    Layout lay;
    int nx, ny, nz;      // interior sizes
    int nx_, ny_;        // padded sizes
    int nPML, fs, pml, irk;
    int FB1, FB2, FB3;
    float rDH, DT, Crho;
    // Decomposition of the interior (stage_boxes):
    //  - the bulk box [bi0,bi1)x[bj0,bj1)x[bk0,bk1): no PML slab, no free-surface layer, covered
    //    by the tile kernel;
    //  - npbox PML boxes outside the free-surface layers, each inside the slabs of one set of
    //    normals (a face: one slab; an edge: two; a corner: three), covered by the tile kernel
    //    with those slabs' auxiliary updates;
    //  - the free-surface box [bi0,bi1)x[bj0,bj1)x[fk0,fk1): the three free-surface layers above
    //    the bulk columns (no PML), covered by the free-surface tile kernel;
    //  - nbox general boxes, covered point by point: the free-surface layers of the PML columns.
    int bulk_on, fs_on;
    int bulk_cfg, bulk_kc;  // tile shape index and z-chunk length (0: automatic)
    int bi0, bi1, bj0, bj1, bk0, bk1, fk0, fk1;
    int npbox;
    int pbox[kMaxPBoxes][8];  // i0,i1,j0,j1,k0,k1, normals mask, slab per axis (4 bits each, 15: none)
    int nbox;
    int box[kMaxGenBoxes][6];            // i0,i1,j0,j1,k0,k1
    long long boxStart[kMaxGenBoxes + 1];  // prefix sums of box sizes
};

struct SetupParams {
    int nx, ny, nz, nx_, ny_, nz_;
    Layout lay;
    double DH;
    int use_terrain;
    double depth;
    int centerX, centerY;
    const double* terrain;  // padded (ny_, nx_)
    int fs;
    double Cv, Cs, Crho;
    int chi_mode;  // 0: chi = 2*lambda + mu (reference), 1: lambda + 2*mu
};

// Source of the medium for the setup kernels: 3-D arrays on the interior grid (vs != null), or a
// layer table of nl rows [top depth (m) | Vs | Vp | rho] stored as four arrays of nl floats.
struct MediumSrc {
    const float* vs;
    const float* vp;
    const float* rho;
    const float* layers;
    int nl;
};

enum Prec { P_F32 = 1, P_F16 = 2, P_BF16 = 3 };

// ---- launchers (defined in the .hip files) ----
int launch_setup_medium(int prec, const SetupParams& q, const MediumSrc& ms, void* cjm, hipStream_t s);
int launch_vpmax(const SetupParams& q, const MediumSrc& ms, unsigned* out, hipStream_t s);
// The metric array and the per-column zeta values of the tile kernels (zcol); m0 + 4 is raised
// where a point's zeta does not have the per-column form.
int launch_setup_metric(int prec, const SetupParams& q, void* cjm, float* zcol, void* m0, hipStream_t s);
// Whether the first metric pair is the same at every interior point (m0: see setup.hip).
int launch_pair0_check(int prec, const SetupParams& q, const void* cjm, void* m0, hipStream_t s);
int launch_setup_rdz(const SetupParams& q, const MediumSrc& ms, float* rdz, hipStream_t s);
int launch_hmin(const SetupParams& q, unsigned long long* out, hipStream_t s);

// One RK stage: the bulk kernel (use_fast != 0) then the general kernel over the remaining
// boxes (skipped when use_fast == 2, used to time the bulk kernel alone). `mid`, if given, is
// recorded between the two launches.
// One fused RK stage on stream s. With a side stream (and two events), the PML, free-surface and
// general kernels run on it concurrently with the bulk kernel (they write disjoint points and
// read only the stage input), joined back into s at the end; mid (if given) is recorded on s
// after the bulk kernel.
// The serving engine may also give two more side streams (side2, with their join events): the
// boundary kernels are then spread over three side streams, which run concurrently with each
// other as well (every kernel of a stage writes its own points and their auxiliaries only).
struct StageFork {
    hipStream_t side;
    hipEvent_t fork, join;
    hipStream_t side2[2] = {nullptr, nullptr};
    hipEvent_t join2[2] = {nullptr, nullptr};
};
int launch_stage(int prec, const StageParams& p, int use_fast, hipStream_t s, hipEvent_t mid,
                 const StageFork* fk = nullptr);

// Point lists (sources, stations) carry padded (i, j, k) triples.
int launch_point_source(int prec, void* A, void* B, void* C, const Layout& L, int SX, int SY, int SZ,
                        const float* coef, float fc, int it, float DT, hipStream_t s);
int launch_multi_source(int prec, void* A, const Layout& L, const int* ijk, const float* mrate, long long npts,
                        int snt, float sdt, int it, float DT, hipStream_t s);
int launch_source_coef(int prec, const void* cjm, const Layout& L, int SX, int SY, int SZ, const float* amp, float DH,
                       double CsDT, float* coef, hipStream_t s);
int launch_scale_mrate(int prec, const void* cjm, const Layout& L, const int* ijk, float* mrate, long long npts,
                       int nt, float DH, float Cv, hipStream_t s);
int launch_station(int prec, const void* A, const Layout& L, const int* ijk, int nst, float* wave, int NT, int it,
                   float c1, float c2, hipStream_t s);
int launch_pgv(int prec, const void* A, const Layout& L, int nx, int ny, int k, float* pgv, float c, hipStream_t s);
// Sets *flag to code (if still 0) when the 9-variable register A holds a value that is not finite.
int launch_nonfinite(int prec, const void* A, const Layout& L, int* flag, int code, hipStream_t s);
int launch_slice(int prec, const void* A, const Layout& L, int nx, int ny, int nz, int axis, int index, int var,
                 float c, float* out, hipStream_t s);
// interior <-> float array of shape (V, NZ, NY, NX)
int launch_gather(int prec, const void* A, const Layout& L, int V, int nx, int ny, int nz, float* out,
                  hipStream_t s);
int launch_scatter(int prec, void* A, const Layout& L, int V, int nx, int ny, int nz, const float* in,
                   hipStream_t s);
// whole padded array -> float array in the reference order (V, nz_, ny_, nx_)
int launch_unpack(int prec, const void* src, const Layout& L, int V, float* dst, hipStream_t s);
int launch_decode(int prec, const void* src, float* dst, long long n, hipStream_t s);

inline int storage_bytes(int prec) { return prec == P_F32 ? 4 : 2; }

// Workgroups of a launch are dealt to the XCDs round-robin (hardware id b runs on XCD b % 8 on
// MI355X, each XCD with its own L2). xcd_order maps hardware id b to a work index so that each
// XCD receives one contiguous run of work items: neighbouring tiles, whose halos overlap, then
// share an L2 instead of fetching the same lines from HBM on different XCDs. A bijection on
// [0, n) for any n (correctness does not depend on the XCD count, only locality does).
constexpr int kXcds = 8;
__host__ __device__ __forceinline__ long long xcd_order(long long b, long long n) {
    const long long q = n / kXcds, r = n % kXcds;
    const long long x = b % kXcds, local = b / kXcds;
    return x * q + (x < r ? x : r) + local;
}

}  // namespace cg
