// stage_bulk.hpp — the hot kernel: one fused RK stage over the bulk of the grid (every point that
// is neither in a PML slab nor in the top three free-surface layers), and the same tile code for
// the single-slab PML boxes.
//
// Work decomposition (CDNA4, wave64):
//  - a workgroup owns a TX x TY column tile in (x, y) and marches a chunk of planes along z in the
//    direction FB3, so the one-sided z stencil (planes k-FB3 .. k+3FB3) lives in a 5-plane
//    register queue and each plane of W is read from HBM once per chunk;
//  - the current plane of the 9 fields, with its (4-wide) x/y halo, is staged in LDS as FP32 and
//    double-buffered, so one barrier per plane suffices; x and y neighbours are LDS reads;
//  - global accesses are buffer loads/stores of variable pairs (the layout stores the two
//    variables of a pair next to each other per point): one 32-bit (16-bit storage) or 64-bit
//    (FP32 storage) access per lane brings two variables, with one descriptor per array and plane,
//    a per-lane byte offset that does not change along the march and constant per-pair offsets;
//  - 16-bit storage keeps the pairs packed in registers (two values per VGPR, widened to FP32 by
//    the mixed-precision FMAs that consume them) and uses the registers this frees to request the
//    next plane's halo cells and metric/medium values one iteration ahead, so the barrier of an
//    iteration never waits on a load issued in the same iteration;
//  - no load sits in a branch (the compiler would wait for it inside the branch); the RK stage is
//    a template parameter; the z loop is unrolled by 5 (the queue rotates by renaming) and always
//    runs a multiple of 5 iterations, the ones past the chunk loading clamped planes and storing
//    nothing, so there is a single copy of the unrolled body.
// Arithmetic is FP32 with the stencil coefficients pre-multiplied by FB*rDH and DT folded into
// lambda, mu and the buoyancy.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace cg {

// First column of the x tiles of a box with columns [ci0, ci1), chosen so that no tile row
// straddles two layout blocks (TX divides the block size): a box no wider than a tile that fits
// in one block from ci0 gets one tile starting at ci0; otherwise the tiles start on the TX-aligned
// run of the block that holds ci0.
__host__ __device__ __forceinline__ int pml_tile_x0(const Layout& L, int ci0, int ci1, int TX) {
    const int off = (ci0 + L.lead) & (kBlk - 1);
    if (ci1 - ci0 <= TX && off + TX <= kBlk) return ci0;
    return ci0 - ((ci0 + L.lead) & (TX - 1));
}

// LDS copy of a plane with its halo, double-buffered. FP32 storage: the 9 values as FP32, one
// word per variable. 16-bit storage: the stored bits themselves, one 32-bit word per variable
// pair (4 words) and the unpaired variable in the low half of a fifth word, so staging needs no
// conversion and one LDS read brings both variables of a pair (the FMAs widen the halves; the
// arithmetic is the same as on FP32 copies of these values).
template <class S, int TX, int TY> struct BulkLds {
    static constexpr int LX = TX + 4, LY = TY + 4, VS = LX * LY;  // LDS row pitch and word stride
    static constexpr bool PACKED = sizeof(S) == 2;
    static constexpr int NW = PACKED ? 5 : 9;  // words per cell
    using word = std::conditional_t<PACKED, unsigned, float>;
    using array = word[2][NW * VS];
    array v;
};

// Cache policy of a buffer access: kCpNt marks a stream the kernel touches once (non-temporal:
// its lines are placed for early eviction from L2), kCpDefault one that neighbours re-read.
constexpr int kCpDefault = 0, kCpNt = 2;

// Raw buffer words with a compile-time cache policy.
template <int CP> __device__ __forceinline__ unsigned short ldw16(Rsrc r, unsigned voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, CP);
}
template <int CP> __device__ __forceinline__ unsigned ldw32(Rsrc r, unsigned voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, CP);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
template <int CP> __device__ __forceinline__ u32x2 ldw64(Rsrc r, unsigned voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, CP);
}
template <int CP> __device__ __forceinline__ void stw16(Rsrc r, unsigned voff, int soff, unsigned short x) {
    __builtin_amdgcn_raw_buffer_store_b16(x, r, (int)voff, soff, CP);
}
template <int CP> __device__ __forceinline__ void stw32(Rsrc r, unsigned voff, int soff, unsigned x) {
    __builtin_amdgcn_raw_buffer_store_b32(x, r, (int)voff, soff, CP);
}
template <int CP> __device__ __forceinline__ void stw64(Rsrc r, unsigned voff, int soff, u32x2 x) {
    __builtin_amdgcn_raw_buffer_store_b64(x, r, (int)voff, soff, CP);
}

// A pair of stored values (two variables of one point) as loaded by one lane, and a single one.
template <class S> struct Pair;
template <> struct Pair<_Float16> {
    typedef _Float16 type __attribute__((ext_vector_type(2)));
    typedef _Float16 one;
    template <int CP> __device__ __forceinline__ static type ld(Rsrc r, unsigned voff, int soff) {
        return __builtin_bit_cast(type, ldw32<CP>(r, voff, soff));
    }
    template <int CP> __device__ __forceinline__ static one ld1(Rsrc r, unsigned voff, int soff) {
        return __builtin_bit_cast(one, ldw16<CP>(r, voff, soff));
    }
    __device__ __forceinline__ static float f1(one x) { return (float)x; }
    __device__ __forceinline__ static float lo(type x) { return (float)x.x; }
    __device__ __forceinline__ static float hi(type x) { return (float)x.y; }
    template <int CP> __device__ __forceinline__ static void st(Rsrc r, unsigned voff, int soff, float a, float b) {
        type x;
        x.x = (_Float16)a;
        x.y = (_Float16)b;
        stw32<CP>(r, voff, soff, __builtin_bit_cast(unsigned, x));
    }
    template <int CP> __device__ __forceinline__ static void st1(Rsrc r, unsigned voff, int soff, float a) {
        stw16<CP>(r, voff, soff, __builtin_bit_cast(unsigned short, (_Float16)a));
    }
};
template <> struct Pair<__bf16> {
    typedef unsigned type;
    typedef unsigned short one;
    template <int CP> __device__ __forceinline__ static type ld(Rsrc r, unsigned voff, int soff) {
        return ldw32<CP>(r, voff, soff);
    }
    template <int CP> __device__ __forceinline__ static one ld1(Rsrc r, unsigned voff, int soff) {
        return ldw16<CP>(r, voff, soff);
    }
    __device__ __forceinline__ static float f1(one x) { return __builtin_bit_cast(float, (unsigned)x << 16); }
    __device__ __forceinline__ static float lo(type x) { return __builtin_bit_cast(float, x << 16); }
    __device__ __forceinline__ static float hi(type x) { return __builtin_bit_cast(float, x & 0xffff0000u); }
    template <int CP> __device__ __forceinline__ static void st(Rsrc r, unsigned voff, int soff, float a, float b) {
        const unsigned ua = __builtin_bit_cast(unsigned short, (__bf16)a);
        const unsigned ub = __builtin_bit_cast(unsigned short, (__bf16)b);
        stw32<CP>(r, voff, soff, ua | (ub << 16));
    }
    template <int CP> __device__ __forceinline__ static void st1(Rsrc r, unsigned voff, int soff, float a) {
        stw16<CP>(r, voff, soff, __builtin_bit_cast(unsigned short, (__bf16)a));
    }
};
template <> struct Pair<float> {
    typedef float type __attribute__((ext_vector_type(2)));
    typedef float one;
    template <int CP> __device__ __forceinline__ static type ld(Rsrc r, unsigned voff, int soff) {
        return __builtin_bit_cast(type, ldw64<CP>(r, voff, soff));
    }
    template <int CP> __device__ __forceinline__ static one ld1(Rsrc r, unsigned voff, int soff) {
        return __builtin_bit_cast(float, ldw32<CP>(r, voff, soff));
    }
    __device__ __forceinline__ static float f1(one x) { return x; }
    __device__ __forceinline__ static float lo(type x) { return x.x; }
    __device__ __forceinline__ static float hi(type x) { return x.y; }
    template <int CP> __device__ __forceinline__ static void st(Rsrc r, unsigned voff, int soff, float a, float b) {
        type x;
        x.x = a;
        x.y = b;
        stw64<CP>(r, voff, soff, __builtin_bit_cast(u32x2, x));
    }
    template <int CP> __device__ __forceinline__ static void st1(Rsrc r, unsigned voff, int soff, float a) {
        stw32<CP>(r, voff, soff, __builtin_bit_cast(unsigned, a));
    }
};

// The 9 stored wavefield values of one point: four pairs and the unpaired last variable.
template <class S> struct W9 {
    typename Pair<S>::type p[4];
    typename Pair<S>::one s;
    __device__ __forceinline__ float operator[](int v) const {
        return v == 8 ? Pair<S>::f1(s) : (v & 1) ? Pair<S>::hi(p[v >> 1]) : Pair<S>::lo(p[v >> 1]);
    }
    // the stored bits as the packed LDS words of one cell (16-bit storage; word stride vs)
    __device__ __forceinline__ void put(unsigned* L, int vs, int cell) const {
#pragma unroll
        for (int c = 0; c < 4; ++c) L[c * vs + cell] = __builtin_bit_cast(unsigned, p[c]);
        L[4 * vs + cell] = (unsigned)__builtin_bit_cast(unsigned short, s);
    }
    // PB: bytes from one pair to the next; col/col1: byte offsets of pair 0 and of variable 8
    template <int PB, int CP = kCpDefault> __device__ __forceinline__ void load(Rsrc r, unsigned col, unsigned col1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) p[c] = Pair<S>::template ld<CP>(r, col, c * PB);
        s = Pair<S>::template ld1<CP>(r, col1, 0);
    }
};
template <class S, int PB, int CP>
__device__ __forceinline__ void store9(Rsrc r, unsigned col, unsigned col1, const float* o) {
#pragma unroll
    for (int c = 0; c < 4; ++c) Pair<S>::template st<CP>(r, col, c * PB, o[2 * c], o[2 * c + 1]);
    Pair<S>::template st1<CP>(r, col1, 0, o[8]);
}

// One tile: columns [ci0 + bx*TX, +TX) x [cj0 + by*TY, +TY) of the box with columns
// [ci0,ci1) x [cj0,cj1), planes [kb, ke). Every thread of the workgroup calls it with the same
// arguments.
//
// MASK = 0: the bulk box, where every PML factor is 1. MASK != 0: a PML box inside the slabs
// whose normals are the set bits of MASK (bit 0 x, 1 y, 2 z; `slabs` holds the slab index of each
// axis in 4 bits) and outside the free-surface layers: the derivatives along each such axis are
// scaled by 1/beta of the point, and each slab's auxiliary variables get their update
// (pml_deriv.cpp:176-237: -beta*Aux*DT into the increment, the single-direction right-hand side
// times d minus (alpha+d)*Aux into the auxiliary increment; pml_rk.cpp:14-70: RK4 on the
// auxiliaries), slab after slab in axis order as the reference's x, y, z launches.
template <class S, int TX, int TY, int IRK, int MASK = 0>
__device__ __forceinline__ void bulk_tile(const StageParams& p, int FB1, int FB2, int FB3, int ci0, int ci1, int cj0,
                                          int cj1, int slabs, int bx, int by, int kb, int ke,
                                          typename BulkLds<S, TX, TY>::array& lds) {
    constexpr bool PML = MASK != 0;
    // a wave (64 consecutive threads of a tile row) must lie inside one layout block: PML tiles
    // divide a block; bulk tiles, which start on a block, may also span whole blocks
    static_assert(kBlk % TX == 0 || (!PML && TX % kBlk == 0), "a wave must lie inside one layout block");
    using PR = Pair<S>;
    using P2 = typename PR::type;
    constexpr bool PREFETCH = sizeof(S) == 2;  // packed pairs leave room for the next plane
    constexpr int LX = BulkLds<S, TX, TY>::LX, VS = BulkLds<S, TX, TY>::VS;
    constexpr bool PACKED = BulkLds<S, TX, TY>::PACKED;
    constexpr int NT = TX * TY;
    constexpr int PB = 2 * kBlk * (int)sizeof(S);  // bytes from one variable pair to the next
    // cache policies (measured): the metric/medium values, the RK registers A and T and the
    // stage output are touched once per point by this kernel and go non-temporal (12 % less
    // time than cached); the wavefield planes, whose halo cells neighbouring tiles re-read, and
    // the PML auxiliaries stay cached (non-temporal made both slower)
    constexpr int CPM = kCpNt, CPS = kCpNt, CPO = kCpNt, CPW = kCpDefault, CPH = kCpDefault, CPX = kCpDefault;

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const Layout L = p.lay;
    // x tiles of a PML box start where no tile row straddles two layout blocks (pml_tile_x0); the
    // bulk box starts a block
    const int i0 = (PML ? pml_tile_x0(L, ci0, ci1, TX) : ci0) + bx * TX, j0 = cj0 + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int hx = FB1 > 0 ? 1 : 3, hy = FB2 > 0 ? 1 : 3;  // halo width on the low side
    const bool has = i >= 0 && i < L.nx_ && j < L.ny_;       // a grid point (possibly outside the box)
    const bool inside = i >= ci0 && i < ci1 && j < cj1;      // computed by this kernel
    // byte offsets of the point in a k-plane of a 9-variable array (W, A, T, Wout: first pair and
    // variable 8) and of the metric/medium array; variable pair q adds q*PB bytes
    const unsigned col = has ? (unsigned)L.col(9, i, j) * (unsigned)sizeof(S) : 0u;
    const unsigned col1 = has ? (unsigned)L.col1(9, i, j) * (unsigned)sizeof(S) : 0u;
    const unsigned colc = has ? (unsigned)L.col(13, i, j) * (unsigned)sizeof(S) : 0u;
    const unsigned colc1 = has ? (unsigned)L.col1(13, i, j) * (unsigned)sizeof(S) : 0u;  // buoyancy
    // zeta of the lane's column in its per-column form (k_zcol): zeta_x = zax (k - 3),
    // zeta_y = zay (k - 3), zeta_z = zgz
    const long long zpl = (long long)L.nx_ * L.ny_, zcx = has ? i + (long long)j * L.nx_ : 0;
    const float zax = p.zcol[zcx], zay = p.zcol[zpl + zcx], zgz = p.zcol[2 * zpl + zcx];

    // z chunk, traversed in the direction FB3
    const int nk = ke - kb;
    const int k0 = FB3 > 0 ? kb : ke - 1;
    const int kmax = L.nz_ - 1;
    // The loop runs a multiple of 5 iterations; the ones past the chunk store nothing and load
    // planes an active iteration has just loaded (served by L2): the last centre plane, and for
    // the queue the last plane an active iteration's stencil reaches (requested at iteration
    // nk - 2). An iteration requests the next queue plane and the next plane's halo and metric
    // only for an active successor.
    const int klast = min(max(k0 + (nk - 1) * FB3, 0), kmax);
    const int kqlast = min(max(k0 + (nk + 2) * FB3, 0), kmax);

    const char* __restrict__ Wb = (const char*)p.Win;
    const char* __restrict__ Cb = (const char*)p.cjm;
    const char* __restrict__ Ab = (const char*)p.A;
    char* __restrict__ Tb = (char*)p.T;
    char* __restrict__ Ob = (char*)p.Wout;
    const long long wrange = L.kplane(9), crange = L.kplane(13);  // elements per k-plane
    const long long wpl = wrange * (long long)sizeof(S), cpl = crange * (long long)sizeof(S);

    // stencil weights with direction and 1/DH folded in
    const float rDH = p.rDH;
    const float cx0 = FB1 * af_1 * rDH, cx1 = FB1 * af0 * rDH, cx2 = FB1 * af1 * rDH, cx3 = FB1 * af2 * rDH,
                cx4 = FB1 * af3 * rDH;
    const float cy0 = FB2 * af_1 * rDH, cy1 = FB2 * af0 * rDH, cy2 = FB2 * af1 * rDH, cy3 = FB2 * af2 * rDH,
                cy4 = FB2 * af3 * rDH;
    const float cz0 = FB3 * af_1 * rDH, cz1 = FB3 * af0 * rDH, cz2 = FB3 * af1 * rDH, cz3 = FB3 * af2 * rDH,
                cz4 = FB3 * af3 * rDH;

    // register queue: slot (m + R) % 5 holds W(k + FB3*(m-1)), m = 0..4, at unrolled step R
    W9<S> q[5];
#pragma unroll
    for (int m = 0; m < 5; ++m)
        q[m].template load<PB, CPW>(make_rsrc((const S*)(Wb + (long long)(k0 + FB3 * (m - 1)) * wpl), wrange), col, col1);

    // halo cells of this thread: cell c = tid + r*NT (r < HPT) of the list x-halo (4 columns x TY
    // rows) then y-halo (4 rows x TX columns: the derivatives are one-dimensional, so the corner
    // cells, which lie on the diagonal neighbours' lines, are never read); a thread without a
    // cell loads element 0 in vain
    constexpr int NHX = 4 * TY, NHY = 4 * TX, NH = NHX + NHY;
    constexpr int HPT = (NH + NT - 1) / NT;
    int hl[HPT];
    unsigned hcol[HPT], hcol1[HPT];
    bool hown[HPT];
#pragma unroll
    for (int r = 0; r < HPT; ++r) {
        const int c = tid + r * NT;
        hl[r] = 0;
        hcol[r] = 0;
        hcol1[r] = 0;
        hown[r] = c < NH;
        if (c < NH) {
            int hr, hc;
            if (c < NHX) {
                const int rr = c / 4, c4 = c % 4;
                hr = hy + rr;
                hc = c4 < hx ? c4 : TX + c4;  // low-side columns [0,hx), high-side [hx+TX, LX)
            } else {
                const int t = c - NHX;
                const int r4 = t / TX;
                hc = hx + t % TX;
                hr = r4 < hy ? r4 : TY + r4;
            }
            const int ig = i0 - hx + hc, jg = j0 - hy + hr;
            const bool hhas = ig >= 0 && ig < L.nx_ && jg >= 0 && jg < L.ny_;
            hcol[r] = hhas ? (unsigned)L.col(9, ig, jg) * (unsigned)sizeof(S) : 0u;
            hcol1[r] = hhas ? (unsigned)L.col1(9, ig, jg) * (unsigned)sizeof(S) : 0u;
            hl[r] = hr * LX + hc;
        }
    }
    const int lc = (ty + hy) * LX + tx + hx;
    const float DT = p.DT;
    const float Crho = p.Crho;
    int k = k0;

    // halo cells and metric/medium values of a plane: the metric array's first two pairs,
    // (xi_x, eta_y) and (mu, lambda), and its unpaired last slot, the buoyancy; xi_y, xi_z, eta_x
    // and eta_z are exactly zero on the terrain-following grids this library builds, so their
    // terms are left out, and zeta comes from the per-column values above
    using P1 = typename PR::one;
    W9<S> hv[HPT];
    P2 M[2];
    P1 Mb;
    [[maybe_unused]] W9<S> hvN[HPT];
    [[maybe_unused]] P2 MN[2];
    [[maybe_unused]] P1 MbN;
    // The first pair (xi_x, eta_y) comes from the copy in p.m0 while the set-up found it uniform
    // over the grid: every lane and plane then loads the same cached line instead of streaming
    // the pair. The choice is of a descriptor and an offset, so no load sits in a branch; the
    // pair's load is cached (the copy is re-read by every wave).
    const int uni = p.m0 ? __builtin_amdgcn_readfirstlane(*(const int*)p.m0) : 0;
    const Rsrc rU = make_rsrc((const S*)(p.m0 ? (const char*)p.m0 + 8 : Cb), 2);
    const unsigned colc0 = uni ? 0u : colc;
    auto fetch_plane = [&](int kk, W9<S> (&h)[HPT], P2 (&m)[2], P1& mb) {
        const Rsrc rW = make_rsrc((const S*)(Wb + (long long)kk * wpl), wrange);
#pragma unroll
        for (int r = 0; r < HPT; ++r) h[r].template load<PB, CPH>(rW, hcol[r], hcol1[r]);
        const Rsrc rC = make_rsrc((const S*)(Cb + (long long)kk * cpl), crange);
        m[0] = PR::template ld<kCpDefault>(uni ? rU : rC, colc0, 0);
        m[1] = PR::template ld<CPM>(rC, colc, PB);
        mb = PR::template ld1<CPM>(rC, colc1, 0);
    };
    if constexpr (PREFETCH) fetch_plane(min(max(k, 0), kmax), hv, M, Mb);
    // PML slabs of the box, per normal axis a: the x/y profile values of the lane, and the
    // auxiliary positions pos = base + kk*kstride with kk the plane index inside the slab's k range
    [[maybe_unused]] float pBeta[3] = {1.f, 1.f, 1.f}, pD[3] = {0.f, 0.f, 0.f}, pAD[3] = {0.f, 0.f, 0.f};
    [[maybe_unused]] long long pBase[3] = {0, 0, 0}, pKs[3] = {0, 0, 0}, xN[3] = {0, 0, 0};
    [[maybe_unused]] int pK0[3] = {0, 0, 0}, pKn[3] = {1, 1, 1}, xn4[3] = {0, 0, 0};
    [[maybe_unused]] S* xA[3] = {nullptr, nullptr, nullptr};
    [[maybe_unused]] S* xM[3] = {nullptr, nullptr, nullptr};
    [[maybe_unused]] S* xT[3] = {nullptr, nullptr, nullptr};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!((MASK >> a) & 1)) continue;
        const Slab& X = p.slab[(slabs >> (4 * a)) & 15];
        const int ii = a == 0 ? i - X.lo : i - HALO;
        const int jj = a == 1 ? j - X.lo : j - HALO;
        pBase[a] = inside ? (long long)ii + (long long)jj * X.n0 : 0;
        pKs[a] = (long long)X.n0 * X.n1;
        pK0[a] = a == 2 ? X.lo : HALO;
        pKn[a] = a == 2 ? p.nPML : p.nz;
        if (a < 2) {
            const int c = a == 0 ? (inside ? i : HALO) : (inside ? j : HALO);
            pBeta[a] = p.pb[a][c];
            pD[a] = p.pd[a][c];
            pAD[a] = pD[a] + p.pa[a][c];
        }
        xA[a] = (S*)X.a;
        xM[a] = (S*)X.m;
        xT[a] = (S*)X.t;
        xN[a] = X.n;
        xn4[a] = (int)(X.n * (long long)sizeof(S));  // bytes per auxiliary variable
    }

    auto step = [&](auto Rc, int it) {
        constexpr int R = decltype(Rc)::value;
        W9<S>& q0 = q[(0 + R) % 5];
        const W9<S>& q1 = q[(1 + R) % 5];
        const W9<S>& q2 = q[(2 + R) % 5];
        const W9<S>& q3 = q[(3 + R) % 5];
        const W9<S>& q4 = q[(4 + R) % 5];
        const bool active = it < nk;  // iterations past the chunk store nothing
        const bool next = it + 1 < nk;
        auto* Lb = lds[it & 1];
        const int kc = active ? min(max(k, 0), kmax) : klast;
        const long long pw = (long long)kc * wpl;
        if constexpr (!PREFETCH) fetch_plane(kc, hv, M, Mb);
        // stage the centre plane and its halo in LDS
        [[maybe_unused]] float cen[9];
        if constexpr (PACKED) {
            q1.put(Lb, VS, lc);
#pragma unroll
            for (int r = 0; r < HPT; ++r)
                if (hown[r]) hv[r].put(Lb, VS, hl[r]);
        } else {
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                cen[v] = q1[v];
                Lb[v * VS + lc] = cen[v];
            }
#pragma unroll
            for (int r = 0; r < HPT; ++r)
                if (hown[r]) {
#pragma unroll
                    for (int v = 0; v < 9; ++v) Lb[v * VS + hl[r]] = hv[r][v];
                }
        }
        __syncthreads();
        // requests for later iterations: the plane entering the queue and, with prefetch, the
        // next plane's halo and metric/medium pairs (clamped at the grid end; unused there)
        W9<S> nq;
        const int kq = next ? min(max(k + 4 * FB3, 0), kmax) : kqlast;
        nq.template load<PB, CPW>(make_rsrc((const S*)(Wb + (long long)kq * wpl), wrange), col, col1);
        if constexpr (PREFETCH) fetch_plane(next ? min(max(k + FB3, 0), kmax) : klast, hvN, MN, MbN);
        // RK registers of this plane: needed only by the final update
        const Rsrc rT = make_rsrc((const S*)(Tb + pw), wrange);
        W9<S> ra, rt;
        if constexpr (IRK == 1 || IRK == 2) ra.template load<PB, CPS>(make_rsrc((const S*)(Ab + pw), wrange), col, col1);
        if constexpr (IRK >= 2) rt.template load<PB, CPS>(rT, col, col1);
        // auxiliaries of this plane, Aux*DT in the storage type (lanes outside the box read
        // element 0); the z profile values belong to the plane
        [[maybe_unused]] float xa[3][9], xm[3][9], xt[3][9];
        [[maybe_unused]] unsigned xoff[3] = {0u, 0u, 0u};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!((MASK >> a) & 1)) continue;
            const int kk = min(max(kc - pK0[a], 0), pKn[a] - 1);
            xoff[a] = inside ? (unsigned)((pBase[a] + (long long)kk * pKs[a]) * (long long)sizeof(S)) : 0u;
            const Rsrc rXa = make_rsrc((const S*)xA[a], 9 * xN[a]);
#pragma unroll
            for (int v = 0; v < 9; ++v) xa[a][v] = PR::f1(PR::template ld1<CPX>(rXa, xoff[a], v * xn4[a]));
            if constexpr (IRK == 1 || IRK == 2) {
                const Rsrc rXm = make_rsrc((const S*)xM[a], 9 * xN[a]);
#pragma unroll
                for (int v = 0; v < 9; ++v) xm[a][v] = PR::f1(PR::template ld1<CPX>(rXm, xoff[a], v * xn4[a]));
            }
            if constexpr (IRK >= 2) {
                const Rsrc rXt = make_rsrc((const S*)xT[a], 9 * xN[a]);
#pragma unroll
                for (int v = 0; v < 9; ++v) xt[a][v] = PR::f1(PR::template ld1<CPX>(rXt, xoff[a], v * xn4[a]));
            }
            if (a == 2) {
                pBeta[2] = p.pb[2][kc];
                pD[2] = p.pd[2][kc];
                pAD[2] = pD[2] + p.pa[2][kc];
            }
        }
        if (inside && active) {
            const float xix = PR::lo(M[0]), ety = PR::hi(M[0]);
            const float kz = (float)(kc - HALO);
            const float ztx = zax * kz, zty = zay * kz, ztz = zgz;
            const float mu = PR::lo(M[1]) * DT, lam = PR::hi(M[1]) * DT, bu = PR::f1(Mb) * (Crho * DT);
            // physical derivative d/dx_c from the three one-sided ones
            auto phys = [&](int c, float d0, float d1, float d2) {
                return c == 0 ? xix * d0 + ztx * d2 : c == 1 ? ety * d1 + zty * d2 : ztz * d2;
            };
            const int ox0 = lc - FB1, ox2 = lc + FB1, ox3 = lc + 2 * FB1, ox4 = lc + 3 * FB1;
            const int oy0 = lc - FB2 * LX, oy2 = lc + FB2 * LX, oy3 = lc + 2 * FB2 * LX, oy4 = lc + 3 * FB2 * LX;
            // The 27 one-sided derivatives are folded into the physical velocity gradient
            // G[a][c] = sum_d m^d_c dV_a/dxi_d and the stress divergence dv[a] as they are formed,
            // so only the three derivatives of one variable are live at a time.
            float G[3][3];
            float dv[3] = {0.f, 0.f, 0.f};
            [[maybe_unused]] float draw[3][9];  // raw derivatives along the slab normals
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                // value of variable v at LDS cell o
                auto Lv = [&](int o) -> float {
                    if constexpr (PACKED) {
                        if (v == 8) return PR::f1(__builtin_bit_cast(typename PR::one, (unsigned short)Lb[4 * VS + o]));
                        const P2 x = __builtin_bit_cast(P2, Lb[(v >> 1) * VS + o]);
                        return (v & 1) ? PR::hi(x) : PR::lo(x);
                    } else {
                        return Lb[v * VS + o];
                    }
                };
                float c;
                if constexpr (PACKED) c = q1[v];
                else c = cen[v];
                float d0, d1, d2;
                if constexpr (PACKED) {
                    // FMA chains from the centre term (c converted once for the three axes), so
                    // every other stored half enters a mixed-precision FMA without a conversion
                    d0 = __builtin_fmaf(cx0, Lv(ox0), cx1 * c);
                    d0 = __builtin_fmaf(cx2, Lv(ox2), d0);
                    d0 = __builtin_fmaf(cx3, Lv(ox3), d0);
                    d0 = __builtin_fmaf(cx4, Lv(ox4), d0);
                    d1 = __builtin_fmaf(cy0, Lv(oy0), cy1 * c);
                    d1 = __builtin_fmaf(cy2, Lv(oy2), d1);
                    d1 = __builtin_fmaf(cy3, Lv(oy3), d1);
                    d1 = __builtin_fmaf(cy4, Lv(oy4), d1);
                    d2 = __builtin_fmaf(cz0, q0[v], cz1 * c);
                    d2 = __builtin_fmaf(cz2, q2[v], d2);
                    d2 = __builtin_fmaf(cz3, q3[v], d2);
                    d2 = __builtin_fmaf(cz4, q4[v], d2);
                } else {
                    d0 = cx0 * Lv(ox0) + cx1 * c + cx2 * Lv(ox2) + cx3 * Lv(ox3) + cx4 * Lv(ox4);
                    d1 = cy0 * Lv(oy0) + cy1 * c + cy2 * Lv(oy2) + cy3 * Lv(oy3) + cy4 * Lv(oy4);
                    d2 = cz0 * q0[v] + cz1 * c + cz2 * q2[v] + cz3 * q3[v] + cz4 * q4[v];
                }
                if constexpr ((MASK & 1) != 0) { draw[0][v] = d0; d0 *= pBeta[0]; }
                if constexpr ((MASK & 2) != 0) { draw[1][v] = d1; d1 *= pBeta[1]; }
                if constexpr ((MASK & 4) != 0) { draw[2][v] = d2; d2 *= pBeta[2]; }
                if (v < 3) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) G[v][cc] = phys(cc, d0, d1, d2);
                } else {
                    // stress component (row r0, column c0) contributes d/dx_c0 to dv[r0] and, off the
                    // diagonal, d/dx_r0 to dv[c0]
                    const int r0 = v == 3 ? 0 : v == 4 ? 1 : v == 5 ? 2 : v == 6 ? 0 : v == 7 ? 0 : 1;
                    const int c0 = v == 3 ? 0 : v == 4 ? 1 : v == 5 ? 2 : v == 6 ? 1 : v == 7 ? 2 : 2;
                    dv[r0] += phys(c0, d0, d1, d2);
                    if (v >= 6) dv[c0] += phys(r0, d0, d1, d2);
                }
            }
            float h[9];
            h[0] = dv[0] * bu;
            h[1] = dv[1] * bu;
            h[2] = dv[2] * bu;
            const float div = lam * (G[0][0] + G[1][1] + G[2][2]);
            const float mu2 = 2.0f * mu;
            h[3] = div + mu2 * G[0][0];
            h[4] = div + mu2 * G[1][1];
            h[5] = div + mu2 * G[2][2];
            h[6] = mu * (G[0][1] + G[1][0]);
            h[7] = mu * (G[0][2] + G[2][0]);
            h[8] = mu * (G[1][2] + G[2][1]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (!((MASK >> a) & 1)) continue;
                // -beta*Aux*DT into the increment (the registers hold Aux*DT), then the
                // single-direction right-hand side (with d, no beta) of the slab normal
#pragma unroll
                for (int v = 0; v < 9; ++v) h[v] += -pBeta[a] * xa[a][v];
                // metric row of the slab normal (zero entries as above)
                const float mx = a == 0 ? xix : a == 2 ? ztx : 0.f;
                const float my = a == 1 ? ety : a == 2 ? zty : 0.f;
                const float mz = a == 2 ? ztz : 0.f;
                float b[9];
#pragma unroll
                for (int v = 0; v < 9; ++v) b[v] = draw[a][v] * pD[a];
                float ha[9];
                ha[0] = (mx * b[3] + my * b[6] + mz * b[7]) * bu;
                ha[1] = (mx * b[6] + my * b[4] + mz * b[8]) * bu;
                ha[2] = (mx * b[7] + my * b[8] + mz * b[5]) * bu;
                const float pdv = (mx * b[0] + my * b[1] + mz * b[2]) * lam;
                ha[3] = pdv + mu2 * (mx * b[0]);
                ha[4] = pdv + mu2 * (my * b[1]);
                ha[5] = pdv + mu2 * (mz * b[2]);
                ha[6] = (my * b[0] + mx * b[1]) * mu;
                ha[7] = (mz * b[0] + mx * b[2]) * mu;
                ha[8] = (mz * b[1] + my * b[2]) * mu;
                const Rsrc rXa = make_rsrc((const S*)xA[a], 9 * xN[a]);
                const Rsrc rXm = make_rsrc((const S*)xM[a], 9 * xN[a]);
                const Rsrc rXt = make_rsrc((const S*)xT[a], 9 * xN[a]);
                const unsigned xo = xoff[a];
                const int x4 = xn4[a];
#pragma unroll
                for (int v = 0; v < 9; ++v) {
                    // increment of Aux*DT: DT * (rhs*DT - (alpha+d)*Aux*DT)
                    const float hav = (ha[v] - pAD[a] * xa[a][v]) * DT;
                    // as for the wavefield, the accumulator after stage 0 is rebuilt in stage 1
                    if constexpr (IRK == 0) {
                        PR::template st1<CPX>(rXm, xo, v * x4, xa[a][v]);
                        PR::template st1<CPX>(rXa, xo, v * x4, xa[a][v] + alpha2 * hav);
                    } else if constexpr (IRK == 1) {
                        PR::template st1<CPX>(rXt, xo, v * x4, (xm[a][v] + (xa[a][v] - xm[a][v]) * kT0) + beta2 * hav);
                        PR::template st1<CPX>(rXa, xo, v * x4, xm[a][v] + alpha3 * hav);
                    } else if constexpr (IRK == 2) {
                        PR::template st1<CPX>(rXt, xo, v * x4, xt[a][v] + beta3 * hav);
                        PR::template st1<CPX>(rXa, xo, v * x4, xm[a][v] + hav);
                    } else {
                        PR::template st1<CPX>(rXa, xo, v * x4, xt[a][v] + beta4 * hav);
                    }
                }
            }
            // RK registers (wave_rk0..3); the accumulator after stage 0 is rebuilt in stage 1
            // from this stage's input (kT0)
            const Rsrc rO = make_rsrc((const S*)(Ob + pw), wrange);
            float o[9], t[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                // the stage input is taken again from the packed queue slot (q1), so the FP32
                // copies of the centre values need not live through the derivative phase
                if constexpr (IRK == 0) {
                    o[v] = q1[v] + alpha2 * h[v];
                } else if constexpr (IRK == 1) {
                    o[v] = ra[v] + alpha3 * h[v];
                    t[v] = (ra[v] + (q1[v] - ra[v]) * kT0) + beta2 * h[v];
                } else if constexpr (IRK == 2) {
                    o[v] = ra[v] + h[v];
                    t[v] = rt[v] + beta3 * h[v];
                } else {
                    o[v] = rt[v] + beta4 * h[v];
                }
            }
            store9<S, PB, CPO>(rO, col, col1, o);
            if constexpr (IRK == 1 || IRK == 2) store9<S, PB, CPS>(rT, col, col1, t);
        }
        // the slot of the oldest plane receives the new one
        q0 = nq;
        if constexpr (PREFETCH) {
#pragma unroll
            for (int r = 0; r < HPT; ++r) hv[r] = hvN[r];
#pragma unroll
            for (int c = 0; c < 2; ++c) M[c] = MN[c];
            Mb = MbN;
        }
        k += FB3;
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    for (int it = 0; it < nk; it += 5) {
        step(I0{}, it);
        step(I1{}, it + 1);
        step(I2{}, it + 2);
        step(I3{}, it + 3);
        step(I4{}, it + 4);
    }
}

// Single-domain launch: a 1-D grid of gx*gy*nch workgroups (x tiles fastest, then y tiles, then
// z chunks of kc planes), dealt to the XCDs in contiguous runs (xcd_order). One instantiation per
// RK stage.
template <class S, int TX, int TY, int IRK>
// Occupancy at least 4 waves per SIMD (at most 128 VGPRs; the launch bound's second argument is
// the minimum number of waves per execution unit): two 64x8 workgroups per CU, as the LDS allows;
// left to itself the compiler gave one RK stage 129 VGPRs and occupancy 3 (6 % slower).
__global__ void __launch_bounds__(TX* TY, 4)
k_stage_bulk(StageParams p, int kc, int gx, int gy) {
    __shared__ BulkLds<S, TX, TY> lds;
    const long long t = xcd_order(blockIdx.x, (long long)gridDim.x);
    const int bx = (int)(t % gx), by = (int)((t / gx) % gy), bz = (int)(t / ((long long)gx * gy));
    const int kb = p.bk0 + bz * kc;
    const int ke = min(kb + kc, p.bk1);
    bulk_tile<S, TX, TY, IRK>(p, p.FB1, p.FB2, p.FB3, p.bi0, p.bi1, p.bj0, p.bj1, 0, bx, by, kb, ke, lds.v);
}

// Tiling of the PML boxes: entry e is box p.pbox[box[e]], cut into gx*gy column tiles and nch
// chunks of kc planes; entries are grouped by box kind (pml_mask_at order, group g holds entries
// first[g] .. first[g+1]-1) and tile indices run consecutively over all entries (start[e]). One
// launch per kind covers the tiles start[first[g]] .. start[first[g+1]]-1.
struct PmlTiles {
    int n;
    int first[kNumPmlMasks + 1];
    int box[kMaxPBoxes];
    int gx[kMaxPBoxes], gy[kMaxPBoxes], nch[kMaxPBoxes], kc[kMaxPBoxes];
    int start[kMaxPBoxes + 1];
};

// Tile shape of a PML box kind: 16x8 tiles aligned inside a layout block for boxes in an x slab
// (12 of 16 columns used instead of 12 of 64), 64x4 otherwise. 16x8 rather than 16x16: twice the
// workgroups over the narrow x slabs of small grids (one served request alone 4 % faster, the
// batched stream 1 %; 1001x1001x401 unchanged); 16x4 and 16x32 were slower.
__host__ __device__ __forceinline__ constexpr int pml_tx(int mask) { return (mask & 1) ? 16 : 64; }
__host__ __device__ __forceinline__ constexpr int pml_ty(int mask) { return (mask & 1) ? 8 : 4; }

// Entry of tile t (a global tile index) among entries e0 .. e1-1.
__device__ __forceinline__ int pml_entry(const PmlTiles& tl, int e0, int e1, int t) {
    int e = e0;
    while (e + 1 < e1 && t >= tl.start[e + 1]) ++e;
    return e;
}

// Tile (bx, by, chunk) of entry e.
template <class S, int TX, int TY, int IRK, int MASK>
__device__ __forceinline__ void pml_tile(const StageParams& p, const PmlTiles& tl, int e, int r,
                                         typename BulkLds<S, TX, TY>::array& lds) {
    const int bx = r % tl.gx[e], by = (r / tl.gx[e]) % tl.gy[e], bz = r / (tl.gx[e] * tl.gy[e]);
    const int* B = p.pbox[tl.box[e]];
    const int kb = B[4] + bz * tl.kc[e];
    const int ke = min(kb + tl.kc[e], B[5]);
    bulk_tile<S, TX, TY, IRK, MASK>(p, p.FB1, p.FB2, p.FB3, B[0], B[1], B[2], B[3], B[7], bx, by, kb, ke, lds);
}

template <class S, int TX, int TY, int IRK, int MASK>
__global__ void __launch_bounds__(TX* TY) k_stage_pml(StageParams p, PmlTiles tl) {
    __shared__ BulkLds<S, TX, TY> lds;
    constexpr int g = pml_group_of(MASK);
    const int e0 = tl.first[g], e1 = tl.first[g + 1];
    const int t = tl.start[e0] + (int)xcd_order(blockIdx.x, (long long)gridDim.x);
    const int e = pml_entry(tl, e0, e1, t);
    pml_tile<S, TX, TY, IRK, MASK>(p, tl, e, t - tl.start[e], lds.v);
}

}  // namespace cg
