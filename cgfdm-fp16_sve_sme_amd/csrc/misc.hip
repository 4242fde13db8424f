// misc.hip — per-step side kernels: sources, stations, PGV, slices and state transfer.
// Point lists carry padded (i, j, k) triples; addresses come from the Layout.
#include "side.hpp"

namespace cg {

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
#define CG_LAUNCH_OK (hipGetLastError() == hipSuccess ? 0 : -3)

// Explosive point source before stage 0 of step it; the Ricker value is formed on the device.
template <class S>
__global__ void k_point_source(S* A, S* B, S* C, Layout L, int SX, int SY, int SZ, const float* __restrict__ coef,
                               float fc, int it, float DT) {
    const int q = threadIdx.x;
    if (q >= 343) return;
    point_source_at<S>(A, B, C, L, SX, SY, SZ, coef, ricker_value(fc, it, DT), q);
}

// Moment-rate table source before stage 0 of step it.
template <class S>
__global__ void k_multi_source(S* A, Layout L, const int* __restrict__ ijk, const float* __restrict__ mrate,
                               long long npts, int snt, float sdt, int it, float DT) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int srcIt;
    float tw;
    if (p >= npts || !table_time(it, DT, sdt, snt, srcIt, tw)) return;
    table_source_at<S>(A, L, ijk, mrate, npts, srcIt, tw, DT, p);
}

// Scan of a wavefield register for values that are not finite (infinite or NaN: the exponent
// field all ones), as raw 32-bit words: one FP32 value or two 16-bit values per word, emask the
// exponent field of every value in a word. A word with such a value writes `code` into *flag
// unless the flag is already set (all lanes of one launch write the same code; launches on one
// stream are ordered, so the first step found is kept). The padding and halo elements of the
// register are zero or finite, so the whole array is scanned without its layout.
__global__ void k_nonfinite(const unsigned* __restrict__ w, long long nwords, unsigned emask, int* flag, int code) {
    const Rsrc rf = make_rsrc(flag, 1);
    bool bad = false;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nwords;
         q += (long long)gridDim.x * blockDim.x) {
        const unsigned x = w[q];
        const unsigned lo = emask & 0xffffu, hi = emask & 0xffff0000u;
        bad |= (lo != 0u && (x & lo) == lo) || (hi != 0u && (x & hi) == hi);
    }
    if (bad && *(volatile const int*)flag == 0) __builtin_amdgcn_raw_buffer_store_b32((unsigned)code, rf, 0, 0, 0);
}

// Coefficients of the 7^3 point-source stencil from the stored Jacobian (no host round trip).
template <class S>
__global__ void k_source_coef(const S* __restrict__ cjm, Layout L, int SX, int SY, int SZ,
                              const float* __restrict__ amp, float DH, double CsDT, float* __restrict__ coef) {
    const int q = threadIdx.x;
    if (q < 343) coef[q] = source_coef_at<S>(cjm, L, SX, SY, SZ, amp[q], DH, CsDT);
}

// One-time scaling of the table by mu*Cv/(J*DH^3) read from the stored CJM
// (calculate_MomentRate, addMoment.cpp:11-65).
template <class S>
__global__ void k_scale_mrate(const S* __restrict__ cjm, Layout L, const int* __restrict__ ijk, float* mrate,
                              long long npts, int nt, float DH, float Cv) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= npts * nt) return;
    const long long p = t % npts;
    const int i = ijk[3 * p], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
    const float jacb = SV<S>::ld(cjm[L.atc(9, i, j, k)]);
    const float mu = SV<S>::ld(cjm[L.atc(10, i, j, k)]);
    float V = jacb * DH * DH * DH;
    V = 1.0f / V;
    V = V * Cv;
    for (int c = 0; c < 6; ++c) mrate[t * 6 + c] *= mu * V;
}

// Station sampling after step it.
template <class S>
__global__ void k_station(const S* __restrict__ A, Layout L, const int* __restrict__ ijk, int nst, float* wave,
                          int NT, int it, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nst * 9) return;
    station_at<S>(A, L, ijk, wave, NT, it, c1, c2, t);
}

// PGV update after a step.
template <class S>
__global__ void k_pgv(const S* __restrict__ A, Layout L, int nx, int ny, int k, float* __restrict__ pgv, float c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    pgv_at<S>(A, L, nx, k, pgv, c, t);
}

// Plane gather with unscaling (pack_iodata_x/y/z, data_io.cpp:86-231).
template <class S>
__global__ void k_slice(const S* __restrict__ A, Layout L, int nx, int ny, int nz, int axis, int index, int var,
                        float c, float* __restrict__ out) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    long long n;
    if (axis == 0) { n = (long long)ny * nz; i = index; j = (int)(t % ny); k = (int)(t / ny); }
    else if (axis == 1) { n = (long long)nx * nz; j = index; i = (int)(t % nx); k = (int)(t / nx); }
    else { n = (long long)nx * ny; k = index; i = (int)(t % nx); j = (int)(t / nx); }
    if (t >= n) return;
    out[t] = SV<S>::ld(A[L.at(9, var, i + HALO, j + HALO, k + HALO)]) * c;
}

template <class S>
__global__ void k_gather(const S* __restrict__ A, Layout L, int V, int nx, int ny, int nz, float* __restrict__ out) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    if (t >= n * V) return;
    const int v = (int)(t / n);
    const long long r = t % n;
    const int i = (int)(r % nx), j = (int)((r / nx) % ny), k = (int)(r / ((long long)nx * ny));
    out[t] = SV<S>::ld(A[L.at(V, v, i + HALO, j + HALO, k + HALO)]);
}

template <class S>
__global__ void k_scatter(S* A, Layout L, int V, int nx, int ny, int nz, const float* __restrict__ in) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    if (t >= n * V) return;
    const int v = (int)(t / n);
    const long long r = t % n;
    const int i = (int)(r % nx), j = (int)((r / nx) % ny), k = (int)(r / ((long long)nx * ny));
    A[L.at(V, v, i + HALO, j + HALO, k + HALO)] = SV<S>::st(in[t]);
}

// Whole padded array to the reference order (V, nz_, ny_, nx_).
// V == 13 is the metric/medium array: out in the reference component order.
template <class S>
__global__ void k_unpack(const S* __restrict__ A, Layout L, int V, float* __restrict__ out) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)L.nx_ * L.ny_ * L.nz_;
    if (t >= n * V) return;
    const int v = (int)(t / n);
    const long long r = t % n;
    const int i = (int)(r % L.nx_), j = (int)((r / L.nx_) % L.ny_), k = (int)(r / ((long long)L.nx_ * L.ny_));
    out[t] = SV<S>::ld(A[V == 13 ? L.atc(v, i, j, k) : L.at(V, v, i, j, k)]);
}

template <class S>
__global__ void k_decode(const S* __restrict__ src, float* __restrict__ dst, long long n) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n) dst[t] = SV<S>::ld(src[t]);
}

#define CG_DISPATCH(prec, KERNEL, GRID, BLOCK, STREAM, PTRS_F32, PTRS_F16, PTRS_BF16, ...)              \
    switch (prec) {                                                                                   \
        case P_F32: KERNEL<float><<<GRID, BLOCK, 0, STREAM>>>(PTRS_F32, __VA_ARGS__); break;         \
        case P_F16: KERNEL<_Float16><<<GRID, BLOCK, 0, STREAM>>>(PTRS_F16, __VA_ARGS__); break;      \
        case P_BF16: KERNEL<__bf16><<<GRID, BLOCK, 0, STREAM>>>(PTRS_BF16, __VA_ARGS__); break;      \
        default: return -1;                                                                           \
    }

int launch_point_source(int prec, void* A, void* B, void* C, const Layout& L, int SX, int SY, int SZ,
                        const float* coef, float fc, int it, float DT, hipStream_t s) {
    switch (prec) {
        case P_F32: k_point_source<float><<<1, 384, 0, s>>>((float*)A, (float*)B, (float*)C, L, SX, SY, SZ, coef, fc, it, DT); break;
        case P_F16: k_point_source<_Float16><<<1, 384, 0, s>>>((_Float16*)A, (_Float16*)B, (_Float16*)C, L, SX, SY, SZ, coef, fc, it, DT); break;
        case P_BF16: k_point_source<__bf16><<<1, 384, 0, s>>>((__bf16*)A, (__bf16*)B, (__bf16*)C, L, SX, SY, SZ, coef, fc, it, DT); break;
        default: return -1;
    }
    return CG_LAUNCH_OK;
}

int launch_multi_source(int prec, void* A, const Layout& L, const int* ijk, const float* mrate, long long npts,
                        int snt, float sdt, int it, float DT, hipStream_t s) {
    CG_DISPATCH(prec, k_multi_source, nblk(npts, 256), 256, s, (float*)A, (_Float16*)A, (__bf16*)A, L, ijk, mrate,
                npts, snt, sdt, it, DT);
    return CG_LAUNCH_OK;
}

int launch_source_coef(int prec, const void* cjm, const Layout& L, int SX, int SY, int SZ, const float* amp, float DH,
                       double CsDT, float* coef, hipStream_t s) {
    CG_DISPATCH(prec, k_source_coef, 1, 384, s, (const float*)cjm, (const _Float16*)cjm, (const __bf16*)cjm, L, SX,
                SY, SZ, amp, DH, CsDT, coef);
    return CG_LAUNCH_OK;
}

int launch_scale_mrate(int prec, const void* cjm, const Layout& L, const int* ijk, float* mrate, long long npts,
                       int nt, float DH, float Cv, hipStream_t s) {
    CG_DISPATCH(prec, k_scale_mrate, nblk(npts * nt, 256), 256, s, (const float*)cjm, (const _Float16*)cjm,
                (const __bf16*)cjm, L, ijk, mrate, npts, nt, DH, Cv);
    return CG_LAUNCH_OK;
}

int launch_station(int prec, const void* A, const Layout& L, const int* ijk, int nst, float* wave, int NT, int it,
                   float c1, float c2, hipStream_t s) {
    CG_DISPATCH(prec, k_station, nblk((long long)nst * 9, 256), 256, s, (const float*)A, (const _Float16*)A,
                (const __bf16*)A, L, ijk, nst, wave, NT, it, c1, c2);
    return CG_LAUNCH_OK;
}

int launch_nonfinite(int prec, const void* A, const Layout& L, int* flag, int code, hipStream_t s) {
    const long long nwords = L.size(9) * storage_bytes(prec) / 4;
    // FP32: one exponent field per word; FP16 (5 bits) and BF16 (8 bits): one per half word
    const unsigned emask = prec == P_F32 ? 0x7f800000u : prec == P_F16 ? 0x7c007c00u : 0x7f807f80u;
    const long long want = 2048LL * 256;  // grid-stride: at most 2048 workgroups
    const unsigned grid = nblk(nwords < want ? nwords : want, 256);
    k_nonfinite<<<grid, 256, 0, s>>>((const unsigned*)A, nwords, emask, flag, code);
    return CG_LAUNCH_OK;
}

int launch_pgv(int prec, const void* A, const Layout& L, int nx, int ny, int k, float* pgv, float c, hipStream_t s) {
    CG_DISPATCH(prec, k_pgv, nblk((long long)nx * ny, 256), 256, s, (const float*)A, (const _Float16*)A,
                (const __bf16*)A, L, nx, ny, k, pgv, c);
    return CG_LAUNCH_OK;
}

int launch_slice(int prec, const void* A, const Layout& L, int nx, int ny, int nz, int axis, int index, int var,
                 float c, float* out, hipStream_t s) {
    long long n = axis == 0 ? (long long)ny * nz : (axis == 1 ? (long long)nx * nz : (long long)nx * ny);
    CG_DISPATCH(prec, k_slice, nblk(n, 256), 256, s, (const float*)A, (const _Float16*)A, (const __bf16*)A, L, nx,
                ny, nz, axis, index, var, c, out);
    return CG_LAUNCH_OK;
}

int launch_gather(int prec, const void* A, const Layout& L, int V, int nx, int ny, int nz, float* out,
                  hipStream_t s) {
    long long n = (long long)nx * ny * nz * V;
    CG_DISPATCH(prec, k_gather, nblk(n, 256), 256, s, (const float*)A, (const _Float16*)A, (const __bf16*)A, L, V,
                nx, ny, nz, out);
    return CG_LAUNCH_OK;
}

int launch_scatter(int prec, void* A, const Layout& L, int V, int nx, int ny, int nz, const float* in,
                   hipStream_t s) {
    long long n = (long long)nx * ny * nz * V;
    CG_DISPATCH(prec, k_scatter, nblk(n, 256), 256, s, (float*)A, (_Float16*)A, (__bf16*)A, L, V, nx, ny, nz, in);
    return CG_LAUNCH_OK;
}

int launch_unpack(int prec, const void* src, const Layout& L, int V, float* dst, hipStream_t s) {
    long long n = (long long)L.nx_ * L.ny_ * L.nz_ * V;
    CG_DISPATCH(prec, k_unpack, nblk(n, 256), 256, s, (const float*)src, (const _Float16*)src, (const __bf16*)src,
                L, V, dst);
    return CG_LAUNCH_OK;
}

int launch_decode(int prec, const void* src, float* dst, long long n, hipStream_t s) {
    CG_DISPATCH(prec, k_decode, nblk(n, 256), 256, s, (const float*)src, (const _Float16*)src, (const __bf16*)src, dst,
                n);
    return CG_LAUNCH_OK;
}

}  // namespace cg
