// Host build of the library's stage kernels for the memory-access checker (bounds_check.cpp).
//
// The kernel headers of the library (common.hpp, stage_bulk.hpp, stage_fs.hpp,
// stage_general.hpp) are compiled for the CPU with this file in place of the HIP runtime header:
// the device vocabulary they use is mapped to plain C++, a kernel runs one thread at a time with
// threadIdx/blockIdx set by the caller, and barriers are no-ops (the LDS contents, which no
// address depends on, are then meaningless). Buffer loads and stores check their offsets against
// the size the descriptor was made with and the address against the registered device arrays,
// and move no data (loads return 0); plain pointer accesses touch real host memory laid out as
// the device arrays, with inaccessible guard pages around each.
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
// LDS: one copy per CPU thread; EMU_LDS_STATIC makes it one ordinary static array, which
// AddressSanitizer instruments (for a single-threaded run that checks LDS indices as well)
#ifdef EMU_LDS_STATIC
#define __shared__ static
#else
#define __shared__ static thread_local
#endif
#define __launch_bounds__(...)

struct emu_dim {
    unsigned x = 0, y = 0, z = 0;
};
extern thread_local emu_dim threadIdx, blockIdx, blockDim, gridDim;
inline void __syncthreads() {}

typedef struct emu_stream_* hipStream_t;
typedef struct emu_event_* hipEvent_t;
typedef int hipError_t;

template <class T> inline T min(T a, T b) { return b < a ? b : a; }
template <class T> inline T max(T a, T b) { return a < b ? b : a; }

// ---- buffer resources
struct emu_rsrc {
    const char* base;
    long long bytes;
};
#define __amdgpu_buffer_rsrc_t emu_rsrc

// Access log: counts and the first reports of offsets past a descriptor's size (the hardware
// drops such accesses; this library never intends one) and of addresses outside every
// registered array (a wild descriptor base).
struct emu_log {
    std::atomic<long long> oob{0}, wild{0}, accesses{0};
    const char* kernel = "";
};
extern emu_log g_emu;
extern thread_local long long t_emu_accesses;  // this thread's count, folded into g_emu by the runner
bool emu_registered(const char* p, long long n);  // [p, p+n) inside one registered array
void emu_report(const char* what, const emu_rsrc& r, long long off, int size);

inline emu_rsrc emu_make_rsrc(const void* p, int stride, int bytes, int) {
    (void)stride;
    return emu_rsrc{(const char*)p, (long long)bytes};
}
inline void emu_access(const emu_rsrc& r, int voff, int soff, int size) {
    ++t_emu_accesses;
    const long long off = (long long)(unsigned)voff + (long long)soff;
    if (soff < 0 || off + size > r.bytes) {
        if (g_emu.oob.fetch_add(1) < 8) emu_report("offset past the descriptor", r, off, size);
        return;
    }
    if (!emu_registered(r.base + off, size)) {
        if (g_emu.wild.fetch_add(1) < 8) emu_report("address outside every device array", r, off, size);
    }
}
#define __builtin_amdgcn_make_buffer_rsrc(p, stride, n, flags) emu_make_rsrc((p), (stride), (n), (flags))
#define __builtin_amdgcn_raw_buffer_load_b16(r, v, s, cp) (emu_access((r), (v), (s), 2), (unsigned short)0)
#define __builtin_amdgcn_raw_buffer_load_b32(r, v, s, cp) (emu_access((r), (v), (s), 4), 0u)
#define __builtin_amdgcn_raw_buffer_load_b64(r, v, s, cp) \
    (emu_access((r), (v), (s), 8), (unsigned __attribute__((ext_vector_type(2)))){0u, 0u})
// Stores: with g_emu_store_data set (the write-coverage mode) an in-range store also writes its
// bytes, so the caller can see which elements a stage wrote.
extern bool g_emu_store_data;
template <class T> inline void emu_store(const emu_rsrc& r, int voff, int soff, T x) {
    emu_access(r, voff, soff, (int)sizeof(T));
    const long long off = (long long)(unsigned)voff + (long long)soff;
    if (g_emu_store_data && soff >= 0 && off + (long long)sizeof(T) <= r.bytes &&
        emu_registered(r.base + off, (long long)sizeof(T)))
        std::memcpy((void*)(r.base + off), &x, sizeof(T));
}
#define __builtin_amdgcn_raw_buffer_store_b16(x, r, v, s, cp) emu_store((r), (v), (s), (unsigned short)(x))
#define __builtin_amdgcn_raw_buffer_store_b32(x, r, v, s, cp) emu_store((r), (v), (s), (unsigned)(x))
#define __builtin_amdgcn_raw_buffer_store_b64(x, r, v, s, cp) emu_store((r), (v), (s), (x))
#define __builtin_amdgcn_readfirstlane(x) (x)
