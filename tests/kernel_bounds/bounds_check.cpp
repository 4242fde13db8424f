// Memory-access checker of the fused RK-stage kernels, on the CPU.
//
// Builds a domain's device arrays the way the library lays them out (make_layout, slab_points,
// assign_slabs, domain_stage_params), decomposes the interior as a stage launch does
// (stage_boxes, bulk_geom, pml_tiles) and runs every workgroup and thread of the bulk, PML,
// free-surface and general kernels of all four RK stages, with the kernel code of the library
// compiled for the host (hip/hip_runtime.h beside this file). Every buffer access is checked
// against its descriptor and the registered arrays; every plain pointer access touches host
// memory laid out like the device arrays, flush against an inaccessible guard region, so an
// access outside an array stops the run with a report of the address, the nearest array and the
// kernel. Exit status 0: no access outside any array.
//
// usage: bounds_check NX NY NZ nPML {fp32|fp16|bf16} free_surface [tile_cfg|-1] [kc] [--boundary-only]
//                     [--steps N] [--shrink-arrays]
// --shrink-arrays registers the wavefield arrays one k-plane short of their layout (a planted
// mistake: the run must then report accesses outside the arrays).
// --coverage also checks what the stage kernels write: the stage output register is filled with
// a sentinel byte pattern before every stage, and afterwards every interior element of its 9
// variables must have been written and no other element (halo, row padding) touched; likewise the
// PML auxiliaries' stage-0 copy in stage 0 and their accumulator in stage 1, at every slab point.
// Stores move their bytes in this mode; loads through descriptors still return 0.
// --plant-coverage (with --coverage) resets one interior output element to the sentinel after
// each stage (a planted mistake: the check must then report an element that was not written).
#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "domain.hpp"
#include "stage_bulk.hpp"
#include "stage_fs.hpp"
#include "stage_general.hpp"
#include "stage_geom.hpp"

thread_local emu_dim threadIdx, blockIdx, blockDim, gridDim;
thread_local long long t_emu_accesses = 0;
emu_log g_emu;
bool g_emu_store_data = false;

using namespace cg;

namespace {

struct Arr {
    const char* p;
    long long n;
    const char* name;
};
std::vector<Arr> g_arrs;
const size_t kGuard = 64ull << 20;  // inaccessible bytes before and after every array

// A device array of n bytes: flush against the guard region above it, 256-byte aligned.
void* dev_alloc(size_t n, const char* name) {
    const size_t page = 4096, body = (n + page - 1) / page * page;
    char* m = (char*)mmap(nullptr, body + 2 * kGuard, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) { perror("mmap"); exit(2); }
    if (mprotect(m + kGuard, body, PROT_READ | PROT_WRITE) != 0) { perror("mprotect"); exit(2); }
    char* p = m + kGuard + ((body - n) & ~(size_t)255);
    g_arrs.push_back({p, (long long)n, name});
    return p;
}

const Arr* nearest(const char* a) {
    const Arr* best = nullptr;
    long long bd = -1;
    for (const Arr& r : g_arrs) {
        const long long d = a < r.p ? r.p - a : (a >= r.p + r.n ? a - (r.p + r.n) + 1 : 0);
        if (!best || d < bd) { best = &r; bd = d; }
    }
    return best;
}

void on_segv(int, siginfo_t* si, void*) {
    const char* a = (const char*)si->si_addr;
    const Arr* r = nearest(a);
    char buf[512];
    const int n = snprintf(buf, sizeof buf,
                           "FAULT: plain access at %p in kernel %s; nearest array %s [%p, +%lld), offset %lld\n",
                           (const void*)a, g_emu.kernel, r ? r->name : "-", r ? (const void*)r->p : nullptr,
                           r ? r->n : 0LL, r ? (long long)(a - r->p) : 0LL);
    if (write(2, buf, n) < 0) {}
    _exit(3);
}

}  // namespace

bool emu_registered(const char* p, long long n) {
    for (const Arr& r : g_arrs)
        if (p >= r.p && p + n <= r.p + r.n) return true;
    return false;
}

void emu_report(const char* what, const emu_rsrc& r, long long off, int size) {
    const Arr* a = nearest(r.base);
    fprintf(stderr, "%s: kernel %s, descriptor base %p (%s%+lld) size %lld, offset %lld, %d bytes, block %u thread %u\n",
            what, g_emu.kernel, (const void*)r.base, a ? a->name : "-", a ? (long long)(r.base - a->p) : 0LL,
            r.bytes, off, size, blockIdx.x, threadIdx.x);
}

namespace {

template <class F>
void run_grid(const char* name, long long grid, int threads, F&& body) {
    g_emu.kernel = name;
#pragma omp parallel
    {
#pragma omp for schedule(dynamic, 4)
        for (long long b = 0; b < grid; ++b) {
            blockIdx.x = (unsigned)b;
            gridDim.x = (unsigned)grid;
            blockDim.x = (unsigned)threads;
            for (int t = 0; t < threads; ++t) {
                threadIdx.x = (unsigned)t;
                body();
            }
        }
        g_emu.accesses += t_emu_accesses;
        t_emu_accesses = 0;
    }
}

template <class S, int TX, int TY, int IRK>
void bulk(const StageParams& q) {
    const BulkGeom g = bulk_geom(q);
    run_grid("k_stage_bulk", (long long)g.gx * g.gy * g.nch, TX * TY,
             [&] { k_stage_bulk<S, TX, TY, IRK>(q, g.kc, g.gx, g.gy); });
}

template <class S, int IRK, int MASK>
void pml(const StageParams& q, const PmlTiles& t) {
    constexpr int g = pml_group_of(MASK), TX = pml_tx(MASK), TY = pml_ty(MASK);
    const int tiles = t.start[t.first[g + 1]] - t.start[t.first[g]];
    if (tiles <= 0) return;
    static const char* names[8] = {"", "k_stage_pml x faces", "k_stage_pml y faces", "k_stage_pml xy edges",
                                   "k_stage_pml z face", "k_stage_pml xz edges", "k_stage_pml yz edges",
                                   "k_stage_pml corners"};
    run_grid(names[MASK], tiles, TX * TY, [&] { k_stage_pml<S, TX, TY, IRK, MASK>(q, t); });
}

template <class S, int IRK>
void stage(const StageParams& p, bool boundary_only) {
    StageParams q = p;
    stage_boxes(q, 1);
    if (q.bulk_on && !boundary_only) {
        switch (bulk_geom(q).cfg) {
            case 0: bulk<S, 64, 8, IRK>(q); break;
            case 1: bulk<S, 64, 4, IRK>(q); break;
            case 2: bulk<S, 64, 16, IRK>(q); break;
            default: bulk<S, 128, 8, IRK>(q); break;
        }
    }
    const PmlTiles t = pml_tiles(q, 1024);
    pml<S, IRK, 1>(q, t);
    pml<S, IRK, 2>(q, t);
    pml<S, IRK, 4>(q, t);
    pml<S, IRK, 3>(q, t);
    pml<S, IRK, 5>(q, t);
    pml<S, IRK, 6>(q, t);
    pml<S, IRK, 7>(q, t);
    if (q.fs_on) {
        const int gx = (q.bi1 - q.bi0 + kFsTx - 1) / kFsTx, gy = (q.bj1 - q.bj0 + kFsTy - 1) / kFsTy;
        run_grid("k_stage_fs", (long long)gx * gy, kFsTx * kFsTy, [&] { k_stage_fs<S, kFsTx, kFsTy>(q, gx, gy); });
    }
    const long long n = q.boxStart[q.nbox];
    if (n > 0) run_grid("k_stage_general", (n + 255) / 256, 256, [&] { k_stage_general<S>(q); });
}

const int FB_TABLE[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                            {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};

// Write coverage (--coverage): sentinel-filled output, checked after each stage.
bool g_cover = false, g_plant_cover = false;
long long g_cover_bad = 0;

void fill_sentinel(void* p, size_t n) { std::memset(p, 0xff, n); }

// Elements of an array of `n` elements of `sb` bytes that still hold the sentinel.
inline bool is_sentinel(const char* e, int sb) {
    for (int b = 0; b < sb; ++b)
        if ((unsigned char)e[b] != 0xff) return false;
    return true;
}

// The stage output register W (9 variables): every interior element written, nothing else.
// Afterwards the whole array is zeroed, so no sentinel reaches a later stage as input.
void check_wout(const cgfdm_domain& h, void* W, const char* what) {
    const Layout& L = h.lay;
    const int sb = h.sb;
    char* b = (char*)W;
    if (g_plant_cover) {
        const long long q = L.at(9, 4, HALO + h.nx / 2, HALO + h.ny / 2, HALO + h.nz / 2);
        std::memset(b + q * sb, 0xff, sb);
    }
    long long miss = 0, written = 0;
    for (int k = HALO; k < HALO + h.nz; ++k)
        for (int j = HALO; j < HALO + h.ny; ++j)
            for (int i = HALO; i < HALO + h.nx; ++i)
                for (int v = 0; v < 9; ++v)
                    if (is_sentinel(b + L.at(9, v, i, j, k) * sb, sb)) {
                        if (miss++ < 4)
                            fprintf(stderr, "coverage: %s: interior element (v %d, i %d, j %d, k %d) not written\n", what, v,
                                    i, j, k);
                    }
    const long long n = L.size(9);
    for (long long e = 0; e < n; ++e)
        if (!is_sentinel(b + e * sb, sb)) ++written;
    const long long interior = 9LL * h.nx * h.ny * h.nz;
    if (written != interior - miss) {
        fprintf(stderr, "coverage: %s: %lld elements written, %lld interior elements\n", what, written, interior - miss);
        ++g_cover_bad;
    }
    if (miss) ++g_cover_bad;
    std::memset(b, 0, (size_t)n * sb);
}

// One PML register (9 variables at every point of every slab that is on) fully written.
void check_aux(const cgfdm_domain& h, int reg, const char* what) {
    for (int s = 0; s < 6; ++s) {
        const Slab& X = h.slab[s];
        if (!X.on) continue;
        char* b = (char*)(reg == 1 ? X.m : X.t);
        long long miss = 0;
        for (long long e = 0; e < 9 * X.n; ++e)
            if (is_sentinel(b + e * h.sb, h.sb)) ++miss;
        if (miss) {
            fprintf(stderr, "coverage: %s of slab %d: %lld of %lld elements not written\n", what, s, miss, 9 * X.n);
            ++g_cover_bad;
        }
        std::memset(b, 0, (size_t)(9 * X.n) * h.sb);
    }
}

void fill_aux(const cgfdm_domain& h, int reg) {
    for (int s = 0; s < 6; ++s) {
        const Slab& X = h.slab[s];
        if (X.on) fill_sentinel(reg == 1 ? X.m : X.t, (size_t)(9 * X.n) * h.sb);
    }
}

template <class S>
int run(cgfdm_domain& h, int cfg, int kc, bool boundary_only, int steps) {
    StageParams p = domain_stage_params(&h);
    for (int uni = 0; uni < 2; ++uni) {
        *(int*)h.m0 = uni;  // the first metric pair from its one copy, or per point
        for (int it = 0; it < steps; ++it) {
            for (int c = 0; c < 4; ++c) {
                if (cfg >= 0 && c != cfg) continue;
                p.bulk_cfg = c;
                p.bulk_kc = kc;
                int FB1 = FB_TABLE[it % 8][0], FB2 = FB_TABLE[it % 8][1], FB3 = FB_TABLE[it % 8][2];
                for (int irk = 0; irk < 4; ++irk) {
                    p.Win = stage_in(&h, irk);
                    p.Wout = stage_out(&h, irk);
                    p.irk = irk;
                    p.FB1 = FB1; p.FB2 = FB2; p.FB3 = FB3;
                    const size_t wbytes = (size_t)h.lay.size(9) * h.sb;
                    if (g_cover) {
                        fill_sentinel(p.Wout, wbytes);
                        if (irk == 0) fill_aux(h, 1);
                        if (irk == 1) fill_aux(h, 2);
                    }
                    switch (irk) {
                        case 0: stage<S, 0>(p, boundary_only); break;
                        case 1: stage<S, 1>(p, boundary_only); break;
                        case 2: stage<S, 2>(p, boundary_only); break;
                        default: stage<S, 3>(p, boundary_only); break;
                    }
                    if (g_cover) {
                        static const char* wn[4] = {"stage 0 output", "stage 1 output", "stage 2 output", "stage 3 output"};
                        check_wout(h, p.Wout, wn[irk]);
                        if (irk == 0) check_aux(h, 1, "stage 0: PML stage-0 copy");
                        if (irk == 1) check_aux(h, 2, "stage 1: PML accumulator");
                    }
                    FB1 = -FB1; FB2 = -FB2; FB3 = -FB3;
                }
                fprintf(stderr, "  step %d (FB %+d%+d%+d), tile cfg %d, uniform pair %d: %lld accesses checked\n", it,
                        FB_TABLE[it % 8][0], FB_TABLE[it % 8][1], FB_TABLE[it % 8][2], c, uni,
                        g_emu.accesses.load());
            }
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) {
        fprintf(stderr, "usage: %s NX NY NZ nPML {fp32|fp16|bf16} free_surface [tile_cfg|-1] [kc] [--boundary-only] "
                        "[--steps N]\n", argv[0]);
        return 2;
    }
    struct sigaction sa {};
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);

    cgfdm_desc d{};
    d.NX = atoi(argv[1]); d.NY = atoi(argv[2]); d.NZ = atoi(argv[3]); d.nPML = atoi(argv[4]);
    const std::string prec = argv[5];
    d.precision = prec == "fp32" ? CGFDM_FP32 : prec == "fp16" ? CGFDM_FP16 : CGFDM_BF16;
    d.free_surface = atoi(argv[6]);
    d.pml = 1;
    d.DH = 100.0;
    d.DT = 0.005;
    d.NT = 1;
    const int cfg = argc > 7 ? atoi(argv[7]) : -1;
    const int kc = argc > 8 ? atoi(argv[8]) : 0;
    bool boundary_only = false, shrink = false;
    int steps = 1;
    for (int a = 9; a < argc; ++a) {
        if (!strcmp(argv[a], "--boundary-only")) boundary_only = true;
        if (!strcmp(argv[a], "--shrink-arrays")) shrink = true;
        if (!strcmp(argv[a], "--coverage")) g_cover = g_emu_store_data = true;
        if (!strcmp(argv[a], "--plant-coverage")) g_plant_cover = true;
        if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps = atoi(argv[++a]);
    }

    cgfdm_domain h;
    h.d = d;
    h.prec = d.precision;
    h.sb = storage_bytes(h.prec);
    h.nx = d.NX; h.ny = d.NY; h.nz = d.NZ;
    h.nx_ = h.nx + 6; h.ny_ = h.ny + 6; h.nz_ = h.nz + 6;
    h.lay = make_layout(d);
    h.Crho = 1.0;
    static const char* wn[4] = {"W_A", "W_B", "W_C", "W_T"};
    const long long wbytes = (h.lay.size(9) - (shrink ? h.lay.kplane(9) : 0)) * h.sb;
    for (int b = 0; b < 4; ++b) h.W[b] = dev_alloc((size_t)wbytes, wn[b]);
    h.cjm = dev_alloc((size_t)h.lay.size(13) * h.sb, "cjm");
    h.m0 = dev_alloc(16, "m0");
    h.zcol = (float*)dev_alloc(sizeof(float) * 3 * (size_t)h.nx_ * h.ny_, "zcol");
    long long auxTotal = 0;
    for (int s = 0; s < 6; ++s) auxTotal += slab_points(d, s);
    h.auxmem = (char*)dev_alloc((size_t)auxTotal * 9 * 3 * h.sb, "aux");
    h.prof = (float*)dev_alloc(sizeof(float) * 3 * (h.nx_ + h.ny_ + h.nz_), "profiles");
    h.rdz = (float*)dev_alloc(sizeof(float) * 18 * (size_t)h.nx_ * h.ny_, "rdz");
    assign_slabs(&h);
    fprintf(stderr, "bounds_check %dx%dx%d nPML %d %s fs %d: layout px %d lead %d, arrays %.1f GiB\n", d.NX, d.NY,
            d.NZ, d.nPML, prec.c_str(), d.free_surface, h.lay.px, h.lay.lead,
            (4.0 * h.lay.size(9) + h.lay.size(13)) * h.sb / 1073741824.0);
    int rc = 0;
    if (h.prec == P_F32) rc = run<float>(h, cfg, kc, boundary_only, steps);
    else if (h.prec == P_F16) rc = run<_Float16>(h, cfg, kc, boundary_only, steps);
    else rc = run<__bf16>(h, cfg, kc, boundary_only, steps);
    const long long oob = g_emu.oob.load(), wild = g_emu.wild.load();
    printf("accesses %lld, past a descriptor %lld, outside every array %lld\n", g_emu.accesses.load(), oob, wild);
    if (g_cover) printf("coverage: %lld failed checks\n", g_cover_bad);
    return rc || oob || wild || g_cover_bad ? 1 : 0;
}
