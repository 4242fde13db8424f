// server.hip — continuous-batching engine for many independent simulations on one GPU.
//
// A server owns one HIP stream, one device arena (a single allocation sized by the caller's HBM
// budget), one pinned host arena and a batch of up to kMaxSlots resident domains. A worker
// thread runs the loop
//
//   admit:   take submitted requests (in arrival order, or the least work first) while a slot is
//            free and the arena has a block for the request's whole device footprint; set the
//            domain up in that block
//            (metric, medium, sources, stations: all enqueued, no host wait) and write its slot;
//   advance: replay a HIP graph of `steps_per_launch` whole time steps over every resident
//            domain (batch.hip), or launch the same kernels directly;
//   retire:  for domains that reached their NT, enqueue the copies of their seismograms and
//            PGV, clear their slot and return their arena block; once the copies have landed
//            (an event one launch later, so the GPU never idles on the host) the request is
//            complete and its latency stamped.
//
// Everything runs in stream order on one stream, so an arena block freed by a retired domain
// can be handed to a new one at once: the new domain's set-up is enqueued after the old
// domain's last use.
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "domain.hpp"

using namespace cg;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator of offsets in [0, size), 256-byte granules, coalescing frees.
class Arena {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t n, size_t& off) {
        n = round256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used_ += n;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = round256(n);
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t size() const { return size_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Req {
    long long id = 0;
    cgfdm_desc desc{};
    std::vector<double> terrain;
    std::vector<float> layers;  // depth | vs | vp | rho, nl each
    int nl = 0;
    std::vector<float> vs, vp, rho;
    int source = 0;
    int sxyz[3] = {0, 0, 0};
    float fc = 0.f;
    std::vector<int> tabXyz;
    std::vector<float> tabRate;
    long long tabNpts = 0;
    int tabNt = 0;
    float tabDt = 0.f;
    std::vector<int> st;
    Clock::time_point tSubmit, tAdmit, tDone;
    int status = CGFDM_OK;
    std::string err;
    std::vector<float> stations, pgv;
    size_t devBytes = 0, hostBytes = 0;
    int rescaled = 0;  // runs repeated with smaller scales after non-finite results
};

struct Resident {
    std::unique_ptr<Req> req;
    cgfdm_domain* dom = nullptr;
    int slot = -1;
    size_t devOff = 0, hostOff = 0;
    int itHost = 0;
    float* resSt = nullptr;   // pinned result regions inside the host block
    float* resPgv = nullptr;
};

// Device and pinned-host bytes one request needs for its whole life (see domain_init,
// domain_set_model, the source and station set-up); every piece is rounded to 256 bytes.
void footprint(Req& r) {
    const cgfdm_desc& d = r.desc;
    const size_t nx_ = d.NX + 6, ny_ = d.NY + 6, nz_ = d.NZ + 6;
    const size_t nin = (size_t)d.NX * d.NY * d.NZ;
    size_t dev = (size_t)domain_estimate(d) + 32 * 256;
    size_t host = round256(3 * (nx_ + ny_ + nz_) * 4);
    if (r.nl > 0) {
        dev += round256(16 * (size_t)r.nl);
        host += round256(16 * (size_t)r.nl);
    } else {
        dev += 3 * round256(4 * nin);
        host += 3 * round256(4 * nin);
    }
    if (d.use_terrain) {
        dev += round256(8 * nx_ * ny_);
        host += round256(8 * nx_ * ny_);
    }
    if (r.source == 1) {
        dev += 2 * round256(343 * 4);
        host += round256(343 * 4);
    } else {
        dev += round256(12 * (size_t)r.tabNpts) + round256(24 * (size_t)r.tabNpts * r.tabNt);
        host += round256(12 * (size_t)r.tabNpts) + round256(24 * (size_t)r.tabNpts * r.tabNt);
    }
    const size_t nst = r.st.size() / 3;
    dev += round256(12 * nst) + round256(36 * nst * d.NT);
    host += round256(12 * nst);
    host += round256(36 * nst * d.NT) + round256(8 * (size_t)d.NX * d.NY);  // results
    r.devBytes = dev;
    r.hostBytes = host + 4 * 256;
}

}  // namespace

// Least-work admission passes a waiting request over for at most this long.
constexpr std::chrono::seconds kAdmitAge{60};

struct cgfdm_server {
    cgfdm_server_config cfg{};
    int cfgTile = 1;  // bulk tile shape of the batch kernels (0: 64x8, 1: 64x4)
    hipStream_t stream = nullptr;
    char* devBase = nullptr;
    char* hostBase = nullptr;
    Arena devArena, hostArena;
    BatchDev b{};
    BatchGrid grid{};
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // side streams of the boundary kernels, forked from and joined into `stream` every stage
    hipStream_t side = nullptr;
    hipStream_t side2[2] = {nullptr, nullptr};
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    hipEvent_t evJoin2[2] = {nullptr, nullptr};
    StageFork fork() const {
        StageFork f{side, evFork, evJoin};
        for (int q = 0; q < 2; ++q) {
            f.side2[q] = side2[q];
            f.join2[q] = evJoin2[q];
        }
        return f;
    }
    Clock::time_point t0;

    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::unique_ptr<Req>> pending;
    std::map<long long, std::unique_ptr<Req>> done;
    std::vector<Resident> resident;     // worker thread only
    std::atomic<int> nResident{0};      // its size, for other threads
    std::vector<int> freeSlots;
    long long nextId = 1;
    bool stop = false;
    int broken = 0;
    std::string brokenMsg;
    cgfdm_server_stats stats{};
    std::thread worker;

    int setup_resident(Resident& R);
    int capture_graph();
    void fail_all(int code, const std::string& msg);
    void loop();
};

namespace {

int hip_fail(const char* what, hipError_t e) {
    set_err("%s: %s", what, hipGetErrorString(e));
    return CGFDM_EHIP;
}

}  // namespace

// Builds one resident domain inside its blocks and writes its slot. Only enqueues device work.
int cgfdm_server::setup_resident(Resident& R) {
    Req& r = *R.req;
    cgfdm_domain* h = new cgfdm_domain();
    R.dom = h;
    h->pooled = true;
    h->async = true;
    h->ownStream = false;
    h->stream = stream;
    h->devPool.base = devBase + R.devOff;
    h->devPool.size = r.devBytes;
    h->hostPool.base = hostBase + R.hostOff;
    h->hostPool.size = r.hostBytes;
    int rc = domain_init(h, &r.desc, cfg.device);
    if (rc) return rc;
    MediumSrc ms{};
    if (r.nl > 0) {
        ms.layers = r.layers.data();
        ms.nl = r.nl;
    } else {
        ms.vs = r.vs.data();
        ms.vp = r.vp.data();
        ms.rho = r.rho.data();
    }
    if ((rc = domain_set_model(h, r.desc.use_terrain ? r.terrain.data() : nullptr, ms))) return rc;
    if (r.source == 1) rc = cgfdm_set_source_point(h, r.sxyz[0], r.sxyz[1], r.sxyz[2], r.fc);
    else rc = cgfdm_set_source_table(h, r.tabXyz.data(), r.tabRate.data(), r.tabNpts, r.tabNt, r.tabDt);
    if (rc) return rc;
    if ((rc = cgfdm_set_stations(h, r.st.data(), (int)(r.st.size() / 3)))) return rc;
    const size_t nst = r.st.size() / 3;
    R.resSt = (float*)h->hostPool.take(36 * nst * r.desc.NT + 1);
    R.resPgv = (float*)h->hostPool.take(8 * (size_t)r.desc.NX * r.desc.NY);
    if (!R.resSt || !R.resPgv) { set_err("pinned result block exhausted"); return CGFDM_ENOMEM; }

    StageParams base = domain_stage_params(h);
    stage_boxes(base, 1);
    SlotDev sd{};
    sd.active = 1;
    sd.it = 0;
    sd.NT = r.desc.NT;
    sd.fs = r.desc.free_surface;
    if (base.bulk_on) {
        const int TX = 64, TY = cfgTile ? 4 : 8;
        const int nzb = base.bk1 - base.bk0;
        sd.gx = (base.bi1 - base.bi0 + TX - 1) / TX;
        sd.gy = (base.bj1 - base.bj0 + TY - 1) / TY;
        sd.kc = std::min(nzb, 16);
        sd.nch = (nzb + sd.kc - 1) / sd.kc;
    }
    if (base.fs_on) {
        sd.fgx = (base.bi1 - base.bi0 + kFsTx - 1) / kFsTx;
        sd.fgy = (base.bj1 - base.bj0 + kFsTy - 1) / kFsTy;
    }
    // the PML tiles of one domain are its share of a launch shared by up to 64 domains: ~96 per
    // box kind, so chunks stay long (a domain's own 1024 would make them 5 planes)
    sd.pt = pml_tiles(base, 96);
    sd.nbox = base.nbox;
    for (int x = 0; x < base.nbox; ++x)
        sd.boxChunks[x] = (base.boxStart[x + 1] - base.boxStart[x] + kGenChunk - 1) / kGenChunk;
    sd.lay = h->lay;
    for (int q = 0; q < 4; ++q) sd.W[q] = h->W[q];
    sd.DT = (float)r.desc.DT;
    sd.srcMode = h->srcMode;
    sd.SX = h->SX; sd.SY = h->SY; sd.SZ = h->SZ;
    sd.fc = h->fc;
    sd.coef = h->srcCoef;
    sd.srcIdx = h->srcIdx;
    sd.mrate = h->mrate;
    sd.npts = h->npts;
    sd.snt = h->snt;
    sd.sdt = h->sdt;
    sd.stIdx = h->stIdx;
    sd.stWave = h->stWave;
    sd.nst = h->nst;
    sd.pgv = h->pgv;
    sd.nx = h->nx;
    sd.ny = h->ny;
    sd.ktop = h->nz_ - HALO - 1;
    sd.rCv = (float)(1.0 / h->Cv);
    sd.rCs = (float)(1.0 / h->Cs);
    if (batch_write_slot(b, R.slot, sd, base, stream)) return hip_fail("batch_write_slot", hipGetLastError());
    return CGFDM_OK;
}

void cgfdm_server::fail_all(int code, const std::string& msg) {
    std::lock_guard<std::mutex> g(mu);
    broken = code;
    brokenMsg = msg;
    auto finish = [&](std::unique_ptr<Req> r) {
        r->status = code;
        r->err = msg;
        r->tDone = Clock::now();
        ++stats.failed;
        done[r->id] = std::move(r);
    };
    for (auto& R : resident) {
        if (R.req) finish(std::move(R.req));
        delete R.dom;
        R.dom = nullptr;
    }
    resident.clear();
    nResident = 0;
    while (!pending.empty()) {
        finish(std::move(pending.front()));
        pending.pop_front();
    }
    cvDone.notify_all();
}

// Captures steps_per_launch whole batch steps (every kernel over the device tables, the boundary
// kernels forked onto the side stream and joined back) into a graph replayed by every launch.
int cgfdm_server::capture_graph() {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { set_err("capture: %s", hipGetErrorString(e)); return CGFDM_EHIP; }
    int rc = 0;
    const StageFork fk = fork();
    for (int s = 0; s < cfg.steps_per_launch && rc == 0; ++s)
        rc = batch_step(cfg.precision, cfgTile, b, grid, stream, &fk);
    e = hipStreamEndCapture(stream, &g);
    if (rc || e != hipSuccess) {
        set_err("graph capture failed: %s", hipGetErrorString(e));
        if (g) (void)hipGraphDestroy(g);
        return CGFDM_EHIP;
    }
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { set_err("graph instantiate: %s", hipGetErrorString(e)); return CGFDM_EHIP; }
    return CGFDM_OK;
}

void cgfdm_server::loop() {
    if (hipSetDevice(cfg.device) != hipSuccess) { fail_all(CGFDM_EHIP, "hipSetDevice failed in the worker"); return; }
    std::vector<Resident> finishing[2];  // retired at launch n, completed when ev[n & 1] has passed
    long long n = 0;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // admit in arrival order (the earliest submitted request, which waits for room if it does
        // not fit), or the least work first: requests waiting longer than kAdmitAge go first,
        // oldest first (a bound on how long a large request can be passed over), then the
        // rest by increasing NX*NY*NZ*NT; the first of them that fits the arenas is admitted
        while (!pending.empty() && !freeSlots.empty()) {
            std::vector<size_t> order(cfg.admit_order == 1 ? pending.size() : 1);
            for (size_t q = 0; q < order.size(); ++q) order[q] = q;
            if (cfg.admit_order == 1) {
                const auto now = Clock::now();
                auto aged = [&](size_t q) { return now - pending[q]->tSubmit > kAdmitAge; };
                auto work = [&](size_t q) {
                    const cgfdm_desc& d = pending[q]->desc;
                    return (long long)d.NX * d.NY * d.NZ * (long long)d.NT;
                };
                std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
                    const bool ax = aged(x), ay = aged(y);
                    if (ax != ay) return ax;
                    return ax ? false : work(x) < work(y);  // aged ones keep arrival order
                });
            }
            size_t pick = pending.size(), doff = 0, hoff = 0;
            for (size_t q : order) {
                const Req& r = *pending[q];
                if (!devArena.alloc(r.devBytes, doff)) continue;
                if (!hostArena.alloc(r.hostBytes, hoff)) { devArena.release(doff, r.devBytes); continue; }
                pick = q;
                break;
            }
            if (pick == pending.size()) break;
            Resident R;
            R.req = std::move(pending[pick]);
            pending.erase(pending.begin() + (long)pick);
            R.devOff = doff;
            R.hostOff = hoff;
            R.slot = freeSlots.back();
            freeSlots.pop_back();
            R.req->tAdmit = Clock::now();
            lk.unlock();
            const int rc = setup_resident(R);
            lk.lock();
            if (rc) {
                // the block may still be written by enqueued set-up work; later users come after it
                devArena.release(R.devOff, R.req->devBytes);
                hostArena.release(R.hostOff, R.req->hostBytes);
                freeSlots.push_back(R.slot);
                delete R.dom;
                R.req->status = rc;
                R.req->err = g_err;
                R.req->tDone = Clock::now();
                ++stats.failed;
                done[R.req->id] = std::move(R.req);
                cvDone.notify_all();
                if (rc == CGFDM_EHIP) { lk.unlock(); fail_all(rc, g_err); return; }
                continue;
            }
            resident.push_back(std::move(R));
            nResident = (int)resident.size();
            stats.peak_resident = std::max<int>(stats.peak_resident, (int)resident.size());
        }
        stats.arena_peak = (long long)devArena.peak();
        if (resident.empty() && finishing[0].empty() && finishing[1].empty()) {
            if (stop && pending.empty()) break;
            cvWork.wait_for(lk, std::chrono::milliseconds(20));
            continue;
        }
        lk.unlock();
        hipError_t e = hipSuccess;
        const int G = cfg.steps_per_launch;
        if (!resident.empty()) {
            if (exec) {
                e = hipGraphLaunch(exec, stream);
            } else {
                for (int s = 0; s < G && e == hipSuccess; ++s)
                    if (const StageFork fk = fork(); batch_step(cfg.precision, cfgTile, b, grid, stream, &fk)) {
                        e = hipGetLastError();
                        if (e == hipSuccess) e = hipErrorUnknown;
                    }
            }
            if (e != hipSuccess) { fail_all(CGFDM_EHIP, std::string("batch launch: ") + hipGetErrorString(e)); return; }
        }
        const bool launched = !resident.empty();
        // retire the domains that reach their end in this launch
        std::vector<Resident>& fin = finishing[n & 1];
        for (size_t q = 0; q < resident.size();) {
            Resident& R = resident[q];
            R.itHost = std::min(R.req->desc.NT, R.itHost + G);
            if (R.itHost < R.req->desc.NT) { ++q; continue; }
            cgfdm_domain* h = R.dom;
            const size_t nst = (size_t)h->nst;
            if (nst) e = hipMemcpyAsync(R.resSt, h->stWave, 36 * nst * h->d.NT, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(R.resPgv, h->pgv, 8 * (size_t)h->nx * h->ny, hipMemcpyDeviceToHost, stream);
            if (e != hipSuccess) { fail_all(CGFDM_EHIP, std::string("result copy: ") + hipGetErrorString(e)); return; }
            if (batch_clear_slot(b, R.slot, stream)) { fail_all(CGFDM_EHIP, "batch_clear_slot failed"); return; }
            fin.push_back(std::move(R));
            resident.erase(resident.begin() + q);
            nResident = (int)resident.size();
        }
        if ((e = hipEventRecord(ev[n & 1], stream)) != hipSuccess) { fail_all(CGFDM_EHIP, "hipEventRecord failed"); return; }
        // complete what was retired one launch earlier (its copies precede that event)
        std::vector<Resident>& prev = finishing[(n + 1) & 1];
        if (n > 0 && (e = hipEventSynchronize(ev[(n + 1) & 1])) != hipSuccess) {
            fail_all(CGFDM_EHIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
            return;
        }
        lk.lock();
        if (launched) ++stats.launches;
        for (Resident& R : prev) {
            Req& r = *R.req;
            const size_t nst = r.st.size() / 3;
            r.stations.assign(R.resSt, R.resSt + nst * 9 * (size_t)r.desc.NT);
            r.pgv.assign(R.resPgv, R.resPgv + 2 * (size_t)r.desc.NX * r.desc.NY);
            // non-finite guard of served requests: decided on the copied results
            bool finite = true;
            for (float x : r.stations) finite = finite && std::isfinite(x);
            for (float x : r.pgv) finite = finite && std::isfinite(x);
            devArena.release(R.devOff, r.devBytes);
            hostArena.release(R.hostOff, r.hostBytes);
            freeSlots.push_back(R.slot);
            delete R.dom;
            if (!finite && r.desc.precision != CGFDM_FP32 && r.rescaled < cfg.rescale_retries) {
                // 16-bit storage out of range: run the request again from its start with scales
                // 8x smaller (Cv/Cs unchanged, so the stored medium is the same), ahead of the
                // waiting requests
                const double Cv = r.desc.Cv > 0 ? r.desc.Cv : kRefCv, Cs = r.desc.Cs > 0 ? r.desc.Cs : kRefCs;
                r.desc.Cv = Cv / 8.0;
                r.desc.Cs = Cs / 8.0;
                ++r.rescaled;
                ++stats.rescaled;
                r.stations.clear();
                r.pgv.clear();
                pending.push_front(std::move(R.req));
                continue;
            }
            if (!finite) {
                r.status = CGFDM_ENONFINITE;
                r.err = "non-finite seismogram or PGV value (16-bit storage out of range?)";
            }
            r.tDone = Clock::now();
            ++stats.completed;
            done[r.id] = std::move(R.req);
        }
        if (!prev.empty()) cvDone.notify_all();
        prev.clear();
        ++n;
    }
}

extern "C" {

int cgfdm_server_create(const cgfdm_server_config* c, cgfdm_server** out) {
    if (!c || !out) { set_err("null argument"); return CGFDM_EINVAL; }
    *out = nullptr;
    if (c->precision < CGFDM_FP32 || c->precision > CGFDM_BF16) { set_err("bad precision"); return CGFDM_EINVAL; }
    if (c->max_domains < 1 || c->max_domains > kMaxSlots) {
        set_err("max_domains must be in [1, %d]", kMaxSlots);
        return CGFDM_EINVAL;
    }
    if (c->steps_per_launch < 1 || c->steps_per_launch > 64) { set_err("steps_per_launch must be in [1, 64]"); return CGFDM_EINVAL; }
    if (c->admit_order != 0 && c->admit_order != 1) { set_err("admit_order must be 0 or 1"); return CGFDM_EINVAL; }
    if (c->rescale_retries < 0 || c->rescale_retries > 8) { set_err("rescale_retries must be in [0, 8]"); return CGFDM_EINVAL; }
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail("hipSetDevice", e);
    auto S = std::make_unique<cgfdm_server>();
    S->cfg = *c;
    if (const char* v = std::getenv("CGFDM_BULK_CFG")) S->cfgTile = std::atoi(v) ? 1 : 0;
    size_t freeB = 0, totalB = 0;
    if ((e = hipMemGetInfo(&freeB, &totalB)) != hipSuccess) return hip_fail("hipMemGetInfo", e);
    size_t budget = c->hbm_budget > 0 ? (size_t)c->hbm_budget : (size_t)(0.8 * (double)freeB);
    if (budget > freeB) { set_err("HBM budget %zu exceeds free memory %zu", budget, freeB); return CGFDM_ENOMEM; }
    size_t hostBudget = c->host_budget > 0 ? (size_t)c->host_budget : ((size_t)1 << 30);
    if ((e = hipStreamCreate(&S->stream)) != hipSuccess) return hip_fail("hipStreamCreate", e);
    auto undo = [&](int rc) {
        if (S->devBase) (void)hipFree(S->devBase);
        if (S->hostBase) (void)hipHostFree(S->hostBase);
        if (S->b.slots) (void)hipFree(S->b.slots);
        if (S->b.sp) (void)hipFree(S->b.sp);
        if (S->b.tab) (void)hipFree(S->b.tab);
        for (auto& v : S->ev) if (v) (void)hipEventDestroy(v);
        if (S->evFork) (void)hipEventDestroy(S->evFork);
        if (S->evJoin) (void)hipEventDestroy(S->evJoin);
        if (S->side) (void)hipStreamDestroy(S->side);
        for (int q = 0; q < 2; ++q) {
            if (S->evJoin2[q]) (void)hipEventDestroy(S->evJoin2[q]);
            if (S->side2[q]) (void)hipStreamDestroy(S->side2[q]);
        }
        (void)hipStreamDestroy(S->stream);
        return rc;
    };
    if ((e = hipStreamCreate(&S->side)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&S->evFork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&S->evJoin, hipEventDisableTiming)) != hipSuccess) {
        set_err("side stream: %s", hipGetErrorString(e));
        return undo(CGFDM_EHIP);
    }
    for (int q = 0; q < 2; ++q)
        if ((e = hipStreamCreate(&S->side2[q])) != hipSuccess ||
            (e = hipEventCreateWithFlags(&S->evJoin2[q], hipEventDisableTiming)) != hipSuccess) {
            set_err("side stream: %s", hipGetErrorString(e));
            return undo(CGFDM_EHIP);
        }
    if ((e = hipMalloc((void**)&S->devBase, budget)) != hipSuccess) { (void)hipGetLastError(); set_err("arena of %zu bytes: %s", budget, hipGetErrorString(e)); return undo(CGFDM_ENOMEM); }
    if ((e = hipHostMalloc((void**)&S->hostBase, hostBudget, 0)) != hipSuccess) { set_err("pinned arena: %s", hipGetErrorString(e)); return undo(CGFDM_ENOMEM); }
    S->devArena.reset(budget);
    S->hostArena.reset(hostBudget);
    S->b.maxSlots = c->max_domains;
    if ((e = hipMalloc((void**)&S->b.slots, sizeof(SlotDev) * kMaxSlots)) != hipSuccess ||
        (e = hipMalloc((void**)&S->b.sp, sizeof(StageParams) * 4 * kMaxSlots)) != hipSuccess ||
        (e = hipMalloc((void**)&S->b.tab, sizeof(BatchTables))) != hipSuccess) {
        set_err("batch tables: %s", hipGetErrorString(e));
        return undo(CGFDM_ENOMEM);
    }
    if ((e = hipMemsetAsync(S->b.slots, 0, sizeof(SlotDev) * kMaxSlots, S->stream)) != hipSuccess ||
        (e = hipMemsetAsync(S->b.tab, 0, sizeof(BatchTables), S->stream)) != hipSuccess) {
        set_err("batch tables: %s", hipGetErrorString(e));
        return undo(CGFDM_EHIP);
    }
    for (auto& v : S->ev)
        if ((e = hipEventCreateWithFlags(&v, hipEventDisableTiming)) != hipSuccess) { set_err("events"); return undo(CGFDM_EHIP); }
    if (batch_grid(c->precision, S->cfgTile, S->grid)) { set_err("occupancy query failed"); return undo(CGFDM_EHIP); }
    if ((e = hipStreamSynchronize(S->stream)) != hipSuccess) { set_err("init sync: %s", hipGetErrorString(e)); return undo(CGFDM_EHIP); }
    for (int s = c->max_domains - 1; s >= 0; --s) S->freeSlots.push_back(s);
    S->t0 = Clock::now();
    // The worker thread captures the step graph (use_graph) before it serves anything, so the graph
    // is captured, instantiated and replayed on the one thread that launches work on the stream;
    // creation waits for that and reports its failure.
    auto ready = std::make_shared<std::promise<std::pair<int, std::string>>>();
    auto started = ready->get_future();
    cgfdm_server* raw = S.get();
    S->worker = std::thread([raw, ready]() {
        int rc = 0;
        if (hipSetDevice(raw->cfg.device) != hipSuccess) {
            set_err("hipSetDevice failed in the worker");
            rc = CGFDM_EHIP;
        }
        if (rc == 0 && raw->cfg.use_graph) rc = raw->capture_graph();
        ready->set_value({rc, rc ? g_err : std::string()});
        if (rc == 0) raw->loop();
    });
    const auto res = started.get();
    if (res.first) {
        S->worker.join();
        if (S->exec) { (void)hipGraphExecDestroy(S->exec); S->exec = nullptr; }
        set_err("%s", res.second.c_str());
        return undo(res.first);
    }
    *out = S.release();
    return CGFDM_OK;
}

int cgfdm_server_submit(cgfdm_server* s, const cgfdm_request* q, long long* id) {
    if (!s || !q || !id) { set_err("null argument"); return CGFDM_EINVAL; }
    int rc = validate_desc(&q->desc);
    if (rc) return rc;
    if (q->desc.precision != s->cfg.precision) { set_err("request precision differs from the server's"); return CGFDM_EINVAL; }
    if (q->desc.check_cfl) { set_err("served requests cannot use the CFL check (it needs a host wait)"); return CGFDM_EINVAL; }
    if (q->desc.NT < 1) { set_err("NT must be >= 1"); return CGFDM_EINVAL; }
    auto r = std::make_unique<Req>();
    r->desc = q->desc;
    const cgfdm_desc& d = q->desc;
    const size_t nin = (size_t)d.NX * d.NY * d.NZ;
    if (d.use_terrain) {
        if (!q->terrain) { set_err("use_terrain without terrain"); return CGFDM_EINVAL; }
        r->terrain.assign(q->terrain, q->terrain + (size_t)(d.NX + 6) * (d.NY + 6));
    }
    if (q->nlayer > 0) {
        if (!q->layer_depth || !q->layer_vs || !q->layer_vp || !q->layer_rho) { set_err("layer arrays missing"); return CGFDM_EINVAL; }
        r->nl = q->nlayer;
        r->layers.insert(r->layers.end(), q->layer_depth, q->layer_depth + q->nlayer);
        r->layers.insert(r->layers.end(), q->layer_vs, q->layer_vs + q->nlayer);
        r->layers.insert(r->layers.end(), q->layer_vp, q->layer_vp + q->nlayer);
        r->layers.insert(r->layers.end(), q->layer_rho, q->layer_rho + q->nlayer);
    } else {
        if (!q->vs || !q->vp || !q->rho) { set_err("no medium given"); return CGFDM_EINVAL; }
        r->vs.assign(q->vs, q->vs + nin);
        r->vp.assign(q->vp, q->vp + nin);
        r->rho.assign(q->rho, q->rho + nin);
    }
    r->source = q->source;
    if (q->source == 1) {
        std::memcpy(r->sxyz, q->src_xyz, sizeof(r->sxyz));
        r->fc = q->fc;
    } else if (q->source == 2) {
        if (!q->tab_xyz || !q->tab_rate || q->tab_npts < 1 || q->tab_nt < 2) { set_err("bad source table"); return CGFDM_EINVAL; }
        r->tabXyz.assign(q->tab_xyz, q->tab_xyz + 3 * q->tab_npts);
        r->tabRate.assign(q->tab_rate, q->tab_rate + 6 * (size_t)q->tab_npts * q->tab_nt);
        r->tabNpts = q->tab_npts;
        r->tabNt = q->tab_nt;
        r->tabDt = q->tab_dt;
    } else {
        set_err("source must be 1 (point) or 2 (table)");
        return CGFDM_EINVAL;
    }
    if (d.precision != CGFDM_FP32 && !(d.Cv > 0) && !(d.Cs > 0)) {
        // per-request scales from the source (cgfdm_source_scales; point source:
        // cgfdm_point_source_scales with the smallest shear modulus of the medium), with the
        // largest shear modulus of the request's medium
        double mu = 0.0, muMin = 1e300;
        if (r->nl > 0) {
            for (int l = 0; l < r->nl; ++l) {
                const double m = (double)r->layers[3 * r->nl + l] * r->layers[r->nl + l] * r->layers[r->nl + l];
                mu = std::max(mu, m);
                muMin = std::min(muMin, m);
            }
            mu = std::max(mu, 2670.0 * 3464.0 * 3464.0);  // points outside every layer keep the default medium
        } else {
            for (size_t t = 0; t < nin; ++t) {
                const double m = (double)r->rho[t] * r->vs[t] * r->vs[t];
                mu = std::max(mu, m);
                muMin = std::min(muMin, m);
            }
        }
        double Cv = 0.0, Cs = 0.0;
        if (q->source == 2)
            source_scales(d.DH, r->tabXyz.data(), r->tabRate.data(), r->tabNpts, r->tabNt, r->tabDt, mu, Cv, Cs);
        else if (muMin > 0 && muMin < 1e300)
            point_source_scales(d.DH, r->fc, muMin, mu, Cv, Cs);
        if (Cv > 0 && std::isfinite(Cv)) {
            r->desc.Cv = Cv;
            r->desc.Cs = Cs;
        }
    }
    if (q->nst < 0 || (q->nst > 0 && !q->st_xyz)) { set_err("bad station list"); return CGFDM_EINVAL; }
    r->st.assign(q->st_xyz, q->st_xyz + 3 * (size_t)q->nst);
    footprint(*r);
    std::lock_guard<std::mutex> g(s->mu);
    if (s->broken) { set_err("server stopped after an error: %s", s->brokenMsg.c_str()); return s->broken; }
    if (s->stop) { set_err("server is shutting down"); return CGFDM_ESTATE; }
    if (r->devBytes > s->devArena.size() || r->hostBytes > s->hostArena.size()) {
        set_err("request needs %zu device / %zu host bytes, more than the server's arenas", r->devBytes, r->hostBytes);
        return CGFDM_ENOMEM;
    }
    r->id = s->nextId++;
    r->tSubmit = Clock::now();
    *id = r->id;
    ++s->stats.submitted;
    s->pending.push_back(std::move(r));
    s->cvWork.notify_all();
    return CGFDM_OK;
}

int cgfdm_server_wait(cgfdm_server* s, long long* ids, int max, int timeout_ms) {
    if (!s || (max > 0 && !ids)) { set_err("null argument"); return CGFDM_EINVAL; }
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&]() { return !s->done.empty(); };
    if (timeout_ms < 0) s->cvDone.wait(lk, ready);
    else s->cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    int n = 0;
    for (auto& kv : s->done) {
        if (n >= max) break;
        ids[n++] = kv.first;
    }
    return n;
}

int cgfdm_server_result(cgfdm_server* s, long long id, float* stations, float* pgv, cgfdm_request_info* info) {
    if (!s) { set_err("null server"); return CGFDM_EINVAL; }
    std::unique_ptr<Req> r;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto f = s->done.find(id);
        if (f == s->done.end()) { set_err("request %lld is not complete (or already fetched)", id); return CGFDM_ESTATE; }
        r = std::move(f->second);
        s->done.erase(f);
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = r->status;
        info->NT = r->desc.NT;
        info->nst = (int)(r->st.size() / 3);
        info->NX = r->desc.NX;
        info->NY = r->desc.NY;
        info->t_submit = secs(s->t0, r->tSubmit);
        info->t_admit = r->tAdmit.time_since_epoch().count() ? secs(s->t0, r->tAdmit) : -1.0;
        info->t_done = secs(s->t0, r->tDone);
        info->device_bytes = (long long)r->devBytes;
        std::snprintf(info->error, sizeof(info->error), "%s", r->err.c_str());
        const bool low = r->desc.precision != CGFDM_FP32;
        info->Cv = r->desc.Cv > 0 ? r->desc.Cv : (low ? kRefCv : 1.0);
        info->Cs = r->desc.Cs > 0 ? r->desc.Cs : (low ? kRefCs : 1.0);
        info->rescaled = r->rescaled;
    }
    if (r->status == CGFDM_OK || r->status == CGFDM_ENONFINITE) {
        if (stations && !r->stations.empty()) std::memcpy(stations, r->stations.data(), r->stations.size() * 4);
        if (pgv) std::memcpy(pgv, r->pgv.data(), r->pgv.size() * 4);
    }
    return r->status;
}

int cgfdm_server_stats_get(cgfdm_server* s, cgfdm_server_stats* st) {
    if (!s || !st) { set_err("null argument"); return CGFDM_EINVAL; }
    std::lock_guard<std::mutex> g(s->mu);
    *st = s->stats;
    st->resident = s->nResident.load();
    st->pending = (int)s->pending.size();
    st->arena_bytes = (long long)s->devArena.size();
    st->arena_used = (long long)s->devArena.used();
    st->arena_peak = (long long)s->devArena.peak();
    st->graph = s->exec ? 1 : 0;
    st->tile_cfg = s->cfgTile;
    st->bulk_blocks = s->grid.bulkBlocks;
    st->general_blocks = s->grid.genBlocks;
    return CGFDM_OK;
}

int cgfdm_server_destroy(cgfdm_server* s) {
    if (!s) return CGFDM_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cvWork.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->stream);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    for (auto& R : s->resident) delete R.dom;
    (void)hipFree(s->devBase);
    (void)hipHostFree(s->hostBase);
    (void)hipFree(s->b.slots);
    (void)hipFree(s->b.sp);
    (void)hipFree(s->b.tab);
    for (auto& v : s->ev) (void)hipEventDestroy(v);
    (void)hipStreamSynchronize(s->side);
    for (int q = 0; q < 2; ++q) (void)hipStreamSynchronize(s->side2[q]);
    (void)hipEventDestroy(s->evFork);
    (void)hipEventDestroy(s->evJoin);
    (void)hipStreamDestroy(s->side);
    for (int q = 0; q < 2; ++q) {
        (void)hipEventDestroy(s->evJoin2[q]);
        (void)hipStreamDestroy(s->side2[q]);
    }
    (void)hipStreamDestroy(s->stream);
    delete s;
    return CGFDM_OK;
}

}  // extern "C"
