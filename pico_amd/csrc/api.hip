// api.hip -- the C ABI declared in include/pico_gpu.h: contexts, device
// matrices, HBM budget accounting, twiddle / coset-scale caches, and the host
// orchestration of the MMCS, PCS and FRI prover steps.
//
// Host-side order of operations follows the reference's prover:
//   Mmcs::commit       p3 MerkleTreeMmcs (stable sort by height desc, leaf
//                      layer over the tallest matrices, compress + inject;
//                      pinned by vm/src/compiler/recursion/circuit/fri.rs:389-436)
//   Pcs::commit        LDE shift = GENERATOR / domain.shift, bit-reversed rows
//                      (vm/src/machine/prover.rs:238-244, circuit/fri.rs:151-171)
//   FRI commit phase   commit -> observe root -> sample beta -> fold -> add the
//                      reduced opening of the new length (circuit/fri.rs:36-73,
//                      :288-387); PoW check_witness (circuit/challenger.rs:210-222)
#include <hip/hip_runtime.h>
#ifndef __HIP_DEVICE_COMPILE__
#include <immintrin.h>
#endif

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <deque>
#include <thread>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pico_gpu.h"
#include "kernels.h"
#include "poseidon2.h"

namespace pg {
hipError_t launch_leaf_hash_w8(int field, const uint32_t* data, uint64_t stride, uint32_t width,
                               size_t rows, uint32_t* digests, hipStream_t stream);
}

using namespace pg;

// ---------------------------------------------------------------- errors --
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PG_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define PG_TRY(expr)          \
    do {                      \
        int r_ = (expr);      \
        if (r_ != PG_OK) return r_; \
    } while (0)

extern "C" const char* pg_last_error(void) { return g_err.c_str(); }
extern "C" int pg_version(void) { return 1; }

// ---------------------------------------------------------------- context --
struct DevMem {  // device allocation accounted against the context budget
    pg_ctx* ctx = nullptr;
    void* ptr = nullptr;
    size_t bytes = 0;
};

// column blocks of pg_pcs_commit_host (see there)
static constexpr size_t HOST_BLOCK_BYTES = 256ull << 20, HOST_BLOCK_MIN_COLS = 64;

struct pg_ctx {
    int device = 0;
    size_t budget = 0;
    std::atomic<size_t> in_use{0};
    std::mutex mu;
    std::vector<hipStream_t> streams;
    size_t next_stream = 0;
    // caches: twiddles per (field, log n); coset scale per (field, log h, b, shift)
    std::map<std::pair<int, int>, std::shared_ptr<DevMem>> twiddles;
    std::map<std::tuple<int, int, int, uint32_t>, std::shared_ptr<DevMem>> scales;
    std::vector<std::tuple<int, int, int, uint32_t>> scale_order;  // insertion order for eviction
    size_t scale_bytes = 0;
    // device memory pool: segments from hipMalloc, carved into blocks
    struct PoolSeg {
        char* base;
        size_t size;
    };
    struct FreeBlk {
        size_t size;
        int seg;
    };
    std::mutex pool_mu;
    std::vector<PoolSeg> segs;                       // base == nullptr once released
    std::map<char*, FreeBlk> free_by_addr;           // free blocks, address order
    std::multimap<size_t, char*> free_by_size;       // the same blocks, size order
    std::map<char*, std::pair<size_t, int>> live;    // block -> (size, segment)
    size_t pool_bytes = 0;                           // sum of live segment sizes
    // per-kernel-family event timings (profiling mode only)
    std::atomic<bool> profiling{false};
    std::map<std::string, std::pair<double, uint64_t>> prof;
    // fault injection: the allocation this many allocations from now fails
    // once with PG_E_OOM; negative = off
    std::atomic<long long> fail_countdown{-1};
    // column-block sizing of pg_pcs_commit_host (pg_debug_host_blocks)
    std::atomic<size_t> host_block_bytes{HOST_BLOCK_BYTES}, host_block_min_cols{HOST_BLOCK_MIN_COLS};
    // page-locked double buffers for uploads from pageable memory (struct
    // Stage), created on first use and freed by pg_destroy
    std::mutex stage_mu;
    std::vector<void*> stages;
};

// Times each launch with a pair of HIP events recorded on the launch stream
// and waits for the second one, so per-kernel durations are exact (at the
// cost of one host sync per launch; enabled only by pg_profile_enable).
struct EventTimer : KernelTimer {
    pg_ctx* ctx;
    hipEvent_t a = nullptr, b = nullptr;
    std::string cur;
    explicit EventTimer(pg_ctx* c) : ctx(c) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~EventTimer() override {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void begin(const char* name, hipStream_t s) override {
        cur = name;
        (void)hipEventRecord(a, s);
    }
    void end(hipStream_t s) override {
        (void)hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) return;
        float ms = 0;
        if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return;
        std::lock_guard<std::mutex> g(ctx->mu);
        auto& e = ctx->prof[cur];
        e.first += ms;
        e.second += 1;
    }
};

static std::unique_ptr<EventTimer> make_timer(pg_ctx* ctx) {
    return ctx->profiling.load() ? std::unique_ptr<EventTimer>(new EventTimer(ctx)) : nullptr;
}

// time one launch expression when profiling
#define TIMED(kt, name, s, expr)            \
    do {                                    \
        if (kt) (kt)->begin(name, s);       \
        hipError_t te_ = (expr);            \
        if (kt) (kt)->end(s);               \
        HIP_TRY(te_);                       \
    } while (0)

// Device memory: live bytes are charged against the budget.  Blocks are
// carved from large hipMalloc segments (at least POOL_SEG_MIN) by best fit;
// a freed block is merged with free neighbours of the same segment and
// reused for any later request it can hold.  hipMalloc / hipFree of
// multi-GiB buffers costs tens of ms per GiB on this device, so a stream of
// commits of varying shapes would otherwise spend most of its time there.
// Wholly free segments go back to the driver only when a new segment does
// not fit, or when the context is destroyed.  Callers free a block only
// after the work using it has completed.
static constexpr size_t POOL_ALIGN = 512;
static constexpr size_t POOL_SEG_MIN = 1ull << 30;
static constexpr size_t POOL_SEG_ROUND = 2ull << 20;

static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static void pool_erase_size(pg_ctx* c, size_t n, char* p) {  // pool_mu held
    auto r = c->free_by_size.equal_range(n);
    for (auto it = r.first; it != r.second; ++it)
        if (it->second == p) {
            c->free_by_size.erase(it);
            return;
        }
}

static void pool_put_free(pg_ctx* c, char* p, size_t n, int seg) {  // pool_mu held
    auto next = c->free_by_addr.find(p + n);
    if (next != c->free_by_addr.end() && next->second.seg == seg) {
        n += next->second.size;
        pool_erase_size(c, next->second.size, next->first);
        c->free_by_addr.erase(next);
    }
    auto after = c->free_by_addr.upper_bound(p);
    if (after != c->free_by_addr.begin()) {
        auto prev = std::prev(after);
        if (prev->second.seg == seg && prev->first + prev->second.size == p) {
            pool_erase_size(c, prev->second.size, prev->first);
            p = prev->first;
            n += prev->second.size;
            c->free_by_addr.erase(prev);
        }
    }
    c->free_by_addr.emplace(p, pg_ctx::FreeBlk{n, seg});
    c->free_by_size.emplace(n, p);
}

static char* pool_take(pg_ctx* c, size_t n) {  // pool_mu held; n aligned
    auto it = c->free_by_size.lower_bound(n);
    if (it == c->free_by_size.end()) return nullptr;
    char* p = it->second;
    const size_t have = it->first;
    c->free_by_size.erase(it);
    auto fa = c->free_by_addr.find(p);
    const int seg = fa->second.seg;
    c->free_by_addr.erase(fa);
    if (have > n) {
        c->free_by_addr.emplace(p + n, pg_ctx::FreeBlk{have - n, seg});
        c->free_by_size.emplace(have - n, p + n);
    }
    c->live.emplace(p, std::make_pair(n, seg));
    return p;
}

static void pool_release_free_segments(pg_ctx* c) {  // pool_mu held
    for (size_t k = 0; k < c->segs.size(); k++) {
        auto& sg = c->segs[k];
        if (!sg.base) continue;
        auto it = c->free_by_addr.find(sg.base);
        if (it == c->free_by_addr.end() || it->second.seg != (int)k || it->second.size != sg.size) continue;
        pool_erase_size(c, sg.size, sg.base);
        c->free_by_addr.erase(it);
        (void)hipFree(sg.base);
        c->pool_bytes -= sg.size;
        sg.base = nullptr;
        sg.size = 0;
    }
}

static hipError_t pool_grow(pg_ctx* c, size_t n) {  // pool_mu held
    const size_t need = round_up(n, POOL_SEG_ROUND);
    size_t sz = std::max(need, c->budget ? std::min(POOL_SEG_MIN, round_up(c->budget, POOL_SEG_ROUND)) : POOL_SEG_MIN);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess && need < sz) {  // the minimum segment may not fit where n does
        (void)hipGetLastError();
        sz = need;
        e = hipMalloc(&p, sz);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    c->segs.push_back({(char*)p, sz});
    c->pool_bytes += sz;
    pool_put_free(c, (char*)p, sz, (int)c->segs.size() - 1);
    return hipSuccess;
}

static int ctx_alloc(pg_ctx* ctx, size_t bytes, void** out) {
    *out = nullptr;
    if (ctx->fail_countdown.load() >= 0 && ctx->fail_countdown.fetch_sub(1) == 0)
        return fail(PG_E_OOM, "injected allocation failure (%zu bytes)", bytes);
    const size_t n = round_up(bytes ? bytes : 1, POOL_ALIGN);
    size_t prev = ctx->in_use.fetch_add(n);
    if (ctx->budget && prev + n > ctx->budget) {
        ctx->in_use.fetch_sub(n);
        return fail(PG_E_ADMISSION, "allocation of %zu bytes exceeds the HBM budget (%zu of %zu in use)",
                    bytes, prev, ctx->budget);
    }
    std::lock_guard<std::mutex> g(ctx->pool_mu);
    char* p = pool_take(ctx, n);
    if (!p) {
        hipError_t e = pool_grow(ctx, n);
        if (e != hipSuccess) {  // hand wholly free segments back and retry once
            pool_release_free_segments(ctx);
            e = pool_grow(ctx, n);
        }
        if (e != hipSuccess) {
            ctx->in_use.fetch_sub(n);
            return fail(PG_E_OOM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        }
        p = pool_take(ctx, n);
    }
    *out = p;
    return PG_OK;
}

static void ctx_free(pg_ctx* ctx, void* ptr, size_t) {
    if (!ptr) return;
    std::lock_guard<std::mutex> g(ctx->pool_mu);
    auto it = ctx->live.find((char*)ptr);
    if (it == ctx->live.end()) return;
    const size_t n = it->second.first;
    const int seg = it->second.second;
    ctx->live.erase(it);
    ctx->in_use.fetch_sub(n);
    pool_put_free(ctx, (char*)ptr, n, seg);
}

static std::shared_ptr<DevMem> make_devmem(pg_ctx* ctx, size_t bytes, int* rc) {
    void* p = nullptr;
    *rc = ctx_alloc(ctx, bytes, &p);
    if (*rc != PG_OK) return nullptr;
    auto d = std::shared_ptr<DevMem>(new DevMem{ctx, p, bytes}, [](DevMem* m) {
        ctx_free(m->ctx, m->ptr, m->bytes);
        delete m;
    });
    return d;
}

// Each API call leases one stream and synchronises it before returning.
struct Lease {
    pg_ctx* ctx;
    hipStream_t s;
    explicit Lease(pg_ctx* c) : ctx(c) {
        std::lock_guard<std::mutex> g(c->mu);
        s = c->streams[c->next_stream++ % c->streams.size()];
    }
};

// Synchronises a stream when it goes out of scope unless disarmed: declared
// after the owners of device memory that queued kernels use, it runs first
// on an early error return, so nothing is freed while still in use.
struct SyncGuard {
    hipStream_t s;
    bool armed = true;
    explicit SyncGuard(hipStream_t st) : s(st) {}
    ~SyncGuard() {
        if (armed) (void)hipStreamSynchronize(s);
    }
};

static int bind_device(pg_ctx* ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return PG_OK;
}

// ------------------------------------------------ allocation-free scratch --
// pg_pcs_open_values_noalloc works in this array instead of device memory it
// allocates.  Every GPU the library runs on holds its own copy (placed when
// the code object is loaded on that device) until the process ends.  In host
// code the name is only a stand-in: the copy's device address on each GPU is
// looked up once per device by pg_create, with that device current, and only
// that address is passed to kernels and copies.
static constexpr size_t NOALLOC_SCRATCH_BYTES = PG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t noalloc_scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoallocSlot {
    std::mutex mu;  // held for a whole call, so calls on one device run one at a time
    uint8_t* ptr = nullptr;
};
static NoallocSlot g_noalloc[NOALLOC_MAX_DEVICES];

// `device` must be current
static int noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(PG_E_UNSUPPORTED, "device %d: the allocation-free scratch covers devices 0..%d", device,
                    NOALLOC_MAX_DEVICES - 1);
    NoallocSlot& sl = g_noalloc[device];
    std::lock_guard<std::mutex> g(sl.mu);
    if (sl.ptr) return PG_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    if (!p) return fail(PG_E_DEVICE, "hipGetSymbolAddress returned a null address for the scratch array");
    sl.ptr = (uint8_t*)p;
    return PG_OK;
}

// A caller-independent device region that a call works in instead of
// allocating (null base: allocate as usual).
struct FixedScratch {
    uint8_t* base = nullptr;
    size_t bytes = 0;
};

// --------------------------------------------------------------- staging --
// Copies between device memory and pageable host memory, and gathers of
// column blocks.  The runtime stages pageable copies itself at ~26-30 GB/s
// on MI355X hosts; copying 16 MiB chunks with 8 host threads through a
// page-locked double buffer, each chunk's DMA overlapping the copy of the
// next, reaches ~54 GB/s up and ~51 GB/s down of the link's 57.6.
class HostPool {  // persistent threads running the pieces of one copy at a time
  public:
    explicit HostPool(int nt) {
        for (int i = 0; i < nt; i++) workers_.emplace_back([this] { loop(); });
    }
    int size() const { return (int)workers_.size(); }
    // f(0) .. f(n-1) on the workers; returns when all are done
    void run(int n, const std::function<void(int)>& f) {
        std::lock_guard<std::mutex> one(run_mu_);
        std::unique_lock<std::mutex> lk(mu_);
        fn_ = &f;
        n_ = n;
        next_ = 0;
        left_ = n;
        cv_.notify_all();
        done_.wait(lk, [&] { return left_ == 0; });
        fn_ = nullptr;
        n_ = next_ = 0;
    }

  private:
    void loop() {
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_.wait(lk, [&] { return next_ < n_; });
            const int i = next_++;
            const std::function<void(int)>* f = fn_;
            lk.unlock();
            (*f)(i);
            lk.lock();
            if (--left_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int n_ = 0, next_ = 0, left_ = 0;
};

static HostPool& host_pool() {  // lives until the process ends
    static HostPool* p = new HostPool((int)std::max(1u, std::min(8u, std::thread::hardware_concurrency())));
    return *p;
}

struct Stage {
    static constexpr size_t CHUNK = 16u << 20;
    // uploads rotate through NB buffers, downloads use two; 3 or 4 measured
    // no faster than 2 (2^20 x 256 from host memory: 22.7-24.8 ms for each)
    static constexpr int NB = 2;
    static_assert(NB >= 2, "double buffering at least");
    void* buf[NB] = {};
    hipEvent_t ev[NB] = {};
};
static constexpr size_t STAGE_MIN_BYTES = 4u << 20;  // smaller uploads go direct

static void stage_destroy(Stage* st) {
    for (int b = 0; b < Stage::NB; b++) {
        if (st->buf[b]) (void)hipHostFree(st->buf[b]);
        if (st->ev[b]) (void)hipEventDestroy(st->ev[b]);
    }
    delete st;
}

static int stage_get(pg_ctx* ctx, Stage** out) {
    {
        std::lock_guard<std::mutex> g(ctx->stage_mu);
        if (!ctx->stages.empty()) {
            *out = (Stage*)ctx->stages.back();
            ctx->stages.pop_back();
            return PG_OK;
        }
    }
    Stage* st = new Stage;
    for (int b = 0; b < Stage::NB; b++) {
        hipError_t e = hipHostMalloc(&st->buf[b], Stage::CHUNK, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&st->ev[b], hipEventDisableTiming);
        if (e != hipSuccess) {
            stage_destroy(st);
            return fail(PG_E_OOM, "staging buffer: %s", hipGetErrorString(e));
        }
    }
    *out = st;
    return PG_OK;
}

static void stage_put(pg_ctx* ctx, Stage* st) {
    std::lock_guard<std::mutex> g(ctx->stage_mu);
    ctx->stages.push_back(st);
}

// Whether p is page-locked memory the DMA engines can read directly.
static bool host_is_locked(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) == hipSuccess) return a.type == hipMemoryTypeHost;
    (void)hipGetLastError();  // pageable memory is unknown to the runtime
    return false;
}

// `bytes` contiguous bytes from device memory src to pageable host memory
// dst through a stage: the DMA of chunk k+1 runs while the host threads copy
// chunk k out of the other buffer.  Work queued on s before the call is
// ordered before the copies; returns when dst holds everything.
static int stage_download(pg_ctx* ctx, hipStream_t s, void* dst, const void* src, size_t bytes) {
    Stage* st;
    PG_TRY(stage_get(ctx, &st));
    HostPool& pool = host_pool();
    const int nt = pool.size();
    const size_t nchunks = (bytes + Stage::CHUNK - 1) / Stage::CHUNK;
    hipError_t e = hipSuccess;
    auto drain = [&](size_t k) {  // chunk k's DMA is done: copy it out
        const size_t off = k * Stage::CHUNK, n = std::min(Stage::CHUNK, bytes - off);
        const char* from = (const char*)st->buf[k & 1];
        char* to = (char*)dst + off;
        const size_t per = (n + nt - 1) / nt;
        pool.run(nt, [&](int i) {
            const size_t a = (size_t)i * per, z = std::min(n, a + per);
            if (a < z) memcpy(to + a, from + a, z - a);
        });
    };
    for (size_t k = 0; k < nchunks && e == hipSuccess; k++) {
        const size_t off = k * Stage::CHUNK, n = std::min(Stage::CHUNK, bytes - off);
        // chunk k goes into buffer k & 1, whose previous chunk (k - 2) was
        // drained in the previous iteration
        e = hipMemcpyAsync(st->buf[k & 1], (const char*)src + off, n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(st->ev[k & 1], s);
        if (e == hipSuccess && k > 0) {
            e = hipEventSynchronize(st->ev[(k - 1) & 1]);
            if (e == hipSuccess) drain(k - 1);
        }
    }
    if (e == hipSuccess && nchunks) {
        e = hipEventSynchronize(st->ev[(nchunks - 1) & 1]);
        if (e == hipSuccess) drain(nchunks - 1);
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        stage_destroy(st);
        return fail(PG_E_DEVICE, "staged download: %s", hipGetErrorString(e));
    }
    stage_put(ctx, st);
    return PG_OK;
}

// `rows` rows of `width` bytes from src (pitch spitch, pageable host memory)
// to dst (device, pitch dpitch), queued on s through a stage: whole chunks
// when both sides are contiguous, else groups of rows gathered on the host
// and written with a pitched DMA.  Returns once the last DMA has read its
// buffer (the copies into dst are complete by then too).
// A stage kept in use across several uploads on one stream: the buffers keep
// rotating and nothing drains between the uploads, so the gather of the next
// upload's first chunk overlaps the DMA of the previous upload's last one
// (pg_pcs_commit_host uploads a matrix as several column blocks).
// stage_run_finish waits for the last DMAs and returns the stage.
struct StageRun {
    Stage* st = nullptr;
    size_t k = 0;
    bool used[Stage::NB] = {};
};

static int stage_upload(pg_ctx* ctx, hipStream_t s, void* dst, size_t dpitch, const void* src, size_t spitch,
                        size_t width, size_t rows, StageRun* run = nullptr) {
    const bool dense = dpitch == width && spitch == width;
    if (!dense && width > Stage::CHUNK) {
        HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyHostToDevice, s));
        return PG_OK;
    }
    StageRun local;
    StageRun& R = run ? *run : local;
    if (!R.st) PG_TRY(stage_get(ctx, &R.st));
    Stage* st = R.st;
    HostPool& pool = host_pool();
    const int nt = pool.size();
    const size_t total = dense ? width * rows : rows;  // bytes, or rows
    const size_t step = dense ? Stage::CHUNK : Stage::CHUNK / width;
    bool* used = R.used;
    hipError_t e = hipSuccess;
    size_t k = R.k;
    for (size_t done = 0; done < total && e == hipSuccess; done += step, k++) {
        const int b = (int)(k % Stage::NB);
        if (used[b]) e = hipEventSynchronize(st->ev[b]);  // its previous DMA is done
        if (e != hipSuccess) break;
        const size_t n = std::min(step, total - done);
        char* buf = (char*)st->buf[b];
        if (dense) {
            const char* from = (const char*)src + done;
            const size_t per = (n + nt - 1) / nt;
            pool.run(nt, [&](int i) {
                const size_t a = (size_t)i * per, z = std::min(n, a + per);
                if (a < z) memcpy(buf + a, from + a, z - a);
            });
            e = hipMemcpyAsync((char*)dst + done, buf, n, hipMemcpyHostToDevice, s);
        } else {
            const size_t per = (n + nt - 1) / nt;
            pool.run(nt, [&](int i) {
                for (size_t r = (size_t)i * per; r < std::min(n, (size_t)(i + 1) * per); r++)
                    memcpy(buf + r * width, (const char*)src + (done + r) * spitch, width);
            });
            e = dpitch == width
                    ? hipMemcpyAsync((char*)dst + done * width, buf, n * width, hipMemcpyHostToDevice, s)
                    : hipMemcpy2DAsync((char*)dst + done * dpitch, dpitch, buf, width, width, n, hipMemcpyHostToDevice,
                                       s);
        }
        if (e == hipSuccess) e = hipEventRecord(st->ev[b], s);
        used[b] = e == hipSuccess;
    }
    R.k = k;
    if (e == hipSuccess && run) return PG_OK;  // the run goes on; its owner finishes it
    for (int b = 0; b < Stage::NB; b++)
        if (used[b]) {
            const hipError_t e2 = hipEventSynchronize(st->ev[b]);
            if (e == hipSuccess) e = e2;
        }
    R.st = nullptr;
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);  // nothing may still read the buffers
        stage_destroy(st);
        return fail(PG_E_DEVICE, "staged upload: %s", hipGetErrorString(e));
    }
    stage_put(ctx, st);
    return PG_OK;
}

// Waits for a run's last DMAs and returns its stage (destroyed when a wait
// fails).  Safe on a run that never started or already ended.
static void stage_run_finish(pg_ctx* ctx, StageRun* run) {
    if (!run->st) return;
    bool ok = true;
    for (int b = 0; b < Stage::NB; b++)
        if (run->used[b] && hipEventSynchronize(run->st->ev[b]) != hipSuccess) ok = false;
    if (ok)
        stage_put(ctx, run->st);
    else
        stage_destroy(run->st);
    run->st = nullptr;
}

extern "C" int pg_create(int device, size_t budget, pg_ctx** out) {
    if (!out) return fail(PG_E_INVALID, "out is null");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(PG_E_INVALID, "device %d of %d", device, ndev);
    HIP_TRY(hipSetDevice(device));
    PG_TRY(noalloc_scratch_init(device));
    std::unique_ptr<pg_ctx> c(new pg_ctx);
    c->device = device;
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        budget = free_b;
    }
    c->budget = budget;
    for (int i = 0; i < 4; i++) {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        c->streams.push_back(s);
    }
    *out = c.release();
    return PG_OK;
}

extern "C" void pg_destroy(pg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (auto s : ctx->streams) (void)hipStreamSynchronize(s);
    ctx->twiddles.clear();
    ctx->scales.clear();
    for (void* st : ctx->stages) stage_destroy((Stage*)st);
    ctx->stages.clear();
    {
        std::lock_guard<std::mutex> g(ctx->pool_mu);
        for (auto& sg : ctx->segs)
            if (sg.base) (void)hipFree(sg.base);
        ctx->segs.clear();
    }
    for (auto s : ctx->streams) (void)hipStreamDestroy(s);
    delete ctx;
}

extern "C" size_t pg_bytes_in_use(const pg_ctx* ctx) { return ctx ? ctx->in_use.load() : 0; }

extern "C" int pg_device_synchronize(int device) {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceSynchronize());
    return PG_OK;
}

extern "C" int pg_profile_enable(pg_ctx* ctx, int enable) {
    if (!ctx) return fail(PG_E_INVALID, "null context");
    ctx->profiling.store(enable != 0);
    return PG_OK;
}

extern "C" int pg_profile_reset(pg_ctx* ctx) {
    if (!ctx) return fail(PG_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->prof.clear();
    return PG_OK;
}

extern "C" int pg_profile_get(pg_ctx* ctx, const char* name, double* total_ms, uint64_t* launches) {
    if (!ctx || !name || !total_ms || !launches) return fail(PG_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->prof.find(name);
    *total_ms = it == ctx->prof.end() ? 0.0 : it->second.first;
    *launches = it == ctx->prof.end() ? 0 : it->second.second;
    return PG_OK;
}

extern "C" int pg_debug_fail_alloc(pg_ctx* ctx, long long after) {
    if (!ctx) return fail(PG_E_INVALID, "null context");
    ctx->fail_countdown.store(after < 0 ? -1 : after);
    return PG_OK;
}

extern "C" int pg_debug_host_blocks(pg_ctx* ctx, size_t block_bytes, size_t min_cols) {
    if (!ctx) return fail(PG_E_INVALID, "null context");
    ctx->host_block_bytes.store(block_bytes ? block_bytes : HOST_BLOCK_BYTES);
    ctx->host_block_min_cols.store(min_cols ? min_cols : HOST_BLOCK_MIN_COLS);
    return PG_OK;
}

// Twiddle table w_{2^log_n}^e, e < 2^log_n; built once per (field, log_n).
static int get_twiddles(pg_ctx* ctx, hipStream_t s, int field, int log_n, std::shared_ptr<DevMem>* out) {
    std::lock_guard<std::mutex> g(ctx->mu);
    auto key = std::make_pair(field, log_n);
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) {
        *out = it->second;
        return PG_OK;
    }
    int rc;
    auto m = make_devmem(ctx, sizeof(uint32_t) << log_n, &rc);
    if (!m) return rc;
    HIP_TRY(launch_twiddles(field, log_n, (uint32_t*)m->ptr, s));
    HIP_TRY(hipStreamSynchronize(s));
    ctx->twiddles[key] = m;
    *out = m;
    return PG_OK;
}

static constexpr size_t SCALE_CACHE_BYTES = 512ull << 20;

static int get_scale(pg_ctx* ctx, hipStream_t s, int field, int log_h, int b, uint32_t shift,
                     std::shared_ptr<DevMem>* out) {
    std::lock_guard<std::mutex> g(ctx->mu);
    auto key = std::make_tuple(field, log_h, b, shift);
    auto it = ctx->scales.find(key);
    if (it != ctx->scales.end()) {
        *out = it->second;
        return PG_OK;
    }
    const size_t bytes = sizeof(uint32_t) << (log_h + b);
    while (ctx->scale_bytes + bytes > SCALE_CACHE_BYTES && !ctx->scale_order.empty()) {
        auto old = ctx->scale_order.front();
        ctx->scale_order.erase(ctx->scale_order.begin());
        auto o = ctx->scales.find(old);
        if (o != ctx->scales.end()) {
            ctx->scale_bytes -= o->second->bytes;
            ctx->scales.erase(o);  // freed once the last user drops it
        }
    }
    int rc;
    auto m = make_devmem(ctx, bytes, &rc);
    if (!m) return rc;
    HIP_TRY(launch_coset_scale(field, log_h, b, shift, (uint32_t*)m->ptr, s));
    HIP_TRY(hipStreamSynchronize(s));
    ctx->scales[key] = m;
    ctx->scale_order.push_back(key);
    ctx->scale_bytes += bytes;
    *out = m;
    return PG_OK;
}

// ---------------------------------------------------------------- buffers --
struct pg_buf {
    pg_ctx* ctx = nullptr;
    uint32_t* ptr = nullptr;
    size_t rows = 0, cols = 0;
    size_t ld = 0;  // row pitch in words: cols, or more for a strided view
    size_t bytes = 0;
    bool owned = false;
};

// Entry points whose kernels read a matrix as rows * cols contiguous words
// turn strided views away; LDE inputs and downloads take them.
static int need_dense(const pg_buf* b, const char* what) {
    if (b && b->ld != b->cols && b->rows > 1)
        return fail(PG_E_INVALID, "%s: a strided view is accepted only as an LDE input or for download", what);
    return PG_OK;
}

static int is_pow2(size_t x) { return x && !(x & (x - 1)); }
static int log2_of(size_t x) {
    int l = 0;
    while (((size_t)1 << l) < x) l++;
    return l;
}

static int buf_new(pg_ctx* ctx, size_t rows, size_t cols, pg_buf** out) {
    if (rows && cols > (SIZE_MAX / 4) / rows) return fail(PG_E_INVALID, "matrix too large");
    std::unique_ptr<pg_buf> b(new pg_buf);
    b->ctx = ctx;
    b->rows = rows;
    b->cols = cols;
    b->ld = cols;
    b->bytes = rows * cols * 4;
    b->owned = true;
    void* p;
    PG_TRY(ctx_alloc(ctx, b->bytes, &p));
    b->ptr = (uint32_t*)p;
    *out = b.release();
    return PG_OK;
}

extern "C" int pg_buf_alloc(pg_ctx* ctx, size_t rows, size_t cols, pg_buf** out) {
    if (!ctx || !out) return fail(PG_E_INVALID, "null argument");
    PG_TRY(bind_device(ctx));
    return buf_new(ctx, rows, cols, out);
}

extern "C" int pg_buf_wrap(pg_ctx* ctx, void* dev, size_t rows, size_t cols, pg_buf** out) {
    if (!ctx || !out || (!dev && rows * cols)) return fail(PG_E_INVALID, "null argument");
    pg_buf* b = new pg_buf;
    b->ctx = ctx;
    b->ptr = (uint32_t*)dev;
    b->rows = rows;
    b->cols = cols;
    b->ld = cols;
    b->bytes = rows * cols * 4;
    b->owned = false;
    *out = b;
    return PG_OK;
}

extern "C" int pg_buf_view(pg_ctx* ctx, const pg_buf* parent, size_t row0, size_t row_step, size_t col0,
                           size_t rows, size_t cols, pg_buf** out) {
    if (!ctx || !parent || !out) return fail(PG_E_INVALID, "null argument");
    *out = nullptr;
    if (row_step == 0 || cols == 0) return fail(PG_E_INVALID, "row step and width must be positive");
    if (col0 > parent->cols || cols > parent->cols - col0)
        return fail(PG_E_INVALID, "columns %zu..%zu outside width %zu", col0, col0 + cols, parent->cols);
    if (rows && (row0 >= parent->rows || (rows - 1) > (parent->rows - 1 - row0) / row_step))
        return fail(PG_E_INVALID, "rows %zu + %zu * k (k < %zu) outside height %zu", row0, row_step, rows,
                    parent->rows);
    if (parent->ld && row_step > SIZE_MAX / parent->ld) return fail(PG_E_INVALID, "row step too large");
    pg_buf* b = new pg_buf;
    b->ctx = ctx;
    b->ptr = parent->ptr + (rows ? row0 * parent->ld + col0 : 0);
    b->rows = rows;
    b->cols = cols;
    b->ld = rows > 1 ? parent->ld * row_step : cols;
    b->bytes = rows * cols * 4;
    b->owned = false;
    *out = b;
    return PG_OK;
}

extern "C" void pg_buf_free(pg_buf* b) {
    if (!b) return;
    if (b->owned) {
        (void)hipSetDevice(b->ctx->device);
        ctx_free(b->ctx, b->ptr, b->bytes);
    }
    delete b;
}

extern "C" void* pg_buf_device_ptr(const pg_buf* b) { return b ? b->ptr : nullptr; }
extern "C" size_t pg_buf_rows(const pg_buf* b) { return b ? b->rows : 0; }
extern "C" size_t pg_buf_cols(const pg_buf* b) { return b ? b->cols : 0; }
extern "C" size_t pg_buf_ld(const pg_buf* b) { return b ? b->ld : 0; }

static bool field_ok(int f) { return f == PG_KOALABEAR || f == PG_BABYBEAR; }

extern "C" int pg_upload(pg_ctx* ctx, pg_buf* b, const uint32_t* host, int field, int canonical) {
    if (!ctx || !b || (!host && b->bytes)) return fail(PG_E_INVALID, "null argument");
    if (canonical && !field_ok(field)) return fail(PG_E_INVALID, "unknown field %d", field);
    PG_TRY(need_dense(b, "pg_upload"));
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    if (b->bytes >= STAGE_MIN_BYTES && !host_is_locked(host))
        PG_TRY(stage_upload(ctx, l.s, b->ptr, b->bytes, host, b->bytes, b->bytes, 1));
    else
        HIP_TRY(hipMemcpyAsync(b->ptr, host, b->bytes, hipMemcpyHostToDevice, l.s));
    if (canonical) HIP_TRY(launch_convert(field, b->ptr, b->rows * b->cols, 1, l.s));
    HIP_TRY(hipStreamSynchronize(l.s));
    return PG_OK;
}

extern "C" int pg_host_alloc(pg_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(PG_E_INVALID, "null argument");
    *out = nullptr;
    PG_TRY(bind_device(ctx));
    hipError_t e = hipHostMalloc(out, bytes ? bytes : 4, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(PG_E_OOM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return PG_OK;
}

extern "C" void pg_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

extern "C" int pg_download(pg_ctx* ctx, const pg_buf* b, size_t row0, size_t nrows, uint32_t* host,
                           int field, int canonical) {
    if (!ctx || !b || !host) return fail(PG_E_INVALID, "null argument");
    if (row0 > b->rows || nrows > b->rows - row0) return fail(PG_E_INVALID, "row range out of bounds");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const size_t n = nrows * b->cols;
    if (n == 0) return PG_OK;
    const uint32_t* src = b->ptr + row0 * b->ld;
    // rows x cols words out of a pitch-ld source into a dense destination
    auto copy_rows = [&](void* dst, hipMemcpyKind kind) {
        return b->ld == b->cols ? hipMemcpyAsync(dst, src, n * 4, kind, l.s)
                                : hipMemcpy2DAsync(dst, b->cols * 4, src, b->ld * 4, b->cols * 4, nrows, kind, l.s);
    };
    const bool staged = n * 4 >= STAGE_MIN_BYTES && !host_is_locked(host);
    if (canonical) {
        if (!field_ok(field)) return fail(PG_E_INVALID, "unknown field %d", field);
        void* t;
        PG_TRY(ctx_alloc(ctx, n * 4, &t));
        hipError_t e = copy_rows(t, hipMemcpyDeviceToDevice);
        if (e == hipSuccess) e = launch_convert(field, (uint32_t*)t, n, 0, l.s);
        int rc = PG_OK;
        if (e == hipSuccess) {
            if (staged)
                rc = stage_download(ctx, l.s, host, t, n * 4);
            else
                e = hipMemcpyAsync(host, t, n * 4, hipMemcpyDeviceToHost, l.s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(l.s);
        ctx_free(ctx, t, n * 4);
        if (rc != PG_OK) return rc;
        HIP_TRY(e);
        return PG_OK;
    }
    if (staged && b->ld == b->cols) {
        PG_TRY(stage_download(ctx, l.s, host, src, n * 4));
        return PG_OK;
    }
    HIP_TRY(copy_rows(host, hipMemcpyDeviceToHost));
    HIP_TRY(hipStreamSynchronize(l.s));
    return PG_OK;
}

extern "C" int pg_fill_random(pg_ctx* ctx, pg_buf* b, int field, uint64_t seed) {
    if (!ctx || !b || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    PG_TRY(need_dense(b, "pg_fill_random"));
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    HIP_TRY(launch_fill_random(field, seed, b->ptr, b->rows * b->cols, 1, l.s));
    HIP_TRY(hipStreamSynchronize(l.s));
    return PG_OK;
}

// ------------------------------------------------------------------- LDE --
static uint32_t gen_monty(int field) {
    return field == PG_KOALABEAR ? to_monty<KoalaBear>(KoalaBear::GEN) : to_monty<BabyBear>(BabyBear::GEN);
}
static int two_adicity(int field) { return field == PG_KOALABEAR ? KoalaBear::TWO_ADICITY : BabyBear::TWO_ADICITY; }

// Shape check of an LDE input; *L_out = log2(rows).
static int lde_check(int field, size_t rows, size_t cols, uint32_t b, int* L_out) {
    if (!is_pow2(rows)) return fail(PG_E_INVALID, "height %zu is not a power of two", rows);
    if (cols == 0) return fail(PG_E_INVALID, "width is zero");
    const int L = log2_of(rows);
    if (L > lde_max_log_h()) return fail(PG_E_UNSUPPORTED, "log height %d > %d", L, lde_max_log_h());
    if (L + (int)b > two_adicity(field))
        return fail(PG_E_INVALID, "LDE size 2^%d exceeds the field's two-adicity %d", L + (int)b, two_adicity(field));
    if (b > 8) return fail(PG_E_UNSUPPORTED, "added_bits %u > 8", b);
    *L_out = L;
    return PG_OK;
}

// Device descriptor tables built on the host, uploaded with one copy.
struct Staging {
    std::vector<uint8_t> bytes;
    size_t put(const void* p, size_t n) {  // p == nullptr reserves n zero bytes
        const size_t off = round_up(bytes.size(), 256);
        bytes.resize(off + n, 0);
        if (n && p) memcpy(bytes.data() + off, p, n);
        return off;
    }
};

// LDE of `in` into a new N x w buffer on stream s; `scratch` (may be null)
// is reused when large enough.
static int lde_on_stream(pg_ctx* ctx, hipStream_t s, int field, const pg_buf* in, uint32_t b,
                         uint32_t shift, int bitrev, pg_buf** out, std::shared_ptr<DevMem>* scratch) {
    *out = nullptr;
    int L;
    PG_TRY(lde_check(field, in->rows, in->cols, b, &L));
    std::shared_ptr<DevMem> tw, sc;
    PG_TRY(get_twiddles(ctx, s, field, L, &tw));
    PG_TRY(get_scale(ctx, s, field, L, (int)b, shift, &sc));
    const size_t N = in->rows << b;
    pg_buf* o;
    PG_TRY(buf_new(ctx, N, in->cols, &o));
    uint32_t* tmp = nullptr;
    if (L > 11) {
        const size_t need = N * in->cols * 4;
        if (!*scratch || (*scratch)->bytes < need) {
            scratch->reset();
            int rc;
            *scratch = make_devmem(ctx, need, &rc);
            if (!*scratch) {
                pg_buf_free(o);
                return rc;
            }
        }
        tmp = (uint32_t*)(*scratch)->ptr;
    }
    auto kt = make_timer(ctx);
    hipError_t e = launch_coset_lde(field, in->ptr, in->ld, L, (uint32_t)in->cols, (int)b,
                                    (const uint32_t*)tw->ptr, (const uint32_t*)sc->ptr, tmp, o->ptr, in->cols,
                                    bitrev, s, kt.get());
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // tables and scratch stay referenced until done
    if (e != hipSuccess) {
        pg_buf_free(o);
        return fail(PG_E_DEVICE, "coset LDE failed: %s", hipGetErrorString(e));
    }
    *out = o;
    return PG_OK;
}

extern "C" int pg_coset_lde(pg_ctx* ctx, int field, const pg_buf* evals, uint32_t added_bits,
                            uint32_t shift_mont, int bitrev, pg_buf** out) {
    if (!ctx || !evals || !out || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    std::shared_ptr<DevMem> scratch;
    return lde_on_stream(ctx, l.s, field, evals, added_bits, shift_mont, bitrev, out, &scratch);
}

// ------------------------------------------------------------------ MMCS --
struct pg_tree {
    pg_ctx* ctx = nullptr;
    int field = 0;
    size_t H = 0;
    int log_h = 0;
    std::vector<pg_buf*> mats;  // original order
    std::vector<bool> owned;
    size_t total_width = 0;
    uint32_t* digests = nullptr;  // layers 0..log_h, (2H - 1) * 8 words
    size_t digest_bytes = 0;
    void* meta = nullptr;  // device: MatRef orig[n], shift[n], coloff[n], layer_off[log_h + 1]
    size_t meta_bytes = 0;
    MatRef* d_mats = nullptr;
    uint32_t* d_shift = nullptr;
    uint32_t* d_coloff = nullptr;
    uint64_t* d_layer_off = nullptr;
    uint32_t root[8] = {0};
};

extern "C" void pg_tree_free(pg_tree* t) {
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    for (size_t k = 0; k < t->mats.size(); k++)
        if (t->owned[k]) pg_buf_free(t->mats[k]);
    ctx_free(t->ctx, t->digests, t->digest_bytes);
    ctx_free(t->ctx, t->meta, t->meta_bytes);
    delete t;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static void tree_discard(pg_tree* x) {  // release only what the tree allocated itself
    for (size_t k = 0; k < x->owned.size(); k++) x->owned[k] = false;
    pg_tree_free(x);
}
using TreePtr = std::unique_ptr<pg_tree, void (*)(pg_tree*)>;

// Allocates a tree's digest layers and device metadata (matrix descriptors
// in original and height-sorted order, height shifts, column offsets, layer
// offsets) without hashing anything.  The tree owns none of `mats` yet.
static int tree_prepare(pg_ctx* ctx, hipStream_t s, int field, const std::vector<pg_buf*>& mats, TreePtr* out,
                        std::vector<size_t>* order_out, MatRef** d_sorted_out) {
    const size_t n = mats.size();
    if (n == 0) return fail(PG_E_INVALID, "no matrices to commit");
    size_t H = 0;
    for (auto* m : mats) {
        if (!m) return fail(PG_E_INVALID, "null matrix");
        if (!is_pow2(m->rows)) return fail(PG_E_INVALID, "height %zu is not a power of two", m->rows);
        H = std::max(H, m->rows);
    }
    TreePtr t(new pg_tree, tree_discard);
    t->ctx = ctx;
    t->field = field;
    t->H = H;
    t->log_h = log2_of(H);
    t->mats = mats;
    t->owned = std::vector<bool>(n, false);
    // stable sort by height, tallest first
    std::vector<size_t> order(n);
    for (size_t k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return mats[a]->rows > mats[b]->rows; });

    // device metadata block
    const int lh = t->log_h;
    const size_t off_sorted = 0;
    const size_t off_orig = align_up(off_sorted + n * sizeof(MatRef), 256);
    const size_t off_shift = align_up(off_orig + n * sizeof(MatRef), 256);
    const size_t off_col = align_up(off_shift + n * 4, 256);
    const size_t off_layer = align_up(off_col + n * 4, 256);
    const size_t meta_bytes = off_layer + (lh + 1) * 8;
    std::vector<uint8_t> hm(meta_bytes, 0);
    MatRef* h_sorted = (MatRef*)(hm.data() + off_sorted);
    MatRef* h_orig = (MatRef*)(hm.data() + off_orig);
    uint32_t* h_shift = (uint32_t*)(hm.data() + off_shift);
    uint32_t* h_col = (uint32_t*)(hm.data() + off_col);
    uint64_t* h_layer = (uint64_t*)(hm.data() + off_layer);
    size_t col = 0;
    for (size_t k = 0; k < n; k++) {
        h_orig[k] = MatRef{mats[k]->ptr, (uint64_t)mats[k]->cols, (uint32_t)mats[k]->cols, 0};
        h_shift[k] = (uint32_t)(lh - log2_of(mats[k]->rows));
        h_col[k] = (uint32_t)col;
        col += mats[k]->cols;
    }
    if (col > 0xffffffffull) return fail(PG_E_INVALID, "total width too large");
    t->total_width = col;
    for (size_t k = 0; k < n; k++) h_sorted[k] = h_orig[order[k]];
    {
        uint64_t off = 0;
        for (int j = 0; j <= lh; j++) {
            h_layer[j] = off;
            off += H >> j;
        }
    }
    PG_TRY(ctx_alloc(ctx, meta_bytes, &t->meta));
    t->meta_bytes = meta_bytes;
    uint8_t* dm = (uint8_t*)t->meta;
    MatRef* d_sorted = (MatRef*)(dm + off_sorted);
    t->d_mats = (MatRef*)(dm + off_orig);
    t->d_shift = (uint32_t*)(dm + off_shift);
    t->d_coloff = (uint32_t*)(dm + off_col);
    t->d_layer_off = (uint64_t*)(dm + off_layer);
    HIP_TRY(hipMemcpyAsync(t->meta, hm.data(), meta_bytes, hipMemcpyHostToDevice, s));

    t->digest_bytes = (2 * H - 1) * 32;
    {
        void* p;
        PG_TRY(ctx_alloc(ctx, t->digest_bytes, &p));
        t->digests = (uint32_t*)p;
    }
    *order_out = std::move(order);
    *d_sorted_out = d_sorted;
    *out = std::move(t);
    return PG_OK;
}

// `stage`, when given, is called with each run [g0, g1) of equal-height
// matrices in height-sorted order right before the first kernel that reads
// them is enqueued on s: it may enqueue their production on s (or make s
// wait for it), so a commit can start hashing tall matrices while shorter
// ones are still arriving.  For the tallest run it also gets the leaf
// digest layer; returning STAGE_HASHED says it has queued those digests
// itself (and the leaf hash is skipped).  For a shorter run it may set *inj
// to the digests of the run's concatenated rows, queued on s already; the
// layer then only compresses.  `sorted` is the device descriptor array of
// the matrices in height order.
static constexpr int STAGE_HASHED = 1;
using StageFn = std::function<int(const std::vector<size_t>& order, size_t g0, size_t g1, uint32_t* leaves,
                                  const MatRef* sorted, const uint32_t** inj)>;

static int mmcs_commit_on_stream(pg_ctx* ctx, hipStream_t s, int field, const std::vector<pg_buf*>& mats,
                                 const std::vector<bool>& owned, uint32_t root[8], pg_tree** out,
                                 const StageFn& stage = nullptr) {
    *out = nullptr;
    const size_t n = mats.size();
    TreePtr t(nullptr, tree_discard);
    std::vector<size_t> order;
    MatRef* d_sorted = nullptr;
    PG_TRY(tree_prepare(ctx, s, field, mats, &t, &order, &d_sorted));
    SyncGuard guard(s);  // on an error below: wait for queued kernels before `t` frees
    const size_t H = t->H;
    // layer 0: every matrix of height H
    size_t g0 = 0, g1 = 0;
    uint32_t wsum = 0;
    while (g1 < n && mats[order[g1]]->rows == H) wsum += (uint32_t)mats[order[g1++]]->cols;
    const pg_buf* first = mats[order[0]];
    auto kt = make_timer(ctx);
    bool hashed = false;
    if (stage) {
        const int r = stage(order, g0, g1, t->digests, d_sorted, nullptr);
        if (r < 0) return r;
        hashed = r == STAGE_HASHED;
    }
    if (hashed) {
    } else if (g1 - g0 == 1 && first->cols % 8 == 0 && first->cols > 0 && ((uintptr_t)first->ptr % 16) == 0)
        TIMED(kt, "leaf_hash_w8", s,
              launch_leaf_hash_w8(field, first->ptr, first->cols, (uint32_t)first->cols, H, t->digests, s));
    else
        TIMED(kt, "leaf_hash", s, launch_leaf_hash(field, d_sorted + g0, (int)(g1 - g0), wsum, H, t->digests, s));
    // compress + inject
    size_t len = H, off = 0, gi = g1;
    while (len > 1) {
        const size_t nl = len / 2;
        size_t gj = gi;
        uint32_t iw = 0;
        while (gj < n && mats[order[gj]]->rows == nl) iw += (uint32_t)mats[order[gj++]]->cols;
        const uint32_t* inj = nullptr;
        if (stage && gj > gi) PG_TRY(stage(order, gi, gj, nullptr, d_sorted, &inj));
        if (inj)
            TIMED(kt, "compress_digest", s,
                  launch_compress_digest(field, t->digests + 8 * off, t->digests + 8 * (off + len), nl, inj, s));
        else
            TIMED(kt, gj > gi ? "compress_inject" : "compress", s,
                  launch_compress(field, t->digests + 8 * off, t->digests + 8 * (off + len), nl,
                                  gj > gi ? d_sorted + gi : nullptr, (int)(gj - gi), iw, s));
        off += len;
        len = nl;
        gi = gj;
    }
    HIP_TRY(hipMemcpyAsync(t->root, t->digests + 8 * off, 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    guard.armed = false;
    memcpy(root, t->root, 32);
    t->owned = owned;
    *out = t.release();
    return PG_OK;
}

extern "C" int pg_mmcs_commit(pg_ctx* ctx, int field, pg_buf* const* mats, size_t n, int take,
                              uint32_t root[8], pg_tree** out) {
    if (!ctx || !mats || !root || !out || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    for (size_t k = 0; k < n; k++) PG_TRY(need_dense(mats[k], "pg_mmcs_commit"));
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    std::vector<pg_buf*> v(mats, mats + n);
    return mmcs_commit_on_stream(ctx, l.s, field, v, std::vector<bool>(n, take != 0), root, out);
}

static int open_on_stream(pg_ctx* ctx, hipStream_t s, const pg_tree* t, const uint64_t* idx, size_t nq,
                          uint32_t* rows_out, uint32_t* paths_out) {
    for (size_t q = 0; q < nq; q++)
        if (idx[q] >= t->H) return fail(PG_E_INVALID, "query index %llu >= height %zu",
                                        (unsigned long long)idx[q], t->H);
    if (nq == 0) return PG_OK;
    const size_t rw = nq * t->total_width * 4, pw = nq * t->log_h * 32;
    const size_t o_rows = align_up(nq * 8, 256), o_paths = align_up(o_rows + rw, 256);
    const size_t bytes = o_paths + pw;
    void* d;
    PG_TRY(ctx_alloc(ctx, bytes, &d));
    uint8_t* dm = (uint8_t*)d;
    hipError_t e = hipMemcpyAsync(dm, idx, nq * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_open_gather(t->d_mats, t->d_shift, t->d_coloff, (int)t->mats.size(), (uint32_t)t->total_width,
                               t->digests, t->d_layer_off, t->log_h, (const uint64_t*)dm, (int)nq,
                               (uint32_t*)(dm + o_rows), (uint32_t*)(dm + o_paths), s);
    if (e == hipSuccess && rows_out && rw) e = hipMemcpyAsync(rows_out, dm + o_rows, rw, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && paths_out && pw) e = hipMemcpyAsync(paths_out, dm + o_paths, pw, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(ctx, d, bytes);
    HIP_TRY(e);
    return PG_OK;
}

extern "C" int pg_mmcs_open_batch(pg_ctx* ctx, const pg_tree* t, const uint64_t* idx, size_t nq,
                                  uint32_t* rows_out, uint32_t* paths_out) {
    if (!ctx || !t || (nq && !idx)) return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    return open_on_stream(ctx, l.s, t, idx, nq, rows_out, paths_out);
}

extern "C" int pg_tree_root(const pg_tree* t, uint32_t root[8]) {
    if (!t || !root) return fail(PG_E_INVALID, "null argument");
    memcpy(root, t->root, 32);
    return PG_OK;
}
extern "C" size_t pg_tree_num_matrices(const pg_tree* t) { return t ? t->mats.size() : 0; }
extern "C" const pg_buf* pg_tree_matrix(const pg_tree* t, size_t k) {
    return (t && k < t->mats.size()) ? t->mats[k] : nullptr;
}
extern "C" int pg_tree_log_height(const pg_tree* t) { return t ? t->log_h : -1; }
extern "C" size_t pg_tree_total_width(const pg_tree* t) { return t ? t->total_width : 0; }

extern "C" int pg_tree_digest_layer(pg_ctx* ctx, const pg_tree* t, int layer, uint32_t* host) {
    if (!ctx || !t || !host || layer < 0 || layer > t->log_h) return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    size_t off = 0;
    for (int j = 0; j < layer; j++) off += t->H >> j;
    HIP_TRY(hipMemcpyAsync(host, t->digests + 8 * off, (t->H >> layer) * 32, hipMemcpyDeviceToHost, l.s));
    HIP_TRY(hipStreamSynchronize(l.s));
    return PG_OK;
}

// ---------------------------------------------- host Poseidon2 (AVX2) --
// The challenger duplexes on the host (every opened value of Pcs::open is
// observed: ~2200 permutations for a 14-matrix shard), and host verification
// hashes too.  One width-16 state is two 8-lane vectors: s0..s7 and s8..s15,
// each 128-bit half one block of the external layer.  Montgomery products
// take the even and odd lanes through vpmuludq; M4 is in-lane shuffles
// (y_i = (x_0+x_1+x_2+x_3) + x_i + 2 x_{i+1}), the sum over blocks one
// 128-bit swap.  All values stay reduced, so the result is the permutation
// word for word (the CPU tests compare the challenger and verify_batch with
// the oracle).  Used when the CPU has AVX2, else the scalar permute.
// Measured per permutation: 0.49 against 0.54 us on an MI355X host's CPU,
// 0.92 against 1.67 us on a slower build host; the 20 partial rounds, a
// dependent chain through lane 0, bound both.
#ifndef __HIP_DEVICE_COMPILE__
namespace hostp2 {
#define PG_AVX2 __attribute__((target("avx2"))) static inline
template <class F>
struct C {
    __m256i p, mu, diag_lo, diag_hi, rc_lo[8], rc_hi[8];
    __attribute__((target("avx2"))) void init() {
        p = _mm256_set1_epi32((int)F::P);
        mu = _mm256_set1_epi32((int)F::MU);  // -p^-1 mod 2^32
        diag_lo = _mm256_loadu_si256((const __m256i*)&P2<F>::K.diag[0]);
        diag_hi = _mm256_loadu_si256((const __m256i*)&P2<F>::K.diag[8]);
        for (int r = 0; r < 8; r++) {
            rc_lo[r] = _mm256_loadu_si256((const __m256i*)&P2<F>::K.ext_rc[r][0]);
            rc_hi[r] = _mm256_loadu_si256((const __m256i*)&P2<F>::K.ext_rc[r][8]);
        }
    }
};
PG_AVX2 __m256i add(__m256i a, __m256i b, __m256i p) {
    const __m256i r = _mm256_add_epi32(a, b);
    return _mm256_min_epu32(r, _mm256_sub_epi32(r, p));
}
// a b 2^-32 mod p for a, b < p: t = a b; m = t (-p^-1) mod 2^32; (t + m p) / 2^32 < 2p
PG_AVX2 __m256i mul(__m256i a, __m256i b, __m256i p, __m256i neg_pinv) {
    __m256i te = _mm256_mul_epu32(a, b);
    __m256i to = _mm256_mul_epu32(_mm256_srli_epi64(a, 32), _mm256_srli_epi64(b, 32));
    const __m256i me = _mm256_mul_epu32(_mm256_mul_epu32(te, neg_pinv), p);
    const __m256i mo = _mm256_mul_epu32(_mm256_mul_epu32(to, neg_pinv), p);
    te = _mm256_srli_epi64(_mm256_add_epi64(te, me), 32);
    to = _mm256_add_epi64(to, mo);
    const __m256i r = _mm256_blend_epi32(te, to, 0xAA);
    return _mm256_min_epu32(r, _mm256_sub_epi32(r, p));
}
template <class F>
PG_AVX2 __m256i sbox(__m256i x, const C<F>& c) {
    const __m256i x2 = mul(x, x, c.p, c.mu), x3 = mul(x2, x, c.p, c.mu);
    if constexpr (F::SBOX == 3) return x3;
    return mul(mul(x3, x3, c.p, c.mu), x, c.p, c.mu);
}
template <class F>
PG_AVX2 __m256i m4(__m256i x, const C<F>& c) {
    const __m256i a = add(x, _mm256_shuffle_epi32(x, _MM_SHUFFLE(2, 3, 0, 1)), c.p);
    const __m256i s4 = add(a, _mm256_shuffle_epi32(a, _MM_SHUFFLE(1, 0, 3, 2)), c.p);
    const __m256i nx = _mm256_shuffle_epi32(x, _MM_SHUFFLE(0, 3, 2, 1));
    return add(add(s4, x, c.p), add(nx, nx, c.p), c.p);
}
template <class F>
PG_AVX2 void external(__m256i& lo, __m256i& hi, const C<F>& c) {
    lo = m4<F>(lo, c);
    hi = m4<F>(hi, c);
    __m256i t = add(lo, hi, c.p);
    t = add(t, _mm256_permute2x128_si256(t, t, 0x01), c.p);
    lo = add(lo, t, c.p);
    hi = add(hi, t, c.p);
}
template <class F>
PG_AVX2 void internal(__m256i& lo, __m256i& hi, const C<F>& c) {
    __m256i t = add(lo, hi, c.p);
    t = add(t, _mm256_shuffle_epi32(t, _MM_SHUFFLE(2, 3, 0, 1)), c.p);
    t = add(t, _mm256_shuffle_epi32(t, _MM_SHUFFLE(1, 0, 3, 2)), c.p);
    t = add(t, _mm256_permute2x128_si256(t, t, 0x01), c.p);
    lo = add(mul(lo, c.diag_lo, c.p, c.mu), t, c.p);
    hi = add(mul(hi, c.diag_hi, c.p, c.mu), t, c.p);
}
template <class F>
__attribute__((target("avx2"))) static void permute_avx2(uint32_t st[16]) {
    static const C<F> c = [] {
        C<F> v;
        v.init();
        return v;
    }();
    __m256i lo = _mm256_loadu_si256((const __m256i*)st), hi = _mm256_loadu_si256((const __m256i*)(st + 8));
    external<F>(lo, hi, c);
    for (int r = 0; r < 4; r++) {
        lo = sbox<F>(add(lo, c.rc_lo[r], c.p), c);
        hi = sbox<F>(add(hi, c.rc_hi[r], c.p), c);
        external<F>(lo, hi, c);
    }
    for (int r = 0; r < F::ROUNDS_P; r++) {
        uint32_t s0 = (uint32_t)_mm256_cvtsi256_si32(lo);
        s0 = fadd<F>(s0, P2<F>::K.int_rc[r]);
        s0 = F::SBOX == 3 ? fmul<F>(fmul<F>(s0, s0), s0)
                          : fmul<F>(fmul<F>(fmul<F>(s0, s0), s0), fmul<F>(fmul<F>(fmul<F>(s0, s0), s0), s0));
        lo = _mm256_blend_epi32(lo, _mm256_castsi128_si256(_mm_cvtsi32_si128((int)s0)), 0x01);
        internal<F>(lo, hi, c);
    }
    for (int r = 4; r < 8; r++) {
        lo = sbox<F>(add(lo, c.rc_lo[r], c.p), c);
        hi = sbox<F>(add(hi, c.rc_hi[r], c.p), c);
        external<F>(lo, hi, c);
    }
    _mm256_storeu_si256((__m256i*)st, lo);
    _mm256_storeu_si256((__m256i*)(st + 8), hi);
}
#undef PG_AVX2
}  // namespace hostp2
#endif

template <class F>
static void host_permute(uint32_t (&st)[16]) {
#ifndef __HIP_DEVICE_COMPILE__
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2) return hostp2::permute_avx2<F>(st);
#endif
    permute<F>(st);
}

// host sponge / compress for verify_batch and the challenger
template <class F>
static void host_sponge(const uint32_t* in, size_t n, uint32_t out[8]) {
    uint32_t st[16] = {0};
    for (size_t o = 0; o < n; o += 8) {
        const size_t len = std::min<size_t>(8, n - o);
        memcpy(st, in + o, len * 4);
        host_permute<F>(st);
    }
    memcpy(out, st, 32);
}
template <class F>
static void host_compress(const uint32_t* l, const uint32_t* r, uint32_t out[8]) {
    uint32_t st[16];
    memcpy(st, l, 32);
    memcpy(st + 8, r, 32);
    host_permute<F>(st);
    memcpy(out, st, 32);
}

template <class F>
static int verify_t(const size_t* heights, const size_t* widths, size_t n, const uint32_t root[8],
                    uint64_t index, const uint32_t* rows, const uint32_t* path) {
    size_t H = 0;
    for (size_t k = 0; k < n; k++) {
        if (!is_pow2(heights[k])) return fail(PG_E_INVALID, "height not a power of two");
        H = std::max(H, heights[k]);
    }
    const int lh = log2_of(H);
    std::vector<size_t> start(n + 1, 0);
    for (size_t k = 0; k < n; k++) start[k + 1] = start[k] + widths[k];
    std::vector<uint32_t> buf;
    auto gather = [&](size_t hgt) {
        buf.clear();
        bool any = false;
        for (size_t k = 0; k < n; k++)
            if (heights[k] == hgt) {
                any = true;
                buf.insert(buf.end(), rows + start[k], rows + start[k + 1]);
            }
        return any;
    };
    uint32_t cur[8], t[8];
    gather(H);
    host_sponge<F>(buf.data(), buf.size(), cur);
    size_t hc = H;
    for (int j = 0; j < lh; j++) {
        if ((index >> j) & 1)
            host_compress<F>(path + 8 * j, cur, t);
        else
            host_compress<F>(cur, path + 8 * j, t);
        memcpy(cur, t, 32);
        hc >>= 1;
        if (gather(hc)) {
            uint32_t rh[8];
            host_sponge<F>(buf.data(), buf.size(), rh);
            host_compress<F>(cur, rh, t);
            memcpy(cur, t, 32);
        }
    }
    return memcmp(cur, root, 32) == 0 ? 1 : 0;
}

extern "C" int pg_mmcs_verify_batch(int field, const size_t* heights, const size_t* widths, size_t n,
                                    const uint32_t root[8], uint64_t index, const uint32_t* rows,
                                    const uint32_t* path) {
    if (!heights || !widths || !root || n == 0 || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    return field == PG_KOALABEAR ? verify_t<KoalaBear>(heights, widths, n, root, index, rows, path)
                                 : verify_t<BabyBear>(heights, widths, n, root, index, rows, path);
}

// ------------------------------------------------------------------- PCS --
// Pcs::commit of device matrices.  The LDEs of all matrices are queued on
// one stream without a host synchronisation between them: matrices of one
// height share each pass launch (a device table of jobs, as in batched
// commits), a height held by a single matrix runs the plain per-matrix
// launches, and the tree follows on the same stream.  Heights run one after
// another and reuse one scratch buffer (the largest group's).  The
// reference commits all chips of a shard in one call, sorted by height
// (vm/src/machine/prover.rs:186-244).
extern "C" int pg_pcs_commit(pg_ctx* ctx, int field, const pg_buf* const* evals, const uint32_t* dshift,
                             size_t n, uint32_t log_blowup, uint32_t root[8], pg_tree** out) {
    if (!ctx || !evals || !root || !out || !field_ok(field) || n == 0) return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const uint32_t g = gen_monty(field);
    std::vector<int> L(n);
    std::vector<uint32_t> shift(n, g);
    for (size_t k = 0; k < n; k++) {
        if (!evals[k]) return fail(PG_E_INVALID, "null matrix %zu", k);
        PG_TRY(lde_check(field, evals[k]->rows, evals[k]->cols, log_blowup, &L[k]));
        if (dshift) {
            if (dshift[k] == 0) return fail(PG_E_INVALID, "domain shift is zero");
            shift[k] = field == PG_KOALABEAR ? fmul<KoalaBear>(g, finv<KoalaBear>(dshift[k]))
                                             : fmul<BabyBear>(g, finv<BabyBear>(dshift[k]));
        }
    }
    std::vector<pg_buf*> ldes;
    auto cleanup = [&]() {
        for (auto* b : ldes) pg_buf_free(b);
    };
    for (size_t k = 0; k < n; k++) {
        pg_buf* o;
        const int rc = buf_new(ctx, evals[k]->rows << log_blowup, evals[k]->cols, &o);
        if (rc != PG_OK) {
            cleanup();
            return rc;
        }
        ldes.push_back(o);
    }
    // tables (kept referenced until the stream is synchronised), scratch and
    // the job tables of the shared launches
    std::map<int, std::vector<size_t>, std::greater<int>> by_L;
    for (size_t k = 0; k < n; k++) by_L[L[k]].push_back(k);
    std::vector<std::shared_ptr<DevMem>> keep;
    std::map<int, const uint32_t*> tw_of;
    std::vector<const uint32_t*> sc_of(n);
    uint64_t tmp_words = 0;
    int rc = PG_OK;
    for (auto& kv : by_L) {
        std::shared_ptr<DevMem> tw;
        if (rc == PG_OK) rc = get_twiddles(ctx, l.s, field, kv.first, &tw);
        if (rc != PG_OK) break;
        keep.push_back(tw);
        tw_of[kv.first] = (const uint32_t*)tw->ptr;
        uint64_t words = 0;
        for (size_t k : kv.second) {
            std::shared_ptr<DevMem> sc;
            rc = get_scale(ctx, l.s, field, kv.first, (int)log_blowup, shift[k], &sc);
            if (rc != PG_OK) break;
            keep.push_back(sc);
            sc_of[k] = (const uint32_t*)sc->ptr;
            if (kv.first > 11) words += (uint64_t)ldes[k]->rows * ldes[k]->cols;
        }
        tmp_words = std::max(tmp_words, words);
    }
    std::shared_ptr<DevMem> scratch;
    if (rc == PG_OK && tmp_words) scratch = make_devmem(ctx, tmp_words * 4, &rc);
    if (rc != PG_OK) {
        cleanup();
        return rc;
    }
    uint32_t* const tmp = scratch ? (uint32_t*)scratch->ptr : nullptr;
    struct PassRec {
        int kind, log_s, variant, njobs;
        size_t off_args, off_prefix;
        uint64_t total;
    };
    std::map<int, std::vector<PassRec>> shared;  // log height -> its pass launches
    Staging st;
    for (auto& kv : by_L) {
        if (kv.second.size() < 2) continue;
        const int Lg = kv.first;
        int k1, k2;
        lde_split(Lg, &k1, &k2);
        std::vector<LdeArgs> args;
        uint64_t off = 0;  // each job's scratch rows within the group's share
        uint32_t wmax = 0;
        for (size_t k : kv.second) {
            LdeArgs a{};
            a.in = evals[k]->ptr;
            a.ld_in = evals[k]->ld;
            a.out = ldes[k]->ptr;
            a.ld_out = ldes[k]->cols;
            a.tmp = Lg > 11 ? tmp + off : nullptr;
            a.tw = tw_of[Lg];
            a.scale = sc_of[k];
            a.w = (uint32_t)ldes[k]->cols;
            a.L = Lg;
            a.k1 = k1;
            a.k2 = k2;
            a.b = (int)log_blowup;
            a.bitrev = 1;
            args.push_back(a);
            if (Lg > 11) off += (uint64_t)ldes[k]->rows * ldes[k]->cols;
            wmax = std::max(wmax, a.w);
        }
        const int kinds3[3] = {KIND_A, KIND_BC, KIND_D};
        for (int i = 0; i < (Lg <= 11 ? 1 : 3); i++) {
            const int kind = Lg <= 11 ? KIND_SINGLE : kinds3[i];
            const int log_s = (kind == KIND_BC || kind == KIND_SINGLE) ? k2 : k1;
            const int variant = lde_variant(field, kind, log_s, wmax);  // one tile shape per launch
            const int C = lde_pass_cols(kind, log_s, variant);
            std::vector<uint64_t> prefix(1, 0);
            for (auto& a : args)
                prefix.push_back(prefix.back() + lde_pass_row_tiles(kind, Lg, a.b) * ((a.w + C - 1) / C));
            shared[Lg].push_back(PassRec{kind, log_s, variant, (int)args.size(),
                                         st.put(args.data(), args.size() * sizeof(LdeArgs)),
                                         st.put(prefix.data(), prefix.size() * 8), prefix.back()});
        }
    }
    std::shared_ptr<DevMem> desc;
    if (!st.bytes.empty()) {
        desc = make_devmem(ctx, st.bytes.size(), &rc);
        if (!desc) {
            cleanup();
            return rc;
        }
    }
    auto kt = make_timer(ctx);
    hipError_t e = hipSuccess;
    if (desc) e = hipMemcpyAsync(desc->ptr, st.bytes.data(), st.bytes.size(), hipMemcpyHostToDevice, l.s);
    for (auto& kv : by_L) {
        if (e != hipSuccess) break;
        const int Lg = kv.first;
        if (kv.second.size() == 1) {
            const size_t k = kv.second[0];
            e = launch_coset_lde(field, evals[k]->ptr, evals[k]->ld, Lg, (uint32_t)evals[k]->cols, (int)log_blowup,
                                 tw_of[Lg], sc_of[k], Lg > 11 ? tmp : nullptr, ldes[k]->ptr, ldes[k]->cols, 1, l.s,
                                 kt.get());
            continue;
        }
        const uint8_t* dd = (const uint8_t*)desc->ptr;
        for (const PassRec& pr : shared[Lg]) {
            e = launch_lde_pass_batch(field, pr.kind, pr.log_s, pr.variant, (const LdeArgs*)(dd + pr.off_args),
                                      (const uint64_t*)(dd + pr.off_prefix), pr.njobs, pr.total, (int)log_blowup,
                                      l.s, kt.get());
            if (e != hipSuccess) break;
        }
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(l.s);  // queued work still reads the tables and writes the LDEs
        cleanup();
        return fail(PG_E_DEVICE, "coset LDE launch failed: %s", hipGetErrorString(e));
    }
    // the tree consumes the LDEs (it owns them from here) and synchronises
    rc = mmcs_commit_on_stream(ctx, l.s, field, ldes, std::vector<bool>(n, true), root, out);
    if (rc != PG_OK) {
        (void)hipStreamSynchronize(l.s);
        cleanup();
    }
    return rc;
}

// Pcs::commit of matrices in host memory.  Uploads run on a second stream,
// one height group at a time in the order the tree consumes them (tallest
// first); the compute stream waits for a group's upload event, runs its
// LDEs, then hashes or injects it, so the upload of shorter matrices overlaps
// the LDE and hashing of taller ones.  With pageable memory the host thread
// blocks inside each upload while the device works on the previous group.
extern "C" int pg_pcs_commit_host(pg_ctx* ctx, int field, const uint32_t* const* host, const size_t* rows,
                                  const size_t* cols, const uint32_t* dshift, size_t n, uint32_t log_blowup,
                                  int canonical, uint32_t root[8], pg_tree** out) {
    if (!ctx || !host || !rows || !cols || !root || !out || !field_ok(field) || n == 0)
        return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    std::vector<int> L(n);
    std::vector<uint32_t> shift(n, gen_monty(field));
    for (size_t k = 0; k < n; k++) {
        if (!host[k]) return fail(PG_E_INVALID, "null matrix %zu", k);
        PG_TRY(lde_check(field, rows[k], cols[k], log_blowup, &L[k]));
        if (rows[k] && cols[k] > (SIZE_MAX / 4) / (rows[k] << log_blowup)) return fail(PG_E_INVALID, "matrix too large");
        if (dshift) {
            if (dshift[k] == 0) return fail(PG_E_INVALID, "domain shift is zero");
            shift[k] = field == PG_KOALABEAR ? fmul<KoalaBear>(shift[k], finv<KoalaBear>(dshift[k]))
                                             : fmul<BabyBear>(shift[k], finv<BabyBear>(dshift[k]));
        }
    }
    PG_TRY(bind_device(ctx));
    Lease l(ctx), lc(ctx);
    std::vector<pg_buf*> in(n, nullptr), ldes(n, nullptr);
    std::vector<std::shared_ptr<DevMem>> tw(n), sc(n);
    std::shared_ptr<DevMem> scratch;
    std::vector<hipEvent_t> events;
    bool committed = false;
    StageRun run;  // every staged upload of this call, on lc.s
    auto cleanup = [&]() {  // after an error some work may still be queued
        (void)hipStreamSynchronize(lc.s);
        (void)hipStreamSynchronize(l.s);
        stage_run_finish(ctx, &run);
        for (auto* b : in) pg_buf_free(b);
        if (!committed)
            for (auto* b : ldes) pg_buf_free(b);
        for (auto e : events) (void)hipEventDestroy(e);
    };
    int rc = PG_OK;
    size_t tmp_bytes = 0;
    for (size_t k = 0; k < n && rc == PG_OK; k++) {
        rc = buf_new(ctx, rows[k], cols[k], &in[k]);
        if (rc == PG_OK) rc = buf_new(ctx, rows[k] << log_blowup, cols[k], &ldes[k]);
        // tables are built (and their stream synchronised) before anything is queued
        if (rc == PG_OK) rc = get_twiddles(ctx, l.s, field, L[k], &tw[k]);
        if (rc == PG_OK) rc = get_scale(ctx, l.s, field, L[k], (int)log_blowup, shift[k], &sc[k]);
        if (L[k] > 11) tmp_bytes = std::max(tmp_bytes, ldes[k] ? ldes[k]->bytes : 0);
    }
    if (rc == PG_OK && tmp_bytes) {
        scratch = make_devmem(ctx, tmp_bytes, &rc);
    }
    if (rc != PG_OK) {
        cleanup();
        return rc;
    }
    auto kt = make_timer(ctx);
    // The tallest matrices are cut into column blocks (Montgomery input only:
    // the canonical conversion works on whole matrices): each block is
    // uploaded, extended into its columns of the LDE and absorbed into every
    // row's sponge, so the leaf hash runs behind the upload instead of after
    // it.  On the device a block is stored dense (rows x cb words at word
    // rows * c0 of the input buffer), so its upload is one contiguous DMA
    // from a stage into which host threads gather the block's row segments
    // (a DMA that reads the pitched rows itself is slower, from page-locked
    // memory too).  Blocks of about HOST_BLOCK_BYTES keep the part of the work
    // that cannot overlap an upload (the last block's LDE and absorb) short;
    // at least HOST_BLOCK_MIN_COLS columns wide, as every block costs a pass
    // over the sponge states (128 bytes per LDE row) and narrow LDE tiles
    // leave lanes idle.  Measured on MI355X, 2^20 x 256 at blowup 2: 25.4 ms
    // from page-locked and 25.3 ms from pageable memory with 64-column
    // blocks, 26.8 / 28.6 ms with 128-column blocks written by pitched DMAs.
    constexpr size_t SPLIT_BYTES = 64ull << 20;
    const size_t block_bytes = ctx->host_block_bytes.load(), min_cols = ctx->host_block_min_cols.load();
    size_t Hmax = 0, segments = 0;
    for (size_t k = 0; k < n; k++) Hmax = std::max(Hmax, rows[k]);
    std::vector<bool> locked(n, false);
    for (size_t k = 0; k < n; k++) locked[k] = host_is_locked(host[k]);
    auto block_cols = [&](size_t k) {
        if (canonical || rows[k] * cols[k] * 4 < SPLIT_BYTES) return cols[k];
        const size_t cb = std::max<size_t>(min_cols, block_bytes / (rows[k] * 4) / 8 * 8);
        return std::min(cb, cols[k]);
    };
    for (size_t k = 0; k < n; k++)
        if (rows[k] == Hmax) segments += (cols[k] + block_cols(k) - 1) / block_cols(k);
    std::shared_ptr<DevMem> state;
    if (segments > 1) {
        state = make_devmem(ctx, (Hmax << log_blowup) * 64, &rc);
        if (!state) {
            cleanup();
            return rc;
        }
    }
    // block (k, c0): dense at word rows * c0 of in[k] (the whole matrix, in
    // its own layout, when it is not split)
    auto block_ptr = [&](size_t k, size_t c0) { return in[k]->ptr + rows[k] * c0; };
    auto upload_wait = [&](size_t k, size_t c0, size_t cb) -> int {
        if (cb != cols[k]) {  // row segments gathered into a stage
            PG_TRY(stage_upload(ctx, lc.s, block_ptr(k, c0), cb * 4, host[k] + c0, cols[k] * 4, cb * 4, rows[k],
                                &run));
        } else if (!locked[k] && in[k]->bytes >= STAGE_MIN_BYTES) {  // pageable: through a stage
            PG_TRY(stage_upload(ctx, lc.s, in[k]->ptr, in[k]->bytes, host[k], in[k]->bytes, in[k]->bytes, 1, &run));
        } else {
            HIP_TRY(hipMemcpyAsync(in[k]->ptr, host[k], in[k]->bytes, hipMemcpyHostToDevice, lc.s));
        }
        if (canonical) HIP_TRY(launch_convert(field, block_ptr(k, c0), rows[k] * cb, 1, lc.s));
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        events.push_back(e);
        HIP_TRY(hipEventRecord(e, lc.s));
        HIP_TRY(hipStreamWaitEvent(l.s, e, 0));
        return PG_OK;
    };
    auto lde_cols = [&](size_t k, size_t c0, size_t cb) -> hipError_t {
        return launch_coset_lde(field, block_ptr(k, c0), cb, L[k], (uint32_t)cb, (int)log_blowup,
                                (const uint32_t*)tw[k]->ptr, (const uint32_t*)sc[k]->ptr,
                                L[k] > 11 ? (uint32_t*)scratch->ptr : nullptr, ldes[k]->ptr + c0, cols[k], 1, l.s,
                                kt.get());
    };
    // The matrices below the tallest height are uploaded, extended and
    // row-hashed first: the upload of the tallest ones bounds the call, and
    // the GPU has time to spare meanwhile.  Left to their level, their
    // uploads, LDEs and the hashing of their rows (2 + w/8 permutations per
    // node at the injection layers) would all run after that upload.  Their
    // digests (32 bytes per LDE row) wait in HBM for their layer.
    std::map<size_t, std::shared_ptr<DevMem>> run_digests;  // run start -> digests of its rows
    auto shorter_first = [&](const std::vector<size_t>& order, size_t g1, const MatRef* sorted) -> int {
        for (size_t a = g1, b; a < order.size(); a = b) {
            uint64_t wsum = 0;
            for (b = a; b < order.size() && rows[order[b]] == rows[order[a]]; b++) {
                const size_t k = order[b];
                PG_TRY(upload_wait(k, 0, cols[k]));
                HIP_TRY(lde_cols(k, 0, cols[k]));
                wsum += cols[k];
            }
            if (wsum > 0xffffffffull) return fail(PG_E_INVALID, "total width too large");
            const size_t nl = rows[order[a]] << log_blowup;
            int mrc = PG_OK;
            auto d = make_devmem(ctx, nl * 32, &mrc);
            if (!d) return mrc;
            TIMED(kt, "leaf_hash", l.s,
                  launch_leaf_hash(field, sorted + a, (int)(b - a), (uint32_t)wsum, nl, (uint32_t*)d->ptr, l.s));
            run_digests[a] = d;
        }
        return PG_OK;
    };
    StageFn stage = [&](const std::vector<size_t>& order, size_t g0, size_t g1, uint32_t* leaves,
                        const MatRef* sorted, const uint32_t** inj) -> int {
        if (leaves) PG_TRY(shorter_first(order, g1, sorted));
        auto it = run_digests.find(g0);
        if (!leaves && it != run_digests.end()) {
            *inj = (const uint32_t*)it->second->ptr;
            return PG_OK;
        }
        if (leaves && segments > 1) {
            uint64_t wsum = 0, done = 0;
            for (size_t j = g0; j < g1; j++) wsum += cols[order[j]];
            const size_t N = Hmax << log_blowup;
            // (matrix, first column, width) of every block, in upload order
            std::vector<std::array<size_t, 3>> blocks;
            for (size_t j = g0; j < g1; j++) {
                const size_t k = order[j], bc = block_cols(k);
                for (size_t c0 = 0; c0 < cols[k]; c0 += bc) blocks.push_back({k, c0, std::min(bc, cols[k] - c0)});
            }
            // The last block's LDE and absorb run after the last upload; a
            // block of at most block_bytes is cut into halves (32 + 16 + 16
            // columns of 64) so that only a narrow piece is left behind it
            // (2^20 x 256: 25.0 -> 24.4 ms).  Taller blocks stay whole: the
            // narrow pieces' row gathers and extra passes over the sponge
            // states cost more there than the shorter tail saves (the shard,
            // 2^22 rows: 342 -> 348 ms).
            if (blocks.size() > 1 &&
                blocks.back()[2] * rows[blocks.back()[0]] * 4 <= block_bytes) {
                const std::array<size_t, 3> last = blocks.back();
                blocks.pop_back();
                size_t c0 = last[1], rem = last[2];
                while (rem > 16) {
                    const size_t w = std::max<size_t>(8, rem / 2 / 8 * 8);
                    blocks.push_back({last[0], c0, w});
                    c0 += w;
                    rem -= w;
                }
                blocks.push_back({last[0], c0, rem});
            }
            for (const auto& bl : blocks) {
                const size_t k = bl[0], c0 = bl[1], cb = bl[2];
                PG_TRY(upload_wait(k, c0, cb));
                HIP_TRY(lde_cols(k, c0, cb));
                TIMED(kt, "absorb_segment", l.s,
                      launch_absorb_segment(field, ldes[k]->ptr + c0, cols[k], (uint32_t)cb, (uint32_t)(done % 8),
                                            done == 0, done + cb == wsum, N, (uint32_t*)state->ptr, leaves, l.s));
                done += cb;
            }
            return STAGE_HASHED;
        }
        // one event per matrix: its LDE starts as soon as its own upload is
        // done, overlapping the uploads of the rest of its group
        for (size_t j = g0; j < g1; j++) {
            const size_t k = order[j];
            PG_TRY(upload_wait(k, 0, cols[k]));
            HIP_TRY(lde_cols(k, 0, cols[k]));
        }
        return PG_OK;
    };
    rc = mmcs_commit_on_stream(ctx, l.s, field, ldes, std::vector<bool>(n, true), root, out, stage);
    committed = rc == PG_OK;  // the tree owns the LDEs from here
    cleanup();
    return rc;
}

extern "C" int pg_pcs_evaluations(pg_ctx* ctx, const pg_tree* t, size_t k, size_t rows, uint32_t* host) {
    if (!ctx || !t || !host || k >= t->mats.size()) return fail(PG_E_INVALID, "bad argument");
    const pg_buf* m = t->mats[k];
    if (!is_pow2(rows) || rows > m->rows) return fail(PG_E_INVALID, "domain size %zu invalid", rows);
    // natural order gathered on the device a block of rows at a time (at most
    // 256 MiB of scratch), each block downloaded as one contiguous run
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const size_t w = m->cols, q = (size_t)log2_of(rows);
    if (w == 0 || rows == 0) return PG_OK;
    const size_t block = std::max<size_t>(1, std::min(rows, (size_t)(256u << 20) / (w * 4)));
    void* tmp;
    PG_TRY(ctx_alloc(ctx, block * w * 4, &tmp));
    int rc = PG_OK;
    for (size_t r0 = 0; r0 < rows && rc == PG_OK; r0 += block) {
        const size_t nr = std::min(block, rows - r0);
        hipError_t e = launch_rows_bitrev(m->ptr, m->ld, (uint32_t)w, (int)q, r0, nr, (uint32_t*)tmp, l.s);
        if (e != hipSuccess) {
            rc = fail(PG_E_DEVICE, "row gather: %s", hipGetErrorString(e));
            break;
        }
        uint32_t* dst = host + r0 * w;
        if (nr * w * 4 >= STAGE_MIN_BYTES && !host_is_locked(host)) {
            rc = stage_download(ctx, l.s, dst, tmp, nr * w * 4);
        } else {
            e = hipMemcpyAsync(dst, tmp, nr * w * 4, hipMemcpyDeviceToHost, l.s);
            if (e == hipSuccess) e = hipStreamSynchronize(l.s);
            if (e != hipSuccess) rc = fail(PG_E_DEVICE, "download: %s", hipGetErrorString(e));
        }
    }
    (void)hipStreamSynchronize(l.s);
    ctx_free(ctx, tmp, block * w * 4);
    return rc;
}

// ------------------------------------------------------------ PCS open --
// Device part of TwoAdicFriPcs::open; the caller's challenger observes the
// values in between (circuit/fri.rs:84-99 is the verifier's order).

template <class F>
static Ext4 ext_from(const uint32_t* w) {
    return Ext4{{w[0], w[1], w[2], w[3]}};
}

// factor = (z^h - s^h) / (h s^(h-1)) of the barycentric form on s <w_h>
template <class F>
static Ext4 bary_factor(const Ext4& z, uint32_t s_mont, int log_h) {
    const uint64_t h = 1ull << log_h;
    Ext4 zh = epow<F>(z, h);
    zh.c[0] = fsub<F>(zh.c[0], fpow<F>(s_mont, h));
    const uint32_t hm = to_monty<F>((uint32_t)(h % F::P));
    const uint32_t den = fmul<F>(hm, fpow<F>(s_mont, h - 1));
    return escale<F>(zh, finv<F>(den));
}

// Barycentric weights depend only on the trace height and the points: the
// matrices of one Pcs::open that share both (every matrix of a round's
// height, opened at zeta and its own next point) share one computation.
// Key: log height, point count, the points' words.
struct WeightCache {
    std::map<std::vector<uint32_t>, std::shared_ptr<DevMem>> bufs;
};

template <class F>
static int open_values_t(pg_ctx* ctx, hipStream_t s, const pg_tree* t, size_t k, uint32_t log_blowup,
                         const uint32_t* points, size_t npts, uint32_t* out, WeightCache* wc = nullptr,
                         const FixedScratch* fixed = nullptr) {
    const pg_buf* m = t->mats[k];
    if (!is_pow2(m->rows) || (m->rows >> log_blowup) == 0 || log2_of(m->rows) < (int)log_blowup)
        return fail(PG_E_INVALID, "matrix %zu: height %zu has no trace coset at log_blowup %u", k, m->rows,
                    log_blowup);
    const int log_h = log2_of(m->rows) - (int)log_blowup;
    const size_t h = (size_t)1 << log_h, w = m->cols;
    if (npts == 0 || w == 0) return PG_OK;
    // In a fixed region the twiddle table is built in the region itself, two
    // points go per launch (weights of two points at a time), and chunks are
    // made taller until the partial sums fit.  Any chunking gives the same
    // values: every sum is exact mod p.
    const size_t per = fixed ? 2 : 4;  // points per launch
    const size_t q_max = std::min(npts, per);
    // about 2^17 (chunk, column) threads, at least 16 rows per chunk: taller
    // chunks write fewer partial sums (2^20 x 256 at two points: 0.29 ms
    // against 0.30 with 2^18 threads, 0.34 with 2^20, 0.39 with 2^16)
    const uint64_t col_threads = open_values_vec4(m->ptr, m->cols, (uint32_t)w, (int)q_max) ? w / 4 : w;
    uint32_t chunk_rows = 16;
    while ((uint64_t)chunk_rows * (1u << 17) < (uint64_t)h * col_threads) chunk_rows <<= 1;
    struct Layout {
        uint32_t nchunks, nslices;
        size_t b_t, b_w, b_p, b_s, b_o;  // twiddles, weights, partials, slice sums, output
        size_t total() const { return b_t + b_w + b_p + b_s + b_o; }
    };
    auto layout = [&](uint32_t cr) {
        Layout y;
        y.nchunks = (uint32_t)((h + cr - 1) / cr);
        // column sums in slices of about 64 chunks (at most 64 slices)
        y.nslices = std::max<uint32_t>(1, std::min<uint32_t>(64, y.nchunks / 64));
        y.b_t = fixed ? round_up(4 * h, 256) : 0;
        y.b_w = wc ? 0 : round_up(q_max * h * 16, 256);
        y.b_p = round_up(q_max * y.nchunks * w * 16, 256);
        y.b_s = round_up(q_max * y.nslices * w * 16, 256);
        y.b_o = npts * w * 16;
        return y;
    };
    Layout y = layout(chunk_rows);
    if (fixed) {
        while (y.total() > fixed->bytes && chunk_rows < h) y = layout(chunk_rows <<= 1);
        if (y.total() > fixed->bytes)
            return fail(PG_E_INVALID,
                        "matrix %zu (trace height 2^%d, width %zu, %zu points) needs %zu bytes of scratch; the "
                        "allocation-free form has %zu",
                        k, log_h, w, npts, y.total(), fixed->bytes);
    }
    std::shared_ptr<DevMem> tw, scratch;
    uint8_t* base;
    const uint32_t* twp;
    if (fixed) {
        base = fixed->base;
        twp = (const uint32_t*)base;
    } else {
        PG_TRY(get_twiddles(ctx, s, t->field, log_h, &tw));
        twp = (const uint32_t*)tw->ptr;
        int rc;
        scratch = make_devmem(ctx, y.total(), &rc);
        if (!scratch) return rc;
        base = (uint8_t*)scratch->ptr;
    }
    SyncGuard guard(s);  // an early return waits for queued kernels before the scratch is released
    if (fixed) HIP_TRY(launch_twiddles(t->field, log_h, (uint32_t*)base, s));
    uint8_t* const wts = base + y.b_t;
    uint8_t* const partial = wts + y.b_w;
    uint8_t* const sums = partial + y.b_p;
    uint8_t* const outs = sums + y.b_s;
    const uint32_t s_mont = to_monty<F>(F::GEN);
    std::unique_ptr<EventTimer> kt = fixed ? nullptr : make_timer(ctx);
    for (size_t q0 = 0; q0 < npts; q0 += per) {
        const int nq = (int)std::min(per, npts - q0);
        Ext4 z[4], fac[4];
        for (int q = 0; q < nq; q++) {
            z[q] = ext_from<F>(points + 4 * (q0 + q));
            fac[q] = bary_factor<F>(z[q], s_mont, log_h);
        }
        Ext4* weights = (Ext4*)wts;
        int ready = 0;
        if (wc) {
            std::vector<uint32_t> key{(uint32_t)log_h, (uint32_t)nq};
            for (int q = 0; q < nq; q++) key.insert(key.end(), z[q].c, z[q].c + 4);
            auto it = wc->bufs.find(key);
            if (it != wc->bufs.end()) {
                weights = (Ext4*)it->second->ptr;
                ready = 1;
            } else {
                int rc;
                auto wb = make_devmem(ctx, (size_t)nq * h * 16, &rc);
                if (!wb) return rc;
                weights = (Ext4*)wb->ptr;
                wc->bufs.emplace(std::move(key), std::move(wb));
            }
        }
        TIMED(kt, "open_values", s,
              launch_open_values(t->field, m->ptr, m->cols, (uint32_t)w, log_h, twp, z, fac, nq, weights, ready,
                                 (Ext4*)partial, y.nchunks, chunk_rows, (Ext4*)sums, y.nslices,
                                 (Ext4*)(outs + q0 * w * 16), s));
    }
    HIP_TRY(hipMemcpyAsync(out, outs, y.b_o, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    guard.armed = false;
    return PG_OK;
}

static int open_values_cached(pg_ctx* ctx, const pg_tree* t, size_t k, uint32_t log_blowup,
                              const uint32_t* points_mont, size_t npts, uint32_t* out_mont, WeightCache* wc) {
    if (!ctx || !t || (npts && (!points_mont || !out_mont)) || k >= t->mats.size())
        return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    return t->field == PG_KOALABEAR
               ? open_values_t<KoalaBear>(ctx, l.s, t, k, log_blowup, points_mont, npts, out_mont, wc)
               : open_values_t<BabyBear>(ctx, l.s, t, k, log_blowup, points_mont, npts, out_mont, wc);
}

extern "C" int pg_pcs_open_values(pg_ctx* ctx, const pg_tree* t, size_t k, uint32_t log_blowup,
                                  const uint32_t* points_mont, size_t npts, uint32_t* out_mont) {
    return open_values_cached(ctx, t, k, log_blowup, points_mont, npts, out_mont, nullptr);
}

// The same values with the device's static scratch array instead of
// allocated memory (twiddles, weights, partial and slice sums, output
// staging): no hipMalloc, hipFree or pool allocation.
extern "C" int pg_pcs_open_values_noalloc(pg_ctx* ctx, const pg_tree* t, size_t k, uint32_t log_blowup,
                                          const uint32_t* points_mont, size_t npts, uint32_t* out_mont) {
    if (!ctx || !t || (npts && (!points_mont || !out_mont)) || k >= t->mats.size())
        return fail(PG_E_INVALID, "bad argument");
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return fail(PG_E_UNSUPPORTED, "device %d has no allocation-free scratch", ctx->device);
    PG_TRY(bind_device(ctx));
    NoallocSlot& sl = g_noalloc[ctx->device];
    std::lock_guard<std::mutex> g(sl.mu);
    if (!sl.ptr) return fail(PG_E_DEVICE, "device %d: the scratch address was not looked up", ctx->device);
    const FixedScratch fx{sl.ptr, NOALLOC_SCRATCH_BYTES};
    Lease l(ctx);
    return t->field == PG_KOALABEAR
               ? open_values_t<KoalaBear>(ctx, l.s, t, k, log_blowup, points_mont, npts, out_mont, nullptr, &fx)
               : open_values_t<BabyBear>(ctx, l.s, t, k, log_blowup, points_mont, npts, out_mont, nullptr, &fx);
}

template <class F>
static int reduce_openings_t(pg_ctx* ctx, hipStream_t s, const pg_tree* const* rounds, size_t nrounds,
                             const uint32_t* npoints, const uint32_t* points, const uint32_t* values,
                             const uint32_t* alpha_w, pg_buf** out_by_log) {
    const int field = rounds[0]->field;
    const Ext4 alpha = ext_from<F>(alpha_w);
    // walk the batches in round / matrix / point order, as the verifier does
    struct Item {
        const pg_buf* m;
        int lh;
        RoMat rm;
        Ext4 z[2];
    };
    std::vector<Item> items;
    size_t maxw = 0, pi = 0, vi = 0, mi = 0;
    uint64_t reduced[64] = {0};
    for (size_t r = 0; r < nrounds; r++) {
        const pg_tree* t = rounds[r];
        if (!t || t->field != field) return fail(PG_E_INVALID, "round %zu: missing tree or mixed fields", r);
        for (size_t k = 0; k < t->mats.size(); k++, mi++) {
            const pg_buf* m = t->mats[k];
            const uint32_t np = npoints[mi];
            if (np > 2) return fail(PG_E_UNSUPPORTED, "more than 2 points on one matrix");
            if (!is_pow2(m->rows)) return fail(PG_E_INVALID, "height %zu is not a power of two", m->rows);
            maxw = std::max(maxw, m->cols);
            Item it{};
            it.m = m;
            it.lh = log2_of(m->rows);
            it.rm.data = m->ptr;
            it.rm.ld = m->cols;
            it.rm.width = (uint32_t)m->cols;
            it.rm.vec4 = (m->cols % 4 == 0) && ((uintptr_t)m->ptr % 16 == 0);
            it.rm.npts = (int)np;
            for (uint32_t q = 0; q < np; q++, pi++) {
                it.z[q] = ext_from<F>(points + 4 * pi);
                // alpha^off and alpha^off * sum_c alpha^c p_c(z)
                const Ext4 aoff = epow<F>(alpha, reduced[it.lh]);
                Ext4 acc{{0, 0, 0, 0}}, ap{{F::R1, 0, 0, 0}};
                for (size_t c = 0; c < m->cols; c++, vi++) {
                    acc = eadd<F>(acc, emul<F>(ap, ext_from<F>(values + 4 * vi)));
                    ap = emul<F>(ap, alpha);
                }
                it.rm.aoff[q] = aoff;
                it.rm.a[q] = emul<F>(aoff, acc);
                reduced[it.lh] += m->cols;
            }
            if (np) items.push_back(it);
        }
    }
    for (int lh = 0; lh < 33; lh++) out_by_log[lh] = nullptr;
    // alpha^c for c < the widest matrix
    std::vector<Ext4> apows(std::max<size_t>(maxw, 1));
    apows[0] = Ext4{{F::R1, 0, 0, 0}};
    for (size_t c = 1; c < apows.size(); c++) apows[c] = emul<F>(apows[c - 1], alpha);
    std::vector<int> heights;
    for (auto& it : items)
        if (std::find(heights.begin(), heights.end(), it.lh) == heights.end()) heights.push_back(it.lh);
    std::sort(heights.rbegin(), heights.rend());
    // one device block: alpha powers, then per height the point and matrix tables
    size_t bytes = align_up(apows.size() * 16, 256);
    for (int lh : heights) {
        size_t nm = 0;
        for (auto& it : items) nm += it.lh == lh;
        bytes += align_up(RO_MAX_POINTS * 16, 256) + align_up(nm * sizeof(RoMat), 256);
    }
    std::vector<uint8_t> host(bytes, 0);
    memcpy(host.data(), apows.data(), apows.size() * 16);
    struct Launch {
        int lh, nm, np;
        size_t off_pts, off_mats;
    };
    std::vector<Launch> launches;
    size_t off = align_up(apows.size() * 16, 256);
    for (int lh : heights) {
        std::vector<Ext4> pts;
        std::vector<RoMat> rms;
        for (auto& it : items) {
            if (it.lh != lh) continue;
            RoMat rm = it.rm;
            for (int q = 0; q < rm.npts; q++) {
                int found = -1;
                for (size_t u = 0; u < pts.size(); u++)
                    if (!memcmp(pts[u].c, it.z[q].c, 16)) found = (int)u;
                if (found < 0) {
                    if (pts.size() == RO_MAX_POINTS)
                        return fail(PG_E_UNSUPPORTED, "more than %d distinct points at one height", RO_MAX_POINTS);
                    pts.push_back(it.z[q]);
                    found = (int)pts.size() - 1;
                }
                rm.pt[q] = found;
            }
            rms.push_back(rm);
        }
        Launch la{lh, (int)rms.size(), (int)pts.size(), off, off + align_up(RO_MAX_POINTS * 16, 256)};
        memcpy(host.data() + la.off_pts, pts.data(), pts.size() * 16);
        memcpy(host.data() + la.off_mats, rms.data(), rms.size() * sizeof(RoMat));
        off = la.off_mats + align_up(rms.size() * sizeof(RoMat), 256);
        launches.push_back(la);
    }
    int rc;
    auto tables = make_devmem(ctx, bytes, &rc);
    if (!tables) return rc;
    uint8_t* dt = (uint8_t*)tables->ptr;
    HIP_TRY(hipMemcpyAsync(dt, host.data(), bytes, hipMemcpyHostToDevice, s));
    // every table and output is in place before the first launch, so a
    // failed allocation never frees memory a queued kernel still uses
    std::vector<pg_buf*> made;
    std::vector<std::shared_ptr<DevMem>> keep;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        for (auto* b : made) pg_buf_free(b);
        for (int lh = 0; lh < 33; lh++) out_by_log[lh] = nullptr;
    };
    for (auto& la : launches) {
        std::shared_ptr<DevMem> tw;
        rc = get_twiddles(ctx, s, field, la.lh, &tw);
        if (rc == PG_OK) {
            keep.push_back(tw);
            pg_buf* o;
            rc = buf_new(ctx, (size_t)1 << la.lh, 4, &o);
            if (rc == PG_OK) made.push_back(o);
        }
        if (rc != PG_OK) {
            cleanup();
            return rc;
        }
    }
    auto kt = make_timer(ctx);
    for (size_t i = 0; i < launches.size(); i++) {
        const Launch& la = launches[i];
        if (kt) kt->begin("reduce_openings", s);
        hipError_t e = launch_reduce_rows(field, (const RoMat*)(dt + la.off_mats), la.nm, (const Ext4*)(dt + la.off_pts),
                                          la.np, (const Ext4*)dt, la.lh, (const uint32_t*)keep[i]->ptr,
                                          (Ext4*)made[i]->ptr, s);
        if (kt) kt->end(s);
        if (e != hipSuccess) {
            cleanup();
            return fail(PG_E_DEVICE, "reduce_rows: %s", hipGetErrorString(e));
        }
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        cleanup();
        return fail(PG_E_DEVICE, "reduce_openings: %s", hipGetErrorString(e));
    }
    for (size_t i = 0; i < launches.size(); i++) out_by_log[launches[i].lh] = made[i];
    return PG_OK;
}

extern "C" int pg_pcs_reduce_openings(pg_ctx* ctx, const pg_tree* const* rounds, size_t nrounds,
                                      const uint32_t* npoints, const uint32_t* points_mont,
                                      const uint32_t* values_mont, const uint32_t alpha_mont[4],
                                      pg_buf** out_by_log) {
    if (!ctx || !rounds || nrounds == 0 || !npoints || !alpha_mont || !out_by_log || !rounds[0])
        return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    return rounds[0]->field == PG_KOALABEAR
               ? reduce_openings_t<KoalaBear>(ctx, l.s, rounds, nrounds, npoints, points_mont, values_mont,
                                              alpha_mont, out_by_log)
               : reduce_openings_t<BabyBear>(ctx, l.s, rounds, nrounds, npoints, points_mont, values_mont,
                                             alpha_mont, out_by_log);
}

// ----------------------------------------------------------------- batch --
// Many independent single-matrix PCS commits as one launch sequence: each
// LDE pass runs once per distinct log-height over all jobs of that height,
// the leaf layer runs once over every job's rows, each Merkle level once over
// every job that has it, and a final kernel gathers the roots.  All outputs
// and descriptors are allocated up front; the sequence can be captured as a
// hipGraph and replayed.
struct pg_batch {
    pg_ctx* ctx = nullptr;
    int field = 0;
    size_t njobs = 0;
    std::vector<pg_tree*> trees;
    std::vector<bool> taken;
    std::vector<std::shared_ptr<DevMem>> keep;  // tables, scratch, descriptors
    uint32_t* d_roots = nullptr;
    std::vector<std::function<hipError_t(hipStream_t)>> launches;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" void pg_batch_free(pg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (size_t j = 0; j < b->trees.size(); j++)
        if (b->trees[j] && !b->taken[j]) pg_tree_free(b->trees[j]);
    b->keep.clear();
    delete b;
}

// host staging for one device descriptor block
extern "C" int pg_batch_create(pg_ctx* ctx, int field, const pg_buf* const* evals, const uint32_t* dshift,
                               const uint32_t* log_blowups, size_t njobs, int use_graph, pg_batch** out) {
    if (!ctx || !evals || !log_blowups || !out || !field_ok(field) || njobs == 0)
        return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    std::unique_ptr<pg_batch, void (*)(pg_batch*)> b(new pg_batch, pg_batch_free);
    b->ctx = ctx;
    b->field = field;
    b->njobs = njobs;
    b->trees.assign(njobs, nullptr);
    b->taken.assign(njobs, false);
    const uint32_t g = gen_monty(field);
    std::vector<int> L(njobs), lb(njobs);
    std::vector<uint64_t> N(njobs), W(njobs), tmp_off(njobs, 0);
    uint64_t tmp_words = 0;
    for (size_t j = 0; j < njobs; j++) {
        const pg_buf* m = evals[j];
        if (!m || !is_pow2(m->rows) || m->cols == 0) return fail(PG_E_INVALID, "job %zu: bad matrix", j);
        L[j] = log2_of(m->rows);
        lb[j] = (int)log_blowups[j];
        if (L[j] > lde_max_log_h()) return fail(PG_E_UNSUPPORTED, "job %zu: log height %d", j, L[j]);
        if (L[j] + lb[j] > two_adicity(field) || lb[j] > 8) return fail(PG_E_INVALID, "job %zu: blowup", j);
        N[j] = m->rows << lb[j];
        W[j] = m->cols;
        if (L[j] > 11) {
            tmp_off[j] = tmp_words;
            tmp_words += N[j] * W[j];
        }
    }
    // outputs: one LDE buffer and one tree per job
    std::vector<uint32_t*> digests(njobs);
    for (size_t j = 0; j < njobs; j++) {
        pg_buf* o;
        PG_TRY(buf_new(ctx, N[j], W[j], &o));
        TreePtr t(nullptr, tree_discard);
        std::vector<size_t> order;
        MatRef* ds;
        const int rc = tree_prepare(ctx, l.s, field, {o}, &t, &order, &ds);
        if (rc != PG_OK) {
            pg_buf_free(o);
            return rc;
        }
        t->owned = {true};
        digests[j] = t->digests;
        b->trees[j] = t.release();
    }
    std::shared_ptr<DevMem> scratch;
    if (tmp_words) {
        int rc;
        scratch = make_devmem(ctx, tmp_words * 4, &rc);
        if (!scratch) return rc;
        b->keep.push_back(scratch);
    }
    // descriptors
    Staging st;
    struct PassRec {
        int kind, log_s, variant, njobs, max_b;
        size_t off_args, off_prefix;
        uint64_t total;
    };
    std::vector<PassRec> passes;
    std::map<int, std::vector<size_t>> by_L;
    for (size_t j = 0; j < njobs; j++) by_L[L[j]].push_back(j);
    for (auto& kv : by_L) {
        const int Lg = kv.first;
        int k1, k2;
        lde_split(Lg, &k1, &k2);
        std::shared_ptr<DevMem> tw;
        PG_TRY(get_twiddles(ctx, l.s, field, Lg, &tw));
        b->keep.push_back(tw);
        std::vector<LdeArgs> args;
        for (size_t j : kv.second) {
            uint32_t shift = g;
            if (dshift) {
                if (dshift[j] == 0) return fail(PG_E_INVALID, "job %zu: domain shift is zero", j);
                shift = field == PG_KOALABEAR ? fmul<KoalaBear>(g, finv<KoalaBear>(dshift[j]))
                                              : fmul<BabyBear>(g, finv<BabyBear>(dshift[j]));
            }
            std::shared_ptr<DevMem> sc;
            PG_TRY(get_scale(ctx, l.s, field, Lg, lb[j], shift, &sc));
            b->keep.push_back(sc);
            LdeArgs a{};
            a.in = evals[j]->ptr;
            a.ld_in = evals[j]->ld;
            a.out = b->trees[j]->mats[0]->ptr;
            a.ld_out = W[j];
            a.tmp = scratch && Lg > 11 ? (uint32_t*)scratch->ptr + tmp_off[j] : nullptr;
            a.tw = (const uint32_t*)tw->ptr;
            a.scale = (const uint32_t*)sc->ptr;
            a.w = (uint32_t)W[j];
            a.L = Lg;
            a.k1 = k1;
            a.k2 = k2;
            a.b = lb[j];
            a.bitrev = 1;
            args.push_back(a);
        }
        const int kinds3[3] = {KIND_A, KIND_BC, KIND_D};
        for (int i = 0; i < (Lg <= 11 ? 1 : 3); i++) {
            const int kind = Lg <= 11 ? KIND_SINGLE : kinds3[i];
            const int log_s = (kind == KIND_BC || kind == KIND_SINGLE) ? k2 : k1;
            uint32_t wmax = 0;  // one tile shape per launch: the widest job's
            for (auto& a : args) wmax = std::max(wmax, a.w);
            const int variant = lde_variant(field, kind, log_s, wmax);
            const int C = lde_pass_cols(kind, log_s, variant);
            std::vector<uint64_t> prefix(1, 0);
            for (auto& a : args)
                prefix.push_back(prefix.back() + lde_pass_row_tiles(kind, Lg, a.b) * ((a.w + C - 1) / C));
            int max_b = 0;
            for (auto& a : args) max_b = std::max(max_b, a.b);
            PassRec pr{kind, log_s, variant, (int)args.size(), max_b, st.put(args.data(), args.size() * sizeof(LdeArgs)),
                       st.put(prefix.data(), prefix.size() * 8), prefix.back()};
            passes.push_back(pr);
        }
    }
    // Merkle: leaf layer, then levels over the jobs (tallest first) that have them
    std::vector<size_t> tall(njobs);
    for (size_t j = 0; j < njobs; j++) tall[j] = j;
    std::stable_sort(tall.begin(), tall.end(), [&](size_t a, size_t c) { return N[a] > N[c]; });
    std::vector<LeafJob> leaves;
    std::vector<uint64_t> leaf_prefix(1, 0);
    for (size_t j : tall) {
        leaves.push_back(LeafJob{b->trees[j]->mats[0]->ptr, W[j], (uint32_t)W[j], 0, digests[j]});
        leaf_prefix.push_back(leaf_prefix.back() + N[j]);
    }
    const size_t off_leaf = st.put(leaves.data(), leaves.size() * sizeof(LeafJob));
    const size_t off_leaf_prefix = st.put(leaf_prefix.data(), leaf_prefix.size() * 8);
    struct LevelRec {
        size_t off_jobs, off_prefix;
        int njobs;
        uint64_t total;
    };
    std::vector<LevelRec> levels;
    const int max_log = log2_of(N[tall[0]]);
    for (int lv = 1; lv <= max_log; lv++) {
        std::vector<CompressJob> cj;
        std::vector<uint64_t> pf(1, 0);
        for (size_t j : tall) {
            const int lj = log2_of(N[j]);
            if (lj < lv) break;
            uint64_t off_prev = 0;
            for (int t = 0; t < lv - 1; t++) off_prev += N[j] >> t;
            const uint64_t off_next = off_prev + (N[j] >> (lv - 1));
            cj.push_back(CompressJob{digests[j] + 8 * off_prev, digests[j] + 8 * off_next});
            pf.push_back(pf.back() + (N[j] >> lv));
        }
        levels.push_back(LevelRec{st.put(cj.data(), cj.size() * sizeof(CompressJob)), st.put(pf.data(), pf.size() * 8),
                                  (int)cj.size(), pf.back()});
    }
    std::vector<const uint32_t*> root_src(njobs);
    for (size_t j = 0; j < njobs; j++) root_src[j] = digests[j] + 8 * (2 * N[j] - 2);
    const size_t off_root_src = st.put(root_src.data(), njobs * sizeof(void*));
    const size_t off_roots = st.put(nullptr, njobs * 32);
    int rc;
    auto desc = make_devmem(ctx, st.bytes.size(), &rc);
    if (!desc) return rc;
    b->keep.push_back(desc);
    uint8_t* dd = (uint8_t*)desc->ptr;
    HIP_TRY(hipMemcpyAsync(dd, st.bytes.data(), st.bytes.size(), hipMemcpyHostToDevice, l.s));
    HIP_TRY(hipStreamSynchronize(l.s));
    b->d_roots = (uint32_t*)(dd + off_roots);
    // launch list
    for (const auto& pr : passes) {
        const LdeArgs* jobs = (const LdeArgs*)(dd + pr.off_args);
        const uint64_t* pf = (const uint64_t*)(dd + pr.off_prefix);
        b->launches.push_back([=](hipStream_t s) {
            return launch_lde_pass_batch(field, pr.kind, pr.log_s, pr.variant, jobs, pf, pr.njobs, pr.total, pr.max_b,
                                         s);
        });
    }
    {
        const LeafJob* lj = (const LeafJob*)(dd + off_leaf);
        const uint64_t* lp = (const uint64_t*)(dd + off_leaf_prefix);
        const uint64_t total = leaf_prefix.back();
        const int nj = (int)njobs;
        b->launches.push_back([=](hipStream_t s) { return launch_leaf_hash_batch(field, lj, lp, nj, total, s); });
    }
    for (const auto& lv : levels) {
        const CompressJob* cj = (const CompressJob*)(dd + lv.off_jobs);
        const uint64_t* pf = (const uint64_t*)(dd + lv.off_prefix);
        b->launches.push_back([=](hipStream_t s) { return launch_compress_batch(field, cj, pf, lv.njobs, lv.total, s); });
    }
    {
        const uint32_t* const* rs = (const uint32_t* const*)(dd + off_root_src);
        uint32_t* dst = b->d_roots;
        const int nj = (int)njobs;
        b->launches.push_back([=](hipStream_t s) { return launch_gather_roots(rs, dst, nj, s); });
    }
    if (use_graph) {
        // Captured on a stream of its own: the context's streams are shared
        // round-robin by concurrent calls, whose work must not be recorded
        // into this graph.
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (auto& fn : b->launches)
                if (e == hipSuccess) e = fn(cap);
            const hipError_t e2 = hipStreamEndCapture(cap, &b->graph);
            if (e == hipSuccess) e = e2;
        }
        (void)hipStreamDestroy(cap);
        HIP_TRY(e);
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    *out = b.release();
    return PG_OK;
}

extern "C" int pg_batch_run(pg_ctx* ctx, pg_batch* b, uint32_t* roots_out) {
    if (!ctx || !b || b->ctx != ctx) return fail(PG_E_INVALID, "bad argument");
    for (size_t j = 0; j < b->njobs; j++)  // its buffers now belong to the caller
        if (b->taken[j]) return fail(PG_E_INVALID, "tree %zu was taken; the batch cannot run again", j);
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    if (b->exec) {
        HIP_TRY(hipGraphLaunch(b->exec, l.s));
    } else {
        for (auto& fn : b->launches) HIP_TRY(fn(l.s));
    }
    std::vector<uint32_t> roots(b->njobs * 8);
    HIP_TRY(hipMemcpyAsync(roots.data(), b->d_roots, roots.size() * 4, hipMemcpyDeviceToHost, l.s));
    HIP_TRY(hipStreamSynchronize(l.s));
    for (size_t j = 0; j < b->njobs; j++)
        if (b->trees[j]) memcpy(b->trees[j]->root, &roots[8 * j], 32);
    if (roots_out) memcpy(roots_out, roots.data(), roots.size() * 4);
    return PG_OK;
}

extern "C" int pg_batch_take_tree(pg_batch* b, size_t j, pg_tree** out) {
    if (!b || !out || j >= b->njobs) return fail(PG_E_INVALID, "bad argument");
    if (b->taken[j]) return fail(PG_E_INVALID, "tree %zu already taken", j);
    b->taken[j] = true;
    *out = b->trees[j];
    return PG_OK;
}

extern "C" size_t pg_batch_num_launches(const pg_batch* b) { return b ? b->launches.size() : 0; }

// --------------------------------------------------------------- service --
// Device-resident commit service.  Jobs (one trace each) queue on submit; a
// job is admitted when its HBM footprint fits the service budget minus what
// admitted-but-unreleased jobs hold (first fit, with a bound on how often the
// oldest job may be passed over).  Worker threads, each with its own stream,
// run admitted jobs concurrently, so jobs arriving while others are on the
// GPU overlap with them.  Prover data stays on the device for openings until
// the job is released.
struct SvcJob {
    uint64_t id = 0;
    int field = 0;
    uint64_t seed = 0;
    pg_buf* trace = nullptr;  // owned input, or null for a seeded synthetic trace
    const uint32_t* host = nullptr;  // or a caller-held host trace
    int canonical = 0;
    size_t rows = 0, cols = 0;
    int lb = 1;
    size_t footprint = 0;
    int skipped = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int state = 0;  // 0 pending, 1 running, 2 done
    int rc = PG_OK;
    std::string err;
    pg_tree* tree = nullptr;
    uint32_t root[8] = {0};
};

struct pg_service {
    pg_ctx* ctx = nullptr;
    size_t budget = 0, reserved = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::map<uint64_t, std::unique_ptr<SvcJob>> jobs;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    bool stop = false;
    uint64_t next_id = 1;
    size_t running = 0, peak_running = 0;
};

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static size_t job_footprint(size_t rows, size_t cols, int lb, bool owns_trace) {
    const size_t N = rows << lb;
    const size_t lde = N * cols * 4, digests = 2 * N * 32;
    const size_t tmp = log2_of(rows) > 11 ? lde : 0;
    return lde + digests + tmp + (owns_trace ? rows * cols * 4 : 0);
}

static void run_svc_job(pg_service* sv, SvcJob* j, hipStream_t s) {
    pg_ctx* ctx = sv->ctx;
    (void)hipSetDevice(ctx->device);
    if (j->host) {
        // a trace in host memory: the host-memory commit (column blocks staged
        // and uploaded while earlier blocks are extended and hashed)
        const size_t rows = j->rows, cols = j->cols;
        j->rc = pg_pcs_commit_host(ctx, j->field, &j->host, &rows, &cols, nullptr, 1, (uint32_t)j->lb, j->canonical,
                                   j->root, &j->tree);
        j->host = nullptr;  // the caller may reuse its memory once the job is done
        if (j->rc != PG_OK) j->err = g_err;
        return;
    }
    pg_buf* tr = j->trace;
    int rc = PG_OK;
    if (!tr) {
        rc = buf_new(ctx, j->rows, j->cols, &tr);
        if (rc == PG_OK) {
            hipError_t e = launch_fill_random(j->field, j->seed, tr->ptr, j->rows * j->cols, 1, s);
            if (e != hipSuccess) rc = fail(PG_E_DEVICE, "fill: %s", hipGetErrorString(e));
        }
    }
    pg_buf* lde = nullptr;
    std::shared_ptr<DevMem> scratch;
    if (rc == PG_OK) rc = lde_on_stream(ctx, s, j->field, tr, j->lb, gen_monty(j->field), 1, &lde, &scratch);
    if (rc != PG_OK) (void)hipStreamSynchronize(s);  // the fill may still be writing the trace
    scratch.reset();
    if (tr) pg_buf_free(tr);
    j->trace = nullptr;
    if (rc == PG_OK) {
        rc = mmcs_commit_on_stream(ctx, s, j->field, {lde}, {true}, j->root, &j->tree);
        if (rc != PG_OK) {
            (void)hipStreamSynchronize(s);
            pg_buf_free(lde);
        }
    }
    j->rc = rc;
    if (rc != PG_OK) j->err = g_err;
}

static void svc_worker(pg_service* sv, int wi) {
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        SvcJob* pick = nullptr;
        while (!sv->stop) {
            // first fit over the queue; once the oldest job has been passed
            // over 8 times nothing else is admitted until it fits
            bool rescan = false;
            for (auto it = sv->pending.begin(); it != sv->pending.end(); ++it) {
                SvcJob* j = sv->jobs[*it].get();
                if (j->footprint > sv->budget) {  // can never fit: fail it
                    j->rc = fail(PG_E_ADMISSION, "job needs %zu bytes, budget %zu", j->footprint, sv->budget);
                    j->err = g_err;
                    j->state = 2;
                    j->t_done = now_s();
                    j->footprint = 0;
                    sv->pending.erase(it);
                    sv->cv_done.notify_all();
                    rescan = true;
                    break;
                }
                if (sv->reserved + j->footprint <= sv->budget) {
                    pick = j;
                    sv->pending.erase(it);
                    break;
                }
                if (it == sv->pending.begin() && ++j->skipped > 8) break;
            }
            if (pick) break;
            if (!rescan) sv->cv_work.wait(lk);
        }
        if (!pick) return;  // stopping
        sv->reserved += pick->footprint;
        pick->state = 1;
        pick->t_start = now_s();
        sv->running++;
        sv->peak_running = std::max(sv->peak_running, sv->running);
        lk.unlock();
        run_svc_job(sv, pick, sv->streams[wi]);
        lk.lock();
        sv->running--;
        pick->state = 2;
        pick->t_done = now_s();
        if (pick->rc != PG_OK) {  // nothing stays resident for a failed job
            sv->reserved -= pick->footprint;
            pick->footprint = 0;
            sv->cv_work.notify_all();
        }
        sv->cv_done.notify_all();
    }
}

extern "C" int pg_service_create(pg_ctx* ctx, size_t budget, int workers, pg_service** out) {
    if (!ctx || !out || workers < 1 || workers > 16) return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    PG_TRY(bind_device(ctx));
    std::unique_ptr<pg_service> sv(new pg_service);
    sv->ctx = ctx;
    sv->budget = budget ? budget : ctx->budget;
    for (int i = 0; i < workers; i++) {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        sv->streams.push_back(s);
    }
    pg_service* p = sv.get();
    for (int i = 0; i < workers; i++) sv->workers.emplace_back(svc_worker, p, i);
    *out = sv.release();
    return PG_OK;
}

static int svc_enqueue(pg_service* sv, std::unique_ptr<SvcJob> j, uint64_t* id_out) {
    std::lock_guard<std::mutex> g(sv->mu);
    j->id = sv->next_id++;
    j->t_submit = now_s();
    *id_out = j->id;
    sv->pending.push_back(j->id);
    sv->jobs[j->id] = std::move(j);
    sv->cv_work.notify_one();
    return PG_OK;
}

extern "C" int pg_service_submit(pg_service* sv, int field, pg_buf* trace, uint32_t log_blowup, uint64_t* id_out) {
    if (!sv || !trace || !id_out || !field_ok(field) || !is_pow2(trace->rows) || trace->cols == 0)
        return fail(PG_E_INVALID, "bad argument");
    PG_TRY(need_dense(trace, "pg_service_submit"));
    if (log2_of(trace->rows) + (int)log_blowup > two_adicity(field)) return fail(PG_E_INVALID, "LDE too large");
    std::unique_ptr<SvcJob> j(new SvcJob);
    j->field = field;
    j->trace = trace;  // the service owns it from here
    j->rows = trace->rows;
    j->cols = trace->cols;
    j->lb = (int)log_blowup;
    j->footprint = job_footprint(j->rows, j->cols, j->lb, true);
    return svc_enqueue(sv, std::move(j), id_out);
}

extern "C" int pg_service_submit_synthetic(pg_service* sv, int field, uint64_t seed, size_t rows, size_t cols,
                                           uint32_t log_blowup, uint64_t* id_out) {
    if (!sv || !id_out || !field_ok(field) || !is_pow2(rows) || cols == 0) return fail(PG_E_INVALID, "bad argument");
    if (log2_of(rows) + (int)log_blowup > two_adicity(field)) return fail(PG_E_INVALID, "LDE too large");
    std::unique_ptr<SvcJob> j(new SvcJob);
    j->field = field;
    j->seed = seed;
    j->rows = rows;
    j->cols = cols;
    j->lb = (int)log_blowup;
    j->footprint = job_footprint(rows, cols, j->lb, true);
    return svc_enqueue(sv, std::move(j), id_out);
}

extern "C" int pg_service_submit_host(pg_service* sv, int field, const uint32_t* host, size_t rows, size_t cols,
                                      int canonical, uint32_t log_blowup, uint64_t* id_out) {
    if (!sv || !host || !id_out || !field_ok(field) || !is_pow2(rows) || cols == 0)
        return fail(PG_E_INVALID, "bad argument");
    int L = 0;
    PG_TRY(lde_check(field, rows, cols, log_blowup, &L));
    std::unique_ptr<SvcJob> j(new SvcJob);
    j->field = field;
    j->host = host;
    j->canonical = canonical != 0;
    j->rows = rows;
    j->cols = cols;
    j->lb = (int)log_blowup;
    // plus the sponge states (64 bytes per LDE row) of a commit that absorbs
    // the trace in column blocks
    j->footprint = job_footprint(rows, cols, j->lb, true) + (rows << j->lb) * 64;
    return svc_enqueue(sv, std::move(j), id_out);
}

extern "C" int pg_service_wait(pg_service* sv, uint64_t id, uint32_t root[8], double times[3]) {
    if (!sv) return fail(PG_E_INVALID, "null service");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return fail(PG_E_INVALID, "unknown job %llu", (unsigned long long)id);
    SvcJob* j = it->second.get();
    sv->cv_done.wait(lk, [&] { return j->state == 2; });
    if (times) {
        times[0] = j->t_submit;
        times[1] = j->t_start;
        times[2] = j->t_done;
    }
    if (j->rc != PG_OK) {
        g_err = j->err;
        return j->rc;
    }
    if (root) memcpy(root, j->root, 32);
    return PG_OK;
}

extern "C" int pg_service_open(pg_service* sv, uint64_t id, const uint64_t* idx, size_t nq, uint32_t* rows_out,
                               uint32_t* paths_out) {
    if (!sv) return fail(PG_E_INVALID, "null service");
    pg_tree* t = nullptr;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end() || it->second->state != 2 || !it->second->tree)
            return fail(PG_E_INVALID, "job %llu has no resident prover data", (unsigned long long)id);
        t = it->second->tree;
    }
    return pg_mmcs_open_batch(sv->ctx, t, idx, nq, rows_out, paths_out);
}

extern "C" int pg_service_cancel(pg_service* sv, uint64_t id) {
    if (!sv) return fail(PG_E_INVALID, "null service");
    pg_buf* trace = nullptr;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return fail(PG_E_INVALID, "unknown job %llu", (unsigned long long)id);
        SvcJob* j = it->second.get();
        if (j->state != 0)
            return fail(PG_E_INVALID, "job %llu is already %s", (unsigned long long)id,
                        j->state == 1 ? "running" : "finished");
        for (auto q = sv->pending.begin(); q != sv->pending.end(); ++q)
            if (*q == id) {
                sv->pending.erase(q);
                break;
            }
        j->state = 2;
        j->rc = PG_E_CANCELLED;
        j->err = "job cancelled before it ran";
        j->t_done = now_s();
        j->footprint = 0;  // never reserved
        trace = j->trace;
        j->trace = nullptr;
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();  // the queue head may have changed
    }
    pg_buf_free(trace);
    return PG_OK;
}

extern "C" int pg_service_release(pg_service* sv, uint64_t id) {
    if (!sv) return fail(PG_E_INVALID, "null service");
    std::unique_ptr<SvcJob> j;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return fail(PG_E_INVALID, "unknown job %llu", (unsigned long long)id);
        if (it->second->state != 2) return fail(PG_E_INVALID, "job %llu is not finished", (unsigned long long)id);
        j = std::move(it->second);
        sv->jobs.erase(it);
    }
    if (j->tree) pg_tree_free(j->tree);
    std::lock_guard<std::mutex> g(sv->mu);
    sv->reserved -= j->footprint;
    sv->cv_work.notify_all();
    return PG_OK;
}

extern "C" int pg_service_stats(pg_service* sv, size_t* reserved, size_t* pending, size_t* peak_running) {
    if (!sv) return fail(PG_E_INVALID, "null service");
    std::lock_guard<std::mutex> g(sv->mu);
    if (reserved) *reserved = sv->reserved;
    if (pending) *pending = sv->pending.size();
    if (peak_running) *peak_running = sv->peak_running;
    return PG_OK;
}

extern "C" void pg_service_destroy(pg_service* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto& t : sv->workers) t.join();
    (void)hipSetDevice(sv->ctx->device);
    for (auto& kv : sv->jobs) {
        if (kv.second->tree) pg_tree_free(kv.second->tree);
        if (kv.second->trace) pg_buf_free(kv.second->trace);
    }
    for (auto s : sv->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    delete sv;
}

// ------------------------------------------------------------ challenger --
template <class F>
static void ch_duplex(pg_challenger* c) {
    for (int i = 0; i < c->n_input; i++) c->state[i] = c->input[i];
    c->n_input = 0;
    host_permute<F>(c->state);
    memcpy(c->output, c->state, 32);
    c->n_output = 8;
}
static void duplex(pg_challenger* c) {
    if (c->field == PG_KOALABEAR)
        ch_duplex<KoalaBear>(c);
    else
        ch_duplex<BabyBear>(c);
}
extern "C" void pg_challenger_init(pg_challenger* c, int field) {
    memset(c, 0, sizeof *c);
    c->field = field;
}
extern "C" void pg_challenger_observe(pg_challenger* c, uint32_t v) {
    c->n_output = 0;
    c->input[c->n_input++] = v;
    if (c->n_input == 8) duplex(c);
}
extern "C" uint32_t pg_challenger_sample(pg_challenger* c) {
    if (c->n_input > 0 || c->n_output == 0) duplex(c);
    return c->output[--c->n_output];
}
extern "C" void pg_challenger_sample_ext(pg_challenger* c, uint32_t out[4]) {
    for (int k = 0; k < 4; k++) out[k] = pg_challenger_sample(c);
}
extern "C" uint32_t pg_challenger_sample_bits(pg_challenger* c, int bits) {
    const uint32_t v = pg_challenger_sample(c);
    const uint32_t canon = c->field == PG_KOALABEAR ? from_monty<KoalaBear>(v) : from_monty<BabyBear>(v);
    return bits >= 32 ? canon : (canon & ((1u << bits) - 1u));
}
extern "C" int pg_challenger_check_witness(pg_challenger* c, int bits, uint32_t w) {
    const uint32_t wm = c->field == PG_KOALABEAR ? to_monty<KoalaBear>(w) : to_monty<BabyBear>(w);
    pg_challenger_observe(c, wm);
    return pg_challenger_sample_bits(c, bits) == 0;
}

// ------------------------------------------------------------------- FRI --
extern "C" int pg_fri_fold(pg_ctx* ctx, int field, const pg_buf* v, const uint32_t beta[4], pg_buf** out) {
    if (!ctx || !v || !beta || !out || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    if (v->cols != 4 || v->rows < 2 || !is_pow2(v->rows)) return fail(PG_E_INVALID, "v must be n x 4, n a power of two >= 2");
    PG_TRY(need_dense(v, "pg_fri_fold"));
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const int ln = log2_of(v->rows);
    if (ln > two_adicity(field)) return fail(PG_E_INVALID, "vector too long");
    std::shared_ptr<DevMem> tw;
    PG_TRY(get_twiddles(ctx, l.s, field, ln, &tw));
    pg_buf* o;
    PG_TRY(buf_new(ctx, v->rows / 2, 4, &o));
    hipError_t e = launch_fri_fold(field, v->ptr, v->rows / 2, beta, (const uint32_t*)tw->ptr, o->ptr, l.s);
    if (e == hipSuccess) e = hipStreamSynchronize(l.s);
    if (e != hipSuccess) {
        pg_buf_free(o);
        return fail(PG_E_DEVICE, "fold failed: %s", hipGetErrorString(e));
    }
    *out = o;
    return PG_OK;
}

struct pg_fri {
    pg_ctx* ctx = nullptr;
    int field = 0;
    int log_max = 0;
    std::vector<pg_tree*> trees;  // round r: (2^(log_max - r - 1)) x 8 leaves, owning its vector
};

extern "C" void pg_fri_free(pg_fri* f) {
    if (!f) return;
    for (auto* t : f->trees) pg_tree_free(t);
    delete f;
}
extern "C" int pg_fri_num_rounds(const pg_fri* f) { return f ? (int)f->trees.size() : 0; }
extern "C" const pg_tree* pg_fri_round_tree(const pg_fri* f, int r) {
    return (f && r >= 0 && r < (int)f->trees.size()) ? f->trees[r] : nullptr;
}

extern "C" int pg_fri_commit_phase(pg_ctx* ctx, int field, pg_challenger* ch, const pg_buf* const* inputs,
                                   int max_log, int log_stop, int final_mode, uint32_t* roots_out,
                                   uint32_t* betas_out, uint32_t* final_out, int* rounds_out, pg_fri** out) {
    if (!ctx || !ch || !inputs || !out || !field_ok(field) || ch->field != field || max_log < 0 ||
        max_log > two_adicity(field) || log_stop < 0)
        return fail(PG_E_INVALID, "bad argument");
    int L0 = -1;
    for (int k = max_log; k >= 0; k--)
        if (inputs[k]) {
            L0 = k;
            break;
        }
    if (L0 < 0) return fail(PG_E_INVALID, "no input vectors");
    for (int k = 0; k <= max_log; k++)
        if (inputs[k] && (inputs[k]->rows != ((size_t)1 << k) || inputs[k]->cols != 4))
            return fail(PG_E_INVALID, "input %d must be 2^%d x 4", k, k);
    for (int k = 0; k <= max_log; k++) PG_TRY(need_dense(inputs[k], "pg_fri_commit_phase"));
    if (log_stop > L0) return fail(PG_E_INVALID, "log_stop > input length");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    std::unique_ptr<pg_fri, void (*)(pg_fri*)> f(new pg_fri, pg_fri_free);
    f->ctx = ctx;
    f->field = field;
    f->log_max = L0;
    pg_buf* cur;
    PG_TRY(buf_new(ctx, (size_t)1 << L0, 4, &cur));
    HIP_TRY(hipMemcpyAsync(cur->ptr, inputs[L0]->ptr, cur->bytes, hipMemcpyDeviceToDevice, l.s));
    int r = 0;
    int len_log = L0;
    while (len_log > log_stop) {
        // commit the current vector as (n/2) x 8 leaves; the tree owns it
        pg_buf* leaves = new pg_buf(*cur);
        leaves->rows = cur->rows / 2;
        leaves->cols = 8;
        cur->owned = false;  // ownership moves to `leaves`
        pg_buf_free(cur);
        pg_tree* t;
        int rc = mmcs_commit_on_stream(ctx, l.s, field, {leaves}, {true}, roots_out + 8 * r, &t);
        if (rc != PG_OK) {
            (void)hipStreamSynchronize(l.s);  // the first round's copy may still be writing it
            pg_buf_free(leaves);
            return rc;
        }
        f->trees.push_back(t);
        for (int k = 0; k < 8; k++) pg_challenger_observe(ch, roots_out[8 * r + k]);
        uint32_t beta[4];
        pg_challenger_sample_ext(ch, beta);
        if (betas_out) memcpy(betas_out + 4 * r, beta, 16);
        std::shared_ptr<DevMem> tw;
        PG_TRY(get_twiddles(ctx, l.s, field, len_log, &tw));
        PG_TRY(buf_new(ctx, (size_t)1 << (len_log - 1), 4, &cur));
        auto kt = make_timer(ctx);
        if (kt) kt->begin("fri_fold", l.s);
        hipError_t e = launch_fri_fold(field, leaves->ptr, cur->rows, beta, (const uint32_t*)tw->ptr, cur->ptr, l.s);
        if (kt) kt->end(l.s);
        len_log--;
        if (e == hipSuccess && inputs[len_log])
            e = launch_field_add(field, cur->ptr, inputs[len_log]->ptr, cur->rows * 4, l.s);
        const hipError_t es = hipStreamSynchronize(l.s);
        if (e == hipSuccess) e = es;
        if (e != hipSuccess) {
            pg_buf_free(cur);
            return fail(PG_E_DEVICE, "FRI fold round %d failed: %s", r, hipGetErrorString(e));
        }
        r++;
    }
    const size_t nf = (size_t)1 << len_log;
    std::vector<uint32_t> fin(nf * 4);
    hipError_t e = hipMemcpyAsync(fin.data(), cur->ptr, nf * 16, hipMemcpyDeviceToHost, l.s);
    if (e == hipSuccess) e = hipStreamSynchronize(l.s);
    pg_buf_free(cur);
    HIP_TRY(e);
    if (final_out) memcpy(final_out, fin.data(), nf * 16);
    if (final_mode == 0) {
        for (size_t i = 1; i < nf; i++)
            if (memcmp(&fin[4 * i], &fin[0], 16))
                return fail(PG_E_INVALID, "final FRI vector is not constant (input not low degree)");
        for (int k = 0; k < 4; k++) pg_challenger_observe(ch, fin[k]);
    } else {
        for (size_t i = 0; i < nf * 4; i++) pg_challenger_observe(ch, fin[i]);
    }
    if (rounds_out) *rounds_out = r;
    *out = f.release();
    return PG_OK;
}

extern "C" int pg_fri_grind(pg_ctx* ctx, int field, pg_challenger* ch, int bits, uint32_t* witness_out) {
    if (!ctx || !ch || !witness_out || !field_ok(field) || ch->field != field || bits < 0 || bits > 30)
        return fail(PG_E_INVALID, "bad argument");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const uint32_t P = field == PG_KOALABEAR ? KoalaBear::P : BabyBear::P;
    void* d;
    PG_TRY(ctx_alloc(ctx, 256, &d));
    uint32_t* dstate = (uint32_t*)d;       // 16 words
    uint32_t* din = dstate + 16;           // 8 words
    uint32_t* dres = dstate + 32;          // 1 word
    uint32_t hres = 0xffffffffu;
    hipError_t e = hipMemcpyAsync(dstate, ch->state, 64, hipMemcpyHostToDevice, l.s);
    if (e == hipSuccess) e = hipMemcpyAsync(din, ch->input, 32, hipMemcpyHostToDevice, l.s);
    if (e == hipSuccess) e = hipMemcpyAsync(dres, &hres, 4, hipMemcpyHostToDevice, l.s);
    const uint32_t chunk = 1u << 22;
    for (uint32_t start = 0; e == hipSuccess && start < P; start += chunk) {
        const uint32_t count = std::min<uint32_t>(chunk, P - start);
        e = launch_pow_search(field, dstate, din, ch->n_input, bits, start, count, dres, l.s);
        if (e == hipSuccess) e = hipMemcpyAsync(&hres, dres, 4, hipMemcpyDeviceToHost, l.s);
        if (e == hipSuccess) e = hipStreamSynchronize(l.s);
        if (hres != 0xffffffffu) break;
    }
    ctx_free(ctx, d, 256);
    HIP_TRY(e);
    if (hres == 0xffffffffu) return fail(PG_E_DEVICE, "no proof-of-work witness found");
    if (!pg_challenger_check_witness(ch, bits, hres))
        return fail(PG_E_DEVICE, "device witness %u fails the host check", hres);
    *witness_out = hres;
    return PG_OK;
}

extern "C" int pg_fri_answer_queries(pg_ctx* ctx, const pg_fri* f, const uint64_t* idx, size_t nq,
                                     uint32_t* siblings_out, uint32_t* paths_out) {
    if (!ctx || !f || (nq && !idx)) return fail(PG_E_INVALID, "bad argument");
    for (size_t q = 0; q < nq; q++)
        if (idx[q] >> f->log_max) return fail(PG_E_INVALID, "query index out of range");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    const int R = (int)f->trees.size();
    size_t path_words_per_q = 0;
    for (int r = 0; r < R; r++) path_words_per_q += 8 * (size_t)f->trees[r]->log_h;
    std::vector<uint64_t> leaf(nq);
    std::vector<uint32_t> rows, paths;
    size_t path_off = 0;
    for (int r = 0; r < R; r++) {
        const pg_tree* t = f->trees[r];
        for (size_t q = 0; q < nq; q++) leaf[q] = (idx[q] >> r) >> 1;
        rows.assign(nq * 8, 0);
        paths.assign(nq * t->log_h * 8, 0);
        PG_TRY(open_on_stream(ctx, l.s, t, leaf.data(), nq, rows.data(), paths.data()));
        for (size_t q = 0; q < nq; q++) {
            const bool odd = (idx[q] >> r) & 1;
            if (siblings_out) memcpy(siblings_out + (q * R + r) * 4, rows.data() + q * 8 + (odd ? 0 : 4), 16);
            if (paths_out)
                memcpy(paths_out + q * path_words_per_q + path_off, paths.data() + q * t->log_h * 8,
                       t->log_h * 32);
        }
        path_off += 8 * (size_t)t->log_h;
    }
    return PG_OK;
}

// ------------------------------------------------------------ Pcs::open --
// The whole of TwoAdicFriPcs::open in one call, in the order the verifier
// replays it (vm/src/compiler/recursion/circuit/fri.rs:36-73, 84-99,
// 118-241): the values of every matrix at its points (round -> matrix ->
// point -> column), each observed; alpha; the reduced openings per height;
// the FRI commit phase (a commitment, an observation and a beta per round,
// the final vector observed); the proof-of-work grind; the query indices;
// every input opening and every commit-phase opening.
struct pg_open_proof {
    int field = 0;
    std::vector<uint32_t> values;  // Montgomery words, round -> matrix -> point -> column -> 4
    uint32_t alpha[4] = {0, 0, 0, 0};
    int fri_rounds = 0;
    std::vector<uint32_t> commits, betas, final_poly;
    uint32_t pow_witness = 0;
    std::vector<uint64_t> queries;
    size_t input_words = 0;  // per query: for each round, its rows then its path
    std::vector<uint32_t> inputs;
    size_t fri_path_words = 0;  // per query
    std::vector<uint32_t> siblings, fri_paths;
    // shapes of the input openings: per round, its matrices' widths (the
    // tree's original order) and its log height (path length)
    std::vector<std::vector<size_t>> round_widths;
    std::vector<int> round_log_h;
    std::vector<int> fri_path_levels;  // per commit-phase round
    size_t num_queries = 0;
};

extern "C" void pg_open_proof_free(pg_open_proof* p) { delete p; }

extern "C" int pg_pcs_open(pg_ctx* ctx, const pg_tree* const* rounds, size_t nrounds, const uint32_t* npoints,
                           const uint32_t* points_mont, uint32_t log_blowup, int log_final, size_t num_queries,
                           int pow_bits, pg_challenger* ch, pg_open_proof** out) {
    if (!ctx || !rounds || nrounds == 0 || !npoints || !ch || !out || !rounds[0])
        return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    const int field = rounds[0]->field;
    if (ch->field != field) return fail(PG_E_INVALID, "challenger field differs from the commitments'");
    size_t nvals = 0, npts_total = 0, mi = 0;
    for (size_t r = 0; r < nrounds; r++) {
        if (!rounds[r] || rounds[r]->field != field) return fail(PG_E_INVALID, "round %zu: missing tree or mixed fields", r);
        for (size_t k = 0; k < rounds[r]->mats.size(); k++, mi++) {
            nvals += (size_t)npoints[mi] * rounds[r]->mats[k]->cols * 4;
            npts_total += npoints[mi];
        }
    }
    if (npts_total == 0) return fail(PG_E_INVALID, "no opening points");
    if (!points_mont) return fail(PG_E_INVALID, "null points");
    std::unique_ptr<pg_open_proof> pf(new pg_open_proof);
    pf->field = field;
    // 1. values, observed in order
    pf->values.resize(nvals);
    size_t voff = 0, poff = 0;
    mi = 0;
    {
        WeightCache wc;  // freed once every value is on the host (each call synchronises)
        for (size_t r = 0; r < nrounds; r++)
            for (size_t k = 0; k < rounds[r]->mats.size(); k++, mi++) {
                const uint32_t np = npoints[mi];
                if (np)
                    PG_TRY(open_values_cached(ctx, rounds[r], k, log_blowup, points_mont + 4 * poff, np,
                                              pf->values.data() + voff, &wc));
                voff += (size_t)np * rounds[r]->mats[k]->cols * 4;
                poff += np;
            }
    }
    for (uint32_t v : pf->values) pg_challenger_observe(ch, v);
    // 2. alpha, reduced openings per height
    pg_challenger_sample_ext(ch, pf->alpha);
    pg_buf* by_log[33];
    PG_TRY(pg_pcs_reduce_openings(ctx, rounds, nrounds, npoints, points_mont, pf->values.data(), pf->alpha, by_log));
    auto free_reduced = [&]() {
        for (auto*& b : by_log) {
            pg_buf_free(b);
            b = nullptr;
        }
    };
    int log_max = -1;
    for (int lh = 0; lh < 33; lh++)
        if (by_log[lh]) log_max = lh;
    const int log_stop = log_final < 0 ? (int)log_blowup : log_final;
    if (log_max < 0 || log_stop > log_max) {
        free_reduced();
        return fail(PG_E_INVALID, "FRI stop 2^%d above the tallest reduced opening 2^%d", log_stop, log_max);
    }
    for (size_t r = 0; r < nrounds; r++)
        if (rounds[r]->log_h > log_max) {
            free_reduced();
            return fail(PG_E_INVALID, "round %zu is taller than every opened matrix", r);
        }
    // 3. FRI commit phase
    const int max_rounds = log_max - log_stop;
    pf->commits.resize(8 * std::max(max_rounds, 1));
    pf->betas.resize(4 * std::max(max_rounds, 1));
    pf->final_poly.resize(4 * ((size_t)1 << log_stop));
    pg_fri* fri = nullptr;
    int rc = pg_fri_commit_phase(ctx, field, ch, by_log, log_max, log_stop, log_final < 0 ? 0 : 1,
                                 pf->commits.data(), pf->betas.data(), pf->final_poly.data(), &pf->fri_rounds, &fri);
    free_reduced();
    if (rc != PG_OK) return rc;
    std::unique_ptr<pg_fri, void (*)(pg_fri*)> fri_guard(fri, pg_fri_free);
    pf->commits.resize(8 * (size_t)pf->fri_rounds);
    pf->betas.resize(4 * (size_t)pf->fri_rounds);
    // 4. proof of work, query indices
    PG_TRY(pg_fri_grind(ctx, field, ch, pow_bits, &pf->pow_witness));
    pf->queries.resize(num_queries);
    pf->num_queries = num_queries;
    for (auto& q : pf->queries) q = pg_challenger_sample_bits(ch, log_max);
    // 5. input openings, round by round, interleaved per query
    for (size_t r = 0; r < nrounds; r++) {
        pf->input_words += rounds[r]->total_width + 8 * (size_t)rounds[r]->log_h;
        std::vector<size_t> ws;
        for (const pg_buf* m : rounds[r]->mats) ws.push_back(m->cols);
        pf->round_widths.push_back(std::move(ws));
        pf->round_log_h.push_back(rounds[r]->log_h);
    }
    pf->inputs.resize(num_queries * pf->input_words);
    size_t roff = 0;
    std::vector<uint64_t> idx(num_queries);
    for (size_t r = 0; r < nrounds; r++) {
        const pg_tree* t = rounds[r];
        const size_t tw = t->total_width, pw = 8 * (size_t)t->log_h;
        for (size_t q = 0; q < num_queries; q++) idx[q] = pf->queries[q] >> (log_max - t->log_h);
        std::vector<uint32_t> rows(num_queries * tw), paths(num_queries * pw);
        PG_TRY(pg_mmcs_open_batch(ctx, t, idx.data(), num_queries, rows.data(), paths.data()));
        for (size_t q = 0; q < num_queries; q++) {
            uint32_t* dst = pf->inputs.data() + q * pf->input_words + roff;
            if (tw) memcpy(dst, rows.data() + q * tw, tw * 4);
            if (pw) memcpy(dst + tw, paths.data() + q * pw, pw * 4);
        }
        roff += tw + pw;
    }
    // 6. commit-phase openings
    for (auto* t : fri->trees) {
        pf->fri_path_words += 8 * (size_t)t->log_h;
        pf->fri_path_levels.push_back(t->log_h);
    }
    pf->siblings.resize(num_queries * (size_t)pf->fri_rounds * 4);
    pf->fri_paths.resize(num_queries * pf->fri_path_words);
    PG_TRY(pg_fri_answer_queries(ctx, fri, pf->queries.data(), num_queries, pf->siblings.data(),
                                 pf->fri_paths.data()));
    *out = pf.release();
    return PG_OK;
}

extern "C" int pg_open_proof_sizes(const pg_open_proof* p, size_t* n_values, int* fri_rounds, size_t* final_words,
                                   size_t* num_queries, size_t* input_words_per_query,
                                   size_t* fri_path_words_per_query) {
    if (!p) return fail(PG_E_INVALID, "null proof");
    if (n_values) *n_values = p->values.size();
    if (fri_rounds) *fri_rounds = p->fri_rounds;
    if (final_words) *final_words = p->final_poly.size();
    if (num_queries) *num_queries = p->num_queries;
    if (input_words_per_query) *input_words_per_query = p->input_words;
    if (fri_path_words_per_query) *fri_path_words_per_query = p->fri_path_words;
    return PG_OK;
}

extern "C" int pg_open_proof_read(const pg_open_proof* p, uint32_t* values, uint32_t alpha[4], uint32_t* commits,
                                  uint32_t* betas, uint32_t* final_poly, uint32_t* pow_witness,
                                  uint64_t* query_indices, uint32_t* input_openings, uint32_t* fri_siblings,
                                  uint32_t* fri_paths) {
    if (!p) return fail(PG_E_INVALID, "null proof");
    auto put = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) memcpy(dst, src, bytes);
    };
    put(values, p->values.data(), p->values.size() * 4);
    put(alpha, p->alpha, 16);
    put(commits, p->commits.data(), p->commits.size() * 4);
    put(betas, p->betas.data(), p->betas.size() * 4);
    put(final_poly, p->final_poly.data(), p->final_poly.size() * 4);
    if (pow_witness) *pow_witness = p->pow_witness;
    put(query_indices, p->queries.data(), p->queries.size() * 8);
    put(input_openings, p->inputs.data(), p->inputs.size() * 4);
    put(fri_siblings, p->siblings.data(), p->siblings.size() * 4);
    put(fri_paths, p->fri_paths.data(), p->fri_paths.size() * 4);
    return PG_OK;
}

// ------------------------------------------------------- proof encoding --
// Bincode 1 (little-endian, fixed-width integers, u64 sequence lengths) of
// p3-fri's FriProof<Challenge, ChallengeMmcs, Val, Vec<BatchOpening>>, the
// PcsProof TwoAdicFriPcs::open returns.  Field order as the reference's
// recursion circuit reads it (FriProofVariable, QueryProofVariable,
// BatchOpeningVariable, FriCommitPhaseProofStepVariable at
// vm/src/compiler/recursion/circuit/types.rs:29-54, filled in that order by
// circuit/witness/embed.rs:55-165):
//   commit_phase_commits  Vec<Hash<Val, Val, 8>>  u64 n, n x 8 words
//   query_proofs          Vec<QueryProof>         u64 n, then per query:
//     input_proof         Vec<BatchOpening>       u64 rounds, then per round:
//       opened_values     Vec<Vec<Val>>           u64 matrices, each u64 w + w words
//       opening_proof     Vec<[Val; 8]>           u64 levels, levels x 8 words
//     commit_phase_openings Vec<CommitPhaseProofStep>  u64 rounds, each:
//       sibling_value     Challenge               4 words
//       opening_proof     Vec<[Val; 8]>           u64 levels, levels x 8 words
//   final_poly            Challenge               4 words
//   pow_witness           Val                     1 word
// Arrays and extension elements are fixed-size tuples: no length.  A word is
// a field element as a u32: the Montgomery value p3 keeps in memory, or with
// canonical != 0 the canonical value.
struct ByteSink {
    uint8_t* out;
    size_t cap, n = 0;
    void put(const void* v, size_t k) {
        if (out && n + k <= cap) memcpy(out + n, v, k);
        n += k;
    }
    void u64(uint64_t v) { put(&v, 8); }
    void u32(uint32_t v) { put(&v, 4); }
};

static uint32_t word_out(int field, uint32_t v, int canonical) {
    if (!canonical) return v;
    return field == PG_KOALABEAR ? from_monty<KoalaBear>(v) : from_monty<BabyBear>(v);
}

extern "C" int pg_open_proof_encode(const pg_open_proof* p, int canonical, uint8_t* out, size_t cap, size_t* len) {
    if (!p || !len) return fail(PG_E_INVALID, "bad argument");
    if (p->final_poly.size() < 4 || p->final_poly.size() % 4) return fail(PG_E_INVALID, "proof has no final value");
    for (size_t i = 4; i < p->final_poly.size(); i++)
        if (p->final_poly[i] != p->final_poly[i % 4])
            return fail(PG_E_UNSUPPORTED,
                        "the final vector is not constant; p3's FriProof holds one final value (open with "
                        "log_final < 0)");
    if ((int)p->fri_path_levels.size() != p->fri_rounds || p->round_widths.size() != p->round_log_h.size())
        return fail(PG_E_INVALID, "proof shapes are incomplete");
    const int f = p->field;
    ByteSink b{out, cap};
    auto word = [&](uint32_t v) { b.u32(word_out(f, v, canonical)); };
    b.u64((uint64_t)p->fri_rounds);
    for (size_t i = 0; i < 8 * (size_t)p->fri_rounds; i++) word(p->commits[i]);
    b.u64(p->num_queries);
    for (size_t q = 0; q < p->num_queries; q++) {
        const uint32_t* in = p->inputs.data() + q * p->input_words;
        b.u64(p->round_widths.size());
        for (size_t r = 0; r < p->round_widths.size(); r++) {
            b.u64(p->round_widths[r].size());
            for (size_t w : p->round_widths[r]) {
                b.u64(w);
                for (size_t j = 0; j < w; j++) word(*in++);
            }
            b.u64((uint64_t)p->round_log_h[r]);
            for (size_t j = 0; j < 8 * (size_t)p->round_log_h[r]; j++) word(*in++);
        }
        b.u64((uint64_t)p->fri_rounds);
        const uint32_t* fp = p->fri_paths.data() + q * p->fri_path_words;
        for (int r = 0; r < p->fri_rounds; r++) {
            for (int e = 0; e < 4; e++) word(p->siblings[(q * (size_t)p->fri_rounds + r) * 4 + e]);
            b.u64((uint64_t)p->fri_path_levels[r]);
            for (size_t j = 0; j < 8 * (size_t)p->fri_path_levels[r]; j++) word(*fp++);
        }
    }
    for (int e = 0; e < 4; e++) word(p->final_poly[e]);
    // the witness is held canonical
    b.u32(canonical ? p->pow_witness
                    : (f == PG_KOALABEAR ? to_monty<KoalaBear>(p->pow_witness) : to_monty<BabyBear>(p->pow_witness)));
    *len = b.n;
    if (out && b.n > cap) return fail(PG_E_INVALID, "buffer of %zu bytes; the encoding needs %zu", cap, b.n);
    return PG_OK;
}

struct ByteSource {
    const uint8_t* p;
    size_t len, n = 0;
    bool ok = true;
    bool take(void* v, size_t k) {
        if (!ok || n + k > len) return ok = false;
        memcpy(v, p + n, k);
        n += k;
        return true;
    }
    uint64_t u64() {
        uint64_t v = 0;
        take(&v, 8);
        return v;
    }
    // a sequence length whose elements take at least `unit` bytes each
    uint64_t count(size_t unit) {
        const uint64_t v = u64();
        if (ok && unit && v > (len - n) / unit) ok = false;
        return ok ? v : 0;
    }
};

extern "C" int pg_open_proof_decode(int field, const uint8_t* bytes, size_t len, int canonical,
                                    pg_open_proof** out) {
    if (!out || (!bytes && len) || !field_ok(field)) return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    const uint32_t P = field == PG_KOALABEAR ? KoalaBear::P : BabyBear::P;
    ByteSource src{bytes, len};
    bool bad_word = false;
    auto word = [&]() -> uint32_t {
        uint32_t v = 0;
        src.take(&v, 4);
        if (v >= P) bad_word = true;
        if (!canonical) return v;
        return field == PG_KOALABEAR ? to_monty<KoalaBear>(v % P) : to_monty<BabyBear>(v % P);
    };
    std::unique_ptr<pg_open_proof> pf(new pg_open_proof);
    pf->field = field;
    const uint64_t R = src.count(32);
    pf->fri_rounds = (int)R;
    for (uint64_t i = 0; i < 8 * R && src.ok; i++) pf->commits.push_back(word());
    const uint64_t nq = src.count(16);
    pf->num_queries = nq;
    bool shape_bad = false;
    for (uint64_t q = 0; q < nq && src.ok && !shape_bad; q++) {
        std::vector<std::vector<size_t>> widths;
        std::vector<int> log_h;
        const uint64_t nr = src.count(16);
        for (uint64_t r = 0; r < nr && src.ok; r++) {
            std::vector<size_t> ws;
            const uint64_t nm = src.count(8);
            for (uint64_t k = 0; k < nm && src.ok; k++) {
                const uint64_t w = src.count(4);
                ws.push_back(w);
                for (uint64_t j = 0; j < w && src.ok; j++) pf->inputs.push_back(word());
            }
            const uint64_t levels = src.count(32);
            for (uint64_t j = 0; j < 8 * levels && src.ok; j++) pf->inputs.push_back(word());
            widths.push_back(std::move(ws));
            log_h.push_back((int)levels);
        }
        const uint64_t nfr = src.count(24);
        std::vector<int> fl;
        for (uint64_t r = 0; r < nfr && src.ok; r++) {
            for (int e = 0; e < 4; e++) pf->siblings.push_back(word());
            const uint64_t levels = src.count(32);
            for (uint64_t j = 0; j < 8 * levels && src.ok; j++) pf->fri_paths.push_back(word());
            fl.push_back((int)levels);
        }
        if (q == 0) {
            pf->round_widths = widths;
            pf->round_log_h = log_h;
            pf->fri_path_levels = fl;
        } else if (widths != pf->round_widths || log_h != pf->round_log_h || fl != pf->fri_path_levels) {
            shape_bad = true;
        }
        if (nfr != R) shape_bad = true;
    }
    pf->final_poly.resize(4);
    for (int e = 0; e < 4; e++) pf->final_poly[e] = word();
    const uint32_t wit = word();  // held canonical
    pf->pow_witness = field == PG_KOALABEAR ? from_monty<KoalaBear>(wit) : from_monty<BabyBear>(wit);
    if (!src.ok) return fail(PG_E_INVALID, "encoding ends early or declares more data than it holds");
    if (src.n != len) return fail(PG_E_INVALID, "%zu bytes after the end of the proof", len - src.n);
    if (bad_word) return fail(PG_E_INVALID, "a field element is not below p");
    if (shape_bad)
        return fail(PG_E_INVALID, "queries differ in shape (every query opens the same rounds, matrices and rounds "
                                  "of the commit phase)");
    for (size_t r = 0; r < pf->round_widths.size(); r++) {
        for (size_t w : pf->round_widths[r]) pf->input_words += w;
        pf->input_words += 8 * (size_t)pf->round_log_h[r];
    }
    for (int l : pf->fri_path_levels) pf->fri_path_words += 8 * (size_t)l;
    *out = pf.release();
    return PG_OK;
}

// ------------------------------------------------------------- autotune --
// Per-pass durations of one LDE, by pass name (HIP events around each launch).
struct PassTimer : KernelTimer {
    hipEvent_t a = nullptr, b = nullptr;
    std::string cur;
    std::map<std::string, double> ms;
    PassTimer() {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~PassTimer() override {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void begin(const char* name, hipStream_t s) override {
        cur = name;
        (void)hipEventRecord(a, s);
    }
    void end(hipStream_t s) override {
        (void)hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) return;
        float t = 0;
        if (hipEventElapsedTime(&t, a, b) == hipSuccess) ms[cur] += t;
    }
};

extern "C" int pg_autotune(pg_ctx* ctx, int field, int log_h_min, int log_h_max, const uint32_t* widths, size_t nw,
                           uint32_t log_blowup, int reps, char* report, size_t report_len) {
    if (!ctx || !field_ok(field) || !widths || nw == 0 || log_h_min < 1 || log_h_max < log_h_min || reps < 1)
        return fail(PG_E_INVALID, "bad argument");
    for (size_t i = 0; i < nw; i++)
        if (widths[i] == 0) return fail(PG_E_INVALID, "width zero");
    if (log_h_max > lde_max_log_h()) return fail(PG_E_UNSUPPORTED, "log height %d > %d", log_h_max, lde_max_log_h());
    if (log_h_max + (int)log_blowup > two_adicity(field) || log_blowup > 8)
        return fail(PG_E_INVALID, "LDE beyond the two-adic subgroup");
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    static const char* const names[4] = {"lde_single", "lde_pass_a", "lde_pass_bc", "lde_pass_d"};
    // total ms per (kind, log_s, bucket, variant) over every tuned shape
    std::map<std::tuple<int, int, int>, double> total[2];
    for (int L = log_h_min; L <= log_h_max; L++)
        for (size_t wi = 0; wi < nw; wi++) {
            const uint32_t w = widths[wi];
            const int bucket = lde_width_bucket(w);
            int k1, k2;
            lde_split(L, &k1, &k2);
            std::vector<std::pair<int, int>> passes;  // (kind, log_s)
            if (L <= 11)
                passes = {{KIND_SINGLE, L}};
            else
                passes = {{KIND_A, k1}, {KIND_BC, k2}, {KIND_D, k1}};
            pg_buf* in = nullptr;
            int rc = buf_new(ctx, (size_t)1 << L, w, &in);
            if (rc != PG_OK) return rc;
            std::unique_ptr<pg_buf, void (*)(pg_buf*)> in_guard(in, pg_buf_free);
            HIP_TRY(launch_fill_random(field, 0x7475 + (uint64_t)L * 4096 + w, in->ptr, in->rows * in->cols, 1, l.s));
            std::shared_ptr<DevMem> scratch;
            for (int v = 0; v < 2; v++) {
                // every pass of this LDE in variant v where it has one
                std::vector<int> saved;
                for (auto& ps : passes) {
                    saved.push_back(lde_variant(field, ps.first, ps.second, w));
                    lde_set_variant(field, ps.first, ps.second, bucket, v < lde_num_variants(ps.first, ps.second) ? v : 0);
                }
                PassTimer pt;
                for (int r = 0; r <= reps; r++) {  // r = 0 warms up
                    pg_buf* o = nullptr;
                    pt.ms.clear();
                    const uint32_t shift = gen_monty(field);
                    int L2;
                    rc = lde_check(field, in->rows, in->cols, log_blowup, &L2);
                    std::shared_ptr<DevMem> tw, sc;
                    if (rc == PG_OK) rc = get_twiddles(ctx, l.s, field, L, &tw);
                    if (rc == PG_OK) rc = get_scale(ctx, l.s, field, L, (int)log_blowup, shift, &sc);
                    if (rc == PG_OK) rc = buf_new(ctx, in->rows << log_blowup, w, &o);
                    if (rc == PG_OK && L > 11 && (!scratch || scratch->bytes < o->bytes)) {
                        scratch.reset();
                        scratch = make_devmem(ctx, o->bytes, &rc);
                    }
                    hipError_t e = hipSuccess;
                    if (rc == PG_OK)
                        e = launch_coset_lde(field, in->ptr, w, L, w, (int)log_blowup, (const uint32_t*)tw->ptr,
                                             (const uint32_t*)sc->ptr, L > 11 ? (uint32_t*)scratch->ptr : nullptr,
                                             o->ptr, w, 1, l.s, &pt);
                    if (e == hipSuccess && rc == PG_OK) e = hipStreamSynchronize(l.s);
                    pg_buf_free(o);
                    if (rc == PG_OK && e != hipSuccess) rc = fail(PG_E_DEVICE, "autotune LDE: %s", hipGetErrorString(e));
                    if (rc != PG_OK) {
                        for (size_t i = 0; i < passes.size(); i++)
                            lde_set_variant(field, passes[i].first, passes[i].second, bucket, saved[i]);
                        return rc;
                    }
                    if (r == 0) continue;
                    for (auto& ps : passes)
                        if (v < lde_num_variants(ps.first, ps.second))
                            total[v][std::make_tuple(ps.first, ps.second, bucket)] += pt.ms[names[ps.first]];
                }
                for (size_t i = 0; i < passes.size(); i++)
                    lde_set_variant(field, passes[i].first, passes[i].second, bucket, saved[i]);
            }
        }
    // the faster variant wherever both were timed
    std::string rep = "# field kind log_s width_bucket ms_default ms_alternative -> variant\n";
    for (auto& kv : total[1]) {
        auto it = total[0].find(kv.first);
        if (it == total[0].end()) continue;
        const int kind = std::get<0>(kv.first), log_s = std::get<1>(kv.first), bucket = std::get<2>(kv.first);
        const int choice = kv.second < it->second ? 1 : 0;
        lde_set_variant(field, kind, log_s, bucket, choice);
        char line[160];
        snprintf(line, sizeof(line), "%d %s %d %d %.4f %.4f -> %d\n", field, names[kind], log_s, bucket,
                 it->second / reps, kv.second / reps, choice);
        rep += line;
    }
    if (report && report_len) {
        strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return PG_OK;
}

extern "C" int pg_tune_set(int field, int kind, int log_s, int width_bucket, int variant) {
    if (lde_set_variant(field, kind, log_s, width_bucket, variant) != 0)
        return fail(PG_E_INVALID, "no tile-shape variant %d for field %d kind %d log size %d bucket %d", variant, field,
                    kind, log_s, width_bucket);
    return PG_OK;
}

extern "C" int pg_tune_get(int field, int kind, int log_s, uint32_t width, int* variant) {
    if (!variant || !field_ok(field) || kind < 0 || kind > 3 || log_s < 0 || log_s > 11)
        return fail(PG_E_INVALID, "bad argument");
    *variant = lde_variant(field, kind, log_s, width);
    return PG_OK;
}

extern "C" int pg_tune_export(char* buf, size_t len, size_t* needed) {
    std::string s;
    for (int f = 0; f < 2; f++)
        for (int k = 0; k < 4; k++)
            for (int ls = 0; ls < 12; ls++) {
                if (lde_num_variants(k, ls) < 2) continue;
                for (int b = 0; b < LDE_WIDTH_BUCKETS; b++) {
                    char line[64];
                    snprintf(line, sizeof(line), "%d %d %d %d %d\n", f, k, ls, b,
                             lde_variant(f, k, ls, b == 0 ? 1u : b == 1 ? 9u : b == 2 ? 33u : 129u));
                    s += line;
                }
            }
    if (needed) *needed = s.size() + 1;
    if (buf && len) {
        if (len < s.size() + 1) return fail(PG_E_INVALID, "buffer of %zu bytes, %zu needed", len, s.size() + 1);
        memcpy(buf, s.c_str(), s.size() + 1);
    }
    return PG_OK;
}

extern "C" int pg_tune_import(const char* text) {
    if (!text) return fail(PG_E_INVALID, "null text");
    // validate every line before applying any
    std::vector<std::array<int, 5>> rows;
    const char* p = text;
    int lineno = 0;
    while (*p) {
        const char* e = strchr(p, '\n');
        std::string line(p, e ? (size_t)(e - p) : strlen(p));
        p = e ? e + 1 : p + line.size();
        lineno++;
        const size_t c = line.find('#');
        if (c != std::string::npos) line.resize(c);
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::array<int, 5> v{};
        char tail = 0;
        if (sscanf(line.c_str(), "%d %d %d %d %d %c", &v[0], &v[1], &v[2], &v[3], &v[4], &tail) != 5)
            return fail(PG_E_INVALID, "line %d: expected 'field kind log_s bucket variant'", lineno);
        if (v[0] < 0 || v[0] > 1 || v[1] < 0 || v[1] > 3 || v[2] < 0 || v[2] > 11 || v[3] < 0 ||
            v[3] >= LDE_WIDTH_BUCKETS || v[4] < 0 || v[4] >= lde_num_variants(v[1], v[2]))
            return fail(PG_E_INVALID, "line %d: no such entry or variant", lineno);
        rows.push_back(v);
    }
    for (auto& v : rows) lde_set_variant(v[0], v[1], v[2], v[3], v[4]);
    return PG_OK;
}

// ----------------------------------------------------- quotient values --
struct pg_air_program {
    int field = 0;
    std::vector<QInstr> code;
    // Montgomery; at launch the public values, challenge words, regional sum
    // and global-sum coordinates follow them
    std::vector<uint32_t> consts;
    uint32_t nregs = 0, nslots = 0, nasserts = 0, nstages = 0;
    // per base assert: (component e, original assert j), folded with X^e alpha^(m-1-j)
    std::vector<std::pair<uint32_t, uint32_t>> afactor;
    uint32_t norig = 0;  // m, the asserts as given
    bool need_selectors = false, reads_pre = false, reads_perm = false, reads_rsum = false;
    // one past the largest index read
    uint32_t max_main_col = 0, max_pre_col = 0, max_perm_col = 0, max_public = 0, max_chal_word = 0, max_gsum = 0;
};

// Leaf kinds of the lowered (base-field) DAG beyond the ABI's: a base column
// of the permutation trace, a word of the challenges, of the regional sum,
// a global-sum coordinate.
enum : uint32_t { LK_PERM = 32, LK_CHAL, LK_RSUM, LK_GSUM };
struct LAssert {
    uint32_t node, comp, orig;
};
static int air_compile_lowered(int field, const std::vector<pg_air_node>& lnodes, const std::vector<LAssert>& las,
                               uint32_t norig, pg_air_program** out, uint32_t* nregs_out, uint32_t* ninstr_out);

extern "C" void pg_air_program_free(pg_air_program* prog) { delete prog; }

extern "C" int pg_air_compile(int field, const pg_air_node* nodes, size_t nnodes, const uint32_t* asserts,
                              size_t nasserts, pg_air_program** out, uint32_t* nregs_out, uint32_t* ninstr_out) {
    if (!out || !field_ok(field) || (nnodes && !nodes) || (nasserts && !asserts))
        return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    if (nnodes > (1u << 26) || nasserts > (1u << 26)) return fail(PG_E_INVALID, "too many nodes");
    for (size_t k = 0; k < nnodes; k++) {
        const pg_air_node& nd = nodes[k];
        switch (nd.kind) {
            case PG_AIR_MAIN:
            case PG_AIR_PREPROCESSED:
                if (nd.a > 1) return fail(PG_E_INVALID, "node %zu: row offset %u (0 local, 1 next)", k, nd.a);
                if (nd.b >= (1u << 24)) return fail(PG_E_INVALID, "node %zu: column %u", k, nd.b);
                break;
            case PG_AIR_PERMUTATION:
                if (nd.a > 1) return fail(PG_E_INVALID, "node %zu: row offset %u (0 local, 1 next)", k, nd.a);
                if (nd.b >= (1u << 22)) return fail(PG_E_INVALID, "node %zu: permutation column %u", k, nd.b);
                break;
            case PG_AIR_ADD:
            case PG_AIR_SUB:
            case PG_AIR_MUL:
                if (nd.a >= k || nd.b >= k) return fail(PG_E_INVALID, "node %zu: operands must be earlier nodes", k);
                break;
            case PG_AIR_NEG:
                if (nd.a >= k) return fail(PG_E_INVALID, "node %zu: operand must be an earlier node", k);
                break;
            case PG_AIR_PUBLIC:
                if (nd.b >= (1u << 24)) return fail(PG_E_INVALID, "node %zu: public value %u", k, nd.b);
                break;
            case PG_AIR_PERM_CHALLENGE:
                if (nd.b >= (1u << 20)) return fail(PG_E_INVALID, "node %zu: challenge %u", k, nd.b);
                break;
            case PG_AIR_GLOBAL_SUM:
                if (nd.b >= 14) return fail(PG_E_INVALID, "node %zu: global-sum coordinate %u (0..13)", k, nd.b);
                break;
            case PG_AIR_IS_FIRST_ROW:
            case PG_AIR_IS_LAST_ROW:
            case PG_AIR_IS_TRANSITION:
            case PG_AIR_CONSTANT:
            case PG_AIR_REGIONAL_SUM:
                break;
            default:
                return fail(PG_E_INVALID, "node %zu: unknown kind %u", k, nd.kind);
        }
    }
    for (size_t j = 0; j < nasserts; j++)
        if (asserts[j] >= nnodes) return fail(PG_E_INVALID, "assert %zu: node %u out of range", j, asserts[j]);
    // Lowering to the base field.  A node is extension-valued if it is an
    // ext4 leaf or has an extension-valued operand; such a node becomes its
    // four coordinates (base nodes; -1 is a zero coordinate), a base node one
    // node.  In F[X]/(X^4 - W): sums and negations coordinate-wise, a base
    // factor scales every coordinate, and a product of two extension values
    // is c_k = sum_{i+j=k} a_i b_j + W sum_{i+j=k+4} a_i b_j.
    const int64_t NONE = -1;
    std::vector<pg_air_node> ln;
    std::vector<std::array<int64_t, 4>> lo(nnodes, {NONE, NONE, NONE, NONE});
    std::vector<uint8_t> is_ext(nnodes, 0);
    std::map<std::array<uint32_t, 3>, uint32_t> memo;  // hash-consing of the lowered nodes
    auto node = [&](uint32_t kind, uint32_t a, uint32_t b) -> int64_t {
        const std::array<uint32_t, 3> key{kind, a, b};
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        ln.push_back(pg_air_node{kind, a, b});
        memo[key] = (uint32_t)(ln.size() - 1);
        return (int64_t)(ln.size() - 1);
    };
    auto add = [&](int64_t x, int64_t y) -> int64_t {
        return x == NONE ? y : y == NONE ? x : node(PG_AIR_ADD, (uint32_t)x, (uint32_t)y);
    };
    auto neg = [&](int64_t x) -> int64_t { return x == NONE ? NONE : node(PG_AIR_NEG, (uint32_t)x, 0); };
    auto sub = [&](int64_t x, int64_t y) -> int64_t {
        return y == NONE ? x : x == NONE ? neg(y) : node(PG_AIR_SUB, (uint32_t)x, (uint32_t)y);
    };
    auto mul = [&](int64_t x, int64_t y) -> int64_t {
        return x == NONE || y == NONE ? NONE : node(PG_AIR_MUL, (uint32_t)x, (uint32_t)y);
    };
    const uint32_t W = field == PG_KOALABEAR ? 3u : 11u;  // X^4 = W (vm/src/primitives/consts.rs:107,134)
    for (size_t k = 0; k < nnodes; k++) {
        const pg_air_node& nd = nodes[k];
        auto& o = lo[k];
        switch (nd.kind) {
            case PG_AIR_PERMUTATION:
                is_ext[k] = 1;
                for (uint32_t e = 0; e < 4; e++) o[e] = node(LK_PERM, nd.a, 4 * nd.b + e);
                break;
            case PG_AIR_PERM_CHALLENGE:
                is_ext[k] = 1;
                for (uint32_t e = 0; e < 4; e++) o[e] = node(LK_CHAL, 0, 4 * nd.b + e);
                break;
            case PG_AIR_REGIONAL_SUM:
                is_ext[k] = 1;
                for (uint32_t e = 0; e < 4; e++) o[e] = node(LK_RSUM, 0, e);
                break;
            case PG_AIR_GLOBAL_SUM: o[0] = node(LK_GSUM, 0, nd.b); break;
            case PG_AIR_ADD:
            case PG_AIR_SUB: {
                is_ext[k] = is_ext[nd.a] | is_ext[nd.b];
                for (int e = 0; e < 4; e++)
                    o[e] = nd.kind == PG_AIR_ADD ? add(lo[nd.a][e], lo[nd.b][e]) : sub(lo[nd.a][e], lo[nd.b][e]);
                break;
            }
            case PG_AIR_NEG:
                is_ext[k] = is_ext[nd.a];
                for (int e = 0; e < 4; e++) o[e] = neg(lo[nd.a][e]);
                break;
            case PG_AIR_MUL: {
                is_ext[k] = is_ext[nd.a] | is_ext[nd.b];
                const auto& x = lo[nd.a];
                const auto& y = lo[nd.b];
                if (!is_ext[nd.a] || !is_ext[nd.b]) {  // a base factor scales every coordinate
                    const int64_t f = !is_ext[nd.a] ? x[0] : y[0];
                    const auto& v = !is_ext[nd.a] ? y : x;
                    for (int e = 0; e < 4; e++) o[e] = mul(f, v[e]);
                    break;
                }
                int64_t wnode = NONE;
                for (int c = 0; c < 4; c++) {
                    int64_t lowp = NONE, highp = NONE;
                    for (int i = 0; i < 4; i++) {
                        const int j = c - i;
                        if (j >= 0) lowp = add(lowp, mul(x[i], y[j]));
                        if (c + 4 - i < 4) highp = add(highp, mul(x[i], y[c + 4 - i]));
                    }
                    if (highp != NONE && wnode == NONE) wnode = node(PG_AIR_CONSTANT, W, 0);
                    o[c] = add(lowp, mul(highp, wnode));
                }
                break;
            }
            default: o[0] = node(nd.kind, nd.a, nd.b); break;  // base leaves as they are
        }
    }
    // Asserts: a base value folds as itself, an extension value as its
    // coordinates with the factors X^e (zero coordinates drop out).
    std::vector<LAssert> las;
    for (size_t j = 0; j < nasserts; j++) {
        const uint32_t k = asserts[j];
        for (uint32_t e = 0; e < (is_ext[k] ? 4u : 1u); e++)
            if (lo[k][e] != NONE) las.push_back(LAssert{(uint32_t)lo[k][e], e, (uint32_t)j});
    }
    return air_compile_lowered(field, ln, las, (uint32_t)nasserts, out, nregs_out, ninstr_out);
}

static int air_compile_lowered(int field, const std::vector<pg_air_node>& lnodes, const std::vector<LAssert>& las,
                               uint32_t norig, pg_air_program** out, uint32_t* nregs_out, uint32_t* ninstr_out) {
    const pg_air_node* const nodes = lnodes.data();
    const size_t nnodes = lnodes.size(), nasserts = las.size();
    // Leaves (columns, constants, selectors, public values) are operands read
    // in place: columns from staged windows, the rest from the constant table
    // or the selectors.  Only inner nodes take registers, so the register
    // count follows the depth of the expressions, not the columns read.
    auto is_leaf = [&](uint32_t k) {
        const uint32_t kd = nodes[k].kind;
        return kd != PG_AIR_ADD && kd != PG_AIR_SUB && kd != PG_AIR_NEG && kd != PG_AIR_MUL;
    };
    auto is_binary = [&](uint32_t k) {
        return nodes[k].kind == PG_AIR_ADD || nodes[k].kind == PG_AIR_SUB || nodes[k].kind == PG_AIR_MUL;
    };
    // emission order: per assert, its not yet computed inner nodes (operands
    // first), then the fold
    struct Item {
        bool is_assert;
        uint32_t node;
    };
    std::vector<Item> order;
    std::vector<uint8_t> state(nnodes, 0);  // 0 new, 1 on stack, 2 emitted
    std::vector<uint32_t> stack;
    for (size_t j = 0; j < nasserts; j++) {
        if (!is_leaf(las[j].node)) stack.push_back(las[j].node);
        while (!stack.empty()) {
            const uint32_t k = stack.back();
            if (state[k] == 2) {
                stack.pop_back();
                continue;
            }
            const pg_air_node& nd = nodes[k];
            if (state[k] == 0) {
                state[k] = 1;
                if (is_binary(k) && !is_leaf(nd.b) && state[nd.b] != 2) stack.push_back(nd.b);
                if (!is_leaf(nd.a) && state[nd.a] != 2) stack.push_back(nd.a);
                continue;
            }
            state[k] = 2;
            stack.pop_back();
            order.push_back(Item{false, k});
        }
        order.push_back(Item{true, las[j].node});
    }
    // last reader of every computed inner node
    std::vector<size_t> last_use(nnodes, 0);
    for (size_t t = 0; t < order.size(); t++) {
        const uint32_t k = order[t].node;
        if (order[t].is_assert) {
            last_use[k] = t;
        } else {
            if (!is_leaf(nodes[k].a)) last_use[nodes[k].a] = t;
            if (is_binary(k) && !is_leaf(nodes[k].b)) last_use[nodes[k].b] = t;
        }
    }
    // A value used only by the item right after the one computing it is
    // forwarded in a VGPR and takes no register; any value is forwarded to
    // that next item.
    std::vector<size_t> item_of(nnodes, SIZE_MAX);
    for (size_t t = 0; t < order.size(); t++)
        if (!order[t].is_assert) item_of[order[t].node] = t;
    auto forward_only = [&](uint32_t k) { return last_use[k] == item_of[k] + 1; };
    // the program for S window slots
    auto build = [&](uint32_t S) {
        std::unique_ptr<pg_air_program> prog(new pg_air_program);
        prog->field = field;
        prog->nasserts = (uint32_t)nasserts;
        prog->norig = norig;
        for (const LAssert& la : las) prog->afactor.push_back({la.comp, la.orig});
        std::vector<uint32_t> reg(nnodes, 0), free_regs;
        auto take_reg = [&]() {
            if (free_regs.empty()) return prog->nregs++;
            auto it = std::min_element(free_regs.begin(), free_regs.end());
            const uint32_t r = *it;
            free_regs.erase(it);
            return r;
        };
        // Symbolic operands, encoded once the register count (which places
        // the windows after the registers) and the constant count are known.
        enum : uint8_t { OP_REG, OP_WIN, OP_WIN_NEXT, OP_CONST, OP_PUBLIC, OP_SEL, OP_CHAL, OP_RSUM, OP_GSUM };
        enum : uint32_t { P_STAGE, P_ADD, P_SUB, P_MUL, P_NEG, P_ASSERT };  // pending instructions
        enum : uint32_t { S_FIRST, S_LAST, S_TRANS, S_PREV };               // OP_SEL operands
        struct Op {
            uint8_t kind = OP_REG;
            uint32_t x = 0;  // register, window word offset, constant, public or selector index
        };
        struct Pend {
            uint32_t op, dst;
            Op a, b;
        };
        std::vector<Pend> code;
        std::map<uint32_t, uint32_t> const_index;
        auto to_m = [&](uint32_t canon) {
            return field == PG_KOALABEAR ? to_monty<KoalaBear>(canon % KoalaBear::P)
                                         : to_monty<BabyBear>(canon % BabyBear::P);
        };
        // Staged windows: least recently used replaced; `keep` is a slot the
        // instruction being emitted already reads (never replaced under it).
        struct Win {
            bool used = false;
            uint32_t src = 0, col0 = 0;
            uint64_t stamp = 0;
        };
        Win win[QUOTIENT_SLOTS];
        uint64_t clock = 0;
        auto window = [&](uint32_t src, uint32_t col, int keep) {
            const uint32_t col0 = col / QW * QW;
            int sl = -1;
            for (int k = 0; k < (int)S; k++)
                if (win[k].used && win[k].src == src && win[k].col0 == col0) sl = k;
            if (sl < 0) {
                for (int k = 0; k < (int)S && sl < 0; k++)
                    if (!win[k].used) sl = k;
                if (sl < 0)
                    for (int k = 0; k < (int)S; k++)
                        if (k != keep && (sl < 0 || win[k].stamp < win[sl].stamp)) sl = k;
                win[sl].used = true;
                win[sl].src = src;
                win[sl].col0 = col0;
                Pend st{P_STAGE, (uint32_t)sl, {}, {}};
                st.a.x = col0;
                st.b.x = src;
                code.push_back(st);
                prog->nstages++;
                prog->nslots = std::max(prog->nslots, (uint32_t)sl + 1);
            }
            win[sl].stamp = ++clock;
            return sl;
        };
        // operand for node k read at item t; `slot` receives the window slot
        // it reads (or -1)
        auto operand = [&](uint32_t k, size_t t, int keep, int* slot) {
            const pg_air_node& nd = nodes[k];
            Op o;
            *slot = -1;
            if (!is_leaf(k)) {
                if (item_of[k] + 1 == t) {
                    o.kind = OP_SEL;
                    o.x = S_PREV;
                } else {
                    o.kind = OP_REG;
                    o.x = reg[k];
                }
                return o;
            }
            switch (nd.kind) {
                case PG_AIR_MAIN:
                case PG_AIR_PREPROCESSED:
                case LK_PERM: {
                    uint32_t src = 0;  // window source: main, preprocessed or permutation rows
                    if (nd.kind == PG_AIR_PREPROCESSED) {
                        prog->reads_pre = true;
                        prog->max_pre_col = std::max(prog->max_pre_col, nd.b + 1);
                        src = 2;
                    } else if (nd.kind == LK_PERM) {
                        prog->reads_perm = true;
                        prog->max_perm_col = std::max(prog->max_perm_col, nd.b + 1);
                        src = 4;
                    } else {
                        prog->max_main_col = std::max(prog->max_main_col, nd.b + 1);
                    }
                    const int sl = window(src + nd.a, nd.b, keep);
                    *slot = sl;
                    o.kind = nd.a ? OP_WIN_NEXT : OP_WIN;
                    o.x = (uint32_t)sl * QSLOT + (nd.b % QW) * QRS;
                    return o;
                }
                case PG_AIR_PUBLIC:
                    prog->max_public = std::max(prog->max_public, nd.b + 1);
                    o.kind = OP_PUBLIC;
                    o.x = nd.b;
                    return o;
                case LK_CHAL:
                    prog->max_chal_word = std::max(prog->max_chal_word, nd.b + 1);
                    o.kind = OP_CHAL;
                    o.x = nd.b;
                    return o;
                case LK_RSUM:
                    prog->reads_rsum = true;
                    o.kind = OP_RSUM;
                    o.x = nd.b;
                    return o;
                case LK_GSUM:
                    prog->max_gsum = std::max(prog->max_gsum, nd.b + 1);
                    o.kind = OP_GSUM;
                    o.x = nd.b;
                    return o;
                case PG_AIR_IS_FIRST_ROW:
                case PG_AIR_IS_LAST_ROW:
                case PG_AIR_IS_TRANSITION:
                    prog->need_selectors = true;
                    o.kind = OP_SEL;
                    o.x = nd.kind == PG_AIR_IS_FIRST_ROW ? S_FIRST : nd.kind == PG_AIR_IS_LAST_ROW ? S_LAST : S_TRANS;
                    return o;
                default: {
                    const uint32_t v = to_m(nd.a);
                    auto it = const_index.find(v);
                    if (it == const_index.end()) {
                        it = const_index.emplace(v, (uint32_t)prog->consts.size()).first;
                        prog->consts.push_back(v);
                    }
                    o.kind = OP_CONST;
                    o.x = it->second;
                    return o;
                }
            }
        };
        for (size_t t = 0; t < order.size(); t++) {
            const uint32_t k = order[t].node;
            const pg_air_node& nd = nodes[k];
            int sa = -1, sb = -1;
            if (order[t].is_assert) {
                Pend in{P_ASSERT, QO_NO_DST, operand(k, t, -1, &sa), {}};
                code.push_back(in);
                if (!is_leaf(k) && last_use[k] == t && !forward_only(k)) free_regs.push_back(reg[k]);
                continue;
            }
            const Op oa = operand(nd.a, t, -1, &sa);
            const Op ob = is_binary(k) ? operand(nd.b, t, sa, &sb) : Op{};
            const uint32_t op = nd.kind == PG_AIR_ADD ? P_ADD
                                : nd.kind == PG_AIR_SUB ? P_SUB
                                : nd.kind == PG_AIR_MUL ? P_MUL
                                                        : P_NEG;
            // operands are read before the result is written: registers whose
            // values die here may take the result
            if (!is_leaf(nd.a) && last_use[nd.a] == t && !forward_only(nd.a)) free_regs.push_back(reg[nd.a]);
            if (is_binary(k) && nd.b != nd.a && !is_leaf(nd.b) && last_use[nd.b] == t && !forward_only(nd.b))
                free_regs.push_back(reg[nd.b]);
            const uint32_t dst = forward_only(k) ? QO_NO_DST : (reg[k] = take_reg());
            code.push_back(Pend{op, dst, oa, ob});
        }
        // the LDS layout (registers, windows, selectors, constants) and the
        // device encoding: operand kinds folded into the opcode
        const uint32_t nconst = (uint32_t)prog->consts.size();
        const uint32_t wbase = prog->nregs * QB;
        const uint32_t sbase = wbase + prog->nslots * QSLOT;
        const uint32_t cbase = sbase + (prog->need_selectors ? 3u * QB : 0u);
        // the constant table at launch: constants, public values, challenge
        // words, the regional sum, the global-sum coordinates
        const uint32_t pub0 = cbase + nconst, chal0 = pub0 + prog->max_public;
        const uint32_t rsum0 = chal0 + prog->max_chal_word, gsum0 = rsum0 + (prog->reads_rsum ? 4u : 0u);
        auto kind = [&](const Op& o) -> uint32_t {
            switch (o.kind) {
                case OP_WIN_NEXT: return QK_NEXT;
                case OP_CONST:
                case OP_PUBLIC:
                case OP_CHAL:
                case OP_RSUM:
                case OP_GSUM: return QK_BCAST;
                case OP_SEL: return o.x == S_PREV ? QK_PREV : QK_ROW;
                default: return QK_ROW;
            }
        };
        auto index = [&](const Op& o) -> uint32_t {
            switch (o.kind) {
                case OP_REG: return o.x * QB;
                case OP_WIN:
                case OP_WIN_NEXT: return wbase + o.x;
                case OP_CONST: return cbase + o.x;
                case OP_PUBLIC: return pub0 + o.x;
                case OP_CHAL: return chal0 + o.x;
                case OP_RSUM: return rsum0 + o.x;
                case OP_GSUM: return gsum0 + o.x;
                default: return o.x == S_PREV ? 0u : sbase + o.x * QB;
            }
        };
        prog->code.reserve(code.size());
        for (const Pend& in : code) {
            const uint32_t dst = in.dst == QO_NO_DST ? QO_NO_DST : in.dst * QB;
            switch (in.op) {
                case P_STAGE: prog->code.push_back(QInstr{QOP_STAGE, in.dst, in.a.x, in.b.x}); break;
                case P_NEG: prog->code.push_back(QInstr{QOP_NEG + kind(in.a), dst, index(in.a), 0}); break;
                case P_ASSERT: prog->code.push_back(QInstr{QOP_ASSERT + kind(in.a), 0, index(in.a), 0}); break;
                default: {
                    const uint32_t op = in.op == P_ADD ? 0u : in.op == P_SUB ? 1u : 2u;
                    prog->code.push_back(
                        QInstr{16 * op + 4 * kind(in.a) + kind(in.b), dst, index(in.a), index(in.b)});
                }
            }
        }
        return prog;
    };
    // Fewest window slots (the most workgroups per CU) that stage at most
    // 1/8 more windows than the most slots do.
    std::unique_ptr<pg_air_program> prog = build(QUOTIENT_SLOTS);
    for (uint32_t S = 2; S < QUOTIENT_SLOTS; S++) {
        std::unique_ptr<pg_air_program> p = build(S);
        if (p->nstages * 8 <= prog->nstages * 9) {
            prog = std::move(p);
            break;
        }
    }
    if (prog->nregs > QUOTIENT_MAX_REGS)
        return fail(PG_E_UNSUPPORTED, "the program needs %u registers; at most %u are supported", prog->nregs,
                    QUOTIENT_MAX_REGS);
    if (nregs_out) *nregs_out = prog->nregs;
    if (ninstr_out) *ninstr_out = (uint32_t)prog->code.size();
    *out = prog.release();
    return PG_OK;
}

template <class F>
static int quotient_t(pg_ctx* ctx, hipStream_t s, const pg_air_program* prog, const pg_buf* mm, const pg_buf* pm,
                      const pg_buf* qm, int log_n, uint32_t qdb, const uint32_t* publics, const uint32_t* chal,
                      const uint32_t* rsum, const uint32_t* gsum, const uint32_t* alpha_w, pg_buf** out) {
    const int log_q = log_n + (int)qdb;
    const uint32_t q = 1u << log_q;
    // constant table (the program's constants, then the public values, the
    // challenge words, the regional sum, the global-sum coordinates), the Z_H
    // table and the powers of alpha
    std::vector<uint32_t> consts = prog->consts;
    consts.insert(consts.end(), publics, publics + prog->max_public);
    consts.insert(consts.end(), chal, chal + prog->max_chal_word);
    if (prog->reads_rsum) consts.insert(consts.end(), rsum, rsum + 4);
    consts.insert(consts.end(), gsum, gsum + prog->max_gsum);
    const uint32_t nz = 1u << qdb;
    std::vector<uint32_t> zh(2 * nz);
    const uint32_t gen = to_monty<F>(F::GEN);
    const uint32_t gn = fpow<F>(gen, 1ull << log_n);  // GEN^n
    const uint32_t wq = to_monty<F>(two_adic_gen_canon<F>((int)qdb));
    uint32_t x = gn;
    for (uint32_t j = 0; j < nz; j++) {  // x^n at x_i, i = j (mod 2^qdb): GEN^n w_{2^qdb}^j
        zh[j] = fsub<F>(x, F::R1);
        zh[nz + j] = finv<F>(zh[j]);
        x = fmul<F>(x, wq);
    }
    if (quotient_lds_bytes(prog->nregs, prog->nslots, prog->need_selectors, (uint32_t)consts.size()) > 65536)
        return fail(PG_E_UNSUPPORTED, "the program's registers, windows and %zu constants need more than 64 KiB of LDS",
                    consts.size());
    // base assert i folds coordinate e of original assert j: X^e alpha^(m-1-j)
    const uint32_t m = prog->nasserts;
    std::vector<uint32_t> apow((size_t)m * 4);
    {
        std::vector<Ext4> pw(prog->norig);
        const Ext4 al = ext_from<F>(alpha_w);
        Ext4 cur{{F::R1, 0u, 0u, 0u}};
        for (uint32_t j = prog->norig; j-- > 0;) {
            pw[j] = cur;
            cur = emul<F>(cur, al);
        }
        const uint32_t W = to_monty<F>(F::P == KoalaBear::P ? 3u : 11u);  // X^4 = W
        for (uint32_t i = 0; i < m; i++) {
            Ext4 v = pw[prog->afactor[i].second];
            for (uint32_t e = 0; e < prog->afactor[i].first; e++)  // times X
                v = Ext4{{fmul<F>(v.c[3], W), v.c[0], v.c[1], v.c[2]}};
            for (int c = 0; c < 4; c++) apow[(size_t)i * 4 + c] = v.c[c];
        }
    }
    // the kernel reads one instruction past the end (zero padding)
    const size_t b_code = round_up((prog->code.size() + 1) * sizeof(QInstr), 256), b_zh = round_up(zh.size() * 4, 256),
                 b_const = round_up(consts.size() * 4, 256), b_pow = apow.size() * 4;
    int rc;
    std::shared_ptr<DevMem> tw;
    PG_TRY(get_twiddles(ctx, s, prog->field, log_q, &tw));
    auto desc = make_devmem(ctx, b_code + b_zh + b_const + b_pow + 256, &rc);
    if (!desc) return rc;
    pg_buf* o;
    PG_TRY(buf_new(ctx, q, 4, &o));
    SyncGuard guard(s);
    uint8_t* dd = (uint8_t*)desc->ptr;
    hipError_t e = hipSuccess;
    auto put = [&](size_t off, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(dd + off, src, bytes, hipMemcpyHostToDevice, s);
    };
    std::vector<QInstr> code = prog->code;
    code.push_back(QInstr{QOP_STAGE, 0, 0, 0});
    put(0, code.data(), code.size() * sizeof(QInstr));
    put(b_code, zh.data(), zh.size() * 4);
    put(b_code + b_zh, consts.data(), consts.size() * 4);
    put(b_code + b_zh + b_const, apow.data(), b_pow);
    QuotientArgs a{};
    a.main = mm->ptr;
    a.ld_main = mm->ld;
    a.w_main = (uint32_t)mm->cols;
    a.pre = pm ? pm->ptr : nullptr;
    a.ld_pre = pm ? pm->ld : 0;
    a.w_pre = pm ? (uint32_t)pm->cols : 0;
    a.perm = qm ? qm->ptr : nullptr;
    a.ld_perm = qm ? qm->ld : 0;
    a.w_perm = qm ? (uint32_t)qm->cols : 0;
    a.prog = (const QInstr*)dd;
    a.nprog = (uint32_t)prog->code.size();
    a.nregs = prog->nregs;
    a.nslots = prog->nslots;
    a.zh = (const uint32_t*)(dd + b_code);
    a.consts = (const uint32_t*)(dd + b_code + b_zh);
    a.nconsts = (uint32_t)consts.size();
    a.nasserts = m;
    a.apow = (const uint32_t*)(dd + b_code + b_zh + b_const);
    a.tw = (const uint32_t*)tw->ptr;
    a.zh_mask = nz - 1;
    a.q = q;
    a.log_q = (uint32_t)log_q;
    a.next_step = nz;
    a.shift = gen;
    a.omega_n_inv = finv<F>(to_monty<F>(two_adic_gen_canon<F>(log_n)));
    a.need_selectors = prog->need_selectors;
    a.out = o->ptr;
    auto kt = make_timer(ctx);
    if (e == hipSuccess) {
        if (kt) kt->begin("quotient", s);
        e = launch_quotient(prog->field, a, s);
        if (kt) kt->end(s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        pg_buf_free(o);
        return fail(PG_E_DEVICE, "quotient evaluation failed: %s", hipGetErrorString(e));
    }
    guard.armed = false;
    *out = o;
    return PG_OK;
}

extern "C" int pg_quotient_values_perm(pg_ctx* ctx, const pg_air_program* prog, const pg_tree* main, size_t main_k,
                                       const pg_tree* pre, size_t pre_k, const pg_tree* perm, size_t perm_k,
                                       uint32_t log_blowup, uint32_t log_qd, const uint32_t* publics, size_t npublics,
                                       const uint32_t* chal, size_t nchal, const uint32_t rsum[4],
                                       const uint32_t* gsum, size_t nglobal, const uint32_t alpha[4], pg_buf** out) {
    if (!ctx || !prog || !main || !alpha || !out || main_k >= main->mats.size())
        return fail(PG_E_INVALID, "bad argument");
    *out = nullptr;
    if (main->field != prog->field) return fail(PG_E_INVALID, "the program and the trace are over different fields");
    const pg_buf* mm = main->mats[main_k];
    if (!is_pow2(mm->rows) || log2_of(mm->rows) < (int)log_blowup)
        return fail(PG_E_INVALID, "main matrix height %zu has no trace domain at log_blowup %u", mm->rows, log_blowup);
    const int log_n = log2_of(mm->rows) - (int)log_blowup;
    if (log_qd > log_blowup)
        return fail(PG_E_UNSUPPORTED, "log_quotient_degree %u > log_blowup %u: the quotient domain is not inside "
                                      "the committed LDE", log_qd, log_blowup);
    if (log_n + (int)log_qd > 32 - 1) return fail(PG_E_UNSUPPORTED, "quotient domain too large");
    if (prog->max_main_col > mm->cols)
        return fail(PG_E_INVALID, "the program reads main column %u of %zu", prog->max_main_col - 1, mm->cols);
    // the preprocessed and permutation traces: committed at the same height
    auto side = [&](bool reads, const pg_tree* t, size_t k, uint32_t max_col, const char* what,
                    const pg_buf** m) -> int {
        *m = nullptr;
        if (!reads) return PG_OK;
        if (!t || k >= t->mats.size()) return fail(PG_E_INVALID, "the program reads a %s trace", what);
        if (t->field != prog->field) return fail(PG_E_INVALID, "%s trace over another field", what);
        *m = t->mats[k];
        if ((*m)->rows != mm->rows)
            return fail(PG_E_INVALID, "%s height %zu differs from the main trace's %zu", what, (*m)->rows, mm->rows);
        if (max_col > (*m)->cols)
            return fail(PG_E_INVALID, "the program reads %s column %u of %zu", what, max_col - 1, (*m)->cols);
        return PG_OK;
    };
    const pg_buf *pm, *qm;
    PG_TRY(side(prog->reads_pre, pre, pre_k, prog->max_pre_col, "preprocessed", &pm));
    PG_TRY(side(prog->reads_perm, perm, perm_k, prog->max_perm_col, "permutation", &qm));
    if (prog->max_public > npublics || (npublics && !publics))
        return fail(PG_E_INVALID, "the program reads public value %u of %zu", prog->max_public - 1, npublics);
    if (prog->max_chal_word > 4 * nchal || (nchal && !chal))
        return fail(PG_E_INVALID, "the program reads challenge %u of %zu", (prog->max_chal_word - 1) / 4, nchal);
    if (prog->reads_rsum && !rsum) return fail(PG_E_INVALID, "the program reads the regional cumulative sum");
    if (prog->max_gsum > nglobal || (nglobal && !gsum))
        return fail(PG_E_INVALID, "the program reads global-sum coordinate %u of %zu", prog->max_gsum - 1, nglobal);
    PG_TRY(bind_device(ctx));
    Lease l(ctx);
    return prog->field == PG_KOALABEAR
               ? quotient_t<KoalaBear>(ctx, l.s, prog, mm, pm, qm, log_n, log_qd, publics, chal, rsum, gsum, alpha, out)
               : quotient_t<BabyBear>(ctx, l.s, prog, mm, pm, qm, log_n, log_qd, publics, chal, rsum, gsum, alpha, out);
}

extern "C" int pg_quotient_values(pg_ctx* ctx, const pg_air_program* prog, const pg_tree* main, size_t main_k,
                                  const pg_tree* pre, size_t pre_k, uint32_t log_blowup, uint32_t log_qd,
                                  const uint32_t* publics, size_t npublics, const uint32_t alpha[4], pg_buf** out) {
    return pg_quotient_values_perm(ctx, prog, main, main_k, pre, pre_k, nullptr, 0, log_blowup, log_qd, publics,
                                   npublics, nullptr, 0, nullptr, nullptr, 0, alpha, out);
}
