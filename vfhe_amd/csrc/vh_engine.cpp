// vh_engine.cpp -- host runtime of the hip engine.
//
// Implements the reference's C ABI (include/vfhe_hip.h cites each function's
// reference counterpart) over device-resident polynomials.  Host code only
// builds work lists and constant tables; all coefficient arithmetic runs in
// the kernels of vh_kernels.hip, enqueued on one in-order HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <map>
#include <tuple>
#include <mutex>
#include <sys/random.h>
#include <vector>

#include "../../include/vfhe_hip.h"
#include "vh_kernels.h"

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// Errors: the reference aborts on a failed assert (live in release builds,
// meson.build:17) and exits on allocation failure (misc_tp.c:37-46).
// ---------------------------------------------------------------------------
[[noreturn]] static void vh_die(const char *what, const char *file, int line)
{
    fprintf(stderr, "vfhe hip engine: %s (%s:%d)\n", what, file, line);
    fflush(stderr);
    abort();
}
[[noreturn]] static void vh_hip_fail(hipError_t e, const char *call, const char *file, int line)
{
    fprintf(stderr, "vfhe hip engine: %s failed: %s (%s:%d)\n", call, hipGetErrorString(e), file, line);
    fflush(stderr);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) exit(EXIT_FAILURE);
    abort();
}
#define VH_CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) vh_hip_fail(e_, #x, __FILE__, __LINE__); } while (0)
#define VH_REQUIRE(c) do { if (!(c)) vh_die("precondition failed: " #c, __FILE__, __LINE__); } while (0)
#define ACTIVE(mask, i) ((((mask) >> (i)) & 1ULL) != 0)

// ---------------------------------------------------------------------------
// Graph capture state (vfhe_hip_mlwe_multiply_batch_graph).  A plan holds one
// batch shape's captured launch sequence.  Modes:
//   LOG     the batch runs eagerly; the sizes of its staged work lists and
//           device allocations are logged in order;
//   RECORD  the host logic runs again under stream capture: work lists are
//           written into the plan's pinned buffer (copied to the device by the
//           graph's first node), allocations come from the plan's own blocks,
//           frees are no-ops;
//   DRY     the host logic runs once more to rebuild the work lists for new
//           operands, with every launch skipped; the graph is then replayed.
// A staged size, allocation size or launch count that differs from the plan
// raises vh_plan_mismatch (the caller falls back to the eager path).
// ---------------------------------------------------------------------------
enum { VH_CAP_OFF = 0, VH_CAP_LOG = 1, VH_CAP_RECORD = 2, VH_CAP_DRY = 3 };
struct vh_plan_mismatch {};
struct vh_graph_plan {
    std::vector<uint64_t> key;
    std::vector<size_t> stage_sz, block_sz;
    std::vector<void *> blocks;
    size_t wl_bytes = 0;
    char *h_wl = nullptr, *d_wl = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    long launches = 0;
};
static int g_cap = VH_CAP_OFF;
static vh_graph_plan *g_plan = nullptr;
static size_t g_cap_stage = 0, g_cap_off = 0, g_cap_block = 0;
static long g_cap_launches = 0;
static std::vector<vh_graph_plan *> g_plans;
static long g_graph_captures = 0, g_graph_replays = 0, g_graph_fallbacks = 0;
// Why the last plan mismatch happened: kind * 1000000 + detail (1 staged work list
// size at index, 2 device block size at index, 3 upload under capture, 4 launch
// count, 5 staged list count, 6 block count).
static long g_graph_why = 0, g_graph_why_want = 0, g_graph_why_got = 0;
// In DRY mode the launch expression still runs (its work lists are staged into the
// plan's buffer) but the launch functions launch nothing (g_vh_suppress_launch).
#define VH_LAUNCH(x) do { if (g_cap >= VH_CAP_RECORD) g_cap_launches++; VH_CHECK(x); } while (0)

// ---------------------------------------------------------------------------
// Engine-private tails of the public structs
// ---------------------------------------------------------------------------
struct _NTT_proc {
    uint64_t n, q, root_of_unity, inv_root_of_unity;  // as the reference's first fields
    vh_limb h;          // host copy; its table pointers are device addresses
    vh_limb *d_self;    // device copy (a one-entry limb table)
    void *d_tab;        // device tables
};

struct vh_pool {
    struct _incNTT pub;
    vh_limb *d_limbs;                // device limb table, one entry per pool prime
    std::vector<vh_limb> h_limbs;    // host copy
    std::vector<uint64_t *> d_w;     // device w rows (split_degree > 1 only)
    uint64_t n, logn;                // transform size N/split and its log
    // division constants per (active mask, divide-prime chunk, rounding), built once
    std::map<std::tuple<uint64_t, uint64_t, int>, void *> div_consts;
};

struct vh_poly {
    struct _RNS_Polynomial pub;
    uint64_t *d;        // device rows [allocated_l][N]
    uint64_t *hblock;   // host mirror block behind coeffs[]
    int host_dirty;
    int borrowed;       // a view: d and hblock belong to another polynomial
};

static inline vh_poly *VP(RNS_Polynomial p) { return reinterpret_cast<vh_poly *>(p); }
static inline vh_pool *VPOOL(incNTT t) { return reinterpret_cast<vh_pool *>(t); }

// ---------------------------------------------------------------------------
// Runtime context: one device, one stream, a staging ring for work lists.
// ---------------------------------------------------------------------------
// A lane = one HIP stream (default priority) with its own staging ring for work
// lists.  Lane 0 is the engine's; the serving runtime runs batches on others so
// independent requests overlap on the device.  A thread selects its lane with
// vfhe_hip_set_lane; every launch and copy it makes goes to that lane's stream.
#define VH_MAX_LANES 8
struct vh_lane {
    hipStream_t stream = nullptr;
    char *h_ring = nullptr, *d_ring = nullptr;
    char *hd_ring = nullptr;   // device address of the pinned host ring (small arrays are read in place)
    size_t ring_size = 0, ring_off = 0;
    int cur_half = 0;
    hipEvent_t half_ev[2];
    bool half_used[2] = {false, false};
};

struct vh_ctx {
    bool ready = false;
    int device = 0;
    vh_lane lanes[VH_MAX_LANES];
    char name[128] = {0};
    int cus = 0;
    unsigned *d_err = nullptr;   // device error word written by kernels with bounded waits
    bool xcd_ok = false;         // the placement probe saw exactly the XCC ids 0..7
    unsigned xcd_hist[8] = {0};
    unsigned pipe_blocks = 0;    // persistent workgroups of the pipelined NTT pass (CUs x occupancy)
};
static vh_ctx C;
// Device address of each device's scratch array (vh_kernels.hip), resolved at set-up.
static void *g_scratch_of[64] = {nullptr};
// While an allocation-free call runs: work lists and scratch blocks are carved from
// g_na_base (that device's scratch array) by a bump offset, nothing is allocated.
static char *g_na_base = nullptr;
static size_t g_na_off = 0;
static thread_local int g_lane = 0;
static inline vh_lane &LN() { return C.lanes[g_lane]; }
static inline hipStream_t S() { return C.lanes[g_lane].stream; }

static void lane_init(vh_lane &L, size_t ring)
{
    VH_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    L.ring_size = ring;
    VH_CHECK(hipHostMalloc((void **)&L.h_ring, L.ring_size, hipHostMallocMapped));
    VH_CHECK(hipHostGetDevicePointer((void **)&L.hd_ring, L.h_ring, 0));
    VH_CHECK(hipMalloc((void **)&L.d_ring, L.ring_size));
    VH_CHECK(hipEventCreateWithFlags(&L.half_ev[0], hipEventDisableTiming));
    VH_CHECK(hipEventCreateWithFlags(&L.half_ev[1], hipEventDisableTiming));
}

static void sync_all_lanes()
{
    for (auto &L : C.lanes)
        if (L.stream) VH_CHECK(hipStreamSynchronize(L.stream));
}

// Tunables (vfhe_hip_set_option): NTT schedule for N = 2^13..2^16 on the FP64 path.
struct vh_opts {
    long ntt_mode = 2;        // 0: two launches per direction, 1: persistent per-XCD kernel,
                              // 2: two launches of the pipelined pass kernel
    long ntt_chunk_mb = 0;    // two-launch mode: rows per sub-batch (MiB), 0 = whole batch
    long ntt_col_wide = 1;    // pipelined column passes: 512-thread workgroups over 32-column tiles
    long ntt_tile_grid = 15;   // pipelined passes launched one tile per workgroup instead of
                              // persistent: bit 0 forward column, 1 forward row, 2 inverse row,
                              // 3 inverse column, 4 the fused product's row kernel
    long xcd_lag = 3;         // persistent mode: queue slots between an item's two passes
    long xcd_occ = 2;         // persistent mode: workgroups per CU (2 or 4)
    long xcd_l2 = 1;          // persistent mode: the first pass stores its intermediate with
                              // ordinary (L2-allocating) stores; 0 streams it non-temporally
    long spin_limit = 1 << 22;
    long ks_fused = 1;        // key-switch: fuse the forward row pass with the key inner product
    long mdr_fp64 = 1;        // ModDown + rescale on the FP64 path when every limb allows it
    long ntt_small = 1;       // N = 2^9 .. 2^12 (FP64): whole transforms in registers (k_ntt_small);
                              // 0: the LDS radix-2 kernel
    long br_fused = 1;        // CGGI blind rotation at N = 1024: every key index in one launch
                              // (0: a launch sequence per index)
    long ks_fold = 1;         // multiply: pass-through components enter the fused key inner
                              // product as P * d (NTT form) instead of after the ModDown
    long fft_fma = 1;         // complex FFT: 1 = avx512ifma engine's fused complex product, 0 = portable's
};
static vh_opts OPT;
static std::recursive_mutex vh_mu;

static void vh_ensure()
{
    if (C.ready) return;
    int n = 0;
    VH_CHECK(hipGetDeviceCount(&n));
    VH_REQUIRE(n > 0);
    VH_CHECK(hipGetDevice(&C.device));
    hipDeviceProp_t prop;
    VH_CHECK(hipGetDeviceProperties(&prop, C.device));
    snprintf(C.name, sizeof(C.name), "%s", prop.gcnArchName);
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, C.device) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
    lane_init(C.lanes[0], (size_t)64 << 20);
    C.cus = prop.multiProcessorCount;
    C.pipe_blocks = (unsigned)(C.cus * vh_ntt16p_occupancy());
    if (const char *e = getenv("VFHE_HIP_NTT_MODE")) OPT.ntt_mode = atol(e);
    if (const char *e = getenv("VFHE_HIP_NTT_CHUNK_MB")) OPT.ntt_chunk_mb = atol(e);
    if (const char *e = getenv("VFHE_HIP_NTT_TILE_GRID")) OPT.ntt_tile_grid = atol(e);
    if (const char *e = getenv("VFHE_HIP_NTT_COL_WIDE")) OPT.ntt_col_wide = atol(e);
    if (const char *e = getenv("VFHE_HIP_XCD_LAG")) OPT.xcd_lag = atol(e);
    if (const char *e = getenv("VFHE_HIP_XCD_OCC")) OPT.xcd_occ = atol(e);
    if (const char *e = getenv("VFHE_HIP_XCD_L2")) OPT.xcd_l2 = atol(e);
    if (const char *e = getenv("VFHE_HIP_KS_FUSED")) OPT.ks_fused = atol(e);
    if (const char *e = getenv("VFHE_HIP_MDR_FP64")) OPT.mdr_fp64 = atol(e);
    if (const char *e = getenv("VFHE_HIP_KS_FOLD")) OPT.ks_fold = atol(e);
    if (const char *e = getenv("VFHE_HIP_NTT_SMALL")) OPT.ntt_small = atol(e);
    if (const char *e = getenv("VFHE_HIP_BR_FUSED")) OPT.br_fused = atol(e);
    if (const char *e = getenv("VFHE_HIP_FFT_FMA")) OPT.fft_fma = atol(e);
    VH_REQUIRE(C.device >= 0 && C.device < 64);
    VH_CHECK(vh_scratch_resolve(&g_scratch_of[C.device]));
    VH_CHECK(hipMalloc((void **)&C.d_err, 256));
    VH_CHECK(hipMemset(C.d_err, 0, 256));
    {
        // Placement probe for the persistent per-XCD NTT: it relies on every XCC id
        // 0..7 hosting workgroups (k_ntt_xcd_verify still guards each launch).
        const unsigned nb = (unsigned)C.cus * 4;
        unsigned *d = nullptr;
        VH_CHECK(hipMalloc((void **)&d, nb * sizeof(unsigned)));
        VH_LAUNCH(vh_launch_xcc_probe(S(), d, nb));
        std::vector<unsigned> h(nb);
        VH_CHECK(hipMemcpyAsync(h.data(), d, nb * sizeof(unsigned), hipMemcpyDeviceToHost, S()));
        VH_CHECK(hipStreamSynchronize(S()));
        VH_CHECK(hipFree(d));
        bool ok = true;
        for (unsigned x : h) { if (x < 8) C.xcd_hist[x]++; else ok = false; }
        for (int k = 0; k < 8; k++) ok = ok && C.xcd_hist[k] > 0;
        C.xcd_ok = ok;
    }
    C.ready = true;
}

#define VH_ENTER std::lock_guard<std::recursive_mutex> vh_lock_(vh_mu); vh_ensure()

// Copy a small host array to device memory, ordered on the stream.
static const size_t VH_STAGE_INPLACE = 65536;
static void *vh_stage(const void *src, size_t bytes)
{
    vh_lane &L = LN();
    const size_t need = (bytes + 255) & ~(size_t)255, half = L.ring_size / 2;
    if (g_cap == VH_CAP_LOG) {
        g_plan->stage_sz.push_back(need);
    } else if (g_cap >= VH_CAP_RECORD) {
        if (g_cap_stage >= g_plan->stage_sz.size() || g_plan->stage_sz[g_cap_stage] != need) {
            g_graph_why = 1000000L + (long)g_cap_stage;
            g_graph_why_want = g_cap_stage < g_plan->stage_sz.size() ? (long)g_plan->stage_sz[g_cap_stage] : -1;
            g_graph_why_got = (long)need;
            throw vh_plan_mismatch();
        }
        memcpy(g_plan->h_wl + g_cap_off, src, bytes);
        char *d = g_plan->d_wl + g_cap_off;
        g_cap_off += need;
        g_cap_stage++;
        return d;
    }
    VH_REQUIRE(need <= half);
    if (L.ring_off + need > (size_t)(L.cur_half + 1) * half) {
        VH_CHECK(hipEventRecord(L.half_ev[L.cur_half], L.stream));
        L.half_used[L.cur_half] = true;
        const int nh = 1 - L.cur_half;
        if (L.half_used[nh]) VH_CHECK(hipEventSynchronize(L.half_ev[nh]));
        L.cur_half = nh;
        L.ring_off = (size_t)nh * half;
    }
    const size_t off = L.ring_off;
    memcpy(L.h_ring + off, src, bytes);
    L.ring_off += need;
    if (g_na_base) {
        // allocation-free call: the list goes to the device's scratch array
        VH_REQUIRE(g_na_off + need <= VH_SCRATCH_BYTES);
        char *d = g_na_base + g_na_off;
        g_na_off += need;
        VH_CHECK(hipMemcpyAsync(d, L.h_ring + off, bytes, hipMemcpyHostToDevice, L.stream));
        return d;
    }
    // Small work lists are read by the kernels straight from pinned host memory
    // (one read per workgroup); larger ones are copied to HBM first.
    if (bytes <= VH_STAGE_INPLACE) return L.hd_ring + off;
    VH_CHECK(hipMemcpyAsync(L.d_ring + off, L.h_ring + off, bytes, hipMemcpyHostToDevice, L.stream));
    return L.d_ring + off;
}

template <typename T> static T *vh_stage_vec(const std::vector<T> &v)
{
    return reinterpret_cast<T *>(vh_stage(v.data(), v.size() * sizeof(T)));
}

// ---------------------------------------------------------------------------
// Launch profiler: when on, every kernel launch is bracketed by two HIP events
// on the engine stream and its duration is summed per kernel family.
// ---------------------------------------------------------------------------
struct vh_prof_rec { int tag; hipEvent_t a, b; };
static bool g_prof = false;
static std::vector<vh_prof_rec> g_prof_recs;
static std::vector<hipEvent_t> g_prof_free;

static hipEvent_t prof_event()
{
    if (!g_prof_free.empty()) { hipEvent_t e = g_prof_free.back(); g_prof_free.pop_back(); return e; }
    hipEvent_t e;
    VH_CHECK(hipEventCreate(&e));
    return e;
}

struct vh_prof_scope {
    int tag; hipEvent_t a = nullptr, b = nullptr;
    explicit vh_prof_scope(int t) : tag(t)
    {
        if (!g_prof || g_cap >= VH_CAP_RECORD) return;
        a = prof_event(); b = prof_event();
        VH_CHECK(hipEventRecord(a, S()));
    }
    ~vh_prof_scope()
    {
        if (!a) return;
        VH_CHECK(hipEventRecord(b, S()));
        g_prof_recs.push_back({tag, a, b});
    }
};

// Device blocks: a cache in front of the stream-ordered allocator.  A freed
// block is parked, by size class, on the list of the lane that frees it, with
// an event recorded on that lane's stream at the free.  An allocation takes a
// parked block that fits within twice its class (best fit): from its own lane
// at once (the reuse is ordered after the free by the stream, as with
// hipFreeAsync + hipMallocAsync on one stream), from another lane once the
// free's event has completed.  Size classes: powers of two up to 1 MiB, then
// multiples of 1 MiB.  At most `cache.max_mb` (default 32 GiB) stays parked; a
// free that would exceed it first returns the least recently parked blocks to
// the allocator (hipFreeAsync on the lane that parked them, so the free stays
// ordered after that lane's last use), so a workload that changes its block
// sizes does not keep missing a cache full of the old sizes.  (hipMallocAsync /
// hipFreeAsync cost tens to hundreds of microseconds per call in a serving loop.)
static size_t g_cache_max = (size_t)32 << 30;
struct vh_parked {
    void *p;
    hipEvent_t ev;
    uint64_t seq;                       // park order (least recently parked first out)
};
struct vh_block_cache {
    std::map<size_t, std::vector<vh_parked>> free_[VH_MAX_LANES];
    std::map<void *, size_t> size_of;   // every live or parked block -> its class size
    std::map<uint64_t, std::pair<int, size_t>> order;   // seq -> (lane, class) of a parked block
    std::vector<hipEvent_t> events;     // idle events
    size_t cached = 0;
    uint64_t seq = 0;
};
static vh_block_cache BC;
// Allocation counters (vfhe_hip_get_option "stat.cache_*"): blocks handed out from the
// cache, blocks taken from the stream-ordered allocator, out-of-memory flushes of the
// cache (each a synchronisation of every lane under the engine lock).
static long g_cache_hits = 0, g_cache_misses = 0, g_cache_flushes = 0;

static size_t vh_size_class(size_t bytes)
{
    if (bytes <= 256) return 256;
    if (bytes <= ((size_t)1 << 20)) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t mib = (size_t)1 << 20;
    return (bytes + mib - 1) / mib * mib;
}

// Takes a block of class in [c, 2c] from lane `ln`; other lanes' blocks only when
// their free has completed on the device.
static void *vh_take_parked(int ln, size_t c)
{
    auto &lane = BC.free_[ln];
    for (auto it = lane.lower_bound(c); it != lane.end() && it->first <= 2 * c; ++it) {
        auto &v = it->second;
        for (size_t k = v.size(); k-- > 0;) {
            if (ln != g_lane) {
                const hipError_t q = hipEventQuery(v[k].ev);
                if (q == hipErrorNotReady) continue;
                VH_CHECK(q);
            }
            const vh_parked b = v[k];
            v.erase(v.begin() + (long)k);
            BC.order.erase(b.seq);
            BC.events.push_back(b.ev);
            BC.cached -= it->first;
            return b.p;
        }
    }
    return nullptr;
}

// Returns the least recently parked block to the allocator; false when nothing is parked.
static bool vh_evict_oldest()
{
    if (BC.order.empty()) return false;
    const auto o = BC.order.begin();
    const uint64_t seq = o->first;
    const int ln = o->second.first;
    const size_t c = o->second.second;
    BC.order.erase(o);
    auto &v = BC.free_[ln][c];
    for (size_t k = 0; k < v.size(); k++) {
        if (v[k].seq != seq) continue;
        VH_CHECK(hipFreeAsync(v[k].p, C.lanes[ln].stream));   // after the lane's last use of it
        BC.size_of.erase(v[k].p);
        BC.events.push_back(v[k].ev);
        v.erase(v.begin() + (long)k);
        break;
    }
    if (v.empty()) BC.free_[ln].erase(c);
    BC.cached -= c;
    return true;
}

static void *vh_dmalloc(size_t bytes)
{
    const size_t c = vh_size_class(bytes);
    if (g_cap == VH_CAP_LOG) {
        g_plan->block_sz.push_back(c);
    } else if (g_cap >= VH_CAP_RECORD) {
        if (g_cap_block >= g_plan->block_sz.size() || g_plan->block_sz[g_cap_block] != c) {
            g_graph_why = 2000000L + (long)g_cap_block;
            g_graph_why_want = g_cap_block < g_plan->block_sz.size() ? (long)g_plan->block_sz[g_cap_block] : -1;
            g_graph_why_got = (long)c;
            throw vh_plan_mismatch();
        }
        return g_plan->blocks[g_cap_block++];
    }
    if (g_na_base) {
        // allocation-free call: a block of the scratch array (sized by the entry point)
        const size_t need = (bytes + 255) & ~(size_t)255;
        VH_REQUIRE(g_na_off + need <= VH_SCRATCH_BYTES);
        void *p = g_na_base + g_na_off;
        g_na_off += need;
        return p;
    }
    if (void *p = vh_take_parked(g_lane, c)) { g_cache_hits++; return p; }
    for (int ln = 0; ln < VH_MAX_LANES; ln++)
        if (ln != g_lane && !BC.free_[ln].empty())
            if (void *p = vh_take_parked(ln, c)) { g_cache_hits++; return p; }
    void *p = nullptr;
    g_cache_misses++;
    if (hipMallocAsync(&p, c, S()) != hipSuccess) {
        // out of memory with blocks parked in the cache: give them back and retry once
        (void)hipGetLastError();
        g_cache_flushes++;
        sync_all_lanes();
        for (auto &lane : BC.free_)
            for (auto &kv : lane)
                for (auto &b : kv.second) { VH_CHECK(hipFree(b.p)); BC.size_of.erase(b.p); BC.events.push_back(b.ev); }
        for (auto &lane : BC.free_) lane.clear();
        BC.order.clear();
        BC.cached = 0;
        VH_CHECK(hipMallocAsync(&p, c, S()));
    }
    BC.size_of[p] = c;
    return p;
}

static void vh_dfree(void *p)
{
    if (!p || g_cap >= VH_CAP_RECORD) return;   // a plan's blocks stay with the plan
    if (g_na_base && (char *)p >= g_na_base && (char *)p < g_na_base + VH_SCRATCH_BYTES) return;
    auto it = BC.size_of.find(p);
    VH_REQUIRE(it != BC.size_of.end());
    const size_t c = it->second;
    while (c <= g_cache_max && BC.cached + c > g_cache_max && vh_evict_oldest()) {}
    if (BC.cached + c <= g_cache_max) {
        hipEvent_t ev;
        if (!BC.events.empty()) { ev = BC.events.back(); BC.events.pop_back(); }
        else VH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        VH_CHECK(hipEventRecord(ev, S()));
        const uint64_t seq = ++BC.seq;
        BC.free_[g_lane][c].push_back({p, ev, seq});
        BC.order[seq] = {g_lane, c};
        BC.cached += c;
    } else {
        BC.size_of.erase(it);
        VH_CHECK(hipFreeAsync(p, S()));
    }
}

// ---------------------------------------------------------------------------
// Number theory (host): the reference's number_theory.c contract
// ---------------------------------------------------------------------------
static inline uint64_t mulmod_h(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }
static inline uint64_t shoup_of(uint64_t w, uint64_t q) { return (uint64_t)(((u128)(w % q) << 64) / q); }

extern "C" uint64_t power_mod(uint64_t base, uint64_t exp, uint64_t mod)
{
    uint64_t r = 1 % mod;
    base %= mod;
    while (exp) {
        if (exp & 1) r = mulmod_h(r, base, mod);
        base = mulmod_h(base, base, mod);
        exp >>= 1;
    }
    return r;
}

extern "C" uint64_t inverse_mod_eea(uint64_t a, uint64_t p)
{
    __int128 t = 0, nt = 1, r = p, nr = a % p;
    while (nr != 0) {
        __int128 qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += p;
    return (uint64_t)t;
}

extern "C" uint64_t inverse_mod(uint64_t a, uint64_t m)
{
    a %= m;
    if (a == 0) return 0;
    return inverse_mod_eea(a, m);
}

extern "C" bool is_prime(uint64_t n)
{
    static const uint64_t B[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : B) {
        if (n == b) return true;
        if (n % b == 0) return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (uint64_t b : B) {
        uint64_t x = power_mod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s && comp; r++) {
            x = mulmod_h(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

extern "C" uint64_t generate_Nth_root_of_unity(uint64_t q, uint64_t n)
{
    for (uint64_t g = 2;; g++) {
        uint64_t root = power_mod(g, (q - 1) / n, q);
        if (power_mod(root, n / 2, q) == q - 1) return root;
    }
}

extern "C" uint64_t next_special_prime(uint64_t x, uint64_t n, bool primitive)
{
    uint64_t ord = 2 * n, a = x / ord;
    if (a * ord + 1 <= x) a++;
    if (primitive && a % 2 == 0) a++;
    for (;;) {
        uint64_t c = a * ord + 1;
        if (is_prime(c) && (!primitive || c % (4 * n) != 1)) return c;
        a += primitive ? 2 : 1;
    }
}

static uint64_t brv(uint64_t x, int bits)
{
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

extern "C" void bit_rev(uint64_t *out, uint64_t *in, uint64_t n, uint64_t log_n)
{
    for (uint64_t i = 0; i < n; i++) out[i] = in[brv(i, (int)log_n)];
}

extern "C" uint64_t add_modq(uint64_t a, uint64_t b, uint64_t q) { uint64_t s = a + b; return s >= q ? s - q : s; }
extern "C" uint64_t sub_modq(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
extern "C" uint64_t negate_modq(uint64_t a, uint64_t q) { return a == 0 ? 0 : q - a; }
extern "C" uint64_t mul_modq(uint64_t a, uint64_t b, NTT_proc proc) { return mulmod_h(a, b, proc->q); }

static int ilog2(uint64_t n)
{
    int l = 0;
    while ((1ULL << l) < n) l++;
    return l;
}

// ---------------------------------------------------------------------------
// NTT_proc: roots (ntt.c:265-313) and the engine's device tables
// ---------------------------------------------------------------------------
static double centred(uint64_t w, uint64_t q) { return w > q / 2 ? -(double)(q - w) : (double)w; }

static void fill_limb_consts(vh_limb &L, uint64_t n, uint64_t q)
{
    memset(&L, 0, sizeof(L));
    L.q = q;
    L.mu64 = UINT64_MAX / q;
    uint32_t s = 0;
    while (s < 64 && (q >> s)) s++;
    L.bar_s = s;
    L.bar_mu = (uint64_t)(((u128)1 << (2 * s)) / q);
    L.fp64 = q < (1ULL << 50);
    L.qd = (double)q;
    L.qinv = 1.0 / (double)q;
    L.ninv = inverse_mod(n, q);
    L.ninv_sh = shoup_of(L.ninv, q);
    L.ninv_d = centred(L.ninv, q);
    L.ninv_q = L.ninv_d / (double)q;
}

extern "C" NTT_proc ntt_new_proc(uint64_t n, uint64_t q)
{
    VH_ENTER;
    VH_REQUIRE(n >= 2 && (n & (n - 1)) == 0 && n <= (1ULL << 17));
    VH_REQUIRE(q > 2 && q <= UINT64_MAX / 3 && (q - 1) % (2 * n) == 0);  // 3q < 2^64 (Barrett bound)
    NTT_proc p = new _NTT_proc();
    p->n = n;
    p->q = q;
    // ntt.c:271-281: first g = 2, 3, ... whose g^((q-1)/2n) has order exactly 2n
    uint64_t root = 0;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t c = power_mod(g, (q - 1) / (2 * n), q);
        if (power_mod(c, n, q) == q - 1) { root = c; break; }
    }
    VH_REQUIRE(root != 0);
    p->root_of_unity = root;
    p->inv_root_of_unity = inverse_mod(root, q);
    fill_limb_consts(p->h, n, q);

    // Tables in the reference's ws order: tab[brv(i)] = psi^(+-i) (ntt_portable.c:25-33)
    const int logn = ilog2(n);
    std::vector<uint64_t> fw(n), iw(n);
    uint64_t a = 1, b = 1;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = brv(i, logn);
        fw[r] = a;
        iw[r] = b;
        a = mulmod_h(a, root, q);
        b = mulmod_h(b, p->inv_root_of_unity, q);
    }
    // Row factors of the pipelined row pass (FP64 path, n = 2^13 .. 2^16).  With
    // tw[i] = psi^brv(i), brv(2^s + k) = 2^(logn-1-s) + 2^(logn-s) brv_s(k) for k < 2^s, so
    // for k = (sub << s') + local (s = s0 + s'):
    //   tw[2^s + k] / tw[2^s' + local] = psi^(2^(logn-1-s) - 2^(logn-1-s') + 2^(logn-s) brv_s0(sub)).
    const bool rowf = p->h.fp64 && logn >= 13 && logn <= 16;
    const uint64_t s0 = rowf ? (uint64_t)logn - 8 : 0, nrow = rowf ? ((uint64_t)8 << s0) : 0;
    const size_t bytes = 2 * n * 16 + 2 * nrow * 16;
    VH_CHECK(hipMalloc(&p->d_tab, bytes));
    if (p->h.fp64) {
        std::vector<double2> t(2 * n + 2 * nrow);
        for (uint64_t i = 0; i < n; i++) {
            double wf = centred(fw[i], q), wi = centred(iw[i], q);
            t[i] = make_double2(wf, wf / (double)q);
            t[n + i] = make_double2(wi, wi / (double)q);
        }
        for (uint64_t sub = 0; sub < ((uint64_t)1 << s0) && rowf; sub++) {
            for (uint64_t sp = 0; sp < 8; sp++) {
                const uint64_t s = s0 + sp, two_n = 2 * n;
                const uint64_t e = ((n >> (s + 1)) + two_n - (n >> (sp + 1)) + (brv(sub, (int)s0) << (logn - s))) % two_n;
                const uint64_t bf = power_mod(root, e, q), bi = power_mod(p->inv_root_of_unity, e, q);
                for (uint64_t local = 0; local < ((uint64_t)1 << sp); local++) {   // the identity, checked
                    const uint64_t k = (sub << sp) + local;
                    VH_REQUIRE(fw[((uint64_t)1 << s) + k] == mulmod_h(fw[((uint64_t)1 << sp) + local], bf, q));
                    VH_REQUIRE(iw[((uint64_t)1 << s) + k] == mulmod_h(iw[((uint64_t)1 << sp) + local], bi, q));
                }
                const double cf = centred(bf, q), ci = centred(bi, q);
                t[2 * n + 8 * sub + sp] = make_double2(cf, cf / (double)q);
                t[2 * n + nrow + 8 * sub + sp] = make_double2(ci, ci / (double)q);
            }
        }
        VH_CHECK(hipMemcpy(p->d_tab, t.data(), bytes, hipMemcpyHostToDevice));
        p->h.tw_fwd = reinterpret_cast<const double2 *>(p->d_tab);
        p->h.tw_inv = p->h.tw_fwd + n;
        p->h.rowf_fwd = rowf ? p->h.tw_fwd + 2 * n : nullptr;
        p->h.rowf_inv = rowf ? p->h.tw_fwd + 2 * n + nrow : nullptr;
    } else {
        std::vector<ulonglong2> t(2 * n);
        for (uint64_t i = 0; i < n; i++) {
            t[i] = make_ulonglong2(fw[i], shoup_of(fw[i], q));
            t[n + i] = make_ulonglong2(iw[i], shoup_of(iw[i], q));
        }
        VH_CHECK(hipMemcpy(p->d_tab, t.data(), bytes, hipMemcpyHostToDevice));
        p->h.twi_fwd = reinterpret_cast<const ulonglong2 *>(p->d_tab);
        p->h.twi_inv = p->h.twi_fwd + n;
    }
    VH_CHECK(hipMalloc((void **)&p->d_self, sizeof(vh_limb)));
    VH_CHECK(hipMemcpy(p->d_self, &p->h, sizeof(vh_limb), hipMemcpyHostToDevice));
    return p;
}

extern "C" void ntt_free_proc(NTT_proc proc)
{
    if (!proc) return;
    VH_ENTER;
    VH_CHECK(hipStreamSynchronize(S()));
    VH_CHECK(hipFree(proc->d_tab));
    VH_CHECK(hipFree(proc->d_self));
    delete proc;
}

extern "C" NTT_proc *new_ntt_list(uint64_t *primes, uint64_t N, uint64_t l)
{
    NTT_proc *v = (NTT_proc *)malloc(sizeof(NTT_proc) * l);
    if (!v) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < l; i++) v[i] = ntt_new_proc(N, primes[i]);
    return v;
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
static const unsigned VH_MAX_GRID_Y = 65535;

// Checks the device error word at a synchronisation point (the stream is idle).
static void vh_check_device_error()
{
    unsigned e = 0;
    VH_CHECK(hipMemcpy(&e, C.d_err, sizeof(e), hipMemcpyDeviceToHost));
    if (e & 1) vh_die("a persistent NTT wait exceeded its spin limit (results are invalid)", __FILE__, __LINE__);
    if (e) vh_die("a persistent NTT launch left transforms incomplete (results are invalid)", __FILE__, __LINE__);
}

// Column passes (N = 2^13..2^16) use 512-thread workgroups over twice the columns per tile
// (ntt.col_wide = 1, 2^(logn-13) tiles per row item) when that still gives every CU two
// workgroups; a smaller launch keeps the 256-thread tiles, twice as many workgroups.
// ntt.col_wide = 2 takes the wide tiles at every launch size (tests use it to reach the
// 512-thread instances with small batches).  Every wide launch is counted
// (vfhe_hip_get_option("stat.col_wide_launches")).
static long g_col_wide_launches = 0;
static long g_mdr_fp64_launches = 0;   // launches of the FP64 ModDown + rescale kernel
static long g_ntt_small_launches = 0;  // launches of the register NTT for N = 2^9 .. 2^12
static long g_br_fused_launches = 0;   // launches of the one-launch blind rotation
static long g_ks_fold_jobs = 0;        // multiply jobs whose pass-through went into the key MAC
static int col_wide(unsigned items, int logn)
{
    const uint64_t tiles = (uint64_t)items << (logn - 13);
    const int w = OPT.ntt_col_wide >= 2 || (OPT.ntt_col_wide == 1 && tiles >= 2 * (uint64_t)C.cus) ? 1 : 0;
    g_col_wide_launches += w;
    return w;
}

// Forward (dir 0) or inverse (dir 1) transforms of n-point rows.
// Launches the transforms of `cnt` staged work items (device list d_items, all of one
// arithmetic path: fp = FP64 limbs) of n-point rows, direction dir.
static void ntt_dispatch(const vh_item *d_items, unsigned cnt, const vh_limb *d_limbs, int fp, uint64_t n, int dir)
{
    const int logn = ilog2(n);
    if (fp && logn >= 13 && logn <= 16 && OPT.ntt_mode == 1 && C.xcd_ok) {
        // Persistent kernel: both passes of every transform on one XCD, the
        // intermediate exchanged through that XCD's L2.
        unsigned *ctl = (unsigned *)vh_dmalloc((16 + 2 * cnt) * sizeof(unsigned));
        VH_LAUNCH(vh_launch_zero_u32(S(), ctl, 16 + 2 * cnt));
        const int occ = OPT.xcd_occ >= 4 ? 4 : 2;
        {
            vh_prof_scope ps_(VH_PROF_NTT_FP64);
            VH_LAUNCH(vh_launch_ntt_xcd(S(), d_items, cnt, d_limbs, dir, logn, ctl, C.d_err,
                                       (unsigned)OPT.xcd_lag, (unsigned)OPT.spin_limit,
                                       (unsigned)(C.cus * occ), occ, OPT.xcd_l2 ? 1 : 0));
        }
        VH_LAUNCH(vh_launch_ntt_xcd_verify(S(), ctl, cnt, logn, C.d_err));
        vh_dfree(ctl);
    } else if (fp && logn >= 13 && logn <= 16) {
        // Two register radix-16 passes (column pass of R1 = logn - 8 stages, row pass of 8),
        // optionally over sub-batches of ntt_chunk_mb MiB of rows.
        const int R1 = logn - 8;
        const uint64_t chunk = OPT.ntt_chunk_mb > 0 ? (uint64_t)OPT.ntt_chunk_mb << 20 : UINT64_MAX;
        const unsigned per = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(cnt, chunk / (n * 8)));
        const bool pipe = OPT.ntt_mode == 2 && C.pipe_blocks > 0;
        for (unsigned c0 = 0; c0 < cnt; c0 += per) {
            const unsigned cc = std::min(per, cnt - c0);
            const vh_item *ci = d_items + c0;
            if (pipe) {
                auto nb = [&](int bit) -> unsigned { return (OPT.ntt_tile_grid >> bit) & 1 ? UINT_MAX : C.pipe_blocks; };
                if (dir == 0) {
                    { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16p(S(), ci, cc, d_limbs, 0, logn, 0, R1, 1, 1, 0, nb(0), col_wide(cc, logn))); }
                    { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16p(S(), ci, cc, d_limbs, 0, logn, R1, 8, 0, 0, 1, nb(1))); }
                } else {
                    { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16p(S(), ci, cc, d_limbs, 1, logn, R1, 8, 0, 1, 0, nb(2))); }
                    { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16p(S(), ci, cc, d_limbs, 1, logn, 0, R1, 1, 0, 1, nb(3), col_wide(cc, logn))); }
                }
            } else if (dir == 0) {
                { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16(S(), ci, cc, d_limbs, 0, logn, 0, R1, 1, 1, 0)); }
                { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16(S(), ci, cc, d_limbs, 0, logn, R1, 8, 0, 0, 1)); }
            } else {
                { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16(S(), ci, cc, d_limbs, 1, logn, R1, 8, 0, 1, 0)); }
                { vh_prof_scope ps_(VH_PROF_NTT_FP64); VH_LAUNCH(vh_launch_ntt16(S(), ci, cc, d_limbs, 1, logn, 0, R1, 1, 0, 1)); }
            }
        }
    } else if (fp && logn >= 9 && logn <= 12 && OPT.ntt_small) {
        // whole transforms in registers, 4096/N of them per workgroup
        vh_prof_scope ps_(VH_PROF_NTT_FP64);
        VH_LAUNCH(vh_launch_ntt_small(S(), d_items, cnt, d_limbs, dir, logn));
        g_ntt_small_launches++;
    } else if (logn <= 12) {
        { vh_prof_scope ps_(fp ? VH_PROF_NTT_FP64 : VH_PROF_NTT_INT); VH_LAUNCH(vh_launch_ntt_pass(S(), d_items, cnt, d_limbs, fp, dir, logn, 0, logn, 0, 1, 1, 1)); }
    } else {
        const int R1 = logn - 8, R2 = 8;
        const int Gc = std::min(1 << (logn - R1), 4096 >> R1);
        const int Gr = std::min(1 << R1, 16);
        if (dir == 0) {
            { vh_prof_scope ps_(fp ? VH_PROF_NTT_FP64 : VH_PROF_NTT_INT); VH_LAUNCH(vh_launch_ntt_pass(S(), d_items, cnt, d_limbs, fp, 0, logn, 0, R1, 1, Gc, 1, 0)); }
            { vh_prof_scope ps_(fp ? VH_PROF_NTT_FP64 : VH_PROF_NTT_INT); VH_LAUNCH(vh_launch_ntt_pass(S(), d_items, cnt, d_limbs, fp, 0, logn, R1, R2, 0, Gr, 0, 1)); }
        } else {
            { vh_prof_scope ps_(fp ? VH_PROF_NTT_FP64 : VH_PROF_NTT_INT); VH_LAUNCH(vh_launch_ntt_pass(S(), d_items, cnt, d_limbs, fp, 1, logn, R1, R2, 0, Gr, 1, 0)); }
            { vh_prof_scope ps_(fp ? VH_PROF_NTT_FP64 : VH_PROF_NTT_INT); VH_LAUNCH(vh_launch_ntt_pass(S(), d_items, cnt, d_limbs, fp, 1, logn, 0, R1, 1, Gc, 0, 1)); }
        }
    }
}

// Forward (dir 0) or inverse (dir 1) transforms of n-point rows.
static void run_ntt(std::vector<vh_item> &items, const vh_limb *d_limbs, const vh_limb *h_limbs, uint64_t n, int dir)
{
    if (items.empty()) return;
    std::vector<vh_item> grp[2];
    for (auto &it : items) grp[h_limbs[it.limb].fp64 ? 1 : 0].push_back(it);
    // Limb-major order: workgroups running together share one limb's twiddle table in L2.
    for (auto &g : grp)
        std::stable_sort(g.begin(), g.end(), [](const vh_item &a, const vh_item &b) { return a.limb < b.limb; });
    for (int fp = 0; fp < 2; fp++) {
        auto &v = grp[fp];
        for (size_t off = 0; off < v.size(); off += VH_MAX_GRID_Y) {
            const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, v.size() - off);
            std::vector<vh_item> chunk(v.begin() + off, v.begin() + off + cnt);
            ntt_dispatch(vh_stage_vec(chunk), cnt, d_limbs, fp, n, dir);
        }
    }
}

static void run_eltwise(std::vector<vh_item> &items, const vh_limb *d_limbs, int op, uint64_t n)
{
    for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
        std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
        {
            vh_prof_scope ps_(VH_PROF_ELTWISE);
            VH_LAUNCH(vh_launch_eltwise(S(), vh_stage_vec(chunk), cnt, d_limbs, op, n));
        }
    }
}

static vh_item mk_item(const uint64_t *src, uint64_t *dst, uint32_t limb, const uint64_t *src2 = nullptr,
                       uint64_t scalar = 0, uint64_t scalar_sh = 0, uint32_t aux = 0)
{
    vh_item it;
    it.src = src;
    it.dst = dst;
    it.limb = limb;
    it.aux = aux;
    it.src2 = src2;
    it.scalar = scalar;
    it.scalar_sh = scalar_sh;
    return it;
}

// ---------------------------------------------------------------------------
// Standalone single-limb kernels on host or device pointers
// ---------------------------------------------------------------------------
static bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice;
}

// Runs fn(device pointers...) on arrays of n uint64 that may live on the host.
struct vh_args {
    uint64_t *dev[3] = {nullptr, nullptr, nullptr};
    uint64_t *host[3] = {nullptr, nullptr, nullptr};
    bool staged = false;
    uint64_t n = 0;
};

static vh_args stage_args(uint64_t n, uint64_t *out, const uint64_t *in1, const uint64_t *in2, bool out_is_input)
{
    vh_args a;
    a.n = n;
    const uint64_t *ptrs[3] = {out, in1, in2};
    bool host = false;
    for (int i = 0; i < 3; i++) if (ptrs[i] && !is_device_ptr(ptrs[i])) host = true;
    if (!host) {
        a.dev[0] = out; a.dev[1] = const_cast<uint64_t *>(in1); a.dev[2] = const_cast<uint64_t *>(in2);
        return a;
    }
    a.staged = true;
    for (int i = 0; i < 3; i++) {
        if (!ptrs[i]) continue;
        a.host[i] = const_cast<uint64_t *>(ptrs[i]);
        a.dev[i] = (uint64_t *)vh_dmalloc(n * 8);
        if (i > 0 || out_is_input)
            VH_CHECK(hipMemcpyAsync(a.dev[i], ptrs[i], n * 8, hipMemcpyHostToDevice, S()));
    }
    return a;
}

static void finish_args(vh_args &a)
{
    if (!a.staged) return;
    VH_CHECK(hipMemcpyAsync(a.host[0], a.dev[0], a.n * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    for (int i = 0; i < 3; i++) vh_dfree(a.dev[i]);
}

static void single_ntt(uint64_t *out, uint64_t *in, NTT_proc proc, int dir)
{
    VH_ENTER;
    vh_args a = stage_args(proc->n, out, in, nullptr, false);
    std::vector<vh_item> items{mk_item(a.dev[1], a.dev[0], 0)};
    run_ntt(items, proc->d_self, &proc->h, proc->n, dir);
    finish_args(a);
}

extern "C" void ntt_forward(uint64_t *out, uint64_t *in, NTT_proc proc) { single_ntt(out, in, proc, 0); }
extern "C" void ntt_reverse(uint64_t *out, uint64_t *in, NTT_proc proc) { single_ntt(out, in, proc, 1); }

static void single_eltwise(int op, uint64_t *out, const uint64_t *in1, const uint64_t *in2, uint64_t scalar,
                           uint64_t n, NTT_proc proc, bool out_is_input)
{
    VH_ENTER;
    if (n == 0) return;
    const uint64_t q = proc->q;
    vh_args a = stage_args(n, out, in1, in2, out_is_input);
    uint64_t s = scalar, ssh = 0;
    if (op == VH_OP_SCALE || op == VH_OP_FMA) { s = scalar % q; ssh = shoup_of(s, q); }
    if (op == VH_OP_ADD_SCALAR || op == VH_OP_SUB_SCALAR) s = scalar % q;
    if (op == VH_OP_REDUCE_MP) { s = (uint64_t)(((u128)1 << 64) % q); ssh = shoup_of(s, q); }
    std::vector<vh_item> items{mk_item(a.dev[1], a.dev[0], 0, a.dev[2], s, ssh)};
    run_eltwise(items, proc->d_self, op, n);
    finish_args(a);
}

extern "C" void mod_eltwise_mul(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_MUL, out, in1, in2, 0, n, proc, false); }
extern "C" void mod_eltwise_mul_addto(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_MUL_ADDTO, out, in1, in2, 0, n, proc, true); }
extern "C" void mod_eltwise_mul_subto(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_MUL_SUBTO, out, in1, in2, 0, n, proc, true); }
extern "C" void mod_eltwise_scale(uint64_t *out, uint64_t *in, uint64_t scale, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_SCALE, out, in, nullptr, scale, n, proc, false); }
extern "C" void mod_eltwise_fma(uint64_t *out, uint64_t *in, uint64_t scale, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_FMA, out, in, nullptr, scale, n, proc, true); }
extern "C" void mod_eltwise_add_scalar(uint64_t *out, uint64_t *in, uint64_t scalar, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_ADD_SCALAR, out, in, nullptr, scalar, n, proc, false); }
extern "C" void mod_eltwise_sub_scalar(uint64_t *out, uint64_t *in, uint64_t scalar, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_SUB_SCALAR, out, in, nullptr, scalar, n, proc, false); }
extern "C" void mod_eltwise_negate(uint64_t *out, uint64_t *in, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_NEGATE, out, in, nullptr, 0, n, proc, false); }
extern "C" void mod_eltwise_add(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_ADD, out, in1, in2, 0, n, proc, false); }
extern "C" void mod_eltwise_sub(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_SUB, out, in1, in2, 0, n, proc, false); }
extern "C" void mod_eltwise_reduce(uint64_t *out, uint64_t *in, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_REDUCE, out, in, nullptr, 0, n, proc, false); }
extern "C" void mod_eltwise_reduce_signed(uint64_t *out, int64_t *in, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_REDUCE_SIGNED, out, (const uint64_t *)in, nullptr, 0, n, proc, false); }
extern "C" void mod_reduce_array_mp(uint64_t *out, uint64_t *in_high, uint64_t *in_low, uint64_t n, NTT_proc proc)
{ single_eltwise(VH_OP_REDUCE_MP, out, in_low, in_high, 0, n, proc, false); }

// ---------------------------------------------------------------------------
// incNTT pool (rns_polynomial.c:7-61)
// ---------------------------------------------------------------------------
static void pool_upload(vh_pool *P)
{
    vh_limb *old = P->d_limbs;
    P->d_limbs = (vh_limb *)vh_dmalloc(P->h_limbs.size() * sizeof(vh_limb));
    VH_CHECK(hipMemcpyAsync(P->d_limbs, P->h_limbs.data(), P->h_limbs.size() * sizeof(vh_limb),
                            hipMemcpyHostToDevice, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(old);
}

static void pool_add_primes(vh_pool *P, const uint64_t *primes, uint64_t count)
{
    incNTT t = &P->pub;
    const uint64_t n = P->n, old_l = t->l, new_l = old_l + count;
    if (!P->div_consts.empty()) {
        // cached division constants are laid out for the old limb count
        sync_all_lanes();
        for (auto &kv : P->div_consts) VH_CHECK(hipFree(kv.second));
        P->div_consts.clear();
    }
    t->ntt = (NTT_proc *)realloc(t->ntt, sizeof(NTT_proc) * new_l);
    t->w = (uint64_t **)realloc(t->w, sizeof(uint64_t *) * new_l);
    if (!t->ntt || !t->w) exit(EXIT_FAILURE);
    std::vector<uint64_t> wp(2 * n);
    for (uint64_t i = old_l; i < new_l; i++) {
        const uint64_t q = primes[i - old_l];
        t->ntt[i] = ntt_new_proc(n, q);
        // rns_polynomial.c:21-27: w[k] = psi^(2 brv(k) + 1)
        const uint64_t psi = t->ntt[i]->root_of_unity;
        wp[0] = psi;
        for (uint64_t j = 1; j < 2 * n; j++) wp[j] = mulmod_h(wp[j - 1], psi, q);
        uint64_t *w = (uint64_t *)aligned_alloc(64, ((n * 8 + 63) / 64) * 64);
        if (!w) exit(EXIT_FAILURE);
        for (uint64_t k = 0; k < n; k++) w[k] = wp[2 * brv(k, (int)P->logn)];
        t->w[i] = w;
        P->h_limbs.push_back(t->ntt[i]->h);
        uint64_t *dw = nullptr;
        if (t->split_degree > 1) {
            VH_CHECK(hipMalloc((void **)&dw, n * 8));
            VH_CHECK(hipMemcpy(dw, w, n * 8, hipMemcpyHostToDevice));
        }
        P->d_w.push_back(dw);
    }
    t->l = new_l;
    pool_upload(P);
}

extern "C" incNTT new_incomplete_ntt_list(uint64_t *primes, uint64_t split_degree, uint64_t N, uint64_t l)
{
    VH_ENTER;
    VH_REQUIRE(split_degree >= 1 && (split_degree & (split_degree - 1)) == 0 && N % split_degree == 0);
    VH_REQUIRE(l <= VH_MAX_LIMBS);
    vh_pool *P = new vh_pool();
    P->pub.ntt = nullptr;
    P->pub.w = nullptr;
    P->pub.split_degree = split_degree;
    P->pub.N = N;
    P->pub.l = 0;
    P->n = N / split_degree;
    P->logn = ilog2(P->n);
    P->d_limbs = nullptr;
    pool_add_primes(P, primes, l);
    return &P->pub;
}

extern "C" void incNTT_extend_with_primes(incNTT ntt, uint64_t *new_primes, uint64_t count)
{
    if (count == 0) return;
    VH_ENTER;
    VH_REQUIRE(ntt->l + count <= VH_MAX_LIMBS);
    pool_add_primes(VPOOL(ntt), new_primes, count);
}

extern "C" uint64_t **incNTT_get_rou_matrix(incNTT ntt) { return ntt->w; }

// ---------------------------------------------------------------------------
// Polynomials (rns_polynomial.c:65-181)
// ---------------------------------------------------------------------------
// Allocates the host mirror of a polynomial created without one.
static void ensure_mirror(vh_poly *v)
{
    if (v->hblock) return;
    VH_REQUIRE(!v->borrowed);
    const uint64_t l = v->pub.allocated_l, N = v->pub.ntt->N;
    const size_t hb = ((l * N * 8 + 63) / 64) * 64;
    v->hblock = (uint64_t *)aligned_alloc(64, hb ? hb : 64);
    if (!v->hblock) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < l; i++) v->pub.coeffs[i] = v->hblock + i * N;
}

static void dev_ready(RNS_Polynomial p)
{
    vh_poly *v = VP(p);
    if (!v->host_dirty) return;
    if (g_cap >= VH_CAP_RECORD) { g_graph_why = 3000000L; throw vh_plan_mismatch(); }   // uploads happen before capture
    VH_CHECK(hipMemcpyAsync(v->d, v->hblock, p->allocated_l * p->ntt->N * 8, hipMemcpyHostToDevice, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    v->host_dirty = 0;
}

static inline uint64_t *drow(RNS_Polynomial p, uint64_t limb) { return VP(p)->d + limb * p->ntt->N; }
static inline const vh_limb *DL(RNS_Polynomial p) { return VPOOL(p->ntt)->d_limbs; }

// Every active limb must have a row (the reference indexes coeffs[i] unchecked).
static void check_rows(RNS_Polynomial p, uint64_t mask)
{
    VH_REQUIRE(p->allocated_l >= 64 || (mask >> p->allocated_l) == 0);
}

// zero = false leaves the device rows uninitialised, as the reference's allocation
// does (rns_polynomial.c:65-78); the public entry point clears them.  mirror =
// false defers the host mirror (and the coeffs[] row pointers) to the first
// polynomial_sync_to_host: engine temporaries never need it, and allocating a
// host block per temporary cost a third of a serving worker's host time.
static RNS_Polynomial new_poly(uint64_t N, uint64_t rns_mask, incNTT ntt, bool zero, bool mirror)
{
    VH_REQUIRE(N == ntt->N);
    vh_poly *v = new vh_poly();
    const uint64_t l = ntt->l;
    v->pub.ntt = ntt;
    v->pub.rns_mask = rns_mask;
    v->pub.allocated_l = l;
    v->pub.coeffs = (uint64_t **)malloc(sizeof(uint64_t *) * (l ? l : 1));
    if (!v->pub.coeffs) exit(EXIT_FAILURE);
    v->hblock = nullptr;
    for (uint64_t i = 0; i < l; i++) v->pub.coeffs[i] = nullptr;
    if (mirror) ensure_mirror(v);
    v->d = (uint64_t *)vh_dmalloc(l * N * 8);
    if (zero && g_cap >= VH_CAP_RECORD) VH_LAUNCH(vh_launch_zero_u32(S(), (unsigned *)v->d, (unsigned)(l * N * 2)));
    else if (zero) VH_CHECK(hipMemsetAsync(v->d, 0, l * N * 8, S()));
    v->host_dirty = 0;
    return &v->pub;
}

extern "C" RNS_Polynomial polynomial_new_RNS_polynomial(uint64_t N, uint64_t rns_mask, incNTT ntt)
{
    VH_ENTER;
    return new_poly(N, rns_mask, ntt, true, true);
}

extern "C" RNS_Polynomial *polynomial_new_RNS_polynomial_array(uint64_t size, uint64_t N, uint64_t rns_mask, incNTT ntt)
{
    RNS_Polynomial *r = (RNS_Polynomial *)malloc(sizeof(RNS_Polynomial) * (size ? size : 1));
    if (!r) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < size; i++) r[i] = polynomial_new_RNS_polynomial(N, rns_mask, ntt);
    return r;
}

extern "C" RNS_Polynomial *polynomial_new_array_of_RNS_polynomials(uint64_t N, uint64_t rns_mask, uint64_t size, incNTT ntt)
{
    return polynomial_new_RNS_polynomial_array(size, N, rns_mask, ntt);
}

extern "C" void free_RNS_polynomial(void *p)
{
    if (!p) return;
    VH_ENTER;
    vh_poly *v = reinterpret_cast<vh_poly *>(p);
    if (!v->borrowed) {
        vh_dfree(v->d);
        free(v->hblock);
    }
    free(v->pub.coeffs);
    delete v;
}

extern "C" void free_RNS_polynomial_array(uint64_t size, RNS_Polynomial *p)
{
    for (uint64_t i = 0; i < size; i++) free_RNS_polynomial(p[i]);
    free(p);
}

extern "C" void polynomial_sync_to_host(RNS_Polynomial p)
{
    VH_ENTER;
    vh_poly *v = VP(p);
    if (v->host_dirty) return;
    ensure_mirror(v);
    sync_all_lanes();   // the rows may have been written on any lane
    VH_CHECK(hipMemcpyAsync(v->hblock, v->d, p->allocated_l * p->ntt->N * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_check_device_error();
}

extern "C" void polynomial_mark_host_dirty(RNS_Polynomial p)
{
    VH_ENTER;
    VH_REQUIRE(VP(p)->hblock);   // rows written through coeffs[] exist only once the mirror does
    VP(p)->host_dirty = 1;
}

extern "C" uint64_t *polynomial_device_row(RNS_Polynomial p, uint64_t limb)
{
    VH_ENTER;
    VH_REQUIRE(limb < p->allocated_l);
    dev_ready(p);
    return drow(p, limb);
}

extern "C" void vfhe_hip_synchronize(void)
{
    VH_ENTER;
    sync_all_lanes();
    vh_check_device_error();
}

extern "C" int vfhe_hip_set_lane(int lane)
{
    VH_ENTER;
    if (lane < 0 || lane >= VH_MAX_LANES) return -1;
    if (!C.lanes[lane].stream) lane_init(C.lanes[lane], (size_t)16 << 20);
    const int prev = g_lane;
    g_lane = lane;
    return prev;
}

extern "C" int vfhe_hip_set_option(const char *key, long value)
{
    VH_ENTER;
    if (!key) return -1;
    if (!strcmp(key, "ntt.mode")) OPT.ntt_mode = value;
    else if (!strcmp(key, "ntt.chunk_mb")) OPT.ntt_chunk_mb = value;
    else if (!strcmp(key, "ntt.tile_grid")) OPT.ntt_tile_grid = value;
    else if (!strcmp(key, "ntt.col_wide")) OPT.ntt_col_wide = value;
    else if (!strcmp(key, "ntt.xcd_lag")) OPT.xcd_lag = value < 0 ? 0 : value;
    else if (!strcmp(key, "ntt.xcd_occupancy")) OPT.xcd_occ = value;
    else if (!strcmp(key, "ntt.xcd_l2")) OPT.xcd_l2 = value;
    else if (!strcmp(key, "ntt.spin_limit")) OPT.spin_limit = value < 1 ? 1 : value;
    else if (!strcmp(key, "ks.fused")) OPT.ks_fused = value;
    else if (!strcmp(key, "ks.mdr_fp64")) OPT.mdr_fp64 = value;
    else if (!strcmp(key, "ks.fold")) OPT.ks_fold = value;
    else if (!strcmp(key, "ntt.small")) OPT.ntt_small = value;
    else if (!strcmp(key, "br.fused")) OPT.br_fused = value;
    else if (!strcmp(key, "cache.max_mb")) {
        if (value < 0) return -1;
        g_cache_max = (size_t)value << 20;
        while (BC.cached > g_cache_max && vh_evict_oldest()) {}
    }
    else if (!strcmp(key, "fft.fma")) OPT.fft_fma = value;
    else return -1;
    return 0;
}

extern "C" uint64_t vfhe_hip_release_cached(void)
{
    VH_ENTER;
    const uint64_t bytes = BC.cached;
    while (vh_evict_oldest()) {}
    // and the stream-ordered pool's unused memory: after large, varied earlier work its free
    // space is fragmented and every later allocation searches it (measured ~1.6 ms per miss)
    sync_all_lanes();
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, C.device) == hipSuccess) VH_CHECK(hipMemPoolTrimTo(pool, 0));
    (void)hipGetLastError();
    return bytes;
}

extern "C" long vfhe_hip_get_option(const char *key)
{
    VH_ENTER;
    if (!key) return -1;
    if (!strcmp(key, "ntt.xcd_available")) return C.xcd_ok ? 1 : 0;
    if (!strncmp(key, "probe.xcd", 9) && key[9] >= '0' && key[9] <= '7' && !key[10]) return C.xcd_hist[key[9] - '0'];
    if (!strcmp(key, "ntt.mode")) return OPT.ntt_mode;
    if (!strcmp(key, "ntt.chunk_mb")) return OPT.ntt_chunk_mb;
    if (!strcmp(key, "ntt.tile_grid")) return OPT.ntt_tile_grid;
    if (!strcmp(key, "ntt.col_wide")) return OPT.ntt_col_wide;
    if (!strcmp(key, "ntt.xcd_lag")) return OPT.xcd_lag;
    if (!strcmp(key, "ntt.xcd_occupancy")) return OPT.xcd_occ;
    if (!strcmp(key, "ntt.xcd_l2")) return OPT.xcd_l2;
    if (!strcmp(key, "ntt.spin_limit")) return OPT.spin_limit;
    if (!strcmp(key, "ks.fused")) return OPT.ks_fused;
    if (!strcmp(key, "ks.mdr_fp64")) return OPT.mdr_fp64;
    if (!strcmp(key, "ks.fold")) return OPT.ks_fold;
    if (!strcmp(key, "ntt.small")) return OPT.ntt_small;
    if (!strcmp(key, "stat.ntt_small_launches")) return g_ntt_small_launches;
    if (!strcmp(key, "br.fused")) return OPT.br_fused;
    if (!strcmp(key, "stat.br_fused_launches")) return g_br_fused_launches;
    if (!strcmp(key, "cache.max_mb")) return (long)(g_cache_max >> 20);
    if (!strcmp(key, "stat.ks_fold_jobs")) return g_ks_fold_jobs;
    if (!strcmp(key, "fft.fma")) return OPT.fft_fma;
    if (!strcmp(key, "stat.col_wide_launches")) return g_col_wide_launches;
    if (!strcmp(key, "stat.mdr_fp64_launches")) return g_mdr_fp64_launches;
    if (!strcmp(key, "stat.cache_hits")) return g_cache_hits;
    if (!strcmp(key, "stat.cache_misses")) return g_cache_misses;
    if (!strcmp(key, "stat.cache_flushes")) return g_cache_flushes;
    if (!strcmp(key, "stat.cache_parked_bytes")) return (long)BC.cached;
    if (!strcmp(key, "stat.graph_captures")) return g_graph_captures;
    if (!strcmp(key, "stat.graph_replays")) return g_graph_replays;
    if (!strcmp(key, "stat.graph_fallbacks")) return g_graph_fallbacks;
    if (!strcmp(key, "stat.graph_why")) return g_graph_why;
    if (!strcmp(key, "stat.graph_why_want")) return g_graph_why_want;
    if (!strcmp(key, "stat.graph_why_got")) return g_graph_why_got;
    if (!strcmp(key, "stat.graph_plans")) return (long)g_plans.size();
    return -1;
}

extern "C" int vfhe_hip_last_error(void)
{
    VH_ENTER;
    return (int)hipGetLastError();
}

extern "C" const char *vfhe_hip_device_name(void)
{
    VH_ENTER;
    return C.name;
}

extern "C" const char *vfhe_engine_active(void) { return "hip"; }

// cpu_probe.c:14-33 semantics (empty = no requirement, unknown = no), plus "hip".
extern "C" int vfhe_cpu_supports(const char *capability)
{
    if (!capability || !capability[0]) return 1;
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
    __builtin_cpu_init();
    if (!strcmp(capability, "avx512ifma")) return __builtin_cpu_supports("avx512ifma") ? 1 : 0;
    if (!strcmp(capability, "avx512f")) return __builtin_cpu_supports("avx512f") ? 1 : 0;
    if (!strcmp(capability, "avx2")) return __builtin_cpu_supports("avx2") ? 1 : 0;
#endif
    if (strcmp(capability, "hip") != 0) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { (void)hipGetLastError(); return 0; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Row-wise helpers over active limbs ------------------------------------------------
template <typename F> static void for_active(uint64_t mask, uint64_t l, F f)
{
    for (uint64_t i = 0; i < l; i++) if (ACTIVE(mask, i)) f(i);
}

static void copy_rows(RNS_Polynomial out, RNS_Polynomial in, uint64_t mask)
{
    if (out == in) return;
    const uint64_t N = out->ntt->N;
    std::vector<vh_item> items;
    for_active(mask, out->ntt->l, [&](uint64_t i) { items.push_back(mk_item(drow(in, i), drow(out, i), i)); });
    run_eltwise(items, DL(out), VH_OP_COPY, N);
}

extern "C" void polynomial_copy_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in)
{
    VH_ENTER;
    dev_ready(in);
    dev_ready(out);
    out->rns_mask = in->rns_mask;
    check_rows(out, out->rns_mask);
    copy_rows(out, in, out->rns_mask);
}

extern "C" void polynomial_copy_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in)
{
    polynomial_copy_RNS_polynomial(out, in);
}

static void zero_rows(RNS_Polynomial p, uint64_t mask)
{
    std::vector<vh_item> items;
    for_active(mask, std::min(p->ntt->l, p->allocated_l), [&](uint64_t i) { items.push_back(mk_item(nullptr, drow(p, i), i)); });
    run_eltwise(items, DL(p), VH_OP_ZERO, p->ntt->N);
}

extern "C" void polynomial_RNS_zero(RNS_Polynomial p)
{
    VH_ENTER;
    dev_ready(p);
    zero_rows(p, p->rns_mask);
}

extern "C" bool polynomial_eq(RNS_Polynomial a, RNS_Polynomial b)
{
    VH_ENTER;
    polynomial_sync_to_host(a);
    polynomial_sync_to_host(b);
    const uint64_t la = a->ntt->l, lb = b->ntt->l, ml = std::max(la, lb);
    for (uint64_t i = 0; i < ml; i++) {
        bool aa = i < la && ACTIVE(a->rns_mask, i), ab = i < lb && ACTIVE(b->rns_mask, i);
        if (aa && ab) {
            if (memcmp(a->coeffs[i], b->coeffs[i], a->ntt->N * 8) != 0) return false;
        } else if (aa) {
            for (uint64_t j = 0; j < a->ntt->N; j++) if (a->coeffs[i][j]) return false;
        } else if (ab) {
            for (uint64_t j = 0; j < b->ntt->N; j++) if (b->coeffs[i][j]) return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Domain conversion (rns_polynomial.c:582-614)
// ---------------------------------------------------------------------------
static void ntt_rows(std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs, int dir)
{
    if (pairs.empty()) return;
    incNTT t = pairs[0].first->ntt;
    vh_pool *P = VPOOL(t);
    std::vector<vh_item> items;
    for (auto &pr : pairs) {
        RNS_Polynomial out = pr.first, in = pr.second;
        VH_REQUIRE(out->ntt == t && in->ntt == t);
        check_rows(out, in->rns_mask);
        for_active(in->rns_mask, t->l, [&](uint64_t i) {
            for (uint64_t k = 0; k < t->split_degree; k++)
                items.push_back(mk_item(drow(in, i) + k * P->n, drow(out, i) + k * P->n, i));
        });
    }
    run_ntt(items, P->d_limbs, P->h_limbs.data(), P->n, dir);
}

extern "C" void polynomial_RNSc_to_RNS(RNS_Polynomial out, RNSc_Polynomial in)
{
    VH_ENTER;
    dev_ready(in);
    dev_ready(out);
    out->rns_mask = in->rns_mask;
    ntt_rows({{out, in}}, 0);
}

extern "C" void polynomial_RNS_to_RNSc(RNSc_Polynomial out, RNS_Polynomial in)
{
    VH_ENTER;
    dev_ready(in);
    dev_ready(out);
    out->rns_mask = in->rns_mask;
    ntt_rows({{out, in}}, 1);
}

// Signed int64 rows already on the device -> every active limb (split placement),
// then forward NTT: the device half of int_array_to_RNS for many polynomials.
static void int_rows_to_RNS(const std::vector<std::pair<RNS_Polynomial, const uint64_t *>> &jobs)
{
    if (jobs.empty()) return;
    std::vector<vh_item> items;
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    for (auto &jb : jobs) {
        RNS_Polynomial out = jb.first;
        dev_ready(out);
        for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
            items.push_back(mk_item(jb.second, drow(out, i), i, nullptr, out->ntt->split_degree));
        });
        pairs.push_back({out, out});
    }
    RNS_Polynomial p0 = jobs[0].first;
    for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
        std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
        vh_prof_scope ps_(VH_PROF_OTHER);
        VH_LAUNCH(vh_launch_scatter_split(S(), vh_stage_vec(chunk), cnt, DL(p0), p0->ntt->N, 1));
    }
    ntt_rows(pairs, 0);
}

// int_array_to_RNS (rns_polynomial.c:204-223) and array_to_RNS (:225-240)
extern "C" void int_array_to_RNS(RNS_Polynomial out, uint64_t *in)
{
    VH_ENTER;
    const uint64_t N = out->ntt->N;
    uint64_t *tmp = (uint64_t *)vh_dmalloc(N * 8);
    VH_CHECK(hipMemcpyAsync(tmp, in, N * 8, hipMemcpyHostToDevice, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    int_rows_to_RNS({{out, tmp}});
    vh_dfree(tmp);
}

extern "C" void array_to_RNS(RNS_Polynomial out, uint64_t **in)
{
    VH_ENTER;
    dev_ready(out);
    const uint64_t N = out->ntt->N;
    std::vector<uint64_t> hostbuf;
    std::vector<vh_item> items;
    uint64_t cnt = 0;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t) { cnt++; });
    uint64_t *tmp = (uint64_t *)vh_dmalloc(cnt * N * 8);
    uint64_t k = 0;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        VH_CHECK(hipMemcpyAsync(tmp + k * N, in[i], N * 8, hipMemcpyHostToDevice, S()));
        items.push_back(mk_item(tmp + k * N, drow(out, i), i, nullptr, out->ntt->split_degree));
        k++;
    });
    VH_CHECK(hipStreamSynchronize(S()));
    if (!items.empty())
        {
            vh_prof_scope ps_(VH_PROF_OTHER);
            VH_LAUNCH(vh_launch_scatter_split(S(), vh_stage_vec(items), (unsigned)items.size(), DL(out), N, 0));
        }
    vh_dfree(tmp);
    ntt_rows({{out, out}}, 0);
}

// ---------------------------------------------------------------------------
// Pointwise products with the split-degree wrap (rns_polynomial.c:273-435)
// ---------------------------------------------------------------------------
// mode 0: out = a*b, 1: out += a*b, 2: out -= a*b, on the active limbs of mask.
struct vh_mulreq { RNS_Polynomial out, a, b; uint64_t mask; };

static void run_muls(const std::vector<vh_mulreq> &reqs, int mode)
{
    if (reqs.empty()) return;
    incNTT t = reqs[0].out->ntt;
    vh_pool *P = VPOOL(t);
    std::vector<vh_item> items;
    for (auto &r : reqs) {
        check_rows(r.out, r.mask);
        for_active(r.mask, t->l, [&](uint64_t i) {
            vh_item it = mk_item(drow(r.a, i), drow(r.out, i), i, drow(r.b, i));
            if (t->split_degree > 1) {
                it.aux = (uint32_t)t->split_degree;
                it.scalar = (uint64_t)(uintptr_t)P->d_w[i];
            }
            items.push_back(it);
        });
    }
    if (t->split_degree == 1) {
        const int op = mode == 0 ? VH_OP_MUL : (mode == 1 ? VH_OP_MUL_ADDTO : VH_OP_MUL_SUBTO);
        run_eltwise(items, P->d_limbs, op, t->N);
    } else {
        for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
            const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
            std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
            {
                vh_prof_scope ps_(VH_PROF_ELTWISE);
                VH_LAUNCH(vh_launch_split_mul(S(), vh_stage_vec(chunk), cnt, P->d_limbs, mode, P->n, t->split_degree));
            }
        }
    }
}

extern "C" void polynomial_mul_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2)
{
    VH_ENTER;
    VH_REQUIRE(out != in1 && out != in2);
    dev_ready(in1); dev_ready(in2); dev_ready(out);
    out->rns_mask = in1->rns_mask & in2->rns_mask;
    run_muls({{out, in1, in2, out->rns_mask}}, 0);
}

extern "C" void polynomial_mul_addto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2)
{
    VH_ENTER;
    VH_REQUIRE(out != in1 && out != in2);
    dev_ready(in1); dev_ready(in2); dev_ready(out);
    out->rns_mask = in1->rns_mask & in2->rns_mask;
    run_muls({{out, in1, in2, out->rns_mask}}, 1);
}

extern "C" void polynomial_mul_subto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2)
{
    VH_ENTER;
    VH_REQUIRE(out != in1 && out != in2);
    dev_ready(in1); dev_ready(in2); dev_ready(out);
    out->rns_mask = in1->rns_mask & in2->rns_mask;
    run_muls({{out, in1, in2, out->rns_mask}}, 2);
}

extern "C" void polynomial_multo_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in)
{
    VH_ENTER;
    dev_ready(in); dev_ready(out);
    out->rns_mask &= in->rns_mask;
    if (out->ntt->split_degree == 1) {
        run_muls({{out, out, in, out->rns_mask}}, 0);
        return;
    }
    RNS_Polynomial tmp = polynomial_new_RNS_polynomial(out->ntt->N, out->rns_mask, out->ntt);
    copy_rows(tmp, out, out->rns_mask);
    run_muls({{out, in, tmp, out->rns_mask}}, 0);
    free_RNS_polynomial(tmp);
}

// ---------------------------------------------------------------------------
// Additive and scalar ops (rns_polynomial.c:437-580)
// ---------------------------------------------------------------------------
static void binop(RNS_Polynomial out, RNS_Polynomial a, RNS_Polynomial b, int op)
{
    dev_ready(a); dev_ready(b); dev_ready(out);
    const uint64_t m = a->rns_mask & b->rns_mask;
    out->rns_mask = m;
    check_rows(out, m);
    std::vector<vh_item> items;
    for_active(m, out->ntt->l, [&](uint64_t i) { items.push_back(mk_item(drow(a, i), drow(out, i), i, drow(b, i))); });
    run_eltwise(items, DL(out), op, out->ntt->N);
}

extern "C" void polynomial_add_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, RNSc_Polynomial in2)
{ VH_ENTER; binop(out, in1, in2, VH_OP_ADD); }
extern "C" void polynomial_add_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2)
{ VH_ENTER; binop(out, in1, in2, VH_OP_ADD); }
extern "C" void polynomial_sub_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2)
{ VH_ENTER; binop(out, in1, in2, VH_OP_SUB); }
extern "C" void polynomial_sub_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, RNSc_Polynomial in2)
{ VH_ENTER; binop(out, in1, in2, VH_OP_SUB); }

extern "C" void polynomial_RNSc_negate(RNSc_Polynomial out, RNSc_Polynomial in)
{
    VH_ENTER;
    dev_ready(in); dev_ready(out);
    out->rns_mask = in->rns_mask;
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) { items.push_back(mk_item(drow(in, i), drow(out, i), i)); });
    run_eltwise(items, DL(out), VH_OP_NEGATE, out->ntt->N);
}
extern "C" void polynomial_RNS_negate(RNS_Polynomial out, RNS_Polynomial in) { polynomial_RNSc_negate(out, in); }

static void scale_op(RNS_Polynomial out, RNS_Polynomial in, const uint64_t *scales, uint64_t scale, int op)
{
    dev_ready(in); dev_ready(out);
    out->rns_mask = in->rns_mask;
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        const uint64_t q = out->ntt->ntt[i]->q, s = (scales ? scales[i] : scale) % q;
        items.push_back(mk_item(drow(in, i), drow(out, i), i, nullptr, s, shoup_of(s, q)));
    });
    run_eltwise(items, DL(out), op, out->ntt->N);
}

extern "C" void polynomial_scale_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t scale)
{ VH_ENTER; scale_op(out, in1, nullptr, scale, VH_OP_SCALE); }
extern "C" void polynomial_scale_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, uint64_t scale)
{ VH_ENTER; scale_op(out, in1, nullptr, scale, VH_OP_SCALE); }
extern "C" void polynomial_scale_RNS_polynomial_RNS(RNS_Polynomial out, RNS_Polynomial in1, uint64_t *scale)
{ VH_ENTER; scale_op(out, in1, scale, 0, VH_OP_SCALE); }
extern "C" void polynomial_scale_addto_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t scale)
{ VH_ENTER; scale_op(out, in1, nullptr, scale, VH_OP_FMA); }
extern "C" void polynomial_scale_addto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, uint64_t scale)
{ VH_ENTER; scale_op(out, in1, nullptr, scale, VH_OP_FMA); }

extern "C" void polynomial_RNSc_add_integer(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t in2)
{
    VH_ENTER;
    dev_ready(in1); dev_ready(out);
    out->rns_mask = in1->rns_mask;
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        const uint64_t q = out->ntt->ntt[i]->q;
        const uint64_t v = (in2 >> 63) ? (q - ((0 - in2) % q)) % q : in2 % q;
        items.push_back(mk_item(drow(in1, i), drow(out, i), i, nullptr, v));
    });
    run_eltwise(items, DL(out), VH_OP_ADD_AT0, out->ntt->N);
}

extern "C" void polynomial_RNS_add_integer(RNS_Polynomial out, RNS_Polynomial in1, uint64_t in2)
{
    VH_ENTER;
    dev_ready(in1); dev_ready(out);
    out->rns_mask = in1->rns_mask;
    copy_rows(out, in1, out->rns_mask);
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        const uint64_t q = out->ntt->ntt[i]->q;
        const uint64_t v = (in2 >> 63) ? negate_modq((0 - in2) % q, q) : in2 % q;
        items.push_back(mk_item(drow(out, i), drow(out, i), i, nullptr, v));
    });
    run_eltwise(items, DL(out), VH_OP_ADD_SCALAR, VPOOL(out->ntt)->n);
}

// ---------------------------------------------------------------------------
// ModUp digit and limb copies (rns_polynomial.c:824-854)
// ---------------------------------------------------------------------------
extern "C" void polynomial_RNSc_mod_reduce_lifted(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t idx)
{
    VH_ENTER;
    dev_ready(in); dev_ready(out);
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) { items.push_back(mk_item(drow(in, idx), drow(out, i), i)); });
    run_eltwise(items, DL(out), VH_OP_REDUCE, out->ntt->N);
}

extern "C" void polynomial_RNSc_mod_reduce(RNSc_Polynomial out, RNSc_Polynomial in)
{
    VH_ENTER;
    dev_ready(in); dev_ready(out);
    std::vector<vh_item> items;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        if (drow(in, i) != drow(out, i)) items.push_back(mk_item(drow(in, i), drow(out, i), i));
    });
    run_eltwise(items, DL(out), VH_OP_COPY, out->ntt->N);
}

// ---------------------------------------------------------------------------
// Rescale / ModDown: sequential division by the primes of divide_mask
// (rns_polynomial.c:905-976), all passes fused into one launch per <=16 primes.
// ---------------------------------------------------------------------------
struct vh_divc_h { uint64_t h, inv, inv_sh; };

// Device table of division constants for dividing the limbs of `mask` by the
// primes `divs` (= the bits of `chunk`, ascending): entry [k][i] = {half of p_k
// mod q_i (0 for floor division), p_k^-1 mod q_i, its Shoup factor}.  Cached
// per pool for the lifetime of the pool's prime list.
static const void *div_consts(incNTT t, uint64_t mask, uint64_t chunk, const std::vector<int> &divs, bool round)
{
    const uint64_t l = t->l;
    void *&d_consts = VPOOL(t)->div_consts[std::make_tuple(mask, chunk, round ? 1 : 0)];
    if (!d_consts) {
        std::vector<vh_divc_h> consts(divs.size() * l);
        for (size_t k = 0; k < divs.size(); k++) {
            const uint64_t p = t->ntt[divs[k]]->q, half = round ? p / 2 : 0;
            for (uint64_t i = 0; i < l; i++) {
                if (!ACTIVE(mask, i)) continue;
                const uint64_t q = t->ntt[i]->q;
                vh_divc_h &c = consts[k * l + i];
                c.h = half % q;
                c.inv = (i == (uint64_t)divs[k]) ? 0 : inverse_mod(p, q);
                c.inv_sh = shoup_of(c.inv, q);
            }
        }
        VH_CHECK(hipMalloc(&d_consts, consts.size() * sizeof(vh_divc_h)));
        VH_CHECK(hipMemcpy(d_consts, consts.data(), consts.size() * sizeof(vh_divc_h), hipMemcpyHostToDevice));
    }
    return d_consts;
}

// Constant rows of k_moddown_rescale_fp for the limbs of `mask`, dividing by the special
// primes dv (ascending, the limbs of `down`) and then by limb `resc`: the sequential round
// divisions (rns_polynomial.c:939-976) written as one affine map per limb.  With
// inv_j(q) = p_j^-1 mod q and h_j(q) = floor(p_j/2) mod q:
//   special k:  G = prod_{j<k} inv_j,  g_kj = -prod_{j<=j'<k} inv_j',
//               c = floor(p_k/2) + sum_{j<k} h_j prod_{j<=j'<k} inv_j'            (mod p_k)
//   limb i:     G_k = prod_{k<=k'<K} inv_k'  (mod q_i);  for i = resc: X = G_0, T_k = -G_k,
//               C = floor(q_r/2) + sum_k h_k G_k;  otherwise, with A = q_r^-1 mod q_i:
//               X = G_0 A, T_k = -G_k A, C = (h_r + sum_k h_k G_k) A, R = -A.
static const void *mdr_fp_consts(incNTT t, uint64_t mask, uint64_t down, const std::vector<int> &dv, int resc)
{
    const uint64_t l = t->l;
    void *&d = VPOOL(t)->div_consts[std::make_tuple(mask, down, 0x100 + resc)];
    if (d) return d;
    const int W = VH_MDR_W, K = (int)dv.size();
    VH_REQUIRE(K >= 1 && K <= 4 && W >= 4 + 4);
    std::vector<double2> cs((size_t)(4 + l) * W, make_double2(0.0, 0.0));
    auto put = [&](size_t at, uint64_t w, uint64_t q) {
        const double c = centred(w % q, q);
        cs[at] = make_double2(c, c / (double)q);
    };
    auto neg = [](uint64_t w, uint64_t q) { return w % q ? q - w % q : 0; };
    for (int k = 0; k < K; k++) {
        const uint64_t pk = t->ntt[dv[k]]->q;
        uint64_t G = 1, c = pk / 2;
        for (int j = 0; j < k; j++) G = mulmod_h(G, inverse_mod(t->ntt[dv[j]]->q % pk, pk), pk);
        for (int j = 0; j < k; j++) {
            uint64_t g = 1;
            for (int jj = j; jj < k; jj++) g = mulmod_h(g, inverse_mod(t->ntt[dv[jj]]->q % pk, pk), pk);
            c = (c + mulmod_h((t->ntt[dv[j]]->q / 2) % pk, g, pk)) % pk;
            put((size_t)k * W + 2 + j, neg(g, pk), pk);
        }
        put((size_t)k * W, G, pk);
        cs[(size_t)k * W + 1] = make_double2(centred(c, pk), 0.0);
    }
    const uint64_t qr = t->ntt[resc]->q;
    for (uint64_t i = 0; i < l; i++) {
        if (!ACTIVE(mask, i) || ACTIVE(down, i)) continue;
        const uint64_t qi = t->ntt[i]->q;
        std::vector<uint64_t> Gk(K + 1, 1);
        for (int k = K - 1; k >= 0; k--) Gk[k] = mulmod_h(Gk[k + 1], inverse_mod(t->ntt[dv[k]]->q % qi, qi), qi);
        uint64_t hs = 0;
        for (int k = 0; k < K; k++) hs = (hs + mulmod_h((t->ntt[dv[k]]->q / 2) % qi, Gk[k], qi)) % qi;
        const size_t row = (size_t)(4 + i) * W;
        if ((int)i == resc) {
            put(row, Gk[0], qi);
            cs[row + 2] = make_double2(centred((hs + qr / 2) % qi, qi), 0.0);
            for (int k = 0; k < K; k++) put(row + 4 + k, neg(Gk[k], qi), qi);
        } else {
            const uint64_t A = inverse_mod(qr % qi, qi);
            put(row, mulmod_h(Gk[0], A, qi), qi);
            put(row + 1, A, qi);
            cs[row + 2] = make_double2(centred(mulmod_h((hs + (qr / 2) % qi) % qi, A, qi), qi), 0.0);
            put(row + 3, neg(A, qi), qi);
            for (int k = 0; k < K; k++) put(row + 4 + k, neg(mulmod_h(Gk[k], A, qi), qi), qi);
        }
    }
    VH_CHECK(hipMalloc(&d, cs.size() * sizeof(double2)));
    VH_CHECK(hipMemcpy(d, cs.data(), cs.size() * sizeof(double2), hipMemcpyHostToDevice));
    return d;
}

static void run_division(const std::vector<RNS_Polynomial> &polys, uint64_t divide_mask, bool round)
{
    if (polys.empty()) return;
    incNTT t = polys[0]->ntt;
    const uint64_t l = t->l;
    // Group polynomials by active mask (they normally share one).
    std::map<uint64_t, std::vector<RNS_Polynomial>> groups;
    for (auto p : polys) {
        VH_REQUIRE(p->ntt == t);
        dev_ready(p);
        groups[p->rns_mask].push_back(p);
    }
    for (auto &g : groups) {
        uint64_t mask = g.first;
        uint64_t dm = divide_mask & mask;
        while (dm) {
            std::vector<int> divs;
            for (uint64_t i = 0; i < l && divs.size() < VH_MAX_DIV; i++) if (ACTIVE(dm, i)) divs.push_back((int)i);
            uint64_t chunk = 0;
            for (int d : divs) chunk |= 1ULL << d;
            std::vector<int> outs;
            for (uint64_t i = 0; i < l; i++) if (ACTIVE(mask, i) && !ACTIVE(chunk, i)) outs.push_back((int)i);
            const void *d_consts = div_consts(t, mask, chunk, divs, round);
            std::vector<vh_div_job> jobs;
            for (auto p : g.second) {
                vh_div_job jb;
                memset(&jb, 0, sizeof(jb));
                jb.base = VP(p)->d;
                jb.stride = t->N;
                jb.K = (int)divs.size();
                jb.nl = (int)l;
                for (size_t k = 0; k < divs.size(); k++) jb.div_limb[k] = divs[k];
                for (size_t k = 0; k < outs.size(); k++) jb.out_limb[k] = outs[k];
                jb.nout = (int)outs.size();
                jobs.push_back(jb);
            }
            if (!outs.empty()) {
                for (size_t off = 0; off < jobs.size(); off += VH_MAX_GRID_Y) {
                    const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, jobs.size() - off);
                    std::vector<vh_div_job> chunkv(jobs.begin() + off, jobs.begin() + off + cnt);
                    {
                        vh_prof_scope ps_(VH_PROF_DIVISION);
                        VH_LAUNCH(vh_launch_division(S(), vh_stage_vec(chunkv), cnt, (unsigned)outs.size(),
                                                VPOOL(t)->d_limbs, d_consts, t->N, (int)divs.size()));
                    }
                }
            }
            std::vector<vh_item> zi;
            for (auto p : g.second) for (int d : divs) zi.push_back(mk_item(nullptr, drow(p, d), d));
            run_eltwise(zi, VPOOL(t)->d_limbs, VH_OP_ZERO, t->N);
            mask &= ~chunk;
            dm &= ~chunk;
        }
        for (auto p : g.second) p->rns_mask = mask;
    }
}

extern "C" void polynomial_round_division_RNSc_wo_free(RNSc_Polynomial out, uint64_t divide_mask)
{
    VH_ENTER;
    if ((divide_mask & out->rns_mask) == 0) return;
    run_division({out}, divide_mask, true);
}

extern "C" void polynomial_floor_division_RNSc_wo_free(RNSc_Polynomial out, uint64_t divide_mask)
{
    VH_ENTER;
    if ((divide_mask & out->rns_mask) == 0) return;
    run_division({out}, divide_mask, false);
}

static int last_active(uint64_t mask)
{
    for (int i = 63; i >= 0; i--) if (ACTIVE(mask, i)) return i;
    return -1;
}

extern "C" void polynomial_round_division_RNSc(RNSc_Polynomial out)
{
    int la = last_active(out->rns_mask);
    if (la >= 0) polynomial_round_division_RNSc_wo_free(out, 1ULL << la);
}

extern "C" void polynomial_floor_division_RNSc(RNSc_Polynomial out)
{
    int la = last_active(out->rns_mask);
    if (la >= 0) polynomial_floor_division_RNSc_wo_free(out, 1ULL << la);
}

// ---------------------------------------------------------------------------
// Fast base extension (rns_polynomial.c:638-822) and scaled lift (:1463-1525)
// ---------------------------------------------------------------------------
extern "C" RNS_BaseConversionParams init_base_conversion_params(incNTT ntt, uint64_t in_mask, uint64_t out_mask)
{
    RNS_BaseConversionParams p = (RNS_BaseConversionParams)calloc(1, sizeof(*p));
    if (!p) exit(EXIT_FAILURE);
    p->in_mask = in_mask;
    p->out_mask = out_mask;
    p->D = (uint32_t *)malloc(sizeof(uint32_t) * (ntt->l ? ntt->l : 1));
    p->P = (uint32_t *)malloc(sizeof(uint32_t) * (ntt->l ? ntt->l : 1));
    for (uint64_t i = 0; i < ntt->l; i++) if (ACTIVE(in_mask, i)) p->D[p->w++] = (uint32_t)i;
    for (uint64_t i = 0; i < ntt->l; i++) if (ACTIVE(out_mask, i) && !ACTIVE(in_mask, i)) p->P[p->v++] = (uint32_t)i;
    if (p->v == 0) return p;
    VH_REQUIRE(p->w > 0);
    p->Dhat = (uint64_t *)malloc(sizeof(uint64_t) * p->w);
    for (uint32_t j = 0; j < p->w; j++) {
        const uint64_t qj = ntt->ntt[p->D[j]]->q;
        uint64_t prod = 1;
        for (uint32_t k = 0; k < p->w; k++) if (k != j) prod = mulmod_h(prod, ntt->ntt[p->D[k]]->q % qj, qj);
        p->Dhat[j] = inverse_mod(prod, qj);
    }
    p->D_mod_p = (uint64_t **)malloc(sizeof(uint64_t *) * p->v);
    for (uint32_t i = 0; i < p->v; i++) {
        p->D_mod_p[i] = (uint64_t *)malloc(sizeof(uint64_t) * p->w);
        const uint64_t pi = ntt->ntt[p->P[i]]->q;
        for (uint32_t j = 0; j < p->w; j++) {
            uint64_t prod = 1;
            for (uint32_t k = 0; k < p->w; k++) if (k != j) prod = mulmod_h(prod, ntt->ntt[p->D[k]]->q % pi, pi);
            p->D_mod_p[i][j] = prod;
        }
    }
    return p;
}

extern "C" void free_base_conversion_params(RNS_BaseConversionParams p)
{
    if (!p) return;
    if (p->D_mod_p) {
        for (uint32_t i = 0; i < p->v; i++) free(p->D_mod_p[i]);
        free(p->D_mod_p);
    }
    free(p->Dhat);
    free(p->D);
    free(p->P);
    free(p);
}

extern "C" void polynomial_base_conversion_RNSc(RNSc_Polynomial out, RNSc_Polynomial in, RNS_BaseConversionParams params)
{
    VH_ENTER;
    VH_REQUIRE(in->ntt->l == out->ntt->l && in->ntt->N == out->ntt->N);
    dev_ready(in); dev_ready(out);
    const uint64_t in_mask = in->rns_mask, out_mask = out->rns_mask;
    copy_rows(out, in, in_mask & out_mask);
    RNS_BaseConversionParams p = params;
    if (!p) p = init_base_conversion_params(in->ntt, in_mask, out_mask);
    else VH_REQUIRE(p->in_mask == in_mask && p->out_mask == out_mask);
    if (p->v > 0) {
        std::vector<ulonglong2> consts(p->w + (size_t)p->v * p->w);
        for (uint32_t j = 0; j < p->w; j++) {
            const uint64_t qj = in->ntt->ntt[p->D[j]]->q;
            consts[j] = make_ulonglong2(p->Dhat[j], shoup_of(p->Dhat[j], qj));
        }
        for (uint32_t t = 0; t < p->v; t++) {
            const uint64_t pt = out->ntt->ntt[p->P[t]]->q;
            for (uint32_t j = 0; j < p->w; j++) {
                const uint64_t d = p->D_mod_p[t][j] % pt;
                consts[p->w + (size_t)t * p->w + j] = make_ulonglong2(d, shoup_of(d, pt));
            }
        }
        vh_bconv_job jb;
        memset(&jb, 0, sizeof(jb));
        jb.in = VP(in)->d;
        jb.out = VP(out)->d;
        jb.stride = in->ntt->N;
        jb.w = (int)p->w;
        jb.v = (int)p->v;
        for (uint32_t j = 0; j < p->w; j++) jb.D[j] = (int)p->D[j];
        for (uint32_t t = 0; t < p->v; t++) jb.P[t] = (int)p->P[t];
        const void *d_consts = vh_stage(consts.data(), consts.size() * sizeof(ulonglong2));
        const vh_bconv_job *d_job = (const vh_bconv_job *)vh_stage(&jb, sizeof(jb));
        {
            vh_prof_scope ps_(VH_PROF_OTHER);
            VH_LAUNCH(vh_launch_base_conversion(S(), d_job, 1, p->v, DL(out), d_consts, out->ntt->N));
        }
    }
    if (!params) free_base_conversion_params(p);
    out->rns_mask = out_mask;
}

extern "C" void rns_compute_scaling_factors(uint64_t *delta_out, incNTT ntt, uint64_t in_mask, uint64_t out_mask)
{
    const uint64_t diff = out_mask & ~in_mask;
    for (uint64_t i = 0; i < ntt->l; i++) {
        delta_out[i] = 0;
        if (!ACTIVE(out_mask, i) || !ACTIVE(in_mask, i)) continue;
        const uint64_t qi = ntt->ntt[i]->q;
        uint64_t d = 1 % qi;
        for (uint64_t j = 0; j < ntt->l; j++) if (ACTIVE(diff, j)) d = mulmod_h(d, ntt->ntt[j]->q % qi, qi);
        delta_out[i] = d;
    }
}

extern "C" void polynomial_RNSc_scaled_lift(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t *delta)
{
    VH_ENTER;
    VH_REQUIRE(in->ntt->N == out->ntt->N);
    dev_ready(in); dev_ready(out);
    uint64_t local[64];
    if (!delta) { rns_compute_scaling_factors(local, out->ntt, in->rns_mask, out->rns_mask); delta = local; }
    std::vector<vh_item> scale, zero;
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t i) {
        const uint64_t q = out->ntt->ntt[i]->q;
        if (ACTIVE(in->rns_mask, i)) {
            const uint64_t s = delta[i] % q;
            scale.push_back(mk_item(drow(in, i), drow(out, i), i, nullptr, s, shoup_of(s, q)));
        } else {
            zero.push_back(mk_item(nullptr, drow(out, i), i));
        }
    });
    run_eltwise(scale, DL(out), VH_OP_SCALE, out->ntt->N);
    run_eltwise(zero, DL(out), VH_OP_ZERO, out->ntt->N);
}

// ---------------------------------------------------------------------------
// Automorphism (rns_polynomial.c:996-1035)
// ---------------------------------------------------------------------------
static void permute_into(RNS_Polynomial out, RNS_Polynomial in, uint64_t gen)
{
    VH_REQUIRE(out != in && gen > 0 && gen < 2 * out->ntt->N);
    dev_ready(in); dev_ready(out);
    out->rns_mask = in->rns_mask;
    std::vector<vh_item> perm, zero;
    const uint64_t l = std::min(out->ntt->l, out->allocated_l);
    for (uint64_t j = 0; j < l; j++) {
        if (ACTIVE(out->rns_mask, j))
            perm.push_back(mk_item(drow(in, j), drow(out, j), (uint32_t)j, nullptr, out->ntt->split_degree, 0, (uint32_t)gen));
        else
            zero.push_back(mk_item(nullptr, drow(out, j), (uint32_t)j));
    }
    if (!perm.empty())
        {
            vh_prof_scope ps_(VH_PROF_OTHER);
            VH_LAUNCH(vh_launch_permute(S(), vh_stage_vec(perm), (unsigned)perm.size(), DL(out), out->ntt->N));
        }
    run_eltwise(zero, DL(out), VH_OP_ZERO, out->ntt->N);
}

extern "C" void polynomial_RNSc_permute(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t gen)
{
    VH_ENTER;
    permute_into(out, in, gen);
}

// ---------------------------------------------------------------------------
// MLWE samples (mlwe.c:98-312)
// ---------------------------------------------------------------------------
// An MLWE sample whose rows are left uninitialised: engine temporaries and the
// serving runtime's outputs, which are written before they are read.
static RNS_MLWE alloc_sample_raw(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    RNS_MLWE c = (RNS_MLWE)malloc(sizeof(*c));
    if (!c) exit(EXIT_FAILURE);
    c->a = (RNS_Polynomial *)malloc(sizeof(RNS_Polynomial) * (r ? r : 1));
    if (!c->a) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < r; i++) c->a[i] = new_poly(N, mask, ntt, false, false);
    c->b = new_poly(N, mask, ntt, false, false);
    c->r = r;
    return c;
}

extern "C" RNS_MLWE vfhe_hip_mlwe_alloc_RNS_sample_uninit(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    VH_ENTER;
    return alloc_sample_raw(N, r, mask, ntt);
}

extern "C" void vfhe_hip_mlwe_alloc_RNS_samples_uninit(RNS_MLWE *out, uint64_t count, uint64_t N, uint64_t r,
                                                       uint64_t mask, incNTT ntt)
{
    VH_ENTER;
    VH_REQUIRE(out || !count);
    for (uint64_t k = 0; k < count; k++) out[k] = alloc_sample_raw(N, r, mask, ntt);
}

extern "C" RNS_MLWE mlwe_alloc_RNS_sample(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    RNS_MLWE c = (RNS_MLWE)malloc(sizeof(*c));
    if (!c) exit(EXIT_FAILURE);
    c->a = polynomial_new_RNS_polynomial_array(r, N, mask, ntt);
    c->b = polynomial_new_RNS_polynomial(N, mask, ntt);
    c->r = r;
    return c;
}

extern "C" RNSc_MLWE mlwe_alloc_RNSc_sample(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    return mlwe_alloc_RNS_sample(N, r, mask, ntt);
}

extern "C" RNS_MLWE *mlwe_alloc_RNS_sample_array(uint64_t size, uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    RNS_MLWE *v = (RNS_MLWE *)malloc(sizeof(RNS_MLWE) * (size ? size : 1));
    if (!v) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < size; i++) v[i] = mlwe_alloc_RNS_sample(N, r, mask, ntt);
    return v;
}

extern "C" void mlwe_copy_RNS_sample(RNS_MLWE out, RNS_MLWE in)
{
    for (uint64_t i = 0; i < out->r; i++) polynomial_copy_RNS_polynomial(out->a[i], in->a[i]);
    polynomial_copy_RNS_polynomial(out->b, in->b);
}

extern "C" RNS_MLWE *mlwe_create_copy_array(RNS_MLWE *in, uint64_t size)
{
    RNS_MLWE *v = mlwe_alloc_RNS_sample_array(size, in[0]->b->ntt->N, in[0]->r, in[0]->b->rns_mask, in[0]->b->ntt);
    for (uint64_t i = 0; i < size; i++) mlwe_copy_RNS_sample(v[i], in[i]);
    return v;
}

extern "C" void free_mlwe_RNS_sample(void *p)
{
    if (!p) return;
    RNS_MLWE c = (RNS_MLWE)p;
    free_RNS_polynomial_array(c->r, c->a);
    free_RNS_polynomial(c->b);
    free(c);
}

extern "C" void free_RNS_mlwe_array(uint64_t size, RNS_MLWE *v)
{
    for (uint64_t i = 0; i < size; i++) free_mlwe_RNS_sample(v[i]);
    free(v);
}

// Frees `count` samples under one acquisition of the engine lock (a serving
// front end returns a poll's worth of outputs at once).  The array itself stays
// the caller's.
extern "C" void vfhe_hip_free_mlwe_samples(RNS_MLWE *samples, uint64_t count)
{
    VH_ENTER;
    for (uint64_t i = 0; i < count; i++) free_mlwe_RNS_sample(samples[i]);
}

static std::vector<RNS_Polynomial> comps(RNS_MLWE c)
{
    std::vector<RNS_Polynomial> v(c->a, c->a + c->r);
    v.push_back(c->b);
    return v;
}

extern "C" void mlwe_RNS_trivial_sample_of_zero(RNS_MLWE out)
{
    VH_ENTER;
    std::vector<vh_item> items;
    for (auto p : comps(out)) {
        dev_ready(p);
        for (uint64_t i = 0; i < std::min(p->ntt->l, p->allocated_l); i++) items.push_back(mk_item(nullptr, drow(p, i), (uint32_t)i));
    }
    run_eltwise(items, DL(out->b), VH_OP_ZERO, out->b->ntt->N);
}

static void mlwe_ntt(RNS_MLWE out, RNS_MLWE in, int dir)
{
    VH_ENTER;
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    VH_REQUIRE(out->r == in->r);
    auto o = comps(out), i = comps(in);
    for (size_t k = 0; k < o.size(); k++) {
        RNS_Polynomial src = i[k];
        dev_ready(src); dev_ready(o[k]);
        o[k]->rns_mask = src->rns_mask;
        pairs.push_back({o[k], src});
    }
    ntt_rows(pairs, dir);
}

extern "C" void mlwe_RNSc_to_RNS(RNS_MLWE out, RNSc_MLWE in) { mlwe_ntt(out, in, 0); }
extern "C" void mlwe_RNS_to_RNSc(RNSc_MLWE out, RNS_MLWE in) { mlwe_ntt(out, in, 1); }

static void mlwe_mul_by_poly_mode(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly, int mode)
{
    VH_ENTER;
    std::vector<vh_mulreq> reqs;
    auto o = comps(out);
    dev_ready(poly);
    for (size_t k = 0; k < o.size(); k++) {
        RNS_Polynomial src = k < out->r ? in->a[k] : in->b;
        VH_REQUIRE(o[k] != src && o[k] != poly);
        dev_ready(src); dev_ready(o[k]);
        o[k]->rns_mask = src->rns_mask & poly->rns_mask;
        reqs.push_back({o[k], src, poly, o[k]->rns_mask});
    }
    run_muls(reqs, mode);
}

extern "C" void mlwe_RNS_mul_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly) { mlwe_mul_by_poly_mode(out, in, poly, 0); }
extern "C" void mlwe_RNS_mul_addto_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly) { mlwe_mul_by_poly_mode(out, in, poly, 1); }
extern "C" void mlwe_RNS_mul_subto_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly) { mlwe_mul_by_poly_mode(out, in, poly, 2); }

static void mlwe_binop(RNS_MLWE out, RNS_MLWE a, RNS_MLWE b, int op)
{
    VH_ENTER;
    std::vector<vh_item> items;
    for (uint64_t k = 0; k <= out->r; k++) {
        RNS_Polynomial o = k < out->r ? out->a[k] : out->b, x = k < out->r ? a->a[k] : a->b, y = k < out->r ? b->a[k] : b->b;
        dev_ready(o); dev_ready(x); dev_ready(y);
        o->rns_mask = x->rns_mask & y->rns_mask;
        for_active(o->rns_mask, o->ntt->l, [&](uint64_t i) { items.push_back(mk_item(drow(x, i), drow(o, i), (uint32_t)i, drow(y, i))); });
    }
    run_eltwise(items, DL(out->b), op, out->b->ntt->N);
}

extern "C" void mlwe_add_RNSc_sample(RNSc_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2) { mlwe_binop(out, in1, in2, VH_OP_ADD); }
extern "C" void mlwe_sub_RNSc_sample(RNSc_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2) { mlwe_binop(out, in1, in2, VH_OP_SUB); }

// mlwe.c:464-482: only the b component changes; a is copied through when out != in1.
static void mlwe_poly_op(RNS_MLWE out, RNS_MLWE in1, RNS_Polynomial in2, int op)
{
    VH_ENTER;
    if (out != in1)
        for (uint64_t i = 0; i < out->r; i++) polynomial_copy_RNS_polynomial(out->a[i], in1->a[i]);
    binop(out->b, in1->b, in2, op);
}

extern "C" void mlwe_add_RNSc_polynomial(RNSc_MLWE out, RNSc_MLWE in1, RNSc_Polynomial in2) { mlwe_poly_op(out, in1, in2, VH_OP_ADD); }
extern "C" void mlwe_sub_RNSc_polynomial(RNSc_MLWE out, RNSc_MLWE in1, RNSc_Polynomial in2) { mlwe_poly_op(out, in1, in2, VH_OP_SUB); }
extern "C" void mlwe_RNS_add_polynomial(RNS_MLWE out, RNS_MLWE in1, RNS_Polynomial in2) { mlwe_poly_op(out, in1, in2, VH_OP_ADD); }
extern "C" void mlwe_RNS_sub_polynomial(RNS_MLWE out, RNS_MLWE in1, RNS_Polynomial in2) { mlwe_poly_op(out, in1, in2, VH_OP_SUB); }

extern "C" void mlwe_scale_RNS_mlwe_RNS(RNS_MLWE c, uint64_t *scale)
{
    for (uint64_t i = 0; i < c->r; i++) polynomial_scale_RNS_polynomial_RNS(c->a[i], c->a[i], scale);
    polynomial_scale_RNS_polynomial_RNS(c->b, c->b, scale);
}

extern "C" void mlwe_round_division_RNSc(RNSc_MLWE out, uint64_t divide_mask)
{
    VH_ENTER;
    if (divide_mask == 0) return;
    std::vector<RNS_Polynomial> ps;
    for (auto p : comps(out)) if (divide_mask & p->rns_mask) ps.push_back(p);
    run_division(ps, divide_mask, true);
}

extern "C" uint64_t mlwe_extended_rank(uint64_t r) { return r * (r + 3) / 2; }

// mlwe_tensor_product (mlwe.c:665-709): quadratic slots q_ij (i <= j), linear slots, constant.
// Appends the products of one tensor product (mlwe.c:665-708) to `first` (plain
// products) and `second` (products added to a slot of `first`), so a batch of
// tensor products runs as two launches.
static void tensor_reqs(const std::vector<RNS_Polynomial> &out, RNS_MLWE in1, RNS_MLWE in2,
                        std::vector<vh_mulreq> &first, std::vector<vh_mulreq> &second)
{
    const uint64_t r = in1->r, R = mlwe_extended_rank(r);
    VH_REQUIRE(in2->r == r && out.size() == R + 1);
    auto m = [](RNS_Polynomial o, RNS_Polynomial a, RNS_Polynomial b) {
        dev_ready(a); dev_ready(b); dev_ready(o);
        o->rns_mask = a->rns_mask & b->rns_mask;
        return vh_mulreq{o, a, b, o->rns_mask};
    };
    uint64_t k = 0;
    for (uint64_t i = 0; i < r; i++)
        for (uint64_t j = i; j < r; j++) {
            first.push_back(m(out[k], in1->a[i], in2->a[j]));
            if (i != j) second.push_back(m(out[k], in1->a[j], in2->a[i]));
            k++;
        }
    for (uint64_t i = 0; i < r; i++) {
        first.push_back(m(out[k], in1->a[i], in2->b));
        second.push_back(m(out[k], in1->b, in2->a[i]));
        k++;
    }
    first.push_back(m(out[R], in1->b, in2->b));
}

// Rank-1 tensor products whose four inputs share one limb mask, in one launch
// (k_tensor1); returns false, doing nothing, when some job does not qualify.
static bool tensor1_batch(const std::vector<std::vector<RNS_Polynomial>> &outs, const std::vector<RNS_MLWE> &in1,
                          const std::vector<RNS_MLWE> &in2)
{
    for (size_t bi = 0; bi < in1.size(); bi++) {
        const uint64_t m = in1[bi]->b->rns_mask;
        if (in1[bi]->r != 1 || in2[bi]->r != 1 || in1[bi]->a[0]->rns_mask != m || in2[bi]->b->rns_mask != m ||
            in2[bi]->a[0]->rns_mask != m || in1[bi]->b->ntt != in1[0]->b->ntt || in2[bi]->b->ntt != in1[0]->b->ntt)
            return false;
    }
    incNTT t = in1[0]->b->ntt;
    std::vector<vh_tensor_job> jobs;
    for (size_t bi = 0; bi < in1.size(); bi++) {
        RNS_Polynomial a1 = in1[bi]->a[0], b1 = in1[bi]->b, a2 = in2[bi]->a[0], b2 = in2[bi]->b;
        for (RNS_Polynomial p : {a1, b1, a2, b2}) dev_ready(p);
        const uint64_t m = b1->rns_mask;
        for (int k = 0; k < 3; k++) {
            dev_ready(outs[bi][k]);
            outs[bi][k]->rns_mask = m;
            check_rows(outs[bi][k], m);
        }
        for_active(m, t->l, [&](uint64_t i) {
            jobs.push_back({drow(a1, i), drow(b1, i), drow(a2, i), drow(b2, i), drow(outs[bi][0], i),
                            drow(outs[bi][1], i), drow(outs[bi][2], i), (uint32_t)i});
        });
    }
    for (size_t off = 0; off < jobs.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, jobs.size() - off);
        std::vector<vh_tensor_job> chunk(jobs.begin() + off, jobs.begin() + off + cnt);
        vh_prof_scope ps_(VH_PROF_ELTWISE);
        VH_LAUNCH(vh_launch_tensor1(S(), vh_stage_vec(chunk), cnt, DL(in1[0]->b), t->N));
    }
    return true;
}

static void tensor_into(const std::vector<RNS_Polynomial> &out, RNS_MLWE in1, RNS_MLWE in2)
{
    std::vector<vh_mulreq> first, second;
    tensor_reqs(out, in1, in2, first, second);
    run_muls(first, 0);
    run_muls(second, 1);
}

extern "C" void mlwe_tensor_product(RNS_Polynomial *out, RNS_MLWE in1, RNS_MLWE in2)
{
    VH_ENTER;
    const uint64_t R = mlwe_extended_rank(in1->r);
    std::vector<RNS_Polynomial> o(out, out + R + 1);
    tensor_into(o, in1, in2);
}

// ---------------------------------------------------------------------------
// GHS hybrid key-switch (mlwe.c:557-598, gadget.c:22-37), batched over
// `count` inputs that share one key.  For every non-NULL ksk[i] and every
// active limb j of in->a[i] (ascending; the k-th such digit consumes
// ksk[i][k]), the digit is lifted to all extended limbs, forward-transformed
// and accumulated as out -= ksk[i][k] * NTT(lift(a_i mod q_j)).  Then the
// output is inverse-transformed, divided by the special primes in ascending
// order with rounding, and the pass-through components are added.
// ---------------------------------------------------------------------------
// Gadget products (gadget.c:5-37) for a batch of jobs: job i sets
//   outs[i] = (add ? +1 : -1) * sum over its sources (poly, keys) of
//             sum over the active limbs j of poly: NTT(lift_j(poly)) * keys[k++]
// in NTT form on the extended mask of its keys.  Jobs may differ in level and
// key and share the ring and the output rank.  The key-switch and the MGSW
// external product are this plus their own ModDown.
struct vh_gsrc {
    RNS_Polynomial poly;
    RNS_MLWE *keys;
};

// The fused ModUp row pass + key inner product (k_ntt16_row_mac) runs when every limb of
// `mask` is on the FP64 path, the ring is not split, N = 2^13..2^16 and the output has at
// most two components.
static bool fused_mac_ok(incNTT t, uint64_t mask, uint64_t comps)
{
    const vh_pool *P = VPOOL(t);
    bool ok = OPT.ks_fused && OPT.ntt_mode != 1 && t->split_degree == 1 && comps <= 2 && P->logn >= 13 &&
              P->logn <= 16;
    for (uint64_t i = 0; i < t->l && ok; i++) if (ACTIVE(mask, i)) ok = P->h_limbs[i].fp64;
    return ok;
}

// {P mod q_i centred, RN(that / q_i)} for every limb i of `level`, P the product of the
// primes of ext & ~level (the special primes of a key-switch).
static const double2 *pmod_consts(incNTT t, uint64_t level, uint64_t ext)
{
    void *&d = VPOOL(t)->div_consts[std::make_tuple(ext, level, 0x200)];
    if (d) return (const double2 *)d;
    std::vector<double2> v(t->l, make_double2(0.0, 0.0));
    for (uint64_t i = 0; i < t->l; i++) {
        if (!ACTIVE(level, i)) continue;
        const uint64_t qi = t->ntt[i]->q;
        uint64_t pm = 1;
        for (uint64_t k = 0; k < t->l; k++)
            if (ACTIVE(ext, k) && !ACTIVE(level, k)) pm = mulmod_h(pm, t->ntt[k]->q % qi, qi);
        const double c = centred(pm, qi);
        v[i] = make_double2(c, c / (double)qi);
    }
    VH_CHECK(hipMalloc(&d, v.size() * sizeof(double2)));
    VH_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice));
    return (const double2 *)d;
}

// fold (optional, per job): NTT-form polynomials added to the output components as
// P * fold[i][c] on the source's level limbs, inside the fused key inner product (the
// caller has checked fused_mac_ok; null entries add nothing).
static void gadget_batch(const std::vector<RNS_MLWE> &outs, const std::vector<std::vector<vh_gsrc>> &srcs, bool add,
                         const std::vector<std::vector<RNS_Polynomial>> *fold = nullptr)
{
    const size_t B = outs.size();
    VH_REQUIRE(B == srcs.size() && B > 0);
    const uint64_t r = outs[0]->r, Cc = r + 1;
    VH_REQUIRE(Cc <= VH_MAX_COMPS);
    incNTT t = nullptr;
    struct digit { uint64_t src, limb, entry; };
    struct job {
        uint64_t ext, next, nd, kptr_off, modup_off;
        std::vector<int> E;
        std::vector<digit> D;
    };
    std::vector<job> J(B);
    std::vector<const uint64_t *> kptr;
    uint64_t modup_rows = 0;
    for (size_t bi = 0; bi < B; bi++) {
        VH_REQUIRE(!srcs[bi].empty() && outs[bi]->r == r);
        job &jb = J[bi];
        RNS_MLWE first = srcs[bi][0].keys[0];
        jb.ext = first->b->rns_mask;
        if (!t) t = first->b->ntt;
        VH_REQUIRE(first->b->ntt == t);
        for (uint64_t i = 0; i < t->l; i++) if (ACTIVE(jb.ext, i)) jb.E.push_back((int)i);
        jb.next = jb.E.size();
        for (size_t si = 0; si < srcs[bi].size(); si++) {
            RNS_Polynomial poly = srcs[bi][si].poly;
            VH_REQUIRE(poly->ntt == t);
            dev_ready(poly);
            uint64_t k = 0;
            for (uint64_t j = 0; j < t->l; j++) if (ACTIVE(poly->rns_mask, j)) jb.D.push_back({si, j, k++});
        }
        jb.nd = jb.D.size();
        for (auto p : comps(outs[bi])) { dev_ready(p); p->rns_mask = jb.ext; check_rows(p, jb.ext); }
        // key rows: kptr[kptr_off + (ti * nd + d) * Cc + c]
        jb.kptr_off = kptr.size();
        kptr.resize(kptr.size() + jb.next * jb.nd * Cc);
        for (uint64_t d = 0; d < jb.nd; d++) {
            RNS_MLWE e = srcs[bi][jb.D[d].src].keys[jb.D[d].entry];
            VH_REQUIRE(e->r == r && e->b->ntt == t);
            for (uint64_t c = 0; c < Cc; c++) {
                RNS_Polynomial kp = c < r ? e->a[c] : e->b;
                dev_ready(kp);
                for (uint64_t ti = 0; ti < jb.next; ti++)
                    kptr[jb.kptr_off + (ti * jb.nd + d) * Cc + c] = drow(kp, jb.E[ti]);
            }
        }
        jb.modup_off = modup_rows;
        modup_rows += jb.nd * jb.next;
    }
    vh_pool *P = VPOOL(t);
    const uint64_t N = t->N, l = t->l;

    if (modup_rows > 0) {
        // ModUp: lift every digit to every extended limb, then forward NTT.
        uint64_t *modup = (uint64_t *)vh_dmalloc(modup_rows * N * 8);
        std::vector<vh_item> lift;
        lift.reserve(modup_rows);
        bool fuse = true;   // FP64 path everywhere: the lift folds into the NTT's first pass
        uint64_t maxnext = 0;
        for (size_t bi = 0; bi < B; bi++) {
            const job &jb = J[bi];
            maxnext = std::max(maxnext, jb.next);
            for (int e : jb.E) fuse = fuse && P->h_limbs[e].fp64;
            for (auto &dg : jb.D) fuse = fuse && P->h_limbs[dg.limb].fp64;
            for (uint64_t d = 0; d < jb.nd; d++) {
                const uint64_t *src = drow(srcs[bi][jb.D[d].src].poly, jb.D[d].limb);
                for (uint64_t ti = 0; ti < jb.next; ti++)
                    lift.push_back(mk_item(src, modup + (jb.modup_off + d * jb.next + ti) * N, (uint32_t)jb.E[ti]));
            }
        }
        // On the FP64 path (every prime < 2^50) the forward NTT mod q_t accepts any
        // input below 2^50, so the digit row is transformed straight from the input;
        // otherwise it is reduced first.
        if (!fuse) run_eltwise(lift, P->d_limbs, VH_OP_REDUCE, N);
        std::vector<vh_item> nt;
        nt.reserve(lift.size() * t->split_degree);
        for (auto &it : lift)
            for (uint64_t k = 0; k < t->split_degree; k++)
                nt.push_back(mk_item((fuse ? it.src : it.dst) + k * P->n, it.dst + k * P->n, it.limb));
        const int logn = ilog2(P->n);
        // Fused form: column pass here, then the row pass of every digit and the key
        // inner product in one kernel (the transformed digits stay in registers).
        uint64_t umask = 0;
        for (auto &jb : J) umask |= jb.ext;
        const bool fused_mac = fuse && fused_mac_ok(t, umask, Cc);
        VH_REQUIRE(!fold || fused_mac);
        if (fused_mac) {
            for (size_t off = 0; off < nt.size(); off += VH_MAX_GRID_Y) {
                const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, nt.size() - off);
                std::vector<vh_item> chunk(nt.begin() + off, nt.begin() + off + cnt);
                vh_prof_scope ps_(VH_PROF_NTT_FP64);
                // One tile per workgroup (the grid runs tile-major, so the workgroups in
                // flight share two or three digit rows, which stay in the L2).
                if (OPT.ntt_mode == 2 && C.pipe_blocks > 0)
                    VH_LAUNCH(vh_launch_ntt16p(S(), vh_stage_vec(chunk), cnt, P->d_limbs, 0, logn, 0, logn - 8, 1, 2, 0,
                                              UINT_MAX, col_wide(cnt, logn)));
                else
                    VH_LAUNCH(vh_launch_ntt16(S(), vh_stage_vec(chunk), cnt, P->d_limbs, 0, logn, 0, logn - 8, 1, 2, 0));
            }
        } else {
            run_ntt(nt, P->d_limbs, P->h_limbs.data(), P->n, 0);
        }

        const uint64_t *const *d_kptr = vh_stage_vec(kptr);
        if (t->split_degree == 1) {
            // FP64 accumulation needs canonical key rows: engine-made keys are; it is
            // used only when every extended limb is on the FP64 path.
            bool fp_mac = true;
            for (auto &jb : J) for (int e : jb.E) fp_mac = fp_mac && P->h_limbs[e].fp64;
            std::vector<vh_mac_job> jobs(B);
            for (size_t bi = 0; bi < B; bi++) {
                const job &src = J[bi];
                vh_mac_job &jb = jobs[bi];
                memset(&jb, 0, sizeof(jb));
                auto oc = comps(outs[bi]);
                for (uint64_t c = 0; c < Cc; c++) jb.out[c] = VP(oc[c])->d;
                jb.modup = modup + src.modup_off * N;
                jb.kptr_off = src.kptr_off;
                jb.comps = (int)Cc;
                jb.add = add ? 1 : 0;
                jb.digits = (int)src.nd;
                jb.next = (int)src.next;
                for (uint64_t ti = 0; ti < src.next; ti++) jb.limb[ti] = src.E[ti];
                if (fold) {
                    const uint64_t level = srcs[bi][0].poly->rns_mask;
                    jb.pass_mask = level;
                    jb.pmod = pmod_consts(t, level, src.ext);
                    for (uint64_t c = 0; c < Cc; c++) {
                        RNS_Polynomial f = (*fold)[bi][c];
                        if (!f) continue;
                        dev_ready(f);
                        VH_REQUIRE(f->ntt == t && (f->rns_mask & level) == level);
                        jb.pass[c] = VP(f)->d;
                    }
                }
            }
            for (size_t off = 0; off < B; off += VH_MAX_GRID_Y) {
                const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, B - off);
                std::vector<vh_mac_job> chunk(jobs.begin() + off, jobs.begin() + off + cnt);
                {
                    vh_prof_scope ps_(VH_PROF_KSMAC);
                    if (fused_mac)
                        VH_LAUNCH(vh_launch_ntt16_row_mac(S(), vh_stage_vec(chunk), cnt, (unsigned)maxnext, P->d_limbs,
                                                         d_kptr, logn, (int)Cc));
                    else
                        VH_LAUNCH(vh_launch_keyswitch_mac(S(), vh_stage_vec(chunk), cnt, (unsigned)maxnext, P->d_limbs,
                                                         d_kptr, N, fp_mac ? 1 : 0));
                }
            }
        } else {
            // Split-degree rings: zero the outputs, then accumulate one digit at a time
            // with the wrap-aware product (rns_polynomial.c:398-435).
            std::vector<vh_item> z;
            for (size_t bi = 0; bi < B; bi++)
                for (auto p : comps(outs[bi]))
                    for (uint64_t i = 0; i < std::min(l, p->allocated_l); i++) z.push_back(mk_item(nullptr, drow(p, i), (uint32_t)i));
            run_eltwise(z, P->d_limbs, VH_OP_ZERO, N);
            for (size_t bi = 0; bi < B; bi++) {
                const job &jb = J[bi];
                auto oc = comps(outs[bi]);
                // One digit's rows may not be accumulated concurrently into one output row:
                // launch per digit so the read-modify-write of out is ordered.
                for (uint64_t d = 0; d < jb.nd; d++) {
                    std::vector<vh_item> chunk;
                    for (uint64_t c = 0; c < Cc; c++)
                        for (uint64_t ti = 0; ti < jb.next; ti++) {
                            vh_item it = mk_item(kptr[jb.kptr_off + (ti * jb.nd + d) * Cc + c], drow(oc[c], jb.E[ti]),
                                                 (uint32_t)jb.E[ti], modup + (jb.modup_off + d * jb.next + ti) * N);
                            it.aux = (uint32_t)t->split_degree;
                            it.scalar = (uint64_t)(uintptr_t)P->d_w[jb.E[ti]];
                            chunk.push_back(it);
                        }
                    vh_prof_scope ps_(VH_PROF_ELTWISE);
                    VH_LAUNCH(vh_launch_split_mul(S(), vh_stage_vec(chunk), (unsigned)chunk.size(), P->d_limbs,
                                                 add ? 1 : 2, P->n, t->split_degree));
                }
            }
        }
        vh_dfree(modup);
    } else {
        VH_REQUIRE(!fold);
        std::vector<vh_item> z;
        for (size_t bi = 0; bi < B; bi++)
            for (auto p : comps(outs[bi]))
                for (uint64_t i = 0; i < std::min(l, p->allocated_l); i++) z.push_back(mk_item(nullptr, drow(p, i), (uint32_t)i));
        run_eltwise(z, P->d_limbs, VH_OP_ZERO, N);
    }

}

// mlwe_RNSc_GHS_hybrid_keyswitch (mlwe.c:557-598) for a batch whose jobs may
// differ in level (masks) and key: job i switches ins[i] with ksks[i].  All jobs
// share the ring and the output rank.
// rescale (optional, per job): after the key-switch, also round-divide by these
// primes (the CKKS rescale of mlwe_multiply's caller, ckks.py:128-148).  A job
// whose rescale is one prime is finished by k_moddown_rescale: ModDown, the
// pass-through add and the rescale in one pass over its rows.
// pass_ntt: the pass-through components of ins are in NTT form and go into the fused key
// inner product as P * d (ModDown(y + P d) == ModDown(y) + d exactly, limb by limb: every
// round division by a special prime p maps y + p P' d to ModDown_p(y) + P' d); the
// key-switched components are in coefficient form as always.
static void keyswitch_batch(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &ins,
                            const std::vector<RNS_MLWE **> &ksks, const std::vector<uint64_t> &rescale = {},
                            bool pass_ntt = false)
{
    const size_t B = outs.size();
    VH_REQUIRE(B == ins.size() && B == ksks.size() && B > 0);
    std::vector<std::vector<vh_gsrc>> srcs(B);
    for (size_t bi = 0; bi < B; bi++) {
        VH_REQUIRE(ksks[bi] != nullptr && outs[bi] != ins[bi] && outs[bi]->r == outs[0]->r);
        for (uint64_t i = 0; i < ins[bi]->r; i++)
            if (ksks[bi][i]) srcs[bi].push_back({ins[bi]->a[i], ksks[bi][i]});
        VH_REQUIRE(!srcs[bi].empty());
        for (auto p : comps(ins[bi])) dev_ready(p);
    }
    // Pass-through components keep their key: added after the ModDown (mlwe.c:588-597), or
    // folded into the key inner product.  pass[bi][c]: the input added to output component c.
    std::vector<std::vector<RNS_Polynomial>> pass(B);
    for (size_t bi = 0; bi < B; bi++) {
        const auto oc = comps(outs[bi]);
        pass[bi].assign(oc.size(), nullptr);
        uint64_t keep = 0;
        for (uint64_t i = 0; i < ins[bi]->r; i++) if (!ksks[bi][i]) pass[bi][keep++] = ins[bi]->a[i];
        pass[bi][oc.size() - 1] = ins[bi]->b;
    }
    if (pass_ntt) {
        gadget_batch(outs, srcs, false, &pass);
        for (auto &pv : pass) pv.assign(pv.size(), nullptr);
        g_ks_fold_jobs += (long)B;
    } else {
        gadget_batch(outs, srcs, false);
    }
    const uint64_t l = outs[0]->b->ntt->l, N = outs[0]->b->ntt->N;
    vh_pool *P = VPOOL(outs[0]->b->ntt);

    // Back to coefficient form, then ModDown by the special primes of each job.
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    std::map<uint64_t, std::vector<RNS_Polynomial>> by_divide;
    for (size_t bi = 0; bi < B; bi++)
        for (auto p : comps(outs[bi])) {
            pairs.push_back({p, p});
            by_divide[p->rns_mask & ~ins[bi]->b->rns_mask].push_back(p);
        }
    ntt_rows(pairs, 1);

    // Jobs finished in one pass: a one-prime rescale, at most VH_MAX_DIV special primes,
    // one mask for every output component, pass-through inputs covering it.
    std::vector<char> fused(B, 0);
    for (size_t bi = 0; bi < B && !rescale.empty(); bi++) {
        const auto oc = comps(outs[bi]);
        const uint64_t m = oc[0]->rns_mask, down = m & ~ins[bi]->b->rns_mask, keep = m & ~down;
        bool ok = __builtin_popcountll(rescale[bi]) == 1 && (rescale[bi] & keep) && down &&
                  __builtin_popcountll(down) <= VH_MAX_DIV;
        for (size_t c = 0; c < oc.size() && ok; c++)
            ok = oc[c]->rns_mask == m && (!pass[bi][c] || (pass[bi][c]->rns_mask & keep) == keep);
        fused[bi] = ok;
    }
    for (auto &g : by_divide) {
        std::vector<RNS_Polynomial> rest;
        for (auto p : g.second) {
            bool f = false;
            for (size_t bi = 0; bi < B && !f; bi++)
                if (fused[bi]) for (auto o : comps(outs[bi])) f = f || o == p;
            if (!f) rest.push_back(p);
        }
        if (g.first) run_division(rest, g.first, true);
    }
    std::vector<vh_item> adds;
    std::vector<vh_mdr_job> mdr;
    std::vector<vh_item> zero;
    uint64_t mdr_mask = 0, mdr_down = 0, mdr_resc = 0;
    for (size_t bi = 0; bi < B; bi++) {
        const auto oc = comps(outs[bi]);
        if (fused[bi]) {
            const uint64_t m = oc[0]->rns_mask, down = m & ~ins[bi]->b->rns_mask, keep = m & ~down;
            if (mdr.empty()) { mdr_mask = m; mdr_down = down; mdr_resc = rescale[bi]; }
            if (m != mdr_mask || down != mdr_down || rescale[bi] != mdr_resc) { fused[bi] = 0; }
            else {
                for (size_t c = 0; c < oc.size(); c++) {
                    vh_mdr_job jb;
                    memset(&jb, 0, sizeof(jb));
                    jb.base = VP(oc[c])->d;
                    jb.add = pass[bi][c] ? VP(pass[bi][c])->d : nullptr;
                    jb.stride = N;
                    jb.nl = (int)l;
                    for (uint64_t i = 0; i < l; i++) if (ACTIVE(down, i)) jb.div_limb[jb.K++] = (int)i;
                    jb.resc = __builtin_ctzll(rescale[bi]);
                    for (uint64_t i = 0; i < l; i++)
                        if (ACTIVE(keep, i) && !ACTIVE(rescale[bi], i)) jb.out_limb[jb.nout++] = (int)i;
                    mdr.push_back(jb);
                    for_active(down | rescale[bi], l, [&](uint64_t i) { zero.push_back(mk_item(nullptr, drow(oc[c], i), (uint32_t)i)); });
                    oc[c]->rns_mask = keep & ~rescale[bi];
                }
                continue;
            }
            // a job of another mask or prime set: the sequential path (rare)
            run_division(oc, down, true);
        }
        for (size_t c = 0; c < oc.size(); c++) {
            RNS_Polynomial o = oc[c], x = pass[bi][c];
            if (!x) continue;
            const uint64_t m = o->rns_mask & x->rns_mask;
            o->rns_mask = m;
            for_active(m, l, [&](uint64_t i) { adds.push_back(mk_item(drow(o, i), drow(o, i), (uint32_t)i, drow(x, i))); });
        }
    }
    run_eltwise(adds, P->d_limbs, VH_OP_ADD, N);
    if (!mdr.empty()) {
        std::vector<int> dv, rv = {__builtin_ctzll(mdr_resc)};
        for (uint64_t i = 0; i < l; i++) if (ACTIVE(mdr_down, i)) dv.push_back((int)i);
        incNTT t = outs[0]->b->ntt;
        bool fp = OPT.mdr_fp64 && dv.size() <= 4;
        for_active(mdr_mask, l, [&](uint64_t i) { fp = fp && P->h_limbs[i].fp64; });
        const void *cd = fp ? nullptr : div_consts(t, mdr_mask, mdr_down, dv, true);
        const void *cr = fp ? nullptr : div_consts(t, mdr_mask & ~mdr_down, mdr_resc, rv, true);
        const void *cf = fp ? mdr_fp_consts(t, mdr_mask, mdr_down, dv, rv[0]) : nullptr;
        for (size_t off = 0; off < mdr.size(); off += VH_MAX_GRID_Y) {
            const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, mdr.size() - off);
            std::vector<vh_mdr_job> chunk(mdr.begin() + off, mdr.begin() + off + cnt);
            vh_prof_scope ps_(VH_PROF_DIVISION);
            if (fp) {
                VH_LAUNCH(vh_launch_moddown_rescale_fp(S(), vh_stage_vec(chunk), cnt, P->d_limbs, cf, N));
                g_mdr_fp64_launches++;
            }
            else VH_LAUNCH(vh_launch_moddown_rescale(S(), vh_stage_vec(chunk), cnt, P->d_limbs, cd, cr, N, (int)dv.size()));
        }
        run_eltwise(zero, P->d_limbs, VH_OP_ZERO, N);
    }
    // jobs with a rescale not done above
    std::map<uint64_t, std::vector<RNS_Polynomial>> resc_groups;
    for (size_t bi = 0; bi < B && !rescale.empty(); bi++)
        if (!fused[bi] && rescale[bi]) for (auto o : comps(outs[bi])) resc_groups[rescale[bi]].push_back(o);
    for (auto &g : resc_groups) run_division(g.second, g.first, true);
}

extern "C" void mlwe_RNSc_GHS_hybrid_keyswitch(RNSc_MLWE out, RNSc_MLWE in, RNS_MLWE **ksk, uint64_t lvl)
{
    (void)lvl;
    VH_ENTER;
    keyswitch_batch({out}, {in}, {ksk});
}

extern "C" void mlwe_automorphism_RNSc_GHS(RNSc_MLWE out, RNSc_MLWE in, uint64_t gen, RNS_MLWE **ksk, uint64_t lvl)
{
    (void)lvl;
    VH_ENTER;
    RNSc_MLWE tmp = alloc_sample_raw(out->a[0]->ntt->N, out->r, out->a[0]->rns_mask, out->a[0]->ntt);
    for (uint64_t i = 0; i < out->r; i++) permute_into(tmp->a[i], in->a[i], gen);
    permute_into(tmp->b, in->b, gen);
    keyswitch_batch({out}, {tmp}, {ksk});
    free_mlwe_RNS_sample(tmp);
}

// mlwe_automorphism_RNSc_GHS (mlwe.c:314-326) for many inputs, each with its own
// Galois element and key: permute every component, then one batched key-switch.
static void automorphism_batch(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &ins,
                               const std::vector<uint64_t> &gens, const std::vector<RNS_MLWE **> &ksks)
{
    const size_t B = outs.size();
    VH_REQUIRE(B > 0 && ins.size() == B && gens.size() == B && ksks.size() == B);
    std::vector<RNS_MLWE> tmp(B);
    for (size_t bi = 0; bi < B; bi++) {
        RNS_MLWE o = outs[bi], x = ins[bi];
        VH_REQUIRE(o != x && o->r == x->r);
        tmp[bi] = alloc_sample_raw(o->a[0]->ntt->N, o->r, o->a[0]->rns_mask, o->a[0]->ntt);
        for (uint64_t i = 0; i < o->r; i++) permute_into(tmp[bi]->a[i], x->a[i], gens[bi]);
        permute_into(tmp[bi]->b, x->b, gens[bi]);
    }
    keyswitch_batch(outs, tmp, ksks);
    for (auto t : tmp) free_mlwe_RNS_sample(t);
}

// mlwe_RNSc_GHS_hybrid_keyswitch (mlwe.c:557-598) over `count` inputs sharing one key.
extern "C" void vfhe_hip_mlwe_keyswitch_batch(RNS_MLWE *out, RNS_MLWE *in, uint64_t count, RNS_MLWE **ksk)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(out && in && ksk);
    keyswitch_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in, in + count),
                    std::vector<RNS_MLWE **>(count, ksk));
}

extern "C" void vfhe_hip_mlwe_automorphism_batch(RNS_MLWE *out, RNS_MLWE *in, uint64_t count, uint64_t gen,
                                                 RNS_MLWE **ksk)
{
    VH_ENTER;
    if (count == 0) return;
    automorphism_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in, in + count),
                       std::vector<uint64_t>(count, gen), std::vector<RNS_MLWE **>(count, ksk));
}

extern "C" void vfhe_hip_mlwe_automorphism_batch_multi(RNS_MLWE *out, RNS_MLWE *in, uint64_t count,
                                                       const uint64_t *gens, RNS_MLWE ***ksks)
{
    VH_ENTER;
    if (count == 0) return;
    automorphism_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in, in + count),
                       std::vector<uint64_t>(gens, gens + count), std::vector<RNS_MLWE **>(ksks, ksks + count));
}

// A view of p restricted to `mask`: shares p's device rows and host mirror.
static RNS_Polynomial view_poly(RNS_Polynomial p, uint64_t mask)
{
    dev_ready(p);
    check_rows(p, mask);
    vh_poly *src = VP(p);
    vh_poly *v = new vh_poly();
    v->pub = src->pub;
    v->pub.rns_mask = mask;
    const uint64_t l = p->allocated_l;
    v->pub.coeffs = (uint64_t **)malloc(sizeof(uint64_t *) * (l ? l : 1));
    if (!v->pub.coeffs) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < l; i++) v->pub.coeffs[i] = p->coeffs[i];
    v->d = src->d;
    v->hblock = src->hblock;
    v->host_dirty = 0;
    v->borrowed = 1;
    return &v->pub;
}

extern "C" RNS_MLWE *vfhe_hip_mlwe_view_array(RNS_MLWE *in, uint64_t count, uint64_t mask)
{
    VH_ENTER;
    RNS_MLWE *out = (RNS_MLWE *)malloc(sizeof(RNS_MLWE) * (count ? count : 1));
    if (!out) exit(EXIT_FAILURE);
    for (uint64_t k = 0; k < count; k++) {
        RNS_MLWE c = (RNS_MLWE)malloc(sizeof(*c));
        if (!c) exit(EXIT_FAILURE);
        c->r = in[k]->r;
        c->a = (RNS_Polynomial *)malloc(sizeof(RNS_Polynomial) * (c->r ? c->r : 1));
        if (!c->a) exit(EXIT_FAILURE);
        for (uint64_t i = 0; i < c->r; i++) c->a[i] = view_poly(in[k]->a[i], mask);
        c->b = view_poly(in[k]->b, mask);
        out[k] = c;
    }
    return out;
}

extern "C" void *vfhe_hip_current_stream(void)
{
    VH_ENTER;
    return (void *)S();
}

// mlwe_multiply (mlwe.c:711-760), batched: tensor, INTT, key-switch, then either
// NTT (mlwe_multiply's contract) or the CKKS rescale by divide_mask directly
// (INTT(NTT(x)) == x exactly, so the pair the reference runs between the two
// calls is skipped without changing a bit).
static void multiply_batch(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &in1,
                           const std::vector<RNS_MLWE> &in2, const std::vector<RNS_MLWE **> &ksks,
                           const std::vector<uint64_t> &divide_masks)
{
    const size_t B = outs.size();
    VH_REQUIRE(B > 0 && in1.size() == B && in2.size() == B && ksks.size() == B && divide_masks.size() == B);
    const uint64_t N = in1[0]->b->ntt->N, r = in1[0]->r, R = mlwe_extended_rank(r);
    incNTT t = in1[0]->b->ntt;
    std::vector<vh_mulreq> first, second;
    std::vector<std::vector<RNS_Polynomial>> slots(B);
    auto tensors = [&]() {
        if (tensor1_batch(slots, in1, in2)) return;
        for (size_t bi = 0; bi < B; bi++) tensor_reqs(slots[bi], in1[bi], in2[bi], first, second);
        run_muls(first, 0);
        run_muls(second, 1);
    };
    if (!ksks[0]) {
        for (size_t bi = 0; bi < B; bi++) {
            VH_REQUIRE(!ksks[bi] && divide_masks[bi] == 0 && outs[bi]->r == R);
            slots[bi] = comps(outs[bi]);
        }
        tensors();
        return;
    }
    std::vector<RNS_MLWE> ext(B);
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    // With the fused key inner product the pass-through slots stay in NTT form and enter
    // it as P * d: only the key-switched slots go back to coefficient form.
    uint64_t kmask = 0;
    for (size_t bi = 0; bi < B; bi++)
        for (uint64_t i = 0; i < R; i++)
            if (ksks[bi] && ksks[bi][i] && ksks[bi][i][0]) kmask |= ksks[bi][i][0]->b->rns_mask;
    const bool fold = OPT.ks_fold && kmask && fused_mac_ok(t, kmask, r + 1);
    for (size_t bi = 0; bi < B; bi++) {
        VH_REQUIRE(ksks[bi] && outs[bi]->r == r && in1[bi]->r == r && in1[bi]->b->ntt == t);
        ext[bi] = alloc_sample_raw(N, R, in1[bi]->b->rns_mask, t);
        slots[bi] = comps(ext[bi]);
        for (uint64_t i = 0; i < R; i++)
            if (!fold || ksks[bi][i]) pairs.push_back({ext[bi]->a[i], ext[bi]->a[i]});
        if (!fold) pairs.push_back({ext[bi]->b, ext[bi]->b});
    }
    tensors();
    ntt_rows(pairs, 1);
    keyswitch_batch(outs, ext, ksks, divide_masks, fold);   // also rescales the jobs that ask for it
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> o;
    for (size_t bi = 0; bi < B; bi++)
        if (divide_masks[bi] == 0)
            for (auto p : comps(outs[bi])) o.push_back({p, p});
    ntt_rows(o, 0);
    for (auto e : ext) free_mlwe_RNS_sample(e);
}

extern "C" void mlwe_multiply(RNS_MLWE out, RNS_MLWE in1, RNS_MLWE in2, RNS_MLWE **ksk)
{
    VH_ENTER;
    multiply_batch({out}, {in1}, {in2}, {ksk}, {0});
}

extern "C" void vfhe_hip_mlwe_multiply_batch(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                             RNS_MLWE **rlk, uint64_t divide_mask)
{
    VH_ENTER;
    if (count == 0) return;
    multiply_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in1, in1 + count),
                   std::vector<RNS_MLWE>(in2, in2 + count), std::vector<RNS_MLWE **>(count, rlk),
                   std::vector<uint64_t>(count, divide_mask));
}

extern "C" void vfhe_hip_mlwe_multiply_batch_multi(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                                   RNS_MLWE ***rlks, const uint64_t *divide_masks)
{
    VH_ENTER;
    if (count == 0) return;
    multiply_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in1, in1 + count),
                   std::vector<RNS_MLWE>(in2, in2 + count), std::vector<RNS_MLWE **>(rlks, rlks + count),
                   std::vector<uint64_t>(divide_masks, divide_masks + count));
}

// ---------------------------------------------------------------------------
// Graph-captured multiply batch.  The launch sequence of multiply_batch depends
// only on the batch's shape: the ring, the count, the rank, every operand's and
// key entry's limb mask, the rows allocated, the divide mask, the lane and the
// schedule options.  The first call of a shape runs eagerly (LOG) and captures
// the sequence (RECORD) into a plan; later calls rebuild the work lists for
// their operands with launches skipped (DRY) and replay the plan's graph, whose
// first node copies the work lists from the plan's pinned buffer to the device.
// Temporaries (ModUp rows, tensor slots) are the plan's own device blocks, so a
// replay allocates nothing.  Replays of one plan run on one stream in order; the
// host waits for the previous replay of the plan before rewriting its pinned
// work lists.
// ---------------------------------------------------------------------------
static void plan_free(vh_graph_plan *pl)
{
    if (pl->done) { VH_CHECK(hipEventSynchronize(pl->done)); VH_CHECK(hipEventDestroy(pl->done)); }
    if (pl->exec) VH_CHECK(hipGraphExecDestroy(pl->exec));
    if (pl->graph) VH_CHECK(hipGraphDestroy(pl->graph));
    for (void *b : pl->blocks) if (b) VH_CHECK(hipFree(b));
    if (pl->d_wl) VH_CHECK(hipFree(pl->d_wl));
    if (pl->h_wl) VH_CHECK(hipHostFree(pl->h_wl));
    delete pl;
}

static std::vector<uint64_t> multiply_key(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &in1,
                                          const std::vector<RNS_MLWE> &in2, RNS_MLWE **rlk, uint64_t divide_mask)
{
    std::vector<uint64_t> k = {(uint64_t)(uintptr_t)in1[0]->b->ntt, outs.size(), in1[0]->r, divide_mask,
                               (uint64_t)g_lane, (uint64_t)OPT.ntt_mode, (uint64_t)OPT.ntt_chunk_mb,
                               (uint64_t)OPT.ntt_tile_grid, (uint64_t)OPT.ntt_col_wide, (uint64_t)OPT.ks_fused,
                               (uint64_t)OPT.xcd_lag, (uint64_t)OPT.xcd_occ, (uint64_t)OPT.xcd_l2,
                               (uint64_t)OPT.mdr_fp64, (uint64_t)OPT.ks_fold, (uint64_t)OPT.ntt_small};
    auto add = [&](RNS_MLWE c) {
        k.push_back(c->r);
        for (auto p : comps(c)) { k.push_back(p->rns_mask); k.push_back(p->allocated_l); }
    };
    for (size_t i = 0; i < outs.size(); i++) { add(outs[i]); add(in1[i]); add(in2[i]); }
    const uint64_t R = mlwe_extended_rank(in1[0]->r);
    for (uint64_t i = 0; i < R; i++) {
        if (!rlk[i]) { k.push_back(~0ULL); continue; }
        // digit entries the batch reads: one per active limb of the inputs
        const uint64_t nd = (uint64_t)__builtin_popcountll(in1[0]->b->rns_mask);
        for (uint64_t d = 0; d < nd; d++) add(rlk[i][d]);
    }
    return k;
}

// Runs body() through the graph plan of `key` (capturing it on the first call of the
// key, replaying it afterwards).  get_state / set_state read and restore the host-side
// state body() changes (output masks), so the eager run, the capture and every replay
// start from the caller's state.  Inputs must be device-resident before the call.
static void plan_run(const std::vector<uint64_t> &key, const std::function<void()> &body,
                     const std::function<std::vector<uint64_t>()> &get_state,
                     const std::function<void(const std::vector<uint64_t> &)> &set_state)
{
    vh_graph_plan *pl = nullptr;
    for (auto q : g_plans) if (q->key == key) { pl = q; break; }
    const std::vector<uint64_t> state0 = get_state();
    if (pl) {
        // Replay: rebuild the work lists, then launch the graph.
        if (pl->launched) VH_CHECK(hipEventSynchronize(pl->done));
        g_cap = VH_CAP_DRY; g_plan = pl; g_cap_stage = g_cap_off = g_cap_block = 0; g_cap_launches = 0;
        g_vh_suppress_launch = 1;
        try {
            body();
            if (g_cap_launches != pl->launches) g_graph_why = 4000000L + std::min(999999L, std::labs(g_cap_launches - pl->launches));
            else if (g_cap_stage != pl->stage_sz.size()) g_graph_why = 5000000L + (long)g_cap_stage;
            else if (g_cap_block != pl->block_sz.size()) g_graph_why = 6000000L + (long)g_cap_block;
            if (g_cap_stage != pl->stage_sz.size() || g_cap_block != pl->block_sz.size() ||
                g_cap_launches != pl->launches)
                throw vh_plan_mismatch();
        } catch (const vh_plan_mismatch &) {
            g_vh_suppress_launch = 0;
            g_cap = VH_CAP_OFF; g_plan = nullptr;
            g_graph_fallbacks++;
            set_state(state0);
            body();
            return;
        }
        g_vh_suppress_launch = 0;
        g_cap = VH_CAP_OFF; g_plan = nullptr;
        VH_CHECK(hipGraphLaunch(pl->exec, S()));
        VH_CHECK(hipEventRecord(pl->done, S()));
        pl->launched = true;
        g_graph_replays++;
        return;
    }
    // First call of this shape: run eagerly, logging sizes, then capture.
    pl = new vh_graph_plan();
    pl->key = key;
    g_cap = VH_CAP_LOG; g_plan = pl;
    body();
    g_cap = VH_CAP_OFF; g_plan = nullptr;
    const std::vector<uint64_t> state1 = get_state();
    for (size_t b : pl->stage_sz) pl->wl_bytes += b;
    if (pl->wl_bytes) {
        VH_CHECK(hipHostMalloc((void **)&pl->h_wl, pl->wl_bytes, hipHostMallocDefault));
        VH_CHECK(hipMalloc((void **)&pl->d_wl, pl->wl_bytes));
    }
    for (size_t b : pl->block_sz) {
        void *p = nullptr;
        VH_CHECK(hipMalloc(&p, b));
        pl->blocks.push_back(p);
    }
    VH_CHECK(hipEventCreateWithFlags(&pl->done, hipEventDisableTiming));
    set_state(state0);
    VH_CHECK(hipStreamBeginCapture(S(), hipStreamCaptureModeThreadLocal));
    g_cap = VH_CAP_RECORD; g_plan = pl; g_cap_stage = g_cap_off = g_cap_block = 0; g_cap_launches = 0;
    bool ok = true;
    try {
        if (pl->wl_bytes)
            VH_CHECK(hipMemcpyAsync(pl->d_wl, pl->h_wl, pl->wl_bytes, hipMemcpyHostToDevice, S()));
        body();
        ok = g_cap_stage == pl->stage_sz.size() && g_cap_block == pl->block_sz.size();
    } catch (const vh_plan_mismatch &) {
        ok = false;
    }
    pl->launches = g_cap_launches;
    g_cap = VH_CAP_OFF; g_plan = nullptr;
    hipGraph_t graph = nullptr;
    VH_CHECK(hipStreamEndCapture(S(), &graph));
    pl->graph = graph;
    set_state(state1);
    if (!ok) {
        // The host logic did not repeat itself: keep the eager result, drop the plan.
        g_graph_fallbacks++;
        plan_free(pl);
        return;
    }
    VH_CHECK(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
    g_plans.push_back(pl);
    g_graph_captures++;
}

static void multiply_batch_eager_or_plan(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &in1,
                                         const std::vector<RNS_MLWE> &in2, RNS_MLWE **rlk, uint64_t divide_mask)
{
    const size_t B = outs.size();
    const std::vector<RNS_MLWE **> ks(B, rlk);
    const std::vector<uint64_t> dm(B, divide_mask);
    // Uploads and lazily built tables happen outside capture.
    for (size_t i = 0; i < B; i++) {
        for (auto p : comps(outs[i])) dev_ready(p);
        for (auto p : comps(in1[i])) dev_ready(p);
        for (auto p : comps(in2[i])) dev_ready(p);
    }
    const uint64_t R = mlwe_extended_rank(in1[0]->r);
    const uint64_t nd = (uint64_t)__builtin_popcountll(in1[0]->b->rns_mask);
    for (uint64_t i = 0; i < R; i++)
        if (rlk[i]) for (uint64_t d = 0; d < nd; d++) for (auto p : comps(rlk[i][d])) dev_ready(p);
    plan_run(multiply_key(outs, in1, in2, rlk, divide_mask), [&] { multiply_batch(outs, in1, in2, ks, dm); },
             [&] {
                 std::vector<uint64_t> m;
                 for (auto o : outs) for (auto p : comps(o)) m.push_back(p->rns_mask);
                 return m;
             },
             [&](const std::vector<uint64_t> &m) {
                 size_t k = 0;
                 for (auto o : outs) for (auto p : comps(o)) p->rns_mask = m[k++];
             });
}

extern "C" void vfhe_hip_mlwe_multiply_batch_graph(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                                   RNS_MLWE **rlk, uint64_t divide_mask)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(out && in1 && in2 && rlk);
    multiply_batch_eager_or_plan(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in1, in1 + count),
                                 std::vector<RNS_MLWE>(in2, in2 + count), rlk, divide_mask);
}

extern "C" void vfhe_hip_graph_clear(void)
{
    VH_ENTER;
    for (auto pl : g_plans) plan_free(pl);
    g_plans.clear();
}

// ---------------------------------------------------------------------------
// MGSW (mlwe/c/src/mgsw.c) and monomial products (rns_polynomial.c:1048-1168,
// mlwe.c:493-509)
// ---------------------------------------------------------------------------
static void mul_xai_rows(RNS_Polynomial out, RNS_Polynomial in, uint64_t a, bool minus1)
{
    VH_REQUIRE(out != in && out->ntt->split_degree == 1 && out->ntt == in->ntt);
    const uint64_t N = out->ntt->N;
    a &= (N << 1) - 1;
    dev_ready(in); dev_ready(out);
    if (a == 0 && !minus1) { copy_rows(out, in, in->rns_mask); out->rns_mask = in->rns_mask; return; }
    out->rns_mask = in->rns_mask;
    std::vector<vh_item> items;
    if (a == 0) {   // (X^0 - 1) * in = 0 on every row (rns_polynomial.c:1114-1121)
        for (uint64_t j = 0; j < std::min(out->ntt->l, out->allocated_l); j++)
            items.push_back(mk_item(nullptr, drow(out, j), (uint32_t)j));
        run_eltwise(items, DL(out), VH_OP_ZERO, N);
        return;
    }
    for_active(out->rns_mask, out->ntt->l, [&](uint64_t j) {
        items.push_back(mk_item(drow(in, j), drow(out, j), (uint32_t)j, nullptr, a, 0, minus1 ? 1u : 0u));
    });
    if (items.empty()) return;
    vh_prof_scope ps_(VH_PROF_OTHER);
    VH_LAUNCH(vh_launch_mul_xai(S(), vh_stage_vec(items), (unsigned)items.size(), DL(out), N));
}

extern "C" void polynomial_RNSc_mul_by_xai(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t a)
{
    VH_ENTER;
    mul_xai_rows(out, in, a, false);
}

extern "C" void polynomial_RNSc_mul_by_xai_minus1(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t a)
{
    VH_ENTER;
    mul_xai_rows(out, in, a, true);
}

extern "C" void mlwe_RNSc_mul_by_xai(RNSc_MLWE out, RNSc_MLWE in, uint64_t a)
{
    VH_ENTER;
    for (uint64_t i = 0; i < out->r; i++) mul_xai_rows(out->a[i], in->a[i], a, false);
    mul_xai_rows(out->b, in->b, a, false);
}

extern "C" void mlwe_RNSc_mul_by_xai_minus1(RNSc_MLWE out, RNSc_MLWE in, uint64_t a)
{
    VH_ENTER;
    for (uint64_t i = 0; i < out->r; i++) mul_xai_rows(out->a[i], in->a[i], a, true);
    mul_xai_rows(out->b, in->b, a, true);
}

// mgsw.c:6-49: out = sum_j G^-1(in_j) * mgsw[j*ell ..] (NTT form, extended mask),
// then, with special primes, round division by the top `special_primes` active
// limbs (INTT, divide, NTT).
static void external_product_batch(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &ins,
                                   const std::vector<RNS_MLWE *> &mgsws, uint64_t ell, uint64_t special_primes)
{
    const size_t B = outs.size();
    VH_REQUIRE(B > 0 && ins.size() == B && mgsws.size() == B);
    std::vector<std::vector<vh_gsrc>> srcs(B);
    for (size_t bi = 0; bi < B; bi++) {
        const uint64_t r = ins[bi]->r;
        VH_REQUIRE(outs[bi] != ins[bi] && outs[bi]->r == r);
        for (uint64_t j = 0; j < r; j++) srcs[bi].push_back({ins[bi]->a[j], mgsws[bi] + j * ell});
        srcs[bi].push_back({ins[bi]->b, mgsws[bi] + r * ell});
    }
    gadget_batch(outs, srcs, true);
    if (special_primes == 0) return;
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    std::map<uint64_t, std::vector<RNS_Polynomial>> by_divide;
    for (size_t bi = 0; bi < B; bi++) {
        uint64_t dm = 0, cnt = 0;
        const uint64_t m = outs[bi]->b->rns_mask;
        for (int idx = 63; idx >= 0 && cnt < special_primes; idx--)
            if (ACTIVE(m, (uint64_t)idx)) { dm |= 1ULL << idx; cnt++; }
        for (auto p : comps(outs[bi])) {
            pairs.push_back({p, p});
            if (dm) by_divide[dm].push_back(p);
        }
    }
    ntt_rows(pairs, 1);
    for (auto &g : by_divide) run_division(g.second, g.first, true);
    ntt_rows(pairs, 0);
}

extern "C" void mgsw_external_product(RNS_MLWE out, RNS_MLWE *mgsw, RNSc_MLWE in, uint64_t ell, uint64_t special_primes)
{
    VH_ENTER;
    external_product_batch({out}, {in}, {mgsw}, ell, special_primes);
}

// mgsw.c:51-71: out = ExtProd(mgsw, in2 - in1) + NTT(in1).
static void cmux_batch(const std::vector<RNS_MLWE> &outs, const std::vector<RNS_MLWE> &in1,
                       const std::vector<RNS_MLWE> &in2, const std::vector<RNS_MLWE *> &mgsws, uint64_t ell,
                       uint64_t special_primes)
{
    const size_t B = outs.size();
    VH_REQUIRE(B > 0 && in1.size() == B && in2.size() == B && mgsws.size() == B);
    incNTT t = in1[0]->b->ntt;
    const uint64_t N = t->N;
    std::vector<RNS_MLWE> diff(B), in1n(B);
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> fw;
    std::vector<vh_item> subs, copies, adds;   // one launch each over the whole batch
    for (size_t bi = 0; bi < B; bi++) {
        VH_REQUIRE(in1[bi]->b->ntt == t && in2[bi]->b->ntt == t && in2[bi]->r == in1[bi]->r);
        const uint64_t r = in1[bi]->r, mask = in1[bi]->b->rns_mask;
        diff[bi] = alloc_sample_raw(N, r, mask, t);
        in1n[bi] = alloc_sample_raw(N, r, mask, t);
        const auto d = comps(diff[bi]), x = comps(in2[bi]), y = comps(in1[bi]), c = comps(in1n[bi]);
        for (size_t k = 0; k < d.size(); k++) {
            dev_ready(x[k]); dev_ready(y[k]);
            d[k]->rns_mask = x[k]->rns_mask & y[k]->rns_mask;   // diff = in2 - in1
            for_active(d[k]->rns_mask, t->l, [&](uint64_t i) { subs.push_back(mk_item(drow(x[k], i), drow(d[k], i), (uint32_t)i, drow(y[k], i))); });
            c[k]->rns_mask = y[k]->rns_mask;                    // in1n = in1, to be transformed
            for_active(c[k]->rns_mask, t->l, [&](uint64_t i) { copies.push_back(mk_item(drow(y[k], i), drow(c[k], i), (uint32_t)i)); });
            fw.push_back({c[k], c[k]});
        }
    }
    run_eltwise(subs, DL(diff[0]->b), VH_OP_SUB, N);
    run_eltwise(copies, DL(diff[0]->b), VH_OP_COPY, N);
    external_product_batch(outs, diff, mgsws, ell, special_primes);
    ntt_rows(fw, 0);
    for (size_t bi = 0; bi < B; bi++) {   // out += NTT(in1)
        const auto o = comps(outs[bi]), c = comps(in1n[bi]);
        for (size_t k = 0; k < o.size(); k++) {
            o[k]->rns_mask &= c[k]->rns_mask;
            for_active(o[k]->rns_mask, t->l, [&](uint64_t i) { adds.push_back(mk_item(drow(o[k], i), drow(o[k], i), (uint32_t)i, drow(c[k], i))); });
        }
    }
    run_eltwise(adds, DL(diff[0]->b), VH_OP_ADD, N);
    for (size_t bi = 0; bi < B; bi++) { free_mlwe_RNS_sample(diff[bi]); free_mlwe_RNS_sample(in1n[bi]); }
}

extern "C" void mgsw_CMUX(RNS_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2, RNS_MLWE *mgsw, uint64_t ell,
                          uint64_t special_primes)
{
    VH_ENTER;
    cmux_batch({out}, {in1}, {in2}, {mgsw}, ell, special_primes);
}

// mgsw.c:73-88: CMUX(in1, automorphism_{2N-1}(in2)).
extern "C" void mgsw_NCMUX(RNS_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2, RNS_MLWE *mgsw, RNS_MLWE **ksk, uint64_t ell,
                           uint64_t special_primes)
{
    VH_ENTER;
    const uint64_t N = in1->b->ntt->N;
    RNSc_MLWE tmp = alloc_sample_raw(N, in1->r, in1->b->rns_mask, in1->b->ntt);
    mlwe_automorphism_RNSc_GHS(tmp, in2, 2 * N - 1, ksk, ell);
    cmux_batch({out}, {in1}, {tmp}, {mgsw}, ell, special_primes);
    free_mlwe_RNS_sample(tmp);
}

// CMUX for many (in1, in2, mgsw) triples in one launch sequence: the blind
// rotation of a batch of LWE ciphertexts (cggi16.py:70-90) steps all of them
// through the same bootstrapping-key index at once.
extern "C" void vfhe_hip_mgsw_cmux_batch(RNS_MLWE *out, RNSc_MLWE *in1, RNSc_MLWE *in2, RNS_MLWE **mgsw, uint64_t count,
                                         uint64_t ell, uint64_t special_primes)
{
    VH_ENTER;
    if (count == 0) return;
    cmux_batch(std::vector<RNS_MLWE>(out, out + count), std::vector<RNS_MLWE>(in1, in1 + count),
               std::vector<RNS_MLWE>(in2, in2 + count), std::vector<RNS_MLWE *>(mgsw, mgsw + count), ell,
               special_primes);
}

// Blind rotation of a batch (cggi16.py:46-90 for every accumulator at once).  acc[k]
// enters as X^-rot_b * tv in coefficient form; for every bootstrapping-key index i the
// reference sets acc = CMUX(acc, X^a * acc, bk[i]) (mgsw.c:51-71: ExtProd(bk[i],
// X^a acc - acc) + NTT(acc), NTT form) and brings it back to coefficient form for the
// next index.  Per index this runs, over every accumulator at once:
//   diff = (X^a - 1) * acc                      one monomial-product launch (per-row a)
//   lift + NTT of diff's digits                 one NTT launch (the lift folded in on FP64 limbs)
//   out = sum_d NTT(digit_d) * bk[i][d]         one inner-product launch (the external product)
//   INTT of out, round division by the special limbs (mgsw.c:6-49)
//   acc = acc + out                             (INTT(NTT(div) + NTT(acc)) = div + acc exactly)
// An accumulator with a = 0 gets diff = 0, hence out = 0 and acc unchanged, which is the
// reference's skip.  Temporaries are allocated once per call; the work lists that do not
// depend on i are built once.  The results are left in acc (coefficient form).
// tv (optional): accumulator k starts as X^rot_tv[k] * tv (mlwe_RNSc_mul_by_xai, coefficient
// form) instead of holding its start already.
static void blind_rotate_core(RNS_MLWE *acc, uint64_t count, RNS_MLWE tv, const uint64_t *rot_tv, RNS_MLWE **bk,
                              uint64_t n, const uint64_t *steps, uint64_t ell, uint64_t special_primes)
{
    if (count == 0) return;
    VH_REQUIRE(acc && bk && steps && bk[0] && bk[0][0] && (!tv || rot_tv));
    incNTT t = acc[0]->b->ntt;
    const uint64_t N = t->N, r = acc[0]->r, Cc = r + 1, l = t->l;
    VH_REQUIRE(t->split_degree == 1 && Cc <= VH_MAX_COMPS && count <= VH_MAX_GRID_Y);
    if (tv) {
        VH_REQUIRE(tv->r == r && tv->b->ntt == t);
        const auto tc = comps(tv);
        for (auto p : tc) dev_ready(p);
        for (uint64_t k = 0; k < count; k++) {
            const auto ac = comps(acc[k]);
            for (uint64_t c = 0; c < Cc; c++) {
                VH_REQUIRE(ac[c] != tc[c]);
                ac[c]->rns_mask = tc[c]->rns_mask;
                check_rows(ac[c], tc[c]->rns_mask);
            }
        }
    }
    if (n == 0 && !tv) return;
    const uint64_t m = acc[0]->b->rns_mask, ext = bk[0][0]->b->rns_mask;
    for (uint64_t k = 0; k < count; k++) {
        VH_REQUIRE(acc[k]->b->ntt == t && acc[k]->r == r);
        for (auto p : comps(acc[k])) { VH_REQUIRE(p->rns_mask == m); dev_ready(p); }
    }
    // digits: the level limbs of every component; extended limbs: the key's; the
    // division removes the top `special_primes` extended limbs (external_product_batch)
    std::vector<int> J, E, divs, keep;
    for (uint64_t i = 0; i < l; i++) { if (ACTIVE(m, i)) J.push_back((int)i); if (ACTIVE(ext, i)) E.push_back((int)i); }
    uint64_t dm = 0, cnt = 0;
    for (int idx = 63; idx >= 0 && cnt < special_primes; idx--)
        if (ACTIVE(ext, (uint64_t)idx)) { dm |= 1ULL << idx; cnt++; }
    for (int e : E) (ACTIVE(dm, (uint64_t)e) ? divs : keep).push_back(e);
    VH_REQUIRE((m & ~ext) == 0 && (ext & ~dm) == m && J.size() <= ell && divs.size() <= VH_MAX_DIV &&
               E.size() <= VH_MAX_LIMBS);
    vh_pool *P = VPOOL(t);
    const uint64_t nd = Cc * J.size(), next = E.size();
    bool fp = true;   // FP64 limbs only: the lift folds into the NTT and the inner product runs in FP64
    for (int e : E) fp = fp && P->h_limbs[e].fp64;
    for (int j : J) fp = fp && P->h_limbs[j].fp64;
    if (fp && OPT.br_fused && N == 1024 && Cc <= 2 && divs.size() == 1 && J.size() <= VH_BR_MAXJ &&
        next == J.size() + 1) {
        // Every key index in one launch, each accumulator resident in its workgroup's LDS.
        vh_br_params bp;
        memset(&bp, 0, sizeof(bp));
        bp.C = (int)Cc;
        bp.nJ = (int)J.size();
        bp.next = (int)next;
        bp.n = n;
        const uint64_t p = t->ntt[divs[0]]->q;
        bp.hp = (double)(p / 2);
        for (size_t jj = 0; jj < J.size(); jj++) bp.J[jj] = J[jj];
        for (size_t w = 0; w < next; w++) {
            bp.E[w] = E[w];
            bp.jw[w] = -1;
            for (size_t jj = 0; jj < J.size(); jj++) if (J[jj] == E[w]) bp.jw[w] = (int)jj;
            if (E[w] == divs[0]) { bp.sp = (int)w; continue; }
            const uint64_t qe = t->ntt[E[w]]->q, iv = inverse_mod(p % qe, qe);
            bp.hdiv[w] = (double)((p / 2) % qe);
            bp.inv[w] = centred(iv, qe);
            bp.invq[w] = bp.inv[w] / (double)qe;
        }
        std::vector<uint64_t *> accp(count * Cc * J.size());
        for (uint64_t k = 0; k < count; k++) {
            const auto ac = comps(acc[k]);
            for (uint64_t c = 0; c < Cc; c++)
                for (size_t jj = 0; jj < J.size(); jj++) accp[(k * Cc + c) * J.size() + jj] = drow(ac[c], J[jj]);
        }
        std::vector<const uint64_t *> bkp(n * nd * Cc * next);
        for (uint64_t i = 0; i < n; i++) {
            VH_REQUIRE(bk[i]);
            for (uint64_t d = 0; d < nd; d++) {
                const uint64_t src = d / J.size(), jj = d % J.size();
                RNS_MLWE e = bk[i][src * ell + jj];
                VH_REQUIRE(e && e->r == r && e->b->ntt == t && (e->b->rns_mask & ext) == ext);
                for (uint64_t c = 0; c < Cc; c++) {
                    RNS_Polynomial kp = c < r ? e->a[c] : e->b;
                    dev_ready(kp);
                    for (uint64_t w = 0; w < next; w++) bkp[((i * nd + d) * Cc + c) * next + w] = drow(kp, E[w]);
                }
            }
        }
        std::vector<uint64_t> amounts(std::max<uint64_t>(1, count * n));
        for (uint64_t k = 0; k < count * n; k++) amounts[k] = steps[k] & ((N << 1) - 1);
        // The tables go through the lane's pinned staging ring (copied in stream order, the
        // ring's reuse ordered by its events), so the call returns once the launch is queued;
        // tables too large for the ring are copied into blocks of their own and the call waits
        // for the launch before their host copies go out of scope.
        const size_t table_bytes = (accp.size() + bkp.size()) * sizeof(void *) + amounts.size() * 8 +
                                   (tv ? (Cc * J.size() + count) * 8 : 0);
        const bool staged = table_bytes + 4096 <= LN().ring_size / 4;
        std::vector<void *> held;
        auto put = [&](const void *src, size_t bytes) -> void * {
            if (staged) return vh_stage(src, bytes);
            void *d = vh_dmalloc(bytes ? bytes : 8);
            if (bytes) VH_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, S()));
            held.push_back(d);
            return d;
        };
        auto *d_acc = (uint64_t *const *)put(accp.data(), accp.size() * sizeof(uint64_t *));
        auto *d_bk = (const uint64_t *const *)put(bkp.data(), bkp.size() * sizeof(uint64_t *));
        auto *d_am = (const uint64_t *)put(amounts.data(), amounts.size() * 8);
        const uint64_t *const *d_tv = nullptr;
        const uint64_t *d_rot = nullptr;
        std::vector<const uint64_t *> tvp;
        std::vector<uint64_t> rot;
        if (tv) {
            const auto tc = comps(tv);
            for (uint64_t c = 0; c < Cc; c++)
                for (size_t jj = 0; jj < J.size(); jj++) tvp.push_back(drow(tc[c], J[jj]));
            for (uint64_t k = 0; k < count; k++) rot.push_back(rot_tv[k] & ((N << 1) - 1));
            d_tv = (const uint64_t *const *)put(tvp.data(), tvp.size() * sizeof(uint64_t *));
            d_rot = (const uint64_t *)put(rot.data(), rot.size() * 8);
        }
        {
            vh_prof_scope ps_(VH_PROF_KSMAC);
            VH_LAUNCH(vh_launch_blind_rotate(S(), bp, P->d_limbs, d_acc, d_bk, d_am, (unsigned)count, d_tv, d_rot));
        }
        g_br_fused_launches++;
        if (!held.empty()) {
            VH_CHECK(hipStreamSynchronize(S()));   // the host tables go out of scope
            for (void *d : held) vh_dfree(d);
        }
        return;
    }
    if (tv) {   // acc[k] = X^rot_tv[k] * tv, one launch (the exponent table holds rot_tv)
        std::vector<uint64_t> rot(count);
        for (uint64_t k = 0; k < count; k++) rot[k] = rot_tv[k] & ((N << 1) - 1);
        const auto tc = comps(tv);
        std::vector<vh_item> items;
        for (uint64_t k = 0; k < count; k++) {
            const auto ac = comps(acc[k]);
            for (uint64_t c = 0; c < Cc; c++)
                for_active(tc[c]->rns_mask, l, [&](uint64_t j) {
                    items.push_back(mk_item(drow(tc[c], j), drow(ac[c], j), (uint32_t)j, nullptr, k, 0, 0u));
                });
        }
        const uint64_t *d_rot = vh_stage_vec(rot);
        for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
            const unsigned cn = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
            std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cn);
            vh_prof_scope ps_(VH_PROF_OTHER);
            VH_LAUNCH(vh_launch_mul_xai(S(), vh_stage_vec(chunk), cn, P->d_limbs, N, d_rot, 1, 0));
        }
        if (n == 0) return;
    }
    // temporaries
    std::vector<RNS_MLWE> diff(count), out(count);
    for (uint64_t k = 0; k < count; k++) {
        diff[k] = alloc_sample_raw(N, r, m, t);
        out[k] = alloc_sample_raw(N, r, ext, t);
    }
    uint64_t *modup = (uint64_t *)vh_dmalloc(count * nd * next * N * 8);
    // The work lists do not depend on the key index: built and copied to the device once.
    std::vector<vh_item> rot, lift, nt, inv, zero, adds;
    std::vector<vh_mac_job> jobs(count);
    std::vector<vh_div_job> dj;
    for (uint64_t k = 0; k < count; k++) {
        const auto dc = comps(diff[k]), oc = comps(out[k]), ac = comps(acc[k]);
        for (uint64_t c = 0; c < Cc; c++)
            for (size_t jj = 0; jj < J.size(); jj++) {
                // diff = (X^a - 1) * acc, a = the accumulator's entry of the exponent table
                rot.push_back(mk_item(drow(ac[c], J[jj]), drow(dc[c], J[jj]), (uint32_t)J[jj], nullptr, k, 0, 1u));
                const uint64_t d = c * J.size() + jj;
                for (uint64_t ti = 0; ti < next; ti++) {
                    uint64_t *dst = modup + ((k * nd + d) * next + ti) * N;
                    lift.push_back(mk_item(drow(dc[c], J[jj]), dst, (uint32_t)E[ti]));
                    nt.push_back(mk_item(fp ? drow(dc[c], J[jj]) : dst, dst, (uint32_t)E[ti]));
                }
            }
        vh_mac_job &jb = jobs[k];
        memset(&jb, 0, sizeof(jb));
        for (uint64_t c = 0; c < Cc; c++) jb.out[c] = VP(oc[c])->d;
        jb.modup = modup + k * nd * next * N;
        jb.kptr_off = 0;   // every accumulator reads the same key rows
        jb.comps = (int)Cc;
        jb.add = 1;
        jb.digits = (int)nd;
        jb.next = (int)next;
        for (uint64_t ti = 0; ti < next; ti++) jb.limb[ti] = E[ti];
        for (uint64_t c = 0; c < Cc; c++) {
            for (int e : E) inv.push_back(mk_item(drow(oc[c], e), drow(oc[c], e), (uint32_t)e));
            for (int d : divs) zero.push_back(mk_item(nullptr, drow(oc[c], d), (uint32_t)d));
            for (int j : J) adds.push_back(mk_item(drow(ac[c], j), drow(ac[c], j), (uint32_t)j, drow(oc[c], j)));
            vh_div_job x;
            memset(&x, 0, sizeof(x));
            x.base = VP(oc[c])->d;
            x.stride = N;
            x.K = (int)divs.size();
            x.nl = (int)l;
            for (size_t q = 0; q < divs.size(); q++) x.div_limb[q] = divs[q];
            for (size_t q = 0; q < keep.size(); q++) x.out_limb[q] = keep[q];
            x.nout = (int)keep.size();
            dj.push_back(x);
        }
    }
    auto by_limb = [](const vh_item &a, const vh_item &b) { return a.limb < b.limb; };
    std::stable_sort(nt.begin(), nt.end(), by_limb);    // workgroups in flight share a limb's twiddles
    std::stable_sort(inv.begin(), inv.end(), by_limb);
    std::vector<uint64_t> amounts(count * n);
    for (uint64_t k = 0; k < count * n; k++) amounts[k] = steps[k] & ((N << 1) - 1);
    std::vector<void *> held;
    auto persist = [&](const void *src, size_t bytes) -> void * {
        void *d = vh_dmalloc(bytes ? bytes : 8);
        if (bytes) VH_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, S()));
        held.push_back(d);
        return d;
    };
    const vh_item *d_rot = (const vh_item *)persist(rot.data(), rot.size() * sizeof(vh_item));
    const vh_item *d_lift = fp ? nullptr : (const vh_item *)persist(lift.data(), lift.size() * sizeof(vh_item));
    const vh_item *d_nt = (const vh_item *)persist(nt.data(), nt.size() * sizeof(vh_item));
    const vh_item *d_inv = (const vh_item *)persist(inv.data(), inv.size() * sizeof(vh_item));
    const vh_item *d_zero = (const vh_item *)persist(zero.data(), zero.size() * sizeof(vh_item));
    const vh_item *d_adds = (const vh_item *)persist(adds.data(), adds.size() * sizeof(vh_item));
    const vh_mac_job *d_jobs = (const vh_mac_job *)persist(jobs.data(), jobs.size() * sizeof(vh_mac_job));
    const vh_div_job *d_dj = (const vh_div_job *)persist(dj.data(), dj.size() * sizeof(vh_div_job));
    const uint64_t *d_amounts = (const uint64_t *)persist(amounts.data(), amounts.size() * 8);
    VH_CHECK(hipStreamSynchronize(S()));   // the host lists may go once copied
    const void *d_consts = div_consts(t, ext, dm, divs, true);
    auto chunks = [&](size_t total, const std::function<void(size_t, unsigned)> &f) {
        for (size_t off = 0; off < total; off += VH_MAX_GRID_Y) f(off, (unsigned)std::min<size_t>(VH_MAX_GRID_Y, total - off));
    };
    std::vector<const uint64_t *> kptr(next * nd * Cc);
    for (uint64_t i = 0; i < n; i++) {
        VH_REQUIRE(bk[i]);
        // 1. diff = (X^a - 1) * acc (a = 0 gives 0: the reference's skipped index)
        chunks(rot.size(), [&](size_t off, unsigned cn) {
            vh_prof_scope ps_(VH_PROF_OTHER);
            VH_LAUNCH(vh_launch_mul_xai(S(), d_rot + off, cn, P->d_limbs, N, d_amounts, n, i));
        });
        // 2. digits lifted to the extended limbs, forward NTT
        if (!fp) chunks(lift.size(), [&](size_t off, unsigned cn) {
            vh_prof_scope ps_(VH_PROF_ELTWISE);
            VH_LAUNCH(vh_launch_eltwise(S(), d_lift + off, cn, P->d_limbs, VH_OP_REDUCE, N));
        });
        chunks(nt.size(), [&](size_t off, unsigned cn) { ntt_dispatch(d_nt + off, cn, P->d_limbs, fp, N, 0); });
        // 3. external product with bk[i]: row (ti, d, c) of the key table
        for (uint64_t d = 0; d < nd; d++) {
            const uint64_t src = d / J.size(), jj = d % J.size();
            RNS_MLWE e = bk[i][src * ell + jj];
            VH_REQUIRE(e && e->r == r && e->b->ntt == t);
            for (uint64_t c = 0; c < Cc; c++) {
                RNS_Polynomial kp = c < r ? e->a[c] : e->b;
                dev_ready(kp);
                for (uint64_t ti = 0; ti < next; ti++) kptr[(ti * nd + d) * Cc + c] = drow(kp, E[ti]);
            }
        }
        const uint64_t *const *d_kptr = vh_stage_vec(kptr);
        chunks(jobs.size(), [&](size_t off, unsigned cn) {
            vh_prof_scope ps_(VH_PROF_KSMAC);
            VH_LAUNCH(vh_launch_keyswitch_mac(S(), d_jobs + off, cn, (unsigned)next, P->d_limbs, d_kptr, N, fp ? 1 : 0));
        });
        // 4. back to coefficient form, round-divided by the special limbs (their rows cleared)
        chunks(inv.size(), [&](size_t off, unsigned cn) { ntt_dispatch(d_inv + off, cn, P->d_limbs, fp, N, 1); });
        if (!divs.empty()) {
            chunks(dj.size(), [&](size_t off, unsigned cn) {
                vh_prof_scope ps_(VH_PROF_DIVISION);
                VH_LAUNCH(vh_launch_division(S(), d_dj + off, cn, (unsigned)keep.size(), P->d_limbs, d_consts, N,
                                             (int)divs.size()));
            });
            chunks(zero.size(), [&](size_t off, unsigned cn) {
                vh_prof_scope ps_(VH_PROF_ELTWISE);
                VH_LAUNCH(vh_launch_eltwise(S(), d_zero + off, cn, P->d_limbs, VH_OP_ZERO, N));
            });
        }
        // 5. acc += out
        chunks(adds.size(), [&](size_t off, unsigned cn) {
            vh_prof_scope ps_(VH_PROF_ELTWISE);
            VH_LAUNCH(vh_launch_eltwise(S(), d_adds + off, cn, P->d_limbs, VH_OP_ADD, N));
        });
    }
    for (void *d : held) vh_dfree(d);
    vh_dfree(modup);
    for (uint64_t k = 0; k < count; k++) { free_mlwe_RNS_sample(diff[k]); free_mlwe_RNS_sample(out[k]); }
}

extern "C" void vfhe_hip_blind_rotate_batch(RNS_MLWE *acc, uint64_t count, RNS_MLWE **bk, uint64_t n,
                                            const uint64_t *steps, uint64_t ell, uint64_t special_primes)
{
    VH_ENTER;
    if (count == 0 || n == 0) return;
    blind_rotate_core(acc, count, nullptr, nullptr, bk, n, steps, ell, special_primes);
}

extern "C" void vfhe_hip_blind_rotate_tv_batch(RNS_MLWE *acc, uint64_t count, RNS_MLWE tv, const uint64_t *rot_tv,
                                               RNS_MLWE **bk, uint64_t n, const uint64_t *steps, uint64_t ell,
                                               uint64_t special_primes)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(tv && rot_tv);
    blind_rotate_core(acc, count, tv, rot_tv, bk, n, steps, ell, special_primes);
}

// mlwe_extract_LWE (mlwe.c:40-66) for `count` samples of one ring and rank: their rows are
// gathered on the device and read back with one copy.
// Pinned host staging for batched read-backs (grows, kept for the process).
static void *g_pin = nullptr;
static size_t g_pin_bytes = 0;
static void *pinned(size_t bytes)
{
    if (bytes > g_pin_bytes) {
        if (g_pin) VH_CHECK(hipHostFree(g_pin));
        g_pin_bytes = std::max<size_t>(bytes, 2 * g_pin_bytes);
        VH_CHECK(hipHostMalloc(&g_pin, g_pin_bytes, hipHostMallocDefault));
    }
    return g_pin;
}

extern "C" void vfhe_hip_extract_lwe_batch(RNSc_MLWE *in, uint64_t count, uint64_t idx, LWE *out)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(in && out && in[0]);
    incNTT t = in[0]->a[0]->ntt;
    const uint64_t N = t->N, r = in[0]->r, mask = in[0]->a[0]->rns_mask, Cc = r + 1;
    VH_REQUIRE(idx < N && t->split_degree == 1);
    std::vector<int> act;
    for (uint64_t i = 0; i < t->l; i++) if (ACTIVE(mask, i)) act.push_back((int)i);
    const uint64_t na = act.size(), width = r * N + 1;
    // the device forms every sample's rows (mlwe.c:40-66's layout: a'_i = a_(idx-i), then
    // -a_(N+idx-i) mod q, b' = b_idx) into one buffer, read back with one copy
    std::vector<const uint64_t *> src(count * na * Cc);
    std::vector<uint64_t> qs(na);
    for (uint64_t j = 0; j < na; j++) qs[j] = t->ntt[act[j]]->q;
    for (uint64_t k = 0; k < count; k++) {
        VH_REQUIRE(in[k] && in[k]->r == r && in[k]->a[0]->ntt == t);
        const auto ic = comps(in[k]);
        for (auto p : ic) { VH_REQUIRE(p->rns_mask == mask); dev_ready(p); }
        for (uint64_t j = 0; j < na; j++)
            for (uint64_t c = 0; c < Cc; c++) src[(k * na + j) * Cc + c] = drow(ic[c], act[j]);
    }
    const size_t bytes = count * na * width * 8;
    uint64_t *buf = (uint64_t *)vh_dmalloc(bytes);
    VH_LAUNCH(vh_launch_extract_lwe(S(), vh_stage_vec(src), vh_stage_vec(qs), buf, (uint32_t)N, (uint32_t)r,
                                    (uint32_t)na, (uint32_t)idx, count));
    uint64_t *host = (uint64_t *)pinned(bytes);
    VH_CHECK(hipMemcpyAsync(host, buf, bytes, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(buf);
    for (uint64_t k = 0; k < count; k++) {
        LWE res = lwe_alloc_sample(r * N, na, t);
        for (uint64_t j = 0; j < na; j++) {
            const uint64_t *row = host + (k * na + j) * width;
            memcpy(res->a[j], row, r * N * 8);
            res->b[j] = row[r * N];
        }
        out[k] = res;
    }
}

// ---------------------------------------------------------------------------
// Ciphertext and key format.  The reference has none (keys and ciphertexts live only
// in memory); this one lets a server keep evaluation keys and ciphertexts across runs.
// Little-endian words:
//   "VFHEHIP1" | u32 version = 1 | u32 form (0 unstated, 1 NTT, 2 coefficient) |
//   u64 count | u64 N | u64 rank r | u64 split degree | u64 limb mask | u64 limbs n |
//   n x u64 prime (the mask's limbs, ascending) |
//   count x (r + 1) x n x N u64 rows (sample, component a_0 .. a_{r-1}, b, limb) |
//   u64 check: h = 0x9E3779B97F4A7C15; per preceding word w: h = ((h ^ w) * 0x100000001B3) ^ (h >> 29)
// ---------------------------------------------------------------------------
static const uint64_t VH_FMT_MAGIC = 0x3150494845484656ULL;   // "VFHEHIP1"
static uint64_t vh_fmt_check(const uint64_t *w, uint64_t n)
{
    uint64_t h = 0x9E3779B97F4A7C15ULL;
    for (uint64_t i = 0; i < n; i++) h = ((h ^ w[i]) * 0x100000001B3ULL) ^ (h >> 29);
    return h;
}

// Words of a block; UINT64_MAX when that does not fit 64 bits (a header read from a file
// may hold any values).
static uint64_t vh_fmt_words(uint64_t count, uint64_t r, uint64_t nact, uint64_t N)
{
    const u128 w = (u128)8 + nact + (u128)count * (r + 1) * nact * N + 1;
    return w >> 61 ? UINT64_MAX : (uint64_t)w;   // also keeps 8 * words below 2^64
}

extern "C" uint64_t vfhe_hip_mlwe_serialized_size(RNS_MLWE *in, uint64_t count)
{
    VH_ENTER;
    if (count == 0 || !in || !in[0]) return 0;
    const uint64_t nact = (uint64_t)__builtin_popcountll(in[0]->b->rns_mask);
    const uint64_t words = vh_fmt_words(count, in[0]->r, nact, in[0]->b->ntt->N);
    return words == UINT64_MAX ? 0 : 8 * words;
}

extern "C" int64_t vfhe_hip_mlwe_serialize(RNS_MLWE *in, uint64_t count, uint32_t form, void *buf, uint64_t size)
{
    VH_ENTER;
    if (count == 0 || !in || !in[0] || !buf) return -1;
    incNTT t = in[0]->b->ntt;
    const uint64_t N = t->N, r = in[0]->r, mask = in[0]->b->rns_mask, Cc = r + 1;
    std::vector<int> act;
    for (uint64_t i = 0; i < t->l; i++) if (ACTIVE(mask, i)) act.push_back((int)i);
    const uint64_t nact = act.size(), words = vh_fmt_words(count, r, nact, N);
    if (words == UINT64_MAX || size / 8 < words || ((uintptr_t)buf & 7)) return -1;
    for (uint64_t k = 0; k < count; k++) {
        if (!in[k] || in[k]->r != r || in[k]->b->ntt != t) return -1;
        for (auto p : comps(in[k])) if (p->rns_mask != mask) return -1;
    }
    uint64_t *w = (uint64_t *)buf;
    w[0] = VH_FMT_MAGIC;
    w[1] = 1ULL | ((uint64_t)form << 32);
    w[2] = count; w[3] = N; w[4] = r; w[5] = t->split_degree; w[6] = mask; w[7] = nact;
    for (uint64_t j = 0; j < nact; j++) w[8 + j] = t->ntt[act[j]]->q;
    const uint64_t rows = count * Cc * nact;
    uint64_t *dbuf = (uint64_t *)vh_dmalloc(rows * N * 8);
    std::vector<vh_item> items;
    items.reserve(rows);
    for (uint64_t k = 0; k < count; k++) {
        const auto ic = comps(in[k]);
        for (uint64_t c = 0; c < Cc; c++) {
            dev_ready(ic[c]);
            for (uint64_t j = 0; j < nact; j++)
                items.push_back(mk_item(drow(ic[c], act[j]), dbuf + ((k * Cc + c) * nact + j) * N, (uint32_t)act[j]));
        }
    }
    run_eltwise(items, VPOOL(t)->d_limbs, VH_OP_COPY, N);
    VH_CHECK(hipMemcpyAsync(w + 8 + nact, dbuf, rows * N * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(dbuf);
    w[words - 1] = vh_fmt_check(w, words - 1);
    return (int64_t)(8 * words);
}

extern "C" int64_t vfhe_hip_mlwe_deserialize(const void *buf, uint64_t size, incNTT t, RNS_MLWE *out,
                                             uint64_t max_count, uint32_t *form)
{
    VH_ENTER;
    if (!buf || !t || ((uintptr_t)buf & 7) || size < 8 * 9) return -4;
    const uint64_t *w = (const uint64_t *)buf;
    if (w[0] != VH_FMT_MAGIC || (w[1] & 0xFFFFFFFFULL) != 1) return -1;
    const uint64_t count = w[2], N = w[3], r = w[4], split = w[5], mask = w[6], nact = w[7];
    if (nact != (uint64_t)__builtin_popcountll(mask) || r > 64 || N == 0 || N > (1ULL << 24) || count > (1ULL << 32))
        return -1;
    const uint64_t words = vh_fmt_words(count, r, nact, N);
    if (words == UINT64_MAX || size / 8 < words) return -4;
    if (vh_fmt_check(w, words - 1) != w[words - 1]) return -1;
    if (N != t->N || split != t->split_degree || (t->l < 64 && (mask >> t->l))) return -2;
    std::vector<int> act;
    for (uint64_t i = 0; i < t->l; i++) if (ACTIVE(mask, i)) act.push_back((int)i);
    for (uint64_t j = 0; j < nact; j++) if (t->ntt[act[j]]->q != w[8 + j]) return -2;
    if (count > max_count || !out) return -3;
    if (form) *form = (uint32_t)(w[1] >> 32);
    const uint64_t Cc = r + 1, rows = count * Cc * nact;
    uint64_t *dbuf = (uint64_t *)vh_dmalloc(std::max<uint64_t>(1, rows) * N * 8);
    VH_CHECK(hipMemcpyAsync(dbuf, w + 8 + nact, rows * N * 8, hipMemcpyHostToDevice, S()));
    std::vector<vh_item> items;
    items.reserve(rows);
    for (uint64_t k = 0; k < count; k++) {
        out[k] = vfhe_hip_mlwe_alloc_RNS_sample_uninit(N, r, mask, t);
        const auto oc = comps(out[k]);
        for (uint64_t c = 0; c < Cc; c++)
            for (uint64_t j = 0; j < nact; j++)
                items.push_back(mk_item(dbuf + ((k * Cc + c) * nact + j) * N, drow(oc[c], act[j]), (uint32_t)act[j]));
    }
    run_eltwise(items, VPOOL(t)->d_limbs, VH_OP_COPY, N);
    VH_CHECK(hipStreamSynchronize(S()));   // the caller's buffer may go once this returns
    vh_dfree(dbuf);
    return (int64_t)count;
}

// ---------------------------------------------------------------------------
// Keys and phase (mlwe.c:8-38, 243-252)
// ---------------------------------------------------------------------------
extern "C" RNS_MLWE_Key mlwe_new_RNS_key_from_array(uint64_t *array, uint64_t N, uint64_t r, uint64_t l, incNTT ntt, double sigma)
{
    VH_ENTER;
    RNS_MLWE_Key k = (RNS_MLWE_Key)malloc(sizeof(*k));
    if (!k) exit(EXIT_FAILURE);
    k->N = N; k->l = l; k->r = r; k->sigma = sigma;
    k->s = (IntPolynomial *)malloc(sizeof(IntPolynomial) * (r ? r : 1));
    k->s_RNS = polynomial_new_RNS_polynomial_array(r, N, l >= 64 ? ~0ULL : ((1ULL << l) - 1), ntt);
    for (uint64_t i = 0; i < r; i++) {
        k->s[i] = (IntPolynomial)malloc(sizeof(*k->s[i]));
        k->s[i]->N = N;
        k->s[i]->coeffs = (uint64_t *)malloc(N * 8);
        memcpy(k->s[i]->coeffs, array + i * N, N * 8);
        int_array_to_RNS(k->s_RNS[i], k->s[i]->coeffs);
    }
    return k;
}

extern "C" void free_mlwe_RNS_key(RNS_MLWE_Key key)
{
    if (!key) return;
    for (uint64_t i = 0; i < key->r; i++) { free(key->s[i]->coeffs); free(key->s[i]); }
    free(key->s);
    free_RNS_polynomial_array(key->r, key->s_RNS);
    free(key);
}

extern "C" void mlwe_RNS_phase(RNS_Polynomial out, RNS_MLWE in, RNS_MLWE_Key key)
{
    polynomial_mul_RNS_polynomial(out, in->a[0], key->s_RNS[0]);
    for (uint64_t i = 1; i < in->r; i++) polynomial_mul_addto_RNS_polynomial(out, in->a[i], key->s_RNS[i]);
    polynomial_sub_RNS_polynomial(out, in->b, out);
}

// ---------------------------------------------------------------------------
// Batched transforms
// ---------------------------------------------------------------------------
extern "C" void vfhe_hip_ntt_batch(RNS_Polynomial *polys, uint64_t count, int direction)
{
    VH_ENTER;
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    for (uint64_t i = 0; i < count; i++) { dev_ready(polys[i]); pairs.push_back({polys[i], polys[i]}); }
    ntt_rows(pairs, direction ? 1 : 0);
}

static void plan_run(const std::vector<uint64_t> &key, const std::function<void()> &body,
                     const std::function<std::vector<uint64_t>()> &get_state,
                     const std::function<void(const std::vector<uint64_t> &)> &set_state);

// vfhe_hip_ntt_batch through a graph plan: the launches of a batch shape (ring, count,
// direction, every polynomial's limb mask and rows, lane, schedule options) are captured
// once and replayed with the polynomials' work lists rebuilt.
extern "C" void vfhe_hip_ntt_batch_graph(RNS_Polynomial *polys, uint64_t count, int direction)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(polys);
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    std::vector<uint64_t> key = {3, (uint64_t)(uintptr_t)polys[0]->ntt, count, (uint64_t)(direction ? 1 : 0),
                                 (uint64_t)g_lane, (uint64_t)OPT.ntt_mode, (uint64_t)OPT.ntt_chunk_mb,
                                 (uint64_t)OPT.ntt_tile_grid, (uint64_t)OPT.ntt_col_wide, (uint64_t)OPT.ntt_small,
                                 (uint64_t)OPT.xcd_lag, (uint64_t)OPT.xcd_occ, (uint64_t)OPT.xcd_l2};
    for (uint64_t i = 0; i < count; i++) {
        dev_ready(polys[i]);
        pairs.push_back({polys[i], polys[i]});
        key.push_back((uint64_t)(uintptr_t)polys[i]->ntt);
        key.push_back(polys[i]->rns_mask);
        key.push_back(polys[i]->allocated_l);
    }
    plan_run(key, [&] { ntt_rows(pairs, direction ? 1 : 0); }, [] { return std::vector<uint64_t>(); },
             [](const std::vector<uint64_t> &) {});
}

// Allocation-free form of vfhe_hip_ntt_batch: the same transforms, run by the
// persistent per-XCD kernel (one launch per direction) wherever it applies (FP64
// limbs, N = 2^13..2^16, XCD placement probe passed), with that kernel's work list
// and queue counters, and every other path's work lists, in this device's scratch
// array (vh_kernels.hip) instead of engine allocations.  The counters are cleared
// by a fill kernel before each launch.  Calls that overlap on one device (other
// lanes, other threads' streams) share the array and race.  Returns 0, or -1 with
// nothing done when the batch needs more than VH_SCRATCH_BYTES.
extern "C" int vfhe_hip_ntt_batch_noalloc(RNS_Polynomial *polys, uint64_t count, int direction)
{
    VH_ENTER;
    char *base = (char *)g_scratch_of[C.device];
    VH_REQUIRE(base != nullptr && (count == 0 || polys));
    uint64_t rows = 0;
    for (uint64_t i = 0; i < count; i++) {
        const incNTT t = polys[i]->ntt;
        const uint64_t m = polys[i]->rns_mask & (t->l >= 64 ? ~0ULL : ((1ULL << t->l) - 1));
        rows += (uint64_t)__builtin_popcountll(m) * t->split_degree;
    }
    // per row: a 48-byte work item and two 4-byte counters; per launch group (at most
    // two: FP64 and integer limbs) the 256-byte roundings and the 16 queue tickets
    const size_t need = rows * (sizeof(vh_item) + 8) + 2048;
    if (need > VH_SCRATCH_BYTES) {
        fprintf(stderr, "vfhe hip engine: vfhe_hip_ntt_batch_noalloc: %lu limb rows need %zu bytes of scratch, "
                        "more than its %zu\n", (unsigned long)rows, need, (size_t)VH_SCRATCH_BYTES);
        return -1;
    }
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    for (uint64_t i = 0; i < count; i++) { dev_ready(polys[i]); pairs.push_back({polys[i], polys[i]}); }
    const long mode = OPT.ntt_mode;
    if (C.xcd_ok) OPT.ntt_mode = 1;
    g_na_base = base;
    g_na_off = 0;
    try {
        ntt_rows(pairs, direction ? 1 : 0);
    } catch (...) {
        g_na_base = nullptr;
        OPT.ntt_mode = mode;
        throw;
    }
    g_na_base = nullptr;
    OPT.ntt_mode = mode;
    return 0;
}

// The fused product's three launches (N = 2^13..2^16, FP64 limbs): forward column
// passes of x into z and of y into its temporary row (col), forward row stages of
// both, pointwise product and inverse row stages in one kernel (mid), inverse
// column pass (inv).
static void fused_product_launches(std::vector<vh_item> &col, std::vector<vh_item> &mid, std::vector<vh_item> &inv,
                                   const vh_pool *P, int logn)
{
    const int R1 = logn - 8;
    auto launch = [&](std::vector<vh_item> &v, int which) {
        for (size_t off = 0; off < v.size(); off += VH_MAX_GRID_Y) {
            const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, v.size() - off);
            std::vector<vh_item> chunk(v.begin() + off, v.begin() + off + cnt);
            const vh_item *d = vh_stage_vec(chunk);
            vh_prof_scope ps_(VH_PROF_NTT_FP64);
            const bool pipe = OPT.ntt_mode == 2 && C.pipe_blocks > 0;
            const unsigned nb0 = (OPT.ntt_tile_grid & 1) ? UINT_MAX : C.pipe_blocks;
            const unsigned nb3 = ((OPT.ntt_tile_grid >> 3) & 1) ? UINT_MAX : C.pipe_blocks;
            if (which == 0 && pipe) VH_LAUNCH(vh_launch_ntt16p(S(), d, cnt, P->d_limbs, 0, logn, 0, R1, 1, 1, 0, nb0, col_wide(cnt, logn)));
            else if (which == 0) VH_LAUNCH(vh_launch_ntt16(S(), d, cnt, P->d_limbs, 0, logn, 0, R1, 1, 1, 0));
            else if (which == 1 && pipe)
                VH_LAUNCH(vh_launch_ntt16p_row_mul(S(), d, cnt, P->d_limbs, logn, C.cus, (int)((OPT.ntt_tile_grid >> 4) & 1)));
            else if (which == 1) VH_LAUNCH(vh_launch_ntt16_row_mul_inv(S(), d, cnt, P->d_limbs, logn));
            else if (pipe) VH_LAUNCH(vh_launch_ntt16p(S(), d, cnt, P->d_limbs, 1, logn, 0, R1, 1, 0, 1, nb3, col_wide(cnt, logn)));
            else VH_LAUNCH(vh_launch_ntt16(S(), d, cnt, P->d_limbs, 1, logn, 0, R1, 1, 0, 1));
        }
    };
    launch(col, 0);
    launch(mid, 1);
    launch(inv, 2);
}

static void ntt_mul_intt(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count)
{
    incNTT t = z[0]->ntt;
    const uint64_t N = t->N;
    std::vector<RNS_Polynomial> tmp(count);
    std::vector<vh_mulreq> muls;
    for (uint64_t i = 0; i < count; i++) {
        VH_REQUIRE(z[i] != x[i] && z[i] != y[i]);
        dev_ready(x[i]); dev_ready(y[i]); dev_ready(z[i]);
        const uint64_t m = x[i]->rns_mask & y[i]->rns_mask;
        check_rows(z[i], m);
        tmp[i] = new_poly(N, m, t, false, false);
        z[i]->rns_mask = m;
    }
    vh_pool *P = VPOOL(t);
    const int logn = ilog2(P->n);
    bool fused = t->split_degree == 1 && OPT.ntt_mode != 1 && logn >= 13 && logn <= 16;
    for (uint64_t i = 0; i < count && fused; i++)
        for_active(z[i]->rns_mask, t->l, [&](uint64_t li) { fused = fused && P->h_limbs[li].fp64; });
    if (fused) {
        // Three launches: forward column passes of x (into z) and y (into tmp); forward
        // row stages, product and inverse row stages in one kernel; inverse column pass.
        // Items limb-major: the pipelined passes reload their twiddle table head when the
        // limb changes.
        std::vector<vh_item> col, mid, inv;
        for (uint64_t li = 0; li < t->l; li++)
            for (uint64_t i = 0; i < count; i++) {
                if (!((z[i]->rns_mask >> li) & 1)) continue;
                col.push_back(mk_item(drow(x[i], li), drow(z[i], li), (uint32_t)li));
                col.push_back(mk_item(drow(y[i], li), drow(tmp[i], li), (uint32_t)li));
                mid.push_back(mk_item(drow(z[i], li), drow(z[i], li), (uint32_t)li, drow(tmp[i], li)));
                inv.push_back(mk_item(drow(z[i], li), drow(z[i], li), (uint32_t)li));
            }
        fused_product_launches(col, mid, inv, P, logn);
        for (auto p : tmp) free_RNS_polynomial(p);
        return;
    }
    {
        // forward transforms of x into z and of y into a temporary, common limbs only
        std::vector<vh_item> items;
        for (uint64_t i = 0; i < count; i++) {
            const uint64_t m = z[i]->rns_mask;
            for_active(m, t->l, [&](uint64_t li) {
                for (uint64_t k = 0; k < t->split_degree; k++) {
                    items.push_back(mk_item(drow(x[i], li) + k * P->n, drow(z[i], li) + k * P->n, (uint32_t)li));
                    items.push_back(mk_item(drow(y[i], li) + k * P->n, drow(tmp[i], li) + k * P->n, (uint32_t)li));
                }
            });
        }
        run_ntt(items, P->d_limbs, P->h_limbs.data(), P->n, 0);
    }
    for (uint64_t i = 0; i < count; i++) muls.push_back({z[i], z[i], tmp[i], z[i]->rns_mask});
    if (t->split_degree == 1) {
        run_muls(muls, 0);
    } else {
        // the wrap-aware product needs an operand distinct from its output
        for (uint64_t i = 0; i < count; i++) {
            RNS_Polynomial zz = polynomial_new_RNS_polynomial(N, z[i]->rns_mask, t);
            copy_rows(zz, z[i], z[i]->rns_mask);
            run_muls({{z[i], zz, tmp[i], z[i]->rns_mask}}, 0);
            free_RNS_polynomial(zz);
        }
    }
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> inv;
    for (uint64_t i = 0; i < count; i++) inv.push_back({z[i], z[i]});
    ntt_rows(inv, 1);
    for (auto p : tmp) free_RNS_polynomial(p);
}

extern "C" void vfhe_hip_ntt_mul_intt_batch(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count)
{
    VH_ENTER;
    if (count == 0) return;
    ntt_mul_intt(z, x, y, count);
}

// The same through a graph plan (plan_run): the three launches of a batch shape are
// captured once and replayed with the operands' work lists rebuilt; the temporaries
// are the plan's own blocks.
extern "C" void vfhe_hip_ntt_mul_intt_batch_graph(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y,
                                                  uint64_t count)
{
    VH_ENTER;
    if (count == 0) return;
    VH_REQUIRE(z && x && y);
    for (uint64_t i = 0; i < count; i++) { dev_ready(x[i]); dev_ready(y[i]); dev_ready(z[i]); }
    std::vector<uint64_t> key = {2, (uint64_t)(uintptr_t)z[0]->ntt, count, (uint64_t)g_lane, (uint64_t)OPT.ntt_mode,
                                 (uint64_t)OPT.ntt_chunk_mb, (uint64_t)OPT.ntt_tile_grid, (uint64_t)OPT.ntt_col_wide,
                                 (uint64_t)OPT.ntt_small};
    for (uint64_t i = 0; i < count; i++) {
        VH_REQUIRE(z[i]->ntt == z[0]->ntt);
        for (RNS_Polynomial p : {x[i], y[i], z[i]}) { key.push_back(p->rns_mask); key.push_back(p->allocated_l); }
    }
    plan_run(key, [&] { ntt_mul_intt(z, x, y, count); },
             [&] {
                 std::vector<uint64_t> m;
                 for (uint64_t i = 0; i < count; i++) m.push_back(z[i]->rns_mask);
                 return m;
             },
             [&](const std::vector<uint64_t> &m) {
                 for (uint64_t i = 0; i < count; i++) z[i]->rns_mask = m[i];
             });
}



// ---------------------------------------------------------------------------
// CKKS encoder FFT (modules/arith/c/src/complex.c).  Twiddle tables live on the
// device; the host-buffer entry points copy the row in, transform it on the
// device and copy it back, keeping the reference's in-place contract.
// ---------------------------------------------------------------------------
struct vh_rous {
    double2 *d;
    uint64_t size;
};

static void *load_rous(const double *re, const double *im, uint64_t size)
{
    VH_REQUIRE(re && im && size >= 2);
    std::vector<double2> h(size);
    for (uint64_t i = 0; i < size; i++) h[i] = make_double2(re[i], im[i]);
    vh_rous *r = new vh_rous();
    r->size = size;
    VH_CHECK(hipMalloc((void **)&r->d, size * sizeof(double2)));
    VH_CHECK(hipMemcpy(r->d, h.data(), size * sizeof(double2), hipMemcpyHostToDevice));
    return r;
}

extern "C" void *load_rous_CT(double *rous_real, double *rous_imag, uint64_t size)
{
    VH_ENTER;
    return load_rous(rous_real, rous_imag, size);
}

extern "C" void *load_rous_GS(double *rous_real, double *rous_imag, uint64_t size)
{
    VH_ENTER;
    return load_rous(rous_real, rous_imag, size);
}

static int log2_exact(uint64_t n)
{
    VH_REQUIRE(n >= 2 && (n & (n - 1)) == 0);
    return ilog2(n);
}

// Transforms `count` device rows of 2n doubles in place.
static void fft_rows(double *const *d_rows_dev, unsigned count, const vh_rous *ws, uint64_t n, int dir)
{
    VH_REQUIRE(ws && n <= ws->size);
    vh_prof_scope ps_(VH_PROF_OTHER);
    VH_LAUNCH(vh_launch_fft(S(), d_rows_dev, count, ws->d, log2_exact(n), dir, OPT.fft_fma ? 1 : 0));
}

static void fft_host_row(double *x, void *ws, uint64_t n, int dir)
{
    double *d = (double *)vh_dmalloc(2 * n * 8);
    VH_CHECK(hipMemcpyAsync(d, x, 2 * n * 8, hipMemcpyHostToDevice, S()));
    std::vector<double *> rows = {d};
    fft_rows(vh_stage_vec(rows), 1, (const vh_rous *)ws, n, dir);
    VH_CHECK(hipMemcpyAsync(x, d, 2 * n * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
}

extern "C" void CT_NR(double *x, void *ws, uint64_t n)
{
    VH_ENTER;
    fft_host_row(x, ws, n, 0);
}

extern "C" void GS_RN(double *x, void *ws, uint64_t n)
{
    VH_ENTER;
    fft_host_row(x, ws, n, 1);
}

static void bitrev_rows(double *const *d_rows, double *const *d_tmp, const std::vector<double *> &rows,
                        const std::vector<double *> &tmp, uint64_t N, uint32_t prec)
{
    VH_REQUIRE(prec <= 20 && N == ((uint64_t)1 << prec));
    {
        vh_prof_scope ps_(VH_PROF_OTHER);
        VH_LAUNCH(vh_launch_fft_bitrev(S(), d_rows, d_tmp, (unsigned)rows.size(), N, prec));
    }
    for (size_t k = 0; k < rows.size(); k++)
        VH_CHECK(hipMemcpyAsync(rows[k], tmp[k], 2 * N * 8, hipMemcpyDeviceToDevice, S()));
}

extern "C" void bit_reverse_array(double *v, uint64_t N, uint32_t prec)
{
    VH_ENTER;
    double *d = (double *)vh_dmalloc(4 * N * 8);
    VH_CHECK(hipMemcpyAsync(d, v, 2 * N * 8, hipMemcpyHostToDevice, S()));
    std::vector<double *> rows = {d}, tmp = {d + 2 * N};
    bitrev_rows(vh_stage_vec(rows), vh_stage_vec(tmp), rows, tmp, N, prec);
    VH_CHECK(hipMemcpyAsync(v, d, 2 * N * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
}

extern "C" void complex_poly_scale_double(double *v, double scale, uint64_t N)
{
    VH_ENTER;
    double *d = (double *)vh_dmalloc(2 * N * 8);
    VH_CHECK(hipMemcpyAsync(d, v, 2 * N * 8, hipMemcpyHostToDevice, S()));
    std::vector<double *> rows = {d};
    VH_LAUNCH(vh_launch_fft_scale(S(), vh_stage_vec(rows), 1, scale, 2 * N));
    VH_CHECK(hipMemcpyAsync(v, d, 2 * N * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
}

// complex.c:39-50: the first out->ntt->N doubles of `in`, rounded half-to-even,
// as signed coefficients of `out` (int_array_to_RNS: every active limb, NTT form).
extern "C" void complex_poly_round_to_RNS(RNS_Polynomial out, double *in, uint64_t N)
{
    (void)N;
    VH_ENTER;
    const uint64_t n = out->ntt->N;
    double *d = (double *)vh_dmalloc(n * 8);
    uint64_t *di = (uint64_t *)vh_dmalloc(n * 8);
    VH_CHECK(hipMemcpyAsync(d, in, n * 8, hipMemcpyHostToDevice, S()));
    std::vector<double *> rows = {d};
    std::vector<uint64_t *> outs = {di};
    VH_LAUNCH(vh_launch_fft_round(S(), vh_stage_vec(rows), vh_stage_vec(outs), 1, n));
    int_rows_to_RNS({{out, di}});
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
    vh_dfree(di);
}

// complex.c:455-500: per row, bit-reverse, GS_RN, scale by temp_delta / n_complex,
// round into outs_rns[i] (NTT form).  Rows are host buffers of 2 n_complex doubles
// and, as in the reference, hold the scaled transform afterwards.  All rows are
// transformed in one launch sequence.
extern "C" void complex_polys_ifft_scale_round_to_RNS_batch(void **rows_in, void **outs_rns, uint64_t count,
                                                           uint64_t n_complex, uint32_t log_prec, void *gs_ws,
                                                           double temp_delta)
{
    VH_ENTER;
    if (count == 0) return;
    const uint64_t n2 = 2 * n_complex;
    for (uint64_t i = 0; i < count; i++)
        VH_REQUIRE(((RNS_Polynomial)outs_rns[i])->ntt->N <= n2);
    double *d = (double *)vh_dmalloc(2 * count * n2 * 8);
    uint64_t *di = (uint64_t *)vh_dmalloc(count * n2 * 8);
    std::vector<double *> rows(count), tmp(count);
    std::vector<uint64_t *> ints(count);
    for (uint64_t i = 0; i < count; i++) {
        rows[i] = d + i * n2;
        tmp[i] = d + (count + i) * n2;
        ints[i] = di + i * n2;
        VH_CHECK(hipMemcpyAsync(rows[i], rows_in[i], n2 * 8, hipMemcpyHostToDevice, S()));
    }
    for (uint64_t off = 0; off < count; off += VH_MAX_GRID_Y) {
        const uint64_t cnt = std::min<uint64_t>(VH_MAX_GRID_Y, count - off);
        std::vector<double *> r(rows.begin() + off, rows.begin() + off + cnt), t(tmp.begin() + off, tmp.begin() + off + cnt);
        std::vector<uint64_t *> o(ints.begin() + off, ints.begin() + off + cnt);
        double *const *d_rows = vh_stage_vec(r);
        bitrev_rows(d_rows, vh_stage_vec(t), r, t, n_complex, log_prec);
        fft_rows(d_rows, (unsigned)cnt, (const vh_rous *)gs_ws, n_complex, 1);
        VH_LAUNCH(vh_launch_fft_scale(S(), d_rows, (unsigned)cnt, temp_delta / (double)n_complex, n2));
        VH_LAUNCH(vh_launch_fft_round(S(), d_rows, vh_stage_vec(o), (unsigned)cnt, n2));
    }
    std::vector<std::pair<RNS_Polynomial, const uint64_t *>> jobs;
    for (uint64_t i = 0; i < count; i++) jobs.push_back({(RNS_Polynomial)outs_rns[i], ints[i]});
    int_rows_to_RNS(jobs);
    for (uint64_t i = 0; i < count; i++)
        VH_CHECK(hipMemcpyAsync(rows_in[i], rows[i], n2 * 8, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
    vh_dfree(di);
}

// ---------------------------------------------------------------------------
// On-device sampling.  Keystream: ChaCha20 under a 256-bit engine key drawn
// from the OS (getrandom) at first use, or derived from a deterministic seed
// (prng.c:132-143 semantics: tests only); every sampling call takes a fresh
// nonce, and rows/jobs inside a call take disjoint counter ranges.
// ---------------------------------------------------------------------------
struct vh_rng_state {
    bool init = false;
    uint32_t key[8];
    uint64_t calls = 0;
};
static vh_rng_state RNG;

static void rng_seed_os()
{
    size_t got = 0;
    while (got < sizeof(RNG.key)) {
        const ssize_t r = getrandom((char *)RNG.key + got, sizeof(RNG.key) - got, 0);
        if (r <= 0) vh_die("getrandom failed", __FILE__, __LINE__);
        got += (size_t)r;
    }
    RNG.calls = 0;
    RNG.init = true;
}

static vh_rng_key rng_next()
{
    if (!RNG.init) rng_seed_os();
    vh_rng_key k;
    memcpy(k.k, RNG.key, sizeof(k.k));
    k.nonce = RNG.calls++;
    return k;
}

extern "C" void vfhe_prng_set_deterministic_seed(uint64_t seed)
{
    VH_ENTER;
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) {   // splitmix64 expansion of the seed into the key
        uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        RNG.key[2 * i] = (uint32_t)z;
        RNG.key[2 * i + 1] = (uint32_t)(z >> 32);
    }
    RNG.calls = 0;
    RNG.init = true;
}

extern "C" void vfhe_prng_clear_deterministic_seed(void)
{
    VH_ENTER;
    rng_seed_os();
}

extern "C" void vfhe_hip_chacha20_blocks(const uint32_t *key, uint64_t nonce, uint64_t ctr0, uint64_t nblocks,
                                         uint32_t *out)
{
    VH_ENTER;
    vh_rng_key k;
    memcpy(k.k, key, sizeof(k.k));
    k.nonce = nonce;
    uint32_t *d = (uint32_t *)vh_dmalloc(nblocks * 64);
    VH_LAUNCH(vh_launch_chacha_raw(S(), k, ctr0, nblocks, d));
    VH_CHECK(hipMemcpyAsync(out, d, nblocks * 64, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(d);
}

static void sample_uniform_rows(const std::vector<std::pair<RNS_Polynomial, uint64_t>> &rows)
{
    if (rows.empty()) return;
    RNS_Polynomial p0 = rows[0].first;
    const uint64_t N = p0->ntt->N;
    VH_REQUIRE(N % 8 == 0);
    std::vector<vh_item> items;
    uint32_t k = 0;
    for (auto &pr : rows) {
        vh_item it = mk_item(nullptr, drow(pr.first, pr.second), (uint32_t)pr.second);
        it.aux = k++;
        items.push_back(it);
    }
    const vh_rng_key key = rng_next();
    for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
        std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
        vh_prof_scope ps_(VH_PROF_OTHER);
        VH_LAUNCH(vh_launch_sample_uniform(S(), vh_stage_vec(chunk), cnt, DL(p0), key, N));
    }
}

// One noise vector per polynomial (shared by its active limbs); mode 1 writes
// truncated raw int64 values into limb row 0 instead.
static void sample_gaussian_polys(const std::vector<RNS_Polynomial> &polys, double sigma, uint32_t mode)
{
    if (polys.empty()) return;
    const uint64_t N = polys[0]->ntt->N;
    VH_REQUIRE(N % 4 == 0);
    std::vector<vh_item> items;
    std::vector<vh_gauss_job> jobs;
    for (auto p : polys) {
        VH_REQUIRE(p->ntt == polys[0]->ntt);
        vh_gauss_job jb = {(uint32_t)items.size(), 0, (uint32_t)jobs.size(), mode};
        if (mode) {
            items.push_back(mk_item(nullptr, drow(p, 0), 0));
        } else {
            for_active(p->rns_mask, std::min(p->ntt->l, p->allocated_l),
                       [&](uint64_t i) { items.push_back(mk_item(nullptr, drow(p, i), (uint32_t)i)); });
        }
        jb.count = (uint32_t)items.size() - jb.first;
        jobs.push_back(jb);
    }
    const vh_rng_key key = rng_next();
    const vh_item *d_items = vh_stage_vec(items);
    for (size_t off = 0; off < jobs.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, jobs.size() - off);
        std::vector<vh_gauss_job> chunk(jobs.begin() + off, jobs.begin() + off + cnt);
        vh_prof_scope ps_(VH_PROF_OTHER);
        VH_LAUNCH(vh_launch_sample_gaussian(S(), vh_stage_vec(chunk), cnt, d_items, DL(polys[0]), key, sigma, N));
    }
}

extern "C" void polynomial_gen_random_RNSc_polynomial(RNSc_Polynomial out)
{
    VH_ENTER;
    dev_ready(out);
    std::vector<std::pair<RNS_Polynomial, uint64_t>> rows;
    for_active(out->rns_mask, std::min(out->ntt->l, out->allocated_l), [&](uint64_t i) { rows.push_back({out, i}); });
    sample_uniform_rows(rows);
}

extern "C" void polynomial_gen_gaussian_RNSc_polynomial(RNSc_Polynomial out, double sigma)
{
    VH_ENTER;
    dev_ready(out);
    sample_gaussian_polys({out}, sigma, 0);
}

// mlwe.c:190-200: b = e + sum_i s_i * a_i with a_i uniform, all in NTT form.
static void sample_of_zero(RNS_MLWE out, RNS_MLWE_Key key)
{
    for (auto p : comps(out)) dev_ready(p);
    sample_gaussian_polys({out->b}, key->sigma, 0);
    std::vector<std::pair<RNS_Polynomial, uint64_t>> rows;
    for (uint64_t i = 0; i < out->r; i++)
        for_active(out->a[i]->rns_mask, std::min(out->a[i]->ntt->l, out->a[i]->allocated_l),
                   [&](uint64_t j) { rows.push_back({out->a[i], j}); });
    sample_uniform_rows(rows);
    std::vector<std::pair<RNS_Polynomial, RNS_Polynomial>> pairs;
    for (auto p : comps(out)) pairs.push_back({p, p});
    ntt_rows(pairs, 0);
    for (uint64_t i = 0; i < out->r; i++) polynomial_mul_addto_RNS_polynomial(out->b, key->s_RNS[i], out->a[i]);
}

extern "C" void mlwe_RNS_sample_of_zero(RNS_MLWE out, RNS_MLWE_Key key)
{
    VH_ENTER;
    sample_of_zero(out, key);
}

extern "C" void mlwe_RNSc_sample_of_zero(RNSc_MLWE out, RNS_MLWE_Key key)
{
    VH_ENTER;
    sample_of_zero(out, key);
    mlwe_RNS_to_RNSc(out, out);
}

// mlwe.c:281-293 allocate with mask = key->l (the limb count used as a mask),
// kept as the reference has it so masks match.
extern "C" RNS_MLWE mlwe_new_RNS_sample_of_zero(RNS_MLWE_Key key)
{
    VH_ENTER;
    RNS_MLWE res = mlwe_alloc_RNS_sample(key->N, key->r, key->l, key->s_RNS[0]->ntt);
    sample_of_zero(res, key);
    return res;
}

extern "C" RNSc_MLWE mlwe_new_RNSc_sample_of_zero(RNS_MLWE_Key key)
{
    VH_ENTER;
    RNSc_MLWE res = mlwe_alloc_RNS_sample(key->N, key->r, key->l, key->s_RNS[0]->ntt);
    mlwe_RNSc_sample_of_zero(res, key);
    return res;
}

extern "C" RNS_MLWE mlwe_new_RNS_trivial_sample_of_zero(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt)
{
    VH_ENTER;
    RNS_MLWE res = mlwe_alloc_RNS_sample(N, r, mask, ntt);
    mlwe_RNS_trivial_sample_of_zero(res);
    return res;
}

// mlwe.c:236-241
extern "C" void mlwe_RNSc_sample(RNSc_MLWE out, RNS_MLWE_Key key, RNSc_Polynomial m)
{
    VH_ENTER;
    VH_REQUIRE(m->rns_mask == out->b->rns_mask);
    mlwe_RNSc_sample_of_zero(out, key);
    polynomial_add_RNSc_polynomial(out->b, out->b, m);
}

// mlwe.c:70-83: secret coefficients (int64)N(0, key_sigma) (truncated), then RNS form.
extern "C" RNS_MLWE_Key mlwe_new_RNS_gaussian_key(uint64_t N, uint64_t r, uint64_t l, double key_sigma, incNTT ntt,
                                                  double sigma)
{
    VH_ENTER;
    std::vector<uint64_t> arr(r * N);
    RNS_Polynomial tmp = polynomial_new_RNS_polynomial(N, 1, ntt);
    for (uint64_t i = 0; i < r; i++) {
        sample_gaussian_polys({tmp}, key_sigma, 1);
        VH_CHECK(hipMemcpyAsync(arr.data() + i * N, VP(tmp)->d, N * 8, hipMemcpyDeviceToHost, S()));
    }
    VH_CHECK(hipStreamSynchronize(S()));
    free_RNS_polynomial(tmp);
    return mlwe_new_RNS_key_from_array(arr.data(), N, r, l, ntt, sigma);
}

// ---------------------------------------------------------------------------
// NTT schedule autotuner: times forward+inverse round trips of `rows` scratch
// rows (limbs of `ntt` in turn) under each candidate schedule and keeps the
// fastest in the engine options.  Candidates: two launches per transform of the
// pipelined pass kernel (every pass one tile per workgroup with 512-thread column
// passes; row passes one tile per workgroup; every pass persistent), of the plain
// pass kernel (whole batch or 256 / 128 MiB sub-batches), and (when the XCC
// placement probe passed and 2^13 <= N <= 2^16) the persistent per-XCD kernel
// with queue lags 3, 5 and 8.
// us_out[k] receives candidate k's microseconds per round trip (-1 if skipped);
// returns the index of the chosen candidate.
// ---------------------------------------------------------------------------
struct vh_sched { long mode, chunk_mb, lag, occ, tile_grid, col_wide; };
static const vh_sched VH_SCHEDS[] = {{2, 0, 3, 2, 15, 1},  {2, 0, 3, 2, 6, 0},  {2, 0, 3, 2, 0, 0},
                                     {0, 0, 3, 2, 15, 1},  {0, 256, 3, 2, 15, 1}, {0, 128, 3, 2, 15, 1},
                                     {1, 0, 3, 2, 15, 1},  {1, 0, 5, 2, 15, 1},  {1, 0, 8, 2, 15, 1}};

extern "C" long vfhe_hip_autotune_ntt(incNTT ntt, uint64_t rows, int reps, double *us_out)
{
    VH_ENTER;
    const long nc = (long)(sizeof(VH_SCHEDS) / sizeof(VH_SCHEDS[0]));
    VH_REQUIRE(ntt && rows > 0 && reps > 0);
    vh_pool *P = VPOOL(ntt);
    const uint64_t n = P->n;
    const int logn = ilog2(n);
    uint64_t *d = (uint64_t *)vh_dmalloc(rows * n * 8);
    std::vector<vh_item> items;
    for (uint64_t i = 0; i < rows; i++) {
        vh_item it = mk_item(nullptr, d + i * n, (uint32_t)(i % ntt->l));
        it.aux = (uint32_t)i;
        items.push_back(it);
    }
    {
        const vh_rng_key key = rng_next();
        for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
            const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
            std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
            VH_LAUNCH(vh_launch_sample_uniform(S(), vh_stage_vec(chunk), cnt, P->d_limbs, key, n));
        }
    }
    for (auto &it : items) it.src = it.dst;
    const vh_opts saved = OPT;
    hipEvent_t e0, e1;
    VH_CHECK(hipEventCreate(&e0));
    VH_CHECK(hipEventCreate(&e1));
    long best = 0;
    double best_us = 1e300;
    std::vector<double> us(nc, -1.0);
    std::vector<bool> wide_ran(nc, false);
    for (int round = 0; round < 2; round++)   // two interleaved rounds, the faster of each kept
        for (long k = 0; k < nc; k++) {
            const vh_sched &sc = VH_SCHEDS[k];
            if (sc.mode == 1 && !(C.xcd_ok && logn >= 13 && logn <= 16)) continue;
            OPT.ntt_mode = sc.mode; OPT.ntt_chunk_mb = sc.chunk_mb; OPT.xcd_lag = sc.lag; OPT.xcd_occ = sc.occ;
            OPT.ntt_tile_grid = (saved.ntt_tile_grid & ~15L) | sc.tile_grid;
            OPT.ntt_col_wide = sc.col_wide;
            const long wide0 = g_col_wide_launches;
            for (int w = 0; w < (round == 0 && k == 0 ? 3 : 1); w++) {   // warm-up
                run_ntt(items, P->d_limbs, P->h_limbs.data(), n, 0);
                run_ntt(items, P->d_limbs, P->h_limbs.data(), n, 1);
            }
            VH_CHECK(hipEventRecord(e0, S()));
            for (int r = 0; r < reps; r++) {
                run_ntt(items, P->d_limbs, P->h_limbs.data(), n, 0);
                run_ntt(items, P->d_limbs, P->h_limbs.data(), n, 1);
            }
            VH_CHECK(hipEventRecord(e1, S()));
            VH_CHECK(hipEventSynchronize(e1));
            vh_check_device_error();
            float ms = 0;
            VH_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double t = ms * 1e3 / reps;
            if (us[k] < 0 || t < us[k]) us[k] = t;
            wide_ran[k] = g_col_wide_launches > wide0;
        }
    for (long k = 0; k < nc; k++) {
        if (us_out) us_out[k] = us[k];
        if (us[k] > 0 && us[k] < best_us) { best_us = us[k]; best = k; }
    }
    VH_CHECK(hipEventDestroy(e0));
    VH_CHECK(hipEventDestroy(e1));
    // Adopt the winner's timed settings only: the fused product's row-kernel bit (tile_grid
    // bit 4) keeps its value, and the 512-thread column rule is enabled only when the winning
    // candidate's timed launches actually ran wide (otherwise the caller's setting stays).
    OPT = saved;
    OPT.ntt_mode = VH_SCHEDS[best].mode; OPT.ntt_chunk_mb = VH_SCHEDS[best].chunk_mb;
    OPT.xcd_lag = VH_SCHEDS[best].lag; OPT.xcd_occ = VH_SCHEDS[best].occ;
    OPT.ntt_tile_grid = (saved.ntt_tile_grid & ~15L) | VH_SCHEDS[best].tile_grid;
    if (VH_SCHEDS[best].col_wide == 0 || wide_ran[best]) OPT.ntt_col_wide = VH_SCHEDS[best].col_wide;
    vh_dfree(d);
    return best;
}

// Times `run` under each candidate setting of the options `apply` sets (two interleaved
// rounds of `reps` calls after a warm-up, the faster round kept), adopts the fastest and
// returns its index; us_out[k] = microseconds per call of candidate k.  Options the
// candidates do not set keep the caller's values.
static long autotune_generic(long nc, const std::function<void(long)> &apply, const std::function<void()> &run,
                             int reps, double *us_out)
{
    VH_REQUIRE(nc > 0 && reps > 0);
    const vh_opts saved = OPT;
    hipEvent_t e0, e1;
    VH_CHECK(hipEventCreate(&e0));
    VH_CHECK(hipEventCreate(&e1));
    std::vector<double> us(nc, -1.0);
    for (int round = 0; round < 2; round++)
        for (long k = 0; k < nc; k++) {
            OPT = saved;
            apply(k);
            for (int w = 0; w < (round == 0 ? 2 : 1); w++) run();   // warm-up (tables, caches)
            VH_CHECK(hipEventRecord(e0, S()));
            for (int r = 0; r < reps; r++) run();
            VH_CHECK(hipEventRecord(e1, S()));
            VH_CHECK(hipEventSynchronize(e1));
            vh_check_device_error();
            float ms = 0;
            VH_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const double t = ms * 1e3 / reps;
            if (us[k] < 0 || t < us[k]) us[k] = t;
        }
    VH_CHECK(hipEventDestroy(e0));
    VH_CHECK(hipEventDestroy(e1));
    long best = 0;
    for (long k = 0; k < nc; k++) {
        if (us_out) us_out[k] = us[k];
        if (us[k] < us[best]) best = k;
    }
    OPT = saved;
    apply(best);
    return best;
}

// The fused product's schedule: its row kernel launched one tile per workgroup or
// persistent (ntt.tile_grid bit 4) x its column passes at 256 threads, 512 threads by the
// occupancy rule, or 512 threads always (ntt.col_wide 0 / 1 / 2): 6 candidates, timed on
// the caller's operands (z is overwritten).
extern "C" long vfhe_hip_autotune_product(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count,
                                          int reps, double *us_out)
{
    VH_ENTER;
    VH_REQUIRE(z && x && y && count > 0);
    return autotune_generic(
        6,
        [&](long k) {
            OPT.ntt_tile_grid = (OPT.ntt_tile_grid & ~16L) | ((k & 1) ? 0L : 16L);
            OPT.ntt_col_wide = k >> 1;
        },
        [&] { ntt_mul_intt(z, x, y, count); }, reps, us_out);
}

// The CKKS multiply batch's schedule: column passes (the ModUp pass and the inverse ones)
// at 256 / 512-by-rule / 512 threads (ntt.col_wide), the inverse passes one tile per
// workgroup or persistent (ntt.tile_grid bits 2 and 3), and the pass-through components
// folded into the key inner product or added after the ModDown (ks.fold): 12 candidates,
// timed on the caller's operands (out is overwritten).  Every candidate is bit-exact.
extern "C" long vfhe_hip_autotune_multiply(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                           RNS_MLWE **rlk, uint64_t divide_mask, int reps, double *us_out)
{
    VH_ENTER;
    VH_REQUIRE(out && in1 && in2 && rlk && count > 0);
    const std::vector<RNS_MLWE> o(out, out + count), a(in1, in1 + count), b(in2, in2 + count);
    const std::vector<RNS_MLWE **> ks(count, rlk);
    const std::vector<uint64_t> dm(count, divide_mask);
    std::vector<uint64_t> masks;
    for (auto c : o) for (auto p : comps(c)) masks.push_back(p->rns_mask);
    return autotune_generic(
        12,
        [&](long k) {
            OPT.ntt_col_wide = k % 3;
            OPT.ntt_tile_grid = (OPT.ntt_tile_grid & ~12L) | (((k / 3) & 1) ? 0L : 12L);
            OPT.ks_fold = (k / 6) ? 0 : 1;
        },
        [&] {
            size_t i = 0;
            for (auto c : o) for (auto p : comps(c)) p->rns_mask = masks[i++];   // every call starts alike
            multiply_batch(o, a, b, ks, dm);
        },
        reps, us_out);
}

// ---------------------------------------------------------------------------
// CGGI over the 32-bit torus with an FP64 FFT (vh_torus.hip).  No counterpart in
// the reference (its CGGI16 runs over RNS primes, cggi16.py:46-140); the entry
// points return an error code instead of aborting on bad arguments.
// ---------------------------------------------------------------------------
struct vfhe_torus_key {
    uint64_t n, N, ell;
    uint32_t base_log;
    double2 *bsk;
};
struct vfhe_torus_batch {
    uint64_t count, dim;
    uint32_t *d;
};
static double2 *g_torus_consts = nullptr;

// The transform constants (twiddles and twist), from long double cos / sin rounded once.
static const double2 *torus_consts()
{
    if (g_torus_consts) return g_torus_consts;
    std::vector<double2> h(576);
    const long double pi = 3.141592653589793238462643383279502884L;
    auto root = [&](long double num, long double den) {
        const long double a = pi * num / den;
        return make_double2((double)cosl(a), (double)sinl(a));
    };
    for (int k = 0; k < 8; k++)   // zeta^r w^(r k): angle pi r / 1024 - 2 pi r k / 512
        for (int r = 0; r < 64; r++) h[k * 64 + r] = root((long double)r - 4.0L * ((r * k) % 512), 1024);
    for (int a = 0; a < 8; a++)
        for (int s = 0; s < 8; s++) h[512 + a * 8 + s] = root(-2.0L * ((8 * a * s) % 512), 512);
    double2 *d = nullptr;
    VH_CHECK(hipMalloc(&d, h.size() * sizeof(double2)));
    VH_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
    g_torus_consts = d;
    return d;
}

extern "C" vfhe_torus_batch *vfhe_hip_torus_batch_new(const uint32_t *samples, uint64_t count, uint64_t dim)
{
    VH_ENTER;
    if (!count || !dim || dim > (1ull << 32) || count > (1ull << 38) / dim) return nullptr;
    auto *b = new vfhe_torus_batch{count, dim, nullptr};
    VH_CHECK(hipMalloc(&b->d, count * dim * 4));
    if (samples) VH_CHECK(hipMemcpyAsync(b->d, samples, count * dim * 4, hipMemcpyHostToDevice, S()));
    else VH_CHECK(hipMemsetAsync(b->d, 0, count * dim * 4, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    return b;
}

extern "C" int vfhe_hip_torus_batch_read(const vfhe_torus_batch *b, uint32_t *samples)
{
    VH_ENTER;
    if (!b || !samples) return -1;
    VH_CHECK(hipMemcpyAsync(samples, b->d, b->count * b->dim * 4, hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    return 0;
}

extern "C" void vfhe_hip_torus_batch_free(vfhe_torus_batch *b)
{
    VH_ENTER;
    if (!b) return;
    sync_all_lanes();
    VH_CHECK(hipFree(b->d));
    delete b;
}

extern "C" vfhe_torus_key *vfhe_hip_torus_key_new(const uint32_t *bsk, uint64_t n, uint64_t N, uint64_t ell,
                                                  uint32_t base_log)
{
    VH_ENTER;
    if (!bsk || N != 1024 || !n || n > (1u << 20) || !ell || ell > 4 || !base_log || base_log > 16 ||
        ell * base_log > 32)
        return nullptr;
    const uint64_t polys = n * 4 * ell;
    uint32_t *d_in = nullptr;
    auto *k = new vfhe_torus_key{n, N, ell, base_log, nullptr};
    VH_CHECK(hipMalloc(&d_in, polys * N * 4));
    VH_CHECK(hipMalloc(&k->bsk, polys * (N / 2) * sizeof(double2)));
    VH_CHECK(hipMemcpyAsync(d_in, bsk, polys * N * 4, hipMemcpyHostToDevice, S()));
    VH_LAUNCH(vh_launch_torus_fft(S(), d_in, k->bsk, torus_consts(), polys, 0));
    VH_CHECK(hipStreamSynchronize(S()));
    VH_CHECK(hipFree(d_in));
    return k;
}

extern "C" void vfhe_hip_torus_key_free(vfhe_torus_key *k)
{
    VH_ENTER;
    if (!k) return;
    sync_all_lanes();
    VH_CHECK(hipFree(k->bsk));
    delete k;
}

extern "C" int vfhe_hip_torus_bootstrap(vfhe_torus_batch *out, vfhe_torus_batch *acc_out, const vfhe_torus_key *key,
                                        const vfhe_torus_batch *in, const vfhe_torus_batch *tv, uint64_t torus_base)
{
    VH_ENTER;
    if (!key || !in || !tv || (!out && !acc_out) || !torus_base || torus_base > (1ull << 30)) return -1;
    const uint64_t N = key->N;
    if (in->dim != key->n + 1 || tv->count != 1 || tv->dim != N) return -2;
    if (out && (out->count != in->count || out->dim != N + 1)) return -3;
    if (acc_out && (acc_out->count != in->count || acc_out->dim != 2 * N)) return -3;
    if (in->count > 0x7fffffffull) return -1;
    uint32_t dec_off = 0;
    for (uint64_t l = 1; l <= key->ell; l++) dec_off += 1u << (32 - key->base_log * l + key->base_log - 1);
    if (key->base_log * key->ell < 32) dec_off += 1u << (32 - key->base_log * key->ell - 1);
    vt_br_args A{};
    A.lwe = in->d;
    A.tv = tv->d;
    A.bsk = key->bsk;
    A.consts = torus_consts();
    A.out = out ? out->d : nullptr;
    A.acc_out = acc_out ? acc_out->d : nullptr;
    A.n = (uint32_t)key->n;
    A.ell = (uint32_t)key->ell;
    A.base_log = key->base_log;
    A.offset = (uint32_t)((1ull << 32) / (4 * torus_base));
    A.dec_off = dec_off;
    vh_prof_scope ps_(VH_PROF_OTHER);
    VH_LAUNCH(vh_launch_torus_blind_rotate(S(), A, in->count));
    return 0;
}

extern "C" int vfhe_hip_torus_fft(const uint32_t *in, uint64_t count, uint64_t N, double *out)
{
    VH_ENTER;
    if (!in || !out || N != 1024 || !count || count > (1ull << 24)) return -1;
    uint32_t *d_in = nullptr;
    double2 *d_out = nullptr;
    VH_CHECK(hipMalloc(&d_in, count * N * 4));
    VH_CHECK(hipMalloc(&d_out, count * (N / 2) * sizeof(double2)));
    VH_CHECK(hipMemcpyAsync(d_in, in, count * N * 4, hipMemcpyHostToDevice, S()));
    VH_LAUNCH(vh_launch_torus_fft(S(), d_in, d_out, torus_consts(), count, 1));
    VH_CHECK(hipMemcpyAsync(out, d_out, count * (N / 2) * sizeof(double2), hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    VH_CHECK(hipFree(d_in));
    VH_CHECK(hipFree(d_out));
    return 0;
}

// Measured per-lane operation rates of the device (operations per second over the whole
// chip): rates[0] FP64 FMA, [1] 32-bit multiply-low + add, [2] 32-bit multiply-high,
// [3] 32x32->64 multiply-add, [4] FP64 modular products (vh_mulmod_tw, q < 2^50),
// [5] 64-bit Shoup modular products.  Each kind runs on 8 workgroups of 256 threads per
// CU, timed with events after one warm-up launch.
extern "C" void vfhe_hip_probe_alu_rates(double *rates)
{
    VH_ENTER;
    VH_REQUIRE(rates);
    const unsigned blocks = (unsigned)C.cus * 8;
    const int iters = 4096;
    uint64_t *d = (uint64_t *)vh_dmalloc((size_t)blocks * 256 * 8);
    hipEvent_t e0, e1;
    VH_CHECK(hipEventCreate(&e0));
    VH_CHECK(hipEventCreate(&e1));
    for (int kind = 0; kind < 6; kind++) {
        VH_LAUNCH(vh_launch_alu_rate(S(), kind, iters, blocks, d));
        VH_CHECK(hipEventRecord(e0, S()));
        VH_LAUNCH(vh_launch_alu_rate(S(), kind, iters, blocks, d));
        VH_CHECK(hipEventRecord(e1, S()));
        VH_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        VH_CHECK(hipEventElapsedTime(&ms, e0, e1));
        rates[kind] = (double)blocks * 256 * 8 * iters / (ms * 1e-3);
    }
    VH_CHECK(hipEventDestroy(e0));
    VH_CHECK(hipEventDestroy(e1));
    vh_dfree(d);
}

// ---------------------------------------------------------------------------
// Profiler API
// ---------------------------------------------------------------------------
extern "C" void vfhe_hip_profile_start(void)
{
    VH_ENTER;
    sync_all_lanes();
    for (auto &r : g_prof_recs) { g_prof_free.push_back(r.a); g_prof_free.push_back(r.b); }
    g_prof_recs.clear();
    g_prof = true;
}

extern "C" void vfhe_hip_profile_stop(double *ms, uint64_t *launches)
{
    VH_ENTER;
    sync_all_lanes();
    for (int t = 0; t < VH_PROF_TAGS; t++) { ms[t] = 0; launches[t] = 0; }
    for (auto &r : g_prof_recs) {
        float e = 0;
        VH_CHECK(hipEventElapsedTime(&e, r.a, r.b));
        ms[r.tag] += e;
        launches[r.tag]++;
        g_prof_free.push_back(r.a);
        g_prof_free.push_back(r.b);
    }
    g_prof_recs.clear();
    g_prof = false;
}

// ---------------------------------------------------------------------------
// Host allocation helper of the reference (misc_tp.c:48-62): 64-byte aligned,
// exit on failure.  Buffers are released with libc free (misc.cdef).
// ---------------------------------------------------------------------------
extern "C" void *safe_aligned_malloc(uint64_t size)
{
    void *p = nullptr;
    if (posix_memalign(&p, 64, size ? size : 64) != 0 || !p) exit(EXIT_FAILURE);
    return p;
}

// ---------------------------------------------------------------------------
// Slot operations on NTT-form polynomials (rns_polynomial.c:1228-1300) and the
// slot-wise inverse (rns_polynomial.c:1384-1432), one launch per call.
// ---------------------------------------------------------------------------
static void slot_rows(RNS_Polynomial out, RNS_Polynomial in, int op, uint64_t scalar, uint32_t aux, unsigned *flag)
{
    incNTT t = in->ntt;
    const uint64_t ps = t->N / t->split_degree;
    std::vector<vh_item> items;
    for_active(in->rns_mask, t->l, [&](uint64_t i) {
        vh_item it = mk_item(drow(in, i), drow(out, i), (uint32_t)i, (const uint64_t *)flag, scalar);
        it.aux = aux;
        items.push_back(it);
    });
    for (size_t off = 0; off < items.size(); off += VH_MAX_GRID_Y) {
        const unsigned cnt = (unsigned)std::min<size_t>(VH_MAX_GRID_Y, items.size() - off);
        std::vector<vh_item> chunk(items.begin() + off, items.begin() + off + cnt);
        vh_prof_scope ps_(VH_PROF_ELTWISE);
        VH_LAUNCH(vh_launch_slot(S(), vh_stage_vec(chunk), cnt, DL(in), op, t->N, ps));
    }
}

static void slot_prepare(RNS_Polynomial out, RNS_Polynomial in)
{
    VH_REQUIRE(out && in && out->ntt->N == in->ntt->N && out->ntt == in->ntt);
    dev_ready(in);
    dev_ready(out);
    out->rns_mask = in->rns_mask;
    check_rows(out, out->rns_mask);
}

extern "C" void polynomial_RNS_broadcast_slot(RNS_Polynomial out, RNS_Polynomial in, uint64_t slot_idx)
{
    VH_ENTER;
    VH_REQUIRE(slot_idx < in->ntt->N / in->ntt->split_degree);
    slot_prepare(out, in);
    slot_rows(out, in, 0, slot_idx, 0, nullptr);
}

extern "C" void polynomial_RNS_rotate_slot(RNS_Polynomial out, RNS_Polynomial in, uint64_t rot)
{
    VH_ENTER;
    VH_REQUIRE(out != in && rot <= in->ntt->N / in->ntt->split_degree);
    slot_prepare(out, in);
    slot_rows(out, in, 1, rot, 0, nullptr);
}

extern "C" void polynomial_RNS_copy_slot(RNS_Polynomial out, uint64_t dst, RNS_Polynomial in, uint64_t src)
{
    VH_ENTER;
    const uint64_t ps = in->ntt->N / in->ntt->split_degree;
    VH_REQUIRE(dst < ps && src < ps);
    slot_prepare(out, in);
    slot_rows(out, in, 2, src, (uint32_t)dst, nullptr);
}

// Returns 0, -2 when a slot is zero (the reference's code, polynomial.py:236-245), or
// -1 for a split degree above 1 (whose slots are extension-field elements; the
// extension-field arithmetic, field.c, is not part of this engine).
extern "C" int polynomial_RNS_inverse(RNS_Polynomial out, RNS_Polynomial in)
{
    VH_ENTER;
    if (in->ntt->split_degree != 1) {
        fprintf(stderr, "vfhe hip engine: polynomial_RNS_inverse supports split degree 1 only\n");
        return -1;
    }
    slot_prepare(out, in);
    unsigned *flag = (unsigned *)vh_dmalloc(sizeof(unsigned));
    VH_LAUNCH(vh_launch_zero_u32(S(), flag, 1));
    slot_rows(out, in, 3, 0, 0, flag);
    unsigned h = 0;
    VH_CHECK(hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, S()));
    VH_CHECK(hipStreamSynchronize(S()));
    vh_dfree(flag);
    return h ? -2 : 0;
}

// ---------------------------------------------------------------------------
// LWE samples and keys (modules/mlwe/c/src/lwe.c): host-resident arrays of a
// few hundred words per limb, handled on the host like the reference does.
// Randomness: the engine's ChaCha20 key (vfhe_prng_set_deterministic_seed pins
// it), one nonce per call, expanded on the host by the same RFC 8439 block
// function the device samplers use.
// ---------------------------------------------------------------------------
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void chacha20_host(const vh_rng_key &k, uint64_t ctr, uint32_t o[16])
{
    uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      k.k[0], k.k[1], k.k[2], k.k[3], k.k[4], k.k[5], k.k[6], k.k[7],
                      (uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)k.nonce, (uint32_t)(k.nonce >> 32)};
    uint32_t s[16];
    memcpy(s, x, sizeof(s));
    auto qr = [&](int a, int b, int c, int d) {
        s[a] += s[b]; s[d] = rotl32(s[d] ^ s[a], 16);
        s[c] += s[d]; s[b] = rotl32(s[b] ^ s[c], 12);
        s[a] += s[b]; s[d] = rotl32(s[d] ^ s[a], 8);
        s[c] += s[d]; s[b] = rotl32(s[b] ^ s[c], 7);
    };
    for (int r = 0; r < 10; r++) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) o[i] = s[i] + x[i];
}

struct vh_host_stream {
    vh_rng_key key;
    uint64_t ctr = 0;
    uint32_t buf[16];
    int pos = 16;
    vh_host_stream() : key(rng_next()) {}
    uint64_t next64()
    {
        if (pos >= 16) { chacha20_host(key, ctr++, buf); pos = 0; }
        const uint64_t v = (uint64_t)buf[pos] | ((uint64_t)buf[pos + 1] << 32);
        pos += 2;
        return v;
    }
    uint64_t residue(uint64_t q) { return (uint64_t)(((u128)next64() * q) >> 64); }
    // rounded-toward-zero Box-Muller normal sample (misc_tp.c:26-31 with double2int, misc.c:22)
    int64_t normal_trunc(double sigma)
    {
        const double u0 = ((double)(next64() >> 11) + 0.5) * 0x1p-53, u1 = ((double)(next64() >> 11) + 0.5) * 0x1p-53;
        return (int64_t)(cos(6.283185307179586476925286766559 * u0) * sqrt(-2.0 * log(u1)) * sigma);
    }
};

static inline uint64_t signed_mod(int64_t v, uint64_t q)
{
    const uint64_t mag = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
    const uint64_t a = mag % q;
    return (v < 0 && a) ? q - a : a;
}

extern "C" LWE_Key lwe_alloc_key(uint64_t n, uint64_t l, incNTT ntt)
{
    LWE_Key key = (LWE_Key)malloc(sizeof(*key));
    if (!key) exit(EXIT_FAILURE);
    key->s = (uint64_t **)malloc(sizeof(uint64_t *) * (l ? l : 1));
    if (!key->s) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < l; i++) key->s[i] = (uint64_t *)safe_aligned_malloc(sizeof(uint64_t) * n);
    key->n = n;
    key->l = l;
    key->ntt = ntt;
    key->sigma = 0;
    return key;
}

extern "C" LWE lwe_alloc_sample(uint64_t n, uint64_t l, incNTT ntt)
{
    LWE c = (LWE)malloc(sizeof(*c));
    if (!c) exit(EXIT_FAILURE);
    c->a = (uint64_t **)malloc(sizeof(uint64_t *) * (l ? l : 1));
    if (!c->a) exit(EXIT_FAILURE);
    for (uint64_t i = 0; i < l; i++) c->a[i] = (uint64_t *)safe_aligned_malloc(sizeof(uint64_t) * n);
    c->b = (uint64_t *)safe_aligned_malloc(sizeof(uint64_t) * (l ? l : 1));
    c->n = n;
    c->l = l;
    c->ntt = ntt;
    return c;
}

extern "C" void free_lwe_sample(LWE c)
{
    if (!c) return;
    for (uint64_t i = 0; i < c->l; i++) free(c->a[i]);
    free(c->a);
    free(c->b);
    free(c);
}

extern "C" void free_lwe_key(LWE_Key key)
{
    if (!key) return;
    for (uint64_t i = 0; i < key->l; i++) free(key->s[i]);
    free(key->s);
    free(key);
}

// Key coefficients: rounded normal of deviation sec_sigma, the same integer in every limb.
extern "C" LWE_Key lwe_new_key(uint64_t n, uint64_t l, incNTT ntt, double sec_sigma, double err_sigma)
{
    VH_ENTER;
    VH_REQUIRE(ntt && l <= ntt->l);
    LWE_Key key = lwe_alloc_key(n, l, ntt);
    vh_host_stream rs;
    for (uint64_t i = 0; i < n; i++) {
        const int64_t s = rs.normal_trunc(sec_sigma);
        for (uint64_t j = 0; j < l; j++) key->s[j][i] = signed_mod(s, ntt->ntt[j]->q);
    }
    key->sigma = err_sigma;
    return key;
}

// Sparse ternary key of Hamming weight h, alternating +1 / -1 (misc.c:32-62: h/2 of each sign).
extern "C" LWE_Key lwe_new_sparse_ternary_key(uint64_t n, uint64_t l, incNTT ntt, uint64_t h, double err_sigma)
{
    VH_ENTER;
    VH_REQUIRE(ntt && l <= ntt->l && h <= n);
    LWE_Key key = lwe_alloc_key(n, l, ntt);
    std::vector<int> tern(n, 0);
    vh_host_stream rs;
    uint64_t placed = 0;
    int sign = 1;
    while (placed < h) {
        const uint64_t k = rs.residue(n);
        if (tern[k]) continue;
        tern[k] = sign;
        sign = -sign;
        placed++;
    }
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t j = 0; j < l; j++) {
            const uint64_t q = ntt->ntt[j]->q;
            key->s[j][i] = tern[i] == 1 ? 1 : (tern[i] == -1 ? q - 1 : 0);
        }
    key->sigma = err_sigma;
    return key;
}

// b = e + <a, s> (+ m) per limb, a uniform (lwe.c:97-123); one error value for every limb.
extern "C" void lwe_sample(LWE c, uint64_t *m, LWE_Key key)
{
    VH_ENTER;
    VH_REQUIRE(c && key && c->n == key->n && c->l == key->l);
    vh_host_stream rs;
    const int64_t e = rs.normal_trunc(key->sigma);
    for (uint64_t i = 0; i < key->l; i++) {
        const uint64_t q = key->ntt->ntt[i]->q;
        u128 acc = 0;
        for (uint64_t j = 0; j < key->n; j++) {
            c->a[i][j] = rs.residue(q);
            acc += (u128)c->a[i][j] * key->s[i][j] % q;
        }
        uint64_t b = (uint64_t)((acc + signed_mod(e, q)) % q);
        if (m) b = add_modq(b, m[i] % q, q);
        c->b[i] = b;
    }
}

extern "C" LWE lwe_new_sample(uint64_t *m, LWE_Key key)
{
    LWE c = lwe_alloc_sample(key->n, key->l, key->ntt);
    lwe_sample(c, m, key);
    return c;
}

extern "C" LWE lwe_new_trivial_sample(uint64_t *m, uint64_t n, uint64_t l, incNTT ntt)
{
    LWE c = lwe_alloc_sample(n, l, ntt);
    for (uint64_t i = 0; i < l; i++) {
        memset(c->a[i], 0, sizeof(uint64_t) * n);
        c->b[i] = m ? m[i] : 0;
    }
    return c;
}

// out[i] = b - <a, s> mod q_i (lwe.c:138-153)
extern "C" void lwe_phase(uint64_t *out, LWE c, LWE_Key key)
{
    VH_REQUIRE(c && key && out && c->n == key->n && c->l <= key->l);
    for (uint64_t i = 0; i < c->l; i++) {
        const uint64_t q = key->ntt->ntt[i]->q;
        u128 acc = 0;
        for (uint64_t j = 0; j < key->n; j++) acc += (u128)(c->a[i][j] % q) * key->s[i][j] % q;
        out[i] = sub_modq(c->b[i] % q, (uint64_t)(acc % q), q);
    }
}

// out -= in (lwe.c:155-164)
extern "C" void lwe_subto(LWE out, LWE in)
{
    VH_REQUIRE(out && in && out->n == in->n && out->l == in->l);
    for (uint64_t i = 0; i < out->l; i++) {
        const uint64_t q = out->ntt->ntt[i]->q;
        for (uint64_t j = 0; j < out->n; j++) out->a[i][j] = sub_modq(out->a[i][j] % q, in->a[i][j] % q, q);
        out->b[i] = sub_modq(out->b[i] % q, in->b[i] % q, q);
    }
}

// ---------------------------------------------------------------------------
// Rotation amounts of a CGGI16 blind rotation (cggi16.py:54-86) for a batch of LWE
// samples: with q the product of the sample's primes and b, a_i recomposed by CRT,
//   rot_tv = (2N - round((b + q // (4 torus_base)) * 2N / q) mod 2N) mod 2N,
//   steps[i] = round(a_i * 2N / q) mod 2N,
// where the reference divides with Python's correctly rounded true division and rounds
// half to even.  With v * 2N = f q + r (q odd, so never an exact tie), the double rounds
// to f + (2r > q) unless f + r/q lies within the double's rounding error of the half;
// such an entry (about one in 2^37 for N = 1024) is marked UINT64_MAX for the caller
// to compute exactly.  Needs (q + q/4) * 2N < 2^128; returns the number of marked
// entries, or -1 when q is too large.
// ---------------------------------------------------------------------------
static int clz128(u128 x)
{
    const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    return hi ? __builtin_clzll(hi) : 64 + (lo ? __builtin_clzll(lo) : 64);
}

static uint64_t round_2n_over_q(u128 v, u128 q, uint64_t N, bool *ambiguous)
{
    const u128 X = v * (u128)(2 * N);
    const u128 f = X / q, r = X % q;
    const u128 twor = r << 1;
    const u128 gap = twor > q ? twor - q : q - twor;
    // A double of magnitude below 2^e carries a rounding error up to 2^(e-53); the float quotient
    // can round across the half only when |frac - 1/2| is within that, i.e. |2r - q| <= q 2^(e-52).
    // Such samples are marked (margin 2^3) and left to an exact restatement of the float path.
    const int e = f ? 128 - clz128(f) : 0;
    if (e >= 48 || gap <= (q >> (49 - e)) + 1) { *ambiguous = true; return UINT64_MAX; }
    return (uint64_t)((f + (twor > q ? 1 : 0)) % (u128)(2 * N));
}

extern "C" int64_t vfhe_hip_cggi_rotations(LWE *in, uint64_t count, uint64_t N, uint64_t torus_base,
                                           uint64_t *rot_tv, uint64_t *steps)
{
    VH_REQUIRE(in && rot_tv && steps && torus_base > 0 && N > 0);
    for (uint64_t k = 0; k < count; k++) VH_REQUIRE(in[k] && in[k]->l >= 1 && in[k]->ntt && in[k]->l <= in[k]->ntt->l);
    // samples are independent: split over host threads (exact integer work, any order)
    auto range = [&](uint64_t k0, uint64_t k1) -> int64_t {
        int64_t marked = 0;
        for (uint64_t k = k0; k < k1; k++) {
            LWE c = in[k];
            // q and the CRT basis of the sample's limbs
            u128 q = 1;
            int bits = 0;
            for (uint64_t j = 0; j < c->l; j++) {
                const uint64_t p = c->ntt->ntt[j]->q;
                bits += 64 - __builtin_clzll(p);
                // (value + q/4) * 2N must stay below 2^128
                if (bits + 1 + (64 - __builtin_clzll(2 * N)) > 127) return -1;
                q *= p;
            }
            // Garner's mixed-radix recomposition: x = sum_j c_j R_j, R_j = q_0 ... q_{j-1},
            // c_j = (x_j - (sum_{i<j} c_i R_i mod q_j)) * R_j^-1 mod q_j; every product fits 128 bits
            std::vector<uint64_t> qs(c->l), rinv(c->l);
            std::vector<u128> R(c->l);
            for (uint64_t j = 0; j < c->l; j++) qs[j] = c->ntt->ntt[j]->q;
            for (uint64_t j = 0; j < c->l; j++) {
                R[j] = j ? R[j - 1] * qs[j - 1] : 1;
                rinv[j] = inverse_mod((uint64_t)(R[j] % qs[j]), qs[j]);
            }
            std::vector<uint64_t> cj(c->l);
            auto crt = [&](auto limb) -> u128 {
                u128 x = 0;
                for (uint64_t j = 0; j < c->l; j++) {
                    const uint64_t qj = qs[j];
                    u128 vmod = 0, w = 1;
                    for (uint64_t i = 0; i < j; i++) {
                        vmod = (vmod + (u128)cj[i] * (uint64_t)w) % qj;
                        w = w * qs[i] % qj;
                    }
                    const uint64_t d = (uint64_t)(((u128)(limb(j) % qj) + qj - (uint64_t)vmod) % qj);
                    cj[j] = (uint64_t)((u128)d * rinv[j] % qj);
                    x += (u128)cj[j] * R[j];
                }
                return x;
            };
            bool amb = false;
            const u128 b = crt([&](uint64_t j) { return c->b[j]; });
            const u128 off = q / (4 * torus_base);
            const uint64_t rb = round_2n_over_q(b + off, q, N, &amb);
            rot_tv[k] = amb ? UINT64_MAX : (2 * N - rb) % (2 * N);
            marked += amb;
            for (uint64_t i = 0; i < c->n; i++) {
                bool ai = false;
                const u128 a = crt([&](uint64_t j) { return c->a[j][i]; });
                steps[k * c->n + i] = round_2n_over_q(a, q, N, &ai);
                marked += ai;
            }
        }
        return marked;
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint64_t nt = std::min<uint64_t>(hw, (count + 31) / 32);
    if (nt <= 1) return range(0, count);
    std::vector<int64_t> res(nt, 0);
    std::vector<std::thread> th;
    for (uint64_t w = 0; w < nt; w++)
        th.emplace_back([&, w] { res[w] = range(count * w / nt, count * (w + 1) / nt); });
    for (auto &x : th) x.join();
    int64_t marked = 0;
    for (int64_t r : res) {
        if (r < 0) return -1;
        marked += r;
    }
    return marked;
}

// ---------------------------------------------------------------------------
// Sample extraction (mlwe.c:40-68): the LWE sample of coefficient idx of a
// coefficient-form MLWE sample, over the active limbs of a[0] (limb j of the
// LWE = the j-th active limb).  The sample's rows are read back to the host.
// ---------------------------------------------------------------------------
extern "C" LWE mlwe_extract_LWE(RNSc_MLWE in, uint64_t idx)
{
    VH_ENTER;
    const uint64_t N = in->a[0]->ntt->N, r = in->r, mask = in->a[0]->rns_mask;
    VH_REQUIRE(idx < N);
    std::vector<int> act;
    for (uint64_t i = 0; i < in->a[0]->ntt->l; i++) if (ACTIVE(mask, i)) act.push_back((int)i);
    LWE res = lwe_alloc_sample(r * N, act.size(), in->a[0]->ntt);
    for (auto p : comps(in)) polynomial_sync_to_host(p);
    for (size_t j = 0; j < act.size(); j++) {
        const int g = act[j];
        const uint64_t q = in->a[0]->ntt->ntt[g]->q;
        for (uint64_t k = 0; k < r; k++) {
            const uint64_t *a = in->a[k]->coeffs[g];
            for (uint64_t i = 0; i <= idx; i++) res->a[j][k * N + i] = a[idx - i];
            for (uint64_t i = idx + 1; i < N; i++) res->a[j][k * N + i] = negate_modq(a[N + idx - i], q);
        }
        res->b[j] = in->b->coeffs[g][idx];
    }
    return res;
}

// ---------------------------------------------------------------------------
// Trace and packing (mlwe.c:328-445, 600-657): sequences of automorphisms with
// key-switching, additions and monomial products on the device.
// ---------------------------------------------------------------------------
extern "C" void mlwe_partial_trace(RNSc_MLWE out, RNSc_MLWE in, uint64_t *gens, RNS_MLWE ***ksks, uint64_t size,
                                   uint64_t lvl)
{
    VH_ENTER;
    RNSc_MLWE acc = alloc_sample_raw(out->a[0]->ntt->N, out->r, out->a[0]->rns_mask, out->a[0]->ntt);
    mlwe_copy_RNS_sample(acc, in);
    for (uint64_t i = 0; i < size; i++) {
        mlwe_automorphism_RNSc_GHS(out, acc, gens[i], ksks[i], lvl);
        mlwe_add_RNSc_sample(acc, acc, out);
    }
    mlwe_copy_RNS_sample(out, acc);
    free_mlwe_RNS_sample(acc);
}

extern "C" void mlwe_trace(RNSc_MLWE out, RNSc_MLWE in, RNS_MLWE ***ksks, uint64_t lvl)
{
    const uint64_t logN = (uint64_t)ilog2(in->a[0]->ntt->N);
    std::vector<uint64_t> gens(logN);
    for (uint64_t i = 1; i <= logN; i++) gens[i - 1] = (1ULL << (logN - i + 1)) + 1;
    mlwe_partial_trace(out, in, gens.data(), ksks, logN, lvl);
}

// Packs `size` LWE samples (dimension n = in[k]->n, limbs = the active limbs of
// out) into one MLWE sample: out = -sum_i ksk[i] (x) a_i(X) + sum_k b_k X^k, with
// a_i(X) = sum_k in[k]->a[.][i] X^k and the gadget product of gadget.c:22-37,
// all n gadget products in one batched ModUp + key inner product.
extern "C" void mlwe_full_packing_keyswitch(RNS_MLWE out, LWE *in, uint64_t size, RNS_MLWE **ksk, uint64_t lvl)
{
    (void)lvl;
    VH_ENTER;
    VH_REQUIRE(out && in && ksk && size > 0);
    incNTT t = out->b->ntt;
    const uint64_t N = t->N, n = in[0]->n, lwe_l = in[0]->l;
    const uint64_t target = out->b->rns_mask, ext = ksk[0][0]->b->rns_mask, divide = ext & ~target;
    VH_REQUIRE(size <= N);
    std::vector<int> act;
    for (uint64_t i = 0; i < t->l; i++) if (ACTIVE(target, i)) act.push_back((int)i);
    VH_REQUIRE(lwe_l <= act.size());
    // a_i(X) for every LWE coordinate i, coefficient form on the target limbs
    std::vector<RNS_Polynomial> polys(n);
    for (uint64_t i = 0; i < n; i++) {
        polys[i] = new_poly(N, target, t, false, true);
        for (int g : act) memset(polys[i]->coeffs[g], 0, N * 8);
        for (uint64_t limb = 0; limb < lwe_l; limb++)
            for (uint64_t k = 0; k < size; k++) polys[i]->coeffs[act[limb]][k] = in[k]->a[limb][i];
        VP(polys[i])->host_dirty = 1;
    }
    std::vector<vh_gsrc> srcs;
    for (uint64_t i = 0; i < n; i++) srcs.push_back({polys[i], ksk[i]});
    gadget_batch({out}, {srcs}, false);        // out = -sum_i <lift(a_i), ksk[i]>, NTT form on ext
    RNS_Polynomial body = new_poly(N, target, t, false, true);
    for (int g : act) memset(body->coeffs[g], 0, N * 8);
    for (uint64_t limb = 0; limb < lwe_l; limb++)
        for (uint64_t k = 0; k < size; k++) body->coeffs[act[limb]][k] = in[k]->b[limb];
    VP(body)->host_dirty = 1;
    dev_ready(body);
    mlwe_ntt(out, out, 1);
    if (divide) run_division(comps(out), divide, true);
    binop(out->b, out->b, body, VH_OP_ADD);
    mlwe_ntt(out, out, 0);
    for (auto p : polys) free_RNS_polynomial(p);
    free_RNS_polynomial(body);
}

static void packing_rec(std::vector<RNSc_MLWE> vec, uint64_t ell, RNS_MLWE ***ksks, uint64_t lvl)
{
    if (ell == 0) return;
    const uint64_t half = 1ULL << (ell - 1);
    std::vector<RNSc_MLWE> even(half), odd(half);
    for (uint64_t i = 0; i < half; i++) { even[i] = vec[2 * i]; odd[i] = vec[2 * i + 1]; }
    packing_rec(even, ell - 1, ksks, lvl);
    packing_rec(odd, ell - 1, ksks, lvl);
    RNSc_MLWE c = even[0];
    const uint64_t N = vec[0]->b->ntt->N, r = vec[0]->r;
    incNTT t = vec[0]->b->ntt;
    RNSc_MLWE tmp = alloc_sample_raw(N, r, vec[0]->b->rns_mask, t);
    RNSc_MLWE tmp2 = alloc_sample_raw(N, r, ksks[ell - 1][0][0]->b->rns_mask, t);
    mlwe_RNSc_mul_by_xai(tmp, odd[0], N >> ell);            // tmp = odd * X^(N / 2^ell)
    mlwe_sub_RNSc_sample(c, even[0], tmp);                  // c = even - tmp
    mlwe_automorphism_RNSc_GHS(tmp2, c, (1ULL << ell) + 1, ksks[ell - 1], lvl);
    for (auto p : comps(tmp)) scale_op(p, p, nullptr, 2, VH_OP_SCALE);   // tmp = 2 tmp
    mlwe_add_RNSc_sample(c, c, tmp2);
    mlwe_add_RNSc_sample(c, c, tmp);
    free_mlwe_RNS_sample(tmp);
    free_mlwe_RNS_sample(tmp2);
}

extern "C" void mlwe_full_packing_keyswitch_scaled(RNSc_MLWE *vec, uint64_t ell, RNS_MLWE ***ksks, uint64_t lvl)
{
    VH_ENTER;
    if (ell == 0) return;
    packing_rec(std::vector<RNSc_MLWE>(vec, vec + (1ULL << ell)), ell, ksks, lvl);
}
