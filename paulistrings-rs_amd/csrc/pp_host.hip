// pp_host.hip -- the host side of the MI355X rotation-layer engine and its C ABI (include/pp_hip.h).
//
// A pp_sum is resident in HBM.  Its device bucketing (GF(2) hash rows of the reference, at a bucket count chosen
// here for LDS-sized pairs) is independent of the reference's bucket count, which this file tracks alongside
// exactly as the reference's layer loop moves it (engine/driver.rs:186-258, :319-592), because that count fixes
// the canonical output order and the order in which TopN keeps tied terms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pp_hip.h"
#include "pp_class.h"
#include "pp_kernels.h"
#include "pp_learned.h"
#include "pp_membench.h"
#include "pp_topn.h"

namespace {

using namespace pp;

constexpr int kIndexBits = 31;                    // pauli_sum/hash.rs:7
constexpr size_t kMinTermsPerTask = 64;           // pauli_sum/storage.rs:48
constexpr size_t kCoarsenBelow = 8;               // pauli_sum/storage.rs:44
constexpr uint32_t kRedrawAttempts = 4096;        // pauli_sum/hash.rs:18
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;  // rng.rs:5
constexpr size_t kRunLayers = 8;                  // engine/bucketed/rotation_layer.rs:508
constexpr int kRunDimension = 3;                  // engine/bucketed/rotation_layer.rs:511
constexpr uint32_t kMaxPairRows = 1024;           // the largest pair a join's redo launch stages
// Device counters of a sum: [0] rows, [1] longest bucket, [2] join overflow, [3] offsets total, [4] redo pairs.
constexpr int kStatSlots = 8;

// Defaults of the per-context tuning, each overridable from the environment.
// The join's grid (PP_JOIN_GROUPS_PER_CU): one workgroup per pair by default (measured fastest on the saturated
// layer); 1..32 selects a persistent grid of at most that many resident workgroups per CU, 0 as many as fit.
uint32_t default_groups_per_cu() {
    const char* e = std::getenv("PP_JOIN_GROUPS_PER_CU");
    if (!e || std::strcmp(e, "pairs") == 0) return kJoinOnePairPerGroup;
    const long n = std::strtol(e, nullptr, 10);
    return n >= 0 && n <= 32 ? (uint32_t)n : kJoinOnePairPerGroup;
}

// Anticommuting rows one join workgroup stages in its first launch (PP_JOIN_STAGE_ROWS, 64..1024; 0 or unset: the
// largest pair, so that no pair is redone; fewer rows raise the resident workgroups per CU, which measured no faster
// on the saturated layer than the redo launch costs).
uint32_t default_stage_rows() {
    const char* e = std::getenv("PP_JOIN_STAGE_ROWS");
    const long t = e ? std::strtol(e, nullptr, 10) : 0;
    return (uint32_t)(t >= 64 && t <= 1024 ? t : 0);
}

// Threads per join workgroup (PP_JOIN_BLOCK, 256 or 512; default 256): the rows of one pass over a pair.
uint32_t default_join_block() {
    const char* e = std::getenv("PP_JOIN_BLOCK");
    return e && std::strtol(e, nullptr, 10) == 512 ? 512u : 256u;
}

// Layers in place when the segments have room (PP_JOIN_IN_PLACE=0 turns it off; default on).
uint32_t default_join_in_place() {
    const char* e = std::getenv("PP_JOIN_IN_PLACE");
    return e && std::strcmp(e, "0") == 0 ? 0u : 1u;
}

// Rows per device bucket (PP_DEVICE_BUCKET_LEN, 16..512; default 256).
uint32_t default_bucket_target() {
    const char* e = std::getenv("PP_DEVICE_BUCKET_LEN");
    const long t = e ? std::strtol(e, nullptr, 10) : 0;
    return (uint32_t)(t >= 16 && t <= 512 ? t : 256);
}

bool valid_tuning(const pp_tuning& t) {
    return t.bucket_target >= 16 && t.bucket_target <= 512 && (t.stage_rows == 0 || (t.stage_rows >= 64 && t.stage_rows <= 1024)) &&
           (t.groups_per_cu <= 32 || t.groups_per_cu == kJoinOnePairPerGroup) &&
           (t.join_block == 256 || t.join_block == 512) && t.join_in_place <= 1;
}

// The staging rows of a join whose largest pair is lcap rows: the tuning's (or lcap), a multiple of 64, at most lcap.
uint32_t join_stage_rows(uint32_t tuned, uint32_t lcap) {
    uint32_t rows = tuned ? tuned : lcap;
    rows = (rows + 63) & ~63u;
    return std::min(rows, lcap);
}

// ---- the reference's hash rows and bucket-count policy ---------------------------------------------------------

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Host threads for a host-side pass over `items` elements: up to 16, one per 2^16 items.
unsigned host_threads(uint64_t items) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(hw, items >> 16));
}

// fn(t, T) on T host threads (the calling thread runs t = 0).
template <class F>
void host_parallel(unsigned T, F fn) {
    std::vector<std::thread> pool;
    pool.reserve(T > 0 ? T - 1 : 0);
    for (unsigned t = 1; t < T; ++t) pool.emplace_back(fn, t, T);
    fn(0u, std::max(T, 1u));
    for (auto& th : pool) th.join();
}

// [lo, hi) of part t of T equal parts of n.
inline void part_range(uint64_t n, unsigned t, unsigned T, uint64_t& lo, uint64_t& hi) {
    lo = n * t / T;
    hi = n * (t + 1) / T;
}

// Stable counting sort of the indices 0..n-1 by key[i] < K on T threads: order[] lists the indices bucket by bucket,
// each bucket in index order (each thread scatters a contiguous index range after the threads before it), and
// start[k] .. start[k + 1] is bucket k's range.
void parallel_counting_sort(const uint32_t* key, uint64_t n, uint32_t K, unsigned T, std::vector<uint64_t>& order,
                            std::vector<uint64_t>& start) {
    std::vector<uint64_t> cnt((size_t)T * K, 0);
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(n, t, TT, lo, hi);
        uint64_t* c = &cnt[(size_t)t * K];
        for (uint64_t i = lo; i < hi; ++i) ++c[key[i]];
    });
    start.assign((size_t)K + 1, 0);
    uint64_t acc = 0;
    for (uint32_t k = 0; k < K; ++k) {
        start[k] = acc;
        for (unsigned t = 0; t < T; ++t) {
            const uint64_t c = cnt[(size_t)t * K + k];
            cnt[(size_t)t * K + k] = acc;  // thread t's first slot in bucket k
            acc += c;
        }
    }
    start[K] = acc;
    order.resize(n);
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(n, t, TT, lo, hi);
        uint64_t* c = &cnt[(size_t)t * K];
        for (uint64_t i = lo; i < hi; ++i) order[c[key[i]]++] = i;
    });
}

uint64_t live_mask(uint32_t num_qubits, uint32_t word) {
    const uint32_t first = 64 * word;
    if (num_qubits >= first + 64) return ~0ull;
    if (num_qubits <= first) return 0;
    return (1ull << (num_qubits - first)) - 1ull;
}

// Gf2Hash::new rows: splitmix64 draws, redrawn while dependent (pauli_sum/hash.rs:22-130).  Returns the row
// count; rows_x/rows_z receive count * W words.
int hash_rows(uint32_t num_qubits, uint64_t seed, uint32_t W, uint64_t* rows_x, uint64_t* rows_z) {
    const int rows = (int)std::min<uint32_t>(2 * num_qubits, kIndexBits);
    std::vector<std::vector<uint64_t>> basis;  // reduced rows, each with a distinct leading column
    std::vector<int> leads;
    auto leading = [&](const std::vector<uint64_t>& r) {
        for (size_t w = 0; w < r.size(); ++w)
            if (r[w]) return (int)(64 * w + __builtin_clzll(r[w]));
        return -1;
    };
    auto insert = [&](std::vector<uint64_t> r) {
        for (;;) {
            const int l = leading(r);
            if (l < 0) return false;
            auto it = std::find(leads.begin(), leads.end(), l);
            if (it == leads.end()) {
                leads.push_back(l);
                basis.push_back(r);
                return true;
            }
            const auto& p = basis[it - leads.begin()];
            for (size_t w = 0; w < r.size(); ++w) r[w] ^= p[w];
        }
    };
    for (int r = 0; r < rows; ++r) {
        std::vector<uint64_t> key(2 * W);
        for (uint32_t attempt = 0;; ++attempt) {
            bool zero = true;
            for (uint32_t w = 0; w < W; ++w) {
                const uint64_t m = live_mask(num_qubits, w);
                for (uint32_t half = 0; half < 2; ++half) {
                    const uint64_t pos = ((uint64_t)attempt << 32) | ((uint64_t)r << 16) | ((uint64_t)w << 1) | half;
                    key[half * W + w] = mix64(seed + kGamma * (pos + 1)) & m;
                    zero = zero && key[half * W + w] == 0;
                }
            }
            const bool accepted = (int)basis.size() < rows ? insert(key) : !zero;
            if (accepted || attempt == kRedrawAttempts) break;
        }
        for (uint32_t w = 0; w < W; ++w) {
            rows_x[r * W + w] = key[w];
            rows_z[r * W + w] = key[W + w];
        }
    }
    return rows;
}

int desired_bits(size_t len, size_t target, size_t min_buckets) {
    int floor_bits = 0;
    if (len >= min_buckets * kMinTermsPerTask)
        while (((size_t)1 << floor_bits) < min_buckets && floor_bits < kIndexBits) ++floor_bits;
    int b = floor_bits;
    while (b < kIndexBits && len > target * ((size_t)1 << b)) ++b;
    return b;
}

int rebucketed_bits(size_t len, int bits, size_t target, size_t min_buckets) {
    const int want = desired_bits(len, target, min_buckets);
    const bool sparse = len * kCoarsenBelow < target * ((size_t)1 << bits);
    return (want > bits || sparse) ? want : bits;
}

// sin_cos with quarter turns snapped to exact 0 / +-1 (channel/rotation.rs:70-78), from the host libm.  Rust's
// f64::sin_cos reaches glibc as one sincos call on x86_64 Linux (LLVM fuses the sin and cos intrinsics of one
// argument); glibc's sincos and its separate sin and cos differ in the last bit on rare arguments, so sincos is
// called by name.
void snapped_sin_cos(double theta, double* s, double* c) {
    const double snap = 1e-15;
    double sv, cv;
    ::sincos(theta, &sv, &cv);
    if (std::fabs(sv) < snap) {
        *s = 0.0;
        *c = std::copysign(1.0, cv);
    } else if (std::fabs(cv) < snap) {
        *s = std::copysign(1.0, sv);
        *c = 0.0;
    } else {
        *s = sv;
        *c = cv;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// Context and sum
// ---------------------------------------------------------------------------------------------------------------

// What one captured out-of-place layer (pp_apply_rotation_to inside a graph capture) baked into the graph: the two
// sums, the column sets and device buffers its kernels read and write, and the input bounds its launch was sized for.
// A replay is refused unless every one of them is still what the sums hold.
struct CapturedJoin {
    pp_sum* src;
    pp_sum* dst;
    int src_cur, dst_cur, device_bits;
    const void* src_ptrs[5];  // x, z, c, bucket start, bucket len
    const void* dst_ptrs[7];  // x, z, c, bucket start, bucket len, counters, redo queue
    uint64_t dst_cap;         // rows of dst's written column set
    uint64_t src_len;         // rows src held at capture: the output columns and staging were sized from them
    uint32_t src_max_bucket;
};

struct pp_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    bool capturing = false;  // between pp_graph_capture_begin and _end: no allocation, no synchronization
    std::vector<CapturedJoin> capture_log;  // the layers of the open capture
    std::vector<pp_graph*> graphs;          // live graphs of this context (pp_sum_free refuses their sums)
    pp_tuning tuning{};      // sizes of the device path (pp_set_tuning, pp_autotune)
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    uint64_t prof_rows_in = 0;
    uint64_t prof_rows_out = 0;  // output rows of the profiled layers the host counted (in-place layers)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> score_events;  // the learned pass's score kernel
    uint64_t prof_scored = 0;
    unsigned long long* d_prof = nullptr;  // [0]: output rows of the profiled joins
    void* stage[2] = {nullptr, nullptr};   // pinned host chunks of the bulk transfers, allocated on first use
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    void* class_pinned = nullptr;  // a class pass's program on its way up and its counters on their way down
    hipEvent_t lag_ev[2] = {nullptr, nullptr};  // the counters of a layer queued ahead of them have arrived
};

struct pp_graph {
    pp_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<CapturedJoin> layers;  // what the replay reads and writes, checked before every launch
};

struct Columns {
    uint64_t* x = nullptr;
    uint64_t* z = nullptr;
    void* c = nullptr;
    uint64_t cap = 0;
};

struct BucketArrays {
    uint64_t* start = nullptr;
    uint32_t* len = nullptr;
    uint32_t* rows = nullptr;  // each segment's room in rows (valid when the sum's has_room says so)
    uint64_t cap = 0;
};

enum ScratchSlot {
    kScrKeyX,
    kScrKeyZ,
    kScrPos,
    kScrBucket,
    kScrFlags,
    kScrCounts,
    kScrGathered,
    kScrRank,
    kScrDest,        // a permutation run's per-row output bucket
    kScrRunCounts,   // its per-bucket destination counts
    kScrDeaths,      // its per-step drops
    kScrSteps,       // its steps
    kScrClassCount,  // a class pass's counters, outputs per bucket and touched rows per class bucket
    kScrClassStart,  // the class buckets' scan
    kScrClassKeep,   // its untouched rows kept per bucket
    kScrClassLayers, // its program
    kScrClassSX,     // its touched rows in class-bucket order
    kScrClassSZ,
    kScrClassSC,
    kScrClassOX,     // its outputs
    kScrClassOZ,
    kScrClassOC,
    kScrClassOB,     // their buckets
    kScratchSlots
};

struct pp_sum {
    pp_context* ctx = nullptr;
    uint32_t num_qubits = 0;
    uint32_t W = 1;
    pp_coeff_kind kind = PP_COEFF_F64;
    uint64_t seed = 0;
    int max_bits = 0;
    std::vector<uint64_t> rows;  // [row][x|z][W]
    uint64_t* d_rows = nullptr;
    Columns col[2];
    BucketArrays bk[2];
    int cur = 0;
    int device_bits = 0;
    int reference_bits = 0;
    uint64_t len = 0;
    uint32_t max_bucket = 0;
    bool stale = false;  // len and max_bucket await the last layer's counters (pp_apply_rotation_to)
    // The current column set's free room: its segments' rooms (bk[cur].rows) are valid, and every row at or after
    // `tail` is free.  Layouts from an offset scan, an upload, a refinement of one, and in-place layers keep both.
    bool has_room = false;
    uint64_t tail = 0;
    // A TopN selection whose retain is left to the next join (inside pp_propagate only): len is then the kept
    // count, phys_len the rows the columns still hold, and max_bucket their longest bucket.
    bool pending_topn = false;
    uint64_t class_out_hint = 0;   // the last class pass's outputs: the next one's first guess at its staging
    uint64_t topn_seed = 0;        // the last selection's boundary prefix, which seeds the next one's first pass
    bool topn_seed_valid = false;
    bool topn_seed_score = false;  // the seed is a score key (the learned pass), not a |c|^2 one
    uint64_t phys_len = 0;
    const uint8_t* pending_flags = nullptr;
    const float* pending_score = nullptr;  // the ranking score when the selection was by score
    std::vector<uint64_t> tie_kept;  // TopN's boundary positions in rank order, the source of an asynchronous copy
    uint64_t* d_tiles = nullptr;  // scratch of the bucket-offset scan
    uint64_t tiles_cap = 0;
    unsigned long long* d_stats = nullptr;
    unsigned long long* h_stats = nullptr;  // pinned
    // Two pinned, device-visible slots of kStatSlots words for the counters of layers queued ahead of them (lagged
    // lengths), allocated on first use; the device address of the same memory; the small join's sequence number.
    unsigned long long* h_lag = nullptr;
    unsigned long long* d_lag_host = nullptr;
    unsigned long long lag_seq = 0;
    unsigned long long* d_lag_counters = nullptr;  // the small joins' counters: two sets of 4, by ring slot
    pp::TopnState* d_topn = nullptr;
    unsigned long long* d_hist = nullptr;
    uint32_t* d_redo = nullptr;  // pairs a join could not stage, for its redo launch
    uint64_t redo_cap = 0;
    float* d_score = nullptr;  // the learned pass's score per column position
    uint64_t score_cap = 0;
    // grow-only scratch of the truncation passes (boundary terms, their flags and counts), kept between layers so
    // that a pass run after every layer allocates nothing once the sizes have settled
    void* scratch[kScratchSlots] = {};
    size_t scratch_cap[kScratchSlots] = {};
    size_t coeff_bytes() const { return kind == PP_COEFF_F64 ? 8 : 4; }
};

namespace {

pp_status fail(pp_context* ctx, pp_status s, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return s;
}

// Entry points that synchronize or allocate refuse to run inside a graph capture.
#define PP_NOT_CAPTURING(ctx, what)                                                                     \
    do {                                                                                               \
        if ((ctx) && (ctx)->capturing)                                                                 \
            return fail((ctx), PP_ERR_UNSUPPORTED,                                                     \
                        std::string(what) + " is not capturable: close the capture first");           \
    } while (0)

#define PP_HIP(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail((ctx), PP_ERR_DRIVER, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

pp_status dev_alloc(pp_context* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (ctx->capturing)
        return fail(ctx, PP_ERR_UNSUPPORTED,
                    "an allocation during graph capture: run the same calls once before capturing so that every "
                    "buffer has its size");
    if (bytes == 0) bytes = 64;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, PP_ERR_OUT_OF_MEMORY,
                    "device " + std::to_string(ctx->device) + ": out of memory allocating " + std::to_string(bytes) +
                        " bytes");
    }
    return PP_OK;
}

// A column set is one allocation (x, then z, then the coefficients), and so are a layout's bucket arrays (start,
// then len, then rows): one allocation call, and one free, each.
void free_columns(Columns& c) {
    if (c.x) (void)hipFree(c.x);
    c = Columns{};
}

void free_buckets(BucketArrays& b) {
    if (b.start) (void)hipFree(b.start);
    b = BucketArrays{};
}

pp_status ensure_columns(pp_sum* s, int which, uint64_t need) {
    Columns& c = s->col[which];
    if (c.cap >= need && c.x) return PP_OK;
    free_columns(c);
    // Headroom of a quarter, and at small sizes of as much again (up to 2^20 rows): a growing sum then reallocates a
    // few times rather than every few layers (each hipFree waits for the device).  A multiple of kSegAlign, so that
    // every word column starts as aligned as the first.
    const uint64_t extra = std::max<uint64_t>(need / 4, std::min<uint64_t>(need, 1ull << 20));
    const uint64_t cap = (std::max<uint64_t>(need + extra, 1024) + (kSegAlign - 1)) & ~(uint64_t)(kSegAlign - 1);
    pp_status st;
    unsigned char* base = nullptr;
    if ((st = dev_alloc(s->ctx, (void**)&base, cap * (2 * s->W * 8 + s->coeff_bytes()))) != PP_OK) return st;
    c.x = (uint64_t*)base;
    c.z = (uint64_t*)(base + cap * s->W * 8);  // cap is a multiple of kSegAlign: every array starts 128-B aligned
    c.c = base + 2 * cap * s->W * 8;
    c.cap = cap;
    return PP_OK;
}

pp_status ensure_buckets(pp_sum* s, int which, uint64_t need) {
    BucketArrays& b = s->bk[which];
    if (b.cap >= need && b.start) return PP_OK;
    free_buckets(b);
    const uint64_t cap = std::max<uint64_t>(2 * need, 256);  // room for the next refinement's count
    pp_status st;
    unsigned char* base = nullptr;
    if ((st = dev_alloc(s->ctx, (void**)&base, cap * 16)) != PP_OK) return st;
    b.start = (uint64_t*)base;
    b.len = (uint32_t*)(base + cap * 8);
    b.rows = (uint32_t*)(base + cap * 12);
    b.cap = cap;
    return PP_OK;
}

// Dispatch on (W, coefficient type).
template <typename F>
auto dispatch(const pp_sum* s, F&& f) {
    if (s->W == 1) {
        if (s->kind == PP_COEFF_F64) return f(std::integral_constant<int, 1>{}, double{});
        return f(std::integral_constant<int, 1>{}, float{});
    }
    if (s->kind == PP_COEFF_F64) return f(std::integral_constant<int, 2>{}, double{});
    return f(std::integral_constant<int, 2>{}, float{});
}

uint32_t bucket_of_host(const pp_sum* s, const uint64_t* gx, const uint64_t* gz, int bits) {
    uint32_t b = 0;
    for (int r = 0; r < bits; ++r) {
        uint64_t f = 0;
        for (uint32_t w = 0; w < s->W; ++w)
            f ^= (gx[w] & s->rows[(2 * r) * s->W + w]) ^ (gz[w] & s->rows[(2 * r + 1) * s->W + w]);
        b |= (uint32_t)(__builtin_popcountll(f) & 1) << r;
    }
    return b;
}

pp_status read_stats(pp_sum* s) {
    pp_context* ctx = s->ctx;
    if (ctx->capturing)
        return fail(ctx, PP_ERR_UNSUPPORTED,
                    "a host read-back during graph capture: only pp_apply_rotation_to from a settled, non-pending "
                    "input is capturable");
    PP_HIP(ctx, hipMemcpyAsync(s->h_stats, s->d_stats, kStatSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               ctx->stream));
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

pp_status clear_stats(pp_sum* s) {
    PP_HIP(s->ctx, launch_fill_u64(s->ctx->stream, s->d_stats, kStatSlots, 0ull));
    return PP_OK;
}

// Scratch for the bucket-offset scan of a layout of `out_n` buckets.
pp_status ensure_tiles(pp_sum* s, uint32_t out_n) {
    const uint64_t need = std::max<uint32_t>(offsets_tiles(out_n), 1);
    if (s->tiles_cap >= need && s->d_tiles) return PP_OK;
    if (s->d_tiles) (void)hipFree(s->d_tiles);
    s->d_tiles = nullptr;
    s->tiles_cap = 0;
    pp_status st = dev_alloc(s->ctx, (void**)&s->d_tiles, need * 8);
    if (st != PP_OK) return st;
    s->tiles_cap = need;
    return PP_OK;
}

// Slot `slot` of the sum's truncation scratch with room for `bytes` (grown to 1.5x the request when it must grow).
template <typename P>
pp_status slot_scratch(pp_sum* s, ScratchSlot slot, size_t bytes, P** out) {
    if (s->scratch_cap[slot] < bytes || !s->scratch[slot]) {
        if (s->scratch[slot]) (void)hipFree(s->scratch[slot]);
        s->scratch[slot] = nullptr;
        s->scratch_cap[slot] = 0;
        const size_t grown = std::max<size_t>(256, 2 * bytes);
        pp_status st = dev_alloc(s->ctx, &s->scratch[slot], grown);
        if (st != PP_OK) return st;
        s->scratch_cap[slot] = grown;
    }
    *out = (P*)s->scratch[slot];
    return PP_OK;
}

// With clear_counters the scan also zeroes the sum's counters (before it writes the total into slot 3), in place of
// clear_stats.
// out_rows (a resident layout's segment rooms) is written with the starts; the total lands in counter slot 3.
pp_status offsets(pp_sum* s, const uint32_t* len, uint32_t num_buckets, uint32_t mode, uint32_t delta,
                  uint32_t coarse_buckets, uint32_t fold, uint64_t* out_start, bool clear_counters = false,
                  uint32_t* out_rows = nullptr) {
    pp_status st;
    if ((st = ensure_tiles(s, mode == 2 ? coarse_buckets : num_buckets)) != PP_OK) return st;
    PP_HIP(s->ctx, launch_bucket_offsets(s->ctx->stream, len, num_buckets, mode, delta, coarse_buckets, fold,
                                         out_start, s->d_stats + 3, s->d_tiles,
                                         clear_counters ? s->d_stats : nullptr, kStatSlots, out_rows));
    return PP_OK;
}

// len and max_bucket of the current layout, read back (one synchronization).  Counter slot 3 (the last offset
// scan's total, the end of the layout it made) is kept.
pp_status count_layout(pp_sum* s, int which, uint32_t num_buckets) {
    pp_status st;
    PP_HIP(s->ctx, launch_fill_u64(s->ctx->stream, s->d_stats, 3, 0ull));
    PP_HIP(s->ctx, launch_reduce_lens(s->ctx->stream, s->bk[which].len, num_buckets, s->d_stats, nullptr));
    if ((st = read_stats(s)) != PP_OK) return st;
    s->len = s->h_stats[0];
    s->max_bucket = (uint32_t)s->h_stats[1];
    return PP_OK;
}

// ---- pinned, chunked host transfers (the reference's staging of uploads and downloads, engine/gpu/staging.rs) -------

constexpr size_t kStageChunk = (size_t)32 << 20;  // bytes per pinned chunk; two alternate

pp_status ensure_staging(pp_context* ctx) {
    if (ctx->stage[0]) return PP_OK;
    for (int k = 0; k < 2; ++k) {
        if (hipHostMalloc(&ctx->stage[k], kStageChunk, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ctx->stage[k] = nullptr;
            return fail(ctx, PP_ERR_OUT_OF_MEMORY, "pinned staging allocation");
        }
        PP_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
        PP_HIP(ctx, hipEventRecord(ctx->stage_ev[k], ctx->stream));  // recorded once, so a first wait is defined
    }
    return PP_OK;
}

// The host's copy into or out of a pinned chunk, on up to 16 threads for a full chunk.
void stage_memcpy(void* dst, const void* src, size_t n) {
    const unsigned T = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(host_threads(n >> 3), 8));
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(n, t, TT, lo, hi);
        std::memcpy((char*)dst + lo, (const char*)src + lo, hi - lo);
    });
}

// host -> device through the two pinned chunks: the host fills one while the other's copy runs.  Queued on the
// context's stream; the caller synchronizes before it reads the result.
pp_status staged_h2d(pp_context* ctx, void* dev, const void* host, size_t bytes) {
    pp_status st;
    if ((st = ensure_staging(ctx)) != PP_OK) return st;
    int k = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, k ^= 1) {
        const size_t n = std::min(kStageChunk, bytes - off);
        PP_HIP(ctx, hipEventSynchronize(ctx->stage_ev[k]));  // the chunk's previous copy has drained
        stage_memcpy(ctx->stage[k], (const char*)host + off, n);
        PP_HIP(ctx, hipMemcpyAsync((char*)dev + off, ctx->stage[k], n, hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(ctx, hipEventRecord(ctx->stage_ev[k], ctx->stream));
    }
    return PP_OK;
}

// device -> host through the two pinned chunks: chunk i + 1's copy runs while the host drains chunk i.  Returns once
// every byte is in `host`.
pp_status staged_d2h(pp_context* ctx, void* host, const void* dev, size_t bytes) {
    pp_status st;
    if ((st = ensure_staging(ctx)) != PP_OK) return st;
    const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
    auto queue = [&](size_t i) -> pp_status {
        const size_t off = i * kStageChunk, n = std::min(kStageChunk, bytes - off);
        PP_HIP(ctx, hipMemcpyAsync(ctx->stage[i & 1], (const char*)dev + off, n, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(ctx, hipEventRecord(ctx->stage_ev[i & 1], ctx->stream));
        return PP_OK;
    };
    if (chunks && (st = queue(0)) != PP_OK) return st;
    for (size_t i = 0; i < chunks; ++i) {
        if (i + 1 < chunks && (st = queue(i + 1)) != PP_OK) return st;  // its chunk was drained one step ago
        PP_HIP(ctx, hipEventSynchronize(ctx->stage_ev[i & 1]));
        const size_t off = i * kStageChunk;
        stage_memcpy((char*)host + off, ctx->stage[i & 1], std::min(kStageChunk, bytes - off));
    }
    return PP_OK;
}

// ---- device rebucketing ------------------------------------------------------------------------------------------

pp_status refine_once(pp_sum* s) {
    const int in = s->cur, out = 1 - s->cur;
    const uint32_t B = 1u << s->device_bits;
    pp_status st;
    if ((st = ensure_columns(s, out, s->col[in].cap)) != PP_OK) return st;
    if ((st = ensure_buckets(s, out, 2ull * B)) != PP_OK) return st;
    if ((st = clear_stats(s)) != PP_OK) return st;
    hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        RefineArgs<W, T> a{};
        a.x = s->col[in].x;
        a.z = s->col[in].z;
        a.c = (const T*)s->col[in].c;
        a.cap = s->col[in].cap;
        a.start = s->bk[in].start;
        a.len = s->bk[in].len;
        a.ox = s->col[out].x;
        a.oz = s->col[out].z;
        a.oc = (T*)s->col[out].c;
        a.ocap = s->col[out].cap;
        a.ostart = s->bk[out].start;
        a.olen = s->bk[out].len;
        a.num_buckets = B;
        a.row = s->device_bits;
        a.rows = s->d_rows;
        a.stats = s->d_stats;
        a.bcap = s->has_room ? s->bk[in].rows : nullptr;  // the halves share the segment's room (the tail stays)
        a.obcap = s->has_room ? s->bk[out].rows : nullptr;
        return launch_refine<W, T>(s->ctx->stream, a);
    });
    PP_HIP(s->ctx, e);
    if ((st = count_layout(s, out, 2 * B)) != PP_OK) return st;
    s->device_bits += 1;
    s->cur = out;
    return PP_OK;
}

pp_status coarsen_to(pp_sum* s, int bits) {
    const int in = s->cur, out = 1 - s->cur;
    const uint32_t B = 1u << s->device_bits, CB = 1u << bits;
    pp_status st;
    if ((st = ensure_columns(s, out, s->len + (uint64_t)kSegAlign * CB)) != PP_OK) return st;
    if ((st = ensure_buckets(s, out, CB)) != PP_OK) return st;
    if ((st = offsets(s, s->bk[in].len, B, 2, 0, CB, B / CB, s->bk[out].start, false, s->bk[out].rows)) != PP_OK)
        return st;
    hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        CoarsenArgs<W, T> a{};
        a.x = s->col[in].x;
        a.z = s->col[in].z;
        a.c = (const T*)s->col[in].c;
        a.cap = s->col[in].cap;
        a.start = s->bk[in].start;
        a.len = s->bk[in].len;
        a.ox = s->col[out].x;
        a.oz = s->col[out].z;
        a.oc = (T*)s->col[out].c;
        a.ocap = s->col[out].cap;
        a.ostart = s->bk[out].start;
        a.olen = s->bk[out].len;
        a.num_buckets = B;
        a.coarse_buckets = CB;
        a.stats = s->d_stats;
        return launch_coarsen<W, T>(s->ctx->stream, a);
    });
    PP_HIP(s->ctx, e);
    if ((st = count_layout(s, out, CB)) != PP_OK) return st;
    s->tail = s->h_stats[3];
    s->has_room = true;
    s->device_bits = bits;
    s->cur = out;
    return PP_OK;
}

pp_status materialize_topn(pp_sum* s);  // pp_topn_impl.inc

// The device's own bucket count: about the context's bucket_target rows per bucket, with a factor-8 hysteresis, and
// fine enough that the largest pair fits one workgroup's LDS.
pp_status settle_device_bits(pp_sum* s) {
    const uint64_t target = s->ctx->tuning.bucket_target;
    int want = 0;
    while (want < s->max_bits && s->len > target * (1ull << want)) ++want;
    pp_status st;
    // refine and coarsen copy every row they hold: a pending selection is applied first
    const bool coarsen = s->len * 8 < target * (1ull << s->device_bits) && want < s->device_bits;
    if ((want > s->device_bits || coarsen || 2ull * s->max_bucket > kMaxPairRows) &&
        (st = materialize_topn(s)) != PP_OK)
        return st;
    if (want > s->device_bits) {
        while (s->device_bits < want)
            if ((st = refine_once(s)) != PP_OK) return st;
    } else if (coarsen) {
        if ((st = coarsen_to(s, want)) != PP_OK) return st;
    }
    while (2ull * s->max_bucket > kMaxPairRows) {
        if (s->device_bits >= s->max_bits)
            return fail(s->ctx, PP_ERR_UNSUPPORTED,
                        "a bucket of " + std::to_string(s->max_bucket) + " terms cannot be split further");
        if ((st = refine_once(s)) != PP_OK) return st;
    }
    return PP_OK;
}

// ---- one prepared channel ------------------------------------------------------------------------------------------

// The device's reading of a Prepared (channel/prepared.rs:343-350): which kernel applies it, and the run planner's
// questions about it (max_fanout engine/bucketed.rs:628-641, PermutationStep::of permutation_run.rs:53-121,
// RotationLayer::of rotation_layer.rs:101-140).
struct Prepared {
    pp_prepared table{};          // the channel's prepared form as the seam carries it
    bool local = true;            // Prepared::Local (support <= 2) rather than Prepared::Rotation
    bool rotation = false;        // the pair join applies it from (g, cos, sin): a rotation's own table or a wider one
    uint64_t gx[4] = {}, gz[4] = {};
    double sin = 0, cos = 1;
    int weight = 0;
    bool key_preserving = false;  // one identity entry: an in-place rescale by amplitude[s]
    bool clifford = false;        // rotation-shaped with |a0|^2 <= eps^2 |a1|^2 wherever a1 != 0 (drops the residue)
    bool rotation_shaped = false; // RotationLayer::of: a rotation run may hold it
    bool permutation = false;     // PermutationStep::of: one output per pattern, distinct targets
    uint32_t fanout = 2;          // max_fanout
    uint8_t step_delta[16] = {};  // the permutation step: per pattern its local delta and amplitude
    double step_amp[16] = {};
};

// Lift a local delta to full-width masks on the table's support (LocalPtm::lift, prepared.rs:284-299).
void lift_local(const pp_prepared& t, uint32_t d, uint64_t* mx, uint64_t* mz) {
    for (int w = 0; w < 4; ++w) mx[w] = mz[w] = 0;
    for (uint32_t j = 0; j < t.k; ++j) {
        const uint32_t q = t.qubits[j];
        if ((d >> (2 * j)) & 1u) mx[q / 64] |= 1ull << (q % 64);
        if ((d >> (2 * j + 1)) & 1u) mz[q / 64] |= 1ull << (q % 64);
    }
}

// The phase exponent of P_s . g for local patterns s and g on the same support (pauli_string.rs:73-93), mod 4.
uint32_t local_phase(uint32_t s, uint32_t g) {
    uint32_t d = 0;
    for (int j = 0; j < 2; ++j) {
        const uint32_t a = (s >> (2 * j)) & 1u, b = (s >> (2 * j + 1)) & 1u;
        const uint32_t c = (g >> (2 * j)) & 1u, e = (g >> (2 * j + 1)) & 1u;
        d += 2u * (b & c) + (a & b) + (c & e) - ((a ^ c) & (b ^ e));
    }
    return d & 3u;
}

uint32_t local_anticommutes(uint32_t s, uint32_t g) {
    uint32_t f = 0;
    for (int j = 0; j < 2; ++j) f += (((s >> (2 * j)) & 1u) & ((g >> (2 * j + 1)) & 1u)) ^
                                     (((s >> (2 * j + 1)) & 1u) & ((g >> (2 * j)) & 1u));
    return f & 1u;
}

// The amplitude an f32 sum's table holds: the f64 value rounded to float, widened back.
double as_coeff(const pp_sum* s, double v) { return s->kind == PP_COEFF_F64 ? v : (double)(float)v; }

// PauliRotation::prepare_for_layer (channel/rotation.rs:94-123): weight <= 2 tabulates the rotation on its support
// as Prepared::derive_local would probe it (prepared.rs:401-426, push_deltas :257-281): pattern s commuting with g
// stays with amplitude 1; an anticommuting one stays with cos and moves to s ^ g with i sin (s g)'s real sign
// (rotation.rs:43-67); entries whose amplitudes are all zero are not made.  Wider generators keep cos and sin.
pp_prepared prepare_channel_table(const pp_sum* s, const pp_channel& ch, bool adjoint) {
    pp_prepared t{};
    double sv, cv;
    snapped_sin_cos(adjoint ? -ch.theta : ch.theta, &sv, &cv);
    int weight = 0;
    for (uint32_t w = 0; w < s->W; ++w) weight += __builtin_popcountll(ch.gx[w] | ch.gz[w]);
    if (weight > 2) {
        t.kind = PP_PREPARED_ROTATION;
        for (uint32_t w = 0; w < 4; ++w) {
            t.gx[w] = w < s->W ? ch.gx[w] : 0;
            t.gz[w] = w < s->W ? ch.gz[w] : 0;
        }
        t.cos = cv;
        t.sin = sv;
        return t;
    }
    t.kind = PP_PREPARED_LOCAL;
    for (uint32_t w = 0; w < s->W; ++w) {
        uint64_t live = ch.gx[w] | ch.gz[w];
        while (live) {
            t.qubits[t.k++] = 64 * w + (uint32_t)__builtin_ctzll(live);
            live &= live - 1;
        }
    }
    uint32_t g = 0;  // the generator's own pattern
    for (uint32_t j = 0; j < t.k; ++j) {
        const uint32_t q = t.qubits[j];
        g |= (uint32_t)((ch.gx[q / 64] >> (q % 64)) & 1u) << (2 * j);
        g |= (uint32_t)((ch.gz[q / 64] >> (q % 64)) & 1u) << (2 * j + 1);
    }
    const uint32_t dim = 1u << (2 * t.k);
    double ident[16] = {}, move[16] = {};
    bool moves = false;
    for (uint32_t p = 0; p < dim; ++p) {
        if (!local_anticommutes(p, g)) {
            ident[p] = 1.0;
            continue;
        }
        ident[p] = cv;
        const uint32_t total = (1u + local_phase(p, g)) & 3u;
        move[p] = (total & 2u) ? -sv : sv;
        moves = moves || move[p] != 0.0;
    }
    bool any_ident = false;
    for (uint32_t p = 0; p < dim; ++p) any_ident = any_ident || ident[p] != 0.0;
    if (any_ident) {
        pp_delta_entry& e = t.deltas[t.n_deltas++];
        e.local_delta = 0;
        for (uint32_t p = 0; p < dim; ++p) e.amplitude[p] = as_coeff(s, ident[p]);
    }
    if (moves && g != 0) {
        pp_delta_entry& e = t.deltas[t.n_deltas++];
        e.local_delta = g;
        for (uint32_t p = 0; p < dim; ++p) e.amplitude[p] = as_coeff(s, move[p]);
    }
    return t;
}

// Check a prepared channel against the sum and read it (Prepared::Local's invariants, prepared.rs:79-86).
pp_status classify(pp_sum* s, const pp_prepared& t, Prepared* out) {
    Prepared p;
    p.table = t;
    const double eps = 2.220446049250313e-16;
    if (t.kind == PP_PREPARED_ROTATION) {
        p.local = false;
        p.rotation = true;
        for (uint32_t w = 0; w < 4; ++w) {
            const uint64_t live = w < s->W ? live_mask(s->num_qubits, w) : 0;
            if ((t.gx[w] | t.gz[w]) & ~live)
                return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "prepared rotation acts outside the register");
            p.gx[w] = t.gx[w];
            p.gz[w] = t.gz[w];
            p.weight += __builtin_popcountll(t.gx[w] | t.gz[w]);
        }
        p.cos = t.cos;
        p.sin = t.sin;
        // a wide rotation is never key-preserving for the planner, even at sin = 0: the join then drops every zero
        // partner and rescales the anticommuting rows, as the reference's Prepared::Rotation layer does
        const double a0 = as_coeff(s, t.cos), a1 = as_coeff(s, t.sin);
        p.clifford = a1 != 0.0 && a0 * a0 <= eps * eps * (a1 * a1);
        p.fanout = 2;
        *out = p;
        return PP_OK;
    }
    if (t.kind != PP_PREPARED_LOCAL || t.k > 2 || t.n_deltas > 16)
        return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "prepared table: kind, support or entry count out of range");
    for (uint32_t j = 0; j < t.k; ++j)
        if (t.qubits[j] >= s->num_qubits || (j > 0 && t.qubits[j] <= t.qubits[j - 1]))
            return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "prepared table: support qubits not ascending in range");
    const uint32_t dim = 1u << (2 * t.k);
    for (uint32_t e = 0; e < t.n_deltas; ++e) {
        if (t.deltas[e].local_delta >= dim || (e > 0 && t.deltas[e].local_delta <= t.deltas[e - 1].local_delta))
            return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "prepared table: local deltas not ascending within the support");
        for (uint32_t q = 0; q < 16; ++q)
            if (t.deltas[e].amplitude[q] != as_coeff(s, t.deltas[e].amplitude[q]) ||
                (q >= dim && t.deltas[e].amplitude[q] != 0.0))
                return fail(s->ctx, PP_ERR_INVALID_ARGUMENT,
                            "prepared table: an amplitude outside the support's patterns or the coefficient type");
    }
    p.weight = (int)t.k;
    // max_fanout (engine/bucketed.rs:628-641)
    uint32_t fan = 0;
    for (uint32_t q = 0; q < dim; ++q) {
        uint32_t f = 0;
        for (uint32_t e = 0; e < t.n_deltas; ++e) f += t.deltas[e].amplitude[q] != 0.0;
        fan = std::max(fan, f);
    }
    p.fanout = std::max(fan, 1u);
    p.key_preserving = t.n_deltas == 0 || (t.n_deltas == 1 && t.deltas[0].local_delta == 0);
    // RotationLayer::of: an identity entry and one generator entry; where a1 = 0 the identity amplitude is 1, and the
    // patterns a1 reaches are closed under ^ g
    if (t.n_deltas == 2 && t.deltas[0].local_delta == 0) {
        const pp_delta_entry &e0 = t.deltas[0], &e1 = t.deltas[1];
        const uint32_t g = e1.local_delta;
        bool shaped = true, clifford = true;
        for (uint32_t q = 0; q < dim && shaped; ++q) {
            const double a0 = e0.amplitude[q], a1 = e1.amplitude[q];
            if (a1 == 0.0) {
                shaped = a0 == 1.0;
            } else {
                shaped = e1.amplitude[q ^ g] != 0.0;
                clifford = clifford && a0 * a0 <= eps * eps * (a1 * a1);
            }
        }
        p.rotation_shaped = shaped;
        p.clifford = shaped && clifford;
        if (shaped && p.clifford) p.fanout = 1;
        // a rotation's own table: a0 = cos on every pattern that anticommutes with g, a1 = +-sin by (s g)'s phase,
        // and a1 = 0, a0 = 1 on the others; the join then needs only (g, cos, sin)
        if (shaped) {
            int first = -1;
            for (uint32_t q = 0; q < dim; ++q)
                if (local_anticommutes(q, g)) {
                    first = (int)q;
                    break;
                }
            bool exact = first >= 0;
            double cv = 1.0, sv = 0.0;
            if (exact) {
                cv = e0.amplitude[first];
                sv = (((1u + local_phase((uint32_t)first, g)) & 3u) & 2u) ? -e1.amplitude[first] : e1.amplitude[first];
            }
            for (uint32_t q = 0; q < dim && exact; ++q) {
                if (!local_anticommutes(q, g)) {
                    exact = e0.amplitude[q] == 1.0 && e1.amplitude[q] == 0.0;
                } else {
                    const double a1 = (((1u + local_phase(q, g)) & 3u) & 2u) ? -sv : sv;
                    exact = e0.amplitude[q] == cv && e1.amplitude[q] == a1 && std::signbit(e1.amplitude[q]) == std::signbit(a1);
                }
            }
            if (exact && sv != 0.0) {
                p.rotation = true;
                p.cos = cv;
                p.sin = sv;
                lift_local(t, g, p.gx, p.gz);
            }
        }
    }
    // PermutationStep::of: per pattern one emitting entry (for a Clifford rotation table the generator entry where
    // it emits, else the identity), reaching distinct patterns
    if (t.k > 0) {
        bool one_to_one = true;
        bool reached[16] = {};
        for (uint32_t q = 0; q < dim && one_to_one; ++q) {
            int emit = -1;
            if (p.rotation_shaped) {
                if (!p.clifford) {
                    one_to_one = false;
                    break;
                }
                emit = t.deltas[1].amplitude[q] != 0.0 ? 1 : 0;
            } else {
                for (uint32_t e = 0; e < t.n_deltas; ++e)
                    if (t.deltas[e].amplitude[q] != 0.0) {
                        if (emit >= 0) one_to_one = false;
                        emit = (int)e;
                    }
            }
            if (!one_to_one) break;
            if (emit < 0) continue;  // the pattern is dropped (an exact zero)
            const uint32_t to = q ^ t.deltas[emit].local_delta;
            if (reached[to]) one_to_one = false;
            reached[to] = true;
            p.step_delta[q] = (uint8_t)t.deltas[emit].local_delta;
            p.step_amp[q] = t.deltas[emit].amplitude[q];
        }
        p.permutation = one_to_one;
    }
    if (p.key_preserving) {
        for (uint32_t q = 0; q < 16; ++q) {
            p.step_delta[q] = 0;
            p.step_amp[q] = t.n_deltas ? t.deltas[0].amplitude[q] : 0.0;
        }
    }
    *out = p;
    return PP_OK;
}

// The prepared form of a PauliRotation channel, read as the device applies it.
Prepared prepare(pp_sum* s, const pp_channel& ch, bool adjoint) {
    Prepared p;
    const pp_status st = classify(s, prepare_channel_table(s, ch, adjoint), &p);
    (void)st;  // a rotation's own table always classifies
    return p;
}

// The device form of a one-output-per-pattern table under the current device bucketing.
TableStep table_step(const pp_sum* s, const Prepared& p) {
    TableStep st{};
    const pp_prepared& t = p.table;
    const uint32_t q0 = t.k ? t.qubits[0] : 0, q1 = t.k ? t.qubits[t.k - 1] : 0;
    st.word[0] = q0 / 64;
    st.bit[0] = q0 % 64;
    st.word[1] = q1 / 64;
    st.bit[1] = q1 % 64;
    st.mask = (1u << (2 * t.k)) - 1u;
    const uint32_t dim = 1u << (2 * t.k);
    for (uint32_t q = 0; q < 16; ++q) {
        const uint32_t d = q < dim ? p.step_delta[q] : 0u;
        st.delta[q] = (uint8_t)d;
        st.amp[q] = q < dim ? p.step_amp[q] : 1.0;
        uint64_t mx[4], mz[4];
        lift_local(t, d, mx, mz);
        st.bucket[q] = bucket_of_host(s, mx, mz, s->device_bits);
        if (d == 0 && st.amp[q] == 1.0) st.inert |= 1u << q;
    }
    return st;
}

size_t fanout(const Prepared& p) { return p.fanout; }  // max_fanout, engine/bucketed.rs:628-641

bool is_permutation_step(const Prepared& p) { return p.local && p.permutation; }  // PermutationStep::of

bool is_rotation_table(const Prepared& p) { return p.local && p.rotation_shaped; }  // RotationLayer::of

int span_dimension(std::vector<uint32_t> v) {
    int rank = 0;
    for (int bit = 31; bit >= 0; --bit) {
        auto it = std::find_if(v.begin(), v.end(), [&](uint32_t d) { return (d >> bit) & 1u; });
        if (it == v.end()) continue;
        const uint32_t p = *it;
        v.erase(it);
        for (auto& d : v)
            if ((d >> bit) & 1u) d ^= p;
        ++rank;
    }
    return rank;
}

KeepProgram lower_policy(const pp_policy* pol) {
    KeepProgram k{};
    if (!pol) return k;
    k.n = std::min<int32_t>(pol->n_nodes, kMaxKeepNodes);
    for (int i = 0; i < k.n; ++i) {
        k.op[i] = pol->nodes[i].op;
        k.k[i] = pol->nodes[i].k;
        k.disabled[i] = pol->nodes[i].eps < 0.0;
        k.eps2[i] = pol->nodes[i].eps * pol->nodes[i].eps;
    }
    return k;
}

bool valid_policy(const pp_policy* pol) {
    if (!pol) return true;
    if (pol->n_nodes < 0 || pol->n_nodes > 15) return false;
    if (pol->has_topn < 0 || pol->has_topn > 2) return false;
    int depth = 0;
    for (int i = 0; i < pol->n_nodes; ++i) {
        const int op = pol->nodes[i].op;
        if (op == PP_KEEP_AND || op == PP_KEEP_OR) {
            if (depth < 2) return false;
            --depth;
        } else if (op == PP_KEEP_TRUE || op == PP_KEEP_COEFF || op == PP_KEEP_WEIGHT) {
            ++depth;
        } else {
            return false;
        }
    }
    return pol->n_nodes == 0 || depth == 1;
}

// ---- one layer on the device ---------------------------------------------------------------------------------------

pp_status refresh(pp_sum* s);

// The rotation join from `src`'s current columns into `dst`'s column set `out` (dst may be src, with out its
// spare set).  Asynchronous: the counters land in dst->d_stats.
// Device memory a join uses beyond its input and output: the offset scan's tile sums, the counters (kStatSlots) and the
// queue of pairs for the redo launch.  By default dst's own; the no-allocation entry point passes a fixed buffer.
struct JoinScratch {
    uint64_t* tiles;
    unsigned long long* stats;
    uint32_t* redo;
};

// The output rows a join from src needs in dst's column set, and the buckets.
uint64_t join_output_rows(const pp_sum* src) {
    const uint64_t rows_held = src->pending_topn ? src->phys_len : src->len;
    return 2 * rows_held + (uint64_t)kSegAlign * (1ull << src->device_bits);
}

// LDS hash-table slots of a join stage of `rows` rows: a power of two at least twice the rows.
uint32_t join_table_bits(uint32_t rows) {
    uint32_t b = 1;
    while ((1u << b) < 2u * rows) ++b;
    return b;
}

// The largest pair a join from src can meet, a multiple of 64 rows (at most kMaxPairRows).
uint32_t join_pair_bound(const pp_sum* src, uint32_t delta) {
    uint32_t lcap = delta ? 2u * src->max_bucket : src->max_bucket;
    lcap = std::max<uint32_t>(64, (lcap + 63) & ~63u);
    return std::min(lcap, kMaxPairRows);
}

// A join's arguments from src's current columns into dst's column set `out`, staging `scap` rows; the counters,
// redo queue, totals mode and filter are left for the caller.
template <int W, typename T>
JoinArgs<W, T> join_args(const pp_sum* src, const pp_sum* dst, int out, const Prepared& p, const KeepProgram& keep,
                         uint32_t delta, uint32_t scap) {
    const int in = src->cur;
    const uint32_t B = 1u << src->device_bits;
    JoinArgs<W, T> a{};
    a.x = src->col[in].x;
    a.z = src->col[in].z;
    a.c = (const T*)src->col[in].c;
    a.cap = src->col[in].cap;
    a.start = src->bk[in].start;
    a.len = src->bk[in].len;
    a.ox = dst->col[out].x;
    a.oz = dst->col[out].z;
    a.oc = (T*)dst->col[out].c;
    a.ocap = dst->col[out].cap;
    a.ostart = dst->bk[out].start;
    a.olen = dst->bk[out].len;
    a.delta = delta;
    a.hb = delta ? 31u - (uint32_t)__builtin_clz(delta) : 0u;
    a.scap = scap;
    a.tcap_bits = join_table_bits(scap);
    a.num_cosets = delta ? B / 2 : B;
    uint32_t gphase = 0;
    for (uint32_t w = 0; w < W; ++w) {
        a.gx[w] = p.gx[w];
        a.gz[w] = p.gz[w];
        gphase += (uint32_t)__builtin_popcountll(p.gx[w] & p.gz[w]);
    }
    a.gphase = gphase & 3u;
    a.cos_t = (T)p.cos;
    a.sin_t = (T)p.sin;
    a.clifford = p.clifford ? 1 : 0;
    a.keep = keep;
    return a;
}

pp_status launch_join(pp_sum* src, pp_sum* dst, int out, const Prepared& p, const KeepProgram& keep,
                      const JoinScratch* fixed = nullptr) {
    pp_context* ctx = src->ctx;
    pp_status st;
    const int in = src->cur;
    const uint32_t B = 1u << src->device_bits;
    if (!fixed) {
        if ((st = ensure_columns(dst, out, join_output_rows(src))) != PP_OK) return st;
        if ((st = ensure_buckets(dst, out, B)) != PP_OK) return st;
    }
    const uint32_t delta = bucket_of_host(src, p.gx, p.gz, src->device_bits);
    // the scan also clears dst's counters
    if (fixed) {
        PP_HIP(ctx, launch_bucket_offsets(ctx->stream, src->bk[in].len, B, 1, delta, 0, 0, dst->bk[out].start,
                                          fixed->stats + 3, fixed->tiles, fixed->stats, kStatSlots,
                                          dst->bk[out].rows));
    } else if ((st = offsets(dst, src->bk[in].len, B, 1, delta, 0, 0, dst->bk[out].start, true,
                             dst->bk[out].rows)) != PP_OK) {
        return st;
    }
    // lcap bounds the largest pair; the first launch stages scap anticommuting rows per pair, and the pairs with
    // more (rare: a generator anticommutes with a small fraction of a random bucket) go to a redo launch that
    // stages lcap.  LDS per workgroup, and so the resident workgroups per CU, follow scap.
    const uint32_t lcap = join_pair_bound(src, delta);
    const uint32_t scap = join_stage_rows(ctx->tuning.stage_rows, lcap);
    const uint32_t cosets = delta ? B / 2 : B;
    // sized for B whatever this channel's cosets, so that no later layer on the same bucketing reallocates it (a
    // captured graph holds the pointer)
    if (!fixed && dst->redo_cap < B) {
        if (dst->d_redo) (void)hipFree(dst->d_redo);
        dst->d_redo = nullptr;
        dst->redo_cap = 0;
        if ((st = dev_alloc(ctx, (void**)&dst->d_redo, (size_t)B * 8)) != PP_OK) return st;
        dst->redo_cap = 2ull * B;  // and the next refinement's
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const bool timed = ctx->profiling && !ctx->capturing && !fixed;
    unsigned long long* const stats = fixed ? fixed->stats : dst->d_stats;
    uint32_t* const redo = fixed ? fixed->redo : dst->d_redo;
    // Up to kJoinCountPairs pairs, the join's workgroups add the totals themselves (one same-address atomic each)
    // and the length-reduce launch is skipped; measurement keeps the reduce, which also feeds the profile counter.
    constexpr uint32_t kJoinCountPairs = 512;
    const bool count_in_join = !timed && cosets <= kJoinCountPairs;
    if (timed) {
        PP_HIP(ctx, hipEventCreate(&ev0));
        PP_HIP(ctx, hipEventCreate(&ev1));
        PP_HIP(ctx, hipEventRecord(ev0, ctx->stream));
    }
    hipError_t e = dispatch(src, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        JoinArgs<W, T> a = join_args<W, T>(src, dst, out, p, keep, delta, scap);
        a.stats = stats;
        a.redo = redo;
        a.redo_count = stats + 4;
        a.count_totals = count_in_join ? 1u : 0u;
        a.filter = src->pending_topn ? src->d_topn : nullptr;
        a.filter_flags = src->pending_topn ? src->pending_flags : nullptr;
        a.filter_score = src->pending_topn ? src->pending_score : nullptr;
        hipError_t le = launch_rotation_join<W, T>(ctx->stream, a, ctx->tuning.groups_per_cu, ctx->tuning.join_block);
        if (le != hipSuccess || scap >= lcap) return le;  // a first launch staging lcap never queues a pair
        a.redo_mode = 1;
        a.scap = lcap;
        a.tcap_bits = join_table_bits(lcap);
        return launch_rotation_join<W, T>(ctx->stream, a, 0, ctx->tuning.join_block);
    });
    PP_HIP(ctx, e);
    if (timed) {
        PP_HIP(ctx, hipEventRecord(ev1, ctx->stream));
        ctx->events.push_back({ev0, ev1});
        ctx->prof_rows_in += src->len;
    }
    if (!count_in_join)
        PP_HIP(ctx, launch_reduce_lens(ctx->stream, dst->bk[out].len, B, stats, timed ? ctx->d_prof : nullptr));
    return PP_OK;
}

// Whether a join layer runs in place (k_rotation_inplace) rather than into the spare columns: the layout's segment
// rooms are known, the layer is not a Clifford arm (every anticommuting row would leave a hole) and earlier moves have
// not left more dead rows than live ones.  A pending TopN selection is read through, its dropped rows holes that the
// new rows and the moved-down rows above each side's final length fill.
// Below 2^18 rows the layer is launch-bound and the out-of-place join, one synchronization, is the faster
// (PP_IN_PLACE_MIN_ROWS overrides the bound, for tests of the in-place path on small sums).
uint64_t in_place_min_rows() {
    const char* e = std::getenv("PP_IN_PLACE_MIN_ROWS");
    return e ? std::strtoull(e, nullptr, 10) : (1ull << 18);
}

// Whether a side of an in-place layer that loses more rows than it gains closes its holes in place (default) or
// sends its pair to be moved whole to fresh segments (PP_IN_PLACE_COMPACT=0, for comparison: the same result).
bool in_place_compaction() {
    const char* e = std::getenv("PP_IN_PLACE_COMPACT");
    return !(e && std::strcmp(e, "0") == 0);
}

bool join_in_place(const pp_sum* s, const Prepared& p) {
    if (!s->ctx->tuning.join_in_place || !s->has_room || p.clifford || s->len == 0 || s->len < in_place_min_rows())
        return false;
    const uint64_t B = 1ull << s->device_bits;
    return s->tail <= 3 * s->len + 4ull * kSegAlign * B;
}

// Grow column set `which` to at least `need` rows, keeping rows [0, keep) where they are (the bucket arrays stay
// valid).
pp_status grow_columns_keeping(pp_sum* s, int which, uint64_t need, uint64_t keep) {
    Columns& c = s->col[which];
    if (c.cap >= need) return PP_OK;
    Columns old = c;
    c = Columns{};
    pp_status st = ensure_columns(s, which, need);
    if (st != PP_OK) {
        free_columns(c);
        c = old;
        return st;
    }
    hipStream_t stream = s->ctx->stream;
    for (uint32_t w = 0; w < s->W; ++w) {
        PP_HIP(s->ctx, hipMemcpyAsync(c.x + w * c.cap, old.x + w * old.cap, keep * 8, hipMemcpyDeviceToDevice, stream));
        PP_HIP(s->ctx, hipMemcpyAsync(c.z + w * c.cap, old.z + w * old.cap, keep * 8, hipMemcpyDeviceToDevice, stream));
    }
    PP_HIP(s->ctx, hipMemcpyAsync(c.c, old.c, keep * s->coeff_bytes(), hipMemcpyDeviceToDevice, stream));
    PP_HIP(s->ctx, hipStreamSynchronize(stream));
    free_columns(old);
    return PP_OK;
}

// A join layer in place: the first launch does every pair that fits its segments and queues the others untouched;
// then the columns are grown if the queued pairs' moves need it, and the redo launch does the queued pairs, moving
// those that do not fit to fresh segments after the last used row.
pp_status layer_in_place(pp_sum* s, const Prepared& p, const KeepProgram& keep) {
    pp_context* ctx = s->ctx;
    pp_status st;
    const int cur = s->cur;
    const uint32_t B = 1u << s->device_bits;
    const uint32_t delta = bucket_of_host(s, p.gx, p.gz, s->device_bits);
    const uint32_t cosets = delta ? B / 2 : B;
    uint32_t lcap = delta ? 2u * s->max_bucket : s->max_bucket;
    lcap = std::max<uint32_t>(64, (lcap + 63) & ~63u);
    if (lcap > kMaxPairRows) lcap = kMaxPairRows;
    // a generator anticommutes with a fraction of a bucket: 256 staged rows leave room for six workgroups per CU
    const uint32_t scap = std::min<uint32_t>(ctx->tuning.stage_rows ? join_stage_rows(ctx->tuning.stage_rows, lcap)
                                                                     : 256u,
                                             lcap);
    auto table_bits_for = [](uint32_t rows) {
        uint32_t b = 1;
        while ((1u << b) < 2u * rows) ++b;
        return b;
    };
    if (s->redo_cap < B) {
        if (s->d_redo) (void)hipFree(s->d_redo);
        s->d_redo = nullptr;
        s->redo_cap = 0;
        if ((st = dev_alloc(ctx, (void**)&s->d_redo, (size_t)B * 8)) != PP_OK) return st;
        s->redo_cap = 2ull * B;  // and the next refinement's
    }
    PP_HIP(ctx, launch_fill_u64(ctx->stream, s->d_stats, kStatSlots, 0ull));
    const bool timed = ctx->profiling && !ctx->capturing;
    constexpr uint32_t kCountPairs = 512;
    const bool count_in_join = !timed && cosets <= kCountPairs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (timed) {
        PP_HIP(ctx, hipEventCreate(&ev0));
        PP_HIP(ctx, hipEventCreate(&ev1));
        PP_HIP(ctx, hipEventRecord(ev0, ctx->stream));
    }
    auto launch = [&](uint32_t redo_mode, uint32_t stage) {
        return dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            InplaceArgs<W, T> a{};
            a.x = s->col[cur].x;
            a.z = s->col[cur].z;
            a.c = (T*)s->col[cur].c;
            a.cap = s->col[cur].cap;
            a.start = s->bk[cur].start;
            a.len = s->bk[cur].len;
            a.bcap = s->bk[cur].rows;
            a.tail_base = s->tail;
            a.tail_limit = s->col[cur].cap;
            a.delta = delta;
            a.hb = delta ? 31u - (uint32_t)__builtin_clz(delta) : 0u;
            a.scap = stage;
            a.tcap_bits = table_bits_for(stage);
            a.num_cosets = cosets;
            uint32_t gphase = 0;
            for (uint32_t w = 0; w < W; ++w) {
                a.gx[w] = p.gx[w];
                a.gz[w] = p.gz[w];
                gphase += (uint32_t)__builtin_popcountll(p.gx[w] & p.gz[w]);
            }
            a.gphase = gphase & 3u;
            a.cos_t = (T)p.cos;
            a.sin_t = (T)p.sin;
            a.clifford = p.clifford ? 1 : 0;
            a.keep = keep;
            a.stats = s->d_stats;
            a.redo = s->d_redo;
            a.redo_count = s->d_stats + 4;
            a.redo_mode = redo_mode;
            a.count_totals = count_in_join ? 1u : 0u;
            a.compact = in_place_compaction() ? 1u : 0u;
            if (s->pending_topn) {  // the layer reads its input through the selection
                a.filter = s->d_topn;
                a.filter_flags = s->pending_flags;
                a.filter_score = s->pending_score;
            }
            return launch_rotation_inplace<W, T>(ctx->stream, a);
        });
    };
    // the lengths come from the workgroups' own totals (few pairs) or one reduce over the buckets
    auto count = [&]() -> pp_status {
        if (!count_in_join) PP_HIP(ctx, launch_reduce_lens(ctx->stream, s->bk[cur].len, B, s->d_stats, nullptr));
        return read_stats(s);
    };
    PP_HIP(ctx, launch(0, scap));
    if (timed) {
        PP_HIP(ctx, hipEventRecord(ev1, ctx->stream));
        ctx->events.push_back({ev0, ev1});
        ctx->prof_rows_in += s->len;
    }
    if ((st = count()) != PP_OK) return st;  // the one synchronization when every pair fits
    if (s->h_stats[4]) {  // queued pairs: room for their moves, then the redo launch, then the lengths again
        if ((st = grow_columns_keeping(s, cur, s->tail + s->h_stats[7], s->tail)) != PP_OK) return st;
        if (!count_in_join) PP_HIP(ctx, launch_fill_u64(ctx->stream, s->d_stats, 2, 0ull));
        hipEvent_t r0 = nullptr, r1 = nullptr;
        if (timed) {
            PP_HIP(ctx, hipEventCreate(&r0));
            PP_HIP(ctx, hipEventCreate(&r1));
            PP_HIP(ctx, hipEventRecord(r0, ctx->stream));
        }
        PP_HIP(ctx, launch(1, lcap));
        if (timed) {
            PP_HIP(ctx, hipEventRecord(r1, ctx->stream));
            ctx->events.push_back({r0, r1});
        }
        if ((st = count()) != PP_OK) return st;
    }
    if (s->h_stats[2])
        return fail(ctx, PP_ERR_DRIVER,
                    s->h_stats[2] & 2 ? "in-place layer: no room to move a pair" : "in-place layer: staging overflow");
    s->len = s->h_stats[0];
    s->max_bucket = (uint32_t)s->h_stats[1];
    s->tail += s->h_stats[5];
    s->pending_topn = false;  // the layer read its input through the selection
    if (timed) ctx->prof_rows_out += s->len;
    return PP_OK;
}

pp_status class_pass(pp_sum* s, const Prepared* layers, uint32_t n, const KeepProgram& keep, uint64_t* len_after,
                     bool* done);  // below

pp_status device_layer(pp_sum* s, const Prepared& p, const KeepProgram& keep) {
    pp_context* ctx = s->ctx;
    pp_status st;
    if (!p.key_preserving && !p.rotation) {  // any other local table: a one-layer class pass
        bool done = false;
        if ((st = class_pass(s, &p, 1, keep, nullptr, &done)) != PP_OK) return st;
        if (!done) return fail(ctx, PP_ERR_DRIVER, "class pass: a class bucket outgrew the largest LDS set");
        return PP_OK;
    }
    if (p.key_preserving && (st = materialize_topn(s)) != PP_OK) return st;
    if ((st = settle_device_bits(s)) != PP_OK) return st;
    const uint32_t B = 1u << s->device_bits;
    if (p.key_preserving) {
        if ((st = clear_stats(s)) != PP_OK) return st;
        const TableStep step = table_step(s, p);
        hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            RescaleArgs<W, T> a{};
            a.x = s->col[s->cur].x;
            a.z = s->col[s->cur].z;
            a.c = (T*)s->col[s->cur].c;
            a.cap = s->col[s->cur].cap;
            a.start = s->bk[s->cur].start;
            a.len = s->bk[s->cur].len;
            a.step = step;
            a.keep = keep;
            a.stats = s->d_stats;
            return launch_rescale<W, T>(ctx->stream, a, B);
        });
        PP_HIP(ctx, e);
        return count_layout(s, s->cur, B);
    }
    if (join_in_place(s, p)) return layer_in_place(s, p, keep);
    const int out = 1 - s->cur;
    if ((st = launch_join(s, s, out, p, keep)) != PP_OK) return st;
    s->pending_topn = false;  // the join read its input through the selection
    if ((st = read_stats(s)) != PP_OK) return st;
    if (s->h_stats[2]) return fail(ctx, PP_ERR_DRIVER, "rotation join: a pair exceeded the largest-pair bound");
    s->len = s->h_stats[0];
    s->max_bucket = (uint32_t)s->h_stats[1];
    s->tail = s->h_stats[3];
    s->has_room = true;
    s->cur = out;
    return PP_OK;
}

// Whether a TopN selection seeds its first pass with the previous boundary (default) or runs its three passes
// (PP_TOPN_SEEDING=0, for comparison: both decide the same digits).
bool topn_seeding() {
    const char* e = std::getenv("PP_TOPN_SEEDING");
    return !(e && std::strcmp(e, "0") == 0);
}

// Whether pp_propagate walks permutation runs in one pass (default) or layer by layer (PP_PERMUTATION_RUNS=layers,
// for comparison: both give the same keys and coefficients).
bool fused_permutation_runs() {
    const char* e = std::getenv("PP_PERMUTATION_RUNS");
    return !(e && std::strcmp(e, "layers") == 0);
}

// A permutation run on the device (engine/bucketed/permutation_run.rs:297-...): k_perm_walk takes every row through
// the steps in registers, in place, counting each output bucket and each step's drops; an offset scan lays out the
// destinations; k_perm_scatter moves the rows into the spare columns.  len_after[j] receives the length after step
// j, as the per-layer path would report it.
pp_status permutation_run_device(pp_sum* s, const Prepared* steps, uint32_t n_steps, const KeepProgram& keep,
                                 uint64_t* len_after) {
    pp_context* ctx = s->ctx;
    pp_status st;
    if ((st = materialize_topn(s)) != PP_OK || (st = settle_device_bits(s)) != PP_OK) return st;
    const uint32_t B = 1u << s->device_bits;
    uint32_t *d_dest = nullptr, *d_counts = nullptr;
    unsigned long long* d_deaths = nullptr;
    TableStep* d_steps = nullptr;
    if ((st = slot_scratch(s, kScrRunCounts, (size_t)B * 4, &d_counts)) != PP_OK ||
        (st = slot_scratch(s, kScrDeaths, (size_t)kMaxPermSteps * 8, &d_deaths)) != PP_OK ||
        (st = slot_scratch(s, kScrSteps, (size_t)kMaxPermSteps * sizeof(TableStep), &d_steps)) != PP_OK)
        return st;
    const uint64_t len_in = s->len;
    std::vector<unsigned long long> deaths(n_steps, 0);
    for (uint32_t first = 0; first < n_steps; first += kMaxPermSteps) {
        const uint32_t m = std::min(kMaxPermSteps, n_steps - first);
        const int in = s->cur, out = 1 - s->cur;
        if ((st = slot_scratch(s, kScrDest, s->col[in].cap * 4, &d_dest)) != PP_OK) return st;
        // every step's table with its device bucket deltas under the current bucketing
        std::vector<TableStep> chunk(m);
        for (uint32_t j = 0; j < m; ++j) chunk[j] = table_step(s, steps[first + j]);
        PP_HIP(ctx, hipMemcpyAsync(d_steps, chunk.data(), m * sizeof(TableStep), hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(ctx, hipMemsetAsync(d_counts, 0, (size_t)B * 4, ctx->stream));
        PP_HIP(ctx, hipMemsetAsync(d_deaths, 0, (size_t)m * 8, ctx->stream));
        hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            PermWalkArgs<W, T> a{};
            a.x = s->col[in].x;
            a.z = s->col[in].z;
            a.c = (T*)s->col[in].c;
            a.cap = s->col[in].cap;
            a.start = s->bk[in].start;
            a.len = s->bk[in].len;
            a.steps = d_steps;
            a.num_steps = m;
            a.keep = keep;
            a.dest = d_dest;
            a.counts = d_counts;
            a.deaths = d_deaths;
            return launch_perm_walk<W, T>(ctx->stream, a, B);
        });
        PP_HIP(ctx, e);
        if ((st = ensure_columns(s, out, s->len + (uint64_t)kSegAlign * B)) != PP_OK) return st;
        if ((st = ensure_buckets(s, out, B)) != PP_OK) return st;
        if ((st = offsets(s, d_counts, B, 3, 0, 0, 0, s->bk[out].start, false, s->bk[out].rows)) != PP_OK) return st;
        PP_HIP(ctx, hipMemsetAsync(s->bk[out].len, 0, (size_t)B * 4, ctx->stream));
        e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            PermScatterArgs<W, T> a{};
            a.x = s->col[in].x;
            a.z = s->col[in].z;
            a.c = (const T*)s->col[in].c;
            a.cap = s->col[in].cap;
            a.start = s->bk[in].start;
            a.len = s->bk[in].len;
            a.dest = d_dest;
            a.ox = s->col[out].x;
            a.oz = s->col[out].z;
            a.oc = (T*)s->col[out].c;
            a.ocap = s->col[out].cap;
            a.ostart = s->bk[out].start;
            a.olen = s->bk[out].len;
            return launch_perm_scatter<W, T>(ctx->stream, a, B);
        });
        PP_HIP(ctx, e);
        PP_HIP(ctx, hipMemcpyAsync(deaths.data() + first, d_deaths, m * 8, hipMemcpyDeviceToHost, ctx->stream));
        s->cur = out;
        if ((st = count_layout(s, s->cur, B)) != PP_OK) return st;  // synchronizes
        s->tail = s->h_stats[3];
        s->has_room = true;
    }
    uint64_t len = len_in;
    for (uint32_t j = 0; j < n_steps; ++j) {
        len -= deaths[j];
        if (len_after) len_after[j] = len;
    }
    if (len != s->len)
        return fail(ctx, PP_ERR_DRIVER, "permutation run: " + std::to_string(s->len) + " rows placed, " +
                                            std::to_string(len) + " survived the walk");
    return PP_OK;
}

// ---- the class pass: a program of local tables in one pass (pp_class.hip) ----------------------------------------

constexpr uint32_t kClassMaxQubits = 16;   // a program's qubits together (RR_MAX_QUBITS of the reference's run)
constexpr uint32_t kClassMinRunLayers = 3; // rotation runs at least this long take the class pass

// The device form of a local table: its slots, its entries (a Clifford rotation table without the identity residue
// the pair join drops), and the patterns it changes.  A pattern is left alone when the identity entry keeps it with
// amplitude exactly 1, no other entry emits from it and none emits into it.
ClassLayer class_layer(const Prepared& p) {
    ClassLayer L{};
    const pp_prepared& t = p.table;
    const uint32_t q0 = t.k ? t.qubits[0] : 0, q1 = t.k ? t.qubits[t.k - 1] : 0;
    L.word[0] = q0 / 64;
    L.bit[0] = q0 % 64;
    L.word[1] = q1 / 64;
    L.bit[1] = q1 % 64;
    L.mask = (1u << (2 * t.k)) - 1u;
    const uint32_t dim = 1u << (2 * t.k);
    L.n_entries = t.n_deltas;
    int identity = -1;
    for (uint32_t e = 0; e < t.n_deltas; ++e) {
        L.delta[e] = t.deltas[e].local_delta;
        if (L.delta[e] == 0) identity = (int)e;
        for (uint32_t q = 0; q < 16; ++q) L.amp[e][q] = q < dim ? t.deltas[e].amplitude[q] : 0.0;
    }
    if (p.rotation_shaped && p.clifford && identity == 0 && t.n_deltas == 2)
        for (uint32_t q = 0; q < dim; ++q)
            if (L.amp[1][q] != 0.0) L.amp[0][q] = 0.0;
    for (uint32_t q = 0; q < dim; ++q) {
        bool still = identity >= 0 && L.amp[identity][q] == 1.0;
        for (uint32_t e = 0; e < t.n_deltas && still; ++e)
            if ((int)e != identity) still = L.amp[e][q] == 0.0 && L.amp[e][q ^ L.delta[e]] == 0.0;
        if (!still) L.active |= 1u << q;
    }
    return L;
}

// Whether layers[0..n) fit one class pass: local tables whose qubits number at most kClassMaxQubits.
bool class_program_fits(const Prepared* layers, uint32_t n) {
    if (n == 0 || n > kClassMaxLayers) return false;
    std::vector<uint32_t> qs;
    for (uint32_t l = 0; l < n; ++l) {
        if (!layers[l].local) return false;
        for (uint32_t j = 0; j < layers[l].table.k; ++j) qs.push_back(layers[l].table.qubits[j]);
    }
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    return qs.size() <= kClassMaxQubits;
}

// Whether pp_propagate and pp_apply_run give rotation runs to the class pass (PP_ROTATION_RUNS=class) or apply them
// layer by layer (the default): both give the same keys, coefficients and per-layer lengths.  Layer by layer is the
// faster on this device at every size measured: at 2^24 terms an eight-layer run takes 12 ms as in-place layers and
// 26-41 ms as a class pass, whose routing and placement pay one global atomic per row; on the 6x6 TFIM (about 3*10^4
// terms) 29.5 against 41.7 ms.
bool class_rotation_runs() {
    const char* e = std::getenv("PP_ROTATION_RUNS");
    return e && std::strcmp(e, "class") == 0;
}

// layers[0..n) in one class pass, the sum's length after each into len_after (optional).  *done = false, the sum as it
// was, when a class bucket outgrew the LDS sets at every class bucketing tried: the caller applies the layers one by
// one.  Two synchronizations: the counters after the apply, the layout's length at the end.
pp_status class_pass(pp_sum* s, const Prepared* layers, uint32_t n, const KeepProgram& keep, uint64_t* len_after,
                     bool* done) {
    pp_context* ctx = s->ctx;
    pp_status st;
    *done = false;
    if (!class_program_fits(layers, n)) return PP_OK;
    if ((st = materialize_topn(s)) != PP_OK || (st = settle_device_bits(s)) != PP_OK) return st;
    const uint64_t len_in = s->len;
    if (len_in == 0) {
        for (uint32_t l = 0; l < n && len_after; ++l) len_after[l] = 0;
        *done = true;
        return PP_OK;
    }
    const uint32_t B = 1u << s->device_bits;
    // the program and the counters pass through one pinned block of the context (the previous pass has drained it:
    // every pass ends with a synchronization)
    constexpr size_t kPinnedStats = kClassMaxLayers * sizeof(ClassLayer);
    if (!ctx->class_pinned &&
        hipHostMalloc(&ctx->class_pinned, kPinnedStats + kClassStatSlots * 8, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        ctx->class_pinned = nullptr;
        return fail(ctx, PP_ERR_OUT_OF_MEMORY, "class pass: pinned host block");
    }
    ClassLayer* prog = static_cast<ClassLayer*>(ctx->class_pinned);
    unsigned long long* h = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->class_pinned) + kPinnedStats);
    uint64_t qmask[4] = {};
    uint32_t prog_words = 0;
    for (uint32_t l = 0; l < n; ++l) {
        prog[l] = class_layer(layers[l]);
        prog_words += prog[l].n_entries * 16u;
        for (uint32_t j = 0; j < layers[l].table.k; ++j)
            qmask[layers[l].table.qubits[j] / 64] |= 1ull << (layers[l].table.qubits[j] % 64);
    }
    ClassLayer* d_prog = nullptr;
    uint32_t* d_ukeep = nullptr;
    if ((st = slot_scratch(s, kScrClassLayers, n * sizeof(ClassLayer), &d_prog)) != PP_OK ||
        (st = slot_scratch(s, kScrClassKeep, (size_t)B * 4, &d_ukeep)) != PP_OK)
        return st;
    PP_HIP(ctx, hipMemcpyAsync(d_prog, prog, n * sizeof(ClassLayer), hipMemcpyHostToDevice, ctx->stream));
    const size_t cb = s->coeff_bytes();
    uint64_t *sx = nullptr, *sz = nullptr;
    void* sc = nullptr;
    if ((st = slot_scratch(s, kScrClassSX, (size_t)s->W * len_in * 8, &sx)) != PP_OK ||
        (st = slot_scratch(s, kScrClassSZ, (size_t)s->W * len_in * 8, &sz)) != PP_OK ||
        (st = slot_scratch(s, kScrClassSC, len_in * cb, &sc)) != PP_OK)
        return st;
    // class buckets of about 16 rows of the input, packed into chunks of a quarter of the LDS set; on an overflow
    // smaller chunks and finer class buckets, then a larger set
    int cbits = 0;
    while (cbits < 22 && ((uint64_t)16 << cbits) < len_in) ++cbits;
    uint64_t growth = 1;  // the program's worst fanout, capped: a chunk of lcap / growth rows fits its outputs
    for (uint32_t l = 0; l < n && growth < 16; ++l) growth *= std::max<uint32_t>(layers[l].fanout, 1u);
    uint32_t lcap = 512, tbits = 10, chunk = (uint32_t)std::max<uint64_t>(16, 512 / std::min<uint64_t>(growth, 16));
    uint64_t ocap = std::max<uint64_t>({2 * len_in, (uint64_t)4096, s->class_out_hint});
    uint64_t *ox = nullptr, *oz = nullptr;
    void* oc = nullptr;
    uint32_t* ob = nullptr;
    unsigned long long* d_cst = nullptr;
    uint32_t* d_ocount = nullptr;
    int attempt = 0;
    for (;;) {
        const uint32_t C = 1u << cbits;
        // the counters, the outputs per bucket and the rows per class bucket: one block, one clear
        unsigned char* d_counts = nullptr;
        uint64_t* d_cstart = nullptr;
        const size_t count_bytes = kClassStatSlots * 8 + (size_t)B * 4 + (size_t)C * 4;
        if ((st = slot_scratch(s, kScrClassCount, count_bytes, &d_counts)) != PP_OK ||
            (st = slot_scratch(s, kScrClassStart, (size_t)C * 8, &d_cstart)) != PP_OK ||
            (st = slot_scratch(s, kScrClassOX, (size_t)s->W * ocap * 8, &ox)) != PP_OK ||
            (st = slot_scratch(s, kScrClassOZ, (size_t)s->W * ocap * 8, &oz)) != PP_OK ||
            (st = slot_scratch(s, kScrClassOC, ocap * cb, &oc)) != PP_OK ||
            (st = slot_scratch(s, kScrClassOB, ocap * 4, &ob)) != PP_OK || (st = ensure_tiles(s, C)) != PP_OK)
            return st;
        d_cst = reinterpret_cast<unsigned long long*>(d_counts);
        d_ocount = reinterpret_cast<uint32_t*>(d_counts + kClassStatSlots * 8);
        uint32_t* const d_ccount = d_ocount + B;
        PP_HIP(ctx, hipMemsetAsync(d_counts, 0, count_bytes, ctx->stream));
        hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            ClassArgs<W, T> a{};
            a.x = s->col[s->cur].x;
            a.z = s->col[s->cur].z;
            a.c = (const T*)s->col[s->cur].c;
            a.cap = s->col[s->cur].cap;
            a.start = s->bk[s->cur].start;
            a.len = s->bk[s->cur].len;
            a.num_buckets = B;
            a.bits = s->device_bits;
            a.rows = s->d_rows;
            a.cbits = cbits;
            a.num_classes = C;
            for (int w = 0; w < W; ++w) a.qx[w] = a.qz[w] = qmask[w];
            a.layers = d_prog;
            a.n_layers = n;
            a.prog_words = prog_words;
            a.keep = keep;
            a.ccount = d_ccount;
            a.cstart = d_cstart;
            a.ukeep = d_ukeep;
            a.sx = sx;
            a.sz = sz;
            a.sc = (T*)sc;
            a.scap = len_in;
            a.ox = ox;
            a.oz = oz;
            a.oc = (T*)oc;
            a.obucket = ob;
            a.ocap = ocap;
            a.ocount = d_ocount;
            a.stats = d_cst;
            a.lcap = lcap;
            a.tbits = tbits;
            a.chunk_rows = chunk;
            hipError_t r = launch_class_route<W, T>(ctx->stream, a);
            if (r == hipSuccess)
                r = launch_bucket_offsets(ctx->stream, d_ccount, C, 0, 0, 0, 0, d_cstart, d_cst + kClassStaged,
                                          s->d_tiles);
            if (r == hipSuccess) r = launch_class_gather<W, T>(ctx->stream, a);
            if (r == hipSuccess) r = launch_class_apply<W, T>(ctx->stream, a);
            return r;
        });
        PP_HIP(ctx, e);
        PP_HIP(ctx, hipMemcpyAsync(h, d_cst, kClassStatSlots * 8, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const unsigned long long flags = h[kClassFlags];
        if (flags == 0) break;
        if ((flags & 3ull) == 0) {  // only the output staging ran out: its exact need is known now
            ocap = h[kClassOutRows];
        } else if (attempt == 0) {
            cbits = std::min(cbits + 2, 24);
            chunk = 32;
        } else if (attempt == 1) {
            cbits = std::min(cbits + 1, 24);
            lcap = 1024;
            tbits = 11;
            chunk = 32;
        } else {
            return PP_OK;  // a class that outgrows the largest set: layer by layer
        }
        if (flags & 3ull) ++attempt;
    }
    // the output layout in the spare columns: each bucket's untouched rows, then its class outputs
    const uint64_t out_rows = h[kClassOutRows], total = h[kClassUntouched] + out_rows;
    s->class_out_hint = out_rows + out_rows / 8;
    const int in = s->cur, out = 1 - s->cur;
    if ((st = ensure_columns(s, out, total + (uint64_t)kSegAlign * B)) != PP_OK) return st;
    if ((st = ensure_buckets(s, out, B)) != PP_OK) return st;
    hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        ClassArgs<W, T> a{};
        a.num_buckets = B;
        a.ukeep = d_ukeep;
        a.ocount = d_ocount;
        a.plen = s->bk[out].len;
        a.stats = d_cst;
        return launch_class_lens<W, T>(ctx->stream, a);
    });
    PP_HIP(ctx, e);
    if ((st = offsets(s, s->bk[out].len, B, 3, 0, 0, 0, s->bk[out].start, false, s->bk[out].rows)) != PP_OK) return st;
    e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        ClassArgs<W, T> a{};
        a.x = s->col[in].x;
        a.z = s->col[in].z;
        a.c = (const T*)s->col[in].c;
        a.cap = s->col[in].cap;
        a.start = s->bk[in].start;
        a.len = s->bk[in].len;
        a.num_buckets = B;
        for (int w = 0; w < W; ++w) a.qx[w] = a.qz[w] = qmask[w];
        a.layers = d_prog;
        a.n_layers = n;
        a.keep = keep;
        a.ukeep = d_ukeep;
        a.ocount = d_ocount;
        a.ox = ox;
        a.oz = oz;
        a.oc = (T*)oc;
        a.obucket = ob;
        a.ocap = ocap;
        a.px = s->col[out].x;
        a.pz = s->col[out].z;
        a.pc = (T*)s->col[out].c;
        a.pcap = s->col[out].cap;
        a.pstart = s->bk[out].start;
        a.plen = s->bk[out].len;
        return launch_class_place<W, T>(ctx->stream, a, out_rows);
    });
    PP_HIP(ctx, e);
    // the longest bucket (from the lens kernel) and the layout's end (the scan's total): one synchronization
    PP_HIP(ctx, hipMemcpyAsync(h + kClassLongest, d_cst + kClassLongest, 8, hipMemcpyDeviceToHost, ctx->stream));
    if ((st = read_stats(s)) != PP_OK) return st;
    s->cur = out;
    s->len = total;
    s->max_bucket = (uint32_t)h[kClassLongest];
    s->tail = s->h_stats[3];
    s->has_room = true;
    for (uint32_t l = 0; l < n && len_after; ++l) len_after[l] = h[kClassUntouched] + h[kClassLayerRows + l];
    if (h[kClassUntouched] + h[kClassLayerRows + n - 1] != total)
        return fail(ctx, PP_ERR_DRIVER, "class pass: the last layer's count disagrees with its outputs");
    *done = true;
    return PP_OK;
}

// The counts of an asynchronous out-of-place layer, without applying a selection left pending for the next layer.
pp_status refresh_counts(pp_sum* s) {
    if (!s->stale) return PP_OK;
    pp_status st = read_stats(s);
    if (st != PP_OK) return st;
    s->stale = false;
    if (s->h_stats[2]) return fail(s->ctx, PP_ERR_DRIVER, "rotation join: a pair exceeded the largest-pair bound");
    s->len = s->h_stats[0];
    s->max_bucket = (uint32_t)s->h_stats[1];
    s->tail = s->h_stats[3];  // the out-of-place layer's offset scan laid the segments out up to here
    s->has_room = true;
    return PP_OK;
}

// Bring len and max_bucket up to date after an asynchronous out-of-place layer (and apply a selection an
// interrupted propagation left pending).
pp_status refresh(pp_sum* s) {
    pp_status st;
    if (s->pending_topn && (st = materialize_topn(s)) != PP_OK) return st;
    return refresh_counts(s);
}

pp_status apply_one(pp_sum* s, const Prepared& p, const KeepProgram& keep) { return device_layer(s, p, keep); }

pp_status batch_topn(pp_batch* b, const std::vector<uint32_t>& which, const std::vector<uint64_t>& keep);  // below
// Whether pp_batch_propagate selects the TopN of every instance that finishes a run in the same step together
// (default) or one instance after another (PP_BATCH_TOPN=0, for comparison: the same kept terms).
bool batch_topn_enabled() {
    const char* e = std::getenv("PP_BATCH_TOPN");
    return !(e && std::strcmp(e, "0") == 0);
}
pp_status topn_device(pp_sum* s, uint64_t n, const float* score = nullptr,
                      bool defer = false);  // pp_topn_impl.inc
pp_status approx_topn_device(pp_sum* s, uint64_t n);  // pp_topn_impl.inc

pp_status check_channels(pp_sum* s, const pp_channel* channels, uint64_t n_channels) {
    for (uint64_t i = 0; i < n_channels; ++i)
        for (uint32_t w = 0; w < 4; ++w) {
            const uint64_t live = w < s->W ? live_mask(s->num_qubits, w) : 0;
            if ((channels[i].gx[w] | channels[i].gz[w]) & ~live)
                return fail(s->ctx, PP_ERR_INVALID_ARGUMENT,
                            "channel " + std::to_string(i) + " acts outside the " + std::to_string(s->num_qubits) +
                                "-qubit register");
        }
    return PP_OK;
}

// run_layers' schedule (engine/driver.rs:319-592) for one sum: the reference's bucket count moves at the start of
// every run (refine_unprepared, engine/backend.rs:155-165), Runs::take (engine/driver.rs:202-258) joins
// permutation steps or rotation tables into runs that share that count, and the layer pass runs once after each.
struct RunPlanner {
    const pp_channel* channels;
    uint64_t n_channels;
    bool adjoint;
    const pp_policy* policy;
    size_t target, min_buckets;
    bool finalizes = false;
    uint64_t limit = UINT64_MAX;  // run_length_limit of TopN and ApproxTopN (builtin.rs:147-149, :462-464)

    RunPlanner(const pp_channel* c, uint64_t n, bool adj, const pp_policy* pol, size_t t, size_t mb)
        : channels(c), n_channels(n), adjoint(adj), policy(pol), target(t), min_buckets(mb) {
        finalizes = pol && pol->has_topn;
        limit = finalizes ? pol->topn : UINT64_MAX;
    }

    const pp_channel& channel_at(uint64_t i) const { return channels[adjoint ? n_channels - 1 - i : i]; }

    // Layer i's prepared table.  The planner, the run and the layer loop ask for the same layers within a few steps
    // of each other, so the last ones are kept (direct-mapped by layer index; a table depends on the sum's width,
    // register and coefficient type only).
    static constexpr uint32_t kPrepCache = 16;
    mutable std::vector<Prepared> prep_cache = std::vector<Prepared>(kPrepCache);
    mutable std::vector<uint64_t> prep_index = std::vector<uint64_t>(kPrepCache, UINT64_MAX);
    Prepared prepared(pp_sum* s, uint64_t i) const {
        const uint32_t slot = (uint32_t)(i % kPrepCache);
        if (prep_index[slot] != i) {
            prep_cache[slot] = prepare(s, channel_at(i), adjoint);
            prep_index[slot] = i;
        }
        return prep_cache[slot];
    }

    // The reference's count for the run starting at layer i, and the run's end.
    uint64_t start_run(pp_sum* s, uint64_t i) const {
        const int want = rebucketed_bits(s->len, s->reference_bits, target, min_buckets);
        s->reference_bits = std::min(want, s->max_bits);
        const Prepared p = prepared(s, i);
        uint64_t end = i + 1;
        if (limit == 0) return end;
        if (s->len <= limit && is_permutation_step(p)) {
            uint64_t j = i + 1;
            while (j < n_channels && is_permutation_step(prepared(s, j))) ++j;
            if (j > i + 1) return j;
        }
        auto mul = [](uint64_t a, uint64_t b) { return (a != 0 && b > UINT64_MAX / a) ? UINT64_MAX : a * b; };
        uint64_t worst = mul(s->len, fanout(p));
        if (worst <= limit && is_rotation_table(p)) {
            std::vector<uint32_t> deltas{0, bucket_of_host(s, channel_at(i).gx, channel_at(i).gz, s->reference_bits)};
            uint64_t j = i + 1;
            while (j < n_channels && j < i + kRunLayers) {
                const Prepared q = prepared(s, j);
                const uint64_t grown = mul(worst, fanout(q));
                if (grown > limit || !is_rotation_table(q)) break;
                auto next = deltas;
                next.push_back(0);
                next.push_back(bucket_of_host(s, channel_at(j).gx, channel_at(j).gz, s->reference_bits));
                if (span_dimension(next) > kRunDimension) break;
                deltas = next;
                worst = grown;
                ++j;
            }
            if (j > i + 1) end = j;
        }
        return end;
    }

    // Layers i .. end - 1 (one run), the sum's length after each into len_after[0 .. end - i) when given.  A
    // permutation run of two or more steps is walked in one pass; any other run layer by layer.
    pp_status apply_run(pp_sum* s, uint64_t i, uint64_t end, const KeepProgram& keep, uint64_t* len_after) const {
        pp_status st;
        bool walk = end > i + 1 && s->len > 0 && fused_permutation_runs();
        std::vector<Prepared> steps;
        for (uint64_t j = i; walk && j < end; ++j) {
            steps.push_back(prepared(s, j));
            walk = is_permutation_step(steps.back());
        }
        if (walk) return permutation_run_device(s, steps.data(), (uint32_t)steps.size(), keep, len_after);
        if (end - i >= kClassMinRunLayers && s->len > 0 && class_rotation_runs()) {
            std::vector<Prepared> run;
            for (uint64_t j = i; j < end; ++j) run.push_back(prepared(s, j));
            bool done = false;
            if ((st = class_pass(s, run.data(), (uint32_t)run.size(), keep, len_after, &done)) != PP_OK) return st;
            if (done) return PP_OK;
        }
        for (uint64_t j = i; j < end; ++j) {
            const Prepared q = prepared(s, j);
            if ((st = apply_one(s, q, keep)) != PP_OK) return st;
            if (len_after) len_after[j - i] = s->len;
        }
        return PP_OK;
    }

    // The layer pass after a run.  With defer, an exact TopN leaves its retain to the next join (the caller applies
    // it with materialize_topn before anything else reads the rows).
    pp_status finish_run(pp_sum* s, bool defer = false) const;

    // A class-pass window from layer i: rotation tables that are not permutation steps, at most kClassMaxLayers of
    // them on at most kClassMaxQubits qubits, whatever the planner's run boundaries among them.  Only under a policy
    // with no layer pass: finish_run then does nothing, so those boundaries change nothing but the reference's bucket
    // count, which the caller replays from the window's per-layer lengths.  Returns the window's end (i: none).
    uint64_t window_end(pp_sum* s, uint64_t i, std::vector<Prepared>* out) const {
        out->clear();
        if (finalizes) return i;
        std::vector<uint32_t> qs;
        uint64_t j = i;
        for (; j < n_channels && j < i + kClassMaxLayers; ++j) {
            const Prepared q = prepared(s, j);
            if (!is_rotation_table(q) || is_permutation_step(q)) break;
            std::vector<uint32_t> more = qs;
            for (uint32_t k = 0; k < q.table.k; ++k) more.push_back(q.table.qubits[k]);
            std::sort(more.begin(), more.end());
            more.erase(std::unique(more.begin(), more.end()), more.end());
            if (more.size() > kClassMaxQubits) break;
            qs = more;
            out->push_back(q);
        }
        return j;
    }
};

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// The no-allocation join's scratch: one fixed array per GPU, for callers that cannot allocate device memory
// (pp_apply_rotation_to_noalloc).  Every call on a GPU shares that GPU's copy, which keeps the last call's contents;
// the counters and the redo count are cleared by each call's offset scan before the join reads them.  Calls that
// overlap on one GPU (two contexts' streams) race on it.
// ---------------------------------------------------------------------------------------------------------------

constexpr uint32_t kNoAllocMaxBuckets = 1u << 18;          // device buckets of the largest input (~2^26 terms)
constexpr size_t kNoAllocStatsOffset = 0;                  // kStatSlots counters
constexpr size_t kNoAllocTilesOffset = 256;                // offsets_tiles(kNoAllocMaxBuckets) tile sums
constexpr size_t kNoAllocRedoOffset = 2048;                // kNoAllocMaxBuckets queued pairs
constexpr size_t kNoAllocScratchBytes = kNoAllocRedoOffset + (size_t)kNoAllocMaxBuckets * 4;
static_assert(kStatSlots * 8 <= kNoAllocTilesOffset, "counters overlap the tile sums");
static_assert(kNoAllocTilesOffset + (kNoAllocMaxBuckets / 2048) * 8 <= kNoAllocRedoOffset, "tiles overlap the queue");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};  // the device address of `scratch` on each GPU

// Look up this GPU's copy of the array once, with the device current (pp_context_create).  The array's name in host
// code is only a stand-in for the symbol; kernels and copies take the address found here.
pp_status bind_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return PP_OK;  // that GPU's calls of the entry point are refused
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return PP_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return PP_ERR_DRIVER;
    }
    g_scratch_ptr[device] = (unsigned char*)ptr;
    return PP_OK;
}

unsigned char* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

// Whether settle_device_bits would leave the sum as it is (no refinement or coarsening, which copy and allocate).
bool device_bits_settled(const pp_sum* s) {
    const uint64_t target = s->ctx->tuning.bucket_target;
    int want = 0;
    while (want < s->max_bits && s->len > target * (1ull << want)) ++want;
    const bool coarsen = s->len * 8 < target * (1ull << s->device_bits) && want < s->device_bits;
    return !(want > s->device_bits || coarsen || 2ull * s->max_bucket > kMaxPairRows);
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------------------

extern "C" {

pp_status pp_context_create(int device, pp_context** out) {
    if (!out) return PP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return PP_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) return PP_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return PP_ERR_DRIVER;
    auto* ctx = new (std::nothrow) pp_context();
    if (!ctx) return PP_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    ctx->tuning.bucket_target = default_bucket_target();
    ctx->tuning.stage_rows = default_stage_rows();
    ctx->tuning.groups_per_cu = default_groups_per_cu();
    ctx->tuning.join_block = default_join_block();
    ctx->tuning.join_in_place = default_join_in_place();
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PP_ERR_DRIVER;
    }
    if (bind_noalloc_scratch(device) != PP_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return PP_ERR_DRIVER;
    }
    *out = ctx;
    return PP_OK;
}

void pp_context_destroy(pp_context* ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto& e : ctx->score_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->d_prof) (void)hipFree(ctx->d_prof);
    for (int k = 0; k < 2; ++k) {
        if (ctx->stage[k]) (void)hipHostFree(ctx->stage[k]);
        if (ctx->stage_ev[k]) (void)hipEventDestroy(ctx->stage_ev[k]);
    }
    if (ctx->class_pinned) (void)hipHostFree(ctx->class_pinned);
    for (hipEvent_t e : ctx->lag_ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
}

const char* pp_last_error(const pp_context* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

pp_status pp_sync(pp_context* ctx) {
    if (!ctx) return PP_ERR_INVALID_ARGUMENT;
    PP_NOT_CAPTURING(ctx, "pp_sync");
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PP_OK;
}

pp_status pp_context_stream(pp_context* ctx, void** stream) {
    if (!ctx || !stream) return PP_ERR_INVALID_ARGUMENT;
    *stream = (void*)ctx->stream;
    return PP_OK;
}

int32_t pp_hash_rows(uint32_t num_qubits, uint64_t seed, uint32_t width, uint64_t* rows_x, uint64_t* rows_z) {
    if (width < 1 || width > 4 || num_qubits > 64 * width) return -1;
    return hash_rows(num_qubits, seed, width, rows_x, rows_z);
}

pp_status pp_sum_upload(pp_context* ctx, uint32_t num_qubits, uint32_t width, pp_coeff_kind kind, uint64_t n,
                        const uint64_t* x, const uint64_t* z, const void* coeff, uint64_t hash_seed,
                        pp_sum** out) {
    PP_NOT_CAPTURING(ctx, "pp_sum_upload");
    if (!ctx || !out) return PP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (width != 1 && width != 2) return fail(ctx, PP_ERR_UNSUPPORTED, "width must be 1 or 2");
    if (num_qubits > 64 * width) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "num_qubits exceeds 64 * width");
    if (kind != PP_COEFF_F64 && kind != PP_COEFF_F32) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "coefficient kind");
    if (n && (!x || !z || !coeff)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "null term arrays");
    PP_HIP(ctx, hipSetDevice(ctx->device));
    auto* s = new (std::nothrow) pp_sum();
    if (!s) return fail(ctx, PP_ERR_OUT_OF_MEMORY, "host allocation");
    s->ctx = ctx;
    s->num_qubits = num_qubits;
    s->W = width;
    s->kind = kind;
    s->seed = hash_seed;
    std::vector<uint64_t> rx(kIndexBits * width), rz(kIndexBits * width);
    s->max_bits = hash_rows(num_qubits, hash_seed, width, rx.data(), rz.data());
    s->rows.assign((size_t)kMaxHashRows * 2 * width, 0);
    for (int r = 0; r < s->max_bits; ++r)
        for (uint32_t w = 0; w < width; ++w) {
            s->rows[(2 * r) * width + w] = rx[r * width + w];
            s->rows[(2 * r + 1) * width + w] = rz[r * width + w];
        }
    auto bail = [&](pp_status st) {
        pp_sum_free(s);
        return st;
    };
    pp_status st;
    if ((st = dev_alloc(ctx, (void**)&s->d_rows, s->rows.size() * 8)) != PP_OK) return bail(st);
    if ((st = dev_alloc(ctx, (void**)&s->d_stats, kStatSlots * sizeof(unsigned long long))) != PP_OK) return bail(st);
    if (hipHostMalloc((void**)&s->h_stats, kStatSlots * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return bail(fail(ctx, PP_ERR_OUT_OF_MEMORY, "pinned host allocation"));
    if (hipMemcpy(s->d_rows, s->rows.data(), s->rows.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(ctx, PP_ERR_DRIVER, "hash rows upload"));

    // Keep nonzero terms, bucket them at the device bucket count (counting sort on the host), transpose to
    // word-major columns.
    const size_t cb = s->coeff_bytes();
    std::vector<uint64_t> keep_idx;
    keep_idx.reserve(n);
    for (uint64_t i = 0; i < n; ++i) {
        const bool nz = kind == PP_COEFF_F64 ? ((const double*)coeff)[i] != 0.0 : ((const float*)coeff)[i] != 0.0f;
        if (nz) keep_idx.push_back(i);
    }
    const uint64_t m = keep_idx.size();
    s->len = m;
    s->reference_bits = std::min(desired_bits(m, 1024, 128), s->max_bits);
    const uint64_t target = ctx->tuning.bucket_target;
    int dbits = 0;
    while (dbits < s->max_bits && m > target * (1ull << dbits)) ++dbits;
    s->device_bits = dbits;
    const uint32_t B = 1u << dbits;
    // The bucket of every kept term, then a stable counting sort by bucket, on up to 16 host threads (the same
    // layout as one thread would make).
    const unsigned T = host_threads(m);
    std::vector<uint32_t> bucket(m);
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(m, t, TT, lo, hi);
        for (uint64_t k = lo; k < hi; ++k) {
            const uint64_t i = keep_idx[k];
            bucket[k] = bucket_of_host(s, x + i * width, z + i * width, dbits);
        }
    });
    std::vector<uint64_t> order, bstart;
    parallel_counting_sort(bucket.data(), m, B, T, order, bstart);
    std::vector<uint64_t> start(B);
    std::vector<uint32_t> lens(B), room(B);
    uint64_t acc = 0;
    uint32_t maxb = 0;
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t cnt = (uint32_t)(bstart[b + 1] - bstart[b]);
        start[b] = acc;
        lens[b] = cnt;
        // a quarter more room, for the partners a first layer in place adds to the bucket
        room[b] = (cnt + cnt / 4 + (kSegAlign - 1)) & ~(kSegAlign - 1);
        acc += room[b];
        maxb = std::max(maxb, cnt);
    }
    s->max_bucket = maxb;
    if ((st = ensure_columns(s, 0, std::max<uint64_t>(acc, 1))) != PP_OK) return bail(st);
    if ((st = ensure_buckets(s, 0, B)) != PP_OK) return bail(st);
    const uint64_t cap = s->col[0].cap;
    // host staging of the word-major columns, left uninitialised: the rows between segments are never read
    std::unique_ptr<uint64_t[]> hx(new (std::nothrow) uint64_t[(size_t)width * cap]);
    std::unique_ptr<uint64_t[]> hz(new (std::nothrow) uint64_t[(size_t)width * cap]);
    std::unique_ptr<unsigned char[]> hc(new (std::nothrow) unsigned char[cap * cb]);
    if (!hx || !hz || !hc) return bail(fail(ctx, PP_ERR_OUT_OF_MEMORY, "host staging allocation"));
    // word-major columns: bucket b's k-th term (in input order) at start[b] + k
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(B, t, TT, lo, hi);
        for (uint64_t b = lo; b < hi; ++b)
            for (uint64_t r = bstart[b]; r < bstart[b + 1]; ++r) {
                const uint64_t i = keep_idx[order[r]];
                const uint64_t pos = start[b] + (r - bstart[b]);
                for (uint32_t w = 0; w < width; ++w) {
                    hx[w * cap + pos] = x[i * width + w];
                    hz[w * cap + pos] = z[i * width + w];
                }
                std::memcpy(&hc[pos * cb], (const unsigned char*)coeff + i * cb, cb);
            }
    });
    // the rows in use of every word plane and the coefficients, through the pinned chunks
    for (uint32_t w = 0; w < width; ++w) {
        if ((st = staged_h2d(ctx, s->col[0].x + w * cap, hx.get() + w * cap, acc * 8)) != PP_OK ||
            (st = staged_h2d(ctx, s->col[0].z + w * cap, hz.get() + w * cap, acc * 8)) != PP_OK)
            return bail(st);
    }
    if ((st = staged_h2d(ctx, s->col[0].c, hc.get(), acc * cb)) != PP_OK) return bail(st);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(s->bk[0].start, start.data(), B * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->bk[0].len, lens.data(), B * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->bk[0].rows, room.data(), B * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(ctx, PP_ERR_DRIVER, "upload copy failed"));
    s->cur = 0;
    s->tail = acc;
    s->has_room = true;
    *out = s;
    return PP_OK;
}

pp_status pp_sum_free(pp_sum* s) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_sum_free");
    if (!s) return PP_OK;
    if (s->ctx)
        for (const pp_graph* g : s->ctx->graphs)
            for (const CapturedJoin& r : g->layers)
                if (r.src == s || r.dst == s)
                    return fail(s->ctx, PP_ERR_INVALID_ARGUMENT,
                                "pp_sum_free: a captured graph reads or writes this sum; free the graph first");
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    for (int i = 0; i < 2; ++i) {
        free_columns(s->col[i]);
        free_buckets(s->bk[i]);
    }
    if (s->d_rows) (void)hipFree(s->d_rows);
    if (s->d_stats) (void)hipFree(s->d_stats);
    if (s->h_stats) (void)hipHostFree(s->h_stats);
    if (s->h_lag) (void)hipHostFree(s->h_lag);
    if (s->d_lag_counters) (void)hipFree(s->d_lag_counters);
    if (s->d_topn) (void)hipFree(s->d_topn);
    if (s->d_redo) (void)hipFree(s->d_redo);
    if (s->d_score) (void)hipFree(s->d_score);
    if (s->d_hist) (void)hipFree(s->d_hist);
    if (s->d_tiles) (void)hipFree(s->d_tiles);
    for (void* p : s->scratch)
        if (p) (void)hipFree(p);
    delete s;
    return PP_OK;
}

pp_status pp_sum_clone(const pp_sum* src_c, pp_sum** out) {
    if (src_c) PP_NOT_CAPTURING(src_c->ctx, "pp_sum_clone");
    if (!src_c || !out) return PP_ERR_INVALID_ARGUMENT;
    pp_sum* src = const_cast<pp_sum*>(src_c);
    pp_context* ctx = src->ctx;
    pp_status rs = refresh(src);
    if (rs != PP_OK) return rs;
    auto* s = new (std::nothrow) pp_sum();
    if (!s) return fail(ctx, PP_ERR_OUT_OF_MEMORY, "host allocation");
    s->ctx = ctx;
    s->num_qubits = src->num_qubits;
    s->W = src->W;
    s->kind = src->kind;
    s->seed = src->seed;
    s->max_bits = src->max_bits;
    s->rows = src->rows;
    s->device_bits = src->device_bits;
    s->reference_bits = src->reference_bits;
    s->len = src->len;
    s->max_bucket = src->max_bucket;
    s->has_room = src->has_room;
    s->tail = src->tail;
    auto bail = [&](pp_status st) {
        pp_sum_free(s);
        return st;
    };
    pp_status st;
    if ((st = dev_alloc(ctx, (void**)&s->d_rows, s->rows.size() * 8)) != PP_OK) return bail(st);
    if ((st = dev_alloc(ctx, (void**)&s->d_stats, kStatSlots * sizeof(unsigned long long))) != PP_OK) return bail(st);
    if (hipHostMalloc((void**)&s->h_stats, kStatSlots * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return bail(fail(ctx, PP_ERR_OUT_OF_MEMORY, "pinned host allocation"));
    const Columns& sc = src->col[src->cur];
    const BucketArrays& sb = src->bk[src->cur];
    if ((st = ensure_columns(s, 0, sc.cap)) != PP_OK) return bail(st);
    if ((st = ensure_buckets(s, 0, 1ull << src->device_bits)) != PP_OK) return bail(st);
    const uint64_t B = 1ull << src->device_bits;
    const uint64_t cap = s->col[0].cap;
    hipStream_t stream = ctx->stream;
    // word-major columns: copy each word plane at the new capacity stride
    for (uint32_t w = 0; w < s->W; ++w) {
        if (hipMemcpyAsync(s->col[0].x + w * cap, sc.x + w * sc.cap, sc.cap * 8, hipMemcpyDeviceToDevice, stream) !=
                hipSuccess ||
            hipMemcpyAsync(s->col[0].z + w * cap, sc.z + w * sc.cap, sc.cap * 8, hipMemcpyDeviceToDevice, stream) !=
                hipSuccess)
            return bail(fail(ctx, PP_ERR_DRIVER, "clone copy failed"));
    }
    if (hipMemcpyAsync(s->col[0].c, sc.c, sc.cap * s->coeff_bytes(), hipMemcpyDeviceToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(s->bk[0].start, sb.start, B * 8, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(s->bk[0].len, sb.len, B * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
        (src->has_room &&
         hipMemcpyAsync(s->bk[0].rows, sb.rows, B * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess) ||
        hipMemcpyAsync(s->d_rows, s->rows.data(), s->rows.size() * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return bail(fail(ctx, PP_ERR_DRIVER, "clone copy failed"));
    s->cur = 0;
    *out = s;
    return PP_OK;
}

pp_status pp_sum_copy(pp_sum* dst, pp_sum* src) {
    if (src) PP_NOT_CAPTURING(src->ctx, "pp_sum_copy");
    if (!dst || !src || dst == src) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = src->ctx;
    if (dst->ctx != ctx || dst->W != src->W || dst->kind != src->kind || dst->num_qubits != src->num_qubits ||
        dst->seed != src->seed)
        return fail(ctx, PP_ERR_INVALID_ARGUMENT, "pp_sum_copy: the sums differ in shape or context");
    pp_status st;
    if ((st = refresh(src)) != PP_OK) return st;
    for (const pp_graph* g : ctx->graphs)
        for (const CapturedJoin& r : g->layers)
            if (r.dst == dst) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "pp_sum_copy: a live graph writes dst");
    const int d = dst->cur;
    const Columns& sc = src->col[src->cur];
    const BucketArrays& sb = src->bk[src->cur];
    const uint64_t B = 1ull << src->device_bits;
    // every row a segment of src may hold: the tail when known, else the whole column
    const uint64_t rows = src->has_room ? src->tail : sc.cap;
    if ((st = ensure_columns(dst, d, rows)) != PP_OK) return st;
    if ((st = ensure_buckets(dst, d, B)) != PP_OK) return st;
    const Columns& dc = dst->col[d];
    const BucketArrays& db = dst->bk[d];
    hipStream_t stream = ctx->stream;
    for (uint32_t w = 0; w < src->W; ++w) {
        PP_HIP(ctx, hipMemcpyAsync(dc.x + w * dc.cap, sc.x + w * sc.cap, rows * 8, hipMemcpyDeviceToDevice, stream));
        PP_HIP(ctx, hipMemcpyAsync(dc.z + w * dc.cap, sc.z + w * sc.cap, rows * 8, hipMemcpyDeviceToDevice, stream));
    }
    PP_HIP(ctx, hipMemcpyAsync(dc.c, sc.c, rows * src->coeff_bytes(), hipMemcpyDeviceToDevice, stream));
    PP_HIP(ctx, hipMemcpyAsync(db.start, sb.start, B * 8, hipMemcpyDeviceToDevice, stream));
    PP_HIP(ctx, hipMemcpyAsync(db.len, sb.len, B * 4, hipMemcpyDeviceToDevice, stream));
    if (src->has_room) PP_HIP(ctx, hipMemcpyAsync(db.rows, sb.rows, B * 4, hipMemcpyDeviceToDevice, stream));
    dst->pending_topn = false;
    dst->pending_flags = nullptr;
    dst->pending_score = nullptr;
    dst->phys_len = 0;
    dst->stale = false;
    dst->len = src->len;
    dst->max_bucket = src->max_bucket;
    dst->device_bits = src->device_bits;
    dst->reference_bits = src->reference_bits;
    dst->has_room = src->has_room;
    dst->tail = src->tail;
    return PP_OK;
}

pp_status pp_sum_len(pp_sum* s, uint64_t* len) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_sum_len");
    if (!s || !len) return PP_ERR_INVALID_ARGUMENT;
    pp_status st = refresh_counts(s);  // exact also while a selection waits for its retain: len is the kept count
    if (st != PP_OK) return st;
    *len = s->len;
    return PP_OK;
}

pp_status pp_sum_check(pp_sum* s, uint64_t* out) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_sum_check");
    if (!s || !out) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = s->ctx;
    pp_status st;
    if ((st = refresh(s)) != PP_OK) return st;
    const uint32_t B = 1u << s->device_bits;
    const int cur = s->cur;
    if ((st = clear_stats(s)) != PP_OK) return st;
    if (s->len > 0 || s->col[cur].x) {
        hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int W = decltype(Wc)::value;
            using T = decltype(Tv);
            CheckArgs<W, T> a{};
            a.x = s->col[cur].x;
            a.z = s->col[cur].z;
            a.c = (const T*)s->col[cur].c;
            a.cap = s->col[cur].cap;
            a.start = s->bk[cur].start;
            a.len = s->bk[cur].len;
            a.room = s->has_room ? s->bk[cur].rows : nullptr;
            a.end = s->has_room ? s->tail : s->col[cur].cap;
            a.num_buckets = B;
            a.rows = s->d_rows;
            a.bits = s->device_bits;
            a.out = s->d_stats;
            return launch_check_layout<W, T>(ctx->stream, a);
        });
        PP_HIP(ctx, e);
    }
    if ((st = read_stats(s)) != PP_OK) return st;
    for (int i = 0; i < 5; ++i) out[i] = s->h_stats[i];
    // overlaps and the total, on the host from the bucket arrays
    std::vector<uint64_t> start(B);
    std::vector<uint32_t> len(B), room(B);
    PP_HIP(ctx, hipMemcpy(start.data(), s->bk[cur].start, (size_t)B * 8, hipMemcpyDeviceToHost));
    PP_HIP(ctx, hipMemcpy(len.data(), s->bk[cur].len, (size_t)B * 4, hipMemcpyDeviceToHost));
    if (s->has_room) PP_HIP(ctx, hipMemcpy(room.data(), s->bk[cur].rows, (size_t)B * 4, hipMemcpyDeviceToHost));
    std::vector<std::pair<uint64_t, uint64_t>> seg;
    uint64_t total = 0;
    for (uint32_t b = 0; b < B; ++b) {
        total += len[b];
        const uint64_t extent = s->has_room ? room[b] : len[b];
        if (extent) seg.push_back({start[b], extent});
    }
    std::sort(seg.begin(), seg.end());
    uint64_t overlaps = 0;
    for (size_t i = 1; i < seg.size(); ++i) overlaps += seg[i - 1].first + seg[i - 1].second > seg[i].first;
    out[5] = overlaps;
    out[6] = total != s->len;
    return PP_OK;
}

// PP_CHECK_INVARIANTS=1: every layer entry point checks the layout it leaves.
pp_status debug_check(pp_sum* s) {
    const char* e = std::getenv("PP_CHECK_INVARIANTS");
    if (!e || std::strcmp(e, "1") != 0) return PP_OK;
    uint64_t out[7] = {};
    pp_status st = pp_sum_check(s, out);
    if (st != PP_OK) return st;
    for (int i = 0; i < 7; ++i)
        if (out[i] && i != 4)
            return fail(s->ctx, PP_ERR_DRIVER,
                        "layout invariant " + std::to_string(i) + " violated " + std::to_string(out[i]) + " times");
    return PP_OK;
}

pp_status pp_sum_bucket_bits(pp_sum* s, int32_t* reference_bits, int32_t* device_bits) {
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    if (reference_bits) *reference_bits = s->reference_bits;
    if (device_bits) *device_bits = s->device_bits;
    return PP_OK;
}

pp_status pp_sum_download(pp_sum* s, uint64_t* x, uint64_t* z, void* coeff) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_sum_download");
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = s->ctx;
    {
        pp_status rs = refresh(s);
        if (rs != PP_OK) return rs;
    }
    const uint64_t n = s->len;
    if (n == 0) return PP_OK;
    if (!x || !z || !coeff) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "null output arrays");
    const uint32_t W = s->W;
    const size_t cb = s->coeff_bytes();
    const uint32_t B = 1u << s->device_bits;
    uint64_t *d_dense_start = nullptr, *dx = nullptr, *dz = nullptr;
    void* dc = nullptr;
    uint32_t* db = nullptr;
    pp_status st = PP_OK;
    auto cleanup = [&] {
        if (d_dense_start) (void)hipFree(d_dense_start);
        if (dx) (void)hipFree(dx);
        if (dz) (void)hipFree(dz);
        if (dc) (void)hipFree(dc);
        if (db) (void)hipFree(db);
    };
    if ((st = dev_alloc(ctx, (void**)&d_dense_start, (size_t)B * 8)) != PP_OK ||
        (st = dev_alloc(ctx, (void**)&dx, n * W * 8)) != PP_OK || (st = dev_alloc(ctx, (void**)&dz, n * W * 8)) != PP_OK ||
        (st = dev_alloc(ctx, &dc, n * cb)) != PP_OK || (st = dev_alloc(ctx, (void**)&db, n * 4)) != PP_OK) {
        cleanup();
        return st;
    }
    if ((st = offsets(s, s->bk[s->cur].len, B, 0, 0, 0, 0, d_dense_start)) != PP_OK) {
        cleanup();
        return st;
    }
    hipError_t e = hipSuccess;
    if (e == hipSuccess)
        e = dispatch(s, [&](auto Wc, auto Tv) {
            constexpr int WW = decltype(Wc)::value;
            using T = decltype(Tv);
            ExportArgs<WW, T> a{};
            a.x = s->col[s->cur].x;
            a.z = s->col[s->cur].z;
            a.c = (const T*)s->col[s->cur].c;
            a.cap = s->col[s->cur].cap;
            a.start = s->bk[s->cur].start;
            a.len = s->bk[s->cur].len;
            a.dense_start = d_dense_start;
            a.rows = s->d_rows;
            a.ref_bits = s->reference_bits;
            a.out_x = dx;
            a.out_z = dz;
            a.out_c = (T*)dc;
            a.out_bucket = db;
            return launch_export<WW, T>(ctx->stream, a, B);
        });
    // host staging, left uninitialised: the copies fill every element
    std::unique_ptr<uint64_t[]> hx(new (std::nothrow) uint64_t[n * W]), hz(new (std::nothrow) uint64_t[n * W]);
    std::unique_ptr<unsigned char[]> hc(new (std::nothrow) unsigned char[n * cb]);
    std::unique_ptr<uint32_t[]> hb(new (std::nothrow) uint32_t[n]);
    if (!hx || !hz || !hc || !hb) {
        cleanup();
        return fail(ctx, PP_ERR_OUT_OF_MEMORY, "host staging allocation");
    }
    if (e != hipSuccess) {
        cleanup();
        return fail(ctx, PP_ERR_DRIVER, std::string("download: ") + hipGetErrorString(e));
    }
    // through the pinned chunks (each call drains its own copies)
    if ((st = staged_d2h(ctx, hx.get(), dx, n * W * 8)) != PP_OK || (st = staged_d2h(ctx, hz.get(), dz, n * W * 8)) != PP_OK ||
        (st = staged_d2h(ctx, hc.get(), dc, n * cb)) != PP_OK || (st = staged_d2h(ctx, hb.get(), db, n * 4)) != PP_OK) {
        cleanup();
        return st;
    }
    cleanup();
    // canonical order: the reference's bucket ascending (counting sort), then lexicographic key per bucket; each
    // step on up to 16 host threads (keys are distinct, so the order is the same however the work is split)
    const uint32_t RB = 1u << s->reference_bits;
    const unsigned T = host_threads(n);
    std::vector<uint64_t> order, rstart;
    parallel_counting_sort(hb.get(), n, RB, T, order, rstart);
    auto less = [&](uint64_t a, uint64_t b) {
        for (uint32_t w = 0; w < W; ++w)
            if (hx[a * W + w] != hx[b * W + w]) return hx[a * W + w] < hx[b * W + w];
        for (uint32_t w = 0; w < W; ++w)
            if (hz[a * W + w] != hz[b * W + w]) return hz[a * W + w] < hz[b * W + w];
        return false;
    };
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(RB, t, TT, lo, hi);
        for (uint64_t b = lo; b < hi; ++b) std::sort(order.begin() + rstart[b], order.begin() + rstart[b + 1], less);
    });
    host_parallel(T, [&](unsigned t, unsigned TT) {
        uint64_t lo, hi;
        part_range(n, t, TT, lo, hi);
        for (uint64_t k = lo; k < hi; ++k) {
            const uint64_t i = order[k];
            for (uint32_t w = 0; w < W; ++w) {
                x[k * W + w] = hx[i * W + w];
                z[k * W + w] = hz[i * W + w];
            }
            std::memcpy((unsigned char*)coeff + k * cb, &hc[i * cb], cb);
        }
    });
    return PP_OK;
}

pp_status pp_apply_rotation(pp_sum* s, const pp_channel* ch, int adjoint, const pp_policy* policy) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_apply_rotation");
    if (!s || !ch) return PP_ERR_INVALID_ARGUMENT;
    if (!valid_policy(policy)) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    pp_status st = refresh(s);
    if (st != PP_OK) return st;
    const Prepared p = prepare(s, *ch, adjoint != 0);
    if ((st = apply_one(s, p, lower_policy(policy))) != PP_OK) return st;
    return debug_check(s);
}

pp_status pp_apply_rotation_to(pp_sum* src, pp_sum* dst, const pp_channel* ch, int adjoint, const pp_policy* policy) {
    if (!src || !dst || !ch || src == dst) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = src->ctx;
    if (dst->ctx != ctx || dst->W != src->W || dst->kind != src->kind || dst->num_qubits != src->num_qubits ||
        dst->seed != src->seed)
        return fail(ctx, PP_ERR_INVALID_ARGUMENT, "pp_apply_rotation_to: the sums differ in shape or context");
    if (!valid_policy(policy)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    pp_status st;
    if (ctx->capturing && (src->stale || src->pending_topn))
        return fail(ctx, PP_ERR_UNSUPPORTED,
                    "pp_apply_rotation_to: the input's counts or a selection are pending during capture");
    if ((st = refresh(src)) != PP_OK) return st;
    if ((st = settle_device_bits(src)) != PP_OK) return st;
    const Prepared p = prepare(src, *ch, adjoint != 0);
    if (p.key_preserving)
        return fail(ctx, PP_ERR_UNSUPPORTED, "pp_apply_rotation_to: a key-preserving channel runs in place");
    // dst is overwritten whole: a selection it still had pending no longer describes its rows
    dst->pending_topn = false;
    dst->pending_flags = nullptr;
    dst->pending_score = nullptr;
    dst->phys_len = 0;
    dst->stale = false;
    dst->device_bits = src->device_bits;
    dst->reference_bits = src->reference_bits;
    const int out = dst->cur;
    if ((st = launch_join(src, dst, out, p, lower_policy(policy))) != PP_OK) return st;
    dst->stale = true;
    if (ctx->capturing) {
        const int in = src->cur;
        CapturedJoin r{};
        r.src = src;
        r.dst = dst;
        r.src_cur = in;
        r.dst_cur = out;
        r.device_bits = src->device_bits;
        const void* sp[5] = {src->col[in].x, src->col[in].z, src->col[in].c, src->bk[in].start, src->bk[in].len};
        const void* dp[7] = {dst->col[out].x, dst->col[out].z, dst->col[out].c, dst->bk[out].start,
                             dst->bk[out].len, dst->d_stats, dst->d_redo};
        std::copy(sp, sp + 5, r.src_ptrs);
        std::copy(dp, dp + 7, r.dst_ptrs);
        r.dst_cap = dst->col[out].cap;
        r.src_len = src->len;
        r.src_max_bucket = src->max_bucket;
        ctx->capture_log.push_back(r);
    }
    return PP_OK;
}

pp_status pp_apply_rotation_to_noalloc(pp_sum* src, pp_sum* dst, const pp_channel* ch, int adjoint,
                                       const pp_policy* policy) {
    if (!src || !dst || !ch || src == dst) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = src->ctx;
    PP_NOT_CAPTURING(ctx, "pp_apply_rotation_to_noalloc");
    if (dst->ctx != ctx || dst->W != src->W || dst->kind != src->kind || dst->num_qubits != src->num_qubits ||
        dst->seed != src->seed)
        return fail(ctx, PP_ERR_INVALID_ARGUMENT, "pp_apply_rotation_to_noalloc: the sums differ in shape or context");
    if (!valid_policy(policy)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    unsigned char* base = noalloc_scratch(ctx->device);
    if (!base) return fail(ctx, PP_ERR_UNSUPPORTED, "pp_apply_rotation_to_noalloc: no scratch on this device");
    pp_status st;
    if ((st = refresh(src)) != PP_OK) return st;  // a read-back or an in-place retain: neither allocates
    if (!device_bits_settled(src))
        return fail(ctx, PP_ERR_UNSUPPORTED,
                    "pp_apply_rotation_to_noalloc: the input's device bucketing must be refined first, which "
                    "allocates (one pp_apply_rotation_to from it does so)");
    const uint32_t B = 1u << src->device_bits;
    if (B > kNoAllocMaxBuckets)
        return fail(ctx, PP_ERR_INVALID_ARGUMENT,
                    "pp_apply_rotation_to_noalloc: " + std::to_string(B) + " device buckets exceed the scratch's " +
                        std::to_string(kNoAllocMaxBuckets));
    const Prepared p = prepare(src, *ch, adjoint != 0);
    if (p.key_preserving)
        return fail(ctx, PP_ERR_UNSUPPORTED, "pp_apply_rotation_to_noalloc: a key-preserving channel runs in place");
    const int out = dst->cur;
    if (!dst->col[out].x || dst->col[out].cap < join_output_rows(src) || !dst->bk[out].start || dst->bk[out].cap < B)
        return fail(ctx, PP_ERR_INVALID_ARGUMENT,
                    "pp_apply_rotation_to_noalloc: the output has room for " + std::to_string(dst->col[out].cap) +
                        " rows and " + std::to_string(dst->bk[out].cap) + " buckets; this layer needs " +
                        std::to_string(join_output_rows(src)) + " and " + std::to_string(B) +
                        " (size it with one pp_apply_rotation_to of the same input)");
    dst->pending_topn = false;
    dst->pending_flags = nullptr;
    dst->pending_score = nullptr;
    dst->phys_len = 0;
    dst->stale = false;
    dst->device_bits = src->device_bits;
    dst->reference_bits = src->reference_bits;
    const JoinScratch fixed{(uint64_t*)(base + kNoAllocTilesOffset),
                            (unsigned long long*)(base + kNoAllocStatsOffset), (uint32_t*)(base + kNoAllocRedoOffset)};
    if ((st = launch_join(src, dst, out, p, lower_policy(policy), &fixed)) != PP_OK) return st;
    // the counters where pp_sum_len(dst) reads them, in stream order
    PP_HIP(ctx, hipMemcpyAsync(dst->d_stats, fixed.stats, kStatSlots * sizeof(unsigned long long),
                               hipMemcpyDeviceToDevice, ctx->stream));
    dst->stale = true;
    return PP_OK;
}

// ---- the layer seam ------------------------------------------------------------------------------------------------

pp_status pp_prepare_channel(uint32_t num_qubits, uint32_t width, pp_coeff_kind kind, const pp_channel* ch,
                             int adjoint, pp_prepared* out) {
    if (!ch || !out || num_qubits == 0 || num_qubits > 128 || (width != 1 && width != 2) ||
        (width == 1) != (num_qubits <= 64) || (kind != PP_COEFF_F64 && kind != PP_COEFF_F32))
        return PP_ERR_INVALID_ARGUMENT;
    for (uint32_t w = 0; w < 4; ++w) {
        const uint64_t live = w < width ? live_mask(num_qubits, w) : 0;
        if ((ch->gx[w] | ch->gz[w]) & ~live) return PP_ERR_INVALID_ARGUMENT;
    }
    pp_sum shape;  // the preparation reads only the register's width and the coefficient type
    shape.num_qubits = num_qubits;
    shape.W = width;
    shape.kind = kind;
    *out = prepare_channel_table(&shape, *ch, adjoint != 0);
    return PP_OK;
}

pp_status pp_apply_prepared(pp_sum* s, const pp_prepared* prepared, const pp_policy* policy) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_apply_prepared");
    if (!s || !prepared) return PP_ERR_INVALID_ARGUMENT;
    if (!valid_policy(policy)) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    pp_status st;
    if ((st = refresh_counts(s)) != PP_OK) return st;
    Prepared p;
    if ((st = classify(s, *prepared, &p)) != PP_OK) return st;
    if ((st = apply_one(s, p, lower_policy(policy))) != PP_OK) return st;
    return debug_check(s);
}

pp_status pp_apply_run(pp_sum* s, pp_run_kind kind, const pp_prepared* layers, uint32_t n, const pp_policy* policy,
                       uint64_t* layer_len) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_apply_run");
    if (!s || (n && !layers)) return PP_ERR_INVALID_ARGUMENT;
    if (!valid_policy(policy)) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    if (kind != PP_RUN_ROTATIONS && kind != PP_RUN_PERMUTATIONS)
        return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "run kind");
    if (kind == PP_RUN_ROTATIONS && n > kRunLayers)
        return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "a rotation run holds at most 8 layers");
    pp_status st;
    if ((st = refresh_counts(s)) != PP_OK) return st;
    std::vector<Prepared> preps(n);
    for (uint32_t j = 0; j < n; ++j) {
        if ((st = classify(s, layers[j], &preps[j])) != PP_OK) return st;
        const bool fits = kind == PP_RUN_PERMUTATIONS ? is_permutation_step(preps[j]) : is_rotation_table(preps[j]);
        if (!fits)
            return fail(s->ctx, PP_ERR_INVALID_ARGUMENT,
                        "run layer " + std::to_string(j) +
                            (kind == PP_RUN_PERMUTATIONS ? " is not a one-to-one step" : " is not rotation-shaped"));
    }
    const KeepProgram keep = lower_policy(policy);
    if (kind == PP_RUN_PERMUTATIONS && n > 0 && s->len > 0) {
        if ((st = permutation_run_device(s, preps.data(), n, keep, layer_len)) != PP_OK) return st;
        return debug_check(s);
    }
    if (n >= kClassMinRunLayers && s->len > 0 && class_rotation_runs()) {
        bool done = false;
        if ((st = class_pass(s, preps.data(), n, keep, layer_len, &done)) != PP_OK) return st;
        if (done) return debug_check(s);
    }
    for (uint32_t j = 0; j < n; ++j) {
        if (s->len == 0) {  // nothing left to transform: the remaining layers keep it empty
            if (layer_len) layer_len[j] = 0;
            continue;
        }
        if ((st = apply_one(s, preps[j], keep)) != PP_OK) return st;
        if (layer_len) layer_len[j] = s->len;
    }
    return debug_check(s);
}

pp_status pp_finalize_layer(pp_sum* s, const pp_policy* policy) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_finalize_layer");
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    if (!valid_policy(policy)) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    if (!policy || !policy->has_topn) return PP_OK;
    pp_status st;
    if ((st = refresh_counts(s)) != PP_OK) return st;
    if (policy->has_topn == 2) return pp_truncate_approx_topn(s, policy->topn);
    if (s->pending_topn && (st = materialize_topn(s)) != PP_OK) return st;  // two passes in a row
    if (s->len <= policy->topn) return PP_OK;
    return topn_device(s, policy->topn, nullptr, true);
}

pp_status pp_settle(pp_sum* s) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_settle");
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    return refresh(s);
}

pp_status pp_sum_refine(pp_sum* s) {
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    if (s->reference_bits >= s->max_bits)
        return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "refine: the hash has no further row");
    s->reference_bits += 1;
    return PP_OK;
}

pp_status pp_sum_coarsen_to(pp_sum* s, int32_t bits) {
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    if (bits < 0 || bits > s->reference_bits)
        return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "coarsen_to: bits above the current count");
    s->reference_bits = bits;
    return PP_OK;
}

namespace {

// ---- lagged lengths (engine/gpu/length.rs:1-4) ------------------------------------------------------------------
//
// A small rotation layer's exact length and longest bucket reach the host one layer late.  Each queued layer's
// counters reach a pinned slot of their own: a small join's from the next launch's extra workgroup (or, before a
// drain, a one-workgroup publication), a larger layer's by a copy behind its kernels.  The host waits for the
// previous layer's slot while the current layer runs, so the GPU is not left idle across a host round trip per
// layer.  The next layer is sized from
// bounds: a rotation layer at most doubles the rows, and an output bucket draws on two input buckets, so at most
// doubles the longest bucket.  A layer whose bounds do not fit (a pair past the largest a join stages, a column set
// that must grow), a sum large enough for the in-place layer, or a table that is not a rotation drains the queue and
// takes the synchronous path.  The results are the same either way: only when the host learns the counts changes.

constexpr uint64_t kLagMaxRows = 1ull << 20;  // above this a layer is bandwidth-bound and its synchronization noise
constexpr uint32_t kLagDepth = 1;             // layers whose counts may still be on their way when one is queued

// Whether pp_propagate queues small layers ahead of their counts (default) or reads every layer's counts before the
// next (PP_LAGGED_LENGTHS=0, for comparison: the same keys, coefficients and per-layer lengths).
bool lagged_lengths() {
    const char* e = std::getenv("PP_LAGGED_LENGTHS");
    return !(e && std::strcmp(e, "0") == 0);
}

// Whether layer p from s can take the one-launch small join: a layout of at most kSmallJoinBuckets buckets, every
// pair staged whole (no redo launch), and no measurement (which times the usual launches).
bool small_join_fits(const pp_sum* s, const Prepared& p) {
    const uint32_t B = 1u << s->device_bits;
    if (B > kSmallJoinBuckets || s->ctx->profiling) return false;
    const uint32_t delta = bucket_of_host(s, p.gx, p.gz, s->device_bits);
    const uint32_t lcap = join_pair_bound(s, delta);
    return join_stage_rows(s->ctx->tuning.stage_rows, lcap) >= lcap;
}

// Queue the small join of layer p from s's current columns into column set `out`, adding into `counters` (zero);
// its extra workgroup publishes `prev` (the counters of the small join queued just before, or null) to the pinned
// slot whose device address is `prev_host`, followed by `prev_seq`.
pp_status launch_join_small(pp_sum* s, int out, const Prepared& p, const KeepProgram& keep,
                            unsigned long long* counters, unsigned long long* prev, unsigned long long* prev_host,
                            unsigned long long prev_seq) {
    pp_context* ctx = s->ctx;
    const uint32_t B = 1u << s->device_bits;
    const uint32_t delta = bucket_of_host(s, p.gx, p.gz, s->device_bits);
    const uint32_t lcap = join_pair_bound(s, delta);
    hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        SmallJoinArgs<W, T> a{};
        a.j = join_args<W, T>(s, s, out, p, keep, delta, lcap);
        a.j.stats = counters;
        a.j.count_totals = 1;
        a.orows = s->bk[out].rows;
        a.num_buckets = B;
        a.prev_counters = prev;
        a.prev_host = prev_host;
        a.prev_seq = prev_seq;
        return launch_rotation_join_small<W, T>(ctx->stream, a);
    });
    PP_HIP(ctx, e);
    return PP_OK;
}

// run_layers under a policy with no layer pass, its rotation layers queued ahead of their counts.  The planner's
// bookkeeping (the reference's bucket count at every run start) runs as each length arrives.
pp_status propagate_lagged(pp_sum* s, const RunPlanner& plan, const KeepProgram& keep, uint64_t* layer_len) {
    pp_context* ctx = s->ctx;
    pp_status st;
    for (hipEvent_t& e : ctx->lag_ev)
        if (!e) PP_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!s->h_lag) {
        if (hipHostMalloc((void**)&s->h_lag, 2 * kStatSlots * sizeof(unsigned long long),
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return fail(ctx, PP_ERR_OUT_OF_MEMORY, "pinned host allocation");
        std::memset(s->h_lag, 0, 2 * kStatSlots * sizeof(unsigned long long));
        PP_HIP(ctx, hipHostGetDevicePointer((void**)&s->d_lag_host, s->h_lag, 0));
    }
    if (!s->d_lag_counters &&
        (st = dev_alloc(ctx, (void**)&s->d_lag_counters, 8 * sizeof(unsigned long long))) != PP_OK)
        return st;
    // zero from here on, each set cleared when its counters are published (and here, in case a propagation that
    // failed left one unpublished)
    PP_HIP(ctx, launch_fill_u64(ctx->stream, s->d_lag_counters, 8, 0ull));
    const uint64_t n = plan.n_channels;
    const uint64_t target = ctx->tuning.bucket_target;
    uint64_t i = 0, run_end = 0;  // layer i is the next to queue; the planner's current run ends at run_end
    uint32_t queued = 0, head = 0;  // layers whose counters are on their way; ring slot of the oldest
    uint64_t queued_layer[2] = {};
    unsigned long long queued_seq[2] = {};  // a small join's sequence number in its slot, or 0 (a copied slot)
    int unpublished = -1;  // the slot of a small join whose counters no launch publishes yet
    auto counters_of = [&](int k) { return s->d_lag_counters + 4 * k; };
    auto host_of = [&](int k) { return s->d_lag_host + kStatSlots * k; };
    // the last small join's counters, when no small join follows it to publish them
    auto publish = [&]() -> pp_status {
        if (unpublished < 0) return PP_OK;
        PP_HIP(ctx, launch_publish_counters(ctx->stream, counters_of(unpublished), host_of(unpublished),
                                            queued_seq[unpublished]));
        unpublished = -1;
        return PP_OK;
    };
    uint64_t known_len = s->len, known_mb = s->max_bucket;  // counts after the last layer read
    // the planner's bookkeeping at layer r, once the length before it is known
    auto plan_at = [&](uint64_t r, uint64_t len_before) {
        if (r != run_end || r >= n) return;
        const uint64_t len_now = s->len;
        s->len = len_before;
        run_end = plan.start_run(s, r);
        s->len = len_now;
    };
    auto read_oldest = [&]() -> pp_status {
        const uint32_t k = head;
        volatile const unsigned long long* h = s->h_lag + kStatSlots * k;
        if (queued_seq[k]) {
            // the next launch (a small join, or the publication at a drain) writes the slot, then the sequence
            // number: wait for it on the host, and after a long wait for the stream itself (which reports a failed
            // launch)
            const auto t0 = std::chrono::steady_clock::now();
            while (h[4] != queued_seq[k])
                if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
                    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
                    if (h[4] != queued_seq[k]) return fail(ctx, PP_ERR_DRIVER, "small join: counters never arrived");
                    break;
                }
        } else {
            PP_HIP(ctx, hipEventSynchronize(ctx->lag_ev[k]));
        }
        if (h[2]) return fail(ctx, PP_ERR_DRIVER, "rotation join: a pair exceeded the largest-pair bound");
        known_len = h[0];
        known_mb = h[1];
        s->tail = h[3];
        head ^= 1u;
        --queued;
        if (layer_len) layer_len[queued_layer[k]] = known_len;
        plan_at(queued_layer[k] + 1, known_len);
        return PP_OK;
    };
    auto drain = [&]() -> pp_status {
        pp_status ps = publish();
        if (ps != PP_OK) return ps;
        while (queued) {
            pp_status rs = read_oldest();
            if (rs != PP_OK) return rs;
        }
        s->len = known_len;
        s->max_bucket = (uint32_t)known_mb;
        return PP_OK;
    };
    while (i < n) {
        if (queued == 0 && i == run_end) run_end = plan.start_run(s, i);
        const Prepared p = plan.prepared(s, i);
        bool lag = p.rotation && !p.key_preserving;
        if (lag && is_permutation_step(p) && i + 1 < n &&
            is_permutation_step(plan.prepared(s, i + 1)))
            lag = false;  // the start of a permutation run: walked in one pass
        if (lag && queued == 0) {
            // exact counts: the device bucketing for the layers to come, fine enough that a pair stays within the
            // join's largest stage under the doubled bound (but not so fine that the next settle coarsens it back)
            if ((st = settle_device_bits(s)) != PP_OK) return st;
            while (4ull * s->max_bucket > kMaxPairRows && s->device_bits < s->max_bits &&
                   s->len * 8 >= target << (s->device_bits + 1))
                if ((st = refine_once(s)) != PP_OK) return st;
            known_len = s->len;
            known_mb = s->max_bucket;
        }
        const uint64_t len_hi = known_len << queued, mb_hi = known_mb << queued;
        lag = lag && len_hi <= kLagMaxRows && 2 * mb_hi <= kMaxPairRows &&
              (!ctx->tuning.join_in_place || len_hi < in_place_min_rows());
        const int out = 1 - s->cur;
        const uint32_t B = 1u << s->device_bits;
        const uint64_t need = 2 * len_hi + (uint64_t)kSegAlign * B;
        // with nothing queued, this column set also takes room for a layer queued behind one more (the doubled
        // bound), so that the queue does not drain again at the next layer to write here
        const uint64_t want = queued ? need : 4 * len_hi + (uint64_t)kSegAlign * B;
        if (lag && (s->col[out].cap < want || !s->col[out].x || s->bk[out].cap < B || !s->bk[out].start)) {
            // the output columns must grow: not under a queued layer, which may still read them
            if (queued) {
                if ((st = drain()) != PP_OK) return st;
                continue;
            }
            if ((st = ensure_columns(s, out, 2 * want)) != PP_OK) return st;  // room for the layers that follow
            if ((st = ensure_buckets(s, out, B)) != PP_OK) return st;
        }
        if (!lag) {
            if ((st = drain()) != PP_OK) return st;  // the planner has now seen every length before layer i
            if (i == run_end) run_end = plan.start_run(s, i);
            const uint64_t end = run_end;
            if ((st = plan.apply_run(s, i, end, keep, layer_len ? layer_len + i : nullptr)) != PP_OK) return st;
            if ((st = plan.finish_run(s, end < n)) != PP_OK) return st;
            if ((st = debug_check(s)) != PP_OK) return st;
            known_len = s->len;
            known_mb = s->max_bucket;
            i = end;
            continue;
        }
        // queue layer i, sized from the bounds; its counters follow it to the next ring slot
        s->len = len_hi;
        s->max_bucket = (uint32_t)mb_hi;
        const uint32_t k = (head + queued) & 1u;
        if (small_join_fits(s, p)) {  // one launch: the join lays out its output, and publishes the one before
            const int prev = unpublished == (int)(k ^ 1u) ? (int)(k ^ 1u) : -1;
            if (unpublished >= 0 && prev < 0 && (st = publish()) != PP_OK) return st;
            queued_seq[k] = ++s->lag_seq;
            if ((st = launch_join_small(s, out, p, keep, counters_of((int)k), prev >= 0 ? counters_of(prev) : nullptr,
                                        prev >= 0 ? host_of(prev) : nullptr,
                                        prev >= 0 ? queued_seq[prev] : 0ull)) != PP_OK)
                return st;
            unpublished = (int)k;
        } else {
            if ((st = publish()) != PP_OK) return st;
            if ((st = launch_join(s, s, out, p, keep)) != PP_OK) return st;
            PP_HIP(ctx, hipMemcpyAsync(s->h_lag + kStatSlots * k, s->d_stats, kStatSlots * sizeof(unsigned long long),
                                       hipMemcpyDeviceToHost, ctx->stream));
            PP_HIP(ctx, hipEventRecord(ctx->lag_ev[k], ctx->stream));  // a small join's slot carries its own mark
            queued_seq[k] = 0;
        }
        queued_layer[k] = i;
        ++queued;
        s->cur = out;
        s->has_room = true;
        s->pending_topn = false;
        ++i;
        if (queued > kLagDepth && (st = read_oldest()) != PP_OK) return st;
    }
    if ((st = drain()) != PP_OK) return st;
    return debug_check(s);
}

}  // namespace

pp_status pp_propagate(pp_sum* s, const pp_channel* channels, uint64_t n_channels, pp_direction direction,
                       const pp_policy* policy, const pp_options* options, uint64_t* layer_len) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_propagate");
    if (!s || (n_channels && !channels)) return PP_ERR_INVALID_ARGUMENT;
    {
        pp_status rs = refresh(s);
        if (rs != PP_OK) return rs;
    }
    if (!valid_policy(policy)) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    const size_t target = options ? options->target_bucket_len : 1024;
    const size_t min_buckets = options ? options->min_buckets : 128;
    if (target < 1 || min_buckets < 1) return fail(s->ctx, PP_ERR_INVALID_ARGUMENT, "bucket options");
    pp_status st;
    if ((st = check_channels(s, channels, n_channels)) != PP_OK) return st;
    const RunPlanner plan{channels, n_channels, direction == PP_HEISENBERG, policy, target, min_buckets};
    const KeepProgram keep = lower_policy(policy);
    if (!plan.finalizes && !class_rotation_runs() && lagged_lengths())
        return propagate_lagged(s, plan, keep, layer_len);
    uint64_t i = 0, run_end = 0;  // layer i belongs to the planner's run ending at run_end
    std::vector<Prepared> window;
    std::vector<uint64_t> lens;
    while (i < n_channels) {
        if (i == run_end) run_end = plan.start_run(s, i);
        // rotation tables across the planner's run boundaries in one class pass, under a policy with no layer pass
        const uint64_t e = class_rotation_runs() && s->len > 0 ? plan.window_end(s, i, &window) : i;
        if (e - i >= kClassMinRunLayers) {
            lens.assign(e - i, 0);
            bool done = false;
            if ((st = class_pass(s, window.data(), (uint32_t)(e - i), keep, lens.data(), &done)) != PP_OK) return st;
            if (done) {
                const uint64_t len_now = s->len;
                for (uint64_t j = i + 1; j < e; ++j)
                    if (j == run_end) {  // a run starts inside the window: its bucket count from the length there
                        s->len = lens[j - 1 - i];
                        run_end = plan.start_run(s, j);
                    }
                s->len = len_now;
                if (layer_len) std::copy(lens.begin(), lens.end(), layer_len + i);
                if ((st = debug_check(s)) != PP_OK) return st;
                i = e;
                continue;
            }
        }
        const uint64_t end = run_end;  // the planner's run, or what a window left of it
        if ((st = plan.apply_run(s, i, end, keep, layer_len ? layer_len + i : nullptr)) != PP_OK) return st;
        if ((st = plan.finish_run(s, end < n_channels)) != PP_OK) return st;
        if (layer_len && plan.finalizes) layer_len[end - 1] = s->len;
        if ((st = debug_check(s)) != PP_OK) return st;
        i = end;
    }
    return PP_OK;
}

pp_status pp_truncate_topn(pp_sum* s, uint64_t n) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_truncate_topn");
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    pp_status rs = refresh(s);
    if (rs != PP_OK) return rs;
    if (s->len <= n) return PP_OK;
    return topn_device(s, n);
}

pp_status pp_truncate_approx_topn(pp_sum* s, uint64_t n) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_truncate_approx_topn");
    if (!s) return PP_ERR_INVALID_ARGUMENT;
    pp_status rs = refresh(s);
    if (rs != PP_OK) return rs;
    if (s->len <= n) return PP_OK;  // EdgeDecision::KeepAll
    return approx_topn_device(s, n);
}

pp_status RunPlanner::finish_run(pp_sum* s, bool defer) const {
    if (!finalizes) return PP_OK;
    if (policy->has_topn == 2) return pp_truncate_approx_topn(s, policy->topn);
    if (!defer) return pp_truncate_topn(s, policy->topn);
    if (s->len <= policy->topn) return PP_OK;
    return topn_device(s, policy->topn, nullptr, true);
}

pp_status pp_expectation_product(pp_sum* s, const uint64_t* axis_x, const uint64_t* axis_z, const uint64_t* negative,
                                 double* out) {
    if (s) PP_NOT_CAPTURING(s->ctx, "pp_expectation_product");
    if (!s || !axis_x || !axis_z || !negative || !out) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = s->ctx;
    {
        pp_status rs = refresh(s);
        if (rs != PP_OK) return rs;
    }
    const uint32_t B = 1u << s->device_bits;
    const uint32_t per = 64;
    const uint32_t blocks = (B + per - 1) / per;
    double* d_partial = nullptr;
    pp_status st;
    if ((st = dev_alloc(ctx, (void**)&d_partial, blocks * 8)) != PP_OK) return st;
    hipError_t e = dispatch(s, [&](auto Wc, auto Tv) {
        constexpr int W = decltype(Wc)::value;
        using T = decltype(Tv);
        ExpectationArgs<W, T> a{};
        a.x = s->col[s->cur].x;
        a.z = s->col[s->cur].z;
        a.c = (const T*)s->col[s->cur].c;
        a.cap = s->col[s->cur].cap;
        a.start = s->bk[s->cur].start;
        a.len = s->bk[s->cur].len;
        a.num_buckets = B;
        a.buckets_per_block = per;
        for (int w = 0; w < W; ++w) {
            a.axis_x[w] = axis_x[w];
            a.axis_z[w] = axis_z[w];
            a.negative[w] = negative[w];
        }
        a.partial = d_partial;
        return launch_expectation<W, T>(ctx->stream, a, blocks);
    });
    std::vector<double> partial(blocks);
    if (e == hipSuccess) e = hipMemcpyAsync(partial.data(), d_partial, blocks * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_partial);
    if (e != hipSuccess) return fail(ctx, PP_ERR_DRIVER, std::string("expectation: ") + hipGetErrorString(e));
    double total = 0.0;
    for (double v : partial) total += v;
    *out = total;
    return PP_OK;
}

pp_status pp_profile_enable(pp_context* ctx, int enable) {
    PP_NOT_CAPTURING(ctx, "pp_profile_enable");
    if (!ctx) return PP_ERR_INVALID_ARGUMENT;
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& e : ctx->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    ctx->events.clear();
    ctx->prof_rows_in = 0;
    ctx->prof_rows_out = 0;
    for (auto& e : ctx->score_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    ctx->score_events.clear();
    ctx->prof_scored = 0;
    if (!ctx->d_prof) {
        pp_status st = dev_alloc(ctx, (void**)&ctx->d_prof, 2 * sizeof(unsigned long long));
        if (st != PP_OK) return st;
    }
    PP_HIP(ctx, hipMemsetAsync(ctx->d_prof, 0, 2 * sizeof(unsigned long long), ctx->stream));
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->profiling = enable != 0;
    return PP_OK;
}

pp_status pp_profile_read(pp_context* ctx, uint64_t* launches, double* total_ms, uint64_t* rows_in,
                          uint64_t* rows_out) {
    PP_NOT_CAPTURING(ctx, "pp_profile_read");
    if (!ctx) return PP_ERR_INVALID_ARGUMENT;
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double ms = 0.0;
    for (auto& e : ctx->events) {
        float t = 0.f;
        PP_HIP(ctx, hipEventElapsedTime(&t, e.first, e.second));
        ms += t;
    }
    if (launches) *launches = ctx->events.size();
    if (total_ms) *total_ms = ms;
    if (rows_in) *rows_in = ctx->prof_rows_in;
    unsigned long long out_rows = 0;
    if (ctx->d_prof) PP_HIP(ctx, hipMemcpy(&out_rows, ctx->d_prof, sizeof(out_rows), hipMemcpyDeviceToHost));
    if (rows_out) *rows_out = out_rows + ctx->prof_rows_out;
    return PP_OK;
}

pp_status pp_stream_probe(pp_context* ctx, uint64_t bytes, uint32_t reps, double* gbs) {
    if (!ctx || !gbs || reps == 0 || bytes < 4096) return PP_ERR_INVALID_ARGUMENT;
    PP_NOT_CAPTURING(ctx, "pp_stream_probe");
    const uint64_t n2 = bytes / 16;
    double* arr[3] = {};
    double* partial = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    pp_status st = PP_OK;
    auto cleanup = [&]() {
        for (double* p : arr)
            if (p) (void)hipFree(p);
        if (partial) (void)hipFree(partial);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    for (double*& p : arr)
        if ((st = dev_alloc(ctx, (void**)&p, n2 * 16)) != PP_OK) {
            cleanup();
            return st;
        }
    if ((st = dev_alloc(ctx, (void**)&partial, stream_blocks(n2) * 8)) != PP_OK ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return st != PP_OK ? st : fail(ctx, PP_ERR_DRIVER, "pp_stream_probe: events");
    }
    const double moved[4] = {1, 1, 2, 3};
    for (int pass = 0; pass < 4; ++pass) {
        double best = 0.0;
        for (uint32_t r = 0; r <= reps; ++r) {  // the first pass warms up
            hipError_t e = hipEventRecord(e0, ctx->stream);
            if (e == hipSuccess) e = launch_stream_pass(ctx->stream, pass, arr[0], arr[1], arr[2], partial, n2);
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                cleanup();
                return fail(ctx, PP_ERR_DRIVER, std::string("pp_stream_probe: ") + hipGetErrorString(e));
            }
            if (r > 0 && ms > 0.f) best = std::max(best, moved[pass] * (double)(n2 * 16) / (ms * 1e-3) / 1e9);
        }
        gbs[pass] = best;
    }
    cleanup();
    return PP_OK;
}

pp_status pp_profile_read_score(pp_context* ctx, uint64_t* launches, double* total_ms, uint64_t* terms) {
    PP_NOT_CAPTURING(ctx, "pp_profile_read_score");
    if (!ctx) return PP_ERR_INVALID_ARGUMENT;
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double ms = 0.0;
    for (auto& e : ctx->score_events) {
        float t = 0.f;
        PP_HIP(ctx, hipEventElapsedTime(&t, e.first, e.second));
        ms += t;
    }
    if (launches) *launches = ctx->score_events.size();
    if (total_ms) *total_ms = ms;
    if (terms) *terms = ctx->prof_scored;
    return PP_OK;
}

}  // extern "C"

// ---- batches: many sums stepped together, one join launch per layer step -----------------------------------------

struct pp_batch {
    pp_context* ctx = nullptr;
    std::vector<pp_sum*> sums;
    uint32_t W = 1;
    pp_coeff_kind kind = PP_COEFF_F64;
    void* d_items = nullptr;  // JoinArgs per item of a step
    void* h_items = nullptr;  // pinned staging of the same
    uint32_t* d_pair_start = nullptr;
    uint32_t* h_pair_start = nullptr;
    unsigned long long* d_out = nullptr;  // 4 counters per item
    unsigned long long* h_out = nullptr;
    // The batched TopN selection (batch_topn): its item arrays (pinned staging and device copy, one block each),
    // and its boundary-bin arrays (grow-only; the host copy of the gathered part).
    void* d_topn_items = nullptr;
    void* h_topn_items = nullptr;
    size_t topn_items_bytes = 0;
    unsigned char* d_ties = nullptr;
    size_t ties_cap = 0;  // terms
    std::vector<unsigned char> h_ties;
    std::vector<uint64_t> h_sorted;  // the host side of the sorted positions' copy (alive until the next sync)
};

pp_status pp_batch_create(pp_context* ctx, pp_sum* const* sums, uint32_t count, pp_batch** out) {
    if (!ctx || !out || (count && !sums)) return PP_ERR_INVALID_ARGUMENT;
    PP_NOT_CAPTURING(ctx, "pp_batch_create");
    *out = nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        if (!sums[i] || sums[i]->ctx != ctx) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "batch: a sum of another context");
        if (sums[i]->W != sums[0]->W || sums[i]->kind != sums[0]->kind)
            return fail(ctx, PP_ERR_UNSUPPORTED, "batch: the sums differ in width or coefficient type");
        for (uint32_t j = 0; j < i; ++j)
            if (sums[j] == sums[i]) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "batch: a sum appears twice");
    }
    auto* b = new (std::nothrow) pp_batch();
    if (!b) return fail(ctx, PP_ERR_OUT_OF_MEMORY, "host allocation");
    b->ctx = ctx;
    b->sums.assign(sums, sums + count);
    if (count) {
        b->W = sums[0]->W;
        b->kind = sums[0]->kind;
    }
    const size_t item_bytes = b->W == 1 ? (b->kind == PP_COEFF_F64 ? sizeof(JoinArgs<1, double>) : sizeof(JoinArgs<1, float>))
                                        : (b->kind == PP_COEFF_F64 ? sizeof(JoinArgs<2, double>) : sizeof(JoinArgs<2, float>));
    const size_t n = std::max<uint32_t>(count, 1);
    pp_status st;
    auto bail = [&](pp_status e) {
        pp_batch_free(b);
        return e;
    };
    if ((st = dev_alloc(ctx, &b->d_items, n * item_bytes)) != PP_OK ||
        (st = dev_alloc(ctx, (void**)&b->d_pair_start, (n + 1) * 4)) != PP_OK ||
        (st = dev_alloc(ctx, (void**)&b->d_out, n * 4 * sizeof(unsigned long long))) != PP_OK)
        return bail(st);
    if (hipHostMalloc(&b->h_items, n * item_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_pair_start, (n + 1) * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_out, n * 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return bail(fail(ctx, PP_ERR_OUT_OF_MEMORY, "pinned host allocation"));
    *out = b;
    return PP_OK;
}

pp_status pp_batch_free(pp_batch* b) {
    if (!b) return PP_OK;
    PP_NOT_CAPTURING(b->ctx, "pp_batch_free");
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_pair_start) (void)hipFree(b->d_pair_start);
    if (b->d_out) (void)hipFree(b->d_out);
    if (b->h_items) (void)hipHostFree(b->h_items);
    if (b->h_pair_start) (void)hipHostFree(b->h_pair_start);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->d_topn_items) (void)hipFree(b->d_topn_items);
    if (b->h_topn_items) (void)hipHostFree(b->h_topn_items);
    if (b->d_ties) (void)hipFree(b->d_ties);
    delete b;
    return PP_OK;
}

namespace {

// One layer step of a batch: item k joins sums[idx[k]] through chans[k] into its spare column set, all in one
// launch of each batched kernel, then one read-back of every item's counters.
template <int W, typename T>
pp_status batch_step(pp_batch* b, const std::vector<uint32_t>& idx, const std::vector<pp_channel>& chans,
                     const std::vector<Prepared>& preps, const std::vector<KeepProgram>& keeps) {
    pp_context* ctx = b->ctx;
    pp_status st;
    auto* items = (JoinArgs<W, T>*)b->h_items;
    uint32_t total_pairs = 0, scap_max = 64, tbits_max = 1;
    for (size_t k = 0; k < idx.size(); ++k) {
        pp_sum* s = b->sums[idx[k]];
        const int in = s->cur, out = 1 - s->cur;
        const uint32_t B = 1u << s->device_bits;
        const uint64_t rows_held = s->pending_topn ? s->phys_len : s->len;
        if ((st = ensure_columns(s, out, 2 * rows_held + (uint64_t)kSegAlign * B)) != PP_OK) return st;
        if ((st = ensure_buckets(s, out, B)) != PP_OK) return st;
        const Prepared& p = preps[k];
        const uint32_t delta = bucket_of_host(s, p.gx, p.gz, s->device_bits);
        uint32_t lcap = delta ? 2u * s->max_bucket : s->max_bucket;
        lcap = std::max<uint32_t>(64, (lcap + 63) & ~63u);
        if (lcap > kMaxPairRows) lcap = kMaxPairRows;
        uint32_t tb = 1;
        while ((1u << tb) < 2u * lcap) ++tb;
        JoinArgs<W, T> a{};
        a.x = s->col[in].x;
        a.z = s->col[in].z;
        a.c = (const T*)s->col[in].c;
        a.cap = s->col[in].cap;
        a.start = s->bk[in].start;
        a.len = s->bk[in].len;
        a.ox = s->col[out].x;
        a.oz = s->col[out].z;
        a.oc = (T*)s->col[out].c;
        a.ocap = s->col[out].cap;
        a.ostart = s->bk[out].start;
        a.olen = s->bk[out].len;
        a.delta = delta;
        a.hb = delta ? 31u - (uint32_t)__builtin_clz(delta) : 0u;
        a.scap = lcap;
        a.tcap_bits = tb;
        a.num_cosets = delta ? B / 2 : B;
        uint32_t gphase = 0;
        for (uint32_t w = 0; w < W; ++w) {
            a.gx[w] = p.gx[w];
            a.gz[w] = p.gz[w];
            gphase += (uint32_t)__builtin_popcountll(p.gx[w] & p.gz[w]);
        }
        a.gphase = gphase & 3u;
        a.cos_t = (T)preps[k].cos;
        a.sin_t = (T)preps[k].sin;
        a.clifford = preps[k].clifford ? 1 : 0;
        a.keep = keeps[k];
        a.stats = s->d_stats;
        a.redo = nullptr;  // every item stages its largest pair
        a.filter = s->pending_topn ? s->d_topn : nullptr;  // a TopN whose retain this join performs
        a.filter_flags = s->pending_topn ? s->pending_flags : nullptr;
        a.filter_score = s->pending_topn ? s->pending_score : nullptr;
        items[k] = a;
        b->h_pair_start[k] = total_pairs;
        total_pairs += a.num_cosets;
        scap_max = std::max(scap_max, lcap);
        tbits_max = std::max(tbits_max, tb);
    }
    const uint32_t n = (uint32_t)idx.size();
    b->h_pair_start[n] = total_pairs;
    PP_HIP(ctx, hipMemcpyAsync(b->d_items, items, n * sizeof(JoinArgs<W, T>), hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(ctx, hipMemcpyAsync(b->d_pair_start, b->h_pair_start, (n + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    const hipError_t le = launch_join_batch<W, T>(ctx->stream, (const JoinArgs<W, T>*)b->d_items, b->d_pair_start,
                                                  n, total_pairs, scap_max, tbits_max, b->d_out);
    PP_HIP(ctx, le);
    PP_HIP(ctx, hipMemcpyAsync(b->h_out, b->d_out, n * 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               ctx->stream));
    PP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < idx.size(); ++k) {
        pp_sum* s = b->sums[idx[k]];
        if (b->h_out[4 * k + 2]) return fail(ctx, PP_ERR_DRIVER, "rotation join: a pair exceeded the largest-pair bound");
        s->pending_topn = false;
        s->len = b->h_out[4 * k];
        s->max_bucket = (uint32_t)b->h_out[4 * k + 1];
        s->cur = 1 - s->cur;
        s->has_room = false;  // the batched offsets lay out starts only: the next layer runs out of place
    }
    return PP_OK;
}

}  // namespace

pp_status pp_batch_propagate(pp_batch* b, const pp_channel* channels, const uint64_t* offsets,
                             const pp_direction* directions, const pp_policy* policies, const pp_options* options) {
    if (!b || !offsets || !directions) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = b->ctx;
    PP_NOT_CAPTURING(ctx, "pp_batch_propagate");
    const uint32_t count = (uint32_t)b->sums.size();
    const size_t target = options ? options->target_bucket_len : 1024;
    const size_t min_buckets = options ? options->min_buckets : 128;
    if (target < 1 || min_buckets < 1) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "bucket options");
    pp_status st;
    std::vector<RunPlanner> plans;
    std::vector<KeepProgram> keeps(count);
    plans.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
        pp_sum* s = b->sums[i];
        if (offsets[i + 1] < offsets[i] || (offsets[i + 1] > offsets[i] && !channels))
            return fail(ctx, PP_ERR_INVALID_ARGUMENT, "batch: channel offsets");
        const pp_policy* pol = policies ? &policies[i] : nullptr;
        if (!valid_policy(pol)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
        if ((st = refresh(s)) != PP_OK) return st;
        if ((st = check_channels(s, channels + offsets[i], offsets[i + 1] - offsets[i])) != PP_OK) return st;
        plans.emplace_back(channels + offsets[i], offsets[i + 1] - offsets[i], directions[i] == PP_HEISENBERG, pol,
                           target, min_buckets);
        keeps[i] = lower_policy(pol);
    }
    std::vector<uint64_t> next(count, 0), run_end(count, 0);
    for (;;) {
        std::vector<uint32_t> idx;
        std::vector<pp_channel> chans;
        std::vector<Prepared> preps;
        std::vector<KeepProgram> ks;
        for (uint32_t i = 0; i < count; ++i) {
            const RunPlanner& plan = plans[i];
            if (next[i] >= plan.n_channels) continue;
            pp_sum* s = b->sums[i];
            if (next[i] == run_end[i]) run_end[i] = plan.start_run(s, next[i]);
            const pp_channel& ch = plan.channel_at(next[i]);
            if (run_end[i] - next[i] >= 2 && s->len > 0 && fused_permutation_runs() &&
                is_permutation_step(plan.prepared(s, next[i]))) {
                // a permutation run: walked in one pass for this instance (RunPlanner::apply_run), not step by step
                // in the batched join
                if ((st = plan.apply_run(s, next[i], run_end[i], keeps[i], nullptr)) != PP_OK) return st;
                next[i] = run_end[i];
                if ((st = plan.finish_run(s, next[i] < plan.n_channels)) != PP_OK) return st;
                continue;
            }
            if ((st = settle_device_bits(s)) != PP_OK) return st;
            const Prepared p = plan.prepared(s, next[i]);
            if (p.key_preserving) {  // in place, on its own
                if ((st = device_layer(s, p, keeps[i])) != PP_OK) return st;
                if (++next[i] == run_end[i] && (st = plan.finish_run(s, next[i] < plan.n_channels)) != PP_OK)
                    return st;
                continue;
            }
            idx.push_back(i);
            chans.push_back(ch);
            preps.push_back(p);
            ks.push_back(keeps[i]);
        }
        if (idx.empty()) {
            bool left = false;
            for (uint32_t i = 0; i < count; ++i) left = left || next[i] < plans[i].n_channels;
            if (!left) break;
            continue;  // every instance took a key-preserving layer this round
        }
        st = b->W == 1 ? (b->kind == PP_COEFF_F64 ? batch_step<1, double>(b, idx, chans, preps, ks)
                                                 : batch_step<1, float>(b, idx, chans, preps, ks))
                       : (b->kind == PP_COEFF_F64 ? batch_step<2, double>(b, idx, chans, preps, ks)
                                                 : batch_step<2, float>(b, idx, chans, preps, ks));
        if (st != PP_OK) return st;
        // the layer pass of every instance whose run ended: deferred exact TopN selections of sums above their cap
        // together, anything else on its own
        std::vector<uint32_t> together;
        std::vector<uint64_t> keep;
        for (uint32_t i : idx) {
            if (++next[i] != run_end[i]) continue;
            const RunPlanner& plan = plans[i];
            pp_sum* s = b->sums[i];
            const bool defer = next[i] < plan.n_channels;
            if (batch_topn_enabled() && defer && plan.finalizes && plan.policy->has_topn == 1 &&
                plan.policy->topn > 0 && s->len > plan.policy->topn && !s->pending_topn) {
                together.push_back(i);
                keep.push_back(plan.policy->topn);
            } else if ((st = plan.finish_run(s, defer)) != PP_OK) {
                return st;
            }
        }
        if (!together.empty() && (st = batch_topn(b, together, keep)) != PP_OK) return st;
    }
    return PP_OK;
}

// ---- tuning -------------------------------------------------------------------------------------------------------

pp_status pp_get_tuning(pp_context* ctx, pp_tuning* out) {
    if (!ctx || !out) return PP_ERR_INVALID_ARGUMENT;
    *out = ctx->tuning;
    return PP_OK;
}

pp_status pp_set_tuning(pp_context* ctx, const pp_tuning* t) {
    if (!ctx || !t) return PP_ERR_INVALID_ARGUMENT;
    PP_NOT_CAPTURING(ctx, "pp_set_tuning");
    if (!valid_tuning(*t)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "tuning out of range");
    ctx->tuning = *t;
    return PP_OK;
}

pp_status pp_autotune(pp_sum* sample, const pp_channel* channels, uint64_t n_channels, const pp_policy* policy,
                      pp_tuning* out) {
    if (!sample || (!channels && n_channels)) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = sample->ctx;
    PP_NOT_CAPTURING(ctx, "pp_autotune");
    if (!valid_policy(policy)) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "malformed keep program");
    pp_status st = refresh(sample);
    if (st != PP_OK) return st;
    // the channels the out-of-place layer takes (a key-preserving one runs in place)
    std::vector<pp_channel> chans;
    for (uint64_t i = 0; i < n_channels; ++i)
        if (!prepare(sample, channels[i], false).key_preserving) chans.push_back(channels[i]);
    const pp_tuning saved = ctx->tuning;
    const bool profiling = ctx->profiling;
    ctx->profiling = false;  // the trials' joins stay out of the measurement counters
    pp_tuning best = saved;
    double best_ms = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PP_HIP(ctx, hipEventCreate(&e0));
    PP_HIP(ctx, hipEventCreate(&e1));
    auto trial = [&](const pp_tuning& t, double* ms) -> pp_status {
        ctx->tuning = t;
        pp_sum *a = nullptr, *b = nullptr;
        pp_status rs = pp_sum_clone(sample, &a);
        if (rs == PP_OK) rs = settle_device_bits(a);  // the candidate's bucket count
        if (rs == PP_OK) rs = pp_sum_clone(a, &b);
        for (int pass = 0; pass < 2 && rs == PP_OK; ++pass) {  // the first pass sizes b's buffers
            if (pass == 1 && hipEventRecord(e0, ctx->stream) != hipSuccess) rs = PP_ERR_DRIVER;
            for (const pp_channel& ch : chans)
                if (rs == PP_OK) rs = pp_apply_rotation_to(a, b, &ch, 0, policy);
            if (rs == PP_OK && pass == 1 && hipEventRecord(e1, ctx->stream) != hipSuccess) rs = PP_ERR_DRIVER;
            if (rs == PP_OK) rs = refresh(b);
        }
        float f = 0.f;
        if (rs == PP_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&f, e0, e1) != hipSuccess))
            rs = fail(ctx, PP_ERR_DRIVER, "autotune timing");
        *ms = f;
        if (b) pp_sum_free(b);
        if (a) pp_sum_free(a);
        return rs;
    };
    std::vector<pp_tuning> candidates;
    for (uint32_t bt : {128u, 192u, 256u, 384u})
        for (uint32_t sr : {0u, 128u, 256u})
            for (uint32_t gp : {kJoinOnePairPerGroup, 0u})
                for (uint32_t jb : {256u, 512u}) candidates.push_back(pp_tuning{bt, sr, gp, jb, saved.join_in_place});
    for (const pp_tuning& t : candidates) {
        double ms = 0;
        if ((st = trial(t, &ms)) != PP_OK) break;
        if (best_ms < 0 || ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->profiling = profiling;
    ctx->tuning = st == PP_OK ? best : saved;
    if (st != PP_OK) return st;
    if (out) *out = best;
    return PP_OK;
}

// ---- layer sequences captured into a HIP graph ----------------------------------------------------------------

namespace {

// Whether the sum's current state still matches what a captured layer baked in; `what` names the first mismatch.
bool captured_join_valid(const CapturedJoin& r, std::string& what) {
    const pp_sum* src = r.src;
    const pp_sum* dst = r.dst;
    const int in = r.src_cur, out = r.dst_cur;
    const void* sp[5] = {src->col[in].x, src->col[in].z, src->col[in].c, src->bk[in].start, src->bk[in].len};
    const void* dp[7] = {dst->col[out].x, dst->col[out].z, dst->col[out].c, dst->bk[out].start, dst->bk[out].len,
                         dst->d_stats, dst->d_redo};
    if (src->cur != in || src->device_bits != r.device_bits || !std::equal(sp, sp + 5, r.src_ptrs)) {
        what = "the input's columns or device bucketing changed since the capture";
        return false;
    }
    if (dst->cur != out || dst->col[out].cap != r.dst_cap || !std::equal(dp, dp + 7, r.dst_ptrs)) {
        what = "the output's current columns changed since the capture (a layer in place, or a reallocation)";
        return false;
    }
    if (src->pending_topn) {
        what = "the input has a pending selection";
        return false;
    }
    // the captured launch sized the output segments, the staging and the table from these bounds
    const uint64_t B = 1ull << src->device_bits;
    if (2 * src->len + (uint64_t)kSegAlign * B > r.dst_cap || src->max_bucket > r.src_max_bucket) {
        what = "the input holds more rows (" + std::to_string(src->len) + ", longest bucket " +
               std::to_string(src->max_bucket) + ") than the captured launch was sized for (" +
               std::to_string(r.src_len) + ", " + std::to_string(r.src_max_bucket) + ")";
        return false;
    }
    return true;
}

}  // namespace

pp_status pp_graph_capture_begin(pp_context* ctx) {
    if (!ctx) return PP_ERR_INVALID_ARGUMENT;
    if (ctx->capturing) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "a capture is already open on this context");
    PP_HIP(ctx, hipSetDevice(ctx->device));
    PP_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    ctx->capture_log.clear();
    return PP_OK;
}

pp_status pp_graph_capture_end(pp_context* ctx, pp_graph** out) {
    if (!ctx || !out) return PP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!ctx->capturing) return fail(ctx, PP_ERR_INVALID_ARGUMENT, "no capture is open on this context");
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (e != hipSuccess || !graph) {
        (void)hipGetLastError();
        return fail(ctx, PP_ERR_DRIVER, std::string("graph capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(ctx, PP_ERR_DRIVER, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    auto* g = new (std::nothrow) pp_graph();
    if (!g) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return fail(ctx, PP_ERR_OUT_OF_MEMORY, "host allocation");
    }
    g->ctx = ctx;
    g->graph = graph;
    g->exec = exec;
    g->layers.swap(ctx->capture_log);
    ctx->graphs.push_back(g);
    *out = g;
    return PP_OK;
}

pp_status pp_graph_launch(pp_graph* g) {
    if (g) PP_NOT_CAPTURING(g->ctx, "pp_graph_launch");
    if (!g) return PP_ERR_INVALID_ARGUMENT;
    pp_context* ctx = g->ctx;
    // Every input's counts, then every captured buffer and bound, before anything is queued: a refused replay
    // writes nothing.  An input that is also an earlier replay's output is read after that replay's counters land.
    for (const CapturedJoin& r : g->layers) {
        pp_status st = refresh(r.src);
        if (st != PP_OK) return st;
    }
    for (size_t i = 0; i < g->layers.size(); ++i) {
        std::string what;
        if (!captured_join_valid(g->layers[i], what))
            return fail(ctx, PP_ERR_INVALID_ARGUMENT, "pp_graph_launch: layer " + std::to_string(i) + ": " + what);
    }
    PP_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    for (const CapturedJoin& r : g->layers) {
        r.dst->stale = true;  // its counts come from this replay
        r.dst->device_bits = r.device_bits;
        r.dst->reference_bits = r.src->reference_bits;
    }
    return PP_OK;
}

pp_status pp_graph_free(pp_graph* g) {
    if (g) PP_NOT_CAPTURING(g->ctx, "pp_graph_free");
    if (!g) return PP_OK;
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    auto& live = g->ctx->graphs;
    live.erase(std::remove(live.begin(), live.end(), g), live.end());
    delete g;
    return PP_OK;
}

// The TopN pass lives in pp_topn.hip and the learned network in pp_learned.hip; their host halves need the sum's
// internals.
#include "pp_topn_impl.inc"
#include "pp_learned_impl.inc"
