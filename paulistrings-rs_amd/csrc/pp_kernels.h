// pp_kernels.h -- argument blocks and launchers of the rotation-layer kernels (pp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pp_device.h"
#include "pp_topn.h"

namespace pp {

// Columns are word-major: x[w * cap + i], z[w * cap + i], c[i].  Buckets are CSR segments start[b], len[b].

template <int W, typename T>
struct JoinArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    uint64_t* ox;
    uint64_t* oz;
    T* oc;
    uint64_t ocap;
    const uint64_t* ostart;
    uint32_t* olen;
    uint32_t delta;      // h(g) under the device bucket bits; 0 makes every coset a single bucket
    uint32_t hb;         // highest set bit of delta
    uint32_t scap;       // staged (anticommuting) rows one workgroup holds in LDS; a pair with more is redone
    uint32_t tcap_bits;  // LDS table slots: 1 << tcap_bits >= 2 * scap
    uint32_t num_cosets; // pairs {b, b ^ delta} (every bucket on its own when delta = 0)
    uint32_t gphase;     // sum over words of |gx & gz|, the generator's own term of the product phase
    uint64_t gx[W];
    uint64_t gz[W];
    T cos_t;
    T sin_t;
    int32_t clifford;
    KeepProgram keep;
    unsigned long long* stats;  // [2] |= overflow in the redo launch (never expected: it stages the largest pair)
    uint32_t* redo;                  // pairs the first launch could not stage, for the redo launch
    unsigned long long* redo_count;  // their number
    uint32_t redo_mode;              // 0: every pair; 1: the queued pairs, staging up to scap = the largest pair
    uint32_t count_totals;           // each pair adds its rows to stats[0] and maxes stats[1] (few buckets: no
                                     // k_reduce_lens launch)
    // A finished TopN selection not yet applied to the input (null: none): rows it drops are not read further, as
    // though its retain had run.
    const TopnState* filter;
    const uint8_t* filter_flags;  // its boundary-bin flags per input position, or null
    const float* filter_score;    // its ranking score per input position when it ranked by score, else null (|c|)
};

// A small layer's join in one launch (k_rotation_join_small, a layout of at most kSmallJoinBuckets buckets, one
// workgroup per pair): every workgroup lays out its own pair's output segments from the input lengths and joins the
// pair, adding its rows and longest bucket to j.stats[0, 1] (zero before the launch; [2] flags an overflow, and
// workgroup 0 writes the layout's end to [3]).  One workgroup more publishes the counters of the launch before
// (prev_counters, complete at this launch's start) to that launch's pinned host slot, then a sequence number the
// host waits on, and clears them for the launch after this one.
constexpr uint32_t kSmallJoinBuckets = 2048;
template <int W, typename T>
struct SmallJoinArgs {
    JoinArgs<W, T> j;          // j.ostart and j.olen are written; j.count_totals set; no redo queue
    uint32_t* orows;           // each output segment's room in rows
    uint32_t num_buckets;
    unsigned long long* prev_counters;  // 4 counters of the launch before, or null
    unsigned long long* prev_host;      // pinned, device-visible: [0] rows, [1] longest bucket, [2] overflow, [3]
                                        // layout end, [4] prev_seq once [0 .. 3] are written
    unsigned long long prev_seq;
};

// The rotation layer in place (k_rotation_inplace): the sum's own columns and bucket arrays are read and written.
template <int W, typename T>
struct InplaceArgs {
    uint64_t* x;
    uint64_t* z;
    T* c;
    uint64_t cap;
    uint64_t* start;
    uint32_t* len;
    uint32_t* bcap;      // each segment's rows
    uint64_t tail_base;  // the first row after every segment: moved pairs allocate from tail_base + counter 5
    uint64_t tail_limit; // rows of the columns: a moved pair's segments stay below it
    uint32_t delta;
    uint32_t hb;
    uint32_t scap;
    uint32_t tcap_bits;
    uint32_t num_cosets;
    uint32_t gphase;
    uint64_t gx[W];
    uint64_t gz[W];
    T cos_t;
    T sin_t;
    int32_t clifford;
    KeepProgram keep;
    // [0] rows and [1] longest bucket when count_totals; [2] |= 1 redo overflow, |= 2 no room to move a pair;
    // [4] queued pairs; [5] rows moved pairs took after tail_base; [6] moved pairs; [7] rows the queued pairs' moves
    // may take
    unsigned long long* stats;
    uint32_t* redo;
    unsigned long long* redo_count;
    uint32_t redo_mode;
    uint32_t count_totals;
    uint32_t compact;  // a side left with more holes than new rows closes them in place (else the pair moves)
    // a TopN selection the layer reads its input through (the filtered launch): rows it drops are holes
    const TopnState* filter;
    const uint8_t* filter_flags;
    const float* filter_score;
};

// A local table applied row by row, one output per row (a permutation step, PermutationStep::of,
// engine/bucketed/permutation_run.rs:53-121, or a key-preserving rescale): the key's pattern on the table's support
// (bit 2j the x bit of support slot j, bit 2j+1 its z bit, channel/prepared.rs:141-151; a one-qubit table repeats
// its qubit in slot 1, masked off) selects the local delta it moves by and the amplitude it takes.
struct TableStep {
    double amp[16];       // exactly 0 drops the row
    uint32_t bucket[16];  // the device bucket delta of lift(delta[s]) under the bucketing of the launch
    uint8_t delta[16];
    uint32_t word[2];
    uint32_t bit[2];
    uint32_t mask;    // the patterns' live bits: 15 with two support qubits, 3 with one, 0 with none
    uint32_t inert;   // bit s: pattern s keeps its key and its coefficient exactly
};
constexpr uint32_t kMaxPermSteps = 256;  // steps per walk launch (a longer run takes several)
constexpr uint32_t kWalkChunk = 32;      // steps a walk workgroup holds in LDS at once
constexpr uint32_t kDeadRow = 0xFFFFFFFFu;

template <int W, typename T>
struct RescaleArgs {
    uint64_t* x;
    uint64_t* z;
    T* c;
    uint64_t cap;
    const uint64_t* start;
    uint32_t* len;
    TableStep step;  // key-preserving: every delta 0
    KeepProgram keep;
    unsigned long long* stats;
};

template <int W, typename T>
struct PermWalkArgs {
    uint64_t* x;  // walked in place
    uint64_t* z;
    T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    const TableStep* steps;
    uint32_t num_steps;
    KeepProgram keep;
    uint32_t* dest;                 // per column position: the row's output bucket, or kDeadRow
    uint32_t* counts;               // per output bucket: rows that land there
    unsigned long long* deaths;     // per step: rows that step drops
};

template <int W, typename T>
struct PermScatterArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    const uint32_t* dest;
    uint64_t* ox;
    uint64_t* oz;
    T* oc;
    uint64_t ocap;
    const uint64_t* ostart;
    uint32_t* olen;  // zero on entry: the write cursors, and the output lengths on exit
};

template <int W, typename T>
struct RefineArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    uint64_t* ox;
    uint64_t* oz;
    T* oc;
    uint64_t ocap;
    uint64_t* ostart;  // 2 * num_buckets entries
    uint32_t* olen;
    uint32_t num_buckets;  // before the split
    int32_t row;           // the hash row that splits
    const uint64_t* rows;  // [row][x|z][W]
    unsigned long long* stats;
    const uint32_t* bcap;  // the input segments' rows, or null (then no output segment rows)
    uint32_t* obcap;
};

template <int W, typename T>
struct CoarsenArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    uint64_t* ox;
    uint64_t* oz;
    T* oc;
    uint64_t ocap;
    const uint64_t* ostart;  // coarse_buckets entries, from k_bucket_offsets mode 2
    uint32_t* olen;
    uint32_t num_buckets;
    uint32_t coarse_buckets;
    unsigned long long* stats;
};

template <int W, typename T>
struct ExpectationArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    uint32_t num_buckets;
    uint32_t buckets_per_block;
    uint64_t axis_x[W];
    uint64_t axis_z[W];
    uint64_t negative[W];
    double* partial;
};

template <int W, typename T>
struct ExportArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    const uint64_t* dense_start;
    const uint64_t* rows;
    int32_t ref_bits;
    uint64_t* out_x;  // term-major [n][W]
    uint64_t* out_z;
    T* out_c;
    uint32_t* out_bucket;
};

// The resident layout's invariants (the reference's k_check_invariants, kernels/invariants.cu): per bucket, the
// segment aligned and inside the columns (within its room when the layout records one), every row in the bucket
// its key hashes to, no zero coefficient, no key twice.
enum CheckCount : uint32_t {
    kCheckMisplaced = 0,  // rows whose key hashes to another bucket
    kCheckZero = 1,       // rows with a coefficient of exactly zero
    kCheckSegments = 2,   // buckets whose segment is misaligned, outside the columns or longer than its room
    kCheckDuplicates = 3, // rows whose key another row of the bucket holds
    kCheckUnscanned = 4,  // buckets too long for the duplicate scan (more than 1024 rows)
    kCheckCounts = 5
};
template <int W, typename T>
struct CheckArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    const uint32_t* room;  // per-bucket segment rows, or null
    uint64_t end;          // rows of the columns in use (segments end at or before it)
    uint32_t num_buckets;
    const uint64_t* rows;  // hash rows
    int bits;
    unsigned long long* out;  // CheckCount slots
};

// Exclusive scan of per-bucket sizes into out_start (one launch for a layout of one tile, else three); tile_sums
// holds offsets_tiles(n) words.  When clear is set, clear[0 .. clear_n) is zeroed before the total is written (a
// join's counters, without a launch of their own).
uint32_t offsets_tiles(uint32_t out_n);
// out_cap (modes 1-3) receives each resident segment's rows.
hipError_t launch_bucket_offsets(hipStream_t st, const uint32_t* len, uint32_t num_buckets, uint32_t mode,
                                 uint32_t delta, uint32_t coarse_buckets, uint32_t fold, uint64_t* out_start,
                                 unsigned long long* total, uint64_t* tile_sums, unsigned long long* clear = nullptr,
                                 uint32_t clear_n = 0, uint32_t* out_cap = nullptr);
// p[0 .. n) = v, as a kernel (capturable).
hipError_t launch_fill_u64(hipStream_t st, unsigned long long* p, uint32_t n, unsigned long long v);
// out[0] += total rows, out[1] = max(out[1], longest bucket); prof[0] += total rows when prof is given.
hipError_t launch_reduce_lens(hipStream_t st, const uint32_t* len, uint32_t num_buckets, unsigned long long* out,
                              unsigned long long* prof);

template <int W, typename T>
size_t join_lds_bytes(uint32_t scap, uint32_t tcap_bits);
// groups_per_cu caps the resident workgroups per CU of the persistent grid (0: as many as fit;
// kJoinOnePairPerGroup: a workgroup per pair, no persistence).
constexpr uint32_t kJoinOnePairPerGroup = 0xFFFFFFFFu;
// block: 256 or 512 threads per workgroup (a chunk of that many rows per pass).
template <int W, typename T>
hipError_t launch_rotation_join(hipStream_t st, const JoinArgs<W, T>& a, uint32_t groups_per_cu, uint32_t block);
template <int W, typename T>
hipError_t launch_rotation_join_small(hipStream_t st, const SmallJoinArgs<W, T>& a);
// The publication of a small join's counters with no launch after it (one workgroup): as the extra workgroup of
// k_rotation_join_small does.
hipError_t launch_publish_counters(hipStream_t st, unsigned long long* counters, unsigned long long* host,
                                   unsigned long long seq);
// One layer for each of num_items sums (device arrays: items, pair_start = exclusive prefix of their pair counts,
// total_pairs in all): their output offsets, the joins (each item staging its largest pair, scap = lcap), and
// their counters, also gathered into out[4 i + 0..2] = rows, longest bucket, overflow flag.
template <int W, typename T>
hipError_t launch_join_batch(hipStream_t st, const JoinArgs<W, T>* items, const uint32_t* pair_start,
                             uint32_t num_items, uint32_t total_pairs, uint32_t scap_max, uint32_t tcap_bits_max,
                             unsigned long long* out);
template <int W, typename T>
hipError_t launch_rescale(hipStream_t st, const RescaleArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
size_t inplace_lds_bytes(uint32_t scap, uint32_t tcap_bits);
template <int W, typename T>
hipError_t launch_rotation_inplace(hipStream_t st, const InplaceArgs<W, T>& a);
template <int W, typename T>
hipError_t launch_perm_walk(hipStream_t st, const PermWalkArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_perm_scatter(hipStream_t st, const PermScatterArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_refine(hipStream_t st, const RefineArgs<W, T>& a);
template <int W, typename T>
hipError_t launch_coarsen(hipStream_t st, const CoarsenArgs<W, T>& a);
template <int W, typename T>
hipError_t launch_expectation(hipStream_t st, const ExpectationArgs<W, T>& a, uint32_t blocks);
template <int W, typename T>
hipError_t launch_export(hipStream_t st, const ExportArgs<W, T>& a, uint32_t num_buckets);
template <int W, typename T>
hipError_t launch_check_layout(hipStream_t st, const CheckArgs<W, T>& a);

}  // namespace pp
