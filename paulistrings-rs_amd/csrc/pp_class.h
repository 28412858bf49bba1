// pp_class.h -- the class pass: up to eight local tables (rotations, permutation steps, rescales, general Pauli
// transfer matrices on one or two qubits) applied to a resident sum in one pass, every layer inside LDS.
//
// A string's class is its key with the program's qubits cleared: every table of the program maps a string to
// strings of its own class, so classes are independent.  The class bucket of a string is the device hash of its
// class key (the GF(2)-linear rows of the sum, `cbits` of them), so one class never straddles two class buckets.
// The pass routes the strings some layer touches into class buckets (k_class_route counts, k_class_gather places),
// applies the whole program to one class bucket at a time in LDS (k_class_apply: every layer gathers each output
// from its sources in entry order and asks `keep`), and lays the result out in the sum's spare columns
// (k_class_place): the untouched strings that `keep` passes stay in their bucket, the class outputs are scattered to
// theirs.
//
// The reference's counterparts: the fused rotation run (engine/gpu/kernels/rotation_run.cu, launched from
// engine/gpu/layer/rotation_run.rs) and the general local-table layer (engine/gpu/kernels/layer.cu with
// engine/gpu/layer/fused.rs; the host's engine/bucketed/coset_fill.rs).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_device.h"

namespace pp {

constexpr uint32_t kClassMaxLayers = 8;
constexpr uint32_t kClassMaxEntries = 16;
constexpr uint32_t kClassBlock = 256;

// Counter slots of a class pass.
enum ClassStat : uint32_t {
    kClassOutRows = 0,    // rows the apply emitted (its cursor into the output staging; may pass its capacity)
    kClassFlags = 1,      // bit 0: a class bucket larger than the LDS set; bit 1: a layer outgrew it; bit 2: staging full
    kClassUntouched = 2,  // untouched rows that `keep` passes
    kClassStaged = 3,     // touched rows (the class-bucket scan's total)
    kClassLayerRows = 4,  // [4 + l]: class rows after layer l
    kClassLongest = 4 + kClassMaxLayers,  // the output layout's longest bucket
    kClassStatSlots = 5 + kClassMaxLayers
};

// One layer: a local table (LocalPtm, channel/prepared.rs:79-86) on the support slots word/bit (pattern_of).
struct ClassLayer {
    uint32_t word[2];
    uint32_t bit[2];
    uint32_t mask;       // the patterns' live bits: 15 with two support qubits, 3 with one
    uint32_t active;     // bit s: the layer moves, rescales or fills pattern s (else s keeps key and coefficient)
    uint32_t n_entries;  // entries ascending by local delta, the identity first when present
    uint32_t delta[kClassMaxEntries];
    double amp[kClassMaxEntries][16];  // entry e's amplitude on source pattern s, rounded to the coefficient type
};

template <int W, typename T>
struct ClassArgs {
    // the input: the sum's current columns and buckets
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    uint64_t cap;
    const uint64_t* start;
    const uint32_t* len;
    uint32_t num_buckets;
    int bits;               // device bucket bits of the input and of the output layout
    const uint64_t* rows;   // the sum's hash rows
    int cbits;              // class bucket bits
    uint32_t num_classes;   // 1 << cbits
    uint64_t qx[W], qz[W];  // the program's qubits (cleared in a class key)
    const ClassLayer* layers;
    uint32_t n_layers;
    uint32_t prog_words;  // sum over the layers of n_entries * 16: the amplitudes the apply holds in LDS
    KeepProgram keep;
    // routing
    uint32_t* ccount;  // [num_classes] touched rows per class bucket
    uint64_t* cstart;  // [num_classes] their exclusive scan
    uint32_t* ukeep;   // [num_buckets] untouched rows of each input bucket that keep passes
    // touched rows in class-bucket order
    uint64_t* sx;
    uint64_t* sz;
    T* sc;
    uint64_t scap;
    // the apply's outputs, each with its output bucket
    uint64_t* ox;
    uint64_t* oz;
    T* oc;
    uint32_t* obucket;
    uint64_t ocap;
    uint32_t* ocount;  // [num_buckets] outputs per output bucket
    unsigned long long* stats;  // ClassStat slots
    uint32_t lcap;              // rows of one LDS set
    uint32_t chunk_rows;        // staged rows one apply iteration aims for (whole class buckets)
    uint32_t tbits;             // LDS table slots: 1 << tbits >= 2 lcap
    // the output layout (the spare columns)
    uint64_t* px;
    uint64_t* pz;
    T* pc;
    uint64_t pcap;
    const uint64_t* pstart;
    uint32_t* plen;
};

template <int W, typename T>
size_t class_lds_bytes(uint32_t lcap, uint32_t tbits, uint32_t prog_words);

// Each launcher queues its kernel on st.
template <int W, typename T>
hipError_t launch_class_route(hipStream_t st, const ClassArgs<W, T>& a);  // counts (ccount, ukeep)
template <int W, typename T>
hipError_t launch_class_gather(hipStream_t st, const ClassArgs<W, T>& a);  // staging in class-bucket order
template <int W, typename T>
hipError_t launch_class_apply(hipStream_t st, const ClassArgs<W, T>& a);  // the program, one class bucket at a time
template <int W, typename T>
hipError_t launch_class_lens(hipStream_t st, const ClassArgs<W, T>& a);  // plen = ukeep + ocount, the longest
template <int W, typename T>
hipError_t launch_class_place(hipStream_t st, const ClassArgs<W, T>& a, uint64_t out_rows);  // the output layout

}  // namespace pp
