// pp_learned.h -- the learned truncation policy's network on the device: features of a term, a 35-128-128-1 MLP
// with bf16 MFMA GEMMs and fp32 accumulation, its backward pass and an Adam step, and the per-term score that the
// learned TopN pass keeps the largest of.  No counterpart in the reference; it plugs into the TruncationPolicy seam
// (crates/paulistrings/src/truncation/mod.rs:30-113), with plain |c| TopN as fallback and judge.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pp_device.h"

namespace pp {

constexpr int kMlpIn = 35, kMlpInPad = 64, kMlpHid = 128;
constexpr int kMlpTile = 64;  // terms per workgroup pass
// flat fp32 parameter layout: W1 [64][128] (input-major, rows >= 35 zero), b1 [128], W2 [128][128] (input-major),
// b2 [128], w3 [128], b3
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpInPad * kMlpHid, kOffW2 = kOffB1 + kMlpHid,
              kOffB2 = kOffW2 + kMlpHid * kMlpHid, kOffW3 = kOffB2 + kMlpHid, kOffB3 = kOffW3 + kMlpHid,
              kMlpParams = kOffB3 + 1;

struct MlpDev {
    const float* params;   // fp32 master, kMlpParams
    const uint16_t* w1t;   // bf16 W1 transposed: [128 outputs][64 inputs]
    const uint16_t* w2t;   // bf16 W2 transposed: [128 outputs][128 inputs]
    const uint16_t* w2;    // bf16 W2 input-major: [128 inputs][128 outputs] (the backward's operand)
    uint64_t groups[16];   // qubit groups (lattice rows, then columns) whose X and Z occupancies are features
    float depth_scale;     // remaining layers are divided by it, the quotient capped at 1 (the horizon)
};

// Samples of one word: keys, coefficients (f64), remaining layers, and (training) labels.
struct MlpBatch {
    const uint64_t* x;
    const uint64_t* z;
    const double* c;
    const float* remaining;
    const float* label;  // training only
    uint64_t count;
};

hipError_t launch_mlp_refresh(hipStream_t st, const float* params, uint16_t* w1t, uint16_t* w2t, uint16_t* w2);
hipError_t launch_mlp_predict(hipStream_t st, const MlpDev& m, const MlpBatch& b, float* out);
// The gradient of the batch's mean squared error into grads (kMlpParams floats) and the sum of its squared errors
// into loss[0], through per-tile partials (partial: partial_tiles x kMlpParams floats, loss_part: partial_tiles
// doubles, at least one tile per kMlpTile rows) summed in a fixed order.
hipError_t launch_mlp_grad(hipStream_t st, const MlpDev& m, const MlpBatch& b, float* partial, double* loss_part,
                           uint64_t partial_tiles, float* grads, double* loss);
hipError_t launch_mlp_adam(hipStream_t st, float* params, float* adam_m, float* adam_v, const float* grads,
                           float lr, float beta1, float beta2, float eps, int step);

// The score of every term of a resident one-word sum, ln|c| + the network's output (the predicted ln of the term's
// read-out factor |<psi| P propagated |psi>|), so the predicted ln|contribution|; written at the term's column
// position: score[start[b] + r] for r < len[b].  A network that outputs a constant ranks exactly as |c| does.
template <typename T>
struct ScoreArgs {
    const uint64_t* x;
    const uint64_t* z;
    const T* c;
    const uint64_t* start;
    const uint32_t* len;
    uint32_t num_buckets;
    float remaining;
    float* score;
};
template <typename T>
hipError_t launch_mlp_score(hipStream_t st, const MlpDev& m, const ScoreArgs<T>& a);

}  // namespace pp
