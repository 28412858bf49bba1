// pp_membench.h -- the HBM ceiling probe's launcher (pp_membench.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

// Workgroups of one pass over n2 16-byte vectors (the read pass writes one partial per workgroup).
uint64_t stream_blocks(uint64_t n2);
// pass 0: read a; 1: write a; 2: copy b -> a; 3: triad a = b + 3 c.  n2 = 16-byte vectors per array.
hipError_t launch_stream_pass(hipStream_t st, int pass, double* a, double* b, double* c, double* partial, uint64_t n2);

}  // namespace pp
