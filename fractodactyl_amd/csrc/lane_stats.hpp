// Per-lane pass statistics for the persistent perturbation kernels.
//
// A sample's counts (iterations, blocks, skipped steps, plain steps, ...) are added at its
// end into an LDS slot that only its lane touches: a non-returning ds_add_u64 / ds_max_u64
// at a lane-private address.  An add at one address shared by the block would make the
// compiler fold the wave's active lanes first, one lane at a time (a scalar loop of about
// eight instructions per lane), and in frames where every lane of a wave finishes on the
// same step that loop ran 64 times per counter at every refill.  The slots are folded once,
// when the block ends.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace fdg {

// ST_ITERATIONS .. ST_MAXOPS are the per-sample counters.
constexpr int kLaneStats = ST_MAXOPS + 1;

template <int BLOCK>
struct LaneStats {
    unsigned long long* v;  // kLaneStats * BLOCK slots in LDS, counter-major
    __device__ void clear() {
        for (int q = 0; q < kLaneStats; q++) v[q * BLOCK + threadIdx.x] = 0;
    }
    __device__ void add(int q, unsigned long long x) {
        atomicAdd(&v[q * BLOCK + threadIdx.x], x);
    }
    __device__ void max(int q, unsigned long long x) {
        atomicMax(&v[q * BLOCK + threadIdx.x], x);
    }
    // Folds this lane's slots over its wave (sums; the maximum for ST_MAXOPS) and adds
    // the wave's totals into the block's counters `stat` (LDS).  Every lane of the block
    // calls it (each reads only its own slots, so no barrier is needed before).
    __device__ void fold(unsigned long long* stat) const {
        const unsigned lane = __lane_id();
        for (int q = 0; q < kLaneStats; q++) {
            unsigned long long x = v[q * BLOCK + threadIdx.x];
            for (int o = 32; o > 0; o >>= 1) {
                unsigned long long y = __shfl_xor(x, o, 64);
                x = q == ST_MAXOPS ? (x > y ? x : y) : x + y;
            }
            if (lane == 0 && x) {
                if (q == ST_MAXOPS)
                    atomicMax(&stat[q], x);
                else
                    atomicAdd(&stat[q], x);
            }
        }
    }
};

}  // namespace fdg
