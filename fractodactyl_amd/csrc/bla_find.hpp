// Bla::find on the device (crates/fd-kernel/src/bla.rs:234-249), shared by the FP64 and
// FP32 perturbation kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fdg {

// p[i] for an array in global memory below 4 GiB: the byte offset is formed in 32 bits, so
// the access takes the scalar-base + 32-bit vector-offset form (p wave-uniform) instead of
// a 64-bit address computed per lane.
template <class T>
__device__ inline T load32(const __attribute__((address_space(1))) T* p, uint32_t i) {
    typedef const __attribute__((address_space(1))) char* Bytes;
    typedef const __attribute__((address_space(1))) T* Elem;
    return *(Elem)((Bytes)p + i * (uint32_t)sizeof(T));
}

// p[i] for a table that may exceed 4 GiB (WIDE: batches of frames, each with its own
// table, in one array), else load32.
template <bool WIDE, class T>
__device__ inline T table_load(const __attribute__((address_space(1))) T* p, uint32_t i) {
    return WIDE ? p[(uint64_t)i] : load32(p, i);
}

// Levels whose radii one lookup step loads together.
constexpr int kFindGroup = 4;

// The highest candidate level of a lookup at k = m - 1 (k even): a level-i block starts
// where 2^(i+1) divides k (level_top = min(trailing zeros of k, stored levels); k = 0
// allows every level), is at most `budget` steps long (2 << i <= budget, budget >= 2), and
// exists below the table's end (checked by the loads).  -1 when no level qualifies.
// (Tables hold at most 26 levels, below the 31 trailing zeros that k = 0 is given here;
// budget | 1 gives 0 and 1 the same floor(log2) of 0, hence -1.)
__device__ inline int find_cap(uint32_t k, uint32_t nlevels, uint32_t budget) {
    uint32_t tz = (uint32_t)__builtin_ctz(k | 0x80000000u);
    int top = (int)(tz < nlevels ? tz : nlevels);
    int by_budget = 30 - __builtin_clz(budget | 1u);
    int cap = top - 1;
    return cap < by_budget ? cap : by_budget;
}

// Bla::find's level once level 0 is known valid (d2 < its r2): the highest level i <= cap
// whose block exists and admits d2 < r2.  The reference walks up from level 1 and stops at
// the first failure; validity only shrinks going up (a level-i block's radius is at most
// its first half's, the level i-1 block at the same start, bla.rs:91-111, and blocks exist
// for a prefix of levels), so the first failure bounds the same prefix a top-down search
// finds.  Searching down from cap in groups of kFindGroup radii loaded together decides
// most lookups with one memory latency: the level taken is usually within three of cap
// (the alignment and budget allow long blocks, the radius trims the last one or two).
//   r2: the table's squared radii in the level-major layout of BlaDev (level i at
//   tb + bla_level_offset(stride, i)), whose padding past each level's last block holds
//   0 (the builder writes it), so a block past the table's end simply fails the test.  A
//   lookup happens at m <= len - 2 (k <= steps - 1), so k >> (i + 1) stays inside level
//   i's slot of stride >> (i + 1) entries.  Indices are 32-bit (tables below 2^32 blocks).
//   The last group of a search may reach below level 1: those loads are not taken, and
//   their shift counts (the level's low five bits) land them on a level-0 block or on
//   entry tb + stride, one past the table, which every r2 buffer allocates.
//   Every level searched divides k by 2^(i+1), so block k >> (i + 1) of level i sits at
//   tb + stride - (stride >> i) + (k >> (i + 1)) = (tb + stride) - ((stride - k/2) >> i):
//   one shift and one subtraction per level.
__device__ inline uint32_t bla_block_index(uint32_t tb, uint32_t stride, uint32_t k, int i) {
    return (tb + stride) - ((stride - (k >> 1)) >> i);
}

template <bool WIDE>
__device__ inline int find_level_down(const __attribute__((address_space(1))) double* r2,
                                      uint32_t tb, uint32_t stride, uint32_t k, int cap,
                                      double d2) {
    const uint32_t top = tb + stride, span = stride - (k >> 1);
    int lvl = 0;
    for (int hi = cap; hi >= 1; hi -= kFindGroup) {
        double rr[kFindGroup];
#pragma unroll
        for (int q = 0; q < kFindGroup; q++) {
            uint32_t lv = (uint32_t)(hi - q) & 31u;  // (levels below 1: not taken)
            rr[q] = table_load<WIDE>(r2, top - (span >> lv));
        }
        int got = -1;
#pragma unroll
        for (int q = kFindGroup - 1; q >= 0; q--)
            got = (d2 < rr[q] && hi - q >= 1) ? hi - q : got;
        if (got >= 1) {
            lvl = got;
            break;
        }
    }
    return lvl;
}

// The same search when every searching lane of the wave has the same table, start and cap
// (lanes in lockstep: deep frames' samples take the same blocks at the same orbit indices).
// The radii are then the same words for every lane, so they are loaded once through the
// scalar unit (r2 in the constant address space, indices in SGPRs) and only the
// comparisons with each lane's |delta|^2 run on the vector unit.  tb, k and cap must be
// wave-uniform.
typedef const __attribute__((address_space(4))) double* ConstR2;

__device__ inline int find_level_down_uniform(ConstR2 r2, uint32_t tb, uint32_t stride,
                                              uint32_t k, int cap, double d2) {
    const uint32_t top = tb + stride, span = stride - (k >> 1);
    int lvl = 0;
    for (int hi = cap; hi >= 1; hi -= kFindGroup) {
        double rr[kFindGroup];
#pragma unroll
        for (int q = 0; q < kFindGroup; q++) {
            int lv = hi - q;
            rr[q] = lv >= 1 ? r2[top - (span >> lv)] : -1.0;
        }
        int got = -1;
#pragma unroll
        for (int q = kFindGroup - 1; q >= 0; q--)
            got = d2 < rr[q] ? hi - q : got;
        if (got >= 1) {
            lvl = got;
            break;
        }
    }
    return lvl;
}

}  // namespace fdg
