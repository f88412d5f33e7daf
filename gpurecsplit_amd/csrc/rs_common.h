// rs_common.h -- definitions shared by the gfx950 kernels and the host engine.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RS_HD __host__ __device__ __forceinline__
#else
#define RS_HD inline
#endif

namespace rsa {

// Bucket sizes are bounded by the memo table (AbstractParallelRecSplit.hpp:60).
constexpr uint32_t kMaxBucket = 3000;
constexpr uint32_t kMaxLeaf = 24;    // AbstractParallelRecSplit.hpp:62
constexpr uint32_t kMaxFanout = 9;   // AbstractParallelRecSplit.hpp:63
constexpr uint32_t kMaxDepth = 17;   // fanout-2 levels use start_seed[0..16]
// Seeds at or above this are reported as RS_ERR_SEED_OVERFLOW; it bounds every
// search so that each wave reaches its exit even on adversarial input.
constexpr uint64_t kSeedLimit = uint64_t(1) << 44;

// Bits of a build's device error word.
constexpr uint32_t kErrDuplicateKeys = 1;   // two keys of one bucket share `second`
constexpr uint32_t kErrSeedLimit = 2;       // a search passed kSeedLimit
constexpr uint32_t kErrEfJump = 4;          // an Elias-Fano jump offset exceeds 16 bits
constexpr uint32_t kErrTimeout = 8;         // the context's time limit passed
constexpr uint32_t kErrBucketTooLarge = 16; // a bucket of kMaxBucket or more keys (build plans)
constexpr uint32_t kErrRiceCap = 32;        // descriptors exceed a plan's Golomb-Rice bound
constexpr uint32_t kErrEfCap = 64;          // Elias-Fano arrays exceed a plan's bounds
// Errors after which a plan's search and encoding kernels write nothing.
constexpr uint32_t kErrStopSearch = kErrDuplicateKeys | kErrBucketTooLarge;
constexpr uint32_t kErrStopEncode = kErrStopSearch | kErrRiceCap | kErrEfCap;

// Starting seed of each tree level (AbstractParallelRecSplit.hpp:95-98): fanout-2
// nodes use [depth], upper aggregation [17], lower aggregation [18], leaves [19].
#define RS_START_SEEDS                                                                          \
    {0x106393c187cae21aULL, 0x6453cec3f7376937ULL, 0x643e521ddbd2be98ULL, 0x3740c6412f6572cbULL, \
     0x717d47562f1ce470ULL, 0x4cd6eb4c63befb7cULL, 0x9bfd8c5e18c8da73ULL, 0x082f20e10092a9a3ULL, \
     0x2ada2ce68d21defcULL, 0xe33cb4f3e7c6466bULL, 0x3980be458c509c59ULL, 0xc466fd9584828e8cULL, \
     0x45f0aabe1a61ede6ULL, 0xf6e7b8b33ad9b98dULL, 0x4ef95e25f4b4983dULL, 0x81175195173b92d3ULL, \
     0x4e50927d8dd15978ULL, 0x1ea2099d1fafae7fULL, 0x425c8a06fbaaa815ULL, 0xcd4216006c74052aULL}
constexpr int kSeedUpper = 17, kSeedLower = 18, kSeedLeaf = 19;

// Node kinds; also the index into rs_build_stats::search_ms / evals / nodes.
enum NodeKind : uint32_t { kHigher = 0, kUpper = 1, kLower = 2, kLeafRot = 3, kLeafBrute = 4, kNumKinds = 5 };

// Phase lists of a build: fanout-2 nodes by depth, then upper aggregation, lower
// aggregation, rotation leaves, brute-force leaves.
constexpr int kNumLists = kMaxDepth + 4;

// One node of a bucket's split tree, as laid out in HBM (16 B).
struct NodeDesc {
    uint64_t key_off;  // first key of the node in the bucket-sorted key array
    uint32_t slot;     // global DFS-preorder index of the node's seed
    uint16_t m;        // number of keys
    uint8_t depth;     // fanout-2 depth (start_seed index) for kHigher
    uint8_t kind;
};

// Search state of one node (32 B): lowest valid seed found so far, the next
// unclaimed chunk of candidate seeds, and the number of waves working on it.
struct NodeState {
    unsigned long long result;
    unsigned long long next_chunk;
    unsigned int waves;
    unsigned int pad0;
    unsigned long long pad1;
};

// Split parameters of a leaf size (AbstractParallelRecSplit.hpp:133-148).
struct SplitParams {
    uint32_t leaf, lower, upper;
};

RS_HD uint32_t higher_split(const SplitParams &P, uint32_t m) {
    return ((uint16_t)(m / 2 + P.upper - 1) / P.upper) * P.upper;
}

RS_HD uint32_t classify(const SplitParams &P, uint32_t m, bool rotate) {
    if (m > P.upper) return kHigher;
    if (m > P.lower) return kUpper;
    if (m > P.leaf) return kLower;
    return (rotate && m == P.leaf) ? kLeafRot : kLeafBrute;
}

// MurmurHash3 fmix32 of the folded 64-bit word, low 16 bits
// (remix32 + remixAndRemap, AbstractParallelRecSplit.hpp:277-284, 327-337).
RS_HD uint32_t fold_h16(uint64_t z) {
    uint32_t t = (uint32_t)(z >> 32) ^ (uint32_t)z;
    t ^= t >> 16;
    t *= 0x85ebca6bu;
    t ^= t >> 13;
    t *= 0xc2b2ae35u;
    t ^= t >> 16;
    return t & 0xFFFFu;
}

RS_HD uint32_t remap16(uint64_t z, uint32_t n) { return (fold_h16(z) * n) >> 16; }

// Flags word of a build (the `rotate` argument of the C ABI): rotation fitting for
// full leaves (USE_BIJECTIONS_ROTATE), and the 64-bit remap of the RecSplitRotate
// variant (USE_64_BIT_REMAP, RecSplitRotate.hpp:115).
constexpr int kFlagRotate = 1, kFlagRemap64 = 2;

// remixAndRemap with USE_64_BIT_REMAP (AbstractParallelRecSplit.hpp:327-332): the
// 64-bit Stafford finalizer (remix, :265-269), its low 48 bits scaled to [0, n).
// With hi16 and lo the bits 32..47 and 0..31 of the finalizer:
// (hi16 * 2^32 + lo) * n >> 48 = (hi16 * n + floor(lo * n / 2^32)) >> 16.
RS_HD uint32_t remap64(uint64_t z, uint32_t n) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    const uint32_t lo = (uint32_t)z, hi16 = (uint32_t)(z >> 32) & 0xFFFFu;
    return (hi16 * n + (uint32_t)(((uint64_t)lo * n) >> 32)) >> 16;
}

RS_HD uint32_t remap_any(uint64_t z, uint32_t n, bool r64) { return r64 ? remap64(z, n) : remap16(z, n); }

}  // namespace rsa
