// rs_kernels.hip -- gfx950 kernels of the RecSplit construction engine.
//
// Data layout in HBM (one build):
//   keys_in   n x 16 B   (first, second) pairs, read twice (histogram, scatter)
//   sorted    n x 8 B    `second` of every key, grouped by bucket; inside a bucket
//                        sorted ascending, then permuted in place as the split
//                        tree is descended (each node's keys stay contiguous)
//   nodes     N x 16 B   NodeDesc of every tree node, grouped by search phase
//   state     16 B/node  NodeState of the phase being searched
//   seeds     N x 8 B    final seed of every node, by DFS-preorder slot
//   node_m    N x 2 B    size of every node, by slot
//
// Seed search (the hot loop): a persistent kernel per phase.  Each wavefront
// owns one node at a time, claims chunks of candidate seeds from the node's
// NodeState with one atomic, and evaluates them with 64 lanes x U seeds in
// flight per lane while streaming the node's keys through scalar registers
// (wave-uniform loads).  A ballot picks the lowest valid seed of a group; the
// node's result is the atomic minimum over all chunks, and a chunk is only
// skipped once its first seed is not below that minimum, so the result is the
// lowest valid seed.  Waves with no fresh node help the nodes still running.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "rs_common.h"
#include "rs_spooky.h"

namespace rsa {

__constant__ uint64_t c_start_seed[20] = RS_START_SEEDS;

constexpr unsigned long long kNone = ~0ull;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-uniform values, kept in scalar registers.
// __builtin_amdgcn_readfirstlane returns int: every result goes through uint32_t
// before it is widened, or a low word with bit 31 set would be sign-extended into
// the high word of a 64-bit value.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t bcast64(uint64_t v, int src) {
    const uint32_t lo = __shfl((uint32_t)v, src), hi = __shfl((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t bcast64_xor(uint64_t v, int mask) {
    const uint32_t lo = __shfl_xor((uint32_t)v, mask), hi = __shfl_xor((uint32_t)(v >> 32), mask);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t bcast64_up(uint64_t v, int delta) {
    const uint32_t lo = __shfl_up((uint32_t)v, delta), hi = __shfl_up((uint32_t)(v >> 32), delta);
    return ((uint64_t)hi << 32) | lo;
}

// Search-side hash: fold_h16 (rs_common.h) split so that its last step fuses
// with the consumer.  hash_pre is MurmurHash3 fmix32 of the folded word up to the
// second multiply; h16_of finishes it, (t ^ (t >> 16)) & 0xFFFF, as one
// word-select xor whose upper half is written as zero.
__device__ __forceinline__ uint32_t hash_pre32(uint32_t t) {
    t ^= t >> 16;
    t *= 0x85ebca6bu;
    t ^= t >> 13;
    return t * 0xc2b2ae35u;
}
__device__ __forceinline__ uint32_t hash_pre(uint64_t z) { return hash_pre32((uint32_t)(z >> 32) ^ (uint32_t)z); }
// hash_pre32 after its first step: t already holds t ^ (t >> 16).
__device__ __forceinline__ uint32_t hash_post16(uint32_t t) {
    t *= 0x85ebca6bu;
    t ^= t >> 13;
    return t * 0xc2b2ae35u;
}
__device__ __forceinline__ uint32_t h16_of(uint32_t t) {
    uint32_t x;
    __asm__("v_xor_b32_sdwa %0, %1, %1 dst_sel:WORD_0 dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:WORD_0"
            : "=v"(x)
            : "v"(t));
    return x;
}
__device__ __forceinline__ uint32_t search_h16(uint64_t z) { return h16_of(hash_pre(z)); }

// LDS byte address of entry (h16 * m) >> 16 of a table of 2^shift-byte entries
// that starts at entry `base_hi >> 16`: one 24-bit multiply-add, then the upper
// half of the sum shifted left by `shift` in one word-select shift.  h16 * m <
// 2^28 and the table base is below 2^15 entries, so the sum fits 32 bits and its
// upper half is the entry index.
__device__ __forceinline__ uint32_t lut_byte_addr(uint32_t h16, uint32_t m, uint32_t base_hi, uint32_t two) {
    const uint32_t p = __umul24(h16, m) + base_hi;
    uint32_t a;
    __asm__("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
            : "=v"(a)
            : "v"(two), "v"(p));
    return a;
}

// Elias-Fano geometry (DoubleEF.hpp:63-69): q = 256 elements per sample,
// super blocks of 2^14 elements, 17 words per super block and sequence.
constexpr uint64_t kEfQMaskDev = 255;
constexpr uint64_t kEfSuperSizeDev = 17;

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ partition

// A contiguous range of keys of one key set: its keys map to buckets
// bbase + floor(first * nb / 2^64) (common.hpp:273-280, with the set's own nb).
// Every set's keys may live in their own device allocation.
struct KeyTile {
    const uint64_t *keys;  // the set's (first, second) pairs
    uint64_t k0, k1;       // key range [k0, k1) of the set
    uint64_t nb;           // buckets of the set
    uint64_t bbase;        // first global bucket of the set
};

// Bucket histogram, one workgroup per key tile.  Tiles of sets with at most
// `priv_max` buckets count in LDS first; larger sets count with global atomics.
__global__ __launch_bounds__(256) void k_bucket_hist(const KeyTile *__restrict__ tiles, uint32_t *__restrict__ counts,
                                                     uint32_t priv_max) {
    extern __shared__ uint32_t s_cnt[];
    const KeyTile t = tiles[blockIdx.x];
    const uint64_t *__restrict__ keys = t.keys;
    if (t.nb <= priv_max) {
        for (uint32_t i = threadIdx.x; i < t.nb; i += blockDim.x) s_cnt[i] = 0;
        __syncthreads();
        for (uint64_t i = t.k0 + threadIdx.x; i < t.k1; i += blockDim.x) atomicAdd(&s_cnt[__umul64hi(keys[2 * i], t.nb)], 1u);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < t.nb; i += blockDim.x)
            if (s_cnt[i]) atomicAdd(&counts[t.bbase + i], s_cnt[i]);
    } else {
        for (uint64_t i = t.k0 + threadIdx.x; i < t.k1; i += blockDim.x)
            atomicAdd(&counts[t.bbase + __umul64hi(keys[2 * i], t.nb)], 1u);
    }
}

// Per bucket: node count (tree_nodes table by bucket size), size histogram
// (counted in LDS per workgroup, so that the few common sizes do not contend on
// global atomics), oversize flag.
__global__ __launch_bounds__(256) void k_bucket_meta(const uint32_t *__restrict__ counts, uint64_t nb,
                                                     const uint32_t *__restrict__ tree_nodes, uint64_t *__restrict__ counts64,
                                                     uint64_t *__restrict__ nodes, uint32_t *__restrict__ size_hist) {
    __shared__ uint32_t s_hist[kMaxBucket + 1];
    for (uint32_t i = threadIdx.x; i <= kMaxBucket; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t m = counts[b];
        counts64[b] = m;
        if (m >= kMaxBucket) {
            atomicAdd(&s_hist[kMaxBucket], 1u);
            nodes[b] = 0;
        } else {
            atomicAdd(&s_hist[m], 1u);
            nodes[b] = tree_nodes[m];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i <= kMaxBucket; i += blockDim.x)
        if (s_hist[i]) atomicAdd(&size_hist[i], s_hist[i]);
}

// Each key's `second` to a free slot of its bucket's range (any order: the sort
// that follows makes it canonical).  Tiles of sets with at most `priv_max` buckets
// count their keys per bucket in LDS, reserve each touched bucket's slots with one
// global atomic and place their keys by LDS atomics at set-relative positions.
// Larger sets take one global atomic per key.
__global__ __launch_bounds__(256) void k_scatter(const KeyTile *__restrict__ tiles, const uint64_t *__restrict__ bucket_acc,
                                                 uint32_t *__restrict__ cursor, uint64_t *__restrict__ sorted,
                                                 uint32_t priv_max) {
    extern __shared__ uint32_t s_pos[];  // priv_max entries
    const KeyTile t = tiles[blockIdx.x];
    const uint64_t *__restrict__ keys = t.keys;
    const uint64_t set0 = bucket_acc[t.bbase];  // the set's first slot
    if (t.nb <= priv_max && bucket_acc[t.bbase + t.nb] - set0 <= 0xffffffffull) {
        const uint32_t nb = (uint32_t)t.nb;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) s_pos[i] = 0;
        __syncthreads();
        for (uint64_t i = t.k0 + threadIdx.x; i < t.k1; i += blockDim.x)
            atomicAdd(&s_pos[__umul64hi(keys[2 * i], t.nb)], 1u);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) {
            const uint32_t k = s_pos[i];
            if (k) s_pos[i] = (uint32_t)(bucket_acc[t.bbase + i] - set0) + atomicAdd(&cursor[t.bbase + i], k);
        }
        __syncthreads();
        uint64_t *__restrict__ out = sorted + set0;
        for (uint64_t i = t.k0 + threadIdx.x; i < t.k1; i += blockDim.x)
            out[atomicAdd(&s_pos[__umul64hi(keys[2 * i], t.nb)], 1u)] = keys[2 * i + 1];
        return;
    }
    for (uint64_t i = t.k0 + threadIdx.x; i < t.k1; i += blockDim.x) {
        const uint64_t b = t.bbase + __umul64hi(keys[2 * i], t.nb);
        const uint32_t p = atomicAdd(&cursor[b], 1u);
        sorted[bucket_acc[b] + p] = keys[2 * i + 1];
    }
}

// One workgroup per bucket: bitonic sort of the bucket's `second` values in LDS
// (canonical order, independent of the scatter's atomics) and detection of equal
// values, which no seed can ever separate.
__global__ __launch_bounds__(256) void k_sort_buckets(uint64_t *__restrict__ sorted, const uint64_t *__restrict__ acc,
                                                      uint32_t *__restrict__ err, uint32_t rot_leaf,
                                                      uint32_t min_m = 0, uint32_t max_m = kMaxBucket - 1) {
    extern __shared__ uint64_t s[];  // next power of two of max_m (the largest bucket it sorts)
    const uint64_t b = blockIdx.x;
    const uint64_t off = acc[b];
    const uint32_t m = (uint32_t)(acc[b + 1] - off);
    if (m < 2 || m >= kMaxBucket || m < min_m || m > max_m) return;
    uint32_t P = 1;
    while (P < m) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) s[i] = i < m ? sorted[off + i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = s[i], c = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        s[i] = c;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
        if (i > 0 && s[i] == s[i - 1]) atomicOr(err, 1u);
    if (m == rot_leaf) {
        // a bucket that is one rotation leaf: even keys first (see k_split)
        if (threadIdx.x == 0) {
            uint32_t o = 0;
            for (uint32_t par = 0; par < 2; par++)
                for (uint32_t i = 0; i < m; i++)
                    if ((s[i] & 1) == par) sorted[off + o++] = s[i];
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) sorted[off + i] = s[i];
}

// One thread per bucket: walks the bucket's split tree in DFS preorder and emits
// every node (key range, size, slot, fanout-2 depth) into its phase list.  The
// tree shape depends only on the bucket size (AbstractParallelRecSplit.hpp:137-148,
// SIMDRecSplit.hpp:624-693), so all nodes are known before any search.
struct ExpandArgs {
    SplitParams P;
    int rotate;
    NodeDesc *pool;
    const uint64_t *list_base;  // [kMaxDepth + 4]
    uint32_t *list_fill;        // [kMaxDepth + 4]
    uint16_t *node_m;
};

__device__ __forceinline__ uint32_t list_of(uint32_t kind, uint32_t depth) {
    return kind == kHigher ? depth : kMaxDepth + kind - 1;
}

// DFS preorder walk of a bucket's split tree: calls f(list, offset, size, depth,
// kind) for every node, in slot order.
template <typename F>
__device__ __forceinline__ void walk_tree(const SplitParams &P, bool rotate, uint32_t m, F &&f) {
    struct Item {
        uint32_t off;
        uint16_t m;
        uint8_t depth;
    };
    Item stack[48];
    int sp = 0;
    stack[sp++] = Item{0, (uint16_t)m, 0};
    while (sp > 0) {
        const Item it = stack[--sp];
        const uint32_t kind = classify(P, it.m, rotate);
        f(list_of(kind, it.depth), it.off, (uint32_t)it.m, (uint32_t)it.depth, kind);
        if (it.m <= P.leaf) continue;
        uint32_t unit, fan;
        if (kind == kHigher) {
            unit = higher_split(P, it.m);
            fan = 2;
        } else if (kind == kUpper) {
            unit = P.lower;
            fan = (uint16_t)(it.m + P.lower - 1) / P.lower;
        } else {
            unit = P.leaf;
            fan = (uint16_t)(it.m + P.leaf - 1) / P.leaf;
        }
        // push children right to left so that the leftmost is visited next
        for (int c = (int)fan - 1; c >= 0; c--) {
            const uint32_t cs = (c == (int)fan - 1) ? it.m - (fan - 1) * unit : unit;
            if (cs > 1) stack[sp++] = Item{it.off + (uint32_t)c * unit, (uint16_t)cs, (uint8_t)(it.depth + 1)};
        }
    }
}

// One wavefront per 64 buckets: every bucket's nodes go to their phase lists.
// A first walk counts each bucket's nodes per list; a wave-wide exclusive scan per
// list gives each bucket its place, so a wave reserves a list's space with one
// atomic instead of one per node (millions of nodes contend on a few lists).
__global__ __launch_bounds__(64) void k_expand(const uint64_t *__restrict__ acc, const uint64_t *__restrict__ node_base,
                                               uint64_t nb, ExpandArgs a) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t m = b < nb ? (uint32_t)(acc[b + 1] - acc[b]) : 0;
    const bool active = m >= 2 && m < kMaxBucket;
    const SplitParams P = a.P;
    const bool rotate = a.rotate != 0;
    uint32_t cnt[kNumLists];
#pragma unroll
    for (int l = 0; l < kNumLists; l++) cnt[l] = 0;
    if (active) walk_tree(P, rotate, m, [&](uint32_t list, uint32_t, uint32_t, uint32_t, uint32_t) { cnt[list]++; });
    uint32_t pos[kNumLists];
#pragma unroll
    for (int l = 0; l < kNumLists; l++) {
        pos[l] = 0;
        if (!__ballot(cnt[l] != 0)) continue;
        uint32_t inc = cnt[l];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(inc, o);
            if (lane >= (uint32_t)o) inc += t;
        }
        const uint32_t total = __shfl(inc, 63);
        uint32_t wbase = 0;
        if (lane == 0) wbase = atomicAdd(&a.list_fill[l], total);
        pos[l] = __shfl(wbase, 0) + inc - cnt[l];
    }
    if (!active) return;
    uint32_t slot = (uint32_t)node_base[b];
    const uint64_t koff = acc[b];
    walk_tree(P, rotate, m, [&](uint32_t list, uint32_t off, uint32_t sz, uint32_t depth, uint32_t kind) {
        NodeDesc d;
        d.key_off = koff + off;
        d.slot = slot;
        d.m = (uint16_t)sz;
        d.depth = (uint8_t)depth;
        d.kind = (uint8_t)kind;
        a.pool[a.list_base[list] + pos[list]++] = d;
        a.node_m[slot] = (uint16_t)sz;
        slot++;
    });
}

// ------------------------------------------------------------------ seed search

struct SearchArgs {
    const NodeDesc *nodes;
    NodeState *state;
    int *running;
    uint32_t *ctl;               // [0] fresh-node cursor
    unsigned long long *work;    // executed lane evaluations of this launch (diagnostic)
    uint32_t n_nodes;
    uint32_t n_waves;
    uint32_t groups;      // groups of 64*U candidates per claimed chunk
    uint32_t leaf;        // leaf size
    uint32_t split_unit;  // part size of an aggregation level
    uint32_t field_bits;  // packed counter field width
    uint32_t magic;       // ceil(2^20 / split_unit)
    uint32_t upper;       // upper aggregation bound (fanout-2 split size)
    uint32_t lut_entries;  // entries of the aggregation increment table (largest node size of the phase)
    uint32_t help_cap;    // helpers join only nodes with fewer waves than this (0: no cap)
    // Fresh nodes are claimed by tickets from ctl[0]: tickets below batch_tickets
    // take claim_batch consecutive nodes each, later tickets one node each (the
    // last nodes of the level, so no wave holds unstarted nodes at the end).
    uint32_t claim_batch;
    uint32_t batch_tickets;
    // Time limit: 0 = none; otherwise waves leave the search once the device wall
    // clock passes *clock0 + limit_ticks, setting bit 3 of *err.
    const unsigned long long *clock0;
    unsigned long long limit_ticks;
    uint32_t *err;
    // Optional (nullptr: off): [2 * wave] start and [2 * wave + 1] exit of every
    // wave on the device wall clock, for the idle share of the launch's tail.
    unsigned long long *wave_clock;
    // Device-driven launch (build plans, nullptr otherwise): the level's node count
    // is count_dev[list] and its nodes start at nodes + lbase_dev[list]; n_nodes,
    // claim_batch and batch_tickets are then derived in the kernel.
    const uint64_t *count_dev;
    const uint64_t *lbase_dev;
    uint32_t list;
    // Optional: the launch does nothing when *stop has a bit of stop_mask set
    // (duplicate keys or an oversized bucket found by the partition).
    const uint32_t *stop;
    uint32_t stop_mask;
    // Aggregation levels: nodes of exactly bt_m keys (the level's full size, most
    // of its nodes) read their counter shift from a byte table indexed by the
    // 16-bit hash (0: no such table; the kernel instance BT builds it).
    uint32_t bt_m;
};

// Node claim parameters of a level of n_nodes nodes for n_waves waves: tickets take
// up to 16 nodes, except for the last 8 per wave (one node per ticket).
__host__ __device__ inline void claim_params(uint32_t n_nodes, uint32_t n_waves, uint32_t &claim_batch,
                                             uint32_t &batch_tickets) {
    claim_batch = 1;
    batch_tickets = 0;
    const uint64_t tail_nodes = (uint64_t)n_waves * 8;
    if (n_nodes > tail_nodes) {
        const uint64_t b = (uint64_t)n_nodes / ((uint64_t)n_waves * 64);
        claim_batch = (uint32_t)(b < 1 ? 1 : (b > 16 ? 16 : b));
        if (claim_batch > 1) batch_tickets = (uint32_t)((n_nodes - tail_nodes) / claim_batch);
    }
}

// Build start on the device wall clock (100 MHz constant clock), the reference
// point of the time limit.
__global__ void k_clock_mark(unsigned long long *t0) {
    if (threadIdx.x == 0) *t0 = (unsigned long long)wall_clock64();
}

// Opaque per-lane zero: indexes derived from it are not provably uniform, so key
// loads become vector loads (vmcnt, L1/L2) instead of scalar loads, whose lgkmcnt
// waits would also block on the LDS table reads of the aggregation search.
__device__ __forceinline__ uint32_t lane_zero() {
    uint32_t z;
    __asm__("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// Streams keys kp[begin, end) (the same for every lane) into registers; the next
// four keys are loaded while the current four are hashed.  The key array carries
// 8 entries of padding, so the look-ahead loads (up to kp[end + 3]) stay inside it.
template <typename F>
__device__ __forceinline__ void stream_keys(const uint64_t *__restrict__ kp, uint32_t begin, uint32_t end, F &&f) {
    const uint64_t *p = kp + lane_zero();
    uint32_t i = begin;
    uint64_t n0 = p[i], n1 = p[i + 1], n2 = p[i + 2], n3 = p[i + 3];
    for (; i + 4 <= end; i += 4) {
        const uint64_t c0 = n0, c1 = n1, c2 = n2, c3 = n3;
        n0 = p[i + 4];
        n1 = p[i + 5];
        n2 = p[i + 6];
        n3 = p[i + 7];
        f(c0);
        f(c1);
        f(c2);
        f(c3);
    }
    if (i < end) f(n0);
    if (i + 1 < end) f(n1);
    if (i + 2 < end) f(n2);
}

// Scalar loads of wave-uniform keys into scalar registers.  Each asm block waits
// for its own load (s_waitcnt in the same block), so no load is in flight when the
// block ends and every output register holds its final value.
__device__ __forceinline__ void sload_keys8(const uint64_t *p, uint64_t (&k)[8]) {
    typedef unsigned int u32x16 __attribute__((ext_vector_type(16)));
    u32x16 r;
    __asm__ volatile("s_load_dwordx16 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=&s"(r) : "s"(p) : "memory");
#pragma unroll
    for (int j = 0; j < 8; j++) k[j] = ((uint64_t)r[2 * j + 1] << 32) | r[2 * j];
}

// Streams keys kp[begin, end) (kp wave-uniform) through scalar registers, eight
// keys per scalar load.  The wait of each load also waits for the LDS reads of
// the aggregation table issued before it (both share one counter), so loads are
// not issued ahead: other waves on the SIMD cover their latency.
template <typename F>
__device__ __forceinline__ void stream_keys_scalar(const uint64_t *__restrict__ kp_any, uint32_t begin_any,
                                                   uint32_t end_any, F &&f) {
    // wave-uniform by construction; readfirstlane lets the compiler keep them
    // (and the loop) in scalar registers
    const uint64_t *kp = (const uint64_t *)uni64((uint64_t)kp_any);
    const uint32_t end = uni(end_any);
    uint32_t i = uni(begin_any);
    for (; i + 8 <= end; i += 8) {
        uint64_t k[8];
        sload_keys8(kp + i, k);
#pragma unroll
        for (int j = 0; j < 8; j++) f(k[j]);
    }
    if (i < end) {
        // the last 1..7 keys in one load: the key array carries 8 entries of
        // padding, so kp[end + 6] is inside it; keys from `end` on are not used
        uint64_t k[8];
        sload_keys8(kp + i, k);
        const uint32_t rem = end - i;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((uint32_t)j < rem) f(k[j]);
    }
}

// Candidate seeds of one group: lane-u candidate = base + off[u] (mod 2^64), base
// wave-uniform, off[u] < 2^32 per lane, max_off >= every off[u].
template <int U>
struct Seeds {
    uint64_t base;
    uint32_t off[U];
    uint32_t max_off;
    // candidate j = u * 64 + lane sits at base + j * stride
    __device__ __forceinline__ Seeds(uint64_t b, uint32_t stride) : base(b) {
        const uint32_t lane = lane_id();
#pragma unroll
        for (int u = 0; u < U; u++) off[u] = (u * 64 + lane) * stride;
        max_off = (U * 64 - 1) * stride;
    }
};

// Calls g(u, h16) for every key of kp[begin, end) and every candidate u, where
// h16 = fold_h16(key + base + off[u]) (rs_common.h).  SK: keys in scalar
// registers; S = key + base is one scalar 64-bit add.  With lo, hi the halves of
// S: while (lo mod 2^16) + max_off < 2^16, adding off[u] changes neither hi nor
// lo >> 16, so the first two steps of the fold, t = (lo+off) ^ hi and
// t ^ (t >> 16), reduce to (lo + off) ^ D with the wave-uniform
// D = hi ^ (hi >> 16) ^ (lo >> 16): one 32-bit add and one xor per candidate.
// Past a 2^16 boundary but not a 2^32 one, t is (lo + off) ^ hi; otherwise (within
// max_off of a 2^32 boundary) the full 64-bit add.  Without SK: vector key
// loads and per-lane 64-bit seeds.
template <int U, bool SK, typename G>
__device__ __forceinline__ void each_hash(const uint64_t *__restrict__ kp, uint32_t begin, uint32_t end,
                                          const Seeds<U> &sd, G &&g) {
    if (SK) {
        const uint64_t base = uni64(sd.base);
        const uint32_t max_off = uni(sd.max_off);
        stream_keys_scalar(kp, begin, end, [&](uint64_t k) {
            const uint64_t S = k + base;
            const uint32_t lo = (uint32_t)S, hi = (uint32_t)(S >> 32);
            if ((lo & 0xFFFFu) <= 0xFFFFu - max_off) {
                const uint32_t D = hi ^ (hi >> 16) ^ (lo >> 16);
#pragma unroll
                for (int u = 0; u < U; u++) g(u, h16_of(hash_post16((lo + sd.off[u]) ^ D)));
            } else if (lo <= ~max_off) {
#pragma unroll
                for (int u = 0; u < U; u++) g(u, h16_of(hash_pre32((lo + sd.off[u]) ^ hi)));
            } else {
#pragma unroll
                for (int u = 0; u < U; u++) g(u, search_h16(S + sd.off[u]));
            }
        });
    } else {
        uint64_t s[U];
#pragma unroll
        for (int u = 0; u < U; u++) s[u] = sd.base + sd.off[u];
        stream_keys(kp, begin, end, [&](uint64_t k) {
#pragma unroll
            for (int u = 0; u < U; u++) g(u, search_h16(k + s[u]));
        });
    }
}

// 64-bit-remap form of each_hash: calls g(u, r) with r = remap64(key + base +
// off[u], m) (rs_common.h), keys in scalar registers.
template <int U, typename G>
__device__ __forceinline__ void each_r64(const uint64_t *__restrict__ kp, uint32_t begin, uint32_t end,
                                         const Seeds<U> &sd, uint32_t m, G &&g) {
    const uint64_t base = uni64(sd.base);
    stream_keys_scalar(kp, begin, end, [&](uint64_t k) {
        const uint64_t S = k + base;
#pragma unroll
        for (int u = 0; u < U; u++) g(u, remap64(S + sd.off[u], m));
    });
}

// Fanout-2 node: lowest x with #{remap(k+S[depth]+x, m) < split} == split
// (SIMDRecSplit.hpp:250-317).  remap < split <=> h16 < ceil(split*2^16/m).
template <int U, bool SK, bool R64>
__device__ __forceinline__ uint64_t group_higher(const uint64_t *__restrict__ kp, uint32_t m, uint64_t sconst,
                                                 uint32_t split, uint32_t H, uint64_t u0) {
    uint32_t cnt[U];
#pragma unroll
    for (int u = 0; u < U; u++) cnt[u] = 0;
    if constexpr (R64)
        each_r64<U>(kp, 0, m, Seeds<U>(u0 + sconst, 1), m, [&](int u, uint32_t r) { cnt[u] += r < split; });
    else
        each_hash<U, SK>(kp, 0, m, Seeds<U>(u0 + sconst, 1), [&](int u, uint32_t h) { cnt[u] += h < H; });
#pragma unroll
    for (int u = 0; u < U; u++) {
        const unsigned long long msk = __ballot(cnt[u] == split);
        if (msk) return u0 + (uint64_t)(u * 64 + __ffsll((long long)msk) - 1);
    }
    return kNone;
}

// Aggregation node: lowest x whose parts floor(remap/SPLIT) have exactly the
// expected counts (SIMDRecSplit.hpp:319-423).  Counts are packed in fields of
// `fb` bits (2^fb >= SPLIT + 2), the last part uncounted; a carry out of any field
// forces another count negative, so equality of the counted fields is
// equality of all counts.  With LUT, the increment 1 << (fb * part) of every
// remapped value r < m is read from the workgroup's LDS table instead of computed.
template <int U, bool WIDE, bool LUT, bool SK, bool R64>
__device__ __forceinline__ uint64_t group_aggr(const uint64_t *__restrict__ kp, uint32_t m, uint64_t sconst,
                                               uint32_t fb, uint32_t magic, uint64_t T, uint64_t cmask,
                                               uint32_t lut_hi, uint32_t two, uint64_t u0) {
    using C = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    C cnt[U];
#pragma unroll
    for (int u = 0; u < U; u++) cnt[u] = 0;
    using LutT = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    if constexpr (R64) {
        // LDS table at entry lut_hi >> 16, indexed by the remapped value
        const uint32_t lut_b = (lut_hi >> 16) * (uint32_t)sizeof(LutT);
        each_r64<U>(kp, 0, m, Seeds<U>(u0 + sconst, 1), m, [&](int u, uint32_t r) {
            if (LUT) {
                cnt[u] += *(const __attribute__((address_space(3))) LutT *)(uintptr_t)(lut_b + (uint32_t)sizeof(LutT) * r);
            } else {
                const uint32_t part = __umul24(r, magic) >> 20;
                cnt[u] += (C)1 << __umul24(part, fb);
            }
        });
    } else {
        each_hash<U, SK>(kp, 0, m, Seeds<U>(u0 + sconst, 1), [&](int u, uint32_t h) {
            if (LUT) {
                cnt[u] += *(const __attribute__((address_space(3))) LutT *)(uintptr_t)lut_byte_addr(h, m, lut_hi, two);
            } else {
                const uint32_t part = __umul24(__umul24(h, m) >> 16, magic) >> 20;
                cnt[u] += (C)1 << __umul24(part, fb);
            }
        });
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const unsigned long long msk = __ballot(((uint64_t)cnt[u] & cmask) == T);
        if (msk) return u0 + (uint64_t)(u * 64 + __ffsll((long long)msk) - 1);
    }
    return kNone;
}

// Byte table of the BT instance, 65536 entries in the workgroup's static LDS
// (two workgroups of kBtThreads threads per CU).  Aggregation levels: entry h (a
// 16-bit hash) holds fb * part for the part floor(((h * bt_m) >> 16) / SPLIT) of a
// node of bt_m keys, so one LDS byte read replaces the remap and the increment
// table lookup.  Rotation leaves: entry h holds the position (h * L) >> 16.
// Workgroup size measured at C2: 768 threads (six waves per SIMD) beat 512 (four)
// by 3%; 640 (ten waves, which do not split evenly over four SIMDs) ran 17% slower
// than 512.
#ifndef RS_BT_THREADS
#define RS_BT_THREADS 768
#endif
constexpr uint32_t kBtThreads = RS_BT_THREADS;
#ifndef RS_BT_THREADS_LEAF
#define RS_BT_THREADS_LEAF 768
#endif
// the rotation-leaf instances: 768 as well (512, four waves per SIMD without the
// spills of the 80-register bound, measured slower: C3 leaves 319 -> 351 ms)
constexpr uint32_t kBtThreadsLeaf = RS_BT_THREADS_LEAF;
__host__ __device__ constexpr uint32_t bt_threads(int kind) { return kind == kLeafRot ? kBtThreadsLeaf : kBtThreads; }

template <bool BT>
__device__ __forceinline__ const uint8_t *bt_table() {
    if constexpr (BT) {
        __shared__ uint32_t s_bt[65536 / 4];
        return (const uint8_t *)s_bt;
    } else {
        return nullptr;
    }
}

// Aggregation node of the byte table's size: the same predicate as group_aggr,
// the increment of every key 1 << bt[h16] (h16 < 2^16; 64-bit counters for WIDE).
template <int U, bool SK, bool WIDE>
__device__ __forceinline__ uint64_t group_aggr_bt(const uint64_t *__restrict__ kp, uint32_t m, uint64_t sconst,
                                                  const uint8_t *bt, uint64_t T, uint64_t cmask, uint64_t u0) {
    using C = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    C cnt[U];
#pragma unroll
    for (int u = 0; u < U; u++) cnt[u] = 0;
    each_hash<U, SK>(kp, 0, m, Seeds<U>(u0 + sconst, 1), [&](int u, uint32_t h) { cnt[u] += (C)1 << bt[h]; });
#pragma unroll
    for (int u = 0; u < U; u++) {
        const unsigned long long msk = __ballot(((uint64_t)cnt[u] & cmask) == T);
        if (msk) return u0 + (uint64_t)(u * 64 + __ffsll((long long)msk) - 1);
    }
    return kNone;
}

// Rotation leaves of at most this many keys test candidates through the table of
// least rotations (2^L two-byte entries of LDS).
constexpr uint32_t kNecklaceMax = 12;

// Full leaf with rotation fitting (SIMDRecSplit.hpp:452-509): unit j is the base
// seed j*L; keys with even `second` set bits of the left mask, odd ones of the
// right mask; the lowest r with left | rotl(right, r) == 2^L-1 gives seed j*L+r.
// The partition places a full leaf's even keys first: kp[0, nl) even, kp[nl, L) odd.
template <int U, bool SK, bool R64, int RB, bool BT = false>
__device__ __forceinline__ uint64_t group_leaf_rot(const uint64_t *__restrict__ kp, uint32_t L, uint32_t nl,
                                                   uint32_t nr, uint64_t u0, const uint16_t *canon) {
    // BT: position (h16 * L) >> 16 of every 16-bit hash read from the byte table
    const uint8_t *bt = bt_table<BT>();
    const uint64_t S = c_start_seed[kSeedLeaf];
    const uint32_t full = (uint32_t)((1ull << L) - 1);
    const Seeds<U> sd(u0 * L + S, L);
    const uint32_t lane = lane_id();
    if constexpr (!RB) {
        // Leaves of 17 to 24 keys: two masks per candidate.  Few candidates per
        // group are collision-free (1.6 of 256 at L = 18); they are tested in index
        // order one at a time across the wave: lane r tests rotation r.
        uint32_t ml[U], mr[U];
#pragma unroll
        for (int u = 0; u < U; u++) ml[u] = mr[u] = 0;
        if constexpr (R64) {
            each_r64<U>(kp, 0, nl, sd, L, [&](int u, uint32_t r) { ml[u] |= 1u << r; });
            each_r64<U>(kp, nl, L, sd, L, [&](int u, uint32_t r) { mr[u] |= 1u << r; });
        } else if constexpr (BT) {
            each_hash<U, SK>(kp, 0, nl, sd, [&](int u, uint32_t h) { ml[u] |= 1u << bt[h]; });
            each_hash<U, SK>(kp, nl, L, sd, [&](int u, uint32_t h) { mr[u] |= 1u << bt[h]; });
        } else {
            each_hash<U, SK>(kp, 0, nl, sd, [&](int u, uint32_t h) { ml[u] |= 1u << (__umul24(h, L) >> 16); });
            each_hash<U, SK>(kp, nl, L, sd, [&](int u, uint32_t h) { mr[u] |= 1u << (__umul24(h, L) >> 16); });
        }
        const uint32_t r = lane < L ? lane : 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            unsigned long long cand = __ballot((uint32_t)__popc(ml[u]) == nl && (uint32_t)__popc(mr[u]) == nr);
            while (cand) {
                const int l = __ffsll((long long)cand) - 1;
                cand &= cand - 1;
                const uint32_t bl = (uint32_t)__builtin_amdgcn_readlane((int)ml[u], l);
                const uint32_t br = (uint32_t)__builtin_amdgcn_readlane((int)mr[u], l);
                const uint32_t rot = r == 0 ? br : (((br << r) | (br >> (L - r))) & full);
                const unsigned long long fit = __ballot(lane < L && (bl | rot) == full);
                if (fit) return (u0 + (uint64_t)(u * 64 + l)) * L + (uint64_t)(__ffsll((long long)fit) - 1);
            }
        }
        return kNone;
    } else {
        // Leaves of at most 16 keys (RB): both masks in one word, the left one in
        // bits 0..L-1 and the right one in bits 16..16+L-1 (the 16 folds into the
        // remap's multiply-add); both halves are collision-free exactly when L bits
        // are set.  Many candidates pass (13 of 256 at L = 12, 43 at L = 8): they go
        // to a per-wave LDS list in unit order (u, then lane), and each step tests
        // floor(64 / L) of them: lane i tests rotation i mod L of candidate i / L,
        // so the lowest set bit of the ballot is the lowest fitting rotation of the
        // first candidate that has one.
        uint32_t mk[U];
#pragma unroll
        for (int u = 0; u < U; u++) mk[u] = 0;
        if constexpr (R64) {
            each_r64<U>(kp, 0, nl, sd, L, [&](int u, uint32_t r) { mk[u] |= 1u << r; });
            each_r64<U>(kp, nl, L, sd, L, [&](int u, uint32_t r) { mk[u] |= 1u << (r + 16); });
        } else if constexpr (BT) {
            uint32_t mr[U];
#pragma unroll
            for (int u = 0; u < U; u++) mr[u] = 0;
            each_hash<U, SK>(kp, 0, nl, sd, [&](int u, uint32_t h) { mk[u] |= 1u << bt[h]; });
            each_hash<U, SK>(kp, nl, L, sd, [&](int u, uint32_t h) { mr[u] |= 1u << bt[h]; });
#pragma unroll
            for (int u = 0; u < U; u++) mk[u] |= mr[u] << 16;
        } else {
            each_hash<U, SK>(kp, 0, nl, sd, [&](int u, uint32_t h) { mk[u] |= 1u << (__umul24(h, L) >> 16); });
            each_hash<U, SK>(kp, nl, L, sd,
                             [&](int u, uint32_t h) { mk[u] |= 1u << ((__umul24(h, L) + (16u << 16)) >> 16); });
        }
        if constexpr (RB == 2) {
            // Leaves of at most kNecklaceMax keys: canon[x] is the least rotation of
            // the L-bit mask x.  Some rotation of the right mask equals the left
            // mask's complement exactly when both have the same least rotation
            // (which also makes both halves collision-free: a collision changes a
            // popcount), so one pair of table reads per candidate replaces the list
            // and its steps; the first matching candidate in unit order then finds
            // its lowest fitting rotation across the wave (lane r tests rotation r).
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool match = canon[mk[u] >> 16] == canon[~mk[u] & full];
                const unsigned long long bal = __ballot(match);
                if (bal) {
                    const int l = __ffsll((long long)bal) - 1;
                    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)mk[u], l);
                    const uint32_t bl = e & 0xFFFFu, br = e >> 16;
                    const uint32_t r = lane < L ? lane : 0;
                    const uint32_t rot = r == 0 ? br : (((br << r) | (br >> (L - r))) & full);
                    const unsigned long long fit = __ballot(lane < L && (bl | rot) == full);
                    return (u0 + (uint64_t)(u * 64 + l)) * L + (uint64_t)(__ffsll((long long)fit) - 1);
                }
            }
            return kNone;
        }
        // one list per wave of the workgroup: masks, and unit indexes (< 256) as bytes
        constexpr int kWaves = (BT ? kBtThreadsLeaf : 256) / 64;
        __shared__ uint32_t s_cand_m[kWaves][U * 64];
        __shared__ uint8_t s_cand_u[kWaves][U * 64];
        uint32_t *cm = s_cand_m[threadIdx.x >> 6];
        uint8_t *cu = s_cand_u[threadIdx.x >> 6];
        uint32_t count = 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool ok = (uint32_t)__popc(mk[u]) == L;
            const unsigned long long bal = __ballot(ok);
            const uint32_t below =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (ok) {
                cm[count + below] = mk[u];
                cu[count + below] = (uint8_t)(u * 64 + lane);
            }
            count += (uint32_t)__popcll(bal);
        }
        if (count == 0) return kNone;
        // the list is read back by other lanes of this wave: wait for the writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t per_step = 64 / L;
        const uint32_t g = (lane * ((65536 + L - 1) / L)) >> 16;  // lane / L (lane < 64, L <= 16)
        const uint32_t r = lane - g * L;
        for (uint32_t base = 0; base < count; base += per_step) {
            const uint32_t idx = base + g;
            const bool valid = g < per_step && idx < count;
            const uint32_t e = valid ? cm[idx] : 0;
            const uint32_t bl = e & 0xFFFFu, br = e >> 16;
            const uint32_t rot = r == 0 ? br : (((br << r) | (br >> (L - r))) & full);
            const unsigned long long fit = __ballot(valid && (bl | rot) == full);
            if (fit) {
                const uint32_t first = (uint32_t)__ffsll((long long)fit) - 1;
                const uint32_t gi = first / L;
                const uint32_t unit = cu[base + gi];
                return (u0 + (uint64_t)unit) * L + (uint64_t)(first - gi * L);
            }
        }
        return kNone;
    }
}

// Leaf without rotation (SIMDRecSplit.hpp:510-599): lowest x whose m keys land on
// m distinct positions.  A prefix of ceil(3 sqrt(m)) keys must already be
// collision-free; waves skip the rest when no lane passes (lossless).
template <int U, bool SK, bool R64>
__device__ __forceinline__ uint64_t group_leaf_brute(const uint64_t *__restrict__ kp, uint32_t m, uint32_t mid,
                                                     uint64_t u0) {
    const uint64_t S = c_start_seed[kSeedLeaf];
    const uint32_t full = (uint32_t)((1ull << m) - 1);
    uint32_t mk[U];
#pragma unroll
    for (int u = 0; u < U; u++) mk[u] = 0;
    const Seeds<U> sd(u0 + S, 1);
    auto part = [&](uint32_t b, uint32_t e) {
        if constexpr (R64) each_r64<U>(kp, b, e, sd, m, [&](int u, uint32_t r) { mk[u] |= 1u << r; });
        else each_hash<U, SK>(kp, b, e, sd, [&](int u, uint32_t h) { mk[u] |= 1u << (__umul24(h, m) >> 16); });
    };
    part(0, mid);
    if (mid < m) {
        bool any = false;
#pragma unroll
        for (int u = 0; u < U; u++) any |= (uint32_t)__popc(mk[u]) == mid;
        if (!__ballot(any)) return kNone;
        part(mid, m);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const unsigned long long msk = __ballot(mk[u] == full);
        if (msk) return u0 + (uint64_t)(u * 64 + __ffsll((long long)msk) - 1);
    }
    return kNone;
}

__device__ __forceinline__ bool has_work(const NodeState *st, uint64_t units_per_chunk, uint64_t scale) {
    const unsigned long long nc = ld_relaxed(&st->next_chunk);
    const unsigned long long res = ld_relaxed(&st->result);
    return nc * units_per_chunk * scale < res;
}

// RB (rotation leaves of at most 16 keys): packed masks, and the candidate test
// batched through LDS (RB = 1) or, for at most kNecklaceMax keys, through the
// table of least rotations (RB = 2), with registers held to five waves per SIMD.  BT
// (aggregation levels, 32-bit counters, 16-bit remap): the byte table for nodes of
// bt_m keys, workgroups of kBtThreads threads, two per CU (registers held to the
// kBtThreads / 128 waves per SIMD that two such workgroups need).
template <int KIND, int U, bool WIDE, bool LUT, bool SK, bool R64, int RB = 0, bool BT = false>
__global__ __launch_bounds__(BT ? bt_threads(KIND) : 256, BT ? bt_threads(KIND) / 128 : (RB ? 5 : 1)) void k_search(const uint64_t *__restrict__ keys,
                                                                                  SearchArgs a) {
    static_assert(!BT || ((((KIND == kUpper || KIND == kLower) && LUT && !RB) || KIND == kLeafRot) && !R64),
                  "the byte table serves aggregation levels and rotation leaves over the 16-bit remap");
    extern __shared__ uint32_t s_lut[];
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (a.wave_clock && lane == 0) a.wave_clock[2 * wave] = (unsigned long long)wall_clock64();
    if (a.stop && (uni(*a.stop) & a.stop_mask)) return;
    uint32_t n_nodes = a.n_nodes, claim_batch = a.claim_batch, batch_tickets = a.batch_tickets;
    const NodeDesc *nodes = a.nodes;
    if (a.count_dev) {
        n_nodes = uni((uint32_t)a.count_dev[a.list]);
        nodes = a.nodes + uni64(a.lbase_dev[a.list]);
        claim_params(n_nodes, a.n_waves, claim_batch, batch_tickets);
    }
    const uint64_t units_per_chunk = (uint64_t)a.groups * U * 64;
    const uint64_t scale = KIND == kLeafRot ? a.leaf : 1;
    // Aggregation increment table, one per workgroup: entry r (a remapped value)
    // holds 1 << (fb * floor(r / split_unit)).  It depends on r alone, not on the
    // node, so it is written once before the first claim.  4-byte entries, 8-byte
    // ones for 64-bit counters (WIDE).
    constexpr uint32_t kLutShift = WIDE ? 3 : 2;
    if constexpr (LUT && (KIND == kUpper || KIND == kLower)) {
        for (uint32_t r = threadIdx.x; r < a.lut_entries; r += blockDim.x) {
            if (WIDE) ((uint64_t *)s_lut)[r] = 1ull << (a.field_bits * (r / a.split_unit));
            else s_lut[r] = 1u << (a.field_bits * (r / a.split_unit));
        }
        if constexpr (BT) {
            uint32_t *bt_w = (uint32_t *)bt_table<BT>();
            const uint32_t bm = a.bt_m, fbits = a.field_bits, magic = a.magic;
            for (uint32_t w = threadIdx.x; w < 65536 / 4; w += blockDim.x) {
                uint32_t v = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++) {
                    const uint32_t r = ((w * 4 + j) * bm) >> 16;  // < bm < 3000
                    v |= (fbits * ((r * magic) >> 20)) << (8 * j);  // r / SPLIT
                }
                bt_w[w] = v;
            }
        }
        __syncthreads();
    }
    if constexpr (BT && KIND == kLeafRot) {
        uint32_t *bt_w = (uint32_t *)bt_table<BT>();
        const uint32_t L = a.bt_m;
        for (uint32_t w = threadIdx.x; w < 65536 / 4; w += blockDim.x) {
            uint32_t v = 0;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) v |= (((w * 4 + j) * L) >> 16) << (8 * j);
            bt_w[w] = v;
        }
        __syncthreads();
    }
    if constexpr (KIND == kLeafRot && RB == 2) {
        // least rotation of every L-bit mask, in the dynamic LDS (2^L entries)
        uint16_t *canon = (uint16_t *)s_lut;
        const uint32_t L = a.leaf, full = (1u << L) - 1;
        for (uint32_t x = threadIdx.x; x <= full; x += blockDim.x) {
            uint32_t c = x, y = x;
            for (uint32_t t = 1; t < L; t++) {
                y = ((y << 1) | (y >> (L - 1))) & full;
                c = y < c ? y : c;
            }
            canon[x] = (uint16_t)c;
        }
        __syncthreads();
    }
    // table start in entries, in the upper half (the low 32 bits of a flat LDS
    // address are its LDS offset)
    const uint32_t lut_hi = uni(((uint32_t)(uintptr_t)s_lut >> kLutShift) << 16);
    const uint32_t two = kLutShift;
    int node = -1;
    bool helping = false;
    bool fresh = false;                     // node came from the fresh-node tickets
    uint32_t span_next = 1;                 // chunks to claim at once on this node
    uint32_t span_cap = 1;                  // at most 2^19 key evaluations per claim, at most 128 chunks
    uint32_t batch_next = 0, batch_end = 0;  // claimed fresh nodes not yet started
    uint32_t scan = wave;
    uint64_t work = 0;
    // cached per-node values (wave-uniform)
    uint64_t key_off = 0, sconst = 0, T = 0, cmask = 0;
    uint32_t m = 0, aux0 = 0, aux1 = 0;
    const unsigned long long t_end = a.limit_ticks ? uni64(*a.clock0) + a.limit_ticks : ~0ull;

    for (;;) {
        if (node < 0) {
            fresh = false;
            if (!helping) {
                if (batch_next >= batch_end) {
                    // tickets come from one counter, so only resident waves hold
                    // nodes (workgroups beyond what fits start as others leave)
                    uint32_t t = 0;
                    if (lane == 0) t = atomicAdd(&a.ctl[0], 1u);
                    t = uni(t);
                    if (t < batch_tickets) {
                        batch_next = t * claim_batch;
                        batch_end = batch_next + claim_batch;
                    } else {
                        const uint32_t c = batch_tickets * claim_batch + (t - batch_tickets);
                        batch_next = c;
                        batch_end = c < n_nodes ? c + 1 : c;
                    }
                }
                if (batch_next < batch_end) {
                    node = (int)batch_next++;
                    fresh = true;
                } else {
                    helping = true;
                }
            }
            if (helping) {
                int found = -1;
                // Among the running entries of a window of 256 (four per lane, the
                // loads of a window issued together) with work left, join the node
                // with the fewest waves (spreads helpers evenly over the nodes still
                // open; a node's chunk overshoot grows with its waves).
                for (uint32_t k = 0; k < a.n_waves && found < 0; k += 256) {
                    int cand[4];
                    uint32_t idx[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        idx[j] = scan + k + 4 * lane + j;
                        while (idx[j] >= a.n_waves) idx[j] -= a.n_waves;
                        cand[j] = k + 4 * lane + j < a.n_waves
                                      ? __hip_atomic_load(&a.running[idx[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : -1;
                    }
                    uint32_t best = 0xFFFFFFFFu;
                    int best_cand = -1;
                    uint32_t best_idx = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        bool ok = cand[j] >= 0 && has_work(&a.state[cand[j]], units_per_chunk, scale);
                        uint32_t w = 0;
                        if (ok) {
                            w = __hip_atomic_load(&a.state[cand[j]].waves, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            ok = a.help_cap == 0 || w < a.help_cap;
                        }
                        if (ok && min(w, 0xFFFFu) < best) {
                            best = min(w, 0xFFFFu);
                            best_cand = cand[j];
                            best_idx = idx[j];
                        }
                    }
                    if (__ballot(best_cand >= 0)) {
                        uint32_t key = best_cand >= 0 ? (best << 16) | lane : 0xFFFFFFFFu;
#pragma unroll
                        for (int o = 32; o > 0; o >>= 1) key = min(key, (uint32_t)__shfl_xor(key, o));
                        const int l = (int)(uni(key) & 0xFFFF);
                        found = (int)uni((uint32_t)__shfl(best_cand, l));
                        scan = uni(__shfl(best_idx, l));
                    }
                }
                if (found < 0) break;
                node = found;
            }
            node = (int)uni((uint32_t)node);
            span_next = 1;
            span_cap = 0;  // set below from the node's size
            if (lane == 0) {
                __hip_atomic_store(&a.running[wave], node, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicAdd(&a.state[node].waves, 1u);
            }
            const NodeDesc d = nodes[node];
            key_off = uni64(d.key_off);
            m = uni(d.m);
            {
                // the power of two at most 2^19 / (key evaluations per chunk), 1..128
                const uint32_t e = uni((uint32_t)units_per_chunk * m);  // < 2^22
                const int lg = 32 - __clz(e);                           // e < 2^lg
                span_cap = lg >= 19 ? 1u : min(128u, 1u << (19 - lg));
                span_cap = uni(span_cap);
            }
            if (KIND == kHigher) {
                sconst = c_start_seed[uni(d.depth)];
                aux0 = higher_split(SplitParams{a.leaf, 0, a.upper}, m);           // split
                aux1 = (uint32_t)(((uint64_t)aux0 * 65536 + m - 1) / m);            // H
            } else if (KIND == kUpper || KIND == kLower) {
                sconst = c_start_seed[KIND == kUpper ? kSeedUpper : kSeedLower];
                const uint32_t su = a.split_unit;
                const uint32_t fan = (m + su - 1) / su;
                T = 0;
                for (uint32_t p = 0; p + 1 < fan; p++) T |= (uint64_t)su << (p * a.field_bits);
                const uint32_t w = (fan - 1) * a.field_bits;
                cmask = w >= 64 ? ~0ull : ((1ull << w) - 1);
            } else if (KIND == kLeafRot) {
                // m <= 24 keys: lane i reads key i, a ballot counts the odd ones
                const bool odd_key = lane < m && (keys[key_off + lane] & 1) != 0;
                const uint32_t odd = (uint32_t)__popcll(__ballot(odd_key));
                aux0 = m - odd;  // nl
                aux1 = odd;      // nr
            } else {
                uint32_t mid = m;
                if (m >= 14) {
                    // ceil(3 sqrt(m)) for m in [14, 24)
                    mid = 1;
                    while (mid * mid < 9 * m) mid++;
                    if (mid > m) mid = m;
                }
                aux0 = mid;
            }
        }
        if (a.limit_ticks) {
            // past the time limit: every wave leaves at its next claim, so the
            // launch drains within one chunk of work per wave
            uint32_t late = 0;
            if (lane == 0) late = (unsigned long long)wall_clock64() > t_end;
            if (uni(late)) {
                if (lane == 0) {
                    atomicOr(a.err, 8u);
                    __hip_atomic_store(&a.running[wave], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    atomicSub(&a.state[node].waves, 1u);
                }
                break;
            }
        }
        unsigned long long chunk = 0, res = kNone;
        uint32_t span = 1;  // consecutive chunks taken by this claim
        if (fresh) {
            // chunk 0 belongs to the wave that starts the node (helpers claim from
            // 1 on), so it needs no claim
            fresh = false;
        } else {
            // A node shared by many waves would be bound by its one chunk counter
            // (thousands of waves on the last heavy nodes of a level): the span
            // grows with the node's waves seen at the previous claim, 1 below 32.
            span = span_next;
            unsigned long long w = 0;
            if (lane == 0) {
                chunk = atomicAdd(&a.state[node].next_chunk, (unsigned long long)span);
                res = ld_relaxed(&a.state[node].result);
                w = __hip_atomic_load(&a.state[node].waves, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            chunk = uni64(chunk);
            res = uni64(res);
            span_next = uni(min(max(uni((uint32_t)w) >> 4, 1u), span_cap));
        }
        const uint64_t unit0 = chunk * units_per_chunk;
        if (unit0 * scale >= res || unit0 * scale >= kSeedLimit) {
            if (unit0 * scale >= kSeedLimit && lane == 0) atomicMin(&a.state[node].result, (unsigned long long)kSeedLimit);
            if (lane == 0) {
                __hip_atomic_store(&a.running[wave], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicSub(&a.state[node].waves, 1u);
            }
            node = -1;
            continue;
        }
        const uint64_t *kp = keys + key_off;
        uint64_t hit = kNone;
        for (uint32_t g = 0; g < a.groups * span; g++) {
            const uint64_t u0 = unit0 + (uint64_t)g * U * 64;
            if (g > 0) {
                unsigned long long r = 0;
                if (lane == 0) r = ld_relaxed(&a.state[node].result);
                r = uni64(r);
                if (u0 * scale >= r) break;
            }
            work += (uint64_t)U * 64 * m;
            if (KIND == kHigher) hit = group_higher<U, SK, R64>(kp, m, sconst, aux0, aux1, u0);
            else if ((KIND == kUpper || KIND == kLower) && BT && m == a.bt_m)
                hit = group_aggr_bt<U, SK, WIDE>(kp, m, sconst, bt_table<BT>(), T, cmask, u0);
            else if (KIND == kUpper || KIND == kLower)
                hit = group_aggr<U, WIDE, LUT, SK, R64>(kp, m, sconst, a.field_bits, a.magic, T, cmask, lut_hi, two,
                                                        u0);
            else if (KIND == kLeafRot)
                hit = group_leaf_rot<U, SK, R64, RB, BT>(kp, m, aux0, aux1, u0, (const uint16_t *)s_lut);
            else hit = group_leaf_brute<U, SK, R64>(kp, m, aux0, u0);
            if (hit != kNone) break;
        }
        if (hit != kNone) {
            // every chunk this wave could still claim lies above this one, so
            // above the hit: leave the node
            if (lane == 0) {
                atomicMin(&a.state[node].result, (unsigned long long)hit);
                __hip_atomic_store(&a.running[wave], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicSub(&a.state[node].waves, 1u);
            }
            node = -1;
        }
    }
    if (lane == 0) {
        __hip_atomic_store(&a.running[wave], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (work) atomicAdd(a.work, (unsigned long long)work);
        if (a.wave_clock) a.wave_clock[2 * wave + 1] = (unsigned long long)wall_clock64();
    }
}

__global__ __launch_bounds__(256) void k_state_init(NodeState *st, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st[i].result = kNone;
        st[i].next_chunk = 1;  // chunk 0 is taken by the wave that starts the node
        st[i].waves = 0;
    }
}

// Leaves: copy each node's final seed to its slot.  With count_dev (build plans)
// the level's size and first node come from the device, grid-stride.
__global__ __launch_bounds__(256) void k_finalize(const NodeDesc *__restrict__ nodes, const NodeState *__restrict__ st,
                                                  uint32_t n, uint64_t *__restrict__ seeds, uint32_t *__restrict__ err,
                                                  const uint64_t *__restrict__ count_dev = nullptr,
                                                  const uint64_t *__restrict__ lbase_dev = nullptr, uint32_t list = 0) {
    if (count_dev) {
        n = (uint32_t)count_dev[list];
        nodes += lbase_dev[list];
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t x = st[i].result;
        seeds[nodes[i].slot] = x;
        if (x >= kSeedLimit) atomicOr(err, 2u);
    }
}

// Split nodes: record the seed and partition the node's keys in place by child,
// one wavefront per node, keys staged in LDS.  Inside every child the keys with
// even `second` come first (stable within each parity): full leaves then split
// into their left and right rotation sets by position.  Predicates are order-free,
// so this layout does not change any seed.
__device__ __forceinline__ void split_node(uint64_t *__restrict__ keys, const NodeDesc *__restrict__ nodes,
                                           const NodeState *__restrict__ st, uint64_t *__restrict__ seeds,
                                           uint32_t *__restrict__ err, SplitParams P, uint32_t kind, int r64,
                                           uint32_t node, uint64_t *s_keys) {
    const uint32_t lane = threadIdx.x;
    const NodeDesc d = nodes[node];
    const uint64_t x = st[node].result;
    if (lane == 0) seeds[d.slot] = x;
    if (x >= kSeedLimit) {
        if (lane == 0) atomicOr(err, 2u);
        return;
    }
    const uint32_t m = d.m;
    uint64_t *kp = keys + d.key_off;
    for (uint32_t i = lane; i < m; i += 64) s_keys[i] = kp[i];
    __syncthreads();
    uint64_t seed_abs;
    uint32_t unit, fan;
    if (kind == kHigher) {
        seed_abs = c_start_seed[d.depth] + x;
        unit = higher_split(P, m);
        fan = 2;
    } else {
        seed_abs = c_start_seed[kind == kUpper ? kSeedUpper : kSeedLower] + x;
        unit = kind == kUpper ? P.lower : P.leaf;
        fan = (m + unit - 1) / unit;
    }
    auto child_of = [&](uint64_t key) -> uint32_t {
        const uint32_t r = remap_any(key + seed_abs, m, r64 != 0);
        return kind == kHigher ? (r >= unit ? 1u : 0u) : r / unit;
    };
    // pass 1: even keys per child
    uint32_t evens[kMaxFanout];
#pragma unroll
    for (uint32_t c = 0; c < kMaxFanout; c++) evens[c] = 0;
    for (uint32_t j0 = 0; j0 < m; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool valid = j < m;
        const uint64_t key = valid ? s_keys[j] : 0;
        const uint32_t cid = child_of(key);
        const bool even = (key & 1) == 0;
#pragma unroll
        for (uint32_t c = 0; c < kMaxFanout; c++)
            if (c < fan) evens[c] += (uint32_t)__popcll(__ballot(valid && even && cid == c));
    }
    // pass 2: scatter
    uint32_t base_e[kMaxFanout], base_o[kMaxFanout];
#pragma unroll
    for (uint32_t c = 0; c < kMaxFanout; c++) {
        base_e[c] = c * unit;
        base_o[c] = c * unit + evens[c];
    }
    const unsigned long long lt = (1ull << lane) - 1;
    for (uint32_t j0 = 0; j0 < m; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool valid = j < m;
        const uint64_t key = valid ? s_keys[j] : 0;
        const uint32_t cid = child_of(key);
        const bool even = (key & 1) == 0;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t c = 0; c < kMaxFanout; c++) {
            if (c < fan) {
                const unsigned long long me = __ballot(valid && even && cid == c);
                const unsigned long long mo = __ballot(valid && !even && cid == c);
                if (valid && cid == c) pos = even ? base_e[c] + (uint32_t)__popcll(me & lt) : base_o[c] + (uint32_t)__popcll(mo & lt);
                base_e[c] += (uint32_t)__popcll(me);
                base_o[c] += (uint32_t)__popcll(mo);
            }
        }
        if (valid) kp[pos] = key;
    }
}

__global__ __launch_bounds__(64) void k_split(uint64_t *__restrict__ keys, const NodeDesc *__restrict__ nodes,
                                              const NodeState *__restrict__ st, uint64_t *__restrict__ seeds,
                                              uint32_t *__restrict__ err, SplitParams P, uint32_t kind, int r64,
                                              const uint64_t *__restrict__ count_dev = nullptr,
                                              const uint64_t *__restrict__ lbase_dev = nullptr, uint32_t list = 0) {
    extern __shared__ uint64_t s_keys[];
    // one node per workgroup; with count_dev (build plans) the level's size and
    // first node come from the device and workgroups stride over the nodes
    uint32_t n_nodes = gridDim.x;
    if (count_dev) {
        n_nodes = (uint32_t)count_dev[list];
        nodes += lbase_dev[list];
    }
    for (uint32_t node = blockIdx.x; node < n_nodes; node += gridDim.x) {
        split_node(keys, nodes, st, seeds, err, P, kind, r64, node, s_keys);
        __syncthreads();  // s_keys is reused by the next node
    }
}


// ------------------------------------------------------------------ encoding

struct EncodeArgs {
    const uint64_t *acc;        // bucket key offsets, nb+1
    const uint64_t *node_base;  // bucket slot offsets, nb+1
    const uint64_t *seeds;
    const uint16_t *node_m;
    const uint32_t *memo;
    const uint32_t *tree_nodes;  // nodes of a bucket by size (unmasked)
    uint64_t nb;
    SplitParams P;
    int rotate;
    // Optional (build plans): the encoding kernels write nothing when *stop has a
    // bit of kErrStopEncode set (a size bound or the partition failed).
    const uint32_t *stop;
};

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += bcast64_xor(v, o);
    return v;
}

// Per bucket (one wavefront): bits of the bucket's descriptor = fixed bits of the
// whole subtree (memo) + unary bits of every node; algorithmic evaluations per kind
// in evals[0, kNumKinds) and candidate seeds tested per kind in evals[kNumKinds,
// 2 kNumKinds) (x + 1 per split or brute-force node; (floor(x / L) + 1) L per
// rotation leaf, whose L rotations of one pass over its keys test L seeds).
__global__ __launch_bounds__(256) void k_bucket_bits(EncodeArgs a, uint64_t *__restrict__ bits,
                                                     unsigned long long *__restrict__ evals) {
    // waves stride over the buckets and keep their evaluation counts in registers:
    // one atomic per kind per wave at the end, not per bucket
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t n_waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t lane = lane_id();
    uint64_t ev[kNumKinds] = {0, 0, 0, 0, 0}, sd[kNumKinds] = {0, 0, 0, 0, 0};
    for (uint64_t b = wave; b < a.nb; b += n_waves) {
        const uint32_t m = (uint32_t)(a.acc[b + 1] - a.acc[b]);
        if (m < 2 || m >= kMaxBucket) {  // oversized buckets fail the build; no table entry
            if (lane == 0) bits[b] = 0;
            continue;
        }
        const uint64_t base = a.node_base[b];
        const uint32_t nn = a.tree_nodes[m];
        uint64_t un = 0;
        for (uint32_t j = lane; j < nn; j += 64) {
            const uint32_t s = a.node_m[base + j];
            const uint32_t g = a.memo[s] >> 27;
            const uint64_t x = a.seeds[base + j];
            un += (x >> g) + 1;
            const uint32_t kind = classify(a.P, s, a.rotate != 0);
            const uint64_t e = kind == kLeafRot ? (x / s + 1) * s : (x + 1) * s;
            const uint64_t t = kind == kLeafRot ? e : x + 1;
#pragma unroll
            for (int k = 0; k < kNumKinds; k++) {
                ev[k] += (uint32_t)k == kind ? e : 0;
                sd[k] += (uint32_t)k == kind ? t : 0;
            }
        }
        un = wave_sum64(un);
        if (lane == 0) bits[b] = (a.memo[m] & 0xFFFF) + un;
    }
#pragma unroll
    for (int k = 0; k < kNumKinds; k++) {
        ev[k] = wave_sum64(ev[k]);
        sd[k] = wave_sum64(sd[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kNumKinds; k++) {
            if (ev[k]) atomicAdd(&evals[k], (unsigned long long)ev[k]);
            if (sd[k]) atomicAdd(&evals[kNumKinds + k], (unsigned long long)sd[k]);
        }
    }
}

__device__ __forceinline__ void or_bits(unsigned long long *w, uint64_t pos, uint64_t val, uint32_t width) {
    if (width == 0 || val == 0) return;
    const uint64_t wi = pos >> 6, sh = pos & 63;
    atomicOr(&w[wi], (unsigned long long)(val << sh));
    if (sh + width > 64) atomicOr(&w[wi + 1], (unsigned long long)(val >> (64 - sh)));
}

template <typename T>
__device__ __forceinline__ T wave_excl_scan(T v, T &total) {
    T inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T t = (T)bcast64_up(inc, o);
        if (lane_id() >= (uint32_t)o) inc += t;
    }
    total = (T)bcast64(inc, 63);
    return inc - v;
}

// Per bucket (one wavefront): write fixed parts in DFS order, then the unary
// parts (RiceBitVector.hpp:73-106, SIMDRecSplit.hpp:707-714).
__global__ __launch_bounds__(256) void k_rice_write(EncodeArgs a, const uint64_t *__restrict__ pos,
                                                    const long long *__restrict__ bit_off,
                                                    unsigned long long *__restrict__ words) {
    const uint64_t b = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = lane_id();
    if (b >= a.nb) return;
    if (a.stop && (*a.stop & kErrStopEncode)) return;
    const uint32_t m = (uint32_t)(a.acc[b + 1] - a.acc[b]);
    if (m < 2 || m >= kMaxBucket) return;
    const uint64_t base = a.node_base[b];
    const uint32_t nn = a.tree_nodes[m];
    const uint64_t p0 = pos[b] + (uint64_t)bit_off[b];
    const uint64_t fixed_total = a.memo[m] & 0xFFFF;
    uint64_t fpre = 0, upre = 0;
    for (uint32_t j0 = 0; j0 < nn; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool valid = j < nn;
        uint32_t g = 0;
        uint64_t x = 0;
        if (valid) {
            g = a.memo[a.node_m[base + j]] >> 27;
            x = a.seeds[base + j];
        }
        uint64_t ftot, utot;
        const uint64_t fo = wave_excl_scan<uint64_t>(g, ftot);
        const uint64_t ul = valid ? (x >> g) + 1 : 0;
        const uint64_t uo = wave_excl_scan<uint64_t>(ul, utot);
        if (valid) {
            or_bits(words, p0 + fpre + fo, x & ((1ull << g) - 1), g);
            const uint64_t up = p0 + fixed_total + upre + uo + (x >> g);
            atomicOr(&words[up >> 6], 1ull << (up & 63));
        }
        fpre += ftot;
        upre += utot;
    }
}

// Per-set layout of the encoded structures (filled by the host between phases).
struct SetEnc {
    uint64_t bbase, nb;          // global bucket range [bbase, bbase + nb)
    uint64_t rice_word;          // first word of the set's Golomb-Rice vector
    uint64_t lower_word, uc_word, up_word, jump_word;  // EF arrays
    uint64_t jump_words;         // words of the set's jump table
    uint64_t lc, lp;             // EF lower-bit widths
};

// One workgroup per set: bit offset of every bucket of the set inside the global
// Golomb-Rice buffer, and the sentinel bit after the set's last bucket
// (appendFixed(1, 1), SIMDRecSplit.hpp:810).
__global__ __launch_bounds__(256) void k_set_offsets(const SetEnc *__restrict__ sets, const uint64_t *__restrict__ pos,
                                                     long long *__restrict__ bit_off,
                                                     unsigned long long *__restrict__ words,
                                                     const uint32_t *__restrict__ stop = nullptr) {
    if (stop && (*stop & kErrStopEncode)) return;
    const SetEnc e = sets[blockIdx.x];
    const long long off = (long long)(e.rice_word * 64) - (long long)pos[e.bbase];
    for (uint64_t i = threadIdx.x; i < e.nb; i += blockDim.x) bit_off[e.bbase + i] = off;
    if (threadIdx.x == 0) {
        const uint64_t p = pos[e.bbase + e.nb] - pos[e.bbase] + e.rice_word * 64;
        atomicOr(&words[p >> 6], 1ull << (p & 63));
    }
}

struct EfScalars {
    uint64_t bpk;
    long long min_delta;
    long long min_diff;
    uint64_t total_bits;
};

// One workgroup per set: bits_per_key_fixed_point (IEEE double, DoubleEF.hpp:144)
// and the two minima (DoubleEF.hpp:146-179) over the set's buckets, with key
// counts and bit positions relative to the set.
__global__ __launch_bounds__(256) void k_ef_scalars(const SetEnc *__restrict__ sets, const uint64_t *__restrict__ acc,
                                                    const uint64_t *__restrict__ pos, EfScalars *__restrict__ out) {
    __shared__ uint64_t s_bpk;
    __shared__ long long s_a[256], s_b[256];
    const SetEnc e = sets[blockIdx.x];
    const uint64_t *cum = acc + e.bbase, *ps = pos + e.bbase;
    const uint64_t c0 = cum[0], p0 = ps[0], nb = e.nb;
    if (threadIdx.x == 0) s_bpk = (uint64_t)((double)(1ull << 20) * ((double)(ps[nb] - p0) / (double)(cum[nb] - c0)));
    __syncthreads();
    const uint64_t bpk = s_bpk;
    long long md = 0x3FFFFFFFFFFFFFFFll, mf = 0x3FFFFFFFFFFFFFFFll;
    for (uint64_t i = 1 + threadIdx.x; i <= nb; i += blockDim.x) {
        const uint64_t ci = cum[i] - c0, cp = cum[i - 1] - c0;
        const long long d = (long long)(ci - cp);
        md = d < md ? d : md;
        const long long bb = (long long)(ps[i] - p0) - (long long)(bpk * ci >> 20);
        const long long bp = (long long)(ps[i - 1] - p0) - (long long)(bpk * cp >> 20);
        mf = (bb - bp) < mf ? (bb - bp) : mf;
    }
    s_a[threadIdx.x] = md;
    s_b[threadIdx.x] = mf;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_a[threadIdx.x] = s_a[threadIdx.x + o] < s_a[threadIdx.x] ? s_a[threadIdx.x + o] : s_a[threadIdx.x];
            s_b[threadIdx.x] = s_b[threadIdx.x + o] < s_b[threadIdx.x] ? s_b[threadIdx.x + o] : s_b[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[blockIdx.x].bpk = bpk;
        out[blockIdx.x].min_delta = s_a[0];
        out[blockIdx.x].min_diff = s_b[0];
        out[blockIdx.x].total_bits = ps[nb] - p0;
    }
}

// One workgroup per set, DoubleEF.hpp:204-349: lower bits of both sequences
// interleaved per bucket, one upper bit per bucket in each upper-bits array, and
// the jump table (every q-th element: 64-bit head per super block, 16-bit
// offsets), computed directly from each sampled element's upper-bit position.
__global__ __launch_bounds__(256) void k_ef_fill(const SetEnc *__restrict__ sets, const EfScalars *__restrict__ sc,
                                                 const uint64_t *__restrict__ acc, const uint64_t *__restrict__ pos,
                                                 unsigned long long *__restrict__ lower, unsigned long long *__restrict__ upper_c,
                                                 unsigned long long *__restrict__ upper_p, uint8_t *__restrict__ jump,
                                                 uint32_t *__restrict__ err, const uint32_t *__restrict__ stop = nullptr) {
    if (stop && (*stop & kErrStopEncode)) return;
    const SetEnc e = sets[blockIdx.x];
    const EfScalars f = sc[blockIdx.x];
    const uint64_t *cum = acc + e.bbase, *ps = pos + e.bbase;
    const uint64_t c0 = cum[0], p0 = ps[0];
    unsigned long long *lw = lower + e.lower_word, *ucw = upper_c + e.uc_word, *upw = upper_p + e.up_word;
    uint8_t *jb = jump + e.jump_word * 8;
    const uint64_t jump_bytes = e.jump_words * 8;
    const uint64_t lc = e.lc, lp = e.lp;
    auto upper_c_pos = [&](uint64_t i) { return (((cum[i] - c0) - i * (uint64_t)f.min_delta) >> lc) + i; };
    auto upper_p_pos = [&](uint64_t i) {
        const uint64_t ci = cum[i] - c0;
        const uint64_t pv = (uint64_t)((long long)(ps[i] - p0) - (long long)(f.bpk * ci >> 20)) - i * (uint64_t)f.min_diff;
        return (pv >> lp) + i;
    };
    for (uint64_t i = threadIdx.x; i <= e.nb; i += blockDim.x) {
        const uint64_t ci = cum[i] - c0;
        const uint64_t cv = ci - i * (uint64_t)f.min_delta;
        const uint64_t pval = (uint64_t)((long long)(ps[i] - p0) - (long long)(f.bpk * ci >> 20)) - i * (uint64_t)f.min_diff;
        const uint64_t start = i * (lc + lp);
        if (lc) or_bits(lw, start, cv & ((1ull << lc) - 1), (uint32_t)lc);
        if (lp) or_bits(lw, start + lc, pval & ((1ull << lp) - 1), (uint32_t)lp);
        const uint64_t uc = (cv >> lc) + i, up = (pval >> lp) + i;
        atomicOr(&ucw[uc >> 6], 1ull << (uc & 63));
        atomicOr(&upw[up >> 6], 1ull << (up & 63));
        if ((i & kEfQMaskDev) == 0) {
            const uint64_t idx = (i >> 14) * (kEfSuperSizeDev * 2);
            const uint64_t head = i & ~(uint64_t)((1 << 14) - 1);
            const uint64_t head_c = upper_c_pos(head), head_p = upper_p_pos(head);
            if (i == head && (idx + 2) * 8 <= jump_bytes) {
                *(uint64_t *)(jb + idx * 8) = head_c;
                *(uint64_t *)(jb + idx * 8 + 8) = head_p;
            }
            const uint64_t oc = uc - head_c, op = up - head_p;
            if (oc >= 65536 || op >= 65536) atomicOr(err, 4u);
            const uint64_t slot = (idx + 2) * 8 + 4 * ((i & ((1 << 14) - 1)) >> 8);
            if (slot + 4 <= jump_bytes) {
                *(uint16_t *)(jb + slot) = (uint16_t)oc;
                *(uint16_t *)(jb + slot + 2) = (uint16_t)op;
            }
        }
    }
}

// ------------------------------------------------------------------ string keys

// One thread per string key: the reference's first_hash (SpookyHash V2 128-bit,
// seeds 0; AbstractParallelRecSplit.hpp:250, include/util/Hash128.h:32-36).  Key i
// is bytes [offsets[i], offsets[i+1]) of `bytes`; out receives (first, second).
__global__ __launch_bounds__(256) void k_spooky_keys(const uint8_t *__restrict__ bytes,
                                                     const uint64_t *__restrict__ offsets, uint64_t n,
                                                     uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = offsets[i];
    uint64_t first, second;
    spooky::first_hash(bytes + b, (size_t)(offsets[i + 1] - b), first, second);
    out[2 * i] = first;
    out[2 * i + 1] = second;
}

// ------------------------------------------------------------------ queries

struct QueryArgs {
    const uint64_t *rice;
    const uint64_t *lower, *upper_c, *upper_p, *jump;
    const uint32_t *memo;
    uint64_t nbuckets, bpk, lc, lp;
    long long min_delta, min_diff;
    // readable words of each array; reads beyond see all-ones words, so a
    // malformed deserialized structure cannot drive a scan out of its buffer
    uint64_t n_rice, n_lower, n_uc, n_up, n_jump;
    uint64_t keys_count;
    SplitParams P;
    int rotate;  // flags: kFlagRotate, kFlagRemap64
};

__device__ __forceinline__ uint64_t rd_word(const uint64_t *base, uint64_t n, uint64_t i) {
    return i < n ? base[i] : ~0ull;
}

__device__ __forceinline__ uint64_t ld_bytes64(const uint64_t *base, uint64_t n, uint64_t byte) {
    const uint64_t w = byte >> 3, sh = (byte & 7) * 8;
    const uint64_t lo = rd_word(base, n, w);
    return sh ? (lo >> sh) | (rd_word(base, n, w + 1) << (64 - sh)) : lo;
}

__device__ __forceinline__ uint64_t select64(uint64_t x, uint64_t k) {
    for (uint64_t i = 0; i < k; i++) x &= x - 1;
    return (uint64_t)__ffsll((long long)x) - 1;
}

// Batched operator() (AbstractParallelRecSplit.hpp:409-466): one thread per key.
__global__ __launch_bounds__(256) void k_query(QueryArgs q, const uint64_t *__restrict__ keys, uint64_t n,
                                               uint64_t *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint64_t first = keys[2 * t], second = keys[2 * t + 1];
    const uint64_t i = __umul64hi(first, q.nbuckets);
    // DoubleEF::get (DoubleEF.hpp:367-402)
    const uint64_t lc = q.lc, lp = q.lp;
    const uint64_t mask_c = (1ull << lc) - 1, mask_p = (1ull << lp) - 1;
    const uint64_t pos_lower = i * (lc + lp);
    uint64_t lower = ld_bytes64(q.lower, q.n_lower, pos_lower / 8) >> (pos_lower % 8);
    const uint64_t jsq = (i >> 14) * kEfSuperSizeDev * 2;
    const uint64_t jin = (i & ((1 << 14) - 1)) >> 8;
    const uint64_t pair = ld_bytes64(q.jump, q.n_jump, (jsq + 2) * 8 + 4 * jin);
    const uint64_t jump_c = rd_word(q.jump, q.n_jump, jsq) + (pair & 0xFFFF);
    const uint64_t jump_p = rd_word(q.jump, q.n_jump, jsq + 1) + ((pair >> 16) & 0xFFFF);
    uint64_t wc = jump_c / 64, wp = jump_p / 64;
    uint64_t win_c = rd_word(q.upper_c, q.n_uc, wc) & (~0ull << (jump_c % 64));
    uint64_t win_p = rd_word(q.upper_p, q.n_up, wp) & (~0ull << (jump_p % 64));
    uint64_t dc = i & 255, dp = i & 255;
    for (uint64_t c; (c = (uint64_t)__popcll(win_c)) <= dc; dc -= c) win_c = rd_word(q.upper_c, q.n_uc, ++wc);
    for (uint64_t c; (c = (uint64_t)__popcll(win_p)) <= dp; dp -= c) win_p = rd_word(q.upper_p, q.n_up, ++wp);
    const uint64_t sel_c = select64(win_c, dc);
    const uint64_t cum_delta = i * (uint64_t)q.min_delta;
    uint64_t cum_keys = ((wc * 64 + sel_c - i) << lc | (lower & mask_c)) + cum_delta;
    lower >>= lc;
    const uint64_t bit_pos = ((wp * 64 + select64(win_p, dp) - i) << lp | (lower & mask_p)) + i * (uint64_t)q.min_diff +
                             (uint64_t)(long long)(q.bpk * cum_keys >> 20);
    win_c &= (~0ull << sel_c) << 1;
    while (win_c == 0) win_c = rd_word(q.upper_c, q.n_uc, ++wc);
    lower >>= lp;
    const uint64_t cum_next = ((wc * 64 + (uint64_t)__ffsll((long long)win_c) - 1 - i - 1) << lc | (lower & mask_c)) +
                              cum_delta + (uint64_t)q.min_delta;
    uint64_t m = cum_next - cum_keys;
    if (m >= kMaxBucket) {  // malformed structure: no valid answer
        out[t] = q.keys_count;
        return;
    }
    // Golomb-Rice reader (RiceBitVector.hpp:155-218)
    uint64_t fixed_off = bit_pos;
    uint64_t up = bit_pos + (q.memo[m] & 0xFFFF);
    uint64_t uw = up / 64;
    uint64_t window = rd_word(q.rice, q.n_rice, uw++) >> (up & 63);
    int valid = 64 - (int)(up & 63);
    auto read_next = [&](uint32_t g) -> uint64_t {
        uint64_t result = 0;
        if (window == 0) {
            result += (uint64_t)valid;
            window = rd_word(q.rice, q.n_rice, uw++);
            valid = 64;
            while (window == 0) {
                result += 64;
                window = rd_word(q.rice, q.n_rice, uw++);
            }
        }
        const uint64_t p = (uint64_t)__ffsll((long long)window) - 1;
        window >>= p;
        window >>= 1;
        valid -= (int)p + 1;
        result = (result + p) << g;
        result |= (ld_bytes64(q.rice, q.n_rice, fixed_off / 8) >> (fixed_off % 8)) & ((1ull << g) - 1);
        fixed_off += g;
        return result;
    };
    auto skip = [&](uint64_t nodes, uint64_t flen) {
        uint64_t missing = nodes, c;
        while ((c = (uint64_t)__popcll(window)) < missing) {
            window = rd_word(q.rice, q.n_rice, uw++);
            missing -= c;
            valid = 64;
        }
        c = select64(window, missing - 1);
        window >>= c;
        window >>= 1;
        valid -= (int)c + 1;
        fixed_off += flen;
    };
    const SplitParams P = q.P;
    const bool r64 = (q.rotate & kFlagRemap64) != 0;
    int level = 0;
    while (m > P.upper) {
        const uint64_t d = read_next(q.memo[m] >> 27);
        const uint32_t hmod = remap_any(second + d + c_start_seed[level], (uint32_t)m, r64);
        const uint32_t split = higher_split(P, (uint32_t)m);
        if (hmod < split) {
            m = split;
        } else {
            skip((q.memo[split] >> 16) & 0x7FF, q.memo[split] & 0xFFFF);
            m -= split;
            cum_keys += split;
        }
        level++;
    }
    if (m > P.lower) {
        const uint64_t d = read_next(q.memo[m] >> 27);
        const uint32_t hmod = remap_any(second + d + c_start_seed[kSeedUpper], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / P.lower;
        m = min((uint64_t)P.lower, m - part * P.lower);
        cum_keys += P.lower * part;
        if (part) skip(((q.memo[P.lower] >> 16) & 0x7FF) * part, (q.memo[P.lower] & 0xFFFF) * part);
    }
    if (m > P.leaf) {
        const uint64_t d = read_next(q.memo[m] >> 27);
        const uint32_t hmod = remap_any(second + d + c_start_seed[kSeedLower], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / P.leaf;
        m = min((uint64_t)P.leaf, m - part * P.leaf);
        cum_keys += (uint64_t)P.leaf * part;
        if (part) skip(part, (q.memo[P.leaf] & 0xFFFF) * part);
    }
    const uint64_t b = read_next(q.memo[m] >> 27);
    const uint64_t x = second + b + c_start_seed[kSeedLeaf];
    uint64_t v;
    if ((q.rotate & kFlagRotate) && m == P.leaf) {
        const uint64_t rot = b % P.leaf;
        v = remap_any(x - rot, P.leaf, r64);
        if (second & 1) v = (v + rot) % P.leaf;
    } else {
        v = remap_any(x, (uint32_t)m, r64);
    }
    out[t] = cum_keys + v;
}

}  // namespace rsa
