/*
 * recsplit_oracle.c -- CPU ORACLE for RecSplit MPHF construction.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpurecsplit_amd/, include/)
 * may link, load or call this file.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline ("port" of SimdRecSplit's algorithm).
 *
 * It restates, in plain C, the construction semantics of the reference's
 * SIMDRecSplit (the path whose MPHF bytes the GPU build must reproduce), the
 * shared base class, the Golomb-Rice builder/reader and the double Elias-Fano
 * structure.  Every function cites the reference lines it follows
 * (paths relative to the reference repository root).
 *
 * Pinning (what checks each part; the tests live in tests/test_oracle.py):
 *   - golomb/skip memo formula ........ per-leaf-size FNV-64 checksums of the
 *                                       reference's committed table
 *                                       (include/function/golombRiceMemo.hpp),
 *                                       tests/golden/golomb_memo.json
 *   - remix32 (MurmurHash3 fmix32) .... published SMHasher vectors
 *   - remix (Stafford mix13) / keygen . published SplitMix64 outputs
 *   - node searches ................... brute-force minimality on tiny nodes
 *                                       (independent Python predicate)
 *   - packed aggregation counters ..... exhaustive/adversarial equivalence with
 *                                       exact part counts
 *   - rotation / remap / seed offsets . MPHF bijection through the restated
 *                                       query operator() (a different
 *                                       reference function than the builder)
 *   - Elias-Fano ...................... get(i) round trip on every bucket
 *   - thread-count invariance ......... byte equality across thread counts
 *   - 64-bit remap (RecSplitRotate) ... Stafford finalizer pinned by the
 *                                       SplitMix64 outputs (same function);
 *                                       the 48-bit scaling against 128-bit
 *                                       arithmetic; searches by brute-force
 *                                       minimality; bijection via the query
 *   - serialized framing of sux::util::Vector (u64 length + raw words) and
 *     its resize() semantics: PARITY UNPINNED (sux is absent from the
 *     reference checkout; no reference test or fixture pins it).
 *
 * The search loops evaluate 16 consecutive seeds per key in an inner loop that
 * gcc vectorises, the same lane layout SIMDRecSplit uses (lanes = consecutive
 * seeds, SIMDRecSplit.hpp:259-264), so the oracle doubles as the CPU baseline.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_BUCKET_SIZE 3000 /* AbstractParallelRecSplit.hpp:60 */
#define ORC_MAX_LEAF_SIZE 24     /* AbstractParallelRecSplit.hpp:62 */
#define ORC_MAX_FANOUT 9         /* AbstractParallelRecSplit.hpp:63 */
#define LANES 16

/* `rotate` arguments and fields are a flags word: bit 0 rotation fitting for full
 * leaves (USE_BIJECTIONS_ROTATE), bit 1 the 64-bit remap of the RecSplitRotate
 * variant (USE_64_BIT_REMAP). */
#define ORC_FLAG_ROTATE 1
#define ORC_FLAG_REMAP64 2

/* AbstractParallelRecSplit.hpp:95-98 */
static const uint64_t start_seed[20] = {
    0x106393c187cae21aULL, 0x6453cec3f7376937ULL, 0x643e521ddbd2be98ULL, 0x3740c6412f6572cbULL,
    0x717d47562f1ce470ULL, 0x4cd6eb4c63befb7cULL, 0x9bfd8c5e18c8da73ULL, 0x082f20e10092a9a3ULL,
    0x2ada2ce68d21defcULL, 0xe33cb4f3e7c6466bULL, 0x3980be458c509c59ULL, 0xc466fd9584828e8cULL,
    0x45f0aabe1a61ede6ULL, 0xf6e7b8b33ad9b98dULL, 0x4ef95e25f4b4983dULL, 0x81175195173b92d3ULL,
    0x4e50927d8dd15978ULL, 0x1ea2099d1fafae7fULL, 0x425c8a06fbaaa815ULL, 0xcd4216006c74052aULL};
#define NUM_START_SEEDS 20

/* AbstractParallelRecSplit.hpp:106 */
static const uint8_t bij_memo[25] = {0, 0, 0, 1, 3, 4, 5, 7, 8, 10, 11, 12, 14,
                                     15, 16, 18, 19, 21, 22, 23, 25, 26, 28, 29, 30};

/* ------------------------------------------------------------------------ */
/* hashing primitives                                                        */
/* ------------------------------------------------------------------------ */

/* AbstractParallelRecSplit.hpp:265-269 (Stafford variant 13 finalizer) */
uint64_t orc_remix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* AbstractParallelRecSplit.hpp:277-284 (MurmurHash3 fmix32) */
uint32_t orc_remix32(uint32_t z) {
    z ^= z >> 16;
    z *= 0x85ebca6bU;
    z ^= z >> 13;
    z *= 0xc2b2ae35U;
    z ^= z >> 16;
    return z;
}

/* remixAndRemap with USE_64_BIT_REMAP=false: AbstractParallelRecSplit.hpp:327-337
 * (same as the SIMD remap, SIMDRecSplit.hpp:109-122). */
static inline uint32_t remap32(uint64_t x, uint32_t n) {
    uint32_t h = orc_remix32((uint32_t)(x >> 32) ^ (uint32_t)x) & 0xFFFFu;
    return (h * n) >> 16;
}
uint32_t orc_remap(uint64_t x, uint32_t n) { return remap32(x, n); }

/* remixAndRemap with USE_64_BIT_REMAP=true (the RecSplitRotate variant,
 * RecSplitRotate.hpp:115 -> AbstractParallelRecSplit.hpp:327-332): the 64-bit
 * Stafford finalizer, its low 48 bits scaled to [0, n). */
static inline uint32_t remap64s(uint64_t x, uint32_t n) {
    return (uint32_t)(((orc_remix(x) & ((UINT64_C(1) << 48) - 1)) * n) >> 48);
}
uint32_t orc_remap64(uint64_t x, uint32_t n) { return remap64s(x, n); }
static inline uint32_t remap_any(uint64_t x, uint32_t n, int r64) { return r64 ? remap64s(x, n) : remap32(x, n); }

/* common.hpp:273-280 */
static inline uint64_t remap128(uint64_t x, uint64_t n) {
    return (uint64_t)(((unsigned __int128)x * (unsigned __int128)n) >> 64);
}
uint64_t orc_remap128(uint64_t x, uint64_t n) { return remap128(x, n); }

/* SplitMix64 key generator for the synthetic workloads: state += golden gamma,
 * output remix(state).  Each key takes first = next(), then second = next(). */
void orc_splitmix64_keys(uint64_t seed, size_t n, uint64_t *out /* 2n */) {
    uint64_t s = seed;
    for (size_t i = 0; i < 2 * n; i++) {
        s += 0x9e3779b97f4a7c15ULL;
        out[i] = orc_remix(s);
    }
}

/* ------------------------------------------------------------------------ */
/* splitting strategy and Golomb-Rice memo                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int leaf;
    uint32_t lower_aggr, upper_aggr;
    uint32_t memo[ORC_MAX_BUCKET_SIZE];
} orc_params;

static size_t smax(size_t a, size_t b) { return a > b ? a : b; }

/* AbstractParallelRecSplit.hpp:133-135 */
static void aggr_bounds(int leaf, uint32_t *lower, uint32_t *upper) {
    size_t lo = (size_t)leaf * smax(2, (size_t)ceil(0.35 * leaf + 1. / 2));
    size_t up = lo * (leaf < 7 ? 2 : (size_t)ceil(0.21 * leaf + 9. / 10));
    *lower = (uint32_t)lo;
    *upper = (uint32_t)up;
}

/* AbstractParallelRecSplit.hpp:137-148 */
static void split_params(const orc_params *P, size_t m, size_t *fanout, size_t *unit) {
    if (m > P->upper_aggr) {
        *unit = P->upper_aggr * ((uint16_t)(m / 2 + P->upper_aggr - 1) / P->upper_aggr);
        *fanout = 2;
    } else if (m > P->lower_aggr) {
        *unit = P->lower_aggr;
        *fanout = (uint16_t)(m + P->lower_aggr - 1) / P->lower_aggr;
    } else {
        *unit = P->leaf;
        *fanout = (uint16_t)(m + P->leaf - 1) / P->leaf;
    }
}

/* AbstractParallelRecSplit.hpp:191-227 */
static void fill_golomb_rice(orc_params *P) {
    uint32_t *memo = P->memo;
    memset(memo, 0, sizeof(P->memo));
    size_t s = 0;
    for (; s <= (size_t)P->leaf; ++s)
        memo[s] = (uint32_t)bij_memo[s] << 27 | (uint32_t)(s > 1) << 16 | bij_memo[s];
    for (; s < ORC_MAX_BUCKET_SIZE; ++s) {
        size_t fanout = 0, unit = 0;
        split_params(P, s, &fanout, &unit);
        int k[ORC_MAX_FANOUT] = {0};
        k[fanout - 1] = (int)s;
        for (size_t i = 0; i < fanout - 1; ++i) {
            k[i] = (int)unit;
            k[fanout - 1] -= k[i];
        }
        double sqrt_prod = 1;
        for (size_t i = 0; i < fanout; ++i) sqrt_prod *= sqrt((double)k[i]);
        const double p = sqrt((double)s) / (pow(2 * M_PI, (fanout - 1.) / 2) * sqrt_prod);
        uint32_t grl = (uint32_t)ceil(log2(-log((sqrt(5) + 1) / 2) / log1p(-p)));
        memo[s] = grl << 27;
        for (size_t i = 0; i < fanout; ++i) grl += memo[k[i]] & 0xFFFF;
        memo[s] |= grl;
        uint32_t nodes = 1;
        for (size_t i = 0; i < fanout; ++i) nodes += (memo[k[i]] >> 16) & 0x7FF;
        memo[s] |= nodes << 16;
    }
}

int orc_params_init(int leaf, orc_params *P) {
    if (leaf < 2 || leaf > ORC_MAX_LEAF_SIZE) return -1;
    P->leaf = leaf;
    aggr_bounds(leaf, &P->lower_aggr, &P->upper_aggr);
    fill_golomb_rice(P);
    return 0;
}

/* Exposed for tests: the memo table (golomb_param << 27 | nodes << 16 | bits). */
int orc_golomb_memo(int leaf, uint32_t *out /* 3000 */) {
    orc_params P;
    if (orc_params_init(leaf, &P)) return -1;
    memcpy(out, P.memo, sizeof(P.memo));
    return 0;
}
int orc_aggr_bounds(int leaf, uint32_t *lower, uint32_t *upper) {
    if (leaf < 2 || leaf > ORC_MAX_LEAF_SIZE) return -1;
    aggr_bounds(leaf, lower, upper);
    return 0;
}

#define golomb_param(P, m) ((P)->memo[m] >> 27)
#define skip_bits(P, m) ((P)->memo[m] & 0xFFFF)
#define skip_nodes(P, m) (((P)->memo[m] >> 16) & 0x7FF)

/* ------------------------------------------------------------------------ */
/* node searches: lowest seed offset satisfying each node predicate          */
/* ------------------------------------------------------------------------ */

/* Split size of a fanout-2 node: SIMDRecSplit.hpp:632 (== query :424). */
static inline uint32_t higher_split(const orc_params *P, uint32_t m) {
    return ((uint16_t)(m / 2 + P->upper_aggr - 1) / P->upper_aggr) * P->upper_aggr;
}

/* Lane vectors (GCC vector extensions; element-wise C semantics): 16 candidate
 * seeds per step as two groups of 8, hashed in 32-bit lanes. */
typedef uint64_t u64x8 __attribute__((vector_size(64)));
typedef uint32_t u32x8 __attribute__((vector_size(32)));
typedef int32_t i32x8 __attribute__((vector_size(32)));
typedef int64_t i64x8 __attribute__((vector_size(64)));

static const u64x8 LANE_OFF0 = {0, 1, 2, 3, 4, 5, 6, 7};

/* remix32 of the folded seed-key sums, low 16 bits, for 8 lanes */
static inline u32x8 h16x8(u64x8 z) {
    u32x8 t = __builtin_convertvector(z, u32x8) ^ __builtin_convertvector(z >> 32, u32x8);
    t ^= t >> 16;
    t *= 0x85ebca6bU;
    t ^= t >> 13;
    t *= 0xc2b2ae35U;
    t ^= t >> 16;
    return t & 0xFFFFu;
}
static inline u32x8 remap16_x8(u64x8 z, uint32_t n) { return (h16x8(z) * n) >> 16; }
static inline u32x8 remap64_x8(u64x8 z, uint32_t n) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return __builtin_convertvector(((z & ((UINT64_C(1) << 48) - 1)) * n) >> 48, u32x8);
}
/* the search loops below are specialised for r64 = 0 (16-bit remap) and r64 = 1 */
#define remap_x8(z, n) (r64 ? remap64_x8((z), (n)) : remap16_x8((z), (n)))
static inline u64x8 bcast64(uint64_t v) { return (u64x8){v, v, v, v, v, v, v, v}; }

/* higherLevel, SIMDRecSplit.hpp:250-317: lowest x >= 0 with
 * #{ remap(key + start_seed[level] + x, m) < split } == split. */
static inline uint64_t search_higher(const uint64_t *keys, uint32_t m, uint32_t split, int level,
                                     uint64_t *evals, const int r64) {
    const uint64_t base = start_seed[level];
    const u64x8 off1 = LANE_OFF0 + 8;
    for (uint64_t x = 0;; x += LANES) {
        i32x8 c0 = {0}, c1 = {0};
        for (uint32_t i = 0; i < m; i++) {
            const u64x8 k = bcast64(keys[i] + base + x);
            c0 -= (i32x8)(remap_x8(k + LANE_OFF0, m) < split);
            c1 -= (i32x8)(remap_x8(k + off1, m) < split);
        }
        for (int j = 0; j < LANES; j++)
            if ((uint32_t)(j < 8 ? c0[j] : c1[j - 8]) == split) {
                if (evals) *evals += (x + j + 1) * (uint64_t)m;
                return x + j;
            }
    }
}

/* aggrLevel, SIMDRecSplit.hpp:319-423.  Parts are counted the way the reference
 * does: 8-bit fields packed into 32-bit words, parts 0-3 in the low word, parts
 * 4-7 in the high word, the last part of a fanout-5 (or fanout-9) node left
 * uncounted (SIMDRecSplit.hpp:144-148, 342-388).  With SPLIT < 255 this is the
 * exact-part-count predicate (tested against plain counting). */
static inline void aggr_targets(uint32_t m, uint32_t split, uint32_t fanout, uint32_t *lo,
                                uint32_t *hi) {
    uint32_t t[ORC_MAX_FANOUT + 1] = {0};
    for (uint32_t p = 0; p + 1 < fanout; p++) t[p] = split;
    t[fanout - 1] = m - (fanout - 1) * split;
    uint32_t L = 0, H = 0;
    for (uint32_t p = 0; p < 4 && p < fanout; p++)
        if (!(fanout == 5 && p == 4)) L |= t[p] << (8 * p);
    for (uint32_t p = 4; p < 8 && p < fanout; p++) H |= t[p] << (8 * (p - 4));
    if (fanout <= 5) H = 0;
    *lo = L;
    *hi = H;
}

static inline uint32_t part_inc_lo(uint32_t part) { return part < 4 ? 1u << (8 * part) : 0u; }
static inline uint32_t part_inc_hi(uint32_t part, uint32_t fanout) {
    return (fanout > 5 && part >= 4 && part < 8) ? 1u << (8 * (part - 4)) : 0u;
}

/* Per-lane search form of the same predicate: one 64-bit word of 8-bit fields
 * for parts 0-7, part 8 (fanout 9) uncounted.  The carry argument that makes the
 * reference's two-word form exact applies unchanged (fields hold SPLIT <= 254,
 * the uncounted part's target is <= 254): a carry out of any field lowers some
 * other count below zero.  part = floor(r / SPLIT) uses M = ceil(2^20 / SPLIT):
 * (r * M) >> 20 is exact for r < 3000 whenever SPLIT < 349. */
static inline uint64_t aggr_target64(uint32_t m, uint32_t split, uint32_t fanout) {
    uint64_t T = 0;
    for (uint32_t p = 0; p < fanout && p < 8; p++)
        T |= (uint64_t)(p + 1 < fanout ? split : m - (fanout - 1) * split) << (8 * p);
    return T;
}
static inline uint32_t split_magic(uint32_t split) { return ((1u << 20) + split - 1) / split; }

/* The reference's lookup (SIMDRecSplit.hpp:144-148, 354, 378-379): increments of
 * the low word (parts 0-3) and of the high word (parts 4-7) by part index; parts
 * 8..15 add nothing (the uncounted last part of a fanout-9 node). */
static const u32x8 INC_LO_A = {1u, 1u << 8, 1u << 16, 1u << 24, 0, 0, 0, 0}, INC_LO_B = {0};
static const u32x8 INC_HI_A = {0, 0, 0, 0, 1u, 1u << 8, 1u << 16, 1u << 24}, INC_HI_B = {0};

static inline uint64_t search_aggr(const uint64_t *keys, uint32_t m, uint32_t split, uint64_t seed,
                                   uint64_t *evals, const int r64) {
    const uint32_t fanout = (uint16_t)(m + split - 1) / split;
    uint32_t tlo, thi;
    aggr_targets(m, split, fanout, &tlo, &thi);
    const uint32_t magic = split_magic(split);
    const u64x8 off1 = LANE_OFF0 + 8;
    /* fanout <= 5: parts 0-3 in one word, part 4 uncounted (its lookup is 0) */
    const int two_words = fanout > 5;
    for (uint64_t x = 0;; x += LANES) {
        u32x8 lo0 = {0}, lo1 = {0}, hi0 = {0}, hi1 = {0};
        for (uint32_t i = 0; i < m; i++) {
            const u64x8 k = bcast64(keys[i] + seed + x);
            const u32x8 p0 = (remap_x8(k + LANE_OFF0, m) * magic) >> 20;
            const u32x8 p1 = (remap_x8(k + off1, m) * magic) >> 20;
            lo0 += __builtin_shuffle(INC_LO_A, INC_LO_B, p0);
            lo1 += __builtin_shuffle(INC_LO_A, INC_LO_B, p1);
            if (two_words) {
                hi0 += __builtin_shuffle(INC_HI_A, INC_HI_B, p0);
                hi1 += __builtin_shuffle(INC_HI_A, INC_HI_B, p1);
            }
        }
        for (int j = 0; j < LANES; j++) {
            const uint32_t lo = j < 8 ? lo0[j] : lo1[j - 8], hi = j < 8 ? hi0[j] : hi1[j - 8];
            if (lo == tlo && hi == thi) {
                if (evals) *evals += (x + j + 1) * (uint64_t)m;
                return x + j;
            }
        }
    }
}

int orc_search_form_aggr_predicate(const uint32_t *parts, uint32_t m, uint32_t split) {
    const uint32_t fanout = (uint16_t)(m + split - 1) / split;
    uint64_t c = 0;
    for (uint32_t i = 0; i < m; i++) c += parts[i] < 8 ? UINT64_C(1) << (8 * parts[i]) : 0;
    return c == aggr_target64(m, split, fanout);
}

/* Exposed for the packed-counter exactness test: evaluates the reference's own
 * two-word packed predicate on an explicit list of part indexes. */
int orc_packed_aggr_predicate(const uint32_t *parts, uint32_t m, uint32_t split) {
    const uint32_t fanout = (uint16_t)(m + split - 1) / split;
    uint32_t tlo, thi, clo = 0, chi = 0;
    aggr_targets(m, split, fanout, &tlo, &thi);
    for (uint32_t i = 0; i < m; i++) {
        clo += part_inc_lo(parts[i]);
        chi += part_inc_hi(parts[i], fanout);
    }
    return clo == tlo && chi == thi;
}

/* leafLevel rotation branch, SIMDRecSplit.hpp:452-509 (rotateRight :437-439):
 * keys with even `second` form the left set, odd the right set.  Bases step by
 * L from start_seed[19]; at a base, the lowest rotation r in [0, L) with
 * mask_left | rotl(mask_right, r) == 2^L-1 wins; the stored seed is base + r. */
static inline uint64_t search_leaf_rot(const uint64_t *keys, uint32_t L, uint64_t *evals, const int r64) {
    const uint64_t SEED = start_seed[NUM_START_SEEDS - 1];
    const uint32_t full = (uint32_t)((1ULL << L) - 1);
    uint64_t left[ORC_MAX_LEAF_SIZE], right[ORC_MAX_LEAF_SIZE];
    int nl = 0, nr = 0;
    for (uint32_t i = 0; i < L; i++) {
        if (keys[i] % 2 == 0) left[nl++] = keys[i];
        else right[nr++] = keys[i];
    }
    const u64x8 step0 = LANE_OFF0 * L, step1 = (LANE_OFF0 + 8) * L;
    for (uint64_t b = 0;; b += (uint64_t)LANES * L) {
        u32x8 ml0 = {0}, ml1 = {0}, mr0 = {0}, mr1 = {0};
        for (int i = 0; i < nl; i++) {
            const u64x8 k = bcast64(left[i] + SEED + b);
            ml0 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + step0, L);
            ml1 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + step1, L);
        }
        for (int i = 0; i < nr; i++) {
            const u64x8 k = bcast64(right[i] + SEED + b);
            mr0 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + step0, L);
            mr1 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + step1, L);
        }
        for (int j = 0; j < LANES; j++) {
            const uint32_t ml = j < 8 ? ml0[j] : ml1[j - 8], mr = j < 8 ? mr0[j] : mr1[j - 8];
            /* lossless prefilter: a cover needs both masks collision-free */
            if (__builtin_popcount(ml) != nl || __builtin_popcount(mr) != nr) continue;
            for (uint32_t r = 0; r < L; r++) {
                const uint32_t rot = r == 0 ? mr : (((mr << r) | (mr >> (L - r))) & full);
                if ((ml | rot) == full) {
                    const uint64_t x = b + (uint64_t)j * L + r;
                    if (evals) *evals += (b / L + j + 1) * (uint64_t)L;
                    return x;
                }
            }
        }
    }
}

/* Midstop for the brute-force leaf prefilter, SIMDRecSplit.hpp:126-132 with the
 * AVX-512 factor (:65).  Lossless: any prefix of distinct images is necessary. */
static uint32_t bij_midstop(uint32_t s) {
    const int v = (int)ceil(2.6 * sqrt((double)s));
    return (int)s < v ? s : (uint32_t)v;
}

/* leafLevel brute-force branch, SIMDRecSplit.hpp:510-599: lowest x with
 * OR_i 1 << remap(key_i + start_seed[19] + x, m) == 2^m - 1.  The midstop +
 * backlog of the reference is a lossless filter; this restatement applies the
 * midstop per lane and then finishes the survivors. */
static inline uint64_t search_leaf_brute(const uint64_t *keys, uint32_t m, uint64_t *evals, const int r64) {
    const uint64_t SEED = start_seed[NUM_START_SEEDS - 1];
    const uint32_t found = (uint32_t)((1ULL << m) - 1);
    const uint32_t midstop = m >= 12 ? bij_midstop(m) : m;
    const u64x8 off1 = LANE_OFF0 + 8;
    for (uint64_t x = 0;; x += LANES) {
        u32x8 mk0 = {0}, mk1 = {0};
        for (uint32_t i = 0; i < midstop; i++) {
            const u64x8 k = bcast64(keys[i] + SEED + x);
            mk0 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + LANE_OFF0, m);
            mk1 |= (u32x8){1, 1, 1, 1, 1, 1, 1, 1} << remap_x8(k + off1, m);
        }
        for (int j = 0; j < LANES; j++) {
            uint32_t mk = j < 8 ? mk0[j] : mk1[j - 8];
            if ((uint32_t)__builtin_popcount(mk) != midstop) continue;
            for (uint32_t i = midstop; i < m; i++) mk |= 1u << remap_any(keys[i] + SEED + x + j, m, r64);
            if (mk == found) {
                if (evals) *evals += (x + j + 1) * (uint64_t)m;
                return x + j;
            }
        }
    }
}

/* Exported searches: flags bit 1 selects the 64-bit remap (r64); each pair of
 * calls below is specialised at compile time. */
uint64_t orc_search_higher(const uint64_t *keys, uint32_t m, uint32_t split, int level, uint64_t *evals) {
    return search_higher(keys, m, split, level, evals, 0);
}
uint64_t orc_search_aggr(const uint64_t *keys, uint32_t m, uint32_t split, uint64_t seed, uint64_t *evals) {
    return search_aggr(keys, m, split, seed, evals, 0);
}
uint64_t orc_search_leaf_rot(const uint64_t *keys, uint32_t L, uint64_t *evals) {
    return search_leaf_rot(keys, L, evals, 0);
}
uint64_t orc_search_leaf_brute(const uint64_t *keys, uint32_t m, uint64_t *evals) {
    return search_leaf_brute(keys, m, evals, 0);
}
uint64_t orc_search_higher_r64(const uint64_t *keys, uint32_t m, uint32_t split, int level, uint64_t *evals) {
    return search_higher(keys, m, split, level, evals, 1);
}
uint64_t orc_search_aggr_r64(const uint64_t *keys, uint32_t m, uint32_t split, uint64_t seed, uint64_t *evals) {
    return search_aggr(keys, m, split, seed, evals, 1);
}
uint64_t orc_search_leaf_rot_r64(const uint64_t *keys, uint32_t L, uint64_t *evals) {
    return search_leaf_rot(keys, L, evals, 1);
}
uint64_t orc_search_leaf_brute_r64(const uint64_t *keys, uint32_t m, uint64_t *evals) {
    return search_leaf_brute(keys, m, evals, 1);
}

/* ------------------------------------------------------------------------ */
/* Golomb-Rice bit vector builder, RiceBitVector.hpp:58-134                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t *data;
    size_t size, cap; /* words; `size` follows sux::util::Vector::resize */
    uint64_t bit_count;
} rbuilder;

static void rb_init(rbuilder *b) {
    b->cap = 16;
    b->size = 16; /* Builder() : Builder(16), RiceBitVector.hpp:63-65 */
    b->data = (uint64_t *)calloc(b->cap, 8);
    b->bit_count = 0;
}

/* sux::util::Vector::resize: grow (zero-filled) when beyond capacity, then set
 * the size (also when smaller).  PARITY UNPINNED (sux is not in the checkout). */
static void rb_resize(rbuilder *b, size_t words) {
    if (words > b->cap) {
        size_t nc = b->cap + b->cap / 2;
        if (nc < words) nc = words;
        b->data = (uint64_t *)realloc(b->data, nc * 8);
        memset(b->data + b->cap, 0, (nc - b->cap) * 8);
        b->cap = nc;
    }
    b->size = words;
}

/* RiceBitVector.hpp:73-90 */
static void rb_append_fixed(rbuilder *b, uint64_t v, int log2golomb) {
    const uint64_t lower = v & ((UINT64_C(1) << log2golomb) - 1);
    int used = (int)(b->bit_count & 63);
    rb_resize(b, (((b->bit_count + log2golomb + 7) / 8) + 7 + 7) / 8);
    uint64_t *p = b->data + b->bit_count / 64;
    uint64_t cur = *p;
    cur |= lower << used;
    if (used + log2golomb > 64) {
        *(p++) = cur;
        cur = lower >> (64 - used);
    }
    *p = cur;
    b->bit_count += log2golomb;
}

/* RiceBitVector.hpp:92-106 */
static void rb_append_unary_all(rbuilder *b, const uint64_t *unary, size_t n) {
    uint64_t inc = 0;
    for (size_t i = 0; i < n; i++) inc += unary[i] + 1;
    rb_resize(b, (((b->bit_count + inc + 7) / 8) + 7 + 7) / 8);
    for (size_t i = 0; i < n; i++) {
        b->bit_count += unary[i];
        b->data[b->bit_count / 64] |= UINT64_C(1) << (b->bit_count & 63);
        ++b->bit_count;
    }
}

/* ------------------------------------------------------------------------ */
/* recursive construction of one bucket, SIMDRecSplit.hpp:624-693            */
/* ------------------------------------------------------------------------ */

typedef struct {
    const orc_params *P;
    int rotate;
    uint64_t *temp;    /* MAX_BUCKET_SIZE scratch */
    uint64_t *fixed_v; /* seeds in DFS preorder */
    uint32_t *fixed_m; /* node sizes in DFS preorder */
    size_t nseeds;
    uint64_t evals;
} bucket_ctx;

static void push_seed(bucket_ctx *c, uint64_t x, uint32_t m) {
    c->fixed_v[c->nseeds] = x;
    c->fixed_m[c->nseeds] = m;
    c->nseeds++;
}

static void leaf_level(bucket_ctx *c, uint64_t *bucket, uint32_t m) {
    uint64_t x;
    const int r64 = (c->rotate & ORC_FLAG_REMAP64) != 0;
    if ((c->rotate & ORC_FLAG_ROTATE) && m == (uint32_t)c->P->leaf)
        x = r64 ? orc_search_leaf_rot_r64(bucket, m, &c->evals) : orc_search_leaf_rot(bucket, m, &c->evals);
    else x = r64 ? orc_search_leaf_brute_r64(bucket, m, &c->evals) : orc_search_leaf_brute(bucket, m, &c->evals);
    push_seed(c, x, m);
}

/* Stable partition of a node's keys by child index (SIMDRecSplit.hpp:288-306,
 * 395-412). */
static void partition(bucket_ctx *c, uint64_t *bucket, uint32_t m, uint64_t seed_abs,
                      uint32_t split_or_unit, int is_higher) {
    uint64_t *temp = c->temp;
    uint32_t fan = is_higher ? 2 : (uint16_t)(m + split_or_unit - 1) / split_or_unit;
    uint32_t pos[ORC_MAX_FANOUT + 1];
    for (uint32_t p = 0; p < fan; p++) pos[p] = is_higher ? (p ? split_or_unit : 0) : p * split_or_unit;
    for (uint32_t i = 0; i < m; i++) {
        const uint32_t r = remap_any(bucket[i] + seed_abs, m, (c->rotate & ORC_FLAG_REMAP64) != 0);
        const uint32_t p = is_higher ? (r >= split_or_unit) : r / split_or_unit;
        temp[pos[p]++] = bucket[i];
    }
    memcpy(bucket, temp, (size_t)m * 8);
}

static void rec_split(bucket_ctx *c, uint64_t *bucket, uint32_t m, int level) {
    const orc_params *P = c->P;
    const int r64 = (c->rotate & ORC_FLAG_REMAP64) != 0;
    if (m <= (uint32_t)P->leaf) {
        leaf_level(c, bucket, m);
        return;
    }
    if (m > P->upper_aggr) { /* fanout 2: SIMDRecSplit.hpp:631-639 */
        const uint32_t split = higher_split(P, m);
        const uint64_t x = r64 ? orc_search_higher_r64(bucket, m, split, level, &c->evals)
                               : orc_search_higher(bucket, m, split, level, &c->evals);
        push_seed(c, x, m);
        partition(c, bucket, m, start_seed[level] + x, split, 1);
        rec_split(c, bucket, split, level + 1);
        if (m - split > 1) rec_split(c, bucket + split, m - split, level + 1);
    } else if (m > P->lower_aggr) { /* SIMDRecSplit.hpp:640-650 */
        const uint64_t S = start_seed[NUM_START_SEEDS - 3];
        const uint64_t x = r64 ? orc_search_aggr_r64(bucket, m, P->lower_aggr, S, &c->evals)
                               : orc_search_aggr(bucket, m, P->lower_aggr, S, &c->evals);
        push_seed(c, x, m);
        partition(c, bucket, m, S + x, P->lower_aggr, 0);
        uint32_t i;
        for (i = 0; i < m - P->lower_aggr; i += P->lower_aggr)
            rec_split(c, bucket + i, P->lower_aggr, level + 1);
        if (m - i > 1) rec_split(c, bucket + i, m - i, level + 1);
    } else { /* SIMDRecSplit.hpp:651-661 */
        const uint64_t S = start_seed[NUM_START_SEEDS - 2];
        const uint32_t L = (uint32_t)P->leaf;
        const uint64_t x = r64 ? orc_search_aggr_r64(bucket, m, L, S, &c->evals) : orc_search_aggr(bucket, m, L, S, &c->evals);
        push_seed(c, x, m);
        partition(c, bucket, m, S + x, L, 0);
        uint32_t i;
        for (i = 0; i < m - L; i += L) leaf_level(c, bucket + i, L);
        if (m - i > 1) leaf_level(c, bucket + i, m - i);
    }
}

/* ------------------------------------------------------------------------ */
/* double Elias-Fano, DoubleEF.hpp:60-437                                    */
/* ------------------------------------------------------------------------ */

#define EF_LOG2Q 8
#define EF_Q (1u << EF_LOG2Q)
#define EF_QMASK (EF_Q - 1)
#define EF_SUPER_Q (1u << 14)
#define EF_SUPER_QMASK (EF_SUPER_Q - 1)
#define EF_Q_PER_SUPER (EF_SUPER_Q / EF_Q)
#define EF_SUPER_SIZE (1 + EF_Q_PER_SUPER / 4)

typedef struct {
    uint64_t num_buckets, u_cum_keys, u_position;
    uint64_t l_position, l_cum_keys;
    int64_t cum_keys_min_delta, min_diff;
    uint64_t bits_per_key_fixed_point;
    uint64_t *lower_bits, *upper_cum, *upper_pos, *jump;
    size_t n_lower, n_upper_cum, n_upper_pos, n_jump;
    uint64_t mask_c, mask_p;
} orc_ef;

static int lambda64(uint64_t w) { return 63 ^ __builtin_clzll(w); }

/* common.hpp:417-437 (any correct select works) */
static uint64_t select64(uint64_t x, uint64_t k) {
    for (uint64_t i = 0; i < k; i++) x &= x - 1;
    return (uint64_t)__builtin_ctzll(x);
}

static void ef_set(uint64_t *bits, uint64_t pos) { bits[pos / 64] |= UINT64_C(1) << (pos % 64); }

/* DoubleEF.hpp:80-86 */
static void ef_set_bits(uint64_t *bits, uint64_t start, int width, uint64_t value) {
    const uint64_t mask = ((UINT64_C(1) << width) - 1) << (start % 8);
    uint64_t t;
    memcpy(&t, (uint8_t *)bits + start / 8, 8);
    t = (t & ~mask) | value << (start % 8);
    memcpy((uint8_t *)bits + start / 8, &t, 8);
}

/* DoubleEF.hpp:94-98 */
static size_t ef_jump_words(uint64_t nb) {
    size_t size = (nb / EF_SUPER_Q) * EF_SUPER_SIZE * 2;
    if (nb % EF_SUPER_Q != 0) size += (1 + ((nb % EF_SUPER_Q + EF_Q - 1) / EF_Q + 3) / 4) * 2;
    return size;
}

static void ef_derive(orc_ef *e) {
    e->l_position = e->u_position / (e->num_buckets + 1) == 0 ? 0 : lambda64(e->u_position / (e->num_buckets + 1));
    e->l_cum_keys = e->u_cum_keys / (e->num_buckets + 1) == 0 ? 0 : lambda64(e->u_cum_keys / (e->num_buckets + 1));
    e->mask_c = (UINT64_C(1) << e->l_cum_keys) - 1;
    e->mask_p = (UINT64_C(1) << e->l_position) - 1;
}

/* Writes one u16 jump entry if it lies inside the allocated table.  Entries past
 * the end (possible when the last super block is exactly full) fall outside
 * the serialized words in the reference too. */
static int ef_put16(orc_ef *e, uint64_t word_idx, uint64_t half_idx, uint64_t val) {
    if (val >= (1u << 16)) return -1; /* DoubleEF.hpp:300,317,328,342 */
    const uint64_t byte = word_idx * 8 + half_idx * 2;
    if (byte + 2 <= e->n_jump * 8) memcpy((uint8_t *)e->jump + byte, &(uint16_t){(uint16_t)val}, 2);
    return 0;
}
static void ef_put64(orc_ef *e, uint64_t word_idx, uint64_t val) {
    if (word_idx < e->n_jump) e->jump[word_idx] = val;
}

/* DoubleEF.hpp:140-365 (scalar construction; the jump table follows the
 * FAST_CONSTRUCTION loop :286-320, equivalent to the bit-scan loop :322-348). */
static int ef_build(orc_ef *e, const uint64_t *cum_keys, const uint64_t *position, uint64_t nb) {
    memset(e, 0, sizeof(*e));
    e->num_buckets = nb;
    e->bits_per_key_fixed_point = (uint64_t)((double)(UINT64_C(1) << 20) * (position[nb] / (double)cum_keys[nb]));
    const uint64_t bpk = e->bits_per_key_fixed_point;
    e->min_diff = INT64_MAX / 2;
    e->cum_keys_min_delta = INT64_MAX / 2;
    int64_t prev_bucket_bits = 0;
    for (uint64_t i = 1; i <= nb; ++i) {
        const int64_t nkeys_delta = (int64_t)(cum_keys[i] - cum_keys[i - 1]);
        if (nkeys_delta < e->cum_keys_min_delta) e->cum_keys_min_delta = nkeys_delta;
        const int64_t bucket_bits = (int64_t)position[i] - (int64_t)(bpk * cum_keys[i] >> 20);
        if (bucket_bits - prev_bucket_bits < e->min_diff) e->min_diff = bucket_bits - prev_bucket_bits;
        prev_bucket_bits = bucket_bits;
    }
    e->u_position = (uint64_t)((int64_t)position[nb] - (int64_t)(bpk * cum_keys[nb] >> 20) -
                               (int64_t)(nb * (uint64_t)e->min_diff) + 1);
    e->u_cum_keys = cum_keys[nb] - nb * (uint64_t)e->cum_keys_min_delta + 1;
    ef_derive(e);
    const uint64_t lc = e->l_cum_keys, lp = e->l_position;
    e->n_lower = ((nb + 1) * (lc + lp) + 63) / 64 + 1;
    e->n_upper_cum = (nb + 1 + (e->u_cum_keys >> lc) + 63) / 64;
    e->n_upper_pos = (nb + 1 + (e->u_position >> lp) + 63) / 64;
    e->lower_bits = (uint64_t *)calloc(e->n_lower + 1, 8);
    e->upper_cum = (uint64_t *)calloc(e->n_upper_cum + 1, 8);
    e->upper_pos = (uint64_t *)calloc(e->n_upper_pos + 1, 8);
    uint64_t cum_delta = 0, bit_delta = 0;
    for (uint64_t i = 0; i <= nb; i++, cum_delta += (uint64_t)e->cum_keys_min_delta, bit_delta += (uint64_t)e->min_diff) {
        if (lc != 0) ef_set_bits(e->lower_bits, i * (lc + lp), (int)lc, (cum_keys[i] - cum_delta) & e->mask_c);
        ef_set(e->upper_cum, ((cum_keys[i] - cum_delta) >> lc) + i);
        const int64_t pval = (int64_t)position[i] - (int64_t)(bpk * cum_keys[i] >> 20);
        if (lp != 0) ef_set_bits(e->lower_bits, i * (lc + lp) + lc, (int)lp, ((uint64_t)pval - bit_delta) & e->mask_p);
        ef_set(e->upper_pos, (((uint64_t)pval - bit_delta) >> lp) + i);
    }
    e->n_jump = ef_jump_words(nb);
    e->jump = (uint64_t *)calloc(e->n_jump + 1, 8);
    for (int which = 0; which < 2; which++) {
        const uint64_t *ub = which ? e->upper_pos : e->upper_cum;
        const size_t nw = which ? e->n_upper_pos : e->n_upper_cum;
        for (uint64_t i = 0, c = 0, q_counter = EF_QMASK, last_super_q = 0; i < nw; i++) {
            const uint64_t bits = ub[i];
            const uint64_t pc = (uint64_t)__builtin_popcountll(bits);
            q_counter += pc;
            c += pc;
            if (q_counter >= EF_Q) {
                q_counter -= EF_Q;
                const uint64_t bit_num = i * 64 + select64(bits, pc - q_counter - 1);
                const uint64_t idx = (c / EF_SUPER_Q) * (EF_SUPER_SIZE * 2);
                const uint64_t exact_c = c - q_counter - 1;
                if ((exact_c & EF_SUPER_QMASK) == 0) {
                    ef_put64(e, idx + which, bit_num);
                    last_super_q = bit_num;
                }
                if (ef_put16(e, idx + 2, 2 * ((exact_c & EF_SUPER_QMASK) / EF_Q) + which, bit_num - last_super_q))
                    return -1;
            }
        }
    }
    return 0;
}

/* DoubleEF.hpp:367-402 */
static void ef_get(const orc_ef *e, uint64_t i, uint64_t *cum_keys, uint64_t *cum_keys_next, uint64_t *position) {
    const uint64_t lc = e->l_cum_keys, lp = e->l_position;
    const uint64_t pos_lower = i * (lc + lp);
    uint64_t lower;
    memcpy(&lower, (const uint8_t *)e->lower_bits + pos_lower / 8, 8);
    lower >>= pos_lower % 8;
    const uint64_t jsq = (i / EF_SUPER_Q) * EF_SUPER_SIZE * 2;
    const uint64_t jin = (i % EF_SUPER_Q) / EF_Q;
    uint16_t o_c, o_p;
    memcpy(&o_c, (const uint8_t *)(e->jump + jsq + 2) + 2 * (2 * jin), 2);
    memcpy(&o_p, (const uint8_t *)(e->jump + jsq + 2) + 2 * (2 * jin + 1), 2);
    const uint64_t jump_c = e->jump[jsq] + o_c;
    const uint64_t jump_p = e->jump[jsq + 1] + o_p;
    uint64_t wc = jump_c / 64, wp = jump_p / 64;
    uint64_t win_c = e->upper_cum[wc] & (UINT64_C(-1) << (jump_c % 64));
    uint64_t win_p = e->upper_pos[wp] & (UINT64_C(-1) << (jump_p % 64));
    uint64_t dc = i & EF_QMASK, dp = i & EF_QMASK;
    for (uint64_t bc; (bc = (uint64_t)__builtin_popcountll(win_c)) <= dc; dc -= bc) win_c = e->upper_cum[++wc];
    for (uint64_t bc; (bc = (uint64_t)__builtin_popcountll(win_p)) <= dp; dp -= bc) win_p = e->upper_pos[++wp];
    const uint64_t sel_c = select64(win_c, dc);
    const int64_t cum_delta = (int64_t)(i * (uint64_t)e->cum_keys_min_delta);
    *cum_keys = ((wc * 64 + sel_c - i) << lc | (lower & e->mask_c)) + (uint64_t)cum_delta;
    lower >>= lc;
    const int64_t bit_delta = (int64_t)(i * (uint64_t)e->min_diff);
    *position = ((wp * 64 + select64(win_p, dp) - i) << lp | (lower & e->mask_p)) + (uint64_t)bit_delta +
                (uint64_t)(int64_t)(e->bits_per_key_fixed_point * *cum_keys >> 20);
    win_c &= (UINT64_C(-1) << sel_c) << 1;
    while (win_c == 0) win_c = e->upper_cum[++wc];
    lower >>= lp;
    *cum_keys_next = ((wc * 64 + (uint64_t)__builtin_ctzll(win_c) - i - 1) << lc | (lower & e->mask_c)) +
                     (uint64_t)cum_delta + (uint64_t)e->cum_keys_min_delta;
}

/* ------------------------------------------------------------------------ */
/* full construction (SIMDRecSplit.hpp:695-812)                              */
/* ------------------------------------------------------------------------ */

typedef struct orc_mphf {
    orc_params P;
    int rotate;
    uint64_t bucket_size, keys_count, nbuckets;
    uint64_t *rice; /* Golomb-Rice words */
    size_t rice_words;
    uint64_t rice_bits;
    orc_ef ef;
    uint64_t evals; /* algorithmic key-seed evaluations */
    uint64_t nseeds;
} orc_mphf;

typedef struct {
    const orc_params *P;
    int rotate;
    const uint64_t *bucket_acc;
    uint64_t *sorted;
    uint64_t b0, b1;
    /* outputs: seeds + sizes per bucket concatenated in bucket order */
    uint64_t *seeds;
    uint32_t *sizes;
    uint64_t *seed_off; /* per bucket (relative to thread start), b1-b0+1 entries */
    uint64_t evals;
} thread_job;

static void *thread_main(void *arg) {
    thread_job *J = (thread_job *)arg;
    bucket_ctx c;
    memset(&c, 0, sizeof(c));
    c.P = J->P;
    c.rotate = J->rotate;
    c.temp = (uint64_t *)malloc(ORC_MAX_BUCKET_SIZE * 8);
    uint64_t total = 0;
    J->seed_off[0] = 0;
    for (uint64_t b = J->b0; b < J->b1; b++) {
        const uint32_t s = (uint32_t)(J->bucket_acc[b + 1] - J->bucket_acc[b]);
        c.fixed_v = J->seeds + total;
        c.fixed_m = J->sizes + total;
        c.nseeds = 0;
        if (s > 1) rec_split(&c, J->sorted + J->bucket_acc[b], s, 0);
        total += c.nseeds;
        J->seed_off[b - J->b0 + 1] = total;
    }
    J->evals = c.evals;
    free(c.temp);
    return NULL;
}

typedef struct { uint64_t bucket, second; } bk_pair;
static int cmp_pair(const void *a, const void *b) {
    const bk_pair *x = (const bk_pair *)a, *y = (const bk_pair *)b;
    if (x->bucket != y->bucket) return x->bucket < y->bucket ? -1 : 1;
    return x->second < y->second ? -1 : x->second > y->second;
}

/* Error codes: -1 bad args, -2 bucket >= 3000 keys, -3 equal `second` inside a
 * bucket (the reference never terminates on such input), -4 EF offset overflow. */
orc_mphf *orc_build(const uint64_t *keys, size_t n, int leaf, size_t bucket_size, int rotate,
                    int threads, int *err) {
    *err = 0;
    if (n == 0 || bucket_size == 0 || leaf < 2 || leaf > ORC_MAX_LEAF_SIZE) {
        *err = -1;
        return NULL;
    }
    if (threads < 1) threads = 1;
    orc_mphf *R = (orc_mphf *)calloc(1, sizeof(orc_mphf));
    orc_params_init(leaf, &R->P);
    R->rotate = rotate;
    R->bucket_size = bucket_size;
    R->keys_count = n;
    /* SIMDRecSplit.hpp:774 */
    const uint64_t nb = (n + bucket_size - 1) / bucket_size > 1 ? (n + bucket_size - 1) / bucket_size : 1;
    R->nbuckets = nb;
    /* parallelPartition, AbstractParallelRecSplit.hpp:339-359: only each bucket's
     * multiset of `second` matters (all node predicates are order-free). */
    bk_pair *pairs = (bk_pair *)malloc(n * sizeof(bk_pair));
    for (size_t i = 0; i < n; i++) {
        pairs[i].bucket = remap128(keys[2 * i], nb);
        pairs[i].second = keys[2 * i + 1];
    }
    qsort(pairs, n, sizeof(bk_pair), cmp_pair);
    uint64_t *bucket_acc = (uint64_t *)calloc(nb + 1, 8);
    uint64_t *sorted = (uint64_t *)malloc(n * 8);
    for (size_t i = 0; i < n; i++) {
        sorted[i] = pairs[i].second;
        bucket_acc[pairs[i].bucket + 1]++;
        if (i && pairs[i].bucket == pairs[i - 1].bucket && pairs[i].second == pairs[i - 1].second) *err = -3;
    }
    free(pairs);
    for (uint64_t b = 0; b < nb; b++) {
        if (bucket_acc[b + 1] >= ORC_MAX_BUCKET_SIZE) *err = -2;
        bucket_acc[b + 1] += bucket_acc[b];
    }
    if (*err) {
        free(bucket_acc);
        free(sorted);
        free(R);
        return NULL;
    }
    /* compute_thread, SIMDRecSplit.hpp:695-745: contiguous bucket ranges per thread */
    thread_job *jobs = (thread_job *)calloc((size_t)threads, sizeof(thread_job));
    pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        thread_job *J = &jobs[t];
        J->P = &R->P;
        J->rotate = rotate;
        J->bucket_acc = bucket_acc;
        J->sorted = sorted;
        J->b0 = (uint64_t)t * nb / (uint64_t)threads;
        J->b1 = (t == threads - 1) ? nb : (uint64_t)(t + 1) * nb / (uint64_t)threads;
        const uint64_t keys_in = bucket_acc[J->b1] - bucket_acc[J->b0];
        J->seeds = (uint64_t *)malloc((keys_in + 1) * 8);
        J->sizes = (uint32_t *)malloc((keys_in + 1) * 4);
        J->seed_off = (uint64_t *)malloc((J->b1 - J->b0 + 1) * 8);
        if (threads == 1) thread_main(J);
        else pthread_create(&tids[t], NULL, thread_main, J);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(tids[t], NULL);

    /* Encoding: per bucket, fixed parts in DFS preorder then all unary parts
     * (SIMDRecSplit.hpp:707-714).  Concatenating per-thread builders in thread
     * order (:726-744) equals appending sequentially. */
    rbuilder B;
    rb_init(&B);
    uint64_t *bucket_pos = (uint64_t *)calloc(nb + 1, 8);
    uint64_t *unary = (uint64_t *)malloc(ORC_MAX_BUCKET_SIZE * 8);
    for (int t = 0; t < threads; t++) {
        thread_job *J = &jobs[t];
        R->evals += J->evals;
        for (uint64_t b = J->b0; b < J->b1; b++) {
            const uint64_t s0 = J->seed_off[b - J->b0], s1 = J->seed_off[b - J->b0 + 1];
            for (uint64_t k = s0; k < s1; k++) {
                const int g = (int)golomb_param(&R->P, J->sizes[k]);
                rb_append_fixed(&B, J->seeds[k], g);
                unary[k - s0] = J->seeds[k] >> g;
            }
            if (s1 > s0) rb_append_unary_all(&B, unary, s1 - s0);
            R->nseeds += s1 - s0;
            bucket_pos[b + 1] = B.bit_count;
        }
        free(J->seeds);
        free(J->sizes);
        free(J->seed_off);
    }
    free(unary);
    free(jobs);
    free(tids);
    rb_append_fixed(&B, 1, 1); /* sentinel, SIMDRecSplit.hpp:810 */
    R->rice = B.data;          /* build(): trimToFit keeps `size` words */
    R->rice_words = B.size;
    R->rice_bits = B.bit_count;
    if (ef_build(&R->ef, bucket_acc, bucket_pos, nb)) *err = -4;
    free(bucket_acc);
    free(bucket_pos);
    free(sorted);
    if (*err) {
        free(R->rice);
        free(R);
        return NULL;
    }
    return R;
}

/* ------------------------------------------------------------------------ */
/* serialisation (AbstractParallelRecSplit.hpp:480-488, DoubleEF.hpp:100-113) */
/* ------------------------------------------------------------------------ */

static size_t vec_bytes(size_t words) { return 8 + words * 8; }

size_t orc_serialized_size(const orc_mphf *R) {
    return 24 + vec_bytes(R->rice_words) + 48 + vec_bytes(R->ef.n_lower) + vec_bytes(R->ef.n_upper_cum) +
           vec_bytes(R->ef.n_upper_pos) + vec_bytes(R->ef.n_jump);
}

static uint8_t *put_u64(uint8_t *p, uint64_t v) {
    memcpy(p, &v, 8);
    return p + 8;
}
static uint8_t *put_vec(uint8_t *p, const uint64_t *w, size_t n) {
    p = put_u64(p, n);
    memcpy(p, w, n * 8);
    return p + n * 8;
}

void orc_serialize(const orc_mphf *R, uint8_t *out) {
    uint8_t *p = out;
    p = put_u64(p, (uint64_t)R->P.leaf);
    p = put_u64(p, R->bucket_size);
    p = put_u64(p, R->keys_count);
    p = put_vec(p, R->rice, R->rice_words);
    p = put_u64(p, R->ef.num_buckets);
    p = put_u64(p, R->ef.u_cum_keys);
    p = put_u64(p, R->ef.u_position);
    p = put_u64(p, (uint64_t)R->ef.cum_keys_min_delta);
    p = put_u64(p, (uint64_t)R->ef.min_diff);
    p = put_u64(p, R->ef.bits_per_key_fixed_point);
    p = put_vec(p, R->ef.lower_bits, R->ef.n_lower);
    p = put_vec(p, R->ef.upper_cum, R->ef.n_upper_cum);
    p = put_vec(p, R->ef.upper_pos, R->ef.n_upper_pos);
    put_vec(p, R->ef.jump, R->ef.n_jump);
}

/* operator>> (AbstractParallelRecSplit.hpp:490-504, DoubleEF.hpp:115-135).
 * Returns NULL on a truncated/malformed stream or a leaf-size mismatch. */
static const uint8_t *get_u64(const uint8_t *p, const uint8_t *end, uint64_t *v) {
    if (!p || end - p < 8) return NULL;
    memcpy(v, p, 8);
    return p + 8;
}
static const uint8_t *get_vec(const uint8_t *p, const uint8_t *end, uint64_t **w, size_t *n) {
    uint64_t len;
    p = get_u64(p, end, &len);
    if (!p || len > (uint64_t)(end - p) / 8) return NULL;
    *n = (size_t)len;
    *w = (uint64_t *)calloc(len + 2, 8); /* +2: unaligned 8-byte reads past the end */
    memcpy(*w, p, len * 8);
    return p + len * 8;
}

orc_mphf *orc_load(const uint8_t *buf, size_t len, int leaf, int rotate) {
    const uint8_t *p = buf, *end = buf + len;
    uint64_t v;
    orc_mphf *R = (orc_mphf *)calloc(1, sizeof(orc_mphf));
    p = get_u64(p, end, &v);
    if (!p || (int)v != leaf || orc_params_init(leaf, &R->P)) goto fail;
    R->rotate = rotate;
    p = get_u64(p, end, &R->bucket_size);
    p = get_u64(p, end, &R->keys_count);
    if (!p || R->bucket_size == 0) goto fail;
    R->nbuckets = (R->keys_count + R->bucket_size - 1) / R->bucket_size;
    if (R->nbuckets < 1) R->nbuckets = 1;
    p = get_vec(p, end, &R->rice, &R->rice_words);
    p = get_u64(p, end, &R->ef.num_buckets);
    p = get_u64(p, end, &R->ef.u_cum_keys);
    p = get_u64(p, end, &R->ef.u_position);
    p = get_u64(p, end, &v);
    R->ef.cum_keys_min_delta = (int64_t)v;
    p = get_u64(p, end, &v);
    R->ef.min_diff = (int64_t)v;
    p = get_u64(p, end, &R->ef.bits_per_key_fixed_point);
    p = get_vec(p, end, &R->ef.lower_bits, &R->ef.n_lower);
    p = get_vec(p, end, &R->ef.upper_cum, &R->ef.n_upper_cum);
    p = get_vec(p, end, &R->ef.upper_pos, &R->ef.n_upper_pos);
    p = get_vec(p, end, &R->ef.jump, &R->ef.n_jump);
    if (!p) goto fail;
    ef_derive(&R->ef);
    return R;
fail:
    free(R->rice);
    free(R->ef.lower_bits);
    free(R->ef.upper_cum);
    free(R->ef.upper_pos);
    free(R->ef.jump);
    free(R);
    return NULL;
}

/* ------------------------------------------------------------------------ */
/* query, AbstractParallelRecSplit.hpp:409-466 + RiceBitVector.hpp:155-218    */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint64_t *data;
    uint64_t fixed_off;
    uint64_t window;
    const uint64_t *uptr;
    int valid;
} rreader;

static void rr_reset(rreader *r, const uint64_t *data, uint64_t bit_pos, uint64_t unary_offset) {
    r->data = data;
    r->fixed_off = bit_pos;
    const uint64_t up = bit_pos + unary_offset;
    r->uptr = data + up / 64;
    r->window = *(r->uptr++) >> (up & 63);
    r->valid = 64 - (int)(up & 63);
}

static uint64_t rr_next(rreader *r, int log2golomb) {
    uint64_t result = 0;
    if (r->window == 0) {
        result += (uint64_t)r->valid;
        r->window = *(r->uptr++);
        r->valid = 64;
        while (r->window == 0) {
            result += 64;
            r->window = *(r->uptr++);
        }
    }
    const uint64_t pos = (uint64_t)__builtin_ctzll(r->window);
    r->window >>= pos;
    r->window >>= 1;
    r->valid -= (int)pos + 1;
    result += pos;
    result <<= log2golomb;
    uint64_t fixed;
    memcpy(&fixed, (const uint8_t *)r->data + r->fixed_off / 8, 8);
    result |= (fixed >> (r->fixed_off % 8)) & ((UINT64_C(1) << log2golomb) - 1);
    r->fixed_off += (uint64_t)log2golomb;
    return result;
}

static void rr_skip(rreader *r, uint64_t nodes, uint64_t fixed_len) {
    uint64_t missing = nodes, cnt;
    while ((cnt = (uint64_t)__builtin_popcountll(r->window)) < missing) {
        r->window = *(r->uptr++);
        missing -= cnt;
        r->valid = 64;
    }
    cnt = select64(r->window, missing - 1);
    r->window >>= cnt;
    r->window >>= 1;
    r->valid -= (int)cnt + 1;
    r->fixed_off += fixed_len;
}

uint64_t orc_query(const orc_mphf *R, uint64_t first, uint64_t second) {
    const orc_params *P = &R->P;
    const int r64 = (R->rotate & ORC_FLAG_REMAP64) != 0;
    const uint64_t bucket = remap128(first, R->nbuckets);
    uint64_t cum_keys, cum_keys_next, bit_pos;
    ef_get(&R->ef, bucket, &cum_keys, &cum_keys_next, &bit_pos);
    uint64_t m = cum_keys_next - cum_keys;
    rreader rd;
    rr_reset(&rd, R->rice, bit_pos, skip_bits(P, m));
    int level = 0;
    while (m > P->upper_aggr) {
        const uint64_t d = rr_next(&rd, (int)golomb_param(P, m));
        const uint32_t hmod = remap_any(second + d + start_seed[level], (uint32_t)m, r64);
        const uint32_t split = ((uint16_t)(m / 2 + P->upper_aggr - 1) / P->upper_aggr) * P->upper_aggr;
        if (hmod < split) {
            m = split;
        } else {
            rr_skip(&rd, skip_nodes(P, split), skip_bits(P, split));
            m -= split;
            cum_keys += split;
        }
        level++;
    }
    if (m > P->lower_aggr) {
        const uint64_t d = rr_next(&rd, (int)golomb_param(P, m));
        const uint32_t hmod = remap_any(second + d + start_seed[NUM_START_SEEDS - 3], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / P->lower_aggr;
        m = P->lower_aggr < m - part * P->lower_aggr ? P->lower_aggr : m - part * P->lower_aggr;
        cum_keys += P->lower_aggr * part;
        if (part) rr_skip(&rd, skip_nodes(P, P->lower_aggr) * part, skip_bits(P, P->lower_aggr) * part);
    }
    if (m > (uint64_t)P->leaf) {
        const uint64_t d = rr_next(&rd, (int)golomb_param(P, m));
        const uint32_t hmod = remap_any(second + d + start_seed[NUM_START_SEEDS - 2], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / (uint64_t)P->leaf;
        m = (uint64_t)P->leaf < m - part * P->leaf ? (uint64_t)P->leaf : m - part * P->leaf;
        cum_keys += (uint64_t)P->leaf * part;
        if (part) rr_skip(&rd, part, skip_bits(P, P->leaf) * part);
    }
    const uint64_t b = rr_next(&rd, (int)golomb_param(P, m));
    const uint64_t x = second + b + start_seed[NUM_START_SEEDS - 1];
    uint64_t leaf_val;
    if ((R->rotate & ORC_FLAG_ROTATE) && m == (uint64_t)P->leaf) {
        const uint64_t rotation = b % (uint64_t)P->leaf;
        leaf_val = remap_any(x - rotation, (uint32_t)P->leaf, r64);
        if (second % 2 != 0) leaf_val = (leaf_val + rotation) % (uint64_t)P->leaf;
    } else {
        leaf_val = remap_any(x, (uint32_t)m, r64);
    }
    return cum_keys + leaf_val;
}

void orc_query_batch(const orc_mphf *R, const uint64_t *keys, size_t n, uint64_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = orc_query(R, keys[2 * i], keys[2 * i + 1]);
}

/* Bijection check (test/recsplitCorrectness.hpp:10-35): returns 0 if the keys map
 * onto [0, n) exactly, else the number of collisions/out-of-range results. */
uint64_t orc_check_bijection(const orc_mphf *R, const uint64_t *keys, size_t n) {
    uint8_t *taken = (uint8_t *)calloc(n, 1);
    uint64_t bad = 0;
    for (size_t i = 0; i < n; i++) {
        const uint64_t v = orc_query(R, keys[2 * i], keys[2 * i + 1]);
        if (v >= n || taken[v]) bad++;
        else taken[v] = 1;
    }
    free(taken);
    return bad;
}

/* EF round trip (DoubleEF.hpp:351-364): returns the number of mismatching buckets. */
uint64_t orc_ef_roundtrip(const uint64_t *cum_keys, const uint64_t *position, uint64_t nb) {
    orc_ef e;
    if (ef_build(&e, cum_keys, position, nb)) return UINT64_MAX;
    uint64_t bad = 0;
    for (uint64_t i = 0; i < nb; i++) {
        uint64_t c, cn, p;
        ef_get(&e, i, &c, &cn, &p);
        if (c != cum_keys[i] || cn != cum_keys[i + 1] || p != position[i]) bad++;
    }
    free(e.lower_bits);
    free(e.upper_cum);
    free(e.upper_pos);
    free(e.jump);
    return bad;
}

uint64_t orc_evals(const orc_mphf *R) { return R->evals; }
uint64_t orc_nseeds(const orc_mphf *R) { return R->nseeds; }
uint64_t orc_rice_bits(const orc_mphf *R) { return R->rice_bits; }
uint64_t orc_nbuckets(const orc_mphf *R) { return R->nbuckets; }

/* Bucket sizes after partition (test helper). */
void orc_bucket_sizes(const uint64_t *keys, size_t n, size_t bucket_size, uint32_t *out, uint64_t nb) {
    memset(out, 0, nb * 4);
    for (size_t i = 0; i < n; i++) out[remap128(keys[2 * i], nb)]++;
    (void)bucket_size;
}

/* Seeds of every coded node in the whole MPHF in encoding order (DFS per bucket),
 * decoded back from the Golomb-Rice stream: used by tests to compare node seeds. */
size_t orc_decode_seeds(const orc_mphf *R, const uint64_t *bucket_sizes_acc, uint64_t *out, size_t cap) {
    (void)bucket_sizes_acc;
    size_t k = 0;
    for (uint64_t b = 0; b < R->nbuckets; b++) {
        uint64_t c0, c1, pos;
        ef_get(&R->ef, b, &c0, &c1, &pos);
        const uint64_t m = c1 - c0;
        if (m <= 1) continue;
        const uint64_t nodes = skip_nodes(&R->P, m);
        rreader rd;
        rr_reset(&rd, R->rice, pos, skip_bits(&R->P, m));
        /* Walk sizes in DFS preorder to know each node's golomb parameter. */
        uint32_t stack[4096];
        int sp = 0;
        stack[sp++] = (uint32_t)m;
        while (sp > 0 && k < cap) {
            const uint32_t s = stack[--sp];
            out[k++] = rr_next(&rd, (int)golomb_param(&R->P, s));
            if (s <= (uint32_t)R->P.leaf) continue;
            size_t fan, unit;
            split_params(&R->P, s, &fan, &unit);
            uint32_t child[ORC_MAX_FANOUT + 1];
            uint32_t rest = s;
            for (size_t i = 0; i < fan; i++) {
                child[i] = (i + 1 < fan) ? (uint32_t)unit : rest;
                rest -= child[i];
            }
            for (size_t i = fan; i-- > 0;)
                if (child[i] > 1) stack[sp++] = child[i];
        }
        (void)nodes;
    }
    return k;
}

void orc_free(orc_mphf *R) {
    if (!R) return;
    free(R->rice);
    free(R->ef.lower_bits);
    free(R->ef.upper_cum);
    free(R->ef.upper_pos);
    free(R->ef.jump);
    free(R);
}
