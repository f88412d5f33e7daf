// rs_engine.hip -- host orchestration of a RecSplit build on one gfx950 GPU and
// the build / device-query entry points of the C ABI.
//
// Pipeline of one build (all on one HIP stream; three small device->host reads):
//   1. bucket histogram, per-bucket node counts, size histogram, scans
//      -> read the size histogram: reject buckets >= 3000 keys, size the node lists
//   2. scatter `second` by bucket, sort each bucket in LDS (detects equal keys)
//      -> read the error flag
//   3. expand every bucket's split tree into per-phase node lists
//   4. per phase (fanout-2 depth 0, 1, ..., upper aggregation, lower aggregation,
//      rotation leaves, brute-force leaves): persistent seed-search kernel, then
//      partition of the keys between the children
//   5. per-bucket descriptor lengths, scan -> read the total bit count
//   6. Golomb-Rice write; Elias-Fano scalars -> read them; Elias-Fano fill
//   7. copy the structure to the host (and keep a device copy for queries)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "rs_internal.h"
#include "rs_kernels.hip"

namespace rsa {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char *what, int line)
        : std::runtime_error(std::string("HIP error ") + hipGetErrorString(c) + " at rs_engine.hip:" +
                             std::to_string(line) + ": " + what),
          code(c) {}
};
struct BuildError : std::runtime_error {
    int code;
    BuildError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(x)                                        \
    do {                                                    \
        hipError_t e_ = (x);                                \
        if (e_ != hipSuccess) throw HipError(e_, #x, __LINE__); \
    } while (0)

// Worst-case device bytes of a build of n keys in nb buckets (keys excluded):
// 8 B/key of bucket-sorted keys, <= n tree nodes of 58 B (descriptor, size, seed,
// search state), 8 B/key of Golomb-Rice words, 64 B/bucket of counts, offsets and
// scans, 25 B/bucket + 1 B/key of Elias-Fano arrays, 1 MiB of fixed-size buffers.
static uint64_t working_set_bound(uint64_t n, uint64_t nb) { return 75 * n + 89 * nb + (uint64_t(1) << 20); }

// Bump allocator over a fixed device region (the allocation-free build): every
// buffer of the build is carved from it in order; nothing is freed.
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0, used = 0;
    template <typename T>
    T *take(size_t count) {
        const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        const size_t at = (used + 255) & ~size_t(255);
        if (at + bytes > size)
            throw BuildError(RS_ERR_ARG, "the build needs more than the " + std::to_string(size >> 20) +
                                             " MiB scratch of the allocation-free entry point");
        used = at + bytes;
        return (T *)(base + at);
    }
};
// Set while the calling thread runs an allocation-free build: DevBuf requests are
// then served from the arena and the context's own buffers are left untouched.
static thread_local Arena *t_arena = nullptr;

// Grow-only device allocation reused across builds of a context.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    template <typename T>
    T *get(size_t count) {
        if (t_arena) return t_arena->take<T>(count);
        const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        if (bytes > cap) {
            if (p) HIP_CHECK(hipFree(p));
            p = nullptr;
            cap = 0;
            HIP_CHECK(hipMalloc(&p, bytes));
            cap = bytes;
        }
        return (T *)p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Grow-only pinned host memory: results come back from HBM at the DMA engines'
// rate, with no bounce buffer and no first-touch page faults on fresh memory.
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    uint64_t *get(size_t bytes) {
        if (bytes > cap) {
            release();
            HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
            cap = bytes;
        }
        return (uint64_t *)p;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Device-resident copy of a structure for batched queries; word counts are the
// readable words of each array.
struct DeviceCopy {
    int device = 0;
    uint64_t *rice = nullptr, *lower = nullptr, *upper_c = nullptr, *upper_p = nullptr, *jump = nullptr;
    uint32_t *memo = nullptr;
    uint64_t n_rice = 0, n_lower = 0, n_uc = 0, n_up = 0, n_jump = 0;
};

void device_copy_free(DeviceCopy *d) {
    if (!d) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(d->device);
    for (void *p : {(void *)d->rice, (void *)d->lower, (void *)d->upper_c, (void *)d->upper_p, (void *)d->jump,
                    (void *)d->memo})
        if (p) (void)hipFree(p);
    (void)hipSetDevice(prev);
    delete d;
}

constexpr int kScanItems = 2048;          // items per workgroup of the scan

// ------------------------------------------------------------------ scans

template <typename T>
__global__ __launch_bounds__(256) void k_scan_blocks(const T *__restrict__ in, uint64_t n, uint64_t *__restrict__ out,
                                                     uint64_t *__restrict__ block_sums) {
    // exclusive scan of in[0..n) extended by one zero (out has n+1 entries)
    __shared__ uint64_t s_warp[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanItems + threadIdx.x * 8;
    uint64_t v[8], run = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t i = base + j;
        v[j] = i < n ? (uint64_t)in[i] : 0;
        run += v[j];
    }
    // wave inclusive scan of thread totals
    uint64_t inc = run;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = bcast64_up(inc, o);
        if (lane >= (uint32_t)o) inc += t;
    }
    if (lane == 63) s_warp[w] = inc;
    __syncthreads();
    uint64_t woff = 0;
    for (uint32_t k = 0; k < w; k++) woff += s_warp[k];
    uint64_t ex = woff + inc - run;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t i = base + j;
        if (i <= n) out[i] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 255) block_sums[blockIdx.x] = woff + inc;
}

__global__ __launch_bounds__(256) void k_scan_add(uint64_t *__restrict__ out, uint64_t n,
                                                  const uint64_t *__restrict__ offs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] += offs[i / kScanItems];
}

// ------------------------------------------------------------------ census

// Node counts per phase list for one bucket-size histogram (tree shapes follow
// AbstractParallelRecSplit.hpp:137-148), with the expected key evaluations of
// each list's search.
struct Census {
    uint64_t count[kNumLists] = {};
    double evals[kNumLists] = {};
    uint64_t full[kNumLists] = {};  // aggregation nodes of the level's full size
    uint32_t max_depth = 0;
};

// Subtree below an aggregation node or leaf of a given size: nodes and expected
// evaluations per kind (the same at every depth).
struct SubCensus {
    bool done = false;
    uint64_t n[4] = {};
    double e[4] = {};
    uint64_t f[4] = {};
};

// Expected key evaluations of the search of one node of s keys.  The Golomb-Rice
// parameter g of its size satisfies 2^g ~ -ln(phi) / ln(1 - p) for the success
// probability p of one candidate (AbstractParallelRecSplit.hpp:191-227), so
// 1/p ~ 2^g / 0.48; each candidate reads the s keys, and a rotation leaf tests
// all l rotations of a candidate with one pass over its keys.
static double node_evals(const Memo &memo, uint32_t s, uint32_t kind, uint32_t leaf) {
    const double e = (double)s * std::ldexp(2.08, (int)golomb_param(memo, s));
    return kind == kLeafRot ? e / leaf : e;
}

static void census_node(const SplitParams &P, const Memo &memo, bool rotate, uint32_t s, uint32_t depth,
                        uint64_t mult, Census &c, std::vector<SubCensus> &small) {
    const uint32_t kind = classify(P, s, rotate);
    if (kind != kHigher) {
        SubCensus &sub = small[s];
        if (!sub.done) {
            SubCensus acc;
            acc.n[kind - 1] += 1;
            acc.e[kind - 1] += node_evals(memo, s, kind, P.leaf);
            acc.f[kind - 1] += (kind == kUpper && s == P.upper) || (kind == kLower && s == P.lower);
            if (s > P.leaf) {
                const uint32_t unit = kind == kUpper ? P.lower : P.leaf;
                const uint32_t fan = (uint16_t)(s + unit - 1) / unit;
                for (uint32_t i = 0; i < fan; i++) {
                    const uint32_t cs = (i + 1 == fan) ? s - (fan - 1) * unit : unit;
                    if (cs > 1) {
                        Census part;
                        census_node(P, memo, rotate, cs, 0, 1, part, small);
                        for (int k = 0; k < 4; k++) {
                            acc.n[k] += part.count[kMaxDepth + k];
                            acc.e[k] += part.evals[kMaxDepth + k];
                            acc.f[k] += part.full[kMaxDepth + k];
                        }
                    }
                }
            }
            acc.done = true;
            sub = acc;
        }
        for (int k = 0; k < 4; k++) {
            c.count[kMaxDepth + k] += sub.n[k] * mult;
            c.evals[kMaxDepth + k] += sub.e[k] * (double)mult;
            c.full[kMaxDepth + k] += sub.f[k] * mult;
        }
        return;
    }
    if (depth >= kMaxDepth) throw BuildError(RS_ERR_ARG, "split tree deeper than the start-seed table");
    c.count[depth] += mult;
    c.evals[depth] += node_evals(memo, s, kHigher, P.leaf) * (double)mult;
    c.max_depth = std::max(c.max_depth, depth);
    const uint32_t split = higher_split(P, s);
    census_node(P, memo, rotate, split, depth + 1, mult, c, small);
    if (s - split > 1) census_node(P, memo, rotate, s - split, depth + 1, mult, c, small);
}

// Workgroups of one search launch.  A wave costs more the more waves share the
// level's nodes (claims and result checks on the nodes' counters, scans of the
// running table by waves without a node), so a level of little work runs on fewer
// waves.  Fitted on MI355X (l = 8..16, 10K..1M keys): w = sqrt((E + 5000 n) /
// 1000) waves for E expected evaluations over n nodes (5000 evaluations stand for
// a node's fixed cost), at least n / 8; levels of more than ~6e10 evaluations or
// 64K nodes take the whole grid.
static uint32_t level_blocks(double evals, double nodes, uint32_t full_blocks) {
    const double waves = std::max(std::sqrt((std::max(evals, 0.0) + 5000.0 * nodes) / 1000.0), nodes / 8.0);
    const double b = std::ceil(waves / 4.0);
    return (uint32_t)std::max(std::min(b, (double)full_blocks), std::min(8.0, (double)full_blocks));
}

}  // namespace rsa

using namespace rsa;

// Launch configuration of the seed search, per node kind (index = NodeKind).
// Defaults are the measured best on MI355X; RS_TUNE overrides them, e.g.
// RS_TUNE="u=4,u0=1,lut=1,g1=1,g2=2,g3=4,bpc=8,sk=1" (u: candidates in flight per
// lane for every kind, uK: for kind K only, lut: LDS increment table for aggregation,
// sk / skK: keys in scalar registers for every kind / kind K,
// gK: 64*u-candidate groups per claimed chunk of kind K, hK: helper cap of kind K
// (0: none), bpc: workgroups of 4 waves per CU, bt: byte table for aggregation
// levels of mostly full-size nodes, ubt: candidates per lane there, btw / ubtw: the
// same for 64-bit counters, btl: byte table of positions for rotation leaves).
// Fanout-2 nodes need ~40 candidates, so they run one per lane.
struct Tune {
    // upper aggregation one per lane: its increment-table launches are the levels
    // whose nodes are mostly below the full size (whole buckets of about b keys),
    // which need few candidates (4096-set batch at l = 12: 20.2 -> 15.8 ms); the
    // full-size levels run the byte table with u_bt
    int u[kNumKinds] = {1, 1, 2, 4, 4};
    int sk[kNumKinds] = {1, 1, 1, 1, 1};  // keys in scalar registers
    int lut = 1;
    // groups per claimed chunk: a wave checks the node's result between groups, so
    // long chunks cost no overshoot on nodes a wave searches alone and save claims
    // (leaves 4 -> 16 groups: C2 leaves 68.0 -> 64.2 ms, C3 319 -> 309 ms; lower
    // aggregation 2 -> 8: 0.6% at C2 but up to 25% slower on small builds' levels,
    // whose few nodes are shared by helpers)
    int groups[kNumKinds] = {1, 1, 2, 16, 16};
    // helpers join a node only while it has fewer waves than this (0: no cap);
    // fanout-2 nodes need ~40 candidates, so extra waves mostly overshoot
    int help_cap[kNumKinds] = {2, 0, 0, 0, 0};
    int blocks_per_cu = 8;
    int bt = 1;      // byte table for the full-size nodes of aggregation levels
    int u_bt = 4;    // candidates in flight per lane in those launches
    int bt_leaf = 1; // byte table of positions for rotation leaves
    int bt_wide = 1; // byte table also for 64-bit counters (leaf sizes above 18)
    int u_bt_wide = 2; // candidates per lane there (four spill at 80 registers)
};

static Tune parse_tune(const char *env, Tune t = Tune()) {
    if (!env) return t;
    std::string s(env);
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find(',', p);
        if (q == std::string::npos) q = s.size();
        const std::string kv = s.substr(p, q - p);
        const size_t e = kv.find('=');
        if (e != std::string::npos) {
            const std::string k = kv.substr(0, e);
            const int v = std::atoi(kv.c_str() + e + 1);
            const bool uv = v == 1 || v == 2 || v == 4;
            if (k == "u" && uv)
                for (int &x : t.u) x = v;
            else if (k.size() == 2 && k[0] == 'u' && k[1] >= '0' && k[1] < '0' + (int)kNumKinds && uv)
                t.u[k[1] - '0'] = v;
            else if (k == "lut") t.lut = v != 0;
            else if (k == "bt") t.bt = v != 0;
            else if (k == "ubt" && uv) t.u_bt = v;
            else if (k == "btl") t.bt_leaf = v != 0;
            else if (k == "btw") t.bt_wide = v != 0;
            else if (k == "ubtw" && uv) t.u_bt_wide = v;
            else if (k == "sk")
                for (int &x : t.sk) x = v != 0;
            else if (k.size() == 3 && k[0] == 's' && k[1] == 'k' && k[2] >= '0' && k[2] < '0' + (int)kNumKinds)
                t.sk[k[2] - '0'] = v != 0;
            else if (k == "bpc" && v >= 1 && v <= 32) t.blocks_per_cu = v;
            else if (k.size() == 2 && k[0] == 'g' && k[1] >= '0' && k[1] < '0' + (int)kNumKinds && v >= 1 && v <= 64)
                t.groups[k[1] - '0'] = v;
            else if (k.size() == 2 && k[0] == 'h' && k[1] >= '0' && k[1] < '0' + (int)kNumKinds && v >= 0)
                t.help_cap[k[1] - '0'] = v;
        }
        p = q + 1;
    }
    return t;
}

// The same settings as an RS_TUNE string (every field).
static std::string format_tune(const Tune &t) {
    std::string s;
    auto add = [&](const std::string &k, int v) { s += (s.empty() ? "" : ",") + k + "=" + std::to_string(v); };
    for (int k = 0; k < kNumKinds; k++) add("u" + std::to_string(k), t.u[k]);
    for (int k = 0; k < kNumKinds; k++) add("g" + std::to_string(k), t.groups[k]);
    for (int k = 0; k < kNumKinds; k++) add("sk" + std::to_string(k), t.sk[k]);
    for (int k = 0; k < kNumKinds; k++) add("h" + std::to_string(k), t.help_cap[k]);
    add("lut", t.lut);
    add("bt", t.bt);
    add("ubt", t.u_bt);
    add("btl", t.bt_leaf);
    add("btw", t.bt_wide);
    add("ubtw", t.u_bt_wide);
    add("bpc", t.blocks_per_cu);
    return s;
}

// Scratch of the allocation-free build (rs_build_device_noalloc): a fixed array in
// the library's code object, one copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel: each device's address comes
// from hipGetSymbolAddress once, at the first context creation on that device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_use[kMaxDevices];  // one allocation-free build per GPU at a time

struct rs_ctx {
    Tune tune;                      // every leaf size without its own tuning
    Tune tune_leaf[kMaxLeaf + 1];   // per leaf size, from rs_autotune / rs_ctx_set_tune
    bool tuned[kMaxLeaf + 1] = {};
    bool autotune = false;          // RS_AUTOTUNE: tune each leaf size at its first build
    int device = 0;
    uint8_t *scratch = nullptr;  // this device's address of `scratch`
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevBuf keys_in, counts, counts64, cursor, acc, node_cnt, node_base, size_hist, scan_tmp, sorted, pool, list_base,
        list_fill, node_m, seeds, state, running, ctl, err, bits, pos, evals, work, rice, ef_scalars, ef_lower, ef_uc,
        ef_up, ef_jump, tiles, set_idx, set_gather, set_enc, bit_off, clock, str_bytes, str_off, tune_keys, wave_clock;
    double time_limit_s = 0;  // 0: none
    bool wave_times = false;  // RS_WAVE_TIMES: record search waves' start and exit times
    int wall_khz = 0;         // device wall-clock rate
    uint32_t *memo[kMaxLeaf + 1] = {};
    uint32_t *h_pinned = nullptr;  // small pinned staging area
    HostBuf h_result;              // pinned staging of every set's words on their way to the handles
    hipEvent_t ev[128];
    int n_ev = 0;
};

namespace rsa {

// out[0..n] = exclusive scan of in[0..n), out[n] = total.  Block sums are
// scanned recursively on the device; tmp holds every level's sums and offsets
// (scan_tmp_words(n) words).
static uint64_t scan_tmp_words(uint64_t n) {
    uint64_t words = 0;
    for (uint64_t items = n + 1;;) {
        const uint64_t nblk = (items + kScanItems - 1) / kScanItems;
        if (nblk <= 1) break;
        words += 2 * nblk + 1;
        items = nblk + 1;
    }
    return words + 1;
}

static void scan_level(hipStream_t s, const void *in, bool in32, uint64_t n, uint64_t *out, uint64_t *tmp) {
    const uint64_t items = n + 1;
    const uint64_t nblk = (items + kScanItems - 1) / kScanItems;
    uint64_t *sums = tmp, *offs = tmp + nblk;  // nblk block sums, then their scan (nblk + 1)
    if (in32)
        k_scan_blocks<uint32_t><<<nblk, 256, 0, s>>>((const uint32_t *)in, n, out, sums);
    else
        k_scan_blocks<uint64_t><<<nblk, 256, 0, s>>>((const uint64_t *)in, n, out, sums);
    HIP_CHECK(hipGetLastError());
    if (nblk > 1) {
        scan_level(s, sums, false, nblk, offs, tmp + 2 * nblk + 1);
        k_scan_add<<<(items + 255) / 256, 256, 0, s>>>(out, n, offs);
        HIP_CHECK(hipGetLastError());
    }
}

static void scan_u64(rs_ctx *c, const void *in, bool in32, uint64_t n, uint64_t *out) {
    scan_level(c->stream, in, in32, n, out, c->scan_tmp.get<uint64_t>(scan_tmp_words(n)));
}

// Device tables of a leaf size: the Golomb-Rice memo, then the unmasked node
// count of a bucket by size (kMaxBucket entries each).
static const uint32_t *ctx_memo(rs_ctx *c, uint32_t leaf) {
    if (t_arena) {
        // allocation-free build: the tables are copied into the scratch per call
        uint32_t *t = t_arena->take<uint32_t>(kMaxBucket * 2);
        HIP_CHECK(hipMemcpyAsync(t, golomb_memo(leaf).data(), sizeof(uint32_t) * kMaxBucket, hipMemcpyHostToDevice,
                                 c->stream));
        HIP_CHECK(hipMemcpyAsync(t + kMaxBucket, tree_nodes(leaf).data(), sizeof(uint32_t) * kMaxBucket,
                                 hipMemcpyHostToDevice, c->stream));
        return t;
    }
    if (!c->memo[leaf]) {
        const Memo &m = golomb_memo(leaf);
        const Memo &t = tree_nodes(leaf);
        HIP_CHECK(hipMalloc(&c->memo[leaf], sizeof(uint32_t) * kMaxBucket * 2));
        HIP_CHECK(hipMemcpy(c->memo[leaf], m.data(), sizeof(uint32_t) * kMaxBucket, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(c->memo[leaf] + kMaxBucket, t.data(), sizeof(uint32_t) * kMaxBucket, hipMemcpyHostToDevice));
    }
    return c->memo[leaf];
}

// Chooses the packed-counter layout of an aggregation level: the narrowest field
// width with 2^fb >= split + 2; 32-bit counters when (fanout-1) fields fit.
struct AggrLayout {
    uint32_t fb, magic;
    bool wide;
};
static AggrLayout aggr_layout(uint32_t split, uint32_t max_fanout) {
    uint32_t fb = 1;
    while ((1u << fb) < split + 2) fb++;
    AggrLayout L{fb, ((1u << 20) + split - 1) / split, false};
    if (fb * (max_fanout - 1) > 31) L.wide = true;
    if (fb * (max_fanout - 1) > 63) throw BuildError(RS_ERR_ARG, "aggregation counters do not fit 64 bits");
    // (r * magic) >> 20 == r / split for every remapped value r < 3000
    for (uint32_t r = 0; r < kMaxBucket; r++)
        if (((r * L.magic) >> 20) != r / split) throw BuildError(RS_ERR_ARG, "aggregation divisor not exact");
    return L;
}

static int ev_record(rs_ctx *c) {
    if (c->n_ev >= 128) return -1;
    HIP_CHECK(hipEventRecord(c->ev[c->n_ev], c->stream));
    return c->n_ev++;
}
static double ev_ms(rs_ctx *c, int a, int b) {
    if (a < 0 || b < 0) return 0;
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, c->ev[a], c->ev[b]));
    return ms;
}

template <int KIND, int U, bool SK, bool R64>
static void launch_kind(hipStream_t s, const uint64_t *keys, const SearchArgs &a, bool wide, bool lut, uint32_t blocks,
                        size_t lds) {
    if constexpr (KIND == kUpper || KIND == kLower) {
        if (a.bt_m && wide && lut && !R64)
            k_search<KIND, U, true, true, SK, false, 0, true><<<blocks, kBtThreads, lds, s>>>(keys, a);
        else if (a.bt_m && lut && !R64)
            k_search<KIND, U, false, true, SK, false, 0, true><<<blocks, kBtThreads, lds, s>>>(keys, a);
        else if (wide && lut) k_search<KIND, U, true, true, SK, R64><<<blocks, 256, lds, s>>>(keys, a);
        else if (wide) k_search<KIND, U, true, false, SK, R64><<<blocks, 256, 0, s>>>(keys, a);
        else if (lut) k_search<KIND, U, false, true, SK, R64><<<blocks, 256, lds, s>>>(keys, a);
        else k_search<KIND, U, false, false, SK, R64><<<blocks, 256, 0, s>>>(keys, a);
    } else if constexpr (KIND == kLeafRot) {
        const size_t canon_lds = ((size_t)2) << a.leaf;  // least-rotation table, leaves of <= kNecklaceMax keys
        if (a.bt_m && !R64 && a.leaf <= kNecklaceMax)
            k_search<KIND, U, false, false, SK, false, 2, true><<<blocks, kBtThreadsLeaf, canon_lds, s>>>(keys, a);
        else if (a.bt_m && !R64 && a.leaf <= 16)
            k_search<KIND, U, false, false, SK, false, 1, true><<<blocks, kBtThreadsLeaf, 0, s>>>(keys, a);
        else if (a.bt_m && !R64)
            k_search<KIND, U, false, false, SK, false, 0, true><<<blocks, kBtThreadsLeaf, 0, s>>>(keys, a);
        else if (a.leaf <= kNecklaceMax)
            k_search<KIND, U, false, false, SK, R64, 2><<<blocks, 256, canon_lds, s>>>(keys, a);
        else if (a.leaf <= 16)
            k_search<KIND, U, false, false, SK, R64, 1><<<blocks, 256, 0, s>>>(keys, a);
        else
            k_search<KIND, U, false, false, SK, R64><<<blocks, 256, 0, s>>>(keys, a);
    } else {
        k_search<KIND, U, false, false, SK, R64><<<blocks, 256, 0, s>>>(keys, a);
    }
}

template <int U, bool SK, bool R64>
static void launch_u(hipStream_t s, uint32_t kind, const uint64_t *keys, const SearchArgs &a, bool wide, bool lut,
                     uint32_t blocks, size_t lds) {
    switch (kind) {
    case kHigher: launch_kind<kHigher, U, SK, R64>(s, keys, a, wide, lut, blocks, lds); break;
    case kUpper: launch_kind<kUpper, U, SK, R64>(s, keys, a, wide, lut, blocks, lds); break;
    case kLower: launch_kind<kLower, U, SK, R64>(s, keys, a, wide, lut, blocks, lds); break;
    case kLeafRot: launch_kind<kLeafRot, U, SK, R64>(s, keys, a, wide, lut, blocks, lds); break;
    default: launch_kind<kLeafBrute, U, SK, R64>(s, keys, a, wide, lut, blocks, lds); break;
    }
}

template <bool SK, bool R64>
static void launch_sk(hipStream_t s, int u, uint32_t kind, const uint64_t *keys, const SearchArgs &a, bool wide,
                      bool lut, uint32_t blocks, size_t lds) {
    switch (u) {
    case 1: launch_u<1, SK, R64>(s, kind, keys, a, wide, lut, blocks, lds); break;
    case 4: launch_u<4, SK, R64>(s, kind, keys, a, wide, lut, blocks, lds); break;
    default: launch_u<2, SK, R64>(s, kind, keys, a, wide, lut, blocks, lds); break;
    }
}

// Grid of one search level from its size in 4-wave workgroups: the byte-table
// instance (aggregation levels with 32-bit counters whose nodes are mostly of the
// full size, and rotation leaves, both over the 16-bit remap) runs
// kBtThreads-thread workgroups, two per CU, since its 64 KiB table is per
// workgroup.
struct SearchShape {
    uint32_t grid, waves;
    bool bt;
};
static SearchShape search_shape(const Tune &tn, uint32_t kind, bool wide, bool lut, bool r64, bool mostly_full,
                                uint32_t blocks4, uint32_t full_blocks4, uint32_t num_cus) {
    const bool aggr = tn.bt && (kind == kUpper || kind == kLower) && lut && (!wide || tn.bt_wide) && mostly_full;
    const bool leaf = tn.bt_leaf && kind == kLeafRot;
    const uint32_t wpg = bt_threads((int)kind) / 64;  // waves per workgroup
    // A leaf level of fewer waves than half a byte-table workgroup per CU would
    // put its waves on a few CUs (sharing their SIMDs) while others idle; it runs
    // the 4-wave workgroups instead.  Measured rotation levels: 828 waves (10K keys,
    // l = 16) 0.84 -> 0.67 ms, 984 waves (100K keys, l = 12) 0.41 -> 0.29 ms; at
    // 2612 waves (100K keys, l = 16) the byte table is faster (2.32 against 2.53 ms).
    // Small aggregation levels keep the byte table, whose instruction savings
    // outweigh that (10K keys at l = 16: upper 1.68 ms with it, 2.54 ms without).
    const bool fills = (uint64_t)blocks4 * 4 * 2 >= (uint64_t)num_cus * wpg;
    if (!((aggr || (leaf && fills)) && !r64)) return {blocks4, blocks4 * 4, false};
    const uint32_t g = std::max<uint32_t>(1, std::min((blocks4 * 4 + wpg - 1) / wpg, full_blocks4 * 4 / wpg));
    return {g, g * wpg, true};  // never past the full grid
}

// The 64-bit remap runs with keys in scalar registers only.
static void launch_search(hipStream_t s, const Tune &tn, uint32_t kind, const uint64_t *keys, const SearchArgs &a,
                          bool wide, bool lut, uint32_t blocks, size_t lds, bool r64) {
    const int u = a.bt_m && (kind == kUpper || kind == kLower) ? (wide ? tn.u_bt_wide : tn.u_bt) : tn.u[kind];
    if (r64) launch_sk<true, true>(s, u, kind, keys, a, wide, lut, blocks, lds);
    else if (tn.sk[kind]) launch_sk<true, false>(s, u, kind, keys, a, wide, lut, blocks, lds);
    else launch_sk<false, false>(s, u, kind, keys, a, wide, lut, blocks, lds);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint64_t *__restrict__ idx, uint64_t n,
                             uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

constexpr uint64_t kTileKeys = 1 << 14;  // keys per histogram / scatter workgroup

static std::string autotune(rs_ctx *c, uint32_t leaf, uint64_t bucket, int flags, uint64_t sample_keys);

// Where a single-set build writes its serialized structure instead of returning a
// handle (the allocation-free entry point): the caller's buffer.
struct SerialOut {
    uint8_t *buf;
    size_t cap;
    size_t *len;
};

// The tuning a build of this leaf size uses: the leaf size's own (rs_autotune,
// rs_ctx_set_tune), else the context's.
static Tune tune_for(const rs_ctx *c, uint32_t leaf) { return c->tuned[leaf] ? c->tune_leaf[leaf] : c->tune; }

// Idle share of every search launch's tail from the waves' start and exit times
// (RS_WAVE_TIMES): per level, busy = sum of the waves' times in the search loop /
// (most waves resident at once * (last exit - first start)); tail = last exit -
// median exit.  Kinds with several levels (fanout-2 depths) report the
// duration-weighted busy share and the longest tail.
static void wave_time_stats(rs_ctx *c, const unsigned long long *d_clock, uint32_t n_waves, const uint64_t *count,
                            rs_build_stats &st) {
    std::vector<unsigned long long> t((size_t)kNumLists * 2 * n_waves);
    HIP_CHECK(hipMemcpyAsync(t.data(), d_clock, t.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double busy_w[kNumKinds] = {}, dur_w[kNumKinds] = {};
    const double ms_per_tick = 1.0 / c->wall_khz;
    for (int l = 0; l < kNumLists; l++) {
        if (!count[l]) continue;
        const uint32_t kind = l < (int)kMaxDepth ? kHigher : (uint32_t)(l - kMaxDepth + 1);
        std::vector<std::pair<unsigned long long, int>> ev;
        std::vector<unsigned long long> exits;
        unsigned long long t0 = ~0ull, t1 = 0;
        double inside = 0;
        for (uint32_t w = 0; w < n_waves; w++) {
            const unsigned long long a = t[((size_t)l * n_waves + w) * 2], b = t[((size_t)l * n_waves + w) * 2 + 1];
            if (!a || b < a) continue;
            t0 = std::min(t0, a);
            t1 = std::max(t1, b);
            inside += (double)(b - a);
            exits.push_back(b);
            ev.push_back({a, 1});
            ev.push_back({b, -1});
        }
        if (exits.empty() || t1 <= t0) continue;
        std::sort(ev.begin(), ev.end());
        int cur = 0, resident = 0;
        for (auto &e : ev) resident = std::max(resident, cur += e.second);
        std::nth_element(exits.begin(), exits.begin() + exits.size() / 2, exits.end());
        const double dur = (double)(t1 - t0);
        busy_w[kind] += inside / resident;
        dur_w[kind] += dur;
        st.search_tail_ms[kind] = std::max(st.search_tail_ms[kind], (t1 - exits[exits.size() / 2]) * ms_per_tick);
    }
    for (int k = 0; k < kNumKinds; k++) st.search_busy[k] = dur_w[k] > 0 ? busy_w[k] / dur_w[k] : 0;
}

// Builds one MPHF per key set; set s has set_n[s] keys at set_keys[s] (device).
// All sets share one leaf size and bucket size; their buckets, tree nodes and
// seed searches run together.  With device_copy, the (single) result also keeps
// a device-resident copy for batched GPU queries.  With so (one set only), the
// structure is serialized straight from device memory into so->buf and no handle
// is made.
static std::vector<rs_mphf *> build_sets(rs_ctx *c, const std::vector<const uint64_t *> &set_keys,
                                         const std::vector<uint64_t> &set_n, uint32_t leaf, uint64_t bucket_size,
                                         int flags, bool device_copy,
                                         const std::chrono::steady_clock::time_point t_start,
                                         const SerialOut *so = nullptr) {
    if (c->autotune && !c->tuned[leaf] && !t_arena) autotune(c, leaf, bucket_size, flags, 0);
    const bool rotate = (flags & kFlagRotate) != 0, r64 = (flags & kFlagRemap64) != 0;
    std::vector<uint64_t> set_off(set_n.size() + 1, 0);  // key offsets of the sets in the sorted array
    for (size_t i = 0; i < set_n.size(); i++) set_off[i + 1] = set_off[i] + set_n[i];
    hipStream_t s = c->stream;
    c->n_ev = 0;
    const size_t S = set_off.size() - 1;
    const SplitParams P = split_params(leaf);
    const uint32_t *memo = ctx_memo(c, leaf);
    const Memo &hmemo = golomb_memo(leaf);
    rs_build_stats st{};

    // ---- sets, buckets and key tiles
    std::vector<uint64_t> nbs(S), bbase(S + 1);
    std::vector<KeyTile> tiles;
    uint64_t NB = 0, max_nb = 0;
    // larger tiles when there are keys for two per CU: fewer global atomics per key
    const uint64_t tile_keys = set_off[S] - set_off[0] >= 8 * kTileKeys * (uint64_t)c->num_cus ? 4 * kTileKeys : kTileKeys;
    for (size_t i = 0; i < S; i++) {
        const uint64_t n = set_off[i + 1] - set_off[i];
        nbs[i] = std::max<uint64_t>(1, (n + bucket_size - 1) / bucket_size);
        bbase[i] = NB;
        NB += nbs[i];
        max_nb = std::max(max_nb, nbs[i]);
        for (uint64_t k = 0; k < n; k += tile_keys)
            tiles.push_back(KeyTile{set_keys[i], k, std::min(k + tile_keys, n), nbs[i], bbase[i]});
    }
    bbase[S] = NB;
    const uint64_t n_total = set_off[S] - set_off[0];
    st.num_buckets = NB;

    // ---- 0. time limit reference point
    unsigned long long *clock0 = nullptr, limit_ticks = 0;
    if (c->time_limit_s > 0) {
        clock0 = c->clock.get<unsigned long long>(1);
        limit_ticks = std::max<unsigned long long>(1, (unsigned long long)(c->time_limit_s * c->wall_khz * 1e3));
        k_clock_mark<<<1, 64, 0, s>>>(clock0);
        HIP_CHECK(hipGetLastError());
    }

    // ---- 1. bucket histogram and per-bucket metadata
    const int e_part0 = ev_record(c);
    KeyTile *d_tiles = c->tiles.get<KeyTile>(tiles.size());
    HIP_CHECK(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(KeyTile), hipMemcpyHostToDevice, s));
    uint32_t *counts = c->counts.get<uint32_t>(NB);
    uint64_t *counts64 = c->counts64.get<uint64_t>(NB);
    uint64_t *acc = c->acc.get<uint64_t>(NB + 1);
    uint64_t *node_cnt = c->node_cnt.get<uint64_t>(NB);
    uint64_t *node_base = c->node_base.get<uint64_t>(NB + 1);
    uint32_t *size_hist = c->size_hist.get<uint32_t>(kMaxBucket + 1);
    uint32_t *err = c->err.get<uint32_t>(4);
    HIP_CHECK(hipMemsetAsync(counts, 0, NB * 4, s));
    HIP_CHECK(hipMemsetAsync(size_hist, 0, (kMaxBucket + 1) * 4, s));
    HIP_CHECK(hipMemsetAsync(err, 0, 16, s));
    const uint32_t priv_max = (uint32_t)std::min<uint64_t>(max_nb, 16384);
    k_bucket_hist<<<tiles.size(), 256, (size_t)priv_max * 4, s>>>(d_tiles, counts, priv_max);
    HIP_CHECK(hipGetLastError());
    k_bucket_meta<<<(uint32_t)std::min<uint64_t>((NB + 255) / 256, (uint64_t)c->num_cus * 4), 256, 0, s>>>(
        counts, NB, memo + kMaxBucket, counts64, node_cnt, size_hist);
    HIP_CHECK(hipGetLastError());
    scan_u64(c, counts, true, NB, acc);
    scan_u64(c, node_cnt, false, NB, node_base);
    uint32_t *h = c->h_pinned;
    HIP_CHECK(hipMemcpyAsync(h, size_hist, (kMaxBucket + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h[kMaxBucket]) throw BuildError(RS_ERR_BUCKET_TOO_LARGE, "a bucket holds 3000 or more keys");
    Census cen;
    {
        std::vector<SubCensus> small(kMaxBucket);
        for (uint32_t m = 2; m < kMaxBucket; m++) {
            if (!h[m]) continue;
            st.max_bucket_size = m;
            census_node(P, hmemo, rotate, m, 0, h[m], cen, small);
        }
        if (!st.max_bucket_size) st.max_bucket_size = h[1] ? 1 : 0;
    }
    uint64_t total_nodes = 0;
    std::vector<uint64_t> lbase(kNumLists);
    for (int l = 0; l < kNumLists; l++) {
        lbase[l] = total_nodes;
        total_nodes += cen.count[l];
    }
    st.num_nodes = total_nodes;
    for (int l = 0; l < kNumLists; l++) st.nodes[l < (int)kMaxDepth ? kHigher : l - kMaxDepth + 1] += cen.count[l];

    // ---- 2. scatter and per-bucket sort
    uint64_t *sorted = c->sorted.get<uint64_t>(n_total + 8);  // +8: look-ahead loads of the search
    uint32_t *cursor = c->cursor.get<uint32_t>(NB);
    HIP_CHECK(hipMemsetAsync(cursor, 0, NB * 4, s));
    k_scatter<<<tiles.size(), 256, (size_t)priv_max * 4, s>>>(d_tiles, acc, cursor, sorted, priv_max);
    HIP_CHECK(hipGetLastError());
    uint32_t sort_p = 1;
    while (sort_p < st.max_bucket_size) sort_p <<= 1;
    k_sort_buckets<<<NB, st.max_bucket_size > 512 ? 256 : 64, (size_t)sort_p * 8, s>>>(sorted, acc, err,
                                                                                       rotate ? leaf : 0);
    HIP_CHECK(hipGetLastError());

    // ---- 3. node lists
    const uint64_t pool_n = std::max<uint64_t>(total_nodes, 1);
    NodeDesc *pool = c->pool.get<NodeDesc>(pool_n);
    uint64_t *d_lbase = c->list_base.get<uint64_t>(kNumLists);
    uint32_t *d_lfill = c->list_fill.get<uint32_t>(kNumLists);
    uint16_t *node_m = c->node_m.get<uint16_t>(pool_n);
    uint64_t *seeds = c->seeds.get<uint64_t>(pool_n);
    HIP_CHECK(hipMemcpyAsync(d_lbase, lbase.data(), kNumLists * 8, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(d_lfill, 0, kNumLists * 4, s));
    ExpandArgs ea{P, rotate ? 1 : 0, pool, d_lbase, d_lfill, node_m};
    k_expand<<<(NB + 63) / 64, 64, 0, s>>>(acc, node_base, NB, ea);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync((uint64_t *)h + 1, node_base + NB, 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h[0] & 1) throw BuildError(RS_ERR_DUPLICATE_KEYS, "two keys of one bucket share the same `second` half");
    if (((uint64_t *)h)[1] != total_nodes)
        throw BuildError(RS_ERR_INTERNAL, "seed slots " + std::to_string(((uint64_t *)h)[1]) + " != tree nodes " +
                                              std::to_string(total_nodes));
    const int e_part1 = ev_record(c);

    // ---- 4. search phases
    uint64_t max_list = 1;
    for (int l = 0; l < kNumLists; l++) max_list = std::max(max_list, cen.count[l]);
    NodeState *state = c->state.get<NodeState>(max_list);
    const Tune tn = tune_for(c, leaf);
    const uint32_t blocks = (uint32_t)c->num_cus * (uint32_t)tn.blocks_per_cu;
    const uint32_t n_waves = blocks * 4;
    st.search_waves = n_waves;
    int *running = c->running.get<int>(n_waves);
    uint32_t *ctl = c->ctl.get<uint32_t>(16);
    unsigned long long *work = c->work.get<unsigned long long>(kNumLists);
    HIP_CHECK(hipMemsetAsync(work, 0, kNumLists * 8, s));
    unsigned long long *wave_clock = nullptr;
    if (c->wave_times && !t_arena) {
        wave_clock = c->wave_clock.get<unsigned long long>((size_t)kNumLists * 2 * n_waves);
        HIP_CHECK(hipMemsetAsync(wave_clock, 0, (size_t)kNumLists * 2 * n_waves * 8, s));
    }
    const AggrLayout up_l = aggr_layout(P.lower, P.upper / P.lower);
    const AggrLayout lo_l = aggr_layout(P.leaf, P.lower / P.leaf);
    struct PhaseEvents {
        uint32_t kind;
        int search0, search1, split1;
    };
    std::vector<PhaseEvents> phase_ev;
    for (int l = 0; l < kNumLists; l++) {
        const uint32_t cnt = (uint32_t)cen.count[l];
        if (!cnt) continue;
        const uint32_t kind = l < (int)kMaxDepth ? kHigher : (uint32_t)(l - kMaxDepth + 1);
        const NodeDesc *nodes = pool + lbase[l];
        const uint32_t lblocks = level_blocks(cen.evals[l], (double)cnt, blocks);
        k_state_init<<<(cnt + 255) / 256, 256, 0, s>>>(state, cnt);
        HIP_CHECK(hipMemsetAsync(ctl, 0, 64, s));
        SearchArgs a{};
        a.nodes = nodes;
        a.state = state;
        a.running = running;
        a.ctl = ctl;
        a.n_nodes = cnt;
        a.leaf = P.leaf;
        a.upper = P.upper;
        a.work = work + l;
        a.clock0 = clock0;
        a.limit_ticks = limit_ticks;
        a.err = err;
        a.wave_clock = wave_clock ? wave_clock + (size_t)l * 2 * n_waves : nullptr;
        a.groups = (uint32_t)tn.groups[kind];
        a.help_cap = (uint32_t)tn.help_cap[kind];
        bool wide = false, lut = false;
        size_t lds = 0;
        if (kind == kUpper || kind == kLower) {
            const AggrLayout &L = kind == kUpper ? up_l : lo_l;
            a.split_unit = kind == kUpper ? P.lower : P.leaf;
            a.field_bits = L.fb;
            a.magic = L.magic;
            wide = L.wide;
            // one increment table per workgroup, indexed by the remapped value
            // (below the largest node size of the phase): 4-byte entries, 8-byte
            // ones for 64-bit counters; tables above 32 KiB would cut occupancy
            a.lut_entries = kind == kUpper ? P.upper : P.lower;
            lds = (size_t)a.lut_entries * (wide ? 8 : 4);
            lut = tn.lut != 0 && lds <= 32768;
            if (!lut) lds = 0;
        }
        const SearchShape sh =
            search_shape(tn, kind, wide, lut, r64, 2 * cen.full[l] >= cnt, lblocks, blocks, (uint32_t)c->num_cus);
        a.bt_m = sh.bt ? (kind == kLeafRot ? P.leaf : a.lut_entries) : 0;
        a.n_waves = sh.waves;
        if (sh.waves > n_waves) throw BuildError(RS_ERR_INTERNAL, "search launch beyond the running table");
        HIP_CHECK(hipMemsetAsync(running, 0xFF, (size_t)sh.waves * 4, s));
        st.kind_waves[kind] = std::max(st.kind_waves[kind], sh.waves);
        // Levels of many small nodes (millions of leaves of many small sets) would
        // otherwise be bound by the one fresh-node counter (one atomic per node).
        claim_params(cnt, sh.waves, a.claim_batch, a.batch_tickets);
        PhaseEvents pe{kind, ev_record(c), -1, -1};
        launch_search(s, tn, kind, sorted, a, wide, lut, sh.grid, lds, r64);
        pe.search1 = ev_record(c);
        if (kind == kHigher || kind == kUpper || kind == kLower) {
            const uint32_t max_m = kind == kHigher ? st.max_bucket_size : (kind == kUpper ? P.upper : P.lower);
            k_split<<<cnt, 64, (size_t)max_m * 8, s>>>(sorted, nodes, state, seeds, err, P, kind, r64 ? 1 : 0);
        } else {
            k_finalize<<<(cnt + 255) / 256, 256, 0, s>>>(nodes, state, cnt, seeds, err);
        }
        HIP_CHECK(hipGetLastError());
        pe.split1 = ev_record(c);
        phase_ev.push_back(pe);
    }
    const int e_search_end = ev_record(c);

    // ---- 5. descriptor lengths (global scan; each set reads its segment)
    uint64_t *bits = c->bits.get<uint64_t>(NB);
    uint64_t *pos = c->pos.get<uint64_t>(NB + 1);
    unsigned long long *evals = c->evals.get<unsigned long long>(2 * kNumKinds);  // evaluations, then seeds
    HIP_CHECK(hipMemsetAsync(evals, 0, 2 * kNumKinds * 8, s));
    EncodeArgs ca{acc, node_base, seeds, node_m, memo, memo + kMaxBucket, NB, P, rotate ? 1 : 0};
    k_bucket_bits<<<(uint32_t)std::min<uint64_t>((NB + 3) / 4, (uint64_t)c->num_cus * 8), 256, 0, s>>>(ca, bits, evals);
    HIP_CHECK(hipGetLastError());
    scan_u64(c, bits, false, NB, pos);
    uint64_t *d_bidx = c->set_idx.get<uint64_t>(S + 1);
    uint64_t *d_gath = c->set_gather.get<uint64_t>(S + 1);
    HIP_CHECK(hipMemcpyAsync(d_bidx, bbase.data(), (S + 1) * 8, hipMemcpyHostToDevice, s));
    k_gather_u64<<<(S + 1 + 255) / 256, 256, 0, s>>>(pos, d_bidx, S + 1, d_gath);
    HIP_CHECK(hipGetLastError());
    std::vector<uint64_t> pos_at(S + 1);
    uint64_t *h64 = (uint64_t *)c->h_pinned;
    HIP_CHECK(hipMemcpyAsync(pos_at.data(), d_gath, (S + 1) * 8, hipMemcpyDeviceToHost, s));
    // pinned staging: error flags at byte 64, work counters at 128..296, evaluations
    // and seeds at 512..592
    HIP_CHECK(hipMemcpyAsync(h64 + 64, evals, 2 * kNumKinds * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h64 + 16, work, kNumLists * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h + 16, err, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < kNumKinds; k++) {
        st.evals[k] = h64[64 + k];
        st.seeds[k] = h64[64 + kNumKinds + k];
    }
    static_assert(16 + kNumLists <= 64, "work counters overlap the evaluation counts");
    for (int l = 0; l < kNumLists; l++)
        st.executed_evals[l < (int)kMaxDepth ? kHigher : l - kMaxDepth + 1] += h64[16 + l];
    if (wave_clock) wave_time_stats(c, wave_clock, n_waves, cen.count, st);
    if (h[16] & 8) throw BuildError(RS_ERR_TIMEOUT, "the build passed the context's time limit during the seed search");
    if (h[16] & 2) throw BuildError(RS_ERR_SEED_OVERFLOW, "a node search passed the 2^44 seed limit");

    // ---- 6. Golomb-Rice: per-set vectors packed one after another
    std::vector<SetEnc> enc(S);
    std::vector<uint64_t> rice_words(S);
    uint64_t rice_total = 0;
    for (size_t i = 0; i < S; i++) {
        const uint64_t B = pos_at[i + 1] - pos_at[i] + 1;  // with the sentinel
        rice_words[i] = ((B + 7) / 8 + 14) / 8;
        enc[i].bbase = bbase[i];
        enc[i].nb = nbs[i];
        enc[i].rice_word = rice_total;
        rice_total += rice_words[i];
        st.rice_bits += B;
    }
    SetEnc *d_enc = c->set_enc.get<SetEnc>(S);
    HIP_CHECK(hipMemcpyAsync(d_enc, enc.data(), S * sizeof(SetEnc), hipMemcpyHostToDevice, s));
    unsigned long long *rice = c->rice.get<unsigned long long>(rice_total + 2);
    long long *bit_off = c->bit_off.get<long long>(NB);
    HIP_CHECK(hipMemsetAsync(rice, 0, (rice_total + 2) * 8, s));
    k_set_offsets<<<S, 256, 0, s>>>(d_enc, pos, bit_off, rice);
    k_rice_write<<<(NB + 3) / 4, 256, 0, s>>>(ca, pos, bit_off, rice);
    EfScalars *d_efs = c->ef_scalars.get<EfScalars>(S);
    k_ef_scalars<<<S, 256, 0, s>>>(d_enc, acc, pos, d_efs);
    HIP_CHECK(hipGetLastError());
    std::vector<EfScalars> he(S);
    HIP_CHECK(hipMemcpyAsync(he.data(), d_efs, S * sizeof(EfScalars), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));

    // ---- 7. Elias-Fano: per-set widths and sizes (DoubleEF.hpp:181-202)
    std::vector<EliasFano> efs(S);
    std::vector<std::array<uint64_t, 4>> ef_words(S);  // lower, upper (keys), upper (positions), jump
    uint64_t tot_lower = 0, tot_uc = 0, tot_up = 0, tot_jump = 0;
    for (size_t i = 0; i < S; i++) {
        EliasFano &ef = efs[i];
        const uint64_t nb = nbs[i], n = set_off[i + 1] - set_off[i];
        ef.num_buckets = nb;
        ef.bits_per_key_fixed_point = he[i].bpk;
        ef.cum_keys_min_delta = he[i].min_delta;
        ef.min_diff = he[i].min_diff;
        ef.u_position = (uint64_t)((int64_t)he[i].total_bits - (int64_t)(he[i].bpk * n >> 20) -
                                   (int64_t)(nb * (uint64_t)he[i].min_diff) + 1);
        ef.u_cum_keys = n - nb * (uint64_t)he[i].min_delta + 1;
        ef.derive_widths();
        enc[i].lc = ef.l_cum_keys;
        enc[i].lp = ef.l_position;
        enc[i].lower_word = tot_lower;
        enc[i].uc_word = tot_uc;
        enc[i].up_word = tot_up;
        enc[i].jump_word = tot_jump;
        enc[i].jump_words = ef_jump_words(nb);
        ef_words[i][0] = ef_lower_words(nb, ef.l_cum_keys, ef.l_position);
        ef_words[i][1] = ef_upper_words(nb, ef.u_cum_keys, ef.l_cum_keys);
        ef_words[i][2] = ef_upper_words(nb, ef.u_position, ef.l_position);
        ef_words[i][3] = enc[i].jump_words;
        tot_lower += ef_words[i][0];
        tot_uc += ef_words[i][1];
        tot_up += ef_words[i][2];
        tot_jump += ef_words[i][3];
    }
    HIP_CHECK(hipMemcpyAsync(d_enc, enc.data(), S * sizeof(SetEnc), hipMemcpyHostToDevice, s));
    unsigned long long *d_lower = c->ef_lower.get<unsigned long long>(tot_lower + 1);
    unsigned long long *d_uc = c->ef_uc.get<unsigned long long>(tot_uc + 1);
    unsigned long long *d_up = c->ef_up.get<unsigned long long>(tot_up + 1);
    uint8_t *d_jump = c->ef_jump.get<uint8_t>(tot_jump * 8 + 8);
    HIP_CHECK(hipMemsetAsync(d_lower, 0, (tot_lower + 1) * 8, s));
    HIP_CHECK(hipMemsetAsync(d_uc, 0, (tot_uc + 1) * 8, s));
    HIP_CHECK(hipMemsetAsync(d_up, 0, (tot_up + 1) * 8, s));
    HIP_CHECK(hipMemsetAsync(d_jump, 0, tot_jump * 8 + 8, s));
    k_ef_fill<<<S, 256, 0, s>>>(d_enc, d_efs, acc, pos, d_lower, d_uc, d_up, d_jump, err);
    HIP_CHECK(hipGetLastError());
    const int e_enc_end = ev_record(c);

    // ---- 8. results to the host
    if (so) {
        // serialized layout (AbstractParallelRecSplit.hpp:480-488, DoubleEF.hpp:100-113):
        // three header words, the descriptor vector, six EF scalars, four EF vectors,
        // each vector a word count then the words, copied straight from HBM
        const EliasFano &ef = efs[0];
        const uint64_t need = 24 + 8 + 8 * rice_total + 48 + 4 * 8 + 8 * (tot_lower + tot_uc + tot_up + tot_jump);
        *so->len = need;
        if (so->cap < need)
            throw BuildError(RS_ERR_ARG, "output buffer of " + std::to_string(so->cap) + " bytes; the structure needs " +
                                             std::to_string(need));
        uint8_t *p = so->buf;
        auto word = [&](uint64_t v) {
            memcpy(p, &v, 8);
            p += 8;
        };
        auto vec = [&](const void *src, uint64_t words) {
            word(words);
            if (words) HIP_CHECK(hipMemcpyAsync(p, src, words * 8, hipMemcpyDeviceToHost, s));
            p += 8 * words;
        };
        word(leaf);
        word(bucket_size);
        word(set_off[1] - set_off[0]);
        vec(rice, rice_total);
        word(ef.num_buckets);
        word(ef.u_cum_keys);
        word(ef.u_position);
        word((uint64_t)ef.cum_keys_min_delta);
        word((uint64_t)ef.min_diff);
        word(ef.bits_per_key_fixed_point);
        vec(d_lower, tot_lower);
        vec(d_uc, tot_uc);
        vec(d_up, tot_up);
        vec(d_jump, tot_jump);
        HIP_CHECK(hipMemcpyAsync(h + 16, err, 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (h[16] & 4) throw BuildError(RS_ERR_EF_OVERFLOW, "Elias-Fano jump offset does not fit 16 bits");
        return {};
    }
    // One copy per array, all sets at once, into the context's pinned staging (a
    // copy into fresh pageable memory took 27 ms longer for the 126 MB of 4096
    // sets of 590M keys); the handles take their words from there.
    uint64_t *h_rice = c->h_result.get(8 * (rice_total + tot_lower + tot_uc + tot_up + tot_jump + 5));
    uint64_t *h_lower = h_rice + rice_total + 1, *h_uc = h_lower + tot_lower + 1, *h_up = h_uc + tot_uc + 1,
             *h_jump = h_up + tot_up + 1;
    HIP_CHECK(hipMemcpyAsync(h_rice, rice, rice_total * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h_lower, d_lower, tot_lower * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h_uc, d_uc, tot_uc * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h_up, d_up, tot_up * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h_jump, d_jump, tot_jump * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h + 16, err, 4, hipMemcpyDeviceToHost, s));
    DeviceCopy *dc = nullptr;
    if (device_copy && S == 1) {
        dc = new DeviceCopy();
        dc->device = c->device;
        dc->n_rice = rice_total;
        dc->n_lower = ef_words[0][0];
        dc->n_uc = ef_words[0][1];
        dc->n_up = ef_words[0][2];
        dc->n_jump = ef_words[0][3];
        HIP_CHECK(hipMalloc(&dc->rice, (rice_total + 2) * 8));
        HIP_CHECK(hipMalloc(&dc->lower, (dc->n_lower + 1) * 8));
        HIP_CHECK(hipMalloc(&dc->upper_c, (dc->n_uc + 1) * 8));
        HIP_CHECK(hipMalloc(&dc->upper_p, (dc->n_up + 1) * 8));
        HIP_CHECK(hipMalloc(&dc->jump, dc->n_jump * 8 + 8));
        HIP_CHECK(hipMalloc(&dc->memo, kMaxBucket * 4));
        HIP_CHECK(hipMemcpyAsync(dc->rice, rice, (rice_total + 2) * 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dc->lower, d_lower, (dc->n_lower + 1) * 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dc->upper_c, d_uc, (dc->n_uc + 1) * 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dc->upper_p, d_up, (dc->n_up + 1) * 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dc->jump, d_jump, dc->n_jump * 8 + 8, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(dc->memo, memo, kMaxBucket * 4, hipMemcpyDeviceToDevice, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (h[16] & 4) {
        if (dc) device_copy_free(dc);
        throw BuildError(RS_ERR_EF_OVERFLOW, "Elias-Fano jump offset does not fit 16 bits");
    }

    // ---- timings
    st.partition_ms = ev_ms(c, e_part0, e_part1);
    for (const PhaseEvents &pe : phase_ev) {
        st.search_ms[pe.kind] += ev_ms(c, pe.search0, pe.search1);
        st.split_ms += ev_ms(c, pe.search1, pe.split1);
    }
    st.encode_ms = ev_ms(c, e_search_end, e_enc_end);
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();

    // One handle per set, each with its own copy of its words; many sets are
    // assembled on several host threads.
    std::vector<rs_mphf *> out(S, nullptr);
    auto make = [&](size_t i) {
        rs_mphf *r = new rs_mphf();
        r->leaf = leaf;
        r->rotate = flags & (kFlagRotate | kFlagRemap64);
        r->bucket_size = bucket_size;
        r->keys_count = set_off[i + 1] - set_off[i];
        r->nbuckets = nbs[i];
        r->rice_words = rice_words[i];
        r->rice.reserve(rice_words[i] + 2);
        r->rice.assign(h_rice + enc[i].rice_word, h_rice + enc[i].rice_word + rice_words[i]);
        r->rice.resize(rice_words[i] + 2, 0);
        EliasFano &ef = efs[i];
        ef.lower_bits.assign(h_lower + enc[i].lower_word, h_lower + enc[i].lower_word + ef_words[i][0]);
        ef.upper_cum.assign(h_uc + enc[i].uc_word, h_uc + enc[i].uc_word + ef_words[i][1]);
        ef.upper_pos.assign(h_up + enc[i].up_word, h_up + enc[i].up_word + ef_words[i][2]);
        ef.jump.assign(h_jump + enc[i].jump_word, h_jump + enc[i].jump_word + ef_words[i][3]);
        r->ef = std::move(ef);
        r->memo = &hmemo;
        r->P = P;
        r->stats = st;
        out[i] = r;
    };
    const size_t n_threads = S >= 256 ? std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency())) : 1;
    bool failed = false;
    if (n_threads <= 1) {
        try {
            for (size_t i = 0; i < S; i++) make(i);
        } catch (const std::bad_alloc &) {
            failed = true;
        }
    } else {
        std::vector<std::thread> pool;
        std::vector<char> bad(n_threads, 0);
        for (size_t t = 0; t < n_threads; t++)
            pool.emplace_back([&, t] {
                try {
                    for (size_t i = t; i < S; i += n_threads) make(i);
                } catch (const std::bad_alloc &) {
                    bad[t] = 1;
                }
            });
        for (std::thread &th : pool) th.join();
        for (char b : bad) failed |= b != 0;
    }
    if (failed) {
        for (rs_mphf *r : out) delete r;
        if (dc) device_copy_free(dc);
        throw std::bad_alloc();
    }
    out[0]->dev = dc;
    return out;
}

static rs_mphf *build(rs_ctx *c, const uint64_t *d_keys, uint64_t n, uint32_t leaf, uint64_t bucket_size,
                      int flags, const std::chrono::steady_clock::time_point t_start) {
    return build_sets(c, {d_keys}, {n}, leaf, bucket_size, flags, true, t_start)[0];
}

// Every reusable device buffer of a context.
static std::vector<DevBuf *> ctx_bufs(rs_ctx *c) {
    return {&c->keys_in, &c->counts, &c->counts64, &c->cursor, &c->acc, &c->node_cnt, &c->node_base,
            &c->size_hist, &c->scan_tmp, &c->sorted, &c->pool, &c->list_base, &c->list_fill, &c->node_m,
            &c->seeds, &c->state, &c->running, &c->ctl, &c->err, &c->bits, &c->pos, &c->evals, &c->work, &c->rice,
            &c->ef_scalars, &c->ef_lower, &c->ef_uc, &c->ef_up, &c->ef_jump, &c->tiles, &c->set_idx,
            &c->set_gather, &c->set_enc, &c->bit_off, &c->clock, &c->str_bytes, &c->str_off, &c->tune_keys};
}

// SplitMix64 keys on the device (the generator of the synthetic workloads): key i
// is (next(), next()) of the stream from `seed`.
__global__ __launch_bounds__(256) void k_splitmix_keys(uint64_t seed, uint64_t n, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (2 * i + 1 + j);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        out[2 * i + j] = z ^ (z >> 31);
    }
}

// Autotuner: picks, per node kind, the candidates per lane (U) and candidate
// groups per claimed chunk, then the workgroups per CU, by timing builds of a
// synthetic sample (SplitMix64 keys) of this leaf size, bucket size and flags.
// Every kind has its own search launch, so one trial measures all kinds at once;
// the sample is sized so that a build takes about 40 ms.  The result becomes
// the context's tuning for the leaf size.  Returns a JSON report.
static std::string autotune(rs_ctx *c, uint32_t leaf, uint64_t bucket, int flags, uint64_t sample_keys) {
    const Tune base = tune_for(c, leaf);
    c->tuned[leaf] = true;  // builds below run with explicit tunings (no recursion)
    const uint64_t n_max = uint64_t(1) << 22;
    uint64_t *keys = c->tune_keys.get<uint64_t>(2 * n_max);
    k_splitmix_keys<<<(n_max + 255) / 256, 256, 0, c->stream>>>(0x5eed, n_max, keys);
    HIP_CHECK(hipGetLastError());
    auto trial = [&](const Tune &t, uint64_t n) {
        c->tune_leaf[leaf] = t;
        std::vector<rs_mphf *> r = build_sets(c, {keys}, {n}, leaf, bucket, flags, false, std::chrono::steady_clock::now());
        const rs_build_stats st = r[0]->stats;
        for (rs_mphf *h : r) rs_mphf_free(h);
        return st;
    };
    uint64_t n = sample_keys ? std::min(sample_keys, n_max) : std::max<uint64_t>(2048, bucket);
    rs_build_stats st0 = trial(base, n);
    auto search_total = [](const rs_build_stats &st) {
        double t = 0;
        for (int k = 0; k < kNumKinds; k++) t += st.search_ms[k];
        return t;
    };
    // without a given size, the sample doubles until one search takes >= 40 ms (or
    // reaches 4M keys): small samples time the launch tails, not the throughput
    while (!sample_keys && search_total(st0) < 40.0 && n < n_max) {
        n = std::min<uint64_t>(n_max, 2 * n);
        st0 = trial(base, n);
    }
    const int reps = st0.total_ms > 100 ? 1 : 2;
    auto timed = [&](const Tune &t) {
        rs_build_stats best = trial(t, n);
        for (int r = 1; r < reps; r++) {
            const rs_build_stats s2 = trial(t, n);
            for (int k = 0; k < kNumKinds; k++) best.search_ms[k] = std::min(best.search_ms[k], s2.search_ms[k]);
        }
        return best;
    };
    auto total = [](const rs_build_stats &st) {
        double t = 0;
        for (int k = 0; k < kNumKinds; k++) t += st.search_ms[k];
        return t;
    };
    const rs_build_stats ref = timed(base);
    std::string rep = "{\"leaf\":" + std::to_string(leaf) + ",\"bucket\":" + std::to_string(bucket) +
                      ",\"flags\":" + std::to_string(flags) + ",\"sample_keys\":" + std::to_string(n) +
                      ",\"search_ms_before\":" + std::to_string(total(ref)) + ",\"trials\":[";
    Tune best = base;
    double best_ms[kNumKinds];
    for (int k = 0; k < kNumKinds; k++) best_ms[k] = ref.search_ms[k];
    const int us[3] = {1, 2, 4}, gs[3] = {1, 2, 4};
    bool first = true;
    for (int u : us)
        for (int g : gs) {
            Tune t = base;
            for (int k = 0; k < kNumKinds; k++) {
                t.u[k] = u;
                t.groups[k] = g;
            }
            t.u_bt = u;
            t.u_bt_wide = u;
            const rs_build_stats st = timed(t);
            rep += std::string(first ? "" : ",") + "{\"u\":" + std::to_string(u) + ",\"g\":" + std::to_string(g) +
                   ",\"search_ms\":[";
            for (int k = 0; k < kNumKinds; k++) {
                rep += (k ? "," : "") + std::to_string(st.search_ms[k]);
                if (st.nodes[k] && st.search_ms[k] < best_ms[k]) {
                    best_ms[k] = st.search_ms[k];
                    best.u[k] = u;
                    best.groups[k] = g;
                    if (k == kLower) best.u_bt = best.u_bt_wide = u;  // byte-table launches (lower level dominates)
                }
            }
            rep += "]}";
            first = false;
        }
    double best_total = total(timed(best));
    const int bpcs[3] = {6, 8, 10};
    for (int b : bpcs) {
        if (b == best.blocks_per_cu) continue;
        Tune t = best;
        t.blocks_per_cu = b;
        const double ms = total(timed(t));
        rep += ",{\"bpc\":" + std::to_string(b) + ",\"search_ms_total\":" + std::to_string(ms) + "}";
        if (ms < best_total) {
            best_total = ms;
            best.blocks_per_cu = b;
        }
    }
    c->tune_leaf[leaf] = best;
    rep += "],\"search_ms_after\":" + std::to_string(best_total) + ",\"tune\":\"" + format_tune(best) + "\"}";
    return rep;
}

static int guarded(const char *what, rs_ctx *c, const std::function<void()> &fn) {
    try {
        std::lock_guard<std::mutex> g(c->mu);
        HIP_CHECK(hipSetDevice(c->device));
        fn();
        return RS_OK;
    } catch (const BuildError &e) {
        set_error(std::string(what) + ": " + e.what());
        return e.code;
    } catch (const HipError &e) {
        set_error(std::string(what) + ": " + e.what());
        return RS_ERR_HIP;
    } catch (const std::bad_alloc &) {
        set_error(std::string(what) + ": out of host memory");
        return RS_ERR_NOMEM;
    }
}

// The structure's device copy on ctx's GPU, made from its host arrays on first
// use (deserialized handles and batch results have none).  A copy on another
// device is replaced.  Uploads of one handle are serialised by a process-wide lock.
static const DeviceCopy *device_copy_of(rs_ctx *c, rs_mphf *h) {
    static std::mutex upload_mu;
    std::lock_guard<std::mutex> g(upload_mu);
    if (h->dev && h->dev->device == c->device) return h->dev;
    if (h->dev) {
        device_copy_free(h->dev);
        h->dev = nullptr;
    }
    DeviceCopy *dc = new DeviceCopy();
    dc->device = c->device;
    const EliasFano &ef = h->ef;
    auto up = [&](uint64_t **dst, const uint64_t *src, uint64_t words) {
        // one zero word of padding beyond the readable words
        HIP_CHECK(hipMalloc(dst, (words + 1) * 8));
        HIP_CHECK(hipMemsetAsync(*dst, 0, (words + 1) * 8, c->stream));
        if (words) HIP_CHECK(hipMemcpyAsync(*dst, src, words * 8, hipMemcpyHostToDevice, c->stream));
    };
    try {
        up(&dc->rice, h->rice.data(), h->rice_words);
        up(&dc->lower, ef.lower_bits.data(), ef.lower_bits.size());
        up(&dc->upper_c, ef.upper_cum.data(), ef.upper_cum.size());
        up(&dc->upper_p, ef.upper_pos.data(), ef.upper_pos.size());
        up(&dc->jump, ef.jump.data(), ef.jump.size());
        HIP_CHECK(hipMalloc(&dc->memo, kMaxBucket * 4));
        HIP_CHECK(hipMemcpyAsync(dc->memo, h->memo->data(), kMaxBucket * 4, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
        device_copy_free(dc);
        throw;
    }
    dc->n_rice = h->rice_words;
    dc->n_lower = ef.lower_bits.size();
    dc->n_uc = ef.upper_cum.size();
    dc->n_up = ef.upper_pos.size();
    dc->n_jump = ef.jump.size();
    h->dev = dc;
    return dc;
}

static int check_args(rs_ctx *c, const void *keys, size_t n, uint32_t leaf, size_t bucket, rs_mphf **out) {
    if (!c || !keys || !out) return set_error("null argument"), RS_ERR_ARG;
    if (n == 0) return set_error("no keys"), RS_ERR_ARG;
    if (leaf < 2 || leaf > kMaxLeaf) return set_error("leaf size must be in [2, 24]"), RS_ERR_ARG;
    if (bucket == 0) return set_error("bucket size must be positive"), RS_ERR_ARG;
    if (n >= (uint64_t(1) << 40)) return set_error("more than 2^40 keys"), RS_ERR_ARG;
    return RS_OK;
}

// Hooks of the construction server (rs_server.cpp).
int build_sets_status(rs_ctx *c, const std::vector<const uint64_t *> &set_keys, const std::vector<uint64_t> &set_n,
                      uint32_t leaf, uint64_t bucket, int flags, std::vector<rs_mphf *> &out) {
    const auto t0 = std::chrono::steady_clock::now();
    return guarded("build", c, [&] { out = build_sets(c, set_keys, set_n, leaf, bucket, flags, false, t0); });
}

uint64_t build_bound_bytes(uint64_t n, uint64_t nb) { return working_set_bound(n, nb); }

int ctx_device(const rs_ctx *c) { return c->device; }

}  // namespace rsa

extern "C" {

int rs_ctx_create(int device, rs_ctx **out) {
    if (!out) return RS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        set_error("rs_ctx_create: no such HIP device");
        return RS_ERR_HIP;
    }
    rs_ctx *c = new rs_ctx();
    c->device = device;
    try {
        HIP_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        c->tune = parse_tune(getenv("RS_TUNE"));
        if (const char *at = getenv("RS_AUTOTUNE")) c->autotune = atoi(at) != 0;
        HIP_CHECK(hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device));
        if (c->wall_khz <= 0) c->wall_khz = 100000;  // s_memrealtime: 100 MHz on gfx9 parts
        if (const char *tl = getenv("RS_TIME_LIMIT")) c->time_limit_s = std::max(0.0, atof(tl));
        if (const char *wt = getenv("RS_WAVE_TIMES")) c->wave_times = atoi(wt) != 0;
        {
            // the allocation-free entry point's scratch: this device's address, once
            std::lock_guard<std::mutex> g(g_scratch_mu);
            if (device >= kMaxDevices) throw HipError(hipErrorInvalidDevice, "device ordinal above 63", __LINE__);
            if (!g_scratch_ptr[device]) {
                void *p = nullptr;
                HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                g_scratch_ptr[device] = (uint8_t *)p;
            }
            c->scratch = g_scratch_ptr[device];
        }
        HIP_CHECK(hipStreamCreate(&c->stream));  // blocking: ordered after null-stream work
        HIP_CHECK(hipHostMalloc((void **)&c->h_pinned, 64 * 1024, hipHostMallocDefault));
        for (auto &e : c->ev) HIP_CHECK(hipEventCreate(&e));
    } catch (const HipError &e) {
        set_error(std::string("rs_ctx_create: ") + e.what());
        delete c;
        return RS_ERR_HIP;
    }
    *out = c;
    return RS_OK;
}

int rs_ctx_set_time_limit(rs_ctx *c, double seconds) {
    if (!c || !(seconds >= 0)) {
        set_error("rs_ctx_set_time_limit: null context or negative limit");
        return RS_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(c->mu);
    c->time_limit_s = seconds;
    return RS_OK;
}

int rs_autotune(rs_ctx *c, uint32_t leaf, size_t bucket, int flags, uint64_t sample_keys, char *report, size_t cap) {
    if (!c || leaf < 2 || leaf > kMaxLeaf || bucket == 0) return set_error("rs_autotune: bad arguments"), RS_ERR_ARG;
    std::string rep;
    const int rc = guarded("rs_autotune", c, [&] { rep = autotune(c, leaf, bucket, flags & 3, sample_keys); });
    if (rc) {
        c->tuned[leaf] = false;
        return rc;
    }
    if (report && cap) {
        const size_t k = std::min(cap - 1, rep.size());
        memcpy(report, rep.data(), k);
        report[k] = 0;
    }
    return RS_OK;
}

int rs_ctx_set_tune(rs_ctx *c, uint32_t leaf, const char *tune) {
    if (!c || !tune || leaf == 1 || leaf > kMaxLeaf) return set_error("rs_ctx_set_tune: bad arguments"), RS_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    if (leaf == 0) {
        c->tune = parse_tune(tune, c->tune);
    } else {
        c->tune_leaf[leaf] = parse_tune(tune, tune_for(c, leaf));
        c->tuned[leaf] = true;
    }
    return RS_OK;
}

int rs_ctx_get_tune(rs_ctx *c, uint32_t leaf, char *buf, size_t cap) {
    if (!c || !buf || !cap || leaf == 1 || leaf > kMaxLeaf) return set_error("rs_ctx_get_tune: bad arguments"), RS_ERR_ARG;
    std::lock_guard<std::mutex> g(c->mu);
    const std::string t = format_tune(leaf ? tune_for(c, leaf) : c->tune);
    const size_t k = std::min(cap - 1, t.size());
    memcpy(buf, t.data(), k);
    buf[k] = 0;
    return RS_OK;
}

uint64_t rs_ctx_device_bytes(rs_ctx *c) {
    if (!c) return 0;
    std::lock_guard<std::mutex> g(c->mu);
    uint64_t t = 0;
    for (DevBuf *b : ctx_bufs(c)) t += b->cap;
    for (auto &m : c->memo)
        if (m) t += sizeof(uint32_t) * kMaxBucket * 2;
    return t;
}

int rs_ctx_release(rs_ctx *c) {
    if (!c) return set_error("rs_ctx_release: null context"), RS_ERR_ARG;
    return guarded("rs_ctx_release", c, [&] {
        HIP_CHECK(hipStreamSynchronize(c->stream));
        for (DevBuf *b : ctx_bufs(c)) b->release();
        c->h_result.release();
    });
}

void rs_ctx_destroy(rs_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : ctx_bufs(c)) b->release();
    for (auto &m : c->memo)
        if (m) (void)hipFree(m);
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    c->h_result.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int rs_build_device(rs_ctx *c, const uint64_t *d_keys, size_t n, uint32_t leaf, size_t bucket, int rotate,
                    rs_mphf **out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (int e = check_args(c, d_keys, n, leaf, bucket, out)) return e;
    return guarded("rs_build_device", c, [&] { *out = build(c, d_keys, n, leaf, bucket, rotate & 3, t0); });
}

int rs_build_device_noalloc(rs_ctx *c, const uint64_t *d_keys, size_t n, uint32_t leaf, size_t bucket, int rotate,
                            uint8_t *out, size_t cap, size_t *out_len) {
    const auto t0 = std::chrono::steady_clock::now();
    rs_mphf *unused = nullptr;
    if (int e = check_args(c, d_keys, n, leaf, bucket, &unused)) return e;
    if (!out || !out_len) return set_error("rs_build_device_noalloc: null output"), RS_ERR_ARG;
    const uint64_t nb = std::max<uint64_t>(1, (n + bucket - 1) / bucket);
    const uint64_t need = working_set_bound(n, nb);
    if (need > kNoAllocScratchBytes) {
        set_error("rs_build_device_noalloc: " + std::to_string(n) + " keys in " + std::to_string(nb) +
                  " buckets may need " + std::to_string(need) + " bytes of scratch; the maximum is " +
                  std::to_string(kNoAllocScratchBytes) + " (75 B per key + 89 B per bucket + 1 MiB)");
        return RS_ERR_ARG;
    }
    return guarded("rs_build_device_noalloc", c, [&] {
        std::lock_guard<std::mutex> g(g_scratch_use[c->device]);
        Arena a;
        a.base = c->scratch;
        a.size = kNoAllocScratchBytes;
        t_arena = &a;
        struct Reset {
            ~Reset() { t_arena = nullptr; }
        } reset;
        const SerialOut so{out, cap, out_len};
        build_sets(c, {d_keys}, {n}, leaf, bucket, rotate & 3, false, t0, &so);
    });
}

int rs_hash_strings_device(rs_ctx *c, const uint8_t *d_bytes, const uint64_t *d_offsets, size_t n, uint64_t *d_keys) {
    if (!c || (n && (!d_bytes || !d_offsets || !d_keys))) return set_error("rs_hash_strings_device: null argument"), RS_ERR_ARG;
    if (n == 0) return RS_OK;
    return guarded("rs_hash_strings_device", c, [&] {
        k_spooky_keys<<<(n + 255) / 256, 256, 0, c->stream>>>(d_bytes, d_offsets, n, d_keys);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int rs_build_strings(rs_ctx *c, const uint8_t *bytes, const uint64_t *offsets, size_t n, uint32_t leaf, size_t bucket,
                     int rotate, rs_mphf **out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!offsets || (!bytes && n && offsets[n] > offsets[0])) return set_error("rs_build_strings: null argument"), RS_ERR_ARG;
    if (int e = check_args(c, offsets, n, leaf, bucket, out)) return e;
    for (size_t i = 0; i < n; i++)
        if (offsets[i + 1] < offsets[i]) return set_error("rs_build_strings: decreasing offsets"), RS_ERR_ARG;
    return guarded("rs_build_strings", c, [&] {
        const uint64_t base = offsets[0], total = offsets[n] - base;
        uint8_t *d_bytes = c->str_bytes.get<uint8_t>(total + 8);
        uint64_t *d_off = c->str_off.get<uint64_t>(n + 1);
        uint64_t *d_keys = c->keys_in.get<uint64_t>(2 * n);
        std::vector<uint64_t> rel(offsets, offsets + n + 1);
        for (uint64_t &o : rel) o -= base;
        if (total) HIP_CHECK(hipMemcpyAsync(d_bytes, bytes + base, total, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(d_off, rel.data(), (n + 1) * 8, hipMemcpyHostToDevice, c->stream));
        k_spooky_keys<<<(n + 255) / 256, 256, 0, c->stream>>>(d_bytes, d_off, n, d_keys);
        HIP_CHECK(hipGetLastError());
        *out = build(c, d_keys, n, leaf, bucket, rotate & 3, t0);
    });
}

int rs_build(rs_ctx *c, const uint64_t *keys, size_t n, uint32_t leaf, size_t bucket, int rotate, rs_mphf **out) {
    const auto t0 = std::chrono::steady_clock::now();
    if (int e = check_args(c, keys, n, leaf, bucket, out)) return e;
    return guarded("rs_build", c, [&] {
        uint64_t *d = c->keys_in.get<uint64_t>(2 * n);
        HIP_CHECK(hipMemcpyAsync(d, keys, n * 16, hipMemcpyHostToDevice, c->stream));
        *out = build(c, d, n, leaf, bucket, rotate & 3, t0);
    });
}

int rs_build_batch(rs_ctx *c, const uint64_t *d_keys, const uint64_t *set_offsets, size_t num_sets, uint32_t leaf,
                   size_t bucket, int rotate, rs_mphf **out, rs_build_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!set_offsets || num_sets == 0) return set_error("rs_build_batch: no key sets"), RS_ERR_ARG;
    for (size_t i = 0; i < num_sets; i++)
        if (set_offsets[i + 1] <= set_offsets[i]) return set_error("rs_build_batch: every key set needs a key"), RS_ERR_ARG;
    if (int e = check_args(c, d_keys, set_offsets[num_sets] - set_offsets[0], leaf, bucket, out)) return e;
    return guarded("rs_build_batch", c, [&] {
        std::vector<const uint64_t *> ptrs(num_sets);
        std::vector<uint64_t> ns(num_sets);
        for (size_t i = 0; i < num_sets; i++) {
            ptrs[i] = d_keys + 2 * set_offsets[i];
            ns[i] = set_offsets[i + 1] - set_offsets[i];
        }
        std::vector<rs_mphf *> r = build_sets(c, ptrs, ns, leaf, bucket, rotate & 3, false, t0);
        for (size_t i = 0; i < num_sets; i++) out[i] = r[i];
        if (stats) *stats = r[0]->stats;
    });
}

int rs_query_device(rs_ctx *c, const rs_mphf *h, const uint64_t *d_keys, size_t n, uint64_t *d_out) {
    if (!c || !h || (!d_keys && n) || (!d_out && n)) return set_error("rs_query_device: null argument"), RS_ERR_ARG;
    if (n == 0) return RS_OK;
    return guarded("rs_query_device", c, [&] {
        const DeviceCopy *dc = device_copy_of(c, const_cast<rs_mphf *>(h));
        QueryArgs q{};
        q.rice = dc->rice;
        q.lower = dc->lower;
        q.upper_c = dc->upper_c;
        q.upper_p = dc->upper_p;
        q.jump = dc->jump;
        q.memo = dc->memo;
        q.n_rice = dc->n_rice;
        q.n_lower = dc->n_lower;
        q.n_uc = dc->n_uc;
        q.n_up = dc->n_up;
        q.n_jump = dc->n_jump;
        q.keys_count = h->keys_count;
        q.nbuckets = h->nbuckets;
        q.bpk = h->ef.bits_per_key_fixed_point;
        q.lc = h->ef.l_cum_keys;
        q.lp = h->ef.l_position;
        q.min_delta = h->ef.cum_keys_min_delta;
        q.min_diff = h->ef.min_diff;
        q.P = h->P;
        q.rotate = h->rotate;
        k_query<<<(n + 255) / 256, 256, 0, c->stream>>>(q, d_keys, n, d_out);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"

#include "rs_plan.hip"
