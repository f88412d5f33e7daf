// rs_plan.hip -- build plans: the whole construction of one MPHF of a fixed key
// count, queued on a caller's stream with no device allocation and no host
// synchronisation, so that it can run under stream capture (hipGraph) and be
// replayed for new keys.  Included by rs_engine.hip (one translation unit with
// the kernels).
//
// Where the ordinary build reads sizes back to the host (the bucket-size census
// that sizes the node lists, the descriptor length, the Elias-Fano widths), a plan
// computes them on the device: every buffer is sized at creation from the key
// count (node lists by n, Golomb-Rice and Elias-Fano arrays by bounds checked on
// the device), the census comes from a per-leaf table of node counts by bucket
// size, and the search, split and encoding kernels read their counts from device
// memory.  A build whose descriptors or Elias-Fano arrays would pass the bounds
// stops writing and reports RS_ERR_ARG; the bounds are several times the sizes
// RecSplit reaches.

namespace rsa {

// Sizes the ordinary build reads back, computed on the device by the plan kernels.
struct PlanMeta {
    uint64_t rice_words;                                       // Golomb-Rice vector words
    uint64_t num_buckets, u_cum_keys, u_position;              // Elias-Fano header (DoubleEF.hpp:100-113)
    long long min_delta, min_diff;
    uint64_t bpk;
    uint64_t lower_words, uc_words, up_words, jump_words;      // Elias-Fano vectors
    uint64_t out_words;                                        // serialized words
};

__global__ __launch_bounds__(256) void k_plan_tiles(KeyTile *__restrict__ tiles, const uint64_t *keys, uint64_t n,
                                                    uint64_t nb, uint32_t n_tiles) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const uint64_t k0 = (uint64_t)t * kTileKeys;
    tiles[t] = KeyTile{keys, k0, k0 + kTileKeys < n ? k0 + kTileKeys : n, nb, 0};
}

// Fills p[0, n) with v, n = count ? min(*count + extra, cap) : cap (grid-stride).
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, uint64_t cap, T v, const uint64_t *count = nullptr,
                                              uint64_t extra = 0) {
    uint64_t n = cap;
    if (count) n = *count + extra < cap ? *count + extra : cap;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Node counts per list of the whole build from the bucket-size histogram and the
// per-size table (table[m * kNumLists + l] = nodes of list l in a bucket of m
// keys), their prefix sums (list bases), and the oversized-bucket error.  One
// workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_plan_census(const uint32_t *__restrict__ hist,
                                                     const uint32_t *__restrict__ table, uint64_t *__restrict__ count,
                                                     uint64_t *__restrict__ lbase, uint32_t *__restrict__ err) {
    __shared__ uint64_t s[kNumLists][256];
    uint64_t acc[kNumLists];
#pragma unroll
    for (int l = 0; l < kNumLists; l++) acc[l] = 0;
    for (uint32_t m = threadIdx.x; m < kMaxBucket; m += blockDim.x) {
        const uint64_t h = hist[m];
        if (!h) continue;
#pragma unroll
        for (int l = 0; l < kNumLists; l++) acc[l] += h * table[m * kNumLists + l];
    }
#pragma unroll
    for (int l = 0; l < kNumLists; l++) s[l][threadIdx.x] = acc[l];
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int l = 0; l < kNumLists; l++) s[l][threadIdx.x] += s[l][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint64_t b = 0;
        for (int l = 0; l < kNumLists; l++) {
            count[l] = s[l][0];
            lbase[l] = b;
            b += s[l][0];
        }
        if (hist[kMaxBucket]) atomicOr(err, kErrBucketTooLarge);
    }
}

// Search state of the level's nodes, the waves' node entries and the ticket
// counters (kernels, not memsets, so the plan records no memset node).
__global__ __launch_bounds__(256) void k_level_init(NodeState *__restrict__ st, int *__restrict__ running,
                                                    uint32_t *__restrict__ ctl, uint32_t n_waves,
                                                    const uint64_t *__restrict__ count, uint32_t list) {
    const uint64_t n = count[list];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < n; i += stride) {
        st[i].result = kNone;
        st[i].next_chunk = 1;  // chunk 0 is taken by the wave that starts the node
        st[i].waves = 0;
    }
    for (uint64_t i = i0; i < n_waves; i += stride) running[i] = -1;
    if (i0 < 16) ctl[i0] = 0;
}

// Golomb-Rice vector of the single set: its word count (RiceBitVector.hpp:77-109,
// with the sentinel bit) and the set's encoding record.
__global__ void k_plan_rice(const uint64_t *__restrict__ pos, uint64_t nb, uint64_t cap_words,
                            SetEnc *__restrict__ enc, PlanMeta *__restrict__ meta, uint32_t *__restrict__ err) {
    if (threadIdx.x != 0) return;
    const uint64_t B = pos[nb] - pos[0] + 1;
    const uint64_t words = ((B + 7) / 8 + 14) / 8;
    if (words + 2 > cap_words) atomicOr(err, kErrRiceCap);
    meta->rice_words = words;
    SetEnc e{};
    e.bbase = 0;
    e.nb = nb;
    e.rice_word = 0;
    *enc = e;
}

__device__ __forceinline__ uint64_t ef_width_dev(uint64_t u, uint64_t nb) {
    const uint64_t q = u / (nb + 1);
    return q == 0 ? 0 : 63 - (uint64_t)__clzll(q);
}

// Elias-Fano header and vector sizes (DoubleEF.hpp:140-202) from the device-side
// scalars, as the ordinary build derives them on the host.
__global__ void k_plan_ef(const EfScalars *__restrict__ sc, uint64_t n, uint64_t nb, uint64_t cap_lower,
                          uint64_t cap_upper, SetEnc *__restrict__ enc, PlanMeta *__restrict__ meta,
                          uint32_t *__restrict__ err) {
    if (threadIdx.x != 0) return;
    const EfScalars f = *sc;
    const uint64_t u_position = (uint64_t)((long long)f.total_bits - (long long)(f.bpk * n >> 20) -
                                           (long long)(nb * (uint64_t)f.min_diff) + 1);
    const uint64_t u_cum_keys = n - nb * (uint64_t)f.min_delta + 1;
    const uint64_t lp = ef_width_dev(u_position, nb), lc = ef_width_dev(u_cum_keys, nb);
    const uint64_t lower = ((nb + 1) * (lc + lp) + 63) / 64 + 1;
    const uint64_t uc = (nb + 1 + (u_cum_keys >> lc) + 63) / 64, up = (nb + 1 + (u_position >> lp) + 63) / 64;
    constexpr uint64_t kSuperQ = 1 << 14, kQ = 256;  // DoubleEF.hpp:63-69
    uint64_t jump = (nb / kSuperQ) * kEfSuperSizeDev * 2;
    if (nb % kSuperQ != 0) jump += (1 + ((nb % kSuperQ + kQ - 1) / kQ + 3) / 4) * 2;
    if (lower + 1 > cap_lower || uc + 1 > cap_upper || up + 1 > cap_upper) atomicOr(err, kErrEfCap);
    meta->num_buckets = nb;
    meta->u_cum_keys = u_cum_keys;
    meta->u_position = u_position;
    meta->min_delta = f.min_delta;
    meta->min_diff = f.min_diff;
    meta->bpk = f.bpk;
    meta->lower_words = lower;
    meta->uc_words = uc;
    meta->up_words = up;
    meta->jump_words = jump;
    meta->out_words = 3 + 1 + meta->rice_words + 6 + 4 + lower + uc + up + jump;
    SetEnc e = *enc;
    e.lc = lc;
    e.lp = lp;
    e.lower_word = e.uc_word = e.up_word = e.jump_word = 0;
    e.jump_words = jump;
    *enc = e;
}

// The serialized structure (AbstractParallelRecSplit.hpp:480-488, DoubleEF.hpp:
// 100-113): three header words, the descriptor vector, six Elias-Fano scalars and
// four vectors, each vector a word count then the words.  One workgroup.
__global__ __launch_bounds__(256) void k_plan_serialize(const PlanMeta *__restrict__ meta, uint64_t leaf,
                                                        uint64_t bucket, uint64_t n,
                                                        const unsigned long long *__restrict__ rice,
                                                        const unsigned long long *__restrict__ lower,
                                                        const unsigned long long *__restrict__ uc,
                                                        const unsigned long long *__restrict__ up,
                                                        const uint64_t *__restrict__ jump, uint64_t cap_out,
                                                        uint64_t *__restrict__ out, uint64_t *__restrict__ out_words,
                                                        const uint32_t *__restrict__ err) {
    const PlanMeta m = *meta;
    if ((*err & kErrStopEncode) || m.out_words > cap_out) {
        if (threadIdx.x == 0) *out_words = 0;
        return;
    }
    uint64_t o = 0;
    auto word = [&](uint64_t v) {
        if (threadIdx.x == 0) out[o] = v;
        o++;
    };
    auto vec = [&](const auto *src, uint64_t words) {
        word(words);
        for (uint64_t i = threadIdx.x; i < words; i += blockDim.x) out[o + i] = (uint64_t)src[i];
        o += words;
    };
    word(leaf);
    word(bucket);
    word(n);
    vec(rice, m.rice_words);
    word(m.num_buckets);
    word(m.u_cum_keys);
    word(m.u_position);
    word((uint64_t)m.min_delta);
    word((uint64_t)m.min_diff);
    word(m.bpk);
    vec(lower, m.lower_words);
    vec(uc, m.uc_words);
    vec(up, m.up_words);
    vec(jump, m.jump_words);
    if (threadIdx.x == 0) *out_words = o;
}

}  // namespace rsa

struct rs_plan {
    int device = 0, num_cus = 0;
    uint64_t n = 0, nb = 0, bucket = 0;
    uint32_t leaf = 0;
    int flags = 0;
    uint8_t *mem = nullptr;  // one device allocation holding every buffer below
    size_t mem_bytes = 0;
    Tune tune;
    uint32_t blocks = 0, n_waves = 0, n_tiles = 0;
    std::vector<int> lists;     // tree levels some bucket size can have
    std::vector<char> likely;   // ... and whether a bucket of the likely sizes has it
    std::vector<uint32_t> level_blocks;  // ... and its search grid (expected work)
    std::vector<char> mostly_full;       // ... and whether most of its nodes are of full size
    uint32_t likely_max_bucket = 0;  // bucket sizes above this are vanishingly rare
    // device buffers
    KeyTile *tiles = nullptr;
    uint32_t *counts = nullptr, *size_hist = nullptr, *err = nullptr, *cursor = nullptr, *list_fill = nullptr,
             *ctl = nullptr, *memo = nullptr, *census = nullptr;
    uint64_t *counts64 = nullptr, *acc = nullptr, *node_cnt = nullptr, *node_base = nullptr, *sorted = nullptr,
             *list_base = nullptr, *list_count = nullptr, *seeds = nullptr, *bits = nullptr, *pos = nullptr,
             *scan_tmp = nullptr, *out = nullptr, *out_words = nullptr;
    unsigned long long *work = nullptr, *evals = nullptr, *rice = nullptr, *lower = nullptr, *uc = nullptr,
                       *up = nullptr, *jump = nullptr;
    NodeDesc *pool = nullptr;
    uint16_t *node_m = nullptr;
    NodeState *state = nullptr;
    int *running = nullptr;
    long long *bit_off = nullptr;
    SetEnc *enc = nullptr;
    EfScalars *efs = nullptr;
    PlanMeta *meta = nullptr;
    uint64_t cap_rice = 0, cap_lower = 0, cap_upper = 0, cap_jump = 0, cap_out = 0;
    std::mutex mu;  // one enqueue at a time per plan (the buffers are shared)
};

namespace rsa {

// Node counts per list of one bucket of each size m < kMaxBucket (the census of
// the ordinary build, one bucket at a time), the expected key evaluations of
// those nodes' searches, and how many of them are aggregation nodes of full size.
static std::vector<uint32_t> census_table(const SplitParams &P, bool rotate, std::vector<double> &evals,
                                          std::vector<uint32_t> &full) {
    std::vector<uint32_t> t((size_t)kMaxBucket * kNumLists, 0);
    evals.assign((size_t)kMaxBucket * kNumLists, 0.0);
    full.assign((size_t)kMaxBucket * kNumLists, 0);
    const Memo &memo = golomb_memo(P.leaf);
    std::vector<SubCensus> small(kMaxBucket);
    for (uint32_t m = 2; m < kMaxBucket; m++) {
        Census c;
        census_node(P, memo, rotate, m, 0, 1, c, small);
        for (int l = 0; l < kNumLists; l++) {
            t[(size_t)m * kNumLists + l] = (uint32_t)c.count[l];
            evals[(size_t)m * kNumLists + l] = c.evals[l];
            full[(size_t)m * kNumLists + l] = (uint32_t)c.full[l];
        }
    }
    return t;
}

static void plan_enqueue(rs_plan *p, const uint64_t *keys, hipStream_t s) {
    const SplitParams P = split_params(p->leaf);
    const bool rotate = (p->flags & kFlagRotate) != 0, r64 = (p->flags & kFlagRemap64) != 0;
    const uint64_t NB = p->nb;
    const uint32_t fill_grid = (uint32_t)p->num_cus * 4;
    // ---- partition (the ordinary build's steps 1-3)
    k_plan_tiles<<<(p->n_tiles + 255) / 256, 256, 0, s>>>(p->tiles, keys, p->n, NB, p->n_tiles);
    k_fill<uint32_t><<<fill_grid, 256, 0, s>>>(p->counts, NB, 0u);
    k_fill<uint32_t><<<1, 256, 0, s>>>(p->size_hist, kMaxBucket + 1, 0u);
    k_fill<uint32_t><<<1, 64, 0, s>>>(p->err, 16, 0u);
    k_fill<uint32_t><<<fill_grid, 256, 0, s>>>(p->cursor, NB, 0u);
    k_fill<uint32_t><<<1, 64, 0, s>>>(p->list_fill, kNumLists, 0u);
    k_fill<unsigned long long><<<1, 64, 0, s>>>(p->work, kNumLists, 0ull);
    k_fill<unsigned long long><<<1, 64, 0, s>>>(p->evals, 2 * kNumKinds, 0ull);
    const uint32_t priv_max = (uint32_t)std::min<uint64_t>(NB, 8192);
    k_bucket_hist<<<p->n_tiles, 256, (size_t)priv_max * 4, s>>>(p->tiles, p->counts, priv_max);
    k_bucket_meta<<<(uint32_t)std::min<uint64_t>((NB + 255) / 256, (uint64_t)p->num_cus * 4), 256, 0, s>>>(
        p->counts, NB, p->memo + kMaxBucket, p->counts64, p->node_cnt, p->size_hist);
    scan_level(s, p->counts, true, NB, p->acc, p->scan_tmp);
    scan_level(s, p->node_cnt, false, NB, p->node_base, p->scan_tmp);
    k_plan_census<<<1, 256, 0, s>>>(p->size_hist, p->census, p->list_count, p->list_base, p->err);
    k_scatter<<<p->n_tiles, 256, (size_t)priv_max * 4, s>>>(p->tiles, p->acc, p->cursor, p->sorted, priv_max);
    // buckets up to the likely largest size with a small LDS footprint, larger
    // ones (possible, vanishingly rare) by a second pass with room for 4096 keys
    uint32_t sort_p = 1;
    while (sort_p < p->likely_max_bucket) sort_p <<= 1;
    k_sort_buckets<<<NB, p->likely_max_bucket > 512 ? 256 : 64, (size_t)sort_p * 8, s>>>(
        p->sorted, p->acc, p->err, rotate ? p->leaf : 0, 0, p->likely_max_bucket);
    if (p->likely_max_bucket < kMaxBucket - 1)
        k_sort_buckets<<<NB, 256, (size_t)4096 * 8, s>>>(p->sorted, p->acc, p->err, rotate ? p->leaf : 0,
                                                         p->likely_max_bucket + 1, kMaxBucket - 1);
    ExpandArgs ea{P, rotate ? 1 : 0, p->pool, p->list_base, p->list_fill, p->node_m};
    k_expand<<<(NB + 63) / 64, 64, 0, s>>>(p->acc, p->node_base, NB, ea);
    HIP_CHECK(hipGetLastError());

    // ---- search levels, sizes from the device census
    const AggrLayout up_l = aggr_layout(P.lower, P.upper / P.lower);
    const AggrLayout lo_l = aggr_layout(P.leaf, P.lower / P.leaf);
    const Tune &tn = p->tune;
    for (size_t li = 0; li < p->lists.size(); li++) {
        const int l = p->lists[li];
        // levels only buckets far above the likely largest size reach run on a
        // small grid: correct if such a bucket appears, nearly free when none does
        const bool likely = p->likely[li];
        const uint32_t blocks = likely ? p->level_blocks[li] : std::max<uint32_t>(1, p->blocks / 64);
        const uint32_t kind = l < (int)kMaxDepth ? kHigher : (uint32_t)(l - kMaxDepth + 1);
        SearchArgs a{};
        a.nodes = p->pool;
        a.state = p->state;
        a.running = p->running;
        a.ctl = p->ctl;
        a.leaf = P.leaf;
        a.upper = P.upper;
        a.work = p->work + l;
        a.err = p->err;
        a.groups = (uint32_t)tn.groups[kind];
        a.help_cap = (uint32_t)tn.help_cap[kind];
        a.count_dev = p->list_count;
        a.lbase_dev = p->list_base;
        a.list = (uint32_t)l;
        a.stop = p->err;
        a.stop_mask = kErrStopSearch;
        bool wide = false, lut = false;
        size_t lds = 0;
        if (kind == kUpper || kind == kLower) {
            const AggrLayout &L = kind == kUpper ? up_l : lo_l;
            a.split_unit = kind == kUpper ? P.lower : P.leaf;
            a.field_bits = L.fb;
            a.magic = L.magic;
            wide = L.wide;
            a.lut_entries = kind == kUpper ? P.upper : P.lower;
            lds = (size_t)a.lut_entries * (wide ? 8 : 4);
            lut = tn.lut != 0 && lds <= 32768;
            if (!lut) lds = 0;
        }
        const SearchShape sh =
            search_shape(tn, kind, wide, lut, r64, p->mostly_full[li] != 0, blocks, p->blocks, (uint32_t)p->num_cus);
        a.bt_m = sh.bt ? (kind == kLeafRot ? P.leaf : a.lut_entries) : 0;
        a.n_waves = sh.waves;
        k_level_init<<<likely ? fill_grid : 4, 256, 0, s>>>(p->state, p->running, p->ctl, sh.waves, p->list_count,
                                                            (uint32_t)l);
        launch_search(s, tn, kind, p->sorted, a, wide, lut, sh.grid, lds, r64);
        if (kind == kHigher || kind == kUpper || kind == kLower) {
            const uint32_t max_m = kind == kHigher ? kMaxBucket : (kind == kUpper ? P.upper : P.lower);
            k_split<<<likely ? (uint32_t)p->num_cus * 8 : 8, 64, (size_t)max_m * 8, s>>>(
                p->sorted, p->pool, p->state, p->seeds, p->err, P, kind, r64 ? 1 : 0, p->list_count, p->list_base,
                (uint32_t)l);
        } else {
            k_finalize<<<likely ? fill_grid : 4, 256, 0, s>>>(p->pool, p->state, 0, p->seeds, p->err, p->list_count,
                                                              p->list_base, (uint32_t)l);
        }
        HIP_CHECK(hipGetLastError());
    }

    // ---- encoding (the ordinary build's steps 5-7 for one set)
    const uint32_t *memo = p->memo;
    EncodeArgs ca{p->acc, p->node_base, p->seeds, p->node_m, memo, memo + kMaxBucket, NB, P, rotate ? 1 : 0, p->err};
    k_bucket_bits<<<(uint32_t)std::min<uint64_t>((NB + 3) / 4, (uint64_t)p->num_cus * 8), 256, 0, s>>>(
        ca, p->bits, p->evals);
    scan_level(s, p->bits, false, NB, p->pos, p->scan_tmp);
    k_plan_rice<<<1, 64, 0, s>>>(p->pos, NB, p->cap_rice, p->enc, p->meta, p->err);
    k_fill<unsigned long long><<<fill_grid, 256, 0, s>>>(p->rice, p->cap_rice, 0ull, &p->meta->rice_words, 2);
    k_set_offsets<<<1, 256, 0, s>>>(p->enc, p->pos, p->bit_off, p->rice, p->err);
    k_rice_write<<<(uint32_t)((NB + 3) / 4), 256, 0, s>>>(ca, p->pos, p->bit_off, p->rice);
    k_ef_scalars<<<1, 256, 0, s>>>(p->enc, p->acc, p->pos, p->efs);
    k_plan_ef<<<1, 64, 0, s>>>(p->efs, p->n, NB, p->cap_lower, p->cap_upper, p->enc, p->meta, p->err);
    k_fill<unsigned long long><<<fill_grid, 256, 0, s>>>(p->lower, p->cap_lower, 0ull, &p->meta->lower_words, 1);
    k_fill<unsigned long long><<<fill_grid, 256, 0, s>>>(p->uc, p->cap_upper, 0ull, &p->meta->uc_words, 1);
    k_fill<unsigned long long><<<fill_grid, 256, 0, s>>>(p->up, p->cap_upper, 0ull, &p->meta->up_words, 1);
    k_fill<unsigned long long><<<fill_grid, 256, 0, s>>>(p->jump, p->cap_jump, 0ull, &p->meta->jump_words, 1);
    k_ef_fill<<<1, 256, 0, s>>>(p->enc, p->efs, p->acc, p->pos, p->lower, p->uc, p->up, (uint8_t *)p->jump, p->err,
                                p->err);
    k_plan_serialize<<<1, 256, 0, s>>>(p->meta, p->leaf, p->bucket, p->n, p->rice, p->lower, p->uc, p->up,
                                       (const uint64_t *)p->jump, p->cap_out, p->out, p->out_words, p->err);
    HIP_CHECK(hipGetLastError());
}

}  // namespace rsa

int rs_plan_create(rs_ctx *c, size_t n, uint32_t leaf, size_t bucket, int flags, rs_plan **out) {
    if (!c || !out || n < 1 || leaf < 2 || leaf > kMaxLeaf || bucket < 1 || bucket >= kMaxBucket || (flags & ~3))
        return set_error("rs_plan_create: bad arguments"), RS_ERR_ARG;
    *out = nullptr;
    rs_plan *p = new rs_plan();
    const int rc = guarded("rs_plan_create", c, [&] {
        p->device = c->device;
        p->num_cus = c->num_cus;
        p->n = n;
        p->leaf = leaf;
        p->bucket = bucket;
        p->flags = flags;
        p->nb = std::max<uint64_t>(1, (n + bucket - 1) / bucket);
        p->tune = tune_for(c, leaf);
        p->blocks = (uint32_t)c->num_cus * (uint32_t)p->tune.blocks_per_cu;
        p->n_waves = p->blocks * 4;
        p->n_tiles = (uint32_t)((n + kTileKeys - 1) / kTileKeys);
        const uint64_t NB = p->nb;
        const SplitParams P = split_params(leaf);
        (void)P;
        // bounds: nodes < n (a tree over m keys has at most m - 1 nodes); descriptors
        // at most 8 bits per key plus a word per bucket (RecSplit stores ~1.6);
        // Elias-Fano lower bits at most 128 per bucket, upper bits at most 3 per bucket
        p->cap_rice = n / 8 + NB + 16;
        p->cap_lower = 2 * (NB + 1) + 4;
        p->cap_upper = 3 * (NB + 1) / 64 + 4;
        p->cap_jump = ef_jump_words(NB) + 2;
        p->cap_out = 3 + 1 + p->cap_rice + 6 + 4 + p->cap_lower + 2 * p->cap_upper + p->cap_jump;
        // layout: one allocation, carved in a fixed order
        Arena a;
        auto plan_layout = [&](Arena &ar) {
            p->tiles = ar.take<KeyTile>(p->n_tiles);
            p->counts = ar.take<uint32_t>(NB);
            p->counts64 = ar.take<uint64_t>(NB);
            p->acc = ar.take<uint64_t>(NB + 1);
            p->node_cnt = ar.take<uint64_t>(NB);
            p->node_base = ar.take<uint64_t>(NB + 1);
            p->size_hist = ar.take<uint32_t>(kMaxBucket + 1);
            p->err = ar.take<uint32_t>(16);
            p->cursor = ar.take<uint32_t>(NB);
            p->sorted = ar.take<uint64_t>(n + 8);
            p->pool = ar.take<NodeDesc>(n + 1);
            p->list_base = ar.take<uint64_t>(kNumLists);
            p->list_count = ar.take<uint64_t>(kNumLists);
            p->list_fill = ar.take<uint32_t>(kNumLists);
            p->node_m = ar.take<uint16_t>(n + 1);
            p->seeds = ar.take<uint64_t>(n + 1);
            p->state = ar.take<NodeState>(n + 1);
            p->running = ar.take<int>(p->n_waves);
            p->ctl = ar.take<uint32_t>(16);
            p->work = ar.take<unsigned long long>(kNumLists);
            p->evals = ar.take<unsigned long long>(2 * kNumKinds);
            p->bits = ar.take<uint64_t>(NB);
            p->pos = ar.take<uint64_t>(NB + 1);
            p->scan_tmp = ar.take<uint64_t>(scan_tmp_words(NB));
            p->bit_off = ar.take<long long>(NB);
            p->rice = ar.take<unsigned long long>(p->cap_rice);
            p->enc = ar.take<SetEnc>(1);
            p->efs = ar.take<EfScalars>(1);
            p->meta = ar.take<PlanMeta>(1);
            p->lower = ar.take<unsigned long long>(p->cap_lower);
            p->uc = ar.take<unsigned long long>(p->cap_upper);
            p->up = ar.take<unsigned long long>(p->cap_upper);
            p->jump = ar.take<unsigned long long>(p->cap_jump);
            p->out = ar.take<uint64_t>(p->cap_out);
            p->out_words = ar.take<uint64_t>(1);
            p->memo = ar.take<uint32_t>(2 * kMaxBucket);
            p->census = ar.take<uint32_t>((size_t)kMaxBucket * kNumLists);
        };
        a.size = ~size_t(0) >> 1;
        plan_layout(a);  // measure
        p->mem_bytes = a.used;
        HIP_CHECK(hipMalloc(&p->mem, p->mem_bytes));
        Arena real;
        real.base = p->mem;
        real.size = p->mem_bytes;
        plan_layout(real);
        // tables, and the tree levels some bucket size can reach
        std::vector<double> evals;
        std::vector<uint32_t> full;
        const std::vector<uint32_t> table = census_table(split_params(leaf), (flags & kFlagRotate) != 0, evals, full);
        HIP_CHECK(hipMemcpy(p->memo, golomb_memo(leaf).data(), sizeof(uint32_t) * kMaxBucket, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(p->memo + kMaxBucket, tree_nodes(leaf).data(), sizeof(uint32_t) * kMaxBucket,
                            hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(p->census, table.data(), table.size() * 4, hipMemcpyHostToDevice));
        // bucket sizes are Poisson(n / nb) for uniform keys: 10 standard deviations
        // above the mean (plus a margin) bounds every bucket of every practical build
        const double mean = (double)n / (double)NB;
        p->likely_max_bucket = (uint32_t)std::min<double>(kMaxBucket - 1, mean + 10.0 * std::sqrt(mean) + 32);
        // expected evaluations per list over Poisson bucket sizes size each
        // level's grid as the ordinary build does from its census
        std::vector<double> pmf(kMaxBucket, 0.0);
        for (uint32_t m = 0; m < kMaxBucket; m++)
            pmf[m] = std::exp((double)m * std::log(mean) - mean - std::lgamma((double)m + 1.0));
        for (int l = 0; l < kNumLists; l++) {
            bool any = false, lik = false;
            double e = 0, cnt = 0, nfull = 0;
            for (uint32_t m = 2; m < kMaxBucket; m++)
                if (table[(size_t)m * kNumLists + l]) {
                    any = true;
                    lik |= m <= p->likely_max_bucket;
                    e += pmf[m] * evals[(size_t)m * kNumLists + l];
                    cnt += pmf[m] * table[(size_t)m * kNumLists + l];
                    nfull += pmf[m] * full[(size_t)m * kNumLists + l];
                }
            if (any) {
                p->lists.push_back(l);
                p->likely.push_back(lik ? 1 : 0);
                p->level_blocks.push_back(rsa::level_blocks(e * (double)NB, cnt * (double)NB, p->blocks));
                p->mostly_full.push_back(2 * nfull >= cnt ? 1 : 0);
            }
        }
    });
    if (rc != RS_OK) {
        if (p->mem) (void)hipFree(p->mem);
        delete p;
        return rc;
    }
    *out = p;
    return RS_OK;
}

int rs_plan_enqueue(rs_plan *p, const uint64_t *d_keys, void *stream) {
    if (!p || !d_keys) return set_error("rs_plan_enqueue: null argument"), RS_ERR_ARG;
    try {
        std::lock_guard<std::mutex> g(p->mu);
        int prev = 0;
        HIP_CHECK(hipGetDevice(&prev));
        if (prev != p->device) HIP_CHECK(hipSetDevice(p->device));
        plan_enqueue(p, d_keys, (hipStream_t)stream);
        if (prev != p->device) HIP_CHECK(hipSetDevice(prev));
        return RS_OK;
    } catch (const BuildError &e) {
        set_error(std::string("rs_plan_enqueue: ") + e.what());
        return e.code;
    } catch (const HipError &e) {
        set_error(std::string("rs_plan_enqueue: ") + e.what());
        return RS_ERR_HIP;
    }
}

int rs_plan_result(rs_plan *p, uint8_t *out, size_t cap, size_t *len) {
    if (!p || !len) return set_error("rs_plan_result: null argument"), RS_ERR_ARG;
    try {
        std::lock_guard<std::mutex> g(p->mu);
        int prev = 0;
        HIP_CHECK(hipGetDevice(&prev));
        if (prev != p->device) HIP_CHECK(hipSetDevice(p->device));
        uint32_t err = 0;
        uint64_t words = 0;
        HIP_CHECK(hipMemcpy(&err, p->err, 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(&words, p->out_words, 8, hipMemcpyDeviceToHost));
        int rc = RS_OK;
        auto fail = [&](int code, const char *msg) {
            set_error(std::string("rs_plan_result: ") + msg);
            rc = code;
        };
        if (err & kErrDuplicateKeys) fail(RS_ERR_DUPLICATE_KEYS, "two keys of one bucket share `second`");
        else if (err & kErrBucketTooLarge) fail(RS_ERR_BUCKET_TOO_LARGE, "a bucket holds 3000 or more keys");
        else if (err & kErrSeedLimit) fail(RS_ERR_SEED_OVERFLOW, "a node search passed the 2^44 seed limit");
        else if (err & kErrEfJump) fail(RS_ERR_EF_OVERFLOW, "Elias-Fano jump offset does not fit 16 bits");
        else if (err & (kErrRiceCap | kErrEfCap))
            fail(RS_ERR_ARG, "the structure passed the plan's size bounds; build it with rs_build_device");
        *len = rc == RS_OK ? words * 8 : 0;
        if (rc == RS_OK && out) {
            if (cap < *len) fail(RS_ERR_ARG, "output buffer too small");
            else HIP_CHECK(hipMemcpy(out, p->out, *len, hipMemcpyDeviceToHost));
        }
        if (prev != p->device) HIP_CHECK(hipSetDevice(prev));
        return rc;
    } catch (const HipError &e) {
        set_error(std::string("rs_plan_result: ") + e.what());
        return RS_ERR_HIP;
    }
}

size_t rs_plan_device_bytes(const rs_plan *p) { return p ? p->mem_bytes : 0; }

void rs_plan_destroy(rs_plan *p) {
    if (!p) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    (void)hipDeviceSynchronize();
    if (p->mem) (void)hipFree(p->mem);
    (void)hipSetDevice(prev);
    delete p;
}
