// rs_host.cpp -- host side of the engine: Golomb-Rice memo, split strategy,
// serialization, and the host query path (Elias-Fano lookup + Rice reader).
#include <cmath>
#include <cstring>
#include <mutex>

#include "rs_internal.h"
#include "rs_spooky.h"

namespace rsa {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

// Leaf bijection Golomb-Rice parameters (AbstractParallelRecSplit.hpp:106).
static const uint8_t kBijMemo[25] = {0, 0, 0, 1, 3, 4, 5, 7, 8, 10, 11, 12, 14, 15, 16, 18, 19, 21, 22, 23, 25, 26, 28, 29, 30};

SplitParams split_params(uint32_t leaf) {
    // AbstractParallelRecSplit.hpp:133-135
    const double l = leaf;
    const uint64_t lo = (uint64_t)leaf * std::max<uint64_t>(2, (uint64_t)std::ceil(0.35 * l + 1. / 2));
    const uint64_t up = lo * (leaf < 7 ? 2 : (uint64_t)std::ceil(0.21 * l + 9. / 10));
    return SplitParams{leaf, (uint32_t)lo, (uint32_t)up};
}

// Child sizes of a node (AbstractParallelRecSplit.hpp:137-148).
static uint32_t children(const SplitParams &P, uint32_t m, uint32_t *out) {
    uint32_t unit, fan;
    if (m > P.upper) {
        unit = higher_split(P, m);
        fan = 2;
    } else if (m > P.lower) {
        unit = P.lower;
        fan = (uint16_t)(m + P.lower - 1) / P.lower;
    } else {
        unit = P.leaf;
        fan = (uint16_t)(m + P.leaf - 1) / P.leaf;
    }
    for (uint32_t i = 0; i + 1 < fan; i++) out[i] = unit;
    out[fan - 1] = m - (fan - 1) * unit;
    return fan;
}

static Memo build_memo(uint32_t leaf) {
    // AbstractParallelRecSplit.hpp:191-227: golomb parameter from the success
    // probability of the split, plus subtree node counts and fixed-bit sums.
    const SplitParams P = split_params(leaf);
    Memo memo{};
    uint32_t s = 0;
    for (; s <= leaf; ++s) memo[s] = (uint32_t)kBijMemo[s] << 27 | (uint32_t)(s > 1) << 16 | kBijMemo[s];
    for (; s < kMaxBucket; ++s) {
        uint32_t k[kMaxFanout];
        const uint32_t fan = children(P, s, k);
        double sqrt_prod = 1;
        for (uint32_t i = 0; i < fan; ++i) sqrt_prod *= std::sqrt((double)k[i]);
        const double p = std::sqrt((double)s) / (std::pow(2 * M_PI, (fan - 1.) / 2) * sqrt_prod);
        uint32_t len = (uint32_t)std::ceil(std::log2(-std::log((std::sqrt(5) + 1) / 2) / std::log1p(-p)));
        memo[s] = len << 27;
        for (uint32_t i = 0; i < fan; ++i) len += memo[k[i]] & 0xFFFF;
        memo[s] |= len;
        uint32_t nodes = 1;
        for (uint32_t i = 0; i < fan; ++i) nodes += (memo[k[i]] >> 16) & 0x7FF;
        memo[s] |= nodes << 16;
    }
    return memo;
}

const Memo &golomb_memo(uint32_t leaf) {
    static std::mutex mu;
    static Memo *tables[kMaxLeaf + 1] = {};
    std::lock_guard<std::mutex> g(mu);
    if (!tables[leaf]) tables[leaf] = new Memo(build_memo(leaf));
    return *tables[leaf];
}

static Memo build_tree_nodes(uint32_t leaf) {
    const SplitParams P = split_params(leaf);
    Memo nodes{};
    for (uint32_t s = 2; s < kMaxBucket; ++s) {
        if (s <= leaf) {
            nodes[s] = 1;
            continue;
        }
        uint32_t k[kMaxFanout];
        const uint32_t fan = children(P, s, k);
        uint32_t t = 1;
        for (uint32_t i = 0; i < fan; ++i) t += nodes[k[i]];
        nodes[s] = t;
    }
    return nodes;
}

const Memo &tree_nodes(uint32_t leaf) {
    static std::mutex mu;
    static Memo *tables[kMaxLeaf + 1] = {};
    std::lock_guard<std::mutex> g(mu);
    if (!tables[leaf]) tables[leaf] = new Memo(build_tree_nodes(leaf));
    return *tables[leaf];
}

// ---------------------------------------------------------------- Elias-Fano

void EliasFano::derive_widths() {
    l_position = ef_width(u_position, num_buckets);
    l_cum_keys = ef_width(u_cum_keys, num_buckets);
}

static inline uint64_t select_in_word(uint64_t x, uint64_t k) {
    for (uint64_t i = 0; i < k; i++) x &= x - 1;
    return (uint64_t)__builtin_ctzll(x);
}

static inline uint64_t load_u64_at_byte(const std::vector<uint64_t> &v, uint64_t byte) {
    uint64_t t = 0;
    const uint64_t avail = v.size() * 8;
    if (byte < avail) memcpy(&t, (const uint8_t *)v.data() + byte, std::min<uint64_t>(8, avail - byte));
    return t;
}

// DoubleEF.hpp:367-402.  Word reads past the end of an array (possible only for
// a malformed deserialized structure) read all-ones words, so every scan ends.
static inline uint64_t word_at(const std::vector<uint64_t> &v, uint64_t i) { return i < v.size() ? v[i] : ~uint64_t(0); }

void EliasFano::get(uint64_t i, uint64_t &cum_keys, uint64_t &cum_keys_next, uint64_t &position) const {
    const uint64_t lc = l_cum_keys, lp = l_position;
    const uint64_t mask_c = (uint64_t(1) << lc) - 1, mask_p = (uint64_t(1) << lp) - 1;
    const uint64_t pos_lower = i * (lc + lp);
    uint64_t lower = load_u64_at_byte(lower_bits, pos_lower / 8) >> (pos_lower % 8);
    const uint64_t jsq = (i / kEfSuperQ) * kEfSuperSize * 2;
    const uint64_t jin = (i % kEfSuperQ) / kEfQ;
    const uint64_t pair = load_u64_at_byte(jump, (jsq + 2) * 8 + 4 * jin);
    const uint64_t jump_c = word_at(jump, jsq) + (pair & 0xFFFF);
    const uint64_t jump_p = word_at(jump, jsq + 1) + ((pair >> 16) & 0xFFFF);
    uint64_t wc = jump_c / 64, wp = jump_p / 64;
    uint64_t win_c = word_at(upper_cum, wc) & (~uint64_t(0) << (jump_c % 64));
    uint64_t win_p = word_at(upper_pos, wp) & (~uint64_t(0) << (jump_p % 64));
    uint64_t dc = i & kEfQMask, dp = i & kEfQMask;
    for (uint64_t c; (c = (uint64_t)__builtin_popcountll(win_c)) <= dc; dc -= c) win_c = word_at(upper_cum, ++wc);
    for (uint64_t c; (c = (uint64_t)__builtin_popcountll(win_p)) <= dp; dp -= c) win_p = word_at(upper_pos, ++wp);
    const uint64_t sel_c = select_in_word(win_c, dc);
    const uint64_t cum_delta = i * (uint64_t)cum_keys_min_delta;
    cum_keys = ((wc * 64 + sel_c - i) << lc | (lower & mask_c)) + cum_delta;
    lower >>= lc;
    const uint64_t bit_delta = i * (uint64_t)min_diff;
    position = ((wp * 64 + select_in_word(win_p, dp) - i) << lp | (lower & mask_p)) + bit_delta +
               (uint64_t)(int64_t)(bits_per_key_fixed_point * cum_keys >> 20);
    win_c &= (~uint64_t(0) << sel_c) << 1;
    while (win_c == 0) win_c = word_at(upper_cum, ++wc);
    lower >>= lp;
    cum_keys_next = ((wc * 64 + (uint64_t)__builtin_ctzll(win_c) - i - 1) << lc | (lower & mask_c)) + cum_delta +
                    (uint64_t)cum_keys_min_delta;
}

// DoubleEF.hpp:435-437
uint64_t EliasFano::bit_count_cum_keys() const {
    return (num_buckets + 1) * l_cum_keys + num_buckets + 1 + (u_cum_keys >> l_cum_keys) + ef_jump_words(num_buckets) / 2;
}
uint64_t EliasFano::bit_count_position() const {
    return (num_buckets + 1) * l_position + num_buckets + 1 + (u_position >> l_position) + ef_jump_words(num_buckets) / 2;
}

// ---------------------------------------------------------------- Rice reader

// Golomb-Rice descriptor reader (RiceBitVector.hpp:155-218): fixed parts are
// read at curr_fixed, unary parts from a 64-bit window advancing over the words.
// Reads past the last word (a malformed deserialized structure only) see
// all-ones words, so every scan ends inside the buffer.
struct RiceReader {
    const uint64_t *data;
    uint64_t words;  // readable words
    uint64_t fixed_off = 0;
    uint64_t window = 0;
    uint64_t uidx = 0;  // next unary word
    int valid = 0;

    RiceReader(const uint64_t *d, uint64_t n) : data(d), words(n) {}

    uint64_t word(uint64_t i) const { return i < words ? data[i] : ~uint64_t(0); }

    void reset(uint64_t bit_pos, uint64_t unary_offset) {
        fixed_off = bit_pos;
        const uint64_t up = bit_pos + unary_offset;
        uidx = up / 64;
        window = word(uidx++) >> (up & 63);
        valid = 64 - (int)(up & 63);
    }

    uint64_t next(int log2golomb) {
        uint64_t result = 0;
        if (window == 0) {
            result += (uint64_t)valid;
            window = word(uidx++);
            valid = 64;
            while (window == 0) {
                result += 64;
                window = word(uidx++);
            }
        }
        const uint64_t pos = (uint64_t)__builtin_ctzll(window);
        window >>= pos;
        window >>= 1;
        valid -= (int)pos + 1;
        result += pos;
        result <<= log2golomb;
        const uint64_t w = fixed_off / 64, sh = fixed_off % 64;
        const uint64_t fixed = sh ? (word(w) >> sh) | (word(w + 1) << (64 - sh)) : word(w);
        result |= fixed & ((uint64_t(1) << log2golomb) - 1);
        fixed_off += (uint64_t)log2golomb;
        return result;
    }

    void skip_subtree(uint64_t nodes, uint64_t fixed_len) {
        uint64_t missing = nodes, cnt;
        while ((cnt = (uint64_t)__builtin_popcountll(window)) < missing) {
            window = word(uidx++);
            missing -= cnt;
            valid = 64;
        }
        cnt = select_in_word(window, missing - 1);
        window >>= cnt;
        window >>= 1;
        valid -= (int)cnt + 1;
        fixed_off += fixed_len;
    }
};

static const uint64_t kStartSeed[20] = RS_START_SEEDS;

static inline uint64_t remap128(uint64_t x, uint64_t n) { return (uint64_t)(((unsigned __int128)x * n) >> 64); }

// operator(), AbstractParallelRecSplit.hpp:409-466
uint64_t query(const rs_mphf &h, uint64_t first, uint64_t second) {
    const Memo &memo = *h.memo;
    const SplitParams &P = h.P;
    const bool r64 = (h.rotate & kFlagRemap64) != 0;
    const uint64_t bucket = remap128(first, h.nbuckets);
    uint64_t cum_keys, cum_keys_next, bit_pos;
    h.ef.get(bucket, cum_keys, cum_keys_next, bit_pos);
    uint64_t m = cum_keys_next - cum_keys;
    if (m >= kMaxBucket) return h.keys_count;  // malformed structure: no valid answer
    RiceReader rd(h.rice.data(), h.rice.size());
    rd.reset(bit_pos, skip_bits(memo, m));
    int level = 0;
    while (m > P.upper) {
        const uint64_t d = rd.next((int)golomb_param(memo, m));
        const uint32_t hmod = remap_any(second + d + kStartSeed[level], (uint32_t)m, r64);
        const uint32_t split = higher_split(P, (uint32_t)m);
        if (hmod < split) {
            m = split;
        } else {
            rd.skip_subtree(skip_nodes(memo, split), skip_bits(memo, split));
            m -= split;
            cum_keys += split;
        }
        level++;
    }
    if (m > P.lower) {
        const uint64_t d = rd.next((int)golomb_param(memo, m));
        const uint32_t hmod = remap_any(second + d + kStartSeed[kSeedUpper], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / P.lower;
        m = std::min<uint64_t>(P.lower, m - part * P.lower);
        cum_keys += P.lower * part;
        if (part) rd.skip_subtree(skip_nodes(memo, P.lower) * part, skip_bits(memo, P.lower) * part);
    }
    if (m > P.leaf) {
        const uint64_t d = rd.next((int)golomb_param(memo, m));
        const uint32_t hmod = remap_any(second + d + kStartSeed[kSeedLower], (uint32_t)m, r64);
        const uint64_t part = (uint16_t)hmod / P.leaf;
        m = std::min<uint64_t>(P.leaf, m - part * P.leaf);
        cum_keys += (uint64_t)P.leaf * part;
        if (part) rd.skip_subtree(part, skip_bits(memo, P.leaf) * part);
    }
    const uint64_t b = rd.next((int)golomb_param(memo, m));
    const uint64_t x = second + b + kStartSeed[kSeedLeaf];
    uint64_t leaf_val;
    if ((h.rotate & kFlagRotate) && m == P.leaf) {
        const uint64_t rotation = b % P.leaf;
        leaf_val = remap_any(x - rotation, P.leaf, r64);
        if (second % 2 != 0) leaf_val = (leaf_val + rotation) % P.leaf;
    } else {
        leaf_val = remap_any(x, (uint32_t)m, r64);
    }
    return cum_keys + leaf_val;
}

// ---------------------------------------------------------------- serialization

// Stream layout: AbstractParallelRecSplit.hpp:480-488 (three size_t), the Rice
// vector (RiceBitVector.hpp:139-142), then DoubleEF.hpp:100-113 (six 8-byte
// scalars, four vectors).  A vector is its 8-byte element count followed by the
// raw little-endian words.
static size_t vec_bytes(uint64_t words) { return 8 + 8 * words; }

size_t serialized_size(const rs_mphf &h) {
    return 24 + vec_bytes(h.rice_words) + 48 + vec_bytes(h.ef.lower_bits.size()) + vec_bytes(h.ef.upper_cum.size()) +
           vec_bytes(h.ef.upper_pos.size()) + vec_bytes(h.ef.jump.size());
}

namespace {
struct Writer {
    uint8_t *p;
    void u64(uint64_t v) {
        memcpy(p, &v, 8);
        p += 8;
    }
    void vec(const uint64_t *w, uint64_t n) {
        u64(n);
        if (n) memcpy(p, w, 8 * n);
        p += 8 * n;
    }
};
struct Reader {
    const uint8_t *p, *end;
    bool ok = true;
    uint64_t u64() {
        uint64_t v = 0;
        if (end - p < 8) {
            ok = false;
            return 0;
        }
        memcpy(&v, p, 8);
        p += 8;
        return v;
    }
    void vec(std::vector<uint64_t> &w, uint64_t pad) {
        const uint64_t n = u64();
        if (!ok || n > (uint64_t)(end - p) / 8) {
            ok = false;
            return;
        }
        w.assign(n + pad, 0);
        if (n) memcpy(w.data(), p, 8 * n);
        w.resize(n + pad);
        p += 8 * n;
    }
};
}  // namespace

void serialize(const rs_mphf &h, uint8_t *out) {
    Writer w{out};
    w.u64(h.leaf);
    w.u64(h.bucket_size);
    w.u64(h.keys_count);
    w.vec(h.rice.data(), h.rice_words);
    w.u64(h.ef.num_buckets);
    w.u64(h.ef.u_cum_keys);
    w.u64(h.ef.u_position);
    w.u64((uint64_t)h.ef.cum_keys_min_delta);
    w.u64((uint64_t)h.ef.min_diff);
    w.u64(h.ef.bits_per_key_fixed_point);
    w.vec(h.ef.lower_bits.data(), h.ef.lower_bits.size());
    w.vec(h.ef.upper_cum.data(), h.ef.upper_cum.size());
    w.vec(h.ef.upper_pos.data(), h.ef.upper_pos.size());
    w.vec(h.ef.jump.data(), h.ef.jump.size());
}

}  // namespace rsa

using namespace rsa;

extern "C" {

const char *rs_last_error(void) { return g_last_error.c_str(); }

int rs_golomb_memo(uint32_t leaf_size, uint32_t *out3000) {
    if (leaf_size < 2 || leaf_size > kMaxLeaf || !out3000) {
        set_error("rs_golomb_memo: leaf size must be in [2, 24]");
        return RS_ERR_ARG;
    }
    const Memo &m = golomb_memo(leaf_size);
    memcpy(out3000, m.data(), sizeof(uint32_t) * kMaxBucket);
    return RS_OK;
}

size_t rs_mphf_serialized_size(const rs_mphf *h) { return h ? serialized_size(*h) : 0; }

int rs_mphf_serialize(const rs_mphf *h, uint8_t *buf, size_t cap) {
    if (!h || !buf || cap < serialized_size(*h)) {
        set_error("rs_mphf_serialize: null handle or buffer too small");
        return RS_ERR_ARG;
    }
    serialize(*h, buf);
    return RS_OK;
}

// operator>> (GPURecSplit.cuh:827-841, DoubleEF.hpp:115-135)
int rs_mphf_deserialize(const uint8_t *buf, size_t len, uint32_t leaf_size, int rotate, rs_mphf **out) {
    if (!buf || !out || leaf_size < 2 || leaf_size > kMaxLeaf) {
        set_error("rs_mphf_deserialize: bad arguments");
        return RS_ERR_ARG;
    }
    Reader r{buf, buf + len};
    const uint64_t leaf = r.u64();
    if (!r.ok || leaf != leaf_size) {
        set_error("rs_mphf_deserialize: serialized leaf size " + std::to_string(leaf) + ", expected " +
                  std::to_string(leaf_size));
        return RS_ERR_FORMAT;
    }
    rs_mphf *h = new rs_mphf();
    h->leaf = leaf_size;
    h->rotate = rotate & (kFlagRotate | kFlagRemap64);
    h->bucket_size = r.u64();
    h->keys_count = r.u64();
    r.vec(h->rice, 2);
    h->rice_words = h->rice.size() >= 2 ? h->rice.size() - 2 : 0;
    h->ef.num_buckets = r.u64();
    h->ef.u_cum_keys = r.u64();
    h->ef.u_position = r.u64();
    h->ef.cum_keys_min_delta = (int64_t)r.u64();
    h->ef.min_diff = (int64_t)r.u64();
    h->ef.bits_per_key_fixed_point = r.u64();
    r.vec(h->ef.lower_bits, 0);
    r.vec(h->ef.upper_cum, 0);
    r.vec(h->ef.upper_pos, 0);
    r.vec(h->ef.jump, 0);
    std::string bad;
    if (!r.ok) bad = "truncated stream";
    else if (h->bucket_size == 0 || h->keys_count == 0) bad = "zero bucket size or key count";
    else if (h->rice_words == 0) bad = "empty descriptor vector";
    if (bad.empty()) {
        h->nbuckets = std::max<uint64_t>(1, (h->keys_count + h->bucket_size - 1) / h->bucket_size);
        h->ef.derive_widths();
        const EliasFano &ef = h->ef;
        const uint64_t nb = h->nbuckets;
        if (ef.num_buckets != nb)
            bad = "Elias-Fano bucket count " + std::to_string(ef.num_buckets) + " != ceil(keys / bucket size) " +
                  std::to_string(nb);
        else if (ef.l_cum_keys + ef.l_position > 56)
            bad = "Elias-Fano lower-bit widths out of range";
        else if (ef.lower_bits.size() != ef_lower_words(nb, ef.l_cum_keys, ef.l_position) ||
                 ef.upper_cum.size() != ef_upper_words(nb, ef.u_cum_keys, ef.l_cum_keys) ||
                 ef.upper_pos.size() != ef_upper_words(nb, ef.u_position, ef.l_position) ||
                 ef.jump.size() != ef_jump_words(nb))
            bad = "Elias-Fano array sizes do not match its parameters";
    }
    if (bad.empty()) {
        // first and last bucket: offsets start at 0, keys add up, and the last
        // bucket's descriptor starts inside the descriptor vector
        uint64_t c0, c1, p0, cl, cn, pl;
        h->ef.get(0, c0, c1, p0);
        h->ef.get(h->nbuckets - 1, cl, cn, pl);
        if (c0 != 0 || p0 != 0 || cn != h->keys_count || cl > cn || pl >= h->rice_words * 64)
            bad = "Elias-Fano offsets inconsistent with the key count or descriptor length";
    }
    if (!bad.empty()) {
        delete h;
        set_error("rs_mphf_deserialize: malformed stream: " + bad);
        return RS_ERR_FORMAT;
    }
    h->memo = &golomb_memo(leaf_size);
    h->P = split_params(leaf_size);
    *out = h;
    return RS_OK;
}

uint64_t rs_mphf_query(const rs_mphf *h, uint64_t first, uint64_t second) { return query(*h, first, second); }

void rs_spooky128(const void *data, size_t len, uint64_t *hash1, uint64_t *hash2) {
    spooky::hash128((const uint8_t *)data, len, *hash1, *hash2);
}

void rs_first_hash(const void *data, size_t len, uint64_t *out2) {
    spooky::first_hash((const uint8_t *)data, len, out2[0], out2[1]);
}

void rs_hash_strings(const uint8_t *bytes, const uint64_t *offsets, size_t n, uint64_t *out) {
    for (size_t i = 0; i < n; i++)
        spooky::first_hash(bytes + offsets[i], offsets[i + 1] - offsets[i], out[2 * i], out[2 * i + 1]);
}

uint64_t rs_mphf_query_string(const rs_mphf *h, const void *data, size_t len) {
    uint64_t first, second;
    spooky::first_hash((const uint8_t *)data, len, first, second);
    return query(*h, first, second);
}

void rs_mphf_query_batch(const rs_mphf *h, const uint64_t *keys, size_t n, uint64_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = query(*h, keys[2 * i], keys[2 * i + 1]);
}

uint64_t rs_mphf_bits(const rs_mphf *h) {
    return h->ef.bit_count_cum_keys() + h->ef.bit_count_position() + h->rice_words * 64;
}

uint64_t rs_mphf_num_keys(const rs_mphf *h) { return h ? h->keys_count : 0; }

int rs_mphf_stats(const rs_mphf *h, rs_build_stats *out) {
    if (!h || !out) return RS_ERR_ARG;
    *out = h->stats;
    return RS_OK;
}

void rs_mphf_free(rs_mphf *h) {
    if (!h) return;
    if (h->dev) device_copy_free(h->dev);
    delete h;
}

}  // extern "C"
