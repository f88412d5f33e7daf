// rs_internal.h -- host-side structures of the engine (not part of the C ABI).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gpurecsplit_amd.h"
#include "rs_common.h"

namespace rsa {

using Memo = std::array<uint32_t, kMaxBucket>;

// Golomb-Rice memo of a leaf size (formula of AbstractParallelRecSplit.hpp:191-227).
const Memo &golomb_memo(uint32_t leaf);
SplitParams split_params(uint32_t leaf);

// Nodes of the split tree of a bucket of m keys (0 for m < 2).  The memo's 11-bit
// node field wraps for leaf size 2 and buckets of 2049 or more keys (the reference
// skips its assert for leaf sizes below 3, AbstractParallelRecSplit.hpp:217); the
// build sizes its per-bucket seed slots from this unmasked table instead.
const Memo &tree_nodes(uint32_t leaf);

inline uint32_t golomb_param(const Memo &memo, uint64_t m) { return memo[m] >> 27; }
inline uint32_t skip_bits(const Memo &memo, uint64_t m) { return memo[m] & 0xFFFF; }
inline uint32_t skip_nodes(const Memo &memo, uint64_t m) { return (memo[m] >> 16) & 0x7FF; }

// Double Elias-Fano fields (DoubleEF.hpp:63-76).
struct EliasFano {
    uint64_t num_buckets = 0, u_cum_keys = 0, u_position = 0;
    int64_t cum_keys_min_delta = 0, min_diff = 0;
    uint64_t bits_per_key_fixed_point = 0;
    uint64_t l_cum_keys = 0, l_position = 0;
    std::vector<uint64_t> lower_bits, upper_cum, upper_pos, jump;

    void derive_widths();
    void get(uint64_t i, uint64_t &cum_keys, uint64_t &cum_keys_next, uint64_t &position) const;
    uint64_t bit_count_cum_keys() const;
    uint64_t bit_count_position() const;
};

// EF geometry (DoubleEF.hpp:63-69, 88-98).
constexpr uint64_t kEfQ = 256, kEfQMask = kEfQ - 1, kEfSuperQ = 1 << 14, kEfSuperQMask = kEfSuperQ - 1;
constexpr uint64_t kEfSuperSize = 1 + (kEfSuperQ / kEfQ) / 4;  // u64 words per super block half
inline uint64_t ef_lower_words(uint64_t nb, uint64_t lc, uint64_t lp) { return ((nb + 1) * (lc + lp) + 63) / 64 + 1; }
inline uint64_t ef_upper_words(uint64_t nb, uint64_t u, uint64_t l) { return (nb + 1 + (u >> l) + 63) / 64; }
inline uint64_t ef_jump_words(uint64_t nb) {
    uint64_t size = (nb / kEfSuperQ) * kEfSuperSize * 2;
    if (nb % kEfSuperQ != 0) size += (1 + ((nb % kEfSuperQ + kEfQ - 1) / kEfQ + 3) / 4) * 2;
    return size;
}
inline uint64_t lambda64(uint64_t w) { return 63 ^ (uint64_t)__builtin_clzll(w); }
inline uint64_t ef_width(uint64_t u, uint64_t nb) { return u / (nb + 1) == 0 ? 0 : lambda64(u / (nb + 1)); }

struct DeviceCopy;  // device-resident query structures (rs_engine.hip)

}  // namespace rsa

struct rs_mphf {
    uint32_t leaf = 0;
    int rotate = 1;  // flags: kFlagRotate, kFlagRemap64
    uint64_t bucket_size = 0, keys_count = 0, nbuckets = 0;
    std::vector<uint64_t> rice;  // Golomb-Rice descriptor words (with read padding beyond size)
    uint64_t rice_words = 0;     // serialized word count
    rsa::EliasFano ef;
    const rsa::Memo *memo = nullptr;
    rsa::SplitParams P{};
    rs_build_stats stats{};
    rsa::DeviceCopy *dev = nullptr;
};

namespace rsa {
void set_error(const std::string &msg);
void device_copy_free(DeviceCopy *d);

// Builds one MPHF per key set on ctx (set i: set_n[i] keys at set_keys[i], device
// memory); returns an rs_status and, on success, the handles in out.
int build_sets_status(rs_ctx *c, const std::vector<const uint64_t *> &set_keys, const std::vector<uint64_t> &set_n,
                      uint32_t leaf, uint64_t bucket, int flags, std::vector<rs_mphf *> &out);
// Upper bound of the device bytes a build of n keys in nb buckets uses.
uint64_t build_bound_bytes(uint64_t n, uint64_t nb);
int ctx_device(const rs_ctx *c);
}  // namespace rsa
