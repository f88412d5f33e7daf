// GPURecSplit.hpp -- header-only C++ drop-in for the reference's GPURecSplit class,
// backed by the gfx950 construction engine through the C ABI (gpurecsplit_amd.h).
//
// Reference surface (paths relative to the ByteHamster/GpuRecSplit root):
//   template <size_t LEAF_SIZE, sux::util::AllocType AT, bool USE_BIJECTIONS_ROTATE>
//   class GPURecSplit                         include/function/GPURecSplit.cuh:409-411
//   class recsplit_rotate::RecSplit           include/function/RecSplitRotate.hpp:115 (64-bit remap)
//   GPURecSplit()                             :432
//   GPURecSplit(vector<hash128_t>&, size_t bucket_size, int numThreads)   :484-489
//   size_t operator()(const hash128_t&)       include/function/AbstractParallelRecSplit.hpp:409-466
//   size_t getBits()                          AbstractParallelRecSplit.hpp:476-478
//   operator<< / operator>>                   GPURecSplit.cuh:817-841
//   hash128_t {first, second}                 include/util/Hash128.h:14-23
//
// Differences, all deliberate:
//   - numThreads defaults to 1 (the two-argument form); construction runs on the GPU.
//   - The caller's key vector is not reordered (the reference sorts it in place).
//   - Failures throw gpurecsplit_amd::Error instead of exit()/abort() or looping
//     forever on duplicate keys.
//   - String keys are hashed with the library's SpookyHash V2 (rs_first_hash), on the
//     GPU for construction; the result is the reference's first_hash.
//
// Link with -lrecsplit_amd (gpurecsplit_amd/librecsplit_amd.so).  The GPU is the
// one selected by the environment variable RS_DEVICE (default 0).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <istream>
#include <iterator>
#include <mutex>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gpurecsplit_amd.h"

#if __has_include(<sux/util/Vector.hpp>)
#include <sux/util/Vector.hpp>
#else
namespace sux::util {
// Same enumerators as sux::util::AllocType; the engine keeps its own buffers, so
// the value only selects the template instance.
enum AllocType { MALLOC, SMALLPAGE, TRANSHUGEPAGE, FORCEHUGEPAGE };
}  // namespace sux::util
#endif

namespace gpurecsplit_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int code, const char *where) {
    if (code != RS_OK) throw Error(code, std::string(where) + ": " + rs_last_error());
}

// One build context per process and device, created on first use.
inline rs_ctx *default_context() {
    static std::once_flag once;
    static rs_ctx *ctx = nullptr;
    std::call_once(once, [] {
        const char *d = std::getenv("RS_DEVICE");
        check(rs_ctx_create(d ? std::atoi(d) : 0, &ctx), "rs_ctx_create");
    });
    return ctx;
}

}  // namespace gpurecsplit_amd

namespace bez::function {

#ifndef GPURECSPLIT_AMD_HASH128_DEFINED
#define GPURECSPLIT_AMD_HASH128_DEFINED
/** 128-bit key hash; `first` selects the bucket, `second` is hashed with the seeds. */
typedef struct __hash128_t {
    uint64_t first, second;

    bool operator<(const __hash128_t &o) const { return first < o.first || second < o.second; }

    __hash128_t() : first(0), second(0) {}
    __hash128_t(const uint64_t first, const uint64_t second) {
        this->first = first;
        this->second = second;
    }
} hash128_t;
static_assert(sizeof(hash128_t) == 16, "hash128_t must be two packed uint64");
#endif

namespace detail {

// The construction engine behind both reference classes: GPURecSplit (16-bit remap)
// and recsplit_rotate::RecSplit (64-bit remap); REMAP64 selects the hash-to-child
// mapping (AbstractParallelRecSplit.hpp:327-337).
template <size_t LEAF_SIZE, sux::util::AllocType AT, bool USE_BIJECTIONS_ROTATE, bool REMAP64>
class RecSplitImpl {
    static_assert(LEAF_SIZE >= 2 && LEAF_SIZE <= 24, "leaf size must be in [2, 24]");
    // the C ABI's flags word: 1 rotation fitting, 2 the 64-bit remap
    static constexpr int kFlags = (USE_BIJECTIONS_ROTATE ? 1 : 0) | (REMAP64 ? 2 : 0);
    rs_mphf *h = nullptr;

  public:
    RecSplitImpl() {}

    /** Builds the MPHF of `keys` (distinct 128-bit hashes) with the given bucket size. */
    RecSplitImpl(std::vector<hash128_t> &keys, const size_t bucket_size, const int numThreads = 1) {
        (void)numThreads;
        gpurecsplit_amd::check(rs_build(gpurecsplit_amd::default_context(),
                                        reinterpret_cast<const uint64_t *>(keys.data()), keys.size(), LEAF_SIZE,
                                        bucket_size, kFlags, &h),
                               "GPURecSplit");
    }

    /** Builds the MPHF of distinct string keys (GPURecSplit.cuh:443-478): every key is
     *  replaced by its 128-bit SpookyHash V2 first_hash, computed on the GPU. */
    RecSplitImpl(const std::vector<std::string> &keys, const size_t bucket_size, const int numThreads = 1) {
        (void)numThreads;
        build_strings(keys, bucket_size);
    }

    /** Builds from the keys of a stream, one per line (GPURecSplit.cuh:498-505). */
    RecSplitImpl(std::ifstream &input, const size_t bucket_size, const int numThreads = 1) {
        (void)numThreads;
        std::vector<std::string> keys;
        for (std::string key; std::getline(input, key);) keys.push_back(std::move(key));
        build_strings(keys, bucket_size);
    }

    /** Builds from keys already resident in device memory of the context's GPU. */
    static RecSplitImpl fromDevice(const hash128_t *d_keys, size_t n, size_t bucket_size) {
        RecSplitImpl r;
        gpurecsplit_amd::check(rs_build_device(gpurecsplit_amd::default_context(),
                                               reinterpret_cast<const uint64_t *>(d_keys), n, LEAF_SIZE, bucket_size,
                                               kFlags, &r.h),
                               "GPURecSplit::fromDevice");
        return r;
    }

    RecSplitImpl(const RecSplitImpl &) = delete;
    RecSplitImpl &operator=(const RecSplitImpl &) = delete;
    RecSplitImpl(RecSplitImpl &&o) noexcept : h(o.h) { o.h = nullptr; }
    RecSplitImpl &operator=(RecSplitImpl &&o) noexcept {
        if (this != &o) {
            rs_mphf_free(h);
            h = o.h;
            o.h = nullptr;
        }
        return *this;
    }
    ~RecSplitImpl() { rs_mphf_free(h); }

    /** Returns the value in [0, n) associated with the given 128-bit hash. */
    size_t operator()(const hash128_t &hash) { return (size_t)rs_mphf_query(h, hash.first, hash.second); }

    /** Returns the value in [0, n) associated with a string key (AbstractParallelRecSplit.hpp:473). */
    size_t operator()(const std::string &key) { return (size_t)rs_mphf_query_string(h, key.data(), key.size()); }

    /** Bytes of the reference's AbstractParallelRecSplit object, the sizeof term of its
     *  getBits(): bucket_size, nbuckets, keys_count (3 words), the RiceBitVector (one
     *  sux::util::Vector) and the DoubleEF (4 Vectors and 10 words), with a
     *  sux::util::Vector taken as pointer + size + capacity (24 bytes). */
    static constexpr size_t kReferenceObjectBytes = 3 * 8 + 24 + (4 * 24 + 10 * 8);

    /** Space of the structure in bits (AbstractParallelRecSplit.hpp:476-478). */
    size_t getBits() { return (size_t)rs_mphf_bits(h) + 8 * kReferenceObjectBytes; }

    /** Number of keys of the function. */
    size_t size() const { return h ? (size_t)rs_mphf_num_keys(h) : 0; }

    /** Batched queries on the GPU: d_keys (n hashes) and d_out (n values) in device memory. */
    void queryDevice(const hash128_t *d_keys, size_t n, uint64_t *d_out) const {
        gpurecsplit_amd::check(rs_query_device(gpurecsplit_amd::default_context(), h,
                                               reinterpret_cast<const uint64_t *>(d_keys), n, d_out),
                               "GPURecSplit::queryDevice");
    }

    const rs_mphf *handle() const { return h; }

  private:
    void build_strings(const std::vector<std::string> &keys, const size_t bucket_size) {
        std::vector<uint64_t> offsets(keys.size() + 1, 0);
        for (size_t i = 0; i < keys.size(); i++) offsets[i + 1] = offsets[i] + keys[i].size();
        std::vector<uint8_t> bytes(offsets.back());
        for (size_t i = 0; i < keys.size(); i++)
            if (!keys[i].empty()) std::memcpy(bytes.data() + offsets[i], keys[i].data(), keys[i].size());
        gpurecsplit_amd::check(rs_build_strings(gpurecsplit_amd::default_context(), bytes.data(), offsets.data(),
                                                keys.size(), LEAF_SIZE, bucket_size, kFlags, &h),
                               "GPURecSplit(strings)");
    }

  public:

    friend std::ostream &operator<<(std::ostream &os, const RecSplitImpl &rs) {
        std::vector<uint8_t> buf(rs_mphf_serialized_size(rs.h));
        gpurecsplit_amd::check(rs_mphf_serialize(rs.h, buf.data(), buf.size()), "operator<<");
        os.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)buf.size());
        return os;
    }

    /** Reads exactly one serialized structure (GPURecSplit.cuh:827-841): the three
     *  header words, the descriptor vector, the six Elias-Fano scalars and its four
     *  vectors, each vector as a word count followed by the words (DoubleEF.hpp:115-135).
     *  Bytes after the structure stay in the stream. */
    friend std::istream &operator>>(std::istream &is, RecSplitImpl &rs) {
        std::vector<uint8_t> buf;
        auto take = [&](size_t bytes) {
            const size_t at = buf.size();
            buf.resize(at + bytes);
            if (bytes && !is.read(reinterpret_cast<char *>(buf.data() + at), (std::streamsize)bytes))
                throw gpurecsplit_amd::Error(RS_ERR_FORMAT, "operator>>: truncated stream");
            return at;
        };
        auto word = [&]() {
            uint64_t v;
            std::memcpy(&v, buf.data() + take(8), 8);
            return v;
        };
        auto vec = [&]() {
            uint64_t n = word();
            while (n) {  // in bounded chunks: a corrupt count fails at the end of the stream
                const uint64_t k = n < (uint64_t(1) << 20) ? n : (uint64_t(1) << 20);
                take(8 * (size_t)k);
                n -= k;
            }
        };
        for (int i = 0; i < 3; i++) word();
        vec();
        for (int i = 0; i < 6; i++) word();
        for (int i = 0; i < 4; i++) vec();
        rs_mphf *nh = nullptr;
        gpurecsplit_amd::check(rs_mphf_deserialize(buf.data(), buf.size(), LEAF_SIZE, kFlags, &nh),
                               "operator>>");
        rs_mphf_free(rs.h);
        rs.h = nh;
        return is;
    }
};

}  // namespace detail

/** The reference's GPURecSplit<LEAF_SIZE, AT, USE_BIJECTIONS_ROTATE> (GPURecSplit.cuh:409-411). */
template <size_t LEAF_SIZE, sux::util::AllocType AT = sux::util::AllocType::MALLOC, bool USE_BIJECTIONS_ROTATE = true>
using GPURecSplit = detail::RecSplitImpl<LEAF_SIZE, AT, USE_BIJECTIONS_ROTATE, false>;

namespace recsplit_rotate {
/** The reference's recsplit_rotate::RecSplit<LEAF_SIZE, AT, USE_BIJECTIONS_ROTATE>
 *  (RecSplitRotate.hpp:115): the same construction with the 64-bit remap; built on
 *  the GPU here (the reference builds it on host threads). */
template <size_t LEAF_SIZE, sux::util::AllocType AT = sux::util::AllocType::MALLOC, bool USE_BIJECTIONS_ROTATE = true>
using RecSplit = detail::RecSplitImpl<LEAF_SIZE, AT, USE_BIJECTIONS_ROTATE, true>;
}  // namespace recsplit_rotate

}  // namespace bez::function
