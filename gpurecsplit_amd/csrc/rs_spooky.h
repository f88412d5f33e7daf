// rs_spooky.h -- SpookyHash V2 (Bob Jenkins, 2012, public domain), 128-bit form,
// for the string-key front end: the reference hashes every string key with
// SpookyHash::Hash128(data, len, &h0, &h1), h0 = h1 = 0, and uses {first = h1,
// second = h0} as its 128-bit key (include/util/Hash128.h:32-36,
// AbstractParallelRecSplit.hpp:250).  Written from the published algorithm:
// inputs under 192 bytes take the short path (4-word state, 32-byte steps);
// longer inputs run the 12-word Mix over 96-byte blocks and End over the zero-
// padded tail whose last byte holds the tail length.  Words are read
// little-endian.  Shared by the host (query, host hashing) and the device kernel.
#pragma once

#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define RS_SPOOKY_HD __host__ __device__ __forceinline__
#else
#define RS_SPOOKY_HD inline
#endif

namespace rsa {
namespace spooky {

constexpr uint64_t kConst = 0xdeadbeefdeadbeefULL;
constexpr size_t kNumVars = 12;
constexpr size_t kBlock = kNumVars * 8;  // 96 bytes
constexpr size_t kBuf = 2 * kBlock;      // 192 bytes: the short/long threshold

RS_SPOOKY_HD uint64_t rot(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Little-endian reads at any byte offset.
RS_SPOOKY_HD uint64_t rd64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
RS_SPOOKY_HD uint64_t rd32(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 4; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

RS_SPOOKY_HD void short_mix(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h2 = rot(h2, 50); h2 += h3; h0 ^= h2;
    h3 = rot(h3, 52); h3 += h0; h1 ^= h3;
    h0 = rot(h0, 30); h0 += h1; h2 ^= h0;
    h1 = rot(h1, 41); h1 += h2; h3 ^= h1;
    h2 = rot(h2, 54); h2 += h3; h0 ^= h2;
    h3 = rot(h3, 48); h3 += h0; h1 ^= h3;
    h0 = rot(h0, 38); h0 += h1; h2 ^= h0;
    h1 = rot(h1, 37); h1 += h2; h3 ^= h1;
    h2 = rot(h2, 62); h2 += h3; h0 ^= h2;
    h3 = rot(h3, 34); h3 += h0; h1 ^= h3;
    h0 = rot(h0, 5);  h0 += h1; h2 ^= h0;
    h1 = rot(h1, 36); h1 += h2; h3 ^= h1;
}

RS_SPOOKY_HD void short_end(uint64_t &h0, uint64_t &h1, uint64_t &h2, uint64_t &h3) {
    h3 ^= h2; h2 = rot(h2, 15); h3 += h2;
    h0 ^= h3; h3 = rot(h3, 52); h0 += h3;
    h1 ^= h0; h0 = rot(h0, 26); h1 += h0;
    h2 ^= h1; h1 = rot(h1, 51); h2 += h1;
    h3 ^= h2; h2 = rot(h2, 28); h3 += h2;
    h0 ^= h3; h3 = rot(h3, 9);  h0 += h3;
    h1 ^= h0; h0 = rot(h0, 47); h1 += h0;
    h2 ^= h1; h1 = rot(h1, 54); h2 += h1;
    h3 ^= h2; h2 = rot(h2, 32); h3 += h2;
    h0 ^= h3; h3 = rot(h3, 25); h0 += h3;
    h1 ^= h0; h0 = rot(h0, 63); h1 += h0;
}

// Inputs shorter than 192 bytes.
RS_SPOOKY_HD void hash_short(const uint8_t *p, size_t length, uint64_t &hash1, uint64_t &hash2) {
    size_t remainder = length % 32;
    uint64_t a = hash1, b = hash2, c = kConst, d = kConst;
    if (length > 15) {
        const uint8_t *end = p + (length / 32) * 32;
        for (; p < end; p += 32) {
            c += rd64(p);
            d += rd64(p + 8);
            short_mix(a, b, c, d);
            a += rd64(p + 16);
            b += rd64(p + 24);
        }
        if (remainder >= 16) {
            c += rd64(p);
            d += rd64(p + 8);
            short_mix(a, b, c, d);
            p += 16;
            remainder -= 16;
        }
    }
    d += (uint64_t)length << 56;
    // the last 0..15 bytes: bytes 0-7 into c, bytes 8-14 into d
    if (remainder == 0) {
        c += kConst;
        d += kConst;
    } else {
        for (size_t i = 0; i < remainder && i < 8; i++) c += (uint64_t)p[i] << (8 * i);
        for (size_t i = 8; i < remainder; i++) d += (uint64_t)p[i] << (8 * (i - 8));
    }
    short_end(a, b, c, d);
    hash1 = a;
    hash2 = b;
}

RS_SPOOKY_HD void mix(const uint64_t *data, uint64_t (&s)[kNumVars]) {
    s[0] += data[0];   s[2] ^= s[10];  s[11] ^= s[0];  s[0] = rot(s[0], 11);   s[11] += s[1];
    s[1] += data[1];   s[3] ^= s[11];  s[0] ^= s[1];   s[1] = rot(s[1], 32);   s[0] += s[2];
    s[2] += data[2];   s[4] ^= s[0];   s[1] ^= s[2];   s[2] = rot(s[2], 43);   s[1] += s[3];
    s[3] += data[3];   s[5] ^= s[1];   s[2] ^= s[3];   s[3] = rot(s[3], 31);   s[2] += s[4];
    s[4] += data[4];   s[6] ^= s[2];   s[3] ^= s[4];   s[4] = rot(s[4], 17);   s[3] += s[5];
    s[5] += data[5];   s[7] ^= s[3];   s[4] ^= s[5];   s[5] = rot(s[5], 28);   s[4] += s[6];
    s[6] += data[6];   s[8] ^= s[4];   s[5] ^= s[6];   s[6] = rot(s[6], 39);   s[5] += s[7];
    s[7] += data[7];   s[9] ^= s[5];   s[6] ^= s[7];   s[7] = rot(s[7], 57);   s[6] += s[8];
    s[8] += data[8];   s[10] ^= s[6];  s[7] ^= s[8];   s[8] = rot(s[8], 55);   s[7] += s[9];
    s[9] += data[9];   s[11] ^= s[7];  s[8] ^= s[9];   s[9] = rot(s[9], 54);   s[8] += s[10];
    s[10] += data[10]; s[0] ^= s[8];   s[9] ^= s[10];  s[10] = rot(s[10], 22); s[9] += s[11];
    s[11] += data[11]; s[1] ^= s[9];   s[10] ^= s[11]; s[11] = rot(s[11], 46); s[10] += s[0];
}

RS_SPOOKY_HD void end_partial(uint64_t (&h)[kNumVars]) {
    h[11] += h[1];  h[2] ^= h[11];  h[1] = rot(h[1], 44);
    h[0] += h[2];   h[3] ^= h[0];   h[2] = rot(h[2], 15);
    h[1] += h[3];   h[4] ^= h[1];   h[3] = rot(h[3], 34);
    h[2] += h[4];   h[5] ^= h[2];   h[4] = rot(h[4], 21);
    h[3] += h[5];   h[6] ^= h[3];   h[5] = rot(h[5], 38);
    h[4] += h[6];   h[7] ^= h[4];   h[6] = rot(h[6], 33);
    h[5] += h[7];   h[8] ^= h[5];   h[7] = rot(h[7], 10);
    h[6] += h[8];   h[9] ^= h[6];   h[8] = rot(h[8], 13);
    h[7] += h[9];   h[10] ^= h[7];  h[9] = rot(h[9], 38);
    h[8] += h[10];  h[11] ^= h[8];  h[10] = rot(h[10], 53);
    h[9] += h[11];  h[0] ^= h[9];   h[11] = rot(h[11], 42);
    h[10] += h[0];  h[1] ^= h[10];  h[0] = rot(h[0], 54);
}

// SpookyHash::Hash128: hash1, hash2 are the seeds on entry and the hash on exit.
RS_SPOOKY_HD void hash128(const uint8_t *p, size_t length, uint64_t &hash1, uint64_t &hash2) {
    if (length < kBuf) {
        hash_short(p, length, hash1, hash2);
        return;
    }
    uint64_t h[kNumVars];
    for (size_t i = 0; i < kNumVars; i += 3) {
        h[i] = hash1;
        h[i + 1] = hash2;
        h[i + 2] = kConst;
    }
    uint64_t w[kNumVars];
    const size_t blocks = length / kBlock;
    for (size_t b = 0; b < blocks; b++, p += kBlock) {
        for (size_t i = 0; i < kNumVars; i++) w[i] = rd64(p + 8 * i);
        mix(w, h);
    }
    // the tail (0..95 bytes), zero-padded, its length in the last byte
    const size_t remainder = length - blocks * kBlock;
    uint8_t buf[kBlock];
    for (size_t i = 0; i < kBlock; i++) buf[i] = i < remainder ? p[i] : 0;
    buf[kBlock - 1] = (uint8_t)remainder;
    for (size_t i = 0; i < kNumVars; i++) h[i] += rd64(buf + 8 * i);
    end_partial(h);
    end_partial(h);
    end_partial(h);
    hash1 = h[0];
    hash2 = h[1];
}

// The reference's first_hash(key, len) = spooky(key, len, 0): {first = h1, second = h0}.
RS_SPOOKY_HD void first_hash(const uint8_t *p, size_t length, uint64_t &first, uint64_t &second) {
    uint64_t h0 = 0, h1 = 0;
    hash128(p, length, h0, h1);
    first = h1;
    second = h0;
}

}  // namespace spooky
}  // namespace rsa
