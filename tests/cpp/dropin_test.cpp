// Exercises the C++ drop-in header (include/gpurecsplit_amd/GPURecSplit.hpp) the
// way the reference's benchmark and tests use GPURecSplit:
//   dropin_test load  <mphf.bin> <keys.bin> <leaf>   operator>> + operator() bijection (no GPU)
//   dropin_test load2 <stream.bin> <keysA.bin> <keysB.bin> <leaf>
//                                                   two structures back to back, then trailing
//                                                   bytes: operator>> reads exactly one each (no GPU)
//   dropin_test loadstr <mphf.bin> <keys.txt> <leaf>   operator>> + operator()(string) (no GPU)
//   dropin_test buildstr <keys.txt> <leaf> <bucket> <out.bin>
//                                                   string constructor from an ifstream (GPU)
//   dropin_test build <keys.bin> <leaf> <bucket> <out.bin>
//   dropin_test load64 / build64 ...  the same with recsplit_rotate::RecSplit (64-bit remap)
//                                                   constructor + operator() + operator<< (GPU)
// keys.bin holds n little-endian (first, second) uint64 pairs.
#include <gpurecsplit_amd/GPURecSplit.hpp>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

using bez::function::GPURecSplit;
using bez::function::hash128_t;
namespace rr = bez::function::recsplit_rotate;

static std::vector<hash128_t> read_keys(const char *path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<hash128_t> keys(raw.size() / 16);
    std::memcpy(keys.data(), raw.data(), keys.size() * 16);
    return keys;
}

template <class RS>
static int check_bijection(RS &rs, const std::vector<hash128_t> &keys) {
    // test/recsplitCorrectness.hpp:10-35, benchmark_construction.hpp:68-76
    std::vector<bool> taken(keys.size(), false);
    for (const auto &k : keys) {
        const size_t v = rs(k);
        if (v >= keys.size() || taken[v]) {
            std::printf("collision or out of range for key (%llu, %llu): %zu\n", (unsigned long long)k.first,
                        (unsigned long long)k.second, v);
            return 1;
        }
        taken[v] = true;
    }
    return 0;
}

template <size_t L, template <size_t, sux::util::AllocType, bool> class RS = GPURecSplit>
static int run(int argc, char **argv) {
    using GPURecSplit = RS<L, sux::util::AllocType::MALLOC, true>;
    std::string mode = argv[1];
    if (mode.size() > 2 && mode.compare(mode.size() - 2, 2, "64") == 0) mode = mode.substr(0, mode.size() - 2);
    if (mode == "load") {
        GPURecSplit rs;
        std::ifstream in(argv[2], std::ios::binary);
        in >> rs;
        auto keys = read_keys(argv[3]);
        if (check_bijection(rs, keys)) return 1;
        std::printf("load ok: n=%zu bits/key=%.4f\n", keys.size(), (double)rs.getBits() / keys.size());
        return 0;
    }
    if (mode == "loadstr" || mode == "buildstr") {
        std::vector<std::string> keys;
        {
            std::ifstream f(mode == "loadstr" ? argv[3] : argv[2]);
            for (std::string k; std::getline(f, k);) keys.push_back(k);
        }
        GPURecSplit rs;
        if (mode == "loadstr") {
            std::ifstream in(argv[2], std::ios::binary);
            in >> rs;
        } else {
            std::ifstream in(argv[2]);
            rs = GPURecSplit(in, std::stoul(argv[4]));
            std::ofstream out(argv[5], std::ios::binary);
            out << rs;
        }
        std::vector<bool> taken(keys.size(), false);
        for (const auto &k : keys) {
            const size_t v = rs(k);
            if (v >= keys.size() || taken[v]) {
                std::printf("collision or out of range for key %s: %zu\n", k.c_str(), v);
                return 1;
            }
            taken[v] = true;
        }
        std::printf("%s ok: n=%zu\n", mode.c_str(), keys.size());
        return 0;
    }
    if (mode == "load2") {
        GPURecSplit a, b;
        std::ifstream in(argv[2], std::ios::binary);
        in >> a >> b;
        std::string rest((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        auto ka = read_keys(argv[3]), kb = read_keys(argv[4]);
        if (check_bijection(a, ka) || check_bijection(b, kb)) return 1;
        std::printf("load2 ok: n=%zu,%zu trailing=%s\n", ka.size(), kb.size(), rest.c_str());
        return 0;
    }
    auto keys = read_keys(argv[2]);
    const auto before = keys;
    GPURecSplit rs(keys, std::stoul(argv[4]));
    if (std::memcmp(before.data(), keys.data(), keys.size() * 16) != 0) {
        std::printf("caller keys were modified\n");
        return 1;
    }
    if (check_bijection(rs, keys)) return 1;
    std::ofstream out(argv[5], std::ios::binary);
    out << rs;
    std::printf("build ok: n=%zu bits/key=%.4f\n", keys.size(), (double)rs.getBits() / keys.size());
    (void)argc;
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: see source\n");
        return 2;
    }
    std::string mode = argv[1];
    const bool r64 = mode.size() > 2 && mode.compare(mode.size() - 2, 2, "64") == 0;
    if (r64) mode = mode.substr(0, mode.size() - 2);
    const int leaf = std::stoi(mode == "load" || mode == "loadstr" ? argv[4] : mode == "load2" ? argv[5] : argv[3]);
    try {
        switch (leaf) {
        case 8: return r64 ? run<8, rr::RecSplit>(argc, argv) : run<8>(argc, argv);
        case 12: return r64 ? run<12, rr::RecSplit>(argc, argv) : run<12>(argc, argv);
        case 16: return r64 ? run<16, rr::RecSplit>(argc, argv) : run<16>(argc, argv);
        default: std::fprintf(stderr, "leaf size %d not compiled in\n", leaf); return 2;
        }
    } catch (const gpurecsplit_amd::Error &e) {
        std::printf("error %d: %s\n", e.code, e.what());
        return 3;
    }
}
