// Drives the non-GPU entry points of include/altopir/engine.hpp (the C++ mirror of the
// reference's src/engine.rs) through libaltopir_hip.so, for tests/test_abi.py.
//
//   engine_hpp_check <out_dir>
//
// Writes <out_dir>/tiled.bin (TiledDb over a seeded column-major database) and
// <out_dir>/ub.bin (build_packing_matrix over seeded inputs) for the test to compare with
// the oracle, and prints one "ok <check>" line per check of the wrapper's own behaviour.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "altopir/engine.hpp"

static uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static bool write_file(const std::string &path, const void *p, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool good = fwrite(p, 1, n, f) == n;
    return fclose(f) == 0 && good;
}

template <class F>
static bool throws(F f, const std::string &part) {
    try {
        f();
    } catch (const std::runtime_error &e) {
        return std::string(e.what()).find(part) != std::string::npos;
    }
    return false;
}

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    using namespace altopir;

    Method m;
    if (!parse_method("amx", m) || m != Method::Amx || !parse_method("vnni", m) || m != Method::Vnni ||
        !parse_method("hip", m) || m != Method::Hip || parse_method("avx2", m))
        return 3;
    printf("ok parse_method\n");
    init_amx();
    printf("ok init_amx\n");

    {
        MappedBytes z = MappedBytes::zeroed(3 << 20);
        for (size_t i = 0; i < z.size(); i++)
            if (z.data()[i]) return 4;
        if ((uintptr_t)z.data() % (2 << 20)) return 5;
        MappedBytes moved(std::move(z));
        if (moved.size() != (3u << 20)) return 6;
    }
    if (!throws([] { MappedBytes::zeroed(0); }, "len > 0")) return 7;
    printf("ok MappedBytes\n");

    // TiledDb over 256 records x 64 columns with a column pitch of 320 bytes
    const size_t rows = 256, cols = 64, ld = 320;
    std::vector<uint8_t> db(cols * ld);
    uint64_t s = 4;
    for (auto &b : db) b = (uint8_t)splitmix(s);
    if (!throws([&] { TiledDb bad(db.data(), db.size(), ld, rows + 1, cols, 2); }, "rows % 64 == 0")) return 8;
    TiledDb tiled(db.data(), db.size(), ld, rows, cols, 3);
    if (!write_file(dir + "/db.bin", db.data(), db.size()) ||
        !write_file(dir + "/tiled.bin", tiled.data(), rows * cols))
        return 9;
    printf("ok TiledDb\n");

    // build_packing_matrix for n = 16, tg = 8, two outputs
    const size_t n = 16, tg = 8, rho = 2, num_iter = n / 2 - 1;
    std::vector<uint32_t> br(n), tables(n * n), gen_pows(n);
    for (size_t j = 0; j < n; j++) {
        size_t r = 0;
        for (size_t b = 0, x = j; b < 4; b++, x >>= 1) r = (r << 1) | (x & 1);
        br[j] = (uint32_t)r;
    }
    for (size_t t = 1, row = 0; t < 2 * n; t += 2, row++)
        for (size_t j = 0; j < n; j++) tables[row * n + j] = br[((t * (2 * br[j] + 1)) % (2 * n) - 1) / 2];
    for (size_t i = 0, g = 1; i < n; i++, g = g * 5 % (2 * n)) gen_pows[i] = (uint32_t)g;
    std::vector<std::vector<uint64_t>> bt(rho), bb(rho);
    std::vector<const uint64_t *> pt, pb;
    for (size_t o = 0; o < rho; o++) {
        bt[o].resize(num_iter * tg * n);
        bb[o].resize(num_iter * tg * n);
        uint64_t s1 = 70 + o, s2 = 80 + o;
        for (auto &v : bt[o]) v = splitmix(s1);
        for (auto &v : bb[o]) v = splitmix(s2);
        pt.push_back(bt[o].data());
        pb.push_back(bb[o].data());
    }
    MappedBytes ub = MappedBytes::zeroed(4 * tg * n * rho * n);
    MappedBytes small = MappedBytes::zeroed(16);
    if (!throws([&] { build_packing_matrix(small, pt, pb, tables, gen_pows, num_iter, tg, n, 4294955009ull, 2); },
                "packing matrix size"))
        return 10;
    build_packing_matrix(ub, pt, pb, tables, gen_pows, num_iter, tg, n, 4294955009ull, 3);
    if (!write_file(dir + "/ub.bin", ub.data(), ub.size())) return 11;
    for (size_t o = 0; o < rho; o++)
        if (!write_file(dir + "/bt" + std::to_string(o) + ".bin", bt[o].data(), bt[o].size() * 8) ||
            !write_file(dir + "/bb" + std::to_string(o) + ".bin", bb[o].data(), bb[o].size() * 8))
            return 12;
    printf("ok build_packing_matrix\n");

    // Engine: an unsupported configuration (q != Q) throws with the library's message, before
    // any device is touched
    std::vector<int8_t> u(4 * 4 * 16 * 64);
    std::vector<std::vector<uint8_t>> prefixes(4, std::vector<uint8_t>(112));
    std::vector<std::vector<uint64_t>> th(4, std::vector<uint64_t>(64));
    std::vector<const uint64_t *> that;
    for (auto &t : th) that.push_back(t.data());
    std::vector<uint64_t> inv(16);
    if (!throws(
            [&] {
                Engine e(db.data(), db.size(), tiled, ld, u.data(), u.size(), 16, 4, prefixes, 288, 1024, 10, that,
                         inv.data(), 4294955011ull, 2);
            },
            "unsupported engine configuration"))
        return 13;
    if (!throws(
            [&] {
                Engine e(db.data(), db.size(), tiled, ld, u.data(), u.size() - 1, 16, 4, prefixes, 288, 1024, 10,
                         that, inv.data(), 4294955009ull, 2);
            },
            "engine inputs"))
        return 14;
    printf("ok Engine checks\n");
    return 0;
}
