// altopir/engine.hpp -- C++ mirror of the reference's Rust engine wrapper
// (src/engine.rs:22-195) over the C ABI of altopir_engine.h.
//
// Same names and checks as the Rust module: Method, MappedBytes, TiledDb,
// build_packing_matrix and Engine{new, request_len, response_len, answer}.  Where the
// Rust code asserts or panics, this header throws std::runtime_error with the same
// message.  Header-only; link against libaltopir_hip.so.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../altopir_engine.h"

namespace altopir {

inline void check(bool ok, const std::string &msg) {
    if (!ok) throw std::runtime_error(msg);
}

// src/engine.rs:22-50; Hip is this engine's own code.  All three run the HIP engine.
enum class Method : int { Amx = ALTO_METHOD_AMX, Vnni = ALTO_METHOD_VNNI, Hip = ALTO_METHOD_HIP };

inline bool parse_method(const std::string &s, Method &m) {
    if (s == "amx") m = Method::Amx;
    else if (s == "vnni") m = Method::Vnni;
    else if (s == "hip") m = Method::Hip;
    else return false;
    return true;
}

// src/engine.rs:66-69
inline void init_amx() { check(sw_amx_init() == 0, "the kernel did not grant AMX tile state"); }

// src/engine.rs:71-100
class MappedBytes {
  public:
    static MappedBytes zeroed(size_t len) {
        check(len > 0, "assertion failed: len > 0");
        void *p = sw_alloc_zeroed(len);
        check(p != nullptr, "out of memory");
        return MappedBytes((int8_t *)p, len);
    }
    MappedBytes(MappedBytes &&o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), len_(o.len_) {}
    MappedBytes(const MappedBytes &) = delete;
    ~MappedBytes() { sw_free(ptr_); }
    int8_t *data() { return ptr_; }
    const int8_t *data() const { return ptr_; }
    size_t size() const { return len_; }

  private:
    MappedBytes(int8_t *p, size_t n) : ptr_(p), len_(n) {}
    int8_t *ptr_;
    size_t len_;
};

// src/engine.rs:102-125
class TiledDb {
  public:
    TiledDb(const uint8_t *db, size_t db_len, size_t ld, size_t rows, size_t cols, int threads)
        : rows(rows), cols(cols) {
        check(rows % 64 == 0 && cols % 32 == 0 && db_len >= (cols - 1) * ld + rows,
              "assertion failed: rows % 64 == 0 && cols % 32 == 0 && db.len() >= (cols - 1) * ld + rows");
        ptr_ = (uint8_t *)sw_alloc(rows * cols);
        check(ptr_ != nullptr, "out of memory");
        sw_tile_db(ptr_, db, ld, rows, cols, threads);
    }
    TiledDb(const TiledDb &) = delete;
    ~TiledDb() { sw_free(ptr_); }
    const uint8_t *data() const { return ptr_; }
    const size_t rows, cols;

  private:
    uint8_t *ptr_ = nullptr;
};

// src/engine.rs:127-138
inline void build_packing_matrix(MappedBytes &ub, const std::vector<const uint64_t *> &bold_t,
                                 const std::vector<const uint64_t *> &bold_t_bar, const std::vector<uint32_t> &tables,
                                 const std::vector<uint32_t> &gen_pows, size_t num_iter, size_t tg, size_t n,
                                 uint64_t qmod, int threads) {
    const size_t rho = bold_t.size();
    check(ub.size() == 4 * tg * n * rho * n && bold_t_bar.size() == rho, "packing matrix size");
    check(tables.size() == n * n && gen_pows.size() >= num_iter, "tables / gen_pows size");
    sw_pack_build(ub.data(), bold_t.data(), bold_t_bar.data(), tables.data(), gen_pows.data(), num_iter, tg, n, rho,
                  qmod, threads);
}

// src/engine.rs:140-195
class Engine {
  public:
    Engine(const uint8_t *db_col, size_t db_len, const TiledDb &tiled, size_t ld, const int8_t *ub, size_t ub_len,
           size_t n, size_t tg, const std::vector<std::vector<uint8_t>> &prefixes, size_t row1_off, uint64_t q1,
           size_t q1_bits, const std::vector<const uint64_t *> &that, const uint64_t *inv_table, uint64_t q,
           int threads) {
        const size_t rows = tiled.rows, cols = tiled.cols;
        check(db_len >= (cols - 1) * ld + rows && ub_len == 4 * tg * n * cols && prefixes.size() == cols / n,
              "engine inputs");
        check(that.size() == cols / n, "that.len() == cols / n");
        const size_t part_len = prefixes[0].size();
        std::vector<const uint8_t *> pp;
        for (auto &p : prefixes) {
            check(p.size() == part_len, "prefix lengths differ");
            pp.push_back(p.data());
        }
        ptr_ = alto_engine_new(db_col, tiled.data(), ld, rows, cols, ub, n, tg, pp.data(), part_len, row1_off, q1,
                               (int)q1_bits, that.data(), inv_table, q, threads);
        check(ptr_ != nullptr, std::string("unsupported engine configuration: ") + alto_last_error());
    }
    Engine(const Engine &) = delete;
    ~Engine() { alto_engine_free(ptr_); }

    size_t request_len() const { return alto_request_len(ptr_); }
    size_t response_len() const { return alto_response_len(ptr_); }

    // Returns (answered, stage_ms); out[c] must hold response_len() bytes.
    std::pair<size_t, std::array<double, 4>> answer(const std::vector<const uint8_t *> &req,
                                                    const std::vector<size_t> &len, std::vector<uint8_t *> &out,
                                                    std::vector<int> &status, Method method) {
        const size_t b = req.size();
        check(out.size() == b && status.size() == b && len.size() == b, "batch sizes differ");
        std::array<double, 4> stage_ms{};
        const int r = alto_answer(ptr_, req.data(), len.data(), b, out.data(), status.data(), (int)method,
                                  stage_ms.data());
        check(r >= 0, "engine error " + std::to_string(r));
        return {(size_t)r, stage_ms};
    }

    void *raw() { return ptr_; }

  private:
    void *ptr_ = nullptr;
};

}  // namespace altopir
