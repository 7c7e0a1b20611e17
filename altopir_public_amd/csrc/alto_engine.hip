// alto_engine.hip -- host side of the MI355X answer engine: the reference's C ABI
// (src/engine.rs:4-20, src/amx_kernels.cpp:97-615) over HBM-resident setup state and the
// kernels of alto_kernels.hpp.
#include <hip/hip_runtime.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/altopir_engine.h"
#include "alto_kernels.hpp"

using namespace alto;

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define HIP_OK(expr)                                                                   \
    ([&]() -> bool {                                                                   \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return false;                                                              \
        }                                                                              \
        return true;                                                                   \
    }())

uint64_t host_mulmod(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % kQ); }
uint64_t host_powmod(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= kQ;
    for (; e; e >>= 1, b = host_mulmod(b, b))
        if (e & 1) r = host_mulmod(r, b);
    return r;
}

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

constexpr size_t kMaxLds = 160 * 1024;  // LDS per workgroup on gfx950

// ------------------------------------------------------------- launch shapes --

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Split the k-range so that no workgroup accumulates more than kFoldSteps k-steps in
// i32, and further (down to min_steps per split) until about `target` workgroups exist.
int choose_splits(int KC, long long tiles, int target, int min_steps) {
    int gz = ceil_div(KC, kFoldSteps);
    while (tiles * gz < target && KC / (gz * 2) >= min_steps) gz *= 2;
    return gz;
}

// Split count for shapes that hold one workgroup per CU: from ceil(KC / kFoldSteps) up to
// four times that (and at least min_steps per split), the count that finishes in the least
// time when the workgroups run in rounds of 256 and each costs its k-steps plus a start-up
// of 16 steps; the smallest count within 2% of the best.  A grid of 1.6 rounds otherwise
// leaves 40% of the CUs idle in its second round.
int balanced_splits(int KC, long long tiles, int min_steps) {
    const int gz0 = ceil_div(KC, kFoldSteps);
    double cost[64];
    int last = gz0;
    double best = 1e300;
    for (int gz = gz0; gz <= 4 * gz0 && gz - gz0 < 64 && (gz == gz0 || KC / gz >= min_steps); ++gz) {
        const long long rounds = (tiles * gz + 255) / 256;
        cost[gz - gz0] = (double)rounds * ((double)KC / gz + 16.0);
        if (cost[gz - gz0] < best) best = cost[gz - gz0];
        last = gz;
    }
    for (int gz = gz0; gz <= last; ++gz)
        if (cost[gz - gz0] <= 1.02 * best) return gz;
    return gz0;
}

// Workgroup grid of one contraction (sets g.gx, g.gy, g.gz; returns the workgroup count):
// streaming shapes split the record range until about WGS workgroups (1024 unless a variant
// says otherwise: measured best or level at 1-64 GiB, 7% faster than 2048 at 8 GiB), ring
// shapes by balanced_splits (at 8 GiB the scan of 37 queries on 12 x 20 tiles takes 7 splits,
// 721 workgroups in 2.8 rounds, instead of 4 in 1.6 rounds: 4-9% faster on two boxes, and
// 3-7% at 33 and 48 queries); the packing contraction is never split (KP fits one i32 range).
template <int MB, int NBT, int WGS = 1024>
int plan_stream(GemmArgs &g, bool pack) {
    g.gx = ceil_div(g.NT, 4 * NBT);
    g.gy = ceil_div(g.MT, MB);
    g.gz = pack ? 1 : choose_splits(g.KC, (long long)g.gx * g.gy, WGS, 16);
    return 8 * ceil_div(g.gx, 8) * g.gy * g.gz;
}
template <int TM, int TN, int WM, int WN>
int plan_ring(GemmArgs &g, bool pack) {
    g.gx = ceil_div(g.NT, WN * TN);
    g.gy = ceil_div(g.MT, WM * TM);
    g.gz = pack ? 1 : balanced_splits(g.KC, (long long)g.gx * g.gy, 64);
    return 8 * ceil_div(g.gx, 8) * g.gy * g.gz;
}

template <int MB, int NBT, int U, bool PACK, int WGS = 1024>
void launch_stream(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    const int grid = plan_stream<MB, NBT, WGS>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((stream_kernel<MB, NBT, U, PACK, true>), dim3(grid), dim3(256), 0, st, (const i32x4 *)a,
                           (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((stream_kernel<MB, NBT, U, PACK, false>), dim3(grid), dim3(256), 0, st, (const i32x4 *)a,
                           (const i32x4 *)b, g);
}

template <int TM, int TN, int WM, int WN, bool PACK>
void launch_lds(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    const int grid = plan_ring<TM, TN, WM, WN>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((lds_kernel<TM, TN, WM, WN, PACK, true>), dim3(grid), dim3(256), 0, st, (const i32x4 *)a,
                           (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((lds_kernel<TM, TN, WM, WN, PACK, false>), dim3(grid), dim3(256), 0, st, (const i32x4 *)a,
                           (const i32x4 *)b, g);
}

// Allows the dynamic LDS of a dma_kernel instantiation (done once per engine creation,
// outside any stream capture).
template <int TM, int TN, int WM, int WN, int S, bool PACK, int KPB = 1, bool PRIO = false>
void prepare_dma() {
    constexpr size_t lds = (size_t)S * (WM * TM + WN * TN) * 1024;
    (void)hipFuncSetAttribute((const void *)dma_kernel<TM, TN, WM, WN, S, PACK, true, KPB, PRIO>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)dma_kernel<TM, TN, WM, WN, S, PACK, false, KPB, PRIO>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}
void prepare_none() {}

template <int TM, int TN, int WM, int WN, int S, bool PACK, int KPB = 1, bool PRIO = false>
void launch_dma(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    constexpr int NBLK = WM * TM + WN * TN;
    constexpr size_t lds = (size_t)S * NBLK * 1024;
    const int grid = plan_ring<TM, TN, WM, WN>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((dma_kernel<TM, TN, WM, WN, S, PACK, true, KPB, PRIO>), dim3(grid), dim3(64 * WM * WN),
                           lds, st, (const i32x4 *)a, (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((dma_kernel<TM, TN, WM, WN, S, PACK, false, KPB, PRIO>), dim3(grid), dim3(64 * WM * WN),
                           lds, st, (const i32x4 *)a, (const i32x4 *)b, g);
}

template <int TM, int TN, int WM, int WN, int S, bool PACK>
void prepare_dma2() {
    constexpr size_t lds = (size_t)S * (WM * TM + WN * TN) * 1024;
    (void)hipFuncSetAttribute((const void *)dma2_kernel<TM, TN, WM, WN, S, PACK, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)dma2_kernel<TM, TN, WM, WN, S, PACK, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <int TM, int TN, int WM, int WN, int S, bool PACK>
void launch_dma2(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    constexpr size_t lds = (size_t)S * (WM * TM + WN * TN) * 1024;
    const int grid = plan_ring<TM, TN, WM, WN>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((dma2_kernel<TM, TN, WM, WN, S, PACK, true>), dim3(grid), dim3(64 * WM * WN), lds, st,
                           (const i32x4 *)a, (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((dma2_kernel<TM, TN, WM, WN, S, PACK, false>), dim3(grid), dim3(64 * WM * WN), lds, st,
                           (const i32x4 *)a, (const i32x4 *)b, g);
}

template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool WIDE = false>
void prepare_dmab() {
    constexpr size_t lds = DmabShape<TM, TN, WM, WN, D, UNR, PACK, false, WIDE>::kLds;
    (void)hipFuncSetAttribute((const void *)dmab_kernel<TM, TN, WM, WN, D, UNR, PACK, true, WIDE>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)dmab_kernel<TM, TN, WM, WN, D, UNR, PACK, false, WIDE>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool WIDE = false>
void launch_dmab(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    constexpr size_t lds = DmabShape<TM, TN, WM, WN, D, UNR, PACK, false, WIDE>::kLds;
    const int grid = plan_ring<TM, TN, WM, WN>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((dmab_kernel<TM, TN, WM, WN, D, UNR, PACK, true, WIDE>), dim3(grid),
                           dim3(64 * WM * WN), lds, st, (const i32x4 *)a, (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((dmab_kernel<TM, TN, WM, WN, D, UNR, PACK, false, WIDE>), dim3(grid),
                           dim3(64 * WM * WN), lds, st, (const i32x4 *)a, (const i32x4 *)b, g);
}

template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK>
constexpr size_t dmac_lds() {
    return DmacShape<TM, TN, WM, WN, D, UNR, PACK, false>::kLds;
}
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK>
void prepare_dmac() {
    constexpr size_t lds = dmac_lds<TM, TN, WM, WN, D, UNR, PACK>();
    (void)hipFuncSetAttribute((const void *)dmac_kernel<TM, TN, WM, WN, D, UNR, PACK, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void *)dmac_kernel<TM, TN, WM, WN, D, UNR, PACK, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK>
void launch_dmac(GemmArgs g, const int8_t *a, const int8_t *b, hipStream_t st) {
    constexpr size_t lds = dmac_lds<TM, TN, WM, WN, D, UNR, PACK>();
    const int grid = plan_ring<TM, TN, WM, WN>(g, PACK);
    if (g.gz > 1)
        hipLaunchKernelGGL((dmac_kernel<TM, TN, WM, WN, D, UNR, PACK, true>), dim3(grid), dim3(64 * WM * WN), lds,
                           st, (const i32x4 *)a, (const i32x4 *)b, g);
    else
        hipLaunchKernelGGL((dmac_kernel<TM, TN, WM, WN, D, UNR, PACK, false>), dim3(grid), dim3(64 * WM * WN), lds,
                           st, (const i32x4 *)a, (const i32x4 *)b, g);
}

// Kernel variants of a contraction.  Every variant computes the full result for any number
// of A tiles (ragged tiles are clamped and masked); they differ in tile shapes and in how
// operands are staged, which the tuner chooses per batch size.
typedef void (*LaunchFn)(GemmArgs, const int8_t *, const int8_t *, hipStream_t);
typedef void (*PrepareFn)();
struct Variant {
    const char *name;
    LaunchFn scan, pack;
    PrepareFn prepare_scan, prepare_pack;
};
const Variant kVariants[] = {
    {"stream_m1_n4_u4", launch_stream<1, 4, 4, false>, launch_stream<1, 4, 4, true>, prepare_none, prepare_none},
    {"stream_m2_n4_u4", launch_stream<2, 4, 4, false>, launch_stream<2, 4, 4, true>, prepare_none, prepare_none},
    // the same with the record range split for about 2048 or 4096 workgroups
    {"stream_m2_n4_u4_z2k", launch_stream<2, 4, 4, false, 2048>, launch_stream<2, 4, 4, true>, prepare_none,
     prepare_none},
    {"stream_m2_n4_u4_z4k", launch_stream<2, 4, 4, false, 4096>, launch_stream<2, 4, 4, true>, prepare_none,
     prepare_none},
    {"stream_m4_n4_u2", launch_stream<4, 4, 2, false>, launch_stream<4, 4, 2, true>, prepare_none, prepare_none},
    {"stream_m8_n4_u2", launch_stream<8, 4, 2, false>, launch_stream<8, 4, 2, true>, prepare_none, prepare_none},
    {"lds_m8_n8", launch_lds<8, 8, 2, 2, false>, launch_lds<8, 8, 2, 2, true>, prepare_none, prepare_none},
    {"dma_m4_n8_s4", launch_dma<4, 8, 2, 2, 4, false>, launch_dma<4, 8, 2, 2, 4, true>,
     prepare_dma<4, 8, 2, 2, 4, false>, prepare_dma<4, 8, 2, 2, 4, true>},
    {"dma_m8_n8_s3", launch_dma<8, 8, 2, 2, 3, false>, launch_dma<8, 8, 2, 2, 3, true>,
     prepare_dma<8, 8, 2, 2, 3, false>, prepare_dma<8, 8, 2, 2, 3, true>},
    {"dma_m8_n8_s4", launch_dma<8, 8, 2, 2, 4, false>, launch_dma<8, 8, 2, 2, 4, true>,
     prepare_dma<8, 8, 2, 2, 4, false>, prepare_dma<8, 8, 2, 2, 4, true>},
    {"dma8_m4_n8_s3", launch_dma<4, 8, 4, 2, 3, false>, launch_dma<4, 8, 4, 2, 3, true>,
     prepare_dma<4, 8, 4, 2, 3, false>, prepare_dma<4, 8, 4, 2, 3, true>},
    {"dma8_m4_n8_s4", launch_dma<4, 8, 4, 2, 4, false>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<4, 8, 4, 2, 4, false>, prepare_dma<4, 8, 4, 2, 4, true>},
    {"dma8_m8_n4_s3", launch_dma<8, 4, 2, 4, 3, false>, launch_dma<8, 4, 2, 4, 3, true>,
     prepare_dma<8, 4, 2, 4, 3, false>, prepare_dma<8, 4, 2, 4, 3, true>},
    {"dma8_m4_n4_s4", launch_dma<4, 4, 4, 2, 4, false>, launch_dma<4, 4, 4, 2, 4, true>,
     prepare_dma<4, 4, 4, 2, 4, false>, prepare_dma<4, 4, 4, 2, 4, true>},
    {"dma8k2_m4_n8_s4", launch_dma<4, 8, 4, 2, 4, false, 2>, launch_dma<4, 8, 4, 2, 4, true, 2>,
     prepare_dma<4, 8, 4, 2, 4, false, 2>, prepare_dma<4, 8, 4, 2, 4, true, 2>},
    {"dma8k2_m8_n4_s4", launch_dma<8, 4, 2, 4, 4, false, 2>, launch_dma<8, 4, 2, 4, 4, true, 2>,
     prepare_dma<8, 4, 2, 4, 4, false, 2>, prepare_dma<8, 4, 2, 4, 4, true, 2>},
    // 16 x 24-tile workgroups (384 accumulator tiles per CU): 17% fewer operand bytes per
    // MFMA than 16 x 16; scan only (packing tiles come in fours), packing runs dma8_m4_n8_s4
    {"dma8_m8_n6_s3", launch_dma<8, 6, 2, 4, 3, false>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<8, 6, 2, 4, 3, false>, prepare_dma<4, 8, 4, 2, 4, true>},
    {"dma8_m8_n6_s4", launch_dma<8, 6, 2, 4, 4, false>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<8, 6, 2, 4, 4, false>, prepare_dma<4, 8, 4, 2, 4, true>},
    {"dma8k2_m8_n6_s4", launch_dma<8, 6, 2, 4, 4, false, 2>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<8, 6, 2, 4, 4, false, 2>, prepare_dma<4, 8, 4, 2, 4, true>},
    {"dma8_m4_n12_s4", launch_dma<8, 6, 2, 4, 4, false>, launch_dma<4, 12, 4, 2, 4, true>,
     prepare_dma<8, 6, 2, 4, 4, false>, prepare_dma<4, 12, 4, 2, 4, true>},
    {"dma8_m4_n12_s3", launch_dma<8, 6, 2, 4, 3, false>, launch_dma<4, 12, 4, 2, 3, true>,
     prepare_dma<8, 6, 2, 4, 3, false>, prepare_dma<4, 12, 4, 2, 3, true>},
    // 32 x 8-tile workgroups: all 128 queries of a batch in one row block, so every database
    // tile is streamed once (25% more operand bytes per MFMA than 16 x 16)
    {"dma8_m4_n8_w81_s4", launch_dma<4, 8, 8, 1, 4, false>, launch_dma<4, 8, 8, 1, 4, true>,
     prepare_dma<4, 8, 8, 1, 4, false>, prepare_dma<4, 8, 8, 1, 4, true>},
    {"dma8_m4_n8_w81_s3", launch_dma<4, 8, 8, 1, 3, false>, launch_dma<4, 8, 8, 1, 3, true>,
     prepare_dma<4, 8, 8, 1, 3, false>, prepare_dma<4, 8, 8, 1, 3, true>},
    // 12-row-tile workgroups (48 queries): less padding for batches of 33-48 queries; scan
    // only (their column tiles do not come in fours), packing runs dma8_m4_n8_s4
    {"dma8_m3_n6_s4", launch_dma<3, 6, 4, 2, 4, false>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<3, 6, 4, 2, 4, false>, prepare_dma<4, 8, 4, 2, 4, true>},
    {"dma8_m3_n10_s4", launch_dma<3, 10, 4, 2, 4, false>, launch_dma<4, 8, 4, 2, 4, true>,
     prepare_dma<3, 10, 4, 2, 4, false>, prepare_dma<4, 8, 4, 2, 4, true>},
    // 8-row-tile workgroups of eight waves (32 queries), for batches of 17-32 queries
    {"dma8_m2_n8_s4", launch_dma<2, 8, 4, 2, 4, false>, launch_dma<2, 8, 4, 2, 4, true>,
     prepare_dma<2, 8, 4, 2, 4, false>, prepare_dma<2, 8, 4, 2, 4, true>},
    {"dma8_m2_n12_s4", launch_dma<2, 12, 4, 2, 4, false>, launch_dma<2, 12, 4, 2, 4, true>,
     prepare_dma<2, 12, 4, 2, 4, false>, prepare_dma<2, 12, 4, 2, 4, true>},
    {"dma8_m4_n8_w24_s4", launch_dma<4, 8, 2, 4, 4, false>, launch_dma<4, 8, 2, 4, 4, true>,
     prepare_dma<4, 8, 2, 4, 4, false>, prepare_dma<4, 8, 2, 4, 4, true>},
    // the same shapes with the matrix work of a k-step issued at raised wave priority
    {"dma8p_m4_n8_s4", launch_dma<4, 8, 4, 2, 4, false, 1, true>, launch_dma<4, 8, 4, 2, 4, true, 1, true>,
     prepare_dma<4, 8, 4, 2, 4, false, 1, true>, prepare_dma<4, 8, 4, 2, 4, true, 1, true>},
    {"dma8k2p_m8_n6_s4", launch_dma<8, 6, 2, 4, 4, false, 2, true>, launch_dma<4, 8, 4, 2, 4, true, 2, true>,
     prepare_dma<8, 6, 2, 4, 4, false, 2, true>, prepare_dma<4, 8, 4, 2, 4, true, 2, true>},
    // A blocks through an LDS ring, each wave's B blocks straight into registers (16 x 16 tiles)
    {"dmab8_m8_n4_u4", launch_dmab<8, 4, 2, 4, 2, 4, false>, launch_dmab<8, 4, 2, 4, 2, 4, true>,
     prepare_dmab<8, 4, 2, 4, 2, 4, false>, prepare_dmab<8, 4, 2, 4, 2, 4, true>},
    {"dmab8_m8_n4_u8", launch_dmab<8, 4, 2, 4, 2, 8, false>, launch_dmab<8, 4, 2, 4, 2, 8, true>,
     prepare_dmab<8, 4, 2, 4, 2, 8, false>, prepare_dmab<8, 4, 2, 4, 2, 8, true>},
    {"dmab8_m8_n4_u16", launch_dmab<8, 4, 2, 4, 2, 16, false>, launch_dmab<8, 4, 2, 4, 2, 16, true>,
     prepare_dmab<8, 4, 2, 4, 2, 16, false>, prepare_dmab<8, 4, 2, 4, 2, 16, true>},
    // the same on 32 x 32 x 32 MFMAs (tile pairs): the same bytes, 25% slower at 1 and 8 GiB
    // (batches of 40-2048)
    {"dmaw8_m8_n4_u16", launch_dmab<8, 4, 2, 4, 2, 16, false, true>, launch_dmab<8, 4, 2, 4, 2, 16, true, true>,
     prepare_dmab<8, 4, 2, 4, 2, 16, false, true>, prepare_dmab<8, 4, 2, 4, 2, 16, true, true>},
    // the same with the A fragments read from LDS just before their MFMAs
    {"dmac8_m8_n4_d2u16", launch_dmac<8, 4, 2, 4, 2, 16, false>, launch_dmac<8, 4, 2, 4, 2, 16, true>,
     prepare_dmac<8, 4, 2, 4, 2, 16, false>, prepare_dmac<8, 4, 2, 4, 2, 16, true>},
    // 12 x 16 tiles, one wave per column of two B tiles (every B block loaded once per
    // workgroup), four k-steps of B loads in flight; scan only (packing on dmab8_m6_n4_u16)
    {"dmac8_w18_m12_n2_d4u16", launch_dmac<12, 2, 1, 8, 4, 16, false>, launch_dmab<6, 4, 2, 4, 2, 16, true>,
     prepare_dmac<12, 2, 1, 8, 4, 16, false>, prepare_dmab<6, 4, 2, 4, 2, 16, true>},
    // the same with 12 x 16 and 8 x 32-tile workgroups (the 8 x 32 one with one wave per
    // column of four B tiles), for 17-48 queries
    {"dmab8_m6_n4_u16", launch_dmab<6, 4, 2, 4, 2, 16, false>, launch_dmab<6, 4, 2, 4, 2, 16, true>,
     prepare_dmab<6, 4, 2, 4, 2, 16, false>, prepare_dmab<6, 4, 2, 4, 2, 16, true>},
    {"dmab8_w18_m8_n4_u16", launch_dmab<8, 4, 1, 8, 2, 16, false>, launch_dmab<8, 4, 1, 8, 2, 16, true>,
     prepare_dmab<8, 4, 1, 8, 2, 16, false>, prepare_dmab<8, 4, 1, 8, 2, 16, true>},
    {"dma2_8w_m4_n8_s4", launch_dma2<4, 8, 4, 2, 4, false>, launch_dma2<4, 8, 4, 2, 4, true>,
     prepare_dma2<4, 8, 4, 2, 4, false>, prepare_dma2<4, 8, 4, 2, 4, true>},
    {"dma2_8w_m4_n8_s3", launch_dma2<4, 8, 4, 2, 3, false>, launch_dma2<4, 8, 4, 2, 3, true>,
     prepare_dma2<4, 8, 4, 2, 3, false>, prepare_dma2<4, 8, 4, 2, 3, true>},
    {"dma2_4w_m8_n8_s4", launch_dma2<8, 8, 2, 2, 4, false>, launch_dma2<8, 8, 2, 2, 4, true>,
     prepare_dma2<8, 8, 2, 2, 4, false>, prepare_dma2<8, 8, 2, 2, 4, true>},
    {"dma2_8w_m8_n4_s4", launch_dma2<8, 4, 2, 4, 4, false>, launch_dma2<8, 4, 2, 4, 4, true>,
     prepare_dma2<8, 4, 2, 4, 4, false>, prepare_dma2<8, 4, 2, 4, 4, true>},
    // 12 x 24 and 12 x 32 tiles, one wave per column of three or four B tiles, A fragments
    // just in time: a third and half the LDS reads per MFMA of the 12 x 16 form; scan only
    // (packing on dmab8_m6_n4_u16)
    {"dmac8_w18_m12_n3_d2u16", launch_dmac<12, 3, 1, 8, 2, 16, false>, launch_dmab<6, 4, 2, 4, 2, 16, true>,
     prepare_dmac<12, 3, 1, 8, 2, 16, false>, prepare_dmab<6, 4, 2, 4, 2, 16, true>},
    {"dmac8_w18_m12_n3_d3u15", launch_dmac<12, 3, 1, 8, 3, 15, false>, launch_dmab<6, 4, 2, 4, 2, 16, true>,
     prepare_dmac<12, 3, 1, 8, 3, 15, false>, prepare_dmab<6, 4, 2, 4, 2, 16, true>},
    {"dmac8_w18_m12_n4_d2u16", launch_dmac<12, 4, 1, 8, 2, 16, false>, launch_dmab<6, 4, 2, 4, 2, 16, true>,
     prepare_dmac<12, 4, 1, 8, 2, 16, false>, prepare_dmab<6, 4, 2, 4, 2, 16, true>},
    // the same, 32 k-steps per loop iteration (the compiler's full wait for loads falls on every
    // 32nd step): 0.3-1.9% faster at batch 128 on 1, 8 and 64 GiB; built-in above 48 queries
    {"dmab8_w18_m8_n4_u32", launch_dmab<8, 4, 1, 8, 2, 32, false>, launch_dmab<8, 4, 1, 8, 2, 32, true>,
     prepare_dmab<8, 4, 1, 8, 2, 32, false>, prepare_dmab<8, 4, 1, 8, 2, 32, true>},
};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));

// Scan and packing in one launch (scan_pack_kernel) for the shape pairs of the built-in
// policy; other pairs run as two launches.
typedef void (*FusedFn)(GemmArgs, const int8_t *, const int8_t *, GemmArgs, const int8_t *, const int8_t *,
                        hipStream_t);

template <int MB, int NBT, int U, int MBp, int NBTp, int Up>
void launch_fused_stream(GemmArgs gs, const int8_t *qa, const int8_t *db, GemmArgs gp, const int8_t *ya,
                         const int8_t *u, hipStream_t st) {
    const int sb = plan_stream<MB, NBT>(gs, false), pb = plan_stream<MBp, NBTp>(gp, true);
    using P = StreamShape<MBp, NBTp, Up, true, false>;
    if (gs.gz > 1)
        hipLaunchKernelGGL((scan_pack_kernel<StreamShape<MB, NBT, U, false, true>, P>), dim3(sb + pb), dim3(256), 0, st,
                           (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
    else
        hipLaunchKernelGGL((scan_pack_kernel<StreamShape<MB, NBT, U, false, false>, P>), dim3(sb + pb), dim3(256), 0,
                           st, (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
}

template <int TM, int TN, int WM, int WN, int S, int KPB, int TMp, int TNp, int WMp, int WNp, int Sp, int KPBp>
struct FusedRing {
    using SA = DmaShape<TM, TN, WM, WN, S, false, true, KPB>;
    using SN = DmaShape<TM, TN, WM, WN, S, false, false, KPB>;
    using P = DmaShape<TMp, TNp, WMp, WNp, Sp, true, false, KPBp>;
    static constexpr size_t kLds = SA::kLds > P::kLds ? SA::kLds : P::kLds;
    static void prepare() {
        (void)hipFuncSetAttribute((const void *)scan_pack_kernel<SA, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLds);
        (void)hipFuncSetAttribute((const void *)scan_pack_kernel<SN, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLds);
    }
    static void launch(GemmArgs gs, const int8_t *qa, const int8_t *db, GemmArgs gp, const int8_t *ya,
                       const int8_t *u, hipStream_t st) {
        const int sb = plan_ring<TM, TN, WM, WN>(gs, false), pb = plan_ring<TMp, TNp, WMp, WNp>(gp, true);
        if (gs.gz > 1)
            hipLaunchKernelGGL((scan_pack_kernel<SA, P>), dim3(sb + pb), dim3(SA::kThreads), kLds, st,
                               (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
        else
            hipLaunchKernelGGL((scan_pack_kernel<SN, P>), dim3(sb + pb), dim3(SN::kThreads), kLds, st,
                               (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
    }
};
// Both contractions on the register-streaming 16 x 16-tile workgroups.
template <int TM, int TN, int WM, int WN, int D, int UNR>
struct FusedRegB {
    using SA = DmabShape<TM, TN, WM, WN, D, UNR, false, true>;
    using SN = DmabShape<TM, TN, WM, WN, D, UNR, false, false>;
    using P = DmabShape<TM, TN, WM, WN, D, UNR, true, false>;
    static constexpr size_t kLds = SA::kLds;
    static void prepare() {
        (void)hipFuncSetAttribute((const void *)scan_pack_kernel<SA, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLds);
        (void)hipFuncSetAttribute((const void *)scan_pack_kernel<SN, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLds);
    }
    static void launch(GemmArgs gs, const int8_t *qa, const int8_t *db, GemmArgs gp, const int8_t *ya,
                       const int8_t *u, hipStream_t st) {
        const int sb = plan_ring<TM, TN, WM, WN>(gs, false), pb = plan_ring<TM, TN, WM, WN>(gp, true);
        if (gs.gz > 1)
            hipLaunchKernelGGL((scan_pack_kernel<SA, P>), dim3(sb + pb), dim3(SA::kThreads), kLds, st,
                               (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
        else
            hipLaunchKernelGGL((scan_pack_kernel<SN, P>), dim3(sb + pb), dim3(SN::kThreads), kLds, st,
                               (const i32x4 *)qa, (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb);
    }
};
using FusedRegB16 = FusedRegB<8, 4, 2, 4, 2, 16>;
using FusedRegBW18 = FusedRegB<8, 4, 1, 8, 2, 16>;
using FusedRegBW18u32 = FusedRegB<8, 4, 1, 8, 2, 32>;
using FusedRing8 = FusedRing<4, 8, 4, 2, 4, 1, 4, 8, 4, 2, 4, 1>;
using FusedRingWide = FusedRing<8, 6, 2, 4, 4, 2, 4, 8, 4, 2, 4, 1>;

struct FusedPair {
    const char *scan, *pack;
    FusedFn launch;
    PrepareFn prepare;
};
const FusedPair kFused[] = {
    {"stream_m1_n4_u4", "stream_m1_n4_u4", launch_fused_stream<1, 4, 4, 1, 4, 4>, prepare_none},
    {"stream_m2_n4_u4", "stream_m2_n4_u4", launch_fused_stream<2, 4, 4, 2, 4, 4>, prepare_none},
    {"stream_m2_n4_u4", "stream_m1_n4_u4", launch_fused_stream<2, 4, 4, 1, 4, 4>, prepare_none},
    {"dma8_m4_n8_s4", "dma8_m4_n8_s4", FusedRing8::launch, FusedRing8::prepare},
    {"dma8k2_m8_n6_s4", "dma8_m4_n8_s4", FusedRingWide::launch, FusedRingWide::prepare},
    {"dmab8_m8_n4_u16", "dmab8_m8_n4_u16", FusedRegB16::launch, FusedRegB16::prepare},
    {"dmab8_w18_m8_n4_u16", "dmab8_w18_m8_n4_u16", FusedRegBW18::launch, FusedRegBW18::prepare},
    {"dmab8_w18_m8_n4_u32", "dmab8_w18_m8_n4_u32", FusedRegBW18u32::launch, FusedRegBW18u32::prepare},
};
constexpr int kNumFused = (int)(sizeof(kFused) / sizeof(kFused[0]));

// Scan, packing and tail in one launch (step_kernel) for the shape pairs of the built-in
// policy at every batch size (the pairs need one workgroup size).  PlanS / PlanP size the
// two grids, BTS / BTP are the B tiles of one x tile of each (a scan B tile is 16 columns,
// a packing one a quarter of that: four byte planes per column tile).
typedef int (*StepFn)(GemmArgs, const int8_t *, const int8_t *, GemmArgs, const int8_t *, const int8_t *,
                      const TailArgs &, StepArgs, hipStream_t);
template <class SA, class SN, class P, int (*PlanS)(GemmArgs &, bool), int (*PlanP)(GemmArgs &, bool), int BTS,
          int BTP>
struct StepPairT {
    static constexpr int kH = SA::kThreads / kTailThreads;
    static constexpr size_t kTileLds = SA::kLds > P::kLds ? SA::kLds : P::kLds;
    static void prepare() {
        (void)hipFuncSetAttribute((const void *)step_kernel<SA, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kMaxLds);
        (void)hipFuncSetAttribute((const void *)step_kernel<SN, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kMaxLds);
    }
    static size_t lds(int n, int part_words) {
        return std::max(kTileLds, step_lds_bytes(kH, n, part_words));
    }
    static int launch(GemmArgs gs, const int8_t *qa, const int8_t *db, GemmArgs gp, const int8_t *ya, const int8_t *u,
                      const TailArgs &ta, StepArgs sa, hipStream_t st) {
        const size_t l = lds(ta.n, ta.part_words);
        if (l > kMaxLds) return -1;
        const int sb = PlanS(gs, false), pb = PlanP(gp, true);
        sa.scan_cols = 16 * BTS;
        sa.pack_cols = 4 * BTP;
        if (gs.gz > 1)
            hipLaunchKernelGGL((step_kernel<SA, P>), dim3(sb + pb), dim3(SA::kThreads), l, st, (const i32x4 *)qa,
                               (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb, ta, sa);
        else
            hipLaunchKernelGGL((step_kernel<SN, P>), dim3(sb + pb), dim3(SN::kThreads), l, st, (const i32x4 *)qa,
                               (const i32x4 *)db, gs, (const i32x4 *)ya, (const i32x4 *)u, gp, sb, ta, sa);
        return 0;
    }
};
template <int MB, int NBT, int U, int MBp, int NBTp, int Up>
using StepStream = StepPairT<StreamShape<MB, NBT, U, false, true>, StreamShape<MB, NBT, U, false, false>,
                             StreamShape<MBp, NBTp, Up, true, false>, plan_stream<MB, NBT>, plan_stream<MBp, NBTp>,
                             4 * NBT, 4 * NBTp>;
using StepRows8 = StepPairT<DmaShape<4, 8, 2, 4, 4, false, true, 1>, DmaShape<4, 8, 2, 4, 4, false, false, 1>,
                            DmaShape<2, 8, 4, 2, 4, true, false, 1>, plan_ring<4, 8, 2, 4>, plan_ring<2, 8, 4, 2>, 32,
                            16>;
using StepRows12c = StepPairT<DmacShape<12, 2, 1, 8, 4, 16, false, true>, DmacShape<12, 2, 1, 8, 4, 16, false, false>,
                              DmabShape<6, 4, 2, 4, 2, 16, true, false>, plan_ring<12, 2, 1, 8>,
                              plan_ring<6, 4, 2, 4>, 16, 16>;
using StepRows12w = StepPairT<DmacShape<12, 4, 1, 8, 2, 16, false, true>, DmacShape<12, 4, 1, 8, 2, 16, false, false>,
                              DmabShape<6, 4, 2, 4, 2, 16, true, false>, plan_ring<12, 4, 1, 8>,
                              plan_ring<6, 4, 2, 4>, 32, 16>;
using StepRows12 = StepPairT<DmaShape<3, 10, 4, 2, 4, false, true, 1>, DmaShape<3, 10, 4, 2, 4, false, false, 1>,
                             DmabShape<6, 4, 2, 4, 2, 16, true, false>, plan_ring<3, 10, 4, 2>, plan_ring<6, 4, 2, 4>,
                             20, 16>;
using StepRegB16 = StepPairT<DmabShape<8, 4, 2, 4, 2, 16, false, true>, DmabShape<8, 4, 2, 4, 2, 16, false, false>,
                             DmabShape<8, 4, 2, 4, 2, 16, true, false>, plan_ring<8, 4, 2, 4>, plan_ring<8, 4, 2, 4>,
                             16, 16>;
using StepRegBW18u32 = StepPairT<DmabShape<8, 4, 1, 8, 2, 32, false, true>, DmabShape<8, 4, 1, 8, 2, 32, false, false>,
                                 DmabShape<8, 4, 1, 8, 2, 32, true, false>, plan_ring<8, 4, 1, 8>, plan_ring<8, 4, 1, 8>,
                                 32, 32>;
using StepRegBW18 = StepPairT<DmabShape<8, 4, 1, 8, 2, 16, false, true>, DmabShape<8, 4, 1, 8, 2, 16, false, false>,
                              DmabShape<8, 4, 1, 8, 2, 16, true, false>, plan_ring<8, 4, 1, 8>, plan_ring<8, 4, 1, 8>,
                              32, 32>;
struct StepPair {
    const char *scan, *pack;
    StepFn launch;
    PrepareFn prepare;
    size_t (*lds)(int, int);
};
const StepPair kStep[] = {
    {"stream_m2_n4_u4", "stream_m1_n4_u4", StepStream<2, 4, 4, 1, 4, 4>::launch, StepStream<2, 4, 4, 1, 4, 4>::prepare,
     StepStream<2, 4, 4, 1, 4, 4>::lds},
    {"stream_m2_n4_u4", "stream_m2_n4_u4", StepStream<2, 4, 4, 2, 4, 4>::launch, StepStream<2, 4, 4, 2, 4, 4>::prepare,
     StepStream<2, 4, 4, 2, 4, 4>::lds},
    {"stream_m1_n4_u4", "stream_m1_n4_u4", StepStream<1, 4, 4, 1, 4, 4>::launch, StepStream<1, 4, 4, 1, 4, 4>::prepare,
     StepStream<1, 4, 4, 1, 4, 4>::lds},
    {"dma8_m4_n8_w24_s4", "dma8_m2_n8_s4", StepRows8::launch, StepRows8::prepare, StepRows8::lds},
    {"dmac8_w18_m12_n2_d4u16", "dmab8_m6_n4_u16", StepRows12c::launch, StepRows12c::prepare, StepRows12c::lds},
    {"dmac8_w18_m12_n4_d2u16", "dmab8_m6_n4_u16", StepRows12w::launch, StepRows12w::prepare, StepRows12w::lds},
    {"dma8_m3_n10_s4", "dmab8_m6_n4_u16", StepRows12::launch, StepRows12::prepare, StepRows12::lds},
    {"dmab8_m8_n4_u16", "dmab8_m8_n4_u16", StepRegB16::launch, StepRegB16::prepare, StepRegB16::lds},
    {"dmab8_w18_m8_n4_u16", "dmab8_w18_m8_n4_u16", StepRegBW18::launch, StepRegBW18::prepare, StepRegBW18::lds},
    {"dmab8_w18_m8_n4_u32", "dmab8_w18_m8_n4_u32", StepRegBW18u32::launch, StepRegBW18u32::prepare,
     StepRegBW18u32::lds},
};
constexpr int kNumStep = (int)(sizeof(kStep) / sizeof(kStep[0]));
// Polls a helper workgroup of the step kernel makes at most while it waits for parts to be
// published (about 1 ms each 1,000; the waits end long before in every correct run).
constexpr int kStepSpin = 1 << 18;

// Built-in shape policy (used until the tuner records a choice for a batch-size bucket),
// from sweeps over batch sizes 1-2048 on 1-64 GiB databases: streaming kernels up to 8
// queries, and for packing up to 16 (the scan of 1-4 queries runs the 8-query streaming
// shape, whose loads issue in a better order: 2% faster at 1-4 GiB, even at 64 GiB); the
// scan of 9-32 queries on 8 x 32-tile workgroups (each streamed query row meets twice the
// database columns of the square tile) and its packing on 8 x 16 tiles; the scan of 33-48
// queries on 12-row tiles and its packing on 12 x 16 tiles that stream the B blocks into
// registers (0.231 against 0.253 ms on 16 x 16 at 8 GiB and 37 queries); above that both
// contractions on 8 x 32-tile workgroups that stream the B blocks into registers, one wave per
// column of four B tiles (every B block loaded by one wave), 16 k-steps per loop iteration:
// against the 16 x 16 form, whose B blocks two waves load, batch 128 4.26 against 4.72 ms per
// batch at 8 GiB, 28.7 against 29.2 ms at 64 GiB, 1.034 against 1.042 ms at 1 GiB, and 1-1.5%
// ahead at 8 GiB batches 500 and 2048 (the 16 x 16 form took 0.432 ms for the 1 GiB scan at
// batch 128 against 0.492 ms for the best all-LDS ring).  The scan of 33-48 queries takes
// 12 x 32-tile workgroups, one wave per column of four B tiles, that read their A fragments
// just before their MFMAs (a quarter KiB of LDS reads per MFMA, against half a KiB on 12 x 16
// tiles and 0.43 on the 12 x 20 ring): at 33, 37, 44 and 48 queries 4-6% faster than the
// 12 x 16 form at 1 GiB and 6-11% faster than the 12 x 20 ring at 8, 32 and 64 GiB (8 GiB,
// 37 queries: 1.53 against 1.69 ms), in alternating order on one box.  The same tiles scan
// more than 256 queries over record ranges of 4,096 k-steps and more (8 GiB and up): 2-9%
// faster than the 8 x 32 tiles at 8, 32 and 64 GiB and 260-2048 queries (64 GiB, 1024
// queries: 205.8 against 213.5 ms); at 1 GiB, and at 128 queries (32 A tiles padded to 36 in
// three row blocks), the 8 x 32 tiles stay ahead.
int variant_index(const char *name);
int default_variant(int which, int MT, int KC) {
    static const int stream1 = variant_index("stream_m1_n4_u4"), stream2 = variant_index("stream_m2_n4_u4"),
                     stream4 = variant_index("stream_m4_n4_u2"), rows8 = variant_index("dma8_m4_n8_w24_s4"),
                     rows8p = variant_index("dma8_m2_n8_s4"),
                     regb = variant_index("dmab8_w18_m8_n4_u32"), regb12 = variant_index("dmab8_m6_n4_u16"),
                     regc12w = variant_index("dmac8_w18_m12_n4_d2u16");
    if (MT <= 2) return which == 1 && MT == 1 ? stream1 : stream2;
    if (which == 1) return MT <= 4 ? stream4 : MT <= 8 ? rows8p : MT <= 12 ? regb12 : regb;
    if (MT <= 8) return rows8;
    if (MT <= 12 || (MT > 64 && KC >= 4096)) return regc12w;
    return regb;
}
int variant_index(const char *name) {
    for (int i = 0; i < kNumVariants; i++)
        if (!strcmp(kVariants[i].name, name)) return i;
    return 0;
}

// Index into kStep of the pair (scan variant vs, packing variant vp), or -1.
int step_index(int vs, int vp) {
    for (int i = 0; i < kNumStep; i++)
        if (variant_index(kStep[i].scan) == vs && variant_index(kStep[i].pack) == vp) return i;
    return -1;
}

// Index into kFused of the pair (scan variant vs, packing variant vp), or -1.
int fused_index(int vs, int vp) {
    for (int i = 0; i < kNumFused; i++)
        if (variant_index(kFused[i].scan) == vs && variant_index(kFused[i].pack) == vp) return i;
    return -1;
}

// Batch-size buckets of the tuning table, in half octaves of A tiles (4 queries each):
// bucket b holds MT up to bucket_top(b) = 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, ... (2^k and
// 3 * 2^k tiles), so that the 12-, 16-, 24- and 32-tile row blocks of the ring kernels each
// end a bucket; the last bucket ends at 2^15 tiles and takes every larger batch.
constexpr int kBuckets = 30;
constexpr long bucket_top(int b) { return b < 2 ? b + 1 : (b & 1) ? 2L << (b >> 1) : 3L << ((b >> 1) - 1); }
int bucket_of(int MT) {
    int b = 0;
    while (b + 1 < kBuckets && bucket_top(b) < MT) b++;
    return b;
}
constexpr int kMaxBucketTiles = 512;  // bucket graphs up to 2048 queries
// Queries a workspace needs so that every batch of up to B queries runs without growing it:
// B rounded up to the top of its bucket where bucket graphs are captured (those are sized for
// the bucket's largest batch), else to a multiple of 4.
size_t bucket_queries(size_t B) {
    const size_t MT = (std::max<size_t>(B, 1) + 3) / 4;
    if (MT > (size_t)kMaxBucketTiles) return 4 * MT;
    return 4 * (size_t)bucket_top(bucket_of((int)MT));
}

// -------------------------------------------------------------------- engine --

struct Context;

// Host threads for the copies of the synchronous entry point (as many as the `threads` the
// caller gives the engine, at most 16): run(n, f) calls f(0) .. f(n - 1) on the workers and the
// calling thread and returns when all are done.
class HostPool {
  public:
    explicit HostPool(int threads) {
        for (int i = 1; i < threads; i++) workers_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    void run(size_t n, const std::function<void(size_t)> &f) {
        if (workers_.empty() || n < 2) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &f;
            n_ = n;
            next_ = 0;
            active_ = (int)workers_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return active_ == 0; });
    }

  private:
    void work() {
        for (size_t i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
    }
    void loop() {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    int active_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

// Host side of the synchronous entry point: a ring of pinned buffers through which the host
// threads stage the frames, a chunk at a time (the DMA copy of one chunk runs while the next is
// staged), and mapped pinned memory that the tail writes the responses into.
struct HostStage {
    static constexpr int kBufs = 4;
    HostPool pool;
    size_t buf_bytes = 0;
    uint8_t *buf[kBufs] = {};
    hipEvent_t copied[kBufs] = {};  // recorded after the DMA copy that reads buf[i]
    bool pending[kBufs] = {};
    uint8_t *resp = nullptr, *resp_dev = nullptr;
    size_t resp_bytes = 0;
    explicit HostStage(int threads) : pool(threads) {}
    ~HostStage() {
        for (int i = 0; i < kBufs; i++) {
            if (pending[i]) (void)hipEventSynchronize(copied[i]);
            if (copied[i]) (void)hipEventDestroy(copied[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
        if (resp) (void)hipHostFree(resp);
    }
    bool reserve(size_t chunk, size_t rbytes) {
        bool ok = true;
        if (chunk > buf_bytes) {
            for (int i = 0; i < kBufs; i++) {
                if (pending[i]) (void)hipEventSynchronize(copied[i]);
                pending[i] = false;
                if (buf[i]) (void)hipHostFree(buf[i]);
                buf[i] = nullptr;
            }
            buf_bytes = 0;
            for (int i = 0; ok && i < kBufs; i++) {
                ok = HIP_OK(hipHostMalloc((void **)&buf[i], chunk, hipHostMallocDefault));
                if (ok && !copied[i]) ok = HIP_OK(hipEventCreateWithFlags(&copied[i], hipEventDisableTiming));
            }
            if (!ok) return false;
            buf_bytes = chunk;
        }
        if (rbytes > resp_bytes) {
            if (resp) (void)hipHostFree(resp);
            resp = resp_dev = nullptr;
            resp_bytes = 0;
            ok = HIP_OK(hipHostMalloc((void **)&resp, rbytes, hipHostMallocMapped)) &&
                 HIP_OK(hipHostGetDevicePointer((void **)&resp_dev, resp, 0));
            if (!ok) return false;
            resp_bytes = rbytes;
        }
        return true;
    }
};

struct Engine {
    int device = 0;
    size_t ld = 0, rows = 0, cols = 0, n = 0, tg = 0, kp = 0, outputs = 0;
    size_t part_len = 0, part_words = 0, row1_off = 0;
    uint64_t q1 = 0;
    int q1_bits = 0, threads = 1, logn = 0;
    bool tail_ok = false;
    size_t KC = 0, KCp = 0, CT = 0;
    int8_t *db = nullptr;        // CT x KC blocks
    int8_t *u = nullptr;         // CT x KCp x 4 blocks
    uint32_t *ucorr = nullptr;   // cols
    uint32_t *that = nullptr;    // outputs x kp
    uint32_t *wtab = nullptr;    // n
    uint8_t *prefix = nullptr;   // outputs x part_words * 4
    size_t device_bytes = 0;
    uint32_t k128e = 0, ycorr0 = 0;
    int tuned[2][kBuckets];      // [scan, pack][bucket] -> variant, -1 = built-in policy
    bool use_graphs = true;
    int fuse = -1;               // -1 built-in, 0 never, 1 scan and packing in one launch for every pair,
                                 // 2 scan, packing and tail in one launch (step pairs; else as 1)
    bool time_graphs = false;    // graph replays record the stage events too
    Context *ctx0 = nullptr;     // the context of alto_answer / alto_answer_device
    HostStage *host = nullptr;   // host staging of alto_answer (created on its first call)

    Engine() {
        for (auto &t : tuned)
            for (int &v : t) v = -1;
    }
    ~Engine();
    int variant_for(int which, int MT) const {
        const int v = tuned[which][bucket_of(MT)];
        return v >= 0 ? v : default_variant(which, MT, (int)(which ? KCp : KC));
    }
    // kFused index used for a batch of MT A tiles, or -1 for two launches
    // Built-in: fused only for batches of one A tile (1-4 queries) on databases of at most
    // 4 KP records, whose scan streams no more than twice the bytes of the packing matrix:
    // there one launch of the two streaming contractions measured 4% faster than two (1 GiB:
    // 0.325 against 0.339 ms); at 8 and 64 GiB it was 2% and 6% slower (the packing tiles
    // then interleave with a much longer scan), at two A tiles 8% slower, from 16 A tiles on
    // level.
    int fused_for(int MT) const {
        if (fuse == 0) return -1;
        if (fuse == 2 || (fuse < 0 && step_builtin(MT))) {
            const int k = step_for(MT);
            if (k >= 0) return kNumFused + k;
        }
        if (fuse < 0 && (MT != 1 || rows > 4 * kp)) return -1;
        return fused_index(variant_for(0, MT), variant_for(1, MT));
    }
    // Built-in use of the one-launch step: never (measured per batch size; see step_for).
    bool step_builtin(int) const { return false; }
    // kStep index for a batch of MT A tiles, or -1: the pair of the chosen scan and packing
    // variants; when the packing variant is the built-in one and has no step pair with the
    // scan (its workgroup size differs), the first step pair of that scan variant.
    int step_for(int MT) const {
        const int vs = variant_for(0, MT), vp = variant_for(1, MT);
        int k = step_index(vs, vp);
        if (k < 0 && tuned[1][bucket_of(MT)] < 0)
            for (int i = 0; i < kNumStep && k < 0; i++)
                if (variant_index(kStep[i].scan) == vs) k = i;
        if (k >= 0 && kStep[k].lds((int)n, (int)part_words) > kMaxLds) return -1;
        return k;
    }
    size_t request_len() const { return 9 + 4 * (2 * kp + rows); }
    size_t response_len() const { return outputs * part_len; }
};

// ------------------------------------------------------------------ contexts --

// Per-batch device workspace of one batch of up to `cap` queries (cap a multiple of 4).
struct Workspace {
    int8_t *qa;                // scan limbs, cap * 4 * rows bytes
    int8_t *ya;                // packing limbs, cap * 4 * KP bytes
    uint32_t *acc;             // scan accumulators, cap * C
    uint32_t *beta;            // packing results, cap * C
    uint32_t *qsum;            // cap
    unsigned long long *ysum;  // cap
    uint32_t *sync;            // 3 * outputs + 4 words: part arrivals and tail claims of the step kernel
};

// Bytes of the workspace of a batch of Bp queries (Bp a multiple of 4), each buffer starting
// on a 256-byte boundary, and its layout inside one block of memory at `base`.
size_t workspace_bytes(const Engine &e, size_t Bp);
Workspace workspace_at(const Engine &e, size_t Bp, uint8_t *base);
// Words of the step kernel's synchronisation area (StepArgs::sync).
size_t sync_words(const Engine &e) { return 3 * e.outputs + 4; }

// One batch in flight: its workspace, stream, timing events and captured graphs.  Batches
// on different contexts of one engine may run concurrently; one context runs one batch at
// a time.
struct Context {
    Engine *e = nullptr;
    size_t cap = 0;              // queries (multiple of 4)
    int8_t *qa = nullptr;        // scan limbs, cap * 4 * rows bytes
    int8_t *ya = nullptr;        // packing limbs, cap * 4 * KP bytes
    uint32_t *acc = nullptr;     // cap * C
    uint32_t *beta = nullptr;    // cap * C
    uint32_t *qsum = nullptr;    // cap
    unsigned long long *ysum = nullptr;
    uint32_t *sync = nullptr;    // 3 * outputs + 4 (step kernel)
    // arenas of the synchronous entry point (allocated on first use)
    size_t slot_stride = 0;
    uint8_t *slots = nullptr, *out = nullptr;
    int *status = nullptr;
    hipStream_t stream = nullptr;
    // Timing events of eager batches: a ring of event sets, event i closing stage i - 1 of
    // (validate, scan prep, scan, pack prep, pack, tail) after event 0 at the start, so that
    // back-to-back batches are timed without the host waiting for each one; a set is folded
    // into the running sums before it is reused.  A batch records only the events of its
    // launch boundaries (ring_mask: bit i = event i recorded); the time between two recorded
    // events counts for the later stage and the skipped stages read zero.  Both limb
    // preparations run in one launch (under "scan prep"), so event 4 is never recorded; with
    // the one-launch contractions neither is event 5 (the pair reads as "scan").
    static constexpr int kRing = 32;
    static constexpr int kGraphRing = 4;  // sets cycled by timed graph replays (one graph per set)
    static constexpr uint8_t kMaskSplit = 0x6f, kMaskFused = 0x4f;
    hipEvent_t ring[kRing][7] = {};
    bool ring_pending[kRing] = {};
    uint8_t ring_mask[kRing] = {};
    int ring_head = -1;          // set of the most recent timed batch
    hipEvent_t *ev = nullptr;    // = ring[ring_head]
    double stage_sum[6] = {};    // ms, over the folded sets
    long long stage_count = 0;
    bool timed_valid = false;
    void record(hipEvent_t x, hipStream_t st) { (void)hipEventRecord(x, st); }
    bool add_stage_events(hipGraph_t graph, int plan, int set);
    std::vector<int> hstatus;
    // captured batches, replayed when the same batch (pointers, size, stream, kernel
    // variants) comes again
    // A bucket graph (bucket >= 0) is captured for the largest batch of its bucket (B = 4 *
    // bucket_top) behind a batch_dims_kernel node, whose arguments every replay sets to the
    // batch's own {B, MT}: the kernels read them and skip the rows past the batch, so one
    // graph answers every batch size of the bucket.  Batches beyond kMaxBucketTiles tiles,
    // and timed replays, are captured per exact size (bucket = -1).
    struct GraphEntry {
        FrameSet fs;
        size_t B;   // batch the chain is sized for
        uint8_t *out;
        size_t out_stride;
        const int *oidx;
        int *status;
        hipStream_t st;
        int vs, vp, fz;
        int ring;  // timing event set recorded by the graph, -1 for none
        int bucket;
        hipGraph_t graph;            // kept for the node updates of a bucket graph
        hipGraphNode_t dims_node;
        hipKernelNodeParams dims_params;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<GraphEntry> graphs;
    int graph_mode = -1;  // -1: the engine's setting, 0: eager, 1: capture and replay
    uint64_t tick = 0;
    static constexpr size_t kMaxGraphs = 64;
    int *dims = nullptr;  // device {B, MT} of the batch a bucket graph answers
    hipEvent_t null_in = nullptr, null_out = nullptr;  // joins of a null-stream call

    bool init(Engine *eng) {
        e = eng;
        bool ok = HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) &&
                  HIP_OK(hipMalloc(&dims, 2 * sizeof(int))) &&
                  HIP_OK(hipEventCreateWithFlags(&null_in, hipEventDisableTiming)) &&
                  HIP_OK(hipEventCreateWithFlags(&null_out, hipEventDisableTiming));
        for (auto &set : ring)
            // device-scope release: the stage events only timestamp the stream (nothing reads
            // device memory on the host through them), so recording one need not write the L2
            // back to memory the way a system-scope release does
            for (auto &x : set) ok = ok && HIP_OK(hipEventCreateWithFlags(&x, hipEventReleaseToDevice));
        ev = ring[0];
        return ok;
    }
    // Adds the stage times of ring set k (waiting for it) to the running sums.
    bool fold(int k) {
        if (!ring_pending[k]) return true;
        ring_pending[k] = false;
        double t[6];
        if (!stage_times(k, t)) return false;
        for (int i = 0; i < 6; i++) stage_sum[i] += t[i];
        stage_count++;
        return true;
    }
    // The six stage times (ms) of event set k, waiting for its last event.
    bool stage_times(int k, double *t) {
        if (!HIP_OK(hipEventSynchronize(ring[k][6]))) return false;
        for (int i = 1, from = 0; i < 7; i++) {
            t[i - 1] = 0;
            if (!(ring_mask[k] >> i & 1)) continue;
            float f = 0;
            (void)hipEventElapsedTime(&f, ring[k][from], ring[k][i]);
            t[i - 1] = f;
            from = i;
        }
        return true;
    }
    // Moves to the next event set for a timed batch (folding what it held).
    hipEvent_t *next_set(int span = kRing) {
        ring_head = (ring_head + 1) % span;
        (void)fold(ring_head);
        ring_pending[ring_head] = true;
        ev = ring[ring_head];
        return ev;
    }
    bool fold_all() {
        bool ok = true;
        for (int k = 0; k < kRing; k++) ok = fold(k) && ok;
        return ok;
    }
    // A batch asked for on the null stream runs on the context's own stream (the null stream
    // cannot be captured into a graph), joined to the null stream on both sides: it starts
    // after the work enqueued there before the call, and the work enqueued there after the
    // call starts after the batch.
    bool join_null_before() {
        return HIP_OK(hipEventRecord(null_in, nullptr)) && HIP_OK(hipStreamWaitEvent(stream, null_in, 0));
    }
    bool join_null_after() {
        return HIP_OK(hipEventRecord(null_out, stream)) && HIP_OK(hipStreamWaitEvent(nullptr, null_out, 0));
    }
    static void destroy(GraphEntry &g) {
        (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    void drop_graphs() {
        for (auto &g : graphs) destroy(g);
        graphs.clear();
    }
    void release_buffers() {
        for (void *p : {(void *)qa, (void *)ya, (void *)acc, (void *)beta, (void *)qsum, (void *)ysum, (void *)sync,
                        (void *)slots,
                        (void *)out, (void *)status})
            if (p) (void)hipFree(p);
        qa = ya = nullptr;
        acc = beta = qsum = nullptr;
        ysum = nullptr;
        sync = nullptr;
        slots = out = nullptr;
        status = nullptr;
        cap = 0;
    }
    ~Context() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        release_buffers();
        for (auto &set : ring)
            for (auto &x : set)
                if (x) (void)hipEventDestroy(x);
        if (dims) (void)hipFree(dims);
        for (hipEvent_t x : {null_in, null_out})
            if (x) (void)hipEventDestroy(x);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Grows the per-batch buffers to hold `B` queries; host_arenas also sizes the request
    // and response arenas of the synchronous entry point.
    bool reserve(size_t B, bool host_arenas) {
        const size_t Bp = round_up(std::max<size_t>(B, 1), 4);
        const bool grow = Bp > cap;
        const bool add_host = host_arenas && !slots;
        if (!grow && !add_host) return true;
        if (cap || slots) {
            // old buffers are freed below: a caller stream may still be using them.  A fresh
            // context has none and synchronises nothing, so that admitting one while another
            // thread captures a graph on its own stream cannot disturb that capture.
            (void)hipStreamSynchronize(stream);
            (void)hipDeviceSynchronize();
        }
        drop_graphs();  // they hold the old workspace addresses
        const size_t ncap = grow ? round_up(std::max(Bp, cap + cap / 2), 4) : cap;
        const bool with_host = host_arenas || slots;
        release_buffers();
        cap = ncap;
        slot_stride = round_up(ALTO_FRAME_OFFSET + e->request_len(), 256);
        bool ok = HIP_OK(hipMalloc(&qa, cap * 4 * e->rows)) && HIP_OK(hipMalloc(&ya, cap * 4 * e->kp)) &&
                  HIP_OK(hipMalloc(&acc, cap * e->cols * 4)) && HIP_OK(hipMalloc(&beta, cap * e->cols * 4)) &&
                  HIP_OK(hipMalloc(&qsum, cap * 4)) && HIP_OK(hipMalloc(&ysum, cap * 8)) &&
                  HIP_OK(hipMalloc(&sync, sync_words(*e) * 4));
        if (ok && with_host)
            ok = HIP_OK(hipMalloc(&slots, cap * slot_stride)) && HIP_OK(hipMalloc(&out, cap * e->response_len())) &&
                 HIP_OK(hipMalloc(&status, cap * 4));
        if (!ok) release_buffers();
        return ok;
    }

    Workspace ws() const { return Workspace{qa, ya, acc, beta, qsum, ysum, sync}; }
    int enqueue(const Workspace &w, FrameSet fs, size_t B, uint8_t *out, size_t out_stride, const int *oidx,
                int *status, hipStream_t st, bool timed, bool record_start = true);
    int answer_device(FrameSet fs, size_t B, uint8_t *out, size_t out_stride, const int *oidx, int *status,
                      hipStream_t st);
};

Engine::~Engine() {
    (void)hipSetDevice(device);
    delete ctx0;
    delete host;
    for (void *p : {(void *)db, (void *)u, (void *)ucorr, (void *)that, (void *)wtab, (void *)prefix})
        if (p) (void)hipFree(p);
}

size_t workspace_bytes(const Engine &e, size_t Bp) {
    return round_up(Bp * 4 * e.rows, 256) + round_up(Bp * 4 * e.kp, 256) + 2 * round_up(Bp * e.cols * 4, 256) +
           round_up(Bp * 4, 256) + round_up(Bp * 8, 256) + round_up(sync_words(e) * 4, 256);
}

Workspace workspace_at(const Engine &e, size_t Bp, uint8_t *base) {
    Workspace w;
    uint8_t *p = base;
    auto take = [&](size_t bytes) {
        uint8_t *q = p;
        p += round_up(bytes, 256);
        return q;
    };
    w.qa = (int8_t *)take(Bp * 4 * e.rows);
    w.ya = (int8_t *)take(Bp * 4 * e.kp);
    w.acc = (uint32_t *)take(Bp * e.cols * 4);
    w.beta = (uint32_t *)take(Bp * e.cols * 4);
    w.qsum = (uint32_t *)take(Bp * 4);
    w.ysum = (unsigned long long *)take(Bp * 8);
    w.sync = (uint32_t *)take(sync_words(e) * 4);
    return w;
}

// The launch chain of one batch on stream st with workspace w: validate, limb preparation,
// the two contractions (or both in one launch), tail.  ev: the stage events to record at the
// launch boundaries (null: none); *fused is set when both contractions ran in one launch.
int launch_batch(const Engine &E, const Workspace &w, FrameSet fs, size_t B, uint8_t *outp, size_t out_stride,
                 const int *oidx, int *st_dev, hipStream_t st, hipEvent_t *ev, bool record_start, int *plan,
                 const int *dims = nullptr) {
    int8_t *const qa = w.qa, *const ya = w.ya;
    uint32_t *const acc = w.acc, *const beta = w.beta, *const qsum = w.qsum;
    unsigned long long *const ysum = w.ysum;
    const size_t Bp = round_up(B, 4), MT = Bp / 4;
    if (ev && record_start) (void)hipEventRecord(ev[0], st);
    hipLaunchKernelGGL(validate_kernel, dim3((unsigned)B, validate_slices(E.kp)), dim3(256), 0, st, fs, st_dev,
                       (uint32_t)E.kp, qsum, ysum, dims);
    if (ev) (void)hipEventRecord(ev[1], st);
    GemmArgs gs{};  // scan: query limbs x database -> acc
    gs.KC = (int)E.KC;
    gs.MT = (int)MT;
    gs.NT = (int)E.CT;
    gs.C = (int)E.cols;
    gs.out = acc;
    gs.dims = dims;
    GemmArgs gp{};  // packing: key limbs x packing matrix -> beta
    gp.KC = (int)E.KCp;
    gp.MT = (int)MT;
    gp.NT = (int)(4 * E.CT);
    gp.C = (int)E.cols;
    gp.out = beta;
    gp.ucorr = E.ucorr;
    gp.ysum = ysum;
    gp.k128e = E.k128e;
    gp.ycorr0 = E.ycorr0;
    gp.dims = dims;
    // one launch prepares both limb sets and clears the scan accumulators
    PrepArgs pa;
    pa.fs = fs;
    pa.status = st_dev;
    pa.B = (int)B;
    pa.scan_off = 2 * E.kp;
    pa.KC = E.KC;
    pa.KCp = E.KCp;
    pa.qa = qa;
    pa.ya = ya;
    pa.qsum = qsum;
    pa.ysum = ysum;
    pa.acc = (uint4 *)acc;
    pa.acc4 = Bp * E.cols / 4;
    pa.sx = (unsigned)((E.KC * 4 + 255) / 256);
    pa.sy = (unsigned)(Bp / kPrepClients);
    pa.px = (unsigned)((E.KCp * 4 + 255) / 256);
    pa.py = (unsigned)(Bp / kPrepClients);
    pa.nfill = (unsigned)std::min<size_t>(1024, (pa.acc4 + 1023) / 1024);
    pa.sync = w.sync;
    pa.nsync = (unsigned)sync_words(E);
    pa.dims = dims;
    const unsigned prep_blocks = pa.sx * pa.sy + pa.px * pa.py + pa.nfill;
    hipLaunchKernelGGL(prep_kernel, dim3(prep_blocks), dim3(256), 0, st, pa);
    if (ev) (void)hipEventRecord(ev[2], st);
    const int fz = E.fused_for((int)MT);
    *plan = fz;
    TailArgs ta;
    ta.fs = fs;
    ta.status = st_dev;
    ta.acc = acc;
    ta.qsum = qsum;
    ta.beta = beta;
    ta.that = E.that;
    ta.wtab = E.wtab;
    ta.prefix = E.prefix;
    ta.out = outp;
    ta.out_stride = out_stride;
    ta.oidx = oidx;
    ta.n = (int)E.n;
    ta.logn = E.logn;
    ta.tg = (int)E.tg;
    ta.C = (int)E.cols;
    ta.kp = (int)E.kp;
    ta.part_len = (int)E.part_len;
    ta.part_words = (int)E.part_words;
    ta.row1_off = (long long)E.row1_off;
    ta.q1_bits = E.q1_bits;
    ta.q1 = (uint32_t)E.q1;
    ta.dims = dims;
    if (fz >= kNumFused) {
        // scan, packing and tail in one launch; the stage times show it under "scan"
        StepArgs sa{w.sync, (int)E.outputs, (int)E.n, 0, 0, (int)B, kStepSpin};
        if (kStep[fz - kNumFused].launch(gs, qa, E.db, gp, ya, E.u, ta, sa, st) != 0) {
            set_err("step kernel: LDS of the tail tasks exceeds the workgroup limit");
            return -1;
        }
        if (ev) {
            (void)hipEventRecord(ev[3], st);
            (void)hipEventRecord(ev[6], st);
        }
    } else {
        if (fz >= 0) {
            // one launch for both contractions; the stage times show the pair under "scan"
            // (packing preparation and packing read as zero)
            kFused[fz].launch(gs, qa, E.db, gp, ya, E.u, st);
            if (ev) (void)hipEventRecord(ev[3], st);
        } else {
            kVariants[E.variant_for(0, (int)MT)].scan(gs, qa, E.db, st);
            if (ev) (void)hipEventRecord(ev[3], st);
            kVariants[E.variant_for(1, (int)MT)].pack(gp, ya, E.u, st);
            if (ev) (void)hipEventRecord(ev[5], st);
        }
        const size_t lds = tail_lds_bytes(E.n, E.part_words);
        hipLaunchKernelGGL(tail_kernel, dim3((unsigned)E.outputs, (unsigned)B), dim3(kTailThreads), lds, st, ta);
        if (ev) (void)hipEventRecord(ev[6], st);
    }
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        set_err("kernel launch failed: %s", hipGetErrorString(le));
        return -1;
    }
    return 0;
}

int Context::enqueue(const Workspace &w, FrameSet fs, size_t B, uint8_t *outp, size_t out_stride, const int *oidx,
                     int *st_dev, hipStream_t st, bool timed, bool record_start) {
    int plan = -1;
    const int rc = launch_batch(*e, w, fs, B, outp, out_stride, oidx, st_dev, st, timed ? ev : nullptr, record_start,
                                &plan);
    if (timed) ring_mask[ring_head] = plan >= 0 ? kMaskFused : kMaskSplit;
    timed_valid = timed;
    return rc;
}

// Inserts the stage events of a batch into its captured launch chain as event-record nodes,
// at the positions where an eager batch records them: the chain is validate, prep, scan,
// packing, tail (fused: validate, prep, scan + packing, tail).
bool Context::add_stage_events(hipGraph_t graph, int plan, int set) {
    const bool fused = plan >= 0, step = plan >= kNumFused;
    hipEvent_t *evs = ring[set];
    std::vector<hipGraphNode_t> chain;
    hipGraphNode_t node = nullptr;
    size_t n = 1;
    if (!HIP_OK(hipGraphGetRootNodes(graph, &node, &n)) || n != 1) return false;
    while (true) {
        chain.push_back(node);
        size_t nd = 0;
        if (!HIP_OK(hipGraphNodeGetDependentNodes(node, nullptr, &nd))) return false;
        if (nd == 0) break;
        if (nd != 1 || chain.size() > 16) return false;
        if (!HIP_OK(hipGraphNodeGetDependentNodes(node, &node, &nd))) return false;
    }
    const int want = step ? 3 : fused ? 4 : 5;
    if ((int)chain.size() != want) {
        set_err("graph timing: %d nodes in the captured batch, expected %d", (int)chain.size(), want);
        return false;
    }
    // event k is recorded after chain node after[k] (-1: before the first, -2: not recorded)
    static const int after_split[7] = {-1, 0, 1, 2, -2, 3, 4}, after_fused[7] = {-1, 0, 1, 2, -2, -2, 3},
                     after_step[7] = {-1, 0, 1, 2, -2, -2, 2};
    const int *after = step ? after_step : fused ? after_fused : after_split;
    ring_mask[set] = fused ? kMaskFused : kMaskSplit;
    for (size_t i = 0; i + 1 < chain.size(); i++)
        if (!HIP_OK(hipGraphRemoveDependencies(graph, &chain[i], &chain[i + 1], 1))) return false;
    hipGraphNode_t prev = nullptr;
    int recorded = 0;
    for (int pos = -1; pos < (int)chain.size(); pos++) {
        if (pos >= 0) {
            if (prev && !HIP_OK(hipGraphAddDependencies(graph, &prev, &chain[pos], 1))) return false;
            prev = chain[pos];
        }
        for (int k = 0; k < 7; k++) {
            if (after[k] != pos) continue;
            hipGraphNode_t ev_node = nullptr;
            if (!HIP_OK(hipGraphAddEventRecordNode(&ev_node, graph, prev ? &prev : nullptr, prev ? 1 : 0, evs[k])))
                return false;
            prev = ev_node;
            recorded++;
        }
    }
    return recorded == __builtin_popcount(ring_mask[set]);
}

// The device entry points: eager with stage events when graphs are off, otherwise captured
// once per distinct batch and replayed.
int Context::answer_device(FrameSet fs, size_t B, uint8_t *outp, size_t out_stride, const int *oidx, int *st_dev,
                           hipStream_t st) {
    if (graph_mode == 0 || (graph_mode < 0 && !e->use_graphs)) {
        next_set();
        return enqueue(ws(), fs, B, outp, out_stride, oidx, st_dev, st, true);
    }
    const int MT = (int)(round_up(B, 4) / 4);
    // timed replays: the batch is captured once per timing event set, each graph recording
    // its set's events, and the replays cycle through kGraphRing sets
    const bool tg = e->time_graphs;
    const int bk = bucket_of(MT);
    const bool bucketed = !tg && bucket_top(bk) <= kMaxBucketTiles;
    const int MTg = bucketed ? (int)bucket_top(bk) : MT;  // A tiles the launch chain is sized for
    const size_t Bg = bucketed ? 4 * (size_t)MTg : B;
    if (bucketed && !reserve(Bg, false)) return -4;
    // the built-in policy and the tuning table choose per bucket, so a bucket graph's variants
    // are those of every batch of the bucket
    const int vs = e->variant_for(0, MTg), vp = e->variant_for(1, MTg), fz = e->fused_for(MTg);
    int rs = -1;
    if (tg) {
        next_set(kGraphRing);
        rs = ring_head;
    }
    timed_valid = tg;
    auto set_dims = [&](GraphEntry &g) {
        int *d = dims;
        int b = (int)B, mt = MT;
        void *args[] = {&d, &b, &mt};
        hipKernelNodeParams kp = g.dims_params;
        kp.kernelParams = args;
        kp.extra = nullptr;
        return HIP_OK(hipGraphExecKernelNodeSetParams(g.exec, g.dims_node, &kp));
    };
    for (auto &g : graphs)
        if (g.fs.slots == fs.slots && g.fs.stride == fs.stride && g.fs.sidx == fs.sidx && g.B == Bg && g.out == outp &&
            g.out_stride == out_stride && g.oidx == oidx && g.status == st_dev && g.st == st && g.vs == vs &&
            g.vp == vp && g.fz == fz && g.ring == rs && g.bucket == (bucketed ? bk : -1)) {
            g.used = ++tick;
            if (rs >= 0) ring_mask[rs] = g.fz >= 0 ? kMaskFused : kMaskSplit;
            if (bucketed && !set_dims(g)) return -1;
            return HIP_OK(hipGraphLaunch(g.exec, st)) ? 0 : -1;
        }
    if (graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(graphs.begin(), graphs.end(),
                                    [](const GraphEntry &a, const GraphEntry &b) { return a.used < b.used; });
        destroy(*lru);
        graphs.erase(lru);
    }
    hipGraph_t graph = nullptr;
    if (!HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) return -1;
    int rc;
    if (bucketed) {
        hipLaunchKernelGGL(batch_dims_kernel, dim3(1), dim3(64), 0, st, dims, (int)Bg, MTg);
        int plan = -1;
        rc = launch_batch(*e, ws(), fs, Bg, outp, out_stride, oidx, st_dev, st, nullptr, false, &plan, dims);
    } else {
        rc = enqueue(ws(), fs, B, outp, out_stride, oidx, st_dev, st, false);
    }
    const bool ended = HIP_OK(hipStreamEndCapture(st, &graph));
    if (rc || !ended || (tg && !add_stage_events(graph, fz, rs))) {
        if (graph) (void)hipGraphDestroy(graph);
        return -1;
    }
    GraphEntry g{fs, Bg, outp, out_stride, oidx, st_dev, st, vs, vp, fz, rs, bucketed ? bk : -1, nullptr, nullptr, {},
                 nullptr, 0};
    if (bucketed) {
        // the chain's root is the dims node
        size_t nroot = 1;
        if (!HIP_OK(hipGraphGetRootNodes(graph, &g.dims_node, &nroot)) || nroot != 1 ||
            !HIP_OK(hipGraphKernelNodeGetParams(g.dims_node, &g.dims_params))) {
            set_err("bucket graph: no dims node at the root of the captured chain");
            (void)hipGraphDestroy(graph);
            return -1;
        }
    }
    const bool inst = HIP_OK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    if (bucketed && inst) g.graph = graph;  // the dims node's graph stays alive for the updates
    else (void)hipGraphDestroy(graph);
    if (!inst) return -1;
    g.used = ++tick;
    graphs.push_back(g);
    if (bucketed && !set_dims(graphs.back())) return -1;
    return HIP_OK(hipGraphLaunch(g.exec, st)) ? 0 : -1;
}

// out[i * bc + c] += acc[c][i] + 128 * sum(b_c)   (the scan's offset correction, mod 2^32)
__global__ void packed_finish_kernel(uint32_t *out, const uint32_t *acc, const uint32_t *qsum, size_t rows,
                                     size_t bc, size_t C) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * bc) return;
    const size_t i = t / bc, c = t % bc;
    out[t] += acc[c * C + i] + (qsum[c] << 7);
}

// The reference's SimplePIR packed product (src/matmul.cpp:105-218, 220-388, 391-674,
// 1548-1559): out[i * bc + c] += sum_j sum_t byte_t(a[i * aCols + j]) * b_c[4 j + t] (mod
// 2^32), b_c = b + c * 4 * aCols.  This is the scan of this engine with the bytes of row i
// of `a` as database column i and the columns of b as queries, so it runs on the same
// limb preparation and contraction kernels.  Host buffers in and out, one call at a time.
int packed_product(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t rows, size_t cols, size_t bc) {
    if (!out || !a || !b || !rows || !cols || !bc) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return -1;
    }
    const size_t P = round_up(4 * cols, 64), R16 = round_up(rows, 16), KC = P / 64, CT = R16 / 16;
    const size_t Bp = round_up(bc, 4), MT = Bp / 4;
    const size_t S = round_up(16 + 4 * P, 256);
    uint8_t *stage = nullptr, *slots = nullptr;
    int8_t *db = nullptr, *qa = nullptr;
    uint32_t *acc = nullptr, *qsum = nullptr, *dout = nullptr;
    unsigned long long *ysum = nullptr;
    int *status = nullptr;
    hipStream_t st = nullptr;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && HIP_OK(hipMalloc(&stage, R16 * P)) &&
              HIP_OK(hipMalloc(&db, R16 * P)) && HIP_OK(hipMalloc(&slots, bc * S)) &&
              HIP_OK(hipMalloc(&qa, Bp * 4 * P)) && HIP_OK(hipMalloc(&acc, Bp * R16 * 4)) &&
              HIP_OK(hipMalloc(&qsum, Bp * 4)) && HIP_OK(hipMalloc(&ysum, Bp * 8)) &&
              HIP_OK(hipMalloc(&status, Bp * 4)) && HIP_OK(hipMalloc(&dout, rows * bc * 4));
    ok = ok && HIP_OK(hipMemsetAsync(stage, 0, R16 * P, st)) && HIP_OK(hipMemsetAsync(slots, 0, bc * S, st)) &&
         HIP_OK(hipMemsetAsync(qsum, 0, Bp * 4, st)) && HIP_OK(hipMemsetAsync(status, 0, Bp * 4, st)) &&
         HIP_OK(hipMemsetAsync(acc, 0, Bp * R16 * 4, st)) &&
         HIP_OK(hipMemcpy2DAsync(stage, P, a, 4 * cols, 4 * cols, rows, hipMemcpyHostToDevice, st)) &&
         HIP_OK(hipMemcpy2DAsync(slots + 16, S, b, 16 * cols, 16 * cols, bc, hipMemcpyHostToDevice, st)) &&
         HIP_OK(hipMemcpyAsync(dout, out, rows * bc * 4, hipMemcpyHostToDevice, st));
    if (ok) {
        const size_t work = CT * KC * 64;
        hipLaunchKernelGGL(tile_db_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, (const uint8_t *)stage,
                           P, db, CT, KC, (size_t)0);
        hipLaunchKernelGGL((limb_prep_kernel<true>), dim3((unsigned)((KC * 4 + 255) / 256), (unsigned)(Bp / kPrepClients)), dim3(256), 0,
                           st, FrameSet{slots, S, nullptr}, (size_t)0, (const int *)status, (int)bc, KC, qa, qsum, ysum);
        GemmArgs g{};
        g.KC = (int)KC;
        g.MT = (int)MT;
        g.NT = (int)CT;
        g.C = (int)R16;
        g.out = acc;
        const Variant &v = kVariants[default_variant(0, (int)MT, (int)KC)];
        v.prepare_scan();  // the ring kernels' LDS size, which engine set-up otherwise sets
        v.scan(g, qa, db, st);
        hipLaunchKernelGGL(packed_finish_kernel, dim3((unsigned)((rows * bc + 255) / 256)), dim3(256), 0, st, dout,
                           (const uint32_t *)acc, (const uint32_t *)qsum, rows, bc, R16);
        ok = HIP_OK(hipGetLastError()) && HIP_OK(hipMemcpyAsync(out, dout, rows * bc * 4, hipMemcpyDeviceToHost, st)) &&
             HIP_OK(hipStreamSynchronize(st));
    }
    for (void *p : {(void *)stage, (void *)db, (void *)slots, (void *)qa, (void *)acc, (void *)qsum, (void *)ysum,
                    (void *)status, (void *)dout})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

bool tail_config_ok(size_t n, uint64_t q1, int q1_bits, size_t row1_off, size_t part_len) {
    // sw_intt_q (src/amx_kernels.cpp:463) and sw_switch_row1 (src/amx_kernels.cpp:517)
    if (n < 16 || (n & (n - 1))) return false;
    if (q1 >= (1ull << 21) || n % 8 || (row1_off + n * (size_t)q1_bits + 7) / 8 > part_len || q1_bits > 24) return false;
    return true;
}

// -------------------------------------------------------- allocation-free batches --

// The workspace of alto_answer_device_noalloc: one fixed array on every device the library
// runs on, held until the process ends.  In host code `scratch` only names the variable; the
// device address comes from hipGetSymbolAddress, looked up once per device (with that device
// current) when the first engine on it is created, and only that address is handed to
// kernels.  Every allocation-free call on a device shares its array.
constexpr size_t kScratchBytes = (size_t)ALTO_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};
int g_scratch_state[kMaxDevices] = {};  // 0 not looked up, 1 found, -1 lookup failed

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// Looks up the device address of `scratch` on the current device `dev` (once per device).
void scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_state[dev]) return;
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch[dev] = rc == hipSuccess ? ptr : nullptr;
    g_scratch_state[dev] = rc == hipSuccess && ptr ? 1 : -1;
}

void *scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_state[dev] == 1 ? g_scratch[dev] : nullptr;
}

}  // namespace

// ======================================================================== C ABI ==

extern "C" {

const char *alto_last_error(void) { return g_err.c_str(); }

int sw_amx_init(void) { return 0; }

// 2 MiB-aligned anonymous mapping with its extent stored just before the pointer.
struct MapHdr {
    void *base;
    size_t len;
};
static void *map_aligned(size_t bytes) {
    const size_t al = 2u << 20, body = round_up(std::max<size_t>(bytes, 1), al), len = body + 2 * al;
    void *base = mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (base == MAP_FAILED) return nullptr;
    uint8_t *p = (uint8_t *)round_up((uintptr_t)base + sizeof(MapHdr), al);
    ((MapHdr *)p)[-1] = MapHdr{base, len};
    return p;
}
void *sw_alloc(size_t bytes) { return map_aligned(bytes); }
void *sw_alloc_zeroed(size_t bytes) { return map_aligned(bytes); /* fresh anonymous pages are zero */ }
void sw_free(void *p) {
    if (!p) return;
    MapHdr h = ((MapHdr *)p)[-1];
    munmap(h.base, h.len);
}

static void parallel_for(size_t n, int threads, const std::function<void(size_t, size_t)> &fn) {
    threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(n, 1)));
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) th.emplace_back([&, t] { fn(n * t / threads, n * (t + 1) / threads); });
    for (auto &x : th) x.join();
}

void sw_tile_db(uint8_t *dst, const uint8_t *db, size_t ld, size_t k, size_t m, int threads) {
    const size_t KC = k / 64;
    parallel_for(m / 16, threads, [&](size_t lo, size_t hi) {
        for (size_t cb = lo; cb < hi; cb++)
            for (size_t r = 0; r < 16; r++)
                for (size_t kc = 0; kc < KC; kc++)
                    memcpy(dst + ((cb / 2) * KC + kc) * 2048 + (cb % 2) * 1024 + r * 64, db + (cb * 16 + r) * ld + kc * 64,
                           64);
    });
}

void sw_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                   const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg, size_t n, size_t rho,
                   uint64_t qmod, int threads) {
    const size_t K = tg * n, KC = K / 64, CT = rho * n / 16;
    parallel_for(rho * n, threads, [&](size_t lo, size_t hi) {
        std::vector<uint64_t> col(K);
        for (size_t oz = lo; oz < hi; oz++) {
            const size_t o = oz / n, z = oz % n;
            std::fill(col.begin(), col.end(), 0);
            for (size_t i = 0; i < num_iter; i++) {
                const uint32_t g = gen_pows[i];
                const size_t zp = tables[((g - 1) / 2) * n + z], zb = tables[((2 * n - g - 1) / 2) * n + z];
                const uint64_t *bt = bold_t[o] + i * tg * n + z, *bb = bold_t_bar[o] + i * tg * n + z;
                for (size_t k = 0; k < tg; k++) {
                    col[k * n + zp] += bt[k * n];
                    col[k * n + zb] += bb[k * n];
                }
            }
            const size_t ct = oz / 16, nn = oz % 16;
            for (size_t kk = 0; kk < K; kk++) {
                const uint32_t v = (uint32_t)(col[kk] % qmod);
                const size_t kc = kk / 64, r = (kk % 64) / 4, t = kk % 4;
                for (size_t j = 0; j < 4; j++)
                    ub[((j * (CT / 2) + ct / 2) * KC + kc) * 2048 + (ct % 2) * 1024 + r * 64 + 4 * nn + t] =
                        (int8_t)((int)((v >> (8 * j)) & 0xff) - 128);
            }
        }
    });
}

}  // extern "C"

namespace {

// Orders the work enqueued on st after the work the caller has enqueued on the null stream:
// the set-up entry points read device inputs on streams of their own, and a caller on the null
// stream (torch's default stream) may still be writing them.
bool after_null_stream(hipStream_t st) {
    hipEvent_t ev = nullptr;
    const bool ok = HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming)) &&
                    HIP_OK(hipEventRecord(ev, nullptr)) && HIP_OK(hipStreamWaitEvent(st, ev, 0));
    if (ev) (void)hipEventDestroy(ev);
    return ok;
}

// A column-major database (`ld`-strided columns of `rows` bytes, host or device memory) into
// the tiled layout at dst, through `stage` (stage_cols columns at a time).
bool tile_db_columns(const uint8_t *db_col, bool on_device, size_t ld, size_t rows, size_t cols, int8_t *dst,
                     uint8_t *stage, size_t stage_cols, hipStream_t st) {
    const size_t KC = rows / 64;
    bool ok = !on_device || after_null_stream(st);
    for (size_t c0 = 0; ok && c0 < cols; c0 += stage_cols) {
        const size_t nc = std::min(stage_cols, cols - c0);
        ok = HIP_OK(hipMemcpy2DAsync(stage, rows, db_col + c0 * ld, ld, rows, nc,
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        const size_t work = nc / 16 * KC * 64;
        hipLaunchKernelGGL(tile_db_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, stage, rows, dst,
                           nc / 16, KC, c0 / 16);
        ok = ok && HIP_OK(hipGetLastError()) && HIP_OK(hipStreamSynchronize(st));
    }
    return ok;
}

// The offline hint (src/server.rs:656-709, 738-768) of a tiled database of `rows` records x
// `cols` columns on the current device: the n x rows expansion of the query polynomials as
// limb rows (hint_limb_kernel), the scan contraction with the exact mod-Q store, and the
// 128 S_z correction (hint_finish_kernel).
// The hint into device memory (dout: n x cols u64), enqueued on st and waited for.
bool hint_device(const int8_t *db, size_t rows, size_t cols, size_t n, const uint64_t *a, size_t npolys,
                 uint64_t *dout, hipStream_t st) {
    const size_t C = cols, KC = rows / 64, MT = n / 4;
    uint64_t *da = nullptr;
    int8_t *qa = nullptr;
    unsigned long long *acc = nullptr, *ssum = nullptr;
    bool ok = HIP_OK(hipMalloc(&da, npolys * n * 8)) && HIP_OK(hipMalloc(&qa, MT * KC * kBlockBytes)) &&
              HIP_OK(hipMalloc(&acc, n * C * 8)) && HIP_OK(hipMalloc(&ssum, n * 8));
    ok = ok && HIP_OK(hipMemcpyAsync(da, a, npolys * n * 8, hipMemcpyHostToDevice, st)) &&
         HIP_OK(hipMemsetAsync(acc, 0, n * C * 8, st)) && HIP_OK(hipMemsetAsync(ssum, 0, n * 8, st));
    if (ok) {
        hipLaunchKernelGGL(hint_limb_kernel, dim3((unsigned)((KC * 4 + 255) / 256), (unsigned)n), dim3(256), 0, st,
                           (const uint64_t *)da, (int)n, KC, qa, ssum);
        GemmArgs g{};
        g.KC = (int)KC;
        g.MT = (int)MT;
        g.NT = (int)(C / 16);
        g.C = (int)C;
        g.out64 = acc;
        const Variant &v = kVariants[default_variant(0, (int)MT, (int)KC)];
        v.prepare_scan();
        v.scan(g, qa, db, st);
        hipLaunchKernelGGL(hint_finish_kernel, dim3((unsigned)((n * C + 255) / 256)), dim3(256), 0, st, dout,
                           (const unsigned long long *)acc, (const unsigned long long *)ssum, C, n * C);
        ok = HIP_OK(hipGetLastError()) && HIP_OK(hipStreamSynchronize(st));
    }
    for (void *q : {(void *)da, (void *)qa, (void *)acc, (void *)ssum})
        if (q) (void)hipFree(q);
    return ok;
}

bool hint_from_tiled(const int8_t *db, size_t rows, size_t cols, size_t n, const uint64_t *a, size_t npolys,
                     uint64_t *hint) {
    uint64_t *dout = nullptr;
    hipStream_t st = nullptr;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && HIP_OK(hipMalloc(&dout, n * cols * 8)) &&
              hint_device(db, rows, cols, n, a, npolys, dout, st) &&
              HIP_OK(hipMemcpyAsync(hint, dout, n * cols * 8, hipMemcpyDeviceToHost, st)) &&
              HIP_OK(hipStreamSynchronize(st));
    if (dout) (void)hipFree(dout);
    if (st) (void)hipStreamDestroy(st);
    return ok;
}

// The packing matrix (tile layout of sw_pack_build) from device bold_t / bold_t_bar (part-major,
// per values a part), device tables (n x n) and generator powers; enqueued on st.
bool pack_build_device(const uint64_t *dbt, const uint64_t *dbb, size_t per, const uint32_t *dtab,
                       const uint32_t *dpow, size_t num_iter, size_t tg, size_t n, size_t cols, uint64_t qmod,
                       int8_t *dub, hipStream_t st) {
    const size_t lds = tg * n * 8;
    if (!HIP_OK(hipFuncSetAttribute((const void *)pack_build_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds)))
        return false;
    hipLaunchKernelGGL(pack_build_kernel, dim3((unsigned)cols), dim3(256), lds, st, dbt, dbb, per, dtab, dpow,
                       (int)num_iter, (int)tg, (int)n, (int)cols, qmod, dub);
    return HIP_OK(hipGetLastError());
}

bool hint_shape_ok(const char *fn, size_t rows, size_t cols, size_t n, size_t npolys) {
    if (n < 16 || (n & (n - 1)) || !rows || rows % 64 || !cols || cols % 16 || rows % n || npolys != rows / n ||
        rows / 64 > (size_t)INT32_MAX || cols > (size_t)INT32_MAX) {
        set_err("%s: %zu polynomials of n = %zu for %zu records x %zu columns (n a power of two >= 16, records a "
                "multiple of n and of 64, columns of 16)", fn, npolys, n, rows, cols);
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

int alto_engine_hint0(void *p, const uint64_t *a, size_t npolys, uint64_t *hint) {
    auto *e = (Engine *)p;
    if (!e || !a || !hint) {
        set_err("alto_engine_hint0: null argument");
        return -1;
    }
    if (!hint_shape_ok("alto_engine_hint0", e->rows, e->cols, e->n, npolys)) return -1;
    if (!HIP_OK(hipSetDevice(e->device))) return -1;
    return hint_from_tiled(e->db, e->rows, e->cols, e->n, a, npolys, hint) ? 0 : -1;
}

int alto_hint0(const uint8_t *db_col, int on_device, size_t ld, size_t rows, size_t cols, size_t n, const uint64_t *a,
               size_t npolys, uint64_t *hint) {
    if (!db_col || !a || !hint || ld < rows) {
        set_err("alto_hint0: null argument or ld < rows");
        return -1;
    }
    if (!hint_shape_ok("alto_hint0", rows, cols, n, npolys)) return -1;
    int dev = 0;
    if (!HIP_OK(hipGetDevice(&dev))) return -1;
    int8_t *db = nullptr;
    uint8_t *stage = nullptr;
    hipStream_t st = nullptr;
    const size_t stage_cols = std::max<size_t>(16, std::min<size_t>(cols, ((size_t)256 << 20) / rows / 16 * 16));
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && HIP_OK(hipMalloc(&db, rows * cols)) &&
              HIP_OK(hipMalloc(&stage, stage_cols * rows));
    ok = ok && tile_db_columns(db_col, on_device != 0, ld, rows, cols, db, stage, stage_cols, st);
    if (stage) (void)hipFree(stage);
    ok = ok && hint_from_tiled(db, rows, cols, n, a, npolys, hint);
    if (db) (void)hipFree(db);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

int alto_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                    const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg, size_t n, size_t rho,
                    uint64_t qmod, int threads) {
    (void)threads;
    const size_t K = tg * n, cols = rho * n, per = num_iter * tg * n;
    if (!ub || !bold_t || !bold_t_bar || !tables || !gen_pows || !num_iter || !tg || !n || !rho || !qmod || K % 64 ||
        cols % 32 || K * 8 > kMaxLds || cols > (size_t)INT32_MAX || num_iter * tg > (size_t)INT32_MAX) {
        set_err("alto_pack_build: unsupported shape (needs tg n %% 64 == 0, rho n %% 32 == 0, tg n <= %zu)", kMaxLds / 8);
        return -1;
    }
    for (size_t o = 0; o < rho; o++)
        if (!bold_t[o] || !bold_t_bar[o]) {
            set_err("alto_pack_build: null bold_t / bold_t_bar");
            return -1;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return -1;
    }
    uint64_t *dbt = nullptr, *dbb = nullptr;
    uint32_t *dtab = nullptr, *dpow = nullptr;
    int8_t *dub = nullptr;
    hipStream_t st = nullptr;
    const size_t ubytes = 4 * K * cols;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
              HIP_OK(hipMalloc(&dbt, rho * per * 8)) && HIP_OK(hipMalloc(&dbb, rho * per * 8)) &&
              HIP_OK(hipMalloc(&dtab, n * n * 4)) && HIP_OK(hipMalloc(&dpow, num_iter * 4)) &&
              HIP_OK(hipMalloc(&dub, ubytes));
    for (size_t o = 0; ok && o < rho; o++)
        ok = HIP_OK(hipMemcpyAsync(dbt + o * per, bold_t[o], per * 8, hipMemcpyHostToDevice, st)) &&
             HIP_OK(hipMemcpyAsync(dbb + o * per, bold_t_bar[o], per * 8, hipMemcpyHostToDevice, st));
    ok = ok && HIP_OK(hipMemcpyAsync(dtab, tables, n * n * 4, hipMemcpyHostToDevice, st)) &&
         HIP_OK(hipMemcpyAsync(dpow, gen_pows, num_iter * 4, hipMemcpyHostToDevice, st));
    ok = ok && pack_build_device(dbt, dbb, per, dtab, dpow, num_iter, tg, n, cols, qmod, dub, st);
    ok = ok && HIP_OK(hipMemcpyAsync(ub, dub, ubytes, hipMemcpyDeviceToHost, st)) && HIP_OK(hipStreamSynchronize(st));
    for (void *p : {(void *)dbt, (void *)dbb, (void *)dtab, (void *)dpow, (void *)dub})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

static void *engine_create(const uint8_t *db_col, bool db_on_device, bool db_rows, const uint8_t *db_tiled, size_t ld,
                           size_t rows,
                           size_t cols, const int8_t *ub, size_t n, size_t tg, const uint8_t *const *prefix,
                           size_t part_len, size_t row1_off, uint64_t q1, int q1_bits, const uint64_t *const *that,
                           const uint64_t *inv_table, uint64_t q, int threads) {
    // configuration checks of src/amx_kernels.cpp:551-552 (the L2-size probe has no GPU
    // meaning), plus empty shapes, which the reference accepts and later aborts on
    if (!db_col || !db_tiled || !ub || !prefix || !that || !inv_table || q != kQ || !n || n > 4096 || n % 16 ||
        cols % n || rows % 64 || cols % 32 || threads <= 0 || (tg * n) % 128 || !rows || !cols || !tg) {
        set_err("unsupported engine configuration");
        return nullptr;
    }
    // the packing contraction accumulates all of KP in i32: |y' u'| <= 2^14, so KP <= 2^16
    if (tg * n > ((size_t)kFoldSteps << 6)) {
        set_err("unsupported engine configuration: tg * n > 65536");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return nullptr;
    }
    auto *e = new Engine;
    if (!HIP_OK(hipGetDevice(&e->device))) { delete e; return nullptr; }
    scratch_setup(e->device);  // device address of the allocation-free entry point's workspace
    e->ld = ld; e->rows = rows; e->cols = cols; e->n = n; e->tg = tg; e->kp = tg * n; e->outputs = cols / n;
    e->part_len = part_len; e->part_words = (part_len + 3) / 4; e->row1_off = row1_off; e->q1 = q1;
    e->q1_bits = q1_bits; e->threads = threads;
    e->tail_ok = tail_config_ok(n, q1, q1_bits, row1_off, part_len);
    while ((size_t)1 << e->logn < n) e->logn++;
    e->KC = rows / 64; e->KCp = e->kp / 64; e->CT = cols / 16;
    e->k128e = (uint32_t)((128ull * kE) % kQ);
    e->ycorr0 = (uint32_t)((kQ - host_mulmod(host_mulmod(e->k128e, e->k128e), e->kp % kQ)) % kQ);
    const size_t lds = tail_lds_bytes(n, e->part_words);
    e->ctx0 = new Context;
    bool ok = e->ctx0->init(e);
    hipStream_t setup = e->ctx0->stream;
    // the attribute is process-wide per kernel, so it is set to the most any engine may ask
    // for (an engine's own need would lower it under an earlier engine that needs more)
    ok = ok && HIP_OK(hipFuncSetAttribute((const void *)tail_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)kMaxLds));
    if (e->tail_ok && lds > kMaxLds) e->tail_ok = false;
    for (const auto &v : kVariants) {
        v.prepare_scan();
        v.prepare_pack();
    }
    for (const FusedPair &f : kFused) {
        f.prepare();
    }
    for (const StepPair &f : kStep) {
        f.prepare();
    }
    // resident state
    const size_t dbb = rows * cols, ubb = 4 * e->kp * cols;
    ok = ok && HIP_OK(hipMalloc(&e->db, dbb)) && HIP_OK(hipMalloc(&e->u, ubb)) &&
         HIP_OK(hipMalloc(&e->ucorr, cols * 4)) && HIP_OK(hipMalloc(&e->that, e->outputs * e->kp * 4)) &&
         HIP_OK(hipMalloc(&e->wtab, n * 4)) && HIP_OK(hipMalloc(&e->prefix, e->outputs * e->part_words * 4 + 4));
    e->device_bytes = dbb + ubb + cols * 4 + e->outputs * e->kp * 4 + n * 4 + e->outputs * e->part_words * 4;
    // database: column slices (or record slices) through a staging buffer, tiled on the device
    uint8_t *stage = nullptr;
    const size_t stage_cols = std::max<size_t>(16, std::min<size_t>(cols, ((size_t)256 << 20) / rows / 16 * 16));
    const size_t stage_rows = std::max<size_t>(64, std::min<size_t>(rows, ((size_t)256 << 20) / cols / 64 * 64));
    const size_t stage_bytes = std::max(db_rows ? stage_rows * cols : stage_cols * rows, e->kp * cols);
    ok = ok && HIP_OK(hipMalloc(&stage, stage_bytes));
    if (db_rows && db_on_device) ok = ok && after_null_stream(setup);
    for (size_t r0 = 0; ok && db_rows && r0 < rows; r0 += stage_rows) {
        const size_t nr = std::min(stage_rows, rows - r0);
        ok = HIP_OK(hipMemcpy2DAsync(stage, cols, db_col + r0 * ld, ld, cols, nr,
                                     db_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, setup));
        hipLaunchKernelGGL(tile_rows_kernel, dim3((unsigned)ceil_div(e->CT, 16), (unsigned)(nr / 64)), dim3(256), 0,
                           setup, (const uint8_t *)stage, cols, e->db, e->CT, e->KC, r0 / 64);
        ok = ok && HIP_OK(hipGetLastError()) && HIP_OK(hipStreamSynchronize(setup));
    }
    if (ok && !db_rows) ok = tile_db_columns(db_col, db_on_device, ld, rows, cols, e->db, stage, stage_cols, setup);
    // packing matrix: one byte plane at a time
    for (int j = 0; ok && j < 4; j++) {
        ok = HIP_OK(hipMemcpyAsync(stage, ub + (size_t)j * e->kp * cols, e->kp * cols, hipMemcpyHostToDevice, setup));
        const size_t work = e->CT * e->KCp * 64;
        hipLaunchKernelGGL(tile_u_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, setup,
                           (const int8_t *)stage, e->u, e->CT, e->KCp, j);
        ok = ok && HIP_OK(hipGetLastError()) && HIP_OK(hipStreamSynchronize(setup));
    }
    if (stage) (void)hipFree(stage);
    if (ok) {
        hipLaunchKernelGGL(u_colsum_kernel, dim3((unsigned)e->CT), dim3(256), 0, setup, (const int8_t *)e->u,
                           e->ucorr, e->KCp, e->k128e);
        ok = HIP_OK(hipGetLastError());
    }
    // tail keys reduced mod Q (the reference reduces them per use, src/amx_kernels.cpp:507),
    // twiddles as the low 32 bits the reference multiplies with (_mm512_mul_epu32), prefixes
    // zero-padded to whole words
    if (ok) {
        std::vector<uint32_t> th(e->outputs * e->kp), w(n), pre(e->outputs * e->part_words, 0);
        for (size_t o = 0; o < e->outputs; o++)
            for (size_t i = 0; i < e->kp; i++) th[o * e->kp + i] = (uint32_t)(that[o][i] % kQ);
        for (size_t i = 0; i < n; i++) w[i] = (uint32_t)(inv_table[i] & 0xffffffffull);
        for (size_t o = 0; o < e->outputs; o++) memcpy((uint8_t *)pre.data() + o * e->part_words * 4, prefix[o], part_len);
        ok = HIP_OK(hipMemcpy(e->that, th.data(), th.size() * 4, hipMemcpyHostToDevice)) &&
             HIP_OK(hipMemcpy(e->wtab, w.data(), n * 4, hipMemcpyHostToDevice)) &&
             HIP_OK(hipMemcpy(e->prefix, pre.data(), pre.size() * 4, hipMemcpyHostToDevice)) &&
             HIP_OK(hipStreamSynchronize(setup));
    }
    if (!ok) {
        const std::string msg = g_err;
        delete e;
        g_err = msg;
        return nullptr;
    }
    return e;
}

void *alto_engine_new(const uint8_t *db_col, const uint8_t *db_tiled, size_t ld, size_t rows, size_t cols,
                      const int8_t *ub, size_t n, size_t tg, const uint8_t *const *prefix, size_t part_len,
                      size_t row1_off, uint64_t q1, int q1_bits, const uint64_t *const *that,
                      const uint64_t *inv_table, uint64_t q, int threads) {
    return engine_create(db_col, false, false, db_tiled, ld, rows, cols, ub, n, tg, prefix, part_len, row1_off, q1, q1_bits,
                         that, inv_table, q, threads);
}

void *alto_engine_new_device(const uint8_t *db_col_device, size_t ld, size_t rows, size_t cols, const int8_t *ub,
                             size_t n, size_t tg, const uint8_t *const *prefix, size_t part_len, size_t row1_off,
                             uint64_t q1, int q1_bits, const uint64_t *const *that, const uint64_t *inv_table,
                             uint64_t q, int threads) {
    return engine_create(db_col_device, true, false, db_col_device, ld, rows, cols, ub, n, tg, prefix, part_len,
                         row1_off, q1, q1_bits, that, inv_table, q, threads);
}

void *alto_engine_new_records(const uint8_t *records, int on_device, size_t ld, size_t rows, size_t cols,
                              const int8_t *ub, size_t n, size_t tg, const uint8_t *const *prefix, size_t part_len,
                              size_t row1_off, uint64_t q1, int q1_bits, const uint64_t *const *that,
                              const uint64_t *inv_table, uint64_t q, int threads) {
    if (ld < cols) {
        set_err("alto_engine_new_records: record stride below the record size");
        return nullptr;
    }
    return engine_create(records, on_device != 0, true, records, ld, rows, cols, ub, n, tg, prefix, part_len,
                         row1_off, q1, q1_bits, that, inv_table, q, threads);
}

void alto_engine_free(void *p) { delete (Engine *)p; }

void matMulVecPacked(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t aRows, size_t aCols) {
    if (packed_product(out, a, b, aRows, aCols, 1)) fprintf(stderr, "matMulVecPacked: %s\n", g_err.c_str());
}
void matMulVecPacked2(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols) {
    if (packed_product(out, a, b_full, aRows, aCols, 2)) fprintf(stderr, "matMulVecPacked2: %s\n", g_err.c_str());
}
void matMulVecPacked4(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols) {
    if (packed_product(out, a, b_full, aRows, aCols, 4)) fprintf(stderr, "matMulVecPacked4: %s\n", g_err.c_str());
}
void matMulVecPacked8(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols) {
    if (packed_product(out, a, b_full, aRows, aCols, 8)) fprintf(stderr, "matMulVecPacked8: %s\n", g_err.c_str());
}
int alto_matmul_vec_packed(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t a_rows, size_t a_cols,
                           size_t b_cols) {
    return packed_product(out, a, b, a_rows, a_cols, b_cols);
}

size_t alto_request_len(const void *p) { return ((const Engine *)p)->request_len(); }
size_t alto_response_len(const void *p) { return ((const Engine *)p)->response_len(); }
size_t alto_engine_device_bytes(const void *p) { return ((const Engine *)p)->device_bytes; }

int alto_answer_device(void *p, const uint8_t *slots, size_t slot_stride, size_t B, uint8_t *out, size_t out_stride,
                       int *status, void *stream) {
    auto *e = (Engine *)p;
    if (!e) {
        set_err("alto_answer_device: null engine");
        return -1;
    }
    return alto_answer_gather(e->ctx0, slots, slot_stride, nullptr, B, out, out_stride, nullptr, status, stream);
}

size_t alto_context_bytes(const void *p, size_t B) {
    auto *e = (const Engine *)p;
    if (!e) return 0;
    return workspace_bytes(*e, bucket_queries(B)) + 2 * sizeof(int);
}

int alto_answer_device_noalloc(void *p, const uint8_t *slots, size_t slot_stride, size_t B, uint8_t *out,
                               size_t out_stride, int *status, void *stream) {
    auto *e = (Engine *)p;
    if (!e || !slots || !out || !status || ((uintptr_t)slots & 15) || (slot_stride & 15) ||
        slot_stride < ALTO_FRAME_OFFSET + e->request_len() || out_stride < e->response_len()) {
        set_err("alto_answer_device_noalloc: bad arguments");
        return -1;
    }
    if (!B) return 0;
    if (!e->tail_ok) {
        set_err("unsupported tail configuration");
        return -3;
    }
    const size_t Bp = round_up(B, 4), need = workspace_bytes(*e, Bp);
    if (need > kScratchBytes) {
        set_err("alto_answer_device_noalloc: a batch of %zu queries needs %zu bytes of workspace on this engine, "
                "more than the %zu of the fixed scratch array",
                B, need, kScratchBytes);
        return -1;
    }
    void *base = scratch_address(e->device);
    if (!base) {
        set_err("alto_answer_device_noalloc: no device address for the scratch array on device %d", e->device);
        return -1;
    }
    if (!HIP_OK(hipSetDevice(e->device))) return -1;
    hipStream_t st = stream ? (hipStream_t)stream : e->ctx0->stream;
    if (!stream && !e->ctx0->join_null_before()) return -1;
    int plan = -1;
    const int rc = launch_batch(*e, workspace_at(*e, Bp, (uint8_t *)base), FrameSet{slots, slot_stride, nullptr}, B,
                                out, out_stride, nullptr, status, st, nullptr, false, &plan);
    if (!stream && !e->ctx0->join_null_after()) return -1;
    return rc;
}

void *alto_context_new(void *p) {
    auto *e = (Engine *)p;
    if (!e) return nullptr;
    if (!HIP_OK(hipSetDevice(e->device))) return nullptr;
    auto *c = new Context;
    if (!c->init(e)) {
        delete c;
        return nullptr;
    }
    return c;
}

void alto_context_free(void *c) {
    auto *ctx = (Context *)c;
    if (ctx && ctx->e && ctx == ctx->e->ctx0) return;  // the default context lives with its engine
    delete ctx;
}

void *alto_engine_context(void *p) { return p ? ((Engine *)p)->ctx0 : nullptr; }

int alto_context_set_graphs(void *c, int mode) {
    auto *ctx = (Context *)c;
    if (!ctx || mode < -1 || mode > 1) return -1;
    if (mode == 0) ctx->drop_graphs();
    ctx->graph_mode = mode;
    return 0;
}

int alto_context_graphs(const void *c) { return c ? (int)((const Context *)c)->graphs.size() : -1; }

int alto_context_reserve(void *c, size_t B) {
    auto *ctx = (Context *)c;
    if (!ctx) return -1;
    if (!HIP_OK(hipSetDevice(ctx->e->device))) return -1;
    return ctx->reserve(bucket_queries(B), false) ? 0 : -4;
}

int alto_answer_gather(void *c, const uint8_t *slots, size_t slot_stride, const int *slot_index, size_t B,
                       uint8_t *out, size_t out_stride, const int *out_index, int *status, void *stream) {
    auto *ctx = (Context *)c;
    if (!ctx || !slots || !out || !status || ((uintptr_t)slots & 15) || (slot_stride & 15) ||
        slot_stride < ALTO_FRAME_OFFSET + ctx->e->request_len() || out_stride < ctx->e->response_len()) {
        set_err("alto_answer_device: bad arguments");
        return -1;
    }
    Engine *e = ctx->e;
    if (!B) return 0;
    if (!e->tail_ok) {
        set_err("unsupported tail configuration");
        return -3;
    }
    if (!HIP_OK(hipSetDevice(e->device))) return -1;
    if (!ctx->reserve(B, false)) return -4;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    if (!stream && !ctx->join_null_before()) return -1;
    const int rc =
        ctx->answer_device(FrameSet{slots, slot_stride, slot_index}, B, out, out_stride, out_index, status, st);
    if (!stream && !ctx->join_null_after()) return -1;
    return rc;
}

int alto_answer(void *p, const uint8_t *const *req, const size_t *len, size_t B, uint8_t *const *out, int *status,
                int method, double *stage_ms) {
    auto *e = (Engine *)p;
    if (!e || !req || !len || !out || !status ||
        (method != ALTO_METHOD_AMX && method != ALTO_METHOD_VNNI && method != ALTO_METHOD_HIP)) {
        set_err("alto_answer: bad arguments or method");
        return -1;
    }
    if (stage_ms) std::fill(stage_ms, stage_ms + 4, 0.0);
    if (!B) return 0;
    if (!HIP_OK(hipSetDevice(e->device))) return -1;
    Context *ctx = e->ctx0;
    if (!ctx->reserve(B, true)) return -1;
    const size_t want = e->request_len(), rl = e->response_len(), S = ctx->slot_stride;
    ctx->hstatus.assign(B, 0);
    for (size_t c = 0; c < B; c++) ctx->hstatus[c] = (!req[c] || len[c] != want) ? 1 : 0;
    hipStream_t st = ctx->stream;
    const FrameSet fs{ctx->slots, S, nullptr};
    if (!e->host) e->host = new HostStage(std::max(1, std::min(e->threads, 16)));
    HostStage &h = *e->host;
    // frames: staged by the host threads into pinned buffers, about 4 MiB of slots at a time
    // (pieces of up to 1 MiB per thread), each chunk's DMA copy overlapping the next chunk's
    // staging; a pinned buffer is reused once its copy has completed
    constexpr size_t kPiece = (size_t)1 << 20;
    const size_t per = std::max<size_t>(1, ((size_t)4 << 20) / S), pieces = (want + kPiece - 1) / kPiece;
    if (!h.reserve(per * S, B * rl)) return -1;
    bool ok = HIP_OK(hipEventRecord(ctx->next_set()[0], st));
    for (size_t c0 = 0, ci = 0; ok && c0 < B; c0 += per, ci++) {
        const size_t c1 = std::min(B, c0 + per);
        const int b = (int)(ci % HostStage::kBufs);
        if (h.pending[b]) ok = HIP_OK(hipEventSynchronize(h.copied[b]));
        h.pending[b] = false;
        uint8_t *dst = h.buf[b];
        h.pool.run((c1 - c0) * pieces, [&](size_t t) {
            const size_t c = c0 + t / pieces, off = (t % pieces) * kPiece;
            if (!ctx->hstatus[c]) memcpy(dst + (c - c0) * S + ALTO_FRAME_OFFSET + off, req[c] + off, std::min(kPiece, want - off));
        });
        ok = ok && HIP_OK(hipMemcpyAsync(ctx->slots + c0 * S, dst, (c1 - c0) * S, hipMemcpyHostToDevice, st)) &&
             HIP_OK(hipEventRecord(h.copied[b], st));
        h.pending[b] = ok;
    }
    ok = ok && HIP_OK(hipMemcpyAsync(ctx->status, ctx->hstatus.data(), B * sizeof(int), hipMemcpyHostToDevice, st));
    if (!ok) return -1;
    if (!e->tail_ok) {
        // validation only: the reference fails the whole batch with -3 once any request is valid
        hipLaunchKernelGGL(validate_kernel, dim3((unsigned)B, validate_slices(e->kp)), dim3(256), 0, st, fs, ctx->status,
                           (uint32_t)e->kp, ctx->qsum, ctx->ysum, (const int *)nullptr);
        if (!HIP_OK(hipMemcpyAsync(status, ctx->status, B * sizeof(int), hipMemcpyDeviceToHost, st)) ||
            !HIP_OK(hipStreamSynchronize(st)))
            return -1;
        for (size_t c = 0; c < B; c++)
            if (status[c] == 0) return -3;
        return 0;
    }
    // the upload is timed as part of the validation stage (ev[0] was recorded before it); the
    // tail writes the responses into the mapped pinned buffer, from which the host threads copy
    // them to the caller's buffers
    if (ctx->enqueue(ctx->ws(), fs, B, h.resp_dev, rl, nullptr, ctx->status, st, true, false)) return -1;
    if (!HIP_OK(hipMemcpyAsync(status, ctx->status, B * sizeof(int), hipMemcpyDeviceToHost, st)) ||
        !HIP_OK(hipStreamSynchronize(st)))
        return -1;
    for (int b = 0; b < HostStage::kBufs; b++) h.pending[b] = false;  // the stream has drained
    int answered = 0;
    for (size_t c = 0; c < B; c++) answered += status[c] == 0;
    h.pool.run(B, [&](size_t c) {
        if (status[c] == 0) memcpy(out[c], h.resp + c * rl, rl);
    });
    if (stage_ms) {
        // {validate, scan, pack, tail} as in src/amx_kernels.cpp:613, from the six stages
        double t[6];
        if (!ctx->stage_times(ctx->ring_head, t)) return -1;
        stage_ms[0] = t[0];
        stage_ms[1] = t[1] + t[2];
        stage_ms[2] = t[3] + t[4];
        stage_ms[3] = t[5];
    }
    return answered;
}

int alto_set_graphs(void *p, int on) {
    auto *e = (Engine *)p;
    if (!e) return -1;
    if (!on) e->ctx0->drop_graphs();
    e->use_graphs = on != 0;
    return 0;
}

int alto_set_graph_timing(void *p, int on) {
    auto *e = (Engine *)p;
    if (!e) return -1;
    e->time_graphs = on != 0;
    return 0;
}

int alto_set_fusion(void *p, int on) {
    auto *e = (Engine *)p;
    if (!e) return -1;
    if (on < -1 || on > 2) return -1;
    const int prev = e->fuse;
    e->fuse = on;
    return prev;
}

int alto_fused(const void *p, size_t B) {
    auto *e = (const Engine *)p;
    if (!e || !B) return -1;
    const int fz = e->fused_for((int)(round_up(B, 4) / 4));
    return fz >= kNumFused ? 2 : fz >= 0 ? 1 : 0;
}

int alto_variant_count(void) { return kNumVariants; }

const char *alto_variant_name(int idx) { return idx >= 0 && idx < kNumVariants ? kVariants[idx].name : nullptr; }

int alto_tune_set(void *p, int which, size_t B, int idx) {
    auto *e = (Engine *)p;
    if (!e || which < 0 || which > 1 || idx < -1 || idx >= kNumVariants || !B) return -1;
    e->tuned[which][bucket_of((int)(round_up(B, 4) / 4))] = idx;
    return 0;
}

int alto_tune_get(const void *p, int which, size_t B) {
    auto *e = (const Engine *)p;
    if (!e || which < 0 || which > 1 || !B) return -1;
    return e->variant_for(which, (int)(round_up(B, 4) / 4));
}

int alto_context_stage_totals(void *c, double *ms, int count, int reset) {
    auto *ctx = (Context *)c;
    if (!ctx || !ms || count <= 0) return -1;
    if (!ctx->fold_all()) return -1;
    const int k = std::min(count, 6);
    for (int i = 0; i < k; i++) ms[i] = ctx->stage_sum[i];
    const long long n = ctx->stage_count;
    if (reset) {
        std::fill(ctx->stage_sum, ctx->stage_sum + 6, 0.0);
        ctx->stage_count = 0;
    }
    return (int)std::min<long long>(n, INT32_MAX);
}

int alto_stage_totals(void *p, double *ms, int count, int reset) {
    return p ? alto_context_stage_totals(((Engine *)p)->ctx0, ms, count, reset) : -1;
}

int alto_context_timings(void *c, double *ms, int count) {
    auto *ctx = (Context *)c;
    if (!ctx || !ms || count <= 0 || !ctx->timed_valid || ctx->ring_head < 0) return 0;
    double t[6];
    if (!ctx->stage_times(ctx->ring_head, t)) return -1;
    const int k = std::min(count, 6);
    for (int i = 0; i < k; i++) ms[i] = t[i];
    return k;
}

int alto_last_timings(void *p, double *ms, int count) { return p ? alto_context_timings(((Engine *)p)->ctx0, ms, count) : 0; }

}  // extern "C"

// the offline packing precomputation on the device (alto_inspiring_offline)
#include "alto_offline.hpp"
