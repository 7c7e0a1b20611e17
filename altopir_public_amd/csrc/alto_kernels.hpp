// alto_kernels.hpp -- HIP kernels of the MI355X answer engine (gfx950 only).
//
// Arithmetic of the path (reference: src/amx_kernels.cpp:564-615):
//   scan   acc[c][col]  = sum_k D[col][k] * q_c[k]              mod 2^32
//          inter        = floor((acc * Q + 2^31) / 2^32)
//   pack   beta[c][col] = sum_kk y_c[kk] * U[kk][col]           mod Q
//   tail   a = beta + sum_k z_c * that_o  mod Q,  coef = INTT(a),
//          v = inter + coef mod Q, rescaled to q1 and bit-packed after the prefix.
//
// Both contractions run on the i8 matrix cores (v_mfma_i32_16x16x64_i8) with exact
// offset corrections:
//   * database bytes are stored as d' = d - 128; query words are split into signed
//     digits s_i = byte_i(q + 0x80808080) - 128, so sum_i 256^i s_i == q (mod 2^32) and
//     acc = sum_i 256^i P_i + 128 * sum_k q_k  (mod 2^32), P_i = sum_k d'_k s_{i,k};
//   * packing keys and the packing matrix are split into bytes offset by -128:
//     y'_i = y_i - 128, u'_j = u_j - 128, and
//     sum y*U = sum_{i,j} 2^{8(i+j)} G_ij + 128E (Usum + Ysum) - (128E)^2 KP  (mod Q),
//     G_ij = sum_kk y'_i u'_j, E = 0x01010101, Usum/Ysum = plain sums of U / y.
// Limb products never leave the registers: each lane of an MFMA output holds the four
// limbs of one query for one column (rows 4c..4c+3 of the A operand), so the limbs are
// recombined in place.
//
// MFMA operand tiles ("blocks") are 1 KiB: 16 rows x 64 k-bytes, stored in lane order so
// that lane l loads the 16 bytes of row (l % 16), k-group (l / 16) with one dwordx4:
//   block[16 * l + j] = X[row0 + l % 16][k0 + 16 * (l / 16) + j]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace alto {

constexpr uint64_t kQ = 4294955009ull;  // 2^32 - 12287
constexpr uint32_t kQ32 = 4294955009u;
constexpr uint64_t kFold = 12287;       // 2^32 mod Q
constexpr uint32_t kE = 0x01010101u;
constexpr int kChunk = 64;              // k-bytes per block
constexpr int kBlockBytes = 1024;
// Accumulate at most this many 64-deep MFMA steps in i32 before folding into a u32:
// |d' * s| <= 2^14, so 1024 * 64 * 2^14 = 2^30 stays clear of i32 overflow.
constexpr int kFoldSteps = 1024;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// ------------------------------------------------------------------ mod-Q helpers --

// x < 2^32 + 2^28 -> x mod Q
__device__ __forceinline__ uint32_t red_once(uint64_t x) { return (uint32_t)(x >= kQ ? x - kQ : x); }

// s + carry * 2^32 (< 2Q) -> mod Q in 32-bit arithmetic: subtracting Q is adding 12287 mod 2^32
__device__ __forceinline__ uint32_t red_carry(uint32_t s, bool carry) { return (carry || s >= kQ) ? s + (uint32_t)kFold : s; }

// any u64 -> mod Q, using 2^32 == 12287 (mod Q) twice: p = H 2^32 + L == H 12287 + L = t (one
// 32 x 32 + 32 -> 64 multiply-add), t = T1 2^32 + T0 == T1 12287 + T0 with T1 <= 2^14 (a 24-bit
// multiply), < 2^32 + 2^28 < 2Q
__device__ __forceinline__ uint32_t red64(uint64_t p) {
    const uint64_t t = (p >> 32) * kFold + (p & 0xffffffffull);
    uint32_t s;
    const bool carry = __builtin_add_overflow((uint32_t)t, __umul24((uint32_t)(t >> 32), (uint32_t)kFold), &s);
    return red_carry(s, carry);
}

__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b) { return red64((uint64_t)a * b); }

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b) {
    uint32_t s;
    const bool carry = __builtin_add_overflow(a, b, &s);
    return red_carry(s, carry);
}

// signed |t| < 2^62 -> t mod Q in [0, Q)
__device__ __forceinline__ uint32_t red_i64(int64_t t) {
    int64_t hi = t >> 32;
    int64_t lo = (int64_t)(uint32_t)t;
    t = hi * (int64_t)kFold + lo;
    hi = t >> 32;
    lo = (int64_t)(uint32_t)t;
    t = hi * (int64_t)kFold + lo;
    if (t < 0) t += (int64_t)kQ;
    else if (t >= (int64_t)kQ) t -= (int64_t)kQ;
    return (uint32_t)t;
}

// ------------------------------------------------------------- setup: layouts ----

// Column-major database slice (pitch bytes per column, `rows` records) -> 16-column x
// 64-record blocks of d - 128, block (ct, kc) at ((ct0 + ct) * KC + kc) * 1 KiB.
__global__ void tile_db_kernel(const uint8_t *__restrict__ src, size_t pitch, int8_t *__restrict__ dst,
                               size_t ntiles, size_t KC, size_t ct0) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ntiles * KC * 64) return;
    const int lane = (int)(gid & 63);
    const size_t blk = gid >> 6, kc = blk % KC, ct = blk / KC;
    const uint4 v = *(const uint4 *)(src + (ct * 16 + (lane & 15)) * pitch + kc * 64 + (lane >> 4) * 16);
    const uint4 f = make_uint4(v.x ^ 0x80808080u, v.y ^ 0x80808080u, v.z ^ 0x80808080u, v.w ^ 0x80808080u);
    *(uint4 *)(dst + ((ct0 + ct) * KC + kc) * kBlockBytes + lane * 16) = f;
}

// Database ingest from records (row-major: record k = `pitch`-strided row of bytes) into the
// same blocks as tile_db_kernel.  A workgroup takes 64 records (one k-step) x 256 columns (16
// column tiles): the records are read row-wise into LDS (rows padded to 65 words, so the
// column-wise reads of the four 16-record groups of a block fall in different banks) and
// written as 16 blocks of 1 KiB.  Records kc0 * 64 .. (kc0 + gridDim.y) * 64 of the source.
__global__ __launch_bounds__(256) void tile_rows_kernel(const uint8_t *__restrict__ src, size_t pitch,
                                                        int8_t *__restrict__ dst, size_t CT, size_t KC,
                                                        size_t kc0) {
    __shared__ uint32_t t[64][65];
    const size_t kcl = blockIdx.y, ct0 = (size_t)blockIdx.x * 16;
    const int nct = (int)min((size_t)16, CT - ct0);
    {
        const int r = threadIdx.x >> 2, q = threadIdx.x & 3;  // record r, columns 64q .. 64q + 63
        const uint8_t *row = src + (kcl * 64 + r) * pitch + ct0 * 16;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int c16 = q * 4 + e;  // 16-column chunk = column tile ct0 + c16
            if (c16 < nct) {
                const uint4 v = *(const uint4 *)(row + c16 * 16);
                t[r][c16 * 4] = v.x;
                t[r][c16 * 4 + 1] = v.y;
                t[r][c16 * 4 + 2] = v.z;
                t[r][c16 * 4 + 3] = v.w;
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int it = 0; it < 4; it++) {
        const int b = it * 4 + (threadIdx.x >> 6);  // column tile ct0 + b
        if (b >= nct) continue;
        // block[16 lane + j] = byte (16 b + lane % 16) of record 16 (lane / 16) + j
        const int col = b * 16 + (lane & 15), r0 = 16 * (lane >> 4);
        uint32_t w[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint32_t x = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint8_t *rowb = (const uint8_t *)&t[r0 + 4 * i + j][0];
                x |= (uint32_t)rowb[col] << (8 * j);
            }
            w[i] = x ^ 0x80808080u;
        }
        *(uint4 *)(dst + ((ct0 + b) * KC + kc0 + kcl) * kBlockBytes + lane * 16) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// One byte plane j of the packing matrix, in the reference's AMX B-tile layout
// (src/amx_kernels.cpp:39-41, 331-336: plane region of KP * cols bytes, 32-column pairs,
// row r of a 1 KiB tile = 16 columns x 4 consecutive k), already offset by -128 ->
// blocks (ct, kc, j) at ((ct * KCp + kc) * 4 + j) * 1 KiB of this engine's layout.
__global__ void tile_u_kernel(const int8_t *__restrict__ plane, int8_t *__restrict__ dst, size_t CT,
                              size_t KCp, int j) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= CT * KCp * 64) return;
    const int lane = (int)(gid & 63);
    const size_t blk = gid >> 6, kc = blk % KCp, ct = blk / KCp;
    const int8_t *base = plane + ((ct / 2) * KCp + kc) * 2048 + (ct % 2) * 1024;
    const int g = lane >> 4, nn = lane & 15;
    uint4 v;
    v.x = *(const uint32_t *)(base + (4 * g + 0) * 64 + 4 * nn);
    v.y = *(const uint32_t *)(base + (4 * g + 1) * 64 + 4 * nn);
    v.z = *(const uint32_t *)(base + (4 * g + 2) * 64 + 4 * nn);
    v.w = *(const uint32_t *)(base + (4 * g + 3) * 64 + 4 * nn);
    *(uint4 *)(dst + ((ct * KCp + kc) * 4 + j) * kBlockBytes + lane * 16) = v;
}

// Packing-matrix build (src/amx_kernels.cpp:311-341) on the device, one workgroup per output
// column oz = (o, z) of the rho * n columns:
//   col[k n + zp(i, z)] += bold_t[o][(i tg + k) n + z],  col[k n + zb(i, z)] += bold_t_bar[...]
// for i < num_iter, k < tg, with zp = tables[((g - 1) / 2) n + z], zb = tables[((2n - g - 1) / 2) n
// + z], g = gen_pows[i]; sums wrap mod 2^64 as the reference's u64 column does (LDS atomics), then
// each entry is reduced mod q and its four bytes are stored as byte - 128 in the reference's
// plane layout (byte plane j of U[kk][oz] at ((j (CT/2) + ct/2) KC + kc) 2048 + (ct % 2) 1024 +
// r 64 + 4 nn + t, ct = oz / 16, nn = oz % 16, kc = kk / 64, r = (kk % 64) / 4, t = kk % 4), one u32
// of four consecutive kk per plane.  Table entries or generator powers outside [0, n) / [1, 2n),
// which the reference would use as out-of-range indices, are skipped.  Groups of 16 consecutive
// columns (one tile row of the output) run on one XCD, so the strided reads of bold_t share L2
// lines and the 16 u32 writes of a tile row meet in one L2.  LDS: K u64.
__global__ __launch_bounds__(256) void pack_build_kernel(const uint64_t *__restrict__ bt, const uint64_t *__restrict__ bb,
                                                         size_t out_stride, const uint32_t *__restrict__ tables,
                                                         const uint32_t *__restrict__ gen_pows, int num_iter, int tg,
                                                         int n, int cols, uint64_t q, int8_t *__restrict__ ub) {
    extern __shared__ unsigned long long colk[];  // K = tg * n
    const int L = blockIdx.x;
    const int oz = cols % 128 == 0 ? ((L >> 3) & 15) + 16 * ((L & 7) + 8 * (L >> 7)) : L;
    const int o = oz / n, z = oz % n, K = tg * n, KC = K / 64, CT = cols / 16;
    for (int r = threadIdx.x; r < K; r += blockDim.x) colk[r] = 0;
    __syncthreads();
    const uint64_t *bto = bt + (size_t)o * out_stride + z, *bbo = bb + (size_t)o * out_stride + z;
    for (int w = threadIdx.x; w < num_iter * tg; w += blockDim.x) {
        const int i = w / tg, k = w % tg;
        const uint32_t g = gen_pows[i];
        if (g < 1 || g >= 2u * (uint32_t)n) continue;
        const uint32_t zp = tables[(size_t)((g - 1) / 2) * n + z], zb = tables[(size_t)((2 * n - g - 1) / 2) * n + z];
        // (i tg + k) n + z = w n + z
        if (zp < (uint32_t)n) atomicAdd(&colk[k * n + zp], (unsigned long long)bto[(size_t)w * n]);
        if (zb < (uint32_t)n) atomicAdd(&colk[k * n + zb], (unsigned long long)bbo[(size_t)w * n]);
    }
    __syncthreads();
    const int ct = oz / 16, nn = oz % 16;
    for (int q4 = threadIdx.x; q4 < K / 4; q4 += blockDim.x) {
        const int kk0 = 4 * q4, kc = kk0 / 64, r = (kk0 % 64) / 4;
        uint32_t wj[4] = {0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const uint64_t v = colk[kk0 + t] % q;
#pragma unroll
            for (int j = 0; j < 4; j++) wj[j] |= ((uint32_t)((v >> (8 * j)) & 0xff) ^ 0x80u) << (8 * t);
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            *(uint32_t *)(ub + (((size_t)j * (CT / 2) + ct / 2) * KC + kc) * 2048 + (ct % 2) * 1024 + r * 64 + 4 * nn) =
                wj[j];
    }
}

__device__ __forceinline__ int sum_i8x4(uint32_t w) {
    return (int)(int8_t)(w & 0xff) + (int)(int8_t)((w >> 8) & 0xff) + (int)(int8_t)((w >> 16) & 0xff) +
           (int)(int8_t)(w >> 24);
}

// Column correction of the packing: ucorr[col] = 128E * sum_kk U[kk][col] mod Q.
// One workgroup per 16-column tile, wave j sums plane j.
__global__ __launch_bounds__(256) void u_colsum_kernel(const int8_t *__restrict__ ut, uint32_t *__restrict__ ucorr,
                                                       size_t KCp, uint32_t k128e) {
    __shared__ long long part[4][16];
    const size_t ct = blockIdx.x;
    const int j = threadIdx.x >> 6, lane = threadIdx.x & 63;
    long long s = 0;
    for (size_t kc = 0; kc < KCp; kc++) {
        const uint4 v = *(const uint4 *)(ut + ((ct * KCp + kc) * 4 + j) * kBlockBytes + lane * 16);
        s += sum_i8x4(v.x) + sum_i8x4(v.y) + sum_i8x4(v.z) + sum_i8x4(v.w) + 16 * 128;
    }
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    if (lane < 16) part[j][lane] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        const int nn = threadIdx.x;
        const uint64_t usum = (uint64_t)part[0][nn] + ((uint64_t)part[1][nn] << 8) + ((uint64_t)part[2][nn] << 16) +
                              ((uint64_t)part[3][nn] << 24);
        ucorr[ct * 16 + nn] = mulmod(red64(usum), k128e);
    }
}

// ------------------------------------------------------------ per batch: checks ----

// Where the frames of a batch live: frame c (batch position) is slot sidx[c] of the arena,
// or slot c when no index array is given; each slot holds one frame at byte 7 so that the
// u32 payload starting at frame byte 9 is 16-byte aligned.
struct FrameSet {
    const uint8_t *slots;
    size_t stride;
    const int *sidx;
    __device__ __forceinline__ const uint8_t *slot(size_t c) const {
        return slots + (size_t)(sidx ? sidx[c] : (int)c) * stride;
    }
};


__device__ __forceinline__ uint32_t ld32_bytes(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// Request validation (src/amx_kernels.cpp:571-581): grid (B, S), workgroup (c, s) checks
// slice s of frame c's 2 KP key words.  Frames with a nonzero status on entry (rejected by
// the host: null or wrong length) keep it.  Every workgroup of a frame reads the same
// header, so a bad header gives status 2 from all of them and no key is checked; a key >= Q
// turns a 0 into 3 (compare-and-swap, so a status set on entry stands).  A workgroup that
// starts after another one's 3 returns at once.  Workgroup (c, 0) also clears the per-query
// sums that the prep kernel accumulates.
// dims: the batch dimensions {B, MT} in device memory when the launch is sized for a larger
// batch of the same bucket (workgroups past B return at once), else null.
__global__ __launch_bounds__(256) void validate_kernel(FrameSet fs, int *__restrict__ status, uint32_t kp,
                                                       uint32_t *__restrict__ qsum,
                                                       unsigned long long *__restrict__ ysum,
                                                       const int *__restrict__ dims) {
    __shared__ int st;
    const size_t c = blockIdx.x;
    if (dims && (int)c >= dims[0]) return;
    const unsigned s = blockIdx.y, S = gridDim.y;
    if (threadIdx.x == 0) {
        if (s == 0) {
            qsum[c] = 0;
            ysum[c] = 0;
        }
        st = __hip_atomic_load(&status[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (st != 0) return;
    const uint8_t *f = fs.slot(c) + 7;
    if (f[0] != 0 || ld32_bytes(f + 1) != kp || ld32_bytes(f + 5) != kp) {
        if (threadIdx.x == 0) status[c] = 2;
        return;
    }
    const uint4 *keys = (const uint4 *)(f + 9);
    const uint32_t nk = kp / 2, per = (nk + S - 1) / S, k0 = s * per, k1 = min(nk, k0 + per);
    int bad = 0;
    for (uint32_t i = k0 + threadIdx.x; i < k1; i += blockDim.x) {
        const uint4 v = keys[i];
        bad |= (v.x >= kQ32) | (v.y >= kQ32) | (v.z >= kQ32) | (v.w >= kQ32);
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0 && bad) atomicCAS(&status[c], 0, 3);
}
// Slices per frame of validate_kernel: about one 16-byte key load per thread.
inline unsigned validate_slices(size_t kp) {
    const size_t s = kp / 2 / 256;
    return (unsigned)(s < 1 ? 1 : s > 16 ? 16 : s);
}

// ---------------------------------------------------------- per batch: limb prep ----

__device__ __forceinline__ uint32_t pick_byte(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, int i) {
    const int s = 8 * i;
    return ((w0 >> s) & 0xffu) | (((w1 >> s) & 0xffu) << 8) | (((w2 >> s) & 0xffu) << 16) | (((w3 >> s) & 0xffu) << 24);
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// Clients whose limb rows one thread of the limb preparation writes: 2 makes each thread's
// stores one whole 128-byte line of an A block.
constexpr int kPrepClients = 2;

// Splits u32 words into four byte-limb rows of an MFMA A operand (rows 4c+i, K = nwords).
// SCAN: signed digits of q + 0x80808080 and sum(q) mod 2^32 into qsum.
// PACK: bytes of y offset by -128 and the exact sum(y) into ysum.
// grid = (ceil(KC * 4 / 256), Bp / kPrepClients): block (bx, by) takes the clients
// kPrepClients * by + q (q < kPrepClients); thread p of the x range takes k-step p >> 2, k-group
// p & 3 of each of them, 16 words of each client, and writes the 16 * kPrepClients contiguous
// 4-byte words those make in the A block (whole cache lines).  All the frame loads of a thread
// are issued before its first store.  Clients c >= B or with status != 0 get zero rows.
// limb_prep_kernel runs the grid, prep_kernel runs both limb preparations and the accumulator
// clear in one launch.
template <bool SCAN>
__device__ __forceinline__ void limb_prep_block(FrameSet fs, size_t word_off, const int *__restrict__ status, int B,
                                                size_t KC, int8_t *__restrict__ dst, uint32_t *__restrict__ qsum,
                                                unsigned long long *__restrict__ ysum, unsigned bx, unsigned by) {
    constexpr int CP = kPrepClients;
    static_assert(4 % CP == 0, "clients of a thread lie in one A tile");
    __shared__ unsigned long long red[16][CP];  // [wave][client of the block]
    const int c0 = CP * (int)by;
    const size_t p = (size_t)bx * blockDim.x + threadIdx.x;
    const size_t kc = p >> 2;
    const int g = (int)(p & 3);
    unsigned long long s[CP];
#pragma unroll
    for (int q = 0; q < CP; q++) s[q] = 0;
    if (kc < KC) {
        bool ok[CP];
        uint4 v[CP][4];
#pragma unroll
        for (int q = 0; q < CP; q++) {
            const int c = c0 + q;
            ok[q] = c < B && status[c] == 0;
#pragma unroll
            for (int e = 0; e < 4; e++) v[q][e] = make_uint4(0, 0, 0, 0);
            if (ok[q]) {
                const uint4 *src = (const uint4 *)(fs.slot(c) + 16 + word_off * 4) + (kc * 64 + g * 16) / 4;
#pragma unroll
                for (int e = 0; e < 4; e++) v[q][e] = src[e];
            }
        }
        uint4 *blk = (uint4 *)(dst + ((size_t)(c0 >> 2) * KC + kc) * kBlockBytes) + 16 * g + 4 * (c0 & 3);
#pragma unroll
        for (int q = 0; q < CP; q++) {
            uint32_t w[16];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                w[4 * e] = v[q][e].x; w[4 * e + 1] = v[q][e].y; w[4 * e + 2] = v[q][e].z; w[4 * e + 3] = v[q][e].w;
            }
            uint4 o[4] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
            if (ok[q]) {
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    s[q] += w[e];
                    if (SCAN) w[e] += 0x80808080u;
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    o[i].x = pick_byte(w[0], w[1], w[2], w[3], i) ^ 0x80808080u;
                    o[i].y = pick_byte(w[4], w[5], w[6], w[7], i) ^ 0x80808080u;
                    o[i].z = pick_byte(w[8], w[9], w[10], w[11], i) ^ 0x80808080u;
                    o[i].w = pick_byte(w[12], w[13], w[14], w[15], i) ^ 0x80808080u;
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) blk[4 * q + i] = o[i];
        }
    }
    // the clients' sums over the block: waves by shuffles, then across waves
#pragma unroll
    for (int q = 0; q < CP; q++)
        for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_xor(s[q], o);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < CP; q++) red[threadIdx.x >> 6][q] = s[q];
    __syncthreads();
    if (threadIdx.x < CP) {
        const int c = c0 + (int)threadIdx.x;
        unsigned long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += red[i][threadIdx.x];
        if (c < B && status[c] == 0) {
            if (SCAN) atomicAdd(&qsum[c], (uint32_t)t);
            else atomicAdd(&ysum[c], t);
        }
    }
}
template <bool SCAN>
__global__ __launch_bounds__(256) void limb_prep_kernel(FrameSet fs, size_t word_off, const int *__restrict__ status, int B,
                                                        size_t KC, int8_t *__restrict__ dst,
                                                        uint32_t *__restrict__ qsum,
                                                        unsigned long long *__restrict__ ysum) {
    limb_prep_block<SCAN>(fs, word_off, status, B, KC, dst, qsum, ysum, blockIdx.x, blockIdx.y);
}

// One launch for the per-batch preparation after validation: blocks [0, sx * sy) prepare
// the query limbs (limb_prep_block<true>, grid sx x sy), the next px * py the key limbs
// (limb_prep_block<false>), the last nfill clear the scan accumulators (acc4 16-byte
// words) and the first of them the step kernel's counters.  The three parts are
// independent, so they run side by side.
struct PrepArgs {
    FrameSet fs;
    const int *status;
    int B;
    size_t scan_off, KC, KCp;
    int8_t *qa, *ya;
    uint32_t *qsum;
    unsigned long long *ysum;
    uint4 *acc;
    size_t acc4;
    unsigned sx, sy, px, py, nfill;
    uint32_t *sync;   // the step kernel's counters, cleared by the first fill block
    unsigned nsync;
    const int *dims;  // {B, MT} in device memory (launch sized for the batch's bucket), or null
};
__global__ __launch_bounds__(256) void prep_kernel(PrepArgs p) {
    unsigned b = blockIdx.x;
    const int B = p.dims ? p.dims[0] : p.B;  // rows from B on get zero limbs
    if (b < p.sx * p.sy) {
        limb_prep_block<true>(p.fs, p.scan_off, p.status, B, p.KC, p.qa, p.qsum, p.ysum, b % p.sx, b / p.sx);
        return;
    }
    b -= p.sx * p.sy;
    if (b < p.px * p.py) {
        limb_prep_block<false>(p.fs, 0, p.status, B, p.KCp, p.ya, p.qsum, p.ysum, b % p.px, b / p.px);
        return;
    }
    b -= p.px * p.py;
    if (b == 0)
        for (unsigned i = threadIdx.x; i < p.nsync; i += blockDim.x) p.sync[i] = 0;
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (size_t i = (size_t)b * blockDim.x + threadIdx.x; i < p.acc4; i += (size_t)p.nfill * blockDim.x) p.acc[i] = z;
}

// ------------------------------------------------------------- setup: the hint --

// Query limbs of the offline hint (the reference's answer_hint_ring, src/server.rs:738-768,
// with multiply_with_db_ring, :656-709): hint[z][col] = sum_r (a_r * d_{col,r})[z] in
// Z_Q[x]/(x^n + 1), where d_{col,r}[j] is the byte of column col in record r n + j.  Over the
// records this is a contraction with n queries z whose words are q_z[r n + j] = a_r[z - j] for
// j <= z and -a_r[n + z - j] otherwise (mod Q).  Each word is centred (q' = q - Q above
// 0x7F7F7F7F, so q' + 0x80808080 lies in [0, 2^32)) and split into four signed base-256 digits
// whose weighted sum is q' exactly, the limb rows of query z in the scan's A-operand layout;
// the exact sum S_z of the centred words goes to ssum[z] (two's complement).
// Grid (ceil(KC * 4 / 256), n); n a power of two of at least 16, so the 16 words of a thread
// lie in one polynomial.
__global__ __launch_bounds__(256) void hint_limb_kernel(const uint64_t *__restrict__ a, int n, size_t KC,
                                                        int8_t *__restrict__ dst,
                                                        unsigned long long *__restrict__ ssum) {
    __shared__ unsigned long long red[8];
    const int z = (int)blockIdx.y;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t kc = p >> 2;
    const int g = (int)(p & 3);
    unsigned long long s = 0;
    if (kc < KC) {
        const size_t k0 = kc * 64 + (size_t)g * 16;
        const size_t r = k0 / (size_t)n;
        const int j0 = (int)(k0 - r * (size_t)n);
        const uint64_t *ar = a + r * (size_t)n;
        uint32_t w[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int j = j0 + e;
            uint64_t v;
            if (j <= z) {
                v = ar[z - j] % kQ;
            } else {
                v = ar[n + z - j] % kQ;
                v = v ? kQ - v : 0;
            }
            const int64_t c = v <= 0x7F7F7F7Full ? (int64_t)v : (int64_t)v - (int64_t)kQ;
            s += (unsigned long long)c;
            w[e] = (uint32_t)(c + 0x80808080ll);
        }
        uint4 o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            o[i].x = pick_byte(w[0], w[1], w[2], w[3], i) ^ 0x80808080u;
            o[i].y = pick_byte(w[4], w[5], w[6], w[7], i) ^ 0x80808080u;
            o[i].z = pick_byte(w[8], w[9], w[10], w[11], i) ^ 0x80808080u;
            o[i].w = pick_byte(w[12], w[13], w[14], w[15], i) ^ 0x80808080u;
        }
        uint4 *blk = (uint4 *)(dst + ((size_t)(z >> 2) * KC + kc) * kBlockBytes);
#pragma unroll
        for (int i = 0; i < 4; i++) blk[(z & 3) * 4 + i + 16 * g] = o[i];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) atomicAdd(&ssum[z], s);
}

// hint[z * C + col] = acc + 128 S_z (mod Q): acc holds sum_k d'_k q'_k mod Q (d' = d - 128, the
// database as stored), so adding 128 sum_k q'_k gives sum_k d_k q_k.
__global__ void hint_finish_kernel(uint64_t *__restrict__ out, const unsigned long long *__restrict__ acc,
                                   const unsigned long long *__restrict__ ssum, size_t C, size_t total) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t S = (int64_t)ssum[t / C] % (int64_t)kQ;
    const uint32_t sq = (uint32_t)(S < 0 ? S + (int64_t)kQ : S);
    out[t] = addmod((uint32_t)(acc[t] % kQ), mulmod(sq, 128u));
}

// Writes the dimensions {B, MT} of the batch a bucket-sized launch chain answers: the first
// node of a captured chain, whose arguments each replay sets to the batch's own.
__global__ void batch_dims_kernel(int *__restrict__ dims, int B, int MT) {
    if (threadIdx.x == 0) {
        dims[0] = B;
        dims[1] = MT;
    }
}

// ------------------------------------------------------------------- contractions ----

// Maps a 1-D workgroup id onto (x, y, z) tiles so that the workgroups of one column tile
// x (all y, same z) run back to back on the same XCD (dispatch round-robins workgroups
// over the 8 XCDs): they then share the database slice in that XCD's L2.  With more than
// eight row blocks (and whole groups of them and of columns) the workgroups an XCD runs
// together cover 8 row blocks x 4 column tiles instead of every row block of one column:
// 12 operand streams per k-step for 32 workgroups instead of gy + 1.
struct TileMap {
    int x, y, z;
    bool valid;
};
__device__ __forceinline__ TileMap map_tile_at(int L, int gx, int gy, int gz) {
    const int xcd = L & 7;
    int t = L >> 3;
    const int gxg = (gx + 7) >> 3;
    TileMap m;
    if (gy > 8 && gy % 8 == 0 && gxg % 4 == 0) {
        const int yy = t & 7;
        t >>= 3;
        const int xx = t & 3;
        t >>= 2;
        const int ry = t % (gy >> 3);
        t /= gy >> 3;
        const int cx = t % (gxg >> 2);
        m.y = ry * 8 + yy;
        m.x = ((cx << 2) + xx) * 8 + xcd;
        m.z = t / (gxg >> 2);
    } else {
        m.y = t % gy;
        t /= gy;
        m.x = (t % gxg) * 8 + xcd;
        m.z = t / gxg;
    }
    m.valid = m.x < gx && m.z < gz;
    return m;
}
__device__ __forceinline__ TileMap map_tile(int gx, int gy, int gz) { return map_tile_at(blockIdx.x, gx, gy, gz); }

struct GemmArgs {
    int KC;        // 64-deep k-steps of the contraction
    int MT;        // A row tiles (16 rows = 4 queries)
    int NT;        // B tiles: scan = column tiles, pack = column tiles x 4 byte planes
    int gx, gy, gz;
    int C;         // columns (row stride of the outputs)
    uint32_t *out;                      // scan: acc (u32, += when split), pack: beta (mod Q)
    const uint32_t *ucorr;              // pack: column corrections
    const unsigned long long *ysum;     // pack: per-query key sums
    uint32_t k128e, ycorr0;
    // batch dimensions in device memory {B, MT} (a launch sized for a larger batch of the same
    // bucket, from a captured graph), or null: MT above is the batch's own
    const int *dims;
    // scan only: when set, the exact limb sum of each (query, column) goes here mod Q (added
    // when split) instead of mod 2^32 to `out` (the offline hint, whose query limbs sum to the
    // centred words exactly)
    unsigned long long *out64;
};
// A tiles of the batch: from the device dimensions when given, else the launch's own MT.
__device__ __forceinline__ int batch_tiles(const GemmArgs &g) { return g.dims ? g.dims[1] : g.MT; }

// B operand block of tile t at k-step k (i32x4 units).  Scan: database tile t.  Pack: tile
// t = 4 * column tile + plane j of the packing matrix.
template <bool PACK>
__device__ __forceinline__ const i32x4 *b_block(const i32x4 *b, int t, int KC) {
    if (PACK) return b + (((size_t)(t >> 2) * KC) * 4 + (t & 3)) * 64;
    return b + (size_t)t * KC * 64;
}
template <bool PACK>
constexpr int b_kstride() { return PACK ? 256 : 64; }

__device__ __forceinline__ uint32_t recombine_u32(i32x4 a) {
    return (uint32_t)a.x + ((uint32_t)a.y << 8) + ((uint32_t)a.z << 16) + ((uint32_t)a.w << 24);
}

// The four limb sums of query c against column col.
template <bool ATOMIC>
__device__ __forceinline__ void scan_store_at(const GemmArgs &g, int c, int col, i32x4 a) {
    if (g.out64) {  // sum_i 256^i P_i exactly (|P_i| <= 2^30 per split), reduced mod Q
        const int64_t t = (int64_t)a.x + ((int64_t)a.y << 8) + ((int64_t)a.z << 16) + ((int64_t)a.w << 24);
        unsigned long long *d = g.out64 + (size_t)c * g.C + col;
        if (ATOMIC) atomicAdd(d, (unsigned long long)red_i64(t));
        else *d = red_i64(t);
        return;
    }
    uint32_t *dst = g.out + (size_t)c * g.C + col;
    if (ATOMIC) atomicAdd(dst, recombine_u32(a));
    else *dst = recombine_u32(a);
}
// A 16 x 16 output tile (A tile rt, B tile ct): lane l holds the limb rows 4 (l >> 4) .. + 3
// of column l & 15.
template <bool ATOMIC>
__device__ __forceinline__ void scan_store(const GemmArgs &g, int rt, int ct, i32x4 a) {
    const int lane = threadIdx.x & 63;
    scan_store_at<ATOMIC>(g, 4 * rt + (lane >> 4), 16 * ct + (lane & 15), a);
}

// beta = sum_{ij} 2^{8(i+j)} G_ij + ucorr[col] + 128E * Ysum[c] - (128E)^2 KP  (mod Q), G_ij the
// product of limb row i of query c with byte plane j of the packing matrix.
// pack_plane: sum_i 2^{8(i+j)} G_ij over the four limb rows i (a = G_0j .. G_3j), with 2^{8s}
// for s = 4, 5, 6 replaced by its residue 12287 * 2^{8(s-4)} mod Q; |G_ij| <= 2^30 (tg * n <=
// 65536), so one plane's sum stays below 2^60 and the four planes' below 2^62.
__device__ __forceinline__ int64_t pack_plane(i32x4 a, int j) {
    const int64_t w[7] = {1, 1ll << 8, 1ll << 16, 1ll << 24, 12287, 12287ll << 8, 12287ll << 16};
    return a.x * w[j] + a.y * w[j + 1] + a.z * w[j + 2] + a.w * w[j + 3];
}
__device__ __forceinline__ void pack_store_at(const GemmArgs &g, int c, int col, int64_t t) {
    const uint32_t yc = addmod(mulmod(red64(g.ysum[c]), g.k128e), g.ycorr0);
    g.out[(size_t)c * g.C + col] = addmod(addmod(red_i64(t), g.ucorr[col]), yc);
}
// a[j] holds G_ij of plane j for the four limb rows i of this lane's query.
__device__ __forceinline__ void pack_store(const GemmArgs &g, int rt, int ct, const i32x4 *a) {
    const int lane = threadIdx.x & 63;
    const int64_t t = pack_plane(a[0], 0) + pack_plane(a[1], 1) + pack_plane(a[2], 2) + pack_plane(a[3], 3);
    pack_store_at(g, 4 * rt + (lane >> 4), 16 * ct + (lane & 15), t);
}

// Streaming contraction for few queries (HBM-bound): operands straight from memory into
// registers, U k-steps of loads in flight per wave.  Workgroup = 4 waves; wave w of tile
// (x, y, z) computes A tiles [y*MB, y*MB+MB) x B tiles [(4x+w)*NBT, +NBT) over the k-range
// of split z (at most kFoldSteps steps, so the i32 accumulators cannot overflow).
template <int MB, int NBT, int U, bool PACK, bool ATOMIC>
__device__ __forceinline__ void stream_tile(const i32x4 *__restrict__ a_src, const i32x4 *__restrict__ b_src,
                                            const GemmArgs &g, int L) {
    static_assert(!PACK || NBT % 4 == 0, "packing tiles come in groups of four planes");
    const TileMap tm = map_tile_at(L, g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bt0 = (tm.x * 4 + wave) * NBT, mt0 = tm.y * MB;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    constexpr int BS = b_kstride<PACK>();
    const i32x4 *ap[MB];
    const i32x4 *bp[NBT];
#pragma unroll
    for (int m = 0; m < MB; m++) ap[m] = a_src + (size_t)min(mt0 + m, MT - 1) * g.KC * 64 + lane;
#pragma unroll
    for (int n = 0; n < NBT; n++) bp[n] = b_block<PACK>(b_src, min(bt0 + n, g.NT - 1), g.KC) + lane;
    i32x4 acc[MB][NBT];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NBT; n++) acc[m][n] = i32x4{0, 0, 0, 0};
    int k = kb;
    for (; k + U <= ke; k += U) {
        i32x4 a[U][MB], b[U][NBT];
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int m = 0; m < MB; m++) a[u][m] = ap[m][(size_t)(k + u) * 64];
#pragma unroll
            for (int n = 0; n < NBT; n++) b[u][n] = __builtin_nontemporal_load(bp[n] + (size_t)(k + u) * BS);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int n = 0; n < NBT; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u][m], b[u][n], acc[m][n], 0, 0, 0);
    }
    for (; k < ke; k++) {
        i32x4 a[MB], b[NBT];
#pragma unroll
        for (int m = 0; m < MB; m++) a[m] = ap[m][(size_t)k * 64];
#pragma unroll
        for (int n = 0; n < NBT; n++) b[n] = __builtin_nontemporal_load(bp[n] + (size_t)k * BS);
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NBT; n++) acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[m], b[n], acc[m][n], 0, 0, 0);
    }
#pragma unroll
    for (int m = 0; m < MB; m++) {
        if (mt0 + m >= MT) continue;
        if (PACK) {
#pragma unroll
            for (int n = 0; n < NBT; n += 4)
                if (bt0 + n < g.NT) pack_store(g, mt0 + m, (bt0 + n) >> 2, &acc[m][n]);
        } else {
#pragma unroll
            for (int n = 0; n < NBT; n++)
                if (bt0 + n < g.NT) scan_store<ATOMIC>(g, mt0 + m, bt0 + n, acc[m][n]);
        }
    }
}
template <int MB, int NBT, int U, bool PACK, bool ATOMIC>
__global__ __launch_bounds__(256) void stream_kernel(const i32x4 *__restrict__ a_src,
                                                     const i32x4 *__restrict__ b_src, GemmArgs g) {
    stream_tile<MB, NBT, U, PACK, ATOMIC>(a_src, b_src, g, blockIdx.x);
}

// LDS-staged contraction for many queries (matrix-core bound).  Workgroup = WM x WN waves,
// wave tile TM x TN blocks, workgroup tile BM = WM*TM A tiles x BN = WN*TN B tiles.  Each
// k-step the workgroup copies its BM + BN operand blocks (1 KiB each) into one half of a
// double-buffered LDS area while the matrix cores work on the other half; one barrier per
// k-step.
template <int TM, int TN, int WM, int WN, bool PACK, bool ATOMIC>
__global__ __launch_bounds__(256, 1) void lds_kernel(const i32x4 *__restrict__ a_src, const i32x4 *__restrict__ b_src,
                                                     GemmArgs g) {
    static_assert(WM * WN == 4, "four waves");
    constexpr int BM = WM * TM, BN = WN * TN, NBLK = BM + BN;
    static_assert(NBLK % 4 == 0, "blocks split evenly over the waves");
    static_assert(!PACK || TN % 4 == 0, "packing tiles come in groups of four planes");
    constexpr int PW = NBLK / 4;
    constexpr int BS = b_kstride<PACK>();
    __shared__ i32x4 lds[2][NBLK][64];
    const TileMap tm = map_tile(g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int mt0 = tm.y * BM, bt0 = tm.x * BN;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    // this wave copies blocks wave, wave + 4, ... of the workgroup's NBLK
    const i32x4 *src[PW];
    int kst[PW];
#pragma unroll
    for (int p = 0; p < PW; p++) {
        const int blk = wave + 4 * p;
        if (blk < BM) {
            src[p] = a_src + (size_t)min(mt0 + blk, MT - 1) * g.KC * 64 + lane;
            kst[p] = 64;
        } else {
            src[p] = b_block<PACK>(b_src, min(bt0 + blk - BM, g.NT - 1), g.KC) + lane;
            kst[p] = BS;
        }
    }
    i32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int n = 0; n < TN; n++) acc[m][n] = i32x4{0, 0, 0, 0};
    {
        i32x4 pre[PW];
#pragma unroll
        for (int p = 0; p < PW; p++) pre[p] = src[p][(size_t)kb * kst[p]];
#pragma unroll
        for (int p = 0; p < PW; p++) lds[0][wave + 4 * p][lane] = pre[p];
    }
    __syncthreads();
    for (int k = kb; k < ke; k++) {
        const int buf = (k - kb) & 1;
        const bool more = k + 1 < ke;
        i32x4 pre[PW];
        if (more) {
#pragma unroll
            for (int p = 0; p < PW; p++) pre[p] = src[p][(size_t)(k + 1) * kst[p]];
        }
        i32x4 b[TN];
#pragma unroll
        for (int n = 0; n < TN; n++) b[n] = lds[buf][BM + wn * TN + n][lane];
#pragma unroll
        for (int m = 0; m < TM; m++) {
            const i32x4 a = lds[buf][wm * TM + m][lane];
#pragma unroll
            for (int n = 0; n < TN; n++) acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[n], acc[m][n], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int p = 0; p < PW; p++) lds[buf ^ 1][wave + 4 * p][lane] = pre[p];
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int rt = mt0 + wm * TM + m;
        if (rt >= MT) continue;
        if (PACK) {
#pragma unroll
            for (int n = 0; n < TN; n += 4) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) pack_store(g, rt, t >> 2, &acc[m][n]);
            }
        } else {
#pragma unroll
            for (int n = 0; n < TN; n++) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) scan_store<ATOMIC>(g, rt, t, acc[m][n]);
            }
        }
    }
}

// vmcnt(n) with the other counters left alone (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] |
// lgkmcnt[11:8] | vmcnt[5:4] << 14).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier without the memory fence of __syncthreads(): the fence would wait for
// every outstanding global load, including the prefetches of later stages.  The empty asm
// statements keep the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void stage_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

typedef __attribute__((address_space(3))) void lds_void;

// Multi-stage contraction for many queries (matrix-core bound).  Like lds_kernel (with 4 or
// 8 waves), but the operand blocks travel from memory straight into LDS (global_load_lds_dwordx4: lane l's
// 16 bytes land at 16 * l of the destination, which is exactly the block layout) through a
// ring of S stages, so S - 1 k-steps of loads are in flight while the matrix cores work
// and no VGPRs hold prefetched data.  Needs S * (BM + BN) KiB of dynamic LDS.
template <int TM, int TN, int WM, int WN, int S, bool PACK, bool ATOMIC, int KPB = 1, bool PRIO = false>
__device__ __forceinline__ void dma_tile(const i32x4 *__restrict__ a_src, const i32x4 *__restrict__ b_src,
                                         const GemmArgs &g, int L) {
    constexpr int NW = WM * WN;  // waves: 4 (one per SIMD) or 8 (two per SIMD)
    static_assert(NW == 4 || NW == 8, "four or eight waves");
    static_assert(S >= 2 && S <= 4, "stages");
    static_assert(KPB >= 1 && S - KPB >= KPB && S - 2 * KPB <= 2, "k-steps per barrier");
    constexpr int BM = WM * TM, BN = WN * TN, NBLK = BM + BN;
    static_assert(NBLK % NW == 0, "blocks split evenly over the waves");
    constexpr int AHEAD = S - KPB;  // stages in flight when a group of KPB steps is computed
    static_assert(!PACK || TN % 4 == 0, "packing tiles come in groups of four planes");
    constexpr int PW = NBLK / NW;
    constexpr int BS = b_kstride<PACK>();
    extern __shared__ i32x4 ring[];  // [S][NBLK][64]
    const TileMap tm = map_tile_at(L, g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int mt0 = tm.y * BM, bt0 = tm.x * BN;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    const i32x4 *src[PW];
    int kst[PW];
#pragma unroll
    for (int p = 0; p < PW; p++) {
        const int blk = wave + NW * p;
        if (blk < BM) {
            src[p] = a_src + (size_t)min(mt0 + blk, MT - 1) * g.KC * 64 + lane;
            kst[p] = 64;
        } else {
            src[p] = b_block<PACK>(b_src, min(bt0 + blk - BM, g.NT - 1), g.KC) + lane;
            kst[p] = BS;
        }
    }
    auto issue = [&](int k) {
        i32x4 *dst = ring + (size_t)(((k - kb) % S) * NBLK) * 64;
#pragma unroll
        for (int p = 0; p < PW; p++)
            __builtin_amdgcn_global_load_lds((const void *)(src[p] + (size_t)k * kst[p]),
                                             (lds_void *)(dst + (wave + NW * p) * 64), 16, 0, 0);
    };
    i32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int n = 0; n < TN; n++) acc[m][n] = i32x4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < AHEAD; s++)
        if (kb + s < ke) issue(kb + s);
    // Steps go in groups of KPB per barrier.  Group k0 needs stages k0 .. k0+KPB-1; the
    // stages issued after them (k0+KPB .. k0+AHEAD-1, those that exist) may stay in flight,
    // PW loads each.  After the barrier every wave is past the previous group, so its KPB
    // ring slots take the next stages.
    for (int k0 = kb; k0 < ke; k0 += KPB) {
        const int young = max(0, min(AHEAD - KPB, ke - k0 - KPB));
        if (young >= 2) wait_vmcnt<2 * PW>();
        else if (young == 1) wait_vmcnt<PW>();
        else wait_vmcnt<0>();
        stage_barrier();
#pragma unroll
        for (int j = 0; j < KPB; j++)
            if (k0 + AHEAD + j < ke) issue(k0 + AHEAD + j);
#pragma unroll
        for (int j = 0; j < KPB; j++) {
            if (KPB > 1 && k0 + j >= ke) break;
            const i32x4 *cur = ring + (size_t)(((k0 + j - kb) % S) * NBLK) * 64;
            i32x4 b[TN];
#pragma unroll
            for (int n = 0; n < TN; n++) b[n] = cur[(BM + wn * TN + n) * 64 + lane];
            // PRIO: a wave that has its operands issues its matrix work ahead of the other
            // wave of its SIMD (which is still waiting on LDS or at the barrier)
            if (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int m = 0; m < TM; m++) {
                const i32x4 a = cur[(wm * TM + m) * 64 + lane];
#pragma unroll
                for (int n = 0; n < TN; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[n], acc[m][n], 0, 0, 0);
            }
            if (PRIO) __builtin_amdgcn_s_setprio(0);
        }
    }
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int rt = mt0 + wm * TM + m;
        if (rt >= MT) continue;
        if (PACK) {
#pragma unroll
            for (int n = 0; n < TN; n += 4) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) pack_store(g, rt, t >> 2, &acc[m][n]);
            }
        } else {
#pragma unroll
            for (int n = 0; n < TN; n++) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) scan_store<ATOMIC>(g, rt, t, acc[m][n]);
            }
        }
    }
}
template <int TM, int TN, int WM, int WN, int S, bool PACK, bool ATOMIC, int KPB = 1, bool PRIO = false>
__global__ __launch_bounds__(64 * WM * WN, 1) void dma_kernel(const i32x4 *__restrict__ a_src,
                                                              const i32x4 *__restrict__ b_src, GemmArgs g) {
    dma_tile<TM, TN, WM, WN, S, PACK, ATOMIC, KPB, PRIO>(a_src, b_src, g, blockIdx.x);
}

// dma_kernel with the LDS reads of a k-step overlapped with the matrix work of the step
// before: the barrier that publishes stage k+1 is taken before the MFMAs of stage k, the
// fragments of stage k+1 are read into a second register set while those MFMAs run, and
// the loop alternates between the two sets.  Stage k+S-1 is then issued into the buffer of
// stage k-1, whose fragments were consumed before that barrier.  Needs S >= 3.
template <int TM, int TN, int WM, int WN, int S, bool PACK, bool ATOMIC>
__global__ __launch_bounds__(64 * WM * WN, 1) void dma2_kernel(const i32x4 *__restrict__ a_src,
                                                               const i32x4 *__restrict__ b_src, GemmArgs g) {
    constexpr int NW = WM * WN;
    static_assert(NW == 4 || NW == 8, "four or eight waves");
    static_assert(S >= 3 && S <= 4, "stages");
    constexpr int BM = WM * TM, BN = WN * TN, NBLK = BM + BN;
    static_assert(NBLK % NW == 0, "blocks split evenly over the waves");
    static_assert(!PACK || TN % 4 == 0, "packing tiles come in groups of four planes");
    constexpr int PW = NBLK / NW;
    constexpr int BS = b_kstride<PACK>();
    extern __shared__ i32x4 ring[];  // [S][NBLK][64]
    const TileMap tm = map_tile(g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int mt0 = tm.y * BM, bt0 = tm.x * BN;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    const i32x4 *src[PW];
    int kst[PW];
#pragma unroll
    for (int p = 0; p < PW; p++) {
        const int blk = wave + NW * p;
        if (blk < BM) {
            src[p] = a_src + (size_t)min(mt0 + blk, MT - 1) * g.KC * 64 + lane;
            kst[p] = 64;
        } else {
            src[p] = b_block<PACK>(b_src, min(bt0 + blk - BM, g.NT - 1), g.KC) + lane;
            kst[p] = BS;
        }
    }
    auto issue = [&](int k) {
        i32x4 *dst = ring + (size_t)(((k - kb) % S) * NBLK) * 64;
#pragma unroll
        for (int p = 0; p < PW; p++)
            __builtin_amdgcn_global_load_lds((const void *)(src[p] + (size_t)k * kst[p]),
                                             (lds_void *)(dst + (wave + NW * p) * 64), 16, 0, 0);
    };
    // waits until stage k is in LDS for every wave (`later` = stages issued after k)
    auto publish = [&](int later) {
        if (later >= 2) wait_vmcnt<2 * PW>();
        else if (later == 1) wait_vmcnt<PW>();
        else wait_vmcnt<0>();
        stage_barrier();
    };
    i32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int n = 0; n < TN; n++) acc[m][n] = i32x4{0, 0, 0, 0};
    i32x4 ca[TM], cb[TN], na[TM], nb[TN];
#define ALTO_READ(K, FA, FB)                                                             \
    {                                                                                    \
        const i32x4 *cur = ring + (size_t)((((K) - kb) % S) * NBLK) * 64;                \
        _Pragma("unroll") for (int n = 0; n < TN; n++) FB[n] = cur[(BM + wn * TN + n) * 64 + lane]; \
        _Pragma("unroll") for (int m = 0; m < TM; m++) FA[m] = cur[(wm * TM + m) * 64 + lane];      \
    }
#pragma unroll
    for (int s = 0; s < S - 1; s++)
        if (kb + s < ke) issue(kb + s);
    publish(min(S - 2, ke - 1 - kb));
    ALTO_READ(kb, ca, cb)
    // everything read so far (these fragments, the kernel arguments) is complete before the
    // loop, so that inside it only the next step's LDS reads are outstanding and the
    // matrix work on the current fragments does not wait for them
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    for (int k = kb; k < ke; k++) {
        if (k + 1 < ke) {
            publish(min(S - 3, ke - 2 - k));  // step k+1 visible; step k-1 fully consumed
            if (k + S - 1 < ke) issue(k + S - 1);
            ALTO_READ(k + 1, na, nb)
        }
#pragma unroll
        for (int m = 0; m < TM; m++)
#pragma unroll
            for (int n = 0; n < TN; n++) acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[m], cb[n], acc[m][n], 0, 0, 0);
#pragma unroll
        for (int m = 0; m < TM; m++) ca[m] = na[m];
#pragma unroll
        for (int n = 0; n < TN; n++) cb[n] = nb[n];
    }
#undef ALTO_READ
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int rt = mt0 + wm * TM + m;
        if (rt >= MT) continue;
        if (PACK) {
#pragma unroll
            for (int n = 0; n < TN; n += 4) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) pack_store(g, rt, t >> 2, &acc[m][n]);
            }
        } else {
#pragma unroll
            for (int n = 0; n < TN; n++) {
                const int t = bt0 + wn * TN + n;
                if (t < g.NT) scan_store<ATOMIC>(g, rt, t, acc[m][n]);
            }
        }
    }
}

// One 1 KiB operand block of zeros: the B block of the k-steps that the register-streaming
// contraction runs past the end of its range (their products are then zero).
__device__ __align__(16) i32x4 kZeroBlock[64];

// Contraction with the A blocks through LDS and the B blocks straight into registers, for
// many queries.  Workgroup = WM x WN waves (two per SIMD), wave tile TM A x TN B tiles (its
// accumulators in ordinary registers), workgroup tile BM = WM * TM A tiles x BN = WN * TN B
// tiles.  Each k-step the A blocks go from memory into an LDS ring (global_load_lds, shared
// by the WN waves of a row), while every wave loads its own TN B blocks into a register ring
// D k-steps ahead (the WM waves of a column load the same blocks; the second request hits the
// CU's L1).  Per k-step a CU moves BM KiB into LDS and reads WM * WN * TM KiB out of it:
// with 16 x 16 tiles 80 KiB of LDS traffic against 128 KiB for the all-LDS ring.  Each wave
// copies the same number of A blocks per step (the wait counts below assume it), so a ring
// slot holds BM rounded up to a multiple of the eight waves; the rows past BM are loaded
// from clamped addresses and never read.
//
// The A fragments of step k + 1 are read into the registers of step k's fragments right
// after the MFMAs that use them, so the barrier that publishes step k + 1 comes before step
// k's matrix work and the MFMAs do not wait on LDS.  Loads are issued in pairs P(t) = (A(t +
// 1) into the ring, B(t) into the register ring), one pair per step and D pairs ahead; at the
// top of step k, P(k) must have landed, so the D - 1 younger pairs stay in flight.  Ring
// slots: S = D + 2 (step k's fragments may still be in flight from LDS, step k + 1 is read, D
// are loading).  The loop runs whole groups of UNR steps (straight-line code: the compiler
// keeps the accumulators in place, and its own wait for the B registers is exact except at
// the top of each group, where it waits for every load): steps past the end of the range
// reload the last A block into slots no remaining step reads and take their B blocks from
// kZeroBlock, so they add nothing; all loads are waited for before the workgroup ends.
//
// WIDE: the same data on 32 x 32 x 32 MFMAs, each on a pair of A tiles and a pair of B tiles
// (half the instructions, each twice the work; the same operand bytes per wave and step).
// A 1 KiB block holds a 16 x 64 operand in four k-groups of 16 bytes (lane l: row l & 15,
// k-group l >> 4).  The 32 x 32 x 32 operand of half h of a step takes, in lane l, row (or
// column) l & 15 of tile (l >> 4) & 1 of the pair and k-group (l >> 5) + 2h: both operands use
// the same k-group for the same lane half, so the two halves sum over all 64 k of the step
// whatever order the hardware gives the bytes of a k-group.  Lane l of the i32x16 result holds
// column l & 31 and, in registers 4G .. 4G + 3, the limb rows 8G + 4 (l >> 5) .. + 3: query
// 2 (G & 1) + (l >> 5) of A tile G >> 1 of the pair.
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC, bool WIDE = false>
__device__ __forceinline__ void dmab_tile(const i32x4 *__restrict__ a_src, const i32x4 *__restrict__ b_src,
                                          const GemmArgs &g, int L) {
    constexpr int NW = WM * WN;
    static_assert(NW == 8, "two waves per SIMD");
    constexpr int BM = WM * TM;
    static_assert(!PACK || TN % 4 == 0, "packing tiles come in groups of four planes");
    static_assert(!WIDE || (TM % 2 == 0 && TN % 2 == 0), "32 x 32 tiles pair the 16 x 16 tiles");
    static_assert(D >= 2 && D <= 3, "prefetch distance");
    static_assert(UNR % D == 0, "whole register rings per loop iteration");
    constexpr int S = D + 2;
    constexpr int PA = (BM + NW - 1) / NW;  // A blocks each wave copies per step
    constexpr int BMP = PA * NW;            // A blocks per ring slot
    constexpr int PAIR = PA + TN;       // vector-memory instructions per step and wave
    static_assert(D * PAIR < 64, "vmcnt range");
    constexpr int BS = b_kstride<PACK>();
    extern __shared__ i32x4 ring[];     // [S][BMP][64]
    const TileMap tm = map_tile_at(L, g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int mt0 = tm.y * BM, bt0 = (tm.x * WN + wn) * TN;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    // fragment m (B fragment n) of a step: WIDE, half m & 1 of the tile pair m >> 1 (tile of
    // the pair: sel), else tile m; frag_at gives its 16-byte unit within the wave's blocks
    const int sel = WIDE ? (lane >> 4) & 1 : 0;
    const int e0 = WIDE ? (lane & 15) + 16 * (lane >> 5) : lane;
    auto frag_at = [&](int m) { return WIDE ? (2 * (m >> 1) + sel) * 64 + e0 + 32 * (m & 1) : m * 64 + lane; };
    const i32x4 *asrc[PA];
#pragma unroll
    for (int p = 0; p < PA; p++) asrc[p] = a_src + (size_t)min(mt0 + wave + NW * p, MT - 1) * g.KC * 64 + lane;
    const i32x4 *bsrc[TN];
#pragma unroll
    for (int n = 0; n < TN; n++) {
        const int t = WIDE ? bt0 + 2 * (n >> 1) + sel : bt0 + n;
        bsrc[n] = b_block<PACK>(b_src, min(t, g.NT - 1), g.KC) + (WIDE ? e0 + 32 * (n & 1) : lane);
    }
    const i32x4 *zero = kZeroBlock + lane;
    auto issue_a = [&](int t) {  // A(t) into its slot (t past the range: the last block)
        i32x4 *dst = ring + (size_t)(((t - kb) % S) * BMP) * 64;
        const size_t kk = (size_t)min(t, ke - 1);
#pragma unroll
        for (int p = 0; p < PA; p++)
            __builtin_amdgcn_global_load_lds((const void *)(asrc[p] + kk * 64), (lds_void *)(dst + (wave + NW * p) * 64),
                                             16, 0, 0);
    };
    auto load_b = [&](int t, i32x4 *b) {  // B(t) (t past the range: zeros)
        const bool live = t < ke;
#pragma unroll
        for (int n = 0; n < TN; n++) b[n] = *(live ? bsrc[n] + (size_t)t * BS : zero);
    };
    constexpr int AM = WIDE ? TM / 2 : TM, AN = WIDE ? TN / 2 : TN;
    typedef int AccT __attribute__((ext_vector_type(WIDE ? 16 : 4)));
    AccT acc[AM][AN];
#pragma unroll
    for (int m = 0; m < AM; m++)
#pragma unroll
        for (int n = 0; n < AN; n++) acc[m][n] = (AccT)0;
    if (ke > kb) {  // an empty record range leaves the accumulators zero
        i32x4 a[TM];
        i32x4 breg[D][TN];
        // prologue: A(kb), then the pairs P(kb) .. P(kb + D - 1), in this order (the wait
        // counts below and in the loop rely on it)
        issue_a(kb);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int d = 0; d < D; d++) {
            issue_a(kb + 1 + d);
            load_b(kb + d, breg[d]);
            __builtin_amdgcn_sched_barrier(0);
        }
        // every prologue load landed: once per workgroup, and the loop then starts from the
        // same pending loads on every entry (the compiler's own wait counts for the B
        // registers stay exact instead of draining every iteration)
        wait_vmcnt<0>();
        stage_barrier();
        {
            const i32x4 *cur = ring + (size_t)(wm * TM) * 64;
#pragma unroll
            for (int m = 0; m < TM; m++) a[m] = cur[frag_at(m)];
        }
        const int kend = kb + (ke - kb + UNR - 1) / UNR * UNR;
        for (int k0 = kb; k0 < kend; k0 += UNR) {
#pragma unroll
            for (int jj = 0; jj < UNR; jj++) {
                const int j = jj % D;
                const int k = k0 + jj;
                wait_vmcnt<(D - 1) * PAIR>();  // P(k): A(k + 1) and B(k)
                stage_barrier();               // A(k + 1) visible; slot of step k - 1 free
                const i32x4 *nxt = ring + (size_t)(((k + 1 - kb) % S) * BMP + wm * TM) * 64;
                // kept in this order (scheduling barriers): hoisting a fragment read or the B
                // loads above the MFMAs that free their registers would double the live
                // operands
                if constexpr (WIDE) {
#pragma unroll
                    for (int h = 0; h < 2; h++)
#pragma unroll
                        for (int mp = 0; mp < AM; mp++) {
#pragma unroll
                            for (int np = 0; np < AN; np++)
                                acc[mp][np] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[2 * mp + h], breg[j][2 * np + h],
                                                                                    acc[mp][np], 0, 0, 0);
                            a[2 * mp + h] = nxt[frag_at(2 * mp + h)];
                            __builtin_amdgcn_sched_barrier(0);
                        }
                } else {
#pragma unroll
                    for (int m = 0; m < TM; m++) {
#pragma unroll
                        for (int n = 0; n < TN; n++)
                            acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[m], breg[j][n], acc[m][n], 0, 0, 0);
                        a[m] = nxt[frag_at(m)];  // step k + 1's fragment (past the range: unused)
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                issue_a(k + 1 + D);  // into the slot of step k - 1
                load_b(k + D, breg[j]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wait_vmcnt<0>();  // loads past the range land before the workgroup's LDS is released
    }
    if constexpr (WIDE) {
        const int h = lane >> 5, col = lane & 15;
#pragma unroll
        for (int mp = 0; mp < AM; mp++)
#pragma unroll
            for (int G = 0; G < 4; G++) {
                const int rt = mt0 + wm * TM + 2 * mp + (G >> 1);
                const int c = 4 * rt + 2 * (G & 1) + h;
                if (PACK) {
                    // B pair np holds planes 2 (np & 1) + sel of column tile (bt0 + 2 np) / 4: a
                    // lane has two of the four planes, its partner (lane ^ 16) the other two
#pragma unroll
                    for (int np = 0; np < AN; np += 2) {
                        const i32x4 p0 = {acc[mp][np][4 * G], acc[mp][np][4 * G + 1], acc[mp][np][4 * G + 2],
                                          acc[mp][np][4 * G + 3]};
                        const i32x4 p1 = {acc[mp][np + 1][4 * G], acc[mp][np + 1][4 * G + 1],
                                          acc[mp][np + 1][4 * G + 2], acc[mp][np + 1][4 * G + 3]};
                        int64_t t = pack_plane(p0, sel) + pack_plane(p1, 2 + sel);
                        t += __shfl_xor(t, 16);
                        const int tt = bt0 + 2 * np;
                        if (sel == 0 && rt < MT && tt < g.NT) pack_store_at(g, c, 16 * (tt >> 2) + col, t);
                    }
                } else {
                    if (rt >= MT) continue;
#pragma unroll
                    for (int np = 0; np < AN; np++) {
                        const int t = bt0 + 2 * np + sel;
                        const i32x4 v = {acc[mp][np][4 * G], acc[mp][np][4 * G + 1], acc[mp][np][4 * G + 2],
                                         acc[mp][np][4 * G + 3]};
                        if (t < g.NT) scan_store_at<ATOMIC>(g, c, 16 * t + col, v);
                    }
                }
            }
    } else {
#pragma unroll
        for (int m = 0; m < TM; m++) {
            const int rt = mt0 + wm * TM + m;
            if (rt >= MT) continue;
            if (PACK) {
#pragma unroll
                for (int n = 0; n < TN; n += 4) {
                    const int t = bt0 + n;
                    if (t < g.NT) pack_store(g, rt, t >> 2, &acc[m][n]);
                }
            } else {
#pragma unroll
                for (int n = 0; n < TN; n++) {
                    const int t = bt0 + n;
                    if (t < g.NT) scan_store<ATOMIC>(g, rt, t, acc[m][n]);
                }
            }
        }
    }
}
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC, bool WIDE = false>
__global__ __launch_bounds__(64 * WM * WN, 1) void dmab_kernel(const i32x4 *__restrict__ a_src,
                                                               const i32x4 *__restrict__ b_src, GemmArgs g) {
    dmab_tile<TM, TN, WM, WN, D, UNR, PACK, ATOMIC, WIDE>(a_src, b_src, g, blockIdx.x);
}

// The register-streaming contraction with the A fragments read from LDS just before their
// MFMAs (two in flight) instead of a whole k-step ahead: 24 fewer registers per wave (a deeper
// B register ring, D = 3 or 4, fits in them, but measured no faster than D = 2).  Loads are
// issued in pairs P(t) = (A(t) into ring slot t % S, B(t) into register slot t % D), D pairs
// ahead; at the top of step k P(k) must have landed ((D - 1) pairs stay in flight), and the
// barrier then publishes A(k) and frees the slot of step k - 1, into which A(k + D) goes at
// the end of step k (S = D + 1 slots).  Steps past the end of the range work as in dmab_tile.
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC>
__device__ __forceinline__ void dmac_tile(const i32x4 *__restrict__ a_src, const i32x4 *__restrict__ b_src,
                                          const GemmArgs &g, int L) {
    constexpr int NW = WM * WN;
    static_assert(NW == 8, "two waves per SIMD");
    constexpr int BM = WM * TM;
    static_assert(!PACK || TN % 4 == 0, "packing tiles come in groups of four planes");
    static_assert(D >= 2 && D <= 4, "prefetch distance");
    static_assert(UNR % D == 0, "whole register rings per loop iteration");
    constexpr int S = D + 1;
    constexpr int PA = (BM + NW - 1) / NW;  // A blocks each wave copies per step
    constexpr int BMP = PA * NW;            // A blocks per ring slot
    constexpr int PAIR = PA + TN;           // vector-memory instructions per step and wave
    static_assert(D * PAIR < 64, "vmcnt range");
    constexpr int BS = b_kstride<PACK>();
    extern __shared__ i32x4 ring[];         // [S][BMP][64]
    const TileMap tm = map_tile_at(L, g.gx, g.gy, g.gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = wave / WN, wn = wave % WN;
    const int mt0 = tm.y * BM, bt0 = (tm.x * WN + wn) * TN;
    const int MT = batch_tiles(g);
    if (mt0 >= MT) return;  // a row block past the batch (launch sized for its bucket)
    const int kb = (int)((long long)g.KC * tm.z / g.gz), ke = (int)((long long)g.KC * (tm.z + 1) / g.gz);
    const i32x4 *asrc[PA];
#pragma unroll
    for (int p = 0; p < PA; p++) asrc[p] = a_src + (size_t)min(mt0 + wave + NW * p, MT - 1) * g.KC * 64 + lane;
    const i32x4 *bsrc[TN];
#pragma unroll
    for (int n = 0; n < TN; n++) bsrc[n] = b_block<PACK>(b_src, min(bt0 + n, g.NT - 1), g.KC) + lane;
    const i32x4 *zero = kZeroBlock + lane;
    auto issue_a = [&](int t) {  // A(t) into its slot (t past the range: the last block)
        i32x4 *dst = ring + (size_t)(((t - kb) % S) * BMP) * 64;
        const size_t kk = (size_t)min(t, ke - 1);
#pragma unroll
        for (int p = 0; p < PA; p++)
            __builtin_amdgcn_global_load_lds((const void *)(asrc[p] + kk * 64), (lds_void *)(dst + (wave + NW * p) * 64),
                                             16, 0, 0);
    };
    auto load_b = [&](int t, i32x4 *b) {  // B(t) (t past the range: zeros)
        const bool live = t < ke;
#pragma unroll
        for (int n = 0; n < TN; n++) b[n] = *(live ? bsrc[n] + (size_t)t * BS : zero);
    };
    i32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; m++)
#pragma unroll
        for (int n = 0; n < TN; n++) acc[m][n] = i32x4{0, 0, 0, 0};
    if (ke > kb) {
        i32x4 breg[D][TN];
#pragma unroll
        for (int d = 0; d < D; d++) {
            issue_a(kb + d);
            load_b(kb + d, breg[d]);
            __builtin_amdgcn_sched_barrier(0);
        }
        const int kend = kb + (ke - kb + UNR - 1) / UNR * UNR;
        for (int k0 = kb; k0 < kend; k0 += UNR) {
#pragma unroll
            for (int jj = 0; jj < UNR; jj++) {
                const int j = jj % D;
                const int k = k0 + jj;
                wait_vmcnt<(D - 1) * PAIR>();  // P(k): A(k) and B(k)
                stage_barrier();               // A(k) visible; slot of step k - 1 free
                const i32x4 *cur = ring + (size_t)(((k - kb) % S) * BMP + wm * TM) * 64 + lane;
                i32x4 af[2];
                af[0] = cur[0];
#pragma unroll
                for (int m = 0; m < TM; m++) {
                    if (m + 1 < TM) af[(m + 1) & 1] = cur[(m + 1) * 64];
#pragma unroll
                    for (int n = 0; n < TN; n++)
                        acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[m & 1], breg[j][n], acc[m][n], 0, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                issue_a(k + D);  // into the slot of step k - 1
                load_b(k + D, breg[j]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        wait_vmcnt<0>();  // loads past the range land before the workgroup's LDS is released
    }
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int rt = mt0 + wm * TM + m;
        if (rt >= MT) continue;
        if (PACK) {
#pragma unroll
            for (int n = 0; n < TN; n += 4) {
                const int t = bt0 + n;
                if (t < g.NT) pack_store(g, rt, t >> 2, &acc[m][n]);
            }
        } else {
#pragma unroll
            for (int n = 0; n < TN; n++) {
                const int t = bt0 + n;
                if (t < g.NT) scan_store<ATOMIC>(g, rt, t, acc[m][n]);
            }
        }
    }
}
template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC>
__global__ __launch_bounds__(64 * WM * WN, 1) void dmac_kernel(const i32x4 *__restrict__ a_src,
                                                               const i32x4 *__restrict__ b_src, GemmArgs g) {
    dmac_tile<TM, TN, WM, WN, D, UNR, PACK, ATOMIC>(a_src, b_src, g, blockIdx.x);
}

// ---------------------------------------------------- both contractions in one launch --

// Tile shapes as types, so that one kernel can run the tiles of two contractions.
template <int MB, int NBT, int U, bool PACK, bool ATOMIC>
struct StreamShape {
    static constexpr int kThreads = 256;
    static constexpr size_t kLds = 0;
    static __device__ __forceinline__ void tile(const i32x4 *a, const i32x4 *b, const GemmArgs &g, int L) {
        stream_tile<MB, NBT, U, PACK, ATOMIC>(a, b, g, L);
    }
};
template <int TM, int TN, int WM, int WN, int S, bool PACK, bool ATOMIC, int KPB>
struct DmaShape {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr size_t kLds = (size_t)S * (WM * TM + WN * TN) * kBlockBytes;
    static __device__ __forceinline__ void tile(const i32x4 *a, const i32x4 *b, const GemmArgs &g, int L) {
        dma_tile<TM, TN, WM, WN, S, PACK, ATOMIC, KPB>(a, b, g, L);
    }
};

template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC, bool WIDE = false>
struct DmabShape {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr size_t kLds = (size_t)(D + 2) * ((WM * TM + 7) / 8 * 8) * kBlockBytes;
    static __device__ __forceinline__ void tile(const i32x4 *a, const i32x4 *b, const GemmArgs &g, int L) {
        dmab_tile<TM, TN, WM, WN, D, UNR, PACK, ATOMIC, WIDE>(a, b, g, L);
    }
};

template <int TM, int TN, int WM, int WN, int D, int UNR, bool PACK, bool ATOMIC>
struct DmacShape {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr size_t kLds = (size_t)(D + 1) * ((WM * TM + 7) / 8 * 8) * kBlockBytes;
    static __device__ __forceinline__ void tile(const i32x4 *a, const i32x4 *b, const GemmArgs &g, int L) {
        dmac_tile<TM, TN, WM, WN, D, UNR, PACK, ATOMIC>(a, b, g, L);
    }
};

// The scan and the packing contraction of a batch are independent (the tail needs both), so
// one launch can run them side by side: workgroups [0, scan_blocks) take scan tiles, the rest
// packing tiles.  Both shapes have the same workgroup size; the dynamic LDS is the larger
// of the two.  scan_blocks is a multiple of 8, so the XCD-aware tile map of each part sees
// the same XCD for a workgroup as a launch of its own would.
template <class Scan, class Pack>
__global__ __launch_bounds__(Scan::kThreads, 1) void scan_pack_kernel(const i32x4 *__restrict__ qa,
                                                                      const i32x4 *__restrict__ db, GemmArgs gs,
                                                                      const i32x4 *__restrict__ ya,
                                                                      const i32x4 *__restrict__ u, GemmArgs gp,
                                                                      int scan_blocks) {
    static_assert(Scan::kThreads == Pack::kThreads, "one workgroup size");
    if ((int)blockIdx.x < scan_blocks) Scan::tile(qa, db, gs, blockIdx.x);
    else Pack::tile(ya, u, gp, blockIdx.x - scan_blocks);
}

// ---------------------------------------------------------------------------- tail --

struct TailArgs {
    FrameSet fs;
    const int *status;
    const uint32_t *acc;    // scan accumulators, B x C
    const uint32_t *qsum;   // per query sum(q) mod 2^32
    const uint32_t *beta;   // packing result mod Q, B x C
    const uint32_t *that;   // O x KP, reduced mod Q
    const uint32_t *wtab;   // n twiddles (low 32 bits of the inverse table)
    const uint8_t *prefix;  // O x part_words * 4 (each prefix zero-padded to whole words)
    uint8_t *out;
    size_t out_stride;
    const int *oidx;        // output slot of each batch position (null: position itself)
    int n, logn, tg, C, kp;
    int part_len, part_words;
    long long row1_off;
    int q1_bits;
    uint32_t q1;
    const int *dims;        // {B, MT} in device memory (launch sized for the batch's bucket), or null
};

// floor(x / Q) for 0 <= x < 2^53, exactly: x = T0 2^32 + L = T0 Q + (T0 12287 + L), the
// remainder form r = T1 2^32 + L1 (T1 <= 8) = T1 Q + (T1 12287 + L1), and the last term is
// below 2Q.
__device__ __forceinline__ uint64_t div_q_floor(uint64_t x) {
    const uint64_t t0 = x >> 32;
    const uint64_t r = t0 * kFold + (x & 0xffffffffull);
    const uint32_t t1 = (uint32_t)(r >> 32);
    uint32_t s;
    const bool carry = __builtin_add_overflow((uint32_t)r, __umul24(t1, (uint32_t)kFold), &s);
    return t0 + t1 + ((carry || s >= kQ) ? 1 : 0);
}

// Row-1 rescale Q -> q1 (src/amx_kernels.cpp:522-529): centre, round half away from zero,
// wrap into [0, q1).  Exact integer form of the reference's double-precision division: with
// the centred value c and h = floor(Q / 2), trunc((c q1 + sign(c) h) / Q) = sign(c) floor((|c|
// q1 + h) / Q), and |c| q1 + h < 2^53 for q1 < 2^21.
__device__ __forceinline__ uint32_t rescale_q1(uint32_t v, uint32_t q1) {
    const uint32_t half = (uint32_t)(kQ / 2);
    const bool neg = v >= half;
    const uint32_t a = neg ? (uint32_t)kQ - v : v;
    const uint32_t u = (uint32_t)div_q_floor((uint64_t)a * q1 + half);
    uint32_t t = neg ? (u ? q1 - u : 0) : u;
    if (t >= q1) t -= q1;
    return t;
}

// Gentleman-Sande butterfly with the halving: x <- (x + y) / 2, y <- (x - y) w (mod Q), in 32-bit
// arithmetic (x + y < 2Q as a carry and 32 bits; (s + Q) / 2 = s / 2 + (Q + 1) / 2 for odd s).
__device__ __forceinline__ void gs_butterfly(uint32_t &x, uint32_t &y, uint32_t w) {
    uint32_t s;
    const bool carry = __builtin_add_overflow(x, y, &s);
    s = red_carry(s, carry);
    s = (s >> 1) + ((s & 1) ? (uint32_t)(kQ / 2 + 1) : 0u);
    const uint32_t d = x >= y ? x - y : x - y - (uint32_t)kFold;  // x - y + Q mod 2^32
    x = s;
    y = mulmod(d, w);
}

__device__ __forceinline__ uint4 ld4(const uint32_t *p) { return *(const uint4 *)p; }

// One workgroup per (output o, query c): pointwise key product, Gentleman-Sande inverse
// NTT (src/amx_kernels.cpp:453-513), row-1 switch and LSB-first bit packing ORed onto the
// precomputed prefix (src/amx_kernels.cpp:515-537, 604-607).
//
// The transform runs in radix-8 phases: in a phase a thread holds a group of 8 elements
// base + 2^ls * m (m < 8) in registers and applies the stages t = 2^ls, 2^(ls+1), 2^(ls+2)
// that stay inside the group (phase 0 only the first 1-3 stages); groups are exchanged
// through LDS between phases, so n = 2048 needs three barriers instead of eleven.  Phase 0
// takes its group straight from the pointwise product and the last phase feeds the switch
// directly.  The twiddles are copied into LDS at the start, next to the part's prefix, so
// that the later phases wait on LDS rather than on L2; the last phase loads its group's scan
// accumulators before its butterflies, so that their latency hides behind them.
// LDS: a[n] | twiddles[n] | part[part_words] (tail_lds_bytes).
__host__ __device__ constexpr size_t tail_lds_bytes(size_t n, size_t part_words) { return (2 * n + part_words) * 4; }

constexpr int kTailThreads = 256;  // threads of one tail task

// The tail of (output o, query c) on kTailThreads threads (tid < kTailThreads) of a
// workgroup, with its LDS at lds (tail_lds_bytes).  The barriers are the workgroup's: every
// thread of the workgroup calls this the same number of times; a group of threads whose task
// is inactive (a row past the batch, a rejected request, no task) only takes part in them.
// KB: keys whose loads are issued together in the pointwise step.
template <int KB = 4>
__device__ __forceinline__ void tail_task(const TailArgs &p, int o, int c, bool active, uint32_t *lds, int tid) {
    constexpr int NTH = kTailThreads;
    const int n = p.n, logn = p.logn, groups = n >> 3;
    uint32_t *a = lds;
    uint32_t *tw = lds + n;
    uint32_t *part = lds + 2 * n;
    const size_t col0 = (size_t)o * n;
    const uint32_t *pre = (const uint32_t *)(p.prefix + (size_t)o * p.part_words * 4);
    const uint32_t *zk = active ? (const uint32_t *)(p.fs.slot(c) + 16) + p.kp : nullptr;
    const uint32_t *th = p.that + (size_t)o * p.kp;
    const uint32_t *bc = p.beta + (size_t)c * p.C + col0;
    const uint32_t *accc = p.acc + (size_t)c * p.C + col0;
    const uint32_t qs = active ? p.qsum[c] << 7 : 0;
    if (active) {
        for (int w = tid; w < p.part_words; w += NTH) part[w] = pre[w];
        // n is a power of two of at least 16
        for (int w = tid; w < (n >> 2); w += NTH) ((uint4 *)tw)[w] = ld4(p.wtab + 4 * w);
    }
    // phase 0 takes the r0 = logn - 3 (nph - 1) first stages (1..3) on contiguous groups, so
    // that every later phase has exactly three stages and its groups (stride 2^ls, 8
    // elements) stay inside the n points
    const int nph = (logn + 2) / 3, r0 = logn - 3 * (nph - 1);
    for (int ph = 0; ph < nph; ph++) {
        const int ls = ph == 0 ? 0 : r0 + 3 * (ph - 1), sp = 1 << ls, nst = ph == 0 ? r0 : 3;
        for (int g = tid; active && g < groups; g += NTH) {
            const int base = ((g >> ls) << (ls + 3)) | (g & (sp - 1));
            uint32_t v[8];
            if (ph == 0) {
                // pointwise: a = beta + sum_k z * that (mod Q), elements 8g .. 8g + 7; the
                // products' high and low words summed apart (each sum < 2^32 (tg + 1)) and
                // reduced once: sum hi 2^32 + sum lo == sum hi 12287 + sum lo, below 2^59 for
                // tg * n <= 65536
                uint64_t s8[8], h8[8];
                const uint4 b0 = ld4(bc + base), b1 = ld4(bc + base + 4);
                s8[0] = b0.x; s8[1] = b0.y; s8[2] = b0.z; s8[3] = b0.w;
                s8[4] = b1.x; s8[5] = b1.y; s8[6] = b1.z; s8[7] = b1.w;
#pragma unroll
                for (int m = 0; m < 8; m++) h8[m] = 0;
                auto mac = [&](int m, uint32_t z, uint32_t t) {
                    const uint64_t pr = (uint64_t)z * t;
                    s8[m] += (uint32_t)pr;
                    h8[m] += pr >> 32;
                };
                // keys KB at a time: all their loads are issued before the first product
                for (int k0 = 0; k0 < p.tg; k0 += KB) {
                    uint4 z0[KB], z1[KB], t0[KB], t1[KB];
#pragma unroll
                    for (int j = 0; j < KB; j++) {
                        const int k = min(k0 + j, p.tg - 1);
                        z0[j] = ld4(zk + k * n + base);
                        z1[j] = ld4(zk + k * n + base + 4);
                        t0[j] = ld4(th + k * n + base);
                        t1[j] = ld4(th + k * n + base + 4);
                    }
#pragma unroll
                    for (int j = 0; j < KB; j++) {
                        if (k0 + j >= p.tg) break;
                        mac(0, z0[j].x, t0[j].x); mac(1, z0[j].y, t0[j].y);
                        mac(2, z0[j].z, t0[j].z); mac(3, z0[j].w, t0[j].w);
                        mac(4, z1[j].x, t1[j].x); mac(5, z1[j].y, t1[j].y);
                        mac(6, z1[j].z, t1[j].z); mac(7, z1[j].w, t1[j].w);
                    }
                }
#pragma unroll
                for (int m = 0; m < 8; m++) v[m] = red64(h8[m] * kFold + s8[m]);
            } else {
#pragma unroll
                for (int m = 0; m < 8; m++) v[m] = a[base + (m << ls)];
            }
            uint32_t accv[8];
            if (ph + 1 == nph) {
#pragma unroll
                for (int m = 0; m < 8; m++) accv[m] = accc[base + (m << ls)] + qs;
            }
#pragma unroll
            for (int q = 0; q < 3; q++) {
                if (q >= nst) break;
                const int lt = ls + q;
                const int h = n >> (lt + 1);
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    if (m & (1 << q)) continue;
                    const int xi = base + (m << ls);
                    // LDS copies are visible from the first barrier on
                    const int wi = h + (xi >> (lt + 1));
                    gs_butterfly(v[m], v[m | (1 << q)], ph == 0 ? p.wtab[wi] : tw[wi]);
                }
            }
            if (ph + 1 < nph) {
#pragma unroll
                for (int m = 0; m < 8; m++) a[base + (m << ls)] = v[m];
            } else {
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    const int z = base + (m << ls);
                    const uint32_t inter = (uint32_t)(((uint64_t)accv[m] * kQ + (1ull << 31)) >> 32);
                    const uint32_t r = rescale_q1(addmod(inter, v[m]), p.q1);
                    const unsigned long long bit = (unsigned long long)p.row1_off + (unsigned long long)z * p.q1_bits;
                    const unsigned long long byte0 = bit >> 3;
                    // the reference ORs (r << bit%8) into the four bytes from byte0 on, inside the part
                    const uint64_t vv = ((uint64_t)r << (bit & 7)) << ((byte0 & 3) * 8);  // < 2^52
                    const unsigned long long wi = byte0 >> 2;
                    if (wi < (unsigned long long)p.part_words) atomicOr(&part[wi], (uint32_t)vv);
                    if (wi + 1 < (unsigned long long)p.part_words && (vv >> 32)) atomicOr(&part[wi + 1], (uint32_t)(vv >> 32));
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    uint8_t *dst = p.out + (size_t)(p.oidx ? p.oidx[c] : c) * p.out_stride + (size_t)o * p.part_len;
    if ((((uintptr_t)dst) & 3) == 0 && (p.part_len & 3) == 0) {
        uint32_t *d32 = (uint32_t *)dst;
        for (int w = tid; w < p.part_words; w += NTH) d32[w] = part[w];
    } else {
        for (int i = tid; i < p.part_len; i += NTH) dst[i] = (uint8_t)(part[i >> 2] >> (8 * (i & 3)));
    }
}

// Whether the tail of query c runs: inside the batch and accepted by validation.
__device__ __forceinline__ bool tail_wanted(const TailArgs &p, int c) {
    return !(p.dims && c >= p.dims[0]) && p.status[c] == 0;
}

__global__ __launch_bounds__(kTailThreads) void tail_kernel(TailArgs p) {
    extern __shared__ uint32_t lds[];
    const int o = blockIdx.x, c = blockIdx.y;
    if (!tail_wanted(p, c)) return;  // the whole workgroup: no barrier is left waiting
    tail_task(p, o, c, true, lds, threadIdx.x);
}

// ------------------------------------------------------- scan, packing and tail in one launch --

// A batch as one launch: the workgroups of both contractions (as scan_pack_kernel) and,
// behind them, the tails.  Output part o (columns [o n, (o + 1) n)) is complete when every
// workgroup whose column range meets it has stored its results: each workgroup, once its
// tile is stored, releases its stores and adds one arrival to each part its columns meet.
// The workgroup that brings a part's arrivals to their total (known in closed form from the
// two grids) publishes the part's B tail tasks, one per query.  Tail tasks are then claimed
// with an atomic counter per part, kTailThreads threads per task (two tasks at a time in a
// workgroup of 512 threads), by
//   * the publisher, which claims from its own part first and only leaves once no published
//     part has a task left (so every published task is run even if nobody else helps), and
//   * every other workgroup once its tile is done, which takes published tasks before it
//     leaves and, once every workgroup of the grid has started, waits for parts still to be
//     published (polling relaxed flags with s_sleep, at most `spin` polls) until every
//     part's tasks are claimed; only when a part has more tasks than one claim takes.
// No workgroup waits while any workgroup of the grid has yet to start, so the waits never
// hold a CU that an unstarted tile needs, and the tiles themselves never wait: every wait
// ends, and the bounded poll count is only a fail-safe (a helper that gives up changes no
// result, the publishers drain their parts).  Part completion is known per part over all
// row blocks.
// sync (3 parts + 4 words, zero before the launch): [0, O) arrivals, [O, 2O) claimed tasks,
// [2O, 3O) published flags, [3O] workgroups started, [3O + 1] parts whose tasks are all
// claimed, [3O + 2] set when a helper gave up waiting.
struct StepArgs {
    uint32_t *sync;
    int parts, part_cols;       // O, n
    int scan_cols, pack_cols;   // columns of one x tile of the scan / packing grid
    int B;                      // tail tasks per part (batch size; dims[0] when given)
    int spin;                   // polls a waiting helper makes at most
};

// LDS of one tail task inside the step kernel, in words, rounded to 16 bytes; the kernel
// needs step_lds_bytes(H, ...) of dynamic LDS at least (and the tiles' own).
__host__ __device__ constexpr int tail_step_words(int n, int part_words) { return (2 * n + part_words + 3) & ~3; }
__host__ __device__ constexpr size_t step_lds_bytes(int H, int n, int part_words) {
    return ((size_t)H * tail_step_words(n, part_words) + 4) * 4;
}

__device__ __forceinline__ int tiles_meeting_part(int p, int n, int cw, int gx) {
    const int x0 = (p * n) / cw, x1 = min((p * n + n - 1) / cw, gx - 1);
    return x1 >= x0 ? x1 - x0 + 1 : 0;
}

template <class Scan, class Pack>
__global__ __launch_bounds__(Scan::kThreads, 1) void step_kernel(const i32x4 *__restrict__ qa,
                                                                 const i32x4 *__restrict__ db, GemmArgs gs,
                                                                 const i32x4 *__restrict__ ya,
                                                                 const i32x4 *__restrict__ u, GemmArgs gp,
                                                                 int scan_blocks, TailArgs ta, StepArgs sa) {
    static_assert(Scan::kThreads == Pack::kThreads, "one workgroup size");
    static_assert(Scan::kThreads % kTailThreads == 0, "whole tail tasks per workgroup");
    constexpr int H = Scan::kThreads / kTailThreads;  // tail tasks a workgroup runs at once
    // dynamic LDS: the tiles' rings while they run, then H tail regions of tail_step_words
    // words each and the claimed group {part, first query} (-1: none) behind them
    extern __shared__ uint32_t step_lds[];
    const int tw_words = tail_step_words(ta.n, ta.part_words);
    int *task = (int *)(step_lds + H * tw_words);
    const int O = sa.parts, n = sa.part_cols;
    uint32_t *arrive = sa.sync, *claimed = sa.sync + O, *ready = sa.sync + 2 * O;
    uint32_t *started = sa.sync + 3 * O, *parts_done = started + 1, *gave_up = started + 2;
    const int L = blockIdx.x;
    const bool scan = L < scan_blocks;
    if (threadIdx.x == 0) atomicAdd(started, 1u);
    if (scan) Scan::tile(qa, db, gs, L);
    else Pack::tile(ya, u, gp, L - scan_blocks);
    const int B = ta.dims ? ta.dims[0] : sa.B;
    // every wave's stores are done (the barrier), then one release at agent scope before
    // the arrivals
    __syncthreads();
    int own = -1;
    if (threadIdx.x == 0) {
        const GemmArgs &g = scan ? gs : gp;
        const TileMap tm = map_tile_at(scan ? L : L - scan_blocks, g.gx, g.gy, g.gz);
        if (tm.valid) {
            const int cw = scan ? sa.scan_cols : sa.pack_cols;
            const int c0 = tm.x * cw, c1 = min((tm.x + 1) * cw, gs.C) - 1;
            __threadfence();
            for (int p = c0 / n; p <= c1 / n; p++) {
                const uint32_t total = (uint32_t)(gs.gy * gs.gz * tiles_meeting_part(p, n, sa.scan_cols, gs.gx) +
                                                  gp.gy * gp.gz * tiles_meeting_part(p, n, sa.pack_cols, gp.gx));
                if (atomicAdd(&arrive[p], 1u) + 1 == total) {
                    __threadfence();  // acquire the other workgroups' results, release them on
                    __hip_atomic_store(&ready[p], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    if (own < 0) own = p;
                }
            }
        }
    }
    // claim groups of H tasks from published parts until none is left (and, once the whole
    // grid has started, until every part is claimed)
    const int start = own >= 0 ? own : L % O;
    int scanned = 0, polls = 0;
    while (true) {
        if (threadIdx.x == 0) {
            int tp = -1, tc = -1;
            while (tp < 0) {
                for (; scanned < O; scanned++) {
                    const int p = (start + scanned) % O;
                    // relaxed: an acquiring load invalidates the L2 lines the other workgroups
                    // of this XCD stream through; the acquire follows a successful claim
                    if (!__hip_atomic_load(&ready[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
                    if (__hip_atomic_load(&claimed[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)B)
                        continue;
                    const uint32_t t = atomicAdd(&claimed[p], (uint32_t)H);
                    if (t >= (uint32_t)B) continue;
                    if (t + H >= (uint32_t)B) atomicAdd(parts_done, 1u);
                    tp = p;
                    tc = (int)t;
                    break;
                }
                if (tp >= 0) break;
                // nothing published is left: wait only when a part holds more than one claim of
                // tasks (else its publisher runs them all), once every workgroup has started,
                // and while some part still has tasks to publish or claim
                if (B <= H ||
                    __hip_atomic_load(parts_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)O ||
                    __hip_atomic_load(started, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x)
                    break;
                if (++polls > sa.spin) {
                    atomicOr(gave_up, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(32);
                scanned = 0;
            }
            // acquire (after the relaxed read of the published flag): the part's results
            // before its tails read them
            if (tp >= 0) __threadfence();
            task[0] = tp;
            task[1] = tc;
        }
        __syncthreads();
        const int tp = task[0], tc = task[1];
        if (tp < 0) break;
        const int h = threadIdx.x / kTailThreads, c = tc + h;
        const bool act = c < B && tail_wanted(ta, c);
        tail_task(ta, tp, act ? c : 0, act, step_lds + (size_t)h * tw_words, threadIdx.x % kTailThreads);
        __syncthreads();  // the task slot and the LDS are reused by the next group
    }
}

}  // namespace alto
