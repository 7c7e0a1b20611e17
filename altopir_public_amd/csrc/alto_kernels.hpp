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

// ------------------------------------------------------------------ mod-Q helpers --

// x < 2^32 + 2^28 -> x mod Q
__device__ __forceinline__ uint32_t red_once(uint64_t x) { return (uint32_t)(x >= kQ ? x - kQ : x); }

// any u64 -> mod Q, using 2^32 == 12287 (mod Q) twice
__device__ __forceinline__ uint32_t red64(uint64_t p) {
    uint64_t t = (p >> 32) * kFold + (p & 0xffffffffull);
    t = (t >> 32) * kFold + (t & 0xffffffffull);
    return red_once(t);
}

__device__ __forceinline__ uint32_t mulmod(uint32_t a, uint32_t b) { return red64((uint64_t)a * b); }

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b) { return red_once((uint64_t)a + b); }

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

__device__ __forceinline__ uint32_t ld32_bytes(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// Request validation (src/amx_kernels.cpp:571-581): one workgroup per frame.  Frames with
// a nonzero status on entry (rejected by the host: null or wrong length) keep it.
// Also clears the per-query sums that the prep kernels accumulate.
__global__ __launch_bounds__(256) void validate_kernel(const uint8_t *__restrict__ slots, size_t stride,
                                                       int *__restrict__ status, uint32_t kp,
                                                       uint32_t *__restrict__ qsum,
                                                       unsigned long long *__restrict__ ysum) {
    __shared__ int st;
    const size_t c = blockIdx.x;
    if (threadIdx.x == 0) {
        qsum[c] = 0;
        ysum[c] = 0;
        st = status[c];
    }
    __syncthreads();
    if (st != 0) return;
    const uint8_t *f = slots + c * stride + 7;
    if (f[0] != 0 || ld32_bytes(f + 1) != kp || ld32_bytes(f + 5) != kp) {
        if (threadIdx.x == 0) status[c] = 2;
        return;
    }
    const uint4 *keys = (const uint4 *)(f + 9);
    int bad = 0;
    for (uint32_t i = threadIdx.x; i < kp / 2; i += blockDim.x) {
        const uint4 v = keys[i];
        bad |= (v.x >= kQ32) | (v.y >= kQ32) | (v.z >= kQ32) | (v.w >= kQ32);
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) status[c] = bad ? 3 : 0;
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

// Splits u32 words into four byte-limb rows of an MFMA A operand (rows 4c+i, K = nwords).
// SCAN: signed digits of q + 0x80808080 and sum(q) mod 2^32 into qsum.
// PACK: bytes of y offset by -128 and the exact sum(y) into ysum.
// grid = (ceil(KC * 4 / 256), Bp); clients c >= B or with status != 0 get zero rows.
template <bool SCAN>
__global__ __launch_bounds__(256) void limb_prep_kernel(const uint8_t *__restrict__ slots, size_t stride,
                                                        size_t word_off, const int *__restrict__ status, int B,
                                                        size_t KC, int8_t *__restrict__ dst,
                                                        uint32_t *__restrict__ qsum,
                                                        unsigned long long *__restrict__ ysum) {
    __shared__ unsigned long long red[8];
    const int c = blockIdx.y;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t kc = p >> 2;
    const int g = (int)(p & 3);
    const bool ok = c < B && status[c] == 0;
    unsigned long long s = 0;
    if (kc < KC) {
        uint4 o[4] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
        if (ok) {
            const uint4 *src = (const uint4 *)(slots + (size_t)c * stride + 16 + word_off * 4) + (kc * 64 + g * 16) / 4;
            uint32_t w[16];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint4 v = src[e];
                w[4 * e] = v.x; w[4 * e + 1] = v.y; w[4 * e + 2] = v.z; w[4 * e + 3] = v.w;
            }
#pragma unroll
            for (int e = 0; e < 16; e++) {
                s += w[e];
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
        uint4 *blk = (uint4 *)(dst + ((size_t)(c >> 2) * KC + kc) * kBlockBytes);
#pragma unroll
        for (int i = 0; i < 4; i++) blk[(c & 3) * 4 + i + 16 * g] = o[i];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0 && ok) {
        if (SCAN) atomicAdd(&qsum[c], (uint32_t)s);
        else atomicAdd(&ysum[c], s);
    }
}

// ------------------------------------------------------------------- contractions ----

// Maps a 1-D workgroup id onto (x, y, z) tiles so that the workgroups of one column tile
// x (all y, same z) run back to back on the same XCD (dispatch round-robins workgroups
// over the 8 XCDs): they then share the database slice in that XCD's L2.
struct TileMap {
    int x, y, z;
    bool valid;
};
__device__ __forceinline__ TileMap map_tile(int gx, int gy, int gz) {
    const int L = blockIdx.x, xcd = L & 7;
    int t = L >> 3;
    const int gxg = (gx + 7) >> 3;
    TileMap m;
    m.y = t % gy;
    t /= gy;
    m.x = (t % gxg) * 8 + xcd;
    m.z = t / gxg;
    m.valid = m.x < gx && m.z < gz;
    return m;
}

// Scan contraction.  A = query limbs (MT tiles of 16 rows = 4 queries), B = database
// (CT tiles of 16 columns), K = KC blocks of 64 records.  Workgroup = 4 waves; wave w
// computes rows of tiles [y*MB, y*MB+MB) x columns of tiles [(4x+w)*NB, +NB) over the
// K range of split z, and adds sum_i 256^i P_i (mod 2^32) into acc[c * C + col]
// (atomically when the K range is split).
template <int MB, int NB, bool ATOMIC>
__global__ __launch_bounds__(256) void scan_kernel(const i32x4 *__restrict__ qa, const i32x4 *__restrict__ db,
                                                   uint32_t *__restrict__ acc_out, int KC, int CT, int MT, int gx,
                                                   int gy, int gz, int C) {
    const TileMap tm = map_tile(gx, gy, gz);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ct0 = (tm.x * 4 + wave) * NB, mt0 = tm.y * MB;
    const int kb = (int)((long long)KC * tm.z / gz), ke = (int)((long long)KC * (tm.z + 1) / gz);
    const i32x4 *ap[MB];
    const i32x4 *bp[NB];
#pragma unroll
    for (int m = 0; m < MB; m++) ap[m] = qa + (size_t)min(mt0 + m, MT - 1) * KC * 64 + lane;
#pragma unroll
    for (int n = 0; n < NB; n++) bp[n] = db + (size_t)min(ct0 + n, CT - 1) * KC * 64 + lane;
    uint32_t run[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) run[m][n] = 0;
    for (int k = kb; k < ke;) {
        const int kend = min(ke, k + kFoldSteps);
        i32x4 acc[MB][NB];
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++) acc[m][n] = i32x4{0, 0, 0, 0};

        for (; k < kend; k++) {
            i32x4 a[MB], b[NB];
#pragma unroll
            for (int m = 0; m < MB; m++) a[m] = ap[m][(size_t)k * 64];
#pragma unroll
            for (int n = 0; n < NB; n++) b[n] = bp[n][(size_t)k * 64];
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int n = 0; n < NB; n++) acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[m], b[n], acc[m][n], 0, 0, 0);
        }
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++)
                run[m][n] += (uint32_t)acc[m][n].x + ((uint32_t)acc[m][n].y << 8) + ((uint32_t)acc[m][n].z << 16) +
                             ((uint32_t)acc[m][n].w << 24);
    }
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++) {
            if (mt0 + m >= MT || ct0 + n >= CT) continue;
            const int c = 4 * (mt0 + m) + (lane >> 4), col = 16 * (ct0 + n) + (lane & 15);
            if (ATOMIC) atomicAdd(&acc_out[(size_t)c * C + col], run[m][n]);
            else acc_out[(size_t)c * C + col] = run[m][n];
        }
}

// Packing contraction.  A = y limbs (MT tiles), B = the four byte planes of U (block
// (ct, kc, j)), K = KCp blocks.  Each lane ends with G_ij (i = limb row, j = plane) of one
// query and one column, recombines them mod Q and adds the corrections:
//   beta = sum_{ij} 2^{8(i+j)} G_ij + ucorr[col] + ycorr(c)   (mod Q)
// ycorr(c) = 128E * Ysum[c] - (128E)^2 * KP, the constant passed as ycorr0.
template <int MB, int NB>
__global__ __launch_bounds__(256) void pack_kernel(const i32x4 *__restrict__ ya, const i32x4 *__restrict__ ut,
                                                   const uint32_t *__restrict__ ucorr,
                                                   const unsigned long long *__restrict__ ysum,
                                                   uint32_t *__restrict__ beta, int KCp, int CT, int MT, int gx, int gy,
                                                   int C, uint32_t k128e, uint32_t ycorr0) {
    const TileMap tm = map_tile(gx, gy, 1);
    if (!tm.valid) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ct0 = (tm.x * 4 + wave) * NB, mt0 = tm.y * MB;
    const i32x4 *ap[MB];
    const i32x4 *bp[NB];
#pragma unroll
    for (int m = 0; m < MB; m++) ap[m] = ya + (size_t)min(mt0 + m, MT - 1) * KCp * 64 + lane;
#pragma unroll
    for (int n = 0; n < NB; n++) bp[n] = ut + (size_t)min(ct0 + n, CT - 1) * KCp * 256 + lane;
    i32x4 acc[MB][NB][4];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int n = 0; n < NB; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[m][n][j] = i32x4{0, 0, 0, 0};

    for (int k = 0; k < KCp; k++) {
        i32x4 a[MB], b[NB][4];
#pragma unroll
        for (int m = 0; m < MB; m++) a[m] = ap[m][(size_t)k * 64];
#pragma unroll
        for (int n = 0; n < NB; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[n][j] = bp[n][(size_t)k * 256 + j * 64];
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int n = 0; n < NB; n++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[m][n][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[m], b[n][j], acc[m][n][j], 0, 0, 0);
    }
    // 2^{8s} mod Q for s = 4, 5, 6 (s <= 3 are exact powers)
    const int64_t w4 = 12287, w5 = 12287ll << 8, w6 = 12287ll << 16;
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int c = 4 * (mt0 + m) + (lane >> 4);
        const uint32_t yc = m + mt0 < MT ? addmod(mulmod(red64(ysum[c]), k128e), ycorr0) : 0;
#pragma unroll
        for (int n = 0; n < NB; n++) {
            if (mt0 + m >= MT || ct0 + n >= CT) continue;
            const int col = 16 * (ct0 + n) + (lane & 15);
            int64_t V[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                V[j + 0] += acc[m][n][j].x;
                V[j + 1] += acc[m][n][j].y;
                V[j + 2] += acc[m][n][j].z;
                V[j + 3] += acc[m][n][j].w;
            }
            const int64_t t = V[0] + (V[1] << 8) + (V[2] << 16) + (V[3] << 24) + V[4] * w4 + V[5] * w5 + V[6] * w6;
            beta[(size_t)c * C + col] = addmod(addmod(red_i64(t), ucorr[col]), yc);
        }
    }
}

// ---------------------------------------------------------------------------- tail --

struct TailArgs {
    const uint8_t *slots;
    size_t stride;
    const int *status;
    const uint32_t *acc;    // scan accumulators, B x C
    const uint32_t *qsum;   // per query sum(q) mod 2^32
    const uint32_t *beta;   // packing result mod Q, B x C
    const uint32_t *that;   // O x KP, reduced mod Q
    const uint32_t *wtab;   // n twiddles (low 32 bits of the inverse table)
    const uint8_t *prefix;  // O x part_words * 4 (each prefix zero-padded to whole words)
    uint8_t *out;
    size_t out_stride;
    int n, logn, tg, C, kp;
    int part_len, part_words;
    long long row1_off;
    int q1_bits;
    uint32_t q1;
};

// Row-1 rescale Q -> q1 (src/amx_kernels.cpp:522-529): centre, round half away from zero,
// wrap into [0, q1).  Exact integer form of the reference's double-precision division.
__device__ __forceinline__ uint32_t rescale_q1(uint32_t v, uint32_t q1) {
    const int64_t half = (int64_t)(kQ / 2);
    const bool neg = (int64_t)v >= half;
    const int64_t c = neg ? (int64_t)v - (int64_t)kQ : (int64_t)v;
    const int64_t x = c * (int64_t)q1 + (neg ? -half : half);
    int64_t t = x / (int64_t)kQ;
    if (t < 0) t += q1;
    if (t >= (int64_t)q1) t -= q1;
    return (uint32_t)t;
}

// One workgroup per (output o, query c): pointwise key product, Gentleman-Sande inverse
// NTT in LDS (src/amx_kernels.cpp:453-513), row-1 switch and LSB-first bit packing ORed
// onto the precomputed prefix (src/amx_kernels.cpp:515-537, 604-607).
__global__ __launch_bounds__(256) void tail_kernel(TailArgs p) {
    extern __shared__ uint32_t lds[];
    const int o = blockIdx.x, c = blockIdx.y;
    if (p.status[c] != 0) return;
    uint32_t *a = lds;
    uint32_t *part = lds + p.n;
    const int n = p.n;
    const size_t col0 = (size_t)o * n;
    const uint32_t *zk = (const uint32_t *)(p.slots + (size_t)c * p.stride + 16) + p.kp;
    const uint32_t *th = p.that + (size_t)o * p.kp;
    const uint32_t *bc = p.beta + (size_t)c * p.C + col0;
    for (int z = threadIdx.x; z < n; z += blockDim.x) {
        uint64_t s = bc[z];
        for (int k = 0; k < p.tg; k++) s += mulmod(zk[k * n + z], th[k * n + z]);
        a[z] = red64(s);
    }
    const uint32_t *pre = (const uint32_t *)(p.prefix + (size_t)o * p.part_words * 4);
    for (int w = threadIdx.x; w < p.part_words; w += blockDim.x) part[w] = pre[w];
    __syncthreads();
    for (int lt = 0; lt < p.logn; lt++) {
        const int t = 1 << lt, h = n >> (lt + 1);
        for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
            const int i = b >> lt, j = b & (t - 1);
            const int xi = 2 * t * i + j;
            const uint32_t x = a[xi], y = a[xi + t];
            uint64_t s = (uint64_t)x + y;
            if (s >= kQ) s -= kQ;
            s = (s & 1) ? (s + kQ) >> 1 : s >> 1;
            const uint32_t d = red_once((uint64_t)x + kQ - y);
            a[xi] = (uint32_t)s;
            a[xi + t] = mulmod(d, p.wtab[h + i]);
        }
        __syncthreads();
    }
    const uint32_t *accc = p.acc + (size_t)c * p.C + col0;
    const uint32_t qs = p.qsum[c] << 7;
    for (int z = threadIdx.x; z < n; z += blockDim.x) {
        const uint32_t acc = accc[z] + qs;
        const uint32_t inter = (uint32_t)(((uint64_t)acc * kQ + (1ull << 31)) >> 32);
        const uint32_t r = rescale_q1(addmod(inter, a[z]), p.q1);
        const unsigned long long bit = (unsigned long long)p.row1_off + (unsigned long long)z * p.q1_bits;
        const unsigned long long byte0 = bit >> 3;
        // the reference ORs (r << bit%8) into the four bytes from byte0 on, inside the part
        const uint64_t v = (uint64_t)r << (bit & 7);
        const unsigned long long wi = byte0 >> 2;
        const int sh = (int)(byte0 & 3) * 8;
        const uint64_t vv = v << sh;  // r < 2^21, so vv < 2^52
        if (wi < (unsigned long long)p.part_words) atomicOr(&part[wi], (uint32_t)vv);
        if (wi + 1 < (unsigned long long)p.part_words && (vv >> 32)) atomicOr(&part[wi + 1], (uint32_t)(vv >> 32));
    }
    __syncthreads();
    uint8_t *dst = p.out + (size_t)c * p.out_stride + (size_t)o * p.part_len;
    if ((((uintptr_t)dst) & 3) == 0 && (p.part_len & 3) == 0) {
        uint32_t *d32 = (uint32_t *)dst;
        for (int w = threadIdx.x; w < p.part_words; w += blockDim.x) d32[w] = part[w];
    } else {
        for (int i = threadIdx.x; i < p.part_len; i += blockDim.x) dst[i] = (uint8_t)(part[i >> 2] >> (8 * (i & 3)));
    }
}

}  // namespace alto
