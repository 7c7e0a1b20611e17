// alto_offline.hpp -- the InspiRING offline precomputation on the device (gfx950).
//
// Included at the end of alto_engine.hip (it uses that file's error reporting).  Computes, for
// every output part o (columns [o n, (o + 1) n) of the hint), what the reference's
// full_packing_with_preprocessing_offline produces on host threads (src/packing.rs:910-942,
// 2260-2359):
//   a~_j   = NTT(negacyclic_perm(hint[., o n + j]))                     j < n
//   E[v]   = NTT(row_v),  row_v[j] = a~_j[v]                            v < n
//   r_i    = n^-1 E[table_{g^i}[s]][s],  rbar_i = n^-1 E[table_{2n-g^i}[s]][s]   i < n/2, g = 5
//   walking i = n/2 - 2 .. 0 (each chain sequential, both chains and every part in parallel):
//     bold_t[i]     = NTT(G^-1(INTT(r_{i+1}))),     r_i    += sum_k tau_{g^i}(w_k) . bold_t[i][k]
//     bold_t_bar[i] = NTT(G^-1(INTT(rbar_{i+1}))),  rbar_i += sum_k tau_{2n-g^i}(w_k) . bold_t_bar[i][k]
//   bold_t_hat = NTT(G^-1(INTT(rbar_0))),  a_hat = INTT(r_0 + sum_k v_k . bold_t_hat[k])
// with the engine's NTT (slot j = evaluation at psi^(2 bitrev(j) + 1)), G^-1 the base-2^ceil(32/tg)
// digits of each coefficient, least significant first.
//
// Kernels: a tiled transpose (hint columns and the a~ matrices into rows), a forward NTT per
// row in LDS (256 threads, radix-2 stages), the gather of the r's, the walk (one workgroup per
// part and chain: 2 parts x n/2 - 1 sequential steps of one inverse and tg forward transforms
// held in LDS), and the finish (a_hat).  Work: 2 n + (n - 1)(1 + tg) ... transforms of n points
// per part; at n = 2048, tg = 4, 16 parts the walk holds 32 CUs for about n/2 steps.

namespace offline {

// (x + y) / 2 mod Q
__device__ __forceinline__ uint32_t half_addmod(uint32_t x, uint32_t y) {
    uint64_t s = (uint64_t)x + y;
    if (s >= kQ) s -= kQ;
    return (uint32_t)((s & 1) ? (s + kQ) >> 1 : s >> 1);
}
__device__ __forceinline__ uint32_t submod(uint32_t x, uint32_t y) { return red_once((uint64_t)x + kQ - y); }

__device__ __forceinline__ uint32_t brev(uint32_t x, int logn) { return __brev(x) >> (32 - logn); }

// tau_t in the NTT domain: tau_t(a)[s] = a[auto_slot(t, s)]  (src/packing.rs:1082-1126)
__device__ __forceinline__ uint32_t auto_slot(uint32_t t, uint32_t s, int logn) {
    const uint32_t n2 = 2u << logn;
    const uint32_t e = (uint32_t)(((uint64_t)t * (2 * brev(s, logn) + 1)) % n2);
    return brev((e - 1) / 2, logn);
}

// In-place forward transforms of `cnt` consecutive rows of n values in LDS (Cooley-Tukey, the
// powers psi^bitrev(m + i) in fwd), by nth threads (tid < nth); ends with a barrier.
__device__ void ntt_lds(uint32_t *a, int cnt, int n, const uint32_t *__restrict__ fwd, int tid, int nth) {
    const int hn = n >> 1;
    for (int m = 1, t = hn; m < n; m <<= 1, t >>= 1) {
        for (int b = tid; b < cnt * hn; b += nth) {
            uint32_t *r = a + (b / hn) * n;
            const int bb = b % hn, i = bb / t, j = 2 * i * t + (bb % t);
            const uint32_t u = r[j], v = mulmod(r[j + t], fwd[m + i]);
            r[j] = addmod(u, v);
            r[j + t] = submod(u, v);
        }
        __syncthreads();
    }
}

// In-place inverse transform of one row (Gentleman-Sande with psi^-bitrev(k) / 2, the engine's
// table); ends with a barrier.
__device__ void intt_lds(uint32_t *a, int n, const uint32_t *__restrict__ inv, int tid, int nth) {
    const int hn = n >> 1;
    for (int t = 1; t < n; t <<= 1) {
        const int h = n / (2 * t);
        for (int b = tid; b < hn; b += nth) {
            const int i = b / t, j = 2 * i * t + (b % t);
            const uint32_t x = a[j], y = a[j + t];
            a[j] = half_addmod(x, y);
            a[j + t] = mulmod(submod(x, y), inv[h + i]);
        }
        __syncthreads();
    }
}

// hint (n x hcols u64, row z) -> ht (ncols x n u32, row col): 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void hint_transpose_kernel(const uint64_t *__restrict__ hint, size_t hcols,
                                                             uint32_t *__restrict__ ht, int n, int ncols) {
    __shared__ uint32_t tile[32][33];
    const int c0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int z = z0 + r, c = c0 + tx;
        tile[r][tx] = (z < n && c < ncols) ? (uint32_t)(hint[(size_t)z * hcols + c] % kQ) : 0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int c = c0 + r, z = z0 + tx;
        if (c < ncols && z < n) ht[(size_t)c * n + z] = tile[tx][r];
    }
}

// Per part (blockIdx.z): rows x cols u32 -> cols x rows
__global__ __launch_bounds__(256) void transpose_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                        int n) {
    __shared__ uint32_t tile[32][33];
    const size_t part = (size_t)blockIdx.z * n * n;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) tile[r][tx] = in[part + (size_t)(r0 + r) * n + c0 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) out[part + (size_t)(c0 + r) * n + r0 + tx] = tile[tx][r];
}

// One forward transform per row (blockIdx.x), in place; SIGMA applies negacyclic_perm on
// load (x_0, -x_{n-1}, ..., -x_1).  LDS n words.
template <bool SIGMA>
__global__ __launch_bounds__(256) void ntt_rows_kernel(uint32_t *__restrict__ rows, int n,
                                                       const uint32_t *__restrict__ fwd) {
    extern __shared__ uint32_t a[];
    uint32_t *row = rows + (size_t)blockIdx.x * n;
    for (int z = threadIdx.x; z < n; z += blockDim.x) {
        if (SIGMA) a[z] = z == 0 ? row[0] : (row[n - z] ? kQ32 - row[n - z] : 0u);
        else a[z] = row[z];
    }
    __syncthreads();
    ntt_lds(a, 1, n, fwd, threadIdx.x, blockDim.x);
    for (int z = threadIdx.x; z < n; z += blockDim.x) row[z] = a[z];
}

// r[o][i][s] = n^-1 E[o][table_{g_i}[s]][s], rbar with 2n - g_i; grid (n/2, parts)
__global__ __launch_bounds__(256) void gather_kernel(const uint32_t *__restrict__ E, uint32_t *__restrict__ r,
                                                     uint32_t *__restrict__ rb, const uint32_t *__restrict__ gp,
                                                     int n, int logn, uint32_t ninv) {
    const int i = blockIdx.x, o = blockIdx.y;
    const uint32_t g = gp[i], gbar = 2 * n - g;
    const uint32_t *e = E + (size_t)o * n * n;
    uint32_t *ro = r + ((size_t)o * (n / 2) + i) * n, *rbo = rb + ((size_t)o * (n / 2) + i) * n;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        ro[s] = mulmod(ninv, e[(size_t)auto_slot(g, s, logn) * n + s]);
        rbo[s] = mulmod(ninv, e[(size_t)auto_slot(gbar, s, logn) * n + s]);
    }
}

struct WalkArgs {
    const uint32_t *r, *rb;          // parts x n/2 x n (initial aggregates)
    const uint32_t *w;               // tg x n mask slots
    const uint32_t *fwd, *inv;       // psi^bitrev(k), psi^-bitrev(k) / 2
    const uint32_t *gp;              // 5^i mod 2n
    uint64_t *bold_t, *bold_t_bar;   // parts x (n/2 - 1) tg n
    uint64_t *bold_t_hat;            // parts x tg n
    uint32_t *r0;                    // parts x n: the final r_0
    int n, logn, tg, bits;
};

// Digits (base 2^bits, least significant first) of the coefficients in c into tg rows of d.
__device__ __forceinline__ void gadget_digits(uint32_t *d, const uint32_t *c, int n, int tg, int bits, int tid,
                                              int nth) {
    const uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
    for (int idx = tid; idx < tg * n; idx += nth) {
        const int k = idx / n, z = idx % n;
        const int sh = k * bits;
        d[idx] = sh < 32 ? (c[z] >> sh) & mask : 0u;
    }
    __syncthreads();
}

// One workgroup per (part, chain): blockIdx.x = 2 o + bar.  LDS: cur[n] | dig[tg n].
__global__ __launch_bounds__(512) void walk_kernel(WalkArgs a) {
    extern __shared__ uint32_t lds[];
    const int o = blockIdx.x >> 1, bar = blockIdx.x & 1;
    const int n = a.n, tg = a.tg, half = n / 2, tid = threadIdx.x, nth = blockDim.x;
    uint32_t *cur = lds, *dig = lds + n;
    const uint32_t *rbase = (bar ? a.rb : a.r) + (size_t)o * half * n;
    uint64_t *bt = (bar ? a.bold_t_bar : a.bold_t) + (size_t)o * (half - 1) * tg * n;
    for (int s = tid; s < n; s += nth) cur[s] = rbase[(size_t)(half - 1) * n + s];
    __syncthreads();
    for (int i = half - 2; i >= 0; i--) {
        intt_lds(cur, n, a.inv, tid, nth);
        gadget_digits(dig, cur, n, tg, a.bits, tid, nth);
        ntt_lds(dig, tg, n, a.fwd, tid, nth);
        const uint32_t t = bar ? 2 * n - a.gp[i] : a.gp[i];
        uint64_t *bti = bt + (size_t)i * tg * n;
        for (int s = tid; s < n; s += nth) {
            const uint32_t ts = auto_slot(t, s, a.logn);
            uint64_t acc = rbase[(size_t)i * n + s];
            for (int k = 0; k < tg; k++) {
                const uint32_t d = dig[k * n + s];
                bti[(size_t)k * n + s] = d;
                acc += mulmod(a.w[(size_t)k * n + ts], d);
            }
            cur[s] = red64(acc);
        }
        __syncthreads();
    }
    if (!bar) {
        for (int s = tid; s < n; s += nth) a.r0[(size_t)o * n + s] = cur[s];
        return;
    }
    intt_lds(cur, n, a.inv, tid, nth);
    gadget_digits(dig, cur, n, tg, a.bits, tid, nth);
    ntt_lds(dig, tg, n, a.fwd, tid, nth);
    for (int idx = tid; idx < tg * n; idx += nth) a.bold_t_hat[(size_t)o * tg * n + idx] = dig[idx];
}

// a_hat[o] = INTT(r0[o] + sum_k v_k . bold_t_hat[o][k]); one workgroup per part.
__global__ __launch_bounds__(256) void finish_kernel(const uint32_t *__restrict__ r0,
                                                     const uint64_t *__restrict__ bth,
                                                     const uint32_t *__restrict__ v, const uint32_t *__restrict__ inv,
                                                     uint64_t *__restrict__ a_hat, int n, int tg) {
    extern __shared__ uint32_t a[];
    const int o = blockIdx.x;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        uint64_t acc = r0[(size_t)o * n + s];
        for (int k = 0; k < tg; k++) acc += mulmod(v[(size_t)k * n + s], (uint32_t)bth[((size_t)o * tg + k) * n + s]);
        a[s] = red64(acc);
    }
    __syncthreads();
    intt_lds(a, n, inv, threadIdx.x, blockDim.x);
    for (int s = threadIdx.x; s < n; s += blockDim.x) a_hat[(size_t)o * n + s] = a[s];
}

}  // namespace offline

namespace offline {

bool shape_ok(const char *fn, size_t parts, size_t n, size_t tg, size_t hint_cols) {
    const size_t lds_walk = (1 + tg) * n * 4;
    if (!parts || !tg || tg > 32 || n < 32 || n > 4096 || (n & (n - 1)) || (kQ - 1) % (2 * n) ||
        parts * n > hint_cols || lds_walk > kMaxLds) {
        set_err("%s: unsupported shape (n a power of two in [32, 4096], parts n <= hint columns, (1 + tg) n words "
                "of LDS)", fn);
        return false;
    }
    return true;
}

// The precomputation of `parts` output parts from a hint in device memory (n x hint_cols u64),
// outputs in device memory (dbt, dbb: parts x (n/2 - 1) tg n; dbh: parts x tg n; dah: parts x
// n); enqueued on st and waited for.
bool offline_device(const uint64_t *dhint, size_t hint_cols, size_t parts, const uint64_t *w, const uint64_t *v,
                    size_t n, size_t tg, uint64_t *dbt, uint64_t *dbb, uint64_t *dbh, uint64_t *dah, hipStream_t st) {
    int logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    // the engine's NTT: psi the smallest primitive 2n-th root of unity
    uint64_t psi = 0;
    for (uint64_t g = 2; !psi && g < (1u << 16); g++) {
        const uint64_t c = host_powmod(g, (kQ - 1) / (2 * n));
        if (host_powmod(c, n) == kQ - 1) psi = c;
    }
    const uint64_t psi_inv = host_powmod(psi, kQ - 2), half_inv = (kQ + 1) / 2;
    std::vector<uint32_t> tab(2 * n + n / 2 + 2 * tg * n);
    uint32_t *fwd = tab.data(), *inv = fwd + n, *gp = inv + n, *wv = gp + n / 2;
    for (size_t k = 0; k < n; k++) {
        size_t b = 0;
        for (int i = 0; i < logn; i++) b |= ((k >> i) & 1) << (logn - 1 - i);
        fwd[k] = (uint32_t)host_powmod(psi, b);
        inv[k] = (uint32_t)host_mulmod(host_powmod(psi_inv, b), half_inv);
    }
    gp[0] = 1;
    for (size_t i = 1; i < n / 2; i++) gp[i] = (uint32_t)(gp[i - 1] * 5u % (2 * n));
    for (size_t i = 0; i < tg * n; i++) {
        wv[i] = (uint32_t)(w[i] % kQ);
        wv[tg * n + i] = (uint32_t)(v[i] % kQ);
    }
    const size_t O = parts, nn = n * n, half = n / 2, lds_walk = (1 + tg) * n * 4;
    uint32_t *dA = nullptr, *dB = nullptr, *dr = nullptr, *drb = nullptr, *dr0 = nullptr, *dtab = nullptr;
    bool ok = HIP_OK(hipMalloc(&dA, O * nn * 4)) && HIP_OK(hipMalloc(&dB, O * nn * 4)) &&
              HIP_OK(hipMalloc(&dr, O * half * n * 4)) && HIP_OK(hipMalloc(&drb, O * half * n * 4)) &&
              HIP_OK(hipMalloc(&dr0, O * n * 4)) && HIP_OK(hipMalloc(&dtab, tab.size() * 4)) &&
              HIP_OK(hipMemcpyAsync(dtab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
    uint32_t *dfwd = dtab, *dinv = dtab + n, *dgp = dtab + 2 * n, *dw = dgp + n / 2, *dv = dw + tg * n;
    if (ok) {
        const unsigned nt = (unsigned)(n / 32);
        // hint columns as rows (A: O n rows of n), a~ = NTT(sigma(row)) in place
        hipLaunchKernelGGL(hint_transpose_kernel, dim3((unsigned)(O * n / 32), nt), dim3(256), 0, st, dhint,
                           hint_cols, dA, (int)n, (int)(O * n));
        hipLaunchKernelGGL(ntt_rows_kernel<true>, dim3((unsigned)(O * n)), dim3(256), n * 4, st, dA, (int)n,
                           (const uint32_t *)dfwd);
        // rows v of each part's a~ matrix (B), E = NTT(row_v) in place
        hipLaunchKernelGGL(transpose_kernel, dim3(nt, nt, (unsigned)O), dim3(256), 0, st, (const uint32_t *)dA, dB,
                           (int)n);
        hipLaunchKernelGGL(ntt_rows_kernel<false>, dim3((unsigned)(O * n)), dim3(256), n * 4, st, dB, (int)n,
                           (const uint32_t *)dfwd);
        hipLaunchKernelGGL(gather_kernel, dim3((unsigned)half, (unsigned)O), dim3(256), 0, st, (const uint32_t *)dB,
                           dr, drb, (const uint32_t *)dgp, (int)n, logn, (uint32_t)host_powmod(n, kQ - 2));
        WalkArgs wa{dr, drb, dw, dfwd, dinv, dgp, dbt, dbb, dbh, dr0, (int)n, logn, (int)tg, (int)((32 + tg - 1) / tg)};
        ok = HIP_OK(hipFuncSetAttribute((const void *)walk_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds_walk));
        if (ok) {
            hipLaunchKernelGGL(walk_kernel, dim3((unsigned)(2 * O)), dim3(512), lds_walk, st, wa);
            hipLaunchKernelGGL(finish_kernel, dim3((unsigned)O), dim3(256), n * 4, st, (const uint32_t *)dr0,
                               (const uint64_t *)dbh, (const uint32_t *)dv, (const uint32_t *)dinv, dah, (int)n,
                               (int)tg);
            ok = HIP_OK(hipGetLastError()) && HIP_OK(hipStreamSynchronize(st));
        }
    }
    for (void *p : {(void *)dA, (void *)dB, (void *)dr, (void *)drb, (void *)dr0, (void *)dtab})
        if (p) (void)hipFree(p);
    return ok;
}

// Every automorphism's slot table, table (t - 1) / 2 for odd t < 2n (n x n u32)
__global__ __launch_bounds__(256) void tables_kernel(uint32_t *__restrict__ tab, int n, int logn) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    const uint32_t ti = (uint32_t)(i / n), s = (uint32_t)(i % n);
    tab[i] = auto_slot(2 * ti + 1, s, logn);
}

}  // namespace offline

extern "C" int alto_inspiring_offline(uint64_t *bold_t, uint64_t *bold_t_bar, uint64_t *bold_t_hat, uint64_t *a_hat,
                                      const uint64_t *hint, size_t hint_cols, size_t parts, const uint64_t *w,
                                      const uint64_t *v, size_t n, size_t tg) {
    using namespace offline;
    if (!bold_t || !bold_t_bar || !bold_t_hat || !a_hat || !hint || !w || !v) {
        set_err("alto_inspiring_offline: null argument");
        return -1;
    }
    if (!shape_ok("alto_inspiring_offline", parts, n, tg, hint_cols)) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return -1;
    }
    const size_t O = parts, per = (n / 2 - 1) * tg * n;
    uint64_t *dhint = nullptr, *dbt = nullptr, *dbb = nullptr, *dbh = nullptr, *dah = nullptr;
    hipStream_t st = nullptr;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
              HIP_OK(hipMalloc(&dhint, n * hint_cols * 8)) && HIP_OK(hipMalloc(&dbt, O * per * 8)) &&
              HIP_OK(hipMalloc(&dbb, O * per * 8)) && HIP_OK(hipMalloc(&dbh, O * tg * n * 8)) &&
              HIP_OK(hipMalloc(&dah, O * n * 8)) &&
              HIP_OK(hipMemcpyAsync(dhint, hint, n * hint_cols * 8, hipMemcpyHostToDevice, st)) &&
              offline_device(dhint, hint_cols, parts, w, v, n, tg, dbt, dbb, dbh, dah, st);
    ok = ok && HIP_OK(hipMemcpyAsync(bold_t, dbt, O * per * 8, hipMemcpyDeviceToHost, st)) &&
         HIP_OK(hipMemcpyAsync(bold_t_bar, dbb, O * per * 8, hipMemcpyDeviceToHost, st)) &&
         HIP_OK(hipMemcpyAsync(bold_t_hat, dbh, O * tg * n * 8, hipMemcpyDeviceToHost, st)) &&
         HIP_OK(hipMemcpyAsync(a_hat, dah, O * n * 8, hipMemcpyDeviceToHost, st)) && HIP_OK(hipStreamSynchronize(st));
    for (void *p : {(void *)dhint, (void *)dbt, (void *)dbb, (void *)dbh, (void *)dah})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

// The whole setup from a database: hint, precomputation and packing matrix on the device, only
// the database (when on the host), the query polynomials and masks in and U, the tail keys and
// a_hat out.
extern "C" int alto_scheme_setup(int8_t *ub, uint64_t *that, uint64_t *a_hat, const uint8_t *db_col, int on_device,
                                 size_t ld, size_t rows, size_t cols, const uint64_t *a, size_t npolys,
                                 const uint64_t *w, const uint64_t *v, size_t n, size_t tg) {
    using namespace offline;
    if (!ub || !that || !a_hat || !db_col || !a || !w || !v || ld < rows) {
        set_err("alto_scheme_setup: null argument or ld < rows");
        return -1;
    }
    if (!hint_shape_ok("alto_scheme_setup", rows, cols, n, npolys) || cols % n ||
        !shape_ok("alto_scheme_setup", cols / n, n, tg, cols))
        return -1;
    const size_t K = tg * n, O = cols / n, num_iter = n / 2 - 1, per = num_iter * tg * n;
    if (K % 64 || cols % 32 || K * 8 > kMaxLds) {
        set_err("alto_scheme_setup: unsupported packing shape (tg n %% 64 == 0, columns %% 32 == 0)");
        return -1;
    }
    int dev = 0;
    if (!HIP_OK(hipGetDevice(&dev))) return -1;
    int logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    int8_t *db = nullptr, *dub = nullptr;
    uint8_t *stage = nullptr;
    uint64_t *dhint = nullptr, *dbt = nullptr, *dbb = nullptr, *dbh = nullptr, *dah = nullptr;
    uint32_t *dtab = nullptr, *dpow = nullptr;
    hipStream_t st = nullptr;
    std::vector<uint32_t> gp(num_iter);
    for (size_t i = 0; i < num_iter; i++) gp[i] = (uint32_t)(i ? gp[i - 1] * 5u % (2 * n) : 1u);
    const size_t stage_cols = std::max<size_t>(16, std::min<size_t>(cols, ((size_t)256 << 20) / rows / 16 * 16));
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && HIP_OK(hipMalloc(&db, rows * cols)) &&
              HIP_OK(hipMalloc(&stage, stage_cols * rows));
    ok = ok && tile_db_columns(db_col, on_device != 0, ld, rows, cols, db, stage, stage_cols, st);
    if (stage) (void)hipFree(stage);
    ok = ok && HIP_OK(hipMalloc(&dhint, n * cols * 8)) && hint_device(db, rows, cols, n, a, npolys, dhint, st);
    if (db) (void)hipFree(db);
    ok = ok && HIP_OK(hipMalloc(&dbt, O * per * 8)) && HIP_OK(hipMalloc(&dbb, O * per * 8)) &&
         HIP_OK(hipMalloc(&dbh, O * K * 8)) && HIP_OK(hipMalloc(&dah, O * n * 8)) &&
         offline_device(dhint, cols, O, w, v, n, tg, dbt, dbb, dbh, dah, st);
    if (dhint) (void)hipFree(dhint);
    ok = ok && HIP_OK(hipMalloc(&dtab, n * n * 4)) && HIP_OK(hipMalloc(&dpow, num_iter * 4)) &&
         HIP_OK(hipMalloc(&dub, 4 * K * cols)) &&
         HIP_OK(hipMemcpyAsync(dpow, gp.data(), num_iter * 4, hipMemcpyHostToDevice, st));
    if (ok) {
        hipLaunchKernelGGL(tables_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, st, dtab, (int)n, logn);
        ok = HIP_OK(hipGetLastError()) &&
             pack_build_device(dbt, dbb, per, dtab, dpow, num_iter, tg, n, cols, kQ, dub, st) &&
             HIP_OK(hipMemcpyAsync(ub, dub, 4 * K * cols, hipMemcpyDeviceToHost, st)) &&
             HIP_OK(hipMemcpyAsync(that, dbh, O * K * 8, hipMemcpyDeviceToHost, st)) &&
             HIP_OK(hipMemcpyAsync(a_hat, dah, O * n * 8, hipMemcpyDeviceToHost, st)) &&
             HIP_OK(hipStreamSynchronize(st));
    }
    for (void *p : {(void *)dbt, (void *)dbb, (void *)dbh, (void *)dah, (void *)dtab, (void *)dpow, (void *)dub})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}

// ------------------------------------------------------------------------------------------
// The two-modulus batched dot product of the reference's SimplePIR server path
// (fast_batched_dot_product_avx512, src/kernel.rs:14-192, called from src/server.rs:593-618 and
// :1446-1547), off the engine path: for batch row r < K and column j,
//   lo = (sum_k (a[r][k] & 0xffffffff) * b[j][k] mod 2^64) mod m0
//   hi = (sum_k (a[r][k] >> 32)        * b[j][k] mod 2^64) mod m1
//   c[r][j] = (c[r][j] + crt(lo, hi)) mod m0 m1,   crt(lo, hi) = lo + m0 ((hi - lo) m0^-1 mod m1)
// (a CRT-packed: a = (A mod m0) | (A mod m1) << 32 gives c + sum_k A b mod m0 m1).  b is
// b_cols columns of b_rows elements of 1, 2 or 4 bytes (b_t[j * b_rows + k]).  One workgroup per
// column j, every batch row at once: 256 threads split k, wrapping u64 partial sums, a tree sum.
namespace offline {

template <typename T>
__global__ __launch_bounds__(256) void batched_dot_kernel(uint64_t *__restrict__ c, size_t c_stride,
                                                          const uint64_t *__restrict__ a, size_t a_stride,
                                                          const T *__restrict__ b, size_t b_rows, int K, uint64_t m0,
                                                          uint64_t m1, uint64_t m0_inv) {
    constexpr int kMaxK = 8;
    __shared__ uint64_t red[2 * kMaxK][256];
    const size_t j = blockIdx.x;
    const T *bj = b + j * b_rows;
    for (int r0 = 0; r0 < K; r0 += kMaxK) {
        const int nk = min(kMaxK, K - r0);
        uint64_t lo[kMaxK] = {}, hi[kMaxK] = {};
        for (size_t k = threadIdx.x; k < b_rows; k += blockDim.x) {
            const uint64_t bv = (uint64_t)bj[k];
            for (int r = 0; r < nk; r++) {
                const uint64_t av = a[(size_t)(r0 + r) * a_stride + k];
                lo[r] += (av & 0xffffffffull) * bv;
                hi[r] += (av >> 32) * bv;
            }
        }
        for (int r = 0; r < nk; r++) {
            red[2 * r][threadIdx.x] = lo[r];
            red[2 * r + 1][threadIdx.x] = hi[r];
        }
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
                for (int r = 0; r < 2 * nk; r++) red[r][threadIdx.x] += red[r][threadIdx.x + s];
            __syncthreads();
        }
        if ((int)threadIdx.x < nk) {
            const int r = threadIdx.x;
            const uint64_t l = red[2 * r][0] % m0, h = red[2 * r + 1][0] % m1;
            const uint64_t d = (h + m1 - l % m1) % m1;
            const uint64_t t = (uint64_t)(((unsigned __int128)d * m0_inv) % m1);
            const unsigned __int128 M = (unsigned __int128)m0 * m1;
            const unsigned __int128 res = (unsigned __int128)l + (unsigned __int128)m0 * t;  // < M
            uint64_t *cr = c + (size_t)(r0 + r) * c_stride + j;
            *cr = (uint64_t)(((unsigned __int128)*cr + res) % M);
        }
        __syncthreads();
    }
}

}  // namespace offline

extern "C" int alto_fast_batched_dot_product(uint64_t *c, size_t c_stride, const uint64_t *a, size_t a_stride,
                                             size_t a_elems, const void *b_t, int b_bytes, size_t b_rows,
                                             size_t b_cols, size_t K, uint64_t m0, uint64_t m1) {
    using namespace offline;
    const unsigned __int128 M = (unsigned __int128)m0 * m1;
    if (!c || !a || !b_t || !K || !b_cols || a_elems != b_rows || !b_rows || m0 < 2 || m1 < 2 ||
        M >> 64 || c_stride < b_cols || a_stride < a_elems || (b_bytes != 1 && b_bytes != 2 && b_bytes != 4) ||
        b_cols > 0x7fffffffu) {
        set_err("alto_fast_batched_dot_product: bad arguments (a_elems == b_rows, m0 m1 < 2^64, b of 1, 2 or 4 bytes)");
        return -1;
    }
    // m0^-1 mod m1 (extended Euclid); the moduli must be coprime
    long long t0 = 0, t1 = 1;
    unsigned long long r0 = m1, r1 = m0 % m1;
    while (r1) {
        const unsigned long long q = r0 / r1, r2 = r0 - q * r1;
        const long long t2 = t0 - (long long)q * t1;
        r0 = r1;
        r1 = r2;
        t0 = t1;
        t1 = t2;
    }
    if (r0 != 1) {
        set_err("alto_fast_batched_dot_product: the moduli are not coprime");
        return -1;
    }
    const uint64_t m0_inv = (uint64_t)((t0 % (long long)m1 + (long long)m1) % (long long)m1);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return -1;
    }
    uint64_t *dc = nullptr, *da = nullptr;
    void *db = nullptr;
    hipStream_t st = nullptr;
    const size_t cb = K * c_stride * 8, ab = K * a_stride * 8, bb = b_cols * b_rows * (size_t)b_bytes;
    bool ok = HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && HIP_OK(hipMalloc(&dc, cb)) &&
              HIP_OK(hipMalloc(&da, ab)) && HIP_OK(hipMalloc(&db, bb)) &&
              HIP_OK(hipMemcpyAsync(dc, c, cb, hipMemcpyHostToDevice, st)) &&
              HIP_OK(hipMemcpyAsync(da, a, ab, hipMemcpyHostToDevice, st)) &&
              HIP_OK(hipMemcpyAsync(db, b_t, bb, hipMemcpyHostToDevice, st));
    if (ok) {
        if (b_bytes == 1)
            hipLaunchKernelGGL(batched_dot_kernel<uint8_t>, dim3((unsigned)b_cols), dim3(256), 0, st, dc, c_stride,
                               (const uint64_t *)da, a_stride, (const uint8_t *)db, b_rows, (int)K, m0, m1, m0_inv);
        else if (b_bytes == 2)
            hipLaunchKernelGGL(batched_dot_kernel<uint16_t>, dim3((unsigned)b_cols), dim3(256), 0, st, dc, c_stride,
                               (const uint64_t *)da, a_stride, (const uint16_t *)db, b_rows, (int)K, m0, m1, m0_inv);
        else
            hipLaunchKernelGGL(batched_dot_kernel<uint32_t>, dim3((unsigned)b_cols), dim3(256), 0, st, dc, c_stride,
                               (const uint64_t *)da, a_stride, (const uint32_t *)db, b_rows, (int)K, m0, m1, m0_inv);
        ok = HIP_OK(hipGetLastError()) && HIP_OK(hipMemcpyAsync(c, dc, cb, hipMemcpyDeviceToHost, st)) &&
             HIP_OK(hipStreamSynchronize(st));
    }
    for (void *p : {(void *)dc, (void *)da, db})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return ok ? 0 : -1;
}
