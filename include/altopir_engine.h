/*
 * altopir_engine.h -- C ABI of the MI355X answer engine.
 *
 * The first eleven entry points are the reference engine's ABI, with the same
 * signatures, argument meaning, status codes and error returns, so that the
 * reference's Rust wrapper (src/engine.rs:4-20) links against this library
 * unchanged.  Each declaration names the reference interface it replaces.
 *
 * Differences in behaviour that a caller can observe:
 *   - alto_engine_new copies (uploads) every setup buffer into HBM; the caller may
 *     release its host copies afterwards (the reference borrows them).
 *   - method: 0 (amx) and 2 (vnni) are accepted for compatibility and, like
 *     ALTO_METHOD_HIP, run the HIP engine; anything else returns -1.
 *   - on -3 (unsupported tail configuration) no output buffer is written.
 *
 * Everything below "Extensions" has no counterpart in the reference.
 */
#ifndef ALTOPIR_ENGINE_H
#define ALTOPIR_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALTO_METHOD_AMX 0  /* src/engine.rs:44-49 */
#define ALTO_METHOD_VNNI 2 /* src/engine.rs:44-49 */
#define ALTO_METHOD_HIP 4  /* new: the HIP engine under its own name */

/* replaces sw_amx_init, src/amx_kernels.cpp:151 (src/engine.rs:5).  Returns 0: the GPU
 * engine needs no CPU tile state. */
int sw_amx_init(void);

/* replace sw_alloc_zeroed / sw_alloc / sw_free, src/amx_kernels.cpp:97-113
 * (src/engine.rs:6-8).  Host memory, 2 MiB aligned; nullptr on failure. */
void *sw_alloc_zeroed(size_t bytes);
void *sw_alloc(size_t bytes);
void sw_free(void *p);

/* replaces sw_tile_db, src/amx_kernels.cpp:138-149 (src/engine.rs:9).  Writes the same
 * 16x64 tile order as the reference so that any consumer of the tiled copy keeps
 * working; this engine itself uploads from db_col. */
void sw_tile_db(uint8_t *dst, const uint8_t *db, size_t ld, size_t k, size_t m, int threads);

/* replaces sw_pack_build, src/amx_kernels.cpp:311-341 (src/engine.rs:10-11).  Builds the
 * packing matrix in the reference's byte-plane tile layout (setup work, on `threads`
 * host threads). */
void sw_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                   const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg,
                   size_t n, size_t rho, uint64_t qmod, int threads);

/* No counterpart in the reference: sw_pack_build computed on the device.  Same arguments and
 * the same bytes in `ub` (host memory, 4 * tg * n * rho * n bytes) as src/amx_kernels.cpp:311-341;
 * the inputs are copied to the device, each output column is summed in LDS (u64 wrap-around, as
 * the reference's column) and reduced mod qmod there, and U is copied back.  `threads` is unused.
 * Table entries outside [0, n) and generator powers outside [1, 2n), which the reference would
 * use as out-of-range indices, are skipped.  Returns 0, or -1 (alto_last_error says why) for a
 * shape it does not take (tg * n a multiple of 64 of at most 20,480, rho * n a multiple of 32),
 * null pointers, no device or a failed device call. */
int alto_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                    const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg, size_t n, size_t rho,
                    uint64_t qmod, int threads);

/* No counterpart in the reference's C ABI: the offline hint of the reference's setup
 * (YServer::answer_hint_ring, src/server.rs:738-768, with multiply_with_db_ring, :656-709)
 * computed on the device from the engine's resident database:
 *   hint[z * cols + col] = sum_r (a_r * d_{col,r})[z]  in Z_Q[x]/(x^n + 1),  z < n,
 * where d_{col,r}[j] is the byte of column col in record r * n + j and a_r (n coefficients,
 * a[r * n + j]) is query polynomial r in coefficient form: the reference's
 * query_raw_transformed (after negacyclic_perm, before its NTT), reduced mod Q.  a: npolys =
 * rows / n polynomials (host); hint: n * cols values < Q (host).  Returns 0, or -1
 * (alto_last_error says why) when rows is not npolys * n or a device call fails. */
int alto_engine_hint0(void *e, const uint64_t *a, size_t npolys, uint64_t *hint);
/* The same hint from a column-major database (column col at db_col + col * ld, rows bytes; host
 * memory, or device memory when on_device != 0) on the current device, without an engine (the
 * reference computes the hint before it can build the packing matrix an engine needs): the
 * columns are tiled on the device in slices, then as above.  rows a multiple of n and of 64,
 * cols of 16.  Returns 0, or -1 (alto_last_error says why). */
int alto_hint0(const uint8_t *db_col, int on_device, size_t ld, size_t rows, size_t cols, size_t n, const uint64_t *a,
               size_t npolys, uint64_t *hint);

/* No counterpart in the reference's C ABI: the offline packing precomputation of the
 * reference's setup (full_packing_with_preprocessing_offline with prep_pack_many_lwes and
 * precompute_aggregates_ntt, src/packing.rs:910-974, 2260-2359, called from
 * perform_offline_precomputation_simplepir, src/server.rs:947-1021) for `parts` output parts on
 * the device.  hint: n rows of hint_cols values (the offline hint, row z = coefficient z, as
 * alto_hint0 writes it); part o takes columns [o n, (o + 1) n).  w, v: tg x n NTT slots of the
 * two packing keys' masks.  Outputs (host memory, part-major, u64 values < Q): bold_t and
 * bold_t_bar ((n/2 - 1) tg n per part, [(i tg + k) n + slot], what sw_pack_build /
 * alto_pack_build read), bold_t_hat (tg n per part: the tail keys `that`), a_hat (n per
 * part: the packed ciphertext's first row, coefficients).  The NTT is the engine's (slot j =
 * evaluation at psi^(2 bitrev(j) + 1), psi the smallest primitive 2n-th root of unity mod Q,
 * so the inverse table is psi^-bitrev(k) / 2); the gadget digits are base 2^ceil(32 / tg),
 * least significant first.  n a power of two in [32, 4096], parts n <= hint_cols, tg <= 32
 * with (1 + tg) n words within a workgroup's LDS.  Returns 0, or -1 (alto_last_error says
 * why). */
int alto_inspiring_offline(uint64_t *bold_t, uint64_t *bold_t_bar, uint64_t *bold_t_hat, uint64_t *a_hat,
                           const uint64_t *hint, size_t hint_cols, size_t parts, const uint64_t *w,
                           const uint64_t *v, size_t n, size_t tg);

/* No counterpart in the reference's C ABI: the server's whole offline setup in one call
 * (src/service.rs:129-157, Prepared::new: perform_offline_precomputation_simplepir, then
 * build_packing_matrix), on the device: the hint as alto_hint0 computes it (db_col, on_device,
 * ld, rows, cols, a, npolys as there), the packing precomputation as alto_inspiring_offline
 * (w, v, n, tg as there, parts = cols / n) and the packing matrix from it as alto_pack_build
 * (the automorphism tables of the same NTT, generator 5, num_iter = n/2 - 1, Q), without the
 * hint or the aggregates leaving the device.  Outputs (host): ub (4 tg n cols bytes, the layout
 * of sw_pack_build), that (tg n per part: bold_t_hat, the engine's tail keys), a_hat (n per
 * part).  Returns 0, or -1 (alto_last_error says why). */
int alto_scheme_setup(int8_t *ub, uint64_t *that, uint64_t *a_hat, const uint8_t *db_col, int on_device, size_t ld,
                      size_t rows, size_t cols, const uint64_t *a, size_t npolys, const uint64_t *w,
                      const uint64_t *v, size_t n, size_t tg);

/* replaces alto_engine_new, src/amx_kernels.cpp:546-560 (src/engine.rs:12-14).
 * Returns nullptr for the configurations the reference rejects (Q != 4294955009,
 * n == 0 or n > 4096 or n % 16, cols % n, rows % 64, cols % 32, threads <= 0,
 * (tg * n) % 128), for rows == 0, cols == 0 or tg == 0, and when no GPU or not
 * enough device memory is available. */
void *alto_engine_new(const uint8_t *db_col, const uint8_t *db_tiled, size_t ld, size_t rows,
                      size_t cols, const int8_t *ub, size_t n, size_t tg,
                      const uint8_t *const *prefix, size_t part_len, size_t row1_off, uint64_t q1,
                      int q1_bits, const uint64_t *const *that, const uint64_t *inv_table,
                      uint64_t q, int threads);

/* replaces alto_engine_free, src/amx_kernels.cpp:561 (src/engine.rs:15). */
void alto_engine_free(void *e);

/* replace alto_request_len / alto_response_len, src/amx_kernels.cpp:562-563
 * (src/engine.rs:16-17): 9 + 4 * (2 * tg * n + rows) and (cols / n) * part_len. */
size_t alto_request_len(const void *e);
size_t alto_response_len(const void *e);

/* replaces alto_answer, src/amx_kernels.cpp:564-615 (src/engine.rs:18-19).
 * Synchronous.  status[c]: 1 null frame or wrong length, 2 flag or key-length field,
 * 3 key word >= Q, 0 answered.  Only answered outputs are written.  Returns the number
 * answered, -1 on bad arguments or method, -3 on an unsupported tail configuration.
 * stage_ms (optional) receives {validate, scan, pack, tail} in milliseconds (device
 * time of each stage; request upload is counted in validate, response download in
 * tail).  Not re-entrant on one engine. */
int alto_answer(void *e, const uint8_t *const *req, const size_t *len, size_t B,
                uint8_t *const *out, int *status, int method, double *stage_ms);

/* replace matMulVecPacked / matMulVecPacked2 / 4 / 8, src/matmul.cpp:105, 220, 676, 1548
 * (declared at src/matmul.rs:3-26), the SimplePIR packed product off the engine path:
 * out[i * bc + c] += sum_j sum_{t<4} byte_t(a[i * aCols + j]) * b[c * 4 * aCols + 4 j + t]
 * (mod 2^32) for bc = 1, 2, 4, 8 columns of b.  Host buffers; runs on the scan kernels;
 * exactly aRows rows are written (the reference works in groups of eight rows). */
void matMulVecPacked(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t aRows, size_t aCols);
void matMulVecPacked2(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols);
void matMulVecPacked4(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols);
void matMulVecPacked8(uint32_t *out, const uint32_t *a, const uint32_t *b_full, size_t aRows, size_t aCols);

/* replaces fast_batched_dot_product_avx512<K, T> (src/kernel.rs:14-22, 143-151), the
 * two-modulus (CRT) product of the reference's SimplePIR server path, off the engine path
 * (callers src/server.rs:593-618, 1446-1547).  The reference passes spiral-rs Params; here the
 * two moduli come explicitly (m0 m1 < 2^64, coprime) and the strides of a and c (the
 * reference's a.len() / K and c.len() / K).  For r < K, j < b_cols:
 *   c[r c_stride + j] = (c + crt(sum_k lo(a[r a_stride + k]) b[j b_rows + k] mod 2^64 mod m0,
 *                                sum_k hi(a[...]) b[...] mod 2^64 mod m1)) mod m0 m1,
 * lo / hi the low and high 32 bits, b elements of b_bytes = 1, 2 or 4 bytes (the reference's T),
 * a_elems == b_rows.  Host buffers; computed on the device.  0, or -1 (alto_last_error). */
int alto_fast_batched_dot_product(uint64_t *c, size_t c_stride, const uint64_t *a, size_t a_stride, size_t a_elems,
                                  const void *b_t, int b_bytes, size_t b_rows, size_t b_cols, size_t K, uint64_t m0,
                                  uint64_t m1);

/* ---------------------------------------------------------------- Extensions -- */

/* The packed product for any number of columns of b; 0, or -1 (bad arguments, no device,
 * device memory). */
int alto_matmul_vec_packed(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t a_rows, size_t a_cols,
                           size_t b_cols);

/* Device-resident batch.  Frame c lies at slots + c * slot_stride + ALTO_FRAME_OFFSET in
 * device memory (slots 16-byte aligned, slot_stride a multiple of 16 and at least
 * ALTO_FRAME_OFFSET + request_len); every frame is request_len bytes long.  status is a
 * device array of B ints: on entry 0 for a frame to check, nonzero for a frame the
 * caller already rejected (it is left as is); on return the answer statuses.  Output c
 * is written at out + c * out_stride (device memory) when its status is 0.  Everything
 * is enqueued on `stream` (a hipStream_t) and nothing is waited for.  With stream 0 the
 * batch keeps the null stream's order: it runs on the engine's own stream (which, unlike
 * the null stream, can be captured into the batch's HIP graph), after the work enqueued on
 * the null stream before the call, and work enqueued there after the call runs after the
 * batch (so two batches on the null stream run one after the other; pass streams of one's
 * own to overlap them).  Returns 0, -1 on bad arguments, -3 on an unsupported tail
 * configuration, -4 when device memory for the batch cannot be allocated. */
#define ALTO_FRAME_OFFSET 7
int alto_answer_device(void *e, const uint8_t *slots, size_t slot_stride, size_t B, uint8_t *out,
                       size_t out_stride, int *status, void *stream);

/* Contexts: the per-batch state of an engine (workspace, stream, events, captured graphs).
 * alto_answer and alto_answer_device use the engine's default context
 * (alto_engine_context); more contexts let several batches of one engine be in flight at
 * once (one batch per context at a time).  alto_context_reserve sizes a context's
 * workspace ahead of time so that no batch of up to B queries grows it (B rounded up to the
 * top of its batch-size bucket, whose captured graph is sized for that bucket's largest
 * batch); 0, or -4 when device memory is short. */
/* No counterpart in the reference: alto_answer_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, statuses, returns and results as alto_answer_device, but it
 * calls no hipMalloc, hipFree or other allocating routine: the batch workspace (limbs,
 * accumulators, per-query sums) lives in one fixed device array of
 * ALTO_NOALLOC_SCRATCH_BYTES per device, held until the process ends, and the kernels clear
 * what must start at zero on every call.  The launches are eager, on `stream` (0: the
 * engine's own stream, in the null stream's order as for alto_answer_device), and never
 * captured by the library.  A batch of B queries needs
 *   W(B) = sum over {4 Bp rows, 4 Bp tg n, 4 Bp cols, 4 Bp cols, 4 Bp, 8 Bp} of each size
 *          rounded up to 256 bytes,  Bp = B rounded up to a multiple of 4,
 * and is refused (-1, alto_last_error says why) when W(B) > ALTO_NOALLOC_SCRATCH_BYTES
 * (B <= 156 at 1 GiB of 32 KB records, 48 at 8 GiB, 4 at 64 GiB).  Every allocation-free call on
 * one device shares the array, which keeps the last call's contents: calls that overlap on
 * one device (several streams or threads) race and must be serialised by the caller. */
#define ALTO_NOALLOC_SCRATCH_BYTES (64u << 20)
int alto_answer_device_noalloc(void *e, const uint8_t *slots, size_t slot_stride, size_t B, uint8_t *out,
                               size_t out_stride, int *status, void *stream);

/* Device bytes a context of engine e holds once reserved for B queries (its per-batch
 * workspace, sized as alto_context_reserve sizes it; the serving loop admits contexts by
 * it).  No counterpart in the reference. */
size_t alto_context_bytes(const void *e, size_t B);
void *alto_context_new(void *e);
void alto_context_free(void *ctx);
void *alto_engine_context(void *e);
int alto_context_reserve(void *ctx, size_t B);
/* Graph mode of one context: 1 capture and replay, 0 eager (drops its graphs), -1 follow
 * the engine's alto_set_graphs (the default).  Returns 0, -1 on bad arguments. */
int alto_context_set_graphs(void *ctx, int mode);
/* Number of launch chains the context holds captured.  With graphs on, a batch of up to 2048
 * queries replays the graph of its batch-size bucket (captured once, for the bucket's largest
 * batch, and told the batch's own size at each replay); larger batches and timed replays are
 * captured per exact size.  Returns -1 for a null context. */
int alto_context_graphs(const void *ctx);

/* alto_answer_device on a context, with frames and outputs gathered through optional
 * device index arrays: frame c at slots + slot_index[c] * slot_stride + ALTO_FRAME_OFFSET,
 * output c at out + out_index[c] * out_stride (null index = position c).  Stream 0: the
 * context's own stream, in the null stream's order. */
int alto_answer_gather(void *ctx, const uint8_t *slots, size_t slot_stride, const int *slot_index,
                       size_t B, uint8_t *out, size_t out_stride, const int *out_index, int *status,
                       void *stream);

/* Stage timings of a context's last eager batch (same layout as alto_last_timings). */
int alto_context_timings(void *ctx, double *ms, int count);

/* Graph mode of alto_answer_device (on by default): the first batch with a given set of
 * pointers, size, stream and kernel variants is captured into a HIP graph and every later
 * identical batch replays it (at most 64 graphs per engine, least recently used dropped;
 * growing the batch workspace drops them all).  Stage timings are recorded with graphs off,
 * or with graphs on when alto_set_graph_timing is on.  Returns 0, -1 on a null engine. */
int alto_set_graphs(void *e, int on);

/* Stage timing of graph replays (off by default): when on, a batch is captured once per
 * timing event set (4 sets, so up to 4 graphs per distinct batch) with the stage events
 * inserted as event-record nodes, and each replay feeds alto_stage_totals and
 * alto_last_timings as an eager batch does.  Each event node adds about 8 us to a replay
 * on MI355X (ROCm 7.2), so this is a diagnostic, not the way to time throughput.  Returns
 * 0, -1 on a null engine. */
int alto_set_graph_timing(void *e, int on);

/* One launch for both contractions of a batch: when the scan and packing shapes chosen for
 * a batch size are a pair the library fuses (the built-in policy's pairs), the packing
 * tiles run in the same launch as the scan tiles (they are independent), after both limb
 * preparations.  The stage timings then report the pair under "scan" and zero for packing.
 * on: 1 fuse every such pair; 0 never; -1 (the default) only batches of 1-4 queries on
 * databases of at most 4 * tg * n records, where one launch of the two streaming
 * contractions measured 4% faster on MI355X (elsewhere it was level or slower: each
 * contraction alone fills the chip).
 * on = 2: the whole batch after validation and limb preparation as one persistent launch
 * (the step kernel): scan and packing tiles, and, as each output part's tiles have all
 * stored their results (counted per part in the workspace), that part's tails, one per
 * query, claimed by the workgroups whose tiles are done; the stage timings report it under
 * "scan" and zero for packing and tail.  Every batch size has a step pair in the built-in
 * policy (for 9-16 queries the packing then runs on 8 x 16-tile workgroups, the scan's
 * workgroup size); batch sizes whose tuned shapes have none run as with on = 1.
 * alto_set_fusion returns the previous setting (-1 to 2), -1 also on a null engine or a
 * bad value; alto_fused returns 2 if a batch of B queries runs as one step launch, 1 if its
 * two contractions run in one launch, 0 if not, -1 on bad arguments. */
int alto_set_fusion(void *e, int on);
int alto_fused(const void *e, size_t B);

/* Device time in milliseconds of the stages of the last batch (synchronous entry point, or
 * device entry point with graphs off), waiting for it to finish: {validate, limb preparation
 * (query and key limbs, one launch), scan contraction, packing limb preparation (always 0:
 * it runs in the launch before), packing contraction, tail}.  Writes up to
 * `count` values and returns how many were written (0 if no batch ran yet). */
int alto_last_timings(void *e, double *ms, int count);

/* Running sums of the same six stage times over every timed batch of a context since the
 * last reset (timed batches: the synchronous entry point, and the device entry points with
 * graphs off).  Each batch's events are kept in a ring of 32 and folded in when the ring
 * comes round or here, so back-to-back batches are timed without the host waiting on each.
 * Waits for the timed batches still in flight; reset != 0 zeroes the sums afterwards.
 * Returns the number of batches summed, -1 on bad arguments. */
int alto_context_stage_totals(void *ctx, double *ms, int count, int reset);
/* alto_context_stage_totals on the engine's own context. */
int alto_stage_totals(void *e, double *ms, int count, int reset);

/* alto_engine_new with the column-major database already in device memory
 * (db_col_device[col * ld + k]); the engine tiles it into its own copy, after which the
 * caller may free it.  Device inputs of every set-up entry point (here, alto_engine_new_records,
 * alto_hint0 and alto_scheme_setup with on_device) are read after the work already enqueued
 * on the null stream; work of the caller's own streams must be complete before the call.
 * The other arguments and checks are those of alto_engine_new. */
void *alto_engine_new_device(const uint8_t *db_col_device, size_t ld, size_t rows, size_t cols,
                             const int8_t *ub, size_t n, size_t tg, const uint8_t *const *prefix,
                             size_t part_len, size_t row1_off, uint64_t q1, int q1_bits,
                             const uint64_t *const *that, const uint64_t *inv_table, uint64_t q,
                             int threads);

/* alto_engine_new with the database given as records (row-major, as a database file holds
 * them: record k, `cols` bytes, at records + k * ld, ld >= cols), in host memory
 * (on_device = 0) or device memory; the records are moved to the device in slices and
 * transposed into the engine's layout there (the reference transposes on the host,
 * src/server.rs:251-366).  The other arguments and checks are those of alto_engine_new. */
void *alto_engine_new_records(const uint8_t *records, int on_device, size_t ld, size_t rows, size_t cols,
                              const int8_t *ub, size_t n, size_t tg, const uint8_t *const *prefix,
                              size_t part_len, size_t row1_off, uint64_t q1, int q1_bits,
                              const uint64_t *const *that, const uint64_t *inv_table, uint64_t q, int threads);

/* Kernel variants of the two contractions (scan and packing) and the per-engine tuning
 * table that picks one per batch-size bucket (half octaves of MT = ceil(B / 4): bucket b
 * ends at MT = 1, 2, 3, 4, 6, 8, 12, 16, 24, ... for b = 0, 1, 2, ...; 30 buckets).
 * which: 0 = scan, 1 = packing; idx -1 restores the built-in choice.  Every variant
 * computes the same bytes; they differ only in tile shapes and operand staging. */
int alto_variant_count(void);
const char *alto_variant_name(int idx);
int alto_tune_set(void *e, int which, size_t B, int idx);
int alto_tune_get(const void *e, int which, size_t B);

/* ------------------------------------------------------------- serving loop -- */

/* A device-resident request loop over one or more engines (resident databases) with
 * continuous batching: a submitted frame is copied into a pinned slot and (by default) into
 * HBM by the database's upload thread, with DMA copies through the HSA runtime (the library
 * links libhsa-runtime64 and holds one reference on it for the life of the process), and
 * whenever one of a database's contexts is free, every waiting request of that database
 * (up to max_batch) is answered as one batch; batches of one database overlap on the
 * device as far as its contexts allow.  Each context is reserved for max_batch queries; a
 * database starts with one and gets another, up to contexts_per_db, when requests wait, all
 * of its contexts are busy and the device memory then free is at least twice one context's
 * footprint (memory admission at run time; a refused admission is retried after the next
 * batch completes; alto_server_get_stats counts both).  Responses are written by the device straight into
 * pinned host memory and stay valid until alto_server_release.  Creating a server writes
 * every page of its pinned slot pools once (slots_per_db x frame size per database), so
 * that no submit takes first-touch page faults. */
typedef struct {
    uint64_t tag;          /* the caller's tag from alto_server_submit */
    int db;                /* database (engine index) */
    int status;            /* 0 answered; 1 wrong length; 2, 3 as alto_answer; 9 no free slot;
                              -1 engine failure */
    int slot;              /* response slot to release (status 0), else -1 */
    const uint8_t *response;
    size_t response_len;
    int64_t t_submit_ns;   /* steady-clock time of submit and of completion */
    int64_t t_done_ns;
    int batch;             /* size of the batch that answered it (0 if never batched) */
} alto_completion;

typedef struct {
    long long submitted, queue_full, batches, answered_in_batches, max_batch_seen;
    long long contexts_admitted, contexts_refused;  /* run-time memory admission (beyond the first) */
} alto_server_stats;

void *alto_server_new(void *const *engines, int n_engines, int max_batch, int slots_per_db,
                      int contexts_per_db);
/* Returns 0 (queued), 1 (wrong length) or 9 (no free slot; both completed at once with that
 * status), -1 on bad arguments.  A frame whose upload fails later completes with status -1
 * (alto_server_error names the cause). */
int alto_server_submit(void *srv, int db, const uint8_t *frame, size_t len, uint64_t tag);
/* Waits up to timeout_ms for completions; returns how many were written to out. */
int alto_server_poll(void *srv, alto_completion *out, int max, int timeout_ms);
int alto_server_release(void *srv, int db, int slot);
int alto_server_contexts(void *srv, int db);
int alto_server_get_stats(void *srv, alto_server_stats *out);
const char *alto_server_error(void *srv);
void alto_server_free(void *srv);

/* Native load driver for the serving loop, standing in for a server's connection readers:
 * `threads` submitter threads send `total` requests for database `db` (frame i % nframes,
 * tag i), keeping at most `window` outstanding, while a collector polls completions and
 * releases their slots.  The caller must not poll the server meanwhile.  Fills `out` with
 * the answered rate over the run and the submit-to-completion latency percentiles.
 * Returns 0, -1 on bad arguments. */
typedef struct {
    long long answered, rejected, batches;
    double wall_s, qps, mean_batch, p50_ms, p90_ms, p99_ms, max_ms;
} alto_replay_result;
int alto_server_replay(void *srv, int db, const uint8_t *const *frames, const size_t *lens, int nframes,
                       long long total, int threads, int window, alto_replay_result *out);
/* Graph mode of every context of the server (see alto_context_set_graphs; the server's
 * contexts start in -1, following their engine).  Returns 0, -1 on bad arguments. */
int alto_server_set_graphs(void *srv, int mode);
/* How frames reach the device: 0 (the default, also chosen by -1) frames are copied into the
 * device arena as they arrive, by an upload thread of the database (one DMA copy per run of
 * consecutive slots, waited for on the host, so that no upload waits behind a batch's
 * kernels); 1 the batch's kernels read the frames straight from the pinned slots over PCIe
 * (no per-frame copy); 2 a batch's frames are copied into the device arena when it is
 * dispatched, one copy per run of consecutive slots, on the batch's stream; 3 as 2, and the
 * batch's responses are written to device memory and copied back the same way (in modes 0-2
 * the tail writes them to the pinned response slots over PCIe).  Only while no request is
 * pending, uploading or in flight; returns 0, -1 otherwise or on bad arguments. */
int alto_server_set_upload(void *srv, int mode);
/* The upload mode in force for database db (0-3 as above), -1 on bad arguments. */
int alto_server_upload_mode(void *srv, int db);

/* TCP front end of the serving loop.  No counterpart in the reference's C ABI: it replaces
 * the reference server's connection readers and writers (src/bin/altopir.rs:403-442, :383-400)
 * and speaks their wire protocol (src/bin/altopir.rs:282-308), little-endian:
 *   request [u32 len][u64 id][len bytes]; reply [u32 len][u64 id][i32 status][len bytes]
 *   (status 0 with the response, otherwise len = 0); stop frame len = 0, id = 2^64 - 1.
 * Database i of `srv` listens on host:ports[i] (IPv4; port 0 takes a free port, see
 * alto_net_port); `engines` is the array given to alto_server_new (the request length of each
 * database comes from its engine).  A frame longer than the request length is discarded
 * without buffering and answered with status 1; a frame of another wrong length gets status 1
 * and a full slot pool status 9 (the reference's busy status) from alto_server_submit; the
 * device's validation gives 2 and 3.  Replies of one connection go out in completion order.
 * The front end polls the server's completions itself: the caller must not call
 * alto_server_poll or alto_server_replay on that server while it runs.  Returns null on bad
 * arguments or when a port cannot be bound. */
typedef struct {
    long long connections, frames, answered, rejected, busy, discarded, stop_frames, bytes_in, bytes_out;
} alto_net_stats;
void *alto_net_new(void *srv, void *const *engines, int n_engines, const char *host, const int *ports);
/* The port database db listens on, -1 on bad arguments. */
int alto_net_port(const void *net, int db);
/* Waits up to timeout_ms (< 0: without limit) for a stop frame on any connection; returns 1
 * once one has arrived, 0 on time-out, -1 on bad arguments. */
int alto_net_wait(void *net, int timeout_ms);
int alto_net_get_stats(const void *net, alto_net_stats *out);
/* Stops listening, closes every connection (replies not yet written are dropped and their
 * response slots released) and joins the front end's threads.  Free it before its server;
 * requests in flight at that moment complete into the server's completion queue. */
void alto_net_free(void *net);

/* Native open-loop load client for the front end above, after the reference's `load` command
 * (src/bin/altopir.rs:594-739): requests are due at exponential gaps of mean 1/rate over
 * `seconds` (generator seeded with `seed`), request i carries frame idx_i (uniform over the
 * nframes frames) and id i, and is sent at its due time over connection i % conns, each
 * connection with a sender and a receiver thread; senders half-close their connection
 * drain_ms after their last request.  Latency runs from the due time to the reply.
 * expect_status (per frame, may be null) and expect (per frame a response of expect_len bytes
 * or null, array may be null) check the replies: wrong_status counts replies other than the
 * expected status or 9, checked/bad the answered replies compared.  Returns 0, -1 on bad
 * arguments or when a connection cannot be opened. */
typedef struct {
    long long requests, answered, busy, rejected, missing, wrong_status, checked, bad;
    double seconds, answered_qps, p50_ms, p90_ms, p99_ms, max_ms;
} alto_net_load_result;
int alto_net_load(const char *host, int port, const uint8_t *const *frames, const size_t *lens, int nframes,
                  const int *expect_status, const uint8_t *const *expect, size_t expect_len, double rate,
                  double seconds, int conns, uint64_t seed, int drain_ms, alto_net_load_result *out);

/* Number of HBM bytes the engine holds for its resident state (database, packing
 * matrix, tail keys, prefixes). */
size_t alto_engine_device_bytes(const void *e);

/* Last error message of this thread ("" if none). */
const char *alto_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
