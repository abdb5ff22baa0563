/*
 * pico_gpu.h -- C ABI of the MI355X (gfx950) backend for the commitment hot
 * path of Pico's STARK prover over KoalaBear / BabyBear.
 *
 * This is the boundary a Plonky3-side adapter binds (see INTEGRATION.md for
 * the Rust extern block and trait impls).  It replaces, behind the traits
 * that Pico's StarkGenericConfig plugs in
 * (vm/src/configs/stark_config/kb_poseidon2.rs:15-27, built at :82-93):
 *   TwoAdicSubgroupDft::coset_lde_batch    -> pg_coset_lde
 *   Mmcs::commit / open_batch / get_matrices / verify_batch
 *                                          -> pg_mmcs_commit, pg_mmcs_open_batch,
 *                                             pg_tree_matrix, pg_mmcs_verify_batch
 *   Pcs::commit / get_evaluations_on_domain -> pg_pcs_commit, pg_pcs_evaluations
 *   FRI commit phase, fold, PoW, query openings (p3-fri prover)
 *                                          -> pg_fri_commit_phase, pg_fri_fold,
 *                                             pg_fri_grind, pg_fri_answer_queries
 *   DuplexChallenger                       -> pg_challenger_*
 *
 * Conventions
 *   - Field elements are u32 in Montgomery form (x * 2^32 mod p), the
 *     in-memory layout of p3 MontyField31, unless a `canonical` flag says
 *     otherwise.  Matrices are row-major.
 *   - Every function returning int returns PG_OK (0) or a negative PG_E_*
 *     code; pg_last_error() then holds a message for the calling thread.  The
 *     reference panics on every failure (e.g. vm/src/machine/prover.rs:593);
 *     an adapter maps non-zero codes to panic!(pg_last_error()).
 *   - All entry points are thread-safe.  Each call runs on a stream of its
 *     own and returns when its device work is complete; device buffers and
 *     trees returned by a call are ready for use by any other call.
 *   - Heights must be powers of two (all Pico trace and LDE heights are).
 */
#ifndef PICO_GPU_H
#define PICO_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_E_INVALID (-1)     /* bad argument or shape */
#define PG_E_OOM (-2)         /* device allocation failed */
#define PG_E_DEVICE (-3)      /* HIP runtime / kernel error */
#define PG_E_ADMISSION (-4)   /* would exceed the context's HBM budget */
#define PG_E_UNSUPPORTED (-5) /* valid request outside the implemented range */
#define PG_E_CANCELLED (-6)   /* service job cancelled before it ran */

#define PG_KOALABEAR 0 /* p = 2^31 - 2^24 + 1 */
#define PG_BABYBEAR 1  /* p = 2^31 - 2^27 + 1 */

typedef struct pg_ctx pg_ctx;   /* one device + HBM budget + table caches */
typedef struct pg_buf pg_buf;   /* device matrix (rows x cols u32) */
typedef struct pg_tree pg_tree; /* MMCS prover data: matrices + digest layers */
typedef struct pg_fri pg_fri;   /* FRI commit-phase prover data */
typedef struct pg_open_proof pg_open_proof; /* everything Pcs::open returns */

/* DuplexChallenger<F, Poseidon2, 16, 8> state (Montgomery values).
 * Replaces the reference's DuplexChallenger (restated in-tree at
 * vm/src/compiler/recursion/circuit/challenger.rs:57-227). */
typedef struct {
    int32_t field;
    uint32_t state[16];
    uint32_t input[8];
    int32_t n_input;
    uint32_t output[8];
    int32_t n_output;
} pg_challenger;

const char *pg_last_error(void);
int pg_version(void);

/* ---- context ------------------------------------------------------------ */
/* hbm_budget_bytes = 0 means "all free device memory". */
int pg_create(int device, size_t hbm_budget_bytes, pg_ctx **out);
void pg_destroy(pg_ctx *ctx);
/* wait for all work on the device (timing brackets; not needed for
 * correctness, every call below completes its own work) */
int pg_device_synchronize(int device);
size_t pg_bytes_in_use(const pg_ctx *ctx);
/* Per-kernel timing with HIP events on the launch stream (adds one host sync
 * per launch while enabled).  Families: lde_single, lde_pass_a, lde_pass_bc,
 * lde_pass_d, leaf_hash, leaf_hash_w8, compress, compress_inject, fri_fold.
 * No counterpart in the reference (its spans time whole phases,
 * vm/src/machine/prover.rs:101,414,539,588); used by the benchmark. */
int pg_profile_enable(pg_ctx *ctx, int enable);
int pg_profile_reset(pg_ctx *ctx);
int pg_profile_get(pg_ctx *ctx, const char *family, double *total_ms, uint64_t *launches);
/* Fault injection for tests of the error paths: the device allocation
 * `after` allocations from now (0 = the next one) fails once with
 * PG_E_OOM; a negative value disarms.  No counterpart in the reference
 * (which panics on allocation failure); every entry point must then return
 * the error and leave no device memory behind. */
int pg_debug_fail_alloc(pg_ctx *ctx, long long after);
/* Column-block sizing of pg_pcs_commit_host, for tests that exercise the
 * blocked path on small matrices: blocks of about `block_bytes`, at least
 * `min_cols` columns wide (0 restores the defaults, 256 MiB and 64).  No
 * counterpart in the reference. */
int pg_debug_host_blocks(pg_ctx *ctx, size_t block_bytes, size_t min_cols);

/* ---- LDE tile-shape autotuning ------------------------------------------------
 * Each LDE pass (single, A, BC, D) of a size-2^s transform runs one of up to
 * two tile shapes (register-DFT size x threads per tile).  The choice is kept
 * per (field, pass, s, width bucket: w <= 8, <= 32, <= 128, larger) in a
 * process-wide table read by every LDE launch (batches fix theirs at
 * pg_batch_create); variant 0 is the shape measured best on MI355X when the
 * library was written.  pg_autotune times both shapes of every pass of LDEs
 * of heights 2^log_h_min .. 2^log_h_max and the given widths (synthetic
 * inputs, `reps` timed runs after one warm-up) and keeps the faster per
 * entry; `report` (may be null) receives one line per decision.  Run it
 * before other work: the table is shared by every context and thread.
 * Export/import persist the table as text lines
 * "field kind log_s width_bucket variant" (kind: 0 single, 1 A, 2 BC, 3 D).
 * Results do not depend on the table: every shape computes the same LDE.
 * No counterpart in the reference (its CPU DFT has no tile shapes). */
int pg_autotune(pg_ctx *ctx, int field, int log_h_min, int log_h_max, const uint32_t *widths, size_t n_widths,
                uint32_t log_blowup, int reps, char *report, size_t report_len);
int pg_tune_set(int field, int kind, int log_s, int width_bucket, int variant);
int pg_tune_get(int field, int kind, int log_s, uint32_t width, int *variant);
int pg_tune_export(char *buf, size_t len, size_t *needed);
int pg_tune_import(const char *text);

/* ---- device matrices ------------------------------------------------------ */
int pg_buf_alloc(pg_ctx *ctx, size_t rows, size_t cols, pg_buf **out);
/* non-owning view of caller memory (e.g. a torch tensor) */
int pg_buf_wrap(pg_ctx *ctx, void *device_ptr, size_t rows, size_t cols, pg_buf **out);
void pg_buf_free(pg_buf *buf);
void *pg_buf_device_ptr(const pg_buf *buf);
size_t pg_buf_rows(const pg_buf *buf);
size_t pg_buf_cols(const pg_buf *buf);
/* Non-owning strided view: row r of the view is row row0 + r * row_step of
 * `parent`, columns col0 .. col0 + cols - 1.  It replaces the host-side
 * re-layouts that sit in front of Pcs::commit in the reference:
 *  - quotient chunks, `domain.split_evals(d, quotient_flat)`
 *    (vm/src/machine/prover.rs:513-523): chunk i of an (n d) x 4 quotient is
 *    view(row0 = i, row_step = d, col0 = 0, rows = n, cols = 4), committed on
 *    `domain.split_domains(d)[i]` (shift s * w_{n d}^i);
 *  - a column block of a wider device matrix.
 * (`flatten_to_base` of an extension-field trace, prover.rs:404-411, needs no
 * view: an h x w ext4 matrix in this layout is the h x 4w base matrix.)
 * Views are accepted as LDE inputs (pg_coset_lde, pg_pcs_commit,
 * pg_batch_create) and by pg_download; other entry points return
 * PG_E_INVALID for them.  The parent must outlive the view and anything
 * committed from it. */
int pg_buf_view(pg_ctx *ctx, const pg_buf *parent, size_t row0, size_t row_step, size_t col0,
                size_t rows, size_t cols, pg_buf **out);
/* row pitch in words (cols for a dense matrix) */
size_t pg_buf_ld(const pg_buf *buf);
/* whole-buffer upload / row-range download.  canonical != 0 means the host
 * side holds canonical values of `field`, converted on the device. */
int pg_upload(pg_ctx *ctx, pg_buf *buf, const uint32_t *host, int field, int canonical);
/* Page-locked host memory (hipHostMalloc) for traces produced on the host:
 * uploads from it are DMA'd directly (measured on MI355X, 1 GiB: 57.5 GB/s).
 * Uploads of 4 MiB or more from pageable memory go through the library's
 * page-locked double buffers, filled 16 MiB at a time by a pool of host
 * threads while the previous chunk's DMA runs (54 GB/s, against 26-30 GB/s
 * through the runtime's own staging).  Trace generation writes rows
 * straight into this memory; no function here registers memory that
 * already exists. */
int pg_host_alloc(pg_ctx *ctx, size_t bytes, void **out);
void pg_host_free(void *p);
int pg_download(pg_ctx *ctx, const pg_buf *buf, size_t row0, size_t nrows, uint32_t *host,
                int field, int canonical);
/* Seeded synthetic input: element i of stream `seed` is the first of
 * SplitMix64(seed) outputs 4i..4i+3 whose top 31 bits are < p (else the last
 * one mod p), stored in Montgomery form. */
int pg_fill_random(pg_ctx *ctx, pg_buf *buf, int field, uint64_t seed);

/* ---- TwoAdicSubgroupDft::coset_lde_batch ----------------------------------
 * evals: h x w values on <w_h>; out: (h << added_bits) x w, the evaluations on
 * shift * <w_N>, in bit-reversed row order when bit_reversed_rows != 0 (the
 * order TwoAdicFriPcs commits after .bit_reverse_rows()), else natural order.
 * h <= 2^22. */
int pg_coset_lde(pg_ctx *ctx, int field, const pg_buf *evals, uint32_t added_bits,
                 uint32_t shift_mont, int bit_reversed_rows, pg_buf **out);

/* ---- Mmcs (MerkleTreeMmcs<.., PaddingFreeSponge<16,8,8>, TruncatedPermutation<2,8,16>, 8>)
 * commit: matrices in caller order; take_ownership != 0 makes the tree free
 * them (p3 commit consumes its inputs). */
int pg_mmcs_commit(pg_ctx *ctx, int field, pg_buf *const *mats, size_t n, int take_ownership,
                   uint32_t root[8], pg_tree **out);
/* open_batch for n_idx query indices at once.  rows_out: n_idx * total_width
 * words (each query: the opened row of every matrix, original order);
 * paths_out: n_idx * log_height * 8 words (sibling digests, leaf level first). */
int pg_mmcs_open_batch(pg_ctx *ctx, const pg_tree *tree, const uint64_t *indices, size_t n_idx,
                       uint32_t *rows_out, uint32_t *paths_out);
/* host-side verify_batch; returns 1 if the opening re-derives root, 0 if not,
 * negative on bad arguments */
int pg_mmcs_verify_batch(int field, const size_t *heights, const size_t *widths, size_t n,
                         const uint32_t root[8], uint64_t index, const uint32_t *rows,
                         const uint32_t *path);
int pg_tree_root(const pg_tree *tree, uint32_t root[8]);
size_t pg_tree_num_matrices(const pg_tree *tree);
const pg_buf *pg_tree_matrix(const pg_tree *tree, size_t k);
int pg_tree_log_height(const pg_tree *tree);
size_t pg_tree_total_width(const pg_tree *tree);
/* copy digest layer `layer` (H >> layer digests of 8 words) to host */
int pg_tree_digest_layer(pg_ctx *ctx, const pg_tree *tree, int layer, uint32_t *host);
void pg_tree_free(pg_tree *tree);

/* ---- Pcs (TwoAdicFriPcs) ----------------------------------------------------
 * commit: for each matrix k, LDE with shift = GENERATOR / domain_shift[k]
 * (trace domains have shift 1), bit-reversed rows, then one MMCS tree over
 * all LDEs (which the tree owns). */
int pg_pcs_commit(pg_ctx *ctx, int field, const pg_buf *const *evals,
                  const uint32_t *domain_shifts_mont, size_t n, uint32_t log_blowup,
                  uint32_t root[8], pg_tree **out);
/* Pcs::commit as the reference calls it, on matrices in host memory
 * (vm/src/machine/prover.rs:244 commits the traces `generate_main` built on
 * the host, :169-223, sorted by (height desc, name) at :192).  Matrix k is
 * rows[k] x cols[k] words at host[k] (Montgomery form, or canonical values
 * when canonical != 0).  Same result as pg_upload of each matrix followed by
 * pg_pcs_commit; the uploads run on their own stream, tallest height group
 * first, overlapping the LDE and hashing of the groups before them.  Host
 * memory from pg_host_alloc uploads at the link's rate. */
int pg_pcs_commit_host(pg_ctx *ctx, int field, const uint32_t *const *host, const size_t *rows,
                       const size_t *cols, const uint32_t *domain_shifts_mont, size_t n,
                       uint32_t log_blowup, int canonical, uint32_t root[8], pg_tree **out);
/* get_evaluations_on_domain (p3 TwoAdicFriPcs, prover.rs:467-490 is the
 * caller): the first `rows` committed rows of matrix k, i.e. its evaluations
 * on the sub-coset of `rows` points, returned in natural order (rows * width
 * Montgomery words).  The rows are put in natural order on the device; a
 * pageable destination is filled through the page-locked stage. */
int pg_pcs_evaluations(pg_ctx *ctx, const pg_tree *tree, size_t k, size_t rows, uint32_t *host);

/* ---- Pcs::open (device part) -----------------------------------------------
 * p3 TwoAdicFriPcs::open evaluates every committed matrix at its points,
 * has the challenger observe the values (round -> matrix -> point -> column,
 * circuit/fri.rs:84-99), samples alpha and reduces all committed rows to one
 * ext4 vector per LDE height (the verifier side: circuit/fri.rs:118-241),
 * which FRI then folds (pg_fri_commit_phase takes these buffers as inputs).
 *
 * pg_pcs_open_values: matrix k of tree t at npts ext4 points (4 Montgomery
 * words each), from the trace-domain coset GEN <w_h> held in the first
 * height >> log_blowup rows of its LDE.  out: npts x width x 4 words,
 * point-major.  Points must lie off that coset (an ext4 point with a
 * non-zero higher coefficient always does).
 *
 * pg_pcs_reduce_openings: rounds[r] are the committed trees in opening order;
 * npoints[] (0..2) gives the point count of every matrix of every round in
 * that order, points_mont the points (4 words each) and values_mont the
 * opened values (per point, width x 4 words) in the same order.  Writes a
 * new 2^lh x 4 buffer (ext4, committed row order) to out_by_log[lh] for each
 * LDE log-height lh that has points, NULL elsewhere (33 entries).  At most
 * 4 distinct points per height. */
int pg_pcs_open_values(pg_ctx *ctx, const pg_tree *t, size_t k, uint32_t log_blowup,
                       const uint32_t *points_mont, size_t npts, uint32_t *out_mont);
/* pg_pcs_open_values_noalloc: the same values as pg_pcs_open_values, bit for
 * bit, computed without allocating device memory.  It has no counterpart in
 * the reference; it is for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It calls
 * no hipMalloc, hipFree or pool allocation.  Instead it works in a static
 * device array of PG_NOALLOC_SCRATCH_BYTES, one per GPU, which each GPU keeps
 * until the process ends.  The array holds the twiddle table (4h bytes,
 * h = trace height), the barycentric weights of two points at a time
 * (32h bytes), the partial column sums and the output staging.  So trace
 * heights reach 2^20 at two points.  A request that does not fit returns
 * PG_E_INVALID with the byte count it needs.  Calls on one GPU share the
 * array and are serialised by a per-device lock held for the whole call.
 * Each call rewrites every part of the array it reads, so nothing is left
 * over from the previous call. */
#define PG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int pg_pcs_open_values_noalloc(pg_ctx *ctx, const pg_tree *t, size_t k, uint32_t log_blowup,
                               const uint32_t *points_mont, size_t npts, uint32_t *out_mont);
int pg_pcs_reduce_openings(pg_ctx *ctx, const pg_tree *const *rounds, size_t nrounds,
                           const uint32_t *npoints, const uint32_t *points_mont,
                           const uint32_t *values_mont, const uint32_t alpha_mont[4],
                           pg_buf **out_by_log);
/* The whole of TwoAdicFriPcs::open (the reference calls it at
 * vm/src/machine/prover.rs:588-589; the verifier replays it at
 * vm/src/compiler/recursion/circuit/fri.rs:36-73,84-99,118-241) in one call:
 * values of every matrix at its points, each observed on `ch` (round ->
 * matrix -> point -> column, 4 words); alpha = ch.sample_ext(); reduced
 * openings per LDE height; the FRI commit phase down to 2^log_blowup (the
 * final vector must be constant and one value is observed) or, with
 * log_final >= 0, down to 2^log_final observing all of it; the PoW grind;
 * num_queries indices of log_max bits; every input opening and every
 * commit-phase opening.  npoints[m] points (<= 2) for matrix m in round ->
 * matrix order, points_mont their ext4 words in the same order.  The
 * challenger ends in the state the verifier's replay reaches. */
int pg_pcs_open(pg_ctx *ctx, const pg_tree *const *rounds, size_t nrounds, const uint32_t *npoints,
                const uint32_t *points_mont, uint32_t log_blowup, int log_final, size_t num_queries,
                int pow_bits, pg_challenger *ch, pg_open_proof **out);
/* Sizes of the arrays pg_open_proof_read fills (any pointer may be null):
 * values n_values words; alpha 4; commits 8 per FRI round; betas 4 per
 * round; final_poly final_words; pow witness 1 (canonical); query_indices
 * num_queries (u64); input_openings input_words_per_query per query (for
 * each round in order: its rows, total-width words, then its path, 8 words
 * per level); fri_siblings 4 per round per query; fri_paths
 * fri_path_words_per_query per query (round 0's path first).  All other
 * words are Montgomery. */
int pg_open_proof_sizes(const pg_open_proof *p, size_t *n_values, int *fri_rounds, size_t *final_words,
                        size_t *num_queries, size_t *input_words_per_query, size_t *fri_path_words_per_query);
int pg_open_proof_read(const pg_open_proof *p, uint32_t *values, uint32_t alpha[4], uint32_t *commits,
                       uint32_t *betas, uint32_t *final_poly, uint32_t *pow_witness, uint64_t *query_indices,
                       uint32_t *input_openings, uint32_t *fri_siblings, uint32_t *fri_paths);
/* The PcsProof part of a pg_open_proof as bytes: bincode 1 (little-endian,
 * fixed-width integers, u64 sequence lengths) of p3-fri's
 * FriProof<Challenge, ChallengeMmcs, Val, Vec<BatchOpening<Val, ValMmcs>>>,
 * which the reference stores in every BaseProof::opening_proof and writes
 * with bincode in MetaProof::save_to_file (vm/src/machine/proof.rs:86-101,
 * :159-171).  Field order as the recursion verifier reads it
 * (vm/src/compiler/recursion/circuit/types.rs:29-54): commit_phase_commits
 * (u64 count, 8 words each); query_proofs (u64 count, then per query:
 * input_proof = u64 rounds, per round u64 matrices, per matrix u64 width +
 * row words, then u64 levels + 8 words per level; commit_phase_openings =
 * u64 rounds, per round 4 sibling words + u64 levels + 8 words per level);
 * final_poly (4 words); pow_witness (1 word).  A word is a u32 field element:
 * the Montgomery value (p3's in-memory form), or the canonical value when
 * canonical != 0.  encode: out may be null to get *len; the final vector must
 * be constant (pg_pcs_open with log_final < 0).  decode: the bytes of one
 * proof, every query of one shape; the result holds the proof parts only (no
 * values, alpha, betas or query indices: the verifier's transcript derives
 * those), readable with pg_open_proof_sizes / pg_open_proof_read and
 * encodable again.  No counterpart beyond the reference's bincode of the same
 * struct. */
int pg_open_proof_encode(const pg_open_proof *p, int canonical, uint8_t *out, size_t cap, size_t *len);
int pg_open_proof_decode(int field, const uint8_t *bytes, size_t len, int canonical, pg_open_proof **out);
void pg_open_proof_free(pg_open_proof *p);

/* ---- quotient evaluation (compute_quotient_values) ------------------------
 * The reference evaluates every chip's constraints on its quotient domain
 * Q = GEN <w_|Q|>, |Q| = n 2^log_quotient_degree, from the main (and
 * preprocessed) trace read back with get_evaluations_on_domain, folding them
 * as acc = acc * alpha + c in assert order and multiplying by 1 / Z_H(x)
 * (vm/src/machine/utils.rs:201-356, called per chip at
 * vm/src/machine/prover.rs:441-508; folder vm/src/machine/folder.rs:222-226;
 * selectors vm/src/compiler/recursion/circuit/domain.rs:46-60:
 * is_first_row = Z_H(x) / (x - 1), is_last_row = Z_H(x) / (x - w_n^-1),
 * is_transition = x - w_n^-1, Z_H(x) = x^n - 1).  Here the constraints cross
 * the boundary as the expression DAG Pico's SymbolicConstraintFolder
 * collects (p3 SymbolicExpression, vm/src/machine/folder.rs:30-120): one
 * pg_air_node per expression node, operands before their users, and the
 * asserted nodes in assert order.  pg_air_compile turns them into a
 * register-allocated program; pg_quotient_values runs it on the device for
 * every row of Q, reading the local row and the row 2^log_quotient_degree
 * steps ahead (the reference's next_step) straight from the committed LDEs
 * (no readback), and writes the |Q| x 4 quotient values (ext4, natural
 * order on Q: exactly compute_quotient_values' output, whose
 * flatten_to_base is the same memory; split it with pg_buf_view for the
 * quotient commit).  Constraints may be base-field (assert_zero) or
 * extension-field (assert_zero_ext): the permutation trace's LogUp
 * constraints (vm/src/machine/permutation.rs:183-346) read its ext4 columns
 * (kind PG_AIR_PERMUTATION: column c is base columns 4c..4c+3 of the
 * committed, flattened permutation trace, vm/src/machine/prover.rs:404-415),
 * the permutation challenges, the regional cumulative sum and the global
 * cumulative sum's 14 coordinates (ProverConstraintFolder, folder.rs:
 * 178-300).  A node is extension-valued when it is one of those leaves or
 * has an extension-valued operand; an asserted extension value folds like
 * assert_zero_ext (acc = acc * alpha + c, c in ext4). */
#define PG_AIR_MAIN 0       /* a = row offset (0 local, 1 next), b = column */
#define PG_AIR_PREPROCESSED 1 /* a = row offset, b = column */
#define PG_AIR_PUBLIC 2     /* b = public value index */
#define PG_AIR_IS_FIRST_ROW 3
#define PG_AIR_IS_LAST_ROW 4
#define PG_AIR_IS_TRANSITION 5
#define PG_AIR_CONSTANT 6   /* a = canonical value */
#define PG_AIR_ADD 7        /* a, b = operand node indices (earlier nodes) */
#define PG_AIR_SUB 8
#define PG_AIR_NEG 9        /* a = operand node index */
#define PG_AIR_MUL 10
#define PG_AIR_PERMUTATION 11   /* ext4: a = row offset, b = ext column */
#define PG_AIR_PERM_CHALLENGE 12 /* ext4: b = challenge index */
#define PG_AIR_REGIONAL_SUM 13  /* ext4: the regional cumulative sum */
#define PG_AIR_GLOBAL_SUM 14    /* b = coordinate 0..13 (x_0..x_6, y_0..y_6) */
typedef struct {
    uint32_t kind, a, b;
} pg_air_node;
typedef struct pg_air_program pg_air_program;
/* Checks the DAG (kinds, operand order, asserted nodes in range) and
 * allocates registers over the emission order: for each assert in turn, the
 * nodes it needs that are not yet computed, then the fold.  A register is
 * reused once its value has no later reader.  nregs / ninstr may be null. */
int pg_air_compile(int field, const pg_air_node *nodes, size_t nnodes, const uint32_t *asserts, size_t nasserts,
                   pg_air_program **out, uint32_t *nregs, uint32_t *ninstr);
void pg_air_program_free(pg_air_program *prog);
/* main: matrix main_k of a committed tree at blowup 2^log_blowup (trace
 * height n = its height >> log_blowup); pre: the preprocessed matrix (may be
 * null when the program reads none), same n.  Requires log_quotient_degree <=
 * log_blowup (Q is then a sub-coset of the committed LDE coset).  publics:
 * Montgomery words.  out: a new |Q| x 4 buffer. */
int pg_quotient_values(pg_ctx *ctx, const pg_air_program *prog, const pg_tree *main, size_t main_k,
                       const pg_tree *pre, size_t pre_k, uint32_t log_blowup, uint32_t log_quotient_degree,
                       const uint32_t *publics_mont, size_t npublics, const uint32_t alpha_mont[4],
                       pg_buf **out);
/* The same with the permutation side of compute_quotient_values
 * (vm/src/machine/utils.rs:201-356): perm is matrix perm_k of the committed
 * permutation trace (flattened to base columns, the same height; may be null
 * when the program reads none), challenges_mont the permutation challenges
 * (nchallenges x 4 words), regional_sum_mont the regional cumulative sum (4
 * words, may be null when unread), global_sum_mont the global cumulative
 * sum's coordinates (nglobal words). */
int pg_quotient_values_perm(pg_ctx *ctx, const pg_air_program *prog, const pg_tree *main, size_t main_k,
                            const pg_tree *pre, size_t pre_k, const pg_tree *perm, size_t perm_k,
                            uint32_t log_blowup, uint32_t log_quotient_degree, const uint32_t *publics_mont,
                            size_t npublics, const uint32_t *challenges_mont, size_t nchallenges,
                            const uint32_t regional_sum_mont[4], const uint32_t *global_sum_mont, size_t nglobal,
                            const uint32_t alpha_mont[4], pg_buf **out);

/* ---- batched commits ------------------------------------------------------
 * Many independent single-matrix Pcs::commit jobs (the reference proves the
 * shards of a batch in parallel rayon workers, vm/src/instances/machine/
 * riscv.rs:235-309) as one launch sequence: each LDE pass once per distinct
 * log-height, one leaf-hash launch, one launch per Merkle level, one root
 * gather.  create allocates every output and uploads all descriptors;
 * use_graph != 0 captures the sequence as a hipGraph.  run executes it (it
 * may run repeatedly while no tree has been taken) and returns njobs x 8 root
 * words.  The input buffers must stay alive until the batch is freed.
 * take_tree hands job j's prover data to the caller (free with pg_tree_free). */
typedef struct pg_batch pg_batch;
int pg_batch_create(pg_ctx *ctx, int field, const pg_buf *const *evals,
                    const uint32_t *domain_shifts_mont, const uint32_t *log_blowups, size_t njobs,
                    int use_graph, pg_batch **out);
int pg_batch_run(pg_ctx *ctx, pg_batch *batch, uint32_t *roots_out);
int pg_batch_take_tree(pg_batch *batch, size_t job, pg_tree **out);
size_t pg_batch_num_launches(const pg_batch *batch);
void pg_batch_free(pg_batch *batch);

/* ---- commit service -------------------------------------------------------
 * The GPU counterpart of the reference's emulator -> bounded channel -> prover
 * pipeline (vm/src/instances/machine/riscv.rs:93-129, shards proven in
 * parallel at :235): shard commit jobs are submitted as they are produced,
 * admitted by HBM footprint (LDE + scratch + digests + trace) against the
 * service budget (0 = the context's budget), run `workers` at a time on
 * their own streams, and keep their LDE and Merkle tree on the device for
 * openings until released.  Jobs submitted while others run are admitted as
 * memory frees and overlap with them on the device.
 * submit takes ownership of `trace`; submit_synthetic fills the trace on the
 * device from a seed (pg_fill_random's stream).  wait blocks until the job
 * is done and returns its root; times (may be null) = submit, start and done
 * times in seconds (steady clock). */
typedef struct pg_service pg_service;
int pg_service_create(pg_ctx *ctx, size_t budget_bytes, int workers, pg_service **out);
int pg_service_submit(pg_service *svc, int field, pg_buf *trace, uint32_t log_blowup, uint64_t *job_id);
int pg_service_submit_synthetic(pg_service *svc, int field, uint64_t seed, size_t rows, size_t cols,
                                uint32_t log_blowup, uint64_t *job_id);
/* A trace in host memory, rows x cols words (Montgomery form, or canonical
 * values when canonical != 0), as the reference's emulator hands a shard's
 * records to the prover over its channel (riscv.rs:93-129; the traces are
 * generated on the host, vm/src/machine/prover.rs:169-223).  The job runs
 * pg_pcs_commit_host: the upload is part of the job and overlaps its LDE and
 * hashing.  The caller keeps `host` unchanged until pg_service_wait returns
 * for the job (or it is cancelled). */
int pg_service_submit_host(pg_service *svc, int field, const uint32_t *host, size_t rows, size_t cols,
                           int canonical, uint32_t log_blowup, uint64_t *job_id);
int pg_service_wait(pg_service *svc, uint64_t job_id, uint32_t root[8], double times[3]);
int pg_service_open(pg_service *svc, uint64_t job_id, const uint64_t *indices, size_t n_idx,
                    uint32_t *rows_out, uint32_t *paths_out);
int pg_service_release(pg_service *svc, uint64_t job_id);
/* Takes a job that has not started off the queue: it never runs, its trace
 * is freed and pg_service_wait returns PG_E_CANCELLED for it (release it as
 * any finished job).  A running or finished job: PG_E_INVALID.  No
 * counterpart in the reference (it has no job queue); for a prover that
 * abandons a shard. */
int pg_service_cancel(pg_service *svc, uint64_t job_id);
int pg_service_stats(pg_service *svc, size_t *reserved_bytes, size_t *pending_jobs, size_t *peak_running);
void pg_service_destroy(pg_service *svc);

/* ---- FRI ----------------------------------------------------------------- */
/* out = fold_even_odd(v, beta): v is n x 4 (ext4, bit-reversed order), out is
 * (n/2) x 4. */
int pg_fri_fold(pg_ctx *ctx, int field, const pg_buf *v_ext4, const uint32_t beta_mont[4],
                pg_buf **out);
/* Commit phase.  inputs[k] (k <= max_log, may be NULL) is the reduced-opening
 * vector of length 2^k (ext4 rows); the largest one starts the fold.  Loop
 * while the length exceeds 2^log_stop: commit rows of 2 ext4 elements,
 * observe the root, sample beta, fold, add inputs of the new length.
 * final_mode 0: the final vector must be constant and one element is
 * observed (reference behaviour); 1: every final element is observed.
 * roots_out: rounds * 8; final_out: 2^log_stop * 4 words.  Returns the number
 * of rounds through *rounds_out. */
int pg_fri_commit_phase(pg_ctx *ctx, int field, pg_challenger *ch, const pg_buf *const *inputs,
                        int max_log, int log_stop, int final_mode, uint32_t *roots_out,
                        uint32_t *betas_out, uint32_t *final_out, int *rounds_out,
                        pg_fri **out);
/* PoW: find the smallest witness w with observe(w); sample_bits(bits) == 0,
 * then apply exactly that observe + sample to *ch. */
int pg_fri_grind(pg_ctx *ctx, int field, pg_challenger *ch, int bits, uint32_t *witness_out);
/* For each query index q (log_max bits) and round r: the sibling value
 * v_r[(q >> r) ^ 1] (4 words) and the Merkle path of leaf (q >> r) >> 1 in
 * round r's tree (log_max - 1 - r digests).  siblings_out: nq * rounds * 4;
 * paths_out: nq * sum_r (log_max - 1 - r) * 8, query-major, round-major. */
int pg_fri_answer_queries(pg_ctx *ctx, const pg_fri *fri, const uint64_t *indices, size_t nq,
                          uint32_t *siblings_out, uint32_t *paths_out);
int pg_fri_num_rounds(const pg_fri *fri);
const pg_tree *pg_fri_round_tree(const pg_fri *fri, int round);
void pg_fri_free(pg_fri *fri);

/* ---- DuplexChallenger (host) --------------------------------------------- */
void pg_challenger_init(pg_challenger *ch, int field);
void pg_challenger_observe(pg_challenger *ch, uint32_t value_mont);
uint32_t pg_challenger_sample(pg_challenger *ch);
void pg_challenger_sample_ext(pg_challenger *ch, uint32_t out_mont[4]);
/* low `bits` bits of the canonical value of one sample */
uint32_t pg_challenger_sample_bits(pg_challenger *ch, int bits);
/* observe(witness) then sample_bits(bits) == 0; returns 1 if it holds */
int pg_challenger_check_witness(pg_challenger *ch, int bits, uint32_t witness_canonical);

#ifdef __cplusplus
}
#endif
#endif /* PICO_GPU_H */
