/*
 * oracle.h -- CPU oracle for the Pico STARK commitment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against.  It may be loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py; no product
 * code path may link, load or call it.
 *
 * Every value crossing this interface is a CANONICAL field element (u32 in
 * [0, p)).  Arithmetic is plain u64 "% p" reduction: deliberately a different
 * representation from the product's Montgomery-form kernels, so that a
 * Montgomery mistake cannot be shared by both sides.
 *
 * What each function restates (reference file:line, paths relative to the
 * reference repository root; "p3" = the un-vendored Plonky3 fork
 * brevis-network/Plonky3@f5056c9, whose semantics are pinned by the in-tree
 * restatements cited):
 *   or_poseidon2          gnark/poseidon2/poseidon2_koalabear.go:71-196,
 *                         gnark/poseidon2/poseidon2_babybear.go:71-185,
 *                         vm/src/primitives/mod.rs:305-346 (round split),
 *                         vm/src/primitives/consts.rs:107-154,185-203
 *   or_sponge             vm/src/compiler/recursion/circuit/hash.rs:40-52
 *   or_compress           vm/src/compiler/recursion/circuit/hash.rs:89-94
 *   or_coset_lde          p3 Radix2DitParallel::coset_lde_batch + bit_reverse_rows
 *                         (row order pinned by circuit/fri.rs:151-171)
 *   or_mmcs_*             p3 MerkleTreeMmcs (pinned by circuit/fri.rs:389-436)
 *   or_fri_fold           p3 fold_even_odd (pinned by circuit/fri.rs:288-387)
 *   or_ch_*               p3 DuplexChallenger (circuit/challenger.rs:57-227)
 *   or_pcs_open_values    p3 TwoAdicFriPcs::open, evaluation at the opening points
 *   or_pcs_reduce         p3 TwoAdicFriPcs::open, reduced openings (the verifier's
 *                         loop, circuit/fri.rs:118-241, applied to every row)
 *   or_pcs_commit_seeded  p3 TwoAdicFriPcs::commit (vm/src/machine/prover.rs:238-244)
 *                         = or_coset_lde + or_mmcs_commit, without holding LDEs
 *   cb_pcs_commit         the same commit for one trace, written for speed
 *                         (cpu_baseline.c): the CPU baseline, not a checker
 */
#ifndef PICO_ORACLE_H
#define PICO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field ids (same numbering as the product ABI) */
#define OR_KOALABEAR 0
#define OR_BABYBEAR 1

uint32_t or_prime(int field);
uint32_t or_add(int field, uint32_t a, uint32_t b);
uint32_t or_sub(int field, uint32_t a, uint32_t b);
uint32_t or_mul(int field, uint32_t a, uint32_t b);
uint32_t or_pow(int field, uint32_t a, uint64_t e);
uint32_t or_inv(int field, uint32_t a);
uint32_t or_generator(int field);
uint32_t or_two_adic_generator(int field, int log_n);
int or_two_adicity(int field);
/* canonical <-> Montgomery (R = 2^32) conversions, for boundary tests */
uint32_t or_to_monty(int field, uint32_t a);
uint32_t or_from_monty(int field, uint32_t a);

/* degree-4 binomial extension x^4 = W (KoalaBear W=3, BabyBear W=11) */
void or_ext_mul(int field, const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void or_ext_inv(int field, const uint32_t a[4], uint32_t out[4]);

/* Poseidon2 width 16 permutation, in place */
void or_poseidon2(int field, uint32_t state[16]);
/* PaddingFreeSponge<16, rate 8, out 8> */
void or_sponge(int field, const uint32_t *in, size_t n, uint32_t out[8]);
/* TruncatedPermutation<2, 8, 16> */
void or_compress(int field, const uint32_t left[8], const uint32_t right[8], uint32_t out[8]);

/* Seeded synthetic inputs: element i of stream `seed` (see or_fill_random). */
void or_fill_random(int field, uint64_t seed, uint32_t *out, size_t n);

/* Coset LDE of an h x w row-major matrix (values on <w_h>) to N = h << added_bits
 * rows: row i of `out` is f(shift * w_N^{rev_N(i)}) when bit_reverse != 0, else
 * row j is f(shift * w_N^j).  Textbook radix-2 per column, OpenMP over columns. */
void or_coset_lde(int field, const uint32_t *in, size_t h, size_t w, int added_bits,
                  uint32_t shift, int bit_reverse, uint32_t *out);
/* Same result by direct O(h*N) Lagrange-free evaluation (tiny sizes only). */
void or_pcs_open_values(int field, const uint32_t *lde, size_t n, size_t w, int log_blowup,
                        const uint32_t z[4], uint32_t *out);
void or_pcs_reduce(int field, size_t nmats, const uint32_t *const *mats, const size_t *heights,
                   const size_t *widths, const uint32_t *npts, const uint32_t *points,
                   const uint32_t *values, const uint32_t alpha[4], uint32_t *const *out);
void or_coset_lde_naive(int field, const uint32_t *in, size_t h, size_t w, int added_bits,
                        uint32_t shift, int bit_reverse, uint32_t *out);

/* ---- mixed-height MMCS (MerkleTreeMmcs) ----
 * Matrices are row-major; heights must be powers of two.  The digest store
 * holds every layer, layer 0 (H = max height digests) first, then H/2, ..., 1;
 * (2H-1)*8 u32 in total.  */
typedef struct {
    const uint32_t *data;
    size_t height;
    size_t width;
} or_matrix;

size_t or_mmcs_digest_words(const or_matrix *mats, size_t n);
int or_mmcs_commit(int field, const or_matrix *mats, size_t n, uint32_t *digests, uint32_t root[8]);
/* rows_out receives every matrix's opened row, concatenated in original order;
 * path_out receives log2(H) sibling digests. */
int or_mmcs_open(const or_matrix *mats, size_t n, const uint32_t *digests, size_t index,
                 uint32_t *rows_out, uint32_t *path_out);
/* returns 1 if the opening re-derives `root` */
int or_mmcs_verify(int field, const size_t *heights, const size_t *widths, size_t n,
                   const uint32_t root[8], size_t index, const uint32_t *rows,
                   const uint32_t *path);

/* TwoAdicFriPcs::commit root of seeded matrices: matrix k is or_fill_random(
 * seeds[k]) of heights[k] x widths[k] on the domain domain_shifts[k] <w_h>
 * (canonical), LDE shift GEN / domain_shifts[k], one mixed-height tree.  LDEs
 * are produced block_cols columns at a time and absorbed into per-row sponge
 * states, so no LDE is held whole.  Returns 0, or -1 on bad shapes. */
int or_pcs_commit_seeded(int field, size_t n, const uint64_t *seeds, const size_t *heights,
                         const size_t *widths, const uint32_t *domain_shifts, int log_blowup,
                         size_t block_cols, uint32_t root[8]);

/* ---- FRI ---- */
/* v: n ext4 elements (bit-reversed evaluation order), n even; out: n/2 */
void or_fri_fold(int field, const uint32_t *v, size_t n, const uint32_t beta[4], uint32_t *out);

/* DuplexChallenger<F, Poseidon2, 16, 8> */
typedef struct {
    int field;
    uint32_t state[16];
    uint32_t in[8];
    int n_in;
    uint32_t out[8];
    int n_out;
} or_challenger;

void or_ch_init(or_challenger *c, int field);
void or_ch_observe(or_challenger *c, uint32_t v);
uint32_t or_ch_sample(or_challenger *c);
void or_ch_sample_ext(or_challenger *c, uint32_t out[4]);
uint32_t or_ch_sample_bits(or_challenger *c, int bits);
int or_ch_check_witness(or_challenger *c, int bits, uint32_t witness);

/* FRI commit phase over one ext4 vector of length 2^log_n (bit-reversed order),
 * folding while the length exceeds 2^log_stop.  `inputs[k]` (may be NULL) is an
 * extra reduced-opening vector of length 2^k added after the fold that reaches
 * that length.  Writes, per round r: root r (roots[8r..]), the committed vector
 * of that round into layers (concatenated, lengths 2^log_n, 2^(log_n-1), ...),
 * the final vector (2^log_stop ext4) and the betas.  Returns the round count. */
int or_fri_commit_phase(or_challenger *c, const uint32_t *v, int log_n, int log_stop,
                        const uint32_t *const *inputs, uint32_t *roots, uint32_t *layers,
                        uint32_t *final_vec, uint32_t *betas);

/* OpenMP thread count actually used by the parallel loops */
int or_num_threads(void);

/* ---- cpu_baseline.c: the CPU baseline bench.py times ---------------------
 * TwoAdicFriPcs::commit of one h x w trace (vm/src/machine/prover.rs:238-244)
 * as the reference's CPU path computes it: Radix2DitParallel coset LDE at
 * blowup 2^log_blowup with shift GENERATOR, rows bit-reversed, MerkleTreeMmcs
 * root.  Unlike the functions above, it takes and returns MONTGOMERY words
 * (p3's in-memory form) and is written for speed (AVX-512 or AVX2, OpenMP).  lde_out
 * (N x w, committed order) may be NULL.  Returns 0, or -1 on a bad shape or
 * a failed allocation.  Pinned to or_coset_lde + or_mmcs_commit by
 * tests/test_oracle.py. */
int cb_pcs_commit(int field, const uint32_t *in_mont, size_t h, size_t w, int log_blowup,
                  uint32_t root_mont[8], uint32_t *lde_out);
/* the same with the instruction set chosen: 0 the best the host has, 1 AVX2,
 * 2 AVX-512 (-1 if the host lacks it); cb_isa(): what 0 picks (1 or 2) */
int cb_pcs_commit_isa(int isa, int field, const uint32_t *in_mont, size_t h, size_t w, int log_blowup,
                      uint32_t root_mont[8], uint32_t *lde_out);
int cb_isa(void);
void cb_to_monty(int field, const uint32_t *in, uint32_t *out, size_t n);
int cb_num_threads(void);
/* LDE and Merkle seconds of the last cb_pcs_commit */
void cb_last_phase_seconds(double *lde_s, double *merkle_s);

#ifdef __cplusplus
}
#endif
#endif
