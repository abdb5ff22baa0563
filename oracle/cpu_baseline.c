/*
 * cpu_baseline.c -- the reference's CPU commit path restated for speed, timed
 * by bench.py as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/: loaded by tests/ and by
 * bench.py's cpu_baseline leg; nothing in the product links it.
 *
 * What it restates: TwoAdicFriPcs::commit of one trace
 * (vm/src/machine/prover.rs:238-244) with the reference's CPU components,
 * p3 Radix2DitParallel::coset_lde_batch + bit_reverse_rows and
 * MerkleTreeMmcs<PaddingFreeSponge<Poseidon2, 16, 8, 8>,
 * TruncatedPermutation<Poseidon2, 2, 8, 16>> (vm/src/configs/stark_config/
 * kb_poseidon2.rs:15-27), in the form the reference runs them: values in
 * Montgomery form (p3 MontyField31), row-parallel radix-2 butterflies
 * vectorised across the columns of a row-major matrix, and Poseidon2 on
 * packed lanes (the reference builds with `nightly-features`, i.e. p3's
 * packed AVX fields, vm/Cargo.toml:117-121).  Here: AVX-512 (16 lanes) where
 * the host has it, else AVX2 (8 lanes), chosen at run time; OpenMP over
 * rows.  The vector code is cpu_baseline_simd.inc, included once per
 * instruction set.  Unlike oracle.c (canonical "% p" arithmetic, the checker), it
 * is written for speed; tests/test_oracle.py pins its roots and LDEs to
 * oracle.c's.
 *
 * LDE: inverse DIF transform of the h rows (natural in, bit-reversed out),
 * rows swapped into natural order while scaled by shift^i / h, zero rows up
 * to N = h 2^b, forward DIF (natural in, bit-reversed out: the committed row
 * order, row j = f(shift w_N^rev(j))).  A DIF layer whose butterflies span
 * fewer than CB_BLOCK_ROWS rows runs block by block with all the remaining
 * layers of a block in cache; wider layers are whole-matrix passes.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define KB_P 2130706433u
#define BB_P 2013265921u
#define INL static inline __attribute__((always_inline))

static const uint32_t CB_RC16_RAW[30][16] = {
#include "rc16.inc"
};

/* ------------------------------------------------------- scalar helpers -- */
static uint32_t cb_pow(uint32_t p, uint32_t a, uint64_t e) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
static void *alloc64(size_t bytes) { return aligned_alloc(64, (bytes + 63) / 64 * 64); }
static uint32_t mont_of(uint32_t p, uint32_t x) { return (uint32_t)(((uint64_t)(x % p) << 32) % p); }

typedef struct {
    uint32_t p, pinv; /* pinv = p^-1 mod 2^32 */
    int rounds_p, sbox7;
    uint32_t rc_full[8][16], rc_part[20], diag[16]; /* Montgomery */
} CbField;

static void cb_field_init(int f, CbField *k) {
    const uint32_t p = f == OR_KOALABEAR ? KB_P : BB_P;
    k->p = p;
    uint32_t x = p;
    for (int i = 0; i < 5; i++) x *= 2u - p * x; /* Newton: x = p^-1 mod 2^32 */
    k->pinv = x;
    k->rounds_p = f == OR_KOALABEAR ? 20 : 13;
    k->sbox7 = f != OR_KOALABEAR;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < 16; i++) {
            k->rc_full[r][i] = mont_of(p, CB_RC16_RAW[r][i]);
            k->rc_full[4 + r][i] = mont_of(p, CB_RC16_RAW[4 + k->rounds_p + r][i]);
        }
    for (int r = 0; r < k->rounds_p; r++) k->rc_part[r] = mont_of(p, CB_RC16_RAW[4 + r][0]);
    /* internal diagonals (vm/src/primitives/consts.rs:110-154) */
    const uint32_t kb[16] = {p - 2, 1, 2, (p + 1) >> 1, 3, 4, (p - 1) >> 1, p - 3, p - 4, p - ((p - 1) >> 8),
                             p - ((p - 1) >> 3), p - 127, (p - 1) >> 8, (p - 1) >> 3, (p - 1) >> 4, 127};
    const uint32_t bb[16] = {p - 2, 1, 2, (p + 1) >> 1, 3, 4, (p - 1) >> 1, p - 3, p - 4, p - ((p - 1) >> 8),
                             p - ((p - 1) >> 2), p - ((p - 1) >> 3), p - 15, (p - 1) >> 8, (p - 1) >> 4, 15};
    for (int i = 0; i < 16; i++) k->diag[i] = mont_of(p, f == OR_KOALABEAR ? kb[i] : bb[i]);
}

/* ------------------------------------------------------ shared helpers -- */
#define CB_BLOCK_ROWS 512

static size_t rev_bits(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

/* twiddles w^e (Montgomery), e < n/2, w = the order-n root (or its inverse) */
static uint32_t *make_twiddles(uint32_t p, uint32_t w, size_t n) {
    const size_t half = n / 2 ? n / 2 : 1;
    uint32_t *t = alloc64(half * 4);
    if (!t) return NULL;
    uint64_t cur = 1;
    for (size_t e = 0; e < half; e++) {
        t[e] = mont_of(p, (uint32_t)cur);
        cur = cur * w % p;
    }
    return t;
}

static double g_last_lde_s, g_last_merkle_s;

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ---------------------------------------------------------- AVX2, 8 lanes -- */
#define VT __m256i
#define VL 8
#define SFX _avx2
#define V_SET1(x) _mm256_set1_epi32((int)(x))
#define V_ADD _mm256_add_epi32
#define V_SUB _mm256_sub_epi32
#define V_MIN _mm256_min_epu32
#define V_MULE _mm256_mul_epu32
#define V_AND _mm256_and_si256
#define V_SRLI64(a, n) _mm256_srli_epi64(a, n)
#define V_SRLI32(a, n) _mm256_srli_epi32(a, n)
#define V_BLEND_ODD(a, b) _mm256_blend_epi32(a, b, 0xAA) /* odd lanes from b */
#define V_ZERO() _mm256_setzero_si256()
#define V_LOAD(p) _mm256_load_si256((const __m256i *)(p))
#define V_STORE(p, x) _mm256_store_si256((__m256i *)(p), x)

/* 8 x 8 transpose of 32-bit words: in[r] = row r; out[c] = column c. */
INL void transpose8(__m256i m[8]) {
    const __m256i a0 = _mm256_unpacklo_epi32(m[0], m[1]), a1 = _mm256_unpackhi_epi32(m[0], m[1]);
    const __m256i a2 = _mm256_unpacklo_epi32(m[2], m[3]), a3 = _mm256_unpackhi_epi32(m[2], m[3]);
    const __m256i a4 = _mm256_unpacklo_epi32(m[4], m[5]), a5 = _mm256_unpackhi_epi32(m[4], m[5]);
    const __m256i a6 = _mm256_unpacklo_epi32(m[6], m[7]), a7 = _mm256_unpackhi_epi32(m[6], m[7]);
    const __m256i b0 = _mm256_unpacklo_epi64(a0, a2), b1 = _mm256_unpackhi_epi64(a0, a2);
    const __m256i b2 = _mm256_unpacklo_epi64(a1, a3), b3 = _mm256_unpackhi_epi64(a1, a3);
    const __m256i b4 = _mm256_unpacklo_epi64(a4, a6), b5 = _mm256_unpackhi_epi64(a4, a6);
    const __m256i b6 = _mm256_unpacklo_epi64(a5, a7), b7 = _mm256_unpackhi_epi64(a5, a7);
    m[0] = _mm256_permute2x128_si256(b0, b4, 0x20);
    m[1] = _mm256_permute2x128_si256(b1, b5, 0x20);
    m[2] = _mm256_permute2x128_si256(b2, b6, 0x20);
    m[3] = _mm256_permute2x128_si256(b3, b7, 0x20);
    m[4] = _mm256_permute2x128_si256(b0, b4, 0x31);
    m[5] = _mm256_permute2x128_si256(b1, b5, 0x31);
    m[6] = _mm256_permute2x128_si256(b2, b6, 0x31);
    m[7] = _mm256_permute2x128_si256(b3, b7, 0x31);
}

/* Load words [c0, c0 + len) of rows r[0..8) (len <= 8) as 8 lane vectors
 * (vectors past len are zero). */
INL void load_cols_avx2(const uint32_t *const r[8], size_t c0, size_t len, __m256i out[8]) {
    if (len == 8) {
        for (int i = 0; i < 8; i++) out[i] = _mm256_loadu_si256((const __m256i *)(r[i] + c0));
    } else {
        uint32_t tmp[8][8] __attribute__((aligned(32)));
        memset(tmp, 0, sizeof tmp);
        for (int i = 0; i < 8; i++) memcpy(tmp[i], r[i] + c0, len * 4);
        for (int i = 0; i < 8; i++) out[i] = _mm256_load_si256((const __m256i *)tmp[i]);
    }
    transpose8(out);
}

/* digest words s[0..8) of lane l to d[l] (lanes past nvalid go to the
 * caller's spill rows) */
INL void store_digests_avx2(__m256i s[8], uint32_t *const d[8], int nvalid) {
    (void)nvalid;
    __m256i m[8];
    for (int i = 0; i < 8; i++) m[i] = s[i];
    transpose8(m);
    for (int i = 0; i < 8; i++) _mm256_storeu_si256((__m256i *)d[i], m[i]);
}

#include "cpu_baseline_simd.inc"

#undef VT
#undef VL
#undef SFX
#undef V_SET1
#undef V_ADD
#undef V_SUB
#undef V_MIN
#undef V_MULE
#undef V_AND
#undef V_SRLI64
#undef V_SRLI32
#undef V_BLEND_ODD
#undef V_ZERO
#undef V_LOAD
#undef V_STORE

/* ------------------------------------------------------ AVX-512, 16 lanes -- */
#pragma GCC push_options
#pragma GCC target("avx512f,avx512bw,avx512dq,avx512vl")
#define VT __m512i
#define VL 16
#define SFX _avx512
#define V_SET1(x) _mm512_set1_epi32((int)(x))
#define V_ADD _mm512_add_epi32
#define V_SUB _mm512_sub_epi32
#define V_MIN _mm512_min_epu32
#define V_MULE _mm512_mul_epu32
#define V_AND _mm512_and_si512
#define V_SRLI64(a, n) _mm512_srli_epi64(a, n)
#define V_SRLI32(a, n) _mm512_srli_epi32(a, n)
#define V_BLEND_ODD(a, b) _mm512_mask_blend_epi32((__mmask16)0xAAAA, a, b) /* odd lanes from b */
#define V_ZERO() _mm512_setzero_si512()
#define V_LOAD(p) _mm512_load_si512((const void *)(p))
#define V_STORE(p, x) _mm512_store_si512((void *)(p), x)

/* Load words [c0, c0 + len) of rows r[0..16) (len <= 8) as lane vectors:
 * one gather per word, the rows addressed relative to r[0] (the callers'
 * rows lie within one array, at or after r[0]); vectors past len are zero. */
INL void load_cols_avx512(const uint32_t *const r[16], size_t c0, size_t len, __m512i out[8]) {
    int32_t off[16] __attribute__((aligned(64)));
    for (int l = 0; l < 16; l++) off[l] = (int32_t)(r[l] - r[0]);
    const __m512i vi = _mm512_load_si512((const void *)off);
    for (size_t i = 0; i < 8; i++)
        out[i] = i < len ? _mm512_i32gather_epi32(vi, (const void *)(r[0] + c0 + i), 4) : _mm512_setzero_si512();
}

/* digest words s[0..8) of lane l to d[l] for the first nvalid lanes (one
 * masked scatter per word; the valid destinations are consecutive digests) */
INL void store_digests_avx512(__m512i s[8], uint32_t *const d[16], int nvalid) {
    int32_t off[16] __attribute__((aligned(64)));
    for (int l = 0; l < 16; l++) off[l] = l < nvalid ? (int32_t)(d[l] - d[0]) : 0;
    const __m512i vi = _mm512_load_si512((const void *)off);
    const __mmask16 mk = (__mmask16)((1u << nvalid) - 1u);
    for (int i = 0; i < 8; i++) _mm512_mask_i32scatter_epi32((void *)(d[0] + i), mk, vi, s[i], 4);
}

#include "cpu_baseline_simd.inc"
#pragma GCC pop_options

/* ---------------------------------------------------------------- entry -- */
/* seconds of the last cb_pcs_commit's LDE (input copy through the forward
 * transform) and Merkle tree (leaves and every compress layer) */
void cb_last_phase_seconds(double *lde_s, double *merkle_s) {
    *lde_s = g_last_lde_s;
    *merkle_s = g_last_merkle_s;
}

/* 2 when this host runs the AVX-512 code, else 1 (AVX2) */
int cb_isa(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                   __builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl")
               ? 2
               : 1;
}

/* isa 0: the best this host has; 1: AVX2; 2: AVX-512 (-1 if the host lacks it) */
int cb_pcs_commit_isa(int isa, int f, const uint32_t *in_mont, size_t h, size_t w, int lb, uint32_t root_mont[8],
                      uint32_t *lde_out) {
    if (h == 0 || (h & (h - 1)) || w == 0 || lb < 0 || lb > 8) return -1;
    const int best = cb_isa();
    if (isa == 0) isa = best;
    if (isa == 2 && best < 2) return -1;
    return isa == 2 ? commit_avx512(f, in_mont, h, w, lb, root_mont, lde_out)
                    : commit_avx2(f, in_mont, h, w, lb, root_mont, lde_out);
}

int cb_pcs_commit(int f, const uint32_t *in_mont, size_t h, size_t w, int lb, uint32_t root_mont[8],
                  uint32_t *lde_out) {
    return cb_pcs_commit_isa(0, f, in_mont, h, w, lb, root_mont, lde_out);
}

/* canonical -> Montgomery, in parallel (input preparation, outside timing) */
void cb_to_monty(int f, const uint32_t *in, uint32_t *out, size_t n) {
    const uint32_t p = f == OR_KOALABEAR ? KB_P : BB_P;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) out[i] = mont_of(p, in[i]);
}

int cb_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
