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
 * packed AVX fields, vm/Cargo.toml:117-121).  Here: AVX2, 8 lanes, OpenMP
 * over rows.  Unlike oracle.c (canonical "% p" arithmetic, the checker), it
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

/* -------------------------------------------------------- AVX2 field ops -- */
typedef struct {
    __m256i p, pinv;
} V;

INL __m256i vadd(const V *v, __m256i a, __m256i b) {
    const __m256i t = _mm256_add_epi32(a, b);
    return _mm256_min_epu32(t, _mm256_sub_epi32(t, v->p));
}
INL __m256i vsub(const V *v, __m256i a, __m256i b) {
    const __m256i t = _mm256_sub_epi32(a, b);
    return _mm256_min_epu32(t, _mm256_add_epi32(t, v->p));
}
/* Montgomery product a b 2^-32 for a, b < p: with q = lo(a b) p^-1 mod 2^32,
 * a b - q p vanishes mod 2^32 and (a b - q p) / 2^32 = hi(a b) - hi(q p) lies
 * in (-p, p); one conditional addition of p ends it. */
INL __m256i vmul(const V *v, __m256i a, __m256i b) {
    const __m256i ao = _mm256_srli_epi64(a, 32), bo = _mm256_srli_epi64(b, 32);
    const __m256i te = _mm256_mul_epu32(a, b), to = _mm256_mul_epu32(ao, bo);
    const __m256i qe = _mm256_mul_epu32(te, v->pinv), qo = _mm256_mul_epu32(to, v->pinv);
    const __m256i qpe = _mm256_mul_epu32(qe, v->p), qpo = _mm256_mul_epu32(qo, v->p);
    const __m256i ht = _mm256_blend_epi32(_mm256_srli_epi64(te, 32), to, 0xAA);
    const __m256i hq = _mm256_blend_epi32(_mm256_srli_epi64(qpe, 32), qpo, 0xAA);
    const __m256i d = _mm256_sub_epi32(ht, hq);
    return _mm256_min_epu32(d, _mm256_add_epi32(d, v->p));
}
INL __m256i vset(uint32_t x) { return _mm256_set1_epi32((int)x); }
/* x / 2 for x < p: (x >> 1) + (x odd) (p + 1) / 2 */
INL __m256i v_halve(const V *v, __m256i x) {
    const __m256i one = _mm256_set1_epi32(1);
    const __m256i odd = _mm256_sub_epi32(_mm256_setzero_si256(), _mm256_and_si256(x, one));
    const __m256i half_p1 = _mm256_srli_epi32(_mm256_add_epi32(v->p, one), 1);
    return _mm256_add_epi32(_mm256_srli_epi32(x, 1), _mm256_and_si256(odd, half_p1));
}

/* ------------------------------------------------------ Poseidon2, 8 lanes -- */
/* Lane l of s[i] is element i of permutation l. */
INL void v_ext_layer(const V *v, __m256i s[16]) {
    /* M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] per 4-block, written as
     * y_i = (a+b+c+d) + x_i + x_{i+1} + x_{i+1} (indices mod 4): row i of M4
     * is 1 everywhere plus 1 at column i and 2 at column i+1. */
    for (int b = 0; b < 16; b += 4) {
        const __m256i x0 = s[b], x1 = s[b + 1], x2 = s[b + 2], x3 = s[b + 3];
        const __m256i t = vadd(v, vadd(v, x0, x1), vadd(v, x2, x3));
        s[b] = vadd(v, t, vadd(v, x0, vadd(v, x1, x1)));
        s[b + 1] = vadd(v, t, vadd(v, x1, vadd(v, x2, x2)));
        s[b + 2] = vadd(v, t, vadd(v, x2, vadd(v, x3, x3)));
        s[b + 3] = vadd(v, t, vadd(v, x3, vadd(v, x0, x0)));
    }
    for (int i = 0; i < 4; i++) {
        const __m256i col = vadd(v, vadd(v, s[i], s[4 + i]), vadd(v, s[8 + i], s[12 + i]));
        for (int b = 0; b < 16; b += 4) s[b + i] = vadd(v, s[b + i], col);
    }
}

INL __m256i v_sbox(const V *v, int sbox7, __m256i x) {
    const __m256i x2 = vmul(v, x, x);
    if (!sbox7) return vmul(v, x2, x);
    const __m256i x3 = vmul(v, x2, x), x4 = vmul(v, x2, x2);
    return vmul(v, x4, x3);
}

static void v_permute(const V *v, const CbField *k, const __m256i diag[16], __m256i s[16]) {
    v_ext_layer(v, s);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 16; i++) s[i] = v_sbox(v, k->sbox7, vadd(v, s[i], vset(k->rc_full[r][i])));
        v_ext_layer(v, s);
    }
    for (int r = 0; r < k->rounds_p; r++) {
        s[0] = v_sbox(v, k->sbox7, vadd(v, s[0], vset(k->rc_part[r])));
        __m256i sum = s[0];
        for (int i = 1; i < 16; i++) sum = vadd(v, sum, s[i]);
        /* s_i D_i + sum.  Both fields' first nine diagonal entries are
         * -2, 1, 2, 1/2, 3, 4, -1/2, -3, -4 (consts.rs:110-154): additions
         * and a halving (linear maps, so valid on Montgomery values); the
         * other entries are Montgomery products. */
        const __m256i x2 = vadd(v, s[2], s[2]);
        const __m256i x4a = vadd(v, s[4], s[4]), x5a = vadd(v, s[5], s[5]);
        const __m256i x7a = vadd(v, s[7], s[7]), x8a = vadd(v, s[8], s[8]);
        s[0] = vsub(v, sum, vadd(v, s[0], s[0]));
        s[1] = vadd(v, s[1], sum);
        s[2] = vadd(v, x2, sum);
        s[3] = vadd(v, v_halve(v, s[3]), sum);
        s[4] = vadd(v, vadd(v, x4a, s[4]), sum);
        s[5] = vadd(v, vadd(v, x5a, x5a), sum);
        s[6] = vsub(v, sum, v_halve(v, s[6]));
        s[7] = vsub(v, sum, vadd(v, x7a, s[7]));
        s[8] = vsub(v, sum, vadd(v, x8a, x8a));
        for (int i = 9; i < 16; i++) s[i] = vadd(v, vmul(v, s[i], diag[i]), sum);
    }
    for (int r = 4; r < 8; r++) {
        for (int i = 0; i < 16; i++) s[i] = v_sbox(v, k->sbox7, vadd(v, s[i], vset(k->rc_full[r][i])));
        v_ext_layer(v, s);
    }
}

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

/* Load words [c0, c0 + len) of rows r[0..8) (len <= 8) as 8 lane vectors. */
INL void load_cols(const uint32_t *const r[8], size_t c0, size_t len, __m256i out[8]) {
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

INL void store_digests(__m256i s[8], uint32_t *const d[8]) {
    __m256i m[8];
    for (int i = 0; i < 8; i++) m[i] = s[i];
    transpose8(m);
    for (int i = 0; i < 8; i++) _mm256_storeu_si256((__m256i *)d[i], m[i]);
}

/* ------------------------------------------------------------------ LDE -- */
#define CB_BLOCK_ROWS 512

/* One DIF layer of half-width m on rows [0, n) of a (n x wp) matrix:
 * (x_a, x_b) -> (x_a + x_b, (x_a - x_b) tw[j * step]) for a = blk + j,
 * b = a + m.  tw holds w^e for e < n / 2 (w of order n). */
INL void dif_pair(const V *v, uint32_t *a, uint32_t *b, size_t wp, __m256i t) {
    for (size_t c = 0; c < wp; c += 8) {
        const __m256i x = _mm256_load_si256((const __m256i *)(a + c));
        const __m256i y = _mm256_load_si256((const __m256i *)(b + c));
        _mm256_store_si256((__m256i *)(a + c), vadd(v, x, y));
        _mm256_store_si256((__m256i *)(b + c), vmul(v, vsub(v, x, y), t));
    }
}

/* Two DIF layers (half-widths m and m/2) in one pass over the rows: rows
 * r, r + m/2, r + m, r + 3m/2 (r = blk + j, j < m/2) take the layer-m
 * butterflies (r, r + m) at w^j and (r + m/2, r + 3m/2) at w^(j + m/2), then
 * the layer-m/2 butterflies (r, r + m/2) and (r + m, r + 3m/2) at w'^j. */
static void dif_quad_pass(const V *v, uint32_t *x, size_t n, size_t m, size_t wp, const uint32_t *tw) {
    const size_t q = m / 2, s1 = n / (2 * m), s2 = n / m;
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < n / 4; t++) {
        const size_t j = t % q, r = (t / q) * 2 * m + j;
        uint32_t *a = x + r * wp, *b = x + (r + q) * wp, *c = x + (r + m) * wp, *d = x + (r + m + q) * wp;
        const __m256i t1 = vset(tw[j * s1]), t2 = vset(tw[(j + q) * s1]), t3 = vset(tw[j * s2]);
        for (size_t k = 0; k < wp; k += 8) {
            const __m256i va = _mm256_load_si256((const __m256i *)(a + k));
            const __m256i vb = _mm256_load_si256((const __m256i *)(b + k));
            const __m256i vc = _mm256_load_si256((const __m256i *)(c + k));
            const __m256i vd = _mm256_load_si256((const __m256i *)(d + k));
            const __m256i a1 = vadd(v, va, vc), c1 = vmul(v, vsub(v, va, vc), t1);
            const __m256i b1 = vadd(v, vb, vd), d1 = vmul(v, vsub(v, vb, vd), t2);
            _mm256_store_si256((__m256i *)(a + k), vadd(v, a1, b1));
            _mm256_store_si256((__m256i *)(b + k), vmul(v, vsub(v, a1, b1), t3));
            _mm256_store_si256((__m256i *)(c + k), vadd(v, c1, d1));
            _mm256_store_si256((__m256i *)(d + k), vmul(v, vsub(v, c1, d1), t3));
        }
    }
}

static void dif_rows(const V *v, uint32_t *x, size_t n, size_t wp, const uint32_t *tw) {
    size_t m = n / 2;
    /* whole-matrix passes while a butterfly spans more than a block: two
     * layers per pass where both are that wide, else one */
    while (m >= 1 && 2 * m > CB_BLOCK_ROWS) {
        if (m >= 2 && m > CB_BLOCK_ROWS) {
            dif_quad_pass(v, x, n, m, wp, tw);
            m /= 4;
            continue;
        }
        const size_t step = n / (2 * m);
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < n / 2; t++) {
            const size_t j = t % m, a = (t / m) * 2 * m + j;
            dif_pair(v, x + a * wp, x + (a + m) * wp, wp, vset(tw[j * step]));
        }
        m /= 2;
    }
    if (m == 0) return;
    /* the rest block by block: every remaining layer stays inside a block */
    const size_t blk = 2 * m;
#pragma omp parallel for schedule(static)
    for (size_t b0 = 0; b0 < n; b0 += blk)
        for (size_t mm = m; mm >= 1; mm /= 2) {
            const size_t step = n / (2 * mm);
            for (size_t g = b0; g < b0 + blk; g += 2 * mm)
                for (size_t j = 0; j < mm; j++)
                    dif_pair(v, x + (g + j) * wp, x + (g + j + mm) * wp, wp, vset(tw[j * step]));
        }
}

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

/* ---------------------------------------------------------------- entry -- */
static double g_last_lde_s, g_last_merkle_s;

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* seconds of the last cb_pcs_commit's LDE (input copy through the forward
 * transform) and Merkle tree (leaves and every compress layer) */
void cb_last_phase_seconds(double *lde_s, double *merkle_s) {
    *lde_s = g_last_lde_s;
    *merkle_s = g_last_merkle_s;
}

int cb_pcs_commit(int f, const uint32_t *in_mont, size_t h, size_t w, int lb, uint32_t root_mont[8],
                  uint32_t *lde_out) {
    if (h == 0 || (h & (h - 1)) || w == 0 || lb < 0 || lb > 8) return -1;
    CbField k;
    cb_field_init(f, &k);
    const V v = {vset(k.p), vset(k.pinv)};
    const uint32_t p = k.p;
    int L = 0;
    while (((size_t)1 << L) < h) L++;
    const size_t N = h << lb, wp = (w + 7) & ~(size_t)7;
    uint32_t *x = alloc64(N * wp * 4);
    const uint32_t g = or_generator(f);
    const uint32_t wh = or_two_adic_generator(f, L), wn = or_two_adic_generator(f, L + lb);
    uint32_t *tw_inv = make_twiddles(p, cb_pow(p, wh, p - 2), h), *tw_fwd = make_twiddles(p, wn, N);
    uint32_t *digests = alloc64(2 * N * 32);
    if (!x || !tw_inv || !tw_fwd || !digests) {
        free(x), free(tw_inv), free(tw_fwd), free(digests);
        return -1;
    }
    const double t_start = now_s();
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < N; r++) {
        if (r < h) {
            memcpy(x + r * wp, in_mont + r * w, w * 4);
            memset(x + r * wp + w, 0, (wp - w) * 4);
        } else {
            memset(x + r * wp, 0, wp * 4);
        }
    }
    /* coefficients: inverse DIF leaves c_{rev(k)} at row k */
    dif_rows(&v, x, h, wp, tw_inv);
    /* natural order, each c_i times shift^i / h (pairs k < rev(k) swap) */
    const uint32_t hinv = cb_pow(p, (uint32_t)(h % p), p - 2);
#pragma omp parallel for schedule(static)
    for (size_t kk = 0; kk < h; kk++) {
        const size_t rk = rev_bits(kk, L);
        if (rk < kk) continue;
        const __m256i sk = vset(mont_of(p, (uint32_t)((uint64_t)cb_pow(p, g, rk) * hinv % p)));
        const __m256i sr = vset(mont_of(p, (uint32_t)((uint64_t)cb_pow(p, g, kk) * hinv % p)));
        uint32_t *a = x + kk * wp, *b = x + rk * wp;
        for (size_t c = 0; c < wp; c += 8) {
            const __m256i va = _mm256_load_si256((const __m256i *)(a + c));
            const __m256i vb = _mm256_load_si256((const __m256i *)(b + c));
            /* row k holds c_{rk}: it goes to row rk scaled by g^rk / h */
            _mm256_store_si256((__m256i *)(b + c), vmul(&v, va, sk));
            if (rk != kk) _mm256_store_si256((__m256i *)(a + c), vmul(&v, vb, sr));
        }
    }
    /* evaluations on g <w_N>, committed (bit-reversed) order */
    dif_rows(&v, x, N, wp, tw_fwd);
    if (lde_out) {
#pragma omp parallel for schedule(static)
        for (size_t r = 0; r < N; r++) memcpy(lde_out + r * w, x + r * wp, w * 4);
    }
    const double t_lde = now_s();
    /* Merkle: leaves (8 rows per vector group), then compress layers */
    __m256i diag[16];
    for (int i = 0; i < 16; i++) diag[i] = vset(k.diag[i]);
#pragma omp parallel for schedule(static)
    for (size_t r0 = 0; r0 < N; r0 += 8) {
        const uint32_t *rows[8];
        uint32_t *dst[8];
        uint32_t spill[8][8];
        for (int i = 0; i < 8; i++) {
            const size_t r = r0 + i < N ? r0 + i : N - 1;
            rows[i] = x + r * wp;
            dst[i] = r0 + i < N ? digests + 8 * (r0 + i) : spill[i];
        }
        __m256i s[16];
        for (int i = 0; i < 16; i++) s[i] = _mm256_setzero_si256();
        for (size_t c0 = 0; c0 < w; c0 += 8) {
            const size_t len = w - c0 < 8 ? w - c0 : 8;
            __m256i m[8];
            load_cols(rows, c0, len, m);
            for (size_t i = 0; i < len; i++) s[i] = m[i]; /* overwrite mode */
            v_permute(&v, &k, diag, s);
        }
        store_digests(s, dst);
    }
    size_t off = 0, n = N;
    while (n > 1) {
        const size_t next = off + n;
#pragma omp parallel for schedule(static)
        for (size_t i0 = 0; i0 < n / 2; i0 += 8) {
            const uint32_t *lft[8], *rgt[8];
            uint32_t *dst[8];
            uint32_t spill[8][8];
            for (int i = 0; i < 8; i++) {
                const size_t ii = i0 + i < n / 2 ? i0 + i : n / 2 - 1;
                lft[i] = digests + 8 * (off + 2 * ii);
                rgt[i] = digests + 8 * (off + 2 * ii + 1);
                dst[i] = i0 + i < n / 2 ? digests + 8 * (next + i0 + i) : spill[i];
            }
            __m256i s[16];
            load_cols(lft, 0, 8, s);
            load_cols(rgt, 0, 8, s + 8);
            v_permute(&v, &k, diag, s);
            store_digests(s, dst);
        }
        off = next;
        n /= 2;
    }
    memcpy(root_mont, digests + 8 * off, 32);
    g_last_lde_s = t_lde - t_start;
    g_last_merkle_s = now_s() - t_lde;
    free(x), free(tw_inv), free(tw_fwd), free(digests);
    return 0;
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
