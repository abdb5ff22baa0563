/*
 * oracle.c -- CPU oracle (test infrastructure only; see oracle.h).
 *
 * Canonical-form arithmetic with u64 "% p" reduction.  Functions take the
 * field id and dispatch to always-inlined bodies in which p is a compile-time
 * constant, so the compiler turns "% p" into multiply-shift sequences.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KB_P 2130706433u /* 2^31 - 2^24 + 1, gnark/koalabear/koalabear.go:20 */
#define BB_P 2013265921u /* 2^31 - 2^27 + 1 */

#define INL static inline __attribute__((always_inline))

static const uint32_t RC16_RAW[30][16] = {
#include "rc16.inc"
};

/* ---------------------------------------------------------------- field -- */
INL uint32_t P_OF(int f) { return f == OR_KOALABEAR ? KB_P : BB_P; }
INL uint32_t fadd(uint32_t p, uint32_t a, uint32_t b) {
    uint32_t r = a + b; /* a, b < p < 2^31: no overflow */
    return r >= p ? r - p : r;
}
INL uint32_t fsub(uint32_t p, uint32_t a, uint32_t b) { return a >= b ? a - b : a + p - b; }
INL uint32_t fmul(uint32_t p, uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * b) % p);
}
static uint32_t fpow(uint32_t p, uint32_t a, uint64_t e) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t or_prime(int f) { return P_OF(f); }
uint32_t or_add(int f, uint32_t a, uint32_t b) { return fadd(P_OF(f), a, b); }
uint32_t or_sub(int f, uint32_t a, uint32_t b) { return fsub(P_OF(f), a, b); }
uint32_t or_mul(int f, uint32_t a, uint32_t b) { return fmul(P_OF(f), a, b); }
uint32_t or_pow(int f, uint32_t a, uint64_t e) { return fpow(P_OF(f), a, e); }
uint32_t or_inv(int f, uint32_t a) { return fpow(P_OF(f), a, P_OF(f) - 2); }
/* Multiplicative generators 3 (KoalaBear) and 31 (BabyBear): the recursion
 * builder's `generator()` (vm/src/compiler/recursion/ir/utils.rs:13-19), the
 * g of the verifier's x = g w^rev(index) (circuit/fri.rs:151-171).  The tests
 * also check that both are primitive roots. */
uint32_t or_generator(int f) { return f == OR_KOALABEAR ? 3u : 31u; }
int or_two_adicity(int f) { return f == OR_KOALABEAR ? 24 : 27; }
/* two_adic_generator(k) = GEN^((p-1) >> k): order exactly 2^k, and
 * two_adic_generator(k)^2 = two_adic_generator(k-1).  Which primitive 2^k-th
 * root p3 uses (its TWO_ADIC_GENERATORS table) is not in any file of the
 * reference: parity unpinned beyond these properties. */
uint32_t or_two_adic_generator(int f, int log_n) {
    uint32_t p = P_OF(f);
    return fpow(p, or_generator(f), (uint64_t)(p - 1) >> log_n);
}
uint32_t or_to_monty(int f, uint32_t a) { return (uint32_t)(((uint64_t)a << 32) % P_OF(f)); }
uint32_t or_from_monty(int f, uint32_t a) {
    uint32_t p = P_OF(f);
    /* a * 2^-32 mod p */
    return fmul(p, a, fpow(p, fpow(p, 2, 32), p - 2));
}

/* ------------------------------------------------------------ extension -- */
static uint32_t ext_w(int f) { return f == OR_KOALABEAR ? 3u : 11u; }

void or_ext_mul(int f, const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) {
    uint32_t p = P_OF(f), w = ext_w(f);
    uint64_t c[7] = {0};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) c[i + j] = (c[i + j] + (uint64_t)a[i] * b[j]) % p;
    /* x^4 = W: fold c[4..6] down */
    for (int k = 0; k < 4; k++) {
        uint64_t v = c[k];
        if (k + 4 < 7) v = (v + (uint64_t)w * c[k + 4]) % p;
        out[k] = (uint32_t)v;
    }
}

void or_ext_inv(int f, const uint32_t a[4], uint32_t out[4]) {
    /* a^(p^4 - 2) by square-and-multiply over a 128-bit exponent */
    uint32_t p = P_OF(f);
    unsigned __int128 q = p;
    unsigned __int128 e = q * q * q * q - 2;
    uint32_t r[4] = {1, 0, 0, 0}, x[4] = {a[0], a[1], a[2], a[3]}, t[4];
    while (e) {
        if ((uint64_t)e & 1) {
            or_ext_mul(f, r, x, t);
            memcpy(r, t, sizeof t);
        }
        or_ext_mul(f, x, x, t);
        memcpy(x, t, sizeof t);
        e >>= 1;
    }
    memcpy(out, r, sizeof r);
}

/* ------------------------------------------------------------ poseidon2 -- */
/* Internal-layer diagonals (the "M - 1" diagonal), canonical values.
 * KoalaBear: vm/src/primitives/consts.rs:136-154 and
 *            gnark/poseidon2/poseidon2_koalabear.go:162-179.
 * BabyBear:  consts.rs:110-127, gnark/poseidon2/poseidon2_babybear.go:165-180. */
static const uint32_t DIAG_KB[16] = {
    KB_P - 2, 1, 2, (KB_P + 1) >> 1, 3, 4, (KB_P - 1) >> 1, KB_P - 3, KB_P - 4,
    KB_P - ((KB_P - 1) >> 8), KB_P - ((KB_P - 1) >> 3), KB_P - 127, (KB_P - 1) >> 8,
    (KB_P - 1) >> 3, (KB_P - 1) >> 4, 127};
static const uint32_t DIAG_BB[16] = {
    BB_P - 2, 1, 2, (BB_P + 1) >> 1, 3, 4, (BB_P - 1) >> 1, BB_P - 3, BB_P - 4,
    BB_P - ((BB_P - 1) >> 8), BB_P - ((BB_P - 1) >> 2), BB_P - ((BB_P - 1) >> 3), BB_P - 15,
    (BB_P - 1) >> 8, (BB_P - 1) >> 4, 15};

/* M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] applied to each 4-block, then
 * every element gets the sum of the elements in the same block position
 * (poseidon2_koalabear.go:126-159).  Written here as an explicit matrix
 * product, not as the reference's addition chain. */
INL void ext_layer(uint32_t p, uint32_t s[16]) {
    static const uint32_t M4[4][4] = {{2, 3, 1, 1}, {1, 2, 3, 1}, {1, 1, 2, 3}, {3, 1, 1, 2}};
    uint32_t t[16];
    for (int b = 0; b < 4; b++)
        for (int i = 0; i < 4; i++) {
            uint64_t acc = 0;
            for (int j = 0; j < 4; j++) acc += (uint64_t)M4[i][j] * s[4 * b + j];
            t[4 * b + i] = (uint32_t)(acc % p);
        }
    for (int i = 0; i < 4; i++) {
        uint64_t col = 0;
        for (int b = 0; b < 4; b++) col += t[4 * b + i];
        col %= p;
        for (int b = 0; b < 4; b++) s[4 * b + i] = (uint32_t)((t[4 * b + i] + col) % p);
    }
}

INL void int_layer(uint32_t p, const uint32_t *diag, uint32_t s[16]) {
    uint64_t sum = 0;
    for (int i = 0; i < 16; i++) sum += s[i];
    sum %= p;
    for (int i = 0; i < 16; i++) s[i] = (uint32_t)(((uint64_t)s[i] * diag[i] + sum) % p);
}

INL uint32_t sbox(int f, uint32_t p, uint32_t x) {
    uint64_t x2 = ((uint64_t)x * x) % p;
    if (f == OR_KOALABEAR) return (uint32_t)((x2 * x) % p); /* x^3 */
    uint64_t x4 = (x2 * x2) % p;
    return (uint32_t)((((x4 * x2) % p) * x) % p); /* x^7 */
}

/* Round structure: initial external layer; 4 full rounds on RC rows 0..3;
 * RP partial rounds on RC rows 4..4+RP-1 (column 0 only); 4 full rounds on
 * the next 4 rows (mod.rs:305-346 drains rows [4, 4+RP) for the internal
 * constants).  RP = 20 (KoalaBear, x^3), 13 (BabyBear, x^7); consts.rs:185-203. */
INL void permute_impl(int f, uint32_t s[16]) {
    const uint32_t p = P_OF(f);
    const int rp = f == OR_KOALABEAR ? 20 : 13;
    const uint32_t *diag = f == OR_KOALABEAR ? DIAG_KB : DIAG_BB;
    ext_layer(p, s);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 16; i++) s[i] = sbox(f, p, fadd(p, s[i], RC16_RAW[r][i] % p));
        ext_layer(p, s);
    }
    for (int r = 4; r < 4 + rp; r++) {
        s[0] = sbox(f, p, fadd(p, s[0], RC16_RAW[r][0] % p));
        int_layer(p, diag, s);
    }
    for (int r = 4 + rp; r < 8 + rp; r++) {
        for (int i = 0; i < 16; i++) s[i] = sbox(f, p, fadd(p, s[i], RC16_RAW[r][i] % p));
        ext_layer(p, s);
    }
}

void or_poseidon2(int f, uint32_t s[16]) {
    if (f == OR_KOALABEAR)
        permute_impl(OR_KOALABEAR, s);
    else
        permute_impl(OR_BABYBEAR, s);
}

/* Overwrite-mode sponge: state starts at zero; each chunk of up to 8 inputs
 * overwrites state[0..len) and the state is permuted; output state[0..8).
 * (circuit/hash.rs:40-52).  Empty input returns zeros. */
void or_sponge(int f, const uint32_t *in, size_t n, uint32_t out[8]) {
    uint32_t s[16] = {0};
    for (size_t off = 0; off < n; off += 8) {
        size_t len = n - off < 8 ? n - off : 8;
        memcpy(s, in + off, len * sizeof(uint32_t));
        or_poseidon2(f, s);
    }
    memcpy(out, s, 8 * sizeof(uint32_t));
}

/* C(a, b) = P(a || b)[0..8] (circuit/hash.rs:89-94) */
void or_compress(int f, const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
    uint32_t s[16];
    memcpy(s, l, 32);
    memcpy(s + 8, r, 32);
    or_poseidon2(f, s);
    memcpy(out, s, 32);
}

/* --------------------------------------------------------------- inputs -- */
/* Element i of stream `seed` is the first of SplitMix64 outputs 4i..4i+3 of
 * that seed whose top 31 bits are below p; if all four are rejected the last
 * candidate is reduced mod p.  Output j of SplitMix64 is the finaliser applied
 * to seed + (j+1)*golden, so every element can be drawn independently. */
static uint64_t splitmix_out(uint64_t seed, uint64_t j) {
    uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_fill_random(int f, uint64_t seed, uint32_t *out, size_t n) {
    uint32_t p = P_OF(f);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        uint32_t x = 0;
        int ok = 0;
        for (int t = 0; t < 4 && !ok; t++) {
            x = (uint32_t)(splitmix_out(seed, 4 * (uint64_t)i + t) >> 33);
            ok = x < p;
        }
        out[i] = ok ? x : x % p;
    }
}

/* ------------------------------------------------------------------ LDE -- */
static size_t rev_bits(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}
static int log2_exact(size_t n) {
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return ((size_t)1 << l) == n ? l : -1;
}

/* Textbook iterative radix-2 DFT (bit-reversal permutation, then DIT levels):
 * a[j] <- sum_k a[k] * root^(jk), root of order n. */
static void dft_inplace(uint32_t p, uint32_t *a, size_t n, uint32_t root) {
    int lg = log2_exact(n);
    for (size_t i = 0; i < n; i++) {
        size_t j = rev_bits(i, lg);
        if (i < j) {
            uint32_t t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        uint32_t wl = fpow(p, root, n / len);
        for (size_t i = 0; i < n; i += len) {
            uint32_t wv = 1;
            for (size_t j = 0; j < len / 2; j++) {
                uint32_t u = a[i + j], v = fmul(p, a[i + j + len / 2], wv);
                a[i + j] = fadd(p, u, v);
                a[i + j + len / 2] = fsub(p, u, v);
                wv = fmul(p, wv, wl);
            }
        }
    }
}

void or_coset_lde(int f, const uint32_t *in, size_t h, size_t w, int added_bits,
                  uint32_t shift, int bit_reverse, uint32_t *out) {
    uint32_t p = P_OF(f);
    int lh = log2_exact(h);
    size_t n_out = h << added_bits;
    int ln = lh + added_bits;
    uint32_t wh = or_two_adic_generator(f, lh), wn = or_two_adic_generator(f, ln);
    uint32_t wh_inv = fpow(p, wh, p - 2), h_inv = fpow(p, (uint32_t)(h % p), p - 2);
#pragma omp parallel
    {
        uint32_t *col = (uint32_t *)malloc(n_out * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 1)
        for (size_t c = 0; c < w; c++) {
            for (size_t r = 0; r < h; r++) col[r] = in[r * w + c];
            dft_inplace(p, col, h, wh_inv); /* h * coefficients */
            uint32_t sp = h_inv;
            for (size_t k = 0; k < h; k++) {
                col[k] = fmul(p, col[k], sp);
                sp = fmul(p, sp, shift);
            }
            memset(col + h, 0, (n_out - h) * sizeof(uint32_t));
            dft_inplace(p, col, n_out, wn);
            for (size_t r = 0; r < n_out; r++) {
                size_t dst = bit_reverse ? rev_bits(r, ln) : r;
                out[dst * w + c] = col[r];
            }
        }
        free(col);
    }
}

/* ------------------------------------------------------------ pcs open -- */
/* Values of a committed matrix at ext4 points (p3 TwoAdicFriPcs::open's
 * evaluation step; the points are then observed in circuit/fri.rs:84-99
 * order).  The first h = N >> log_blowup rows of the LDE (committed order)
 * hold f on GEN <w_h> in bit-reversed order.  Method: bring them to natural
 * order, inverse DFT to the coefficients of g(X) = f(GEN X), then Horner at
 * z / GEN in the extension. */
void or_pcs_open_values(int f, const uint32_t *lde, size_t n, size_t w, int log_blowup,
                        const uint32_t z[4], uint32_t *out) {
    uint32_t p = P_OF(f);
    size_t h = n >> log_blowup;
    int lh = log2_exact(h);
    uint32_t wh_inv = fpow(p, or_two_adic_generator(f, lh), p - 2), h_inv = fpow(p, (uint32_t)(h % p), p - 2);
    uint32_t g_inv = fpow(p, or_generator(f), p - 2);
    uint32_t u[4] = {fmul(p, z[0], g_inv), fmul(p, z[1], g_inv), fmul(p, z[2], g_inv), fmul(p, z[3], g_inv)};
    uint32_t *col = (uint32_t *)malloc(h * sizeof(uint32_t));
    for (size_t c = 0; c < w; c++) {
        for (size_t j = 0; j < h; j++) col[j] = lde[rev_bits(j, lh) * w + c];
        dft_inplace(p, col, h, wh_inv);
        uint32_t acc[4] = {0, 0, 0, 0}, t[4];
        for (size_t k = h; k-- > 0;) { /* acc = acc * u + coef_k */
            or_ext_mul(f, acc, u, t);
            t[0] = fadd(p, t[0], fmul(p, col[k], h_inv));
            memcpy(acc, t, sizeof t);
        }
        memcpy(out + 4 * c, acc, sizeof acc);
    }
    free(col);
}

/* Reduced openings, restated from the verifier's loop (circuit/fri.rs:
 * 118-241) applied to every row: matrices in the order given, each with its
 * points; one alpha-power counter per LDE log-height, advanced per column;
 * row i of a height-N matrix is x = GEN w_N^rev(i);
 *   ro[lh][i] += (sum_k alpha^pow_k (p_k(z) - M[i][k])) / (z - x).
 * mats[m] is N_m x w_m canonical; npts[m] points, points and values
 * consecutive over (m, point) (values: w_m x 4 per point).  out[lh] is
 * (2^lh x 4) or NULL; rows of heights that occur are overwritten. */
void or_pcs_reduce(int f, size_t nmats, const uint32_t *const *mats, const size_t *heights,
                   const size_t *widths, const uint32_t *npts, const uint32_t *points,
                   const uint32_t *values, const uint32_t alpha[4], uint32_t *const *out) {
    uint32_t p = P_OF(f);
    for (size_t m = 0; m < nmats; m++)
        if (out[log2_exact(heights[m])]) memset(out[log2_exact(heights[m])], 0, heights[m] * 16);
    size_t pow_ctr[64] = {0};
    size_t pi = 0, vi = 0;
    for (size_t m = 0; m < nmats; m++) {
        size_t n = heights[m], w = widths[m];
        int lh = log2_exact(n);
        uint32_t wn = or_two_adic_generator(f, lh);
        for (uint32_t q = 0; q < npts[m]; q++, pi++) {
            const uint32_t *z = points + 4 * pi;
            const uint32_t *pz = values + 4 * vi;
            /* alpha^pow for this (matrix, point)'s columns */
            uint32_t *ap = (uint32_t *)malloc(w * 16 + 16);
            uint32_t a0[4] = {1, 0, 0, 0}, t[4];
            for (size_t e = 0; e < pow_ctr[lh]; e++) {
                or_ext_mul(f, a0, alpha, t);
                memcpy(a0, t, sizeof t);
            }
            for (size_t k = 0; k < w; k++) {
                memcpy(ap + 4 * k, a0, 16);
                or_ext_mul(f, a0, alpha, t);
                memcpy(a0, t, sizeof t);
            }
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < n; i++) {
                uint32_t x = fmul(p, or_generator(f), fpow(p, wn, rev_bits(i, lh)));
                uint32_t acc[4] = {0, 0, 0, 0}, d[4], tt[4];
                for (size_t k = 0; k < w; k++) {
                    uint32_t diff[4] = {fsub(p, pz[4 * k], mats[m][i * w + k]), pz[4 * k + 1], pz[4 * k + 2],
                                        pz[4 * k + 3]};
                    or_ext_mul(f, ap + 4 * k, diff, tt);
                    for (int c = 0; c < 4; c++) acc[c] = fadd(p, acc[c], tt[c]);
                }
                uint32_t den[4] = {fsub(p, z[0], x), z[1], z[2], z[3]};
                or_ext_inv(f, den, d);
                or_ext_mul(f, acc, d, tt);
                uint32_t *o = out[lh] + 4 * i;
                for (int c = 0; c < 4; c++) o[c] = fadd(p, o[c], tt[c]);
            }
            free(ap);
            pow_ctr[lh] += w;
            vi += w;
        }
    }
}

void or_coset_lde_naive(int f, const uint32_t *in, size_t h, size_t w, int added_bits,
                        uint32_t shift, int bit_reverse, uint32_t *out) {
    /* coefficients by the O(h^2) inverse-DFT sum, then O(h*N) evaluation */
    uint32_t p = P_OF(f);
    int lh = log2_exact(h);
    size_t n_out = h << added_bits;
    int ln = lh + added_bits;
    uint32_t wh = or_two_adic_generator(f, lh), wn = or_two_adic_generator(f, ln);
    uint32_t wh_inv = fpow(p, wh, p - 2), h_inv = fpow(p, (uint32_t)(h % p), p - 2);
    uint32_t *coef = (uint32_t *)malloc(h * sizeof(uint32_t));
    for (size_t c = 0; c < w; c++) {
        for (size_t k = 0; k < h; k++) {
            uint64_t acc = 0;
            for (size_t j = 0; j < h; j++)
                acc = (acc + (uint64_t)in[j * w + c] * fpow(p, wh_inv, (uint64_t)j * k)) % p;
            coef[k] = fmul(p, (uint32_t)acc, h_inv);
        }
        for (size_t r = 0; r < n_out; r++) {
            size_t e = bit_reverse ? rev_bits(r, ln) : r;
            uint32_t x = fmul(p, shift, fpow(p, wn, e));
            uint64_t acc = 0; /* Horner */
            for (size_t k = h; k-- > 0;) acc = (acc * x + coef[k]) % p;
            out[r * w + c] = (uint32_t)acc;
        }
    }
    free(coef);
}

/* ----------------------------------------------------------------- MMCS -- */
static int mmcs_shape(const or_matrix *m, size_t n, size_t *max_h, int *log_max) {
    if (n == 0) return -1;
    size_t H = 0;
    for (size_t k = 0; k < n; k++) {
        if (m[k].height == 0 || log2_exact(m[k].height) < 0) return -1;
        if (m[k].height > H) H = m[k].height;
    }
    *max_h = H;
    *log_max = log2_exact(H);
    return 0;
}

size_t or_mmcs_digest_words(const or_matrix *m, size_t n) {
    size_t H;
    int lm;
    if (mmcs_shape(m, n, &H, &lm)) return 0;
    return (2 * H - 1) * 8;
}

/* hash of row i of every matrix with height hgt, in original order (which is
 * the stable height-descending order restricted to one height) */
static void hash_rows_at_height(int f, const or_matrix *m, size_t n, size_t hgt, size_t i,
                                uint32_t *buf, uint32_t out[8]) {
    size_t len = 0;
    for (size_t k = 0; k < n; k++)
        if (m[k].height == hgt) {
            memcpy(buf + len, m[k].data + i * m[k].width, m[k].width * sizeof(uint32_t));
            len += m[k].width;
        }
    or_sponge(f, buf, len, out);
}

int or_mmcs_commit(int f, const or_matrix *m, size_t n, uint32_t *dig, uint32_t root[8]) {
    size_t H;
    int lm;
    if (mmcs_shape(m, n, &H, &lm)) return -1;
    size_t wtot = 0;
    for (size_t k = 0; k < n; k++) wtot += m[k].width;
    /* layer 0 */
#pragma omp parallel
    {
        uint32_t *buf = (uint32_t *)malloc((wtot + 8) * sizeof(uint32_t));
#pragma omp for schedule(static)
        for (size_t i = 0; i < H; i++) hash_rows_at_height(f, m, n, H, i, buf, dig + 8 * i);
        free(buf);
    }
    size_t off = 0, len = H;
    while (len > 1) {
        size_t nl = len / 2;
        uint32_t *prev = dig + 8 * off, *next = dig + 8 * (off + len);
        int inject = 0;
        for (size_t k = 0; k < n; k++) inject |= m[k].height == nl;
#pragma omp parallel
        {
            uint32_t *buf = (uint32_t *)malloc((wtot + 8) * sizeof(uint32_t));
#pragma omp for schedule(static)
            for (size_t i = 0; i < nl; i++) {
                uint32_t d[8];
                or_compress(f, prev + 16 * i, prev + 16 * i + 8, d);
                if (inject) {
                    uint32_t rh[8];
                    hash_rows_at_height(f, m, n, nl, i, buf, rh);
                    or_compress(f, d, rh, d);
                }
                memcpy(next + 8 * i, d, 32);
            }
            free(buf);
        }
        off += len;
        len = nl;
    }
    memcpy(root, dig + 8 * off, 32);
    return 0;
}

int or_mmcs_open(const or_matrix *m, size_t n, const uint32_t *dig, size_t index,
                 uint32_t *rows_out, uint32_t *path_out) {
    size_t H;
    int lm;
    if (mmcs_shape(m, n, &H, &lm) || index >= H) return -1;
    size_t pos = 0;
    for (size_t k = 0; k < n; k++) {
        int lh = log2_exact(m[k].height);
        size_t r = index >> (lm - lh);
        memcpy(rows_out + pos, m[k].data + r * m[k].width, m[k].width * sizeof(uint32_t));
        pos += m[k].width;
    }
    size_t off = 0, len = H;
    for (int j = 0; j < lm; j++) {
        memcpy(path_out + 8 * j, dig + 8 * (off + ((index >> j) ^ 1)), 32);
        off += len;
        len >>= 1;
    }
    return 0;
}

/* Restates verify_batch (circuit/fri.rs:389-436) for power-of-two heights. */
int or_mmcs_verify(int f, const size_t *heights, const size_t *widths, size_t n,
                   const uint32_t root[8], size_t index, const uint32_t *rows,
                   const uint32_t *path) {
    size_t H = 0, wtot = 0;
    for (size_t k = 0; k < n; k++) {
        if (log2_exact(heights[k]) < 0) return 0;
        if (heights[k] > H) H = heights[k];
        wtot += widths[k];
    }
    int lm = log2_exact(H);
    size_t *start = (size_t *)malloc((n + 1) * sizeof(size_t));
    start[0] = 0;
    for (size_t k = 0; k < n; k++) start[k + 1] = start[k] + widths[k];
    uint32_t *buf = (uint32_t *)calloc(wtot + 8, sizeof(uint32_t));
    size_t len = 0;
    for (size_t k = 0; k < n; k++)
        if (heights[k] == H) {
            memcpy(buf + len, rows + start[k], widths[k] * 4);
            len += widths[k];
        }
    uint32_t cur[8];
    or_sponge(f, buf, len, cur);
    size_t hcur = H;
    for (int j = 0; j < lm; j++) {
        uint32_t t[8];
        if ((index >> j) & 1)
            or_compress(f, path + 8 * j, cur, t);
        else
            or_compress(f, cur, path + 8 * j, t);
        memcpy(cur, t, 32);
        hcur >>= 1;
        len = 0;
        for (size_t k = 0; k < n; k++)
            if (heights[k] == hcur) {
                memcpy(buf + len, rows + start[k], widths[k] * 4);
                len += widths[k];
            }
        int any = 0;
        for (size_t k = 0; k < n; k++) any |= heights[k] == hcur;
        if (any) {
            uint32_t rh[8];
            or_sponge(f, buf, len, rh);
            or_compress(f, cur, rh, t);
            memcpy(cur, t, 32);
        }
    }
    free(buf);
    free(start);
    return memcmp(cur, root, 32) == 0;
}

/* ------------------------------------------------ seeded PCS commit root -- */
/* TwoAdicFriPcs::commit of seeded matrices (matrix k = or_fill_random(seed k),
 * h_k x w_k, on the domain s_k <w_h>) computed without holding any LDE whole,
 * so that commitments of tens of GB of LDE fit in host memory:
 *   - LDE shift GEN / s_k, rows bit-reversed (as or_coset_lde);
 *   - the leaves of LDE height N hash the concatenated rows of every matrix of
 *     that height in original order (the stable height-descending order of
 *     MerkleTreeMmcs restricted to one height).  The overwrite-mode sponge is
 *     run incrementally: a row's state and its position in the current chunk
 *     are kept while the matrices' LDEs are produced `block_cols` columns at a
 *     time; each input overwrites state[pos], a full chunk permutes, and a
 *     partial last chunk permutes once at the end -- the same calls as
 *     or_sponge on the whole concatenated row;
 *   - layers as or_mmcs_commit: next[i] = C(prev[2i], prev[2i+1]), then
 *     C(next[i], leaf digest of height len(next)) when that height exists.
 * Pinned against or_mmcs_commit of or_coset_lde outputs (tests). */
int or_pcs_commit_seeded(int f, size_t n, const uint64_t *seeds, const size_t *heights,
                         const size_t *widths, const uint32_t *domain_shifts, int log_blowup,
                         size_t block_cols, uint32_t root[8]) {
    uint32_t p = P_OF(f);
    if (n == 0 || block_cols == 0) return -1;
    size_t H = 0;
    for (size_t k = 0; k < n; k++) {
        if (log2_exact(heights[k]) < 0) return -1;
        size_t N = heights[k] << log_blowup;
        if (N > H) H = N;
    }
    int lm = log2_exact(H);
    /* leaves[l]: digests of the rows of LDE height 2^l (NULL when absent) */
    uint32_t *leaves[64] = {0};
    for (int l = lm; l >= 0; l--) {
        size_t N = (size_t)1 << l;
        int any = 0;
        for (size_t k = 0; k < n; k++) any |= (heights[k] << log_blowup) == N;
        if (!any) continue;
        uint32_t *st = (uint32_t *)calloc(N * 16, sizeof(uint32_t));
        size_t pos = 0; /* inputs absorbed so far, mod 8: the same for every row */
        for (size_t k = 0; k < n; k++) {
            size_t h = heights[k], w = widths[k];
            if ((h << log_blowup) != N || w == 0) continue;
            uint32_t *tr = (uint32_t *)malloc(h * w * sizeof(uint32_t));
            or_fill_random(f, seeds[k], tr, h * w);
            uint32_t shift = fmul(p, or_generator(f), fpow(p, domain_shifts[k], p - 2));
            size_t cbmax = block_cols < w ? block_cols : w;
            uint32_t *blk = (uint32_t *)malloc(h * cbmax * sizeof(uint32_t));
            uint32_t *lde = (uint32_t *)malloc(N * cbmax * sizeof(uint32_t));
            for (size_t c0 = 0; c0 < w; c0 += cbmax) {
                size_t cb = w - c0 < cbmax ? w - c0 : cbmax;
#pragma omp parallel for schedule(static)
                for (size_t r = 0; r < h; r++) memcpy(blk + r * cb, tr + r * w + c0, cb * sizeof(uint32_t));
                or_coset_lde(f, blk, h, cb, log_blowup, shift, 1, lde);
#pragma omp parallel for schedule(static)
                for (size_t r = 0; r < N; r++) {
                    uint32_t *s = st + 16 * r;
                    size_t q = pos;
                    for (size_t c = 0; c < cb; c++) {
                        s[q++] = lde[r * cb + c];
                        if (q == 8) {
                            or_poseidon2(f, s);
                            q = 0;
                        }
                    }
                }
                pos = (pos + cb) % 8;
            }
            free(lde);
            free(blk);
            free(tr);
        }
        if (pos != 0) {
#pragma omp parallel for schedule(static)
            for (size_t r = 0; r < N; r++) or_poseidon2(f, st + 16 * r);
        }
        leaves[l] = (uint32_t *)malloc(N * 32);
        for (size_t r = 0; r < N; r++) memcpy(leaves[l] + 8 * r, st + 16 * r, 32);
        free(st);
    }
    uint32_t *cur = leaves[lm];
    leaves[lm] = NULL;
    for (int l = lm - 1; l >= 0; l--) {
        size_t len = (size_t)1 << l;
        uint32_t *nx = (uint32_t *)malloc(len * 32);
        const uint32_t *inj = leaves[l];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < len; i++) {
            uint32_t d[8];
            or_compress(f, cur + 16 * i, cur + 16 * i + 8, d);
            if (inj) or_compress(f, d, inj + 8 * i, d);
            memcpy(nx + 8 * i, d, 32);
        }
        free(cur);
        free(leaves[l]);
        leaves[l] = NULL;
        cur = nx;
    }
    memcpy(root, cur, 32);
    free(cur);
    return 0;
}

/* ------------------------------------------------------------------ FRI -- */
/* v'[i] = (1/2 + beta/(2 x_i)) v[2i] + (1/2 - beta/(2 x_i)) v[2i+1],
 * x_i = w_n^{rev_{log(n/2)}(i)}: the line through (x_i, v[2i]) and
 * (-x_i, v[2i+1]) evaluated at beta (circuit/fri.rs:288-387). */
void or_fri_fold(int f, const uint32_t *v, size_t n, const uint32_t beta[4], uint32_t *out) {
    uint32_t p = P_OF(f);
    size_t half = n / 2;
    int lh = log2_exact(half);
    uint32_t wn = or_two_adic_generator(f, lh + 1);
    uint32_t inv2 = fpow(p, 2, p - 2);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < half; i++) {
        uint32_t x = fpow(p, wn, rev_bits(i, lh));
        uint32_t c = fmul(p, inv2, fpow(p, x, p - 2)); /* 1/(2x) */
        const uint32_t *a = v + 8 * i, *b = v + 8 * i + 4;
        for (int k = 0; k < 4; k++) {
            /* (a+b)/2 + beta * (a-b)/(2x), ext4 beta times base (a-b)c */
            uint32_t s = fmul(p, fadd(p, a[k], b[k]), inv2);
            out[4 * i + k] = s;
        }
        uint32_t d[4], t[4];
        for (int k = 0; k < 4; k++) d[k] = fmul(p, fsub(p, a[k], b[k]), c);
        or_ext_mul(f, beta, d, t);
        for (int k = 0; k < 4; k++) out[4 * i + k] = fadd(p, out[4 * i + k], t[k]);
    }
}

/* ----------------------------------------------------------- challenger -- */
void or_ch_init(or_challenger *c, int f) {
    memset(c, 0, sizeof *c);
    c->field = f;
}
static void ch_duplex(or_challenger *c) {
    for (int i = 0; i < c->n_in; i++) c->state[i] = c->in[i];
    c->n_in = 0;
    or_poseidon2(c->field, c->state);
    memcpy(c->out, c->state, 32);
    c->n_out = 8;
}
void or_ch_observe(or_challenger *c, uint32_t v) {
    c->n_out = 0;
    c->in[c->n_in++] = v;
    if (c->n_in == 8) ch_duplex(c);
}
uint32_t or_ch_sample(or_challenger *c) {
    if (c->n_in > 0 || c->n_out == 0) ch_duplex(c);
    return c->out[--c->n_out];
}
void or_ch_sample_ext(or_challenger *c, uint32_t out[4]) {
    for (int k = 0; k < 4; k++) out[k] = or_ch_sample(c);
}
uint32_t or_ch_sample_bits(or_challenger *c, int bits) {
    return or_ch_sample(c) & (uint32_t)((1ull << bits) - 1);
}
int or_ch_check_witness(or_challenger *c, int bits, uint32_t witness) {
    or_ch_observe(c, witness);
    return or_ch_sample_bits(c, bits) == 0;
}

int or_fri_commit_phase(or_challenger *c, const uint32_t *v, int log_n, int log_stop,
                        const uint32_t *const *inputs, uint32_t *roots, uint32_t *layers,
                        uint32_t *final_vec, uint32_t *betas) {
    int f = c->field;
    uint32_t p = P_OF(f);
    size_t n = (size_t)1 << log_n;
    uint32_t *cur = (uint32_t *)malloc(n * 16);
    memcpy(cur, v, n * 16);
    size_t lay_off = 0;
    int r = 0;
    while (n > ((size_t)1 << log_stop)) {
        size_t half = n / 2;
        memcpy(layers + lay_off, cur, n * 16);
        lay_off += 4 * n;
        /* commit rows of 2 ext4 = 8 base elements (ExtensionMmcs flattening) */
        or_matrix m = {cur, half, 8};
        uint32_t *dig = (uint32_t *)malloc((2 * half - 1) * 32);
        or_mmcs_commit(f, &m, 1, dig, roots + 8 * r);
        free(dig);
        for (int k = 0; k < 8; k++) or_ch_observe(c, roots[8 * r + k]);
        or_ch_sample_ext(c, betas + 4 * r);
        uint32_t *nx = (uint32_t *)malloc(half * 16);
        or_fri_fold(f, cur, n, betas + 4 * r, nx);
        free(cur);
        cur = nx;
        n = half;
        r++;
        int lg = log_n - r;
        if (inputs && inputs[lg])
            for (size_t i = 0; i < 4 * n; i++) cur[i] = fadd(p, cur[i], inputs[lg][i]);
    }
    memcpy(final_vec, cur, n * 16);
    free(cur);
    return r;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
