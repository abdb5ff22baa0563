// fri.hip -- FRI arity-2 fold, vector add, proof-of-work search, and the
// element-wise helpers (representation conversion, seeded synthetic inputs).
//
// Fold semantics (p3 fold_even_odd; pinned in-tree by the query check at
// vm/src/compiler/recursion/circuit/fri.rs:288-387): the committed vector is
// in bit-reversed evaluation order over <w_n>, so v[2i] and v[2i+1] are the
// values at x_i and -x_i with x_i = w_n^{rev_{log(n/2)}(i)}; the folded value
// is the line through both points evaluated at beta:
//   out[i] = (v[2i] + v[2i+1]) / 2 + beta/2 * (v[2i] - v[2i+1]) * x_i^{-1}.
// PoW semantics: DuplexChallenger::check_witness (circuit/challenger.rs:210-222):
// observe(w) then sample_bits(bits) == 0.
#include "kernels.h"
#include "poseidon2.h"

namespace pg {

template <class F>
__global__ void __launch_bounds__(256) fri_fold_kernel(const uint32_t* __restrict__ v, uint64_t n_out,
                                                       Ext4 half_beta, const uint32_t* __restrict__ tw,
                                                       int log_n_out, uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const uint4 a4 = reinterpret_cast<const uint4*>(v)[2 * i];
    const uint4 b4 = reinterpret_cast<const uint4*>(v)[2 * i + 1];
    const Ext4 a = {{a4.x, a4.y, a4.z, a4.w}}, b = {{b4.x, b4.y, b4.z, b4.w}};
    const uint32_t mask = (uint32_t)(2 * n_out - 1);
    const uint32_t r = log_n_out == 0 ? 0u : (__brev((uint32_t)i) >> (32 - log_n_out));
    const uint32_t x_inv = tw[(0u - r) & mask];  // w_n^{-rev(i)}
    Ext4 s = eadd<F>(a, b);
#pragma unroll
    for (int k = 0; k < 4; k++) s.c[k] = fhalve<F>(s.c[k]);
    const Ext4 d = escale<F>(esub<F>(a, b), x_inv);
    const Ext4 o = eadd<F>(s, emul<F>(d, half_beta));
    reinterpret_cast<uint4*>(out)[i] = make_uint4(o.c[0], o.c[1], o.c[2], o.c[3]);
}

template <class F>
__global__ void field_add_kernel(uint32_t* out, const uint32_t* add, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fadd<F>(out[i], add[i]);
}

template <class F>
__global__ void __launch_bounds__(256) pow_kernel(const uint32_t* __restrict__ state,
                                                  const uint32_t* __restrict__ in, int n_in, int bits,
                                                  uint32_t start, uint32_t count, uint32_t* result) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t w = start + i;
    if (w >= F::P) return;
    // observe(w) then sample(): either way the pending inputs and w overwrite
    // state[0..n_in], one permutation runs, and the sample is state[7].
    uint32_t st[16];
#pragma unroll
    for (int j = 0; j < 16; j++) st[j] = state[j];
    const uint32_t wm = to_monty<F>(w);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (j < n_in) st[j] = in[j];
        if (j == n_in) st[j] = wm;
    }
    permute<F>(st);
    const uint32_t v = from_monty<F>(st[7]);
    if ((v & ((1u << bits) - 1u)) == 0u) atomicMin(result, w);
}

template <class F>
__global__ void convert_kernel(uint32_t* data, uint64_t n, int to) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] = to ? to_monty<F>(data[i]) : from_monty<F>(data[i]);
}

// Element i of seeded stream `seed`: the first of SplitMix64 outputs
// 4i .. 4i+3 of that seed whose top 31 bits are below p, else the last
// candidate reduced mod p.  Output j of SplitMix64 is its finaliser applied to
// seed + (j + 1) * golden, so elements are drawn independently in parallel.
__device__ __forceinline__ uint64_t splitmix_out(uint64_t seed, uint64_t j) {
    uint64_t z = seed + (j + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class F>
__global__ void fill_random_kernel(uint64_t seed, uint32_t* out, uint64_t n, int monty) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x = 0;
    bool ok = false;
    for (int t = 0; t < 4 && !ok; t++) {
        x = (uint32_t)(splitmix_out(seed, 4 * i + t) >> 33);
        ok = x < F::P;
    }
    if (!ok) x %= F::P;
    out[i] = monty ? to_monty<F>(x) : x;
}

static unsigned blocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_fri_fold(int field, const uint32_t* v, size_t n_out, const uint32_t beta[4],
                           const uint32_t* tw, uint32_t* out, hipStream_t stream) {
    if (n_out == 0) return hipSuccess;
    int log_n_out = 0;
    while ((1ull << log_n_out) < n_out) log_n_out++;
    Ext4 hb;
    if (field == FIELD_KOALABEAR) {
        for (int k = 0; k < 4; k++) hb.c[k] = fhalve<KoalaBear>(beta[k]);
        fri_fold_kernel<KoalaBear><<<blocks(n_out), 256, 0, stream>>>(v, n_out, hb, tw, log_n_out, out);
    } else {
        for (int k = 0; k < 4; k++) hb.c[k] = fhalve<BabyBear>(beta[k]);
        fri_fold_kernel<BabyBear><<<blocks(n_out), 256, 0, stream>>>(v, n_out, hb, tw, log_n_out, out);
    }
    return hipGetLastError();
}

// dst row r (r < nrows) = src row rev_q(r0 + r), w words per row: a run of
// natural-order rows of a matrix stored in bit-reversed row order.  One
// thread per word, adjacent threads on adjacent columns (coalesced).
__global__ void rows_bitrev_kernel(const uint32_t* __restrict__ src, uint64_t ld, uint32_t w, int q, uint64_t r0,
                                   uint64_t nrows, uint32_t* __restrict__ dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * w) return;
    const uint64_t r = t / w, c = t % w;
    const uint64_t j = r0 + r;
    const uint64_t sr = q ? (uint64_t)(__brevll(j) >> (64 - q)) : 0;
    dst[t] = ((gcptr)src)[sr * ld + c];
}

hipError_t launch_rows_bitrev(const uint32_t* src, uint64_t ld, uint32_t w, int q, uint64_t r0, uint64_t nrows,
                              uint32_t* dst, hipStream_t stream) {
    if (nrows == 0 || w == 0) return hipSuccess;
    rows_bitrev_kernel<<<blocks(nrows * w), 256, 0, stream>>>(src, ld, w, q, r0, nrows, dst);
    return hipGetLastError();
}

hipError_t launch_field_add(int field, uint32_t* out, const uint32_t* add, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (field == FIELD_KOALABEAR)
        field_add_kernel<KoalaBear><<<blocks(n), 256, 0, stream>>>(out, add, n);
    else
        field_add_kernel<BabyBear><<<blocks(n), 256, 0, stream>>>(out, add, n);
    return hipGetLastError();
}

hipError_t launch_pow_search(int field, const uint32_t* state, const uint32_t* in, int n_in, int bits,
                             uint32_t start, uint32_t count, uint32_t* result, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (field == FIELD_KOALABEAR)
        pow_kernel<KoalaBear><<<blocks(count), 256, 0, stream>>>(state, in, n_in, bits, start, count, result);
    else
        pow_kernel<BabyBear><<<blocks(count), 256, 0, stream>>>(state, in, n_in, bits, start, count, result);
    return hipGetLastError();
}

hipError_t launch_convert(int field, uint32_t* data, size_t n, int to_monty_, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (field == FIELD_KOALABEAR)
        convert_kernel<KoalaBear><<<blocks(n), 256, 0, stream>>>(data, n, to_monty_);
    else
        convert_kernel<BabyBear><<<blocks(n), 256, 0, stream>>>(data, n, to_monty_);
    return hipGetLastError();
}

hipError_t launch_fill_random(int field, uint64_t seed, uint32_t* out, size_t n, int monty,
                              hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (field == FIELD_KOALABEAR)
        fill_random_kernel<KoalaBear><<<blocks(n), 256, 0, stream>>>(seed, out, n, monty);
    else
        fill_random_kernel<BabyBear><<<blocks(n), 256, 0, stream>>>(seed, out, n, monty);
    return hipGetLastError();
}

}  // namespace pg
