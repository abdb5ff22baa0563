// kernels.h -- launchers shared between the kernel translation units and the
// C-ABI layer (api.hip).  Every launcher enqueues on `stream` and returns the
// hipError_t of the launch; none of them synchronises.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pg {

// Optional per-launch timing hook: begin() before a kernel launch and end()
// after it, on the launch stream (the implementation records HIP events).
struct KernelTimer {
    virtual void begin(const char* name, hipStream_t s) = 0;
    virtual void end(hipStream_t s) = 0;
    virtual ~KernelTimer() {}
};

// One matrix as seen by a hashing kernel: row-major, `stride` elements
// between rows, Montgomery-form values.
struct MatRef {
    const uint32_t* data;
    uint64_t stride;
    uint32_t width;
    uint32_t pad_;
};

// ---- Merkle (merkle.hip) -------------------------------------------------
// digests[i] = sponge(concat_k row i of mats[k]) for i < rows.  `mats` is a
// device array of n MatRef; total_width = sum of widths.
hipError_t launch_leaf_hash(int field, const MatRef* mats, int n, uint32_t total_width,
                            size_t rows, uint32_t* digests, hipStream_t stream);
// next[i] = C(prev[2i], prev[2i+1]), then, if n_inj > 0,
// next[i] = C(next[i], sponge(concat_k row i of inj[k])), for i < n_out.
hipError_t launch_compress(int field, const uint32_t* prev, uint32_t* next, size_t n_out,
                           const MatRef* inj, int n_inj, uint32_t inj_width, hipStream_t stream);
// Opening gather: for each query q < nq and matrix k, copy row
// (idx[q] >> shift[k]) of mats[k] to rows_out[q * total_width + col_off[k] ..];
// for each layer j < log_h, copy digest layer_j[(idx[q] >> j) ^ 1] to
// paths_out[(q * log_h + j) * 8 ..].  layer_off[j] = digest offset of layer j.
hipError_t launch_open_gather(const MatRef* mats, const uint32_t* shift, const uint32_t* col_off,
                              int n_mats, uint32_t total_width, const uint32_t* digests,
                              const uint64_t* layer_off, int log_h, const uint64_t* idx,
                              int nq, uint32_t* rows_out, uint32_t* paths_out,
                              hipStream_t stream);

// ---- NTT / LDE (ntt.hip) --------------------------------------------------
// tw[e] = w_{2^log_n}^e, e < 2^log_n (Montgomery)
hipError_t launch_twiddles(int field, int log_n, uint32_t* tw, hipStream_t stream);
// scale[c*h + k] = (shift * w_N^{rev_b(c)})^k / h for c < 2^b, k < h
hipError_t launch_coset_scale(int field, int log_h, int added_bits, uint32_t shift,
                              uint32_t* scale, hipStream_t stream);
// Coset LDE of an h x w matrix (row stride ld_in) into out (N x w, stride w).
// tmp must hold N*w elements when log_h > 11 (unused otherwise).  The
// output rows are bit-reversed (committed order) when bit_reverse != 0.
hipError_t launch_coset_lde(int field, const uint32_t* in, size_t ld_in, int log_h, uint32_t w,
                            int added_bits, const uint32_t* tw, const uint32_t* scale,
                            uint32_t* tmp, uint32_t* out, int bit_reverse, hipStream_t stream,
                            KernelTimer* kt = nullptr);
int lde_max_log_h();

// ---- FRI (fri.hip) ----------------------------------------------------------
// out[i] = (1/2 + beta/(2 x_i)) v[2i] + (1/2 - beta/(2 x_i)) v[2i+1],
// x_i = w_n^{rev(i)}; v and out are ext4 arrays; tw is the twiddle table of
// size n = 2 * n_out.
hipError_t launch_fri_fold(int field, const uint32_t* v, size_t n_out, const uint32_t beta[4],
                           const uint32_t* tw, uint32_t* out, hipStream_t stream);
// out[i] += add[i] for i < n words (field addition)
hipError_t launch_field_add(int field, uint32_t* out, const uint32_t* add, size_t n,
                            hipStream_t stream);
// Proof-of-work search: smallest witness w in [start, start + count) such
// that the challenger state after observe(w) samples a value whose low
// `bits` bits are zero.  state: 16 words; in: n_in pending inputs.
hipError_t launch_pow_search(int field, const uint32_t* state, const uint32_t* in, int n_in,
                             int bits, uint32_t start, uint32_t count, uint32_t* result,
                             hipStream_t stream);

// ---- misc (api.hip) ------------------------------------------------------
hipError_t launch_convert(int field, uint32_t* data, size_t n, int to_monty, hipStream_t stream);
hipError_t launch_fill_random(int field, uint64_t seed, uint32_t* out, size_t n, int monty,
                              hipStream_t stream);

}  // namespace pg
