// kernels.h -- launchers shared between the kernel translation units and the
// C-ABI layer (api.hip).  Every launcher enqueues on `stream` and returns the
// hipError_t of the launch; none of them synchronises.
#pragma once
#include <hip/hip_runtime.h>

#include "field.h"
#include <stddef.h>
#include <stdint.h>

namespace pg {

// Global-memory pointers.  A pointer that reaches a kernel inside a struct or
// a device table is generic to the compiler, which then emits FLAT accesses:
// those pay the LDS-aperture check and count against the LDS wait counter, so
// every LDS exchange would also wait for outstanding HBM loads.  Kernels cast
// such pointers to address space 1 once.
#define GLOBAL __attribute__((address_space(1)))
typedef GLOBAL const uint32_t* gcptr;
typedef GLOBAL uint32_t* gptr;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // copyable across address spaces

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
// One segment of every row's leaf sponge (state kept in `state`, 16 words
// per row, between segments); `last` writes the digests.
hipError_t launch_absorb_segment(int field, const uint32_t* data, uint64_t stride, uint32_t width, uint32_t pos0,
                                 int first, int last, uint64_t rows, uint32_t* state, uint32_t* digests,
                                 hipStream_t stream);
hipError_t launch_leaf_hash(int field, const MatRef* mats, int n, uint32_t total_width,
                            size_t rows, uint32_t* digests, hipStream_t stream);
// next[i] = C(prev[2i], prev[2i+1]), then, if n_inj > 0,
// next[i] = C(next[i], sponge(concat_k row i of inj[k])), for i < n_out.
hipError_t launch_compress(int field, const uint32_t* prev, uint32_t* next, size_t n_out,
                           const MatRef* inj, int n_inj, uint32_t inj_width, hipStream_t stream);
// The same with the injected rows already hashed: next[i] = C(C(prev[2i],
// prev[2i+1]), inj[i]) for 8-word digests inj[i].
hipError_t launch_compress_digest(int field, const uint32_t* prev, uint32_t* next, size_t n_out,
                                  const uint32_t* inj, hipStream_t stream);
// Opening gather: for each query q < nq and matrix k, copy row
// (idx[q] >> shift[k]) of mats[k] to rows_out[q * total_width + col_off[k] ..];
// for each layer j < log_h, copy digest layer_j[(idx[q] >> j) ^ 1] to
// paths_out[(q * log_h + j) * 8 ..].  layer_off[j] = digest offset of layer j.
hipError_t launch_open_gather(const MatRef* mats, const uint32_t* shift, const uint32_t* col_off,
                              int n_mats, uint32_t total_width, const uint32_t* digests,
                              const uint64_t* layer_off, int log_h, const uint64_t* idx,
                              int nq, uint32_t* rows_out, uint32_t* paths_out,
                              hipStream_t stream);

// ---- batched Merkle (merkle.hip): many single-matrix commits per launch ---
// Jobs are found by binary search of the thread's global row / node index in
// a prefix array (njobs + 1 entries, prefix[0] = 0).
struct LeafJob {
    const uint32_t* data;  // rows x width, row stride `stride`
    uint64_t stride;
    uint32_t width;
    uint32_t pad_;
    uint32_t* digests;  // rows x 8
};
struct CompressJob {
    const uint32_t* prev;  // 2 n x 8
    uint32_t* next;        // n x 8
};
hipError_t launch_leaf_hash_batch(int field, const LeafJob* jobs, const uint64_t* row_prefix, int njobs,
                                  uint64_t total_rows, hipStream_t stream);
hipError_t launch_compress_batch(int field, const CompressJob* jobs, const uint64_t* node_prefix,
                                 int njobs, uint64_t total_nodes, hipStream_t stream);
// dst[8 j + k] = src[j][k] for j < n (root digests of n trees)
hipError_t launch_gather_roots(const uint32_t* const* src, uint32_t* dst, int n, hipStream_t stream);

// ---- NTT / LDE (ntt.hip) --------------------------------------------------
// One LDE job as the kernels see it (see ntt.hip for the pass structure).
struct LdeArgs {
    const uint32_t* in;    // h x w input, row stride ld_in
    uint64_t ld_in;
    uint32_t* out;         // N x w output, row stride ld_out (also pass-A scratch)
    uint64_t ld_out;
    uint32_t* tmp;         // N x w scratch (multi-pass only)
    const uint32_t* tw;    // w_h^e, e < h
    const uint32_t* scale; // (shift_c)^k / h, [c h + k]
    uint64_t ntiles;       // tiles of this job in the current pass
    uint32_t w;
    int L, k1, k2, b, bitrev;
};
enum LdeKind { KIND_SINGLE = 0, KIND_A = 1, KIND_BC = 2, KIND_D = 3 };
// Split of a size-2^L transform: L <= 11 is one pass (k1 = 0, k2 = L);
// otherwise k1 = ceil(L/2) strided bits (10 at L = 21), k2 = L - k1 block bits.
void lde_split(int L, int* k1, int* k2);
// Columns per tile of the kernel that runs pass `kind` with tile size 2^log_s
// in tile-shape variant `variant`.
int lde_pass_cols(int kind, int log_s, int variant);
// Tile-shape variants: 1 or 2 per (pass, log size); the choice is kept per
// (field, pass, log size, width bucket) and read by every LDE launch.
constexpr int LDE_WIDTH_BUCKETS = 4;  // widths <= 8, <= 32, <= 128, larger
int lde_width_bucket(uint32_t w);
int lde_num_variants(int kind, int log_s);
int lde_variant(int field, int kind, int log_s, uint32_t w);
int lde_set_variant(int field, int kind, int log_s, int bucket, int variant);
// Row tiles (before the column-block factor) of pass `kind` for one job.
uint64_t lde_pass_row_tiles(int kind, int L, int b);
// One pass over a device table of jobs (all with the same L); prefix holds
// each job's first tile (row tiles x column blocks), prefix[njobs] = total;
// max_b = the largest added_bits among the jobs.
hipError_t launch_lde_pass_batch(int field, int kind, int log_s, int variant, const LdeArgs* jobs,
                                 const uint64_t* prefix, int njobs, uint64_t total_tiles, int max_b,
                                 hipStream_t stream, KernelTimer* kt = nullptr);
// tw[e] = w_{2^log_n}^e, e < 2^log_n (Montgomery)
hipError_t launch_twiddles(int field, int log_n, uint32_t* tw, hipStream_t stream);
// scale[c*h + k] = (shift * w_N^{rev_b(c)})^k / h for c < 2^b, k < h
hipError_t launch_coset_scale(int field, int log_h, int added_bits, uint32_t shift,
                              uint32_t* scale, hipStream_t stream);
// Coset LDE of an h x w matrix (row stride ld_in) into out (N x w, row
// stride ld_out >= w, so a column slice of a wider matrix can be written).
// tmp must hold N*w elements (stride w) when log_h > 11 (unused otherwise).
// The output rows are bit-reversed (committed order) when bit_reverse != 0.
hipError_t launch_coset_lde(int field, const uint32_t* in, size_t ld_in, int log_h, uint32_t w,
                            int added_bits, const uint32_t* tw, const uint32_t* scale,
                            uint32_t* tmp, uint32_t* out, size_t ld_out, int bit_reverse,
                            hipStream_t stream, KernelTimer* kt = nullptr);
int lde_max_log_h();

// ---- Pcs::open (open.hip) ------------------------------------------------
// Values of one matrix at npts <= 4 points z[q]: weights (npts x h Ext4) and
// partial sums (npts x nchunks x w Ext4) are caller-provided scratch; out
// gets npts x w Ext4, factor[q] * sum_r M[r] c_q[r] over the first
// h = 2^log_h rows, tw = w_h^e table.
// whether launch_open_values reads four columns per thread (then a thread
// slot covers 4 columns when sizing the row chunks)
bool open_values_vec4(const uint32_t* m, uint64_t ld, uint32_t w, int npts);
hipError_t launch_open_values(int field, const uint32_t* m, uint64_t ld, uint32_t w, int log_h,
                              const uint32_t* tw, const Ext4* z, const Ext4* factor, int npts, Ext4* weights,
                              int weights_ready, Ext4* partial, uint32_t nchunks, uint32_t chunk_rows, Ext4* sums,
                              uint32_t nslices, Ext4* out, hipStream_t stream);
// Reduced opening of one LDE log-height (see open.hip).  mats / pts are
// device tables; alpha_pows[k] = alpha^k for k < the widest matrix.
constexpr int RO_MAX_POINTS = 4;  // distinct points per log-height
struct RoMat {
    const uint32_t* data;
    uint64_t ld;
    uint32_t width;
    int vec4;           // rows 16-byte aligned (width, ld multiples of 4)
    int npts;           // points opened on this matrix (<= 2)
    int pt[2];          // index into the height's distinct points
    Ext4 a[2];          // alpha^off * sum_k alpha^k p_k(z)
    Ext4 aoff[2];       // alpha^off
};
hipError_t launch_reduce_rows(int field, const RoMat* mats, int nmats, const Ext4* pts, int npts,
                              const Ext4* alpha_pows, int log_n, const uint32_t* tw, Ext4* out,
                              hipStream_t stream);

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
// dst row r = src row rev_q(r0 + r) for r < nrows (w words per row, src pitch ld)
hipError_t launch_rows_bitrev(const uint32_t* src, uint64_t ld, uint32_t w, int q, uint64_t r0, uint64_t nrows,
                              uint32_t* dst, hipStream_t stream);
hipError_t launch_fill_random(int field, uint64_t seed, uint32_t* out, size_t n, int monty,
                              hipStream_t stream);


// ---- quotient (quotient.hip) ----------------------------------------------
// A workgroup (one wave of QT threads) evaluates QB = QT QR consecutive rows
// of the committed LDE; thread t owns rows t + QT k, k < QR.  Its LDS holds,
// in this order: the program's registers (register r of row u at word
// r QB + u); `nslots` staged windows (QW columns of the block's QB rows, or
// of the block of rows that are their next rows; column c of a window at
// word c QRS + row); the selectors is_first_row, is_last_row, is_transition
// (QB words each) when the program reads them; the constant table (the
// AIR's constants, then the public values).
//
// An operand is an LDS word index read with one of four row offsets: QK_ROW
// (the row itself: registers, windows, selectors), QK_NEXT (the slot of the
// row's next row in a next-row window), QK_BCAST (none: every row reads the
// same constant) or QK_PREV (no read: the previous arithmetic result, held
// in VGPRs).  The host folds the operation and the operand kinds into the
// opcode, so the device runs one straight-line handler per instruction:
//   16 op + 4 ka + kb   ADD (op 0), SUB (1), MUL (2) of operands a, b
//   QOP_NEG + ka        -a
//   QOP_ASSERT + ka     the accumulator gains a times the next power of alpha
//   QOP_STAGE           stage columns [a, a + QW) of source b (0 main local,
//                       1 main next, 2 preprocessed local, 3 preprocessed
//                       next, 4 permutation local, 5 permutation next) into
//                       window slot dst
// dst of an arithmetic instruction is the LDS word of its register, or
// QO_NO_DST for a value only the next instruction reads (forwarded).
constexpr uint32_t QK_ROW = 0, QK_NEXT = 1, QK_BCAST = 2, QK_PREV = 3;
constexpr uint32_t QOP_NEG = 48, QOP_ASSERT = 52, QOP_STAGE = 56;
constexpr uint32_t QO_NO_DST = 0xFFFFFFFFu;
struct QInstr {
    uint32_t op, dst, a, b;
};
constexpr int QT = 64;               // threads per workgroup (one wave)
constexpr int QR = 2;                // rows per thread
constexpr int QB = QT * QR;          // rows per workgroup
constexpr int QW = 8;                // columns per staged window (32-byte row segments)
constexpr int QRS = QB + QT / QW;    // LDS words per window column (conflict-free staging writes)
constexpr int QSLOT = QW * QRS;      // LDS words per window
constexpr uint32_t QUOTIENT_SLOTS = 4;
constexpr uint32_t QUOTIENT_MAX_REGS = (65536 / 4 - QUOTIENT_SLOTS * QSLOT) / QB;  // < 64 KiB of LDS
struct QuotientArgs {
    const uint32_t* main;  // committed LDE, bit-reversed rows
    uint64_t ld_main;
    uint32_t w_main;
    const uint32_t* pre;   // may be null when the program reads no preprocessed column
    uint64_t ld_pre;
    uint32_t w_pre;
    const uint32_t* perm;  // the permutation trace, flattened to base columns (or null)
    uint64_t ld_perm;
    uint32_t w_perm;
    const QInstr* prog;
    uint32_t nprog, nregs, nslots;
    const uint32_t* consts;  // constant table (Montgomery), copied to LDS after the selectors
    uint32_t nconsts;
    const uint32_t* apow;    // alpha^(m-1-j) for assert j (ext4, Montgomery)
    uint32_t nasserts;
    const uint32_t* tw;    // w_Q^e, e < |Q|
    const uint32_t* zh;    // Z_H at the 2^qdb values of x^n, then their inverses
    uint32_t zh_mask;      // 2^qdb - 1
    uint32_t q, log_q, next_step;
    uint32_t shift, omega_n_inv;  // GEN and w_n^-1 (Montgomery)
    int need_selectors;
    uint32_t* out;         // |Q| x 4
};
size_t quotient_lds_bytes(uint32_t nregs, uint32_t nslots, bool selectors, uint32_t nconsts);
hipError_t launch_quotient(int field, const QuotientArgs& a, hipStream_t stream);
}  // namespace pg
