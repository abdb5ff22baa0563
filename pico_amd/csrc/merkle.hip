// merkle.hip -- Poseidon2 Merkle kernels for the mixed-height MMCS:
// leaf hashing (sponge over concatenated same-height rows), compress/inject
// layers and the opening gather.
//
// Semantics (p3 MerkleTreeMmcs, pinned in-tree by the inverse verify_batch at
// vm/src/compiler/recursion/circuit/fri.rs:389-436):
//   layer0[i] = H(concat of row i of every matrix of the tallest height)
//   next[i]   = C(prev[2i], prev[2i+1]); if matrices of height len(next)
//               exist: next[i] = C(next[i], H(concat of their row i))
//   H = PaddingFreeSponge<16, 8, 8> (overwrite mode, circuit/hash.rs:40-52)
//   C = TruncatedPermutation (circuit/hash.rs:89-94)
// One thread owns one permutation state for the whole row / node.
#include "kernels.h"
#include "poseidon2.h"

namespace pg {

// Absorb the concatenation of row `row` of mats[0..n) into st (which the
// caller zeroes).  The cursor (k, j) is wave-uniform, so the matrix
// descriptors are read through the scalar cache.
template <class F>
__device__ __forceinline__ void absorb_rows(uint32_t (&st)[16], const MatRef* mats, int n,
                                            uint32_t total, uint64_t row) {
    if (total == 0) return;
    int k = 0;
    uint32_t j = 0;
    gcptr base = (gcptr)mats[0].data + row * mats[0].stride;
    uint32_t wk = mats[0].width;
    for (uint32_t off = 0; off < total; off += 8) {
        uint32_t cnt = total - off < 8u ? total - off : 8u;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if ((uint32_t)t < cnt) {
                while (j >= wk && k + 1 < n) {
                    k++;
                    j = 0;
                    base = (gcptr)mats[k].data + row * mats[k].stride;
                    wk = mats[k].width;
                }
                st[t] = base[j];
                j++;
            }
        }
        permute<F>(st);
    }
}

// ---- one permutation over 16 lanes ---------------------------------------
// A layer with few nodes cannot fill the GPU, so with one state per thread it
// costs one permutation's latency (about 5200 dependent VALU instructions,
// ~11 us on MI355X) however small it is: the top ~15 layers of every tree
// and the small FRI rounds.  For those layers the 16 lanes of a DPP row
// (lanes 16r .. 16r+15 of a wave) hold one state, lane l holding s_l: the
// S-boxes of a full round run side by side, and the linear layers take their
// cross-lane terms from DPP moves (quad_perm inside groups of 4 lanes,
// row_ror across the groups): about 35 dependent VALU instructions per
// round, 983 per permutation against ~5200.  All arithmetic is on reduced
// values, so the result is the same permutation word for word.
static __constant__ P2Const<KoalaBear> kLaneKb = make_p2const<KoalaBear>();
static __constant__ P2Const<BabyBear> kLaneBb = make_p2const<BabyBear>();

template <class F>
__device__ __forceinline__ const P2Const<F>& lane_consts() {
    if constexpr (F::P == KoalaBear::P)
        return kLaneKb;
    else
        return kLaneBb;
}

// v from lane CTRL selects, within the lane's DPP row (every lane active)
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_from(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
constexpr int DPP_QUAD_SWAP1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_SWAP2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_QUAD_NEXT = 0x39;   // quad_perm [1,2,3,0]: lane i of a quad reads lane i+1
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;

// Sum over the lane's group of 4, then over the 4 groups: both orbits are
// complete whichever way the moves rotate.
template <class F>
__device__ __forceinline__ uint32_t quad_sum(uint32_t x) {
    const uint32_t a = fadd<F>(x, dpp_from<DPP_QUAD_SWAP1>(x));
    return fadd<F>(a, dpp_from<DPP_QUAD_SWAP2>(a));
}
template <class F>
__device__ __forceinline__ uint32_t across_quads(uint32_t y) {
    const uint32_t u = fadd<F>(y, dpp_from<DPP_ROW_ROR4>(y));
    return fadd<F>(u, dpp_from<DPP_ROW_ROR8>(u));
}

// External layer: M4 row i is (2, 3, 1, 1) rotated by i, i.e.
// y_i = (x_0 + x_1 + x_2 + x_3) + x_i + 2 x_{i+1 mod 4}; then
// s_i = y_i + sum over the four groups of y at position i mod 4.
template <class F>
__device__ __forceinline__ uint32_t external_lanes(uint32_t x) {
    const uint32_t nx = dpp_from<DPP_QUAD_NEXT>(x);
    const uint32_t y = fadd<F>(fadd<F>(quad_sum<F>(x), x), fadd<F>(nx, nx));
    return fadd<F>(y, across_quads<F>(y));
}

// Internal layer: s_i = s_i diag_i + sum of all 16
template <class F>
__device__ __forceinline__ uint32_t internal_lanes(uint32_t x, uint32_t diag) {
    return fadd<F>(fmul<F>(x, diag), across_quads<F>(quad_sum<F>(x)));
}

// x = s_l in, P(s)_l out; l = lane index within the DPP row.
template <class F>
__device__ __forceinline__ uint32_t permute_lanes(uint32_t x, int l) {
    const P2Const<F>& K = lane_consts<F>();
    uint32_t rc[8];
#pragma unroll
    for (int r = 0; r < 8; r++) rc[r] = K.ext_rc[r][l];
    const uint32_t diag = K.diag[l];
    x = external_lanes<F>(x);
#pragma unroll
    for (int r = 0; r < 4; r++) x = external_lanes<F>(sbox<F>(fadd<F>(x, rc[r])));
#pragma unroll
    for (int r = 0; r < F::ROUNDS_P; r++) {
        const uint32_t v = sbox<F>(fadd<F>(x, P2<F>::K.int_rc[r]));
        x = internal_lanes<F>(l == 0 ? v : x, diag);
    }
#pragma unroll
    for (int r = 4; r < 8; r++) x = external_lanes<F>(sbox<F>(fadd<F>(x, rc[r])));
    return x;
}

// Layers of at most this many nodes (or leaves) take the 16-lane kernels:
// above it the 16x larger thread count costs more than the latency saved.
constexpr uint64_t LANES_MAX_NODES = 1u << 14;

// next[i] = C(prev[2i], prev[2i+1]): the two child digests are the state.
// Thread 16 i + l is lane l of node i; a node's 16 lanes leave together.
template <class F>
__global__ void __launch_bounds__(256) compress_lanes_kernel(const uint32_t* prev, uint32_t* next, uint64_t n_out) {
    const uint64_t node = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int l = (int)(threadIdx.x & 15u);
    if (node >= n_out) return;
    const uint32_t x = permute_lanes<F>(((gcptr)prev)[16 * node + l], l);
    if (l < 8) ((gptr)next)[8 * node + l] = x;
}

// next[i] = C(C(prev[2i], prev[2i+1]), inj[i]) with the digests inj[i] of the
// injected rows computed beforehand: the second state is the first one's
// output with its upper half replaced by inj[i].
template <class F>
__global__ void __launch_bounds__(256) compress_digest_lanes_kernel(const uint32_t* prev, const uint32_t* inj,
                                                                    uint32_t* next, uint64_t n_out) {
    const uint64_t node = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int l = (int)(threadIdx.x & 15u);
    if (node >= n_out) return;
    const uint32_t d = l < 8 ? 0u : ((gcptr)inj)[8 * node + (l - 8)];
    uint32_t x = permute_lanes<F>(((gcptr)prev)[16 * node + l], l);
    x = permute_lanes<F>(l < 8 ? x : d, l);
    if (l < 8) ((gptr)next)[8 * node + l] = x;
}

// Rows of exactly 8 words: one sponge chunk, st = (row, 0^8).
template <class F>
__global__ void __launch_bounds__(256) leaf8_lanes_kernel(const uint32_t* data, uint64_t stride, uint64_t rows,
                                                          uint32_t* digests) {
    const uint64_t row = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int l = (int)(threadIdx.x & 15u);
    if (row >= rows) return;
    const uint32_t x = permute_lanes<F>(l < 8 ? ((gcptr)data)[row * stride + l] : 0u, l);
    if (l < 8) ((gptr)digests)[8 * row + l] = x;
}

// General leaf hash: any number of matrices, any widths.
template <class F>
__global__ void __launch_bounds__(256) leaf_hash_kernel(const MatRef* mats, int n, uint32_t total,
                                                        uint64_t rows, uint32_t* digests) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    uint32_t st[16];
#pragma unroll
    for (int t = 0; t < 16; t++) st[t] = 0;
    absorb_rows<F>(st, mats, n, total, i);
    uint4* o = reinterpret_cast<uint4*>(digests + 8 * i);
    o[0] = make_uint4(st[0], st[1], st[2], st[3]);
    o[1] = make_uint4(st[4], st[5], st[6], st[7]);
}

// One matrix whose width is a multiple of 8 and whose rows are 32-byte
// aligned: every sponge chunk is two 16-byte loads.
template <class F>
__global__ void __launch_bounds__(256) leaf_hash_w8_kernel(const uint32_t* data, uint64_t stride,
                                                           uint32_t chunks, uint64_t rows,
                                                           uint32_t* digests) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const uint4* rp = reinterpret_cast<const uint4*>(data + i * stride);
    uint32_t st[16];
#pragma unroll
    for (int t = 0; t < 16; t++) st[t] = 0;
    for (uint32_t c = 0; c < chunks; c++) {
        uint4 a = rp[2 * c], b = rp[2 * c + 1];
        st[0] = a.x; st[1] = a.y; st[2] = a.z; st[3] = a.w;
        st[4] = b.x; st[5] = b.y; st[6] = b.z; st[7] = b.w;
        permute<F>(st);
    }
    uint4* o = reinterpret_cast<uint4*>(digests + 8 * i);
    o[0] = make_uint4(st[0], st[1], st[2], st[3]);
    o[1] = make_uint4(st[4], st[5], st[6], st[7]);
}

// One segment of the leaf sponge: columns [0, width) of a matrix view are
// the next `width` inputs of every row's sponge, whose state lives in
// `state` (16 words per row) between segments.  pos0 = inputs already
// written into the pending chunk (the running input count mod 8): in
// overwrite mode a chunk's inputs go to st[0..8) and the permutation runs
// when it is full, so st itself carries a partial chunk across segments.
// first: start from the zero state; last: run the final partial chunk's
// permutation (if any) and write the digest instead of the state.  The
// segments of a row, concatenated, give the same digest as one pass of
// absorb_rows over the whole row.
template <class F, bool VEC>
__global__ void __launch_bounds__(256) absorb_segment_kernel(const uint32_t* data, uint64_t stride,
                                                             uint32_t width, uint32_t pos0, int first,
                                                             int last, uint64_t rows, uint32_t* state,
                                                             uint32_t* digests) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    uint32_t st[16];
    if (first) {
#pragma unroll
        for (int t = 0; t < 16; t++) st[t] = 0;
    } else {
        const GLOBAL u32x4* sp = reinterpret_cast<const GLOBAL u32x4*>((gcptr)state + 16 * i);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            u32x4 v = sp[q];
            st[4 * q] = v.x; st[4 * q + 1] = v.y; st[4 * q + 2] = v.z; st[4 * q + 3] = v.w;
        }
    }
    gcptr row = (gcptr)data + i * stride;
    uint32_t pos = pos0;
    if constexpr (VEC) {  // pos0 == 0, width % 8 == 0, 16-byte aligned rows
        const GLOBAL u32x4* rp = reinterpret_cast<const GLOBAL u32x4*>(row);
        for (uint32_t c = 0; c < width / 8; c++) {
            u32x4 a = rp[2 * c], b = rp[2 * c + 1];
            st[0] = a.x; st[1] = a.y; st[2] = a.z; st[3] = a.w;
            st[4] = b.x; st[5] = b.y; st[6] = b.z; st[7] = b.w;
            permute<F>(st);
        }
    } else {
        uint32_t e = 0;
        while (e < width) {  // pos and e are the same in every lane
            const uint32_t take = width - e < 8 - pos ? width - e : 8 - pos;
#pragma unroll
            for (int t = 0; t < 8; t++)
                if ((uint32_t)t >= pos && (uint32_t)t < pos + take) st[t] = row[e + (uint32_t)t - pos];
            e += take;
            pos += take;
            if (pos == 8) {
                permute<F>(st);
                pos = 0;
            }
        }
    }
    if (last) {
        if (pos > 0) permute<F>(st);
        GLOBAL u32x4* o = reinterpret_cast<GLOBAL u32x4*>((gptr)digests + 8 * i);
        o[0] = u32x4{st[0], st[1], st[2], st[3]};
        o[1] = u32x4{st[4], st[5], st[6], st[7]};
    } else {
        GLOBAL u32x4* o = reinterpret_cast<GLOBAL u32x4*>((gptr)state + 16 * i);
#pragma unroll
        for (int q = 0; q < 4; q++) o[q] = u32x4{st[4 * q], st[4 * q + 1], st[4 * q + 2], st[4 * q + 3]};
    }
}

template <class F>
__global__ void __launch_bounds__(256) compress_kernel(const uint32_t* prev, uint32_t* next,
                                                       uint64_t n_out, const MatRef* inj,
                                                       int n_inj, uint32_t inj_width) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const uint4* p = reinterpret_cast<const uint4*>(prev + 16 * i);
    uint32_t st[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint4 v = p[q];
        st[4 * q] = v.x; st[4 * q + 1] = v.y; st[4 * q + 2] = v.z; st[4 * q + 3] = v.w;
    }
    permute<F>(st);
    if (n_inj > 0) {
        uint32_t h[16];
#pragma unroll
        for (int t = 0; t < 16; t++) h[t] = 0;
        absorb_rows<F>(h, inj, n_inj, inj_width, i);
#pragma unroll
        for (int t = 0; t < 8; t++) st[8 + t] = h[t];
        permute<F>(st);
    }
    uint4* o = reinterpret_cast<uint4*>(next + 8 * i);
    o[0] = make_uint4(st[0], st[1], st[2], st[3]);
    o[1] = make_uint4(st[4], st[5], st[6], st[7]);
}

__global__ void open_gather_kernel(const MatRef* mats, const uint32_t* shift,
                                   const uint32_t* col_off, int n_mats, uint32_t total_width,
                                   const uint32_t* digests, const uint64_t* layer_off, int log_h,
                                   const uint64_t* idx, uint32_t* rows_out, uint32_t* paths_out) {
    const int q = blockIdx.x;
    const uint64_t index = idx[q];
    for (int k = 0; k < n_mats; k++) {
        const uint64_t r = index >> shift[k];
        const MatRef m = mats[k];
        uint32_t* dst = rows_out + (uint64_t)q * total_width + col_off[k];
        for (uint32_t c = threadIdx.x; c < m.width; c += blockDim.x) dst[c] = m.data[r * m.stride + c];
    }
    for (int t = threadIdx.x; t < log_h * 8; t += blockDim.x) {
        const int j = t >> 3, e = t & 7;
        const uint64_t sib = (index >> j) ^ 1u;
        paths_out[((uint64_t)q * log_h + j) * 8 + e] = digests[(layer_off[j] + sib) * 8 + e];
    }
}

template <class F>
__global__ void __launch_bounds__(256) compress_digest_kernel(const uint32_t* prev, const uint32_t* inj,
                                                              uint32_t* next, uint64_t n_out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const GLOBAL u32x4* p = reinterpret_cast<const GLOBAL u32x4*>((gcptr)prev + 16 * i);
    uint32_t st[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const u32x4 v = p[q];
        st[4 * q] = v.x; st[4 * q + 1] = v.y; st[4 * q + 2] = v.z; st[4 * q + 3] = v.w;
    }
    permute<F>(st);
    const GLOBAL u32x4* d = reinterpret_cast<const GLOBAL u32x4*>((gcptr)inj + 8 * i);
    const u32x4 d0 = d[0], d1 = d[1];
    st[8] = d0.x; st[9] = d0.y; st[10] = d0.z; st[11] = d0.w;
    st[12] = d1.x; st[13] = d1.y; st[14] = d1.z; st[15] = d1.w;
    permute<F>(st);
    GLOBAL u32x4* o = reinterpret_cast<GLOBAL u32x4*>((gptr)next + 8 * i);
    o[0] = u32x4{st[0], st[1], st[2], st[3]};
    o[1] = u32x4{st[4], st[5], st[6], st[7]};
}

// ---- batched single-matrix commits --------------------------------------
__device__ __forceinline__ int find_job(const uint64_t* __restrict__ prefix, int njobs, uint64_t i) {
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Leaf layer of many jobs: thread -> (job, row); rows of any width.
template <class F>
__global__ void __launch_bounds__(256) leaf_hash_batch_kernel(const LeafJob* __restrict__ jobs,
                                                              const uint64_t* __restrict__ prefix,
                                                              int njobs, uint64_t total) {
    const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const int j = find_job(prefix, njobs, gi);
    const LeafJob jb = jobs[j];
    const uint64_t row = gi - prefix[j];
    const gcptr rp = (gcptr)jb.data + row * jb.stride;
    uint32_t st[16];
#pragma unroll
    for (int t = 0; t < 16; t++) st[t] = 0;
    const uint32_t full = jb.width / 8, tail = jb.width % 8;
    if ((jb.width & 3u) == 0 && (jb.stride & 3u) == 0) {  // 16-byte aligned rows
        const GLOBAL u32x4* r4 = reinterpret_cast<const GLOBAL u32x4*>(rp);
        for (uint32_t c = 0; c < full; c++) {
            const u32x4 a = r4[2 * c], b = r4[2 * c + 1];
            st[0] = a.x; st[1] = a.y; st[2] = a.z; st[3] = a.w;
            st[4] = b.x; st[5] = b.y; st[6] = b.z; st[7] = b.w;
            permute<F>(st);
        }
    } else {
        for (uint32_t c = 0; c < full; c++) {
#pragma unroll
            for (int t = 0; t < 8; t++) st[t] = rp[8 * c + t];
            permute<F>(st);
        }
    }
    if (tail) {  // overwrite only the first `tail` state words
        const gcptr tp = rp + 8 * full;
#pragma unroll
        for (int t = 0; t < 7; t++)
            if ((uint32_t)t < tail) st[t] = tp[t];
        permute<F>(st);
    }
    GLOBAL u32x4* o = reinterpret_cast<GLOBAL u32x4*>((gptr)jb.digests + 8 * row);
    o[0] = u32x4{st[0], st[1], st[2], st[3]};
    o[1] = u32x4{st[4], st[5], st[6], st[7]};
}

template <class F>
__global__ void __launch_bounds__(256) compress_batch_kernel(const CompressJob* __restrict__ jobs,
                                                             const uint64_t* __restrict__ prefix,
                                                             int njobs, uint64_t total) {
    const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const int j = find_job(prefix, njobs, gi);
    const CompressJob jb = jobs[j];
    const uint64_t i = gi - prefix[j];
    const GLOBAL u32x4* p = reinterpret_cast<const GLOBAL u32x4*>((gcptr)jb.prev + 16 * i);
    uint32_t st[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const u32x4 v = p[q];
        st[4 * q] = v.x; st[4 * q + 1] = v.y; st[4 * q + 2] = v.z; st[4 * q + 3] = v.w;
    }
    permute<F>(st);
    GLOBAL u32x4* o = reinterpret_cast<GLOBAL u32x4*>((gptr)jb.next + 8 * i);
    o[0] = u32x4{st[0], st[1], st[2], st[3]};
    o[1] = u32x4{st[4], st[5], st[6], st[7]};
}

template <class F>
__global__ void __launch_bounds__(256) compress_batch_lanes_kernel(const CompressJob* __restrict__ jobs,
                                                                   const uint64_t* __restrict__ prefix,
                                                                   int njobs, uint64_t total) {
    const uint64_t gi = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int l = (int)(threadIdx.x & 15u);
    if (gi >= total) return;
    const int j = find_job(prefix, njobs, gi);
    const CompressJob jb = jobs[j];
    const uint64_t i = gi - prefix[j];
    const uint32_t x = permute_lanes<F>(((gcptr)jb.prev)[16 * i + l], l);
    if (l < 8) ((gptr)jb.next)[8 * i + l] = x;
}

__global__ void gather_roots_kernel(const uint32_t* const* __restrict__ src, uint32_t* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * 8) dst[i] = src[i >> 3][i & 7];
}

static unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_gather_roots(const uint32_t* const* src, uint32_t* dst, int n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    gather_roots_kernel<<<grid_for((uint64_t)n * 8, 256), 256, 0, stream>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t launch_leaf_hash_batch(int field, const LeafJob* jobs, const uint64_t* prefix, int njobs,
                                  uint64_t total, hipStream_t stream) {
    if (total == 0) return hipSuccess;
    if (field == FIELD_KOALABEAR)
        leaf_hash_batch_kernel<KoalaBear><<<grid_for(total, 256), 256, 0, stream>>>(jobs, prefix, njobs, total);
    else
        leaf_hash_batch_kernel<BabyBear><<<grid_for(total, 256), 256, 0, stream>>>(jobs, prefix, njobs, total);
    return hipGetLastError();
}

hipError_t launch_compress_batch(int field, const CompressJob* jobs, const uint64_t* prefix, int njobs,
                                 uint64_t total, hipStream_t stream) {
    if (total == 0) return hipSuccess;
    if (total <= LANES_MAX_NODES) {
        const unsigned g = grid_for(total * 16, 256);
        if (field == FIELD_KOALABEAR)
            compress_batch_lanes_kernel<KoalaBear><<<g, 256, 0, stream>>>(jobs, prefix, njobs, total);
        else
            compress_batch_lanes_kernel<BabyBear><<<g, 256, 0, stream>>>(jobs, prefix, njobs, total);
        return hipGetLastError();
    }
    if (field == FIELD_KOALABEAR)
        compress_batch_kernel<KoalaBear><<<grid_for(total, 256), 256, 0, stream>>>(jobs, prefix, njobs, total);
    else
        compress_batch_kernel<BabyBear><<<grid_for(total, 256), 256, 0, stream>>>(jobs, prefix, njobs, total);
    return hipGetLastError();
}

hipError_t launch_leaf_hash(int field, const MatRef* mats, int n, uint32_t total_width,
                            size_t rows, uint32_t* digests, hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    const unsigned g = grid_for(rows, 256);
    if (field == FIELD_KOALABEAR)
        leaf_hash_kernel<KoalaBear><<<g, 256, 0, stream>>>(mats, n, total_width, rows, digests);
    else
        leaf_hash_kernel<BabyBear><<<g, 256, 0, stream>>>(mats, n, total_width, rows, digests);
    return hipGetLastError();
}

// Fast path entry used by the host when a height group is one matrix with
// width % 8 == 0 and 32-byte aligned rows (declared here, not in the shared
// header, because only api.hip's commit planner chooses it).
hipError_t launch_absorb_segment(int field, const uint32_t* data, uint64_t stride, uint32_t width, uint32_t pos0,
                                 int first, int last, uint64_t rows, uint32_t* state, uint32_t* digests,
                                 hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    if (pos0 > 7 || (!last && !state) || (last && !digests) || (!first && !state)) return hipErrorInvalidValue;
    const bool vec = pos0 == 0 && width % 8 == 0 && stride % 4 == 0 && ((uintptr_t)data % 16) == 0;
    const unsigned g = grid_for(rows, 256);
    if (field == FIELD_KOALABEAR) {
        if (vec)
            absorb_segment_kernel<KoalaBear, true><<<g, 256, 0, stream>>>(data, stride, width, 0, first, last, rows, state, digests);
        else
            absorb_segment_kernel<KoalaBear, false><<<g, 256, 0, stream>>>(data, stride, width, pos0, first, last, rows, state, digests);
    } else {
        if (vec)
            absorb_segment_kernel<BabyBear, true><<<g, 256, 0, stream>>>(data, stride, width, 0, first, last, rows, state, digests);
        else
            absorb_segment_kernel<BabyBear, false><<<g, 256, 0, stream>>>(data, stride, width, pos0, first, last, rows, state, digests);
    }
    return hipGetLastError();
}

hipError_t launch_leaf_hash_w8(int field, const uint32_t* data, uint64_t stride, uint32_t width,
                               size_t rows, uint32_t* digests, hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    if (width == 8 && rows <= LANES_MAX_NODES) {
        const unsigned g = grid_for(rows * 16, 256);
        if (field == FIELD_KOALABEAR)
            leaf8_lanes_kernel<KoalaBear><<<g, 256, 0, stream>>>(data, stride, rows, digests);
        else
            leaf8_lanes_kernel<BabyBear><<<g, 256, 0, stream>>>(data, stride, rows, digests);
        return hipGetLastError();
    }
    const unsigned g = grid_for(rows, 256);
    if (field == FIELD_KOALABEAR)
        leaf_hash_w8_kernel<KoalaBear><<<g, 256, 0, stream>>>(data, stride, width / 8, rows, digests);
    else
        leaf_hash_w8_kernel<BabyBear><<<g, 256, 0, stream>>>(data, stride, width / 8, rows, digests);
    return hipGetLastError();
}

hipError_t launch_compress(int field, const uint32_t* prev, uint32_t* next, size_t n_out,
                           const MatRef* inj, int n_inj, uint32_t inj_width, hipStream_t stream) {
    if (n_out == 0) return hipSuccess;
    if (n_inj == 0 && n_out <= LANES_MAX_NODES) {
        const unsigned g = grid_for(n_out * 16, 256);
        if (field == FIELD_KOALABEAR)
            compress_lanes_kernel<KoalaBear><<<g, 256, 0, stream>>>(prev, next, n_out);
        else
            compress_lanes_kernel<BabyBear><<<g, 256, 0, stream>>>(prev, next, n_out);
        return hipGetLastError();
    }
    const unsigned g = grid_for(n_out, 256);
    if (field == FIELD_KOALABEAR)
        compress_kernel<KoalaBear><<<g, 256, 0, stream>>>(prev, next, n_out, inj, n_inj, inj_width);
    else
        compress_kernel<BabyBear><<<g, 256, 0, stream>>>(prev, next, n_out, inj, n_inj, inj_width);
    return hipGetLastError();
}

hipError_t launch_compress_digest(int field, const uint32_t* prev, uint32_t* next, size_t n_out,
                                  const uint32_t* inj, hipStream_t stream) {
    if (n_out == 0) return hipSuccess;
    if (n_out <= LANES_MAX_NODES) {
        const unsigned g = grid_for(n_out * 16, 256);
        if (field == FIELD_KOALABEAR)
            compress_digest_lanes_kernel<KoalaBear><<<g, 256, 0, stream>>>(prev, inj, next, n_out);
        else
            compress_digest_lanes_kernel<BabyBear><<<g, 256, 0, stream>>>(prev, inj, next, n_out);
        return hipGetLastError();
    }
    const unsigned g = grid_for(n_out, 256);
    if (field == FIELD_KOALABEAR)
        compress_digest_kernel<KoalaBear><<<g, 256, 0, stream>>>(prev, inj, next, n_out);
    else
        compress_digest_kernel<BabyBear><<<g, 256, 0, stream>>>(prev, inj, next, n_out);
    return hipGetLastError();
}

hipError_t launch_open_gather(const MatRef* mats, const uint32_t* shift, const uint32_t* col_off,
                              int n_mats, uint32_t total_width, const uint32_t* digests,
                              const uint64_t* layer_off, int log_h, const uint64_t* idx, int nq,
                              uint32_t* rows_out, uint32_t* paths_out, hipStream_t stream) {
    if (nq == 0) return hipSuccess;
    open_gather_kernel<<<nq, 128, 0, stream>>>(mats, shift, col_off, n_mats, total_width, digests,
                                               layer_off, log_h, idx, rows_out, paths_out);
    return hipGetLastError();
}

}  // namespace pg
