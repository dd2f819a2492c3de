// gfx950 kernels of one EZLDA Gibbs sweep.
//
// Stage order per sweep (reference EZLDA/mainKernel.cu:158-380):
//   zero_counts -> topic_hist (n_k) -> wt_head / wt_tail (WT from z)
//   -> inv + word_top (phi top-3) -> dt_build (DT rows + per-token aux)
//   -> sample (three-bucket draw, writes z) -> advance_sweep.
// Every float operation follows the canonical order of the CPU oracle
// (oracle/ezlda_oracle.c); builds use -ffp-contract=off, so the topics agree
// bit for bit.  Wave64: one wavefront owns a doc (dt_build) or a tail word
// (wt_tail, word_top); in the sampler one lane owns a token.
#include <hip/hip_runtime.h>

#include "ezlda_device.h"
#include "ezlda_philox.h"

// Sampler row blocking (16-byte loads per block, checkpoints per lane).
// DT rows are padded to multiples of 32 entries, so blocks up to 8 loads
// never read past a row's padding.
#ifndef EZ_SAMPLE_BLK
#define EZ_SAMPLE_BLK 4
#endif
#ifndef EZ_SAMPLE_CP
#define EZ_SAMPLE_CP 12
#endif
// Largest K whose sampler keeps a table of the Q running sums inside each
// 64-topic slab in LDS (a Q draw then bisects it); above it a Q draw
// recomputes its slab's running sums, which halves the sampler's LDS.
#ifndef EZ_Q_TABLE_MAX_K
#define EZ_Q_TABLE_MAX_K 1024
#endif
static_assert(EZ_Q_TABLE_MAX_K <= 8192, "the Q-table sampler variants assume compact (K <= 8192) DT rows");

namespace ez {

static constexpr float kMargin = 0.998046875f;  // 1 - 2^-9
static constexpr uint32_t kNone = 0xFFFFu;

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ int first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
__device__ __forceinline__ int last_lane(uint64_t m) { return 63 - __clzll((long long)m); }
__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

// Orders LDS traffic between the phases of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

static int grid_for(size_t n, int bd, int cap) {
    size_t g = (n + (size_t)bd - 1) / (size_t)bd;
    if (g < 1) g = 1;
    if (g > (size_t)cap) g = (size_t)cap;
    return (int)g;
}

static size_t wave_hist_bytes(int K) { return (size_t)((K + 1) >> 1) * sizeof(uint32_t); }
// DT rebuild: 16-bit counts plus a topic bitmap per wavefront
static size_t dt_wave_bytes(int K) { return wave_hist_bytes(K) + (size_t)((K + 31) >> 5) * sizeof(uint32_t); }

// Launch of a kernel whose wavefronts each own wb bytes of LDS (cleared
// when the wavefront starts) and stride over n units (words, docs).  When
// at least 16 wavefronts' LDS fits on a CU: workgroups of four wavefronts,
// one unit each (up to cap workgroups), which the dispatcher balances.
// Otherwise (large K, where clearing the LDS costs more than a unit):
// persistent one-wavefront workgroups, as many as the CUs' LDS holds.
struct WaveGrid {
    int grid, nwv;
};
static WaveGrid wave_grid(const DevState &s, size_t n, size_t wb, int cap) {
    const size_t fit = wb ? (size_t)s.lds_cu / wb : 32;
    if (fit >= 16) return {grid_for(n, 4, cap), 4};
    return {grid_for(n, 1, s.n_cu * (int)std::max<size_t>(1, fit)), 1};
}

// A DT entry is (topic << 2) | (count << 16): its low half is the byte
// offset of phi[topic] in a float array, so a gather needs one AND.
__device__ __forceinline__ uint32_t dt_topic(uint32_t e) { return (e & 0xFFFFu) >> 2; }
__device__ __forceinline__ float dt_count(uint32_t e) { return (float)(e >> 16); }
__device__ __forceinline__ float phi_of(const float *phi, uint32_t e) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(phi) + (e & 0xFFFCu));
}

// Compact DT rows (K <= 8192): 16-bit entries topic | piece << tb, tb =
// max(10, ceil(log2 K)) topic bits; a count stored in pieces of at most
// 2^(16 - tb) - 1.  Decoders take the low 16 bits of e.
__device__ __forceinline__ float c16_count(uint32_t e, uint32_t tb) { return (float)((e & 0xFFFFu) >> tb); }
__device__ __forceinline__ uint32_t c16_topic(uint32_t e, uint32_t tb) { return e & ((1u << tb) - 1u); }

// The same read for a phi table at LDS address 0 (the sampler's dynamic LDS
// when the kernel has no static LDS, which sampler_static_lds() lets the host
// check): the entry's low half is the address itself, no base to add.
typedef __attribute__((address_space(3))) const float lds_cfloat;
__device__ __forceinline__ float phi_at0(uint32_t e) { return *(lds_cfloat *)(e & 0xFFFCu); }
// phi of a compact entry (the 16-bit entry in the low half of e)
__device__ __forceinline__ float phi_c16(uint32_t e, uint32_t tb) { return *(lds_cfloat *)(c16_topic(e, tb) << 2); }

// 16-bit counters, two per LDS word (counts < 65536 by construction).
__device__ __forceinline__ void h16_add(uint32_t *h, uint32_t k) {
    atomicAdd(&h[k >> 1], 1u << ((k & 1u) << 4));
}

__device__ __forceinline__ uint32_t h16_get(const uint32_t *h, uint32_t k) {
    return (h[k >> 1] >> ((k & 1u) << 4)) & 0xFFFFu;
}

// ------------------------------------------------------------ count reset
__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }
__global__ void k_inc_u32(uint32_t *p) { *p = *p + 1u; }

// ------------------------------------------------------- n_k histogram
// Adds the topics of tokens [0, n) of zs to the LDS histogram sh: lanes take
// groups of eight tokens (one 16-byte load) i0, i0 + stride, ..., and add
// each run of equal neighbouring topics among them with one atomic (word-
// major neighbours often share a topic, and equal topics in one LDS
// instruction contend for one address).  The lane with i0 == 0 also adds the
// tokens before the first 16-byte boundary and after the last whole group.
__device__ __forceinline__ void hist_add_tokens(uint32_t *sh, const uint16_t *__restrict__ zs, size_t n, size_t i0,
                                                size_t stride) {
    const size_t mis = ((16u - ((uintptr_t)zs & 15u)) & 15u) >> 1;
    const size_t head = mis < n ? mis : n;
    const size_t n8 = (n - head) >> 3;
    const uint4 *z8 = reinterpret_cast<const uint4 *>(zs + head);
    for (size_t i = i0; i < n8; i += stride) {
        const uint4 v = z8[i];
        const uint32_t t[8] = {v.x & 0xFFFFu, v.x >> 16, v.y & 0xFFFFu, v.y >> 16,
                               v.z & 0xFFFFu, v.z >> 16, v.w & 0xFFFFu, v.w >> 16};
        uint32_t c = 1;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            if (t[j + 1] != t[j]) {
                atomicAdd(&sh[t[j]], c);
                c = 1;
            } else {
                ++c;
            }
        }
        atomicAdd(&sh[t[7]], c);
    }
    if (i0 == 0) {
        for (size_t i = 0; i < head; ++i) atomicAdd(&sh[zs[i]], 1u);
        for (size_t i = head + (n8 << 3); i < n; ++i) atomicAdd(&sh[zs[i]], 1u);
    }
}

// n_k = column sums of WT (WTDen_Sum_Update_Kernel utility.cu:1104-1126 plus
// the tail's atomicAdd in sparseMatrixAdd :1132-1204): one LDS histogram of
// z per workgroup, flushed with one atomic per nonzero topic.
// (block `blk` of `nblk`; the kernel below, or a role of k_wt_all)
__device__ __forceinline__ void topic_hist_block(uint32_t *sh, const uint16_t *__restrict__ z, uint32_t N, int K,
                                                 int32_t *__restrict__ nk, uint32_t blk, uint32_t nblk) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    hist_add_tokens(sh, z, N, (size_t)blk * blockDim.x + threadIdx.x, (size_t)nblk * blockDim.x);
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        uint32_t c = sh[k];
        if (c) atomicAdd(&nk[k], (int32_t)c);
    }
}
__global__ void k_topic_hist(const uint16_t *__restrict__ z, uint32_t N, int K, int32_t *__restrict__ nk) {
    extern __shared__ uint32_t sh[];
    topic_hist_block(sh, z, N, K, nk, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------- head WT
// Dense head rows (WTDen_Update_Kernel, utility.cu:1067-1102): one workgroup
// per segment of a head word's tokens, LDS histogram, atomic flush into the
// (pre-zeroed) row.
__device__ __forceinline__ void wt_head_block(uint32_t *sh, const uint16_t *__restrict__ z,
                                              const uint4 *__restrict__ segs, int K, int32_t *__restrict__ wt_head,
                                              uint32_t seg) {
    const uint4 sg = segs[seg];
    for (int k = threadIdx.x; k < K; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    hist_add_tokens(sh, z + sg.y, sg.z - sg.y, threadIdx.x, blockDim.x);
    __syncthreads();
    int32_t *row = wt_head + (size_t)sg.x * K;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        uint32_t c = sh[k];
        if (c) atomicAdd(&row[k], (int32_t)c);
    }
}
__global__ void k_wt_head(const uint16_t *__restrict__ z, const uint4 *__restrict__ segs, int K,
                          int32_t *__restrict__ wt_head) {
    extern __shared__ uint32_t sh[];
    wt_head_block(sh, z, segs, K, wt_head, blockIdx.x);
}

// ------------------------------------------------------------- tail WT
// Sorted sparse tail rows (WT_Update_Kernel, utility.cu:10-76): one
// wavefront per tail word; a 16-bit LDS histogram and a bitmap of the
// word's topics, then a compaction over the bitmap in ascending topic order
// (lane j owns topics [32j, 32j + 32) of each 64-word chunk and takes its
// place from a wave prefix sum), clearing its counts and bitmap as it goes.
__device__ __forceinline__ void wt_tail_block(uint32_t *sh, const uint16_t *__restrict__ z,
                                              const uint32_t *__restrict__ word_off, int VD, int V, int K,
                                              const uint32_t *__restrict__ tail_off, uint32_t *__restrict__ wt_tail,
                                              uint32_t *__restrict__ tail_nnz, int blk, int nblk) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int KH = (K + 1) >> 1, NW = (K + 31) >> 5;
    uint32_t *h = sh + wv * (KH + NW);  // 16-bit counts
    uint32_t *bm = h + KH;              // bitmap of the word's topics
    for (int k = lane; k < KH + NW; k += 64) h[k] = 0;
    wave_lds_sync();
    const int ntail = V - VD;
    for (int tw = blk * nwv + wv; tw < ntail; tw += nblk * nwv) {
        const int w = VD + tw;
        const uint32_t b = word_off[w], e = word_off[w + 1];
        if (b == e) {
            if (lane == 0) tail_nnz[tw] = 0;
            continue;
        }
        for (uint32_t i = b + lane; i < e; i += 64) {
            const uint32_t t = z[i];
            h16_add(h, t);
            atomicOr(&bm[t >> 5], 1u << (t & 31u));
        }
        wave_lds_sync();
        const uint32_t base = tail_off[tw];
        uint32_t nz = 0;
        for (int j = 0; j < NW; j += 64) {
            const int wi = j + lane;
            const uint32_t word = wi < NW ? bm[wi] : 0u;
            const uint32_t m = (uint32_t)__popc(word);
            uint32_t incl = m;  // inclusive scan over lanes
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t v = __shfl_up(incl, d);
                if (lane >= d) incl += v;
            }
            uint32_t pos = nz + incl - m;
            for (uint32_t x = word; x; x &= x - 1u) {
                const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                wt_tail[base + pos++] = k | (h16_get(h, k) << 16);
            }
            for (uint32_t x = word; x; x &= x - 1u) {  // the lane's own 16 count words
                const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                h[k >> 1] = 0;
            }
            if (wi < NW) bm[wi] = 0;
            nz += __shfl(incl, 63);
        }
        wave_lds_sync();
        if (lane == 0) tail_nnz[tw] = nz;
    }
}
__global__ void k_wt_tail(const uint16_t *__restrict__ z, const uint32_t *__restrict__ word_off, int VD, int V,
                          int K, const uint32_t *__restrict__ tail_off, uint32_t *__restrict__ wt_tail,
                          uint32_t *__restrict__ tail_nnz) {
    extern __shared__ uint32_t sh[];
    wt_tail_block(sh, z, word_off, VD, V, K, tail_off, wt_tail, tail_nnz, blockIdx.x, gridDim.x);
}

// The three count builds of a sweep in one launch (they touch disjoint
// outputs): blocks [0, nb_hist) build n_k, the next nb_head the head rows,
// the rest the tail rows, so each role's last blocks overlap the next role's
// instead of draining the device between launches.
__global__ void __launch_bounds__(256) k_wt_all(int nb_hist, int nb_head, const uint16_t *__restrict__ z,
                                                uint32_t N, int K, int32_t *__restrict__ nk,
                                                const uint4 *__restrict__ segs, int32_t *__restrict__ wt_head,
                                                const uint32_t *__restrict__ word_off, int VD, int V,
                                                const uint32_t *__restrict__ tail_off, uint32_t *__restrict__ wt_tail,
                                                uint32_t *__restrict__ tail_nnz) {
    extern __shared__ uint32_t sh[];
    const int b = (int)blockIdx.x;
    if (b < nb_hist)
        topic_hist_block(sh, z, N, K, nk, (uint32_t)b, (uint32_t)nb_hist);
    else if (b < nb_hist + nb_head)
        wt_head_block(sh, z, segs, K, wt_head, (uint32_t)(b - nb_hist));
    else
        wt_tail_block(sh, z, word_off, VD, V, K, tail_off, wt_tail, tail_nnz, b - nb_hist - nb_head,
                      (int)gridDim.x - nb_hist - nb_head);
}

// ----------------------------------------------------- inv and top-3
__global__ void k_inv(const int32_t *__restrict__ nk, int K, float vbeta, float *__restrict__ inv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) inv[k] = 1.0f / ((float)nk[k] + vbeta);
}

struct Top3 {
    float v0, v1, v2;
    uint32_t k0, k1, k2;
};

__device__ __forceinline__ bool better(float va, uint32_t ka, float vb, uint32_t kb) {
    return va > vb || (va == vb && ka < kb);
}

// A key whose unsigned order is better()'s: value bits (v >= 0, so their
// order is the value's) above the complemented topic; empty (v < 0) is 0.
__device__ __forceinline__ uint64_t top_key(float v, uint32_t k) {
    return v < 0.0f ? 0ull : ((uint64_t)(__float_as_uint(v) + 1u) << 32) | (uint64_t)(0xFFFFFFFFu - k);
}

__device__ __forceinline__ void top3_insert(Top3 &t, float v, uint32_t k) {
    if (!better(v, k, t.v2, t.k2)) return;
    if (better(v, k, t.v1, t.k1)) {
        t.v2 = t.v1; t.k2 = t.k1;
        if (better(v, k, t.v0, t.k0)) {
            t.v1 = t.v0; t.k1 = t.k0;
            t.v0 = v; t.k0 = k;
        } else {
            t.v1 = v; t.k1 = k;
        }
    } else {
        t.v2 = v; t.k2 = k;
    }
}

// Topics by beta * inv[k] (the value of a topic a word has no count of),
// largest first, ties to the lower topic.  A workgroup of four wavefronts
// ranks 64 topics: lane i's topic is compared with a quarter of all topics
// by each wavefront, and the four counts are added in LDS.
__global__ void __launch_bounds__(256) k_zero_order(const float *__restrict__ inv, int K, float beta,
                                                    uint32_t *__restrict__ zero_order) {
    extern __shared__ float sv[];  // K values, then 64 ranks
    uint32_t *rank = reinterpret_cast<uint32_t *>(sv + K);
    for (int k = threadIdx.x; k < K; k += blockDim.x) sv[k] = ((float)0u + beta) * inv[k];
    if (threadIdx.x < 64) rank[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int q = (K + 3) >> 2, j0 = wv * q, j1 = min(K, j0 + q);
    if (k < K) {
        const float v = sv[k];
        uint32_t r = 0;
        for (int j = j0; j < j1; ++j) r += better(sv[j], (uint32_t)j, v, (uint32_t)k) ? 1u : 0u;
        atomicAdd(&rank[lane], r);
    }
    __syncthreads();
    if (wv == 0 && k < K) zero_order[rank[lane]] = (uint32_t)k;
}

// Per word the three largest phi_wk, ties to the lower topic
// (MaxTopicDense/Sparse_Update_Kernel, utility.cu:305-678, :679-1062).  One
// wavefront per word; a tail word's counts are scattered into a 16-bit LDS
// row first.  Per-lane lists are merged with a butterfly; the order is total,
// so the merge order does not matter.
// Words [w0, w1) of one launch.  Only tail words use the LDS count table:
// a launch with w0 < VD and no LDS (w1 <= VD) runs head words at full
// occupancy; a launch that reaches tail words has the table.
__global__ void k_word_top(int w0, int w1, int VD, int K, float beta, const int32_t *__restrict__ wt_head,
                           const uint32_t *__restrict__ tail_off, const uint32_t *__restrict__ wt_tail,
                           const uint32_t *__restrict__ tail_nnz, const float *__restrict__ inv,
                           const uint32_t *__restrict__ zero_order, ushort4 *__restrict__ wtop) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int KH = (K + 1) >> 1;
    uint32_t *h = sh + wv * KH;
    if (w1 > VD) {
        for (int k = lane; k < KH; k += 64) h[k] = 0;
        wave_lds_sync();
    }
    for (int w = w0 + blockIdx.x * nwv + wv; w < w1; w += gridDim.x * nwv) {
        Top3 t{-1.f, -1.f, -1.f, kNone, kNone, kNone};
        if (w < VD) {
            // four loads of the row and of inv in flight per lane
            const int32_t *row = wt_head + (size_t)w * K;
            for (int k0 = lane; k0 < K; k0 += 256) {
                int32_t c[4];
                float iv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + 64 * u;
                    c[u] = k < K ? row[k] : 0;
                    iv[u] = k < K ? inv[k] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (k0 + 64 * u < K) top3_insert(t, ((float)c[u] + beta) * iv[u], (uint32_t)(k0 + 64 * u));
            }
        } else {
            // the word's topics with a count, then those without one: their
            // value is beta * inv[k], so the first three of zero_order that
            // have no count are the best of them
            const int tw = w - VD;
            const uint32_t b = tail_off[tw], n = tail_nnz[tw];
            for (uint32_t i = lane; i < n; i += 64) {
                const uint32_t e = wt_tail[b + i];
                const uint32_t k = e & 0xFFFFu;
                atomicAdd(&h[k >> 1], (e >> 16) << ((k & 1u) << 4));
                top3_insert(t, ((float)(e >> 16) + beta) * inv[k], k);
            }
            wave_lds_sync();
            uint32_t got = 0;  // wavefront-uniform
            for (int j0 = 0; j0 < K && got < 3u; j0 += 64) {
                const int j = j0 + lane;
                const uint32_t k = j < K ? zero_order[j] : 0u;
                const bool free_k = j < K && h16_get(h, k) == 0u;
                const uint64_t m = __ballot(free_k);
                if (free_k && got + (uint32_t)__popcll(m & lanemask_lt(lane)) < 3u)
                    top3_insert(t, ((float)0u + beta) * inv[k], k);
                got += (uint32_t)__popcll(m);
            }
            wave_lds_sync();
            for (uint32_t i = lane; i < n; i += 64) h[(wt_tail[b + i] & 0xFFFFu) >> 1] = 0;
            wave_lds_sync();
        }
        // the wavefront's best three: three max-reductions of the lanes'
        // list heads, the winner's lane popping its head each time (a lane's
        // topics are its own, so the winner is one lane)
        uint64_t a0 = top_key(t.v0, t.k0), a1 = top_key(t.v1, t.k1), a2 = top_key(t.v2, t.k2);
        uint32_t best[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            uint64_t m = a0;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const uint64_t o = __shfl_xor(m, off);
                m = o > m ? o : m;
            }
            best[r] = m ? 0xFFFFFFFFu - (uint32_t)m : kNone;
            if (a0 == m) {
                a0 = a1;
                a1 = a2;
                a2 = 0;
            }
        }
        if (lane == 0)
            wtop[w] = make_ushort4((unsigned short)best[0], (unsigned short)best[1], (unsigned short)best[2], 0);
    }
}

// Doc-slot copy of z from z (after a layout change: slots of words that
// were tail words hold stale topics there).
__global__ void k_z_to_doc(uint32_t N, const uint32_t *__restrict__ slot_tok, const uint16_t *__restrict__ z,
                           uint16_t *__restrict__ z_doc) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x)
        z_doc[p] = z[slot_tok[p]];
}

hipError_t launch_z_to_doc(const DevState &s, hipStream_t st) {
    if (s.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_z_to_doc, dim3(grid_for(s.N, 256, 1 << 16)), dim3(256), 0, st, s.N, s.slot_tok, s.z, s.z_doc);
    return hipGetLastError();
}

// --------------------------------------------------- DT rows and aux
// DT_Update_Kernel (utility.cu:79-284), on the doc-major copy of z: doc q
// (docs in descending length order) owns slots [soff[q], soff[q+1]), its
// tokens grouped by word, and the sampler keeps z_doc in step with z, so
// every read and write here is a coalesced stream.  One wavefront per doc
// builds a 16-bit LDS histogram, compacts the row in ascending topic order
// (:183-238), and writes each slot's record: n_dk1 at its word's top topic
// with the row's length (the sampler's), or with n_dk2 = n_dk at its word's
// second topic (the re-pack of :125-161, for the getter).  A row starts on a
// 128-byte boundary and is zero-padded to a multiple of 32 entries.
static constexpr int kDtU = 8;  // slots per lane per chunk in k_dt_build

template <bool COMPACT, bool N_DK2>
__global__ void k_dt_build(int q0, int q1, int K, int VD, uint32_t tb, const uint32_t *__restrict__ soff,
                           const uint32_t *__restrict__ slot_wrd, const uint32_t *__restrict__ slot_tok,
                           const uint16_t *__restrict__ z, const uint16_t *__restrict__ z_doc,
                           const ushort4 *__restrict__ wtop,
                           uint32_t *__restrict__ aux_doc, uint2 *__restrict__ doc_info, uint32_t *__restrict__ dt) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int KH = (K + 1) >> 1, NW = (K + 31) >> 5;
    const uint32_t piece = (1u << (16u - tb)) - 1u;  // compact rows: largest count of one entry
    uint32_t *h = sh + wv * (KH + NW);  // 16-bit counts
    uint32_t *bm = h + KH;              // bitmap of the doc's topics
    for (int k = lane; k < KH + NW; k += 64) h[k] = 0;
    wave_lds_sync();
    for (int q = q0 + blockIdx.x * nwv + wv; q < q1; q += gridDim.x * nwv) {
        const uint32_t b = soff[q], e = soff[q + 1];
        // The doc in chunks of kDtU slots per lane, every load of a chunk
        // issued before the first is used.  A doc of one chunk keeps its
        // words' top topics in registers for the aux pass.
        const bool one = e - b <= 64u * kDtU;
        uint32_t tp[kDtU];  // top-1 | top-2 << 16 of each slot's word (one-chunk docs)
        for (uint32_t c = b; c < e; c += 64u * kDtU) {
            uint32_t w[kDtU], t[kDtU];
#pragma unroll
            for (int u = 0; u < kDtU; ++u) {
                const uint32_t i = c + lane + 64u * u;
                w[u] = i < e ? slot_wrd[i] : 0u;
                t[u] = i < e ? z_doc[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kDtU; ++u) {
                const uint32_t i = c + lane + 64u * u;
                if (i < e && w[u] >= (uint32_t)VD) t[u] = z[slot_tok[i]];
            }
            if (one) {
#pragma unroll
                for (int u = 0; u < kDtU; ++u) {
                    const ushort4 x = wtop[w[u]];
                    tp[u] = x.x | ((uint32_t)x.y << 16);
                }
            }
#pragma unroll
            for (int u = 0; u < kDtU; ++u) {
                if (c + lane + 64u * u < e) {
                    h16_add(h, t[u]);
                    atomicOr(&bm[t[u] >> 5], 1u << (t[u] & 31u));
                }
            }
        }
        wave_lds_sync();
        // Compaction over the bitmap: lane j owns topics [32j, 32j + 32) of
        // each 64-word chunk, counts its entries, takes its place from a
        // wave prefix sum and writes them in ascending topic order.  Cost
        // follows the doc's topics, not K.
        const uint32_t base = doc_info[q].x;
        uint32_t nz = 0;
        for (int j = 0; j < NW; j += 64) {
            const int wi = j + lane;
            const uint32_t word = wi < NW ? bm[wi] : 0u;
            uint32_t m = 0;  // entries of this lane's topics
            if (COMPACT) {
                for (uint32_t x = word; x; x &= x - 1u) {
                    const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                    m += (h16_get(h, k) + piece - 1u) / piece;
                }
            } else {
                m = (uint32_t)__popc(word);
            }
            uint32_t incl = m;  // inclusive scan over lanes
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t v = __shfl_up(incl, d);
                if (lane >= d) incl += v;
            }
            uint32_t pos = nz + incl - m;
            for (uint32_t x = word; x; x &= x - 1u) {
                const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                uint32_t c = h16_get(h, k);
                if (COMPACT) {
                    uint16_t *dt16 = reinterpret_cast<uint16_t *>(dt) + base;
                    for (; c > piece; c -= piece) dt16[pos++] = (uint16_t)(k | (piece << tb));
                    dt16[pos++] = (uint16_t)(k | (c << tb));
                } else {
                    dt[base + pos++] = (k << 2) | (c << 16);
                }
            }
            nz += __shfl(incl, 63);
        }
        // zero entries up to the next 128 bytes: the sampler reads rows in
        // 128-byte blocks and a zero entry adds nothing
        if (e > b) {
            if (COMPACT) {
                uint16_t *dt16 = reinterpret_cast<uint16_t *>(dt) + base;
                if ((uint32_t)lane < ((64u - (nz & 63u)) & 63u)) dt16[nz + lane] = 0;
            } else {
                if ((uint32_t)lane < ((32u - (nz & 31u)) & 31u)) dt[base + nz + lane] = 0u;
            }
        }
        if (lane == 0) doc_info[q].y = nz | ((e - b) << 16);
        // each slot's record: n_dk1 and the row's length (a tail word's run
        // is in a random doc, whose doc_info would cost a cache line of its
        // own), or n_dk1 and n_dk2
        if (one) {
#pragma unroll
            for (int u = 0; u < kDtU; ++u) {
                const uint32_t i = b + lane + 64u * u;
                if (i < e) {
                    const uint32_t k2 = tp[u] >> 16;
                    const uint32_t hi = N_DK2 ? (k2 != kNone ? h16_get(h, k2) : 0u) : nz;
                    aux_doc[i] = h16_get(h, tp[u] & 0xFFFFu) | (hi << 16);
                }
            }
        } else {
            for (uint32_t c = b; c < e; c += 64u * kDtU) {
                uint32_t x[kDtU];
#pragma unroll
                for (int u = 0; u < kDtU; ++u) {
                    const uint32_t i = c + lane + 64u * u;
                    x[u] = i < e ? slot_wrd[i] : 0u;
                }
#pragma unroll
                for (int u = 0; u < kDtU; ++u) {
                    const ushort4 y = wtop[x[u]];
                    x[u] = y.x | ((uint32_t)y.y << 16);
                }
#pragma unroll
                for (int u = 0; u < kDtU; ++u) {
                    const uint32_t i = c + lane + 64u * u;
                    if (i < e) {
                        const uint32_t k2 = x[u] >> 16;
                        const uint32_t hi = N_DK2 ? (k2 != kNone ? h16_get(h, k2) : 0u) : nz;
                        aux_doc[i] = h16_get(h, x[u] & 0xFFFFu) | (hi << 16);
                    }
                }
            }
        }
        wave_lds_sync();
        // each lane clears its own count words and bitmap words
        for (int wi = lane; wi < NW; wi += 64) {
            for (uint32_t x = bm[wi]; x; x &= x - 1u) {
                const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                h[k >> 1] = 0;
            }
            bm[wi] = 0;
        }
        wave_lds_sync();
    }
}

// The same DT rows and aux for short documents (n <= kDtShort slots) at
// large K, where a K-sized LDS histogram per wavefront limits a CU to a few
// wavefronts: one wavefront per doc sorts its topics in 2 KB of LDS (a
// bitonic network over the next power of two >= n, 64..512 keys), reads the
// runs of equal topics off the sorted keys (ascending topic, as the
// histogram's compaction writes them) and finds each slot's (n_dk1, n_dk2)
// by bisection.  Rows, pieces, padding and aux are those of k_dt_build.

template <bool COMPACT, bool N_DK2>
__global__ void __launch_bounds__(256) k_dt_sort(int q0, int q1, int VD, uint32_t tb,
                                                 const uint32_t *__restrict__ soff, const uint32_t *__restrict__ slot_wrd,
                                                 const uint32_t *__restrict__ slot_tok, const uint16_t *__restrict__ z,
                                                 const uint16_t *__restrict__ z_doc, const ushort4 *__restrict__ wtop,
                                                 uint32_t *__restrict__ aux_doc,
                                                 uint2 *__restrict__ doc_info, uint32_t *__restrict__ dt) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    uint16_t *key = reinterpret_cast<uint16_t *>(sh + wv * kDtShort);  // kDtShort sorted topics
    uint16_t *run = key + kDtShort;                                      // run length at each run's first key
    const uint32_t piece = (1u << (16u - tb)) - 1u;
    constexpr int kU = kDtShort / 64;
    for (int q = q0 + blockIdx.x * nwv + wv; q < q1; q += gridDim.x * nwv) {
        const uint32_t b = soff[q], e = soff[q + 1], n = e - b;  // n <= kDtShort
        uint32_t P = 64;
        while (P < n) P <<= 1;
        const int nu = (int)(P >> 6);
        uint32_t wrd[kU];
        // topics (the doc-slot copy for head words, z by token id for tail
        // words), padded with 0xFFFF, which sorts last
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t i = (uint32_t)lane + 64u * u;
            wrd[u] = i < n ? slot_wrd[b + i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t i = (uint32_t)lane + 64u * u;
            if (u < nu) {
                uint32_t t = 0xFFFFu;
                if (i < n) t = wrd[u] < (uint32_t)VD ? z_doc[b + i] : z[slot_tok[b + i]];
                key[i] = (uint16_t)t;
            }
        }
        wave_lds_sync();
        // bitonic sort, ascending: P / 2 compare-exchanges per pass
        for (uint32_t k = 2; k <= P; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t p = (uint32_t)lane; p < (P >> 1); p += 64) {
                    const uint32_t i = (p / j) * (j << 1) + (p % j), l = i + j;
                    const uint16_t a = key[i], c = key[l];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up && a != c) {
                        key[i] = c;
                        key[l] = a;
                    }
                }
                wave_lds_sync();
            }
        // run heads and their lengths: the next head (or n) minus the head
        uint64_t hm[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t i = (uint32_t)lane + 64u * u;
            const bool head = u < nu && i < n && (i == 0 || key[i - 1] != key[i]);
            hm[u] = __ballot(head);
        }
        const uint32_t base = doc_info[q].x;
        uint32_t nz = 0;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (u >= nu) break;
            const uint32_t i = (uint32_t)lane + 64u * u;
            const bool head = (hm[u] >> lane) & 1ull;
            uint32_t len = 0, m = 0;
            if (head) {
                uint32_t nxt = n;  // the first head after i
                const uint64_t above = hm[u] & ~((2ull << lane) - 1ull);
                if (above) {
                    nxt = 64u * u + (uint32_t)first_lane(above);
                } else {
#pragma unroll
                    for (int v = kU - 1; v > u; --v)
                        if (v < nu && hm[v]) nxt = 64u * v + (uint32_t)first_lane(hm[v]);
                }
                len = nxt - i;
                m = COMPACT ? (len + piece - 1u) / piece : 1u;
                run[i] = (uint16_t)len;
            }
            uint32_t incl = m;  // inclusive scan over lanes
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t v = __shfl_up(incl, d);
                if (lane >= d) incl += v;
            }
            uint32_t pos = nz + incl - m;
            if (head) {
                const uint32_t k = key[i];
                if (COMPACT) {
                    uint16_t *dt16 = reinterpret_cast<uint16_t *>(dt) + base;
                    uint32_t c = len;
                    for (; c > piece; c -= piece) dt16[pos++] = (uint16_t)(k | (piece << tb));
                    dt16[pos++] = (uint16_t)(k | (c << tb));
                } else {
                    dt[base + pos] = (k << 2) | (len << 16);
                }
            }
            nz += __shfl(incl, 63);
        }
        // zero entries up to the next 128 bytes, as k_dt_build
        if (n > 0) {
            if (COMPACT) {
                uint16_t *dt16 = reinterpret_cast<uint16_t *>(dt) + base;
                if ((uint32_t)lane < ((64u - (nz & 63u)) & 63u)) dt16[nz + lane] = 0;
            } else {
                if ((uint32_t)lane < ((32u - (nz & 31u)) & 31u)) dt[base + nz + lane] = 0u;
            }
        }
        if (lane == 0) doc_info[q].y = nz | (n << 16);
        wave_lds_sync();  // run lengths visible
        // (n_dk1, n_dk2) of each slot's word: its top topics' runs, found by
        // bisection of the sorted keys
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t i = (uint32_t)lane + 64u * u;
            if (u < nu && i < n) {
                const ushort4 tp = wtop[wrd[u]];
                uint32_t cnt[2] = {0u, 0u};
                const uint32_t want[2] = {tp.x, tp.y};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (want[h] == kNone) continue;
                    uint32_t lo = 0, hi = n;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (key[mid] < want[h]) lo = mid + 1; else hi = mid;
                    }
                    if (lo < n && key[lo] == want[h]) cnt[h] = run[lo];
                }
                aux_doc[b + i] = cnt[0] | ((N_DK2 ? cnt[1] : nz) << 16);
            }
        }
        wave_lds_sync();  // the next doc reuses the keys
    }
}

// ------------------------------------------------------------- sampler
// Checkpoints of the S running sum kept per lane (one per block of kBlkU4
// 16-byte loads, i.e. 4*kBlkU4 DT entries, and with kHalfCp, in 16-bit rows,
// one more in the middle of each block, so the S-bucket search re-reads half a
// block: 6.09 -> 5.98 ms a sweep on the 99M-token corpus at K = 1024; with
// 32-bit rows it cost 1.5 ms of 63.6 at K = 10000, so they search whole blocks).
static constexpr int kCP = EZ_SAMPLE_CP;
static constexpr int kBlkU4 = EZ_SAMPLE_BLK;
#ifndef EZ_SAMPLE_HALF_CP
#define EZ_SAMPLE_HALF_CP 1
#endif
static constexpr bool kHalfCp = EZ_SAMPLE_HALF_CP != 0 && kBlkU4 % 2 == 0;
// Uniforms kept in registers per run between the shortcut test and the draw
// (tokens past them draw theirs again).
// (32-bit rows keep one: at K > 8192 the registers of four cost more than the
// draws they save, 63.9 against 63.5 ms a sweep on 180M tokens at K = 10000)
static constexpr int kUKeep = 4;    // 16-bit rows
static constexpr int kUKeep32 = 1;  // 32-bit rows

// Running FMA sum over one block of a DT row: kBlkU4 16-byte loads issued
// together, then the entries in order.
struct RowBlock {
    uint4 e[kBlkU4];
};

__device__ __forceinline__ RowBlock load_block(const uint4 *__restrict__ p) {
    RowBlock r;
#pragma unroll
    for (int i = 0; i < kBlkU4; ++i) r.e[i] = p[i];
    return r;
}

// The two compact entries of a dword (low half first) decoded together: both
// counts by one packed 16-bit shift, both topics by one mask (tm2 = the topic
// mask in each half), each phi address (topic * 4) by one shift that reads
// its half of the word (SDWA word select; the compiler forms it for the high
// half by itself, the low half is written out).
typedef unsigned short ez_us2 __attribute__((ext_vector_type(2)));
struct C16x2 {
    float c0, c1;
    uint32_t a0, a1;
};
__device__ __forceinline__ uint32_t lo16_times4(uint32_t t) {
    uint32_t r;
    asm("v_lshlrev_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
        : "=v"(r)
        : "v"(2u), "v"(t));
    return r;
}
__device__ __forceinline__ C16x2 c16_x2(uint32_t w, uint32_t tb, uint32_t tm2) {
    const ez_us2 c = __builtin_bit_cast(ez_us2, w) >> (ez_us2)((unsigned short)tb);
    const uint32_t t = w & tm2;
    return C16x2{(float)c.x, (float)c.y, lo16_times4(t), (t >> 16) << 2};
}
__device__ __forceinline__ float phi_a(uint32_t a) { return *(lds_cfloat *)a; }

// The entries of one 16-byte load, in order.
template <bool COMPACT>
__device__ __forceinline__ float sum_u4(const uint4 &e, float s, uint32_t tb) {
    if (COMPACT) {  // two 16-bit entries per dword, low half first
        const uint32_t w[4] = {e.x, e.y, e.z, e.w};
        const uint32_t tm2 = ((1u << tb) - 1u) * 0x10001u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const C16x2 d = c16_x2(w[q], tb, tm2);
            s = __builtin_fmaf(d.c0, phi_a(d.a0), s);
            s = __builtin_fmaf(d.c1, phi_a(d.a1), s);
        }
    } else {
        s = __builtin_fmaf(dt_count(e.x), phi_at0(e.x), s);
        s = __builtin_fmaf(dt_count(e.y), phi_at0(e.y), s);
        s = __builtin_fmaf(dt_count(e.z), phi_at0(e.z), s);
        s = __builtin_fmaf(dt_count(e.w), phi_at0(e.w), s);
    }
    return s;
}

// The block's entries in order; `half` gets the running sum after its
// first kBlkU4 / 2 loads.
template <bool COMPACT>
__device__ __forceinline__ float sum_block(const RowBlock &r, float s, uint32_t tb, float &half) {
#pragma unroll
    for (int i = 0; i < kBlkU4; ++i) {
        s = sum_u4<COMPACT>(r.e[i], s, tb);
        if (i == kBlkU4 / 2 - 1) half = s;
    }
    return s;
}

// S-bucket search over n consecutive 16-byte loads of a row (issued
// together) from running sum `run`: the first entry whose running sum
// exceeds t2 (running sums never decrease).
template <bool COMPACT, int N>
__device__ __forceinline__ void search_u4(const uint4 *__restrict__ p, float &run, float t2, uint32_t tb,
                                          uint32_t &newz, bool &found) {
    uint4 blk[N];
#pragma unroll
    for (int i = 0; i < N; ++i) blk[i] = p[i];
#pragma unroll
    for (int gi = 0; gi < N; ++gi) {
        const uint4 e = blk[gi];
        if (COMPACT) {
            const uint32_t w[4] = {e.x, e.y, e.z, e.w};
            const uint32_t tm2 = ((1u << tb) - 1u) * 0x10001u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const C16x2 d = c16_x2(w[q], tb, tm2);
                const float ra = __builtin_fmaf(d.c0, phi_a(d.a0), run);
                const float rb = __builtin_fmaf(d.c1, phi_a(d.a1), ra);
                if (!found && ra > t2) {
                    newz = d.a0 >> 2;
                    found = true;
                } else if (!found && rb > t2) {
                    newz = d.a1 >> 2;
                    found = true;
                }
                run = rb;
            }
        } else if (!found) {
            const float r0 = __builtin_fmaf(dt_count(e.x), phi_at0(e.x), run);
            const float r1 = __builtin_fmaf(dt_count(e.y), phi_at0(e.y), r0);
            const float r2 = __builtin_fmaf(dt_count(e.z), phi_at0(e.z), r1);
            const float r3 = __builtin_fmaf(dt_count(e.w), phi_at0(e.w), r2);
            if (r3 > t2) {
                newz = dt_topic(r0 > t2 ? e.x : r1 > t2 ? e.y : r2 > t2 ? e.z : e.w);
                found = true;
            }
            run = r3;
        }
    }
}

// Sampler work item: a word, or a doc tile / token segment of a head word,
// given as a range of runs, and the offset from the layout's token ids to
// the corpus-wide ones that key the draws (0 unless the corpus is chunked).  A run is up to run_cap consecutive tokens of the
// word in one document (tokens of a word are numbered by document, so the
// repeats of a word in a document are adjacent); its tokens share n_d, the
// max-topic counts and the DT row, hence the S sum and its checkpoints.  One
// lane per run: the shortcut test of each token, then, if any token needs
// the full draw, one pass over the doc's DT row (16 bytes per load) for S and
// the checkpoints, and per such token the bucket choice and either a re-read
// of one row block to the first running sum above the target (S) or a
// bisection of the monotone Q running sums (Q).  Every token's draw is the
// one it would get on its own; only the row scan is shared.
// Replaces UpdateProbKernelTrainD1 (utility.cu:2680-3312), LDAKernelTrainD
// (:1625-2103) and LDAKernelTrain (:1217-1611); one kernel serves head and
// tail words, which differ only in how phi_w is staged.
struct SampleArgs {
    const uint4 *segs;
    const uint4 *runs;  // (first token, first doc slot, the doc's DT row offset, tokens | doc length << 16)
    uint32_t n_items;
    int VD, K;
    float alpha, beta;
    const int32_t *wt_head;
    const uint32_t *tail_off, *wt_tail, *tail_nnz;
    const float *inv;
    const ushort4 *wtop;
    const uint32_t *aux_doc;  // by doc slot: n_dk1 | the doc's row length (nnz) << 16
    const uint32_t *dt;
    uint16_t *z, *z_doc;      // by token (word-major) and by doc slot
    const uint32_t *counters;
    uint32_t *queue;  // persistent mode: queue q's item counter at queue[q * kQueueStride]
    uint32_t n_queues;
    uint32_t tb;  // compact rows: topic bits
    unsigned long long *stats;
    uint64_t seed;
    uint32_t mode;  // kSampleHead | kSampleTail: which words' items; | kShortcutOnly: no full draws
};

template <bool SHORTCUT, bool COMPACT, bool QTAB>
__device__ __forceinline__ void sample_item(const SampleArgs &A, uint32_t item, float *smf, float *f123,
                                            uint32_t sweep, uint32_t &n_full, uint32_t &n_scanned) {
    const int K = A.K, NS = (K + 63) >> 6;
    const float alpha = A.alpha, beta = A.beta;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
    constexpr bool qtab = QTAB;  // K <= EZ_Q_TABLE_MAX_K
    float *phi = smf;                         // NS*64
    float *qs = phi + NS * 64;                // NS*64 when qtab: running sums inside each slab
    float *qt = qs + (qtab ? NS * 64 : 0);    // NS: slab totals
    float *qacc = qt + NS;                    // NS: running sum of the slab totals
    const uint4 sg = A.segs[item];
    if (sg.x == 0xFFFFFFFFu) return;  // padding item (uniform over the block)
    const int w = (int)sg.x;
    if (!(A.mode & (w < A.VD ? kSampleHead : kSampleTail))) return;  // the other words' stage
    const bool sc_only = (A.mode & kShortcutOnly) != 0;
    const ushort4 tp = A.wtop[w];

    if (w < A.VD) {
        const int32_t *row = A.wt_head + (size_t)w * K;
        for (int k = tid; k < NS * 64; k += blockDim.x) phi[k] = k < K ? ((float)row[k] + beta) * A.inv[k] : 0.0f;
    } else {
        for (int k = tid; k < NS * 64; k += blockDim.x) phi[k] = k < K ? beta * A.inv[k] : 0.0f;
        __syncthreads();
        const int tw = w - A.VD;
        const uint32_t b = A.tail_off[tw], n = A.tail_nnz[tw];
        for (uint32_t i = tid; i < n; i += blockDim.x) {
            const uint32_t e = A.wt_tail[b + i];
            const uint32_t k = e & 0xFFFFu;
            phi[k] = ((float)(e >> 16) + beta) * A.inv[k];
        }
    }
    __syncthreads();
    if (tid == 0) {
        f123[0] = phi[tp.x];
        f123[1] = tp.y != kNone ? phi[tp.y] : 0.0f;
        f123[2] = tp.z != kNone ? phi[tp.z] : 0.0f;
        phi[tp.x] = 0.0f;
    }
    __syncthreads();
    // Q: the running sums inside each 64-topic slab, left to right (phi is 0
    // past K, which leaves the sum unchanged); the slab's 64 values are
    // loaded at once, then summed in order.  With qtab the running sums are
    // kept; else only the slab total, and a Q draw recomputes its slab's
    // running sums in the same order
    for (int j = tid; j < NS; j += blockDim.x) {
        const float4 *p4 = reinterpret_cast<const float4 *>(phi + j * 64);
        float4 v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = p4[i];
        float run = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            run = run + alpha * v[i].x;
            v[i].x = run;
            run = run + alpha * v[i].y;
            v[i].y = run;
            run = run + alpha * v[i].z;
            v[i].z = run;
            run = run + alpha * v[i].w;
            v[i].w = run;
        }
        if (qtab) {
            float4 *q4 = reinterpret_cast<float4 *>(qs + j * 64);
#pragma unroll
            for (int i = 0; i < 16; ++i) q4[i] = v[i];
        }
        qt[j] = run;
    }
    __syncthreads();
    if (tid == 0) {
        float acc = 0.0f;
        for (int j0 = 0; j0 < NS; j0 += 8) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = j0 + i < NS ? qt[j0 + i] : 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (j0 + i < NS) {
                    acc = acc + v[i];
                    qacc[j0 + i] = acc;
                }
            }
        }
    }
    __syncthreads();
    const float Q = qacc[NS - 1];
    const float f1 = f123[0], f2 = f123[1];
    const uint32_t k1 = tp.x;
    // head words also write the doc-slot copy of z (their items share a doc
    // tile, so the copy's lines fill in L2); tail tokens, in random docs,
    // are read back by token id in the DT rebuild instead
    const bool head = w < A.VD;
    const uint32_t tb = A.tb;

    for (uint32_t base = sg.y + (uint32_t)wv * 64u; base < sg.z; base += (uint32_t)nwv * 64u) {
        const uint32_t r = base + (uint32_t)lane;
        if (r >= sg.z) break;
        const uint4 rn = A.runs[r];
        const uint32_t t0 = rn.x, p0 = rn.y, nd = rn.w >> 16;
        uint32_t len = rn.w & 0xFFFFu;
        // (the empty asm hides len's bound from the optimizer, whose
        // rewrites of the token loops from it otherwise cost a spill)
        asm volatile("" : "+v"(len));
        // n_dk1 and the row's length in one load (a tail word's docs are
        // random ones: a per-doc record would cost a cache line of its own);
        // the row's offset and the doc's length come with the run
        const uint32_t a = A.aux_doc[p0];
        const uint2 di = make_uint2(rn.z, a >> 16);  // row offset, nnz
        const uint32_t n1 = a & 0xFFFFu;
        const float P1 = ((float)n1 + alpha) * f1;
        // tokens of the run that need the full draw; shortcut tokens take k1.
        // The uniforms of the first kUKeep tokens are kept for the draw.
        uint32_t need = 0;
        constexpr int UK = COMPACT ? kUKeep : kUKeep32;
        float uk[UK];
        // S <= (n_d - n_dk1) * phi_k2: every other topic's phi is at most the
        // second largest.  Any bound >= S leaves every draw as the full path
        // makes it (the margin covers the rounding), so the shortcut only
        // skips work; this one needs no n_dk2.
        const float maxS = (float)(nd - n1) * f2;
        const float B = (P1 + maxS) + Q;
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)UK; ++j) {
            uk[j] = 0.0f;
            if (j < len) {
                const float u = ez_u24(ez_philox_x0(A.seed, t0 + j + sg.w, sweep, EZ_TAG_SAMPLE));
                uk[j] = u;
                if (!SHORTCUT || !(u * B < P1 * kMargin)) {
                    need |= 1u << j;
                } else {
                    A.z[t0 + j] = (uint16_t)k1;
                    if (head) A.z_doc[p0 + j] = (uint16_t)k1;
                }
            }
        }
        for (uint32_t j = UK; j < len; ++j) {
            bool full = true;
            if (SHORTCUT) {
                const float u = ez_u24(ez_philox_x0(A.seed, t0 + j + sg.w, sweep, EZ_TAG_SAMPLE));
                full = !(u * B < P1 * kMargin);
            }
            if (full) {
                need |= 1u << j;
            } else {
                A.z[t0 + j] = (uint16_t)k1;
                if (head) A.z_doc[p0 + j] = (uint16_t)k1;
            }
        }
        if (!need || sc_only) continue;
        // the row: 32-bit entries, or 16-bit entries at a 16-bit offset
        const uint4 *rp = COMPACT ? reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(A.dt) + di.x)
                                  : reinterpret_cast<const uint4 *>(A.dt + di.x);
        const uint32_t nnz = di.y & 0xFFFFu;
        constexpr uint32_t kEPB = (COMPACT ? 8u : 4u) * kBlkU4;  // entries per block
        const uint32_t nb = (nnz + kEPB - 1u) / kEPB;            // blocks of the row
        n_full += (uint32_t)__popc(need);
        n_scanned += nnz;
        // S: one FMA per entry, left to right; cp[b] = running sum at the
        // end of block b (the first kCP blocks), so the S-bucket search
        // re-reads one block instead of the row up to the target
        // Block b + 1 is loaded before block b is summed (only by lanes
        // whose row has it, so a wavefront stops loading at its longest row).
        float cp[kCP], ch[kCP];  // running sum at the end / the middle of block b
        float S = 0.0f;
        RowBlock cur = load_block(rp);
#pragma unroll
        for (int b = 0; b < kCP; ++b) {
            RowBlock nxt;
            if ((uint32_t)b + 1u < nb) nxt = load_block(rp + kBlkU4 * (b + 1));
            ch[b] = S;
            if ((uint32_t)b < nb) S = sum_block<COMPACT>(cur, S, tb, ch[b]);
            cp[b] = S;
            cur = nxt;
        }
        for (uint32_t b = kCP; b < nb; ++b) {
            RowBlock nxt;
            if (b + 1u < nb) nxt = load_block(rp + kBlkU4 * (b + 1u));
            float unused;
            S = sum_block<COMPACT>(cur, S, tb, unused);
            cur = nxt;
        }
        const float Aa = P1 + S;
        const float total = Aa + Q;
        while (need) {
            const uint32_t j = (uint32_t)__ffs(need) - 1u;
            need &= need - 1u;
            const uint32_t t = t0 + j;
            float u = uk[0];
#pragma unroll
            for (int i = 1; i < UK; ++i)
                if (j == (uint32_t)i) u = uk[i];
            if (j >= (uint32_t)UK) u = ez_u24(ez_philox_x0(A.seed, t + sg.w, sweep, EZ_TAG_SAMPLE));
            const float target = u * total;
            uint32_t newz = k1;
            if (target < P1) {
                newz = k1;
            } else if (target < Aa) {
                const float t2 = target - P1;
                int bs = -1;
#pragma unroll
                for (int b = kCP - 1; b >= 0; --b)
                    if (cp[b] > t2) bs = b;
                float run = 0.0f;
                bool found = false;
                if (bs >= 0) {
                    // the checkpointed block that holds the target: from the
                    // running sum before it, or with kHalfCp from the one in
                    // its middle, at most half the block's loads
                    float mid = ch[0];
#pragma unroll
                    for (int b = 0; b < kCP - 1; ++b)
                        if (b + 1 == bs) {
                            run = cp[b];
                            mid = ch[b + 1];
                        }
                    const uint4 *bp = rp + kBlkU4 * (uint32_t)bs;
                    if (kHalfCp && COMPACT) {
                        if (mid > t2) {
                            search_u4<COMPACT, kBlkU4 / 2>(bp, run, t2, tb, newz, found);
                        } else {
                            run = mid;
                            search_u4<COMPACT, kBlkU4 / 2>(bp + kBlkU4 / 2, run, t2, tb, newz, found);
                        }
                    } else {
                        search_u4<COMPACT, kBlkU4>(bp, run, t2, tb, newz, found);
                    }
                } else if (nb > (uint32_t)kCP) {
                    // past the checkpoints: the remaining blocks in order
                    run = cp[kCP - 1];
                    for (uint32_t gb = kCP; gb < nb && !found; ++gb)
                        search_u4<COMPACT, kBlkU4>(rp + kBlkU4 * gb, run, t2, tb, newz, found);
                }
                if (!found) {  // rounding left every running sum at or below t2
                    if (COMPACT) {
                        const uint16_t *rr = reinterpret_cast<const uint16_t *>(A.dt) + di.x;
                        for (uint32_t i = 0; i < nnz; ++i)
                            if (c16_count(rr[i], tb) > 0.0f && phi_c16(rr[i], tb) > 0.0f) newz = c16_topic(rr[i], tb);
                    } else {
                        const uint32_t *rr = A.dt + di.x;
                        for (uint32_t i = 0; i < nnz; ++i)
                            if (dt_count(rr[i]) > 0.0f && phi_at0(rr[i]) > 0.0f) newz = dt_topic(rr[i]);
                    }
                }
            } else {
                const float t2 = target - Aa;
                int lo = 0, hi = NS;  // first slab whose running sum exceeds t2
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (qacc[mid] > t2) hi = mid; else lo = mid + 1;
                }
                int js = lo < NS ? lo : -1;
                if (js < 0)
                    for (int jj = NS - 1; jj >= 0; --jj)
                        if (qt[jj] > 0.0f) {
                            js = jj;
                            break;
                        }
                if (js >= 0) {
                    const float t3 = t2 - (js > 0 ? qacc[js - 1] : 0.0f);
                    const int k0 = js * 64, ke = min(k0 + 64, K);
                    lo = ke;
                    if (qtab) {  // first running sum above t3, by bisection
                        lo = k0;
                        hi = ke;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (qs[mid] > t3) hi = mid; else lo = mid + 1;
                        }
                    }
                    // else the slab's running sums, recomputed in the order
                    // its total was summed: the first above t3
                    const float4 *p4 = reinterpret_cast<const float4 *>(phi + k0);
                    float run = 0.0f;
                    for (int i4 = 0; i4 < 16 && !qtab && lo == ke; i4 += 4) {
                        float4 v[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) v[u] = p4[i4 + u];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float x[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int k = k0 + 4 * (i4 + u) + c;
                                run = run + alpha * x[c];
                                if (lo == ke && k < ke && run > t3) lo = k;
                            }
                        }
                    }
                    if (lo < ke) {
                        newz = (uint32_t)lo;
                    } else {
                        for (int k = ke - 1; k >= k0; --k)
                            if (alpha * phi[k] > 0.0f) {
                                newz = (uint32_t)k;
                                break;
                            }
                    }
                }
            }
            A.z[t] = (uint16_t)newz;
            if (head) A.z_doc[p0 + j] = (uint16_t)newz;
        }
    }
}

// One workgroup per item (grid = items), or persistent workgroups pulling
// items from n_queues counters: queue q holds the items with index = q mod
// n_queues (the items of one XCD when items are dealt round-robin to XCDs),
// a workgroup drains its own queue (blockIdx mod n_queues) and then the
// others.  Every workgroup leaves once every queue is past its last item.
template <bool SHORTCUT, bool COMPACT, bool QTAB>
__global__ void __launch_bounds__(1024) k_sample(const SampleArgs A) {
    // the only LDS of the kernel, so phi starts at address 0 and a DT entry's
    // low half is directly the LDS address of its phi value
    extern __shared__ float smf[];
    const int NS = (A.K + 63) >> 6;
    float *f123 = smf + (QTAB ? 2 : 1) * NS * 64 + 2 * NS;  // after phi, qs, qt, qacc
    uint32_t &s_item = *reinterpret_cast<uint32_t *>(f123 + 3);
    const uint32_t sweep = A.counters[0];
    uint32_t n_full = 0, n_scanned = 0;
    if (A.n_queues == 0) {
        sample_item<SHORTCUT, COMPACT, QTAB>(A, blockIdx.x, smf, f123, sweep, n_full, n_scanned);
    } else {
        const uint32_t nq = A.n_queues, x0 = blockIdx.x % nq;
        for (uint32_t h = 0; h < nq; ++h) {
            const uint32_t q = (x0 + h) % nq;
            while (true) {
                __syncthreads();  // the previous item is done with s_item and LDS
                if (threadIdx.x == 0) s_item = atomicAdd(&A.queue[q * kQueueStride], 1u);
                __syncthreads();
                const uint64_t item = (uint64_t)q + (uint64_t)s_item * nq;
                if (item >= A.n_items) break;
                sample_item<SHORTCUT, COMPACT, QTAB>(A, (uint32_t)item, smf, f123, sweep, n_full, n_scanned);
            }
        }
    }
    // per-wavefront totals of the counters
    const int lane = threadIdx.x & 63;
    for (int off = 32; off >= 1; off >>= 1) {
        n_full += __shfl_xor(n_full, off);
        n_scanned += __shfl_xor(n_scanned, off);
    }
    if (lane == 0 && n_full) {
        atomicAdd(&A.stats[0], (unsigned long long)n_full);
        atomicAdd(&A.stats[1], (unsigned long long)n_scanned);
    }
}

// ------------------------------------------------ FP64 log-likelihood
// sum_t log( sum_k (n_dk + alpha)/(n_d + K alpha) * (n_wk + beta)/(n_k + V beta) )
// per token, the term of profilePerp/utility.cu:2398,2810 reduced in FP64
// (profilePerp/utility.cu:2858-2928).  One workgroup per sampler item (phi_w
// in FP64 in LDS), one lane per run; the per-item partials are summed in
// fixed order by k_ll_final.
__global__ void __launch_bounds__(1024) k_ll(const uint4 *__restrict__ segs, const uint4 *__restrict__ runs, int VD,
                                             int K, double alpha, double beta,
                                             double vbeta, const int32_t *__restrict__ wt_head,
                                             const uint32_t *__restrict__ tail_off,
                                             const uint32_t *__restrict__ wt_tail,
                                             const uint32_t *__restrict__ tail_nnz, const int32_t *__restrict__ nk,
                                             const uint32_t *__restrict__ aux_doc, const uint32_t *__restrict__ dt,
                                             int compact, uint32_t tb, double *__restrict__ part) {
    extern __shared__ double smd[];
    __shared__ double red[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
    double *phi = smd;
    const uint4 sg = segs[blockIdx.x];
    if (sg.x == 0xFFFFFFFFu) {  // padding item
        if (tid == 0) part[blockIdx.x] = 0.0;
        return;
    }
    const int w = (int)sg.x;
    if (w < VD) {
        const int32_t *row = wt_head + (size_t)w * K;
        for (int k = tid; k < K; k += blockDim.x) phi[k] = ((double)row[k] + beta) / ((double)nk[k] + vbeta);
    } else {
        for (int k = tid; k < K; k += blockDim.x) phi[k] = beta / ((double)nk[k] + vbeta);
        __syncthreads();
        const int tw = w - VD;
        const uint32_t b = tail_off[tw], n = tail_nnz[tw];
        for (uint32_t i = tid; i < n; i += blockDim.x) {
            const uint32_t e = wt_tail[b + i];
            const uint32_t k = e & 0xFFFFu;
            phi[k] = ((double)(e >> 16) + beta) / ((double)nk[k] + vbeta);
        }
    }
    __syncthreads();
    double q = 0.0;
    for (int k = tid; k < K; k += blockDim.x) q += phi[k];
    for (int off = 32; off >= 1; off >>= 1) q += __shfl_xor(q, off);
    if (lane == 0) red[wv] = q;
    __syncthreads();
    q = 0.0;
    for (int i = 0; i < nwv; ++i) q += red[i];
    q *= alpha;
    __syncthreads();
    // one lane per run (a word's tokens in one doc share the term): its DT
    // row in 16-byte loads (rows are 128-byte aligned and zero padded; a zero
    // entry adds nothing)
    double acc = 0.0;
    const double kalpha = (double)K * alpha;
    const uint32_t tmask = (1u << tb) - 1u;
    for (uint32_t r = sg.y + (uint32_t)tid; r < sg.z; r += (uint32_t)blockDim.x) {
        const uint4 rn = runs[r];
        const uint2 di = make_uint2(rn.z, aux_doc[rn.y] >> 16);  // row offset, nnz
        const uint32_t nnz = di.y & 0xFFFFu;
        double s = 0.0;
        if (compact) {
            const uint4 *rp = reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(dt) + di.x);
            for (uint32_t i = 0; i < (nnz + 7u) / 8u; ++i) {
                const uint4 v = rp[i];
                const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t a = e[u] & 0xFFFFu, b = e[u] >> 16;
                    s += (double)(a >> tb) * phi[a & tmask];
                    s += (double)(b >> tb) * phi[b & tmask];
                }
            }
        } else {
            const uint4 *rp = reinterpret_cast<const uint4 *>(dt + di.x);
            for (uint32_t i = 0; i < (nnz + 3u) / 4u; ++i) {
                const uint4 v = rp[i];
                const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) s += (double)(e[u] >> 16) * phi[dt_topic(e[u])];
            }
        }
        acc += (double)(rn.w & 0xFFFFu) * log((s + q) / ((double)(rn.w >> 16) + kalpha));
    }
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (tid == 0) {
        double tot = 0.0;
        for (int i = 0; i < nwv; ++i) tot += red[i];
        part[blockIdx.x] = tot;
    }
}

__global__ void k_ll_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// --------------------------------------------------------------- fold-in
// Frozen-model tables for fold-in (no reference counterpart): per word the
// phi row and the Q bucket's running sums (inside 64-topic slabs, and over the
// slab totals), from the counts of the model at freezing time.
__global__ void __launch_bounds__(256) k_foldin_tables(int V, int VD, int K, float alpha, float beta,
                                                       const int32_t *__restrict__ wt_head,
                                                       const uint32_t *__restrict__ tail_off,
                                                       const uint32_t *__restrict__ wt_tail,
                                                       const uint32_t *__restrict__ tail_nnz,
                                                       const float *__restrict__ inv, float *__restrict__ phi_out,
                                                       float *__restrict__ qs_out, float *__restrict__ qacc_out) {
    extern __shared__ float ph[];
    const int NS = (K + 63) >> 6;
    const int tid = threadIdx.x;
    for (int w = blockIdx.x; w < V; w += gridDim.x) {
        if (w < VD) {
            const int32_t *row = wt_head + (size_t)w * K;
            for (int k = tid; k < K; k += blockDim.x) ph[k] = ((float)row[k] + beta) * inv[k];
        } else {
            for (int k = tid; k < K; k += blockDim.x) ph[k] = beta * inv[k];
            __syncthreads();
            const int tw = w - VD;
            const uint32_t b = tail_off[tw], n = tail_nnz[tw];
            for (uint32_t i = tid; i < n; i += blockDim.x) {
                const uint32_t e = wt_tail[b + i];
                const uint32_t k = e & 0xFFFFu;
                ph[k] = ((float)(e >> 16) + beta) * inv[k];
            }
        }
        __syncthreads();
        float *po = phi_out + (size_t)w * K;
        for (int k = tid; k < K; k += blockDim.x) po[k] = ph[k];
        float *qo = qs_out + (size_t)w * K;
        __shared__ float qt[1024];  // slab totals (NS <= 256 for K <= 16383)
        for (int j = tid; j < NS; j += blockDim.x) {
            const int k0 = j * 64, ke = min(k0 + 64, K);
            float run = 0.0f;
            for (int k = k0; k < ke; ++k) {
                run = run + alpha * ph[k];
                qo[k] = run;
            }
            qt[j] = run;
        }
        __syncthreads();
        if (tid == 0) {
            float acc = 0.0f;
            for (int j = 0; j < NS; ++j) {
                acc = acc + qt[j];
                qacc_out[(size_t)w * NS + j] = acc;
            }
        }
        __syncthreads();
    }
}

static constexpr int kFB = 8, kFCP = 16;     // fold-in S blocks and checkpoints
static constexpr uint32_t kFoldSeg = 256;     // tokens per big fold-in work item at most
static constexpr uint32_t kFoldHot = 2048;    // word ranks counted in LDS before the global atomics
static constexpr uint32_t kFoldStageQ = 64;   // tokens of a word in a step that make big items

// A fold-in step runs in word-major order, as the training sweep does: the
// resident documents' rows are built (k_fi_rows), their tokens are bucketed
// by word (k_fi_count, the three-pass scan, k_fi_scatter), and persistent
// workgroups take one (word, segment) item at a time, stage the frozen phi
// row and Q running sums in LDS and draw its tokens (k_fi_sample); documents
// finishing their last sweep get their topic counts and topics packed for
// one copy to the host (k_fi_final).  The order tokens are drawn in never
// changes a draw, which is keyed by (seed, request id, position, sweep).
// Slots are (request id, pool offset, tokens, sweep); pool offsets are
// multiples of 8, and a document's row lives in the row pool at its offset
// (at most min(n, K) entries, zero padded to a multiple of 8 <= round8(n)).

// Copies admitted documents' words from the upload buffer into the pool:
// records (source offset, pool offset, tokens, 0).
__global__ void __launch_bounds__(256) k_fi_admit(const uint4 *__restrict__ adm, int nadm,
                                                  const uint32_t *__restrict__ src, uint32_t *__restrict__ wpool) {
    for (int a = blockIdx.x; a < nadm; a += gridDim.x) {
        const uint4 r = adm[a];
        for (uint32_t i = threadIdx.x; i < r.z; i += blockDim.x) wpool[r.y + i] = src[r.x + i];
    }
}

// f(value, i) for every i < n of a document's pool range, eight loads per
// thread issued before their uses (long documents would otherwise chain one
// load latency per 256 tokens).
template <class T, class F>
__device__ __forceinline__ void for_doc_tokens(const T *__restrict__ p, uint32_t n, F f) {
    constexpr int U = 8;
    for (uint32_t i0 = threadIdx.x; i0 < n; i0 += U * blockDim.x) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + (uint32_t)u * blockDim.x;
            v[u] = i < n ? p[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + (uint32_t)u * blockDim.x;
            if (i < n) f(v[u], i);
        }
    }
}

// Rows of the resident documents (one workgroup per slot): topics drawn from
// the init stream on sweep 0, a 16-bit LDS histogram, a row of topic | count
// << 16 in ascending topic order, zero padded to a multiple of kFB.
__global__ void __launch_bounds__(256) k_fi_rows(const uint4 *__restrict__ slots, int nslots, int K, uint64_t seed,
                                                 uint16_t *__restrict__ zpool, uint32_t *__restrict__ rows,
                                                 uint32_t *__restrict__ row_nnz) {
    extern __shared__ uint32_t shr[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, bd = blockDim.x;
    const int KH = (K + 1) >> 1;
    uint32_t *h = shr;
    for (int k = tid; k < KH; k += bd) h[k] = 0;
    __syncthreads();
    for (int j = blockIdx.x; j < nslots; j += gridDim.x) {
        const uint4 sl = slots[j];
        const uint32_t rid = sl.x, off = sl.y, n = sl.z, sweep = sl.w;
        uint16_t *z = zpool + off;
        if (sweep == 0) {
            for (uint32_t i = tid; i < n; i += bd)
                z[i] = (uint16_t)(ez_philox_x0(seed, ((uint64_t)rid << 32) | i, 0u, EZ_TAG_FOLDIN_INIT) % (uint32_t)K);
            __syncthreads();
        }
        for_doc_tokens(z, n, [&](uint16_t t, uint32_t) { h16_add(h, t); });
        __syncthreads();
        if (wv == 0) {  // one wavefront compacts the row (and clears the counts)
            uint32_t *row = rows + off;
            uint32_t nz = 0;
            for (int c0 = 0; c0 < K; c0 += 64) {
                const int k = c0 + lane;
                const uint32_t c = k < K ? h16_get(h, (uint32_t)k) : 0u;
                const uint64_t m = __ballot(c != 0);
                if (c) row[nz + __popcll(m & lanemask_lt(lane))] = (uint32_t)k | (c << 16);
                nz += (uint32_t)__popcll(m);
                wave_lds_sync();
                if (k < K && !(k & 1)) h[k >> 1] = 0;
            }
            // zero entries up to the next block: count 0 adds exactly nothing
            if ((uint32_t)lane < ((kFB - (nz % kFB)) % kFB)) row[nz + lane] = 0u;
            if (lane == 0) row_nnz[j] = nz;
        }
        __syncthreads();
    }
}

// Tokens per word of the resident documents; the most frequent ranks are
// summed in LDS first so they take one global atomic per workgroup.
__global__ void __launch_bounds__(256) k_fi_count(const uint4 *__restrict__ slots, int nslots,
                                                  const uint32_t *__restrict__ wpool, uint32_t *__restrict__ wcount) {
    __shared__ uint32_t hc[kFoldHot];
    const int tid = threadIdx.x, bd = blockDim.x;
    for (uint32_t k = tid; k < kFoldHot; k += bd) hc[k] = 0;
    __syncthreads();
    for (int j = blockIdx.x; j < nslots; j += gridDim.x) {
        const uint4 sl = slots[j];
        for_doc_tokens(wpool + sl.y, sl.z, [&](uint32_t w, uint32_t) {
            if (w < kFoldHot) atomicAdd(&hc[w], 1u);
            else atomicAdd(&wcount[w], 1u);
        });
    }
    __syncthreads();
    for (uint32_t k = tid; k < kFoldHot; k += bd)
        if (hc[k]) atomicAdd(&wcount[k], hc[k]);
}

// Exclusive scan of M values per thread over a workgroup of up to 1024
// threads; the totals go to tot.  sh holds 16 * M words.
template <int M>
__device__ __forceinline__ void block_scan(uint32_t (&v)[M], uint32_t *sh, uint32_t (&tot)[M]) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) inc[m] = v[m];
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const uint32_t x = __shfl_up(inc[m], d);
            if (lane >= d) inc[m] += x;
        }
    }
    if (lane == 63)
#pragma unroll
        for (int m = 0; m < M; ++m) sh[16 * m + wv] = inc[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        uint32_t o = 0, t = 0;
        for (int w = 0; w < nw; ++w) {
            const uint32_t x = sh[16 * m + w];
            if (w < wv) o += x;
            t += x;
        }
        tot[m] = t;
        v[m] = o + inc[m] - v[m];
    }
    __syncthreads();
}

// A word with at least kFoldStageQ tokens in the step is big: its tokens
// form items of at most kFoldSeg tokens, a workgroup each, with its phi
// row staged in LDS.  The other words' tokens follow all big words' tokens
// in the bucketed list and are drawn one lane per token straight from the
// tables.
__device__ __forceinline__ uint32_t fold_big(uint32_t c) { return c >= kFoldStageQ ? (c + kFoldSeg - 1u) / kFoldSeg : 0u; }

// Scan pass 1: per 1024 words, their big-word tokens, big items and
// small-word tokens.
__global__ void __launch_bounds__(1024) k_fi_bsum(int V, const uint32_t *__restrict__ wcount,
                                                  uint32_t *__restrict__ bsum) {
    __shared__ uint32_t sh[48];
    const uint32_t w = blockIdx.x * 1024u + threadIdx.x;
    const uint32_t c = w < (uint32_t)V ? wcount[w] : 0u;
    const bool bg = c >= kFoldStageQ;
    uint32_t v[3] = {bg ? c : 0u, fold_big(c), bg ? 0u : c}, t[3];
    block_scan<3>(v, sh, t);
    if (threadIdx.x == 0)
        for (int m = 0; m < 3; ++m) bsum[3 * blockIdx.x + m] = t[m];
}

// Scan pass 2 (one workgroup): exclusive offsets of the per-block sums; ctr
// = [big items, big-word tokens, small-word tokens, 0].
__global__ void __launch_bounds__(1024) k_fi_btop(int nb, uint32_t *__restrict__ bsum, uint32_t *__restrict__ ctr) {
    __shared__ uint32_t sh[48];
    uint32_t carry[3] = {0u, 0u, 0u};
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int i = c0 + (int)threadIdx.x;
        uint32_t v[3], t[3];
        for (int m = 0; m < 3; ++m) v[m] = i < nb ? bsum[3 * i + m] : 0u;
        block_scan<3>(v, sh, t);
        for (int m = 0; m < 3; ++m) {
            if (i < nb) bsum[3 * i + m] = carry[m] + v[m];
            carry[m] += t[m];
        }
    }
    if (threadIdx.x == 0) {
        ctr[0] = carry[1];
        ctr[1] = carry[0];
        ctr[2] = carry[2];
        ctr[3] = 0u;
    }
}

// Scan pass 3: bucket starts (the scatter's cursors: big words first, then
// small words, each in word order) and the big items (word, first, end, 0).
__global__ void __launch_bounds__(1024) k_fi_bfin(int V, const uint32_t *__restrict__ wcount,
                                                  const uint32_t *__restrict__ bsum, const uint32_t *__restrict__ ctr,
                                                  uint32_t *__restrict__ wcursor, uint4 *__restrict__ big) {
    __shared__ uint32_t sh[48];
    const uint32_t w = blockIdx.x * 1024u + threadIdx.x;
    const uint32_t c = w < (uint32_t)V ? wcount[w] : 0u;
    const bool bg = c >= kFoldStageQ;
    uint32_t v[3] = {bg ? c : 0u, fold_big(c), bg ? 0u : c}, t[3];
    block_scan<3>(v, sh, t);
    if (w >= (uint32_t)V) return;
    if (bg) {
        const uint32_t base = bsum[3 * blockIdx.x] + v[0];
        wcursor[w] = base;
        uint32_t ib = bsum[3 * blockIdx.x + 1] + v[1];
        for (uint32_t x = 0; x < c; x += kFoldSeg) big[ib++] = make_uint4(w, base + x, base + min(c, x + kFoldSeg), 0u);
    } else {
        wcursor[w] = ctr[1] + bsum[3 * blockIdx.x + 2] + v[2];
    }
}

// Token references (slot << 16 | position) bucketed by word.  Ranks below
// kFoldHot are counted in LDS, reserve one range per workgroup and fill it
// through LDS cursors; the rest take a global cursor per token.
__global__ void __launch_bounds__(256) k_fi_scatter(const uint4 *__restrict__ slots, int nslots,
                                                    const uint32_t *__restrict__ wpool, uint32_t *__restrict__ wcursor,
                                                    uint32_t *__restrict__ tokw) {
    __shared__ uint32_t hc[kFoldHot];
    const int tid = threadIdx.x, bd = blockDim.x;
    for (uint32_t k = tid; k < kFoldHot; k += bd) hc[k] = 0;
    __syncthreads();
    for (int j = blockIdx.x; j < nslots; j += gridDim.x) {
        const uint4 sl = slots[j];
        for_doc_tokens(wpool + sl.y, sl.z, [&](uint32_t w, uint32_t i) {
            if (w < kFoldHot) atomicAdd(&hc[w], 1u);
            else tokw[atomicAdd(&wcursor[w], 1u)] = ((uint32_t)j << 16) | i;
        });
    }
    __syncthreads();
    {  // one range per hot word of the workgroup; the atomics issued together
        constexpr uint32_t U = kFoldHot / 256;
        uint32_t c[U], b[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) c[u] = tid + u * bd < kFoldHot ? hc[tid + u * bd] : 0u;
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) b[u] = c[u] ? atomicAdd(&wcursor[tid + u * bd], c[u]) : 0u;
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            if (c[u]) hc[tid + u * bd] = b[u];
        for (uint32_t k = tid + U * bd; k < kFoldHot; k += bd) {  // block sizes below 256
            const uint32_t cc = hc[k];
            if (cc) hc[k] = atomicAdd(&wcursor[k], cc);
        }
    }
    __syncthreads();
    for (int j = blockIdx.x; j < nslots; j += gridDim.x) {
        const uint4 sl = slots[j];
        for_doc_tokens(wpool + sl.y, sl.z, [&](uint32_t w, uint32_t i) {
            if (w < kFoldHot) tokw[atomicAdd(&hc[w], 1u)] = ((uint32_t)j << 16) | i;
        });
    }
}

__device__ __forceinline__ float fi_sum8(const uint4 &a, const uint4 &b, const float *ph, float s) {
    const uint32_t e[kFB] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int u = 0; u < kFB; ++u) s = __builtin_fmaf((float)(e[u] >> 16), ph[e[u] & 0xFFFFu], s);
    return s;
}

// The entry of a block (zero padded) at which the running sum from `run`
// first exceeds the target; -1 if none does.  Zero entries add exactly
// nothing, so they never cross the target first.
__device__ __forceinline__ int fi_scan8(const uint4 &a, const uint4 &b, const float *ph, float run, float target) {
    const uint32_t e[kFB] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    int zz = -1;
#pragma unroll
    for (int q = 0; q < kFB; ++q) {
        run = __builtin_fmaf((float)(e[q] >> 16), ph[e[q] & 0xFFFFu], run);
        if (zz < 0 && run > target) zz = (int)(e[q] & 0xFFFFu);
    }
    return zz;
}

// First index in [k0, ke) whose running sum exceeds t, ke if none: the
// running sums are nondecreasing, so this is the index a bisection finds.
// Rounds of 8 independent loads: the last entries of up to 8 groups of 8,
// then the entries of the first group whose last one exceeds t.
__device__ __forceinline__ int first_above(const float *q, int k0, int ke, float t) {
    for (int c0 = k0; c0 < ke; c0 += 64) {
        const int c1 = min(c0 + 64, ke);
        int grp = -1;
#pragma unroll
        for (int g = 7; g >= 0; --g) {
            const int s0 = c0 + 8 * g;
            if (s0 < c1 && q[min(s0 + 7, c1 - 1)] > t) grp = g;
        }
        if (grp >= 0) {
            const int g0 = c0 + 8 * grp, g1 = min(g0 + 8, c1);
            int lo = g1;
#pragma unroll
            for (int u = 7; u >= 0; --u)
                if (g0 + u < g1 && q[g0 + u] > t) lo = g0 + u;
            return lo;
        }
    }
    return ke;
}

// One fold-in draw: token i of slot sl (its row of nz entries at pool
// offset sl.y) for a word whose phi row (ph) and Q running sums (qa over
// the slab totals, qw within slabs) are given, in LDS or in the tables.  S
// over the row in blocks of kFB entries, four blocks' loads issued together
// (the first kFCP block sums kept), then the S entry or the Q topic whose
// running sum first exceeds the target.
__device__ __forceinline__ void fi_draw(const uint4 sl, uint32_t i, uint32_t nz, const uint32_t *__restrict__ rows,
                                        const float *ph, const float *qa, const float *qw, int K, int NS, float alpha,
                                        uint64_t seed, uint16_t *__restrict__ zpool) {
    const uint32_t *row = rows + sl.y;
    const float u = ez_u24(ez_philox_x0(seed, ((uint64_t)sl.x << 32) | i, sl.w, EZ_TAG_FOLDIN));
    const uint32_t nb = (nz + kFB - 1) / kFB;
    float cp[kFCP];
    float S = 0.0f;
#pragma unroll
    for (int g = 0; g < kFCP; g += 4) {
        uint4 e[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            e[2 * q] = e[2 * q + 1] = make_uint4(0u, 0u, 0u, 0u);
            if ((uint32_t)(g + q) < nb) {
                e[2 * q] = *(const uint4 *)(row + kFB * (g + q));
                e[2 * q + 1] = *(const uint4 *)(row + kFB * (g + q) + 4);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if ((uint32_t)(g + q) < nb) S = fi_sum8(e[2 * q], e[2 * q + 1], ph, S);
            cp[g + q] = S;
        }
    }
    for (uint32_t b = kFCP; b < nb; b += 4) {
        uint4 e[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            e[2 * q] = e[2 * q + 1] = make_uint4(0u, 0u, 0u, 0u);
            if (b + q < nb) {
                e[2 * q] = *(const uint4 *)(row + kFB * (b + q));
                e[2 * q + 1] = *(const uint4 *)(row + kFB * (b + q) + 4);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (b + q < nb) S = fi_sum8(e[2 * q], e[2 * q + 1], ph, S);
    }
    const float target = u * (S + qa[NS - 1]);
    int zz = -1;
    if (target < S) {
        // the first block whose running sum exceeds the target (one exists:
        // the last running sum is S), then its entries in order
        int bs = -1;
#pragma unroll
        for (int b = kFCP - 1; b >= 0; --b)
            if (cp[b] > target) bs = b;
        float run = 0.0f;
        if (bs >= 0) {
#pragma unroll
            for (int b = 0; b < kFCP - 1; ++b)
                if (b + 1 == bs) run = cp[b];
            const uint32_t g0 = (uint32_t)bs * kFB;
            zz = fi_scan8(*(const uint4 *)(row + g0), *(const uint4 *)(row + g0 + 4), ph, run, target);
        } else {
            // past the checkpoints: blocks four at a time until the block
            // whose running sum exceeds the target, then its entries
            run = cp[kFCP - 1];
            for (uint32_t b = kFCP; b < nb && zz < 0; b += 4) {
                uint4 e[8];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    e[2 * q] = e[2 * q + 1] = make_uint4(0u, 0u, 0u, 0u);
                    if (b + q < nb) {
                        e[2 * q] = *(const uint4 *)(row + kFB * (b + q));
                        e[2 * q + 1] = *(const uint4 *)(row + kFB * (b + q) + 4);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (zz < 0 && b + q < nb) {
                        const float r2 = fi_sum8(e[2 * q], e[2 * q + 1], ph, run);
                        if (r2 > target) zz = fi_scan8(e[2 * q], e[2 * q + 1], ph, run, target);
                        else run = r2;
                    }
                }
            }
        }
        if (zz < 0)  // not reached: kept as the canonical fallback
            for (uint32_t g = 0; g < nz; ++g)
                if ((row[g] >> 16) && ph[row[g] & 0xFFFFu] > 0.0f) zz = (int)(row[g] & 0xFFFFu);
    } else {
        const float t2 = target - S;
        const int lo = first_above(qa, 0, NS, t2);
        int js = lo < NS ? lo : -1;
        if (js < 0)
            for (int jj = NS - 1; jj >= 0; --jj)
                if (qa[jj] > (jj > 0 ? qa[jj - 1] : 0.0f)) {
                    js = jj;
                    break;
                }
        if (js >= 0) {
            const float t3 = t2 - (js > 0 ? qa[js - 1] : 0.0f);
            const int k0 = js * 64, ke = min(k0 + 64, K);
            const int kq = first_above(qw, k0, ke, t3);
            if (kq < ke) {
                zz = kq;
            } else {
                for (int k = ke - 1; k >= k0; --k)
                    if (alpha * ph[k] > 0.0f) {
                        zz = k;
                        break;
                    }
            }
        }
    }
    zpool[sl.y + i] = (uint16_t)(zz < 0 ? 0 : zz);
}

// Workgroups of 4 wavefronts.  First the big items, assigned by striding
// (each is one pass of similar cost; a shared queue head would serialize on
// one address): phi_w, the slab running sums and the running sums of the
// slab totals staged in LDS, one lane per token.  Then the small words'
// tokens, one lane each, the word read from the pool and its tables from
// HBM (their phi rows are touched by few tokens each, and staging them
// would leave most lanes idle).
__global__ void __launch_bounds__(256) k_fi_sample(const uint4 *__restrict__ slots, int K, float alpha,
                                                   const float *__restrict__ phi, const float *__restrict__ qs,
                                                   const float *__restrict__ qacc, const uint32_t *__restrict__ wpool,
                                                   const uint32_t *__restrict__ rows,
                                                   const uint32_t *__restrict__ row_nnz,
                                                   const uint32_t *__restrict__ tokw, const uint4 *__restrict__ big,
                                                   const uint32_t *__restrict__ ctr, uint16_t *__restrict__ zpool,
                                                   uint64_t seed) {
    extern __shared__ float shs[];
    const int NS = (K + 63) >> 6, tid = threadIdx.x, bd = blockDim.x;
    const uint32_t n_big = ctr[0], t_big = ctr[1], t_all = ctr[1] + ctr[2];
    float *ph = shs;      // K: phi_w
    float *qa = ph + K;   // NS: running sums of the slab totals
    float *ql = qa + NS;  // K: slab running sums
    for (uint32_t it = blockIdx.x; it < n_big; it += gridDim.x) {
        __syncthreads();  // the previous item's reads are done
        const uint4 item = big[it];
        const size_t w = item.x;
        for (int k = tid; k < K; k += bd) {
            ph[k] = phi[w * K + k];
            ql[k] = qs[w * K + k];
        }
        for (int j = tid; j < NS; j += bd) qa[j] = qacc[w * NS + j];
        __syncthreads();
        for (uint32_t t = item.y + tid; t < item.z; t += bd) {
            const uint32_t ref = tokw[t], j = ref >> 16;
            fi_draw(slots[j], ref & 0xFFFFu, row_nnz[j], rows, ph, qa, ql, K, NS, alpha, seed, zpool);
        }
    }
    for (uint32_t t = t_big + blockIdx.x * (uint32_t)bd + tid; t < t_all; t += gridDim.x * (uint32_t)bd) {
        const uint32_t ref = tokw[t], j = ref >> 16, i = ref & 0xFFFFu;
        const uint4 sl = slots[j];
        const uint32_t nz = row_nnz[j];
        const size_t w = wpool[sl.y + i];
        fi_draw(sl, i, nz, rows, phi + w * K, qacc + w * NS, qs + w * K, K, NS, alpha, seed, zpool);
    }
}

// The plain fold-in path: one launch runs every sweep of one document.  One
// workgroup: the document's topics live in z (device memory, n values), its
// 16-bit topic histogram and its row (topic | count << 16, ascending topic,
// zero padded to a multiple of kFB) in LDS.  Each sweep rebuilds the row from
// z, then every token is drawn against it (Jacobi: the row is complete before
// any token of the sweep moves) with the draw the batched step uses, so the
// topics are the server's and the oracle's.
static size_t foldin_doc_lds(int K) {
    const size_t kh4 = (((size_t)K + 1) / 2 + 3) & ~(size_t)3;  // 16-byte aligned row after the counts
    return (kh4 + (((size_t)K + 7) & ~(size_t)7) + 8 + 4) * sizeof(uint32_t);
}

__global__ void __launch_bounds__(1024) k_fi_doc(const uint32_t *__restrict__ words, uint32_t n, uint32_t rid,
                                                 int sweeps, int K, float alpha, const float *__restrict__ phi,
                                                 const float *__restrict__ qs, const float *__restrict__ qacc,
                                                 uint64_t seed, uint16_t *__restrict__ z) {
    extern __shared__ uint32_t shd[];
    const int NS = (K + 63) >> 6, tid = threadIdx.x, bd = blockDim.x, lane = tid & 63, wv = tid >> 6;
    const int KH = (K + 1) >> 1, KH4 = (KH + 3) & ~3;
    uint32_t *h = shd;                                   // KH words: 16-bit counts
    uint32_t *row = shd + KH4;                           // the row, 16-byte aligned
    uint32_t *s_nz = row + ((K + 7) & ~7) + 8;           // the row's entries
    for (uint32_t i = tid; i < n; i += bd)
        z[i] = (uint16_t)(ez_philox_x0(seed, ((uint64_t)rid << 32) | i, 0u, EZ_TAG_FOLDIN_INIT) % (uint32_t)K);
    for (int k = tid; k < KH; k += bd) h[k] = 0;
    __syncthreads();
    for (int sw = 0; sw < sweeps; ++sw) {
        for (uint32_t i = tid; i < n; i += bd) h16_add(h, z[i]);
        __syncthreads();
        if (wv == 0) {  // one wavefront compacts the row in ascending topic order and clears the counts
            uint32_t nz = 0;
            for (int c0 = 0; c0 < K; c0 += 64) {
                const int k = c0 + lane;
                const uint32_t c = k < K ? h16_get(h, (uint32_t)k) : 0u;
                const uint64_t m = __ballot(c != 0);
                if (c) row[nz + __popcll(m & lanemask_lt(lane))] = (uint32_t)k | (c << 16);
                nz += (uint32_t)__popcll(m);
                wave_lds_sync();
                if (k < K && !(k & 1)) h[k >> 1] = 0;
            }
            // zero entries up to the next block: a count of 0 adds exactly nothing
            if ((uint32_t)lane < ((kFB - (nz % kFB)) % kFB)) row[nz + lane] = 0u;
            if (lane == 0) *s_nz = nz;
        }
        __syncthreads();
        const uint32_t nz = *s_nz;
        const uint4 sl = make_uint4(rid, 0u, n, (uint32_t)sw);
        for (uint32_t i = tid; i < n; i += bd) {
            const size_t w = words[i];
            fi_draw(sl, i, nz, row, phi + w * K, qacc + w * NS, qs + w * K, K, NS, alpha, seed, z);
        }
        __syncthreads();  // every topic of the sweep is written before the next histogram
    }
}

hipError_t launch_foldin_doc(const uint32_t *words, uint32_t n, uint32_t rid, int sweeps, int K, float alpha,
                             const float *phi, const float *qs, const float *qacc, uint64_t seed, uint16_t *z,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fi_doc, dim3(1), dim3(1024), foldin_doc_lds(K), st, words, n, rid, sweeps, K, alpha, phi, qs,
                       qacc, seed, z);
    return hipGetLastError();
}

// Documents that just drew their last sweep (fin[j] = (output index, topic
// offset), output index ~0 for the others): their topics packed into zout at
// the offset, for one copy to the host (which counts them).
__global__ void __launch_bounds__(256) k_fi_final(const uint4 *__restrict__ slots, const uint2 *__restrict__ fin,
                                                  int nslots, const uint16_t *__restrict__ zpool,
                                                  uint16_t *__restrict__ zout) {
    for (int j = blockIdx.x; j < nslots; j += gridDim.x) {
        const uint2 f = fin[j];
        if (f.x == 0xFFFFFFFFu) continue;  // uniform over the workgroup
        const uint4 sl = slots[j];
        for_doc_tokens(zpool + sl.y, sl.z, [&](uint16_t t, uint32_t i) { zout[f.y + i] = t; });
    }
}

hipError_t launch_foldin_tables(const DevState &s, float *phi, float *qs, float *qacc, hipStream_t st) {
    hipLaunchKernelGGL(k_foldin_tables, dim3(grid_for(s.V, 1, 1 << 16)), dim3(256), (size_t)s.K * sizeof(float), st,
                       s.V, s.VD, s.K, s.alpha, s.beta, s.wt_head, s.tail_off, s.wt_tail, s.tail_nnz, s.inv, phi, qs,
                       qacc);
    return hipGetLastError();
}

static size_t foldin_sample_lds(int K) { return (2 * (size_t)K + (size_t)((K + 63) >> 6)) * sizeof(float); }

size_t foldin_lds_bytes(int K) {
    return std::max(foldin_sample_lds(K) + 16, (size_t)((K + 1) >> 1) * sizeof(uint32_t));
}

uint32_t foldin_max_items(int V, uint32_t max_tokens) { return (uint32_t)V + max_tokens / kFoldSeg + 1u; }

hipError_t launch_foldin_admit(const uint4 *adm, int nadm, const uint32_t *src, uint32_t *wpool, hipStream_t st) {
    if (nadm <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fi_admit, dim3(std::min(nadm, 1024)), dim3(256), 0, st, adm, nadm, src, wpool);
    return hipGetLastError();
}

hipError_t launch_foldin_step(const FoldinBatch &fb, const uint4 *slots, const uint2 *fin, int nslots, uint32_t ntok,
                              int V, int K, float alpha, const float *phi, const float *qs, const float *qacc,
                              const uint32_t *wpool, uint16_t *zpool, uint64_t seed, hipStream_t st) {
    if (nslots <= 0) return hipSuccess;
    // count and scatter: one workgroup per CU (fb.grid is 8 per CU), each
    // flushing its LDS table of the most frequent words once for many slots
    const unsigned gs = (unsigned)std::min(nslots, 1 << 16);
    const unsigned gc = (unsigned)std::max(1, std::min(nslots, std::max(1, fb.grid / 8)));
    const unsigned nb = (unsigned)((V + 1023) / 1024);
    const size_t lds_rows = (size_t)((K + 1) >> 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((size_t)V, 256, 4096)), dim3(256), 0, st, fb.wcount, (size_t)V, 0u);
    hipLaunchKernelGGL(k_fi_rows, dim3(gs), dim3(256), lds_rows, st, slots, nslots, K, seed, zpool, fb.rows,
                       fb.row_nnz);
    hipLaunchKernelGGL(k_fi_count, dim3(gc), dim3(256), 0, st, slots, nslots, wpool, fb.wcount);
    hipLaunchKernelGGL(k_fi_bsum, dim3(nb), dim3(1024), 0, st, V, fb.wcount, fb.bsum);
    hipLaunchKernelGGL(k_fi_btop, dim3(1), dim3(1024), 0, st, (int)nb, fb.bsum, fb.ctr);
    hipLaunchKernelGGL(k_fi_bfin, dim3(nb), dim3(1024), 0, st, V, fb.wcount, fb.bsum, fb.ctr, fb.wcursor,
                       fb.items);
    hipLaunchKernelGGL(k_fi_scatter, dim3(gc), dim3(256), 0, st, slots, nslots, wpool, fb.wcursor, fb.tokw);
    // no more workgroups than a lane per token could use
    const unsigned gsmp = (unsigned)std::max<int64_t>(1, std::min<int64_t>(fb.grid, ((int64_t)ntok + 255) / 256));
    hipLaunchKernelGGL(k_fi_sample, dim3(gsmp), dim3(256), foldin_sample_lds(K), st, slots, K, alpha, phi, qs, qacc,
                       wpool, fb.rows, fb.row_nnz, fb.tokw, fb.items, fb.ctr, zpool, seed);
    hipLaunchKernelGGL(k_fi_final, dim3(gs), dim3(256), 0, st, slots, fin, nslots, zpool, fb.zout);
    return hipGetLastError();
}

// ------------------------------------------------------------ self-tests
__global__ void k_selftest_philox(const uint4 *__restrict__ ctr, const uint2 *__restrict__ key,
                                  uint4 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 c = ctr[i];
    const ez_u4 o = ez_philox4x32_10(ez_u4{c.x, c.y, c.z, c.w}, key[i].x, key[i].y);
    out[i] = make_uint4(o.x, o.y, o.z, o.w);
}

// Plain streaming copy (16 B per lane per step): the known-size reference
// for scaling HBM counters.
__global__ void k_copy_u4(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Scattered 64-byte reads (the sampler's row-block pattern: each lane loads
// four 16-byte pieces of one 64-byte block at an address no neighbouring lane
// shares), n_reads of them over a buffer of nblk = 2^m 64-byte blocks: read i
// takes block (i * an odd constant) mod nblk, a bijection, so n_reads <= nblk
// reads touch n_reads distinct blocks, each once (64 * n_reads bytes, none of
// them cached by an earlier read).  The known-size workload that scales the
// traffic counters for that request mix.  Nothing is written unless the data
// holds an impossible value (the loads cannot be removed).
__global__ void k_gather64(const uint4 *__restrict__ src, uint64_t blk_mask, uint64_t n_reads,
                           uint32_t *__restrict__ sink) {
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_reads;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t blk = (i * 0x9E3779B97F4A7C15ull) & blk_mask;
        const uint4 *p = src + blk * 4u;
        const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
        acc ^= a.x ^ a.w ^ b.y ^ c.z ^ d.x ^ d.w;
    }
    if (acc == 0x9E3779B9u) sink[threadIdx.x & 63u] = acc;
}

hipError_t launch_gather64(const void *src, size_t bytes, uint64_t n_reads, uint32_t *sink, hipStream_t st) {
    uint64_t nblk = 1;
    while (nblk * 2 * 64 <= bytes) nblk *= 2;
    if (bytes < 64 || n_reads > nblk) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather64, dim3(grid_for((size_t)n_reads, 256, 8192)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(src), nblk - 1, n_reads, sink);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_u4, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

// ------------------------------------ allocation-free log-likelihood memory
// The device memory of ezlda_loglik_noalloc: kLLScratchItems per-item
// log-likelihood partials, then the result.  A file-scope array, so the call
// needs no allocation; every GPU holds its own copy (8 MiB) until the process
// ends, and every call on a GPU shares it.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): scratch_address() returns the device address, once per device.
static __device__ __align__(256) double scratch[kLLScratchItems + 1];

hipError_t scratch_address(double **p) {
    *p = nullptr;
    return hipGetSymbolAddress(reinterpret_cast<void **>(p), HIP_SYMBOL(scratch));
}

hipError_t launch_selftest_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_selftest_philox, dim3((n + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(ctr), reinterpret_cast<const uint2 *>(key),
                       reinterpret_cast<uint4 *>(out), n);
    return hipGetLastError();
}

// ------------------------------------------------------------ launchers

hipError_t launch_set_u32(uint32_t *p, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, p, v);
    return hipGetLastError();
}

hipError_t launch_zero_counts(const DevState &s, hipStream_t st) {
    const size_t nh = (size_t)s.VD * s.K;
    if (nh) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nh, 256, 4096)), dim3(256), 0, st,
                               reinterpret_cast<uint32_t *>(s.wt_head), nh, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(s.K, 256, 64)), dim3(256), 0, st,
                       reinterpret_cast<uint32_t *>(s.nk), (size_t)s.K, 0u);
    return hipGetLastError();
}

hipError_t launch_build_wt(const DevState &s, hipStream_t st) {
    hipError_t e = launch_zero_counts(s, st);
    if (e != hipSuccess) return e;
    return launch_accumulate_wt(s, st);
}

hipError_t launch_accumulate_wt(const DevState &s, hipStream_t st) {
    const size_t ldsK = (size_t)s.K * sizeof(uint32_t);
    const int ntail = s.V - s.VD;
    if (s.N && s.n_wt_segs && ntail > 0 && (size_t)s.lds_cu / dt_wave_bytes(s.K) >= 16) {
        // one launch of 256-thread blocks (the tail rows' four wavefronts a
        // block, as their own launch takes when 16 wavefronts' tables fit a CU)
        const int nb_hist = (int)grid_for(s.N, 256 * 16, 2048);
        const int nb_tail = (int)grid_for((size_t)ntail, 4, 1 << 16);
        const size_t lds = std::max(ldsK, 4 * dt_wave_bytes(s.K));
        hipLaunchKernelGGL(k_wt_all, dim3(nb_hist + s.n_wt_segs + nb_tail), dim3(256), lds, st, nb_hist, s.n_wt_segs,
                           s.z, s.N, s.K, s.nk, s.wt_segs, s.wt_head, s.word_off, s.VD, s.V, s.tail_off, s.wt_tail,
                           s.tail_nnz);
        return hipGetLastError();
    }
    if (s.N) hipLaunchKernelGGL(k_topic_hist, dim3(grid_for(s.N, 256 * 16, 2048)), dim3(256), ldsK, st, s.z, s.N, s.K, s.nk);
    if (s.n_wt_segs)
        hipLaunchKernelGGL(k_wt_head, dim3(s.n_wt_segs), dim3(256), ldsK, st, s.z, s.wt_segs, s.K, s.wt_head);
    if (ntail > 0) {
        const WaveGrid g = wave_grid(s, (size_t)ntail, dt_wave_bytes(s.K), 1 << 16);
        hipLaunchKernelGGL(k_wt_tail, dim3(g.grid), dim3(64 * g.nwv), g.nwv * dt_wave_bytes(s.K), st,
                           s.z, s.word_off, s.VD, s.V, s.K, s.tail_off, s.wt_tail, s.tail_nnz);
    }
    return hipGetLastError();
}

hipError_t launch_head_rows(const DevState &s, hipStream_t st) {
    const size_t nh = (size_t)s.VD * s.K;
    if (nh) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nh, 256, 4096)), dim3(256), 0, st,
                               reinterpret_cast<uint32_t *>(s.wt_head), nh, 0u);
    if (s.n_wt_segs)
        hipLaunchKernelGGL(k_wt_head, dim3(s.n_wt_segs), dim3(256), (size_t)s.K * sizeof(uint32_t), st, s.z, s.wt_segs,
                           s.K, s.wt_head);
    return hipGetLastError();
}

hipError_t launch_topic_hist(const DevState &s, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(s.K, 256, 64)), dim3(256), 0, st, reinterpret_cast<uint32_t *>(s.nk),
                       (size_t)s.K, 0u);
    if (s.N)
        hipLaunchKernelGGL(k_topic_hist, dim3(grid_for(s.N, 256 * 16, 2048)), dim3(256), (size_t)s.K * sizeof(uint32_t),
                           st, s.z, s.N, s.K, s.nk);
    return hipGetLastError();
}

hipError_t launch_topk(const DevState &s, hipStream_t st) {
    hipLaunchKernelGGL(k_inv, dim3(grid_for(s.K, 256, 1 << 16)), dim3(256), 0, st, s.nk, s.K, s.vbeta, s.inv);
    if (s.V > s.VD)
        hipLaunchKernelGGL(k_zero_order, dim3(grid_for(s.K, 64, 1 << 16)), dim3(256),
                           (size_t)s.K * sizeof(float) + 64 * sizeof(uint32_t), st, s.inv, s.K, s.beta, s.zero_order);
    // head words need no LDS: where the tail words' count tables limit the
    // wavefronts per CU, they get a launch of their own at full occupancy
    const bool split = s.VD > 0 && s.V > s.VD && (size_t)s.lds_cu / wave_hist_bytes(s.K) < 16;
    const int w_tail = split || s.V == s.VD ? s.VD : 0;  // first word of the LDS launch
    if (w_tail > 0) {
        const WaveGrid g = wave_grid(s, (size_t)w_tail, 0, 1 << 16);
        hipLaunchKernelGGL(k_word_top, dim3(g.grid), dim3(64 * g.nwv), 0, st, 0, w_tail, s.VD, s.K, s.beta,
                           s.wt_head, s.tail_off, s.wt_tail, s.tail_nnz, s.inv, s.zero_order, s.wtop);
    }
    if (s.V > w_tail) {
        const WaveGrid g = wave_grid(s, (size_t)(s.V - w_tail), wave_hist_bytes(s.K), 1 << 16);
        hipLaunchKernelGGL(k_word_top, dim3(g.grid), dim3(64 * g.nwv), g.nwv * wave_hist_bytes(s.K), st, w_tail, s.V,
                           s.VD, s.K, s.beta, s.wt_head, s.tail_off, s.wt_tail, s.tail_nnz, s.inv, s.zero_order,
                           s.wtop);
    }
    return hipGetLastError();
}

template <bool COMPACT, bool N_DK2>
static void build_dt_kernels(const DevState &s, hipStream_t st) {
    // where a K-sized histogram per wavefront leaves fewer than 16 on a CU,
    // the short documents (the last ones: docs come longest first) are sorted
    // instead
    const bool sort_short = (size_t)s.lds_cu / dt_wave_bytes(s.K) < 16;
    const int q1 = sort_short ? s.D_long : s.D;
    if (q1 > 0) {
        const WaveGrid g = wave_grid(s, (size_t)q1, dt_wave_bytes(s.K), 1 << 18);
        hipLaunchKernelGGL((k_dt_build<COMPACT, N_DK2>), dim3(g.grid), dim3(64 * g.nwv), g.nwv * dt_wave_bytes(s.K), st,
                           0, q1, s.K, s.VD, s.topic_bits, s.soff, s.slot_wrd, s.slot_tok, s.z, s.z_doc, s.wtop,
                           s.aux_doc, s.doc_info, s.dt);
    }
    if (q1 < s.D) {
        const size_t lds = 4 * kDtShort * sizeof(uint32_t);  // four wavefronts, 2 KB each
        const unsigned grid = (unsigned)grid_for((size_t)(s.D - q1), 4, 1 << 18);
        hipLaunchKernelGGL((k_dt_sort<COMPACT, N_DK2>), dim3(grid), dim3(256), lds, st, q1, s.D, s.VD, s.topic_bits,
                           s.soff, s.slot_wrd, s.slot_tok, s.z, s.z_doc, s.wtop, s.aux_doc, s.doc_info, s.dt);
    }
}

hipError_t launch_build_dt(const DevState &s, hipStream_t st, bool n_dk2) {
    if (s.D == 0) return hipSuccess;
    if (s.compact) {
        if (n_dk2) build_dt_kernels<true, true>(s, st);
        else build_dt_kernels<true, false>(s, st);
    } else {
        if (n_dk2) build_dt_kernels<false, true>(s, st);
        else build_dt_kernels<false, false>(s, st);
    }
    return hipGetLastError();
}

size_t sample_lds_bytes(int K, int bd) {
    const int NS = (K + 63) >> 6;
    (void)bd;
    const size_t tab = K <= EZ_Q_TABLE_MAX_K ? 2 : 1;  // phi (and the slab running sums)
    return (tab * (size_t)NS * 64 + 2 * (size_t)NS + 4) * sizeof(float);  // + qt, qacc, f123[3], item
}

size_t sampler_static_lds() {
    const void *fs[6] = {reinterpret_cast<const void *>(&k_sample<true, true, true>),
                         reinterpret_cast<const void *>(&k_sample<true, true, false>),
                         reinterpret_cast<const void *>(&k_sample<true, false, false>),
                         reinterpret_cast<const void *>(&k_sample<false, true, true>),
                         reinterpret_cast<const void *>(&k_sample<false, true, false>),
                         reinterpret_cast<const void *>(&k_sample<false, false, false>)};
    size_t m = 0;
    for (const void *f : fs) {
        hipFuncAttributes a;
        if (hipFuncGetAttributes(&a, f) != hipSuccess) return (size_t)-1;
        m = std::max(m, a.sharedSizeBytes);
    }
    return m;
}

size_t max_lds_bytes(int K, int bd) {
    size_t m = sample_lds_bytes(K, bd);
    m = std::max(m, 4 * dt_wave_bytes(K));                           // WT tail, top-3, DT build
    m = std::max(m, (size_t)K * sizeof(double));                     // log-likelihood
    m = std::max(m, (size_t)K * sizeof(float) + 1024 * sizeof(float));  // fold-in tables
    m = std::max(m, foldin_lds_bytes(K));
    m = std::max(m, foldin_doc_lds(K));
    return m;
}

hipError_t launch_sample(const DevState &s, bool shortcut, hipStream_t st, uint32_t mode) {
    if (s.n_smp_segs == 0) return hipSuccess;
    const int bd = s.block_dim;
    const size_t lds = sample_lds_bytes(s.K, bd);
    SampleArgs A;
    A.segs = s.smp_segs;
    A.runs = s.runs;
    A.n_items = (uint32_t)s.n_smp_segs;
    A.VD = s.VD;
    A.K = s.K;
    A.alpha = s.alpha;
    A.beta = s.beta;
    A.wt_head = s.wt_head;
    A.tail_off = s.tail_off;
    A.wt_tail = s.wt_tail;
    A.tail_nnz = s.tail_nnz;
    A.inv = s.inv;
    A.wtop = s.wtop;
    A.aux_doc = s.aux_doc;
    A.dt = s.dt;
    A.z = s.z;
    A.z_doc = s.z_doc;
    A.counters = s.counters;
    A.queue = s.counters + kQueueStride;
    A.n_queues = s.grid_dim > 0 ? (uint32_t)s.n_queues : 0u;
    A.stats = s.stats;
    A.seed = s.seed;
    A.tb = s.topic_bits;
    A.mode = mode;
    unsigned grid = (unsigned)s.n_smp_segs;
    if (A.n_queues) {
        // queue heads start at zero (a fill kernel, so a captured sweep
        // replays it)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, s.counters + kQueueStride,
                           (size_t)kQueueStride * A.n_queues, 0u);
        grid = (unsigned)s.grid_dim;
    }
    const bool qtab = s.K <= EZ_Q_TABLE_MAX_K;  // implies compact rows
    if (shortcut && qtab)
        hipLaunchKernelGGL((k_sample<true, true, true>), dim3(grid), dim3(bd), lds, st, A);
    else if (shortcut && s.compact)
        hipLaunchKernelGGL((k_sample<true, true, false>), dim3(grid), dim3(bd), lds, st, A);
    else if (shortcut)
        hipLaunchKernelGGL((k_sample<true, false, false>), dim3(grid), dim3(bd), lds, st, A);
    else if (qtab)
        hipLaunchKernelGGL((k_sample<false, true, true>), dim3(grid), dim3(bd), lds, st, A);
    else if (s.compact)
        hipLaunchKernelGGL((k_sample<false, true, false>), dim3(grid), dim3(bd), lds, st, A);
    else
        hipLaunchKernelGGL((k_sample<false, false, false>), dim3(grid), dim3(bd), lds, st, A);
    return hipGetLastError();
}

hipError_t launch_advance_sweep(const DevState &s, hipStream_t st) {
    hipLaunchKernelGGL(k_inc_u32, dim3(1), dim3(1), 0, st, s.counters);
    return hipGetLastError();
}

hipError_t launch_loglik(const DevState &s, hipStream_t st) {
    if (s.n_smp_segs == 0) {
        hipLaunchKernelGGL(k_ll_final, dim3(1), dim3(256), 0, st, s.ll_part, 0, s.ll_out);
        return hipGetLastError();
    }
    const int bd = 256;
    hipLaunchKernelGGL(k_ll, dim3(s.n_smp_segs), dim3(bd), (size_t)s.K * sizeof(double), st, s.smp_segs, s.runs, s.VD,
                       s.K,
                       (double)s.alpha, (double)s.beta, (double)s.V * (double)s.beta, s.wt_head, s.tail_off,
                       s.wt_tail, s.tail_nnz, s.nk, s.aux_doc, s.dt, (int)s.compact, s.topic_bits, s.ll_part);
    hipLaunchKernelGGL(k_ll_final, dim3(1), dim3(256), 0, st, s.ll_part, s.n_smp_segs, s.ll_out);
    return hipGetLastError();
}

}  // namespace ez
