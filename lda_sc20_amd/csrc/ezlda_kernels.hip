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
// n_k = column sums of WT (WTDen_Sum_Update_Kernel utility.cu:1104-1126 plus
// the tail's atomicAdd in sparseMatrixAdd :1132-1204): one LDS histogram of
// z per workgroup, flushed with one atomic per nonzero topic.
__global__ void k_topic_hist(const uint16_t *__restrict__ z, uint32_t N, int K, int32_t *__restrict__ nk) {
    extern __shared__ uint32_t sh[];
    for (int k = threadIdx.x; k < K; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x * 2;
    const uint32_t *z2 = reinterpret_cast<const uint32_t *>(z);
    for (size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < N; i += stride) {
        if (i + 1 < N) {
            uint32_t p = z2[i >> 1];
            atomicAdd(&sh[p & 0xFFFFu], 1u);
            atomicAdd(&sh[p >> 16], 1u);
        } else {
            atomicAdd(&sh[z[i]], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        uint32_t c = sh[k];
        if (c) atomicAdd(&nk[k], (int32_t)c);
    }
}

// ------------------------------------------------------------- head WT
// Dense head rows (WTDen_Update_Kernel, utility.cu:1067-1102): one workgroup
// per segment of a head word's tokens, LDS histogram, atomic flush into the
// (pre-zeroed) row.
__global__ void k_wt_head(const uint16_t *__restrict__ z, const uint4 *__restrict__ segs, int K,
                          int32_t *__restrict__ wt_head) {
    extern __shared__ uint32_t sh[];
    const uint4 sg = segs[blockIdx.x];
    for (int k = threadIdx.x; k < K; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    for (uint32_t i = sg.y + threadIdx.x; i < sg.z; i += blockDim.x) atomicAdd(&sh[z[i]], 1u);
    __syncthreads();
    int32_t *row = wt_head + (size_t)sg.x * K;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        uint32_t c = sh[k];
        if (c) atomicAdd(&row[k], (int32_t)c);
    }
}

// ------------------------------------------------------------- tail WT
// Sorted sparse tail rows (WT_Update_Kernel, utility.cu:10-76): one
// wavefront per tail word; 16-bit LDS histogram, then a ballot/popcount
// compaction in ascending topic order, clearing the histogram as it goes.
__global__ void k_wt_tail(const uint16_t *__restrict__ z, const uint32_t *__restrict__ word_off, int VD, int V,
                          int K, const uint32_t *__restrict__ tail_off, uint32_t *__restrict__ wt_tail,
                          uint32_t *__restrict__ tail_nnz) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int KH = (K + 1) >> 1;
    uint32_t *h = sh + wv * KH;
    for (int k = lane; k < KH; k += 64) h[k] = 0;
    wave_lds_sync();
    const int ntail = V - VD;
    for (int tw = blockIdx.x * nwv + wv; tw < ntail; tw += gridDim.x * nwv) {
        const int w = VD + tw;
        const uint32_t b = word_off[w], e = word_off[w + 1];
        if (b == e) {
            if (lane == 0) tail_nnz[tw] = 0;
            continue;
        }
        for (uint32_t i = b + lane; i < e; i += 64) h16_add(h, z[i]);
        wave_lds_sync();
        const uint32_t base = tail_off[tw];
        uint32_t nz = 0;
        for (int j = 0; j < K; j += 64) {
            const int k = j + lane;
            const uint32_t c = k < K ? h16_get(h, (uint32_t)k) : 0u;
            const uint64_t m = __ballot(c != 0);
            if (c) wt_tail[base + nz + __popcll(m & lanemask_lt(lane))] = (uint32_t)k | (c << 16);
            nz += (uint32_t)__popcll(m);
            wave_lds_sync();
            if (k < K && !(k & 1)) h[k >> 1] = 0;
        }
        wave_lds_sync();
        if (lane == 0) tail_nnz[tw] = nz;
    }
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

// Per word the three largest phi_wk, ties to the lower topic
// (MaxTopicDense/Sparse_Update_Kernel, utility.cu:305-678, :679-1062).  One
// wavefront per word; a tail word's counts are scattered into a 16-bit LDS
// row first.  Per-lane lists are merged with a butterfly; the order is total,
// so the merge order does not matter.
__global__ void k_word_top(int V, int VD, int K, float beta, const int32_t *__restrict__ wt_head,
                           const uint32_t *__restrict__ tail_off, const uint32_t *__restrict__ wt_tail,
                           const uint32_t *__restrict__ tail_nnz, const float *__restrict__ inv,
                           ushort4 *__restrict__ wtop) {
    extern __shared__ uint32_t sh[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int KH = (K + 1) >> 1;
    uint32_t *h = sh + wv * KH;
    for (int k = lane; k < KH; k += 64) h[k] = 0;
    wave_lds_sync();
    for (int w = blockIdx.x * nwv + wv; w < V; w += gridDim.x * nwv) {
        Top3 t{-1.f, -1.f, -1.f, kNone, kNone, kNone};
        if (w < VD) {
            const int32_t *row = wt_head + (size_t)w * K;
            for (int k = lane; k < K; k += 64) top3_insert(t, ((float)row[k] + beta) * inv[k], (uint32_t)k);
        } else {
            const int tw = w - VD;
            const uint32_t b = tail_off[tw], n = tail_nnz[tw];
            for (uint32_t i = lane; i < n; i += 64) {
                const uint32_t e = wt_tail[b + i];
                atomicAdd(&h[(e & 0xFFFFu) >> 1], (e >> 16) << ((e & 1u) << 4));
            }
            wave_lds_sync();
            for (int k = lane; k < K; k += 64)
                top3_insert(t, ((float)h16_get(h, (uint32_t)k) + beta) * inv[k], (uint32_t)k);
            wave_lds_sync();
            for (uint32_t i = lane; i < n; i += 64) h[(wt_tail[b + i] & 0xFFFFu) >> 1] = 0;
            wave_lds_sync();
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            Top3 o;
            o.v0 = __shfl_xor(t.v0, off); o.k0 = __shfl_xor(t.k0, off);
            o.v1 = __shfl_xor(t.v1, off); o.k1 = __shfl_xor(t.k1, off);
            o.v2 = __shfl_xor(t.v2, off); o.k2 = __shfl_xor(t.k2, off);
            top3_insert(t, o.v0, o.k0);
            top3_insert(t, o.v1, o.k1);
            top3_insert(t, o.v2, o.k2);
        }
        if (lane == 0) wtop[w] = make_ushort4((unsigned short)t.k0, (unsigned short)t.k1, (unsigned short)t.k2, 0);
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
// builds a 16-bit LDS histogram, writes each slot's (n_dk1, n_dk2) at its
// word's top-2 topics (the re-pack of :125-161), and compacts the row in
// ascending topic order (:183-238).  A row starts on a 128-byte boundary and
// is zero-padded to a multiple of 32 entries.
template <bool COMPACT>
__global__ void k_dt_build(int D, int K, int VD, uint32_t tb, const uint32_t *__restrict__ soff,
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
    for (int q = blockIdx.x * nwv + wv; q < D; q += gridDim.x * nwv) {
        const uint32_t b = soff[q], e = soff[q + 1];
        for (uint32_t i = b + lane; i < e; i += 64) {
            const uint32_t t = slot_wrd[i] < (uint32_t)VD ? z_doc[i] : z[slot_tok[i]];
            h16_add(h, t);
            atomicOr(&bm[t >> 5], 1u << (t & 31u));
        }
        wave_lds_sync();
        for (uint32_t i = b + lane; i < e; i += 64) {
            const ushort4 tp = wtop[slot_wrd[i]];
            const uint32_t n1 = h16_get(h, tp.x);
            const uint32_t n2 = tp.y != kNone ? h16_get(h, tp.y) : 0u;
            aux_doc[i] = n1 | (n2 << 16);
        }
        wave_lds_sync();
        // Compaction over the bitmap: lane j owns topics [32j, 32j + 32) of
        // each 64-word chunk, counts its entries, takes its place from a
        // wave prefix sum and writes them in ascending topic order; then it
        // clears its counts and bitmap word.  Cost follows the doc's topics,
        // not K.
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
            for (uint32_t x = word; x; x &= x - 1u) {  // the lane's own 16 count words
                const uint32_t k = ((uint32_t)wi << 5) + (uint32_t)__ffs(x) - 1u;
                h[k >> 1] = 0;
            }
            if (wi < NW) bm[wi] = 0;
            nz += __shfl(incl, 63);
        }
        wave_lds_sync();
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
    }
}

// ------------------------------------------------------------- sampler
// Checkpoints of the S running sum kept per lane (one per block of kBlkU4
// 16-byte loads, i.e. 4*kBlkU4 DT entries).
static constexpr int kCP = EZ_SAMPLE_CP;
static constexpr int kBlkU4 = EZ_SAMPLE_BLK;

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

template <bool COMPACT>
__device__ __forceinline__ float sum_block(const RowBlock &r, float s, uint32_t tb) {
    const uint4 *e = r.e;
#pragma unroll
    for (int i = 0; i < kBlkU4; ++i) {
        if (COMPACT) {  // two 16-bit entries per dword, low half first
            const uint32_t w[4] = {e[i].x, e[i].y, e[i].z, e[i].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s = __builtin_fmaf(c16_count(w[q], tb), phi_c16(w[q], tb), s);
                s = __builtin_fmaf((float)(w[q] >> (16u + tb)), phi_c16(w[q] >> 16, tb), s);
            }
        } else {
            s = __builtin_fmaf(dt_count(e[i].x), phi_at0(e[i].x), s);
            s = __builtin_fmaf(dt_count(e[i].y), phi_at0(e[i].y), s);
            s = __builtin_fmaf(dt_count(e[i].z), phi_at0(e[i].z), s);
            s = __builtin_fmaf(dt_count(e[i].w), phi_at0(e[i].w), s);
        }
    }
    return s;
}

// Sampler work item: a word, or a doc tile / token segment of a head word,
// given as a range of runs.  A run is up to run_cap consecutive tokens of the
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
    const uint4 *runs;  // (first token, first doc slot, doc, tokens)
    uint32_t n_items;
    int VD, K;
    float alpha, beta;
    const int32_t *wt_head;
    const uint32_t *tail_off, *wt_tail, *tail_nnz;
    const float *inv;
    const ushort4 *wtop;
    const uint32_t *aux_doc;  // by doc slot: n_dk1 | n_dk2 << 16
    const uint2 *doc_info;
    const uint32_t *dt;
    uint16_t *z, *z_doc;      // by token (word-major) and by doc slot
    const uint32_t *counters;
    uint32_t *queue;  // persistent mode: one item counter per queue
    uint32_t n_queues;
    uint32_t tb;  // compact rows: topic bits
    unsigned long long *stats;
    uint64_t seed;
};

template <bool SHORTCUT, bool COMPACT>
__device__ __forceinline__ void sample_item(const SampleArgs &A, uint32_t item, float *smf, float *f123,
                                            uint32_t sweep, uint32_t &n_full, uint32_t &n_scanned) {
    const int K = A.K, NS = (K + 63) >> 6;
    const float alpha = A.alpha, beta = A.beta;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
    float *phi = smf;              // NS*64
    float *qs = phi + NS * 64;     // NS*64: running sums inside each slab
    float *qt = qs + NS * 64;      // NS: slab totals
    float *qacc = qt + NS;         // NS: running sum of the slab totals
    const uint4 sg = A.segs[item];
    if (sg.x == 0xFFFFFFFFu) return;  // padding item (uniform over the block)
    const int w = (int)sg.x;
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
    for (int j = tid; j < NS; j += blockDim.x) {
        const int k0 = j * 64, ke = min(k0 + 64, K);
        float run = 0.0f;
        for (int k = k0; k < ke; ++k) {
            run = run + alpha * phi[k];
            qs[k] = run;
        }
        qt[j] = run;
    }
    __syncthreads();
    if (tid == 0) {
        float acc = 0.0f;
        for (int j = 0; j < NS; ++j) {
            acc = acc + qt[j];
            qacc[j] = acc;
        }
    }
    __syncthreads();
    const float Q = qacc[NS - 1];
    const float f1 = f123[0], f2 = f123[1], f3 = f123[2];
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
        const uint32_t t0 = rn.x, p0 = rn.y, len = rn.w;
        const uint32_t a = A.aux_doc[p0];
        const uint2 di = A.doc_info[rn.z];
        const uint32_t n1 = a & 0xFFFFu, n2 = a >> 16, nd = di.y >> 16;
        const float P1 = ((float)n1 + alpha) * f1;
        // tokens of the run that need the full draw; shortcut tokens take k1
        uint32_t need = 0;
        for (uint32_t j = 0; j < len; ++j) {
            bool full = true;
            if (SHORTCUT) {
                const float u = ez_u24(ez_philox_x0(A.seed, t0 + j, sweep, EZ_TAG_SAMPLE));
                const float maxS = (float)n2 * f2 + (float)(nd - n1 - n2) * f3;
                const float B = (P1 + maxS) + Q;
                full = !(u * B < P1 * kMargin);
            }
            if (full) {
                need |= 1u << j;
            } else {
                A.z[t0 + j] = (uint16_t)k1;
                if (head) A.z_doc[p0 + j] = (uint16_t)k1;
            }
        }
        if (!need) continue;
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
        float cp[kCP];
        float S = 0.0f;
        RowBlock cur = load_block(rp);
#pragma unroll
        for (int b = 0; b < kCP; ++b) {
            RowBlock nxt;
            if ((uint32_t)b + 1u < nb) nxt = load_block(rp + kBlkU4 * (b + 1));
            if ((uint32_t)b < nb) S = sum_block<COMPACT>(cur, S, tb);
            cp[b] = S;
            cur = nxt;
        }
        for (uint32_t b = kCP; b < nb; ++b) {
            RowBlock nxt;
            if (b + 1u < nb) nxt = load_block(rp + kBlkU4 * (b + 1u));
            S = sum_block<COMPACT>(cur, S, tb);
            cur = nxt;
        }
        const float Aa = P1 + S;
        const float total = Aa + Q;
        while (need) {
            const uint32_t j = (uint32_t)__ffs(need) - 1u;
            need &= need - 1u;
            const uint32_t t = t0 + j;
            const float u = ez_u24(ez_philox_x0(A.seed, t, sweep, EZ_TAG_SAMPLE));
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
                uint32_t b0 = 0, b1 = nb;
                if (bs >= 0) {
#pragma unroll
                    for (int b = 0; b < kCP - 1; ++b)
                        if (b + 1 == bs) run = cp[b];
                    b0 = (uint32_t)bs;
                    b1 = b0 + 1u;
                } else if (nb > (uint32_t)kCP) {
                    run = cp[kCP - 1];
                    b0 = kCP;
                } else {
                    b1 = 0;  // no running sum exceeds t2
                }
                bool found = false;
                for (uint32_t g = b0 * kBlkU4; g < b1 * kBlkU4 && !found; ++g) {
                    const uint4 e = rp[g];
                    if (COMPACT) {
                        const uint32_t w[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (found) break;
                            const float ra = __builtin_fmaf(c16_count(w[q], tb), phi_c16(w[q], tb), run);
                            const float rb =
                                __builtin_fmaf((float)(w[q] >> (16u + tb)), phi_c16(w[q] >> 16, tb), ra);
                            if (ra > t2) {
                                newz = c16_topic(w[q], tb);
                                found = true;
                            } else if (rb > t2) {
                                newz = c16_topic(w[q] >> 16, tb);
                                found = true;
                            }
                            run = rb;
                        }
                    } else {
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
                    lo = k0;
                    hi = ke;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (qs[mid] > t3) hi = mid; else lo = mid + 1;
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
template <bool SHORTCUT, bool COMPACT>
__global__ void __launch_bounds__(1024) k_sample(const SampleArgs A) {
    // the only LDS of the kernel, so phi starts at address 0 and a DT entry's
    // low half is directly the LDS address of its phi value
    extern __shared__ float smf[];
    const int NS = (A.K + 63) >> 6;
    float *f123 = smf + 2 * NS * 64 + 2 * NS;  // after phi, qs, qt, qacc
    uint32_t &s_item = *reinterpret_cast<uint32_t *>(f123 + 3);
    const uint32_t sweep = A.counters[0];
    uint32_t n_full = 0, n_scanned = 0;
    if (A.n_queues == 0) {
        sample_item<SHORTCUT, COMPACT>(A, blockIdx.x, smf, f123, sweep, n_full, n_scanned);
    } else {
        const uint32_t nq = A.n_queues, x0 = blockIdx.x % nq;
        for (uint32_t h = 0; h < nq; ++h) {
            const uint32_t q = (x0 + h) % nq;
            while (true) {
                __syncthreads();  // the previous item is done with s_item and LDS
                if (threadIdx.x == 0) s_item = atomicAdd(&A.queue[q], 1u);
                __syncthreads();
                const uint64_t item = (uint64_t)q + (uint64_t)s_item * nq;
                if (item >= A.n_items) break;
                sample_item<SHORTCUT, COMPACT>(A, (uint32_t)item, smf, f123, sweep, n_full, n_scanned);
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
// (profilePerp/utility.cu:2858-2928).  One workgroup per sampler item; the
// per-item partials are summed in fixed order by k_ll_final.
__global__ void __launch_bounds__(1024) k_ll(const uint4 *__restrict__ segs, const uint4 *__restrict__ runs, int VD,
                                             int K, double alpha, double beta,
                                             double vbeta, const int32_t *__restrict__ wt_head,
                                             const uint32_t *__restrict__ tail_off,
                                             const uint32_t *__restrict__ wt_tail,
                                             const uint32_t *__restrict__ tail_nnz, const int32_t *__restrict__ nk,
                                             const uint2 *__restrict__ doc_info, const uint32_t *__restrict__ dt,
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
    double acc = 0.0;
    for (uint32_t r = sg.y + (uint32_t)wv; r < sg.z; r += (uint32_t)nwv) {
        const uint4 rn = runs[r];  // rn.w tokens of this word in doc rn.z share the term
        const uint2 di = doc_info[rn.z];
        const uint32_t nnz = di.y & 0xFFFFu;
        double s = 0.0;
        if (compact) {
            const uint16_t *r16 = reinterpret_cast<const uint16_t *>(dt) + di.x;
            for (uint32_t i = lane; i < nnz; i += 64) {
                const uint32_t e = r16[i];
                s += (double)(e >> tb) * phi[e & ((1u << tb) - 1u)];
            }
        } else {
            for (uint32_t i = lane; i < nnz; i += 64) {
                const uint32_t e = dt[di.x + i];
                s += (double)(e >> 16) * phi[dt_topic(e)];
            }
        }
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
        acc += (double)rn.w * log((s + q) / ((double)(di.y >> 16) + (double)K * alpha));
    }
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

static constexpr int kFB = 8, kFCP = 16;  // fold-in S blocks and checkpoints

// Running FMA sum over kFB row entries (LDS) with phi gathered from the
// frozen table: the gathers are issued together, the FMAs in order.
__device__ __forceinline__ float foldin_block(const uint32_t *row, const float *__restrict__ ph, float s) {
    uint32_t e[kFB];
    float f[kFB];
#pragma unroll
    for (int u = 0; u < kFB; ++u) e[u] = row[u];
#pragma unroll
    for (int u = 0; u < kFB; ++u) f[u] = ph[e[u] & 0xFFFFu];
#pragma unroll
    for (int u = 0; u < kFB; ++u) s = __builtin_fmaf((float)(e[u] >> 16), f[u], s);
    return s;
}

// One fold-in sweep for every active request slot (rid, first token, n,
// sweep): one workgroup per slot.  The doc's topic counts are histogrammed
// in LDS from its current z and compacted into an ascending row; then one
// lane per token draws from S (the row, FMA running sum, phi gathered from
// the frozen table) and Q (the word's precomputed running sums).  The first
// sweep draws the initial topics; the last one writes the doc's counts.
__global__ void __launch_bounds__(256) k_foldin_step(
    const uint4 *__restrict__ slots, int nslots, int K, int sweeps_total, float alpha,
    const float *__restrict__ phi, const float *__restrict__ qs, const float *__restrict__ qacc,
    const uint32_t *__restrict__ wpool, uint16_t *__restrict__ zpool, int32_t *__restrict__ counts_out,
    uint64_t seed) {
    extern __shared__ uint32_t shf[];
    __shared__ uint32_t s_nz;
    const int NS = (K + 63) >> 6;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, bd = blockDim.x;
    const int KH = (K + 1) >> 1, KR = (K + kFB - 1) / kFB * kFB;
    uint32_t *h = shf;        // the doc's 16-bit topic counts
    uint32_t *row = h + KH;   // its sorted row, zero padded to a block
    for (int k = tid; k < KH; k += bd) h[k] = 0;
    __syncthreads();
    for (int si = blockIdx.x; si < nslots; si += gridDim.x) {
        const uint4 sl = slots[si];
        const uint32_t rid = sl.x, off = sl.y, n = sl.z, sweep = sl.w;
        const uint64_t key = (uint64_t)rid << 32;
        uint16_t *z = zpool + off;
        const uint32_t *wd = wpool + off;
        if (sweep == 0) {
            for (uint32_t i = tid; i < n; i += bd)
                z[i] = (uint16_t)(ez_philox_x0(seed, key | i, 0u, EZ_TAG_FOLDIN_INIT) % (uint32_t)K);
            __syncthreads();
        }
        for (uint32_t i = tid; i < n; i += bd) h16_add(h, z[i]);
        __syncthreads();
        if (wv == 0) {  // one wavefront compacts the row (and clears the counts)
            uint32_t nz = 0;
            for (int j = 0; j < K; j += 64) {
                const int k = j + lane;
                const uint32_t c = k < K ? h16_get(h, (uint32_t)k) : 0u;
                const uint64_t m = __ballot(c != 0);
                if (c) row[nz + __popcll(m & lanemask_lt(lane))] = (uint32_t)k | (c << 16);
                nz += (uint32_t)__popcll(m);
                wave_lds_sync();
                if (k < K && !(k & 1)) h[k >> 1] = 0;
            }
            // zero entries up to the next block: count 0 adds exactly nothing
            if ((uint32_t)lane < ((kFB - (nz % kFB)) % kFB)) row[nz + lane] = 0u;
            if (lane == 0) s_nz = nz;
        }
        __syncthreads();
        const uint32_t nz = s_nz;
        for (uint32_t base = 0; base < n; base += (uint32_t)bd) {
            const uint32_t i = base + (uint32_t)tid;
            if (i < n) {
                const uint32_t w = wd[i];
                const float *ph = phi + (size_t)w * K;
                const float *qa = qacc + (size_t)w * NS;
                const float u = ez_u24(ez_philox_x0(seed, key | i, sweep, EZ_TAG_FOLDIN));
                // S over the row in blocks of kFB entries (8 independent phi
                // gathers in flight per lane), zero padding adding nothing;
                // cp[b] = running sum after block b, for the first kFCP blocks
                const uint32_t nb = (nz + kFB - 1) / kFB;
                float cp[kFCP];
                float S = 0.0f;
#pragma unroll
                for (int b = 0; b < kFCP; ++b) {
                    if ((uint32_t)b < nb) S = foldin_block(row + kFB * b, ph, S);
                    cp[b] = S;
                }
                for (uint32_t b = kFCP; b < nb; ++b) S = foldin_block(row + kFB * b, ph, S);
                const float target = u * (S + qa[NS - 1]);
                int zz = -1;
                if (target < S) {
                    // the first block whose running sum exceeds the target,
                    // then its entries in order (a running sum above the
                    // target exists: the last one is S)
                    int bs = -1;
#pragma unroll
                    for (int b = kFCP - 1; b >= 0; --b)
                        if (cp[b] > target) bs = b;
                    float run = 0.0f;
                    uint32_t g0 = 0, g1 = nz;
                    if (bs >= 0) {
#pragma unroll
                        for (int b = 0; b < kFCP - 1; ++b)
                            if (b + 1 == bs) run = cp[b];
                        g0 = (uint32_t)bs * kFB;
                        g1 = min(nz, g0 + kFB);
                    } else {
                        run = cp[kFCP - 1];
                        g0 = kFCP * kFB;
                    }
                    for (uint32_t j = g0; j < g1; ++j) {
                        const uint32_t e = row[j];
                        run = __builtin_fmaf((float)(e >> 16), ph[e & 0xFFFFu], run);
                        if (run > target) {
                            zz = (int)(e & 0xFFFFu);
                            break;
                        }
                    }
                    if (zz < 0)  // not reached: kept as the canonical fallback
                        for (uint32_t j = 0; j < nz; ++j)
                            if (row[j] >> 16 && ph[row[j] & 0xFFFFu] > 0.0f) zz = (int)(row[j] & 0xFFFFu);
                } else {
                    const float t2 = target - S;
                    int lo = 0, hi = NS;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (qa[mid] > t2) hi = mid; else lo = mid + 1;
                    }
                    int js = lo < NS ? lo : -1;
                    if (js < 0)
                        for (int j = NS - 1; j >= 0; --j)
                            if (qa[j] > (j > 0 ? qa[j - 1] : 0.0f)) {
                                js = j;
                                break;
                            }
                    if (js >= 0) {
                        const float t3 = t2 - (js > 0 ? qa[js - 1] : 0.0f);
                        const float *qw = qs + (size_t)w * K;
                        const int k0 = js * 64, ke = min(k0 + 64, K);
                        lo = k0;
                        hi = ke;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (qw[mid] > t3) hi = mid; else lo = mid + 1;
                        }
                        if (lo < ke) {
                            zz = lo;
                        } else {
                            for (int k = ke - 1; k >= k0; --k)
                                if (alpha * ph[k] > 0.0f) {
                                    zz = k;
                                    break;
                                }
                        }
                    }
                }
                z[i] = (uint16_t)(zz < 0 ? 0 : zz);
            }
        }
        if (sweep + 1 == (uint32_t)sweeps_total) {  // final counts of the doc
            __syncthreads();
            for (uint32_t i = tid; i < n; i += bd) h16_add(h, z[i]);
            __syncthreads();
            int32_t *co = counts_out + (size_t)si * K;
            for (int k = tid; k < K; k += bd) co[k] = (int32_t)h16_get(h, (uint32_t)k);
            __syncthreads();
            for (int k = tid; k < KH; k += bd) h[k] = 0;
        }
        __syncthreads();
    }
}

hipError_t launch_foldin_tables(const DevState &s, float *phi, float *qs, float *qacc, hipStream_t st) {
    hipLaunchKernelGGL(k_foldin_tables, dim3(grid_for(s.V, 1, 1 << 16)), dim3(256), (size_t)s.K * sizeof(float), st,
                       s.V, s.VD, s.K, s.alpha, s.beta, s.wt_head, s.tail_off, s.wt_tail, s.tail_nnz, s.inv, phi, qs,
                       qacc);
    return hipGetLastError();
}

static size_t foldin_wave_lds(int K) { return (size_t)(((K + 1) >> 1) + (K + kFB - 1) / kFB * kFB) * sizeof(uint32_t); }

int foldin_waves_per_block(int K) {
    (void)K;
    return 4;  // one workgroup of 4 wavefronts per document
}

hipError_t launch_foldin_step(const uint4 *slots, int nslots, int K, int sweeps_total, float alpha, const float *phi,
                              const float *qs, const float *qacc, const uint32_t *wpool, uint16_t *zpool,
                              int32_t *counts_out, uint64_t seed, hipStream_t st) {
    if (nslots <= 0) return hipSuccess;
    const int nwv = foldin_waves_per_block(K);
    const size_t lds = foldin_wave_lds(K);  // one document per workgroup
    hipLaunchKernelGGL(k_foldin_step, dim3(grid_for(nslots, 1, 1 << 16)), dim3(64 * nwv), lds, st, slots, nslots, K,
                       sweeps_total, alpha, phi, qs, qacc, wpool, zpool, counts_out, seed);
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

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_u4, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(dst), n);
    return hipGetLastError();
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
    const size_t ldsK = (size_t)s.K * sizeof(uint32_t);
    if (s.N) hipLaunchKernelGGL(k_topic_hist, dim3(grid_for(s.N, 256 * 16, 2048)), dim3(256), ldsK, st, s.z, s.N, s.K, s.nk);
    if (s.n_wt_segs)
        hipLaunchKernelGGL(k_wt_head, dim3(s.n_wt_segs), dim3(256), ldsK, st, s.z, s.wt_segs, s.K, s.wt_head);
    const int ntail = s.V - s.VD;
    if (ntail > 0) {
        const int nwv = 4;
        hipLaunchKernelGGL(k_wt_tail, dim3(grid_for(ntail, nwv, 1 << 16)), dim3(64 * nwv), nwv * wave_hist_bytes(s.K), st,
                           s.z, s.word_off, s.VD, s.V, s.K, s.tail_off, s.wt_tail, s.tail_nnz);
    }
    return hipGetLastError();
}

hipError_t launch_topk(const DevState &s, hipStream_t st) {
    hipLaunchKernelGGL(k_inv, dim3(grid_for(s.K, 256, 1 << 16)), dim3(256), 0, st, s.nk, s.K, s.vbeta, s.inv);
    const int nwv = 4;
    hipLaunchKernelGGL(k_word_top, dim3(grid_for(s.V, nwv, 1 << 16)), dim3(64 * nwv), nwv * wave_hist_bytes(s.K), st,
                       s.V, s.VD, s.K, s.beta, s.wt_head, s.tail_off, s.wt_tail, s.tail_nnz, s.inv, s.wtop);
    return hipGetLastError();
}

hipError_t launch_build_dt(const DevState &s, hipStream_t st) {
    if (s.D == 0) return hipSuccess;
    const int nwv = 4;
    if (s.compact)
        hipLaunchKernelGGL(k_dt_build<true>, dim3(grid_for(s.D, nwv, 1 << 18)), dim3(64 * nwv),
                           nwv * dt_wave_bytes(s.K), st, s.D, s.K, s.VD, s.topic_bits, s.soff, s.slot_wrd, s.slot_tok, s.z,
                           s.z_doc, s.wtop, s.aux_doc, s.doc_info, s.dt);
    else
        hipLaunchKernelGGL(k_dt_build<false>, dim3(grid_for(s.D, nwv, 1 << 18)), dim3(64 * nwv),
                           nwv * dt_wave_bytes(s.K), st, s.D, s.K, s.VD, s.topic_bits, s.soff, s.slot_wrd, s.slot_tok, s.z,
                           s.z_doc, s.wtop, s.aux_doc, s.doc_info, s.dt);
    return hipGetLastError();
}

size_t sample_lds_bytes(int K, int bd) {
    const int NS = (K + 63) >> 6;
    (void)bd;
    return (2 * (size_t)NS * 64 + 2 * (size_t)NS + 4) * sizeof(float);  // + f123[3], item
}

size_t sampler_static_lds() {
    const void *fs[4] = {reinterpret_cast<const void *>(&k_sample<true, true>),
                         reinterpret_cast<const void *>(&k_sample<true, false>),
                         reinterpret_cast<const void *>(&k_sample<false, true>),
                         reinterpret_cast<const void *>(&k_sample<false, false>)};
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
    m = std::max(m, foldin_wave_lds(K) + 16);
    return m;
}

hipError_t launch_sample(const DevState &s, bool shortcut, hipStream_t st) {
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
    A.doc_info = s.doc_info;
    A.dt = s.dt;
    A.z = s.z;
    A.z_doc = s.z_doc;
    A.counters = s.counters;
    A.queue = s.counters + 8;  // counters[8 .. 8 + n_queues)
    A.n_queues = s.grid_dim > 0 ? (uint32_t)s.n_queues : 0u;
    A.stats = s.stats;
    A.seed = s.seed;
    A.tb = s.topic_bits;
    unsigned grid = (unsigned)s.n_smp_segs;
    if (A.n_queues) {
        // queue heads start at zero (a fill kernel, so a captured sweep
        // replays it)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, s.counters + 8, (size_t)A.n_queues, 0u);
        grid = (unsigned)s.grid_dim;
    }
    if (shortcut && s.compact)
        hipLaunchKernelGGL((k_sample<true, true>), dim3(grid), dim3(bd), lds, st, A);
    else if (shortcut)
        hipLaunchKernelGGL((k_sample<true, false>), dim3(grid), dim3(bd), lds, st, A);
    else if (s.compact)
        hipLaunchKernelGGL((k_sample<false, true>), dim3(grid), dim3(bd), lds, st, A);
    else
        hipLaunchKernelGGL((k_sample<false, false>), dim3(grid), dim3(bd), lds, st, A);
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
                       s.wt_tail, s.tail_nnz, s.nk, s.doc_info, s.dt, (int)s.compact, s.topic_bits, s.ll_part);
    hipLaunchKernelGGL(k_ll_final, dim3(1), dim3(256), 0, st, s.ll_part, s.n_smp_segs, s.ll_out);
    return hipGetLastError();
}

}  // namespace ez
