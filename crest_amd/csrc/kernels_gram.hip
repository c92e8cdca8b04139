// Gram SYRK G += S^T S on MFMA over a densified chunk tile, deterministic split-K
// reduction into the FP64 Gram, FP64 centring into the covariance, and the
// post-processing of the eigensolve (descending order, floor at 0, sign rule).
//
// Tile layout (written by k_densify): tile[(g*d_pad + h)*8 + r] holds S[8g + r, h].
// G is d_pad x d_pad row-major FP64; only blocks bi <= bj (of 128 x 128) are
// produced, which covers every a <= b entry the reference reads (kernels.rs:319-322).
#include "common.hpp"

namespace crest {

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) double f64x4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

__device__ inline bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}
__device__ inline bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// One K-step of 16 on v_mfma_f32_32x32x16_{bf16,f16}: TY 0 = bf16, 1 = fp16.
template <int TY> __device__ inline f32x16 mfma_x16(u32x4 a, u32x4 b, f32x16 c) {
    if constexpr (TY == 0)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
}

// blk -> (bi, bj) over the upper triangle of nb x nb blocks, row by row.
__device__ inline void tri_block(int blk, int nb, int &bi, int &bj) {
    int r = 0;
    while (blk >= nb - r) {
        blk -= nb - r;
        ++r;
    }
    bi = r;
    bj = r + blk;
}

// Split-K partition: workgroup w handles split w % n_split of block w / n_split, so
// with n_split a multiple of 8 every split (one range of tile rows) lives on one XCD
// and the row slab it streams is shared through that XCD's L2 by all its blocks.
__device__ inline void gram_work(int n_split, int nb, int &blk, int &split, int &bi, int &bj) {
    split = blockIdx.x % n_split;
    blk = blockIdx.x / n_split;
    tri_block(blk, nb, bi, bj);
}

// Output of a 32x32 MFMA tile: register r of lane l is C[8*(r/4) + 4*(l/32) + r%4][l%32],
// stored (or, with add, added) at dst[m * ld + n].
__device__ inline void store_partial32(double *dst, const double (&v)[16], int m0, int n0, int lane, int ld = kGramBlock,
                                       bool add = false) {
    const int col = lane & 31, half = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + 8 * (r / 4) + 4 * half + (r % 4);
        double *o = dst + (int64_t)m * ld + n0 + col;
        *o = add ? *o + v[r] : v[r];
    }
}

// Destination of a work item: its partial slot, or (slot < 0: the item is the only
// one for its block in this launch) the block of its subset's Gram, added in place.
__device__ inline double *item_dst(const GramItem &w, int bi, int bj, int d_pad, double *partial, double *gram,
                                   int &ld, bool &add) {
    if (w.slot >= 0) {
        ld = kGramBlock;
        add = false;
        return partial + w.slot * kGramBlock * kGramBlock;
    }
    ld = d_pad;
    add = true;
    return gram + (int64_t)w.subset * d_pad * d_pad + (int64_t)bi * kGramBlock * d_pad + bj * kGramBlock;
}

// ---------------------------------------------------------------- split 16-bit (BF16 or FP16)
// S = H + L (16-bit hi and the 16-bit rounding of the residual); S^T S = H^T H +
// H^T L + L^T H + L^T L.  The first three run on MFMA.  L^T L is ~2^-2b of |s_a s_b|
// per row (b = 8 for bf16, 11 for fp16): off the diagonal its signs vary and it sums
// to nothing measurable, but on the diagonal it is a positive bias, so k_densify adds
// the exact diagonal Σ l² into the Gram instead.  FP32 accumulators are added into
// FP64 every FLUSH K-steps (16 rows each).
// Block (bi, bj) over the K-steps [p0, p1) of 16 rows into dst (128 x 128 FP64).
template <int TY, int FLUSH>
__device__ inline void gram_x16_block(const uint4 *__restrict__ th, const uint4 *__restrict__ tl, int d_pad, int bi,
                                      int bj, int64_t p0, int64_t p1, double *dst, int ld = kGramBlock,
                                      bool add = false) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wm = wave >> 1, wn = wave & 1;
    const int col = lane & 31, half = lane >> 5;
    const int ga = bi * kGramBlock + wm * 64 + col, gb = bj * kGramBlock + wn * 64 + col;
    f32x16 acc[2][2];
    double dacc[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x16){};
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] = 0.0;
        }
    // Operands of K-step pair p (rows 32p..32p+31): hi and lo of the wave's two A and two
    // B column runs.  The loads of step p+1 are issued before the MFMAs of step p
    // (two register sets, alternating), so their latency hides behind matrix work.
    struct Ops {
        u32x4 ah[2], al[2], bh[2], bl[2];
    };
    const u32x4 *th4 = reinterpret_cast<const u32x4 *>(th), *tl4 = reinterpret_cast<const u32x4 *>(tl);
    auto load = [&](int64_t p, Ops &o) {
        const int64_t u = (2 * p + half) * (int64_t)d_pad;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            o.ah[i] = th4[u + ga + 32 * i];
            o.al[i] = tl4[u + ga + 32 * i];
            o.bh[i] = th4[u + gb + 32 * i];
            o.bl[i] = tl4[u + gb + 32 * i];
        }
    };
    int cnt = 0;
    auto step = [&](const Ops &o) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = mfma_x16<TY>(o.ah[i], o.bh[j], acc[i][j]);
                acc[i][j] = mfma_x16<TY>(o.ah[i], o.bl[j], acc[i][j]);
                acc[i][j] = mfma_x16<TY>(o.al[i], o.bh[j], acc[i][j]);
            }
        if (++cnt == FLUSH) {
            cnt = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
                    acc[i][j] = (f32x16){};
                }
        }
    };
    // Prefetches are unconditional (clamped to the last step; the extra copy is never
    // used), so no branch sits between a load and the wait for it; the empty asm
    // statement with a memory clobber keeps the compiler from sinking the loads down
    // to their use (it emits no instruction).
    if (p0 < p1) {
        Ops oa, ob;
        load(p0, oa);
        for (int64_t p = p0;; p += 2) {
            load(p + 1 < p1 ? p + 1 : p1 - 1, ob);
            asm volatile("" ::: "memory");
            step(oa);
            if (p + 1 >= p1) break;
            load(p + 2 < p1 ? p + 2 : p1 - 1, oa);
            asm volatile("" ::: "memory");
            step(ob);
            if (p + 2 >= p1) break;
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
            store_partial32(dst, dacc[i][j], wm * 64 + i * 32, wn * 64 + j * 32, lane, ld, add);
        }
}

// LDS-staged form of gram_x16_block with the same accumulation order and result.  Per
// K-step the workgroup's distinct operands are hi and lo of its 128 A and 128 B
// columns over both 8-row groups: 1,024 uint4 (16 KB), where the per-wave form loads
// each of them twice (two waves share every A run, two every B run).  Here they go
// from global memory straight into LDS (global_load_lds_dwordx4: each lane's 16 B
// land at the wave's LDS base + 16·lane, no registers held), NB stages of KS K-steps
// in a ring with NB-1 stages in flight, and the waves read their MFMA operands from
// LDS.  LDS index of an operand within a stage: ks*1024 + q*256 + h*128 + c, q = 0 A
// hi, 1 A lo, 2 B hi, 3 B lo; h = row group of the K-step; c = column of the run.
// Waits are explicit: vmcnt counts this wave's loads in issue order, so waiting for
// all but the (NB-2)·4·KS most recent ones completes stage st; the barrier after it
// makes every wave's part of the stage visible.  A workgroup fence would wait for all
// loads in flight (vmcnt(0)), so the barrier is a bare s_barrier between compiler-only
// memory barriers.
template <int KS, int NB> constexpr int gram_dma_lds_bytes() { return NB * KS * 1024 * 16; }
// one K-step per stage, three stages (48 KB): measured against (K-steps, stages) =
// (2, 3), (1, 4), (2, 4), (1, 2) and (1, 5) at 1,000,000 x 2,000, the fastest Gram stage
// (23.6 ms against 24.0-27.9; a smaller ring also leaves LDS for the densify of the
// next tile, which runs beside the Gram)
constexpr int kGramKS = 1, kGramNB = 3;

__device__ inline void lds_dma16(const u32x4 *src, u32x4 *dst) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// s_waitcnt immediate for vmcnt(n), expcnt and lgkmcnt not waited on (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x70 | 0xF00; }

__device__ inline void bare_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int TY, int FLUSH, int KS, int NB>
__device__ inline void gram_x16_block_dma(const uint4 *__restrict__ th, const uint4 *__restrict__ tl, int d_pad,
                                          int bi, int bj, int64_t p0, int64_t p1, double *dst, u32x4 *lds,
                                          int ld = kGramBlock, bool add = false) {
    static_assert(FLUSH % KS == 0, "the FP64 flush falls on stage boundaries");
    static_assert(NB >= 2 && (NB - 2) * 4 * KS < 64, "vmcnt holds at most 63 loads");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
    const int half = lane >> 5, col = lane & 31;
    f32x16 acc[2][2];
    double dacc[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x16){};
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] = 0.0;
        }
    int cnt = 0;
    auto mma = [&](const u32x4 *L) {
        u32x4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ah[i] = L[0 * 256 + half * 128 + wm * 64 + 32 * i + col];
            al[i] = L[1 * 256 + half * 128 + wm * 64 + 32 * i + col];
            bh[i] = L[2 * 256 + half * 128 + wn * 64 + 32 * i + col];
            bl[i] = L[3 * 256 + half * 128 + wn * 64 + 32 * i + col];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = mfma_x16<TY>(ah[i], bh[j], acc[i][j]);
                acc[i][j] = mfma_x16<TY>(ah[i], bl[j], acc[i][j]);
                acc[i][j] = mfma_x16<TY>(al[i], bh[j], acc[i][j]);
            }
    };
    // after every FLUSH K-steps counted from p0, as in gram_x16_block
    auto flush = [&](int done) {
        cnt += done;
        if (cnt == FLUSH) {
            cnt = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
                    acc[i][j] = (f32x16){};
                }
        }
    };
    if (p0 < p1) {
        const u32x4 *th4 = reinterpret_cast<const u32x4 *>(th), *tl4 = reinterpret_cast<const u32x4 *>(tl);
        const int64_t nst = (p1 - p0 + KS - 1) / KS;
        const int h = tid >> 7;
        const int wu = __builtin_amdgcn_readfirstlane(wave);
        const int64_t ca = (int64_t)bi * kGramBlock + (tid & 127), cb = (int64_t)bj * kGramBlock + (tid & 127);
        // stage st into ring slot st % NB; K-steps past p1 (the tail, and the stages
        // issued past the last one so that every iteration waits on the same count)
        // load the last K-step again: valid memory, never multiplied
        auto issue = [&](int64_t st) {
            u32x4 *B = lds + (int)(st % NB) * KS * 1024 + wu * 64;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                int64_t k = p0 + st * KS + ks;
                k = k < p1 ? k : p1 - 1;
                const int64_t u = (2 * k + h) * (int64_t)d_pad;
                lds_dma16(th4 + u + ca, B + ks * 1024 + 0 * 256);
                lds_dma16(tl4 + u + ca, B + ks * 1024 + 1 * 256);
                lds_dma16(th4 + u + cb, B + ks * 1024 + 2 * 256);
                lds_dma16(tl4 + u + cb, B + ks * 1024 + 3 * 256);
            }
        };
#pragma unroll
        for (int st = 0; st < NB - 1; ++st) issue(st);
        for (int64_t st = 0; st < nst; ++st) {
            // stages st .. st+NB-2 are in flight: complete stage st
            __builtin_amdgcn_s_waitcnt(waitcnt_vm((NB - 2) * 4 * KS));
            // every wave's part of stage st is in LDS, and every wave has consumed
            // stage st-1, whose slot the next issue refills
            bare_barrier();
            issue(st + NB - 1);
            const u32x4 *L = lds + (int)(st % NB) * KS * 1024;
            const int64_t left = p1 - p0 - st * KS;
            if (left >= KS) {
                // a full stage: straight-line code, so the LDS reads of K-step ks+1
                // overlap the MFMAs of K-step ks
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) mma(L + ks * 1024);
                flush(KS);
            } else {
                // the tail stage (left < KS K-steps; cnt stays below FLUSH)
                for (int ks = 0; ks < left; ++ks) mma(L + ks * 1024);
            }
            asm volatile("" ::: "memory");
        }
        // no load may land in LDS after the workgroup ends
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
            store_partial32(dst, dacc[i][j], wm * 64 + i * 32, wn * 64 + j * 32, lane, ld, add);
        }
}

template <int TY, int FLUSH, int KS, int NB>
__global__ __launch_bounds__(256) void k_gram_x16_dma(const uint4 *__restrict__ th, const uint4 *__restrict__ tl,
                                                         int64_t n_groups, int d_pad, int n_split, double *partial) {
    extern __shared__ u32x4 gram_lds[];
    const int nb = d_pad / kGramBlock;
    int blk, split, bi, bj;
    gram_work(n_split, nb, blk, split, bi, bj);
    const int64_t n_pairs = n_groups / 2;
    const int64_t per = (n_pairs + n_split - 1) / n_split;
    const int64_t p0 = (int64_t)split * per;
    const int64_t p1 = p0 + per < n_pairs ? p0 + per : n_pairs;
    gram_x16_block_dma<TY, FLUSH, KS, NB>(th, tl, d_pad, bi, bj, p0, p1,
                                          partial + ((int64_t)blk * n_split + split) * kGramBlock * kGramBlock,
                                          gram_lds);
}

template <int TY, int FLUSH>
__global__ __launch_bounds__(256) void k_gram_x16(const uint4 *__restrict__ th, const uint4 *__restrict__ tl,
                                                     int64_t n_groups, int d_pad, int n_split, double *partial) {
    const int nb = d_pad / kGramBlock;
    int blk, split, bi, bj;
    gram_work(n_split, nb, blk, split, bi, bj);
    const int64_t n_pairs = n_groups / 2;
    const int64_t per = (n_pairs + n_split - 1) / n_split;
    const int64_t p0 = (int64_t)split * per;
    const int64_t p1 = p0 + per < n_pairs ? p0 + per : n_pairs;
    gram_x16_block<TY, FLUSH>(th, tl, d_pad, bi, bj, p0, p1,
                              partial + ((int64_t)blk * n_split + split) * kGramBlock * kGramBlock);
}

// Persistent form over a work list (many independent cell subsets in one tile):
// workgroup w takes items w, w + gridDim.x, ...; an item is one Gram block over one
// row range of one subset's rows, written to its own partial slot.
template <int TY, int FLUSH>
__global__ __launch_bounds__(256) void k_gram_x16_items(const uint4 *__restrict__ th, const uint4 *__restrict__ tl,
                                                           int d_pad, const GramItem *items, int64_t n_items,
                                                           double *partial, double *gram) {
    const int nb = d_pad / kGramBlock;
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const GramItem w = items[it];
        if (w.blk < 0) continue;
        int bi, bj, ld;
        bool add;
        tri_block(w.blk, nb, bi, bj);
        double *dst = item_dst(w, bi, bj, d_pad, partial, gram, ld, add);
        gram_x16_block<TY, FLUSH>(th, tl, d_pad, bi, bj, w.g0 / 2, w.g1 / 2, dst, ld, add);
    }
}

// Work-list form of k_gram_x16_dma.  Between two items every wave has consumed the
// previous item's last stage (its MFMAs read it before the barrier) before the next
// item's first loads refill the ring.
template <int TY, int FLUSH, int KS, int NB>
__global__ __launch_bounds__(256) void k_gram_x16_items_dma(const uint4 *__restrict__ th,
                                                               const uint4 *__restrict__ tl, int d_pad,
                                                               const GramItem *items, int64_t n_items, double *partial,
                                                               double *gram) {
    extern __shared__ u32x4 gram_lds[];
    const int nb = d_pad / kGramBlock;
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const GramItem w = items[it];
        if (w.blk < 0) continue;
        int bi, bj, ld;
        bool add;
        tri_block(w.blk, nb, bi, bj);
        double *dst = item_dst(w, bi, bj, d_pad, partial, gram, ld, add);
        __syncthreads();
        gram_x16_block_dma<TY, FLUSH, KS, NB>(th, tl, d_pad, bi, bj, w.g0 / 2, w.g1 / 2, dst, gram_lds, ld, add);
    }
}

// ---------------------------------------------------------------- FP32
// v_mfma_f32_32x32x2_f32; lane half h takes rows 4h..4h+3 of each row group, so the
// four MFMAs of a group cover its eight rows.
template <int FLUSH>
__device__ inline void gram_f32_block(const float4 *__restrict__ t, int d_pad, int bi, int bj, int64_t g0, int64_t g1,
                                      double *dst, int ld = kGramBlock, bool add = false) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wm = wave >> 1, wn = wave & 1;
    const int col = lane & 31, half = lane >> 5;
    const int ga = bi * kGramBlock + wm * 64 + col, gb = bj * kGramBlock + wn * 64 + col;
    f32x16 acc[2][2];
    double dacc[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x16){};
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] = 0.0;
        }
    // Same operand prefetch as the split-BF16 kernel: group g+1's loads are issued
    // (unconditionally, clamped) before group g's MFMAs.
    typedef __attribute__((ext_vector_type(4))) float f32x4;
    const f32x4 *t4 = reinterpret_cast<const f32x4 *>(t);
    struct Ops {
        f32x4 a[2], b[2];
    };
    auto load = [&](int64_t g, Ops &o) {
        const int64_t u = g * (int64_t)d_pad;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            o.a[i] = t4[(u + ga + 32 * i) * 2 + half];
            o.b[i] = t4[(u + gb + 32 * i) * 2 + half];
        }
    };
    int cnt = 0;
    auto step = [&](const Ops &o) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(o.a[i][e], o.b[j][e], acc[i][j], 0, 0, 0);
        if (++cnt == FLUSH) {
            cnt = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
                    acc[i][j] = (f32x16){};
                }
        }
    };
    if (g0 < g1) {
        Ops oa, ob;
        load(g0, oa);
        for (int64_t g = g0;; g += 2) {
            load(g + 1 < g1 ? g + 1 : g1 - 1, ob);
            asm volatile("" ::: "memory");
            step(oa);
            if (g + 1 >= g1) break;
            load(g + 2 < g1 ? g + 2 : g1 - 1, oa);
            asm volatile("" ::: "memory");
            step(ob);
            if (g + 2 >= g1) break;
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) dacc[i][j][r] += (double)acc[i][j][r];
            store_partial32(dst, dacc[i][j], wm * 64 + i * 32, wn * 64 + j * 32, lane, ld, add);
        }
}

template <int FLUSH>
__global__ __launch_bounds__(256) void k_gram_f32(const float4 *__restrict__ t, int64_t n_groups, int d_pad,
                                                  int n_split, double *partial) {
    const int nb = d_pad / kGramBlock;
    int blk, split, bi, bj;
    gram_work(n_split, nb, blk, split, bi, bj);
    const int64_t per = (n_groups + n_split - 1) / n_split;
    const int64_t g0 = (int64_t)split * per;
    const int64_t g1 = g0 + per < n_groups ? g0 + per : n_groups;
    gram_f32_block<FLUSH>(t, d_pad, bi, bj, g0, g1, partial + ((int64_t)blk * n_split + split) * kGramBlock * kGramBlock);
}

template <int FLUSH>
__global__ __launch_bounds__(256) void k_gram_f32_items(const float4 *__restrict__ t, int d_pad, const GramItem *items,
                                                        int64_t n_items, double *partial, double *gram) {
    const int nb = d_pad / kGramBlock;
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const GramItem w = items[it];
        if (w.blk < 0) continue;
        int bi, bj, ld;
        bool add;
        tri_block(w.blk, nb, bi, bj);
        double *dst = item_dst(w, bi, bj, d_pad, partial, gram, ld, add);
        gram_f32_block<FLUSH>(t, d_pad, bi, bj, w.g0, w.g1, dst, ld, add);
    }
}

// ---------------------------------------------------------------- FP64
// v_mfma_f64_16x16x4_f64; lane quarter q = l/16 takes rows 2q, 2q+1 of each row
// group.  Output (FP64 layout, rows interleaved across the lane quarters): register
// r of lane l is C[4*r + l/16][l%16].
__device__ inline void gram_f64_block(const double2 *__restrict__ t, int d_pad, int bi, int bj, int64_t g0, int64_t g1,
                                      double *dst, int ld = kGramBlock, bool add = false) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wm = wave >> 1, wn = wave & 1;
    const int c16 = lane & 15, q = lane >> 4;
    const int ga = bi * kGramBlock + wm * 64 + c16, gb = bj * kGramBlock + wn * 64 + c16;
    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (f64x4){};
    for (int64_t g = g0; g < g1; ++g) {
        const int64_t u = g * (int64_t)d_pad;
        double2 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = t[(u + ga + 16 * i) * 4 + q];
            b[i] = t[(u + gb + 16 * i) * 4 + q];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, acc[i][j], 0, 0, 0);
            }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = wm * 64 + i * 16 + 4 * r + q;
                const int n = wn * 64 + j * 16 + c16;
                double *o = dst + (int64_t)m * ld + n;
                *o = add ? *o + acc[i][j][r] : acc[i][j][r];
            }
}

__global__ __launch_bounds__(256) void k_gram_f64(const double2 *__restrict__ t, int64_t n_groups, int d_pad,
                                                  int n_split, double *partial) {
    const int nb = d_pad / kGramBlock;
    int blk, split, bi, bj;
    gram_work(n_split, nb, blk, split, bi, bj);
    const int64_t per = (n_groups + n_split - 1) / n_split;
    const int64_t g0 = (int64_t)split * per;
    const int64_t g1 = g0 + per < n_groups ? g0 + per : n_groups;
    gram_f64_block(t, d_pad, bi, bj, g0, g1, partial + ((int64_t)blk * n_split + split) * kGramBlock * kGramBlock);
}

__global__ __launch_bounds__(256) void k_gram_f64_items(const double2 *__restrict__ t, int d_pad, const GramItem *items,
                                                        int64_t n_items, double *partial, double *gram) {
    const int nb = d_pad / kGramBlock;
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const GramItem w = items[it];
        if (w.blk < 0) continue;
        int bi, bj, ld;
        bool add;
        tri_block(w.blk, nb, bi, bj);
        double *dst = item_dst(w, bi, bj, d_pad, partial, gram, ld, add);
        gram_f64_block(t, d_pad, bi, bj, w.g0, w.g1, dst, ld, add);
    }
}

void launch_gram(int fmt, const void *tile_a, const void *tile_b, int64_t rows_pad, int d_pad, int n_split,
                 double *partial, hipStream_t s) {
    const int nb = d_pad / kGramBlock;
    const int nblk = nb * (nb + 1) / 2;
    const int64_t n_groups = rows_pad / kRowGroup;
    dim3 grid(nblk * n_split), block(256);
    switch (fmt) {
    case TILE_F64:
        hipLaunchKernelGGL(k_gram_f64, grid, block, 0, s, (const double2 *)tile_a, n_groups, d_pad, n_split, partial);
        break;
    case TILE_F32:
        hipLaunchKernelGGL(k_gram_f32<32>, grid, block, 0, s, (const float4 *)tile_a, n_groups, d_pad, n_split, partial);
        break;
    default: {
        const uint4 *ta = (const uint4 *)tile_a, *tb = (const uint4 *)tile_b;
        const bool dma = gram_dma_enabled();
#define CREST_GRAM_DMA_LAUNCH(T, KS, NB)                                                                               \
    do {                                                                                                               \
        hipFuncSetAttribute((const void *)k_gram_x16_dma<T, 16, KS, NB>, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                            gram_dma_lds_bytes<KS, NB>());                                                             \
        hipLaunchKernelGGL((k_gram_x16_dma<T, 16, KS, NB>), grid, block, (gram_dma_lds_bytes<KS, NB>()), s, ta, tb,   \
                           n_groups, d_pad, n_split, partial);                                                         \
    } while (0)
        if (fmt == TILE_F16X2) {
            if (dma) CREST_GRAM_DMA_LAUNCH(1, kGramKS, kGramNB);
            else hipLaunchKernelGGL((k_gram_x16<1, 16>), grid, block, 0, s, ta, tb, n_groups, d_pad, n_split, partial);
        } else {
            if (dma) CREST_GRAM_DMA_LAUNCH(0, kGramKS, kGramNB);
            else hipLaunchKernelGGL((k_gram_x16<0, 16>), grid, block, 0, s, ta, tb, n_groups, d_pad, n_split, partial);
        }
#undef CREST_GRAM_DMA_LAUNCH
        break;
    }
    }
}

// Form of the split-16-bit Gram: LDS-staged through direct loads unless
// CREST_GRAM_DMA=0, which selects the form that loads operands per wave.
bool gram_dma_enabled() {
    static const bool on = [] {
        const char *e = getenv("CREST_GRAM_DMA");
        return !(e && e[0] == '0' && e[1] == 0);
    }();
    return on;
}

// Deterministic split-K sum into the FP64 Gram, splits in order.
__global__ void k_gram_reduce(const double *partial, int n_split, int nb, int d_pad, double *gram) {
    const int nblk = nb * (nb + 1) / 2;
    const int64_t total = (int64_t)nblk * kGramBlock * kGramBlock;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int blk = (int)(e / (kGramBlock * kGramBlock));
        const int rem = (int)(e % (kGramBlock * kGramBlock));
        int bi, bj;
        tri_block(blk, nb, bi, bj);
        const double *src = partial + (int64_t)blk * n_split * kGramBlock * kGramBlock + rem;
        double s = 0.0;
        for (int sp = 0; sp < n_split; ++sp) s += src[(int64_t)sp * kGramBlock * kGramBlock];
        const int m = rem / kGramBlock, n = rem % kGramBlock;
        gram[(int64_t)(bi * kGramBlock + m) * d_pad + bj * kGramBlock + n] += s;
    }
}

// gram[h, h] += src[h, h] for h < d_pad (both d_pad x d_pad): the densify's diagonal
// terms accumulated apart from the split-K sums.
__global__ void k_add_diag(const double *src, int d_pad, double *gram) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < d_pad) gram[(int64_t)h * d_pad + h] += src[(int64_t)h * d_pad + h];
}
void launch_add_diag(const double *src, int d_pad, double *gram, hipStream_t s) {
    hipLaunchKernelGGL(k_add_diag, dim3((unsigned)ceil_div(d_pad, 256)), dim3(256), 0, s, src, d_pad, gram);
}

// Work-list form: item r sums slots slot0 .. slot0 + n_slots - 1 (in order) of one
// block and adds them into that block of its subset's Gram (gram + subset d_pad^2).
// No two items of one launch name the same (subset, block), so the adds do not race.
__global__ void k_gram_reduce_items(const double *partial, const RedItem *items, int64_t n_items, int nb, int d_pad,
                                    double *gram) {
    const int64_t total = n_items * kGramBlock * kGramBlock;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const RedItem r = items[e / (kGramBlock * kGramBlock)];
        const int rem = (int)(e % (kGramBlock * kGramBlock));
        int bi, bj;
        tri_block(r.blk, nb, bi, bj);
        const double *src = partial + r.slot0 * kGramBlock * kGramBlock + rem;
        double s = 0.0;
        for (int sp = 0; sp < r.n_slots; ++sp) s += src[(int64_t)sp * kGramBlock * kGramBlock];
        const int m = rem / kGramBlock, n = rem % kGramBlock;
        gram[(int64_t)r.subset * d_pad * d_pad + (int64_t)(bi * kGramBlock + m) * d_pad + bj * kGramBlock + n] += s;
    }
}

void launch_gram_items(int fmt, const void *tile_a, const void *tile_b, int d_pad, const GramItem *items,
                       int64_t n_items, int grid, double *partial, double *gram, hipStream_t s) {
    if (n_items <= 0) return;
    const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, n_items))), block(256);
    switch (fmt) {
    case TILE_F64:
        hipLaunchKernelGGL(k_gram_f64_items, g, block, 0, s, (const double2 *)tile_a, d_pad, items, n_items, partial,
                           gram);
        break;
    case TILE_F32:
        hipLaunchKernelGGL(k_gram_f32_items<32>, g, block, 0, s, (const float4 *)tile_a, d_pad, items, n_items,
                           partial, gram);
        break;
    default: {
        const uint4 *ta = (const uint4 *)tile_a, *tb = (const uint4 *)tile_b;
        const int lds = gram_dma_lds_bytes<kGramKS, kGramNB>();
        if (gram_dma_enabled()) {
            if (fmt == TILE_F16X2) {
                hipFuncSetAttribute((const void *)k_gram_x16_items_dma<1, 16, kGramKS, kGramNB>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                hipLaunchKernelGGL((k_gram_x16_items_dma<1, 16, kGramKS, kGramNB>), g, block, lds, s, ta, tb, d_pad,
                                   items, n_items, partial, gram);
            } else {
                hipFuncSetAttribute((const void *)k_gram_x16_items_dma<0, 16, kGramKS, kGramNB>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds);
                hipLaunchKernelGGL((k_gram_x16_items_dma<0, 16, kGramKS, kGramNB>), g, block, lds, s, ta, tb, d_pad,
                                   items, n_items, partial, gram);
            }
        } else if (fmt == TILE_F16X2) {
            hipLaunchKernelGGL((k_gram_x16_items<1, 16>), g, block, 0, s, ta, tb, d_pad, items, n_items, partial, gram);
        } else {
            hipLaunchKernelGGL((k_gram_x16_items<0, 16>), g, block, 0, s, ta, tb, d_pad, items, n_items, partial, gram);
        }
        break;
    }
    }
}

int gram_items_occupancy(int fmt) {
    int n = 0;
    const bool x16 = fmt != TILE_F64 && fmt != TILE_F32, dma = x16 && gram_dma_enabled();
    const void *f = fmt == TILE_F64   ? (const void *)k_gram_f64_items
                    : fmt == TILE_F32 ? (const void *)k_gram_f32_items<32>
                    : dma ? (fmt == TILE_F16X2 ? (const void *)k_gram_x16_items_dma<1, 16, kGramKS, kGramNB>
                                               : (const void *)k_gram_x16_items_dma<0, 16, kGramKS, kGramNB>)
                          : (fmt == TILE_F16X2 ? (const void *)k_gram_x16_items<1, 16>
                                               : (const void *)k_gram_x16_items<0, 16>);
    const size_t lds = dma ? (size_t)gram_dma_lds_bytes<kGramKS, kGramNB>() : 0;
    if (dma) hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 256, lds) != hipSuccess || n < 1) n = 1;
    return n;
}

void launch_gram_reduce_items(const double *partial, const RedItem *items, int64_t n_items, int d_pad, double *gram,
                              hipStream_t s) {
    if (n_items <= 0) return;
    const int64_t total = n_items * kGramBlock * kGramBlock;
    const int g = (int)std::min<int64_t>(ceil_div(total, 256), 16384);
    hipLaunchKernelGGL(k_gram_reduce_items, dim3(g), dim3(256), 0, s, partial, items, n_items, d_pad / kGramBlock,
                       d_pad, gram);
}

void launch_gram_reduce(const double *partial, int n_split, int d_pad, double *gram, hipStream_t s) {
    const int nb = d_pad / kGramBlock;
    const int64_t total = (int64_t)nb * (nb + 1) / 2 * kGramBlock * kGramBlock;
    const int g = (int)std::min<int64_t>(ceil_div(total, 256), 8192);
    hipLaunchKernelGGL(k_gram_reduce, dim3(g), dim3(256), 0, s, partial, n_split, nb, d_pad, gram);
}

// ---------------------------------------------------------------- covariance
// cov[i, j] = (G[a, b] - c_a c_b / n) / (n - 1) with a = min, b = max (kernels.rs:318-322),
// in the reference's operation order; written full and column-major (symmetric).
// blockIdx.y = subset of a batch (gram + y ld_g^2, colsum + y ld_g, cov + y d^2), with
// n = cell_off[y+1] - cell_off[y] when cell_off is given.
__global__ void k_cov(const double *gram, const double *colsum, int d, int ld_g, double n, double *cov,
                      const int64_t *cell_off, const int32_t *ids) {
    const int y = blockIdx.y, sub = ids ? ids[y] : y;  // ids: the subsets of this group
    gram += (int64_t)sub * ld_g * ld_g;
    colsum += (int64_t)sub * ld_g;
    cov += (int64_t)y * d * d;
    if (cell_off) n = (double)(cell_off[sub + 1] - cell_off[sub]);
    const int64_t total = (int64_t)d * d;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / d), i = (int)(e % d);
        const int a = i <= j ? i : j, b = i <= j ? j : i;
        const double cc = colsum[a] * colsum[b];
        const double t = cc / n;
        const double u = gram[(int64_t)a * ld_g + b] - t;
        cov[e] = u / (n - 1.0);
    }
}
// total = Σ_i cov[i, i] in index order (kernels.rs:323): the strided diagonal is
// gathered into LDS by the whole block, then summed sequentially by one thread.
__global__ __launch_bounds__(256) void k_trace(const double *cov, int d, double *total, const int32_t *ids) {
    __shared__ double diag[4096];
    cov += (int64_t)blockIdx.x * d * d;  // one block per subset of a batch
    total += ids ? ids[blockIdx.x] : blockIdx.x;
    double s = 0.0;
    for (int i0 = 0; i0 < d; i0 += 4096) {
        const int m = d - i0 < 4096 ? d - i0 : 4096;
        for (int i = threadIdx.x; i < m; i += blockDim.x) diag[i] = cov[(int64_t)(i0 + i) * d + i0 + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; ++i) s += diag[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s;
}

void launch_cov(const double *gram, const double *colsum, int d, int ld_g, double n, double *cov, double *total,
                hipStream_t s) {
    const int g = (int)std::min<int64_t>(ceil_div((int64_t)d * d, 256), 8192);
    hipLaunchKernelGGL(k_cov, dim3(g), dim3(256), 0, s, gram, colsum, d, ld_g, n, cov, (const int64_t *)nullptr,
                       (const int32_t *)nullptr);
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, s, cov, d, total, (const int32_t *)nullptr);
}

void launch_cov_batch(const double *gram, const double *colsum, int d, int ld_g, const int64_t *dev_cell_off,
                      int n_sub, double *cov, double *total, hipStream_t s, const int32_t *ids) {
    const int g = (int)std::min<int64_t>(ceil_div((int64_t)d * d, 256), 2048);
    hipLaunchKernelGGL(k_cov, dim3(g, n_sub), dim3(256), 0, s, gram, colsum, d, ld_g, 0.0, cov, dev_cell_off, ids);
    hipLaunchKernelGGL(k_trace, dim3(n_sub), dim3(256), 0, s, cov, d, total, ids);
}

// total = trace of the covariance, straight from the Gram diagonal and the column sums
// with k_cov's operations and k_trace's order (for subsets whose covariance is never
// formed): one block per subset ids[b].
__global__ __launch_bounds__(256) void k_trace_gram(const double *gram, const double *colsum, int d, int ld_g,
                                                    const int64_t *cell_off, const int32_t *ids, double *total) {
    __shared__ double diag[4096];
    const int sub = ids[blockIdx.x];
    gram += (int64_t)sub * ld_g * ld_g;
    colsum += (int64_t)sub * ld_g;
    const double n = (double)(cell_off[sub + 1] - cell_off[sub]);
    double s = 0.0;
    for (int i0 = 0; i0 < d; i0 += 4096) {
        const int m = d - i0 < 4096 ? d - i0 : 4096;
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            const int a = i0 + i;
            const double cc = colsum[a] * colsum[a];
            const double t = cc / n;
            const double u = gram[(int64_t)a * ld_g + a] - t;
            diag[i] = u / (n - 1.0);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; ++i) s += diag[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[sub] = s;
}
void launch_trace_gram(const double *gram, const double *colsum, int d, int ld_g, const int64_t *dev_cell_off,
                       const int32_t *ids, int n_ids, double *total, hipStream_t s) {
    if (n_ids <= 0) return;
    hipLaunchKernelGGL(k_trace_gram, dim3(n_ids), dim3(256), 0, s, gram, colsum, d, ld_g, dev_cell_off, ids, total);
}

// Eigenpairs of C = Zc^T Zc / (n - 1) from those of K = Zc Zc^T: column j of V (d x p,
// = Zc^T U) normalised, eigenvalue W[j] / (n - 1).  One block per (column, subset
// ids[y]); zero columns stay zero.
__global__ __launch_bounds__(256) void k_dual_finish(double *V, const double *W, int d, int p,
                                                     const int64_t *cell_off, const int32_t *ids, double *evals) {
    __shared__ double red[256];
    const int j = blockIdx.x, y = blockIdx.y, sub = ids[y];
    double *v = V + ((int64_t)y * p + j) * d;
    double ss = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) ss += v[i] * v[i];
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double nrm = sqrt(red[0]);
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) v[i] *= inv;
    if (threadIdx.x == 0) {
        const double n = (double)(cell_off[sub + 1] - cell_off[sub]);
        evals[(int64_t)y * p + j] = W[(int64_t)y * p + j] / (n - 1.0);
    }
}
void launch_dual_finish(double *V, const double *W, int d, int p, const int64_t *dev_cell_off, const int32_t *ids,
                        int n_ids, double *evals, hipStream_t s) {
    if (n_ids <= 0 || p <= 0) return;
    hipLaunchKernelGGL(k_dual_finish, dim3(p, n_ids), dim3(256), 0, s, V, W, d, p, dev_cell_off, ids, evals);
}

// ---------------------------------------------------------------- eigen selection
// Component c takes eigenpair ncols-1-c of the ascending solver output (ncols = d for
// a full eigensolve, the number of computed eigenpairs otherwise); variance is
// floored at 0; the sign makes the first largest-|loading| entry positive
// (kernels.rs:328-346).  evecs is column-major (column j = eigenvector j).
__global__ __launch_bounds__(256) void k_eig_select(const double *evecs, const double *evals, int d, int k,
                                                    double *variance, double *loadings, int ncols, const int32_t *ids) {
    __shared__ double sbest[256];
    __shared__ int sidx[256];
    const int y = blockIdx.y, sub = ids ? ids[y] : y;  // subset of a batch
    evecs += (int64_t)y * d * ncols;
    evals += (int64_t)y * ncols;
    variance += (int64_t)sub * k;
    loadings += (int64_t)sub * d * k;
    const int c = blockIdx.x;
    const int src = ncols - 1 - c;
    const double *u = evecs + (int64_t)src * d;
    double best = 0.0;
    int bidx = -1;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        const double a = fabs(u[i]);
        if (a > best) {
            best = a;
            bidx = i;
        }
    }
    sbest[threadIdx.x] = best;
    sidx[threadIdx.x] = bidx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double b2 = sbest[threadIdx.x + o];
            const int i2 = sidx[threadIdx.x + o];
            const double b1 = sbest[threadIdx.x];
            const int i1 = sidx[threadIdx.x];
            if (i2 >= 0 && (i1 < 0 || b2 > b1 || (b2 == b1 && i2 < i1))) {
                sbest[threadIdx.x] = b2;
                sidx[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    double sign = 1.0;
    if (sidx[0] >= 0) sign = signbit(u[sidx[0]]) ? -1.0 : 1.0;
    if (threadIdx.x == 0) {
        const double v = evals[src];
        variance[c] = v > 0.0 ? v : 0.0;
    }
    for (int i = threadIdx.x; i < d; i += blockDim.x) loadings[(int64_t)i * k + c] = u[i] * sign;
}

void launch_eig_select(const double *evecs, const double *evals, int d, int k, double *variance, double *loadings,
                       hipStream_t s, int ncols) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_eig_select, dim3(k), dim3(256), 0, s, evecs, evals, d, k, variance, loadings,
                       ncols > 0 ? ncols : d, (const int32_t *)nullptr);
}
void launch_eig_select_batch(const double *evecs, const double *evals, int d, int k, int n_sub, double *variance,
                             double *loadings, hipStream_t s, int ncols, const int32_t *ids) {
    if (k <= 0 || n_sub <= 0) return;
    hipLaunchKernelGGL(k_eig_select, dim3(k, n_sub), dim3(256), 0, s, evecs, evals, d, k, variance, loadings,
                       ncols > 0 ? ncols : d, ids);
}

// ---------------------------------------------------------------- shift prefill
// shift = (colsum / n) @ V (tl.py:76), one block per component (a matrix-vector
// product, so the summation order is the block reduction's), then
// X[:] = -(f32)shift (tl.py:77-78).
__global__ __launch_bounds__(256) void k_shift(const double *colsum, const double *V, int d, int k, double n,
                                               float *shift_f, int ld_c, const int64_t *cell_off) {
    __shared__ double part[256];
    const int y = blockIdx.y;  // subset of a batch
    colsum += (int64_t)y * ld_c;
    V += (int64_t)y * d * k;
    shift_f += (int64_t)y * k;
    if (cell_off) n = (double)(cell_off[y + 1] - cell_off[y]);
    const int c = blockIdx.x;
    double s = 0.0;
    for (int a = threadIdx.x; a < d; a += blockDim.x) {
        const double m = colsum[a] / n;
        const double p = m * V[(int64_t)a * k + c];
        s = s + p;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) shift_f[c] = -(float)part[0];
}
__global__ void k_fill_rows(const float *shift_f, int k, int64_t n_rows, float *x, const int64_t *cell_off) {
    if (cell_off) {  // blockIdx.y = subset: its rows cell_off[y] .. cell_off[y+1]-1
        const int y = blockIdx.y;
        shift_f += (int64_t)y * k;
        x += cell_off[y] * k;
        n_rows = cell_off[y + 1] - cell_off[y];
    }
    const int64_t total = n_rows * k;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = shift_f[e % k];
}

void launch_shift_fill(const double *colsum, const double *V, int d, int k, double n, int64_t n_rows, float *x,
                       float *shift_f, hipStream_t s) {
    hipLaunchKernelGGL(k_shift, dim3(k), dim3(256), 0, s, colsum, V, d, k, n, shift_f, 0, (const int64_t *)nullptr);
    const int g = (int)std::min<int64_t>(ceil_div(n_rows * k, 256), 8192);
    hipLaunchKernelGGL(k_fill_rows, dim3(g), dim3(256), 0, s, shift_f, k, n_rows, x, (const int64_t *)nullptr);
}

void launch_fill_rows(const float *shift_f, int k, int64_t n_rows, float *x, hipStream_t s) {
    const int g = (int)std::min<int64_t>(ceil_div(n_rows * k, 256), 8192);
    if (n_rows * k > 0)
        hipLaunchKernelGGL(k_fill_rows, dim3(g), dim3(256), 0, s, shift_f, k, n_rows, x, (const int64_t *)nullptr);
}

void launch_shift_fill_batch(const double *colsum, int ld_c, const double *V, int d, int k,
                             const int64_t *dev_cell_off, int n_sub, int64_t max_rows, float *x, float *shift_f,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_shift, dim3(k, n_sub), dim3(256), 0, s, colsum, V, d, k, 0.0, shift_f, ld_c, dev_cell_off);
    const int g = (int)std::min<int64_t>(ceil_div(max_rows * k, 256), 1024);
    hipLaunchKernelGGL(k_fill_rows, dim3(g, n_sub), dim3(256), 0, s, shift_f, k, (int64_t)0, x, dev_cell_off);
}

}  // namespace crest
