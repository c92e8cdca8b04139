// The FP64 Gram of the scaled entries by sparse pair products (kernels.rs:221-309).
//
// The scaled matrix S is sparse: its entries are the non-zero counts of HVGs, each
// s = clip((x - mu)/sd) - clip(-mu/sd) (kernels.rs:211-218), and zeros stay zero.
// The reference accumulates the upper triangle G[a][b] += s_a s_b row by row over
// each row's (HVG, s) list; so does this path, exactly in FP64 (one rounding per
// product, FP64 sums), instead of a dense MFMA product over a tile that is 92% zeros.
//
// k_sparse_pack (one workgroup per group of 64 rows): the row's HVG entries, scaled,
// written grouped by 128-column bucket (bucket = hvg >> 7): within a group of 64 rows
// all entries of bucket 0 (row by row, each row's entries in entry order), then
// bucket 1, ...  The rows take their places in the group by descending HVG entry
// count (the Gram is a sum over rows, so their order is free).  The group's region starts at the group's own CSR offset (it holds at
// most that many entries).  Per group a record of bucket starts (nb + 1 words) and
// per-bucket row offsets (65 u16 per bucket, by place).  Entry = FP64 value + the
// column within its bucket (one byte).
//
// k_gram_sparse (one 1024-thread workgroup per (128 x 128 Gram block, row split)): the
// block lives in LDS as FP64 (row stride 129, so the rows of one add fall in different
// banks).  For each group of its split, each wave takes 4 rows, one per 16-lane row of
// the wave: lane t loads a-entry t and b-entry t of its row straight into registers
// (the next group's rows loading while this group's pairs run), and step i broadcasts
// a-entry i (its value and its block-row offset) across the row's lanes (DPP
// row_newbcast), every lane adding s_a s_b into the block with an LDS FP64 atomic add.
// A diagonal block adds both orders of every pair and drops its lower triangle when
// it is written out, so no lane compares columns; its entries' s go into the colsum
// once each.  The adds are the only LDS traffic and no result feeds another, so they
// issue back to back.  The block (upper triangle, as the reference accumulates it) and
// colsum go to a per-split partial slot, reduced over splits in order afterwards.  The order of
// the adds within one workgroup is the hardware's, so results agree with any other
// order to FP64 rounding (not bitwise between runs).
#include "common.hpp"

namespace crest {

constexpr int kSpGroup = kRowGroupSparse;  // rows per group
constexpr int kSpMaxNb = 32;      // buckets (d_pad <= 4096)
constexpr int kSpThreads = 1024;  // block kernel (16 waves: 4 rows each of a 64-row group)
constexpr int kSpWaves = kSpThreads / kWave;
constexpr int kSpLd = 129;        // LDS row stride of the Gram block (doubles)

__device__ inline double sp_cell_total(const DevChunk &c, int64_t lo, int64_t hi, int lane) {
    double t = 0.0;
    for (int64_t e = lo + lane; e < hi; e += kWave)
        if (chunk_var(c, c.genes[e]) >= 0) t += (double)c.values[e];
    return wave_sum_f64(t);
}

// words of one group record: bucket starts [nb + 1] (u32), then row offsets
// [nb][65] (u16, two per word), padded to a multiple of 16 words
__host__ __device__ inline int sp_rec_words(int nb) { return ((nb + 1) + (nb * 65 + 1) / 2 + 15) / 16 * 16; }

__global__ __launch_bounds__(256) void k_sparse_pack(DevChunk c, DevScale sc0, int nb, int64_t n_groups,
                                                     const SegDesc *grp, double *vals, uint8_t *cols, uint32_t *meta) {
    __shared__ int cnt[kSpGroup][kSpMaxNb + 1];  // padded: a wave reads one bucket of 64 rows
    __shared__ int bstart[kSpMaxNb + 1];
    __shared__ int run[4][kSpMaxNb];
    __shared__ int tot[kSpGroup], inv[kSpGroup], btot[kSpMaxNb];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int rw = sp_rec_words(nb);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int64_t g = blockIdx.x; g < n_groups; g += gridDim.x) {
        int64_t r0 = g * kSpGroup;
        int nr = (int)((c.n_spans - r0) < kSpGroup ? (c.n_spans - r0) : kSpGroup);
        DevScale sc = sc0;
        if (grp) {  // a batch: the group's own rows and its subset's scale parameters
            const SegDesc gd = grp[g];
            r0 = gd.cell0;
            nr = (int)gd.n;
            sc.mean += (int64_t)gd.subset * sc.d;
            sc.sd += (int64_t)gd.subset * sc.d;
        }
        const int64_t gbase = c.span_off[r0] - c.base;
        for (int i = threadIdx.x; i < kSpGroup * kSpMaxNb; i += blockDim.x) cnt[i / kSpMaxNb][i % kSpMaxNb] = 0;
        if (threadIdx.x < kSpGroup) tot[threadIdx.x] = 0;
        __syncthreads();
        // entries per (row, bucket), and per row
        for (int rl = wave; rl < nr; rl += 4) {
            if (chunk_row(c, r0 + rl) < 0) continue;
            const int64_t lo = c.span_off[r0 + rl] - c.base, hi = c.span_off[r0 + rl + 1] - c.base;
            int n_row = 0;
            for (int64_t e0 = lo; e0 < hi; e0 += kWave) {
                const int64_t e = e0 + lane;
                bool in = false;
                if (e < hi) {
                    const int32_t m = chunk_var(c, c.genes[e]);
                    const int32_t h = m >= 0 ? sc.sub[m] : -1;
                    in = h >= 0 && h < sc.d;
                    if (in) atomicAdd(&cnt[rl][h >> 7], 1);
                }
                n_row += __popcll(__ballot(in));
            }
            if (lane == 0) tot[rl] = n_row;
        }
        __syncthreads();
        // the group's rows take their places by descending entry count (ties in row
        // order), so the four rows a wave of the block kernel takes have similar lists:
        // wave 0, lane r = row r, ranks its row against every other by register reads
        if (wave == 0) {
            const int t = tot[lane];
            int k = 0;
#pragma unroll
            for (int q = 0; q < kSpGroup; ++q) {
                const int tq = __builtin_amdgcn_readlane(t, q);
                k += (tq > t || (tq == t && q < lane)) ? 1 : 0;
            }
            inv[k] = lane;
        }
        __syncthreads();
        // per bucket (one wave each), lane k = the row at place k: the exclusive prefix of
        // the bucket's counts over the places, the row's offset within the bucket
        uint32_t *rec = meta + g * rw;
        uint16_t *roff = reinterpret_cast<uint16_t *>(rec + nb + 1);
        for (int b = wave; b < nb; b += 4) {
            const int rl = inv[lane], t = cnt[rl][b];
            int x = t;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(x, o, kWave);
                if (lane >= o) x += y;
            }
            cnt[rl][b] = x - t;
            roff[b * 65 + lane] = (uint16_t)(x - t);
            if (lane == kWave - 1) {
                roff[b * 65 + kSpGroup] = (uint16_t)x;
                btot[b] = x;
            }
        }
        __syncthreads();
        // bucket starts: the exclusive prefix of the bucket totals
        if (wave == 0) {
            const int t = lane < nb ? btot[lane] : 0;
            int x = t;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(x, o, kWave);
                if (lane >= o) x += y;
            }
            if (lane <= nb) {
                bstart[lane] = x - t;
                rec[lane] = (uint32_t)(x - t);
            }
        }
        __syncthreads();
        // scaled entries at their positions (entry order within each (row, bucket)); the
        // waves take the rows in place order, so their stores fill neighbouring lines
        for (int k = wave; k < kSpGroup; k += 4) {
            const int rl = inv[k];
            if (rl >= nr || chunk_row(c, r0 + rl) < 0) continue;
            const int64_t lo = c.span_off[r0 + rl] - c.base, hi = c.span_off[r0 + rl + 1] - c.base;
            const double ns = norm_factor(sp_cell_total(c, lo, hi, lane), c.target_sum);
            for (int b = lane; b < nb; b += kWave) run[wave][b] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int64_t e0 = lo; e0 < hi; e0 += kWave) {
                const int64_t e = e0 + lane;
                int h = -1;
                double s = 0.0;
                if (e < hi) {
                    const int32_t m = chunk_var(c, c.genes[e]);
                    if (m >= 0) {
                        h = sc.sub[m];
                        if (h >= sc.d) h = -1;
                        if (h >= 0) {
                            const float v = lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
                            s = sc.scale ? scale_value(v, sc.mean[h], sc.sd[h], sc.max_value, 1) : (double)v;
                        }
                    }
                }
                const int b = h >= 0 ? (h >> 7) : -1;
                int pos = -1;
                for (int bb = 0; bb < nb; ++bb) {
                    const uint64_t msk = __ballot(b == bb);
                    if (msk == 0) continue;
                    const int base_bb = run[wave][bb];
                    if (b == bb) pos = base_bb + __popcll(msk & lt);
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) run[wave][bb] = base_bb + __popcll(msk);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
                if (b >= 0) {
                    const int64_t at = gbase + bstart[b] + cnt[rl][b] + pos;
                    vals[at] = s;
                    cols[at] = (uint8_t)(h & 127);
                }
            }
        }
        __syncthreads();
    }
}

// lane N of each row of 16 lanes, to every lane of that row (DPP row_newbcast; every
// lane is written, so no previous value is needed; a double moves in one v_mov_b64_dpp)
template <int N, class T> __device__ inline T row_bcast(T v) { return __builtin_amdgcn_mov_dpp(v, 0x150 + N, 0xf, 0xf, false); }

// Pairs of one 16 x 16 chunk: lane t of row group r holds broadcast entry t (value sa,
// byte offset ao of its part of the block address) and lane entry t (value sb; gb =
// the block element its own part addresses) of its row; step i adds broadcast entry i
// (over the row's 16 lanes) times every lane entry of the row.  na: the row's broadcast
// entries in this chunk.  In a diagonal block both orders of every pair are added (the
// block's lower triangle is dropped when the block is written out), so no lane tests
// its columns.
template <int I>
__device__ inline void sp_step(double *gb, double sa, int ao, double sb, bool ok_b, int na) {
    const double ai = row_bcast<I>(sa);
    const int ai_off = row_bcast<I>(ao);
    if (ok_b && I < na) atomicAdd(reinterpret_cast<double *>(reinterpret_cast<char *>(gb) + ai_off), ai * sb);
}
__device__ inline void sp_chunk(double *gb, double sa, int ao, double sb, bool ok_b, int na, int n_i) {
    sp_step<0>(gb, sa, ao, sb, ok_b, na);
#define CREST_SP_STEP(I) \
    if (n_i > I) sp_step<I>(gb, sa, ao, sb, ok_b, na);
    CREST_SP_STEP(1) CREST_SP_STEP(2) CREST_SP_STEP(3) CREST_SP_STEP(4) CREST_SP_STEP(5) CREST_SP_STEP(6)
    CREST_SP_STEP(7) CREST_SP_STEP(8) CREST_SP_STEP(9) CREST_SP_STEP(10) CREST_SP_STEP(11) CREST_SP_STEP(12)
    CREST_SP_STEP(13) CREST_SP_STEP(14) CREST_SP_STEP(15)
#undef CREST_SP_STEP
}

// wave-uniform maximum of a per-row value held by every lane of each 16-lane row
__device__ inline int rows_max(int v) {
    const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return max(max(a, b), max(c, d));
}

// One wave's view of its 4 rows of a group: row r = lane / 16 of the wave.  Of the
// row's a-list (entries of bucket bi: the block's rows) and b-list (bucket bj: its
// columns) the shorter one is broadcast entry by entry and the longer one spread over
// the row's 16 lanes (a diagonal block: the a-list both ways), so a row takes
// min(na, nb) steps with max(na, nb) lanes busy; each row chooses for itself, as the
// broadcast stays within its 16 lanes.  Per lane: the broadcast list's start, length
// and first 16 entries, the lane list's the same, and whether the roles are swapped
// (the broadcast list is the b-list).
struct SpRows {
    int64_t k_at, l_at;  // first broadcast / lane entry of this lane's row
    int nk, nl;
    double vk, vl;
    int ck, cl;
    bool swap;
};

template <bool DIAG>
__device__ inline void sp_load_rows(SpRows &R, const double *vals, const uint8_t *cols, const uint32_t *meta,
                                    const int64_t *span_off, int64_t base, const SegDesc *grp, int rw, int nb, int bi,
                                    int bj, int64_t g, int row, int t) {
    const uint32_t *rec = meta + g * rw;
    const uint16_t *roff = reinterpret_cast<const uint16_t *>(rec + nb + 1);
    const int64_t gbase = span_off[grp ? grp[g].cell0 : g * kSpGroup] - base;
    const int a0 = roff[bi * 65 + row], a1 = roff[bi * 65 + row + 1];
    const int64_t a_at = gbase + rec[bi] + a0;
    const int na = a1 - a0;
    int64_t b_at = a_at;
    int nbr = na;
    if (!DIAG) {
        const int b0 = roff[bj * 65 + row], b1 = roff[bj * 65 + row + 1];
        b_at = gbase + rec[bj] + b0;
        nbr = b1 - b0;
    }
    R.swap = !DIAG && na > nbr;
    R.k_at = R.swap ? b_at : a_at;
    R.nk = R.swap ? nbr : na;
    R.l_at = R.swap ? a_at : b_at;
    R.nl = R.swap ? na : nbr;
    R.vk = t < R.nk ? vals[R.k_at + t] : 0.0;
    R.ck = t < R.nk ? (int)cols[R.k_at + t] : 0;
    R.vl = t < R.nl ? vals[R.l_at + t] : 0.0;
    R.cl = t < R.nl ? (int)cols[R.l_at + t] : 0;
}

template <bool DIAG>
__device__ inline void sp_pairs(double *G, double *cs, const SpRows &R, const double *vals, const uint8_t *cols,
                                int t) {
    const int mx_k = rows_max(R.nk), mx_l = rows_max(R.nl);
    // the product of a-column ca and b-column cb goes to element ca kSpLd + cb: the
    // broadcast entry gives its byte offset, the lane entry its element pointer
    const int k_scale = R.swap ? (int)sizeof(double) : kSpLd * (int)sizeof(double);
    const int l_scale = R.swap ? kSpLd : 1;
    for (int kc = 0; kc < mx_k; kc += 16) {
        double vk = R.vk;
        int ck = R.ck;
        if (kc > 0) {  // further broadcast entries of rows with more than 16 in this bucket
            const bool h = kc + t < R.nk;
            vk = h ? vals[R.k_at + kc + t] : 0.0;
            ck = h ? (int)cols[R.k_at + kc + t] : 0;
        }
        const int ko = ck * k_scale;
        const int nk_c = R.nk - kc, n_i = min(mx_k - kc, 16);
        for (int lc = 0; lc < mx_l; lc += 16) {
            double vl = R.vl;
            int cl = R.cl;
            if (lc > 0) {
                const bool h = lc + t < R.nl;
                vl = h ? vals[R.l_at + lc + t] : 0.0;
                cl = h ? (int)cols[R.l_at + lc + t] : 0;
            }
            const bool ok_l = lc + t < R.nl;
            // a diagonal block's entries are its column sums' terms, each once
            if (DIAG && kc == 0 && ok_l) atomicAdd(&cs[cl], vl);
            sp_chunk(G + cl * l_scale, vk, ko, vl, ok_l, nk_c, n_i);
        }
    }
}

template <bool DIAG>
__device__ __attribute__((always_inline)) inline void sp_block(double *G, double *cs, const double *vals, const uint8_t *cols, const uint32_t *meta,
                         const int64_t *span_off, int64_t base, const SegDesc *grp, int rw, int nb, int bi, int bj,
                         int64_t g0, int64_t g1, int wave, int lane) {
    const int t = lane & 15;
    // a group's rows sit by descending entry count, so wave w takes places
    // 4((w + g) mod 16) .. +3 of group g: rows of similar lists together, and every
    // wave meets every size class in turn
    auto row_of = [&](int64_t g) { return 4 * (int)((wave + g) % kSpWaves) + (lane >> 4); };
    if (g0 >= g1) return;
    SpRows cur, nxt;
    sp_load_rows<DIAG>(cur, vals, cols, meta, span_off, base, grp, rw, nb, bi, bj, g0, row_of(g0), t);
    for (int64_t g = g0; g < g1; ++g) {
        if (g + 1 < g1)
            sp_load_rows<DIAG>(nxt, vals, cols, meta, span_off, base, grp, rw, nb, bi, bj, g + 1, row_of(g + 1), t);
        sp_pairs<DIAG>(G, cs, cur, vals, cols, t);
        cur = nxt;
    }
}

__global__ __launch_bounds__(kSpThreads) void k_gram_sparse(const double *__restrict__ vals,
                                                          const uint8_t *__restrict__ cols,
                                                          const uint32_t *__restrict__ meta,
                                                          const int64_t *__restrict__ span_off, int64_t base,
                                                          int64_t n_groups, int nb, int n_split, int split0,
                                                          int n_sl, double *partial, double *cs_partial) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double *G = reinterpret_cast<double *>(lds_raw);  // [128][kSpLd]
    double *cs = G + 128 * kSpLd;                       // [128]
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    // this launch's splits split0 .. split0 + n_sl - 1 of n_split over the groups
    const int split = split0 + (int)(blockIdx.x % n_sl), blk = (int)(blockIdx.x / n_sl);
    int bi = 0, bj = 0;
    {
        int r = 0, t = blk;
        while (t >= nb - r) {
            t -= nb - r;
            ++r;
        }
        bi = r;
        bj = r + t;
    }
    const int rw = sp_rec_words(nb);
    for (int i = tid; i < 128 * kSpLd; i += kSpThreads) G[i] = 0.0;
    for (int i = tid; i < 128; i += kSpThreads) cs[i] = 0.0;
    __syncthreads();
    const int64_t g0 = split * n_groups / n_split, g1 = (split + 1) * n_groups / n_split;
    if (bi == bj)
        sp_block<true>(G, cs, vals, cols, meta, span_off, base, nullptr, rw, nb, bi, bj, g0, g1, wave, lane);
    else
        sp_block<false>(G, cs, vals, cols, meta, span_off, base, nullptr, rw, nb, bi, bj, g0, g1, wave, lane);
    __syncthreads();
    double *dst = partial + ((int64_t)blk * n_split + split) * kGramBlock * kGramBlock;
    for (int i = tid; i < kGramBlock * kGramBlock; i += kSpThreads) {
        const int m = i / kGramBlock, n = i % kGramBlock;
        dst[i] = (bi == bj && m > n) ? 0.0 : G[m * kSpLd + n];  // upper triangle only, as the reference's
    }
    if (bi == bj)
        for (int i = tid; i < 128; i += kSpThreads) cs_partial[((int64_t)bi * n_split + split) * 128 + i] = cs[i];
}

// Batch form: workgroup b takes item b -- one 128 x 128 block of one subset's Gram
// over that subset's groups [g0, g1) of the group table (64-row groups that never
// straddle two subsets).  An item with slot < 0 is the only one for its (subset,
// block) and adds the block into the subset's Gram; the others write partial slot
// `slot`, which a reduce item sums into the Gram in range order afterwards.  The
// diagonal blocks' entry sums go into the subset's colsum with FP64 atomics (one
// per column and item).
__global__ __launch_bounds__(kSpThreads) void k_gram_sparse_items(const double *__restrict__ vals,
                                                                const uint8_t *__restrict__ cols,
                                                                const uint32_t *__restrict__ meta,
                                                                const int64_t *__restrict__ span_off, int64_t base,
                                                                const SegDesc *__restrict__ grp, const GramItem *items,
                                                                int nb, int d_pad, double *partial, double *gram,
                                                                double *colsum) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double *G = reinterpret_cast<double *>(lds_raw);  // [128][kSpLd]
    double *cs = G + 128 * kSpLd;                       // [128]
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const GramItem it = items[blockIdx.x];
    if (it.blk < 0) return;
    int bi = 0, bj = 0;
    {
        int r = 0, t = it.blk;
        while (t >= nb - r) {
            t -= nb - r;
            ++r;
        }
        bi = r;
        bj = r + t;
    }
    const int rw = sp_rec_words(nb);
    for (int i = tid; i < 128 * kSpLd; i += kSpThreads) G[i] = 0.0;
    for (int i = tid; i < 128; i += kSpThreads) cs[i] = 0.0;
    __syncthreads();
    if (bi == bj)
        sp_block<true>(G, cs, vals, cols, meta, span_off, base, grp, rw, nb, bi, bj, it.g0, it.g1, wave, lane);
    else
        sp_block<false>(G, cs, vals, cols, meta, span_off, base, grp, rw, nb, bi, bj, it.g0, it.g1, wave, lane);
    __syncthreads();
    if (it.slot < 0) {
        double *dst = gram + (int64_t)it.subset * d_pad * d_pad + (int64_t)bi * kGramBlock * d_pad + bj * kGramBlock;
        for (int i = tid; i < kGramBlock * kGramBlock; i += kSpThreads) {
            const int m = i / kGramBlock, n = i % kGramBlock;
            if (bi != bj || m <= n) dst[(int64_t)m * d_pad + n] += G[m * kSpLd + n];
        }
    } else {
        double *dst = partial + it.slot * kGramBlock * kGramBlock;
        for (int i = tid; i < kGramBlock * kGramBlock; i += kSpThreads) {
            const int m = i / kGramBlock, n = i % kGramBlock;
            dst[i] = (bi == bj && m > n) ? 0.0 : G[m * kSpLd + n];
        }
    }
    if (bi == bj)
        for (int i = tid; i < 128; i += kSpThreads)
            atomicAdd(&colsum[(int64_t)it.subset * d_pad + bi * kGramBlock + i], cs[i]);
}

// colsum[h] += sum over splits (in order) of the diagonal blocks' entry sums
__global__ void k_sparse_colsum(const double *cs_partial, int nb, int n_split, double *colsum) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nb * 128) return;
    const int b = h / 128, i = h % 128;
    double s = 0.0;
    for (int sp = 0; sp < n_split; ++sp) s += cs_partial[((int64_t)b * n_split + sp) * 128 + i];
    colsum[h] += s;
}

bool sparse_gram_supported(int d_pad) { return d_pad > 0 && d_pad <= kSpMaxNb * kGramBlock; }
size_t sparse_meta_words(int64_t n_rows, int d_pad) {
    return (size_t)ceil_div(n_rows, kSpGroup) * sp_rec_words(d_pad / kGramBlock);
}
size_t sparse_gram_lds() { return (size_t)128 * kSpLd * 8 + 128 * 8; }

void launch_sparse_pack(const DevChunk &c, const DevScale &sc, int d_pad, double *vals, uint8_t *cols, uint32_t *meta,
                        hipStream_t s) {
    const int64_t n_groups = ceil_div(c.n_spans, kSpGroup);
    if (n_groups <= 0) return;
    const int g = (int)std::min<int64_t>(n_groups, 8192);
    const SegDesc *no_groups = nullptr;
    hipLaunchKernelGGL(k_sparse_pack, dim3(g), dim3(256), 0, s, c, sc, d_pad / kGramBlock, n_groups, no_groups, vals,
                       cols, meta);
}

void launch_sparse_pack_groups(const DevChunk &c, const DevScale &sc, int d_pad, const SegDesc *groups,
                               int64_t n_groups, double *vals, uint8_t *cols, uint32_t *meta, hipStream_t s) {
    if (n_groups <= 0) return;
    const int g = (int)std::min<int64_t>(n_groups, 8192);
    hipLaunchKernelGGL(k_sparse_pack, dim3(g), dim3(256), 0, s, c, sc, d_pad / kGramBlock, n_groups, groups, vals, cols,
                       meta);
}

void sparse_set_attrs() {
    const int lds = (int)sparse_gram_lds();
    hipFuncSetAttribute((const void *)k_gram_sparse, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    hipFuncSetAttribute((const void *)k_gram_sparse_items, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

void launch_gram_sparse_items(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                              int64_t base, const SegDesc *groups, const GramItem *items, int64_t n_items, int d_pad,
                              double *partial, double *gram, double *colsum, hipStream_t s) {
    if (n_items <= 0) return;
    hipLaunchKernelGGL(k_gram_sparse_items, dim3((unsigned)n_items), dim3(kSpThreads), sparse_gram_lds(), s, vals, cols,
                       meta, span_off, base, groups, items, d_pad / kGramBlock, d_pad, partial, gram, colsum);
}

void launch_gram_sparse_part(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                             int64_t base, int64_t n_rows, int d_pad, int n_split, int split0, int n_sl, double *partial,
                             double *cs_partial, hipStream_t s) {
    const int64_t n_groups = ceil_div(n_rows, kSpGroup);
    const int nb = d_pad / kGramBlock, nblk = nb * (nb + 1) / 2;
    if (n_groups <= 0 || n_sl <= 0) return;
    const size_t lds = sparse_gram_lds();
    hipFuncSetAttribute((const void *)k_gram_sparse, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_gram_sparse, dim3(nblk * n_sl), dim3(kSpThreads), lds, s, vals, cols, meta, span_off, base,
                       n_groups, nb, n_split, split0, n_sl, partial, cs_partial);
}
void launch_sparse_colsum(const double *cs_partial, int d_pad, int n_split, double *colsum, hipStream_t s) {
    hipLaunchKernelGGL(k_sparse_colsum, dim3((unsigned)ceil_div(d_pad, 256)), dim3(256), 0, s, cs_partial,
                       d_pad / kGramBlock, n_split, colsum);
}
void launch_gram_sparse(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                        int64_t base, int64_t n_rows, int d_pad, int n_split, double *partial, double *cs_partial,
                        double *colsum, hipStream_t s) {
    if (n_rows <= 0) return;
    launch_gram_sparse_part(vals, cols, meta, span_off, base, n_rows, d_pad, n_split, 0, n_split, partial, cs_partial,
                            s);
    launch_sparse_colsum(cs_partial, d_pad, n_split, colsum, s);
}

}  // namespace crest
