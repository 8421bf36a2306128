// The reference's comparison quicksort, restated data-parallel on the GPU.
//
// The reference orders B/J/H/F/E/M/D/Z/U/W/T columns with an unstable first-element-pivot
// Hoare quicksort (lib_partition / lib_quicksort_other, backend/h_libs.cc:409-442), reached from
// lib_order_by_vector (:730-736) for one column and, run by run, from lib_quicksort_list
// (:637-697) for lists; pfnGroupBasic orders such keys with it before cutting runs
// (h_primitive.cc:1013-1029).  Rows whose keys tie come out in whatever order that exact
// sequence of partitions leaves them, and group keys (the first row of each run) and group row
// lists inherit it, so only the same partitions reproduce them.
//
// One partition of a segment [lo, hi) with pivot p = key[lo] and the reference's comparator
// (lib_quicksort_cmp_item, h_libs.cc:451-482): a position is G when cmp(key, p) > 0, else LE.
// lib_partition's two scans only ever meet positions no swap has touched, so its k-th swap
// exchanges the k-th G from the left with the k-th LE from the right while the first lies left
// of the second, and the loop ends with
//     j = lo + (number of LE in [lo, hi), the pivot included) - 1,
//     [lo, j] all LE, (j, hi) all G,
// after which key[lo] and key[j] swap -- unless every key equals the pivot (leftSame), which
// finishes the segment.  The recursion goes on with [lo, j) and [j + 1, hi).  So one partition
// is: count G, place j, pair the G's at or left of j with the LE's right of j in rank order,
// swap the pairs, swap the pivot.  Segments never interact, so all of one level run together.
//
//   big segments (more than QS_SMALL positions): one partition per level over the whole device,
//     every big segment of the level at once, in tiles of QS_TILE positions;
//   small segments: one workgroup each, keys and rows held in LDS; its waves partition the
//     segments of one level (64-position chunks, ballots), and segments of at most QS_TINY
//     positions finish on one thread each with the sequential algorithm itself.
//
// Keys are the column's values widened to int64 (integer-like types) or double (F, E), so every
// comparison is the reference's own.  NaN keys are rejected with HPA_E_DOMAIN: a NaN pivot
// compares greater than itself, and lib_partition's j scan then leaves the segment.
#include "hpa_internal.h"

#define QS_BLOCK 256
#define QS_ROWS 16
#define QS_TILE (QS_BLOCK * QS_ROWS)   // positions per tile of a big-segment pass
#define QS_SMALL 4096                  // largest segment finished in LDS
#define QS_TINY 32                     // segments finished by one thread
#define QS_LIST 256                    // LDS list of segments above QS_TINY (<= QS_SMALL / 33)
#define QS_MAX_LEVELS 16384            // device-wide levels before the sort gives up

// cmp(v, p) > 0 of lib_quicksort_cmp_item with flag `up` (h_libs.cc:461-482)
template<typename KT> __device__ __forceinline__ bool qs_gt(KT v, KT p, bool up){
    return v != p && (up ? !(v < p) : (v < p));
}

struct QsState { unsigned nbig[2]; unsigned nsmall; int nan; };

// ---------------------------------------------------------------- keys
template<typename T, typename KT>
__global__ void __launch_bounds__(256) k_qs_keys(const T *col, const uint32_t *perm, KT *K, L n, QsState *st){
    bool nan = false;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const KT v = (KT)col[perm[i]];
        nan |= v != v;
        K[i] = v;
    }
    if(__ballot(nan) && (threadIdx.x & 63) == 0) atomicOr(&st->nan, 1);
}

// ---------------------------------------------------------------- segment lists
__device__ __forceinline__ void qs_push(unsigned lo, unsigned hi, unsigned *blo, unsigned *bhi, unsigned *nb,
                                        unsigned *slo, unsigned *shi, unsigned *ns){
    const unsigned m = hi - lo;
    if(m < 2) return;
    if(m > QS_SMALL){ const unsigned k = atomicAdd(nb, 1u); blo[k] = lo; bhi[k] = hi; }
    else { const unsigned k = atomicAdd(ns, 1u); slo[k] = lo; shi[k] = hi; }
}
__global__ void k_qs_init(const uint32_t *lo, const uint32_t *hi, L nseg, L n, QsState *st,
                          unsigned *blo, unsigned *bhi, unsigned *slo, unsigned *shi){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (L)gridDim.x * blockDim.x){
        const unsigned a = lo ? lo[s] : 0u, b = hi ? hi[s] : (unsigned)n;
        qs_push(a, b, blo, bhi, &st->nbig[0], slo, shi, &st->nsmall);
    }
}

// ---------------------------------------------------------------- big segments, one level
// The level's count of big segments is read on the device (*nbp): the host queues a level
// before it knows the count, and grids are sized for the most segments a level can hold.
// tiles per segment: the positions lo+1 .. hi-1 in tiles of QS_TILE; 0 past the level's segments
// (up to nbmax), so a scan of nbmax entries puts the tile total at tstart[nb]
__global__ void k_qs_ntiles(const unsigned *blo, const unsigned *bhi, const unsigned *nbp, L nbmax, int64_t *nt){
    const L nb = *nbp;
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < nbmax; s += (L)gridDim.x * blockDim.x)
        nt[s] = s < nb ? ((L)(bhi[s] - blo[s] - 1) + QS_TILE - 1) / QS_TILE : 0;
}
// segment of tile t: the last s with tstart[s] <= t (every big segment has at least one tile)
__device__ __forceinline__ L qs_tile_seg(const int64_t *tstart, L nb, L t){
    L a = 0, b = nb - 1;
    while(a < b){ const L m = (a + b + 1) >> 1; if(tstart[m] <= t) a = m; else b = m - 1; }
    return a;
}
struct QsTile { L s; unsigned lo, hi, a, b; L r; };
__device__ __forceinline__ bool qs_tile(const int64_t *tstart, const unsigned *nbp, const unsigned *blo, const unsigned *bhi, QsTile &T){
    __shared__ L s_seg;
    const L t = blockIdx.x, nb = *nbp;
    if(nb == 0 || t >= tstart[nb]) return false;
    if(threadIdx.x == 0) s_seg = qs_tile_seg(tstart, nb, t);
    __syncthreads();
    T.s = s_seg; T.lo = blo[T.s]; T.hi = bhi[T.s]; T.r = t - tstart[T.s];
    T.a = T.lo + 1 + (unsigned)(T.r * QS_TILE);
    T.b = min(T.a + (unsigned)QS_TILE, T.hi);
    return true;
}
// per tile: number of G and of keys unequal to the pivot
template<typename KT>
__global__ void __launch_bounds__(QS_BLOCK) k_qs_count(const KT *K, const int64_t *tstart, const unsigned *nbp, const unsigned *blo,
                                                       const unsigned *bhi, bool up, unsigned *tg, unsigned *tne){
    QsTile T; if(!qs_tile(tstart, nbp, blo, bhi, T)) return;
    const KT pv = K[T.lo];
    unsigned g = 0, ne = 0;
    for(int i = 0; i < QS_ROWS; i++){
        const unsigned p = T.a + i * QS_BLOCK + threadIdx.x;
        if(p < T.b){ const KT v = K[p]; g += qs_gt(v, pv, up); ne += v != pv; }
    }
    for(int d = 32; d > 0; d >>= 1){ g += __shfl_down(g, d, 64); ne += __shfl_down(ne, d, 64); }
    __shared__ unsigned sg[QS_BLOCK / 64], sn[QS_BLOCK / 64];
    if((threadIdx.x & 63) == 0){ sg[threadIdx.x >> 6] = g; sn[threadIdx.x >> 6] = ne; }
    __syncthreads();
    if(threadIdx.x == 0){
        unsigned a = 0, b = 0;
        for(int w = 0; w < QS_BLOCK / 64; w++){ a += sg[w]; b += sn[w]; }
        tg[tstart[T.s] + T.r] = a; tne[tstart[T.s] + T.r] = b;
    }
}
// inclusive scan of one value per thread over a workgroup of 64 * NW threads
template<int NW>
__device__ __forceinline__ unsigned qs_block_incl(unsigned x, unsigned *sw, unsigned *total){
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for(int d = 1; d < 64; d <<= 1){ const unsigned y = __shfl_up(x, d, 64); if(lane >= d) x += y; }
    if(lane == 63) sw[w] = x;
    __syncthreads();
    unsigned base = 0, tot = 0;
    for(int k = 0; k < NW; k++){ if(k < w) base += sw[k]; tot += sw[k]; }
    __syncthreads();
    *total = tot;
    return x + base;
}
// per segment: exclusive scan of its tiles' G counts (gpre), the segment's G and unequal
// totals, and S = 0 (the tile holding position j overwrites it)
__global__ void __launch_bounds__(1024) k_qs_segscan(const int64_t *tstart, const unsigned *nbp, const unsigned *tg, const unsigned *tne,
                                                     unsigned *gpre, unsigned *segG, unsigned *segNE, unsigned *segS){
    __shared__ unsigned sw[16];
    const L s = blockIdx.x;
    if(s >= (L)*nbp) return;
    const L t0 = tstart[s], t1 = tstart[s + 1];
    unsigned run = 0, ne = 0;
    for(L base = t0; base < t1; base += 1024){
        const L t = base + threadIdx.x;
        const unsigned x = t < t1 ? tg[t] : 0u, y = t < t1 ? tne[t] : 0u;
        unsigned tot;
        const unsigned incl = qs_block_incl<16>(x, sw, &tot);
        if(t < t1) gpre[t] = run + incl - x;
        run += tot;
        unsigned tn;
        qs_block_incl<16>(y, sw, &tn);
        ne += tn;
    }
    if(threadIdx.x == 0){ segG[s] = run; segNE[s] = ne; segS[s] = 0; }
}
// Each G at or left of j writes its position at its rank from the left (PG), each LE right of j
// at its rank from the right (PL); the thread holding j writes S, the number of pairs.  A tile
// is staged in LDS and each thread takes 16 consecutive positions, so ranks follow position.
template<typename KT>
__global__ void __launch_bounds__(QS_BLOCK) k_qs_place(const KT *K, const int64_t *tstart, const unsigned *nbp, const unsigned *blo,
                                                       const unsigned *bhi, bool up, const unsigned *gpre,
                                                       const unsigned *segG, unsigned *segS, unsigned *PG, unsigned *PL){
    __shared__ KT sk[QS_TILE];
    __shared__ unsigned sw[QS_BLOCK / 64];
    QsTile T; if(!qs_tile(tstart, nbp, blo, bhi, T)) return;
    const KT pv = K[T.lo];
    const unsigned cnt = T.b - T.a;
    for(int i = 0; i < QS_ROWS; i++){
        const unsigned o = i * QS_BLOCK + threadIdx.x;
        if(o < cnt) sk[o] = K[T.a + o];
    }
    __syncthreads();
    const unsigned o0 = threadIdx.x * QS_ROWS;
    unsigned myg = 0;
    for(int i = 0; i < QS_ROWS; i++){ const unsigned o = o0 + i; if(o < cnt) myg += qs_gt(sk[o], pv, up); }
    unsigned tot;
    const unsigned incl = qs_block_incl<QS_BLOCK / 64>(myg, sw, &tot);
    const unsigned G = segG[T.s];
    const unsigned j = T.hi - 1 - G;
    unsigned before = gpre[tstart[T.s] + T.r] + incl - myg;     // G's in [lo+1, p)
    for(int i = 0; i < QS_ROWS; i++){
        const unsigned o = o0 + i;
        if(o >= cnt) break;
        const unsigned p = T.a + o;
        const bool g = qs_gt(sk[o], pv, up);
        if(g){ if(p <= j) PG[T.lo + before] = p; }
        else if(p > j){
            const unsigned rank_r = (T.hi - 1 - p) - (G - before);   // LE's in (p, hi)
            PL[T.lo + rank_r] = p;
        }
        before += g;
        if(p == j) segS[T.s] = before;
    }
}
// the S pairs of each segment swap keys and rows
template<typename KT>
__global__ void __launch_bounds__(QS_BLOCK) k_qs_swap(KT *K, uint32_t *P, const int64_t *tstart, const unsigned *nbp, const unsigned *blo,
                                                      const unsigned *bhi, const unsigned *segS, const unsigned *PG, const unsigned *PL){
    QsTile T; if(!qs_tile(tstart, nbp, blo, bhi, T)) return;
    const unsigned S = segS[T.s];
    const L k0 = T.r * QS_TILE;
    for(int i = 0; i < QS_ROWS; i++){
        const L k = k0 + i * QS_BLOCK + threadIdx.x;
        if(k >= S) break;
        const unsigned p = PG[T.lo + k], q = PL[T.lo + k];
        const KT a = K[p]; K[p] = K[q]; K[q] = a;
        const uint32_t b = P[p]; P[p] = P[q]; P[q] = b;
    }
}
// pivot swap and the children of every segment that is not all equal to its pivot
template<typename KT>
__global__ void k_qs_finish(KT *K, uint32_t *P, const unsigned *nbp, const unsigned *blo, const unsigned *bhi, const unsigned *segG,
                            const unsigned *segNE, QsState *st, int nxt, unsigned *nblo, unsigned *nbhi, unsigned *slo, unsigned *shi){
    const L nb = *nbp;
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < nb; s += (L)gridDim.x * blockDim.x){
        if(segNE[s] == 0) continue;                       // leftSame: every key equals the pivot
        const unsigned lo = blo[s], hi = bhi[s], j = hi - 1 - segG[s];
        const KT a = K[lo]; K[lo] = K[j]; K[j] = a;
        const uint32_t b = P[lo]; P[lo] = P[j]; P[j] = b;
        qs_push(lo, j, nblo, nbhi, &st->nbig[nxt], slo, shi, &st->nsmall);
        qs_push(j + 1, hi, nblo, nbhi, &st->nbig[nxt], slo, shi, &st->nsmall);
    }
}

// ---------------------------------------------------------------- small segments in LDS
__device__ __forceinline__ void qs_wave_sync(){
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// position of the k-th (0-based) set bit of m
__device__ __forceinline__ int qs_select(uint64_t m, int k){
    int pos = 0;
    for(int w = 32; w >= 1; w >>= 1){
        const uint64_t low = m & ((1ull << w) - 1);
        const int c = __popcll(low);
        if(k >= c){ k -= c; m >>= w; pos += w; } else m = low;
    }
    return pos;
}
template<typename KT>
__device__ __forceinline__ void qs_lds_swap(KT *sK, uint32_t *sP, int a, int b){
    const KT x = sK[a]; sK[a] = sK[b]; sK[b] = x;
    const uint32_t y = sP[a]; sP[a] = sP[b]; sP[b] = y;
}
// lib_quicksort_other itself on one thread, over [lo0, hi0) of the LDS copy
template<typename KT>
__device__ void qs_seq(KT *sK, uint32_t *sP, int lo0, int hi0, bool up){
    int st_lo[QS_TINY / 2 + 2], st_hi[QS_TINY / 2 + 2], sp = 0;
    st_lo[sp] = lo0; st_hi[sp] = hi0; sp++;
    while(sp > 0){
        sp--; const int low = st_lo[sp], high = st_hi[sp];
        if(low >= high) continue;
        const KT pv = sK[low];
        bool same = true;
        int i = low, j = high;
        for(;;){
            while(++i < high){
                const KT v = sK[i];
                if(v != pv) same = false;
                if(qs_gt(v, pv, up)) break;
            }
            do{ --j; } while(j > low && qs_gt(sK[j], pv, up));
            if(i >= j) break;
            qs_lds_swap(sK, sP, i, j);
        }
        // segments under two positions are finished, so only longer ones are kept: at most
        // (hi0 - lo0) / 2 are pending at once
        if(!same){ qs_lds_swap(sK, sP, low, j); if(j - low >= 2){ st_lo[sp] = low; st_hi[sp] = j; sp++; } }
        if(high - (j + 1) >= 2){ st_lo[sp] = j + 1; st_hi[sp] = high; sp++; }
    }
}
struct QsLds {
    unsigned list[2][QS_LIST];
    unsigned tiny[QS_SMALL / 2];
    uint64_t gm[QS_BLOCK / 64][64];
    unsigned pre[QS_BLOCK / 64][64];
    unsigned ncur, nnext, ntiny;
};
__device__ __forceinline__ void qs_lds_push(QsLds &s, int nxt, int lo, int hi){
    const int m = hi - lo;
    if(m < 2) return;
    const unsigned v = ((unsigned)lo << 16) | (unsigned)hi;
    if(m > QS_TINY) s.list[nxt][atomicAdd(&s.nnext, 1u)] = v;
    else s.tiny[atomicAdd(&s.ntiny, 1u)] = v;
}
// one partition of [lo, hi) (more than QS_TINY positions) by one wave
template<typename KT>
__device__ void qs_wave_partition(KT *sK, uint32_t *sP, QsLds &s, int nxt, int lo, int hi, bool up){
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const KT pv = sK[lo];
    const int m = hi - lo, nch = (m - 1 + 63) >> 6;
    bool anyne = false;
    for(int ch = 0; ch < nch; ch++){
        const int p = lo + 1 + (ch << 6) + lane;
        const bool valid = p < hi;
        const KT v = valid ? sK[p] : pv;
        const bool ne = valid && v != pv;
        const uint64_t gm = __ballot(ne && qs_gt(v, pv, up));
        anyne |= __ballot(ne) != 0;
        if(lane == 0) s.gm[w][ch] = gm;
    }
    if(!anyne) return;                                    // leftSame: finished
    qs_wave_sync();
    const uint64_t mygm = lane < nch ? s.gm[w][lane] : 0ull;
    const unsigned c = __popcll(mygm);
    unsigned incl = c;
    for(int d = 1; d < 64; d <<= 1){ const unsigned y = __shfl_up(incl, d, 64); if(lane >= d) incl += y; }
    const int G = (int)__shfl(incl, 63, 64);
    if(lane < nch) s.pre[w][lane] = incl - c;
    qs_wave_sync();
    const int j = hi - 1 - G;
    for(int ch = (j - lo) >> 6; ch < nch; ch++){
        const int p = lo + 1 + (ch << 6) + lane;
        const uint64_t gm = s.gm[w][ch];
        if(p > j && p < hi && !((gm >> lane) & 1ull)){
            const int ginc = (int)s.pre[w][ch] + __popcll(gm & ((2ull << lane) - 1ull));
            const int rank_r = (hi - 1 - p) - (G - ginc);
            int a = 0, b = nch - 1;
            while(a < b){ const int mid = (a + b + 1) >> 1; if((int)s.pre[w][mid] <= rank_r) a = mid; else b = mid - 1; }
            const int q = lo + 1 + (a << 6) + qs_select(s.gm[w][a], rank_r - (int)s.pre[w][a]);
            qs_lds_swap(sK, sP, p, q);
        }
    }
    qs_wave_sync();
    if(lane == 0){
        qs_lds_swap(sK, sP, lo, j);
        qs_lds_push(s, nxt, lo, j);
        qs_lds_push(s, nxt, j + 1, hi);
    }
    qs_wave_sync();
}
template<typename KT>
__global__ void __launch_bounds__(QS_BLOCK) k_qs_small(KT *K, uint32_t *P, const unsigned *slo, const unsigned *shi,
                                                       const QsState *st, bool up){
    __shared__ KT sK[QS_SMALL];
    __shared__ uint32_t sP[QS_SMALL];
    __shared__ QsLds s;
    const unsigned ns = st->nsmall;
    for(unsigned seg = blockIdx.x; seg < ns; seg += gridDim.x){
        const unsigned base = slo[seg], m = shi[seg] - base;
        for(unsigned i = threadIdx.x; i < m; i += QS_BLOCK){ sK[i] = K[base + i]; sP[i] = P[base + i]; }
        if(threadIdx.x == 0){
            s.ncur = 0; s.nnext = 0; s.ntiny = 0;
            if(m > QS_TINY) s.list[0][s.ncur++] = m;     // (0 << 16) | m
            else s.tiny[s.ntiny++] = m;
        }
        __syncthreads();
        int cur = 0;
        for(;;){
            const unsigned ncur = s.ncur;
            if(ncur == 0) break;
            for(unsigned it = threadIdx.x >> 6; it < ncur; it += QS_BLOCK / 64){
                const unsigned v = s.list[cur][it];
                qs_wave_partition(sK, sP, s, cur ^ 1, (int)(v >> 16), (int)(v & 0xffffu), up);
            }
            __syncthreads();
            if(threadIdx.x == 0){ s.ncur = s.nnext; s.nnext = 0; }
            cur ^= 1;
            __syncthreads();
        }
        for(unsigned it = threadIdx.x; it < s.ntiny; it += QS_BLOCK){
            const unsigned v = s.tiny[it];
            qs_seq(sK, sP, (int)(v >> 16), (int)(v & 0xffffu), up);
        }
        __syncthreads();
        for(unsigned i = threadIdx.x; i < m; i += QS_BLOCK){ K[base + i] = sK[i]; P[base + i] = sP[i]; }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- driver
// Levels are queued one ahead of what the host knows: after level L the count of level L + 1's
// big segments is published into mapped host memory (a begun read) and level L + 1 is queued at
// once; the host then waits for that count while level L + 1 runs, and stops when it is zero
// (that last queued level finds no segment and does nothing).  So the device never idles for a
// read between levels.
template<typename KT>
static int32_t qsort_levels(hpa_ctx *c, KT *K, uint32_t *P, L n, QsState *st, unsigned nbig, unsigned *bl[2][2],
                            unsigned *slo, unsigned *shi, bool up){
    int32_t e;
    const L nbmax = n / QS_SMALL + 2, tmax = n / QS_TILE + nbmax + 1;
    int64_t *tstart = (int64_t*)dalloc(c, 8 * (nbmax + 1));
    unsigned *tg = (unsigned*)dalloc(c, 4 * tmax), *tne = (unsigned*)dalloc(c, 4 * tmax), *gpre = (unsigned*)dalloc(c, 4 * tmax);
    unsigned *segG = (unsigned*)dalloc(c, 4 * nbmax), *segNE = (unsigned*)dalloc(c, 4 * nbmax), *segS = (unsigned*)dalloc(c, 4 * nbmax);
    unsigned *PG = (unsigned*)dalloc(c, 4 * n), *PL = (unsigned*)dalloc(c, 4 * n);
    if(!tstart || !tg || !tne || !gpre || !segG || !segNE || !segS || !PG || !PL) R HPA_E_MEMORY;
    if(nbig == 0) R 0;
    int cur = 0;
    auto queue_level = [&](int cu) -> int32_t {
        unsigned *blo = bl[cu][0], *bhi = bl[cu][1];
        const unsigned *nbp = &st->nbig[cu];
        hipMemsetAsync(&st->nbig[cu ^ 1], 0, 4, c->stream);
        hipLaunchKernelGGL(k_qs_ntiles, dim3(grid_for(nbmax, 256, 1)), dim3(256), 0, c->stream, blo, bhi, nbp, nbmax, tstart);
        int32_t e2;
        if((e2 = check_launch("order_qsort"))) R e2;
        if((e2 = scan_i64(c, tstart, tstart, nbmax, tstart + nbmax))) R e2;
        hipLaunchKernelGGL(k_qs_count<KT>, dim3((unsigned)tmax), dim3(QS_BLOCK), 0, c->stream, (const KT*)K, (const int64_t*)tstart, nbp,
                           (const unsigned*)blo, (const unsigned*)bhi, up, tg, tne);
        hipLaunchKernelGGL(k_qs_segscan, dim3((unsigned)nbmax), dim3(1024), 0, c->stream, (const int64_t*)tstart, nbp, (const unsigned*)tg,
                           (const unsigned*)tne, gpre, segG, segNE, segS);
        hipLaunchKernelGGL(k_qs_place<KT>, dim3((unsigned)tmax), dim3(QS_BLOCK), 0, c->stream, (const KT*)K, (const int64_t*)tstart, nbp,
                           (const unsigned*)blo, (const unsigned*)bhi, up, (const unsigned*)gpre, (const unsigned*)segG, segS, PG, PL);
        hipLaunchKernelGGL(k_qs_swap<KT>, dim3((unsigned)tmax), dim3(QS_BLOCK), 0, c->stream, K, P, (const int64_t*)tstart, nbp,
                           (const unsigned*)blo, (const unsigned*)bhi, (const unsigned*)segS, (const unsigned*)PG, (const unsigned*)PL);
        hipLaunchKernelGGL(k_qs_finish<KT>, dim3(grid_for(nbmax, 256, 1)), dim3(256), 0, c->stream, K, P, nbp, (const unsigned*)blo,
                           (const unsigned*)bhi, (const unsigned*)segG, (const unsigned*)segNE, st, cu ^ 1,
                           bl[cu ^ 1][0], bl[cu ^ 1][1], slo, shi);
        R check_launch("order_qsort");
    };
    // level 0 (nbig > 0 known), then: count of the next level begun, next level queued, count awaited
    if((e = queue_level(cur))) R e;
    for(int level = 1; ; level++){
        if(level >= QS_MAX_LEVELS){
            set_error_detail("order: the reference's quicksort recursion is deeper than %d levels of segments over %d rows on this input", QS_MAX_LEVELS, QS_SMALL);
            R HPA_E_NOT_IMPL;
        }
        uint64_t ticket;
        if((e = d2h_begin(c, &st->nbig[cur ^ 1], 4, &ticket))) R e;
        cur ^= 1;
        if((e = queue_level(cur))) R e;
        if((e = d2h_end(c, ticket, &nbig, 4))) R e;
        if(nbig == 0) break;
    }
    R 0;
}

template<typename T, typename KT>
static int32_t qsort_typed(hpa_ctx *c, const T *col, uint32_t *perm, L n, const uint32_t *seg_lo, const uint32_t *seg_hi,
                           L nseg, bool up){
    int32_t e;
    const L mark = c->dcur;
    const L nbmax = n / QS_SMALL + 2, nsmax = n / 2 + 2;
    KT *K = (KT*)dalloc(c, sizeof(KT) * n);
    QsState *st = (QsState*)dalloc(c, sizeof(QsState));
    unsigned *bl[2][2];
    for(int a = 0; a < 2; a++) for(int b = 0; b < 2; b++) bl[a][b] = (unsigned*)dalloc(c, 4 * nbmax);
    unsigned *slo = (unsigned*)dalloc(c, 4 * nsmax), *shi = (unsigned*)dalloc(c, 4 * nsmax);
    if(!K || !st || !bl[0][0] || !bl[0][1] || !bl[1][0] || !bl[1][1] || !slo || !shi){ arena_set(c, mark); R HPA_E_MEMORY; }
    hipMemsetAsync(st, 0, sizeof(QsState), c->stream);
    hipLaunchKernelGGL((k_qs_keys<T, KT>), dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, col, (const uint32_t*)perm, K, n, st);
    const L ns0 = seg_lo ? nseg : 1;
    hipLaunchKernelGGL(k_qs_init, dim3(grid_for(ns0, 256, 1)), dim3(256), 0, c->stream, seg_lo, seg_hi, ns0, n, st,
                       bl[0][0], bl[0][1], slo, shi);
    if((e = check_launch("order_qsort"))){ arena_set(c, mark); R e; }
    QsState h;
    if((e = d2h(c, &h, st, sizeof(QsState)))){ arena_set(c, mark); R e; }
    if(h.nan){ arena_set(c, mark); set_error_detail("order: NaN key in a quicksort-ordered column"); R HPA_E_DOMAIN; }
    if((e = qsort_levels<KT>(c, K, perm, n, st, h.nbig[0], bl, slo, shi, up))){ arena_set(c, mark); R e; }
    if((e = d2h(c, &h, st, sizeof(QsState)))){ arena_set(c, mark); R e; }
    if(h.nsmall > 0){
        int g = num_cus() * 2; if((L)g > (L)h.nsmall) g = (int)h.nsmall;
        hipLaunchKernelGGL(k_qs_small<KT>, dim3(g), dim3(QS_BLOCK), 0, c->stream, K, perm, (const unsigned*)slo, (const unsigned*)shi,
                           (const QsState*)st, up);
        if((e = check_launch("order_qsort_small"))){ arena_set(c, mark); R e; }
    }
    arena_set(c, mark);                                   // later work on this stream runs after
    R 0;
}

// True when the reference orders a column of this type with lib_quicksort_other
// (lib_quicksort, h_libs.cc:375-384; in lists every type but I/L/C, h_libs.cc:650-656).
bool qsort_ref_type(int32_t t){
    switch(t){
        case HPA_B: case HPA_J: case HPA_H: case HPA_F: case HPA_E: case HPA_M: case HPA_D:
        case HPA_Z: case HPA_U: case HPA_W: case HPA_T: R true;
    }
    R false;
}

// perm (n rows, n < 2^32) holds the current order; each segment [seg_lo[s], seg_hi[s]) of it
// (or [0, n) when seg_lo is null) is reordered exactly as lib_quicksort_other orders it by the
// column `col` (device payload of type typ, indexed by row), with flag `up`.
int32_t qsort_ref_perm(hpa_ctx *c, int32_t typ, const void *col, uint32_t *perm, L n, const uint32_t *seg_lo,
                       const uint32_t *seg_hi, L nseg, bool up){
    if(n < 2 || (seg_lo && nseg == 0)) R 0;
    if(n >= (L)UINT32_MAX){ set_error_detail("order: more than 2^32 rows"); R HPA_E_NOT_IMPL; }
    ProfScope ps("order_qsort", c->stream);
    switch(typ){
        case HPA_B: R qsort_typed<uint8_t, int64_t>(c, (const uint8_t*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_J: R qsort_typed<int8_t, int64_t>(c, (const int8_t*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_H: R qsort_typed<int16_t, int64_t>(c, (const int16_t*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T:
            R qsort_typed<int32_t, int64_t>(c, (const int32_t*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_Z: R qsort_typed<int64_t, int64_t>(c, (const int64_t*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_F: R qsort_typed<float, double>(c, (const float*)col, perm, n, seg_lo, seg_hi, nseg, up);
        case HPA_E: R qsort_typed<double, double>(c, (const double*)col, perm, n, seg_lo, seg_hi, nseg, up);
    }
    R HPA_E_NOT_IMPL;
}

// ---------------------------------------------------------------- list orders, column by column
// lib_order_by_list / lib_quicksort_list (h_libs.cc:619-697): starting from the identity, each
// column in turn orders every run of rows that the earlier columns left equal.  For one run:
//   lib_order_by_sorted (:604-617): when the column is already in the flag's order along the
//     run -- every adjacent pair compares <= 0 and some < 0 (ascending), or >= 0 and some > 0
//     (descending flag: a non-decreasing run) -- it is kept, and reversed in place for a
//     descending flag;
//   otherwise the run is sorted: I/L by the stable merge sort (lib_mergesort, :490-536),
//     C by lib_quicksort_char (:386-407), every other type by lib_quicksort_other;
// then the run is cut into runs of equal values for the next column.  Runs are independent, so
// every run of one column is handled at once.  Positions of finished runs are left alone.
#define QS_NONE 0xffffffffu

struct QsCol { int32_t typ; const void *p; };
// lib_quicksort_cmp_item (h_libs.cc:451-482) on rows a and b
__device__ __forceinline__ int qs_cmp_rows(QsCol c, uint32_t a, uint32_t b, bool up){
#define QS_CMP(T) { const T x = ((const T*)c.p)[a], y = ((const T*)c.p)[b]; if(x != y) return x < y ? (up ? -1 : 1) : (up ? 1 : -1); return 0; }
    switch(c.typ){
        case HPA_B: QS_CMP(uint8_t)
        case HPA_J: QS_CMP(int8_t)
        case HPA_C: QS_CMP(char)
        case HPA_H: QS_CMP(int16_t)
        case HPA_L: case HPA_Z: QS_CMP(int64_t)
        case HPA_F: QS_CMP(float)
        case HPA_E: QS_CMP(double)
        default: QS_CMP(int32_t)
    }
#undef QS_CMP
}
// segment of every position (segments are sorted by start), QS_NONE outside them
__global__ void k_ql_sid(const uint32_t *slo, const uint32_t *shi, L ns, L n, uint32_t *sid){
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        L a = 0, b = ns - 1;
        while(a < b){ const L m = (a + b + 1) >> 1; if((L)slo[m] <= p) a = m; else b = m - 1; }
        sid[p] = (ns > 0 && (L)slo[a] <= p && p < (L)shi[a]) ? (uint32_t)a : QS_NONE;
    }
}
// lib_order_by_sorted's two facts per segment: some adjacent pair compares as the flag's
// order (bit 0) and some pair against it (bit 1).  A wave whose positions share one segment
// sends one atomic.
__global__ void __launch_bounds__(256) k_ql_sorted(QsCol col, const uint32_t *perm, const uint32_t *sid, const uint32_t *slo,
                                                   L n, bool up, uint32_t *sflag){
    for(L p0 = (L)blockIdx.x * blockDim.x; p0 < n; p0 += (L)gridDim.x * blockDim.x){
        const L p = p0 + threadIdx.x;
        uint32_t s = QS_NONE, bits = 0;
        if(p < n){
            s = sid[p];
            if(s != QS_NONE && (L)slo[s] < p){
                const int t = qs_cmp_rows(col, perm[p - 1], perm[p], up);
                const int want = up ? -1 : 1;
                bits = t == want ? 1u : t == -want ? 2u : 0u;
            }
        }
        const uint32_t s0 = __shfl(s, 0, 64);
        const bool same = __ballot(s != s0) == 0;
        if(same){
            uint32_t b = bits;
            for(int d = 32; d > 0; d >>= 1) b |= __shfl_down(b, d, 64);
            if((threadIdx.x & 63) == 0 && b && s0 != QS_NONE) atomicOr(&sflag[s0], b);
        } else if(bits) atomicOr(&sflag[s], bits);
    }
}
// kept runs with a descending flag are reversed in place; the runs to sort are listed
__global__ void k_ql_keep(L ns, const uint32_t *sflag, uint32_t *todo){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (L)gridDim.x * blockDim.x)
        todo[s] = !(sflag[s] == 1u);                     // isOrder: some pair in order, none against
}
__global__ void k_ql_init(uint32_t *slo, uint32_t *shi, L n){ slo[0] = 0; shi[0] = (uint32_t)n; }
__global__ void k_ql_reverse(uint32_t *perm, const uint32_t *sid, const uint32_t *slo, const uint32_t *shi, const uint32_t *todo, L n){
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        const uint32_t s = sid[p];
        if(s == QS_NONE || todo[s]) continue;
        const L lo = slo[s], hi = shi[s];
        if(p - lo < (hi - lo) / 2){ const L q = hi - 1 - (p - lo); const uint32_t t = perm[p]; perm[p] = perm[q]; perm[q] = t; }
    }
}
// compaction of the listed segments, in segment order
__global__ void k_ql_pick(const uint32_t *slo, const uint32_t *shi, L ns, const uint32_t *todo, const int64_t *pos,
                          uint32_t *olo, uint32_t *ohi){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (L)gridDim.x * blockDim.x)
        if(todo[s]){ olo[pos[s]] = slo[s]; ohi[pos[s]] = shi[s]; }
}
// stable order keys of I/L/C rows inside the runs being sorted (descending: complemented), and
// each position's block: the run's start inside a run, the position itself outside.  Char runs
// under ten rows stay out: lib_quicksort_char exchanges them in place instead.
__global__ void k_ql_keys(QsCol col, const uint32_t *perm, const uint32_t *sid, const uint32_t *slo, const uint32_t *shi,
                          const uint32_t *todo, L n, bool desc, int bits, uint64_t *key, uint32_t *blk){
    const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        const uint32_t s = sid[p];
        const bool in = s != QS_NONE && todo[s] && !(col.typ == HPA_C && shi[s] - slo[s] < 10);
        uint64_t k = 0;
        if(in){
            const uint32_t r = perm[p];
            switch(col.typ){
                case HPA_I: k = (uint32_t)((const int32_t*)col.p)[r] ^ 0x80000000u; break;
                case HPA_L: k = (uint64_t)((const int64_t*)col.p)[r] ^ 0x8000000000000000ull; break;
                default:    k = (uint8_t)((const char*)col.p)[r]; break;          // C: counting sort on the byte
            }
            if(desc) k = ~k & mask;
        }
        key[p] = k;
        blk[p] = in ? slo[s] : (uint32_t)p;
    }
}
__global__ void k_ql_gather_u32(const uint32_t *src, const uint32_t *idx, L n, uint32_t *dst){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}
// lib_quicksort_char below ten rows (h_libs.cc:390-392): the all-pairs exchange with `>`, one
// thread per run, then the run reversed for a descending flag
__global__ void k_ql_char_small(QsCol col, uint32_t *perm, const uint32_t *olo, const uint32_t *ohi, L no, bool up){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < no; s += (L)gridDim.x * blockDim.x){
        const L lo = olo[s], hi = ohi[s];
        if(hi - lo >= 10) continue;
        const char *str = (const char*)col.p;
        for(L i = lo; i < hi; i++) for(L j = lo; j < hi; j++)
            if(str[perm[i]] > str[perm[j]]){ const uint32_t t = perm[i]; perm[i] = perm[j]; perm[j] = t; }
        if(!up) for(L i = lo, k = hi - 1; i < k; i++, k--){ const uint32_t t = perm[i]; perm[i] = perm[k]; perm[k] = t; }
    }
}
// char runs of ten rows or more: reversed after the ascending counting sort for a descending flag
__global__ void k_ql_char_rev(uint32_t *perm, const uint32_t *sid, const uint32_t *slo, const uint32_t *shi, const uint32_t *todo, L n){
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        const uint32_t s = sid[p];
        if(s == QS_NONE || !todo[s]) continue;
        const L lo = slo[s], hi = shi[s];
        if(hi - lo < 10) continue;
        if(p - lo < (hi - lo) / 2){ const L q = hi - 1 - (p - lo); const uint32_t t = perm[p]; perm[p] = perm[q]; perm[q] = t; }
    }
}
// run heads of the column inside the current segments
__global__ void k_ql_heads(QsCol col, const uint32_t *perm, const uint32_t *sid, const uint32_t *slo, L n, uint32_t *head){
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        const uint32_t s = sid[p];
        head[p] = s != QS_NONE && ((L)slo[s] == p || qs_cmp_rows(col, perm[p - 1], perm[p], true) != 0);
    }
}
// run starts in position order, then each run's end: the next start when it lies inside the
// run's segment, else the segment's end; runs of two rows or more are kept
__global__ void k_ql_runs(const uint32_t *head, const int64_t *hpos, L n, uint32_t *rlo){
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x)
        if(head[p]) rlo[hpos[p]] = (uint32_t)p;
}
__global__ void k_ql_runs_end(const uint32_t *rlo, L nr, const uint32_t *sid, const uint32_t *shi, uint32_t *rhi, uint32_t *keep){
    for(L k = (L)blockIdx.x * blockDim.x + threadIdx.x; k < nr; k += (L)gridDim.x * blockDim.x){
        const uint32_t a = rlo[k], se = shi[sid[a]];
        const uint32_t e = (k + 1 < nr && rlo[k + 1] < se) ? rlo[k + 1] : se;
        rhi[k] = e; keep[k] = e - a >= 2;
    }
}

int32_t order_list_ref(hpa_ctx *c, int nc, const int32_t *typ, const void *const *cols, const uint8_t *flags, L n,
                       uint32_t **perm_out){
    int32_t e;
    if(n >= (L)UINT32_MAX){ set_error_detail("order: more than 2^32 rows"); R HPA_E_NOT_IMPL; }
    ProfScope ps("order_list_ref", c->stream);
    const L cap = n / 2 + 2;
    uint32_t *perm = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)), *sid = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1));
    uint32_t *slo = (uint32_t*)dalloc(c, 4 * cap), *shi = (uint32_t*)dalloc(c, 4 * cap);
    uint32_t *sflag = (uint32_t*)dalloc(c, 4 * cap), *todo = (uint32_t*)dalloc(c, 4 * cap);
    uint32_t *olo = (uint32_t*)dalloc(c, 4 * cap), *ohi = (uint32_t*)dalloc(c, 4 * cap);
    int64_t *pos = (int64_t*)dalloc(c, 8 * (n + 1));
    uint32_t *head = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1));
    if(!perm || !sid || !slo || !shi || !sflag || !todo || !olo || !ohi || !pos || !head) R HPA_E_MEMORY;
    if((e = iota_u32(c, perm, n))) R e;
    *perm_out = perm;
    if(n < 2) R 0;
    hipLaunchKernelGGL(k_ql_init, dim3(1), dim3(1), 0, c->stream, slo, shi, n);
    L ns = 1;
    const int gn = grid_for(n, 256, 4);
    for(int q = 0; q < nc && ns > 0; q++){
        const QsCol col{typ[q], cols[q]};
        const bool up = flags[q] != 0;
        auto gs = [&](L k){ R dim3(grid_for(k, 256, 1)); };
        hipLaunchKernelGGL(k_ql_sid, dim3(gn), dim3(256), 0, c->stream, (const uint32_t*)slo, (const uint32_t*)shi, ns, n, sid);
        hipMemsetAsync(sflag, 0, 4 * ns, c->stream);
        hipLaunchKernelGGL(k_ql_sorted, dim3(grid_for(n, 256, 1)), dim3(256), 0, c->stream, col, (const uint32_t*)perm,
                           (const uint32_t*)sid, (const uint32_t*)slo, n, up, sflag);
        hipLaunchKernelGGL(k_ql_keep, gs(ns), dim3(256), 0, c->stream, ns, (const uint32_t*)sflag, todo);
        if(!up) hipLaunchKernelGGL(k_ql_reverse, dim3(gn), dim3(256), 0, c->stream, perm, (const uint32_t*)sid, (const uint32_t*)slo,
                                   (const uint32_t*)shi, (const uint32_t*)todo, n);
        if((e = check_launch("order_list_ref"))) R e;
        // the runs to sort
        if((e = scan_u32_to_i64(c, todo, pos, ns, pos + ns))) R e;
        int64_t no = 0; if((e = d2h(c, &no, pos + ns, 8))) R e;
        if(no > 0){
            hipLaunchKernelGGL(k_ql_pick, gs(ns), dim3(256), 0, c->stream, (const uint32_t*)slo, (const uint32_t*)shi, ns,
                               (const uint32_t*)todo, (const int64_t*)pos, olo, ohi);
            if(qsort_ref_type(col.typ)){
                if((e = qsort_ref_perm(c, col.typ, col.p, perm, n, olo, ohi, no, up))) R e;
            } else if(col.typ == HPA_I || col.typ == HPA_L || col.typ == HPA_C){
                const L mark = c->dcur;
                const int bits = col.typ == HPA_I ? 32 : col.typ == HPA_L ? 64 : 8;
                uint64_t *key = (uint64_t*)dalloc(c, 8 * n);
                uint32_t *blk = (uint32_t*)dalloc(c, 4 * n), *blk2 = (uint32_t*)dalloc(c, 4 * n), *np = (uint32_t*)dalloc(c, 4 * n);
                if(!key || !blk || !blk2 || !np) R HPA_E_MEMORY;
                // merge sort: stable in the flag's order; counting sort: stable ascending
                hipLaunchKernelGGL(k_ql_keys, dim3(gn), dim3(256), 0, c->stream, col, (const uint32_t*)perm, (const uint32_t*)sid,
                                   (const uint32_t*)slo, (const uint32_t*)shi, (const uint32_t*)todo, n, col.typ != HPA_C && !up,
                                   bits, key, blk);
                if((e = check_launch("order_list_ref"))) R e;
                uint64_t *sk; uint32_t *sv;
                if((e = radix_sort_pairs_u64(c, key, nullptr, n, 0, bits, &sk, &sv))) R e;
                const uint32_t *byk = sv;
                if(!byk){ uint32_t *id = (uint32_t*)dalloc(c, 4 * n); if(!id) R HPA_E_MEMORY; if((e = iota_u32(c, id, n))) R e; byk = id; }
                hipLaunchKernelGGL(k_ql_gather_u32, dim3(gn), dim3(256), 0, c->stream, (const uint32_t*)blk, byk, n, blk2);
                if((e = check_launch("order_list_ref"))) R e;
                uint32_t *sb, *sv2;
                if((e = radix_sort_pairs_u32(c, blk2, byk, n, 0, 32, &sb, &sv2))) R e;
                hipLaunchKernelGGL(k_ql_gather_u32, dim3(gn), dim3(256), 0, c->stream, (const uint32_t*)perm, (const uint32_t*)sv2, n, np);
                hipMemcpyAsync(perm, np, 4 * n, hipMemcpyDeviceToDevice, c->stream);
                if(col.typ == HPA_C){
                    hipLaunchKernelGGL(k_ql_char_small, gs(no), dim3(256), 0, c->stream, col, perm, (const uint32_t*)olo, (const uint32_t*)ohi, (L)no, up);
                    if(!up) hipLaunchKernelGGL(k_ql_char_rev, dim3(gn), dim3(256), 0, c->stream, perm, (const uint32_t*)sid,
                                               (const uint32_t*)slo, (const uint32_t*)shi, (const uint32_t*)todo, n);
                }
                if((e = check_launch("order_list_ref"))) R e;
                arena_set(c, mark);
            } else {
                set_error_detail("order: type %d in a list order", (int)col.typ);
                R HPA_E_NOT_IMPL;
            }
        }
        if(q + 1 == nc) break;
        // runs of equal values of this column become the next column's segments
        hipLaunchKernelGGL(k_ql_heads, dim3(gn), dim3(256), 0, c->stream, col, (const uint32_t*)perm, (const uint32_t*)sid,
                           (const uint32_t*)slo, n, head);
        if((e = scan_u32_to_i64(c, head, pos, n, pos + n))) R e;
        int64_t nr = 0; if((e = d2h(c, &nr, pos + n, 8))) R e;
        const L mark = c->dcur;
        uint32_t *rlo = (uint32_t*)dalloc(c, 4 * (nr > 0 ? nr : 1)), *rhi = (uint32_t*)dalloc(c, 4 * (nr > 0 ? nr : 1));
        uint32_t *keep = (uint32_t*)dalloc(c, 4 * (nr > 0 ? nr : 1));
        int64_t *kpos = (int64_t*)dalloc(c, 8 * (nr + 1));
        if(!rlo || !rhi || !keep || !kpos) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_ql_runs, dim3(gn), dim3(256), 0, c->stream, (const uint32_t*)head, (const int64_t*)pos, n, rlo);
        if(nr > 0) hipLaunchKernelGGL(k_ql_runs_end, gs(nr), dim3(256), 0, c->stream, (const uint32_t*)rlo, (L)nr, (const uint32_t*)sid,
                                      (const uint32_t*)shi, rhi, keep);
        if((e = check_launch("order_list_ref"))) R e;
        if((e = scan_u32_to_i64(c, keep, kpos, nr, kpos + nr))) R e;
        int64_t nk = 0; if((e = d2h(c, &nk, kpos + nr, 8))) R e;
        if(nk > cap){ set_error_detail("order: run list overflow"); R HPA_E_GENERAL; }
        if(nk > 0) hipLaunchKernelGGL(k_ql_pick, gs(nr), dim3(256), 0, c->stream, (const uint32_t*)rlo, (const uint32_t*)rhi, (L)nr,
                                      (const uint32_t*)keep, (const int64_t*)kpos, slo, shi);
        if((e = check_launch("order_list_ref"))) R e;
        arena_set(c, mark);
        ns = nk;
    }
    R 0;
}
