// Batched execution: many independent filter -> compress -> group -> sum queries in one
// captured launch sequence.
//
// Query q reads its own table (i32 key, f64 value, i32 filter column) and computes, over the
// rows with `flt <op> c`, per distinct key: the key, the row count and the FP64 sum of the
// values, groups in the order pfnGroup gives i32 keys (ascending as unsigned 32-bit values,
// h_libs.cc:1295-1329).  That is the result of the per-query primitive sequence
//   compare -> compress(key), compress(val) -> group -> keys/values -> index(key, keys),
//   raze(each(sum, each_right(index, val, groups))), raze(each(len, ...))
// (the FP1/FP7 patterns of optimizer/pattern.cc:180-219 over a q3.hir-style group-by).
//
// Slots are dense over each query's key range [kmin, kmax], measured when the plan is built.
// Three aggregation paths, chosen per query by its key range and the tunables:
//   LDS         range <= lds_keys: a workgroup aggregates a row tile in LDS (one replica per
//               wave when the range is small) and flushes one global add per touched key.
//   partitioned range > lds_keys with partitioning on: each row tile counts its rows per key
//               partition (part_keys consecutive keys), a per-query scan of the (partition,
//               tile) counts gives every tile its place in each partition, the tile scatters
//               its rows there with LDS cursors, and one workgroup per partition aggregates it
//               in LDS and owns its slots outright (no global atomics anywhere on this path).
//   global      otherwise: compensated global atomics straight into the slots.
// Kernels, captured once into a hipGraph and replayed (all persistent over an atomic queue of
// items of every query):
//   k_batch_clear    zero the slots that are accumulated atomically (a fill kernel: no memset node)
//   k_batch_agg      row tiles: LDS / global aggregation, or partition counts
//   k_batch_pscan    (partition, tile) offsets (one workgroup per partitioned query)
//   k_batch_scatter  row tiles of partitioned queries into their partitions
//   k_batch_pagg     partitions (split over several workgroups when large): LDS aggregation
//   k_batch_emit     slot tiles: non-empty slots compacted in key order into one output of all
//                    queries (one decoupled look-back across every query's slot tiles)
//
// FP64 sums are compensated: every atomic add returns the previous value, so the rounding
// error of that add is recovered exactly (Knuth's TwoSum) and accumulated beside the sum;
// the group's result is hi + lo.  The error of the result is about one rounding of the exact
// sum whatever order the adds land in, the accuracy class of the compensated sums of the
// per-query path.
#include "hpa_internal.h"

#define ST_AGG  (1ull << 62)
#define ST_PRE  (2ull << 62)
#define ST_MASK ((1ull << 62) - 1)
#define EMIT_ITEMS 8
#define EMIT_TILE (256 * EMIT_ITEMS)
#define BATCH_MAX_RANGE (1ll << 26)
#define PART_MAX 4096          // partitions per query (LDS counters of the count and scatter kernels)
#define PART_MIN_KEYS 1024     // smallest partition width the tunables may choose
#define PART_MIN_TILE 8192     // smallest row tile the tunables may choose (sizes the count matrix)
#define BATCH_CTRS 4           // queue counters of a run

struct BQuery {
    const int32_t *key; const double *val; const int32_t *flt;
    int64_t n; int32_t op; int32_t thr;
    int32_t kmin; int32_t range; int32_t rot; int32_t part;   // slot = key - kmin; emit starts at rot
    double *hi; double *lo; unsigned int *cnt;                // dense slots [range]
    int64_t *ng;                                              // [0] end of its groups in the output, [1] keys out of range
    int64_t tile0, ntiles, etile0, entiles;
    // partitioned aggregation
    int32_t nparts, psplit;                                   // partitions; workgroups per partition
    unsigned *pcnt;                                           // rows per (partition, tile), partition-major
    int64_t *poff;                                            // exclusive scan of pcnt [nparts * ntiles + 1]
    uint16_t *skey; double *sval;                             // partition-contiguous rows (capacity n)
    int64_t pitem0;
    int64_t stile0;                                           // first scatter item (tiles of partitioned queries)
    unsigned long long *pcsum;                                // per PS_CHUNK-entry chunk of pcnt: its sum
    int64_t pch0;                                             // first scan chunk of the query
};

struct BPlan {
    BQuery *q; int nq;
    int64_t total_tiles, total_etiles, total_pitems, total_stiles;
    unsigned *ctr;                                            // [0] agg [1] emit [2] scatter [3] pagg
    uint64_t *estatus;                                        // emit look-back, per slot tile of all queries
    int32_t *gkey; double *gsum; int64_t *gcnt;               // groups of every query, query after query
    int tile_rows, lds_keys, part_keys, part_shift;
    const int *pq; int npq;                                   // partitioned queries
    int64_t total_pchunks;                                    // scan chunks of all partitioned queries
};

enum { SEL_TILE, SEL_EMIT, SEL_PART, SEL_STILE };
__device__ __forceinline__ int64_t item0_of(const BQuery &q, int sel){
    return sel == SEL_TILE ? q.tile0 : sel == SEL_EMIT ? q.etile0 : sel == SEL_PART ? q.pitem0 : q.stile0;
}
// the query owning queue item t: the last query whose first item is <= t
__device__ __forceinline__ int find_query(const BQuery *q, int nq, int64_t t, int sel){
    int lo = 0, hi = nq - 1;
    while(lo < hi){
        int mid = (lo + hi + 1) >> 1;
        if(item0_of(q[mid], sel) <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// A query's columns and slot arrays are reached through its descriptor in device memory, so the
// compiler cannot tell that they are global memory and would issue flat accesses, each waiting
// for the one before; these views say that they are.  Row loads go to a clamped index (the last
// row of the range) instead of under a branch, so a thread's loads are all in flight together.
template<typename T> using gptr = T __attribute__((address_space(1))) *;
template<typename T> __device__ __forceinline__ gptr<T> G(T *p){ return (gptr<T>)p; }

// The persistent kernels look an item's query up in an LDS copy of every query's first item
// (up to QTAB queries; more are searched in the descriptors in global memory), filled before the
// barrier of the first next_item.  A query with no items of a queue shares its first item with
// the next query and never owns one: the search takes the last query whose first item is <= t.
#define QTAB 256
__device__ __forceinline__ void load_qtab(const BPlan &p, int sel, int64_t *t0){
    if(p.nq <= QTAB) for(int k = threadIdx.x; k < p.nq; k += blockDim.x) t0[k] = item0_of(p.q[k], sel);
}
__device__ __forceinline__ int find_query_tab(const BPlan &p, const int64_t *t0, int64_t t, int sel){
    if(p.nq > QTAB) return find_query(p.q, p.nq, t, sel);
    int lo = 0, hi = p.nq - 1;
    while(lo < hi){
        const int mid = (lo + hi + 1) >> 1;
        if(t0[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool batch_pass(int op, int32_t f, int32_t c){
    switch(op){
    case HPA_OP_LT:  return f <  c;
    case HPA_OP_LEQ: return f <= c;
    case HPA_OP_GT:  return f >  c;
    case HPA_OP_GEQ: return f >= c;
    case HPA_OP_EQ:  return f == c;
    default:         return f != c;
    }
}

__device__ __forceinline__ double two_sum_err(double a, double b, double s){
    double bp = s - a;
    return (a - (s - bp)) + (b - bp);
}

// next item of one queue for the whole workgroup
__device__ __forceinline__ int64_t next_item(unsigned *ctr, int64_t *s_item){
    if(threadIdx.x == 0) *s_item = (int64_t)atomicAdd(ctr, 1u);
    __syncthreads();
    const int64_t item = *s_item;
    __syncthreads();
    return item;
}


__global__ void k_batch_minmax(BQuery *qs, int nq, int32_t *mm){
    __shared__ int32_t smin[4], smax[4];
    for(int qi = blockIdx.y; qi < nq; qi += gridDim.y){
        const BQuery &q = qs[qi];
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q.n; i += (int64_t)gridDim.x * blockDim.x){
            int32_t k = q.key[i]; lo = k < lo ? k : lo; hi = k > hi ? k : hi;
        }
        for(int o = 32; o; o >>= 1){
            int32_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
            lo = a < lo ? a : lo; hi = b > hi ? b : hi;
        }
        if((threadIdx.x & 63) == 0){ smin[threadIdx.x >> 6] = lo; smax[threadIdx.x >> 6] = hi; }
        __syncthreads();
        if(threadIdx.x == 0){
            for(int w = 1; w < (int)(blockDim.x >> 6); w++){ lo = smin[w] < lo ? smin[w] : lo; hi = smax[w] > hi ? smax[w] : hi; }
            atomicMin(&mm[2 * qi], lo); atomicMax(&mm[2 * qi + 1], hi);
        }
        __syncthreads();
    }
}

__global__ void k_batch_clear(BPlan p){
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for(int qi = blockIdx.y; qi < p.nq; qi += gridDim.y){
        const BQuery &q = p.q[qi];
        const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if(!q.part || q.psplit > 1) for(int64_t i = t0; i < q.range; i += stride){ q.hi[i] = 0.0; q.lo[i] = 0.0; q.cnt[i] = 0u; }
        if(blockIdx.x == 0 && threadIdx.x < 2) q.ng[threadIdx.x] = 0;
    }
    if(blockIdx.y == 0) for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.total_etiles; i += stride) p.estatus[i] = 0;
    if(blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < BATCH_CTRS) p.ctr[threadIdx.x] = 0;
}

extern __shared__ double g_batch_lds[];

__global__ void k_batch_agg(BPlan p){
    __shared__ int64_t s_item;
    __shared__ int64_t s_t0[QTAB];
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
    load_qtab(p, SEL_TILE, s_t0);
    while(true){
        const int64_t item = next_item(&p.ctr[0], &s_item);
        if(item >= p.total_tiles) break;
        const BQuery &q = p.q[find_query_tab(p, s_t0, item, SEL_TILE)];
        const int64_t r0 = (item - q.tile0) * p.tile_rows;
        const int64_t r1 = r0 + p.tile_rows < q.n ? r0 + p.tile_rows : q.n;
        const int32_t kmin = q.kmin, op = q.op, thr = q.thr;
        const uint32_t range = (uint32_t)q.range;
        bool bad = false;
        if(q.part){
            // partitioned query: rows per key partition (reads the filter and key columns only)
            unsigned *lc = (unsigned*)g_batch_lds;
            for(int k = threadIdx.x; k < q.nparts; k += blockDim.x) lc[k] = 0u;
            __syncthreads();
            // four rows per thread per step, their loads issued together
            const int64_t step = (int64_t)blockDim.x;
            const gptr<const int32_t> flt = G(q.flt), key = G(q.key);
            // the next step's rows are loaded before this step's counts (clamped to the last row)
            int32_t nf[4], nk[4];
            #pragma unroll
            for(int u = 0; u < 4; u++){ const int64_t i = r0 + threadIdx.x + u * step, ic = i < r1 ? i : r1 - 1; nf[u] = flt[ic]; nk[u] = key[ic]; }
            for(int64_t i0 = r0 + threadIdx.x; i0 < r1; i0 += 4 * step){
                int32_t f[4], k[4];
                #pragma unroll
                for(int u = 0; u < 4; u++){ f[u] = nf[u]; k[u] = nk[u]; }
                #pragma unroll
                for(int u = 0; u < 4; u++){ const int64_t i = i0 + (4 + u) * step, ic = i < r1 ? i : r1 - 1; nf[u] = flt[ic]; nk[u] = key[ic]; }
                #pragma unroll
                for(int u = 0; u < 4; u++){
                    if(i0 + u * step < r1 && batch_pass(op, f[u], thr)){
                        const uint32_t s = (uint32_t)k[u] - (uint32_t)kmin;
                        if(s < range) atomicAdd(&lc[s >> p.part_shift], 1u); else bad = true;
                    }
                }
            }
            __syncthreads();
            const int64_t tile = item - q.tile0;
            const gptr<unsigned> pcnt = G(q.pcnt);
            for(int k = threadIdx.x; k < q.nparts; k += blockDim.x) pcnt[(int64_t)k * q.ntiles + tile] = lc[k];
            __syncthreads();
        } else if(q.range <= p.lds_keys){
            int reps = p.lds_keys / q.range; reps = reps > nw ? nw : reps;
            const int tot = reps * q.range;
            double *lhi = g_batch_lds, *llo = lhi + p.lds_keys;
            unsigned *lcnt = (unsigned*)(llo + p.lds_keys);
            for(int k = threadIdx.x; k < tot; k += blockDim.x){ lhi[k] = 0.0; llo[k] = 0.0; lcnt[k] = 0u; }
            __syncthreads();
            const int base = (w % reps) * q.range;
            const int64_t step = (int64_t)blockDim.x;
            const gptr<const int32_t> flt = G(q.flt), key = G(q.key);
            const gptr<const double> val = G(q.val);
            // the next step's rows are loaded before this step's adds (clamped to the last row)
            int32_t nf[4], nk[4]; double nv[4];
            #pragma unroll
            for(int u = 0; u < 4; u++){
                const int64_t i = r0 + threadIdx.x + u * step, ic = i < r1 ? i : r1 - 1;
                nf[u] = flt[ic]; nk[u] = key[ic]; nv[u] = val[ic];
            }
            for(int64_t i0 = r0 + threadIdx.x; i0 < r1; i0 += 4 * step){
                int32_t f[4], k[4]; double v[4];
                #pragma unroll
                for(int u = 0; u < 4; u++){ f[u] = nf[u]; k[u] = nk[u]; v[u] = nv[u]; }
                #pragma unroll
                for(int u = 0; u < 4; u++){
                    const int64_t i = i0 + (4 + u) * step, ic = i < r1 ? i : r1 - 1;
                    nf[u] = flt[ic]; nk[u] = key[ic]; nv[u] = val[ic];
                }
                // the step's returning adds first, then their rounding errors and counts
                bool ok[4]; uint32_t sl[4]; double old[4];
                #pragma unroll
                for(int u = 0; u < 4; u++){
                    sl[u] = (uint32_t)k[u] - (uint32_t)kmin;
                    const bool pass = i0 + u * step < r1 && batch_pass(op, f[u], thr);
                    ok[u] = pass && sl[u] < range;
                    if(pass && !ok[u]) bad = true;
                    if(ok[u]) old[u] = atomicAdd(&lhi[base + sl[u]], v[u]);
                }
                #pragma unroll
                for(int u = 0; u < 4; u++){
                    if(!ok[u]) continue;
                    atomicAdd(&llo[base + sl[u]], two_sum_err(old[u], v[u], old[u] + v[u]));
                    atomicAdd(&lcnt[base + sl[u]], 1u);
                }
            }
            __syncthreads();
            for(int k = threadIdx.x; k < q.range; k += blockDim.x){
                unsigned c = 0; double h = 0.0, l = 0.0;
                for(int r = 0; r < reps; r++){
                    int j = r * q.range + k;
                    if(lcnt[j]){ c += lcnt[j]; double s = h + lhi[j]; l += two_sum_err(h, lhi[j], s) + llo[j]; h = s; }
                }
                if(c){
                    double old = unsafeAtomicAdd(&q.hi[k], h);
                    unsafeAtomicAdd(&q.lo[k], two_sum_err(old, h, old + h) + l);
                    atomicAdd(&q.cnt[k], c);
                }
            }
            __syncthreads();
        } else {
            const gptr<const int32_t> flt = G(q.flt), key = G(q.key);
            const gptr<const double> val = G(q.val);
            for(int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x){
                if(batch_pass(op, flt[i], thr)){
                    uint32_t s = (uint32_t)key[i] - (uint32_t)kmin;
                    if(s < range){
                        double v = val[i];
                        double old = unsafeAtomicAdd(&q.hi[s], v);
                        unsafeAtomicAdd(&q.lo[s], two_sum_err(old, v, old + v));
                        atomicAdd(&q.cnt[s], 1u);
                    } else bad = true;
                }
            }
        }
        if(bad) atomicAdd((unsigned long long*)&q.ng[1], 1ull);
    }
}

// Exclusive scan of one partitioned query's (partition, tile) counts per workgroup, in rounds of
// 16384 entries (sixteen consecutive per thread: a wave scan of the thread sums, then the 16
// wave totals), carrying the running total between rounds.
#define PS_PER 16
// The (partition, tile) count matrices of the partitioned queries are scanned in chunks of
// PS_CHUNK entries, one workgroup a chunk over every query at once: k_batch_psum sums each
// chunk, k_batch_pscan scans its chunk from the sum of the query's earlier chunks (a few values),
// so the largest query's matrix no longer runs through one workgroup.
#define PS_CHUNK (1024 * PS_PER)
// the partitioned query (index into p.q) owning scan chunk t
__device__ __forceinline__ int find_pchunk(const BPlan &p, int64_t t){
    int lo = 0, hi = p.npq - 1;
    while(lo < hi){
        const int mid = (lo + hi + 1) >> 1;
        if(p.q[p.pq[mid]].pch0 <= t) lo = mid; else hi = mid - 1;
    }
    return p.pq[lo];
}
__global__ void __launch_bounds__(1024) k_batch_psum(BPlan p){
    __shared__ unsigned long long s_w[16];
    const BQuery &q = p.q[find_pchunk(p, blockIdx.x)];
    const int64_t ch = (int64_t)blockIdx.x - q.pch0, N = (int64_t)q.nparts * q.ntiles, r0 = ch * PS_CHUNK;
    const int t = threadIdx.x;
    const gptr<const unsigned> pcnt = G((const unsigned*)q.pcnt);
    unsigned long long s = 0;
    unsigned v[PS_PER];
    #pragma unroll
    for(int k = 0; k < PS_PER; k++){ const int64_t i = r0 + (int64_t)k * 1024 + t; v[k] = pcnt[i < N ? i : N - 1]; }
    #pragma unroll
    for(int k = 0; k < PS_PER; k++) s += r0 + (int64_t)k * 1024 + t < N ? v[k] : 0u;
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if((t & 63) == 0) s_w[t >> 6] = s;
    __syncthreads();
    if(t == 0){ unsigned long long a = 0; for(int k = 0; k < 16; k++) a += s_w[k]; q.pcsum[ch] = a; }
}
__global__ void __launch_bounds__(1024) k_batch_pscan(BPlan p){
    __shared__ unsigned long long s_w[16];
    __shared__ unsigned long long s_carry;
    const BQuery &q = p.q[find_pchunk(p, blockIdx.x)];
    const int64_t ch = (int64_t)blockIdx.x - q.pch0, N = (int64_t)q.nparts * q.ntiles, r0 = ch * PS_CHUNK;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if(t == 0){ unsigned long long a = 0; for(int64_t k = 0; k < ch; k++) a += q.pcsum[k]; s_carry = a; }
    const gptr<const unsigned> pcnt = G((const unsigned*)q.pcnt);
    const gptr<int64_t> poff = G(q.poff);
    unsigned v[PS_PER]; unsigned long long s = 0;
    #pragma unroll
    for(int k = 0; k < PS_PER; k++){ const int64_t i = r0 + (int64_t)t * PS_PER + k; v[k] = pcnt[i < N ? i : N - 1]; }
    #pragma unroll
    for(int k = 0; k < PS_PER; k++){ if(r0 + (int64_t)t * PS_PER + k >= N) v[k] = 0u; s += v[k]; }
    unsigned long long inc = s;
    for(int o = 1; o < 64; o <<= 1){ const unsigned long long y = __shfl_up(inc, o, 64); if(lane >= o) inc += y; }
    if(lane == 63) s_w[w] = inc;
    __syncthreads();
    unsigned long long pre = s_carry + inc - s, tot = 0;
    for(int k = 0; k < 16; k++){ if(k < w) pre += s_w[k]; tot += s_w[k]; }
    #pragma unroll
    for(int k = 0; k < PS_PER; k++){ const int64_t i = r0 + (int64_t)t * PS_PER + k; if(i < N) poff[i] = (int64_t)pre; pre += v[k]; }
    if(t == 0 && r0 + PS_CHUNK >= N) poff[N] = (int64_t)(s_carry + tot);
}

// Scatter of a partitioned query's row tile, staged through LDS so that the writes go out as
// runs: the tile is taken 8192 selected-or-not rows at a time (eight per thread); each row gets
// its rank inside its partition for this chunk from an LDS counter, a block scan of the
// partitions' chunk counts places every partition's rows contiguously in LDS, and the block
// then writes the chunk out slot by slot -- consecutive threads to consecutive places of one
// partition's run -- at each partition's running output position (its tile base from the
// (partition, tile) scan plus the rows of earlier chunks).  The next chunk continues the same
// runs, so a partition's bytes reach L2 in order from one workgroup instead of one 2-byte key
// and one 8-byte value per lane to as many lines.  The queue holds the partitioned queries'
// tiles only.  One workgroup per CU: the next chunk's filter and key columns are loaded while
// the current chunk is ranked and written, and its values with its filter.
#define SC_BLOCK 1024
#define SC_ROWS 8
#define SC_CHUNK (SC_ROWS * SC_BLOCK)
__global__ void __launch_bounds__(SC_BLOCK) k_batch_scatter(BPlan p){
    // 146 KB of LDS: one workgroup per CU
    __shared__ int64_t s_item;
    __shared__ int64_t s_t0[QTAB];
    __shared__ uint32_t s_wsum[SC_BLOCK / 64];
    __shared__ int64_t pos0[PART_MAX];                                   // next output place per partition
    __shared__ uint32_t coff[PART_MAX];                                  // chunk counts, then offsets
    __shared__ double st_val[SC_CHUNK];
    __shared__ uint16_t st_key[SC_CHUNK], st_part[SC_CHUNK];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    load_qtab(p, SEL_STILE, s_t0);
    while(true){
        const int64_t item = next_item(&p.ctr[2], &s_item);
        if(item >= p.total_stiles) break;
        const BQuery &q = p.q[find_query_tab(p, s_t0, item, SEL_STILE)];
        if(!q.part) continue;                                            // uniform over the workgroup
        const int64_t tile = item - q.stile0;
        const int64_t r0 = tile * p.tile_rows;
        const int64_t r1 = r0 + p.tile_rows < q.n ? r0 + p.tile_rows : q.n;
        const uint32_t range = (uint32_t)q.range, mask = (uint32_t)p.part_keys - 1u;
        const gptr<const int32_t> flt = G(q.flt), key = G(q.key);
        const gptr<const double> val = G(q.val);
        const gptr<uint16_t> skey = G(q.skey);
        const gptr<double> sval = G(q.sval);
        const gptr<const int64_t> poff = G((const int64_t*)q.poff);
        const int32_t op = q.op, thr = q.thr, kmin = q.kmin;
        const int P = q.nparts;
        for(int k = threadIdx.x; k < P; k += SC_BLOCK) pos0[k] = poff[(int64_t)k * q.ntiles + tile];
        int32_t nf[SC_ROWS], nk[SC_ROWS];
        #pragma unroll
        for(int u = 0; u < SC_ROWS; u++){
            const int64_t i = r0 + (int64_t)u * SC_BLOCK + threadIdx.x, ic = i < r1 ? i : r1 - 1;
            nf[u] = flt[ic]; nk[u] = key[ic];
        }
        for(int64_t c0 = r0; c0 < r1; c0 += SC_CHUNK){
            for(int k = threadIdx.x; k < P; k += SC_BLOCK) coff[k] = 0u;
            uint32_t s[SC_ROWS], rk[SC_ROWS]; double v[SC_ROWS]; bool ok[SC_ROWS];
            #pragma unroll
            for(int u = 0; u < SC_ROWS; u++){
                const int64_t i = c0 + (int64_t)u * SC_BLOCK + threadIdx.x;
                s[u] = (uint32_t)nk[u] - (uint32_t)kmin;
                ok[u] = i < r1 && batch_pass(op, nf[u], thr) && s[u] < range;
                v[u] = val[i < r1 ? i : r1 - 1];
            }
            // the next chunk's filter and keys, in flight during this chunk's LDS work (the
            // last chunk loads its own last row again)
            const int64_t c1 = c0 + SC_CHUNK;
            #pragma unroll
            for(int u = 0; u < SC_ROWS; u++){
                const int64_t i = c1 + (int64_t)u * SC_BLOCK + threadIdx.x, ic = i < r1 ? i : r1 - 1;
                nf[u] = flt[ic]; nk[u] = key[ic];
            }
            __syncthreads();
            #pragma unroll
            for(int u = 0; u < SC_ROWS; u++) rk[u] = ok[u] ? atomicAdd(&coff[s[u] >> p.part_shift], 1u) : 0u;
            __syncthreads();
            // exclusive scan of the P chunk counts: thread t owns partitions [4t, 4t + 4)
            uint32_t c[4], tot = 0;
            #pragma unroll
            for(int j = 0; j < 4; j++){ const int k = 4 * threadIdx.x + j; c[j] = k < P ? coff[k] : 0u; tot += c[j]; }
            uint32_t inc = tot;
            for(int o = 1; o < 64; o <<= 1){ const uint32_t y = __shfl_up(inc, o, 64); if(lane >= o) inc += y; }
            if(lane == 63) s_wsum[w] = inc;
            __syncthreads();
            uint32_t pre = inc - tot, total = 0;
            for(int q2 = 0; q2 < SC_BLOCK / 64; q2++){ if(q2 < w) pre += s_wsum[q2]; total += s_wsum[q2]; }
            #pragma unroll
            for(int j = 0; j < 4; j++){ const int k = 4 * threadIdx.x + j; if(k < P) coff[k] = pre; pre += c[j]; }
            __syncthreads();
            #pragma unroll
            for(int u = 0; u < SC_ROWS; u++){
                if(!ok[u]) continue;
                const uint32_t part = s[u] >> p.part_shift, slot = coff[part] + rk[u];
                st_key[slot] = (uint16_t)(s[u] & mask); st_part[slot] = (uint16_t)part; st_val[slot] = v[u];
            }
            __syncthreads();
            for(uint32_t j = threadIdx.x; j < total; j += SC_BLOCK){
                const uint32_t k = st_part[j];
                const int64_t pos = pos0[k] + (int64_t)(j - coff[k]);
                skey[pos] = st_key[j]; sval[pos] = st_val[j];
            }
            __syncthreads();
            for(int k = threadIdx.x; k < P; k += SC_BLOCK) pos0[k] += (int64_t)((k + 1 < P ? coff[k + 1] : total) - coff[k]);
            __syncthreads();
        }
    }
}

// One slice of a partition per item (a partition of a query with many rows per partition is
// split over psplit workgroups): LDS aggregation of its rows, then its part_keys slots are
// written directly when the partition is whole, or added with compensated atomics when split.
__global__ void k_batch_pagg(BPlan p){
    __shared__ int64_t s_item;
    double *lhi = g_batch_lds, *llo = lhi + p.part_keys;
    unsigned *lcnt = (unsigned*)(llo + p.part_keys);
    __shared__ int64_t s_t0[QTAB];
    load_qtab(p, SEL_PART, s_t0);
    while(true){
        const int64_t item = next_item(&p.ctr[3], &s_item);
        if(item >= p.total_pitems) break;
        const BQuery &q = p.q[find_query_tab(p, s_t0, item, SEL_PART)];
        const int64_t pt = (item - q.pitem0) / q.psplit, sl = (item - q.pitem0) % q.psplit;
        for(int k = threadIdx.x; k < p.part_keys; k += blockDim.x){ lhi[k] = 0.0; llo[k] = 0.0; lcnt[k] = 0u; }
        __syncthreads();
        const gptr<const int64_t> poff = G((const int64_t*)q.poff);
        const gptr<const uint16_t> skey = G((const uint16_t*)q.skey);
        const gptr<const double> sval = G((const double*)q.sval);
        const int64_t pa = poff[pt * q.ntiles], pb = poff[(pt + 1) * q.ntiles];
        const int64_t a = pa + (pb - pa) * sl / q.psplit, b = pa + (pb - pa) * (sl + 1) / q.psplit;
        const int64_t step = (int64_t)blockDim.x;
        if(a < b){
            // eight rows a thread per step; the next step's rows are loaded before this step's adds
            // (clamped to the slice's last row)
            constexpr int PU = 8;
            uint32_t ns[PU]; double nv[PU];
            #pragma unroll
            for(int u = 0; u < PU; u++){ const int64_t i = a + threadIdx.x + u * step, ic = i < b ? i : b - 1; ns[u] = skey[ic]; nv[u] = sval[ic]; }
            for(int64_t i0 = a + threadIdx.x; i0 < b; i0 += PU * step){
                uint32_t s[PU]; double v[PU];
                #pragma unroll
                for(int u = 0; u < PU; u++){ s[u] = ns[u]; v[u] = nv[u]; }
                #pragma unroll
                for(int u = 0; u < PU; u++){ const int64_t i = i0 + (PU + u) * step, ic = i < b ? i : b - 1; ns[u] = skey[ic]; nv[u] = sval[ic]; }
                // the step's returning adds first, then their rounding errors and counts
                double old[PU];
                #pragma unroll
                for(int u = 0; u < PU; u++) if(i0 + u * step < b) old[u] = atomicAdd(&lhi[s[u]], v[u]);
                #pragma unroll
                for(int u = 0; u < PU; u++){
                    if(i0 + u * step >= b) continue;
                    atomicAdd(&llo[s[u]], two_sum_err(old[u], v[u], old[u] + v[u]));
                    atomicAdd(&lcnt[s[u]], 1u);
                }
            }
        }
        __syncthreads();
        const int64_t base = pt * (int64_t)p.part_keys;
        for(int k = threadIdx.x; k < p.part_keys; k += blockDim.x){
            const int64_t s = base + k;
            if(s >= q.range) continue;
            if(q.psplit == 1){ q.hi[s] = lhi[k]; q.lo[s] = llo[k]; q.cnt[s] = lcnt[k]; }
            else if(lcnt[k]){
                const double h = lhi[k];
                double old = unsafeAtomicAdd(&q.hi[s], h);
                unsafeAtomicAdd(&q.lo[s], two_sum_err(old, h, old + h) + llo[k]);
                atomicAdd(&q.cnt[s], lcnt[k]);
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_batch_emit(BPlan p){
    __shared__ int64_t s_item;
    __shared__ uint32_t s_cnt[EMIT_ITEMS * 4];
    __shared__ int64_t s_pre;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    __shared__ int64_t s_t0[QTAB];
    load_qtab(p, SEL_EMIT, s_t0);
    while(true){
        const int64_t item = next_item(&p.ctr[1], &s_item);
        if(item >= p.total_etiles) break;
        const BQuery &q = p.q[find_query_tab(p, s_t0, item, SEL_EMIT)];
        const int64_t tile = item - q.etile0;
        const int64_t base = tile * EMIT_TILE;
        uint64_t bal[EMIT_ITEMS];
        const gptr<const unsigned> cnt = G((const unsigned*)q.cnt);
        unsigned cv[EMIT_ITEMS];
        #pragma unroll
        for(int k = 0; k < EMIT_ITEMS; k++){
            const int64_t e = base + (int64_t)k * 256 + threadIdx.x, ec = e < q.range ? e : q.range - 1;
            int64_t s = ec + q.rot; if(s >= q.range) s -= q.range;
            cv[k] = cnt[s];
        }
        #pragma unroll
        for(int k = 0; k < EMIT_ITEMS; k++){
            int64_t e = base + (int64_t)k * 256 + threadIdx.x;
            bool m = e < q.range && cv[k] != 0u;
            bal[k] = __ballot(m);
            if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
        }
        __syncthreads();
        if(threadIdx.x == 0){
            uint32_t run = 0;
            for(int k = 0; k < EMIT_ITEMS * 4; k++){ uint32_t t = s_cnt[k]; s_cnt[k] = run; run += t; }
            // queue items are dispensed in order, so the predecessor of this tile was taken
            // earlier by a running workgroup and its status always arrives
            uint64_t pre = 0;
            if(item == 0) __hip_atomic_store(&p.estatus[0], ST_PRE | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else {
                __hip_atomic_store(&p.estatus[item], ST_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int64_t j = item - 1;
                while(true){
                    uint64_t st = __hip_atomic_load(&p.estatus[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    while((st >> 62) == 0){
                        __builtin_amdgcn_s_sleep(1);
                        st = __hip_atomic_load(&p.estatus[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    pre += st & ST_MASK;
                    if((st >> 62) == 2) break;
                    j--;
                }
                __hip_atomic_store(&p.estatus[item], ST_PRE | (pre + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_pre = (int64_t)pre;
            if(tile == q.entiles - 1) q.ng[0] = (int64_t)(pre + run);    // end of the query's groups
        }
        __syncthreads();
        const int64_t pre = s_pre;
        #pragma unroll
        for(int k = 0; k < EMIT_ITEMS; k++){
            if((bal[k] >> lane) & 1ull){
                int64_t e = base + (int64_t)k * 256 + threadIdx.x;
                int64_t s = e + q.rot; if(s >= q.range) s -= q.range;
                int64_t pos = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
                p.gkey[pos] = (int32_t)(s + q.kmin); p.gsum[pos] = G((const double*)q.hi)[s] + G((const double*)q.lo)[s]; p.gcnt[pos] = (int64_t)cv[k];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host side
struct hpa_batch {
    hpa_ctx *c = nullptr;
    int nq = 0;
    std::vector<BQuery> hq;
    BQuery *dq = nullptr;
    unsigned *ctr = nullptr;
    int64_t *ngdev = nullptr;
    uint64_t *estatus = nullptr;
    int32_t *gkey = nullptr; double *gsum = nullptr; int64_t *gcnt = nullptr;
    int64_t slots = 0;             // sum of the key ranges: capacity of the output
    int *dpq = nullptr; int npq = 0;
    int64_t total_tiles = 0, total_etiles = 0, total_pitems = 0, total_pchunks = 0, total_stiles = 0;
    int block = 256, tile_rows = 8192, lds_keys = 2730, part_keys = 2048, grid = 0, max_parts = 0;
    hipGraph_t graph = nullptr; hipGraphExec_t exec = nullptr; bool captured = false;
    std::vector<int64_t> ng; bool ng_valid = false;
    double scan_bytes = 0;         // algorithmic bytes of one run: the three input columns, once
    char *pinned = nullptr;        // host staging of every query's results (hpaBatchFetch)
    // pipelined fetch (hpaBatchFetchBegin / End): two snapshot slots in HBM and pinned memory,
    // copied to the host on a stream of their own while the next run computes
    hipStream_t cstream = nullptr;
    hipEvent_t ev_run[2] = {nullptr, nullptr}, ev_copy[2] = {nullptr, nullptr};
    char *dsnap[2] = {nullptr, nullptr}, *hsnap[2] = {nullptr, nullptr};
    int64_t snap_cap = 0;          // groups a slot holds
    int64_t gknown = -1;           // groups of the last run whose count was read
    int64_t runs = 0, begun = 0, ended = 0, slot_run[2] = {-1, -1};
};

static BPlan plan_of(hpa_batch *b){
    BPlan p;
    p.q = b->dq; p.nq = b->nq; p.total_tiles = b->total_tiles; p.total_etiles = b->total_etiles;
    p.total_pitems = b->total_pitems; p.total_stiles = b->total_stiles;
    p.ctr = b->ctr; p.tile_rows = b->tile_rows; p.lds_keys = b->lds_keys;
    p.estatus = b->estatus; p.gkey = b->gkey; p.gsum = b->gsum; p.gcnt = b->gcnt;
    p.part_keys = b->part_keys; p.part_shift = 0;
    while(b->part_keys > 0 && (1 << p.part_shift) < b->part_keys) p.part_shift++;
    p.pq = b->dpq; p.npq = b->npq; p.total_pchunks = b->total_pchunks;
    return p;
}

static size_t lds_bytes(int keys){ return (size_t)keys * (2 * sizeof(double) + sizeof(unsigned)); }

static void drop_graph(hpa_batch *b){
    if(b->captured){ hipGraphExecDestroy(b->exec); hipGraphDestroy(b->graph); }
    b->captured = false; b->exec = nullptr; b->graph = nullptr;
}

// Tile and partition bookkeeping for the tunables; uploads the descriptors.
static int32_t batch_layout(hpa_batch *b, int block, int tile_rows, int lds_keys, int part_keys){
    if(block != 64 && block != 128 && block != 256 && block != 512 && block != 1024){ set_error_detail("batch block %d", block); R HPA_E_DOMAIN; }
    if(tile_rows < 256 || lds_keys < 0 || lds_bytes(lds_keys) > 64 * 1024){ set_error_detail("batch tile %d lds %d", tile_rows, lds_keys); R HPA_E_DOMAIN; }
    if(part_keys != 0 && (part_keys < PART_MIN_KEYS || part_keys > 2048 || (part_keys & (part_keys - 1)))){
        set_error_detail("batch partition width %d (0, or a power of two in [%d, 2048])", part_keys, PART_MIN_KEYS); R HPA_E_DOMAIN;
    }
    if(part_keys != 0 && tile_rows < PART_MIN_TILE){
        set_error_detail("batch tile %d: partitioning needs tiles of at least %d rows", tile_rows, PART_MIN_TILE); R HPA_E_DOMAIN;
    }
    b->block = block; b->tile_rows = tile_rows; b->lds_keys = lds_keys; b->part_keys = part_keys;
    int64_t t = 0, pi = 0, pc = 0, st = 0;
    std::vector<int> pq;
    b->max_parts = 0;
    for(int qi = 0; qi < b->nq; qi++){
        BQuery &q = b->hq[qi];
        q.ntiles = ceil_div(q.n, tile_rows); q.tile0 = t; t += q.ntiles;
        const int64_t P = part_keys ? ceil_div(q.range, part_keys) : 0;
        // the count matrix was sized for the narrowest partitions and the smallest tiles
        q.part = (part_keys && q.range > lds_keys && q.skey && P <= PART_MAX) ? 1 : 0;
        q.nparts = q.part ? (int32_t)P : 0;
        // about 64K input rows per aggregating workgroup, at most 64 workgroups per partition
        int64_t sp = q.part ? ceil_div(ceil_div(q.n, P), 65536) : 1;
        q.psplit = (int32_t)(sp < 1 ? 1 : sp > 64 ? 64 : sp);
        q.pitem0 = pi; pi += (int64_t)q.nparts * q.psplit;
        q.pch0 = pc;
        q.stile0 = st;
        if(q.part){
            st += q.ntiles;
            pq.push_back(qi); b->max_parts = q.nparts > b->max_parts ? q.nparts : b->max_parts;
            pc += ceil_div((int64_t)q.nparts * q.ntiles, PS_CHUNK);
        }
    }
    b->total_tiles = t; b->total_pitems = pi; b->total_pchunks = pc; b->total_stiles = st;
    b->npq = (int)pq.size();
    b->grid = num_cus() * (2048 / block);
    drop_graph(b);
    if(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * b->nq, hipMemcpyHostToDevice, b->c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(b->npq && hipMemcpyAsync(b->dpq, pq.data(), sizeof(int) * b->npq, hipMemcpyHostToDevice, b->c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamSynchronize(b->c->stream) != hipSuccess) R HPA_E_DEVICE;
    R 0;
}

static int32_t batch_enqueue(hpa_batch *b, hipStream_t s){
    BPlan p = plan_of(b);
    int64_t maxr = 1; for(auto &q : b->hq) maxr = q.range > maxr ? q.range : maxr;
    int64_t gx = ceil_div(maxr, 256); gx = gx > 256 ? 256 : gx;
    hipLaunchKernelGGL(k_batch_clear, dim3((unsigned)gx, (unsigned)(b->nq < 64 ? b->nq : 64)), dim3(256), 0, s, p);
    size_t agg_lds = lds_bytes(b->lds_keys);
    if((size_t)b->max_parts * 4 > agg_lds) agg_lds = (size_t)b->max_parts * 4;
    hipLaunchKernelGGL(k_batch_agg, dim3(b->grid), dim3(b->block), agg_lds, s, p);
    if(b->npq){
        hipLaunchKernelGGL(k_batch_psum, dim3((unsigned)b->total_pchunks), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(k_batch_pscan, dim3((unsigned)b->total_pchunks), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(k_batch_scatter, dim3(num_cus()), dim3(SC_BLOCK), 0, s, p);
        hipLaunchKernelGGL(k_batch_pagg, dim3(num_cus() * 4), dim3(256), lds_bytes(b->part_keys), s, p);
    }
    hipLaunchKernelGGL(k_batch_emit, dim3(num_cus() * 8), dim3(256), 0, s, p);
    R check_launch("batch");
}

extern "C" void hpaBatchDestroy(hpa_batch *b){
    if(!b) R;
    drop_graph(b);
    if(b->cstream) hipStreamSynchronize(b->cstream);
    hipStreamSynchronize(b->c->stream);
    for(int k = 0; k < 2; k++){
        if(b->dsnap[k]) hipFree(b->dsnap[k]);
        if(b->hsnap[k]) hipHostFree(b->hsnap[k]);
        if(b->ev_run[k]) hipEventDestroy(b->ev_run[k]);
        if(b->ev_copy[k]) hipEventDestroy(b->ev_copy[k]);
    }
    if(b->cstream) hipStreamDestroy(b->cstream);
    if(b->pinned) hipHostFree(b->pinned);
    delete b;
}

extern "C" hpa_batch *hpaBatchCreate(int32_t nq, hpa_V *keys, hpa_V *vals, hpa_V *flts,
                                     const int32_t *ops, const int32_t *consts){
    hpa_ctx *c = ctx_cur();
    if(nq <= 0){ set_error_detail("batch of %d queries", nq); R nullptr; }
    hpa_batch *b = new hpa_batch();
    b->c = c; b->nq = nq;
    b->hq.resize(nq);
    auto fail = [&](const char *why, int qi) -> hpa_batch* { set_error_detail("batch query %d: %s", qi, why); delete b; R nullptr; };
    for(int qi = 0; qi < nq; qi++){
        BQuery &q = b->hq[qi];
        memset(&q, 0, sizeof(q));
        V k = keys[qi], v = vals[qi], f = flts[qi];
        if(k->typ != HPA_I || f->typ != HPA_I || v->typ != HPA_E) R fail("key/filter columns must be i32, values f64", qi);
        if(k->len != v->len || k->len != f->len) R fail("column lengths differ", qi);
        if(ops[qi] < HPA_OP_LT || ops[qi] > HPA_OP_NEQ) R fail("filter op", qi);
        const void *pk, *pv, *pf;
        if(dev_view(c, k, &pk) || dev_view(c, v, &pv) || dev_view(c, f, &pf)) R fail("arena full", qi);
        q.key = (const int32_t*)pk; q.val = (const double*)pv; q.flt = (const int32_t*)pf;
        q.n = k->len; q.op = ops[qi]; q.thr = consts[qi];
        b->scan_bytes += 16.0 * (double)q.n;
    }
    // key ranges: one min/max pass over every query's keys
    b->dq = (BQuery*)dalloc(c, sizeof(BQuery) * nq);
    int32_t *mm = (int32_t*)dalloc(c, 8 * (L)nq);
    b->dpq = (int*)dalloc(c, sizeof(int) * nq);
    if(!b->dq || !mm || !b->dpq) R fail("arena full", 0);
    std::vector<int32_t> hmm(2 * nq);
    for(int qi = 0; qi < nq; qi++){ hmm[2 * qi] = INT32_MAX; hmm[2 * qi + 1] = INT32_MIN; }
    hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * nq, hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(mm, hmm.data(), 8 * (L)nq, hipMemcpyHostToDevice, c->stream);
    hipLaunchKernelGGL(k_batch_minmax, dim3(64, (unsigned)(nq < 1024 ? nq : 1024)), dim3(256), 0, c->stream, b->dq, nq, mm);
    if(check_launch("batch_minmax") || d2h(c, hmm.data(), mm, 8 * (L)nq)) R fail("device", 0);
    int64_t et = 0;
    for(int qi = 0; qi < nq; qi++){
        BQuery &q = b->hq[qi];
        int32_t lo = hmm[2 * qi], hi = hmm[2 * qi + 1];
        if(q.n == 0){ lo = 0; hi = 0; }
        int64_t range = (int64_t)hi - lo + 1;
        if(range > BATCH_MAX_RANGE) R fail("key range exceeds the dense slot budget", qi);
        q.kmin = lo; q.range = (int32_t)range;
        // unsigned key order: non-negative keys first, then the negative ones
        q.rot = (lo < 0 && hi >= 0) ? -lo : 0;
        q.hi = (double*)dalloc(c, 8 * range); q.lo = (double*)dalloc(c, 8 * range);
        q.cnt = (unsigned*)dalloc(c, 4 * range);
        q.entiles = ceil_div(range, EMIT_TILE);
        b->slots += range;
        if(!q.hi || !q.lo || !q.cnt) R fail("arena full", qi);
        q.etile0 = et; et += q.entiles;
        // partition buffers for any partition width and tile size the tunables allow
        if(ceil_div(range, PART_MIN_KEYS) <= PART_MAX && range > 256 && q.n > 0){
            const int64_t M = ceil_div(range, PART_MIN_KEYS) * ceil_div(q.n, PART_MIN_TILE);
            q.pcnt = (unsigned*)dalloc(c, 4 * M); q.poff = (int64_t*)dalloc(c, 8 * (M + 1));
            q.skey = (uint16_t*)dalloc(c, 2 * q.n); q.sval = (double*)dalloc(c, 8 * q.n);
            q.pcsum = (unsigned long long*)dalloc(c, 8 * ceil_div(M, PS_CHUNK));
            if(!q.pcnt || !q.poff || !q.skey || !q.sval || !q.pcsum) R fail("arena full", qi);
        }
    }
    b->total_etiles = et;
    b->estatus = (uint64_t*)dalloc(c, 8 * et);
    b->gkey = (int32_t*)dalloc(c, 4 * b->slots); b->gsum = (double*)dalloc(c, 8 * b->slots);
    b->gcnt = (int64_t*)dalloc(c, 8 * b->slots);
    if(!b->estatus || !b->gkey || !b->gsum || !b->gcnt) R fail("arena full", 0);
    b->ngdev = (int64_t*)dalloc(c, 16 * (L)nq);
    b->ctr = (unsigned*)dalloc(c, 4 * BATCH_CTRS);
    if(!b->ngdev || !b->ctr) R fail("arena full", 0);
    for(int qi = 0; qi < nq; qi++) b->hq[qi].ng = b->ngdev + 2 * qi;
    b->ng.assign(2 * (size_t)nq, 0);
    if(batch_layout(b, 256, 8192, 2730, 2048)){ delete b; R nullptr; }
    R b;
}

extern "C" int32_t hpaBatchSetConfig(hpa_batch *b, int32_t block, int32_t tile_rows, int32_t lds_keys, int32_t part_keys){
    R batch_layout(b, block, tile_rows, lds_keys, part_keys);
}

extern "C" int32_t hpaBatchGetConfig(hpa_batch *b, int32_t *block, int32_t *tile_rows, int32_t *lds_keys, int32_t *part_keys){
    *block = b->block; *tile_rows = b->tile_rows; *lds_keys = b->lds_keys; *part_keys = b->part_keys;
    R 0;
}

extern "C" int32_t hpaBatchRun(hpa_batch *b, int32_t use_graph){
    hpa_ctx *c = b->c;
    int32_t e;
    b->ng_valid = false;
    b->runs++;
    if(use_graph){
        if(!b->captured){
            if(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) R HPA_E_DEVICE;
            e = batch_enqueue(b, c->stream);
            hipGraph_t g = nullptr;
            if(hipStreamEndCapture(c->stream, &g) != hipSuccess || e){ if(g) hipGraphDestroy(g); R e ? e : HPA_E_DEVICE; }
            if(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess){ hipGraphDestroy(g); R HPA_E_DEVICE; }
            b->graph = g; b->captured = true;
        }
        ProfScope ps("batch_graph", c->stream, b->scan_bytes);
        if(hipGraphLaunch(b->exec, c->stream) != hipSuccess) R HPA_E_DEVICE;
        R 0;
    }
    ProfScope ps("batch", c->stream, b->scan_bytes);
    R batch_enqueue(b, c->stream);
}

static int32_t batch_ends(hpa_batch *b){
    if(b->ng_valid) R 0;
    int32_t e = d2h(b->c, b->ng.data(), b->ngdev, 16 * (L)b->nq);
    if(e) R e;
    b->ng_valid = true;
    for(int qi = 0; qi < b->nq; qi++)
        if(b->ng[2 * qi + 1]){ set_error_detail("batch query %d: key outside the planned range", qi); R HPA_E_DOMAIN; }
    R 0;
}
// groups of query qi in the output: [start, end)
static void batch_span(hpa_batch *b, int qi, int64_t *start, int64_t *end){
    *start = qi > 0 ? b->ng[2 * (qi - 1)] : 0; *end = b->ng[2 * qi];
}

// Result of query qi of the last run: zkey (I), zcnt (L), zsum (E) over the groups in the
// pfnGroup order.  The vectors point into the plan's output (valid until the next run).
extern "C" int32_t hpaBatchResult(hpa_batch *b, int32_t qi, hpa_V zkey, hpa_V zcnt, hpa_V zsum){
    hpa_ctx *c = b->c;
    if(qi < 0 || qi >= b->nq) R HPA_E_INDEX;
    int32_t e;
    if((e = batch_ends(b))) R e;
    int64_t a, z; batch_span(b, qi, &a, &z);
    const int64_t g = z - a;
    auto setv = [&](V v, int32_t typ, void *p, L w) -> int32_t {
        v->typ = typ; v->len = g;
        if(g <= 1){ v->g = (char*)&v->b; v->i64 = 0; R g == 1 ? d2h(c, &v->b, p, w) : 0; }
        v->g = (char*)p; R 0;
    };
    if((e = setv(zkey, HPA_I, b->gkey + a, 4)) || (e = setv(zcnt, HPA_L, b->gcnt + a, 8)) || (e = setv(zsum, HPA_E, b->gsum + a, 8))) R e;
    R 0;
}

// All queries' results of the last run, copied into pinned host memory owned by the plan: one
// read of the query ends, three copies (keys, counts, sums of every query, query after query)
// and one synchronisation.  For query q: ng[q] groups; keys[q] (i32), cnts[q] (i64) and
// sums[q] (f64) point into the staging, valid until the next fetch or the plan's destruction.
extern "C" int32_t hpaBatchFetch(hpa_batch *b, int64_t *ng, void **keys, void **cnts, void **sums){
    hpa_ctx *c = b->c;
    int32_t e;
    if(!b->pinned){
        if(hipHostMalloc((void**)&b->pinned, (size_t)(20 * (b->slots > 0 ? b->slots : 1)), hipHostMallocDefault) != hipSuccess){
            b->pinned = nullptr; set_error_detail("batch: pinned staging of %lld bytes", (long long)(20 * b->slots)); R HPA_E_MEMORY;
        }
    }
    if((e = batch_ends(b))) R e;
    const int64_t G = b->ng[2 * (b->nq - 1)];
    int32_t *hk = (int32_t*)(b->pinned + 16 * b->slots);
    int64_t *hc = (int64_t*)b->pinned;
    double *hs = (double*)(b->pinned + 8 * b->slots);
    if(G > 0){
        if(hipMemcpyAsync(hk, b->gkey, 4 * G, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
           hipMemcpyAsync(hc, b->gcnt, 8 * G, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
           hipMemcpyAsync(hs, b->gsum, 8 * G, hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
    }
    for(int qi = 0; qi < b->nq; qi++){
        int64_t a, z; batch_span(b, qi, &a, &z);
        ng[qi] = z - a; keys[qi] = hk + a; cnts[qi] = hc + a; sums[qi] = hs + a;
    }
    R hipStreamSynchronize(c->stream) == hipSuccess ? 0 : HPA_E_DEVICE;
}

// Pipelined fetch.  Begin: on the plan's stream, a snapshot of the last run's group counts and
// results (as many groups as the last known count; the first call reads that count) into slot
// `begun & 1` in HBM (a device copy), then on the copy stream the snapshot's copy into the
// slot's pinned buffer; returns at once, so the next hpaBatchRun computes while the results
// travel.  End: waits for the oldest begun slot and hands out its results as hpaBatchFetch does
// (valid until that slot is begun again, two begins later).  A run with more groups than its
// snapshot took (its tables changed since the count was read) fails with HPA_E_LENGTH.
static int32_t ensure_snap(hpa_batch *b, int64_t cap){
    if(!b->cstream){
        if(hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking) != hipSuccess) R HPA_E_DEVICE;
        for(int k = 0; k < 2; k++)
            if(hipEventCreateWithFlags(&b->ev_run[k], hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&b->ev_copy[k], hipEventDisableTiming) != hipSuccess) R HPA_E_DEVICE;
    }
    if(cap <= b->snap_cap) R 0;
    hipStreamSynchronize(b->cstream); hipStreamSynchronize(b->c->stream);
    for(int k = 0; k < 2; k++){
        if(b->dsnap[k]){ hipFree(b->dsnap[k]); b->dsnap[k] = nullptr; }
        if(b->hsnap[k]){ hipHostFree(b->hsnap[k]); b->hsnap[k] = nullptr; }
    }
    const size_t bytes = (size_t)(16 * b->nq + 20 * (cap > 0 ? cap : 1) + 64);
    for(int k = 0; k < 2; k++){
        if(hipMalloc((void**)&b->dsnap[k], bytes) != hipSuccess){ set_error_detail("batch snapshot of %zu bytes", bytes); R HPA_E_MEMORY; }
        if(hipHostMalloc((void**)&b->hsnap[k], bytes, hipHostMallocDefault) != hipSuccess){ set_error_detail("batch pinned snapshot"); R HPA_E_MEMORY; }
    }
    b->snap_cap = cap;
    R 0;
}
// slot layout: ng (16 nq bytes), then keys (4 cap), counts (8 cap), sums (8 cap)
static void snap_parts(hpa_batch *b, char *base, int64_t **ng, int32_t **k, int64_t **c, double **s){
    *ng = (int64_t*)base;
    char *p = base + 16 * b->nq;
    *c = (int64_t*)p; *s = (double*)(p + 8 * b->snap_cap); *k = (int32_t*)(p + 16 * b->snap_cap);
}
extern "C" int32_t hpaBatchFetchBegin(hpa_batch *b){
    hpa_ctx *c = b->c;
    int32_t e;
    if(b->begun - b->ended >= 2){ set_error_detail("batch: two fetches begun and not ended"); R HPA_E_DOMAIN; }
    if(b->gknown < 0){
        if((e = batch_ends(b))) R e;
        b->gknown = b->ng[2 * (b->nq - 1)];
    }
    const int64_t cap = b->gknown;
    if((e = ensure_snap(b, cap))) R e;
    const int slot = (int)(b->begun & 1);
    // the slot's previous copy to the host has finished before the snapshot overwrites it
    if(b->slot_run[slot] >= 0 && hipStreamWaitEvent(c->stream, b->ev_copy[slot], 0) != hipSuccess) R HPA_E_DEVICE;
    int64_t *ng; int32_t *k; int64_t *cn; double *sm;
    snap_parts(b, b->dsnap[slot], &ng, &k, &cn, &sm);
    if(hipMemcpyAsync(ng, b->ngdev, (size_t)(16 * b->nq), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(cap > 0){
        if(hipMemcpyAsync(k, b->gkey, (size_t)(4 * cap), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
           hipMemcpyAsync(cn, b->gcnt, (size_t)(8 * cap), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
           hipMemcpyAsync(sm, b->gsum, (size_t)(8 * cap), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
    }
    if(hipEventRecord(b->ev_run[slot], c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamWaitEvent(b->cstream, b->ev_run[slot], 0) != hipSuccess) R HPA_E_DEVICE;
    const size_t bytes = (size_t)(16 * b->nq + 20 * cap);
    if(hipMemcpyAsync(b->hsnap[slot], b->dsnap[slot], bytes, hipMemcpyDeviceToHost, b->cstream) != hipSuccess) R HPA_E_DEVICE;
    if(hipEventRecord(b->ev_copy[slot], b->cstream) != hipSuccess) R HPA_E_DEVICE;
    b->slot_run[slot] = b->runs;
    b->begun++;
    R 0;
}
extern "C" int32_t hpaBatchFetchEnd(hpa_batch *b, int64_t *ng, void **keys, void **cnts, void **sums){
    if(b->ended >= b->begun){ set_error_detail("batch: no fetch begun"); R HPA_E_DOMAIN; }
    const int slot = (int)(b->ended & 1);
    if(hipEventSynchronize(b->ev_copy[slot]) != hipSuccess) R HPA_E_DEVICE;
    b->ended++;
    int64_t *hng; int32_t *hk; int64_t *hc; double *hs;
    snap_parts(b, b->hsnap[slot], &hng, &hk, &hc, &hs);
    for(int qi = 0; qi < b->nq; qi++)
        if(hng[2 * qi + 1]){ set_error_detail("batch query %d: key outside the planned range", qi); R HPA_E_DOMAIN; }
    const int64_t G = hng[2 * (b->nq - 1)];
    if(G > b->snap_cap){
        b->gknown = G;
        set_error_detail("batch: the run made %lld groups, its snapshot took %lld", (long long)G, (long long)b->snap_cap);
        R HPA_E_LENGTH;
    }
    b->gknown = G;
    for(int qi = 0; qi < b->nq; qi++){
        const int64_t a = qi > 0 ? hng[2 * (qi - 1)] : 0, z = hng[2 * qi];
        ng[qi] = z - a; keys[qi] = hk + a; cnts[qi] = hc + a; sums[qi] = hs + a;
    }
    R 0;
}

// Autotuner: time each candidate (block size, tile rows, LDS key budget, partition width) over
// `iters` runs of this plan and keep the fastest.  Writes the timings of all candidates to `ms`
// (capacity `cap`, may be null) in candidate order; returns their number through *ncand.
extern "C" int32_t hpaBatchAutotune(hpa_batch *b, int32_t iters, float *ms, int32_t cap, int32_t *ncand){
    static const int blocks[] = {256, 512, 1024};
    static const int tiles[] = {8192, 32768};
    static const int modes[][2] = {{2730, 0}, {2730, 2048}, {1024, 1024}, {1024, 2048}};
    hpa_ctx *c = b->c;
    hipEvent_t a, z;
    hipEventCreate(&a); hipEventCreate(&z);
    float best = 1e30f; int bb = b->block, bt = b->tile_rows, bl = b->lds_keys, bp = b->part_keys, k = 0;
    int32_t e = 0;
    for(int bi : blocks) for(int ti : tiles) for(auto &md : modes){
        if((e = batch_layout(b, bi, ti, md[0], md[1]))) goto done;
        if((e = batch_enqueue(b, c->stream))) goto done;           // warm
        hipEventRecord(a, c->stream);
        for(int it = 0; it < iters; it++) if((e = batch_enqueue(b, c->stream))) goto done;
        hipEventRecord(z, c->stream);
        if(hipEventSynchronize(z) != hipSuccess){ e = HPA_E_DEVICE; goto done; }
        float t = 0; hipEventElapsedTime(&t, a, z); t /= iters > 0 ? iters : 1;
        if(ms && k < cap) ms[k] = t;
        k++;
        if(t < best){ best = t; bb = bi; bt = ti; bl = md[0]; bp = md[1]; }
    }
done:
    hipEventDestroy(a); hipEventDestroy(z);
    if(ncand) *ncand = k;
    int32_t e2 = batch_layout(b, bb, bt, bl, bp);
    b->ng_valid = false;
    R e ? e : e2;
}
