// Batched execution: many independent filter -> compress -> group -> sum queries in one
// persistent launch.
//
// Query q reads its own table (i32 key, f64 value, i32 filter column) and computes, over the
// rows with `flt <op> c`, per distinct key: the key, the row count and the FP64 sum of the
// values, groups in the order pfnGroup gives i32 keys (ascending as unsigned 32-bit values,
// h_libs.cc:1295-1329).  That is the result of the per-query primitive sequence
//   compare -> compress(key), compress(val) -> group -> keys/values -> index(key, keys),
//   raze(each(sum, each_right(index, val, groups))), raze(each(len, ...))
// (the FP1/FP7 patterns of optimizer/pattern.cc:180-219 over a q3.hir-style group-by).
//
// Plan, built once and kept resident (its kernels are captured once into a hipGraph and
// replayed):
//   k_batch_clear  zero the per-query dense slot arrays (a fill kernel: no memset node in the
//                  graph)
//   k_batch_agg    persistent workgroups take (query, row tile) items from an atomic queue.
//                  Queries whose key range fits the LDS budget aggregate a tile in LDS (one
//                  replica per wave when the range is small, against same-address contention)
//                  and flush one global add per touched key; larger ranges add straight into
//                  the query's slots in HBM.
//   k_batch_emit   persistent workgroups take (query, slot tile) items and compact the
//                  non-empty slots in key order (per-query decoupled look-back).
// Slots are dense over the key range [kmin, kmax] measured when the plan is built.
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

struct BQuery {
    const int32_t *key; const double *val; const int32_t *flt;
    int64_t n; int32_t op; int32_t thr;
    int32_t kmin; int32_t range; int32_t rot; int32_t _pad;  // slot = key - kmin; emit starts at rot
    double *hi; double *lo; unsigned int *cnt;               // dense slots [range]
    int32_t *okey; double *osum; int64_t *ocnt;              // compacted outputs (capacity range)
    int64_t *ng;                                             // [0] groups, [1] keys out of range
    uint64_t *status;                                        // emit look-back, per slot tile
    int64_t tile0, ntiles, etile0, entiles;
};

struct BPlan {
    BQuery *q; int nq;
    int64_t total_tiles, total_etiles;
    unsigned *ctr;
    int tile_rows, lds_keys;
};

__device__ __forceinline__ int find_query(const BQuery *q, int nq, int64_t t, bool emit){
    int lo = 0, hi = nq - 1;
    while(lo < hi){
        int mid = (lo + hi + 1) >> 1;
        int64_t s = emit ? q[mid].etile0 : q[mid].tile0;
        if(s <= t) lo = mid; else hi = mid - 1;
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
        for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q.range; i += stride){
            q.hi[i] = 0.0; q.lo[i] = 0.0; q.cnt[i] = 0u;
        }
        for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q.entiles; i += stride) q.status[i] = 0;
        if(blockIdx.x == 0 && threadIdx.x < 2) q.ng[threadIdx.x] = 0;
    }
    if(blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2) p.ctr[threadIdx.x] = 0;
}

extern __shared__ double g_batch_lds[];

__global__ void k_batch_agg(BPlan p){
    __shared__ int64_t s_item;
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
    while(true){
        if(threadIdx.x == 0) s_item = (int64_t)atomicAdd(&p.ctr[0], 1u);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();
        if(item >= p.total_tiles) break;
        const BQuery &q = p.q[find_query(p.q, p.nq, item, false)];
        const int64_t r0 = (item - q.tile0) * p.tile_rows;
        const int64_t r1 = r0 + p.tile_rows < q.n ? r0 + p.tile_rows : q.n;
        const int32_t kmin = q.kmin, op = q.op, thr = q.thr;
        const uint32_t range = (uint32_t)q.range;
        bool bad = false;
        if(q.range <= p.lds_keys){
            int reps = p.lds_keys / q.range; reps = reps > nw ? nw : reps;
            const int tot = reps * q.range;
            double *lhi = g_batch_lds, *llo = lhi + p.lds_keys;
            unsigned *lcnt = (unsigned*)(llo + p.lds_keys);
            for(int k = threadIdx.x; k < tot; k += blockDim.x){ lhi[k] = 0.0; llo[k] = 0.0; lcnt[k] = 0u; }
            __syncthreads();
            const int base = (w % reps) * q.range;
            for(int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x){
                if(batch_pass(op, q.flt[i], thr)){
                    uint32_t s = (uint32_t)q.key[i] - (uint32_t)kmin;
                    if(s < range){
                        double v = q.val[i];
                        double old = atomicAdd(&lhi[base + s], v);
                        atomicAdd(&llo[base + s], two_sum_err(old, v, old + v));
                        atomicAdd(&lcnt[base + s], 1u);
                    } else bad = true;
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
            for(int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x){
                if(batch_pass(op, q.flt[i], thr)){
                    uint32_t s = (uint32_t)q.key[i] - (uint32_t)kmin;
                    if(s < range){
                        double v = q.val[i];
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

__global__ void __launch_bounds__(256) k_batch_emit(BPlan p){
    __shared__ int64_t s_item;
    __shared__ uint32_t s_cnt[EMIT_ITEMS * 4];
    __shared__ int64_t s_pre;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    while(true){
        if(threadIdx.x == 0) s_item = (int64_t)atomicAdd(&p.ctr[1], 1u);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();
        if(item >= p.total_etiles) break;
        const BQuery &q = p.q[find_query(p.q, p.nq, item, true)];
        const int64_t tile = item - q.etile0;
        const int64_t base = tile * EMIT_TILE;
        uint64_t bal[EMIT_ITEMS];
        #pragma unroll
        for(int k = 0; k < EMIT_ITEMS; k++){
            int64_t e = base + (int64_t)k * 256 + threadIdx.x;
            int64_t s = e + q.rot; if(s >= q.range) s -= q.range;
            bool m = e < q.range && q.cnt[s] != 0u;
            bal[k] = __ballot(m);
            if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
        }
        __syncthreads();
        if(threadIdx.x == 0){
            uint32_t run = 0;
            for(int k = 0; k < EMIT_ITEMS * 4; k++){ uint32_t t = s_cnt[k]; s_cnt[k] = run; run += t; }
            // tiles of one query are consecutive queue items, so the predecessor of this tile
            // was taken earlier by a running workgroup and its status always arrives
            uint64_t pre = 0;
            if(tile == 0) __hip_atomic_store(&q.status[0], ST_PRE | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else {
                __hip_atomic_store(&q.status[tile], ST_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int64_t j = tile - 1;
                while(true){
                    uint64_t st = __hip_atomic_load(&q.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    while((st >> 62) == 0){
                        __builtin_amdgcn_s_sleep(1);
                        st = __hip_atomic_load(&q.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    pre += st & ST_MASK;
                    if((st >> 62) == 2) break;
                    j--;
                }
                __hip_atomic_store(&q.status[tile], ST_PRE | (pre + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_pre = (int64_t)pre;
            if(tile == q.entiles - 1) q.ng[0] = (int64_t)(pre + run);
        }
        __syncthreads();
        const int64_t pre = s_pre;
        #pragma unroll
        for(int k = 0; k < EMIT_ITEMS; k++){
            if((bal[k] >> lane) & 1ull){
                int64_t e = base + (int64_t)k * 256 + threadIdx.x;
                int64_t s = e + q.rot; if(s >= q.range) s -= q.range;
                int64_t pos = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
                q.okey[pos] = (int32_t)(s + q.kmin); q.osum[pos] = q.hi[s] + q.lo[s]; q.ocnt[pos] = (int64_t)q.cnt[s];
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
    int64_t total_tiles = 0, total_etiles = 0;
    int block = 256, tile_rows = 8192, lds_keys = 2048, grid = 0;
    hipGraph_t graph = nullptr; hipGraphExec_t exec = nullptr; bool captured = false;
    std::vector<int64_t> ng; bool ng_valid = false;
    double scan_bytes = 0;         // algorithmic bytes of one run: the three input columns, once
    char *pinned = nullptr;        // host staging of every query's results (hpaBatchFetch)
    std::vector<int64_t> pin_off;  // per query: offset of its (keys, counts, sums) block
};

static BPlan plan_of(hpa_batch *b){
    BPlan p; p.q = b->dq; p.nq = b->nq; p.total_tiles = b->total_tiles; p.total_etiles = b->total_etiles;
    p.ctr = b->ctr; p.tile_rows = b->tile_rows; p.lds_keys = b->lds_keys;
    return p;
}

static size_t lds_bytes(int lds_keys){ return (size_t)lds_keys * (2 * sizeof(double) + sizeof(unsigned)); }

static void drop_graph(hpa_batch *b){
    if(b->captured){ hipGraphExecDestroy(b->exec); hipGraphDestroy(b->graph); }
    b->captured = false; b->exec = nullptr; b->graph = nullptr;
}

// Tile bookkeeping for the tunables; uploads the descriptors.
static int32_t batch_layout(hpa_batch *b, int block, int tile_rows, int lds_keys){
    if(block != 64 && block != 128 && block != 256 && block != 512 && block != 1024){ set_error_detail("batch block %d", block); R HPA_E_DOMAIN; }
    if(tile_rows < 256 || lds_keys < 0 || lds_bytes(lds_keys) > 64 * 1024){ set_error_detail("batch tile %d lds %d", tile_rows, lds_keys); R HPA_E_DOMAIN; }
    b->block = block; b->tile_rows = tile_rows; b->lds_keys = lds_keys;
    int64_t t = 0;
    for(auto &q : b->hq){ q.ntiles = ceil_div(q.n, tile_rows); q.tile0 = t; t += q.ntiles; }
    b->total_tiles = t;
    b->grid = num_cus() * (2048 / block);
    drop_graph(b);
    if(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * b->nq, hipMemcpyHostToDevice, b->c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamSynchronize(b->c->stream) != hipSuccess) R HPA_E_DEVICE;
    R 0;
}

static int32_t batch_enqueue(hpa_batch *b, hipStream_t s){
    BPlan p = plan_of(b);
    int64_t maxr = 1; for(auto &q : b->hq) maxr = q.range > maxr ? q.range : maxr;
    int64_t gx = ceil_div(maxr, 256); gx = gx > 256 ? 256 : gx;
    hipLaunchKernelGGL(k_batch_clear, dim3((unsigned)gx, (unsigned)(b->nq < 64 ? b->nq : 64)), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_batch_agg, dim3(b->grid), dim3(b->block), lds_bytes(b->lds_keys), s, p);
    hipLaunchKernelGGL(k_batch_emit, dim3(num_cus() * 8), dim3(256), 0, s, p);
    R check_launch("batch");
}

extern "C" void hpaBatchDestroy(hpa_batch *b){
    if(!b) R;
    drop_graph(b);
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
    if(!b->dq || !mm) R fail("arena full", 0);
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
        q.okey = (int32_t*)dalloc(c, 4 * range); q.osum = (double*)dalloc(c, 8 * range);
        q.ocnt = (int64_t*)dalloc(c, 8 * range);
        q.entiles = ceil_div(range, EMIT_TILE);
        q.status = (uint64_t*)dalloc(c, 8 * q.entiles);
        if(!q.hi || !q.lo || !q.cnt || !q.okey || !q.osum || !q.ocnt || !q.status) R fail("arena full", qi);
        q.etile0 = et; et += q.entiles;
    }
    b->total_etiles = et;
    b->ngdev = (int64_t*)dalloc(c, 16 * (L)nq);
    b->ctr = (unsigned*)dalloc(c, 16);
    if(!b->ngdev || !b->ctr) R fail("arena full", 0);
    for(int qi = 0; qi < nq; qi++) b->hq[qi].ng = b->ngdev + 2 * qi;
    b->ng.assign(2 * (size_t)nq, 0);
    if(batch_layout(b, 256, 8192, 2048)){ delete b; R nullptr; }
    R b;
}

// All queries' results of the last run, copied into pinned host memory owned by the plan with
// one read of the group counts, one copy per output vector and one synchronisation.  For query
// q: ng[q] groups; keys[q] (i32), cnts[q] (i64) and sums[q] (f64) point into the staging, valid
// until the next fetch or the plan's destruction.
extern "C" int32_t hpaBatchFetch(hpa_batch *b, int64_t *ng, void **keys, void **cnts, void **sums){
    hpa_ctx *c = b->c;
    int32_t e;
    if(!b->pinned){
        int64_t tot = 0;
        b->pin_off.resize(b->nq);
        for(int qi = 0; qi < b->nq; qi++){ b->pin_off[qi] = tot; tot += 24 * (int64_t)b->hq[qi].range; }
        if(hipHostMalloc((void**)&b->pinned, (size_t)(tot > 0 ? tot : 8), hipHostMallocDefault) != hipSuccess){
            b->pinned = nullptr; set_error_detail("batch: pinned staging of %lld bytes", (long long)tot); R HPA_E_MEMORY;
        }
    }
    if(!b->ng_valid){
        if((e = d2h(c, b->ng.data(), b->ngdev, 16 * (L)b->nq))) R e;
        b->ng_valid = true;
    }
    for(int qi = 0; qi < b->nq; qi++){
        if(b->ng[2 * qi + 1]){ set_error_detail("batch query %d: key outside the planned range", qi); R HPA_E_DOMAIN; }
        const BQuery &q = b->hq[qi];
        const int64_t g = b->ng[2 * qi];
        char *base = b->pinned + b->pin_off[qi];
        char *hk = base, *hc = base + 8 * (int64_t)q.range, *hs = base + 16 * (int64_t)q.range;
        ng[qi] = g; keys[qi] = hk; cnts[qi] = hc; sums[qi] = hs;
        if(g > 0){
            if(hipMemcpyAsync(hk, q.okey, 4 * g, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
               hipMemcpyAsync(hc, q.ocnt, 8 * g, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
               hipMemcpyAsync(hs, q.osum, 8 * g, hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
        }
    }
    R hipStreamSynchronize(c->stream) == hipSuccess ? 0 : HPA_E_DEVICE;
}

extern "C" int32_t hpaBatchSetConfig(hpa_batch *b, int32_t block, int32_t tile_rows, int32_t lds_keys){
    R batch_layout(b, block, tile_rows, lds_keys);
}

extern "C" int32_t hpaBatchGetConfig(hpa_batch *b, int32_t *block, int32_t *tile_rows, int32_t *lds_keys){
    *block = b->block; *tile_rows = b->tile_rows; *lds_keys = b->lds_keys;
    R 0;
}

extern "C" int32_t hpaBatchRun(hpa_batch *b, int32_t use_graph){
    hpa_ctx *c = b->c;
    int32_t e;
    b->ng_valid = false;
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

// Result of query qi of the last run: zkey (I), zcnt (L), zsum (E) over the groups in the
// pfnGroup order.  The vectors point into the plan's buffers (valid until the next run).
extern "C" int32_t hpaBatchResult(hpa_batch *b, int32_t qi, hpa_V zkey, hpa_V zcnt, hpa_V zsum){
    hpa_ctx *c = b->c;
    if(qi < 0 || qi >= b->nq) R HPA_E_INDEX;
    int32_t e;
    if(!b->ng_valid){
        if((e = d2h(c, b->ng.data(), b->ngdev, 16 * (L)b->nq))) R e;
        b->ng_valid = true;
    }
    if(b->ng[2 * qi + 1]){ set_error_detail("batch query %d: key outside the planned range", qi); R HPA_E_DOMAIN; }
    const int64_t g = b->ng[2 * qi];
    const BQuery &q = b->hq[qi];
    auto setv = [&](V z, int32_t typ, void *p, L w) -> int32_t {
        z->typ = typ; z->len = g;
        if(g <= 1){ z->g = (char*)&z->b; z->i64 = 0; R g == 1 ? d2h(c, &z->b, p, w) : 0; }
        z->g = (char*)p; R 0;
    };
    if((e = setv(zkey, HPA_I, q.okey, 4)) || (e = setv(zcnt, HPA_L, q.ocnt, 8)) || (e = setv(zsum, HPA_E, q.osum, 8))) R e;
    R 0;
}

// Autotuner: time each candidate (block size, tile rows, LDS key budget) over `iters` runs of
// this plan and keep the fastest.  Writes the timings of all candidates to `ms` (capacity
// `cap`, may be null) in candidate order; returns the number of candidates through *ncand.
extern "C" int32_t hpaBatchAutotune(hpa_batch *b, int32_t iters, float *ms, int32_t cap, int32_t *ncand){
    static const int blocks[] = {256, 512, 1024};
    static const int tiles[] = {4096, 16384, 65536};
    static const int ldsk[] = {0, 512, 2730};
    hpa_ctx *c = b->c;
    hipEvent_t a, z;
    hipEventCreate(&a); hipEventCreate(&z);
    float best = 1e30f; int bb = b->block, bt = b->tile_rows, bl = b->lds_keys, k = 0;
    int32_t e = 0;
    for(int bi : blocks) for(int ti : tiles) for(int li : ldsk){
        if((e = batch_layout(b, bi, ti, li))) goto done;
        if((e = batch_enqueue(b, c->stream))) goto done;           // warm
        hipEventRecord(a, c->stream);
        for(int it = 0; it < iters; it++) if((e = batch_enqueue(b, c->stream))) goto done;
        hipEventRecord(z, c->stream);
        if(hipEventSynchronize(z) != hipSuccess){ e = HPA_E_DEVICE; goto done; }
        float t = 0; hipEventElapsedTime(&t, a, z); t /= iters > 0 ? iters : 1;
        if(ms && k < cap) ms[k] = t;
        k++;
        if(t < best){ best = t; bb = bi; bt = ti; bl = li; }
    }
done:
    hipEventDestroy(a); hipEventDestroy(z);
    if(ncand) *ncand = k;
    int32_t e2 = batch_layout(b, bb, bt, bl);
    b->ng_valid = false;
    R e ? e : e2;
}
