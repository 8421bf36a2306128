// Order-preserving stream compaction: compress, where, the multi-column peephole compress and
// the FA fused compress scan.  One pass: each 4096-row tile counts its selected rows with
// wavefront ballots, publishes its count, looks back over earlier tiles for its output offset
// (decoupled look-back), then scatters the selected rows of every column.  Rows keep their
// input order, as the reference's DOT count + ordered scatter does
// (h_primitive.cc:1479-1521, backend/h_system.cc:304-310, optimizer/autofusion.cc:886-899).
#include "hpa_internal.h"
#include "hpa_expr.h"

#define CP_BLOCK 256
#define CP_ITEMS 16
#define CP_TILE  (CP_BLOCK * CP_ITEMS)
#define CP_MAXC  8
static_assert(CP_BLOCK == 256 && CP_ITEMS * 4 == 64, "the tile scan gives one (item, wave) count to each lane of one wave");

#define ST_AGG  (1ull << 62)
#define ST_PRE  (2ull << 62)
#define ST_MASK ((1ull << 62) - 1)

struct CpCols { int n; int w[CP_MAXC]; const char *src[CP_MAXC]; char *dst[CP_MAXC]; };

__global__ void k_fill_u64(uint64_t *p, L n, uint64_t v){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) p[i] = v;
}
int32_t fill_u64(hpa_ctx *c, uint64_t *p, L n, uint64_t v){
    if(n <= 0) R 0;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, p, n, v);
    R check_launch("fill");
}

struct MaskPred {
    const uint8_t *m;
    __device__ __forceinline__ bool operator()(L i) const { return m[i] != 0; }
};
struct ExprPred {
    DevExpr e; DevCols x;
    __device__ __forceinline__ bool operator()(L i) const { return expr_true(e, eval_expr(e, x, i)); }
};

// Tile-order decoupled look-back, run by one whole wave: publishes the tile's aggregate, then
// reads the statuses of 64 predecessors at once (lane l reads tile - 1 - l - 64k), sums the
// aggregates up to the nearest published inclusive prefix and publishes its own inclusive
// prefix.  Returns the tile's exclusive prefix in every lane.  Statuses before tile 0 read as a
// prefix of 0.
__device__ __forceinline__ int64_t lookback_wave(uint64_t *status, L tile, uint64_t agg){
    const int lane = threadIdx.x & 63;
    if(tile == 0){
        if(lane == 0) __hip_atomic_store(&status[0], ST_PRE | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if(lane == 0) __hip_atomic_store(&status[tile], ST_AGG | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t pre = 0;
    L j = tile - 1;
    while(true){
        const L idx = j - lane;
        const uint64_t s = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ST_PRE;
        const uint64_t is_pre = __ballot((s >> 62) == 2);
        const uint64_t not_ready = __ballot((s >> 62) == 0);
        const int p = is_pre ? __ffsll((long long)is_pre) - 1 : 63;
        const uint64_t upto = p == 63 ? ~0ull : ((2ull << p) - 1);        // lanes 0..p
        if(not_ready & upto){ __builtin_amdgcn_s_sleep(1); continue; }
        uint64_t v = lane <= p ? (s & ST_MASK) : 0;
        for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        pre += v;
        if(is_pre) break;
        j -= 64;
    }
    if(lane == 0) __hip_atomic_store(&status[tile], ST_PRE | (pre + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (int64_t)pre;
}

__device__ __forceinline__ void copy_elem(char *dst, const char *src, int w, int64_t pos, L i){
    switch(w){
        case 1: ((uint8_t*)dst)[pos] = ((const uint8_t*)src)[i]; break;
        case 2: ((uint16_t*)dst)[pos] = ((const uint16_t*)src)[i]; break;
        case 4: ((uint32_t*)dst)[pos] = ((const uint32_t*)src)[i]; break;
        default: ((uint64_t*)dst)[pos] = ((const uint64_t*)src)[i]; break;
    }
}

template<typename Pred>
__global__ void __launch_bounds__(CP_BLOCK) k_compress(Pred pred, L n, CpCols cols, int64_t *idx_out,
                                                       uint64_t *status, unsigned *tile_ctr, L ntiles, int64_t *total){
    __shared__ unsigned s_tile;
    __shared__ uint32_t s_cnt[CP_ITEMS * 4];
    __shared__ int64_t s_pre;
    if(threadIdx.x == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const L tile = s_tile;
    const L base = tile * CP_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t bal[CP_ITEMS];
    #pragma unroll
    for(int k = 0; k < CP_ITEMS; k++){
        L i = base + (L)k * CP_BLOCK + threadIdx.x;
        bool m = i < n && pred(i);
        bal[k] = __ballot(m);
        if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    if(w == 0){
        // exclusive scan of the 64 (item, wave) counts in row order, one per lane
        const uint32_t mine = s_cnt[lane];
        uint32_t inc = mine;
        for(int o = 1; o < 64; o <<= 1){ uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        s_cnt[lane] = inc - mine;
        const uint32_t run = __shfl(inc, 63, 64);
        const int64_t pre = lookback_wave(status, tile, run);
        if(lane == 0){
            s_pre = pre;
            if(tile == ntiles - 1 && total) *total = pre + run;
        }
    }
    __syncthreads();
    const int64_t pre = s_pre;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    #pragma unroll
    for(int k = 0; k < CP_ITEMS; k++){
        if((bal[k] >> lane) & 1ull){
            L i = base + (L)k * CP_BLOCK + threadIdx.x;
            int64_t pos = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
            for(int c = 0; c < cols.n; c++) copy_elem(cols.dst[c], cols.src[c], cols.w[c], pos, i);
            if(idx_out) idx_out[pos] = i;
        }
    }
}

// per_row / per_sel: algorithmic bytes read per input row and written per selected row
template<typename Pred>
static int32_t run_compress(hpa_ctx *c, const char *name, Pred pred, L n, CpCols cols, int64_t *idx_out, int64_t *k_host,
                            double per_row, double per_sel){
    L ntiles = ceil_div(n, CP_TILE);
    if(ntiles < 1){ *k_host = 0; R 0; }
    uint64_t *status = (uint64_t*)dalloc(c, 8 * (ntiles + 2));
    if(!status) R HPA_E_MEMORY;
    unsigned *ctr = (unsigned*)(status + ntiles);
    int64_t *total = (int64_t*)(status + ntiles + 1);
    int32_t e = fill_u64(c, status, ntiles + 2, 0); if(e) R e;
    ProfScope ps(name, c->stream);
    hipLaunchKernelGGL(k_compress<Pred>, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream,
                       pred, n, cols, idx_out, status, ctr, ntiles, total);
    ps.stop();
    if((e = check_launch(name))) R e;
    e = d2h(c, k_host, total, 8);
    ps.bytes = per_row * (double)n + per_sel * (double)*k_host;
    R e;
}

// pfnCompress (h_primitive.cc:1479-1521)
static int32_t compress_many(hpa_ctx *c, int ny, V *z, V x, V *y, bool expr_mode, const ExprPred *ep, L n_expr){
    L n = expr_mode ? n_expr : x->len;
    for(int k = 0; k < ny; k++){
        if(y[k]->len != n) R HPA_E_LENGTH;
        switch(y[k]->typ){
            case HPA_B: case HPA_J: case HPA_H: case HPA_I: case HPA_L: case HPA_F: case HPA_E:
            case HPA_Q: case HPA_C: case HPA_S: case HPA_D: case HPA_Y: break;
            default: R HPA_E_NOT_IMPL;
        }
    }
    if(ny > CP_MAXC) R HPA_E_NOT_IMPL;
    L mark = c->dcur;
    CpCols cols; cols.n = ny;
    int32_t e;
    for(int k = 0; k < ny; k++){
        const void *p = nullptr;
        if(n > 0 && (e = dev_view(c, y[k], &p))) R e;
        cols.src[k] = (const char*)p; cols.w[k] = (int)type_width(y[k]->typ);
    }
    const uint8_t *m = nullptr;
    if(!expr_mode && n > 0){ const void *p; if((e = dev_view(c, x, &p))) R e; m = (const uint8_t*)p; }
    // worst-case outputs; the unused tail of each stays in the arena until the next reset
    L outmark = c->dcur;
    for(int k = 0; k < ny; k++){
        cols.dst[k] = (char*)dalloc(c, type_bytes(y[k]->typ, n > 0 ? n : 1));
        if(!cols.dst[k]){ c->dcur = mark; R HPA_E_MEMORY; }
    }
    int64_t kk = 0;
    double wsum = 0; for(int k = 0; k < ny; k++) wsum += cols.w[k];
    double wcond = 0;
    if(expr_mode) for(int k = 0; k < ep->x.n; k++) wcond += (double)type_width(ep->x.typ[k]);
    if(expr_mode) e = run_compress(c, "fused_compress", *ep, n, cols, nullptr, &kk, wcond + wsum, wsum);
    else e = run_compress(c, "compress", MaskPred{m}, n, cols, nullptr, &kk, 1.0 + wsum, wsum);
    if(e) R e;
    for(int k = 0; k < ny; k++){
        V zk = z[k];
        if(kk == n && n > 0 && !expr_mode){ *zk = *y[k]; continue; }   // copyV alias (h_primitive.cc:1513-1515)
        zk->typ = y[k]->typ; zk->len = kk;
        if(kk <= 1){
            zk->g = (char*)&zk->b; zk->i64 = 0;
            if(kk == 1 && (e = d2h(c, &zk->b, cols.dst[k], cols.w[k]))) R e;
        } else {
            zk->g = cols.dst[k];
            if(zk->typ == HPA_Y) zk->p = y[k]->p;            // enum keeps its target (vy(z)=vy(y))
        }
    }
    if(kk <= 1 || (kk == n && !expr_mode)) c->dcur = outmark;  // nothing of the scratch outputs is referenced
    else if(ny == 1) c->dcur = outmark + ((type_bytes(y[0]->typ, kk) + 255) & ~(L)255);
    R 0;
}

extern "C" int32_t pfnCompress(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(x->typ != HPA_B) R HPA_E_DOMAIN;
    if(x->len != y->len) R HPA_E_LENGTH;
    R compress_many(c, 1, &z, x, &y, false, nullptr, 0);
}

extern "C" int32_t hpaCompressMany(int32_t ny, hpa_V *z, hpa_V mask, hpa_V *y){
    hpa_ctx *c = ctx_cur();
    if(mask->typ != HPA_B) R HPA_E_DOMAIN;
    for(int k = 0; k < ny; k++) if(y[k]->len != mask->len) R HPA_E_LENGTH;
    R compress_many(c, ny, z, mask, y, false, nullptr, 0);
}

// pfnWhere (h_primitive.cc:728-762): indices of true rows, ascending
extern "C" int32_t pfnWhere(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(x->typ != HPA_B) R HPA_E_DOMAIN;
    L n = x->len;
    L mark = c->dcur;
    const void *p = nullptr; int32_t e;
    if(n > 0 && (e = dev_view(c, x, &p))) R e;
    int64_t *out = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!out) R HPA_E_MEMORY;
    CpCols cols; cols.n = 0;
    int64_t kk = 0;
    if((e = run_compress(c, "where", MaskPred{(const uint8_t*)p}, n, cols, out, &kk, 1.0, 8.0))) R e;
    z->typ = HPA_L; z->len = kk;
    note_bound(c, out, n);
    if(kk <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(kk == 1 && (e = d2h(c, &z->b, out, 8))) R e;
        c->dcur = mark;
    } else z->g = (char*)out;
    R 0;
}

// ---------------------------------------------------------------- expressions
int32_t expr_compile(const hpa_expr *e, int nx, const int32_t *coltyp, DevExpr *out){
    if(!e || e->n <= 0 || e->n > EXPR_MAX_INS) R HPA_E_DOMAIN;
    out->n = e->n;
    int8_t kind[EXPR_MAX_DEPTH]; int sp = 0, depth = 0;
    for(int k = 0; k < e->n; k++){
        const hpa_ins &in = e->code[k];
        DevIns &d = out->ins[k];
        d.op = (int16_t)in.op; d.arg = (int16_t)in.arg; d.imm = in.imm_i; d.af = d.bf = d.kf = 0; d.pad = 0;
        switch(in.op){
            case HPA_OP_COL: {
                if(in.arg < 0 || in.arg >= nx) R HPA_E_DOMAIN;
                int32_t t = coltyp[in.arg];
                if(!(is_basic(t)) || t == HPA_S || t == HPA_X) R HPA_E_TYPE;
                d.kf = (t == HPA_F || t == HPA_E);
                if(sp >= EXPR_MAX_DEPTH) R HPA_E_DOMAIN;
                kind[sp++] = d.kf; break;
            }
            case HPA_OP_CONST_I: if(sp >= EXPR_MAX_DEPTH) R HPA_E_DOMAIN; kind[sp++] = 0; break;
            case HPA_OP_CONST_F: { if(sp >= EXPR_MAX_DEPTH) R HPA_E_DOMAIN; memcpy(&d.imm, &in.imm_f, 8); d.kf = 1; kind[sp++] = 1; } break;
            case HPA_OP_NOT: if(sp < 1) R HPA_E_DOMAIN; d.af = kind[sp-1]; kind[sp-1] = 0; d.kf = 0; break;
            case HPA_OP_NEG: if(sp < 1) R HPA_E_DOMAIN; d.af = kind[sp-1]; d.kf = d.af; break;
            case HPA_OP_ADD: case HPA_OP_SUB: case HPA_OP_MUL: case HPA_OP_DIV:
            case HPA_OP_LT: case HPA_OP_LEQ: case HPA_OP_GT: case HPA_OP_GEQ: case HPA_OP_EQ: case HPA_OP_NEQ:
            case HPA_OP_AND: case HPA_OP_OR: {
                if(sp < 2) R HPA_E_DOMAIN;
                d.af = kind[sp-2]; d.bf = kind[sp-1]; sp--;
                bool arith = in.op <= HPA_OP_DIV;
                d.kf = arith ? (d.af || d.bf) : 0;
                kind[sp-1] = d.kf;
            } break;
            default: R HPA_E_DOMAIN;
        }
        if(sp > depth) depth = sp;
    }
    if(sp != 1) R HPA_E_DOMAIN;
    out->depth = depth;
    R 0;
}

int32_t cols_view(hpa_ctx *c, int nx, V *x, DevCols *cols, L *n){
    if(nx > EXPR_MAX_COLS) R HPA_E_NOT_IMPL;
    cols->n = nx; *n = nx > 0 ? x[0]->len : 0;
    for(int k = 0; k < nx; k++){
        if(x[k]->len != *n) R HPA_E_LENGTH;
        cols->typ[k] = x[k]->typ;
        const void *p = nullptr;
        if(*n > 0){ int32_t e = dev_view(c, x[k], &p); if(e) R e; }
        cols->p[k] = p;
    }
    R 0;
}

// FA compress scan: z[k] = y[k][cond], one evaluation of cond per row
extern "C" int32_t hpaFusedCompress(int32_t ny, hpa_V *z, hpa_V *y, int32_t nx, hpa_V *x, const hpa_expr *cond){
    hpa_ctx *c = ctx_cur();
    ExprPred ep; L n; int32_t e;
    if((e = cols_view(c, nx, x, &ep.x, &n))) R e;
    int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    if((e = expr_compile(cond, nx, ct, &ep.e))) R e;
    R compress_many(c, ny, z, nullptr, y, true, &ep, n);
}
