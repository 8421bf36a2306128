// Order-preserving stream compaction: compress, where, the multi-column peephole compress and
// the FA fused compress scan.  Two passes over 4096-row tiles, as the reference's DOT count +
// ordered scatter (h_primitive.cc:1479-1521, backend/h_system.cc:304-310,
// optimizer/autofusion.cc:886-899): count the selected rows of each tile (a byte mask is read
// 16 rows per thread with one vector load), scan the tile counts, then each tile ranks its
// selected rows with wave ballots and scatters every column at its offset.  The total is read
// between the passes, so outputs are allocated at their exact size, and an all-true mask
// aliases its input without a second pass.
#include "hpa_internal.h"
#include "hpa_expr.h"

#define CP_BLOCK 256
#define CP_ITEMS 16
#define CP_TILE  (CP_BLOCK * CP_ITEMS)
#define CP_MAXC  8
static_assert(CP_BLOCK == 256 && CP_ITEMS * 4 == 64, "the tile scan gives one (item, wave) count to each lane of one wave");


struct CpCols { int n; int w[CP_MAXC]; const char *src[CP_MAXC]; char *dst[CP_MAXC]; };
// FE map outputs of a compress: dst[j][k] = e[j](x) at the k-th selected row, as type typ[j]
#define CP_MAXM 2
struct CpMaps { int n; int typ[CP_MAXM]; char *dst[CP_MAXM]; DevCols x; DevExpr e[CP_MAXM]; const hpa_expr *hx[CP_MAXM]; };
// the FE map's conversion of a result to its output type (k_fused_map)
__device__ __forceinline__ uint64_t cp_map_bits(int typ, bool fk, Slot s){
    switch(typ){
        case HPA_B: return (uint64_t)(fk ? s.f != 0.0 : s.i != 0);
        case HPA_I: return (uint64_t)(uint32_t)(fk ? (int32_t)s.f : (int32_t)s.i);
        case HPA_L: return (uint64_t)(fk ? (int64_t)s.f : s.i);
        default: { const double d = fk ? s.f : (double)s.i; return (uint64_t)__double_as_longlong(d); }
    }
}

struct MaskPred {
    const uint8_t *m;
    __device__ __forceinline__ bool operator()(L i) const { return m[i] != 0; }
};
struct ExprPred {
    DevExpr e; DevCols x;
    __device__ __forceinline__ bool operator()(L i) const { return expr_true(e, eval_expr(e, x, i)); }
};
// a condition's byte mask, interpreted (when no generated code is available)
__global__ void k_cond_mask(ExprPred p, L n, uint8_t *m){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) m[i] = p(i) ? 1 : 0;
}

__device__ __forceinline__ void copy_elem(char *dst, const char *src, int w, int64_t pos, L i){
    switch(w){
        case 1: ((uint8_t*)dst)[pos] = ((const uint8_t*)src)[i]; break;
        case 2: ((uint16_t*)dst)[pos] = ((const uint16_t*)src)[i]; break;
        case 4: ((uint32_t*)dst)[pos] = ((const uint32_t*)src)[i]; break;
        default: ((uint64_t*)dst)[pos] = ((const uint64_t*)src)[i]; break;
    }
}

// Pass 1: selected rows per 4096-row tile.
template<typename Pred>
__global__ void __launch_bounds__(CP_BLOCK) k_cp_count(Pred pred, L n, uint32_t *tcount){
    __shared__ uint32_t sh[CP_BLOCK / 64];
    const L base = (L)blockIdx.x * CP_TILE;
    uint32_t c = 0;
    #pragma unroll
    for(int k = 0; k < CP_ITEMS; k++){
        L i = base + (L)k * CP_BLOCK + threadIdx.x;
        c += (i < n && pred(i)) ? 1u : 0u;
    }
    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0) tcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
// bit k of the result: row i0 + k of the 16 consecutive rows at i0 is selected
__device__ __forceinline__ uint32_t mask16(const uint8_t *mask, L i0, L n){
    uint32_t m = 0;
    if(i0 + 16 <= n){
        const uint4 v = *(const uint4*)(mask + i0);
        const uint32_t x[4] = {v.x, v.y, v.z, v.w};
        #pragma unroll
        for(int q = 0; q < 4; q++)
            #pragma unroll
            for(int b = 0; b < 4; b++) if((x[q] >> (8 * b)) & 0xFFu) m |= 1u << (4 * q + b);
    } else {
        for(int k = 0; k < 16 && i0 + k < n; k++) if(mask[i0 + k]) m |= 1u << k;
    }
    return m;
}
// the same for a byte mask: 16 consecutive mask bytes per thread, one 16-byte load.  With `mb`
// the thread's 16 selection bits are kept (mb[tile * 256 + thread]), so the write pass reads
// n / 8 bytes instead of the n-byte mask again.
__global__ void __launch_bounds__(CP_BLOCK) k_cp_count_mask(const uint8_t *m, L n, uint32_t *tcount, uint16_t *mb){
    __shared__ uint32_t sh[CP_BLOCK / 64];
    const L i0 = (L)blockIdx.x * CP_TILE + (L)threadIdx.x * 16;
    const uint32_t bits = mask16(m, i0, n);
    if(mb) mb[(L)blockIdx.x * CP_BLOCK + threadIdx.x] = (uint16_t)bits;
    uint32_t c = (uint32_t)__popc(bits);
    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0) tcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Pass 2: each tile takes its output offset from the scanned tile counts, ranks its selected
// rows with wave ballots (row order) and scatters every column (and/or the row index).
template<typename Pred, bool MAPS = false>
__global__ void __launch_bounds__(CP_BLOCK) k_cp_write(Pred pred, L n, CpCols cols, int64_t *idx_out, const int64_t *toff,
                                                       CpMaps maps = CpMaps()){
    __shared__ uint32_t s_cnt[CP_ITEMS * 4];
    const L base = (L)blockIdx.x * CP_TILE;
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
        const uint32_t mine = s_cnt[lane];
        uint32_t inc = mine;
        for(int o = 1; o < 64; o <<= 1){ uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        s_cnt[lane] = inc - mine;
    }
    __syncthreads();
    const int64_t pre = toff[blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    #pragma unroll
    for(int k = 0; k < CP_ITEMS; k++){
        if((bal[k] >> lane) & 1ull){
            L i = base + (L)k * CP_BLOCK + threadIdx.x;
            int64_t pos = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
            for(int c = 0; c < cols.n; c++) copy_elem(cols.dst[c], cols.src[c], cols.w[c], pos, i);
            if(idx_out) idx_out[pos] = i;
            if(MAPS) for(int j = 0; j < maps.n; j++){
                // lanes hold consecutive rows, so the map's column loads are coalesced
                const uint64_t v = cp_map_bits(maps.typ[j], maps.e[j].ins[maps.e[j].n - 1].kf, eval_expr(maps.e[j], maps.x, i));
                switch(maps.typ[j]){
                    case HPA_B: ((uint8_t*)maps.dst[j])[pos] = (uint8_t)v; break;
                    case HPA_I: ((uint32_t*)maps.dst[j])[pos] = (uint32_t)v; break;
                    default: ((uint64_t*)maps.dst[j])[pos] = v; break;
                }
            }
        }
    }
}

// Pass 2 for a byte mask: thread t owns the 16 consecutive rows [16t, 16t + 16) of the tile and
// reads their mask bytes with one 16-byte load and a selected run's values with vector loads;
// the tile's selected values are compacted in LDS (at the thread's offset from a scan of the
// threads' counts) and then written out by all threads together, consecutive threads to
// consecutive places.  Requires the mask and every column to start on a 16-element boundary,
// as allocations do.
template<int W>
__device__ __forceinline__ void cp_stage(const char *src, uint64_t *buf, L i0, uint32_t off, uint32_t m, bool full){
    typedef typename std::conditional<W == 1, uint8_t, typename std::conditional<W == 2, uint16_t,
            typename std::conditional<W == 4, uint32_t, uint64_t>::type>::type>::type E;
    E *lb = (E*)buf;
    if(full){
        E v[16];
        const uint4 *s4 = (const uint4*)((const E*)src + i0);
        #pragma unroll
        for(int q = 0; q < (16 * W) / 16; q++) ((uint4*)v)[q] = s4[q];
        #pragma unroll
        for(int k = 0; k < 16; k++) if((m >> k) & 1u) lb[off++] = v[k];
    } else {
        for(int k = 0; k < 16; k++) if((m >> k) & 1u) lb[off++] = ((const E*)src)[i0 + k];
    }
}
template<int W>
__device__ __forceinline__ void cp_flush(const uint64_t *buf, char *dst, int64_t pre, uint32_t cnt){
    typedef typename std::conditional<W == 1, uint8_t, typename std::conditional<W == 2, uint16_t,
            typename std::conditional<W == 4, uint32_t, uint64_t>::type>::type>::type E;
    const E *lb = (const E*)buf;
    for(uint32_t j = threadIdx.x; j < cnt; j += CP_BLOCK) ((E*)dst)[pre + j] = lb[j];
}

__global__ void __launch_bounds__(CP_BLOCK) k_cp_write_mask(const uint8_t *mask, const uint16_t *mb, L n, CpCols cols,
                                                          int64_t *idx_out, const int64_t *toff){
    __shared__ uint64_t buf[CP_TILE];
    __shared__ uint32_t sw[CP_BLOCK / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const L i0 = (L)blockIdx.x * CP_TILE + (L)t * 16;
    const bool full = i0 + 16 <= n;
    const uint32_t m = mb ? (uint32_t)mb[(L)blockIdx.x * CP_BLOCK + t] : mask16(mask, i0, n);
    // exclusive scan of the per-thread counts in thread (= row) order
    const uint32_t c = (uint32_t)__popc(m);
    uint32_t inc = c;
    for(int o = 1; o < 64; o <<= 1){ uint32_t y = __shfl_up(inc, o, 64); if(lane >= o) inc += y; }
    if(lane == 63) sw[w] = inc;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for(int q = 0; q < CP_BLOCK / 64; q++){ if(q < w) wbase += sw[q]; total += sw[q]; }
    if(total == 0) return;                                   // uniform over the workgroup
    const uint32_t off = wbase + inc - c;
    const int64_t pre = toff[blockIdx.x];
    for(int col = 0; col < cols.n; col++){
        if(m) switch(cols.w[col]){
            case 1: cp_stage<1>(cols.src[col], buf, i0, off, m, full); break;
            case 2: cp_stage<2>(cols.src[col], buf, i0, off, m, full); break;
            case 4: cp_stage<4>(cols.src[col], buf, i0, off, m, full); break;
            default: cp_stage<8>(cols.src[col], buf, i0, off, m, full); break;
        }
        __syncthreads();
        switch(cols.w[col]){
            case 1: cp_flush<1>(buf, cols.dst[col], pre, total); break;
            case 2: cp_flush<2>(buf, cols.dst[col], pre, total); break;
            case 4: cp_flush<4>(buf, cols.dst[col], pre, total); break;
            default: cp_flush<8>(buf, cols.dst[col], pre, total); break;
        }
        __syncthreads();
    }
    if(idx_out){
        uint32_t o = off;
        for(int k = 0; k < 16; k++) if((m >> k) & 1u) buf[o++] = (uint64_t)(i0 + k);
        __syncthreads();
        cp_flush<8>(buf, (char*)idx_out, pre, total);
    }
}

// Counting pass: tile offsets on the device and the total on the host.
template<typename Pred>
// Per-tile counts and their scan: toff[t] = rows selected before tile t, toff[ntiles] = total.
// With k_host the total is also read to the host (a synchronisation); without, only *toff_out.
// With mb_out, the byte-mask count also keeps the selection bits for the write pass (*mb_out;
// null when the mask is not 16-byte aligned).
static int32_t compress_count(hpa_ctx *c, Pred pred, const uint8_t *mask, L n, int64_t **toff_out, int64_t *k_host,
                              uint16_t **mb_out = nullptr){
    L ntiles = ceil_div(n, CP_TILE);
    if(k_host) *k_host = 0;
    *toff_out = nullptr;
    if(mb_out) *mb_out = nullptr;
    if(ntiles < 1) R 0;
    uint32_t *tcount = (uint32_t*)dalloc(c, 4 * ntiles);
    int64_t *toff = (int64_t*)dalloc(c, 8 * (ntiles + 1));
    if(!tcount || !toff) R HPA_E_MEMORY;
    if(mask && ((uintptr_t)mask & 15) == 0){
        uint16_t *mb = nullptr;
        if(mb_out){ mb = (uint16_t*)dalloc(c, 2 * CP_BLOCK * ntiles); if(!mb) R HPA_E_MEMORY; *mb_out = mb; }
        hipLaunchKernelGGL(k_cp_count_mask, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, mask, n, tcount, mb);
    } else
        hipLaunchKernelGGL(k_cp_count<Pred>, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred, n, tcount);
    int32_t e;
    if((e = check_launch("compress_count"))) R e;
    if((e = scan_u32_to_i64(c, tcount, toff, ntiles, toff + ntiles))) R e;
    *toff_out = toff;
    R k_host ? d2h(c, k_host, toff + ntiles, 8) : 0;
}
template<typename Pred>
static int32_t compress_write(hpa_ctx *c, Pred pred, L n, CpCols cols, int64_t *idx_out, const int64_t *toff){
    L ntiles = ceil_div(n, CP_TILE);
    if(ntiles < 1) R 0;
    hipLaunchKernelGGL(k_cp_write<Pred>, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred, n, cols, idx_out, toff);
    R check_launch("compress_write");
}
// the byte-mask form: 16 consecutive rows per thread when every operand is 16-element aligned
// (mb: the selection bits the count pass kept, or null)
static int32_t compress_write(hpa_ctx *c, MaskPred pred, L n, CpCols cols, int64_t *idx_out, const int64_t *toff,
                              const uint16_t *mb = nullptr, const CpMaps *maps = nullptr){
    L ntiles = ceil_div(n, CP_TILE);
    if(ntiles < 1) R 0;
    bool aligned = ((uintptr_t)pred.m & 15) == 0;
    for(int k = 0; k < cols.n; k++)
        aligned = aligned && ((uintptr_t)cols.src[k] % (uintptr_t)(16 * cols.w[k])) == 0 && cols.w[k] <= 8;
    if(maps && aligned && mb){
        // generated map kernels over the kept selection bits, then the plain columns; with the
        // code generator switched off, the interpreted write pass below
        int32_t e = 0;
        for(int j = 0; j < maps->n && !e; j++) e = jit_compress_map(c, maps->typ[j], maps->x, maps->hx[j], mb, toff, n, maps->dst[j]);
        if(!e){
            if(cols.n > 0 || idx_out)
                hipLaunchKernelGGL(k_cp_write_mask, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred.m, mb, n, cols, idx_out, toff);
            R check_launch("compress_write");
        }
        if(e != HPA_E_NOT_IMPL) R e;
    }
    if(maps) hipLaunchKernelGGL((k_cp_write<MaskPred, true>), dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred, n, cols, idx_out, toff, *maps);
    else if(aligned) hipLaunchKernelGGL(k_cp_write_mask, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred.m, mb, n, cols, idx_out, toff);
    else hipLaunchKernelGGL(k_cp_write<MaskPred>, dim3((unsigned)ntiles), dim3(CP_BLOCK), 0, c->stream, pred, n, cols, idx_out, toff);
    R check_launch("compress_write");
}

// pfnCompress (h_primitive.cc:1479-1521)
// nm > 0: also zm[j] = compress(x, maps->e[j](maps->x)) (no aliasing)
// cond (expression mode, the source of ep): the count pass is generated code that evaluates it on
// 16-row vector loads and keeps the selection bits, and the write pass runs from those bits as
// for a byte mask -- no mask is written; needs 16-element aligned columns.  Without that path the
// call returns HPA_E_NOT_IMPL with nothing done when `gen_only`, else evaluates cond per row.
static int32_t compress_many(hpa_ctx *c, int ny, V *z, V x, V *y, bool expr_mode, const ExprPred *ep, L n_expr,
                             int nm = 0, V *zm = nullptr, CpMaps *maps = nullptr, const hpa_expr *cond = nullptr,
                             bool gen_only = false){
    L n = expr_mode ? n_expr : x->len;
    const bool alias_ok = nm == 0;
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
    double wsum = 0; for(int k = 0; k < ny; k++) wsum += cols.w[k];
    for(int j = 0; j < nm; j++) wsum += (double)type_width(maps->typ[j]);
    double wcond = expr_mode ? 0 : 1.0;
    if(expr_mode) for(int k = 0; k < ep->x.n; k++) wcond += (double)type_width(ep->x.typ[k]);
    for(int k = 0; k < (nm > 0 ? maps->x.n : 0); k++) wcond += (double)type_width(maps->x.typ[k]);
    ProfScope ps(expr_mode ? "fused_compress" : "compress", c->stream);
    int64_t kk = 0, *toff = nullptr;
    // When the arena has room for every output at full length, the write pass is launched right
    // behind the count pass and the host reads the count while it runs; the arena top is then
    // pulled back to the end of the last output's selected rows.
    L maxb = 0; for(int k = 0; k < ny; k++) maxb += (type_bytes(y[k]->typ, n > 0 ? n : 1) + 255) & ~(L)255;
    for(int j = 0; j < nm; j++) maxb += (type_bytes(maps->typ[j], n > 0 ? n : 1) + 255) & ~(L)255;
    const bool spec = n > 0 && c->dcur + maxb + 4096 + (16 + 2 * CP_BLOCK) * ceil_div(n, CP_TILE) <= c->dtop;
    uint16_t *mb = nullptr;
    // expression mode through the generated count: every column 16-element aligned (the write
    // pass from selection bits reads 16 rows of a column per thread)
    bool gen = expr_mode && cond && n > 0 && jit_enabled();
    for(int k = 0; k < ny && gen; k++) gen = ((uintptr_t)cols.src[k] % (uintptr_t)(16 * cols.w[k])) == 0 && cols.w[k] <= 8;
    for(int k = 0; k < ep->x.n && gen; k++) gen = ((uintptr_t)ep->x.p[k] & 15) == 0;
    for(int k = 0; k < (nm > 0 ? maps->x.n : 0) && gen; k++) gen = ((uintptr_t)maps->x.p[k] & 15) == 0;
    if(expr_mode && gen_only && !gen){ arena_set(c, mark); R HPA_E_NOT_IMPL; }
    if(gen){
        const L ntiles = ceil_div(n, CP_TILE);
        uint32_t *tcount = (uint32_t*)dalloc(c, 4 * ntiles);
        toff = (int64_t*)dalloc(c, 8 * (ntiles + 1));
        mb = (uint16_t*)dalloc(c, 2 * CP_BLOCK * ntiles);
        if(!tcount || !toff || !mb){ arena_set(c, mark); R HPA_E_MEMORY; }
        e = jit_compress_count(c, ep->x, cond, n, mb, tcount);
        if(e == HPA_E_NOT_IMPL && !gen_only){ gen = false; mb = nullptr; toff = nullptr; }
        else if(e){ arena_set(c, mark); R e; }
        else if((e = scan_u32_to_i64(c, tcount, toff, ntiles, toff + ntiles)) || (!spec && (e = d2h(c, &kk, toff + ntiles, 8)))){
            arena_set(c, mark); R e;
        }
    }
    if(!gen){
        e = expr_mode ? compress_count(c, *ep, nullptr, n, &toff, spec ? nullptr : &kk)
                      : compress_count(c, MaskPred{m}, m, n, &toff, spec ? nullptr : &kk, &mb);
        if(e){ arena_set(c, mark); R e; }
    }
    if(!spec && kk == n && n > 0 && !expr_mode && alias_ok){   // copyV alias (h_primitive.cc:1513-1515)
        ps.stop(); ps.bytes = (double)n;
        arena_set(c, mark);
        for(int k = 0; k < ny; k++) *z[k] = *y[k];
        R 0;
    }
    for(int k = 0; k < ny; k++){
        cols.dst[k] = (char*)dalloc(c, type_bytes(y[k]->typ, spec ? n : (kk > 0 ? kk : 1)));
        if(!cols.dst[k]){ arena_set(c, mark); R HPA_E_MEMORY; }
    }
    for(int j = 0; j < nm; j++){
        maps->dst[j] = (char*)dalloc(c, type_bytes(maps->typ[j], spec ? n : (kk > 0 ? kk : 1)));
        if(!maps->dst[j]){ arena_set(c, mark); R HPA_E_MEMORY; }
    }
    e = expr_mode && !gen ? compress_write(c, *ep, n, cols, nullptr, toff)
                          : compress_write(c, MaskPred{m}, n, cols, nullptr, toff, mb, nm > 0 ? maps : nullptr);
    if(e) R e;
    if(spec){
        if((e = d2h(c, &kk, toff + ceil_div(n, CP_TILE), 8))) R e;
        if(kk == n && !expr_mode && alias_ok){              // copyV alias (h_primitive.cc:1513-1515)
            ps.stop(); ps.bytes = (double)n;
            arena_set(c, mark);
            for(int k = 0; k < ny; k++) *z[k] = *y[k];
            R 0;
        }
        // pull the arena top back to the end of the last output's selected rows
        const char *last = nm > 0 ? maps->dst[nm - 1] : cols.dst[ny - 1];
        const int32_t lt = nm > 0 ? maps->typ[nm - 1] : y[ny - 1]->typ;
        arena_set(c, (L)(last - c->dbase) + ((type_bytes(lt, kk > 0 ? kk : 1) + 255) & ~(L)255));
    }
    ps.stop();
    ps.bytes = wcond * (double)n + wsum * (double)kk;
    for(int k = 0; k < ny; k++){
        V zk = z[k];
        zk->typ = y[k]->typ; zk->len = kk;
        if(kk <= 1){
            zk->g = (char*)&zk->b; zk->i64 = 0;
            if(kk == 1 && (e = d2h(c, &zk->b, cols.dst[k], cols.w[k]))) R e;
        } else {
            zk->g = cols.dst[k];
            if(zk->typ == HPA_Y) zk->p = y[k]->p;            // enum keeps its target (vy(z)=vy(y))
        }
    }
    for(int j = 0; j < nm; j++){
        V zj = zm[j];
        const L w = type_width(maps->typ[j]);
        zj->typ = maps->typ[j]; zj->len = kk;
        if(kk <= 1){
            zj->g = (char*)&zj->b; zj->i64 = 0;
            if(kk == 1 && (e = d2h(c, &zj->b, maps->dst[j], w))) R e;
        } else zj->g = maps->dst[j];
    }
    R 0;
}

extern "C" int32_t pfnCompress(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(x->typ != HPA_B) R HPA_E_DOMAIN;
    if(x->len != y->len) R HPA_E_LENGTH;
    R compress_many(c, 1, &z, x, &y, false, nullptr, 0);
}

// The FA compress scan with FE maps fused (optimizer/autofusion.cc:814-817, 886-899): plain
// columns z[k] = compress(mask, y[k]) and computed ones zm[j] = compress(mask, maps[j](x)) in the
// one write pass, each map evaluated at the selected rows only.
extern "C" int32_t hpaCompressMap(int32_t ny, hpa_V *z, hpa_V mask, hpa_V *y, int32_t nm, hpa_V *zm,
                                  const int32_t *mtyp, const hpa_expr *maps, int32_t nx, hpa_V *x){
    hpa_ctx *c = ctx_cur();
    if(mask->typ != HPA_B) R HPA_E_DOMAIN;
    if(nm < 0 || nm > CP_MAXM || ny < 0) R HPA_E_NOT_IMPL;
    for(int k = 0; k < ny; k++) if(y[k]->len != mask->len) R HPA_E_LENGTH;
    CpMaps mp; memset(&mp, 0, sizeof(mp)); mp.n = nm;
    L n; int32_t e;
    if((e = cols_view(c, nx, x, &mp.x, &n))) R e;
    if(nx > 0 && n != mask->len) R HPA_E_LENGTH;
    int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    for(int j = 0; j < nm; j++){
        if(!(mtyp[j] == HPA_B || mtyp[j] == HPA_I || mtyp[j] == HPA_L || mtyp[j] == HPA_E)) R HPA_E_TYPE;
        mp.typ[j] = mtyp[j];
        mp.hx[j] = maps + j;
        if((e = expr_compile(maps + j, nx, ct, &mp.e[j]))) R e;
    }
    if(ny == 0 && nm == 0) R 0;
    R compress_many(c, ny, z, mask, y, false, nullptr, 0, nm, zm, &mp);
}

// hpaCompressMap with the mask's condition as an expression over x in place of the mask: the
// FA compress scan of one condition, with its plain columns and FE maps (optimizer/autofusion.cc:
// 814-817, 886-899).  The condition is evaluated by the count pass, which keeps each row's
// selection for the write pass, so no mask is materialised.
extern "C" int32_t hpaFusedCompressMap(int32_t ny, hpa_V *z, hpa_V *y, int32_t nm, hpa_V *zm, const int32_t *mtyp,
                                       const hpa_expr *maps, int32_t nx, hpa_V *x, const hpa_expr *cond){
    hpa_ctx *c = ctx_cur();
    if(nm < 0 || nm > CP_MAXM || ny < 0 || !cond) R HPA_E_NOT_IMPL;
    CpMaps mp; memset(&mp, 0, sizeof(mp)); mp.n = nm;
    L n; int32_t e;
    if((e = cols_view(c, nx, x, &mp.x, &n))) R e;
    for(int k = 0; k < ny; k++) if(y[k]->len != n) R HPA_E_LENGTH;
    int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    for(int j = 0; j < nm; j++){
        if(!(mtyp[j] == HPA_B || mtyp[j] == HPA_I || mtyp[j] == HPA_L || mtyp[j] == HPA_E)) R HPA_E_TYPE;
        mp.typ[j] = mtyp[j];
        mp.hx[j] = maps + j;
        if((e = expr_compile(maps + j, nx, ct, &mp.e[j]))) R e;
    }
    ExprPred ep; ep.x = mp.x;
    if((e = expr_compile(cond, nx, ct, &ep.e))) R e;
    if(ny == 0 && nm == 0) R 0;
    if(n > 1){
        e = compress_many(c, ny, z, nullptr, y, true, &ep, n, nm, zm, &mp, cond, true);
        if(e != HPA_E_NOT_IMPL) R e;
    }
    // otherwise the condition into a byte mask, then hpaCompressMap's path
    L mark = c->dcur;
    uint8_t *mask = (uint8_t*)dalloc(c, n > 0 ? n : 1);
    if(!mask) R HPA_E_MEMORY;
    V0 m; memset(&m, 0, sizeof(m));
    m.typ = HPA_B; m.len = n; m.g = (char*)mask;
    if(n > 0 && (e = jit_fused_map(c, HPA_B, nx, x, cond, mp.x, n, mask))){
        if(e != HPA_E_NOT_IMPL){ arena_set(c, mark); R e; }
        hipLaunchKernelGGL(k_cond_mask, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, ep, n, mask);
        if((e = check_launch("fused_compress_map"))) R e;
    }
    if(n == 1){ uint8_t b = 0; if((e = d2h(c, &b, mask, 1))) R e; m.g = (char*)&m.b; m.i64 = 0; m.b = b; }
    R compress_many(c, ny, z, &m, y, false, nullptr, 0, nm, zm, &mp);
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
    const uint8_t *m = (const uint8_t*)p;
    ProfScope ps("where", c->stream);
    int64_t kk = 0, *toff = nullptr;
    // as in compress: with room for n indices the write pass runs while the host reads the count
    const bool spec = n > 0 && c->dcur + 8 * n + 4096 + (16 + 2 * CP_BLOCK) * ceil_div(n, CP_TILE) <= c->dtop;
    uint16_t *mb = nullptr;
    if((e = compress_count(c, MaskPred{m}, m, n, &toff, spec ? nullptr : &kk, &mb))){ arena_set(c, mark); R e; }
    int64_t *out = (int64_t*)dalloc(c, 8 * (spec ? n : (kk > 0 ? kk : 1))); if(!out){ arena_set(c, mark); R HPA_E_MEMORY; }
    CpCols cols; cols.n = 0;
    if((e = compress_write(c, MaskPred{m}, n, cols, out, toff, mb))) R e;
    if(spec){
        if((e = d2h(c, &kk, toff + ceil_div(n, CP_TILE), 8))) R e;
        arena_set(c, (L)((char*)out - c->dbase) + ((8 * (kk > 0 ? kk : 1) + 255) & ~(L)255));
    }
    ps.stop(); ps.bytes = (double)n + 8.0 * (double)kk;
    z->typ = HPA_L; z->len = kk;
    if(kk <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(kk == 1 && (e = d2h(c, &z->b, out, 8))) R e;
        arena_set(c, mark);
    } else { z->g = (char*)out; note_bound(c, out, n); }
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
    for(int k = 0; k < ny; k++) if(y[k]->len != n) R HPA_E_LENGTH;
    if(n > 1){
        // the condition evaluated by the generated count pass (no mask written)
        e = compress_many(c, ny, z, nullptr, y, true, &ep, n, 0, nullptr, nullptr, cond, true);
        if(e != HPA_E_NOT_IMPL) R e;
    }
    if(n > 1){
        // the condition through generated code into a byte mask, then the mask compress
        L mark = c->dcur;
        uint8_t *mask = (uint8_t*)dalloc(c, n);
        if(!mask) R HPA_E_MEMORY;
        if(jit_fused_map(c, HPA_B, nx, x, cond, ep.x, n, mask) == 0){
            V0 m; memset(&m, 0, sizeof(m));
            m.typ = HPA_B; m.len = n; m.g = (char*)mask;
            R compress_many(c, ny, z, &m, y, false, nullptr, 0);
        }
        arena_set(c, mark);
    }
    R compress_many(c, ny, z, nullptr, y, true, &ep, n);
}
