// Sorting, grouping and ordering on a stable LSD radix sort.
//
// The reference groups and orders with serial radix/merge/quick sorts (backend/h_libs.cc:375-736,
// 1184-1403, 1912-1940; h_primitive.cc:993-1057, 1760-1816, 2656-2681).  Every group and order
// output is a function of a *stable* sort on a per-case key:
//   - i32 group / i32 order:  unsigned 32-bit key (4 byte-digit LSD on two's complement bytes)
//   - i64 group / i64 order:  unsigned 64-bit key, or its low 32 bits when no key has
//                             (x >> 32) > 0 (lib_radixsort_core_long, h_libs.cc:1240-1266)
//   - two i32 columns, equal flags: packed ((x0 << 32) | x1), x1 sign-extended
//   - char pair group: (c0 << 8) + c1 bucket order (pfnGroupTrie)
//   - other types: an order-preserving key of the type's `<`
// Descending single-column and packed orders reverse the stable ascending result, so ties come
// out in descending row order as in the reference.
//
// Radix pass: 4096-key tiles; per-tile digit histogram; exclusive scan in (digit, tile) order;
// scatter with each key's stable rank inside its tile computed from eight wavefront ballots
// (lanes sharing the digit) and per-wave LDS counters.  Digits that are constant over all keys
// are skipped.
#include "hpa_internal.h"

#define RS_BLOCK 256
#define RS_ROWS  16
#define RS_TILE  (RS_BLOCK * RS_ROWS)

// Digit histogram of one 4096-key tile.  With `bits` set (the first pass) it also writes the
// tile's OR and AND of all keys to bits[2*tile], bits[2*tile+1]: which key bits vary.
template<typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const K *keys, L n, int shift, uint32_t *hist, L ntiles,
                                                      unsigned long long *bits){
    __shared__ uint32_t cnt[256];
    __shared__ unsigned long long so[RS_BLOCK / 64], sa[RS_BLOCK / 64];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    L base = (L)blockIdx.x * RS_TILE;
    unsigned long long o = 0, a = ~0ull;
    #pragma unroll 4
    for(int r = 0; r < RS_ROWS; r++){
        L i = base + (L)r * RS_BLOCK + threadIdx.x;
        if(i < n){
            const K k = keys[i];
            unsigned d = (unsigned)((k >> shift) & 255); atomicAdd(&cnt[d], 1u);
            o |= (unsigned long long)k; a &= (unsigned long long)k;
        }
    }
    if(bits){
        for(int s = 32; s > 0; s >>= 1){ o |= __shfl_down(o, s, 64); a &= __shfl_down(a, s, 64); }
        if((threadIdx.x & 63) == 0){ so[threadIdx.x >> 6] = o; sa[threadIdx.x >> 6] = a; }
    }
    __syncthreads();
    hist[(L)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
    if(bits && threadIdx.x == 0){
        for(int w = 1; w < RS_BLOCK / 64; w++){ o |= so[w]; a &= sa[w]; }
        bits[2 * blockIdx.x] = o; bits[2 * blockIdx.x + 1] = a;
    }
}
// OR / AND of the tiles' partials (one workgroup)
__global__ void __launch_bounds__(256) k_rs_bits_final(const unsigned long long *part, L ntiles, unsigned long long *out){
    __shared__ unsigned long long so[4], sa[4];
    unsigned long long o = 0, a = ~0ull;
    for(L t = threadIdx.x; t < ntiles; t += 256){ o |= part[2 * t]; a &= part[2 * t + 1]; }
    for(int s = 32; s > 0; s >>= 1){ o |= __shfl_down(o, s, 64); a &= __shfl_down(a, s, 64); }
    if((threadIdx.x & 63) == 0){ so[threadIdx.x >> 6] = o; sa[threadIdx.x >> 6] = a; }
    __syncthreads();
    if(threadIdx.x == 0){
        for(int w = 1; w < 4; w++){ o |= so[w]; a &= sa[w]; }
        out[0] = o; out[1] = a;
    }
}

template<typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                                                         L n, int shift, const int64_t *off, L ntiles){
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t wbase[4][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for(int q = 0; q < 4; q++) wcnt[q][threadIdx.x] = 0;
    __syncthreads();
    const L tbase = (L)blockIdx.x * RS_TILE + (L)w * (RS_TILE / 4);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    K kr[RS_ROWS]; uint32_t vr[RS_ROWS]; uint32_t rk[RS_ROWS];
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        bool valid = i < n;
        K key = valid ? kin[i] : (K)0;
        unsigned d = (unsigned)((key >> shift) & 255);
        uint64_t peers = __ballot(valid);
        #pragma unroll
        for(int b = 0; b < 8; b++){
            uint64_t bb = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cur = wcnt[w][d];
        if(valid && before == 0) wcnt[w][d] = cur + (uint32_t)__popcll(peers);
        kr[r] = key; vr[r] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u; rk[r] = cur + before;
    }
    __syncthreads();
    {
        unsigned d = threadIdx.x;
        int64_t s = off[(L)d * ntiles + blockIdx.x];
        for(int q = 0; q < 4; q++){ wbase[q][d] = (uint32_t)s; s += wcnt[q][d]; }
    }
    __syncthreads();
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        if(i < n){
            unsigned d = (unsigned)((kr[r] >> shift) & 255);
            uint32_t pos = wbase[w][d] + rk[r];
            kout[pos] = kr[r]; vout[pos] = vr[r];
        }
    }
}

// Stable sort of (keys, vals) by key bits [lo, hi).  vals may be null (row ids 0..n-1).
template<typename K>
static int32_t radix_sort_t(hpa_ctx *c, K *keys, uint32_t *vals, L n, int lo, int hi, K **kout, uint32_t **vout){
    if(n >= (L)UINT32_MAX){ set_error_detail("radix sort of %lld keys exceeds 32-bit row ids", (long long)n); R HPA_E_NOT_IMPL; }
    *kout = keys; *vout = vals;
    if(n <= 0) R 0;
    int32_t e;
    L ntiles = ceil_div(n, RS_TILE);
    unsigned long long *bits = (unsigned long long*)dalloc(c, 16 + 16 * ntiles);
    K *k2 = (K*)dalloc(c, sizeof(K) * n);
    uint32_t *v2 = (uint32_t*)dalloc(c, 4 * n);
    uint32_t *v1 = vals;
    if(!v1){ v1 = (uint32_t*)dalloc(c, 4 * n); if(!v1) R HPA_E_MEMORY; }
    uint32_t *hist = (uint32_t*)dalloc(c, 4 * 256 * ntiles);
    int64_t *off = (int64_t*)dalloc(c, 8 * 256 * ntiles);
    if(!bits || !k2 || !v2 || !hist || !off) R HPA_E_MEMORY;
    // ping-pong: va == nullptr means "row ids" (first pass without given values)
    K *ka = keys, *kb = k2; uint32_t *va = vals, *vb = v2;
    uint32_t *spare = vals ? vals : v1;
    bool any = false;
    unsigned long long vary = ~0ull;
    bool first = true;
    for(int s = lo; s < hi; s += 8){
        if(!first && ((vary >> s) & 255ull) == 0) continue;
        {
            ProfScope ps("radix_hist", c->stream, (double)n * sizeof(K) + 4.0 * 256 * ntiles);
            hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream, ka, n, s, hist, ntiles,
                               first ? bits + 2 : nullptr);
        }
        if(first){
            // which bits vary, from the first histogram pass: constant digits are skipped
            hipLaunchKernelGGL(k_rs_bits_final, dim3(1), dim3(256), 0, c->stream, bits + 2, ntiles, bits);
            unsigned long long hb[2]; if((e = d2h(c, hb, bits, 16))) R e;
            vary = hb[0] ^ hb[1];
            first = false;
            if(((vary >> s) & 255ull) == 0) continue;
        }
        if((e = scan_u32_to_i64(c, hist, off, 256 * ntiles, nullptr))) R e;
        {
            ProfScope ps("radix_scatter", c->stream, (double)n * (2 * sizeof(K) + (va ? 8 : 4)) + 8.0 * 256 * ntiles);
            hipLaunchKernelGGL(k_rs_scatter<K>, dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream,
                               ka, va, kb, vb, n, s, off, ntiles);
        }
        if((e = check_launch("radix_sort"))) R e;
        any = true;
        K *tk = ka; ka = kb; kb = tk;
        uint32_t *written = vb;
        vb = (va != nullptr) ? va : spare;
        va = written;
    }
    if(!any){
        // every digit is constant: the input order is already sorted
        if(!vals){ if((e = iota_u32(c, v1, n))) R e; va = v1; }
    }
    *kout = ka; *vout = va;
    R 0;
}
int32_t radix_sort_pairs_u64(hpa_ctx *c, uint64_t *keys, uint32_t *vals, L n, int lo, int hi, uint64_t **ko, uint32_t **vo){
    R radix_sort_t<uint64_t>(c, keys, vals, n, lo, hi, ko, vo);
}
int32_t radix_sort_pairs_u32(hpa_ctx *c, uint32_t *keys, uint32_t *vals, L n, int lo, int hi, uint32_t **ko, uint32_t **vo){
    R radix_sort_t<uint32_t>(c, keys, vals, n, lo, hi, ko, vo);
}

// ---------------------------------------------------------------- runs -> dict(keys, flat list)
// Runs of equal keys in a sorted sequence, in two light passes over 4096-row tiles: count the
// run heads of each tile (a head is a row whose key differs from the previous row's), scan the
// tile counts, then write every row's original index into the flat i64 list and, for each head,
// the run's original first index (its key) and its start position.
#define RUN_ITEMS 16
#define RUN_TILE (256 * RUN_ITEMS)
template<typename K>
__device__ __forceinline__ bool run_head(const K *sk, const uint32_t *sidx, const uint64_t *eq, L i){
    if(i == 0) return true;
    return eq ? eq[sidx[i]] != eq[sidx[i - 1]] : sk[i] != sk[i - 1];
}
template<typename K>
__global__ void __launch_bounds__(256) k_run_tilecount(const K *sk, const uint32_t *sidx, const uint64_t *eq, L n, uint32_t *tcount){
    const L base = (L)blockIdx.x * RUN_TILE;
    uint32_t c = 0;
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){
        L i = base + (L)k * 256 + threadIdx.x;
        c += (i < n && run_head(sk, sidx, eq, i)) ? 1u : 0u;
    }
    __shared__ uint32_t sh[4];
    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0) tcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
template<typename K>
__global__ void __launch_bounds__(256) k_run_write(const K *sk, const uint32_t *sidx, const uint64_t *eq, L n,
                                                   const int64_t *toff, int64_t *keys, int64_t *start, int64_t *flat){
    __shared__ uint32_t s_cnt[RUN_ITEMS * 4];
    const L base = (L)blockIdx.x * RUN_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t bal[RUN_ITEMS];
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){
        L i = base + (L)k * 256 + threadIdx.x;
        bool h = false;
        if(i < n){ flat[i] = (int64_t)sidx[i]; h = run_head(sk, sidx, eq, i); }
        bal[k] = __ballot(h);
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
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){
        if((bal[k] >> lane) & 1ull){
            L i = base + (L)k * 256 + threadIdx.x;
            int64_t r = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
            keys[r] = (int64_t)sidx[i]; start[r] = i;
        }
    }
}
__global__ void k_run_counts(const int64_t *start, L g, L n, int64_t *cnt){
    for(L r = (L)blockIdx.x * blockDim.x + threadIdx.x; r < g; r += (L)gridDim.x * blockDim.x)
        cnt[r] = (r + 1 < g ? start[r + 1] : n) - start[r];
}

int32_t make_group_dict(hpa_ctx *c, V z, int64_t *keys, int64_t *counts, int64_t *flat, L g, L n){
    V cells = (V)halloc(c, 2 * sizeof(V0));
    V inner = node_alloc(c);
    if(!cells || !inner) R HPA_E_MEMORY;
    z->typ = HPA_N; z->len = 2; z->g = (char*)cells;
    V kz = cells, vz = cells + 1;
    kz->typ = HPA_L; kz->len = g;
    if(g <= 1){ kz->g = (char*)&kz->b; kz->i64 = 0; if(g == 1){ int32_t e = d2h(c, &kz->b, keys, 8); if(e) R e; } }
    else kz->g = (char*)keys;
    inner->typ = HPA_L; inner->len = n;
    if(n <= 1){ inner->g = (char*)&inner->b; inner->i64 = 0; if(n == 1){ int32_t e = d2h(c, &inner->b, flat, 8); if(e) R e; } }
    else inner->g = (char*)flat;
    vz->typ = HPA_G; vz->len = g; vz->g = (char*)counts; vz->g2 = inner;
    note_bound(c, keys, n); note_bound(c, flat, n);   // row numbers of the grouped vector
    R 0;
}

template<typename K>
static int32_t runs_to_group_t(hpa_ctx *c, V z, const K *sk, const uint32_t *sidx, L n, const uint64_t *eq){
    int32_t e;
    const L nt = ceil_div(n, RUN_TILE) > 0 ? ceil_div(n, RUN_TILE) : 1;
    uint32_t *tcount = (uint32_t*)dalloc(c, 4 * nt);
    int64_t *toff = (int64_t*)dalloc(c, 8 * (nt + 1));
    if(!tcount || !toff) R HPA_E_MEMORY;
    int64_t g = 0;
    if(n > 0){
        hipLaunchKernelGGL(k_run_tilecount<K>, dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, tcount);
        if((e = scan_u32_to_i64(c, tcount, toff, nt, toff + nt))) R e;
        if((e = d2h(c, &g, toff + nt, 8))) R e;
    }
    int64_t *keys = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *start = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *cnt = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *flat = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1));
    if(!keys || !start || !cnt || !flat) R HPA_E_MEMORY;
    if(n > 0){
        hipLaunchKernelGGL(k_run_write<K>, dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, toff, keys, start, flat);
        hipLaunchKernelGGL(k_run_counts, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, start, g, n, cnt);
    }
    if((e = check_launch("group_runs"))) R e;
    R make_group_dict(c, z, keys, cnt, flat, g, n);
}
int32_t runs_to_group_u64(hpa_ctx *c, V z, const uint64_t *sk, const uint32_t *sidx, L n, const uint64_t *eq){
    R runs_to_group_t<uint64_t>(c, z, sk, sidx, n, eq);
}
int32_t runs_to_group_u32(hpa_ctx *c, V z, const uint32_t *sk, const uint32_t *sidx, L n){
    R runs_to_group_t<uint32_t>(c, z, sk, sidx, n, nullptr);
}

// ---------------------------------------------------------------- key builders
__global__ void k_key_i32(const int32_t *x, uint32_t *k, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) k[i] = (uint32_t)x[i];
}
__global__ void k_key_trie(const uint8_t *a, const uint8_t *b, uint32_t *k, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) k[i] = ((uint32_t)a[i] << 8) | b[i];
}
__global__ void k_key_i64(const int64_t *x, uint64_t *k, L n, int low32){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        k[i] = low32 ? ((uint64_t)x[i] & 0xffffffffull) : (uint64_t)x[i];
}
__global__ void k_any_high(const int64_t *x, L n, int *flag){
    int f = 0;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) f |= (x[i] >> 32) > 0;
    if(__any(f) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}
// checkOrderCase1 (h_libs.cc:1912-1940): ((0|x0) << 32) | x1 with both sign-extended
__global__ void k_key_case1(const int32_t *x0, const int32_t *x1, int64_t *k, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        k[i] = (int64_t)(((uint64_t)(int64_t)x0[i]) << 32) | (int64_t)x1[i];
}

// order-preserving keys of `<` for the types the reference orders by comparison
template<typename T> __device__ __forceinline__ uint64_t okey(T v);
template<> __device__ __forceinline__ uint64_t okey<uint8_t>(uint8_t v){ return v; }
template<> __device__ __forceinline__ uint64_t okey<int8_t>(int8_t v){ return (uint64_t)(uint8_t)((uint8_t)v ^ 0x80); }
template<> __device__ __forceinline__ uint64_t okey<char>(char v){ return (uint64_t)(uint8_t)((uint8_t)v ^ 0x80); }
template<> __device__ __forceinline__ uint64_t okey<int16_t>(int16_t v){ return (uint64_t)(uint16_t)((uint16_t)v ^ 0x8000); }
template<> __device__ __forceinline__ uint64_t okey<int32_t>(int32_t v){ return (uint64_t)((uint32_t)v ^ 0x80000000u); }
template<> __device__ __forceinline__ uint64_t okey<int64_t>(int64_t v){ return (uint64_t)v ^ 0x8000000000000000ull; }
template<> __device__ __forceinline__ uint64_t okey<float>(float v){
    uint32_t b = __float_as_uint(v); if(v == 0.0f) b = 0;     // -0 == +0 under <
    return (b >> 31) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
}
template<> __device__ __forceinline__ uint64_t okey<double>(double v){
    uint64_t b = (uint64_t)__double_as_longlong(v); if(v == 0.0) b = 0;
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
template<typename T>
__global__ void k_okey(const T *x, uint64_t *k, L n, int desc, int bits){
    uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        uint64_t v = okey<T>(x[i]);
        k[i] = desc ? (~v & mask) : v;
    }
}
static int key_bits_of(int32_t t){
    switch(t){ case HPA_B: case HPA_J: case HPA_C: R 8; case HPA_H: R 16;
        case HPA_I: case HPA_F: case HPA_D: case HPA_M: case HPA_U: case HPA_W: case HPA_T: R 32; }
    R 64;
}
// key (into a fresh u64 buffer) for one column in comparison order; gathered through perm when
// given (LSD over columns sorts the current permutation by the next column)
template<typename T>
__global__ void k_okey_perm(const T *x, const uint32_t *perm, uint64_t *k, L n, int desc, int bits){
    uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        uint64_t v = okey<T>(x[perm ? perm[i] : i]);
        k[i] = desc ? (~v & mask) : v;
    }
}
static int32_t column_key(hpa_ctx *c, V col, const uint32_t *perm, int desc, uint64_t *k, int *bits){
    L n = col->len; const void *p; int32_t e;
    if((e = dev_view(c, col, &p))) R e;
    int b = key_bits_of(col->typ); *bits = b;
    dim3 g(grid_for(n, 256, 4)), bl(256);
    switch(col->typ){
        case HPA_B: hipLaunchKernelGGL(k_okey_perm<uint8_t>, g, bl, 0, c->stream, (const uint8_t*)p, perm, k, n, desc, b); break;
        case HPA_J: hipLaunchKernelGGL(k_okey_perm<int8_t>, g, bl, 0, c->stream, (const int8_t*)p, perm, k, n, desc, b); break;
        case HPA_C: hipLaunchKernelGGL(k_okey_perm<char>, g, bl, 0, c->stream, (const char*)p, perm, k, n, desc, b); break;
        case HPA_H: hipLaunchKernelGGL(k_okey_perm<int16_t>, g, bl, 0, c->stream, (const int16_t*)p, perm, k, n, desc, b); break;
        case HPA_I: case HPA_D: case HPA_M: case HPA_U: case HPA_W: case HPA_T:
            hipLaunchKernelGGL(k_okey_perm<int32_t>, g, bl, 0, c->stream, (const int32_t*)p, perm, k, n, desc, b); break;
        case HPA_L: case HPA_Z: hipLaunchKernelGGL(k_okey_perm<int64_t>, g, bl, 0, c->stream, (const int64_t*)p, perm, k, n, desc, b); break;
        case HPA_F: hipLaunchKernelGGL(k_okey_perm<float>, g, bl, 0, c->stream, (const float*)p, perm, k, n, desc, b); break;
        case HPA_E: hipLaunchKernelGGL(k_okey_perm<double>, g, bl, 0, c->stream, (const double*)p, perm, k, n, desc, b); break;
        default: R HPA_E_NOT_IMPL;
    }
    R check_launch("order_key");
}

// ---------------------------------------------------------------- group
// pfnGroupTrie (h_primitive.cc:2656-2681): buckets by (c0<<8)+c1 in ascending code order
extern "C" int32_t pfnGroupTrie(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(!(x->typ == HPA_G && x->len == 2)) R HPA_E_DOMAIN;
    V a = (V)x->g, b = a + 1;
    if(a->typ != HPA_C || b->typ != HPA_C || a->len != b->len) R HPA_E_DOMAIN;
    L n = a->len; int32_t e;
    const void *pa = nullptr, *pb = nullptr;
    if(n > 0){ if((e = dev_view(c, a, &pa)) || (e = dev_view(c, b, &pb))) R e; }
    uint32_t *k = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_trie, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)pa, (const uint8_t*)pb, k, n);
    uint32_t *sk, *sv;
    if((e = radix_sort_pairs_u32(c, k, nullptr, n, 0, 16, &sk, &sv))) R e;
    R runs_to_group_u32(c, z, sk, sv, n);
}

static int32_t group_int(hpa_ctx *c, V z, V x){      // lib_group_by_normal_int, h_libs.cc:1295-1329
    L n = x->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, x, &p))) R e;
    uint32_t *k = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_i32, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)p, k, n);
    uint32_t *sk, *sv;
    if((e = radix_sort_pairs_u32(c, k, nullptr, n, 0, 32, &sk, &sv))) R e;
    R runs_to_group_u32(c, z, sk, sv, n);
}

static int32_t any_high_i64(hpa_ctx *c, const int64_t *p, L n, bool *hi){
    int *f = (int*)dalloc(c, 16); if(!f) R HPA_E_MEMORY;
    hipMemsetAsync(f, 0, 4, c->stream);
    if(n > 0) hipLaunchKernelGGL(k_any_high, dim3(grid_for(n, 256, 8)), dim3(256), 0, c->stream, p, n, f);
    int h = 0; int32_t e = d2h(c, &h, f, 4); *hi = h != 0; R e;
}

static int32_t group_long(hpa_ctx *c, V z, V x){     // lib_group_by_normal_long, h_libs.cc:1369-1403
    L n = x->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, x, &p))) R e;
    bool hi; if((e = any_high_i64(c, (const int64_t*)p, n, &hi))) R e;
    uint64_t *k = (uint64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_i64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)p, k, n, hi ? 0 : 1);
    uint64_t *sk; uint32_t *sv;
    if((e = radix_sort_pairs_u64(c, k, nullptr, n, 0, hi ? 64 : 32, &sk, &sv))) R e;
    // runs on the full 64-bit value (h_libs.cc:1378: pos[i-1].x != pos[i].x)
    R runs_to_group_u64(c, z, sk, sv, n, (const uint64_t*)p);
}

// runs of equal rows (all columns) along a permutation
__global__ void k_perm_flags(int nc, const void *const *cols, const int *w, const uint32_t *perm, L n, uint32_t *flag){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        if(i == 0){ flag[0] = 1; continue; }
        uint32_t a = perm[i - 1], b = perm[i]; bool d = false;
        for(int q = 0; q < nc && !d; q++){
            switch(w[q]){
                case 1: d = ((const uint8_t*)cols[q])[a] != ((const uint8_t*)cols[q])[b]; break;
                case 2: d = ((const uint16_t*)cols[q])[a] != ((const uint16_t*)cols[q])[b]; break;
                case 4: d = ((const uint32_t*)cols[q])[a] != ((const uint32_t*)cols[q])[b]; break;
                default: d = ((const uint64_t*)cols[q])[a] != ((const uint64_t*)cols[q])[b]; break;
            }
        }
        flag[i] = d ? 1u : 0u;
    }
}
// value equality of the comparator (floats: == so that -0 and +0 tie)
__global__ void k_perm_flags_f(int nc, const void *const *cols, const int *typ, const uint32_t *perm, L n, uint32_t *flag){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        if(i == 0){ flag[0] = 1; continue; }
        uint32_t a = perm[i - 1], b = perm[i]; bool d = false;
        for(int q = 0; q < nc && !d; q++){
            switch(typ[q]){
                case HPA_B: case HPA_J: case HPA_C: d = ((const uint8_t*)cols[q])[a] != ((const uint8_t*)cols[q])[b]; break;
                case HPA_H: d = ((const uint16_t*)cols[q])[a] != ((const uint16_t*)cols[q])[b]; break;
                case HPA_F: d = ((const float*)cols[q])[a] != ((const float*)cols[q])[b]; break;
                case HPA_E: d = ((const double*)cols[q])[a] != ((const double*)cols[q])[b]; break;
                case HPA_L: case HPA_Z: case HPA_Q: d = ((const uint64_t*)cols[q])[a] != ((const uint64_t*)cols[q])[b]; break;
                default: d = ((const uint32_t*)cols[q])[a] != ((const uint32_t*)cols[q])[b]; break;
            }
        }
        flag[i] = d ? 1u : 0u;
    }
}

static int32_t order_perm(hpa_ctx *c, V x, const uint8_t *flags, uint32_t **perm_out, bool *ties);
// run heads flagged per row (list keys): scatter each head's original index and position
__global__ void k_run_emit(const uint32_t *flag, const int64_t *pos, const uint32_t *sidx, L n,
                           int64_t *keys, int64_t *start, int64_t *flat){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        flat[i] = sidx[i];
        if(flag[i]){ int64_t r = pos[i]; keys[r] = sidx[i]; start[r] = i; }
    }
}

static int32_t runs_along_perm(hpa_ctx *c, V z, int nc, V cols, const uint32_t *perm, L n){
    int32_t e;
    const void **hp = (const void**)malloc(sizeof(void*) * nc);
    int *ht = (int*)malloc(sizeof(int) * nc);
    for(int q = 0; q < nc; q++){ const void *p = nullptr; if(n > 0 && (e = dev_view(c, cols + q, &p))){ free(hp); free(ht); R e; } hp[q] = p; ht[q] = cols[q].typ; }
    void **dp = (void**)dalloc(c, sizeof(void*) * nc); int *dt = (int*)dalloc(c, sizeof(int) * nc);
    uint32_t *flag = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1));
    int64_t *pos = (int64_t*)dalloc(c, 8 * (n + 1));
    if(!dp || !dt || !flag || !pos){ free(hp); free(ht); R HPA_E_MEMORY; }
    hipMemcpyAsync(dp, hp, sizeof(void*) * nc, hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(dt, ht, sizeof(int) * nc, hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    free(hp); free(ht);
    if(n > 0) hipLaunchKernelGGL(k_perm_flags_f, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, nc, (const void* const*)dp, dt, perm, n, flag);
    if((e = scan_u32_to_i64(c, flag, pos, n, pos + n))) R e;
    int64_t g = 0; if((e = d2h(c, &g, pos + n, 8))) R e;
    int64_t *keys = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *start = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *cnt = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *flat = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1));
    if(!keys || !start || !cnt || !flat) R HPA_E_MEMORY;
    if(n > 0){
        hipLaunchKernelGGL(k_run_emit, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, flag, pos, perm, n, keys, start, flat);
        hipLaunchKernelGGL(k_run_counts, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, start, g, n, cnt);
    }
    if((e = check_launch("group_runs"))) R e;
    R make_group_dict(c, z, keys, cnt, flat, g, n);
}

// isOrdered (backend/h_system.cc:678-698) on the device: monotone (either direction, non-strict)
template<typename T>
__global__ void __launch_bounds__(256) k_monotone(const T *x, L n, int *flags){
    int up = 0, down = 0;
    for(L i = 1 + (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T a = x[i - 1], b = x[i]; up |= b > a; down |= b < a;
    }
    up = __syncthreads_or(up); down = __syncthreads_or(down);
    if(threadIdx.x == 0){ if(up) atomicOr(&flags[0], 1); if(down) atomicOr(&flags[1], 1); }
}
int32_t is_ordered_dev(hpa_ctx *c, V x, bool *ord){
    *ord = false;
    if(!(x->typ == HPA_I || x->typ == HPA_L)) R 0;
    L n = x->len; if(n <= 2){ *ord = true; if(n <= 1) R 0; }
    const void *p; int32_t e = dev_view(c, x, &p); if(e) R e;
    int *f = (int*)dalloc(c, 16); if(!f) R HPA_E_MEMORY;
    hipMemsetAsync(f, 0, 8, c->stream);
    if(x->typ == HPA_I) hipLaunchKernelGGL(k_monotone<int32_t>, dim3(reduce_grid(n, 256)), dim3(256), 0, c->stream, (const int32_t*)p, n, f);
    else hipLaunchKernelGGL(k_monotone<int64_t>, dim3(reduce_grid(n, 256)), dim3(256), 0, c->stream, (const int64_t*)p, n, f);
    int h[2]; if((e = d2h(c, h, f, 8))) R e;
    *ord = !(h[0] && h[1]);
    R 0;
}

// pfnGroupBasic (h_primitive.cc:1002-1057)
extern "C" int32_t pfnGroupBasic(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(x->typ == HPA_I) R group_int(c, z, x);
    if(x->typ == HPA_L) R group_long(c, z, x);
    if(x->typ == HPA_G){
        L nc = x->len;
        if(nc == 0) R HPA_E_DOMAIN;
        L n = ((V)x->g)->len;
        for(L q = 0; q < nc; q++){ V col = ((V)x->g) + q; if(!is_basic(col->typ)) R HPA_E_DOMAIN; if(col->len != n) R HPA_E_MATCH; }
        std::vector<uint8_t> up((size_t)nc, 1);
        uint32_t *perm; bool ties; int32_t e;
        if((e = order_perm(c, x, up.data(), &perm, &ties))) R e;
        R runs_along_perm(c, z, (int)nc, (V)x->g, perm, n);
    }
    if(is_basic(x->typ)){
        if(x->typ == HPA_S || x->typ == HPA_X || x->typ == HPA_Q) R HPA_E_NOT_IMPL;
        V0 lst; memset(&lst, 0, sizeof(lst));
        lst.typ = HPA_G; lst.len = 1; lst.g = (char*)x;
        uint8_t up = 1; uint32_t *perm; bool ties; int32_t e;
        if((e = order_perm(c, &lst, &up, &perm, &ties))) R e;
        R runs_along_perm(c, z, 1, x, perm, x->len);
    }
    R HPA_E_DOMAIN;
}

extern "C" int32_t pfnGroup(hpa_V z, hpa_V x){      // h_primitive.cc:993-999
    if(x->typ == HPA_G && x->len == 2){
        V a = (V)x->g, b = a + 1;
        if(a->typ == HPA_C && b->typ == HPA_C && a->len == b->len) R pfnGroupTrie(z, x);
    }
    R pfnGroupBasic(z, x);
}

// ---------------------------------------------------------------- order
__global__ void k_perm_to_i64(const uint32_t *perm, int64_t *z, L n, int reverse){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        z[i] = reverse ? perm[n - 1 - i] : perm[i];
}
// adjacent full-key equality along the final permutation (ties the reference may order
// differently through its quicksort)
__global__ void k_count_ties(const uint64_t *k, L n, unsigned long long *cnt){
    unsigned long long c0 = 0;
    for(L i = 1 + (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) c0 += k[i] == k[i - 1];
    for(int o = 32; o > 0; o >>= 1) c0 += __shfl_down(c0, o, 64);
    if((threadIdx.x & 63) == 0 && c0) atomicAdd(cnt, c0);
}

// Stable permutation of x (list of columns, flags per column) in lexicographic order, ties by
// ascending row.  LSD over columns: sort by the last column, then stably by each earlier one.
// Small inputs (n <= 1024, <= 8 columns): one workgroup ranks every row against every other on
// the same order keys -- ties by row -- in a single launch with no host round trip.
#define SMALL_ORDER 1024
#define SMALL_COLS 8
struct SmallCols { int nc; int typ[SMALL_COLS]; int desc[SMALL_COLS]; const void *p[SMALL_COLS]; };
__device__ __forceinline__ uint64_t okey_any(int t, const void *p, L i){
    switch(t){
        case HPA_B: return okey<uint8_t>(((const uint8_t*)p)[i]);
        case HPA_J: return okey<int8_t>(((const int8_t*)p)[i]);
        case HPA_C: return okey<char>(((const char*)p)[i]);
        case HPA_H: return okey<int16_t>(((const int16_t*)p)[i]);
        case HPA_L: case HPA_Z: return okey<int64_t>(((const int64_t*)p)[i]);
        case HPA_F: return okey<float>(((const float*)p)[i]);
        case HPA_E: return okey<double>(((const double*)p)[i]);
        default: return okey<int32_t>(((const int32_t*)p)[i]);
    }
}
__global__ void __launch_bounds__(1024) k_order_small(SmallCols s, int n, uint32_t *perm){
    __shared__ uint64_t k[SMALL_COLS][SMALL_ORDER];
    for(int q = 0; q < s.nc; q++)
        for(int i = threadIdx.x; i < n; i += blockDim.x){ uint64_t v = okey_any(s.typ[q], s.p[q], i); k[q][i] = s.desc[q] ? ~v : v; }
    __syncthreads();
    for(int i = threadIdx.x; i < n; i += blockDim.x){
        int rank = 0;
        for(int j = 0; j < n; j++){
            int less = j < i, q = 0;
            for(; q < s.nc; q++){ uint64_t a = k[q][j], b = k[q][i]; if(a != b){ less = a < b; break; } }
            rank += less;
        }
        perm[rank] = (uint32_t)i;
    }
}

static int32_t order_perm(hpa_ctx *c, V x, const uint8_t *flags, uint32_t **perm_out, bool *ties){
    L nc = x->len; L n = ((V)x->g)->len; int32_t e;
    *ties = false;
    uint32_t *perm = nullptr;
    if(n > 0 && n <= SMALL_ORDER && nc <= SMALL_COLS){
        SmallCols s; memset(&s, 0, sizeof(s)); s.nc = (int)nc;
        for(L q = 0; q < nc; q++){
            V col = ((V)x->g) + q;
            if(col->typ == HPA_S || col->typ == HPA_X || col->typ == HPA_Q) R HPA_E_NOT_IMPL;
            const void *p; if((e = dev_view(c, col, &p))) R e;
            s.typ[q] = col->typ; s.p[q] = p; s.desc[q] = flags[q] ? 0 : 1;
        }
        perm = (uint32_t*)dalloc(c, 4 * n); if(!perm) R HPA_E_MEMORY;
        ProfScope ps("order_small", c->stream);
        hipLaunchKernelGGL(k_order_small, dim3(1), dim3(1024), 0, c->stream, s, (int)n, perm);
        if((e = check_launch("order_small"))) R e;
        *perm_out = perm;
        R 0;
    }
    uint64_t *k = (uint64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    for(L q = nc - 1; q >= 0; q--){
        V col = ((V)x->g) + q;
        int bits;
        if((e = column_key(c, col, perm, flags[q] ? 0 : 1, k, &bits))) R e;
        uint64_t *sk; uint32_t *sv;
        // the sort reorders the current permutation: values = perm (or identity)
        uint32_t *vals = perm;
        if((e = radix_sort_pairs_u64(c, k, vals, n, 0, bits, &sk, &sv))) R e;
        perm = sv;
        if(sk != k) k = sk;   // keep the sorted keys buffer for reuse
        if(!perm){ perm = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!perm) R HPA_E_MEMORY; if((e = iota_u32(c, perm, n))) R e; }
    }
    if(!perm){ perm = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!perm) R HPA_E_MEMORY; if((e = iota_u32(c, perm, n))) R e; }
    *perm_out = perm;
    R 0;
}

static int32_t order_emit(hpa_ctx *c, V z, const uint32_t *perm, L n, bool reverse){
    void *dz; int32_t e;
    L mark = c->dcur;
    if((e = out_init(c, z, HPA_L, n, &dz))) R e;
    if(n > 0) hipLaunchKernelGGL(k_perm_to_i64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, perm, (int64_t*)dz, n, reverse ? 1 : 0);
    if((e = check_launch("order"))) R e;
    if(is_inline(z)){ e = commit_inline(c, z, dz); c->dcur = mark; }
    else note_bound(c, dz, n);                   // a permutation of 0..n-1
    R e;
}

// single i32 / i64 column or the packed pair: stable unsigned radix, desc = reversal
static int32_t order_radix_i32(hpa_ctx *c, V z, V col, bool up){
    L n = col->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, col, &p))) R e;
    uint32_t *k = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_i32, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)p, k, n);
    uint32_t *sk, *sv;
    if((e = radix_sort_pairs_u32(c, k, nullptr, n, 0, 32, &sk, &sv))) R e;
    R order_emit(c, z, sv, n, !up);
}
static int32_t order_radix_i64(hpa_ctx *c, V z, const int64_t *p, L n, bool up){
    int32_t e; bool hi;
    if((e = any_high_i64(c, p, n, &hi))) R e;
    uint64_t *k = (uint64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_i64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, p, k, n, hi ? 0 : 1);
    uint64_t *sk; uint32_t *sv;
    if((e = radix_sort_pairs_u64(c, k, nullptr, n, 0, hi ? 64 : 32, &sk, &sv))) R e;
    R order_emit(c, z, sv, n, !up);
}

// lib_quicksort_char on fewer than 10 rows (h_libs.cc:390-392): the all-pairs exchange with
// `>` leaves the range non-increasing; one thread restates it.
__global__ void k_char_small(const char *s, int64_t *z, int n, int desc){
    if(threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t r[10];
    for(int i = 0; i < n; i++) r[i] = i;
    for(int i = 0; i < n; i++) for(int j = 0; j < n; j++)
        if(s[r[i]] > s[r[j]]){ int64_t t = r[i]; r[i] = r[j]; r[j] = t; }
    if(desc) for(int i = 0, k = n - 1; i < k; i++, k--){ int64_t t = r[i]; r[i] = r[k]; r[k] = t; }
    for(int i = 0; i < n; i++) z[i] = r[i];
}

static int32_t order_vector(hpa_ctx *c, V z, V col, bool up){   // lib_order_by_vector, h_libs.cc:730-736
    L n = col->len; int32_t e;
    if(col->typ == HPA_I) R order_radix_i32(c, z, col, up);
    if(col->typ == HPA_L){
        const void *p = nullptr;
        if(n > 0 && (e = dev_view(c, col, &p))) R e;
        R order_radix_i64(c, z, (const int64_t*)p, n, up);
    }
    if(col->typ == HPA_S || col->typ == HPA_X || col->typ == HPA_Q) R HPA_E_NOT_IMPL;
    if(col->typ == HPA_C){
        if(n < 10){
            const void *p = nullptr;
            if(n > 0 && (e = dev_view(c, col, &p))) R e;
            void *dz; L mark = c->dcur;
            if((e = out_init(c, z, HPA_L, n, &dz))) R e;
            if(n > 0) hipLaunchKernelGGL(k_char_small, dim3(1), dim3(64), 0, c->stream, (const char*)p, (int64_t*)dz, (int)n, up ? 0 : 1);
            if((e = check_launch("order"))) R e;
            if(is_inline(z)){ e = commit_inline(c, z, dz); c->dcur = mark; }
            R e;
        }
        // counting sort by byte (stable, ascending), then reversal for desc
        uint64_t *k = (uint64_t*)dalloc(c, 8 * n); if(!k) R HPA_E_MEMORY;
        const void *p; if((e = dev_view(c, col, &p))) R e;
        hipLaunchKernelGGL(k_okey_perm<uint8_t>, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)p, (const uint32_t*)nullptr, k, n, 0, 8);
        uint64_t *sk; uint32_t *sv;
        if((e = radix_sort_pairs_u64(c, k, nullptr, n, 0, 8, &sk, &sv))) R e;
        R order_emit(c, z, sv, n, !up);
    }
    // comparison-sorted types (quicksort in the reference): stable sort on the order key
    V0 lst; memset(&lst, 0, sizeof(lst)); lst.typ = HPA_G; lst.len = 1; lst.g = (char*)col;
    uint8_t f = up ? 1 : 0; uint32_t *perm; bool ties;
    if((e = order_perm(c, &lst, &f, &perm, &ties))) R e;
    R order_emit(c, z, perm, n, false);
}

// pfnOrderBy (h_primitive.cc:1769-1816)
extern "C" int32_t pfnOrderBy(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    int32_t e;
    if(x->typ == HPA_G && y->typ == HPA_B && x->len == y->len){
        L nc = x->len;
        for(L q = 0; q < nc; q++) if(!is_basic(((V)x->g)[q].typ)) R HPA_E_DOMAIN;
        if(nc > 0){ L n = ((V)x->g)->len; for(L q = 0; q < nc; q++) if(((V)x->g)[q].len != n) R HPA_E_MATCH; }
        std::vector<uint8_t> fl((size_t)(nc > 0 ? nc : 1));
        if(nc > 0){ if((e = hpaReadPayload(y, fl.data(), nc))) R e; }
        V a = (V)x->g;
        if(nc == 2 && a[0].typ == HPA_I && a[1].typ == HPA_I && fl[0] == fl[1]){
            // checkOrderCase1 (h_libs.cc:1912-1940)
            L n = a[0].len;
            const void *p0 = nullptr, *p1 = nullptr;
            if(n > 0){ if((e = dev_view(c, a, &p0)) || (e = dev_view(c, a + 1, &p1))) R e; }
            int64_t *k = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
            if(n > 0) hipLaunchKernelGGL(k_key_case1, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)p0, (const int32_t*)p1, k, n);
            R order_radix_i64(c, z, k, n, fl[0] != 0);
        }
        if(nc == 1) R order_vector(c, z, a, fl[0] != 0);
        L n = nc == 0 ? 0 : a[0].len;
        uint32_t *perm = nullptr; bool ties;
        if(nc == 0 || n == 0){ void *dz; R out_init(c, z, HPA_L, 0, &dz); }
        if((e = order_perm(c, x, fl.data(), &perm, &ties))) R e;
        R order_emit(c, z, perm, n, false);
    }
    // lib_order_basic (h_primitive.cc:1760-1767)
    if(is_basic(x->typ) && y->typ == HPA_B && y->len == 1){
        uint8_t f; if((e = hpaReadPayload(y, &f, 1))) R e;
        R order_vector(c, z, x, f != 0);
    }
    R HPA_E_DOMAIN;
}
