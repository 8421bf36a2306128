// Sorting, grouping and ordering on a stable LSD radix sort.
//
// The reference groups and orders with serial radix/merge/quick sorts (backend/h_libs.cc:375-736,
// 1184-1403, 1912-1940; h_primitive.cc:993-1057, 1760-1816, 2656-2681).  Every group and order
// output is a function of a *stable* sort on a per-case key:
//   - i32 group / i32 order:  unsigned 32-bit key (the bit pattern: the reference's byte-digit
//                             LSD on two's complement bytes orders the same way)
//   - i64 group / i64 order:  unsigned 64-bit key, or its low 32 bits when no key has
//                             (x >> 32) > 0 (lib_radixsort_core_long, h_libs.cc:1240-1266)
//   - two i32 columns, equal flags: packed ((x0 << 32) | x1), x1 sign-extended
//   - char pair group: (c0 << 8) + c1 bucket order (pfnGroupTrie)
//   - other types: an order-preserving key of the type's `<`
// Descending single-column and packed orders reverse the stable ascending result, so ties come
// out in descending row order as in the reference.
//
// Radix pass: 4096-key tiles; per-tile digit histogram; exclusive scan in (digit, tile) order;
// scatter with each key's stable rank inside its tile computed from one wavefront ballot per
// digit bit (lanes sharing the digit), staged through LDS so the writes go out as digit runs.
// 32-bit keys take digits of up to 10 bits (a 20-bit key range sorts in two passes), 64-bit keys
// bytes; digits that are constant over all keys are skipped.  (A single-pass variant with decoupled look-back
// between tiles was measured slower here: its look-back chains wait on device-scope atomics.)
#include "hpa_internal.h"

#define RS_BLOCK 256
#define RS_ROWS  16
#define RS_TILE  (RS_BLOCK * RS_ROWS)

// Digit histogram of one 4096-key tile (2^DB bins; the digit is (key >> shift) & dmask, dmask at
// most 2^DB - 1), counted in per-wave LDS histograms; a wave whose valid keys all share the digit
// adds them with one increment.  hist[tile * 2^DB + d]: each tile writes one contiguous row.  With
// `bits` set (the first pass) it also writes the tile's OR and AND of all keys, and whether some
// key is below the key before it (on the sort's bits, `kmask`), to bits[4*tile .. 4*tile+2]:
// which key bits vary, and whether the input is already in key order.
template<typename K, int DB>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const K *keys, L n, int shift, unsigned dmask, uint32_t *hist, L ntiles,
                                                      unsigned long long *bits, K kmask){
    constexpr int NB = 1 << DB;
    __shared__ uint32_t cnt[4][NB];
    __shared__ unsigned long long so[RS_BLOCK / 64], sa[RS_BLOCK / 64];
    __shared__ int sd[RS_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for(int q = 0; q < 4; q++) for(int d = threadIdx.x; d < NB; d += RS_BLOCK) cnt[q][d] = 0;
    __syncthreads();
    const L base = (L)blockIdx.x * RS_TILE + (L)w * (RS_TILE / 4);
    unsigned long long o = 0, a = ~0ull;
    bool desc = false;
    K kr[RS_ROWS];
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){ const L i = base + (L)r * 64 + lane; kr[r] = i < n ? keys[i] : (K)0; }
    // key of the row before this wave's first row (from the wave or tile before)
    K prev = (bits && base > 0 && base <= n) ? keys[base - 1] : (K)0;
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        const L i = base + (L)r * 64 + lane;
        const bool valid = i < n;
        const K k = kr[r];
        const unsigned d = (unsigned)((k >> shift) & (K)dmask);
        const unsigned d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t vm = __ballot(valid);
        if(vm == 0) continue;
        if(__ballot(valid && d == d0) == vm){ if(lane == 0) cnt[w][d0] += (uint32_t)__popcll(vm); }
        else if(valid) atomicAdd(&cnt[w][d], 1u);
        if(valid){ o |= (unsigned long long)k; a &= (unsigned long long)k; }
        if(bits){
            // the row before row i: lane - 1 of this row, or lane 63 of the row before
            const K up = __shfl_up(k, 1, 64);
            const K kp = lane == 0 ? prev : up;
            if(valid && i > 0 && (k & kmask) < (kp & kmask)) desc = true;
            prev = __shfl(k, 63, 64);
        }
    }
    if(bits){
        for(int s = 32; s > 0; s >>= 1){ o |= __shfl_down(o, s, 64); a &= __shfl_down(a, s, 64); }
        const bool anyd = __ballot(desc) != 0;
        if(lane == 0){ so[w] = o; sa[w] = a; sd[w] = anyd ? 1 : 0; }
    }
    __syncthreads();
    for(int d = threadIdx.x; d < NB; d += RS_BLOCK) hist[(L)blockIdx.x * NB + d] = cnt[0][d] + cnt[1][d] + cnt[2][d] + cnt[3][d];
    if(bits && threadIdx.x == 0){
        int dd = sd[0];
        for(int q = 1; q < RS_BLOCK / 64; q++){ o |= so[q]; a &= sa[q]; dd |= sd[q]; }
        bits[4 * blockIdx.x] = o; bits[4 * blockIdx.x + 1] = a; bits[4 * blockIdx.x + 2] = (unsigned long long)dd;
    }
}
// OR / AND / any-descent of the tiles' partials (one workgroup)
__global__ void __launch_bounds__(256) k_rs_bits_final(const unsigned long long *part, L ntiles, unsigned long long *out){
    __shared__ unsigned long long so[4], sa[4], sd[4];
    unsigned long long o = 0, a = ~0ull, dd = 0;
    for(L t = threadIdx.x; t < ntiles; t += 256){ o |= part[4 * t]; a &= part[4 * t + 1]; dd |= part[4 * t + 2]; }
    for(int s = 32; s > 0; s >>= 1){ o |= __shfl_down(o, s, 64); a &= __shfl_down(a, s, 64); dd |= __shfl_down(dd, s, 64); }
    if((threadIdx.x & 63) == 0){ so[threadIdx.x >> 6] = o; sa[threadIdx.x >> 6] = a; sd[threadIdx.x >> 6] = dd; }
    __syncthreads();
    if(threadIdx.x == 0){
        for(int w = 1; w < 4; w++){ o |= so[w]; a &= sa[w]; dd |= sd[w]; }
        out[0] = o; out[1] = a; out[2] = dd;
    }
}

// Scatter of one tile: every key's stable rank inside its wave's quarter of the tile comes from
// DB ballots (lanes sharing the digit); the tile is then laid out digit by digit in LDS and
// written out in that order, so consecutive threads write consecutive addresses of one digit's
// run instead of 4096 scattered 4/8-byte stores.  off[tile * 2^DB + d]: output start of the
// tile's keys of digit d (positions fit 32 bits: sorts take fewer than 2^32 keys).  Block b takes
// tile xcd_tile(b): the tiles in flight on one XCD are neighbours, and the runs they write for one
// digit are adjacent in the output.  MODE 0: keys and values to kout / vout; MODE 1: only the
// values, widened to i64 into out64 at position pos, or n - 1 - pos with `reverse`; MODE 2: the
// keys to kout and the widened values to out64 (no reversal).
template<typename K, int MODE, int DB>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                                                         int64_t *out64, int reverse, L n, int shift, unsigned dmask,
                                                         const uint32_t *off, L ntiles){
    constexpr int NB = 1 << DB, PER = NB / RS_BLOCK;   // digits per thread in the tile's digit scan
    // LDS kept small for occupancy: 16-bit per-wave counts (at most 4096), one staging buffer
    // that holds the tile's keys and then its values
    __shared__ uint16_t wcnt[4][NB];           // per-wave digit counts, then per-wave digit starts
    __shared__ uint32_t gbase[NB];             // output position of a digit's run minus its tile-local start (>= 0)
    __shared__ uint32_t wsum[RS_BLOCK / 64];
    __shared__ K sk[RS_TILE];
    uint32_t *sv = (uint32_t*)sk;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for(int q = 0; q < 4; q++) for(int d = threadIdx.x; d < NB; d += RS_BLOCK) wcnt[q][d] = 0;
    __syncthreads();
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const L t0 = (L)tile * RS_TILE;
    const L tbase = t0 + (L)w * (RS_TILE / 4);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    K kr[RS_ROWS]; uint32_t vr[RS_ROWS]; uint32_t rk[RS_ROWS];
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        bool valid = i < n;
        kr[r] = valid ? kin[i] : (K)0;
        vr[r] = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        bool valid = i < n;
        unsigned d = (unsigned)((kr[r] >> shift) & (K)dmask);
        const uint64_t peers = match_digit(d, valid, DB);
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cur = wcnt[w][d];
        if(valid && before == 0) wcnt[w][d] = (uint16_t)(cur + (uint32_t)__popcll(peers));
        rk[r] = cur + before;
    }
    __syncthreads();
    {
        // tile-local start of each digit (exclusive scan over the digit totals; thread t owns
        // digits [t * PER, t * PER + PER)), then the start of each wave's share inside its digit
        uint32_t tot = 0;
        #pragma unroll
        for(int j = 0; j < PER; j++){ const int d = threadIdx.x * PER + j; tot += (uint32_t)wcnt[0][d] + wcnt[1][d] + wcnt[2][d] + wcnt[3][d]; }
        uint32_t inc = tot;
        for(int o = 1; o < 64; o <<= 1){ uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        if(lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t pre = inc - tot;
        for(int q = 0; q < w; q++) pre += wsum[q];
        #pragma unroll
        for(int j = 0; j < PER; j++){
            const int d = threadIdx.x * PER + j;
            const uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d], c3 = wcnt[3][d];
            wcnt[0][d] = (uint16_t)pre; wcnt[1][d] = (uint16_t)(pre + c0);
            wcnt[2][d] = (uint16_t)(pre + c0 + c1); wcnt[3][d] = (uint16_t)(pre + c0 + c1 + c2);
            gbase[d] = off[(L)tile * NB + d] - pre;
            pre += c0 + c1 + c2 + c3;
        }
    }
    __syncthreads();
    // keys laid out digit by digit; each thread then takes slots tid + 256 k, keeping their output
    // positions for the values
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        if(i < n){
            unsigned d = (unsigned)((kr[r] >> shift) & (K)dmask);
            sk[wcnt[w][d] + rk[r]] = kr[r];
        }
    }
    __syncthreads();
    const int cnt = (int)(n - t0 < (L)RS_TILE ? n - t0 : (L)RS_TILE);
    uint32_t pos[RS_ROWS];
    #pragma unroll
    for(int k = 0; k < RS_ROWS; k++){
        const int j = threadIdx.x + k * RS_BLOCK;
        if(j < cnt){
            const K key = sk[j];
            pos[k] = gbase[(unsigned)((key >> shift) & (K)dmask)] + (uint32_t)j;
            if(MODE != 1) kout[pos[k]] = key;
        }
    }
    __syncthreads();
    #pragma unroll
    for(int r = 0; r < RS_ROWS; r++){
        L i = tbase + (L)r * 64 + lane;
        if(i < n){
            unsigned d = (unsigned)((kr[r] >> shift) & (K)dmask);
            sv[wcnt[w][d] + rk[r]] = vr[r];
        }
    }
    __syncthreads();
    #pragma unroll
    for(int k = 0; k < RS_ROWS; k++){
        const int j = threadIdx.x + k * RS_BLOCK;
        if(j < cnt){
            const uint32_t v = sv[j];
            const L p = (L)pos[k];
            if(MODE == 1) out64[reverse ? n - 1 - p : p] = (int64_t)v;
            else if(MODE == 2) out64[p] = (int64_t)v;
            else vout[p] = v;
        }
    }
}
// Exclusive scan of tile-major digit counts (NB digits) in digit-major order:
//   off[t * NB + d] = (counts of all digits below d) + (counts of digit d in tiles before t).
// Thread (d, q) of a 1024-thread block owns digit d and the q-th of Q = 1024 / NB slices of the
// block's tiles, so every access is one contiguous row of NB counts, and each thread issues its
// loads together.  Three launches: per-chunk sums of each digit (32 tiles a chunk), chunk bases
// per digit, then each chunk writes its tiles' offsets.
#define DS_PER 32
template<int NB>
__global__ void __launch_bounds__(1024) k_ds_part(const uint32_t *cnt, L ntiles, uint32_t *part){
    constexpr int Q = 1024 / NB, TQ = DS_PER / Q;
    __shared__ uint32_t ps[Q][NB];
    const unsigned d = threadIdx.x % NB, q = threadIdx.x / NB;
    const L t0 = (L)blockIdx.x * DS_PER + q * TQ;
    uint32_t v[TQ];
    #pragma unroll
    for(int j = 0; j < TQ; j++){ const L t = t0 + j; v[j] = t < ntiles ? cnt[t * NB + d] : 0u; }
    uint32_t s = 0;
    #pragma unroll
    for(int j = 0; j < TQ; j++) s += v[j];
    ps[q][d] = s;
    __syncthreads();
    if(q == 0){
        uint32_t t = 0;
        #pragma unroll
        for(int r = 0; r < Q; r++) t += ps[r][d];
        part[(L)blockIdx.x * NB + d] = t;
    }
}
// chunk bases relative to their digit: cbase[c * NB + d] = counts of digit d in chunks before
// c, and tot[d] = all counts of digit d.  Block g takes digits 32g..32g+31; thread (dl, q) sums
// digit 32g + dl over the q-th of 32 slices of the chunks.
template<int NB>
__global__ void __launch_bounds__(1024) k_ds_mid(const uint32_t *part, int nchunks, uint32_t *cbase, uint32_t *tot){
    __shared__ uint32_t qs[32][33];
    const unsigned dl = threadIdx.x & 31, q = threadIdx.x >> 5, d = blockIdx.x * 32 + dl;
    const int cq = (nchunks + 31) / 32, c0 = (int)q * cq, c1 = c0 + cq < nchunks ? c0 + cq : nchunks;
    uint32_t s = 0;
    for(int cb = c0; cb < c1; cb += 8){
        uint32_t v[8];
        #pragma unroll
        for(int j = 0; j < 8; j++) v[j] = cb + j < c1 ? part[(L)(cb + j) * NB + d] : 0u;
        #pragma unroll
        for(int j = 0; j < 8; j++) s += v[j];
    }
    qs[q][dl] = s;
    __syncthreads();
    uint32_t run = 0;
    for(unsigned r = 0; r < q; r++) run += qs[r][dl];
    if(q == 31) tot[d] = run + s;
    for(int cb = c0; cb < c1; cb += 8){
        uint32_t v[8];
        #pragma unroll
        for(int j = 0; j < 8; j++) v[j] = cb + j < c1 ? part[(L)(cb + j) * NB + d] : 0u;
        #pragma unroll
        for(int j = 0; j < 8; j++) if(cb + j < c1){ cbase[(L)(cb + j) * NB + d] = run; run += v[j]; }
    }
}
// off of the chunk's 32 tiles: the digit base (exclusive scan of the NB digit totals, done by
// every block), the chunk base, and the counts of earlier tiles in the chunk
template<int NB>
__global__ void __launch_bounds__(1024) k_ds_down(const uint32_t *cnt, L ntiles, const uint32_t *cbase, const uint32_t *tot,
                                                  uint32_t *off){
    constexpr int Q = 1024 / NB, TQ = DS_PER / Q;
    __shared__ uint32_t ps[Q][NB];
    __shared__ uint32_t wtot[NB / 64], dbase[NB];
    const unsigned d = threadIdx.x % NB, q = threadIdx.x / NB, lane = threadIdx.x & 63;
    const L t0 = (L)blockIdx.x * DS_PER + q * TQ;
    const uint32_t cb = cbase[(L)blockIdx.x * NB + d];
    const uint32_t td = q == 0 ? tot[d] : 0u;
    uint32_t v[TQ];
    #pragma unroll
    for(int j = 0; j < TQ; j++){ const L t = t0 + j; v[j] = t < ntiles ? cnt[t * NB + d] : 0u; }
    uint32_t s = 0;
    #pragma unroll
    for(int j = 0; j < TQ; j++) s += v[j];
    ps[q][d] = s;
    uint32_t inc = 0;
    if(q == 0){
        inc = td;
        for(int o = 1; o < 64; o <<= 1){ const uint32_t t = __shfl_up(inc, o, 64); if((int)lane >= o) inc += t; }
        if(lane == 63) wtot[d >> 6] = inc;
    }
    __syncthreads();
    if(q == 0){
        uint32_t pre = inc - td;
        for(unsigned w = 0; w < (d >> 6); w++) pre += wtot[w];
        dbase[d] = pre;
    }
    __syncthreads();
    uint32_t run = dbase[d] + cb;
    for(unsigned r = 0; r < q; r++) run += ps[r][d];
    #pragma unroll
    for(int j = 0; j < TQ; j++){
        const L t = t0 + j;
        if(t < ntiles){ off[t * NB + d] = run; run += v[j]; }
    }
}
static L digit_scan_chunks(L ntiles){ R ceil_div(ntiles, DS_PER); }
// part, cbase: NB * chunks u32 (chunks = digit_scan_chunks), tot: NB u32
template<int NB>
static int32_t digit_scan(hpa_ctx *c, const uint32_t *cnt, L ntiles, uint32_t *part, uint32_t *cbase, uint32_t *tot, uint32_t *off){
    const L nchunks = digit_scan_chunks(ntiles);
    if(nchunks > INT32_MAX){ set_error_detail("digit scan of %lld tiles", (long long)ntiles); R HPA_E_NOT_IMPL; }
    ProfScope ps("radix_scan", c->stream, 12.0 * NB * ntiles);
    hipLaunchKernelGGL(k_ds_part<NB>, dim3((unsigned)nchunks), dim3(1024), 0, c->stream, cnt, ntiles, part);
    hipLaunchKernelGGL(k_ds_mid<NB>, dim3(NB / 32), dim3(1024), 0, c->stream, (const uint32_t*)part, (int)nchunks, cbase, tot);
    hipLaunchKernelGGL(k_ds_down<NB>, dim3((unsigned)nchunks), dim3(1024), 0, c->stream, cnt, ntiles, (const uint32_t*)cbase,
                       (const uint32_t*)tot, off);
    R check_launch("radix_scan");
}

__global__ void k_iota_i64(int64_t *z, L n, int reverse){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) z[i] = reverse ? n - 1 - i : i;
}

// One LSD pass over digit (key >> s) & dmask with 2^DB bins: histogram (unless `have_hist`), digit
// scan, scatter (to kb / vb; with out64 the values to out64 as i64, and the keys to kb when kb is
// set).
template<typename K, int DB>
static int32_t radix_pass(hpa_ctx *c, const K *ka, const uint32_t *va, K *kb, uint32_t *vb, int64_t *out64, bool reverse,
                          L n, L ntiles, int s, unsigned dmask, bool have_hist, uint32_t *hist, uint32_t *part, uint32_t *cbase,
                          uint32_t *tot, uint32_t *off){
    constexpr int NB = 1 << DB;
    int32_t e;
    if(!have_hist){
        ProfScope ps("radix_hist", c->stream, (double)n * sizeof(K) + 4.0 * NB * ntiles);
        hipLaunchKernelGGL((k_rs_hist<K, DB>), dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream, ka, n, s, dmask, hist, ntiles,
                           (unsigned long long*)nullptr, (K)0);
    }
    if((e = digit_scan<NB>(c, hist, ntiles, part, cbase, tot, off))) R e;
    const double rd = (double)n * (sizeof(K) + (va ? 4 : 0)),
                 wr = out64 ? 8.0 * n + (kb ? (double)n * sizeof(K) : 0.0) : (double)n * (sizeof(K) + 4);
    ProfScope ps("radix_scatter", c->stream, rd + wr);
    if(out64 && kb)
        hipLaunchKernelGGL((k_rs_scatter<K, 2, DB>), dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream,
                           ka, va, kb, (uint32_t*)nullptr, out64, 0, n, s, dmask, (const uint32_t*)off, ntiles);
    else if(out64)
        hipLaunchKernelGGL((k_rs_scatter<K, 1, DB>), dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream,
                           ka, va, (K*)nullptr, (uint32_t*)nullptr, out64, reverse ? 1 : 0, n, s, dmask, (const uint32_t*)off, ntiles);
    else
        hipLaunchKernelGGL((k_rs_scatter<K, 0, DB>), dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream,
                           ka, va, kb, vb, (int64_t*)nullptr, 0, n, s, dmask, (const uint32_t*)off, ntiles);
    R check_launch("radix_sort");
}

// Stable sort of (keys, vals) by key bits [lo, hi); vals may be null (row ids 0..n-1).  The
// inputs are only read.  Without out64 the sorted keys and values come back in *kout / *vout
// (the input keys, and the input values or fresh row ids, when every digit is constant).  With
// out64 the sorted values are written there as i64, reversed with `reverse`, and *kout / *vout
// are not set -- unless `keep_keys` (no reversal): the sorted keys then come back in *kout too.
//
// Digits: 32-bit keys are sorted with digits of up to RS_WIDE bits (1024 bins), 64-bit keys with
// bytes.  The first histogram pass, over the lowest digit, also finds the key bits that vary and
// whether the keys are already in order; the remaining varying bits are then covered by as few
// passes as possible (bytes when wide digits would not save a pass), and digits whose bits are
// all constant are skipped.  Each pass: per-tile digit histograms, a scan in (digit, tile) order,
// a scatter.
#define RS_WIDE 10
// hint (optional): the OR and AND of all keys and whether some key is below the one before it,
// known from an earlier pass -- the first histogram pass then finds nothing and no read is made
struct RsHint { unsigned long long orv, andv; bool desc; };
template<typename K>
static int32_t radix_sort_t(hpa_ctx *c, const K *keys, const uint32_t *vals, L n, int lo, int hi, K **kout, uint32_t **vout,
                            int64_t *out64, bool reverse, bool keep_keys = false, const RsHint *hint = nullptr){
    if(keep_keys && (!out64 || reverse)) R HPA_E_DOMAIN;
    if(n >= (L)UINT32_MAX){ set_error_detail("radix sort of %lld keys exceeds 32-bit row ids", (long long)n); R HPA_E_NOT_IMPL; }
    int32_t e;
    if(n <= 0){ if(!out64){ *kout = (K*)keys; *vout = (uint32_t*)vals; } else if(keep_keys) *kout = (K*)keys; R 0; }
    constexpr int kw = (int)(8 * sizeof(K)), wide = sizeof(K) == 4 ? RS_WIDE : 8, NBmax = 1 << wide;
    if(lo < 0 || hi > kw || lo >= hi){ set_error_detail("radix sort of key bits [%d, %d)", lo, hi); R HPA_E_DOMAIN; }
    const L ntiles = ceil_div(n, RS_TILE);
    unsigned long long *bits = (unsigned long long*)dalloc(c, 32 + 32 * ntiles);
    uint32_t *hist = (uint32_t*)dalloc(c, 4 * NBmax * ntiles);
    uint32_t *off = (uint32_t*)dalloc(c, 4 * NBmax * ntiles);
    uint32_t *part = (uint32_t*)dalloc(c, 4 * NBmax * digit_scan_chunks(ntiles));
    uint32_t *cbase = (uint32_t*)dalloc(c, 4 * NBmax * digit_scan_chunks(ntiles));
    uint32_t *tot = (uint32_t*)dalloc(c, 4 * NBmax);
    if(!bits || !hist || !off || !part || !cbase || !tot) R HPA_E_MEMORY;
    const K kmask = (K)((hi - lo >= kw ? ~(K)0 : (((K)1 << (hi - lo)) - 1)) << lo);
    auto dmask_of = [&](int s, int db) -> unsigned { const int w = hi - s < db ? hi - s : db; R (1u << w) - 1u; };
    unsigned long long hb[3];
    if(hint){ hb[0] = hint->orv; hb[1] = hint->andv; hb[2] = hint->desc ? 1 : 0; }
    else {
        {
            ProfScope ps("radix_hist", c->stream, (double)n * sizeof(K) + 4.0 * NBmax * ntiles);
            hipLaunchKernelGGL((k_rs_hist<K, wide>), dim3((unsigned)ntiles), dim3(RS_BLOCK), 0, c->stream, keys, n, lo, dmask_of(lo, wide),
                               hist, ntiles, bits + 4, kmask);
            hipLaunchKernelGGL(k_rs_bits_final, dim3(1), dim3(256), 0, c->stream, bits + 4, ntiles, bits);
        }
        if((e = check_launch("radix_hist"))) R e;
        if((e = d2h(c, hb, bits, 24))) R e;
    }
    const unsigned long long vary = (hb[0] ^ hb[1]) & (unsigned long long)kmask;
    if(vary == 0 || hb[2] == 0){
        // every digit is constant, or no key is below the one before: the input order is the
        // stable sorted order
        if(out64){
            if(keep_keys) *kout = (K*)keys;
            hipLaunchKernelGGL(k_iota_i64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, out64, n, reverse ? 1 : 0);
            R check_launch("radix_sort");
        }
        *kout = (K*)keys;
        if(vals){ *vout = (uint32_t*)vals; R 0; }
        uint32_t *v1 = (uint32_t*)dalloc(c, 4 * n); if(!v1) R HPA_E_MEMORY;
        if((e = iota_u32(c, v1, n))) R e;
        *vout = v1; R 0;
    }
    // the digit plan: the first digit [lo, lo + wide), then digits of `step` bits up to the top
    // varying bit
    const int vhi = 64 - __builtin_clzll(vary);
    struct Digit { int s, db; };
    Digit plan[64]; int nrun = 0;
    if((vary >> lo) & dmask_of(lo, wide)) plan[nrun++] = {lo, wide};
    const int r0 = lo + wide, rest = vhi - r0;
    if(rest > 0){
        const int step = (int)ceil_div(rest, 8) == (int)ceil_div(rest, wide) ? 8 : wide;
        for(int s = r0; s < vhi; s += step) if((vary >> s) & dmask_of(s, step)) plan[nrun++] = {s, step};
    }
    K *kb[2] = {nullptr, nullptr}; uint32_t *vb[2] = {nullptr, nullptr};
    const int nout = out64 ? nrun - 1 : nrun, nbuf = nout < 2 ? nout : 2, nkb = keep_keys ? (nrun < 2 ? nrun : 2) : nbuf;
    for(int q = 0; q < nkb; q++){ kb[q] = (K*)dalloc(c, sizeof(K) * n); if(!kb[q]) R HPA_E_MEMORY; }
    for(int q = 0; q < nbuf; q++){ vb[q] = (uint32_t*)dalloc(c, 4 * n); if(!vb[q]) R HPA_E_MEMORY; }
    const K *ka = keys; const uint32_t *va = vals;
    for(int r = 0; r < nrun; r++){
        const int s = plan[r].s, db = plan[r].db;
        const bool last = r == nrun - 1, fin = last && out64 != nullptr;
        const bool have = !hint && r == 0 && s == lo && db == wide;
        K *kd = fin && !keep_keys ? nullptr : kb[r & 1];
        e = db == 8 ? radix_pass<K, 8>(c, ka, va, kd, vb[r & 1], fin ? out64 : nullptr, reverse, n, ntiles, s, dmask_of(s, 8),
                                       have, hist, part, cbase, tot, off)
                    : radix_pass<K, wide>(c, ka, va, kd, vb[r & 1], fin ? out64 : nullptr, reverse, n, ntiles, s,
                                          dmask_of(s, wide), have, hist, part, cbase, tot, off);
        if(e) R e;
        if(!fin){ ka = kb[r & 1]; va = vb[r & 1]; }
        else if(keep_keys) *kout = kd;
    }
    if(!out64){ *kout = (K*)ka; *vout = (uint32_t*)va; }
    R 0;
}
int32_t radix_sort_pairs_u64(hpa_ctx *c, const uint64_t *keys, const uint32_t *vals, L n, int lo, int hi, uint64_t **ko, uint32_t **vo){
    R radix_sort_t<uint64_t>(c, keys, vals, n, lo, hi, ko, vo, nullptr, false);
}
int32_t radix_sort_pairs_u32(hpa_ctx *c, const uint32_t *keys, const uint32_t *vals, L n, int lo, int hi, uint32_t **ko, uint32_t **vo){
    R radix_sort_t<uint32_t>(c, keys, vals, n, lo, hi, ko, vo, nullptr, false);
}
// the sorted keys in *ko and the stable permutation (row ids) as i64 in flat[0, n)
static int32_t radix_sort_flat_u32(hpa_ctx *c, const uint32_t *keys, L n, int lo, int hi, uint32_t **ko, int64_t *flat,
                                   const RsHint *hint = nullptr){
    uint32_t *vo = nullptr;
    R radix_sort_t<uint32_t>(c, keys, nullptr, n, lo, hi, ko, &vo, flat, false, true, hint);
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
__device__ __forceinline__ K shfl_up1(K v){
    if(sizeof(K) == 8){
        const uint64_t u = (uint64_t)v;
        const uint32_t lo = __shfl_up((uint32_t)u, 1, 64), hi = __shfl_up((uint32_t)(u >> 32), 1, 64);
        return (K)(((uint64_t)hi << 32) | lo);
    }
    return (K)__shfl_up((uint32_t)v, 1, 64);
}
// run head of row i = iw + lane where the lanes of the wave hold consecutive rows from iw (iw
// wave-uniform, n > 0): the previous key comes from the lane below, lane 0's from one
// wave-uniform load.  Keys are loaded at indices clamped to the last row instead of under a
// branch, so the loads of a thread's rows are in flight together.  Every lane of the wave must
// call it; *cur is row i's key (the last row's beyond n).
template<typename K>
__device__ __forceinline__ bool run_head_w(const K *sk, const uint32_t *sidx, const uint64_t *eq, L i, L iw, L n, K *cur){
    if(eq){ *cur = (K)0; return i < n && run_head(sk, sidx, eq, i); }
    const K k = sk[i < n ? i : n - 1];
    const K pw = iw > 0 ? sk[iw - 1 < n ? iw - 1 : n - 1] : (K)0;
    K prev = shfl_up1(k);
    if((threadIdx.x & 63) == 0) prev = pw;
    *cur = k;
    return i < n && (i == 0 || k != prev);
}
// EQ: runs of equal eq[sidx[i]] (the key words behind the sorted row ids) instead of equal sk[i]
// The run heads of ITEMS steps of 256 consecutive rows (row base + 256 k + threadIdx.x): every
// key load is issued before the first comparison.  bal[k]: the wave's heads of step k; kv[k]:
// the thread's key of step k.
template<typename K, int ITEMS>
__device__ __forceinline__ void run_heads(const K *sk, L base, L n, uint64_t *bal, K *kv){
    const L wbase = base + (L)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    K pw[ITEMS];
    #pragma unroll
    for(int k = 0; k < ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x, iw = wbase + (L)k * 256;
        kv[k] = sk[i < n ? i : n - 1];
        pw[k] = iw > 0 ? sk[iw - 1 < n ? iw - 1 : n - 1] : (K)0;
    }
    #pragma unroll
    for(int k = 0; k < ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x;
        K prev = shfl_up1(kv[k]);
        if((threadIdx.x & 63) == 0) prev = pw[k];
        bal[k] = __ballot(i < n && (i == 0 || kv[k] != prev));
    }
}
template<typename K, bool EQ>
__global__ void __launch_bounds__(256) k_run_tilecount(const K *sk, const uint32_t *sidx, const uint64_t *eq, L n, uint32_t *tcount){
    const L base = (L)blockIdx.x * RUN_TILE;
    const L wbase = base + (L)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    uint32_t c = 0;
    if(EQ){
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++){
            const L i = base + (L)k * 256 + threadIdx.x;
            K cur;
            c += run_head_w(sk, sidx, eq, i, wbase + (L)k * 256, n, &cur) ? 1u : 0u;
        }
    } else {
        uint64_t bal[RUN_ITEMS]; K kv[RUN_ITEMS];
        run_heads<K, RUN_ITEMS>(sk, base, n, bal, kv);
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++) c += (uint32_t)((bal[k] >> (threadIdx.x & 63)) & 1ull);
    }
    __shared__ uint32_t sh[4];
    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0) tcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
// (with flat_in: the flat list is already written -- the sort's last pass wrote it -- and a
// head's key is read from it)
template<typename K, bool EQ>
__global__ void __launch_bounds__(256) k_run_write(const K *sk, const uint32_t *sidx, const uint64_t *eq, L n,
                                                   const int64_t *toff, int64_t *keys, int64_t *start, int64_t *flat,
                                                   uint32_t *kval, const int64_t *flat_in){
    __shared__ uint32_t s_cnt[RUN_ITEMS * 4];
    const L base = (L)blockIdx.x * RUN_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const L wbase = base + (L)__builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
    uint64_t bal[RUN_ITEMS];
    K kv[RUN_ITEMS];
    int64_t hv[RUN_ITEMS];                      // each head's key: its first row's original index
    // every load of the step first, then the stores (which may alias the inputs for the compiler)
    if(!flat_in){
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; hv[k] = (int64_t)sidx[i < n ? i : n - 1]; }
    }
    if(EQ){
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++){
            const L i = base + (L)k * 256 + threadIdx.x;
            bal[k] = __ballot(run_head_w(sk, sidx, eq, i, wbase + (L)k * 256, n, &kv[k]));
        }
    } else run_heads<K, RUN_ITEMS>(sk, base, n, bal, kv);
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++) if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    if(!flat_in){
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; if(i < n) flat[i] = hv[k]; }
    }
    // with flat_in, the heads' entries are read together: lanes that are no head read the wave's
    // first row of the step (a line the heads' loads share or one more), so nothing waits under a branch
    if(flat_in){
        #pragma unroll
        for(int k = 0; k < RUN_ITEMS; k++){
            const L i = base + (L)k * 256 + threadIdx.x, iw = wbase + (L)k * 256;
            const L j = ((bal[k] >> lane) & 1ull) ? i : (iw < n ? iw : n - 1);
            hv[k] = flat_in[j];
        }
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
            keys[r] = hv[k]; start[r] = i;
            if(kval) kval[r] = (uint32_t)kv[k];
        }
    }
}
// cnt[r] = start[r + 1] - start[r] with start[g] = n, which is written too (start holds g + 1
// entries: the groups' offsets in the flat list)
__global__ void k_run_counts(int64_t *start, L g, L n, int64_t *cnt){
    for(L r = (L)blockIdx.x * blockDim.x + threadIdx.x; r < g; r += (L)gridDim.x * blockDim.x){
        cnt[r] = (r + 1 < g ? start[r + 1] : n) - start[r];
        if(r == 0) start[g] = n;
    }
}
// the same with the group count read on the device (*gp), for outputs allocated at a bound
__global__ void k_run_counts_dev(int64_t *start, const int64_t *gp, L n, int64_t *cnt){
    const L g = *gp;
    for(L r = (L)blockIdx.x * blockDim.x + threadIdx.x; r < g; r += (L)gridDim.x * blockDim.x)
        cnt[r] = (r + 1 < g ? start[r + 1] : n) - start[r];
    if(blockIdx.x == 0 && threadIdx.x == 0) start[g] = n;
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

// src: the i32 column whose bit patterns sk holds (null otherwise); its groups' key values are
// then kept beside the keys for index(column, keys)
// gmax > 0: a bound on the group count; when the arena holds the outputs at that size, every pass
// is queued at once and the group count is read once at the end
template<typename K>
static int32_t runs_to_group_t(hpa_ctx *c, V z, const K *sk, const uint32_t *sidx, L n, const uint64_t *eq, const void *src = nullptr,
                               int64_t *flat_in = nullptr, L gmax = 0){
    int32_t e;
    const L nt = ceil_div(n, RUN_TILE) > 0 ? ceil_div(n, RUN_TILE) : 1;
    uint32_t *tcount = (uint32_t*)dalloc(c, 4 * nt);
    int64_t *toff = (int64_t*)dalloc(c, 8 * (nt + 1));
    if(!tcount || !toff) R HPA_E_MEMORY;
    if(gmax > 0 && n > 0 && c->dcur + 36 * gmax + (flat_in ? 0 : 8 * n) + 4096 <= c->dtop){
        uint32_t *kval = src ? (uint32_t*)dalloc(c, 4 * gmax) : nullptr;          // before keys: dies with them
        int64_t *keys = (int64_t*)dalloc(c, 8 * gmax), *start = (int64_t*)dalloc(c, 8 * (gmax + 1));
        int64_t *cnt = (int64_t*)dalloc(c, 8 * gmax);
        int64_t *flat = flat_in ? flat_in : (int64_t*)dalloc(c, 8 * n);
        if(!keys || !start || !cnt || !flat || (src && !kval)) R HPA_E_MEMORY;
        if(eq) hipLaunchKernelGGL((k_run_tilecount<K, true>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, tcount);
        else hipLaunchKernelGGL((k_run_tilecount<K, false>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, tcount);
        if((e = scan_u32_to_i64(c, tcount, toff, nt, toff + nt))) R e;
        if(eq) hipLaunchKernelGGL((k_run_write<K, true>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, (const int64_t*)toff,
                                  keys, start, flat, kval, (const int64_t*)flat_in);
        else hipLaunchKernelGGL((k_run_write<K, false>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, (const int64_t*)toff,
                                keys, start, flat, kval, (const int64_t*)flat_in);
        hipLaunchKernelGGL(k_run_counts_dev, dim3(grid_for(gmax, 256, 1)), dim3(256), 0, c->stream, start, (const int64_t*)(toff + nt), n, cnt);
        if((e = check_launch("group_runs"))) R e;
        int64_t g = 0; if((e = d2h(c, &g, toff + nt, 8))) R e;
        if(g > gmax){ set_error_detail("group count %lld above its bound %lld", (long long)g, (long long)gmax); R HPA_E_DOMAIN; }
        if(kval && g > 1) note_group_vals(c, keys, src, n, kval);
        note_group_offsets(c, cnt, g, start);
        R make_group_dict(c, z, keys, cnt, flat, g, n);
    }
    int64_t g = 0;
    if(n > 0){
        if(eq) hipLaunchKernelGGL((k_run_tilecount<K, true>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, tcount);
        else hipLaunchKernelGGL((k_run_tilecount<K, false>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, tcount);
        if((e = scan_u32_to_i64(c, tcount, toff, nt, toff + nt))) R e;
        if((e = d2h(c, &g, toff + nt, 8))) R e;
    }
    uint32_t *kval = src && g > 1 ? (uint32_t*)dalloc(c, 4 * g) : nullptr;       // before keys: dies with them
    int64_t *keys = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *start = (int64_t*)dalloc(c, 8 * (g + 1));
    int64_t *cnt = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *flat = flat_in ? flat_in : (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1));
    if(!keys || !start || !cnt || !flat) R HPA_E_MEMORY;
    if(n > 0){
        if(eq) hipLaunchKernelGGL((k_run_write<K, true>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, toff, keys, start,
                                  flat, kval, (const int64_t*)flat_in);
        else hipLaunchKernelGGL((k_run_write<K, false>), dim3((unsigned)nt), dim3(256), 0, c->stream, sk, sidx, eq, n, toff, keys, start,
                                flat, kval, (const int64_t*)flat_in);
        hipLaunchKernelGGL(k_run_counts, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, start, g, n, cnt);
    }
    if((e = check_launch("group_runs"))) R e;
    if(kval) note_group_vals(c, keys, src, n, kval);
    note_group_offsets(c, cnt, g, start);
    R make_group_dict(c, z, keys, cnt, flat, g, n);
}
int32_t runs_to_group_u64(hpa_ctx *c, V z, const uint64_t *sk, const uint32_t *sidx, L n, const uint64_t *eq){
    R runs_to_group_t<uint64_t>(c, z, sk, sidx, n, eq);
}
int32_t runs_to_group_u32(hpa_ctx *c, V z, const uint32_t *sk, const uint32_t *sidx, L n, const void *src){
    R runs_to_group_t<uint32_t>(c, z, sk, sidx, n, nullptr, src);
}
// the same over a flat list the sort already wrote (radix_sort_flat_u32)
static int32_t runs_to_group_flat_u32(hpa_ctx *c, V z, const uint32_t *sk, int64_t *flat, L n, const void *src, L gmax = 0){
    R runs_to_group_t<uint32_t>(c, z, sk, nullptr, n, nullptr, src, flat, gmax);
}

// ---------------------------------------------------------------- key builders
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
    if(__any(f) && (threadIdx.x & 63) == 0 && !*(volatile int*)flag) atomicOr(flag, 1);
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

// ---------------------------------------------------------------- small-code group
// Groups whose keys map onto at most 256 dense bins in key order are one stable counting sort
// straight from the key columns, with no key buffer, radix passes or run detection:
//   trie (two char columns): bin = rank(c0) * nb + rank(c1), ranks among the byte values present,
//                            nb = number of distinct c1 values (ascending (c0<<8)+c1 order);
//   i32: bin = (uint32)x - min, when max - min < 256 in the unsigned order the radix key uses.
// Bins with no rows produce no group.  Passes: presence / range, per-tile bin counts (one read),
// scan of the (bin, tile) matrix, scatter of row numbers staged in LDS by bin so the 8-byte
// writes go out as runs, and a one-block pass for the group keys and counts.
// On large inputs the bins come from a strided sample (one 16-row chunk in SG_SAMPLE_STRIDE);
// the count pass checks every row against them and raises a flag on a row outside the bins,
// the scatter then does nothing and the bins are rebuilt from the whole input.
#define SG_TILE 4096
#define SG_BINS 256
#define SG_ABSENT 0x100u                // rank-table entry of a byte value outside the bins
#define SG_MISS 0x1000u                 // bin of a row outside the bins
#define SG_SAMPLE_MIN (1 << 20)
#define SG_SAMPLE_STRIDE 64
struct SgSrc {
    const uint8_t *a, *b;        // trie columns (a null: i32 source)
    const int32_t *x;
    uint32_t base;               // i32: smallest key in the unsigned order
    uint32_t nb;                 // trie: distinct c1 values
    uint32_t nbins;
    int vec;                     // 16-byte aligned sources: vector loads
    uint16_t ra[256], rb[256];   // trie: dense rank of each c0 / c1 byte value, or SG_ABSENT
};
__device__ __forceinline__ unsigned sg_trie_bin(const uint16_t *ra, const uint16_t *rb, unsigned nb, unsigned va, unsigned vb){
    const unsigned r1 = ra[va], r2 = rb[vb];
    R ((r1 | r2) & SG_ABSENT) ? SG_MISS : r1 * nb + r2;
}
__device__ __forceinline__ unsigned sg_i32_bin(const SgSrc &s, unsigned x){
    const unsigned d = x - s.base;
    R d < s.nbins ? d : SG_MISS;
}
// The raw key bytes of one thread's 16 consecutive rows of a tile travel in four uint4
// registers q0..q3 (trie: q0 = c0 bytes, q1 = c1 bytes; i32: the 16 keys).
__device__ __forceinline__ bool sg_full(const SgSrc &s, L t0, L n){ R s.vec && t0 + (L)(threadIdx.x + 1) * 16 <= n; }
__device__ __forceinline__ void sg_fetch(const SgSrc &s, L t0, L n, uint4 &q0, uint4 &q1, uint4 &q2, uint4 &q3){
    const L r0 = t0 + (L)threadIdx.x * 16;
    if(!sg_full(s, t0, n)) R;
    if(s.a){ q0 = *(const uint4*)(s.a + r0); q1 = *(const uint4*)(s.b + r0); }
    else {
        q0 = *(const uint4*)(s.x + r0); q1 = *(const uint4*)(s.x + r0 + 4);
        q2 = *(const uint4*)(s.x + r0 + 8); q3 = *(const uint4*)(s.x + r0 + 12);
    }
}
__device__ __forceinline__ unsigned sg_word(const uint4 &q0, const uint4 &q1, const uint4 &q2, const uint4 &q3, int j){
    const uint4 v = (j >> 2) == 0 ? q0 : (j >> 2) == 1 ? q1 : (j >> 2) == 2 ? q2 : q3;      // j: a constant after unrolling
    R (j & 3) == 0 ? v.x : (j & 3) == 1 ? v.y : (j & 3) == 2 ? v.z : v.w;
}
__device__ __forceinline__ unsigned sg_byte(const uint4 &v, int j){
    const unsigned w = j < 4 ? v.x : j < 8 ? v.y : j < 12 ? v.z : v.w;
    R (w >> (8 * (j & 3))) & 255u;
}
// dense bins of the thread's 16 rows into sd[16 * threadIdx.x ...] (rows past n and rows outside
// the bins: bin 0; the latter set *miss)
__device__ __forceinline__ void sg_bins(const SgSrc &s, const uint16_t *ra, const uint16_t *rb, L t0, L n,
                                        const uint4 &q0, const uint4 &q1, const uint4 &q2, const uint4 &q3, uint8_t *sd, bool *miss){
    const L r0 = t0 + (L)threadIdx.x * 16;
    unsigned pk[4] = {0, 0, 0, 0};
    const bool full = sg_full(s, t0, n);
    #pragma unroll
    for(int j = 0; j < 16; j++){
        unsigned v;
        if(s.a) v = full ? sg_trie_bin(ra, rb, s.nb, sg_byte(q0, j), sg_byte(q1, j))
                         : (r0 + j < n ? sg_trie_bin(ra, rb, s.nb, s.a[r0 + j], s.b[r0 + j]) : 0u);
        else v = full ? sg_i32_bin(s, sg_word(q0, q1, q2, q3, j)) : (r0 + j < n ? sg_i32_bin(s, (uint32_t)s.x[r0 + j]) : 0u);
        if(v == SG_MISS){ *miss = true; v = 0; }
        pk[j >> 2] |= (v & 255u) << (8 * (j & 3));
    }
    *(uint4*)(sd + threadIdx.x * 16) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
}
// lanes of this wave holding the same bin (kb ballots), among valid lanes
__device__ __forceinline__ uint64_t sg_peers(unsigned d, bool valid, int kb){ R match_digit(d, valid, kb); }
// 9..256 bins: one 4096-row tile per workgroup; a row's stable rank inside its wave's quarter of
// the tile comes from kb ballots (lanes sharing the bin)
__global__ void __launch_bounds__(256) k_sg_hist(SgSrc s, L n, int nbins, int kb, uint32_t *cnt, L ntiles, int *miss){
    __shared__ __align__(16) uint8_t sd[SG_TILE];
    __shared__ uint32_t wc[4][SG_BINS];
    __shared__ uint16_t ra[256], rb[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const L t0 = (L)blockIdx.x * SG_TILE;
    uint4 q0 = make_uint4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
    sg_fetch(s, t0, n, q0, q1, q2, q3);
    ra[threadIdx.x] = s.ra[threadIdx.x]; rb[threadIdx.x] = s.rb[threadIdx.x];
    for(int q = 0; q < 4; q++) wc[q][threadIdx.x] = 0;
    __syncthreads();
    bool m = false;
    sg_bins(s, ra, rb, t0, n, q0, q1, q2, q3, sd, &m);
    if(__syncthreads_or(m)){ if(threadIdx.x == 0) atomicOr(miss, 1); R; }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for(int r = 0; r < SG_TILE / 256; r++){
        const int j = w * (SG_TILE / 4) + r * 64 + lane;
        const bool valid = t0 + j < n;
        const unsigned d = sd[j];
        const uint64_t peers = sg_peers(d, valid, kb);
        if(valid && (peers & lt) == 0) wc[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for(int d = threadIdx.x; d < nbins; d += 256) cnt[(L)d * ntiles + blockIdx.x] = wc[0][d] + wc[1][d] + wc[2][d] + wc[3][d];
}
__global__ void __launch_bounds__(256) k_sg_scatter(SgSrc s, L n, int nbins, int kb, const int64_t *off, L ntiles, int64_t *flat,
                                                    const int *miss){
    __shared__ __align__(16) uint8_t sd[SG_TILE];
    __shared__ uint16_t srow[SG_TILE];
    __shared__ uint32_t wc[4][SG_BINS];
    __shared__ int64_t gbase[SG_BINS];
    __shared__ uint32_t wsum[4];
    __shared__ uint16_t ra[256], rb[256];
    if(*miss) R;                                    // bins from a sample missed a row: nothing to place
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const L t0 = (L)blockIdx.x * SG_TILE;
    uint4 q0 = make_uint4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
    sg_fetch(s, t0, n, q0, q1, q2, q3);
    ra[threadIdx.x] = s.ra[threadIdx.x]; rb[threadIdx.x] = s.rb[threadIdx.x];
    for(int q = 0; q < 4; q++) wc[q][threadIdx.x] = 0;
    __syncthreads();
    bool m = false;
    sg_bins(s, ra, rb, t0, n, q0, q1, q2, q3, sd, &m);
    __syncthreads();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint16_t rk[SG_TILE / 256];
    #pragma unroll
    for(int r = 0; r < SG_TILE / 256; r++){
        const int j = w * (SG_TILE / 4) + r * 64 + lane;
        const bool valid = t0 + j < n;
        const unsigned d = sd[j];
        const uint64_t peers = sg_peers(d, valid, kb);
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        const uint32_t cur = wc[w][d];
        if(valid && before == 0) wc[w][d] = cur + (uint32_t)__popcll(peers);
        rk[r] = (uint16_t)(cur + before);
    }
    __syncthreads();
    {
        const unsigned d = threadIdx.x;
        const uint32_t c0 = wc[0][d], c1 = wc[1][d], c2 = wc[2][d], c3 = wc[3][d];
        const uint32_t tot = c0 + c1 + c2 + c3;
        uint32_t inc = tot;
        for(int o = 1; o < 64; o <<= 1){ uint32_t u = __shfl_up(inc, o, 64); if(lane >= o) inc += u; }
        if(lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t pre = inc - tot;
        for(int q = 0; q < w; q++) pre += wsum[q];
        wc[0][d] = pre; wc[1][d] = pre + c0; wc[2][d] = pre + c0 + c1; wc[3][d] = pre + c0 + c1 + c2;
        if((int)d < nbins) gbase[d] = off[(L)d * ntiles + blockIdx.x] - (int64_t)pre;
    }
    __syncthreads();
    #pragma unroll
    for(int r = 0; r < SG_TILE / 256; r++){
        const int j = w * (SG_TILE / 4) + r * 64 + lane;
        if(t0 + j < n) srow[wc[w][sd[j]] + rk[r]] = (uint16_t)j;
    }
    __syncthreads();
    const int cnt = (int)(n - t0 < (L)SG_TILE ? n - t0 : (L)SG_TILE);
    for(int j = threadIdx.x; j < cnt; j += 256){
        const int tr = srow[j];
        flat[gbase[sd[tr]] + j] = t0 + tr;
    }
}

// At most 8 bins: a 16384-row tile is four groups of 4096 rows; in group g thread t owns the 16
// consecutive rows g*4096 + 16t ... (each load instruction reads 1 KiB contiguous per wave).  A
// thread counts the bins of each chunk in the 8-bit fields of one 64-bit word (bin d in bits
// 8d..8d+7; a chunk has 16 rows).  For the workgroup scans the word is split into two words of
// 16-bit fields (even bins, odd bins), wide enough for a whole tile.  Adding each bin's tile start
// to the scanned prefix gives, per chunk, the tile-local position of the chunk's first row of
// every bin; a row's stable position is that plus the chunk's rows of its bin already placed.
#define SG8_G 4
#define SG8_TILE (SG8_G * 4096)
// 8-bit fields of bins 0..7 -> 16-bit fields (field i = bin 2i + odd)
__device__ __forceinline__ unsigned long long sg8_half(unsigned long long c8, int odd){
    R (c8 >> (8 * odd)) & 0x00FF00FF00FF00FFull;
}
__device__ __forceinline__ unsigned sg8_f16(unsigned long long w, unsigned i){ R (unsigned)(w >> (16 * i)) & 0xFFFFu; }
// bins of the thread's SG8_G chunks (rows past n: 0xFF; rows outside the bins: 0xFF and *miss),
// counted into c8[g] and, with STORE, kept byte-packed in pk[g] (row j of chunk g in byte j%4 of
// pk[g][j/4]).  All loads are issued before the first is used.
template<bool STORE>
__device__ __forceinline__ void sg8_bins(const SgSrc &s, const uint16_t *ra, const uint16_t *rb, L t0, L n,
                                         unsigned (*pk)[4], unsigned long long *c8, bool *miss){
    #pragma unroll
    for(int g = 0; g < SG8_G; g++) c8[g] = 0;
    auto put = [&](int g, int j, unsigned v){          // g, j: constants after unrolling
        if(v == SG_MISS){ *miss = true; v = 0xFFu; }
        if(v != 0xFFu) c8[g] += 1ull << (8 * v);
        if(STORE){ if((j & 3) == 0) pk[g][j >> 2] = v & 255u; else pk[g][j >> 2] |= (v & 255u) << (8 * (j & 3)); }
    };
    const L rt = t0 + (L)threadIdx.x * 16;
    // the vector loads of all groups go out together: a group past the end loads the last whole
    // aligned 16 rows instead (and is then read row by row), so no load waits under a branch
    const bool vok = s.vec && n >= 16;
    const L rlast = vok ? ((n - 16) & ~(L)15) : 0;
    if(s.a){
        uint4 qa[SG8_G], qb[SG8_G];
        if(vok){
            #pragma unroll
            for(int g = 0; g < SG8_G; g++){
                const L r = rt + g * 4096, rc = r + 16 <= n ? r : rlast;
                qa[g] = *(const uint4*)(s.a + rc); qb[g] = *(const uint4*)(s.b + rc);
            }
        }
        #pragma unroll
        for(int g = 0; g < SG8_G; g++){
            const L r = rt + g * 4096;
            if(s.vec && r + 16 <= n){
                #pragma unroll
                for(int j = 0; j < 16; j++) put(g, j, sg_trie_bin(ra, rb, s.nb, sg_byte(qa[g], j), sg_byte(qb[g], j)));
            } else {
                #pragma unroll
                for(int j = 0; j < 16; j++) put(g, j, r + j < n ? sg_trie_bin(ra, rb, s.nb, s.a[r + j], s.b[r + j]) : 0xFFu);
            }
        }
    } else {
        uint4 qx[SG8_G][4];
        if(vok){
            #pragma unroll
            for(int g = 0; g < SG8_G; g++){
                const L r = rt + g * 4096, rc = r + 16 <= n ? r : rlast;
                #pragma unroll
                for(int k = 0; k < 4; k++) qx[g][k] = *(const uint4*)(s.x + rc + 4 * k);
            }
        }
        #pragma unroll
        for(int g = 0; g < SG8_G; g++){
            const L r = rt + g * 4096;
            if(s.vec && r + 16 <= n){
                #pragma unroll
                for(int j = 0; j < 16; j++){
                    const uint4 &q = qx[g][j >> 2];
                    put(g, j, sg_i32_bin(s, (j & 3) == 0 ? q.x : (j & 3) == 1 ? q.y : (j & 3) == 2 ? q.z : q.w));
                }
            } else {
                #pragma unroll
                for(int j = 0; j < 16; j++) put(g, j, r + j < n ? sg_i32_bin(s, (uint32_t)s.x[r + j]) : 0xFFu);
            }
        }
    }
}
__global__ void __launch_bounds__(256) k_sg8_hist(SgSrc s, L n, int nbins, uint32_t *cnt, L ntiles, int *miss){
    __shared__ unsigned long long ws[4][2];
    __shared__ uint16_t ra[256], rb[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    ra[threadIdx.x] = s.ra[threadIdx.x]; rb[threadIdx.x] = s.rb[threadIdx.x];
    __syncthreads();
    unsigned long long c8[SG8_G];
    bool m = false;
    sg8_bins<false>(s, ra, rb, (L)blockIdx.x * SG8_TILE, n, nullptr, c8, &m);
    if(__syncthreads_or(m)){ if(threadIdx.x == 0) atomicOr(miss, 1); R; }
    unsigned long long e = 0, o = 0;
    #pragma unroll
    for(int g = 0; g < SG8_G; g++){ e += sg8_half(c8[g], 0); o += sg8_half(c8[g], 1); }
    for(int k = 32; k > 0; k >>= 1){ e += __shfl_xor(e, k, 64); o += __shfl_xor(o, k, 64); }
    if(lane == 0){ ws[w][0] = e; ws[w][1] = o; }
    __syncthreads();
    if((int)threadIdx.x < nbins){
        const unsigned d = threadIdx.x;
        unsigned t = 0;
        for(int q = 0; q < 4; q++) t += sg8_f16(ws[q][d & 1], d >> 1);
        cnt[(L)d * ntiles + blockIdx.x] = t;
    }
}
__global__ void __launch_bounds__(256) k_sg8_scatter(SgSrc s, L n, int nbins, const int64_t *off, L ntiles, int64_t *flat,
                                                     const int *miss){
    __shared__ uint16_t srow[SG8_TILE];
    __shared__ unsigned long long ws[SG8_G][4][2];
    __shared__ unsigned tst[9];
    __shared__ int64_t gbase[8];
    __shared__ uint16_t ra[256], rb[256];
    if(*miss) R;                                    // bins from a sample missed a row: nothing to place
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    ra[threadIdx.x] = s.ra[threadIdx.x]; rb[threadIdx.x] = s.rb[threadIdx.x];
    __syncthreads();
    const L t0 = (L)blockIdx.x * SG8_TILE;
    unsigned long long c8[SG8_G];
    unsigned pk[SG8_G][4];
    bool m = false;
    sg8_bins<true>(s, ra, rb, t0, n, pk, c8, &m);
    // per group: exclusive scans of the even-bin and odd-bin words over the workgroup's threads
    unsigned long long P[SG8_G][2];
    #pragma unroll
    for(int g = 0; g < SG8_G; g++){
        #pragma unroll
        for(int h = 0; h < 2; h++){
            const unsigned long long v = sg8_half(c8[g], h);
            unsigned long long inc = v;
            for(int k = 1; k < 64; k <<= 1){ const unsigned long long u = __shfl_up(inc, k, 64); if(lane >= k) inc += u; }
            if(lane == 63) ws[g][w][h] = inc;
            P[g][h] = inc - v;
        }
    }
    __syncthreads();
    // rows of each bin before the thread's chunk: earlier groups, earlier waves; then each bin's
    // tile start (exclusive prefix over bins of the tile totals), all in 16-bit fields
    unsigned long long tot[2] = {0, 0};
    #pragma unroll
    for(int g = 0; g < SG8_G; g++){
        #pragma unroll
        for(int h = 0; h < 2; h++){
            unsigned long long before = 0, gt = 0;
            #pragma unroll
            for(int q = 0; q < 4; q++){ const unsigned long long v = ws[g][q][h]; before += q < w ? v : 0ull; gt += v; }
            P[g][h] += tot[h] + before;
            tot[h] += gt;
        }
    }
    unsigned long long start[2] = {0, 0};
    unsigned run = 0;
    #pragma unroll
    for(int d = 0; d < 8; d++){
        start[d & 1] |= (unsigned long long)run << (16 * (d >> 1));
        run += sg8_f16(tot[d & 1], d >> 1);
    }
    if((int)threadIdx.x <= nbins){
        const unsigned d = threadIdx.x;
        tst[d] = d < (unsigned)nbins ? sg8_f16(start[d & 1], d >> 1) : 0x7FFFFFFFu;
        if(d < (unsigned)nbins) gbase[d] = off[(L)d * ntiles + blockIdx.x] - (int64_t)sg8_f16(start[d & 1], d >> 1);
    }
    // stable tile-local position of each of the thread's rows
    #pragma unroll
    for(int g = 0; g < SG8_G; g++){
        const unsigned long long pe = P[g][0] + start[0], po = P[g][1] + start[1];
        const int li = g * 4096 + threadIdx.x * 16;
        const uint4 b4 = make_uint4(pk[g][0], pk[g][1], pk[g][2], pk[g][3]);
        unsigned long long runc = 0;
        #pragma unroll
        for(int j = 0; j < 16; j++){
            const unsigned d = sg_byte(b4, j);
            if(d != 0xFFu){
                const unsigned p = sg8_f16(d & 1 ? po : pe, d >> 1) + ((unsigned)(runc >> (8 * d)) & 255u);
                srow[p] = (uint16_t)(li + j);
                runc += 1ull << (8 * d);
            }
        }
    }
    __syncthreads();
    // positions [tst[d], tst[d+1]) hold bin d: each thread follows its bin as j grows
    const int cnt = (int)(n - t0 < (L)SG8_TILE ? n - t0 : (L)SG8_TILE);
    int d = 0;
    #pragma unroll 4
    for(int j = threadIdx.x; j < cnt; j += 256){
        while(j >= (int)tst[d + 1]) d++;
        flat[gbase[d] + j] = t0 + srow[j];
    }
}
// group keys (first row of each non-empty bin) and counts; g = number of groups
__global__ void __launch_bounds__(256) k_sg_final(const int64_t *off, L ntiles, int nbins, L n, const int64_t *flat,
                                                  int64_t *keys, int64_t *counts, int64_t *g, const int *miss){
    __shared__ uint32_t wsum[4];
    if(*miss){ if(threadIdx.x == 0) *g = -1; R; }
    const int d = threadIdx.x, lane = d & 63, w = d >> 6;
    int64_t st = 0, c = 0;
    if(d < nbins){ st = off[(L)d * ntiles]; c = (d + 1 < nbins ? off[(L)(d + 1) * ntiles] : n) - st; }
    const uint32_t nz = c > 0 ? 1u : 0u;
    uint32_t inc = nz;
    for(int o = 1; o < 64; o <<= 1){ uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
    if(lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = inc - nz;
    for(int q = 0; q < w; q++) pre += wsum[q];
    if(nz){ keys[pre] = flat[st]; counts[pre] = c; }
    if(d == 255) *g = (int64_t)(pre + nz);
}

// Presence of byte values in two char columns (4 x 64-bit masks per column), and the unsigned
// range of an i32 column: per-workgroup partials, then one workgroup folds them.  Each thread
// has several 16-byte loads in flight; every load instruction reads one contiguous 1 KiB per wave.
#define SG_SCAN_UNROLL 4
// Presence: every row sets its two bytes' flags in LDS (plain stores of 1; no read-modify-write),
// then each wave turns 64 flags into one mask word with a ballot.
#define SG_PRESENT_UNROLL 4
__global__ void __launch_bounds__(256) k_sg_present(const uint8_t *a, const uint8_t *b, L n, int vec, unsigned long long *part){
    __shared__ uint8_t fa[256], fb[256];
    fa[threadIdx.x] = 0; fb[threadIdx.x] = 0;
    __syncthreads();
    const L S = (L)gridDim.x * 256 * 16;
    L i = ((L)blockIdx.x * 256 + threadIdx.x) * 16;
    if(vec){
        for(; i + (SG_PRESENT_UNROLL - 1) * S + 16 <= n; i += SG_PRESENT_UNROLL * S){
            uint4 wa[SG_PRESENT_UNROLL], wb[SG_PRESENT_UNROLL];
            #pragma unroll
            for(int u = 0; u < SG_PRESENT_UNROLL; u++){ wa[u] = *(const uint4*)(a + i + u * S); wb[u] = *(const uint4*)(b + i + u * S); }
            #pragma unroll
            for(int u = 0; u < SG_PRESENT_UNROLL; u++){
                #pragma unroll
                for(int j = 0; j < 16; j++){ fa[sg_byte(wa[u], j)] = 1; fb[sg_byte(wb[u], j)] = 1; }
            }
        }
    }
    for(; i < n; i += S){
        if(vec && i + 16 <= n){
            const uint4 wa = *(const uint4*)(a + i), wb = *(const uint4*)(b + i);
            #pragma unroll
            for(int j = 0; j < 16; j++){ fa[sg_byte(wa, j)] = 1; fb[sg_byte(wb, j)] = 1; }
        } else
            for(int j = 0; j < 16 && i + j < n; j++){ fa[a[i + j]] = 1; fb[b[i + j]] = 1; }
    }
    __syncthreads();
    const unsigned long long ma = __ballot(fa[threadIdx.x] != 0), mb = __ballot(fb[threadIdx.x] != 0);
    if((threadIdx.x & 63) == 0){
        part[(L)blockIdx.x * 8 + (threadIdx.x >> 6)] = ma;
        part[(L)blockIdx.x * 8 + 4 + (threadIdx.x >> 6)] = mb;
    }
}
__global__ void __launch_bounds__(256) k_sg_present_final(const unsigned long long *part, int nblocks, unsigned long long *out){
    __shared__ unsigned long long sm[4][8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    #pragma unroll
    for(int q = 0; q < 8; q++){
        unsigned long long v = 0;
        for(int k = threadIdx.x; k < nblocks; k += 256) v |= part[(L)k * 8 + q];
        for(int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
        if(lane == 0) sm[w][q] = v;
    }
    __syncthreads();
    if(threadIdx.x < 8) out[threadIdx.x] = sm[0][threadIdx.x] | sm[1][threadIdx.x] | sm[2][threadIdx.x] | sm[3][threadIdx.x];
}
// The same over a strided sample: thread k looks at rows [16 * stride * k, +16)
__global__ void __launch_bounds__(256) k_sg_present_sample(const uint8_t *a, const uint8_t *b, L n, L stride, int vec,
                                                           unsigned long long *part){
    __shared__ uint8_t fa[256], fb[256];
    fa[threadIdx.x] = 0; fb[threadIdx.x] = 0;
    __syncthreads();
    const L i = ((L)blockIdx.x * 256 + threadIdx.x) * stride * 16;
    if(i < n){
        if(vec && i + 16 <= n){
            const uint4 wa = *(const uint4*)(a + i), wb = *(const uint4*)(b + i);
            #pragma unroll
            for(int j = 0; j < 16; j++){ fa[sg_byte(wa, j)] = 1; fb[sg_byte(wb, j)] = 1; }
        } else
            for(int j = 0; j < 16 && i + j < n; j++){ fa[a[i + j]] = 1; fb[b[i + j]] = 1; }
    }
    __syncthreads();
    const unsigned long long ma = __ballot(fa[threadIdx.x] != 0), mb = __ballot(fb[threadIdx.x] != 0);
    if((threadIdx.x & 63) == 0){
        part[(L)blockIdx.x * 8 + (threadIdx.x >> 6)] = ma;
        part[(L)blockIdx.x * 8 + 4 + (threadIdx.x >> 6)] = mb;
    }
}
// workgroups of a whole-column pass that keeps SG_SCAN_UNROLL loads per thread in flight
static int sg_scan_grid(L n, int rows_per_load){
    L g = ceil_div(n, (L)256 * rows_per_load * SG_SCAN_UNROLL), cap = (L)num_cus() * 8;
    R (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// *missed: a row fell outside bins taken from a sample (z untouched)
static int32_t small_group_run(hpa_ctx *c, V z, SgSrc &s, L n, int nbins, bool *missed){
    int32_t e;
    *missed = false;
    s.nbins = (uint32_t)nbins;
    int kb = 0; while((1 << kb) < nbins) kb++;
    const bool few = nbins <= 8;                   // packed-count kernels
    const L ntiles = ceil_div(n, few ? SG8_TILE : SG_TILE);
    uint32_t *cnt = (uint32_t*)dalloc(c, 4 * nbins * ntiles);
    int64_t *off = (int64_t*)dalloc(c, 8 * nbins * ntiles);
    int64_t *flat = (int64_t*)dalloc(c, 8 * n);
    int64_t *keys = (int64_t*)dalloc(c, 8 * nbins), *counts = (int64_t*)dalloc(c, 8 * nbins);
    int64_t *gd = (int64_t*)dalloc(c, 16);
    if(!cnt || !off || !flat || !keys || !counts || !gd) R HPA_E_MEMORY;
    int *miss = (int*)(gd + 1);
    hipMemsetAsync(miss, 0, sizeof(int), c->stream);
    for(int pass = 0; pass < 2; pass++){
        if(pass == 1 && (e = scan_u32_to_i64(c, cnt, off, (L)nbins * ntiles, nullptr))) R e;
        ProfScope ps(pass == 0 ? "group_small_hist" : "group_small_scatter", c->stream,
                     (double)n * ((s.a ? 2 : 4) + (pass == 0 ? 0 : 8)));
        if(few){
            if(pass == 0) hipLaunchKernelGGL(k_sg8_hist, dim3((unsigned)ntiles), dim3(256), 0, c->stream, s, n, nbins, cnt, ntiles, miss);
            else hipLaunchKernelGGL(k_sg8_scatter, dim3((unsigned)ntiles), dim3(256), 0, c->stream, s, n, nbins, off, ntiles, flat, miss);
        } else {
            if(pass == 0) hipLaunchKernelGGL(k_sg_hist, dim3((unsigned)ntiles), dim3(256), 0, c->stream, s, n, nbins, kb, cnt, ntiles, miss);
            else hipLaunchKernelGGL(k_sg_scatter, dim3((unsigned)ntiles), dim3(256), 0, c->stream, s, n, nbins, kb, off, ntiles, flat, miss);
        }
    }
    hipLaunchKernelGGL(k_sg_final, dim3(1), dim3(256), 0, c->stream, off, ntiles, nbins, n, flat, keys, counts, gd, miss);
    if((e = check_launch("group_small"))) R e;
    int64_t g = 0; if((e = d2h(c, &g, gd, 8))) R e;
    if(g < 0){ *missed = true; R 0; }
    R make_group_dict(c, z, keys, counts, flat, g, n);
}
// presence masks (4 x 64 bits per column) of the two char columns, over all rows or a sample
static int32_t trie_presence(hpa_ctx *c, const uint8_t *a, const uint8_t *b, L n, int vec, bool sample, unsigned long long *hm){
    int32_t e;
    int grid;
    if(sample) grid = (int)ceil_div(ceil_div(n, 16 * (L)SG_SAMPLE_STRIDE), 256);
    else grid = sg_scan_grid(n, 16);
    unsigned long long *pm = (unsigned long long*)dalloc(c, 64 * (L)(grid + 1)); if(!pm) R HPA_E_MEMORY;
    {
        ProfScope ps("group_small_present", c->stream, (double)n * (sample ? 2.0 / SG_SAMPLE_STRIDE : 2.0));
        if(sample) hipLaunchKernelGGL(k_sg_present_sample, dim3(grid), dim3(256), 0, c->stream, a, b, n, (L)SG_SAMPLE_STRIDE, vec, pm + 8);
        else hipLaunchKernelGGL(k_sg_present, dim3(grid), dim3(256), 0, c->stream, a, b, n, vec, pm + 8);
        hipLaunchKernelGGL(k_sg_present_final, dim3(1), dim3(256), 0, c->stream, pm + 8, grid, pm);
    }
    if((e = check_launch("group_small_present"))) R e;
    R d2h(c, hm, pm, 64);
}
// trie group on at most 256 (c0, c1) bins; *done false when the byte values present need more
static int32_t small_group_trie(hpa_ctx *c, V z, const uint8_t *a, const uint8_t *b, L n, bool *done){
    *done = false;
    int32_t e;
    const int vec = ((uintptr_t)a % 16 == 0 && (uintptr_t)b % 16 == 0) ? 1 : 0;
    for(int sample = n >= SG_SAMPLE_MIN ? 1 : 0; sample >= 0; sample--){
        unsigned long long hm[8];
        if((e = trie_presence(c, a, b, n, vec, sample != 0, hm))) R e;
        SgSrc s; memset(&s, 0, sizeof(s));
        int na = 0, nb = 0;
        for(int v = 0; v < 256; v++){
            s.ra[v] = (hm[v >> 6] >> (v & 63)) & 1 ? (uint16_t)(na++) : (uint16_t)SG_ABSENT;
            s.rb[v] = (hm[4 + (v >> 6)] >> (v & 63)) & 1 ? (uint16_t)(nb++) : (uint16_t)SG_ABSENT;
        }
        if(na * nb > SG_BINS) R 0;                  // the whole input holds at least the sample's values
        s.a = a; s.b = b; s.nb = (uint32_t)nb; s.vec = vec;
        bool missed;
        if((e = small_group_run(c, z, s, n, na * nb, &missed))) R e;
        if(!missed){ *done = true; R 0; }
    }
    R 0;
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
    if(n > 0){
        bool done; if((e = small_group_trie(c, z, (const uint8_t*)pa, (const uint8_t*)pb, n, &done)) || done) R e;
    }
    uint32_t *k = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_trie, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)pa, (const uint8_t*)pb, k, n);
    uint32_t *sk, *sv;
    if((e = radix_sort_pairs_u32(c, k, nullptr, n, 0, 16, &sk, &sv))) R e;
    R runs_to_group_u32(c, z, sk, sv, n);
}
// ---------------------------------------------------------------- i32 group: one scan of the keys
// Before an i32 group every key is read once: per 4096-row tile the unsigned minimum and maximum,
// the OR and AND of the keys, whether some key is below the key before it, how many consecutive
// rows hold different keys, and the tile's run heads (a row whose key differs from the row
// before, or whose row number is a multiple of the run cap).  One read returns the summary and
// the run count, and picks the path: a counting sort over the key range when it spans fewer than
// SG_BINS values, the run-compressed group below, or the radix sort -- whose digit plan (which
// key bits vary, whether the input is already in key order) then needs no pass or read of its own.
struct GiSum { unsigned lo, hi, orv, andv; unsigned long long ch; unsigned desc, pad; };
__device__ __forceinline__ void gi_fold(GiSum &a, const GiSum &b){
    a.lo = b.lo < a.lo ? b.lo : a.lo; a.hi = b.hi > a.hi ? b.hi : a.hi;
    a.orv |= b.orv; a.andv &= b.andv; a.ch += b.ch; a.desc |= b.desc;
}
__device__ __forceinline__ void gi_wave_fold(GiSum &a){
    for(int o = 32; o > 0; o >>= 1){
        GiSum b;
        b.lo = __shfl_down(a.lo, o, 64); b.hi = __shfl_down(a.hi, o, 64);
        b.orv = __shfl_down(a.orv, o, 64); b.andv = __shfl_down(a.andv, o, 64);
        b.ch = __shfl_down(a.ch, o, 64); b.desc = __shfl_down(a.desc, o, 64);
        gi_fold(a, b);
    }
}
// Chunks of GI_CHUNK rows, one wave each (lane l holds rows [16 l, 16 l + 16) of the chunk: four
// 16-byte loads when aligned; the key before its first row is the lane below's last, lane 0
// reads it); waves take chunks independently, so no barrier holds back the next loads.  The run
// heads of each chunk go to ccount[chunk]; the summary is folded per workgroup into part.
#define GI_CHUNK (64 * RUN_ITEMS)
__global__ void __launch_bounds__(256) k_gi_scan(const int32_t *x, L n, int vec, unsigned cap_mask, uint32_t *ccount, L nchunks,
                                                 GiSum *part){
    __shared__ GiSum sp[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    GiSum a = {0xFFFFFFFFu, 0u, 0u, 0xFFFFFFFFu, 0ull, 0u, 0u};
    const L nw = (L)gridDim.x * 4;
    for(L ch = (L)blockIdx.x * 4 + w; ch < nchunks; ch += nw){
        const L i0 = ch * GI_CHUNK + (L)lane * RUN_ITEMS;
        unsigned v[RUN_ITEMS];
        if(vec && i0 + RUN_ITEMS <= n){
            #pragma unroll
            for(int q = 0; q < RUN_ITEMS / 4; q++){
                const uint4 u = *(const uint4*)(x + i0 + 4 * q);
                v[4 * q] = u.x; v[4 * q + 1] = u.y; v[4 * q + 2] = u.z; v[4 * q + 3] = u.w;
            }
        } else {
            #pragma unroll
            for(int j = 0; j < RUN_ITEMS; j++) v[j] = i0 + j < n ? (unsigned)x[i0 + j] : 0u;
        }
        unsigned prev = __shfl_up(v[RUN_ITEMS - 1], 1, 64);
        if(lane == 0) prev = (i0 > 0 && i0 <= n) ? (unsigned)x[i0 - 1] : 0u;
        uint32_t hc = 0;
        #pragma unroll
        for(int j = 0; j < RUN_ITEMS; j++){
            const L i = i0 + j;
            if(i < n){
                const unsigned k = v[j], pk = j == 0 ? prev : v[j - 1];
                a.lo = k < a.lo ? k : a.lo; a.hi = k > a.hi ? k : a.hi; a.orv |= k; a.andv &= k;
                const bool dif = i > 0 && k != pk;
                a.ch += dif ? 1ull : 0ull;
                a.desc |= (i > 0 && k < pk) ? 1u : 0u;
                hc += (((uint64_t)i & cap_mask) == 0 || dif) ? 1u : 0u;
            }
        }
        for(int o = 32; o > 0; o >>= 1) hc += __shfl_down(hc, o, 64);
        if(lane == 0) ccount[ch] = hc;
    }
    gi_wave_fold(a);
    if(lane == 0) sp[w] = a;
    __syncthreads();
    if(threadIdx.x == 0){
        for(int q = 1; q < 4; q++) gi_fold(a, sp[q]);
        part[blockIdx.x] = a;
    }
}
__global__ void __launch_bounds__(256) k_gi_final(const GiSum *part, int nb, GiSum *out){
    __shared__ GiSum sp[4];
    GiSum a = {0xFFFFFFFFu, 0u, 0u, 0xFFFFFFFFu, 0ull, 0u, 0u};
    for(int k = threadIdx.x; k < nb; k += 256) gi_fold(a, part[k]);
    gi_wave_fold(a);
    if((threadIdx.x & 63) == 0) sp[threadIdx.x >> 6] = a;
    __syncthreads();
    if(threadIdx.x == 0){ for(int q = 1; q < 4; q++) gi_fold(a, sp[q]); *out = a; }
}
struct GiScan { GiSum s; int64_t nr; int64_t *toff; };
// summary and run count of x (cap_mask: run cap - 1); toff[k] = runs before chunk k, so the
// runs before a RUN_TILE tile t are toff[t * (RUN_TILE / GI_CHUNK)]
static int32_t gi_scan(hpa_ctx *c, const int32_t *x, L n, unsigned cap_mask, GiScan *r){
    int32_t e;
    const L nch = ceil_div(n, GI_CHUNK);
    const L cap = (L)num_cus() * 8;
    const L nb = ceil_div(nch, 4);
    const int grid = (int)(nb < cap ? nb : cap);
    uint32_t *ccount = (uint32_t*)dalloc(c, 4 * nch);
    int64_t *toff = (int64_t*)dalloc(c, 8 * (nch + 1));
    GiSum *part = (GiSum*)dalloc(c, (L)sizeof(GiSum) * (grid + 1));
    if(!ccount || !toff || !part) R HPA_E_MEMORY;
    {
        ProfScope ps("group_scan", c->stream, 4.0 * n);
        hipLaunchKernelGGL(k_gi_scan, dim3(grid), dim3(256), 0, c->stream, x, n, (uintptr_t)x % 16 == 0 ? 1 : 0, cap_mask, ccount, nch, part + 1);
        hipLaunchKernelGGL(k_gi_final, dim3(1), dim3(256), 0, c->stream, (const GiSum*)(part + 1), grid, part);
    }
    if((e = check_launch("group_scan"))) R e;
    if((e = scan_u32_to_i64(c, ccount, toff, nch, toff + nch))) R e;
    const void *src[2] = {part, toff + nch};
    void *dst[2] = {&r->s, &r->nr};
    const int64_t bytes[2] = {(int64_t)sizeof(GiSum), 8};
    if((e = read_staged(c, 2, src, dst, bytes))) R e;
    r->toff = toff;
    R 0;
}

// Group of an i32 column whose equal keys sit in runs of consecutive rows (a fact table stored in
// key order of another table, say).  The stable sort that defines the group only has to order
// the runs: runs of one key keep their input order, and each run expands to ascending row
// numbers.  A run is cut at every cap-th row, cap = 2^cb with cb = min(6, 32 - bits of the row
// count), so that (start << cb) | (length - 1) fits the 32-bit value the sort carries.  Passes
// after the scan: the runs' (key, packed start and length) written at the scan's run offsets; a
// stable radix sort of the runs (digit plan from the scan); per tile of sorted runs, its group
// heads and rows; one launch that expands every run into the flat list at its position and
// writes each group's first row, flat offset and key value; the counts.  The outputs are
// allocated for as many groups as runs, so the only read is the group count at the end.  Taken
// from 2^24 rows when consecutive rows show few key changes and the runs are at most
// GR_RUNS_MAX_SHARE of the rows.
#define GR_RUNS_MIN (1 << 24)
#define GR_RUNS_MAX_CHANGE 0.4
#define GR_RUNS_MAX_SHARE 0.55
// Run heads of row i (threads of a wave hold consecutive rows): the previous row's key comes from
// the lane below; lane 0's row is a multiple of 64, so a forced head, and needs none.
__device__ __forceinline__ bool gr_head(int32_t xi, L i, L n, unsigned cap_mask){
    const int32_t xp = __shfl_up(xi, 1, 64);
    return i < n && (((uint64_t)i & cap_mask) == 0 || xi != xp);
}
// run r (input order): rkey[r] = its key, rval[r] = (first row << cb) | (length - 1).  A head is
// forced at every multiple of the cap (a divisor of 64), so the next head after a row is in the
// rest of its wave's 64-row ballot word or at the start of the next word (the next tile's first
// row, for the tile's last word).
__global__ void __launch_bounds__(256) k_gr_head_write(const int32_t *x, L n, unsigned cap_mask, int cb, const int64_t *toff, int tstride,
                                                       uint32_t *rkey, uint32_t *rval){
    __shared__ uint32_t s_cnt[RUN_ITEMS * 4];
    const L base = (L)blockIdx.x * RUN_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t bal[RUN_ITEMS];
    int32_t xv[RUN_ITEMS];
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; xv[k] = x[i < n ? i : n - 1]; }
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){
        bal[k] = __ballot(gr_head(xv[k], base + (L)k * 256 + threadIdx.x, n, cap_mask));
        if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    if(w == 0){
        const uint32_t mine = s_cnt[lane];
        uint32_t inc = mine;
        for(int o = 1; o < 64; o <<= 1){ const uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        s_cnt[lane] = inc - mine;
    }
    __syncthreads();
    const int64_t pre = toff[(L)blockIdx.x * tstride];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    #pragma unroll
    for(int k = 0; k < RUN_ITEMS; k++){
        if((bal[k] >> lane) & 1ull){
            const L i = base + (L)k * 256 + threadIdx.x;
            const int64_t r = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
            const int word = k * 4 + w;                // rows base + 64 * word ..
            const uint64_t above = lane == 63 ? 0ull : (bal[k] & (~0ull << (lane + 1)));
            L next = above ? base + 64 * (L)word + __builtin_ctzll(above) : base + 64 * (L)(word + 1);
            if(next > n) next = n;
            rkey[r] = (uint32_t)xv[k]; rval[r] = ((uint32_t)i << cb) | (uint32_t)(next - i - 1);
        }
    }
}
// Sorted runs (sk: keys, sv: (first row << cb) | (length - 1)) in tiles of GRF_TILE runs, thread
// t taking runs t, t + 256, ... of its tile.  k_grf_count: the group heads (a run whose key
// differs from the run before) and the rows of each tile.  k_grf_write: every run's position in
// the flat list (the tile's row offset, then the runs before it in the tile: a wave scan per
// step and a scan over the (step, wave) sums), the run expanded there as ascending row numbers,
// and for each group head the group's first row, flat offset and key value at its rank.
#define GRF_ITEMS 8
#define GRF_TILE (256 * GRF_ITEMS)
__global__ void __launch_bounds__(256) k_grf_count(const uint32_t *sk, const uint32_t *sv, L nr, unsigned len_mask,
                                                   uint32_t *gcount, uint32_t *rcount){
    __shared__ uint32_t sh[4], sr[4];
    const L base = (L)blockIdx.x * GRF_TILE;
    uint32_t h = 0, rows = 0;
    uint32_t vv[GRF_ITEMS];
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; vv[k] = sv[i < nr ? i : nr - 1]; }
    uint64_t bal[GRF_ITEMS]; uint32_t kv[GRF_ITEMS];
    run_heads<uint32_t, GRF_ITEMS>(sk, base, nr, bal, kv);
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x;
        h += (uint32_t)((bal[k] >> (threadIdx.x & 63)) & 1ull);
        if(i < nr) rows += (vv[k] & len_mask) + 1u;
    }
    for(int o = 32; o > 0; o >>= 1){ h += __shfl_down(h, o, 64); rows += __shfl_down(rows, o, 64); }
    if((threadIdx.x & 63) == 0){ sh[threadIdx.x >> 6] = h; sr[threadIdx.x >> 6] = rows; }
    __syncthreads();
    if(threadIdx.x == 0){ gcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3]; rcount[blockIdx.x] = sr[0] + sr[1] + sr[2] + sr[3]; }
}
// Tiles holding at most GRF_STAGE rows expand through LDS: a bitmap of the run starts over the
// tile's output range and a prefix count per 64-bit word give each output row its run, so the
// flat list is written with consecutive threads on consecutive rows; larger tiles (long runs)
// let each lane write its own run.  A head whose group ends inside the tile writes its count
// (the next head's position minus its own); the last group of each tile gets its count from
// k_grf_fix.
#define GRF_STAGE 16384
__global__ void __launch_bounds__(256) k_grf_write(const uint32_t *sk, const uint32_t *sv, L nr, int cb, unsigned len_mask,
                                                   const int64_t *goff, const int64_t *roff, int64_t *flat, int64_t *keys,
                                                   int64_t *start, int64_t *cnt, uint32_t *kval, uint32_t *grun){
    __shared__ uint32_t s_h[GRF_ITEMS * 4], s_r[GRF_ITEMS * 4];
    __shared__ uint64_t s_hb[GRF_ITEMS * 4];              // head flags of the tile's runs, 64 a word
    __shared__ uint32_t s_fpos[GRF_TILE], s_row[GRF_TILE];  // each run's tile-relative flat position, first row
    __shared__ unsigned long long s_bm[GRF_STAGE / 64];   // run starts over the tile's output range
    __shared__ uint32_t s_pre[GRF_STAGE / 64], s_wsum[4];
    const L base = (L)blockIdx.x * GRF_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t gpre = goff[blockIdx.x], rpre = roff[blockIdx.x];
    const L rows = roff[blockIdx.x + 1] - rpre;
    const bool staged = rows <= GRF_STAGE;
    if(staged) for(int q = threadIdx.x; q < GRF_STAGE / 64; q += 256) s_bm[q] = 0ull;
    uint32_t kv[GRF_ITEMS], vv[GRF_ITEMS], ex[GRF_ITEMS];
    uint64_t bal[GRF_ITEMS];
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; vv[k] = sv[i < nr ? i : nr - 1]; }
    run_heads<uint32_t, GRF_ITEMS>(sk, base, nr, bal, kv);
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x;
        const bool valid = i < nr;
        if(!valid){ kv[k] = 0u; vv[k] = 0u; }
        const uint32_t len = valid ? (vv[k] & len_mask) + 1u : 0u;
        uint32_t inc = len;
        for(int o = 1; o < 64; o <<= 1){ const uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        ex[k] = inc - len;
        if(lane == 63) s_r[k * 4 + w] = inc;
        if(lane == 0){ s_h[k * 4 + w] = (uint32_t)__popcll(bal[k]); s_hb[k * 4 + w] = bal[k]; }
    }
    __syncthreads();
    if(w == 0){
        const uint32_t mh = lane < GRF_ITEMS * 4 ? s_h[lane] : 0u, mr = lane < GRF_ITEMS * 4 ? s_r[lane] : 0u;
        uint32_t ih = mh, ir = mr;
        for(int o = 1; o < 64; o <<= 1){
            const uint32_t a = __shfl_up(ih, o, 64), b = __shfl_up(ir, o, 64);
            if(lane >= o){ ih += a; ir += b; }
        }
        if(lane < GRF_ITEMS * 4){ s_h[lane] = ih - mh; s_r[lane] = ir - mr; }
    }
    __syncthreads();
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x;
        const int idx = k * 256 + threadIdx.x;
        if(i < nr){
            const uint32_t fr = s_r[k * 4 + w] + ex[k];
            s_fpos[idx] = fr; s_row[idx] = vv[k] >> cb;
            if(staged) atomicOr(&s_bm[fr >> 6], 1ull << (fr & 63));
        }
    }
    __syncthreads();
    if(staged){
        // run count before each 64-row word of the output range: thread t owns word t
        static_assert(GRF_STAGE / 64 == 256, "one bitmap word per thread");
        const int nwd = (int)((rows + 63) >> 6);
        const uint32_t pc = (int)threadIdx.x < nwd ? (uint32_t)__popcll(s_bm[threadIdx.x]) : 0u;
        uint32_t inc = pc;
        for(int o = 1; o < 64; o <<= 1){ const uint32_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        if(lane == 63) s_wsum[w] = inc;
        __syncthreads();
        uint32_t pre = inc - pc;
        for(int q = 0; q < w; q++) pre += s_wsum[q];
        s_pre[threadIdx.x] = pre;
        __syncthreads();
        for(L p = threadIdx.x; p < rows; p += 256){
            const int wi = (int)(p >> 6);
            const unsigned long long m = s_bm[wi] & (~0ull >> (63 - (p & 63)));
            const int r = (int)s_pre[wi] + __popcll(m) - 1;
            flat[rpre + p] = (int64_t)s_row[r] + (p - (L)s_fpos[r]);
        }
    }
    #pragma unroll
    for(int k = 0; k < GRF_ITEMS; k++){
        const L i = base + (L)k * 256 + threadIdx.x;
        if(i < nr){
            const int idx = k * 256 + threadIdx.x;
            const int64_t fpos = rpre + (int64_t)s_fpos[idx];
            const int64_t row = (int64_t)(vv[k] >> cb);
            if(!staged){
                const uint32_t len = (vv[k] & len_mask) + 1u;
                for(uint32_t t = 0; t < len; t++) flat[fpos + t] = row + (int64_t)t;
            }
            if((bal[k] >> lane) & 1ull){
                const int64_t g = gpre + (int64_t)s_h[k * 4 + w] + __popcll(bal[k] & lt);
                keys[g] = row; start[g] = fpos; kval[g] = kv[k]; grun[g] = (uint32_t)i;
                // the next head of the tile, if any: later in this word, else the first later word
                int nx = -1;
                const uint64_t above = lane == 63 ? 0ull : (bal[k] & (~0ull << (lane + 1)));
                if(above) nx = k * 256 + w * 64 + __builtin_ctzll(above);
                else for(int q = k * 4 + w + 1; q < GRF_ITEMS * 4; q++) if(s_hb[q]){ nx = (q >> 2) * 256 + (q & 3) * 64 + __builtin_ctzll(s_hb[q]); break; }
                if(nx >= 0) cnt[g] = (int64_t)s_fpos[nx] - (int64_t)s_fpos[idx];
            }
        }
    }
}
// count of each tile's last group (its end lies in a later tile): start[g + 1] - start[g], with
// start[G] = n written here
__global__ void k_grf_fix(const int64_t *goff, L gt, int64_t *start, int64_t *cnt, L n, uint32_t *grun, L nr){
    const L G = goff[gt];
    for(L t = (L)blockIdx.x * blockDim.x + threadIdx.x; t < gt; t += (L)gridDim.x * blockDim.x){
        if(goff[t + 1] > goff[t]){
            const L g = goff[t + 1] - 1;
            cnt[g] = (g + 1 < G ? start[g + 1] : n) - start[g];
        }
    }
    if(blockIdx.x == 0 && threadIdx.x == 0){ start[G] = n; grun[G] = (uint32_t)nr; }
}
static int32_t group_i32_runs(hpa_ctx *c, V z, const int32_t *x, L n, int cb, const GiScan &sc){
    int32_t e;
    const unsigned cap_mask = (1u << cb) - 1;
    const L nt = ceil_div(n, RUN_TILE), nr = sc.nr;
    ProfScope ps("group_runs", c->stream);
    uint32_t *rkey = (uint32_t*)dalloc(c, 4 * nr), *rval = (uint32_t*)dalloc(c, 4 * nr);
    if(!rkey || !rval) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_gr_head_write, dim3((unsigned)nt), dim3(256), 0, c->stream, x, n, cap_mask, cb, (const int64_t*)sc.toff,
                       (int)(RUN_TILE / GI_CHUNK), rkey, rval);
    if((e = check_launch("group_runs"))) R e;
    // the runs' keys are some of the rows' keys: the rows' varying bits cover theirs, and rows out
    // of order (a key below the one before is always a new run) leave the runs out of order
    const RsHint h{sc.s.orv, sc.s.andv, true};
    uint32_t *sk, *sv;
    if((e = radix_sort_t<uint32_t>(c, rkey, rval, nr, 0, 32, &sk, &sv, nullptr, false, false, &h))) R e;
    const L gt = ceil_div(nr, GRF_TILE);
    uint32_t *gcount = (uint32_t*)dalloc(c, 4 * gt), *rcount = (uint32_t*)dalloc(c, 4 * gt);
    int64_t *goff = (int64_t*)dalloc(c, 8 * (gt + 1)), *roff = (int64_t*)dalloc(c, 8 * (gt + 1));
    if(!gcount || !rcount || !goff || !roff) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_grf_count, dim3((unsigned)gt), dim3(256), 0, c->stream, (const uint32_t*)sk, (const uint32_t*)sv, (L)nr, cap_mask,
                       gcount, rcount);
    if((e = check_launch("group_runs"))) R e;
    if((e = scan_u32_to_i64(c, gcount, goff, gt, goff + gt))) R e;
    if((e = scan_u32_to_i64(c, rcount, roff, gt, roff + gt))) R e;
    // outputs for as many groups as runs (the first run of each group before the counts: the
    // run side table dies with them)
    uint32_t *grun = (uint32_t*)dalloc(c, 4 * (nr + 1));
    uint32_t *kval = (uint32_t*)dalloc(c, 4 * nr);           // before keys: dies with them
    int64_t *keys = (int64_t*)dalloc(c, 8 * nr), *start = (int64_t*)dalloc(c, 8 * (nr + 1)), *cnt = (int64_t*)dalloc(c, 8 * nr);
    int64_t *flat = (int64_t*)dalloc(c, 8 * n);
    if(!grun || !kval || !keys || !start || !cnt || !flat) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_grf_write, dim3((unsigned)gt), dim3(256), 0, c->stream, (const uint32_t*)sk, (const uint32_t*)sv, (L)nr, cb, cap_mask,
                       (const int64_t*)goff, (const int64_t*)roff, flat, keys, start, cnt, kval, grun);
    hipLaunchKernelGGL(k_grf_fix, dim3(grid_for(gt, 256, 1)), dim3(256), 0, c->stream, (const int64_t*)goff, gt, start, cnt, n,
                       grun, (L)nr);
    if((e = check_launch("group_runs"))) R e;
    int64_t G = 0; if((e = d2h(c, &G, goff + gt, 8))) R e;
    ps.stop(); ps.bytes = 12.0 * n + 8.0 * G;        // int group: keys read, row list and keys written
    if(G > 1) note_group_vals(c, keys, x, n, kval);
    note_group_offsets(c, cnt, G, start);
    note_group_runs(c, cnt, hpa_ctx::GroupRuns{G, n, nr, sv, grun, cb});
    R make_group_dict(c, z, keys, cnt, flat, G, n);
}

static int32_t group_int(hpa_ctx *c, V z, V x){      // lib_group_by_normal_int, h_libs.cc:1295-1329
    L n = x->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, x, &p))) R e;
    if(n <= 0){
        int64_t *flat = (int64_t*)dalloc(c, 8); if(!flat) R HPA_E_MEMORY;
        uint32_t *sk = nullptr;
        if((e = radix_sort_flat_u32(c, (const uint32_t*)p, n, 0, 32, &sk, flat))) R e;
        R runs_to_group_flat_u32(c, z, sk, flat, n, p);
    }
    const int sbits = n > 1 ? 64 - __builtin_clzll((unsigned long long)(n - 1)) : 1;
    const int cb = 32 - sbits < 6 ? 32 - sbits : 6;
    GiScan sc;
    if((e = gi_scan(c, (const int32_t*)p, n, cb >= 2 ? (1u << cb) - 1 : 63u, &sc))) R e;
    // keys on fewer than SG_BINS consecutive unsigned values: counting sort
    if(sc.s.hi - sc.s.lo < (unsigned)SG_BINS){
        SgSrc s; memset(&s, 0, sizeof(s));
        s.x = (const int32_t*)p; s.base = sc.s.lo; s.vec = (uintptr_t)p % 16 == 0 ? 1 : 0;
        bool missed;
        if((e = small_group_run(c, z, s, n, (int)(sc.s.hi - sc.s.lo) + 1, &missed))) R e;
        if(!missed) R 0;
        set_error_detail("group: a key outside the scanned range"); R HPA_E_DEVICE;
    }
    const double change = n > 1 ? (double)sc.s.ch / (double)(n - 1) : 1.0;
    if(sc.s.desc && n >= GR_RUNS_MIN && cb >= 2 && change <= GR_RUNS_MAX_CHANGE && (double)sc.nr <= GR_RUNS_MAX_SHARE * (double)n)
        R group_i32_runs(c, z, (const int32_t*)p, n, cb, sc);
    // the unsigned radix key of an i32 is its bit pattern: the column is sorted in place of keys;
    // the sort's last pass scatters 4-byte row ids, which the run pass widens into the flat list
    // with sequential writes (an i64 scatter half-fills its lines) and takes each run's key from;
    // at most one group per key value of the range
    const RsHint h{sc.s.orv, sc.s.andv, sc.s.desc != 0};
    uint32_t *sk = nullptr, *sv = nullptr;
    if((e = radix_sort_t<uint32_t>(c, (const uint32_t*)p, nullptr, n, 0, 32, &sk, &sv, nullptr, false, false, &h))) R e;
    const L span = (L)(sc.s.hi - sc.s.lo) + 1;
    R runs_to_group_t<uint32_t>(c, z, sk, sv, n, nullptr, p, nullptr, span < n ? span : n);
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

static int32_t order_perm(hpa_ctx *c, V x, const uint8_t *flags, uint32_t **perm_out, bool *ties, bool *known_ties);
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
    int64_t *start = (int64_t*)dalloc(c, 8 * (g + 1));
    int64_t *cnt = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1));
    int64_t *flat = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1));
    if(!keys || !start || !cnt || !flat) R HPA_E_MEMORY;
    if(n > 0){
        hipLaunchKernelGGL(k_run_emit, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, flag, pos, perm, n, keys, start, flat);
        hipLaunchKernelGGL(k_run_counts, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, start, g, n, cnt);
    }
    if((e = check_launch("group_runs"))) R e;
    if(n > 0) note_group_offsets(c, cnt, g, start);
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

extern "C" int32_t pfnOrderBy(hpa_V z, hpa_V x, hpa_V y);
static int32_t order_vector_qsort(hpa_ctx *c, V z, V col, bool up, bool expect_ties);
bool qsort_ref_type(int32_t t);
__global__ void k_perm64_to_u32(const int64_t *p, L n, uint32_t *q){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) q[i] = (uint32_t)p[i];
}
// pfnGroupBasic (h_primitive.cc:1002-1057).  Lists and the types other than I/L are grouped along
// the order pfnOrderBy gives them with every flag true (h_primitive.cc:1017-1020) -- its case 1
// for two i32 columns, its single-vector rules for one column -- cut into runs of equal rows
// (lib_get_group_by, h_libs.cc:776-809).
static int32_t group_along_order(hpa_ctx *c, V z, V x, int nc, V cols, L n){
    int32_t e;
    std::vector<uint8_t> up((size_t)nc, 1);
    V0 y; memset(&y, 0, sizeof(y)); y.typ = HPA_B; y.len = nc; y.g = (char*)up.data();
    V0 t; memset(&t, 0, sizeof(t));
    if((e = pfnOrderBy(&t, x, &y))) R e;
    uint32_t *perm = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!perm) R HPA_E_MEMORY;
    if(n > 1) hipLaunchKernelGGL(k_perm64_to_u32, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)t.g, n, perm);
    else if(n == 1 && (e = iota_u32(c, perm, 1))) R e;
    if((e = check_launch("group"))) R e;
    R runs_along_perm(c, z, nc, cols, perm, n);
}
extern "C" int32_t pfnGroupBasic(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(x->typ == HPA_I) R group_int(c, z, x);
    if(x->typ == HPA_L) R group_long(c, z, x);
    if(x->typ == HPA_G){
        L nc = x->len;
        if(nc == 0) R HPA_E_DOMAIN;
        L n = ((V)x->g)->len;
        for(L q = 0; q < nc; q++){ V col = ((V)x->g) + q; if(!is_basic(col->typ)) R HPA_E_DOMAIN; if(col->len != n) R HPA_E_MATCH; }
        R group_along_order(c, z, x, (int)nc, (V)x->g, n);
    }
    if(is_basic(x->typ)){
        if(x->typ == HPA_S || x->typ == HPA_X || x->typ == HPA_Q) R HPA_E_NOT_IMPL;
        if(qsort_ref_type(x->typ)){
            // pfnOrderBy of the vector with a true flag (h_primitive.cc:1017-1020): a group's keys
            // repeat, so the reference's partitions run at once
            V0 t; memset(&t, 0, sizeof(t)); int32_t e;
            if((e = order_vector_qsort(c, &t, x, true, true))) R e;
            L n = x->len;
            uint32_t *perm = (uint32_t*)dalloc(c, 4 * (n > 0 ? n : 1)); if(!perm) R HPA_E_MEMORY;
            if(n > 1) hipLaunchKernelGGL(k_perm64_to_u32, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)t.g, n, perm);
            else if(n == 1 && (e = iota_u32(c, perm, 1))) R e;
            if((e = check_launch("group"))) R e;
            R runs_along_perm(c, z, 1, x, perm, n);
        }
        R group_along_order(c, z, x, 1, x, x->len);
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

// Stable permutation of x (list of columns, flags per column) in lexicographic order, ties by
// ascending row.  LSD over columns: sort by the last column, then stably by each earlier one.
// Small inputs (n <= 1024, <= 8 columns): one workgroup ranks every row against every other on
// the same order keys -- ties by row -- in a single launch with no host round trip.
#define SMALL_ORDER 1024
#define SMALL_COLS 8
struct SmallCols { int nc; int typ[SMALL_COLS]; int desc[SMALL_COLS]; int bits[SMALL_COLS]; const void *p[SMALL_COLS]; };
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

// Larger inputs: rows are bucketed on the most significant varying bits of the first column's
// key -- bucket order is key order -- and placed bucket by bucket; a row's place inside its
// bucket is its rank among the bucket's rows on the full key, ties by row.  Only rows that share
// a bucket are compared.  The bucket window (nb bits ending at the highest bit that varies) is
// worked out on the device from the OR and AND of all keys, so the only host read is the final
// check: a bucket of more than BR_BUCKET_MAX rows (the placing and ranking then write nothing)
// sends the order to the radix passes instead.
#define BR_BUCKET_MAX 1024
__device__ __forceinline__ uint64_t br_key(const SmallCols &s, int q, L r){
    const uint64_t v = okey_any(s.typ[q], s.p[q], r);
    const uint64_t mask = s.bits[q] >= 64 ? ~0ull : ((1ull << s.bits[q]) - 1);
    return s.desc[q] ? (~v & mask) : v;
}
// k0 = first column's key; or_and[0] |= every key, or_and[1] &= every key
__global__ void __launch_bounds__(256) k_br_key(SmallCols s, L n, uint64_t *k0, unsigned long long *or_and){
    __shared__ unsigned long long so[4], sa[4];
    unsigned long long o = 0, a = ~0ull;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const uint64_t v = br_key(s, 0, i);
        k0[i] = v; o |= v; a &= v;
    }
    for(int d = 32; d > 0; d >>= 1){ o |= __shfl_down(o, d, 64); a &= __shfl_down(a, d, 64); }
    if((threadIdx.x & 63) == 0){ so[threadIdx.x >> 6] = o; sa[threadIdx.x >> 6] = a; }
    __syncthreads();
    // one pair of atomics per workgroup: the waves' atomics on two words serialised the kernel
    if(threadIdx.x == 0){
        for(int q = 1; q < 4; q++){ o |= so[q]; a &= sa[q]; }
        atomicOr(or_and, o); atomicAnd(or_and + 1, a);
    }
}
__global__ void k_br_init(uint32_t *p, L n, unsigned long long *or_and){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) p[i] = 0;
    if(or_and && blockIdx.x == 0 && threadIdx.x == 0){ or_and[0] = 0; or_and[1] = ~0ull; }
}
// low bit of the bucket window: nb bits ending at the highest varying key bit
__device__ __forceinline__ int br_shift(const unsigned long long *or_and, int nb){
    const unsigned long long vary = or_and[0] ^ or_and[1];
    const int top = vary ? 64 - __clzll(vary) : 0;
    return top > nb ? top - nb : 0;
}
// bucket counts; a wave adds the lanes that share its leading lanes' buckets with one atomic each
// (up to four rounds), so a few crowded buckets do not serialise the kernel
__global__ void __launch_bounds__(256) k_br_hist(const uint64_t *k0, L n, const unsigned long long *or_and, int nb, uint32_t *cnt, int *over){
    const int shift = br_shift(or_and, nb);
    const uint32_t bmask = (1u << nb) - 1;
    const int lane = threadIdx.x & 63;
    for(L i0 = (L)blockIdx.x * blockDim.x; i0 < n; i0 += (L)gridDim.x * blockDim.x){
        const L i = i0 + threadIdx.x;
        bool pending = i < n;
        const uint32_t b = pending ? (uint32_t)(k0[i] >> shift) & bmask : 0u;
        for(int it = 0; it < 4; it++){
            const uint64_t pm = __ballot(pending);
            if(!pm) break;
            const int leader = __ffsll((unsigned long long)pm) - 1;
            const uint32_t bl = __shfl(b, leader, 64);
            const bool mine = pending && b == bl;
            const uint64_t mm = __ballot(mine);
            if(lane == leader){
                const uint32_t k = (uint32_t)__popcll(mm);
                if(atomicAdd(&cnt[bl], k) + k > BR_BUCKET_MAX) atomicOr(over, 1);
            }
            if(mine) pending = false;
        }
        if(pending && atomicAdd(&cnt[b], 1u) + 1 > BR_BUCKET_MAX) atomicOr(over, 1);
    }
}
__global__ void __launch_bounds__(256) k_br_place(const uint64_t *k0, L n, const unsigned long long *or_and, int nb, const int *over,
                                                  const int64_t *start, uint32_t *cursor, uint32_t *slot_row, uint64_t *slot_k0){
    if(*over) return;
    const int shift = br_shift(or_and, nb);
    const uint32_t bmask = (1u << nb) - 1;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const uint64_t v = k0[i];
        const uint32_t b = (uint32_t)(v >> shift) & bmask;
        const int64_t p = start[b] + atomicAdd(&cursor[b], 1u);
        slot_row[p] = (uint32_t)i; slot_k0[p] = v;
    }
}
// over[1] is set when two rows are equal on every column (a full tie)
__global__ void __launch_bounds__(256) k_br_rank(SmallCols s, L n, const unsigned long long *or_and, int nb, int *over,
                                                 const int64_t *start, const uint32_t *cnt, const uint32_t *slot_row,
                                                 const uint64_t *slot_k0, uint32_t *perm){
    if(over[0]) return;
    const int shift = br_shift(or_and, nb);
    const uint32_t bmask = (1u << nb) - 1;
    for(L p = (L)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (L)gridDim.x * blockDim.x){
        const uint32_t r = slot_row[p];
        const uint64_t a = slot_k0[p];
        const uint32_t b = (uint32_t)(a >> shift) & bmask;
        const int64_t s0 = start[b], e = s0 + cnt[b];
        uint32_t rank = 0;
        bool tie = false;
        for(int64_t j = s0; j < e; j++){
            const uint64_t kj = slot_k0[j];
            if(kj != a){ rank += kj < a; continue; }
            const uint32_t rj = slot_row[j];
            if(rj == r) continue;
            int less = rj < r, q = 1;
            for(; q < s.nc; q++){
                const uint64_t x = br_key(s, q, rj), y = br_key(s, q, r);
                if(x != y){ less = x < y; break; }
            }
            tie |= q == s.nc;
            rank += less;
        }
        perm[s0 + rank] = r;
        if(tie) over[1] = 1;
    }
}
// *fallback: a bucket is too large and nothing was written; *ties: two rows are equal on every
// column
static int32_t order_bucket_rank(hpa_ctx *c, const SmallCols &s, L n, uint32_t **perm_out, bool *fallback, bool *ties){
    int32_t e;
    *fallback = false;
    // about two buckets per row: real keys crowd into a few stretches of the window (a sum's
    // exponent takes a handful of values), and a row's ranking work is its bucket's size
    int lg = 0; while((1ll << lg) < n) lg++;
    int nb = lg + 1; if(nb < 8) nb = 8; if(nb > 22) nb = 22;
    const L B = 1ll << nb;
    unsigned long long *or_and = (unsigned long long*)dalloc(c, 16);
    int *over = (int*)dalloc(c, 16);
    uint64_t *k0 = (uint64_t*)dalloc(c, 8 * n), *slot_k0 = (uint64_t*)dalloc(c, 8 * n);
    uint32_t *cnt = (uint32_t*)dalloc(c, 8 * B), *cursor = cnt + B;       // zeroed together
    int64_t *start = (int64_t*)dalloc(c, 8 * (B + 1));
    uint32_t *slot_row = (uint32_t*)dalloc(c, 4 * n), *perm = (uint32_t*)dalloc(c, 4 * n);
    if(!or_and || !over || !k0 || !slot_k0 || !cnt || !start || !slot_row || !perm) R HPA_E_MEMORY;
    const int g = grid_for(n, 256, 4);
    hipLaunchKernelGGL(k_br_init, dim3(grid_for(2 * B, 256, 4)), dim3(256), 0, c->stream, cnt, 2 * B, or_and);
    hipLaunchKernelGGL(k_br_init, dim3(1), dim3(64), 0, c->stream, (uint32_t*)over, (L)2, (unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_br_key, dim3(g), dim3(256), 0, c->stream, s, n, k0, or_and);
    hipLaunchKernelGGL(k_br_hist, dim3(g), dim3(256), 0, c->stream, (const uint64_t*)k0, n, (const unsigned long long*)or_and, nb, cnt, over);
    if((e = check_launch("order_bucket"))) R e;
    if((e = scan_u32_to_i64(c, cnt, start, B, start + B))) R e;
    hipLaunchKernelGGL(k_br_place, dim3(g), dim3(256), 0, c->stream, (const uint64_t*)k0, n, (const unsigned long long*)or_and, nb,
                       (const int*)over, (const int64_t*)start, cursor, slot_row, slot_k0);
    hipLaunchKernelGGL(k_br_rank, dim3(grid_for(n, 256, 1)), dim3(256), 0, c->stream, s, n, (const unsigned long long*)or_and, nb, over,
                       (const int64_t*)start, (const uint32_t*)cnt, (const uint32_t*)slot_row, (const uint64_t*)slot_k0, perm);
    if((e = check_launch("order_bucket"))) R e;
    int ov[2] = {0, 0}; if((e = d2h(c, ov, over, 8))) R e;
    if(ov[0]){ *fallback = true; R 0; }
    *ties = ov[1] != 0;
    *perm_out = perm;
    R 0;
}

// *ties / *known_ties: whether two rows are equal on every column, when the path found out
static int32_t order_perm(hpa_ctx *c, V x, const uint8_t *flags, uint32_t **perm_out, bool *ties, bool *known_ties){
    L nc = x->len; L n = ((V)x->g)->len; int32_t e;
    *ties = false; *known_ties = false;
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
    if(n > SMALL_ORDER && n < (L)UINT32_MAX && nc <= SMALL_COLS){
        SmallCols s; memset(&s, 0, sizeof(s)); s.nc = (int)nc;
        bool ok = true;
        for(L q = 0; q < nc; q++){
            V col = ((V)x->g) + q;
            if(col->typ == HPA_S || col->typ == HPA_X || col->typ == HPA_Q){ ok = false; break; }
            const void *p; if((e = dev_view(c, col, &p))) R e;
            s.typ[q] = col->typ; s.p[q] = p; s.desc[q] = flags[q] ? 0 : 1; s.bits[q] = key_bits_of(col->typ);
        }
        if(ok){
            const L mark = c->dcur;
            bool fallback;
            ProfScope ps("order_bucket", c->stream);
            bool bt = true;
            if((e = order_bucket_rank(c, s, n, perm_out, &fallback, &bt))) R e;
            if(!fallback){ *ties = bt; *known_ties = true; R 0; }
            arena_set(c, mark);
        }
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
    if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
    else note_bound(c, dz, n);                   // a permutation of 0..n-1
    R e;
}

// single i32 / i64 column or the packed pair: stable unsigned radix, desc = reversal
// order of a single key column by its radix key: the last radix pass writes the i64 permutation
// (reversed for a descending order) straight into the result
template<typename K>
static int32_t order_radix_out(hpa_ctx *c, V z, const K *keys, L n, int hi, bool up){
    int32_t e;
    if(n <= 1){
        uint32_t *pm = (uint32_t*)dalloc(c, 4); if(!pm) R HPA_E_MEMORY;
        if(n == 1 && (e = iota_u32(c, pm, 1))) R e;
        R order_emit(c, z, pm, n, !up);
    }
    void *dz;
    if((e = out_init(c, z, HPA_L, n, &dz))) R e;
    K *sk; uint32_t *sv;
    if((e = radix_sort_t<K>(c, keys, nullptr, n, 0, hi, &sk, &sv, (int64_t*)dz, !up))) R e;
    note_bound(c, dz, n);                        // a permutation of 0..n-1
    R 0;
}
static int32_t order_radix_i32(hpa_ctx *c, V z, V col, bool up){
    L n = col->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, col, &p))) R e;
    R order_radix_out<uint32_t>(c, z, (const uint32_t*)p, n, 32, up);     // key = the i32 bit pattern
}
static int32_t order_radix_i64(hpa_ctx *c, V z, const int64_t *p, L n, bool up){
    int32_t e; bool hi;
    if((e = any_high_i64(c, p, n, &hi))) R e;
    if(hi) R order_radix_out<uint64_t>(c, z, (const uint64_t*)p, n, 64, up);
    uint64_t *k = (uint64_t*)dalloc(c, 8 * (n > 0 ? n : 1)); if(!k) R HPA_E_MEMORY;
    if(n > 0) hipLaunchKernelGGL(k_key_i64, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, p, k, n, 1);
    R order_radix_out<uint64_t>(c, z, k, n, 32, up);
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

// the reference's quicksort restated on the device (hpa_qsort.hip)
int32_t qsort_ref_perm(hpa_ctx *c, int32_t typ, const void *col, uint32_t *perm, L n, const uint32_t *seg_lo,
                       const uint32_t *seg_hi, L nseg, bool up);
int32_t order_list_ref(hpa_ctx *c, int nc, const int32_t *typ, const void *const *cols, const uint8_t *flags, L n,
                       uint32_t **perm_out);

// whether two rows adjacent in perm are equal on every column (flag[0]), and whether an F/E
// column holds a NaN (flag[1])
__global__ void __launch_bounds__(256) k_full_ties(SmallCols s, const uint32_t *perm, L n, int *flag){
    bool tie = false, nan = false;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const L r = perm ? perm[i] : i, r0 = i > 0 ? (perm ? perm[i - 1] : i - 1) : 0;
        bool eq = i > 0;
        for(int q = 0; q < s.nc; q++){
            if(s.typ[q] == HPA_E){ const double v = ((const double*)s.p[q])[r]; nan |= v != v; }
            if(s.typ[q] == HPA_F){ const float v = ((const float*)s.p[q])[r]; nan |= v != v; }
            if(eq) eq = okey_any(s.typ[q], s.p[q], r0) == okey_any(s.typ[q], s.p[q], r);
        }
        tie |= eq;
    }
    if(__ballot(tie) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
    if(__ballot(nan) && (threadIdx.x & 63) == 0) atomicOr(flag + 1, 1);
}
// *ties: some adjacent rows of perm tie on every column; HPA_E_DOMAIN for a NaN key
static int32_t full_ties(hpa_ctx *c, int nc, V cols, const uint32_t *perm, L n, bool *ties){
    int32_t e;
    *ties = false;
    if(n < 1) R 0;
    SmallCols s; memset(&s, 0, sizeof(s)); s.nc = nc;
    for(int q = 0; q < nc; q++){ const void *p; if((e = dev_view(c, cols + q, &p))) R e; s.typ[q] = cols[q].typ; s.p[q] = p; }
    int *f = (int*)dalloc(c, 16); if(!f) R HPA_E_MEMORY;
    hipMemsetAsync(f, 0, 8, c->stream);
    hipLaunchKernelGGL(k_full_ties, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, s, perm, n, f);
    if((e = check_launch("order_ties"))) R e;
    int h[2]; if((e = d2h(c, h, f, 8))) R e;
    if(h[1]){ set_error_detail("order: NaN key in a quicksort-ordered column"); R HPA_E_DOMAIN; }
    *ties = h[0] != 0;
    R 0;
}
// A column ordered by comparison in the reference (lib_quicksort_other).  Without ties every
// correct order is its order, so the stable sort on the order key answers; keys with ties run
// the reference's partitions (hpa_qsort.hip) from the identity, as lib_order_by_vector starts.
// expect_ties (groups, whose keys repeat by design) goes to the partitions directly.
static int32_t order_vector_qsort(hpa_ctx *c, V z, V col, bool up, bool expect_ties){
    L n = col->len; int32_t e;
    const void *p = nullptr;
    if(n > 0 && (e = dev_view(c, col, &p))) R e;
    const L mark = c->dcur;
    if(!expect_ties || n < 2){
        V0 lst; memset(&lst, 0, sizeof(lst)); lst.typ = HPA_G; lst.len = 1; lst.g = (char*)col;
        uint8_t f = up ? 1 : 0; uint32_t *perm; bool ties, known;
        if((e = order_perm(c, &lst, &f, &perm, &ties, &known))) R e;
        bool t2; if((e = full_ties(c, 1, col, perm, n, &t2))) R e;
        if(!t2) R order_emit(c, z, perm, n, false);
        arena_set(c, mark);
    } else {
        bool t2; if((e = full_ties(c, 1, col, nullptr, n, &t2))) R e;      // the NaN check
    }
    uint32_t *perm = (uint32_t*)dalloc(c, 4 * n); if(!perm) R HPA_E_MEMORY;
    if((e = iota_u32(c, perm, n))) R e;
    if((e = qsort_ref_perm(c, col->typ, p, perm, n, nullptr, nullptr, 0, up))) R e;
    R order_emit(c, z, perm, n, false);
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
            if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
            R e;
        }
        // counting sort by byte (stable, ascending), then reversal for desc
        uint64_t *k = (uint64_t*)dalloc(c, 8 * n); if(!k) R HPA_E_MEMORY;
        const void *p; if((e = dev_view(c, col, &p))) R e;
        hipLaunchKernelGGL(k_okey_perm<uint8_t>, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)p, (const uint32_t*)nullptr, k, n, 0, 8);
        R order_radix_out<uint64_t>(c, z, k, n, 8, up);
    }
    if(qsort_ref_type(col->typ)) R order_vector_qsort(c, z, col, up, false);
    R HPA_E_NOT_IMPL;
}

// ---------------------------------------------------------------- list order: tie direction
// The reference orders a list column by column (lib_quicksort_list, h_libs.cc:637-697): within a
// run of rows equal on the earlier columns, a column that is already non-decreasing and not all
// equal is kept for an ascending flag and *reversed in place* for a descending one
// (lib_order_by_sorted, h_libs.cc:604-617); otherwise the run is sorted stably (integers) and
// split into runs of equal values.  So the rows of every run stay in ascending or descending row
// order, and the result is the lexicographic order with rows equal on every column in ascending
// or descending row order -- descending where an odd number of such reversals lies above them.
// order_perm breaks those full ties by ascending row; this pass finds the reversals and turns the
// affected tie groups around.  It runs only for lists with a descending column and some full tie.
//   lvl[i]: first column where the rows at positions i-1 and i of the order differ (nc: a full
//   tie; 0 at i = 0).  Positions i-1, i share a run of level c (equal on columns < c) iff
//   lvl[i] >= c.
__global__ void k_ot_lvl(SmallCols s, const uint32_t *perm, L n, uint8_t *lvl, unsigned long long *nties){
    unsigned long long t = 0;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        int l = 0;
        if(i > 0){
            const uint32_t a = perm[i - 1], b = perm[i];
            while(l < s.nc && okey_any(s.typ[l], s.p[l], a) == okey_any(s.typ[l], s.p[l], b)) l++;
        }
        lvl[i] = (uint8_t)l;
        t += l == s.nc;
    }
    for(int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if((threadIdx.x & 63) == 0 && t) atomicAdd(nties, t);
}
// b[i] = 1 where position i starts a run of level c (i > 0)
__global__ void k_ot_bound(const uint8_t *lvl, L n, int c, uint32_t *b){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) b[i] = i > 0 && lvl[i] < c;
}
// run id of each row (inclusive count of run starts up to its position)
__global__ void k_ot_rowrun(const uint32_t *perm, const uint32_t *b, const int64_t *ex, L n, uint32_t *rowrun){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) rowrun[perm[i]] = (uint32_t)(ex[i] + b[i]);
}
// rows sorted by (run, row): consecutive rows of one run in ascending row order; bit 1 of
// rflag[run]: column c increases somewhere along them, bit 2: decreases somewhere
// (one run can hold every row: a wave whose rows share a run ORs its bits first, and a bit
// already set is not set again, so the flag word does not serialise the kernel)
__global__ void k_ot_check(SmallCols s, int c, const uint32_t *srun, const uint32_t *srow, L n, uint32_t *rflag){
    const L stride = (L)gridDim.x * blockDim.x;
    for(L j0 = (L)blockIdx.x * blockDim.x; j0 < n; j0 += stride){
        const L j = j0 + threadIdx.x + 1;                     // rows j-1, j: the pair at hand
        uint32_t bit = 0, run = 0xFFFFFFFFu;
        if(j < n && srun[j] == srun[j - 1]){
            run = srun[j];
            const uint64_t a = okey_any(s.typ[c], s.p[c], srow[j - 1]), b = okey_any(s.typ[c], s.p[c], srow[j]);
            if(a != b) bit = a < b ? 1u : 2u;
        }
        const uint32_t r0 = __builtin_amdgcn_readfirstlane(run);
        if(__all(run == r0)){
            uint32_t w = bit;
            for(int o = 32; o > 0; o >>= 1) w |= __shfl_xor(w, o, 64);
            if((threadIdx.x & 63) == 0 && w && r0 != 0xFFFFFFFFu && (rflag[r0] & w) != w) atomicOr(&rflag[r0], w);
        } else if(bit && (rflag[run] & bit) == 0) atomicOr(&rflag[run], bit);
    }
}
// a run whose rows run in direction d (0 ascending rows, 1 descending) is reversed when column
// c, read in that direction, is non-decreasing and not all equal; its sub-runs take the new
// direction
__global__ void k_ot_flip(const uint32_t *b, const int64_t *ex, const uint32_t *rflag, L n, uint32_t *dir){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const uint32_t f = rflag[ex[i] + b[i]];
        const bool inc = f & 1u, dec = f & 2u, d = dir[i] != 0;
        const bool flip = d ? (dec && !inc) : (inc && !dec);
        if(flip) dir[i] = d ? 0u : 1u;
    }
}
// group starts of the full-tie groups (runs of level nc)
__global__ void k_ot_gstart(const uint32_t *b, const int64_t *ex, L n, int64_t *gs){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        if(i == 0 || b[i]) gs[ex[i] + b[i]] = i;
}
// tie groups in descending direction come out reversed
__global__ void k_ot_apply(const uint32_t *perm, const uint32_t *b, const int64_t *ex, const int64_t *gs, const uint32_t *dir,
                           L n, L ng, uint32_t *out){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const L g = ex[i] + b[i];
        const L lo = gs[g], hi = g + 1 < ng ? gs[g + 1] : n;
        out[dir[i] ? lo + hi - 1 - i : i] = perm[i];
    }
}
// known: whether the order found full ties (1), found none (0), or did not look (-1)
static int32_t order_list_ties(hpa_ctx *c, V x, const uint8_t *flags, uint32_t **perm_io, L n, int known){
    const L nc = x->len;
    bool any_desc = false;
    for(L q = 0; q < nc; q++) any_desc |= flags[q] == 0;
    if(!any_desc || n < 2 || nc > SMALL_COLS || known == 0) R 0;
    int32_t e;
    SmallCols s; memset(&s, 0, sizeof(s)); s.nc = (int)nc;
    for(L q = 0; q < nc; q++){
        V col = ((V)x->g) + q;
        if(col->typ == HPA_S || col->typ == HPA_X || col->typ == HPA_Q) R 0;
        const void *p; if((e = dev_view(c, col, &p))) R e;
        s.typ[q] = col->typ; s.p[q] = p; s.desc[q] = flags[q] ? 0 : 1; s.bits[q] = key_bits_of(col->typ);
    }
    const uint32_t *perm = *perm_io;
    const int g = grid_for(n, 256, 4);
    uint8_t *lvl = (uint8_t*)dalloc(c, n);
    uint32_t *dir = (uint32_t*)dalloc(c, 4 * n);
    unsigned long long *nt = (unsigned long long*)dalloc(c, 16);
    if(!lvl || !dir || !nt) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_br_init, dim3(1), dim3(64), 0, c->stream, (uint32_t*)nt, (L)2, (unsigned long long*)nullptr);
    hipLaunchKernelGGL(k_ot_lvl, dim3(g), dim3(256), 0, c->stream, s, perm, n, lvl, nt);
    if((e = check_launch("order_ties"))) R e;
    if(known < 0){
        unsigned long long ties = 0; if((e = d2h(c, &ties, nt, 8))) R e;
        if(ties == 0) R 0;                                    // the lexicographic order is unique
    }
    ProfScope ps("order_ties", c->stream);
    hipLaunchKernelGGL(k_br_init, dim3(g), dim3(256), 0, c->stream, dir, n, (unsigned long long*)nullptr);
    uint32_t *b = (uint32_t*)dalloc(c, 4 * n), *rowrun = (uint32_t*)dalloc(c, 4 * n), *rflag = (uint32_t*)dalloc(c, 4 * n);
    int64_t *ex = (int64_t*)dalloc(c, 8 * (n + 1));
    if(!b || !rowrun || !rflag || !ex) R HPA_E_MEMORY;
    for(int q = 0; q < (int)nc; q++){
        if(flags[q]) continue;                               // ascending columns never reverse a run
        hipLaunchKernelGGL(k_ot_bound, dim3(g), dim3(256), 0, c->stream, (const uint8_t*)lvl, n, q, b);
        if((e = scan_u32_to_i64(c, b, ex, n, ex + n))) R e;
        hipLaunchKernelGGL(k_ot_rowrun, dim3(g), dim3(256), 0, c->stream, perm, (const uint32_t*)b, (const int64_t*)ex, n, rowrun);
        hipLaunchKernelGGL(k_br_init, dim3(g), dim3(256), 0, c->stream, rflag, n, (unsigned long long*)nullptr);
        const L mark = c->dcur;
        uint32_t *srun, *srow;
        if((e = radix_sort_pairs_u32(c, rowrun, nullptr, n, 0, 32, &srun, &srow))) R e;
        hipLaunchKernelGGL(k_ot_check, dim3(g), dim3(256), 0, c->stream, s, q, (const uint32_t*)srun, (const uint32_t*)srow, n, rflag);
        hipLaunchKernelGGL(k_ot_flip, dim3(g), dim3(256), 0, c->stream, (const uint32_t*)b, (const int64_t*)ex, (const uint32_t*)rflag,
                           n, dir);
        if((e = check_launch("order_ties"))) R e;
        arena_set(c, mark);
    }
    // the full-tie groups, each turned around where its direction is descending
    hipLaunchKernelGGL(k_ot_bound, dim3(g), dim3(256), 0, c->stream, (const uint8_t*)lvl, n, (int)nc, b);
    if((e = scan_u32_to_i64(c, b, ex, n, ex + n))) R e;
    int64_t starts = 0; if((e = d2h(c, &starts, ex + n, 8))) R e;
    const L ng = starts + 1;                                 // group starts after position 0, plus the first
    int64_t *gs = (int64_t*)dalloc(c, 8 * ng);
    uint32_t *out = (uint32_t*)dalloc(c, 4 * n);
    if(!gs || !out) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_ot_gstart, dim3(g), dim3(256), 0, c->stream, (const uint32_t*)b, (const int64_t*)ex, n, gs);
    hipLaunchKernelGGL(k_ot_apply, dim3(g), dim3(256), 0, c->stream, perm, (const uint32_t*)b, (const int64_t*)ex, (const int64_t*)gs,
                       (const uint32_t*)dir, n, ng, out);
    if((e = check_launch("order_ties"))) R e;
    *perm_io = out;
    R 0;
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
        uint32_t *perm = nullptr; bool ties, known;
        if(nc == 0 || n == 0){ void *dz; R out_init(c, z, HPA_L, 0, &dz); }
        // columns sorted by lib_quicksort_other or lib_quicksort_char: with rows tied on every
        // column the reference's column-by-column passes decide their order (hpa_qsort.hip)
        bool ref_cols = false;
        for(L q = 0; q < nc; q++) ref_cols |= qsort_ref_type(a[q].typ) || a[q].typ == HPA_C;
        if((e = order_perm(c, x, fl.data(), &perm, &ties, &known))) R e;
        if(ref_cols){
            bool t2 = known ? ties : true;
            if(!known && nc <= SMALL_COLS && (e = full_ties(c, (int)nc, a, perm, n, &t2))) R e;
            if(!t2) R order_emit(c, z, perm, n, false);          // the lexicographic order is unique
            std::vector<int32_t> ty((size_t)nc); std::vector<const void*> cp((size_t)nc);
            for(L q = 0; q < nc; q++){ ty[q] = a[q].typ; if((e = dev_view(c, a + q, &cp[q]))) R e; }
            if((e = order_list_ref(c, (int)nc, ty.data(), cp.data(), fl.data(), n, &perm))) R e;
            R order_emit(c, z, perm, n, false);
        }
        if((e = order_list_ties(c, x, fl.data(), &perm, n, known ? (ties ? 1 : 0) : -1))) R e;
        R order_emit(c, z, perm, n, false);
    }
    // lib_order_basic (h_primitive.cc:1760-1767)
    if(is_basic(x->typ) && y->typ == HPA_B && y->len == 1){
        uint8_t f; if((e = hpaReadPayload(y, &f, 1))) R e;
        R order_vector(c, z, x, f != 0);
    }
    R HPA_E_DOMAIN;
}
