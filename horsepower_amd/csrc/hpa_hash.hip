// Hash primitives: join_index, index_of, member.
//
// join_index (h_primitive.cc:1944-2155, backend/h_hash_v3.cc:230-277): the shorter side is the
// build side (ties: x), the longer side probes.  Pairs come out probe-row ascending and, within
// one probe row, build rows ascending -- the reference's insertion-order duplicate arrays.  The
// build side is radix-sorted into runs of equal keys (rows ascending inside a run); an
// open-addressing table maps each distinct key to its run.  Probe pass 1 stores each probe row's
// run and match count, an exclusive scan places the rows, pass 2 writes the pairs.
// index_of (h_primitive.cc:1525-1579, h_libs.cc:299-373): first occurrence via atomicMin on the
// slot of each key; an ordered (monotone) i32/i64 x is searched with the reference's binary
// search (backend/h_system.cc:715-745), which returns whichever equal element it meets first.
// member (h_primitive.cc:2288-2336, h_libs.cc:897-1018): set membership.
// Float keys compare with == (so -0 matches +0 and NaN matches nothing).
#include "hpa_internal.h"

#define EMPTY_KEY 0xffffffffffffffffull

__device__ __forceinline__ uint64_t mix64(uint64_t z){
    z ^= z >> 33; z *= 0xff51afd7ed558ccdull; z ^= z >> 33; z *= 0xc4ceb9fe1a85ec53ull; z ^= z >> 33; return z;
}

// key bits of element i of a typed column (after promotion)
template<typename T> __device__ __forceinline__ uint64_t kbits(T v){ return (uint64_t)(int64_t)v; }
template<> __device__ __forceinline__ uint64_t kbits<float>(float v){ if(v == 0.0f) v = 0.0f; return (uint64_t)__float_as_uint(v); }
template<> __device__ __forceinline__ uint64_t kbits<double>(double v){ if(v == 0.0) v = 0.0; return (uint64_t)__double_as_longlong(v); }
template<typename T> __device__ __forceinline__ bool knan(T v){ return false; }
template<> __device__ __forceinline__ bool knan<float>(float v){ return v != v; }
template<> __device__ __forceinline__ bool knan<double>(double v){ return v != v; }

struct HTab { uint64_t *keys; unsigned long long *vals; uint64_t mask; unsigned long long *ones_val; };

__device__ __forceinline__ unsigned long long *ht_slot(const HTab &t, uint64_t k, bool insert){
    if(k == EMPTY_KEY) return t.ones_val;
    uint64_t h = mix64(k) & t.mask;
    while(true){
        uint64_t cur = t.keys[h];
        if(cur == k) return &t.vals[h];
        if(cur == EMPTY_KEY){
            if(!insert) return nullptr;
            unsigned long long prev = atomicCAS((unsigned long long*)&t.keys[h], (unsigned long long)EMPTY_KEY, (unsigned long long)k);
            if(prev == EMPTY_KEY || prev == k) return &t.vals[h];
            continue;    // lost the race to another key: re-examine this slot
        }
        h = (h + 1) & t.mask;
    }
}
__device__ __forceinline__ const unsigned long long *ht_find(const HTab &t, uint64_t k){
    if(k == EMPTY_KEY) return (*t.ones_val != ~0ull) ? t.ones_val : nullptr;
    uint64_t h = mix64(k) & t.mask;
    while(true){
        uint64_t cur = t.keys[h];
        if(cur == k) return &t.vals[h];
        if(cur == EMPTY_KEY) return nullptr;
        h = (h + 1) & t.mask;
    }
}

__global__ void k_ht_clear(uint64_t *keys, unsigned long long *vals, L cap, unsigned long long *ones){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (L)gridDim.x * blockDim.x){ keys[i] = EMPTY_KEY; vals[i] = ~0ull; }
    if(blockIdx.x == 0 && threadIdx.x == 0) *ones = ~0ull;
}
static int32_t ht_create(hpa_ctx *c, L n, HTab *t){
    L cap = 16; while(cap < 2 * n) cap <<= 1;
    t->keys = (uint64_t*)dalloc(c, 8 * cap); t->vals = (unsigned long long*)dalloc(c, 8 * cap);
    t->ones_val = (unsigned long long*)dalloc(c, 16);
    if(!t->keys || !t->vals || !t->ones_val) R HPA_E_MEMORY;
    t->mask = (uint64_t)cap - 1;
    hipLaunchKernelGGL(k_ht_clear, dim3(grid_for(cap, 256, 4)), dim3(256), 0, c->stream, t->keys, t->vals, cap, t->ones_val);
    R check_launch("hash_clear");
}

__global__ void k_fill_i32(int *p, int n, int v){ for(int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v; }
// Packed table for 4-byte keys: one 8-byte slot per distinct key, (key << 32) | row, inserted
// with a single CAS; a key met again keeps the smaller row (first occurrence, as index_of needs)
// and raises `dup` (a join whose build keys repeat then takes its sort-based path, whose output
// this path equals whenever the keys are unique: each probe row matches at most one build row).
#define PK_EMPTY 0xffffffffffffffffull
__global__ void k_pk_clear(unsigned long long *slots, L cap){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (L)gridDim.x * blockDim.x) slots[i] = PK_EMPTY;
}
// Lanes of a wave hold consecutive rows.  Returns whether this lane is the lowest of the lanes
// holding its key (up to four keys a wave are matched this way; lanes left over return true),
// and sets *d for a lane whose key a lower lane holds: one table update per key and wave
// instead of one per row, so a key repeated in many rows does not serialise on its slot.
__device__ __forceinline__ bool wave_key_leader(bool valid, uint32_t k, int *d){
    const int lane = threadIdx.x & 63;
    bool pending = valid, lead = false;
    for(int it = 0; it < 4; it++){
        const uint64_t pm = __ballot(pending);
        if(!pm) break;
        const int leader = __ffsll((unsigned long long)pm) - 1;
        const uint32_t kl = __shfl(k, leader, 64);
        if(pending && k == kl){
            if(lane == leader) lead = true; else *d = 1;
            pending = false;
        }
    }
    return lead || pending;
}
__global__ void k_pk_insert(const int32_t *x, L n, unsigned long long *slots, uint64_t mask, int *dup){
    int d = 0;
    for(L i0 = (L)blockIdx.x * blockDim.x; i0 < n; i0 += (L)gridDim.x * blockDim.x){
        const L i = i0 + threadIdx.x;
        const bool valid = i < n;
        const uint32_t k = valid ? (uint32_t)x[i] : 0u;
        if(!wave_key_leader(valid, k, &d)) continue;
        const unsigned long long want = ((unsigned long long)k << 32) | (unsigned long long)(uint32_t)i;
        uint64_t h = mix64(k) & mask;
        while(true){
            // the CAS returns the slot's content: no separate read before it
            const unsigned long long cur = atomicCAS(&slots[h], PK_EMPTY, want);
            if(cur == PK_EMPTY) break;
            if((uint32_t)(cur >> 32) == k){ d = 1; if(cur > want) atomicMin(&slots[h], want); break; }
            h = (h + 1) & mask;
        }
    }
    if(__any(d) && (threadIdx.x & 63) == 0 && !*(volatile int*)dup) atomicOr(dup, 1);
}
__global__ void k_pk_probe(const int32_t *y, L ny, const unsigned long long *slots, uint64_t mask, uint32_t *cnt, uint32_t *rid){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const uint32_t k = (uint32_t)y[j];
        uint64_t h = mix64(k) & mask;
        uint32_t c = 0, r = 0;
        while(true){
            const unsigned long long cur = slots[h];
            if(cur == PK_EMPTY) break;
            if((uint32_t)(cur >> 32) == k){ c = 1; r = (uint32_t)cur; break; }
            h = (h + 1) & mask;
        }
        cnt[j] = c; rid[j] = r;
    }
}
__global__ void k_pk_write(L ny, const uint32_t *cnt, const uint32_t *rid, const int64_t *off, int64_t *z0, int64_t *z1, int swap){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        if(!cnt[j]) continue;
        const int64_t o = off[j];
        if(swap){ z0[o] = j; z1[o] = rid[j]; } else { z0[o] = rid[j]; z1[o] = j; }
    }
}

__global__ void k_pk_probe_index(const int32_t *y, L ny, const unsigned long long *slots, uint64_t mask, int64_t *z, int64_t miss){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const uint32_t k = (uint32_t)y[j];
        uint64_t h = mix64(k) & mask;
        int64_t r = miss;
        while(true){
            const unsigned long long cur = slots[h];
            if(cur == PK_EMPTY) break;
            if((uint32_t)(cur >> 32) == k){ r = (int64_t)(uint32_t)cur; break; }
            h = (h + 1) & mask;
        }
        z[j] = r;
    }
}
__global__ void k_pk_probe_member(const int32_t *y, L ny, const unsigned long long *slots, uint64_t mask, uint8_t *z){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const uint32_t k = (uint32_t)y[j];
        uint64_t h = mix64(k) & mask;
        uint8_t r = 0;
        while(true){
            const unsigned long long cur = slots[h];
            if(cur == PK_EMPTY) break;
            if((uint32_t)(cur >> 32) == k){ r = 1; break; }
            h = (h + 1) & mask;
        }
        z[j] = r;
    }
}
// packed table over x (4-byte keys): clear + insert; returns the slots and mask
static int32_t pk_build(hpa_ctx *c, const int32_t *x, L n, unsigned long long **slots, uint64_t *mask, int **dup){
    L cap = 16; while(cap < 2 * n) cap <<= 1;
    *slots = (unsigned long long*)dalloc(c, 8 * cap);
    *dup = (int*)dalloc(c, 16);
    if(!*slots || !*dup) R HPA_E_MEMORY;
    *mask = (uint64_t)(cap - 1);
    hipLaunchKernelGGL(k_pk_clear, dim3(grid_for(cap, 256, 4)), dim3(256), 0, c->stream, *slots, cap);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, *dup, 4, 0);
    if(n > 0) hipLaunchKernelGGL(k_pk_insert, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, x, n, *slots, *mask, *dup);
    R check_launch("hash_build");
}

// ---------------------------------------------------------------- dense i32 key ranges
// Build keys that span at most DENSE_PER_KEY slots per key (surrogate keys, dates, codes) are
// addressed directly by k - lo: no hashing, no probe chains, and no atomics for unique keys.
//   index_of / join: tab[k - lo] = first row of k.  Every row stores its index (any one wins),
//     then every row whose key shows another row lowers it with atomicMin: only repeated keys
//     pay an atomic, and the table ends at the first occurrence.  `dup` is raised then.  (A
//     ranked form, a bitmap plus prefix counts indexing 4 bytes per distinct key, probed slower:
//     its three dependent reads cost more than one read of the larger table.)
//   member: one flag byte per key value (plain stores), packed into a bitmap of the range that
//     stays in every XCD's L2 (750 KB for 6M values).
// Sparser ranges, up to GUARD_PER_KEY slots per key and GUARD_MAX_RANGE slots, keep the direct
// table for index_of and join but guard it with a presence bitmap: only the bitmap is cleared
// (one bit per slot), the table is written at the keys present and read only where the bit is
// set.  The bitmap is small enough to stay in cache (7.5 MB for a 60M-key range), and probes
// that miss touch nothing else.  Taken when the arena has room for the table's address range.
#define DENSE_PER_KEY 8
#define GUARD_PER_KEY 512
#define GUARD_MAX_RANGE (1ll << 30)
// smallest and largest key: per-workgroup partials part[2 b], part[2 b + 1] (16-byte loads
// when aligned), folded by one workgroup into out[0] min, out[1] max
__device__ __forceinline__ void i32_minmax_block(int lo, int hi, int *part){
    __shared__ int slo[4], shi[4];
    for(int o = 32; o > 0; o >>= 1){ int t = __shfl_xor(lo, o, 64); lo = t < lo ? t : lo; t = __shfl_xor(hi, o, 64); hi = t > hi ? t : hi; }
    if((threadIdx.x & 63) == 0){ slo[threadIdx.x >> 6] = lo; shi[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if(threadIdx.x == 0){
        for(int q = 1; q < 4; q++){ lo = slo[q] < lo ? slo[q] : lo; hi = shi[q] > hi ? shi[q] : hi; }
        part[0] = lo; part[1] = hi;
    }
}
__global__ void __launch_bounds__(256) k_i32_range(const int32_t *x, L n, int vec, int *part){
    int lo = 2147483647, hi = -2147483647 - 1;
    const L S = (L)gridDim.x * 256;
    const L nv = vec ? n / 4 : 0;
    for(L i = (L)blockIdx.x * 256 + threadIdx.x; i < nv; i += S){
        const int4 q = ((const int4*)x)[i];
        lo = min(lo, min(min(q.x, q.y), min(q.z, q.w))); hi = max(hi, max(max(q.x, q.y), max(q.z, q.w)));
    }
    for(L i = 4 * nv + (L)blockIdx.x * 256 + threadIdx.x; i < n; i += S){ const int v = x[i]; lo = v < lo ? v : lo; hi = v > hi ? v : hi; }
    i32_minmax_block(lo, hi, part + 2 * blockIdx.x);
}
__global__ void __launch_bounds__(256) k_i32_range_final(const int *part, int np, int *out){
    int lo = 2147483647, hi = -2147483647 - 1;
    for(int b = threadIdx.x; b < np; b += 256){ lo = min(lo, part[2 * b]); hi = max(hi, part[2 * b + 1]); }
    i32_minmax_block(lo, hi, out);
}
// Monotonicity (isOrdered, backend/h_system.cc:678-698) and range of an i32 column in one read:
// part[4 b ...] = (rises, falls, min, max) of workgroup b's rows and the pair before each
__global__ void __launch_bounds__(256) k_i32_order_range(const int32_t *x, L n, int *part){
    __shared__ int s[4][4];
    int up = 0, down = 0, lo = 2147483647, hi = -2147483647 - 1;
    for(L i = (L)blockIdx.x * 256 + threadIdx.x; i < n; i += (L)gridDim.x * 256){
        const int v = x[i];
        lo = v < lo ? v : lo; hi = v > hi ? v : hi;
        if(i > 0){ const int a = x[i - 1]; up |= v > a; down |= v < a; }
    }
    for(int o = 32; o > 0; o >>= 1){
        up |= __shfl_xor(up, o, 64); down |= __shfl_xor(down, o, 64);
        int t = __shfl_xor(lo, o, 64); lo = t < lo ? t : lo; t = __shfl_xor(hi, o, 64); hi = t > hi ? t : hi;
    }
    const int w = threadIdx.x >> 6;
    if((threadIdx.x & 63) == 0){ s[w][0] = up; s[w][1] = down; s[w][2] = lo; s[w][3] = hi; }
    __syncthreads();
    if(threadIdx.x == 0){
        for(int q = 1; q < 4; q++){ up |= s[q][0]; down |= s[q][1]; lo = min(lo, s[q][2]); hi = max(hi, s[q][3]); }
        part[4 * blockIdx.x] = up; part[4 * blockIdx.x + 1] = down; part[4 * blockIdx.x + 2] = lo; part[4 * blockIdx.x + 3] = hi;
    }
}
__global__ void __launch_bounds__(256) k_i32_order_range_final(const int *part, int np, int *out){
    __shared__ int s[4][4];
    int up = 0, down = 0, lo = 2147483647, hi = -2147483647 - 1;
    for(int b = threadIdx.x; b < np; b += 256){ up |= part[4 * b]; down |= part[4 * b + 1]; lo = min(lo, part[4 * b + 2]); hi = max(hi, part[4 * b + 3]); }
    for(int o = 32; o > 0; o >>= 1){
        up |= __shfl_xor(up, o, 64); down |= __shfl_xor(down, o, 64);
        int t = __shfl_xor(lo, o, 64); lo = t < lo ? t : lo; t = __shfl_xor(hi, o, 64); hi = t > hi ? t : hi;
    }
    const int w = threadIdx.x >> 6;
    if((threadIdx.x & 63) == 0){ s[w][0] = up; s[w][1] = down; s[w][2] = lo; s[w][3] = hi; }
    __syncthreads();
    if(threadIdx.x == 0){
        for(int q = 1; q < 4; q++){ up |= s[q][0]; down |= s[q][1]; lo = min(lo, s[q][2]); hi = max(hi, s[q][3]); }
        out[0] = up; out[1] = down; out[2] = lo; out[3] = hi;
    }
}
// ordered (monotone in either direction) and [lo, hi] of an i32 column, with one host read
static int32_t i32_order_range(hpa_ctx *c, const int32_t *x, L n, bool *ord, int *lo, int *hi){
    *ord = n <= 2; *lo = 0; *hi = -1;
    if(n <= 0) R 0;
    const int np = grid_for(n, 256, 8, num_cus() * 4);
    int *r = (int*)dalloc(c, 16 * (L)(np + 1)); if(!r) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_i32_order_range, dim3(np), dim3(256), 0, c->stream, x, n, r + 4);
    hipLaunchKernelGGL(k_i32_order_range_final, dim3(1), dim3(256), 0, c->stream, r + 4, np, r);
    int32_t e;
    if((e = check_launch("order_range"))) R e;
    int h[4]; if((e = d2h(c, h, r, 16))) R e;
    if(n > 2) *ord = !(h[0] && h[1]);
    *lo = h[2]; *hi = h[3];
    R 0;
}
// [lo, hi] of n i32 keys; *dense when they span at most DENSE_PER_KEY slots per key, *guard
// (when asked) when they span at most GUARD_PER_KEY slots per key and GUARD_MAX_RANGE slots; a
// range known already comes in through known (lo, hi)
static int32_t dense_range(hpa_ctx *c, const int32_t *x, L n, int *lo, int *hi, bool *dense, const int *known = nullptr,
                           bool *guard = nullptr){
    *dense = false;
    if(guard) *guard = false;
    if(n < 1024) R 0;                                  // small builds: the hash table is cheap
    if(known){
        *lo = known[0]; *hi = known[1];
        const L R_ = (L)known[1] - (L)known[0] + 1;
        *dense = R_ <= (L)DENSE_PER_KEY * n && R_ < (L)INT32_MAX;
        if(guard) *guard = !*dense && R_ <= (L)GUARD_PER_KEY * n && R_ <= GUARD_MAX_RANGE;
        R 0;
    }
    const int np = grid_for(ceil_div(n, 4), 256, 4, num_cus() * 4);
    int *r = (int*)dalloc(c, 8 * (L)(np + 1)); if(!r) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_i32_range, dim3(np), dim3(256), 0, c->stream, x, n, (uintptr_t)x % 16 == 0 ? 1 : 0, r + 2);
    hipLaunchKernelGGL(k_i32_range_final, dim3(1), dim3(256), 0, c->stream, r + 2, np, r);
    int32_t e;
    if((e = check_launch("dense_range"))) R e;
    int h[2]; if((e = d2h(c, h, r, 8))) R e;
    *lo = h[0]; *hi = h[1];
    const L R_ = (L)h[1] - (L)h[0] + 1;
    *dense = R_ <= (L)DENSE_PER_KEY * n && R_ < (L)INT32_MAX;
    if(guard) *guard = !*dense && R_ <= (L)GUARD_PER_KEY * n && R_ <= GUARD_MAX_RANGE;
    R 0;
}
__global__ void k_dense_flag(const int32_t *x, L n, int lo, uint8_t *flag){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) flag[(uint32_t)(x[i] - lo)] = 1;
}
// bits[w] bit b = flag[32 w + b] (flag zero past m); cnt[w] = its population (cnt may be null)
__global__ void k_dense_pack(const uint8_t *flag, L m, uint32_t *bits, uint32_t *cnt, L nw){
    for(L w = (L)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (L)gridDim.x * blockDim.x){
        uint32_t v = 0;
        if(32 * w + 32 <= m){
            const uint4 *q = (const uint4*)(flag + 32 * w);
            const uint4 a = q[0], b = q[1];
            const unsigned wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            #pragma unroll
            for(int k = 0; k < 32; k++) v |= ((wd[k >> 2] >> (8 * (k & 3))) & 1u) << k;
        } else {
            for(int k = 0; k < 32 && 32 * w + k < m; k++) v |= (uint32_t)(flag[32 * w + k] != 0) << k;
        }
        bits[w] = v;
        if(cnt) cnt[w] = (uint32_t)__popc(v);
    }
}
__global__ void k_dense_clear(int32_t *tab, L m){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (L)gridDim.x * blockDim.x) tab[i] = -1;
}
__global__ void k_dense_store(const int32_t *x, L n, int lo, int32_t *tab){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        tab[(uint32_t)(x[i] - lo)] = (int32_t)i;
}
// gated: run only when *dup is raised already (a guarded build counted fewer keys than rows)
__global__ void k_dense_first(const int32_t *x, L n, int lo, int32_t *tab, int *dup, int gated){
    if(gated && !*(volatile int*)dup) return;
    int d = 0;
    for(L i0 = (L)blockIdx.x * blockDim.x; i0 < n; i0 += (L)gridDim.x * blockDim.x){
        const L i = i0 + threadIdx.x;
        const bool valid = i < n;
        const uint32_t k = valid ? (uint32_t)(x[i] - lo) : 0u;
        const bool lead = wave_key_leader(valid, k, &d);
        if(!valid) continue;
        const int32_t t = tab[k];
        if(t != (int32_t)i){ d = 1; if(lead && t > (int32_t)i) atomicMin(&tab[k], (int32_t)i); }
    }
    if(__any(d) && (threadIdx.x & 63) == 0 && !*(volatile int*)dup) atomicOr(dup, 1);
}
// pres: presence bitmap of a guarded table (null: every slot of tab is written)
struct DenseTab { int lo; L m; uint32_t *bits; int32_t *tab; int *dup; const uint32_t *pres; };
__global__ void k_fill_u32(uint32_t *p, L n, uint32_t v){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) p[i] = v;
}
// set bits of the presence bitmap, one partial per workgroup (no atomics on one address)
#define POPC_PARTS 1024
__global__ void __launch_bounds__(256) k_popc_part(const uint32_t *bits, L nw, unsigned long long *part){
    __shared__ unsigned long long sh[4];
    unsigned long long s = 0;
    for(L w = (L)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (L)gridDim.x * blockDim.x) s += (unsigned)__popc(bits[w]);
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
// fewer distinct keys (set bits) than build rows: some key repeats
__global__ void __launch_bounds__(256) k_dup_if_fewer(const unsigned long long *part, int np, L n, int *dup){
    __shared__ unsigned long long sh[4];
    unsigned long long s = 0;
    for(int b = threadIdx.x; b < np; b += 256) s += part[b];
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0 && sh[0] + sh[1] + sh[2] + sh[3] != (unsigned long long)n) *dup = 1;
}
// guarded table: every build row stores its index and sets its key's presence bit
__global__ void k_guard_store(const int32_t *x, L n, int lo, int32_t *tab, uint32_t *pres){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const uint32_t k = (uint32_t)(x[i] - lo);
        tab[k] = (int32_t)i;
        atomicOr(&pres[k >> 5], 1u << (k & 31));
    }
}
// first-row table (rows) or bitmap (member) over x[0..n), keys in [lo, lo + m)
static int32_t dense_build(hpa_ctx *c, const int32_t *x, L n, int lo, L m, bool rows, DenseTab *t){
    t->lo = lo; t->m = m; t->bits = nullptr; t->tab = nullptr; t->pres = nullptr;
    t->dup = (int*)dalloc(c, 16);
    if(!t->dup) R HPA_E_MEMORY;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, t->dup, 4, 0);
    if(rows){
        t->tab = (int32_t*)dalloc(c, 4 * m); if(!t->tab) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_dense_clear, dim3(grid_for(m, 256, 4)), dim3(256), 0, c->stream, t->tab, m);
        hipLaunchKernelGGL(k_dense_store, dim3(grid_for(n, 256, 1, 1 << 30)), dim3(256), 0, c->stream, x, n, lo, t->tab);
        hipLaunchKernelGGL(k_dense_first, dim3(grid_for(n, 256, 1, 1 << 30)), dim3(256), 0, c->stream, x, n, lo, t->tab, t->dup, 0);
        R check_launch("dense_build");
    }
    const L nw = ceil_div(m, 32);
    uint8_t *flag = (uint8_t*)dalloc(c, m + 32);
    t->bits = (uint32_t*)dalloc(c, 4 * nw);
    if(!flag || !t->bits) R HPA_E_MEMORY;
    hipMemsetAsync(flag, 0, (size_t)(m + 32), c->stream);
    hipLaunchKernelGGL(k_dense_flag, dim3(grid_for(n, 256, 1, 1 << 30)), dim3(256), 0, c->stream, x, n, lo, flag);
    hipLaunchKernelGGL(k_dense_pack, dim3(grid_for(nw, 256, 1)), dim3(256), 0, c->stream, flag, m, t->bits, (uint32_t*)nullptr, nw);
    R check_launch("dense_build");
}
// guarded first-row table over keys [lo, lo + m); *room false (nothing kept) when the arena
// cannot hold the table's address range
static int32_t guard_build(hpa_ctx *c, const int32_t *x, L n, int lo, L m, DenseTab *t, bool *room){
    *room = false;
    const L mark = c->dcur, nw = ceil_div(m, 32);
    t->lo = lo; t->m = m; t->bits = nullptr;
    t->dup = (int*)dalloc(c, 16);
    unsigned long long *part = (unsigned long long*)dalloc(c, 8 * POPC_PARTS);
    uint32_t *pres = (uint32_t*)dalloc(c, 4 * nw);
    t->tab = (int32_t*)dalloc(c, 4 * m);
    if(!t->dup || !part || !pres || !t->tab){ arena_set(c, mark); R 0; }
    *room = true;
    t->pres = pres;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, t->dup, 4, 0);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nw, 256, 4)), dim3(256), 0, c->stream, pres, nw, 0u);
    hipLaunchKernelGGL(k_guard_store, dim3(grid_for(n, 256, 1, 1 << 30)), dim3(256), 0, c->stream, x, n, lo, t->tab, pres);
    // keys are unique exactly when the bitmap holds as many bits as there are rows; only then is
    // the first-row pass skipped
    const int np = grid_for(nw, 256, 8, POPC_PARTS);
    hipLaunchKernelGGL(k_popc_part, dim3(np), dim3(256), 0, c->stream, (const uint32_t*)pres, nw, part);
    hipLaunchKernelGGL(k_dup_if_fewer, dim3(1), dim3(256), 0, c->stream, (const unsigned long long*)part, np, n, t->dup);
    hipLaunchKernelGGL(k_dense_first, dim3(grid_for(n, 256, 1, 1 << 30)), dim3(256), 0, c->stream, x, n, lo, t->tab, t->dup, 1);
    R check_launch("guard_build");
}
// row of key y under the table: -1 when absent (a guarded table is read only where its bit is set)
__device__ __forceinline__ int32_t dense_row(const DenseTab &t, int32_t y){
    const int64_t d = (int64_t)y - t.lo;
    if(d < 0 || d >= t.m) return -1;
    if(t.pres && !((t.pres[d >> 5] >> (d & 31)) & 1u)) return -1;
    return t.tab[d];
}
__global__ void k_dense_probe_index(const int32_t *y, L ny, DenseTab t, int64_t *z, int64_t miss){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const int32_t r = dense_row(t, y[j]);
        z[j] = r >= 0 ? (int64_t)r : miss;
    }
}
__global__ void k_dense_probe_join(const int32_t *y, L ny, DenseTab t, uint32_t *cnt, uint32_t *rid){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const int32_t r = dense_row(t, y[j]);
        cnt[j] = r >= 0 ? 1u : 0u; rid[j] = r >= 0 ? (uint32_t)r : 0u;
    }
}
// Join on a guarded table with unique build keys: every probe row matches at most once, so the
// pairs are a compaction of the probe rows.  k_dj_count counts each 4096-row tile's matches;
// after a scan of the tile counts, k_dj_write probes the tile again and writes its pairs at
// their rank in row order: per row of 64 lanes a wave ballot, the wave counts of the tile's 16
// rows x 4 waves scanned in (row, wave) order.  A probe of a guarded table is a bit test in a
// bitmap that stays in cache, and the table is read only on a hit, so probing twice costs less
// than per-row count and build-row arrays and their scan (Q3's join at 60M: 238 -> 208 us).  On
// a full direct table every probe reads the table, and random probes that mostly hit are
// cheaper probed once (1.5M keys x 6M probes: 241 us once, 291 us twice).
#define DJ_ITEMS 16
#define DJ_TILE (256 * DJ_ITEMS)
// dense_row of the DJ_ITEMS rows base + 256 k + threadIdx.x in three rounds of loads -- keys,
// bitmap words, table entries -- each round's loads at clamped indices and in flight together
// (rows past n and keys out of range read entry 0 and come out -1).  With `need_row` false on
// a guarded table the table is not read: rr[k] is 0 for a hit.
__device__ __forceinline__ void dense_rows(const DenseTab &t, const int32_t *y, L base, L n, bool need_row, int32_t *rr){
    int64_t d[DJ_ITEMS]; bool in[DJ_ITEMS];
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; d[k] = (int64_t)y[i < n ? i : n - 1] - t.lo; }
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++){ const L i = base + (L)k * 256 + threadIdx.x; in[k] = i < n && d[k] >= 0 && d[k] < t.m; }
    if(t.pres){
        uint32_t wv[DJ_ITEMS];
        #pragma unroll
        for(int k = 0; k < DJ_ITEMS; k++) wv[k] = t.pres[in[k] ? (d[k] >> 5) : 0];
        #pragma unroll
        for(int k = 0; k < DJ_ITEMS; k++) in[k] = in[k] && ((wv[k] >> (d[k] & 31)) & 1u);
        if(!need_row){
            #pragma unroll
            for(int k = 0; k < DJ_ITEMS; k++) rr[k] = in[k] ? 0 : -1;
            return;
        }
    }
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++) rr[k] = t.tab[in[k] ? d[k] : 0];
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++) rr[k] = in[k] ? rr[k] : -1;
}
__global__ void __launch_bounds__(256) k_dj_count(const int32_t *y, L n, DenseTab t, uint32_t *tcount){
    const L base = (L)blockIdx.x * DJ_TILE;
    uint32_t c = 0;
    int32_t rr[DJ_ITEMS];
    dense_rows(t, y, base, n, false, rr);
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++) c += rr[k] >= 0 ? 1u : 0u;
    __shared__ uint32_t sh[4];
    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0) tcount[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void __launch_bounds__(256) k_dj_write(const int32_t *y, L n, DenseTab t, const int64_t *toff,
                                                  int64_t *z0, int64_t *z1, int swap){
    __shared__ uint32_t s_cnt[DJ_ITEMS * 4];
    const L base = (L)blockIdx.x * DJ_TILE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t bal[DJ_ITEMS];
    int32_t rr[DJ_ITEMS];
    dense_rows(t, y, base, n, true, rr);
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++){
        bal[k] = __ballot(rr[k] >= 0);
        if(lane == 0) s_cnt[k * 4 + w] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    if(w == 0){
        const uint32_t mine = s_cnt[lane];
        uint32_t inc = mine;
        for(int o = 1; o < 64; o <<= 1){ const uint32_t tt = __shfl_up(inc, o, 64); if(lane >= o) inc += tt; }
        s_cnt[lane] = inc - mine;
    }
    __syncthreads();
    const int64_t pre = toff[blockIdx.x];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    #pragma unroll
    for(int k = 0; k < DJ_ITEMS; k++){
        if(rr[k] >= 0){
            const L i = base + (L)k * 256 + threadIdx.x;
            const int64_t o = pre + s_cnt[k * 4 + w] + __popcll(bal[k] & lt);
            if(swap){ z0[o] = i; z1[o] = rr[k]; } else { z0[o] = rr[k]; z1[o] = i; }
        }
    }
}
__global__ void k_dense_probe_member(const int32_t *y, L ny, int lo, L m, const uint32_t *bits, uint8_t *z){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const int64_t d = (int64_t)y[j] - lo;
        z[j] = (d >= 0 && d < m) ? (uint8_t)((bits[d >> 5] >> (d & 31)) & 1u) : 0;
    }
}

// ---------------------------------------------------------------- index_of / member build+probe
template<typename T>
__global__ void k_ht_insert_min(const T *x, L n, HTab t){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T v = x[i]; if(knan(v)) continue;
        unsigned long long *s = ht_slot(t, kbits(v), true);
        atomicMin(s, (unsigned long long)i);
    }
}
template<typename T>
__global__ void k_ht_probe_index(const T *y, L n, HTab t, int64_t *z, int64_t miss){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T v = y[i];
        const unsigned long long *s = knan(v) ? nullptr : ht_find(t, kbits(v));
        z[i] = (s && *s != ~0ull) ? (int64_t)*s : miss;
    }
}
template<typename T>
__global__ void k_ht_probe_member(const T *y, L n, HTab t, uint8_t *z){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T v = y[i];
        const unsigned long long *s = knan(v) ? nullptr : ht_find(t, kbits(v));
        z[i] = (s && *s != ~0ull) ? 1 : 0;
    }
}
// small sets (<= 2 items): the reference compares with == directly (h_libs.cc:899-944)
template<typename T>
__global__ void k_member_small(const T *y, L n, const T *x, int nx, uint8_t *z){
    T a = nx > 0 ? x[0] : T(), b = nx > 1 ? x[1] : T();
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T v = y[i]; z[i] = (nx > 0 && v == a) || (nx > 1 && v == b);
    }
}
// bool / char lookup tables (lib_index_of_B/C, lib_member_B/C)
__global__ void k_table_probe_index(const uint8_t *y, L n, const int64_t *tab, int64_t *z, int64_t miss, int nt){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){ unsigned v = y[i]; if(nt == 2) v = v != 0; int64_t t = tab[v]; z[i] = t < 0 ? miss : t; }
}
__global__ void k_table_build_first(const uint8_t *x, L n, unsigned long long *tab, int nt){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){ unsigned v = x[i]; if(nt == 2) v = v != 0; atomicMin(&tab[v], (unsigned long long)i); }
}
__global__ void k_table_fix(unsigned long long *tab, int64_t *out, int nt, int bool_bug){
    for(int i = threadIdx.x; i < nt; i += blockDim.x){
        int64_t v = tab[i] == ~0ull ? -1 : (int64_t)tab[i];
        // lib_index_of_B initialises flag[2]={-1} so flag[1] starts at 0 and is never set
        if(bool_bug && i == 1) v = 0;
        out[i] = v;
    }
}

// binarySearch_I/L (backend/h_system.cc:715-729) restated iteratively
template<typename T>
__global__ void k_search_ordered(const T *x, L nx, const T *y, L n, int64_t *z){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        T val = y[i]; L low = 0, high = nx, r = -1;
        if(nx > 0){
            while(true){
                L mid = (low + high) >> 1; T xm = x[mid];
                if(xm == val){ r = mid; break; }
                else if(low == mid){ r = -1; break; }
                else if(xm > val) high = mid;
                else low = mid;
            }
        }
        z[i] = r < 0 ? nx : r;
    }
}

int32_t is_ordered_dev(hpa_ctx *c, V x, bool *ord);

template<typename T>
static int32_t index_of_hash(hpa_ctx *c, const T *px, L nx, const T *py, L ny, int64_t *z, const int *known = nullptr){
    int32_t e;
    if(std::is_same<T, int32_t>::value && nx < (L)INT32_MAX){
        int lo, hi; bool dense, guard;
        if((e = dense_range(c, (const int32_t*)px, nx, &lo, &hi, &dense, known, &guard))) R e;
        DenseTab t; bool room = false;
        if(guard && (e = guard_build(c, (const int32_t*)px, nx, lo, (L)hi - lo + 1, &t, &room))) R e;
        if(dense || room){
            if(dense && (e = dense_build(c, (const int32_t*)px, nx, lo, (L)hi - lo + 1, true, &t))) R e;
            ProfScope ps("index_of", c->stream, 4.0 * (nx + ny) + 8.0 * ny);
            if(ny > 0) hipLaunchKernelGGL(k_dense_probe_index, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)py, ny, t, z, (int64_t)nx);
            R check_launch("index_of");
        }
    }
    if(std::is_same<T, int32_t>::value && nx < (L)UINT32_MAX - 1){   // 4-byte keys: packed (key, first row) slots
        unsigned long long *slots; uint64_t mask; int *dup;
        if((e = pk_build(c, (const int32_t*)px, nx, &slots, &mask, &dup))) R e;
        ProfScope ps("index_of", c->stream, 4.0 * (nx + ny) + 8.0 * ny);
        if(ny > 0) hipLaunchKernelGGL(k_pk_probe_index, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)py, ny, slots, mask, z, (int64_t)nx);
        R check_launch("index_of");
    }
    HTab t;
    if((e = ht_create(c, nx, &t))) R e;
    ProfScope ps("index_of", c->stream, (double)sizeof(T) * (nx + ny) + 8.0 * ny);
    if(nx > 0) hipLaunchKernelGGL(k_ht_insert_min<T>, dim3(grid_for(nx, 256, 4)), dim3(256), 0, c->stream, px, nx, t);
    if(ny > 0) hipLaunchKernelGGL(k_ht_probe_index<T>, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, py, ny, t, z, (int64_t)nx);
    R check_launch("index_of");
}

static int32_t index_of_table(hpa_ctx *c, const uint8_t *px, L nx, const uint8_t *py, L ny, int64_t *z, int nt, bool boolbug){
    unsigned long long *tab = (unsigned long long*)dalloc(c, 8 * nt); int64_t *tt = (int64_t*)dalloc(c, 8 * nt);
    if(!tab || !tt) R HPA_E_MEMORY;
    hipMemsetAsync(tab, 0xff, 8 * nt, c->stream);
    if(nx > 0) hipLaunchKernelGGL(k_table_build_first, dim3(grid_for(nx, 256, 4)), dim3(256), 0, c->stream, px, nx, tab, nt);
    hipLaunchKernelGGL(k_table_fix, dim3(1), dim3(256), 0, c->stream, tab, tt, nt, boolbug ? 1 : 0);
    if(ny > 0) hipLaunchKernelGGL(k_table_probe_index, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, py, ny, tt, z, (int64_t)nx, nt);
    R check_launch("index_of");
}

extern "C" int32_t pfnIndexOf(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    int32_t e;
    L nx = x->len, ny = y->len;
    if(is_real(x->typ) && is_real(y->typ)){
        int32_t t = x->typ > y->typ ? x->typ : y->typ;
        bool ord = false;
        int known[2]; bool have = false;                 // i32 x: order and key range in one read
        if(x->typ == HPA_I){
            const void *p0 = nullptr;
            if(nx > 0 && (e = dev_view(c, x, &p0))) R e;
            if((e = i32_order_range(c, (const int32_t*)p0, nx, &ord, &known[0], &known[1]))) R e;
            have = nx > 0;
        } else if((e = is_ordered_dev(c, x, &ord))) R e;
        const void *px = nullptr, *py = nullptr;
        if(ord){
            // searchOrdered: y compared in x's integer type, widened when y is wider
            int32_t st = (y->typ == HPA_L || x->typ == HPA_L) ? HPA_L : HPA_I;
            if(!is_int(y->typ) || y->typ == HPA_B) st = HPA_L;
            if(!is_int(y->typ)) R HPA_E_NOT_IMPL;
            if(nx > 0 && (e = promote_dev(c, x, st, &px))) R e;
            if(ny > 0 && (e = promote_dev(c, y, st, &py))) R e;
            void *dz; L mark = c->dcur;
            if((e = out_init(c, z, HPA_L, ny, &dz))) R e;
            if(ny > 0){
                ProfScope ps("index_of", c->stream, (double)type_width(st) * (nx + ny) + 8.0 * ny);
                if(st == HPA_I) hipLaunchKernelGGL(k_search_ordered<int32_t>, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)px, nx, (const int32_t*)py, ny, (int64_t*)dz);
                else hipLaunchKernelGGL(k_search_ordered<int64_t>, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)px, nx, (const int64_t*)py, ny, (int64_t*)dz);
                if((e = check_launch("index_of"))) R e;
            }
            if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
            R e;
        }
        if(nx > 0 && (e = promote_dev(c, x, t, &px))) R e;
        if(ny > 0 && (e = promote_dev(c, y, t, &py))) R e;
        void *dz; L mark = c->dcur;
        if((e = out_init(c, z, HPA_L, ny, &dz))) R e;
        int64_t *zz = (int64_t*)dz;
        switch(t){
            case HPA_B: e = index_of_table(c, (const uint8_t*)px, nx, (const uint8_t*)py, ny, zz, 2, true); break;
            case HPA_J: e = index_of_hash<int8_t>(c, (const int8_t*)px, nx, (const int8_t*)py, ny, zz); break;
            case HPA_H: e = index_of_hash<int16_t>(c, (const int16_t*)px, nx, (const int16_t*)py, ny, zz); break;
            case HPA_I: e = index_of_hash<int32_t>(c, (const int32_t*)px, nx, (const int32_t*)py, ny, zz, have ? known : nullptr); break;
            case HPA_L: e = index_of_hash<int64_t>(c, (const int64_t*)px, nx, (const int64_t*)py, ny, zz); break;
            case HPA_F: e = index_of_hash<float>(c, (const float*)px, nx, (const float*)py, ny, zz); break;
            case HPA_E: e = index_of_hash<double>(c, (const double*)px, nx, (const double*)py, ny, zz); break;
        }
        if(e) R e;
        if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
        R e;
    }
    if(x->typ == y->typ && is_basic(x->typ)){
        const void *px = nullptr, *py = nullptr;
        if(nx > 0 && (e = dev_view(c, x, &px))) R e;
        if(ny > 0 && (e = dev_view(c, y, &py))) R e;
        void *dz; L mark = c->dcur;
        if((e = out_init(c, z, HPA_L, ny, &dz))) R e;
        int64_t *zz = (int64_t*)dz;
        switch(x->typ){
            case HPA_C: e = index_of_table(c, (const uint8_t*)px, nx, (const uint8_t*)py, ny, zz, 256, false); break;
            case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T:
                e = index_of_hash<int32_t>(c, (const int32_t*)px, nx, (const int32_t*)py, ny, zz); break;
            case HPA_Z: case HPA_Q:
                e = index_of_hash<int64_t>(c, (const int64_t*)px, nx, (const int64_t*)py, ny, zz); break;
            default: e = HPA_E_NOT_IMPL;
        }
        if(e) R e;
        if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
        R e;
    }
    R HPA_E_TYPE;
}

template<typename T>
static int32_t member_t(hpa_ctx *c, const T *py, L ny, const T *px, L nx, uint8_t *z){
    ProfScope ps("member", c->stream, (double)sizeof(T) * (nx + ny) + (double)ny);
    if(nx <= 2){
        if(ny > 0) hipLaunchKernelGGL(k_member_small<T>, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, py, ny, px, (int)nx, z);
        R check_launch("member");
    }
    int32_t e;
    if(std::is_same<T, int32_t>::value){
        int lo, hi; bool dense;
        if((e = dense_range(c, (const int32_t*)px, nx, &lo, &hi, &dense))) R e;
        if(dense){
            DenseTab t;
            if((e = dense_build(c, (const int32_t*)px, nx, lo, (L)hi - lo + 1, false, &t))) R e;
            if(ny > 0) hipLaunchKernelGGL(k_dense_probe_member, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)py, ny, lo, t.m, t.bits, z);
            R check_launch("member");
        }
    }
    if(std::is_same<T, int32_t>::value && nx < (L)UINT32_MAX - 1){
        unsigned long long *slots; uint64_t mask; int *dup;
        if((e = pk_build(c, (const int32_t*)px, nx, &slots, &mask, &dup))) R e;
        if(ny > 0) hipLaunchKernelGGL(k_pk_probe_member, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)py, ny, slots, mask, z);
        R check_launch("member");
    }
    HTab t;
    if((e = ht_create(c, nx, &t))) R e;
    hipLaunchKernelGGL(k_ht_insert_min<T>, dim3(grid_for(nx, 256, 4)), dim3(256), 0, c->stream, px, nx, t);
    if(ny > 0) hipLaunchKernelGGL(k_ht_probe_member<T>, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, py, ny, t, z);
    R check_launch("member");
}
__global__ void k_member_table(const uint8_t *y, L n, const uint8_t *x, L nx, uint8_t *z, int nt){
    __shared__ uint8_t tab[256];
    for(int i = threadIdx.x; i < 256; i += blockDim.x) tab[i] = 0;
    __syncthreads();
    for(L i = threadIdx.x; i < nx; i += blockDim.x){ unsigned v = x[i]; if(nt == 2) v = v != 0; tab[v] = 1; }
    __syncthreads();
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){ unsigned v = y[i]; if(nt == 2) v = v != 0; z[i] = tab[v]; }
}

// pfnMember(z, y, x): z[i] = y[i] in x, length of y (h_primitive.cc:2288)
extern "C" int32_t pfnMember(hpa_V z, hpa_V y, hpa_V x){
    hpa_ctx *c = ctx_cur();
    int32_t e;
    L nx = x->len, ny = y->len;
    const void *px = nullptr, *py = nullptr;
    if(is_real(x->typ) && is_real(y->typ)){
        int32_t t = x->typ > y->typ ? x->typ : y->typ;
        if(t == HPA_J) R HPA_E_NOT_IMPL;                 // no MEMBER(J) case in the reference
        if(nx > 0 && (e = promote_dev(c, x, t, &px))) R e;
        if(ny > 0 && (e = promote_dev(c, y, t, &py))) R e;
        void *dz; L mark = c->dcur;
        if((e = out_init(c, z, HPA_B, ny, &dz))) R e;
        uint8_t *zz = (uint8_t*)dz;
        switch(t){
            case HPA_B:
                if(ny > 0) hipLaunchKernelGGL(k_member_table, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)py, ny, (const uint8_t*)px, nx, zz, 2);
                e = check_launch("member"); break;
            case HPA_H: e = member_t<int16_t>(c, (const int16_t*)py, ny, (const int16_t*)px, nx, zz); break;
            case HPA_I: e = member_t<int32_t>(c, (const int32_t*)py, ny, (const int32_t*)px, nx, zz); break;
            case HPA_L: e = member_t<int64_t>(c, (const int64_t*)py, ny, (const int64_t*)px, nx, zz); break;
            case HPA_F: e = member_t<float>(c, (const float*)py, ny, (const float*)px, nx, zz); break;
            case HPA_E: e = member_t<double>(c, (const double*)py, ny, (const double*)px, nx, zz); break;
        }
        if(e) R e;
        if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
        R e;
    }
    if(x->typ == HPA_S && y->typ == HPA_S) R member_str(c, z, y, x);   // by the strings' bytes
    if(x->typ == y->typ && is_basic(x->typ)){
        if(nx > 0 && (e = dev_view(c, x, &px))) R e;
        if(ny > 0 && (e = dev_view(c, y, &py))) R e;
        void *dz; L mark = c->dcur;
        if((e = out_init(c, z, HPA_B, ny, &dz))) R e;
        uint8_t *zz = (uint8_t*)dz;
        switch(x->typ){
            case HPA_C:
                if(ny > 0) hipLaunchKernelGGL(k_member_table, dim3(grid_for(ny, 256, 4)), dim3(256), 0, c->stream, (const uint8_t*)py, ny, (const uint8_t*)px, nx, zz, 256);
                e = check_launch("member"); break;
            case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T:
                e = member_t<int32_t>(c, (const int32_t*)py, ny, (const int32_t*)px, nx, zz); break;
            case HPA_Z: case HPA_Q:
                e = member_t<int64_t>(c, (const int64_t*)py, ny, (const int64_t*)px, nx, zz); break;
            default: e = HPA_E_NOT_IMPL;
        }
        if(e) R e;
        if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
        R e;
    }
    R HPA_E_DOMAIN;
}

// ---------------------------------------------------------------- join
__global__ void k_join_build_heads(const uint32_t *sk, L n, const int64_t *pos, const uint32_t *flag, HTab t, int64_t *start){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        if(flag[i]){
            int64_t r = pos[i]; start[r] = i;
            unsigned long long *s = ht_slot(t, (uint64_t)sk[i], true);
            *s = (unsigned long long)r;
        }
    }
}
__global__ void k_join_flags(const uint32_t *sk, L n, uint32_t *flag){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1u : 0u;
}
__global__ void k_join_probe(const int32_t *y, L ny, HTab t, const int64_t *start, L g, L nb, uint32_t *cnt, uint32_t *rid){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        const unsigned long long *s = ht_find(t, (uint64_t)(uint32_t)y[j]);
        if(s){ L r = (L)*s; rid[j] = (uint32_t)r; cnt[j] = (uint32_t)((r + 1 < g ? start[r + 1] : nb) - start[r]); }
        else { rid[j] = 0; cnt[j] = 0; }
    }
}
__global__ void k_join_write(L ny, const uint32_t *cnt, const uint32_t *rid, const int64_t *off, const int64_t *start,
                             const uint32_t *sidx, int64_t *z0, int64_t *z1, int swap){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        uint32_t c0 = cnt[j]; if(!c0) continue;
        int64_t o = off[j]; L s = start[rid[j]];
        for(uint32_t q = 0; q < c0; q++){
            int64_t b = sidx[s + q];
            if(swap){ z0[o + q] = j; z1[o + q] = b; } else { z0[o + q] = b; z1[o + q] = j; }
        }
    }
}
// lib_join_index_linear (h_libs.cc:1542-1575): the first matching build row only
template<typename T>
__global__ void k_join_linear(const T *x, int nx, const T *y, L ny, int eq, uint32_t *cnt, uint32_t *rid){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        T v = y[j]; int hit = -1;
        for(int i = 0; i < nx; i++){ bool m = eq ? (v == x[i]) : (v != x[i]); if(m){ hit = i; break; } }
        cnt[j] = hit >= 0 ? 1u : 0u; rid[j] = hit >= 0 ? (uint32_t)hit : 0u;
    }
}

static int32_t set_l_vec(hpa_ctx *c, V z, int64_t *p, L n){
    z->typ = HPA_L; z->len = n;
    if(n <= 1){ z->g = (char*)&z->b; z->i64 = 0; if(n == 1) R d2h(c, &z->b, p, 8); R 0; }
    z->g = (char*)p; R 0;
}

// build on bx (shorter), probe py; `swap` writes (probe, build) into (z0, z1)
// hash_only: the hash path whatever the build size (joinOneColumn calls lib_join_index_hash
// directly, h_system.cc:860-868, with no linear path below NUM_JOIN_LINEAR rows)
static int32_t join_run(hpa_ctx *c, V z, V bx, V py, int32_t t, int op, bool swap, bool hash_only = false){
    L nb = bx->len, np = py->len; int32_t e;
    const void *pb = nullptr, *pp = nullptr;
    if(nb > 0 && (e = promote_dev(c, bx, t, &pb))) R e;
    if(np > 0 && (e = promote_dev(c, py, t, &pp))) R e;
    uint32_t *cnt = (uint32_t*)dalloc(c, 4 * (np > 0 ? np : 1)), *rid = (uint32_t*)dalloc(c, 4 * (np > 0 ? np : 1));
    int64_t *off = (int64_t*)dalloc(c, 8 * (np + 1));
    if(!cnt || !rid || !off) R HPA_E_MEMORY;
    int64_t *start = nullptr; uint32_t *sidx = nullptr;
    ProfScope ps("join", c->stream);
    if(nb >= 10 || hash_only){                       // NUM_JOIN_LINEAR (common.h:152): hash path
        if(op != 4 || t != HPA_I) R HPA_E_NOT_IMPL;  // v3 radix hash join: i32 equality only (h_hash_v3.cc:233,270-277)
        // unique build keys: a direct first-row table (cleared, or guarded by a presence bitmap)
        // or a packed hash table, no sort
        L mark = c->dcur;
        int lo, hi; bool dense, guard;
        if((e = dense_range(c, (const int32_t*)pb, nb, &lo, &hi, &dense, nullptr, &guard))) R e;
        unsigned long long *slots = nullptr; uint64_t mask = 0; DenseTab dt; int *dup = nullptr;
        bool room = false;
        if(guard && (e = guard_build(c, (const int32_t*)pb, nb, lo, (L)hi - lo + 1, &dt, &room))) R e;
        if(room) dense = true;                       // probed like a dense table
        else if(dense){ if((e = dense_build(c, (const int32_t*)pb, nb, lo, (L)hi - lo + 1, true, &dt))) R e; }
        else if((e = pk_build(c, (const int32_t*)pb, nb, &slots, &mask, &dup))) R e;
        if(dense) dup = dt.dup;
        // Probe, scan and (when the arena has room for np pairs, the most unique build keys can
        // give) the pair writes are queued before the host reads anything: one read then returns
        // the repeated-key flag and the pair count (dup[0], dup[2..3]), and the pair buffers are
        // trimmed.  Repeated keys discard the lot for the sort path below.
        // A guarded table probes each 4096-row tile twice (count, then write at the tile's
        // offset); full direct and hash tables probe once into per-row counts and build rows.
        int64_t *tot = (int64_t*)(dup + 2);          // zero until a scan writes it
        const L ntl = ceil_div(np, DJ_TILE);
        uint32_t *tcount = nullptr; int64_t *toff = nullptr;
        const bool tiled = room;
        if(tiled){
            tcount = (uint32_t*)dalloc(c, 4 * (ntl > 0 ? ntl : 1)); toff = (int64_t*)dalloc(c, 8 * (ntl + 1));
            if(!tcount || !toff) R HPA_E_MEMORY;
            if(np > 0){
                hipLaunchKernelGGL(k_dj_count, dim3((unsigned)ntl), dim3(256), 0, c->stream, (const int32_t*)pp, np, dt, tcount);
                if((e = scan_u32_to_i64(c, tcount, toff, ntl, tot))) R e;
            }
        } else {
            if(np > 0){
                if(dense) hipLaunchKernelGGL(k_dense_probe_join, dim3(grid_for(np, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)pp, np, dt, cnt, rid);
                else hipLaunchKernelGGL(k_pk_probe, dim3(grid_for(np, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)pp, np, slots, mask, cnt, rid);
            }
            if((e = scan_u32_to_i64(c, cnt, off, np, tot))) R e;
        }
        auto write_pairs = [&](int64_t *z0, int64_t *z1){
            if(np <= 0) return;
            if(tiled) hipLaunchKernelGGL(k_dj_write, dim3((unsigned)ntl), dim3(256), 0, c->stream, (const int32_t*)pp, np, dt,
                                         (const int64_t*)toff, z0, z1, swap ? 1 : 0);
            else hipLaunchKernelGGL(k_pk_write, dim3(grid_for(np, 256, 4)), dim3(256), 0, c->stream, np, cnt, rid, off, z0, z1, swap ? 1 : 0);
        };
        const bool spec = c->dcur + 16 * (np > 0 ? np : 1) + 1024 <= c->dtop;
        int64_t *z0 = nullptr, *z1 = nullptr;
        if(spec){
            z0 = (int64_t*)dalloc(c, 8 * (np > 0 ? np : 1)); z1 = (int64_t*)dalloc(c, 8 * (np > 0 ? np : 1));
            if(!z0 || !z1) R HPA_E_MEMORY;
            write_pairs(z0, z1);
        }
        int64_t st[2] = {0, 0};
        if((e = d2h(c, st, dup, 16))) R e;
        const int hd = (int)(st[0] & 0xffffffff);
        const int64_t P = st[1];
        if(!hd){
            if(spec) arena_set(c, (L)((char*)z1 - c->dbase) + ((8 * (P > 0 ? P : 1) + 255) & ~(L)255));
            else {
                z0 = (int64_t*)dalloc(c, 8 * (P > 0 ? P : 1)); z1 = (int64_t*)dalloc(c, 8 * (P > 0 ? P : 1));
                if(!z0 || !z1) R HPA_E_MEMORY;
                write_pairs(z0, z1);
            }
            if((e = check_launch("join"))) R e;
            V cells = (V)halloc(c, 2 * sizeof(V0)); if(!cells) R HPA_E_MEMORY;
            z->typ = HPA_G; z->len = 2; z->g = (char*)cells; z->g2 = nullptr;
            if((e = set_l_vec(c, cells, z0, P))) R e;
            if((e = set_l_vec(c, cells + 1, z1, P))) R e;
            note_bound(c, z0, swap ? np : nb); note_bound(c, z1, swap ? nb : np);
            ps.stop(); ps.bytes = 4.0 * (nb + np) + 16.0 * P;          // keys read once, pairs written
            R 0;
        }
        arena_set(c, mark);                              // duplicates: the sort-based build below
        uint32_t *k = (uint32_t*)dalloc(c, 4 * nb); if(!k) R HPA_E_MEMORY;
        hipMemcpyAsync(k, pb, 4 * nb, hipMemcpyDeviceToDevice, c->stream);
        uint32_t *sk;
        if((e = radix_sort_pairs_u32(c, k, nullptr, nb, 0, 32, &sk, &sidx))) R e;
        uint32_t *flag = (uint32_t*)dalloc(c, 4 * nb); int64_t *pos = (int64_t*)dalloc(c, 8 * (nb + 1));
        if(!flag || !pos) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_join_flags, dim3(grid_for(nb, 256, 4)), dim3(256), 0, c->stream, sk, nb, flag);
        if((e = scan_u32_to_i64(c, flag, pos, nb, pos + nb))) R e;
        int64_t g = 0; if((e = d2h(c, &g, pos + nb, 8))) R e;
        HTab tb; if((e = ht_create(c, g, &tb))) R e;
        start = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1)); if(!start) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_join_build_heads, dim3(grid_for(nb, 256, 4)), dim3(256), 0, c->stream, sk, nb, pos, flag, tb, start);
        if(np > 0) hipLaunchKernelGGL(k_join_probe, dim3(grid_for(np, 256, 4)), dim3(256), 0, c->stream, (const int32_t*)pp, np, tb, start, g, nb, cnt, rid);
    } else {
        if(!(t == HPA_I || t == HPA_L || t == HPA_E)) R HPA_E_NOT_IMPL;
        // the linear path writes the build row itself: start = identity, one match per probe row
        start = (int64_t*)dalloc(c, 8 * 16); sidx = (uint32_t*)dalloc(c, 4 * 16);
        if(!start || !sidx) R HPA_E_MEMORY;
        int64_t hs[16]; uint32_t hi[16]; for(int i = 0; i < 16; i++){ hs[i] = i; hi[i] = (uint32_t)i; }
        hipMemcpyAsync(start, hs, sizeof(hs), hipMemcpyHostToDevice, c->stream);
        hipMemcpyAsync(sidx, hi, sizeof(hi), hipMemcpyHostToDevice, c->stream);
        if(np > 0){
            dim3 gg(grid_for(np, 256, 4)), bb(256);
            int eq = op == 4;
            if(t == HPA_I) hipLaunchKernelGGL(k_join_linear<int32_t>, gg, bb, 0, c->stream, (const int32_t*)pb, (int)nb, (const int32_t*)pp, np, eq, cnt, rid);
            else if(t == HPA_L) hipLaunchKernelGGL(k_join_linear<int64_t>, gg, bb, 0, c->stream, (const int64_t*)pb, (int)nb, (const int64_t*)pp, np, eq, cnt, rid);
            else hipLaunchKernelGGL(k_join_linear<double>, gg, bb, 0, c->stream, (const double*)pb, (int)nb, (const double*)pp, np, eq, cnt, rid);
        }
        hipStreamSynchronize(c->stream);
    }
    if((e = scan_u32_to_i64(c, cnt, off, np, off + np))) R e;
    int64_t P = 0; if((e = d2h(c, &P, off + np, 8))) R e;
    int64_t *z0 = (int64_t*)dalloc(c, 8 * (P > 0 ? P : 1)), *z1 = (int64_t*)dalloc(c, 8 * (P > 0 ? P : 1));
    if(!z0 || !z1) R HPA_E_MEMORY;
    if(np > 0) hipLaunchKernelGGL(k_join_write, dim3(grid_for(np, 256, 4)), dim3(256), 0, c->stream, np, cnt, rid, off, start, sidx, z0, z1, swap ? 1 : 0);
    if((e = check_launch("join"))) R e;
    V cells = (V)halloc(c, 2 * sizeof(V0)); if(!cells) R HPA_E_MEMORY;
    z->typ = HPA_G; z->len = 2; z->g = (char*)cells; z->g2 = nullptr;
    if((e = set_l_vec(c, cells, z0, P))) R e;
    if((e = set_l_vec(c, cells + 1, z1, P))) R e;
    note_bound(c, z0, swap ? np : nb); note_bound(c, z1, swap ? nb : np);   // rows of x and of y
    ps.stop(); ps.bytes = 4.0 * (nb + np) + 16.0 * P;
    R 0;
}

// ---------------------------------------------------------------- multi-column join
// pfnJoinIndexMultiple (h_primitive.cc:2095-2110): the first "eq" column (else the first "neq",
// else column 0; getFirstEqual, :2088-2092) is joined alone by joinOneColumn -- the radix hash
// join built on x, probed with y, no swap of the sides (h_system.cc:860-881) -- and then
// joinOtherColumns keeps, column by column and in order, the pairs whose values satisfy that
// column's operator (compareOneColumn, h_system.cc:819-843: exact comparisons COMP, :791-794, in
// the column pair's larger type).  Keeping the pairs that pass every column at once is the same
// filter.
#define JOIN_MAX_COLS 16
struct JoinPred {
    int n;
    int op[JOIN_MAX_COLS], cls[JOIN_MAX_COLS], tx[JOIN_MAX_COLS], ty[JOIN_MAX_COLS];
    const void *px[JOIN_MAX_COLS], *py[JOIN_MAX_COLS];
};
__device__ __forceinline__ int64_t jp_int(int t, const void *p, int64_t i){
    switch(t){
        case HPA_B: return ((const uint8_t*)p)[i];
        case HPA_J: return ((const int8_t*)p)[i];
        case HPA_H: return ((const int16_t*)p)[i];
        case HPA_I: return ((const int32_t*)p)[i];
        default:    return ((const int64_t*)p)[i];
    }
}
template<typename R_> __device__ __forceinline__ R_ jp_real(int t, const void *p, int64_t i){
    switch(t){
        case HPA_F: return (R_)((const float*)p)[i];
        case HPA_E: return (R_)((const double*)p)[i];
        default:    return (R_)jp_int(t, p, i);
    }
}
template<typename T> __device__ __forceinline__ bool jp_comp(int op, T x, T y){   // COMP, h_system.cc:791-794
    return op < 2 ? (op == 0 ? x < y : x <= y) : op < 4 ? (op == 2 ? x > y : x >= y) : op < 6 ? (op == 4 ? x == y : x != y) : false;
}
__global__ void __launch_bounds__(256) k_join_filter(JoinPred p, const int64_t *z0, const int64_t *z1, L np, uint8_t *keep){
    for(L k = (L)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (L)gridDim.x * blockDim.x){
        const int64_t a = z0[k], b = z1[k];
        bool ok = true;
        for(int q = 0; q < p.n && ok; q++){
            switch(p.cls[q]){
                case 0: ok = jp_comp(p.op[q], jp_int(p.tx[q], p.px[q], a), jp_int(p.ty[q], p.py[q], b)); break;
                case 1: ok = jp_comp(p.op[q], jp_real<float>(p.tx[q], p.px[q], a), jp_real<float>(p.ty[q], p.py[q], b)); break;
                default: ok = jp_comp(p.op[q], jp_real<double>(p.tx[q], p.px[q], a), jp_real<double>(p.ty[q], p.py[q], b)); break;
            }
        }
        keep[k] = ok ? 1 : 0;
    }
}
static int op_from_name(const char *s);
static int32_t join_multi(hpa_ctx *c, V z, V x, V y, V f){
    int32_t e;
    const L nc = f->len;
    if(x->typ != HPA_G || y->typ != HPA_G || x->len != nc || y->len != nc) R HPA_E_LENGTH;
    if(nc > JOIN_MAX_COLS){ set_error_detail("join: more than %d key columns", JOIN_MAX_COLS); R HPA_E_NOT_IMPL; }
    std::vector<int64_t> q((size_t)nc);
    if((e = hpaReadPayload(f, q.data(), 8 * nc))) R e;
    std::vector<int> ops((size_t)nc);
    for(L i = 0; i < nc; i++){ ops[i] = op_from_name(hpaSymbolName(q[i])); if(ops[i] < 0) R HPA_E_DOMAIN; }
    L fx = -1;
    for(L i = 0; i < nc && fx < 0; i++) if(ops[i] == 4) fx = i;
    for(L i = 0; i < nc && fx < 0; i++) if(ops[i] == 5) fx = i;
    if(fx < 0) fx = 0;
    V xc = (V)x->g, yc = (V)y->g;
    const L nx = xc[0].len, ny = yc[0].len;
    for(L i = 0; i < nc; i++){
        if(!is_real(xc[i].typ) || !is_real(yc[i].typ)) R HPA_E_NOT_IMPL;   // compareOpWithIndex: B/H/J/I/L/F/E
        if(xc[i].len != nx || yc[i].len != ny) R HPA_E_LENGTH;
    }
    const int32_t t0 = xc[fx].typ > yc[fx].typ ? xc[fx].typ : yc[fx].typ;
    // joinOneColumn: eq takes the hash join (i32 keys only in v3, h_hash_v3.cc:233-259); its neq
    // and ordered operators leave this path
    if(ops[fx] != 4 || t0 != HPA_I){ set_error_detail("join: the first key column must be an i32 eq"); R HPA_E_NOT_IMPL; }
    V0 one; memset(&one, 0, sizeof(one));
    if((e = join_run(c, &one, xc + fx, yc + fx, t0, 4, false, true))) R e;
    V pc = (V)one.g;                                   // list(x rows, y rows)
    const L P = pc[0].len;
    JoinPred jp; memset(&jp, 0, sizeof(jp));
    for(L i = 0; i < nc; i++){
        if(i == fx) continue;
        const int k = jp.n++;
        const int32_t tm = xc[i].typ > yc[i].typ ? xc[i].typ : yc[i].typ;
        jp.op[k] = ops[i]; jp.tx[k] = xc[i].typ; jp.ty[k] = yc[i].typ;
        jp.cls[k] = tm == HPA_F ? 1 : tm == HPA_E ? 2 : 0;
        if(P > 0){
            if(nx > 0 && (e = dev_view(c, xc + i, &jp.px[k]))) R e;
            if(ny > 0 && (e = dev_view(c, yc + i, &jp.py[k]))) R e;
        }
    }
    if(jp.n == 0 || P == 0){ *z = one; R 0; }
    const void *p0 = nullptr, *p1 = nullptr;
    if((e = dev_view(c, pc, &p0)) || (e = dev_view(c, pc + 1, &p1))) R e;
    V0 mask; memset(&mask, 0, sizeof(mask));
    void *dm; if((e = out_init(c, &mask, HPA_B, P, &dm))) R e;
    {
        ProfScope ps("join_filter", c->stream, 16.0 * P + (double)P);
        hipLaunchKernelGGL(k_join_filter, dim3(grid_for(P, 256, 4)), dim3(256), 0, c->stream, jp, (const int64_t*)p0,
                           (const int64_t*)p1, P, (uint8_t*)dm);
        if((e = check_launch("join_filter"))) R e;
        if(is_inline(&mask) && (e = commit_inline(c, &mask, dm))) R e;
    }
    V cells = (V)halloc(c, 2 * sizeof(V0)); if(!cells) R HPA_E_MEMORY;
    V zs[2] = {cells, cells + 1}, ys[2] = {pc, pc + 1};
    if((e = hpaCompressMany(2, zs, &mask, ys))) R e;
    z->typ = HPA_G; z->len = 2; z->g = (char*)cells; z->g2 = nullptr;
    R 0;
}

static int op_from_name(const char *s){            // backend/h_system.cc:779-789
    if(!s) R -1;
    if(!strcmp(s, "lt")) R 0; if(!strcmp(s, "leq")) R 1; if(!strcmp(s, "gt")) R 2;
    if(!strcmp(s, "geq")) R 3; if(!strcmp(s, "eq")) R 4; if(!strcmp(s, "neq")) R 5;
    R -1;
}

// pfnJoinIndex (h_primitive.cc:2122-2155): z = list(i64 x-rows, i64 y-rows)
extern "C" int32_t pfnJoinIndex(hpa_V z, hpa_V x, hpa_V y, hpa_V f){
    hpa_ctx *c = ctx_cur();
    if(f->typ != HPA_Q || f->len < 1) R HPA_E_DOMAIN;
    if(f->len > 1) R join_multi(c, z, x, y, f);
    int64_t q; int32_t e = read_scalar(c, f, &q, 8); if(e) R e;
    int op = op_from_name(hpaSymbolName(q));
    if(op < 0) R HPA_E_DOMAIN;
    if(!(is_real(x->typ) && is_real(y->typ))) R HPA_E_NOT_IMPL;
    if(op != 4 && op != 5) R HPA_E_NOT_IMPL;         // the reference's brute-force path is a debug stop
    int32_t t = x->typ > y->typ ? x->typ : y->typ;
    // build on the shorter side; after the swap z0 still indexes x (swap2, h_system.cc:47-56)
    if(x->len > y->len) R join_run(c, z, y, x, t, op, true);
    R join_run(c, z, x, y, t, op, false);
}
