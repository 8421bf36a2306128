// Seeded synthetic tables generated directly in HBM, for benchmarks and serving runs whose
// tables are too large to build on the host and upload (600M fact rows, 150M dimension rows).
//
// Every column is a counter-based stream, value(row) = f(splitmix64(seed, stream, row)), the
// same generator as horsepower_amd/datagen.py (gen_serve there is the host twin of
// hpaGenServeTables; tests check equality).  The tables have the shape of the three benchmark
// queries over one fact table: orders(key, order_day, segment) with unique sparse keys, and
// lineitem(fk, ship_day, price, discount, quantity, tax, return_flag, line_status) with 1..7 rows
// per order (tests/hyper/online-hir/q1.hir, q3.hir, q6.hir read these columns).
#include "hpa_internal.h"
#include <numeric>

#define SM_G  0x9E3779B97F4A7C15ull
#define SM_C1 0xBF58476D1CE4E5B9ull
#define SM_C2 0x94D049BB133111EBull

static uint64_t mix_seed(uint64_t seed, uint64_t stream){
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * SM_C1;
    z = (z ^ (z >> 27)) * SM_C2;
    return z ^ (z >> 31);
}

// splitmix64 output for counter i (1-based, as datagen.rand_u64)
__device__ __forceinline__ uint64_t sm_at(uint64_t s0, uint64_t i){
    uint64_t z = s0 + i * SM_G;
    z = (z ^ (z >> 30)) * SM_C1;
    z = (z ^ (z >> 27)) * SM_C2;
    return z ^ (z >> 31);
}
__device__ __forceinline__ int64_t sm_uniform(uint64_t s0, uint64_t i, int64_t lo, uint64_t span){
    return lo + (int64_t)(sm_at(s0, i) % span);
}

template<typename T>
__global__ void k_gen_uniform(T *z, L n, L start, uint64_t s0, int64_t lo, uint64_t span){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        z[i] = (T)sm_uniform(s0, (uint64_t)(start + i + 1), lo, span);
}
__global__ void k_gen_cents(double *z, L n, L start, uint64_t s0, int64_t lo, uint64_t span){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        z[i] = (double)sm_uniform(s0, (uint64_t)(start + i + 1), lo, span) / 100.0;
}

// orders: key = 4 * ((a*i + b) mod n) + 1 + U{0..3}; day U{0..2405}; segment U{0..4};
// per = U{1..7} lineitem rows
__global__ void k_gen_dim(L n, L a, L b, uint64_t s11, uint64_t s12, uint64_t s13, uint64_t s14,
                          int32_t *key, int32_t *day, int32_t *seg, int64_t *per){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        uint64_t c = (uint64_t)i + 1;
        L perm = (L)(((uint64_t)a * (uint64_t)i + (uint64_t)b) % (uint64_t)n);      // a, i < 2^30
        key[i] = (int32_t)(4 * perm + 1 + sm_uniform(s11, c, 0, 4));
        day[i] = (int32_t)sm_uniform(s12, c, 0, 2406);
        seg[i] = (int32_t)sm_uniform(s13, c, 0, 5);
        per[i] = sm_uniform(s14, c, 1, 7);
    }
}
__global__ void k_gen_owner(L n, const int64_t *per, const int64_t *off, int64_t *owner){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x)
        for(int64_t k = 0; k < per[i]; k++) owner[off[i] + k] = i;
}
// days from an order to its lineitems' shipping (datagen.Q3_SHIP_SPAN: about 30K result groups
// of the join + group query per 1.5M orders)
#define GEN_SHIP_SPAN 318
struct FactOut { int32_t *fk, *ship, *qty; double *price, *disc, *tax; char *flag, *status; };
struct FactSeeds { uint64_t s15, s16, s17, s18, s19, s20, s21; };
__global__ void k_gen_fact(L nf, const int64_t *owner, const int32_t *key, const int32_t *day, FactSeeds s, FactOut o){
    for(L r = (L)blockIdx.x * blockDim.x + threadIdx.x; r < nf; r += (L)gridDim.x * blockDim.x){
        const uint64_t c = (uint64_t)r + 1;
        const int64_t w = owner[r];
        const int32_t ship = day[w] + (int32_t)sm_uniform(s.s15, c, 1, GEN_SHIP_SPAN);
        o.fk[r] = key[w];
        o.ship[r] = ship;
        o.price[r] = (double)sm_uniform(s.s16, c, 90000, 10500000 - 90000 + 1) / 100.0;
        o.disc[r] = (double)sm_uniform(s.s17, c, 0, 11) / 100.0;
        o.qty[r] = (int32_t)sm_uniform(s.s18, c, 1, 50);
        o.tax[r] = (double)sm_uniform(s.s19, c, 0, 9) / 100.0;
        const int32_t receipt = ship + (int32_t)sm_uniform(s.s20, c, 1, 30);
        const bool r_flag = sm_uniform(s.s21, c, 0, 2) == 1;
        o.flag[r] = receipt > 1263 ? 'N' : (r_flag ? 'R' : 'A');
        o.status[r] = ship > 1263 ? 'O' : 'F';
    }
}

// One column of n values: typ I or L, uniform integers in [lo, hi]; typ E, k/100.0 for k
// uniform in [lo, hi] (datagen.uniform / datagen.cents).  Counters start at start + 1.
extern "C" int32_t hpaGenUniform(hpa_V z, int32_t typ, int64_t n, int64_t start, uint64_t seed, uint64_t stream,
                                 int64_t lo, int64_t hi){
    hpa_ctx *c = ctx_cur();
    if(n < 0 || hi < lo || !(typ == HPA_I || typ == HPA_L || typ == HPA_E)) R HPA_E_DOMAIN;
    int32_t e;
    if((e = resident_alloc(z, typ, n))) R e;
    const uint64_t s0 = mix_seed(seed, stream), span = (uint64_t)(hi - lo + 1);
    const int G = grid_for(n, 256, 4);
    if(n > 0){
        if(typ == HPA_I) hipLaunchKernelGGL(k_gen_uniform<int32_t>, dim3(G), dim3(256), 0, c->stream, (int32_t*)z->g, n, start, s0, lo, span);
        else if(typ == HPA_L) hipLaunchKernelGGL(k_gen_uniform<int64_t>, dim3(G), dim3(256), 0, c->stream, (int64_t*)z->g, n, start, s0, lo, span);
        else hipLaunchKernelGGL(k_gen_cents, dim3(G), dim3(256), 0, c->stream, (double*)z->g, n, start, s0, lo, span);
    }
    if((e = check_launch("gen_uniform"))) R e;
    R hipStreamSynchronize(c->stream) == hipSuccess ? 0 : HPA_E_DEVICE;
}

// The serving tables: dim[0..2] = orders(key I, order_day I, segment I) with n_dim rows;
// fact[0..7] = lineitem(fk I, ship_day I, price E, discount E, quantity I, tax E,
// return_flag C, line_status C), 1..7 rows per order.  All columns are resident payloads.
extern "C" int32_t hpaGenServeTables(int64_t n_dim, uint64_t seed, hpa_V *dim, hpa_V *fact){
    hpa_ctx *c = ctx_cur();
    if(n_dim <= 0 || n_dim > (1ll << 30)) R HPA_E_DOMAIN;
    int32_t e;
    for(int k = 0; k < 3; k++) if((e = resident_alloc(dim[k], HPA_I, n_dim))) R e;
    L mark = c->dcur;
    int64_t *per = (int64_t*)dalloc(c, 8 * n_dim), *off = (int64_t*)dalloc(c, 8 * n_dim), *tot = (int64_t*)dalloc(c, 8);
    if(!per || !off || !tot){ arena_set(c, mark); R HPA_E_MEMORY; }
    // affine permutation of [0, n): a odd and coprime with n (datagen._affine_perm)
    L a = (L)(mix_seed(seed, 99) % (uint64_t)n_dim) | 1;
    while(std::gcd(a, (L)n_dim) != 1) a += 2;
    L b = (L)(mix_seed(seed, 98) % (uint64_t)n_dim);
    const int G = grid_for(n_dim, 256, 4);
    hipLaunchKernelGGL(k_gen_dim, dim3(G), dim3(256), 0, c->stream, (L)n_dim, a, b, mix_seed(seed, 11), mix_seed(seed, 12),
                       mix_seed(seed, 13), mix_seed(seed, 14), (int32_t*)dim[0]->g, (int32_t*)dim[1]->g, (int32_t*)dim[2]->g, per);
    if((e = check_launch("gen_dim")) || (e = scan_i64(c, per, off, n_dim, tot))) { arena_set(c, mark); R e; }
    int64_t nf = 0;
    if((e = d2h(c, &nf, tot, 8))){ arena_set(c, mark); R e; }
    int64_t *owner = (int64_t*)dalloc(c, 8 * nf);
    if(!owner){ arena_set(c, mark); R HPA_E_MEMORY; }
    hipLaunchKernelGGL(k_gen_owner, dim3(G), dim3(256), 0, c->stream, (L)n_dim, per, off, owner);
    static const int32_t ftyp[8] = {HPA_I, HPA_I, HPA_E, HPA_E, HPA_I, HPA_E, HPA_C, HPA_C};
    for(int k = 0; k < 8; k++) if((e = resident_alloc(fact[k], ftyp[k], nf))){ arena_set(c, mark); R e; }
    FactOut o = {(int32_t*)fact[0]->g, (int32_t*)fact[1]->g, (int32_t*)fact[4]->g, (double*)fact[2]->g,
                 (double*)fact[3]->g, (double*)fact[5]->g, fact[6]->g, fact[7]->g};
    FactSeeds s = {mix_seed(seed, 15), mix_seed(seed, 16), mix_seed(seed, 17), mix_seed(seed, 18),
                   mix_seed(seed, 19), mix_seed(seed, 20), mix_seed(seed, 21)};
    hipLaunchKernelGGL(k_gen_fact, dim3(grid_for(nf, 256, 4)), dim3(256), 0, c->stream, (L)nf, owner, (const int32_t*)dim[0]->g,
                       (const int32_t*)dim[1]->g, s, o);
    e = check_launch("gen_fact");
    if(!e && hipStreamSynchronize(c->stream) != hipSuccess) e = HPA_E_DEVICE;
    arena_set(c, mark);
    R e;
}
