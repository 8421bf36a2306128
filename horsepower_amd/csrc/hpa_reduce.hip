// Reductions: sum/avg/min/max over vectors (h_primitive.cc:764-848), grouped reductions over
// group index lists (the FP1/2/5/6/7 templates, optimizer/pattern.cc:180-219, and the
// interpreter's each_right(@index) + each(@sum/@avg/@len) + raze, h_primitive.cc:1818-1904),
// and the FA/FE fused loops (optimizer/autofusion.cc:770-935, optimizer/elementwise.cc:239-281).
//
// FP64 sums are compensated (Neumaier per thread, TwoSum when partials merge) and merged in a
// fixed order, so a result depends only on the input and its length: it is the same alone and
// inside a batch, and within a few ulps of the exact sum.
#include "hpa_internal.h"
#include "hpa_expr.h"
#include <float.h>
#include <limits.h>

struct KS { double s, c; };   // Neumaier accumulator: value s + c

__device__ __forceinline__ void ks_add(KS &a, double v){
    double t = a.s + v;
    a.c += (fabs(a.s) >= fabs(v)) ? (a.s - t) + v : (v - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ void ks_merge(KS &a, const KS &b){
    double t = a.s + b.s;
    double bb = t - a.s;
    double err = (a.s - (t - bb)) + (b.s - bb);      // TwoSum
    a.s = t; a.c += b.c + err;
}
__device__ __forceinline__ KS ks_shfl_down(const KS &a, int o){
    KS r; r.s = __shfl_down(a.s, o, 64); r.c = __shfl_down(a.c, o, 64); return r;
}
__device__ __forceinline__ KS ks_wave(KS a){
    for(int o = 32; o > 0; o >>= 1){ KS b = ks_shfl_down(a, o); ks_merge(a, b); }
    return a;   // lane 0 holds the wave result
}
__device__ __forceinline__ KS ks_block(KS a, KS *sh){
    a = ks_wave(a);
    int w = threadIdx.x >> 6;
    if((threadIdx.x & 63) == 0) sh[w] = a;
    __syncthreads();
    KS r = {0, 0};
    if(threadIdx.x == 0){ int nw = blockDim.x >> 6; for(int q = 0; q < nw; q++) ks_merge(r, sh[q]); }
    __syncthreads();
    return r;
}

#define RD_BLOCK 256

// ---------------------------------------------------------------- vector reductions
template<typename T>
__device__ __forceinline__ double as_d(T v){ return (double)v; }

// kind: 0 sum, 2 min, 3 max;  F=float results accumulate in double then round (see note)
template<typename T> struct __attribute__((aligned(4 * sizeof(T)))) V4 { T v[4]; };

// Four consecutive elements per thread and iteration (one vector load), four independent
// Neumaier accumulators (no dependency between the adds of one iteration), merged in a fixed
// order: the result depends only on the input and the grid size.
template<typename T>
__global__ void __launch_bounds__(RD_BLOCK) k_sum_f(const T *x, L n, KS *part){
    __shared__ KS sh[RD_BLOCK / 64];
    KS a0 = {0, 0}, a1 = {0, 0}, a2 = {0, 0}, a3 = {0, 0};
    const L stride = (L)gridDim.x * RD_BLOCK * 4;
    L i = ((L)blockIdx.x * RD_BLOCK + threadIdx.x) * 4;
    for(; i + 4 <= n; i += stride){
        const V4<T> w = *(const V4<T>*)(x + i);
        ks_add(a0, as_d(w.v[0])); ks_add(a1, as_d(w.v[1])); ks_add(a2, as_d(w.v[2])); ks_add(a3, as_d(w.v[3]));
    }
    for(L j = i; j < n && j < i + 4; j++) ks_add(a0, as_d(x[j]));
    ks_merge(a0, a1); ks_merge(a2, a3); ks_merge(a0, a2);
    KS r = ks_block(a0, sh);
    if(threadIdx.x == 0) part[blockIdx.x] = r;
}
// Partials merged in a fixed order: thread t folds partials t, t + 1024, ... (coalesced loads,
// all in flight together), then a pairwise tree in LDS.
#define SUM_FINAL 1024
__global__ void __launch_bounds__(SUM_FINAL) k_sum_f_final(const KS *part, int np, double *out){
    __shared__ KS sh[SUM_FINAL];
    KS a = {0, 0};
    for(int q = threadIdx.x; q < np; q += SUM_FINAL) ks_merge(a, part[q]);
    sh[threadIdx.x] = a;
    __syncthreads();
    for(int s = SUM_FINAL / 2; s > 0; s >>= 1){
        if((int)threadIdx.x < s) ks_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if(threadIdx.x == 0) *out = sh[0].s + sh[0].c;
}
// Exact int64 sum (wrapping, as the reference's i64 accumulator): 4 elements per thread and
// iteration, one global atomic per workgroup.
template<typename T>
__global__ void __launch_bounds__(RD_BLOCK) k_sum_i(const T *x, L n, unsigned long long *acc){
    __shared__ unsigned long long sh[RD_BLOCK / 64];
    unsigned long long a = 0;
    const L stride = (L)gridDim.x * RD_BLOCK * 4;
    L i = ((L)blockIdx.x * RD_BLOCK + threadIdx.x) * 4;
    for(; i + 4 <= n; i += stride){
        const V4<T> w = *(const V4<T>*)(x + i);
        a += (unsigned long long)(int64_t)w.v[0] + (unsigned long long)(int64_t)w.v[1]
           + (unsigned long long)(int64_t)w.v[2] + (unsigned long long)(int64_t)w.v[3];
    }
    for(L j = i; j < n && j < i + 4; j++) a += (unsigned long long)(int64_t)x[j];
    for(int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if(threadIdx.x == 0){
        for(int w = 1; w < RD_BLOCK / 64; w++) a += sh[w];
        atomicAdd(acc, a);
    }
}
template<typename T, int MAXK>
__global__ void __launch_bounds__(RD_BLOCK) k_minmax(const T *x, L n, T *part){
    __shared__ T sh[RD_BLOCK / 64];
    T a = x[0];
    for(L i = (L)blockIdx.x * RD_BLOCK + threadIdx.x; i < n; i += (L)gridDim.x * RD_BLOCK){
        T v = x[i]; a = MAXK ? (a > v ? a : v) : (a < v ? a : v);
    }
    for(int o = 32; o > 0; o >>= 1){ T v = __shfl_down(a, o, 64); a = MAXK ? (a > v ? a : v) : (a < v ? a : v); }
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if(threadIdx.x == 0){
        T r = sh[0];
        for(int q = 1; q < RD_BLOCK / 64; q++){ T v = sh[q]; r = MAXK ? (r > v ? r : v) : (r < v ? r : v); }
        part[blockIdx.x] = r;
    }
}
// min / max of the partials: one workgroup, each thread over a strided slice, then a tree (the
// result does not depend on the order)
template<typename T, int MAXK>
__global__ void __launch_bounds__(RD_BLOCK) k_minmax_final(const T *part, int np, T *out){
    __shared__ T sh[RD_BLOCK];
    T r = part[0];
    for(int q = threadIdx.x; q < np; q += RD_BLOCK){ const T v = part[q]; r = MAXK ? (r > v ? r : v) : (r < v ? r : v); }
    sh[threadIdx.x] = r;
    __syncthreads();
    for(int s = RD_BLOCK / 2; s > 0; s >>= 1){
        if((int)threadIdx.x < s){ const T v = sh[threadIdx.x + s], u = sh[threadIdx.x]; sh[threadIdx.x] = MAXK ? (u > v ? u : v) : (u < v ? u : v); }
        __syncthreads();
    }
    if(threadIdx.x == 0) *out = sh[0];
}

static int sum_grid(L n){
    L g = ceil_div(n, RD_BLOCK * 16), cap = (L)num_cus() * 8;
    if(g > cap) g = cap;
    if(g < 1) g = 1;
    R (int)g;
}

static int32_t sum_fp(hpa_ctx *c, V x, const void *p, double *res){
    L n = x->len; int G = sum_grid(n);
    KS *part = (KS*)dalloc(c, sizeof(KS) * G); double *out = (double*)dalloc(c, 16);
    if(!part || !out) R HPA_E_MEMORY;
    {
        ProfScope ps("sum", c->stream, (double)n * type_width(x->typ));
        if(((uintptr_t)p & 31) != 0){ set_error_detail("sum: payload not 32-byte aligned"); R HPA_E_DOMAIN; }
        if(x->typ == HPA_E) hipLaunchKernelGGL(k_sum_f<double>, dim3(G), dim3(RD_BLOCK), 0, c->stream, (const double*)p, n, part);
        else hipLaunchKernelGGL(k_sum_f<float>, dim3(G), dim3(RD_BLOCK), 0, c->stream, (const float*)p, n, part);
    }
    hipLaunchKernelGGL(k_sum_f_final, dim3(1), dim3(SUM_FINAL), 0, c->stream, part, G, out);
    int32_t e = check_launch("sum"); if(e) R e;
    R d2h(c, res, out, 8);
}
static int32_t sum_int(hpa_ctx *c, V x, const void *p, int64_t *res){
    L n = x->len; int G = sum_grid(n);
    unsigned long long *acc = (unsigned long long*)dalloc(c, 16); if(!acc) R HPA_E_MEMORY;
    hipMemsetAsync(acc, 0, 8, c->stream);
    {
        ProfScope ps("sum", c->stream, (double)n * type_width(x->typ));
        if(((uintptr_t)p & 31) != 0){ set_error_detail("sum: payload not 32-byte aligned"); R HPA_E_DOMAIN; }
        dim3 g(reduce_grid(n, RD_BLOCK)), b(RD_BLOCK);
        switch(x->typ){
            case HPA_B: hipLaunchKernelGGL(k_sum_i<uint8_t>, g, b, 0, c->stream, (const uint8_t*)p, n, acc); break;
            case HPA_J: hipLaunchKernelGGL(k_sum_i<int8_t>, g, b, 0, c->stream, (const int8_t*)p, n, acc); break;
            case HPA_H: hipLaunchKernelGGL(k_sum_i<int16_t>, g, b, 0, c->stream, (const int16_t*)p, n, acc); break;
            case HPA_I: hipLaunchKernelGGL(k_sum_i<int32_t>, g, b, 0, c->stream, (const int32_t*)p, n, acc); break;
            case HPA_L: hipLaunchKernelGGL(k_sum_i<int64_t>, g, b, 0, c->stream, (const int64_t*)p, n, acc); break;
        }
    }
    int32_t e = check_launch("sum"); if(e) R e;
    R d2h(c, res, acc, 8);
}

// pfnSum (h_primitive.cc:764-786): integers -> i64 exact, F -> F, E -> E
extern "C" int32_t pfnSum(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(!is_real(x->typ)) R HPA_E_DOMAIN;
    int32_t tz = x->typ == HPA_F ? HPA_F : x->typ == HPA_E ? HPA_E : HPA_L;
    L mark = c->dcur;
    const void *p = nullptr; int32_t e = 0;
    if(x->len > 0 && (e = dev_view(c, x, &p))) R e;
    int64_t iv = 0; double dv = 0;
    if(x->len > 0){
        if(tz == HPA_L) e = sum_int(c, x, p, &iv);
        else e = sum_fp(c, x, p, &dv);
    }
    arena_set(c, mark);
    if(e) R e;
    z->typ = tz; z->len = 1; z->g = (char*)&z->b; z->i64 = 0;
    if(tz == HPA_L) z->i64 = iv;
    else if(tz == HPA_E) z->f64 = dv;
    else z->f32 = (float)dv;
    R 0;
}

// ---------------------------------------------------------------- allocation-free FP64 sum
// hpaSumNoAlloc: pfnSum's kernels for an E vector, with the workgroup partials and the final
// value's staging in `scratch`, a file-scope device array instead of arena memory.  Host code
// never passes the array's name to a kernel (on the host it is only a stand-in for the symbol):
// noalloc_scratch_init resolves the array's address on each device with hipGetSymbolAddress, with
// that device current, and the kernels get that address.  The partials are written whole by every
// call before they are read, so nothing needs clearing; calls that overlap on one GPU race on it.
#define NOALLOC_PARTS 4096
static __device__ __align__(256) KS scratch[NOALLOC_PARTS];
#define NOALLOC_MAX_DEV 64
static KS *g_scratch_dev[NOALLOC_MAX_DEV];

int32_t noalloc_scratch_init(int device){
    if(device < 0 || device >= NOALLOC_MAX_DEV) R HPA_E_DEVICE;
    if(g_scratch_dev[device]) R 0;
    void *ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if(err != hipSuccess || !ptr){
        set_error_detail("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(err));
        R HPA_E_DEVICE;
    }
    g_scratch_dev[device] = (KS*)ptr;
    R 0;
}
// the value k_sum_f and k_sum_f_final give for zero or one element: (0 + v) + 0
__global__ void k_store_sum1(double *out, double v){ *out = (0.0 + v) + 0.0; }

extern "C" int32_t hpaSumNoAlloc(hpa_V x, double *out){
    if(!hpa_is_initialized()){ set_error_detail("hpaSumNoAlloc: call hpaInit first"); R HPA_E_DOMAIN; }
    if(!x || !out) R HPA_E_DOMAIN;
    if(x->typ != HPA_E){ set_error_detail("hpaSumNoAlloc: an E (f64) vector is required"); R HPA_E_TYPE; }
    hpa_ctx *c = ctx_cur();
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV || !g_scratch_dev[dev]){
        set_error_detail("hpaSumNoAlloc: no scratch resolved for the current device");
        R HPA_E_DEVICE;
    }
    const L n = x->len;
    if(n <= 0 || is_inline(x)){                           // no element, or one held in the node
        hipLaunchKernelGGL(k_store_sum1, dim3(1), dim3(1), 0, c->stream, out, n > 0 ? x->f64 : 0.0);
        R check_launch("sum_noalloc");
    }
    if(!is_device(x->g)){ set_error_detail("hpaSumNoAlloc: the vector's payload must be in HBM"); R HPA_E_DOMAIN; }
    if(((uintptr_t)x->g & 31) != 0){ set_error_detail("hpaSumNoAlloc: payload not 32-byte aligned"); R HPA_E_DOMAIN; }
    const int G = sum_grid(n);
    if(G > NOALLOC_PARTS){
        set_error_detail("hpaSumNoAlloc: %d workgroup partials exceed the %d the static scratch holds", G, NOALLOC_PARTS);
        R HPA_E_MEMORY;
    }
    KS *part = g_scratch_dev[dev];
    hipLaunchKernelGGL(k_sum_f<double>, dim3(G), dim3(RD_BLOCK), 0, c->stream, (const double*)x->g, n, part);
    hipLaunchKernelGGL(k_sum_f_final, dim3(1), dim3(SUM_FINAL), 0, c->stream, (const KS*)part, G, out);
    R check_launch("sum_noalloc");
}

// pfnAvg (h_primitive.cc:788-800)
extern "C" int32_t pfnAvg(hpa_V z, hpa_V x){
    V0 t; memset(&t, 0, sizeof(t));
    int32_t e = pfnSum(&t, x); if(e) R e;
    int32_t tz = t.typ == HPA_F ? HPA_F : HPA_E;
    z->typ = tz; z->len = 1; z->g = (char*)&z->b; z->i64 = 0;
    switch(t.typ){
        case HPA_L: z->f64 = t.i64 * 1.0 / x->len; break;
        case HPA_F: z->f32 = t.f32 / x->len; break;
        case HPA_E: z->f64 = t.f64 / x->len; break;
    }
    R 0;
}

template<typename T>
static int32_t minmax_typed(hpa_ctx *c, const void *p, L n, int op, T *res){
    int G = sum_grid(n);
    T *part = (T*)dalloc(c, sizeof(T) * G + 16); T *out = (T*)dalloc(c, 16);
    if(!part || !out) R HPA_E_MEMORY;
    {
        ProfScope ps("minmax", c->stream, (double)n * sizeof(T));
        if(op == 0) hipLaunchKernelGGL((k_minmax<T,0>), dim3(G), dim3(RD_BLOCK), 0, c->stream, (const T*)p, n, part);
        else hipLaunchKernelGGL((k_minmax<T,1>), dim3(G), dim3(RD_BLOCK), 0, c->stream, (const T*)p, n, part);
    }
    if(op == 0) hipLaunchKernelGGL((k_minmax_final<T,0>), dim3(1), dim3(RD_BLOCK), 0, c->stream, part, G, out);
    else hipLaunchKernelGGL((k_minmax_final<T,1>), dim3(1), dim3(RD_BLOCK), 0, c->stream, part, G, out);
    int32_t e = check_launch("minmax"); if(e) R e;
    R d2h(c, res, out, sizeof(T));
}

// pfnReduce (h_primitive.cc:808-840): empty input gives LLONG_MAX/LLONG_MIN, FLT_MAX/FLT_MIN,
// DBL_MAX/DBL_MIN (the reference uses FLT_MIN/DBL_MIN, the smallest positive normals, for max)
extern "C" int32_t pfnReduce(hpa_V z, hpa_V x, int32_t op){
    hpa_ctx *c = ctx_cur();
    if(!is_real(x->typ)) R HPA_E_DOMAIN;
    L n = x->len;
    if(n < 1){
        int32_t tz = is_int(x->typ) ? HPA_L : x->typ;
        z->typ = tz; z->len = 1; z->g = (char*)&z->b; z->i64 = 0;
        switch(tz){
            case HPA_L: z->i64 = op == 0 ? LLONG_MAX : LLONG_MIN; break;
            case HPA_F: z->f32 = op == 0 ? FLT_MAX : FLT_MIN; break;
            case HPA_E: z->f64 = op == 0 ? DBL_MAX : DBL_MIN; break;
        }
        R 0;
    }
    L mark = c->dcur;
    const void *p; int32_t e = dev_view(c, x, &p); if(e) R e;
    int64_t bits = 0;
    switch(x->typ){
        case HPA_B: e = minmax_typed<uint8_t>(c, p, n, op, (uint8_t*)&bits); break;
        case HPA_J: e = minmax_typed<int8_t>(c, p, n, op, (int8_t*)&bits); break;
        case HPA_H: e = minmax_typed<int16_t>(c, p, n, op, (int16_t*)&bits); break;
        case HPA_I: e = minmax_typed<int32_t>(c, p, n, op, (int32_t*)&bits); break;
        case HPA_L: e = minmax_typed<int64_t>(c, p, n, op, (int64_t*)&bits); break;
        case HPA_F: e = minmax_typed<float>(c, p, n, op, (float*)&bits); break;
        case HPA_E: e = minmax_typed<double>(c, p, n, op, (double*)&bits); break;
    }
    arena_set(c, mark);
    if(e) R e;
    z->typ = x->typ; z->len = 1; z->g = (char*)&z->b; z->i64 = bits;
    R 0;
}
extern "C" int32_t pfnMin(hpa_V z, hpa_V x){ R pfnReduce(z, x, 0); }
extern "C" int32_t pfnMax(hpa_V z, hpa_V x){ R pfnReduce(z, x, 1); }

// ---------------------------------------------------------------- segmented (grouped) reductions
// Segment g covers flat positions [off[g], off[g+1]); value(j) = x[idx ? idx[j] : j].
// Three regimes chosen from the average segment length: a thread per segment (sequential,
// ascending -- the FP1 template's order), a wavefront per segment, or several blocks per
// segment with partials merged in chunk order.
enum { GK_SUM = 0, GK_AVG = 1, GK_MIN = 2, GK_MAX = 3, GK_LEN = 4 };

// value j of a segment; indirect reads outside [0, nx) yield T() and raise *bad
template<typename T>
__device__ __forceinline__ T seg_val(const T *x, const int64_t *idx, L j, L nx, int *bad){
    if(!idx) return x[j];
    int64_t k = idx[j];
    if((uint64_t)k >= (uint64_t)nx){ if(bad) *bad = 1; return T(); }
    return x[k];
}

// floating segments -> KS partials; integer segments -> int64 sums; min/max -> T.  Short
// segments: a wave takes 64 consecutive segments, whose values occupy one contiguous flat range;
// the wave loads that range together (coalesced index loads, independent value loads) into LDS,
// SEG_CAP values at a time, and each lane then folds its own segment's values in ascending
// order -- the FP1 template's order, one segment per lane.
#define SEG_CAP 256
template<typename T, bool FLT>
__global__ void __launch_bounds__(256) k_seg_thread(const T *x, const int64_t *idx, const int64_t *off, L g,
                                                    int kind, void *out, L nx, int *bad){
    __shared__ T buf[4][SEG_CAP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const L nw = ((L)gridDim.x * blockDim.x) >> 6;
    for(L base = (((L)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64; base < g; base += nw * 64){
        const L s = base + lane;
        const bool valid = s < g;
        const L a = valid ? off[s] : 0, b = valid ? off[s + 1] : 0;
        const L last = g - base < 64 ? g - base - 1 : 63;
        const L A = __shfl(a, 0, 64), E = __shfl(b, (int)last, 64);
        KS facc = {0, 0};
        unsigned long long iacc = 0;
        T r = T();
        bool first = true;
        for(L c0 = A; c0 < E; c0 += SEG_CAP){
            const L c1 = c0 + SEG_CAP < E ? c0 + SEG_CAP : E;
            // index and value loads at clamped positions, so each set is in flight together (an
            // index out of [0, nx) reads x[0] and is flagged; nx is above 0 whenever a flat list
            // has entries, the indices being rows of x)
            int64_t k[SEG_CAP / 64];
            #pragma unroll
            for(int t = 0; t < SEG_CAP / 64; t++){ const L j = c0 + t * 64 + lane, jc = j < c1 ? j : c1 - 1; k[t] = idx ? idx[jc] : jc; }
            T v[SEG_CAP / 64];
            #pragma unroll
            for(int t = 0; t < SEG_CAP / 64; t++){
                const bool inb = !idx || (uint64_t)k[t] < (uint64_t)nx;
                v[t] = x[inb ? k[t] : 0];
            }
            #pragma unroll
            for(int t = 0; t < SEG_CAP / 64; t++){
                const L j = c0 + t * 64 + lane;
                const bool inb = !idx || (uint64_t)k[t] < (uint64_t)nx;
                if(j < c1 && !inb && bad) *bad = 1;
                buf[w][t * 64 + lane] = j < c1 && inb ? v[t] : T();
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const L j0 = a > c0 ? a : c0, j1 = b < c1 ? b : c1;
            for(L j = j0; j < j1; j++){
                const T v = buf[w][j - c0];
                if(kind == GK_SUM || kind == GK_AVG){
                    if(FLT) ks_add(facc, (double)v); else iacc += (unsigned long long)(int64_t)v;
                } else {
                    if(first){ r = v; first = false; }
                    else r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if(!valid) continue;
        if(kind == GK_SUM || kind == GK_AVG){
            if(FLT){ double q = facc.s + facc.c; ((double*)out)[s] = kind == GK_AVG ? q / (double)(b - a) : q; }
            else if(kind == GK_AVG) ((double*)out)[s] = (int64_t)iacc * 1.0 / (double)(b - a);
            else ((int64_t*)out)[s] = (int64_t)iacc;
        } else ((T*)out)[s] = r;
    }
}

// The same fold over a group made of sorted runs (the run-compressed i32 group): thread s folds
// group s's runs in order, each run's rows read in place (first row, length), so the flat list is
// not read; the rows and their order are the flat list's, so the results equal k_seg_thread's.
template<typename T, bool FLT>
__global__ void __launch_bounds__(256) k_seg_runs(const T *x, const uint32_t *sv, const uint32_t *grun, int cb, L g, int kind, void *out){
    const unsigned lmask = (1u << cb) - 1u;
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < g; s += (L)gridDim.x * blockDim.x){
        const uint32_t r0 = grun[s], r1 = grun[s + 1];
        KS facc = {0, 0};
        unsigned long long iacc = 0;
        T r = T();
        bool first = true;
        L cnt = 0;
        for(uint32_t q = r0; q < r1; q++){
            const uint32_t v = sv[q];
            const L row = (L)(v >> cb);
            const uint32_t len = (v & lmask) + 1u;
            cnt += len;
            for(uint32_t t = 0; t < len; t += 4){
                // four rows of the run loaded together (past its end, its last row again)
                T w[4];
                #pragma unroll
                for(int u = 0; u < 4; u++) w[u] = x[row + (t + u < len ? t + u : len - 1)];
                #pragma unroll
                for(int u = 0; u < 4; u++){
                    if(t + u < len){
                        if(kind == GK_SUM || kind == GK_AVG){
                            if(FLT) ks_add(facc, (double)w[u]); else iacc += (unsigned long long)(int64_t)w[u];
                        } else {
                            if(first){ r = w[u]; first = false; }
                            else r = kind == GK_MAX ? (r > w[u] ? r : w[u]) : (r < w[u] ? r : w[u]);
                        }
                    }
                }
            }
        }
        if(kind == GK_SUM || kind == GK_AVG){
            if(FLT){ double q = facc.s + facc.c; ((double*)out)[s] = kind == GK_AVG ? q / (double)cnt : q; }
            else if(kind == GK_AVG) ((double*)out)[s] = (int64_t)iacc * 1.0 / (double)cnt;
            else ((int64_t*)out)[s] = (int64_t)iacc;
        } else ((T*)out)[s] = r;
    }
}

template<typename T, bool FLT>
__global__ void __launch_bounds__(256) k_seg_wave(const T *x, const int64_t *idx, const int64_t *off, L g,
                                                  int kind, void *out, L nx, int *bad){
    int lane = threadIdx.x & 63;
    L wave = ((L)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    L nw = ((L)gridDim.x * blockDim.x) >> 6;
    for(L s = wave; s < g; s += nw){
        L a = off[s], b = off[s + 1];
        if(kind == GK_SUM || kind == GK_AVG){
            if(FLT){
                KS acc = {0, 0};
                for(L j = a + lane; j < b; j += 64) ks_add(acc, (double)seg_val(x, idx, j, nx, bad));
                acc = ks_wave(acc);
                if(lane == 0){ double r = acc.s + acc.c; ((double*)out)[s] = kind == GK_AVG ? r / (double)(b - a) : r; }
            } else {
                unsigned long long acc = 0;
                for(L j = a + lane; j < b; j += 64) acc += (unsigned long long)(int64_t)seg_val(x, idx, j, nx, bad);
                for(int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
                if(lane == 0){ if(kind == GK_AVG) ((double*)out)[s] = (int64_t)acc * 1.0 / (double)(b - a); else ((int64_t*)out)[s] = (int64_t)acc; }
            }
        } else {
            T r = b > a ? seg_val(x, idx, a, nx, bad) : T();
            for(L j = a + lane; j < b; j += 64){ T v = seg_val(x, idx, j, nx, bad); r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v); }
            for(int o = 32; o > 0; o >>= 1){ T v = __shfl_down(r, o, 64); r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v); }
            if(lane == 0) ((T*)out)[s] = r;
        }
    }
}

// multi-block: grid (chunks, segments); partial per (segment, chunk)
#define SEG_CHUNK 32768
template<typename T, bool FLT>
__global__ void __launch_bounds__(256) k_seg_block(const T *x, const int64_t *idx, const int64_t *off,
                                                   int kind, KS *pf, unsigned long long *pi, T *pm, int nch, L nx, int *bad){
    __shared__ KS sh[4];
    __shared__ T shm[4];
    L s = blockIdx.y; int ch = blockIdx.x;
    L a = off[s] + (L)ch * SEG_CHUNK, b = off[s + 1];
    if(a > b) a = b;
    L e = a + SEG_CHUNK < b ? a + SEG_CHUNK : b;
    L slot = s * nch + ch;
    if(kind == GK_SUM || kind == GK_AVG){
        if(FLT){
            KS acc = {0, 0};
            for(L j = a + threadIdx.x; j < e; j += 256) ks_add(acc, (double)seg_val(x, idx, j, nx, bad));
            KS r = ks_block(acc, sh);
            if(threadIdx.x == 0) pf[slot] = r;
        } else {
            unsigned long long acc = 0;
            for(L j = a + threadIdx.x; j < e; j += 256) acc += (unsigned long long)(int64_t)seg_val(x, idx, j, nx, bad);
            for(int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
            __shared__ unsigned long long si[4];
            if((threadIdx.x & 63) == 0) si[threadIdx.x >> 6] = acc;
            __syncthreads();
            if(threadIdx.x == 0) pi[slot] = si[0] + si[1] + si[2] + si[3];
        }
    } else {
        bool any = e > a;
        T r = any ? seg_val(x, idx, a, nx, bad) : T();
        for(L j = a + threadIdx.x; j < e; j += 256){ T v = seg_val(x, idx, j, nx, bad); r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v); }
        for(int o = 32; o > 0; o >>= 1){ T v = __shfl_down(r, o, 64); r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v); }
        if((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = r;
        __syncthreads();
        if(threadIdx.x == 0){
            T q = shm[0];
            for(int w = 1; w < 4; w++){ T v = shm[w]; q = kind == GK_MAX ? (q > v ? q : v) : (q < v ? q : v); }
            if(any) pm[slot] = q; else pm[slot] = (off[s] < off[s+1]) ? seg_val(x, idx, off[s], nx, bad) : T();
        }
    }
}
template<typename T, bool FLT>
__global__ void k_seg_block_final(const int64_t *off, L g, int kind, const KS *pf, const unsigned long long *pi,
                                  const T *pm, int nch, void *out){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < g; s += (L)gridDim.x * blockDim.x){
        L cnt = off[s + 1] - off[s];
        int used = (int)((cnt + SEG_CHUNK - 1) / SEG_CHUNK); if(used < 1) used = 1;
        if(kind == GK_SUM || kind == GK_AVG){
            if(FLT){ KS a = {0, 0}; for(int q = 0; q < used; q++) ks_merge(a, pf[s * nch + q]);
                     double r = a.s + a.c; ((double*)out)[s] = kind == GK_AVG ? r / (double)cnt : r; }
            else { unsigned long long a = 0; for(int q = 0; q < used; q++) a += pi[s * nch + q];
                   if(kind == GK_AVG) ((double*)out)[s] = (int64_t)a * 1.0 / (double)cnt; else ((int64_t*)out)[s] = (int64_t)a; }
        } else {
            T r = pm[s * nch];
            for(int q = 1; q < used; q++){ T v = pm[s * nch + q]; r = kind == GK_MAX ? (r > v ? r : v) : (r < v ? r : v); }
            ((T*)out)[s] = r;
        }
    }
}

__global__ void k_seg_maxlen(const int64_t *off, L g, unsigned long long *mx){
    unsigned long long m = 0;
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < g; s += (L)gridDim.x * blockDim.x){
        unsigned long long v = (unsigned long long)(off[s + 1] - off[s]); m = v > m ? v : m;
    }
    for(int o = 32; o > 0; o >>= 1){ unsigned long long v = __shfl_down(m, o, 64); m = v > m ? v : m; }
    __shared__ unsigned long long sh[4];
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if(threadIdx.x == 0){
        for(int w = 1; w < (int)(blockDim.x >> 6); w++) m = sh[w] > m ? sh[w] : m;
        atomicMax(mx, m);
    }
}

template<typename T, bool FLT>
static int32_t seg_reduce_t(hpa_ctx *c, const T *x, const int64_t *idx, const int64_t *off, L g, L n, int kind, void *out, L nx, int *bad,
                            const hpa_ctx::GroupRuns *runs = nullptr){
    if(g <= 0) R 0;
    double bytes = (double)n * ((idx ? 8 : 0) + sizeof(T)) + 8.0 * g + 8.0 * (g + 1);
    double avg = (double)n / (double)g;
    if(runs && avg <= 16.0){
        ProfScope ps("group_reduce", c->stream, (double)n * sizeof(T) + 4.0 * runs->nr + 4.0 * (g + 1) + 8.0 * g);
        // one group a thread, no grid-stride loop: each thread's loads depend on its run's first
        // row, so many threads in flight, not a long per-thread chain, cover the latency
        const L blocks = ceil_div(g, 256);
        hipLaunchKernelGGL((k_seg_runs<T,FLT>), dim3((unsigned)(blocks < (1LL << 31) ? blocks : (1LL << 31) - 1)), dim3(256), 0, c->stream,
                           x, runs->sv, runs->grun, runs->cb, g, kind, out);
        R check_launch("group_reduce");
    }
    if(avg <= 16.0){
        ProfScope ps("group_reduce", c->stream, bytes);
        hipLaunchKernelGGL((k_seg_thread<T,FLT>), dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, x, idx, off, g, kind, out, nx, bad);
        R check_launch("group_reduce");
    }
    if(avg <= 16384.0 || g > 65535){
        L waves = g; L blocks = ceil_div(waves * 64, 256);
        if(blocks > (L)num_cus() * 32) blocks = (L)num_cus() * 32;
        ProfScope ps("group_reduce", c->stream, bytes);
        hipLaunchKernelGGL((k_seg_wave<T,FLT>), dim3((unsigned)blocks), dim3(256), 0, c->stream, x, idx, off, g, kind, out, nx, bad);
        R check_launch("group_reduce");
    }
    unsigned long long *mx = (unsigned long long*)dalloc(c, 16); if(!mx) R HPA_E_MEMORY;
    hipMemsetAsync(mx, 0, 8, c->stream);
    hipLaunchKernelGGL(k_seg_maxlen, dim3(reduce_grid(g, 256)), dim3(256), 0, c->stream, off, g, mx);
    unsigned long long maxlen = 0; int32_t e = d2h(c, &maxlen, mx, 8); if(e) R e;
    int nch = (int)((maxlen + SEG_CHUNK - 1) / SEG_CHUNK); if(nch < 1) nch = 1;
    KS *pf = nullptr; unsigned long long *pi = nullptr; T *pm = nullptr;
    L slots = g * nch;
    if(kind == GK_SUM || kind == GK_AVG){ if(FLT) pf = (KS*)dalloc(c, sizeof(KS) * slots); else pi = (unsigned long long*)dalloc(c, 8 * slots); }
    else pm = (T*)dalloc(c, sizeof(T) * slots + 16);
    if(!pf && !pi && !pm) R HPA_E_MEMORY;
    {
        ProfScope ps("group_reduce", c->stream, bytes);
        hipLaunchKernelGGL((k_seg_block<T,FLT>), dim3(nch, (unsigned)g), dim3(256), 0, c->stream, x, idx, off, kind, pf, pi, pm, nch, nx, bad);
    }
    hipLaunchKernelGGL((k_seg_block_final<T,FLT>), dim3(grid_for(g, 64, 1)), dim3(64), 0, c->stream, off, g, kind, pf, pi, pm, nch, out);
    R check_launch("group_reduce");
}

// result type of a grouped reduction of element type t
static int32_t seg_out_type(int32_t t, int kind){
    if(kind == GK_LEN) R HPA_L;
    if(kind == GK_AVG) R t == HPA_F ? HPA_F : HPA_E;
    if(kind == GK_SUM) R t == HPA_F ? HPA_F : t == HPA_E ? HPA_E : HPA_L;
    R t;
}

// Grouped reduction into a fresh device vector of g results.
static int32_t seg_reduce(hpa_ctx *c, V x, const void *px, const int64_t *idx, const int64_t *off, L g, L n, int kind,
                          void **out, int32_t *tout, const hpa_ctx::GroupRuns *runs = nullptr){
    int32_t t = x->typ;
    *tout = seg_out_type(t, kind);
    if(kind == GK_LEN) R HPA_E_DOMAIN;
    void *o = dalloc(c, 8 * (g > 0 ? g : 1)); if(!o) R HPA_E_MEMORY;
    *out = o;
    bool fl_out_f32 = (t == HPA_F && (kind == GK_SUM || kind == GK_AVG));
    void *dst = fl_out_f32 ? dalloc(c, 8 * (g > 0 ? g : 1)) : o;
    if(!dst) R HPA_E_MEMORY;
    int32_t e;
    int *bad = nullptr;
    // row lists made here (a group's flat list) carry a bound on their rows: no check to read back
    if(idx && !within_bound(c, idx, x->len)){ bad = (int*)dalloc(c, 16); if(!bad) R HPA_E_MEMORY; hipMemsetAsync(bad, 0, 4, c->stream); }
    L nx = x->len;
    // rows of an empty vector: every entry is out of range (the kernels read x[0] for them)
    if(n > 0 && nx <= 0) R HPA_E_INDEX;
    switch(t){
        case HPA_B: e = seg_reduce_t<uint8_t,false>(c, (const uint8_t*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_J: e = seg_reduce_t<int8_t,false>(c, (const int8_t*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_H: e = seg_reduce_t<int16_t,false>(c, (const int16_t*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_I: e = seg_reduce_t<int32_t,false>(c, (const int32_t*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_L: e = seg_reduce_t<int64_t,false>(c, (const int64_t*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_F: e = seg_reduce_t<float,true>(c, (const float*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        case HPA_E: e = seg_reduce_t<double,true>(c, (const double*)px, idx, off, g, n, kind, dst, nx, bad, runs); break;
        default: R HPA_E_NOT_IMPL;
    }
    if(e) R e;
    if(bad){ int h = 0; if((e = d2h(c, &h, bad, 4))) R e; if(h) R HPA_E_INDEX; }
    if(fl_out_f32){
        // float sums/avgs are accumulated in double and rounded once
        extern int32_t convert_d2f(hpa_ctx *c, const double *s, float *d, L n);
        e = convert_d2f(c, (const double*)dst, (float*)o, g);
    }
    R e;
}

__global__ void k_d2f(const double *s, float *d, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) d[i] = (float)s[i];
}
int32_t convert_d2f(hpa_ctx *c, const double *s, float *d, L n){
    if(n <= 0) R 0;
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(n, 256, 1)), dim3(256), 0, c->stream, s, d, n);
    R check_launch("d2f");
}

// wrap a device vector as a value node
static void set_vec(V z, int32_t typ, L len, void *p, hpa_ctx *c){
    z->typ = typ; z->len = len;
    if(len <= 1 && is_basic(typ)){
        z->g = (char*)&z->b; z->i64 = 0;
        if(len == 1) d2h(c, &z->b, p, type_width(typ));
    } else z->g = (char*)p;
}

// Grouped reduction over a list of group index vectors (flat or normal list), FP templates
extern "C" int32_t hpaGroupReduce(hpa_V z, int32_t kind, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(y->typ != HPA_G) R HPA_E_DOMAIN;
    if(kind < 0 || kind > 4) R HPA_E_DOMAIN;
    int32_t e;
    L g, n; const int64_t *off, *idx;
    if(is_flat(y)){
        if(y->g2->typ != HPA_L) R HPA_E_TYPE;
        if((e = flat_offsets(c, y, &off, &idx, &g, &n))) R e;
    } else {
        // normal list of index vectors: concatenate on the device
        g = y->len; n = 0;
        for(L s = 0; s < g; s++){ V t = ((V)y->g) + s; if(!is_int(t->typ)) R HPA_E_TYPE; n += t->len; }
        int64_t *cat = (int64_t*)dalloc(c, 8 * (n > 0 ? n : 1)), *o = (int64_t*)dalloc(c, 8 * (g + 1));
        if(!cat || !o) R HPA_E_MEMORY;
        std::vector<int64_t> ho(g + 1); L pos = 0;
        for(L s = 0; s < g; s++){
            V t = ((V)y->g) + s; ho[s] = pos;
            const void *p; if((e = promote_dev(c, t, HPA_L, &p))) R e;
            if(t->len > 0) hipMemcpyAsync(cat + pos, p, 8 * t->len, hipMemcpyDeviceToDevice, c->stream);
            pos += t->len;
        }
        ho[g] = pos;
        hipMemcpyAsync(o, ho.data(), 8 * (g + 1), hipMemcpyHostToDevice, c->stream);
        hipStreamSynchronize(c->stream);
        off = o; idx = cat;
    }
    if(kind == GK_LEN){
        int64_t *o = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1)); if(!o) R HPA_E_MEMORY;
        extern int32_t adjacent_diff(hpa_ctx *c, const int64_t *off, int64_t *out, L g);
        if((e = adjacent_diff(c, off, o, g))) R e;
        set_vec(z, HPA_L, g, o, c);
        R 0;
    }
    const void *px = nullptr;
    if(x->len > 0 && (e = dev_view(c, x, &px))) R e;
    // a group built from sorted runs of x's own length: fold each run's rows in place
    const hpa_ctx::GroupRuns *runs = is_flat(y) ? group_runs(c, y->g, g) : nullptr;
    if(runs && (runs->n != n || x->len != n)) runs = nullptr;
    void *out; int32_t tout;
    if((e = seg_reduce(c, x, px, idx, off, g, n, kind, &out, &tout, runs))) R e;
    set_vec(z, tout, g, out, c);
    R 0;
}

__global__ void k_adjdiff(const int64_t *off, int64_t *out, L g){
    for(L s = (L)blockIdx.x * blockDim.x + threadIdx.x; s < g; s += (L)gridDim.x * blockDim.x) out[s] = off[s + 1] - off[s];
}
int32_t adjacent_diff(hpa_ctx *c, const int64_t *off, int64_t *out, L g){
    if(g <= 0) R 0;
    hipLaunchKernelGGL(k_adjdiff, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, off, out, g);
    R check_launch("adjdiff");
}

// ---------------------------------------------------------------- list primitives
// pfnList (h_primitive.cc:32-36): list cells alias the items
extern "C" int32_t pfnList(hpa_V z, int64_t n, hpa_V x[]){
    hpa_ctx *c = ctx_cur();
    z->typ = HPA_G; z->len = n;
    z->g = (char*)halloc(c, sizeof(V0) * (n > 0 ? n : 1));
    if(!z->g) R HPA_E_MEMORY;
    z->g2 = nullptr;
    for(L i = 0; i < n; i++) ((V)z->g)[i] = *x[i];
    R 0;
}

int32_t enum_values(hpa_ctx *c, V z, V x);          // hpa_enum.hip
int32_t enum_keys(V z, V x);
extern "C" int32_t pfnKeys(hpa_V z, hpa_V x){      // h_primitive.cc:190-203
    if(x->typ == HPA_Y && x->p) R enum_keys(z, x);
    if(x->typ == HPA_N || x->typ == HPA_A){ *z = *((V)x->g); R 0; }
    R HPA_E_DOMAIN;
}
extern "C" int32_t pfnValues(hpa_V z, hpa_V x){    // h_primitive.cc:205-218
    if(x->typ == HPA_Y && x->p) R enum_values(ctx_cur(), z, x);     // copyEnumIndex
    if(x->typ == HPA_N || x->typ == HPA_A){ *z = *(((V)x->g) + 1); R 0; }
    R HPA_E_DOMAIN;
}

// flat list whose groups each hold one value: counts all 1, g2 -> the value vector
__global__ void k_fill_i64(int64_t *p, L n, int64_t v){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) p[i] = v;
}
static int32_t make_flat_singletons(hpa_ctx *c, V z, V vals){
    L g = vals->len;
    int64_t *cnt = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1)); if(!cnt) R HPA_E_MEMORY;
    if(g > 0) hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(g, 256, 1)), dim3(256), 0, c->stream, cnt, g, (int64_t)1);
    V node = node_alloc(c); if(!node) R HPA_E_MEMORY;
    *node = *vals;
    z->typ = HPA_G; z->len = g; z->g = (char*)cnt; z->g2 = node;
    R check_launch("flat");
}

// pfnEach (h_primitive.cc:1818-1832).  On a flat list, len/sum/avg/min/max run as one
// segmented reduction over the flat values and return a flat list of singletons, which
// @raze turns back into a vector (the reference's listFlatEachLen form, h_system.cc:759-768).
extern "C" int32_t pfnEach(hpa_V z, hpa_V x, hpa_fn1 foo){
    hpa_ctx *c = ctx_cur();
    if(x->typ != HPA_G) R HPA_E_DOMAIN;
    int32_t e;
    if(is_flat(x)){
        int kind = foo == pfnSum ? GK_SUM : foo == pfnAvg ? GK_AVG : foo == pfnMin ? GK_MIN :
                   foo == pfnMax ? GK_MAX : foo == pfnLen ? GK_LEN : -1;
        if(kind < 0) R HPA_E_NOT_IMPL;
        L g, n; const int64_t *off, *idx;
        if((e = flat_offsets(c, x, &off, &idx, &g, &n))) R e;
        V vals = node_alloc(c); if(!vals) R HPA_E_MEMORY;
        if(kind == GK_LEN){
            int64_t *o = (int64_t*)dalloc(c, 8 * (g > 0 ? g : 1)); if(!o) R HPA_E_MEMORY;
            if((e = adjacent_diff(c, off, o, g))) R e;
            vals->typ = HPA_L; vals->len = g; vals->g = (char*)o;
            if(g <= 1){ vals->g = (char*)&vals->b; vals->i64 = 0; if(g == 1 && (e = d2h(c, &vals->b, o, 8))) R e; }
        } else {
            V t = x->g2;
            const void *pt = nullptr;
            if(t->len > 0 && (e = dev_view(c, t, &pt))) R e;
            void *out; int32_t tout;
            if((e = seg_reduce(c, t, pt, nullptr, off, g, n, kind, &out, &tout))) R e;
            set_vec(vals, tout, g, out, c);
        }
        R make_flat_singletons(c, z, vals);
    }
    L n = x->len;
    z->typ = HPA_G; z->len = n; z->g = (char*)halloc(c, sizeof(V0) * (n > 0 ? n : 1)); z->g2 = nullptr;
    if(!z->g) R HPA_E_MEMORY;
    for(L i = 0; i < n; i++) if((e = foo(((V)z->g) + i, ((V)x->g) + i))) R e;
    R 0;
}

// pfnEachRight (h_primitive.cc:1882-1904).  each_right(@index, x, groups) over a flat list
// gathers once into a flat list with the same group sizes.
extern "C" int32_t pfnEachRight(hpa_V z, hpa_V x, hpa_V y, hpa_fn2 foo){
    hpa_ctx *c = ctx_cur();
    int32_t e;
    if(y->typ == HPA_G){
        if(is_flat(y) && foo == pfnIndex){
            V t = y->g2;
            if(t->typ != HPA_L) R HPA_E_TYPE;
            if(!is_basic(x->typ)) R HPA_E_NOT_IMPL;
            V node = node_alloc(c); if(!node) R HPA_E_MEMORY;
            if((e = pfnIndex(node, x, t))) R e;
            z->typ = HPA_G; z->len = y->len; z->g = y->g; z->g2 = node;
            R 0;
        }
        if(is_flat(y)) R HPA_E_NOT_IMPL;
        L n = y->len;
        z->typ = HPA_G; z->len = n; z->g = (char*)halloc(c, sizeof(V0) * (n > 0 ? n : 1)); z->g2 = nullptr;
        if(!z->g) R HPA_E_MEMORY;
        for(L i = 0; i < n; i++) if((e = foo(((V)z->g) + i, x, ((V)y->g) + i))) R e;
        R 0;
    }
    R foo(z, x, y);
}

// pfnRaze (h_primitive.cc:935-955): a flat list razes to its value vector (copyV of vg2)
extern "C" int32_t pfnRaze(hpa_V z, hpa_V x){
    hpa_ctx *c = ctx_cur();
    if(x->typ == HPA_G){
        if(x->len == 0) R HPA_E_RAZE_LIST;
        if(is_flat(x)){ *z = *x->g2; R 0; }
        int32_t t = ((V)x->g)->typ; L n = 0;
        for(L i = 0; i < x->len; i++){ V v = ((V)x->g) + i; if(v->typ != t) R HPA_E_TYPE; n += v->len; }
        if(!is_basic(t)) R HPA_E_NOT_IMPL;
        void *dz; int32_t e;
        L mark = c->dcur;
        if((e = out_init(c, z, t, n, &dz))) R e;
        L w = type_width(t), pos = 0;
        for(L i = 0; i < x->len; i++){
            V v = ((V)x->g) + i;
            if(v->len == 0) continue;
            hipMemcpyKind k = (is_inline(v) || !is_device(v->g)) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
            if(hipMemcpyAsync((char*)dz + pos * w, v->g, (size_t)(w * v->len), k, c->stream) != hipSuccess) R HPA_E_DEVICE;
            pos += v->len;
        }
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
        R e;
    }
    if(is_basic(x->typ)){ *z = *x; R 0; }
    R HPA_E_DOMAIN;
}

// ---------------------------------------------------------------- FE / FA fused loops
template<int OT>
__global__ void __launch_bounds__(256) k_fused_map(DevExpr e, DevCols x, L n, void *z){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        Slot s = eval_expr(e, x, i);
        bool fk = e.ins[e.n - 1].kf;
        if(OT == HPA_B) ((uint8_t*)z)[i] = fk ? s.f != 0.0 : s.i != 0;
        else if(OT == HPA_I) ((int32_t*)z)[i] = fk ? (int32_t)s.f : (int32_t)s.i;
        else if(OT == HPA_L) ((int64_t*)z)[i] = fk ? (int64_t)s.f : s.i;
        else ((double*)z)[i] = fk ? s.f : (double)s.i;
    }
}

extern "C" int32_t hpaFusedMap(hpa_V z, int32_t typ, int32_t nx, hpa_V *x, const hpa_expr *ex){
    hpa_ctx *c = ctx_cur();
    if(!(typ == HPA_B || typ == HPA_I || typ == HPA_L || typ == HPA_E)) R HPA_E_TYPE;
    DevCols cols; L n; int32_t e;
    L mark = c->dcur;
    if((e = cols_view(c, nx, x, &cols, &n))) R e;
    DevExpr de; int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    if((e = expr_compile(ex, nx, ct, &de))) R e;
    void *dz;
    if((e = out_init(c, z, typ, n, &dz))) R e;
    if(n > 0 && jit_fused_map(c, typ, nx, x, ex, cols, n, dz) != 0){
        double cb = 0; for(int k = 0; k < nx; k++) cb += (double)type_width(x[k]->typ);
        ProfScope ps("fused_map_interp", c->stream, (double)n * (cb + type_width(typ)));
        dim3 g(grid_for(n, 256, 4)), b(256);
        switch(typ){
            case HPA_B: hipLaunchKernelGGL(k_fused_map<HPA_B>, g, b, 0, c->stream, de, cols, n, dz); break;
            case HPA_I: hipLaunchKernelGGL(k_fused_map<HPA_I>, g, b, 0, c->stream, de, cols, n, dz); break;
            case HPA_L: hipLaunchKernelGGL(k_fused_map<HPA_L>, g, b, 0, c->stream, de, cols, n, dz); break;
            default:    hipLaunchKernelGGL(k_fused_map<HPA_E>, g, b, 0, c->stream, de, cols, n, dz); break;
        }
        if((e = check_launch("fused_map"))) R e;
    }
    if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
    R e;
}

// FA reduction: per thread Neumaier (or exact int64), count of selected rows, min/max
struct FaPart { KS s; unsigned long long i; unsigned long long cnt; double mn, mx; long long imn, imx; };
__global__ void __launch_bounds__(256) k_fused_reduce(DevExpr ce, bool has_cond, DevExpr be, DevCols x, L n, FaPart *part){
    __shared__ FaPart sh[4];
    FaPart a; a.s = {0, 0}; a.i = 0; a.cnt = 0; a.mn = DBL_MAX; a.mx = -DBL_MAX; a.imn = LLONG_MAX; a.imx = LLONG_MIN;
    bool bf = be.ins[be.n - 1].kf;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        if(has_cond && !expr_true(ce, eval_expr(ce, x, i))) continue;
        Slot v = eval_expr(be, x, i);
        a.cnt++;
        if(bf){ ks_add(a.s, v.f); a.mn = v.f < a.mn ? v.f : a.mn; a.mx = v.f > a.mx ? v.f : a.mx; }
        else { a.i += (unsigned long long)v.i; a.imn = v.i < a.imn ? v.i : a.imn; a.imx = v.i > a.imx ? v.i : a.imx; }
    }
    // wave then block merge, fixed order
    for(int o = 32; o > 0; o >>= 1){
        KS b = ks_shfl_down(a.s, o); ks_merge(a.s, b);
        a.i += __shfl_down(a.i, o, 64); a.cnt += __shfl_down(a.cnt, o, 64);
        double m1 = __shfl_down(a.mn, o, 64), m2 = __shfl_down(a.mx, o, 64);
        long long q1 = __shfl_down(a.imn, o, 64), q2 = __shfl_down(a.imx, o, 64);
        a.mn = m1 < a.mn ? m1 : a.mn; a.mx = m2 > a.mx ? m2 : a.mx; a.imn = q1 < a.imn ? q1 : a.imn; a.imx = q2 > a.imx ? q2 : a.imx;
    }
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if(threadIdx.x == 0){
        FaPart r = sh[0];
        for(int w = 1; w < 4; w++){
            const FaPart &b = sh[w];
            ks_merge(r.s, b.s); r.i += b.i; r.cnt += b.cnt;
            r.mn = b.mn < r.mn ? b.mn : r.mn; r.mx = b.mx > r.mx ? b.mx : r.mx;
            r.imn = b.imn < r.imn ? b.imn : r.imn; r.imx = b.imx > r.imx ? b.imx : r.imx;
        }
        part[blockIdx.x] = r;
    }
}
__device__ __forceinline__ void fa_merge(FaPart &r, const FaPart &b){
    ks_merge(r.s, b.s); r.i += b.i; r.cnt += b.cnt;
    r.mn = b.mn < r.mn ? b.mn : r.mn; r.mx = b.mx > r.mx ? b.mx : r.mx;
    r.imn = b.imn < r.imn ? b.imn : r.imn; r.imx = b.imx > r.imx ? b.imx : r.imx;
}
// np partials merged in a fixed order (block q merges part[q * np .. q * np + np)): thread t
// folds partials t, t + 1024, ... (coalesced loads, all in flight together), then a pairwise
// tree in LDS.  Deterministic: the same association on every run.
#define FA_FINAL 1024
__global__ void __launch_bounds__(FA_FINAL) k_fused_reduce_final(const FaPart *part, int np, FaPart *out){
    __shared__ FaPart sh[FA_FINAL];
    const FaPart *p = part + (L)blockIdx.x * np;
    FaPart a; a.s = {0, 0}; a.i = 0; a.cnt = 0; a.mn = DBL_MAX; a.mx = -DBL_MAX; a.imn = LLONG_MAX; a.imx = LLONG_MIN;
    for(int q = threadIdx.x; q < np; q += FA_FINAL) fa_merge(a, p[q]);
    sh[threadIdx.x] = a;
    __syncthreads();
    for(int s = FA_FINAL / 2; s > 0; s >>= 1){
        if((int)threadIdx.x < s) fa_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if(threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Shared-scan FA reductions: z[q] = sum (kind 0) or avg (kind 1) of body over the rows where
// conds[q] holds, for nq <= 8 conditions over the same columns, computed in one pass (each
// result as hpaFusedReduce would give it alone).
extern "C" int32_t hpaFusedReduceMany(int32_t nq, hpa_V *z, int32_t kind, int32_t nx, hpa_V *x,
                                      const hpa_expr *conds, const hpa_expr *body){
    hpa_ctx *c = ctx_cur();
    if(kind < 0 || kind > 1 || nq < 1 || nq > 8) R HPA_E_DOMAIN;
    DevCols cols; L n; int32_t e;
    L mark = c->dcur;
    if((e = cols_view(c, nx, x, &cols, &n))) R e;
    int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    DevExpr be;
    if((e = expr_compile(body, nx, ct, &be))) R e;
    bool bf = be.ins[be.n - 1].kf;
    int G = sum_grid(n);
    FaPart *part = (FaPart*)dalloc(c, sizeof(FaPart) * G * nq), *out = (FaPart*)dalloc(c, sizeof(FaPart) * nq);
    if(!part || !out){ arena_set(c, mark); R HPA_E_MEMORY; }
    if(jit_fused_reduce_many(c, nq, nx, x, conds, body, cols, n, part, G) != 0){
        arena_set(c, mark);
        for(int q = 0; q < nq; q++) if((e = hpaFusedReduce(z[q], kind, nx, x, &conds[q], body))) R e;
        R 0;
    }
    hipLaunchKernelGGL(k_fused_reduce_final, dim3(nq), dim3(FA_FINAL), 0, c->stream, part, G, out);
    if((e = check_launch("fused_reduce_many"))) R e;
    FaPart h[8]; if((e = d2h(c, h, out, sizeof(FaPart) * nq))) R e;
    arena_set(c, mark);
    for(int q = 0; q < nq; q++){
        V zq = z[q];
        zq->len = 1; zq->g = (char*)&zq->b; zq->i64 = 0;
        if(bf){ double t = h[q].s.s + h[q].s.c; zq->typ = HPA_E; zq->f64 = kind == 0 ? t : t / (double)h[q].cnt; }
        else if(kind == 1){ zq->typ = HPA_E; zq->f64 = (double)(int64_t)h[q].i / (double)h[q].cnt; }
        else { zq->typ = HPA_L; zq->i64 = (int64_t)h[q].i; }
    }
    R 0;
}

// FA-reduce (optimizer/autofusion.cc:795-817,886-917): sum c += body over rows where cond
// holds; avg c/tot; min/max seed with the type's extreme as the emitted code does.
extern "C" int32_t hpaFusedReduce(hpa_V z, int32_t kind, int32_t nx, hpa_V *x, const hpa_expr *cond, const hpa_expr *body){
    hpa_ctx *c = ctx_cur();
    if(kind < 0 || kind > 3) R HPA_E_DOMAIN;
    DevCols cols; L n; int32_t e;
    L mark = c->dcur;
    if((e = cols_view(c, nx, x, &cols, &n))) R e;
    int32_t ct[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) ct[k] = x[k]->typ;
    DevExpr ce, be;
    if(cond && (e = expr_compile(cond, nx, ct, &ce))) R e;
    if(!cond){ ce.n = 0; ce.depth = 0; }
    if((e = expr_compile(body, nx, ct, &be))) R e;
    bool bf = be.ins[be.n - 1].kf;
    int G = sum_grid(n);
    FaPart *part = (FaPart*)dalloc(c, sizeof(FaPart) * G), *out = (FaPart*)dalloc(c, sizeof(FaPart));
    if(!part || !out) R HPA_E_MEMORY;
    if(jit_fused_reduce(c, kind, nx, x, cond, body, cols, n, part, G) != 0){
        // generated code unavailable: interpret the expressions on the device
        double cb = 0; for(int k = 0; k < nx; k++) cb += (double)type_width(x[k]->typ);
        ProfScope ps("fused_reduce_interp", c->stream, (double)n * cb);
        hipLaunchKernelGGL(k_fused_reduce, dim3(G), dim3(256), 0, c->stream, ce, cond != nullptr, be, cols, n, part);
    }
    hipLaunchKernelGGL(k_fused_reduce_final, dim3(1), dim3(FA_FINAL), 0, c->stream, part, G, out);
    if((e = check_launch("fused_reduce"))) R e;
    FaPart h; if((e = d2h(c, &h, out, sizeof(FaPart)))) R e;
    arena_set(c, mark);
    z->len = 1; z->g = (char*)&z->b; z->i64 = 0;
    if(bf){
        z->typ = HPA_E;
        double s = h.s.s + h.s.c;
        z->f64 = kind == 0 ? s : kind == 1 ? s / (double)h.cnt : kind == 2 ? h.mn : h.mx;
    } else {
        if(kind == 1){ z->typ = HPA_E; z->f64 = (double)(int64_t)h.i / (double)h.cnt; }
        else { z->typ = HPA_L; z->i64 = kind == 0 ? (int64_t)h.i : kind == 2 ? h.imn : h.imx; }
    }
    R 0;
}
