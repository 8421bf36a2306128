// Runtime of libhorsepower_amd: device/host arenas, contexts, residency, staging, symbols,
// kernel timing and the device-wide scan used by every order-preserving primitive.
#include "hpa_internal.h"
#include <deque>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <mutex>
#include <algorithm>
#include <atomic>
#include <chrono>

static std::mutex g_mu;
static bool g_init = false;
static int  g_device = 0;
static int  g_cus = 256;
static hpa_ctx *g_default = nullptr;
static thread_local hpa_ctx *t_cur = nullptr;
static thread_local char t_err[512];

struct Range { uintptr_t lo, hi; };
static std::vector<Range> g_ranges;            // device ranges: arenas and resident payloads
static std::map<uintptr_t, L> g_resident;       // resident payload -> bytes
static std::map<uintptr_t, uintptr_t> g_res_host; // hpaToHost copy of a resident payload -> the payload

int num_cus(){ R g_cus; }

void set_error_detail(const char *fmt, ...){
    va_list ap; va_start(ap, fmt); vsnprintf(t_err, sizeof(t_err), fmt, ap); va_end(ap);
}
extern "C" const char *hpaLastErrorDetail(void){ R t_err; }

extern "C" const char *hpaErrorString(int32_t e){
    switch(e){
        case 0: R "ok";
        case HPA_E_DOMAIN: R "domain error";
        case HPA_E_GENERAL: R "general error";
        case HPA_E_INDEX: R "index error";
        case HPA_E_DIV_ZERO: R "divide by zero";
        case HPA_E_LENGTH: R "length error";
        case HPA_E_TYPE: R "type error";
        case HPA_E_NULL_VALUE: R "null value";
        case HPA_E_MATCH: R "match error";
        case HPA_E_RAZE_LIST: R "raze list error";
        case HPA_E_UNKNOWN: R "unknown error";
        case HPA_E_NOT_IMPL: R "not implemented";
        case HPA_E_MEMORY: R "device arena exhausted";
        case HPA_E_DEVICE: R "HIP runtime error";
    }
    R "error";
}

int32_t check_launch(const char *what){
    hipError_t e = hipGetLastError();
    if(e != hipSuccess){ set_error_detail("%s: %s", what, hipGetErrorString(e)); R HPA_E_DEVICE; }
    R 0;
}

// ---------------------------------------------------------------- device ranges
void register_device_range(void *p, L bytes){
    std::lock_guard<std::mutex> lk(g_mu);
    g_ranges.push_back({(uintptr_t)p, (uintptr_t)p + (uintptr_t)bytes});
}
void unregister_device_range(void *p){
    std::lock_guard<std::mutex> lk(g_mu);
    for(size_t i=0;i<g_ranges.size();i++) if(g_ranges[i].lo == (uintptr_t)p){ g_ranges.erase(g_ranges.begin()+i); break; }
}
bool is_device(const void *p){
    uintptr_t u = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_mu);
    for(auto &r : g_ranges) if(u >= r.lo && u < r.hi) R true;
    R false;
}
extern "C" int32_t hpaIsDevicePtr(const void *p){ R is_device(p) ? 1 : 0; }

// ---------------------------------------------------------------- contexts
static hpa_ctx *ctx_new(L arena_bytes, L host_bytes){
    hpa_ctx *c = new hpa_ctx();
    c->device = g_device;
    if(hipSetDevice(g_device) != hipSuccess){ delete c; R nullptr; }
    if(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess){ delete c; R nullptr; }
    c->own_stream = true;
    if(arena_bytes < 0){
        c->own_arena = false;                                  // the caller attaches a pool slice
    } else {
    if(arena_bytes == 0) arena_bytes = 17179869184LL;          // 16 GiB (backend/h_memory.cc:11)
    size_t freeb = 0, totalb = 0;
    if(hipMemGetInfo(&freeb, &totalb) == hipSuccess && (size_t)arena_bytes > freeb - (512ull<<20))
        arena_bytes = (L)(freeb > (1ull<<30) ? freeb - (512ull<<20) : freeb/2);
    if(hipMalloc((void**)&c->dbase, (size_t)arena_bytes) != hipSuccess){
        hipStreamDestroy(c->stream); delete c; R nullptr;
    }
    c->dtop = arena_bytes;
    register_device_range(c->dbase, arena_bytes);
    }
    c->hbase = (char*)malloc((size_t)host_bytes); c->htop = host_bytes;
    c->pinned_bytes = 1 << 20;
    if(hipHostMalloc((void**)&c->pinned, (size_t)c->pinned_bytes, hipHostMallocDefault) != hipSuccess) c->pinned = nullptr;
    if(hipHostMalloc((void**)&c->mapped, PUB_HEAD + PUB_BYTES, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) c->mapped = nullptr;
    else if(hipHostGetDevicePointer((void**)&c->mapped_dev, c->mapped, 0) != hipSuccess){ hipHostFree(c->mapped); c->mapped = nullptr; }
    else memset(c->mapped, 0, PUB_HEAD);
    R c;
}

extern "C" int32_t hpaInit(int32_t device, int64_t arena_bytes){
    if(g_init) R 0;
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n <= device){ set_error_detail("no HIP device %d", device); R HPA_E_DEVICE; }
    g_device = device;
    if(hipSetDevice(device) != hipSuccess) R HPA_E_DEVICE;
    hipDeviceProp_t pr;
    if(hipGetDeviceProperties(&pr, device) == hipSuccess) g_cus = pr.multiProcessorCount;
    const char *env = getenv("HPA_ARENA_BYTES");
    if(env && arena_bytes <= 0) arena_bytes = atoll(env);
    g_default = ctx_new(arena_bytes, 256LL << 20);
    if(!g_default){ set_error_detail("default context allocation failed"); R HPA_E_MEMORY; }
    int32_t e = noalloc_scratch_init(device);
    if(e){ hpaCtxDestroy(g_default); g_default = nullptr; R e; }
    g_init = true;
    R 0;
}
bool hpa_is_initialized(){ R g_init; }

extern "C" void hpaCtxDestroy(hpa_ctx *c){
    if(!c) R;
    hipStreamSynchronize(c->stream);
    if(c->dbase && c->own_arena){ unregister_device_range(c->dbase); hipFree(c->dbase); }
    if(c->own_stream) hipStreamDestroy(c->stream);
    if(c->pinned) hipHostFree(c->pinned);
    if(c->mapped) hipHostFree(c->mapped);
    for(auto &hb : c->host_bufs) free(hb.p);
    free(c->hbase);
    if(t_cur == c) t_cur = nullptr;
    delete c;
}

extern "C" void hpaShutdown(void){
    if(!g_init) R;
    hipDeviceSynchronize();
    hpaCtxDestroy(g_default); g_default = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for(auto &kv : g_resident) hipFree((void*)kv.first);
        for(auto &kv : g_res_host) free((void*)kv.first);
        g_resident.clear(); g_res_host.clear(); g_ranges.clear();
    }
    g_init = false;
}

extern "C" hpa_ctx *hpaCtxCreate(int64_t arena_bytes){
    if(!g_init && hpaInit(0, 0)) R nullptr;
    R ctx_new(arena_bytes > 0 ? arena_bytes : arena_bytes < 0 ? -1 : (1LL<<30), 64LL << 20);
}

// A long-lived device allocation registered as device memory (a pool that contexts' arenas are
// carved from, hpaCtxSetArena).
extern "C" void *hpaDeviceAlloc(int64_t bytes){
    if(!g_init && hpaInit(0, 0)) R nullptr;
    void *p = nullptr;
    if(bytes <= 0 || hipMalloc(&p, (size_t)bytes) != hipSuccess){ set_error_detail("hipMalloc of %lld bytes failed", (long long)bytes); R nullptr; }
    register_device_range(p, bytes);
    R p;
}
extern "C" int32_t hpaDeviceFree(void *p){
    if(!p) R 0;
    hipDeviceSynchronize();
    unregister_device_range(p);
    R hipFree(p) == hipSuccess ? 0 : HPA_E_DEVICE;
}

// Point a context created with a negative arena size at [base, base+bytes) of a registered
// device allocation and rewind it (device arena, host nodes, marks): the next request starts on
// an empty arena of exactly that budget.  The context's stream must be idle.
extern "C" int32_t hpaCtxSetArena(hpa_ctx *c, void *base, int64_t bytes){
    if(!c || c->own_arena || bytes < 0){ set_error_detail("hpaCtxSetArena: context owns its arena"); R HPA_E_DOMAIN; }
    if(base && !(is_device(base) && is_device((char*)base + bytes - 1))){ set_error_detail("hpaCtxSetArena: not a device pool range"); R HPA_E_DOMAIN; }
    if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
    c->dbase = (char*)base; c->dtop = base ? bytes : 0; c->dcur = 0; c->dpeak = 0;
    c->hcur = 0; c->marks.clear(); c->bounds.clear(); c->gvals.clear(); c->goffs.clear(); c->gruns.clear();
    for(auto &hb : c->host_bufs) free(hb.p);
    c->host_bufs.clear();
    R 0;
}
extern "C" hpa_ctx *hpaCtxDefault(void){ R g_default; }
extern "C" void hpaCtxSetCurrent(hpa_ctx *c){ t_cur = c; }
extern "C" hpa_ctx *hpaCtxCurrent(void){ R ctx_cur(); }
extern "C" void *hpaCtxStream(hpa_ctx *c){ R c ? (void*)c->stream : nullptr; }
extern "C" int64_t hpaCtxArenaBytes(hpa_ctx *c){ R c ? c->dtop : 0; }
extern "C" int64_t hpaCtxArenaPeak(hpa_ctx *c){ R c ? c->dpeak : 0; }
extern "C" void hpaCtxResetPeak(hpa_ctx *c){ if(c) c->dpeak = c->dcur; }

hpa_ctx *ctx_cur(){
    if(t_cur) R t_cur;
    if(!g_init) hpaInit(0, 0);
    R g_default;
}

void *dalloc(hpa_ctx *c, L bytes){
    L a = (bytes + 255) & ~(L)255;
    if(a == 0) a = 256;
    if(c->dcur + a > c->dtop){
        set_error_detail("device arena full: %lld + %lld > %lld", (long long)c->dcur, (long long)a, (long long)c->dtop);
        R nullptr;
    }
    void *p = c->dbase + c->dcur; c->dcur += a;
    if(c->dcur > c->dpeak) c->dpeak = c->dcur;
    R p;
}

void *halloc(hpa_ctx *c, L bytes){
    L a = (bytes + 15) & ~(L)15;
    if(c->hcur + a > c->htop){ set_error_detail("host node arena full"); R nullptr; }
    void *p = c->hbase + c->hcur; c->hcur += a;
    memset(p, 0, (size_t)a);
    R p;
}

V node_alloc(hpa_ctx *c){ R (V)halloc(c, sizeof(V0)); }

extern "C" hpa_V hpaAllocNode(void){ R node_alloc(ctx_cur()); }
extern "C" int32_t hpaAllocNodes(int32_t n, hpa_V *out){
    hpa_ctx *c = ctx_cur();
    for(int32_t i = 0; i < n; i++){ out[i] = node_alloc(c); if(!out[i]) R HPA_E_MEMORY; }
    R 0;
}

extern "C" int64_t hpaGetHeapOffset(void){
    hpa_ctx *c = ctx_cur();
    c->marks[c->dcur] = {c->hcur, c->host_seq};
    R c->dcur;
}
extern "C" int64_t hpaSetHeapOffset(int64_t cur){
    hpa_ctx *c = ctx_cur();
    hipStreamSynchronize(c->stream);          // outstanding kernels may still use the payloads
    auto it = c->marks.find(cur);
    // host copies made after the mark die with it; without a recorded mark, those made while the
    // arena stood above the new offset
    const bool known = it != c->marks.end();
    const uint64_t seq = known ? it->second.seq : 0;
    if(known) c->hcur = it->second.hcur;
    c->marks.erase(c->marks.upper_bound(cur), c->marks.end());
    size_t k = 0;
    for(auto &hb : c->host_bufs){
        if(known ? hb.seq >= seq : hb.dcur > cur) free(hb.p);
        else c->host_bufs[k++] = hb;
    }
    c->host_bufs.resize(k);
    arena_set(c, cur);
    R cur;
}

// Every move of the arena top goes through here: the side tables keyed by arena address (index
// bounds, group key values, group offsets) lose the entries at or above the new top, so a later
// allocation at a freed address never inherits a dead buffer's facts.
void arena_set(hpa_ctx *c, L cur){
    const void *top = (const void*)(c->dbase + cur);
    c->bounds.erase(c->bounds.lower_bound(top), c->bounds.end());
    c->gvals.erase(c->gvals.lower_bound(top), c->gvals.end());
    c->goffs.erase(c->goffs.lower_bound(top), c->goffs.end());
    c->gruns.erase(c->gruns.lower_bound(top), c->gruns.end());
    str_prune(c, cur);
    c->dcur = cur;
}

void note_bound(hpa_ctx *c, const void *p, L bound){
    if(p && (const char*)p >= c->dbase && (const char*)p < c->dbase + c->dtop) c->bounds[p] = bound;
}
bool within_bound(hpa_ctx *c, const void *p, L limit){
    auto it = c->bounds.find(p);
    R it != c->bounds.end() && it->second <= limit;
}
static bool in_arena(hpa_ctx *c, const void *p){ R p && (const char*)p >= c->dbase && (const char*)p < c->dbase + c->dtop; }
void note_group_vals(hpa_ctx *c, const void *keys, const void *src, L n, const void *vals){
    if(in_arena(c, keys) && in_arena(c, src) && in_arena(c, vals) && vals < keys && src < vals) c->gvals[keys] = {src, n, vals};
}
const void *group_vals(hpa_ctx *c, const void *keys, const void *src, L n){
    auto it = c->gvals.find(keys);
    R it != c->gvals.end() && it->second.src == src && it->second.n == n ? it->second.vals : nullptr;
}
void note_group_runs(hpa_ctx *c, const void *counts, const hpa_ctx::GroupRuns &r){
    if(r.g > 0 && in_arena(c, counts) && in_arena(c, r.sv) && in_arena(c, r.grun) && (const void*)r.sv < counts &&
       (const void*)r.grun < counts) c->gruns[counts] = r;
}
const hpa_ctx::GroupRuns *group_runs(hpa_ctx *c, const void *counts, L g){
    auto it = c->gruns.find(counts);
    R it != c->gruns.end() && it->second.g == g ? &it->second : nullptr;
}
void note_group_offsets(hpa_ctx *c, const void *counts, L g, const int64_t *offs){
    if(g > 0 && in_arena(c, counts) && in_arena(c, offs) && (const void*)offs < counts) c->goffs[counts] = {g, offs};
}
static const int64_t *group_offsets(hpa_ctx *c, const void *counts, L g){
    auto it = c->goffs.find(counts);
    R it != c->goffs.end() && it->second.g == g ? it->second.offs : nullptr;
}

// ---------------------------------------------------------------- value model
L type_width(int32_t typ){
    switch(typ){
        case HPA_B: case HPA_J: case HPA_C: R 1;
        case HPA_H: R 2;
        case HPA_I: case HPA_F: case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T: R 4;
        case HPA_L: case HPA_E: case HPA_X: case HPA_Q: case HPA_S: case HPA_Z: case HPA_Y: R 8;
        case HPA_G: case HPA_N: case HPA_A: case HPA_K: case HPA_V: R (L)sizeof(V0);
    }
    R -1;
}
L type_bytes(int32_t typ, L len){ R type_width(typ) * len + (typ == HPA_C ? 1 : 0); }
bool is_basic(int32_t t){ R t >= HPA_B && t <= HPA_T; }
bool is_int(int32_t t){ R t==HPA_B||t==HPA_J||t==HPA_H||t==HPA_I||t==HPA_L; }
bool is_real(int32_t t){ R is_int(t) || t==HPA_F || t==HPA_E; }
bool is_comparable(int32_t t){ R is_real(t) || t==HPA_C || t==HPA_Q || (t>=HPA_M && t<=HPA_T); }
bool is_inline(V x){ R x->g == (char*)&x->b; }
bool is_flat(V x){ R x->typ == HPA_G && x->g2 != nullptr; }

// ---------------------------------------------------------------- small device->host reads
// A one-workgroup kernel copies the device ranges into the context's coherent mapped host buffer
// and then writes a new sequence number at its head; the host spins on that number and copies
// out.  Measured on MI355X: a kernel followed by an 8-byte read takes 7.3 us a round trip this
// way against 16.6 us with a copy command and a blocking stream synchronisation.  A spin longer
// than PUB_SPIN_US falls back to a blocking synchronisation, which also reports a failed stream.
#define PUB_MAX 32
struct PubList { const char *src[PUB_MAX]; int64_t off[PUB_MAX]; int64_t bytes[PUB_MAX]; int n; };
__global__ void __launch_bounds__(256) k_publish(PubList r, char *dst, unsigned long long *seqp, unsigned long long seq){
    for(int k = 0; k < r.n; k++){
        const char *s = r.src[k]; char *d = dst + r.off[k]; const L nb = r.bytes[k];
        const bool w8 = ((uintptr_t)s % 8 == 0) && ((uintptr_t)d % 8 == 0);
        if(w8) for(L i = (L)threadIdx.x * 8; i + 8 <= nb; i += 256 * 8) *(uint64_t*)(d + i) = *(const uint64_t*)(s + i);
        for(L j = (w8 ? (nb & ~(L)7) : 0) + threadIdx.x; j < nb; j += 256) d[j] = s[j];
    }
    __syncthreads();
    if(threadIdx.x == 0){
        __threadfence_system();
        __hip_atomic_store(seqp, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}
#define PUB_SPIN_US 4000
static int32_t pub_wait(hpa_ctx *c, uint64_t seq){
    volatile uint64_t *f = (volatile uint64_t*)c->mapped;
    const auto t0 = std::chrono::steady_clock::now();
    for(uint64_t it = 1; *f < seq; it++){
        __builtin_ia32_pause();
        if((it & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(PUB_SPIN_US)){
            if(hipStreamSynchronize(c->stream) != hipSuccess){ set_error_detail("stream failed before a device read"); R HPA_E_DEVICE; }
            if(*f < seq){ set_error_detail("device read: completion number %llu not published", (unsigned long long)seq); R HPA_E_DEVICE; }
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    R 0;
}
// n device ranges into the mapped buffer at 16-byte aligned offsets (returned in off[]); false
// when they do not fit one publish
static bool pub_fits(int n, const int64_t *bytes){
    if(n > PUB_MAX) R false;
    L tot = 0; for(int k = 0; k < n; k++) tot += (bytes[k] + 15) & ~(L)15;
    R tot <= PUB_BYTES - PUB_ASYNC_SLOTS * PUB_ASYNC_BYTES;
}
static int32_t pub_read(hpa_ctx *c, int n, const void *const *src, void *const *dst, const int64_t *bytes){
    PubList r; r.n = 0; L off = 0;
    for(int k = 0; k < n; k++){
        if(bytes[k] <= 0) continue;
        r.src[r.n] = (const char*)src[k]; r.off[r.n] = off; r.bytes[r.n] = bytes[k]; r.n++;
        off += (bytes[k] + 15) & ~(L)15;
    }
    if(r.n == 0) R 0;
    const uint64_t seq = ++c->pub_seq;
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(256), 0, c->stream, r, c->mapped_dev + PUB_HEAD,
                       (unsigned long long*)c->mapped_dev, (unsigned long long)seq);
    int32_t e;
    if((e = check_launch("device read"))) R e;
    if((e = pub_wait(c, seq))) R e;
    int j = 0;
    for(int k = 0; k < n; k++) if(bytes[k] > 0){ memcpy(dst[k], c->mapped + PUB_HEAD + r.off[j], (size_t)bytes[k]); j++; }
    R 0;
}

// begun reads land in slot (ticket % PUB_ASYNC_SLOTS) at the end of the data area; the
// sequence number is shared with the synchronous reads, so a later number implies an earlier one
static L pub_async_off(uint64_t ticket){ R PUB_HEAD + PUB_BYTES - (L)PUB_ASYNC_SLOTS * PUB_ASYNC_BYTES + (L)(ticket % PUB_ASYNC_SLOTS) * PUB_ASYNC_BYTES; }
int32_t d2h_begin(hpa_ctx *c, const void *src, L bytes, uint64_t *ticket){
    if(bytes <= 0 || bytes > PUB_ASYNC_BYTES){ set_error_detail("d2h_begin of %lld bytes", (long long)bytes); R HPA_E_DOMAIN; }
    const uint64_t seq = ++c->pub_seq;
    *ticket = seq;
    if(!c->mapped) R d2h(c, c->pub_fallback[seq % PUB_ASYNC_SLOTS], src, bytes);
    PubList r; r.n = 1; r.src[0] = (const char*)src; r.off[0] = 0; r.bytes[0] = bytes;
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(256), 0, c->stream, r, c->mapped_dev + pub_async_off(seq),
                       (unsigned long long*)c->mapped_dev, (unsigned long long)seq);
    R check_launch("device read");
}
int32_t d2h_end(hpa_ctx *c, uint64_t ticket, void *dst, L bytes){
    if(!c->mapped){ memcpy(dst, c->pub_fallback[ticket % PUB_ASYNC_SLOTS], (size_t)bytes); R 0; }
    int32_t e = pub_wait(c, ticket); if(e) R e;
    memcpy(dst, c->mapped + pub_async_off(ticket), (size_t)bytes);
    R 0;
}

int32_t d2h(hpa_ctx *c, void *dst, const void *src, L bytes){
    if(bytes <= 0) R 0;
    if(c->mapped && bytes <= PUB_BYTES - PUB_ASYNC_SLOTS * PUB_ASYNC_BYTES && is_device(src)){
        const void *s1[1] = {src}; void *d1[1] = {dst}; const int64_t b1[1] = {bytes};
        R pub_read(c, 1, s1, d1, b1);
    }
    if(c->pinned && bytes <= c->pinned_bytes){
        if(hipMemcpyAsync(c->pinned, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        memcpy(dst, c->pinned, (size_t)bytes);
        R 0;
    }
    if(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
    R 0;
}

int32_t out_init(hpa_ctx *c, V z, int32_t typ, L len, void **dev){
    z->typ = typ; z->len = len;
    L bytes = type_bytes(typ, len);
    if(is_basic(typ) && len <= 1){
        z->g = (char*)&z->b;
        z->i64 = 0;
        *dev = dalloc(c, 16);
        R *dev ? 0 : HPA_E_MEMORY;
    }
    void *p = dalloc(c, bytes > 0 ? bytes : 1);
    if(!p) R HPA_E_MEMORY;
    z->g = (char*)p;
    if(typ == HPA_G) z->g2 = nullptr;
    *dev = p;
    R 0;
}

int32_t commit_inline(hpa_ctx *c, V z, const void *dev){
    if(!is_inline(z) || z->len < 1) R 0;
    L w = type_width(z->typ);
    R d2h(c, &z->b, dev, w);
}

extern "C" int32_t hpaInitV(hpa_V x, int32_t typ, int64_t len){
    void *d; R out_init(ctx_cur(), x, typ, len, &d);
}

int32_t dev_view(hpa_ctx *c, V x, const void **p){
    L bytes = type_bytes(x->typ, x->len);
    if(x->len == 0){ *p = dalloc(c, 16); R *p ? 0 : HPA_E_MEMORY; }
    if(is_inline(x)){
        void *d = dalloc(c, 16); if(!d) R HPA_E_MEMORY;
        if(hipMemcpyAsync(d, &x->b, 8, hipMemcpyHostToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        *p = d; R 0;
    }
    if(is_device(x->g)){ *p = x->g; R 0; }
    void *d = dalloc(c, bytes); if(!d) R HPA_E_MEMORY;
    if(hipMemcpyAsync(d, x->g, (size_t)bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
    *p = d;
    R 0;
}

int32_t read_scalar(hpa_ctx *c, V x, void *dst, L width){
    if(is_inline(x) || !is_device(x->g)){ memcpy(dst, x->g, (size_t)width); R 0; }
    R d2h(c, dst, x->g, width);
}

// ---------------------------------------------------------------- residency
// A long-lived HBM payload for x (typ, len): kept across heap resets, released by
// hpaFreeResident.  Used by device-side table generation.
int32_t resident_alloc(V x, int32_t typ, L len){
    L bytes = type_bytes(typ, len);
    void *d = nullptr;
    if(hipMalloc(&d, (size_t)(bytes > 0 ? bytes : 1)) != hipSuccess){ set_error_detail("hipMalloc(%lld) failed", (long long)bytes); R HPA_E_MEMORY; }
    register_device_range(d, bytes > 0 ? bytes : 1);
    { std::lock_guard<std::mutex> lk(g_mu); g_resident[(uintptr_t)d] = bytes; }
    x->typ = typ; x->len = len; x->g = (char*)d;
    R 0;
}

static int32_t upload_rec(hpa_ctx *c, V x){
    if(is_basic(x->typ)){
        if(is_inline(x) || x->len <= 1 || is_device(x->g)) R 0;
        if(x->typ == HPA_S){ int32_t e = str_upload(c, x); if(e) R e; }   // the strings' bytes first
        L bytes = type_bytes(x->typ, x->len);
        void *d = nullptr;
        if(hipMalloc(&d, (size_t)bytes) != hipSuccess){ set_error_detail("hipMalloc(%lld) failed", (long long)bytes); R HPA_E_MEMORY; }
        if(hipMemcpy(d, x->g, (size_t)bytes, hipMemcpyHostToDevice) != hipSuccess){ hipFree(d); R HPA_E_DEVICE; }
        register_device_range(d, bytes);
        { std::lock_guard<std::mutex> lk(g_mu); g_resident[(uintptr_t)d] = bytes; }
        x->g = (char*)d;
        R 0;
    }
    if(x->typ == HPA_G || x->typ == HPA_N || x->typ == HPA_A){
        if(is_flat(x)) R upload_rec(c, x->g2);
        L n = x->typ == HPA_G ? x->len : 2;
        for(L i=0;i<n;i++){ int32_t e = upload_rec(c, ((V)x->g) + i); if(e) R e; }
        R 0;
    }
    R HPA_E_NOT_IMPL;
}
extern "C" int32_t hpaUpload(hpa_V x){ R upload_rec(ctx_cur(), x); }

extern "C" int32_t hpaFreeResident(hpa_V x){
    void *p = x->g;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto hc = g_res_host.find((uintptr_t)p);     // a node that hpaToHost moved to host memory
        if(hc != g_res_host.end()){ free(p); p = (void*)hc->second; g_res_host.erase(hc); }
        auto it = g_resident.find((uintptr_t)p);
        if(it == g_resident.end()) R HPA_E_DOMAIN;
        g_resident.erase(it);
    }
    unregister_device_range(p);
    hipDeviceSynchronize();
    hipFree(p);
    x->g = nullptr;
    R 0;
}

// Host copy of one device payload of `bytes`, repointing x->g at it.  A copy of an arena payload
// belongs to the context and dies with the first rewind below its creation (hpaSetHeapOffset);
// a copy of a resident payload (hpaUpload, generated tables) keeps the payload's handle and lives
// until hpaFreeResident of the node.
static int32_t tohost_payload(hpa_ctx *c, V x, L bytes){
    void *h = malloc((size_t)(bytes > 0 ? bytes : 1));
    if(!h){ set_error_detail("hpaToHost: malloc(%lld) failed", (long long)bytes); R HPA_E_MEMORY; }
    if(hipMemcpyAsync(h, x->g, (size_t)bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess){ free(h); R HPA_E_DEVICE; }
    if(hipStreamSynchronize(c->stream) != hipSuccess){ free(h); R HPA_E_DEVICE; }
    bool res;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        res = g_resident.count((uintptr_t)x->g) != 0;
        if(res) g_res_host[(uintptr_t)h] = (uintptr_t)x->g;
    }
    if(!res) c->host_bufs.push_back({h, c->host_seq++, c->dcur});
    x->g = (char*)h;
    R 0;
}
static int32_t tohost_rec(hpa_ctx *c, V x){
    if(is_basic(x->typ)){
        if(x->typ == HPA_S){ int32_t e = str_tohost(c); if(e) R e; }      // bytes the device wrote
        if(is_inline(x) || !is_device(x->g)) R 0;
        R tohost_payload(c, x, type_bytes(x->typ, x->len));
    }
    if(x->typ == HPA_G && is_flat(x)){
        // counts vector lives on the device too
        if(!is_inline(x) && is_device(x->g)){ int32_t e = tohost_payload(c, x, 8 * x->len); if(e) R e; }
        R tohost_rec(c, x->g2);
    }
    if(x->typ == HPA_G || x->typ == HPA_N || x->typ == HPA_A){
        L n = x->typ == HPA_G ? x->len : 2;
        for(L i=0;i<n;i++){ int32_t e = tohost_rec(c, ((V)x->g) + i); if(e) R e; }
        R 0;
    }
    R 0;
}
extern "C" int32_t hpaToHost(hpa_V x){ R tohost_rec(ctx_cur(), x); }

extern "C" int32_t hpaReadPayload(hpa_V x, void *dst, int64_t bytes){
    hpa_ctx *c = ctx_cur();
    if(bytes <= 0) R 0;
    if(is_inline(x) || !is_device(x->g)){ memcpy(dst, x->g, (size_t)bytes); R 0; }
    R d2h(c, dst, x->g, bytes);                 // small reads through the pinned staging buffer
}

// Copy the payloads of n vectors into caller memory with one stream synchronisation.
// Small device ranges copied into the pinned staging buffer by one kernel (the buffer is
// device-accessible host memory): one launch and one synchronisation for all of them, instead of
// one copy command each.  Workgroup k copies entry k, 8 bytes a thread when both ends allow.
#define RM_MAX 32
struct RmList { const char *src[RM_MAX]; char *dst[RM_MAX]; int64_t bytes[RM_MAX]; };
__global__ void __launch_bounds__(256) k_read_many(RmList r){
    const char *s = r.src[blockIdx.x]; char *d = r.dst[blockIdx.x]; const L nb = r.bytes[blockIdx.x];
    const bool w8 = ((uintptr_t)s % 8 == 0) && ((uintptr_t)d % 8 == 0);
    L i = (L)threadIdx.x * 8;
    if(w8) for(; i + 8 <= nb; i += 256 * 8) *(uint64_t*)(d + i) = *(const uint64_t*)(s + i);
    for(L j = (w8 ? (nb & ~(L)7) : 0) + threadIdx.x; j < nb; j += 256) d[j] = s[j];
}
// Device ranges src[k] (bytes[k] each, device pointers) into host memory dst[k], after the work
// queued before: through the pinned staging buffer with one kernel and one synchronisation when
// they fit, else one copy command each and one synchronisation.
int32_t read_staged(hpa_ctx *c, int n, const void *const *src, void *const *dst, const int64_t *bytes){
    L dev_total = 0;
    for(int k = 0; k < n; k++) if(bytes[k] > 0) dev_total += (bytes[k] + 15) & ~(L)15;
    if(dev_total == 0) R 0;
    if(c->mapped && pub_fits(n, bytes)) R pub_read(c, n, src, dst, bytes);
    if(c->pinned && dev_total <= c->pinned_bytes){
        L off = 0;
        std::vector<L> at(n, -1);
        RmList rl; int nr = 0;
        for(int k = 0; k < n; k++){
            if(bytes[k] <= 0) continue;
            rl.src[nr] = (const char*)src[k]; rl.dst[nr] = c->pinned + off; rl.bytes[nr] = bytes[k]; nr++;
            at[k] = off; off += (bytes[k] + 15) & ~(L)15;
            if(nr == RM_MAX){ hipLaunchKernelGGL(k_read_many, dim3(nr), dim3(256), 0, c->stream, rl); nr = 0; }
        }
        if(nr) hipLaunchKernelGGL(k_read_many, dim3(nr), dim3(256), 0, c->stream, rl);
        if(hipGetLastError() != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        for(int k = 0; k < n; k++) if(at[k] >= 0) memcpy(dst[k], c->pinned + at[k], (size_t)bytes[k]);
        R 0;
    }
    for(int k = 0; k < n; k++)
        if(bytes[k] > 0 && hipMemcpyAsync(dst[k], src[k], (size_t)bytes[k], hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
    if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
    R 0;
}
extern "C" int32_t hpaReadMany(int32_t n, hpa_V *x, void **dst, const int64_t *bytes){
    hpa_ctx *c = ctx_cur();
    // small results (the usual case: aggregates over groups) go through the pinned staging
    // buffer: one kernel fills it and one synchronisation covers every read
    std::vector<const void*> s; std::vector<void*> d; std::vector<int64_t> b;
    for(int k = 0; k < n; k++){
        if(bytes[k] <= 0) continue;
        if(is_inline(x[k]) || !is_device(x[k]->g)){ memcpy(dst[k], x[k]->g, (size_t)bytes[k]); continue; }
        s.push_back(x[k]->g); d.push_back(dst[k]); b.push_back(bytes[k]);
    }
    R read_staged(c, (int)s.size(), s.data(), d.data(), b.data());
}

extern "C" int32_t hpaSync(void){
    hpa_ctx *c = ctx_cur();
    R hipStreamSynchronize(c->stream) == hipSuccess ? 0 : HPA_E_DEVICE;
}

// ---------------------------------------------------------------- kernel timing
static bool g_prof = false;        // the library's launch scopes
static bool g_prof_user = false;   // caller scopes (hpaProfileBegin/End)
struct PendingEv { std::string name; hipEvent_t a, b; double bytes; };
static std::vector<PendingEv> g_pending;
struct Stat { int64_t count = 0; double ms = 0, bytes = 0; };
static std::map<std::string, Stat> g_stats;

bool prof_enabled(){ R g_prof; }
// events are recycled: creating one per launch would add its own cost to what is measured
static std::vector<hipEvent_t> g_evpool;
static hipEvent_t ev_get(){
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if(!g_evpool.empty()){ hipEvent_t e = g_evpool.back(); g_evpool.pop_back(); R e; }
    }
    hipEvent_t e = nullptr; hipEventCreate(&e); R e;
}
static void ev_put(hipEvent_t e){ std::lock_guard<std::mutex> lk(g_mu); g_evpool.push_back(e); }
ProfScope::ProfScope(const char *n, hipStream_t st, double algo_bytes, bool user) : name(n), s(st), on(user ? g_prof_user : g_prof), bytes(algo_bytes){
    if(!on) R;
    a = ev_get(); b = ev_get();
    hipEventRecord(a, s);
}
void ProfScope::stop(){
    if(!on || stopped) R;
    hipEventRecord(b, s);
    stopped = true;
}
ProfScope::~ProfScope(){
    if(!on) R;
    stop();
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back({name, a, b, bytes});
}
static void prof_resolve(){
    std::vector<PendingEv> p;
    { std::lock_guard<std::mutex> lk(g_mu); p.swap(g_pending); }
    for(auto &e : p){
        hipEventSynchronize(e.b);
        float ms = 0; hipEventElapsedTime(&ms, e.a, e.b);
        ev_put(e.a); ev_put(e.b);
        std::lock_guard<std::mutex> lk(g_mu);
        auto &s = g_stats[e.name]; s.count++; s.ms += ms; s.bytes += e.bytes;
    }
}
// 0 off, 1 every scope, 2 caller scopes only (no events inside the library's calls)
extern "C" void hpaProfileEnable(int32_t on){ g_prof = on == 1; g_prof_user = on == 1 || on == 2; }

// Caller-named scopes: HIP events on the current context's stream around any sequence of calls
// (a whole primitive, a query), reported beside the library's own launch scopes.
struct UserScope { std::string name; ProfScope *ps; };
static thread_local std::deque<UserScope> t_scopes;     // deque: names keep their address
extern "C" void hpaProfileBegin(const char *name, double algo_bytes){
    if(!g_prof_user) R;
    t_scopes.push_back({name, nullptr});
    t_scopes.back().ps = new ProfScope(t_scopes.back().name.c_str(), ctx_cur()->stream, algo_bytes, true);
}
extern "C" void hpaProfileEnd(void){
    if(t_scopes.empty()) R;
    delete t_scopes.back().ps;          // records the end event and queues the pair
    t_scopes.pop_back();
}
extern "C" void hpaProfileReset(void){ prof_resolve(); std::lock_guard<std::mutex> lk(g_mu); g_stats.clear(); }
extern "C" int32_t hpaProfileStats(const char *kernel, int64_t *count, double *total_ms){
    prof_resolve();
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_stats.find(kernel);
    if(it == g_stats.end()){ *count = 0; *total_ms = 0; R 1; }
    *count = it->second.count; *total_ms = it->second.ms; R 0;
}
extern "C" int32_t hpaProfileList(char *buf, int64_t cap){
    prof_resolve();
    std::lock_guard<std::mutex> lk(g_mu);
    std::string s;
    for(auto &kv : g_stats){ char line[256]; snprintf(line, sizeof(line), "%s %lld %.6f %.0f\n", kv.first.c_str(), (long long)kv.second.count, kv.second.ms, kv.second.bytes); s += line; }
    if((int64_t)s.size() + 1 > cap) R (int32_t)(s.size() + 1);
    memcpy(buf, s.c_str(), s.size() + 1);
    R 0;
}

// ---------------------------------------------------------------- device-wide exclusive scan
// Three-phase reduce-then-scan over int64: per-block sums, a single-block scan of the block
// sums, then per-block scans seeded with the block offset.  Deterministic and order-exact.
// Thread t of block b owns the SCAN_ITEMS consecutive elements from (b * SCAN_BLOCK + t) *
// SCAN_ITEMS, read (and, in the down pass, written) with 16-byte vector accesses when aligned,
// so a block needs one barrier and no dependent loads.
#define SCAN_BLOCK 256
#define SCAN_ITEMS 16
template<typename TI>
__device__ __forceinline__ void scan_load(const TI *in, L i0, L n, int vec, int64_t *v){
    if(vec && i0 + SCAN_ITEMS <= n){
        if(sizeof(TI) == 4){
            #pragma unroll
            for(int q = 0; q < SCAN_ITEMS / 4; q++){
                const uint4 w = *(const uint4*)(in + i0 + 4 * q);
                v[4 * q] = (int64_t)(TI)w.x; v[4 * q + 1] = (int64_t)(TI)w.y; v[4 * q + 2] = (int64_t)(TI)w.z; v[4 * q + 3] = (int64_t)(TI)w.w;
            }
        } else {
            #pragma unroll
            for(int q = 0; q < SCAN_ITEMS / 2; q++){
                const uint4 w = *(const uint4*)(in + i0 + 2 * q);
                v[2 * q] = (int64_t)(((uint64_t)w.y << 32) | w.x); v[2 * q + 1] = (int64_t)(((uint64_t)w.w << 32) | w.z);
            }
        }
    } else {
        #pragma unroll
        for(int k = 0; k < SCAN_ITEMS; k++) v[k] = i0 + k < n ? (int64_t)in[i0 + k] : 0;
    }
}
// exclusive scan of one value per thread over the block; *total = block sum
__device__ __forceinline__ int64_t scan_block_excl(int64_t x, int64_t *wsum, int64_t *total){
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t inc = x;
    for(int o = 1; o < 64; o <<= 1){ int64_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
    if(lane == 63) wsum[w] = inc;
    __syncthreads();
    int64_t pre = inc - x, tot = 0;
    for(int q = 0; q < SCAN_BLOCK / 64; q++){ pre += q < w ? wsum[q] : 0; tot += wsum[q]; }
    *total = tot;
    R pre;
}
template<typename TI>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const TI *in, L n, int vec, int64_t *bsum){
    __shared__ int64_t wsum[SCAN_BLOCK / 64];
    const L i0 = ((L)blockIdx.x * SCAN_BLOCK + threadIdx.x) * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    scan_load<TI>(in, i0, n, vec, v);
    int64_t s = 0;
    #pragma unroll
    for(int k = 0; k < SCAN_ITEMS; k++) s += v[k];
    for(int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0){ int64_t t = 0; for(int w = 0; w < SCAN_BLOCK / 64; w++) t += wsum[w]; bsum[blockIdx.x] = t; }
}
__global__ void __launch_bounds__(1024) k_scan_blocks(int64_t *bsum, L nb, int64_t *total){
    __shared__ int64_t wsum[16];
    __shared__ int64_t carry;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if(threadIdx.x == 0) carry = 0;
    __syncthreads();
    for(L base = 0; base < nb; base += 1024){
        const L i = base + threadIdx.x;
        const int64_t v = i < nb ? bsum[i] : 0;
        int64_t inc = v;
        for(int o = 1; o < 64; o <<= 1){ int64_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
        if(lane == 63) wsum[w] = inc;
        __syncthreads();
        int64_t pre = carry + inc - v, tot = 0;
        for(int q = 0; q < 16; q++){ pre += q < w ? wsum[q] : 0; tot += wsum[q]; }
        if(i < nb) bsum[i] = pre;
        __syncthreads();
        if(threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if(threadIdx.x == 0 && total) *total = carry;
}
template<typename TI>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_down(const TI *in, int64_t *out, L n, int vec, const int64_t *boff){
    __shared__ int64_t wsum[SCAN_BLOCK / 64];
    const L i0 = ((L)blockIdx.x * SCAN_BLOCK + threadIdx.x) * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    scan_load<TI>(in, i0, n, vec, v);
    int64_t s = 0;
    #pragma unroll
    for(int k = 0; k < SCAN_ITEMS; k++){ const int64_t t = v[k]; v[k] = s; s += t; }
    int64_t tot;
    const int64_t pre = boff[blockIdx.x] + scan_block_excl(s, wsum, &tot);
    if(vec && i0 + SCAN_ITEMS <= n){
        #pragma unroll
        for(int q = 0; q < SCAN_ITEMS / 2; q++){
            const uint64_t a = (uint64_t)(pre + v[2 * q]), b = (uint64_t)(pre + v[2 * q + 1]);
            *(uint4*)(out + i0 + 2 * q) = make_uint4((unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32));
        }
    } else {
        #pragma unroll
        for(int k = 0; k < SCAN_ITEMS; k++) if(i0 + k < n) out[i0 + k] = pre + v[k];
    }
}
// inputs of at most 1024 * SCAN_ITEMS elements: one workgroup, one launch.  The results pass
// through LDS (rows padded by one word against bank conflicts) so that consecutive threads write
// consecutive elements.
template<typename TI>
__global__ void __launch_bounds__(1024) k_scan_small(const TI *in, int64_t *out, L n, int vec, int64_t *total){
    __shared__ int64_t wsum[16];
    __shared__ int64_t so[1024 * (SCAN_ITEMS + 1)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const L i0 = (L)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    scan_load<TI>(in, i0, n, vec, v);
    int64_t s = 0;
    #pragma unroll
    for(int k = 0; k < SCAN_ITEMS; k++){ const int64_t t = v[k]; v[k] = s; s += t; }
    int64_t inc = s;
    for(int o = 1; o < 64; o <<= 1){ int64_t t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }
    if(lane == 63) wsum[w] = inc;
    __syncthreads();
    int64_t pre = inc - s;
    for(int q = 0; q < w; q++) pre += wsum[q];
    #pragma unroll
    for(int k = 0; k < SCAN_ITEMS; k++) so[threadIdx.x * (SCAN_ITEMS + 1) + k] = pre + v[k];
    if(threadIdx.x == 1023 && total) *total = pre + s;
    __syncthreads();
    for(L j = threadIdx.x; j < n; j += 1024) out[j] = so[(j / SCAN_ITEMS) * (SCAN_ITEMS + 1) + (j % SCAN_ITEMS)];
}
template<typename TI>
static int32_t scan_impl(hpa_ctx *c, const TI *in, int64_t *out, L n, int64_t *total_dev){
    if(n > 0 && n <= 1024 * SCAN_ITEMS){
        const int vec = ((uintptr_t)in % 16 == 0) ? 1 : 0;
        hipLaunchKernelGGL(k_scan_small<TI>, dim3(1), dim3(1024), 0, c->stream, in, out, n, vec, total_dev);
        R check_launch("scan");
    }
    L per = (L)SCAN_BLOCK * SCAN_ITEMS;
    L nb = ceil_div(n, per); if(nb < 1) nb = 1;
    int64_t *bsum = (int64_t*)dalloc(c, 8 * nb);
    if(!bsum) R HPA_E_MEMORY;
    const int vec = ((uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0) ? 1 : 0;
    if(n > 0) hipLaunchKernelGGL(k_scan_reduce<TI>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, c->stream, in, n, vec, bsum);
    else hipMemsetAsync(bsum, 0, 8, c->stream);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, c->stream, bsum, nb, total_dev);
    if(n > 0) hipLaunchKernelGGL(k_scan_down<TI>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, c->stream, in, out, n, vec, bsum);
    R check_launch("scan");
}
int32_t scan_i64(hpa_ctx *c, const int64_t *in, int64_t *out, L n, int64_t *total_dev){ R scan_impl<int64_t>(c, in, out, n, total_dev); }
int32_t scan_u32_to_i64(hpa_ctx *c, const uint32_t *in, int64_t *out, L n, int64_t *total_dev){ R scan_impl<uint32_t>(c, in, out, n, total_dev); }

__global__ void k_iota_u32(uint32_t *p, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) p[i] = (uint32_t)i;
}
int32_t iota_u32(hpa_ctx *c, uint32_t *p, L n){
    if(n <= 0) R 0;
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, p, n);
    R check_launch("iota");
}

// flat list (CSR) accessors: counts in vg(y) (i64, device), indices in vg2(y) (i64, device)
__global__ void k_dummy(){}
int32_t flat_offsets(hpa_ctx *c, V y, const int64_t **offs, const int64_t **idx, L *g, L *n){
    if(!is_flat(y)) R HPA_E_DOMAIN;
    *g = y->len; *n = y->g2->len;
    const void *pc = nullptr, *pi = nullptr;
    int32_t e;
    if(is_device(y->g)) pc = y->g;
    else {
        // host counts (a flat list built by the caller): stage 8*g bytes
        void *d = dalloc(c, 8 * (*g > 0 ? *g : 1)); if(!d) R HPA_E_MEMORY;
        if(*g > 0 && hipMemcpyAsync(d, y->g, (size_t)(8 * *g), hipMemcpyHostToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        pc = d;
    }
    V t = y->g2;
    // the inner vector is the i64 row list of a group result, or the gathered values of
    // each_right(@index) (any element type); only the former is returned as an index
    if(t->typ == HPA_L){ if((e = dev_view(c, t, &pi))) R e; }
    // a group made here kept its offsets: no scan of the counts
    if(const int64_t *go = group_offsets(c, pc, *g)){ *offs = go; *idx = (const int64_t*)pi; R 0; }
    int64_t *o = (int64_t*)dalloc(c, 8 * (*g + 1));
    if(!o) R HPA_E_MEMORY;
    if((e = scan_i64(c, (const int64_t*)pc, o, *g, o + *g))) R e;
    *offs = o; *idx = (const int64_t*)pi;
    R 0;
}
