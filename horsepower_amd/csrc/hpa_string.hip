// Symbols and strings: the symbol table (h_symbol.cc), like (h_primitive.cc:1669-1757 with the
// pattern translation genLikeString, h_system.cc:313-334), symbol and string compares
// (h_primitive.cc:1121-1160, compareSymbol h_symbol.cc:241-244) and sub_string
// (h_primitive.cc:2468-2489).
//
// Symbols (Q) are ids into one process-wide table.  String work on them runs over a device copy of
// the table (bytes, offsets, lengths and the dense rank of each string in strcmp order): like
// matches every table entry once and the column gathers the result; an order compare compares
// ranks, which agrees with compareSymbol's strcmp.
//
// Strings (S) keep the reference layout: a vector of 8-byte char* values holding HOST addresses.
// The bytes they point to are mirrored in HBM by address ranges (host [lo, hi) -> device copy), so
// a kernel translates each pointer by its range and reads the mirror, and hpaToHost only has to
// copy the pointer vector (and the bytes of ranges the device wrote).  Ranges come from
// hpaUpload of an S column (clusters of its strings, resident), from host S payloads staged for
// one call, and from sub_string's outputs (a host buffer and its device twin in the arena, written
// on the device, copied back by hpaToHost); the arena ones die with the rewind below them.
//
// like: the reference compiles the translated pattern with PCRE2 and matches it anchored.  The
// translation maps % to .*, _ to . and escapes . and \; every other byte is literal unless it is
// one of the regular-expression operators ^ $ [ | ( ) ? * + {, which the translation leaves
// active -- such patterns return HPA_E_NOT_IMPL here.  What remains is a wildcard match with
// PCRE2's default line rules: neither . nor .* matches a newline, and $ matches at the end or
// before a final newline.
#include "hpa_internal.h"
#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>

// ---------------------------------------------------------------- symbol table
static std::mutex g_sym_mu;
static std::vector<std::string> g_syms;
static std::unordered_map<std::string, int64_t> g_sym_ids;
static uint64_t g_sym_version = 1;              // bumped by every change of the table

extern "C" int64_t hpaSymbol(const char *name){
    if(!name) R -1;
    std::lock_guard<std::mutex> lk(g_sym_mu);
    auto it = g_sym_ids.find(name);
    if(it != g_sym_ids.end()) R it->second;
    g_syms.push_back(name);
    g_sym_ids[g_syms.back()] = (int64_t)g_syms.size() - 1;
    g_sym_version++;
    R (int64_t)g_syms.size() - 1;
}
extern "C" const char *hpaSymbolName(int64_t id){
    std::lock_guard<std::mutex> lk(g_sym_mu);
    if(id < 0 || id >= (int64_t)g_syms.size()) R nullptr;
    R g_syms[id].c_str();
}
// The caller's symbol table becomes the library's: names[i] gets id i (the interpreter passes its
// symTable, h_symbol.cc:148-157, so ids agree on both sides).  Names the library had interned that
// are not among them are appended after.
extern "C" int32_t hpaImportSymbols(int64_t n, const char *const *names){
    if(n < 0 || (n > 0 && !names)) R HPA_E_DOMAIN;
    for(int64_t i = 0; i < n; i++) if(!names[i]) R HPA_E_NULL_VALUE;
    std::lock_guard<std::mutex> lk(g_sym_mu);
    std::vector<std::string> old; old.swap(g_syms);
    g_sym_ids.clear();
    for(int64_t i = 0; i < n; i++){
        g_syms.push_back(names[i]);
        g_sym_ids.emplace(g_syms.back(), i);       // a repeated name keeps its first id
    }
    for(auto &s : old) if(!g_sym_ids.count(s)){ g_syms.push_back(s); g_sym_ids[s] = (int64_t)g_syms.size() - 1; }
    g_sym_version++;
    R 0;
}

// device copy of the table
struct SymDict { int64_t n; const uint8_t *bytes; const int64_t *off; const int32_t *len; const int32_t *rank; };
static SymDict g_dict = {0, nullptr, nullptr, nullptr, nullptr};
static void *g_dict_mem = nullptr;
static uint64_t g_dict_version = 0;

static int32_t sym_dict(hpa_ctx *c, SymDict *out){
    std::lock_guard<std::mutex> lk(g_sym_mu);
    if(g_dict_version != g_sym_version){
        const int64_t n = (int64_t)g_syms.size();
        int64_t nb = 0; for(auto &s : g_syms) nb += (int64_t)s.size() + 1;
        const int64_t a8 = (nb + 15) & ~(int64_t)15;
        const size_t tot = (size_t)(a8 + 8 * (n + 1) + 4 * (n + 1) + 4 * (n + 1) + 64);
        std::vector<char> h(tot, 0);
        int64_t *off = (int64_t*)(h.data() + a8);
        int32_t *len = (int32_t*)(off + n + 1), *rank = len + n + 1;
        int64_t p = 0;
        for(int64_t i = 0; i < n; i++){
            off[i] = p; len[i] = (int32_t)g_syms[i].size();
            memcpy(h.data() + p, g_syms[i].c_str(), g_syms[i].size() + 1);
            p += (int64_t)g_syms[i].size() + 1;
        }
        std::vector<int64_t> ord(n);
        for(int64_t i = 0; i < n; i++) ord[i] = i;
        std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b){ R strcmp(g_syms[a].c_str(), g_syms[b].c_str()) < 0; });
        int32_t r = 0;                              // dense rank: equal strings share one
        for(int64_t k = 0; k < n; k++){
            if(k > 0 && strcmp(g_syms[ord[k]].c_str(), g_syms[ord[k - 1]].c_str()) != 0) r++;
            rank[ord[k]] = r;
        }
        void *d = nullptr;
        if(hipMalloc(&d, tot) != hipSuccess){ set_error_detail("symbol table: hipMalloc(%zu) failed", tot); R HPA_E_MEMORY; }
        if(hipMemcpy(d, h.data(), tot, hipMemcpyHostToDevice) != hipSuccess){ hipFree(d); R HPA_E_DEVICE; }
        if(g_dict_mem){ hipDeviceSynchronize(); hipFree(g_dict_mem); }
        g_dict_mem = d;
        char *b = (char*)d;
        g_dict = {n, (const uint8_t*)b, (const int64_t*)(b + a8), (const int32_t*)(b + a8 + 8 * (n + 1)),
                  (const int32_t*)(b + a8 + 8 * (n + 1) + 4 * (n + 1))};
        g_dict_version = g_sym_version;
    }
    (void)c;
    *out = g_dict;
    R 0;
}

// ---------------------------------------------------------------- string ranges
#define SM_MAX (1 << 16)
// the ranges as the kernels see them: sorted starts, ends and device-minus-host offsets in HBM
struct StrMap { int64_t n; const uint64_t *lo, *hi; const int64_t *delta; };
struct StrRange { uintptr_t lo, hi; char *dev; bool dirty; hpa_ctx *owner; };
static std::mutex g_str_mu;
static std::vector<StrRange> g_str;             // sorted by lo, disjoint
static uint64_t g_str_version = 1;              // bumped by every change of g_str

// the range holding address p, or null (g_str is sorted by lo and disjoint)
static StrRange *range_of(uintptr_t p){
    auto it = std::upper_bound(g_str.begin(), g_str.end(), p, [](uintptr_t v, const StrRange &a){ R v < a.lo; });
    if(it == g_str.begin()) R nullptr;
    --it;
    R p < it->hi ? &*it : nullptr;
}
static void add_range(const StrRange &r){
    g_str_version++;
    auto it = std::lower_bound(g_str.begin(), g_str.end(), r.lo, [](const StrRange &a, uintptr_t v){ R a.lo < v; });
    g_str.insert(it, r);
}
// host copies of the ranges the device wrote (their host bytes are stale until then)
static int32_t flush_dirty(hpa_ctx *c){
    for(auto &r : g_str){
        if(!r.dirty) continue;
        if(hipMemcpyAsync((void*)r.lo, r.dev, (size_t)(r.hi - r.lo), hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        r.dirty = false;
    }
    R 0;
}
// the ranges of host S payload pointers p[0..n) not mirrored yet: clusters of strings less than a
// page apart (the bytes between two mapped strings that close are mapped too), each copied to
// HBM -- resident (hipMalloc) for an upload, in the arena (owner c) for a payload staged for one
// call.  A resident mirror replaces the arena mirrors it overlaps (they die with a rewind).
static int32_t mirror_strings(hpa_ctx *c, const char *const *p, L n, bool resident){
    std::lock_guard<std::mutex> lk(g_str_mu);
    int32_t e;
    if((e = flush_dirty(c))) R e;
    std::vector<std::pair<uintptr_t, uintptr_t>> iv;
    for(L i = 0; i < n; i++){
        if(!p[i]) R HPA_E_NULL_VALUE;
        const uintptr_t a = (uintptr_t)p[i];
        const StrRange *r = range_of(a);
        if(r && (!resident || !r->owner)) continue;          // mirrored already
        iv.push_back({a, a + strlen(p[i]) + 1});
    }
    std::sort(iv.begin(), iv.end());
    std::vector<std::pair<uintptr_t, uintptr_t>> cl;
    for(auto &v : iv){
        if(!cl.empty() && v.first < cl.back().second + 4096){ if(v.second > cl.back().second) cl.back().second = v.second; }
        else cl.push_back(v);
    }
    for(auto &k : cl){
        size_t w = 0;
        for(auto &r : g_str){
            const bool over = k.first < r.hi && r.lo < k.second;
            if(over && (!resident || !r.owner)){ set_error_detail("string range overlaps a mirrored one"); R HPA_E_NOT_IMPL; }
            if(!over) g_str[w++] = r;                           // an overlapped arena mirror goes
        }
        if(w != g_str.size()) g_str_version++;
        g_str.resize(w);
        if(g_str.size() >= SM_MAX){ set_error_detail("strings in more than %d address ranges", SM_MAX); R HPA_E_NOT_IMPL; }
        const size_t b = (size_t)(k.second - k.first);
        char *d = nullptr;
        if(resident){
            if(hipMalloc((void**)&d, b) != hipSuccess){ set_error_detail("hipMalloc(%zu) failed", b); R HPA_E_MEMORY; }
            if(hipMemcpy(d, (const void*)k.first, b, hipMemcpyHostToDevice) != hipSuccess){ hipFree(d); R HPA_E_DEVICE; }
        } else {
            d = (char*)dalloc(c, (L)b); if(!d) R HPA_E_MEMORY;
            if(hipMemcpyAsync(d, (const void*)k.first, b, hipMemcpyHostToDevice, c->stream) != hipSuccess) R HPA_E_DEVICE;
            if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        }
        add_range({k.first, k.second, d, false, resident ? nullptr : c});
    }
    R 0;
}
// hpaUpload of an S column: its strings are mirrored (resident) before the pointers move to HBM
int32_t str_upload(hpa_ctx *c, V x){
    if(x->len <= 0) R 0;
    R mirror_strings(c, (const char *const *)x->g, x->len, true);
}
// the arena rewound to `cur`: ranges whose device copy lay above it are gone
void str_prune(hpa_ctx *c, L cur){
    std::lock_guard<std::mutex> lk(g_str_mu);
    size_t k = 0;
    for(auto &r : g_str){
        if(r.owner == c && r.dev >= c->dbase + cur) continue;
        g_str[k++] = r;
    }
    if(k != g_str.size()) g_str_version++;
    g_str.resize(k);
}
// hpaToHost of an S vector: the bytes of the ranges the device wrote are copied back
int32_t str_tohost(hpa_ctx *c){
    std::lock_guard<std::mutex> lk(g_str_mu);
    R flush_dirty(c);
}
static uint64_t g_map_version = 0;
static uint64_t *g_map_mem = nullptr;
static int64_t g_map_cap = 0;
// the device copy of the range table, rewritten when the ranges changed (after the device is
// idle: a kernel in flight may be reading the old one)
static int32_t str_map(StrMap *m){
    std::lock_guard<std::mutex> lk(g_str_mu);
    const int64_t n = (int64_t)g_str.size();
    if(g_map_version != g_str_version){
        hipDeviceSynchronize();
        if(n > g_map_cap){
            if(g_map_mem) hipFree(g_map_mem);
            g_map_cap = n + 64;
            if(hipMalloc((void**)&g_map_mem, (size_t)(24 * g_map_cap)) != hipSuccess){ g_map_mem = nullptr; g_map_cap = 0; R HPA_E_MEMORY; }
        }
        std::vector<uint64_t> h((size_t)(3 * (n > 0 ? n : 1)), 0);
        for(int64_t k = 0; k < n; k++){
            h[k] = g_str[k].lo; h[n + k] = g_str[k].hi;
            h[2 * n + k] = (uint64_t)((uintptr_t)g_str[k].dev - g_str[k].lo);
        }
        if(n > 0 && hipMemcpy(g_map_mem, h.data(), (size_t)(24 * n), hipMemcpyHostToDevice) != hipSuccess) R HPA_E_DEVICE;
        g_map_version = g_str_version;
    }
    m->n = n; m->lo = g_map_mem; m->hi = g_map_mem + n; m->delta = (const int64_t*)(g_map_mem + 2 * n);
    R 0;
}
// device view of an S node's pointer vector, its strings mirrored
static int32_t str_view(hpa_ctx *c, V x, const uint64_t **p){
    int32_t e;
    if(x->len <= 0){ *p = nullptr; R 0; }
    if(is_inline(x) || !is_device(x->g)){
        if((e = mirror_strings(c, (const char *const *)x->g, x->len, false))) R e;
    }
    const void *d; if((e = dev_view(c, x, &d))) R e;
    *p = (const uint64_t*)d;
    R 0;
}

// a string's mirror and length (up to its NUL, or the end of its range); null when the pointer is
// in no range
__device__ __forceinline__ const uint8_t *str_dev(const StrMap &m, uint64_t p, int64_t *len){
    int64_t a = 0, b = m.n;                      // the last range starting at or below p
    while(a < b){ const int64_t mid = (a + b) >> 1; if(m.lo[mid] <= p) a = mid + 1; else b = mid; }
    if(a == 0) return nullptr;
    const int64_t k = a - 1;
    if(p >= m.hi[k]) return nullptr;
    const uint8_t *s = (const uint8_t*)(p + (uint64_t)m.delta[k]);
    const int64_t lim = (int64_t)(m.hi[k] - p);
    int64_t l = 0;
    while(l < lim && s[l]) l++;
    *len = l;
    return s;
}

// ---------------------------------------------------------------- like
#define LK_MAX 256
#define LK_ANY 256
#define LK_STAR 257
struct LikePat { int n; short t[LK_MAX]; };

// the translated pattern as literal bytes, any-byte (_) and any-run (%) tokens
static int32_t like_compile(const uint8_t *s, L n, LikePat *p){
    if(n <= 0) R HPA_E_NULL_VALUE;                // genLikeString returns NULL, pfnLike E_NULL_VALUE
    if(n > LK_MAX){ set_error_detail("like pattern longer than %d bytes", LK_MAX); R HPA_E_NOT_IMPL; }
    p->n = 0;
    for(L i = 0; i < n; i++){
        const uint8_t ch = s[i];
        if(ch == '%') p->t[p->n++] = LK_STAR;
        else if(ch == '_') p->t[p->n++] = LK_ANY;
        else if(strchr("^$[|()?*+{", ch) && ch){ set_error_detail("like pattern with the regex operator '%c'", ch); R HPA_E_NOT_IMPL; }
        else p->t[p->n++] = ch;
    }
    R 0;
}
// anchored wildcard match of s[0, len): a run may not cover a newline, nor may _ match one.
// Only the last run is ever widened: every newline must meet a literal newline of the pattern,
// so a failed widening over one cannot be rescued by an earlier run.
__device__ bool like_span(const LikePat &p, const uint8_t *s, int64_t len){
    int64_t si = 0, mark = 0; int pi = 0, star = -1;
    while(si < len){
        const int t = pi < p.n ? p.t[pi] : -1;
        if(t >= 0 && t < 256 && s[si] == (uint8_t)t){ si++; pi++; }
        else if(t == LK_ANY && s[si] != '\n'){ si++; pi++; }
        else if(t == LK_STAR){ star = pi++; mark = si; }
        else if(star >= 0 && s[mark] != '\n'){ pi = star + 1; si = ++mark; }
        else return false;
    }
    while(pi < p.n && p.t[pi] == LK_STAR) pi++;
    return pi == p.n;
}
__device__ __forceinline__ bool like_match(const LikePat &p, const uint8_t *s, int64_t len){
    return like_span(p, s, len) || (len > 0 && s[len - 1] == '\n' && like_span(p, s, len - 1));
}

__global__ void k_like_dict(SymDict d, LikePat p, uint8_t *m){
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * blockDim.x)
        m[i] = like_match(p, d.bytes + d.off[i], d.len[i]) ? 1 : 0;
}
__global__ void k_like_sym(const int64_t *q, L n, const uint8_t *m, int64_t nd, uint8_t *z, int *bad){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const int64_t v = q[i];
        if(v < 0 || v >= nd){ atomicOr(bad, 1); z[i] = 0; }
        else z[i] = m[v];
    }
}
__global__ void k_like_str(const uint64_t *ps, L n, StrMap sm, LikePat p, uint8_t *z, int *bad){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        int64_t len; const uint8_t *s = str_dev(sm, ps[i], &len);
        if(!s){ atomicOr(bad, 1); z[i] = 0; }
        else z[i] = like_match(p, s, len) ? 1 : 0;
    }
}
// a char vector taken as one string: up to its first NUL (the reference's strlen)
__global__ void k_like_chars(const uint8_t *s, L n, LikePat p, uint8_t *z){
    int64_t len = 0;
    while(len < n && s[len]) len++;
    z[0] = like_match(p, s, len) ? 1 : 0;
}

static int32_t flag_check(hpa_ctx *c, int *bad, int32_t code){
    int h = 0; int32_t e = d2h(c, &h, bad, 4);
    R e ? e : (h ? code : 0);
}

// pfnLike (h_primitive.cc:1670-1757), one pattern: x is a C, Q or S value, y one C/Q/S pattern
extern "C" int32_t pfnLike(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    auto strish = [](int32_t t){ R t == HPA_C || t == HPA_Q || t == HPA_S; };
    if(!strish(x->typ) || !strish(y->typ)) R HPA_E_DOMAIN;
    if(!(y->len == 1 || y->typ == HPA_C)){
        // one pattern per row (getLikeFromString): the reference's own branch is unfinished
        set_error_detail("like with a pattern per row");
        R HPA_E_NOT_IMPL;
    }
    int32_t e;
    // the pattern's bytes on the host
    std::string pat;
    if(y->typ == HPA_C){
        std::vector<char> b((size_t)y->len + 1, 0);
        if(y->len > 0 && (e = read_scalar(c, y, b.data(), y->len))) R e;
        pat.assign(b.data());                         // sC(y): up to the first NUL
    } else if(y->typ == HPA_Q){
        int64_t q; if((e = read_scalar(c, y, &q, 8))) R e;
        const char *s = hpaSymbolName(q); if(!s) R HPA_E_DOMAIN;
        pat = s;
    } else {
        // the pattern's host bytes (written back first if the device made them)
        uint64_t ptr; if((e = read_scalar(c, y, &ptr, 8))) R e;
        if((e = str_tohost(c))) R e;
        if(!ptr) R HPA_E_NULL_VALUE;
        pat = (const char*)ptr;
    }
    LikePat lp;
    if((e = like_compile((const uint8_t*)pat.data(), (L)pat.size(), &lp))) R e;
    const L n = x->typ == HPA_C ? 1 : x->len;
    uint8_t *dz = n > 1 ? (uint8_t*)dalloc(c, n) : nullptr;
    uint8_t one[16];
    int *bad = (int*)dalloc(c, 16);
    if((n > 1 && !dz) || !bad) R HPA_E_MEMORY;
    hipMemsetAsync(bad, 0, 4, c->stream);
    uint8_t *out = dz ? dz : (uint8_t*)dalloc(c, 16);
    if(!out) R HPA_E_MEMORY;
    ProfScope ps("like", c->stream, (double)n * 9);
    if(x->typ == HPA_C){
        const void *p = nullptr;
        if(x->len > 0 && (e = dev_view(c, x, &p))) R e;
        hipLaunchKernelGGL(k_like_chars, dim3(1), dim3(1), 0, c->stream, (const uint8_t*)p, x->len, lp, out);
    } else if(x->typ == HPA_Q && n > 0){
        SymDict d; if((e = sym_dict(c, &d))) R e;
        uint8_t *m = (uint8_t*)dalloc(c, d.n > 0 ? d.n : 1); if(!m) R HPA_E_MEMORY;
        if(d.n > 0) hipLaunchKernelGGL(k_like_dict, dim3(grid_for(d.n, 256, 1)), dim3(256), 0, c->stream, d, lp, m);
        const void *q; if((e = dev_view(c, x, &q))) R e;
        hipLaunchKernelGGL(k_like_sym, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)q, n, (const uint8_t*)m, d.n, out, bad);
    } else if(n > 0){
        const uint64_t *ps_; if((e = str_view(c, x, &ps_))) R e;
        StrMap sm; if((e = str_map(&sm))) R e;
        hipLaunchKernelGGL(k_like_str, dim3(grid_for(n, 256, 1)), dim3(256), 0, c->stream, ps_, n, sm, lp, out, bad);
    }
    if((e = check_launch("like"))) R e;
    if((e = flag_check(c, bad, HPA_E_DOMAIN))) R e;    // a symbol id or string pointer unknown here
    z->typ = HPA_B; z->len = n; z->g2 = nullptr;
    if(n <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(n == 1){ if((e = d2h(c, one, out, 1))) R e; z->b = one[0]; }
    } else z->g = (char*)dz;
    R 0;
}

// ---------------------------------------------------------------- compares
__device__ __forceinline__ bool comp_op(int op, int cmp){
    switch(op){
        case 0: return cmp < 0; case 1: return cmp <= 0; case 2: return cmp > 0; case 3: return cmp >= 0;
        case 4: return cmp == 0; default: return cmp != 0;
    }
}
// compareSymbol (h_symbol.cc:241-244): equal ids compare equal, others by strcmp -- the ranks
__global__ void k_cmp_sym(const int64_t *x, L sx, const int64_t *y, L sy, L n, const int32_t *rank, int64_t nd, int op,
                          uint8_t *z, int *bad){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const int64_t a = x[i * sx], b = y[i * sy];
        int cmp = 0;
        if(a != b){
            if(a < 0 || a >= nd || b < 0 || b >= nd){ atomicOr(bad, 1); z[i] = 0; continue; }
            const int32_t ra = rank[a], rb = rank[b];
            cmp = ra < rb ? -1 : ra > rb ? 1 : 0;
        }
        z[i] = comp_op(op, cmp) ? 1 : 0;
    }
}
__device__ __forceinline__ bool dev_streq(const uint8_t *a, int64_t la, const uint8_t *b, int64_t lb){
    if(la != lb) return false;
    for(int64_t k = 0; k < la; k++) if(a[k] != b[k]) return false;
    return true;
}
// the order ops of the compares of two symbol operands (h_primitive.cc:1121-1160; eq and neq
// compare the ids directly); lengths are checked by the caller (one side of length 1, or equal)
int32_t compare_sym(hpa_ctx *c, V z, V x, V y, int op){
    int32_t e;
    const L n = x->len == 1 ? y->len : x->len;
    const L sx = x->len == 1 ? 0 : 1, sy = y->len == 1 ? 0 : 1;
    uint8_t *dz = (uint8_t*)dalloc(c, n > 0 ? n : 1);
    int *bad = (int*)dalloc(c, 16);
    if(!dz || !bad) R HPA_E_MEMORY;
    hipMemsetAsync(bad, 0, 4, c->stream);
    if(n > 0){
        ProfScope ps("compare_sym", c->stream, (double)n * 17);
        SymDict d; if((e = sym_dict(c, &d))) R e;
        const void *px, *py;
        if((e = dev_view(c, x, &px)) || (e = dev_view(c, y, &py))) R e;
        hipLaunchKernelGGL(k_cmp_sym, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)px, sx,
                           (const int64_t*)py, sy, n, d.rank, d.n, op, dz, bad);
        if((e = check_launch("compare_sym"))) R e;
    }
    if((e = flag_check(c, bad, HPA_E_DOMAIN))) R e;
    z->typ = HPA_B; z->len = n; z->g2 = nullptr;
    if(n <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(n == 1){ uint8_t v; if((e = d2h(c, &v, dz, 1))) R e; z->b = v; }
    } else z->g = (char*)dz;
    R 0;
}

// ---------------------------------------------------------------- member of strings
// lib_member_S (h_libs.cc:1015-1018): a string hash table of the set x (djb2 of the bytes, as
// hash_S, h_hash.h:25, open addressing over slots holding row + 1), probed by every y[j]; equal
// means equal bytes
__device__ __forceinline__ uint32_t str_hash(const uint8_t *s, int64_t l){
    uint32_t h = 5381;
    for(int64_t k = 0; k < l; k++) h = h * 33u + s[k];
    return h * 2654435761u;
}
__global__ void k_mem_str_build(const uint64_t *x, L nx, StrMap sm, uint32_t *tab, uint32_t mask, int *bad){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += (L)gridDim.x * blockDim.x){
        int64_t l; const uint8_t *s = str_dev(sm, x[i], &l);
        if(!s){ atomicOr(bad, 1); continue; }
        uint32_t h = str_hash(s, l) & mask;
        for(uint32_t probe = 0; probe <= mask; probe++){
            if(atomicCAS(&tab[h], 0u, (uint32_t)(i + 1)) == 0u) break;
            h = (h + 1) & mask;
        }
    }
}
__global__ void k_mem_str_probe(const uint64_t *x, const uint64_t *y, L ny, StrMap sm, const uint32_t *tab, uint32_t mask,
                                uint8_t *z, int *bad){
    for(L j = (L)blockIdx.x * blockDim.x + threadIdx.x; j < ny; j += (L)gridDim.x * blockDim.x){
        int64_t l; const uint8_t *s = str_dev(sm, y[j], &l);
        if(!s){ atomicOr(bad, 1); z[j] = 0; continue; }
        uint32_t h = str_hash(s, l) & mask;
        bool f = false;
        for(uint32_t probe = 0; probe <= mask && !f; probe++){
            const uint32_t v = tab[h];
            if(v == 0u) break;
            int64_t lx; const uint8_t *t = str_dev(sm, x[v - 1], &lx);
            f = t && dev_streq(s, l, t, lx);
            h = (h + 1) & mask;
        }
        z[j] = f ? 1 : 0;
    }
}
// pfnMember(z, y, x) of two string operands: z[j] = some x[i] equals y[j] (left shape: y)
int32_t member_str(hpa_ctx *c, V z, V y, V x){
    int32_t e;
    const L ny = y->len, nx = x->len;
    uint8_t *dz = (uint8_t*)dalloc(c, ny > 0 ? ny : 1);
    int *bad = (int*)dalloc(c, 16);
    L slots = 16; while(slots < 2 * nx) slots <<= 1;
    if(slots > (1LL << 31)) R HPA_E_NOT_IMPL;
    uint32_t *tab = (uint32_t*)dalloc(c, 4 * slots);
    if(!dz || !bad || !tab) R HPA_E_MEMORY;
    hipMemsetAsync(bad, 0, 4, c->stream);
    if(ny > 0){
        hipMemsetAsync(dz, 0, (size_t)ny, c->stream);
        if(nx > 0){
            const uint64_t *px, *py;
            if((e = str_view(c, x, &px)) || (e = str_view(c, y, &py))) R e;
            StrMap sm; if((e = str_map(&sm))) R e;
            ProfScope ps("member_str", c->stream, 8.0 * (double)(nx + ny) + (double)ny);
            hipMemsetAsync(tab, 0, (size_t)(4 * slots), c->stream);
            hipLaunchKernelGGL(k_mem_str_build, dim3(grid_for(nx, 256, 1)), dim3(256), 0, c->stream, px, nx, sm, tab, (uint32_t)(slots - 1), bad);
            hipLaunchKernelGGL(k_mem_str_probe, dim3(grid_for(ny, 256, 1)), dim3(256), 0, c->stream, px, py, ny, sm,
                               (const uint32_t*)tab, (uint32_t)(slots - 1), dz, bad);
            if((e = check_launch("member_str"))) R e;
        }
    }
    if((e = flag_check(c, bad, HPA_E_DOMAIN))) R e;
    z->typ = HPA_B; z->len = ny; z->g2 = nullptr;
    if(ny <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(ny == 1){ uint8_t v; if((e = d2h(c, &v, dz, 1))) R e; z->b = v; }
    } else z->g = (char*)dz;
    R 0;
}

// ---------------------------------------------------------------- sub_string
// row i: seg bytes from position start (1-based) of x[i] -- min(seg, strlen - start) of them, as
// the reference counts the room (h_primitive.cc:2477-2478), none when that is negative -- padded
// with spaces to seg, NUL-terminated, at out + i * (seg + 1)
__global__ void k_substr(const uint64_t *x, L n, StrMap sm, int64_t start, int64_t seg, uint8_t *out, uint64_t hbase,
                         uint64_t *z, int *bad){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        uint8_t *p = out + i * (seg + 1);
        z[i] = hbase + (uint64_t)(i * (seg + 1));
        int64_t t0; const uint8_t *t = str_dev(sm, x[i], &t0);
        if(!t){ atomicOr(bad, 1); p[0] = 0; continue; }
        const int64_t space = t0 - start;
        int64_t len = seg < space ? seg : space;
        if(len < 0) len = 0;
        for(int64_t j = 0; j < len; j++) p[j] = t[start + j - 1];
        for(int64_t j = len; j < seg; j++) p[j] = ' ';
        p[seg] = 0;
    }
}
static int64_t int_at(const uint8_t *b, int32_t typ, int k){
    switch(typ){
        case HPA_B: case HPA_J: R typ == HPA_J ? (int64_t)((const int8_t*)b)[k] : (int64_t)b[k];
        case HPA_H: R ((const int16_t*)b)[k];
        case HPA_I: R ((const int32_t*)b)[k];
        default: R ((const int64_t*)b)[k];
    }
}
// pfnSubString (h_primitive.cc:2468-2489): x a string vector, y (start, length) integers.  A start
// below 1 reads before the string in the reference; it is HPA_E_DOMAIN here.
extern "C" int32_t pfnSubString(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(x->typ != HPA_S || !(y->typ == HPA_B || y->typ == HPA_J || y->typ == HPA_H || y->typ == HPA_I || y->typ == HPA_L)) R HPA_E_DOMAIN;
    if(y->len != 2) R HPA_E_NOT_IMPL;
    int32_t e;
    uint8_t yb[16];
    if((e = read_scalar(c, y, yb, type_bytes(y->typ, 2)))) R e;
    const int64_t start = int_at(yb, y->typ, 0), seg = int_at(yb, y->typ, 1);
    if(start < 1 || seg < 0){ set_error_detail("sub_string(start %lld, length %lld)", (long long)start, (long long)seg); R HPA_E_DOMAIN; }
    const L n = x->len;
    const L bytes = n * (seg + 1);
    // the strings' host home (what the pointers hold) and their device twin
    char *h = (char*)malloc((size_t)(bytes > 0 ? bytes : 1));
    if(!h) R HPA_E_MEMORY;
    c->host_bufs.push_back({h, c->host_seq++, c->dcur});
    uint8_t *dv = (uint8_t*)dalloc(c, bytes > 0 ? bytes : 1);
    uint64_t *dz = (uint64_t*)dalloc(c, 8 * (n > 0 ? n : 1));
    int *bad = (int*)dalloc(c, 16);
    if(!dv || !dz || !bad) R HPA_E_MEMORY;
    hipMemsetAsync(bad, 0, 4, c->stream);
    if(n > 0){
        const uint64_t *px; if((e = str_view(c, x, &px))) R e;
        StrMap sm; if((e = str_map(&sm))) R e;
        ProfScope ps("sub_string", c->stream, (double)n * 8 + (double)bytes);
        hipLaunchKernelGGL(k_substr, dim3(grid_for(n, 256, 1)), dim3(256), 0, c->stream, px, n, sm, start, seg, dv,
                           (uint64_t)(uintptr_t)h, dz, bad);
        if((e = check_launch("sub_string"))) R e;
    }
    if((e = flag_check(c, bad, HPA_E_DOMAIN))) R e;
    {
        std::lock_guard<std::mutex> lk(g_str_mu);
        if(g_str.size() >= SM_MAX){ set_error_detail("strings in more than %d address ranges", SM_MAX); R HPA_E_NOT_IMPL; }
        add_range({(uintptr_t)h, (uintptr_t)h + (uintptr_t)(bytes > 0 ? bytes : 1), (char*)dv, true, c});
    }
    z->typ = HPA_S; z->len = n; z->g2 = nullptr;
    if(n <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(n == 1 && (e = d2h(c, &z->b, dz, 8))) R e;
    } else z->g = (char*)dz;
    R 0;
}
