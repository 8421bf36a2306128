// Enumerations and foreign keys: pfnEnum (h_primitive.cc:2198-2222), pfnValues / pfnKeys /
// pfnFetch of an enum (:190-225, copyEnumIndex h_system.cc:501-506), and the table-level
// pfnAddFKey (:2425-2466), pfnColumnValue (:283-313) and pfnLoadTable (:20-30) over tables the
// caller registers (registerTable, h_system.cc:58-62).
//
// An enum of values y over keys x is the index_of of y in x -- on the GPU index_of's direct,
// guarded or hashed tables -- checked so that every value is present (E_ENUM_INDEX otherwise),
// and kept as that i64 index vector in HBM.  Its bound (every index < len(x)) is recorded, so a
// later index(x-column, values(enum)) -- the foreign-key gather of the compiled plans
// (tests/hyper/online-hir/q3.hir:20-23) -- skips its bounds check.
#include "hpa_internal.h"
#include <mutex>

__global__ void __launch_bounds__(256) k_enum_check(const int64_t *ind, L n, int64_t nx, int *bad){
    bool b = false;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) b |= ind[i] >= nx;
    if(__ballot(b) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}
__global__ void k_enum_store(int64_t *p, int64_t v){ *p = v; }

extern "C" int32_t pfnEnum(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    int32_t e;
    if(!is_basic(x->typ)){
        if(x->typ == HPA_G && y->typ == HPA_G){ set_error_detail("enum over key lists"); R HPA_E_NOT_IMPL; }
        R HPA_E_DOMAIN;
    }
    V0 ind; memset(&ind, 0, sizeof(ind));
    if((e = pfnIndexOf(&ind, x, y))) R e;
    const L n = ind.len, nx = x->len;
    int64_t *payload = nullptr;
    if(is_inline(&ind)){
        if(n == 1 && ind.i64 >= nx) R HPA_E_ENUM_INDEX;
        payload = (int64_t*)dalloc(c, 8); if(!payload) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_enum_store, dim3(1), dim3(1), 0, c->stream, payload, n == 1 ? ind.i64 : (int64_t)0);
        if((e = check_launch("enum"))) R e;
    } else {
        const void *p; if((e = dev_view(c, &ind, &p))) R e;
        payload = (int64_t*)p;
        int *bad = (int*)dalloc(c, 16); if(!bad) R HPA_E_MEMORY;
        hipMemsetAsync(bad, 0, 4, c->stream);
        {
            ProfScope ps("enum_check", c->stream, 8.0 * n);
            hipLaunchKernelGGL(k_enum_check, dim3(grid_for(n, 256, 4)), dim3(256), 0, c->stream, (const int64_t*)payload, n, (int64_t)nx, bad);
            if((e = check_launch("enum"))) R e;
        }
        int h = 0; if((e = d2h(c, &h, bad, 4))) R e;
        if(h) R HPA_E_ENUM_INDEX;                       // some value is not among the keys
    }
    hpa_Y0 *yy = (hpa_Y0*)halloc(c, sizeof(hpa_Y0)); if(!yy) R HPA_E_MEMORY;
    yy->source = -1; yy->ek = x; yy->ev = y;            // initEnum(z, -1, x, y, index)
    z->typ = HPA_Y; z->len = n; z->g = (char*)payload; z->p = yy;
    note_bound(c, payload, nx);
    R 0;
}

// values of an enum: the index vector itself (copyEnumIndex aliases it); one value comes back
// inline, as initV keeps basic values of length one
int32_t enum_values(hpa_ctx *c, V z, V x){
    z->typ = HPA_L; z->len = x->len; z->g2 = nullptr;
    if(x->len <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        R x->len == 1 ? d2h(c, &z->b, x->g, 8) : 0;
    }
    z->g = x->g;
    R 0;
}
int32_t enum_keys(V z, V x){ *z = *((const hpa_Y0*)x->p)->ek; R 0; }

extern "C" int32_t pfnFetch(hpa_V z, hpa_V x){
    if(x->typ != HPA_Y || !x->p) R HPA_E_DOMAIN;
    *z = *((const hpa_Y0*)x->p)->ev;
    R 0;
}

// ---------------------------------------------------------------- tables
static std::mutex g_tab_mu;
static std::map<int64_t, V> g_tables;                 // listTable (h_system.cc:58-67)

extern "C" int32_t hpaRegisterTable(int64_t name, hpa_V table){
    if(!table || (table->typ != HPA_A && table->typ != HPA_K)) R HPA_E_DOMAIN;
    std::lock_guard<std::mutex> lk(g_tab_mu);
    g_tables[name] = table;
    R 0;
}
static V find_table(V x){                             // fetchTableByName (h_primitive.cc:257-262)
    if(x->typ != HPA_Q || x->len != 1) R nullptr;
    int64_t q = x->q;
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tables.find(q);
    R it == g_tables.end() ? nullptr : it->second;
}
extern "C" int32_t pfnLoadTable(hpa_V z, hpa_V x){
    if(x->typ != HPA_Q) R HPA_E_DOMAIN;
    V t = find_table(x);
    if(!t) R HPA_E_TABLE_NOT_FOUND;
    *z = *t;
    R 0;
}
static const int64_t *sym_payload(V s){ R is_inline(s) ? &s->q : (const int64_t*)s->g; }
// fetchFKeyByName (h_primitive.cc:264-281): the enum registered for exactly this column-name list
static V fkey_by_name(V t, V m){
    if(t->typ != HPA_A || !t->p || m->typ != HPA_Q) R nullptr;
    for(hpa_ListY *y = ((hpa_A0*)t->p)->fkey; y; y = y->next){
        V col = (V)y->cn;
        if(col->typ != m->typ || col->len != m->len) continue;
        const int64_t *a = sym_payload(m), *b = sym_payload(col);
        bool f = true;
        for(L i = 0; i < m->len && f; i++) f = a[i] == b[i];
        if(f) R (V)y->y;
    }
    R nullptr;
}
static L find_col(V t, int64_t cid){                  // findColFromTable (h_system.cc:157-166)
    V keys = (V)t->g;
    const int64_t *k = sym_payload(keys);
    for(L i = 0; i < keys->len; i++) if(k[i] == cid) R i;
    R -1;
}
// fetchColumnValue (h_primitive.cc:283-308); `fk`: the enum of a foreign-key column first
static int32_t column_value(hpa_ctx *c, V z, V x, V y, bool fk){
    if(!(x->typ == HPA_A || x->typ == HPA_K) || y->typ != HPA_Q) R HPA_E_DOMAIN;
    if(fk){ V e0 = fkey_by_name(x, y); if(e0){ *z = *e0; R 0; } }
    V vals = ((V)x->g) + 1;
    if(y->len == 1){
        L k = find_col(x, y->q);
        if(k < 0) R HPA_E_COL_NOT_FOUND;
        *z = ((V)vals->g)[k];
        R 0;
    }
    const int64_t *ids = sym_payload(y);
    V cells = (V)halloc(c, sizeof(V0) * (y->len > 0 ? y->len : 1)); if(!cells) R HPA_E_MEMORY;
    for(L i = 0; i < y->len; i++){
        L k = find_col(x, ids[i]);
        if(k < 0) R HPA_E_COL_NOT_FOUND;
        cells[i] = ((V)vals->g)[k];
    }
    z->typ = HPA_G; z->len = y->len; z->g = (char*)cells; z->g2 = nullptr;
    R 0;
}
extern "C" int32_t pfnColumnValue(hpa_V z, hpa_V x, hpa_V y){ R column_value(ctx_cur(), z, x, y, true); }

// pfnAddFKey (h_primitive.cc:2425-2466): the enum of table y's column yKey over table x's key
// column xKey, appended to y's foreign-key list (setFKey, h_system.cc:652-675).  The enum and
// the list entry live as long as the tables (the reference allocates them with NEW).
extern "C" int32_t pfnAddFKey(hpa_V x, hpa_V xKey, hpa_V y, hpa_V yKey){
    hpa_ctx *c = ctx_cur();
    V tx = find_table(x); if(!tx) R HPA_E_DOMAIN;      // CHECKE(tableX==NULL)
    V ty = find_table(y); if(!ty) R HPA_E_DOMAIN;
    if(!(xKey->typ == HPA_Q && yKey->typ == xKey->typ && yKey->len == xKey->len)) R HPA_E_DOMAIN;
    if(ty->typ != HPA_A || !ty->p) R HPA_E_DOMAIN;
    V0 xc, yc; memset(&xc, 0, sizeof(xc)); memset(&yc, 0, sizeof(yc));
    int32_t e;
    if((e = column_value(c, &xc, tx, xKey, false)) || (e = column_value(c, &yc, ty, yKey, false))) R e;
    // the key columns are the tables' own nodes (long-lived); the enum node and its hpa_Y0 are
    // made here, outside the arena
    V xcol = (V)malloc(sizeof(V0)), ycol = (V)malloc(sizeof(V0)), fkey = (V)malloc(sizeof(V0));
    hpa_ListY *n = (hpa_ListY*)malloc(sizeof(hpa_ListY));
    if(!xcol || !ycol || !fkey || !n){ free(xcol); free(ycol); free(fkey); free(n); R HPA_E_MEMORY; }
    *xcol = xc; *ycol = yc;
    if(is_inline(&xc)) xcol->g = (char*)&xcol->b;
    if(is_inline(&yc)) ycol->g = (char*)&ycol->b;
    memset(fkey, 0, sizeof(V0));
    const L mark = c->dcur;
    if((e = pfnEnum(fkey, xcol, ycol))){ free(xcol); free(ycol); free(fkey); free(n); R e; }
    // the index vector outlives arena rewinds: move it to a resident allocation
    if(fkey->len > 0){
        void *d = nullptr;
        if(hipMalloc(&d, (size_t)(8 * fkey->len)) != hipSuccess){ free(xcol); free(ycol); free(fkey); free(n); R HPA_E_MEMORY; }
        hipMemcpyAsync(d, fkey->g, (size_t)(8 * fkey->len), hipMemcpyDeviceToDevice, c->stream);
        hipStreamSynchronize(c->stream);
        register_device_range(d, 8 * fkey->len);
        fkey->g = (char*)d;
    }
    hpa_Y0 *yy = (hpa_Y0*)malloc(sizeof(hpa_Y0));
    if(!yy){ free(n); R HPA_E_MEMORY; }
    *yy = *(hpa_Y0*)fkey->p;
    fkey->p = yy;
    arena_set(c, mark);
    n->y = fkey; n->cn = yKey; n->next = nullptr;
    hpa_A0 *a = (hpa_A0*)ty->p;                      // insertFKey: append at the end
    if(!a->fkey) a->fkey = n;
    else { hpa_ListY *t = a->fkey; while(t->next) t = t->next; t->next = n; }
    R 0;
}
