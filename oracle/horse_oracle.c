/*
 * horse_oracle.c -- TEST INFRASTRUCTURE ONLY (see horse_oracle.h).
 *
 * Clean-room C restatement of the HorsePower v3 primitive library.  Citations are to
 * src/horseir/v3/backend/ unless another path is given.  Parallel parts use OpenMP with the
 * reference's schedules: DOP = `omp parallel for simd` (common.h:93-96) and DOT = static
 * contiguous per-thread ranges (common.h:97-104), so order-dependent outputs are identical for
 * any thread count.  The single-threaded reference phases (join, radix group, group trie,
 * sorts, index_of build/probe) are single-threaded here too.
 *
 * Deliberate, documented departures (reference behaviour undefined):
 *  - pfnIndex gathers through the promoted i64 index (the reference reads y as i64 whatever its
 *    type, h_primitive.cc:56; identical for the i64 indices every call site passes) and rejects
 *    negative indices with E_INDEX (unchecked in the reference, h_primitive.cc:50).
 *  - char descending order reverses [low,high) (the reference's swap indexes rtn[high-i+1],
 *    one and two past the range, h_libs.cc:404-406).
 *  - searchOrdered reads y with x's element type only when the types match; mismatched
 *    integer widths are promoted first (the reference reinterprets memory, h_system.cc:731-744).
 */
#include "horse_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>
#include <float.h>
#include <limits.h>
#include <omp.h>
#include <sys/mman.h>

#define R return
#define H_EPS 1e-13                   /* h_system.cc:9 */
#define MINV(a,b) (((a)<(b))?(a):(b))  /* global.h:44 */
#define MAXV(a,b) (((a)>(b))?(a):(b))  /* global.h:45 */

typedef int64_t L;
typedef double  E;

static char   *g_heap = NULL;
static L       g_cur = 0, g_top = 0, g_cap = 0;   /* bump pointer, usable top, reservation */
static int     g_threads = 1;

/* ------------------------------------------------------------------ memory (h_memory.cc) */
static L type_size(int32_t typ, L len){           /* getTypeSize, h_memory.cc:124-153 */
    L r = -1, k = 0;
    switch(typ){
        case HPA_B: case HPA_J: r = 1; break;
        case HPA_C: r = 1; k = 1; break;
        case HPA_H: r = 2; break;
        case HPA_I: case HPA_F: case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T: r = 4; break;
        case HPA_L: case HPA_E: case HPA_X: case HPA_Q: case HPA_S: case HPA_Z: case HPA_Y: r = 8; break;
        case HPA_G: case HPA_N: case HPA_A: case HPA_K: case HPA_V: r = (L)sizeof(O0); break;
        default: R -1;
    }
    R r * len + k;
}

static int is_basic(int32_t t){                    /* isTypeGroupBasic, h_system.cc:967 */
    R (t >= HPA_B && t <= HPA_T);
}
static int is_int(int32_t t){ R t==HPA_B||t==HPA_J||t==HPA_H||t==HPA_I||t==HPA_L; }
static int is_real(int32_t t){ R is_int(t) || t==HPA_F || t==HPA_E; }
static int is_comparable(int32_t t){                /* h_system.cc:963 */
    R is_real(t) || t==HPA_C || t==HPA_Q || (t>=HPA_M && t<=HPA_T);
}

/* The heap is one anonymous mapping reserved without swap accounting: pages are committed as
 * the bump pointer first touches them, so a large reservation (the caller sizes it from its
 * inputs) costs only what is used. */
int32_t orc_init(int64_t heap_bytes, int32_t threads){
    if(!g_heap){
        if(heap_bytes <= 0) heap_bytes = 1LL<<32;
        void *p = mmap(NULL, (size_t)heap_bytes, PROT_READ | PROT_WRITE,
                       MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if(p == MAP_FAILED) R HPA_E_MEMORY;
        g_heap = (char*)p;
        g_top = g_cap = heap_bytes; g_cur = 0;
    }
    if(threads > 0) g_threads = threads;
    else g_threads = omp_get_max_threads();
    omp_set_num_threads(g_threads);
    R 0;
}
int32_t orc_threads(void){ R g_threads; }
int64_t orc_get_heap(void){ R g_cur; }
int64_t orc_heap_capacity(void){ R g_cap; }
/* Lower (or restore) the heap's usable top, within the reservation; returns the previous top.
 * Lets a test run a primitive against a nearly full heap. */
int64_t orc_set_heap_top(int64_t top){
    L old = g_top;
    g_top = (top <= 0 || top > g_cap) ? g_cap : top;
    R old;
}
int64_t orc_set_heap(int64_t cur){ R g_cur = cur; }

static char *heap_alloc(L size){                   /* allocMem, h_memory.cc:93-112 */
    if(size & 0xF) size = ((size>>4)+1)<<4;
    if(g_cur + size >= g_top){ R NULL; }
    char *g = g_heap + g_cur; g_cur += size; R g;
}

OV orc_alloc_node(void){
    OV x = (OV)heap_alloc(sizeof(O0));
    if(x) memset(x, 0, sizeof(O0));
    R x;
}

/* initV (h_memory.cc:182-196).  Returns NULL when the payload does not fit the heap (the
 * reference stops the process with "Heap full!!", h_memory.cc:110); every caller turns that
 * into HPA_E_MEMORY through IV. */
OV orc_initV(OV x, int32_t typ, int64_t len){
    x->typ = typ; x->len = len;
    if(is_basic(typ)) x->g = (len > 1) ? heap_alloc(type_size(typ,len)) : (char*)&x->b;
    else x->g = heap_alloc(type_size(typ,len));
    if(typ==HPA_G) x->g2 = NULL;
    R x->g ? x : NULL;
}
#define IV(x, t, n) do{ if(!orc_initV((x), (t), (n))) R HPA_E_MEMORY; }while(0)
static void copyV(OV z, OV x){ *z = *x; }         /* h_system.cc:70-73 */
/* *z = *x for a node that moves (x is going away): an inline scalar's payload moves with it */
static void move_node(OV z, OV x){ *z = *x; if(x->g == (char*)&x->b) z->g = (char*)&z->b; }
#define vV(x,k) (((OV)(x)->g)+(k))

/* ------------------------------------------------------------------ symbols (h_symbol.cc) */
static char **g_syms = NULL; static int64_t g_nsym = 0, g_capsym = 0;
int64_t orc_symbol(const char *name){               /* getSymbol: one id per distinct name */
    for(int64_t i=0;i<g_nsym;i++) if(!strcmp(g_syms[i],name)) R i;
    if(g_nsym == g_capsym){
        int64_t c = g_capsym ? 2*g_capsym : 64;
        char **t = (char**)realloc(g_syms, sizeof(char*)*(size_t)c); if(!t) R -1;
        g_syms = t; g_capsym = c;
    }
    g_syms[g_nsym] = strdup(name); R g_nsym++;
}
static const char *sym_str(int64_t q){ R (q>=0 && q<g_nsym) ? g_syms[q] : NULL; }
const char *orc_symbol_name(int64_t q){ R sym_str(q); }
static int compare_symbol(int64_t x, int64_t y){   /* compareSymbol, h_symbol.cc:241-244 */
    R x==y ? 0 : strcmp(sym_str(x), sym_str(y));
}
static int op_from_symbol(int64_t q){             /* getOpFromName, h_system.cc:779-789 */
    if(q < 0 || q >= g_nsym) R -1;
    const char *s = g_syms[q];
    if(!strcmp(s,"lt")) R 0; if(!strcmp(s,"leq")) R 1; if(!strcmp(s,"gt")) R 2;
    if(!strcmp(s,"geq")) R 3; if(!strcmp(s,"eq")) R 4; if(!strcmp(s,"neq")) R 5;
    R -1;
}

/* ------------------------------------------------------------------ element access */
#define sB(x) ((uint8_t*)(x)->g)
#define sJ(x) ((int8_t*)(x)->g)
#define sH(x) ((int16_t*)(x)->g)
#define sI(x) ((int32_t*)(x)->g)
#define sL(x) ((int64_t*)(x)->g)
#define sF(x) ((float*)(x)->g)
#define sE(x) ((double*)(x)->g)
#define sC(x) ((char*)(x)->g)

/* promoteValue, h_system.cc:168-252 (without index) */
static int32_t promote(OV z, OV x, int32_t typMax){
    if(x->typ == typMax){ copyV(z,x); R 0; }
    L n = x->len;
    IV(z, typMax, n);
    if(n > 1 && !z->g) R HPA_E_MEMORY;
#define PROM(T, ctype) { ctype *d=(ctype*)z->g; switch(x->typ){ \
        case HPA_B: for(L i=0;i<n;i++) d[i]=(ctype)sB(x)[i]; break; \
        case HPA_J: for(L i=0;i<n;i++) d[i]=(ctype)sJ(x)[i]; break; \
        case HPA_H: for(L i=0;i<n;i++) d[i]=(ctype)sH(x)[i]; break; \
        case HPA_I: for(L i=0;i<n;i++) d[i]=(ctype)sI(x)[i]; break; \
        case HPA_L: for(L i=0;i<n;i++) d[i]=(ctype)sL(x)[i]; break; \
        case HPA_F: for(L i=0;i<n;i++) d[i]=(ctype)sF(x)[i]; break; \
        default: R HPA_E_TYPE; } }
    switch(typMax){
        case HPA_J: PROM(J, int8_t); break;
        case HPA_H: PROM(H, int16_t); break;
        case HPA_I: PROM(I, int32_t); break;
        case HPA_L: PROM(L, int64_t); break;
        case HPA_F: PROM(F, float); break;
        case HPA_E: PROM(E, double); break;
        default: R HPA_E_TYPE;
    }
#undef PROM
    R 0;
}

/* ------------------------------------------------------------------ compare (h_primitive.cc:1061-1200) */
static inline int cmp_float(double x, double y){   /* compareFloat, h_primitive.cc:1061 */
    R x<y ? (y-x<H_EPS?0:-1) : (x-y<H_EPS?0:1);
}
static inline int comp_op(int op, int c){          /* COMPFN on a three-way result */
    switch(op){ case 0: R c<0; case 1: R c<=0; case 2: R c>0; case 3: R c>=0; case 4: R c==0; case 5: R c!=0; }
    R 0;
}
#define COMP(op,a,b) (op==0?(a)<(b):op==1?(a)<=(b):op==2?(a)>(b):op==3?(a)>=(b):op==4?(a)==(b):op==5?(a)!=(b):0)

int32_t orc_pfnCompare(OV z, OV x, OV y, int32_t op){
    if(!(is_comparable(x->typ) && is_comparable(y->typ))) R HPA_E_DOMAIN;
    if(!(x->len==1 || y->len==1 || x->len==y->len)) R HPA_E_LENGTH;
    L n = (x->len==1) ? y->len : x->len;
    L sx = (x->len==1)?0:1, sy = (y->len==1)?0:1;
    if(is_real(x->typ) && is_real(y->typ)){
        int32_t t = MAXV(x->typ, y->typ);
        O0 tx, ty; int32_t e;
        if((e=promote(&tx,x,t))) R e;
        if((e=promote(&ty,y,t))) R e;
        IV(z, HPA_B, n); uint8_t *zz = sB(z);
        switch(t){
#define CMPI(ct) { ct *a=(ct*)tx.g, *b=(ct*)ty.g; _Pragma("omp parallel for simd") for(L i=0;i<n;i++) zz[i]=COMP(op,a[i*sx],b[i*sy]); } break
            case HPA_B: CMPI(uint8_t);
            case HPA_J: CMPI(int8_t);
            case HPA_H: CMPI(int16_t);
            case HPA_I: CMPI(int32_t);
            case HPA_L: CMPI(int64_t);
#undef CMPI
            case HPA_F: { float *a=(float*)tx.g, *b=(float*)ty.g;
                #pragma omp parallel for simd
                for(L i=0;i<n;i++){ float p=a[i*sx], q=b[i*sy];
                    int c = p<q ? ((double)(q-p)<H_EPS?0:-1) : ((double)(p-q)<H_EPS?0:1);
                    zz[i]=comp_op(op,c);} } break;
            case HPA_E: { double *a=(double*)tx.g, *b=(double*)ty.g;
                #pragma omp parallel for simd
                for(L i=0;i<n;i++) zz[i]=comp_op(op,cmp_float(a[i*sx],b[i*sy])); } break;
        }
        R 0;
    }
    else if(x->typ == y->typ){
        IV(z, HPA_B, n); uint8_t *zz = sB(z);
        switch(x->typ){
#define CMPS(ct) { ct *a=(ct*)x->g, *b=(ct*)y->g; for(L i=0;i<n;i++) zz[i]=COMP(op,a[i*sx],b[i*sy]); } break
            case HPA_C: CMPS(char);
            case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T: CMPS(int32_t);
            case HPA_Z: CMPS(int64_t);
            case HPA_Q: { int64_t *a=(int64_t*)x->g, *b=(int64_t*)y->g;
                for(L i=0;i<n;i++){
                    int64_t p=a[i*sx], q=b[i*sy];
                    if(p!=q && (!sym_str(p) || !sym_str(q))) R HPA_E_DOMAIN;
                    zz[i]=comp_op(op, compare_symbol(p,q)); } } break;
#undef CMPS
            default: R HPA_E_NOT_IMPL;
        }
        R 0;
    }
    R HPA_E_TYPE;
}
int32_t orc_pfnLt (OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,0); }
int32_t orc_pfnLeq(OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,1); }
int32_t orc_pfnGt (OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,2); }
int32_t orc_pfnGeq(OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,3); }
int32_t orc_pfnEq (OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,4); }
int32_t orc_pfnNeq(OV z, OV x, OV y){ R orc_pfnCompare(z,x,y,5); }

/* ------------------------------------------------------------------ arith (h_primitive.cc:1227-1309) */
static int check_zero(OV x){                       /* checkZero, h_system.cc:254-267 */
    L n = x->len;
    switch(x->typ){
        case HPA_B: for(L i=0;i<n;i++) if(!sB(x)[i]) R 1; break;
        case HPA_J: for(L i=0;i<n;i++) if(!sJ(x)[i]) R 1; break;
        case HPA_H: for(L i=0;i<n;i++) if(!sH(x)[i]) R 1; break;
        case HPA_I: for(L i=0;i<n;i++) if(!sI(x)[i]) R 1; break;
        case HPA_L: for(L i=0;i<n;i++) if(!sL(x)[i]) R 1; break;
        case HPA_F: for(L i=0;i<n;i++) if(sF(x)[i]==0) R 1; break;
        case HPA_E: for(L i=0;i<n;i++) if(sE(x)[i]==0) R 1; break;
    }
    R 0;
}
#define ARITH2(op,a,b) (op==0?(a)+(b):op==1?(a)-(b):op==2?(a)*(b):(a)/(b))

int32_t orc_pfnArith(OV z, OV x, OV y, int32_t op){
    if(!(is_real(x->typ) && is_real(y->typ))) R HPA_E_DOMAIN;
    if(!(x->len==1 || y->len==1 || x->len==y->len)) R HPA_E_LENGTH;
    if(op==3 && check_zero(y)) R HPA_E_DIV_ZERO;
    L n = (x->len==1) ? y->len : x->len;
    L sx = (x->len==1)?0:1, sy = (y->len==1)?0:1;
    int32_t t = MAXV(x->typ, y->typ);
    O0 tx, ty; int32_t e;
    if((e=promote(&tx,x,t))) R e;
    if((e=promote(&ty,y,t))) R e;
    IV(z, t, n);
    switch(t){
        case HPA_B: { uint8_t *a=sB(&tx),*b=sB(&ty),*d=sB(z);
            #pragma omp parallel for simd
            for(L i=0;i<n;i++){ int v = ARITH2(op,(int)a[i*sx],(int)b[i*sy]); d[i]=(v!=0); } } break;
#define AR(ct) { ct *a=(ct*)tx.g,*b=(ct*)ty.g,*d=(ct*)z->g; _Pragma("omp parallel for simd") for(L i=0;i<n;i++) d[i]=(ct)ARITH2(op,a[i*sx],b[i*sy]); } break
        case HPA_J: AR(int8_t);
        case HPA_H: AR(int16_t);
        case HPA_I: AR(int32_t);
        case HPA_L: AR(int64_t);
        case HPA_F: AR(float);
        case HPA_E: AR(double);
#undef AR
    }
    R 0;
}
int32_t orc_pfnPlus (OV z, OV x, OV y){ R orc_pfnArith(z,x,y,0); }
int32_t orc_pfnMinus(OV z, OV x, OV y){ R orc_pfnArith(z,x,y,1); }
int32_t orc_pfnMul  (OV z, OV x, OV y){ R orc_pfnArith(z,x,y,2); }
int32_t orc_pfnDiv  (OV z, OV x, OV y){ R orc_pfnArith(z,x,y,3); }

/* ------------------------------------------------------------------ logic (h_primitive.cc:1315-1357, h_system.cc:898-916)
 * B is C++ bool: and = x&y, or = x|y, xor = x^y; nand = ~(x&y) and nor = ~(x|y) are ints
 * -1/-2 converted to bool, i.e. always true. */
static int32_t logic(OV z, OV x, OV y, int which){
    if(!(x->typ==HPA_B && y->typ==HPA_B)) R HPA_E_DOMAIN;
    if(!(x->len==1 || y->len==1 || x->len==y->len)) R HPA_E_LENGTH;
    L n = (x->len==1) ? y->len : x->len;
    L sx = (x->len==1)?0:1, sy = (y->len==1)?0:1;
    uint8_t *a=sB(x), *b=sB(y);
    IV(z, HPA_B, n); uint8_t *d = sB(z);
    #pragma omp parallel for simd
    for(L i=0;i<n;i++){
        int p=a[i*sx], q=b[i*sy], r;
        switch(which){ case 0: r=p&q; break; case 1: r=~(p&q); break; case 2: r=p|q; break;
                       case 3: r=~(p|q); break; default: r=p^q; }
        d[i] = (r!=0);
    }
    R 0;
}
int32_t orc_pfnAnd (OV z, OV x, OV y){ R logic(z,x,y,0); }
int32_t orc_pfnNand(OV z, OV x, OV y){ R logic(z,x,y,1); }
int32_t orc_pfnOr  (OV z, OV x, OV y){ R logic(z,x,y,2); }
int32_t orc_pfnNor (OV z, OV x, OV y){ R logic(z,x,y,3); }
int32_t orc_pfnXor (OV z, OV x, OV y){ R logic(z,x,y,4); }

/* ------------------------------------------------------------------ compress / where */
static void dot_counts(const uint8_t *m, L n, L *par, int T){   /* getNumOfNonZero, h_system.cc:304-310 */
    L seg = n / T;
    #pragma omp parallel num_threads(T)
    {
        int tid = omp_get_thread_num(); L c = 0;
        L lo = tid*seg, hi = (tid != T-1) ? lo+seg : n;
        for(L i=lo;i<hi;i++) c += m[i] ? 1 : 0;
        par[tid] = c;
    }
}

int32_t orc_pfnCompress(OV z, OV x, OV y){        /* h_primitive.cc:1479-1521 */
    if(x->typ != HPA_B) R HPA_E_DOMAIN;
    if(x->len != y->len) R HPA_E_LENGTH;
    L n = x->len; int T = g_threads;
    L *par = (L*)calloc(T, sizeof(L)), *off = (L*)calloc(T, sizeof(L));
    dot_counts(sB(x), n, par, T);
    L k = 0; for(int t=0;t<T;t++) k += par[t];
    for(int t=1;t<T;t++) off[t] = off[t-1] + par[t-1];
    int32_t typ = y->typ;
    if(n != k){
        L w;
        switch(typ){
            case HPA_B: case HPA_J: case HPA_C: w=1; break;
            case HPA_H: w=2; break;
            case HPA_I: case HPA_F: case HPA_D: w=4; break;
            case HPA_L: case HPA_E: case HPA_Q: case HPA_S: case HPA_Y: w=8; break;
            default: free(par); free(off); R HPA_E_NOT_IMPL;
        }
        if(!orc_initV(z, typ, k)){ free(par); free(off); R HPA_E_MEMORY; }
        char *dst = z->g; const char *src = y->g; const uint8_t *m = sB(x);
        L seg = n / T;
        #pragma omp parallel num_threads(T)
        {
            int tid = omp_get_thread_num(); L o = off[tid];
            L lo = tid*seg, hi = (tid != T-1) ? lo+seg : n;
            for(L i=lo;i<hi;i++) if(m[i]){ memcpy(dst + o*w, src + i*w, (size_t)w); o++; }
        }
    }
    else if(k > 0) copyV(z, y);
    else if(!orc_initV(z, typ, 0)){ free(par); free(off); R HPA_E_MEMORY; }
    free(par); free(off);
    R 0;
}

int32_t orc_pfnWhere(OV z, OV x){                  /* h_primitive.cc:728-762 */
    if(x->typ != HPA_B) R HPA_E_DOMAIN;
    L n = x->len, k = 0; const uint8_t *m = sB(x);
    for(L i=0;i<n;i++) k += m[i]?1:0;
    IV(z, HPA_L, k);
    L c = 0; for(L i=0;i<n;i++) if(m[i]) sL(z)[c++] = i;
    R 0;
}

/* ------------------------------------------------------------------ index (h_primitive.cc:43-94) */
int32_t orc_pfnIndex(OV z, OV x, OV y){
    if(!is_int(y->typ)) R HPA_E_DOMAIN;
    L nz = y->len, nx = x->len;
    O0 ty; int32_t e;
    if((e=promote(&ty, y, HPA_L))) R e;
    const int64_t *iy = sL(&ty);
    for(L i=0;i<nz;i++) if(iy[i] >= nx || iy[i] < 0) R HPA_E_INDEX;
    if(is_basic(x->typ)){
        L w = type_size(x->typ, 1); if(x->typ==HPA_C) w = 1;
        IV(z, x->typ, nz);
        char *d = z->g; const char *s = x->g;
        #pragma omp parallel for
        for(L i=0;i<nz;i++) memcpy(d + i*w, s + iy[i]*w, (size_t)w);
        R 0;
    }
    else if(x->typ == HPA_G){
        if(nz == 1){ copyV(z, vV(x, iy[0])); R 0; }
        R HPA_E_NOT_IMPL;
    }
    R HPA_E_NOT_IMPL;
}

int32_t orc_pfnList(OV z, int64_t n, OV x[]){      /* h_primitive.cc:32-36 */
    IV(z, HPA_G, n); z->g2 = NULL;
    for(L i=0;i<n;i++) *vV(z,i) = *x[i];
    R 0;
}

int32_t orc_pfnLen(OV z, OV x){                    /* h_primitive.cc:656-662 */
    IV(z, HPA_L, 1);
    z->i64 = x->len;
    R 0;
}

/* ------------------------------------------------------------------ reductions (h_primitive.cc:764-848) */
int32_t orc_pfnSum(OV z, OV x){
    if(!is_real(x->typ)) R HPA_E_DOMAIN;
    int32_t tz = x->typ==HPA_F?HPA_F:x->typ==HPA_E?HPA_E:HPA_L;
    L n = x->len;
    IV(z, tz, 1);
    switch(x->typ){
#define SUMI(ct) { L t=0; ct *a=(ct*)x->g; _Pragma("omp parallel for simd reduction(+:t)") for(L i=0;i<n;i++) t+=a[i]; z->i64=t; } break
        case HPA_B: SUMI(uint8_t);
        case HPA_J: SUMI(int8_t);
        case HPA_H: SUMI(int16_t);
        case HPA_I: SUMI(int32_t);
        case HPA_L: SUMI(int64_t);
#undef SUMI
        case HPA_F: { float t=0; float *a=sF(x);
            #pragma omp parallel for simd reduction(+:t)
            for(L i=0;i<n;i++) t+=a[i];
            z->f32=t; } break;
        case HPA_E: { double t=0; double *a=sE(x);
            #pragma omp parallel for simd reduction(+:t)
            for(L i=0;i<n;i++) t+=a[i];
            z->f64=t; } break;
    }
    R 0;
}

int32_t orc_pfnAvg(OV z, OV x){
    O0 t; int32_t e;
    if((e=orc_pfnSum(&t, x))) R e;
    int32_t tz = t.typ==HPA_F ? HPA_F : HPA_E;
    IV(z, tz, 1);
    switch(t.typ){
        case HPA_L: z->f64 = t.i64*1.0/x->len; break;
        case HPA_F: z->f32 = t.f32/x->len; break;
        case HPA_E: z->f64 = t.f64/x->len; break;
    }
    R 0;
}

int32_t orc_pfnReduce(OV z, OV x, int32_t op){
    if(!is_real(x->typ)) R HPA_E_DOMAIN;
    L n = x->len;
    if(n < 1){
        int32_t tz = is_int(x->typ) ? HPA_L : x->typ;
        IV(z, tz, 1);
        switch(tz){
            case HPA_L: z->i64 = op==0 ? LLONG_MAX : LLONG_MIN; break;
            case HPA_F: z->f32 = op==0 ? FLT_MAX : FLT_MIN; break;
            case HPA_E: z->f64 = op==0 ? DBL_MAX : DBL_MIN; break;
        }
        R 0;
    }
    IV(z, x->typ, 1);
    switch(x->typ){
#define RED(ct, fld) { ct *a=(ct*)x->g; ct t=a[0]; for(L i=0;i<n;i++) t = op==0?MINV(t,a[i]):MAXV(t,a[i]); z->fld=t; } break
        case HPA_B: { uint8_t *a=sB(x); uint8_t t=a[0]; for(L i=0;i<n;i++) t = op==0?MINV(t,a[i]):MAXV(t,a[i]); z->b=t; } break;
        case HPA_J: RED(int8_t, j);
        case HPA_H: RED(int16_t, h);
        case HPA_I: RED(int32_t, i32);
        case HPA_L: RED(int64_t, i64);
        case HPA_F: RED(float, f32);
        case HPA_E: RED(double, f64);
#undef RED
    }
    R 0;
}
int32_t orc_pfnMin(OV z, OV x){ R orc_pfnReduce(z,x,0); }
int32_t orc_pfnMax(OV z, OV x){ R orc_pfnReduce(z,x,1); }

double orc_sum_neumaier(const double *x, int64_t n){
    double s = 0, c = 0;
    for(L i=0;i<n;i++){ double v=x[i], t=s+v; c += (fabs(s)>=fabs(v)) ? (s-t)+v : (v-t)+s; s=t; }
    R s + c;
}
double orc_sum_neumaier_idx(const double *x, const int64_t *idx, int64_t n){
    double s = 0, c = 0;
    for(L i=0;i<n;i++){ double v=x[idx[i]], t=s+v; c += (fabs(s)>=fabs(v)) ? (s-t)+v : (v-t)+s; s=t; }
    R s + c;
}

/* ------------------------------------------------------------------ sorting (h_libs.cc) */
typedef struct { int32_t x; L i; } PosI;   /* Pos_i, h_libs.cc:1182 */
typedef struct { L x; int32_t i; } PosL;   /* Pos,   h_libs.cc:1181 */

static void radix_core_int(PosI *v, L len){        /* lib_radixsort_core_int, h_libs.cc:1184-1209 */
    L *cnt = (L*)malloc(sizeof(L)*256); PosI *tmp = (PosI*)malloc(sizeof(PosI)*(len>0?len:1));
    for(int j=0;j<32;j+=8){
        memset(cnt, 0, sizeof(L)*256);
        for(L i=0;i<len;i++) cnt[(v[i].x>>j)&255]++;
        for(int i=1;i<256;i++) cnt[i]+=cnt[i-1];
        for(L i=len-1;i>=0;i--){ L k=(v[i].x>>j)&255; tmp[--cnt[k]] = v[i]; }
        memcpy(v, tmp, sizeof(PosI)*len);
    }
    free(cnt); free(tmp);
}

static void radix_core_long(PosL *v, L len){       /* lib_radixsort_core_long, h_libs.cc:1240-1266 */
    int32_t *cnt = (int32_t*)malloc(sizeof(int32_t)*256); PosL *tmp = (PosL*)malloc(sizeof(PosL)*(len>0?len:1));
    int maxSize = 32;
    for(L i=0;i<len;i++) if((v[i].x>>32) > 0){ maxSize = 64; break; }
    for(int j=0;j<maxSize;j+=8){
        memset(cnt, 0, sizeof(int32_t)*256);
        for(L i=0;i<len;i++) cnt[(v[i].x>>j)&255]++;
        for(int i=1;i<256;i++) cnt[i]+=cnt[i-1];
        for(L i=len-1;i>=0;i--){ L k=(v[i].x>>j)&255; tmp[--cnt[k]] = v[i]; }
        memcpy(v, tmp, sizeof(PosL)*len);
    }
    free(cnt); free(tmp);
}

static void radixsort_int(L *rtn, OV val, L len, int up){   /* lib_radixsort_int, h_libs.cc:1215-1238 */
    PosI *p = (PosI*)malloc(sizeof(PosI)*(len>0?len:1));
    for(L i=0;i<len;i++){ p[i].x = sI(val)[i]; p[i].i = i; }
    radix_core_int(p, len);
    for(L i=0;i<len;i++) rtn[i] = up ? p[i].i : p[len-i-1].i;
    free(p);
}
static void radixsort_long_keys(L *rtn, const int64_t *keys, L len, int up){ /* lib_radixsort_long, h_libs.cc:1268-1283 */
    PosL *p = (PosL*)malloc(sizeof(PosL)*(len>0?len:1));
    for(L i=0;i<len;i++){ p[i].x = keys[i]; p[i].i = (int32_t)i; }
    radix_core_long(p, len);
    for(L i=0;i<len;i++) rtn[i] = up ? p[i].i : p[len-i-1].i;
    free(p);
}

/* lib_quicksort_cmp_item, h_libs.cc:451-482 (SORT_MERGE off: ties compare 0) */
static L cmp_item(OV t, L a, L b, const uint8_t *isUp){
    int f = isUp ? *isUp : 1;
#define SC(ct) { ct p=((ct*)t->g)[a], q=((ct*)t->g)[b]; if(p!=q) R p<q ? (f?-1:1) : (f?1:-1); } break
    switch(t->typ){
        case HPA_B: SC(uint8_t); case HPA_J: SC(int8_t); case HPA_H: SC(int16_t);
        case HPA_I: case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T: SC(int32_t);
        case HPA_L: case HPA_Z: case HPA_Q: SC(int64_t);
        case HPA_F: SC(float); case HPA_E: SC(double); case HPA_C: SC(char);
        default: break;
    }
#undef SC
    R 0;
}
/* lib_quicksort_cmp, h_libs.cc:445-449 */
static L cmp_list(OV val, L a, L b, const uint8_t *isUp){
    for(L i=0;i<val->len;i++){ uint8_t f = isUp ? isUp[i] : 1; L x = cmp_item(vV(val,i), a, b, &f); if(x) R x; }
    R 0;
}

/* lib_partition / lib_quicksort_other, h_libs.cc:409-442 */
static L partition(L *rtn, OV val, L low, L high, const uint8_t *isUp, int *leftSame){
    L pivot = low, i = low, j = high, t;
    while(1){
        while(++i && i<high){
            L c = cmp_item(val, rtn[i], rtn[pivot], isUp);
            if(c != 0) *leftSame = 0;
            if(c > 0) break;
        }
        do{ --j; } while(j>=0 && cmp_item(val, rtn[j], rtn[pivot], isUp) > 0);
        if(i>=j) break;
        t = rtn[i]; rtn[i] = rtn[j]; rtn[j] = t;
    }
    if(!*leftSame){ t = rtn[low]; rtn[low] = rtn[j]; rtn[j] = t; }
    R j;
}
static void quicksort_other(L *rtn, OV val, L low, L high, const uint8_t *isUp){
    if(low < high){
        int leftSame = 1;
        L pos = partition(rtn, val, low, high, isUp, &leftSame);
        if(!leftSame) quicksort_other(rtn, val, low, pos, isUp);
        if(pos < high) quicksort_other(rtn, val, pos+1, high, isUp);
    }
}
/* lib_quicksort_char, h_libs.cc:386-407: <10 items use the all-pairs exchange with `>` (which
 * leaves the range in non-increasing order), otherwise a stable counting sort */
static void quicksort_char(L *rtn, OV val, L low, L high, const uint8_t *isUp){
    const char *str = sC(val); L len = high-low; int f0 = isUp ? *isUp : 1;
    if(len < 10){
        for(L i=low;i<high;i++) for(L j=low;j<high;j++)
            if(str[rtn[i]] > str[rtn[j]]){ L t=rtn[i]; rtn[i]=rtn[j]; rtn[j]=t; }
    } else {
        L cnt[256]={0}, step[256]={0}, cursor=0;
        L *temp = (L*)malloc(sizeof(L)*len);
        for(L i=low;i<high;i++) cnt[(unsigned char)str[rtn[i]]]++;
        for(int i=0;i<256;i++){ if(cnt[i]!=0) step[i]=cursor; cursor+=cnt[i]; }
        for(L i=low;i<high;i++) temp[step[(unsigned char)str[rtn[i]]]++] = rtn[i];
        memcpy(rtn+low, temp, sizeof(L)*len); free(temp);
    }
    if(!f0){ for(L i=low, k=high-1; i<k; i++, k--){ L t=rtn[i]; rtn[i]=rtn[k]; rtn[k]=t; } }
}
static void lib_quicksort(L *rtn, OV val, L low, L high, const uint8_t *isUp){ /* h_libs.cc:375-384 */
    switch(val->typ){
        case HPA_C: quicksort_char(rtn, val, low, high, isUp); break;
        case HPA_I: radixsort_int(rtn, val, high, isUp ? *isUp : 1); break;
        case HPA_L: radixsort_long_keys(rtn, sL(val), high, isUp ? *isUp : 1); break;
        default: quicksort_other(rtn, val, low, high, isUp); break;
    }
}

/* lib_mergesort, h_libs.cc:490-536 */
static L *g_buff = NULL;
static void mergesort_r(L *rtn, OV val, L low, L high, const uint8_t *isUp){
    if(low < high - 1){
        if(low + 1 == high - 1){
            if(cmp_item(val, rtn[low], rtn[high-1], isUp) > 0){ L t=rtn[low]; rtn[low]=rtn[high-1]; rtn[high-1]=t; }
        } else {
            L mid = (low + high) >> 1, len = high-low;
            mergesort_r(rtn, val, low, mid, isUp);
            mergesort_r(rtn, val, mid, high, isUp);
            if(cmp_item(val, rtn[mid-1], rtn[mid], isUp) > 0){
                L cl = low, cr = mid;
                for(L i=0;i<len;i++){
                    if(cl == mid) g_buff[i] = rtn[cr++];
                    else if(cr == high) g_buff[i] = rtn[cl++];
                    else if(cmp_item(val, rtn[cl], rtn[cr], isUp) > 0) g_buff[i] = rtn[cr++];
                    else g_buff[i] = rtn[cl++];
                }
                memcpy(rtn+low, g_buff, sizeof(L)*len);
            }
        }
    }
}

/* lib_order_by_sorted, h_libs.cc:604-617 */
static int order_by_sorted(L *targ, OV val, const uint8_t *isUp, L low, L high){
    L rtn = 0;
    for(L i=low+1;i<high;i++){
        L t = cmp_item(val, targ[i-1], targ[i], isUp);
        rtn = rtn==0 ? t : rtn==-1 ? (t==1?-2:rtn) : (t==-1?-2:rtn);
        if(rtn == -2) R 0;
    }
    int isOrder = rtn == (*isUp ? -1 : 1);
    if(isOrder && !(*isUp)){
        for(L i=low;i<(low+high)/2;i++){ L pos = high-(i-low)-1; L t=targ[i]; targ[i]=targ[pos]; targ[pos]=t; }
    }
    R isOrder;
}

static int is_integers(OV v){ R v->typ==HPA_I || v->typ==HPA_L; }

/* lib_quicksort_list, h_libs.cc:637-697 */
static void quicksort_list(L *targ, OV val, const uint8_t *isUp, L low, L high, L colId){
    if(colId >= val->len) R;
    OV cur = vV(val, colId); const uint8_t *cb = isUp + colId;
    if(order_by_sorted(targ, cur, cb, low, high)) { ; }
    else if(is_integers(cur)){
        g_buff = (L*)malloc(sizeof(L)*(cur->len>0?cur->len:1));
        mergesort_r(targ, cur, low, high, cb);
        free(g_buff); g_buff = NULL;
    }
    else lib_quicksort(targ, cur, low, high, cb);
    L pos = low;
    do{
        L start = pos, end = high;
        L hend = start + 5, hpos = -1;
        if(hend > high) hend = high;
        for(L i=start+1;i<hend;i++) if(cmp_item(cur, targ[i-1], targ[i], cb) != 0){ hpos = i; break; }
        if(hpos >= 0){ quicksort_list(targ, val, isUp, start, hpos, colId+1); pos = hpos; }
        else if(hend == high){ quicksort_list(targ, val, isUp, start, high, colId+1); pos = high; }
        else {
            L pivot = start;
            do{
                if(0 != cmp_item(cur, targ[pivot], targ[end-1], cb)){
                    L mid = (start + end) >> 1;
                    if(0 != cmp_item(cur, targ[pivot], targ[mid], cb)) end = mid;
                    else start = mid+1;
                } else break;
            } while(start < end);
            quicksort_list(targ, val, isUp, pivot, end, colId+1);
            pos = end;
        }
    } while(pos < high);
}

static void order_by_list(L *targ, OV val, const uint8_t *isUp, L n){ /* h_libs.cc:619-635 */
    for(L i=0;i<n;i++) targ[i] = i;
    quicksort_list(targ, val, isUp, 0, n, 0);
}
static void order_by_vector(L *targ, OV val, const uint8_t *isUp, L n){ /* h_libs.cc:730-736 */
    if(!is_integers(val)) for(L i=0;i<n;i++) targ[i] = i;
    lib_quicksort(targ, val, 0, n, isUp);
}

/* checkOrderCase1, h_libs.cc:1912-1940: both columns i32 and equal flags.  The packed key is
 * ((0|x0)<<32)|x1 with x0, x1 sign-extended, so a negative x1 sets all high bits.
 * Returns -1 when the case does not apply, else 0 or an error code. */
static int32_t check_order_case1(OV z, OV x, OV f){
    if(x->len==2 && vV(x,0)->typ==HPA_I && vV(x,1)->typ==HPA_I && sB(f)[0]==sB(f)[1]){
        OV x0 = vV(x,0), x1 = vV(x,1);
        L n = x0->len;
        int64_t *k = (int64_t*)malloc(sizeof(int64_t)*(n>0?n:1));
        if(!k) R HPA_E_MEMORY;
        for(L i=0;i<n;i++) k[i] = (int64_t)((uint64_t)(int64_t)sI(x0)[i] << 32) | (int64_t)sI(x1)[i];
        if(!orc_initV(z, HPA_L, n)){ free(k); R HPA_E_MEMORY; }
        radixsort_long_keys(sL(z), k, n, sB(f)[0]);
        free(k);
        R 0;
    }
    R -1;
}

static int32_t order_basic(OV z, OV x, OV y){      /* lib_order_basic, h_primitive.cc:1760-1767 */
    if(is_basic(x->typ) && y->typ==HPA_B && y->len==1){
        if(x->typ==HPA_S || x->typ==HPA_X) R HPA_E_NOT_IMPL;
        IV(z, HPA_L, x->len);
        order_by_vector(sL(z), x, sB(y), x->len);
        R 0;
    }
    R 1;
}

/* A NaN key makes lib_partition's j scan leave its segment (a NaN pivot compares greater than
 * itself, h_libs.cc:419), so orders and groups over F/E columns holding NaN are rejected with
 * E_DOMAIN instead (defined behaviour chosen where the reference's is undefined). */
static int has_nan(OV v){
    if(v->typ == HPA_E){ for(L i=0;i<v->len;i++) if(sE(v)[i] != sE(v)[i]) R 1; }
    if(v->typ == HPA_F){ for(L i=0;i<v->len;i++) if(sF(v)[i] != sF(v)[i]) R 1; }
    R 0;
}
int32_t orc_pfnOrderBy(OV z, OV x, OV y){          /* h_primitive.cc:1769-1816 */
    if(x->typ==HPA_G){ for(L i=0;i<x->len;i++) if(has_nan(vV(x,i))) R HPA_E_DOMAIN; }
    else if(has_nan(x)) R HPA_E_DOMAIN;
    if(x->typ==HPA_G && y->typ==HPA_B && x->len==y->len){
        for(L i=0;i<x->len;i++) if(!is_basic(vV(x,i)->typ)) R HPA_E_DOMAIN;
        if(x->len > 0){ L n=vV(x,0)->len; for(L i=0;i<x->len;i++) if(vV(x,i)->len!=n) R HPA_E_MATCH; }
        int32_t c1 = check_order_case1(z, x, y);
        if(c1 >= 0) R c1;
        if(x->len == 1) R order_basic(z, vV(x,0), y);
        else {
            L n = x->len==0 ? 0 : vV(x,0)->len;
            IV(z, HPA_L, n);
            order_by_list(sL(z), x, sB(y), n);
        }
        R 0;
    }
    R order_basic(z, x, y);
}

/* ------------------------------------------------------------------ group */
static int32_t init_dict(OV z, L g){
    IV(z, HPA_N, 2);
    IV(vV(z,0), HPA_L, g);
    IV(vV(z,1), HPA_G, g);
    R 0;
}

/* lib_group_by_normal_int, h_libs.cc:1295-1329 */
static int32_t group_int(OV z, OV x){
    L n = x->len;
    PosI *p = (PosI*)malloc(sizeof(PosI)*(n>0?n:1));
    if(!p) R HPA_E_MEMORY;
    for(L i=0;i<n;i++){ p[i].x = sI(x)[i]; p[i].i = i; }
    radix_core_int(p, n);
    L g = n>0 ? 1 : 0;
    for(L i=1;i<n;i++) if(p[i-1].x != p[i].x) g++;
    if(init_dict(z, g)){ free(p); R HPA_E_MEMORY; }
    OV keys = vV(z,0), vals = vV(z,1);
    L k = 0, s = 0;
    for(L i=1;i<=n;i++){
        if(i==n || p[i-1].x != p[i].x){
            OV d = vV(vals,k); if(!orc_initV(d, HPA_L, i-s)){ free(p); R HPA_E_MEMORY; }
            for(L j=s;j<i;j++) sL(d)[j-s] = p[j].i;
            sL(keys)[k] = p[s].i; k++; s = i;
        }
    }
    free(p); R 0;
}
/* lib_group_by_normal_long, h_libs.cc:1369-1403 */
static int32_t group_long(OV z, OV x){
    L n = x->len;
    PosL *p = (PosL*)malloc(sizeof(PosL)*(n>0?n:1));
    if(!p) R HPA_E_MEMORY;
    for(L i=0;i<n;i++){ p[i].x = sL(x)[i]; p[i].i = (int32_t)i; }
    radix_core_long(p, n);
    L g = n>0 ? 1 : 0;
    for(L i=1;i<n;i++) if(p[i-1].x != p[i].x) g++;
    if(init_dict(z, g)){ free(p); R HPA_E_MEMORY; }
    OV keys = vV(z,0), vals = vV(z,1);
    L k = 0, s = 0;
    for(L i=1;i<=n;i++){
        if(i==n || p[i-1].x != p[i].x){
            OV d = vV(vals,k); if(!orc_initV(d, HPA_L, i-s)){ free(p); R HPA_E_MEMORY; }
            for(L j=s;j<i;j++) sL(d)[j-s] = p[j].i;
            sL(keys)[k] = p[s].i; k++; s = i;
        }
    }
    free(p); R 0;
}
/* lib_get_group_by_other / _order, h_libs.cc:776-846: runs of equal keys in `index` order
 * (index NULL = identity) */
static int32_t group_runs(OV z, OV val, const L *index, L n, int isList){
    L g = n>0 ? 1 : 0;
#define IX(i) (index ? index[i] : (i))
#define CMPR(a,b) (isList ? cmp_list(val,(a),(b),NULL) : cmp_item(val,(a),(b),NULL))
    for(L i=1;i<n;i++) if(CMPR(IX(i-1), IX(i)) != 0) g++;
    if(init_dict(z, g)) R HPA_E_MEMORY;
    OV keys = vV(z,0), vals = vV(z,1);
    L k = 0, s = 0;
    for(L i=1;i<=n;i++){
        if(i==n || CMPR(IX(i-1), IX(i)) != 0){
            OV d = vV(vals,k); IV(d, HPA_L, i-s);
            for(L j=s;j<i;j++) sL(d)[j-s] = IX(j);
            sL(keys)[k] = IX(s); k++; s = i;
        }
    }
#undef IX
#undef CMPR
    R 0;
}

/* isOrdered, h_system.cc:678-698 */
static int is_ordered(OV x){
    L n = x->len;
    if(x->typ == HPA_I || x->typ == HPA_L){
        int order = -2;
#define GETV(i) (x->typ==HPA_I ? (int64_t)sI(x)[i] : sL(x)[i])
        for(L i=1;i<n;i++){ int64_t a=GETV(i), b=GETV(i-1); int v = a>b?1:a<b?-1:0; if(order==-2 && v!=0){ order=v; break; } }
        if(order == -2) R 1;
        for(L i=1;i<n;i++){ int64_t a=GETV(i), b=GETV(i-1); int v = a>b?1:a<b?-1:0; if(v!=0 && order!=v) R 0; }
#undef GETV
        R 1;
    }
    R 0;
}

int32_t orc_pfnGroupTrie(OV z, OV x){              /* h_primitive.cc:2656-2681 */
    OV c0 = vV(x,0), c1 = vV(x,1);
    L n = c0->len;
    L *cnt = (L*)calloc(65536, sizeof(L)), *id = (L*)malloc(sizeof(L)*65536);
    if(!cnt || !id){ free(cnt); free(id); R HPA_E_MEMORY; }
#define CODE(i) ((((int)(unsigned char)sC(c0)[i])<<8) + (int)(unsigned char)sC(c1)[i])
    for(L i=0;i<n;i++) cnt[CODE(i)]++;
    L g = 0; for(int i=0;i<65536;i++) if(cnt[i]) id[i] = g++;
    if(init_dict(z, g)){ free(cnt); free(id); R HPA_E_MEMORY; }
    OV keys = vV(z,0), vals = vV(z,1);
    L c = 0;
    for(int i=0;i<65536;i++) if(cnt[i] && !orc_initV(vV(vals,c++), HPA_L, cnt[i])){ free(cnt); free(id); R HPA_E_MEMORY; }
    L *fill = (L*)calloc(g>0?g:1, sizeof(L));
    for(L i=0;i<n;i++){ L k = id[CODE(i)]; sL(vV(vals,k))[fill[k]++] = i; }
    for(L k=0;k<g;k++) sL(keys)[k] = sL(vV(vals,k))[0];
#undef CODE
    free(cnt); free(id); free(fill);
    R 0;
}

int32_t orc_pfnGroupBasic(OV z, OV x){             /* h_primitive.cc:1002-1057 */
    L *order = NULL; O0 t;
    if(is_ordered(x)) order = NULL;
    else if(is_integers(x)) { ; }
    else {
        L lenZ = x->typ==HPA_G ? x->len : 1;
        O0 y; IV(&y, HPA_B, lenZ);
        for(L i=0;i<lenZ;i++) sB(&y)[i] = 1;
        int32_t e = orc_pfnOrderBy(&t, x, &y); if(e) R e;
        order = sL(&t);
    }
    if(x->typ == HPA_G){
        L n = x->len==0 ? 0 : vV(x,0)->len;
        R group_runs(z, x, order, n, 1);
    }
    else if(x->typ == HPA_I) R group_int(z, x);
    else if(x->typ == HPA_L) R group_long(z, x);
    else if(is_basic(x->typ)) R group_runs(z, x, order, x->len, 0);
    R HPA_E_DOMAIN;
}

int32_t orc_pfnGroup(OV z, OV x){                  /* h_primitive.cc:993-999 */
    if(x->typ==HPA_G && x->len==2 && vV(x,0)->typ==HPA_C && vV(x,1)->typ==HPA_C && vV(x,0)->len==vV(x,1)->len)
        R orc_pfnGroupTrie(z, x);
    R orc_pfnGroupBasic(z, x);
}

int32_t orc_pfnKeys(OV z, OV x){                   /* h_primitive.cc:190-203 */
    if(x->typ==HPA_Y){ copyV(z, ((hpa_Y0*)x->p)->ek); R 0; }      /* getEnumKey */
    if(x->typ==HPA_N || x->typ==HPA_A){ copyV(z, vV(x,0)); R 0; }
    R HPA_E_DOMAIN;
}
int32_t orc_pfnValues(OV z, OV x){                 /* h_primitive.cc:205-218 */
    if(x->typ==HPA_Y){                             /* copyEnumIndex, h_system.cc:501-506 */
        IV(z, HPA_L, x->len);
        if(x->len == 1) z->i64 = *(int64_t*)x->g;
        else z->g = x->g;
        R 0;
    }
    if(x->typ==HPA_N || x->typ==HPA_A){ copyV(z, vV(x,1)); R 0; }
    R HPA_E_DOMAIN;
}
int32_t orc_pfnFetch(OV z, OV x){                  /* h_primitive.cc:220-225 */
    if(x->typ==HPA_Y){ copyV(z, ((hpa_Y0*)x->p)->ev); R 0; }      /* getEnumVal */
    R HPA_E_DOMAIN;
}
/* pfnEnum, h_primitive.cc:2198-2222: index_of the values in the keys; a value that is not
 * among the keys is E_ENUM_INDEX.  The index vector is the enum's payload (initEnum). */
int32_t orc_pfnEnum(OV z, OV x, OV y){
    if(is_basic(x->typ)){
        OV ind = orc_alloc_node(); if(!ind) R HPA_E_MEMORY;
        int32_t e = orc_pfnIndexOf(ind, x, y); if(e) R e;
        L n = ind->len;
        for(L i=0;i<n;i++) if(sL(ind)[i] >= x->len) R HPA_E_ENUM_INDEX;
        hpa_Y0 *yy = (hpa_Y0*)heap_alloc(sizeof(hpa_Y0)); if(!yy) R HPA_E_MEMORY;
        yy->source = -1; yy->ek = x; yy->ev = y;
        z->typ = HPA_Y; z->len = n; z->g = ind->g; z->p = yy;     /* ind is a heap node: inline ok */
        R 0;
    }
    if(x->typ==HPA_G && y->typ==HPA_G) R HPA_E_NOT_IMPL;
    R HPA_E_DOMAIN;
}
/* the catalog (registerTable / findTableByName, h_system.cc:58-67) */
static struct { int64_t sid; OV t; } g_tabs[64]; static int g_ntab = 0;
int32_t orc_register_table(int64_t sid, OV t){
    for(int i=0;i<g_ntab;i++) if(g_tabs[i].sid == sid){ g_tabs[i].t = t; R 0; }
    if(g_ntab >= 64) R HPA_E_MEMORY;
    g_tabs[g_ntab].sid = sid; g_tabs[g_ntab].t = t; g_ntab++;
    R 0;
}
static OV find_table(OV x){                        /* fetchTableByName, h_primitive.cc:257-262 */
    if(x->typ != HPA_Q || x->len != 1) R NULL;
    for(int i=0;i<g_ntab;i++) if(g_tabs[i].sid == x->q) R g_tabs[i].t;
    R NULL;
}
int32_t orc_pfnLoadTable(OV z, OV x){              /* h_primitive.cc:20-30 (catalog lookup only) */
    if(x->typ != HPA_Q) R HPA_E_DOMAIN;
    OV t = find_table(x); if(!t) R HPA_E_TABLE_NOT_FOUND;
    *z = *t; R 0;
}
static const int64_t *sym_of(OV s){ R s->len <= 1 ? &s->q : (const int64_t*)s->g; }
static OV fkey_by_name(OV t, OV m){               /* fetchFKeyByName, h_primitive.cc:264-281 */
    if(t->typ != HPA_A || !t->p || m->typ != HPA_Q) R NULL;
    for(hpa_ListY *y = ((hpa_A0*)t->p)->fkey; y; y = y->next){
        OV col = (OV)y->cn;
        if(col->typ != m->typ || col->len != m->len) continue;
        int f = 1; for(L i=0;i<m->len;i++) if(sym_of(m)[i] != sym_of(col)[i]){ f = 0; break; }
        if(f) R (OV)y->y;
    }
    R NULL;
}
static L find_col(OV t, int64_t cid){              /* findColFromTable, h_system.cc:157-166 */
    OV k = vV(t,0);
    for(L i=0;i<k->len;i++) if(sym_of(k)[i] == cid) R i;
    R -1;
}
static int32_t column_value(OV z, OV x, OV y, int fk){  /* fetchColumnValue, h_primitive.cc:283-308 */
    if(!(x->typ==HPA_A || x->typ==HPA_K) || y->typ != HPA_Q) R HPA_E_DOMAIN;
    if(fk){ OV e0 = fkey_by_name(x, y); if(e0){ copyV(z, e0); R 0; } }
    OV vals = vV(x,1);
    if(y->len == 1){ L k = find_col(x, y->q); if(k < 0) R HPA_E_COL_NOT_FOUND; copyV(z, vV(vals,k)); R 0; }
    IV(z, HPA_G, y->len);
    for(L i=0;i<y->len;i++){ L k = find_col(x, sym_of(y)[i]); if(k < 0) R HPA_E_COL_NOT_FOUND; copyV(vV(z,i), vV(vals,k)); }
    R 0;
}
int32_t orc_pfnColumnValue(OV z, OV x, OV y){ R column_value(z, x, y, 1); }   /* h_primitive.cc:310-313 */
/* pfnAddFKey, h_primitive.cc:2425-2466, with setFKey/insertFKey (h_system.cc:652-675) */
int32_t orc_pfnAddFKey(OV x, OV xKey, OV y, OV yKey){
    OV tx = find_table(x); if(!tx) R HPA_E_DOMAIN;
    OV ty = find_table(y); if(!ty) R HPA_E_DOMAIN;
    if(!(xKey->typ==HPA_Q && yKey->typ==xKey->typ && yKey->len==xKey->len)) R HPA_E_DOMAIN;
    if(ty->typ != HPA_A || !ty->p) R HPA_E_DOMAIN;
    OV xc = orc_alloc_node(), yc = orc_alloc_node(), fk = orc_alloc_node();
    if(!xc || !yc || !fk) R HPA_E_MEMORY;
    int32_t e;
    if((e = column_value(xc, tx, xKey, 0)) || (e = column_value(yc, ty, yKey, 0))) R e;
    if((e = orc_pfnEnum(fk, xc, yc))) R e;
    hpa_ListY *n = (hpa_ListY*)heap_alloc(sizeof(hpa_ListY)); if(!n) R HPA_E_MEMORY;
    n->y = fk; n->cn = yKey; n->next = NULL;
    hpa_A0 *a = (hpa_A0*)ty->p;
    if(!a->fkey) a->fkey = n;
    else { hpa_ListY *t = a->fkey; while(t->next) t = t->next; t->next = n; }
    R 0;
}

/* ------------------------------------------------------------------ each / raze */
int32_t orc_pfnEach(OV z, OV x, int32_t (*foo)(OV, OV)){      /* h_primitive.cc:1818-1832 */
    if(x->typ != HPA_G) R HPA_E_DOMAIN;
    IV(z, HPA_G, x->len);
    for(L i=0;i<x->len;i++){ int32_t e = foo(vV(z,i), vV(x,i)); if(e) R e; }
    R 0;
}
int32_t orc_pfnEachRight(OV z, OV x, OV y, int32_t (*foo)(OV, OV, OV)){ /* h_primitive.cc:1882-1904 */
    if(y->typ == HPA_G){
        IV(z, HPA_G, y->len);
        for(L i=0;i<y->len;i++){ int32_t e = foo(vV(z,i), x, vV(y,i)); if(e) R e; }
        R 0;
    }
    R foo(z, x, y);
}
int32_t orc_pfnRaze(OV z, OV x){                   /* h_primitive.cc:935-955, fillRaze */
    if(x->typ == HPA_G){
        if(x->len == 0) R HPA_E_RAZE_LIST;
        int32_t t = vV(x,0)->typ; L n = 0;
        for(L i=0;i<x->len;i++){ if(vV(x,i)->typ != t) R HPA_E_TYPE; n += vV(x,i)->len; }
        L w = type_size(t,1); if(t==HPA_C) w = 1;
        IV(z, t, n);
        char *d = z->g; L c = 0;
        for(L i=0;i<x->len;i++){ OV v = vV(x,i); memcpy(d + c*w, v->g, (size_t)(w*v->len)); c += v->len; }
        R 0;
    }
    if(is_basic(x->typ)){ copyV(z, x); R 0; }
    R HPA_E_DOMAIN;
}

/* ------------------------------------------------------------------ join (h_primitive.cc:1944-2155, h_hash_v3.cc:230-260) */
static int cmp_keys_i32(const void *a, const void *b){
    const int64_t *p = (const int64_t*)a, *q = (const int64_t*)b;
    if(p[0] != q[0]) R p[0] < q[0] ? -1 : 1;
    R p[1] < q[1] ? -1 : (p[1] > q[1]);
}
/* lib_join_radix_hash_v3: build on x (inserted in row order, duplicates appended in order),
 * probe y in row order; pairs (x-row, y-row) are written probe-major, build rows ascending. */
static int32_t join_hash(OV z0, OV z1, OV x, OV y){
    L nx = x->len, ny = y->len;
    int64_t *b = (int64_t*)malloc(sizeof(int64_t)*2*(nx>0?nx:1));
    for(L i=0;i<nx;i++){ b[2*i] = sI(x)[i]; b[2*i+1] = i; }
    qsort(b, (size_t)nx, 2*sizeof(int64_t), cmp_keys_i32);
    L *lo = (L*)malloc(sizeof(L)*(ny>0?ny:1)), *cnt = (L*)malloc(sizeof(L)*(ny>0?ny:1));
    L c = 0;
    for(L j=0;j<ny;j++){
        int64_t key = sI(y)[j]; L a = 0, h = nx;
        while(a < h){ L m = (a+h)>>1; if(b[2*m] < key) a = m+1; else h = m; }
        L s = a; while(a < nx && b[2*a] == key) a++;
        lo[j] = s; cnt[j] = a - s; c += a - s;
    }
    if(!orc_initV(z0, HPA_L, c) || !orc_initV(z1, HPA_L, c)){ free(b); free(lo); free(cnt); R HPA_E_MEMORY; }
    c = 0;
    for(L j=0;j<ny;j++) for(L k=0;k<cnt[j];k++){ sL(z0)[c] = b[2*(lo[j]+k)+1]; sL(z1)[c] = j; c++; }
    free(b); free(lo); free(cnt);
    R 0;
}
/* lib_join_index_linear, h_libs.cc:1542-1575: for each y row, the first matching x row only */
static int32_t join_linear(OV z0, OV z1, OV x, OV y, int isEq){
    L nx = x->len, ny = y->len, size = 0;
#define KV(v,i) (v->typ==HPA_I ? (double)sI(v)[i] : v->typ==HPA_L ? (double)sL(v)[i] : sE(v)[i])
#define MATCH(i,j) (x->typ==HPA_L ? (isEq ? sL(y)[i]==sL(x)[j] : sL(y)[i]!=sL(x)[j]) : (isEq ? KV(y,i)==KV(x,j) : KV(y,i)!=KV(x,j)))
    if(!(x->typ==HPA_I || x->typ==HPA_L || x->typ==HPA_E)) R HPA_E_NOT_IMPL;
    for(L i=0;i<ny;i++) for(L j=0;j<nx;j++) if(MATCH(i,j)){ size++; break; }
    IV(z0, HPA_L, size); IV(z1, HPA_L, size);
    size = 0;
    for(L i=0;i<ny;i++) for(L j=0;j<nx;j++) if(MATCH(i,j)){ sL(z0)[size] = j; sL(z1)[size++] = i; break; }
#undef KV
#undef MATCH
    R 0;
}
static int32_t join_single(OV z, OV x, OV y, OV f){
    int op = op_from_symbol(f->q);
    if(op < 0) R HPA_E_DOMAIN;
    if(!(is_real(x->typ) && is_real(y->typ))) R HPA_E_NOT_IMPL;
    int32_t t = MAXV(x->typ, y->typ);
    O0 tx, ty; int32_t e;
    if((e=promote(&tx,x,t))) R e;
    if((e=promote(&ty,y,t))) R e;
    IV(z, HPA_G, 2);
    if(op==4 || op==5){
        if(tx.len >= 10){                          /* NUM_JOIN_LINEAR, common.h:152 */
            if(op != 4 || t != HPA_I) R HPA_E_NOT_IMPL;   /* v3 hash path: i32 eq only, h_hash_v3.cc:233,272-275 */
            R join_hash(vV(z,0), vV(z,1), &tx, &ty);
        }
        R join_linear(vV(z,0), vV(z,1), &tx, &ty, op==4);
    }
    R HPA_E_NOT_IMPL;
}
/* COMP, h_system.cc:791-794 */
#define JCOMP(op,a,b) ((op)<2 ? ((op)==0 ? (a)<(b) : (a)<=(b)) : (op)<4 ? ((op)==2 ? (a)>(b) : (a)>=(b)) : (op)<6 ? ((op)==4 ? (a)==(b) : (a)!=(b)) : 0)
static int64_t jval_i(OV v, L k){
    switch(v->typ){ case HPA_B: R sB(v)[k]; case HPA_J: R sJ(v)[k]; case HPA_H: R sH(v)[k]; case HPA_I: R sI(v)[k]; default: R sL(v)[k]; }
}
static double jval_e(OV v, L k){ R v->typ==HPA_E ? sE(v)[k] : v->typ==HPA_F ? (double)sF(v)[k] : (double)jval_i(v,k); }
static float  jval_f(OV v, L k){ R v->typ==HPA_F ? sF(v)[k] : (float)jval_i(v,k); }
/* compareOneColumn, h_system.cc:819-843: keep the pairs (i0[k], i1[k]) with x[i0[k]] op y[i1[k]],
 * both read in the larger of the two types, in pair order */
static int32_t join_compare_column(OV z0, OV z1, OV x, OV y, OV i0, OV i1, int op){
    L size = i0->len, tot = 0;
    int32_t tm = MAXV(x->typ, y->typ);
    uint8_t *t = (uint8_t*)malloc(size > 0 ? size : 1);
    if(!t) R HPA_E_MEMORY;
    for(L k=0;k<size;k++){
        L a = sL(i0)[k], b = sL(i1)[k];
        int r;
        if(tm == HPA_E){ double p = jval_e(x,a), q = jval_e(y,b); r = JCOMP(op,p,q); }
        else if(tm == HPA_F){ float p = jval_f(x,a), q = jval_f(y,b); r = JCOMP(op,p,q); }
        else { int64_t p = jval_i(x,a), q = jval_i(y,b); r = JCOMP(op,p,q); }
        t[k] = (uint8_t)r; tot += r;
    }
    if(!orc_initV(z0, HPA_L, tot) || !orc_initV(z1, HPA_L, tot)){ free(t); R HPA_E_MEMORY; }
    tot = 0;
    for(L k=0;k<size;k++) if(t[k]){ sL(z0)[tot] = sL(i0)[k]; sL(z1)[tot++] = sL(i1)[k]; }
    free(t);
    R 0;
}
/* pfnJoinIndexMultiple, h_primitive.cc:2095-2110: the first eq column (getFirstEqual,
 * :2088-2092; else the first neq, else column 0) through joinOneColumn -- the radix hash join on
 * x and y as given, no side swap and no linear path (h_system.cc:860-881) -- then every other
 * column in order through compareOneColumn (joinOtherColumns, :845-852).  The v3 hash join takes
 * i32 keys only; other first-column cases stop the reference (EP/TODO) and are NOT_IMPL here. */
static int32_t join_multi(OV z, OV x, OV y, OV f){
    L nc = f->len;
    if(x->typ != HPA_G || y->typ != HPA_G || x->len != nc || y->len != nc) R HPA_E_LENGTH;
    int ops[64];
    if(nc > 64) R HPA_E_NOT_IMPL;
    for(L i=0;i<nc;i++){ ops[i] = op_from_symbol(((int64_t*)f->g)[i]); if(ops[i] < 0) R HPA_E_DOMAIN; }
    L fx = -1;
    for(L i=0;i<nc && fx<0;i++) if(ops[i]==4) fx = i;
    for(L i=0;i<nc && fx<0;i++) if(ops[i]==5) fx = i;
    if(fx < 0) fx = 0;
    L nx = vV(x,0)->len, ny = vV(y,0)->len;
    for(L i=0;i<nc;i++){
        if(!is_real(vV(x,i)->typ) || !is_real(vV(y,i)->typ)) R HPA_E_NOT_IMPL;
        if(vV(x,i)->len != nx || vV(y,i)->len != ny) R HPA_E_LENGTH;
    }
    int32_t t0 = MAXV(vV(x,fx)->typ, vV(y,fx)->typ);
    if(ops[fx] != 4 || t0 != HPA_I) R HPA_E_NOT_IMPL;
    O0 tx, ty; int32_t e;
    if((e=promote(&tx, vV(x,fx), t0))) R e;
    if((e=promote(&ty, vV(y,fx), t0))) R e;
    O0 a0, a1;
    if((e=join_hash(&a0, &a1, &tx, &ty))) R e;
    for(L i=0;i<nc;i++){
        if(i == fx) continue;
        O0 b0, b1;
        if((e=join_compare_column(&b0, &b1, vV(x,i), vV(y,i), &a0, &a1, ops[i]))) R e;
        move_node(&a0, &b0); move_node(&a1, &b1);
    }
    IV(z, HPA_G, 2);
    move_node(vV(z,0), &a0); move_node(vV(z,1), &a1);
    R 0;
}
int32_t orc_pfnJoinIndex(OV z, OV x, OV y, OV f){  /* h_primitive.cc:2122-2155 */
    if(f->len == 1){
        if(x->len > y->len){
            int32_t e = join_single(z, y, x, f); if(e) R e;
            O0 a, b;                                     /* swap2, h_system.cc:47-56 */
            move_node(&a, vV(z,0)); move_node(&b, vV(z,1));
            move_node(vV(z,0), &b); move_node(vV(z,1), &a);
            R 0;
        }
        R join_single(z, x, y, f);
    }
    if(f->len > 1) R join_multi(z, x, y, f);
    R HPA_E_DOMAIN;
}

/* ------------------------------------------------------------------ index_of / member */
static int cmp_val_idx(const void *a, const void *b){
    const int64_t *p = (const int64_t*)a, *q = (const int64_t*)b;
    if(p[0] != q[0]) R p[0] < q[0] ? -1 : 1;
    R p[1] < q[1] ? -1 : (p[1] > q[1]);
}
/* keys compared by value bits for floats (the reference hashes float bytes, h_hash_common.h:111-117) */
static int64_t key_bits(OV v, L i){
    switch(v->typ){
        case HPA_B: R sB(v)[i]; case HPA_J: R sJ(v)[i]; case HPA_H: R sH(v)[i];
        case HPA_I: case HPA_D: case HPA_M: case HPA_U: case HPA_W: case HPA_T: R sI(v)[i];
        case HPA_L: case HPA_Z: case HPA_Q: R sL(v)[i];
        case HPA_F: { int32_t b; memcpy(&b, &sF(v)[i], 4); R b; }
        case HPA_E: { int64_t b; memcpy(&b, &sE(v)[i], 8); R b; }
        case HPA_C: R (unsigned char)sC(v)[i];
    }
    R 0;
}
/* binarySearch_I/L + searchOrdered, h_system.cc:715-745 */
static L bsearch_ref(OV x, L low, L high, int64_t val){
    for(;;){
        L mid = (low+high)>>1;
        int64_t xm = x->typ==HPA_I ? sI(x)[mid] : sL(x)[mid];
        if(xm == val) R mid;
        else if(low == mid) R -1;
        else if(xm > val) high = mid;
        else low = mid;
    }
}
int32_t orc_pfnIndexOf(OV z, OV x, OV y){          /* h_primitive.cc:1525-1579 */
    if(is_real(x->typ) && is_real(y->typ)){
        if(is_ordered(x)){
            O0 ty; int32_t e;
            if((e=promote(&ty, y, x->typ > y->typ ? x->typ : y->typ))) R e;
            L n = y->len; IV(z, HPA_L, n);
            for(L i=0;i<n;i++){
                int64_t v = ty.typ==HPA_I ? sI(&ty)[i] : ty.typ==HPA_L ? sL(&ty)[i] : 0;
                if(!is_int(ty.typ)){ R HPA_E_NOT_IMPL; }
                L k = x->len>0 ? bsearch_ref(x, 0, x->len, v) : -1;
                sL(z)[i] = k<0 ? x->len : k;
            }
            R 0;
        }
        int32_t t = MAXV(x->typ, y->typ);
        O0 tx, ty; int32_t e;
        if((e=promote(&tx,x,t))) R e;
        if((e=promote(&ty,y,t))) R e;
        L nx = tx.len, ny = ty.len;
        IV(z, HPA_L, ny);
        if(t == HPA_B){                            /* lib_index_of_B, h_libs.cc:309-314: flag[1] starts at 0 */
            L flag[2] = {-1, 0}; int c = 0;
            for(L i=0;i<nx;i++){ int s = sB(&tx)[i]; if(c<2 && flag[s]==-1){ flag[s]=i; c++; if(c==2) break; } }
            for(L i=0;i<ny;i++){ L q = flag[sB(&ty)[i]]; sL(z)[i] = q<0 ? nx : q; }
            R 0;
        }
        int64_t *b = (int64_t*)malloc(sizeof(int64_t)*2*(nx>0?nx:1));
        for(L i=0;i<nx;i++){ b[2*i] = key_bits(&tx,i); b[2*i+1] = i; }
        qsort(b, (size_t)nx, 2*sizeof(int64_t), cmp_val_idx);
        for(L i=0;i<ny;i++){
            int64_t key = key_bits(&ty,i); L a = 0, h = nx;
            while(a < h){ L m = (a+h)>>1; if(b[2*m] < key) a = m+1; else h = m; }
            sL(z)[i] = (a < nx && b[2*a] == key) ? b[2*a+1] : nx;
        }
        free(b);
        R 0;
    }
    if(x->typ == y->typ && x->typ == HPA_C){       /* lib_index_of_C, h_libs.cc:352-358 */
        L nx = x->len, ny = y->len, buf[256];
        for(int i=0;i<256;i++) buf[i] = -1;
        for(L i=0;i<nx;i++){ unsigned char c = sC(x)[i]; if(buf[c]==-1) buf[c]=i; }
        IV(z, HPA_L, ny);
        for(L i=0;i<ny;i++){ L q = buf[(unsigned char)sC(y)[i]]; sL(z)[i] = q==-1 ? nx : q; }
        R 0;
    }
    R HPA_E_NOT_IMPL;
}

static int cmp_str_ptr(const void *p, const void *q){ R strcmp(*(char *const*)p, *(char *const*)q); }
int32_t orc_pfnMember(OV z, OV y, OV x){           /* h_primitive.cc:2288-2336: z[i] = y[i] in x */
    if(is_real(x->typ) && is_real(y->typ)){
        int32_t t = MAXV(x->typ, y->typ);
        O0 tx, ty; int32_t e;
        if((e=promote(&tx,x,t))) R e;
        if((e=promote(&ty,y,t))) R e;
        L sLen = tx.len, vLen = ty.len;
        IV(z, HPA_B, vLen);
        if(t == HPA_B){                            /* lib_member_B, h_libs.cc:969-974 */
            int flag[2] = {0,0};
            for(L i=0;i<sLen;i++) flag[sB(&tx)[i]] = 1;
            for(L i=0;i<vLen;i++) sB(z)[i] = flag[sB(&ty)[i]];
            R 0;
        }
        if(t == HPA_J) R HPA_E_NOT_IMPL;           /* no MEMBER(J) case: EP exit, h_primitive.cc:2299-2305 */
        if(sLen <= 2){                             /* lib_member_fast1/2, h_libs.cc:899-944: value == */
            for(L i=0;i<vLen;i++){
                int hit = 0;
                for(L k=0;k<sLen;k++){
                    switch(t){
                        case HPA_F: hit |= sF(&ty)[i]==sF(&tx)[k]; break;
                        case HPA_E: hit |= sE(&ty)[i]==sE(&tx)[k]; break;
                        default: hit |= key_bits(&ty,i)==key_bits(&tx,k);
                    }
                }
                sB(z)[i] = (uint8_t)hit;
            }
            R 0;
        }
        int64_t *b = (int64_t*)malloc(sizeof(int64_t)*2*(sLen>0?sLen:1));
        for(L i=0;i<sLen;i++){ b[2*i] = key_bits(&tx,i); b[2*i+1] = i; }
        qsort(b, (size_t)sLen, 2*sizeof(int64_t), cmp_val_idx);
        for(L i=0;i<vLen;i++){
            int64_t key = key_bits(&ty,i); L a = 0, h = sLen;
            while(a < h){ L m = (a+h)>>1; if(b[2*m] < key) a = m+1; else h = m; }
            sB(z)[i] = (a < sLen && b[2*a] == key);
        }
        free(b);
        R 0;
    }
    if(x->typ == y->typ && x->typ == HPA_C){       /* lib_member_C, h_libs.cc:1005-1011 */
        uint8_t buf[256] = {0};
        for(L i=0;i<x->len;i++) buf[(unsigned char)sC(x)[i]] = 1;
        IV(z, HPA_B, y->len);
        for(L i=0;i<y->len;i++) sB(z)[i] = buf[(unsigned char)sC(y)[i]];
        R 0;
    }
    if(x->typ == y->typ && x->typ == HPA_S){       /* lib_member_S, h_libs.cc:1015-1018: equal strings */
        char **b = (char**)y->g;
        char **a = (char**)malloc(sizeof(char*)*(size_t)(x->len>0?x->len:1));
        memcpy(a, x->g, sizeof(char*)*(size_t)x->len);
        qsort(a, (size_t)x->len, sizeof(char*), cmp_str_ptr);
        IV(z, HPA_B, y->len);
        for(L i=0;i<y->len;i++) sB(z)[i] = bsearch(&b[i], a, (size_t)x->len, sizeof(char*), cmp_str_ptr) != NULL;
        free(a);
        R 0;
    }
    R HPA_E_NOT_IMPL;
}

/* ------------------------------------------------------------------ like, sub_string */
/* genLikeString (h_system.cc:319-334): ^ + pattern with % -> .*, _ -> ., . and \ escaped + $.
 * NULL for an empty pattern.  Bytes that are regular-expression operators stay operators in the
 * reference; the oracle (like the GPU library) takes only patterns without them. */
static char *gen_like_string(const char *in, L n){
    if(n <= 0) R NULL;
    char *o = (char*)malloc((size_t)(2*n + 3)); L c = 0;
    o[c++] = '^';
    for(L i=0;i<n;i++){
        if(in[i]=='%'){ o[c++]='.'; o[c++]='*'; }
        else if(in[i]=='.' || in[i]=='\\'){ o[c++]='\\'; o[c++]=in[i]; }
        else o[c++] = in[i]=='_' ? '.' : in[i];
    }
    o[c++] = '$'; o[c] = 0;
    R o;
}
/* backtracking match of the translated expression at s[i..n) (PCRE2 defaults: . and .* never
 * match a newline; $ matches at the end or before a final newline) */
static int re_here(const char *re, const char *s, L i, L n){
    if(re[0] == '$' && re[1] == 0) R i == n || (i == n-1 && s[i] == '\n');
    if(re[0] == '.' && re[1] == '*'){
        for(L k=i; ; k++){                         /* .* as short as possible, then longer */
            if(re_here(re+2, s, k, n)) R 1;
            if(k >= n || s[k] == '\n') R 0;
        }
    }
    if(i >= n) R 0;
    if(re[0] == '.'){ if(s[i] == '\n') R 0; R re_here(re+1, s, i+1, n); }
    if(re[0] == '\\'){ if(s[i] != re[1]) R 0; R re_here(re+2, s, i+1, n); }
    if(s[i] != re[0]) R 0;
    R re_here(re+1, s, i+1, n);
}
static int like_one(const char *re, const char *s, L n){ R re_here(re+1, s, 0, n); }   /* skip ^ */

/* pfnLike (h_primitive.cc:1670-1757), one pattern y (C, Q or S) over x (C: one string, Q, S) */
int32_t orc_pfnLike(OV z, OV x, OV y){
    int strish_x = x->typ==HPA_C || x->typ==HPA_Q || x->typ==HPA_S;
    int strish_y = y->typ==HPA_C || y->typ==HPA_Q || y->typ==HPA_S;
    if(!strish_x || !strish_y) R HPA_E_DOMAIN;
    if(!(y->len == 1 || y->typ == HPA_C)) R HPA_E_NOT_IMPL;
    char *pat; L pn;
    if(y->typ == HPA_C){ pat = (char*)calloc((size_t)y->len+1, 1); memcpy(pat, y->g, (size_t)y->len); }
    else if(y->typ == HPA_Q){ const char *t = sym_str(sL(y)[0]); if(!t) R HPA_E_DOMAIN; pat = strdup(t); }
    else pat = strdup(((char**)y->g)[0]);
    pn = (L)strlen(pat);
    for(L i=0;i<pn;i++) if(strchr("^$[|()?*+{", pat[i])){ free(pat); R HPA_E_NOT_IMPL; }
    char *re = gen_like_string(pat, pn);
    free(pat);
    if(!re) R HPA_E_NULL_VALUE;                    /* getLikePatten NULL -> E_NULL_VALUE */
    L n = x->typ == HPA_C ? 1 : x->len;
    IV(z, HPA_B, n);
    if(x->typ == HPA_C){
        L l = 0; while(l < x->len && ((char*)x->g)[l]) l++;
        sB(z)[0] = (uint8_t)like_one(re, (char*)x->g, l);
    } else if(x->typ == HPA_Q){
        for(L i=0;i<n;i++){ const char *t = sym_str(sL(x)[i]); if(!t){ free(re); R HPA_E_DOMAIN; }
            sB(z)[i] = (uint8_t)like_one(re, t, (L)strlen(t)); }
    } else {
        char **a = (char**)x->g;
        for(L i=0;i<n;i++) sB(z)[i] = (uint8_t)like_one(re, a[i], (L)strlen(a[i]));
    }
    free(re);
    R 0;
}

/* pfnSubString (h_primitive.cc:2468-2489): seg bytes from 1-based start, the room counted as
 * strlen - start, padded with spaces; start < 1 (a read before the string there) is E_DOMAIN */
int32_t orc_pfnSubString(OV z, OV x, OV y){
    if(x->typ != HPA_S || !is_int(y->typ)) R HPA_E_DOMAIN;
    if(y->len != 2) R HPA_E_NOT_IMPL;
    L start = (L)key_bits(y, 0), seg = (L)key_bits(y, 1);
    if(start < 1 || seg < 0) R HPA_E_DOMAIN;
    L n = x->len;
    IV(z, HPA_S, n);
    char **a = (char**)x->g, **o = (char**)z->g;
    for(L i=0;i<n;i++){
        char *p = (char*)heap_alloc(seg+1); if(!p) R HPA_E_MEMORY;
        L t0 = (L)strlen(a[i]), space = t0 - start, len = seg < space ? seg : space;
        if(len < 0) len = 0;
        for(L j=0;j<len;j++) p[j] = a[i][start+j-1];
        while(len < seg) p[len++] = ' ';
        p[seg] = 0; o[i] = p;
    }
    R 0;
}

/* ------------------------------------------------------------------ serializeV / readSerializeV */
/* h_io.cc:630-776: the V0 header as it sits in memory, then the payload per type (see
 * serialize_node); the reader re-makes symbol ids (getSymbol) and string storage (allocStrMem) */
static int ser_node(OV x, FILE *fp){
    if(fwrite(x, sizeof(O0), 1, fp) != 1) R HPA_E_GENERAL;
    L n = x->len;
    switch(x->typ){
        case HPA_Q: case HPA_S: {
            for(L i=0;i<n;i++){
                const char *t = x->typ==HPA_Q ? sym_str(sL(x)[i]) : ((char**)x->g)[i];
                if(!t) R HPA_E_DOMAIN;
                int64_t l = (int64_t)strlen(t); fwrite(&l, 8, 1, fp);
            }
            for(L i=0;i<n;i++){
                const char *t = x->typ==HPA_Q ? sym_str(sL(x)[i]) : ((char**)x->g)[i];
                fwrite(t, 1, strlen(t)+1, fp);
            }
            R 0;
        }
        case HPA_G: for(L i=0;i<n;i++){ int32_t e = ser_node(vV(x,i), fp); if(e) R e; } R 0;
        case HPA_A: {
            hpa_A0 a; memset(&a, 0, sizeof(a)); if(x->p) a = *(hpa_A0*)x->p;
            fwrite(&a, sizeof(a), 1, fp);
            int32_t e = ser_node(vV(x,0), fp); if(e) R e;
            R ser_node(vV(x,1), fp);
        }
        default:
            if(!is_basic(x->typ)) R HPA_E_TYPE;
            fwrite(x->g, 1, (size_t)type_size(x->typ, n), fp);   /* getTypeSize: a char vector + 1 */
            R 0;
    }
}
int32_t orc_serializeV(OV x, const char *path){
    FILE *fp = fopen(path, "wb"); if(!fp) R HPA_E_GENERAL;
    int32_t e = ser_node(x, fp);
    if(fclose(fp) != 0 && !e) e = HPA_E_GENERAL;
    R e;
}
static int rd_exact(void *p, size_t n, FILE *fp){ R (n == 0 || fread(p, 1, n, fp) == n) ? 0 : HPA_E_GENERAL; }
static int rd_node(OV x, FILE *fp, int depth){
    O0 h; int32_t e;
    if(depth > 64) R HPA_E_DOMAIN;
    if((e = rd_exact(&h, sizeof(h), fp))) R e;
    L n = h.len;
    if(n < 0) R HPA_E_DOMAIN;
    if(h.typ == HPA_A){
        x->typ = HPA_A; x->len = h.len;
        x->g = heap_alloc(2*sizeof(O0)); if(!x->g) R HPA_E_MEMORY;
        hpa_A0 *a = (hpa_A0*)heap_alloc(sizeof(hpa_A0)); if(!a) R HPA_E_MEMORY;
        if((e = rd_exact(a, sizeof(*a), fp))) R e;
        a->fkey = NULL;                           /* a pointer from another process */
        x->p = a;
        if((e = rd_node(vV(x,0), fp, depth+1))) R e;
        R rd_node(vV(x,1), fp, depth+1);
    }
    if(h.typ == HPA_G){
        x->typ = HPA_G; x->len = n; x->g2 = NULL;
        x->g = heap_alloc(sizeof(O0)*(size_t)(n>0?n:1)); if(!x->g) R HPA_E_MEMORY;
        for(L i=0;i<n;i++) if((e = rd_node(vV(x,i), fp, depth+1))) R e;
        R 0;
    }
    if(!is_basic(h.typ)) R HPA_E_TYPE;
    IV(x, h.typ, n);
    if(h.typ == HPA_Q || h.typ == HPA_S){
        int64_t *len = (int64_t*)malloc(sizeof(int64_t)*(size_t)(n>0?n:1));
        if((e = rd_exact(len, 8*(size_t)n, fp))){ free(len); R e; }
        for(L i=0;i<n;i++){
            char *t = heap_alloc(len[i]+1); if(!t){ free(len); R HPA_E_MEMORY; }
            if((e = rd_exact(t, (size_t)len[i]+1, fp))){ free(len); R e; }
            if(h.typ == HPA_Q) sL(x)[i] = orc_symbol(t); else ((char**)x->g)[i] = t;
        }
        free(len);
        R 0;
    }
    R rd_exact(x->g, (size_t)type_size(h.typ, n), fp);
}
int32_t orc_readSerializeV(OV z, const char *path){
    FILE *fp = fopen(path, "rb"); if(!fp) R HPA_E_GENERAL;
    int32_t e = rd_node(z, fp, 0);
    fclose(fp);
    R e;
}

/* ------------------------------------------------------------------ fused Q6 (opt/q6.h:31-48) */
double orc_q6_fused(int64_t n, const int32_t *ship, const double *disc, const int32_t *qty,
                    const double *price, int32_t lo, int32_t hi, double dlo, double dhi, int32_t qmax){
    double total = 0;
    #pragma omp parallel for simd reduction(+:total)
    for(L i=0;i<n;i++)
        if(ship[i]>=lo && ship[i]<hi && qty[i]<qmax && disc[i]>=dlo && disc[i]<=dhi)
            total += price[i]*disc[i];
    R total;
}
