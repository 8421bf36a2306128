// Element-wise primitives: compare, arith, logic, type promotion and gather (index).
// One pass over the operands, four consecutive elements per thread (one 16-byte load per
// 4-byte operand), grid-stride over a grid sized to the chip.
#include "hpa_internal.h"

#define H_EPSILON 1e-13   // backend/h_system.cc:9

// Operand: a vector (p != nullptr) or a broadcast scalar s (the reference's isOne(x) case).
template<typename T> struct Src {
    const T *p; T s;
    __device__ __forceinline__ T at(L i) const { return p ? p[i] : s; }
};
template<typename T> struct alignas(4 * sizeof(T)) Vec4 { T v[4]; };

template<typename T>
__device__ __forceinline__ void load4(const Src<T> &a, L i, T out[4]){
    if(a.p){ Vec4<T> q = *(const Vec4<T>*)(a.p + i); out[0]=q.v[0]; out[1]=q.v[1]; out[2]=q.v[2]; out[3]=q.v[3]; }
    else { out[0]=out[1]=out[2]=out[3]=a.s; }
}

template<typename TA, typename TB, typename TZ, typename F>
__global__ void __launch_bounds__(256) k_map2(Src<TA> a, Src<TB> b, TZ *z, L n, F f){
    L nq = (n + 3) >> 2;
    for(L q = (L)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (L)gridDim.x * blockDim.x){
        L i = q << 2;
        if(i + 4 <= n){
            TA x[4]; TB y[4]; Vec4<TZ> r;
            load4(a, i, x); load4(b, i, y);
            #pragma unroll
            for(int k=0;k<4;k++) r.v[k] = f(x[k], y[k]);
            *(Vec4<TZ>*)(z + i) = r;
        } else {
            for(L j=i;j<n;j++) z[j] = f(a.at(j), b.at(j));
        }
    }
}

template<typename TA, typename TB, typename TZ, typename F>
static int32_t launch_map2(hpa_ctx *c, const char *name, Src<TA> a, Src<TB> b, TZ *z, L n, F f){
    if(n <= 0) R 0;
    double bytes = (double)n * ((a.p ? sizeof(TA) : 0) + (b.p ? sizeof(TB) : 0) + sizeof(TZ));
    ProfScope ps(name, c->stream, bytes);
    L nq = (n + 3) >> 2;
    hipLaunchKernelGGL((k_map2<TA,TB,TZ,F>), dim3(grid_for(nq, 256, 1)), dim3(256), 0, c->stream, a, b, z, n, f);
    R check_launch(name);
}

// ---------------------------------------------------------------- compare (h_primitive.cc:1061-1200)
template<int OP> struct CmpInt {
    template<typename T> __device__ uint8_t operator()(T x, T y) const {
        return OP==0 ? x<y : OP==1 ? x<=y : OP==2 ? x>y : OP==3 ? x>=y : OP==4 ? x==y : x!=y;
    }
};
// compareFloat (h_primitive.cc:1061): three-way result with an absolute epsilon, then COMPFN.
template<int OP> struct CmpFlt {
    template<typename T> __device__ uint8_t operator()(T x, T y) const {
        int c = x<y ? ((double)(y-x) < H_EPSILON ? 0 : -1) : ((double)(x-y) < H_EPSILON ? 0 : 1);
        return OP==0 ? c<0 : OP==1 ? c<=0 : OP==2 ? c>0 : OP==3 ? c>=0 : OP==4 ? c==0 : c!=0;
    }
};

template<typename T>
static int32_t make_src(hpa_ctx *c, V x, int32_t typ, Src<T> *s){
    // typ: the compute type; x is promoted when its type differs
    s->p = nullptr; s->s = T();
    if(x->len == 1){
        // broadcast scalar: read on the host and convert as promoteValue would
        int64_t bits = 0;
        int32_t e = read_scalar(c, x, &bits, type_width(x->typ)); if(e) R e;
        double dv = 0; int64_t iv = 0;
        switch(x->typ){
            case HPA_B: iv = *(uint8_t*)&bits; dv = (double)iv; break;
            case HPA_J: iv = *(int8_t*)&bits;  dv = (double)iv; break;
            case HPA_H: iv = *(int16_t*)&bits; dv = (double)iv; break;
            case HPA_I: case HPA_D: case HPA_M: case HPA_U: case HPA_W: case HPA_T: iv = *(int32_t*)&bits; dv = (double)iv; break;
            case HPA_L: case HPA_Z: case HPA_Q: iv = bits; dv = (double)iv; break;
            case HPA_F: dv = *(float*)&bits; break;
            case HPA_E: dv = *(double*)&bits; break;
            case HPA_C: iv = *(char*)&bits; dv = (double)iv; break;
        }
        if(typ == HPA_F) s->s = (T)(x->typ == HPA_F ? *(float*)&bits : (float)iv);
        else if(typ == HPA_E) s->s = (T)(x->typ == HPA_E ? dv : x->typ == HPA_F ? (double)*(float*)&bits : (double)iv);
        else s->s = (T)iv;
        R 0;
    }
    const void *p = nullptr;
    int32_t e;
    if(x->typ == typ || !is_real(x->typ)) e = dev_view(c, x, &p);
    else e = promote_dev(c, x, typ, &p);
    if(e) R e;
    s->p = (const T*)p;
    R 0;
}

template<typename T, bool FLT>
static int32_t cmp_typed(hpa_ctx *c, V x, V y, int32_t typ, int op, uint8_t *z, L n){
    Src<T> a, b; int32_t e;
    if((e = make_src<T>(c, x, typ, &a))) R e;
    if((e = make_src<T>(c, y, typ, &b))) R e;
    if(FLT){
        switch(op){
            case 0: R launch_map2(c, "compare", a, b, z, n, CmpFlt<0>());
            case 1: R launch_map2(c, "compare", a, b, z, n, CmpFlt<1>());
            case 2: R launch_map2(c, "compare", a, b, z, n, CmpFlt<2>());
            case 3: R launch_map2(c, "compare", a, b, z, n, CmpFlt<3>());
            case 4: R launch_map2(c, "compare", a, b, z, n, CmpFlt<4>());
            default: R launch_map2(c, "compare", a, b, z, n, CmpFlt<5>());
        }
    }
    switch(op){
        case 0: R launch_map2(c, "compare", a, b, z, n, CmpInt<0>());
        case 1: R launch_map2(c, "compare", a, b, z, n, CmpInt<1>());
        case 2: R launch_map2(c, "compare", a, b, z, n, CmpInt<2>());
        case 3: R launch_map2(c, "compare", a, b, z, n, CmpInt<3>());
        case 4: R launch_map2(c, "compare", a, b, z, n, CmpInt<4>());
        default: R launch_map2(c, "compare", a, b, z, n, CmpInt<5>());
    }
}

extern "C" int32_t pfnCompare(hpa_V z, hpa_V x, hpa_V y, int32_t op){
    hpa_ctx *c = ctx_cur();
    if(!(is_comparable(x->typ) && is_comparable(y->typ))) R HPA_E_DOMAIN;
    if(!(x->len == 1 || y->len == 1 || x->len == y->len)) R HPA_E_LENGTH;
    if(op < 0 || op > 5) R HPA_E_DOMAIN;
    L n = x->len == 1 ? y->len : x->len;
    int32_t t;
    if(is_real(x->typ) && is_real(y->typ)) t = x->typ > y->typ ? x->typ : y->typ;
    else if(x->typ == y->typ){
        t = x->typ;
        // symbol order: strcmp of the names (compareSymbol), over the symbol table
        if(t == HPA_Q && op != 4 && op != 5) R compare_sym(c, z, x, y, op);
    }
    else R HPA_E_TYPE;
    void *dz; int32_t e;
    L mark = c->dcur;
    if((e = out_init(c, z, HPA_B, n, &dz))) R e;
    uint8_t *zz = (uint8_t*)dz;
    switch(t){
        case HPA_B: e = cmp_typed<uint8_t,false>(c, x, y, t, op, zz, n); break;
        case HPA_J: e = cmp_typed<int8_t,false>(c, x, y, t, op, zz, n); break;
        case HPA_H: e = cmp_typed<int16_t,false>(c, x, y, t, op, zz, n); break;
        case HPA_I: case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T:
                    e = cmp_typed<int32_t,false>(c, x, y, t, op, zz, n); break;
        case HPA_L: case HPA_Z: case HPA_Q:
                    e = cmp_typed<int64_t,false>(c, x, y, t, op, zz, n); break;
        case HPA_F: e = cmp_typed<float,true>(c, x, y, t, op, zz, n); break;
        case HPA_E: e = cmp_typed<double,true>(c, x, y, t, op, zz, n); break;
        case HPA_C: e = cmp_typed<char,false>(c, x, y, t, op, zz, n); break;
        default: e = HPA_E_NOT_IMPL;
    }
    if(e) R e;
    if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
    R e;
}
extern "C" int32_t pfnLt (hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,0); }
extern "C" int32_t pfnLeq(hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,1); }
extern "C" int32_t pfnGt (hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,2); }
extern "C" int32_t pfnGeq(hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,3); }
extern "C" int32_t pfnEq (hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,4); }
extern "C" int32_t pfnNeq(hpa_V z, hpa_V x, hpa_V y){ R pfnCompare(z,x,y,5); }

// ---------------------------------------------------------------- arith (h_primitive.cc:1227-1309)
// ARITH2 in the promoted type; bool results are (x op y) != 0 (C++ bool conversion).
template<int OP> struct Arith {
    template<typename T> __device__ T operator()(T x, T y) const {
        return (T)(OP==0 ? x+y : OP==1 ? x-y : OP==2 ? x*y : x/y);
    }
};
template<int OP> struct ArithB {
    __device__ uint8_t operator()(uint8_t x, uint8_t y) const {
        int a = x, b = y; int r = OP==0 ? a+b : OP==1 ? a-b : OP==2 ? a*b : a/b; return (uint8_t)(r != 0);
    }
};
// integer ops wrap (two's complement), as the reference's compiled C does on x86
template<int OP> struct ArithWrap {
    template<typename T> __device__ T operator()(T x, T y) const {
        typedef typename std::make_unsigned<T>::type U;
        if(OP==0) return (T)((U)x + (U)y);
        if(OP==1) return (T)((U)x - (U)y);
        if(OP==2) return (T)((U)x * (U)y);
        if(y == (T)-1) return (T)(0 - (U)x);
        return (T)(x / y);
    }
};

template<typename T, typename F>
static int32_t arith_run(hpa_ctx *c, V x, V y, int32_t t, T *z, L n, F f){
    Src<T> a, b; int32_t e;
    if((e = make_src<T>(c, x, t, &a))) R e;
    if((e = make_src<T>(c, y, t, &b))) R e;
    R launch_map2(c, "arith", a, b, z, n, f);
}
template<typename T>
static int32_t arith_int(hpa_ctx *c, V x, V y, int32_t t, T *z, L n, int op){
    switch(op){
        case 0: R arith_run<T>(c, x, y, t, z, n, ArithWrap<0>());
        case 1: R arith_run<T>(c, x, y, t, z, n, ArithWrap<1>());
        case 2: R arith_run<T>(c, x, y, t, z, n, ArithWrap<2>());
        default: R arith_run<T>(c, x, y, t, z, n, ArithWrap<3>());
    }
}
template<typename T>
static int32_t arith_flt(hpa_ctx *c, V x, V y, int32_t t, T *z, L n, int op){
    switch(op){
        case 0: R arith_run<T>(c, x, y, t, z, n, Arith<0>());
        case 1: R arith_run<T>(c, x, y, t, z, n, Arith<1>());
        case 2: R arith_run<T>(c, x, y, t, z, n, Arith<2>());
        default: R arith_run<T>(c, x, y, t, z, n, Arith<3>());
    }
}

// checkZero (backend/h_system.cc:254-267) on the device: any element == 0
template<typename T>
__global__ void k_any_zero(const T *p, L n, int *flag){
    int f = 0;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) f |= (p[i] == (T)0);
    if(__any(f) && (threadIdx.x & 63) == 0 && !*(volatile int*)flag) atomicOr(flag, 1);
}
static int32_t check_zero(hpa_ctx *c, V y, bool *found){
    *found = false;
    if(y->len == 0) R 0;
    if(y->len == 1){
        int64_t bits = 0; int32_t e = read_scalar(c, y, &bits, type_width(y->typ)); if(e) R e;
        switch(y->typ){
            case HPA_F: *found = *(float*)&bits == 0; break;
            case HPA_E: *found = *(double*)&bits == 0; break;
            case HPA_B: case HPA_J: *found = (bits & 0xff) == 0; break;
            case HPA_H: *found = (bits & 0xffff) == 0; break;
            case HPA_I: *found = (int32_t)bits == 0; break;
            default: *found = bits == 0;
        }
        R 0;
    }
    const void *p; int32_t e = dev_view(c, y, &p); if(e) R e;
    int *flag = (int*)dalloc(c, 16); if(!flag) R HPA_E_MEMORY;
    hipMemsetAsync(flag, 0, 4, c->stream);
    L n = y->len; dim3 g(grid_for(n, 256, 4)), b(256);
    switch(y->typ){
        case HPA_B: hipLaunchKernelGGL(k_any_zero<uint8_t>, g, b, 0, c->stream, (const uint8_t*)p, n, flag); break;
        case HPA_J: hipLaunchKernelGGL(k_any_zero<int8_t>, g, b, 0, c->stream, (const int8_t*)p, n, flag); break;
        case HPA_H: hipLaunchKernelGGL(k_any_zero<int16_t>, g, b, 0, c->stream, (const int16_t*)p, n, flag); break;
        case HPA_I: hipLaunchKernelGGL(k_any_zero<int32_t>, g, b, 0, c->stream, (const int32_t*)p, n, flag); break;
        case HPA_L: hipLaunchKernelGGL(k_any_zero<int64_t>, g, b, 0, c->stream, (const int64_t*)p, n, flag); break;
        case HPA_F: hipLaunchKernelGGL(k_any_zero<float>, g, b, 0, c->stream, (const float*)p, n, flag); break;
        case HPA_E: hipLaunchKernelGGL(k_any_zero<double>, g, b, 0, c->stream, (const double*)p, n, flag); break;
    }
    int h = 0; if((e = d2h(c, &h, flag, 4))) R e;
    *found = h != 0;
    R 0;
}

extern "C" int32_t pfnArith(hpa_V z, hpa_V x, hpa_V y, int32_t op){
    hpa_ctx *c = ctx_cur();
    if(!(is_real(x->typ) && is_real(y->typ))) R HPA_E_DOMAIN;
    if(!(x->len == 1 || y->len == 1 || x->len == y->len)) R HPA_E_LENGTH;
    if(op < 0 || op > 3) R HPA_E_DOMAIN;
    L mark0 = c->dcur;
    if(op == 3){ bool zf; int32_t e = check_zero(c, y, &zf); if(e) R e; arena_set(c, mark0); if(zf) R HPA_E_DIV_ZERO; }
    L n = x->len == 1 ? y->len : x->len;
    int32_t t = x->typ > y->typ ? x->typ : y->typ;
    void *dz; int32_t e;
    L mark = c->dcur;
    if((e = out_init(c, z, t, n, &dz))) R e;
    switch(t){
        case HPA_B:
            switch(op){
                case 0: e = arith_run<uint8_t>(c, x, y, t, (uint8_t*)dz, n, ArithB<0>()); break;
                case 1: e = arith_run<uint8_t>(c, x, y, t, (uint8_t*)dz, n, ArithB<1>()); break;
                case 2: e = arith_run<uint8_t>(c, x, y, t, (uint8_t*)dz, n, ArithB<2>()); break;
                default: e = arith_run<uint8_t>(c, x, y, t, (uint8_t*)dz, n, ArithB<3>()); break;
            } break;
        case HPA_J: e = arith_int<int8_t>(c, x, y, t, (int8_t*)dz, n, op); break;
        case HPA_H: e = arith_int<int16_t>(c, x, y, t, (int16_t*)dz, n, op); break;
        case HPA_I: e = arith_int<int32_t>(c, x, y, t, (int32_t*)dz, n, op); break;
        case HPA_L: e = arith_int<int64_t>(c, x, y, t, (int64_t*)dz, n, op); break;
        case HPA_F: e = arith_flt<float>(c, x, y, t, (float*)dz, n, op); break;
        case HPA_E: e = arith_flt<double>(c, x, y, t, (double*)dz, n, op); break;
    }
    if(e) R e;
    if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
    R e;
}
extern "C" int32_t pfnPlus (hpa_V z, hpa_V x, hpa_V y){ R pfnArith(z,x,y,0); }
extern "C" int32_t pfnMinus(hpa_V z, hpa_V x, hpa_V y){ R pfnArith(z,x,y,1); }
extern "C" int32_t pfnMul  (hpa_V z, hpa_V x, hpa_V y){ R pfnArith(z,x,y,2); }
extern "C" int32_t pfnDiv  (hpa_V z, hpa_V x, hpa_V y){ R pfnArith(z,x,y,3); }

// ---------------------------------------------------------------- logic (h_primitive.cc:1315-1357)
// logicAnd/Or/Xor are bitwise on 0/1; logicNand/Nor return ~(..) converted to bool, which is
// true for both 0 and 1 inputs (backend/h_system.cc:898-916).
template<int W> struct Logic {
    __device__ uint8_t operator()(uint8_t x, uint8_t y) const {
        return W==0 ? (uint8_t)(x & y) : W==2 ? (uint8_t)(x | y) : W==4 ? (uint8_t)(x ^ y) : (uint8_t)1;
    }
};
static int32_t logic(hpa_V z, hpa_V x, hpa_V y, int w){
    hpa_ctx *c = ctx_cur();
    if(!(x->typ == HPA_B && y->typ == HPA_B)) R HPA_E_DOMAIN;
    if(!(x->len == 1 || y->len == 1 || x->len == y->len)) R HPA_E_LENGTH;
    L n = x->len == 1 ? y->len : x->len;
    void *dz; int32_t e;
    L mark = c->dcur;
    if((e = out_init(c, z, HPA_B, n, &dz))) R e;
    Src<uint8_t> a, b;
    if((e = make_src<uint8_t>(c, x, HPA_B, &a))) R e;
    if((e = make_src<uint8_t>(c, y, HPA_B, &b))) R e;
    uint8_t *zz = (uint8_t*)dz;
    switch(w){
        case 0: e = launch_map2(c, "logic", a, b, zz, n, Logic<0>()); break;
        case 2: e = launch_map2(c, "logic", a, b, zz, n, Logic<2>()); break;
        case 4: e = launch_map2(c, "logic", a, b, zz, n, Logic<4>()); break;
        default: e = launch_map2(c, "logic", a, b, zz, n, Logic<1>()); break;
    }
    if(e) R e;
    if(is_inline(z)){ e = commit_inline(c, z, dz); arena_set(c, mark); }
    R e;
}
extern "C" int32_t pfnAnd (hpa_V z, hpa_V x, hpa_V y){ R logic(z,x,y,0); }
extern "C" int32_t pfnNand(hpa_V z, hpa_V x, hpa_V y){ R logic(z,x,y,1); }
extern "C" int32_t pfnOr  (hpa_V z, hpa_V x, hpa_V y){ R logic(z,x,y,2); }
extern "C" int32_t pfnNor (hpa_V z, hpa_V x, hpa_V y){ R logic(z,x,y,3); }
extern "C" int32_t pfnXor (hpa_V z, hpa_V x, hpa_V y){ R logic(z,x,y,4); }

// ---------------------------------------------------------------- promotion (h_system.cc:168-252)
template<typename TS, typename TD>
__global__ void __launch_bounds__(256) k_convert(const TS *s, TD *d, L n){
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x) d[i] = (TD)s[i];
}
template<typename TS>
static int32_t convert_from(hpa_ctx *c, const TS *s, void *d, int32_t to, L n){
    dim3 g(grid_for(n, 256, 4)), b(256);
    switch(to){
        case HPA_J: hipLaunchKernelGGL((k_convert<TS,int8_t>), g, b, 0, c->stream, s, (int8_t*)d, n); break;
        case HPA_H: hipLaunchKernelGGL((k_convert<TS,int16_t>), g, b, 0, c->stream, s, (int16_t*)d, n); break;
        case HPA_I: hipLaunchKernelGGL((k_convert<TS,int32_t>), g, b, 0, c->stream, s, (int32_t*)d, n); break;
        case HPA_L: hipLaunchKernelGGL((k_convert<TS,int64_t>), g, b, 0, c->stream, s, (int64_t*)d, n); break;
        case HPA_F: hipLaunchKernelGGL((k_convert<TS,float>), g, b, 0, c->stream, s, (float*)d, n); break;
        case HPA_E: hipLaunchKernelGGL((k_convert<TS,double>), g, b, 0, c->stream, s, (double*)d, n); break;
        default: R HPA_E_TYPE;
    }
    R check_launch("promote");
}
int32_t promote_dev(hpa_ctx *c, V x, int32_t to, const void **p){
    const void *s; int32_t e;
    if((e = dev_view(c, x, &s))) R e;
    if(x->typ == to){ *p = s; R 0; }
    if(!(is_real(x->typ) && is_real(to) && to > x->typ)) R HPA_E_TYPE;
    L n = x->len;
    void *d = dalloc(c, type_bytes(to, n)); if(!d) R HPA_E_MEMORY;
    *p = d;
    if(n == 0) R 0;
    switch(x->typ){
        case HPA_B: R convert_from<uint8_t>(c, (const uint8_t*)s, d, to, n);
        case HPA_J: R convert_from<int8_t>(c, (const int8_t*)s, d, to, n);
        case HPA_H: R convert_from<int16_t>(c, (const int16_t*)s, d, to, n);
        case HPA_I: R convert_from<int32_t>(c, (const int32_t*)s, d, to, n);
        case HPA_L: R convert_from<int64_t>(c, (const int64_t*)s, d, to, n);
        case HPA_F: R convert_from<float>(c, (const float*)s, d, to, n);
    }
    R HPA_E_TYPE;
}

// ---------------------------------------------------------------- gather / index (h_primitive.cc:43-94)
template<typename T>
__global__ void __launch_bounds__(256) k_gather(const T *src, const int64_t *idx, T *dst, L n, L nsrc, int *bad){
    int f = 0;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        int64_t k = idx[i];
        if(k < 0 || k >= nsrc){ f = 1; continue; }
        dst[i] = src[k];
    }
    if(bad && __any(f) && (threadIdx.x & 63) == 0 && !*(volatile int*)bad) atomicOr(bad, 1);
}
static int32_t gather_checked(hpa_ctx *c, void *dst, const void *src, const int64_t *idx, L n, L nsrc, L w, int *bad){
    if(n <= 0) R 0;
    ProfScope ps("gather", c->stream, (double)n * (8 + 2 * w));
    dim3 g(grid_for(n, 256, 4)), b(256);
    switch(w){
        case 1: hipLaunchKernelGGL(k_gather<uint8_t>, g, b, 0, c->stream, (const uint8_t*)src, idx, (uint8_t*)dst, n, nsrc, bad); break;
        case 2: hipLaunchKernelGGL(k_gather<uint16_t>, g, b, 0, c->stream, (const uint16_t*)src, idx, (uint16_t*)dst, n, nsrc, bad); break;
        case 4: hipLaunchKernelGGL(k_gather<uint32_t>, g, b, 0, c->stream, (const uint32_t*)src, idx, (uint32_t*)dst, n, nsrc, bad); break;
        case 8: hipLaunchKernelGGL(k_gather<uint64_t>, g, b, 0, c->stream, (const uint64_t*)src, idx, (uint64_t*)dst, n, nsrc, bad); break;
        default: R HPA_E_NOT_IMPL;
    }
    R check_launch("gather");
}
int32_t gather_dev(hpa_ctx *c, void *dst, const void *src, const int64_t *idx, L n, L w){
    R gather_checked(c, dst, src, idx, n, INT64_MAX, w, nullptr);
}

// A few host-held indices (a limit's range, say) travel in the launch arguments: checked on the
// host, no copy to the device and no flag to read back
#define GATHER_ARG_MAX 64
struct GatherArgs { int64_t idx[GATHER_ARG_MAX]; };
__global__ void k_gather_args(const char *src, GatherArgs a, int n, int w, char *dst){
    const int t = threadIdx.x;
    if(t >= n) return;
    const char *s = src + a.idx[t] * w; char *d = dst + (L)t * w;
    for(int b = 0; b < w; b++) d[b] = s[b];
}

extern "C" int32_t pfnIndex(hpa_V z, hpa_V x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(!is_int(y->typ)) R HPA_E_DOMAIN;
    L nz = y->len, nx = x->len;
    if(is_basic(x->typ) && y->typ == HPA_L && nz > 1 && nz <= GATHER_ARG_MAX && nx > 0 && !is_device(y->g) && is_device(x->g)){
        GatherArgs ga;
        memcpy(ga.idx, y->g, (size_t)(8 * nz));
        for(L i = 0; i < nz; i++) if(ga.idx[i] < 0 || ga.idx[i] >= nx) R HPA_E_INDEX;
        const L w = type_width(x->typ);
        void *dst = dalloc(c, type_bytes(x->typ, nz)); if(!dst) R HPA_E_MEMORY;
        hipLaunchKernelGGL(k_gather_args, dim3(1), dim3(GATHER_ARG_MAX), 0, c->stream, (const char*)x->g, ga, (int)nz, (int)w, (char*)dst);
        int32_t e; if((e = check_launch("gather"))) R e;
        z->typ = x->typ; z->len = nz; z->g = (char*)dst;
        if(x->typ == HPA_L){ auto it = c->bounds.find(x->g); if(it != c->bounds.end()) note_bound(c, dst, it->second); }
        R 0;
    }
    if(is_basic(x->typ)){
        if(x->typ == HPA_S || x->typ == HPA_X) { /* host pointers / complex: width-8 copy still valid */ }
        L w = type_width(x->typ);
        L mark = c->dcur;
        const void *py; int32_t e;
        if((e = promote_dev(c, y, HPA_L, &py))) R e;
        const void *px = nullptr;
        if(nx > 0 && (e = dev_view(c, x, &px))) R e;
        // the keys of a group of this very i32 column: the group kept each key's value
        if(x->typ == HPA_I && y->typ == HPA_L && nz > 1){
            const void *gv = group_vals(c, py, px, nx);
            if(gv){ arena_set(c, mark); z->typ = HPA_I; z->len = nz; z->g = (char*)gv; R 0; }
        }
        // index vectors made here (order, group, join, where) are known to be in range
        const bool known = y->typ == HPA_L && within_bound(c, py, nx);
        int *bad = nullptr;
        if(!known){ bad = (int*)dalloc(c, 16); if(!bad) R HPA_E_MEMORY; hipMemsetAsync(bad, 0, 4, c->stream); }
        void *dst = dalloc(c, type_bytes(x->typ, nz > 0 ? nz : 1)); if(!dst) R HPA_E_MEMORY;
        if(nz > 0 && nx == 0){ arena_set(c, mark); R HPA_E_INDEX; }
        if((e = gather_checked(c, dst, px, (const int64_t*)py, nz, nx, w, bad))) R e;
        if(!known){
            int h = 0; if((e = d2h(c, &h, bad, 4))) R e;
            if(h){ arena_set(c, mark); R HPA_E_INDEX; }
        }
        z->typ = x->typ; z->len = nz;
        if(nz <= 1){
            z->g = (char*)&z->b; z->i64 = 0;
            if(nz == 1 && (e = d2h(c, &z->b, dst, w))) R e;
            arena_set(c, mark);
        } else {
            z->g = (char*)dst;
            // gathering a bounded index vector keeps its bound
            if(x->typ == HPA_L){ auto it = c->bounds.find(px); if(it != c->bounds.end()) note_bound(c, dst, it->second); }
        }
        R 0;
    }
    if(x->typ == HPA_G){
        if(nz == 1){
            int64_t k = 0; int32_t e = 0;
            switch(y->typ){
                case HPA_B: { uint8_t v; e = read_scalar(c, y, &v, 1); k = v; } break;
                case HPA_J: { int8_t v; e = read_scalar(c, y, &v, 1); k = v; } break;
                case HPA_H: { int16_t v; e = read_scalar(c, y, &v, 2); k = v; } break;
                case HPA_I: { int32_t v; e = read_scalar(c, y, &v, 4); k = v; } break;
                case HPA_L: { int64_t v; e = read_scalar(c, y, &v, 8); k = v; } break;
            }
            if(e) R e;
            if(k < 0 || k >= nx) R HPA_E_INDEX;
            *z = *(((V)x->g) + k);                 // copyV alias (h_system.cc:70-73)
            R 0;
        }
        R HPA_E_NOT_IMPL;
    }
    R HPA_E_NOT_IMPL;
}

// Gather of several basic columns by one index vector in one launch: z[k] = index(x[k], y).
#define GM_MAXC 16
struct GatherCols { int n; int w[GM_MAXC]; const char *src[GM_MAXC]; char *dst[GM_MAXC]; L nsrc[GM_MAXC]; };
__global__ void __launch_bounds__(256) k_gather_many(GatherCols g, const int64_t *idx, L n, int *bad){
    int f = 0;
    for(L i = (L)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (L)gridDim.x * blockDim.x){
        const int64_t k = idx[i];
        for(int c = 0; c < g.n; c++){
            if(k < 0 || k >= g.nsrc[c]){ f = 1; continue; }
            switch(g.w[c]){
                case 1: ((uint8_t*)g.dst[c])[i] = ((const uint8_t*)g.src[c])[k]; break;
                case 2: ((uint16_t*)g.dst[c])[i] = ((const uint16_t*)g.src[c])[k]; break;
                case 4: ((uint32_t*)g.dst[c])[i] = ((const uint32_t*)g.src[c])[k]; break;
                default: ((uint64_t*)g.dst[c])[i] = ((const uint64_t*)g.src[c])[k]; break;
            }
        }
    }
    if(bad && __any(f) && (threadIdx.x & 63) == 0 && !*(volatile int*)bad) atomicOr(bad, 1);
}

// The multi-column peephole of index (the gathers after a group or an order, as in
// src/horseir/v1/compile-static/cur/q1.h and q3.h): each result is pfnIndex(z[k], x[k], y),
// with one kernel and one bounds check for all columns.
extern "C" int32_t hpaIndexMany(int32_t nx, hpa_V *z, hpa_V *x, hpa_V y){
    hpa_ctx *c = ctx_cur();
    if(!is_int(y->typ)) R HPA_E_DOMAIN;
    if(nx < 1 || nx > GM_MAXC) R HPA_E_DOMAIN;
    for(int k = 0; k < nx; k++) if(!is_basic(x[k]->typ)) R HPA_E_NOT_IMPL;
    const L nz = y->len;
    L mark = c->dcur;
    const void *py; int32_t e;
    if((e = promote_dev(c, y, HPA_L, &py))) R e;
    GatherCols g; memset(&g, 0, sizeof(g)); g.n = nx;
    bool known = y->typ == HPA_L;
    for(int k = 0; k < nx; k++){
        const void *px = nullptr;
        if(x[k]->len > 0 && (e = dev_view(c, x[k], &px))){ arena_set(c, mark); R e; }
        if(nz > 0 && x[k]->len == 0){ arena_set(c, mark); R HPA_E_INDEX; }
        g.src[k] = (const char*)px; g.w[k] = (int)type_width(x[k]->typ); g.nsrc[k] = x[k]->len;
        known = known && within_bound(c, py, x[k]->len);
    }
    for(int k = 0; k < nx; k++){
        g.dst[k] = (char*)dalloc(c, type_bytes(x[k]->typ, nz > 0 ? nz : 1));
        if(!g.dst[k]){ arena_set(c, mark); R HPA_E_MEMORY; }
    }
    int *bad = nullptr;
    if(!known){ bad = (int*)dalloc(c, 16); if(!bad){ arena_set(c, mark); R HPA_E_MEMORY; } hipMemsetAsync(bad, 0, 4, c->stream); }
    if(nz > 0){
        double by = 8.0 * nz; for(int k = 0; k < nx; k++) by += 2.0 * g.w[k] * nz;
        ProfScope ps("gather_many", c->stream, by);
        hipLaunchKernelGGL(k_gather_many, dim3(grid_for(nz, 256, 4)), dim3(256), 0, c->stream, g, (const int64_t*)py, nz, bad);
        if((e = check_launch("gather_many"))) R e;
    }
    if(!known){
        int h = 0; if((e = d2h(c, &h, bad, 4))) R e;
        if(h){ arena_set(c, mark); R HPA_E_INDEX; }
    }
    // one read back for every inline (single-row) result
    if(nz == 1){
        uint64_t tmp[GM_MAXC];
        for(int k = 0; k < nx; k++)
            if(hipMemcpyAsync(&tmp[k], g.dst[k], (size_t)g.w[k], hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        for(int k = 0; k < nx; k++){
            V zk = z[k]; zk->typ = x[k]->typ; zk->len = 1; zk->g = (char*)&zk->b; zk->i64 = 0;
            memcpy(&zk->b, &tmp[k], (size_t)g.w[k]);
        }
        arena_set(c, mark);
        R 0;
    }
    for(int k = 0; k < nx; k++){
        V zk = z[k]; zk->typ = x[k]->typ; zk->len = nz;
        if(nz == 0){ zk->g = (char*)&zk->b; zk->i64 = 0; }
        else zk->g = g.dst[k];
    }
    R 0;
}

extern "C" int32_t pfnLen(hpa_V z, hpa_V x){       // h_primitive.cc:656-662
    z->typ = HPA_L; z->len = 1; z->g = (char*)&z->b;
    z->i64 = x->len;
    R 0;
}
