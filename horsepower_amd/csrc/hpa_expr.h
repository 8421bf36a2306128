// Device evaluator for the postfix expressions of the fused loops (hpa_expr in the public
// header).  The emitted C loops of the reference compare floats exactly and compute with C's
// usual arithmetic conversions (src/horseir/v1/compile-static/fuse.h:13-26,
// optimizer/elementwise.cc:245-276); values here are int64 or float64 per instruction, with
// each instruction's result kind fixed on the host.
#pragma once
#include "hpa_internal.h"

#define EXPR_MAX_INS  48
#define EXPR_MAX_COLS 12
#define EXPR_MAX_DEPTH 12

struct DevIns { int16_t op; int16_t arg; int8_t kf; int8_t af; int8_t bf; int8_t pad; int64_t imm; };
struct DevExpr { int n; int depth; DevIns ins[EXPR_MAX_INS]; };
struct DevCols { int n; int typ[EXPR_MAX_COLS]; const void *p[EXPR_MAX_COLS]; };

union Slot { int64_t i; double f; };

__device__ __forceinline__ Slot load_col(const DevCols &x, int k, L i){
    Slot s; const void *p = x.p[k];
    switch(x.typ[k]){
        case HPA_B: s.i = ((const uint8_t*)p)[i]; break;
        case HPA_J: s.i = ((const int8_t*)p)[i]; break;
        case HPA_H: s.i = ((const int16_t*)p)[i]; break;
        case HPA_I: case HPA_D: case HPA_M: case HPA_U: case HPA_W: case HPA_T: s.i = ((const int32_t*)p)[i]; break;
        case HPA_L: case HPA_Z: case HPA_Q: s.i = ((const int64_t*)p)[i]; break;
        case HPA_F: s.f = (double)((const float*)p)[i]; break;
        case HPA_E: s.f = ((const double*)p)[i]; break;
        case HPA_C: s.i = ((const char*)p)[i]; break;
        default: s.i = 0;
    }
    return s;
}

// Evaluate the program at row i; returns the top of the stack (kind in e.ins[n-1].kf).
__device__ __forceinline__ Slot eval_expr(const DevExpr &e, const DevCols &x, L i){
    Slot st[EXPR_MAX_DEPTH];
    int sp = 0;
    for(int k = 0; k < e.n; k++){
        const DevIns in = e.ins[k];
        switch(in.op){
            case HPA_OP_COL: st[sp++] = load_col(x, in.arg, i); break;
            case HPA_OP_CONST_I: st[sp].i = in.imm; sp++; break;
            case HPA_OP_CONST_F: st[sp].i = in.imm; sp++; break;      // bits of the double
            case HPA_OP_NOT: st[sp-1].i = in.af ? (st[sp-1].f == 0.0) : (st[sp-1].i == 0); break;
            case HPA_OP_NEG: if(in.af) st[sp-1].f = -st[sp-1].f; else st[sp-1].i = (int64_t)(0ull - (uint64_t)st[sp-1].i); break;
            default: {
                Slot b = st[--sp], a = st[sp-1], r;
                bool fl = in.af || in.bf;
                double af = in.af ? a.f : (double)a.i, bf = in.bf ? b.f : (double)b.i;
                switch(in.op){
                    case HPA_OP_ADD: if(fl) r.f = af + bf; else r.i = (int64_t)((uint64_t)a.i + (uint64_t)b.i); break;
                    case HPA_OP_SUB: if(fl) r.f = af - bf; else r.i = (int64_t)((uint64_t)a.i - (uint64_t)b.i); break;
                    case HPA_OP_MUL: if(fl) r.f = af * bf; else r.i = (int64_t)((uint64_t)a.i * (uint64_t)b.i); break;
                    case HPA_OP_DIV: if(fl) r.f = af / bf; else r.i = b.i == 0 ? 0 : (b.i == -1 ? (int64_t)(0ull - (uint64_t)a.i) : a.i / b.i); break;
                    case HPA_OP_LT:  r.i = fl ? af <  bf : a.i <  b.i; break;
                    case HPA_OP_LEQ: r.i = fl ? af <= bf : a.i <= b.i; break;
                    case HPA_OP_GT:  r.i = fl ? af >  bf : a.i >  b.i; break;
                    case HPA_OP_GEQ: r.i = fl ? af >= bf : a.i >= b.i; break;
                    case HPA_OP_EQ:  r.i = fl ? af == bf : a.i == b.i; break;
                    case HPA_OP_NEQ: r.i = fl ? af != bf : a.i != b.i; break;
                    case HPA_OP_AND: r.i = (in.af ? a.f != 0.0 : a.i != 0) && (in.bf ? b.f != 0.0 : b.i != 0); break;
                    case HPA_OP_OR:  r.i = (in.af ? a.f != 0.0 : a.i != 0) || (in.bf ? b.f != 0.0 : b.i != 0); break;
                    default: r.i = 0;
                }
                st[sp-1] = r;
            }
        }
    }
    return st[sp > 0 ? sp - 1 : 0];
}

__device__ __forceinline__ bool expr_true(const DevExpr &e, const Slot &s){
    return e.ins[e.n - 1].kf ? s.f != 0.0 : s.i != 0;
}

// Host: type-check a public expression against its columns and build the device form.
int32_t expr_compile(const hpa_expr *e, int nx, const int32_t *coltyp, DevExpr *out);
// Host: device views of the input columns (all of equal length n).
int32_t cols_view(hpa_ctx *c, int nx, V *x, DevCols *cols, L *n);
// Generated-code versions of the FA reduction and FE map (nonzero: not available, the caller
// interprets the expressions instead).
int32_t jit_fused_reduce(hpa_ctx *c, int kind, int nx, V *x, const hpa_expr *cond, const hpa_expr *body,
                         const DevCols &cols, L n, void *part, int G);
int32_t jit_fused_map(hpa_ctx *c, int otyp, int nx, V *x, const hpa_expr *ex, const DevCols &cols, L n, void *dz);
// whether run-time generated kernels are in use (hipRTC present, not switched off)
bool jit_enabled();
// count pass of a compress by a condition over cols: selection bits per 16 rows into mb, tile counts
int32_t jit_compress_count(hpa_ctx *c, const DevCols &cols, const hpa_expr *cond, L n, uint16_t *mb, uint32_t *tcount);
int32_t jit_compress_map(hpa_ctx *c, int otyp, const DevCols &cols, const hpa_expr *ex, const uint16_t *mb,
                         const int64_t *toff, L n, void *dz);
int32_t jit_fused_reduce_many(hpa_ctx *c, int nq, int nx, V *x, const hpa_expr *conds, const hpa_expr *body,
                              const DevCols &cols, L n, void *part, int G);
