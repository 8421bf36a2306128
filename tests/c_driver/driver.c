/*
 * A C caller of libhorsepower_amd through include/horsepower_amd.h alone, in the reference's
 * calling convention: the caller makes the V0 headers (as allocNode/NEW(V0) do in
 * translator/interpreter.cc:62-81), calls pfn* primitives and reads inline scalar results from
 * the node, resetting the arena per run (src/horseir/v1/compile-static/main.c:20-35).
 *
 *   driver DIR
 *
 * reads little-endian columns from DIR (ship_day.i32, discount.f64, quantity.i32, price.f64,
 * jx_key.i32, jx_val.f64, jy_key.i32, jy_val.i64, gkey.f64) and writes to DIR:
 *   q6.f64         revenue of q6.hir's statement sequence (tests/hyper/online-hir/q6.hir)
 *   q6_fused.f64   the same through the fused FA reduction (src/horseir/v1/compile-static/opt/q6.h)
 *   join0.i64, join1.i64   pfnJoinIndex(list(jx_key, jx_val), list(jy_key, jy_val), `eq`,`lt`)
 *   gkeys.i64, gcounts.i64, grows.i64   pfnGroup(gkey)
 * and exits 0, or prints the failing call and its error and exits 1.
 */
#include "horsepower_amd.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *g_dir;

static void *load(const char *name, size_t w, int64_t *n){
    char path[4096]; snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE *f = fopen(path, "rb");
    if(!f){ fprintf(stderr, "cannot open %s\n", path); exit(1); }
    fseek(f, 0, SEEK_END); long sz = ftell(f); fseek(f, 0, SEEK_SET);
    void *p = malloc(sz > 0 ? (size_t)sz : 1);
    if(!p || fread(p, 1, (size_t)sz, f) != (size_t)sz){ fprintf(stderr, "cannot read %s\n", path); exit(1); }
    fclose(f);
    *n = (int64_t)(sz / (long)w);
    return p;
}
static void save(const char *name, const void *p, size_t bytes){
    char path[4096]; snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE *f = fopen(path, "wb");
    if(!f || fwrite(p, 1, bytes, f) != bytes){ fprintf(stderr, "cannot write %s\n", path); exit(1); }
    fclose(f);
}
static hpa_V0 vec(int32_t typ, void *p, int64_t n){
    hpa_V0 v; memset(&v, 0, sizeof v);
    v.typ = typ; v.len = n; v.g = (char*)p;
    return v;
}
/* inline scalars (initV's len <= 1 rule, backend/h_memory.cc:185-187): vg points into the node
 * itself, so the node is filled in place */
static void scalar_i(hpa_V x, int32_t v){ memset(x, 0, sizeof *x); x->typ = HPA_I; x->len = 1; x->g = (char*)&x->b; x->i32 = v; }
static void scalar_e(hpa_V x, double v){ memset(x, 0, sizeof *x); x->typ = HPA_E; x->len = 1; x->g = (char*)&x->b; x->f64 = v; }
#define CALL(e) do{ int32_t rc_ = (e); if(rc_){ fprintf(stderr, "%s -> %d (%s: %s)\n", #e, rc_, hpaErrorString(rc_), hpaLastErrorDetail()); return 1; } }while(0)

/* read a result vector of 8-byte elements into a fresh buffer (inline scalars included) */
static int64_t *read8(hpa_V x){
    int64_t *p = (int64_t*)malloc(8 * (size_t)(x->len > 0 ? x->len : 1));
    if(x->len > 0 && hpaReadPayload(x, p, 8 * x->len)){ fprintf(stderr, "read failed\n"); exit(1); }
    return p;
}

int main(int argc, char **argv){
    if(argc != 2){ fprintf(stderr, "usage: driver DIR\n"); return 2; }
    g_dir = argv[1];
    CALL(hpaInit(0, 0));
    int64_t n, n2;
    void *ship = load("ship_day.i32", 4, &n), *disc = load("discount.f64", 8, &n2), *qty = load("quantity.i32", 4, &n2),
         *price = load("price.f64", 8, &n2);
    hpa_V0 c_ship = vec(HPA_I, ship, n), c_disc = vec(HPA_E, disc, n), c_qty = vec(HPA_I, qty, n), c_price = vec(HPA_E, price, n);
    CALL(hpaUpload(&c_ship)); CALL(hpaUpload(&c_disc)); CALL(hpaUpload(&c_qty)); CALL(hpaUpload(&c_price));

    int64_t mark = hpaGetHeapOffset();
    /* q6.hir: the five predicates, and-ed, compress of price and discount, mul, sum */
    hpa_V0 lo, hi, q, dlo, dhi;
    scalar_i(&lo, 730); scalar_i(&hi, 1095); scalar_i(&q, 24); scalar_e(&dlo, 0.05); scalar_e(&dhi, 0.07);
    hpa_V0 t0 = {0}, t1 = {0}, t2 = {0}, t3 = {0}, t4 = {0}, t5 = {0}, t6 = {0}, t7 = {0}, t8 = {0};
    hpa_V0 p = {0}, d = {0}, pd = {0}, s = {0};
    CALL(pfnGeq(&t0, &c_ship, &lo)); CALL(pfnLt(&t1, &c_ship, &hi)); CALL(pfnAnd(&t2, &t0, &t1));
    CALL(pfnGeq(&t3, &c_disc, &dlo)); CALL(pfnLeq(&t4, &c_disc, &dhi)); CALL(pfnAnd(&t5, &t3, &t4));
    CALL(pfnAnd(&t6, &t2, &t5)); CALL(pfnLt(&t7, &c_qty, &q)); CALL(pfnAnd(&t8, &t6, &t7));
    CALL(pfnCompress(&p, &t8, &c_price)); CALL(pfnCompress(&d, &t8, &c_disc));
    CALL(pfnMul(&pd, &p, &d)); CALL(pfnSum(&s, &pd));
    save("q6.f64", &s.f64, 8);
    hpaSetHeapOffset(mark);

    /* the same through the fused FA reduction */
    hpa_ins cond[] = {
        {HPA_OP_COL, 0, 0, 0}, {HPA_OP_CONST_I, 0, 730, 0}, {HPA_OP_GEQ, 0, 0, 0},
        {HPA_OP_COL, 0, 0, 0}, {HPA_OP_CONST_I, 0, 1095, 0}, {HPA_OP_LT, 0, 0, 0}, {HPA_OP_AND, 0, 0, 0},
        {HPA_OP_COL, 2, 0, 0}, {HPA_OP_CONST_F, 0, 0, 0.05}, {HPA_OP_GEQ, 0, 0, 0}, {HPA_OP_AND, 0, 0, 0},
        {HPA_OP_COL, 2, 0, 0}, {HPA_OP_CONST_F, 0, 0, 0.07}, {HPA_OP_LEQ, 0, 0, 0}, {HPA_OP_AND, 0, 0, 0},
        {HPA_OP_COL, 1, 0, 0}, {HPA_OP_CONST_I, 0, 24, 0}, {HPA_OP_LT, 0, 0, 0}, {HPA_OP_AND, 0, 0, 0},
    };
    hpa_ins body[] = { {HPA_OP_COL, 3, 0, 0}, {HPA_OP_COL, 2, 0, 0}, {HPA_OP_MUL, 0, 0, 0} };
    hpa_expr ce = {(int32_t)(sizeof cond / sizeof *cond), 0, cond}, be = {3, 0, body};
    hpa_V cols[4] = {&c_ship, &c_qty, &c_disc, &c_price};
    hpa_V0 fz = {0};
    CALL(hpaFusedReduce(&fz, 0, 4, cols, &ce, &be));
    save("q6_fused.f64", &fz.f64, 8);
    hpaSetHeapOffset(mark);

    /* a two-column join: (key eq, value lt) */
    int64_t nx, ny;
    void *jxk = load("jx_key.i32", 4, &nx), *jxv = load("jx_val.f64", 8, &n2), *jyk = load("jy_key.i32", 4, &ny), *jyv = load("jy_val.i64", 8, &n2);
    hpa_V0 xc[2] = {vec(HPA_I, jxk, nx), vec(HPA_E, jxv, nx)}, yc[2] = {vec(HPA_I, jyk, ny), vec(HPA_L, jyv, ny)};
    hpa_V0 xl = vec(HPA_G, xc, 2), yl = vec(HPA_G, yc, 2);
    int64_t ops[2] = {hpaSymbol("eq"), hpaSymbol("lt")};
    hpa_V0 f = vec(HPA_Q, ops, 2), jz = {0};
    CALL(pfnJoinIndex(&jz, &xl, &yl, &f));
    hpa_V cells = (hpa_V)jz.g;
    int64_t *j0 = read8(cells), *j1 = read8(cells + 1);
    save("join0.i64", j0, 8 * (size_t)cells[0].len); save("join1.i64", j1, 8 * (size_t)cells[1].len);
    hpaSetHeapOffset(mark);

    /* a group on an f64 column with repeated values (keys ordered by the reference's quicksort) */
    int64_t ng;
    void *gk = load("gkey.f64", 8, &ng);
    hpa_V0 gcol = vec(HPA_E, gk, ng), gz = {0};
    CALL(pfnGroup(&gz, &gcol));
    hpa_V kv = (hpa_V)gz.g;                            /* dict: keys, values (flat list) */
    hpa_V keys = kv, vals = kv + 1;
    int64_t *k = read8(keys), *cnt = read8(vals);
    int64_t *rows = read8(vals->g2);
    save("gkeys.i64", k, 8 * (size_t)keys->len); save("gcounts.i64", cnt, 8 * (size_t)vals->len);
    save("grows.i64", rows, 8 * (size_t)vals->g2->len);
    hpaSetHeapOffset(mark);
    hpaShutdown();
    return 0;
}
