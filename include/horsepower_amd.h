/*
 * horsepower_amd.h -- C-ABI of the MI355X (gfx950) HorseIR primitive execution layer.
 *
 * The library replaces HorsePower's primitive library (src/horseir/v3/backend/h_primitive.cc,
 * h_libs.cc, h_hash_v3.cc, h_system.cc) behind the same calling convention:
 *
 *     I pfnX(V z, V x [, V y [, V m]])      -- 0 on success, an E_* code otherwise
 *
 * over value nodes whose layout is byte-identical to the reference's V0
 * (src/horseir/v3/backend/common.h:31-55; 32 bytes, `typ` at offset 16).  A caller that was
 * linked against the reference primitives relinks against libhorsepower_amd.so; its V0 headers,
 * type codes and error codes are unchanged.
 *
 * Memory model.  Vector payloads (vg(x) when len > 1) produced by this library live in HBM of
 * the current device, carved from a per-context bump arena (the reference's process heap,
 * backend/h_memory.cc:93-112).  Inputs may be device payloads (produced here, or made resident
 * with hpaUpload) or host payloads (staged to HBM on use).  Basic values with len <= 1 keep the
 * reference's inline-scalar rule (vg(x) == &x->b, backend/h_memory.cc:182-196) and are readable
 * on the host.  List/dict cells (the V0 arrays behind H_G/H_N nodes) stay in host memory, as in
 * the reference.  Group results use the reference's flat-list form (typ H_G, vg = i64 group
 * sizes, vg2 -> one i64 vector of row indices; backend/h_memory.cc:210-215).  hpaToHost copies a
 * device payload back before host code dereferences it.
 *
 * Errors are returned, never turned into process exits (the reference prints and exits through
 * printErrMsg/EP, backend/h_system.cc:1009-1028, global.h:65-66).
 *
 * Threading.  Each host thread issues work on its current context (a HIP stream plus arenas).
 * The plain ABI uses the process default context, so single-threaded callers see the
 * reference's one-primitive-at-a-time behaviour.  Concurrent requests (serving) each use their
 * own context; calls that share one context from several threads race.
 */
#ifndef HORSEPOWER_AMD_H
#define HORSEPOWER_AMD_H

#include <stdint.h>
#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- value node: layout of V0 in src/horseir/v3/backend/common.h:31-55 ---------------- */
typedef struct hpa_V0 {
    int64_t len;          /* vn(x) */
    char   *g;            /* vg(x): payload, or &b for inline scalars */
    int32_t typ;          /* vp(x): H_* type code */
    int32_t _pad;
    union {               /* inline scalar storage / list flat pointer (vg2) */
        bool     b;  int8_t j; int16_t h; int32_t i32; int64_t i64;
        float    f32; double f64; char c; int64_t q; char *s;
        int32_t  d;  int64_t z;
        struct hpa_V0 *g2;
        void    *p;
    };
} hpa_V0, *hpa_V;

/* type codes: src/horseir/v3/backend/h_type.h:1-23 */
#define HPA_B 0
#define HPA_J 1
#define HPA_H 2
#define HPA_I 3
#define HPA_L 4
#define HPA_F 5
#define HPA_E 6
#define HPA_X 7
#define HPA_C 8
#define HPA_Q 9
#define HPA_S 10
#define HPA_M 11
#define HPA_D 12
#define HPA_Z 13
#define HPA_U 14
#define HPA_W 15
#define HPA_T 16
#define HPA_G 17
#define HPA_N 18
#define HPA_Y 19
#define HPA_A 90
#define HPA_K 91
#define HPA_V 93

/* error codes: src/horseir/v3/backend/h_system.h:128-142 (plus HPA_E_MEMORY / HPA_E_DEVICE,
 * which replace the reference's "Heap full!!" exits and report HIP runtime failures) */
#define HPA_E_DOMAIN          1
#define HPA_E_GENERAL         2
#define HPA_E_INDEX           3
#define HPA_E_DIV_ZERO        4
#define HPA_E_LENGTH          5
#define HPA_E_TYPE            6
#define HPA_E_NULL_VALUE      7
#define HPA_E_LIKE_PATTERN    8
#define HPA_E_MATCH           9
#define HPA_E_ENUM_INDEX      10
#define HPA_E_RAZE_LIST       11
#define HPA_E_UNKNOWN         12
#define HPA_E_TABLE_NOT_FOUND 50
#define HPA_E_COL_NOT_FOUND   51
#define HPA_E_NOT_IMPL        99
#define HPA_E_MEMORY          100
#define HPA_E_DEVICE          101

typedef int32_t (*hpa_fn1)(hpa_V, hpa_V);
typedef int32_t (*hpa_fn2)(hpa_V, hpa_V, hpa_V);

/* ---- runtime ----------------------------------------------------------------------------- */
typedef struct hpa_ctx hpa_ctx;

/* Initialise on `device` with a default context whose device arena holds `arena_bytes`
 * (0 = 16 GiB, the reference's INIT_HEAP_SIZE, backend/h_memory.cc:11).  Idempotent. */
int32_t  hpaInit(int32_t device, int64_t arena_bytes);
void     hpaShutdown(void);
const char *hpaErrorString(int32_t code);
const char *hpaLastErrorDetail(void);

/* allocNode (backend/h_memory.cc:155-157): a zeroed host V0 from the current context. */
hpa_V    hpaAllocNode(void);
/* n allocNode calls in one: out[0..n) receive zeroed host V0s (HPA_E_MEMORY when the host arena
 * is full).  No counterpart in the reference; it saves a foreign-function caller n crossings. */
int32_t  hpaAllocNodes(int32_t n, hpa_V *out);
/* getHeapOffset / setHeapOffset (backend/h_memory.cc:147-153): arena watermark of the current
 * context; setting it releases every payload and node allocated after the mark. */
int64_t  hpaGetHeapOffset(void);
int64_t  hpaSetHeapOffset(int64_t cur);
/* initV (backend/h_memory.cc:182-196) with the payload in HBM. */
int32_t  hpaInitV(hpa_V x, int32_t typ, int64_t len);

/* Residency: copy a host payload into a long-lived HBM allocation (kept across heap resets)
 * and repoint vg(x) at it.  Lists/tables are walked recursively. */
int32_t  hpaUpload(hpa_V x);
/* Release an hpaUpload allocation (the node is left pointing at nothing). */
int32_t  hpaFreeResident(hpa_V x);
/* Copy a device payload to a host buffer owned by the context and repoint vg(x) at it
 * (recursively for lists and dicts; flat lists keep their form). */
int32_t  hpaToHost(hpa_V x);
/* Copy `bytes` of x's payload (device or host) into caller memory. */
int32_t  hpaReadPayload(hpa_V x, void *dst, int64_t bytes);
/* Copy the payloads of n vectors into caller buffers with a single synchronisation. */
int32_t  hpaReadMany(int32_t n, hpa_V *x, void **dst, const int64_t *bytes);
int32_t  hpaIsDevicePtr(const void *p);
int32_t  hpaSync(void);

/* Contexts: a stream plus device/host arenas.  The calling thread's current context is the
 * target of every pfn* call. */
hpa_ctx *hpaCtxCreate(int64_t arena_bytes);          /* < 0: no arena until hpaCtxSetArena */
/* Registered device pool, and per-request arena slices of it (memory-budget admission). */
void    *hpaDeviceAlloc(int64_t bytes);
int32_t  hpaDeviceFree(void *p);
int32_t  hpaCtxSetArena(hpa_ctx *c, void *base, int64_t bytes);
void     hpaCtxDestroy(hpa_ctx *c);
hpa_ctx *hpaCtxDefault(void);
void     hpaCtxSetCurrent(hpa_ctx *c);
hpa_ctx *hpaCtxCurrent(void);
void    *hpaCtxStream(hpa_ctx *c);          /* hipStream_t */
int64_t  hpaCtxArenaBytes(hpa_ctx *c);
int64_t  hpaCtxArenaPeak(hpa_ctx *c);
void     hpaCtxResetPeak(hpa_ctx *c);

/* Kernel timing: while enabled (on = 1), launches named by the library are bracketed with HIP
 * events on the launching stream; stats are (launch count, total ms).  on = 2 keeps only the
 * caller's scopes (hpaProfileBegin/End), so no events are recorded inside the library's calls;
 * 0 turns timing off. */
void     hpaProfileEnable(int32_t on);
void     hpaProfileReset(void);
int32_t  hpaProfileStats(const char *kernel, int64_t *count, double *total_ms);
int32_t  hpaProfileList(char *buf, int64_t cap);
/* A named scope around any calls on the current context (e.g. one whole primitive), with
 * its algorithmic bytes; listed by hpaProfileList with the library's own scopes. */
void     hpaProfileBegin(const char *name, double algo_bytes);
void     hpaProfileEnd(void);

/* ---- primitives: src/horseir/v3/backend/h_primitive.h:72-211 --------------------------- */
int32_t pfnList       (hpa_V z, int64_t n, hpa_V x[]);     /* h_primitive.cc:32-36 */
int32_t pfnIndex      (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:43-94 */
int32_t pfnLen        (hpa_V z, hpa_V x);                  /* h_primitive.cc:656-662 */
int32_t pfnWhere      (hpa_V z, hpa_V x);                  /* h_primitive.cc:728-762 */
int32_t pfnSum        (hpa_V z, hpa_V x);                  /* h_primitive.cc:764-786 */
int32_t pfnAvg        (hpa_V z, hpa_V x);                  /* h_primitive.cc:788-800 */
int32_t pfnReduce     (hpa_V z, hpa_V x, int32_t op);      /* h_primitive.cc:808-840 */
int32_t pfnMin        (hpa_V z, hpa_V x);                  /* h_primitive.cc:842-844 */
int32_t pfnMax        (hpa_V z, hpa_V x);                  /* h_primitive.cc:846-848 */
int32_t pfnRaze       (hpa_V z, hpa_V x);                  /* h_primitive.cc:935-955 */
int32_t pfnGroup      (hpa_V z, hpa_V x);                  /* h_primitive.cc:993-999 */
int32_t pfnGroupTrie  (hpa_V z, hpa_V x);                  /* h_primitive.cc:2656-2681 */
int32_t pfnGroupBasic (hpa_V z, hpa_V x);                  /* h_primitive.cc:1002-1057 */
int32_t pfnCompare    (hpa_V z, hpa_V x, hpa_V y, int32_t op); /* h_primitive.cc:1073-1169 */
int32_t pfnLt         (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnLeq        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnGt         (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnGeq        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnEq         (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnNeq        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnArith      (hpa_V z, hpa_V x, hpa_V y, int32_t op); /* h_primitive.cc:1229-1291 */
int32_t pfnPlus       (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnMinus      (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnMul        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnDiv        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnAnd        (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:1315-1357 */
int32_t pfnNand       (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnOr         (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnNor        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnXor        (hpa_V z, hpa_V x, hpa_V y);
int32_t pfnCompress   (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:1479-1521 */
int32_t pfnIndexOf    (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:1525-1579 */
int32_t pfnOrderBy    (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:1769-1816 */
int32_t pfnEach       (hpa_V z, hpa_V x, hpa_fn1 foo);     /* h_primitive.cc:1818-1832 */
int32_t pfnEachRight  (hpa_V z, hpa_V x, hpa_V y, hpa_fn2 foo); /* h_primitive.cc:1882-1904 */
int32_t pfnJoinIndex  (hpa_V z, hpa_V x, hpa_V y, hpa_V f);/* h_primitive.cc:2122-2155 */
int32_t pfnMember     (hpa_V z, hpa_V y, hpa_V x);         /* h_primitive.cc:2288-2336 */
int32_t pfnKeys       (hpa_V z, hpa_V x);                  /* h_primitive.cc:190-203 */
int32_t pfnValues     (hpa_V z, hpa_V x);                  /* h_primitive.cc:205-218 */

/* ---- enumerations and foreign keys ---------------------------------------------------------
 * An enum node (typ HPA_Y) has the reference's layout (backend/common.h:19, h_primitive.h:12-16,
 * h_system.h:12-14): len = number of values, vg = i64 index of each value among the keys (in
 * HBM), and the union's pointer `p` -> hpa_Y0 {source, ek = keys node, ev = values node}.
 * Tables (typ HPA_A) are V0[2] (column-name symbols, list of columns) with p -> hpa_A0 {rows,
 * columns, foreign-key list} (h_system.h:17-31); the list holds the enums pfnAddFKey made. */
typedef struct hpa_Y0 { int64_t source; hpa_V ek; hpa_V ev; } hpa_Y0;
typedef struct hpa_ListY { void *y; void *cn; struct hpa_ListY *next; } hpa_ListY;
typedef struct hpa_A0 { int64_t row; int64_t col; hpa_ListY *fkey; } hpa_A0;
int32_t pfnEnum       (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:2198-2222 */
int32_t pfnFetch      (hpa_V z, hpa_V x);                  /* h_primitive.cc:220-225 */
int32_t pfnAddFKey    (hpa_V x, hpa_V xKey, hpa_V y, hpa_V yKey); /* h_primitive.cc:2425-2466 */
int32_t pfnColumnValue(hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:283-313 */
int32_t pfnLoadTable  (hpa_V z, hpa_V x);                  /* h_primitive.cc:20-30 */
/* registerTable (h_system.cc:58-62): the catalog entry pfnLoadTable and pfnAddFKey look up by
 * name; the reference makes it when it reads a table from disk (load_db.cc), which this library
 * does not do, so the caller registers the tables it holds. */
int32_t hpaRegisterTable(int64_t name, hpa_V table);

/* Table snapshots in the reference's .bin format (serializeV / readSerializeV,
 * backend/h_io.cc:630-776): the reader streams every column payload from the file into resident
 * HBM through pinned staging (the table never exists as host payloads), re-interns symbol names
 * and mirrors strings as hpaUpload does; *payload_bytes (may be null) gets the payload bytes
 * moved.  hpaInitTableFromBin is initTableFromBin (load_db.cc:224-234): read, then registered
 * under `name` for pfnLoadTable, the table node in *z.  hpaWriteSerialized is serializeV of a
 * node whose payloads are in HBM or on the host. */
int32_t hpaReadSerialized(hpa_V z, const char *path, int64_t *payload_bytes);
int32_t hpaInitTableFromBin(const char *name, const char *path, hpa_V *z);
int32_t hpaWriteSerialized(hpa_V x, const char *path);

/* Join operator names.  pfnJoinIndex's `f` is a symbol vector of operator names
 * ("eq", "neq", ...; backend/h_system.cc:779-789).  Symbols are interned ids in the reference
 * (backend/h_symbol.cc); this library interns the names it understands here. */
int64_t hpaSymbol(const char *name);
const char *hpaSymbolName(int64_t id);
/* The caller's symbol table becomes the library's: names[i] gets id i, so the ids of symbol
 * values agree with the interpreter's getSymbol (backend/h_symbol.cc:148-157).  Call it once the
 * host's symbols exist (and again after it makes more); names the library interned before and
 * that are not among them move to ids after n. */
int32_t hpaImportSymbols(int64_t n, const char *const *names);

/* ---- symbols and strings -----------------------------------------------------------------
 * Symbols are ids into that table; strings (HPA_S) are vectors of char* holding host addresses,
 * as in the reference.  hpaUpload of an S column mirrors its strings' bytes in HBM; kernels read
 * the mirror.  Symbol order compares (pfnLt/Leq/Gt/Geq on HPA_Q, compareSymbol,
 * h_symbol.cc:241-244) and every string compare (strcmp, h_primitive.cc:1121-1160) go through
 * pfnCompare as the other types do. */
int32_t pfnLike       (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:1670-1757 */
int32_t pfnSubString  (hpa_V z, hpa_V x, hpa_V y);         /* h_primitive.cc:2468-2489 */

/* ---- fused loops emitted by the HorseIR optimizer ---------------------------------------
 * An expression is postfix code over the input columns x[0..nx).  Each instruction yields an
 * int64 or a float64 value, following C's usual arithmetic conversions as in the emitted
 * loops (optimizer/elementwise.cc:245-276, src/horseir/v1/compile-static/fuse.h:13-26: float
 * compares are exact, arithmetic is per-operation IEEE without contraction). */
enum hpa_op {
    HPA_OP_COL = 1,   /* push x[arg][i] */
    HPA_OP_CONST_I,   /* push imm_i */
    HPA_OP_CONST_F,   /* push imm_f */
    HPA_OP_ADD, HPA_OP_SUB, HPA_OP_MUL, HPA_OP_DIV,
    HPA_OP_LT, HPA_OP_LEQ, HPA_OP_GT, HPA_OP_GEQ, HPA_OP_EQ, HPA_OP_NEQ,
    HPA_OP_AND, HPA_OP_OR, HPA_OP_NOT, HPA_OP_NEG
};
typedef struct hpa_ins { int32_t op; int32_t arg; int64_t imm_i; double imm_f; } hpa_ins;
typedef struct hpa_expr { int32_t n; int32_t _pad; const hpa_ins *code; } hpa_expr;

/* FE (optimizer/elementwise.cc:239-281): z[i] = e(x..)[i] as type `typ` (B, I, L or E). */
int32_t hpaFusedMap(hpa_V z, int32_t typ, int32_t nx, hpa_V *x, const hpa_expr *e);
/* FA reduction with filter (optimizer/autofusion.cc:795-817,886-917):
 * c = reduce over {i : cond(i)} of body(i); kind 0 sum, 1 avg (c/count), 2 min, 3 max.
 * cond may be NULL (all rows).  z is E for an E body, else L (sum) / same as body. */
int32_t hpaFusedReduce(hpa_V z, int32_t kind, int32_t nx, hpa_V *x,
                       const hpa_expr *cond, const hpa_expr *body);
/* Shared scan of nq <= 8 FA reductions (sum kind 0, avg kind 1) of one body under nq
 * conditions (conds[q]) over the same columns: z[q] equals hpaFusedReduce(z[q], kind, nx, x,
 * &conds[q], body), computed in one pass over the columns. */
int32_t hpaFusedReduceMany(int32_t nq, hpa_V *z, int32_t kind, int32_t nx, hpa_V *x,
                           const hpa_expr *conds, const hpa_expr *body);
/* FA compress scan (optimizer/autofusion.cc:886-899) and the multi-column peephole
 * (src/horseir/v1/compile-static/cur/q6.h:12-26): z[k] = compress(cond, y[k]) for k < ny,
 * one mask evaluation, order preserving. */
int32_t hpaFusedCompress(int32_t ny, hpa_V *z, hpa_V *y, int32_t nx, hpa_V *x,
                         const hpa_expr *cond);
/* Fused filter + group + aggregates in one pass over the table: the FA filter, pfnGroupTrie
 * on one or two byte-wide key columns (h_primitive.cc:2656-2681) and the FP grouped reductions
 * (optimizer/pattern.cc:180-219) of the reference fused into a single scan.  Groups come out in
 * ascending key order, lexicographic over the key columns with each byte compared as its type
 * orders it (C and J signed, B unsigned): pfnGroupTrie's bucket order over the byte values its
 * code defines (0..127), and the order pfnOrderBy gives the key columns; zkeys[k] holds key column k per group,
 * zfirst the first row of each group, zaggs[a] aggregate a (kind 0 sum, 1 avg, 2 min, 3 max,
 * 4 len; body unused for len).  Returns HPA_E_NOT_IMPL without writing outputs when one
 * workgroup meets more than 16 distinct key codes (the caller then runs the unfused primitives). */
typedef struct hpa_agg { int32_t kind; int32_t _pad; const hpa_expr *body; } hpa_agg;
int32_t hpaFusedGroupAgg(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                         int32_t na, const hpa_agg *aggs, hpa_V zfirst, hpa_V *zkeys, hpa_V *zaggs);
/* hpaFusedGroupAgg that also delivers its results to host memory, for a plan whose next step is
 * reading them: hkeys[k] (cap bytes each), hfirst (cap i64, used when zfirst is given) and
 * haggs[a] (cap 8-byte values: i64 for len and integer sums, f64 otherwise) receive the first
 * min(cap, groups) groups, and *ng the group count, with one stream synchronisation in all
 * (flag, count and outputs are read together).  The z* nodes are set as by hpaFusedGroupAgg;
 * when *ng > cap the host copies are incomplete and the nodes hold the results.  No counterpart
 * in the reference, whose compiled plans end in host memory already (cur/q1.h). */
int32_t hpaFusedGroupAggHost(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                             int32_t na, const hpa_agg *aggs, hpa_V zfirst, hpa_V *zkeys, hpa_V *zaggs,
                             int64_t cap, void **hkeys, int64_t *hfirst, void **haggs, int64_t *ng);
/* hpaFusedGroupAggHost over groups (bucket, keys): `bucket` is an expression over x giving each
 * row a bucket number below 2^bbits (bbits 1 or 2; the value is taken modulo 2^bbits), and
 * hbucket (cap bytes) receives each group's bucket beside its key bytes; groups come out in
 * (bucket, keys) order.  Requests that differ only in one range bound of a column share one
 * pass this way: with bucket = the number of bounds below the row's value, a request's groups
 * are the sum of the buckets up to its bound.  No counterpart in the reference: it is the
 * serving loop's shared scan. */
int32_t hpaFusedGroupAggBuckets(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                                const hpa_expr *bucket, int32_t bbits, int32_t na, const hpa_agg *aggs,
                                hpa_V *zkeys, hpa_V *zaggs, int64_t cap, void **hkeys, uint8_t *hbucket,
                                void **haggs, int64_t *ng);
/* Diagnostic (no device needed): generate the fused group-aggregate kernel for these column
 * types and expressions and compile it with hipRTC; 0 when the generated code compiles. */
int32_t hpaJitCompileCheck(int32_t nx, const int32_t *coltyp, const hpa_expr *cond,
                           int32_t na, const hpa_agg *aggs);
/* Peephole index: z[k] = index(x[k], y) for nx <= 16 basic columns, one launch and one bounds
 * check (the gathers of every aggregate by a group or order permutation, q1.hir:57-66). */
int32_t hpaIndexMany(int32_t nx, hpa_V *z, hpa_V *x, hpa_V y);
/* Peephole compress with a ready mask: z[k] = compress(mask, y[k]). */
int32_t hpaCompressMany(int32_t ny, hpa_V *z, hpa_V mask, hpa_V *y);
/* The FA compress scan with FE maps fused (optimizer/autofusion.cc:814-817,886-899): z[k] =
 * compress(mask, y[k]) for k < ny and zm[j] = compress(mask, maps[j](x)) as type mtyp[j] (B, I,
 * L or E) for j < nm <= 2, in one write pass; each map is evaluated at the selected rows only.
 * Outputs never alias the inputs (an all-true mask copies). */
int32_t hpaCompressMap(int32_t ny, hpa_V *z, hpa_V mask, hpa_V *y, int32_t nm, hpa_V *zm,
                       const int32_t *mtyp, const hpa_expr *maps, int32_t nx, hpa_V *x);
/* hpaCompressMap with the mask given as its condition: cond, an expression over x (the columns
 * the maps read), selects the rows; z[k] = y[k][cond] and zm[j] = maps[j](x)[cond] as type
 * mtyp[j].  The FA compress scan of one condition (optimizer/autofusion.cc:886-899): the count
 * pass evaluates cond and keeps each row's selection for the write pass, so no byte mask is
 * written.  Results equal hpaCompressMap's with the mask cond(x). */
int32_t hpaFusedCompressMap(int32_t ny, hpa_V *z, hpa_V *y, int32_t nm, hpa_V *zm, const int32_t *mtyp,
                            const hpa_expr *maps, int32_t nx, hpa_V *x, const hpa_expr *cond);
/* FP1/FP2/FP5/FP6/FP7 grouped reductions (optimizer/pattern.cc:180-219): for each group list
 * t of y (flat or normal list of i64 row indices), z[g] = reduce_{j in t}(x[j]);
 * kind 0 sum, 1 avg, 2 min, 3 max, 4 len.  z is a vector over groups. */
int32_t hpaGroupReduce(hpa_V z, int32_t kind, hpa_V x, hpa_V y);

/* ---- allocation-free entry point -----------------------------------------------------------
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * hpaSumNoAlloc is pfnSum (h_primitive.cc:764-786) of an E vector whose payload is in HBM: the
 * same kernels and the same FP64 result, written to the device address `out` on the current
 * context's stream, with no allocation and no host synchronisation.  Its workgroup partials live
 * in a static device scratch of 4096 partials (64 KiB per GPU, held until the process ends);
 * a call that would need more returns HPA_E_MEMORY.  hpaInit must have run on the device.  Calls
 * that overlap on one GPU share that scratch and race. */
int32_t hpaSumNoAlloc(hpa_V x, double *out);

/* ---- synthetic tables generated in HBM (benchmarks, serving): counter-based splitmix64
 * columns, the values of horsepower_amd/datagen.py.  Payloads are resident (hpaFreeResident). */
int32_t hpaGenUniform(hpa_V z, int32_t typ, int64_t n, int64_t start, uint64_t seed, uint64_t stream,
                      int64_t lo, int64_t hi);
/* dim[3] = orders(key, order_day, segment), fact[8] = lineitem(fk, ship_day, price, discount,
 * quantity, tax, return_flag, line_status), 1..7 lineitem rows per order. */
int32_t hpaGenServeTables(int64_t n_dim, uint64_t seed, hpa_V *dim, hpa_V *fact);

/* ---- batched execution (no single reference counterpart: it runs, for many queries at once,
 * the reference's per-query sequence compare -> compress -> group -> FP1/FP7 grouped sum/len,
 * h_primitive.cc:1073-1106,1479-1511,993-999 and optimizer/pattern.cc:180-219).
 * Query q: rows with flt[q] <ops[q]> consts[q] (ops HPA_OP_LT..HPA_OP_NEQ), grouped on the i32
 * column keys[q], with the f64 sum of vals[q] and the row count per group, groups in pfnGroup's
 * order for i32 keys (h_libs.cc:1295-1329).  The plan lives in the current context's arena
 * (free it with hpaSetHeapOffset after hpaBatchDestroy); NULL on error (hpaLastErrorDetail). */
typedef struct hpa_batch hpa_batch;
hpa_batch *hpaBatchCreate(int32_t nq, hpa_V *keys, hpa_V *vals, hpa_V *flts,
                          const int32_t *ops, const int32_t *consts);
void     hpaBatchDestroy(hpa_batch *b);
/* One persistent launch over all queries; use_graph != 0 replays a captured hipGraph. */
int32_t  hpaBatchRun(hpa_batch *b, int32_t use_graph);
/* Query qi of the last run: zkey (I), zcnt (L), zsum (E), one element per group. */
int32_t  hpaBatchResult(hpa_batch *b, int32_t qi, hpa_V zkey, hpa_V zcnt, hpa_V zsum);
/* Every query's results of the last run in pinned host memory owned by the plan (valid until
 * the next fetch): ng[q] groups, keys[q] (i32), cnts[q] (i64), sums[q] (f64). */
int32_t  hpaBatchFetch(hpa_batch *b, int64_t *ng, void **keys, void **cnts, void **sums);
/* The same fetch pipelined with the next run: Begin snapshots the last run's results in HBM and
 * starts their copy to pinned host memory on a stream of its own (it returns at once, so a
 * following hpaBatchRun computes while they travel); End waits for the oldest begun fetch and
 * gives its results as hpaBatchFetch does, valid until two more Begins.  At most two fetches
 * are outstanding.  The snapshot takes as many groups as the last counted run made; a run that
 * makes more (its tables changed) fails in End with HPA_E_LENGTH. */
int32_t  hpaBatchFetchBegin(hpa_batch *b);
int32_t  hpaBatchFetchEnd(hpa_batch *b, int64_t *ng, void **keys, void **cnts, void **sums);
/* Tunables: block size, rows per work item, LDS key budget, and the key-partition width of
 * queries whose key range exceeds the LDS budget (0 = compensated global atomics instead;
 * otherwise a power of two in [1024, 2048]). */
int32_t  hpaBatchSetConfig(hpa_batch *b, int32_t block, int32_t tile_rows, int32_t lds_keys, int32_t part_keys);
int32_t  hpaBatchGetConfig(hpa_batch *b, int32_t *block, int32_t *tile_rows, int32_t *lds_keys, int32_t *part_keys);
/* Autotuner: times every candidate configuration on this plan and keeps the fastest. */
int32_t  hpaBatchAutotune(hpa_batch *b, int32_t iters, float *ms, int32_t cap, int32_t *ncand);

#ifdef __cplusplus
}
#endif
#endif /* HORSEPOWER_AMD_H */
