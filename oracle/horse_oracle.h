/*
 * horse_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room CPU (C11 + OpenMP) restatement of HorsePower's primitive library
 * (src/horseir/v3/backend/h_primitive.cc, h_libs.cc, h_hash_v3.cc, h_system.cc), over host
 * value nodes with the reference's V0 layout.  It is the checker for the HIP library and the
 * timed CPU baseline; nothing in the product links or calls it.  Every function cites the
 * reference lines it follows.  Functions whose reference behaviour is undefined (out-of-bounds
 * reads/writes) state the defined behaviour chosen instead.
 */
#ifndef HORSE_ORACLE_H
#define HORSE_ORACLE_H
#include "../include/horsepower_amd.h"

typedef hpa_V0 O0;
typedef hpa_V  OV;

int32_t orc_init(int64_t heap_bytes, int32_t threads);
int32_t orc_threads(void);
int64_t orc_get_heap(void);
int64_t orc_heap_capacity(void);
int64_t orc_set_heap_top(int64_t top);
int64_t orc_set_heap(int64_t cur);
OV      orc_alloc_node(void);
OV      orc_initV(OV x, int32_t typ, int64_t len);

int32_t orc_pfnList(OV z, int64_t n, OV x[]);
int32_t orc_pfnIndex(OV z, OV x, OV y);
int32_t orc_pfnLen(OV z, OV x);
int32_t orc_pfnWhere(OV z, OV x);
int32_t orc_pfnSum(OV z, OV x);
int32_t orc_pfnAvg(OV z, OV x);
int32_t orc_pfnReduce(OV z, OV x, int32_t op);
int32_t orc_pfnMin(OV z, OV x);
int32_t orc_pfnMax(OV z, OV x);
int32_t orc_pfnRaze(OV z, OV x);
int32_t orc_pfnGroup(OV z, OV x);
int32_t orc_pfnGroupTrie(OV z, OV x);
int32_t orc_pfnGroupBasic(OV z, OV x);
int32_t orc_pfnCompare(OV z, OV x, OV y, int32_t op);
int32_t orc_pfnLt(OV z, OV x, OV y);
int32_t orc_pfnLeq(OV z, OV x, OV y);
int32_t orc_pfnGt(OV z, OV x, OV y);
int32_t orc_pfnGeq(OV z, OV x, OV y);
int32_t orc_pfnEq(OV z, OV x, OV y);
int32_t orc_pfnNeq(OV z, OV x, OV y);
int32_t orc_pfnArith(OV z, OV x, OV y, int32_t op);
int32_t orc_pfnPlus(OV z, OV x, OV y);
int32_t orc_pfnMinus(OV z, OV x, OV y);
int32_t orc_pfnMul(OV z, OV x, OV y);
int32_t orc_pfnDiv(OV z, OV x, OV y);
int32_t orc_pfnAnd(OV z, OV x, OV y);
int32_t orc_pfnNand(OV z, OV x, OV y);
int32_t orc_pfnOr(OV z, OV x, OV y);
int32_t orc_pfnNor(OV z, OV x, OV y);
int32_t orc_pfnXor(OV z, OV x, OV y);
int32_t orc_pfnCompress(OV z, OV x, OV y);
int32_t orc_pfnIndexOf(OV z, OV x, OV y);
int32_t orc_pfnOrderBy(OV z, OV x, OV y);
int32_t orc_pfnEach(OV z, OV x, int32_t (*foo)(OV, OV));
int32_t orc_pfnEachRight(OV z, OV x, OV y, int32_t (*foo)(OV, OV, OV));
int32_t orc_pfnJoinIndex(OV z, OV x, OV y, OV f);
int32_t orc_pfnMember(OV z, OV y, OV x);
int32_t orc_pfnEnum(OV z, OV x, OV y);
int32_t orc_pfnFetch(OV z, OV x);
int32_t orc_pfnAddFKey(OV x, OV xKey, OV y, OV yKey);
int32_t orc_pfnColumnValue(OV z, OV x, OV y);
int32_t orc_pfnLoadTable(OV z, OV x);
int32_t orc_register_table(int64_t sid, OV t);
int32_t orc_pfnKeys(OV z, OV x);
int32_t orc_pfnValues(OV z, OV x);
int64_t orc_symbol(const char *name);
const char *orc_symbol_name(int64_t q);
int32_t orc_pfnLike(OV z, OV x, OV y);
int32_t orc_pfnSubString(OV z, OV x, OV y);
int32_t orc_serializeV(OV x, const char *path);
int32_t orc_readSerializeV(OV z, const char *path);

/* compensated FP64 sums (Neumaier), the accuracy reference for every FP64 aggregate */
double  orc_sum_neumaier(const double *x, int64_t n);
double  orc_sum_neumaier_idx(const double *x, const int64_t *idx, int64_t n);

/* the reference's fused Q6 form (src/horseir/v1/compile-static/opt/q6.h:31-48), exact
 * compares, one OpenMP reduction */
double  orc_q6_fused(int64_t n, const int32_t *ship, const double *disc, const int32_t *qty,
                     const double *price, int32_t lo, int32_t hi, double dlo, double dhi,
                     int32_t qmax);
#endif
