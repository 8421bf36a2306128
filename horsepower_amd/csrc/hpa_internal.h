// Internal declarations shared by the HIP translation units of libhorsepower_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <vector>
#include <type_traits>
#include <map>
#include <string>
#include "../../include/horsepower_amd.h"

typedef hpa_V  V;
typedef hpa_V0 V0;
typedef int64_t L;

#define R return

// ---------------------------------------------------------------- context and arenas
// mapped completion buffer of a context: a 64-byte head holding the sequence number, then the data
#define PUB_HEAD 64
#define PUB_BYTES (64 << 10)
#define PUB_ASYNC_SLOTS 64          // the buffer's last 64 x 64 bytes: reads begun and ended later
#define PUB_ASYNC_BYTES 64
struct hpa_ctx {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool own_arena = true;         // false: the arena is a slice of a pool (hpaCtxSetArena)
    int  device = 0;
    // device bump arena (the reference's process heap, backend/h_memory.cc:93-112)
    char *dbase = nullptr; L dcur = 0, dtop = 0, dpeak = 0;
    // host arena for V0 nodes and list cells
    char *hbase = nullptr; L hcur = 0, htop = 0;
    // device watermark -> (host watermark, host-buffer sequence number) when it was taken
    struct Mark { L hcur; uint64_t seq; };
    std::map<L, Mark> marks;
    // pinned host staging for small device->host reads
    char *pinned = nullptr; L pinned_bytes = 0;
    // coherent host memory mapped into the device: small reads are published there by a kernel
    // followed by a sequence number the host spins on (no blocking stream synchronisation)
    char *mapped = nullptr; char *mapped_dev = nullptr; uint64_t pub_seq = 0;
    char pub_fallback[PUB_ASYNC_SLOTS][PUB_ASYNC_BYTES];      // begun reads when nothing is mapped
    // host buffers handed out by hpaToHost, each with the creation sequence number and the
    // device watermark at creation; a rewind releases the ones made after its mark
    struct HostBuf { void *p; uint64_t seq; L dcur; };
    std::vector<HostBuf> host_bufs;
    uint64_t host_seq = 0;
    // index vectors produced here, with an exclusive upper bound of their (non-negative) values;
    // pfnIndex skips its device bounds check for them (entries die with their arena range)
    std::map<const void*, L> bounds;
    // group keys vectors made here over an arena i32 column, with the key value of each group:
    // index(column, keys) is those values (entries die with their arena range; the values are
    // allocated before the keys, so a rewind that frees either frees the keys)
    struct GroupVals { const void *src; L n; const void *vals; };
    std::map<const void*, GroupVals> gvals;
    // group counts vectors made here, with the g + 1 group offsets a flat list over them needs
    // (allocated before the counts, so a rewind that frees them frees the counts too)
    struct GroupOffs { L g; const int64_t *offs; };
    std::map<const void*, GroupOffs> goffs;
    // group counts vectors made by the run-compressed i32 group, with its sorted runs ((first row
    // << cb) | (length - 1)) and each group's first run (g + 1 entries): grouped reductions over
    // it read each run's rows in place of the flat list (allocated before the counts)
    struct GroupRuns { L g, n, nr; const uint32_t *sv; const uint32_t *grun; int cb; };
    std::map<const void*, GroupRuns> gruns;
};
// set the arena top of c (rewind or truncate), dropping side-table entries above it
void arena_set(hpa_ctx *c, L cur);
void note_bound(hpa_ctx *c, const void *p, L bound);
bool within_bound(hpa_ctx *c, const void *p, L limit);
void note_group_vals(hpa_ctx *c, const void *keys, const void *src, L n, const void *vals);
const void *group_vals(hpa_ctx *c, const void *keys, const void *src, L n);
void note_group_offsets(hpa_ctx *c, const void *counts, L g, const int64_t *offs);
void note_group_runs(hpa_ctx *c, const void *counts, const hpa_ctx::GroupRuns &r);
const hpa_ctx::GroupRuns *group_runs(hpa_ctx *c, const void *counts, L g);
int32_t resident_alloc(V x, int32_t typ, L len);   // long-lived HBM payload (hpaFreeResident)

hpa_ctx *ctx_cur();
bool     hpa_is_initialized();
int32_t  noalloc_scratch_init(int device);        // address of hpaSumNoAlloc's static scratch
void    *dalloc(hpa_ctx *c, L bytes);               // nullptr when the arena is full
void    *halloc(hpa_ctx *c, L bytes);
V        node_alloc(hpa_ctx *c);
void     set_error_detail(const char *fmt, ...);

// ---------------------------------------------------------------- value model helpers
L     type_width(int32_t typ);                      // payload bytes of one element (C: 1)
L     type_bytes(int32_t typ, L len);               // getTypeSize, backend/h_memory.cc:124-153
bool  is_basic(int32_t t);
bool  is_int(int32_t t);
bool  is_real(int32_t t);
bool  is_comparable(int32_t t);
bool  is_inline(V x);                               // vg(x) == &x->b
bool  is_device(const void *p);
void  register_device_range(void *p, L bytes);
void  unregister_device_range(void *p);

// Output allocation following initV's inline-scalar rule.  Returns the device buffer kernels
// write; for inline outputs this is arena scratch and `commit_inline` copies it back.
int32_t out_init(hpa_ctx *c, V z, int32_t typ, L len, void **dev);
int32_t commit_inline(hpa_ctx *c, V z, const void *dev);
// Device view of x's payload (staging host payloads into the arena).  Inline scalars are
// staged too, so every operand has a device address.
int32_t dev_view(hpa_ctx *c, V x, const void **p);
// Scalar read of element 0 of x as int64 / double bits (host or device payload).
int32_t read_scalar(hpa_ctx *c, V x, void *dst, L width);
// Copy small device data to host (synchronous on the context stream).
int32_t d2h(hpa_ctx *c, void *dst, const void *src, L bytes);
// A small read begun now, in stream order, and collected later (at most PUB_ASYNC_BYTES bytes;
// at most PUB_ASYNC_SLOTS begun and not yet ended): the host can queue more work before it waits.
int32_t d2h_begin(hpa_ctx *c, const void *src, L bytes, uint64_t *ticket);
int32_t d2h_end(hpa_ctx *c, uint64_t ticket, void *dst, L bytes);
int32_t check_launch(const char *what);

// ---------------------------------------------------------------- timing
// Brackets one kernel launch with HIP events on its stream while profiling is enabled;
// `bytes` is the launch's algorithmic HBM traffic (inputs read once, outputs written once).
struct ProfScope {
    const char *name; hipStream_t s; hipEvent_t a = nullptr, b = nullptr; bool on; double bytes = 0;
    bool stopped = false;
    ProfScope(const char *n, hipStream_t st, double algo_bytes = 0, bool user = false);
    void stop();                      // record the end event now (before later stream work)
    ~ProfScope();
};
bool prof_enabled();

// ---------------------------------------------------------------- launch geometry
int  num_cus();
inline L ceil_div(L a, L b){ R (a + b - 1) / b; }
// Grid of a whole-array reduction that ends in one global atomic per workgroup: few enough
// workgroups that the atomics to one address do not serialise the kernel.
inline int reduce_grid(L n, int block){
    L g = ceil_div(n, (L)block * 16), cap = (L)num_cus() * 2;
    if(g > cap) g = cap;
    if(g < 1) g = 1;
    R (int)g;
}
inline int grid_for(L n, int block, int per_thread, int max_blocks = 0){
    L g = ceil_div(n, (L)block * per_thread);
    L cap = max_blocks > 0 ? max_blocks : (L)num_cus() * 16;
    if(g > cap) g = cap;
    if(g < 1) g = 1;
    R (int)g;
}
// Workgroups go to the eight XCDs round-robin by block index (block b on XCD b mod 8).  This
// renumbers block b of a g-block grid so that each XCD takes one contiguous range of tiles:
// tiles that run side by side then share an L2, and writes that two neighbouring tiles make to
// one cache line meet there.  A bijection of [0, g).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned g){
    const unsigned x = b & 7u, q = b >> 3, per = g >> 3, rem = g & 7u;
    return x * per + (x < rem ? x : rem) + q;
}

// Lanes of the wave whose `nb`-bit digit d equals this lane's, among `valid` lanes: one ballot
// per digit bit, then per 32-bit half of the mask p &= ~(ballot ^ m), m all ones where this
// lane's bit is set, as one three-input bitwise op (v_bitop3 truth table f(S0, S1, S2) read at
// S0 = 0xF0, S1 = 0xCC, S2 = 0xAA: p & ~(b ^ m) = 0x90).
__device__ __forceinline__ uint64_t match_digit(unsigned d, bool valid, int nb){
    const uint64_t v = __ballot(valid);
    uint32_t plo = (uint32_t)v, phi = (uint32_t)(v >> 32);
    #pragma unroll
    for(int b = 0; b < nb; b++){
        const uint32_t m = (uint32_t)(((int32_t)(d << (31 - b))) >> 31);
        const uint64_t bb = __builtin_amdgcn_uicmp(m, 0u, 33);      // wave compare m != 0 (33: not equal)
        plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bb, m, 0x90);
        phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bb >> 32), m, 0x90);
    }
    return ((uint64_t)phi << 32) | plo;
}

// ---------------------------------------------------------------- shared device algorithms
// exclusive scan of n int64 values (in place allowed); total returned through *total_dev
// (device pointer, may be null).
int32_t scan_i64(hpa_ctx *c, const int64_t *in, int64_t *out, L n, int64_t *total_dev);
int32_t scan_u32_to_i64(hpa_ctx *c, const uint32_t *in, int64_t *out, L n, int64_t *total_dev);
// device ranges into host memory with one synchronisation (pinned staging when they fit)
int32_t read_staged(hpa_ctx *c, int n, const void *const *src, void *const *dst, const int64_t *bytes);
// promote a device vector to type `to` (promoteValue, backend/h_system.cc:168-252)
int32_t promote_dev(hpa_ctx *c, V x, int32_t to, const void **p);
// gather: dst[i] = src[idx[i]] for element width w
int32_t gather_dev(hpa_ctx *c, void *dst, const void *src, const int64_t *idx, L n, L w);
// stable LSD radix sort of (key, value) pairs by unsigned key bits [lo_bit, hi_bit)
// (inputs are only read; vals null = row ids; outputs in fresh buffers, or the inputs
// themselves when every key digit in the range is constant)
int32_t radix_sort_pairs_u64(hpa_ctx *c, const uint64_t *keys, const uint32_t *vals, L n, int lo_bit, int hi_bit,
                             uint64_t **keys_out, uint32_t **vals_out);
int32_t radix_sort_pairs_u32(hpa_ctx *c, const uint32_t *keys, const uint32_t *vals, L n, int lo_bit, int hi_bit,
                             uint32_t **keys_out, uint32_t **vals_out);
int32_t iota_u32(hpa_ctx *c, uint32_t *p, L n);
// strings and symbols (hpa_string.hip)
int32_t compare_sym(hpa_ctx *c, V z, V x, V y, int op);
int32_t member_str(hpa_ctx *c, V z, V y, V x);
int32_t str_upload(hpa_ctx *c, V x);
int32_t str_tohost(hpa_ctx *c);
void str_prune(hpa_ctx *c, L cur);
// Group a sorted key sequence into runs; emits the reference's dict layout with a flat list:
// keys = first original index of each run, values = flat list (counts, i64 indices).
int32_t runs_to_group_u64(hpa_ctx *c, V z, const uint64_t *skeys, const uint32_t *sidx, L n,
                          const uint64_t *eqkeys);
int32_t runs_to_group_u32(hpa_ctx *c, V z, const uint32_t *skeys, const uint32_t *sidx, L n, const void *src = nullptr);
// Build the dict(keys, flat list) node structure from device arrays.
int32_t make_group_dict(hpa_ctx *c, V z, int64_t *keys, int64_t *counts, int64_t *flat, L g, L n);

// flat list helpers (backend/h_memory.cc:210-215)
bool   is_flat(V x);
int32_t flat_offsets(hpa_ctx *c, V y, const int64_t **offs, const int64_t **idx, L *g, L *n);
