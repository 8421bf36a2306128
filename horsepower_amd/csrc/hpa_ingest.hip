// Table snapshots (.bin): readSerializeV / serializeV (backend/h_io.cc:630-776) with the payloads
// in HBM, and initTableFromBin (backend/load_db.cc:224-234).
//
// The format is the reference's, written node by node depth first: the 32-byte V0 header as it
// sits in memory (its pointer fields are meaningless in the file), then
//   basic types  getTypeSize(typ, len) bytes of payload (a char vector carries one more byte)
//   Q            len 8-byte name lengths, then each name with its NUL (ids are re-made on read)
//   S            len 8-byte string lengths, then each string with its NUL
//   G (list)     the len items, each a node
//   A (table)    the 24-byte A0 (row, col, foreign-key list), then the key node (symbols) and
//                the value node (a list of columns)
// The reader streams each basic payload from the file through two pinned staging buffers into
// its resident HBM allocation: a buffer is filled by eight parallel preads while the other one's
// copy to the device runs, so ingest runs at the slower of the file read and the host-to-device
// link, and a table never exists as host payloads.  Symbol names are interned into the library's table;
// strings go to one host pool per column and are mirrored into HBM as hpaUpload mirrors them.
// List cells, table cells and A0 records are long-lived host allocations (the table lives as long
// as the process).  A foreign-key list in the file is dropped (its pointer means nothing there).
#include "hpa_internal.h"
#include <atomic>
#include <stdio.h>
#include <string>
#include <thread>
#include <unistd.h>

#define BIN_STAGE (32 << 20)                 // bytes per pinned staging buffer
#define BIN_THREADS 8                        // parallel reads per staging buffer

struct BinReader {
    FILE *fp; hpa_ctx *c;
    char *stage[2]; hipEvent_t ev[2]; bool busy[2]; int cur;
    int64_t bytes;                            // payload bytes moved to the device
};

static int32_t rd(BinReader &r, void *dst, size_t n){
    if(n && fread(dst, 1, n, r.fp) != n){ set_error_detail("readSerializeV: the file ends early"); R HPA_E_GENERAL; }
    R 0;
}
// k bytes at the file's position into dst by BIN_THREADS parallel preads (one thread copying out
// of the page cache runs well below what the host-to-device link takes); the position moves on
static int32_t rd_parallel(BinReader &r, char *dst, size_t k){
    const int fd = fileno(r.fp);
    const off_t base = ftello(r.fp);
    if(base < 0) R HPA_E_GENERAL;
    std::atomic<int> bad(0);
    std::vector<std::thread> th;
    const size_t per = (k + BIN_THREADS - 1) / BIN_THREADS;
    for(int t = 0; t < BIN_THREADS; t++){
        const size_t lo = (size_t)t * per, hi = lo + per < k ? lo + per : k;
        if(lo >= hi) break;
        th.emplace_back([&, lo, hi](){
            size_t o = lo;
            while(o < hi){
                const ssize_t got = pread(fd, dst + o, hi - o, base + (off_t)o);
                if(got <= 0){ bad = 1; return; }
                o += (size_t)got;
            }
        });
    }
    for(auto &x : th) x.join();
    if(bad){ set_error_detail("readSerializeV: the file ends early"); R HPA_E_GENERAL; }
    if(fseeko(r.fp, base + (off_t)k, SEEK_SET) != 0) R HPA_E_GENERAL;
    R 0;
}
// n bytes of the file into device memory at d, through the staging buffers
static int32_t rd_device(BinReader &r, char *d, size_t n){
    size_t off = 0;
    while(off < n){
        const size_t k = n - off < (size_t)BIN_STAGE ? n - off : (size_t)BIN_STAGE;
        const int b = r.cur;
        if(r.busy[b]){ if(hipEventSynchronize(r.ev[b]) != hipSuccess) R HPA_E_DEVICE; r.busy[b] = false; }
        int32_t e; if((e = k >= (1u << 20) ? rd_parallel(r, r.stage[b], k) : rd(r, r.stage[b], k))) R e;
        if(hipMemcpyAsync(d + off, r.stage[b], k, hipMemcpyHostToDevice, r.c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipEventRecord(r.ev[b], r.c->stream) != hipSuccess) R HPA_E_DEVICE;
        r.busy[b] = true; r.cur ^= 1; off += k;
    }
    r.bytes += (int64_t)n;
    R 0;
}

static int32_t read_node(BinReader &r, V x, int depth){
    if(depth > 64){ set_error_detail("readSerializeV: nesting deeper than 64"); R HPA_E_DOMAIN; }
    hpa_V0 h; int32_t e;
    if((e = rd(r, &h, sizeof(h)))) R e;
    memset(x, 0, sizeof(*x));
    x->len = h.len; x->typ = h.typ;
    if(h.len < 0){ set_error_detail("readSerializeV: negative length"); R HPA_E_DOMAIN; }
    const L n = h.len;
    if(is_basic(x->typ) && x->typ != HPA_Q && x->typ != HPA_S){
        const L size = type_bytes(x->typ, n);
        if(n <= 1){                                         // initV keeps it in the node
            x->g = (char*)&x->b;
            char tmp[16] = {0};
            if((e = rd(r, tmp, (size_t)size))) R e;
            memcpy(&x->b, tmp, 8);
            R 0;
        }
        if((e = resident_alloc(x, x->typ, n))) R e;
        R rd_device(r, x->g, (size_t)size);
    }
    if(x->typ == HPA_Q || x->typ == HPA_S){
        std::vector<int64_t> len((size_t)n);
        if(n > 0 && (e = rd(r, len.data(), 8 * (size_t)n))) R e;
        int64_t tot = 0;
        for(L i = 0; i < n; i++){
            if(len[i] < 0){ set_error_detail("readSerializeV: negative string length"); R HPA_E_DOMAIN; }
            tot += len[i] + 1;
        }
        char *pool = (char*)malloc((size_t)(tot > 0 ? tot : 1));
        if(!pool) R HPA_E_MEMORY;
        if((e = rd(r, pool, (size_t)tot))){ free(pool); R e; }
        std::vector<uint64_t> v((size_t)(n > 0 ? n : 1));
        int64_t p = 0;
        for(L i = 0; i < n; i++){
            if(pool[p + len[i]] != 0){ free(pool); set_error_detail("readSerializeV: a string without its NUL"); R HPA_E_DOMAIN; }
            v[i] = x->typ == HPA_Q ? (uint64_t)hpaSymbol(pool + p) : (uint64_t)(uintptr_t)(pool + p);
            p += len[i] + 1;
        }
        if(x->typ == HPA_Q) free(pool);                    // names live in the symbol table
        if(n <= 1){
            x->g = (char*)&x->b; x->i64 = n == 1 ? (int64_t)v[0] : 0;
            R 0;
        }
        if(x->typ == HPA_Q){
            if((e = resident_alloc(x, HPA_Q, n))) R e;
            if(hipMemcpy(x->g, v.data(), (size_t)(8 * n), hipMemcpyHostToDevice) != hipSuccess) R HPA_E_DEVICE;
            R 0;
        }
        // strings: the pool stays (the values point into it); the pointers and a mirror of the
        // bytes move to HBM
        uint64_t *hv = (uint64_t*)malloc((size_t)(8 * n));
        if(!hv) R HPA_E_MEMORY;
        memcpy(hv, v.data(), (size_t)(8 * n));
        x->g = (char*)hv;
        e = hpaUpload(x);
        if(!e) free(hv);
        R e;
    }
    if(x->typ == HPA_G){
        V cells = (V)calloc((size_t)(n > 0 ? n : 1), sizeof(V0));
        if(!cells) R HPA_E_MEMORY;
        x->g = (char*)cells; x->g2 = nullptr;
        for(L i = 0; i < n; i++) if((e = read_node(r, cells + i, depth + 1))) R e;
        R 0;
    }
    if(x->typ == HPA_A){
        hpa_A0 a0;
        if((e = rd(r, &a0, sizeof(a0)))) R e;
        hpa_A0 *a = (hpa_A0*)malloc(sizeof(hpa_A0));
        V cells = (V)calloc(2, sizeof(V0));
        if(!a || !cells){ free(a); free(cells); R HPA_E_MEMORY; }
        a->row = a0.row; a->col = a0.col; a->fkey = nullptr;
        x->g = (char*)cells; x->p = a;
        if((e = read_node(r, cells, depth + 1)) || (e = read_node(r, cells + 1, depth + 1))) R e;
        R 0;
    }
    set_error_detail("readSerializeV: type %d", (int)x->typ);
    R HPA_E_TYPE;                                          // EP("Type not supported") there
}

// readSerializeV from the file at `path` into z, payloads resident in HBM
extern "C" int32_t hpaReadSerialized(hpa_V z, const char *path, int64_t *payload_bytes){
    hpa_ctx *c = ctx_cur();
    if(!z || !path) R HPA_E_DOMAIN;
    FILE *fp = fopen(path, "rb");
    if(!fp){ set_error_detail("cannot open %s", path); R HPA_E_GENERAL; }
    BinReader r; memset(&r, 0, sizeof(r));
    r.fp = fp; r.c = c;
    int32_t e = 0;
    for(int b = 0; b < 2 && !e; b++){
        if(hipHostMalloc((void**)&r.stage[b], BIN_STAGE, hipHostMallocDefault) != hipSuccess){ r.stage[b] = nullptr; e = HPA_E_MEMORY; }
        else if(hipEventCreateWithFlags(&r.ev[b], hipEventDisableTiming) != hipSuccess) e = HPA_E_DEVICE;
    }
    if(!e) e = read_node(r, z, 0);
    if(hipStreamSynchronize(c->stream) != hipSuccess && !e) e = HPA_E_DEVICE;
    for(int b = 0; b < 2; b++){
        if(r.stage[b]) hipHostFree(r.stage[b]);
        if(r.ev[b]) hipEventDestroy(r.ev[b]);
    }
    fclose(fp);
    if(payload_bytes) *payload_bytes = r.bytes;
    R e;
}

// initTableFromBin (load_db.cc:224-234): the table read from `path`, registered under `name`
// (registerTable) for pfnLoadTable; z receives the table node
extern "C" int32_t hpaInitTableFromBin(const char *name, const char *path, hpa_V *z){
    if(!name || !path) R HPA_E_DOMAIN;
    V t = (V)calloc(1, sizeof(V0));
    if(!t) R HPA_E_MEMORY;
    int32_t e = hpaReadSerialized(t, path, nullptr);
    if(e) R e;
    if(t->typ != HPA_A && t->typ != HPA_K){ set_error_detail("%s holds no table", path); R HPA_E_TYPE; }
    if((e = hpaRegisterTable(hpaSymbol(name), t))) R e;
    if(z) *z = t;
    R 0;
}

// ---------------------------------------------------------------- serializeV
static int32_t wr(FILE *fp, const void *p, size_t n){
    if(n && fwrite(p, 1, n, fp) != n){ set_error_detail("serializeV: write failed"); R HPA_E_GENERAL; }
    R 0;
}
// a payload's bytes on the host (device payloads are copied in staged chunks)
static int32_t wr_payload(hpa_ctx *c, FILE *fp, const char *g, size_t n, char *stage){
    if(!is_device(g)) R wr(fp, g, n);
    for(size_t off = 0; off < n; off += BIN_STAGE){
        const size_t k = n - off < (size_t)BIN_STAGE ? n - off : (size_t)BIN_STAGE;
        if(hipMemcpyAsync(stage, g + off, k, hipMemcpyDeviceToHost, c->stream) != hipSuccess) R HPA_E_DEVICE;
        if(hipStreamSynchronize(c->stream) != hipSuccess) R HPA_E_DEVICE;
        int32_t e; if((e = wr(fp, stage, k))) R e;
    }
    R 0;
}
static int32_t write_node(hpa_ctx *c, FILE *fp, V x, char *stage, int depth){
    if(depth > 64) R HPA_E_DOMAIN;
    int32_t e;
    if((e = wr(fp, x, sizeof(hpa_V0)))) R e;
    const L n = x->len;
    if(x->typ == HPA_Q || x->typ == HPA_S){
        std::vector<uint64_t> v((size_t)(n > 0 ? n : 1));
        if(n > 0){
            if(is_inline(x) || !is_device(x->g)) memcpy(v.data(), x->g, (size_t)(8 * n));
            else if((e = d2h(c, v.data(), x->g, 8 * n))) R e;
        }
        if(x->typ == HPA_S && (e = str_tohost(c))) R e;       // bytes the device wrote
        std::vector<const char*> s((size_t)(n > 0 ? n : 1));
        for(L i = 0; i < n; i++){
            s[i] = x->typ == HPA_Q ? hpaSymbolName((int64_t)v[i]) : (const char*)(uintptr_t)v[i];
            if(!s[i]) R HPA_E_DOMAIN;
        }
        for(L i = 0; i < n; i++){ int64_t l = (int64_t)strlen(s[i]); if((e = wr(fp, &l, 8))) R e; }
        for(L i = 0; i < n; i++) if((e = wr(fp, s[i], strlen(s[i]) + 1))) R e;
        R 0;
    }
    if(is_basic(x->typ)) R wr_payload(c, fp, x->g, (size_t)type_bytes(x->typ, n), stage);
    if(x->typ == HPA_G){
        if(is_flat(x)){ set_error_detail("serializeV of a flat group list"); R HPA_E_NOT_IMPL; }
        for(L i = 0; i < n; i++) if((e = write_node(c, fp, ((V)x->g) + i, stage, depth + 1))) R e;
        R 0;
    }
    if(x->typ == HPA_A){
        hpa_A0 a; memset(&a, 0, sizeof(a));
        if(x->p) a = *(const hpa_A0*)x->p;
        if((e = wr(fp, &a, sizeof(a)))) R e;
        if((e = write_node(c, fp, (V)x->g, stage, depth + 1))) R e;
        R write_node(c, fp, ((V)x->g) + 1, stage, depth + 1);
    }
    R HPA_E_TYPE;
}
// serializeV of x (device or host payloads) to the file at `path`
extern "C" int32_t hpaWriteSerialized(hpa_V x, const char *path){
    hpa_ctx *c = ctx_cur();
    if(!x || !path) R HPA_E_DOMAIN;
    FILE *fp = fopen(path, "wb");
    if(!fp){ set_error_detail("cannot create %s", path); R HPA_E_GENERAL; }
    char *stage = nullptr;
    int32_t e = hipHostMalloc((void**)&stage, BIN_STAGE, hipHostMallocDefault) == hipSuccess ? 0 : HPA_E_MEMORY;
    if(!e) e = write_node(c, fp, x, stage, 0);
    if(stage) hipHostFree(stage);
    if(fclose(fp) != 0 && !e) e = HPA_E_GENERAL;
    R e;
}
