// Run-time code generation of fused loops.
//
// The reference's optimiser turns fused sections into C text (optimizer/elementwise.cc:239-281,
// autofusion.cc:770-935, pattern.cc:180-219) that a C compiler builds offline.  Here each fused
// section is emitted as HIP C++ for the running gfx950 device, compiled once with hipRTC and
// cached by its source text; the postfix expressions of the public API become straight-line
// typed C expressions, so the kernels stream their columns at HBM speed instead of interpreting
// an instruction list per row.
//
// Kernels generated here:
//   fa_reduce  FA reduction with filter (sum / avg / min / max of a body over rows where cond)
//   fe_map     FE element-wise map into a typed output
//   fga        filter + group on one or two byte-wide key columns + N aggregates, one pass:
//              a 64K-entry LDS table maps each key code to a block-local slot (first come,
//              first served), each thread keeps register accumulators per slot, slots are
//              reduced per block into partials; fga_merge combines the partials per key code
//              (eight blocks over slices of the partials, folded by fga_merge2, when there
//              are few groups) in a fixed order (deterministic) and emits groups in ascending
//              code order, the
//              order of pfnGroupTrie (h_primitive.cc:2656-2681).  More than NS distinct codes
//              in one block is reported as an overflow and the caller takes the unfused path.
#include "hpa_internal.h"
#include "hpa_expr.h"
#include <hip/hiprtc.h>
#include <mutex>
#include <sstream>
#include <stdio.h>

#define FGA_NS 16          // largest slot count tried before the unfused fallback
#define FGA_ROWS 4         // consecutive rows per thread and iteration (one vector load per column)
#define FGA_BLOCK 256
#define FGA_MS 8           // merge blocks, each over one slice of the partial slots
#define FGA_MCAP 64        // most groups the sliced merge takes (more: one block merges all); fga_merge2
                           // folds them with one 64-thread block

static std::mutex g_jit_mu;
struct JitFn { hipModule_t mod; hipFunction_t fn; };
static std::map<std::string, JitFn> g_jit;
static int g_jit_disabled = -1;
static double g_jit_compile_ms = 0;

bool jit_enabled(){
    if(g_jit_disabled < 0){ const char *e = getenv("HPA_NO_JIT"); g_jit_disabled = (e && *e && *e != '0') ? 1 : 0; }
    R g_jit_disabled == 0;
}

static std::string arch_name(){
    hipDeviceProp_t p; int d = 0; hipGetDevice(&d);
    if(hipGetDeviceProperties(&p, d) == hipSuccess){
        std::string a = p.gcnArchName;
        size_t k = a.find(':'); if(k != std::string::npos) a = a.substr(0, k);
        if(!a.empty()) R a;
    }
    R "gfx950";
}

static bool g_compile_only = false;       // diagnostic: generate + compile, never load or launch

static int32_t jit_get(const std::string &src, const char *kname, hipFunction_t *fn){
    std::lock_guard<std::mutex> lk(g_jit_mu);
    std::string key = std::string(kname) + "\n" + src;
    auto it = g_jit.find(key);
    if(it != g_jit.end()){ *fn = it->second.fn; R 0; }
    hipEvent_t t0, t1;
    hiprtcProgram prog;
    if(hiprtcCreateProgram(&prog, src.c_str(), "fused.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS){
        set_error_detail("hiprtcCreateProgram failed"); R HPA_E_DEVICE;
    }
    std::string arch = "--offload-arch=" + arch_name();
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    (void)t0; (void)t1;
    hiprtcResult r = hiprtcCompileProgram(prog, 8, opts);
    if(r != HIPRTC_SUCCESS){
        size_t ls = 0; hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0'); hiprtcGetProgramLog(prog, &log[0]);
        set_error_detail("hipRTC compile of %s failed: %.400s", kname, log.c_str());
        hiprtcDestroyProgram(&prog);
        R HPA_E_DEVICE;
    }
    size_t cs = 0; hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs); hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if(g_compile_only){ *fn = nullptr; R cs > 0 ? 0 : HPA_E_DEVICE; }
    JitFn j;
    if(hipModuleLoadData(&j.mod, code.data()) != hipSuccess){ set_error_detail("hipModuleLoadData failed"); R HPA_E_DEVICE; }
    if(hipModuleGetFunction(&j.fn, j.mod, kname) != hipSuccess){ set_error_detail("hipModuleGetFunction %s failed", kname); R HPA_E_DEVICE; }
    g_jit[key] = j;
    *fn = j.fn;
    R 0;
}

// Second-level cache: generated kernels found by a compact description of what they compute
// (expression code with constants as argument slots, column types, shapes), so a repeated call
// builds no source text at all.
static std::map<std::string, hipFunction_t> g_jit_by_key;
template<typename Gen>
static int32_t jit_get_keyed(const std::string &key, const char *kname, Gen gen, hipFunction_t *fn){
    const std::string k = key + "\n" + kname;
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        auto it = g_jit_by_key.find(k);
        if(it != g_jit_by_key.end()){ *fn = it->second; R 0; }
    }
    int32_t e = jit_get(gen(), kname, fn);
    if(!e && !g_compile_only){ std::lock_guard<std::mutex> lk(g_jit_mu); g_jit_by_key[k] = *fn; }
    R e;
}
static std::string types_key(int nx, const int32_t *typ){
    std::string k;
    for(int i = 0; i < nx; i++){ k += std::to_string(typ[i]); k += ','; }
    R k;
}

// A prepared request of one generated-kernel kind: its function and the constants its launch
// passes, kept by the request's raw bytes (kind, column types, expression code), so a repeated
// request skips building the C text and its key.
static std::string raw_key(const char *tag, int nx, const int32_t *typ, int extra, const hpa_expr *e, const hpa_expr *e2 = nullptr){
    std::string k(tag);
    k.append((const char*)&nx, sizeof(nx)); k.append((const char*)typ, sizeof(int32_t) * (size_t)nx);
    k.append((const char*)&extra, sizeof(extra));
    for(const hpa_expr *x : {e, e2}) if(x){ k.append((const char*)&x->n, sizeof(x->n)); k.append((const char*)x->code, sizeof(hpa_ins) * (size_t)x->n); }
    R k;
}

template<typename A>
static int32_t jit_launch(hpa_ctx *c, hipFunction_t fn, unsigned grid, unsigned block, A &args){
    void *params[] = {&args};
    if(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, c->stream, params, nullptr) != hipSuccess)
        R check_launch("jit launch") ? HPA_E_DEVICE : HPA_E_DEVICE;
    R check_launch("jit launch");
}

// ---------------------------------------------------------------- expression -> C
static const char *ctype(int32_t t){
    switch(t){
        case HPA_B: R "unsigned char"; case HPA_J: R "signed char"; case HPA_C: R "char";
        case HPA_H: R "short";
        case HPA_I: case HPA_M: case HPA_D: case HPA_U: case HPA_W: case HPA_T: R "int";
        case HPA_L: case HPA_Z: case HPA_Q: R "long long";
        case HPA_F: R "float"; case HPA_E: R "double";
    }
    R nullptr;
}

struct CExpr { std::string s; bool f; };

// Constants of the expressions are kernel arguments (a.ki / a.kf), not literals, so one compiled
// kernel serves every value of a query's constants (a serving loop sees new predicate constants
// with each request).  With kp == nullptr the constants are written as literals (used as a key to
// recognise equal aggregate bodies).
#define JIT_MAX_CONST 32
#define KONST_FIELDS "i64 ki[32]; double kf[32];"
struct Konst { int ni = 0, nf = 0; long long i[JIT_MAX_CONST]; double f[JIT_MAX_CONST]; };
// prepared requests by raw key (raw_key): the function and the launch's constants
struct JitPrep { hipFunction_t fn = nullptr; Konst kc; std::vector<int> used; };
static std::map<std::string, JitPrep> g_jit_prep;
static bool prep_find(const std::string &k, JitPrep *p){
    std::lock_guard<std::mutex> lk(g_jit_mu);
    auto it = g_jit_prep.find(k);
    if(it == g_jit_prep.end()) R false;
    *p = it->second; R true;
}
static void prep_store(const std::string &k, const JitPrep &p){
    if(g_compile_only) R;
    std::lock_guard<std::mutex> lk(g_jit_mu);
    if(g_jit_prep.size() >= 4096) g_jit_prep.clear();
    g_jit_prep[k] = p;
}

static int32_t expr_to_c(const hpa_expr *e, int nx, const int32_t *typ, CExpr *out, std::vector<int> &used,
                         Konst *kp = nullptr){
    if(!e || e->n <= 0 || e->n > 256) R HPA_E_DOMAIN;
    std::vector<CExpr> st;
    char buf[128];
    for(int k = 0; k < e->n; k++){
        const hpa_ins &in = e->code[k];
        switch(in.op){
            case HPA_OP_COL: {
                if(in.arg < 0 || in.arg >= nx || !ctype(typ[in.arg])) R HPA_E_TYPE;
                used[in.arg] = 1;
                bool f = typ[in.arg] == HPA_F || typ[in.arg] == HPA_E;
                snprintf(buf, sizeof(buf), f ? "((double)v%d)" : "((i64)v%d)", in.arg);
                st.push_back({buf, f}); break;
            }
            case HPA_OP_CONST_I:
                if(kp){
                    if(kp->ni >= JIT_MAX_CONST) R HPA_E_DOMAIN;
                    kp->i[kp->ni] = in.imm_i;
                    snprintf(buf, sizeof(buf), "((i64)a.ki[%d])", kp->ni++);
                } else if(in.imm_i == INT64_MIN) snprintf(buf, sizeof(buf), "(-9223372036854775807LL-1)");
                else snprintf(buf, sizeof(buf), "((i64)%lldLL)", (long long)in.imm_i);
                st.push_back({buf, false}); break;
            case HPA_OP_CONST_F:
                if(kp){
                    if(kp->nf >= JIT_MAX_CONST) R HPA_E_DOMAIN;
                    kp->f[kp->nf] = in.imm_f;
                    snprintf(buf, sizeof(buf), "(a.kf[%d])", kp->nf++);
                } else snprintf(buf, sizeof(buf), "(%a)", in.imm_f);
                st.push_back({buf, true}); break;
            case HPA_OP_NOT: {
                if(st.empty()) R HPA_E_DOMAIN;
                CExpr a = st.back(); st.pop_back();
                st.push_back({"((" + a.s + ") == 0 ? (i64)1 : (i64)0)", false}); break;
            }
            case HPA_OP_NEG: {
                if(st.empty()) R HPA_E_DOMAIN;
                CExpr a = st.back(); st.pop_back();
                st.push_back({a.f ? "(-(" + a.s + "))" : "((i64)(0ULL - (u64)(" + a.s + ")))", a.f}); break;
            }
            default: {
                if(st.size() < 2) R HPA_E_DOMAIN;
                CExpr b = st.back(); st.pop_back(); CExpr a = st.back(); st.pop_back();
                bool f = a.f || b.f;
                std::string A = f && !a.f ? "((double)" + a.s + ")" : a.s;
                std::string B = f && !b.f ? "((double)" + b.s + ")" : b.s;
                const char *op = nullptr;
                switch(in.op){
                    case HPA_OP_ADD: op = "+"; break; case HPA_OP_SUB: op = "-"; break; case HPA_OP_MUL: op = "*"; break;
                    case HPA_OP_DIV: op = "/"; break; case HPA_OP_LT: op = "<"; break; case HPA_OP_LEQ: op = "<="; break;
                    case HPA_OP_GT: op = ">"; break; case HPA_OP_GEQ: op = ">="; break; case HPA_OP_EQ: op = "=="; break;
                    case HPA_OP_NEQ: op = "!="; break; case HPA_OP_AND: op = "&&"; break; case HPA_OP_OR: op = "||"; break;
                    default: R HPA_E_DOMAIN;
                }
                if(in.op <= HPA_OP_MUL){
                    if(f) st.push_back({"(" + A + " " + op + " " + B + ")", true});
                    else st.push_back({"((i64)((u64)(" + A + ") " + op + " (u64)(" + B + ")))", false});
                } else if(in.op == HPA_OP_DIV){
                    if(f) st.push_back({"(" + A + " / " + B + ")", true});
                    else st.push_back({"hdiv(" + A + ", " + B + ")", false});
                } else if(in.op == HPA_OP_AND || in.op == HPA_OP_OR){
                    st.push_back({"(((" + a.s + ") != 0 " + op + " (" + b.s + ") != 0) ? (i64)1 : (i64)0)", false});
                } else {
                    st.push_back({"((" + A + " " + op + " " + B + ") ? (i64)1 : (i64)0)", false});
                }
            }
        }
    }
    if(st.size() != 1) R HPA_E_DOMAIN;
    *out = st[0];
    R 0;
}

static const char *PREAMBLE = R"(
typedef long long i64; typedef unsigned long long u64;
#define NCOL 12
template<typename T> struct __attribute__((aligned(4 * sizeof(T)))) V4 { T v[4]; };
struct KS { double s, c; };
__device__ __forceinline__ i64 hdiv(i64 a, i64 b){ return b == 0 ? 0 : (b == -1 ? (i64)(0ULL - (u64)a) : a / b); }
__device__ __forceinline__ void ks_add(KS &a, double v){
    double t = a.s + v;
    a.c += (__builtin_fabs(a.s) >= __builtin_fabs(v)) ? (a.s - t) + v : (v - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ void ks_merge(KS &a, const KS &b){
    double t = a.s + b.s; double bb = t - a.s;
    double err = (a.s - (t - bb)) + (b.s - bb);
    a.s = t; a.c += b.c + err;
}
__device__ __forceinline__ KS ks_wave(KS a){
    for(int o = 32; o > 0; o >>= 1){ KS b; b.s = __shfl_down(a.s, o, 64); b.c = __shfl_down(a.c, o, 64); ks_merge(a, b); }
    return a;
}
)";

static std::string loads_code(int nx, const int32_t *typ, const std::vector<int> &used){
    std::ostringstream o;
    for(int k = 0; k < nx; k++) if(used[k]) o << "        const " << ctype(typ[k]) << " v" << k << " = ((const " << ctype(typ[k]) << "*)a.c[" << k << "])[i];\n";
    R o.str();
}


static const char *COND_TRUE = "((i64)1)";

// ---------------------------------------------------------------- FA reduce (fa_reduce)
struct JArgs { const void *c[EXPR_MAX_COLS]; long long n; void *out; void *aux; long long ki[JIT_MAX_CONST]; double kf[JIT_MAX_CONST]; };


// columns read only for selected rows (used by the body, not by the condition)
static void split_used(int nx, const std::vector<int> &uc, const std::vector<int> &ub, std::vector<int> &late){
    late.assign(nx, 0);
    for(int k = 0; k < nx; k++) if(ub[k] && !uc[k]) late[k] = 1;
}

// The generated loops read 4 consecutive rows with one vector load: every column they read must
// start on a 4-element boundary (allocation starts always do; anything else takes the
// interpreted path).
static bool cols_aligned(int nx, const int32_t *typ, const std::vector<int> &used, const DevCols &cols){
    for(int k = 0; k < nx; k++)
        if(used[k] && ((uintptr_t)cols.p[k] % (uintptr_t)(4 * type_width(typ[k]))) != 0) R false;
    R true;
}

// Row loop shared by the generated FA reductions: 4 consecutive rows per thread and iteration,
// visited in row order.  Condition columns are read with one vector load per column; columns that
// only the body reads are loaded when some of the four rows passes some condition (at low
// selectivity most of their cache lines are then never fetched).  update(q) is the code run for
// a row that passes condition q, with the row's values bound to v<k> and the body value to `v`.
template<typename F>
static std::string fa_row_loop(int nx, const int32_t *typ, const std::vector<int> &uc, const std::vector<int> &late,
                               const std::vector<CExpr> &conds, const CExpr &body, F update){
    const int nq = (int)conds.size();
    std::ostringstream s;
    auto bind = [&](const std::vector<int> &cols, const char *w, int u){
        std::ostringstream b;
        for(int k = 0; k < nx; k++) if(cols[k]) b << " const " << ctype(typ[k]) << " v" << k << " = " << w << k << ".v[" << u << "];";
        return b.str();
    };
    const char *bt = body.f ? "double" : "i64";
    s << "    const i64 stride = (i64)gridDim.x * 1024;\n"
         "    for(i64 base = ((i64)blockIdx.x * 256 + threadIdx.x) * 4; base < a.n; base += stride){\n"
         "        if(base + 4 <= a.n){\n";
    for(int k = 0; k < nx; k++) if(uc[k])
        s << "            const V4<" << ctype(typ[k]) << "> w" << k << " = *(const V4<" << ctype(typ[k]) << ">*)((const " << ctype(typ[k]) << "*)a.c[" << k << "] + base);\n";
    for(int u = 0; u < 4; u++) for(int q = 0; q < nq; q++) s << "            bool m" << u << "_" << q << ";\n";
    for(int u = 0; u < 4; u++){
        s << "            {" << bind(uc, "w", u);
        for(int q = 0; q < nq; q++) s << " m" << u << "_" << q << " = (" << conds[q].s << ") != 0;";
        s << " }\n";
    }
    s << "            if(";
    for(int u = 0; u < 4; u++) for(int q = 0; q < nq; q++) s << (u || q ? " | " : "") << "m" << u << "_" << q;
    s << "){\n";
    for(int k = 0; k < nx; k++) if(late[k])
        s << "                const V4<" << ctype(typ[k]) << "> l" << k << " = *(const V4<" << ctype(typ[k]) << ">*)((const " << ctype(typ[k]) << "*)a.c[" << k << "] + base);\n";
    for(int u = 0; u < 4; u++){
        s << "                if(";
        for(int q = 0; q < nq; q++) s << (q ? " | " : "") << "m" << u << "_" << q;
        s << "){" << bind(uc, "w", u) << bind(late, "l", u) << "\n"
          << "                    const " << bt << " v = " << body.s << ";\n";
        for(int q = 0; q < nq; q++) s << "                    if(m" << u << "_" << q << "){ " << update(q) << " }\n";
        s << "                }\n";
    }
    s << "            }\n"
         "        } else {\n"
         "            for(i64 i = base; i < a.n; i++){\n";
    std::vector<int> all(nx, 0);
    for(int k = 0; k < nx; k++) all[k] = uc[k] | late[k];
    s << loads_code(nx, typ, all);
    for(int q = 0; q < nq; q++)
        s << "                if((" << conds[q].s << ") != 0){ const " << bt << " v = " << body.s << "; " << update(q) << " }\n";
    s << "            }\n        }\n    }\n";
    R s.str();
}

// FA-reduce through generated code; HPA_E_NOT_IMPL when code generation is unavailable
int32_t jit_fused_reduce(hpa_ctx *c, int kind, int nx, V *x, const hpa_expr *cond, const hpa_expr *body,
                         const DevCols &cols, L n, void *part, int G){
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    int32_t typ[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) typ[k] = x[k]->typ;
    const std::string rk = raw_key("fa", nx, typ, cond ? 1 : 0, body, cond);
    JitPrep prep;
    int32_t e;
    if(prep_find(rk, &prep)){ if(!cols_aligned(nx, typ, prep.used, cols)) R HPA_E_NOT_IMPL; }
    else {
    std::vector<int> uc(nx, 0), ub(nx, 0), late;
    CExpr ce{COND_TRUE, false}, be;
    Konst &kc = prep.kc;
    if(cond && (e = expr_to_c(cond, nx, typ, &ce, uc, &kc))) R e;
    if((e = expr_to_c(body, nx, typ, &be, ub, &kc))) R e;
    split_used(nx, uc, ub, late);
    prep.used.assign(nx, 0);
    for(int k = 0; k < nx; k++) prep.used[k] = uc[k] | ub[k];
    if(!cols_aligned(nx, typ, prep.used, cols)) R HPA_E_NOT_IMPL;
    auto gen = [&](){
    std::ostringstream s;
    s << PREAMBLE;
    s << "struct A { const void *c[" << EXPR_MAX_COLS << "]; i64 n; void *out; void *aux; " KONST_FIELDS " };\n"
         "struct P { KS s; u64 i, cnt; double mn, mx; i64 imn, imx; };\n"
         "extern \"C\" __global__ void __launch_bounds__(256) fa_reduce(A a){\n"
         "    __shared__ P sh[4];\n"
         "    P p; p.s.s = 0; p.s.c = 0; p.i = 0; p.cnt = 0; p.mn = 1.7976931348623157e308; p.mx = -1.7976931348623157e308;\n"
         "    p.imn = 9223372036854775807LL; p.imx = -9223372036854775807LL - 1;\n"
      << fa_row_loop(nx, typ, uc, late, std::vector<CExpr>{ce}, be, [&](int){
             return be.f ? std::string("p.cnt++; ks_add(p.s, v); p.mn = v < p.mn ? v : p.mn; p.mx = v > p.mx ? v : p.mx;")
                         : std::string("p.cnt++; p.i += (u64)v; p.imn = v < p.imn ? v : p.imn; p.imx = v > p.imx ? v : p.imx;"); });
    s << "    for(int o = 32; o > 0; o >>= 1){\n"
         "        KS b; b.s = __shfl_down(p.s.s, o, 64); b.c = __shfl_down(p.s.c, o, 64); ks_merge(p.s, b);\n"
         "        p.i += __shfl_down(p.i, o, 64); p.cnt += __shfl_down(p.cnt, o, 64);\n"
         "        double m1 = __shfl_down(p.mn, o, 64), m2 = __shfl_down(p.mx, o, 64);\n"
         "        i64 q1 = __shfl_down(p.imn, o, 64), q2 = __shfl_down(p.imx, o, 64);\n"
         "        p.mn = m1 < p.mn ? m1 : p.mn; p.mx = m2 > p.mx ? m2 : p.mx; p.imn = q1 < p.imn ? q1 : p.imn; p.imx = q2 > p.imx ? q2 : p.imx;\n"
         "    }\n"
         "    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = p;\n"
         "    __syncthreads();\n"
         "    if(threadIdx.x == 0){\n"
         "        P r = sh[0];\n"
         "        for(int w = 1; w < 4; w++){ P b = sh[w]; ks_merge(r.s, b.s); r.i += b.i; r.cnt += b.cnt;\n"
         "            r.mn = b.mn < r.mn ? b.mn : r.mn; r.mx = b.mx > r.mx ? b.mx : r.mx; r.imn = b.imn < r.imn ? b.imn : r.imn; r.imx = b.imx > r.imx ? b.imx : r.imx; }\n"
         "        ((P*)a.out)[blockIdx.x] = r;\n"
         "    }\n"
         "}\n";
    return s.str();
    };
    const std::string key = "fa|" + std::to_string(nx) + "|" + types_key(nx, typ) + "|" + ce.s + "|" + be.s;
    if((e = jit_get_keyed(key, "fa_reduce", gen, &prep.fn))) R e;
    prep_store(rk, prep);
    }
    JArgs a; memset(&a, 0, sizeof(a));
    for(int k = 0; k < nx; k++) a.c[k] = cols.p[k];
    memcpy(a.ki, prep.kc.i, sizeof(a.ki)); memcpy(a.kf, prep.kc.f, sizeof(a.kf));
    a.n = n; a.out = part;
    double cb = 0; for(int k = 0; k < nx; k++) cb += (double)type_width(typ[k]);
    ProfScope ps("fused_reduce", c->stream, (double)n * cb);
    R jit_launch(c, prep.fn, (unsigned)G, 256, a);
}

// ---------------------------------------------------------------- shared-scan FA reduce (fa_many)
// nq reductions of one body under nq conditions in one pass: each row's columns are loaded once
// and every condition adds the body to its own Neumaier accumulator (sum / avg only).  Partials
// are laid out [query][block] in the FaPart layout of fa_reduce.
int32_t jit_fused_reduce_many(hpa_ctx *c, int nq, int nx, V *x, const hpa_expr *conds, const hpa_expr *body,
                              const DevCols &cols, L n, void *part, int G){
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    int32_t typ[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) typ[k] = x[k]->typ;
    std::vector<int> uc(nx, 0), ub(nx, 0), late, used(nx, 0);
    Konst kc;
    std::vector<CExpr> ce(nq);
    CExpr be; int32_t e;
    for(int q = 0; q < nq; q++) if((e = expr_to_c(&conds[q], nx, typ, &ce[q], uc, &kc))) R e;
    if((e = expr_to_c(body, nx, typ, &be, ub, &kc))) R e;
    split_used(nx, uc, ub, late);
    for(int k = 0; k < nx; k++) used[k] = uc[k] | ub[k];
    if(!cols_aligned(nx, typ, used, cols)) R HPA_E_NOT_IMPL;
    auto gen = [&](){
    std::ostringstream s;
    s << PREAMBLE << "#define NQ " << nq << "\n"
      << "struct A { const void *c[" << EXPR_MAX_COLS << "]; i64 n; void *out; void *aux; " KONST_FIELDS " };\n"
         "struct P { KS s; u64 i, cnt; double mn, mx; i64 imn, imx; };\n"
         "extern \"C\" __global__ void __launch_bounds__(256) fa_many(A a){\n"
         "    __shared__ KS shs[4][NQ]; __shared__ u64 shi[4][NQ], shc[4][NQ];\n";
    for(int q = 0; q < nq; q++) s << "    KS s" << q << "; s" << q << ".s = 0; s" << q << ".c = 0; u64 i" << q << " = 0, c" << q << " = 0;\n";
    s << fa_row_loop(nx, typ, uc, late, ce, be, [&](int q){
        std::string qs = std::to_string(q);
        return be.f ? "c" + qs + "++; ks_add(s" + qs + ", v);" : "c" + qs + "++; i" + qs + " += (u64)v;"; });
    s << "    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;\n";
    for(int q = 0; q < nq; q++){
        s << "    s" << q << " = ks_wave(s" << q << ");\n"
             "    for(int o = 32; o > 0; o >>= 1){ i" << q << " += __shfl_down(i" << q << ", o, 64); c" << q << " += __shfl_down(c" << q << ", o, 64); }\n"
             "    if(lane == 0){ shs[w][" << q << "] = s" << q << "; shi[w][" << q << "] = i" << q << "; shc[w][" << q << "] = c" << q << "; }\n";
    }
    s << "    __syncthreads();\n"
         "    if(threadIdx.x < NQ){\n"
         "        const int q = threadIdx.x;\n"
         "        P r; r.s = shs[0][q]; r.i = shi[0][q]; r.cnt = shc[0][q];\n"
         "        for(int ww = 1; ww < 4; ww++){ ks_merge(r.s, shs[ww][q]); r.i += shi[ww][q]; r.cnt += shc[ww][q]; }\n"
         "        r.mn = 1.7976931348623157e308; r.mx = -1.7976931348623157e308; r.imn = 9223372036854775807LL; r.imx = -9223372036854775807LL - 1;\n"
         "        ((P*)a.out)[(i64)q * gridDim.x + blockIdx.x] = r;\n"
         "    }\n}\n";
    return s.str();
    };
    hipFunction_t fn;
    std::string key = "fam|" + std::to_string(nx) + "|" + types_key(nx, typ) + "|" + be.s;
    for(int q = 0; q < nq; q++) key += "|" + ce[q].s;
    if((e = jit_get_keyed(key, "fa_many", gen, &fn))) R e;
    JArgs a; memset(&a, 0, sizeof(a));
    for(int k = 0; k < nx; k++) a.c[k] = cols.p[k];
    memcpy(a.ki, kc.i, sizeof(a.ki)); memcpy(a.kf, kc.f, sizeof(a.kf));
    a.n = n; a.out = part;
    double cb = 0; for(int k = 0; k < nx; k++) if(used[k]) cb += (double)type_width(typ[k]);
    ProfScope ps("fused_reduce_many", c->stream, (double)n * cb);
    R jit_launch(c, fn, (unsigned)G, 256, a);
}

// ---------------------------------------------------------------- FE map (fe_map)
int32_t jit_fused_map(hpa_ctx *c, int otyp, int nx, V *x, const hpa_expr *ex, const DevCols &cols, L n, void *dz){
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    int32_t typ[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) typ[k] = x[k]->typ;
    const std::string rk = raw_key("fe", nx, typ, otyp, ex);
    JitPrep prep; int32_t e;
    std::vector<int> used(nx, 0);
    const bool have = prep_find(rk, &prep);
    if(!have){
    CExpr be;
    Konst &kc = prep.kc;
    if((e = expr_to_c(ex, nx, typ, &be, used, &kc))) R e;
    const char *ot = otyp == HPA_B ? "unsigned char" : otyp == HPA_I ? "int" : otyp == HPA_L ? "long long" : "double";
    auto gen = [&](){
    std::ostringstream s;
    s << PREAMBLE
      << "struct A { const void *c[" << EXPR_MAX_COLS << "]; i64 n; void *out; void *aux; " KONST_FIELDS " };\n"
         "extern \"C\" __global__ void __launch_bounds__(256) fe_map(A a){\n"
         "    for(i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (i64)gridDim.x * 256){\n"
      << loads_code(nx, typ, used);
    if(otyp == HPA_B) s << "        ((unsigned char*)a.out)[i] = (unsigned char)((" << be.s << ") != 0);\n";
    else s << "        ((" << ot << "*)a.out)[i] = (" << ot << ")(" << be.s << ");\n";
    s << "    }\n}\n";
    return s.str();
    };
    const std::string key = "fe|" + std::to_string(otyp) + "|" + std::to_string(nx) + "|" + types_key(nx, typ) + "|" + be.s;
    if((e = jit_get_keyed(key, "fe_map", gen, &prep.fn))) R e;
    prep_store(rk, prep);
    } else {
        // the columns the expression reads, for the scope's bytes
        for(int i = 0; i < ex->n; i++) if(ex->code[i].op == HPA_OP_COL && ex->code[i].arg >= 0 && ex->code[i].arg < nx) used[ex->code[i].arg] = 1;
    }
    JArgs a; memset(&a, 0, sizeof(a));
    for(int k = 0; k < nx; k++) a.c[k] = cols.p[k];
    memcpy(a.ki, prep.kc.i, sizeof(a.ki)); memcpy(a.kf, prep.kc.f, sizeof(a.kf));
    a.n = n; a.out = dz;
    double cb = 0; for(int k = 0; k < nx; k++) if(used[k]) cb += (double)type_width(typ[k]);
    ProfScope ps("fused_map", c->stream, (double)n * (cb + type_width(otyp)));
    R jit_launch(c, prep.fn, (unsigned)grid_for(n, 256, 4), 256, a);
}

// ---------------------------------------------------------------- FE map of a compress (cp_map)
// The byte-mask write pass of the stream compaction (4096-row tiles, thread t owns the 16
// consecutive rows [16t, 16t + 16), its selection bits `mb` and the tile offsets `toff` from the
// count pass) with each selected row's value computed from the map's columns: the columns are
// read with 16-row vector loads, the selected values compacted in LDS and written out by all
// threads together.  Every column must start on a 16-byte boundary.
int32_t jit_compress_map(hpa_ctx *c, int otyp, const DevCols &cols, const hpa_expr *ex, const uint16_t *mb,
                         const int64_t *toff, L n, void *dz){
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    const int nx = cols.n;
    for(int k = 0; k < nx; k++) if((uintptr_t)cols.p[k] & 15) R HPA_E_NOT_IMPL;
    const std::string rk = raw_key("cpm", nx, cols.typ, otyp, ex);
    JitPrep prep; int32_t e;
    if(!prep_find(rk, &prep)){
    std::vector<int> used(nx, 0);
    CExpr be;
    Konst &kc = prep.kc;
    if((e = expr_to_c(ex, nx, cols.typ, &be, used, &kc))) R e;
    const char *ot = otyp == HPA_B ? "unsigned char" : otyp == HPA_I ? "int" : otyp == HPA_L ? "long long" : "double";
    const std::string val = otyp == HPA_B ? "(unsigned char)((" + be.s + ") != 0)" : "(" + std::string(ot) + ")(" + be.s + ")";
    auto gen = [&](){
    std::ostringstream s;
    s << PREAMBLE
      << "template<typename T> struct __attribute__((aligned(16))) R16 { T v[16]; };\n"
         "struct A { const void *c[" << EXPR_MAX_COLS << "]; i64 n; void *out; const unsigned short *mb; const i64 *toff; " KONST_FIELDS " };\n"
         "extern \"C\" __global__ void __launch_bounds__(256) cp_map(A a){\n"
         "    __shared__ " << ot << " buf[4096];\n"
         "    __shared__ unsigned sw[4];\n"
         "    const int t = threadIdx.x, lane = t & 63, w = t >> 6;\n"
         "    const i64 i0 = (i64)blockIdx.x * 4096 + (i64)t * 16;\n"
         "    const unsigned m = a.mb[(i64)blockIdx.x * 256 + t];\n"
         "    const bool full = i0 + 16 <= a.n;\n";
    // the thread's column loads go out as soon as its selection bits are known, before the scan
    // and the barrier (none for 16 unselected rows)
    for(int k = 0; k < nx; k++) if(used[k])
        s << "    R16<" << ctype(cols.typ[k]) << "> r" << k << ";\n"
          << "    if(m && full) r" << k << " = *(const R16<" << ctype(cols.typ[k]) << ">*)((const " << ctype(cols.typ[k]) << "*)a.c[" << k << "] + i0);\n";
    s << "    const unsigned cnt = __popc(m);\n"
         "    unsigned inc = cnt;\n"
         "    for(int o = 1; o < 64; o <<= 1){ unsigned y = __shfl_up(inc, o, 64); if(lane >= o) inc += y; }\n"
         "    if(lane == 63) sw[w] = inc;\n"
         "    __syncthreads();\n"
         "    unsigned wbase = 0, total = 0;\n"
         "    for(int q = 0; q < 4; q++){ if(q < w) wbase += sw[q]; total += sw[q]; }\n"
         "    if(total == 0) return;\n"
         "    unsigned off = wbase + inc - cnt;\n"
         "    if(m && full){\n";
    s << "        #pragma unroll\n"
         "        for(int k = 0; k < 16; k++) if((m >> k) & 1u){\n";
    for(int k = 0; k < nx; k++) if(used[k]) s << "            const " << ctype(cols.typ[k]) << " v" << k << " = r" << k << ".v[k];\n";
    s << "            buf[off++] = " << val << ";\n"
         "        }\n"
         "    } else if(m){\n"
         "        for(int k = 0; k < 16; k++) if((m >> k) & 1u){\n"
         "            const i64 i = i0 + k;\n"
      << loads_code(nx, cols.typ, used)
      << "            buf[off++] = " << val << ";\n"
         "        }\n"
         "    }\n"
         "    __syncthreads();\n"
         "    const i64 pre = a.toff[blockIdx.x];\n"
         "    for(unsigned j = t; j < total; j += 256) ((" << ot << "*)a.out)[pre + j] = buf[j];\n"
         "}\n";
    return s.str();
    };
    const std::string key = "cpm|" + std::to_string(otyp) + "|" + std::to_string(nx) + "|" + types_key(nx, cols.typ) + "|" + be.s;
    if((e = jit_get_keyed(key, "cp_map", gen, &prep.fn))) R e;
    prep_store(rk, prep);
    }
    struct CArgs { const void *c[EXPR_MAX_COLS]; long long n; void *out; const uint16_t *mb; const int64_t *toff;
                   long long ki[JIT_MAX_CONST]; double kf[JIT_MAX_CONST]; } a;
    memset(&a, 0, sizeof(a));
    for(int k = 0; k < nx; k++) a.c[k] = cols.p[k];
    memcpy(a.ki, prep.kc.i, sizeof(a.ki)); memcpy(a.kf, prep.kc.f, sizeof(a.kf));
    a.n = n; a.out = dz; a.mb = mb; a.toff = toff;
    R jit_launch(c, prep.fn, (unsigned)ceil_div(n, 4096), 256, a);
}

// ---------------------------------------------------------------- condition count of a compress (cp_cnt)
// The count pass of the FA compress scan with its condition evaluated in place of a byte mask
// (4096-row tiles, thread t owns the 16 consecutive rows [16t, 16t + 16)): the used columns are
// read with 16-row vector loads, the condition's 16 selection bits go to mb[tile * 256 + t] for
// the write pass and their count to tcount[tile].  Every column must start on a 16-byte boundary.
int32_t jit_compress_count(hpa_ctx *c, const DevCols &cols, const hpa_expr *cond, L n, uint16_t *mb, uint32_t *tcount){
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    const int nx = cols.n;
    for(int k = 0; k < nx; k++) if((uintptr_t)cols.p[k] & 15) R HPA_E_NOT_IMPL;
    const std::string rk = raw_key("cpc", nx, cols.typ, 0, cond);
    JitPrep prep; int32_t e;
    if(!prep_find(rk, &prep)){
    std::vector<int> used(nx, 0);
    CExpr be;
    Konst &kc = prep.kc;
    if((e = expr_to_c(cond, nx, cols.typ, &be, used, &kc))) R e;
    auto gen = [&](){
    std::ostringstream s;
    s << PREAMBLE
      << "template<typename T> struct __attribute__((aligned(16))) R16 { T v[16]; };\n"
         "struct A { const void *c[" << EXPR_MAX_COLS << "]; i64 n; unsigned short *mb; unsigned *tcount; " KONST_FIELDS " };\n"
         "extern \"C\" __global__ void __launch_bounds__(256) cp_cnt(A a){\n"
         "    __shared__ unsigned sw[4];\n"
         "    const int t = threadIdx.x;\n"
         "    const i64 i0 = (i64)blockIdx.x * 4096 + (i64)t * 16;\n"
         "    unsigned m = 0;\n"
         "    if(i0 + 16 <= a.n){\n";
    for(int k = 0; k < nx; k++) if(used[k])
        s << "        const R16<" << ctype(cols.typ[k]) << "> r" << k << " = *(const R16<" << ctype(cols.typ[k]) << ">*)((const "
          << ctype(cols.typ[k]) << "*)a.c[" << k << "] + i0);\n";
    s << "        #pragma unroll\n"
         "        for(int k = 0; k < 16; k++){\n";
    for(int k = 0; k < nx; k++) if(used[k]) s << "            const " << ctype(cols.typ[k]) << " v" << k << " = r" << k << ".v[k];\n";
    s << "            if(" << be.s << ") m |= 1u << k;\n"
         "        }\n"
         "    } else {\n"
         "        for(int k = 0; k < 16 && i0 + k < a.n; k++){\n"
         "            const i64 i = i0 + k;\n"
      << loads_code(nx, cols.typ, used)
      << "            if(" << be.s << ") m |= 1u << k;\n"
         "        }\n"
         "    }\n"
         "    a.mb[(i64)blockIdx.x * 256 + t] = (unsigned short)m;\n"
         "    unsigned c = __popc(m);\n"
         "    for(int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);\n"
         "    if((t & 63) == 0) sw[t >> 6] = c;\n"
         "    __syncthreads();\n"
         "    if(t == 0) a.tcount[blockIdx.x] = sw[0] + sw[1] + sw[2] + sw[3];\n"
         "}\n";
    return s.str();
    };
    const std::string key = "cpc|" + std::to_string(nx) + "|" + types_key(nx, cols.typ) + "|" + be.s;
    if((e = jit_get_keyed(key, "cp_cnt", gen, &prep.fn))) R e;
    prep_store(rk, prep);
    }
    struct CArgs { const void *c[EXPR_MAX_COLS]; long long n; uint16_t *mb; uint32_t *tcount;
                   long long ki[JIT_MAX_CONST]; double kf[JIT_MAX_CONST]; } a;
    memset(&a, 0, sizeof(a));
    for(int k = 0; k < nx; k++) a.c[k] = cols.p[k];
    memcpy(a.ki, prep.kc.i, sizeof(a.ki)); memcpy(a.kf, prep.kc.f, sizeof(a.kf));
    a.n = n; a.mb = mb; a.tcount = tcount;
    R jit_launch(c, prep.fn, (unsigned)ceil_div(n, 4096), 256, a);
}

// ---------------------------------------------------------------- fused group + aggregates (fga)
// Accumulator classes: sums (and avgs, which share the sum of the same body), mins, maxes.
struct Acc { std::string body; bool f; int cls; };   // cls 0 sum, 2 min, 3 max

// bk (may be null): a bucket number per row, kept in code bits 16 .. 16 + bbits (bbits <= 2), so
// groups are (bucket, key bytes); the merge's code bitmap grows to match
static std::string fga_source(int nk, int ns, int nx, const int32_t *typ, const CExpr &ce,
                              const std::vector<int> &used, const std::vector<Acc> &acc, const CExpr *bk = nullptr, int bbits = 0){
    const int NA = (int)acc.size() > 0 ? (int)acc.size() : 1;
    const int U = FGA_ROWS;
    std::ostringstream s;
    s << PREAMBLE;
    s << "#define NS " << ns << "\n#define NA " << NA << "\n#define HT 64\n#define U " << U << "\n";
    s << "template<typename T> struct __attribute__((aligned(U * sizeof(T)))) VU { T v[U]; };\n"
         // streaming column loads marked nontemporal: each row is read once
         "typedef unsigned nt4 __attribute__((ext_vector_type(4)));\n"
         "typedef unsigned nt2 __attribute__((ext_vector_type(2)));\n"
         "template<typename T> __device__ __forceinline__ VU<T> ldnt(const T *p){\n"
         "    VU<T> r;\n"
         "    if(sizeof(VU<T>) % 16 == 0){\n"
         "        #pragma unroll\n"
         "        for(int k = 0; k < (int)(sizeof(VU<T>) / 16); k++){ nt4 v = __builtin_nontemporal_load((const nt4*)p + k); __builtin_memcpy((char*)&r + 16 * k, &v, 16); }\n"
         "    } else if(sizeof(VU<T>) == 8){ nt2 v = __builtin_nontemporal_load((const nt2*)p); __builtin_memcpy(&r, &v, 8); }\n"
         "    else if(sizeof(VU<T>) == 4){ unsigned v = __builtin_nontemporal_load((const unsigned*)p); __builtin_memcpy(&r, &v, 4); }\n"
         "    else r = *(const VU<T>*)p;\n"
         "    return r;\n"
         "}\n"
         "struct A { const void *c[" << EXPR_MAX_COLS << "]; const void *k0; const void *k1; i64 n; void *part; unsigned *used; unsigned kx0, kx1; " KONST_FIELDS " };\n"
         // a partial slot; the slots' codes live in their own array after the G * NS slots, so the
         // merge tests 64 slots' codes with one line and loads only the slots that match
         "struct P { u64 cnt; i64 first; double f[2 * NA]; i64 q[NA]; };\n"
         // LDS hash of key code -> slot: entry = code << 8 | slot; 0xFFFFFFFF empty; slot 0xFE busy, 0xFD overflow
         "__device__ __forceinline__ unsigned slot_of(unsigned *ht, unsigned *s_code, unsigned *s_used, unsigned code){\n"
         "    unsigned h = ((code * 2654435761u) >> 26) & (HT - 1);\n"
         "    unsigned probes = 0;\n"
         "    while(true){\n"
         "        unsigned e = __hip_atomic_load(&ht[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);\n"
         "        if(e == 0xFFFFFFFFu){\n"
         "            unsigned want = (code << 8) | 0xFEu;\n"
         "            unsigned prev = atomicCAS(&ht[h], 0xFFFFFFFFu, want);\n"
         "            if(prev == 0xFFFFFFFFu){\n"
         "                unsigned n = atomicAdd(s_used, 1u);\n"
         "                unsigned sl = n < NS ? n : 0xFDu;\n"
         "                if(n < NS) s_code[n] = code;\n"
         "                __threadfence_block();\n"
         "                atomicExch(&ht[h], (code << 8) | sl);\n"
         "                return sl;\n"
         "            }\n"
         "            continue;\n"
         "        }\n"
         "        if((e >> 8) == code){\n"
         "            unsigned sl = e & 0xFFu;\n"
         "            if(sl == 0xFEu){ __builtin_amdgcn_s_sleep(1); continue; }\n"
         "            return sl;\n"
         "        }\n"
         "        h = (h + 1) & (HT - 1);\n"
         "        if(++probes >= HT) return 0xFDu;      // table full: more codes than the kernel keeps\n"
         "    }\n"
         "}\n"
         "extern \"C\" __global__ void __launch_bounds__(256) fga(A a){\n"
         "    __shared__ int s_ovf;\n"
         "    __shared__ unsigned ht[HT];\n"
         "    __shared__ unsigned s_code[NS];\n"
         "    __shared__ unsigned s_used;\n"
         "    __shared__ double shf[4][2 * NA + 1];\n"
         "    __shared__ i64 shq[4][NA + 2];\n"
         "    if(threadIdx.x < HT) ht[threadIdx.x] = 0xFFFFFFFFu;\n"
         "    if(threadIdx.x == 0){ s_used = 0; s_ovf = 0; }\n"
         "    __syncthreads();\n";
    // accumulators: count and first row (as the iteration-relative offset it*U+u) per slot
    for(int q = 0; q < ns; q++){
        s << "    unsigned cnt" << q << " = 0, fst" << q << " = 0xFFFFFFFFu;";
        for(int k = 0; k < (int)acc.size(); k++){
            const Acc &A = acc[k];
            const char *init = A.cls == 0 ? "0" : (A.f ? (A.cls == 2 ? "1.7976931348623157e308" : "-1.7976931348623157e308")
                                                       : (A.cls == 2 ? "9223372036854775807LL" : "(-9223372036854775807LL-1)"));
            s << (A.f ? " double" : " i64") << " a" << k << "_" << q << " = " << init << ";";
        }
        s << "\n";
    }
    // one row: values v<k> of the used columns, key bytes kb0/kb1, offset `off` = it*U+u
    std::ostringstream row;
    row << "            if(" << ce.s << "){\n";
    // signed key types have their sign bit flipped, so ascending codes are ascending keys
    // (the bucket is masked to bbits: a value past them cannot index past the merge's code bitmap)
    const std::string bpart = bk ? "((((unsigned)(" + bk->s + ")) & " + std::to_string((1 << bbits) - 1) + "u) << 16) | " : "";
    if(nk == 2) row << "                const unsigned code = " << bpart << "(((unsigned)kb0 ^ a.kx0) << 8) | ((unsigned)kb1 ^ a.kx1);\n";
    else row << "                const unsigned code = " << bpart << "((unsigned)kb0 ^ a.kx0);\n";
    row << "                const unsigned sl = slot_of(ht, s_code, &s_used, code);\n"
           "                if(sl >= NS){ __hip_atomic_store(&s_ovf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); goto fga_done; }\n";
    for(int k = 0; k < (int)acc.size(); k++) row << "                const " << (acc[k].f ? "double" : "i64") << " b" << k << " = " << acc[k].body << ";\n";
    row << "                switch(sl){\n";
    for(int q = 0; q < ns; q++){
        row << "                case " << q << ": if(cnt" << q << "++ == 0) fst" << q << " = off;";
        for(int k = 0; k < (int)acc.size(); k++){
            std::string v = "a" + std::to_string(k) + "_" + std::to_string(q), b = "b" + std::to_string(k);
            if(acc[k].cls == 0) row << (acc[k].f ? " " + v + " += " + b + ";" : " " + v + " = (i64)((u64)" + v + " + (u64)" + b + ");");
            else if(acc[k].cls == 2) row << " " << v << " = " << b << " < " << v << " ? " << b << " : " << v << ";";
            else row << " " << v << " = " << b << " > " << v << " ? " << b << " : " << v << ";";
        }
        row << " break;\n";
    }
    row << "                }\n            }\n";
    s << "    const i64 tid = (i64)blockIdx.x * 256 + threadIdx.x;\n"
         "    const i64 stride = (i64)gridDim.x * 256 * U;\n"
         "    unsigned it = 0;\n"
         "    for(i64 base = tid * U; base < a.n; base += stride, it++){\n"
         "        if(__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;\n"
         "        if(base + U <= a.n){\n";
    // vector loads of U consecutive rows
    for(int k = 0; k < nx; k++) if(used[k])
        s << "            const VU<" << ctype(typ[k]) << "> w" << k << " = ldnt<" << ctype(typ[k]) << ">((const " << ctype(typ[k]) << "*)a.c[" << k << "] + base);\n";
    s << "            const VU<unsigned char> wk0 = ldnt<unsigned char>((const unsigned char*)a.k0 + base);\n";
    if(nk == 2) s << "            const VU<unsigned char> wk1 = ldnt<unsigned char>((const unsigned char*)a.k1 + base);\n";
    for(int u = 0; u < U; u++){
        s << "          {\n            const unsigned off = it * U + " << u << ";\n";
        for(int k = 0; k < nx; k++) if(used[k]) s << "            const " << ctype(typ[k]) << " v" << k << " = w" << k << ".v[" << u << "];\n";
        s << "            const unsigned char kb0 = wk0.v[" << u << "];" << (nk == 2 ? " const unsigned char kb1 = wk1.v[" + std::to_string(u) + "];" : "") << "\n";
        s << row.str() << "          }\n";
    }
    s << "        } else {\n"
         "            for(int u = 0; u < U && base + u < a.n; u++){\n"
         "                const i64 i = base + u; const unsigned off = it * U + u;\n"
      << loads_code(nx, typ, used)
      << "                const unsigned char kb0 = ((const unsigned char*)a.k0)[i];"
      << (nk == 2 ? " const unsigned char kb1 = ((const unsigned char*)a.k1)[i];" : "") << "\n"
      << row.str()
      << "            }\n        }\n    }\n"
         "fga_done:\n"
         "    __syncthreads();\n"
         "    const unsigned used = s_used < NS ? s_used : NS;\n"
         "    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;\n"
         "    P *out = (P*)a.part + (i64)blockIdx.x * NS;\n";
    // per-slot block reduction, fixed order
    for(int q = 0; q < ns; q++){
        s << "    if(" << q << " < used){\n"
             "        u64 c = cnt" << q << "; i64 f = fst" << q << " == 0xFFFFFFFFu ? 9223372036854775807LL : tid * U + (i64)(fst" << q << " / U) * stride + (fst" << q << " % U);\n"
             "        for(int o = 32; o > 0; o >>= 1){ c += __shfl_down(c, o, 64); i64 t = __shfl_down(f, o, 64); f = t < f ? t : f; }\n";
        for(int k = 0; k < (int)acc.size(); k++){
            std::string v = "a" + std::to_string(k) + "_" + std::to_string(q);
            if(acc[k].f && acc[k].cls == 0)
                s << "        KS k" << k << "; k" << k << ".s = " << v << "; k" << k << ".c = 0; k" << k << " = ks_wave(k" << k << ");\n";
            else if(acc[k].f)
                s << "        double k" << k << " = " << v << "; for(int o = 32; o > 0; o >>= 1){ double t = __shfl_down(k" << k << ", o, 64); k" << k << " = "
                  << (acc[k].cls == 2 ? "t < k" : "t > k") << k << " ? t : k" << k << "; }\n";
            else if(acc[k].cls == 0)
                s << "        u64 k" << k << " = (u64)" << v << "; for(int o = 32; o > 0; o >>= 1) k" << k << " += __shfl_down(k" << k << ", o, 64);\n";
            else
                s << "        i64 k" << k << " = " << v << "; for(int o = 32; o > 0; o >>= 1){ i64 t = __shfl_down(k" << k << ", o, 64); k" << k << " = "
                  << (acc[k].cls == 2 ? "t < k" : "t > k") << k << " ? t : k" << k << "; }\n";
        }
        s << "        if(lane == 0){ shq[w][NA] = (i64)c; shq[w][NA + 1] = f;";
        for(int k = 0; k < (int)acc.size(); k++){
            if(acc[k].f && acc[k].cls == 0) s << " shf[w][2*" << k << "] = k" << k << ".s; shf[w][2*" << k << "+1] = k" << k << ".c;";
            else if(acc[k].f) s << " shf[w][2*" << k << "] = k" << k << ";";
            else s << " shq[w][" << k << "] = (i64)k" << k << ";";
        }
        s << " }\n        __syncthreads();\n"
             "        if(threadIdx.x == 0){\n"
             "            P r; r.cnt = 0; r.first = 9223372036854775807LL;\n"
             "            ((unsigned*)((P*)a.part + (i64)gridDim.x * NS))[(i64)blockIdx.x * NS + " << q << "] = s_code[" << q << "];\n"
             "            for(int ww = 0; ww < 4; ww++){ r.cnt += (u64)shq[ww][NA]; r.first = shq[ww][NA + 1] < r.first ? shq[ww][NA + 1] : r.first; }\n";
        for(int k = 0; k < (int)acc.size(); k++){
            if(acc[k].f && acc[k].cls == 0)
                s << "            { KS t; t.s = 0; t.c = 0; for(int ww = 0; ww < 4; ww++){ KS u; u.s = shf[ww][2*" << k << "]; u.c = shf[ww][2*" << k << "+1]; ks_merge(t, u); } r.f[2*" << k << "] = t.s; r.f[2*" << k << "+1] = t.c; }\n";
            else if(acc[k].f)
                s << "            { double t = shf[0][2*" << k << "]; for(int ww = 1; ww < 4; ww++){ double u = shf[ww][2*" << k << "]; t = " << (acc[k].cls == 2 ? "u < t" : "u > t") << " ? u : t; } r.f[2*" << k << "] = t; }\n";
            else if(acc[k].cls == 0)
                s << "            { u64 t = 0; for(int ww = 0; ww < 4; ww++) t += (u64)shq[ww][" << k << "]; r.q[" << k << "] = (i64)t; }\n";
            else
                s << "            { i64 t = shq[0][" << k << "]; for(int ww = 1; ww < 4; ww++){ i64 u = shq[ww][" << k << "]; t = " << (acc[k].cls == 2 ? "u < t" : "u > t") << " ? u : t; } r.q[" << k << "] = t; }\n";
        }
        s << "            out[" << q << "] = r;\n        }\n        __syncthreads();\n    }\n";
    }
    // bit 31 of the used count: this block met more codes than slots (the merge reports it)
    s << "    if(threadIdx.x == 0) a.used[blockIdx.x] = used | (__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ? 0x80000000u : 0u);\n}\n";

    // merge kernels.  fga_merge: FGA_MS blocks of 1024 threads.  Every block lists the codes
    // present (the partials' codes, one thread per slot); with at most FGA_MCAP groups each block
    // takes one slice of the slots and writes a record per group, and fga_merge2 folds a group's
    // FGA_MS records in block order (a fixed association, so runs give identical sums).  More
    // groups: block 0 merges every slot and writes the results itself.
    auto acc_init = [&](std::ostringstream &o, const char *ind){
        o << ind << "u64 c = 0; i64 fs = 9223372036854775807LL;\n";
        for(int k = 0; k < (int)acc.size(); k++){
            if(acc[k].f && acc[k].cls == 0) o << ind << "KS t" << k << "; t" << k << ".s = 0; t" << k << ".c = 0;\n";
            else if(acc[k].f) o << ind << "double t" << k << " = " << (acc[k].cls == 2 ? "1.7976931348623157e308" : "-1.7976931348623157e308") << ";\n";
            else if(acc[k].cls == 0) o << ind << "u64 t" << k << " = 0;\n";
            else o << ind << "i64 t" << k << " = " << (acc[k].cls == 2 ? "9223372036854775807LL" : "(-9223372036854775807LL-1)") << ";\n";
        }
    };
    auto acc_final = [&](std::ostringstream &o, const char *ind){
        o << ind << "m.cnt[r] = c; m.first[r] = fs;\n";
        for(int k = 0; k < (int)acc.size(); k++){
            if(acc[k].f && acc[k].cls == 0) o << ind << "m.f[(i64)r * NA + " << k << "] = t" << k << ".s + t" << k << ".c;\n";
            else if(acc[k].f) o << ind << "m.f[(i64)r * NA + " << k << "] = t" << k << ";\n";
            else o << ind << "m.q[(i64)r * NA + " << k << "] = (i64)t" << k << ";\n";
        }
    };
    const int cwt = 2 << bbits;                              // code-bitmap words per merge thread
    s << "#define MS " << FGA_MS << "\n#define MCAP " << FGA_MCAP << "\n#define CWT " << cwt << "\n"
         "struct M { const void *part; const unsigned *used; int G; unsigned *codes; int *g; u64 *cnt; i64 *first; double *f; i64 *q;\n"
         "           u64 *rc; i64 *rfs; double *rf; i64 *rq; int *ovf; };\n"
         "extern \"C\" __global__ void __launch_bounds__(1024) fga_merge(M m){\n"
         "    __shared__ unsigned bits[1024 * CWT];\n"
         "    __shared__ unsigned wsum[16];\n"
         "    __shared__ int sovf;\n"
         "    __shared__ unsigned scodes[MCAP];\n"
         "    __shared__ u64 mc[16]; __shared__ i64 mfs[16]; __shared__ double mf[16][2 * NA + 1]; __shared__ i64 mq[16][NA + 1];\n"
         "    const P *part = (const P*)m.part;\n"
         "    const unsigned *pcode = (const unsigned*)(part + (i64)m.G * NS);\n"
         "    const int SL = m.G * NS;\n"
         "    for(int k = threadIdx.x; k < 1024 * CWT; k += 1024) bits[k] = 0;\n"
         "    if(threadIdx.x == 0) sovf = 0;\n"
         "    __syncthreads();\n"
         "    // one thread per slot: the used count and the code load together; bit 31 of a used\n"
         "    // count is the block's overflow\n"
         "    #pragma unroll 4\n"
         "    for(int i = threadIdx.x; i < SL; i += 1024){\n"
         "        const unsigned uw = m.used[i / NS], cd = pcode[i];\n"
         "        if((uw >> 31) && (i % NS) == 0) sovf = 1;\n"
         "        if((unsigned)(i % NS) < (uw & 0x7FFFFFFFu)) atomicOr(&bits[cd >> 5], 1u << (cd & 31));\n"
         "    }\n"
         "    __syncthreads();\n"
         "    // list position of this thread's codes: a scan inside each wave, then the wave totals\n"
         "    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;\n"
         "    unsigned cc = 0;\n"
         "    #pragma unroll\n"
         "    for(int h = 0; h < CWT; h++) cc += __popc(bits[CWT * threadIdx.x + h]);\n"
         "    unsigned inc = cc;\n"
         "    for(int o = 1; o < 64; o <<= 1){ const unsigned t = __shfl_up(inc, o, 64); if(lane >= o) inc += t; }\n"
         "    if(lane == 63) wsum[wv] = inc;\n"
         "    __syncthreads();\n"
         "    unsigned pos = inc - cc, gt = 0;\n"
         "    for(int q = 0; q < 16; q++){ const unsigned t = wsum[q]; if(q < wv) pos += t; gt += t; }\n"
         "    const bool b0 = blockIdx.x == 0;\n"
         "    for(int h = 0; h < CWT; h++){ unsigned wd = bits[CWT * threadIdx.x + h]; while(wd){ int bt = __ffs(wd) - 1; wd &= wd - 1;\n"
         "        const unsigned cv = (unsigned)((CWT * threadIdx.x + h) * 32 + bt); if(b0) m.codes[pos] = cv; if(pos < MCAP) scodes[pos] = cv; pos++; } }\n"
         "    const int g = (int)gt;\n"
         "    const bool split = g <= MCAP && gridDim.x > 1;\n"
         "    if(b0 && threadIdx.x == 0){ *m.g = g; *m.ovf = sovf; }\n"
         "    if(!split && !b0) return;\n"
         "    __syncthreads();\n"
         "    __threadfence_block();\n"
         "    const int per = split ? (SL + (int)gridDim.x - 1) / (int)gridDim.x : SL;\n"
         "    const int lo = split ? (int)blockIdx.x * per : 0, hi = lo + per < SL ? lo + per : SL;\n"
         "    // W waves per group (16 / W groups at a time); each wave takes every W-th run of 64\n"
         "    // partial slots of the block's range, reduces it with shuffles, and the group's first\n"
         "    // wave merges the W wave results in wave order\n"
         "    const int W = g >= 16 ? 1 : (g >= 8 ? 2 : (g >= 4 ? 4 : (g >= 2 ? 8 : 16)));\n"
         "    const int sub = wv % W;\n"
         "    for(int r0 = 0; r0 < g; r0 += 16 / W){\n"
         "        const int r = r0 + wv / W;\n"
         "        const unsigned code = r < g ? (split ? scodes[r] : m.codes[r]) : 0xFFFFFFFFu;\n";
    {
        std::ostringstream o; acc_init(o, "        "); s << o.str();
    }
    // partial slots taken in a flat order (slot i = block i / NS, entry i % NS): the loads of
    // successive iterations do not depend on each other and overlap
    s << "        #pragma unroll 4\n"
         "        for(int i = lo + sub * 64 + lane; i < hi; i += 64 * W){\n"
         "            const unsigned u = m.used[i / NS] & 0x7FFFFFFFu, cd = pcode[i];\n"
         "            if(r >= g || (unsigned)(i % NS) >= u || cd != code) continue;\n"
         "            {\n"
         "                const P &p = part[i];\n"
         "                c += p.cnt; fs = p.first < fs ? p.first : fs;\n";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << "                { KS u; u.s = p.f[2*" << k << "]; u.c = p.f[2*" << k << "+1]; ks_merge(t" << k << ", u); }\n";
        else if(acc[k].f) s << "                t" << k << " = p.f[2*" << k << "] " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? p.f[2*" << k << "] : t" << k << ";\n";
        else if(acc[k].cls == 0) s << "                t" << k << " += (u64)p.q[" << k << "];\n";
        else s << "                t" << k << " = p.q[" << k << "] " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? p.q[" << k << "] : t" << k << ";\n";
    }
    s << "            }\n        }\n"
         "        for(int o = 32; o > 0; o >>= 1){\n"
         "            c += __shfl_down(c, o, 64); { i64 t = __shfl_down(fs, o, 64); fs = t < fs ? t : fs; }\n";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << "            { KS u; u.s = __shfl_down(t" << k << ".s, o, 64); u.c = __shfl_down(t" << k << ".c, o, 64); ks_merge(t" << k << ", u); }\n";
        else if(acc[k].f) s << "            { double u = __shfl_down(t" << k << ", o, 64); t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
        else if(acc[k].cls == 0) s << "            t" << k << " += __shfl_down(t" << k << ", o, 64);\n";
        else s << "            { i64 u = __shfl_down(t" << k << ", o, 64); t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
    }
    s << "        }\n"
         "        if(lane == 0){ mc[wv] = c; mfs[wv] = fs;";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << " mf[wv][2*" << k << "] = t" << k << ".s; mf[wv][2*" << k << "+1] = t" << k << ".c;";
        else if(acc[k].f) s << " mf[wv][2*" << k << "] = t" << k << ";";
        else s << " mq[wv][" << k << "] = (i64)t" << k << ";";
    }
    s << " }\n"
         "        __syncthreads();\n"
         "        if(lane == 0 && sub == 0 && r < g){\n"
         "            for(int q = wv + 1; q < wv + W; q++){\n"
         "                c += mc[q]; fs = mfs[q] < fs ? mfs[q] : fs;\n";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << "                { KS u; u.s = mf[q][2*" << k << "]; u.c = mf[q][2*" << k << "+1]; ks_merge(t" << k << ", u); }\n";
        else if(acc[k].f) s << "                { double u = mf[q][2*" << k << "]; t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
        else if(acc[k].cls == 0) s << "                t" << k << " += (u64)mq[q][" << k << "];\n";
        else s << "                { i64 u = mq[q][" << k << "]; t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
    }
    s << "            }\n"
         "            if(split){\n"
         "                const i64 rec = (i64)r * MS + blockIdx.x;\n"
         "                m.rc[rec] = c; m.rfs[rec] = fs;\n";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << "                m.rf[rec * (2 * NA) + 2*" << k << "] = t" << k << ".s; m.rf[rec * (2 * NA) + 2*" << k << "+1] = t" << k << ".c;\n";
        else if(acc[k].f) s << "                m.rf[rec * (2 * NA) + 2*" << k << "] = t" << k << ";\n";
        else s << "                m.rq[rec * NA + " << k << "] = (i64)t" << k << ";\n";
    }
    s << "            } else {\n";
    {
        std::ostringstream o; acc_final(o, "                "); s << o.str();
    }
    s << "            }\n"
         "        }\n        __syncthreads();\n";
    s << "    }\n}\n";
    // fga_merge2: with at most MCAP groups, block 0 folds each group's FGA_MS records in block
    // order (one thread per group); then every group's typed outputs are written (the layout of
    // FgaOut: key bytes, first row, and per output a count, an average, or the accumulator)
    s << "struct EO { int na, stride, nk; unsigned kx0, kx1; int kind[16]; int accid[16]; int accf[16]; void *out[16];\n"
         "            unsigned char *key0, *key1; i64 *first; unsigned char *bucket; };\n"
         "struct M2 { M m; EO o; };\n"
         "__device__ __forceinline__ void fga_emit(const M &m, const EO &o, int r){\n"
         "    const unsigned cd = m.codes[r];\n"
         "    if(o.bucket) o.bucket[r] = (unsigned char)(cd >> 16);\n"
         "    if(o.nk == 2){ o.key0[r] = (unsigned char)(((cd >> 8) & 255) ^ o.kx0); o.key1[r] = (unsigned char)((cd & 255) ^ o.kx1); }\n"
         "    else o.key0[r] = (unsigned char)((cd & 255) ^ o.kx0);\n"
         "    o.first[r] = m.first[r];\n"
         "    for(int a = 0; a < o.na; a++){\n"
         "        const i64 at = (i64)r * o.stride + o.accid[a];\n"
         "        switch(o.kind[a]){\n"
         "            case 4: ((i64*)o.out[a])[r] = (i64)m.cnt[r]; break;\n"
         "            case 1:      // avg = sum / len\n"
         "                if(o.accf[a]) ((double*)o.out[a])[r] = m.f[at] / (double)m.cnt[r];\n"
         "                else ((double*)o.out[a])[r] = (double)m.q[at] * 1.0 / (double)m.cnt[r];\n"
         "                break;\n"
         "            default:\n"
         "                if(o.accf[a]) ((double*)o.out[a])[r] = m.f[at];\n"
         "                else ((i64*)o.out[a])[r] = m.q[at];\n"
         "        }\n"
         "    }\n"
         "}\n"
         "extern \"C\" __global__ void __launch_bounds__(64) fga_merge2(M2 a){\n"
         "    const M &m = a.m;\n"
         "    const int g = *m.g;\n"
         "    if(g > MCAP){\n"
         "        for(int r = blockIdx.x * 64 + threadIdx.x; r < g; r += gridDim.x * 64) fga_emit(m, a.o, r);\n"
         "        return;\n"
         "    }\n"
         "    const int r = threadIdx.x;\n"
         "    if(blockIdx.x != 0 || r >= g) return;\n";
    {
        std::ostringstream o; acc_init(o, "    "); s << o.str();
    }
    s << "    #pragma unroll\n"
         "    for(int b = 0; b < MS; b++){\n"
         "        const i64 rec = (i64)r * MS + b;\n"
         "        c += m.rc[rec]; fs = m.rfs[rec] < fs ? m.rfs[rec] : fs;\n";
    for(int k = 0; k < (int)acc.size(); k++){
        if(acc[k].f && acc[k].cls == 0) s << "        { KS u; u.s = m.rf[rec * (2 * NA) + 2*" << k << "]; u.c = m.rf[rec * (2 * NA) + 2*" << k << "+1]; ks_merge(t" << k << ", u); }\n";
        else if(acc[k].f) s << "        { double u = m.rf[rec * (2 * NA) + 2*" << k << "]; t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
        else if(acc[k].cls == 0) s << "        t" << k << " += (u64)m.rq[rec * NA + " << k << "];\n";
        else s << "        { i64 u = m.rq[rec * NA + " << k << "]; t" << k << " = u " << (acc[k].cls == 2 ? "<" : ">") << " t" << k << " ? u : t" << k << "; }\n";
    }
    s << "    }\n";
    {
        std::ostringstream o; acc_final(o, "    "); s << o.str();
    }
    s << "    fga_emit(m, a.o, r);\n";
    s << "}\n";
    R s.str();
}

// outputs of the merge are laid out [group][accumulator]; fga_merge2 makes typed vectors of
// them through this layout (mirrored by EO in the generated source)
struct FgaOut { int na, stride, nk; unsigned kx0, kx1; int kind[16]; int accid[16]; int accf[16]; void *out[16]; uint8_t *key0, *key1; int64_t *first;
                uint8_t *bucket; };
// a one-element result is inline: its value comes from `host` when given, else from the device
static void set_dev_vec(hpa_ctx *c, V z, int32_t typ, L len, void *p, const void *host = nullptr){
    z->typ = typ; z->len = len;
    if(len <= 1){
        z->g = (char*)&z->b; z->i64 = 0;
        if(len == 1){ if(host) memcpy(&z->b, host, (size_t)type_width(typ)); else d2h(c, &z->b, p, type_width(typ)); }
    }
    else z->g = (char*)p;
}
// host destinations of hpaFusedGroupAggHost
struct FgaHost { L cap; void **keys; int64_t *first; void **aggs; int64_t *ng; uint8_t *bucket = nullptr; };

// hpaFusedGroupAgg: see the public header.  Returns HPA_E_NOT_IMPL (nothing written) when a
// block meets more distinct key codes than it has slots; the caller then runs the unfused plan.
// prepared inputs of one distinct fused group-aggregate request (see hpaFusedGroupAgg)
// fn[i]: the fga, fga_merge and fga_merge2 kernels with 4 << i slots, once compiled
#define FGA_NSV 3
struct FgaPrep { CExpr ce; Konst kc; std::vector<Acc> acc; std::vector<int> used; FgaOut o; hipFunction_t fn[FGA_NSV][3] = {};
                 CExpr bk; bool has_bk = false; };
static std::map<std::string, FgaPrep> g_fga_prep;
static int32_t fga_run(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                       int32_t na, const hpa_agg *aggs, hpa_V zfirst, hpa_V *zkeys, hpa_V *zaggs, const FgaHost *hd,
                       const hpa_expr *bucket = nullptr, int bbits = 0){
    if(bbits < 0 || bbits > 2 || (bucket && bbits == 0) || (!bucket && bbits)) R HPA_E_DOMAIN;
    hpa_ctx *c = ctx_cur();
    if(nk < 1 || nk > 2 || na < 1 || na > 16 || nx < 0 || nx > EXPR_MAX_COLS) R HPA_E_DOMAIN;
    if(!jit_enabled()) R HPA_E_NOT_IMPL;
    L n = keys[0]->len;
    for(int k = 0; k < nk; k++){
        int32_t t = keys[k]->typ;
        if(!(t == HPA_C || t == HPA_B || t == HPA_J)) R HPA_E_TYPE;
        if(keys[k]->len != n) R HPA_E_LENGTH;
    }
    DevCols cols; L nn; int32_t e;
    L mark = c->dcur;
    if((e = cols_view(c, nx, x, &cols, &nn))) R e;
    if(nx > 0 && nn != n) R HPA_E_LENGTH;
    int32_t typ[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) typ[k] = x[k]->typ;
    // The generated-code inputs (condition and accumulator sources, constants, output layout)
    // depend only on the request's bytes: they are prepared once per distinct request and reused.
    std::string raw;
    raw.append((const char*)&nk, 4); raw.append((const char*)&nx, 4); raw.append((const char*)typ, 4 * nx);
    auto add_expr = [&](const hpa_expr *ex){
        int32_t m = ex ? ex->n : -1; raw.append((const char*)&m, 4);
        if(ex && m > 0) raw.append((const char*)ex->code, sizeof(hpa_ins) * (size_t)m);
    };
    add_expr(cond);
    add_expr(bucket); raw.append((const char*)&bbits, 4);
    raw.append((const char*)&na, 4);
    for(int a = 0; a < na; a++){ raw.append((const char*)&aggs[a].kind, 4); add_expr(aggs[a].kind == 4 ? nullptr : aggs[a].body); }
    FgaPrep prep;
    FgaPrep *stored = nullptr;                 // the map entry (std::map entries do not move)
    bool have = false;
    {
        std::lock_guard<std::mutex> lk(g_jit_mu);
        auto it = g_fga_prep.find(raw);
        if(it != g_fga_prep.end()){ prep = it->second; stored = &it->second; have = true; }
    }
    if(!have){
        std::vector<int> uc(nx, 0), ub(nx, 0);
        prep.ce = CExpr{COND_TRUE, false};
        if(cond && (e = expr_to_c(cond, nx, typ, &prep.ce, uc, &prep.kc))) R e;
        if(bucket){ if((e = expr_to_c(bucket, nx, typ, &prep.bk, uc, &prep.kc))) R e; prep.has_bk = true; }
        std::vector<std::string> acc_key;                    // bodies with literal constants
        memset(&prep.o, 0, sizeof(prep.o));
        prep.o.na = na; prep.o.nk = nk;
        for(int a = 0; a < na; a++){
            int kd = aggs[a].kind;
            if(kd < 0 || kd > 4) R HPA_E_DOMAIN;
            prep.o.kind[a] = kd; prep.o.accid[a] = 0; prep.o.accf[a] = 0;
            if(kd == 4) continue;
            CExpr lit; if((e = expr_to_c(aggs[a].body, nx, typ, &lit, ub))) R e;
            int cls = (kd == 0 || kd == 1) ? 0 : kd;
            int id = -1;
            for(int k = 0; k < (int)prep.acc.size(); k++) if(acc_key[k] == lit.s && prep.acc[k].cls == cls){ id = k; break; }
            if(id < 0){
                CExpr be; if((e = expr_to_c(aggs[a].body, nx, typ, &be, ub, &prep.kc))) R e;
                id = (int)prep.acc.size(); prep.acc.push_back({be.s, be.f, cls}); acc_key.push_back(lit.s);
            }
            prep.o.accid[a] = id; prep.o.accf[a] = lit.f ? 1 : 0;
        }
        prep.used.assign(nx, 0);
        for(int k = 0; k < nx; k++) prep.used[k] = uc[k] | ub[k];
        std::lock_guard<std::mutex> lk(g_jit_mu);
        stored = &(g_fga_prep[raw] = prep);
    }
    const CExpr &ce = prep.ce;
    const Konst &kc = prep.kc;
    const std::vector<Acc> &acc = prep.acc;
    const std::vector<int> &used = prep.used;
    FgaOut o = prep.o;
    if(!cols_aligned(nx, typ, used, cols)) R HPA_E_NOT_IMPL;
    const int NA = acc.empty() ? 1 : (int)acc.size();
    o.stride = NA;
    const size_t psize = 8 + 8 + 16 * NA + 8 * NA + 4;     // struct P, and the slot's code
    // Workgroups per CU, measured on MI355X (Q1): with one workgroup merging every partial, two
    // per CU was fastest of 1-12 at 60M rows.  With the sliced merge, three per CU take the 60M
    // query from 0.409-0.414 to 0.399-0.400 ms end to end (the kernel itself is no faster), and
    // four are slower (0.420); at 6M rows two stay best.  So three from 2^25 rows, two below.
    const int per_cu = n >= (1ll << 25) ? 3 : 2;
    int G = (int)ceil_div(n, 256 * FGA_ROWS * 8);
    if(G > num_cus() * per_cu) G = num_cus() * per_cu;
    if(G < 1) G = 1;
    const void *pk0 = nullptr, *pk1 = nullptr;
    if(n > 0){ if((e = dev_view(c, keys[0], &pk0))) R e; if(nk == 2 && (e = dev_view(c, keys[1], &pk1))) R e; }
    // FGA_ROWS-row vector loads of every column
    if(((uintptr_t)pk0 | (uintptr_t)pk1) % FGA_ROWS){ arena_set(c, mark); R HPA_E_NOT_IMPL; }
    for(int k = 0; k < nx; k++)
        if(used[k] && (uintptr_t)cols.p[k] % (uintptr_t)(FGA_ROWS * type_width(typ[k]))){ arena_set(c, mark); R HPA_E_NOT_IMPL; }
    void *part = dalloc(c, (L)psize * FGA_NS * G);
    unsigned *usedb = (unsigned*)dalloc(c, 4 * G);
    int *ovf = (int*)dalloc(c, 16);
    const L GM = (L)G * FGA_NS;                              // groups cannot exceed the partial slots
    unsigned *codes = (unsigned*)dalloc(c, 4 * GM);
    int *gdev = ovf + 1;                                     // overflow flag and group count, one read
    unsigned long long *cnt = (unsigned long long*)dalloc(c, 8 * GM);
    long long *first = (long long*)dalloc(c, 8 * GM);
    double *fo = (double*)dalloc(c, 8 * GM * NA);
    long long *qo = (long long*)dalloc(c, 8 * GM * NA);
    // per-(group, merge block) records of the sliced merge
    const L NREC = (L)FGA_MCAP * FGA_MS;
    unsigned long long *rc = (unsigned long long*)dalloc(c, 8 * NREC);
    long long *rfs = (long long*)dalloc(c, 8 * NREC);
    double *rf = (double*)dalloc(c, 16 * NREC * NA);
    long long *rq = (long long*)dalloc(c, 8 * NREC * NA);
    if(!part || !usedb || !ovf || !codes || !gdev || !cnt || !first || !fo || !qo || !rc || !rfs || !rf || !rq){ arena_set(c, mark); R HPA_E_MEMORY; }
    struct { const void *c[EXPR_MAX_COLS]; const void *k0; const void *k1; long long n; void *part; unsigned *used;
             unsigned kx0, kx1; long long ki[JIT_MAX_CONST]; double kf[JIT_MAX_CONST]; } ga;
    memset(&ga, 0, sizeof(ga));
    for(int k = 0; k < nx; k++) ga.c[k] = cols.p[k];
    memcpy(ga.ki, kc.i, sizeof(ga.ki)); memcpy(ga.kf, kc.f, sizeof(ga.kf));
    ga.k0 = pk0; ga.k1 = pk1; ga.n = n; ga.part = part; ga.used = usedb;
    // key bytes compared as their type orders them: C and J signed, B unsigned
    ga.kx0 = (keys[0]->typ == HPA_B) ? 0u : 0x80u;
    ga.kx1 = (nk == 2 && keys[1]->typ != HPA_B) ? 0x80u : 0u;
    o.kx0 = ga.kx0; o.kx1 = ga.kx1;
    double cb = nk; for(int k = 0; k < nx; k++) if(used[k]) cb += (double)type_width(typ[k]);
    // slot counts tried in turn: fewer slots keep the accumulators in fewer registers; a block
    // that meets more distinct codes than slots reports overflow and the next size reruns
    int hv[2] = {0, 0};
    static_assert((4 << (FGA_NSV - 1)) >= FGA_NS, "a kernel cache entry for every slot count");
    const CExpr *bkp = prep.has_bk ? &prep.bk : nullptr;
    // a bucketed request meets up to 2^bbits times the key codes: start at the largest slot count
    // rather than rerun the pass after an overflow
    const int ns0 = bkp ? FGA_NS : 4, nsi0 = bkp ? FGA_NSV - 1 : 0;
    for(int ns = ns0, nsi = nsi0; ns <= FGA_NS; ns *= 2, nsi++){
        hipFunction_t f1 = prep.fn[nsi][0], f2 = prep.fn[nsi][1], f3 = prep.fn[nsi][2];
        if(!f1 || !f2 || !f3){
            std::string key = "fga|" + std::to_string(nk) + "|" + std::to_string(ns) + "|" + std::to_string(nx) + "|" + types_key(nx, typ) + "|" + ce.s;
            for(int k = 0; k < nx; k++) key += used[k] ? '1' : '0';
            for(auto &ac : acc) key += "|" + ac.body + "#" + std::to_string(ac.cls) + (ac.f ? "f" : "i");
            if(bkp) key += "|bk" + std::to_string(bbits) + ":" + bkp->s;
            auto gen = [&](){ return fga_source(nk, ns, nx, typ, ce, used, acc, bkp, bbits); };
            if((e = jit_get_keyed(key, "fga", gen, &f1)) || (e = jit_get_keyed(key, "fga_merge", gen, &f2)) ||
               (e = jit_get_keyed(key, "fga_merge2", gen, &f3))) R e;
            std::lock_guard<std::mutex> lk(g_jit_mu);
            stored->fn[nsi][0] = f1; stored->fn[nsi][1] = f2; stored->fn[nsi][2] = f3;
        }
        {
            ProfScope ps("fused_group_agg", c->stream, (double)n * cb);
            if((e = jit_launch(c, f1, (unsigned)G, 256, ga))) R e;
        }
        struct { const void *part; const unsigned *used; int G; unsigned *codes; int *g; unsigned long long *cnt; long long *first; double *f; long long *q;
                 unsigned long long *rc; long long *rfs; double *rf; long long *rq; int *ovf; } ma;
        ma.part = part; ma.used = usedb; ma.G = G; ma.codes = codes; ma.g = gdev; ma.cnt = cnt; ma.first = first; ma.f = fo; ma.q = qo;
        ma.rc = rc; ma.rfs = rfs; ma.rf = rf; ma.rq = rq; ma.ovf = ovf;
        if((e = jit_launch(c, f2, FGA_MS, 1024, ma))) R e;
        // the typed outputs are sized for every possible group, so they are written right behind
        // the merge and one read brings back the overflow flag and the group count
        if(!o.key0){
            o.key0 = (uint8_t*)dalloc(c, GM); o.key1 = (uint8_t*)dalloc(c, GM);
            o.first = (int64_t*)dalloc(c, 8 * GM);
            o.bucket = bkp ? (uint8_t*)dalloc(c, GM) : nullptr;
            if(!o.key0 || !o.key1 || !o.first || (bkp && !o.bucket)){ arena_set(c, mark); R HPA_E_MEMORY; }
            for(int a = 0; a < na; a++){ o.out[a] = dalloc(c, 8 * GM); if(!o.out[a]){ arena_set(c, mark); R HPA_E_MEMORY; } }
        }
        struct { decltype(ma) m; FgaOut o; } m2;
        m2.m = ma; m2.o = o;
        if((e = jit_launch(c, f3, (unsigned)grid_for(GM, 64, 1), 64, m2))) R e;
        if(hd){
            // the flag, the group count and the first cap groups of every output in one read
            const L cap = hd->cap < GM ? hd->cap : GM;
            const void *src[4 + 2 * 16]; void *dst[4 + 2 * 16]; int64_t nb[4 + 2 * 16]; int m = 0;
            src[m] = ovf; dst[m] = hv; nb[m++] = 8;
            src[m] = o.key0; dst[m] = hd->keys[0]; nb[m++] = cap;
            if(nk == 2){ src[m] = o.key1; dst[m] = hd->keys[1]; nb[m++] = cap; }
            if(hd->first){ src[m] = o.first; dst[m] = hd->first; nb[m++] = 8 * cap; }
            if(hd->bucket && o.bucket){ src[m] = o.bucket; dst[m] = hd->bucket; nb[m++] = cap; }
            for(int a = 0; a < na; a++){ src[m] = o.out[a]; dst[m] = hd->aggs[a]; nb[m++] = 8 * cap; }
            if((e = read_staged(c, m, src, dst, nb))) R e;
        } else if((e = d2h(c, hv, ovf, 8))) R e;
        if(!hv[0]) break;
    }
    if(hv[0]){ arena_set(c, mark); R HPA_E_NOT_IMPL; }
    const L g = hv[1];
    const bool staged = hd && g <= (hd->cap < GM ? hd->cap : GM);
    uint8_t *k0 = o.key0, *k1 = o.key1;
    int64_t *zf = o.first;
    set_dev_vec(c, zkeys[0], keys[0]->typ, g, k0, staged ? hd->keys[0] : nullptr);
    if(nk == 2) set_dev_vec(c, zkeys[1], keys[1]->typ, g, k1, staged ? hd->keys[1] : nullptr);
    if(zfirst){ set_dev_vec(c, zfirst, HPA_L, g, zf, staged ? hd->first : nullptr); note_bound(c, zf, n); }
    for(int a = 0; a < na; a++){
        int kd = aggs[a].kind;
        int32_t t = kd == 4 ? HPA_L : kd == 1 ? HPA_E : (o.accf[a] ? HPA_E : HPA_L);
        set_dev_vec(c, zaggs[a], t, g, o.out[a], staged ? hd->aggs[a] : nullptr);
    }
    if(hd) *hd->ng = g;
    R 0;
}
extern "C" int32_t hpaFusedGroupAgg(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                                    int32_t na, const hpa_agg *aggs, hpa_V zfirst, hpa_V *zkeys, hpa_V *zaggs){
    R fga_run(nk, keys, nx, x, cond, na, aggs, zfirst, zkeys, zaggs, nullptr);
}
extern "C" int32_t hpaFusedGroupAggHost(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                                        int32_t na, const hpa_agg *aggs, hpa_V zfirst, hpa_V *zkeys, hpa_V *zaggs,
                                        int64_t cap, void **hkeys, int64_t *hfirst, void **haggs, int64_t *ng){
    if(cap < 0 || !hkeys || !haggs || !ng || (nk == 2 && !hkeys[1])) R HPA_E_DOMAIN;
    FgaHost hd{cap, hkeys, zfirst ? hfirst : nullptr, haggs, ng};
    R fga_run(nk, keys, nx, x, cond, na, aggs, zfirst, zkeys, zaggs, &hd);
}

extern "C" int32_t hpaFusedGroupAggBuckets(int32_t nk, hpa_V *keys, int32_t nx, hpa_V *x, const hpa_expr *cond,
                                           const hpa_expr *bucket, int32_t bbits, int32_t na, const hpa_agg *aggs,
                                           hpa_V *zkeys, hpa_V *zaggs, int64_t cap, void **hkeys, uint8_t *hbucket,
                                           void **haggs, int64_t *ng){
    if(!bucket || bbits < 1 || bbits > 2 || cap < 0 || !hkeys || !haggs || !ng || !hbucket || (nk == 2 && !hkeys[1])) R HPA_E_DOMAIN;
    FgaHost hd{cap, hkeys, nullptr, haggs, ng, hbucket};
    R fga_run(nk, keys, nx, x, cond, na, aggs, nullptr, zkeys, zaggs, &hd, bucket, bbits);
}

// Diagnostic for hosts without a GPU: emit the generated sources of one fused group-aggregate,
// one FA reduction and one FE map for the given expressions and compile them with hipRTC.
extern "C" int32_t hpaJitCompileCheck(int32_t nx, const int32_t *coltyp, const hpa_expr *cond,
                                      int32_t na, const hpa_agg *aggs){
    if(nx < 0 || nx > EXPR_MAX_COLS || na < 1 || na > 16) R HPA_E_DOMAIN;
    int32_t typ[EXPR_MAX_COLS]; for(int k = 0; k < nx; k++) typ[k] = coltyp[k];
    std::vector<int> uc(nx, 0), ub(nx, 0);
    CExpr ce{COND_TRUE, false};
    Konst kc;
    int32_t e;
    if(cond && (e = expr_to_c(cond, nx, typ, &ce, uc, &kc))) R e;
    std::vector<Acc> acc;
    for(int a = 0; a < na; a++){
        if(aggs[a].kind == 4) continue;
        CExpr be; if((e = expr_to_c(aggs[a].body, nx, typ, &be, ub, &kc))) R e;
        acc.push_back({be.s, be.f, aggs[a].kind <= 1 ? 0 : aggs[a].kind});
    }
    std::vector<int> used(nx, 0);
    for(int k = 0; k < nx; k++) used[k] = uc[k] | ub[k];
    std::string src = fga_source(2, 4, nx, typ, ce, used, acc);
    g_compile_only = true;
    hipFunction_t f;
    e = jit_get(src, "fga", &f);
    g_compile_only = false;
    R e;
}
