/*
 * hs_codegen.hip -- kernels for networks with code models, compiled at run time.
 *
 * The reference turns every NeuronGroup's equations into generated CUDA
 * (cuda_generator.py:200-808 translating the state updater's abstract code,
 * stateupdate.cu / threshold.cu / reset.cu as the kernel templates) and
 * compiles the whole project with nvcc.  Here the engine's own kernels are the
 * templates: a network's code models become HsCg<R, c> specialisations (the
 * update statements, the threshold expression and the reset statements over
 * the model's named variables), and the engine source -- embedded in the
 * library at build time -- is compiled once with hipRTC for the device with
 * those specialisations in scope.  The update, window, fused and finalize
 * kernels of the result then replace the library's built-in ones for that
 * network; delivery, noise, plasticity and monitors are unchanged.
 *
 * The numerical contract of the built-in models holds: no contraction, IEEE
 * division and square root, the library's deterministic exp and log, one
 * Philox normal draw per neuron-step for noisy models.  So a code model gives
 * the same bits as the same statements compiled for the host against the same
 * math (the oracle's code models).
 */
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hs.h"
#include "hs_codegen.h"

#define HS_CG_TRY(expr)               \
    do {                              \
        const int _r = (expr);        \
        if (_r != HS_OK) return _r;   \
    } while (0)

/* hs_rtc_names[k], hs_rtc_texts[k] (k < hs_rtc_count): the engine's sources by
   the names its #include lines use; written by the library's build script */
#include "hs_rtc_sources.inc"

static bool hs_cg_ident(const std::string &s)
{
    if (s.empty() || s.size() > 64) return false;
    if (!(isalpha((unsigned char)s[0]) || s[0] == '_')) return false;
    for (char c : s)
        if (!(isalnum((unsigned char)c) || c == '_')) return false;
    return true;
}

/* code may name no identifier of the generated program's own (hs_..., __...)
   other than the allowed[0..n) it is documented to call */
static int hs_cg_reserved_ids(const std::string &e, const char *const *allowed, int n, const char *what)
{
    for (size_t k = 0; k < e.size();) {
        if (isalpha((unsigned char)e[k]) || e[k] == '_') {
            size_t q = k;
            while (q < e.size() && (isalnum((unsigned char)e[q]) || e[q] == '_')) ++q;
            const std::string id = e.substr(k, q - k);
            const bool suffix = k > 0 && (isalnum((unsigned char)e[k - 1]) || e[k - 1] == '.'); /* 1e5, 2LL */
            if (!suffix && (id.compare(0, 3, "hs_") == 0 || id.compare(0, 2, "__") == 0)) {
                bool ok = false;
                for (int a = 0; a < n; ++a) ok = ok || id == allowed[a];
                if (!ok) return hs_rt_fail(HS_ERR_ARG, "%s: name '%s' is reserved", what, id.c_str());
            }
            k = q;
        } else {
            ++k;
        }
    }
    return HS_OK;
}

int hs_cg_check(const HsCgModel &m)
{
    static const char *const reserved[] = {"R", "dt", "t", "i", "xi", "not_refractory", "exp", "log", "log10",
                                           "sqrt", "abs", "fabs", "clip", "floor", "ceil", "sign", "int", "sin",
                                           "cos", "tan", "sinh", "cosh", "tanh", "arcsin", "arccos", "arctan",
                                           "expm1", "log1p", "exprel"};
    if (m.names.empty() || (int)m.names.size() > HS_MAX_VARS)
        return hs_rt_fail(HS_ERR_ARG, "a code model has 1 to %d variables", HS_MAX_VARS);
    for (size_t a = 0; a < m.names.size(); ++a) {
        const std::string &n = m.names[a];
        if (!hs_cg_ident(n) || n.compare(0, 3, "hs_") == 0 || n.compare(0, 2, "__") == 0)
            return hs_rt_fail(HS_ERR_ARG, "code model variable name '%s' is not an identifier of its own",
                              n.c_str());
        for (const char *r : reserved)
            if (n == r) return hs_rt_fail(HS_ERR_ARG, "code model variable name '%s' is reserved", n.c_str());
        for (size_t b = 0; b < a; ++b)
            if (m.names[b] == n) return hs_rt_fail(HS_ERR_ARG, "code model variable '%s' twice", n.c_str());
    }
    for (const std::string *c : {&m.update, &m.threshold, &m.reset}) {
        if (c->find('#') != std::string::npos)
            return hs_rt_fail(HS_ERR_ARG, "code model code holds a preprocessor directive");
        int depth = 0;
        for (char ch : *c) {
            depth += ch == '{' ? 1 : (ch == '}' ? -1 : 0);
            if (depth < 0) break;
        }
        if (depth != 0) return hs_rt_fail(HS_ERR_ARG, "code model code has unbalanced braces");
        static const char *const allowed[] = {"hs_rand_u", "hs_rand_n", "hs_rand_xi"};
        HS_CG_TRY(hs_cg_reserved_ids(*c, allowed, 3, "code model code"));
    }
    return HS_OK;
}

/* the names in scope of one synaptic code function, and its statements */
static int hs_cg_check_scope(const std::vector<std::string> &all, const std::string &code)
{
    static const char *const reserved[] = {"R", "dt", "t", "i", "j", "lastupdate", "exp", "log", "log10", "sqrt",
                                           "abs", "fabs", "clip", "floor", "ceil", "sign", "int", "sin", "cos",
                                           "tan", "sinh", "cosh", "tanh", "arcsin", "arccos", "arctan", "expm1",
                                           "log1p", "exprel"};
    for (size_t a = 0; a < all.size(); ++a) {
        const std::string &n = all[a];
        if (!hs_cg_ident(n) || n.compare(0, 3, "hs_") == 0 || n.compare(0, 2, "__") == 0)
            return hs_rt_fail(HS_ERR_ARG, "synaptic code name '%s' is not an identifier of its own", n.c_str());
        for (const char *r : reserved)
            if (n == r) return hs_rt_fail(HS_ERR_ARG, "synaptic code name '%s' is reserved", n.c_str());
        for (size_t b = 0; b < a; ++b)
            if (all[b] == n) return hs_rt_fail(HS_ERR_ARG, "synaptic code name '%s' twice", n.c_str());
    }
    if (code.find('#') != std::string::npos)
        return hs_rt_fail(HS_ERR_ARG, "synaptic code holds a preprocessor directive");
    int depth = 0;
    for (char ch : code) {
        depth += ch == '{' ? 1 : (ch == '}' ? -1 : 0);
        if (depth < 0) break;
    }
    if (depth != 0) return hs_rt_fail(HS_ERR_ARG, "synaptic code has unbalanced braces");
    static const char *const allowed[] = {"hs_srand_u", "hs_srand_n"};
    return hs_cg_reserved_ids(code, allowed, 2, "synaptic code");
}

int hs_cg_check_syn(const HsSynCgModel &m)
{
    if ((int)m.syn_names.size() > HS_MAX_SYN_VARS || (int)m.post_names.size() > HS_MAX_VARS ||
        (int)m.pre_names.size() > HS_MAX_VARS || (int)m.post_read_names.size() > HS_MAX_VARS)
        return hs_rt_fail(HS_ERR_ARG, "synaptic code: at most %d synapse and %d target/source variables",
                          HS_MAX_SYN_VARS, HS_MAX_VARS);
    std::vector<std::string> pre_scope = m.syn_names, post_scope = m.syn_names;
    for (const std::string &n : m.pre_names) {
        pre_scope.push_back(n + "_pre");
        post_scope.push_back(n + "_pre");
    }
    for (const std::string &n : m.post_names) pre_scope.push_back(n + "_post");
    for (const std::string &n : m.post_read_names) post_scope.push_back(n + "_post");
    HS_CG_TRY(hs_cg_check_scope(pre_scope, m.on_pre));
    return hs_cg_check_scope(post_scope, m.on_post);
}

/* one precision's specialisation of synaptic code c: on_pre(sv, pre, eff, ...)
   and on_post(sv, pre, post, ...) over named locals */
static void hs_cg_syn_struct(std::string &o, const char *R, int c, const HsSynCgModel &m)
{
    const int ns = (int)m.syn_names.size(), npost = (int)m.post_names.size(), npre = (int)m.pre_names.size(),
              npr = (int)m.post_read_names.size();
    o += "template <> struct HsSynCg<" + std::string(R) + ", " + std::to_string(c) + "> {\n";
    o += "    static constexpr int NS = " + std::to_string(ns) + ", NPOST = " + std::to_string(npost) +
         ", NPRE = " + std::to_string(npre) + ", NPR = " + std::to_string(npr) +
         ", HAS_POST = " + std::string(m.on_post.empty() ? "0" : "1") + ";\n";
    o += "    typedef " + std::string(R) + " R;\n";
    auto head = [&]() {
        for (int v = 0; v < ns; ++v) o += "        R " + m.syn_names[v] + " = hs_sv[" + std::to_string(v) + "];\n";
        for (int v = 0; v < npre; ++v)
            o += "        const R " + m.pre_names[v] + "_pre = hs_pre[" + std::to_string(v) + "];\n";
    };
    auto tail = [&]() {
        for (int v = 0; v < ns; ++v) o += "        hs_sv[" + std::to_string(v) + "] = " + m.syn_names[v] + ";\n";
        for (int v = 0; v < npre; ++v) o += "        (void)" + m.pre_names[v] + "_pre;\n";
    };
    o += "    static __device__ __forceinline__ void on_pre(R *hs_sv, const R *hs_pre, R *hs_eff, R dt, double t, "
         "int i, int j, double lastupdate, unsigned long long hs_seed, unsigned int hs_path, long long hs_step)\n"
         "    {\n        (void)hs_sv; (void)hs_pre; (void)hs_eff; (void)dt; (void)t; (void)i; (void)j; "
         "(void)lastupdate; (void)hs_seed; (void)hs_path; (void)hs_step;\n";
    head();
    for (int v = 0; v < npost; ++v) o += "        R " + m.post_names[v] + "_post = (R)0;\n";
    o += "        {\n" + m.on_pre + "\n        }\n";
    tail();
    for (int v = 0; v < npost; ++v) o += "        hs_eff[" + std::to_string(v) + "] = " + m.post_names[v] + "_post;\n";
    o += "    }\n";
    o += "    static __device__ __forceinline__ void on_post(R *hs_sv, const R *hs_pre, const R *hs_post, R dt, "
         "double t, int i, int j, double lastupdate, unsigned long long hs_seed, unsigned int hs_path, "
         "long long hs_step)\n    {\n        (void)hs_sv; (void)hs_pre; (void)hs_post; (void)dt; (void)t; "
         "(void)i; (void)j; (void)lastupdate; (void)hs_seed; (void)hs_path; (void)hs_step;\n";
    head();
    for (int v = 0; v < npr; ++v)
        o += "        const R " + m.post_read_names[v] + "_post = hs_post[" + std::to_string(v) + "];\n";
    o += "        {\n" + m.on_post + "\n        }\n";
    tail();
    for (int v = 0; v < npr; ++v) o += "        (void)" + m.post_read_names[v] + "_post;\n";
    o += "    }\n};\n";
}

/* one precision's specialisation of model c */
static void hs_cg_struct(std::string &o, const char *R, int c, const HsCgModel &m)
{
    const int nv = (int)m.names.size();
    o += "template <> struct HsCg<" + std::string(R) + ", " + std::to_string(c) + "> {\n";
    o += "    static constexpr int NV = " + std::to_string(nv) + ", NOISY = " + std::to_string(m.noisy ? 1 : 0) +
         ";\n";
    o += "    typedef " + std::string(R) + " R;\n";
    auto load = [&]() {
        for (int v = 0; v < nv; ++v) o += "        R " + m.names[v] + " = hs_s[" + std::to_string(v) + "];\n";
    };
    auto store = [&]() {
        for (int v = 0; v < nv; ++v) o += "        hs_s[" + std::to_string(v) + "] = " + m.names[v] + ";\n";
    };
    o += "    static __device__ __forceinline__ bool step(R *hs_s, int not_refractory, R xi, R dt, double t, int i, "
         "unsigned long long hs_seed, unsigned int hs_stream, long long hs_step)\n";
    o += "    {\n        (void)not_refractory; (void)xi; (void)dt; (void)t; (void)i; (void)hs_seed; (void)hs_stream; "
         "(void)hs_step;\n";
    load();
    o += "        {\n" + m.update + "\n        }\n";
    store();
    for (int v = 0; v < nv; ++v) o += "        (void)" + m.names[v] + ";\n";
    o += "        return " + (m.threshold.empty() ? std::string("false") : "(bool)(" + m.threshold + ")") + ";\n";
    o += "    }\n";
    o += "    static __device__ __forceinline__ void reset(R *hs_s, R dt, double t, int i, unsigned long long hs_seed, "
         "unsigned int hs_stream, long long hs_step)\n";
    o += "    {\n        (void)dt; (void)t; (void)i; (void)hs_seed; (void)hs_stream; (void)hs_step;\n";
    load();
    o += "        {\n" + m.reset + "\n        }\n";
    store();
    o += "    }\n};\n";
}

/* the functions generated code may call: the library's deterministic exp and
   log, IEEE sqrt; overloads keep the argument's precision.  Included after
   the engine sources, whose own calls are not renamed by the macros */
static const char *const hs_cg_math = R"HSCG(
#include "hs_cgfuncs.h"
__device__ __forceinline__ double hs_cg_exp(double x) { return hs_exp(x); }
__device__ __forceinline__ float hs_cg_exp(float x) { return hs_expf(x); }
__device__ __forceinline__ double hs_cg_exp(int x) { return hs_exp((double)x); }
__device__ __forceinline__ double hs_cg_log(double x) { return hs_log(x); }
__device__ __forceinline__ float hs_cg_log(float x) { return hs_logf(x); }
__device__ __forceinline__ double hs_cg_log(int x) { return hs_log((double)x); }
__device__ __forceinline__ double hs_cg_log10(double x) { return hs_log(x) / 2.302585092994046; }
__device__ __forceinline__ float hs_cg_log10(float x) { return hs_logf(x) / 2.3025851f; }
__device__ __forceinline__ double hs_cg_sqrt(double x) { return hs_sqrt(x); }
__device__ __forceinline__ float hs_cg_sqrt(float x) { return hs_sqrtf(x); }
__device__ __forceinline__ double hs_cg_sqrt(int x) { return hs_sqrt((double)x); }
__device__ __forceinline__ double hs_cg_abs(double x) { return x < 0.0 ? -x : x; }
__device__ __forceinline__ float hs_cg_abs(float x) { return x < 0.0f ? -x : x; }
__device__ __forceinline__ int hs_cg_abs(int x) { return x < 0 ? -x : x; }
__device__ __forceinline__ long long hs_cg_abs(long long x) { return x < 0 ? -x : x; }
__device__ __forceinline__ double hs_cg_exp(long long x) { return hs_exp((double)x); }
__device__ __forceinline__ double hs_cg_log(long long x) { return hs_log((double)x); }
__device__ __forceinline__ double hs_cg_sqrt(long long x) { return hs_sqrt((double)x); }
__device__ __forceinline__ double hs_cg_floor(long long x) { return (double)x; }
__device__ __forceinline__ double hs_cg_ceil(long long x) { return (double)x; }
__device__ __forceinline__ double hs_cg_floor(double x) { return floor(x); }
__device__ __forceinline__ float hs_cg_floor(float x) { return floorf(x); }
__device__ __forceinline__ double hs_cg_ceil(double x) { return ceil(x); }
__device__ __forceinline__ float hs_cg_ceil(float x) { return ceilf(x); }
template <typename T> __device__ __forceinline__ T hs_cg_sign(T x) { return (T)((x > (T)0) - (x < (T)0)); }
template <typename A, typename B, typename C>
__device__ __forceinline__ auto hs_cg_clip(A x, B lo, C hi) -> decltype(x + lo + hi)
{
    typedef decltype(x + lo + hi) T;
    const T v = (T)x;
    return v < (T)lo ? (T)lo : (v > (T)hi ? (T)hi : v);
}
#define exp hs_cg_exp
#define log hs_cg_log
#define log10 hs_cg_log10
#define sqrt hs_cg_sqrt
#define abs hs_cg_abs
#define fabs hs_cg_abs
#define floor hs_cg_floor
#define ceil hs_cg_ceil
#define sign hs_cg_sign
#define clip hs_cg_clip
#define sin hs_cg_sin
#define cos hs_cg_cos
#define tan hs_cg_tan
#define sinh hs_cg_sinh
#define cosh hs_cg_cosh
#define tanh hs_cg_tanh
#define arcsin hs_cg_arcsin
#define arccos hs_cg_arccos
#define arctan hs_cg_arctan
#define expm1 hs_cg_expm1
#define log1p hs_cg_log1p
#define exprel hs_cg_exprel
/* rand() / randn() call site k of a code model: a pure function of the
   network's seed, the group's stream, the neuron and the step */
#define hs_rand_u(k) hs_rand(hs_seed, hs_cg_rnd_stream(hs_stream, k, 0), (uint32_t)i, hs_step)
#define hs_rand_n(k) hs_randn(hs_seed, hs_cg_rnd_stream(hs_stream, k, 1), (uint32_t)i, hs_step)
/* noise variable k >= 1 of an equation with several (xi_1, xi_2, ...; k = 0 is xi) */
#define hs_rand_xi(k) hs_randn(hs_seed, hs_cg_rnd_stream(hs_stream, k, 2), (uint32_t)i, hs_step)
/* rand() / randn() call site k of synaptic code: a pure function of the
   network's seed, the pathway, the synapse's (i, j) and the step */
#define hs_srand_u(k) hs_rand(hs_seed, hs_cg_syn_stream(hs_path, k, 0), (uint32_t)i, hs_cg_syn_ctr(hs_step, j))
#define hs_srand_n(k) hs_randn(hs_seed, hs_cg_syn_stream(hs_path, k, 1), (uint32_t)i, hs_cg_syn_ctr(hs_step, j))
/* rand() / randn() call site k of a connect condition or probability: a pure
   function of the connect call's seed and stream and the pair (i, j) relative to
   the ranges, so both connect passes see the same draw */
#define hs_crand_u(k) hs_rand(hs_seed, hs_cg_conn_stream(hs_cstream, k, 0), (uint32_t)i, hs_cg_conn_ctr(j, hs_ccopy))
#define hs_crand_n(k) hs_randn(hs_seed, hs_cg_conn_stream(hs_cstream, k, 1), (uint32_t)i, hs_cg_conn_ctr(j, hs_ccopy))
)HSCG";

std::string hs_cg_source(int prec, const std::vector<HsCgModel> &models, const std::vector<HsSynCgModel> &syn)
{
    (void)prec;
    std::string o;
    if (!models.empty()) o += "#define HS_CG_MODELS " + std::to_string(models.size()) + "\n";
    if (!syn.empty()) o += "#define HS_SCG_PATHS " + std::to_string(syn.size()) + "\n";
    o += "#include \"hs_engine.hip\"\n";
    o += "#include \"hs_syncode.h\"\n";
    o += hs_cg_math;
    for (size_t c = 0; c < models.size(); ++c) {
        hs_cg_struct(o, "double", (int)c, models[c]);
        hs_cg_struct(o, "float", (int)c, models[c]);
    }
    for (size_t c = 0; c < syn.size(); ++c) {
        hs_cg_syn_struct(o, "double", (int)c, syn[c]);
        hs_cg_syn_struct(o, "float", (int)c, syn[c]);
    }
    return o;
}

struct HsCgEntry {
    hipModule_t mod = nullptr;
    HsCgKernels k;
};

static std::mutex g_cg_mu;
static std::map<std::string, HsCgEntry> g_cg_cache; /* precision + program text -> loaded kernels (device of hs_init) */

/* compiles src (with the engine sources as its headers) for the current device
   and loads the kernels names[k] into *slots[k]; what names the program in
   errors */
static int hs_rtc_build(const std::string &src, const char *prog_name, const char *what,
                        const std::vector<std::string> &names, const std::vector<hipFunction_t *> &slots,
                        hipModule_t *mod)
{
    const int nk = (int)names.size();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), prog_name, hs_rtc_count, hs_rtc_texts, hs_rtc_names) !=
        HIPRTC_SUCCESS)
        return hs_rt_fail(HS_ERR_HIP, "hipRTC: program creation failed");
    for (const std::string &n : names) (void)hiprtcAddNameExpression(prog, n.c_str());
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hiprtcDestroyProgram(&prog);
        return hs_rt_fail(HS_ERR_HIP, "hipRTC: no device properties");
    }
    std::string arch = prop.gcnArchName;
    arch = "--offload-arch=" + arch.substr(0, arch.find(':'));
    /* the flags the library itself is built with: the numerical contract
       (no contraction, no fast math, no packed FP32) and hardware FP atomics */
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (cr != HIPRTC_SUCCESS) {
        size_t n = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
            log.resize(n);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        if (log.size() > 3000) log = "..." + log.substr(log.size() - 3000);
        return hs_rt_fail(HS_ERR_ARG, "%s do not compile: %s", what, log.c_str());
    }
    size_t code_n = 0;
    std::vector<char> code;
    if (hiprtcGetCodeSize(prog, &code_n) == HIPRTC_SUCCESS && code_n) {
        code.resize(code_n);
        (void)hiprtcGetCode(prog, code.data());
    }
    std::vector<std::string> lowered(nk);
    for (int k = 0; k < nk; ++k) {
        const char *ln = nullptr;
        if (hiprtcGetLoweredName(prog, names[k].c_str(), &ln) != HIPRTC_SUCCESS || !ln) {
            (void)hiprtcDestroyProgram(&prog);
            return hs_rt_fail(HS_ERR_HIP, "hipRTC: no symbol for %s", names[k].c_str());
        }
        lowered[k] = ln;
    }
    (void)hiprtcDestroyProgram(&prog);
    if (code.empty()) return hs_rt_fail(HS_ERR_HIP, "hipRTC: empty code object");
    hipError_t he = hipModuleLoadData(mod, code.data());
    if (he != hipSuccess) return hs_rt_fail(HS_ERR_HIP, "%s kernels: %s", what, hipGetErrorString(he));
    for (int k = 0; k < nk; ++k) {
        if ((he = hipModuleGetFunction(slots[k], *mod, lowered[k].c_str())) != hipSuccess) {
            (void)hipModuleUnload(*mod);
            *mod = nullptr;
            return hs_rt_fail(HS_ERR_HIP, "%s kernel %s: %s", what, names[k].c_str(), hipGetErrorString(he));
        }
    }
    return HS_OK;
}

int hs_cg_kernels(int prec, int ms, const std::vector<HsCgModel> &models, const std::vector<HsSynCgModel> &syn,
                  const HsCgKernels **out)
{
    if (ms != 0 && !(HS_IS_CODE_MODEL(ms) && ms - HS_MODEL_CODE < (int)models.size()))
        return hs_rt_fail(HS_ERR_ARG, "bad kernel specialisation %d", ms);
    if ((int)models.size() > HS_MAX_CODE_MODELS || (int)syn.size() > HS_MAX_CODE_MODELS ||
        (models.empty() && syn.empty()))
        return hs_rt_fail(HS_ERR_UNSUPPORTED, "1 to %d code models and synaptic codes per network",
                          HS_MAX_CODE_MODELS);
    const std::string src = hs_cg_source(prec, models, syn);
    /* the kernels are instantiated for one precision and one specialisation */
    const std::string key = std::to_string(prec) + " " + std::to_string(ms) + "\n" + src;
    std::lock_guard<std::mutex> lock(g_cg_mu);
    auto it = g_cg_cache.find(key);
    if (it != g_cg_cache.end()) {
        *out = &it->second.k;
        return HS_OK;
    }
    const char *R = prec == HS_F64 ? "double" : "float";
    HsCgEntry e;
    std::vector<std::string> names;
    std::vector<hipFunction_t *> slots;
    if (!models.empty()) {
        const std::string m = std::to_string(ms);
        names = {std::string("hs_k_update<") + R + ", false, " + m + ">",
                 std::string("hs_k_update<") + R + ", true, " + m + ">",
                 std::string("hs_k_window<") + R + ", " + m + ">",
                 std::string("hs_k_fused<") + R + ", " + m + ">",
                 std::string("hs_k_finalize<") + R + ", false>",
                 std::string("hs_k_finalize<") + R + ", true>"};
        slots = {&e.k.update[0], &e.k.update[1], &e.k.window, &e.k.fused, &e.k.finalize[0], &e.k.finalize[1]};
        if (prec != HS_F64 && ms != 0) {
            names.push_back(std::string("hs_k_window_o8<") + R + ", " + m + ">");
            slots.push_back(&e.k.window_o8);
        }
    }
    if (!syn.empty()) {
        names.push_back(std::string("hs_k_syn_code<") + R + ">");
        slots.push_back(&e.k.syn_code);
        names.push_back(std::string("hs_k_syn_post<") + R + ">");
        slots.push_back(&e.k.syn_post);
    }
    HS_CG_TRY(hs_rtc_build(src, "hs_code_models.hip", "code models", names, slots, &e.mod));
    auto ins = g_cg_cache.emplace(key, e);
    *out = &ins.first->second.k;
    return HS_OK;
}

/* ------------------------------------------------------------------------ */
/* connect conditions and probabilities                                      */
/* ------------------------------------------------------------------------ */
int hs_cg_conn_check(const std::string &cond, const std::string &p, const std::string &w)
{
    for (const std::string *e : {&cond, &p, &w}) {
        if (e->size() > 16384) return hs_rt_fail(HS_ERR_ARG, "connect expression longer than 16 kB");
        if (e->find_first_of("#;{}") != std::string::npos)
            return hs_rt_fail(HS_ERR_ARG, "a connect expression holds '#', ';' or a brace: expressions only");
        int depth = 0;
        for (char ch : *e) {
            depth += ch == '(' ? 1 : (ch == ')' ? -1 : 0);
            if (depth < 0) break;
        }
        if (depth != 0) return hs_rt_fail(HS_ERR_ARG, "a connect expression has unbalanced parentheses");
        /* the generated function's own names (hs_a, hs_c, hs_ai, ...) are not the
           expression's: of the hs_ names only the documented functions */
        static const char *const allowed[] = {"hs_pre",      "hs_post",   "hs_cg_floordiv", "hs_cg_pymod",
                                              "hs_cg_pow",   "hs_crand_u", "hs_crand_n"};
        HS_CG_TRY(hs_cg_reserved_ids(*e, allowed, 7, "connect expression"));
    }
    return HS_OK;
}

/* the program of one condition / probability pair: the filter policy HsConnCg
   of hs_conndev.h and the two passes instantiated with it */
static std::string hs_cg_conn_source(const std::string &cond, const std::string &p, const std::string &w,
                                     bool copies)
{
    std::string o = "#include \"hs_conndev.h\"\n#include \"hs_cgconn.h\"\n";
    o += hs_cg_math;
    o += "#define hs_pre(k) ((double)((const R *)hs_a.pre[k])[hs_ai])\n";
    o += "#define hs_post(k) ((double)((const R *)hs_a.post[k])[hs_aj])\n";
    o += "struct HsConnCg {\n";
    o += "    static constexpr bool COND = " + std::string(cond.empty() ? "false" : "true") +
         ", PEXPR = " + std::string(p.empty() ? "false" : "true") + ", WEXPR = " +
         std::string(w.empty() ? "false" : "true") + ", COPIES = " + std::string(copies ? "true" : "false") + ";\n";
    /* the functions' common head (after their signatures h0 / h1) */
    const char *head =
        "        const long long i = hs_ai - hs_c.pre_lo, j = hs_aj - hs_c.post_lo;\n"
        "        const long long N_pre = (long long)hs_c.pre_hi - hs_c.pre_lo, N_post = (long long)hs_c.post_hi - "
        "hs_c.post_lo;\n"
        "        const unsigned long long hs_seed = hs_c.seed;\n"
        "        const unsigned int hs_cstream = hs_c.stream;\n"
        "        (void)hs_a; (void)i; (void)j; (void)N_pre; (void)N_post; (void)hs_seed; (void)hs_cstream; "
        "(void)hs_ccopy;\n";
    const std::string h0 = "(const HsConnSpec &hs_c, const HsConnVars &hs_a, long long hs_ai, long long hs_aj)\n"
                           "    {\n        const int hs_ccopy = 0;\n";
    const std::string h1 = "(const HsConnSpec &hs_c, const HsConnVars &hs_a, long long hs_ai, long long hs_aj, "
                           "int hs_ccopy)\n    {\n";
    o += "    template <typename R>\n    static __device__ __forceinline__ bool cond" + h0 + std::string(head);
    o += "        return (bool)(" + (cond.empty() ? std::string("true") : cond) + ");\n    }\n";
    o += "    template <typename R>\n    static __device__ __forceinline__ double p" + h0 + std::string(head);
    o += "        return (double)(" + (p.empty() ? std::string("0.0") : p) + ");\n    }\n";
    o += "    template <typename R>\n    static __device__ __forceinline__ double w" + h1 + std::string(head);
    o += "        return (double)(" + (w.empty() ? std::string("0.0") : w) + ");\n    }\n};\n";
    o += R"HSCG(
template <typename R>
__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE)
    hs_k_conn_count_cg(HsConnSpec c, HsConnVars a, int32_t *__restrict__ cnt, int *err)
{
    extern __shared__ int32_t tab[];
    hs_conn_count_body<HsConnCg, R>(c, a, cnt, err, tab);
}

template <typename R>
__global__ void __launch_bounds__(HS_CONN_WAVES * HS_WAVE)
    hs_k_conn_write_cg(HsConnSpec c, HsConnVars a, const int32_t *__restrict__ bstart, int32_t *__restrict__ post,
                       R *__restrict__ w, int *err)
{
    extern __shared__ int32_t cur[];
    hs_conn_write_body<HsConnCg, R>(c, a, bstart, post, w, err, cur);
}
)HSCG";
    return o;
}

struct HsCgConnEntry {
    hipModule_t mod = nullptr;
    HsCgConnKernels k;
};
static std::map<std::string, HsCgConnEntry> g_cg_conn_cache; /* precision + program text -> kernels */

int hs_cg_conn_kernels(int prec, const std::string &cond, const std::string &p, const std::string &w, bool copies,
                       const HsCgConnKernels **out)
{
    HS_CG_TRY(hs_cg_conn_check(cond, p, w));
    const std::string src = hs_cg_conn_source(cond, p, w, copies);
    const std::string key = std::to_string(prec) + "\n" + src;
    std::lock_guard<std::mutex> lock(g_cg_mu);
    auto it = g_cg_conn_cache.find(key);
    if (it != g_cg_conn_cache.end()) {
        *out = &it->second.k;
        return HS_OK;
    }
    const char *R = prec == HS_F64 ? "double" : "float";
    HsCgConnEntry e;
    const std::vector<std::string> names = {std::string("hs_k_conn_count_cg<") + R + ">",
                                            std::string("hs_k_conn_write_cg<") + R + ">"};
    const std::vector<hipFunction_t *> slots = {&e.k.count, &e.k.write};
    HS_CG_TRY(hs_rtc_build(src, "hs_connect_code.hip", "connect expressions", names, slots, &e.mod));
    auto ins = g_cg_conn_cache.emplace(key, e);
    *out = &ins.first->second.k;
    return HS_OK;
}

hipError_t hs_cg_launch_conn_count(const HsCgConnKernels *K, const HsConnSpec &c, const HsConnVars &a, int32_t *cnt,
                                   int *err, hipStream_t s)
{
    if (c.pre_hi <= c.pre_lo) return hipSuccess;
    if (c.D < 1 || c.D > HS_MAX_DELAY_SLOTS || c.fixed_k < 0 || c.fixed_k > HS_CONN_KMAX) return hipErrorInvalidValue;
    HsConnSpec cc = c;
    HsConnVars aa = a;
    void *args[] = {&cc, &aa, &cnt, &err};
    return hipModuleLaunchKernel(K->count, (unsigned)hs_conn_grid(c), 1, 1, (unsigned)(hs_conn_waves(c) * 64),
                                 1, 1, (unsigned)hs_conn_lds(c), s, args, nullptr);
}

hipError_t hs_cg_launch_conn_write(const HsCgConnKernels *K, const HsConnSpec &c, const HsConnVars &a,
                                   const int32_t *bstart, int32_t *post, void *w, int *err, hipStream_t s)
{
    if (c.pre_hi <= c.pre_lo) return hipSuccess;
    if (c.D < 1 || c.D > HS_MAX_DELAY_SLOTS || c.fixed_k < 0 || c.fixed_k > HS_CONN_KMAX) return hipErrorInvalidValue;
    HsConnSpec cc = c;
    HsConnVars aa = a;
    void *args[] = {&cc, &aa, &bstart, &post, &w, &err};
    return hipModuleLaunchKernel(K->write, (unsigned)hs_conn_grid(c), 1, 1, (unsigned)(hs_conn_waves(c) * 64),
                                 1, 1, (unsigned)hs_conn_lds(c), s, args, nullptr);
}

static hipError_t hs_cg_go(hipFunction_t f, int grid, int block, hipStream_t s, void **args)
{
    return hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, 0, s, args, nullptr);
}

hipError_t hs_cg_launch_update(const HsCgKernels *K, const void *groups, int ng, int total_tiles, HsCtl *ctl, int koff,
                               int w, int wmax, int fused, int block, int grid, int n_noise, hipStream_t s)
{
    if (w < 1 || w > HS_MAX_WINDOW || wmax < w || wmax > HS_MAX_WINDOW || block > 256 || grid < 1 || n_noise < 0)
        return hipErrorInvalidValue;
    void *args[] = {&groups, &ng, &total_tiles, &ctl, &koff, &w, &wmax, &grid};
    return hs_cg_go(K->update[fused ? 1 : 0], grid + n_noise, block, s, args);
}

hipError_t hs_cg_launch_window(const HsCgKernels *K, const void *groups, int ng, int total_tiles, const void *paths,
                               int np, HsCtl *ctl, int koff, int w, int W, int block, int n_upd, int n_far,
                               int n_noise, hipStream_t s)
{
    if (w < 1 || w > W || W > HS_MAX_WINDOW || block > 256 || block < 64 || block % 64 || n_upd < 1 || n_far < 0 ||
        n_noise < 0)
        return hipErrorInvalidValue;
    void *args[] = {&groups, &ng, &total_tiles, &paths, &np, &ctl, &koff, &w, &W, &n_upd, &n_far};
    /* as hs_launch_window: the 8-wave kernel when the update waves (64 lanes) exceed 7 per SIMD */
    const bool o8 = K->window_o8 && (long long)n_upd * (block / 64) > 7LL * 4 * hs_rt_cus();
    return hs_cg_go(o8 ? K->window_o8 : K->window, n_upd + n_far + n_noise, block, s, args);
}

hipError_t hs_cg_launch_fused(const HsCgKernels *K, const void *groups, int ng, int total_tiles, const void *paths,
                              int np, HsCtl *ctl, int koff, int block, int n_upd, int n_del, hipStream_t s)
{
    if (block > 256) return hipErrorInvalidValue;
    void *args[] = {&groups, &ng, &total_tiles, &paths, &np, &ctl, &koff, &n_upd};
    return hs_cg_go(K->fused, n_upd + n_del, block, s, args);
}

hipError_t hs_cg_launch_finalize(const HsCgKernels *K, const void *groups, int ng, int total_tiles, HsCtl *ctl,
                                 int fused, int block, hipStream_t s)
{
    int grid = total_tiles < 4096 ? total_tiles : 4096;
    if (grid < 1) grid = 1;
    void *args[] = {&groups, &ng, &total_tiles, &ctl};
    hipError_t e = hs_cg_go(K->finalize[fused ? 1 : 0], grid, block, s, args);
    if (e != hipSuccess) return e;
    return hs_launch_mark_reset(ctl, s);
}

hipError_t hs_cg_launch_syn_code(const HsCgKernels *K, const void *paths, int np, const void *groups, HsCtl *ctl,
                                 int koff, int fused, int grid, hipStream_t s)
{
    if (!K->syn_code || grid < 1) return hipErrorInvalidValue;
    void *args[] = {&paths, &np, &groups, &ctl, &koff, &fused};
    return hs_cg_go(K->syn_code, grid, 256, s, args);
}

hipError_t hs_cg_launch_syn_post(const HsCgKernels *K, const void *paths, int np, const void *groups, HsCtl *ctl,
                                 int koff, int grid, hipStream_t s)
{
    if (!K->syn_post || grid < 1) return hipErrorInvalidValue;
    void *args[] = {&paths, &np, &groups, &ctl, &koff};
    return hs_cg_go(K->syn_post, grid, 256, s, args);
}
