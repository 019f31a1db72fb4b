// jit.hip -- run-time specialised kernels for component-wise programs.
//
// cuMat compiles one CwiseEvaluationKernel per expression type in the
// caller's translation unit (cuMat/src/CwiseOp.h:66-85). The C ABI receives
// the expression as an accumulator program instead, which the interpreter of
// cwise_core.hpp evaluates with a wave-uniform switch per instruction. For
// large evaluations this file turns the program into straight-line HIP
// source, compiles it once with hipRTC for the device's architecture and
// caches the loaded kernel by the program's signature (its code and the
// kinds, dtypes and alignment of leaves and destination -- never pointers or
// sizes). The instruction semantics are the text of cwise_ops.hpp, the same
// header the interpreter is built from, so both paths compute the same bits.
//
// A program the generator does not cover, a failed compilation or a full
// cache leaves the call on the interpreter.
#include "jit.hpp"

#include <unordered_map>

#include <hip/hiprtc.h>

#include <cinttypes>
#include <cstdlib>
#include <map>
#include <memory>

namespace cm {
namespace jit {

static const char kOpsText[] =
#include "cwise_ops.inc"
    ;
static const char kFoldText[] =
#include "fold_ops.inc"
    ;

namespace {

std::mutex g_mu;
int g_mode = -1;  // resolved lazily from CUMAT_JIT
std::map<std::string, Kernel> g_cache;
constexpr size_t CACHE_MAX = 4096;

int mode_locked()
{
    if (g_mode < 0) {
        const char* s = getenv("CUMAT_JIT");
        g_mode = (s && *s) ? atoi(s) : MODE_AUTO;
        if (g_mode < MODE_OFF || g_mode > MODE_ALWAYS) g_mode = MODE_AUTO;
    }
    return g_mode;
}

void appendf(std::string& s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    s += buf;
}

// The constants cwise_ops.hpp and the generated code use, from the C ABI's
// own enumerations.
std::string prelude()
{
    std::string s;
    s += "typedef unsigned int uint32_t;\ntypedef unsigned long long uint64_t;\n";
    const struct {
        const char* name;
        int value;
    } k[] = {
        {"CM_F32", CM_F32}, {"CM_F64", CM_F64}, {"CM_I32", CM_I32}, {"CM_BOOL", CM_BOOL},
        {"CM_BIN_ADD", CM_BIN_ADD}, {"CM_BIN_SUB", CM_BIN_SUB}, {"CM_BIN_MUL", CM_BIN_MUL},
        {"CM_BIN_DIV", CM_BIN_DIV}, {"CM_BIN_MOD", CM_BIN_MOD}, {"CM_BIN_POW", CM_BIN_POW},
        {"CM_BIN_MIN", CM_BIN_MIN}, {"CM_BIN_MAX", CM_BIN_MAX}, {"CM_BIN_AND", CM_BIN_AND},
        {"CM_BIN_OR", CM_BIN_OR}, {"CM_BIN_XOR", CM_BIN_XOR}, {"CM_BIN_LAND", CM_BIN_LAND},
        {"CM_BIN_LOR", CM_BIN_LOR}, {"CM_BIN_EQ", CM_BIN_EQ}, {"CM_BIN_NE", CM_BIN_NE},
        {"CM_BIN_LT", CM_BIN_LT}, {"CM_BIN_GT", CM_BIN_GT}, {"CM_BIN_LE", CM_BIN_LE},
        {"CM_BIN_GE", CM_BIN_GE}, {"CM_UN_NEG", CM_UN_NEG}, {"CM_UN_ABS", CM_UN_ABS},
        {"CM_UN_ABS2", CM_UN_ABS2}, {"CM_UN_INVERSE", CM_UN_INVERSE},
        {"CM_UN_INVERSE_CHECK", CM_UN_INVERSE_CHECK}, {"CM_UN_SQRT", CM_UN_SQRT},
        {"CM_UN_RSQRT", CM_UN_RSQRT}, {"CM_UN_CBRT", CM_UN_CBRT}, {"CM_UN_EXP", CM_UN_EXP},
        {"CM_UN_LOG", CM_UN_LOG}, {"CM_UN_LOG1P", CM_UN_LOG1P}, {"CM_UN_LOG10", CM_UN_LOG10},
        {"CM_UN_SIN", CM_UN_SIN}, {"CM_UN_COS", CM_UN_COS}, {"CM_UN_TAN", CM_UN_TAN},
        {"CM_UN_ASIN", CM_UN_ASIN}, {"CM_UN_ACOS", CM_UN_ACOS}, {"CM_UN_ATAN", CM_UN_ATAN},
        {"CM_UN_SINH", CM_UN_SINH}, {"CM_UN_COSH", CM_UN_COSH}, {"CM_UN_TANH", CM_UN_TANH},
        {"CM_UN_FLOOR", CM_UN_FLOOR}, {"CM_UN_CEIL", CM_UN_CEIL}, {"CM_UN_ROUND", CM_UN_ROUND},
        {"CM_UN_NOT", CM_UN_NOT}, {"CM_UN_BITNOT", CM_UN_BITNOT}, {"CM_UN_SIGN", CM_UN_SIGN},
        {"CM_ASSIGN", CM_ASSIGN}, {"CM_ADD", CM_ADD}, {"CM_SUB", CM_SUB}, {"CM_MUL", CM_MUL},
        {"CM_DIV", CM_DIV}, {"CM_MOD", CM_MOD}, {"CM_AND", CM_AND}, {"CM_OR", CM_OR},
    };
    for (const auto& c : k) appendf(s, "constexpr int %s = %d;\n", c.name, c.value);
    s += kOpsText;
    s += "\n";
    s += kFoldText;
    s += "\nusing namespace cm::ops;\n";
    return s;
}

// dtype class of an instruction for the functor templates (I32 covers BOOL
// and any other integer code, as the interpreter's dispatch does)
const char* dt_class(int dt)
{
    return dt == CM_F32 ? "CM_F32" : (dt == CM_F64 ? "CM_F64" : "CM_I32");
}
const char* dt_name(int dt)
{
    switch (dt) {
    case CM_F32: return "CM_F32";
    case CM_F64: return "CM_F64";
    case CM_I32: return "CM_I32";
    case CM_BOOL: return "CM_BOOL";
    default: return nullptr;
    }
}

// Immediate of an instruction as a register value (the interpreter's imm()).
std::string imm_expr(const cm_instr& I)
{
    std::string s;
    if (I.dtype == CM_F32) {
        const float f = (float)I.fimm;
        uint32_t b;
        memcpy(&b, &f, 4);
        appendf(s, "mkf<R>(__uint_as_float(0x%08xu))", b);
    } else if (I.dtype == CM_F64) {
        uint64_t b;
        memcpy(&b, &I.fimm, 8);
        appendf(s, "mkd<R>(__longlong_as_double((long long)0x%016" PRIx64 "ull))", b);
    } else {
        appendf(s, "mki<R>((int)0x%08xu)", (uint32_t)(int)I.iimm);
    }
    return s;
}

// Conversion expression of register expression 'e' from 'from' to 'to'
// (the interpreter's convert(): equal codes are a no-op, an unknown source
// converts 0, an unknown target leaves the value).
std::string convert_expr(const std::string& e, int from, int to)
{
    if (from == to || !dt_name(to)) return e;
    const char* f = dt_name(from);
    std::string s = "convert1<";
    s += f ? f : "-1";
    s += ", ";
    s += dt_name(to);
    s += ">(" + e + ")";
    return s;
}

}  // namespace

int mode()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return mode_locked();
}

int set_mode(int m)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const int old = mode_locked();
    if (m >= MODE_OFF && m <= MODE_ALWAYS) g_mode = m;
    return old;
}

bool want(long long n)
{
    const int m = mode();
    return m == MODE_ALWAYS || (m == MODE_AUTO && n >= AUTO_MIN_ELEMENTS);
}

bool program_body(const CwProgram& P, const char* const* leafvals, std::string& out)
{
    out.clear();
    for (int pc = 0; pc < P.ncode; ++pc) {
        const cm_instr& I = P.code[pc];
        // source operand (the interpreter's fetch_src)
        std::string src;
        if (I.opcode == CM_OP_LOAD || I.opcode == CM_OP_BINARY) {
            if (I.src == CM_SRC_LEAF) {
                if (I.arg < 0 || I.arg >= P.nleaves) return false;
                src = convert_expr(leafvals[I.arg], P.leaves[I.arg].dtype, I.dtype);
            } else if (I.src == CM_SRC_IMM) {
                src = imm_expr(I);
            } else if (I.src == CM_SRC_STACK) {
                if (I.arg < 0 || I.arg >= CM_CWISE_MAX_STACK) return false;
                appendf(src, "s%d", I.arg);
            } else {
                return false;  // no source: left to the interpreter
            }
        }
        switch (I.opcode) {
        case CM_OP_LOAD:
            out += "acc = " + src + ";\n";
            break;
        case CM_OP_PUSH:
            if (I.arg < 0 || I.arg >= CM_CWISE_MAX_STACK) return false;
            appendf(out, "s%d = acc;\n", I.arg);
            break;
        case CM_OP_BINARY: {
            if (I.fn < 0 || I.fn >= CM_BIN_COUNT) return false;
            const bool rev = (I.flags & 1) != 0;
            out += "t = " + src + ";\n";
            appendf(out, "acc = binary1<%s, %d>(acc, %s, %s);\n", dt_class(I.dtype), I.fn, rev ? "t" : "acc",
                    rev ? "acc" : "t");
            break;
        }
        case CM_OP_UNARY:
            if (I.fn < 0 || I.fn >= CM_UN_COUNT) return false;
            appendf(out, "acc = unary1<%s, %d>(acc);\n", dt_class(I.dtype), I.fn);
            break;
        case CM_OP_CAST:
            out += "acc = " + convert_expr("acc", I.arg, I.dtype) + ";\n";
            break;
        case CM_OP_FMA: {
            if (I.arg < 0 || I.arg >= P.nleaves) return false;
            uint64_t fb;
            memcpy(&fb, &I.fimm, 8);
            appendf(out, "acc = fma1<%s>(acc, %s, __longlong_as_double((long long)0x%016" PRIx64
                         "ull), (long long)0x%016" PRIx64 "ull);\n",
                    dt_class(I.dtype), leafvals[I.arg], fb, (uint64_t)I.iimm);
            break;
        }
        case CM_OP_SELECT: {
            if (I.arg < 0 || I.arg >= CM_CWISE_MAX_STACK) return false;
            std::string b;
            if (I.src == CM_SRC_IMM) {
                b = imm_expr(I);
            } else {
                if (I.iimm < 0 || I.iimm >= P.nleaves) return false;
                b = leafvals[I.iimm];
            }
            appendf(out, "acc = ((uint32_t)acc != 0u) ? s%d : (%s);\n", I.arg, b.c_str());
            break;
        }
        default:
            return false;
        }
    }
    return true;
}

int compile(const std::string& body, const std::string& arch, std::vector<char>& code, std::string& log)
{
    const std::string src = prelude() + body;
    hiprtcProgram prog = nullptr;
    bool ok = hiprtcCreateProgram(&prog, src.c_str(), "cm_jit.hip", 0, nullptr, nullptr) == HIPRTC_SUCCESS;
    if (ok) {
        const std::string arch_opt = "--offload-arch=" + arch;
        // fp-contract=on: a multiply and an add of different instructions
        // stay separate, as in the interpreter; packed FP32 instructions are
        // kept out of the code as in the library build
        const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=on",
                              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
        ok = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts) == HIPRTC_SUCCESS;
        size_t ls = 0;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
            log.assign(ls, '\0');
            hiprtcGetProgramLog(prog, &log[0]);
        }
    }
    if (ok) {
        size_t cs = 0;
        ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
        if (ok) {
            code.resize(cs);
            ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
        }
    }
    if (prog) hiprtcDestroyProgram(&prog);
    if (!ok && getenv("CUMAT_JIT_VERBOSE"))
        fprintf(stderr, "cumat_amd jit: compilation failed:\n%s\n--- source ---\n%s\n", log.c_str(), src.c_str());
    return ok ? 0 : -1;
}

int get(const std::string& key, const std::string& body, const char* name, Kernel* out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    const std::string full_key = std::to_string(dev) + "|" + key;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(full_key);
    if (it != g_cache.end()) {
        *out = it->second;
        return out->fn ? 0 : -1;
    }
    if (g_cache.size() >= CACHE_MAX) return -1;
    Kernel k;
    hipDeviceProp_t prop;
    std::string arch = "gfx950";
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch.resize(colon);
    }
    std::vector<char> code;
    std::string log;
    bool ok = compile(body, arch, code, log) == 0;
    if (ok) ok = hipModuleLoadData(&k.mod, code.data()) == hipSuccess;
    if (ok) ok = hipModuleGetFunction(&k.fn, k.mod, name) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();  // a failed load must not surface in a later check
        k.fn = nullptr;
    } else {
        count(CM_CNT_JIT_COMPILE);
    }
    g_cache[full_key] = k;  // failures are cached too: the interpreter keeps the program
    *out = k;
    return ok ? 0 : -1;
}

// front cache: a byte signature of everything a generator reads (more
// specific than its key) -> the kernel, so a repeated evaluation skips
// building the source and the key
namespace {
std::unordered_map<std::string, Kernel> g_sig;
constexpr size_t SIG_MAX = 8192;
}  // namespace

int find_sig(const std::string& sig, Kernel* out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    std::string full = sig;
    full.append((const char*)&dev, sizeof(dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sig.find(full);
    if (it == g_sig.end()) return -1;
    *out = it->second;
    return 0;
}

void put_sig(const std::string& sig, const Kernel& k)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || !k.fn) return;
    std::string full = sig;
    full.append((const char*)&dev, sizeof(dev));
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_sig.size() >= SIG_MAX) g_sig.clear();  // the kernels stay in the main cache
    g_sig[full] = k;
}

int launch(const Kernel& k, unsigned grid, unsigned block, hipStream_t stream, const void* args, size_t args_bytes,
           unsigned grid_y)
{
    size_t sz = args_bytes;
    void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                   HIP_LAUNCH_PARAM_END};
    if (hipModuleLaunchKernel(k.fn, grid, grid_y, 1, block, 1, 1, 0, stream, nullptr, cfg) != hipSuccess) return -1;
    count(CM_CNT_JIT_LAUNCH);
    return 0;
}

}  // namespace jit
}  // namespace cm

extern "C" int cm_jit_mode(int mode)
{
    return cm::jit::set_mode(mode);
}
