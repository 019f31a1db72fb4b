// jit.hpp -- run-time specialised kernels for component-wise programs
// (hipRTC; see jit.hip).
#pragma once
#include "cwise_core.hpp"

#include <string>
#include <vector>

namespace cm {
namespace jit {

// CUMAT_JIT / cm_jit_mode: 0 never, 1 for evaluations of at least
// AUTO_MIN_ELEMENTS elements (smaller ones are launch-bound and the
// interpreter costs nothing there), 2 always.
enum { MODE_OFF = 0, MODE_AUTO = 1, MODE_ALWAYS = 2 };
constexpr long long AUTO_MIN_ELEMENTS = 1ll << 18;

struct Kernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

int mode();
int set_mode(int m);
// whether an evaluation of n elements should use a generated kernel
bool want(long long n);

// Straight-line statements of the program for one element: they read the
// leaf values through the expressions leafvals[l] (register-typed, raw
// bits) and leave the result in 'acc'; the caller declares R acc, t and the
// stack registers s0..s3. false: an instruction the generator does not
// cover.
bool program_body(const CwProgram& P, const char* const* leafvals, std::string& out);

// hipRTC compilation of 'body' (after the prelude of constants and the
// cwise_ops.hpp text) for 'arch'; 0 on success, 'log' holds the compiler's
// messages.
int compile(const std::string& body, const std::string& arch, std::vector<char>& code, std::string& log);
// Compiled kernel 'name' of 'body' (cached by key on the current device);
// 0 on success.
int get(const std::string& key, const std::string& body, const char* name, Kernel* out);
// Front cache by a byte signature that determines the generated source
// (find_sig: 0 and the kernel on a hit; put_sig after a successful get).
int find_sig(const std::string& sig, Kernel* out);
void put_sig(const std::string& sig, const Kernel& k);
// Append the bytes of a value to a signature.
template <typename V>
inline void sig_add(std::string& sig, const V& v)
{
    sig.append(reinterpret_cast<const char*>(&v), sizeof(V));
}
// Launch (grid x grid_y blocks) with the kernel arguments packed in one
// buffer; 0 on success.
int launch(const Kernel& k, unsigned grid, unsigned block, hipStream_t stream, const void* args, size_t args_bytes,
           unsigned grid_y = 1);

}  // namespace jit
}  // namespace cm
