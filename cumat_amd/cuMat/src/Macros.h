#ifndef CUMAT_AMD_MACROS_H
#define CUMAT_AMD_MACROS_H

// Constants, error types and argument checks of the C++ host layer.
// Names and values follow cuMat's public constants so user code compiles
// unchanged: Dynamic / Flags (cuMat/src/Constants.h:13-31), Axis
// (cuMat/src/Constants.h:94-101), AssignmentMode (cuMat/src/Constants.h:153-163),
// SparseFlags (cuMat/src/Constants.h:168-183). Errors surface as the
// exception kinds of cuMat/src/Macros.h:158-192 and cuMat/src/Errors.h:17-30.

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include <cumat_amd.h>

// Eigen interop (Matrix::toEigen/fromEigen, CSR from Eigen::SparseMatrix):
// off unless the user defines CUMAT_EIGEN_SUPPORT=1 before the first
// include (cuMat/src/Macros.h:65-71)
#ifndef CUMAT_EIGEN_SUPPORT
#define CUMAT_EIGEN_SUPPORT 0
#endif

#define CUMAT_NAMESPACE_BEGIN namespace cuMat {
#define CUMAT_NAMESPACE_END }
#define CUMAT_AMD_STR_(x) #x
#define CUMAT_AMD_STR(x) CUMAT_AMD_STR_(x)

#define CUMAT_ASSERT_ARGUMENT(x) \
    if (!(x)) throw std::invalid_argument(__FILE__ ":" CUMAT_AMD_STR(__LINE__) ": Invalid argument: " #x);
#define CUMAT_ASSERT_DIMENSION(x) \
    if (!(x)) throw std::invalid_argument(__FILE__ ":" CUMAT_AMD_STR(__LINE__) ": Invalid dimensions: " #x);
#define CUMAT_ASSERT_BOUNDS(x) \
    if (!(x)) throw std::out_of_range(__FILE__ ":" CUMAT_AMD_STR(__LINE__) ": Out of bounds: " #x);
#define CUMAT_ASSERT_ERROR(x) \
    if (!(x)) throw std::runtime_error(__FILE__ ":" CUMAT_AMD_STR(__LINE__) ": Runtime Error: " #x);

CUMAT_NAMESPACE_BEGIN

typedef std::ptrdiff_t Index;

const int Dynamic = -1;

enum Flags {
    ColumnMajor = 0x00,
    RowMajor = 0x01
};

enum Axis {
    NoAxis = 0,
    Row = 1,
    Column = 2,
    Batch = 4,
    All = Row | Column | Batch
};

enum class AssignmentMode {
    ASSIGN, ADD, SUB, MUL, DIV, MOD, AND, OR
};

enum SparseFlags {
    CSR = 2,
    ELLPACK = 3
};

// A HIP runtime failure (cuMat throws cuda_error for its CUDA failures; the
// alias keeps `catch (cuMat::cuda_error&)` in user code meaningful).
class hip_error : public std::exception {
    std::string message_;

public:
    explicit hip_error(std::string message) : message_(std::move(message)) {}
    const char* what() const noexcept override { return message_.c_str(); }
};
typedef hip_error cuda_error;

namespace internal {

// Turns a C ABI status into the exception cuMat would have thrown.
inline void check(int status, const char* call)
{
    if (status == CM_OK) return;
    std::string msg = std::string(call) + ": " + cm_last_error();
    switch (status) {
    case CM_E_ARG:
    case CM_E_DIM:
        throw std::invalid_argument(msg);
    case CM_E_NOMEM:
    case CM_E_UNSUPPORTED:
    case CM_E_STATE:
        throw std::runtime_error(msg);
    default:
        throw hip_error(msg);
    }
}

#define CUMAT_AMD_CALL(expr) ::cuMat::internal::check((expr), #expr)

template <typename T> struct DType;
template <> struct DType<int> { static constexpr int value = CM_I32; };
template <> struct DType<float> { static constexpr int value = CM_F32; };
template <> struct DType<double> { static constexpr int value = CM_F64; };
template <> struct DType<bool> { static constexpr int value = CM_BOOL; };

template <typename T> struct IsFloating { static constexpr bool value = false; };
template <> struct IsFloating<float> { static constexpr bool value = true; };
template <> struct IsFloating<double> { static constexpr bool value = true; };

}  // namespace internal

CUMAT_NAMESPACE_END

#endif
