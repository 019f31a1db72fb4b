#ifndef CUMAT_AMD_CWISEOP_H
#define CUMAT_AMD_CWISEOP_H

// Expression templates of the host layer. A tree of MatrixBase nodes
// (leaves = matrices or strided views of them, host scalars, unary / binary /
// casting ops) is lowered at assignment time to the accumulator program of
// cm_cwise_eval and evaluated by ONE kernel launch, the fusion that cuMat gets
// from instantiating CwiseEvaluationKernel over the tree
// (cuMat/src/CwiseOp.h:66-85,130-158). Operator set and typing follow
// cuMat/src/BinaryOps.h:69-293 (broadcasting of size-1 dimensions, no
// implicit casting), cuMat/src/UnaryOps.h:128-312 and HostScalar
// (cuMat/src/NullaryOps.h:78-108).

#include <cmath>
#include <type_traits>
#include <vector>

#include "Context.h"

CUMAT_NAMESPACE_BEGIN

template <typename _Scalar, int _Rows, int _Columns, int _Batches, int _Flags>
class Matrix;
template <typename _Left, typename _Right>
class ProductOp;

namespace internal {

enum NodeKind { KindExpr = 0, KindLeaf = 1, KindScalar = 2 };

struct LeafDesc {
    const void* ptr;
    int dtype;
    int kind;
    int64_t sr, sc, sb;
    Index rows, cols, batches;
};

// Collects distinct leaves and the instruction list of one assignment.
class ProgramBuilder {
public:
    std::vector<LeafDesc> leaves;
    std::vector<cm_instr> code;

    int leaf(const LeafDesc& d)
    {
        for (size_t i = 0; i < leaves.size(); ++i) {
            const LeafDesc& l = leaves[i];
            if (l.ptr == d.ptr && l.dtype == d.dtype && l.kind == d.kind && l.sr == d.sr && l.sc == d.sc &&
                l.sb == d.sb && l.rows == d.rows && l.cols == d.cols && l.batches == d.batches)
                return (int)i;
        }
        if ((int)leaves.size() >= CM_CWISE_MAX_LEAVES)
            throw std::invalid_argument("cuMat: expression has more distinct matrices than one kernel takes");
        leaves.push_back(d);
        return (int)leaves.size() - 1;
    }
    void push(const cm_instr& i)
    {
        if ((int)code.size() >= CM_CWISE_MAX_CODE)
            throw std::invalid_argument("cuMat: expression too long for one kernel");
        code.push_back(i);
    }
    static cm_instr instr(int opcode, int dtype)
    {
        cm_instr i{};
        i.opcode = opcode;
        i.dtype = dtype;
        return i;
    }
};

inline void checkDepth(int depth)
{
    if (depth >= CM_CWISE_MAX_STACK)
        throw std::invalid_argument("cuMat: expression nested too deeply for one kernel");
}

template <int A, int B> struct BroadcastDim {
    static constexpr int value = (A == 1) ? B : A;
};

// functor ids whose result type is bool
constexpr bool boolResult(int fn)
{
    return fn == CM_BIN_LAND || fn == CM_BIN_LOR || fn == CM_BIN_EQ || fn == CM_BIN_NE || fn == CM_BIN_LT ||
           fn == CM_BIN_GT || fn == CM_BIN_LE || fn == CM_BIN_GE;
}
constexpr bool intOnly(int fn)
{
    return fn == CM_BIN_MOD || fn == CM_BIN_AND || fn == CM_BIN_OR || fn == CM_BIN_XOR;
}

template <typename T> struct is_matrix : std::false_type {};
template <typename S, int R, int C, int B, int F> struct is_matrix<Matrix<S, R, C, B, F>> : std::true_type {};

}  // namespace internal

template <typename _Child, int _Fn> class UnaryOp;
template <typename _Left, typename _Right, int _Fn> class BinaryOp;

namespace internal {
template <typename T> struct is_axpy_term : std::false_type {};
template <typename L, typename R, int F>
struct is_axpy_term<BinaryOp<L, R, F>> : std::integral_constant<bool, BinaryOp<L, R, F>::IsAxpyTerm> {};
}  // namespace internal
template <typename _Child, typename _Target> class CastingOp;
template <typename _Scalar> class HostScalar;
template <int _SparseFlags> struct SparsityPattern;
template <typename _Child, int _SparseFlags> class SparseExpressionOp;

template <typename _Derived>
class MatrixBase {
public:
    const _Derived& derived() const { return *static_cast<const _Derived*>(this); }
    _Derived& derived() { return *static_cast<_Derived*>(this); }

    Index rows() const { return derived().rows(); }
    Index cols() const { return derived().cols(); }
    Index batches() const { return derived().batches(); }
    Index size() const { return rows() * cols() * batches(); }

    // named component-wise functors (cuMat/src/UnaryOps.h:128-225)
#define CUMAT_AMD_UNARY(name, fn) \
    UnaryOp<_Derived, fn> name() const { return UnaryOp<_Derived, fn>(derived()); }
    CUMAT_AMD_UNARY(cwiseNegate, CM_UN_NEG)
    CUMAT_AMD_UNARY(cwiseAbs, CM_UN_ABS)
    CUMAT_AMD_UNARY(cwiseAbs2, CM_UN_ABS2)
    CUMAT_AMD_UNARY(cwiseInverse, CM_UN_INVERSE)
    CUMAT_AMD_UNARY(cwiseInverseCheck, CM_UN_INVERSE_CHECK)
    CUMAT_AMD_UNARY(cwiseSqrt, CM_UN_SQRT)
    CUMAT_AMD_UNARY(cwiseRsqrt, CM_UN_RSQRT)
    CUMAT_AMD_UNARY(cwiseCbrt, CM_UN_CBRT)
    CUMAT_AMD_UNARY(cwiseExp, CM_UN_EXP)
    CUMAT_AMD_UNARY(cwiseLog, CM_UN_LOG)
    CUMAT_AMD_UNARY(cwiseLog1p, CM_UN_LOG1P)
    CUMAT_AMD_UNARY(cwiseLog10, CM_UN_LOG10)
    CUMAT_AMD_UNARY(cwiseSin, CM_UN_SIN)
    CUMAT_AMD_UNARY(cwiseCos, CM_UN_COS)
    CUMAT_AMD_UNARY(cwiseTan, CM_UN_TAN)
    CUMAT_AMD_UNARY(cwiseAsin, CM_UN_ASIN)
    CUMAT_AMD_UNARY(cwiseAcos, CM_UN_ACOS)
    CUMAT_AMD_UNARY(cwiseAtan, CM_UN_ATAN)
    CUMAT_AMD_UNARY(cwiseSinh, CM_UN_SINH)
    CUMAT_AMD_UNARY(cwiseCosh, CM_UN_COSH)
    CUMAT_AMD_UNARY(cwiseTanh, CM_UN_TANH)
    CUMAT_AMD_UNARY(cwiseFloor, CM_UN_FLOOR)
    CUMAT_AMD_UNARY(cwiseCeil, CM_UN_CEIL)
    CUMAT_AMD_UNARY(cwiseRound, CM_UN_ROUND)
    CUMAT_AMD_UNARY(cwiseLogicalNot, CM_UN_NOT)
    CUMAT_AMD_UNARY(cwiseBinaryNot, CM_UN_BITNOT)
    CUMAT_AMD_UNARY(cwiseSign, CM_UN_SIGN)
#undef CUMAT_AMD_UNARY

    // named binary functors (cuMat/src/BinaryOps.h:230-243, BinaryOpsPlugin.inl)
#define CUMAT_AMD_BINARY(name, fn) \
    template <typename _Other> \
    BinaryOp<_Derived, _Other, fn> name(const MatrixBase<_Other>& o) const \
    { \
        return BinaryOp<_Derived, _Other, fn>(derived(), o.derived()); \
    }
    CUMAT_AMD_BINARY(cwiseAdd, CM_BIN_ADD)
    CUMAT_AMD_BINARY(cwiseSub, CM_BIN_SUB)
    CUMAT_AMD_BINARY(cwiseMul, CM_BIN_MUL)
    CUMAT_AMD_BINARY(cwiseDot, CM_BIN_MUL)
    CUMAT_AMD_BINARY(cwiseDiv, CM_BIN_DIV)
    CUMAT_AMD_BINARY(cwiseMod, CM_BIN_MOD)
    CUMAT_AMD_BINARY(cwisePow, CM_BIN_POW)
    CUMAT_AMD_BINARY(cwiseMin, CM_BIN_MIN)
    CUMAT_AMD_BINARY(cwiseMax, CM_BIN_MAX)
    CUMAT_AMD_BINARY(cwiseBinaryAnd, CM_BIN_AND)
    CUMAT_AMD_BINARY(cwiseBinaryOr, CM_BIN_OR)
    CUMAT_AMD_BINARY(cwiseBinaryXor, CM_BIN_XOR)
    CUMAT_AMD_BINARY(cwiseLogicalAnd, CM_BIN_LAND)
    CUMAT_AMD_BINARY(cwiseLogicalOr, CM_BIN_LOR)
    CUMAT_AMD_BINARY(cwiseEqual, CM_BIN_EQ)
    CUMAT_AMD_BINARY(cwiseNotEqual, CM_BIN_NE)
#undef CUMAT_AMD_BINARY

    // explicit casting (cuMat/src/UnaryOps.h:248-312)
    template <typename _Target>
    CastingOp<_Derived, _Target> cast() const { return CastingOp<_Derived, _Target>(derived()); }

    // evaluation into a new matrix (column-major unless the expression is a
    // matrix, whose own type is returned)
    auto eval() const;

    // reductions (cuMat/src/ReductionOpsPlugin.inl:124-450); defined in ReductionOps.h
    template <int _Axis> auto sum() const;
    template <int _Axis> auto prod() const;
    template <int _Axis> auto minCoeff() const;
    template <int _Axis> auto maxCoeff() const;
    template <int _Axis> auto all() const;
    template <int _Axis> auto any() const;
    template <int _Axis> auto bitwiseAnd() const;
    template <int _Axis> auto bitwiseOr() const;
    auto sum() const;
    auto prod() const;
    auto minCoeff() const;
    auto maxCoeff() const;
    auto all() const;
    auto any() const;
    auto sum(int axis) const;
    auto prod(int axis) const;
    auto minCoeff(int axis) const;
    auto maxCoeff(int axis) const;
    template <typename _Other> auto dot(const MatrixBase<_Other>& rhs) const;
    auto squaredNorm() const;
    auto norm() const;

    // the expression read only at the stored entries of 'pattern', as the
    // matrix of a sparse product (sparseView,
    // cuMat/src/SparseExpressionOpPlugin.inl:54-57); defined in SparseMatrix.h
    template <int _SparseFlags>
    SparseExpressionOp<_Derived, _SparseFlags> sparseView(const SparsityPattern<_SparseFlags>& pattern) const;
};

// ------------------------------------------------------------------ scalar
template <typename _Scalar>
class HostScalar : public MatrixBase<HostScalar<_Scalar>> {
    _Scalar value_;

public:
    typedef _Scalar Scalar;
    enum { Rows = 1, Columns = 1, Batches = 1, Flags = ColumnMajor, Kind = internal::KindScalar };
    explicit HostScalar(const _Scalar& v) : value_(v) {}
    Index rows() const { return 1; }
    Index cols() const { return 1; }
    Index batches() const { return 1; }
    _Scalar value() const { return value_; }

    void srcOf(cm_instr& i, internal::ProgramBuilder&) const
    {
        i.src = CM_SRC_IMM;
        if (internal::IsFloating<_Scalar>::value) i.fimm = (double)value_;
        else i.iimm = (int64_t)value_;
    }
    void emit(internal::ProgramBuilder& b, int depth) const
    {
        internal::checkDepth(depth);
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_LOAD, internal::DType<_Scalar>::value);
        srcOf(i, b);
        b.push(i);
    }
};

// A constant of given size (cuMat's Matrix::Constant, cuMat/src/NullaryOps.h:78-108).
template <typename _Scalar>
class ConstantOp : public MatrixBase<ConstantOp<_Scalar>> {
    _Scalar value_;
    Index rows_, cols_, batches_;

public:
    typedef _Scalar Scalar;
    enum { Rows = Dynamic, Columns = Dynamic, Batches = Dynamic, Flags = ColumnMajor, Kind = internal::KindScalar };
    ConstantOp(Index r, Index c, Index b, const _Scalar& v) : value_(v), rows_(r), cols_(c), batches_(b) {}
    Index rows() const { return rows_; }
    Index cols() const { return cols_; }
    Index batches() const { return batches_; }
    _Scalar value() const { return value_; }
    void srcOf(cm_instr& i, internal::ProgramBuilder& b) const { HostScalar<_Scalar>(value_).srcOf(i, b); }
    void emit(internal::ProgramBuilder& b, int depth) const { HostScalar<_Scalar>(value_).emit(b, depth); }
};

// ------------------------------------------------------------------ strided view leaf
// Transposed, diagonal and as-diagonal views of a matrix (TransposeOp in cwise
// context, ExtractDiagonalOp / AsDiagonalOp of cuMat/src/UnaryOps.h:361-487).
template <typename _Scalar>
class StridedView : public MatrixBase<StridedView<_Scalar>> {
    DevicePointer<_Scalar> owner_;
    internal::LeafDesc d_;

public:
    typedef _Scalar Scalar;
    enum { Rows = Dynamic, Columns = Dynamic, Batches = Dynamic, Flags = ColumnMajor, Kind = internal::KindLeaf };
    StridedView(const DevicePointer<_Scalar>& owner, const internal::LeafDesc& d) : owner_(owner), d_(d) {}
    Index rows() const { return d_.rows; }
    Index cols() const { return d_.cols; }
    Index batches() const { return d_.batches; }
    internal::LeafDesc leafDesc() const { return d_; }
    void srcOf(cm_instr& i, internal::ProgramBuilder& b) const
    {
        i.src = CM_SRC_LEAF;
        i.arg = b.leaf(d_);
    }
    void emit(internal::ProgramBuilder& b, int depth) const
    {
        internal::checkDepth(depth);
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_LOAD, internal::DType<_Scalar>::value);
        srcOf(i, b);
        b.push(i);
    }
};

// ------------------------------------------------------------------ unary
template <typename _Child, int _Fn>
class UnaryOp : public MatrixBase<UnaryOp<_Child, _Fn>> {
    _Child child_;

public:
    typedef typename _Child::Scalar ChildScalar;
    typedef typename std::conditional<_Fn == CM_UN_NOT, bool, ChildScalar>::type Scalar;
    enum {
        Rows = _Child::Rows, Columns = _Child::Columns, Batches = _Child::Batches, Flags = _Child::Flags,
        Kind = internal::KindExpr
    };
    explicit UnaryOp(const _Child& c) : child_(c) {}
    Index rows() const { return child_.rows(); }
    Index cols() const { return child_.cols(); }
    Index batches() const { return child_.batches(); }
    void emit(internal::ProgramBuilder& b, int depth) const
    {
        child_.emit(b, depth);
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_UNARY, internal::DType<ChildScalar>::value);
        i.fn = _Fn;
        b.push(i);
        if (!std::is_same<Scalar, ChildScalar>::value) {
            cm_instr c = internal::ProgramBuilder::instr(CM_OP_CAST, internal::DType<Scalar>::value);
            c.arg = CM_I32;
            b.push(c);
        }
    }
};

// ------------------------------------------------------------------ cast
template <typename _Child, typename _Target>
class CastingOp : public MatrixBase<CastingOp<_Child, _Target>> {
    _Child child_;

public:
    typedef _Target Scalar;
    enum {
        Rows = _Child::Rows, Columns = _Child::Columns, Batches = _Child::Batches, Flags = _Child::Flags,
        Kind = internal::KindExpr
    };
    explicit CastingOp(const _Child& c) : child_(c) {}
    Index rows() const { return child_.rows(); }
    Index cols() const { return child_.cols(); }
    Index batches() const { return child_.batches(); }
    void emit(internal::ProgramBuilder& b, int depth) const
    {
        child_.emit(b, depth);
        if (std::is_same<_Target, typename _Child::Scalar>::value) return;
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_CAST, internal::DType<_Target>::value);
        i.arg = internal::DType<typename _Child::Scalar>::value;
        b.push(i);
    }
};

// ------------------------------------------------------------------ binary
template <typename _Left, typename _Right, int _Fn>
class BinaryOp : public MatrixBase<BinaryOp<_Left, _Right, _Fn>> {
    _Left left_;
    _Right right_;

    static_assert(std::is_same<typename _Left::Scalar, typename _Right::Scalar>::value,
                  "No implicit casting is allowed in binary operations; use .cast<T>()");
    static_assert(!internal::intOnly(_Fn) || !internal::IsFloating<typename _Left::Scalar>::value,
                  "this functor is only defined for integer types");

public:
    typedef typename _Left::Scalar OperandScalar;
    typedef typename std::conditional<internal::boolResult(_Fn), bool, OperandScalar>::type Scalar;
    enum {
        Rows = internal::BroadcastDim<_Left::Rows, _Right::Rows>::value,
        Columns = internal::BroadcastDim<_Left::Columns, _Right::Columns>::value,
        Batches = internal::BroadcastDim<_Left::Batches, _Right::Batches>::value,
        Flags = (int)_Left::Kind == (int)internal::KindScalar ? (int)_Right::Flags : (int)_Left::Flags,
        Kind = internal::KindExpr
    };

    BinaryOp(const _Left& l, const _Right& r) : left_(l), right_(r)
    {
        CUMAT_ASSERT_DIMENSION(l.rows() == r.rows() || l.rows() == 1 || r.rows() == 1);
        CUMAT_ASSERT_DIMENSION(l.cols() == r.cols() || l.cols() == 1 || r.cols() == 1);
        CUMAT_ASSERT_DIMENSION(l.batches() == r.batches() || l.batches() == 1 || r.batches() == 1);
    }
    Index rows() const { return left_.rows() == 1 ? right_.rows() : left_.rows(); }
    Index cols() const { return left_.cols() == 1 ? right_.cols() : left_.cols(); }
    Index batches() const { return left_.batches() == 1 ? right_.batches() : left_.batches(); }
    const _Left& left() const { return left_; }
    const _Right& right() const { return right_; }

    // true for leaf * scalar, which a sum can take as a multiply-add term
    static constexpr bool IsAxpyTerm = _Fn == CM_BIN_MUL && (int)_Left::Kind == (int)internal::KindLeaf &&
                                       (int)_Right::Kind == (int)internal::KindScalar;
    // fills an FMA instruction with this term's leaf and scalar
    void fmaTerm(internal::ProgramBuilder& b, cm_instr& i) const
    {
        if constexpr (IsAxpyTerm) {
            i.arg = b.leaf(left_.leafDesc());
            if (internal::IsFloating<OperandScalar>::value) i.fimm = (double)right_.value();
            else i.iimm = (int64_t)right_.value();
        }
    }

    void emit(internal::ProgramBuilder& b, int depth) const
    {
        internal::checkDepth(depth);
        const int odt = internal::DType<OperandScalar>::value;
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_BINARY, odt);
        i.fn = _Fn;
        if constexpr (_Fn == CM_BIN_ADD && internal::is_axpy_term<_Right>::value) {
            // X + leaf*scalar -> one FMA instruction
            cm_instr f = internal::ProgramBuilder::instr(CM_OP_FMA, odt);
            right_.fmaTerm(b, f);
            left_.emit(b, depth);
            b.push(f);
            return;
        } else if constexpr ((int)_Right::Kind != (int)internal::KindExpr) {
            left_.emit(b, depth);
            right_.srcOf(i, b);
        } else if constexpr ((int)_Left::Kind != (int)internal::KindExpr) {
            right_.emit(b, depth);
            left_.srcOf(i, b);
            i.flags = 1;
        } else {
            left_.emit(b, depth);
            cm_instr p = internal::ProgramBuilder::instr(CM_OP_PUSH, odt);
            p.arg = depth;
            b.push(p);
            right_.emit(b, depth + 1);
            i.src = CM_SRC_STACK;
            i.arg = depth;
            i.flags = 1;
        }
        b.push(i);
    }
};

// ------------------------------------------------------------------ operators
namespace internal {
template <typename T>
using if_arith = typename std::enable_if<std::is_arithmetic<T>::value>::type;
}

// with a scalar on either side
#define CUMAT_AMD_BINARY_OPERATOR_SCALAR(op, fn) \
    template <typename L, typename S, typename = internal::if_arith<S>> \
    BinaryOp<L, HostScalar<typename L::Scalar>, fn> operator op(const MatrixBase<L>& l, const S& s) \
    { \
        return BinaryOp<L, HostScalar<typename L::Scalar>, fn>(l.derived(), \
                                                               HostScalar<typename L::Scalar>((typename L::Scalar)s)); \
    } \
    template <typename S, typename R, typename = internal::if_arith<S>> \
    BinaryOp<HostScalar<typename R::Scalar>, R, fn> operator op(const S& s, const MatrixBase<R>& r) \
    { \
        return BinaryOp<HostScalar<typename R::Scalar>, R, fn>(HostScalar<typename R::Scalar>((typename R::Scalar)s), \
                                                               r.derived()); \
    }
// between two matrices / expressions, and with a scalar
#define CUMAT_AMD_BINARY_OPERATOR(op, fn) \
    template <typename L, typename R> \
    BinaryOp<L, R, fn> operator op(const MatrixBase<L>& l, const MatrixBase<R>& r) \
    { \
        return BinaryOp<L, R, fn>(l.derived(), r.derived()); \
    } \
    CUMAT_AMD_BINARY_OPERATOR_SCALAR(op, fn)
// As in cuMat (cuMat/src/BinaryOps.h:286-293), * and / between two matrices
// are not component-wise: matrix * matrix is the matrix product (ProductOp.h)
// and the component-wise forms are cwiseMul / cwiseDiv.
CUMAT_AMD_BINARY_OPERATOR(+, CM_BIN_ADD)
CUMAT_AMD_BINARY_OPERATOR(-, CM_BIN_SUB)
CUMAT_AMD_BINARY_OPERATOR_SCALAR(*, CM_BIN_MUL)
CUMAT_AMD_BINARY_OPERATOR_SCALAR(/, CM_BIN_DIV)
CUMAT_AMD_BINARY_OPERATOR(%, CM_BIN_MOD)
CUMAT_AMD_BINARY_OPERATOR(&, CM_BIN_AND)
CUMAT_AMD_BINARY_OPERATOR(|, CM_BIN_OR)
CUMAT_AMD_BINARY_OPERATOR(^, CM_BIN_XOR)
CUMAT_AMD_BINARY_OPERATOR(&&, CM_BIN_LAND)
CUMAT_AMD_BINARY_OPERATOR(||, CM_BIN_LOR)
CUMAT_AMD_BINARY_OPERATOR(==, CM_BIN_EQ)
CUMAT_AMD_BINARY_OPERATOR(!=, CM_BIN_NE)
CUMAT_AMD_BINARY_OPERATOR(<, CM_BIN_LT)
CUMAT_AMD_BINARY_OPERATOR(>, CM_BIN_GT)
CUMAT_AMD_BINARY_OPERATOR(<=, CM_BIN_LE)
CUMAT_AMD_BINARY_OPERATOR(>=, CM_BIN_GE)
#undef CUMAT_AMD_BINARY_OPERATOR
#undef CUMAT_AMD_BINARY_OPERATOR_SCALAR

template <typename E>
UnaryOp<E, CM_UN_NEG> operator-(const MatrixBase<E>& e) { return UnaryOp<E, CM_UN_NEG>(e.derived()); }
template <typename E>
UnaryOp<E, CM_UN_NOT> operator!(const MatrixBase<E>& e) { return UnaryOp<E, CM_UN_NOT>(e.derived()); }
template <typename E>
UnaryOp<E, CM_UN_BITNOT> operator~(const MatrixBase<E>& e) { return UnaryOp<E, CM_UN_BITNOT>(e.derived()); }

// ------------------------------------------------------------------ evaluation
namespace internal {

// dst (mode)= expr over a strided destination in one cm_cwise_eval launch
template <typename E>
void evalCwise(void* dst, int dst_dtype, Index rows, Index cols, Index batches, int64_t sr, int64_t sc, int64_t sb,
               const MatrixBase<E>& expr, AssignmentMode mode)
{
    const E& e = expr.derived();
    CUMAT_ASSERT_DIMENSION(e.rows() == rows || e.rows() == 1);
    CUMAT_ASSERT_DIMENSION(e.cols() == cols || e.cols() == 1);
    CUMAT_ASSERT_DIMENSION(e.batches() == batches || e.batches() == 1);
    if (rows * cols * batches == 0) return;
    ProgramBuilder b;
    e.emit(b, 0);
    std::vector<cm_leaf> leaves(b.leaves.size() ? b.leaves.size() : 1);
    for (size_t k = 0; k < b.leaves.size(); ++k) {
        const LeafDesc& l = b.leaves[k];
        if (l.sr < 0 || l.sc < 0)
            throw std::invalid_argument("cuMat: SparseMatrix::direct() is only valid in assignments to a sparse "
                                        "matrix of the same pattern");
        leaves[k].ptr = l.ptr;
        leaves[k].dtype = l.dtype;
        leaves[k].kind = l.kind;
        leaves[k].sr = (l.rows == 1 && rows != 1) ? 0 : l.sr;
        leaves[k].sc = (l.cols == 1 && cols != 1) ? 0 : l.sc;
        leaves[k].sb = (l.batches == 1 && batches != 1) ? 0 : l.sb;
    }
    CUMAT_AMD_CALL(cm_cwise_eval(Context::current().handle(), leaves.data(), (int)b.leaves.size(), b.code.data(),
                                 (int)b.code.size(), dst, dst_dtype, (int)rows, (int)cols, (int)batches, sr, sc, sb,
                                 (int)mode));
}

inline void layoutStrides(int flags, Index rows, Index cols, int64_t& sr, int64_t& sc, int64_t& sb)
{
    if (flags & RowMajor) {
        sc = 1;
        sr = cols;
    } else {
        sr = 1;
        sc = rows;
    }
    sb = (int64_t)rows * cols;
}

}  // namespace internal

CUMAT_NAMESPACE_END

#endif
