#ifndef CUMAT_AMD_MATRIX_H
#define CUMAT_AMD_MATRIX_H

// Matrix<Scalar, Rows, Columns, Batches, Flags>: a reference-counted device
// array with compile-time or dynamic dimensions, the storage type of cuMat
// (cuMat/src/Matrix.h:30-436,477-1292). Element (r, c, b) lives at
// r + rows*(c + cols*b) (column major) or c + cols*(r + rows*b) (row major)
// (cuMat/src/Matrix.h:587-627). Copies are shallow; operator= from an
// expression reuses the memory only when it is exclusively owned and of the
// same size, inplace() always writes into it (cuMat/src/Matrix.h:1017-1040,
// 1128-1139).

#include <vector>

#include "CwiseOp.h"
#include "EigenInterop.h"

CUMAT_NAMESPACE_BEGIN

namespace internal {

// A writable strided window on an existing allocation: the target of
// Matrix::inplace() and Matrix::slice(b). Assignments write into the
// window; as an expression it is a leaf.
template <typename _Scalar, int _Flags>
class DenseRef : public MatrixBase<DenseRef<_Scalar, _Flags>> {
    DevicePointer<_Scalar> data_;
    Index rows_, cols_, batches_;

public:
    typedef _Scalar Scalar;
    enum { Rows = Dynamic, Columns = Dynamic, Batches = Dynamic, Flags = _Flags, Kind = KindLeaf };
    DenseRef(const DevicePointer<_Scalar>& d, Index r, Index c, Index b) : data_(d), rows_(r), cols_(c), batches_(b) {}
    Index rows() const { return rows_; }
    Index cols() const { return cols_; }
    Index batches() const { return batches_; }
    _Scalar* data() const { return data_.pointer(); }
    const DevicePointer<_Scalar>& dataPointer() const { return data_; }

    LeafDesc leafDesc() const
    {
        LeafDesc d{};
        d.ptr = data_.pointer();
        d.dtype = DType<_Scalar>::value;
        d.kind = CM_LEAF_STRIDED;
        layoutStrides(_Flags, rows_, cols_, d.sr, d.sc, d.sb);
        d.rows = rows_;
        d.cols = cols_;
        d.batches = batches_;
        return d;
    }
    void srcOf(cm_instr& i, ProgramBuilder& b) const
    {
        i.src = CM_SRC_LEAF;
        i.arg = b.leaf(leafDesc());
    }
    void emit(ProgramBuilder& b, int depth) const
    {
        checkDepth(depth);
        cm_instr i = ProgramBuilder::instr(CM_OP_LOAD, DType<_Scalar>::value);
        srcOf(i, b);
        b.push(i);
    }

    template <typename E>
    void assign(const MatrixBase<E>& e, AssignmentMode mode)
    {
        static_assert(std::is_same<typename E::Scalar, _Scalar>::value, "No implicit casting in assignments");
        int64_t sr, sc, sb;
        layoutStrides(_Flags, rows_, cols_, sr, sc, sb);
        evalCwise(data_.pointer(), DType<_Scalar>::value, rows_, cols_, batches_, sr, sc, sb, e, mode);
    }
    template <typename E> DenseRef& operator=(const MatrixBase<E>& e) { assign(e, AssignmentMode::ASSIGN); return *this; }
    DenseRef& operator=(const DenseRef& e) { assign(e, AssignmentMode::ASSIGN); return *this; }
    template <typename E> DenseRef& operator+=(const MatrixBase<E>& e) { assign(e, AssignmentMode::ADD); return *this; }
    template <typename E> DenseRef& operator-=(const MatrixBase<E>& e) { assign(e, AssignmentMode::SUB); return *this; }
    template <typename E> DenseRef& operator*=(const MatrixBase<E>& e) { assign(e, AssignmentMode::MUL); return *this; }
    template <typename E> DenseRef& operator/=(const MatrixBase<E>& e) { assign(e, AssignmentMode::DIV); return *this; }

    // sparse products written into the window (inplace() = A * x)
    template <typename A, typename V>
    DenseRef& operator=(const ProductOp<A, V>& p)
    {
        CUMAT_ASSERT_DIMENSION(cols_ == 1);
        p.evalTo(data(), rows_, batches_, AssignmentMode::ASSIGN);
        return *this;
    }
    template <typename A, typename V>
    DenseRef& operator+=(const ProductOp<A, V>& p)
    {
        CUMAT_ASSERT_DIMENSION(cols_ == 1);
        p.evalTo(data(), rows_, batches_, AssignmentMode::ADD);
        return *this;
    }
};

}  // namespace internal

template <typename _Scalar, int _Rows, int _Columns, int _Batches, int _Flags>
class Matrix : public MatrixBase<Matrix<_Scalar, _Rows, _Columns, _Batches, _Flags>> {
    DevicePointer<_Scalar> data_;
    Index rows_ = 0, cols_ = 0, batches_ = 0;

    static Index fixed(int s, Index v) { return s == Dynamic ? v : (Index)s; }
    void checkDims() const
    {
        CUMAT_ASSERT_DIMENSION(_Rows == Dynamic || rows_ == _Rows);
        CUMAT_ASSERT_DIMENSION(_Columns == Dynamic || cols_ == _Columns);
        CUMAT_ASSERT_DIMENSION(_Batches == Dynamic || batches_ == _Batches);
        CUMAT_ASSERT_DIMENSION(rows_ >= 0 && cols_ >= 0 && batches_ >= 0);
    }
    void allocate()
    {
        checkDims();
        data_ = DevicePointer<_Scalar>(rows_ * cols_ * batches_);
    }

public:
    typedef _Scalar Scalar;
    typedef Matrix Type;
    enum { Rows = _Rows, Columns = _Columns, Batches = _Batches, Flags = _Flags, Kind = internal::KindLeaf };

    // fixed dimensions allocate, dynamic ones start empty
    Matrix() : rows_(fixed(_Rows, 0)), cols_(fixed(_Columns, 0)), batches_(fixed(_Batches, 0)) { allocate(); }
    // vector constructor (compile-time row or column vector)
    explicit Matrix(Index size)
    {
        static_assert(_Rows == 1 || _Columns == 1, "the one-size constructor is for vectors");
        rows_ = _Columns == 1 ? size : 1;
        cols_ = _Columns == 1 ? 1 : size;
        batches_ = _Batches == Dynamic ? 1 : _Batches;
        if (_Rows == 1 && _Columns == 1) rows_ = cols_ = 1;
        allocate();
    }
    Matrix(Index rows, Index cols) : rows_(rows), cols_(cols), batches_(_Batches == Dynamic ? 1 : _Batches)
    {
        allocate();
    }
    Matrix(Index rows, Index cols, Index batches) : rows_(rows), cols_(cols), batches_(batches) { allocate(); }
    Matrix(const DevicePointer<_Scalar>& ptr, Index rows, Index cols, Index batches)
        : data_(ptr), rows_(rows), cols_(cols), batches_(batches)
    {
        checkDims();
    }

    Matrix(const Matrix&) = default;
    Matrix(Matrix&&) = default;
    Matrix& operator=(const Matrix&) = default;  // shallow, as cuMat
    Matrix& operator=(Matrix&&) = default;

    // shallow conversion between types of the same scalar and storage order
    template <int R2, int C2, int B2>
    Matrix(const Matrix<_Scalar, R2, C2, B2, _Flags>& o)
        : data_(o.dataPointer()), rows_(o.rows()), cols_(o.cols()), batches_(o.batches())
    {
        checkDims();
    }
    template <int R2, int C2, int B2>
    Matrix& operator=(const Matrix<_Scalar, R2, C2, B2, _Flags>& o)
    {
        data_ = o.dataPointer();
        rows_ = o.rows();
        cols_ = o.cols();
        batches_ = o.batches();
        checkDims();
        return *this;
    }

    // evaluation of any component-wise expression (one kernel)
    template <typename E>
    Matrix(const MatrixBase<E>& e) : rows_(e.rows()), cols_(e.cols()), batches_(e.batches())
    {
        allocate();
        inplaceAssign(e, AssignmentMode::ASSIGN);
    }
    template <typename E>
    Matrix& operator=(const MatrixBase<E>& e)
    {
        if (!isExclusiveUse() || rows_ != e.rows() || cols_ != e.cols() || batches_ != e.batches()) {
            Matrix fresh(e.rows(), e.cols(), e.batches());
            fresh.inplaceAssign(e, AssignmentMode::ASSIGN);
            *this = fresh;
        } else {
            inplaceAssign(e, AssignmentMode::ASSIGN);
        }
        return *this;
    }

    // y = A * x with a sparse A: one SpMV launch
    template <typename A, typename V>
    Matrix(const ProductOp<A, V>& p) : rows_(p.rows()), cols_(1), batches_(p.batches())
    {
        allocate();
        p.evalTo(data(), rows_, batches_, AssignmentMode::ASSIGN);
    }
    template <typename A, typename V>
    Matrix& operator=(const ProductOp<A, V>& p)
    {
        if (!isExclusiveUse() || rows_ != p.rows() || cols_ != 1 || batches_ != p.batches()) {
            Matrix fresh(p.rows(), 1, p.batches());
            p.evalTo(fresh.data(), fresh.rows(), fresh.batches(), AssignmentMode::ASSIGN);
            *this = fresh;
        } else {
            p.evalTo(data(), rows_, batches_, AssignmentMode::ASSIGN);
        }
        return *this;
    }
    template <typename A, typename V>
    Matrix& operator+=(const ProductOp<A, V>& p)
    {
        CUMAT_ASSERT_DIMENSION(cols_ == 1);
        p.evalTo(data(), rows_, batches_, AssignmentMode::ADD);
        return *this;
    }
    template <typename A, typename V>
    Matrix& operator-=(const ProductOp<A, V>& p)
    {
        CUMAT_ASSERT_DIMENSION(cols_ == 1);
        p.evalTo(data(), rows_, batches_, AssignmentMode::SUB);
        return *this;
    }

    // compound assignment writes into the existing memory (cuMat/src/Matrix.h:1040-1126)
#define CUMAT_AMD_COMPOUND(op, mode) \
    template <typename E> \
    Matrix& operator op(const MatrixBase<E>& e) \
    { \
        inplaceAssign(e, mode); \
        return *this; \
    } \
    template <typename S, typename = internal::if_arith<S>> \
    Matrix& operator op(const S& s) \
    { \
        inplaceAssign(HostScalar<_Scalar>((_Scalar)s), mode); \
        return *this; \
    }
    CUMAT_AMD_COMPOUND(+=, AssignmentMode::ADD)
    CUMAT_AMD_COMPOUND(-=, AssignmentMode::SUB)
    CUMAT_AMD_COMPOUND(*=, AssignmentMode::MUL)
    CUMAT_AMD_COMPOUND(/=, AssignmentMode::DIV)
    CUMAT_AMD_COMPOUND(%=, AssignmentMode::MOD)
    CUMAT_AMD_COMPOUND(&=, AssignmentMode::AND)
    CUMAT_AMD_COMPOUND(|=, AssignmentMode::OR)
#undef CUMAT_AMD_COMPOUND

    template <typename E>
    void inplaceAssign(const MatrixBase<E>& e, AssignmentMode mode)
    {
        static_assert(std::is_same<typename E::Scalar, _Scalar>::value,
                      "No implicit casting is allowed in assignments; use .cast<T>()");
        int64_t sr, sc, sb;
        internal::layoutStrides(_Flags, rows_, cols_, sr, sc, sb);
        internal::evalCwise(data_.pointer(), internal::DType<_Scalar>::value, rows_, cols_, batches_, sr, sc, sb, e,
                            mode);
    }

    // ---- dimensions and storage
    Index rows() const { return rows_; }
    Index cols() const { return cols_; }
    Index batches() const { return batches_; }
    Index size() const { return rows_ * cols_ * batches_; }
    _Scalar* data() { return data_.pointer(); }
    const _Scalar* data() const { return data_.pointer(); }
    const DevicePointer<_Scalar>& dataPointer() const { return data_; }
    DevicePointer<_Scalar>& dataPointer() { return data_; }
    bool isExclusiveUse() const { return data_.getCounter() <= 1; }
    void makeExclusiveUse()
    {
        if (isExclusiveUse()) return;
        *this = deepClone();
    }
    Matrix deepClone() const
    {
        Matrix m(rows_, cols_, batches_);
        if (size())
            CUMAT_AMD_CALL(cm_memcpy_d2d(Context::current().handle(), m.data(), data(), sizeof(_Scalar) * size()));
        return m;
    }

    // ---- leaf protocol
    internal::LeafDesc leafDesc() const
    {
        internal::LeafDesc d{};
        d.ptr = data_.pointer();
        d.dtype = internal::DType<_Scalar>::value;
        d.kind = CM_LEAF_STRIDED;
        internal::layoutStrides(_Flags, rows_, cols_, d.sr, d.sc, d.sb);
        d.rows = rows_;
        d.cols = cols_;
        d.batches = batches_;
        return d;
    }
    void srcOf(cm_instr& i, internal::ProgramBuilder& b) const
    {
        i.src = CM_SRC_LEAF;
        i.arg = b.leaf(leafDesc());
    }
    void emit(internal::ProgramBuilder& b, int depth) const
    {
        internal::checkDepth(depth);
        cm_instr i = internal::ProgramBuilder::instr(CM_OP_LOAD, internal::DType<_Scalar>::value);
        srcOf(i, b);
        b.push(i);
    }

    // ---- views
    internal::DenseRef<_Scalar, _Flags> inplace() { return internal::DenseRef<_Scalar, _Flags>(data_, rows_, cols_, batches_); }
    internal::DenseRef<_Scalar, _Flags> slice(Index batch)
    {
        CUMAT_ASSERT_BOUNDS(batch >= 0 && batch < batches_);
        return internal::DenseRef<_Scalar, _Flags>(DevicePointer<_Scalar>(data_, batch * rows_ * cols_), rows_, cols_,
                                                   1);
    }
    StridedView<_Scalar> transpose() const
    {
        internal::LeafDesc d = leafDesc();
        std::swap(d.sr, d.sc);
        std::swap(d.rows, d.cols);
        return StridedView<_Scalar>(data_, d);
    }
    // the main diagonal as a column vector (ExtractDiagonalOp)
    StridedView<_Scalar> diagonal() const
    {
        internal::LeafDesc d = leafDesc();
        d.sr = d.sr + d.sc;
        d.sc = 0;
        d.rows = rows_ < cols_ ? rows_ : cols_;
        d.cols = 1;
        return StridedView<_Scalar>(data_, d);
    }
    // a vector as a diagonal matrix (AsDiagonalOp)
    StridedView<_Scalar> asDiagonal() const
    {
        internal::LeafDesc d = leafDesc();
        Index n = cols_ == 1 ? rows_ : cols_;
        d.sr = cols_ == 1 ? d.sr : d.sc;
        d.sc = 0;
        d.rows = d.cols = n;
        d.kind = CM_LEAF_AS_DIAGONAL;
        return StridedView<_Scalar>(data_, d);
    }

    // ---- nullary factories (cuMat/src/NullaryOps.h:78-108)
    static ConstantOp<_Scalar> Constant(Index rows, Index cols, Index batches, const _Scalar& v)
    {
        return ConstantOp<_Scalar>(rows, cols, batches, v);
    }
    static ConstantOp<_Scalar> Constant(Index rows, Index cols, const _Scalar& v)
    {
        return ConstantOp<_Scalar>(rows, cols, _Batches == Dynamic ? 1 : _Batches, v);
    }
    static ConstantOp<_Scalar> Constant(Index size, const _Scalar& v)
    {
        return _Columns == 1 ? ConstantOp<_Scalar>(size, 1, _Batches == Dynamic ? 1 : _Batches, v)
                             : ConstantOp<_Scalar>(1, size, _Batches == Dynamic ? 1 : _Batches, v);
    }
    static ConstantOp<_Scalar> Zero()
    {
        return ConstantOp<_Scalar>(fixed(_Rows, 1), fixed(_Columns, 1), fixed(_Batches, 1), _Scalar(0));
    }
    static ConstantOp<_Scalar> Zero(Index size) { return Constant(size, _Scalar(0)); }
    static ConstantOp<_Scalar> Zero(Index rows, Index cols) { return Constant(rows, cols, _Scalar(0)); }
    static ConstantOp<_Scalar> Zero(Index rows, Index cols, Index batches) { return Constant(rows, cols, batches, _Scalar(0)); }
    // identity matrix as an expression: an as-diagonal view of a device 1
    static StridedView<_Scalar> Identity(Index rows, Index cols, Index batches = 1)
    {
        Matrix<_Scalar, 1, 1, 1, ColumnMajor> one;
        one.inplace() = HostScalar<_Scalar>(_Scalar(1));
        internal::LeafDesc d{};
        d.ptr = one.data();
        d.dtype = internal::DType<_Scalar>::value;
        d.kind = CM_LEAF_AS_DIAGONAL;
        d.sr = d.sc = d.sb = 0;
        d.rows = rows;
        d.cols = cols;
        d.batches = batches;
        return StridedView<_Scalar>(one.dataPointer(), d);
    }
    static StridedView<_Scalar> Identity(Index size) { return Identity(size, size, 1); }
    void setZero()
    {
        if (size())
            CUMAT_AMD_CALL(cm_fill_bytes(Context::current().handle(), data(), 0, sizeof(_Scalar) * size()));
    }

    // ---- host transfer (synchronous for device -> host, cuMat/src/Matrix.h:826-858)
    void copyFromHost(const _Scalar* host)
    {
        if (size())
            CUMAT_AMD_CALL(cm_memcpy_h2d(Context::current().handle(), data(), host, sizeof(_Scalar) * size()));
    }
    void copyToHost(_Scalar* host) const
    {
        if (size())
            CUMAT_AMD_CALL(cm_memcpy_d2h(Context::current().handle(), host, data(), sizeof(_Scalar) * size()));
    }
    // element (r, c, b) read back to the host
    _Scalar getCoeff(Index r, Index c, Index b) const
    {
        CUMAT_ASSERT_BOUNDS(r >= 0 && r < rows_ && c >= 0 && c < cols_ && b >= 0 && b < batches_);
        int64_t sr, sc, sb;
        internal::layoutStrides(_Flags, rows_, cols_, sr, sc, sb);
        _Scalar v;
        CUMAT_AMD_CALL(cm_memcpy_d2h(Context::current().handle(), &v, data() + r * sr + c * sc + b * sb,
                                     sizeof(_Scalar)));
        return v;
    }
    // all elements in logical [batch][row][col] order
    std::vector<_Scalar> toHostRowMajor() const
    {
        std::unique_ptr<_Scalar[]> raw(new _Scalar[(size_t)size() + 1]);  // not vector: vector<bool> has no data()
        std::vector<_Scalar> out((size_t)size());
        copyToHost(raw.get());
        int64_t sr, sc, sb;
        internal::layoutStrides(_Flags, rows_, cols_, sr, sc, sb);
        size_t k = 0;
        for (Index b = 0; b < batches_; ++b)
            for (Index r = 0; r < rows_; ++r)
                for (Index c = 0; c < cols_; ++c) out[k++] = raw[(size_t)(r * sr + c * sc + b * sb)];
        return out;
    }
    // the value of a 1x1x1 matrix
    _Scalar getScalar() const
    {
        CUMAT_ASSERT_DIMENSION(size() == 1);
        return getCoeff(0, 0, 0);
    }

    // a row-major matrix with the contents of a [batch][row][col] array
    // (cuMat/src/Matrix.h:806-817)
    template <int R, int C, int B>
    static Matrix<_Scalar, (R > 1) ? Dynamic : R, (C > 1) ? Dynamic : C, (B > 1) ? Dynamic : B, RowMajor>
    fromArray(const _Scalar (&a)[B][R][C])
    {
        Matrix<_Scalar, (R > 1) ? Dynamic : R, (C > 1) ? Dynamic : C, (B > 1) ? Dynamic : B, RowMajor> m(R, C, B);
        m.copyFromHost(&a[0][0][0]);
        return m;
    }
    // from host data in this matrix's storage order
    static Matrix fromHost(const _Scalar* host, Index rows, Index cols, Index batches)
    {
        Matrix m(rows, cols, batches);
        m.copyFromHost(host);
        return m;
    }
    static Matrix fromHost(const std::vector<_Scalar>& v)
    {
        static_assert(_Rows == 1 || _Columns == 1, "fromHost(vector) builds a vector");
        Matrix m((Index)v.size());
        m.copyFromHost(v.data());
        return m;
    }

#if CUMAT_EIGEN_SUPPORT == 1
    // ---- Eigen interop (cuMat/src/Matrix.h:860-930): synchronous copies
    // through host memory; one batch only, Eigen has no batches
    typedef typename eigen::EigenTypeOf<_Scalar, _Rows, _Columns, _Flags>::type EigenMatrix_t;
    EigenMatrix_t toEigen() const
    {
        static_assert(_Batches == 1 || _Batches == Dynamic, "Eigen matrices have one batch");
        CUMAT_ASSERT_ARGUMENT(batches_ == 1);
        EigenMatrix_t m(rows_, cols_);
        copyToHost(m.data());
        return m;
    }
    static Matrix fromEigen(const EigenMatrix_t& m)
    {
        static_assert(_Batches == 1 || _Batches == Dynamic, "Eigen matrices have one batch");
        Matrix out(m.rows(), m.cols(), 1);
        out.copyFromHost(m.data());
        return out;
    }
#endif
};

template <typename _Derived>
auto MatrixBase<_Derived>::eval() const
{
    return Matrix<typename _Derived::Scalar, Dynamic, Dynamic, Dynamic, ColumnMajor>(derived());
}

// typedefs of cuMat/src/Matrix.h:1396-1439 for the scalar types of the path
#define CUMAT_AMD_TYPEDEFS1(S, s, F, o) \
    typedef Matrix<S, 1, 1, 1, F> Scalar##s##o; \
    typedef Matrix<S, 1, 1, Dynamic, F> BScalar##s##o; \
    typedef Matrix<S, 2, 1, 1, F> Vector2##s##o; \
    typedef Matrix<S, 2, 1, Dynamic, F> BVector2##s##o; \
    typedef Matrix<S, 3, 1, 1, F> Vector3##s##o; \
    typedef Matrix<S, 3, 1, Dynamic, F> BVector3##s##o; \
    typedef Matrix<S, 4, 1, 1, F> Vector4##s##o; \
    typedef Matrix<S, 4, 1, Dynamic, F> BVector4##s##o; \
    typedef Matrix<S, Dynamic, 1, 1, F> VectorX##s##o; \
    typedef Matrix<S, Dynamic, 1, Dynamic, F> BVectorX##s##o; \
    typedef Matrix<S, 1, 2, 1, F> RowVector2##s##o; \
    typedef Matrix<S, 1, 2, Dynamic, F> BRowVector2##s##o; \
    typedef Matrix<S, 1, 3, 1, F> RowVector3##s##o; \
    typedef Matrix<S, 1, 3, Dynamic, F> BRowVector3##s##o; \
    typedef Matrix<S, 1, 4, 1, F> RowVector4##s##o; \
    typedef Matrix<S, 1, 4, Dynamic, F> BRowVector4##s##o; \
    typedef Matrix<S, 1, Dynamic, 1, F> RowVectorX##s##o; \
    typedef Matrix<S, 1, Dynamic, Dynamic, F> BRowVectorX##s##o; \
    typedef Matrix<S, 2, 2, 1, F> Matrix2##s##o; \
    typedef Matrix<S, 2, 2, Dynamic, F> BMatrix2##s##o; \
    typedef Matrix<S, 3, 3, 1, F> Matrix3##s##o; \
    typedef Matrix<S, 3, 3, Dynamic, F> BMatrix3##s##o; \
    typedef Matrix<S, 4, 4, 1, F> Matrix4##s##o; \
    typedef Matrix<S, 4, 4, Dynamic, F> BMatrix4##s##o; \
    typedef Matrix<S, Dynamic, Dynamic, 1, F> MatrixX##s##o; \
    typedef Matrix<S, Dynamic, Dynamic, Dynamic, F> BMatrixX##s##o;
#define CUMAT_AMD_TYPEDEFS(S, s) \
    CUMAT_AMD_TYPEDEFS1(S, s, ColumnMajor, C) \
    CUMAT_AMD_TYPEDEFS1(S, s, RowMajor, R) \
    CUMAT_AMD_TYPEDEFS1(S, s, ColumnMajor, )
CUMAT_AMD_TYPEDEFS(bool, b)
CUMAT_AMD_TYPEDEFS(int, i)
CUMAT_AMD_TYPEDEFS(float, f)
CUMAT_AMD_TYPEDEFS(double, d)
#undef CUMAT_AMD_TYPEDEFS
#undef CUMAT_AMD_TYPEDEFS1

CUMAT_NAMESPACE_END

#endif
