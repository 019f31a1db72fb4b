#ifndef CUMAT_AMD_IO_H
#define CUMAT_AMD_IO_H

// operator<< for matrices, for debugging (it copies to the host): a header
// line with the shape and storage, then the values batch by batch, columns
// right-aligned to the widest value, batches wrapped in "[ ... ],[ ... ]"
// when there are several. Same layout as cuMat's printing
// (cuMat/src/Matrix.h:1295-1325, cuMat/src/IO.h:42-230,
// cuMat/src/SparseMatrix.h:417-510), so logs read alike; the sparse form
// ends each section with a newline.

#include <algorithm>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

#include "SparseMatrix.h"

CUMAT_NAMESPACE_BEGIN

namespace io {

struct PrintFormat {
    std::string rowPrefix;
    std::string batchPrefix = "[\n";
    std::string batchSuffix = "\n]";
    std::string batchSeparator = "\n],[\n";
};

template <typename S>
inline void printValue(std::ostream& os, const S& v)
{
    if constexpr (std::is_same<S, bool>::value) os << (v ? 1 : 0);
    else os << v;
}

// v: [batch][row][col] (row-major per batch); transposed prints each column
// of a batch as one line
template <typename S>
void printBatches(std::ostream& os, const std::vector<S>& v, Index rows, Index cols, Index batches,
                  const PrintFormat& f, bool transposed = false)
{
    if (v.empty()) return;
    std::vector<std::string> txt(v.size());
    size_t width = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        std::ostringstream ss;
        ss.copyfmt(os);
        ss.width(0);
        printValue(ss, v[i]);
        txt[i] = ss.str();
        width = std::max(width, txt[i].size());
    }
    const Index lines = transposed ? cols : rows, per = transposed ? rows : cols;
    if (batches > 1) os << f.batchPrefix;
    for (Index b = 0; b < batches; ++b) {
        if (b > 0) os << f.batchSeparator;
        for (Index l = 0; l < lines; ++l) {
            if (l > 0) os << "\n";
            os << f.rowPrefix;
            for (Index k = 0; k < per; ++k) {
                const Index r = transposed ? k : l, c = transposed ? l : k;
                if (k > 0) os << " ";
                const std::string& t = txt[(size_t)((b * rows + r) * cols + c)];
                os << std::string(width - t.size(), ' ') << t;
            }
        }
    }
    if (batches > 1) os << f.batchSuffix;
}

}  // namespace io

template <typename _Scalar, int _Rows, int _Columns, int _Batches, int _Flags>
std::ostream& operator<<(std::ostream& os, const Matrix<_Scalar, _Rows, _Columns, _Batches, _Flags>& m)
{
    os << "Matrix: ";
    os << "rows=" << m.rows() << " (" << (_Rows == Dynamic ? "dynamic" : "compile-time") << ")";
    os << ", cols=" << m.cols() << " (" << (_Columns == Dynamic ? "dynamic" : "compile-time") << ")";
    os << ", batches=" << m.batches() << " (" << (_Batches == Dynamic ? "dynamic" : "compile-time") << ")";
    os << ", storage=" << ((_Flags & Flags::RowMajor) ? "Row-Major" : "Column-Major") << std::endl;
    io::printBatches(os, m.toHostRowMajor(), m.rows(), m.cols(), m.batches(), io::PrintFormat());
    return os;
}

template <typename _Scalar, int _Batches>
std::ostream& operator<<(std::ostream& os, const SparseMatrix<_Scalar, _Batches, SparseFlags::CSR>& m)
{
    os << "SparseMatrix: " << std::endl;
    os << " rows=" << m.rows() << ", cols=" << m.cols();
    os << ", batches=" << m.batches() << " (" << (_Batches == Dynamic ? "dynamic" : "compile-time") << ")";
    os << ", storage=CSR" << std::endl;
    const auto& P = m.getSparsityPattern();
    io::PrintFormat f;
    os << " Outer Indices (row): ";
    io::printBatches(os, P.JA.toHostRowMajor(), P.JA.rows(), 1, 1, f, true);
    os << std::endl << " Inner Indices (column): ";
    io::printBatches(os, P.IA.toHostRowMajor(), P.IA.rows(), 1, 1, f, true);
    os << std::endl << " Data:" << std::endl;
    f.rowPrefix = "   ";
    f.batchPrefix = "  [\n";
    f.batchSuffix = "\n  ]";
    f.batchSeparator = "\n  ],[\n";
    io::printBatches(os, m.getData().toHostRowMajor(), m.getData().rows(), 1, m.getData().batches(), f, true);
    os << std::endl;
    return os;
}

template <typename _Scalar, int _Batches>
std::ostream& operator<<(std::ostream& os, const SparseMatrix<_Scalar, _Batches, SparseFlags::ELLPACK>& m)
{
    os << "SparseMatrix: " << std::endl;
    os << " rows=" << m.rows() << ", cols=" << m.cols();
    os << ", batches=" << m.batches() << " (" << (_Batches == Dynamic ? "dynamic" : "compile-time") << ")";
    os << ", storage=ELLPACK" << std::endl;
    const auto& idx = m.getSparsityPattern().indices;
    os << " Indices:" << std::endl;
    io::printBatches(os, idx.toHostRowMajor(), idx.rows(), idx.cols(), 1, io::PrintFormat());
    os << std::endl;
    const auto& d = m.getData();
    const auto all = d.toHostRowMajor();
    for (Index b = 0; b < d.batches(); ++b) {
        std::vector<_Scalar> one(all.begin() + (size_t)(b * d.rows() * d.cols()),
                                 all.begin() + (size_t)((b + 1) * d.rows() * d.cols()));
        os << " Data (Batch " << b << "):" << std::endl;
        io::printBatches(os, one, d.rows(), d.cols(), 1, io::PrintFormat());
        os << std::endl;
    }
    return os;
}

CUMAT_NAMESPACE_END

#endif
