#ifndef CUMAT_AMD_BLOCKSCALAR_H
#define CUMAT_AMD_BLOCKSCALAR_H

// Block scalars for SparseMatrix and ConjugateGradient: the reference runs
// its CG over block types through template plug-ins
// (cuMat/src/ConjugateGradient.h:47-61), and its blocked CG test defines
// them as a float4 vector entry (real3, w = 0) and a 3x3 block of three
// float4 rows (real3x3) with a four-term dot per row
// (tests/TestBlockedConjugateGradient.cu:12-239). Here the block semantics
// are those of that test, fixed, and a block type only declares its shape:
//
//   template <> struct cuMat::BlockScalarTraits<MyBlock> {
//       static constexpr int rows = 3;   // float4 rows per block (2..4)
//       typedef MyVec4 Vector;           // the 16-byte vector entry type
//   };
//
// The built-in types BlockRows<D> / Float4 have it already. A block type
// must be D float4 rows laid out contiguously (16 * D bytes), a vector type
// four floats (16 bytes, 16-byte aligned in device memory).

#include <type_traits>

#include "Macros.h"

CUMAT_NAMESPACE_BEGIN

struct alignas(16) Float4 {
    float x, y, z, w;
};

template <int D>
struct BlockRows {
    static_assert(D >= 2 && D <= 4, "blocks of 2 to 4 float4 rows");
    Float4 r[D];
};

template <typename T>
struct BlockScalarTraits {
    static constexpr int rows = 0;  // not a block scalar
};
template <int D>
struct BlockScalarTraits<BlockRows<D>> {
    static constexpr int rows = D;
    typedef Float4 Vector;
};

namespace internal {

template <typename T>
struct is_block_scalar : std::integral_constant<bool, (BlockScalarTraits<T>::rows > 0)> {};

template <typename T, bool B = is_block_scalar<T>::value>
struct block_checked {
    typedef T type;
};
template <typename T>
struct block_checked<T, true> {
    typedef typename BlockScalarTraits<T>::Vector Vector;
    static_assert(sizeof(T) == 16 * BlockScalarTraits<T>::rows, "a block is its float4 rows, contiguous");
    static_assert(sizeof(Vector) == 16, "a block vector entry is four floats");
    static_assert(std::is_trivially_copyable<T>::value && std::is_trivially_copyable<Vector>::value,
                  "block types are copied as bytes");
    typedef T type;
};

// the entry type of the vectors a matrix of scalar type S multiplies
template <typename S, bool B = is_block_scalar<S>::value>
struct vector_scalar {
    typedef S type;
};
template <typename S>
struct vector_scalar<S, true> {
    typedef typename BlockScalarTraits<typename block_checked<S>::type>::Vector type;
};

// the scalar type of tolerances and errors (NumTraits<...>::ElementalType of
// the reference's test: float for the block types)
template <typename S, bool B = is_block_scalar<S>::value>
struct elemental_scalar {
    typedef S type;
};
template <typename S>
struct elemental_scalar<S, true> {
    typedef float type;
};

}  // namespace internal

CUMAT_NAMESPACE_END

#endif
