// CabanaPD_Indexing.hpp -- the material-pair indexing functors of the drop-in
// API (CabanaPD_Indexing.hpp, CabanaPD_Types.hpp:164-191): an indexing maps
// the material types (a, b) of a bond's two particles to the position of the
// bond's model in the model list given to createMultiForceModel.
//
//   DiagonalIndexing<N>  symmetric; diagonal order d = |b - a| is stored after
//                        the N + (N - 1) + .. + (N - d + 1) entries of the lower
//                        orders, then min(a, b) along the diagonal.
//   BinaryIndexing       symmetric; 0 for equal types, 1 otherwise (any number
//                        of types).
//   FullIndexing<N>      a * N + b.
//
// Host code only: the engine evaluates the indexing once per type pair when
// the pair table is built, so out-of-range types throw std::out_of_range
// (the reference aborts inside the kernel).
#ifndef CABANAPD_DROPIN_INDEXING_HPP
#define CABANAPD_DROPIN_INDEXING_HPP

#include <limits>
#include <stdexcept>
#include <type_traits>

#include "CabanaPD_Types.hpp"

namespace CabanaPD
{
template <unsigned NumBaseModels>
struct DiagonalIndexing
{
    static_assert( NumBaseModels > 0, "NumBaseModels must be larger than 0" );
    static constexpr unsigned MaxValidIndex = NumBaseModels * ( NumBaseModels + 1 ) / 2;

    unsigned operator()( unsigned a, unsigned b ) const
    {
        if ( a >= NumBaseModels || b >= NumBaseModels )
            throw std::out_of_range( "Index out of range of DiagonalIndexing" );
        const unsigned lo = a < b ? a : b;
        const unsigned d = ( a < b ? b : a ) - lo;
        // entries of the diagonals 0 .. d-1: N + (N-1) + .. + (N-d+1)
        const unsigned before = d * NumBaseModels - d * ( d - 1 ) / 2;
        return before + lo;
    }
};

struct BinaryIndexing
{
    static constexpr unsigned MaxValidIndex = std::numeric_limits<unsigned>::max();
    unsigned operator()( unsigned a, unsigned b ) const { return a != b ? 1u : 0u; }
};

template <unsigned NumBaseModels>
struct FullIndexing
{
    static_assert( NumBaseModels > 0, "NumBaseModels must be larger than 0" );
    static constexpr unsigned MaxValidIndex = NumBaseModels * NumBaseModels;

    unsigned operator()( unsigned a, unsigned b ) const
    {
        if ( a >= NumBaseModels || b >= NumBaseModels )
            throw std::out_of_range( "Index out of range of FullIndexing" );
        return a * NumBaseModels + b;
    }
};

template <unsigned N>
struct is_symmetric<DiagonalIndexing<N>> : std::true_type
{
};
template <>
struct is_symmetric<BinaryIndexing> : std::true_type
{
};

namespace detail
{
// Number of material types an indexing covers; 0 when it covers any number
// (the particles' types then decide).
template <class Indexing>
struct IndexingTypes : std::integral_constant<unsigned, 0>
{
};
template <unsigned N>
struct IndexingTypes<DiagonalIndexing<N>> : std::integral_constant<unsigned, N>
{
};
template <unsigned N>
struct IndexingTypes<FullIndexing<N>> : std::integral_constant<unsigned, N>
{
};
} // namespace detail

} // namespace CabanaPD

#endif
