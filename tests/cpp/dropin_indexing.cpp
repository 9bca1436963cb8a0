// Host-only checks of the drop-in multi-material surface (include/CabanaPD):
// the indexing functors (CabanaPD_Indexing.hpp), the model lists that
// createMultiForceModel builds for AverageTag, the explicit-indexing form and
// the pair table handed to the engine.  Calls no GPU function (the model
// constants come from pdk_model_constants_get, a host routine).
//
//   dropin_indexing      prints one line per check, exits 1 on the first failure
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include <CabanaPD.hpp>

namespace
{
int failures = 0;
void expect( bool ok, const char* what )
{
    std::printf( "%s %s\n", ok ? "ok  " : "FAIL", what );
    if ( !ok )
        failures++;
}
template <class F>
bool throwsOutOfRange( F f )
{
    try
    {
        f();
    }
    catch ( const std::out_of_range& )
    {
        return true;
    }
    return false;
}
bool samePair( const pdk_pair_model& a, const pdk_pair_model& b )
{
    return a.c == b.c && ( a.bond_break_coeff == b.bond_break_coeff ||
                           ( std::isinf( a.bond_break_coeff ) && std::isinf( b.bond_break_coeff ) ) ) &&
           a.theta_coeff[0] == b.theta_coeff[0] && a.theta_coeff[1] == b.theta_coeff[1] &&
           a.s_coeff[0] == b.s_coeff[0] && a.s_coeff[1] == b.s_coeff[1];
}
} // namespace

int main()
{
    // DiagonalIndexing<N>: the N materials first, then each further diagonal
    // |a - b| = d in order of min(a, b); symmetric.
    {
        CabanaPD::DiagonalIndexing<4> idx;
        const unsigned want[4][4] = { { 0, 4, 7, 9 }, { 4, 1, 5, 8 }, { 7, 5, 2, 6 }, { 9, 8, 6, 3 } };
        bool ok = true;
        for ( unsigned a = 0; a < 4; a++ )
            for ( unsigned b = 0; b < 4; b++ )
                ok = ok && idx( a, b ) == want[a][b];
        expect( ok, "DiagonalIndexing<4> table" );
        expect( CabanaPD::DiagonalIndexing<4>::MaxValidIndex == 10, "DiagonalIndexing<4> MaxValidIndex" );
        CabanaPD::DiagonalIndexing<1> one;
        expect( one( 0, 0 ) == 0, "DiagonalIndexing<1>" );
        expect( throwsOutOfRange( [&] { (void)idx( 4, 0 ); } ) && throwsOutOfRange( [&] { (void)idx( 0, 4 ); } ),
                "DiagonalIndexing out of range" );
    }
    {
        CabanaPD::BinaryIndexing idx;
        expect( idx( 0, 0 ) == 0 && idx( 5, 5 ) == 0 && idx( 0, 1 ) == 1 && idx( 7, 2 ) == 1, "BinaryIndexing" );
    }
    {
        CabanaPD::FullIndexing<3> idx;
        bool ok = true;
        for ( unsigned a = 0; a < 3; a++ )
            for ( unsigned b = 0; b < 3; b++ )
                ok = ok && idx( a, b ) == a * 3 + b;
        expect( ok, "FullIndexing<3> table" );
        expect( throwsOutOfRange( [&] { (void)idx( 3, 1 ); } ), "FullIndexing out of range" );
    }
    static_assert( CabanaPD::is_symmetric<CabanaPD::DiagonalIndexing<2>>::value );
    static_assert( CabanaPD::is_symmetric<CabanaPD::BinaryIndexing>::value );
    static_assert( !CabanaPD::is_symmetric<CabanaPD::FullIndexing<2>>::value );
    static_assert( CabanaPD::is_symmetric<CabanaPD::PMB>::value && !CabanaPD::is_symmetric<CabanaPD::LPS>::value );

    const double h = 0.3;
    CabanaPD::Particles particles( Kokkos::HostSpace{}, CabanaPD::PMB{} );

    // AverageTag with three PMB models: DiagonalIndexing<3> over
    // m0, m1, m2, m01, m12, m02; averaged c and K-weighted s0.
    {
        CabanaPD::ForceModel m0( CabanaPD::PMB{}, h, 1.0, 2.0 );
        CabanaPD::ForceModel m1( CabanaPD::PMB{}, h, 3.0, 5.0 );
        CabanaPD::ForceModel m2( CabanaPD::PMB{}, h, 7.0, 11.0 );
        auto models = CabanaPD::createMultiForceModel( particles, CabanaPD::AverageTag{}, m0, m1, m2 );
        static_assert( std::is_same_v<decltype( models )::indexing_type, CabanaPD::DiagonalIndexing<3>> );
        expect( models.models.size() == 6, "AverageTag PMB x3: six models" );
        auto table = models.pairTable( 3 );
        const CabanaPD::ForceModel<CabanaPD::PMB, CabanaPD::Elastic, CabanaPD::Fracture>* ms[3] = { &m0, &m1, &m2 };
        bool ok = true;
        for ( int a = 0; a < 3; a++ )
            for ( int b = 0; b < 3; b++ )
            {
                const pdk_pair_model& p = table[a * 3 + b];
                const double c = a == b ? ms[a]->c : ( ms[a]->c + ms[b]->c ) / 2.0;
                const double s0 =
                    a == b ? ms[a]->s0
                           : std::sqrt( ( ms[a]->s0 * ms[a]->s0 * ms[a]->K + ms[b]->s0 * ms[b]->s0 * ms[b]->K ) /
                                        ( ms[a]->K + ms[b]->K ) );
                ok = ok && p.c == c && p.bond_break_coeff == ( 1.0 + s0 ) * ( 1.0 + s0 );
                ok = ok && samePair( p, table[b * 3 + a] );
            }
        expect( ok, "AverageTag PMB x3: pair table of the averaged models, symmetric" );
    }
    // AverageTag with two LPS models: FullIndexing<2> over m0, m01, m10, m1;
    // the averaged model's theta/s coefficients are [first, second].
    {
        CabanaPD::ForceModel m0( CabanaPD::LPS{}, h, 1.0, 0.4, 2.0 );
        CabanaPD::ForceModel m1( CabanaPD::LPS{}, h, 3.0, 1.1, 5.0 );
        auto models = CabanaPD::createMultiForceModel( particles, CabanaPD::AverageTag{}, m0, m1 );
        static_assert( std::is_same_v<decltype( models )::indexing_type, CabanaPD::FullIndexing<2>> );
        auto table = models.pairTable( 2 );
        const pdk_pair_model& p01 = table[1];
        const pdk_pair_model& p10 = table[2];
        expect( p01.theta_coeff[0] == 3.0 * 1.0 - 5.0 * 0.4 && p01.theta_coeff[1] == 3.0 * 3.0 - 5.0 * 1.1 &&
                    p01.s_coeff[0] == 15.0 * 0.4 && p01.s_coeff[1] == 15.0 * 1.1,
                "AverageTag LPS x2: (0, 1) coefficients [first, second]" );
        expect( p10.theta_coeff[0] == p01.theta_coeff[1] && p10.theta_coeff[1] == p01.theta_coeff[0],
                "AverageTag LPS x2: (1, 0) is the mirror" );
        expect( table[0].theta_coeff[0] == m0.theta_coeff[0] && table[0].theta_coeff[1] == m0.theta_coeff[0] &&
                    table[3].s_coeff[1] == m1.s_coeff[0],
                "AverageTag LPS x2: diagonal entries are the materials" );
        // the same list spelled out for FullIndexing<2> gives the same table
        CabanaPD::ForceModel m01( m0, m1 );
        CabanaPD::ForceModel m10( m1, m0 );
        auto spelled = CabanaPD::createMultiForceModel( particles, CabanaPD::FullIndexing<2>{}, m0, m01, m10, m1 );
        auto t2 = spelled.pairTable( 2 );
        bool ok = true;
        for ( int k = 0; k < 4; k++ )
            ok = ok && samePair( t2[k], table[k] );
        expect( ok, "explicit FullIndexing<2> list equals AverageTag" );
    }
    // BinaryIndexing: one model for like pairs, one for unlike pairs, any
    // number of types (here 3, from the particles' largest type).
    {
        CabanaPD::ForceModel like( CabanaPD::PMB{}, h, 1.0, 2.0 );
        CabanaPD::ForceModel unlike( CabanaPD::PMB{}, h, 4.0, 9.0 );
        auto models = CabanaPD::createMultiForceModel( particles, CabanaPD::BinaryIndexing{}, like, unlike );
        const std::vector<int> types = { 0, 2, 1, 2 };
        expect( models.numTypes( types ) == 3, "BinaryIndexing: types from the particles" );
        auto table = models.pairTable( 3 );
        bool ok = true;
        for ( int a = 0; a < 3; a++ )
            for ( int b = 0; b < 3; b++ )
                ok = ok && samePair( table[a * 3 + b], a == b ? like.pairModel() : unlike.pairModel() );
        expect( ok, "BinaryIndexing: like / unlike table" );
        expect( throwsOutOfRange( [&] {
                    auto short_list = CabanaPD::createMultiForceModel( particles, CabanaPD::BinaryIndexing{}, like );
                    (void)short_list.pairTable( 2 );
                } ),
                "an index past the model list is refused" );
    }
    // NoFracture pair models never break: +inf bond_break_coeff.
    {
        CabanaPD::ForceModel m0( CabanaPD::PMB{}, CabanaPD::NoFracture{}, h, 1.0 );
        expect( std::isinf( m0.pairModel().bond_break_coeff ), "NoFracture pair model never breaks" );
    }
    std::printf( "%d failure(s)\n", failures );
    return failures == 0 ? 0 : 1;
}
