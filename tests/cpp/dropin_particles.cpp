// Particles::create through the drop-in headers (CPU only: no engine call):
// Cabana::InitUniform places one particle at each selected cell centre,
// Cabana::InitRandom one at a random point of each selected cell; volumes are
// the cell volume; the random placement repeats run to run.
#include <cmath>
#include <cstdio>
#include <vector>

#include <CabanaPD.hpp>

namespace
{
int failures = 0;
void expect( bool ok, const char* what )
{
    if ( !ok )
    {
        std::printf( "FAIL: %s\n", what );
        failures++;
    }
}
} // namespace

int main()
{
    using exec_space = Kokkos::DefaultExecutionSpace;
    using memory_space = exec_space::memory_space;
    const std::array<double, 3> lo = { 0.0, -1.0, 0.5 }, hi = { 1.0, 1.0, 2.0 };
    const std::array<int, 3> nc = { 5, 4, 3 };
    const double dx = 0.2, dy = 0.5, dz = 0.5;

    CabanaPD::Particles uni( memory_space{}, CabanaPD::PMB{} );
    uni.domain( lo, hi, nc, 0 );
    uni.create( exec_space{}, Cabana::InitUniform{} );
    auto xu = uni.sliceReferencePosition();
    auto vu = uni.sliceVolume();
    expect( uni.numLocal() == 60, "uniform: one particle per cell" );
    expect( std::fabs( xu( 0, 0 ) - ( lo[0] + 0.5 * dx ) ) < 1e-15 && std::fabs( xu( 59, 2 ) - ( hi[2] - 0.5 * dz ) ) < 1e-15,
            "uniform: cell centres, x fastest" );
    expect( std::fabs( vu( 7 ) - dx * dy * dz ) < 1e-15, "uniform: cell volume" );

    auto left = []( const int, const double x[3] ) { return x[0] < 0.5; };
    std::vector<double> first;
    for ( int pass = 0; pass < 2; pass++ )
    {
        CabanaPD::Particles rnd( memory_space{}, CabanaPD::PMB{} );
        rnd.domain( lo, hi, nc, 0 );
        rnd.create( exec_space{}, Cabana::InitRandom{}, left );
        auto x = rnd.sliceReferencePosition();
        std::vector<int> seen( 60, 0 );
        bool inside = true, centred = true;
        std::vector<double> pos;
        for ( std::size_t p = 0; p < rnd.numLocal(); p++ )
        {
            const int i = static_cast<int>( std::floor( ( x( p, 0 ) - lo[0] ) / dx ) );
            const int j = static_cast<int>( std::floor( ( x( p, 1 ) - lo[1] ) / dy ) );
            const int k = static_cast<int>( std::floor( ( x( p, 2 ) - lo[2] ) / dz ) );
            inside = inside && i >= 0 && i < nc[0] && j >= 0 && j < nc[1] && k >= 0 && k < nc[2] && x( p, 0 ) < 0.5;
            if ( inside )
                seen[i + nc[0] * ( j + nc[1] * k )]++;
            centred = centred && std::fabs( x( p, 0 ) - ( lo[0] + ( i + 0.5 ) * dx ) ) < 1e-12;
            for ( int d = 0; d < 3; d++ )
                pos.push_back( x( p, d ) );
        }
        int multi = 0;
        for ( int c : seen )
            multi += c > 1 ? 1 : 0;
        expect( rnd.numLocal() > 0 && inside, "random: each particle in the domain, its cell and the selection" );
        expect( multi == 0, "random: at most one particle per cell" );
        expect( !centred, "random: not at the cell centres" );
        if ( pass == 0 )
            first = pos;
        else
            expect( pos == first, "random: the same points run to run" );
    }

    CabanaPD::Particles fn( memory_space{}, CabanaPD::PMB{} );
    fn.domain( lo, hi, nc, 0 );
    fn.create( exec_space{}, []( const int, const double x[3] ) { return x[0] < 0.45; } );
    expect( fn.numLocal() == 24, "functor: the cells with centre x < 0.45 (2 of 5 columns)" );
    std::printf( "%s\n", failures ? "PARTICLES_FAIL" : "PARTICLES_OK" );
    return failures ? 1 : 0;
}
