// Force/energy known answers through the drop-in API, in the shape of the
// reference's unit_test/tstForce.hpp: Particles on [-1,1]^3, u = s0 x,
// Force + Neighbor objects, computeForce / computeEnergy / computeStress,
// checks against the single-particle lattice sums (tstForce.hpp:91-155,
// 238-287, 345-422) on interior particles.
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <cstdio>

#include <CabanaPD.hpp>

namespace
{
int failures = 0;
void expect( bool c, const char* what, double a, double b )
{
    if ( !c )
    {
        std::printf( "FAIL %s: %.17g vs %.17g\n", what, a, b );
        failures++;
    }
}

template <class Model>
double latticeW( const Model& model, int m, double s0, bool lps, double K, double G )
{
    const double h = model.force_horizon, dx = h / m, vol = dx * dx * dx;
    const double c = 18.0 * K / ( M_PI * h * h * h * h );
    double wv = 0.0, nn = 0.0;
    for ( int i = -m; i <= m; i++ )
        for ( int j = -m; j <= m; j++ )
            for ( int k = -m; k <= m; k++ )
            {
                double xi = std::sqrt( double( i * i + j * j + k * k ) ) * dx;
                if ( xi > 0.0 && xi < h + 1e-14 )
                {
                    wv += ( 1.0 / xi ) * xi * xi * vol;
                    nn += 1.0;
                }
            }
    double theta = 0.0, W = 0.0;
    for ( int i = -m; i <= m; i++ )
        for ( int j = -m; j <= m; j++ )
            for ( int k = -m; k <= m; k++ )
            {
                double xi = std::sqrt( double( i * i + j * j + k * k ) ) * dx;
                if ( xi > 0.0 && xi < h + 1e-14 )
                    theta += 3.0 / wv * ( 1.0 / xi ) * s0 * xi * xi * vol;
            }
    for ( int i = -m; i <= m; i++ )
        for ( int j = -m; j <= m; j++ )
            for ( int k = -m; k <= m; k++ )
            {
                double xi = std::sqrt( double( i * i + j * j + k * k ) ) * dx;
                if ( !( xi > 0.0 && xi < h + 1e-14 ) )
                    continue;
                if ( lps )
                    W += 1.0 / nn * 0.5 * ( 3.0 * K - 5.0 * G ) / 3.0 * theta * theta +
                         0.5 * ( 15.0 * G / wv ) * ( 1.0 / xi ) * s0 * s0 * xi * xi * vol;
                else
                    W += 0.25 * c * s0 * s0 * xi * vol;
            }
    return W;
}

// Diagonal stress of the PMB lattice sum (tstForce.hpp:91-126); LPS gives
// exactly 3 K s0 on a full neighborhood.
double latticeStressPMB( double h, int m, double s0, double K )
{
    const double dx = h / m, vol = dx * dx * dx;
    const double c = 18.0 * K / ( M_PI * h * h * h * h );
    double sxx = 0.0;
    for ( int i = -m; i <= m; i++ )
        for ( int j = -m; j <= m; j++ )
            for ( int k = -m; k <= m; k++ )
            {
                double xi = std::sqrt( double( i * i + j * j + k * k ) ) * dx;
                if ( xi > 0.0 && xi < h + 1e-14 )
                    sxx += 0.5 * ( c * s0 * vol * ( i * dx ) / xi ) * ( i * dx );
            }
    return sxx;
}

template <class Model, class Tag>
void run( const Model& model, Tag, int nc, double bw, bool lps, double K, double G )
{
    using exec_space = Kokkos::DefaultExecutionSpace;
    using memory_space = exec_space::memory_space;
    const int m = 3;
    const double s0 = 0.1;
    CabanaPD::Particles particles( memory_space{}, Tag{}, CabanaPD::EnergyStressOutput{} );
    static_assert( CabanaPD::is_stress_output<typename decltype( particles )::output_type>::value );
    particles.domain( { -1.0, -1.0, -1.0 }, { 1.0, 1.0, 1.0 }, { nc, nc, nc }, 0 );
    particles.create( exec_space{} );
    auto x = particles.sliceReferencePosition();
    auto u = particles.sliceDisplacement();
    particles.update( exec_space{}, [=]( const int pid ) {
        for ( int d = 0; d < 3; d++ )
            u( pid, d ) = s0 * x( pid, d );
    } );
    CabanaPD::Force<memory_space, typename Model::force_tag, typename Model::fracture_type> force;
    CabanaPD::Neighbor<memory_space, typename Model::fracture_type> neighbor( model, particles );
    force.computeWeightedVolume( model, particles, neighbor );
    force.computeDilatation( model, particles, neighbor );
    CabanaPD::computeForce( model, force, particles, neighbor );
    CabanaPD::computeEnergy( model, force, particles, neighbor );
    CabanaPD::computeStress( model, force, particles, neighbor );

    auto f = particles.sliceForce();
    auto sigma = particles.sliceStress();
    auto W = particles.sliceStrainEnergy();
    auto vol = particles.sliceVolume();
    auto theta = particles.sliceDilatation();
    const double h = model.force_horizon;
    const double refW = latticeW( model, m, s0, lps, K, G );
    const double refS = lps ? 3.0 * K * s0 : latticeStressPMB( model.force_horizon, m, s0, K );
    double Phi = 0.0;
    int checked = 0;
    for ( std::size_t p = 0; p < particles.numLocal(); p++ )
    {
        bool in = true;
        for ( int d = 0; d < 3; d++ )
            in = in && x( p, d ) > -1.0 + h * bw && x( p, d ) < 1.0 - h * bw;
        if ( in )
        {
            for ( int d = 0; d < 3; d++ )
                expect( f( p, d ) <= 1e-13, "f ~ 0", f( p, d ), 0.0 );
            expect( std::fabs( W( p ) - refW ) <= 1e-6 * std::fabs( refW ), "W lattice sum", W( p ), refW );
            if ( lps )
                expect( std::fabs( theta( p ) - 3 * s0 ) <= 1e-6, "theta = 3 s0", theta( p ), 3 * s0 );
            for ( int a = 0; a < 3; a++ )
                for ( int b = 0; b < 3; b++ )
                {
                    if ( a == b )
                        expect( std::fabs( sigma( p, a, b ) - refS ) <= 1e-6 * std::fabs( refS ), "stress diagonal",
                                sigma( p, a, b ), refS );
                    else
                        expect( std::fabs( sigma( p, a, b ) ) <= 1e-13, "stress off-diagonal", sigma( p, a, b ),
                                0.0 );
                }
            checked++;
        }
        Phi += W( p ) * vol( p );
    }
    expect( std::fabs( Phi - force.totalStrainEnergy() ) <= 1e-5, "total strain energy", Phi,
            force.totalStrainEnergy() );
    expect( checked > 0, "particles checked", checked, 0 );
    std::printf( "%s: %d interior particles checked\n", lps ? "LPS" : "PMB", checked );

    // Particle field output (built with CABANAPD_ENABLE_VTK_OUTPUT): read the
    // file back and compare the current position, force and stress of the
    // last particle with the slices (legacy VTK stores big-endian doubles).
    particles.output( 7, 0.25, false );
    std::ifstream in( "particles_7.vtk", std::ios::binary );
    std::string text( ( std::istreambuf_iterator<char>( in ) ), std::istreambuf_iterator<char>() );
    auto readBe = [&]( std::size_t at ) {
        uint64_t b = 0;
        for ( int k = 0; k < 8; k++ )
            b = ( b << 8 ) | static_cast<unsigned char>( text[at + k] );
        double v;
        std::memcpy( &v, &b, 8 );
        return v;
    };
    const std::size_t n = particles.numLocal(), last = n - 1;
    auto fieldAt = [&]( const std::string& header ) {
        const std::size_t pos = text.find( header );
        return pos == std::string::npos ? std::string::npos : pos + header.size();
    };
    const std::size_t pp = fieldAt( "POINTS " + std::to_string( n ) + " double\n" );
    const std::size_t pf = fieldAt( "VECTORS force double\n" );
    const std::size_t ps = fieldAt( "TENSORS stress double\n" );
    const std::size_t pt = fieldAt( "TIME 1 1 double\n" );
    const bool found = pp != std::string::npos && pf != std::string::npos && ps != std::string::npos &&
                       pt != std::string::npos && text.find( lps ? "weighted_volume" : "strain_energy" ) != std::string::npos;
    expect( found, "vtk fields present", found ? 1.0 : 0.0, 1.0 );
    if ( found )
    {
        expect( readBe( pt ) == 0.25, "vtk time", readBe( pt ), 0.25 );
        expect( readBe( pp + 24 * last ) == x( last, 0 ) + u( last, 0 ), "vtk current position", readBe( pp + 24 * last ),
                x( last, 0 ) + u( last, 0 ) );
        expect( readBe( pf + 24 * last + 16 ) == f( last, 2 ), "vtk force", readBe( pf + 24 * last + 16 ), f( last, 2 ) );
        expect( readBe( ps + 72 * last + 8 ) == sigma( last, 0, 1 ), "vtk stress", readBe( ps + 72 * last + 8 ),
                sigma( last, 0, 1 ) );
    }
}
} // namespace

int main()
{
    const double K = 1.0, G = 0.5;
    {
        const int nc = 11;
        const double h = 3 * 2.0 / nc;
        CabanaPD::ForceModel model( CabanaPD::PMB{}, CabanaPD::Elastic{}, CabanaPD::NoFracture{}, h, K );
        run( model, CabanaPD::PMB{}, nc, 1.1, false, K, G );
        CabanaPD::ForceModel dmodel( CabanaPD::PMB{}, h, K, 1000.0 );
        run( dmodel, CabanaPD::PMB{}, nc, 1.1, false, K, G );
    }
    {
        const int nc = 15;
        const double h = 3 * 2.0 / nc;
        CabanaPD::ForceModel model( CabanaPD::LPS{}, CabanaPD::Elastic{}, CabanaPD::NoFracture{}, h, K, G, 1 );
        run( model, CabanaPD::LPS{}, nc, 2.1, true, K, G );
        CabanaPD::ForceModel dmodel( CabanaPD::LPS{}, h, K, G, 1000.0, 1 );
        run( dmodel, CabanaPD::LPS{}, nc, 2.1, true, K, G );
    }
    std::printf( "%s\n", failures ? "DROPIN_FORCE_FAIL" : "DROPIN_FORCE_OK" );
    return failures ? 1 : 0;
}
