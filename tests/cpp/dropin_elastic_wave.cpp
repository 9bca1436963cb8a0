// Elastic wave through the drop-in C++ API: the reference's
// examples/mechanics/elastic_wave.cpp written against include/CabanaPD,
// run on the deck its regression test uses (regression_test/inputs/
// elastic_wave.json).  It writes the x-displacement profile along the x axis
// with createDisplacementProfile; the caller compares that file with the
// polynomial fit of regression_test/tstElasticWave.hpp:54-93.
#include <cmath>
#include <cstdio>
#include <string>

#include <CabanaPD.hpp>

int main( int argc, char* argv[] )
{
    Kokkos::initialize( argc, argv );
    using exec_space = Kokkos::DefaultExecutionSpace;
    using memory_space = typename exec_space::memory_space;

    CabanaPD::Inputs inputs( argv[1] );
    std::string outname = argc > 2 ? argv[2] : "displacement_profile.txt";
    std::remove( outname.c_str() );

    double rho0 = inputs["density"];
    auto K = inputs["bulk_modulus"];
    double G = inputs["shear_modulus"];
    double horizon = inputs["horizon"];
    horizon += 1e-10;

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];
    std::array<int, 3> num_cells = inputs["num_cells"];
    int m = std::floor( horizon / ( ( high_corner[0] - low_corner[0] ) / num_cells[0] ) );
    int halo_width = m + 1;

    using model_type = CabanaPD::LinearLPS;
    CabanaPD::ForceModel force_model( model_type{}, CabanaPD::NoFracture{}, horizon, K, G );
    static_assert( std::is_same_v<decltype( force_model )::force_tag, CabanaPD::LPS>,
                   "ForceModel<LinearLPS> runs the LPS force, as in the reference" );

    CabanaPD::Particles particles( memory_space{}, model_type{} );
    particles.domain( low_corner, high_corner, num_cells, halo_width );
    particles.create( exec_space{} );

    auto rho = particles.sliceDensity();
    auto x = particles.sliceReferencePosition();
    auto u = particles.sliceDisplacement();
    auto v = particles.sliceVelocity();

    auto init_functor = KOKKOS_LAMBDA( const int pid )
    {
        rho( pid ) = rho0;
        double a = 0.001;
        double r0 = 0.25;
        double l = 0.07;
        double norm = std::sqrt( x( pid, 0 ) * x( pid, 0 ) + x( pid, 1 ) * x( pid, 1 ) + x( pid, 2 ) * x( pid, 2 ) );
        double diff = norm - r0;
        double arg = diff * diff / l / l;
        for ( int d = 0; d < 3; d++ )
        {
            double comp = 0.0;
            if ( norm > 0.0 )
                comp = x( pid, d ) / norm;
            u( pid, d ) = a * std::exp( -arg ) * comp;
            v( pid, d ) = 0.0;
        }
    };
    particles.update( exec_space{}, init_functor );

    CabanaPD::Solver solver( inputs, particles, force_model );
    solver.init();
    solver.run();

    CabanaPD::createDisplacementProfile( outname, solver.particles, 0 );
    std::printf( "steps %d\nEW_DONE\n", solver.num_steps );
    Kokkos::finalize();
    return 0;
}
