// Crack-inclusion through the drop-in C++ API: the reference's
// examples/mechanics/crack_inclusion.cpp written against include/CabanaPD
// (two materials via createMultiForceModel with AverageTag, a functor force
// boundary condition, a no-fail band, one pre-notch), on the deck of
// regression_test/tstCrackInclusion.hpp.  It prints the last value of the
// crack-extent time series, which the reference's test compares with
// 5.75e-3 +- 5e-5 (regression_test/tstCrackInclusion.hpp:55).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include <CabanaPD.hpp>

int main( int argc, char* argv[] )
{
    Kokkos::initialize( argc, argv );
    using exec_space = Kokkos::DefaultExecutionSpace;
    using memory_space = typename exec_space::memory_space;

    CabanaPD::Inputs inputs( argv[1] );
    std::string outname = argc > 2 ? argv[2] : "output_incl_crack_y.txt";
    std::remove( outname.c_str() );

    // Plate
    double rho0 = inputs["density"][0];
    double E = inputs["elastic_modulus"][0];
    double nu = inputs["Poisson's_ratio"][0];
    double K = E / ( 3 * ( 1 - 2 * nu ) );
    double G = E / ( 2 * ( 1 + nu ) );
    double G0 = inputs["fracture_energy"][0];
    // Inclusion
    double rho0_I = inputs["density"][1];
    double E_I = inputs["elastic_modulus"][1];
    double nu_I = inputs["Poisson's_ratio"][1];
    double K_I = E_I / ( 3 * ( 1 - 2 * nu_I ) );
    double G_I = E_I / ( 2 * ( 1 + nu_I ) );
    double G0_I = inputs["fracture_energy"][1];

    double horizon = inputs["horizon"];
    horizon += 1e-10;

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];

    double height = inputs["system_size"][0];
    double thickness = inputs["system_size"][2];
    double L_prenotch = height / 2.0;
    double y_prenotch = 0.5 * ( low_corner[1] + high_corner[1] );
    Kokkos::Array<double, 3> p01 = { low_corner[0], y_prenotch, low_corner[2] };
    Kokkos::Array<double, 3> v1 = { L_prenotch, 0, 0 };
    Kokkos::Array<double, 3> v2 = { 0, 0, thickness };
    Kokkos::Array<Kokkos::Array<double, 3>, 1> notch_positions = { p01 };
    CabanaPD::Prenotch<1> prenotch( v1, v2, notch_positions );

    using model_type = CabanaPD::LPS;
    CabanaPD::ForceModel force_model_plate( model_type{}, horizon, K, G, G0 );
    CabanaPD::ForceModel force_model_inclusion( model_type{}, horizon, K_I, G_I, G0_I );

    CabanaPD::Particles particles( memory_space{}, model_type{} );
    particles.domain( inputs );
    particles.create( exec_space{} );

    double dy = particles.dx[1];
    CabanaPD::Region<CabanaPD::RectangularPrism> plane1( low_corner[0], high_corner[0], low_corner[1] - dy,
                                                         low_corner[1] + dy, low_corner[2], high_corner[2] );
    CabanaPD::Region<CabanaPD::RectangularPrism> plane2( low_corner[0], high_corner[0], high_corner[1] - dy,
                                                         high_corner[1] + dy, low_corner[2], high_corner[2] );

    auto rho = particles.sliceDensity();
    auto x = particles.sliceReferencePosition();
    auto v = particles.sliceVelocity();
    auto f = particles.sliceForce();
    auto nofail = particles.sliceNoFail();
    auto type = particles.sliceType();

    double R = inputs["inclusion_radius"];
    double xI = inputs["inclusion_center"][0];
    double yI = inputs["inclusion_center"][1];

    auto init_functor = KOKKOS_LAMBDA( const int pid )
    {
        rho( pid ) = rho0;
        if ( x( pid, 1 ) <= plane1.low[1] + horizon + 1e-10 || x( pid, 1 ) >= plane2.high[1] - horizon - 1e-10 )
            nofail( pid ) = 1;
        double rsq = ( x( pid, 0 ) - xI ) * ( x( pid, 0 ) - xI ) + ( x( pid, 1 ) - yI ) * ( x( pid, 1 ) - yI );
        if ( rsq < R * R )
        {
            type( pid ) = 1;
            rho( pid ) = rho0_I;
        }
        else
        {
            rho( pid ) = rho0;
        }
    };
    particles.update( exec_space{}, init_functor );

    auto models = CabanaPD::createMultiForceModel( particles, CabanaPD::AverageTag{}, force_model_plate,
                                                   force_model_inclusion );
    CabanaPD::Solver solver( inputs, particles, models );

    double sigma0 = inputs["traction"];
    double b0 = sigma0 / dy;
    f = solver.particles.sliceForce();
    x = solver.particles.sliceReferencePosition();
    auto bc_op = KOKKOS_LAMBDA( const int pid, const double )
    {
        auto ypos = x( pid, 1 );
        auto sign = std::abs( ypos ) / ypos;
        f( pid, 1 ) += b0 * sign;
    };
    auto bc = createBoundaryCondition( bc_op, exec_space{}, solver.particles, true, plane1, plane2 );

    CabanaPD::Region<CabanaPD::RectangularPrism> box( low_corner, high_corner );
    auto d = solver.particles.sliceDamage();
    auto crack_y_func = KOKKOS_LAMBDA( const int p, const double )
    {
        if ( d( p ) > 0.3 )
            return x( p, 1 );
        else
            return 0.0;
    };
    auto output_yl = CabanaPD::createOutputTimeSeries<Kokkos::Max<double>>( outname, inputs, exec_space{},
                                                                            solver.particles, crack_y_func, box );

    solver.init( bc, prenotch );
    solver.run( bc, output_yl );

    // regression_test/tstCrackInclusion.hpp:38-55
    std::ifstream outfile( outname );
    std::string line;
    double max_crack_y = 0.0;
    while ( std::getline( outfile, line ) )
        if ( !line.empty() )
            max_crack_y = std::stod( line );
    std::printf( "max_crack_y %.15g\n", max_crack_y );
    const bool ok = std::fabs( max_crack_y - 5.75e-3 ) <= 5e-5;
    std::printf( "%s\n", ok ? "CI_REGRESSION_OK" : "CI_REGRESSION_FAIL" );
    Kokkos::finalize();
    return ok ? 0 : 1;
}
