// Kalthoff-Winkler through the drop-in C++ API (CabanaPD::Solver, ForceModel,
// Prenotch, createBoundaryCondition, createOutputTimeSeries).  The driver
// follows the reference's example program line for line where the API is
// the same; it prints the last value of the crack-extent time series, which
// the reference's regression test compares with 7.42e-2 +- 1.1e-4
// (regression_test/tstKalthoffWinkler.hpp:53).
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
    std::string outname = argc > 2 ? argv[2] : "output_kw_crack_y.txt";
    std::remove( outname.c_str() );

    double rho0 = inputs["density"];
    double E = inputs["elastic_modulus"];
    double nu = 1.0 / 3.0;
    double K = E / ( 3.0 * ( 1.0 - 2.0 * nu ) );
    double G0 = inputs["fracture_energy"];
    double horizon = inputs["horizon"];
    horizon += 1e-10;

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];
    std::array<int, 3> num_cells = inputs["num_cells"];
    int m = std::floor( horizon / ( ( high_corner[0] - low_corner[0] ) / num_cells[0] ) );
    int halo_width = m + 1;

    std::array<double, 3> system_size = inputs["system_size"];
    double height = system_size[0];
    double width = system_size[1];
    double thickness = system_size[2];
    double L_prenotch = height / 2.0;
    double y_prenotch1 = -width / 8.0;
    double y_prenotch2 = width / 8.0;
    double low_x = low_corner[0];
    double low_z = low_corner[2];
    Kokkos::Array<double, 3> p01 = { low_x, y_prenotch1, low_z };
    Kokkos::Array<double, 3> p02 = { low_x, y_prenotch2, low_z };
    Kokkos::Array<double, 3> v1 = { L_prenotch, 0, 0 };
    Kokkos::Array<double, 3> v2 = { 0, 0, thickness };
    Kokkos::Array<Kokkos::Array<double, 3>, 2> notch_positions = { p01, p02 };
    CabanaPD::Prenotch<2> prenotch( v1, v2, notch_positions );

    using model_type = CabanaPD::PMB;
    CabanaPD::ForceModel force_model( model_type{}, horizon, K, G0 );

    CabanaPD::Particles particles( memory_space{}, model_type{} );
    particles.domain( low_corner, high_corner, num_cells, halo_width );
    particles.create( exec_space{} );

    auto rho = particles.sliceDensity();
    auto x = particles.sliceReferencePosition();
    auto v = particles.sliceVelocity();
    auto f = particles.sliceForce();

    double dx = particles.dx[0];
    double v0 = inputs["impactor_velocity"];
    auto init_functor = KOKKOS_LAMBDA( const int pid )
    {
        rho( pid ) = rho0;
        if ( x( pid, 1 ) > y_prenotch1 && x( pid, 1 ) < y_prenotch2 && x( pid, 0 ) < -0.5 * height + dx )
            v( pid, 0 ) = v0;
    };
    particles.update( exec_space{}, init_functor );

    CabanaPD::Solver solver( inputs, particles, force_model );

    double x_bc = -0.5 * height;
    CabanaPD::Region<CabanaPD::RectangularPrism> plane( x_bc - dx, x_bc + dx, y_prenotch1 - 0.25 * dx,
                                                        y_prenotch2 + 0.25 * dx, -thickness, thickness );
    auto bc = createBoundaryCondition( CabanaPD::ForceValueBCTag{}, 0.0, exec_space{}, solver.particles, plane );

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

    // regression_test/tstKalthoffWinkler.hpp:39-53
    std::ifstream outfile( outname );
    std::string line;
    double max_crack_y = 0.0;
    while ( std::getline( outfile, line ) )
        if ( !line.empty() )
            max_crack_y = std::stod( line );
    std::printf( "max_crack_y %.15g\n", max_crack_y );
    const bool ok = std::fabs( max_crack_y - 7.42e-2 ) <= 1.1e-4;
    std::printf( "%s\n", ok ? "KW_REGRESSION_OK" : "KW_REGRESSION_FAIL" );
    Kokkos::finalize();
    return ok ? 0 : 1;
}
