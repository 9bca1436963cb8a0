// Thermal deformation through the drop-in C++ API: the reference's
// examples/thermomechanics/thermal_deformation.cpp written against
// include/CabanaPD (PMB NoFracture TemperatureDependent; the temperature
// temp0 + 5000 (y - y_low) t imposed by a BodyTerm with force = false,
// applied before every force pass).  Writes the final displacement (caller
// order, double[n][3]) and temperature (double[n]) as raw binary to argv[2]
// and argv[2] + ".temp".
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
    const std::string outname = argc > 2 ? argv[2] : "thermal_u.bin";

    double rho0 = inputs["density"];
    double E = inputs["elastic_modulus"];
    double nu = 0.25;
    double K = E / ( 3 * ( 1 - 2 * nu ) );
    double horizon = inputs["horizon"];
    horizon += 1e-10;
    double alpha = inputs["thermal_expansion_coeff"];
    double temp0 = inputs["reference_temperature"];

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];
    std::array<int, 3> num_cells = inputs["num_cells"];
    int m = std::floor( horizon / ( ( high_corner[0] - low_corner[0] ) / num_cells[0] ) );
    int halo_width = m + 1;

    using model_type = CabanaPD::PMB;
    using thermal_type = CabanaPD::TemperatureDependent;
    CabanaPD::Particles particles( memory_space{}, model_type{}, thermal_type{} );
    particles.domain( low_corner, high_corner, num_cells, halo_width );
    particles.create( exec_space{} );

    auto rho = particles.sliceDensity();
    auto init_functor = KOKKOS_LAMBDA( const int pid ) { rho( pid ) = rho0; };
    particles.update( exec_space{}, init_functor );

    auto temp = particles.sliceTemperature();
    CabanaPD::ForceModel force_model( model_type{}, CabanaPD::NoFracture{}, horizon, K, temp, alpha, temp0 );
    CabanaPD::Solver solver( inputs, particles, force_model );

    auto x = solver.particles.sliceReferencePosition();
    temp = solver.particles.sliceTemperature();
    const double low_corner_y = low_corner[1];
    auto temp_func = KOKKOS_LAMBDA( const int pid, const double t )
    {
        temp( pid ) = temp0 + 5000.0 * ( x( pid, 1 ) - low_corner_y ) * t;
    };
    CabanaPD::BodyTerm body_term( temp_func, solver.particles.size(), false );

    solver.init( body_term );
    solver.run( body_term );

    CabanaPD::createDisplacementProfile( "ydisplacement_xaxis_profile.txt", solver.particles, 0, 1 );

    auto u = solver.particles.sliceDisplacement();
    std::ofstream out( outname, std::ios::binary );
    out.write( reinterpret_cast<const char*>( &u( 0, 0 ) ), 3 * u.size() * sizeof( double ) );
    std::ofstream tout( outname + ".temp", std::ios::binary );
    tout.write( reinterpret_cast<const char*>( &temp( 0 ) ), temp.size() * sizeof( double ) );
    std::printf( "particles %zu\nTD_DONE\n", u.size() );
    Kokkos::finalize();
    return 0;
}
