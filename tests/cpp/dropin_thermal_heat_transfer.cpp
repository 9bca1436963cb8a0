// Thermal deformation with heat transfer through the drop-in C++ API: the
// reference's examples/thermomechanics/thermal_deformation_heat_transfer.cpp
// written against include/CabanaPD (PMB NoFracture DynamicTemperature; the
// temperature starts at temp0 and is held at 0 on two y planes by a
// force_update = false functor; heat conduction every thermal_subcycle_steps
// steps).  Writes the final displacement (caller order, double[n][3]) and
// temperature (double[n]) as raw binary to argv[2] and argv[2] + ".temp".
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
    const std::string outname = argc > 2 ? argv[2] : "heat_u.bin";

    double rho0 = inputs["density"];
    double E = inputs["elastic_modulus"];
    double nu = 0.25;
    double K = E / ( 3 * ( 1 - 2 * nu ) );
    double horizon = inputs["horizon"];
    horizon += 1e-10;
    double alpha = inputs["thermal_expansion_coeff"];
    double kappa = inputs["thermal_conductivity"];
    double cp = inputs["specific_heat_capacity"];
    double temp0 = inputs["reference_temperature"];

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];
    std::array<int, 3> num_cells = inputs["num_cells"];
    int m = std::floor( horizon / ( ( high_corner[0] - low_corner[0] ) / num_cells[0] ) );
    int halo_width = m + 1;

    using model_type = CabanaPD::PMB;
    using thermal_type = CabanaPD::DynamicTemperature;
    CabanaPD::Particles particles( memory_space{}, model_type{}, thermal_type{} );
    particles.domain( low_corner, high_corner, num_cells, halo_width );
    particles.create( exec_space{} );

    auto rho = particles.sliceDensity();
    auto temp = particles.sliceTemperature();
    auto init_functor = KOKKOS_LAMBDA( const int pid )
    {
        rho( pid ) = rho0;
        temp( pid ) = temp0;
    };
    particles.update( exec_space{}, init_functor );

    CabanaPD::ForceModel force_model( model_type{}, CabanaPD::NoFracture{}, horizon, K, temp, kappa, cp, alpha,
                                      temp0 );
    CabanaPD::Solver solver( inputs, particles, force_model );

    double dy = solver.particles.dx[1];
    using plane_type = CabanaPD::Region<CabanaPD::RectangularPrism>;
    plane_type plane1( low_corner[0], high_corner[0], high_corner[1] - dy, high_corner[1] + dy, low_corner[2],
                       high_corner[2] );
    plane_type plane2( low_corner[0], high_corner[0], low_corner[1] - dy, low_corner[1] + dy, low_corner[2],
                       high_corner[2] );
    temp = solver.particles.sliceTemperature();
    auto temp_bc = KOKKOS_LAMBDA( const int pid, const double ) { temp( pid ) = 0.0; };
    auto bc = CabanaPD::createBoundaryCondition( temp_bc, exec_space{}, solver.particles, false, plane1, plane2 );

    solver.init( bc );
    solver.run( bc );

    int profile_dim = 1;
    auto value = KOKKOS_LAMBDA( const int pid ) { return temp( pid ); };
    CabanaPD::createOutputProfile( "temperature_yaxis_profile.txt", solver.particles, profile_dim, value );

    auto u = solver.particles.sliceDisplacement();
    std::ofstream out( outname, std::ios::binary );
    out.write( reinterpret_cast<const char*>( &u( 0, 0 ) ), 3 * u.size() * sizeof( double ) );
    std::ofstream tout( outname + ".temp", std::ios::binary );
    tout.write( reinterpret_cast<const char*>( &temp( 0 ) ), temp.size() * sizeof( double ) );
    std::printf( "particles %zu\nHT_DONE\n", u.size() );
    Kokkos::finalize();
    return 0;
}
