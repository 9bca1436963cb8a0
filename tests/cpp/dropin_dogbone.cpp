// Dogbone tensile test through the drop-in C++ API: the reference's
// examples/mechanics/dogbone_tensile_test.cpp written against include/CabanaPD
// (PMB ElasticPerfectlyPlastic with fracture; grips held by a displacement
// functor with force_update = false; grip force time series).  One change:
// the example seeds particles randomly within their cells
// (Cabana::InitRandom); here they sit at the cell centres.  Writes the final
// displacement (caller order, double[n][3]) as raw binary to argv[2].
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
    const std::string outname = argc > 2 ? argv[2] : "dogbone_u.bin";

    double rho0 = inputs["density"];
    double E = inputs["elastic_modulus"];
    double nu = 0.25;
    double K = E / ( 3 * ( 1 - 2 * nu ) );
    double G0 = inputs["fracture_energy"];
    double sigma_y = inputs["yield_stress"];
    double horizon = inputs["horizon"];
    horizon += 1e-10;

    std::array<double, 3> low_corner = inputs["low_corner"];
    std::array<double, 3> high_corner = inputs["high_corner"];
    std::array<int, 3> num_cells = inputs["num_cells"];
    int m = std::floor( horizon / ( ( high_corner[0] - low_corner[0] ) / num_cells[0] ) );
    int halo_width = m + 1;

    using model_type = CabanaPD::PMB;
    using mechanics_type = CabanaPD::ElasticPerfectlyPlastic;

    double G = inputs["gage_length"];
    double D = inputs["distance_between_grips"];
    double W = inputs["width_narrow_section"];
    double R = inputs["fillet_radius"];
    double midx = 0.5 * ( low_corner[0] + high_corner[0] );
    double midy = 0.5 * ( low_corner[1] + high_corner[1] );

    auto init_op = KOKKOS_LAMBDA( const int, const double x[3] )
    {
        double Rsq = R * R;
        double xc_bl = midx - 0.5 * G;
        double yc_bl = midy - 0.5 * W - R;
        double xc_br = midx + 0.5 * G;
        double yc_br = yc_bl;
        double xc_tl = xc_bl;
        double yc_tl = midy + 0.5 * W + R;
        double xc_tr = xc_br;
        double yc_tr = yc_tl;
        if ( Kokkos::abs( x[0] - midx ) < 0.5 * G && Kokkos::abs( x[1] - midy ) > 0.5 * W )
            return false;
        else if ( Kokkos::abs( x[0] - xc_bl ) * Kokkos::abs( x[0] - xc_bl ) +
                      Kokkos::abs( x[1] - yc_bl ) * Kokkos::abs( x[1] - yc_bl ) <
                  Rsq )
            return false;
        else if ( Kokkos::abs( x[0] - xc_br ) * Kokkos::abs( x[0] - xc_br ) +
                      Kokkos::abs( x[1] - yc_br ) * Kokkos::abs( x[1] - yc_br ) <
                  Rsq )
            return false;
        else if ( Kokkos::abs( x[0] - xc_tl ) * Kokkos::abs( x[0] - xc_tl ) +
                      Kokkos::abs( x[1] - yc_tl ) * Kokkos::abs( x[1] - yc_tl ) <
                  Rsq )
            return false;
        else if ( Kokkos::abs( x[0] - xc_tr ) * Kokkos::abs( x[0] - xc_tr ) +
                      Kokkos::abs( x[1] - yc_tr ) * Kokkos::abs( x[1] - yc_tr ) <
                  Rsq )
            return false;
        return true;
    };

    CabanaPD::Particles particles( memory_space{}, model_type{} );
    particles.domain( low_corner, high_corner, num_cells, halo_width );
    particles.create( exec_space{}, init_op );

    auto rho = particles.sliceDensity();
    auto init_functor = KOKKOS_LAMBDA( const int pid ) { rho( pid ) = rho0; };
    particles.update( exec_space{}, init_functor );

    CabanaPD::ForceModel force_model( model_type{}, mechanics_type{}, memory_space{}, horizon, K, G0, sigma_y );
    CabanaPD::Solver solver( inputs, particles, force_model );

    double v0 = inputs["grip_velocity"];
    CabanaPD::Region<CabanaPD::RectangularPrism> right_grip( midx + 0.5 * D, high_corner[0], low_corner[1],
                                                             high_corner[1], low_corner[2], high_corner[2] );
    CabanaPD::Region<CabanaPD::RectangularPrism> left_grip( low_corner[0], midx - 0.5 * D, low_corner[1],
                                                            high_corner[1], low_corner[2], high_corner[2] );
    auto x = solver.particles.sliceReferencePosition();
    auto u = solver.particles.sliceDisplacement();
    auto disp_func = KOKKOS_LAMBDA( const int pid, const double t )
    {
        if ( right_grip.inside( x, pid ) )
        {
            u( pid, 0 ) = v0 * t;
            u( pid, 1 ) = 0.0;
            u( pid, 2 ) = 0.0;
        }
        else if ( left_grip.inside( x, pid ) )
        {
            u( pid, 0 ) = 0.0;
            u( pid, 1 ) = 0.0;
            u( pid, 2 ) = 0.0;
        }
    };
    auto bc = CabanaPD::createBoundaryCondition( disp_func, exec_space{}, solver.particles, false, left_grip,
                                                 right_grip );

    auto dx = solver.particles.dx[0];
    auto dy = solver.particles.dx[1];
    auto dz = solver.particles.dx[2];
    auto f = solver.particles.sliceForce();
    auto force_func_x = KOKKOS_LAMBDA( const int p, const double ) { return f( p, 0 ) * dx * dy * dz; };
    auto output_fx = CabanaPD::createOutputTimeSeries( "output_force_x.txt", inputs, exec_space{}, solver.particles,
                                                       force_func_x, right_grip );

    solver.init( bc );
    solver.run( bc, output_fx );

    u = solver.particles.sliceDisplacement();
    std::ofstream out( outname, std::ios::binary );
    out.write( reinterpret_cast<const char*>( &u( 0, 0 ) ), 3 * u.size() * sizeof( double ) );
    std::printf( "particles %zu\nDB_DONE\n", u.size() );
    Kokkos::finalize();
    return 0;
}
