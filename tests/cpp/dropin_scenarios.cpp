// Scenario driver of the drop-in C++ API (include/CabanaPD): one program,
// one scenario per command line, each built only from the public drop-in
// surface -- Inputs decks, Particles::domain/create/update, the ForceModel
// constructors, Prenotch, createMultiForceModel, Region, force and functor
// boundary conditions (force_update true and false), BodyTerm, output time
// series and profiles, and Solver::init/run/runStep with PD and DEM models.
//
//   dropin_scenarios <scenario> <deck.json> <output>
//
//   notched-impact  PMB fracture plate, two pre-notches, an edge strip struck
//                   between them; prints the last crack-front value of a Max
//                   time series (the deck is the Kalthoff-Winkler one).
//   pulse           LinearLPS model (LPS force) on a cube with a radial
//                   Gaussian displacement shell; writes the x-displacement
//                   profile along the x axis to <output>.
//   inclusion       LPS fracture with a stiffer disc (two materials, averaged
//                   bonds), no-fail bands and a traction pair; prints the last
//                   crack-front value.
//   inclusion-indexed  the same with the averaged models built by hand and
//                   passed with FullIndexing<2> (explicit indexing).
//   grips           ElasticPerfectlyPlastic PMB dogbone held by displacement
//                   grips (force_update = false); writes u to <output>.
//   heated          TemperatureDependent PMB, a temperature ramp imposed by a
//                   BodyTerm (force = false); writes u and T.
//   conduction      DynamicTemperature PMB with two planes held cold; writes u
//                   and T.
//   collide         two PMB fracture blocks that meet only through a contact
//                   model (PD + NormalRepulsion Solver); writes u to <output>.
//   rebound         Hertz contact of two spheres (DEM-only Solver); prints the
//                   ratio of final to initial speed.
//   pulloff         Hertz-JKR contact of two spheres stepped with runStep;
//                   prints the most tensile force over the pull-off force.
//
// Results go to stdout as "key value" lines; the tests hold the tolerances.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include <CabanaPD.hpp>

namespace
{
using ExecSpace = Kokkos::DefaultExecutionSpace;
using MemSpace = ExecSpace::memory_space;
using Box = CabanaPD::Region<CabanaPD::RectangularPrism>;
using Vec3 = std::array<double, 3>;

struct Args
{
    std::string deck;
    std::string out;
};

// Bulk and shear moduli of an isotropic solid.
double bulkModulus( double E, double nu ) { return E / ( 3.0 * ( 1.0 - 2.0 * nu ) ); }
double shearModulus( double E, double nu ) { return E / ( 2.0 * ( 1.0 + nu ) ); }

// The decks' horizon, widened by 1e-10 so that neighbours at exactly the
// horizon distance stay inside under rounding (as the decks are meant).
double widenedHorizon( const CabanaPD::Inputs& deck )
{
    double h = deck["horizon"];
    return h + 1e-10;
}

// Ghost layers of a domain: the horizon in cells, plus one.
int ghostLayers( const CabanaPD::Inputs& deck, double horizon )
{
    Vec3 lo = deck["low_corner"];
    Vec3 hi = deck["high_corner"];
    std::array<int, 3> cells = deck["num_cells"];
    return static_cast<int>( std::floor( horizon / ( ( hi[0] - lo[0] ) / cells[0] ) ) ) + 1;
}

void print( const char* key, double value ) { std::printf( "%s %.17g\n", key, value ); }

void writeDoubles( const std::string& path, const double* data, std::size_t count )
{
    std::ofstream f( path, std::ios::binary );
    f.write( reinterpret_cast<const char*>( data ), static_cast<std::streamsize>( count * sizeof( double ) ) );
}

// Last non-empty line of a time-series file, as a number.
double lastValueIn( const std::string& path )
{
    std::ifstream f( path );
    std::string line, last;
    while ( std::getline( f, line ) )
        if ( !line.empty() )
            last = line;
    return last.empty() ? 0.0 : std::stod( last );
}

// ---------------------------------------------------------------------------
int notchedImpact( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const std::array<int, 3> cells = deck["num_cells"];
    const Vec3 size = deck["system_size"];
    const double horizon = widenedHorizon( deck );
    // the deck's material: steel with Poisson's ratio 1/3 for the bulk modulus
    const double K = bulkModulus( deck["elastic_modulus"], 1.0 / 3.0 );
    const double G0 = deck["fracture_energy"];
    const double rho0 = deck["density"];
    const double strike = deck["impactor_velocity"];

    // two notches, a quarter of the width apart, cut from the struck edge
    // through the whole thickness and half the plate's length
    const double yA = -size[1] / 8.0, yB = size[1] / 8.0;
    const Vec3 along = { size[0] / 2.0, 0.0, 0.0 };
    const Vec3 across = { 0.0, 0.0, size[2] };
    CabanaPD::Prenotch<2> notches( along, across, { Vec3{ lo[0], yA, lo[2] }, Vec3{ lo[0], yB, lo[2] } } );

    CabanaPD::ForceModel model( CabanaPD::PMB{}, horizon, K, G0 );
    CabanaPD::Particles plate( MemSpace{}, CabanaPD::PMB{} );
    plate.domain( lo, hi, cells, ghostLayers( deck, horizon ) );
    plate.create( ExecSpace{} );
    const double h = plate.dx[0];
    const double edge = -0.5 * size[0];
    auto x0 = plate.sliceReferencePosition();
    auto rho = plate.sliceDensity();
    auto vel = plate.sliceVelocity();
    plate.update( ExecSpace{}, [=]( const int p ) {
        rho( p ) = rho0;
        const bool between = x0( p, 1 ) > yA && x0( p, 1 ) < yB;
        if ( between && x0( p, 0 ) < edge + h )
            vel( p, 0 ) = strike;
    } );

    CabanaPD::Solver solver( deck, plate, model );
    // the struck strip carries no internal force
    Box strip( edge - h, edge + h, yA - 0.25 * h, yB + 0.25 * h, -size[2], size[2] );
    auto held = CabanaPD::createBoundaryCondition( CabanaPD::ForceValueBCTag{}, 0.0, ExecSpace{}, solver.particles,
                                                   strip );
    auto dmg = solver.particles.sliceDamage();
    auto xs = solver.particles.sliceReferencePosition();
    auto front = [=]( const int p, const double ) { return dmg( p ) > 0.3 ? xs( p, 1 ) : 0.0; };
    auto series = CabanaPD::createOutputTimeSeries<Kokkos::Max<double>>( a.out, deck, ExecSpace{},
                                                                         solver.particles, front, Box( lo, hi ) );
    solver.init( held, notches );
    solver.run( held, series );
    print( "steps", solver.num_steps );
    print( "crack_extent", lastValueIn( a.out ) );
    return 0;
}

// ---------------------------------------------------------------------------
int pulse( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const double horizon = widenedHorizon( deck );
    const double K = deck["bulk_modulus"];
    const double G = deck["shear_modulus"];
    const double rho0 = deck["density"];
    CabanaPD::ForceModel model( CabanaPD::LinearLPS{}, CabanaPD::NoFracture{}, horizon, K, G );
    static_assert( std::is_same_v<decltype( model )::force_tag, CabanaPD::LPS>,
                   "a LinearLPS model runs the LPS force" );

    CabanaPD::Particles cube( MemSpace{}, CabanaPD::LinearLPS{} );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const std::array<int, 3> cells = deck["num_cells"];
    cube.domain( lo, hi, cells, ghostLayers( deck, horizon ) );
    cube.create( ExecSpace{} );
    // a spherical shell of radial displacement: amplitude 1e-3, radius 0.25,
    // Gaussian width 0.07, at rest
    const double amp = 1e-3, shell = 0.25, width = 0.07;
    auto x0 = cube.sliceReferencePosition();
    auto u = cube.sliceDisplacement();
    auto vel = cube.sliceVelocity();
    auto rho = cube.sliceDensity();
    cube.update( ExecSpace{}, [=]( const int p ) {
        rho( p ) = rho0;
        const double r = std::sqrt( x0( p, 0 ) * x0( p, 0 ) + x0( p, 1 ) * x0( p, 1 ) + x0( p, 2 ) * x0( p, 2 ) );
        const double g = std::exp( -( r - shell ) * ( r - shell ) / width / width );
        for ( int d = 0; d < 3; d++ )
        {
            u( p, d ) = amp * g * ( r > 0.0 ? x0( p, d ) / r : 0.0 );
            vel( p, d ) = 0.0;
        }
    } );
    CabanaPD::Solver solver( deck, cube, model );
    solver.init();
    solver.run();
    CabanaPD::createDisplacementProfile( a.out, solver.particles, 0 );
    print( "steps", solver.num_steps );
    return 0;
}

// ---------------------------------------------------------------------------
int inclusionRun( const Args& a, bool indexed )
{
    CabanaPD::Inputs deck( a.deck );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const double horizon = widenedHorizon( deck );
    // material 0 is the plate, material 1 the disc
    struct Mat
    {
        double rho, K, G, G0;
    };
    Mat mat[2];
    for ( int k = 0; k < 2; k++ )
    {
        const double E = deck["elastic_modulus"][k];
        const double nu = deck["Poisson's_ratio"][k];
        const double rho = deck["density"][k];
        const double G0 = deck["fracture_energy"][k];
        mat[k] = { rho, bulkModulus( E, nu ), shearModulus( E, nu ), G0 };
    }
    CabanaPD::ForceModel plateModel( CabanaPD::LPS{}, horizon, mat[0].K, mat[0].G, mat[0].G0 );
    CabanaPD::ForceModel discModel( CabanaPD::LPS{}, horizon, mat[1].K, mat[1].G, mat[1].G0 );

    CabanaPD::Particles body( MemSpace{}, CabanaPD::LPS{} );
    body.domain( deck );
    body.create( ExecSpace{} );
    const double dy = body.dx[1];
    const double R = deck["inclusion_radius"];
    const double cx = deck["inclusion_center"][0];
    const double cy = deck["inclusion_center"][1];
    // the load is carried by the outer layers of the y faces, which never fail
    const double band_lo = lo[1] - dy + horizon + 1e-10;
    const double band_hi = hi[1] + dy - horizon - 1e-10;
    auto x0 = body.sliceReferencePosition();
    auto rho = body.sliceDensity();
    auto keep = body.sliceNoFail();
    auto kind = body.sliceType();
    body.update( ExecSpace{}, [=]( const int p ) {
        const double ex = x0( p, 0 ) - cx, ey = x0( p, 1 ) - cy;
        const int m = ex * ex + ey * ey < R * R ? 1 : 0;
        kind( p ) = m;
        rho( p ) = mat[m].rho;
        if ( x0( p, 1 ) <= band_lo || x0( p, 1 ) >= band_hi )
            keep( p ) = 1;
    } );
    // the same model table either through AverageTag or spelled out for
    // FullIndexing<2> with the averaged models built by hand
    auto run = [&]( auto models ) {
        CabanaPD::Solver solver( deck, body, models );

        Box bottom( lo[0], hi[0], lo[1] - dy, lo[1] + dy, lo[2], hi[2] );
        Box top( lo[0], hi[0], hi[1] - dy, hi[1] + dy, lo[2], hi[2] );
        const double sigma = deck["traction"];
        auto f = solver.particles.sliceForce();
        auto xs = solver.particles.sliceReferencePosition();
        // +-sigma / dy on the face layers, pulling the faces apart
        auto pull = [=]( const int p, const double ) { f( p, 1 ) += std::copysign( sigma / dy, xs( p, 1 ) ); };
        auto traction = CabanaPD::createBoundaryCondition( pull, ExecSpace{}, solver.particles, true, bottom, top );

        // one notch at mid height from the low-x face, half the length
        const double length = deck["system_size"][0];
        const double thick = deck["system_size"][2];
        const double ymid = 0.5 * ( lo[1] + hi[1] );
        CabanaPD::Prenotch<1> notch( Vec3{ length / 2.0, 0.0, 0.0 }, Vec3{ 0.0, 0.0, thick },
                                     { Vec3{ lo[0], ymid, lo[2] } } );
        auto dmg = solver.particles.sliceDamage();
        auto front = [=]( const int p, const double ) { return dmg( p ) > 0.3 ? xs( p, 1 ) : 0.0; };
        auto series = CabanaPD::createOutputTimeSeries<Kokkos::Max<double>>( a.out, deck, ExecSpace{},
                                                                             solver.particles, front, Box( lo, hi ) );
        solver.init( traction, notch );
        solver.run( traction, series );
        print( "steps", solver.num_steps );
        print( "crack_extent", lastValueIn( a.out ) );
        return 0;
    };
    if ( !indexed )
        return run( CabanaPD::createMultiForceModel( body, CabanaPD::AverageTag{}, plateModel, discModel ) );
    CabanaPD::ForceModel plateDisc( plateModel, discModel );
    CabanaPD::ForceModel discPlate( discModel, plateModel );
    return run( CabanaPD::createMultiForceModel( body, CabanaPD::FullIndexing<2>{}, plateModel, plateDisc, discPlate,
                                                 discModel ) );
}

// ---------------------------------------------------------------------------
// Dogbone outline in the x-y plane: a narrow gage section of the given width
// and length joined to the grip ends by four fillet arcs.
struct Dogbone
{
    double mx, my, gage, width, fillet;
    bool operator()( const int, const double x[3] ) const
    {
        const double ax = std::fabs( x[0] - mx ), ay = std::fabs( x[1] - my );
        if ( ax < 0.5 * gage && ay > 0.5 * width )
            return false;
        // fillet centres sit outside the gage edges, one per corner
        const double fx = ax - 0.5 * gage, fy = ay - ( 0.5 * width + fillet );
        return !( fx * fx + fy * fy < fillet * fillet );
    }
};

int grips( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const double horizon = widenedHorizon( deck );
    const double K = bulkModulus( deck["elastic_modulus"], 0.25 );
    const double rho0 = deck["density"];
    const double mx = 0.5 * ( lo[0] + hi[0] ), my = 0.5 * ( lo[1] + hi[1] );
    const double gage = deck["gage_length"];
    const double width = deck["width_narrow_section"];
    const double fillet = deck["fillet_radius"];
    const Dogbone shape{ mx, my, gage, width, fillet };

    CabanaPD::Particles bone( MemSpace{}, CabanaPD::PMB{} );
    const std::array<int, 3> cells = deck["num_cells"];
    bone.domain( lo, hi, cells, ghostLayers( deck, horizon ) );
    bone.create( ExecSpace{}, shape );
    auto rho = bone.sliceDensity();
    bone.update( ExecSpace{}, [=]( const int p ) { rho( p ) = rho0; } );
    const double G0 = deck["fracture_energy"];
    const double sigma_y = deck["yield_stress"];
    CabanaPD::ForceModel model( CabanaPD::PMB{}, CabanaPD::ElasticPerfectlyPlastic{}, MemSpace{}, horizon, K, G0,
                                sigma_y );
    CabanaPD::Solver solver( deck, bone, model );

    // the grips are everything beyond half the grip distance from the centre
    const double half = 0.5 * static_cast<double>( deck["distance_between_grips"] );
    Box fixed( lo[0], mx - half, lo[1], hi[1], lo[2], hi[2] );
    Box moving( mx + half, hi[0], lo[1], hi[1], lo[2], hi[2] );
    const double speed = deck["grip_velocity"];
    auto xs = solver.particles.sliceReferencePosition();
    auto u = solver.particles.sliceDisplacement();
    auto hold = [=]( const int p, const double t ) {
        const double ux = moving.inside( xs, p ) ? speed * t : 0.0;
        if ( moving.inside( xs, p ) || fixed.inside( xs, p ) )
        {
            u( p, 0 ) = ux;
            u( p, 1 ) = 0.0;
            u( p, 2 ) = 0.0;
        }
    };
    auto bc = CabanaPD::createBoundaryCondition( hold, ExecSpace{}, solver.particles, false, fixed, moving );
    // reaction of the moving grip: sum of f_x times the cell volume
    const double cellVol = solver.particles.dx[0] * solver.particles.dx[1] * solver.particles.dx[2];
    auto f = solver.particles.sliceForce();
    auto reaction = CabanaPD::createOutputTimeSeries( a.out + ".force", deck, ExecSpace{}, solver.particles,
                                                      [=]( const int p, const double ) { return f( p, 0 ) * cellVol; },
                                                      moving );
    solver.init( bc );
    solver.run( bc, reaction );
    u = solver.particles.sliceDisplacement();
    writeDoubles( a.out, &u( 0, 0 ), 3 * u.size() );
    print( "particles", static_cast<double>( u.size() ) );
    print( "last_grip_force", lastValueIn( a.out + ".force" ) );
    return 0;
}

// ---------------------------------------------------------------------------
template <class ThermalTag>
CabanaPD::Particles<MemSpace, CabanaPD::PMB, ThermalTag> heatedBlock( const CabanaPD::Inputs& deck, double horizon )
{
    CabanaPD::Particles block( MemSpace{}, CabanaPD::PMB{}, ThermalTag{} );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const std::array<int, 3> cells = deck["num_cells"];
    block.domain( lo, hi, cells, ghostLayers( deck, horizon ) );
    block.create( ExecSpace{} );
    return block;
}

template <class SolverT>
void writeState( const std::string& out, SolverT& solver )
{
    auto u = solver.particles.sliceDisplacement();
    auto T = solver.particles.sliceTemperature();
    writeDoubles( out, &u( 0, 0 ), 3 * u.size() );
    writeDoubles( out + ".temp", &T( 0 ), T.size() );
    print( "particles", static_cast<double>( u.size() ) );
}

int heated( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const double horizon = widenedHorizon( deck );
    const double K = bulkModulus( deck["elastic_modulus"], 0.25 );
    const double rho0 = deck["density"];
    const double T0 = deck["reference_temperature"];
    const Vec3 lo = deck["low_corner"];
    auto block = heatedBlock<CabanaPD::TemperatureDependent>( deck, horizon );
    auto rho = block.sliceDensity();
    block.update( ExecSpace{}, [=]( const int p ) { rho( p ) = rho0; } );
    const double alpha = deck["thermal_expansion_coeff"];
    CabanaPD::ForceModel model( CabanaPD::PMB{}, CabanaPD::NoFracture{}, horizon, K, block.sliceTemperature(), alpha,
                                T0 );
    CabanaPD::Solver solver( deck, block, model );
    // temperature rising in time, linearly in y from the low face: 5000 K/(m s)
    auto T = solver.particles.sliceTemperature();
    auto xs = solver.particles.sliceReferencePosition();
    const double y0 = lo[1];
    CabanaPD::BodyTerm ramp( [=]( const int p, const double t ) { T( p ) = T0 + 5000.0 * ( xs( p, 1 ) - y0 ) * t; },
                             solver.particles.size(), false );
    solver.init( ramp );
    solver.run( ramp );
    writeState( a.out, solver );
    return 0;
}

int conduction( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const double horizon = widenedHorizon( deck );
    const double K = bulkModulus( deck["elastic_modulus"], 0.25 );
    const double rho0 = deck["density"];
    const double T0 = deck["reference_temperature"];
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    auto block = heatedBlock<CabanaPD::DynamicTemperature>( deck, horizon );
    auto rho = block.sliceDensity();
    auto T = block.sliceTemperature();
    block.update( ExecSpace{}, [=]( const int p ) {
        rho( p ) = rho0;
        T( p ) = T0;
    } );
    const double kappa = deck["thermal_conductivity"];
    const double cp = deck["specific_heat_capacity"];
    const double alpha = deck["thermal_expansion_coeff"];
    CabanaPD::ForceModel model( CabanaPD::PMB{}, CabanaPD::NoFracture{}, horizon, K, T, kappa, cp, alpha, T0 );
    CabanaPD::Solver solver( deck, block, model );
    // both y faces held at zero temperature
    const double dy = solver.particles.dx[1];
    Box lowFace( lo[0], hi[0], lo[1] - dy, lo[1] + dy, lo[2], hi[2] );
    Box highFace( lo[0], hi[0], hi[1] - dy, hi[1] + dy, lo[2], hi[2] );
    auto Ts = solver.particles.sliceTemperature();
    auto cold = CabanaPD::createBoundaryCondition( [=]( const int p, const double ) { Ts( p ) = 0.0; }, ExecSpace{},
                                                   solver.particles, false, highFace, lowFace );
    solver.init( cold );
    solver.run( cold );
    CabanaPD::createOutputProfile( a.out + ".profile", solver.particles, 1,
                                   [=]( const int p ) { return Ts( p ); } );
    writeState( a.out, solver );
    return 0;
}

// ---------------------------------------------------------------------------
// Positions and volumes of a custom particle set (Particles::create(x, vol)).
struct PointSet
{
    std::vector<double> xyz, vol;
    double operator()( const int p, const int d ) const { return xyz[3 * p + d]; }
    double operator()( const int p ) const { return vol[p]; }
    std::size_t size() const { return vol.size(); }
};

// One particle at the centre of each cell of the box [lo, hi] split into n cells.
void addBlock( PointSet& ps, const Vec3& lo, const Vec3& hi, const std::array<int, 3>& n )
{
    const Vec3 d = { ( hi[0] - lo[0] ) / n[0], ( hi[1] - lo[1] ) / n[1], ( hi[2] - lo[2] ) / n[2] };
    for ( int k = 0; k < n[2]; k++ )
        for ( int j = 0; j < n[1]; j++ )
            for ( int i = 0; i < n[0]; i++ )
            {
                ps.xyz.push_back( lo[0] + i * d[0] + 0.5 * d[0] );
                ps.xyz.push_back( lo[1] + j * d[1] + 0.5 * d[1] );
                ps.xyz.push_back( lo[2] + k * d[2] + 0.5 * d[2] );
                ps.vol.push_back( d[0] * d[1] * d[2] );
            }
}

int collide( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    const Vec3 dxv = deck["dx"];
    const double dx = dxv[0];
    const std::array<int, 3> cells = deck["block_cells"];
    const double gap = deck["gap_cells"];
    const double speed = deck["impact_speed"];
    const double K = deck["bulk_modulus"];
    const double rho0 = deck["density"];
    const double sc = deck["critical_stretch"];
    const double horizon = deck["horizon"];
    const double cfac = deck["contact_horizon_factor"];
    const double efac = deck["contact_horizon_extend_factor"];
    // G0 of the PMB relation s_c = sqrt(5 G0 / (9 K delta))
    const double G0 = 9.0 * K * horizon * ( sc * sc ) / 5.0;
    CabanaPD::ForceModel model( CabanaPD::PMB{}, horizon, K, G0 );
    CabanaPD::NormalRepulsionModel contact( horizon, cfac * dx / 2.0, efac * dx, K );

    // block A on the left, block B on the right, gap cell spacings apart
    const double lx = cells[0] * dx;
    PointSet ps;
    addBlock( ps, { -lx - 0.5 * gap * dx, 0.0, 0.0 }, { -0.5 * gap * dx, cells[1] * dx, cells[2] * dx }, cells );
    const std::size_t na = ps.size();
    addBlock( ps, { 0.5 * gap * dx, 0.0, 0.0 }, { lx + 0.5 * gap * dx, cells[1] * dx, cells[2] * dx }, cells );
    CabanaPD::Particles blocks( MemSpace{}, CabanaPD::NormalRepulsionModel{} );
    blocks.domain( deck["low_corner"], deck["high_corner"], deck["num_cells"], ghostLayers( deck, horizon ) );
    blocks.create( ExecSpace{}, ps, ps );
    auto rho = blocks.sliceDensity();
    auto vel = blocks.sliceVelocity();
    blocks.update( ExecSpace{}, [=]( const int p ) {
        rho( p ) = rho0;
        vel( p, 0 ) = static_cast<std::size_t>( p ) < na ? speed : -speed;
    } );
    CabanaPD::Solver solver( deck, blocks, model, contact );
    solver.init();
    solver.run();
    auto u = solver.particles.sliceDisplacement();
    writeDoubles( a.out, &u( 0, 0 ), 3 * u.size() );
    print( "particles", static_cast<double>( u.size() ) );
    print( "steps", solver.num_steps );
    return 0;
}

// ---------------------------------------------------------------------------
// Two equal spheres on the x axis at +-gap, moving toward each other at 1 m/s.
struct TwoSpheres
{
    double gap, vol;
    double operator()( const int p, const int d ) const { return d == 0 ? ( p == 0 ? gap : -gap ) : 0.0; }
};
struct Volumes
{
    double vol;
    double operator()( const int ) const { return vol; }
    std::size_t size() const { return 2; }
};

template <class Model>
auto sphereSolver( const CabanaPD::Inputs& deck, const Model& model, double gap )
{
    const double rho0 = deck["density"];
    const double vol = deck["volume"];
    CabanaPD::Particles pair( MemSpace{}, Model{} );
    const Vec3 lo = deck["low_corner"];
    const Vec3 hi = deck["high_corner"];
    const std::array<int, 3> cells = deck["num_cells"];
    pair.domain( lo, hi, cells, 1 );
    pair.create( ExecSpace{}, TwoSpheres{ gap, vol }, Volumes{ vol } );
    auto rho = pair.sliceDensity();
    auto vel = pair.sliceVelocity();
    pair.update( ExecSpace{}, [=]( const int p ) {
        rho( p ) = rho0;
        vel( p, 0 ) = p == 0 ? -1.0 : 1.0;
    } );
    return CabanaPD::Solver( deck, pair, model );
}

template <class SolverT>
double kineticEnergy( SolverT& solver )
{
    auto v = solver.particles.sliceVelocity();
    auto rho = solver.particles.sliceDensity();
    auto vol = solver.particles.sliceVolume();
    double ke = 0.0;
    for ( std::size_t p = 0; p < v.size(); p++ )
        ke += 0.5 * rho( p ) * vol( p ) * ( v( p, 0 ) * v( p, 0 ) + v( p, 1 ) * v( p, 1 ) + v( p, 2 ) * v( p, 2 ) );
    return ke;
}

int rebound( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    CabanaPD::HertzianModel model( deck["radius"], deck["radius_extend"], deck["poisson_ratio"],
                                   deck["elastic_modulus"], deck["restitution"] );
    auto solver = sphereSolver( deck, model, 5.1e-5 );
    const double before = kineticEnergy( solver );
    solver.init();
    solver.run();
    print( "steps", solver.num_steps );
    print( "speed_ratio", std::sqrt( kineticEnergy( solver ) / before ) );
    return 0;
}

int pulloff( const Args& a )
{
    CabanaPD::Inputs deck( a.deck );
    CabanaPD::HertzianJKRModel model( deck["radius"], deck["radius_extend"], deck["poisson_ratio"],
                                      deck["elastic_modulus"], deck["restitution"], deck["surface_adhesion"], 0.0 );
    auto solver = sphereSolver( deck, model, 5.01e-5 );
    solver.init();
    auto f = solver.particles.sliceForce();
    const double vol = deck["volume"];
    double most_tensile = 0.0;
    for ( int step = 1; step <= solver.num_steps; step++ )
    {
        solver.runStep( step );
        most_tensile = std::min( most_tensile, f( 0, 0 ) * vol );
    }
    print( "steps", solver.num_steps );
    print( "pulloff_ratio", most_tensile / model.fc );
    return 0;
}
} // namespace

int inclusion( const Args& a )
{
    return inclusionRun( a, false );
}
int inclusionIndexed( const Args& a )
{
    return inclusionRun( a, true );
}

int main( int argc, char* argv[] )
{
    const std::map<std::string, std::function<int( const Args& )>> scenarios = {
        { "notched-impact", notchedImpact }, { "pulse", pulse },   { "inclusion", inclusion },
        { "inclusion-indexed", inclusionIndexed },
        { "grips", grips },                  { "heated", heated }, { "conduction", conduction },
        { "rebound", rebound },              { "pulloff", pulloff },    { "collide", collide } };
    if ( argc < 4 || !scenarios.count( argv[1] ) )
    {
        std::fprintf( stderr, "usage: %s <scenario> <deck.json> <output>\n", argv[0] );
        return 2;
    }
    Kokkos::initialize( argc, argv );
    int rc = 1;
    try
    {
        rc = scenarios.at( argv[1] )( Args{ argv[2], argv[3] } );
    }
    catch ( const std::exception& e )
    {
        std::fprintf( stderr, "error: %s\n", e.what() );
        rc = 1;
    }
    Kokkos::finalize();
    std::printf( "%s\n", rc == 0 ? "SCENARIO_DONE" : "SCENARIO_FAILED" );
    return rc;
}
