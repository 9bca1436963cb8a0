// CabanaPD_Boundary.hpp -- regions, boundary conditions, body terms,
// pre-notches, line profiles and region time series of the drop-in API.
//
// Reference: src/CabanaPD_Geometry.hpp:24-168 (Region), :170-273
// (ParticleSteeringVector), src/CabanaPD_Boundary.hpp:26-204
// (BoundaryCondition, createBoundaryCondition), src/CabanaPD_BodyTerm.hpp:25-69,
// src/CabanaPD_Prenotch.hpp:174-281, src/CabanaPD_OutputProfiles.hpp:26-215.
#ifndef CABANAPD_DROPIN_BOUNDARY_HPP
#define CABANAPD_DROPIN_BOUNDARY_HPP

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <string>
#include <vector>

#include <pdk.h>

#include "CabanaPD_Input.hpp"
#include "CabanaPD_Particles.hpp"

namespace CabanaPD
{
struct RectangularPrism
{
};
struct Cylinder
{
};

template <typename UserFunctor>
struct Region
{
    UserFunctor _user_functor;
    Region( UserFunctor user ) : _user_functor( user ) {}
    template <class PositionType>
    bool inside( const PositionType& x, const int pid ) const
    {
        return _user_functor( x, pid );
    }
};

template <>
struct Region<RectangularPrism>
{
    Kokkos::Array<double, 3> low;
    Kokkos::Array<double, 3> high;
    Region( const double low_x, const double high_x, const double low_y, const double high_y, const double low_z,
            const double high_z )
        : low( { low_x, low_y, low_z } ), high( { high_x, high_y, high_z } )
    {
        assert( low[0] < high[0] && low[1] < high[1] && low[2] < high[2] );
    }
    template <class ArrayType>
    Region( const ArrayType _low, const ArrayType _high )
        : low( { _low[0], _low[1], _low[2] } ), high( { _high[0], _high[1], _high[2] } )
    {
    }
    template <class PositionType>
    bool inside( const PositionType& x, const int pid ) const
    {
        for ( int d = 0; d < 3; d++ )
            if ( !( x( pid, d ) >= low[d] && x( pid, d ) <= high[d] ) )
                return false;
        return true;
    }
};

template <>
struct Region<Cylinder>
{
    double radius_in, radius_out, low_z, high_z, x_center, y_center;
    Region( const double rin, const double rout, const double lz, const double hz, double xc = 0.0, double yc = 0.0 )
        : radius_in( rin ), radius_out( rout ), low_z( lz ), high_z( hz ), x_center( xc ), y_center( yc )
    {
        assert( radius_in < radius_out && low_z < high_z );
    }
    template <class PositionType>
    bool inside( const PositionType& x, const int pid ) const
    {
        double rsq = ( x( pid, 0 ) - x_center ) * ( x( pid, 0 ) - x_center ) +
                     ( x( pid, 1 ) - y_center ) * ( x( pid, 1 ) - y_center );
        return rsq >= radius_in * radius_in && rsq <= radius_out * radius_out && x( pid, 2 ) >= low_z &&
               x( pid, 2 ) <= high_z;
    }
};

struct Line
{
};

// Particles on the line through the origin along profile_dim: the other two
// reference coordinates strictly within half a cell of zero.
template <>
struct Region<Line>
{
    Kokkos::Array<double, 2> _half_dx;
    Kokkos::Array<int, 2> _dims;
    template <typename ArrayType>
    Region( const double profile_dim, const ArrayType dx )
    {
        const int p = static_cast<int>( profile_dim );
        _dims = { ( p + 1 ) % 3, ( p + 2 ) % 3 };
        _half_dx = { dx[_dims[0]] / 2.0, dx[_dims[1]] / 2.0 };
    }
    template <class PositionType>
    bool inside( const PositionType& x, const int pid ) const
    {
        return x( pid, _dims[0] ) < _half_dx[0] && x( pid, _dims[0] ) > -_half_dx[0] &&
               x( pid, _dims[1] ) < _half_dx[1] && x( pid, _dims[1] ) > -_half_dx[1];
    }
};

// Particle indices inside any of the regions (reference positions).
struct ParticleSteeringVector
{
    std::vector<int64_t> _view;
    std::size_t particle_count = 0;
    ParticleSteeringVector() = default;
    template <class ExecSpace, class ParticlesT, class... RegionType>
    ParticleSteeringVector( ExecSpace, const ParticlesT& particles, RegionType... regions )
        : particle_count( particles.referenceOffset() )
    {
        auto x = particles.sliceReferencePosition();
        ( add( x, particles.localOffset(), regions ), ... );
    }
    template <class X, class R>
    void add( const X& x, std::size_t n, const R& region )
    {
        for ( std::size_t p = 0; p < n; p++ )
            if ( region.inside( x, static_cast<int>( p ) ) )
                _view.push_back( static_cast<int64_t>( p ) );
    }
    std::size_t size() const { return _view.size(); }
    double time() const { return 0.0; }
};

struct ForceValueBCTag
{
};
struct ForceUpdateBCTag
{
};

// Force boundary condition on an index set: f = value or f += value.
template <class BCTag>
struct BoundaryCondition
{
    double _value;
    ParticleSteeringVector _indices;
    bool forceUpdate() const { return true; }
    double timeInit() const { return 0.0; }
    int32_t kind() const { return std::is_same_v<BCTag, ForceValueBCTag> ? PDK_BC_FORCE_VALUE : PDK_BC_FORCE_UPDATE; }
    BoundaryCondition( BCTag, const double value, ParticleSteeringVector idx ) : _value( value ), _indices( idx ) {}
};

template <class BCTag, class ExecSpace, class ParticlesT, class... RegionType>
auto createBoundaryCondition( BCTag tag, const double value, ExecSpace exec, const ParticlesT& particles,
                              RegionType... regions )
{
    return BoundaryCondition<BCTag>( tag, value, ParticleSteeringVector( exec, particles, regions... ) );
}

// Boundary condition with a user functor (pid, time) over the particles of
// the regions (CabanaPD_Boundary.hpp:26-73, 177-185).
// force_update = true: as for BodyTerm, the engine applies a constant
// per-particle force term after the force pass each step: the functor is
// evaluated once, at time 0, against zeroed force storage on the region's
// particles, so it must add a time-independent force (a traction such as
// f(pid, 1) += b0 * sign(y)).
// force_update = false (displacement or temperature conditions, applied after
// the initial half step and before the force pass): the Solver evaluates the
// functor on the host at every step's time and hands the displacements (all
// three components) or temperatures it wrote to the engine (pdk_set_prescribed);
// such functors may read reference positions and time, not the evolving state.
template <class UserFunctor>
struct FunctorBoundaryCondition
{
    UserFunctor _user_functor;
    ParticleSteeringVector _indices;
    bool _force_update;
    FunctorBoundaryCondition( ParticleSteeringVector idx, UserFunctor user, const bool force_update )
        : _user_functor( user ), _indices( idx ), _force_update( force_update )
    {
    }
    bool forceUpdate() const { return _force_update; }
    double timeInit() const { return 0.0; }
    // Particles the functor may write, and one application at time t.
    std::vector<int64_t> members( std::size_t ) const { return _indices._view; }
    void applyAt( double t ) const
    {
        for ( int64_t p : _indices._view )
            _user_functor( static_cast<int>( p ), t );
    }
    std::vector<double> evaluate( ParticleStorage& s ) const
    {
        if ( !_force_update )
            throw std::runtime_error( "a force_update = false functor is applied through the Solver" );
        std::vector<double> saved = s.f;
        std::fill( s.f.begin(), s.f.end(), 0.0 );
        for ( int64_t p : _indices._view )
            _user_functor( static_cast<int>( p ), 0.0 );
        std::vector<double> b = s.f;
        s.f = saved;
        return b;
    }
};

template <class UserFunctor, class ExecSpace, class ParticlesT, class... RegionType,
          std::enable_if_t<!std::is_arithmetic_v<ExecSpace>, int> = 0>
auto createBoundaryCondition( UserFunctor user, ExecSpace exec, const ParticlesT& particles, const bool force_update,
                              RegionType... regions )
{
    return FunctorBoundaryCondition<UserFunctor>( ParticleSteeringVector( exec, particles, regions... ), user,
                                                  force_update );
}

// BodyTerm with a user functor (pid, time).  force = true: the engine
// applies a constant per-particle body force density each step; the functor
// is evaluated once against zeroed force storage to obtain it, so it must add
// a time-independent term (gravity, constant traction).  force = false: an
// imposed field (such as the temperature of
// examples/thermomechanics/thermal_deformation.cpp) set before each force
// pass, as FunctorBoundaryCondition with force_update = false.
template <class UserFunctor>
struct BodyTerm
{
    UserFunctor _user_functor;
    std::size_t _particle_count;
    bool _force_update;
    bool _update_frozen;
    BodyTerm( UserFunctor user, const std::size_t particle_count, const bool force, const bool update_frozen = false )
        : _user_functor( user ), _particle_count( particle_count ), _force_update( force ),
          _update_frozen( update_frozen )
    {
    }
    bool forceUpdate() const { return _force_update; }
    double timeInit() const { return 0.0; }
    // Particles the functor may write (force = false: applied before the
    // force pass like a displacement condition), and one application at t.
    std::vector<int64_t> members( std::size_t ) const
    {
        std::vector<int64_t> all( _particle_count );
        for ( std::size_t p = 0; p < _particle_count; p++ )
            all[p] = static_cast<int64_t>( p );
        return all;
    }
    void applyAt( double t ) const
    {
        for ( std::size_t p = 0; p < _particle_count; p++ )
            _user_functor( static_cast<int>( p ), t );
    }
    // Evaluate the constant body force density into b[n][3].
    std::vector<double> evaluate( ParticleStorage& s ) const
    {
        std::vector<double> saved = s.f;
        std::fill( s.f.begin(), s.f.end(), 0.0 );
        for ( std::size_t p = 0; p < s.n; p++ )
            _user_functor( static_cast<int>( p ), 0.0 );
        std::vector<double> b = s.f;
        s.f = saved;
        return b;
    }
};

template <class T>
struct is_body_term : std::false_type
{
};
template <class F>
struct is_body_term<BodyTerm<F>> : std::true_type
{
};
template <class T>
struct is_boundary_condition : std::false_type
{
};
template <class F>
struct is_body_term<FunctorBoundaryCondition<F>> : std::true_type
{
};
template <class F>
struct is_boundary_condition<FunctorBoundaryCondition<F>> : std::true_type
{
};
template <class T>
struct is_boundary_condition<BoundaryCondition<T>> : std::true_type
{
};
template <class F>
struct is_boundary_condition<BodyTerm<F>> : std::true_type
{
};

// Prenotch<N> (CabanaPD_Prenotch.hpp:174-281)
template <std::size_t NumNotch, std::size_t NumVector = 1>
struct Prenotch
{
    static constexpr std::size_t num_notch = NumNotch;
    Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> _v1{};
    Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> _v2{};
    Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> _p0{};
    Prenotch() = default;
    Prenotch( Kokkos::Array<double, 3> v1, Kokkos::Array<double, 3> v2,
              Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> p0 )
        : _p0( p0 )
    {
        for ( std::size_t p = 0; p < NumNotch; p++ )
        {
            _v1[p] = v1;
            _v2[p] = v2;
        }
    }
    Prenotch( Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> v1, Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> v2,
              Kokkos::Array<Kokkos::Array<double, 3>, NumNotch> p0 )
        : _v1( v1 ), _v2( v2 ), _p0( p0 )
    {
    }
    // Flattened double[N][3] arrays for pdk_prenotch.
    void flatten( std::vector<double>& p0, std::vector<double>& v1, std::vector<double>& v2 ) const
    {
        for ( std::size_t p = 0; p < NumNotch; p++ )
            for ( int d = 0; d < 3; d++ )
            {
                p0.push_back( _p0[p][d] );
                v1.push_back( _v1[p][d] );
                v2.push_back( _v2[p][d] );
            }
    }
    double time() const { return 0.0; }
};

// Scalar time series over a region (CabanaPD_OutputProfiles.hpp:113-215):
// the functor is evaluated on host mirrors that the Solver refreshes at
// every output step.
template <class ReduceType, class FunctorType>
class OutputTimeSeries
{
  public:
    OutputTimeSeries( std::string name, const Inputs& inputs, ParticleSteeringVector indices, FunctorType output )
        : _indices( std::move( indices ) ), _file( std::move( name ) ), _output( output )
    {
        double time = inputs["final_time"];
        double dt = inputs["timestep"];
        double freq = inputs["output_frequency"];
        _profile.assign( static_cast<int>( time / dt / freq ) + 1, 0.0 );
    }
    void update( const double time )
    {
        constexpr bool is_max = std::is_same_v<ReduceType, Kokkos::Max<double>>;
        constexpr bool is_min = std::is_same_v<ReduceType, Kokkos::Min<double>>;
        double acc = is_max ? std::numeric_limits<double>::lowest()
                            : ( is_min ? std::numeric_limits<double>::max() : 0.0 );
        for ( int64_t p : _indices._view )
        {
            double v = _output( static_cast<int>( p ), time );
            if ( is_max )
                acc = std::max( acc, v );
            else if ( is_min )
                acc = std::min( acc, v );
            else
                acc += v;
        }
        if ( _index < _profile.size() )
            _profile[_index] = acc;
        _index++;
    }
    template <class Comm>
    void print( Comm ) const
    {
        std::fstream fout( _file, std::ios::app | std::ios::out );
        for ( std::size_t t = 0; t < std::min( _index, _profile.size() ); t++ )
            fout << std::fixed << std::setprecision( 15 ) << _profile[t] << "  "
                 << _profile[t] / static_cast<double>( std::max<std::size_t>( _indices.size(), 1 ) ) << std::endl;
    }
    const std::vector<double>& values() const { return _profile; }

  private:
    ParticleSteeringVector _indices;
    std::string _file;
    FunctorType _output;
    std::vector<double> _profile;
    std::size_t _index = 0;
};

template <class FunctorType, class ExecSpace, class ParticlesT, class GeometryType>
auto createOutputTimeSeries( std::string name, const Inputs& inputs, ExecSpace exec, const ParticlesT& particles,
                             FunctorType user, const GeometryType geom )
{
    return OutputTimeSeries<Kokkos::Sum<double>, FunctorType>( name, inputs,
                                                               ParticleSteeringVector( exec, particles, geom ), user );
}
template <class ReducerType, class FunctorType, class ExecSpace, class ParticlesT, class GeometryType>
auto createOutputTimeSeries( std::string name, const Inputs& inputs, ExecSpace exec, const ParticlesT& particles,
                             FunctorType user, const GeometryType geom )
{
    return OutputTimeSeries<ReducerType, FunctorType>( name, inputs, ParticleSteeringVector( exec, particles, geom ),
                                                       user );
}

// ---- profiles along a line (src/CabanaPD_OutputProfiles.hpp:26-93) ----------
// Appends "rank x value" per particle of Region<Line> to file_name, in
// particle order, with the stream's default precision as the reference does.
// The values are read from the host copy the solver refreshes after run().
template <typename ParticlesT, typename UserFunctor>
void createOutputProfile( std::string file_name, ParticlesT particles, const int profile_dim, UserFunctor user )
{
    Region<Line> line( profile_dim, particles.dx );
    ParticleSteeringVector sv( Kokkos::DefaultExecutionSpace{}, particles, line );
    auto x = particles.sliceReferencePosition();
    std::fstream fout( file_name, std::ios::app | std::ios::out );
    for ( std::size_t b = 0; b < sv.size(); b++ )
    {
        const int p = static_cast<int>( sv._view[b] );
        fout << 0 << " " << x( p, profile_dim ) << " " << user( p ) << std::endl;
    }
}

template <typename ParticlesT>
void createDisplacementProfile( std::string file_name, ParticlesT particles, const int profile_dim,
                                int displacement_dim = -1 )
{
    if ( displacement_dim == -1 )
        displacement_dim = profile_dim;
    auto u = particles.sliceDisplacement();
    createOutputProfile( file_name, particles, profile_dim,
                         [=]( const int pid ) { return u( pid, displacement_dim ); } );
}

template <typename ParticlesT>
void createDisplacementMagnitudeProfile( std::string file_name, ParticlesT particles, const int profile_dim )
{
    auto u = particles.sliceDisplacement();
    createOutputProfile( file_name, particles, profile_dim, [=]( const int pid ) {
        return std::sqrt( u( pid, 0 ) * u( pid, 0 ) + u( pid, 1 ) * u( pid, 1 ) + u( pid, 2 ) * u( pid, 2 ) );
    } );
}

} // namespace CabanaPD

#endif
