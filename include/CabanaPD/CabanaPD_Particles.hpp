// CabanaPD_Particles.hpp -- Particles of the drop-in API.
//
// Reference: src/CabanaPD_Particles.hpp:85-687 (fields, domain/create/update,
// offsets, slices), :689-822 (LPS m/theta), :1071-1231 (W/phi).  Storage is
// host-side and shared between copies (the reference's Solver keeps a shallow
// copy, CabanaPD_Solver.hpp:617); the engine holds the device state while a
// Solver or Neighbor is attached and writes results back into these arrays.
#ifndef CABANAPD_DROPIN_PARTICLES_HPP
#define CABANAPD_DROPIN_PARTICLES_HPP

#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <string>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "CabanaPD_Input.hpp"
#include "CabanaPD_Types.hpp"

namespace CabanaPD
{
// Non-owning view with the reference's slice indexing: s(i) or s(i, d).
template <class T, int Dim>
struct Slice
{
    T* data = nullptr;
    std::size_t n = 0;
    T& operator()( const std::size_t i ) const { return data[i]; }
    T& operator()( const std::size_t i, const int d ) const { return data[Dim * i + d]; }
    std::size_t size() const { return n; }
    std::size_t extent( int ) const { return n; }
};

// Stress view: s(i, a, b) over double[n][3][3].
struct StressSlice
{
    double* data = nullptr;
    std::size_t n = 0;
    double& operator()( const std::size_t i, const int a, const int b ) const { return data[9 * i + 3 * a + b]; }
    std::size_t size() const { return n; }
};

struct ParticleStorage
{
    std::size_t n = 0;
    std::vector<double> x, u, v, f, vol, rho, W, phi, m, theta, y, stress, temp;
    std::vector<int> nofail, type;
    void resize( std::size_t nn )
    {
        n = nn;
        for ( auto* a : { &x, &u, &v, &f, &y } )
            a->assign( 3 * nn, 0.0 );
        for ( auto* a : { &vol, &W, &phi, &m, &theta, &temp } )
            a->assign( nn, 0.0 ); // temperature zero at creation (CabanaPD_Particles.hpp:860-864)
        stress.assign( 9 * nn, 0.0 );
        rho.assign( nn, 1.0 );
        nofail.assign( nn, 0 );
        type.assign( nn, 0 );
    }
};

template <class MemorySpace, class ModelType, class ThermalType = TemperatureIndependent,
          class OutputType = EnergyOutput, int Dimension = 3>
class Particles
{
  public:
    using memory_space = MemorySpace;
    using execution_space = typename MemorySpace::execution_space;
    using model_type = ModelType;
    using thermal_type = ThermalType;
    using output_type = OutputType;
    static constexpr int dim = Dimension;

    std::array<double, 3> global_mesh_ext{};
    std::array<double, 3> local_mesh_ext{};
    std::array<double, 3> local_mesh_lo{};
    std::array<double, 3> local_mesh_hi{};
    double ghost_mesh_lo[3] = { 0, 0, 0 };
    double ghost_mesh_hi[3] = { 0, 0, 0 };
    Kokkos::Array<double, 3> dx{};
    int halo_width = 0;

    template <class... Tags>
    Particles( MemorySpace, ModelType, Tags... ) : _s( std::make_shared<ParticleStorage>() )
    {
    }

    // Particles::domain (CabanaPD_Particles.hpp:173-235)
    void domain( std::array<double, 3> low_corner, std::array<double, 3> high_corner,
                 const std::array<int, 3> num_cells, const int max_halo_width )
    {
        halo_width = max_halo_width;
        for ( int d = 0; d < 3; d++ )
        {
            _nc[d] = num_cells[d];
            dx[d] = ( high_corner[d] - low_corner[d] ) / num_cells[d];
            global_mesh_ext[d] = high_corner[d] - low_corner[d];
            local_mesh_lo[d] = low_corner[d];
            local_mesh_hi[d] = high_corner[d];
            local_mesh_ext[d] = global_mesh_ext[d];
            ghost_mesh_lo[d] = low_corner[d] - halo_width * dx[d];
            ghost_mesh_hi[d] = high_corner[d] + halo_width * dx[d];
        }
        _has_domain = true;
    }
    void domain( const Inputs& inputs )
    {
        std::array<double, 3> lo = inputs["low_corner"];
        std::array<double, 3> hi = inputs["high_corner"];
        std::array<int, 3> nc = inputs["num_cells"];
        double horizon = inputs["horizon"];
        int m = static_cast<int>( std::floor( horizon / ( ( hi[0] - lo[0] ) / nc[0] ) ) );
        domain( lo, hi, nc, m + 1 );
    }

    // Grid creation (CabanaPD_Particles.hpp:241-400): one particle per cell,
    // vol = cell volume, rho = 1.  Cabana::InitUniform (the default) places it
    // at the cell centre, Cabana::InitRandom at a uniformly random point of the
    // cell (a fixed-seed generator: the same points run to run, not the ones
    // of Cabana's random pool).  The user functor selects which cells get a
    // particle.
    template <class ExecSpace, class UserFunctor>
    void create( const ExecSpace& exec, Cabana::InitUniform, UserFunctor user_create, const std::size_t num_previous = 0,
                 const bool create_frozen = false )
    {
        createCells( exec, false, user_create, num_previous, create_frozen );
    }
    template <class ExecSpace, class UserFunctor>
    void create( const ExecSpace& exec, Cabana::InitRandom, UserFunctor user_create, const std::size_t num_previous = 0,
                 const bool create_frozen = false )
    {
        createCells( exec, true, user_create, num_previous, create_frozen );
    }
    template <class ExecSpace>
    void create( const ExecSpace& exec, Cabana::InitUniform, const std::size_t num_previous = 0,
                 const bool create_frozen = false )
    {
        createCells( exec, false, all(), num_previous, create_frozen );
    }
    template <class ExecSpace>
    void create( const ExecSpace& exec, Cabana::InitRandom, const std::size_t num_previous = 0,
                 const bool create_frozen = false )
    {
        createCells( exec, true, all(), num_previous, create_frozen );
    }
    template <class ExecSpace, class UserFunctor,
              std::enable_if_t<!std::is_integral_v<UserFunctor>, int> = 0>
    void create( const ExecSpace& exec, UserFunctor user_create, const std::size_t num_previous = 0,
                 const bool create_frozen = false )
    {
        createCells( exec, false, user_create, num_previous, create_frozen );
    }
    template <class ExecSpace>
    void create( const ExecSpace& exec, const std::size_t num_previous = 0, const bool create_frozen = false )
    {
        createCells( exec, false, all(), num_previous, create_frozen );
    }
    // Custom creation from positions and volumes (CabanaPD_Particles.hpp:310-368).
    template <class ExecSpace, class PositionType, class VolumeType,
              class = decltype( std::declval<PositionType>()( 0, 0 ) )>
    void create( const ExecSpace, const PositionType& x, const VolumeType& vol, const std::size_t num_previous = 0,
                 const bool = false )
    {
        const std::size_t n = vol.size();
        std::vector<double> px( 3 * n ), pv( n );
        for ( std::size_t i = 0; i < n; i++ )
        {
            for ( int d = 0; d < 3; d++ )
                px[3 * i + d] = x( i, d );
            pv[i] = vol( i );
        }
        appendParticles( px, pv, num_previous );
    }

    // Particles::update: run the user functor on every local particle (host).
    template <class ExecSpace, class FunctorType>
    void update( const ExecSpace, const FunctorType init_functor, const bool = false )
    {
        for ( std::size_t p = 0; p < _s->n; p++ )
            init_functor( static_cast<int>( p ) );
        _host_newer = true;
    }

    // Offsets (CabanaPD_Particles.hpp:445-455); frozen particles are not used.
    std::size_t numFrozen() const { return 0; }
    std::size_t frozenOffset() const { return 0; }
    std::size_t numLocal() const { return _s->n; }
    std::size_t localOffset() const { return _s->n; }
    std::size_t numGhost() const { return 0; }
    std::size_t referenceOffset() const { return _s->n; }
    std::size_t size() const { return _s->n; }
    unsigned long long numGlobal() const { return _s->n; }

    Slice<double, 3> sliceReferencePosition() const { return { _s->x.data(), _s->n }; }
    Slice<double, 3> sliceCurrentPosition() const
    {
        for ( std::size_t i = 0; i < 3 * _s->n; i++ )
            _s->y[i] = _s->x[i] + _s->u[i];
        return { _s->y.data(), _s->n };
    }
    Slice<double, 3> sliceDisplacement() const { return { _s->u.data(), _s->n }; }
    Slice<double, 3> sliceVelocity() const { return { _s->v.data(), _s->n }; }
    Slice<double, 3> sliceForce() const { return { _s->f.data(), _s->n }; }
    Slice<double, 3> sliceForceAtomic() const { return sliceForce(); }
    Slice<double, 1> sliceVolume() const { return { _s->vol.data(), _s->n }; }
    Slice<double, 1> sliceDensity() const { return { _s->rho.data(), _s->n }; }
    Slice<double, 1> sliceStrainEnergy() const { return { _s->W.data(), _s->n }; }
    Slice<double, 1> sliceDamage() const { return { _s->phi.data(), _s->n }; }
    Slice<double, 1> sliceWeightedVolume() const { return { _s->m.data(), _s->n }; }
    Slice<double, 1> sliceDilatation() const { return { _s->theta.data(), _s->n }; }
    // Particles<..., EnergyStressOutput> (CabanaPD_Particles.hpp:1233-1356)
    StressSlice sliceStress() const { return { _s->stress.data(), _s->n }; }
    // TemperatureDependent particles (CabanaPD_Particles.hpp:884-887)
    Slice<double, 1> sliceTemperature() const { return { _s->temp.data(), _s->n }; }
    Slice<int, 1> sliceNoFail() const { return { _s->nofail.data(), _s->n }; }
    Slice<int, 1> sliceType() const { return { _s->type.data(), _s->n }; }

    ParticleStorage& storage() const { return *_s; }
    std::size_t gridSize() const { return static_cast<std::size_t>( _nc[0] ) * _nc[1] * _nc[2]; }

    double timeInit() { return 0.0; }
    double time() { return 0.0; }
    double timeOutput() { return 0.0; }
    // Particles::output (CabanaPD_Particles.hpp:622-652 and the overrides at
    // :787-792, :926-931, :1194-1199, :1325-1330): position (reference or
    // current), force, displacement, velocity, type, plus the LPS weighted
    // volume and dilatation, the temperature, strain energy and damage, and
    // stress, as the particle type carries them.  The reference writes them
    // with Cabana's HDF5 or SILO particle output when Cabana is built with
    // one of those; HDF5 and SILO are not part of this build, which writes the
    // same fields to particles_<step>.vtk (legacy VTK, binary PolyData) when
    // CABANAPD_ENABLE_VTK_OUTPUT is defined and otherwise prints the
    // reference's "No particle output enabled."  The host copy is refreshed by
    // the Solver before each output.
    template <class... OtherFields>
    void output( const int output_step, const double output_time, const bool use_reference, OtherFields&&... )
    {
#ifdef CABANAPD_ENABLE_VTK_OUTPUT
        writeVtk( "particles_" + std::to_string( output_step ) + ".vtk", output_time, use_reference );
#else
        (void)output_step;
        (void)output_time;
        (void)use_reference;
        log( std::cout, "No particle output enabled." );
#endif
    }

    // Legacy VTK (big-endian binary) PolyData of the output fields.
    void writeVtk( const std::string& name, const double time, const bool use_reference ) const
    {
        const ParticleStorage& s = *_s;
        const std::size_t n = s.n;
        std::ofstream out( name, std::ios::binary );
        if ( !out )
            throw std::runtime_error( "cannot write " + name );
        auto be64 = [&]( double v ) {
            uint64_t b;
            std::memcpy( &b, &v, 8 );
            for ( int k = 7; k >= 0; k-- )
                out.put( static_cast<char>( ( b >> ( 8 * k ) ) & 0xff ) );
        };
        auto be32 = [&]( int32_t v ) {
            const uint32_t b = static_cast<uint32_t>( v );
            for ( int k = 3; k >= 0; k-- )
                out.put( static_cast<char>( ( b >> ( 8 * k ) ) & 0xff ) );
        };
        auto vec3 = [&]( const char* nm, const std::vector<double>& a ) {
            out << "VECTORS " << nm << " double\n";
            for ( std::size_t i = 0; i < 3 * n; i++ )
                be64( a[i] );
            out << "\n";
        };
        auto scalar = [&]( const char* nm, const std::vector<double>& a ) {
            out << "SCALARS " << nm << " double 1\nLOOKUP_TABLE default\n";
            for ( std::size_t i = 0; i < n; i++ )
                be64( a[i] );
            out << "\n";
        };
        out << "# vtk DataFile Version 3.0\nCabanaPD particles\nBINARY\nDATASET POLYDATA\n";
        out << "FIELD FieldData 1\nTIME 1 1 double\n";
        be64( time );
        out << "\nPOINTS " << n << " double\n";
        for ( std::size_t i = 0; i < 3 * n; i++ )
            be64( use_reference ? s.x[i] : s.x[i] + s.u[i] );
        out << "\nVERTICES " << n << " " << 2 * n << "\n";
        for ( std::size_t i = 0; i < n; i++ )
        {
            be32( 1 );
            be32( static_cast<int32_t>( i ) );
        }
        out << "\nPOINT_DATA " << n << "\n";
        vec3( "force", s.f );
        vec3( "displacement", s.u );
        vec3( "velocity", s.v );
        out << "SCALARS type int 1\nLOOKUP_TABLE default\n";
        for ( std::size_t i = 0; i < n; i++ )
            be32( s.type[i] );
        out << "\n";
        if constexpr ( std::is_same_v<typename ModelType::base_type, LPS> )
        {
            scalar( "weighted_volume", s.m );
            scalar( "dilatation", s.theta );
        }
        if constexpr ( is_temperature_dependent<ThermalType>::value )
            scalar( "temperature", s.temp );
        if constexpr ( is_energy_output<OutputType>::value )
        {
            scalar( "strain_energy", s.W );
            scalar( "damage", s.phi );
        }
        if constexpr ( is_stress_output<OutputType>::value )
        {
            out << "TENSORS stress double\n";
            for ( std::size_t i = 0; i < 9 * n; i++ )
                be64( s.stress[i] );
            out << "\n";
        }
    }

  private:
    static auto all()
    {
        return []( const int, const double* ) { return true; };
    }
    template <class ExecSpace, class UserFunctor>
    void createCells( const ExecSpace&, const bool random, UserFunctor user_create, const std::size_t num_previous,
                      const bool create_frozen )
    {
        if ( !_has_domain )
            throw std::runtime_error( "Cannot create() particles without calling domain() first." );
        if ( create_frozen )
            throw std::runtime_error( "frozen particles are not supported by this build" );
        std::vector<double> px;
        const double pv = dx[0] * dx[1] * dx[2];
        std::mt19937_64 gen( 0x5eed );
        std::uniform_real_distribution<double> unit( 0.0, 1.0 );
        for ( int k = 0; k < _nc[2]; k++ )
            for ( int j = 0; j < _nc[1]; j++ )
                for ( int i = 0; i < _nc[0]; i++ )
                {
                    const double o[3] = { random ? unit( gen ) : 0.5, random ? unit( gen ) : 0.5,
                                          random ? unit( gen ) : 0.5 };
                    double p[3] = { local_mesh_lo[0] + i * dx[0] + o[0] * dx[0],
                                    local_mesh_lo[1] + j * dx[1] + o[1] * dx[1],
                                    local_mesh_lo[2] + k * dx[2] + o[2] * dx[2] };
                    if ( user_create( static_cast<int>( px.size() / 3 ), p ) )
                        px.insert( px.end(), p, p + 3 );
                }
        appendParticles( px, std::vector<double>( px.size() / 3, pv ), num_previous );
    }

    void appendParticles( const std::vector<double>& px, const std::vector<double>& pv, std::size_t num_previous )
    {
        ParticleStorage& s = *_s;
        const std::size_t keep = std::min( num_previous, s.n );
        ParticleStorage old = s;
        s.resize( keep + pv.size() );
        for ( std::size_t i = 0; i < keep; i++ )
        {
            for ( int d = 0; d < 3; d++ )
            {
                s.x[3 * i + d] = old.x[3 * i + d];
                s.u[3 * i + d] = old.u[3 * i + d];
                s.v[3 * i + d] = old.v[3 * i + d];
            }
            s.vol[i] = old.vol[i];
            s.rho[i] = old.rho[i];
            s.temp[i] = old.temp[i];
            s.nofail[i] = old.nofail[i];
        }
        for ( std::size_t i = 0; i < pv.size(); i++ )
        {
            for ( int d = 0; d < 3; d++ )
                s.x[3 * ( keep + i ) + d] = px[3 * i + d];
            s.vol[keep + i] = pv[i];
        }
        _host_newer = true;
    }

    std::shared_ptr<ParticleStorage> _s;
    std::array<int, 3> _nc{ 0, 0, 0 };
    bool _has_domain = false;
    bool _host_newer = true;
};

namespace detail
{
// Output type named among the constructor tags (EnergyOutput when none is).
template <class... Tags>
struct OutputOf
{
    using type = EnergyOutput;
};
template <class T, class... Rest>
struct OutputOf<T, Rest...>
{
    using type = std::conditional_t<is_output<T>::value, T, typename OutputOf<Rest...>::type>;
};
// Thermal type named among the constructor tags (TemperatureIndependent when none is).
template <class... Tags>
struct ThermalOf
{
    using type = TemperatureIndependent;
};
template <class T, class... Rest>
struct ThermalOf<T, Rest...>
{
    using type = std::conditional_t<std::is_same_v<T, TemperatureDependent> || std::is_same_v<T, DynamicTemperature>, T,
                                    typename ThermalOf<Rest...>::type>;
};
} // namespace detail

// Deduction (CabanaPD_Particles.hpp:1364-1395): Particles( memory_space{},
// model{} [, thermal] [, output] ), energy output by default.
template <class MemorySpace, class ModelType, class... Tags>
Particles( MemorySpace, ModelType, Tags... )
    -> Particles<MemorySpace, ModelType, typename detail::ThermalOf<Tags...>::type,
                 typename detail::OutputOf<Tags...>::type>;

} // namespace CabanaPD

#endif
