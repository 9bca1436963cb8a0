// CabanaPD_Solver.hpp -- Neighbor, Force, computeForce/computeEnergy and
// Solver of the drop-in API, calling the engine through include/pdk.h.
//
// Reference: src/CabanaPD_Neighbor.hpp:24-220, src/CabanaPD_Force.hpp:140-219,
// src/CabanaPD_Solver.hpp:87-657.  The step order is the reference's
// (Solver::runStep, :332-404); the engine runs the velocity-Verlet halves in
// the force pass epilogue and keeps the state on the GPU between output steps.
#ifndef CABANAPD_DROPIN_SOLVER_HPP
#define CABANAPD_DROPIN_SOLVER_HPP

#include <chrono>
#include <cmath>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <pdk.h>

#include "CabanaPD_Boundary.hpp"
#include "CabanaPD_ForceModels.hpp"
#include "CabanaPD_Input.hpp"
#include "CabanaPD_Particles.hpp"

namespace CabanaPD
{
namespace detail
{
inline void check( int rc )
{
    if ( rc != 0 )
        throw std::runtime_error( pdk_last_error() );
}

// One engine context per process (device 0).
inline pdk_ctx context()
{
    static pdk_ctx ctx = [] {
        pdk_ctx c = nullptr;
        check( pdk_ctx_create( 0, 0, &c ) );
        return c;
    }();
    return ctx;
}

// Owning handle of the engine-side specimen.
struct DeviceSpec
{
    pdk_spec spec = nullptr;
    bool lps = false;
    bool fracture = false;
    bool thermal = false;
    bool stress = false; // a stress pass has run
    DeviceSpec() = default;
    DeviceSpec( const DeviceSpec& ) = delete;
    DeviceSpec& operator=( const DeviceSpec& ) = delete;
    ~DeviceSpec()
    {
        if ( spec )
            (void)pdk_spec_destroy( spec );
    }

    // Particles::create + update -> device; the pair table of a
    // multi-material model over ntypes material types (Particles::sliceType);
    // then the Neighbor build.
    void attach( const pdk_model_desc& d, ParticleStorage& s, const std::vector<pdk_pair_model>& pairs = {},
                 int32_t ntypes = 0 )
    {
        lps = d.model == PDK_MODEL_LPS || d.model == PDK_MODEL_LINEAR_LPS;
        fracture = d.fracture != 0;
        thermal = d.thermal != PDK_TEMP_INDEPENDENT;
        check( pdk_spec_create( context(), &d, static_cast<int64_t>( s.n ), &spec ) );
        check( pdk_spec_upload( spec, s.x.data(), s.u.data(), s.v.data(), s.vol.data(), s.rho.data(),
                                s.nofail.data(), nullptr ) );
        if ( thermal )
            check( pdk_set_field( spec, PDK_FIELD_TEMPERATURE, s.temp.data() ) );
        if ( !pairs.empty() )
        {
            std::vector<int32_t> type( s.type.begin(), s.type.end() );
            check( pdk_set_pair_models( spec, ntypes, pairs.data(), type.data() ) );
        }
        check( pdk_build_bonds( spec ) );
    }

    void pull( ParticleStorage& s, std::initializer_list<int> fields )
    {
        for ( int f : fields )
        {
            double* dst = nullptr;
            switch ( f )
            {
            case PDK_FIELD_U: dst = s.u.data(); break;
            case PDK_FIELD_V: dst = s.v.data(); break;
            case PDK_FIELD_F: dst = s.f.data(); break;
            case PDK_FIELD_W: dst = s.W.data(); break;
            case PDK_FIELD_PHI: dst = s.phi.data(); break;
            case PDK_FIELD_M: dst = s.m.data(); break;
            case PDK_FIELD_THETA: dst = s.theta.data(); break;
            case PDK_FIELD_STRESS: dst = s.stress.data(); break;
            case PDK_FIELD_TEMPERATURE: dst = s.temp.data(); break;
            default: continue;
            }
            if ( f == PDK_FIELD_TEMPERATURE && !thermal )
                continue;
            if ( f == PDK_FIELD_STRESS && !stress )
                continue;
            if ( ( f == PDK_FIELD_M || f == PDK_FIELD_THETA ) && !lps )
                continue;
            check( pdk_download( spec, f, dst ) );
        }
    }
    void pullState( ParticleStorage& s )
    {
        pull( s, { PDK_FIELD_U, PDK_FIELD_V, PDK_FIELD_F, PDK_FIELD_W, PDK_FIELD_PHI, PDK_FIELD_M,
                   PDK_FIELD_THETA, PDK_FIELD_STRESS, PDK_FIELD_TEMPERATURE } );
    }
    // Stress pass (computeStress); pulled from then on.
    void computeStress()
    {
        check( pdk_compute_stress( spec ) );
        stress = true;
    }
};
} // namespace detail

// ---- Neighbor (CabanaPD_Neighbor.hpp:24-220) ----------------------------------
template <class MemorySpace, class FractureType = Fracture>
class Neighbor
{
  public:
    template <class ParticleType, class ModelType>
    Neighbor( const ModelType& model, const ParticleType& particles, const double tol = 1e-14, double dt = 1.0 )
        : _dev( std::make_shared<detail::DeviceSpec>() )
    {
        pdk_model_desc d = model.desc( dt );
        d.neighbor_tol = tol;
        int32_t ntypes = 0;
        auto pairs = detail::pairTableOf( model, particles.storage().type, ntypes );
        _dev->attach( d, particles.storage(), pairs, ntypes );
    }
    unsigned getMaxLocal() const
    {
        int64_t t = 0;
        int32_t m = 0;
        detail::check( pdk_bond_stats( _dev->spec, &t, &m ) );
        return static_cast<unsigned>( m );
    }
    void getStatistics( unsigned& max_neighbors, unsigned long long& total_neighbors ) const
    {
        int64_t t = 0;
        int32_t m = 0;
        detail::check( pdk_bond_stats( _dev->spec, &t, &m ) );
        max_neighbors = static_cast<unsigned>( m );
        total_neighbors = static_cast<unsigned long long>( t );
    }
    template <class ExecSpace, class ParticleType, class PrenotchType>
    void prenotch( ExecSpace, const ParticleType&, PrenotchType& notch )
    {
        std::vector<double> p0, v1, v2;
        notch.flatten( p0, v1, v2 );
        detail::check( pdk_prenotch( _dev->spec, static_cast<int32_t>( PrenotchType::num_notch ), p0.data(),
                                     v1.data(), v2.data() ) );
    }
    pdk_spec spec() const { return _dev->spec; }
    detail::DeviceSpec& device() const { return *_dev; }
    double time() const { return 0.0; }

  private:
    std::shared_ptr<detail::DeviceSpec> _dev;
};

// ---- Force (CabanaPD_Force.hpp:140-171, force/CabanaPD_PMB.hpp, force/CabanaPD_LPS.hpp) ----
template <class MemorySpace, class ModelTag, class FractureTag>
class Force
{
  public:
    Force() = default;
    template <class ModelType, class ParticleType, class NeighborType>
    void computeWeightedVolume( const ModelType&, ParticleType& particles, const NeighborType& neighbor )
    {
        detail::check( pdk_compute_weighted_volume( neighbor.spec() ) );
        neighbor.device().pull( particles.storage(), { PDK_FIELD_M } );
    }
    template <class ModelType, class ParticleType, class NeighborType>
    void computeDilatation( const ModelType&, ParticleType& particles, const NeighborType& neighbor )
    {
        detail::check( pdk_compute_dilatation( neighbor.spec() ) );
        neighbor.device().pull( particles.storage(), { PDK_FIELD_THETA } );
    }
    template <class ModelType, class ParticleType, class NeighborType>
    void computeForceFull( const ModelType&, ParticleType& particles, NeighborType& neighbor, bool reset )
    {
        detail::check( pdk_compute_force( neighbor.spec(), reset ? 1 : 0 ) );
        neighbor.device().pull( particles.storage(), { PDK_FIELD_F } );
    }
    template <class ModelType, class ParticleType, class NeighborType>
    void computeStressFull( const ModelType&, ParticleType& particles, const NeighborType& neighbor )
    {
        neighbor.device().computeStress();
        neighbor.device().pull( particles.storage(), { PDK_FIELD_STRESS } );
    }
    template <class ModelType, class ParticleType, class NeighborType>
    void computeEnergyFull( const ModelType&, ParticleType& particles, const NeighborType& neighbor )
    {
        detail::check( pdk_compute_energy( neighbor.spec(), &_total_strain_energy, &_total_damage ) );
        neighbor.device().pull( particles.storage(), { PDK_FIELD_W, PDK_FIELD_PHI } );
    }
    double totalStrainEnergy() const { return _total_strain_energy; }
    double totalDamage() const { return _total_damage; }
    double time() const { return 0.0; }
    double timeEnergy() const { return 0.0; }

  private:
    double _total_strain_energy = 0.0;
    double _total_damage = 0.0;
};

// computeForce / computeEnergy (CabanaPD_Force.hpp:176-219)
template <class ModelType, class ForceType, class ParticleType, class NeighborType>
void computeForce( const ModelType& model, ForceType& force, ParticleType& particles, NeighborType& neighbor,
                   const bool reset = true )
{
    force.computeForceFull( model, particles, neighbor, reset );
}

template <class ModelType, class ForceType, class ParticleType, class NeighborType>
void computeEnergy( const ModelType& model, ForceType& force, ParticleType& particles, const NeighborType& neighbor )
{
    if constexpr ( is_energy_output<typename ParticleType::output_type>::value )
        force.computeEnergyFull( model, particles, neighbor );
}

// computeStress (CabanaPD_Force.hpp:221-236): only for particles with stress output.
template <class ModelType, class ForceType, class ParticleType, class NeighborType>
void computeStress( const ModelType& model, ForceType& force, ParticleType& particles, const NeighborType& neighbor )
{
    if constexpr ( is_stress_output<typename ParticleType::output_type>::value )
        force.computeStressFull( model, particles, neighbor );
}

// ---- Solver (CabanaPD_Solver.hpp:87-657) ---------------------------------------
template <class ParticleType, class ForceModelType, class ContactModelType = NoContact>
class Solver
{
  public:
    using force_fracture_type = typename ForceModelType::fracture_type;
    using force_type = Force<typename ParticleType::memory_space, typename ForceModelType::force_tag, force_fracture_type>;
    using neighbor_type = Neighbor<typename ParticleType::memory_space, force_fracture_type>;

    Solver( Inputs _inputs, ParticleType _particles, ForceModelType _force_model )
        : particles( _particles ), inputs( _inputs ), force_model( _force_model )
    {
        static_assert( !is_temperature_dependent<typename ForceModelType::thermal_type>::value ||
                           is_temperature_dependent<typename ParticleType::thermal_type>::value,
                       "a TemperatureDependent force model needs TemperatureDependent particles" );
        setup();
    }

    // A PD force model and a contact model (CabanaPD_Solver.hpp:134-151): the
    // contact list is built on the current positions now; each step adds the
    // contact forces after the PD force (pdk_set_contact).
    Solver( Inputs _inputs, ParticleType _particles, ForceModelType _force_model, ContactModelType _contact_model )
        : particles( _particles ), inputs( _inputs ), force_model( _force_model )
    {
        static_assert( std::is_same_v<ContactModelType, NormalRepulsionModel>,
                       "the contact model of a PD Solver is NormalRepulsionModel" );
        static_assert( !is_contact<ForceModelType>::value, "the force model of a PD + contact Solver is a PD model" );
        setup();
        contact_model = _contact_model;
        const pdk_contact_desc cd = contact_model.contactDesc();
        detail::check( pdk_set_contact( spec(), &cd ) );
    }

    void setup()
    {
        // Solver::setup (CabanaPD_Solver.hpp:153-227)
        if constexpr ( !is_contact<ForceModelType>::value )
            inputs.computeCriticalTimeStep( force_model );
        if constexpr ( is_heat_transfer<typename ForceModelType::thermal_type>::value )
        {
            // heat transfer every thermal_subcycle_steps steps (CabanaPD_Solver.hpp:186-197, 340-347)
            double sub = inputs["thermal_subcycle_steps"];
            force_model.thermal_subcycle = static_cast<int>( sub );
        }
        num_steps = inputs["num_steps"];
        output_frequency = inputs["output_frequency"];
        output_reference = inputs["output_reference"];
        dt = inputs["timestep"];
        neighbor = std::make_shared<neighbor_type>( force_model, particles, 1e-14, dt );
        force = std::make_shared<force_type>();
        unsigned max_neighbors = 0;
        unsigned long long total_neighbors = 0;
        neighbor->getStatistics( max_neighbors, total_neighbors );
        log( std::cout, "Local particles: ", particles.numLocal(), ", Maximum neighbors: ", max_neighbors );
        log( std::cout, "#Timestep/Total-steps Simulation-time" );
    }

    // Solver::init overloads (CabanaPD_Solver.hpp:229-297)
    void init( const bool initial_output = true )
    {
        if ( _pre )
            _pre( 0.0 ); // force_update = false conditions at t = 0 (CabanaPD_Solver.hpp:255-257)
        detail::check( pdk_init( spec() ) );
        if constexpr ( is_stress_output<typename ParticleType::output_type>::value )
            neighbor->device().computeStress();
        pullAll();
        if ( initial_output )
            particles.output( 0, 0.0, output_reference ); // CabanaPD_Solver.hpp:245-246, 276-277
    }
    template <class BoundaryType, std::enable_if_t<is_boundary_condition<BoundaryType>::value, int> = 0>
    void init( BoundaryType boundary_condition, const bool initial_output = true )
    {
        registerBoundary( boundary_condition );
        init( initial_output );
    }
    template <std::size_t N, std::size_t V>
    void init( Prenotch<N, V> prenotch, const bool initial_output = true )
    {
        neighbor->prenotch( 0, particles, prenotch );
        init( initial_output );
    }
    template <class BoundaryType, std::size_t N, std::size_t V>
    void init( BoundaryType boundary_condition, Prenotch<N, V> prenotch, const bool initial_output = true )
    {
        neighbor->prenotch( 0, particles, prenotch );
        init( boundary_condition, initial_output );
    }

    // Solver::run (CabanaPD_Solver.hpp:406-445)
    void run() { runImpl(); }
    template <class First, class... Rest>
    void run( First& first, Rest&... rest )
    {
        if constexpr ( is_boundary_condition<First>::value )
        {
            registerBoundary( first );
            runImpl( rest... );
        }
        else
            runImpl( first, rest... );
    }

    // One step driven by the caller: afterwards the displacement, velocity
    // and force slices (and temperature) hold the new state, as the
    // reference's device slices do after Solver::runStep.
    void runStep( const int step )
    {
        if ( _pre )
            _pre( step * dt );
        detail::check( pdk_step( spec(), 1 ) );
        output( step );
        neighbor->device().pull( particles.storage(), { PDK_FIELD_U, PDK_FIELD_V, PDK_FIELD_F,
                                                        PDK_FIELD_TEMPERATURE } );
    }
    template <class BoundaryType>
    void runStep( const int step, BoundaryType boundary_condition )
    {
        registerBoundary( boundary_condition );
        runStep( step );
    }

    // Solver::updateForce (CabanaPD_Solver.hpp:462-477)
    void updateForce()
    {
        if constexpr ( is_fracture<ForceModelType>::value )
            detail::check( pdk_compute_weighted_volume( spec() ) );
        detail::check( pdk_compute_dilatation( spec() ) );
        detail::check( pdk_compute_force( spec(), 1 ) );
        neighbor->device().pull( particles.storage(), { PDK_FIELD_F } );
    }

    int num_steps = 0;
    int output_frequency = 1;
    bool output_reference = true;
    double dt = 0.0;
    ParticleType particles;

  protected:
    pdk_spec spec() const { return neighbor->spec(); }

    void pullAll() { neighbor->device().pullState( particles.storage() ); }

    template <class BC>
    void registerBoundary( BC& bc )
    {
        if constexpr ( is_body_term<std::decay_t<BC>>::value )
        {
            if ( !bc.forceUpdate() )
            {
                // applied before every force pass at that step's time
                _pre = [this, bc]( double t ) { prescribeFrom( bc, t ); };
                return;
            }
            std::vector<double> b = bc.evaluate( particles.storage() );
            detail::check( pdk_set_body_force( spec(), b.data() ) );
        }
        else if constexpr ( is_boundary_condition<std::decay_t<BC>>::value )
        {
            detail::check( pdk_set_bc( spec(), bc.kind(), bc._indices._view.data(),
                                       static_cast<int64_t>( bc._indices._view.size() ), bc._value ) );
        }
    }

    // Evaluate a force_update = false functor at time t on the host: the
    // entries it writes (marked by overwriting NaN placeholders) become the
    // engine's prescribed displacements / temperatures for the next force pass.
    template <class BC>
    void prescribeFrom( const BC& bc, double t )
    {
        ParticleStorage& s = particles.storage();
        const std::vector<int64_t> cand = bc.members( s.n );
        const double nan = std::numeric_limits<double>::quiet_NaN();
        std::vector<double> su( 3 * cand.size() ), st( cand.size() );
        for ( std::size_t m = 0; m < cand.size(); m++ )
        {
            const int64_t p = cand[m];
            for ( int d = 0; d < 3; d++ )
            {
                su[3 * m + d] = s.u[3 * p + d];
                s.u[3 * p + d] = nan;
            }
            st[m] = s.temp[p];
            s.temp[p] = nan;
        }
        bc.applyAt( t );
        std::vector<int64_t> iu, it;
        std::vector<double> au, at;
        for ( std::size_t m = 0; m < cand.size(); m++ )
        {
            const int64_t p = cand[m];
            int written = 0;
            for ( int d = 0; d < 3; d++ )
                written += std::isnan( s.u[3 * p + d] ) ? 0 : 1;
            if ( written != 0 && written != 3 )
                throw std::runtime_error( "a force_update = false functor must set all three displacement components" );
            if ( written == 3 )
            {
                iu.push_back( p );
                for ( int d = 0; d < 3; d++ )
                    au.push_back( s.u[3 * p + d] );
            }
            else
                for ( int d = 0; d < 3; d++ )
                    s.u[3 * p + d] = su[3 * m + d];
            if ( !std::isnan( s.temp[p] ) )
            {
                it.push_back( p );
                at.push_back( s.temp[p] );
            }
            else
                s.temp[p] = st[m];
        }
        if ( !it.empty() && !neighbor->device().thermal )
            throw std::runtime_error( "the functor sets temperatures but the force model is TemperatureIndependent" );
        detail::check( pdk_set_prescribed( spec(), PDK_FIELD_U, iu.data(), static_cast<int64_t>( iu.size() ),
                                           au.data(), nullptr ) );
        if ( neighbor->device().thermal )
            detail::check( pdk_set_prescribed( spec(), PDK_FIELD_TEMPERATURE, it.data(),
                                               static_cast<int64_t>( it.size() ), at.data(), nullptr ) );
    }

    void output( const int step )
    {
        if ( output_frequency > 0 && step % output_frequency == 0 )
        {
            double W = 0.0, D = 0.0;
            detail::check( pdk_compute_energy( spec(), &W, &D ) );
            if constexpr ( is_stress_output<typename ParticleType::output_type>::value )
                neighbor->device().computeStress();
            pullAll();
            particles.output( step / output_frequency, step * dt, output_reference ); // :488-489
            _last_energy = W;
            _last_damage = D;
            auto now = std::chrono::steady_clock::now();
            double sec = std::chrono::duration<double>( now - _t0 ).count();
            double rate = sec > 0 ? static_cast<double>( particles.numGlobal() ) * step / sec : 0.0;
            log( std::cout, step, "/", num_steps, " ", step * dt, " ", W, " ",
                 D / static_cast<double>( particles.numGlobal() ), " ", rate, " particle-steps/s" );
        }
    }

    template <class... OutputType>
    void runImpl( OutputType&... region_output )
    {
        _t0 = std::chrono::steady_clock::now();
        int step = 0;
        while ( step < num_steps )
        {
            int chunk = num_steps - step;
            if ( output_frequency > 0 )
                chunk = std::min( chunk, output_frequency - step % output_frequency );
            if ( _pre )
            {
                // host-evaluated conditions: one step per call
                chunk = 1;
                _pre( ( step + 1 ) * dt );
            }
            detail::check( pdk_step( spec(), chunk ) );
            step += chunk;
            output( step );
            if ( output_frequency > 0 && step % output_frequency == 0 )
                ( region_output.update( step * dt ), ... );
        }
        pullAll();
        ( region_output.print( 0 ), ... );
    }

    Inputs inputs;
    ForceModelType force_model;
    ContactModelType contact_model{};
    std::shared_ptr<neighbor_type> neighbor;
    std::shared_ptr<force_type> force;
    double _last_energy = 0.0;
    double _last_damage = 0.0;
    std::chrono::steady_clock::time_point _t0;
    std::function<void( double )> _pre; // force_update = false condition, if any
};

} // namespace CabanaPD

#endif
