// CabanaPD_Types.hpp -- tags and traits of the drop-in API
// (reference: src/CabanaPD_Types.hpp:19-294, subset on the force/update path).
#ifndef CABANAPD_DROPIN_TYPES_HPP
#define CABANAPD_DROPIN_TYPES_HPP

#include <array>
#include <cmath>
#include <cstddef>
#include <type_traits>

// Host-side names that user drivers written against CabanaPD spell with
// Kokkos: execution/memory space tags, fixed-size arrays, the lambda macro and
// library init.  Nothing here runs device code; the compute path is the
// engine behind include/pdk.h.
#ifndef KOKKOS_LAMBDA
#define KOKKOS_LAMBDA [=]
#endif
#ifndef KOKKOS_INLINE_FUNCTION
#define KOKKOS_INLINE_FUNCTION inline
#endif
#ifndef KOKKOS_FUNCTION
#define KOKKOS_FUNCTION
#endif

namespace Kokkos
{
struct HostSpace;
struct DefaultExecutionSpace
{
    using memory_space = HostSpace;
    using execution_space = DefaultExecutionSpace;
};
struct HostSpace
{
    using memory_space = HostSpace;
    using execution_space = DefaultExecutionSpace;
};
using DefaultHostExecutionSpace = DefaultExecutionSpace;
template <class T, std::size_t N>
using Array = std::array<T, N>;
inline void initialize( int&, char** ) {}
inline void initialize() {}
inline void finalize() {}
// Kokkos math functions used by user drivers
template <class T>
inline T abs( const T v )
{
    return std::abs( v );
}
template <class T>
inline T sqrt( const T v )
{
    return std::sqrt( v );
}
template <class T>
struct Max
{
    using value_type = T;
};
template <class T>
struct Min
{
    using value_type = T;
};
template <class T>
struct Sum
{
    using value_type = T;
};
} // namespace Kokkos

// Particle initialisation tags of Cabana::Grid::createParticles, as the
// reference's Particles::create takes them (CabanaPD_Particles.hpp:241-400).
namespace Cabana
{
struct InitUniform
{
};
struct InitRandom
{
};
} // namespace Cabana

namespace CabanaPD
{
// Model tags (src/CabanaPD_Types.hpp)
struct PMB
{
    using base_type = PMB;
};
struct LPS
{
    using base_type = LPS;
};
// Linearized variants (src/CabanaPD_Types.hpp:147-161): the force model of
// the parent, linearized stretch in the force and energy passes.
struct LinearPMB
{
    using base_type = PMB;
    using model_tag = PMB;
};
struct LinearLPS
{
    using base_type = LPS;
    using model_tag = LPS;
};
// Models whose pair force does not depend on the order of the two particles'
// materials, and indexings that give (a, b) and (b, a) one model
// (src/CabanaPD_Types.hpp:163-191).
template <class T>
struct is_symmetric : std::false_type
{
};
template <>
struct is_symmetric<PMB> : std::true_type
{
};
template <>
struct is_symmetric<LinearPMB> : std::true_type
{
};
struct Elastic
{
};
// ElasticPerfectlyPlastic (CabanaPD_Types.hpp:52): PMB with fracture
struct ElasticPerfectlyPlastic
{
};
struct NoFracture
{
};
struct Fracture
{
};
struct TemperatureIndependent
{
    using base_type = TemperatureIndependent;
};
// TemperatureDependent: an imposed temperature field (CabanaPD_Types.hpp)
struct TemperatureDependent
{
    using base_type = TemperatureDependent;
};
// DynamicTemperature: TemperatureDependent plus heat conduction (CabanaPD_Types.hpp:98-101)
struct DynamicTemperature : public TemperatureDependent
{
    using base_type = TemperatureDependent;
};
template <class T>
struct is_temperature_dependent : std::is_same<typename T::base_type, TemperatureDependent>
{
};
template <class T>
struct is_heat_transfer : std::is_same<T, DynamicTemperature>
{
};
struct SingleMaterial
{
};
struct MultiMaterial
{
};
struct Contact
{
    using base_type = Contact;
};
struct NoContact
{
    using base_type = NoContact;
    using fracture_type = NoFracture;
    using force_tag = NoContact;
};
struct BaseOutput
{
};
struct EnergyOutput
{
};
struct EnergyStressOutput
{
};

template <class T>
struct is_fracture : std::false_type
{
};
template <class T>
struct is_state_based : std::is_same<typename T::base_type, LPS>
{
};
template <class T>
struct is_energy_output
    : std::integral_constant<bool, std::is_same_v<T, EnergyOutput> || std::is_same_v<T, EnergyStressOutput>>
{
};
template <class T>
struct is_stress_output : std::is_same<T, EnergyStressOutput>
{
};
template <class T>
struct is_output
    : std::integral_constant<bool, std::is_same_v<T, BaseOutput> || std::is_same_v<T, EnergyOutput> ||
                                       std::is_same_v<T, EnergyStressOutput>>
{
};

} // namespace CabanaPD

#endif
