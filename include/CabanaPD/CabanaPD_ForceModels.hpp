// CabanaPD_ForceModels.hpp -- ForceModel<PMB|LPS, Elastic, NoFracture|Fracture>,
// ForceModel<LinearPMB|LinearLPS, Elastic, NoFracture>, the PMB
// ElasticPerfectlyPlastic and TemperatureDependent forms,
// and the Hertz / Hertz-JKR contact models of the drop-in API.
//
// Constructors, deduction guides and public members follow the reference:
//   force_models/CabanaPD_PMB.hpp:149-231,277-297
//   force_models/CabanaPD_LPS.hpp:151-253,274-302
//   CabanaPD_ForceModels.hpp:42-129,149-256
//   force_models/CabanaPD_PMB.hpp:74-147,232-305,306-568 (EPP, thermal)
//   force_models/CabanaPD_Contact.hpp:26-51, CabanaPD_Hertzian.hpp:23-56,
//   CabanaPD_HertzianJKR.hpp:23-80
// Derived constants come from the engine (pdk_model_constants_get), which
// evaluates the reference's expressions.
#ifndef CABANAPD_DROPIN_FORCEMODELS_HPP
#define CABANAPD_DROPIN_FORCEMODELS_HPP

#include <array>
#include <cmath>
#include <stdexcept>
#include <algorithm>
#include <tuple>
#include <type_traits>
#include <vector>

#include <pdk.h>

#include "CabanaPD_Indexing.hpp"
#include "CabanaPD_Types.hpp"

namespace CabanaPD
{
namespace detail
{
// Parent model of a (possibly linearized) model tag.
template <class Tag>
struct ParentTag
{
    using type = Tag;
};
template <>
struct ParentTag<LinearPMB>
{
    using type = PMB;
};
template <>
struct ParentTag<LinearLPS>
{
    using type = LPS;
};

inline pdk_model_constants constantsOf( const pdk_model_desc& d )
{
    pdk_model_constants k{};
    if ( pdk_model_constants_get( &d, &k ) != 0 )
        throw std::runtime_error( pdk_last_error() );
    return k;
}
inline pdk_model_desc blankDesc()
{
    pdk_model_desc d{};
    d.neighbor_tol = 1e-14;
    d.nu = 0.25;
    d.restitution = 1.0;
    return d;
}
} // namespace detail

template <class ModelType, class MechanicsType = Elastic, class FractureType = Fracture,
          class ThermalType = TemperatureIndependent, class... FieldTypes>
struct ForceModel;

struct BaseForceModel
{
    using material_type = SingleMaterial;
    using thermal_type = TemperatureIndependent;
    double force_horizon = 0.0;
    double K = 0.0;
    auto cutoff() const { return force_horizon; }
    auto extend() const { return 0.0; }
    void updateBonds( const int, const int ) {}
};

template <class Tag, class FractureType>
struct PDForceModel : BaseForceModel
{
    // As in the reference, a linearized model keeps its parent's model_tag
    // (src/CabanaPD_Types.hpp:147-161) and only LinearPMB replaces the force
    // tag (force_models/CabanaPD_PMB.hpp:584); ForceModel<LinearLPS> keeps
    // force_tag = LPS (force_models/CabanaPD_LPS.hpp:31-32, 255-272), so the
    // solver runs the LPS force for it (CabanaPD_Solver.hpp:96-99).
    using model_tag = typename detail::ParentTag<Tag>::type;
    using force_tag = std::conditional_t<std::is_same_v<Tag, LinearLPS>, LPS, Tag>;
    using base_model = Tag;
    using fracture_type = FractureType;
    using mechanics_type = Elastic;
    double G = 0.0;
    double G0 = 0.0;
    double s0 = 0.0;
    double bond_break_coeff = 0.0;
    double c = 0.0;
    // LPS coefficients on particle i's terms ([0]) and particle j's ([1]);
    // both the model's own unless the model was averaged from two
    // (force_models/CabanaPD_LPS.hpp:41-44, 55-69)
    std::array<double, 2> theta_coeff{};
    std::array<double, 2> s_coeff{};
    int influence_type = 0;

    // This model's constants as one entry of a multi-material pair table.
    pdk_pair_model pairModel() const
    {
        pdk_pair_model p{};
        p.c = c;
        p.bond_break_coeff = std::is_same_v<FractureType, Fracture> ? bond_break_coeff : INFINITY;
        for ( int side = 0; side < 2; side++ )
        {
            p.theta_coeff[side] = theta_coeff[side];
            p.s_coeff[side] = s_coeff[side];
        }
        return p;
    }

    // The engine's description of this model for time step dt.
    pdk_model_desc desc( double dt ) const
    {
        pdk_model_desc d = detail::blankDesc();
        d.model = std::is_same_v<force_tag, LPS>         ? PDK_MODEL_LPS
                  : std::is_same_v<force_tag, LinearPMB> ? PDK_MODEL_LINEAR_PMB
                  : std::is_same_v<force_tag, LinearLPS> ? PDK_MODEL_LINEAR_LPS
                                                         : PDK_MODEL_PMB;
        d.fracture = std::is_same_v<FractureType, Fracture> ? 1 : 0;
        d.influence = influence_type;
        d.precision = PDK_FP64;
        d.horizon = force_horizon;
        d.K = K;
        d.G = G;
        d.G0 = G0;
        d.dt = dt;
        d.mechanics = mechanics_;
        d.thermal = thermal_;
        d.sigma_y = sigma_y_;
        d.alpha = alpha_;
        d.temp0 = temp0_;
        return d;
    }

  protected:
    // EPP and thermal coupling of the PMB forms (the engine's description)
    int32_t mechanics_ = PDK_MECH_ELASTIC;
    int32_t thermal_ = PDK_TEMP_INDEPENDENT;
    double sigma_y_ = 0.0, alpha_ = 0.0, temp0_ = 0.0;

    void derive()
    {
        pdk_model_constants k = detail::constantsOf( desc( 1.0 ) );
        c = k.c;
        s0 = k.s0;
        bond_break_coeff = k.bond_break_coeff;
        theta_coeff = { k.theta_coeff, k.theta_coeff };
        s_coeff = { k.s_coeff, k.s_coeff };
    }

    // The constructor that averages two models (BaseForceModel,
    // CabanaPD_ForceModels.hpp:58-62: first horizon, mean K;
    // force_models/CabanaPD_PMB.hpp:51-55: mean c;
    // force_models/CabanaPD_LPS.hpp:55-69: mean G, [0] coefficients from the
    // first model's K and G, [1] from the second's;
    // BaseFractureModel, CabanaPD_ForceModels.hpp:114-121: mean G0,
    // K-weighted s0).
    template <class M1, class M2>
    void average( const M1& m1, const M2& m2 )
    {
        force_horizon = m1.force_horizon;
        K = ( m1.K + m2.K ) / 2.0;
        G = ( m1.G + m2.G ) / 2.0;
        c = ( m1.c + m2.c ) / 2.0;
        theta_coeff = { 3.0 * m1.K - 5.0 * m1.G, 3.0 * m2.K - 5.0 * m2.G };
        s_coeff = { 15.0 * m1.G, 15.0 * m2.G };
        influence_type = m1.influence_type;
        if ( m2.influence_type != m1.influence_type )
            throw std::runtime_error( "Influence function type for each model must match for multi-material systems" );
        if constexpr ( std::is_same_v<FractureType, Fracture> )
        {
            G0 = ( m1.G0 + m2.G0 ) / 2.0;
            s0 = std::sqrt( ( m1.s0 * m1.s0 * m1.K + m2.s0 * m2.s0 * m2.K ) / ( m1.K + m2.K ) );
            bond_break_coeff = ( 1.0 + s0 ) * ( 1.0 + s0 );
        }
    }
};

namespace detail
{
template <class M>
using IfPDModel = std::enable_if_t<std::is_base_of_v<BaseForceModel, M>, int>;
} // namespace detail

template <class Model, class Mech, class F, class Thermal, class... Fields>
struct is_fracture<ForceModel<Model, Mech, F, Thermal, Fields...>> : std::is_same<F, Fracture>
{
};

// ---- PMB -------------------------------------------------------------------
template <>
struct ForceModel<PMB, Elastic, NoFracture> : PDForceModel<PMB, NoFracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( PMB, NoFracture, const double force_horizon_, const double K_ ) { set( force_horizon_, K_ ); }
    ForceModel( PMB, Elastic, NoFracture, const double force_horizon_, const double K_ )
    {
        set( force_horizon_, K_ );
    }

  private:
    void set( double h, double k )
    {
        force_horizon = h;
        K = k;
        influence_type = 1;
        derive();
    }
};

template <>
struct ForceModel<PMB, Elastic, Fracture> : PDForceModel<PMB, Fracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( PMB, const double h, const double k, const double g0, const int influence = 1 )
    {
        set( h, k, g0, influence );
    }
    ForceModel( PMB, Elastic, const double h, const double k, const double g0, const int influence = 1 )
    {
        set( h, k, g0, influence );
    }
    ForceModel( PMB, Fracture, const double h, const double k, const double g0, const int influence = 1 )
    {
        set( h, k, g0, influence );
    }
    ForceModel( PMB, Elastic, Fracture, const double h, const double k, const double g0, const int influence = 1 )
    {
        set( h, k, g0, influence );
    }

  private:
    void set( double h, double k, double g0, int infl )
    {
        force_horizon = h;
        K = k;
        G0 = g0;
        influence_type = infl;
        derive();
    }
};

// ---- PMB ElasticPerfectlyPlastic (force_models/CabanaPD_PMB.hpp:74-147, 232-284) ----
// One plastic stretch per bond, held by the engine (zero when the bonds are
// built); s_Y = sigma_y / (3 K), s0 = (5 G0 / sigma_y / delta + sigma_y / K) / 6.
template <class MemorySpace>
struct ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureIndependent, MemorySpace>
    : PDForceModel<PMB, Fracture>
{
    using mechanics_type = ElasticPerfectlyPlastic;
    double s_Y = 0.0;
    ForceModel( PMB, ElasticPerfectlyPlastic, MemorySpace, const double h, const double k, const double g0,
                const double sigma_y )
    {
        set( h, k, g0, sigma_y );
    }
    ForceModel( PMB, ElasticPerfectlyPlastic, Fracture, MemorySpace, const double h, const double k, const double g0,
                const double sigma_y )
    {
        set( h, k, g0, sigma_y );
    }
    void updateBonds( const int, const int ) {}

  private:
    void set( double h, double k, double g0, double sy )
    {
        force_horizon = h;
        K = k;
        G0 = g0;
        influence_type = 1;
        mechanics_ = PDK_MECH_EPP;
        sigma_y_ = sy;
        derive();
        s_Y = detail::constantsOf( desc( 1.0 ) ).s_Y;
    }
};

// ---- PMB TemperatureDependent (force_models/CabanaPD_PMB.hpp:306-568) ------------
// The temperature field lives in the particles (sliceTemperature) and on the
// device; alpha, temp0 and the slice are kept as the reference's members.
// DynamicTemperature adds BaseDynamicTemperatureModel's kappa, cp,
// microconductivity (CabanaPD_ForceModels.hpp:262-310); the Solver sets
// thermal_subcycle from the deck's "thermal_subcycle_steps".
template <class Tag, class Mech, class F, class TemperatureType, class ThermalTag = TemperatureDependent>
struct ThermalPMB : PDForceModel<Tag, F>
{
    using thermal_type = ThermalTag;
    using mechanics_type = Mech;
    double alpha = 0.0;
    double temp0 = 0.0;
    double s_Y = 0.0;
    TemperatureType temperature;
    // DynamicTemperature
    double thermal_horizon = 0.0;
    double thermal_coeff = 0.0;
    double kappa = 0.0;
    double cp = 0.0;
    bool constant_microconductivity = true;
    int thermal_subcycle = 1;

    pdk_model_desc desc( double dt ) const
    {
        pdk_model_desc d = PDForceModel<Tag, F>::desc( dt );
        d.kappa = kappa;
        d.cp = cp;
        d.linear_microconductivity = constant_microconductivity ? 0 : 1;
        d.thermal_subcycle = thermal_subcycle;
        return d;
    }

  protected:
    void setHeat( double k, double c, bool constant )
    {
        kappa = k;
        cp = c;
        constant_microconductivity = constant;
    }
    void setThermal( double h, double k, double g0, double sy, const TemperatureType& t, double a, double t0 )
    {
        this->force_horizon = h;
        this->K = k;
        this->G0 = g0;
        this->influence_type = 1; // ThermalFractureModel's default (CabanaPD_ForceModels.hpp:228-236)
        this->thermal_ = std::is_same_v<ThermalTag, DynamicTemperature> ? PDK_TEMP_DYNAMIC : PDK_TEMP_DEPENDENT;
        this->alpha_ = a;
        this->temp0_ = t0;
        if ( std::is_same_v<Mech, ElasticPerfectlyPlastic> )
        {
            this->mechanics_ = PDK_MECH_EPP;
            this->sigma_y_ = sy;
        }
        alpha = a;
        temp0 = t0;
        temperature = t;
        thermal_horizon = h;
        // constants from the full description (the heat terms included)
        const pdk_model_constants kc = detail::constantsOf( desc( 1.0 ) );
        this->c = kc.c;
        this->s0 = kc.s0;
        this->bond_break_coeff = kc.bond_break_coeff;
        s_Y = kc.s_Y;
        thermal_coeff = kc.thermal_coeff;
    }
};

template <class TemperatureType>
struct ForceModel<PMB, Elastic, NoFracture, TemperatureDependent, TemperatureType>
    : ThermalPMB<PMB, Elastic, NoFracture, TemperatureType>
{
    ForceModel( PMB, NoFracture, const double h, const double k, const TemperatureType t, const double a,
                const double t0 = 0.0 )
    {
        this->setThermal( h, k, 0.0, 0.0, t, a, t0 );
    }
};
template <class TemperatureType>
struct ForceModel<PMB, Elastic, Fracture, TemperatureDependent, TemperatureType>
    : ThermalPMB<PMB, Elastic, Fracture, TemperatureType>
{
    ForceModel( PMB, const double h, const double k, const double g0, const TemperatureType t, const double a,
                const double t0 = 0.0 )
    {
        this->setThermal( h, k, g0, 0.0, t, a, t0 );
    }
};
template <class TemperatureType>
struct ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureDependent, TemperatureType>
    : ThermalPMB<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureType>
{
    ForceModel( PMB, ElasticPerfectlyPlastic, const double h, const double k, const double g0, const double sigma_y,
                const TemperatureType t, const double a, const double t0 = 0.0 )
    {
        this->setThermal( h, k, g0, sigma_y, t, a, t0 );
    }
};
// ForceModel<LinearPMB, ...> derives from the PMB form with force_tag
// LinearPMB (force_models/CabanaPD_PMB.hpp:575-597).
template <class TemperatureType>
struct ForceModel<LinearPMB, Elastic, NoFracture, TemperatureDependent, TemperatureType>
    : ThermalPMB<LinearPMB, Elastic, NoFracture, TemperatureType>
{
    ForceModel( LinearPMB, NoFracture, const double h, const double k, const TemperatureType t, const double a,
                const double t0 = 0.0 )
    {
        this->setThermal( h, k, 0.0, 0.0, t, a, t0 );
    }
};

// DynamicTemperature forms (force_models/CabanaPD_PMB.hpp:421-568)
template <class TemperatureType>
struct ForceModel<PMB, Elastic, NoFracture, DynamicTemperature, TemperatureType>
    : ThermalPMB<PMB, Elastic, NoFracture, TemperatureType, DynamicTemperature>
{
    ForceModel( PMB, NoFracture, const double h, const double k, const TemperatureType& t, const double kap,
                const double c, const double a, const double t0 = 0.0, const bool constant = true )
    {
        this->setHeat( kap, c, constant );
        this->setThermal( h, k, 0.0, 0.0, t, a, t0 );
    }
};
template <class TemperatureType>
struct ForceModel<PMB, Elastic, Fracture, DynamicTemperature, TemperatureType>
    : ThermalPMB<PMB, Elastic, Fracture, TemperatureType, DynamicTemperature>
{
    ForceModel( PMB, const double h, const double k, const double g0, const TemperatureType t, const double kap,
                const double c, const double a, const double t0 = 0.0, const bool constant = true )
    {
        this->setHeat( kap, c, constant );
        this->setThermal( h, k, g0, 0.0, t, a, t0 );
    }
};
template <class TemperatureType>
struct ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, DynamicTemperature, TemperatureType>
    : ThermalPMB<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureType, DynamicTemperature>
{
    ForceModel( PMB, ElasticPerfectlyPlastic, const double h, const double k, const double g0, const double sigma_y,
                const TemperatureType& t, const double kap, const double c, const double a, const double t0 = 0.0,
                const bool constant = true )
    {
        this->setHeat( kap, c, constant );
        this->setThermal( h, k, g0, sigma_y, t, a, t0 );
    }
};

namespace detail
{
template <class T>
using IfField = std::enable_if_t<!std::is_arithmetic_v<T>, int>;
}

template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, NoFracture, double, double, TT, double, double, double )
    -> ForceModel<PMB, Elastic, NoFracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, NoFracture, double, double, TT, double, double, double, double )
    -> ForceModel<PMB, Elastic, NoFracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, NoFracture, double, double, TT, double, double, double, double, bool )
    -> ForceModel<PMB, Elastic, NoFracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, double, double, double, TT, double, double, double )
    -> ForceModel<PMB, Elastic, Fracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, double, double, double, TT, double, double, double, double )
    -> ForceModel<PMB, Elastic, Fracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, double, double, double, TT, double, double, double, double, bool )
    -> ForceModel<PMB, Elastic, Fracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, double, double, double, double, TT, double, double, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, double, double, double, double, TT, double, double, double, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, DynamicTemperature, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, double, double, double, double, TT, double, double, double, double, bool )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, DynamicTemperature, TT>;

template <class MS, detail::IfField<MS> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, MS, double, double, double, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureIndependent, MS>;
template <class MS, detail::IfField<MS> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, Fracture, MS, double, double, double, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureIndependent, MS>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, NoFracture, double, double, TT, double ) -> ForceModel<PMB, Elastic, NoFracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, NoFracture, double, double, TT, double, double )
    -> ForceModel<PMB, Elastic, NoFracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, double, double, double, TT, double ) -> ForceModel<PMB, Elastic, Fracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, double, double, double, TT, double, double )
    -> ForceModel<PMB, Elastic, Fracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, double, double, double, double, TT, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( PMB, ElasticPerfectlyPlastic, double, double, double, double, TT, double, double )
    -> ForceModel<PMB, ElasticPerfectlyPlastic, Fracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( LinearPMB, NoFracture, double, double, TT, double )
    -> ForceModel<LinearPMB, Elastic, NoFracture, TemperatureDependent, TT>;
template <class TT, detail::IfField<TT> = 0>
ForceModel( LinearPMB, NoFracture, double, double, TT, double, double )
    -> ForceModel<LinearPMB, Elastic, NoFracture, TemperatureDependent, TT>;

// ---- LPS -------------------------------------------------------------------
template <>
struct ForceModel<LPS, Elastic, NoFracture> : PDForceModel<LPS, NoFracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( LPS, NoFracture, const double h, const double k, const double g, const int influence = 0 )
    {
        set( h, k, g, influence );
    }
    ForceModel( LPS, Elastic, NoFracture, const double h, const double k, const double g, const int influence = 0 )
    {
        set( h, k, g, influence );
    }

  private:
    void set( double h, double k, double g, int infl )
    {
        force_horizon = h;
        K = k;
        G = g;
        influence_type = infl;
        derive();
    }
};

template <>
struct ForceModel<LPS, Elastic, Fracture> : PDForceModel<LPS, Fracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( LPS, const double h, const double k, const double g, const double g0, const int influence = 0 )
    {
        set( h, k, g, g0, influence );
    }
    ForceModel( LPS, Fracture, const double h, const double k, const double g, const double g0,
                const int influence = 0 )
    {
        set( h, k, g, g0, influence );
    }
    ForceModel( LPS, Elastic, const double h, const double k, const double g, const double g0,
                const int influence = 0 )
    {
        set( h, k, g, g0, influence );
    }
    ForceModel( LPS, Elastic, Fracture, const double h, const double k, const double g, const double g0,
                const int influence = 0 )
    {
        set( h, k, g, g0, influence );
    }

  private:
    void set( double h, double k, double g, double g0, int infl )
    {
        force_horizon = h;
        K = k;
        G = g;
        G0 = g0;
        influence_type = infl;
        derive();
    }
};

// ---- linearized models (force_models/CabanaPD_PMB.hpp:575-597,
//      force_models/CabanaPD_LPS.hpp:255-273): NoFracture only -----------------
template <>
struct ForceModel<LinearPMB, Elastic, NoFracture> : PDForceModel<LinearPMB, NoFracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( LinearPMB, NoFracture, const double h, const double k ) { set( h, k ); }
    ForceModel( LinearPMB, Elastic, NoFracture, const double h, const double k ) { set( h, k ); }

  private:
    void set( double h, double k )
    {
        force_horizon = h;
        K = k;
        influence_type = 1;
        derive();
    }
};

template <>
struct ForceModel<LinearLPS, Elastic, NoFracture> : PDForceModel<LinearLPS, NoFracture>
{
    // averaged from two models (multi-material bonds between their materials)
    template <class M1, class M2, detail::IfPDModel<M1> = 0, detail::IfPDModel<M2> = 0>
    ForceModel( const M1& m1, const M2& m2 )
    {
        this->average( m1, m2 );
    }
    ForceModel( LinearLPS, NoFracture, const double h, const double k, const double g, const int influence = 0 )
    {
        set( h, k, g, influence );
    }
    ForceModel( LinearLPS, Elastic, NoFracture, const double h, const double k, const double g,
                const int influence = 0 )
    {
        set( h, k, g, influence );
    }

  private:
    void set( double h, double k, double g, int infl )
    {
        force_horizon = h;
        K = k;
        G = g;
        influence_type = infl;
        derive();
    }
};

// Deduction guides (force_models/CabanaPD_PMB.hpp:277-297, force_models/CabanaPD_LPS.hpp:274-302)
ForceModel( LinearPMB, NoFracture, double, double ) -> ForceModel<LinearPMB, Elastic, NoFracture>;
ForceModel( LinearPMB, Elastic, NoFracture, double, double ) -> ForceModel<LinearPMB, Elastic, NoFracture>;
ForceModel( LinearLPS, NoFracture, double, double, double ) -> ForceModel<LinearLPS, Elastic, NoFracture>;
ForceModel( LinearLPS, NoFracture, double, double, double, int ) -> ForceModel<LinearLPS, Elastic, NoFracture>;
ForceModel( LinearLPS, Elastic, NoFracture, double, double, double ) -> ForceModel<LinearLPS, Elastic, NoFracture>;
ForceModel( LinearLPS, Elastic, NoFracture, double, double, double, int )
    -> ForceModel<LinearLPS, Elastic, NoFracture>;
ForceModel( PMB, NoFracture, double, double ) -> ForceModel<PMB, Elastic, NoFracture>;
ForceModel( PMB, Elastic, NoFracture, double, double ) -> ForceModel<PMB, Elastic, NoFracture>;
ForceModel( PMB, double, double, double ) -> ForceModel<PMB, Elastic, Fracture>;
ForceModel( PMB, double, double, double, int ) -> ForceModel<PMB, Elastic, Fracture>;
ForceModel( PMB, Elastic, double, double, double ) -> ForceModel<PMB, Elastic, Fracture>;
ForceModel( PMB, Fracture, double, double, double ) -> ForceModel<PMB, Elastic, Fracture>;
ForceModel( PMB, Elastic, Fracture, double, double, double ) -> ForceModel<PMB, Elastic, Fracture>;
ForceModel( LPS, NoFracture, double, double, double ) -> ForceModel<LPS, Elastic, NoFracture>;
ForceModel( LPS, NoFracture, double, double, double, int ) -> ForceModel<LPS, Elastic, NoFracture>;
ForceModel( LPS, Elastic, NoFracture, double, double, double ) -> ForceModel<LPS, Elastic, NoFracture>;
ForceModel( LPS, Elastic, NoFracture, double, double, double, int ) -> ForceModel<LPS, Elastic, NoFracture>;
ForceModel( LPS, double, double, double, double ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, double, double, double, double, int ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Fracture, double, double, double, double ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Fracture, double, double, double, double, int ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Elastic, double, double, double, double ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Elastic, double, double, double, double, int ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Elastic, Fracture, double, double, double, double ) -> ForceModel<LPS, Elastic, Fracture>;
ForceModel( LPS, Elastic, Fracture, double, double, double, double, int ) -> ForceModel<LPS, Elastic, Fracture>;

// ---- contact ---------------------------------------------------------------
struct ContactModel
{
    using model_tag = Contact;
    using material_type = SingleMaterial;
    using thermal_type = TemperatureIndependent;
    using fracture_type = NoFracture;
    double radius = 0.0;
    double radius_extend = 0.0;
    double K = 0.0;
    ContactModel() = default;
    ContactModel( const double r, const double e ) : radius( r ), radius_extend( e ) {}
    auto cutoff() const { return 2.0 * radius + radius_extend; }
    auto extend() const { return radius_extend; }
    void updateBonds( const int, const int ) {}
};

struct HertzianModel : ContactModel
{
    using force_tag = HertzianModel;
    using base_model = HertzianModel;
    double nu = 0.25, E = 0.0, e = 1.0;
    double Rs = 0.0, Es = 0.0, beta = 0.0, coeff_h_n = 0.0, coeff_h_d = 0.0;
    HertzianModel() = default;
    HertzianModel( const double r, const double ext, const double nu_, const double E_, const double e_ )
        : ContactModel( r, ext ), nu( nu_ ), E( E_ ), e( e_ )
    {
        auto k = detail::constantsOf( desc( 1.0 ) );
        Rs = k.Rs;
        Es = k.Es;
        beta = k.beta;
        coeff_h_n = k.coeff_h_n;
        coeff_h_d = k.coeff_h_d;
    }
    pdk_model_desc desc( double dt ) const
    {
        pdk_model_desc d = detail::blankDesc();
        d.model = PDK_MODEL_HERTZ;
        d.radius = radius;
        d.radius_extend = radius_extend;
        d.nu = nu;
        d.E = E;
        d.restitution = e;
        d.dt = dt;
        return d;
    }
};

struct HertzianJKRModel : ContactModel
{
    using force_tag = HertzianModel; // reference: HertzianJKR.hpp:29
    using base_model = HertzianJKRModel;
    double nu = 0.25, E = 0.0, mu = 0.0, e = 1.0, gamma = 0.0;
    double Rs = 0.0, Es = 0.0, Gs = 0.0, beta = 0.0, coeff_h_n = 0.0, coeff_h_d = 0.0, Gamma = 0.0, a0 = 0.0,
           delta_tear = 0.0, fc = 0.0;
    HertzianJKRModel() = default;
    HertzianJKRModel( const double r, const double ext, const double nu_, const double E_, const double e_,
                      const double gamma_, const double mu_ )
        : ContactModel( r, ext ), nu( nu_ ), E( E_ ), mu( mu_ ), e( e_ ), gamma( gamma_ )
    {
        auto k = detail::constantsOf( desc( 1.0 ) );
        Rs = k.Rs;
        Es = k.Es;
        Gs = k.Gs;
        beta = k.beta;
        coeff_h_n = k.coeff_h_n;
        coeff_h_d = k.coeff_h_d;
        Gamma = k.Gamma;
        a0 = k.a0;
        delta_tear = k.delta_tear;
        fc = k.fc;
    }
    pdk_model_desc desc( double dt ) const
    {
        pdk_model_desc d = detail::blankDesc();
        d.model = PDK_MODEL_HERTZ_JKR;
        d.radius = radius;
        d.radius_extend = radius_extend;
        d.nu = nu;
        d.E = E;
        d.restitution = e;
        d.gamma = gamma;
        d.mu_friction = mu;
        d.dt = dt;
        return d;
    }
};

// NormalRepulsionModel (force_models/CabanaPD_Contact.hpp:53-90): the contact
// model the reference pairs with a PD force model in one Solver (repulsion
// 15 c ((r - radius) / delta) V_j for r < radius, c = 18 K / (pi delta^4)).
struct NormalRepulsionModel : ContactModel
{
    using base_type = ContactModel;
    using force_tag = NormalRepulsionModel;
    double delta = 0.0;
    double c = 0.0;
    NormalRepulsionModel() = default;
    NormalRepulsionModel( const double delta_, const double r, const double ext, const double K_ )
        : ContactModel( r, ext ), delta( delta_ )
    {
        K = K_;
        c = 18.0 * K / ( 3.141592653589793238462643383279502884 * delta * delta * delta * delta );
    }
    double forceCoeff( const double r, const double vol ) const
    {
        const double sc = ( r - radius ) / delta;
        return 15.0 * c * sc * vol;
    }
    pdk_contact_desc contactDesc() const
    {
        pdk_contact_desc cd{};
        cd.model = PDK_CONTACT_NORMAL_REPULSION;
        cd.radius = radius;
        cd.radius_extend = radius_extend;
        cd.horizon = delta;
        cd.K = K;
        return cd;
    }
};

template <class T>
struct is_contact : std::integral_constant<bool, std::is_base_of_v<ContactModel, T>>
{
};

// ForceModel m12( m1, m2 ): the type of the first model.
template <class T, class Mech, class F, class Th, class... X, class M2, detail::IfPDModel<M2> = 0>
ForceModel( const ForceModel<T, Mech, F, Th, X...>&, const M2& ) -> ForceModel<T, Mech, F, Th, X...>;

// ---- multi-material models (CabanaPD_ForceModelsMulti.hpp:185-358) ------------
// ForceModels< Indexing, M_0, .. >: a list of models of one model kind,
// fracture form and horizon, and an indexing; the bond from a particle of
// material type a (Particles::sliceType) to one of type b uses the model at
// list position indexing(a, b).  The engine receives the table of those
// models' constants per ordered type pair (pdk_set_pair_models).
//
//   createMultiForceModel( particles, indexing, m_0, .., m_{k-1} )
//   createMultiForceModel( particles, AverageTag{}, m_0, .., m_{n-1} ), 2 to 4
//     models: the list m_0 .. m_{n-1} followed by the models averaged between
//     each pair of materials, ordered for DiagonalIndexing<n> (PMB forms,
//     symmetric) or FullIndexing<n> (LPS forms) as the reference orders them
//     (CabanaPD_ForceModelsMulti.hpp:360-502).
struct AverageTag
{
};

template <class Indexing, class M0, class... Ms>
struct ForceModels
{
    static_assert( ( std::is_same_v<typename M0::force_tag, typename Ms::force_tag> && ... ),
                   "All forces need the same base type" );
    static_assert( ( std::is_same_v<typename M0::model_tag, typename Ms::model_tag> && ... ),
                   "All models need the same base model" );
    static_assert( ( std::is_same_v<typename M0::base_model, typename Ms::base_model> && ... ),
                   "this build needs one model form (linearized or not) for all materials" );
    static_assert( ( std::is_same_v<typename M0::fracture_type, typename Ms::fracture_type> && ... ),
                   "this build needs one fracture form for all materials" );
    static_assert( is_symmetric<typename M0::model_tag>::value == is_symmetric<Indexing>::value,
                   "Symmetry of model and indexing must match" );
    using model_tag = typename M0::model_tag;
    using force_tag = typename M0::force_tag;
    using base_model = typename M0::base_model;
    using fracture_type = typename M0::fracture_type;
    using mechanics_type = Elastic;
    using material_type = MultiMaterial;
    using thermal_type = TemperatureIndependent;
    using indexing_type = Indexing;
    static constexpr std::size_t num_models = 1 + sizeof...( Ms );

    double force_horizon = 0.0;
    double K = 0.0;
    Indexing indexing;
    // the constants of each listed model, in list order
    std::vector<pdk_pair_model> models;
    M0 first;

    ForceModels( const Indexing& i, const M0& m0, const Ms&... ms )
        : force_horizon( m0.force_horizon ), K( m0.K ), indexing( i ), first( m0 )
    {
        add( m0 );
        ( add( ms ), ... );
    }
    // A list whose constants are already known (the AverageTag form).
    ForceModels( const Indexing& i, const M0& m0, std::vector<pdk_pair_model> list )
        : force_horizon( m0.force_horizon ), K( m0.K ), indexing( i ), models( std::move( list ) ), first( m0 )
    {
    }
    auto cutoff() const { return force_horizon; }
    void updateBonds( const int, const int ) {}
    pdk_model_desc desc( double dt ) const { return first.desc( dt ); }

    // Material types covered: the indexing's own count, else the particles'.
    int32_t numTypes( const std::vector<int>& type ) const
    {
        unsigned n = detail::IndexingTypes<Indexing>::value;
        if ( n == 0 )
            for ( int t : type )
                n = std::max( n, static_cast<unsigned>( std::max( t, 0 ) ) + 1u );
        if ( n > 4 )
            throw std::runtime_error( "this build supports 1 to 4 material types" );
        return static_cast<int32_t>( std::max( n, 1u ) );
    }
    // The pair table for ntypes types: entry a * ntypes + b holds the model
    // at list position indexing(a, b).
    std::vector<pdk_pair_model> pairTable( int32_t ntypes ) const
    {
        std::vector<pdk_pair_model> t( static_cast<size_t>( ntypes ) * ntypes );
        for ( int32_t a = 0; a < ntypes; a++ )
            for ( int32_t b = 0; b < ntypes; b++ )
            {
                const unsigned k = indexing( static_cast<unsigned>( a ), static_cast<unsigned>( b ) );
                if ( k > Indexing::MaxValidIndex )
                    throw std::out_of_range( "CabanaPD: Indexing computed an index that is out of its valid range" );
                if ( k >= models.size() )
                    throw std::out_of_range( "Functor outside of range requested" );
                t[static_cast<size_t>( a ) * ntypes + b] = models[k];
            }
        return t;
    }

  private:
    template <class M>
    void add( const M& m )
    {
        // CabanaPD_ForceModelsMulti.hpp:295-311: equal horizons required
        if ( std::abs( m.force_horizon - force_horizon ) > 1e-10 )
            throw std::runtime_error( "Horizon for each model must match for multi-material systems." );
        models.push_back( m.pairModel() );
    }
};

template <class Indexing, class M0, class... Ms>
struct is_fracture<ForceModels<Indexing, M0, Ms...>> : std::is_same<typename M0::fracture_type, Fracture>
{
};

template <class ParticleType, class Indexing, class M0, class... Ms>
auto createMultiForceModel( ParticleType, Indexing indexing, M0 m0, Ms... ms )
{
    return ForceModels<Indexing, M0, Ms...>( indexing, m0, ms... );
}

namespace detail
{
template <unsigned N, class Sym>
using AveragedIndexing = std::conditional_t<Sym::value, DiagonalIndexing<N>, FullIndexing<N>>;
} // namespace detail

template <class ParticleType, class M0, class... Ms>
auto createMultiForceModel( ParticleType, AverageTag, M0 m0, Ms... ms )
{
    constexpr unsigned n = 1 + sizeof...( Ms );
    static_assert( n >= 2 && n <= 4, "2 to 4 materials" );
    static_assert( ( std::is_same_v<M0, Ms> && ... ), "this build averages models of one type" );
    static_assert( std::is_constructible_v<M0, const M0&, const M0&>,
                   "this build averages the Elastic PMB, LPS and linearized forms" );
    using Sym = is_symmetric<typename M0::model_tag>;
    using Indexing = detail::AveragedIndexing<n, Sym>;
    const std::vector<M0> mats{ m0, ms... };
    for ( const M0& m : mats )
        // CabanaPD_ForceModelsMulti.hpp:295-311: equal horizons required
        if ( std::abs( m.force_horizon - m0.force_horizon ) > 1e-10 )
            throw std::runtime_error( "Horizon for each model must match for multi-material systems." );
    auto pairOf = [&]( unsigned a, unsigned b ) {
        return a == b ? mats[a].pairModel() : M0( mats[a], mats[b] ).pairModel();
    };
    std::vector<pdk_pair_model> models;
    if constexpr ( Sym::value )
    {
        // diagonal d = 0 (the materials), then d = 1 .. n-1, each by min(a, b)
        for ( unsigned d = 0; d < n; d++ )
            for ( unsigned a = 0; a + d < n; a++ )
                models.push_back( pairOf( a, a + d ) );
    }
    else
    {
        for ( unsigned a = 0; a < n; a++ )
            for ( unsigned b = 0; b < n; b++ )
                models.push_back( pairOf( a, b ) );
    }
    return ForceModels<Indexing, M0, Ms...>( Indexing{}, m0, std::move( models ) );
}

namespace detail
{
// The pair table of a model for pdk_set_pair_models (empty: one material).
template <class Model>
std::vector<pdk_pair_model> pairTableOf( const Model&, const std::vector<int>&, int32_t& ntypes )
{
    ntypes = 0;
    return {};
}
template <class Indexing, class M0, class... Ms>
std::vector<pdk_pair_model> pairTableOf( const ForceModels<Indexing, M0, Ms...>& m, const std::vector<int>& type,
                                         int32_t& ntypes )
{
    ntypes = m.numTypes( type );
    return m.pairTable( ntypes );
}
} // namespace detail

} // namespace CabanaPD

#endif
