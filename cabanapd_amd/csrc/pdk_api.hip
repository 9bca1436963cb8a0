// pdk_api.hip -- the C ABI (include/pdk.h): specimen state on one GPU and the
// host orchestration of the reference's Solver step order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "pdk_internal.hpp"
#include "pdk_serving.hpp"

using namespace pdk;

namespace
{
thread_local std::string g_last_error;

// Kokkos::numbers::pi_v<double> (CabanaPD_Constants.hpp:20)
constexpr double kPi = 3.141592653589793238462643383279502884;

template <class F>
int guarded( F&& f )
{
    try
    {
        f();
        return 0;
    }
    catch ( const std::exception& e )
    {
        g_last_error = e.what();
        return 1;
    }
    catch ( ... )
    {
        g_last_error = "unknown error";
        return 1;
    }
}

void require( bool cond, const char* msg )
{
    if ( !cond )
        throw Error( msg );
}

bool isContact( const pdk_model_desc& d )
{
    return d.model == PDK_MODEL_HERTZ || d.model == PDK_MODEL_HERTZ_JKR;
}
// LinearLPS / LinearPMB derive from the LPS / PMB force models
// (force_models/CabanaPD_LPS.hpp:255-273, force_models/CabanaPD_PMB.hpp:575-597).
bool isLpsModel( int32_t m )
{
    return m == PDK_MODEL_LPS || m == PDK_MODEL_LINEAR_LPS;
}
bool isLinearModel( int32_t m )
{
    return m == PDK_MODEL_LINEAR_PMB || m == PDK_MODEL_LINEAR_LPS;
}

// Derived constants with the reference's expressions.
void modelConstants( const pdk_model_desc& d, pdk_model_constants& k )
{
    std::memset( &k, 0, sizeof( k ) );
    const double delta = d.horizon;
    const double K = d.K;
    if ( d.model == PDK_MODEL_PMB || d.model == PDK_MODEL_LPS || isLinearModel( d.model ) )
    {
        require( delta > 0 && K > 0, "horizon and K must be positive" );
        // Force<LinearPMB|LinearLPS> exists for NoFracture only
        // (force/CabanaPD_PMB.hpp:406-409, force/CabanaPD_LPS.hpp:577-580)
        require( !( isLinearModel( d.model ) && d.fracture ), "the linearized models have no fracture form" );
        // force_models/CabanaPD_PMB.hpp:44-46
        k.c = 18.0 * K / ( kPi * delta * delta * delta * delta );
        if ( d.fracture )
        {
            require( d.G0 > 0, "fracture models need G0 > 0" );
            // CabanaPD_ForceModels.hpp:97-101; LPS override force_models/CabanaPD_LPS.hpp:214-222
            k.s0 = std::sqrt( 5.0 * d.G0 / 9.0 / K / delta );
            if ( d.influence == 0 )
                k.s0 = std::sqrt( 8.0 * d.G0 / 15.0 / K / delta );
            k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
        }
        // force_models/CabanaPD_LPS.hpp:75-82
        k.theta_coeff = 3.0 * K - 5.0 * d.G;
        k.s_coeff = 15.0 * d.G;
        if ( isLpsModel( d.model ) )
            require( d.influence == 0 || d.influence == 1, "Influence function type must be 0 or 1." );
        require( d.mechanics == PDK_MECH_ELASTIC || d.mechanics == PDK_MECH_EPP, "unknown mechanics" );
        require( d.thermal == PDK_TEMP_INDEPENDENT || d.thermal == PDK_TEMP_DEPENDENT || d.thermal == PDK_TEMP_DYNAMIC,
                 "unknown thermal type" );
        if ( d.thermal == PDK_TEMP_DYNAMIC )
        {
            require( d.kappa > 0 && d.cp > 0, "DynamicTemperature needs kappa > 0 and cp > 0" );
            require( d.thermal_subcycle >= 1, "thermal_subcycle must be >= 1" );
            // BaseDynamicTemperatureModel (CabanaPD_ForceModels.hpp:277-288),
            // thermal horizon = force horizon
            const double d3 = delta * delta * delta;
            k.thermal_coeff = 9.0 / 2.0 * d.kappa / kPi / d3;
        }
        if ( d.mechanics == PDK_MECH_EPP )
        {
            // ForceModel<PMB, ElasticPerfectlyPlastic, Fracture> only
            // (force_models/CabanaPD_PMB.hpp:232-284; tstForce.hpp:1048)
            require( d.model == PDK_MODEL_PMB && d.fracture, "ElasticPerfectlyPlastic exists for PMB with fracture only" );
            require( d.sigma_y > 0, "ElasticPerfectlyPlastic needs sigma_y > 0" );
            // force_models/CabanaPD_PMB.hpp:90-96
            k.s_Y = d.sigma_y / 3.0 / K;
            if ( d.thermal != PDK_TEMP_DEPENDENT )
            {
                // force_models/CabanaPD_PMB.hpp:255-264
                k.s0 = ( 5.0 * d.G0 / d.sigma_y / delta + d.sigma_y / K ) / 6.0;
                k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
            }
        }
        if ( d.thermal != PDK_TEMP_INDEPENDENT )
        {
            // PMB and LinearPMB (force_models/CabanaPD_PMB.hpp:306-568, 575-597)
            require( d.model == PDK_MODEL_PMB || d.model == PDK_MODEL_LINEAR_PMB,
                     "TemperatureDependent exists for the PMB models only" );
            if ( d.fracture )
            {
                // ThermalFractureModel: BaseFractureModel( delta, K, G0 ),
                // influence_type 1 (CabanaPD_ForceModels.hpp:228-236, 93-102)
                k.s0 = std::sqrt( 5.0 * d.G0 / 9.0 / K / delta );
                k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
            }
        }
    }
    else if ( d.mechanics != PDK_MECH_ELASTIC || d.thermal != PDK_TEMP_INDEPENDENT )
        throw Error( "the contact models are elastic and temperature independent" );
    else if ( d.model == PDK_MODEL_HERTZ )
    {
        require( d.radius > 0 && d.E > 0 && d.restitution > 0, "Hertz needs radius, E, restitution > 0" );
        // force_models/CabanaPD_Hertzian.hpp:46-55
        k.Rs = 0.5 * d.radius;
        k.Es = d.E / ( 2.0 * ( 1.0 - std::pow( d.nu, 2.0 ) ) );
        double ln_e = std::log( d.restitution );
        k.beta = -ln_e / std::sqrt( std::pow( ln_e, 2.0 ) + std::pow( kPi, 2.0 ) );
        k.coeff_h_n = 4.0 / 3.0 * k.Es * std::sqrt( k.Rs );
        k.coeff_h_d = -2.0 * std::sqrt( 5.0 / 6.0 ) * k.beta;
    }
    else if ( d.model == PDK_MODEL_HERTZ_JKR )
    {
        require( d.radius > 0 && d.E > 0 && d.restitution > 0, "JKR needs radius, E, restitution > 0" );
        // force_models/CabanaPD_HertzianJKR.hpp:59-79
        k.Rs = 0.5 * d.radius;
        k.Es = d.E / ( 2.0 * ( 1.0 - std::pow( d.nu, 2.0 ) ) );
        k.Gs = d.E / ( 4.0 * ( 2 - d.nu ) * ( 1 + d.nu ) );
        double ln_e = std::log( d.restitution );
        k.beta = -ln_e / std::sqrt( std::pow( ln_e, 2.0 ) + std::pow( kPi, 2.0 ) );
        k.coeff_h_n = 4.0 / 3.0 * k.Es / k.Rs;
        k.coeff_h_d = -2.0 * std::sqrt( 5.0 / 6.0 ) * k.beta;
        k.Gamma = 2.0 * d.gamma;
        k.a0 = std::pow( 9.0 / 2.0 * k.Gamma * kPi / k.Es * std::pow( k.Rs, 2.0 ), 1.0 / 3.0 );
        k.delta_tear = -1.0 / 2.0 * std::pow( 6.0, -1.0 / 3.0 ) * std::pow( k.a0, 2.0 ) / k.Rs;
        k.fc = 3.0 / 2.0 * kPi * k.Rs * k.Gamma;
    }
    else
        throw Error( "unknown model kind" );
}

enum KernelClass
{
    kClsForce = 0,
    kClsLps = 1,
    kClsEnergy = 2,
    kClsBuild = 3,
    kClsOther = 4,
    kNumCls = 5
};

} // namespace

struct pdk_ctx_s
{
    int device = 0;
    MemAccount acct;
    std::unique_ptr<BatchEngine> batch;
    std::atomic<int64_t> live_specs{ 0 };
};

struct pdk_series_s;

struct pdk_spec_s
{
    pdk_ctx ctx = nullptr;
    pdk_model_desc d{};
    pdk_model_constants k{};
    int64_t n = 0, nslices = 0;
    hipStream_t st = nullptr;
    DevBuf x, U0, U1, v, f, vol, rho, nofail, bcflag, body, m, theta, lpsA, lpsB, W, phi, perm, inv, cnt,
        slice_ptr, col, intact, radius, u_neigh, partial, args, notch, staging, sid, stab, stab_xi2d, stab_xiv, stress, mtype,
        splas,
        temp, cond;
    // Prescribed values (force_update == false conditions): members (device
    // ids), a and b ([count][ncomp]); one set for U and one for TEMPERATURE.
    struct Prescribed
    {
        DevBuf idx, a, b;
        int64_t count = 0;
        int32_t ncomp = 0;
        bool has_b = false;
    };
    Prescribed pre_u, pre_temp;
    // PD specimen with a contact model (pdk_set_contact): its contact list on
    // current positions, the displacement at the last build and a scratch
    // pair for the largest displacement since then
    struct PdContact
    {
        bool on = false;
        pdk_contact_desc cd{};
        double c = 0.0;
        DevBuf slice_ptr, cnt, col, intact, u_neigh, partial;
        BuildScratch scratch;
        int64_t rows_cap = 0, total_rows = 0;
        int32_t max_cnt = 0;
    } pdc;
    // multi-material (pdk_set_materials): materials and pair tables
    int32_t nmat = 1;
    bool pair_tables = false; // particle types and pair tables on the device
    double mat_c[16] = {}, mat_bcc[16] = {}, mat_tc[4] = {}, mat_sc[4] = {};
    double lat_h[3] = { 0, 0, 0 }; // lattice spacing when lattice_rows
    int64_t lat_dims[3] = { 0, 0, 0 };
    double lat_dev = 0.0;          // largest distance of a position from its lattice site / spacing
    int32_t nsten = 0;             // stencil entries (0 = no stencil)
    SlotForm slot;                 // stencil-slot form (slot.slots == 0: compact form)
    std::vector<int64_t> h_perm, h_inv;
    int64_t layout_version = 0;            // bumped whenever the device order changes
    std::vector<pdk_series_s*> series_list; // live region series on this spec
    PDArgs h{};
    int32_t cur = 0;
    bool uploaded = false, built = false, inited = false;
    bool lattice_rows = false; // device order is lattice rows (else Morton)
    int32_t unroll = -1;       // force-pass variant (tuning knob): -1 default, 0 row windows, 1/2/4/8 gathers
    bool vol_uniform = false;  // every particle has volume vol_value
    double vol_value = 0.0;
    bool any_nofail = false;
    int64_t col_rows_cap = 0, total_rows = 0, total_bonds = 0;
    int32_t max_cnt = 0;
    bool stats_dirty = false; // contact list rebuilt since total_bonds/max_cnt were computed
    int64_t steps_done = 0;
    // contact bookkeeping
    int64_t steps_since_build = 0;
    double max_disp = 0.0;
    // contact step: pinned host copy of the largest displacement since the
    // last build, and the event after its copy (the host waits on that, not
    // on the stream)
    double* disp_pinned = nullptr;
    hipEvent_t disp_ready = nullptr;
    double radius_max = 0.0; // largest per-particle radius (polydisperse contact)
    // timing
    bool timing = false;
    struct Ev
    {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Ev> evs;
    double tsum[kNumCls] = {};
    int64_t tcnt[kNumCls] = {};
    BuildScratch scratch;
    // serving
    bool in_batch = false;

    MemAccount* acct() { return &ctx->acct; }

    // Specs whose force passes need the FP64 gather variant: several
    // materials, EPP (per-bond state) or a temperature field.
    bool gatherOnly() const { return pair_tables || d.mechanics == PDK_MECH_EPP || d.thermal != PDK_TEMP_INDEPENDENT; }
    // the row-window force pass covers the specimen (with its features, if any)
    bool rowsCover() const
    {
        return !gatherOnly() || rowFeaturesSupported( d.model, pair_tables, d.thermal != PDK_TEMP_INDEPENDENT,
                                                      d.mechanics == PDK_MECH_EPP );
    }
    bool hasPrescribed() const { return pre_u.count > 0 || pre_temp.count > 0; }
    bool dynamicTemp() const { return d.thermal == PDK_TEMP_DYNAMIC; }
    // Work between the initial half step and the force pass of step g:
    // heat transfer every thermal_subcycle steps (CabanaPD_Solver.hpp:340-347),
    // then the prescribed values at t = g dt (:350-352).
    bool hasPreForce() const { return hasPrescribed() || dynamicTemp(); }
    void preForce( int64_t g )
    {
        if ( dynamicTemp() && g % d.thermal_subcycle == 0 )
            timed( kClsOther, [&] { launchHeatTransfer( dargs(), h, d.thermal_subcycle * d.dt, st ); } );
        if ( hasPrescribed() )
            applyPrescribed( static_cast<double>( g ) * d.dt );
    }
    // The prescribed values at time t, before a force pass.
    void applyPrescribed( double t )
    {
        if ( pre_u.count )
            launchPrescribed( pre_u.idx.as<int64_t>(), pre_u.count, pre_u.a.as<double>(),
                              pre_u.has_b ? pre_u.b.as<double>() : nullptr, 3, t, Ucur(), n, st );
        if ( pre_temp.count )
            launchPrescribed( pre_temp.idx.as<int64_t>(), pre_temp.count, pre_temp.a.as<double>(),
                              pre_temp.has_b ? pre_temp.b.as<double>() : nullptr, 1, t, temp.as<double>(), n, st );
    }

    // Two consecutive middle steps of pdk_step captured as one HIP graph per
    // starting displacement buffer (after two steps the buffer index is back
    // where it started).  Dropped whenever the launch configuration may have
    // changed.
    hipGraphExec_t step_graph[2] = { nullptr, nullptr };
    // Off by default: on the 41^3 cube (30-80 us steps) eager launches
    // queued on the stream measured 4-8 % faster than graph replays.
    bool use_graph = false;
    void dropGraphs()
    {
        for ( auto& g : step_graph )
            if ( g )
            {
                (void)hipGraphExecDestroy( g );
                g = nullptr;
            }
    }
    // The launches of one middle step (initial half step of the next step
    // fused into the force epilogue), displacement buffer c.
    void middleStep( int32_t c )
    {
        if ( lps() )
            launchLpsState( dargs(), h, d.fracture != 0, true, c, st );
        launchForce( dargs(), h, kEpiFinalInitial, c, st );
    }
    hipGraphExec_t stepPairGraph( int32_t c )
    {
        if ( !step_graph[c] )
        {
            hipGraph_t g = nullptr;
            PDK_HIP( hipStreamBeginCapture( st, hipStreamCaptureModeThreadLocal ) );
            middleStep( c );
            middleStep( c ^ 1 );
            PDK_HIP( hipStreamEndCapture( st, &g ) );
            const hipError_t e = hipGraphInstantiate( &step_graph[c], g, nullptr, nullptr, 0 );
            (void)hipGraphDestroy( g );
            PDK_HIP( e );
        }
        return step_graph[c];
    }

    void pushArgs()
    {
        dropGraphs();
        h.n = n;
        h.nslices = nslices;
        h.x = x.as<double>();
        h.U[0] = U0.as<double>();
        h.U[1] = U1.as<double>();
        h.v = v.as<double>();
        h.f = f.as<double>();
        h.vol = vol.as<double>();
        h.rho = rho.as<double>();
        h.nofail = nofail.as<uint8_t>();
        h.bcflag = bcflag.as<uint8_t>();
        h.body = body ? body.as<double>() : nullptr;
        h.slice_ptr = slice_ptr.as<int64_t>();
        h.cnt = cnt.as<int32_t>();
        h.col = col.as<int32_t>();
        h.intact = intact.as<uint64_t>();
        h.m = m.as<double>();
        h.theta = theta.as<double>();
        h.lpsA = lpsA.as<double>();
        h.lpsB = lpsB.as<double>();
        h.W = W.as<double>();
        h.phi = phi.as<double>();
        h.stress = stress.as<double>();
        h.radius = radius ? radius.as<double>() : nullptr;
        h.u_neigh = u_neigh ? u_neigh.as<double>() : nullptr;
        h.sid = sid ? sid.as<uint8_t>() : nullptr;
        h.stencil = stab ? stab.as<float4>() : nullptr;
        h.stencil_xi2d = stab_xi2d ? stab_xi2d.as<double>() : nullptr;
        h.stencil_xiv = stab_xiv ? stab_xiv.as<double>() : nullptr;
        h.exist = nullptr;
        h.soff = nullptr;
        h.stencil_slots = 0;
        if ( slot.slots > 0 )
        {
            h.sid = nullptr;
            h.stencil = slot.table.as<float4>();
            h.stencil_xi2d = slot.xi2d.as<double>();
            h.stencil_xiv = slot.xiv.as<double>();
            h.exist = slot.exist.as<uint64_t>();
            h.soff = slot.soff.as<int32_t>();
            h.stencil_slots = slot.slots;
            h.rows = slot.nrows > 0 ? slot.rows.as<int4>() : nullptr;
            h.rowstart = slot.nrows > 0 ? slot.rowstart.as<uint32_t>() : nullptr;
            h.nrows = slot.nrows;
            for ( int dd = 0; dd < 3; dd++ )
                h.lat_n[dd] = static_cast<int32_t>( lat_dims[dd] );
        }
        else
        {
            h.rows = nullptr;
            h.rowstart = nullptr;
            h.nrows = 0;
        }
        // tuning knob: -1 = default (row windows where available, else groups
        // of 4, or of 2 for the mixed form in the compact layout, measured
        // faster there); feature combinations the row windows do not cover
        // use the gather variant
        const int32_t gather_unroll = ( d.precision == PDK_MIXED && slot.slots == 0 ) ? 2 : 4;
        h.unroll = unroll;
        if ( h.unroll < 0 )
            h.unroll = ( h.nrows > 0 && rowsCover() ) ? 0 : gather_unroll;
        if ( h.unroll == 0 && ( h.nrows == 0 || !rowsCover() ) )
            h.unroll = gather_unroll;
        h.splas = splas ? splas.as<double>() : nullptr;
        h.sY = k.s_Y;
        h.temp = temp ? temp.as<double>() : nullptr;
        h.alpha = d.alpha;
        h.temp0 = d.temp0;
        h.cond = cond ? cond.as<double>() : nullptr;
        h.thermal_coeff = k.thermal_coeff;
        h.cp = d.cp;
        h.thermal_horizon = d.horizon;
        h.lin_micro = d.linear_microconductivity;
        h.nmat = nmat;
        h.mtype = pair_tables ? mtype.as<uint8_t>() : nullptr;
        // Mixed form on a lattice: bond vectors come from the stencil (k h),
        // which differs from x_j - x_i by up to |e| = 2 lat_dev |h| (|h| the
        // spacing vector's length); the stretch then moves by at most about
        // (2 + s) |e| / |xi| with |xi| >= min h, and the band grows by twice
        // that bound.
        double band_lat = 0.0;
        if ( slot.slots > 0 || sid )
        {
            const double hmin = std::min( lat_h[0], std::min( lat_h[1], lat_h[2] ) );
            const double hn = std::sqrt( lat_h[0] * lat_h[0] + lat_h[1] * lat_h[1] + lat_h[2] * lat_h[2] );
            band_lat = hmin > 0.0 ? 4.0 * lat_dev * hn / hmin : 0.0;
        }
        for ( int t = 0; t < 16; t++ )
        {
            h.mat_c[t] = mat_c[t];
            h.mat_bcc[t] = mat_bcc[t];
            // s0 of the pair from bond_break_coeff = (1 + s0)^2; a pair that
            // never breaks (+inf) gets no band, so its FP32 test is final
            const double s0p = std::sqrt( mat_bcc[t] ) - 1.0;
            h.mat_s0f[t] = static_cast<float>( s0p );
            h.mat_bandK[t] = std::isfinite( s0p ) ? static_cast<float>( 2.44140625e-4 * s0p + 9.094947017729282e-13 +
                                                                        ( 2.0 + s0p ) * band_lat )
                                                  : 0.0f;
        }
        for ( int t = 0; t < 4; t++ )
        {
            h.mat_tc[t] = mat_tc[t];
            h.mat_sc[t] = mat_sc[t];
        }
        h.c = k.c;
        h.bcc = k.bond_break_coeff;
        h.s0 = k.s0;
        h.theta_coeff = k.theta_coeff;
        h.s_coeff = k.s_coeff;
        h.dt = d.dt;
        h.half_dt = 0.5 * d.dt;
        h.s0f = static_cast<float>( k.s0 );
        // The mixed stretch comes from an FP64 numerator |r|^2 - |xi|^2 and a
        // few FP32 operations: its error is ~10 ulp of |s| (2^-20 |s|).  A
        // bond within 2^-12 s0 (+ 2^-40 for s0 = 0) of s0 is re-tested in
        // FP64 -- a margin of 2^8 over that error where it matters (|s| near
        // s0; far from s0 the FP32 decision cannot be wrong).  The gather
        // passes scale bandf = 2^-13 by the bond's own terms (thermal strain).
        h.bandf = 1.220703125e-4f;
        h.bandK = static_cast<float>( 2.44140625e-4 * k.s0 + 9.094947017729282e-13 + ( 2.0 + k.s0 ) * band_lat );
        // uniform volume: c vol (PMB) or vol (LPS) scales each FP32 row sum
        h.row_scale = isLpsModel( d.model ) ? vol_value : k.c * vol_value;
        h.Rs = k.Rs;
        h.Es = k.Es;
        h.coeff_h_n = k.coeff_h_n;
        h.coeff_h_d = k.coeff_h_d;
        h.Gamma = k.Gamma;
        h.a0 = k.a0;
        h.delta_tear = k.delta_tear;
        h.jkr_c6 = std::pow( 6.0, -1.0 / 3.0 );
        h.two_r = 2.0 * d.radius;
        h.mu_f = d.mu_friction;
        h.nu = d.nu;
        h.E = d.E;
        h.model = d.model;
        h.fracture = d.fracture;
        h.influence = d.influence;
        h.precision = d.precision;
        h.polydisperse = d.polydisperse;
        h.vol_uniform = vol_uniform ? 1 : 0;
        h.vol_value = vol_value;
        h.any_nofail = any_nofail ? 1 : 0;
        h.stencil_aux = nullptr;
        h.stencil_rec = nullptr;
        if ( slot.slots > 0 )
        {
            if ( !slot.aux )
            {
                slot.aux.alloc( static_cast<size_t>( slot.slots ) * sizeof( float4 ), acct() );
                launchStencilAux( slot.table.as<float4>(), slot.slots, d.influence, slot.aux.as<float4>(), st );
            }
            if ( !slot.rec )
            {
                slot.rec.alloc( static_cast<size_t>( slot.slots ) * kRecStride * sizeof( double ), acct() );
                launchStencilRec( slot.table.as<float4>(), slot.aux.as<float4>(), slot.xiv.as<double>(), slot.slots,
                                  slot.rec.as<double>(), st );
            }
            h.stencil_aux = slot.aux.as<float4>();
            h.stencil_rec = slot.rec.as<double>();
        }
        PDK_HIP( hipMemcpyAsync( args.as<PDArgs>(), &h, sizeof( PDArgs ), hipMemcpyHostToDevice, st ) );
    }

    const PDArgs* dargs() const { return args.as<PDArgs>(); }

    template <class F>
    void timed( int cls, F&& fn )
    {
        if ( !timing )
        {
            fn();
            return;
        }
        Ev e;
        e.cls = cls;
        PDK_HIP( hipEventCreate( &e.a ) );
        PDK_HIP( hipEventCreate( &e.b ) );
        PDK_HIP( hipEventRecord( e.a, st ) );
        fn();
        PDK_HIP( hipEventRecord( e.b, st ) );
        evs.push_back( e );
    }

    void resolveEvents()
    {
        if ( evs.empty() )
            return;
        PDK_HIP( hipStreamSynchronize( st ) );
        for ( auto& e : evs )
        {
            float ms = 0.f;
            PDK_HIP( hipEventElapsedTime( &ms, e.a, e.b ) );
            tsum[e.cls] += ms;
            tcnt[e.cls] += 1;
            (void)hipEventDestroy( e.a );
            (void)hipEventDestroy( e.b );
        }
        evs.clear();
    }

    void stageUpload( const void* src, size_t bytes )
    {
        if ( staging.bytes() < bytes )
            staging.alloc( bytes, acct() );
        PDK_HIP( hipMemcpyAsync( staging.as<void>(), src, bytes, hipMemcpyHostToDevice, st ) );
    }

    void uploadAos3( const double* src, double* dst_soa )
    {
        stageUpload( src, 3 * n * sizeof( double ) );
        launchGatherAos3( staging.as<double>(), perm.as<int64_t>(), dst_soa, n, st );
    }
    void uploadScalar( const double* src, double* dst )
    {
        stageUpload( src, n * sizeof( double ) );
        launchGatherScalar( staging.as<double>(), perm.as<int64_t>(), dst, n, st );
    }
    void fillScalar( double* dst, double value )
    {
        std::vector<double> hv( n, value );
        PDK_HIP( hipMemcpyAsync( dst, hv.data(), n * sizeof( double ), hipMemcpyHostToDevice, st ) );
        PDK_HIP( hipStreamSynchronize( st ) );
    }

    double* Ucur() { return cur ? U1.as<double>() : U0.as<double>(); }

    // Device array of a pdk_field, [ncomp][n] in device order; null when the
    // spec does not hold that field.
    const double* fieldDev( int32_t field, int32_t* ncomp )
    {
        *ncomp = 1;
        switch ( field )
        {
        case PDK_FIELD_X: *ncomp = 3; return x.as<double>();
        case PDK_FIELD_U: *ncomp = 3; return Ucur();
        case PDK_FIELD_V: *ncomp = 3; return v.as<double>();
        case PDK_FIELD_F: *ncomp = 3; return f.as<double>();
        case PDK_FIELD_STRESS: *ncomp = 9; return stress.as<double>();
        case PDK_FIELD_VOL: return vol.as<double>();
        case PDK_FIELD_RHO: return rho.as<double>();
        case PDK_FIELD_W: return W.as<double>();
        case PDK_FIELD_PHI: return phi.as<double>();
        case PDK_FIELD_M: return m.as<double>();
        case PDK_FIELD_THETA: return theta.as<double>();
        case PDK_FIELD_TEMPERATURE: return temp.as<double>();
        default: throw Error( "unknown field" );
        }
    }

    // Bond count and longest row from the device row lengths.
    void refreshStats()
    {
        std::vector<int32_t> c( n );
        PDK_HIP( hipMemcpyAsync( c.data(), cnt.as<int32_t>(), n * sizeof( int32_t ), hipMemcpyDeviceToHost, st ) );
        PDK_HIP( hipStreamSynchronize( st ) );
        int64_t t = 0;
        int32_t mx = 0;
        for ( int32_t v : c )
        {
            t += v;
            mx = std::max( mx, v );
        }
        total_bonds = t;
        max_cnt = mx;
        stats_dirty = false;
    }

    // ---- PD orchestration ---------------------------------------------------
    bool lps() const { return isLpsModel( d.model ); }

    void lpsState( bool wv, bool dil )
    {
        if ( !lps() )
            return;
        timed( kClsLps, [&] { launchLpsState( dargs(), h, wv, dil, cur, st ); } );
    }

    // Contact list of a PD specimen on the current positions x + U: cutoff
    // 2 radius + radius_extend (+ tol at the first build), neighbors in
    // ascending caller id; u_neigh = U.
    void pdContactBuild( double tol )
    {
        NeighborQuery q{};
        q.x = x.as<double>();
        q.u = Ucur();
        q.radius = nullptr;
        q.rc = 2.0 * pdc.cd.radius + pdc.cd.radius_extend + tol;
        q.extend = pdc.cd.radius_extend;
        q.n = n;
        // after the first build: one capped pass (see contactBuildImpl)
        const int32_t cap = tol > 0.0 || pdc.max_cnt <= 0 ? 0 : ( ( pdc.max_cnt + 4 + 7 ) / 8 ) * 8;
        pdc.total_rows = buildSlicedCsr( q, perm.as<int64_t>(), inv.as<int64_t>(), pdc.slice_ptr.as<int64_t>(),
                                         pdc.cnt.as<int32_t>(), pdc.col, pdc.intact, pdc.rows_cap, pdc.scratch, acct(),
                                         st, &pdc.max_cnt, cap );
        PDK_HIP( hipMemcpyAsync( pdc.u_neigh.as<double>(), Ucur(), 3 * n * sizeof( double ), hipMemcpyDeviceToDevice,
                                 st ) );
    }
    PdContactArgs pdContactArgs() const
    {
        return PdContactArgs{ pdc.slice_ptr.as<int64_t>(), pdc.cnt.as<int32_t>(), pdc.col.as<int32_t>(),
                              pdc.cd.radius, pdc.cd.horizon, pdc.c };
    }

    // Solver::updateForce (CabanaPD_Solver.hpp:462-477) without the epilogue.
    void updateForceStore()
    {
        if ( lps() )
            lpsState( d.fracture != 0, true );
        timed( kClsForce, [&] { launchForce( dargs(), h, kEpiStoreOnly, cur, st ); } );
    }
};

// ---- contact orchestration lives in pdk_contact_api.hip ------------------------
namespace pdk
{
void contactBuild( pdk_spec s, bool initial );
void contactStep( pdk_spec s, int32_t nsteps );
void contactForceStore( pdk_spec s );
} // namespace pdk

extern "C" {

const char* pdk_last_error( void )
{
    return g_last_error.c_str();
}

int pdk_model_constants_get( const pdk_model_desc* desc, pdk_model_constants* out )
{
    return guarded( [&] {
        require( desc && out, "null argument" );
        modelConstants( *desc, *out );
    } );
}

int pdk_critical_timestep( double dx, double dy, double dz, double K, double rho, double horizon,
                           double safety_factor, double* dt_out )
{
    return guarded( [&] {
        require( dt_out && dx > 0 && dy > 0 && dz > 0 && horizon >= dx, "invalid critical time step input" );
        // Inputs::computeCriticalTimeStep (CabanaPD_Input.hpp:217-336); m from
        // the Inputs ctor (CabanaPD_Input.hpp:71-83).
        double v_p = dx * dy * dz;
        double sum = 0;
        int mm = static_cast<int>( std::floor( horizon / dx ) );
        double c = 18.0 * K / ( kPi * horizon * horizon * horizon * horizon );
        for ( int i = -( mm + 1 ); i < mm + 2; i++ )
        {
            double xi_1 = i * dx;
            for ( int j = -( mm + 1 ); j < mm + 2; j++ )
            {
                double xi_2 = j * dy;
                for ( int kk = -( mm + 1 ); kk < mm + 2; kk++ )
                {
                    double xi_3 = kk * dz;
                    double r2 = xi_1 * xi_1 + xi_2 * xi_2 + xi_3 * xi_3;
                    if ( r2 < horizon * horizon + 1e-10 && r2 > 0 )
                        sum += v_p * c / std::sqrt( r2 );
                }
            }
        }
        *dt_out = safety_factor * std::sqrt( 2.0 * rho / sum );
    } );
}

} // extern "C"

// ---- fixed device scratch of pdk_compute_energy_noalloc ------------------------
// Per-block partial totals of the energy pass (2 doubles per block of 4
// slices) followed by the two final totals.  Every GPU the library runs on
// holds one copy until the process ends; all specs on that GPU share it.
namespace
{
constexpr int64_t kNoAllocMaxBlocks = 65536; // up to 65536 * 256 = 16,777,216 particles
constexpr int64_t kNoAllocScratchDoubles = 2 * kNoAllocMaxBlocks + 2;
static __device__ __align__( 256 ) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];

// Device address of this GPU's copy of scratch, looked up once per device
// with that device current (the host-side name is not a device address).
void initScratch( int device )
{
    require( device >= 0 && device < kMaxDevices, "device index beyond the scratch table" );
    std::call_once( g_scratch_once[device], [device] {
        void* p = nullptr;
        PDK_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
        g_scratch_ptr[device] = static_cast<double*>( p );
    } );
    require( g_scratch_ptr[device] != nullptr, "scratch address lookup failed" );
}

// Sums the per-block partials in block order -- the order of the host loop of
// pdk_compute_energy -- so both entry points return the same bits.
__global__ void kSumPartialsInOrder( double* __restrict__ part, int64_t nblocks )
{
#pragma clang fp contract( off )
    if ( threadIdx.x != 0 || blockIdx.x != 0 )
        return;
    double Phi = 0.0, D = 0.0;
    for ( int64_t b = 0; b < nblocks; b++ )
    {
        Phi += part[2 * b];
        D += part[2 * b + 1];
    }
    part[2 * nblocks] = Phi;
    part[2 * nblocks + 1] = D;
}
} // namespace

extern "C" {

int pdk_ctx_create( int device, uint64_t hbm_budget, pdk_ctx* out )
{
    return guarded( [&] {
        require( out != nullptr, "null ctx out" );
        int count = 0;
        PDK_HIP( hipGetDeviceCount( &count ) );
        require( device >= 0 && device < count, "no such HIP device" );
        PDK_HIP( hipSetDevice( device ) );
        initScratch( device );
        auto c = std::make_unique<pdk_ctx_s>();
        c->device = device;
        size_t freeb = 0, total = 0;
        PDK_HIP( hipMemGetInfo( &freeb, &total ) );
        c->acct.budget = hbm_budget ? hbm_budget : static_cast<uint64_t>( 0.9 * freeb );
        c->batch = std::make_unique<BatchEngine>( c.get() );
        *out = c.release();
    } );
}

int pdk_ctx_destroy( pdk_ctx ctx )
{
    return guarded( [&] {
        if ( ctx )
        {
            require( ctx->live_specs == 0, "destroy every spec of a context before the context" );
            ctx->batch.reset();
            delete ctx;
        }
    } );
}

int pdk_ctx_bytes_in_use( pdk_ctx ctx, uint64_t* bytes )
{
    return guarded( [&] {
        require( ctx && bytes, "null argument" );
        *bytes = ctx->acct.used();
    } );
}

int pdk_spec_create( pdk_ctx ctx, const pdk_model_desc* desc, int64_t n, pdk_spec* out )
{
    return guarded( [&] {
        require( ctx && desc && out, "null argument" );
        require( n > 0, "a specimen needs at least one particle" );
        require( n < ( int64_t( 1 ) << 31 ), "at most 2^31-1 particles per specimen" );
        require( desc->precision == PDK_FP64 || desc->precision == PDK_MIXED, "unknown precision" );
        require( !( isContact( *desc ) && desc->precision != PDK_FP64 ), "contact models run in FP64" );
        require( desc->dt > 0, "dt must be positive" );
        PDK_HIP( hipSetDevice( ctx->device ) );
        auto s = std::make_unique<pdk_spec_s>();
        s->ctx = ctx;
        s->d = *desc;
        modelConstants( s->d, s->k );
        s->n = n;
        s->nslices = ( n + kSlice - 1 ) / kSlice;
        PDK_HIP( hipStreamCreateWithFlags( &s->st, hipStreamNonBlocking ) );
        auto* a = s->acct();
        const size_t d3 = 3 * n * sizeof( double ), d1 = n * sizeof( double );
        s->x.alloc( d3, a );
        s->U0.alloc( d3, a );
        s->U1.alloc( d3, a );
        s->v.alloc( d3, a );
        s->f.alloc( d3, a );
        s->vol.alloc( d1, a );
        s->rho.alloc( d1, a );
        s->nofail.alloc( n, a );
        s->bcflag.alloc( n, a );
        s->W.alloc( d1, a );
        s->phi.alloc( d1, a );
        if ( isLpsModel( desc->model ) )
        {
            s->m.alloc( d1, a );
            s->theta.alloc( d1, a );
            s->lpsA.alloc( d1, a );
            s->lpsB.alloc( d1, a );
        }
        s->perm.alloc( n * sizeof( int64_t ), a );
        s->inv.alloc( n * sizeof( int64_t ), a );
        s->cnt.alloc( n * sizeof( int32_t ), a );
        s->slice_ptr.alloc( ( s->nslices + 1 ) * sizeof( int64_t ), a );
        s->partial.alloc( std::max<int64_t>( 2 * ( s->nslices / kSlicesPerBlock + 2 ), 64 ) * sizeof( double ), a );
        s->args.alloc( sizeof( PDArgs ), a );
        if ( desc->thermal != PDK_TEMP_INDEPENDENT )
        {
            // zero at creation (CabanaPD_Particles.hpp:860-864)
            s->temp.alloc( d1, a );
            PDK_HIP( hipMemsetAsync( s->temp.as<double>(), 0, d1, s->st ) );
            if ( desc->thermal == PDK_TEMP_DYNAMIC )
                s->cond.alloc( d1, a );
        }
        if ( isContact( *desc ) )
        {
            s->u_neigh.alloc( d3, a );
            if ( desc->polydisperse )
                s->radius.alloc( d1, a );
        }
        s->h.bc_kind = -1;
        *out = s.release();
        ctx->live_specs++;
    } );
}

// ---- region series -------------------------------------------------------------
struct pdk_series_s
{
    pdk_spec spec = nullptr;
    pdk_quantity q{};
    int32_t red = 0;
    int64_t capacity = 0, used = 0;
    std::vector<int64_t> ids;  // members, ascending caller id
    int64_t version = -1;      // spec layout the device ids were made for
    DevBuf dev_ids, partial, values;
    std::vector<double> times;

    // Device ids of the members in the spec's current order.
    void mapIds()
    {
        if ( version == spec->layout_version )
            return;
        std::vector<int64_t> d( ids.size() );
        for ( size_t k = 0; k < ids.size(); k++ )
            d[k] = spec->h_inv[ids[k]];
        if ( dev_ids.bytes() < d.size() * sizeof( int64_t ) )
            dev_ids.alloc( d.size() * sizeof( int64_t ), spec->acct() );
        if ( !d.empty() )
        {
            PDK_HIP( hipMemcpyAsync( dev_ids.as<int64_t>(), d.data(), d.size() * sizeof( int64_t ),
                                     hipMemcpyHostToDevice, spec->st ) );
            PDK_HIP( hipStreamSynchronize( spec->st ) );
        }
        version = spec->layout_version;
    }

    // Device pointers of the quantity as of now (the displacement buffer
    // alternates between steps).
    QuantityArgs args()
    {
        QuantityArgs Q{};
        int32_t nc = 1;
        const double* src = spec->fieldDev( q.field, &nc );
        require( src != nullptr, "the series' field is not held by this spec" );
        Q.src = src + q.component * spec->n;
        if ( q.gate_field >= 0 )
        {
            Q.gate = spec->fieldDev( q.gate_field, &nc );
            require( Q.gate != nullptr, "the series' gate field is not held by this spec" );
            Q.threshold = q.gate_threshold;
        }
        if ( q.weight_volume )
            Q.vol = spec->vol.as<double>();
        return Q;
    }
};

namespace
{
// The spec is going away: its series keep their host data but lose their
// device buffers (accounted to the spec's context) and refuse further work.
void detachSeries( pdk_spec s )
{
    for ( pdk_series_s* ser : s->series_list )
    {
        ser->dev_ids.release();
        ser->partial.release();
        ser->values.release();
        ser->spec = nullptr;
    }
    s->series_list.clear();
}
} // namespace

int pdk_spec_destroy( pdk_spec s )
{
    return guarded( [&] {
        if ( !s )
            return;
        if ( s->in_batch && s->ctx && s->ctx->batch )
            s->ctx->batch->forget( s );
        s->resolveEvents();
        if ( s->st )
            (void)hipStreamSynchronize( s->st );
        s->dropGraphs();
        detachSeries( s );
        if ( s->disp_ready )
            (void)hipEventDestroy( s->disp_ready );
        if ( s->disp_pinned )
            (void)hipHostFree( s->disp_pinned );
        if ( s->st )
            (void)hipStreamDestroy( s->st );
        if ( s->ctx )
            s->ctx->live_specs--;
        delete s;
    } );
}

int pdk_spec_upload( pdk_spec s, const double* x, const double* u, const double* v, const double* vol,
                     const double* rho, const int32_t* nofail, const double* radius )
{
    return guarded( [&] {
        require( s && x && vol, "x and vol are required" );
        require( !s->in_batch, "spec is running in the batch engine" );
        const int64_t n = s->n;
        // Morton order of the reference positions on a grid of the mean
        // particle spacing (cube root of the mean volume).
        double lo[3] = { x[0], x[1], x[2] };
        double vsum = 0.0;
        for ( int64_t i = 0; i < n; i++ )
        {
            for ( int dd = 0; dd < 3; dd++ )
                lo[dd] = std::min( lo[dd], x[3 * i + dd] );
            vsum += vol[i];
        }
        double h = std::cbrt( vsum / n );
        if ( !( h > 0 ) || !std::isfinite( h ) )
            h = 1.0;
        // Lattice specimens are ordered in lattice rows (x fastest): the 64
        // lanes of a slice then read contiguous runs of x and u at every bond
        // slot.  Other point sets use Morton order.
        std::vector<int64_t> lperm;
        bool lattice = false;
        if ( s->d.ordering != PDK_ORDER_MORTON )
            lattice = latticeOrder( x, n, lperm, s->lat_h, s->lat_dims, &s->lat_dev );
        require( lattice || s->d.ordering != PDK_ORDER_LATTICE_ROWS, "positions do not form a lattice" );
        s->lattice_rows = lattice;
        s->vol_uniform = true;
        s->vol_value = vol[0];
        for ( int64_t i = 1; i < n && s->vol_uniform; i++ )
            s->vol_uniform = vol[i] == vol[0];
        s->any_nofail = false;
        if ( nofail )
            for ( int64_t i = 0; i < n && !s->any_nofail; i++ )
                s->any_nofail = nofail[i] != 0;
        s->stageUpload( x, 3 * n * sizeof( double ) );
        if ( lattice )
        {
            PDK_HIP( hipMemcpyAsync( s->perm.as<int64_t>(), lperm.data(), n * sizeof( int64_t ), hipMemcpyHostToDevice,
                                     s->st ) );
            launchInvertPerm( s->perm.as<int64_t>(), n, s->inv.as<int64_t>(), s->st );
        }
        else
            launchMortonOrder( s->staging.as<double>(), n, lo, 1.0 / h, s->perm.as<int64_t>(), s->inv.as<int64_t>(),
                               s->scratch, s->acct(), s->st );
        s->h_perm.resize( n );
        s->h_inv.resize( n );
        s->layout_version++;
        PDK_HIP( hipMemcpyAsync( s->h_perm.data(), s->perm.as<int64_t>(), n * sizeof( int64_t ), hipMemcpyDeviceToHost,
                                 s->st ) );
        PDK_HIP( hipMemcpyAsync( s->h_inv.data(), s->inv.as<int64_t>(), n * sizeof( int64_t ), hipMemcpyDeviceToHost,
                                 s->st ) );
        launchGatherAos3( s->staging.as<double>(), s->perm.as<int64_t>(), s->x.as<double>(), n, s->st );
        PDK_HIP( hipStreamSynchronize( s->st ) );
        if ( u )
            s->uploadAos3( u, s->U0.as<double>() );
        else
            PDK_HIP( hipMemsetAsync( s->U0.as<double>(), 0, 3 * n * sizeof( double ), s->st ) );
        PDK_HIP( hipMemsetAsync( s->U1.as<double>(), 0, 3 * n * sizeof( double ), s->st ) );
        if ( v )
            s->uploadAos3( v, s->v.as<double>() );
        else
            PDK_HIP( hipMemsetAsync( s->v.as<double>(), 0, 3 * n * sizeof( double ), s->st ) );
        PDK_HIP( hipMemsetAsync( s->f.as<double>(), 0, 3 * n * sizeof( double ), s->st ) );
        s->uploadScalar( vol, s->vol.as<double>() );
        if ( rho )
            s->uploadScalar( rho, s->rho.as<double>() );
        else
            s->fillScalar( s->rho.as<double>(), 1.0 ); // CabanaPD_Particles.hpp:363
        if ( nofail )
        {
            s->stageUpload( nofail, n * sizeof( int32_t ) );
            launchGatherFlags( s->staging.as<int32_t>(), s->perm.as<int64_t>(), s->nofail.as<uint8_t>(), n, s->st );
        }
        else
            PDK_HIP( hipMemsetAsync( s->nofail.as<uint8_t>(), 0, n, s->st ) );
        PDK_HIP( hipMemsetAsync( s->bcflag.as<uint8_t>(), 0, n, s->st ) );
        if ( s->radius )
        {
            require( radius != nullptr, "polydisperse contact needs per-particle radii" );
            s->uploadScalar( radius, s->radius.as<double>() );
            s->radius_max = 0.0;
            for ( int64_t i = 0; i < n; i++ )
                s->radius_max = std::max( s->radius_max, radius[i] );
        }
        PDK_HIP( hipMemsetAsync( s->W.as<double>(), 0, n * sizeof( double ), s->st ) );
        PDK_HIP( hipMemsetAsync( s->phi.as<double>(), 0, n * sizeof( double ), s->st ) );
        if ( s->lps() )
            for ( DevBuf* b : { &s->m, &s->theta, &s->lpsA, &s->lpsB } )
                PDK_HIP( hipMemsetAsync( b->as<double>(), 0, n * sizeof( double ), s->st ) );
        if ( s->u_neigh )
            PDK_HIP( hipMemsetAsync( s->u_neigh.as<double>(), 0, 3 * n * sizeof( double ), s->st ) );
        if ( s->temp )
            PDK_HIP( hipMemsetAsync( s->temp.as<double>(), 0, n * sizeof( double ), s->st ) );
        // prescribed members were mapped to the previous device order
        s->pre_u = pdk_spec_s::Prescribed{};
        s->pre_temp = pdk_spec_s::Prescribed{};
        s->cur = 0;
        s->uploaded = true;
        s->built = false;
        s->inited = false;
        s->steps_done = 0;
        s->pushArgs();
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_spec_set_state( pdk_spec s, const double* u, const double* v )
{
    return guarded( [&] {
        require( s && s->uploaded, "spec not uploaded" );
        require( !s->in_batch, "spec is running in the batch engine" );
        if ( u )
            s->uploadAos3( u, s->Ucur() );
        if ( v )
            s->uploadAos3( v, s->v.as<double>() );
        PDK_HIP( hipMemsetAsync( s->f.as<double>(), 0, 3 * s->n * sizeof( double ), s->st ) );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_spec_copy_state( pdk_spec dst, pdk_spec src )
{
    return guarded( [&] {
        require( dst && src && dst != src, "two distinct specs required" );
        require( dst->ctx == src->ctx, "specs belong to different contexts" );
        require( src->inited && dst->built, "source not initialised or destination bonds not built" );
        require( !src->in_batch && !dst->in_batch, "spec is running in the batch engine" );
        require( !isContact( src->d ) && !isContact( dst->d ) && !src->pdc.on && !dst->pdc.on,
                 "specs with contact lists cannot be copied" );
        require( dst->n == src->n && dst->h_perm == src->h_perm, "specs differ in particle count or device order" );
        require( dst->d.model == src->d.model && dst->d.fracture == src->d.fracture &&
                     dst->d.mechanics == src->d.mechanics && dst->d.thermal == src->d.thermal,
                 "specs differ in model" );
        require( dst->slot.slots == src->slot.slots && dst->total_rows == src->total_rows &&
                     dst->intact.bytes() == src->intact.bytes() && dst->col.bytes() == src->col.bytes(),
                 "specs differ in bond layout" );
        const size_t v3 = 3 * static_cast<size_t>( src->n ) * sizeof( double );
        const size_t v1 = static_cast<size_t>( src->n ) * sizeof( double );
        // src's stream must be done with its state before dst's copies read it
        PDK_HIP( hipStreamSynchronize( src->st ) );
        auto copy = [&]( const DevBuf& a, const DevBuf& b, size_t bytes ) {
            if ( a && b && bytes )
                PDK_HIP( hipMemcpyAsync( a.as<void>(), b.as<void>(), bytes, hipMemcpyDeviceToDevice, dst->st ) );
        };
        PDK_HIP( hipMemcpyAsync( dst->Ucur(), src->Ucur(), v3, hipMemcpyDeviceToDevice, dst->st ) );
        copy( dst->v, src->v, v3 );
        copy( dst->f, src->f, v3 );
        copy( dst->intact, src->intact, src->intact.bytes() );
        copy( dst->W, src->W, v1 );
        copy( dst->phi, src->phi, v1 );
        copy( dst->m, src->m, v1 );
        copy( dst->theta, src->theta, v1 );
        copy( dst->lpsA, src->lpsA, v1 );
        copy( dst->lpsB, src->lpsB, v1 );
        copy( dst->temp, src->temp, v1 );
        if ( dst->splas && src->splas && dst->splas.bytes() == src->splas.bytes() )
            copy( dst->splas, src->splas, src->splas.bytes() );
        dst->steps_done = src->steps_done;
        dst->inited = true;
        PDK_HIP( hipStreamSynchronize( dst->st ) );
    } );
}

int pdk_spec_bytes( pdk_spec s, uint64_t* bytes )
{
    return guarded( [&] {
        require( s && bytes, "null argument" );
        uint64_t t = 0;
        for ( const DevBuf* b : { &s->x, &s->U0, &s->U1, &s->v, &s->f, &s->vol, &s->rho, &s->nofail, &s->bcflag,
                                  &s->body, &s->m, &s->theta, &s->lpsA, &s->lpsB, &s->W, &s->phi, &s->perm, &s->inv,
                                  &s->cnt, &s->slice_ptr, &s->col, &s->intact, &s->radius, &s->u_neigh,
                                  &s->partial, &s->args, &s->sid, &s->stab, &s->stab_xi2d, &s->stab_xiv } )
            t += b->bytes();
        *bytes = t;
    } );
}

int pdk_build_bonds( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->uploaded, "spec not uploaded" );
        require( !s->in_batch, "spec is running in the batch engine" );
        if ( isContact( s->d ) )
        {
            s->timed( kClsBuild, [&] { contactBuild( s, true ); } );
        }
        else
        {
            NeighborQuery q{};
            q.x = s->x.as<double>();
            q.u = nullptr;
            q.radius = nullptr;
            q.rc = s->d.horizon + s->d.neighbor_tol;
            q.extend = 0.0;
            q.n = s->n;
            s->sid.release();
            s->stab.release();
            s->stab_xi2d.release();
            s->stab_xiv.release();
            s->nsten = 0;
            s->slot = SlotForm{};
            // Full lattice: slot q of every row is stencil entry q, so the
            // force pass computes neighbor ids instead of streaming them.
            // Built straight from the lattice where the stencil is clear of
            // the cutoff, else from the cell search below.
            bool direct = false;
            if ( s->lattice_rows && s->d.bond_layout != PDK_BONDS_COMPACT )
                s->timed( kClsBuild, [&] {
                    direct = buildStencilSlotsDirect( s->x.as<double>(), s->n, s->lat_dims, s->lat_h, q.rc, s->h_perm,
                                                      s->perm.as<int64_t>(), s->cnt.as<int32_t>(), s->col, s->intact,
                                                      s->slice_ptr.as<int64_t>(), s->total_rows, s->col_rows_cap,
                                                      s->slot, s->scratch, s->acct(), s->st );
                } );
            if ( direct )
                s->nsten = s->slot.slots;
            else
                s->timed( kClsBuild, [&] {
                    s->total_rows = buildSlicedCsr( q, s->perm.as<int64_t>(), s->inv.as<int64_t>(),
                                                    s->slice_ptr.as<int64_t>(), s->cnt.as<int32_t>(), s->col,
                                                    s->intact, s->col_rows_cap, s->scratch, s->acct(), s->st,
                                                    &s->max_cnt );
                } );
            if ( !direct && s->lattice_rows && s->d.bond_layout != PDK_BONDS_COMPACT )
            {
                s->pushArgs();
                s->timed( kClsBuild, [&] {
                    if ( buildStencilSlots( s->h, s->lat_dims, s->lat_h, q.rc, s->h_perm, s->col, s->intact,
                                            s->slice_ptr.as<int64_t>(), s->total_rows, s->col_rows_cap, s->slot,
                                            s->scratch, s->acct(), s->st ) )
                        s->nsten = s->slot.slots;
                } );
            }
            // Mixed form on other lattices: one stencil id per bond replaces the x_j gather.
            if ( s->slot.slots == 0 && s->d.precision == PDK_MIXED && s->lattice_rows )
            {
                s->pushArgs();
                if ( !buildStencil( s->h, s->lat_h, q.rc, s->total_rows, s->sid, s->stab, s->stab_xi2d, s->stab_xiv, s->scratch,
                                    s->acct(), s->st, &s->nsten ) )
                    s->nsten = 0;
            }
        }
        s->refreshStats();
        if ( s->d.mechanics == PDK_MECH_EPP )
        {
            // BasePlasticity::updateBonds after the neighbor list: zero s_p
            // per bond (CabanaPD_Solver.hpp:183-185, CabanaPD_ForceModels.hpp:157-161)
            const size_t b = static_cast<size_t>( std::max<int64_t>( s->total_rows, 1 ) ) * kSlice * sizeof( double );
            s->splas.alloc( b, s->acct() );
            PDK_HIP( hipMemsetAsync( s->splas.as<double>(), 0, b, s->st ) );
        }
        s->built = true;
        s->inited = false;
        s->pushArgs();
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_bond_stats( pdk_spec s, int64_t* total_bonds, int32_t* max_neighbors )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        if ( s->stats_dirty )
            s->refreshStats();
        if ( total_bonds )
            *total_bonds = s->total_bonds;
        if ( max_neighbors )
            *max_neighbors = s->max_cnt;
    } );
}

int pdk_prenotch( pdk_spec s, int32_t nn, const double* p0, const double* v1, const double* v2 )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        require( s->d.fracture != 0, "Cannot create prenotch in system without fracture." );
        require( nn >= 0 && ( nn == 0 || ( p0 && v1 && v2 ) ), "invalid prenotch arguments" );
        if ( nn == 0 )
            return;
        std::vector<double> nt( 9 * nn );
        for ( int t = 0; t < nn; t++ )
        {
            for ( int dd = 0; dd < 3; dd++ )
            {
                nt[9 * t + dd] = p0[3 * t + dd];
                nt[9 * t + 3 + dd] = v1[3 * t + dd];
                nt[9 * t + 6 + dd] = v2[3 * t + dd];
            }
            // Prenotch.hpp:108 asserts non-parallel orientation vectors.
            const double cx = v1[3 * t + 1] * v2[3 * t + 2] - v1[3 * t + 2] * v2[3 * t + 1];
            const double cy = v1[3 * t + 0] * v2[3 * t + 2] - v1[3 * t + 2] * v2[3 * t + 0];
            const double cz = v1[3 * t + 0] * v2[3 * t + 1] - v1[3 * t + 1] * v2[3 * t + 0];
            require( std::sqrt( cx * cx + cy * cy + cz * cz ) > 1e-10, "prenotch vectors are parallel" );
        }
        s->notch.alloc( nt.size() * sizeof( double ), s->acct() );
        PDK_HIP( hipMemcpyAsync( s->notch.as<double>(), nt.data(), nt.size() * sizeof( double ), hipMemcpyHostToDevice,
                                 s->st ) );
        s->timed( kClsOther, [&] { launchPrenotch( s->h, nn, s->notch.as<double>(), s->st ); } );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_download_bonds( pdk_spec s, int64_t capacity, int64_t* offsets, int64_t* nbr, uint8_t* intact )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        if ( s->stats_dirty )
            s->refreshStats();
        require( ( !nbr && !intact ) || capacity >= s->total_bonds,
                 "pdk_download_bonds: capacity is smaller than the bond count" );
        const int64_t n = s->n, ns = s->nslices;
        std::vector<int64_t> sp( ns + 1 );
        std::vector<int32_t> c( n );
        std::vector<int32_t> col( static_cast<size_t>( s->total_rows ) * kSlice );
        std::vector<uint64_t> bits( s->total_rows );
        const bool slots = s->slot.slots > 0;
        std::vector<uint64_t> ex( slots ? s->total_rows : 0 );
        PDK_HIP( hipMemcpyAsync( sp.data(), s->slice_ptr.as<int64_t>(), ( ns + 1 ) * sizeof( int64_t ),
                                 hipMemcpyDeviceToHost, s->st ) );
        PDK_HIP( hipMemcpyAsync( c.data(), s->cnt.as<int32_t>(), n * sizeof( int32_t ), hipMemcpyDeviceToHost, s->st ) );
        if ( s->total_rows )
        {
            PDK_HIP( hipMemcpyAsync( col.data(), s->col.as<int32_t>(), col.size() * sizeof( int32_t ),
                                     hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipMemcpyAsync( bits.data(), s->intact.as<uint64_t>(), bits.size() * sizeof( uint64_t ),
                                     hipMemcpyDeviceToHost, s->st ) );
            if ( slots )
                PDK_HIP( hipMemcpyAsync( ex.data(), s->slot.exist.as<uint64_t>(), ex.size() * sizeof( uint64_t ),
                                         hipMemcpyDeviceToHost, s->st ) );
        }
        PDK_HIP( hipStreamSynchronize( s->st ) );
        int64_t sum = 0;
        for ( int32_t v : c )
            sum += v;
        require( ( !nbr && !intact ) || sum <= capacity, "pdk_download_bonds: capacity is smaller than the bond count" );
        int64_t off = 0;
        for ( int64_t cid = 0; cid < n; cid++ )
        {
            const int64_t p = s->h_inv[cid];
            const int64_t slice = p / kSlice, lane = p % kSlice;
            if ( offsets )
                offsets[cid] = off;
            // compact form: slots [0, cnt); stencil-slot form: every slot of
            // the row whose exist bit is set, in slot order
            const int64_t qend = slots ? sp[slice + 1] - sp[slice] : c[p];
            int64_t k = 0;
            for ( int64_t q = 0; q < qend; q++ )
            {
                const int64_t row = sp[slice] + q;
                if ( slots && !( ( ex[row] >> lane ) & 1ull ) )
                    continue;
                require( k < c[p], "pdk_download_bonds: inconsistent bond rows" );
                if ( nbr )
                    nbr[off + k] = s->h_perm[col[row * kSlice + lane]];
                if ( intact )
                    intact[off + k] = static_cast<uint8_t>( ( bits[row] >> lane ) & 1ull );
                k++;
            }
            require( k == c[p], "pdk_download_bonds: inconsistent bond rows" );
            off += c[p];
        }
        if ( offsets )
            offsets[n] = off;
    } );
}

int pdk_set_bc( pdk_spec s, int32_t kind, const int64_t* idx, int64_t nidx, double value )
{
    return guarded( [&] {
        require( s && s->uploaded, "spec not uploaded" );
        require( kind == PDK_BC_FORCE_VALUE || kind == PDK_BC_FORCE_UPDATE, "unknown BC kind" );
        require( nidx >= 0 && ( nidx == 0 || idx ), "invalid BC index list" );
        std::vector<uint8_t> flag( s->n, 0 );
        for ( int64_t b = 0; b < nidx; b++ )
        {
            require( idx[b] >= 0 && idx[b] < s->n, "BC particle id out of range" );
            flag[s->h_inv[idx[b]]] = 1;
        }
        PDK_HIP( hipMemcpyAsync( s->bcflag.as<uint8_t>(), flag.data(), s->n, hipMemcpyHostToDevice, s->st ) );
        s->h.bc_kind = kind;
        s->h.bc_value = value;
        s->pushArgs();
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_set_body_force( pdk_spec s, const double* b )
{
    return guarded( [&] {
        require( s && s->uploaded, "spec not uploaded" );
        if ( b )
        {
            if ( !s->body )
                s->body.alloc( 3 * s->n * sizeof( double ), s->acct() );
            s->uploadAos3( b, s->body.as<double>() );
        }
        else
            s->body.release();
        s->pushArgs();
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

namespace
{
int32_t prescribedComponents( pdk_spec s, int32_t field )
{
    if ( field == PDK_FIELD_U )
        return 3;
    require( field == PDK_FIELD_TEMPERATURE, "prescribed values are for PDK_FIELD_U or PDK_FIELD_TEMPERATURE" );
    require( s->d.thermal != PDK_TEMP_INDEPENDENT, "the spec has no temperature field (TemperatureIndependent)" );
    return 1;
}
} // namespace

int pdk_set_prescribed( pdk_spec s, int32_t field, const int64_t* idx, int64_t count, const double* a,
                        const double* b )
{
    return guarded( [&] {
        require( s && s->uploaded, "spec not uploaded" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( !isContact( s->d ), "prescribed values are for the PD models" );
        const int32_t nc = prescribedComponents( s, field );
        require( count >= 0 && ( count == 0 || ( idx && a ) ), "invalid prescribed arguments" );
        auto& pr = field == PDK_FIELD_U ? s->pre_u : s->pre_temp;
        pr.count = 0;
        if ( count == 0 )
            return;
        std::vector<int64_t> dev( count );
        std::vector<uint8_t> seen( s->n, 0 );
        for ( int64_t m = 0; m < count; m++ )
        {
            require( idx[m] >= 0 && idx[m] < s->n, "prescribed particle id out of range" );
            require( !seen[idx[m]], "prescribed particle ids are repeated" );
            seen[idx[m]] = 1;
            dev[m] = s->h_inv[idx[m]];
        }
        // Buffers are reused while they are large enough (a caller may set
        // new values every step); copies are ordered on the spec's stream
        // after the launches that read the previous values.
        const size_t vb = static_cast<size_t>( count ) * nc * sizeof( double );
        const size_t ib = static_cast<size_t>( count ) * sizeof( int64_t );
        auto ensure = [&]( DevBuf& buf, size_t bytes ) {
            if ( buf.bytes() < bytes )
            {
                PDK_HIP( hipStreamSynchronize( s->st ) ); // the old buffer may still be read
                buf.alloc( bytes, s->acct() );
            }
        };
        ensure( pr.idx, ib );
        ensure( pr.a, vb );
        PDK_HIP( hipMemcpyAsync( pr.idx.as<int64_t>(), dev.data(), ib, hipMemcpyHostToDevice, s->st ) );
        PDK_HIP( hipMemcpyAsync( pr.a.as<double>(), a, vb, hipMemcpyHostToDevice, s->st ) );
        pr.has_b = b != nullptr;
        if ( b )
        {
            ensure( pr.b, vb );
            PDK_HIP( hipMemcpyAsync( pr.b.as<double>(), b, vb, hipMemcpyHostToDevice, s->st ) );
        }
        pr.count = count;
        pr.ncomp = nc;
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_set_field( pdk_spec s, int32_t field, const double* values )
{
    return guarded( [&] {
        require( s && s->uploaded && values, "spec not uploaded" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( field == PDK_FIELD_TEMPERATURE, "pdk_set_field sets PDK_FIELD_TEMPERATURE" );
        require( s->d.thermal != PDK_TEMP_INDEPENDENT, "the spec has no temperature field (TemperatureIndependent)" );
        s->uploadScalar( values, s->temp.as<double>() );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_compute_weighted_volume( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        s->lpsState( true, false );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_compute_dilatation( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        s->lpsState( false, true );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_compute_force( pdk_spec s, int32_t reset )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        if ( isContact( s->d ) )
        {
            require( reset != 0, "contact force accumulation without reset is not supported" );
            contactForceStore( s );
        }
        else
            s->timed( kClsForce, [&] { launchForce( s->dargs(), s->h, reset ? kEpiStoreOnly : -1, s->cur, s->st ); } );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_compute_energy( pdk_spec s, double* total_strain_energy, double* total_damage )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        double Phi = 0.0, D = 0.0;
        if ( !isContact( s->d ) )
        {
            const int64_t nb = energyBlocks( s->h );
            if ( s->partial.bytes() < static_cast<size_t>( 2 * nb ) * sizeof( double ) )
                s->partial.alloc( 2 * nb * sizeof( double ), s->acct() );
            s->timed( kClsEnergy, [&] { launchEnergy( s->dargs(), s->h, s->cur, s->partial.as<double>(), nb, s->st ); } );
            std::vector<double> part( 2 * nb );
            PDK_HIP( hipMemcpyAsync( part.data(), s->partial.as<double>(), 2 * nb * sizeof( double ),
                                     hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
            for ( int64_t b = 0; b < nb; b++ )
            {
                Phi += part[2 * b];
                D += part[2 * b + 1];
            }
        }
        if ( total_strain_energy )
            *total_strain_energy = Phi;
        if ( total_damage )
            *total_damage = s->d.fracture ? D : 0.0;
    } );
}

int pdk_set_contact( pdk_spec s, const pdk_contact_desc* cd )
{
    return guarded( [&] {
        require( s && cd, "null argument" );
        require( !isContact( s->d ), "pdk_set_contact takes a PD specimen (DEM specimens are contact models)" );
        require( s->built, "build the bonds before pdk_set_contact" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( cd->model == PDK_CONTACT_NORMAL_REPULSION, "unknown contact model" );
        require( cd->radius > 0.0 && cd->radius_extend >= 0.0 && cd->horizon > 0.0, "invalid contact parameters" );
        s->pdc.on = true;
        s->pdc.cd = *cd;
        // NormalRepulsionModel ctor (force_models/CabanaPD_Contact.hpp:70-80)
        s->pdc.c = 18.0 * cd->K / ( kPi * cd->horizon * cd->horizon * cd->horizon * cd->horizon );
        if ( !s->pdc.slice_ptr )
        {
            s->pdc.slice_ptr.alloc( ( s->nslices + 1 ) * sizeof( int64_t ), s->acct() );
            s->pdc.cnt.alloc( std::max<int64_t>( s->n, 1 ) * sizeof( int32_t ), s->acct() );
            s->pdc.u_neigh.alloc( 3 * std::max<int64_t>( s->n, 1 ) * sizeof( double ), s->acct() );
            s->pdc.partial.alloc( ( maxDisplacementBlocks( s->n ) + 1 ) * sizeof( double ), s->acct() );
        }
        s->dropGraphs();
        // Neighbor ctor with a contact model: current positions, cutoff() + tol
        s->timed( kClsBuild, [&] { s->pdContactBuild( s->d.neighbor_tol ); } );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_contact_pairs( pdk_spec s, int64_t* pairs )
{
    return guarded( [&] {
        require( s && pairs, "null argument" );
        int64_t t = 0;
        if ( s->pdc.on )
        {
            std::vector<int32_t> c( s->n );
            PDK_HIP( hipMemcpyAsync( c.data(), s->pdc.cnt.as<int32_t>(), s->n * sizeof( int32_t ),
                                     hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
            for ( int32_t v : c )
                t += v;
        }
        *pairs = t;
    } );
}

namespace
{
void requireTableSpec( pdk_spec s, int32_t ntypes, const int32_t* type )
{
    require( s->uploaded, "call pdk_spec_upload first" );
    require( !s->in_batch, "spec is running in the batch engine" );
    require( !s->inited, "set the materials before pdk_init" );
    require( ntypes >= 1 && ntypes <= 4, "1 to 4 material types" );
    require( !isContact( s->d ), "the contact models have no materials" );
    require( ntypes == 1 || type != nullptr, "type[] is required for several materials" );
    require( s->d.precision == PDK_FP64 ||
                 ( s->d.model != PDK_MODEL_LINEAR_PMB && s->d.model != PDK_MODEL_LINEAR_LPS ),
             "the mixed form of several materials covers the PMB and LPS models" );
}

// Install the pair tables: c and bond_break_coeff per ordered type pair, the
// LPS coefficients per type (the i-term coefficients of pair (a, b) are those
// of type a, the j-term ones those of type b), and the particle types.
// `tables` false (one type, given as materials) keeps the spec's single model.
void installPairTables( pdk_spec s, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type, bool tables )
{
    for ( int32_t a = 0; a < 4; a++ )
        for ( int32_t b = 0; b < 4; b++ )
        {
            const bool in = a < ntypes && b < ntypes;
            s->mat_c[4 * a + b] = in ? pairs[a * ntypes + b].c : 0.0;
            s->mat_bcc[4 * a + b] = in ? pairs[a * ntypes + b].bond_break_coeff : 0.0;
        }
    for ( int32_t t = 0; t < 4; t++ )
    {
        s->mat_tc[t] = t < ntypes ? pairs[t * ntypes + t].theta_coeff[0] : 0.0;
        s->mat_sc[t] = t < ntypes ? pairs[t * ntypes + t].s_coeff[0] : 0.0;
    }
    if ( tables )
    {
        std::vector<uint8_t> dev( s->n );
        for ( int64_t cidx = 0; cidx < s->n; cidx++ )
        {
            const int32_t t = type ? type[cidx] : 0;
            require( t >= 0 && t < ntypes, "material type out of range" );
            dev[s->h_inv[cidx]] = static_cast<uint8_t>( t );
        }
        if ( s->mtype.bytes() < static_cast<size_t>( s->n ) )
            s->mtype.alloc( std::max<int64_t>( s->n, 1 ), s->acct() );
        PDK_HIP( hipMemcpyAsync( s->mtype.as<uint8_t>(), dev.data(), s->n, hipMemcpyHostToDevice, s->st ) );
    }
    s->nmat = ntypes;
    s->pair_tables = tables;
    s->pushArgs();
    PDK_HIP( hipStreamSynchronize( s->st ) );
}
} // namespace

int pdk_set_materials( pdk_spec s, int32_t nmat, const pdk_material* mats, const int32_t* type )
{
    return guarded( [&] {
        require( s && mats, "null argument" );
        requireTableSpec( s, nmat, type );
        // createMultiForceModel( ..., AverageTag, m_0, .. ) (CabanaPD_ForceModelsMulti.hpp:360-502):
        // pair (a, a) -> material a's model; a != b -> the model averaged from a's
        // and b's (force_models/CabanaPD_PMB.hpp:51-55, force_models/CabanaPD_LPS.hpp:55-69,
        // CabanaPD_ForceModels.hpp:114-121).
        double K[4], c[4], s0[4], bcc[4], tc[4], sc[4];
        for ( int32_t t = 0; t < nmat; t++ )
        {
            pdk_model_desc d = s->d;
            d.K = mats[t].K;
            d.G = mats[t].G;
            d.G0 = mats[t].G0;
            pdk_model_constants k;
            modelConstants( d, k ); // validates K, G, G0 as for one model
            K[t] = d.K;
            c[t] = k.c;
            s0[t] = k.s0;
            bcc[t] = k.bond_break_coeff;
            tc[t] = k.theta_coeff;
            sc[t] = k.s_coeff;
        }
        pdk_pair_model pairs[16];
        for ( int32_t a = 0; a < nmat; a++ )
            for ( int32_t b = 0; b < nmat; b++ )
            {
                pdk_pair_model& p = pairs[a * nmat + b];
                if ( a == b )
                {
                    p.c = c[a];
                    p.bond_break_coeff = bcc[a];
                }
                else
                {
                    p.c = ( c[a] + c[b] ) / 2.0;
                    const double s0ab = std::sqrt( ( s0[a] * s0[a] * K[a] + s0[b] * s0[b] * K[b] ) / ( K[a] + K[b] ) );
                    p.bond_break_coeff = ( 1.0 + s0ab ) * ( 1.0 + s0ab );
                }
                p.theta_coeff[0] = tc[a];
                p.theta_coeff[1] = tc[b];
                p.s_coeff[0] = sc[a];
                p.s_coeff[1] = sc[b];
            }
        installPairTables( s, nmat, pairs, type, nmat > 1 );
    } );
}

int pdk_set_pair_models( pdk_spec s, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type )
{
    return guarded( [&] {
        require( s && pairs, "null argument" );
        requireTableSpec( s, ntypes, type );
        const bool lps = s->d.model == PDK_MODEL_LPS || s->d.model == PDK_MODEL_LINEAR_LPS;
        for ( int32_t a = 0; a < ntypes; a++ )
            for ( int32_t b = 0; b < ntypes; b++ )
            {
                const pdk_pair_model& p = pairs[a * ntypes + b];
                require( std::isfinite( p.c ) && std::isfinite( p.theta_coeff[0] ) &&
                             std::isfinite( p.theta_coeff[1] ) && std::isfinite( p.s_coeff[0] ) &&
                             std::isfinite( p.s_coeff[1] ),
                         "pair model constants must be finite" );
                require( !s->d.fracture || p.bond_break_coeff > 0.0, "bond_break_coeff must be positive" );
                if ( lps )
                {
                    // per-particle folding of the LPS terms: [0] from type a, [1] from type b
                    const pdk_pair_model& ia = pairs[a * ntypes + a];
                    const pdk_pair_model& ib = pairs[b * ntypes + b];
                    require( p.theta_coeff[0] == ia.theta_coeff[0] && p.s_coeff[0] == ia.s_coeff[0] &&
                                 p.theta_coeff[1] == ib.theta_coeff[0] && p.s_coeff[1] == ib.s_coeff[0] &&
                                 ia.theta_coeff[0] == ia.theta_coeff[1] && ia.s_coeff[0] == ia.s_coeff[1],
                             "LPS pair models must take each particle's theta_coeff and s_coeff from its own type" );
                }
            }
        installPairTables( s, ntypes, pairs, type, true );
    } );
}

int pdk_compute_stress( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( !isContact( s->d ), "the contact models have no stress pass" );
        if ( !s->stress )
        {
            s->stress.alloc( static_cast<size_t>( 9 ) * std::max<int64_t>( s->n, 1 ) * sizeof( double ), s->acct() );
            s->pushArgs();
        }
        s->timed( kClsOther, [&] { launchStress( s->dargs(), s->h, s->cur, s->st ); } );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_compute_energy_noalloc( pdk_spec s, double* total_strain_energy, double* total_damage )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        // blocks of the energy pass: kSlicesPerBlock slices of kSlice particles
        const int64_t nb = ( s->nslices + kSlicesPerBlock - 1 ) / kSlicesPerBlock;
        if ( nb > kNoAllocMaxBlocks )
            throw Error( "pdk_compute_energy_noalloc: " + std::to_string( s->n ) +
                         " particles exceed the fixed scratch (at most " +
                         std::to_string( kNoAllocMaxBlocks * kBlock ) + ")" );
        require( s->built, "bonds not built" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( nb == energyBlocks( s->h ), "energy pass block count mismatch" );
        double tot[2] = { 0.0, 0.0 };
        if ( !isContact( s->d ) )
        {
            double* part = g_scratch_ptr[s->ctx->device];
            require( part != nullptr, "scratch not initialised for this device" );
            // every block writes its own pair of partials, then the in-order
            // sum writes the totals: nothing has to start at zero
            launchEnergy( s->dargs(), s->h, s->cur, part, nb, s->st );
            hipLaunchKernelGGL( kSumPartialsInOrder, dim3( 1 ), dim3( 64 ), 0, s->st, part, nb );
            PDK_HIP( hipGetLastError() );
            PDK_HIP( hipMemcpyAsync( tot, part + 2 * nb, 2 * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
        }
        if ( total_strain_energy )
            *total_strain_energy = tot[0];
        if ( total_damage )
            *total_damage = s->d.fracture ? tot[1] : 0.0;
    } );
}

int pdk_compute_energy_async( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        require( !s->in_batch, "spec is running in the batch engine" );
        if ( isContact( s->d ) )
            return;
        const int64_t nb = energyBlocks( s->h );
        if ( s->partial.bytes() < static_cast<size_t>( 2 * nb ) * sizeof( double ) )
            s->partial.alloc( 2 * nb * sizeof( double ), s->acct() );
        s->timed( kClsEnergy, [&] { launchEnergy( s->dargs(), s->h, s->cur, s->partial.as<double>(), nb, s->st ); } );
    } );
}

int pdk_init( pdk_spec s )
{
    return guarded( [&] {
        require( s && s->built, "bonds not built" );
        require( !s->in_batch, "spec is running in the batch engine" );
        // Solver::init(bc) (CabanaPD_Solver.hpp:229-279)
        s->applyPrescribed( 0.0 );
        if ( isContact( s->d ) )
        {
            contactForceStore( s );
        }
        else
        {
            if ( s->lps() && !s->d.fracture )
                s->lpsState( true, false );
            s->updateForceStore();
            const int64_t nb = energyBlocks( s->h );
            if ( s->partial.bytes() < static_cast<size_t>( 2 * nb ) * sizeof( double ) )
                s->partial.alloc( 2 * nb * sizeof( double ), s->acct() );
            s->timed( kClsEnergy, [&] { launchEnergy( s->dargs(), s->h, s->cur, s->partial.as<double>(), nb, s->st ); } );
        }
        s->timed( kClsOther, [&] { launchApplyForceTerms( s->dargs(), s->h, s->st ); } );
        PDK_HIP( hipStreamSynchronize( s->st ) );
        s->inited = true;
    } );
}

int pdk_step( pdk_spec s, int32_t nsteps )
{
    return guarded( [&] {
        require( s && s->inited, "call pdk_init before pdk_step" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( nsteps >= 0, "nsteps must be >= 0" );
        if ( nsteps == 0 )
            return;
        if ( isContact( s->d ) )
        {
            contactStep( s, nsteps );
        }
        else if ( s->pdc.on )
        {
            // Solver::runStep with a contact model (CabanaPD_Solver.hpp:332-404):
            // initial half step and the largest displacement since the last
            // contact list build, heat transfer / prescribed values, the PD
            // force, the contact force (rebuilding the list first when that
            // displacement exceeds radius_extend, Neighbor::update), the force
            // terms and the final half step.
            double* mx_dev = s->pdc.partial.as<double>() + maxDisplacementBlocks( s->n );
            for ( int32_t k = 1; k <= nsteps; k++ )
            {
                s->timed( kClsOther, [&] { launchInitialHalfStep( s->dargs(), s->h, s->cur, s->st ); } );
                s->cur ^= 1;
                launchMaxDisplacement( s->Ucur(), s->pdc.u_neigh.as<double>(), s->n, s->pdc.partial.as<double>(),
                                       mx_dev, s->st );
                double mx = 0.0;
                PDK_HIP( hipMemcpyAsync( &mx, mx_dev, sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
                PDK_HIP( hipStreamSynchronize( s->st ) );
                if ( s->hasPreForce() )
                    s->preForce( s->steps_done + k );
                if ( mx > s->pdc.cd.radius_extend )
                    s->timed( kClsBuild, [&] { s->pdContactBuild( 0.0 ); } );
                if ( s->lps() )
                    s->lpsState( s->d.fracture != 0, true );
                s->timed( kClsForce, [&] { launchForce( s->dargs(), s->h, kEpiStoreOnly, s->cur, s->st ); } );
                s->timed( kClsOther, [&] {
                    launchPdContact( s->dargs(), s->h, s->pdContactArgs(), kEpiFinal, s->cur, s->st );
                } );
            }
        }
        else
        {
            // Solver::runStep (CabanaPD_Solver.hpp:332-375): the initial half
            // step of step 1 runs alone; every later initial half step runs in
            // the previous step's force epilogue.
            s->timed( kClsOther, [&] { launchInitialHalfStep( s->dargs(), s->h, s->cur, s->st ); } );
            s->cur ^= 1;
            // heat transfer / prescribed values of step g, after its initial half step
            const bool pre = s->hasPreForce();
            if ( pre )
                s->preForce( s->steps_done + 1 );
            int32_t k = 1;
            // middle steps in pairs from a captured graph (not while kernels
            // are being timed: events are recorded around each launch)
            if ( s->use_graph && !s->timing && !pre )
                for ( ; k + 1 < nsteps; k += 2 )
                    PDK_HIP( hipGraphLaunch( s->stepPairGraph( s->cur ), s->st ) );
            for ( ; k <= nsteps; k++ )
            {
                if ( s->lps() )
                    s->lpsState( s->d.fracture != 0, true );
                const int32_t epi = k < nsteps ? kEpiFinalInitial : kEpiFinal;
                s->timed( kClsForce, [&] { launchForce( s->dargs(), s->h, epi, s->cur, s->st ); } );
                if ( epi == kEpiFinalInitial )
                {
                    s->cur ^= 1;
                    if ( pre )
                        s->preForce( s->steps_done + k + 1 );
                }
            }
        }
        s->steps_done += nsteps;
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_steps_done( pdk_spec s, int64_t* steps )
{
    return guarded( [&] {
        require( s && steps, "null argument" );
        *steps = s->steps_done;
    } );
}

int pdk_download( pdk_spec s, int32_t field, double* out )
{
    return guarded( [&] {
        require( s && s->uploaded && out, "spec not uploaded" );
        const int64_t n = s->n;
        int32_t nc = 1;
        const double* src = s->fieldDev( field, &nc );
        if ( !src )
        {
            std::fill( out, out + nc * n, 0.0 );
            return;
        }
        const size_t bytes = static_cast<size_t>( nc ) * n * sizeof( double );
        if ( s->staging.bytes() < bytes )
            s->staging.alloc( bytes, s->acct() );
        if ( nc == 3 )
            launchScatterAos3( src, s->perm.as<int64_t>(), s->staging.as<double>(), n, s->st );
        else if ( nc == 1 )
            launchScatterScalar( src, s->perm.as<int64_t>(), s->staging.as<double>(), n, s->st );
        else
            launchScatterRows( src, s->perm.as<int64_t>(), s->staging.as<double>(), n, nc, s->st );
        PDK_HIP( hipMemcpyAsync( out, s->staging.as<double>(), bytes, hipMemcpyDeviceToHost, s->st ) );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_sync( pdk_spec s )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        PDK_HIP( hipStreamSynchronize( s->st ) );
    } );
}

int pdk_series_create( pdk_spec s, const pdk_region_desc* r, const pdk_quantity* q, int32_t reducer,
                       int64_t capacity, pdk_series* out )
{
    return guarded( [&] {
        require( s && r && q && out, "null argument" );
        *out = nullptr;
        require( s->uploaded, "spec not uploaded" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( reducer >= PDK_REDUCE_SUM && reducer <= PDK_REDUCE_MIN, "unknown reducer" );
        require( capacity >= 0, "capacity must be >= 0" );
        int32_t nc = 1;
        (void)s->fieldDev( q->field, &nc );
        require( q->component >= 0 && q->component < nc, "component out of range for the field" );
        if ( q->gate_field >= 0 )
        {
            int32_t gnc = 1;
            (void)s->fieldDev( q->gate_field, &gnc );
            require( gnc == 1, "the gate field must be a scalar field" );
        }
        require( q->weight_volume == 0 || q->weight_volume == 1, "weight_volume must be 0 or 1" );
        const int64_t n = s->n;
        auto ser = std::make_unique<pdk_series_s>();
        switch ( r->kind )
        {
        case PDK_REGION_ALL:
            ser->ids.resize( n );
            for ( int64_t c = 0; c < n; c++ )
                ser->ids[c] = c;
            break;
        case PDK_REGION_INDICES:
        {
            require( r->count >= 0 && ( r->ids || r->count == 0 ), "INDICES region needs ids[count]" );
            ser->ids.assign( r->ids, r->ids + r->count );
            std::sort( ser->ids.begin(), ser->ids.end() );
            for ( size_t k = 0; k < ser->ids.size(); k++ )
            {
                require( ser->ids[k] >= 0 && ser->ids[k] < n, "region id out of range" );
                require( k == 0 || ser->ids[k] != ser->ids[k - 1], "repeated region id" );
            }
            break;
        }
        case PDK_REGION_BOX:
        case PDK_REGION_CYLINDER:
        case PDK_REGION_LINE:
        {
            require( r->kind != PDK_REGION_LINE || ( r->dim >= 0 && r->dim < 3 ), "line dim must be 0, 1 or 2" );
            RegionParams P{};
            P.kind = r->kind;
            P.dim = r->dim;
            for ( int d = 0; d < 3; d++ )
            {
                P.low[d] = r->low[d];
                P.high[d] = r->high[d];
            }
            P.r_in = r->r_in;
            P.r_out = r->r_out;
            P.low_z = r->low_z;
            P.high_z = r->high_z;
            P.x_center = r->x_center;
            P.y_center = r->y_center;
            P.half[0] = r->half[0];
            P.half[1] = r->half[1];
            DevBuf mask;
            mask.alloc( std::max<int64_t>( n, 1 ), s->acct() );
            launchRegionMask( P, s->x.as<double>(), s->inv.as<int64_t>(), n, mask.as<uint8_t>(), s->st );
            std::vector<uint8_t> hm( n );
            if ( n )
                PDK_HIP( hipMemcpyAsync( hm.data(), mask.as<uint8_t>(), n, hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
            for ( int64_t c = 0; c < n; c++ )
                if ( hm[c] )
                    ser->ids.push_back( c );
            break;
        }
        default:
            throw Error( "unknown region kind" );
        }
        ser->spec = s;
        ser->q = *q;
        ser->red = reducer;
        ser->capacity = capacity;
        ser->partial.alloc( kSeriesMaxPartials * sizeof( double ), s->acct() );
        if ( capacity )
            ser->values.alloc( capacity * sizeof( double ), s->acct() );
        ser->times.reserve( static_cast<size_t>( std::min<int64_t>( capacity, 1 << 20 ) ) );
        ser->mapIds();
        s->series_list.push_back( ser.get() );
        *out = ser.release();
    } );
}

int pdk_series_destroy( pdk_series ser )
{
    return guarded( [&] {
        if ( !ser )
            return;
        if ( pdk_spec s = ser->spec )
        {
            (void)hipStreamSynchronize( s->st ); // pending updates write into ser's buffers
            auto& l = s->series_list;
            l.erase( std::remove( l.begin(), l.end(), ser ), l.end() );
        }
        delete ser;
    } );
}

int pdk_series_members( pdk_series ser, int64_t* count )
{
    return guarded( [&] {
        require( ser && count, "null argument" );
        *count = static_cast<int64_t>( ser->ids.size() );
    } );
}

int pdk_series_update( pdk_series ser, double time )
{
    return guarded( [&] {
        require( ser != nullptr, "null series" );
        pdk_spec s = ser->spec;
        require( s != nullptr, "the series' spec was destroyed" );
        require( !s->in_batch, "spec is running in the batch engine" );
        require( ser->used < ser->capacity, "series is full" );
        ser->mapIds();
        const QuantityArgs Q = ser->args();
        s->timed( kClsOther, [&] {
            launchSeriesReduce( ser->dev_ids.as<int64_t>(), static_cast<int64_t>( ser->ids.size() ), Q, ser->red,
                                ser->partial.as<double>(), ser->values.as<double>() + ser->used, s->st );
        } );
        ser->times.push_back( time );
        ser->used++;
    } );
}

int pdk_series_values( pdk_series ser, int64_t capacity, double* times, double* values, int64_t* count )
{
    return guarded( [&] {
        require( ser != nullptr, "null series" );
        const int64_t k = std::min( std::max<int64_t>( capacity, 0 ), ser->used );
        if ( k > 0 && values )
        {
            pdk_spec s = ser->spec;
            require( s != nullptr, "the series' spec was destroyed" );
            PDK_HIP( hipMemcpyAsync( values, ser->values.as<double>(), k * sizeof( double ), hipMemcpyDeviceToHost,
                                     s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
        }
        if ( k > 0 && times )
            std::copy( ser->times.begin(), ser->times.begin() + k, times );
        if ( count )
            *count = ser->used;
    } );
}

int pdk_series_gather( pdk_series ser, int64_t capacity, int64_t* ids, double* values, int64_t* count )
{
    return guarded( [&] {
        require( ser != nullptr, "null series" );
        pdk_spec s = ser->spec;
        require( s != nullptr, "the series' spec was destroyed" );
        const int64_t m = static_cast<int64_t>( ser->ids.size() );
        const int64_t k = std::min( std::max<int64_t>( capacity, 0 ), m );
        if ( k > 0 && values )
        {
            ser->mapIds();
            const QuantityArgs Q = ser->args();
            const size_t bytes = static_cast<size_t>( k ) * sizeof( double );
            if ( s->staging.bytes() < bytes )
                s->staging.alloc( bytes, s->acct() );
            launchGatherQuantity( ser->dev_ids.as<int64_t>(), k, Q, s->staging.as<double>(), s->st );
            PDK_HIP( hipMemcpyAsync( values, s->staging.as<double>(), bytes, hipMemcpyDeviceToHost, s->st ) );
            PDK_HIP( hipStreamSynchronize( s->st ) );
        }
        if ( k > 0 && ids )
            std::copy( ser->ids.begin(), ser->ids.begin() + k, ids );
        if ( count )
            *count = m;
    } );
}

int pdk_spec_layout( pdk_spec s, int32_t* lattice_rows, int32_t* stencil_entries, int32_t* stencil_slots )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        if ( lattice_rows )
            *lattice_rows = s->lattice_rows ? 1 : 0;
        if ( stencil_entries )
            *stencil_entries = s->nsten;
        if ( stencil_slots )
            *stencil_slots = s->slot.slots > 0 ? 1 : 0;
    } );
}

int pdk_set_graph( pdk_spec s, int32_t enable )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        s->use_graph = enable != 0;
        s->dropGraphs();
    } );
}

int pdk_set_tuning( pdk_spec s, int32_t unroll )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        require( unroll == 0 || unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8,
                 "unroll must be 0 (row windows), 1, 2, 4 or 8" );
        require( unroll != 0 || s->rowsCover(),
                 "this combination of materials, EPP and temperature dependence uses the gather variant" );
        require( unroll != 0 || !s->built || s->slot.nrows > 0,
                 "row windows need the stencil-slot layout of a full lattice" );
        s->unroll = unroll;
        if ( s->uploaded )
        {
            s->pushArgs();
            PDK_HIP( hipStreamSynchronize( s->st ) );
        }
    } );
}

int pdk_autotune( pdk_spec s, int32_t launches, int32_t* unroll_out )
{
    return guarded( [&] {
        require( s && s->inited, "call pdk_init before pdk_autotune" );
        require( !s->in_batch, "spec is running in the batch engine" );
        if ( isContact( s->d ) )
        {
            if ( unroll_out )
                *unroll_out = s->h.unroll;
            return;
        }
        launches = std::max( launches, 1 );
        hipEvent_t a, b;
        PDK_HIP( hipEventCreate( &a ) );
        PDK_HIP( hipEventCreate( &b ) );
        float best = 1e30f;
        int32_t best_u = s->unroll;
        std::vector<int32_t> cand;
        if ( s->slot.nrows > 0 && s->rowsCover() )
            cand.push_back( 0 ); // row windows staged in LDS
        for ( int32_t u : { 1, 2, 4, 8 } )
            cand.push_back( u );
        for ( int32_t u : cand )
        {
            s->unroll = u;
            s->pushArgs();
            // warm-up launch, then the timed ones; each repeats the init force
            // pass on the unchanged state
            launchForce( s->dargs(), s->h, kEpiStoreOnly, s->cur, s->st );
            PDK_HIP( hipEventRecord( a, s->st ) );
            for ( int32_t k = 0; k < launches; k++ )
                launchForce( s->dargs(), s->h, kEpiStoreOnly, s->cur, s->st );
            PDK_HIP( hipEventRecord( b, s->st ) );
            PDK_HIP( hipEventSynchronize( b ) );
            float ms = 0.f;
            PDK_HIP( hipEventElapsedTime( &ms, a, b ) );
            if ( ms < best )
            {
                best = ms;
                best_u = u;
            }
        }
        (void)hipEventDestroy( a );
        (void)hipEventDestroy( b );
        s->unroll = best_u;
        s->pushArgs();
        // restore f as pdk_init left it (force terms on the raw force)
        launchForce( s->dargs(), s->h, kEpiStoreOnly, s->cur, s->st );
        launchApplyForceTerms( s->dargs(), s->h, s->st );
        PDK_HIP( hipStreamSynchronize( s->st ) );
        if ( unroll_out )
            *unroll_out = best_u;
    } );
}

int pdk_kernel_timing( pdk_spec s, int32_t enable )
{
    return guarded( [&] {
        require( s != nullptr, "null spec" );
        s->resolveEvents();
        s->timing = enable != 0;
        for ( int c = 0; c < kNumCls; c++ )
        {
            s->tsum[c] = 0.0;
            s->tcnt[c] = 0;
        }
    } );
}

int pdk_kernel_time( pdk_spec s, int32_t cls, double* total_ms, int64_t* launches )
{
    return guarded( [&] {
        require( s && cls >= 0 && cls < kNumCls, "invalid kernel class" );
        s->resolveEvents();
        if ( total_ms )
            *total_ms = s->tsum[cls];
        if ( launches )
            *launches = s->tcnt[cls];
    } );
}

int pdk_kernel_name( pdk_spec s, int32_t cls, char* buf, int32_t buflen )
{
    return guarded( [&] {
        require( s && buf && buflen > 0, "invalid argument" );
        const char* name = "";
        switch ( cls )
        {
        case kClsForce: name = isContact( s->d ) ? contactKernelName( s->h ) : forceKernelName( s->h ); break;
        case kClsLps: name = lpsKernelName( s->h ); break;
        case kClsEnergy: name = energyKernelName( s->h ); break;
        case kClsBuild: name = "kFillNeighbors"; break;
        default: name = "kInitialHalfStep"; break;
        }
        std::snprintf( buf, buflen, "%s", name );
    } );
}

} // extern "C"

// ---- accessors for the contact and serving translation units -----------------
namespace pdk
{
SpecView view( pdk_spec s )
{
    SpecView v;
    v.h = &s->h;
    v.dargs = s->dargs();
    v.st = s->st;
    v.n = s->n;
    v.nslices = s->nslices;
    v.cur = &s->cur;
    v.steps_done = &s->steps_done;
    v.in_batch = &s->in_batch;
    v.inited = s->inited;
    v.model = s->d.model;
    v.fracture = s->d.fracture;
    v.precision = s->d.precision;
    v.bytes = 0;
    pdk_spec_bytes( s, &v.bytes );
    return v;
}

void specPushArgs( pdk_spec s )
{
    s->pushArgs();
}

void contactBuildImpl( pdk_spec s, bool initial, double* max_disp, int64_t* since_build )
{
    NeighborQuery q{};
    q.x = s->x.as<double>();
    q.u = s->Ucur();
    q.radius = s->d.polydisperse ? s->radius.as<double>() : nullptr;
    const double tol = initial ? s->d.neighbor_tol : 0.0;
    q.rc = 2.0 * s->d.radius + s->d.radius_extend + tol;
    q.extend = s->d.radius_extend + tol;
    q.n = s->n;
    q.rmax = s->radius_max;
    // after the first build: one capped pass sized from the last list's
    // largest count (the counting path if a particle outgrows it)
    const int32_t cap = initial || s->max_cnt <= 0 ? 0 : ( ( s->max_cnt + 4 + 7 ) / 8 ) * 8;
    s->total_rows = buildSlicedCsr( q, s->perm.as<int64_t>(), s->inv.as<int64_t>(), s->slice_ptr.as<int64_t>(),
                                    s->cnt.as<int32_t>(), s->col, s->intact, s->col_rows_cap, s->scratch, s->acct(),
                                    s->st, &s->max_cnt, cap );
    PDK_HIP( hipMemcpyAsync( s->u_neigh.as<double>(), s->Ucur(), 3 * s->n * sizeof( double ), hipMemcpyDeviceToDevice,
                             s->st ) );
    s->stats_dirty = true;
    s->pushArgs();
    *max_disp = 0.0;
    *since_build = 0;
    s->max_disp = 0.0;
    s->steps_since_build = 0;
}

ContactState contactState( pdk_spec s )
{
    ContactState c;
    c.max_disp = &s->max_disp;
    c.since_build = &s->steps_since_build;
    c.extend = s->d.radius_extend;
    c.rebuild_every = s->d.rebuild_every;
    c.partial = &s->partial;
    c.acct = s->acct();
    if ( !s->disp_pinned )
    {
        PDK_HIP( hipHostMalloc( reinterpret_cast<void**>( &s->disp_pinned ), sizeof( double ), hipHostMallocDefault ) );
        PDK_HIP( hipEventCreateWithFlags( &s->disp_ready, hipEventDisableTiming ) );
    }
    c.disp_pinned = s->disp_pinned;
    c.disp_ready = s->disp_ready;
    return c;
}

void specTimed( pdk_spec s, int cls, const std::function<void()>& fn )
{
    s->timed( cls, fn );
}

MemAccount& ctxAccount( pdk_ctx_s* c )
{
    return c->acct;
}
} // namespace pdk

extern "C" {

int pdk_submit( pdk_ctx ctx, pdk_spec s, int32_t nsteps, uint64_t* ticket )
{
    return guarded( [&] {
        require( s && !s->pdc.on, "specs with a contact model are stepped with pdk_step" );
        require( ctx && s && ticket, "null argument" );
        require( s->ctx == ctx, "spec belongs to another context" );
        require( !s->gatherOnly(),
                 "multi-material, EPP and temperature-dependent specimens are stepped with pdk_step, not the batch engine" );
        require( !s->hasPreForce(),
                 "specimens with prescribed values or heat transfer are stepped with pdk_step, not the batch engine" );
        *ticket = ctx->batch->submit( s, nsteps );
    } );
}

int pdk_poll( pdk_ctx ctx, uint64_t ticket, int32_t* done )
{
    return guarded( [&] {
        require( ctx && done, "null argument" );
        *done = ctx->batch->done( ticket ) ? 1 : 0;
    } );
}

int pdk_batch_progress( pdk_ctx ctx, int32_t max_steps, int32_t* launched )
{
    return guarded( [&] {
        require( ctx != nullptr, "null ctx" );
        const int32_t l = ctx->batch->progress( max_steps );
        if ( launched )
            *launched = l;
    } );
}

int pdk_batch_set_graph( pdk_ctx ctx, int32_t steps )
{
    return guarded( [&] {
        require( ctx != nullptr, "null ctx" );
        ctx->batch->setGraphSteps( steps );
    } );
}

} // extern "C"
