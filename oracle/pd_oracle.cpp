// pd_oracle.cpp -- CPU oracle for the CabanaPD force/update path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker of tests/ and the CPU
// baseline leg of bench.py.  The product (libpdk.so) never links, loads or
// calls this file.
//
// This is a restatement of the reference's CPU path, loop for loop: one
// thread per particle over [0, n), a serial loop over that particle's
// neighbors, no atomics -- the structure of Kokkos RangePolicy +
// Cabana::neighbor_parallel_for with SerialOpTag (CabanaPD_Neighbor.hpp:31,
// 134-180).  Every floating-point expression keeps the reference's operand
// order; the file is compiled without FMA contraction (-ffp-contract=off,
// no -march) so each line rounds exactly as written.
//
// Neighbor rows are kept in ascending caller id.  The reference's slot order
// comes from Cabana's VerletList fill (atomic counters per cell) and is not
// reproducible; only the set is defined, so the oracle fixes the order.
//
// Pins (tests/test_oracle_*.py): lattice known answers of
// unit_test/tstForce.hpp:93-524,729-1015; neighbor counts (tstForce.hpp:983-1015,
// tstComm.hpp:153-166) and O(N^2) brute force; two-particle closed forms for
// the break test; geometric special cases for the prenotch; closed forms of
// velocity-Verlet; the Kalthoff-Winkler regression value
// (regression_test/tstKalthoffWinkler.hpp:53); Hertz restitution and JKR
// pull-off (regression_test/tstHertz.hpp:135, tstHertzJKR.hpp:145); EPP
// load-cycle closed forms, uniform-stretch scalings and the elastic limit;
// thermal free expansion, held-body equivalence and break threshold;
// prescribed values (tests/test_oracle_plastic_thermal.py).
// Behaviour no reference-held number covers is pinned by what it must be
// (tests/test_oracle_quirks.py): the LPS no-fail asymmetry, the NaN LPS energy
// of a particle with all bonds broken, and the polydisperse contact extension
// against the dissimilar-sphere Hertz and JKR closed forms.

#include "pd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
// Kokkos::numbers::pi_v<double> (CabanaPD_Constants.hpp:20)
constexpr double kPi = 3.141592653589793238462643383279502884;

int g_threads = 0;

struct Sys
{
    pdk_model_desc d;
    pdk_model_constants k;
    int64_t n = 0;
    std::vector<double> x, u, v, f, vol, rho, W, phi, m, theta, radius, u_neigh, stress;
    std::vector<int32_t> nofail;
    // VerletLayout2D-style table: nbr[i * maxn + slot], cnt[i]
    int32_t maxn = 0;
    std::vector<int32_t> cnt;
    std::vector<int32_t> nbr;
    std::vector<uint8_t> mu;
    bool built = false;
    // boundary condition / body term
    int32_t bc_kind = -1;
    std::vector<int64_t> bc_idx;
    double bc_value = 0.0;
    std::vector<double> body;
    // contact bookkeeping
    double max_disp = 0.0;
    int64_t steps_since_build = 0;
    double total_W = 0.0, total_D = 0.0;
    // multi-material (createMultiForceModel): material type per particle and
    // the constants of the model of each ordered type pair (type(i), type(j));
    // LPS pair models carry theta_coeff / s_coeff [0] for particle i's terms
    // and [1] for particle j's (force_models/CabanaPD_LPS.hpp:41-44)
    int32_t nmat = 1;
    bool pair_tables = false;
    std::vector<int32_t> mtype;
    double pair_c[16] = {}, pair_bcc[16] = {}, pair_tc[16][2] = {}, pair_sc[16][2] = {};
    // EPP: plastic stretch per bond sp[i * maxn + slot] (BasePlasticity::_s_p,
    // CabanaPD_ForceModels.hpp:149-163, zeroed when the bonds are built)
    std::vector<double> sp;
    // TemperatureDependent: temperature per particle (zero at creation,
    // CabanaPD_Particles.hpp:860-864)
    std::vector<double> temp;
    // values set before each force pass (boundary conditions and body terms
    // with force_update == false): a + b t per member, field U or TEMPERATURE
    struct Prescribed
    {
        std::vector<int64_t> idx;
        std::vector<double> a, b;
        int32_t ncomp = 0;
    };
    Prescribed pre_u, pre_temp;
    int64_t steps_done = 0;
    // PD specimen with a contact model (Solver with four arguments,
    // CabanaPD_Solver.hpp:134-151): its own neighbor list on current positions
    bool pdc = false;
    pdk_contact_desc cd{};
    double crep = 0.0; // NormalRepulsionModel c (force_models/CabanaPD_Contact.hpp:77-80)
    int32_t cmaxn = 0;
    std::vector<int32_t> ccnt, cnbr;
};

bool isContact( const pdk_model_desc& d )
{
    return d.model == PDK_MODEL_HERTZ || d.model == PDK_MODEL_HERTZ_JKR;
}
// LinearLPS is an LPS model (its force model derives from LPS's,
// force_models/CabanaPD_LPS.hpp:255-273); LinearPMB from PMB's
// (force_models/CabanaPD_PMB.hpp:575-597).
bool isLps( const pdk_model_desc& d )
{
    return d.model == PDK_MODEL_LPS || d.model == PDK_MODEL_LINEAR_LPS;
}
bool isLinear( const pdk_model_desc& d )
{
    return d.model == PDK_MODEL_LINEAR_PMB || d.model == PDK_MODEL_LINEAR_LPS;
}
bool isEpp( const pdk_model_desc& d )
{
    return d.mechanics == PDK_MECH_EPP;
}
bool isThermal( const pdk_model_desc& d )
{
    return d.thermal == PDK_TEMP_DEPENDENT || d.thermal == PDK_TEMP_DYNAMIC;
}
bool isDynamic( const pdk_model_desc& d )
{
    return d.thermal == PDK_TEMP_DYNAMIC;
}

// ---------------------------------------------------------------------------
// Derived model constants.
// ---------------------------------------------------------------------------
int modelConstants( const pdk_model_desc& d, pdk_model_constants& k )
{
    std::memset( &k, 0, sizeof( k ) );
    const double delta = d.horizon;
    const double K = d.K;
    if ( d.model == PDK_MODEL_PMB || d.model == PDK_MODEL_LPS || isLinear( d ) )
    {
        if ( isLinear( d ) && d.fracture )
            return 1; // the linearized models exist without fracture only
        // force_models/CabanaPD_PMB.hpp:44-46
        k.c = 18.0 * K / ( kPi * delta * delta * delta * delta );
        if ( d.fracture )
        {
            // CabanaPD_ForceModels.hpp:97-101
            k.s0 = std::sqrt( 5.0 * d.G0 / 9.0 / K / delta );
            if ( d.influence == 0 )
                k.s0 = std::sqrt( 8.0 * d.G0 / 15.0 / K / delta );
            if ( isLps( d ) )
            {
                // force_models/CabanaPD_LPS.hpp:214-222
                if ( d.influence == 1 )
                    k.s0 = std::sqrt( 5.0 * d.G0 / 9.0 / K / delta );
                else
                    k.s0 = std::sqrt( 8.0 * d.G0 / 15.0 / K / delta );
            }
            k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
        }
        // force_models/CabanaPD_LPS.hpp:75-76
        k.theta_coeff = 3.0 * K - 5.0 * d.G;
        k.s_coeff = 15.0 * d.G;
        if ( isLps( d ) && ( d.influence < 0 || d.influence > 1 ) )
            return 1;
        if ( d.mechanics != PDK_MECH_ELASTIC && d.mechanics != PDK_MECH_EPP )
            return 1;
        if ( d.thermal != PDK_TEMP_INDEPENDENT && d.thermal != PDK_TEMP_DEPENDENT && d.thermal != PDK_TEMP_DYNAMIC )
            return 1;
        if ( isDynamic( d ) )
        {
            if ( !( d.kappa > 0 ) || !( d.cp > 0 ) || d.thermal_subcycle < 1 )
                return 1;
            // BaseDynamicTemperatureModel (CabanaPD_ForceModels.hpp:277-288):
            // thermal horizon = force horizon
            const double d3 = delta * delta * delta;
            k.thermal_coeff = 9.0 / 2.0 * d.kappa / kPi / d3;
        }
        if ( isEpp( d ) )
        {
            // ForceModel<PMB, ElasticPerfectlyPlastic, Fracture> only
            // (force_models/CabanaPD_PMB.hpp:232-284)
            if ( d.model != PDK_MODEL_PMB || !d.fracture || !( d.sigma_y > 0 ) )
                return 1;
            // BaseForceModelPMB<ElasticPerfectlyPlastic> (force_models/CabanaPD_PMB.hpp:90-96)
            k.s_Y = d.sigma_y / 3.0 / K;
            if ( !isThermal( d ) )
            {
                // force_models/CabanaPD_PMB.hpp:255-264
                k.s0 = ( 5.0 * d.G0 / d.sigma_y / delta + d.sigma_y / K ) / 6.0;
                k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
            }
        }
        if ( isThermal( d ) )
        {
            // PMB and LinearPMB only (force_models/CabanaPD_PMB.hpp:306-568, 575-597)
            if ( d.model != PDK_MODEL_PMB && d.model != PDK_MODEL_LINEAR_PMB )
                return 1;
            if ( d.fracture )
            {
                // ThermalFractureModel: BaseFractureModel( delta, K, G0 ) with
                // influence_type 1 (CabanaPD_ForceModels.hpp:228-236, 93-102)
                k.s0 = std::sqrt( 5.0 * d.G0 / 9.0 / K / delta );
                k.bond_break_coeff = ( 1.0 + k.s0 ) * ( 1.0 + k.s0 );
            }
        }
    }
    else if ( d.mechanics != PDK_MECH_ELASTIC || d.thermal != PDK_TEMP_INDEPENDENT )
        return 1;
    else if ( d.model == PDK_MODEL_HERTZ )
    {
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
        return 1;
    return 0;
}

// ---------------------------------------------------------------------------
// getDistance (CabanaPD_Force.hpp:74-93).
// ---------------------------------------------------------------------------
inline void getDistance( const Sys& s, int64_t i, int64_t j, double& xi, double& r, double& st,
                         double& rx, double& ry, double& rz )
{
    const double* x = s.x.data();
    const double* u = s.u.data();
    const double xi_x = x[3 * j + 0] - x[3 * i + 0];
    const double eta_u = u[3 * j + 0] - u[3 * i + 0];
    const double xi_y = x[3 * j + 1] - x[3 * i + 1];
    const double eta_v = u[3 * j + 1] - u[3 * i + 1];
    const double xi_z = x[3 * j + 2] - x[3 * i + 2];
    const double eta_w = u[3 * j + 2] - u[3 * i + 2];
    rx = xi_x + eta_u;
    ry = xi_y + eta_v;
    rz = xi_z + eta_w;
    r = std::sqrt( rx * rx + ry * ry + rz * rz );
    xi = std::sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
    st = ( r - xi ) / xi;
}

// getLinearizedDistance (CabanaPD_Force.hpp:114-138): stretch linearized in
// eta, force along the reference bond; returned as r = xi and (rx, ry, rz) =
// xi so that the force passes' coeff * rx / r is the reference's
// coeff * xi_x / xi.
inline void getLinearizedDistance( const Sys& s, int64_t i, int64_t j, double& xi, double& r, double& st,
                                   double& rx, double& ry, double& rz )
{
    const double* x = s.x.data();
    const double* u = s.u.data();
    const double xi_x = x[3 * j + 0] - x[3 * i + 0];
    const double eta_u = u[3 * j + 0] - u[3 * i + 0];
    const double xi_y = x[3 * j + 1] - x[3 * i + 1];
    const double eta_v = u[3 * j + 1] - u[3 * i + 1];
    const double xi_z = x[3 * j + 2] - x[3 * i + 2];
    const double eta_w = u[3 * j + 2] - u[3 * i + 2];
    xi = std::sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
    st = ( xi_x * eta_u + xi_y * eta_v + xi_z * eta_w ) / ( xi * xi );
    rx = xi_x;
    ry = xi_y;
    rz = xi_z;
    r = xi;
}

// The bond geometry of the force and energy passes for the model.
inline void bondGeometry( const Sys& s, int64_t i, int64_t j, double& xi, double& r, double& st, double& rx,
                          double& ry, double& rz )
{
    if ( isLinear( s.d ) )
        getLinearizedDistance( s, i, j, xi, r, st, rx, ry, rz );
    else
        getDistance( s, i, j, xi, r, st, rx, ry, rz );
}

// ---------------------------------------------------------------------------
// Neighbor list (CabanaPD_Neighbor.hpp:44-66 over Cabana::VerletList with a
// full-neighbor, cell-binned search): j != i with |p_j - p_i|^2 <= rc^2.
// For polydisperse contact the pair cutoff is R_i + R_j + extend.
// ---------------------------------------------------------------------------
void buildNeighbors( Sys& s, const std::vector<double>& p, double rc, bool poly, double extend, int32_t& out_maxn,
                     std::vector<int32_t>& out_cnt, std::vector<int32_t>& out_nbr )
{
    const int64_t n = s.n;
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for ( int64_t i = 0; i < n; i++ )
        for ( int d = 0; d < 3; d++ )
        {
            lo[d] = std::min( lo[d], p[3 * i + d] );
            hi[d] = std::max( hi[d], p[3 * i + d] );
        }
    double rmax = 0.0;
    if ( poly )
        for ( int64_t i = 0; i < n; i++ )
            rmax = std::max( rmax, s.radius[i] );
    const double bin = poly ? ( 2.0 * rmax + extend ) : rc;
    int64_t nc[3];
    for ( int d = 0; d < 3; d++ )
    {
        double ext = hi[d] - lo[d];
        nc[d] = std::max<int64_t>( 1, static_cast<int64_t>( ext / bin ) );
        if ( nc[d] > 1024 )
            nc[d] = 1024;
    }
    auto cellOf = [&]( int64_t i, int d )
    {
        double ext = hi[d] - lo[d];
        int64_t c = ext > 0 ? static_cast<int64_t>( ( p[3 * i + d] - lo[d] ) / ext * nc[d] ) : 0;
        return std::min<int64_t>( std::max<int64_t>( c, 0 ), nc[d] - 1 );
    };
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    std::vector<int64_t> cstart( ncell + 1, 0 ), cell( n );
    for ( int64_t i = 0; i < n; i++ )
    {
        cell[i] = ( cellOf( i, 2 ) * nc[1] + cellOf( i, 1 ) ) * nc[0] + cellOf( i, 0 );
        cstart[cell[i] + 1]++;
    }
    for ( int64_t c = 0; c < ncell; c++ )
        cstart[c + 1] += cstart[c];
    std::vector<int64_t> fill( cstart.begin(), cstart.end() - 1 ), sorted( n );
    for ( int64_t i = 0; i < n; i++ )
        sorted[fill[cell[i]]++] = i;

    const double rc2 = rc * rc;
    auto visit = [&]( int64_t i, auto&& fn )
    {
        int64_t ci[3] = { cellOf( i, 0 ), cellOf( i, 1 ), cellOf( i, 2 ) };
        for ( int64_t cz = std::max<int64_t>( 0, ci[2] - 1 ); cz <= std::min( nc[2] - 1, ci[2] + 1 ); cz++ )
            for ( int64_t cy = std::max<int64_t>( 0, ci[1] - 1 ); cy <= std::min( nc[1] - 1, ci[1] + 1 ); cy++ )
                for ( int64_t cx = std::max<int64_t>( 0, ci[0] - 1 ); cx <= std::min( nc[0] - 1, ci[0] + 1 ); cx++ )
                {
                    int64_t c = ( cz * nc[1] + cy ) * nc[0] + cx;
                    for ( int64_t q = cstart[c]; q < cstart[c + 1]; q++ )
                    {
                        int64_t j = sorted[q];
                        if ( j == i )
                            continue;
                        const double dx = p[3 * j + 0] - p[3 * i + 0];
                        const double dy = p[3 * j + 1] - p[3 * i + 1];
                        const double dz = p[3 * j + 2] - p[3 * i + 2];
                        const double dist2 = dx * dx + dy * dy + dz * dz;
                        double cut2 = rc2;
                        if ( poly )
                        {
                            const double cut = s.radius[i] + s.radius[j] + extend;
                            cut2 = cut * cut;
                        }
                        if ( dist2 <= cut2 )
                            fn( j );
                    }
                }
    };

    out_cnt.assign( n, 0 );
#pragma omp parallel for schedule( dynamic, 256 )
    for ( int64_t i = 0; i < n; i++ )
    {
        int32_t c = 0;
        visit( i, [&]( int64_t ) { c++; } );
        out_cnt[i] = c;
    }
    int32_t maxn = 0;
    for ( int64_t i = 0; i < n; i++ )
        maxn = std::max( maxn, out_cnt[i] );
    out_maxn = maxn;
    out_nbr.assign( static_cast<size_t>( n ) * std::max( maxn, 1 ), 0 );
#pragma omp parallel for schedule( dynamic, 256 )
    for ( int64_t i = 0; i < n; i++ )
    {
        int32_t* row = out_nbr.data() + i * maxn;
        int32_t c = 0;
        visit( i, [&]( int64_t j ) { row[c++] = static_cast<int32_t>( j ); } );
        std::sort( row, row + c );
    }
}
void buildNeighbors( Sys& s, const std::vector<double>& p, double rc, bool poly, double extend )
{
    buildNeighbors( s, p, rc, poly, extend, s.maxn, s.cnt, s.nbr );
    s.built = true;
}

// ---------------------------------------------------------------------------
// Prenotch geometry (CabanaPD_Prenotch.hpp:24-172).
// ---------------------------------------------------------------------------
struct V3
{
    double a[3];
};
inline double dot( V3 a, V3 b )
{
    double v = 0;
    for ( int d = 0; d < 3; d++ )
        v += ( a.a[d] ) * ( b.a[d] );
    return v;
}
inline double norm( V3 a ) { return std::sqrt( dot( a, a ) ); }
inline V3 cross( V3 a, V3 b )
{
    V3 c;
    c.a[0] = a.a[1] * b.a[2] - a.a[2] * b.a[1];
    c.a[1] = -( a.a[0] * b.a[2] - a.a[2] * b.a[0] );
    c.a[2] = a.a[0] * b.a[1] - a.a[1] * b.a[0];
    return c;
}
inline V3 scale( V3 a, double s )
{
    for ( int d = 0; d < 3; d++ )
        a.a[d] *= s;
    return a;
}
inline V3 diff( V3 a, V3 b )
{
    V3 c;
    for ( int d = 0; d < 3; d++ )
        c.a[d] = a.a[d] - b.a[d];
    return c;
}
inline V3 sum( V3 a, V3 b )
{
    V3 c;
    for ( int d = 0; d < 3; d++ )
        c.a[d] = a.a[d] + b.a[d];
    return c;
}
inline int linePlaneIntersection( V3 p0, V3 nrm, V3 l0, V3 l, double tol = 1e-10 )
{
    if ( std::abs( dot( l, nrm ) ) < tol )
    {
        if ( std::abs( dot( diff( p0, l0 ), nrm ) ) < tol )
            return 1;
        else
            return 2;
    }
    return 3;
}
int bondPrenotchIntersection( V3 v1, V3 v2, V3 p0, V3 x_i, V3 x_j, double tol = 1e-10 )
{
    auto cross_v1_v2 = cross( v1, v2 );
    double norm_cross_v1_v2 = norm( cross_v1_v2 );
    auto nrm = scale( cross_v1_v2, 1.0 / norm_cross_v1_v2 );
    auto l0 = x_i;
    auto l = diff( x_j, x_i );
    int case_flag = linePlaneIntersection( p0, nrm, l0, l );
    int keep_bond = 1;
    if ( case_flag == 1 )
    {
        double norm2_cross_v1_v2 = norm_cross_v1_v2 * norm_cross_v1_v2;
        double li1 = dot( cross( diff( x_i, p0 ), v2 ), cross_v1_v2 ) / norm2_cross_v1_v2;
        double li2 = -dot( cross( diff( x_i, p0 ), v1 ), cross_v1_v2 ) / norm2_cross_v1_v2;
        double lj1 = dot( cross( diff( x_j, p0 ), v2 ), cross_v1_v2 ) / norm2_cross_v1_v2;
        double lj2 = -dot( cross( diff( x_j, p0 ), v1 ), cross_v1_v2 ) / norm2_cross_v1_v2;
        if ( !( std::fmin( li1, lj1 ) > 1 + tol || std::fmax( li1, lj1 ) < -tol ||
                std::fmin( li2, lj2 ) > 1 + tol || std::fmax( li2, lj2 ) < -tol ) )
            keep_bond = 0;
    }
    else if ( case_flag == 3 )
    {
        auto dd = dot( diff( p0, l0 ), nrm ) / dot( l, nrm );
        if ( -tol < dd && dd < 1 + tol )
        {
            auto p = sum( l0, scale( l, dd ) );
            double norm2_cross_v1_v2 = norm_cross_v1_v2 * norm_cross_v1_v2;
            double l1 = dot( cross( diff( p, p0 ), v2 ), cross_v1_v2 ) / norm2_cross_v1_v2;
            double l2 = -dot( cross( diff( p, p0 ), v1 ), cross_v1_v2 ) / norm2_cross_v1_v2;
            if ( -tol < std::fmin( l1, l2 ) && std::fmax( l1, l2 ) < 1 + tol )
                keep_bond = 0;
        }
    }
    return keep_bond;
}

// ---------------------------------------------------------------------------
// PD force passes.
// ---------------------------------------------------------------------------
inline double influence( const Sys& s, double xi )
{
    // force_models/CabanaPD_LPS.hpp:85-92
    if ( s.d.influence == 1 )
        return 1.0 / xi;
    else
        return 1.0;
}

// Force<LPS,*>::computeWeightedVolume (force/CabanaPD_LPS.hpp:89-112, 287-329)
void computeWeightedVolume( Sys& s )
{
    if ( !isLps( s.d ) )
        return;
    const bool frac = s.d.fracture != 0;
    std::fill( s.m.begin(), s.m.end(), 0.0 );
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            double xi, r, st, rx, ry, rz;
            getDistance( s, i, j, xi, r, st, rx, ry, rz );
            // model( WeightedVolumeTag ) (force_models/CabanaPD_LPS.hpp:100-106)
            const double wv = influence( s, xi ) * xi * xi * s.vol[j];
            if ( frac )
                s.m[i] += s.mu[i * s.maxn + q] * wv;
            else
                s.m[i] += wv;
        }
    }
}

// Force<LPS,*>::computeDilatation (force/CabanaPD_LPS.hpp:114-141, 331-378)
void computeDilatation( Sys& s )
{
    if ( !isLps( s.d ) )
        return;
    const bool frac = s.d.fracture != 0;
    std::fill( s.theta.begin(), s.theta.end(), 0.0 );
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            double xi, r, st, rx, ry, rz;
            getDistance( s, i, j, xi, r, st, rx, ry, rz );
            // model( DilatationTag ) (force_models/CabanaPD_LPS.hpp:108-116)
            const double theta_i = influence( s, xi ) * st * xi * xi * s.vol[j];
            const double dil = 3.0 * theta_i / s.m[i];
            if ( frac )
            {
                if ( s.m[i] > 0 )
                    s.theta[i] += s.mu[i * s.maxn + q] * dil;
            }
            else
                s.theta[i] += dil;
        }
    }
}

// model( ForceCoeffTag ) for LPS (force_models/CabanaPD_LPS.hpp:118-134)
// Model of the bond (i, j) in a multi-material specimen: the model of the
// material pair, the averaged model between different materials
// (CabanaPD_ForceModelsMulti.hpp:257-273, 338-372).
inline int32_t matOf( const Sys& s, int64_t i )
{
    return s.pair_tables ? s.mtype[i] : 0;
}
inline int32_t pairOf( const Sys& s, int64_t i, int64_t j )
{
    return 4 * matOf( s, i ) + matOf( s, j );
}
inline double pairC( const Sys& s, int64_t i, int64_t j )
{
    return s.pair_tables ? s.pair_c[pairOf( s, i, j )] : s.k.c;
}
inline double pairBcc( const Sys& s, int64_t i, int64_t j )
{
    return s.pair_tables ? s.pair_bcc[pairOf( s, i, j )] : s.k.bond_break_coeff;
}
// LPS coefficients of the bond's model: `side` 0 for particle i's terms, 1 for
// particle j's (force_models/CabanaPD_LPS.hpp:125-128, 143-144).
inline double thetaCoeff( const Sys& s, int64_t i, int64_t j, int side )
{
    return s.pair_tables ? s.pair_tc[pairOf( s, i, j )][side] : s.k.theta_coeff;
}
inline double sCoeff( const Sys& s, int64_t i, int64_t j, int side )
{
    return s.pair_tables ? s.pair_sc[pairOf( s, i, j )][side] : s.k.s_coeff;
}

inline double lpsForceCoeff( const Sys& s, int64_t i, int64_t j, double st, double xi, double vol, double m_i,
                             double m_j, double theta_i, double theta_j )
{
    const double infl = influence( s, xi );
    const double theta_coeff_i = thetaCoeff( s, i, j, 0 );
    const double theta_coeff_j = thetaCoeff( s, i, j, 1 );
    const double s_coeff_i = sCoeff( s, i, j, 0 );
    const double s_coeff_j = sCoeff( s, i, j, 1 );
    return ( theta_coeff_i * theta_i / m_i + theta_coeff_j * theta_j / m_j +
             st * ( s_coeff_i / m_i + s_coeff_j / m_j ) ) *
           infl * xi * vol;
}

// BaseTemperatureModel<TemperatureDependent>: temp_avg and the stretch
// correction (CabanaPD_ForceModels.hpp:203-211); the identity when
// TemperatureIndependent.
inline double thermalStretch( const Sys& s, int64_t i, int64_t j, double st )
{
    if ( !isThermal( s.d ) )
        return st;
    double temp_avg = 0.5 * ( s.temp[i] + s.temp[j] ) - s.d.temp0;
    return st - ( s.d.alpha * temp_avg );
}

// model( CriticalStretchTag, i, j, r, xi ): BaseFractureModel
// (CabanaPD_ForceModels.hpp:131-136), the averaged pair coefficient of a
// multi-material specimen, or ThermalFractureModel (:246-255).
inline bool criticalStretch( const Sys& s, int64_t i, int64_t j, double r, double xi )
{
    if ( isThermal( s.d ) )
    {
        double temp_avg = 0.5 * ( s.temp[i] + s.temp[j] ) - s.d.temp0;
        double bond_break_coeff =
            ( 1.0 + s.k.s0 + s.d.alpha * temp_avg ) * ( 1.0 + s.k.s0 + s.d.alpha * temp_avg );
        return r * r >= bond_break_coeff * xi * xi;
    }
    return r * r >= pairBcc( s, i, j ) * xi * xi;
}

// model( ForceCoeffTag, i, j, s, vol, n ) of the PMB family: Elastic
// (force_models/CabanaPD_PMB.hpp:57-62), or EPP with the update of bond
// (i, slot)'s plastic stretch (:104-121).
inline double pmbForceCoeff( Sys& s, int64_t i, int64_t j, int32_t slot, double st, double vol )
{
    const double c = pairC( s, i, j );
    if ( !isEpp( s.d ) )
        return c * st * vol;
    const double s_Y = s.k.s_Y;
    double& s_p = s.sp[i * s.maxn + slot];
    if ( st >= s_p + s_Y )
        s_p = st - s_Y;
    else if ( st <= s_p - s_Y )
        s_p = st + s_Y;
    return c * ( st - s_p ) * vol;
}

// model( EnergyTag, i, j, s, xi, vol, n ) of the PMB family: Elastic
// (force_models/CabanaPD_PMB.hpp:64-71) or EPP (:123-146).
inline double pmbEnergy( const Sys& s, int64_t i, int64_t j, int32_t slot, double st, double xi, double vol )
{
    const double c = pairC( s, i, j );
    if ( !isEpp( s.d ) )
        return 0.25 * c * st * st * xi * vol;
    const double s_Y = s.k.s_Y;
    const double s_p = s.sp[i * s.maxn + slot];
    double stretch_term;
    if ( st >= s_p + s_Y )
        stretch_term = s_Y * ( 2.0 * st - s_Y );
    else if ( st <= s_p - s_Y )
        stretch_term = s_Y * ( -2.0 * st - s_Y );
    else
        stretch_term = st * st;
    return 0.25 * c * stretch_term * xi * vol;
}

// Force<PMB|LPS, NoFracture|Fracture>::computeForceFull
void computeForceFull( Sys& s )
{
    const bool frac = s.d.fracture != 0;
    const bool lps = isLps( s.d );
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        uint8_t* mu = frac ? s.mu.data() + i * s.maxn : nullptr;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            double fx_i = 0.0, fy_i = 0.0, fz_i = 0.0;
            double xi, r, st, rx, ry, rz;
            bondGeometry( s, i, j, xi, r, st, rx, ry, rz );
            st = thermalStretch( s, i, j, st ); // force/CabanaPD_PMB.hpp:109, 257, 442
            if ( !frac )
            {
                // force/CabanaPD_PMB.hpp:105-118, force/CabanaPD_LPS.hpp:161-174
                double coeff;
                if ( lps )
                    coeff = lpsForceCoeff( s, i, j, st, xi, s.vol[j], s.m[i], s.m[j], s.theta[i], s.theta[j] );
                else
                    coeff = pmbForceCoeff( s, i, j, q, st, s.vol[j] );
                fx_i = coeff * rx / r;
                fy_i = coeff * ry / r;
                fz_i = coeff * rz / r;
                s.f[3 * i + 0] += fx_i;
                s.f[3 * i + 1] += fy_i;
                s.f[3 * i + 2] += fz_i;
            }
            else
            {
                // force/CabanaPD_PMB.hpp:259-279; force/CabanaPD_LPS.hpp:416-438
                // (the LPS no-fail test reads nofail(i) twice, :417-418).
                const bool nofail_j = lps ? ( s.nofail[i] != 0 ) : ( s.nofail[j] != 0 );
                if ( criticalStretch( s, i, j, r, xi ) && !( s.nofail[i] != 0 ) && !nofail_j )
                {
                    mu[q] = 0;
                }
                else if ( mu[q] > 0 )
                {
                    double coeff;
                    if ( lps )
                        coeff = lpsForceCoeff( s, i, j, st, xi, s.vol[j], s.m[i], s.m[j], s.theta[i], s.theta[j] );
                    else
                        coeff = pmbForceCoeff( s, i, j, q, st, s.vol[j] );
                    double muij = mu[q];
                    fx_i = muij * coeff * rx / r;
                    fy_i = muij * coeff * ry / r;
                    fz_i = muij * coeff * rz / r;
                    s.f[3 * i + 0] += fx_i;
                    s.f[3 * i + 1] += fy_i;
                    s.f[3 * i + 2] += fz_i;
                }
            }
        }
    }
}

// Force<*>::computeEnergyFull (force/CabanaPD_PMB.hpp:127-155, 289-339;
// force/CabanaPD_LPS.hpp:183-219, 449-504)
void computeEnergyFull( Sys& s )
{
    const bool frac = s.d.fracture != 0;
    const bool lps = isLps( s.d );
    double Phi = 0.0, D = 0.0;
    std::fill( s.W.begin(), s.W.end(), 0.0 );
#pragma omp parallel for schedule( static ) reduction( + : Phi, D )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        const uint8_t* mu = frac ? s.mu.data() + i * s.maxn : nullptr;
        const int32_t nn = s.cnt[i];
        double num_bonds = 0.0;
        if ( lps )
        {
            if ( frac )
                for ( int32_t q = 0; q < nn; q++ )
                    num_bonds += static_cast<double>( mu[q] );
            else
                num_bonds = static_cast<double>( nn );
        }
        double phi_i = 0.0, vol_H_i = 0.0;
        for ( int32_t q = 0; q < nn; q++ )
        {
            const int64_t j = row[q];
            double xi, r, st, rx, ry, rz;
            bondGeometry( s, i, j, xi, r, st, rx, ry, rz );
            st = thermalStretch( s, i, j, st ); // force/CabanaPD_PMB.hpp:143, 319, 478
            double w;
            if ( lps )
            {
                // model( EnergyTag ) (force_models/CabanaPD_LPS.hpp:136-148)
                const double infl = influence( s, xi );
                const double theta_coeff_i = thetaCoeff( s, i, j, 0 );
                const double s_coeff_i = sCoeff( s, i, j, 0 );
                w = 1.0 / num_bonds * 0.5 * theta_coeff_i / 3.0 * ( s.theta[i] * s.theta[i] ) +
                    0.5 * ( s_coeff_i / s.m[i] ) * infl * st * st * xi * xi * s.vol[j];
            }
            else
            {
                // model( EnergyTag ) (force_models/CabanaPD_PMB.hpp:64-71, 123-146)
                w = pmbEnergy( s, i, j, q, st, xi, s.vol[j] );
            }
            if ( frac )
            {
                w = mu[q] * w;
                s.W[i] += w;
                phi_i += mu[q] * s.vol[j];
                vol_H_i += s.vol[j];
            }
            else
            {
                s.W[i] += w;
                Phi += w * s.vol[i];
            }
        }
        if ( frac )
        {
            Phi += s.W[i] * s.vol[i];
            s.phi[i] = 1 - phi_i / vol_H_i;
            D += s.phi[i];
        }
    }
    s.total_W = Phi;
    s.total_D = frac ? D : 0.0;
}

// Force<*>::computeStressFull (force/CabanaPD_PMB.hpp:157-202, 341-401,
// 492-536; force/CabanaPD_LPS.hpp:221-267, 506-572, 677-723) after
// computeStress's reset (CabanaPD_Force.hpp:221-236); stress(i, a, b) at
// stress[9 i + 3 a + b].
void computeStressFull( Sys& s )
{
    const bool frac = s.d.fracture != 0;
    const bool lps = isLps( s.d );
    s.stress.assign( 9 * s.n, 0.0 );
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        const uint8_t* mu = frac ? s.mu.data() + i * s.maxn : nullptr;
        double* st9 = s.stress.data() + 9 * i;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            // LPS fracture: unbroken bonds only (force/CabanaPD_LPS.hpp:535)
            if ( lps && frac && !( mu[q] > 0 ) )
                continue;
            double xi, r, st, rx, ry, rz;
            bondGeometry( s, i, j, xi, r, st, rx, ry, rz );
            st = thermalStretch( s, i, j, st ); // force/CabanaPD_PMB.hpp:176, 373, 510
            const double xi_x = s.x[3 * j + 0] - s.x[3 * i + 0];
            const double xi_y = s.x[3 * j + 1] - s.x[3 * i + 1];
            const double xi_z = s.x[3 * j + 2] - s.x[3 * i + 2];
            double coeff;
            if ( lps )
            {
                coeff = lpsForceCoeff( s, i, j, st, xi, s.vol[j], s.m[i], s.m[j], s.theta[i], s.theta[j] );
                coeff *= 0.5;
            }
            else
                coeff = 0.5 * pmbForceCoeff( s, i, j, q, st, s.vol[j] ); // EPP: updates s_p of every bond
            double fx_i, fy_i, fz_i;
            if ( frac )
            {
                const double muij = mu[q];
                fx_i = muij * coeff * rx / r;
                fy_i = muij * coeff * ry / r;
                fz_i = muij * coeff * rz / r;
            }
            else
            {
                fx_i = coeff * rx / r;
                fy_i = coeff * ry / r;
                fz_i = coeff * rz / r;
            }
            st9[0] += fx_i * xi_x;
            st9[4] += fy_i * xi_y;
            st9[8] += fz_i * xi_z;
            st9[1] += fx_i * xi_y;
            st9[3] += fy_i * xi_x;
            st9[2] += fx_i * xi_z;
            st9[6] += fz_i * xi_x;
            st9[5] += fy_i * xi_z;
            st9[7] += fz_i * xi_y;
        }
    }
}

// ---------------------------------------------------------------------------
// Correctly rounded powers of the contact models.  The reference evaluates
// Kokkos::pow (the platform's libm, within an ulp of the exact value); the
// oracle and the engine both use these routines, built from IEEE operations
// only (+ - * /, sqrt, fma, frexp/ldexp), so they return the same bits on the
// CPU and the GPU.  Method: double-double arithmetic (Dekker's exact
// products, here through fma) carries each power to ~2^-100 relative before
// the one final rounding (Muller et al., Handbook of Floating-Point
// Arithmetic, ch. 11 on correct rounding): the result is the correctly
// rounded value unless the exact value lies within ~2^-100 of a rounding
// boundary.
//   crSq(x)    = x^2        (one product: exact rounding of x^2)
//   crCube(x)  = x^3        (x^2 as a double-double, times x)
//   crHalf(x)  = x^0.5      (sqrt; +0 at -0, as pow)
//   crPow32(x) = x^1.5      (sqrt with its exact remainder, times x)
//   crThird(x) = x^y, y = the double nearest 1/3 = (1 - 2^-54) / 3:
//                cbrt(x) in double-double (Newton steps, the last with the
//                exact double-double cube) times x^(-2^-54 / 3) =
//                1 - t + t^2/2, t = ln(x) / (3 2^54), ln from an atanh series.
// ---------------------------------------------------------------------------
double crSq( double x )
{
    return x * x;
}
double crCube( double x )
{
    if ( !( std::fabs( x ) < 1e100 ) || x == 0.0 )
        return x * x * x; // overflow, infinities, NaN and signed zeros as pow
    const double ph = x * x, pl = std::fma( x, x, -ph );
    const double qh = ph * x, ql = std::fma( ph, x, -qh );
    return qh + ( ql + pl * x );
}
double crHalf( double x )
{
    return x == 0.0 ? 0.0 : std::sqrt( x );
}
double crPow32( double x )
{
    if ( !( x > 0.0 ) )
        return x == 0.0 ? 0.0 : std::sqrt( -1.0 ); // pow(0, 1.5) = 0; NaN for x < 0 and NaN
    if ( x == HUGE_VAL )
        return x;
    const double sq = std::sqrt( x );
    const double rem = std::fma( -sq, sq, x ); // x - sq^2, exact
    const double sl = rem / ( 2.0 * sq );      // sqrt(x) = sq + sl to ~2^-106
    const double ph = x * sq, pl = std::fma( x, sq, -ph );
    return ph + ( pl + x * sl );
}
double crLog( double x ) // x > 0 normal; ~1e-16 relative, the same operations everywhere
{
    int e = 0;
    double m = std::frexp( x, &e ); // x = m 2^e, m in [0.5, 1)
    if ( m < 0.70710678118654752440 )
    {
        m = 2.0 * m;
        e = e - 1;
    }
    const double z = ( m - 1.0 ) / ( m + 1.0 ); // |z| <= 0.1716
    const double z2 = z * z;
    double sum = 1.0 / 23.0;
    for ( int k = 21; k >= 1; k -= 2 )
        sum = sum * z2 + 1.0 / k;
    return static_cast<double>( e ) * 0.69314718055994530942 + 2.0 * z * sum;
}
double crThird( double x )
{
    if ( x != x || x < 0.0 )
        return std::sqrt( -1.0 );
    if ( x == 0.0 )
        return 0.0; // +0 for -0, as pow
    if ( x == HUGE_VAL )
        return x;
    int e = 0;
    const double m0 = std::frexp( x, &e ); // x = m0 2^e, m0 in [0.5, 1)
    // m = m0 2^r with e - r divisible by 3, m in [0.5, 4)
    int r = e % 3;
    if ( r < 0 )
        r += 3;
    const double m = std::ldexp( m0, r );
    const int q = ( e - r ) / 3;
    // Newton for c^3 = m from a cubic start (cbrt on [0.5, 1), relative
    // error below 1.4e-4, times 2^(r/3)): three steps reach ~1 ulp
    double c = 0.44019728644131806 + m0 * ( 0.9266652504745495 + m0 * ( -0.5110337078357359 + m0 * 0.1442771968559388 ) );
    c = c * ( r == 0 ? 1.0 : ( r == 1 ? 1.2599210498948732 : 1.5874010519681994 ) );
    for ( int it = 0; it < 3; it++ )
        c = c - ( c * c * c - m ) / ( 3.0 * c * c );
    // one Newton step on the double-double cube: m - c^3 exactly to ~2^-106
    const double ph = c * c, pl = std::fma( c, c, -ph );
    const double qh = ph * c, ql = std::fma( ph, c, -qh );
    const double res = ( ( m - qh ) - ql ) - pl * c;
    const double cl = res / ( 3.0 * ph );
    // x^(-2^-54/3) = exp(-t), t = ln(x) / (3 2^54)
    const double t = crLog( x ) * ( 1.0 / 54043195528445952.0 );
    const double k = -t + 0.5 * ( t * t );
    const double chi = std::ldexp( c, q ), clo = std::ldexp( cl, q );
    return chi + ( clo + chi * k );
}

// ---------------------------------------------------------------------------
// Contact (force/CabanaPD_Contact.hpp:109-172).
// ---------------------------------------------------------------------------
struct PairK
{
    double Rs, coeff_h_n, a0, delta_tear, two_r;
};

double jkrPatchRadius( const Sys& s, const PairK& pk, double delta_n )
{
    // force_models/CabanaPD_HertzianJKR.hpp:141-176
    const double Rs = pk.Rs;
    const double nu = s.d.nu, E = s.d.E, Gamma = s.k.Gamma;
    double c0 = crSq( Rs * delta_n );
    double c1 = -4.0 * ( 1.0 - crSq( nu ) ) * kPi * Gamma * crSq( Rs ) / E;
    double c2 = -2.0 * Rs * delta_n;
    double P = -crSq( c2 ) / 12.0 - c0;
    double Q = -crCube( c2 ) / 108.0 + c0 * c2 / 3.0 - crSq( c1 ) / 8.0;
    double U = crThird( -Q / 2.0 + crHalf( std::max( 0.0, crSq( Q ) / 4.0 + crCube( P ) / 27.0 ) ) );
    double sv = 0.0;
    if ( P != 0.0 )
        sv = -5.0 / 6.0 * c2 + U - P / ( 3.0 * U );
    else
        sv = -5.0 / 6.0 * c2 - crThird( Q );
    double w = crHalf( std::max( 0.0, c2 + 2.0 * sv ) );
    double lambda = c1 / ( 2.0 * w + 1e-14 );
    return 1.0 / 2.0 * ( w + crHalf( std::max( 0.0, crSq( w ) - 4.0 * ( c2 + sv + lambda ) ) ) );
}

inline PairK pairConstants( const Sys& s, int64_t i, int64_t j )
{
    PairK pk;
    if ( !s.d.polydisperse )
    {
        pk.Rs = s.k.Rs;
        pk.coeff_h_n = s.k.coeff_h_n;
        pk.a0 = s.k.a0;
        pk.delta_tear = s.k.delta_tear;
        pk.two_r = 2.0 * s.d.radius;
        return pk;
    }
    // Polydisperse extension: R* = R_i R_j / (R_i + R_j); equal radii give
    // the reference's Rs = R/2 exactly only up to rounding, so uniform specs
    // take the branch above.
    const double Ri = s.radius[i], Rj = s.radius[j];
    pk.Rs = Ri * Rj / ( Ri + Rj );
    pk.two_r = Ri + Rj;
    if ( s.d.model == PDK_MODEL_HERTZ )
    {
        pk.coeff_h_n = 4.0 / 3.0 * s.k.Es * std::sqrt( pk.Rs );
        pk.a0 = pk.delta_tear = 0.0;
    }
    else
    {
        pk.coeff_h_n = 4.0 / 3.0 * s.k.Es / pk.Rs;
        pk.a0 = crThird( 9.0 / 2.0 * s.k.Gamma * kPi / s.k.Es * crSq( pk.Rs ) );
        pk.delta_tear = -1.0 / 2.0 * std::pow( 6.0, -1.0 / 3.0 ) * crSq( pk.a0 ) / pk.Rs;
    }
    return pk;
}

// Effective-mass term of the damping: (rho vol)/2 for equal particles
// (Hertzian.hpp:79, HertzianJKR.hpp:101); m_i m_j/(m_i + m_j) when polydisperse.
inline double pairMass( const Sys& s, int64_t i, int64_t j )
{
    if ( !s.d.polydisperse )
        return ( s.rho[i] * s.vol[i] ) / 2.0;
    const double mi = s.rho[i] * s.vol[i], mj = s.rho[j] * s.vol[j];
    return mi * mj / ( mi + mj );
}

double normalForce( const Sys& s, const PairK& pk, double ms, double r, double vn, double vol )
{
    if ( s.d.model == PDK_MODEL_HERTZ )
    {
        // force_models/CabanaPD_Hertzian.hpp:58-82
        const double delta_n = ( r - pk.two_r );
        double coeff = 0.0;
        if ( delta_n < 0.0 )
            coeff = std::min( 0.0, -pk.coeff_h_n * crPow32( std::abs( delta_n ) ) );
        coeff /= vol;
        double Sn = 0.0;
        if ( delta_n < 0.0 )
            Sn = 2.0 * s.k.Es * std::sqrt( pk.Rs * std::abs( delta_n ) );
        coeff -= s.k.coeff_h_d * std::sqrt( Sn * ms ) * vn / vol;
        return coeff;
    }
    // force_models/CabanaPD_HertzianJKR.hpp:82-113
    const double delta_n = ( pk.two_r - r );
    double coeff = 0.0;
    if ( delta_n >= pk.delta_tear )
    {
        auto a = jkrPatchRadius( s, pk, delta_n );
        auto a3 = crCube( a );
        coeff = -1.0 * ( pk.coeff_h_n * a3 - crHalf( 8.0 * s.k.Gamma * kPi * s.k.Es * a3 ) );
        double Sn = 2.0 * s.k.Es * a;
        coeff -= s.k.coeff_h_d * std::sqrt( Sn * ms ) * vn;
        if ( delta_n <= 0.0 && vn <= 0.0 )
            coeff = 0.0;
    }
    coeff /= vol;
    return coeff;
}

void contactUpdateNeighbors( Sys& s, bool require )
{
    // Neighbor::update (CabanaPD_Neighbor.hpp:106-132): rebuild on current
    // positions y = x + u with cutoff() (no tolerance) when the maximum
    // displacement since the last build exceeds radius_extend.
    bool periodic = s.d.rebuild_every > 0 && s.steps_since_build >= s.d.rebuild_every;
    if ( s.max_disp > s.d.radius_extend || require || periodic )
    {
        std::vector<double> y( 3 * s.n );
        for ( int64_t i = 0; i < 3 * s.n; i++ )
            y[i] = s.x[i] + s.u[i];
        const double cutoff = 2.0 * s.d.radius + s.d.radius_extend;
        buildNeighbors( s, y, cutoff, s.d.polydisperse != 0, s.d.radius_extend );
        s.u_neigh = s.u;
        s.steps_since_build = 0;
        s.max_disp = 0.0;
    }
}

void computeContactForceFull( Sys& s )
{
    contactUpdateNeighbors( s, false );
    const bool jkr = s.d.model == PDK_MODEL_HERTZ_JKR;
    const double mu_f = s.d.mu_friction;
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            double xi, r, st, rx, ry, rz;
            getDistance( s, i, j, xi, r, st, rx, ry, rz );
            double nx = rx / r, ny = ry / r, nz = rz / r;
            const double vx = s.v[3 * j + 0] - s.v[3 * i + 0];
            const double vy = s.v[3 * j + 1] - s.v[3 * i + 1];
            const double vz = s.v[3 * j + 2] - s.v[3 * i + 2];
            double vn = vx * nx + vy * ny + vz * nz;
            PairK pk = pairConstants( s, i, j );
            double fn = normalForce( s, pk, pairMass( s, i, j ), r, vn, s.vol[i] );
            s.f[3 * i + 0] += fn * nx;
            s.f[3 * i + 1] += fn * ny;
            s.f[3 * i + 2] += fn * nz;
            double vtx = vx - vn * nx, vty = vy - vn * ny, vtz = vz - vn * nz;
            double ftx = 0.0, fty = 0.0, ftz = 0.0;
            if ( jkr )
            {
                // force_models/CabanaPD_HertzianJKR.hpp:114-129
                auto mu_c = [&]( const double vv ) { return std::abs( mu_f * std::tanh( 10 * vv ) ); };
                auto vt = std::sqrt( vtx * vtx + vty * vty + vtz * vtz );
                auto fna = std::abs( fn );
                ftx = ( vt > 0 ) ? mu_c( vtx ) * fna * vtx / vt : 0.0;
                fty = ( vt > 0 ) ? mu_c( vty ) * fna * vty / vt : 0.0;
                ftz = ( vt > 0 ) ? mu_c( vtz ) * fna * vtz / vt : 0.0;
            }
            s.f[3 * i + 0] += ftx;
            s.f[3 * i + 1] += fty;
            s.f[3 * i + 2] += ftz;
        }
    }
}

// Contact list of a PD specimen on current positions (the Neighbor ctor with
// a contact model, CabanaPD_Neighbor.hpp:60-64, and Neighbor::update,
// :106-132): cutoff 2 radius + radius_extend (+ tol at the first build).
void pdContactBuild( Sys& s, double tol )
{
    std::vector<double> y( 3 * s.n );
    for ( int64_t i = 0; i < 3 * s.n; i++ )
        y[i] = s.x[i] + s.u[i];
    buildNeighbors( s, y, 2.0 * s.cd.radius + s.cd.radius_extend + tol, false, 0.0, s.cmaxn, s.ccnt, s.cnbr );
    s.u_neigh = s.u;
    s.max_disp = 0.0;
}

// Force<NormalRepulsionModel>::computeForceFull (force/CabanaPD_Contact.hpp:44-80)
// with NormalRepulsionModel::forceCoeff (force_models/CabanaPD_Contact.hpp:82-89),
// added to f (computeForce with reset = false, CabanaPD_Solver.hpp:364-366).
void computeNormalRepulsion( Sys& s )
{
    if ( s.max_disp > s.cd.radius_extend )
        pdContactBuild( s, 0.0 );
    const double radius = s.cd.radius, delta = s.cd.horizon, c = s.crep;
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.cnbr.data() + i * s.cmaxn;
        for ( int32_t q = 0; q < s.ccnt[i]; q++ )
        {
            const int64_t j = row[q];
            double fcx_i = 0.0, fcy_i = 0.0, fcz_i = 0.0;
            double xi, r, st, rx, ry, rz;
            getDistance( s, i, j, xi, r, st, rx, ry, rz );
            if ( r < radius )
            {
                const double sc = ( r - radius ) / delta;
                const double coeff = 15.0 * c * sc * s.vol[j];
                fcx_i = coeff * rx / r;
                fcy_i = coeff * ry / r;
                fcz_i = coeff * rz / r;
            }
            s.f[3 * i + 0] += fcx_i;
            s.f[3 * i + 1] += fcy_i;
            s.f[3 * i + 2] += fcz_i;
        }
    }
}

// ---------------------------------------------------------------------------
// Integrator (CabanaPD_Integrate.hpp:88-193).
// ---------------------------------------------------------------------------
void initialHalfStep( Sys& s )
{
    const double dt = s.d.dt;
    const double half_dt = 0.5 * dt;
    // VelocityVerlet<Contact> tracks the largest displacement since the last
    // contact list build (DEM specimens and PD specimens with contact)
    const bool contact = isContact( s.d ) || s.pdc;
    double max_u = 0.0;
#pragma omp parallel for schedule( static ) reduction( max : max_u )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const double half_dt_m = half_dt / s.rho[i];
        for ( int d = 0; d < 3; d++ )
            s.v[3 * i + d] += half_dt_m * s.f[3 * i + d];
        for ( int d = 0; d < 3; d++ )
            s.u[3 * i + d] += dt * s.v[3 * i + d];
        if ( contact )
        {
            const double a = s.u[3 * i + 0] - s.u_neigh[3 * i + 0];
            const double b = s.u[3 * i + 1] - s.u_neigh[3 * i + 1];
            const double c = s.u[3 * i + 2] - s.u_neigh[3 * i + 2];
            const double u_mag = std::sqrt( a * a + b * b + c * c );
            if ( u_mag > max_u )
                max_u = u_mag;
        }
    }
    if ( contact )
        s.max_disp = max_u;
}

void finalHalfStep( Sys& s )
{
    const double half_dt = 0.5 * s.d.dt;
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const double half_dt_m = half_dt / s.rho[i];
        for ( int d = 0; d < 3; d++ )
            s.v[3 * i + d] += half_dt_m * s.f[3 * i + d];
    }
}

// Force BC (CabanaPD_Boundary.hpp:94-113, 138-157) then body term
// (CabanaPD_BodyTerm.hpp:46-63).  The engine applies the body term first.
void applyForceTerms( Sys& s )
{
    if ( !s.body.empty() )
    {
#pragma omp parallel for schedule( static )
        for ( int64_t i = 0; i < 3 * s.n; i++ )
            s.f[i] += s.body[i];
    }
    if ( s.bc_kind >= 0 )
    {
        for ( int64_t b = 0; b < static_cast<int64_t>( s.bc_idx.size() ); b++ )
        {
            const int64_t pid = s.bc_idx[b];
            for ( int d = 0; d < 3; d++ )
            {
                if ( s.bc_kind == PDK_BC_FORCE_VALUE )
                    s.f[3 * pid + d] = s.bc_value;
                else
                    s.f[3 * pid + d] += s.bc_value;
            }
        }
    }
}

void computeForce( Sys& s, bool reset )
{
    // computeForce (CabanaPD_Force.hpp:176-198)
    if ( reset )
        std::fill( s.f.begin(), s.f.end(), 0.0 );
    if ( isContact( s.d ) )
        computeContactForceFull( s );
    else
        computeForceFull( s );
}

// computeHeatTransfer (CabanaPD_HeatTransfer.hpp:155-182): conduction per
// particle over its (intact, with fracture :126) bonds, then forward Euler
// with dt_th (HeatTransfer::forwardEuler, :66-83).
void computeHeatTransfer( Sys& s, double dt )
{
    const bool frac = s.d.fracture != 0;
    const double thermal_coeff = s.k.thermal_coeff;
    const double thermal_horizon = s.d.horizon;
    std::vector<double> conduction( s.n, 0.0 );
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
    {
        const int32_t* row = s.nbr.data() + i * s.maxn;
        const uint8_t* mu = frac ? s.mu.data() + i * s.maxn : nullptr;
        for ( int32_t q = 0; q < s.cnt[i]; q++ )
        {
            const int64_t j = row[q];
            double xi, r, st, rx, ry, rz;
            getDistance( s, i, j, xi, r, st, rx, ry, rz );
            if ( frac && !( mu[q] > 0 ) )
                continue;
            // microconductivity_function (CabanaPD_ForceModels.hpp:302-308)
            const double coeff =
                s.d.linear_microconductivity ? 4.0 * thermal_coeff * ( 1.0 - xi / thermal_horizon ) : thermal_coeff;
            conduction[i] += coeff * ( s.temp[j] - s.temp[i] ) / xi / xi * s.vol[j];
        }
    }
#pragma omp parallel for schedule( static )
    for ( int64_t i = 0; i < s.n; i++ )
        s.temp[i] += dt / s.rho[i] / s.d.cp * conduction[i];
}

// A boundary condition or body term with force_update == false applied at
// time t (CabanaPD_Solver.hpp:255-257, 350-352): member m gets a + b t.
void applyPrescribed( Sys::Prescribed& pr, std::vector<double>& dst, double t )
{
    const int32_t nc = pr.ncomp;
    for ( size_t m = 0; m < pr.idx.size(); m++ )
        for ( int32_t c = 0; c < nc; c++ )
        {
            const size_t e = m * nc + c;
            dst[nc * pr.idx[m] + c] = pr.b.empty() ? pr.a[e] : pr.a[e] + pr.b[e] * t;
        }
}

void applyPrescribed( Sys& s, double t )
{
    applyPrescribed( s.pre_u, s.u, t );
    applyPrescribed( s.pre_temp, s.temp, t );
}

void updateForce( Sys& s )
{
    // Solver::updateForce (CabanaPD_Solver.hpp:462-477)
    if ( s.d.fracture && isLps( s.d ) )
        computeWeightedVolume( s );
    computeDilatation( s );
    computeForce( s, true );
}

} // namespace

struct orc_sys_s : Sys
{
};

extern "C" {

void orc_set_threads( int32_t nthreads )
{
    g_threads = nthreads;
#ifdef _OPENMP
    if ( nthreads > 0 )
        omp_set_num_threads( nthreads );
#endif
}

int32_t orc_get_threads( void )
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_model_constants( const pdk_model_desc* desc, pdk_model_constants* out )
{
    return modelConstants( *desc, *out );
}

int orc_critical_timestep( double dx, double dy, double dz, double K, double rho, double horizon,
                           double safety_factor, double* dt_out )
{
    // Inputs::computeCriticalTimeStep (CabanaPD_Input.hpp:217-336); m from
    // Inputs ctor (CabanaPD_Input.hpp:71-83).
    double v_p = dx * dy * dz;
    double sum = 0;
    int m = static_cast<int>( std::floor( horizon / dx ) );
    double c = 18.0 * K / ( kPi * horizon * horizon * horizon * horizon );
    for ( int i = -( m + 1 ); i < m + 2; i++ )
    {
        double xi_1 = i * dx;
        for ( int j = -( m + 1 ); j < m + 2; j++ )
        {
            double xi_2 = j * dy;
            for ( int k = -( m + 1 ); k < m + 2; k++ )
            {
                double xi_3 = k * dz;
                double r2 = xi_1 * xi_1 + xi_2 * xi_2 + xi_3 * xi_3;
                if ( r2 < horizon * horizon + 1e-10 )
                {
                    if ( r2 > 0 )
                    {
                        double xi = std::sqrt( r2 );
                        sum += v_p * c / xi;
                    }
                }
            }
        }
    }
    double dt_crit = std::sqrt( 2.0 * rho / sum );
    *dt_out = safety_factor * dt_crit;
    return 0;
}

int orc_create( const pdk_model_desc* desc, int64_t n, const double* x, const double* u,
                const double* v, const double* vol, const double* rho, const int32_t* nofail,
                const double* radius, orc_sys* out )
{
    if ( !desc || !out || n < 0 || !x || !vol )
        return 1;
    orc_sys s = new ( std::nothrow ) orc_sys_s();
    if ( !s )
        return 2;
    s->d = *desc;
    if ( modelConstants( s->d, s->k ) )
    {
        delete s;
        return 3;
    }
    s->n = n;
    s->x.assign( x, x + 3 * n );
    s->u.assign( 3 * n, 0.0 );
    s->v.assign( 3 * n, 0.0 );
    if ( u )
        s->u.assign( u, u + 3 * n );
    if ( v )
        s->v.assign( v, v + 3 * n );
    s->f.assign( 3 * n, 0.0 );
    s->vol.assign( vol, vol + n );
    s->rho.assign( n, 1.0 );
    if ( rho )
        s->rho.assign( rho, rho + n );
    s->nofail.assign( n, 0 );
    if ( nofail )
        s->nofail.assign( nofail, nofail + n );
    s->radius.assign( n, desc->radius );
    if ( radius )
        s->radius.assign( radius, radius + n );
    s->W.assign( n, 0.0 );
    s->phi.assign( n, 0.0 );
    s->m.assign( n, 0.0 );
    s->theta.assign( n, 0.0 );
    s->u_neigh.assign( 3 * n, 0.0 );
    s->temp.assign( n, 0.0 );
    *out = s;
    return 0;
}

int orc_destroy( orc_sys s )
{
    delete s;
    return 0;
}

int orc_set_state( orc_sys s, const double* u, const double* v )
{
    if ( u )
        s->u.assign( u, u + 3 * s->n );
    if ( v )
        s->v.assign( v, v + 3 * s->n );
    std::fill( s->f.begin(), s->f.end(), 0.0 );
    return 0;
}

int orc_build_neighbors( orc_sys s )
{
    const double tol = s->d.neighbor_tol;
    if ( isContact( s->d ) )
    {
        // Neighbor ctor on current positions with cutoff() + tol
        // (CabanaPD_Neighbor.hpp:60-64; ContactModel::cutoff, force_models/CabanaPD_Contact.hpp:47).
        std::vector<double> y( 3 * s->n );
        for ( int64_t i = 0; i < 3 * s->n; i++ )
            y[i] = s->x[i] + s->u[i];
        buildNeighbors( *s, y, 2.0 * s->d.radius + s->d.radius_extend + tol, s->d.polydisperse != 0,
                        s->d.radius_extend + tol );
        s->u_neigh = s->u;
        s->steps_since_build = 0;
    }
    else
    {
        buildNeighbors( *s, s->x, s->d.horizon + tol, false, 0.0 );
    }
    if ( s->d.fracture )
        s->mu.assign( static_cast<size_t>( s->n ) * std::max( s->maxn, 1 ), 1 );
    // BasePlasticity::updateBonds after the neighbor list (CabanaPD_Solver.hpp:183-185)
    if ( isEpp( s->d ) )
        s->sp.assign( static_cast<size_t>( s->n ) * std::max( s->maxn, 1 ), 0.0 );
    return 0;
}

int orc_bond_stats( orc_sys s, int64_t* total, int32_t* max_neighbors )
{
    int64_t t = 0;
    for ( int64_t i = 0; i < s->n; i++ )
        t += s->cnt[i];
    if ( total )
        *total = t;
    if ( max_neighbors )
        *max_neighbors = s->maxn;
    return 0;
}

int orc_prenotch( orc_sys s, int32_t nn, const double* p0, const double* v1, const double* v2 )
{
    if ( !s->d.fracture || !s->built )
        return 1;
    for ( int32_t p = 0; p < nn; p++ )
    {
        V3 P0{ { p0[3 * p], p0[3 * p + 1], p0[3 * p + 2] } };
        V3 A{ { v1[3 * p], v1[3 * p + 1], v1[3 * p + 2] } };
        V3 B{ { v2[3 * p], v2[3 * p + 1], v2[3 * p + 2] } };
#pragma omp parallel for schedule( static )
        for ( int64_t i = 0; i < s->n; i++ )
        {
            V3 xi{ { s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2] } };
            const int32_t* row = s->nbr.data() + i * s->maxn;
            for ( int32_t q = 0; q < s->cnt[i]; q++ )
            {
                const int64_t j = row[q];
                V3 xj{ { s->x[3 * j], s->x[3 * j + 1], s->x[3 * j + 2] } };
                if ( !bondPrenotchIntersection( A, B, P0, xi, xj ) )
                    s->mu[i * s->maxn + q] = 0;
            }
        }
    }
    return 0;
}

int orc_get_bonds( orc_sys s, int64_t* offsets, int64_t* nbr, uint8_t* intact )
{
    int64_t off = 0;
    for ( int64_t i = 0; i < s->n; i++ )
    {
        if ( offsets )
            offsets[i] = off;
        for ( int32_t q = 0; q < s->cnt[i]; q++ )
        {
            if ( nbr )
                nbr[off + q] = s->nbr[i * s->maxn + q];
            if ( intact )
                intact[off + q] = s->d.fracture ? s->mu[i * s->maxn + q] : 1;
        }
        off += s->cnt[i];
    }
    if ( offsets )
        offsets[s->n] = off;
    return 0;
}

int orc_set_bc( orc_sys s, int32_t kind, const int64_t* idx, int64_t n, double value )
{
    if ( kind != PDK_BC_FORCE_VALUE && kind != PDK_BC_FORCE_UPDATE )
        return 1;
    s->bc_kind = kind;
    s->bc_idx.assign( idx, idx + n );
    s->bc_value = value;
    return 0;
}

int orc_set_body_force( orc_sys s, const double* b )
{
    if ( b )
        s->body.assign( b, b + 3 * s->n );
    else
        s->body.clear();
    return 0;
}

int orc_compute_weighted_volume( orc_sys s )
{
    computeWeightedVolume( *s );
    return 0;
}
int orc_compute_dilatation( orc_sys s )
{
    computeDilatation( *s );
    return 0;
}
int orc_compute_force( orc_sys s, int32_t reset )
{
    computeForce( *s, reset != 0 );
    return 0;
}
int orc_compute_energy( orc_sys s, double* W, double* D )
{
    if ( !isContact( s->d ) )
        computeEnergyFull( *s );
    if ( W )
        *W = s->total_W;
    if ( D )
        *D = s->total_D;
    return 0;
}

int orc_init( orc_sys s )
{
    // Solver::init(bc) (CabanaPD_Solver.hpp:229-279)
    applyPrescribed( *s, 0.0 );
    if ( !s->d.fracture && isLps( s->d ) )
        computeWeightedVolume( *s );
    updateForce( *s );
    if ( !isContact( s->d ) )
        computeEnergyFull( *s );
    applyForceTerms( *s );
    return 0;
}

int orc_step( orc_sys s, int32_t nsteps )
{
    // Solver::runStep(step, bc) (CabanaPD_Solver.hpp:332-375)
    for ( int32_t k = 0; k < nsteps; k++ )
    {
        s->steps_done++;
        initialHalfStep( *s );
        if ( isContact( s->d ) )
            s->steps_since_build++;
        // heat transfer every thermal_subcycle steps (CabanaPD_Solver.hpp:340-347)
        if ( isDynamic( s->d ) && s->steps_done % s->d.thermal_subcycle == 0 )
            computeHeatTransfer( *s, s->d.thermal_subcycle * s->d.dt );
        applyPrescribed( *s, s->steps_done * s->d.dt );
        updateForce( *s );
        if ( s->pdc )
            computeNormalRepulsion( *s );
        applyForceTerms( *s );
        finalHalfStep( *s );
    }
    return 0;
}

int orc_set_contact( orc_sys s, const pdk_contact_desc* cd )
{
    if ( !s || !cd || isContact( s->d ) || !s->built || cd->model != PDK_CONTACT_NORMAL_REPULSION )
        return 1;
    s->pdc = true;
    s->cd = *cd;
    // NormalRepulsionModel ctor (force_models/CabanaPD_Contact.hpp:70-80)
    s->crep = 18.0 * cd->K / ( kPi * cd->horizon * cd->horizon * cd->horizon * cd->horizon );
    pdContactBuild( *s, s->d.neighbor_tol );
    return 0;
}

int orc_cr_pow( int32_t which, int64_t n, const double* x, double* out )
{
    for ( int64_t i = 0; i < n; i++ )
        out[i] = which == 0 ? crPow32( x[i] ) : which == 1 ? crThird( x[i] ) : which == 2 ? crCube( x[i] )
                                                                                           : crHalf( x[i] );
    return 0;
}

int orc_contact_pairs( orc_sys s, int64_t* pairs )
{
    int64_t t = 0;
    if ( s->pdc )
        for ( int32_t c : s->ccnt )
            t += c;
    *pairs = t;
    return 0;
}

namespace
{
// Install a pair table of ntypes^2 models and the particle types.
int installPairs( orc_sys s, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type, bool tables )
{
    for ( int32_t a = 0; a < ntypes; a++ )
        for ( int32_t b = 0; b < ntypes; b++ )
        {
            const pdk_pair_model& p = pairs[a * ntypes + b];
            s->pair_c[4 * a + b] = p.c;
            s->pair_bcc[4 * a + b] = p.bond_break_coeff;
            for ( int side = 0; side < 2; side++ )
            {
                s->pair_tc[4 * a + b][side] = p.theta_coeff[side];
                s->pair_sc[4 * a + b][side] = p.s_coeff[side];
            }
        }
    s->mtype.assign( s->n, 0 );
    if ( tables )
        for ( int64_t i = 0; i < s->n; i++ )
        {
            const int32_t t = type ? type[i] : 0;
            if ( t < 0 || t >= ntypes )
                return 1;
            s->mtype[i] = t;
        }
    s->nmat = ntypes;
    s->pair_tables = tables;
    return 0;
}
} // namespace

int orc_set_materials( orc_sys s, int32_t nmat, const pdk_material* mats, const int32_t* type )
{
    // createMultiForceModel( particles, AverageTag, m_0, .., m_{nmat-1} )
    // (CabanaPD_ForceModelsMulti.hpp:360-502): the pair (a, a) uses material
    // a's model, a pair a != b the model averaged from a's and b's: PMB c =
    // (c_a + c_b) / 2 (force_models/CabanaPD_PMB.hpp:51-55), s0 = sqrt((s0_a^2
    // K_a + s0_b^2 K_b) / (K_a + K_b)), bond_break_coeff = (1 + s0)^2
    // (CabanaPD_ForceModels.hpp:114-121), LPS theta_coeff / s_coeff [0] from
    // a's model and [1] from b's (force_models/CabanaPD_LPS.hpp:55-69).
    if ( !s || nmat < 1 || nmat > 4 || !mats || ( nmat > 1 && !type ) || isContact( s->d ) )
        return 1;
    double K[4], c[4], s0[4], bcc[4], tc[4], sc[4];
    for ( int32_t t = 0; t < nmat; t++ )
    {
        pdk_model_desc d = s->d;
        d.K = mats[t].K;
        d.G = mats[t].G;
        d.G0 = mats[t].G0;
        pdk_model_constants k;
        if ( modelConstants( d, k ) )
            return 1;
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
    return installPairs( s, nmat, pairs, type, nmat > 1 );
}

int orc_set_pair_models( orc_sys s, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type )
{
    // createMultiForceModel( particles, indexing, m_0, .. )
    // (CabanaPD_ForceModelsMulti.hpp:351-358): the bond (i, j) calls the model
    // m_{indexing(type(i), type(j))} (:228-250), whose constants the caller
    // has placed at pairs[type(i) * ntypes + type(j)].
    if ( !s || ntypes < 1 || ntypes > 4 || !pairs || ( ntypes > 1 && !type ) || isContact( s->d ) )
        return 1;
    return installPairs( s, ntypes, pairs, type, true );
}

int orc_set_prescribed( orc_sys s, int32_t field, const int64_t* idx, int64_t count, const double* a,
                        const double* b )
{
    if ( !s || count < 0 || ( count > 0 && ( !idx || !a ) ) )
        return 1;
    Sys::Prescribed* pr;
    int32_t nc;
    if ( field == PDK_FIELD_U )
    {
        pr = &s->pre_u;
        nc = 3;
    }
    else if ( field == PDK_FIELD_TEMPERATURE && isThermal( s->d ) )
    {
        pr = &s->pre_temp;
        nc = 1;
    }
    else
        return 1;
    std::vector<uint8_t> seen( s->n, 0 );
    for ( int64_t m = 0; m < count; m++ )
    {
        if ( idx[m] < 0 || idx[m] >= s->n || seen[idx[m]] )
            return 1;
        seen[idx[m]] = 1;
    }
    pr->ncomp = nc;
    pr->idx.assign( idx, idx + count );
    pr->a.assign( a, a + nc * count );
    if ( b )
        pr->b.assign( b, b + nc * count );
    else
        pr->b.clear();
    return 0;
}

int orc_set_field( orc_sys s, int32_t field, const double* values )
{
    if ( !s || !values || field != PDK_FIELD_TEMPERATURE || !isThermal( s->d ) )
        return 1;
    s->temp.assign( values, values + s->n );
    return 0;
}

int orc_get_plastic( orc_sys s, double* out )
{
    // EPP plastic stretch per bond, in the order of orc_get_bonds
    if ( !s || !out || !isEpp( s->d ) || !s->built )
        return 1;
    int64_t off = 0;
    for ( int64_t i = 0; i < s->n; i++ )
        for ( int32_t q = 0; q < s->cnt[i]; q++ )
            out[off++] = s->sp[i * s->maxn + q];
    return 0;
}

int orc_compute_stress( orc_sys s )
{
    if ( isContact( s->d ) || !s->built )
        return 1;
    computeStressFull( *s );
    return 0;
}

int orc_get( orc_sys s, int32_t field, double* out )
{
    const std::vector<double>* src = nullptr;
    switch ( field )
    {
    case PDK_FIELD_STRESS:
        if ( s->stress.empty() )
        {
            std::fill( out, out + 9 * s->n, 0.0 );
            return 0;
        }
        src = &s->stress;
        break;
    case PDK_FIELD_X: src = &s->x; break;
    case PDK_FIELD_U: src = &s->u; break;
    case PDK_FIELD_V: src = &s->v; break;
    case PDK_FIELD_F: src = &s->f; break;
    case PDK_FIELD_VOL: src = &s->vol; break;
    case PDK_FIELD_RHO: src = &s->rho; break;
    case PDK_FIELD_W: src = &s->W; break;
    case PDK_FIELD_PHI: src = &s->phi; break;
    case PDK_FIELD_M: src = &s->m; break;
    case PDK_FIELD_THETA: src = &s->theta; break;
    case PDK_FIELD_TEMPERATURE: src = &s->temp; break;
    default: return 1;
    }
    std::copy( src->begin(), src->end(), out );
    return 0;
}

int orc_bond_pmb( const pdk_model_desc* desc, const double* xi3, const double* eta3, double vol_j,
                  double* f3, double* w, int32_t* breaks )
{
    // One bond of Force<PMB,*>::computeForceFull from particle i at the origin.
    pdk_model_constants k;
    if ( modelConstants( *desc, k ) )
        return 1;
    const double xi_x = xi3[0], xi_y = xi3[1], xi_z = xi3[2];
    const double rx = xi_x + eta3[0], ry = xi_y + eta3[1], rz = xi_z + eta3[2];
    const double r = std::sqrt( rx * rx + ry * ry + rz * rz );
    const double xi = std::sqrt( xi_x * xi_x + xi_y * xi_y + xi_z * xi_z );
    const double st = ( r - xi ) / xi;
    const double coeff = k.c * st * vol_j;
    f3[0] = coeff * rx / r;
    f3[1] = coeff * ry / r;
    f3[2] = coeff * rz / r;
    *w = 0.25 * k.c * st * st * xi * vol_j;
    *breaks = desc->fracture ? ( r * r >= k.bond_break_coeff * xi * xi ) : 0;
    return 0;
}

int orc_bond_prenotch( const double* v1, const double* v2, const double* p0, const double* xi,
                       const double* xj, int32_t* keep )
{
    V3 A{ { v1[0], v1[1], v1[2] } }, B{ { v2[0], v2[1], v2[2] } }, P{ { p0[0], p0[1], p0[2] } };
    V3 I{ { xi[0], xi[1], xi[2] } }, J{ { xj[0], xj[1], xj[2] } };
    *keep = bondPrenotchIntersection( A, B, P, I, J );
    return 0;
}

int orc_jkr_normal_force( const pdk_model_desc* desc, double r, double vn, double vol, double rho,
                          double* fn )
{
    Sys s;
    s.d = *desc;
    if ( modelConstants( s.d, s.k ) )
        return 1;
    s.d.polydisperse = 0;
    PairK pk;
    pk.Rs = s.k.Rs;
    pk.coeff_h_n = s.k.coeff_h_n;
    pk.a0 = s.k.a0;
    pk.delta_tear = s.k.delta_tear;
    pk.two_r = 2.0 * desc->radius;
    *fn = normalForce( s, pk, ( rho * vol ) / 2.0, r, vn, vol );
    return 0;
}

int orc_jkr_patch_radius( const pdk_model_desc* desc, double delta_n, double* a )
{
    Sys s;
    s.d = *desc;
    if ( modelConstants( s.d, s.k ) )
        return 1;
    PairK pk;
    pk.Rs = s.k.Rs;
    *a = jkrPatchRadius( s, pk, delta_n );
    return 0;
}

} // extern "C"
