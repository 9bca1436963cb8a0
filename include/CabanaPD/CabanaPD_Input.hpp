// CabanaPD_Input.hpp -- Inputs: JSON input file with derived keys
// (reference: src/CabanaPD_Input.hpp:27-402).  Same file format
// ({"key": {"value": ..., "unit": ...}}), same derived keys and truncations.
#ifndef CABANAPD_DROPIN_INPUT_HPP
#define CABANAPD_DROPIN_INPUT_HPP

#include <array>
#include <cctype>
#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <pdk.h>

namespace CabanaPD
{
namespace json
{
// Minimal JSON value: null, bool, number, string, array, object.
struct Value
{
    enum Kind
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool contains( const std::string& k ) const { return kind == Object && obj.count( k ) > 0; }
    const Value& at( const std::string& k ) const
    {
        auto it = obj.find( k );
        if ( kind != Object || it == obj.end() )
            throw std::runtime_error( "Inputs: missing key '" + k + "'" );
        return it->second;
    }
    Value& operator[]( const std::string& k )
    {
        if ( kind == Null )
            kind = Object;
        return obj[k];
    }
    Value& operator[]( const char* k ) { return ( *this )[std::string( k )]; }
    Value& operator[]( std::size_t i )
    {
        if ( kind == Null )
            kind = Array;
        if ( arr.size() <= i )
            arr.resize( i + 1 );
        return arr[i];
    }
    Value& operator[]( int i ) { return ( *this )[static_cast<std::size_t>( i )]; }
    const Value& operator[]( std::size_t i ) const { return arr.at( i ); }
    const Value& operator[]( int i ) const { return arr.at( static_cast<std::size_t>( i ) ); }
    std::size_t size() const { return kind == Array ? arr.size() : ( kind == Object ? obj.size() : 1 ); }
    bool is_array() const { return kind == Array; }

    static Value number( double v )
    {
        Value x;
        x.kind = Number;
        x.num = v;
        return x;
    }
    static Value string( const std::string& s )
    {
        Value x;
        x.kind = String;
        x.str = s;
        return x;
    }
    static Value boolean( bool v )
    {
        Value x;
        x.kind = Bool;
        x.b = v;
        return x;
    }

    // Conversions used through Inputs::operator[].
    operator double() const
    {
        if ( kind == Number )
            return num;
        if ( kind == Bool )
            return b ? 1.0 : 0.0;
        throw std::runtime_error( "Inputs: value is not a number" );
    }
    operator int() const { return static_cast<int>( static_cast<double>( *this ) ); }
    operator bool() const { return kind == Bool ? b : static_cast<double>( *this ) != 0.0; }
    operator std::string() const
    {
        if ( kind != String )
            throw std::runtime_error( "Inputs: value is not a string" );
        return str;
    }
    template <class T, std::size_t N>
    operator std::array<T, N>() const
    {
        if ( kind != Array || arr.size() != N )
            throw std::runtime_error( "Inputs: value is not an array of the expected size" );
        std::array<T, N> a;
        for ( std::size_t i = 0; i < N; i++ )
            a[i] = static_cast<T>( static_cast<double>( arr[i] ) );
        return a;
    }
};

class Parser
{
  public:
    explicit Parser( const std::string& t ) : s_( t ) {}
    Value parse()
    {
        Value v = value();
        ws();
        if ( p_ != s_.size() )
            fail( "trailing characters" );
        return v;
    }

  private:
    [[noreturn]] void fail( const char* what )
    {
        throw std::runtime_error( std::string( "Inputs: JSON parse error (" ) + what + ") at offset " +
                                  std::to_string( p_ ) );
    }
    void ws()
    {
        while ( p_ < s_.size() && std::isspace( static_cast<unsigned char>( s_[p_] ) ) )
            p_++;
    }
    Value value()
    {
        ws();
        if ( p_ >= s_.size() )
            fail( "unexpected end" );
        char c = s_[p_];
        if ( c == '{' )
            return object();
        if ( c == '[' )
            return array();
        if ( c == '"' )
            return Value::string( string() );
        if ( s_.compare( p_, 4, "true" ) == 0 )
        {
            p_ += 4;
            return Value::boolean( true );
        }
        if ( s_.compare( p_, 5, "false" ) == 0 )
        {
            p_ += 5;
            return Value::boolean( false );
        }
        if ( s_.compare( p_, 4, "null" ) == 0 )
        {
            p_ += 4;
            return Value();
        }
        return number();
    }
    Value object()
    {
        Value v;
        v.kind = Value::Object;
        p_++;
        ws();
        if ( p_ < s_.size() && s_[p_] == '}' )
        {
            p_++;
            return v;
        }
        while ( true )
        {
            ws();
            if ( p_ >= s_.size() || s_[p_] != '"' )
                fail( "expected key" );
            std::string k = string();
            ws();
            if ( p_ >= s_.size() || s_[p_] != ':' )
                fail( "expected ':'" );
            p_++;
            v.obj[k] = value();
            ws();
            if ( p_ < s_.size() && s_[p_] == ',' )
            {
                p_++;
                continue;
            }
            if ( p_ < s_.size() && s_[p_] == '}' )
            {
                p_++;
                return v;
            }
            fail( "expected ',' or '}'" );
        }
    }
    Value array()
    {
        Value v;
        v.kind = Value::Array;
        p_++;
        ws();
        if ( p_ < s_.size() && s_[p_] == ']' )
        {
            p_++;
            return v;
        }
        while ( true )
        {
            v.arr.push_back( value() );
            ws();
            if ( p_ < s_.size() && s_[p_] == ',' )
            {
                p_++;
                continue;
            }
            if ( p_ < s_.size() && s_[p_] == ']' )
            {
                p_++;
                return v;
            }
            fail( "expected ',' or ']'" );
        }
    }
    std::string string()
    {
        std::string out;
        p_++;
        while ( p_ < s_.size() && s_[p_] != '"' )
        {
            if ( s_[p_] == '\\' && p_ + 1 < s_.size() )
                p_++;
            out += s_[p_++];
        }
        if ( p_ >= s_.size() )
            fail( "unterminated string" );
        p_++;
        return out;
    }
    Value number()
    {
        std::size_t used = 0;
        double d = 0.0;
        try
        {
            d = std::stod( s_.substr( p_ ), &used );
        }
        catch ( ... )
        {
            fail( "bad number" );
        }
        p_ += used;
        return Value::number( d );
    }
    const std::string& s_;
    std::size_t p_ = 0;
};

inline void dump( std::ostream& os, const Value& v )
{
    switch ( v.kind )
    {
    case Value::Null: os << "null"; break;
    case Value::Bool: os << ( v.b ? "true" : "false" ); break;
    case Value::Number:
    {
        std::ostringstream t;
        t.precision( 17 );
        t << v.num;
        os << t.str();
        break;
    }
    case Value::String: os << '"' << v.str << '"'; break;
    case Value::Array:
        os << '[';
        for ( std::size_t i = 0; i < v.arr.size(); i++ )
        {
            if ( i )
                os << ',';
            dump( os, v.arr[i] );
        }
        os << ']';
        break;
    case Value::Object:
    {
        os << '{';
        bool first = true;
        for ( const auto& kv : v.obj )
        {
            if ( !first )
                os << ',';
            first = false;
            os << '"' << kv.first << "\":";
            dump( os, kv.second );
        }
        os << '}';
        break;
    }
    }
}
} // namespace json

inline void log( std::ostream& os ) { os << std::endl; }
template <class T, class... R>
void log( std::ostream& os, const T& t, const R&... r )
{
    os << t;
    log( os, r... );
}

class Inputs
{
  public:
    // Inputs( filename ) (CabanaPD_Input.hpp:30-128)
    explicit Inputs( const std::string filename )
    {
        std::ifstream in( filename );
        if ( !in )
            throw std::runtime_error( "Inputs: cannot open " + filename );
        std::stringstream ss;
        ss << in.rdbuf();
        std::string text = ss.str();
        inputs = json::Parser( text ).parse();
        setupSize();
        double tf = inputs.at( "final_time" ).at( "value" );
        double dt = inputs.at( "timestep" ).at( "value" );
        if ( inputs.contains( "mass_scaling_factor" ) && !inputs.contains( "density_input" ) )
        {
            double ms = inputs.at( "mass_scaling_factor" ).at( "value" );
            double rho0 = inputs.at( "density" ).at( "value" );
            inputs["density_input"]["value"] = json::Value::number( rho0 );
            inputs["density"]["value"] = json::Value::number( rho0 * ms );
            inputs["timestep_input"]["value"] = json::Value::number( dt );
            dt *= std::sqrt( ms );
            inputs["timestep"]["value"] = json::Value::number( dt );
        }
        if ( inputs.contains( "horizon" ) && inputs.contains( "m" ) )
            throw std::runtime_error( "Cannot input both horizon and m." );
        double dx = inputs.at( "dx" ).at( "value" )[0];
        if ( inputs.contains( "horizon" ) )
        {
            double horizon = inputs.at( "horizon" ).at( "value" );
            int m = static_cast<int>( std::floor( horizon / dx ) );
            if ( horizon < dx )
                throw std::runtime_error( "Input horizon must be greater than or equal to dx to result in any "
                                          "neighbour-particles" );
            inputs["m"]["value"] = json::Value::number( m );
        }
        else if ( inputs.contains( "m" ) )
        {
            int m = inputs.at( "m" ).at( "value" );
            if ( m < 1 )
                throw std::runtime_error( "Input m must be greater than or equal to 1." );
            inputs["horizon"]["value"] = json::Value::number( static_cast<double>( m ) * dx + 0.01 * dx );
        }
        if ( !inputs.contains( "timestep_safety_factor" ) )
            inputs["timestep_safety_factor"]["value"] = json::Value::number( 0.85 );
        int num_steps = static_cast<int>( tf / dt );
        inputs["num_steps"]["value"] = json::Value::number( num_steps );
        if ( !inputs.contains( "output_file" ) )
            inputs["output_file"]["value"] = json::Value::string( "cabanaPD.out" );
        if ( !inputs.contains( "error_file" ) )
            inputs["error_file"]["value"] = json::Value::string( "cabanaPD.err" );
        inputs["input_file"]["value"] = json::Value::string( filename );
        if ( !inputs.contains( "output_reference" ) )
            inputs["output_reference"]["value"] = json::Value::boolean( true );
        if ( !inputs.contains( "output_frequency" ) )
            inputs["output_frequency"]["value"] = json::Value::number( num_steps > 0 ? num_steps : 1 );
        std::ofstream out( "cabanaPD.in.json" );
        json::dump( out, inputs );
    }

    // Get a single input (CabanaPD_Input.hpp:367-370).
    const json::Value& operator[]( const std::string& label ) const { return inputs.at( label ).at( "value" ); }
    bool contains( const std::string& label ) const { return inputs.contains( label ); }
    std::string units( const std::string& label ) const
    {
        const auto& v = inputs.at( label );
        return v.contains( "unit" ) ? std::string( v.at( "unit" ) ) : std::string();
    }
    void set( const std::string& label, double value ) { inputs[label]["value"] = json::Value::number( value ); }

    // Critical time step (CabanaPD_Input.hpp:217-336, 385-399): warns when
    // dt exceeds the estimate and stores it.
    template <class ForceModel>
    void computeCriticalTimeStep( const ForceModel& )
    {
        // Inputs::computeCriticalTimeStep (CabanaPD_Input.hpp:217-300): K from
        // bulk_modulus, else from elastic_modulus with nu = 0.25; scalars or
        // one value per material, the material of least rho / K decides.
        const json::Value& dxv = inputs.at( "dx" ).at( "value" );
        std::vector<double> Kv;
        if ( inputs.contains( "bulk_modulus" ) )
            Kv = getVector( "bulk_modulus" );
        else
        {
            Kv = getVector( "elastic_modulus" );
            for ( double& k : Kv )
                k = k / ( 3.0 * ( 1.0 - 2.0 * 0.25 ) );
        }
        const std::vector<double> rhov = getVector( "density" );
        if ( rhov.empty() || rhov.size() != Kv.size() )
            throw std::runtime_error( "density and moduli need one value per material" );
        std::size_t mi = 0;
        for ( std::size_t t = 1; t < rhov.size(); t++ )
            if ( rhov[t] / Kv[t] < rhov[mi] / Kv[mi] )
                mi = t;
        double K = Kv[mi];
        double rho = rhov[mi];
        double horizon = ( *this )["horizon"];
        double safety = ( *this )["timestep_safety_factor"];
        double dt_crit = 0.0;
        if ( pdk_critical_timestep( dxv[0], dxv[1], dxv[2], K, rho, horizon, safety, &dt_crit ) != 0 )
            throw std::runtime_error( pdk_last_error() );
        double dt = ( *this )["timestep"];
        if ( dt > dt_crit )
            log( std::cout, "WARNING: timestep (", dt, ") is larger than estimated stable timestep for mechanics (",
                 dt_crit, "), using safety factor of ", safety, ".\n" );
        inputs["critical_timestep_mechanics"]["value"] = json::Value::number( dt_crit );
    }

  private:
    // A scalar input or each entry of an array input.
    std::vector<double> getVector( const std::string& key ) const
    {
        const json::Value& v = inputs.at( key ).at( "value" );
        std::vector<double> out;
        if ( v.is_array() )
            for ( std::size_t i = 0; i < v.arr.size(); i++ )
                out.push_back( double( v.arr[i] ) );
        else
            out.push_back( double( v ) );
        return out;
    }

    void setupSize()
    {
        if ( inputs.contains( "system_size" ) )
        {
            const json::Value ss = inputs.at( "system_size" ).at( "value" );
            if ( ss.size() != 3 )
                throw std::runtime_error( "CabanaPD requires 3d (system_size)." );
            for ( std::size_t d = 0; d < 3; d++ )
            {
                double s = ss[d];
                inputs["low_corner"]["value"][d] = json::Value::number( -0.5 * s );
                inputs["high_corner"]["value"][d] = json::Value::number( 0.5 * s );
            }
        }
        else if ( inputs.contains( "low_corner" ) && inputs.contains( "high_corner" ) )
        {
            const json::Value lo = inputs.at( "low_corner" ).at( "value" );
            const json::Value hi = inputs.at( "high_corner" ).at( "value" );
            for ( std::size_t d = 0; d < 3; d++ )
                inputs["system_size"]["value"][d] = json::Value::number( double( hi[d] ) - double( lo[d] ) );
        }
        else
            throw std::runtime_error( "Must input either system_size or both low_corner and high_corner." );

        if ( inputs.contains( "dx" ) )
        {
            const json::Value dx = inputs.at( "dx" ).at( "value" );
            for ( std::size_t d = 0; d < 3; d++ )
            {
                double size_d = inputs.at( "system_size" ).at( "value" )[d];
                inputs["num_cells"]["value"][d] = json::Value::number( static_cast<int>( size_d / double( dx[d] ) ) );
            }
        }
        else if ( inputs.contains( "num_cells" ) )
        {
            const json::Value nc = inputs.at( "num_cells" ).at( "value" );
            for ( std::size_t d = 0; d < 3; d++ )
            {
                double size_d = inputs.at( "system_size" ).at( "value" )[d];
                inputs["dx"]["value"][d] = json::Value::number( size_d / double( nc[d] ) );
            }
        }
        else
            throw std::runtime_error( "Must input either num_cells or dx." );
    }

    json::Value inputs;
};

} // namespace CabanaPD

#endif
