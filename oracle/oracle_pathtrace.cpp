// CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// The ray-traced passes of a frame restated in FP64 from the reference
// shaders, in the order of VulkanDevice::Render (Source/VulkanDevice.cpp:
// 567-584):
//   primary G-buffer      RaygenPrimary.inl:280-348, HitInfo.inl:204-466
//   light grid            CmLightGridBuild.comp:46-105, LightGrid.h:27-134
//   initial reservoirs    RtInitialReservoirs.rgen:37-51, RaygenCommon.h:396-486
//   direct illumination   RtRaygenDirect.rgen:37-68, RaygenCommon.h:503-674
//   indirect (init/final) RtRaygenIndirect.inl:66-434, ReservoirIndirect.h
//   gradient samples      RtGradients.rgen:44-181
// Lights, BRDF and sampling follow Light.h, BRDF.h and Random.h.  Only the
// primary rays are formed in FP32 (with the kernel's operation order), so
// that both sides trace the same rays; every other quantity is FP64.  Rays
// are traced with the oracle's own LBVH traversal (or_trace_rays).  The
// random decisions use the same integer hashes and the same 16/8-bit
// uniforms (converted exactly as the kernels convert them); a decision whose
// FP64 and FP32 inputs straddle its threshold can differ, so comparisons
// allow a small fraction of pixels to disagree.
#include "oracle_pathtrace.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace
{
const double PI          = 3.14159265358979323846;
const double MAX_RAY_LEN = 10000.0;
const double MIN_ROUGH   = 0.005; // MIN_GGX_ROUGHNESS, BRDF.h:110

// Random.h:24-32
const uint32_t SALT_DIFF_BOUNCE_BASE       = 8;
const uint32_t SALT_SPEC_BOUNCE_BASE       = 12;
const uint32_t SALT_LIGHT_POINT            = 20;
const uint32_t SALT_LIGHT_GRID_BASE        = 24;
const uint32_t SALT_INITIAL_RESERVOIRS     = 48;
const uint32_t SALT_LIGHT_CHOOSE_DIRECT    = 72;
const uint32_t SALT_LIGHT_CHOOSE_INDIRECT  = 96;
const uint32_t SALT_RESAMPLE_INDIRECT_BASE = 132;

// LIGHT_GRID_SIZE_* and LIGHT_GRID_CELL_SIZE (ShaderCommonC.h)
const int      GX = 16, GY = 16, GZ = 16, GCELL = 128;
const uint32_t GRID_ENTRIES = GX * GY * GZ * GCELL;

const uint32_t LT_NONE = 0, LT_DIRECTIONAL = 1, LT_SPHERE = 2, LT_TRIANGLE = 3, LT_SPOT = 4;
const uint32_t LIGHT_NONE = 32767; // LIGHT_INDEX_NONE

struct D3
{
    double x, y, z;
};
D3     mk( double x, double y, double z ) { return { x, y, z }; }
D3     operator+( D3 a, D3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
D3     operator-( D3 a, D3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
D3     operator-( D3 a ) { return { -a.x, -a.y, -a.z }; }
D3     operator*( D3 a, double s ) { return { a.x * s, a.y * s, a.z * s }; }
D3     operator*( D3 a, D3 b ) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
D3     operator/( D3 a, double s ) { return { a.x / s, a.y / s, a.z / s }; }
double dot( D3 a, D3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3     cross( D3 a, D3 b ) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
double length( D3 a ) { return std::sqrt( dot( a, a ) ); }
D3     normalize( D3 a ) { return a / length( a ); }
double clampd( double x, double lo, double hi ) { return std::min( std::max( x, lo ), hi ); }
double saturate( double x ) { return clampd( x, 0.0, 1.0 ); }
double mixd( double a, double b, double t ) { return a * ( 1.0 - t ) + b * t; }
D3     mix3( D3 a, D3 b, double t ) { return { mixd( a.x, b.x, t ), mixd( a.y, b.y, t ), mixd( a.z, b.z, t ) }; }
double lum( D3 c ) { return 0.2125 * c.x + 0.7154 * c.y + 0.0721 * c.z; } // Utils.h:54-57
double safeRcp( double f ) { return f <= 0.0 ? 0.0 : 1.0 / f; }
double square( double x ) { return x * x; }
D3     safeNormalize( D3 v )
{
    const double l = length( v );
    return l > 0.001 ? v / l : mk( 0, 1, 0 );
}
double smoothstep( double e0, double e1, double x )
{
    const double t = clampd( ( x - e0 ) / ( e1 - e0 ), 0.0, 1.0 );
    return t * t * ( 3.0 - 2.0 * t );
}

template< typename F >
void ParallelFor( int n, int threads, F&& f )
{
    threads = std::max( 1, std::min( threads, 64 ) );
    std::vector< std::thread > ts;
    for( int t = 0; t < threads; t++ )
    {
        ts.emplace_back( [ &, t ] {
            for( int i = t; i < n; i += threads )
            {
                f( i );
            }
        } );
    }
    for( auto& t : ts )
    {
        t.join();
    }
}

// ------------------------------------------------------------ random

uint32_t wellons( uint32_t x ) { return or_wellons_low_bias32( x ); }
// rnd16 / rnd16_2 / rnd8_4 (Random.h:242-287): the FP32 quotient, as the kernels form it
double rnd16( uint32_t seed, uint32_t salt ) { return double( float( wellons( seed + salt ) & 0xFFFFu ) / 65535.0f ); }
void   rnd16_2( uint32_t seed, uint32_t salt, double* a, double* b )
{
    const uint32_t r = wellons( seed + salt );
    *a               = double( float( r & 0xFFFFu ) / 65535.0f );
    *b               = double( float( r >> 16 ) / 65535.0f );
}
void rnd8_4( uint32_t seed, uint32_t salt, double o[ 4 ] )
{
    const uint32_t r = wellons( seed + salt );
    for( int k = 0; k < 4; k++ )
    {
        o[ k ] = double( float( ( r >> ( 8 * k ) ) & 0xFFu ) / 255.0f );
    }
}
uint32_t seedOf( int x, int y, uint32_t frame ) { return or_random_seed( x, y, frame ); }

// ------------------------------------------------------------ framebuffers

struct Fb
{
    const OrFramebuffers* f;
    int                   W, H, GW, GH;
    explicit Fb( const OrFramebuffers* fb ) : f( fb ), W( fb->width ), H( fb->height )
    {
        GW = std::max( 1, ( W + 1 ) / 3 );
        GH = std::max( 1, ( H + 1 ) / 3 );
    }
    int     comps( int p ) const { return or_plane_components( p ); }
    int     pw( int p ) const { return or_plane_layout( p ) == 2 ? GW : W; }
    int     ph( int p ) const { return or_plane_layout( p ) == 2 ? GH : H; }
    double* at( int p, int x, int y, bool prev = false ) const
    {
        double* a = prev ? f->prev[ p ] : f->cur[ p ];
        return a + ( size_t( y ) * pw( p ) + x ) * comps( p );
    }
    // texelFetch with robust access: zeros outside
    double get( int p, int x, int y, int c, bool prev = false ) const
    {
        if( x < 0 || y < 0 || x >= pw( p ) || y >= ph( p ) )
        {
            return 0.0;
        }
        return at( p, x, y, prev )[ c ];
    }
    void set4( int p, int x, int y, double a, double b = 0, double c = 0, double d = 0 ) const
    {
        double*   o = at( p, x, y );
        const int n = comps( p );
        o[ 0 ]      = a;
        if( n > 1 ) o[ 1 ] = b;
        if( n > 2 ) o[ 2 ] = c;
        if( n > 3 ) o[ 3 ] = d;
    }
};

D3 DecodeNormal( double packed )
{
    double n[ 3 ];
    or_decode_normal( uint32_t( packed ), n );
    return mk( n[ 0 ], n[ 1 ], n[ 2 ] );
}
double EncodeNormal( D3 n )
{
    const double a[ 3 ] = { n.x, n.y, n.z };
    return double( or_encode_normal( a ) );
}

// ------------------------------------------------------------ scene

struct Light
{
    uint32_t type = LT_NONE;
    D3       color{ 0, 0, 0 };
    double   d0[ 4 ] = { 0, 0, 0, 0 };
    double   d1[ 4 ] = { 0, 0, 0, 0 };
    double   d2[ 4 ] = { 0, 0, 0, 0 };
};

D3 UnpackColor( uint32_t c )
{
    return mk( double( c & 0xFFu ) / 255.0, double( ( c >> 8 ) & 0xFFu ) / 255.0, double( ( c >> 16 ) & 0xFFu ) / 255.0 );
}

// LightManager.cpp:99-148 (light styles not set: multiplier 1)
Light EncodeDirectional( const RgDirectionalLightUploadInfo& i )
{
    Light        l;
    const D3     d   = mk( i.direction.data[ 0 ], i.direction.data[ 1 ], i.direction.data[ 2 ] );
    const double len = length( d );
    l.type           = LT_DIRECTIONAL;
    l.color          = UnpackColor( i.color ) * double( i.intensity );
    const D3 dn      = len > 0 ? d / len : mk( 0, 0, 0 );
    l.d0[ 0 ] = dn.x, l.d0[ 1 ] = dn.y, l.d0[ 2 ] = dn.z;
    l.d0[ 3 ] = 0.5 * double( i.angularDiameterDegrees ) * PI / 180.0;
    return l;
}
Light EncodeSphere( const RgSphericalLightUploadInfo& i )
{
    Light        l;
    const double r    = std::max( 0.005, double( i.radius ) ); // MIN_SPHERE_RADIUS
    const double area = PI * r * r;
    l.type            = LT_SPHERE;
    l.color           = UnpackColor( i.color ) * ( double( i.intensity ) / area );
    l.d0[ 0 ] = i.position.data[ 0 ], l.d0[ 1 ] = i.position.data[ 1 ], l.d0[ 2 ] = i.position.data[ 2 ];
    l.d0[ 3 ] = r;
    return l;
}
// EncodeAsSpotLight (LightManager.cpp:194-234): cone angles clamped to
// [0, 89] degrees, the inner one at least 1 degree inside the outer one.
Light EncodeSpot( const RgSpotLightUploadInfo& i )
{
    Light        l;
    const double r    = std::max( 0.005, double( i.radius ) );
    const double area = PI * r * r;
    const D3     d    = mk( i.direction.data[ 0 ], i.direction.data[ 1 ], i.direction.data[ 2 ] );
    const double len  = length( d );
    const double deg  = PI / 180.0;
    auto         cl   = [ & ]( double a ) { return std::min( std::max( a, 0.0 ), 89.0 * deg ); };
    const double inner = std::min( double( i.angleInner ), double( i.angleOuter ) - 1.0 * deg );
    l.type             = LT_SPOT;
    l.color            = UnpackColor( i.color ) * ( double( i.intensity ) / area );
    l.d0[ 0 ] = i.position.data[ 0 ], l.d0[ 1 ] = i.position.data[ 1 ], l.d0[ 2 ] = i.position.data[ 2 ], l.d0[ 3 ] = r;
    const D3 dn = len > 0 ? d / len : mk( 0, 0, 0 );
    l.d1[ 0 ] = dn.x, l.d1[ 1 ] = dn.y, l.d1[ 2 ] = dn.z;
    l.d2[ 0 ] = std::cos( cl( inner ) ), l.d2[ 1 ] = std::cos( cl( double( i.angleOuter ) ) );
    return l;
}
// EncodeAsTriangleLight (LightManager.cpp:150-192) with the unnormalised
// normal cross(p1 - p0, p2 - p0) (Utils::GetUnnormalizedNormal, Utils.cpp:346-359)
Light EncodeTriangle( const RgPolygonalLightUploadInfo& i )
{
    Light    l;
    const D3 p0 = mk( i.positions[ 0 ].data[ 0 ], i.positions[ 0 ].data[ 1 ], i.positions[ 0 ].data[ 2 ] );
    const D3 p1 = mk( i.positions[ 1 ].data[ 0 ], i.positions[ 1 ].data[ 1 ], i.positions[ 1 ].data[ 2 ] );
    const D3 p2 = mk( i.positions[ 2 ].data[ 0 ], i.positions[ 2 ].data[ 1 ], i.positions[ 2 ].data[ 2 ] );
    const D3 n  = cross( p1 - p0, p2 - p0 );
    l.type      = LT_TRIANGLE;
    l.color     = UnpackColor( i.color ) * ( double( i.intensity ) / ( 0.5 * length( n ) ) );
    l.d0[ 0 ] = p0.x, l.d0[ 1 ] = p0.y, l.d0[ 2 ] = p0.z, l.d0[ 3 ] = n.x;
    l.d1[ 0 ] = p1.x, l.d1[ 1 ] = p1.y, l.d1[ 2 ] = p1.z, l.d1[ 3 ] = n.y;
    l.d2[ 0 ] = p2.x, l.d2[ 1 ] = p2.y, l.d2[ 2 ] = p2.z, l.d2[ 3 ] = n.z;
    return l;
}

struct Surface
{
    D3       position{ 0, 0, 0 };
    D3       normal{ 0, 0, 0 };
    double   roughness = 1;
    D3       albedo{ 0, 0, 0 };
    bool     isSky     = true;
    D3       specColor{ 0, 0, 0 };
    D3       toViewer{ 0, 0, 0 };
    uint32_t custom    = 1; // instanceCustomIndex of the surface's instance (Surface.inl:31)
};
struct LSample
{
    D3     position{ 0, 0, 0 }, color{ 0, 0, 0 };
    double dw = 0;
};
struct Res
{
    uint32_t sel = LIGHT_NONE;
    double   tp = 0, wsum = 0;
    uint32_t M = 0;
};

D3 SpecColor( D3 albedo, double metallic ) { return mix3( mk( 0.04, 0.04, 0.04 ), albedo, metallic ); }

// ------------------------------------------------------------ sampling (Random.h, BRDF.h, Light.h)

void sampleDisk( double radius, double u1, double u2, double* x, double* y )
{
    u1 *= 0.99;
    u2 *= 0.99;
    const double r = radius * std::sqrt( u1 ), phi = 2 * PI * u2;
    *x             = r * std::cos( phi );
    *y             = r * std::sin( phi );
}
void revisedONB( D3 n, D3* b1, D3* b2 )
{
    if( n.z < 0.0 )
    {
        const double a = 1.0 / ( 1.0 - n.z ), b = n.x * n.y * a;
        *b1            = mk( 1.0 - n.x * n.x * a, -b, n.x );
        *b2            = mk( b, n.y * n.y * a - 1.0, -n.y );
    }
    else
    {
        const double a = 1.0 / ( 1.0 + n.z ), b = -n.x * n.y * a;
        *b1            = mk( 1.0 - n.x * n.x * a, b, -n.x );
        *b2            = mk( b, 1.0 - n.y * n.y * a, -n.y );
    }
}
D3 sampleOrientedHemisphere( D3 n, double u1, double u2, double* oneOverPdf )
{
    double       a = 1.0 - 2.0 * u1, b = std::sqrt( 1.0 - a * a );
    const double phi = 2 * PI * u2;
    a *= 0.98;
    b *= 0.98;
    const D3 r  = normalize( mk( n.x + b * std::cos( phi ), n.y + b * std::sin( phi ), n.z + a ) );
    *oneOverPdf = PI * safeRcp( dot( r, n ) );
    return r;
}
double D_GGX( double nm, double alpha, bool loose )
{
    if( loose ) alpha = std::max( 0.2, alpha );
    const double a2 = alpha * alpha;
    nm              = std::max( 0.0, nm );
    return a2 / PI / square( nm * nm * ( a2 - 1.0 ) + 1.0 );
}
double G1_GGX( double ns, double alpha ) { return 2.0 * ns * safeRcp( ns * ( 2.0 - alpha ) + alpha ); }
D3     evalGGX( D3 n, D3 v, D3 l, double alpha, D3 specColor, bool loose )
{
    alpha           = std::max( alpha, MIN_ROUGH );
    const double nl = std::max( dot( n, l ), 0.0 );
    if( nl <= 0.0 ) return mk( 0, 0, 0 );
    const double f  = std::pow( 1.0 - nl, 5.0 );
    const D3     F  = specColor + ( mk( 1, 1, 1 ) - specColor ) * f;
    const D3     h  = normalize( v + l );
    const double D  = D_GGX( dot( n, h ), alpha, loose );
    const double nv = std::max( dot( n, v ), 0.0 );
    const double G2 = 0.5 / mixd( 2.0 * nl * nv, nl + nv, alpha );
    return F * ( G2 * D );
}
D3 sampleGGXVNDF( D3 v, double alpha, double u1, double u2, double* oneOverPdf, bool loose )
{
    alpha = std::max( alpha, MIN_ROUGH );
    u1 *= 0.98;
    u2 *= 0.98;
    const D3     Vh    = normalize( mk( alpha * v.x, alpha * v.y, v.z ) );
    const double lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const D3     T1    = lensq > 0.0 ? mk( -Vh.y, Vh.x, 0.0 ) * ( 1.0 / std::sqrt( lensq ) ) : mk( 1, 0, 0 );
    const D3     T2    = cross( Vh, T1 );
    const double r = std::sqrt( u1 ), phi = 2 * PI * u2;
    const double t1 = r * std::cos( phi );
    double       t2 = r * std::sin( phi );
    const double s  = 0.5 * ( 1.0 + Vh.z );
    t2              = ( 1.0 - s ) * std::sqrt( 1.0 - t1 * t1 ) + s * t2;
    const D3     Nh = T1 * t1 + T2 * t2 + Vh * std::sqrt( std::max( 0.0, 1.0 - t1 * t1 - t2 * t2 ) );
    const D3     Ne = normalize( mk( alpha * Nh.x, alpha * Nh.y, std::max( 0.02, Nh.z ) ) );
    const double D  = D_GGX( Ne.z, alpha, loose );
    const double G1 = G1_GGX( v.z, alpha );
    *oneOverPdf     = v.z * safeRcp( G1 * std::max( 0.0, dot( v, Ne ) ) * D );
    return Ne;
}
D3 sampleSmithGGX( D3 n, D3 v, double alpha, double u1, double u2, double* oneOverPdf, bool loose )
{
    D3 b0, b1;
    revisedONB( n, &b0, &b1 );
    const D3 ve = mk( dot( b0, v ), dot( b1, v ), dot( n, v ) );
    const D3 m  = sampleGGXVNDF( ve, alpha, u1, u2, oneOverPdf, loose );
    const D3 l  = -ve + m * ( 2.0 * dot( m, ve ) );
    *oneOverPdf *= 4.0 * dot( ve, m );
    return b0 * l.x + b1 * l.y + n * l.z;
}
double safeSolidAngle( double a ) { return ( a > 0.0 && std::isfinite( a ) ) ? clampd( a, 0.0, 4 * PI ) : 0.0; }
double solidAngleSphere( double radius, double dist )
{
    const double s = radius / std::max( radius, dist ), c = std::sqrt( 1.0 - s * s );
    return safeSolidAngle( 2 * PI * ( 1.0 - c ) );
}
double colorWeight( D3 c ) { return clampd( lum( c ) * 0.1 + 0.9, 1.0, 10.0 ); }
D3     P0( const Light& l ) { return mk( l.d0[ 0 ], l.d0[ 1 ], l.d0[ 2 ] ); }
D3     P1( const Light& l ) { return mk( l.d1[ 0 ], l.d1[ 1 ], l.d1[ 2 ] ); }
D3     P2( const Light& l ) { return mk( l.d2[ 0 ], l.d2[ 1 ], l.d2[ 2 ] ); }
D3     TriNormal( const Light& l ) { return normalize( mk( l.d0[ 3 ], l.d1[ 3 ], l.d2[ 3 ] ) ); }
// isSphereInFront (Light.h:111-114)
double inFront( D3 planeNormal, D3 planePos, D3 center, double radius ) { return dot( planeNormal, center - planePos ) > -radius ? 1.0 : 0.0; }
// getLightWeight (Light.h:162-204, :286-296)
double lightWeight( const Light& l, D3 cellCenter, double cellRadius )
{
    switch( l.type )
    {
        case LT_DIRECTIONAL: return colorWeight( l.color );
        case LT_SPHERE:
            return colorWeight( l.color ) * solidAngleSphere( l.d0[ 3 ], std::max( length( P0( l ) - cellCenter ), cellRadius ) );
        case LT_TRIANGLE:
        {
            const D3     c = ( P0( l ) + P1( l ) + P2( l ) ) / 3.0;
            const double r = ( length( P0( l ) - c ) + length( P1( l ) - c ) + length( P2( l ) - c ) ) / 3.0;
            return colorWeight( l.color ) * solidAngleSphere( r, std::max( length( c - cellCenter ), cellRadius ) ) *
                   inFront( TriNormal( l ), c, cellCenter, cellRadius );
        }
        case LT_SPOT:
            return colorWeight( l.color ) * solidAngleSphere( l.d0[ 3 ], std::max( length( P0( l ) - cellCenter ), cellRadius ) ) *
                   inFront( P1( l ), P0( l ), cellCenter, cellRadius );
        default: return 0.0;
    }
}
LSample sampleLight( const Light& l, D3 surfPos, double pu, double pv, double polyFactor )
{
    LSample r;
    if( l.type == LT_DIRECTIONAL )
    {
        const D3     dir = mk( l.d0[ 0 ], l.d0[ 1 ], l.d0[ 2 ] );
        const double rad = std::sin( std::max( 0.01, l.d0[ 3 ] ) );
        double       dx, dy;
        sampleDisk( rad, pu, pv, &dx, &dy );
        D3 b0, b1;
        revisedONB( dir, &b0, &b1 );
        const D3 ln = normalize( dir + b0 * dx + b1 * dy );
        r.position  = surfPos - ln * MAX_RAY_LEN;
        r.color     = l.color;
        r.dw        = 1.0;
    }
    else if( l.type == LT_SPHERE )
    {
        const D3     c   = mk( l.d0[ 0 ], l.d0[ 1 ], l.d0[ 2 ] );
        D3           d   = c - surfPos;
        const double len = length( d );
        d                = d / len;
        double   pdf;
        const D3 ln = sampleOrientedHemisphere( -d, pu, pv, &pdf );
        r.position  = c + ln * l.d0[ 3 ];
        r.color     = l.color;
        r.dw        = solidAngleSphere( l.d0[ 3 ], len );
    }
    else if( l.type == LT_TRIANGLE )
    {
        // sampleTriangle (Random.h:58-69): uniform over the triangle's area
        const double u1 = pu * 0.99, u2 = pv * 0.99;
        const double beta = 1.0 - std::sqrt( u1 ), gamma = ( 1.0 - beta ) * u2, alpha = 1.0 - beta - gamma;
        r.position          = P0( l ) * alpha + P1( l ) * beta + P2( l ) * gamma;
        const D3     n      = TriNormal( l );
        const double area   = 0.5 * length( mk( l.d0[ 3 ], l.d1[ 3 ], l.d2[ 3 ] ) );
        D3           toSurf = surfPos - r.position;
        const double dist   = length( toSurf );
        toSurf              = toSurf / dist;
        // getPolySpotFactor (Light.h:101-105); calcSolidAngleForArea (:153-158)
        r.color = l.color * std::pow( std::max( dot( n, toSurf ), 0.0 ), polyFactor );
        r.dw    = safeSolidAngle( area * std::fabs( dot( n, toSurf ) ) / ( dist * dist ) );
    }
    else if( l.type == LT_SPOT )
    {
        // sampleSpotLight (Light.h:282-297): a point of the disk facing the
        // spot direction; the cone factor smoothstep(cosOuter, cosInner, cosA)^2
        // (getSpotFactor, Light.h:107-110) of the direction to the centre
        const D3 dir = P1( l );
        double   dx, dy;
        sampleDisk( l.d0[ 3 ], pu, pv, &dx, &dy );
        D3 b0, b1;
        revisedONB( dir, &b0, &b1 );
        r.position          = P0( l ) + b0 * dx + b1 * dy;
        const D3     d      = P0( l ) - surfPos;
        const double len    = length( d );
        const double cosA   = std::max( -dot( dir, d ) / len, 0.0 );
        const double t      = clampd( ( cosA - l.d2[ 1 ] ) / ( l.d2[ 0 ] - l.d2[ 1 ] ), 0.0, 1.0 );
        const double smooth = t * t * ( 3.0 - 2.0 * t );
        r.color             = l.color * ( smooth * smooth );
        r.dw                = solidAngleSphere( l.d0[ 3 ], len );
    }
    return r;
}
void shade( const Surface& s, const LSample& lt, double oneOverPdf, bool loose, D3* diff, D3* spec )
{
    const D3     l  = safeNormalize( lt.position - s.position );
    const double nl = dot( s.normal, l );
    if( nl <= 0.0 )
    {
        *diff = *spec = mk( 0, 0, 0 );
        return;
    }
    *diff = lt.color * ( lt.dw * nl / PI ) * oneOverPdf;
    *spec = ( lt.color * ( lt.dw * nl ) ) * evalGGX( s.normal, s.toViewer, l, s.roughness, s.specColor, loose ) * oneOverPdf;
}
double targetPdf( const Surface& s, const LSample& lt, bool loose )
{
    D3 d, sp;
    shade( s, lt, 1.0, loose, &d, &sp );
    return lum( d + sp );
}

// Reservoir.h:24-112
bool   valid( const Res& r ) { return r.sel != LIGHT_NONE; }
double selWeight( const Res& r ) { return safeRcp( r.tp ) * ( r.wsum / double( std::max( 1u, r.M ) ) ); }
void   normalizeRes( Res& r, uint32_t maxM )
{
    r.wsum /= double( std::max( r.M, 1u ) );
    r.M = std::min( r.M, maxM );
    r.wsum *= double( r.M );
}
void update( Res& r, uint32_t light, double tp, double oneOverSource, double rnd )
{
    const double w = tp * oneOverSource;
    r.wsum += w;
    r.M += 1;
    if( rnd * r.wsum < w )
    {
        r.sel = light;
        r.tp  = tp;
    }
}
void combine( Res& c, const Res& b, double rnd )
{
    const double w = b.wsum;
    c.wsum += w;
    c.M += b.M;
    if( rnd * c.wsum < w )
    {
        c.sel = b.sel;
        c.tp  = b.tp;
    }
}
void combineNewSurf( Res& c, const Res& b, double tpB, double rnd )
{
    const double w = tpB * safeRcp( b.tp ) * b.wsum;
    c.wsum += w;
    c.M += b.M;
    if( rnd * c.wsum < w )
    {
        c.sel = b.sel;
        c.tp  = tpB;
    }
}
// packReservoir (Reservoir.h:116-146): M and the index share a 32-bit word
double packSelM( const Res& r ) { return double( ( std::min( r.M, 65535u ) << 16 ) | std::min( r.sel, 65535u ) ); }
void   storeRes( const Fb& fb, int plane, int x, int y, const Res& r )
{
    if( std::isfinite( r.wsum ) )
    {
        fb.set4( plane, x, y, packSelM( r ), r.tp, r.wsum, 0 );
    }
    else
    {
        fb.set4( plane, x, y, double( LIGHT_NONE ), 0, 0, 0 );
    }
}
Res loadRes( const Fb& fb, int plane, int x, int y, bool prev )
{
    const uint32_t sm = uint32_t( fb.get( plane, x, y, 0, prev ) );
    Res            r;
    r.sel  = sm & 65535u;
    r.tp   = fb.get( plane, x, y, 1, prev );
    r.wsum = fb.get( plane, x, y, 2, prev );
    r.M    = ( sm >> 16 ) & 65535u;
    return r;
}

} // namespace

// ======================================================================

struct OrPathTracer
{
    int                     W, H;
    std::vector< uint32_t > blueNoise;
    std::vector< double >   grid[ 2 ]; // 3 doubles per entry: selected, targetPdf, weightSum
    int                     gridParity = 0;
    std::vector< uint64_t > prevLightIds;
    std::vector< uint64_t > prevGeomIds;
};

namespace
{

bool IsNull( const RgExtRay& r );

struct Ctx
{
    const OrPtScene*       sc;
    const OrBvh*           bvh;
    RgExtFrameParams       p;
    Fb                     fb;
    OrPathTracer*          pt;
    std::vector< Light >   lights;     // [0] directional or none
    std::vector< uint32_t > prevToCur;  // last frame's light index -> this frame's
    std::vector< uint32_t > geomPrevToCur;
    uint32_t               lightCount = 0;
    bool                   hasDir     = false;
    int                    threads;
    uint32_t               traceFlags;
};

// out-of-range indices read as an empty light, as the kernels' GetLight does
const Light& LightAt( const Ctx& c, uint32_t i )
{
    static const Light none{};
    return i < c.lights.size() ? c.lights[ i ] : none;
}

// ---- instances and cull masks
// The reference puts each primitive into a BLAS by its flags
// (VertexCollectorFilterTypeFlags_GetForGeometry,
// VertexCollectorFilterType.cpp:200-257) and gives that BLAS's TLAS instance
// a mask and custom-index flags (ASManager::SetupTLASInstanceFromBLAS,
// ASManager.cpp:837-963).  Here each triangle is tested against the ray's
// cull mask on its own, over one tree of all triangles: a triangle is hit
// only if its instance is in the TLAS and (instance mask & cull mask) != 0.
const uint32_t M_WORLD_0 = 1, M_WORLD_1 = 2, M_WORLD_2 = 4, M_REFRACT = 32, M_FP = 64, M_FPV = 128; // ShaderCommonC.h:64-71
const uint32_t CI_DYNAMIC = 1, CI_FP = 2, CI_FPV = 4, CI_SKY = 8;                                  // ShaderCommonC.h:60-63

// The instance mask of a primitive's TLAS instance; *inTlas is false when
// the world it belongs to is not in rayCullMaskWorld.
uint32_t InstanceMask( uint32_t flags, uint32_t worldMask, bool* inTlas )
{
    *inTlas = true;
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON ) return M_FP;
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON_VIEWER ) return M_FPV;
    const uint32_t world = ( flags & RG_MESH_PRIMITIVE_SKY_VISIBILITY ) ? M_WORLD_2 : M_WORLD_0;
    *inTlas              = ( worldMask & world ) != 0;
    const bool alpha     = ( flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) != 0;
    const bool refract   = !alpha && ( flags & ( RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH |
                                                 RG_MESH_PRIMITIVE_ACID ) ) != 0;
    return refract ? M_REFRACT : world;
}

uint32_t CustomIndex( uint32_t flags, bool skyGeometry )
{
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON ) return CI_DYNAMIC | CI_FP;
    if( flags & RG_MESH_PRIMITIVE_FIRST_PERSON_VIEWER ) return CI_DYNAMIC | CI_FPV;
    if( skyGeometry && ( flags & RG_MESH_PRIMITIVE_SKY_VISIBILITY ) ) return CI_DYNAMIC | CI_SKY;
    return CI_DYNAMIC;
}

uint32_t TriCustom( const Ctx& c, uint32_t prim )
{
    return CustomIndex( c.sc->geoms[ c.sc->triGeom[ prim ] ].flags, c.p.skyGeometry != 0 );
}

// doesPayloadContainHitInfo (RaygenCommon.h:191-205)
bool IsSkyHit( const Ctx& c, const RgExtHit& h )
{
    return h.primitiveId == RG_EXT_MISS || ( TriCustom( c, h.primitiveId ) & CI_SKY ) != 0;
}

// getPrimaryVisibilityCullMask .. getIndirectIlluminationCullMask
// (RaygenCommon.h:103-177) and traceVisibility's first-person-viewer
// exception (:350-358)
uint8_t PrimaryMask( const RgExtFrameParams& p ) { return uint8_t( p.rayCullMaskWorld | M_REFRACT | M_FP ); }
uint8_t ReflRefrMask( const RgExtFrameParams& p, uint32_t custom, uint32_t flags, bool refraction )
{
    uint32_t w = p.rayCullMaskWorld | M_REFRACT;
    if( flags & RG_MESH_PRIMITIVE_IGNORE_REFRACT_AFTER )
    {
        w &= ~M_REFRACT;
        if( custom & CI_FP ) return uint8_t( w );
    }
    if( custom & CI_FP ) return uint8_t( w | M_FP );
    return uint8_t( w | ( refraction ? M_FP : M_FPV ) );
}
uint8_t ShadowMask( const RgExtFrameParams& p, uint32_t custom, uint32_t lightIndex )
{
    uint32_t m = p.rayCullMaskWorldShadow;
    if( custom & CI_FP ) m |= M_FP;
    else m |= M_FPV; // first-person viewer and world surfaces alike
    if( lightIndex == p.lightIndexIgnoreFPVShadows ) m &= ~M_FPV;
    return uint8_t( m );
}
uint8_t IndirectMask( const RgExtFrameParams& p, uint32_t custom )
{
    return uint8_t( p.rayCullMaskWorld | ( ( custom & CI_FP ) ? M_FP : M_FPV ) );
}

// RtAlphaTest.rahit: alpha-tested primitives with a texture ignore hits
// where the layer-0 colour times the base-level texel has
// (r+g+b)/3 * a + a < 0.5.
bool AlphaPass( const OrPtScene* sc, uint32_t prim, float u, float v )
{
    const OrGeom& g = sc->geoms[ sc->triGeom[ prim ] ];
    if( !( g.flags & RG_MESH_PRIMITIVE_ALPHA_TESTED ) || g.texture == 0 || g.texture > sc->textureCount ) return true;
    const float* t  = sc->texcoords + size_t( prim ) * 6;
    const double b0 = 1.0 - double( u ) - double( v );
    double       c[ 4 ];
    or_texture_sample_lod( sc->textures[ g.texture - 1 ], t[ 0 ] * b0 + t[ 2 ] * u + t[ 4 ] * v, t[ 1 ] * b0 + t[ 3 ] * u + t[ 5 ] * v, 0.0, c );
    const double r = c[ 0 ] * ( g.color & 0xFFu ) / 255.0, gg = c[ 1 ] * ( ( g.color >> 8 ) & 0xFFu ) / 255.0;
    const double b = c[ 2 ] * ( ( g.color >> 16 ) & 0xFFu ) / 255.0, a = c[ 3 ] * ( ( g.color >> 24 ) & 0xFFu ) / 255.0;
    return ( r + gg + b ) / 3.0 * a + a >= 0.5;
}

struct FilterCtx
{
    const OrPtScene* sc;
    uint32_t         worldMask, rayMask;
};

int MaskedAccept( const void* ctx, uint32_t prim, float u, float v )
{
    const FilterCtx* f = static_cast< const FilterCtx* >( ctx );
    bool             inTlas;
    const uint32_t   m = InstanceMask( f->sc->geoms[ f->sc->triGeom[ prim ] ].flags, f->worldMask, &inTlas );
    if( !inTlas || ( m & f->rayMask ) == 0 ) return 0;
    return AlphaPass( f->sc, prim, u, v ) ? 1 : 0;
}

// Traces rays[i] with cull mask masks[i] (a group per distinct mask).
void Trace( const Ctx& c, std::vector< RgExtRay >& rays, const std::vector< uint8_t >& masks, std::vector< RgExtHit >& hits, uint32_t flags )
{
    hits.resize( rays.size() );
    for( auto& h : hits )
    {
        h.t = 0, h.u = h.v = 0, h.primitiveId = RG_EXT_MISS;
    }
    if( c.sc->triCount == 0 ) return;
    std::vector< std::vector< uint32_t > > groups( 256 );
    for( uint32_t i = 0; i < rays.size(); i++ )
    {
        if( !IsNull( rays[ i ] ) ) groups[ masks[ i ] ].push_back( i );
    }
    for( uint32_t m = 0; m < 256; m++ )
    {
        if( groups[ m ].empty() ) continue;
        std::vector< RgExtRay > rr( groups[ m ].size() );
        std::vector< RgExtHit > hh( groups[ m ].size() );
        for( size_t k = 0; k < rr.size(); k++ ) rr[ k ] = rays[ groups[ m ][ k ] ];
        FilterCtx f{ c.sc, c.p.rayCullMaskWorld, m };
        or_trace_rays_filtered( c.bvh, rr.data(), uint32_t( rr.size() ), hh.data(), flags | c.traceFlags, c.threads, nullptr, nullptr, MaskedAccept, &f );
        for( size_t k = 0; k < rr.size(); k++ ) hits[ groups[ m ][ k ] ] = hh[ k ];
    }
}

RgExtRay MakeRay( D3 o, D3 d, double tmin, double tmax )
{
    RgExtRay r;
    r.origin[ 0 ] = float( o.x ), r.origin[ 1 ] = float( o.y ), r.origin[ 2 ] = float( o.z );
    r.direction[ 0 ] = float( d.x ), r.direction[ 1 ] = float( d.y ), r.direction[ 2 ] = float( d.z );
    r.tmin = float( tmin ), r.tmax = float( tmax );
    return r;
}
RgExtRay NullRay() { return MakeRay( mk( 0, 0, 0 ), mk( 0, 1, 0 ), 1.0, -1.0 ); }
bool     IsNull( const RgExtRay& r ) { return r.tmax < r.tmin; }

// traceShadowRay / traceVisibility (RaygenCommon.h:322-359)
RgExtRay ShadowRay( const Surface& s, D3 lightPos )
{
    const D3     start = s.position + s.toViewer * 0.01;
    D3           l     = lightPos - start;
    const double dist  = length( l );
    return MakeRay( start, l / dist, 0.001, dist - 0.01 );
}

// column-major 4x4 x (v, w) in FP64
D3 MulCM( const float* m, D3 v, double w, double* wOut )
{
    double o[ 4 ];
    for( int r = 0; r < 4; r++ )
    {
        o[ r ] = double( m[ r ] ) * v.x + double( m[ 4 + r ] ) * v.y + double( m[ 8 + r ] ) * v.z + double( m[ 12 + r ] ) * w;
    }
    if( wOut ) *wOut = o[ 3 ];
    return mk( o[ 0 ], o[ 1 ], o[ 2 ] );
}
// getRayDir (ShaderCommonGLSLFunc.h:460-468)
D3 RayDir( const RgExtFrameParams& p, double u, double v )
{
    u = u * 2 - 1;
    v = v * 2 - 1;
    double   tw;
    const D3 t     = MulCM( p.invProjection, mk( u, v, 1.0 ), 1.0, &tw );
    const D3 local = std::fabs( tw ) < 0.001 ? t : t / tw;
    return MulCM( p.invView, normalize( local ), 0.0, nullptr );
}
// The same direction formed in FP32 with the kernel's operation order.
void RayDirF( const RgExtFrameParams& p, float u, float v, float d[ 3 ] )
{
    u = u * 2.0f - 1.0f;
    v = v * 2.0f - 1.0f;
    const float* ip = p.invProjection;
    float        t[ 3 ];
    for( int i = 0; i < 3; i++ )
    {
        t[ i ] = ip[ i ] * u + ip[ 4 + i ] * v + ip[ 8 + i ] * 1.0f + ip[ 12 + i ] * 1.0f;
    }
    const float tw = ip[ 3 ] * u + ip[ 7 ] * v + ip[ 11 ] * 1.0f + ip[ 15 ] * 1.0f;
    float       l[ 3 ];
    for( int i = 0; i < 3; i++ )
    {
        l[ i ] = std::fabs( tw ) < 0.001f ? t[ i ] : t[ i ] / tw;
    }
    const float len = std::sqrt( l[ 0 ] * l[ 0 ] + l[ 1 ] * l[ 1 ] + l[ 2 ] * l[ 2 ] );
    for( int i = 0; i < 3; i++ )
    {
        l[ i ] = l[ i ] / len;
    }
    const float* iv = p.invView;
    for( int i = 0; i < 3; i++ )
    {
        d[ i ] = iv[ i ] * l[ 0 ] + iv[ 4 + i ] * l[ 1 ] + iv[ 8 + i ] * l[ 2 ] + iv[ 12 + i ] * 0.0f;
    }
}

D3 SkyColor( const RgExtFrameParams& p )
{
    const D3     c = mk( p.skyColorDefault[ 0 ], p.skyColorDefault[ 1 ], p.skyColorDefault[ 2 ] );
    const double g = lum( c );
    return mix3( mk( g, g, g ), c, p.skyColorSaturation ) * double( p.skyColorMultiplier );
}

void BlueNoise8( const OrPathTracer* pt, uint32_t seed, uint32_t salt, double out[ 4 ] )
{
    const uint32_t layer = ( ( seed >> 14 ) + salt ) % 128u, oy = ( seed >> 7 ) & 127u, ox = seed & 127u;
    const uint32_t t     = pt->blueNoise[ ( layer * 128u + oy ) * 128u + ox ];
    for( int k = 0; k < 4; k++ )
    {
        out[ k ] = double( float( ( t >> ( 8 * k ) ) & 0xFFu ) / 255.0f );
    }
}

struct Hit
{
    D3       position, normal, albedo, emission;
    double   roughness, metallic;
    uint32_t geom, localPrim;
};

// The texture LOD source of a hit (HitInfo.inl): primary rays sample with
// gradients from the neighbouring pixels' rays (barycentric deltas dbx,
// dby), reflection / refraction rays with a ray cone (RayCone.h:86-130),
// bounces at getBounceLOD (HitInfo.inl:160-175).
enum TexMode { TQ_LOD = 0, TQ_GRAD = 1, TQ_CONE = 2, TQ_BOUNCE = 3 };
struct TexQuery
{
    int    mode = TQ_LOD;
    double a = 0, b = 0;
    D3     dbx{}, dby{}, dir{};
};

double BounceLod( double roughness, double viewDist, double hitDist, double screenWidth, double bias )
{
    const double range = 300.0 * std::pow( ( 1.0 - roughness ) * 0.9 + 0.1, 4.0 );
    const double fx = std::sqrt( clampd( viewDist / range, 0, 1 ) ), fy = std::sqrt( clampd( hitDist / range, 0, 1 ) );
    return std::max( std::log2( 3840.0 / screenWidth ), 0.0 ) + fx * 10.0 + fy * 10.0 + bias;
}

// getHitInfo (HitInfo.inl:204-466): layer-0 colour times its albedo texture
Hit GetHit( const Ctx& c, uint32_t prim, double u, double v, D3 rayOrigin, const TexQuery& q = TexQuery{} )
{
    Hit           h;
    const uint32_t g = c.sc->triGeom[ prim ];
    const OrGeom& gr = c.sc->geoms[ g ];
    const float*  t  = c.sc->tris + size_t( prim ) * 9;
    const float*  n  = c.sc->normals + size_t( prim ) * 9;
    const double  b0 = 1.0 - u - v;
    const D3      p0 = mk( t[ 0 ], t[ 1 ], t[ 2 ] ), p1 = mk( t[ 3 ], t[ 4 ], t[ 5 ] ), p2 = mk( t[ 6 ], t[ 7 ], t[ 8 ] );
    h.position       = p0 * b0 + p1 * u + p2 * v;
    if( ( gr.flags & RG_MESH_PRIMITIVE_FORCE_EXACT_NORMALS ) == 0 )
    {
        h.normal = normalize( mk( n[ 0 ], n[ 1 ], n[ 2 ] ) * b0 + mk( n[ 3 ], n[ 4 ], n[ 5 ] ) * u + mk( n[ 6 ], n[ 7 ], n[ 8 ] ) * v );
    }
    else
    {
        h.normal = safeNormalize( cross( p1 - p0, p2 - p0 ) );
        if( dot( h.normal, h.position - rayOrigin ) > 0.0 ) h.normal = -h.normal;
    }
    const D3 col = UnpackColor( gr.color );
    h.albedo     = mk( saturate( col.x ), saturate( col.y ), saturate( col.z ) );
    double rough = gr.roughness, metal = gr.metallic;
    bool   emissiveMap = false;
    auto   valid       = [ & ]( uint32_t id ) { return id != 0 && id <= c.sc->textureCount; };
    if( valid( gr.texture ) || valid( gr.texOrm ) || valid( gr.texNormal ) || valid( gr.texEmissive ) )
    {
        // texture coordinates and the LOD source of the query (HitInfo.inl,
        // RayCone.h:86-130, getBounceLOD)
        const float* tc  = c.sc->texcoords + size_t( prim ) * 6;
        const double tu  = tc[ 0 ] * b0 + tc[ 2 ] * u + tc[ 4 ] * v, tv = tc[ 1 ] * b0 + tc[ 3 ] * u + tc[ 5 ] * v;
        double       g[ 4 ] = { 0, 0, 0, 0 }, lod = q.a;
        if( q.mode == TQ_GRAD )
        {
            g[ 0 ] = tc[ 0 ] * q.dbx.x + tc[ 2 ] * q.dbx.y + tc[ 4 ] * q.dbx.z, g[ 1 ] = tc[ 1 ] * q.dbx.x + tc[ 3 ] * q.dbx.y + tc[ 5 ] * q.dbx.z;
            g[ 2 ] = tc[ 0 ] * q.dby.x + tc[ 2 ] * q.dby.y + tc[ 4 ] * q.dby.z, g[ 3 ] = tc[ 1 ] * q.dby.x + tc[ 3 ] * q.dby.y + tc[ 5 ] * q.dby.z;
        }
        else if( q.mode == TQ_CONE )
        {
            const double area   = length( cross( p1 - p0, p2 - p0 ) );
            const double pcw    = q.a / std::fabs( dot( q.dir, h.normal ) );
            const double uvArea = std::fabs( ( tc[ 2 ] - tc[ 0 ] ) * ( tc[ 5 ] - tc[ 1 ] ) - ( tc[ 4 ] - tc[ 0 ] ) * ( tc[ 3 ] - tc[ 1 ] ) );
            g[ 0 ] = g[ 3 ] = std::sqrt( uvArea * pcw * pcw / area );
        }
        else if( q.mode == TQ_BOUNCE )
        {
            const D3 cam = mk( c.p.cameraPosition[ 0 ], c.p.cameraPosition[ 1 ], c.p.cameraPosition[ 2 ] );
            lod          = BounceLod( q.a, length( h.position - cam ), length( h.position - rayOrigin ), c.p.renderWidth, q.b );
        }
        auto sample = [ & ]( uint32_t id, double scale, double out[ 4 ] ) {
            if( q.mode == TQ_GRAD || q.mode == TQ_CONE )
                or_texture_sample_grad( c.sc->textures[ id - 1 ], tu, tv, g[ 0 ] * scale, g[ 1 ] * scale, g[ 2 ] * scale, g[ 3 ] * scale, out );
            else
                or_texture_sample_lod( c.sc->textures[ id - 1 ], tu, tv, lod, out );
        };
        double s[ 4 ];
        if( valid( gr.texture ) )
        {
            sample( gr.texture, 1.0, s );
            h.albedo = mk( saturate( col.x * s[ 0 ] ), saturate( col.y * s[ 1 ] ), saturate( col.z * s[ 2 ] ) );
        }
        if( valid( gr.texOrm ) )
        {
            sample( gr.texOrm, 1.0, s );
            rough = s[ 1 ], metal = s[ 2 ];
        }
        if( valid( gr.texEmissive ) )
        {
            sample( gr.texEmissive, 1.0, s );
            h.emission  = mk( s[ 0 ], s[ 1 ], s[ 2 ] );
            emissiveMap = true;
        }
        if( ( q.mode == TQ_GRAD || q.mode == TQ_CONE ) && valid( gr.texNormal ) )
        {
            // getTangent (VertexData.inl:249-269) on the world-space triangle,
            // equal to the reference's model-space tangent times mat3(model)
            // for rigid transforms; normal map with gradients x 5 on primary hits
            sample( gr.texNormal, q.mode == TQ_GRAD ? 5.0 : 1.0, s );
            const double nx = s[ 0 ] * 2.0 - 1.0, ny = s[ 1 ] * 2.0 - 1.0;
            const D3     e1 = p1 - p0, e2 = p2 - p0;
            const double u1x = tc[ 2 ] - tc[ 0 ], u1y = tc[ 3 ] - tc[ 1 ], u2x = tc[ 4 ] - tc[ 0 ], u2y = tc[ 5 ] - tc[ 1 ];
            const double invDet = 1.0 / ( u1x * u2y - u2x * u1y );
            const D3     T  = normalize( ( e1 * u2y - e2 * u1y ) * invDet );
            const D3     Bg = normalize( ( e2 * u1x - e1 * u2x ) * invDet );
            const D3     ns = safeNormalize( mk( n[ 0 ] + n[ 3 ] + n[ 6 ], n[ 1 ] + n[ 4 ] + n[ 7 ], n[ 2 ] + n[ 5 ] + n[ 8 ] ) );
            const double w  = dot( cross( ns, T ), Bg ) > 0.0 ? 1.0 : -1.0;
            const D3     B  = cross( h.normal, T ) * w;
            const D3     nn = safeNormalize( T * nx + B * ny + h.normal );
            h.normal        = safeNormalize( mix3( h.normal, nn, c.p.normalMapStrength ) );
        }
    }
    h.roughness  = std::max( rough, std::max( double( c.p.minRoughness ), MIN_ROUGH ) );
    h.metallic   = metal;
    if( !emissiveMap ) h.emission = h.albedo * double( gr.emissive );
    h.geom       = g;
    h.localPrim  = prim - gr.firstTri;
    return h;
}

Surface HitToSurface( const Ctx& c, const Hit& h, D3 rayDir )
{
    Surface s;
    s.custom    = CustomIndex( c.sc->geoms[ h.geom ].flags, c.p.skyGeometry != 0 );
    s.position  = h.position;
    s.normal    = h.normal;
    s.roughness = h.roughness;
    s.albedo    = h.albedo;
    s.isSky     = false;
    s.specColor = SpecColor( h.albedo, h.metallic );
    s.toViewer  = -rayDir;
    return s;
}

Surface FetchSurface( const Ctx& c, int x, int y )
{
    Surface s;
    s.isSky = c.fb.get( RG_EXT_PLANE_IS_SKY, x, y, 0 ) != 0.0;
    if( s.isSky ) return s;
    int rx, ry;
    or_checkerboard_to_regular( x, y, c.fb.W, &rx, &ry );
    s.albedo    = mk( c.fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 0 ), c.fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 1 ), c.fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 2 ) );
    s.position  = mk( c.fb.get( RG_EXT_PLANE_SURFACE_POSITION, x, y, 0 ), c.fb.get( RG_EXT_PLANE_SURFACE_POSITION, x, y, 1 ), c.fb.get( RG_EXT_PLANE_SURFACE_POSITION, x, y, 2 ) );
    s.specColor = SpecColor( s.albedo, c.fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 0 ) );
    s.roughness = c.fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 1 );
    s.normal    = DecodeNormal( c.fb.get( RG_EXT_PLANE_NORMAL, x, y, 0 ) );
    s.toViewer  = -mk( c.fb.get( RG_EXT_PLANE_VIEW_DIRECTION, x, y, 0 ), c.fb.get( RG_EXT_PLANE_VIEW_DIRECTION, x, y, 1 ), c.fb.get( RG_EXT_PLANE_VIEW_DIRECTION, x, y, 2 ) );
    s.custom    = uint32_t( c.fb.get( RG_EXT_PLANE_SURFACE_POSITION, x, y, 3 ) ); // the custom index as a number
    return s;
}

// ---------------------------------------------------------------- reflection / refraction
// RAYGEN_REFL_REFR_SHADER (RaygenPrimary.inl:352-599) in FP64 over FP32
// rays; portals, animated water normals, screen emission and acid fog are
// not part of the path.  Media.h:28-124 and BRDF.h:73-84 give the media,
// Fresnel and refraction terms; GeomInfoManager.cpp:131-186 maps the
// primitive flags.
const uint32_t MEDIA_VACUUM = 0, MEDIA_WATER = 1, MEDIA_GLASS = 2, MEDIA_ACID = 3;

uint32_t MediaOf( uint32_t f, double rough )
{
    if( f & RG_MESH_PRIMITIVE_WATER ) return MEDIA_WATER;
    if( ( f & RG_MESH_PRIMITIVE_GLASS ) || ( ( f & RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH ) && rough < 0.1 ) ) return MEDIA_GLASS;
    if( f & RG_MESH_PRIMITIVE_ACID ) return MEDIA_ACID;
    return MEDIA_VACUUM;
}
bool Refracts( uint32_t f, double rough )
{
    return ( f & ( RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_ACID | RG_MESH_PRIMITIVE_GLASS ) ) ||
           ( ( f & RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH ) && rough < 0.1 );
}
bool Reflects( uint32_t f, double rough )
{
    return ( f & ( RG_MESH_PRIMITIVE_MIRROR | RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_ACID | RG_MESH_PRIMITIVE_GLASS ) ) ||
           ( ( f & RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH ) && rough < 0.1 );
}
double Ior( const RgExtFrameParams& p, uint32_t m )
{
    return m == MEDIA_WATER || m == MEDIA_ACID ? p.indexOfRefractionWater : ( m == MEDIA_GLASS ? p.indexOfRefractionGlass : 1.0 );
}
D3 Transmittance( const RgExtFrameParams& p, uint32_t m, double dist )
{
    if( m != MEDIA_WATER && m != MEDIA_ACID ) return mk( 1, 1, 1 );
    const float* col = m == MEDIA_WATER ? p.waterColorAndDensity : p.acidColorAndDensity;
    const double k   = m == MEDIA_ACID ? std::max( 1.0, std::sqrt( double( col[ 3 ] ) ) ) : 1.0;
    return mk( std::exp( dist * std::log( double( col[ 0 ] ) ) * k ), std::exp( dist * std::log( double( col[ 1 ] ) ) * k ),
               std::exp( dist * std::log( double( col[ 2 ] ) ) * k ) );
}
// getGlowingMediaFog (Media.h:63-77)
D3 GlowingFog( const RgExtFrameParams& p, uint32_t m, double dist )
{
    if( m != MEDIA_ACID ) return mk( 0, 0, 0 );
    const double fog = clampd( 1.0 - std::exp( -dist * 0.00005 * double( p.acidColorAndDensity[ 3 ] ) ), 0.0, 1.0 );
    return mk( p.acidColorAndDensity[ 0 ] * fog, p.acidColorAndDensity[ 1 ] * fog, p.acidColorAndDensity[ 2 ] * fog );
}

// Snell's law in vector form; false on total internal reflection
bool Refract( double n1, double n2, D3 I, D3 N, D3* T )
{
    const double eta = n1 / n2, cosI = -dot( I, N );
    const double k   = 1.0 - eta * eta * ( 1.0 - cosI * cosI );
    if( k < 0.0 ) return false;
    *T = I * eta + N * ( eta * cosI - std::sqrt( k ) );
    return true;
}

struct RrPix
{
    D3       thr, dir, vpos;
    double   fpl;
    double   cone, coneLast; // ray-cone width now and at the last hit (RayCone.h:25-36)
    D3       emis, fog;      // screen emission and acid fog of the loop
    bool     doSplit;
    uint32_t media;
    bool     active, split;
};

void ReflectRefract( Ctx& c, std::vector< RgExtRay >& rays, std::vector< RgExtHit >& hits )
{
    std::vector< uint8_t > masks( rays.size(), 0 );
    const RgExtFrameParams& p = c.p;
    const Fb&               fb = c.fb;
    const int               W = fb.W, H = fb.H;
    bool                    any = false;
    for( uint32_t g = 0; g < c.sc->geomCount; g++ )
    {
        any = any || ( c.sc->geoms[ g ].flags & ( RG_MESH_PRIMITIVE_MIRROR | RG_MESH_PRIMITIVE_WATER | RG_MESH_PRIMITIVE_GLASS | RG_MESH_PRIMITIVE_ACID |
                                                  RG_MESH_PRIMITIVE_GLASS_IF_SMOOTH | RG_MESH_PRIMITIVE_MIRROR_IF_SMOOTH ) ) != 0;
    }
    if( p.reflectRefractMaxDepth == 0 || !any || c.sc->triCount == 0 ) return;
    std::vector< RrPix > st( size_t( W ) * H );
    const D3             up = mk( p.worldUp[ 0 ], p.worldUp[ 1 ], p.worldUp[ 2 ] );
    for( uint32_t depth = 0; depth < std::min( p.reflectRefractMaxDepth, 16u ); depth++ )
    {
        ParallelFor( H, c.threads, [ & ]( int y ) {
            for( int x = 0; x < W; x++ )
            {
                const size_t i = size_t( y ) * W + x;
                int          cx, cy;
                or_checkerboard_from_regular( x, y, W, &cx, &cy );
                RrPix&   s = st[ i ];
                D3       pos, nrm, alb;
                double   rough, metal;
                uint32_t flags, custom;
                if( depth == 0 )
                {
                    s.active = false;
                    rays[ i ] = NullRay();
                    if( fb.get( RG_EXT_PLANE_IS_SKY, cx, cy, 0 ) != 0.0 ) continue;
                    // the primary surface as the G-buffer holds it
                    alb   = mk( fb.get( RG_EXT_PLANE_ALBEDO, x, y, 0 ), fb.get( RG_EXT_PLANE_ALBEDO, x, y, 1 ), fb.get( RG_EXT_PLANE_ALBEDO, x, y, 2 ) );
                    pos   = mk( fb.get( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 0 ), fb.get( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 1 ),
                                fb.get( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 2 ) );
                    nrm   = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, cx, cy, 0 ) );
                    metal = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, 0 );
                    rough = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, 1 );
                    flags  = c.sc->geoms[ c.sc->triGeom[ hits[ i ].primitiveId ] ].flags;
                    custom = uint32_t( fb.get( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 3 ) );
                    s.thr   = mk( fb.get( RG_EXT_PLANE_THROUGHPUT, cx, cy, 0 ), fb.get( RG_EXT_PLANE_THROUGHPUT, cx, cy, 1 ),
                                  fb.get( RG_EXT_PLANE_THROUGHPUT, cx, cy, 2 ) );
                    s.dir   = RayDir( p, ( double( x ) + 0.5 + p.jitterX ) / W, ( double( y ) + 0.5 + p.jitterY ) / H );
                    s.fpl   = fb.get( RG_EXT_PLANE_DEPTH_WORLD, cx, cy, 0 );
                    s.vpos  = pos;
                    s.cone  = double( p.cameraRayConeSpreadAngle ) * s.fpl; // width 0 propagated over the primary hit
                    s.emis  = mk( fb.get( RG_EXT_PLANE_SCREEN_EMISSION, x, y, 0 ), fb.get( RG_EXT_PLANE_SCREEN_EMISSION, x, y, 1 ),
                                  fb.get( RG_EXT_PLANE_SCREEN_EMISSION, x, y, 2 ) );
                    s.fog   = mk( fb.get( RG_EXT_PLANE_ACID_FOG, x, y, 0 ), fb.get( RG_EXT_PLANE_ACID_FOG, x, y, 1 ), fb.get( RG_EXT_PLANE_ACID_FOG, x, y, 2 ) );
                    s.media = p.cameraMediaType & 3u;
                    s.split = false;
                }
                else
                {
                    if( !s.active )
                    {
                        rays[ i ] = NullRay();
                        continue;
                    }
                    const RgExtRay& r = rays[ i ];
                    TexQuery        cq;
                    cq.mode = TQ_CONE, cq.a = s.coneLast, cq.dir = s.dir;
                    const Hit       h = GetHit( c, hits[ i ].primitiveId, hits[ i ].u, hits[ i ].v, mk( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] ), cq );
                    pos = h.position, nrm = h.normal, alb = h.albedo, rough = h.roughness, metal = h.metallic;
                    flags  = c.sc->geoms[ h.geom ].flags;
                    custom = CustomIndex( flags, p.skyGeometry != 0 );
                }
                // getNewRayMedia (RaygenPrimary.inl:141-150)
                uint32_t   newMedia = ( depth == 0 && p.cameraMediaType != MEDIA_VACUUM ) ? MEDIA_VACUUM : MediaOf( flags, rough );
                const bool toRefract = Refracts( flags, rough ), toReflect = Reflects( flags, rough );
                if( !toReflect && !toRefract )
                {
                    s.active  = false;
                    rays[ i ] = NullRay();
                    continue;
                }
                const double n1 = Ior( p, s.media ), n2 = Ior( p, newMedia );
                const bool   water = newMedia == MEDIA_WATER || s.media == MEDIA_WATER || newMedia == MEDIA_ACID || s.media == MEDIA_ACID;
                if( dot( nrm, -s.dir ) < 0.0 ) nrm = -nrm;
                if( water )
                {
                    // getWaterNormal (RaygenPrimary.inl:152-199) over the
                    // default water texel (127, 127, 255): each octave reads
                    // (2 * 127/255 - 1, same, 1) = (-1/255, -1/255, 1)
                    D3 b0, b1;
                    revisedONB( nrm, &b0, &b1 );
                    const double k  = double( p.waterWaveStrength ) * ( 0.25 + 0.2 + 0.1 );
                    const D3     tn = normalize( mk( k * ( 2.0 * 127.0 / 255.0 - 1.0 ), k * ( 2.0 * 127.0 / 255.0 - 1.0 ), 1.0 + k ) );
                    nrm             = b0 * tn.x + b1 * tn.y + nrm * tn.z;
                }
                else
                {
                    nrm = normalize( nrm );
                }
                D3     origin = pos, t;
                bool   doSplit = !s.split, doRefr = false;
                double F       = 1.0;
                if( toRefract && Refract( n1, n2, s.dir, nrm, &t ) )
                {
                    doRefr         = cx >= W / 2; // odd checkerboard pixels refract
                    const double r0 = square( ( n1 - n2 ) / ( n1 + n2 ) );
                    F               = r0 + ( 1.0 - r0 ) * std::pow( 1.0 - std::fabs( dot( nrm, -s.dir ) ), 5.0 );
                }
                else
                {
                    doSplit = false; // total internal reflection
                }
                if( doRefr )
                {
                    s.dir = t;
                    s.thr = s.thr * ( 1.0 - F );
                    if( flags & RG_MESH_PRIMITIVE_THIN_MEDIA )
                    {
                        const double len = p.thinMediaWidth / std::max( 0.001, -dot( nrm, s.dir ) );
                        origin           = origin + s.dir * len;
                        s.thr            = s.thr * Transmittance( p, newMedia, len );
                        s.fpl += len;
                        if( Refract( n2, n1, s.dir, nrm, &t ) ) s.dir = t;
                        newMedia = s.media;
                    }
                    s.media = newMedia;
                }
                else
                {
                    const D3 in = normalize( s.dir );
                    s.dir       = in - nrm * ( 2.0 * dot( nrm, in ) );
                    if( !water )
                    {
                        const D3     f0 = SpecColor( alb, metal );
                        const double k  = std::pow( 1.0 - std::max( dot( nrm, s.dir ), 0.0 ), 5.0 );
                        s.thr           = s.thr * ( f0 + ( mk( 1, 1, 1 ) - f0 ) * k );
                    }
                    else
                    {
                        s.thr  = s.thr * F;
                        origin = origin + s.dir * 0.005; // the reference's SHIPPING_HACK offset
                    }
                }
                s.doSplit = doSplit;
                if( doSplit )
                {
                    s.thr   = s.thr * 2.0;
                    s.split = true;
                }
                s.active   = true;
                rays[ i ]  = MakeRay( origin, s.dir, 0.001, p.rayLength );
                masks[ i ] = ReflRefrMask( p, custom, flags, doRefr );
            }
        } );
        Trace( c, rays, masks, hits, RG_EXT_TRACE_CLOSEST );
        ParallelFor( H, c.threads, [ & ]( int y ) {
            for( int x = 0; x < W; x++ )
            {
                const size_t i = size_t( y ) * W + x;
                RrPix&       s = st[ i ];
                if( !s.active ) continue;
                int cx, cy;
                or_checkerboard_from_regular( x, y, W, &cx, &cy );
                const double thrW = s.split ? 1.0 : -1.0;
                if( IsSkyHit( c, hits[ i ] ) )
                {
                    const double a = std::fabs( dot( s.dir, up ) );
                    s.thr          = s.thr * Transmittance( p, s.media, 1.0 / ( a * a * a ) );
                    // storeSky (RaygenPrimary.inl:90-139)
                    const double cu = ( double( cx ) + 0.5 + p.jitterX ) / W, cv = ( double( cy ) + 0.5 + p.jitterY ) / H;
                    const D3     rd = RayDir( p, cu, cv );
                    const D3     vc = MulCM( p.view, rd, 0.0, nullptr ), vp = MulCM( p.viewPrev, rd, 0.0, nullptr );
                    const double ccx = p.projection[ 0 ] * vc.x + p.projection[ 4 ] * vc.y + p.projection[ 8 ] * vc.z;
                    const double ccy = p.projection[ 1 ] * vc.x + p.projection[ 5 ] * vc.y + p.projection[ 9 ] * vc.z;
                    const double cpx = p.projectionPrev[ 0 ] * vp.x + p.projectionPrev[ 4 ] * vp.y + p.projectionPrev[ 8 ] * vp.z;
                    const double cpy = p.projectionPrev[ 1 ] * vp.x + p.projectionPrev[ 5 ] * vp.y + p.projectionPrev[ 9 ] * vp.z;
                    const D3     sky = SkyColor( p );
                    fb.set4( RG_EXT_PLANE_IS_SKY, cx, cy, 1 );
                    fb.set4( RG_EXT_PLANE_ALBEDO, x, y, sky.x, sky.y, sky.z, 0 );
                    fb.set4( RG_EXT_PLANE_NORMAL, cx, cy, EncodeNormal( mk( 0, 0, 0 ) ) );
                    fb.set4( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, 0, 0 );
                    fb.set4( RG_EXT_PLANE_DEPTH_WORLD, cx, cy, MAX_RAY_LEN * 2.0 );
                    fb.set4( RG_EXT_PLANE_MOTION, cx, cy, ( cpx * 0.5 + 0.5 ) - ( ccx * 0.5 + 0.5 ), ( cpy * 0.5 + 0.5 ) - ( ccy * 0.5 + 0.5 ), 0, 0 );
                    fb.set4( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 1e7, 1e7, 1e7, 1e7 );
                    fb.set4( RG_EXT_PLANE_VISIBILITY_BUFFER, cx, cy, 4294967295.0, 4294967295.0, 4294967295.0, 4294967295.0 );
                    fb.set4( RG_EXT_PLANE_VIEW_DIRECTION, cx, cy, s.dir.x, s.dir.y, s.dir.z, 0 );
                    fb.set4( RG_EXT_PLANE_THROUGHPUT, cx, cy, s.thr.x, s.thr.y, s.thr.z, thrW );
                    fb.set4( RG_EXT_PLANE_SCREEN_EMISSION, x, y, 0, 0, 0, 0 );
                    fb.set4( RG_EXT_PLANE_ACID_FOG, x, y, 0, 0, 0, 0 );
                    s.active = false;
                    continue;
                }
                const RgExtRay& r      = rays[ i ];
                const D3        origin = mk( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] );
                TexQuery        cq;
                cq.mode = TQ_CONE, cq.a = s.cone, cq.dir = s.dir;
                const Hit       h      = GetHit( c, hits[ i ].primitiveId, hits[ i ].u, hits[ i ].v, origin, cq );
                const double    len    = length( h.position - origin );
                // propagateRayCone: the spread doubles at each propagation
                s.coneLast = s.cone;
                s.cone += double( p.cameraRayConeSpreadAngle ) * std::exp2( double( depth + 1 ) ) * len;
                // virtual position for motion: along the camera ray by the path length
                const D3 cam = RayDir( p, ( double( x ) + 0.5 + p.jitterX ) / W, ( double( y ) + 0.5 + p.jitterY ) / H );
                s.vpos       = s.vpos + cam * len;
                double   wc, wp, cwc, cwp;
                const D3 vCur  = MulCM( p.view, s.vpos, 1.0, &wc );
                const D3 vPrev = MulCM( p.viewPrev, s.vpos, 1.0, &wp );
                const D3 cCur  = MulCM( p.projection, vCur, wc, &cwc );
                const D3 cPrev = MulCM( p.projectionPrev, vPrev, wp, &cwp );
                s.thr          = s.thr * Transmittance( p, s.media, len );
                s.fpl += len;
                s.emis         = s.emis + h.emission * s.thr;
                s.fog          = s.fog + GlowingFog( p, s.media, len ) * ( s.doSplit ? 2.0 : 1.0 );
                const D3 semi  = p.cameraMediaType != MEDIA_ACID ? s.emis + s.fog * 0.05 : s.emis;
                fb.set4( RG_EXT_PLANE_SCREEN_EMISSION, x, y, semi.x, semi.y, semi.z, 0 );
                fb.set4( RG_EXT_PLANE_ACID_FOG, x, y, s.fog.x, s.fog.y, s.fog.z, 0 );
                fb.set4( RG_EXT_PLANE_IS_SKY, cx, cy, 0 );
                fb.set4( RG_EXT_PLANE_ALBEDO, x, y, h.albedo.x, h.albedo.y, h.albedo.z, 0 );
                fb.set4( RG_EXT_PLANE_NORMAL, cx, cy, EncodeNormal( h.normal ) );
                fb.set4( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, h.metallic, h.roughness );
                fb.set4( RG_EXT_PLANE_DEPTH_WORLD, cx, cy, s.fpl );
                fb.set4( RG_EXT_PLANE_MOTION, cx, cy, ( cPrev.x / cwp ) * 0.5 - ( cCur.x / cwc ) * 0.5, ( cPrev.y / cwp ) * 0.5 - ( cCur.y / cwc ) * 0.5,
                         length( vPrev ) - length( vCur ), 0 );
                fb.set4( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, h.position.x, h.position.y, h.position.z,
                         CustomIndex( c.sc->geoms[ h.geom ].flags, p.skyGeometry != 0 ) );
                fb.set4( RG_EXT_PLANE_VISIBILITY_BUFFER, cx, cy, 1, double( ( h.localPrim << 12 ) | ( h.geom & 0xFFFu ) ), hits[ i ].u, hits[ i ].v );
                fb.set4( RG_EXT_PLANE_VIEW_DIRECTION, cx, cy, s.dir.x, s.dir.y, s.dir.z, 0 );
                fb.set4( RG_EXT_PLANE_THROUGHPUT, cx, cy, s.thr.x, s.thr.y, s.thr.z, thrW );
            }
        } );
    }
}

// ---------------------------------------------------------------- light grid (LightGrid.h:31-71)

struct Grid
{
    D3     minExt, maxExt, delta;
    double cellRadius;
};
Grid MakeGrid( const RgExtFrameParams& p )
{
    Grid         g;
    const double d = p.cellWorldSize;
    g.delta        = mk( d, d, d );
    const D3 whole = mk( d * GX, d * GY, d * GZ );
    const D3 ctr   = mk( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] ) + g.delta * 0.5;
    g.minExt       = ctr - whole * 0.5;
    g.maxExt       = ctr + whole * 0.5;
    g.cellRadius   = length( g.delta ) * 0.5;
    return g;
}
bool InsideGrid( const Grid& g, D3 p )
{
    return p.x > g.minExt.x && p.y > g.minExt.y && p.z > g.minExt.z && p.x < g.maxExt.x && p.y < g.maxExt.y && p.z < g.maxExt.z;
}
int CellArrayIndex( const Grid& g, D3 wp )
{
    const D3 r = wp - g.minExt;
    int      cx = std::min( std::max( int( r.x / g.delta.x ), 0 ), GX );
    int      cy = std::min( std::max( int( r.y / g.delta.y ), 0 ), GY );
    int      cz = std::min( std::max( int( r.z / g.delta.z ), 0 ), GZ );
    return GCELL * ( cx + cy * GX + cz * GX * GY );
}

void LightGrid( Ctx& c )
{
    if( c.lightCount == 0 )
    {
        return;
    }
    OrPathTracer* pt      = c.pt;
    const auto&   gPrev   = pt->grid[ pt->gridParity ];
    auto&         gOut    = pt->grid[ pt->gridParity ^ 1 ];
    const Grid    g       = MakeGrid( c.p );
    const double  nLights = double( c.lightCount );
    ParallelFor( int( GRID_ENTRIES / 1024 ), c.threads, [ & ]( int chunk ) {
        for( int ai = chunk * 1024; ai < chunk * 1024 + 1024; ai++ )
        {
            const int cc = ai / GCELL;
            const int ci[ 3 ] = { cc % GX, ( cc % ( GX * GY ) ) / GX, cc / ( GX * GY ) };
            const D3  center  = g.minExt + g.delta * mk( ci[ 0 ] + 0.5, ci[ 1 ] + 0.5, ci[ 2 ] + 0.5 );
            const uint32_t seed = seedOf( ai, 0, c.p.frameId );
            uint32_t       salt = SALT_LIGHT_GRID_BASE;
            Res            reg;
            for( int i = 0; i < 8; i++ )
            {
                const double   rnd = rnd16( seed, salt++ );
                const uint32_t xi  = 1u + std::min( uint32_t( float( rnd ) * float( c.lightCount ) ), c.lightCount - 1u );
                const double   tp  = lightWeight( LightAt( c, xi ), center, g.cellRadius );
                update( reg, xi, tp, nLights, rnd16( seed, salt++ ) );
            }
            Res tmp;
            tmp.sel  = uint32_t( gPrev[ size_t( ai ) * 3 + 0 ] );
            tmp.tp   = gPrev[ size_t( ai ) * 3 + 1 ];
            tmp.wsum = gPrev[ size_t( ai ) * 3 + 2 ];
            tmp.M    = 1;
            normalizeRes( tmp, reg.M * 20 );
            double tpTmp = 0.0;
            if( tmp.sel != LIGHT_NONE )
            {
                const uint32_t cur = tmp.sel < c.prevToCur.size() ? c.prevToCur[ tmp.sel ] : 0xFFFFFFFFu;
                if( cur != 0xFFFFFFFFu && cur != LIGHT_NONE )
                {
                    tpTmp   = lightWeight( LightAt( c, cur ), center, g.cellRadius );
                    tmp.sel = cur;
                }
            }
            const double rndRis = rnd16( seed, salt++ );
            Res          comb   = reg;
            combineNewSurf( comb, tmp, tpTmp, rndRis );
            normalizeRes( comb, 1 );
            gOut[ size_t( ai ) * 3 + 0 ] = comb.sel;
            gOut[ size_t( ai ) * 3 + 1 ] = comb.tp;
            gOut[ size_t( ai ) * 3 + 2 ] = comb.wsum;
        }
    } );
    pt->gridParity ^= 1;
}

void LightPointRnd( uint32_t seed, double* u, double* v )
{
    rnd16_2( seed, SALT_LIGHT_POINT, u, v );
    *u = double( float( *u ) * 0.99f );
    *v = double( float( *v ) * 0.99f );
}

bool DirectValid( const Ctx& c, int bounce )
{
    return ( uint32_t( bounce ) < c.p.maxBounceShadowsLights || c.p.maxBounceShadowsLights == 0 ) && ( c.lightCount + ( c.hasDir ? 1 : 0 ) > 0 );
}

// calcInitialReservoir (RaygenCommon.h:396-486) up to the directional visibility test
void InitialCandidates( const Ctx& c, const Surface& s, uint32_t seed, uint32_t& salt, double pu, double pv, bool loose, Res& reg, Res& dir )
{
    const Grid g  = MakeGrid( c.p );
    const auto& G = c.pt->grid[ c.pt->gridParity ];
    reg           = Res{};
    if( InsideGrid( g, s.position ) )
    {
        double r4[ 4 ];
        rnd8_4( seed, salt++, r4 );
        D3 jp = s.position + ( mk( r4[ 0 ], r4[ 1 ], r4[ 2 ] ) * 2.0 - mk( 1, 1, 1 ) ) * g.cellRadius;
        jp    = mk( clampd( jp.x, g.minExt.x, g.maxExt.x ), clampd( jp.y, g.minExt.y, g.maxExt.y ), clampd( jp.z, g.minExt.z, g.maxExt.z ) );
        const int base = CellArrayIndex( g, jp );
        for( uint32_t i = 0; i < c.p.restirInitialCandidates; i++ )
        {
            const double rnd = rnd16( seed, salt++ );
            const int    idx = base + std::min( std::max( int( float( rnd ) * float( GCELL ) ), 0 ), GCELL - 1 );
            Res          r;
            if( idx < int( GRID_ENTRIES ) )
            {
                r.sel  = uint32_t( G[ size_t( idx ) * 3 ] );
                r.tp   = G[ size_t( idx ) * 3 + 1 ];
                r.wsum = G[ size_t( idx ) * 3 + 2 ];
            }
            r.M = 1;
            if( !valid( r ) ) continue;
            const double  oneOverSource = r.wsum * safeRcp( r.tp );
            const LSample ls            = sampleLight( LightAt( c, r.sel ), s.position, pu, pv, c.p.polyLightSpotlightFactor );
            update( reg, r.sel, targetPdf( s, ls, loose ), oneOverSource, rnd16( seed, salt++ ) );
        }
    }
    else
    {
        for( uint32_t i = 0; i < c.p.restirInitialCandidates; i++ )
        {
            const double   rnd = rnd16( seed, salt++ );
            const uint32_t xi  = 1u + std::min( uint32_t( float( rnd ) * float( c.lightCount ) ), c.lightCount - 1u );
            const LSample  ls  = sampleLight( LightAt( c, xi ), s.position, pu, pv, c.p.polyLightSpotlightFactor );
            update( reg, xi, targetPdf( s, ls, loose ), double( float( c.lightCount ) ), rnd16( seed, salt++ ) );
        }
    }
    normalizeRes( reg, 1 );
    dir = Res{};
    if( c.hasDir )
    {
        const LSample ls = sampleLight( c.lights[ 0 ], s.position, pu, pv, c.p.polyLightSpotlightFactor );
        update( dir, 0, targetPdf( s, ls, loose ), 1.0, rnd16( seed, salt++ ) );
    }
    normalizeRes( dir, 1 );
}

Res CombineInitial( Res reg, const Res& dir, uint32_t seed, uint32_t& salt )
{
    combine( reg, dir, rnd16( seed, salt++ ) );
    normalizeRes( reg, 1 );
    return reg;
}

bool ReuseOk( int lo, int hi, int H, int qx, int qy, double dCur, double dOther, D3 nCur, D3 nOther, double dThr, double nThr )
{
    return qy >= 0 && qy < H && qx >= lo && qx < hi && ( std::fabs( dCur - dOther ) / std::fabs( dCur ) < dThr ) && ( dot( nCur, nOther ) > nThr );
}

void RenderArea( int cx, int W, int* lo, int* hi )
{
    const int sep = W / 2, odd = cx >= sep ? 1 : 0;
    *lo           = odd * sep;
    *hi           = ( odd + 1 ) * sep;
}

// Direct light at a bounce hit: RIS without visibility (RaygenCommon.h:701-741)
void IndirectDirect( const Ctx& c, const Surface& hs, uint32_t seed, int bounce, D3* diffOut, bool* isDir, bool* needShadow, D3* lightPos,
                     uint32_t* lightIndex = nullptr )
{
    *diffOut = mk( 0, 0, 0 ), *needShadow = false, *isDir = false;
    if( !DirectValid( c, bounce ) ) return;
    double pu, pv;
    LightPointRnd( seed, &pu, &pv );
    uint32_t salt = SALT_LIGHT_CHOOSE_INDIRECT;
    Res      reg, dir;
    InitialCandidates( c, hs, seed, salt, pu, pv, true, reg, dir );
    const Res r = CombineInitial( reg, dir, seed, salt );
    if( !valid( r ) ) return;
    const LSample ls = sampleLight( LightAt( c, r.sel ), hs.position, pu, pv, c.p.polyLightSpotlightFactor );
    D3            d, sp;
    shade( hs, ls, selWeight( r ), true, &d, &sp );
    if( lum( d + sp ) <= 0.0 ) return;
    *diffOut    = d;
    *isDir      = r.sel == 0u;
    *needShadow = uint32_t( bounce ) < c.p.maxBounceShadowsLights;
    *lightPos   = ls.position;
    if( lightIndex ) *lightIndex = r.sel;
}

double Antilag( double g, double n ) { return clampd( n > 0.0001 ? clampd( std::fabs( g ) / n, 0.0, 1.0 ) : 0.0, 0.0, 1.0 ); }

} // namespace

extern "C" OrPathTracer* or_pt_create( int width, int height )
{
    auto* pt = new OrPathTracer();
    pt->W    = width;
    pt->H    = height;
    pt->blueNoise.resize( 128u * 128u * 128u );
    for( uint32_t i = 0; i < pt->blueNoise.size(); i++ )
    {
        // the kernels' seeded stand-in for the blue-noise blob: SplitMix64 of the texel index
        uint64_t z = 0x9E3779B97F4A7C15ull * ( uint64_t( i ) + 1ull ) + 0xC0FFEEull;
        z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
        z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
        pt->blueNoise[ i ] = uint32_t( z ^ ( z >> 31 ) );
    }
    for( auto& g : pt->grid )
    {
        g.assign( size_t( GRID_ENTRIES ) * 3, 0.0 );
        for( size_t e = 0; e < GRID_ENTRIES; e++ )
        {
            g[ e * 3 ] = LIGHT_NONE; // empty cells
        }
    }
    return pt;
}

extern "C" void or_pt_destroy( OrPathTracer* pt ) { delete pt; }

extern "C" void or_pt_frame( OrPathTracer* pt, const OrPtScene* sc, const OrBvh* bvh, const RgExtFrameParams* pIn, const OrFramebuffers* fbp, int threads )
{
    Ctx c{ sc, bvh, *pIn, Fb( fbp ), pt, {}, {}, {}, 0, false, threads, 0 };
    const int W = c.fb.W, H = c.fb.H, GW = c.fb.GW, GH = c.fb.GH;
    const int n = W * H;
    c.traceFlags = c.p.rayCullBackFaces ? RG_EXT_TRACE_CULL_FRONT_FACE : 0u;

    // ---- lights and the previous-frame matches (LightManager.cpp:322-336, FillMatchPrev)
    std::vector< uint64_t > ids;
    c.lights.push_back( Light{} );
    ids.push_back( 0 );
    if( sc->dirLight && sc->dirLight->intensity > 0.00001f && ( sc->dirLight->color & 0x00FFFFFFu ) != 0 )
    {
        c.lights[ 0 ] = EncodeDirectional( *sc->dirLight );
        ids[ 0 ]      = sc->dirLight->uniqueID;
        c.hasDir      = true;
    }
    for( uint32_t i = 0; i < sc->sphereCount; i++ )
    {
        const auto& s = sc->spheres[ i ];
        if( s.intensity <= 0.00001f || ( s.color & 0x00FFFFFFu ) == 0 ) continue;
        c.lights.push_back( EncodeSphere( s ) );
        ids.push_back( s.uniqueID );
    }
    for( uint32_t i = 0; i < sc->spotCount; i++ )
    {
        const auto& s = sc->spots[ i ];
        if( s.intensity <= 0.00001f || ( s.color & 0x00FFFFFFu ) == 0 || s.radius < 0.0f || s.angleOuter <= 0.0f ) continue;
        c.lights.push_back( EncodeSpot( s ) );
        ids.push_back( s.uniqueID );
    }
    for( uint32_t i = 0; i < sc->polyCount; i++ )
    {
        const auto& s = sc->polys[ i ];
        if( s.intensity <= 0.00001f || ( s.color & 0x00FFFFFFu ) == 0 ) continue;
        const Light l = EncodeTriangle( s );
        if( !( l.d0[ 3 ] * l.d0[ 3 ] + l.d1[ 3 ] * l.d1[ 3 ] + l.d2[ 3 ] * l.d2[ 3 ] > 0.0 ) ) continue;
        c.lights.push_back( l );
        ids.push_back( s.uniqueID );
    }
    c.lightCount = uint32_t( c.lights.size() ) - 1;
    c.prevToCur.assign( std::max< size_t >( 1, pt->prevLightIds.size() ), 0xFFFFFFFFu );
    {
        std::unordered_map< uint64_t, uint32_t > prevIdx;
        for( uint32_t i = 0; i < pt->prevLightIds.size(); i++ )
        {
            if( i == 0 && pt->prevLightIds[ 0 ] == 0 ) continue; // no directional light last frame
            prevIdx[ pt->prevLightIds[ i ] ] = i;
        }
        for( uint32_t i = 0; i < ids.size(); i++ )
        {
            if( i == 0 && !c.hasDir ) continue;
            auto it = prevIdx.find( ids[ i ] );
            if( it != prevIdx.end() ) c.prevToCur[ it->second ] = i;
        }
    }
    pt->prevLightIds = ids;
    if( !c.hasDir ) pt->prevLightIds[ 0 ] = 0;
    c.p.lightCount             = c.lightCount;
    c.p.directionalLightExists = c.hasDir ? 1u : 0u;
    // geometry ids: (object id, primitive index) per geometry
    {
        std::unordered_map< uint64_t, uint32_t > cur;
        for( uint32_t i = 0; i < sc->geomCount; i++ ) cur.emplace( sc->geomIds[ i ], i );
        c.geomPrevToCur.assign( std::max< size_t >( 1, pt->prevGeomIds.size() ), 0xFFFFFFFFu );
        for( uint32_t i = 0; i < pt->prevGeomIds.size(); i++ )
        {
            auto it = cur.find( pt->prevGeomIds[ i ] );
            if( it != cur.end() ) c.geomPrevToCur[ i ] = it->second;
        }
        pt->prevGeomIds.assign( sc->geomIds, sc->geomIds + sc->geomCount );
    }
    const RgExtFrameParams& p = c.p;
    const Fb&               fb = c.fb;
    const D3                camPos = mk( p.cameraPosition[ 0 ], p.cameraPosition[ 1 ], p.cameraPosition[ 2 ] );

    std::vector< RgExtRay > rays( n ), rays2( n ), rays3( n );
    std::vector< RgExtHit > hits, hits2, hits3;
    std::vector< uint8_t >  masks( n, PrimaryMask( p ) ), masks2( n, 0 ), masks3( n, 0 );

    // ---- primary rays (FP32, the kernel's operations) and G-buffer
    for( int y = 0; y < H; y++ )
    {
        for( int x = 0; x < W; x++ )
        {
            const float u = ( float( x ) + 0.5f + p.jitterX ) / float( W );
            const float v = ( float( y ) + 0.5f + p.jitterY ) / float( H );
            float       d[ 3 ];
            RayDirF( p, u, v, d );
            RgExtRay& r = rays[ size_t( y ) * W + x ];
            r.origin[ 0 ] = p.cameraPosition[ 0 ], r.origin[ 1 ] = p.cameraPosition[ 1 ], r.origin[ 2 ] = p.cameraPosition[ 2 ];
            r.direction[ 0 ] = d[ 0 ], r.direction[ 1 ] = d[ 1 ], r.direction[ 2 ] = d[ 2 ];
            r.tmin = p.primaryRayMinDist, r.tmax = p.rayLength;
        }
    }
    Trace( c, rays, masks, hits, RG_EXT_TRACE_CLOSEST );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            int cx, cy;
            or_checkerboard_from_regular( x, y, W, &cx, &cy );
            const RgExtHit& hit = hits[ size_t( y ) * W + x ];
            const double    u   = ( double( x ) + 0.5 + p.jitterX ) / W;
            const double    v   = ( double( y ) + 0.5 + p.jitterY ) / H;
            const D3        dir = RayDir( p, u, v );
            if( IsSkyHit( c, hit ) )
            {
                const double cu = ( double( cx ) + 0.5 + p.jitterX ) / W, cv = ( double( cy ) + 0.5 + p.jitterY ) / H;
                const D3     rd = RayDir( p, cu, cv );
                const D3     vc = MulCM( p.view, rd, 0.0, nullptr ), vp = MulCM( p.viewPrev, rd, 0.0, nullptr );
                const double ccx = p.projection[ 0 ] * vc.x + p.projection[ 4 ] * vc.y + p.projection[ 8 ] * vc.z;
                const double ccy = p.projection[ 1 ] * vc.x + p.projection[ 5 ] * vc.y + p.projection[ 9 ] * vc.z;
                const double cpx = p.projectionPrev[ 0 ] * vp.x + p.projectionPrev[ 4 ] * vp.y + p.projectionPrev[ 8 ] * vp.z;
                const double cpy = p.projectionPrev[ 1 ] * vp.x + p.projectionPrev[ 5 ] * vp.y + p.projectionPrev[ 9 ] * vp.z;
                const D3     sky = SkyColor( p );
                fb.set4( RG_EXT_PLANE_IS_SKY, cx, cy, 1 );
                fb.set4( RG_EXT_PLANE_ALBEDO, x, y, sky.x, sky.y, sky.z, 0 );
                fb.set4( RG_EXT_PLANE_NORMAL, cx, cy, EncodeNormal( mk( 0, 0, 0 ) ) );
                fb.set4( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, 0, 0 );
                fb.set4( RG_EXT_PLANE_DEPTH_WORLD, cx, cy, MAX_RAY_LEN * 2.0 );
                fb.set4( RG_EXT_PLANE_MOTION, cx, cy, ( cpx * 0.5 + 0.5 ) - ( ccx * 0.5 + 0.5 ), ( cpy * 0.5 + 0.5 ) - ( ccy * 0.5 + 0.5 ), 0, 0 );
                fb.set4( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, 1e7, 1e7, 1e7, 1e7 );
                fb.set4( RG_EXT_PLANE_VISIBILITY_BUFFER, cx, cy, 4294967295.0, 4294967295.0, 4294967295.0, 4294967295.0 );
                fb.set4( RG_EXT_PLANE_VIEW_DIRECTION, cx, cy, dir.x, dir.y, dir.z, 0 );
                fb.set4( RG_EXT_PLANE_DEPTH_GRAD, cx, cy, 0 );
                fb.set4( RG_EXT_PLANE_THROUGHPUT, cx, cy, 1, 1, 1, 0 );
                fb.set4( RG_EXT_PLANE_SCREEN_EMISSION, x, y, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_ACID_FOG, x, y, 0, 0, 0, 0 );
                continue;
            }
            const float*  t  = sc->tris + size_t( hit.primitiveId ) * 9;
            const D3      p0 = mk( t[ 0 ], t[ 1 ], t[ 2 ] ), p1 = mk( t[ 3 ], t[ 4 ], t[ 5 ] ), p2 = mk( t[ 6 ], t[ 7 ], t[ 8 ] );
            // ray differentials on the same triangle (HitInfo.inl:181-268)
            auto baryOf = [ & ]( D3 d ) {
                const D3     e1 = p1 - p0, e2 = p2 - p0, pv = cross( d, e2 );
                const double inv = 1.0 / dot( e1, pv );
                const D3     tv  = camPos - p0, qv = cross( tv, e1 );
                const double bu  = dot( tv, pv ) * inv, bv = dot( d, qv ) * inv;
                return mk( 1.0 - bu - bv, bu, bv );
            };
            const D3 bAX = baryOf( RayDir( p, u + 1.0 / W, v ) ), bAY = baryOf( RayDir( p, u, v + 1.0 / H ) );
            const D3 pAX = p0 * bAX.x + p1 * bAX.y + p2 * bAX.z, pAY = p0 * bAY.x + p1 * bAY.y + p2 * bAY.z;
            TexQuery tq;
            tq.mode       = TQ_GRAD;
            const D3 b    = mk( 1.0 - double( hit.u ) - double( hit.v ), hit.u, hit.v );
            tq.dbx        = bAX - b;
            tq.dby        = bAY - b;
            const Hit     h  = GetHit( c, hit.primitiveId, hit.u, hit.v, camPos, tq );
            const OrGeom& gr = sc->geoms[ h.geom ];
            D3       prevPos = h.position;
            if( !gr.isStatic )
            {
                const float* m = gr.prevFromCur;
                prevPos        = mk( m[ 0 ] * h.position.x + m[ 1 ] * h.position.y + m[ 2 ] * h.position.z + m[ 3 ],
                                     m[ 4 ] * h.position.x + m[ 5 ] * h.position.y + m[ 6 ] * h.position.z + m[ 7 ],
                                     m[ 8 ] * h.position.x + m[ 9 ] * h.position.y + m[ 10 ] * h.position.z + m[ 11 ] );
            }
            double   wc, wp, wx, wy, cwc, cwp;
            const D3 vCur  = MulCM( p.view, h.position, 1.0, &wc );
            const D3 vPrev = MulCM( p.viewPrev, prevPos, 1.0, &wp );
            const D3 vAX   = MulCM( p.view, pAX, 1.0, &wx );
            const D3 vAY   = MulCM( p.view, pAY, 1.0, &wy );
            const D3 cCur  = MulCM( p.projection, vCur, wc, &cwc );
            const D3 cPrev = MulCM( p.projectionPrev, vPrev, wp, &cwp );
            // dot(projection[2], v): the matrix's third column (RaygenPrimary.inl, depth gradient)
            const float* col2 = p.projection + 8;
            const double dAX  = col2[ 0 ] * vAX.x + col2[ 1 ] * vAX.y + col2[ 2 ] * vAX.z + col2[ 3 ] * wx;
            const double dAY  = col2[ 0 ] * vAY.x + col2[ 1 ] * vAY.y + col2[ 2 ] * vAY.z + col2[ 3 ] * wy;
            const double gx = dAX - cCur.z, gy = dAY - cCur.z;
            const double sxc = ( cCur.x / cwc ) * 0.5 + 0.5, syc = ( cCur.y / cwc ) * 0.5 + 0.5;
            const double sxp = ( cPrev.x / cwp ) * 0.5 + 0.5, syp = ( cPrev.y / cwp ) * 0.5 + 0.5;
            const double depth = length( vCur );
            fb.set4( RG_EXT_PLANE_IS_SKY, cx, cy, 0 );
            fb.set4( RG_EXT_PLANE_ALBEDO, x, y, h.albedo.x, h.albedo.y, h.albedo.z, 0 );
            fb.set4( RG_EXT_PLANE_NORMAL, cx, cy, EncodeNormal( h.normal ) );
            fb.set4( RG_EXT_PLANE_METALLIC_ROUGHNESS, cx, cy, h.metallic, h.roughness );
            fb.set4( RG_EXT_PLANE_DEPTH_WORLD, cx, cy, depth );
            fb.set4( RG_EXT_PLANE_DEPTH_GRAD, cx, cy, std::sqrt( gx * gx + gy * gy ) );
            fb.set4( RG_EXT_PLANE_MOTION, cx, cy, sxp - sxc, syp - syc, length( vPrev ) - depth, 0 );
            fb.set4( RG_EXT_PLANE_SURFACE_POSITION, cx, cy, h.position.x, h.position.y, h.position.z, CustomIndex( gr.flags, p.skyGeometry != 0 ) );
            fb.set4( RG_EXT_PLANE_VISIBILITY_BUFFER, cx, cy, 1, double( ( h.localPrim << 12 ) | ( h.geom & 0xFFFu ) ), hit.u, hit.v );
            fb.set4( RG_EXT_PLANE_VIEW_DIRECTION, cx, cy, dir.x, dir.y, dir.z, 0 );
            // the camera's medium attenuates the primary segment; screen
            // emission and acid fog (RaygenPrimary.inl:321-328)
            const D3 thr = Transmittance( p, p.cameraMediaType & 3u, depth );
            const D3 fog = GlowingFog( p, p.cameraMediaType & 3u, depth );
            fb.set4( RG_EXT_PLANE_THROUGHPUT, cx, cy, thr.x, thr.y, thr.z, 0 );
            fb.set4( RG_EXT_PLANE_SCREEN_EMISSION, x, y, h.emission.x * thr.x, h.emission.y * thr.y, h.emission.z * thr.z, 0 );
            fb.set4( RG_EXT_PLANE_ACID_FOG, x, y, fog.x, fog.y, fog.z, 0 );
        }
    } );

    // ---- reflection / refraction chain
    ReflectRefract( c, rays, hits );

    // ---- light grid
    LightGrid( c );

    // ---- initial reservoirs (directional visibility)
    std::vector< Res > pendDir( n );
    std::vector< uint32_t > pendSalt( n );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t  i = size_t( y ) * W + x;
            const Surface s = FetchSurface( c, x, y );
            rays[ i ]       = NullRay();
            if( s.isSky )
            {
                storeRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, x, y, Res{} );
                pendDir[ i ] = Res{};
                continue;
            }
            const uint32_t seed = seedOf( x, y, p.frameId );
            uint32_t       salt = SALT_INITIAL_RESERVOIRS;
            double         pu, pv;
            LightPointRnd( seed, &pu, &pv );
            Res reg, dir;
            InitialCandidates( c, s, seed, salt, pu, pv, false, reg, dir );
            storeRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, x, y, reg );
            pendDir[ i ]  = dir;
            pendSalt[ i ] = salt;
            if( valid( dir ) )
            {
                rays[ i ]  = ShadowRay( s, sampleLight( c.lights[ 0 ], s.position, pu, pv, c.p.polyLightSpotlightFactor ).position );
                masks[ i ] = ShadowMask( p, s.custom, 0u );
            }
        }
    } );
    if( p.directionalLightExists )
    {
        Trace( c, rays, masks, hits, RG_EXT_TRACE_ANY_HIT );
    }
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            if( fb.get( RG_EXT_PLANE_IS_SKY, x, y, 0 ) != 0 ) continue;
            Res dir = pendDir[ i ];
            if( valid( dir ) && hits[ i ].primitiveId != RG_EXT_MISS ) dir = Res{};
            uint32_t salt = pendSalt[ i ];
            // the stored FP32 reservoir, as the kernel unpacks it
            Res reg = loadRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, x, y, false );
            storeRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, x, y, CombineInitial( reg, dir, seedOf( x, y, p.frameId ), salt ) );
        }
    } );

    // ---- direct illumination (spatiotemporal reuse, shadow ray)
    std::vector< D3 >     pD( n ), pS( n );
    std::vector< double > pDist( n );
    std::vector< char >   pSky( n );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t  i = size_t( y ) * W + x;
            const Surface s = FetchSurface( c, x, y );
            rays[ i ]       = NullRay();
            pSky[ i ]       = s.isSky;
            if( s.isSky )
            {
                storeRes( fb, RG_EXT_PLANE_RESERVOIRS, x, y, Res{} );
                continue;
            }
            const uint32_t seed = seedOf( x, y, p.frameId );
            double         dist = MAX_RAY_LEN;
            D3             diff = mk( 0, 0, 0 ), spec = mk( 0, 0, 0 );
            Res            res;
            if( DirectValid( c, 0 ) )
            {
                double pu, pv;
                LightPointRnd( seed, &pu, &pv );
                int lo, hi;
                RenderArea( x, W, &lo, &hi );
                const double mx = fb.get( RG_EXT_PLANE_MOTION, x, y, 0 ), my = fb.get( RG_EXT_PLANE_MOTION, x, y, 1 ), mz = fb.get( RG_EXT_PLANE_MOTION, x, y, 2 );
                const double depthCur = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x, y, 0 );
                const double px = ( ( x + 0.5 ) * ( 1.0 / ( W / 2.0 ) ) + mx ) * ( W / 2.0 );
                const double py = ( ( y + 0.5 ) * ( 1.0 / double( H ) ) + my ) * H;
                uint32_t     salt = SALT_LIGHT_CHOOSE_DIRECT;
                const Res    init = loadRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, x, y, false );
                Res          comb = init;
                for( uint32_t k = 0; k < p.restirTemporalSamples; k++ )
                {
                    double r4[ 4 ];
                    rnd8_4( seed, salt++, r4 );
                    const int    qx    = int( std::floor( px + ( r4[ 0 ] * 2.0 - 1.0 ) * 2.0 ) );
                    const int    qy    = int( std::floor( py + ( r4[ 1 ] * 2.0 - 1.0 ) * 2.0 ) );
                    const double dPrev = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0, true );
                    const D3     nPrev = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0, true ) );
                    if( !ReuseOk( lo, hi, H, qx, qy, depthCur, dPrev - mz, s.normal, nPrev, 0.1, 0.5 ) ) continue;
                    Res t = loadRes( fb, RG_EXT_PLANE_RESERVOIRS, qx, qy, true );
                    normalizeRes( t, init.M * 20 );
                    double tpCur = 0.0;
                    if( t.sel != LIGHT_NONE )
                    {
                        const uint32_t cur = t.sel < c.prevToCur.size() ? c.prevToCur[ t.sel ] : 0xFFFFFFFFu;
                        if( cur != 0xFFFFFFFFu && cur != LIGHT_NONE )
                        {
                            tpCur = targetPdf( s, sampleLight( LightAt( c, cur ), s.position, pu, pv, c.p.polyLightSpotlightFactor ), false );
                            t.sel = cur;
                        }
                    }
                    combineNewSurf( comb, t, tpCur, rnd16( seed, salt++ ) );
                }
                for( uint32_t k = 0; k < p.restirSpatialSamples; k++ )
                {
                    double r4[ 4 ];
                    rnd8_4( seed, salt++, r4 );
                    const int qx = x + int( ( float( r4[ 0 ] ) * 2.0f - 1.0f ) * 30.0f );
                    const int qy = y + int( ( float( r4[ 1 ] ) * 2.0f - 1.0f ) * 30.0f );
                    const double dO = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0 );
                    const D3     nO = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0 ) );
                    if( !ReuseOk( lo, hi, H, qx, qy, depthCur, dO, s.normal, nO, 0.1, 0.5 ) ) continue;
                    combine( comb, loadRes( fb, RG_EXT_PLANE_RESERVOIRS_INITIAL, qx, qy, false ), rnd16( seed, salt++ ) );
                }
                if( !( comb.wsum <= 0.0 || comb.sel == LIGHT_NONE ) )
                {
                    res              = comb;
                    const LSample ls = sampleLight( LightAt( c, res.sel ), s.position, pu, pv, c.p.polyLightSpotlightFactor );
                    shade( s, ls, selWeight( res ), false, &diff, &spec );
                    if( lum( diff + spec ) <= 0.0 )
                    {
                        diff = spec = mk( 0, 0, 0 );
                    }
                    else if( 0u < p.maxBounceShadowsLights )
                    {
                        rays[ i ]  = ShadowRay( s, ls.position );
                        masks[ i ] = ShadowMask( p, s.custom, res.sel );
                        dist       = length( ls.position - s.position );
                    }
                }
            }
            storeRes( fb, RG_EXT_PLANE_RESERVOIRS, x, y, res );
            pD[ i ] = diff, pS[ i ] = spec, pDist[ i ] = dist;
        }
    } );
    Trace( c, rays, masks, hits, RG_EXT_TRACE_ANY_HIT );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            if( pSky[ i ] )
            {
                fb.set4( RG_EXT_PLANE_UNFILTERED_DIRECT, x, y, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_UNFILTERED_SPECULAR, x, y, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_GRADIENT_INPUTS, x, y, 0, 0 );
                continue;
            }
            D3 diff = pD[ i ], spec = pS[ i ];
            if( !IsNull( rays[ i ] ) && hits[ i ].primitiveId != RG_EXT_MISS ) diff = spec = mk( 0, 0, 0 );
            int rx, ry;
            or_checkerboard_to_regular( x, y, W, &rx, &ry );
            const D3 alb = mk( fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 0 ), fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 1 ), fb.get( RG_EXT_PLANE_ALBEDO, rx, ry, 2 ) );
            const D3 sc2 = SpecColor( alb, fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 0 ) );
            fb.set4( RG_EXT_PLANE_UNFILTERED_DIRECT, x, y, diff.x, diff.y, diff.z, 0 );
            fb.set4( RG_EXT_PLANE_UNFILTERED_SPECULAR, x, y, spec.x / std::max( 0.01, sc2.x ), spec.y / std::max( 0.01, sc2.y ), spec.z / std::max( 0.01, sc2.z ), 0 );
            fb.at( RG_EXT_PLANE_VIEW_DIRECTION, x, y )[ 3 ] = pDist[ i ];
            fb.set4( RG_EXT_PLANE_GRADIENT_INPUTS, x, y, lum( diff ), lum( spec ) );
        }
    } );

    // ---- indirect, init: bounce 1 (GGX-VNDF), direct light at its hit, bounce 2
    std::vector< D3 >     b1dir( n ), hpos( n ), hnrm( n ), halb( n ), hdiff( n ), b2dir( n );
    std::vector< double > b1pdf( n ), b2pdf( n );
    std::vector< char >   b1sky( n ), hIsDir( n ), hShadow( n ), pxSky( n );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            Surface      s = FetchSurface( c, x, y );
            rays[ i ]      = NullRay();
            pxSky[ i ]     = s.isSky;
            if( s.isSky ) continue;
            s.position          = s.position + s.toViewer * 0.01; // RAY_ORIGIN_LEAK_BIAS
            const uint32_t seed = seedOf( x, y, p.frameId );
            double         bn[ 4 ];
            BlueNoise8( pt, seed, SALT_SPEC_BOUNCE_BASE + 1, bn );
            double   opdf;
            const D3 d = sampleSmithGGX( s.normal, s.toViewer, s.roughness, bn[ 0 ], bn[ 1 ], &opdf, true );
            rays[ i ]  = MakeRay( s.position + s.normal * 0.01, d, 0.001, p.rayLength );
            masks[ i ] = IndirectMask( p, s.custom );
            b1dir[ i ] = d, b1pdf[ i ] = opdf;
        }
    } );
    Trace( c, rays, masks, hits, RG_EXT_TRACE_CLOSEST );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            rays2[ i ] = NullRay(), rays3[ i ] = NullRay();
            b1sky[ i ] = 0;
            if( pxSky[ i ] ) continue;
            const RgExtRay r   = rays[ i ];
            const RgExtHit hit = hits[ i ];
            const D3       dir = mk( r.direction[ 0 ], r.direction[ 1 ], r.direction[ 2 ] );
            const D3       org = mk( r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] );
            if( IsSkyHit( c, hit ) )
            {
                b1sky[ i ] = 1;
                continue;
            }
            TexQuery       bq;
            bq.mode             = TQ_BOUNCE; // FIRST_BOUNCE_MIP_BIAS 0 (RtRaygenIndirect.inl:63)
            bq.a                = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 1 );
            const Hit      h    = GetHit( c, hit.primitiveId, hit.u, hit.v, org, bq );
            const Surface  hs   = HitToSurface( c, h, dir );
            const uint32_t seed = seedOf( x, y, p.frameId );
            D3             diff, lp = mk( 0, 0, 0 );
            bool           isDir, needShadow;
            uint32_t       lightIndex = LIGHT_NONE;
            IndirectDirect( c, hs, seed, 1, &diff, &isDir, &needShadow, &lp, &lightIndex );
            if( needShadow )
            {
                rays2[ i ]  = ShadowRay( hs, lp );
                masks2[ i ] = ShadowMask( p, hs.custom, lightIndex );
            }
            double bn[ 4 ];
            BlueNoise8( pt, seed, SALT_DIFF_BOUNCE_BASE + 2, bn );
            double   pdf2;
            const D3 d2 = sampleOrientedHemisphere( hs.normal, bn[ 0 ], bn[ 1 ], &pdf2 );
            rays3[ i ]  = MakeRay( hs.position + hs.normal * 0.01, d2, 0.001, p.rayLength );
            masks3[ i ] = IndirectMask( p, hs.custom );
            hpos[ i ] = h.position, hnrm[ i ] = h.normal, halb[ i ] = h.albedo, hdiff[ i ] = diff;
            hIsDir[ i ] = isDir, hShadow[ i ] = needShadow, b2dir[ i ] = d2, b2pdf[ i ] = pdf2;
        }
    } );
    Trace( c, rays2, masks2, hits2, RG_EXT_TRACE_ANY_HIT );
    Trace( c, rays3, masks3, hits3, RG_EXT_TRACE_CLOSEST );
    // initial samples (RtRaygenIndirect.inl:221-240); positions / radiance kept in FP64
    std::vector< D3 >     sPos( n ), sNrm( n ), sRad( n );
    std::vector< double > sPdf( n );
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            if( pxSky[ i ] )
            {
                fb.set4( RG_EXT_PLANE_UNFILTERED_INDIR, x, y, 0, 0, 0, 0 );
                sPos[ i ] = sNrm[ i ] = sRad[ i ] = mk( 0, 0, 0 ), sPdf[ i ] = 0;
                continue;
            }
            Surface s  = FetchSurface( c, x, y );
            s.position = s.position + s.toViewer * 0.01;
            if( b1sky[ i ] )
            {
                sPos[ i ] = s.position + b1dir[ i ] * MAX_RAY_LEN;
                sNrm[ i ] = -b1dir[ i ];
                sRad[ i ] = SkyColor( p );
            }
            else
            {
                D3 diffuse = hdiff[ i ];
                if( hShadow[ i ] && hits2[ i ].primitiveId != RG_EXT_MISS ) diffuse = mk( 0, 0, 0 );
                if( !hIsDir[ i ] ) diffuse = mk( clampd( diffuse.x, 0, 20 ), clampd( diffuse.y, 0, 20 ), clampd( diffuse.z, 0, 20 ) ); // SHIPPING_HACK
                const RgExtHit h1   = hits[ i ];
                TexQuery       q1;
                q1.mode             = TQ_BOUNCE;
                q1.a                = s.roughness;
                const Hit      hi1  = GetHit( c, h1.primitiveId, h1.u, h1.v, s.position + s.normal * 0.01, q1 );
                const D3       emis = hi1.emission * double( p.emissionMapBoost );
                D3             second;
                const RgExtHit h2 = hits3[ i ];
                if( IsSkyHit( c, h2 ) )
                {
                    second = SkyColor( p ) * b2pdf[ i ];
                }
                else
                {
                    const RgExtRay br  = rays3[ i ];
                    TexQuery       q2;
                    q2.mode            = TQ_BOUNCE;
                    q2.a               = hi1.roughness;
                    q2.b               = 32.0; // SECOND_BOUNCE_MIP_BIAS (RtRaygenIndirect.inl:64)
                    const Hit      hi2 = GetHit( c, h2.primitiveId, h2.u, h2.v, mk( br.origin[ 0 ], br.origin[ 1 ], br.origin[ 2 ] ), q2 );
                    const Surface  hs2 = HitToSurface( c, hi2, b2dir[ i ] );
                    D3             diff2, lp;
                    bool           isDir2, ns2;
                    IndirectDirect( c, hs2, seedOf( x, y, p.frameId ), 2, &diff2, &isDir2, &ns2, &lp );
                    if( !isDir2 ) diff2 = mk( clampd( diff2.x, 0, 20 ), clampd( diff2.y, 0, 20 ), clampd( diff2.z, 0, 20 ) );
                    second = ( hi2.emission * double( p.emissionMapBoost ) + diff2 ) * hi2.albedo * b2pdf[ i ];
                }
                diffuse   = diffuse + second;
                sPos[ i ] = hpos[ i ];
                sNrm[ i ] = hnrm[ i ];
                sRad[ i ] = ( emis + diffuse ) * halb[ i ];
            }
            sPdf[ i ] = b1pdf[ i ];
            fb.set4( RG_EXT_PLANE_INDIR_INITIAL_SAMPLE, x, y, sPos[ i ].x, sPos[ i ].y, sPos[ i ].z, EncodeNormal( sNrm[ i ] ) );
            const double rgb[ 3 ] = { sRad[ i ].x, sRad[ i ].y, sRad[ i ].z };
            fb.set4( RG_EXT_PLANE_INDIR_INITIAL_RADIANCE, x, y, double( or_encode_e5b9g9r9( rgb ) ), sPdf[ i ] );
        }
    } );

    // ---- indirect, final: ReSTIR GI (RtRaygenIndirect.inl:257-434)
    struct RI
    {
        D3       pos{ 0, 0, 0 }, nrm{ 0, 0, 0 }, rad{ 0, 0, 0 };
        double   tp = 0, wsum = 0;
        uint32_t M = 0;
    };
    auto initialAsRes = [ & ]( int qx, int qy ) {
        // the stored sample: E5B9G9R9 radiance, encoded normal (ReservoirIndirect.h:214-233)
        RI r;
        D3 pos = mk( fb.get( RG_EXT_PLANE_INDIR_INITIAL_SAMPLE, qx, qy, 0 ), fb.get( RG_EXT_PLANE_INDIR_INITIAL_SAMPLE, qx, qy, 1 ), fb.get( RG_EXT_PLANE_INDIR_INITIAL_SAMPLE, qx, qy, 2 ) );
        D3 nrm = DecodeNormal( fb.get( RG_EXT_PLANE_INDIR_INITIAL_SAMPLE, qx, qy, 3 ) );
        double rgb[ 3 ];
        or_decode_e5b9g9r9( uint32_t( fb.get( RG_EXT_PLANE_INDIR_INITIAL_RADIANCE, qx, qy, 0 ) ), rgb );
        const D3     rad = mk( rgb[ 0 ], rgb[ 1 ], rgb[ 2 ] );
        const double oos = fb.get( RG_EXT_PLANE_INDIR_INITIAL_RADIANCE, qx, qy, 1 );
        const double tp  = lum( rad );
        const double w   = tp * oos;
        r.wsum += w;
        r.M += 1;
        if( 0.5 * r.wsum < w )
        {
            r.pos = pos, r.nrm = nrm, r.rad = rad, r.tp = tp;
        }
        return r;
    };
    ParallelFor( H, threads, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            Surface s = FetchSurface( c, x, y );
            if( s.isSky )
            {
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, x, y, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, x, y, 0, 0, 0, 0 );
                fb.set4( OR_PLANE_GI_TIE, x, y, 0.0 );
                continue;
            }
            s.position          = s.position + s.toViewer * 0.01;
            const uint32_t seed = seedOf( x, y, p.frameId );
            uint32_t       salt = SALT_RESAMPLE_INDIRECT_BASE;
            RI             comb = initialAsRes( x, y );
            bool           tie  = false; // see OR_PLANE_GI_TIE
            int            lo, hi;
            RenderArea( x, W, &lo, &hi );
            const double mx = fb.get( RG_EXT_PLANE_MOTION, x, y, 0 ), my = fb.get( RG_EXT_PLANE_MOTION, x, y, 1 ), mz = fb.get( RG_EXT_PLANE_MOTION, x, y, 2 );
            const double depthCur = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x, y, 0 );
            const double px = ( ( x + 0.5 ) * ( 1.0 / ( W / 2.0 ) ) + mx ) * ( W / 2.0 );
            const double py = ( ( y + 0.5 ) * ( 1.0 / double( H ) ) + my ) * H;
            const double dw = smoothstep( MIN_ROUGH, 0.5 + 0.25, s.roughness );
            int          spatialCount = int( 6.0 * dw );
            {
                double r4[ 4 ];
                rnd8_4( seed, salt++, r4 );
                const int qx  = int( std::floor( px + ( r4[ 0 ] * 2.0 - 1.0 ) * square( dw ) * 2.0 ) );
                const int qy  = int( std::floor( py + ( r4[ 1 ] * 2.0 - 1.0 ) * square( dw ) * 2.0 ) );
                bool      use = fb.get( RG_EXT_PLANE_IS_SKY, qx, qy, 0 ) == 0.0;
                if( use )
                {
                    const double dPrev = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0, true );
                    const D3     nPrev = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0, true ) );
                    use                = ReuseOk( lo, hi, H, qx, qy, depthCur, dPrev - mz, s.normal, nPrev, 0.05, 0.0 );
                }
                if( use && fb.get( RG_EXT_PLANE_DIS_GRADIENT_HISTORY, qx / 3, qy / 3, 1 ) > 0.25 )
                {
                    spatialCount++;
                    use = false;
                }
                if( use )
                {
                    tie = tie || fb.get( OR_PLANE_GI_TIE, qx, qy, 0, true ) != 0.0;
                    RI t;
                    t.pos  = mk( fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, qx, qy, 0, true ), fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, qx, qy, 1, true ),
                                 fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, qx, qy, 2, true ) );
                    t.nrm  = DecodeNormal( fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, qx, qy, 3, true ) );
                    double rgb[ 3 ];
                    or_decode_e5b9g9r9( uint32_t( fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, qx, qy, 0, true ) ), rgb );
                    t.rad  = mk( rgb[ 0 ], rgb[ 1 ], rgb[ 2 ] );
                    t.tp   = fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, qx, qy, 1, true );
                    t.wsum = fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, qx, qy, 2, true );
                    t.M    = uint32_t( fb.get( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, qx, qy, 3, true ) );
                    t.wsum /= double( std::max( t.M, 1u ) );
                    t.M = std::min( t.M, 20u );
                    t.wsum *= double( t.M );
                    const double rnd = rnd16( seed, salt++ );
                    comb.wsum += t.wsum;
                    comb.M += t.M;
                    if( rnd * comb.wsum < t.wsum )
                    {
                        comb.pos = t.pos, comb.nrm = t.nrm, comb.rad = t.rad, comb.tp = t.tp;
                    }
                }
            }
            {
                const double radius  = mixd( 8.0, 32.0, clampd( H / 1080.0, 0.0, 1.0 ) );
                uint32_t     nobiasM = comb.M;
                for( int k = 0; k < spatialCount; k++ )
                {
                    double r4[ 4 ];
                    rnd8_4( seed, salt++, r4 );
                    const int qx = x + int( ( float( r4[ 0 ] ) * 2.0f - 1.0f ) * float( radius ) );
                    const int qy = y + int( ( float( r4[ 1 ] ) * 2.0f - 1.0f ) * float( radius ) );
                    if( fb.get( RG_EXT_PLANE_IS_SKY, qx, qy, 0 ) != 0.0 ) continue;
                    const double dO = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0 );
                    const D3     nO = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0 ) );
                    if( !ReuseOk( lo, hi, H, qx, qy, depthCur, dO, s.normal, nO, 0.05, 0.0 ) ) continue;
                    const RI rq  = initialAsRes( qx, qy );
                    const D3 x1q = mk( fb.get( RG_EXT_PLANE_SURFACE_POSITION, qx, qy, 0 ), fb.get( RG_EXT_PLANE_SURFACE_POSITION, qx, qy, 1 ),
                                       fb.get( RG_EXT_PLANE_SURFACE_POSITION, qx, qy, 2 ) );
                    auto     G   = [ & ]( D3 from, D3 to ) {
                        D3           d   = to - from;
                        const double len = std::max( length( d ), 0.001 );
                        d                = d / len;
                        return std::max( 0.0, dot( rq.nrm, d ) ) * safeRcp( square( len ) );
                    };
                    {
                        // The cosines at the sample towards both receivers, and the
                        // error of each when evaluated in FP32: the positions carry
                        // one rounding each (2^-24 of their magnitude), so a receiver
                        // near the sample's plane gets a cosine whose relative error
                        // can exceed 1e-3 (FP32 cancellation), and one at the plane
                        // a cosine of either sign.
                        auto illConditioned = [ & ]( D3 recv ) {
                            const D3     d   = recv - rq.pos;
                            const double len = std::max( length( d ), 1e-30 );
                            const double mag = std::max( { std::fabs( recv.x ), std::fabs( recv.y ), std::fabs( recv.z ) } ) +
                                               std::max( { std::fabs( rq.pos.x ), std::fabs( rq.pos.y ), std::fabs( rq.pos.z ) } );
                            const double err = 6e-8 * ( mag / len + 1.0 );
                            return std::fabs( dot( rq.nrm, d ) / len ) < 1e3 * err;
                        };
                        tie = tie || illConditioned( s.position ) || illConditioned( x1q );
                    }
                    const double ooj   = clampd( safeRcp( G( rq.pos, s.position ) ) * G( rq.pos, x1q ), 0.0, 1.0 ); // SHIPPING_HACK
                    const double tpCur = lum( rq.rad ) * ooj;
                    const double w     = tpCur * safeRcp( rq.tp ) * rq.wsum;
                    const double rnd   = rnd16( seed, salt++ );
                    comb.wsum += w;
                    comb.M += rq.M;
                    if( rnd * comb.wsum < w )
                    {
                        comb.pos = rq.pos, comb.nrm = rq.nrm, comb.rad = rq.rad, comb.tp = tpCur;
                    }
                    if( tpCur > 0.0 ) nobiasM += rq.M;
                }
                comb.M = nobiasM;
            }
            fb.set4( OR_PLANE_GI_TIE, x, y, tie ? 1.0 : 0.0 );
            if( std::isfinite( comb.wsum ) && comb.wsum >= 0.0 )
            {
                const double rgb[ 3 ] = { comb.rad.x, comb.rad.y, comb.rad.z };
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, x, y, comb.pos.x, comb.pos.y, comb.pos.z, EncodeNormal( comb.nrm ) );
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, x, y, double( or_encode_e5b9g9r9( rgb ) ), comb.tp, comb.wsum, double( comb.M ) );
            }
            else
            {
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE, x, y, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT, x, y, 0, 0, 0, 0 );
            }
            const double oneOverPdf = safeRcp( comb.tp ) * ( comb.wsum / double( std::max( 1u, comb.M ) ) );
            D3           diff = mk( 0, 0, 0 ), spec = mk( 0, 0, 0 );
            const D3     l  = safeNormalize( comb.pos - s.position );
            const double nl = dot( s.normal, l );
            if( nl > 0.0 )
            {
                diff = comb.rad * ( nl / PI ) * oneOverPdf;
                spec = ( comb.rad * nl ) * evalGGX( s.normal, s.toViewer, l, s.roughness, s.specColor, true ) * oneOverPdf;
            }
            double*  us     = fb.at( RG_EXT_PLANE_UNFILTERED_SPECULAR, x, y );
            const D3 direct = mk( us[ 0 ], us[ 1 ], us[ 2 ] );
            if( lum( direct ) < lum( spec ) )
            {
                fb.at( RG_EXT_PLANE_VIEW_DIRECTION, x, y )[ 3 ] = length( comb.pos - s.position );
            }
            us[ 0 ] = direct.x + spec.x / std::max( 0.01, s.specColor.x );
            us[ 1 ] = direct.y + spec.y / std::max( 0.01, s.specColor.y );
            us[ 2 ] = direct.z + spec.z / std::max( 0.01, s.specColor.z );
            us[ 3 ] = 0;
            fb.set4( RG_EXT_PLANE_UNFILTERED_INDIR, x, y, diff.x, diff.y, diff.z, 0 );
        }
    } );

    // ---- gradient samples (RtGradients.rgen:44-181)
    std::vector< RgExtRay > grays( size_t( GW ) * GH, NullRay() );
    std::vector< uint8_t >  gmasks( size_t( GW ) * GH, 0 );
    std::vector< RgExtHit > ghits;
    struct GP
    {
        double   lumD = 0, lumS = 0, prevD = 0, prevS = 0;
        int      psx = 0, psy = 0;
        uint32_t strata = 10, state = 0;
    };
    std::vector< GP > gp( size_t( GW ) * GH );
    ParallelFor( GH, threads, [ & ]( int gy ) {
        for( int gx = 0; gx < GW; gx++ )
        {
            const size_t gi = size_t( gy ) * GW + gx;
            GP&          g  = gp[ gi ];
            if( gx * 3 >= W || gy * 3 >= H ) continue;
            if( p.gradientMultDiffuse < 0.001f && p.gradientMultIndirect < 0.001f && p.gradientMultSpecular < 0.001f )
            {
                fb.set4( RG_EXT_PLANE_DIS_PING_GRADIENT, gx, gy, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_GRADIENT_PREV_PIX, gx, gy, 10 );
                continue;
            }
            const uint32_t prevStrata = uint32_t( fb.get( RG_EXT_PLANE_GRADIENT_PREV_PIX, gx, gy, 0 ) );
            int            psx = 0, psy = 0, csx = 0, csy = 0;
            uint32_t       curStrata = 0;
            double         pl[ 2 ]   = { 0, 0 };
            for( int yy = 0; yy < 3; yy++ )
            {
                for( int xx = 0; xx < 3; xx++ )
                {
                    if( uint32_t( yy * 3 + xx ) == prevStrata ) continue;
                    const int    cx = gx * 3 + xx, cy = gy * 3 + yy;
                    const double mx = fb.get( RG_EXT_PLANE_MOTION, cx, cy, 0 ), my = fb.get( RG_EXT_PLANE_MOTION, cx, cy, 1 );
                    const int    qx = int( std::floor( ( ( cx + 0.5 ) * ( 1.0 / ( W / 2.0 ) ) + mx ) * ( W / 2.0 ) ) );
                    const int    qy = int( std::floor( ( ( cy + 0.5 ) * ( 1.0 / double( H ) ) + my ) * H ) );
                    const double l0 = fb.get( RG_EXT_PLANE_GRADIENT_INPUTS, qx, qy, 0, true ), l1 = fb.get( RG_EXT_PLANE_GRADIENT_INPUTS, qx, qy, 1, true );
                    if( std::max( l0, l1 ) > std::max( pl[ 0 ], pl[ 1 ] ) )
                    {
                        psx = qx, psy = qy, csx = cx, csy = cy, curStrata = uint32_t( yy * 3 + xx );
                        pl[ 0 ] = l0, pl[ 1 ] = l1;
                    }
                }
            }
            int lo, hi;
            RenderArea( csx, W, &lo, &hi );
            if( !( psy >= 0 && psy < H && psx >= lo && psx < hi ) || fb.get( RG_EXT_PLANE_IS_SKY, csx, csy, 0 ) != 0.0 || ( pl[ 0 ] <= 0.0 && pl[ 1 ] <= 0.0 ) )
            {
                fb.set4( RG_EXT_PLANE_DIS_PING_GRADIENT, gx, gy, 0, 0, 0, 0 );
                fb.set4( RG_EXT_PLANE_GRADIENT_PREV_PIX, gx, gy, 10 );
                continue;
            }
            const uint32_t oldSeed = seedOf( psx, psy, p.frameId - 1u );
            Surface        ps;
            ps.isSky     = false;
            ps.albedo    = mk( 1, 1, 1 );
            ps.position  = mk( fb.get( RG_EXT_PLANE_SURFACE_POSITION, psx, psy, 0, true ), fb.get( RG_EXT_PLANE_SURFACE_POSITION, psx, psy, 1, true ),
                               fb.get( RG_EXT_PLANE_SURFACE_POSITION, psx, psy, 2, true ) );
            ps.specColor = SpecColor( ps.albedo, fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, psx, psy, 0, true ) );
            ps.roughness = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, psx, psy, 1, true );
            ps.normal    = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, psx, psy, 0, true ) );
            ps.custom    = uint32_t( fb.get( RG_EXT_PLANE_SURFACE_POSITION, psx, psy, 3, true ) );
            bool matched = false;
            if( fb.get( RG_EXT_PLANE_VISIBILITY_BUFFER, psx, psy, 0, true ) == 1.0 )
            {
                const uint32_t packed = uint32_t( fb.get( RG_EXT_PLANE_VISIBILITY_BUFFER, psx, psy, 1, true ) );
                const uint32_t pg = packed & 0xFFFu, local = packed >> 12;
                const uint32_t cg = pg < c.geomPrevToCur.size() ? c.geomPrevToCur[ pg ] : 0xFFFFFFFFu;
                if( cg < sc->geomCount && local < sc->geoms[ cg ].triCount )
                {
                    const float* t = sc->tris + size_t( sc->geoms[ cg ].firstTri + local ) * 9;
                    const double u = fb.get( RG_EXT_PLANE_VISIBILITY_BUFFER, psx, psy, 2, true ), v = fb.get( RG_EXT_PLANE_VISIBILITY_BUFFER, psx, psy, 3, true );
                    ps.position    = mk( t[ 0 ], t[ 1 ], t[ 2 ] ) * ( 1.0 - u - v ) + mk( t[ 3 ], t[ 4 ], t[ 5 ] ) * u + mk( t[ 6 ], t[ 7 ], t[ 8 ] ) * v;
                    matched        = true;
                }
            }
            ps.toViewer = normalize( mk( p.cameraPositionPrev[ 0 ], p.cameraPositionPrev[ 1 ], p.cameraPositionPrev[ 2 ] ) - ps.position );
            g.state     = 1;
            if( matched )
            {
                Res r = loadRes( fb, RG_EXT_PLANE_RESERVOIRS, psx, psy, true );
                if( valid( r ) )
                {
                    const uint32_t cur = r.sel < c.prevToCur.size() ? c.prevToCur[ r.sel ] : 0xFFFFFFFFu;
                    if( cur != 0xFFFFFFFFu && cur != LIGHT_NONE && DirectValid( c, 0 ) )
                    {
                        r.sel = cur;
                        double pu, pv;
                        LightPointRnd( oldSeed, &pu, &pv );
                        const LSample ls = sampleLight( LightAt( c, cur ), ps.position, pu, pv, c.p.polyLightSpotlightFactor );
                        D3            d, sp;
                        shade( ps, ls, selWeight( r ), false, &d, &sp );
                        if( lum( d + sp ) > 0.0 )
                        {
                            g.lumD = lum( d ), g.lumS = lum( sp );
                            if( 0u < p.maxBounceShadowsLights )
                            {
                                grays[ gi ]  = ShadowRay( ps, ls.position );
                                gmasks[ gi ] = ShadowMask( p, ps.custom, cur );
                                g.state      = 2;
                            }
                        }
                    }
                }
            }
            g.prevD = pl[ 0 ], g.prevS = pl[ 1 ], g.psx = psx, g.psy = psy, g.strata = curStrata;
        }
    } );
    Trace( c, grays, gmasks, ghits, RG_EXT_TRACE_ANY_HIT );
    for( int gy = 0; gy < GH; gy++ )
    {
        for( int gx = 0; gx < GW; gx++ )
        {
            const size_t gi = size_t( gy ) * GW + gx;
            const GP&    g  = gp[ gi ];
            if( g.state == 0 ) continue;
            const double vis  = ( g.state == 2 && ghits[ gi ].primitiveId != RG_EXT_MISS ) ? 0.0 : 1.0;
            const double fwdD = g.lumD * vis, fwdS = g.lumS * vis;
            double       gr[ 3 ];
            gr[ 0 ] = Antilag( fwdD - g.prevD, std::max( fwdD, g.prevD ) );
            {
                double     gs = std::max( 0.0, g.prevD - fwdD );
                const D3   ia = mk( fb.get( RG_EXT_PLANE_INDIR_ACCUM, g.psx, g.psy, 0, true ), fb.get( RG_EXT_PLANE_INDIR_ACCUM, g.psx, g.psy, 1, true ),
                                    fb.get( RG_EXT_PLANE_INDIR_ACCUM, g.psx, g.psy, 2, true ) );
                gs *= 1.0 - smoothstep( 1.0, 10.0, lum( ia ) );
                gr[ 1 ] = Antilag( gs, std::max( fwdD, g.prevD ) );
            }
            gr[ 2 ] = Antilag( fwdS - g.prevS, std::max( fwdS, g.prevS ) );
            gr[ 0 ] *= p.gradientMultDiffuse;
            gr[ 1 ] *= p.gradientMultIndirect;
            gr[ 2 ] *= p.gradientMultSpecular;
            const int hx = g.psx / 3, hy = g.psy / 3;
            fb.set4( RG_EXT_PLANE_DIS_PING_GRADIENT, gx, gy, mixd( gr[ 0 ], fb.get( RG_EXT_PLANE_DIS_GRADIENT_HISTORY, hx, hy, 0 ), 0.5 ),
                     mixd( gr[ 1 ], fb.get( RG_EXT_PLANE_DIS_GRADIENT_HISTORY, hx, hy, 1 ), 0.5 ),
                     mixd( gr[ 2 ], fb.get( RG_EXT_PLANE_DIS_GRADIENT_HISTORY, hx, hy, 2 ), 0.5 ), 0 );
            fb.set4( RG_EXT_PLANE_GRADIENT_PREV_PIX, gx, gy, double( g.strata ) );
        }
    }
}

// VertexPreprocessPartial.inl:62-120 (GEOM_INST_FLAG_GENERATE_NORMALS with
// normalSign +1): one pass over the triangles in index order.  The shader
// computes normalize(cross(p1 - p0, p2 - p0)) in FP32; this restates it in
// FP32 with one rounding per operation in the order written (edges, cross
// product, the dot product left to right, a correctly rounded square root,
// three divisions), so the vertex normals are the same floats as the
// kernel's.
extern "C" void or_flat_normals( const float* positions, uint32_t vertexCount, const uint32_t* indices, uint32_t indexCount, float* normals )
{
    const uint32_t triCount = indices ? indexCount / 3 : vertexCount / 3;
    for( uint32_t t = 0; t < triCount; t++ )
    {
        uint32_t v[ 3 ];
        for( int k = 0; k < 3; k++ ) v[ k ] = indices ? indices[ 3 * t + k ] : 3 * t + k;
        const float* p0 = positions + 3 * v[ 0 ];
        const float* p1 = positions + 3 * v[ 1 ];
        const float* p2 = positions + 3 * v[ 2 ];
        const float  ax = p1[ 0 ] - p0[ 0 ], ay = p1[ 1 ] - p0[ 1 ], az = p1[ 2 ] - p0[ 2 ];
        const float  bx = p2[ 0 ] - p0[ 0 ], by = p2[ 1 ] - p0[ 1 ], bz = p2[ 2 ] - p0[ 2 ];
        const float  cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const float  len = std::sqrt( cx * cx + cy * cy + cz * cz );
        for( int k = 0; k < 3; k++ )
        {
            normals[ 3 * v[ k ] + 0 ] = cx / len;
            normals[ 3 * v[ k ] + 1 ] = cy / len;
            normals[ 3 * v[ k ] + 2 ] = cz / len;
        }
    }
}
