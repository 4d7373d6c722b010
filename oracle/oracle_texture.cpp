// CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Albedo textures in FP64.  The sRGB transfer functions are IEC 61966-2-1;
// the mip chain restates a linear-filtered vkCmdBlitImage of each level into
// the next (TextureUploader.cpp:104-175): every destination texel centre is
// mapped into the source level and filtered bilinearly there (clamp to edge),
// in linear space, then re-encoded to 8-bit sRGB; sampling follows the
// Vulkan specification's texel filtering for the reference's samplers
// (SamplerManager.cpp:213-265): magnification filter at lod <= 0,
// minification filter (linear when forced) otherwise, linear between mips,
// repeat or clamp-to-edge wrapping, isotropic LOD from the longer gradient.
#include "oracle_texture.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct OrTexture
{
    std::vector< std::vector< uint32_t > > levels;
    std::vector< int >                     w, h;
    bool                                   magNearest, minNearest, clampU, clampV, srgb;
};

namespace
{
double SrgbDecode( uint32_t c8 )
{
    const double c = c8 / 255.0;
    return c <= 0.04045 ? c / 12.92 : std::pow( ( c + 0.055 ) / 1.055, 2.4 );
}

uint32_t SrgbEncode( double c )
{
    c              = std::min( std::max( c, 0.0 ), 1.0 );
    const double s = c <= 0.0031308 ? c * 12.92 : 1.055 * std::pow( c, 1.0 / 2.4 ) - 0.055;
    return uint32_t( std::min( std::max( std::nearbyint( s * 255.0 ), 0.0 ), 255.0 ) );
}

int Wrap( int i, int n, bool clamp )
{
    if( clamp ) return std::min( std::max( i, 0 ), n - 1 );
    const int m = i % n;
    return m < 0 ? m + n : m;
}

double Decode( bool srgb, uint32_t c8 )
{
    return srgb ? SrgbDecode( c8 ) : c8 / 255.0;
}

void Texel( const OrTexture& t, int l, int x, int y, double out[ 4 ] )
{
    const uint32_t c = t.levels[ l ][ size_t( Wrap( y, t.h[ l ], t.clampV ) ) * t.w[ l ] + Wrap( x, t.w[ l ], t.clampU ) ];
    out[ 0 ] = Decode( t.srgb, c & 0xFFu );
    out[ 1 ] = Decode( t.srgb, ( c >> 8 ) & 0xFFu );
    out[ 2 ] = Decode( t.srgb, ( c >> 16 ) & 0xFFu );
    out[ 3 ] = ( c >> 24 ) / 255.0;
}

void SampleLevel( const OrTexture& t, int l, double u, double v, bool nearest, double out[ 4 ] )
{
    const int W = t.w[ l ], H = t.h[ l ];
    if( nearest )
    {
        Texel( t, l, int( std::floor( u * W ) ), int( std::floor( v * H ) ), out );
        return;
    }
    const double x = u * W - 0.5, y = v * H - 0.5;
    const double fx = std::floor( x ), fy = std::floor( y );
    const double ax = x - fx, ay = y - fy;
    double       c[ 4 ][ 4 ];
    for( int k = 0; k < 4; k++ ) Texel( t, l, int( fx ) + ( k & 1 ), int( fy ) + ( k >> 1 ), c[ k ] );
    for( int i = 0; i < 4; i++ )
    {
        out[ i ] = ( c[ 0 ][ i ] * ( 1 - ax ) + c[ 1 ][ i ] * ax ) * ( 1 - ay ) + ( c[ 2 ][ i ] * ( 1 - ax ) + c[ 3 ][ i ] * ax ) * ay;
    }
}
} // namespace

extern "C" OrTexture* or_texture_create( const uint8_t* rgba, uint32_t width, uint32_t height, int filter, int addressU, int addressV,
                                         int forceMinificationLinear, int srgb )
{
    OrTexture* t   = new OrTexture;
    t->srgb        = srgb != 0;
    const bool nearest = filter == RG_SAMPLER_FILTER_NEAREST;
    t->magNearest  = nearest;
    t->minNearest  = nearest && !forceMinificationLinear;
    t->clampU      = addressU == RG_SAMPLER_ADDRESS_MODE_CLAMP;
    t->clampV      = addressV == RG_SAMPLER_ADDRESS_MODE_CLAMP;
    t->levels.emplace_back( size_t( width ) * height );
    std::copy( reinterpret_cast< const uint32_t* >( rgba ), reinterpret_cast< const uint32_t* >( rgba ) + size_t( width ) * height,
               t->levels[ 0 ].begin() );
    t->w.push_back( int( width ) );
    t->h.push_back( int( height ) );
    while( ( t->w.back() > 1 || t->h.back() > 1 ) && t->levels.size() < 16 )
    {
        const int sw = t->w.back(), sh = t->h.back();
        const int dw = std::max( sw / 2, 1 ), dh = std::max( sh / 2, 1 );
        std::vector< uint32_t > dst( size_t( dw ) * dh );
        const int l = int( t->levels.size() ) - 1;
        for( int y = 0; y < dh; y++ )
        {
            for( int x = 0; x < dw; x++ )
            {
                // blit: destination texel centre in source texel space, bilinear, clamped
                const double sx = ( x + 0.5 ) * double( sw ) / dw - 0.5, sy = ( y + 0.5 ) * double( sh ) / dh - 0.5;
                const double fx = std::floor( sx ), fy = std::floor( sy ), ax = sx - fx, ay = sy - fy;
                double       acc[ 4 ] = { 0, 0, 0, 0 };
                for( int k = 0; k < 4; k++ )
                {
                    const int   xx = std::min( std::max( int( fx ) + ( k & 1 ), 0 ), sw - 1 );
                    const int   yy = std::min( std::max( int( fy ) + ( k >> 1 ), 0 ), sh - 1 );
                    const double wk = ( ( k & 1 ) ? ax : 1 - ax ) * ( ( k >> 1 ) ? ay : 1 - ay );
                    const uint32_t c = t->levels[ l ][ size_t( yy ) * sw + xx ];
                    // sRGB channels filter in linear space; UNORM channels and
                    // alpha filter the 8-bit codes (the same value, without a
                    // /255 * 255 round trip before the rounding)
                    for( int ch = 0; ch < 3; ch++ )
                    {
                        const uint32_t b = ( c >> ( 8 * ch ) ) & 0xFFu;
                        acc[ ch ] += wk * ( t->srgb ? SrgbDecode( b ) : double( b ) );
                    }
                    acc[ 3 ] += wk * double( c >> 24 );
                }
                auto code = []( double v ) { return uint32_t( std::nearbyint( std::min( std::max( v, 0.0 ), 255.0 ) ) ); };
                uint32_t out = code( acc[ 3 ] ) << 24;
                for( int ch = 0; ch < 3; ch++ ) out |= ( t->srgb ? SrgbEncode( acc[ ch ] ) : code( acc[ ch ] ) ) << ( 8 * ch );
                dst[ size_t( y ) * dw + x ] = out;
            }
        }
        t->levels.push_back( std::move( dst ) );
        t->w.push_back( dw );
        t->h.push_back( dh );
    }
    return t;
}

extern "C" void or_texture_free( OrTexture* t )
{
    delete t;
}

extern "C" uint32_t or_texture_mip_count( const OrTexture* t )
{
    return uint32_t( t->levels.size() );
}

extern "C" void or_texture_copy_mip( const OrTexture* t, uint32_t level, uint8_t* out )
{
    std::copy( t->levels[ level ].begin(), t->levels[ level ].end(), reinterpret_cast< uint32_t* >( out ) );
}

extern "C" void or_texture_sample_lod( const OrTexture* t, double u, double v, double lod, double out[ 4 ] )
{
    if( !( lod > 0.0 ) )
    {
        SampleLevel( *t, 0, u, v, t->magNearest, out );
        return;
    }
    const double q  = double( t->levels.size() - 1 );
    const double l  = std::min( lod, q );
    const int    l0 = int( std::floor( l ) );
    const int    l1 = std::min( l0 + 1, int( q ) );
    const double f  = l - l0;
    double       a[ 4 ], b[ 4 ];
    SampleLevel( *t, l0, u, v, t->minNearest, a );
    SampleLevel( *t, l1, u, v, t->minNearest, b );
    for( int i = 0; i < 4; i++ ) out[ i ] = a[ i ] * ( 1 - f ) + b[ i ] * f;
}

extern "C" void or_texture_sample_grad( const OrTexture* t, double u, double v, double dudx, double dvdx, double dudy, double dvdy,
                                        double out[ 4 ] )
{
    const double W = t->w[ 0 ], H = t->h[ 0 ];
    const double rho = std::max( std::hypot( dudx * W, dvdx * H ), std::hypot( dudy * W, dvdy * H ) );
    or_texture_sample_lod( t, u, v, rho > 0.0 ? std::log2( rho ) : -INFINITY, out );
}
